"""prc_d2z_history_noalloc (GetHistory of a device-resident zreion field with
no context and no allocation) against the context-based prc_d2z_history, on
two calls in a row, the second with a smaller field, through the C ABI.
Counts are integers: bit-exact."""
import numpy as np
import pytest

from _cases import field, lib_params
from conftest import announce
from prc_d2z import D2Z, D2ZError, FUSED_FP32, NHISTORY, history_noalloc

pytestmark = pytest.mark.gpu


def test_noalloc_history_matches_context_history():
    results = []
    for N in (64, 32):  # the second call has the smaller input
        announce(f"no-alloc history N={N}")
        with D2Z(lib_params(N, nscales=8, mode=FUSED_FP32)) as d:
            d.set_delta(field(N))
            d.run()
            ref = d.history()
            got = history_noalloc(d.zreion_device_ptr(), N ** 3)
            results.append((N, int(ref.sum())))
            assert got.shape == (NHISTORY,)
            assert np.array_equal(got, ref), N
            assert ref.sum() > 0
            # the same call twice: the scratch's previous counts are cleared
            assert np.array_equal(history_noalloc(d.zreion_device_ptr(), N ** 3), ref)
    print("cells counted per call:", results)


def test_noalloc_history_rejects_bad_arguments():
    announce("no-alloc history errors")
    with D2Z(lib_params(32, nscales=4, mode=FUSED_FP32)) as d:
        d.set_delta(field(32))
        d.run()
        ptr = d.zreion_device_ptr()
        with pytest.raises(D2ZError, match="ncells"):
            history_noalloc(ptr, 2 ** 32)
        with pytest.raises(D2ZError, match="ncells"):
            history_noalloc(ptr, 0)
        with pytest.raises(D2ZError, match="aligned"):
            history_noalloc(ptr + 4, 100)
        host = np.zeros(1024, np.float32)
        with pytest.raises(D2ZError, match="device memory"):
            history_noalloc(host.ctypes.data, 1024)
        # the entry point still works after the rejected calls
        assert np.array_equal(history_noalloc(ptr, 32 ** 3), d.history())
