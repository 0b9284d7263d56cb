"""lbm_order_director_noalloc against lbm_order_director (needs a GPU).

Two calls in a row through the C ABI, the second on a smaller grid, in both
precisions: the no-allocation variant must return exactly what the
allocating entry point returns (the same kernel on the same state). A grid
past the 2^19-cell staging limit is refused with LBM_ERR_INVALID and a
message, without touching the device array.
"""
import ctypes as C

import numpy as np
import pytest

from lbman3d_amd import abi

pytestmark = pytest.mark.gpu


def run(dims, precision, steps):
    from lbman3d_amd.solver import DeviceSolver
    cfg = abi.make_config(*dims, abi.ChannelConfig, abi.Params(), precision=precision)
    with DeviceSolver(cfg) as s:
        s.upload_state(q=abi.seed_director_field(5, *dims))
        s.step(steps)
        o_ref, d_ref = s.order_director()
        o, d = s.order_director(noalloc=True)
    return o_ref, d_ref, o, d


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_SHIFTED])
def test_noalloc_matches_allocating_entry_point_on_two_calls(precision):
    for dims, steps in (((40, 36, 32), 7), ((24, 20, 16), 3)):
        print(f"case order/director noalloc {dims} precision={precision}", flush=True)
        o_ref, d_ref, o, d = run(dims, precision, steps)
        np.testing.assert_array_equal(o, o_ref)
        np.testing.assert_array_equal(d, d_ref)
        assert np.isfinite(o).all() and (o > 0).all()


def test_noalloc_refuses_grids_past_the_staging_limit():
    from lbman3d_amd.solver import DeviceSolver, lib
    print("case order/director noalloc size limit", flush=True)
    dims = (128, 64, 65)   # 532,480 cells > 2^19
    cfg = abi.make_config(*dims, abi.FullyPeriodicConfig, abi.Params(), precision=abi.FP32_SHIFTED)
    with DeviceSolver(cfg) as s:
        o, d = np.empty(s.V), np.empty(3 * s.V)
        p = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        r = lib().lbm_order_director_noalloc(s.h, p(o), p(d))
        assert r == abi.LBM_ERR_INVALID
        assert b"staging limit" in lib().lbm_last_error()
