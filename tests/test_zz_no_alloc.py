"""cola_coarse_grid_noalloc (no counterpart in the reference: the coarse grid
for callers that cannot allocate device memory) against cola_coarse_grid on
the same positions: the same kernels, so the grids are equal bit for bit
(the CIC sums are order-independent 64-bit fixed point) and the two sums
agree to double rounding (they are summed with double atomics).  Two calls in
a row, the second on a smaller grid, so the second must not see the first's
accumulators; a grid above the scratch's 128^3 is refused.  Device memory
for the outputs comes from hipMalloc of the HIP runtime the library links
against (the caller owns its outputs)."""
import ctypes as C

import numpy as np
import pytest

import cola_halo_amd as ch

pytestmark = pytest.mark.gpu

OM = 0.273


def _hip():
    ch.lib()  # the library's HIP runtime is loaded first; this is the same object
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def _dev_alloc(h, nbytes):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), nbytes) == 0
    return p.value


def _to_host(h, ptr, out):
    assert h.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
    return out


def test_noalloc_coarse_grid_equals_allocating_entry_point(eh_table):
    print("coarse grid without allocation: 32^3 particles, n = 128 then 48", flush=True)
    nc, ng, box = 32, 96, 100.0
    k, p = eh_table
    h = _hip()
    delta_dev = _dev_alloc(h, 4 * 128 ** 3)
    sums_dev = _dev_alloc(h, 16)
    try:
        with ch.Context(nc, ng, box) as ctx:
            ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
            ctx.lpt_set_displacement(0.1, OM, 5)
            ctx.set_noncola_initial(float(np.float32(OM)))  # snapshot records of the initial conditions
            for n, src in ((128, "snapshot"), (48, "particles")):
                ref, rsums = ctx.coarse_grid(n, source=src)
                ctx.coarse_grid_noalloc(n, delta_dev, sums_dev, source=src)
                ctx.synchronize()
                got = _to_host(h, delta_dev, np.empty((n, n, n), np.float32))
                sums = _to_host(h, sums_dev, np.zeros(2))
                assert np.array_equal(got, ref), n
                assert sums[0] == pytest.approx(rsums[0], rel=1e-12, abs=1e-9)
                assert sums[1] == pytest.approx(rsums[1], rel=1e-9, abs=1e-6)
                assert abs(sums[0] - nc ** 3) < 0.1 * sums[0]  # coarse_grid.c:84, the caller's check
            # sums kept in the scratch when the caller passes none: the grid is the same
            ctx.coarse_grid_noalloc(40, delta_dev, 0, source="particles")
            ctx.synchronize()
            ref, _ = ctx.coarse_grid(40, source="particles")
            assert np.array_equal(_to_host(h, delta_dev, np.empty((40, 40, 40), np.float32)), ref)
            with pytest.raises(ch.ColaError):
                ctx.coarse_grid_noalloc(129, delta_dev, sums_dev)
    finally:
        h.hipFree(C.c_void_p(delta_dev))
        h.hipFree(C.c_void_p(sums_dev))
