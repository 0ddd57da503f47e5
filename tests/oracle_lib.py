"""ctypes binding of the CPU oracle (oracle/libcola_oracle.so).

Test infrastructure only: tests, the package's smoke check and bench.py's
cpu_baseline leg use it as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libcola_oracle.so")

PARTICLE_DTYPE = np.dtype([("x", "<f4", 3), ("dx1", "<f4", 3), ("dx2", "<f4", 3),
                           ("v", "<f4", 3), ("id", "<i8")])
PARTICLE_MIN_DTYPE = np.dtype([("x", "<f4", 3), ("v", "<f4", 3), ("id", "<i8")])
assert PARTICLE_DTYPE.itemsize == 56 and PARTICLE_MIN_DTYPE.itemsize == 32


class Ranlxd(C.Structure):
    _fields_ = [("x", C.c_double * 12), ("carry", C.c_double), ("ir", C.c_int),
                ("jr", C.c_int), ("ir_old", C.c_int), ("pr", C.c_int)]


def _ptr(a, ctype=C.c_float):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def load():
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
    lib = C.CDLL(LIB_PATH)
    d, f, i, vp = C.c_double, C.c_float, C.c_int, C.c_void_p
    sig = {
        "orc_ranlxd_set": (None, [C.POINTER(Ranlxd), C.c_ulong, i]),
        "orc_ranlxd_uniform": (d, [C.POINTER(Ranlxd)]),
        "orc_seedtable": (None, [i, i, vp]),
        "orc_power_set_table": (i, [vp, vp, i]),
        "orc_power_read_file": (i, [C.c_char_p]),
        "orc_powerspec": (d, [d]),
        "orc_tophat_sigma2": (d, [d]),
        "orc_power_check_sigma8": (i, [d, C.POINTER(d)]),
        "orc_growth_factor": (d, [d, d, d, d]),
        "orc_hyp2f1": (d, [d, d, d, d]),
        "orc_growthDtemp": (f, [f, f]),
        "orc_growthD": (f, [f, f]),
        "orc_growthD2": (f, [f, f]),
        "orc_growthD2v": (f, [f, f]),
        "orc_Qfactor": (f, [f, f]),
        "orc_DprimeQ": (d, [d, f, f]),
        "orc_Sq": (d, [d, d, d, f]),
        "orc_Sphi": (d, [d, d, d, f]),
        "orc_fft_1d": (None, [i, vp, i]),
        "orc_fft_r2c_3d": (None, [i, vp]),
        "orc_fft_c2r_3d": (None, [i, vp]),
        "orc_lpt_set_displacement": (i, [i, d, d, i, d, vp, vp, vp, vp]),
        "orc_lpt_from_modes": (i, [i, d, vp, vp, vp]),
        "orc_lpt_modes": (i, [i, d, i, vp]),
        "orc_pm_readout": (None, [i, f, vp, i, vp, vp]),
        "orc_wrap": (None, [vp, i, f]),
        "orc_pm_deposit": (None, [i, f, f, vp, i, vp]),
        "orc_density_sum": (d, [i, vp]),
        "orc_pm_calculate_forces": (i, [i, f, f, C.c_longlong, vp, i, vp, vp, vp, C.POINTER(d)]),
        "orc_kick": (None, [vp, vp, i, f, C.POINTER(f), C.POINTER(f), f]),
        "orc_drift": (None, [vp, i, f, C.POINTER(f), C.POINTER(f), f]),
        "orc_set_snapshot": (None, [d, vp, vp, i, f, f, f, vp]),
        "orc_set_noncola_initial": (None, [vp, i, f, f, vp]),
        "orc_run": (i, [i, i, d, d, d, i, i, vp, i, vp, vp, C.POINTER(i)]),
        "orc_bf16_round": (d, [d]),
        "orc_conv3d": (None, [i, i, i, vp, vp, vp, vp]),
        "orc_net_forward": (d, [i, vp, vp, vp, vp, i]),
        "orc_net_gradient": (d, [i, vp, vp, vp, vp, i]),
        "orc_fof": (i, [vp, C.c_long, i, f, f, i, vp, vp]),
        "orc_coarse_grid": (None, [vp, C.c_long, i, i, f, vp]),
        "orc_density_contrast": (None, [vp, i, C.c_longlong, vp]),
        "orc_power_spectrum": (None, [vp, i, d, d, d, i, i, d, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def padded_real_shape(n):
    return (n, n, 2 * (n // 2 + 1))


def seedtable(n, seed):
    t = np.zeros(n * n, dtype=np.uint32)
    lib().orc_seedtable(n, seed, t.ctypes.data)
    return t.reshape(n, n)


def set_power_table(k, p):
    k = np.ascontiguousarray(k, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64)
    return lib().orc_power_set_table(k.ctypes.data, p.ctypes.data, len(k))


def lpt(nc, a_init, omega_m, seed, box, want_grids=False):
    parts = np.zeros(nc ** 3, dtype=PARTICLE_DTYPE)
    cs = nc * nc * (nc // 2 + 1) * 2
    cdisp = np.zeros(3 * cs, dtype=np.float32) if want_grids else None
    disp = np.zeros(3 * cs, dtype=np.float32) if want_grids else None
    disp2 = np.zeros(3 * cs, dtype=np.float32) if want_grids else None
    st = lib().orc_lpt_set_displacement(nc, a_init, omega_m, seed, box, parts.ctypes.data,
                                        cdisp.ctypes.data if want_grids else None,
                                        disp.ctypes.data if want_grids else None,
                                        disp2.ctypes.data if want_grids else None)
    assert st == 0
    if want_grids:
        return parts, cdisp.reshape(3, -1), disp.reshape(3, -1), disp2.reshape(3, -1)
    return parts


def lpt_modes(n, box, seed):
    """The white-noise ZA displacement spectra of seed: complex64 [3, n, n, n/2+1]."""
    c = np.zeros((3, n, n, n // 2 + 1), np.complex64)
    assert lib().orc_lpt_modes(n, box, seed, c.ctypes.data) == 0
    return c


def lpt_from_modes(n, box, cdisp):
    """2LPT stages on ZA spectra cdisp (complex [3, n, n, n/2+1]).  Returns
    (disp, disp2) as real [3, n, n, n]: the ZA displacement and the 2LPT
    displacement with lpt.c's 1/n^3 applied."""
    c = np.ascontiguousarray(cdisp, np.complex64)
    assert c.shape == (3, n, n, n // 2 + 1)
    disp = np.zeros((3,) + padded_real_shape(n), np.float32)
    disp2 = np.zeros((3,) + padded_real_shape(n), np.float32)
    assert lib().orc_lpt_from_modes(n, box, c.ctypes.data, disp.ctypes.data, disp2.ctypes.data) == 0
    return disp[..., :n], disp2[..., :n]


def pm_readout(ngrid, box, parts, grid):
    """CIC readout (with ghosts) of a real grid [ngrid, ngrid, ngrid] or its
    padded layout at the particles."""
    g = np.zeros(padded_real_shape(ngrid), np.float32)
    g[..., :ngrid] = np.asarray(grid)[..., :ngrid]
    parts = np.ascontiguousarray(parts)
    out = np.zeros(len(parts), np.float32)
    lib().orc_pm_readout(ngrid, box, parts.ctypes.data, len(parts), g.ctypes.data, out.ctypes.data)
    return out


def pm_forces(ngrid, box, wpar, parts, want_grids=False):
    np_ = len(parts)
    force = np.zeros((np_, 3), dtype=np.float32)
    gsize = ngrid * ngrid * 2 * (ngrid // 2 + 1)
    dens = np.zeros(gsize, dtype=np.float32) if want_grids else None
    fgrid = np.zeros(3 * gsize, dtype=np.float32) if want_grids else None
    msum = C.c_double(0)
    st = lib().orc_pm_calculate_forces(ngrid, box, wpar, np_, parts.ctypes.data, np_, force.ctypes.data,
                                       dens.ctypes.data if want_grids else None,
                                       fgrid.ctypes.data if want_grids else None, C.byref(msum))
    out = {"status": st, "force": force, "mass_sum": msum.value}
    if want_grids:
        out["density"] = dens.reshape(padded_real_shape(ngrid))
        out["force_grid"] = fgrid.reshape((3,) + padded_real_shape(ngrid))
    return out


def coarse_grid(pos, n, box):
    """coarse_grid.c:95-121 then :59-70 on positions pos[np, 3] (or a record
    array with an "x" field).  Returns (delta[n, n, n], (sum rho, sum delta))."""
    if pos.dtype.names:
        stride = pos.dtype.itemsize // 4
        base = pos
    else:
        base = np.ascontiguousarray(pos, dtype=np.float32)
        stride = 3
    grid = np.zeros(n ** 3, np.float32)
    lib().orc_coarse_grid(base.ctypes.data, stride, len(pos), n, box, grid.ctypes.data)
    sums = np.zeros(2)
    lib().orc_density_contrast(grid.ctypes.data, n, len(pos), sums.ctypes.data)
    return grid.reshape(n, n, n), (sums[0], sums[1])


def power_spectrum(delta, box, kmin, kmax, nbins, flags=0, np_shot=1.0):
    """Shell power spectrum of a real grid delta[n, n, n]."""
    delta = np.ascontiguousarray(delta, dtype=np.float32)
    n = delta.shape[0]
    k = np.zeros(nbins)
    p = np.zeros(nbins)
    m = np.zeros(nbins)
    lib().orc_power_spectrum(delta.ctypes.data, n, box, kmin, kmax, nbins, flags, float(np_shot), k.ctypes.data,
                             p.ctypes.data, m.ctypes.data)
    return k, p, m


HALO_DTYPE = np.dtype([("nfof", "<i4"), ("x", "<f4", 3), ("v", "<f4", 3)])


def fof(records, box, ll, nfof_min=32):
    """Friends-of-friends on snapshot records (wrapped in place, like the
    reference).  Returns (labels, halos)."""
    assert records.dtype == PARTICLE_MIN_DTYPE and records.flags.c_contiguous
    n = len(records)
    label = np.zeros(n, np.int32)
    halos = np.zeros(max(n, 1), HALO_DTYPE)
    nh = lib().orc_fof(records.ctypes.data, PARTICLE_MIN_DTYPE.itemsize // 4, n, box, ll, nfof_min,
                       label.ctypes.data, halos.ctypes.data)
    return label, halos[:nh].copy()


NET_CIN = (8, 16, 16, 16)
NET_COUT = (16, 16, 16, 3)
NET_NPARAM = sum(27 * a * b + b for a, b in zip(NET_CIN, NET_COUT))


def _dptr(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data


def conv3d(x, w, b):
    """x [n,n,n,cin], w [27,cin,cout], b [cout] -> [n,n,n,cout] (periodic)."""
    x = np.ascontiguousarray(x, np.float64)
    n, cin = x.shape[0], x.shape[3]
    cout = w.shape[2]
    out = np.zeros((n, n, n, cout))
    w = np.ascontiguousarray(w, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    lib().orc_conv3d(n, cin, cout, x.ctypes.data, w.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def net_forward(params, x, target=None, round_bf16=False):
    params = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    n = x.shape[0]
    y = np.zeros((n, n, n, 3))
    t = np.ascontiguousarray(target, np.float64) if target is not None else None
    loss = lib().orc_net_forward(n, params.ctypes.data, x.ctypes.data, t.ctypes.data if t is not None else None,
                                 y.ctypes.data, 1 if round_bf16 else 0)
    return y, loss


def net_gradient(params, x, target, round_bf16=False):
    params = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(target, np.float64)
    g = np.zeros(NET_NPARAM)
    loss = lib().orc_net_gradient(x.shape[0], params.ctypes.data, x.ctypes.data, t.ctypes.data, g.ctypes.data,
                                  1 if round_bf16 else 0)
    return loss, g
