"""cola_halo_amd: MI355X-native 2LPT / particle-mesh / COLA time stepping.

Python binding of the C ABI in include/cola_halo_amd.h (libcola_halo_amd.so,
built in-tree by ``python -c "import __graft_entry__; __graft_entry__.build()"``
or ``make -C cola_halo_amd/csrc``).  The method names follow the reference's
module functions (power_init, lpt_set_displacement, move_particles2,
pm_calculate_forces, cola_kick, cola_drift, cola_set_snapshot, ...).  Every
call runs on the GPU through the HIP library; there is no CPU fallback, and a
missing library raises immediately.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcola_halo_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "cola_halo_amd.h")

# particle.h:6-12 and particle.h:24-28
PARTICLE_DTYPE = np.dtype([("x", "<f4", 3), ("dx1", "<f4", 3), ("dx2", "<f4", 3),
                           ("v", "<f4", 3), ("id", "<i8")])
PARTICLE_MIN_DTYPE = np.dtype([("x", "<f4", 3), ("v", "<f4", 3), ("id", "<i8")])

# status codes (msg_abort codes of the reference)
OK = 0
ERR_ALLOC = 4001
ERR_POWER_FILE = 3000
ERR_SIGMA8 = 3010
ERR_MASS = 6040
ERR_ARG = 9001
ERR_UNSUPPORTED = 9002
ERR_HIP = 9003
ERR_STATE = 9004


class ColaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cola_halo_amd error {code}: {msg}")
        self.code = code


_lib_handle = None

# name -> (restype, argtypes)
_SIGNATURES = {
    "cola_last_error": (C.c_char_p, []),
    "cola_context_create": (C.c_int, [C.c_int, C.c_int, C.c_double, C.POINTER(C.c_void_p)]),
    "cola_context_destroy": (C.c_int, [C.c_void_p]),
    "cola_synchronize": (C.c_int, [C.c_void_p]),
    "cola_power_init": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double]),
    "cola_power_init_table": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_double,
                                        C.c_double, C.c_double]),
    "cola_power_spec": (C.c_double, [C.c_void_p, C.c_double]),
    "cola_growth_factor": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
    "cola_power_sigma8": (C.c_double, [C.c_void_p]),
    "cola_lpt_set_displacement": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_double]),
    "cola_lpt_modes": (C.c_int, [C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_size_t]),
    "cola_lpt_get_local_nx": (C.c_int, [C.c_void_p]),
    "cola_move_particles": (C.c_int, [C.c_void_p]),
    "cola_pm_calculate_forces": (C.c_int, [C.c_void_p]),
    "cola_pm_assign": (C.c_int, [C.c_void_p]),
    "cola_kick": (C.c_int, [C.c_void_p, C.c_float, C.c_float]),
    "cola_drift": (C.c_int, [C.c_void_p, C.c_float, C.c_float]),
    "cola_step": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float]),
    "cola_set_snapshot": (C.c_int, [C.c_void_p, C.c_double, C.c_float, C.c_void_p, C.c_size_t]),
    "cola_set_noncola_initial": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p, C.c_size_t]),
    "cola_np": (C.c_longlong, [C.c_void_p]),
    "cola_get_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "cola_set_times": (C.c_int, [C.c_void_p, C.c_float, C.c_float]),
    "cola_get_particles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "cola_set_particles": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_float, C.c_float]),
    "cola_get_forces": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "cola_get_density": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "cola_get_force_grid": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]),
    "cola_last_mass_sum": (C.c_double, [C.c_void_p]),
    "cola_read_parameters": (C.c_int, [C.c_char_p, C.c_void_p]),
    "cola_confirm_parameters": (C.c_int, [C.c_void_p]),
    "cola_read_gadget": (C.c_int, [C.c_char_p, C.c_void_p, C.c_longlong, C.c_void_p]),
    "cola_set_snapshot_records": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "cola_write_gadget": (C.c_int, [C.c_char_p, C.c_void_p, C.c_longlong, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int]),
    "cola_write_coarse_grid": (C.c_int, [C.c_char_p, C.c_float, C.c_int, C.c_void_p]),
    "cola_fof_find_halos": (C.c_int, [C.c_void_p, C.c_float, C.c_int, C.POINTER(C.c_longlong)]),
    "cola_fof_get_halos": (C.c_int, [C.c_void_p, C.c_void_p, C.c_longlong]),
    "cola_fof_get_labels": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "cola_get_snapshot": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "cola_write_halos": (C.c_int, [C.c_char_p, C.c_void_p, C.c_longlong]),
    "cola_subsampler_create": (C.c_int, [C.c_double, C.c_uint, C.POINTER(C.c_void_p)]),
    "cola_subsampler_destroy": (C.c_int, [C.c_void_p]),
    "cola_write_subsample": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_longlong, C.c_float, C.c_float,
                                       C.c_float, C.c_float, C.POINTER(C.c_longlong)]),
    "cola_resnet_nparam": (C.c_int, []),
    "cola_resnet_create": (C.c_int, [C.c_int, C.c_uint, C.POINTER(C.c_void_p)]),
    "cola_resnet_destroy": (C.c_int, [C.c_void_p]),
    "cola_resnet_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cola_resnet_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cola_resnet_get_scales": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "cola_resnet_set_scales": (C.c_int, [C.c_void_p, C.c_float, C.c_float]),
    "cola_resnet_fit_scale": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_resnet_forward_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cola_resnet_forward_features_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cola_resnet_time_forward": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "cola_resnet_set_restriction": (C.c_int, [C.c_void_p, C.c_float]),
    "cola_resnet_set_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_resnet_set_plane_order": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_resnet_set_head_fusion": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_resnet_gradient_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                                            C.c_void_p]),
    "cola_resnet_adam_step": (C.c_int, [C.c_void_p, C.c_float]),
    "cola_copy_particles": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cola_resnet_add_pair": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "cola_resnet_num_samples": (C.c_int, [C.c_void_p]),
    "cola_resnet_get_sample": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "cola_resnet_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_uint, C.c_void_p]),
    "cola_resnet_eval": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "cola_set_residual_net": (C.c_int, [C.c_void_p, C.c_void_p]),
    "cola_resnet_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "cola_resnet_load": (C.c_int, [C.c_void_p, C.c_char_p]),
    "cola_set_step_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_step_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_float, C.c_void_p, C.c_void_p]),
    "cola_ensemble_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "cola_ensemble_step": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p]),
    "cola_ensemble_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_ensemble_launches": (C.c_int, [C.c_void_p]),
    "cola_ensemble_set_merge": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_ensemble_destroy": (C.c_int, [C.c_void_p]),
    "cola_set_fft_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "cola_set_fft_tiles_xy": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cola_autotune": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
    "cola_fft_time": (C.c_int, [C.c_int, C.c_int, C.c_void_p]),
    "cola_coarse_grid": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "cola_coarse_grid_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "cola_power_spectrum": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_void_p]),
    "cola_state_save": (C.c_int, [C.c_void_p]),
    "cola_state_restore": (C.c_int, [C.c_void_p]),
    "cola_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_phase_count": (C.c_int, []),
    "cola_phase_name": (C.c_char_p, [C.c_int]),
    "cola_get_phase_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "cola_reset_phase_stats": (C.c_int, [C.c_void_p]),
    "cola_set_tiling": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cola_set_fft_schedule": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_fft_radices": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int]),
    "cola_set_fft_xcd_group": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_set_readout_group": (C.c_int, [C.c_void_p, C.c_int]),
    "cola_eh98_nowiggle_table": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_double, C.c_int, C.c_void_p, C.c_void_p]),
    "cola_bandwidth_probe": (C.c_int, [C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "cola_lpt_seedtable": (C.c_int, [C.c_int, C.c_int, C.c_void_p]),
    "cola_power_table_sigma8": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "cola_power_table_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    "cola_power_file_sigma8": (C.c_int, [C.c_char_p, C.POINTER(C.c_double)]),
    "cola_ranlxd1_uniforms": (C.c_int, [C.c_ulong, C.c_int, C.c_void_p]),
    "cola_growth_scalars": (C.c_int, [C.c_float, C.c_float, C.c_void_p]),
    "cola_step_factors": (C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_void_p]),
    "cola_fft_r2c_3d": (C.c_int, [C.c_int, C.c_void_p]),
    "cola_fft_c2r_3d": (C.c_int, [C.c_int, C.c_void_p]),
}


def lib():
    """Load libcola_halo_amd.so (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with make -C cola_halo_amd/csrc "
                              "(there is no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib_handle = h
    return _lib_handle


def exported_symbols():
    return sorted(_SIGNATURES)


def _check(code):
    if code != 0:
        raise ColaError(code, lib().cola_last_error().decode(errors="replace"))


# Cosmology of the benchmark configurations
EH_COSMOLOGY = dict(omega_m=0.273, omega_b=0.046, h=0.705, ns=0.96, sigma8=0.8)


def eh98_table(omega_m=0.273, omega_b=0.046, h=0.705, ns=0.96, sigma8=0.8, kmin=1e-4, kmax=1e2, n=500):
    """Eisenstein & Hu (1998) no-wiggle linear P(k) table [k h/Mpc, P (Mpc/h)^3],
    normalised to sigma8 by power.c's integral."""
    k = np.zeros(n)
    p = np.zeros(n)
    _check(lib().cola_eh98_nowiggle_table(omega_m, omega_b, h, ns, sigma8, kmin, kmax, n, k.ctypes.data,
                                          p.ctypes.data))
    return k, p


def power_table_sigma8(k, pk):
    """sigma8 of a (k, P) table by power.c's integral (host code)."""
    k = np.ascontiguousarray(k, np.float64)
    pk = np.ascontiguousarray(pk, np.float64)
    s = C.c_double()
    _check(lib().cola_power_table_sigma8(k.ctypes.data, pk.ctypes.data, len(k), C.byref(s)))
    return s.value


def power_table_eval(k, pk, kq):
    """power.c PowerSpec of a (k, P) table at the wave numbers kq (host code)."""
    k = np.ascontiguousarray(k, np.float64)
    pk = np.ascontiguousarray(pk, np.float64)
    kq = np.ascontiguousarray(np.atleast_1d(kq), np.float64)
    out = np.zeros(len(kq))
    _check(lib().cola_power_table_eval(k.ctypes.data, pk.ctypes.data, len(k), kq.ctypes.data, len(kq),
                                       out.ctypes.data))
    return out


def power_file_sigma8(filename):
    s = C.c_double()
    _check(lib().cola_power_file_sigma8(os.fsencode(filename), C.byref(s)))
    return s.value


def seedtable(nmesh, seed):
    """lpt.c:172-207 seed table (host code of the library)."""
    t = np.zeros(nmesh * nmesh, np.uint32)
    _check(lib().cola_lpt_seedtable(nmesh, seed, t.ctypes.data))
    return t.reshape(nmesh, nmesh)


def ranlxd1_uniforms(seed, n):
    out = np.zeros(n, np.float64)
    _check(lib().cola_ranlxd1_uniforms(seed, n, out.ctypes.data))
    return out


def growth_scalars(omega_m, a):
    """(growthD, growthD2, growthD2v, Qfactor, DprimeQ) of cola.c at a."""
    out = np.zeros(5)
    _check(lib().cola_growth_scalars(omega_m, a, out.ctypes.data))
    return out


def step_factors(omega_m, a_x, a_v, avel1, apos1):
    """(dda, q1, q2, dyyy, da1, da2) of cola_kick / cola_drift."""
    out = np.zeros(6)
    _check(lib().cola_step_factors(omega_m, a_x, a_v, avel1, apos1, out.ctypes.data))
    return out


def bandwidth_probe(nbytes=1 << 30, reps=20):
    """Copy bandwidth of a plain streaming kernel on the device (GB/s)."""
    g = C.c_double()
    _check(lib().cola_bandwidth_probe(nbytes, reps, C.byref(g)))
    return g.value


def padded_real_shape(n):
    return (n, n, 2 * (n // 2 + 1))


def fft_r2c_3d(grid):
    """FFTW-convention r2c of a real (n, n, n) array on the GPU; returns the
    (n, n, n//2+1) complex64 half spectrum."""
    g = np.asarray(grid, dtype=np.float32)
    n = g.shape[0]
    buf = np.zeros(padded_real_shape(n), np.float32)
    buf[:, :, :n] = g
    _check(lib().cola_fft_r2c_3d(n, buf.ctypes.data))
    return buf.view(np.complex64)


def fft_c2r_3d(spec):
    """FFTW-convention c2r (unnormalised) of an (n, n, n//2+1) half spectrum."""
    s = np.ascontiguousarray(spec, dtype=np.complex64)
    n = s.shape[0]
    # a private copy: the transform runs in place and must not touch spec
    buf = np.array(s.view(np.float32).reshape(padded_real_shape(n)), copy=True)
    _check(lib().cola_fft_c2r_3d(n, buf.ctypes.data))
    return buf[:, :, :n].copy()


HALO_DTYPE = np.dtype([("nfof", "<i4"), ("x", "<f4", 3), ("v", "<f4", 3)])  # cola_halo, 28 bytes


def write_halos(filename, halos):
    """fof.c:1166-1168 text lines."""
    halos = np.ascontiguousarray(halos, HALO_DTYPE)
    _check(lib().cola_write_halos(os.fsencode(filename), halos.ctypes.data, len(halos)))


class Subsampler:
    """subsample.c: RANLUX stream seeded with 2 * seed, kept across outputs."""

    def __init__(self, factor, seed):
        self._h = C.c_void_p()
        _check(lib().cola_subsampler_create(factor, seed, C.byref(self._h)))

    def write(self, filename, snap, a, boxsize, omega_m, h):
        snap = np.ascontiguousarray(snap, PARTICLE_MIN_DTYPE)
        n = C.c_longlong(0)
        _check(lib().cola_write_subsample(self._h, os.fsencode(filename), snap.ctypes.data, len(snap), a, boxsize,
                                          omega_m, h, C.byref(n)))
        return n.value

    def close(self):
        if self._h:
            lib().cola_subsampler_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


class GadgetInfo(C.Structure):
    """cola_gadget_info."""
    _fields_ = [("nfiles", C.c_int), ("np_read", C.c_longlong), ("np_total", C.c_longlong), ("a", C.c_float),
                ("boxsize", C.c_float), ("omega_m", C.c_float), ("h", C.c_float)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Tuning(C.Structure):
    """cola_tuning: the launch shapes the autotuner chose."""
    _fields_ = [("tile", C.c_int * 3), ("fft_log2_cols", C.c_int), ("fft_log2_rows", C.c_int),
                ("xcd_group", C.c_int), ("fft_log2_xcols", C.c_int), ("ms_before", C.c_double),
                ("ms_after", C.c_double), ("fft_schedule", C.c_int), ("fft_radices", C.c_int * 8)]

    def as_dict(self):
        return {"tile": list(self.tile), "fft_log2_cols": self.fft_log2_cols, "fft_log2_rows": self.fft_log2_rows,
                "fft_log2_xcols": self.fft_log2_xcols, "xcd_group": self.xcd_group,
                "fft_schedule": self.fft_schedule, "fft_radices": [r for r in self.fft_radices if r],
                "ms_before": round(self.ms_before, 4), "ms_after": round(self.ms_after, 4)}


class ResidualNet:
    """The learned coarse-to-fine force residual (cola_resnet_*)."""

    def __init__(self, n, seed=1):
        self.n = n
        self._h = C.c_void_p()
        _check(lib().cola_resnet_create(n, seed, C.byref(self._h)))

    @staticmethod
    def nparam():
        return lib().cola_resnet_nparam()

    def params(self):
        out = np.zeros(self.nparam(), np.float32)
        _check(lib().cola_resnet_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparam()
        _check(lib().cola_resnet_set_params(self._h, p.ctypes.data))

    def scales(self):
        a, b = C.c_float(), C.c_float()
        _check(lib().cola_resnet_get_scales(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_scales(self, in_scale, out_scale):
        _check(lib().cola_resnet_set_scales(self._h, in_scale, out_scale))

    def fit_scale(self, count):
        _check(lib().cola_resnet_fit_scale(self._h, count))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[:3] + (3,), np.float32)
        _check(lib().cola_resnet_forward_host(self._h, x.ctypes.data, y.ctypes.data))
        return y

    def forward_features(self, x):
        """The inference path on x [n,n,n,8]: channel 0 the feature, channel 1
        the scale factor a in every cell, channels 2-7 zero; the first layer
        runs on the feature channel with a folded into its bias."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(x.shape[:3] + (3,), np.float32)
        _check(lib().cola_resnet_forward_features_host(self._h, x.ctypes.data, y.ctypes.data))
        return y

    def time_forward(self, reps=20):
        """Mean device time of one inference forward pass on the feature grid
        of the last host input (ms)."""
        ms = C.c_double()
        _check(lib().cola_resnet_time_forward(self._h, reps, C.byref(ms)))
        return ms.value

    def gradient(self, x, target):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_double()
        g = np.zeros(self.nparam(), np.float32)
        _check(lib().cola_resnet_gradient_host(self._h, x.ctypes.data, t.ctypes.data, C.byref(loss), g.ctypes.data))
        return loss.value, g

    def adam_step(self, lr):
        """One Adam step on the gradient of the last gradient() call."""
        _check(lib().cola_resnet_adam_step(self._h, lr))

    def set_restriction(self, side_weight):
        """Training targets: the fine force averaged over the 3^3 fine
        neighbours of each coarse cell with weights (w, 1 - 2w, w) per axis
        (0 = the point value, 0.25 = full weighting, default 0.125)."""
        _check(lib().cola_resnet_set_restriction(self._h, side_weight))

    def set_fusion(self, on):
        """Inference launch shape in a context's force: features and layers
        1-2 in one pass (True, the default) or three launches (False); the
        forces are the same bit for bit."""
        _check(lib().cola_resnet_set_fusion(self._h, int(bool(on))))

    def set_head_fusion(self, on):
        """Layer 3 and the last layer in one launch after the force
        transforms (True, the default) or two launches through h3 (False);
        the forces are the same bit for bit."""
        _check(lib().cola_resnet_set_head_fusion(self._h, int(bool(on))))

    def set_plane_order(self, by_input_planes):
        """Loop order of the hidden 16 -> 16 layers: by input planes (True,
        the default: k_conv16r) or by output planes (False: k_conv16);
        results agree to float rounding."""
        _check(lib().cola_resnet_set_plane_order(self._h, int(bool(by_input_planes))))

    def add_pair(self, coarse, fine):
        _check(lib().cola_resnet_add_pair(self._h, coarse._h, fine._h))

    def sample(self, i):
        """Stored training pair i: (feature [n,n,n], target [3,n,n,n], a)."""
        n = self.n
        f = np.zeros((n, n, n), np.float32)
        t = np.zeros((3, n, n, n), np.float32)
        a = C.c_float()
        _check(lib().cola_resnet_get_sample(self._h, i, f.ctypes.data, t.ctypes.data, C.byref(a)))
        return f, t, a.value

    def num_samples(self):
        return lib().cola_resnet_num_samples(self._h)

    def train(self, ntrain, epochs, lr=1e-3, seed=1):
        out = np.zeros(epochs)
        _check(lib().cola_resnet_train(self._h, ntrain, epochs, lr, seed, out.ctypes.data))
        return out

    def eval(self, first, count):
        loss = C.c_double()
        _check(lib().cola_resnet_eval(self._h, first, count, C.byref(loss)))
        return loss.value

    def save(self, filename):
        """Write the parameters and scales to a net-weights file."""
        _check(lib().cola_resnet_save(self._h, os.fsencode(filename)))

    def load(self, filename):
        """Parameters, scales and restriction weight from a net-weights file."""
        _check(lib().cola_resnet_load(self._h, os.fsencode(filename)))

    @classmethod
    def from_file(cls, n, filename):
        """A net on an n^3 grid with the weights of a net-weights file."""
        net = cls(n)
        try:
            net.load(filename)
        except Exception:
            net.close()
            raise
        return net

    def close(self):
        if self._h:
            lib().cola_resnet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


PARAM_MAX_ZOUT = 256
PARAM_MAX_STRING = 1024
_PARAM_STRINGS = ("power_spectrum_filename", "fof_filename", "snapshot_filename", "subsample_filename",
                  "cgrid_filename", "init_filename")


class Params(C.Structure):
    """cola_params (parameters.h:4-49 with fixed-size storage)."""
    _fields_ = ([("nc", C.c_int), ("pm_nc_factor", C.c_int), ("np_alloc_factor", C.c_double),
                 ("ntimestep", C.c_int), ("random_seed", C.c_int), ("nrealization", C.c_int),
                 ("loglevel", C.c_int), ("a_final", C.c_double), ("boxsize", C.c_double),
                 ("omega_m", C.c_double), ("sigma8", C.c_double), ("h", C.c_double),
                 ("n_zout", C.c_int), ("zout", C.c_double * PARAM_MAX_ZOUT)]
                + [f for name in _PARAM_STRINGS
                   for f in ((name, C.c_char * PARAM_MAX_STRING), ("strlen_" + name, C.c_int))]
                + [("fof_linking_factor", C.c_double), ("subsample_factor", C.c_double),
                   ("cgrid_nc", C.c_int), ("write_longid", C.c_int), ("pm_ngrid", C.c_int),
                   ("residual_net_filename", C.c_char * PARAM_MAX_STRING),
                   ("strlen_residual_net_filename", C.c_int)])

    def as_dict(self):
        """Plain values; an absent string key is None."""
        out = {}
        for name, _ in self._fields_:
            if name.startswith("strlen_") or name == "n_zout":
                continue
            if name == "zout":
                out["zout"] = [self.zout[i] for i in range(self.n_zout)]
            elif name in _PARAM_STRINGS or name == "residual_net_filename":
                n = getattr(self, "strlen_" + name)
                out[name] = None if n == 0 else getattr(self, name).decode()
            else:
                out[name] = getattr(self, name)
        return out


def read_parameters(filename):
    """read_param.c semantics; raises ColaError with the reference's abort code."""
    p = Params()
    _check(lib().cola_read_parameters(os.fsencode(filename), C.byref(p)))
    return p


def confirm_parameters(p):
    _check(lib().cola_confirm_parameters(C.byref(p)))


def write_gadget(filename, snap, a, boxsize, omega_m, h, long_id=False):
    snap = np.ascontiguousarray(snap, PARTICLE_MIN_DTYPE)
    _check(lib().cola_write_gadget(os.fsencode(filename), snap.ctypes.data, len(snap), a, boxsize, omega_m, h,
                                   1 if long_id else 0))


def read_gadget(filename):
    """GADGET-1 snapshot (one file or filename.0, .1, ...) as read.c reads
    it: (records, info dict); records is None when no such file exists."""
    info = GadgetInfo()
    _check(lib().cola_read_gadget(os.fsencode(filename), None, 0, C.byref(info)))
    if info.nfiles == 0:
        return None, info.as_dict()
    out = np.zeros(info.np_read, PARTICLE_MIN_DTYPE)
    _check(lib().cola_read_gadget(os.fsencode(filename), out.ctypes.data, len(out), C.byref(info)))
    return out, info.as_dict()


def write_coarse_grid(filename, boxsize, delta):
    delta = np.ascontiguousarray(delta, np.float32)
    _check(lib().cola_write_coarse_grid(os.fsencode(filename), boxsize, delta.shape[0], delta.ctypes.data))


class Context:
    """One realization's device state: the reference's Particles, PM grids,
    FFT plans and the a_x / a_v clocks (main.c:52-111 set-up)."""

    def __init__(self, nc, ngrid, boxsize):
        self.nc, self.ngrid, self.boxsize = int(nc), int(ngrid), float(boxsize)
        h = C.c_void_p()
        _check(lib().cola_context_create(self.nc, self.ngrid, self.boxsize, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().cola_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def np(self):
        return lib().cola_np(self._h)

    # power.c
    def power_init(self, filename, a_init, sigma8, omega_m, omega_lambda):
        _check(lib().cola_power_init(self._h, os.fsencode(filename), a_init, sigma8, omega_m, omega_lambda))

    def power_init_table(self, k, pk, a_init, sigma8, omega_m, omega_lambda):
        k = np.ascontiguousarray(k, np.float64)
        pk = np.ascontiguousarray(pk, np.float64)
        _check(lib().cola_power_init_table(self._h, k.ctypes.data, pk.ctypes.data, len(k), a_init, sigma8,
                                           omega_m, omega_lambda))

    def power_spec(self, k):
        return lib().cola_power_spec(self._h, k)

    def growth_factor(self, astart, aend):
        return lib().cola_growth_factor(self._h, astart, aend)

    def sigma8(self):
        return lib().cola_power_sigma8(self._h)

    # lpt.c / move.c / pm.c / cola.c
    def lpt_set_displacement(self, a_init, omega_m, seed, boxsize=None):
        _check(lib().cola_lpt_set_displacement(self._h, a_init, omega_m, int(seed),
                                               self.boxsize if boxsize is None else boxsize))

    def lpt_modes(self, seed, boxsize=None):
        """The device's white-noise ZA displacement spectra of seed
        (lpt.c:172-351): complex64 [3, nc, nc, nc/2+1]."""
        n = self.nc
        out = np.zeros((3, n, n, n // 2 + 1), np.complex64)
        _check(lib().cola_lpt_modes(self._h, int(seed), self.boxsize if boxsize is None else boxsize,
                                    out.ctypes.data, out.size * 2))
        return out

    def move_particles(self):
        _check(lib().cola_move_particles(self._h))

    def pm_calculate_forces(self):
        _check(lib().cola_pm_calculate_forces(self._h))

    def pm_assign(self):
        _check(lib().cola_pm_assign(self._h))

    def kick(self, omega_m, avel1):
        _check(lib().cola_kick(self._h, omega_m, avel1))

    def drift(self, omega_m, apos1):
        _check(lib().cola_drift(self._h, omega_m, apos1))

    def step(self, omega_m, avel1, apos1):
        _check(lib().cola_step(self._h, omega_m, avel1, apos1))

    def synchronize(self):
        _check(lib().cola_synchronize(self._h))

    def set_snapshot(self, aout, omega_m):
        out = np.zeros(self.np, PARTICLE_MIN_DTYPE)
        _check(lib().cola_set_snapshot(self._h, aout, omega_m, out.ctypes.data, len(out)))
        return out

    def set_noncola_initial(self, omega_m):
        out = np.zeros(self.np, PARTICLE_MIN_DTYPE)
        _check(lib().cola_set_noncola_initial(self._h, omega_m, out.ctypes.data, len(out)))
        return out

    # state
    def set_times(self, a_x, a_v):
        _check(lib().cola_set_times(self._h, a_x, a_v))

    def times(self):
        ax, av = C.c_float(), C.c_float()
        _check(lib().cola_get_times(self._h, C.byref(ax), C.byref(av)))
        return ax.value, av.value

    def get_particles(self):
        out = np.zeros(self.np, PARTICLE_DTYPE)
        _check(lib().cola_get_particles(self._h, out.ctypes.data, len(out)))
        return out

    def set_particles(self, parts, a_x, a_v):
        parts = np.ascontiguousarray(parts, PARTICLE_DTYPE)
        _check(lib().cola_set_particles(self._h, parts.ctypes.data, len(parts), a_x, a_v))

    def get_forces(self):
        out = np.zeros((self.np, 3), np.float32)
        _check(lib().cola_get_forces(self._h, out.ctypes.data, out.size))
        return out

    def get_density(self):
        out = np.zeros(padded_real_shape(self.ngrid), np.float32)
        _check(lib().cola_get_density(self._h, out.ctypes.data, out.size))
        return out

    def get_force_grid(self, axis):
        out = np.zeros(padded_real_shape(self.ngrid), np.float32)
        _check(lib().cola_get_force_grid(self._h, axis, out.ctypes.data, out.size))
        return out

    def state_save(self):
        _check(lib().cola_state_save(self._h))

    def state_restore(self):
        _check(lib().cola_state_restore(self._h))

    def last_mass_sum(self):
        return lib().cola_last_mass_sum(self._h)

    # friends-of-friends
    def fof_find_halos(self, ll, nfof_min=32):
        """FoF on the last snapshot (wraps its positions).  Returns the halo
        records (HALO_DTYPE), ordered by the smallest member index."""
        n = C.c_longlong(0)
        _check(lib().cola_fof_find_halos(self._h, ll, nfof_min, C.byref(n)))
        out = np.zeros(n.value, HALO_DTYPE)
        _check(lib().cola_fof_get_halos(self._h, out.ctypes.data, len(out)))
        return out

    def fof_labels(self):
        out = np.zeros(self.np, np.int32)
        _check(lib().cola_fof_get_labels(self._h, out.ctypes.data, len(out)))
        return out

    def get_snapshot(self):
        out = np.zeros(self.np, PARTICLE_MIN_DTYPE)
        _check(lib().cola_get_snapshot(self._h, out.ctypes.data, len(out)))
        return out

    # output grids
    _SOURCES = {"particles": 0, "snapshot": 1}

    def coarse_grid(self, n, source="snapshot"):
        """coarse_grid.c: delta = rho n^3 / np - 1 on an n^3 CIC grid (periodic),
        C order.  Returns (delta[n, n, n], (sum of rho, sum of delta))."""
        out = np.empty((n, n, n), np.float32)
        sums = np.zeros(2, np.float64)
        _check(lib().cola_coarse_grid(self._h, n, self._SOURCES[source], out.ctypes.data, sums.ctypes.data))
        return out, (float(sums[0]), float(sums[1]))

    def coarse_grid_noalloc(self, n, delta_dev, sums_dev=0, source="snapshot"):
        """cola_coarse_grid_noalloc: the coarse grid into caller-owned device
        memory (delta_dev: address of n^3 floats, sums_dev: 0 or the address
        of two doubles), accumulated in the library's static scratch (n <=
        128); enqueued only: synchronize() before reading the outputs."""
        _check(lib().cola_coarse_grid_noalloc(self._h, n, self._SOURCES[source], C.c_void_p(delta_dev),
                                              C.c_void_p(sums_dev) if sums_dev else None))

    def power_spectrum(self, n, kmin, kmax, nbins, deconvolve=False, subtract_shot=False, source="particles"):
        """Shell-averaged P(k) of the particles on an n^3 CIC grid.  Returns
        (mean k, mean P, modes) per bin."""
        k = np.zeros(nbins)
        p = np.zeros(nbins)
        m = np.zeros(nbins)
        flags = (1 if deconvolve else 0) | (2 if subtract_shot else 0)
        _check(lib().cola_power_spectrum(self._h, n, self._SOURCES[source], kmin, kmax, nbins, flags,
                                         k.ctypes.data, p.ctypes.data, m.ctypes.data))
        return k, p, m

    # profiling / tuning
    def set_profiling(self, on=True):
        _check(lib().cola_set_profiling(self._h, 1 if on else 0))

    def phase_stats(self):
        n = lib().cola_phase_count()
        ms = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int64)
        _check(lib().cola_get_phase_stats(self._h, ms.ctypes.data, cnt.ctypes.data, n))
        return {lib().cola_phase_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}

    def reset_phase_stats(self):
        _check(lib().cola_reset_phase_stats(self._h))

    def set_tiling(self, tx, ty, tz, fft_log2_tile=-1):
        _check(lib().cola_set_tiling(self._h, tx, ty, tz, fft_log2_tile))

    def copy_particles_from(self, other):
        _check(lib().cola_copy_particles(self._h, other._h))

    def set_snapshot_records(self, records):
        """Snapshot records (e.g. read_gadget's) for the output functions."""
        records = np.ascontiguousarray(records, PARTICLE_MIN_DTYPE)
        _check(lib().cola_set_snapshot_records(self._h, records.ctypes.data, len(records)))

    def set_residual_net(self, net):
        _check(lib().cola_set_residual_net(self._h, net._h if net is not None else None))

    def set_step_graph(self, on=True):
        _check(lib().cola_set_step_graph(self._h, 1 if on else 0))

    def set_fft_tiles(self, log2_cols=-1, log2_rows=-1, xcd_group=8, log2_xcols=-1):
        """Launch shape of the PM transforms (log2_xcols: the x column pass
        apart from the y pass, -1 = as log2_cols)."""
        _check(lib().cola_set_fft_tiles_xy(self._h, log2_cols, log2_xcols, log2_rows, xcd_group))

    def apply_tuning(self, t):
        """Launch shapes from a tuning dict (autotune's result, e.g. of another
        context of the same sizes)."""
        tx, ty, tz = t["tile"]
        _check(lib().cola_set_tiling(self._h, tx, ty, tz, t["fft_log2_cols"]))
        self.set_fft_schedule(t.get("fft_schedule", 0))
        self.set_fft_tiles(t["fft_log2_cols"], t["fft_log2_rows"], t["xcd_group"], t["fft_log2_xcols"])

    def set_fft_schedule(self, schedule):
        """Radix schedule of the PM transforms (0-4, see fft_radices)."""
        _check(lib().cola_set_fft_schedule(self._h, int(schedule)))

    def autotune(self, reps=3):
        """Pick the fastest deposit tile / FFT launch shapes for the current
        particle state; returns the choice (results do not depend on it)."""
        t = Tuning()
        _check(lib().cola_autotune(self._h, reps, C.byref(t)))
        return t.as_dict()

    def set_readout_group(self, group):
        """Consecutive readout workgroups per XCD (results unchanged)."""
        _check(lib().cola_set_readout_group(self._h, group))

    def set_fft_xcd_group(self, group):
        _check(lib().cola_set_fft_xcd_group(self._h, group))


def fft_radices(n, schedule=0):
    """Radix passes of an n-point transform under a radix schedule (host only)."""
    r = np.zeros(8, np.int32)
    k = lib().cola_fft_radices(int(n), int(schedule), r.ctypes.data, 8)
    return [int(v) for v in r[:max(k, 0)]]


def step_batch(ctxs, omega_m, avel1, apos1):
    """cola_step on every context, back to back on their own streams."""
    n = len(ctxs)
    hs = (C.c_void_p * n)(*[c._h for c in ctxs])
    a1 = np.ascontiguousarray(avel1, np.float32)
    p1 = np.ascontiguousarray(apos1, np.float32)
    _check(lib().cola_step_batch(hs, n, omega_m, a1.ctypes.data, p1.ctypes.data))


class Ensemble:
    """Contexts stepped together with one kernel launch per phase over the
    workgroups of every realization (cola_ensemble_*).  Each context may be
    one member; an ensemble whose member was closed refuses to step."""

    def __init__(self, ctxs, graph=False, merge=False):
        self._h = None
        self.ctxs = list(ctxs)
        n = len(self.ctxs)
        hs = (C.c_void_p * n)(*[c._h for c in self.ctxs])
        h = C.c_void_p()
        _check(lib().cola_ensemble_create(hs, n, C.byref(h)))
        self._h = h.value
        if graph:
            self.set_graph(True)
        if merge:
            self.set_merge(True)

    def step(self, omega_m, avel1, apos1):
        n = len(self.ctxs)
        a1 = np.ascontiguousarray(np.broadcast_to(np.asarray(avel1, np.float32), (n,)))
        p1 = np.ascontiguousarray(np.broadcast_to(np.asarray(apos1, np.float32), (n,)))
        _check(lib().cola_ensemble_step(self._h, omega_m, a1.ctypes.data, p1.ctypes.data))

    def set_graph(self, on):
        _check(lib().cola_ensemble_set_graph(self._h, 1 if on else 0))

    def set_merge(self, on):
        """One launch per transform pass for all members (True) or per pass
        and FFT kernel variant (False, the default)."""
        _check(lib().cola_ensemble_set_merge(self._h, 1 if on else 0))

    def launches(self):
        """Kernel launches per ensemble step (0 before the first step)."""
        return int(lib().cola_ensemble_launches(self._h))

    def close(self):
        if self._h:
            _check(lib().cola_ensemble_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_cola_batch(ctxs, omega_m, a_final, nsteps, seeds, a_out_final=True, ensemble=None):
    """Realizations of the same schedule (main.c:117-207, one output at
    a_final) on several contexts, their steps interleaved so the GPU overlaps
    them (ensemble=None: each context's step on its own stream), or run as
    one ensemble (ensemble=Ensemble over ctxs: one launch per phase).
    Returns one snapshot per context."""
    da = a_final / nsteps
    a_init = da
    Om = float(np.float32(omega_m))
    for ctx, seed in zip(ctxs, seeds):
        ctx.lpt_set_displacement(a_init, omega_m, seed)
        ctx.set_times(ctx.times()[0], da)  # main.c:145
    aout = float(np.float32(a_final))
    for istep in range(1, nsteps + 1):
        avel0, apos0 = (istep - 0.5) * da, istep * da
        avel1, apos1 = (istep + 0.5) * da, (istep + 1.0) * da
        # main.c:176-203, as run_cola: a snapshot inside [avel0, apos0] is taken
        # from the forces; one in (apos0, avel1] after the kick to avel1
        snap_now = avel0 <= aout <= apos0
        snap_mid = (not snap_now) and apos0 < aout <= avel1
        if not (snap_now or snap_mid):
            if ensemble is not None:
                ensemble.step(Om, avel1, apos1)
            else:
                step_batch(ctxs, Om, [avel1] * len(ctxs), [apos1] * len(ctxs))
            continue
        for ctx in ctxs:
            ctx.move_particles()
            ctx.pm_calculate_forces()
            if snap_mid:
                ctx.kick(Om, avel1)
        return [ctx.set_snapshot(aout, Om) for ctx in ctxs]
    return []


def run_cola(ctx, omega_m, a_final, nsteps, seed, zout, fused=True, on_snapshot=None):
    """The reference's realization loop for one seed (main.c:117-207),
    linear-in-a schedule.  Returns the list of (aout, snapshot) pairs.  With
    fused=True the steps without a snapshot run as cola_step (wrap, force,
    fused readout+kick+drift); results equal the unfused calls bit for bit."""
    da = a_final / nsteps
    a_init = da
    aout = [float(np.float32(1.0 / (1.0 + z))) for z in zout]
    ctx.lpt_set_displacement(a_init, omega_m, seed)
    Om = float(np.float32(omega_m))
    ctx.set_times(ctx.times()[0], da)  # main.c:145
    out = []
    iout = 0
    nout = len(aout)
    if nout > 0 and nsteps > 1 and a_final > a_init:
        for istep in range(1, nsteps + 1):
            avel0 = (istep - 0.5) * da
            apos0 = istep * da
            avel1 = (istep + 0.5) * da
            apos1 = (istep + 1.0) * da
            snap_now = iout < nout and avel0 <= aout[iout] <= apos0
            snap_mid = (not snap_now) and iout < nout and apos0 < aout[iout] <= avel1
            if fused and not snap_now and not snap_mid and not (iout >= nout):
                ctx.step(Om, avel1, apos1)
                continue
            ctx.move_particles()
            ctx.pm_calculate_forces()
            while iout < nout and avel0 <= aout[iout] <= apos0:
                s = ctx.set_snapshot(aout[iout], Om)
                out.append((aout[iout], s))
                if on_snapshot:
                    on_snapshot(aout[iout], s)
                iout += 1
            if iout >= nout:
                break
            ctx.kick(Om, avel1)
            while iout < nout and apos0 < aout[iout] <= avel1:
                s = ctx.set_snapshot(aout[iout], Om)
                out.append((aout[iout], s))
                if on_snapshot:
                    on_snapshot(aout[iout], s)
                iout += 1
            if iout >= nout:
                break
            ctx.drift(Om, apos1)
    ctx.synchronize()
    return out

