"""The C ABI: the shared library loads without a GPU and exports every
function include/cola_halo_amd.h declares; the Python binding covers them."""
import ctypes as C

import numpy as np
import re

import cola_halo_amd as ch


def declared_functions():
    text = open(ch.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(cola_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("cola_context_create", "cola_lpt_set_displacement", "cola_move_particles",
                 "cola_pm_calculate_forces", "cola_kick", "cola_drift", "cola_set_snapshot",
                 "cola_set_noncola_initial", "cola_power_init", "cola_step"):
        assert must in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(ch.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert sorted(declared_functions()) == ch.exported_symbols()


def test_struct_layouts():
    assert ch.PARTICLE_DTYPE.itemsize == 56       # particle.h:6-12
    assert ch.PARTICLE_MIN_DTYPE.itemsize == 32   # particle.h:24-28
    assert ch.PARTICLE_DTYPE.fields["id"][1] == 48


def test_error_path_without_gpu_or_args():
    h = C.c_void_p()
    assert ch.lib().cola_context_create(15, 48, 100.0, C.byref(h)) == ch.ERR_ARG  # odd nc
    assert b"even" in ch.lib().cola_last_error()


def test_fft_radix_schedules():
    """Every schedule factors n into supported radices whose product is n."""
    for n in (96, 192, 768, 1536, 160, 250, 24):
        seen = set()
        for sch in range(5):
            r = ch.fft_radices(n, sch)
            assert r and int(np.prod(r)) == n and set(r) <= {2, 3, 4, 5, 8}, (n, sch, r)
            seen.add(tuple(r))
        assert len(seen) >= (2 if n in (96, 192, 768, 1536) else 1)
    assert ch.fft_radices(768, 0) == [8, 8, 4, 3]
    assert ch.fft_radices(768, 1) == [3, 8, 8, 4]
    assert ch.fft_radices(768, 2) == [4, 4, 4, 4, 3]
    assert ch.fft_radices(768, 3) == [4, 8, 8, 3]
    assert ch.fft_radices(768, 4) == [8, 4, 8, 3]
    assert ch.fft_radices(7 * 16, 0) == [] and ch.fft_radices(768, 5) == []


def test_ensemble_argument_errors_without_gpu():
    h = C.c_void_p()
    assert ch.lib().cola_ensemble_create(None, 0, C.byref(h)) == ch.ERR_ARG
    assert ch.lib().cola_ensemble_step(None, C.c_float(0.3), None, None) == ch.ERR_ARG
    assert ch.lib().cola_ensemble_launches(None) == 0
    assert ch.lib().cola_ensemble_destroy(None) == 0
