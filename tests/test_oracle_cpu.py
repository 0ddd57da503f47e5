"""CPU checks that pin the oracle (and the library's host code) to
independent references: GSL known answers, numpy FFT, mpmath / scipy
integrals, closed forms and invariants.  No GPU needed."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import cola_halo_amd as ch
import oracle_lib as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OM32 = float(np.float32(0.273))


def _json(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ------------------------------------------------------------------ RANLUX
@pytest.mark.parametrize("gen,lux", [("ranlxd1", 202), ("ranlxd2", 397)])
def test_oracle_ranlux_gsl_known_answers(gen, lux):
    kat = _json("ranlux_kat.json")[gen]
    r = orc.Ranlxd()
    orc.lib().orc_ranlxd_set(C.byref(r), kat["seed"], lux)
    for _ in range(kat["n"]):
        u = orc.lib().orc_ranlxd_uniform(C.byref(r))
    assert int(u * 4294967296.0) == kat["value"]


def test_library_ranlux_matches_gsl_and_oracle():
    kat = _json("ranlux_kat.json")["ranlxd1"]
    u = ch.ranlxd1_uniforms(kat["seed"], kat["n"])
    assert int(u[-1] * 4294967296.0) == kat["value"]
    r = orc.Ranlxd()
    orc.lib().orc_ranlxd_set(C.byref(r), 12345, 202)
    ref = np.array([orc.lib().orc_ranlxd_uniform(C.byref(r)) for _ in range(1000)])
    assert np.array_equal(ch.ranlxd1_uniforms(12345, 1000), ref)
    # 48-bit values in [0, 1)
    assert np.all((ref >= 0) & (ref < 1))
    assert np.all(ref * 2.0 ** 48 == np.floor(ref * 2.0 ** 48))


@pytest.mark.parametrize("n", [4, 16, 64, 128])
def test_seedtable_library_equals_oracle(n):
    assert np.array_equal(ch.seedtable(n, 100), orc.seedtable(n, 100))
    t = orc.seedtable(n, 7)
    assert t.max() < 2 ** 31 and len(np.unique(t)) > 0.99 * n * n


# ------------------------------------------------------------------ FFT
def test_oracle_fft_golden():
    g = np.load(os.path.join(GOLD, "fft_golden.npz"))
    for n in (7, 12, 30, 96, 105):
        v = g[f"c1d_{n}"]
        buf = np.empty(2 * n)
        buf[0::2], buf[1::2] = v.real, v.imag
        orc.lib().orc_fft_1d(n, buf.ctypes.data, -1)
        assert np.abs(buf[0::2] + 1j * buf[1::2] - g[f"fft1d_{n}"]).max() < 1e-12 * np.abs(g[f"fft1d_{n}"]).max()
    for n in (6, 10, 16):
        pad = np.zeros(orc.padded_real_shape(n), np.float32)
        pad[:, :, :n] = g[f"real_{n}"]
        orc.lib().orc_fft_r2c_3d(n, pad.ctypes.data)
        ref = g[f"rfftn_{n}"]
        assert np.abs(pad.view(np.complex64) - ref).max() <= 1e-6 * np.abs(ref).max()
        buf = np.ascontiguousarray(g[f"half_{n}"].view(np.float32).reshape(orc.padded_real_shape(n)))
        orc.lib().orc_fft_c2r_3d(n, buf.ctypes.data)
        ref = g[f"irfftn_{n}"]
        assert np.abs(buf[:, :, :n] - ref).max() <= 1e-6 * np.abs(ref).max()


def test_oracle_fft_brute_force_dft():
    n = 6
    x = np.random.default_rng(1).standard_normal(n) + 1j * np.random.default_rng(2).standard_normal(n)
    buf = np.empty(2 * n)
    buf[0::2], buf[1::2] = x.real, x.imag
    orc.lib().orc_fft_1d(n, buf.ctypes.data, +1)
    k = np.arange(n)
    dft = np.exp(2j * np.pi * np.outer(k, k) / n) @ x
    assert np.allclose(buf[0::2] + 1j * buf[1::2], dft, atol=1e-12)


# ------------------------------------------------------------------ cosmology scalars
def test_hyp2f1_oracle_and_library_vs_mpmath():
    for case in _json("cosmo_golden.json")["hyp2f1"]:
        v = orc.lib().orc_hyp2f1(case["a"], case["b"], case["c"], case["z"])
        assert v == pytest.approx(case["value"], rel=1e-13)


def test_growth_factor_cola_vs_exact():
    """cola.c growthD against the exact flat-LCDM growth factor.  Above
    a = 0.2 the closed form (2F1 or its 14-term series for x > 30,
    cola.c:184-185) is good to ~4e-6; the reference's a <= 0.2 series is
    accurate to ~1e-3 only (cola.c:193), which the oracle reproduces."""
    for case in _json("cosmo_golden.json")["growth"]:
        a = case["a"]
        for d in (orc.lib().orc_growthD(a, OM32), ch.growth_scalars(OM32, a)[0]):
            tol = 1e-5 if np.float32(a) > 0.2 else 1e-3
            assert d == pytest.approx(case["D"], rel=tol), a


def test_dprimeq_is_q_dd_da():
    for case in _json("cosmo_golden.json")["growth"]:
        a = case["a"]
        if np.float32(a) <= 0.2 or a > 0.99:
            continue
        v = orc.lib().orc_DprimeQ(a, 1.0, OM32)
        assert v == pytest.approx(case["QdDda"], rel=2e-5), a
        assert ch.growth_scalars(OM32, a)[4] == pytest.approx(case["QdDda"], rel=2e-5)


def test_growth_d2v_approximates_derivative():
    """growthD2v = Q (D2/a) 2 Omega^(6/11) approximates Q dD2/da (cola.c:219-223);
    check it against the numerical derivative of growthD2 to 2%."""
    for a in (0.3, 0.5, 0.8):
        h = 1e-3
        d2p = orc.lib().orc_growthD2(a + h, OM32)
        d2m = orc.lib().orc_growthD2(a - h, OM32)
        q = orc.lib().orc_Qfactor(a, OM32)
        assert orc.lib().orc_growthD2v(a, OM32) == pytest.approx(q * (d2p - d2m) / (2 * h), rel=2e-2)


def test_time_integrals_vs_mpmath():
    for case in _json("cosmo_golden.json")["time_integrals"]:
        sq = orc.lib().orc_Sq(case["ai"], case["af"], case["aref"], OM32)
        # the reference integrates a float Qfactor: ~1e-7 relative noise
        assert sq == pytest.approx(case["Sq"], rel=2e-7)
        sphi = orc.lib().orc_Sphi(case["ai"], case["af"], case["aref"], OM32)
        assert sphi == pytest.approx(case["Sphi"], rel=2e-7)


def test_power_c_growth_factor_vs_mpmath():
    for case in _json("cosmo_golden.json")["growth_factor_power_c"]:
        g = orc.lib().orc_growth_factor(case["astart"], case["aend"], OM32, 1 - OM32)
        assert g == pytest.approx(case["G"], rel=1e-9)


def test_host_scalars_equal_oracle_bit_for_bit():
    """The library's host scalars (its own QAG, 2F1 by Pfaff's transform) and
    the oracle's (direct series) give the same float factors."""
    L = orc.lib()
    for a in np.linspace(0.05, 1.0, 40).astype(np.float32):
        a = float(a)
        g = ch.growth_scalars(OM32, a)
        assert np.float32(g[0]) == np.float32(L.orc_growthD(a, OM32))
        assert np.float32(g[1]) == np.float32(L.orc_growthD2(a, OM32))
        assert np.float32(g[2]) == np.float32(L.orc_growthD2v(a, OM32))
        assert np.float32(g[3]) == np.float32(L.orc_Qfactor(a, OM32))
    for istep in range(1, 10):
        da = 0.1
        f = ch.step_factors(OM32, istep * da, (istep - 0.5) * da if istep > 1 else da, (istep + 0.5) * da,
                            (istep + 1) * da)
        ax = np.float32(istep * da)
        av = np.float32((istep - 0.5) * da if istep > 1 else da)
        dda = np.float32(L.orc_Sphi(float(av), float(np.float32((istep + 0.5) * da)), float(ax), OM32))
        assert np.float32(f[0]) == dda
        dyyy = np.float32(L.orc_Sq(float(ax), float(np.float32((istep + 1) * da)),
                                   float(np.float32((istep + 0.5) * da)), OM32))
        assert np.float32(f[3]) == dyyy


# ------------------------------------------------------------------ power spectrum
def test_sigma8_of_reference_table():
    """power.c's sigma8 integral on the reference's own input table
    (data/linear_matterpower.dat) against scipy; the library and the oracle
    agree with it and with the expected 0.8103."""
    gold = _json("sigma8_golden.json")["sigma8_scipy"]
    table = os.path.join(GOLD, "linear_matterpower.dat")
    assert orc.lib().orc_power_read_file(table.encode()) == 0
    s = C.c_double()
    orc.lib().orc_power_check_sigma8(0.81, C.byref(s))
    assert s.value == pytest.approx(gold, rel=1e-4)
    assert gold == pytest.approx(0.8103, abs=2e-4)
    k, p = np.loadtxt(table, unpack=True)
    assert ch.power_file_sigma8(table) == pytest.approx(gold, rel=1e-4)
    assert ch.power_table_sigma8(k, p) == pytest.approx(gold, rel=1e-4)
    # power.c:120-151 log-log interpolation, 0 outside the table
    kq = np.array([2e-4, 0.0123, 0.5, 3.0, 100.0])
    lk, lD = np.log10(k), np.log10(k ** 3 * p / (2 * np.pi ** 2))
    ref = 10 ** np.interp(np.log10(kq), lk, lD) / (4 * np.pi * kq ** 3)
    assert np.allclose(ch.power_table_eval(k, p, kq), ref, rtol=1e-12)
    assert np.allclose([orc.lib().orc_powerspec(x) for x in kq], ref, rtol=1e-12)
    assert np.all(ch.power_table_eval(k, p, [1e-5, 1e5]) == 0.0)


def test_power_file_error_code():
    with pytest.raises(ch.ColaError) as e:
        ch.power_file_sigma8("/nonexistent/pk.dat")
    assert e.value.code == 3000


def test_eh98_table():
    """Eisenstein & Hu no-wiggle table: sigma8 = 0.8 by an independent scipy
    integration, T(k -> 0) -> 1, P ~ k^ns at small k and ~ k^(ns-4) ln^2 k
    at large k."""
    from scipy import integrate

    k, p = ch.eh98_table()
    assert len(k) == 500 and k[0] == pytest.approx(1e-4) and k[-1] == pytest.approx(100.0)
    lk, lp = np.log(k), np.log(p)

    def integrand(x):
        kk = np.exp(x)
        kr = 8 * kk
        w = 3 * (np.sin(kr) / kr ** 3 - np.cos(kr) / kr ** 2)
        return kk ** 3 * w * w * np.exp(np.interp(x, lk, lp)) / (2 * np.pi ** 2)

    s8 = np.sqrt(integrate.quad(integrand, np.log(1e-4), np.log(500 / 8), limit=500, epsrel=1e-6)[0])
    assert s8 == pytest.approx(0.8, rel=2e-3)
    slope = (lp[10] - lp[0]) / (lk[10] - lk[0])
    assert slope == pytest.approx(0.96, abs=2e-3)


# ------------------------------------------------------------------ PM pieces (closed forms)
def test_deposit_single_particle_weights():
    """A particle at (i+dx, j+dy, k+dz) cell units deposits WPAR*(1-dx or dx)...
    into the 8 surrounding cells (pm.c:229-276); at a cell corner it puts all
    its mass into one cell."""
    ng, box = 8, 8.0
    p = np.zeros(1, orc.PARTICLE_DTYPE)
    p["x"][0] = (2.25, 3.5, 4.0)
    d = np.zeros(orc.padded_real_shape(ng), np.float32)
    orc.lib().orc_pm_deposit(ng, box, 27.0, p.ctypes.data, 1, d.ctypes.data)
    dd = d[:, :, :ng] + 1.0
    expect = np.zeros((ng, ng, ng))
    for ix, wx in ((2, 0.75), (3, 0.25)):
        for iy, wy in ((3, 0.5), (4, 0.5)):
            expect[ix, iy, 4] += 27.0 * wx * wy
    assert np.allclose(dd, expect, atol=1e-6)


def test_deposit_mass_conservation_and_ghosts():
    """Total mass WPAR * Np (pm.c:535-562), including particles in the last x
    layer deposited through ghost copies and a particle exactly at x = L."""
    ng, box, nc = 12, 30.0, 4
    rng = np.random.default_rng(0)
    p = np.zeros(nc ** 3, orc.PARTICLE_DTYPE)
    p["x"] = rng.uniform(0, box, (nc ** 3, 3)).astype(np.float32)
    p["x"][0, 0] = np.float32(box - 0.5 * box / ng)
    p["x"][1, 0] = np.nextafter(np.float32(box), np.float32(0))
    d = np.zeros(orc.padded_real_shape(ng), np.float32)
    orc.lib().orc_pm_deposit(ng, box, 27.0, p.ctypes.data, len(p), d.ctypes.data)
    total = d[:, :, :ng].astype(np.float64).sum()
    assert abs(total) < 1e-3
    # the last-layer particle's ghost puts its D1 share into x layer 0
    assert d[0, :, :ng].sum() + ng * ng > 0


def test_plane_wave_poisson_closed_form():
    """delta = A cos(2 pi n.x/L) on the grid nodes -> the PM force grid
    (L/2pi) A n_a/|n|^2 sin(2 pi n.x/L) (pm.c:304-355)."""
    ng, box = 16, 40.0
    n = np.array([1, -2, 3])
    idx = np.indices((ng, ng, ng)).astype(np.float64)
    theta = 2 * np.pi * (n[0] * idx[0] + n[1] * idx[1] + n[2] * idx[2]) / ng
    # place particles so that the CIC density is exactly the sampled field is not
    # possible; drive the force stage directly with the density instead
    dens = np.zeros(orc.padded_real_shape(ng), np.float32)
    dens[:, :, :ng] = 0.5 * np.cos(theta)
    orc.lib().orc_fft_r2c_3d(ng, dens.ctypes.data)
    spec = dens.view(np.complex64)
    w = np.where(np.arange(ng) <= ng // 2, np.arange(ng), np.arange(ng) - ng)
    I0, J0, K = np.meshgrid(w, w, np.arange(ng // 2 + 1), indexing="ij")
    scale = np.float32(2 * np.pi / np.float32(box))
    f1 = np.float32(-1.0 / ng ** 3 / np.float64(scale))
    k2 = (K * K + I0 * I0 + J0 * J0).astype(np.float32)
    k2[0, 0, 0] = 1
    for a, dcomp in enumerate((I0, J0, K)):
        f2 = (f1 / k2 * dcomp.astype(np.float32)).astype(np.float32)
        fk = np.ascontiguousarray((-f2 * spec.imag + 1j * f2 * spec.real).astype(np.complex64))
        fk[0, 0, 0] = 0
        buf = np.ascontiguousarray(fk.view(np.float32).reshape(orc.padded_real_shape(ng)))
        orc.lib().orc_fft_c2r_3d(ng, buf.ctypes.data)
        expect = 0.5 * box / (2 * np.pi) * n[a] / (n @ n) * np.sin(theta)
        assert np.abs(buf[:, :, :ng] - expect).max() < 1e-5


def test_force_of_uniform_lattice_vanishes():
    """A perfect particle lattice (one particle per 3^3 cells, at cell corners)
    has a uniform CIC density and therefore zero PM force."""
    nc, ng, box = 8, 24, 48.0
    q = (np.indices((nc, nc, nc)).reshape(3, -1).T * (box / nc)).astype(np.float32)
    p = np.zeros(nc ** 3, orc.PARTICLE_DTYPE)
    p["x"] = q
    out = orc.pm_forces(ng, box, 27.0, p)
    assert out["status"] == 0
    assert np.abs(out["force"]).max() < 1e-5
    assert abs(out["mass_sum"]) < 1e-3


# ------------------------------------------------------------------ 2LPT invariants
def test_lpt_hermitian_k0_plane_and_nyquist():
    """The ZA displacement spectra are conjugate-symmetric on the kz = 0 plane
    (lpt.c:289-347) and zero on the Nyquist planes."""
    k, p = ch.eh98_table()
    orc.set_power_table(k, p)
    n = 16
    parts, cdisp, disp, disp2 = orc.lpt(n, 0.1, 0.273, 5, 64.0, want_grids=True)
    c = cdisp.view(np.complex64).reshape(3, n, n, n // 2 + 1)
    for a in range(3):
        plane = c[a, :, :, 0]
        ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        mirror = plane[(-ii) % n, (-jj) % n]
        mask = (ii != n // 2) & (jj != n // 2)
        assert np.allclose(plane[mask], np.conj(mirror[mask]), atol=0)
        assert np.all(c[a, n // 2] == 0) and np.all(c[a, :, n // 2] == 0) and np.all(c[a, :, :, n // 2] == 0)
    assert np.all(c[:, 0, 0, 0] == 0)
    # displacement fields are real transforms: dx1 = disp, lattice positions
    assert np.array_equal(parts["id"], np.arange(1, n ** 3 + 1))


def test_lpt_mode_amplitudes_follow_pk():
    """<|delta_k|^2> of the white-noise modes equals (2pi/L)^3 P(k): the
    exponential amplitude draw has unit mean (lpt.c:271-275)."""
    k, p = ch.eh98_table()
    orc.set_power_table(k, p)
    n, box = 32, 200.0
    _, cdisp, _, _ = orc.lpt(n, 0.1, 0.273, 11, box, want_grids=True)
    c = cdisp.view(np.complex64).reshape(3, n, n, n // 2 + 1)
    w = np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n)
    kx, ky, kz = np.meshgrid(w, w, np.arange(n // 2 + 1), indexing="ij")
    kv = np.stack([kx, ky, kz]) * 2 * np.pi / box
    kk = np.sqrt((kv ** 2).sum(0))
    sel = (kz > 0) & (kz < n // 2) & (np.abs(kx) < n // 2) & (np.abs(ky) < n // 2) & (kk > 0)
    # |cdisp_a|^2 summed over a = |delta|^2 / k^2
    d2 = (np.abs(c) ** 2).sum(0)[sel] * kk[sel] ** 2
    pk = np.array([orc.lib().orc_powerspec(x) for x in kk[sel]])
    ratio = d2 / (pk * (2 * np.pi / box) ** 3)
    assert ratio.mean() == pytest.approx(1.0, abs=0.03)


def test_lpt_growth_of_positions():
    """Particle displacement at a_init is D(a_init) dx1 + D2 dx2 with D from
    power.c's growth integral (lpt.c:509-545)."""
    k, p = ch.eh98_table()
    orc.set_power_table(k, p)
    n, box = 16, 64.0
    parts = orc.lpt(n, 0.1, 0.273, 3, box)
    q = ((np.indices((n, n, n)).reshape(3, -1).T + np.float32(0.5)) * np.float32(box / n)).astype(np.float32)
    dplus = np.float32(1.0 / orc.lib().orc_growth_factor(0.1, 1.0, 0.273, 0.727))
    omega = 0.273 / (0.273 + 0.727 * 0.1 ** 3)
    d2 = np.float32(dplus * dplus) * (omega / 0.273) ** (-1 / 143)
    pred = parts["dx1"] * dplus + parts["dx2"] * d2
    assert np.abs(parts["x"] - q - pred).max() < 1e-5 * box
    assert dplus == pytest.approx(0.13117, rel=1e-4)
    # the second-order field is -3/7 of the 2LPT displacement (lpt.c:547)
    assert np.abs(parts["dx2"]).max() < np.abs(parts["dx1"]).max()


# ------------------------------------------------------------------ kick / drift
def test_kick_drift_keep_lpt_trajectory():
    """With the force equal to the 2LPT force, the COLA kick leaves v = 0 and
    the drift moves the particle by dx1 dD + dx2 dD2 (cola.c:104-149)."""
    n = 64
    rng = np.random.default_rng(4)
    p = np.zeros(n, orc.PARTICLE_DTYPE)
    p["dx1"] = rng.normal(0, 1, (n, 3))
    p["dx2"] = rng.normal(0, 0.1, (n, 3))
    ax, av = C.c_float(0.3), C.c_float(0.35)
    f = ch.step_factors(OM32, 0.3, 0.35, 0.45, 0.5)
    q1, q2 = f[1], f[2]
    force = (-(p["dx1"] * q1 + p["dx2"] * q2) / (1.5 * OM32)).astype(np.float32)
    x0 = p["x"].copy()
    orc.lib().orc_kick(p.ctypes.data, np.ascontiguousarray(force).ctypes.data, n, OM32, C.byref(ax), C.byref(av), 0.45)
    assert np.abs(p["v"]).max() < 1e-6
    orc.lib().orc_drift(p.ctypes.data, n, OM32, C.byref(ax), C.byref(av), 0.5)
    gold = {c["a"]: c["D"] for c in _json("cosmo_golden.json")["growth"]}
    dD = gold[0.5] - gold[0.3]
    d2 = orc.lib().orc_growthD2(0.5, OM32) - orc.lib().orc_growthD2(0.3, OM32)
    expect = x0 + p["dx1"] * dD + p["dx2"] * d2
    assert np.abs(p["x"] - expect).max() < 1e-5
    assert (ax.value, av.value) == (np.float32(0.5), np.float32(0.45))


# ------------------------------------------------------------------ CIC readout (pm.c:359-412)
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_readout_interpolates_a_linear_ramp_exactly(axis):
    """A field linear in the cell index along one axis, g = 1 + 0.5 i_axis,
    is read out as 1 + 0.5 X_axis (X = x Ng / L) wherever the stencil does
    not cross the periodic seam of that axis; along the other axes the field
    is constant, so particles in the last x layer (read through their ghost
    copies, pm.c:416-532) get the same exact value.  Swapped D/T weights or a
    wrong stencil index fail this."""
    ng, box = 12, 30.0
    rng = np.random.default_rng(10 + axis)
    n = 400
    p = np.zeros(n, orc.PARTICLE_DTYPE)
    p["x"] = rng.uniform(0, box, (n, 3)).astype(np.float32)
    # stay off the seam of the ramp axis: cell coordinate in [0, ng - 1)
    p["x"][:, axis] = rng.uniform(0, box * (ng - 1) / ng * 0.999, n).astype(np.float32)
    if axis != 0:
        p["x"][:40, 0] = rng.uniform(box - box / ng, box, 40).astype(np.float32)  # ghost layer
    idx = np.indices((ng, ng, ng))[axis].astype(np.float32)
    out = orc.pm_readout(ng, box, p, 1.0 + 0.5 * idx)
    X = p["x"][:, axis].astype(np.float64) * ng / box
    assert np.abs(out - (1.0 + 0.5 * X)).max() < 1e-5 * (1.0 + 0.5 * ng)


def test_readout_is_adjoint_of_deposit():
    """The readout and the deposit use the same CIC weights (pm.c:229-276 and
    pm.c:384-410, ghost copies included): for any grid g,
    sum_p readout_p(g) = sum_c g_c (delta_c + 1) / WPAR."""
    ng, box, wpar = 10, 20.0, 8.0
    rng = np.random.default_rng(3)
    n = 500
    p = np.zeros(n, orc.PARTICLE_DTYPE)
    p["x"] = rng.uniform(0, box, (n, 3)).astype(np.float32)
    p["x"][:50, 0] = rng.uniform(box - box / ng, box, 50).astype(np.float32)
    g = rng.standard_normal((ng, ng, ng)).astype(np.float32)
    lhs = orc.pm_readout(ng, box, p, g).astype(np.float64).sum()
    d = np.zeros(orc.padded_real_shape(ng), np.float32)
    orc.lib().orc_pm_deposit(ng, box, wpar, p.ctypes.data, n, d.ctypes.data)
    w = (d[..., :ng].astype(np.float64) + 1.0) / wpar
    rhs = (g.astype(np.float64) * w).sum()
    assert lhs == pytest.approx(rhs, abs=1e-4 * np.abs(g).max() * n ** 0.5)
    # a single particle: its 8 weights sum to one
    one = orc.pm_readout(ng, box, p[:1], np.ones((ng, ng, ng), np.float32))
    assert one[0] == pytest.approx(1.0, abs=1e-6)


# ------------------------------------------------------------------ 2LPT stages (lpt.c:358-503)
def test_lpt_two_mode_second_order_closed_form():
    """ZA displacement psi = A k1 sin(k1.x + al) + B k2 sin(k2.x + be)
    (curl-free, two non-parallel modes with off-diagonal gradients and
    phases, so the spectra have real and imaginary parts).  The 2LPT source
    of lpt.c:409-422 (sum of the principal 2x2 minors of d_ij = d_j psi_i)
    is A B |k1 x k2|^2 cos(k1.x + al) cos(k2.x + be), and the 2LPT
    displacement of lpt.c:433-503 (div psi2 = source, curl-free) is
    (A B |k1 x k2|^2 / 2) sum_{K = k1 +- k2} K sin(K.x + al +- be) / |K|^2."""
    n, box = 16, 50.0
    kf = 2 * np.pi / box
    n1, n2 = np.array([1, 2, 0]), np.array([0, 1, -1])
    k1, k2 = kf * n1, kf * n2
    A, B, al, be = 2.0, 3.0, 0.3, 1.1
    x = np.indices((n, n, n)).astype(np.float64) * (box / n)
    th1 = np.tensordot(k1, x, 1) + al
    th2 = np.tensordot(k2, x, 1) + be
    psi = A * k1[:, None, None, None] * np.sin(th1) + B * k2[:, None, None, None] * np.sin(th2)
    cdisp = np.fft.rfftn(psi, axes=(1, 2, 3)) / n ** 3  # lpt.c's c2r is unnormalised
    disp, disp2 = orc.lpt_from_modes(n, box, cdisp)
    assert np.abs(disp - psi).max() < 1e-5 * np.abs(psi).max()
    c2 = np.sum(np.cross(k1, k2) ** 2)
    expect = np.zeros_like(psi)
    for K, ph in ((k1 + k2, al + be), (k1 - k2, al - be)):
        expect += (A * B * c2 / 2) * K[:, None, None, None] * np.sin(np.tensordot(K, x, 1) + ph) / (K @ K)
    scale = np.abs(expect).max()
    assert scale > 0
    assert np.abs(disp2 - expect).max() < 2e-5 * scale


def test_lpt_single_mode_has_no_second_order():
    """A single plane wave is one-dimensional: every 2x2 minor of d_ij
    vanishes, so the 2LPT displacement is zero (to float rounding)."""
    n, box = 16, 40.0
    k = 2 * np.pi / box * np.array([1, -2, 3])
    x = np.indices((n, n, n)).astype(np.float64) * (box / n)
    psi = 1.5 * k[:, None, None, None] * np.sin(np.tensordot(k, x, 1))
    cdisp = np.fft.rfftn(psi, axes=(1, 2, 3)) / n ** 3
    _, disp2 = orc.lpt_from_modes(n, box, cdisp)
    assert np.abs(disp2).max() < 1e-6 * np.abs(psi).max() * np.abs(k).max()


# ------------------------------------------------------------------ white-noise modes (lpt.c:172-351)
def _draw_pairs(seed, count):
    """(phase, amplitude) uniform pairs of one column's RANLUX stream with
    the redraw of a zero amplitude (lpt.c:228-233)."""
    r = orc.Ranlxd()
    orc.lib().orc_ranlxd_set(C.byref(r), int(seed), 202)
    out = []
    for _ in range(count):
        u1 = orc.lib().orc_ranlxd_uniform(C.byref(r))
        u2 = orc.lib().orc_ranlxd_uniform(C.byref(r))
        while u2 == 0.0:
            u2 = orc.lib().orc_ranlxd_uniform(C.byref(r))
        out.append((u1, u2))
    return out


@pytest.mark.parametrize("n,box,seed", [(8, 40.0, 100), (16, 64.0, 7)])
def test_lpt_modes_divergence_and_curl(n, box, seed):
    """The ZA displacement modes are curl-free (k x psi_k = 0) and their
    divergence is the white-noise density: i k . psi_k = -delta_k with
    delta_k = (2pi/L)^1.5 sqrt(P(k) (-ln u2)) exp(2 pi i u1), (u1, u2) the
    draw pair of mode kz of column (i, j) seeded from the seed table
    (lpt.c:219-284).  Checked on every mode the reference sets from its own
    column's draws: 0 < kz < N/2, and on the kz = 0 plane the columns it owns
    (the mirrored half is pinned by the Hermitian test)."""
    k, p = ch.eh98_table()
    orc.set_power_table(k, p)
    _, cdisp, _, _ = orc.lpt(n, 0.1, 0.273, seed, box, want_grids=True)
    c = cdisp.view(np.complex64).reshape(3, n, n, n // 2 + 1).astype(np.complex128)
    table = orc.seedtable(n, seed)
    fac = (2 * np.pi / box) ** 1.5
    w = np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n) * 2 * np.pi / box
    checked = 0
    for i in range(n):
        for j in range(n):
            draws = _draw_pairs(table[i, j], n // 2)
            for kz in range(n // 2):
                if i == n // 2 or j == n // 2 or (i == 0 and j == 0 and kz == 0):
                    continue
                if kz == 0 and not (i < n // 2 and (i > 0 or j < n // 2)):
                    continue
                kv = np.array([w[i], w[j], kz * 2 * np.pi / box])
                psi = c[:, i, j, kz]
                kk = np.sqrt(kv @ kv)
                u1, u2 = draws[kz]
                delta = fac * np.sqrt(orc.lib().orc_powerspec(kk) * -np.log(u2)) * np.exp(2j * np.pi * u1)
                assert 1j * (kv @ psi) == pytest.approx(-delta, rel=2e-6, abs=0), (i, j, kz)
                curl = np.cross(kv, psi)
                assert np.abs(curl).max() <= 2e-6 * kk * np.abs(psi).max(), (i, j, kz)
                checked += 1
    assert checked > n * n * (n // 2 - 1) * 0.5


@pytest.mark.parametrize("n1,n2", [(8, 16), (16, 64)])
def test_seedtable_shells_are_resolution_independent(n1, n2):
    """lpt.c:172-207 fills the table in shells of growing max(|i|, |j|) (the
    corner blocks i, N-1-i), so the low-|k| columns keep their seeds when the
    mesh grows: the n1 table is the n2 table's corner blocks; and every cell
    of the table is written exactly once from the first N^2 draws."""
    t1 = orc.seedtable(n1, 5)
    t2 = orc.seedtable(n2, 5)
    h = n1 // 2
    idx = np.r_[np.arange(h), np.arange(n2 - h, n2)]
    assert np.array_equal(t2[np.ix_(idx, idx)], t1)
    u = ch.ranlxd1_uniforms(5, n2 * n2)
    draws = np.sort((0x7fffffff * u).astype(np.uint32))
    assert np.array_equal(np.sort(t2.ravel()), draws)


# ------------------------------------------------------------------ snapshot velocities (cola.c:297-416)
def _growth_gold():
    return {c["a"]: c for c in _json("cosmo_golden.json")["growth"]}


def test_noncola_initial_velocity_closed_form():
    """set_noncola_initial (cola.c:297-338): v = 100/a (dx1 dD/dy + dx2
    dD2/dy) km/s with dD/dy = Q dD/da from the exact growth integral
    (mpmath fixture) and dD2/dy = growthD2v; positions pass through."""
    gold = _growth_gold()
    rng = np.random.default_rng(8)
    n = 50
    p = np.zeros(n, orc.PARTICLE_DTYPE)
    p["x"] = rng.uniform(0, 10, (n, 3))
    p["dx1"] = rng.normal(0, 1, (n, 3))
    p["dx2"] = rng.normal(0, 0.2, (n, 3))
    for a in (0.3, 0.5):
        out = np.zeros(n, orc.PARTICLE_MIN_DTYPE)
        orc.lib().orc_set_noncola_initial(p.ctypes.data, n, OM32, a, out.ctypes.data)
        dv2 = orc.lib().orc_growthD2v(a, OM32)
        want = 100.0 / a * (p["dx1"] * gold[a]["QdDda"] + p["dx2"] * dv2)
        assert np.abs(out["v"] - want).max() < 3e-5 * np.abs(want).max()
        assert np.array_equal(out["x"], p["x"])


def test_snapshot_on_the_lpt_trajectory_closed_form():
    """cola_set_snapshot (cola.c:341-416) with zero COLA velocity and the
    force equal to the 2LPT force (so the kick adds nothing): the snapshot
    sits on the 2LPT trajectory, x(aout) = x(a) + dx1 (D(aout) - D(a)) +
    dx2 (D2(aout) - D2(a)), with the 2LPT velocity 100/aout (dx1 Q dD/da +
    dx2 growthD2v) km/s; D and Q dD/da from the exact growth integral."""
    gold = _growth_gold()
    rng = np.random.default_rng(9)
    n = 50
    p = np.zeros(n, orc.PARTICLE_DTYPE)
    p["x"] = rng.uniform(0, 10, (n, 3))
    p["dx1"] = rng.normal(0, 1, (n, 3))
    p["dx2"] = rng.normal(0, 0.2, (n, 3))
    a_x, a_v, aout = 0.3, 0.35, 0.5
    f = ch.step_factors(OM32, a_x, a_v, 0.45, 0.5)
    force = np.ascontiguousarray((-(p["dx1"] * f[1] + p["dx2"] * f[2]) / (1.5 * OM32)).astype(np.float32))
    out = np.zeros(n, orc.PARTICLE_MIN_DTYPE)
    orc.lib().orc_set_snapshot(aout, p.ctypes.data, force.ctypes.data, n, OM32, a_x, a_v, out.ctypes.data)
    d2 = orc.lib().orc_growthD2(aout, OM32) - orc.lib().orc_growthD2(a_x, OM32)
    want_x = p["x"] + p["dx1"] * (gold[aout]["D"] - gold[a_x]["D"]) + p["dx2"] * d2
    assert np.abs(out["x"] - want_x).max() < 2e-5 * 10
    want_v = 100.0 / aout * (p["dx1"] * gold[aout]["QdDda"] + p["dx2"] * orc.lib().orc_growthD2v(aout, OM32))
    assert np.abs(out["v"] - want_v).max() < 5e-5 * np.abs(want_v).max()
    assert np.array_equal(out["id"], p["id"])


def test_snapshot_kick_and_drift_of_the_cola_velocity():
    """With no LPT displacement, the snapshot is the COLA velocity kicked by
    the force over Sphi(a_v, aout; a) and drifted over Sq(a, aout; a_v)
    (cola.c:376-410), velocity in km/s: v_out = 100/aout (v - 1.5 Om f Sphi),
    x_out = x + v Sq, with Sq, Sphi from the mpmath fixtures."""
    case = [c for c in _json("cosmo_golden.json")["time_integrals"] if c["ai"] == 0.15][0]
    ai, af, aref = case["ai"], case["af"], case["aref"]
    rng = np.random.default_rng(12)
    n = 40
    p = np.zeros(n, orc.PARTICLE_DTYPE)
    p["x"] = rng.uniform(0, 10, (n, 3))
    p["v"] = rng.normal(0, 1, (n, 3))
    force = np.ascontiguousarray(rng.normal(0, 1, (n, 3)).astype(np.float32))
    out = np.zeros(n, orc.PARTICLE_MIN_DTYPE)
    # the kick integrates Sphi(AI = a_v, AF = aout; A = a_x)
    orc.lib().orc_set_snapshot(af, p.ctypes.data, force.ctypes.data, n, OM32, aref, ai, out.ctypes.data)
    want_v = 100.0 / af * (p["v"] - 1.5 * OM32 * force * case["Sphi"])
    assert np.abs(out["v"] - want_v).max() < 3e-5 * np.abs(want_v).max()
    # the drift integrates Sq(A = a_x, AF = aout; AC = a_v)
    orc.lib().orc_set_snapshot(af, p.ctypes.data, force.ctypes.data, n, OM32, ai, aref, out.ctypes.data)
    assert np.abs(out["x"] - (p["x"] + p["v"] * case["Sq"])).max() < 2e-5 * 10
