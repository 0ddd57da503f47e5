"""CPU tests pinning the oracle's coarse density grid (coarse_grid.c) and its
shell power spectrum estimator against closed forms, numpy and statistics."""
import numpy as np
import pytest

import oracle_lib as orc


def numpy_cic(pos, n, box):
    """Independent periodic CIC in float64 with numpy (coarse_grid.c's rule:
    left cell floor(x n / L), weight 1 - frac to the left cell)."""
    pos = np.asarray(pos, np.float64)
    X = pos * n / box
    i0 = np.floor(X).astype(np.int64)
    w = 1.0 - (X - i0)
    grid = np.zeros((n, n, n))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                wx = w[:, 0] if dx == 0 else 1 - w[:, 0]
                wy = w[:, 1] if dy == 0 else 1 - w[:, 1]
                wz = w[:, 2] if dz == 0 else 1 - w[:, 2]
                np.add.at(grid, ((i0[:, 0] + dx) % n, (i0[:, 1] + dy) % n, (i0[:, 2] + dz) % n), wx * wy * wz)
    return grid


def numpy_power(delta, box, kmin, kmax, nbins, deconvolve=False):
    n = delta.shape[0]
    dk = np.fft.fftn(delta.astype(np.float64))
    m = np.fft.fftfreq(n, 1.0 / n)
    mx, my, mz = np.meshgrid(m, m, m, indexing="ij")
    kk = 2 * np.pi / box * np.sqrt(mx ** 2 + my ** 2 + mz ** 2)
    p = box ** 3 * np.abs(dk / n ** 3) ** 2
    if deconvolve:
        w = (np.sinc(mx / n) * np.sinc(my / n) * np.sinc(mz / n)) ** 2
        p = p / w ** 2
    t = (kk - kmin) / ((kmax - kmin) / nbins)
    sel = (kk > 0) & (t >= 0) & (t < nbins)
    b = np.floor(t[sel]).astype(int)
    cnt = np.bincount(b, minlength=nbins).astype(float)
    ks = np.bincount(b, kk[sel], minlength=nbins)
    ps = np.bincount(b, p[sel], minlength=nbins)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(cnt > 0, ks / cnt, 0), np.where(cnt > 0, ps / cnt, 0), cnt


def test_coarse_grid_single_particle_weights():
    n, box = 8, 80.0  # cell 10
    pos = np.array([[12.5, 37.0, 79.0]], np.float32)
    grid, sums = orc.coarse_grid(pos, n, box)
    # x: cell 1 frac 0.25; y: cell 3 frac 0.7; z: cell 7 frac 0.9 -> right neighbour wraps to 0
    wx = {1: 0.75, 2: 0.25}
    wy = {3: 0.3, 4: 0.7}
    wz = {7: 0.1, 0: 0.9}
    rho = np.zeros((n, n, n))
    for i, a in wx.items():
        for j, b in wy.items():
            for k, c in wz.items():
                rho[i, j, k] = a * b * c
    nbar_inv = n ** 3 / 1.0
    np.testing.assert_allclose(grid, nbar_inv * rho - 1.0, rtol=1e-5, atol=1e-3)
    assert abs(sums[0] - 1.0) < 1e-6


def test_coarse_grid_lattice_is_uniform():
    n, box = 16, 100.0
    g = np.arange(n) * (box / n)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    grid, sums = orc.coarse_grid(pos, n, box)
    assert np.all(grid == 0.0)
    assert sums == (float(n ** 3), 0.0)


def test_coarse_grid_matches_numpy_and_conserves_mass():
    rng = np.random.default_rng(5)
    n, box, npart = 24, 50.0, 20000
    pos = rng.uniform(0, box, (npart, 3)).astype(np.float32)
    pos[:10, 0] = -0.3   # slightly outside: the periodic rule still applies
    pos[10:20, 2] = box  # x = L maps to cell 0
    grid, sums = orc.coarse_grid(pos, n, box)
    rho = numpy_cic(pos, n, box)
    np.testing.assert_allclose(grid, (n ** 3 / npart) * rho - 1.0, rtol=1e-5, atol=2e-5)
    assert abs(sums[0] - npart) < 1e-2
    assert abs(sums[1]) < 1e-1


def test_coarse_grid_record_array_source():
    rng = np.random.default_rng(6)
    rec = np.zeros(500, orc.PARTICLE_MIN_DTYPE)
    rec["x"] = rng.uniform(0, 30.0, (500, 3))
    rec["v"] = 1e6  # must be ignored
    a, _ = orc.coarse_grid(rec, 6, 30.0)
    b, _ = orc.coarse_grid(np.ascontiguousarray(rec["x"]), 6, 30.0)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("m", [(1, 0, 0), (2, 3, 1), (0, 0, 5), (3, 4, 8)])
def test_power_spectrum_plane_wave(m):
    """delta = A cos(2 pi m.x / n): two modes +-m of amplitude A/2, so the bin
    holding abs(k) = 2 pi abs(m) / L has P = L^3 A^2 / 4 and 2 modes."""
    n, box, amp = 16, 200.0, 0.3
    i = np.arange(n)
    x, y, z = np.meshgrid(i, i, i, indexing="ij")
    delta = amp * np.cos(2 * np.pi * (m[0] * x + m[1] * y + m[2] * z) / n)
    kf = 2 * np.pi / box
    kabs = kf * np.sqrt(sum(c * c for c in m))
    nbins = 40
    k, p, cnt = orc.power_spectrum(delta, box, 0.0, kf * 20, nbins)
    b = int(np.floor(kabs / (kf * 20 / nbins)))
    assert cnt[b] >= 2
    expect = box ** 3 * amp ** 2 / 4 * 2 / cnt[b]  # bin mean over cnt modes, two of them non-zero
    assert p[b] == pytest.approx(expect, rel=1e-6)
    others = np.delete(p, b)
    assert np.all(np.abs(others) < 1e-9 * expect)
    # deconvolution divides the mode by W^2 = prod sinc^4(m_i / n)
    _, pd, _ = orc.power_spectrum(delta, box, 0.0, kf * 20, nbins, flags=1)
    w = np.prod([np.sinc(c / n) ** 2 for c in m])
    assert pd[b] == pytest.approx(expect / w ** 2, rel=1e-6)
    # shot-noise subtraction is a constant per occupied bin
    _, ps, _ = orc.power_spectrum(delta, box, 0.0, kf * 20, nbins, flags=2, np_shot=1000.0)
    occupied = cnt > 0
    np.testing.assert_allclose(ps[occupied], p[occupied] - box ** 3 / 1000.0, rtol=0, atol=1e-9)


def test_power_spectrum_matches_numpy():
    rng = np.random.default_rng(7)
    n, box = 20, 150.0
    delta = rng.standard_normal((n, n, n)).astype(np.float32)
    kf = 2 * np.pi / box
    for dec in (False, True):
        k, p, c = orc.power_spectrum(delta, box, 0.5 * kf, 12 * kf, 23, flags=1 if dec else 0)
        kn, pn, cn = numpy_power(delta, box, 0.5 * kf, 12 * kf, 23, deconvolve=dec)
        assert np.array_equal(c, cn)
        np.testing.assert_allclose(k, kn, rtol=1e-12)
        np.testing.assert_allclose(p, pn, rtol=1e-10)
    # every mode but k = 0 lands in some bin when the range covers the grid
    _, _, c = orc.power_spectrum(delta, box, 0.0, kf * n, 50)
    assert c.sum() == n ** 3 - 1


def test_power_spectrum_shot_noise_cic_aliasing():
    """Uniform random particles: the expected P of each mode is
    (L^3 / N) prod_i (1 - 2/3 sin^2(pi m_i / n)) (Jing 2005, eq. 20, CIC)."""
    rng = np.random.default_rng(11)
    n, box, npart = 32, 100.0, 200000
    pos = rng.uniform(0, box, (npart, 3)).astype(np.float32)
    delta, _ = orc.coarse_grid(pos, n, box)
    kf = 2 * np.pi / box
    nbins = 8
    kmax = kf * n / 2
    k, p, cnt = orc.power_spectrum(delta, box, 0.0, kmax, nbins)
    m = np.fft.fftfreq(n, 1.0 / n)
    mx, my, mz = np.meshgrid(m, m, m, indexing="ij")
    c = np.prod([1 - 2.0 / 3.0 * np.sin(np.pi * q / n) ** 2 for q in (mx, my, mz)], axis=0)
    kk = kf * np.sqrt(mx ** 2 + my ** 2 + mz ** 2)
    t = kk / (kmax / nbins)
    sel = (kk > 0) & (t < nbins)
    b = np.floor(t[sel]).astype(int)
    cexp = np.bincount(b, c[sel], minlength=nbins) / np.bincount(b, minlength=nbins)
    ratio = p / (box ** 3 / npart * cexp)
    # each mode's P is exponential: the bin mean has relative scatter 1 / sqrt(modes)
    assert np.all(np.abs(ratio - 1) < 5 / np.sqrt(cnt / 2)), ratio
