"""GPU parity of the output grids: the coarse density grid (coarse_grid.c) and
the shell power spectrum, against the CPU oracle on identical positions.

Tolerances:
  * coarse grid: abs(gpu - oracle) <= 1e-5 * (1 + abs(oracle)) per cell (float
    atomics sum in another order than the reference's serial loop);
  * power spectrum: identical mode counts and bin k (same double arithmetic),
    P within 1e-4 relative (float r2c on the device, double DFT in the oracle);
  * 256^3 initial conditions at a = 0.1: P(k) on k in [0.02, 0.2] h/Mpc
    within the cosmic-variance band of D(a)^2 P_lin(k).
"""
import numpy as np
import pytest

import cola_halo_amd as ch
import oracle_lib as orc

pytestmark = pytest.mark.gpu

OM = 0.273


def _say(msg):
    print(msg, flush=True)


def _ctx_with_particles(nc, ng, box, table, seed):
    k, p = table
    ctx = ch.Context(nc, ng, box)
    ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
    rng = np.random.default_rng(seed)
    n = nc ** 3
    parts = np.zeros(n, ch.PARTICLE_DTYPE)
    parts["x"] = rng.uniform(0, box, (n, 3)).astype(np.float32)
    parts["x"][:50] = np.float32(box)                     # x == L maps to cell 0
    parts["x"][50:100, 1] = -np.float32(0.01 * box)        # slightly below 0
    parts["x"][100:150] = parts["x"][100]                  # a stack of identical particles
    parts["id"] = np.arange(1, n + 1)
    ctx.set_particles(parts, 0.1, 0.1)
    return ctx, parts


@pytest.mark.parametrize("nc,n,box", [(16, 16, 64.0), (32, 24, 100.0), (24, 50, 80.0), (32, 7, 30.0)])
def test_coarse_grid_matches_oracle(eh_table, nc, n, box):
    _say(f"coarse grid nc={nc} n={n}")
    ctx, parts = _ctx_with_particles(nc, 3 * nc if nc != 24 else 48, box, eh_table, 17 + n)
    with ctx:
        g, sums = ctx.coarse_grid(n, source="particles")
    ref, rsums = orc.coarse_grid(parts["x"], n, box)
    assert np.all(np.abs(g - ref) <= 1e-5 * (1 + np.abs(ref)))
    # each device cell is its exact weight sum rounded once; the oracle
    # rounds after every add: the totals agree to float rounding
    assert sums[0] == pytest.approx(rsums[0], rel=1e-6)
    assert sums[1] == pytest.approx(rsums[1], abs=1e-6 * n ** 3)


def test_coarse_grid_of_snapshot_records(eh_table):
    _say("coarse grid of the 2LPT snapshot")
    nc, box = 32, 100.0
    k, p = eh_table
    with ch.Context(nc, 96, box) as ctx:
        ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        ctx.lpt_set_displacement(0.1, OM, 5)
        snap = ctx.set_noncola_initial(np.float32(OM))
        g, _ = ctx.coarse_grid(16, source="snapshot")
    ref, _ = orc.coarse_grid(snap, 16, box)
    assert np.all(np.abs(g - ref) <= 1e-5 * (1 + np.abs(ref)))


def test_output_grid_errors(eh_table):
    with ch.Context(16, 48, 64.0) as ctx:
        with pytest.raises(ch.ColaError) as e:
            ctx.coarse_grid(8)  # no snapshot yet
        assert e.value.code == ch.ERR_STATE
        k, p = eh_table
        ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        ctx.lpt_set_displacement(0.1, OM, 1)
        with pytest.raises(ch.ColaError) as e:
            ctx.power_spectrum(49, 0.0, 1.0, 10)  # odd grid
        assert e.value.code == ch.ERR_UNSUPPORTED
        with pytest.raises(ch.ColaError) as e:
            ctx.power_spectrum(32, 0.0, 1.0, 0)
        assert e.value.code == ch.ERR_ARG


@pytest.mark.parametrize("flags", [0, 1, 3])
@pytest.mark.parametrize("nc,n,box", [(16, 32, 64.0), (32, 48, 100.0)])
def test_power_spectrum_matches_oracle(eh_table, nc, n, box, flags):
    _say(f"power spectrum nc={nc} n={n} flags={flags}")
    k, p = eh_table
    with ch.Context(nc, 3 * nc, box) as ctx:
        ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        ctx.lpt_set_displacement(0.1, OM, 3)
        parts = ctx.get_particles()
        kf = 2 * np.pi / box
        kg, pg, cg = ctx.power_spectrum(n, 0.5 * kf, 0.5 * n * kf, 30, deconvolve=bool(flags & 1),
                                        subtract_shot=bool(flags & 2))
    delta, _ = orc.coarse_grid(parts["x"], n, box)
    kr, pr, cr = orc.power_spectrum(delta, box, 0.5 * kf, 0.5 * n * kf, 30, flags=flags, np_shot=nc ** 3)
    assert np.array_equal(cg, cr)
    np.testing.assert_allclose(kg, kr, rtol=1e-12)
    np.testing.assert_allclose(pg, pr, rtol=1e-4, atol=1e-6 * np.abs(pr).max())


def test_power_spectrum_of_initial_conditions_is_linear_theory(eh_table):
    """256^3 particles, L = 500 Mpc/h, 2LPT at a = 0.1: on large scales the
    measured spectrum is D(a)^2 P_lin(k), D = 1 / GrowthFactor(a, 1)."""
    _say("P(k) of the 256^3 initial conditions")
    nc, box, a = 256, 500.0, 0.1
    k, p = eh_table
    with ch.Context(nc, 768, box) as ctx:
        ctx.power_init_table(k, p, a, 0.8, OM, 1.0 - OM)
        ctx.lpt_set_displacement(a, OM, 1)
        d = 1.0 / ctx.growth_factor(a, 1.0)
        kb, pb, cnt = ctx.power_spectrum(512, 0.02, 0.2, 9, deconvolve=True)
        # PowerSpec (power.c:120-151) is P(k) / (2 pi)^3
        plin = np.array([ctx.power_spec(x) for x in kb]) * (2 * np.pi) ** 3
    ratio = pb / (d * d * plin)
    tol = 5 / np.sqrt(cnt / 2) + 0.03
    assert np.all(np.abs(ratio - 1) < tol), (ratio, tol)
