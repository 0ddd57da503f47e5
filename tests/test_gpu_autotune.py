"""The autotuner changes launch shapes only: forces and a full run are bit for
bit the same before and after tuning, and every tuned shape is accepted."""
import numpy as np
import pytest

import cola_halo_amd as ch

pytestmark = pytest.mark.gpu

OM = 0.273


def test_autotune_keeps_results(eh_table):
    print("autotune 64^3 / 192^3", flush=True)
    k, p = eh_table
    with ch.Context(64, 192, 100.0) as ctx:
        ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        ctx.lpt_set_displacement(0.1, OM, 3)
        ctx.state_save()
        ctx.move_particles()
        ctx.pm_calculate_forces()
        f0 = ctx.get_forces()
        tuned = ctx.autotune(reps=2)
        assert tuned["ms_after"] <= tuned["ms_before"]
        assert all(t in (4, 8, 16, 32, 64) for t in tuned["tile"])
        assert tuned["fft_radices"] == ch.fft_radices(192, tuned["fft_schedule"])
        ctx.state_restore()
        ctx.move_particles()
        ctx.pm_calculate_forces()
        f1 = ctx.get_forces()
        # the tuned launch shapes with the default radix schedule: bit for bit
        ctx.set_fft_schedule(0)
        ctx.state_restore()
        ctx.move_particles()
        ctx.pm_calculate_forces()
        f2 = ctx.get_forces()
    assert np.array_equal(f0, f2)
    # another radix schedule rounds differently, by float rounding only
    assert np.abs(f1 - f0).max() <= 2e-6 * np.abs(f0).max()


@pytest.mark.parametrize("ng", [96, 192, 768])
def test_fft_radix_schedules_agree(eh_table, ng):
    """Every radix schedule of the PM transforms gives the same forces up to
    float rounding (2e-6 of the maximum); within a schedule, another tile
    width gives the same forces bit for bit."""
    print(f"radix schedules {ng}^3", flush=True)
    k, p = eh_table
    nc = ng // 3
    with ch.Context(nc, ng, 100.0 * nc / 32) as ctx:
        ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        ctx.lpt_set_displacement(0.1, OM, 9)
        ctx.move_particles()
        ctx.state_save()
        ref = None
        for sch in range(5):
            ctx.set_fft_schedule(sch)
            forces = []
            for shape in ((-1, -1, 8, -1), (2, 0, 4, 1)):
                ctx.set_fft_tiles(*shape)
                ctx.state_restore()
                ctx.pm_calculate_forces()
                forces.append(ctx.get_forces())
            assert np.array_equal(forces[0], forces[1]), sch
            if ref is None:
                ref = forces[0]
            else:
                assert np.abs(forces[0] - ref).max() <= 2e-6 * np.abs(ref).max(), sch


@pytest.mark.parametrize("cols,rows,group,xcols", [(2, 0, 4, -1), (3, 1, 8, 2), (4, 2, 64, 1), (2, 1, 8, 3)])
def test_fft_launch_shapes_agree_bit_for_bit(eh_table, cols, rows, group, xcols):
    k, p = eh_table
    out = []
    for shape in (None, (cols, rows, group, xcols)):
        with ch.Context(32, 96, 100.0) as ctx:
            ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
            if shape:
                ctx.set_fft_tiles(*shape)
            (a, snap), = ch.run_cola(ctx, OM, 1.0, 10, 5, [0.0])
            out.append(snap)
    assert np.array_equal(out[0], out[1])
