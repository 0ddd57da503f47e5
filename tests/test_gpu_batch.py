"""Graph-captured steps and batched realizations against plain per-context
runs: bit for bit equal (the step is deterministic).  The graph test replays
the captured step many times with the program's own synchronising work in
between and checks every replay."""
import numpy as np
import pytest

import cola_halo_amd as ch

pytestmark = pytest.mark.gpu

OM = 0.273
OMF = float(np.float32(OM))


def _ctx(nc, ng, box, table):
    k, p = table
    ctx = ch.Context(nc, ng, box)
    ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
    return ctx


def test_graph_step_equals_eager_every_replay(eh_table):
    print("graph replay 64^3", flush=True)
    a = _ctx(64, 192, 100.0, eh_table)
    b = _ctx(64, 192, 100.0, eh_table)
    with a, b:
        for c in (a, b):
            c.lpt_set_displacement(0.1, OM, 11)
            c.set_times(c.times()[0], 0.1)
        b.set_step_graph(True)
        da = 0.1
        for istep in range(1, 10):
            avel1, apos1 = (istep + 0.5) * da, (istep + 1.0) * da
            a.step(OMF, avel1, apos1)
            b.step(OMF, avel1, apos1)
            pa, pb = a.get_particles(), b.get_particles()   # synchronising work between replays
            assert np.array_equal(pa, pb), istep
            assert a.times() == b.times()
            assert b.last_mass_sum() == a.last_mass_sum()
        # leave graph mode and finish eagerly: still identical
        b.set_step_graph(False)
        for c in (a, b):
            c.move_particles()
            c.pm_calculate_forces()
        assert np.array_equal(a.set_snapshot(1.0, OMF), b.set_snapshot(1.0, OMF))


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single_runs(eh_table, graph):
    print(f"batch graph={graph}", flush=True)
    shapes = [(32, 96, 100.0, 3), (48, 144, 150.0, 4), (64, 192, 200.0, 5), (32, 96, 100.0, 6)]
    ctxs = [_ctx(nc, ng, box, eh_table) for nc, ng, box, _ in shapes]
    try:
        for c in ctxs:
            c.set_step_graph(graph)
        snaps = ch.run_cola_batch(ctxs, OM, 1.0, 10, [s[3] for s in shapes])
    finally:
        for c in ctxs:
            c.close()
    for (nc, ng, box, seed), snap in zip(shapes, snaps):
        with _ctx(nc, ng, box, eh_table) as c:
            (a, ref), = ch.run_cola(c, OM, 1.0, 10, seed, [0.0])
        assert np.array_equal(snap, ref), (nc, seed)


def test_batch_snapshot_after_kick_matches_single_run(eh_table):
    """With 49 steps to a = 1, 49 * (1/49) < float32(1.0) in double, so the
    output time falls between apos0 and avel1 of the last step: main.c kicks
    to avel1 before the snapshot (main.c:195-203).  Batched runs must do the
    same as run_cola."""
    print("batch 49 steps", flush=True)
    da = 1.0 / 49
    assert 49 * da < float(np.float32(1.0)) <= 49.5 * da
    shapes = [(16, 48, 50.0, 8), (24, 72, 75.0, 9)]
    ctxs = [_ctx(nc, ng, box, eh_table) for nc, ng, box, _ in shapes]
    try:
        snaps = ch.run_cola_batch(ctxs, OM, 1.0, 49, [s[3] for s in shapes])
    finally:
        for c in ctxs:
            c.close()
    for (nc, ng, box, seed), snap in zip(shapes, snaps):
        with _ctx(nc, ng, box, eh_table) as c:
            (a, ref), = ch.run_cola(c, OM, 1.0, 49, seed, [0.0])
        assert np.array_equal(snap, ref), (nc, seed)
