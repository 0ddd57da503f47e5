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


def _ens_members(eh_table):
    """Realizations of different sizes (grids 96 to 192, one with radix 5)
    and their own launch shapes and radix schedules."""
    shapes = [(32, 96, 100.0, 3), (40, 120, 125.0, 4), (48, 144, 150.0, 5), (64, 192, 200.0, 6), (32, 96, 100.0, 7)]

    def make(i):
        nc, ng, box, _ = shapes[i]
        c = _ctx(nc, ng, box, eh_table)
        if i == 1:
            c.set_fft_schedule(3)
        if i == 2:
            c.set_tiling(16, 16, 16, 2)
        if i == 3:
            c.set_fft_tiles(4, 2, 4, 2)
        if i == 4:
            c.set_fft_schedule(1)
        return c
    return shapes, make


@pytest.mark.parametrize("graph,merge", [(False, True), (True, True), (False, False), (True, False)])
def test_ensemble_equals_single_runs(eh_table, graph, merge):
    """One launch per phase over all realizations (merge) or per phase and
    FFT kernel variant gives each realization's own run bit for bit."""
    print(f"ensemble graph={graph} merge={merge}", flush=True)
    shapes, make = _ens_members(eh_table)
    ctxs = [make(i) for i in range(len(shapes))]
    try:
        with ch.Ensemble(ctxs, graph=graph, merge=merge) as ens:
            snaps = ch.run_cola_batch(ctxs, OM, 1.0, 10, [s[3] for s in shapes], ensemble=ens)
            if merge:
                assert ens.launches() == 14
            else:
                assert ens.launches() >= 14
    finally:
        for c in ctxs:
            c.close()
    for i, ((nc, ng, box, seed), snap) in enumerate(zip(shapes, snaps)):
        with make(i) as c:
            (a, ref), = ch.run_cola(c, OM, 1.0, 10, seed, [0.0])
        assert np.array_equal(snap, ref), (nc, seed)


def test_ensemble_rebuilds_after_member_changes(eh_table):
    """Retiling a member between ensemble steps rebuilds the tables; every
    step still equals the member's own cola_step."""
    print("ensemble retile", flush=True)
    shapes = [(32, 96, 100.0, 3), (48, 144, 150.0, 4)]
    ctxs = [_ctx(nc, ng, box, eh_table) for nc, ng, box, _ in shapes]
    refs = [_ctx(nc, ng, box, eh_table) for nc, ng, box, _ in shapes]
    try:
        for c, r, s in zip(ctxs, refs, shapes):
            for x in (c, r):
                x.lpt_set_displacement(0.1, OM, s[3])
                x.set_times(x.times()[0], 0.1)
        with ch.Ensemble(ctxs) as ens:
            da = 0.1
            for istep in range(1, 7):
                avel1, apos1 = (istep + 0.5) * da, (istep + 1.0) * da
                if istep == 3:
                    ctxs[0].set_tiling(16, 16, 16, 2)
                    ctxs[1].set_fft_tiles(4, 1, 4, 3)
                    refs[0].set_tiling(16, 16, 16, 2)
                    refs[1].set_fft_tiles(4, 1, 4, 3)
                ens.step(OMF, avel1, apos1)
                for r in refs:
                    r.step(OMF, avel1, apos1)
                for c, r in zip(ctxs, refs):
                    assert np.array_equal(c.get_particles(), r.get_particles()), istep
                    assert c.times() == r.times()
                    assert c.last_mass_sum() == pytest.approx(r.last_mass_sum(), abs=1e-6)
    finally:
        for c in ctxs + refs:
            c.close()


def test_ensemble_refuses_member_with_net(eh_table):
    print("ensemble net refusal", flush=True)
    with _ctx(32, 96, 100.0, eh_table) as a:
        a.lpt_set_displacement(0.1, OM, 3)
        a.set_times(a.times()[0], 0.1)
        net = ch.ResidualNet(96)
        a.set_residual_net(net)
        with ch.Ensemble([a]) as ens:
            with pytest.raises(ch.ColaError):
                ens.step(OMF, 0.15, 0.2)
            a.set_residual_net(None)
            ens.step(OMF, 0.15, 0.2)
            assert a.times() == (np.float32(0.2), np.float32(0.15))
        net.close()


def test_ensemble_members_distinct_and_outlived(eh_table):
    """A context cannot be two members of one ensemble; destroying a member
    first leaves the ensemble refusing to step (no access to freed
    buffers), and the ensemble still closes cleanly."""
    print("ensemble member lifetime", flush=True)
    a = _ctx(32, 96, 100.0, eh_table)
    b = _ctx(32, 96, 100.0, eh_table)
    try:
        for c in (a, b):
            c.lpt_set_displacement(0.1, OM, 3)
            c.set_times(c.times()[0], 0.1)
        with pytest.raises(ch.ColaError) as err:
            ch.Ensemble([a, b, a])
        assert err.value.code == 9001
        ens = ch.Ensemble([a, b], graph=True)
        ens.step(OMF, 0.15, 0.2)
        ens.step(OMF, 0.25, 0.3)
        b.close()
        with pytest.raises(ch.ColaError) as err:
            ens.step(OMF, 0.35, 0.4)
        assert err.value.code == 9004
        ens.close()
        assert a.times() == (np.float32(0.3), np.float32(0.25))
        a.step(OMF, 0.35, 0.4)
    finally:
        a.close()
        b.close()
