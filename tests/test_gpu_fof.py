"""GPU friends-of-friends against the CPU oracle on identical snapshot records.

Bar: the wrap of the records is bit-exact, the group of every particle
(labels) is identical, halo sizes are identical, centres agree within
2e-6 L and mean velocities within 1e-5 relative (fixed-point sums on the
device, double sums in the oracle)."""
import numpy as np
import pytest

import cola_halo_amd as ch
import oracle_lib as orc

pytestmark = pytest.mark.gpu

OM = 0.273


def _say(msg):
    print(msg, flush=True)


def _clustered_ctx(nc, box, seed, table, sigma):
    k, p = table
    ctx = ch.Context(nc, 3 * nc, box)
    ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
    rng = np.random.default_rng(seed)
    n = nc ** 3
    parts = np.zeros(n, ch.PARTICLE_DTYPE)
    nblob = max(4, n // 400)
    centres = rng.uniform(0, box, (nblob, 3))
    centres[0] = [0.01 * box, 0.5 * box, 0.999 * box]  # across two periodic faces
    centres[1] = [box, box, box]
    which = rng.integers(0, nblob, n)
    clustered = rng.uniform(0, 1, n) < 0.7
    x = np.where(clustered[:, None], centres[which] + rng.normal(0, sigma, (n, 3)),
                 rng.uniform(0, box, (n, 3)))
    x[:20, 0] = -0.001 * box          # below 0: wrapped by FoF
    x[20:40, 1] = box                 # exactly L: wrapped to 0
    parts["x"] = x.astype(np.float32)
    parts["dx1"] = rng.normal(0, 1.0, (n, 3)).astype(np.float32)
    parts["dx2"] = rng.normal(0, 0.1, (n, 3)).astype(np.float32)
    parts["id"] = np.arange(1, n + 1)
    ctx.set_particles(parts, 0.5, 0.5)
    snap = ctx.set_noncola_initial(np.float32(OM))
    return ctx, snap


def _compare(ctx, snap, box, ll, nfof_min):
    halos = ctx.fof_find_halos(ll, nfof_min)
    labels = ctx.fof_labels()
    wrapped = ctx.get_snapshot()
    rec = snap.copy()
    lab_o, halos_o = orc.fof(rec, box, ll, nfof_min)
    assert np.array_equal(wrapped, rec)
    assert np.array_equal(labels, lab_o)
    assert len(halos) == len(halos_o) and len(halos) > 0
    assert np.array_equal(halos["nfof"], halos_o["nfof"])
    d = np.abs(halos["x"] - halos_o["x"])
    d = np.minimum(d, box - d)
    assert d.max() <= 2e-6 * box, d.max()
    np.testing.assert_allclose(halos["v"], halos_o["v"], rtol=1e-5, atol=1e-4)
    return halos


@pytest.mark.parametrize("nc,box,ll,nmin", [(16, 64.0, 0.4, 8), (32, 100.0, 0.3, 32), (48, 150.0, 0.25, 20),
                                           (16, 3.0, 0.9, 2)])
def test_fof_matches_oracle(eh_table, nc, box, ll, nmin):
    _say(f"fof nc={nc} box={box} ll={ll}")
    ctx, snap = _clustered_ctx(nc, box, 7 + nc, eh_table, sigma=ll)
    with ctx:
        _compare(ctx, snap, box, ll, nmin)


def test_fof_evolved_128(eh_table):
    """128^3 particles evolved to z = 0 (L = 250, 384^3 PM grid, 10 steps),
    linking length 0.2 of the mean separation (main.c:353)."""
    _say("fof evolved 128^3")
    nc, box = 128, 250.0
    k, p = eh_table
    with ch.Context(nc, 384, box) as ctx:
        ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        (a, snap), = ch.run_cola(ctx, OM, 1.0, 10, 1, [0.0])
        ll = float(np.float32(0.2 * np.float32(box) / nc))
        halos = _compare(ctx, snap, box, ll, 32)
    assert halos["nfof"].min() >= 32
    assert 100 < len(halos) < nc ** 3 // 32


def test_fof_full_size_is_deterministic(eh_table):
    """256^3, L = 500 at z = 0: two searches give identical catalogues."""
    _say("fof 256^3 determinism")
    nc, box = 256, 500.0
    k, p = eh_table
    with ch.Context(nc, 768, box) as ctx:
        ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        ch.run_cola(ctx, OM, 1.0, 10, 1, [0.0])
        ll = float(np.float32(0.2 * np.float32(box) / nc))
        h1 = ctx.fof_find_halos(ll, 32)
        l1 = ctx.fof_labels()
        h2 = ctx.fof_find_halos(ll, 32)
        l2 = ctx.fof_labels()
    assert np.array_equal(l1, l2)
    assert np.array_equal(h1["nfof"], h2["nfof"])
    assert np.abs(h1["x"] - h2["x"]).max() <= 1e-6 * box
    assert len(h1) > 10000
    # every labelled particle's group is a listed halo of the right size
    heads, counts = np.unique(l1[l1 >= 0], return_counts=True)
    assert len(heads) == len(h1)
    assert np.array_equal(np.sort(counts), np.sort(h1["nfof"]))


def test_fof_errors(eh_table):
    with ch.Context(16, 48, 64.0) as ctx:
        with pytest.raises(ch.ColaError) as e:
            ctx.fof_find_halos(0.5)
        assert e.value.code == ch.ERR_STATE
        with pytest.raises(ch.ColaError) as e:
            ctx.fof_labels()
        assert e.value.code == ch.ERR_STATE
    ctx, _ = _clustered_ctx(16, 64.0, 1, eh_table, sigma=0.4)
    with ctx:
        with pytest.raises(ch.ColaError) as e:
            ctx.fof_find_halos(40.0)  # ll >= L/2
        assert e.value.code == ch.ERR_ARG
