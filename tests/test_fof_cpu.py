"""CPU tests pinning the oracle's friends-of-friends groups (the link rule of
fof.c:586-587: squared float distance strictly below ll^2, periodic) against
an all-pairs numpy/scipy search and constructed cases, and the halo text
writer against Python's own %e formatting (fof.c:1166-1168)."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import cola_halo_amd as ch
import oracle_lib as orc


def records(pos, vel=None):
    r = np.zeros(len(pos), orc.PARTICLE_MIN_DTYPE)
    r["x"] = np.asarray(pos, np.float32)
    if vel is not None:
        r["v"] = np.asarray(vel, np.float32)
    r["id"] = np.arange(1, len(pos) + 1)
    return r


def brute_force(pos, box, ll, nfof_min):
    """All pairs in float32 with the minimum-image fold, components by scipy."""
    x = np.asarray(pos, np.float32)
    box32, half = np.float32(box), np.float32(0.5 * box)
    d = x[:, None, :] - x[None, :, :]
    d = np.where(d > half, d - box32, np.where(d < -half, d + box32, d)).astype(np.float32)
    d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]
    i, j = np.nonzero(np.triu(d2 < np.float32(ll) * np.float32(ll), 1))
    n = len(x)
    _, comp = connected_components(coo_matrix((np.ones(len(i)), (i, j)), shape=(n, n)), directed=False)
    size = np.bincount(comp)
    first = np.full(comp.max() + 1, n)
    np.minimum.at(first, comp, np.arange(n))
    return np.where(size[comp] >= nfof_min, first[comp], -1)


def wrap(pos, box):
    p = np.asarray(pos, np.float32).copy()
    box32 = np.float32(box)
    p = np.where(p < 0, p + box32, np.where(p >= box32, p - box32, p))
    return p.astype(np.float32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_groups_match_all_pairs(seed):
    rng = np.random.default_rng(seed)
    box, ll = 50.0, 1.0
    centres = rng.uniform(0, box, (12, 3))
    centres[0] = [0.2, 25.0, 49.9]  # straddles two periodic faces
    pos = np.concatenate([c + rng.normal(0, 0.8, (rng.integers(10, 300), 3)) for c in centres]
                         + [rng.uniform(0, box, (500, 3))])
    pos = pos.astype(np.float32)
    rec = records(pos, rng.normal(0, 100, pos.shape))
    label, halos = orc.fof(rec, box, ll, nfof_min=20)
    expect = brute_force(wrap(pos, box), box, ll, 20)
    assert np.array_equal(label, expect)
    heads = np.unique(expect[expect >= 0])
    assert len(halos) == len(heads)
    for h, head in zip(halos, heads):
        members = np.nonzero(expect == head)[0]
        assert h["nfof"] == len(members)
        x = wrap(pos, box)
        d = x[members] - x[head]
        d = np.where(d > 0.5 * box, d - box, np.where(d < -0.5 * box, d + box, d))
        centre = (x[head].astype(np.float64) + d.astype(np.float64).mean(axis=0)) % box
        np.testing.assert_allclose(h["x"], centre, atol=2e-5 * box)
        np.testing.assert_allclose(h["v"], rec["v"][members].astype(np.float64).mean(axis=0), rtol=1e-5, atol=1e-3)


def test_chain_across_the_boundary():
    box, ll = 10.0, 0.5
    for axis in range(3):
        xs = (np.arange(40) * 0.45 * ll - 5.0) % box  # crosses 0 / L
        pos = np.full((40, 3), 3.0, np.float32)
        pos[:, axis] = xs
        label, halos = orc.fof(records(pos), box, ll, nfof_min=2)
        assert np.all(label == 0) and len(halos) == 1 and halos[0]["nfof"] == 40
        pos = pos[:15].copy()  # 15 x 1.05 ll < L: no overlap around the torus
        pos[:, axis] = (np.arange(15) * 1.05 * ll - 5.0) % box
        label, halos = orc.fof(records(pos), box, ll, nfof_min=2)
        assert np.all(label == -1) and len(halos) == 0


def test_link_is_strict():
    box, ll = 8.0, 0.25
    pos = np.array([[1.0, 1.0, 1.0], [1.25, 1.0, 1.0]], np.float32)  # distance^2 == ll^2 exactly
    label, _ = orc.fof(records(pos), box, ll, nfof_min=2)
    assert np.all(label == -1)
    pos[1, 0] = np.nextafter(np.float32(1.25), np.float32(0))
    label, _ = orc.fof(records(pos), box, ll, nfof_min=2)
    assert np.all(label == 0)


def test_positions_are_wrapped_in_place():
    box = 10.0
    pos = np.array([[-0.5, 10.0, 3.0], [10.5, 2.0, -0.0]], np.float32)
    rec = records(pos)
    orc.fof(rec, box, 0.2, nfof_min=2)
    assert np.array_equal(rec["x"], wrap(pos, box))


def test_halo_centre_wraps_into_box():
    box, ll = 20.0, 0.6
    rng = np.random.default_rng(9)
    pos = (np.array([0.05, 10.0, 10.0]) + rng.normal(0, 0.1, (200, 3))).astype(np.float32)
    _, halos = orc.fof(records(pos), box, ll, nfof_min=32)
    assert len(halos) == 1
    x = halos[0]["x"]
    assert np.all((x >= 0) & (x < box))
    # centre near x = 0.05 (either side of the face): periodic distance small
    dx = min(abs(x[0] - 0.05), box - abs(x[0] - 0.05))
    assert dx < 0.05


def test_halo_file_format(tmp_path):
    h = np.zeros(3, ch.HALO_DTYPE)
    h["nfof"] = [32, 1000, 123456]
    h["x"] = [[0.0, 1.5, 499.99997], [1e-8, 250.0, 3.0], [12.25, 0.1, 7.0]]
    h["v"] = [[-1.5e3, 0.0, 2.5], [3.0, -4.0, 5.0e2], [1e-3, 2e-3, -3e-3]]
    path = str(tmp_path / "fof00001a.txt")
    ch.write_halos(path, h)
    lines = open(path).read().splitlines()
    expect = ["%d %e %e %e %e %e %e" % (r["nfof"], *map(float, r["x"]), *map(float, r["v"])) for r in h]
    assert lines == expect
    with pytest.raises(ch.ColaError) as e:
        ch.write_halos("/nonexistent/fof.txt", h)
    assert e.value.code == 8110
