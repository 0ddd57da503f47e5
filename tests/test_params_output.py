"""The driver's parameter reader against the reference's own (read_param.c,
confirm_param.c, msg.c compiled from the reference tree; fixture
tests/golden/params_golden.json, and the live binary oracle/_ref/ref_params
when it is built), and the GADGET-1 / coarse-grid file writers against
independent numpy readers of the documented layouts."""
import json
import os
import struct

import numpy as np
import pytest

import cola_halo_amd as ch

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "params_golden.json")))["cases"]
REF_BIN = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "ref_params")


def _ours(text, confirm, tmp_path):
    path = tmp_path / "param.lua"
    path.write_bytes(text.encode())
    try:
        p = ch.read_parameters(str(path))
        if confirm:
            ch.confirm_parameters(p)
    except ch.ColaError as e:
        return e.code, None
    d = p.as_dict()
    d["n_zout"] = p.n_zout
    for name in ("power_spectrum_filename", "fof_filename", "snapshot_filename", "subsample_filename",
                 "cgrid_filename", "init_filename"):
        d["strlen_" + name] = getattr(p, "strlen_" + name)
    return 0, d


@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_parameters_match_reference_reader(case, tmp_path):
    g = GOLDEN[case]
    rc, params = _ours(g["text"], g["confirm"], tmp_path)
    assert rc == g["rc"], (case, rc, g["rc"])
    if rc == 0:
        ref = g["params"]
        for key, val in ref.items():
            assert params[key] == val, (case, key, params[key], val)
        assert set(ref) <= set(params)


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference reader not built (oracle/ref.mk)")
def test_fixture_is_current_with_reference_binary(tmp_path):
    """Re-run a few cases through the live reference build."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen", os.path.join(HERE, "golden", "gen_params_golden.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    for case in ("full_param_file", "array_blank_before_comma", "confirm_missing_power", "crlf_line_ends"):
        got = gen.run_reference(GOLDEN[case]["text"], GOLDEN[case]["confirm"])
        assert got["rc"] == GOLDEN[case]["rc"]
        assert got["params"] == GOLDEN[case]["params"]


def test_unreadable_parameter_file():
    with pytest.raises(ch.ColaError) as e:
        ch.read_parameters("/nonexistent/param.lua")
    assert e.value.code == 1001


# ---------------------------------------------------------------- GADGET-1
def read_gadget(path):
    """Independent reader of the GADGET-1 layout of write.c:14-107."""
    raw = open(path, "rb").read()
    off = 0

    def record():
        nonlocal off
        (n,) = struct.unpack_from("<i", raw, off)
        body = raw[off + 4: off + 4 + n]
        (n2,) = struct.unpack_from("<i", raw, off + 4 + n)
        assert n == n2
        off += 8 + n
        return body

    h = record()
    assert len(h) == 256
    hdr = {
        "np": struct.unpack_from("<6i", h, 0),
        "mass": struct.unpack_from("<6d", h, 24),
        "time": struct.unpack_from("<d", h, 72)[0],
        "redshift": struct.unpack_from("<d", h, 80)[0],
        "np_total": struct.unpack_from("<6I", h, 96),
        "num_files": struct.unpack_from("<i", h, 124)[0],
        "boxsize": struct.unpack_from("<d", h, 128)[0],
        "omega0": struct.unpack_from("<d", h, 136)[0],
        "omega_lambda": struct.unpack_from("<d", h, 144)[0],
        "hubble": struct.unpack_from("<d", h, 152)[0],
        "np_total_highword": struct.unpack_from("<6I", h, 168),
    }
    pos = np.frombuffer(record(), "<f4").reshape(-1, 3)
    vel = np.frombuffer(record(), "<f4").reshape(-1, 3)
    ids = record()
    assert off == len(raw)
    return hdr, pos, vel, ids


@pytest.mark.parametrize("long_id", [False, True])
def test_gadget_writer_layout(tmp_path, long_id):
    rng = np.random.default_rng(3)
    n = 1000
    snap = np.zeros(n, ch.PARTICLE_MIN_DTYPE)
    snap["x"] = rng.uniform(0, 100, (n, 3))
    snap["v"] = rng.normal(0, 300, (n, 3))
    snap["id"] = np.arange(1, n + 1) + (2 ** 33 if long_id else 0)
    a, box, om, h = np.float32(0.5780347), np.float32(100.0), np.float32(0.273), np.float32(0.705)
    path = str(tmp_path / "snp00100a.0")
    ch.write_gadget(path, snap, a, box, om, h, long_id=long_id)
    hdr, pos, vel, ids = read_gadget(path)
    assert hdr["np"] == (0, n, 0, 0, 0, 0)
    assert hdr["mass"][1] == float(om) * 27.7455 * float(box) ** 3 / n   # write.c:45-46
    assert hdr["time"] == float(a)
    assert hdr["redshift"] == 1.0 / float(a) - 1
    assert hdr["np_total"] == (0, n, 0, 0, 0, 0)
    assert hdr["np_total_highword"] == (0,) * 6
    assert hdr["num_files"] == 1
    assert hdr["boxsize"] == float(box)
    assert hdr["omega0"] == float(om)
    assert hdr["omega_lambda"] == 1.0 - float(om)
    assert hdr["hubble"] == float(h)
    assert np.array_equal(pos, snap["x"])
    vfac = np.float32(1.0 / np.sqrt(np.float64(a)))   # write.c:74, double sqrt stored to float
    assert np.array_equal(vel, (vfac * snap["v"]).astype(np.float32))
    if long_id:
        assert np.array_equal(np.frombuffer(ids, "<u8"), snap["id"].astype(np.uint64))
    else:
        assert np.array_equal(np.frombuffer(ids, "<u4"), snap["id"].astype(np.uint32))


def test_gadget_writer_unwritable_path():
    with pytest.raises(ch.ColaError) as e:
        ch.write_gadget("/nonexistent/dir/snp.0", np.zeros(2, ch.PARTICLE_MIN_DTYPE), 1.0, 1.0, 0.3, 0.7)
    assert e.value.code == 9000


def write_gadget_parts(base, parts, a, box, om=0.273, h=0.705, bad_marker=False):
    """Independent writer of a snapshot split over len(parts) GADGET-1 files
    base.0, base.1, ... (num_files in every header, np_total the sum)."""
    ntot = sum(len(x) for x, _ in parts)
    for num, (x, v) in enumerate(parts):
        n = len(x)
        hdr = bytearray(256)
        struct.pack_into("<6i", hdr, 0, 0, n, 0, 0, 0, 0)
        struct.pack_into("<d", hdr, 72, float(a))
        struct.pack_into("<6I", hdr, 96, 0, ntot & 0xffffffff, 0, 0, 0, 0)
        struct.pack_into("<i", hdr, 124, len(parts))
        struct.pack_into("<4d", hdr, 128, float(box), float(om), 1.0 - float(om), float(h))
        struct.pack_into("<6I", hdr, 168, 0, ntot >> 32, 0, 0, 0, 0)
        with open(f"{base}.{num}", "wb") as f:
            for body in (bytes(hdr), np.asarray(x, "<f4").tobytes(), np.asarray(v, "<f4").tobytes(),
                         np.arange(n, dtype="<u4").tobytes()):
                f.write(struct.pack("<i", len(body)))
                f.write(body)
                f.write(struct.pack("<i", len(body) + (1 if bad_marker and len(body) == 12 * n else 0)))


def test_gadget_reader_round_trip(tmp_path):
    """read.c on the writer's file: positions as written (already in the
    box), velocities v / sqrt(a) * sqrt(a) with float rounding, ids 0."""
    rng = np.random.default_rng(5)
    n = 777
    snap = np.zeros(n, ch.PARTICLE_MIN_DTYPE)
    snap["x"] = rng.uniform(0, 100, (n, 3))
    snap["v"] = rng.normal(0, 300, (n, 3))
    snap["id"] = np.arange(1, n + 1)
    a = np.float32(0.5780347)
    path = str(tmp_path / "snp_000")
    ch.write_gadget(path, snap, a, 100.0, 0.273, 0.705)
    rec, info = ch.read_gadget(path)
    assert info["nfiles"] == 1 and info["np_read"] == n and info["np_total"] == n
    assert info["a"] == a and info["boxsize"] == np.float32(100.0)
    assert info["omega_m"] == np.float32(0.273) and info["h"] == np.float32(0.705)
    assert np.array_equal(rec["x"], snap["x"])
    vfac = np.float32(1.0 / np.sqrt(np.float64(a)))
    back = np.float32(np.sqrt(np.float64(a)))          # read.c:76, double sqrt stored to float
    assert np.array_equal(rec["v"], back * (vfac * snap["v"]).astype(np.float32))
    assert np.abs(rec["v"] - snap["v"]).max() <= 1e-6 * np.abs(snap["v"]).max()
    assert not rec["id"].any()
    assert ch.read_gadget(str(tmp_path / "snp_001"))[0] is None   # missing: not an error


def test_gadget_reader_multi_file_and_wrap(tmp_path):
    """A snapshot in two files (base.0, base.1) is read in file order;
    positions outside [0, L) are wrapped as read.c:99-102 does."""
    box, a = np.float32(50.0), np.float32(0.25)
    x0 = np.array([[-1.0, 10.0, 49.999], [50.0, 0.0, 25.0]], np.float32)
    x1 = np.array([[51.5, -0.0001, 3.0], [1.0, 2.0, 3.0], [49.0, 60.0, -50.0]], np.float32)
    v0 = np.full((2, 3), 2.0, np.float32)
    v1 = np.full((3, 3), -4.0, np.float32)
    base = str(tmp_path / "snp_000")
    write_gadget_parts(base, [(x0, v0), (x1, v1)], a, box)
    rec, info = ch.read_gadget(base)
    assert info["nfiles"] == 2 and info["np_read"] == 5 and info["np_total"] == 5
    want = np.concatenate([x0, x1])
    want = np.where(want < 0, want + box, want)
    want = np.where(want >= box, want - box, want)
    assert np.array_equal(rec["x"], want)
    assert (rec["x"] >= 0).all() and (rec["x"] < box).all()
    back = np.float32(np.sqrt(np.float64(a)))
    assert np.array_equal(rec["v"], back * np.concatenate([v0, v1]))


def test_gadget_reader_errors(tmp_path):
    base = str(tmp_path / "snp_000")
    two = [(np.zeros((4, 3)), np.zeros((4, 3))), (np.zeros((2, 3)), np.zeros((2, 3)))]
    write_gadget_parts(base, two, 0.5, 10.0, bad_marker=True)
    with pytest.raises(ch.ColaError) as e:
        ch.read_gadget(base)
    assert e.value.code == 15005   # read.c:16-27
    # one file named base.0 with num_files = 1: read.c then opens `base`
    # itself (read.c:60-63), which does not exist
    write_gadget_parts(base, [(np.zeros((4, 3)), np.zeros((4, 3)))], 0.5, 10.0)
    with pytest.raises(ch.ColaError) as e:
        ch.read_gadget(base)
    assert e.value.code == 15000
    write_gadget_parts(base, two, 0.5, 10.0)
    out = np.zeros(5, ch.PARTICLE_MIN_DTYPE)
    info = ch.GadgetInfo()
    code = ch.lib().cola_read_gadget(base.encode(), out.ctypes.data, 5, ch.C.byref(info))
    assert code == 15001           # read.c:80-84: more particles than room
    # a header with a negative type-1 count (np[1] after the 4-byte marker
    # and np[0]) is refused as malformed, before any buffer is sized by it
    raw = bytearray(open(f"{base}.0", "rb").read())
    struct.pack_into("<i", raw, 8, -5)
    open(f"{base}.0", "wb").write(bytes(raw))
    with pytest.raises(ch.ColaError) as e:
        ch.read_gadget(base)
    assert e.value.code == 15005


def test_coarse_grid_file_layout(tmp_path):
    """tools/density_grid.c:1-28: float boxsize, int nc, nc^3 floats in C
    order, int nc."""
    n = 6
    delta = np.random.default_rng(1).standard_normal((n, n, n)).astype(np.float32)
    path = str(tmp_path / "grid00001a.b")
    ch.write_coarse_grid(path, 250.0, delta)
    raw = open(path, "rb").read()
    assert len(raw) == 4 + 4 + 4 * n ** 3 + 4
    assert struct.unpack_from("<f", raw, 0)[0] == 250.0
    assert struct.unpack_from("<i", raw, 4)[0] == n
    assert np.array_equal(np.frombuffer(raw, "<f4", n ** 3, 8).reshape(n, n, n), delta)
    assert struct.unpack_from("<i", raw, 8 + 4 * n ** 3)[0] == n
    with pytest.raises(ch.ColaError) as e:
        ch.write_coarse_grid("/nonexistent/grid.b", 1.0, delta)
    assert e.value.code == 10021


def test_subsample_selection_and_layout(tmp_path):
    """subsample.c:115-159: record i is kept when the next ranlxd1 uniform of
    the stream seeded with 2 * seed is below the factor; the stream carries
    over to the next output."""
    import oracle_lib as orc
    rng = np.random.default_rng(4)
    n, factor, seed = 5000, 0.05, 100
    snap = np.zeros(n, ch.PARTICLE_MIN_DTYPE)
    snap["x"] = rng.uniform(0, 100, (n, 3))
    snap["v"] = rng.normal(0, 300, (n, 3))
    sub = ch.Subsampler(factor, seed)
    r = orc.Ranlxd()
    orc.lib().orc_ranlxd_set(r, 2 * seed, 202)
    a, box, om, h = np.float32(0.5), np.float32(100.0), np.float32(0.273), np.float32(0.705)
    for k in range(2):  # two outputs from one stream
        path = str(tmp_path / f"sub00100{'ab'[k]}.b")
        nsub = sub.write(path, snap, a, box, om, h)
        keep = np.array([orc.lib().orc_ranlxd_uniform(r) < factor for _ in range(n)])
        raw = open(path, "rb").read()
        hdr = np.frombuffer(raw, "<f4", 6)
        assert hdr[0] == box and hdr[2] == om and hdr[3] == h and hdr[4] == a
        assert hdr[1] == np.float32(float(om) * 27.7455 * float(box) ** 3 / n)
        assert hdr[5] == np.float32(1.0 / float(a) - 1.0)
        cnt = struct.unpack_from("<i", raw, 24)[0]
        assert cnt == nsub == keep.sum()
        body = np.frombuffer(raw, "<f4", 6 * cnt, 28).reshape(cnt, 6)
        assert np.array_equal(body[:, :3], snap["x"][keep])
        assert np.array_equal(body[:, 3:], snap["v"][keep])
        assert struct.unpack_from("<i", raw, 28 + 24 * cnt)[0] == cnt
        assert len(raw) == 32 + 24 * cnt
    sub.close()


def test_extension_keys(tmp_path):
    """pm_ngrid and residual_net (keys of this build, unknown to the
    reference reader) parse; files without them read as before (0 / absent);
    an empty residual_net or an odd pm_ngrid fails the confirmation."""
    path = tmp_path / "param.lua"
    base = open(os.path.join(HERE, "golden", "param_full.lua")).read()
    path.write_text(base + 'pm_ngrid = 72\nresidual_net = "nets/x.net"  -- learned residual\n')
    p = ch.read_parameters(str(path))
    d = p.as_dict()
    assert d["pm_ngrid"] == 72 and d["residual_net_filename"] == "nets/x.net"
    assert d["nc"] == 48 and d["pm_nc_factor"] == 2
    ch.confirm_parameters(p)
    path.write_text(base)
    d = ch.read_parameters(str(path)).as_dict()
    assert d["pm_ngrid"] == 0 and d["residual_net_filename"] is None
    for extra, code in (('residual_net = ""\n', 10024), ("pm_ngrid = 73\n", 10025)):
        path.write_text(base + extra)
        q = ch.read_parameters(str(path))
        with pytest.raises(ch.ColaError) as e:
            ch.confirm_parameters(q)
        assert e.value.code == code
