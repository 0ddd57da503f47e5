"""The cola_halo driver program end to end on the GPU: a parameter file in the
reference's syntax, two output redshifts, every output kind, checked file by
file against the same run driven from Python through the library."""
import os
import struct
import subprocess

import numpy as np
import pytest

import cola_halo_amd as ch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "cola_halo_amd", "cola_halo")
PK_FILE = os.path.join(HERE, "golden", "linear_matterpower.dat")

PARAM = """-- cola code parameter file (test)
nc = 32
boxsize = 64.0
random_seed= 7
nrealization= 2
ntimestep= 10
a_final= 1.0
output_redshifts= {{1.0, 0.0}}
omega_m = 0.273
h       = 0.705
sigma8  = 0.812
pm_nc_factor= 3
loglevel=2
powerspectrum= "{pk}"
fof= "{d}/fof"
linking_factor= 0.2
snapshot= "{d}/snp"
subsample= "{d}/sub"
subsample_factor= 0.05
coarse_grid= "{d}/grid"
coarse_grid_nc= 16
initial= "{d}/init"
write_longid= false
"""


def _gadget(path):
    raw = open(path, "rb").read()
    off = 0
    recs = []
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        recs.append(raw[off + 4: off + 4 + n])
        off += 8 + n
    time = struct.unpack_from("<d", recs[0], 72)[0]
    return time, np.frombuffer(recs[1], "<f4").reshape(-1, 3), np.frombuffer(recs[2], "<f4").reshape(-1, 3), \
        np.frombuffer(recs[3], "<u4")


def _wrap(x, box):
    b = np.float32(box)
    return np.where(x < 0, x + b, np.where(x >= b, x - b, x)).astype(np.float32)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_outputs_match_library_run(tmp_path):
    print("driver run", flush=True)
    param = tmp_path / "param.lua"
    param.write_text(PARAM.format(pk=PK_FILE, d=tmp_path))
    r = subprocess.run([DRIVER, str(param)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Timestep 10/10" in r.stdout

    nc, box, om = 32, 64.0, 0.273
    Om = np.float32(om)
    sub = ch.Subsampler(0.05, 7)
    with ch.Context(nc, 3 * nc, box) as ctx:
        ctx.power_init(PK_FILE, 0.1, 0.812, om, 1 - om)
        for seed in (7, 8):
            # initial condition file (main.c:131-137)
            ctx.lpt_set_displacement(0.1, om, seed)
            init = ctx.set_noncola_initial(Om)
            t, pos, vel, ids = _gadget(tmp_path / f"init_{seed:05d}.0")
            assert t == np.float32(0.1)
            assert np.array_equal(pos, init["x"]) and np.array_equal(ids, init["id"].astype(np.uint32))

            snaps = []

            def on_snapshot(a, s):
                ll = float(np.float32(0.2 * np.float32(box) / nc))
                halos = ctx.fof_find_halos(ll, 32)
                snaps.append((a, ctx.get_snapshot(), halos, ctx.coarse_grid(16)[0]))

            ch.run_cola(ctx, om, 1.0, 10, seed, [1.0, 0.0], on_snapshot=on_snapshot)
            assert len(snaps) == 2
            for iout, (a, snap, halos, grid) in enumerate(snaps):
                sfx = "ab"[iout]
                t, pos, vel, ids = _gadget(tmp_path / f"snp{seed:05d}{sfx}.0")
                assert t == np.float32(a)
                assert np.array_equal(pos, snap["x"])                      # wrapped by FoF, as the reference
                assert np.array_equal(pos, _wrap(snap["x"], box))
                vfac = np.float32(1.0 / np.sqrt(np.float64(np.float32(a))))
                assert np.array_equal(vel, (vfac * snap["v"]).astype(np.float32))
                assert np.array_equal(ids, np.arange(1, nc ** 3 + 1, dtype=np.uint32))
                lines = open(tmp_path / f"fof{seed:05d}{sfx}.txt").read().splitlines()
                assert len(lines) == len(halos) > 0
                for line, h in zip(lines, halos):
                    assert line == "%d %e %e %e %e %e %e" % (h["nfof"], *map(float, h["x"]), *map(float, h["v"]))
                raw = open(tmp_path / f"grid{seed:05d}{sfx}.b", "rb").read()
                assert np.array_equal(np.frombuffer(raw, "<f4", 16 ** 3, 8).reshape(16, 16, 16), grid)
                sub_path = str(tmp_path / f"check_sub{seed}{sfx}.b")
                sub.write(sub_path, snap, np.float32(a), np.float32(box), Om, np.float32(0.705))
                assert open(sub_path, "rb").read() == open(tmp_path / f"sub{seed:05d}{sfx}.b", "rb").read()
    sub.close()


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_abort_codes(tmp_path):
    param = tmp_path / "param.lua"
    param.write_text("nc = 16\nfoo = 1\n")
    r = subprocess.run([DRIVER, str(param)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1007 & 0xFF and "unknown variable" in r.stderr
    param.write_text(PARAM.format(pk="/nonexistent/pk.dat", d=tmp_path))
    r = subprocess.run([DRIVER, str(param)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3000 & 0xFF


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_matches_oracle_run(tmp_path):
    """The driver's z = 0 GADGET positions against the CPU oracle's run of the
    same parameters (main.c schedule, the reference's P(k) file): within
    1e-3 of the mean interparticle spacing, velocities within 1e-3 of their
    scale (the GADGET file holds v / sqrt(a), write.c:74)."""
    import oracle_lib as orc

    print("driver vs oracle", flush=True)
    nc, box, seed = 32, 64.0, 7
    param = tmp_path / "param.lua"
    text = PARAM.format(pk=PK_FILE, d=tmp_path).replace("nrealization= 2", "nrealization= 1")
    for key in ("fof", "subsample", "coarse_grid", "initial"):
        text = "\n".join(l for l in text.splitlines() if not l.startswith(key + "=")) + "\n"
    param.write_text(text.replace("output_redshifts= {1.0, 0.0}", "output_redshifts= {0.0}"))
    r = subprocess.run([DRIVER, str(param)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    t, pos, vel, ids = _gadget(tmp_path / f"snp{seed:05d}a.0")
    assert orc.lib().orc_power_read_file(PK_FILE.encode()) == 0
    ref = np.zeros(nc ** 3, orc.PARTICLE_MIN_DTYPE)
    a_out = np.zeros(1, np.float32)
    nw = orc.C.c_int(0)
    zz = np.zeros(1)
    assert orc.lib().orc_run(nc, 3 * nc, box, 0.273, 1.0, 10, seed, zz.ctypes.data, 1, ref.ctypes.data,
                             a_out.ctypes.data, orc.C.byref(nw)) == 0
    assert np.array_equal(ids, ref["id"].astype(np.uint32))
    d = np.abs(pos.astype(np.float64) - ref["x"]) % box
    assert np.minimum(d, box - d).max() <= 1e-3 * box / nc
    v = vel.astype(np.float64) * np.sqrt(np.float64(t))
    assert np.abs(v - ref["v"]).max() <= 1e-3 * np.abs(ref["v"]).max()


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_corrected_run(tmp_path):
    """pm_ngrid and residual_net: the driver runs 64^3 particles on a 96^3
    grid (factor 1.5, cell 1.30 Mpc/h) with the committed net and writes the
    same snapshot as the library run with the net attached, which differs
    from the plain run on the same grid."""
    print("driver corrected run", flush=True)
    net_file = os.path.join(os.path.dirname(HERE), "cola_halo_amd", "nets", "eh98_cell1.30.net")
    nc, ng, box, seed = 64, 96, 125.0, 3
    param = tmp_path / "param.lua"
    param.write_text(f"""nc = {nc}
boxsize = {box}
random_seed = {seed}
ntimestep = 10
output_redshifts = {{0.0}}
omega_m = 0.273
h = 0.705
sigma8 = 0.812
powerspectrum = "{PK_FILE}"
snapshot = "{tmp_path}/snp"
pm_ngrid = {ng}            -- factor 1.5
residual_net = "{net_file}"
""")
    r = subprocess.run([DRIVER, str(param)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "learned force residual" in r.stdout
    t, pos, vel, ids = _gadget(tmp_path / f"snp{seed:05d}a.0")
    with ch.Context(nc, ng, box) as ctx, ch.ResidualNet.from_file(ng, net_file) as net:
        ctx.power_init(PK_FILE, 0.1, 0.812, 0.273, 0.727)
        ctx.set_residual_net(net)
        (a, snap), = ch.run_cola(ctx, 0.273, 1.0, 10, seed, [0.0])
        ctx.set_residual_net(None)
        (a, plain), = ch.run_cola(ctx, 0.273, 1.0, 10, seed, [0.0])
    assert np.array_equal(pos, snap["x"])
    assert not np.array_equal(pos, plain["x"])
    d = np.abs(pos.astype(np.float64) - plain["x"]) % box
    d = np.minimum(d, box - d)  # periodic distance
    assert np.sqrt(np.mean(d ** 2)) < 0.5 * box / nc  # a correction, not a different run


HALO = os.path.join(os.path.dirname(HERE), "cola_halo_amd", "halo")

PARAM_RUN = """nc = 32
boxsize = 64.0
random_seed= 7
nrealization= 1
ntimestep= 10
a_final= 1.0
output_redshifts= {{0.0}}
omega_m = 0.273
h       = 0.705
sigma8  = 0.812
pm_nc_factor= 3
loglevel=2
powerspectrum= "{pk}"
fof= "{d}/fof"
linking_factor= 0.2
snapshot= "{d}/snp"
subsample= "{d}/sub"
subsample_factor= 0.05
coarse_grid= "{d}/grid"
coarse_grid_nc= 16
"""


def _halo_lines(path):
    rows = [line.split() for line in open(path)]
    return np.array([int(r[0]) for r in rows]), np.array([[float(v) for v in r[1:4]] for r in rows]), \
        np.array([[float(v) for v in r[4:7]] for r in rows])


@pytest.mark.skipif(not (os.path.exists(DRIVER) and os.path.exists(HALO)), reason="programs not built")
def test_halo_program_reads_driver_snapshots(tmp_path):
    """halo (halo_main.c) on the driver's own snapshot: the snapshot read
    back (read.c) gives the driver's FoF catalogue (same groups and centres;
    velocities through v / sqrt(a) * sqrt(a) in float), its coarse grid and
    the same subsample records."""
    print("halo program", flush=True)
    run = tmp_path / "run"
    run.mkdir()
    param = tmp_path / "param.lua"
    param.write_text(PARAM_RUN.format(pk=PK_FILE, d=run))
    r = subprocess.run([DRIVER, str(param)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    post = tmp_path / "post"
    post.mkdir()
    # halo_main.c names snapshots <dir>/<snapshot>_NNN; one file, no .0
    os.replace(run / "snp00007a.0", post / "snp_000")
    (tmp_path / "dirs.txt").write_text(f"{post}\n")
    hparam = tmp_path / "halo.lua"
    hparam.write_text(PARAM_RUN.format(pk=PK_FILE, d=run).replace(f'"{run}/fof"', '"fof"')
                      .replace(f'"{run}/snp"', '"snp"').replace(f'"{run}/sub"', '"sub"')
                      .replace(f'"{run}/grid"', '"grid"'))
    r = subprocess.run([HALO, str(hparam)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "halo done" in r.stdout
    n0, x0, v0 = _halo_lines(run / "fof00007a.txt")
    n1, x1, v1 = _halo_lines(post / "fof_000.txt")
    assert len(n0) > 10
    assert np.array_equal(n0, n1)
    assert np.array_equal(x0, x1)
    assert np.allclose(v1, v0, rtol=1e-5, atol=1e-3)
    assert open(run / "grid00007a.b", "rb").read() == open(post / "grid_000.b", "rb").read()
    s0 = np.frombuffer(open(run / "sub00007a.b", "rb").read()[28:-4], "<f4").reshape(-1, 6)
    s1 = np.frombuffer(open(post / "sub_000.b", "rb").read()[28:-4], "<f4").reshape(-1, 6)
    assert s0.shape == s1.shape and len(s0) > 100
    assert np.array_equal(s0[:, :3], s1[:, :3])
    assert np.allclose(s0[:, 3:], s1[:, 3:], rtol=1e-5, atol=1e-3)
    # no snapshot at all: abort code 2 (halo_main.c:139-141)
    (tmp_path / "dirs.txt").write_text(f"{run}\n")
    r = subprocess.run([HALO, str(hparam)], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    assert r.returncode == 2
