"""Write tests/golden/fullsize_<name>.npz: the oracle's complete runs at the
configurations' own sizes, too slow to repeat inside a GPU test.

Each fixture is one oracle run (oracle/libcola_oracle.so: 2LPT from the
seed, ten COLA steps a = 0.1 -> 1, snapshot at a = 1, EH98 no-wiggle table
of the benchmark cosmology) reduced to
  * the snapshot records (x, v) of a fixed random sample of particle IDs,
  * the shell power spectrum of all its particles: CIC on an n^3 grid
    (coarse_grid.c's periodic assignment), CIC window divided out, linear
    bins on [0, k_Nyq / 2], k_Nyq = pi nc / L of the particle lattice.
The GPU tests run the same configuration and compare.

Usage: OMP_NUM_THREADS=8 python tests/golden/gen_fullsize_golden.py [name ...]
(the C2 run takes ~10 minutes on 8 cores).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import cola_halo_amd as ch  # noqa: E402  (host-only table generator)
import oracle_lib as orc  # noqa: E402

OM = 0.273
NSAMPLE = 20000
NBINS = 16
# name: (nc, ngrid, box, seed, estimator grid)
CONFIGS = {
    "c2": (256, 768, 500.0, 1, 256),
    "c3_96": (96, 288, 187.5, 1, 192),
    "c3_128": (128, 384, 250.0, 17, 256),
    "c3_160": (160, 480, 312.5, 33, 320),
}


def sample_ids(np_total):
    rng = np.random.default_rng(1234)
    return np.sort(rng.choice(np_total, size=min(NSAMPLE, np_total), replace=False)) + 1


def run(name):
    nc, ng, box, seed, npk = CONFIGS[name]
    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    orc.set_power_table(k, p)
    np_total = nc ** 3
    snap = np.zeros(np_total, orc.PARTICLE_MIN_DTYPE)
    a_out = np.zeros(1, np.float32)
    nw = orc.C.c_int(0)
    zz = np.zeros(1, np.float64)
    t0 = time.time()
    st = orc.lib().orc_run(nc, ng, box, OM, 1.0, 10, seed, zz.ctypes.data, 1, snap.ctypes.data, a_out.ctypes.data,
                           orc.C.byref(nw))
    assert st == 0 and nw.value == 1, (st, nw.value)
    t_run = time.time() - t0
    assert np.array_equal(snap["id"], np.arange(1, np_total + 1))
    kmax = 0.5 * np.pi * nc / box
    delta, _ = orc.coarse_grid(snap, npk, box)
    kb, pk, nm = orc.power_spectrum(delta, box, 0.0, kmax, NBINS, flags=1)
    ids = sample_ids(np_total)
    rec = snap[ids - 1]
    np.savez_compressed(os.path.join(HERE, f"fullsize_{name}.npz"), nc=nc, ngrid=ng, box=box, seed=seed,
                        a=float(a_out[0]), ids=ids, x=rec["x"], v=rec["v"], pk_grid=npk, kmax=kmax, k=kb, pk=pk,
                        nmodes=nm)
    print(f"{name}: oracle run {t_run:.1f} s, P(k) on {npk}^3", flush=True)


if __name__ == "__main__":
    for name in sys.argv[1:] or list(CONFIGS):
        run(name)
