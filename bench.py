"""Benchmark: particle-steps/s of plain COLA on one MI355X.

Workload (the configuration the metric is quoted on): 256^3 particles,
768^3 PM grid (pm_nc_factor 3), box 500 Mpc/h, 2LPT ICs from seed 1 at
a = 0.1, linear-in-a steps of da = 0.1 to a = 1, EH98 no-wiggle P(k)
(Om 0.273, Ob 0.046, h 0.705, ns 0.96, sigma8 0.8).

A step is one COLA time step of main.c:162-206: wrap, PM force, kick, drift
(fused readout/kick/drift); the tenth step of each 10-step run is the final
force evaluation at a = 1 (the snapshot itself is outside the timed region,
as in the reference's COLA timer category).  particle-steps = nc^3 x force
evaluations.  Inputs are resident in HBM when the timed region starts.

Prints one JSON line.  The roofline object is for the kernel with the most
device time in the timed region, from HIP events on the context's stream;
cpu_baseline times the CPU oracle (C restatement, OpenMP) on a bounded
sample: one step of the same workload from the same initial conditions.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s on 1 MI355X at P(k) err ≤1% (k≤1 h/Mpc) vs plain fine grid"

NC, NG, BOX, SEED, NSTEPS = 256, 768, 500.0, 1, 10
OM = 0.273


def kernel_bytes(phase, nc, ng):
    """Algorithmic HBM bytes of one launch: each input array read once and
    each output array written once (FP32), per the canonical per-phase model."""
    n_p = nc ** 3
    n_g = ng ** 3
    n_c = ng * ng * (ng // 2 + 1)
    table = {
        "fft_z_r2c": 4 * n_g + 8 * n_c,
        "fft_y_fwd": 16 * n_c,
        "fft_x_fwd": 16 * n_c,
        "fft_x_inv": 16 * n_c,   # reads delta_k, writes F_k (Poisson fused on load)
        "fft_x_fused": 32 * n_c,  # reads delta_k once, writes the three F_k
        "fft_y_inv": 16 * n_c,
        "fft_z_c2r": 8 * n_c + 4 * n_g,
        "deposit": 4 * n_g + 12 * n_p,
        "readout_kick_drift": 12 * n_g + 72 * n_p,
        "bin_count": 12 * n_p,
        "bin_scatter": 24 * n_p,
    }
    return table.get(phase)


def measured_traffic(kernel, nc, ng):
    """HBM bytes per launch of `kernel` from the calibrated rocprofv3 counter
    table (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a copy kernel of known size), or None when the table
    does not cover this kernel at this workload."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    for entry in table.get("workloads", []):
        if entry.get("nc") == nc and entry.get("ngrid") == ng:
            k = entry.get("kernels", {}).get(kernel)
            if k is not None:
                return int(k["fetch_bytes"] + k["write_bytes"])
    return None


def cpu_baseline(ctx_parts, nc, ng, box, threads):
    """Oracle on the host: one PM force + kick + drift of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    import ctypes as C

    import oracle_lib as orc

    parts = ctx_parts.copy()
    wpar = float(np.float32((ng / nc) ** 3))
    force = np.zeros((len(parts), 3), np.float32)
    t0 = time.perf_counter()
    lib = orc.lib()
    lib.orc_wrap(parts.ctypes.data, len(parts), box)
    st = lib.orc_pm_calculate_forces(ng, box, wpar, len(parts), parts.ctypes.data, len(parts), force.ctypes.data,
                                     None, None, C.byref(C.c_double()))
    ax, av = C.c_float(0.1), C.c_float(0.1)
    Om = float(np.float32(OM))
    lib.orc_kick(parts.ctypes.data, force.ctypes.data, len(parts), Om, C.byref(ax), C.byref(av), 0.15)
    lib.orc_drift(parts.ctypes.data, len(parts), Om, C.byref(ax), C.byref(av), 0.2)
    dt = time.perf_counter() - t0
    return {"value": nc ** 3 / dt, "unit": "particle-steps/s", "cores": int(threads), "kind": "port",
            "sample": f"1 step (wrap + PM force on {ng}^3 + kick + drift) of the {nc}^3 workload, "
                      f"{dt:.1f} s, oracle status {st}"}


ENSEMBLE = [(96, 288, 187.5)] * 16 + [(128, 384, 250.0)] * 16 + [(160, 480, 312.5)] * 16


def ensemble_bench(args):
    """48 uneven realizations (16 x 96^3/288^3, 16 x 128^3/384^3,
    16 x 160^3/480^3, seeds 1-48, 10 steps each): the step loops run one
    realization after another, then interleaved across all contexts with
    graph-captured steps.  Prints one JSON line."""
    import cola_halo_amd as ch

    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    da = 1.0 / NSTEPS
    Om = float(np.float32(OM))
    ctxs = [ch.Context(nc, ng, box) for nc, ng, box in ENSEMBLE]
    for c in ctxs:
        c.power_init_table(k, p, da, 0.8, OM, 1.0 - OM)

    def prepare(graph):
        for i, c in enumerate(ctxs):
            c.set_step_graph(graph)
            c.lpt_set_displacement(da, OM, i + 1)
            c.set_times(c.times()[0], da)
        for c in ctxs:
            c.synchronize()

    def finish():
        for c in ctxs:
            c.move_particles()
            c.pm_calculate_forces()
        for c in ctxs:
            c.synchronize()

    # launch shapes tuned once per size (on the first context of each size,
    # its initial conditions), outside the timed region
    prepare(False)
    tuned = {}
    for (nc, ng, box), c in zip(ENSEMBLE, ctxs):
        if ng not in tuned:
            tuned[ng] = c.autotune(reps=3)
        else:
            c.apply_tuning(tuned[ng])
    particles = sum(nc ** 3 for nc, _, _ in ENSEMBLE)
    out = {}
    for mode in ("sequential", "batched_graph"):
        prepare(mode != "sequential")
        t0 = time.perf_counter()
        if mode == "sequential":
            for c in ctxs:
                for istep in range(1, NSTEPS):
                    c.step(Om, (istep + 0.5) * da, (istep + 1.0) * da)
                c.move_particles()
                c.pm_calculate_forces()
                c.synchronize()
        else:
            for istep in range(1, NSTEPS):
                ch.step_batch(ctxs, Om, [(istep + 0.5) * da] * len(ctxs), [(istep + 1.0) * da] * len(ctxs))
            finish()
        dt = time.perf_counter() - t0
        out[mode] = {"seconds": round(dt, 4), "particle_steps_per_s": round(particles * NSTEPS / dt, 1)}
    for c in ctxs:
        c.close()
    print(json.dumps({"metric": "ensemble particle-steps/s, 48 uneven realizations", "value":
                      out["batched_graph"]["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "16 x 96^3/288^3 L187.5 + 16 x 128^3/384^3 L250 + 16 x 160^3/480^3 "
                                             "L312.5, seeds 1-48, 10 steps", "particles": particles},
                      "modes": out, "tuning": {str(k): v for k, v in tuned.items()}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nc", type=int, default=NC)
    ap.add_argument("--ng", type=int, default=NG)
    ap.add_argument("--box", type=float, default=BOX)
    ap.add_argument("--tile", type=str, default="", help="deposit tile tx,ty,tz")
    ap.add_argument("--fft-log2-tile", type=int, default=-2, help="FFT transforms per workgroup (log2)")
    ap.add_argument("--xcd-group", type=int, default=0, help="FFT column tiles per XCD group")
    ap.add_argument("--ensemble", action="store_true",
                    help="time the 48-realization ensemble (sequential vs batched graph steps) instead")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of tuning them on the initial conditions")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the plain copy kernel on this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import cola_halo_amd as ch

        print(json.dumps({"copy_probe_mib": args.copy_probe,
                          "gbps": ch.bandwidth_probe(args.copy_probe << 20, 5)}))
        return
    if args.ensemble:
        ensemble_bench(args)
        return
    if args.gpus != 1:
        print("bench.py measures a single GPU: ensembles would shard by seed and one realization by x slabs, "
              "neither of which this benchmark runs",
              file=sys.stderr)
        sys.exit(2)

    import cola_halo_amd as ch

    nc, ng, box = args.nc, args.ng, args.box
    da = 1.0 / NSTEPS
    a_init = da
    Om = float(np.float32(OM))
    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    ctx = ch.Context(nc, ng, box)
    if args.tile or args.fft_log2_tile != -2:
        tx, ty, tz = (int(v) for v in args.tile.split(",")) if args.tile else (8, 8, 32)
        ctx.set_tiling(tx, ty, tz, args.fft_log2_tile if args.fft_log2_tile != -2 else -1)
    if args.xcd_group:
        ctx.set_fft_xcd_group(args.xcd_group)
    ctx.power_init_table(k, p, a_init, 0.8, OM, 1.0 - OM)
    t_ic = time.perf_counter()
    ctx.lpt_set_displacement(a_init, OM, SEED)
    ctx.set_times(a_init, da)  # main.c:145
    ctx.synchronize()
    t_ic = time.perf_counter() - t_ic
    ctx.state_save()
    # launch-shape tuning on the initial state, outside the timed region (the
    # reference plans its FFTs with FFTW_MEASURE at pm_init, pm.c:139-147)
    tuning = None
    if not args.no_autotune and not (args.tile or args.fft_log2_tile != -2 or args.xcd_group):
        tuning = ctx.autotune(reps=3)
        ctx.state_restore()
        ctx.synchronize()

    def run_steps(nsteps):
        forces = 0
        for s in range(nsteps):
            istep = s % NSTEPS + 1
            if istep == 1 and s > 0:
                ctx.state_restore()
            if istep < NSTEPS:
                ctx.step(Om, (istep + 0.5) * da, (istep + 1.0) * da)
            else:
                ctx.move_particles()
                ctx.pm_calculate_forces()
            forces += 1
        return forces

    # warmup, then back to the initial conditions
    run_steps(args.warmup)
    ctx.synchronize()
    ctx.state_restore()
    ctx.synchronize()

    ctx.set_profiling(True)
    ctx.reset_phase_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    forces = run_steps(args.steps)
    ctx.synchronize()
    t1 = time.perf_counter()
    stats = ctx.phase_stats()
    ctx.set_profiling(False)
    elapsed = t1 - t0
    value = nc ** 3 * forces / elapsed

    # roofline of the kernel with the most device time
    dom, (dom_ms, dom_n) = max(stats.items(), key=lambda kv: kv[1][0])
    nbytes = kernel_bytes(dom, nc, ng)
    roof = None
    if nbytes is not None and dom_n > 0:
        avg_s = dom_ms / dom_n / 1e3
        achieved = nbytes / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, nc, ng), "kernel": dom,
                "avg_launch_ms": round(dom_ms / dom_n, 4), "bytes_per_launch": nbytes}
    # whole step against the canonical bytes per step (156 Np + 32 Ng + 64 Nc)
    n_c = ng * ng * (ng // 2 + 1)
    b_step = 156 * nc ** 3 + 32 * ng ** 3 + 64 * n_c
    step_frac = b_step / (elapsed / forces) / 1e9 / HBM_PEAK_GBS

    # output side on the final state (a = 1), outside the timed region
    outputs = None
    if args.steps % NSTEPS == 0:
        ctx.synchronize()
        t = time.perf_counter()
        ctx.set_snapshot(1.0, Om)
        t_snap = time.perf_counter() - t
        ll = float(np.float32(0.2 * np.float32(box) / nc))  # main.c:353, linking factor 0.2
        t = time.perf_counter()
        halos = ctx.fof_find_halos(ll, 32)
        t_fof = time.perf_counter() - t
        t = time.perf_counter()
        ctx.coarse_grid(nc)
        t_grid = time.perf_counter() - t
        t = time.perf_counter()
        ctx.power_spectrum(ng, 0.0, 1.0, 50, deconvolve=True)
        t_pk = time.perf_counter() - t
        outputs = {"snapshot_to_host_ms": round(t_snap * 1e3, 2), "fof_ms": round(t_fof * 1e3, 2),
                   "halos": int(len(halos)), f"coarse_grid_{nc}_ms": round(t_grid * 1e3, 2),
                   f"power_spectrum_{ng}_ms": round(t_pk * 1e3, 2)}

    cpu = None
    if not args.no_cpu_baseline:
        ctx.state_restore()
        parts = ctx.get_particles()
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(parts, nc, ng, box, threads)
    ctx.close()

    phases = {k2: round(v[0] / max(forces, 1), 3) for k2, v in stats.items() if v[1] > 0}
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain COLA {nc}^3 particles, {ng}^3 PM grid, box {box:g} Mpc/h, 2LPT seed {SEED}, "
                               f"a 0.1->1 in {NSTEPS} steps, EH98 no-wiggle P(k) sigma8 0.8",
                   "nc": nc, "ngrid": ng, "boxsize": box, "seed": SEED, "parallelism": "single GPU"},
        "roofline": roof,
        "step_canonical_frac": round(step_frac, 4),
        "phase_ms_per_step": phases,
        "ic_seconds": round(t_ic, 3),
        "tuning": tuning,
        "outputs": outputs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
