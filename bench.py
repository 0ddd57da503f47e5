"""Benchmark: particle-steps/s on one MI355X of the learned-corrected coarse
COLA run, gated on its P(k) agreeing with the plain fine-grid run.

Default workload (the production comparison): 512^3 particles, box 1000
Mpc/h, 2LPT ICs from seed 1 (timed; seeds 1-4 for the P(k) gate) at a =
0.1, ten linear-in-a steps to a = 1,
EH98 no-wiggle P(k) (Om 0.273, Ob 0.046, h 0.705, ns 0.96, sigma8 0.8).
  * corrected run: 768^3 PM grid (factor 1.5) + the learned force residual
    (4-layer 3-D conv net, BF16 MFMA, weights in cola_halo_amd/nets/,
    trained on 128^3-particle force pairs in a 250 Mpc/h box: same cell
    size and particle spacing);
  * plain fine run: 1536^3 PM grid (factor 3), no net.
A step is one COLA time step of main.c:162-206 (wrap, PM force, fused
readout/kick/drift); the tenth step of each 10-step run is the final force
at a = 1.  particle-steps = nc^3 x force evaluations.  Inputs are resident
in HBM when a timed region starts; each run's K steps are timed on their
own (synchronised on both sides) and cycle the 10-step schedule from the
saved initial conditions.

`value` is the corrected run's particle-steps/s when the P(k) gate holds:
after complete untimed 10-step runs of each of seeds 1-4, |P_corrected /
P_fine - 1| <= 1% for every shell with k <= 1 h/Mpc and every seed (768^3
shell estimator, CIC window divided out); it is 0 when the gate fails.
The FoF mass-function ratio counts the halos of all four seeds.  `speedup_vs_fine` is the ratio
of the two runs' throughputs.  `roofline` is for the kernel with the most
device time per corrected step (HIP events on the context's stream during
the timed region); `cpu_baseline` times the CPU restatement (oracle, C +
OpenMP) on one plain step of 128^3 particles on a 384^3 grid in 250 Mpc/h
(the fine run's cell size and particle spacing).  `--workload plain` times
plain COLA at 256^3 / 768^3 (box 500) instead, `--ensemble` the 48
realizations of uneven sizes.
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
MFMA_BF16_PEAK_TFLOPS = 2500.0
METRIC = "particle-steps/s on 1 MI355X at P(k) err ≤1% (k≤1 h/Mpc) vs plain fine grid"

NSTEPS = 10
OM = 0.273
PROD = dict(nc=512, coarse=768, fine=1536, box=1000.0, seed=1, pk_seeds=(1, 2, 3, 4))
PLAIN = dict(nc=256, ng=768, box=500.0, seed=1)
NET_WEIGHTS = os.path.join(ROOT, "cola_halo_amd", "nets", "eh98_cell1.30.net")
PK_KMAX, PK_BINS, PK_TOL = 1.0, 25, 0.01


# launches of one kernel inside one profiling phase: the hidden layers are one
# phase of two k_conv16 launches (layers 2 and 3), or of one (layer 3) when
# layers 1-2 run fused (phase net_conv12); with layers 3-4 fused (phase
# net_conv34) there is none
def launches_per_phase(phase, stats):
    if phase == "net_conv_hidden":
        return 1 if stats.get("net_conv12", (0.0, 0))[1] > 0 else 2
    return 1


def kernel_cost(phase, nc, ng):
    """(algorithmic HBM bytes, MFMA flops, kernel) of one launch: each input
    array read once and each output array written once (FP32 PM arrays,
    bfloat16 activations), per the canonical per-phase model."""
    n_p = nc ** 3
    n_g = ng ** 3
    n_c = ng * ng * (ng // 2 + 1)
    conv16 = 27 * 16 * 16 * 2 * n_g  # useful flops of one 16 -> 16 channel 3^3 layer
    table = {
        "fft_z_r2c": (4 * n_g + 8 * n_c, 0, "fft_tile (z r2c rows)"),
        "fft_y_fwd": (16 * n_c, 0, "fft_tile (y columns)"),
        "fft_x_fwd": (16 * n_c, 0, "fft_tile (x columns)"),
        "fft_x_inv": (16 * n_c, 0, "fft_tile (x columns)"),
        "fft_x_fused": (32 * n_c, 0, "fft_x_fused"),  # reads delta_k once, writes the three F_k
        "fft_y_inv": (16 * n_c, 0, "fft_tile (y columns)"),
        "fft_z_c2r": (8 * n_c + 4 * n_g, 0, "fft_tile (z c2r rows)"),
        "deposit": (4 * n_g + 12 * n_p, 0, "k_deposit_tiles"),
        "readout_kick_drift": (12 * n_g + 72 * n_p, 0, "k_readout_kick_drift"),
        "bin_count": (12 * n_p, 0, "k_bin_count"),
        "bin_scatter": (24 * n_p, 0, "k_bin_scatter"),
        # net: bfloat16 activations of 16 channels = 32 bytes per cell
        "net_features": (4 * n_g + 2 * n_g, 0, "k_features"),
        "net_conv1": (2 * n_g + 32 * n_g, 27 * 16 * 2 * n_g, "k_conv1"),
        "net_conv_hidden": (64 * n_g, conv16, "k_conv16r<kFwdRelu>"),
        # features + layers 1-2 fused: fp32 density in, bf16 h2 out
        "net_conv12": (4 * n_g + 32 * n_g, 27 * 16 * 2 * n_g + conv16, "k_conv12r"),
        "net_conv_out_add": (32 * n_g + 24 * n_g, 27 * 16 * 3 * 2 * n_g, "k_conv_head_r<kFwdAdd>"),
        # layer 3 + last layer fused: bf16 h2 in, plain force in, corrected force out
        "net_conv34": (32 * n_g + 24 * n_g, conv16 + 27 * 16 * 3 * 2 * n_g, "k_conv3h"),
    }
    return table.get(phase)


def kernel_bytes(phase, nc, ng):
    c = kernel_cost(phase, nc, ng)
    return c[0] if c else None


def measured_traffic(kernel, nc, ng, net=False):
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
        if entry.get("nc") == nc and entry.get("ngrid") == ng and entry.get("net", False) == net:
            k = entry.get("kernels", {}).get(kernel)
            if k is not None:
                return int(k["fetch_bytes"] + k["write_bytes"])
    return None


def roofline(stats, nc, ng, net=False):
    """Roofline of the phase with the most device time: algorithmic bytes
    (and MFMA flops) per launch over the mean launch time; the binding roof
    is the larger of bytes / 8 TB/s and flops / 2.5 PFLOP/s (dense BF16)."""
    dom, (dom_ms, dom_n) = max(stats.items(), key=lambda kv: kv[1][0])
    cost = kernel_cost(dom, nc, ng)
    if cost is None or dom_n == 0:
        return None
    nbytes, flops, kname = cost
    launches = dom_n * launches_per_phase(dom, stats)
    avg_s = dom_ms / launches / 1e3
    t_hbm = nbytes / (HBM_PEAK_GBS * 1e9)
    t_mfma = flops / (MFMA_BF16_PEAK_TFLOPS * 1e12)
    out = {"bound": "hbm" if t_hbm >= t_mfma else "mfma"}
    if out["bound"] == "hbm":
        achieved = nbytes / avg_s / 1e9
        out.update({"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4)})
    else:
        achieved = flops / avg_s / 1e12
        out.update({"achieved": round(achieved, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4)})
    out.update({"traffic": measured_traffic(dom, nc, ng, net), "kernel": kname, "phase": dom,
                "avg_launch_ms": round(avg_s * 1e3, 4), "bytes_per_launch": nbytes})
    if flops:
        out["mfma_tflops"] = round(flops / avg_s / 1e12, 1)
        out["mfma_frac"] = round(flops / avg_s / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """The host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where the environment sets the job's CPU share."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


# The CPU baseline's bounded sample: one plain step of 128^3 particles on a
# 384^3 grid in 250 Mpc/h, the cell size (0.651 Mpc/h) and particle spacing
# (1.95 Mpc/h) of the plain fine production run and of the plain 256^3 / 768^3
# run, at a size the host finishes in seconds (particle-steps/s of a PM step
# at fixed cell size depends on the size only through the FFT's log N).
CPU_SAMPLE = (128, 384, 250.0)


def sample_particles(nc, ng, box, seed=1):
    """2LPT particles (seed 1, a = 0.1) of nc^3 in a box of side box, made by
    the library on the GPU (the CPU baseline times only the step)."""
    import cola_halo_amd as ch

    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    with ch.Context(nc, ng, box) as ctx:
        ctx.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        ctx.lpt_set_displacement(0.1, OM, seed)
        return ctx.get_particles()


def cpu_baseline(ctx_parts, nc, ng, box, threads):
    """CPU restatement of the reference's path (the oracle: C + OpenMP) on
    the host: one plain COLA step (wrap, PM force, kick, drift)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
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
            "cpu": cpu_model(),
            "host_cores": os.cpu_count(),
            "sample": f"CPU restatement (the oracle, C + OpenMP on {threads} threads of {os.cpu_count()} host "
                      f"cores), not the reference binary (FFTW/GSL/MPI are absent): 1 plain step (wrap + PM force "
                      f"on {ng}^3 + kick + drift) of {nc}^3 particles in a {box:g} Mpc/h box from the 2LPT state "
                      f"of seed 1, {dt:.1f} s, oracle status {st}"}


def timed_steps(ctx, nsteps, Om):
    """nsteps COLA steps cycling the 10-step schedule from the saved initial
    conditions (the tenth step of a cycle is the force at a = 1); returns the
    number of force evaluations."""
    da = 1.0 / NSTEPS
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


def time_run(ctx, steps, warmup, Om, profile=True):
    """Warm-up, back to the initial conditions, then `steps` timed steps with
    per-phase HIP events; returns (seconds, forces, phase stats)."""
    timed_steps(ctx, warmup, Om)
    ctx.synchronize()
    ctx.state_restore()
    ctx.synchronize()
    ctx.set_profiling(profile)
    ctx.reset_phase_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    forces = timed_steps(ctx, steps, Om)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    stats = ctx.phase_stats()
    ctx.set_profiling(False)
    return dt, forces, stats


def final_state(ctx, Om):
    """An untimed complete 10-step run from the saved initial conditions and
    its snapshot at a = 1."""
    ctx.state_restore()
    timed_steps(ctx, NSTEPS, Om)
    ctx.synchronize()
    return ctx.set_snapshot(1.0, Om)


def tune_mid_run(ctx, Om, steps=5):
    """Launch shapes tuned on the state after `steps` steps (clustered enough
    that the deposit sees late-time contention), then back to the ICs."""
    ctx.state_restore()
    timed_steps(ctx, steps, Om)
    ctx.synchronize()
    t = ctx.autotune(reps=3)
    ctx.state_restore()
    ctx.synchronize()
    return t


def production_bench(args):
    import cola_halo_amd as ch

    P = PROD
    nc, box = P["nc"], P["box"]
    Om = float(np.float32(OM))
    da = 1.0 / NSTEPS
    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    weights = args.weights or NET_WEIGHTS
    net = ch.ResidualNet.from_file(P["coarse"], weights)
    runs = {"corrected": ch.Context(nc, P["coarse"], box), "fine": ch.Context(nc, P["fine"], box)}
    runs["corrected"].set_residual_net(net)
    t_ic = {}
    tuning = {}
    for name, ctx in runs.items():
        ctx.power_init_table(k, p, da, 0.8, OM, 1.0 - OM)
        t = time.perf_counter()
        ctx.lpt_set_displacement(da, OM, P["seed"])
        ctx.set_times(ctx.times()[0], da)  # main.c:145
        ctx.synchronize()
        t_ic[name] = round(time.perf_counter() - t, 3)
        ctx.state_save()
        if not args.no_autotune:
            tuning[name] = tune_mid_run(ctx, Om)
    res = {}
    for name, ctx in runs.items():
        dt, forces, stats = time_run(ctx, args.steps, args.warmup, Om)
        ng = P["coarse"] if name == "corrected" else P["fine"]
        res[name] = {"seconds": dt, "forces": forces, "stats": stats, "ng": ng,
                     "value": nc ** 3 * forces / dt, "ms_per_step": dt / args.steps * 1e3}
    # P(k) gate and mass functions on complete untimed runs of every seed of
    # the comparison (the timed seed's saved ICs first, then the others')
    ll = float(np.float32(0.2 * np.float32(box) / nc))  # main.c:353
    edges = 32 * 2.0 ** np.arange(0, 9)
    mf = {name: np.zeros(len(edges) - 1, np.int64) for name in runs}
    seed_err, ratios = {}, {}
    for seed in P["pk_seeds"]:
        pk = {}
        for name, ctx in runs.items():
            if seed != P["seed"]:
                ctx.lpt_set_displacement(da, OM, seed)
                ctx.set_times(ctx.times()[0], da)
                ctx.state_save()
            final_state(ctx, Om)
            kb, pk[name], nmodes = ctx.power_spectrum(P["coarse"], 0.0, PK_KMAX, PK_BINS, deconvolve=True,
                                                      source="snapshot")
            mf[name] += np.histogram(ctx.fof_find_halos(ll, 32)["nfof"], bins=edges)[0]
        sel = kb > 0
        ratios[seed] = pk["corrected"][sel] / pk["fine"][sel]
        seed_err[seed] = float(np.abs(ratios[seed] - 1.0).max())
    err = max(seed_err.values())
    ratio = ratios[P["seed"]]
    gate = err <= PK_TOL
    cpu = None
    if not args.no_cpu_baseline:
        s_nc, s_ng, s_box = CPU_SAMPLE
        cpu = cpu_baseline(sample_particles(s_nc, s_ng, s_box), s_nc, s_ng, s_box, cpu_threads())
    for ctx in runs.values():
        ctx.close()
    net.close()
    c, f = res["corrected"], res["fine"]
    phases = {k2: round(v[0] / max(c["forces"], 1), 3) for k2, v in c["stats"].items() if v[1] > 0}
    phases_fine = {k2: round(v[0] / max(f["forces"], 1), 3) for k2, v in f["stats"].items() if v[1] > 0}
    n_c = P["coarse"] ** 2 * (P["coarse"] // 2 + 1)
    out = {
        "metric": METRIC,
        "value": round(c["value"], 1) if gate else 0.0,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (PM, particles); bf16 activations with fp32 MFMA accumulation (net)",
        "data": "synthetic (2LPT ICs of seeds 1-4, seed 1 timed; net weights trained by this project on its own "
                "force pairs)",
        "config": {"workload": f"production comparison: {nc}^3 particles, box {box:g} Mpc/h, 2LPT seed {P['seed']} "
                               f"a 0.1->1 in {NSTEPS} steps, EH98 no-wiggle P(k) sigma8 0.8; corrected "
                               f"{P['coarse']}^3 PM + learned residual vs plain {P['fine']}^3 PM",
                   "nc": nc, "coarse_grid": P["coarse"], "fine_grid": P["fine"], "boxsize": box,
                   "seed": P["seed"], "pk_seeds": list(P["pk_seeds"]), "net_weights": os.path.relpath(weights, ROOT),
                   "parallelism": "single GPU"},
        "roofline": roofline(c["stats"], nc, P["coarse"], net=True),
        "pk_gate": {"max_abs_err_k_le_1": round(err, 5), "tolerance": PK_TOL, "pass": bool(gate),
                    "max_abs_err_by_seed": {str(sd): round(e, 5) for sd, e in seed_err.items()},
                    "k": [round(float(v), 4) for v in kb[sel]],
                    "ratio_corrected_to_fine": [round(float(v), 5) for v in ratio],
                    "ratio_seed": P["seed"]},
        "mass_function_ratio_corrected_to_fine": {
            "seeds": list(P["pk_seeds"]),
            "nfof_bins": [int(v) for v in edges],
            "ratio": [round(float(a / b), 4) if b > 0 else None for a, b in zip(mf["corrected"], mf["fine"])],
            "fine_counts": [int(v) for v in mf["fine"]]},
        "fine": {"value": round(f["value"], 1), "ms_per_step": round(f["ms_per_step"], 3),
                 "roofline": roofline(f["stats"], nc, P["fine"]), "phase_ms_per_step": phases_fine},
        "speedup_vs_fine": round(c["value"] / f["value"], 4),
        "step_canonical_frac_fine": round((156 * nc ** 3 + 32 * P["fine"] ** 3 + 64 * P["fine"] ** 2 * (P["fine"] // 2 + 1))
                                          / (f["ms_per_step"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "step_canonical_frac_coarse_pm": round((156 * nc ** 3 + 32 * P["coarse"] ** 3 + 64 * n_c)
                                               / (c["ms_per_step"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "phase_ms_per_step": phases,
        "ic_seconds": t_ic,
        "tuning": tuning,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def plain_bench(args):
    """Plain COLA (no net) of one configuration: nc^3 particles on an ng^3
    grid, seed 1, 10 steps from a = 0.1."""
    import cola_halo_amd as ch

    nc, ng, box = args.nc, args.ng, args.box
    da = 1.0 / NSTEPS
    Om = float(np.float32(OM))
    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    ctx = ch.Context(nc, ng, box)
    if args.tile or args.fft_log2_tile != -2:
        tx, ty, tz = (int(v) for v in args.tile.split(",")) if args.tile else (8, 8, 32)
        ctx.set_tiling(tx, ty, tz, args.fft_log2_tile if args.fft_log2_tile != -2 else -1)
    if args.xcd_group:
        ctx.set_fft_xcd_group(args.xcd_group)
    ctx.power_init_table(k, p, da, 0.8, OM, 1.0 - OM)
    t_ic = time.perf_counter()
    ctx.lpt_set_displacement(da, OM, PLAIN["seed"])
    ctx.set_times(ctx.times()[0], da)  # main.c:145
    ctx.synchronize()
    t_ic = time.perf_counter() - t_ic
    ctx.state_save()
    tuning = None
    if not args.no_autotune and not (args.tile or args.fft_log2_tile != -2 or args.xcd_group):
        tuning = tune_mid_run(ctx, Om)
    elapsed, forces, stats = time_run(ctx, args.steps, args.warmup, Om)
    value = nc ** 3 * forces / elapsed
    n_c = ng * ng * (ng // 2 + 1)
    b_step = 156 * nc ** 3 + 32 * ng ** 3 + 64 * n_c
    step_frac = b_step / (elapsed / forces) / 1e9 / HBM_PEAK_GBS
    cpu = None
    if not args.no_cpu_baseline:
        s_nc, s_ng, s_box = CPU_SAMPLE
        cpu = cpu_baseline(sample_particles(s_nc, s_ng, s_box), s_nc, s_ng, s_box, cpu_threads())
    ctx.close()
    phases = {k2: round(v[0] / max(forces, 1), 3) for k2, v in stats.items() if v[1] > 0}
    print(json.dumps({
        "metric": "plain COLA particle-steps/s on 1 MI355X", "value": round(value, 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain COLA {nc}^3 particles, {ng}^3 PM grid, box {box:g} Mpc/h, 2LPT seed "
                               f"{PLAIN['seed']}, a 0.1->1 in {NSTEPS} steps, EH98 no-wiggle P(k) sigma8 0.8",
                   "nc": nc, "ngrid": ng, "boxsize": box, "seed": PLAIN["seed"], "parallelism": "single GPU"},
        "roofline": roofline(stats, nc, ng), "step_canonical_frac": round(step_frac, 4),
        "phase_ms_per_step": phases, "ic_seconds": round(t_ic, 3), "tuning": tuning, "cpu_baseline": cpu}))


ENSEMBLE = [(96, 288, 187.5)] * 16 + [(128, 384, 250.0)] * 16 + [(160, 480, 312.5)] * 16
# many small realizations at the same cell size, where one realization's
# launches are too small to fill the GPU
ENSEMBLE_SMALL = [(32, 96, 62.5)] * 32 + [(48, 144, 93.75)] * 32


def ensemble_bench(args, members=ENSEMBLE, label="48 uneven realizations",
                   workload="16 x 96^3/288^3 L187.5 + 16 x 128^3/384^3 L250 + 16 x 160^3/480^3 L312.5, seeds 1-48, "
                            "10 steps"):
    """Uneven realizations (by default 16 x 96^3/288^3, 16 x 128^3/384^3,
    16 x 160^3/480^3, seeds 1-48, 10 steps each): the step loops run one
    realization after another; interleaved across all contexts with
    graph-captured steps on their own streams; and as one ensemble (one
    launch per phase over every realization's workgroups), eager and as a
    replayed graph.  Prints one JSON line."""
    import cola_halo_amd as ch

    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    da = 1.0 / NSTEPS
    Om = float(np.float32(OM))
    ctxs = [ch.Context(nc, ng, box) for nc, ng, box in members]
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
    for (nc, ng, box), c in zip(members, ctxs):
        if ng not in tuned:
            tuned[ng] = c.autotune(reps=3)
        else:
            c.apply_tuning(tuned[ng])
    particles = sum(nc ** 3 for nc, _, _ in members)
    out = {}
    ens = ch.Ensemble(ctxs)
    # the ensemble builds its tables at its first step: one untimed step
    prepare(False)
    ens.step(Om, 1.5 * da, 2.0 * da)
    launches = ens.launches()
    merged = {}
    for mode in ("sequential", "batched_graph", "ensemble", "ensemble_graph", "ensemble_merged_graph"):
        if mode.startswith("ensemble"):
            # tables (re)built, the first step eager and the graph captured
            # in untimed steps
            ens.set_merge(mode == "ensemble_merged_graph")
            ens.set_graph(mode != "ensemble")
            prepare(False)
            for istep in (1, 2):
                ens.step(Om, (istep + 0.5) * da, (istep + 1.0) * da)
            if mode == "ensemble_merged_graph":
                merged["launches"] = ens.launches()
        prepare(mode == "batched_graph")
        t0 = time.perf_counter()
        if mode == "sequential":
            for c in ctxs:
                for istep in range(1, NSTEPS):
                    c.step(Om, (istep + 0.5) * da, (istep + 1.0) * da)
                c.move_particles()
                c.pm_calculate_forces()
                c.synchronize()
        elif mode == "batched_graph":
            for istep in range(1, NSTEPS):
                ch.step_batch(ctxs, Om, [(istep + 0.5) * da] * len(ctxs), [(istep + 1.0) * da] * len(ctxs))
            finish()
        else:
            for istep in range(1, NSTEPS):
                ens.step(Om, (istep + 0.5) * da, (istep + 1.0) * da)
            finish()
        dt = time.perf_counter() - t0
        out[mode] = {"seconds": round(dt, 4), "particle_steps_per_s": round(particles * NSTEPS / dt, 1)}
    ens.close()
    for c in ctxs:
        c.close()
    out["ensemble"]["launches_per_step"] = launches
    out["ensemble_graph"]["launches_per_step"] = launches
    out["ensemble_merged_graph"]["launches_per_step"] = merged.get("launches")
    best = max(out, key=lambda m: out[m]["particle_steps_per_s"])
    print(json.dumps({"metric": f"ensemble particle-steps/s, {label}", "value":
                      out[best]["particle_steps_per_s"], "best_mode": best, "unit": "particle-steps/s", "n_gpus": 1,
                      "dtype": "f32", "data": "synthetic",
                      "config": {"workload": workload, "particles": particles},
                      "modes": out, "tuning": {str(k): v for k, v in tuned.items()}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=("production", "plain"), default="production",
                    help="production: corrected 768^3 vs plain 1536^3 at 512^3 particles (the metric); "
                         "plain: plain COLA of --nc/--ng/--box")
    ap.add_argument("--weights", default="", help="net-weights file (default: the committed net)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nc", type=int, default=PLAIN["nc"])
    ap.add_argument("--ng", type=int, default=PLAIN["ng"])
    ap.add_argument("--box", type=float, default=PLAIN["box"])
    ap.add_argument("--tile", type=str, default="", help="deposit tile tx,ty,tz (plain workload)")
    ap.add_argument("--fft-log2-tile", type=int, default=-2, help="FFT transforms per workgroup (log2)")
    ap.add_argument("--xcd-group", type=int, default=0, help="FFT column tiles per XCD group")
    ap.add_argument("--ensemble", action="store_true",
                    help="time the 48-realization ensemble (sequential vs batched graph steps) instead")
    ap.add_argument("--ensemble-small", action="store_true",
                    help="time 64 small realizations (32 x 32^3/96^3, 32 x 48^3/144^3) the same way instead")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of tuning them")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the plain copy kernel on this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box here has one GPU.  Ensembles would shard by seed "
              "and one realization by x slabs across GPUs; neither runs in this benchmark.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        import cola_halo_amd as ch

        print(json.dumps({"copy_probe_mib": args.copy_probe,
                          "gbps": ch.bandwidth_probe(args.copy_probe << 20, 5)}))
        return
    if args.ensemble:
        ensemble_bench(args)
    elif args.ensemble_small:
        ensemble_bench(args, ENSEMBLE_SMALL, "64 small realizations",
                       "32 x 32^3/96^3 L62.5 + 32 x 48^3/144^3 L93.75, seeds 1-64, 10 steps")
    elif args.workload == "plain":
        plain_bench(args)
    else:
        production_bench(args)


if __name__ == "__main__":
    main()
