"""Train and assess the learned coarse-to-fine force residual on one GPU.

Workflow (defaults: 128^3 particles, box 250 Mpc/h, coarse grid 192^3 =
factor 1.5, fine grid 384^3 = factor 3, seeds 1-56 for training and 57-64
held out, 10 COLA steps from a = 0.1):

1. force pairs: each training seed runs plain COLA on the fine grid; before
   every force evaluation the particles are copied to the coarse context and
   the pair (coarse density, fine force restricted to the coarse cells -
   weights (1/8, 3/4, 1/8) per axis by default - minus coarse force, a) is
   stored on the device;
2. training: Adam on the device's BF16 MFMA kernels, one step per pair;
3. assessment on the held-out seeds: plain fine, plain coarse and corrected
   coarse runs to z = 0; P(k) ratios to the fine run for k <= 1 h/Mpc (shell
   estimator on a common grid, CIC window divided out) and FoF mass functions
   (b = 0.2, nfof >= 32); particle-steps/s of the corrected coarse run against
   the plain fine run;
4. optionally (--production-nc 512 --production-box 1000): the trained net,
   on the same cell size, corrects coarse-grid runs of a larger box, assessed
   against plain fine-grid runs of seeds 1..4 the same way.

Usage: python -m cola_halo_amd.learned [--nc 128 --box 250 --coarse 192
       --fine 384 --ntrain 56 --nval 8 --epochs 8] [--save-weights FILE] > result.json
       python -m cola_halo_amd.learned --weights FILE ...   (no training: assess a saved net)
       python -m cola_halo_amd.learned --time-net 192   (forward timing only)
       python -m cola_halo_amd.learned --time-step 512 768 1000   (corrected steps only)
"""
import argparse
import json
import time

import numpy as np

import cola_halo_amd as ch

OM = 0.273
NSTEPS = 10


def pairs_from_run(fine, coarse, net, seed):
    da = 1.0 / NSTEPS
    Om = float(np.float32(OM))
    fine.lpt_set_displacement(da, OM, seed)
    fine.set_times(fine.times()[0], da)
    for istep in range(1, NSTEPS + 1):
        coarse.copy_particles_from(fine)
        net.add_pair(coarse, fine)
        if istep < NSTEPS:
            fine.step(Om, (istep + 0.5) * da, (istep + 1.0) * da)


def timed_run(ctx, seed):
    """Plain schedule to a = 1; returns (snapshot, seconds of the step loop)."""
    da = 1.0 / NSTEPS
    Om = float(np.float32(OM))
    ctx.lpt_set_displacement(da, OM, seed)
    ctx.set_times(ctx.times()[0], da)
    ctx.synchronize()
    t0 = time.perf_counter()
    for istep in range(1, NSTEPS):
        ctx.step(Om, (istep + 0.5) * da, (istep + 1.0) * da)
    ctx.move_particles()
    ctx.pm_calculate_forces()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    return ctx.set_snapshot(1.0, Om), dt


def assess(runs, seeds, nc, box, pk_n):
    """Runs (name, context) on each seed with graph-captured steps and launch
    shapes autotuned on the first seed's initial conditions (one untimed
    warm-up run per context captures the graphs); P(k) ratios to the "fine"
    run for k <= 1 h/Mpc on a pk_n^3 grid, FoF mass-function ratios (b = 0.2,
    nfof >= 32) and particle-steps/s of the step loops."""
    tuning = {}
    for name, ctx in runs:
        ctx.lpt_set_displacement(1.0 / NSTEPS, OM, seeds[0])
        ctx.set_times(ctx.times()[0], 1.0 / NSTEPS)
        tuning[name] = ctx.autotune(reps=3)
        ctx.set_step_graph(True)
        timed_run(ctx, seeds[0])
    kb = None
    ratios = {name: [] for name, _ in runs if name != "fine"}
    mf = {name: [] for name, _ in runs}
    times = {name: 0.0 for name, _ in runs}
    ll = float(np.float32(0.2 * np.float32(box) / nc))
    edges = 32 * 2.0 ** np.arange(0, 12)
    for seed in seeds:
        pk = {}
        for name, ctx in runs:
            snap, dt = timed_run(ctx, seed)
            del snap
            times[name] += dt
            kb, pk[name], cnt = ctx.power_spectrum(pk_n, 0.0, 1.0, 25, deconvolve=True, source="snapshot")
            halos = ctx.fof_find_halos(ll, 32)
            mf[name].append(np.histogram(halos["nfof"], bins=edges)[0])
        for name in ratios:
            ratios[name].append(pk[name] / pk["fine"])
    sel = kb > 0
    res = {}
    for name in ratios:
        r = np.mean(ratios[name], axis=0)
        res[name] = {"pk_ratio": [round(float(v), 5) for v in r[sel]],
                     "max_abs_pk_err_k_le_1": round(float(np.abs(r[sel] - 1).max()), 5),
                     "mass_function_ratio": [round(float(a / b), 4) if b > 0 else None
                                             for a, b in zip(np.sum(mf[name], axis=0), np.sum(mf["fine"], axis=0))]}
    nps = nc ** 3 * NSTEPS * len(seeds)
    out = {"k": [round(float(v), 4) for v in kb[sel]], "assessment": res,
           "mass_function_bins_nfof": [int(v) for v in edges],
           "mass_function_fine_counts": [int(v) for v in np.sum(mf["fine"], axis=0)],
           "particle_steps_per_s": {name: round(nps / t, 1) for name, t in times.items()},
           "tuning": tuning,
           "speedup_corrected_vs_fine": round(times["fine"] / times["corrected"], 3)}
    return out


def production(net, args, k, p):
    """The trained net (same cell size: box / coarse grid) on larger boxes:
    corrected runs on the coarse grid against plain runs on the fine grid."""
    nc, box = args.production_nc, args.production_box
    ng_c = nc * args.coarse // args.nc
    ng_f = nc * args.fine // args.nc
    big = ch.ResidualNet(ng_c, seed=1)
    big.set_params(net.params())
    big.set_scales(*net.scales())
    fine = ch.Context(nc, ng_f, box)
    corrected = ch.Context(nc, ng_c, box)
    for c in (fine, corrected):
        c.power_init_table(k, p, 1.0 / NSTEPS, 0.8, OM, 1.0 - OM)
    corrected.set_residual_net(big)
    res = {"nc": nc, "box": box, "coarse": ng_c, "fine": ng_f, "seeds": args.production_seeds}
    res.update(assess((("fine", fine), ("corrected", corrected)), list(range(1, args.production_seeds + 1)), nc, box,
                     args.production_pk_grid or ng_c))
    corrected.close()
    fine.close()
    big.close()
    return res


def time_net(n, reps=20):
    """Device time of the net's forward pass on an n^3 grid (random weights and
    features): ms per pass and the MFMA rate of the executed products (the
    K chunks of 32 as issued: 9 for the 8-channel layer, 14 for the others)."""
    rng = np.random.default_rng(1)
    x = np.zeros((n, n, n, 8), np.float32)
    x[..., 0] = rng.standard_normal((n, n, n))
    x[..., 1] = 0.5
    with ch.ResidualNet(n) as net:
        net.forward(x)
        ms = net.time_forward(reps)
    cells = n ** 3
    flops = cells * 16 * 32 * 2 * (9 + 3 * 14)
    return {"n": n, "forward_ms": round(ms, 4), "mfma_tflops": round(flops / (ms * 1e-3) / 1e12, 1),
            "cells_per_s": round(cells / (ms * 1e-3), 1)}


def time_step(nc, ng, box, steps=5):
    """Wall time of corrected coarse steps (eager) with a randomly initialised
    net attached: nc^3 particles on an ng^3 grid of side box, 2LPT at a = 0.1
    (seed 1), `steps` COLA steps; for kernel traces of the corrected step."""
    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    Om = float(np.float32(OM))
    with ch.ResidualNet(ng, seed=1) as net, ch.Context(nc, ng, box) as ctx:
        ctx.power_init_table(k, p, 1.0 / NSTEPS, 0.8, OM, 1.0 - OM)
        ctx.set_residual_net(net)
        ctx.lpt_set_displacement(1.0 / NSTEPS, OM, 1)
        ctx.set_times(ctx.times()[0], 1.0 / NSTEPS)
        ctx.synchronize()
        t0 = time.perf_counter()
        for istep in range(1, steps + 1):
            ctx.step(Om, (istep + 0.5) / NSTEPS, (istep + 1.0) / NSTEPS)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
    return {"nc": nc, "ngrid": ng, "box": box, "steps": steps, "ms_per_step": round(ms, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nc", type=int, default=128)
    ap.add_argument("--box", type=float, default=250.0)
    ap.add_argument("--coarse", type=int, default=192)
    ap.add_argument("--fine", type=int, default=384)
    ap.add_argument("--ntrain", type=int, default=56)
    ap.add_argument("--nval", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--pk-grid", type=int, default=0, help="power-spectrum grid (default: the fine grid)")
    ap.add_argument("--restriction", type=float, default=0.125, metavar="W",
                    help="side weight w of the fine-to-coarse restriction (w, 1-2w, w): 0 point values, 0.25 full weighting")
    ap.add_argument("--time-net", type=int, default=0, metavar="N",
                    help="only time the net's forward pass on an N^3 grid and exit")
    ap.add_argument("--time-step", type=float, nargs=3, default=None, metavar=("NC", "NG", "BOX"),
                    help="only time five corrected steps (random net) of NC^3 particles on an NG^3 grid and exit")
    ap.add_argument("--production-nc", type=int, default=0, metavar="NC",
                    help="after training, also assess the net on NC^3 particles (same cell size and particle spacing)")
    ap.add_argument("--production-box", type=float, default=1000.0)
    ap.add_argument("--production-seeds", type=int, default=4)
    ap.add_argument("--production-pk-grid", type=int, default=0)
    ap.add_argument("--save-weights", default="", metavar="FILE",
                    help="write the trained net to a net-weights file")
    ap.add_argument("--weights", default="", metavar="FILE",
                    help="skip pair generation and training: assess the net of a net-weights file")
    args = ap.parse_args()
    if args.time_net:
        print(json.dumps(time_net(args.time_net)))
        return
    if args.time_step:
        nc, ng, box = args.time_step
        print(json.dumps(time_step(int(nc), int(ng), box)))
        return
    nc, box = args.nc, args.box
    pk_n = args.pk_grid or args.fine
    k, p = ch.eh98_table(**ch.EH_COSMOLOGY)
    out = {"config": vars(args)}

    fine = ch.Context(nc, args.fine, box)
    coarse = ch.Context(nc, args.coarse, box)
    for c in (fine, coarse):
        c.power_init_table(k, p, 1.0 / NSTEPS, 0.8, OM, 1.0 - OM)
    if args.weights:
        net = ch.ResidualNet.from_file(args.coarse, args.weights)
        out["weights"] = args.weights
    else:
        net = ch.ResidualNet(args.coarse, seed=1)
        net.set_restriction(args.restriction)

        t0 = time.perf_counter()
        for seed in range(1, args.ntrain + args.nval + 1):
            pairs_from_run(fine, coarse, net, seed)
        out["pairs"] = {"count": net.num_samples(), "seconds": round(time.perf_counter() - t0, 2)}

        ntr = args.ntrain * NSTEPS
        nva = args.nval * NSTEPS
        net.fit_scale(ntr)
        val_before = net.eval(ntr, nva)
        t0 = time.perf_counter()
        losses = []
        for e in range(args.epochs):
            lr = args.lr if e < args.epochs // 2 else args.lr * 0.3
            losses.append(float(net.train(ntr, 1, lr=lr, seed=e + 1)[0]))
        out["training"] = {"epoch_loss": [round(v, 5) for v in losses],
                           "seconds": round(time.perf_counter() - t0, 2),
                           "val_loss": round(net.eval(ntr, nva), 5), "val_loss_untrained": round(val_before, 5),
                           "out_scale": net.scales()[1],
                           "loss_of_zero_correction": 1.0}
        if args.save_weights:
            net.save(args.save_weights)
            out["saved_weights"] = args.save_weights

    # assessment on the held-out seeds
    corrected = ch.Context(nc, args.coarse, box)
    corrected.power_init_table(k, p, 1.0 / NSTEPS, 0.8, OM, 1.0 - OM)
    corrected.set_residual_net(net)
    runs = (("fine", fine), ("coarse", coarse), ("corrected", corrected))
    out.update(assess(runs, list(range(args.ntrain + 1, args.ntrain + args.nval + 1)), nc, box, pk_n))
    corrected.close()
    fine.close()
    coarse.close()
    if args.production_nc:
        out["production"] = production(net, args, k, p)
    net.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
