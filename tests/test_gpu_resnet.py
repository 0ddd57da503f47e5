"""The learned force residual's BF16 MFMA kernels against the double-precision
oracle net (tests/test_net_cpu.py pins the oracle to finite differences).

Tolerances: the oracle rounds the input, the weights and each stored
activation to bfloat16 exactly as the device does, so the forward output
differs only by fp32 vs double accumulation and the occasional bfloat16
rounding flip of an activation: within 1e-2 of the output's maximum.  The
device also stores the layer gradients in bfloat16: each parameter block's
gradient is within 5e-2 (relative 2-norm) of the oracle's."""
import numpy as np
import pytest

import cola_halo_amd as ch
import oracle_lib as orc

pytestmark = pytest.mark.gpu

OM = 0.273


def _blocks():
    off = 0
    out = []
    for ci, co in zip(orc.NET_CIN, orc.NET_COUT):
        out.append(("W", off, off + 27 * ci * co))
        off += 27 * ci * co
        out.append(("b", off, off + co))
        off += co
    return out


def _setup(n, seed):
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.15, orc.NET_NPARAM).astype(np.float32)
    x = np.zeros((n, n, n, 8), np.float32)
    x[..., 0] = rng.standard_normal((n, n, n))
    x[..., 1] = 0.5
    x[..., 2:] = 0.3 * rng.standard_normal((n, n, n, 6))
    return params, x


def test_nparam_matches_oracle():
    assert ch.ResidualNet.nparam() == orc.NET_NPARAM


@pytest.mark.parametrize("n", [32, 64])
def test_forward_matches_oracle(n):
    print(f"net forward n={n}", flush=True)
    params, x = _setup(n, n)
    with ch.ResidualNet(n) as net:
        net.set_params(params)
        y = net.forward(x)
    ref, _ = orc.net_forward(params.astype(np.float64), x.astype(np.float64), round_bf16=True)
    err = np.abs(y - ref).max() / np.abs(ref).max()
    assert err <= 1e-2, err


@pytest.mark.parametrize("n", [32, 64])
def test_feature_path_matches_general_path(n):
    """The inference path (first layer on the feature channel, a folded into
    its bias) against the 8-channel path and the oracle on an input of that
    form (channel 1 = a everywhere, channels 2-7 zero).  Only the fp32
    summation order differs, so outputs agree far inside the oracle
    tolerance; a wrong tap, channel or bias term would not."""
    print(f"net feature path n={n}", flush=True)
    params, x = _setup(n, 100 + n)
    x[..., 1] = 0.37
    x[..., 2:] = 0.0
    with ch.ResidualNet(n) as net:
        net.set_params(params)
        y8 = net.forward(x)
        y1 = net.forward_features(x)
        ms = net.time_forward(2)
    assert ms > 0.0
    scale = np.abs(y8).max()
    assert np.abs(y1 - y8).max() <= 2e-3 * scale, np.abs(y1 - y8).max() / scale
    ref, _ = orc.net_forward(params.astype(np.float64), x.astype(np.float64), round_bf16=True)
    assert np.abs(y1 - ref).max() <= 1e-2 * np.abs(ref).max()


@pytest.mark.parametrize("n", [32, 64, 96])
def test_plane_orders_agree(n):
    """The hidden layers by input planes (k_conv16r: one march of 1, 1 and 4
    output planes per unit at n = 32, 64, 96) against by output planes
    (k_conv16) and the oracle, on both inference paths.  Only the fp32
    order of the same products differs (and a bfloat16 rounding flip of an
    activation where it lands on a tie), so the two agree far inside the
    oracle tolerance; a wrong tap, plane offset or dropped plane would not."""
    print(f"net plane orders n={n}", flush=True)
    params, x = _setup(n, 300 + n)
    x[..., 1] = 0.41
    x[..., 2:] = 0.0
    out = {}
    with ch.ResidualNet(n) as net:
        net.set_params(params)
        for order in (False, True):
            net.set_plane_order(order)
            out[order] = (net.forward(x), net.forward_features(x))
    ref, _ = orc.net_forward(params.astype(np.float64), x.astype(np.float64), round_bf16=True)
    scale = np.abs(ref).max()
    for i in range(2):
        assert np.abs(out[True][i] - out[False][i]).max() <= 2e-3 * scale, (i, np.abs(out[True][i] - out[False][i]).max() / scale)
        assert np.abs(out[True][i] - ref).max() <= 1e-2 * scale


def test_gradient_matches_oracle():
    print("net gradient n=32", flush=True)
    n = 32
    params, x = _setup(n, 7)
    rng = np.random.default_rng(8)
    target = rng.standard_normal((n, n, n, 3)).astype(np.float32)
    with ch.ResidualNet(n) as net:
        net.set_params(params)
        net.set_scales(1.0, 1.0)
        loss, g = net.gradient(x, target)
    ref_loss, ref = orc.net_gradient(params.astype(np.float64), x.astype(np.float64), target.astype(np.float64),
                                     round_bf16=True)
    assert loss == pytest.approx(ref_loss, rel=1e-2)
    for kind, a, b in _blocks():
        rel = np.linalg.norm(g[a:b] - ref[a:b]) / np.linalg.norm(ref[a:b])
        assert rel <= 5e-2, (kind, a, rel)


def test_adam_steps_follow_the_published_update():
    """Two Adam steps (Kingma & Ba 2015, Algorithm 1: beta1 0.9, beta2 0.999,
    eps 1e-8, moments m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, bias
    corrected by 1 - beta^t) on the device's gradients, against the update
    computed here in float64 from the same gradients.  The first step alone
    would be lr sign(g) for any bias-corrected variant; the second pins the
    moment recursions and the t = 2 corrections (a missing correction, a
    swapped beta or a square root in the wrong place moves parameters by far
    more than the 1e-6 lr allowed)."""
    print("net adam n=32", flush=True)
    n, lr = 32, 1e-3
    params, x = _setup(n, 11)
    rng = np.random.default_rng(12)
    t1 = rng.standard_normal((n, n, n, 3)).astype(np.float32)
    t2 = rng.standard_normal((n, n, n, 3)).astype(np.float32)
    b1, b2, eps = 0.9, 0.999, 1e-8
    with ch.ResidualNet(n) as net:
        net.set_params(params)
        net.set_scales(1.0, 1.0)
        p0 = net.params().astype(np.float64)
        _, g1 = net.gradient(x, t1)
        net.adam_step(lr)
        p1 = net.params().astype(np.float64)
        _, g2 = net.gradient(x, t2)
        net.adam_step(lr)
        p2 = net.params().astype(np.float64)
    g1 = g1.astype(np.float64)
    g2 = g2.astype(np.float64)
    m = (1 - b1) * g1
    v = (1 - b2) * g1 ** 2
    want1 = p0 - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    m = b1 * m + (1 - b1) * g2
    v = b2 * v + (1 - b2) * g2 ** 2
    want2 = p1 - lr * (m / (1 - b1 ** 2)) / (np.sqrt(v / (1 - b2 ** 2)) + eps)
    live = np.abs(g1) > 1e-6  # parameters whose step is not dominated by eps
    assert live.sum() > 0.9 * len(g1)
    assert np.abs(p1 - want1)[live].max() <= 1e-6 * lr + 1e-7 * np.abs(p0).max()
    assert np.abs(p2 - want2)[live].max() <= 1e-6 * lr + 1e-7 * np.abs(p0).max()


def test_pairs_training_and_corrected_force(eh_table):
    """Force pairs from a 32^3 run (coarse 64^3, fine 128^3): training lowers
    the held-out loss below 1 (the loss of predicting no correction); an
    attached net changes the coarse force, deterministically, and detaching it
    restores the plain force bit for bit."""
    print("net pairs + training", flush=True)
    k, p = eh_table
    nc, box = 32, 100.0
    with ch.Context(nc, 64, box) as coarse, ch.Context(nc, 128, box) as fine, ch.ResidualNet(64, seed=3) as net:
        for c in (coarse, fine):
            c.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        for seed in (1, 2, 3):
            fine.lpt_set_displacement(0.1, OM, seed)
            fine.set_times(fine.times()[0], 0.1)
            for istep in range(1, 10):
                coarse.copy_particles_from(fine)
                net.add_pair(coarse, fine)
                fine.step(float(np.float32(OM)), (istep + 0.5) * 0.1, (istep + 1.0) * 0.1)
        assert net.num_samples() == 27
        net.fit_scale(18)
        before = net.eval(18, 9)
        assert before == pytest.approx(1.0, rel=0.5)  # targets normalised by the training RMS
        losses = net.train(18, 6, lr=2e-3)
        after = net.eval(18, 9)
        assert losses[-1] < losses[0]
        assert after < before and after < 1.0, (before, after, losses)
        # attached net: force grids change by out_scale * y
        coarse.copy_particles_from(fine)
        coarse.move_particles()
        coarse.pm_calculate_forces()
        f_plain = coarse.get_forces()
        g_plain = np.stack([coarse.get_force_grid(ax)[:, :, :64] for ax in range(3)], axis=-1)
        coarse.pm_assign()
        dens = coarse.get_density()[:, :, :64]
        a = coarse.times()[0]
        coarse.set_residual_net(net)
        coarse.pm_calculate_forces()
        f_net = coarse.get_forces()
        g_net = np.stack([coarse.get_force_grid(ax)[:, :, :64] for ax in range(3)], axis=-1)
        coarse.pm_calculate_forces()
        assert np.array_equal(f_net, coarse.get_forces())  # deterministic
        assert not np.array_equal(f_plain, f_net)
        # the grid correction is out_scale * net(feature of the density, a)
        in_scale, out_scale = net.scales()
        x = np.zeros((64, 64, 64, 8), np.float32)
        x[..., 0] = np.arcsinh(dens.astype(np.float64)) * in_scale
        x[..., 1] = a
        y = net.forward(x)
        corr = (g_net.astype(np.float64) - g_plain) / out_scale
        assert np.abs(corr - y).max() <= 1e-2 * np.abs(y).max(), np.abs(corr - y).max() / np.abs(y).max()
        # the particles read the corrected grids (the oracle's CIC readout)
        # less the correction's mean over the grid: a periodic force has none
        parts = coarse.get_particles()
        for ax in range(3):
            mean = (g_net[..., ax].astype(np.float64) - g_plain[..., ax]).mean()
            want = orc.pm_readout(64, box, parts, g_net[..., ax]).astype(np.float64) - mean
            scale = np.abs(want).max()
            assert np.abs(f_net[:, ax] - want).max() <= 1e-5 * scale, (ax, mean, np.abs(f_net[:, ax] - want).max())
        coarse.set_residual_net(None)
        coarse.pm_calculate_forces()
        assert np.array_equal(f_plain, coarse.get_forces())


def _bf16(x):
    """Round float32 values to bfloat16 (nearest even), as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def _restrict(f, w):
    """Fine grid f [2n]^3 -> coarse cells: weights (w, 1-2w, w) per axis
    around the coincident fine cell 2i, periodic (numpy, float64)."""
    f = f.astype(np.float64)
    wt = (w, 1.0 - 2.0 * w, w)
    out = 0.0
    for a in range(3):
        for b in range(3):
            for c in range(3):
                g = np.roll(f, (1 - a, 1 - b, 1 - c), axis=(0, 1, 2))
                out = out + wt[a] * wt[b] * wt[c] * g[::2, ::2, ::2]
    return out


@pytest.mark.parametrize("w", [0.0, 0.125, 0.25])
def test_pair_targets_match_restricted_forces(eh_table, w):
    """Force-pair generator: the stored target is the fine force restricted to
    the coarse cells (numpy restriction of the fine grids the library left
    behind) minus the coarse force, the feature is asinh(delta) * in_scale of
    the coarse density, both rounded to bfloat16."""
    print(f"pair targets w={w}", flush=True)
    k, p = eh_table
    nc, box = 32, 100.0
    with ch.Context(nc, 64, box) as coarse, ch.Context(nc, 128, box) as fine, ch.ResidualNet(64, seed=3) as net:
        for c in (coarse, fine):
            c.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        fine.lpt_set_displacement(0.1, OM, 5)
        fine.set_times(fine.times()[0], 0.1)
        for istep in range(1, 4):
            fine.step(float(np.float32(OM)), (istep + 0.5) * 0.1, (istep + 1.0) * 0.1)
        coarse.copy_particles_from(fine)
        net.set_restriction(w)
        net.add_pair(coarse, fine)
        fc = [coarse.get_force_grid(a)[:, :, :64] for a in range(3)]
        ff = [fine.get_force_grid(a)[:, :, :128] for a in range(3)]
        coarse.pm_assign()
        delta = coarse.get_density()[:, :, :64]
        feat, target, a = net.sample(0)
        in_scale = net.scales()[0]
    for ax in range(3):
        want = _restrict(ff[ax], w) - fc[ax].astype(np.float64)
        scale = np.abs(want).max()
        err = np.abs(target[ax] - _bf16(want.astype(np.float32))).max()
        # float32 sums on the device vs float64 here can flip a bfloat16 rounding
        assert err <= scale * 2.0 ** -8, (ax, err, scale)
    want_f = _bf16((np.arcsinh(delta.astype(np.float64)) * in_scale).astype(np.float32))
    assert np.abs(feat - want_f).max() <= np.abs(want_f).max() * 2.0 ** -8
    assert 0.39 < a < 0.41  # a_x of the coarse context after three steps from 0.1


def test_corrected_graph_steps_equal_eager_every_replay(eh_table):
    """Corrected coarse steps (net attached: feature pass, first layers,
    force FFTs, last layer added into the force grids) captured as a graph
    and replayed every step, with synchronising host reads in between, equal
    the eager steps bit for bit; the scale factor reaches the net through
    the step's device struct, so every replay sees its own a."""
    print("net graph replay 32^3 / 64^3", flush=True)
    k, p = eh_table
    nc, ng, box = 32, 64, 100.0
    # one net per context: a net's activation buffers serve the context it
    # is attached to; the two nets hold the same parameters
    with ch.ResidualNet(ng, seed=5) as net_a, ch.ResidualNet(ng, seed=5) as net_b, \
            ch.Context(nc, ng, box) as a, ch.Context(nc, ng, box) as b:
        net_b.set_params(net_a.params())
        for c, net in ((a, net_a), (b, net_b)):
            net.set_scales(1.0, 0.5)
            c.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
            c.set_residual_net(net)
            c.lpt_set_displacement(0.1, OM, 21)
            c.set_times(c.times()[0], 0.1)
        with pytest.raises(ch.ColaError) as e:  # a second context may not share net_a
            b.set_residual_net(net_a)
        assert e.value.code == 9004
        b.set_step_graph(True)
        omf = float(np.float32(OM))
        for istep in range(1, 10):
            avel1, apos1 = (istep + 0.5) * 0.1, (istep + 1.0) * 0.1
            a.step(omf, avel1, apos1)
            b.step(omf, avel1, apos1)
            pa, pb = a.get_particles(), b.get_particles()
            assert np.array_equal(pa, pb), istep
        b.set_step_graph(False)
        for c in (a, b):
            c.move_particles()
            c.pm_calculate_forces()
        assert np.array_equal(a.get_forces(), b.get_forces())


def test_destroying_an_attached_net_detaches_it(eh_table):
    """cola_resnet_destroy on a net still attached to a context detaches it
    and drops the context's captured step graph: from the same state, the
    context's next graph steps equal those of a context that never had a
    net."""
    print("net destroy while attached 16^3 / 32^3", flush=True)
    k, p = eh_table
    nc, ng, box = 16, 32, 50.0
    omf = float(np.float32(OM))
    with ch.Context(nc, ng, box) as a, ch.Context(nc, ng, box) as b:
        for c in (a, b):
            c.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
            c.lpt_set_displacement(0.1, OM, 4)
            c.set_times(c.times()[0], 0.1)
            c.set_step_graph(True)
        start = b.get_particles()
        net = ch.ResidualNet(ng, seed=2)
        net.set_scales(1.0, 0.5)
        a.set_residual_net(net)
        a.step(omf, 0.15, 0.2)  # captures a step graph with the net in it
        net.close()
        a.set_particles(start, 0.1, 0.1)
        for istep in range(1, 4):
            for c in (a, b):
                c.step(omf, (istep + 0.5) * 0.1, (istep + 1.0) * 0.1)
        assert np.array_equal(a.get_particles(), b.get_particles())


@pytest.mark.parametrize("order", [False, True])
@pytest.mark.parametrize("nc,ng,box", [(32, 64, 100.0), (256, 768, 500.0)])
def test_fused_first_layers_equal_separate_launches(eh_table, nc, ng, box, order):
    """The inference path's fused pass (features, layer 1 and layer 2 from
    the density grid in one launch) gives the corrected forces of the three
    separate launches bit for bit: at 64^3 every workgroup marches one x
    plane, at 768^3 hundreds of planes (ring slots reused many times)."""
    print(f"fused layers 1-2, {nc}^3 / {ng}^3, by input planes {order}", flush=True)
    k, p = eh_table
    with ch.ResidualNet(ng, seed=9) as net, ch.Context(nc, ng, box) as c:
        net.set_scales(1.3, 0.7)
        net.set_plane_order(order)  # k_conv12 / k_conv12r sum layer 2 as k_conv16 / k_conv16r
        c.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        c.lpt_set_displacement(0.1, OM, 4)
        c.set_times(c.times()[0], 0.1)
        omf = float(np.float32(OM))
        for istep in range(1, 4):
            c.step(omf, (istep + 0.5) * 0.1, (istep + 1.0) * 0.1)
        c.move_particles()
        c.pm_calculate_forces()
        plain = c.get_forces()
        c.set_residual_net(net)
        c.pm_calculate_forces()
        fused = c.get_forces()
        net.set_fusion(False)
        c.pm_calculate_forces()
        separate = c.get_forces()
        net.set_fusion(True)
        c.pm_calculate_forces()
        again = c.get_forces()
    assert not np.array_equal(plain, fused)
    assert np.array_equal(fused, separate)
    assert np.array_equal(fused, again)


@pytest.mark.parametrize("nc,ng,box", [(32, 64, 100.0), (48, 96, 150.0), (256, 768, 500.0)])
def test_fused_last_layers_equal_separate_launches(eh_table, nc, ng, box):
    """Layer 3 and the last layer in one launch (h3 kept in LDS, z blocks of
    30 cells with a ragged last block at every size here) give the corrected
    force grid of the two separate launches bit for bit; the particles'
    forces agree up to the rounding of the residual's mean (its double sums
    run in another order)."""
    print(f"fused layers 3-4, {nc}^3 / {ng}^3", flush=True)
    k, p = eh_table
    with ch.ResidualNet(ng, seed=9) as net, ch.Context(nc, ng, box) as c:
        net.set_scales(1.3, 0.7)
        c.power_init_table(k, p, 0.1, 0.8, OM, 1.0 - OM)
        c.lpt_set_displacement(0.1, OM, 4)
        c.set_times(c.times()[0], 0.1)
        omf = float(np.float32(OM))
        for istep in range(1, 4):
            c.step(omf, (istep + 0.5) * 0.1, (istep + 1.0) * 0.1)
        c.move_particles()
        c.set_residual_net(net)
        grids = {}
        forces = {}
        for fused in (True, False, True):
            net.set_head_fusion(fused)
            c.pm_calculate_forces()
            grids.setdefault(fused, []).append(np.stack([c.get_force_grid(a)[:, :, :ng] for a in range(3)]))
            forces.setdefault(fused, []).append(c.get_forces())
    assert np.array_equal(grids[True][0], grids[False][0])
    assert np.array_equal(grids[True][0], grids[True][1])
    scale = np.abs(forces[False][0]).max()
    assert np.abs(forces[True][0] - forces[False][0]).max() <= 1e-6 * scale
    assert np.array_equal(forces[True][0], forces[True][1])
