"""GPU checks of the learned initial model (BF16 MFMA network kernels).

  * forward and parameter gradient of the device network against the
    oracle's restatement with the same BF16 roundings (oracle/mlp.py), on a
    batch that is not a multiple of the 64-row tiles, through a gathered
    (shuffled) row subset;
  * training targets and the decode into an initial model against the
    oracle's closed forms;
  * training on pairs from the device forward lowers the validation loss, and
    the trained network's initial model, against the default one, starts the
    inversion closer to the truth and ends it at a lower median chi^2.
"""
import numpy as np
import pytest

import oracle
from oracle import mlp as omlp
from firtez_dz_amd import atmos, gpu, learned, lines

pytestmark = pytest.mark.gpu


def report(name, err, tol):
    print("[mlp] %-52s %.3e (tol %.1e)" % (name, err, tol), flush=True)
    assert err <= tol, name


def setup_net(gpu_ctx, nz=64, nh=512, nl=3, seed=5):
    import torch
    assert torch.cuda.is_available()
    ls = lines.fe6301_6302()
    z = np.arange(nz) * 15.0
    gpu_ctx.configure(ls, z)
    gpu_ctx.gain_hsra()
    nin = 4 * gpu_ctx.numw
    gpu_ctx.mlp_create(nin, nh, nl, 4, seed)
    return nin, nh, nl, 20


def random_params(gpu_ctx, nin, nh, nl, nout, rng):
    p = gpu_ctx.mlp_params().astype(np.float64)
    npar = len(p) - 2 * (nin + nout)
    p[:npar] += rng.normal(0, 0.02, npar)        # non-zero biases too
    p[npar:npar + nin] = rng.normal(0.3, 0.2, nin)
    p[npar + nin:npar + 2 * nin] = rng.uniform(0.05, 0.2, nin)
    p[npar + 2 * nin:npar + 2 * nin + nout] = rng.normal(0, 1, nout)
    p[npar + 2 * nin + nout:] = rng.uniform(0.5, 2.0, nout)
    p = p.astype(np.float32)
    gpu_ctx.mlp_set_params(p)
    return p.astype(np.float64)


def per_layer(a, b, nin, nh, nl, nout):
    dims = [nin] + [nh] * nl + [nout]
    o, out = 0, []
    for l in range(nl + 1):
        w = dims[l] * dims[l + 1]
        for sl in (slice(o, o + w), slice(o + w, o + w + dims[l + 1])):
            out.append(np.max(np.abs(a[sl] - b[sl])) / max(np.abs(b[sl]).max(), 1e-30))
        o += w + dims[l + 1]
    return out


def test_mlp_forward_vs_oracle(gpu_ctx):
    import torch
    nin, nh, nl, nout = setup_net(gpu_ctx)
    rng = np.random.default_rng(0)
    p = random_params(gpu_ctx, nin, nh, nl, nout, rng)
    x = rng.normal(0.3, 0.15, (200, nin)).astype(np.float32)
    y = gpu_ctx.mlp_forward_torch(torch.from_numpy(x).cuda()).cpu().numpy()
    ref, _ = omlp.forward(p, x, nin, nh, nl, nout, emulate=True)
    # BF16 activations: a value at a rounding boundary may round the other way (1e-3 measured)
    report("forward (200 rows) vs oracle, rel. to max|y|", np.max(np.abs(y - ref)) / np.abs(ref).max(), 5e-3)
    yd = gpu_ctx.mlp_forward_torch(torch.from_numpy(x.astype(np.float64)).cuda()).cpu().numpy()
    report("forward from float64 inputs == float32 inputs", np.max(np.abs(yd - y)), 0.0)


def test_mlp_gradient_vs_oracle(gpu_ctx):
    import torch
    nin, nh, nl, nout = setup_net(gpu_ctx)
    rng = np.random.default_rng(1)
    p = random_params(gpu_ctx, nin, nh, nl, nout, rng)
    n_all = 300
    x = rng.normal(0.3, 0.15, (n_all, nin)).astype(np.float32)
    t = rng.normal(0, 1, (n_all, nout)).astype(np.float32)
    idx = rng.permutation(n_all)[:100].astype(np.int32)
    loss, g = gpu_ctx.mlp_gradient_torch(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(),
                                         torch.from_numpy(idx).cuda())
    g = g.cpu().numpy()
    rl, rg = omlp.loss_and_grad(p, x[idx], t[idx], nin, nh, nl, nout, emulate=True)
    report("loss vs oracle (relative)", abs(float(loss.item()) - rl) / rl, 2e-6)
    errs = per_layer(g, rg, nin, nh, nl, nout)
    for i, e in enumerate(errs):
        report("gradient layer %d %s vs oracle (rel. to max)" % (i // 2, "Wb"[i % 2]), e, 1e-4)


def test_mlp_targets_and_decode_vs_oracle(gpu_ctx, cpu):
    import torch
    nin, nh, nl, nout = setup_net(gpu_ctx)
    atm, z = atmos.seeded(8, 8, 64, seed=4)
    t = gpu_ctx.mlp_targets_torch(torch.from_numpy(atm).cuda()).cpu().numpy()
    coefs = np.array(cpu.int_coefs(64, 4))
    ref = omlp.targets(atm, coefs)
    report("targets vs oracle (rel. to column range)", np.max(np.abs(t - ref) / (np.abs(ref).max(axis=0) + 1)), 1e-5)
    rng = np.random.default_rng(2)
    p = random_params(gpu_ctx, nin, nh, nl, nout, rng)
    obs = rng.normal(0.3, 0.15, (64, nin))
    a = torch.from_numpy(atm.copy()).cuda()
    gpu_ctx.learned_init_torch(torch.from_numpy(obs).cuda(), a)
    a = a.cpu().numpy()
    y, _ = omlp.forward(p, obs, nin, nh, nl, nout, emulate=True)
    npar = len(p) - 2 * (nin + nout)
    tm, ts = p[npar + 2 * nin:npar + 2 * nin + nout], p[npar + 2 * nin + nout:]
    dec = omlp.decode(y, tm, ts, coefs)
    for row, name in ((0, "T"), (7, "Bz"), (10, "vz")):
        sc = np.abs(dec[name]).max() + 1e-30
        report("decoded %s vs oracle (rel. to max)" % name, np.max(np.abs(a[row].T - dec[name])) / sc, 3e-3)
    bp = np.hypot(a[5], a[6]).T
    report("decoded B_perp vs oracle", np.max(np.abs(bp - np.hypot(dec["Bx"], dec["By"]))) / (bp.max() + 1e-30), 3e-3)
    # rows the network does not propose are left as given
    assert np.array_equal(a[1:5], atm[1:5]) and np.array_equal(a[8:10], atm[8:10]) and np.array_equal(a[11], atm[11])


def test_training_gains_over_default_init(gpu_ctx):
    """Train on 32k pairs from the device forward, then invert a 16x16 patch
    of a cube the network never saw from the learned initial model and from
    the default one (column-mean T, B = (100, 100, 100) G, v = 0)."""
    import torch
    nin, nh, nl, nout = setup_net(gpu_ctx)
    X, T = learned.training_set(gpu_ctx, 8, 64, 64, 64, seed0=500, noise=1e-3)
    rep = learned.train(gpu_ctx, X, T, epochs=12, batch=512, lr=1e-3)
    print("[mlp] train losses", np.round(rep.train_loss, 4), "valid", np.round(rep.valid_loss, 4),
          "untrained %.3f" % rep.valid_loss_initial, "%.1f s" % rep.seconds, flush=True)
    assert rep.valid_loss[-1] < 0.5 * rep.valid_loss_initial and rep.valid_loss[-1] < rep.valid_loss[0]
    nx = ny = 16
    atm, z = atmos.seeded(nx, ny, 64, seed=7)
    truth = torch.from_numpy(atm).cuda()
    assert int(gpu_ctx.hydrostatic_torch(truth).max()) == 0
    st, _, status = gpu_ctx.forward_torch(truth, gpu.RF_NONE)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    obs = (st.double().reshape(nx * ny, -1) + 1e-3 * torch.randn(st.reshape(nx * ny, -1).shape, generator=gen,
                                                                  device="cuda", dtype=torch.float64)).contiguous()
    init = torch.zeros_like(truth)
    init[0] = truth[0].mean(dim=1, keepdim=True)
    init[1, -1] = truth[1, -1].mean()
    init[5:8] = 100.0
    assert int(gpu_ctx.hydrostatic_torch(init).max()) == 0
    prop = init.clone()
    gpu_ctx.learned_init_torch(obs, prop)
    assert int(gpu_ctx.hydrostatic_torch(prop).max()) == 0

    def chi2_of(model):
        s, _, _ = gpu_ctx.forward_torch(model.clone(), gpu.RF_NONE)
        return float(((s.double().reshape(nx * ny, -1) - obs) ** 2).mean(dim=1).median()) / 1e-6

    c_plain, c_learn = chi2_of(init), chi2_of(prop)
    verr_plain = float((init[10] - truth[10]).abs().mean())
    verr_learn = float((prop[10] - truth[10]).abs().mean())
    print("[mlp] initial model: median chi2 default %.4g, learned %.4g; mean |dv_z| %.3g / %.3g cm/s"
          % (c_plain, c_learn, verr_plain, verr_learn), flush=True)
    assert c_learn < 0.5 * c_plain and verr_learn < verr_plain
    ctl = gpu.invert_control({"T": 4, "Bx": 4, "By": 4, "Bz": 4, "vz": 4}, noise=1e-3, smoothing=False)
    _, _, chi_p, rp = learned.invert(gpu_ctx, init, obs, ctl, nx, ny, None, learned=False)
    _, _, chi_l, rl = learned.invert(gpu_ctx, init, obs, ctl, nx, ny, None, learned=True)
    mp, ml = float(chi_p.median()), float(chi_l.median())
    print("[mlp] LM iterations: default %d, learned %d (final cycle only); median chi2 %.4g / %.4g" % (rp.iterations, rl.iterations,
                                                                                    mp, ml), flush=True)
    # the gain: the inversion from the proposal ends at a lower chi^2
    assert ml < mp
