"""GET_PERTURBATION and HESSIAN (lm.f90:30-39, 137-204) through the C ABI
(fz_get_perturbation, fz_get_perturbation_dev, fz_hessian) against the
oracle's FP64 restatement and numpy.

  * HESSIAN: J J^T of random and structured Jacobians vs numpy (the FP64
    matrix-core contraction against a sequential sum: a few ulp of the
    largest entry);
  * GET_PERTURBATION, truncated pseudo-inverse (DTYPE 1): with and without
    the regularisation terms, small (work arrays in LDS) and large (device
    memory) systems, against the oracle and against numpy's pinv with the
    same relative cutoff; a rank-deficient system (two equal columns) gets
    the minimum-norm step, equal on both columns;
  * CG (DTYPE 2) against the oracle's iteration;
  * the device-array entry point equals the host-array one bit for bit;
  * a DTYPE other than 1 or 2 is an argument error (the reference stops).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def report(name, err, tol):
    print("[pert] %-58s max err %.3e (tol %.1e)" % (name, err, tol), flush=True)
    assert err <= tol, name


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.abs(b).max())


def system(np_, nf, seed, regul=False, dup=None):
    rng = np.random.default_rng(seed)
    # columns of different scales and some correlation, like normalised response functions
    base = rng.normal(size=(nf, np_)) * np.exp(rng.uniform(-1.5, 1.5, np_))
    jac = base + 0.3 * base[:, [0]]
    if dup is not None:
        jac[:, dup[1]] = jac[:, dup[0]]
    diff = rng.normal(size=nf) * 0.1
    pen = hss = None
    if regul:
        pen = rng.normal(size=np_) * 0.01
        L = np.eye(np_) - np.eye(np_, k=-1)
        hss = 0.05 * (L.T @ L)
    return jac, diff, pen, hss


def numpy_pinv_step(jac, diff, lam, pen, hss, svdtol):
    H = jac.T @ jac + (0 if hss is None else hss)
    H = H + np.diag(np.diag(H) * lam)
    g = jac.T @ diff - (0 if pen is None else pen)
    return np.linalg.pinv(H, rcond=svdtol) @ g


@pytest.mark.parametrize("np_,nf", [(20, 448), (80, 1200), (300, 4480)])
def test_hessian_vs_numpy(gpu_ctx, np_, nf):
    rng = np.random.default_rng(np_)
    J = rng.normal(size=(np_, nf)) * np.exp(rng.uniform(-3, 3, np_))[:, None]
    H = gpu_ctx.hessian(J)
    ref = J @ J.T
    assert np.array_equal(H, H.T)
    report("HESSIAN %dx%d vs numpy (relative to max)" % (np_, nf), rel(H, ref), 2e-14)
    # each entry against its own sequential FP64 sum: a few ulp of sum |products|
    scale = np.abs(J) @ np.abs(J).T
    report("HESSIAN %dx%d per entry / sum|products|" % (np_, nf), float(np.max(np.abs(H - ref) / scale)), 1e-13)


CASES = [(20, 448, False, 1.0), (20, 448, True, 0.1), (6, 448, True, 10.0), (64, 900, True, 0.01),
         (80, 1200, False, 0.3), (150, 2000, True, 0.1)]


@pytest.mark.parametrize("np_,nf,regul,lam", CASES)
def test_perturbation_svd_vs_oracle(gpu_ctx, cpu, np_, nf, regul, lam):
    jac, diff, pen, hss = system(np_, nf, seed=np_ + nf, regul=regul)
    d_gpu = gpu_ctx.get_perturbation(jac, diff, lam, pen, hss, dtype=1, svdtol=1e-4)
    d_ora = cpu.get_perturbation(jac, diff, lam, pen, hss, dtype=1, svdtol=1e-4)
    d_np = numpy_pinv_step(jac, diff, lam, pen, hss, 1e-4)
    report("DTYPE 1 np=%d nf=%d regul=%d vs oracle" % (np_, nf, regul), rel(d_gpu, d_ora), 1e-9)
    report("DTYPE 1 np=%d nf=%d regul=%d oracle vs numpy pinv" % (np_, nf, regul), rel(d_ora, d_np), 1e-9)


def test_perturbation_rank_deficient(gpu_ctx, cpu):
    """Two equal columns: the damped matrix is singular only at lambda = 0,
    so the truncated pseudo-inverse at lambda = 0 drops the null direction
    and splits the step equally between the two columns."""
    jac, diff, _, _ = system(12, 300, seed=5, dup=(3, 7))
    d_gpu = gpu_ctx.get_perturbation(jac, diff, 0.0, dtype=1, svdtol=1e-4)
    d_ora = cpu.get_perturbation(jac, diff, 0.0, dtype=1, svdtol=1e-4)
    d_np = numpy_pinv_step(jac, diff, 0.0, None, None, 1e-4)
    report("rank-deficient vs numpy pinv", rel(d_gpu, d_np), 1e-8)
    report("rank-deficient vs oracle", rel(d_gpu, d_ora), 1e-8)
    assert abs(d_gpu[3] - d_gpu[7]) <= 1e-9 * np.abs(d_gpu).max()


@pytest.mark.parametrize("np_,nf,regul", [(20, 448, False), (20, 448, True), (100, 1500, True)])
def test_perturbation_cg_vs_oracle(gpu_ctx, cpu, np_, nf, regul):
    """CG stops at a mean squared relative residual of 1e-3, so its iterate
    follows the rounding of every sum along the iteration path.  The
    device's dot products are trees of partial sums; it is held to ten times
    the spread of the oracle itself when only the order of its sums over
    the spectral rows changes (1e-7 on twenty parameters, 2e-5 on a
    hundred), at least 1e-7."""
    jac, diff, pen, hss = system(np_, nf, seed=3 * np_, regul=regul)
    d_gpu = gpu_ctx.get_perturbation(jac, diff, 1.0, pen, hss, dtype=2)
    d_ora = cpu.get_perturbation(jac, diff, 1.0, pen, hss, dtype=2)
    p = np.random.default_rng(0).permutation(nf)
    spread = rel(cpu.get_perturbation(jac[p], diff[p], 1.0, pen, hss, dtype=2), d_ora)
    report("DTYPE 2 np=%d nf=%d regul=%d vs oracle (spread %.1e)" % (np_, nf, regul, spread), rel(d_gpu, d_ora),
           max(1e-7, 10 * spread))
    # and the iterate is a CG solution of the damped system to its own stopping rule
    d_np = numpy_pinv_step(jac, diff, 1.0, pen, hss, 1e-12)
    report("DTYPE 2 np=%d nf=%d regul=%d vs exact solve" % (np_, nf, regul), rel(d_gpu, d_np), 0.05)


def test_perturbation_device_entry_matches_host(gpu_ctx):
    import torch
    jac, diff, pen, hss = system(40, 700, seed=11, regul=True)
    d_host = gpu_ctx.get_perturbation(jac, diff, 0.5, pen, hss, dtype=1, svdtol=1e-4)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    delta = torch.zeros(40, dtype=torch.float64, device="cuda")
    gpu_ctx.get_perturbation_torch(t(jac.T), t(diff), 0.5, delta, t(pen), t(hss.T), dtype=1, svdtol=1e-4)
    torch.cuda.synchronize()
    assert np.array_equal(delta.cpu().numpy(), d_host)


def test_perturbation_rejects_bad_dtype(gpu_ctx):
    from firtez_dz_amd import gpu
    jac, diff, _, _ = system(4, 50, seed=1)
    with pytest.raises(gpu.FzError, match="DTYPE 3"):
        gpu_ctx.get_perturbation(jac, diff, 1.0, dtype=3)
    with pytest.raises(gpu.FzError):
        gpu_ctx.get_perturbation(np.zeros((50, 0)), diff, 1.0)
