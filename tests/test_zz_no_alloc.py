"""fz_get_perturbation_dev_noalloc (the fixed-scratch variant of
fz_get_perturbation_dev) through the C ABI: three calls in a row on
shrinking systems (300 and 120 parameters with the solver's work arrays in
the fixed array, then 40 in LDS), so each later call finds the previous
call's values in the array; each is equal bit for bit to
fz_get_perturbation_dev on the same inputs; more than 512 parameters is an
argument error.  Named to run after the other GPU tests."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _system(np_, nf, seed):
    rng = np.random.default_rng(seed)
    jac = rng.normal(size=(nf, np_)) * np.exp(rng.uniform(-1.5, 1.5, np_))
    diff = rng.normal(size=nf) * 0.1
    pen = rng.normal(size=np_) * 0.01
    L = np.eye(np_) - np.eye(np_, k=-1)
    return jac, diff, pen, 0.05 * (L.T @ L)


@pytest.mark.parametrize("dtype", [1, 2])
def test_noalloc_matches_allocating_entry(dtype):
    import torch
    from firtez_dz_amd import gpu
    ctx = gpu.Context(0)
    try:
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        for np_, nf, lam in ((300, 2000, 0.3), (120, 900, 0.5), (40, 700, 1.0)):
            jac, diff, pen, hss = _system(np_, nf, seed=np_ + dtype)
            args = (t(jac.T), t(diff), lam)
            ref = torch.zeros(np_, dtype=torch.float64, device="cuda")
            out = torch.full((np_,), float("nan"), dtype=torch.float64, device="cuda")
            ctx.get_perturbation_noalloc_torch(*args, out, t(pen), t(hss.T), dtype=dtype, svdtol=1e-4)
            torch.cuda.synchronize()
            ctx.get_perturbation_torch(*args, ref, t(pen), t(hss.T), dtype=dtype, svdtol=1e-4)
            torch.cuda.synchronize()
            o, r = out.cpu().numpy(), ref.cpu().numpy()
            print("[noalloc] dtype %d np=%d nf=%d: max |delta| %.3e, identical %s"
                  % (dtype, np_, nf, np.abs(r).max(), np.array_equal(o, r)), flush=True)
            assert np.all(np.isfinite(o)) and np.array_equal(o, r)
        jac, diff, _, _ = _system(513, 600, seed=1)
        with pytest.raises(gpu.FzError, match="at most 512"):
            ctx.get_perturbation_noalloc_torch(t(jac.T), t(diff), 1.0, torch.zeros(513, dtype=torch.float64,
                                                                                  device="cuda"))
    finally:
        ctx.close()
