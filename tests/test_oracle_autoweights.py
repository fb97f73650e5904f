"""The oracle's STOKES SETUP AUTO weights (oracle.invert.auto_weights, after
SET_WEIGHTS1D get_dmodel.f90:2191-2224 and SET_ISIGMAP :356-373) pinned by
what the definition fixes independently of the restatement: the weights are
inversely proportional to each block's sample standard deviation (numpy's
ddof=1 std), capped at ten times the pixel's smallest, with a root mean square
of one over the inverted blocks, laid out block by block in I, Q, U, V order
and divided by ISIGMA**2."""
import numpy as np

from oracle import invert

NOISE = 1e-3


def _unit(isig):
    """WSTK of every block back from ISIGMAP = WSTK**2 / ISIGMA**2."""
    return np.sqrt(isig * float(np.float32(NOISE)) ** 2)


def test_weights_follow_inverse_std_and_unit_rms():
    rng = np.random.default_rng(3)
    numw, invstk = 40, (1, 1, 1, 1)
    scales = np.array([0.2, 0.05, 0.08, 0.12])
    obs = np.concatenate([rng.normal(1.0 if s == 0 else 0.0, scales[s], numw) for s in range(4)])[None, :]
    w = _unit(invert.auto_weights(obs, invstk, NOISE, numw))[0].reshape(4, numw)
    assert np.all(w == w[:, :1])                       # one value per Stokes block
    inv_std = 1.0 / np.array([np.std(obs[0, b * numw:(b + 1) * numw], ddof=1) for b in range(4)])
    ratio = w[:, 0] / inv_std
    np.testing.assert_allclose(ratio / ratio[0], 1.0, rtol=1e-6)
    np.testing.assert_allclose(np.sqrt(np.mean(w[:, 0] ** 2)), 1.0, rtol=1e-6)


def test_weights_capped_at_ten_times_the_smallest():
    numw = 8
    alt = np.array([1.0, -1.0] * (numw // 2))
    # I varies by 1, V by 1e-3: raw weights 1 and 1000 -> V capped at 10
    obs = np.concatenate([5.0 + alt, 1e-3 * alt])[None, :]
    w = _unit(invert.auto_weights(obs, (1, 0, 0, 1), NOISE, numw))[0]
    wi, wv = w[0], w[numw]
    np.testing.assert_allclose(wv / wi, 10.0, rtol=1e-6)
    # rms over the two inverted blocks is 1: wi^2 + wv^2 = 2
    np.testing.assert_allclose(wi * wi + wv * wv, 2.0, rtol=1e-6)
    np.testing.assert_allclose(wi, np.sqrt(2.0 / 101.0), rtol=1e-6)


def test_single_block_and_noise_scaling():
    """One inverted block: WSTK = 1 whatever the data; ISIGMAP = 1/ISIGMA^2
    with ISIGMA the REAL(4) noise."""
    rng = np.random.default_rng(5)
    obs = rng.normal(0.0, 0.3, (3, 16))
    isig = invert.auto_weights(obs, (0, 0, 1, 0), 4e-3, 16)
    np.testing.assert_allclose(isig, 1.0 / float(np.float32(4e-3)) ** 2, rtol=1e-7)


def test_constant_block_takes_the_cap():
    """A block of equal values has an infinite weight before the cap, so it
    ends at ten times the other block's."""
    numw = 6
    obs = np.concatenate([np.full(numw, 0.9), np.linspace(-0.1, 0.1, numw)])[None, :]
    w = _unit(invert.auto_weights(obs, (1, 1, 0, 0), NOISE, numw))[0]
    np.testing.assert_allclose(w[0] / w[numw], 10.0, rtol=1e-6)
