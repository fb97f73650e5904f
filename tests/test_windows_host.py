"""Host-side helpers of the per-pixel windows (no GPU): the uneven line list,
the window draw, and the per-pixel line setup / index pair that the GPU tests
use to restate a pixel's ragged spectrum as a dense one."""
import numpy as np

from firtez_dz_amd import lines


def test_uneven_list_regions():
    ls = lines.uneven_list()
    assert [l.id for l in ls.lines] == list(lines.UNEVEN_IDS)
    assert len(ls.regions) == 8 and all(r.nwave == 300 for r in ls.regions)
    for l, r in zip(ls.lines, ls.regions):
        # centred grid, step proportional to lambda0 (10 mA at 6300 A)
        assert abs(r.wave_start + r.wave_step * 150) < 1e-9
        assert abs(r.wave_step - 10.0 * l.lambda0 / 6300.0) <= 0.05 + 1e-12


def test_uneven_windows_draw():
    ls = lines.uneven_list()
    w = lines.uneven_windows(ls, 2000, seed=9)
    assert w.shape == (2000, 8, 2) and w.dtype == np.int32
    nobs = (w[:, :, 1] > 0).sum(axis=1)
    assert nobs.min() == 1 and nobs.max() == 4
    cnt = w[:, :, 1][w[:, :, 1] > 0]
    assert cnt.min() >= 60 and cnt.max() <= 300
    # centred: first = (nwave - count) // 2
    obs = w[:, :, 1] > 0
    assert np.array_equal(w[:, :, 0][obs], (300 - w[:, :, 1][obs]) // 2)
    assert np.all(w[:, :, 0][~obs] == 0)
    # seeded
    assert np.array_equal(w, lines.uneven_windows(ls, 2000, seed=9))


def test_window_setup_and_index_agree():
    ls = lines.setup([((4, 5), 112, -624.0, 20.0), ((2,), 80, -400.0, 10.0), ((14,), 96, -1200.0, 25.0)])
    full = ls.wave()
    rng = np.random.default_rng(1)
    for _ in range(50):
        win = np.zeros((3, 2), np.int32)
        for r, reg in enumerate(ls.regions):
            if rng.random() < 0.6:
                c = int(rng.integers(1, reg.nwave + 1))
                win[r] = (int(rng.integers(0, reg.nwave - c + 1)), c)
        if win[:, 1].sum() == 0:
            win[0] = (0, 1)
        idx = lines.window_index(ls, win)
        sub = lines.window_setup(ls, win)
        assert idx.size == win[:, 1].sum() == sub.numw
        # the pixel's dense grid is the union grid at its indices
        assert np.allclose(sub.wave(), full[idx], rtol=0, atol=1e-9)
        # region of each index: the union region it falls in keeps its lines
        kept = [r for r in range(3) if win[r, 1] > 0]
        assert [s.lines for s in sub.regions] == [ls.regions[r].lines for r in kept]
        assert np.all(np.diff(idx) > 0)
