"""Batched ensemble (CPU): the scene generator and the batch layout.

SplitMix64 is pinned by its published test vector (seed 1234567); the scene
draws by their ranges; the batch's concatenated host arrays and offsets by
construction.
"""
import numpy as np

from zsph_amd import FLUID, cases


def test_splitmix64_known_answer():
    state, out = 1234567, []
    for _ in range(5):
        state, z = cases.splitmix64(state)
        out.append(z)
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]


def test_ensemble_draws_are_deterministic_and_in_range():
    draws = np.array([cases.ensemble_draw(20240601 + k) for k in range(2000)])
    assert np.array_equal(draws, np.array([cases.ensemble_draw(20240601 + k) for k in range(2000)]))
    lo, hi = np.array([0.2, 0.3, 0.004]), np.array([0.6, 0.8, 0.01])
    assert (draws >= lo).all() and (draws < hi).all()
    # roughly uniform: each quarter of each range gets 20-30 % of the draws
    for a in range(3):
        q = np.floor((draws[:, a] - lo[a]) / (hi[a] - lo[a]) * 4).astype(int)
        frac = np.bincount(q, minlength=4) / len(draws)
        assert (frac > 0.2).all() and (frac < 0.3).all()


def test_batch_layout():
    b = cases.ensemble2d(n_scenes=6, seed=7)
    assert b.nscenes == 6 and b.n == sum(c.n for c in b.cases)
    assert list(b.offsets) == list(np.cumsum([0] + [c.n for c in b.cases]))
    for k, c in enumerate(b.cases):
        sl = slice(b.offsets[k], b.offsets[k + 1])
        assert np.array_equal(b.host.pos[sl], c.host.pos)
        assert np.array_equal(b.host.type[sl], c.host.type)
        w, H, dx = cases.ensemble_draw(7 + k)
        assert c.json.dx == np.float32(dx)
        nf = int((c.host.type == FLUID).sum())

        def lattice(length):  # sites of the (i + 1/2) dx lattice in a column of this length
            return int(np.floor(float(np.float32(length)) / float(np.float32(dx)) + 0.5))
        assert nf == lattice(w) * lattice(H)
        assert c.cpu.dim == 2
    assert np.array_equal(b.scene_of(np.array([0, b.offsets[1] - 1, b.offsets[1], b.n - 1])), [0, 0, 1, 5])
    assert b.grid_num == sum(c.cpu.grid_num for c in b.cases)
    # per-scene parameters differ (spacing, h, dt)
    assert len({round(c.params.h, 9) for c in b.cases}) == 6
    assert len({c.params.half_dt for c in b.cases}) == 6
