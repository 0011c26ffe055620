"""Host-side logic that needs no GPU: the brick-shape candidates of the tuner."""
import itertools

import numpy as np
import pytest

from flowmol_amd.md import FB_BLOCK, brick_candidates


@pytest.mark.parametrize("ncells,nmol", [((39, 39, 39), 1048576), ((4, 4, 4), 2048), ((9, 9, 9), 2048),
                                         ((3, 3, 3), 500), ((12, 5, 7), 30000)])
def test_brick_candidates_fit_a_workgroup_pass_and_rank_by_halo(ncells, nmol):
    occ = max(1.0, nmol / float(np.prod(ncells)))
    got = brick_candidates(ncells, nmol, max_candidates=6)
    assert 1 <= len(got) <= 6
    eff = [np.prod(b) / np.prod([x + 2 for x in b]) for b in got]
    assert eff == sorted(eff, reverse=True)
    for b in got:
        assert all(1 <= b[d] <= min(ncells[d], 12) for d in range(3))
        assert np.prod(b) == 1 or np.prod(b) * occ <= FB_BLOCK * 15 / 16
    # brute force: no admissible shape has less halo per cell than the first candidate
    best = max(np.prod(b) / np.prod([x + 2 for x in b])
               for b in itertools.product(*[range(1, min(n, 12) + 1) for n in ncells])
               if np.prod(b) == 1 or np.prod(b) * occ <= FB_BLOCK * 15 / 16)
    assert abs(eff[0] - best) < 1e-15


def test_brick_candidates_at_the_1m_workload():
    # 17.7 molecules per cell: 27 cells fit one pass of 480 threads, 3x3x3 has the least halo
    assert brick_candidates((39, 39, 39), 1048576, 1) == [(3, 3, 3)]


def test_random_walk_chains():
    """Seeded random-walk chains: raw walks have every bond exactly 0.97 long (minimum image);
    after the overlap relaxation of the initial state no two non-bonded beads are closer than
    0.8 and every bond is within 0.05 of 0.97; beads inside the box; the same chains for the
    same seed."""
    from scipy.spatial import cKDTree

    from flowmol_amd import config

    p, r, v, mt, mono = config.config_polymer(nunits=(8, 8, 8), init="random_walk", seed=9)
    L = np.array([p.domain[0], p.domain[1], p.domain[2]])
    n = r.shape[0]
    assert n == 2048 and np.all(r >= -0.5 * L) and np.all(r < 0.5 * L)
    raw = config.random_walk_chains(L, n // 8, 8, 9)
    bonded = (np.arange(n - 1) % 8) != 7

    def bonds(x):
        d = x[1:] - x[:-1]
        d -= L * np.round(d / L)
        return np.linalg.norm(d, axis=1)[bonded]

    assert np.abs(bonds(raw) - 0.97).max() < 1e-12
    assert np.abs(bonds(r) - 0.97).max() <= 0.05
    pairs = cKDTree(r + 0.5 * L, boxsize=L).query_pairs(0.8, output_type="ndarray")
    close = [(i, j) for i, j in pairs if not (j == i + 1 and i % 8 != 7)]
    assert not close
    chain, sub, funcy, bflag = mono
    assert list(sub[:8]) == list(range(1, 9)) and list(funcy[:8]) == [1, 2, 2, 2, 2, 2, 2, 1]
    _, r2, _, _, _ = config.config_polymer(nunits=(8, 8, 8), init="random_walk", seed=9)
    assert np.array_equal(r, r2)
