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
