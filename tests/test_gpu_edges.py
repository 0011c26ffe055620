"""Error paths the reference aborts on, through the C ABI on the GPU: a molecule beyond the box
at binning (src/linklist.f90:2665-2671) and a domain that the reference's cell rule rejects; the
run stops with the reference's error instead of binning the molecule anywhere."""
import sys

import numpy as np
import pytest

from flowmol_amd import config
from flowmol_amd._lib import FlowmolHipError
from flowmol_amd.md import MDContext

pytestmark = pytest.mark.gpu
FM_EESCAPE = 4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_escaped_molecule_is_an_error():
    p, r, v = config.config_fcc((6, 6, 6), 0.8, 2.5, seed=3, tplot=10)
    L = np.array(p.domain[:])
    r = r + np.random.default_rng(3).uniform(-1e-3, 1e-3, r.shape)
    say("case: one molecule two box lengths out, 864 LJ")
    g = MDContext(p, r, v)
    g.setup(0)
    bad = r.copy()
    bad[17, 0] = 2.0 * L[0]          # sendmols' single wrap leaves it a box length outside
    g.upload_state(bad, v)
    with pytest.raises(FlowmolHipError) as e:
        g.assign_to_cell()
    assert e.value.code == FM_EESCAPE
    # the same state with the molecule back inside builds normally
    g.upload_state(r, v)
    g.assign_to_cell()
    g.messenger_updateborders(1)
    g.assign_to_neighbourlist()
    g.simulation_compute_forces(0)
    g.close()


def test_domain_below_three_cells_is_rejected():
    """set_parameters_cells aborts when a domain has fewer than 3 cells per axis
    (src/setup_set_parameters.f90:1891-1896); the context refuses it at creation."""
    p, r, v = config.config_fcc((3, 3, 3), 0.8, 2.5, seed=3, tplot=10)   # 5.1 wide: 1 cell of 2.8
    say("case: 108 LJ in a box of one cell per axis")
    with pytest.raises(FlowmolHipError) as e:
        MDContext(p, r, v)
    assert "3 CELLS" in str(e.value).upper()
