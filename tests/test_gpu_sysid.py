"""Batched system identification: Adam on (ln E, nu) per scene from adjoint gradients.

Small version of the identification workload (3 scenes of 8^3 cells x 2
particles at dx = 1/48, 16 substeps, 12 iterations).  Each scene's gradient
comes from the same multi-scene adjoint run; the loss must fall at every
iteration and the fitted moduli must move toward the hidden ones.
"""
import numpy as np
import pytest

from zeno_amd import sysid

pytestmark = pytest.mark.gpu


def test_sysid_loss_decreases():
    batch = sysid.make_scenes(n_scenes=3, cells=8, ppc=2, n_cells=48, seed=7)
    r = sysid.identify(batch, steps=16, iters=12)
    loss = np.array(r["loss"])
    assert np.all(np.diff(loss) < 0), loss
    assert loss[-1] < 0.8 * loss[0], loss
    start = np.abs(np.log(1e5) - np.log(r["hidden_E"]))
    end = np.abs(np.log(r["E"]) - np.log(r["hidden_E"]))
    assert np.all(end < start), (r["E"], r["hidden_E"])
