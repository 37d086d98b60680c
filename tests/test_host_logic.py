"""Host-side logic that needs no GPU: Lame chain rule of the gradients, the
batched and identification scene generators, scene merging."""
import math

import numpy as np
import pytest

from zeno_amd import scenes
from zeno_amd.diff import lame, lame_grad


@pytest.mark.parametrize("E,nu", [(5e4, 0.3), (1e6, 0.15), (1e4, 0.45)])
def test_lame_grad_matches_finite_differences(E, nu):
    rng = np.random.default_rng(1)
    mub, lamb = rng.standard_normal(2)
    L = lambda e, n: mub * lame(e, n)[0] + lamb * lame(e, n)[1]
    gE, gnu = lame_grad(E, nu, mub, lamb)
    hE, hn = 1e-6 * E, 1e-7
    fdE = (L(E + hE, nu) - L(E - hE, nu)) / (2 * hE)
    fdn = (L(E, nu + hn) - L(E, nu - hn)) / (2 * hn)
    assert abs(gE - fdE) <= 1e-7 * abs(fdE) + 1e-12
    assert abs(gnu - fdn) <= 1e-6 * abs(fdn) + 1e-9


def test_lame_matches_oracle_convention():
    mu, lam = lame(5e4, 0.2)
    assert math.isclose(mu, 5e4 / 2.4) and math.isclose(lam, 5e4 * 0.2 / (1.2 * 0.6))


def test_batched_scenes_reproducible_and_in_range():
    a = scenes.config_batched(8)
    b = scenes.config_batched(8)
    for s, t in zip(a, b):
        assert np.array_equal(s["state"]["x"], t["state"]["x"])
        n = s["state"]["x"].shape[1]
        assert 2048 <= n <= 65536
        assert s["n_cells"] in (64, 128)
        m = s["mats"][0]
        assert 1e4 <= m["E"] <= 1e6 and 0.1 <= m["nu"] <= 0.4
        x = s["state"]["x"]
        assert np.all(np.abs(x - 0.5) <= 0.5 * s["side"] + 1e-6)
        assert np.all(np.abs(s["state"]["v"]) <= 1.0)


def test_merge_scenes_offsets_materials_and_ranges():
    batch = [scenes.random_box_scene(i, 100, n_range=(100, 400)) for i in range(3)] + [scenes.small_mixed_scene()]
    descs, mats, st, ranges = scenes.merge_scenes(batch)
    assert len(descs) == 4 and [d["n_cells"] for d in descs] == [b["n_cells"] for b in batch]
    assert [m["scene"] for m in mats] == [0, 1, 2, 3, 3]
    assert ranges[-1][1] == st["x"].shape[1]
    for i, (a, b) in enumerate(ranges):
        assert np.array_equal(st["x"][:, a:b], batch[i]["state"]["x"])
        assert all(mats[k]["scene"] == i for k in np.unique(st["mat"][a:b]))


def test_sysid_scenes():
    from zeno_amd import sysid
    b = sysid.make_scenes(n_scenes=4, cells=4, ppc=2, n_cells=48)
    for s in b:
        assert s["state"]["x"].shape[1] == 4 ** 3 * 2
        assert 1e4 <= s["E"] <= 1e6 and 0.15 <= s["nu"] <= 0.4
        v = s["state"]["v"]
        # drift + stretch along x and squeeze along y about the centre
        assert v[0].std() > 0 and v[1].std() > 0 and np.allclose(v[2], 0.1)
