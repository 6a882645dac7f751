"""Input builders: sizes, densities and topology of the benchmark workloads."""
import numpy as np
import pytest

from pd_amd import systems as S
from pd_amd import units as U


def test_water_box_config():
    w = S.water_box(32768, seed=2024)
    assert w.n == 98304
    assert w.box[0] == pytest.approx(99.346, abs=2e-3)
    grams = w.mass.sum() * 1e3
    assert grams / (w.box.prod() * 1e-24) == pytest.approx(1.0, rel=1e-12)
    oh = np.linalg.norm(w.pos[1::3] - w.pos[0::3], axis=1)
    assert np.allclose(oh, 0.9572, atol=1e-12)
    assert w.charge.sum() == pytest.approx(0.0, abs=1e-9)
    assert len(w.mol_first) == 32768


def test_even_points_lattice():
    pts = S.even_points(np.array([10.0, 10.0, 10.0]), 27)
    assert len(pts) == 27
    assert np.allclose(np.sort(np.unique(pts[:, 0])), [-10 / 3, 0.0, 10 / 3])


@pytest.mark.parametrize("n_res", [20, 60])
def test_peptide_chain(n_res):
    s = S.peptide_chain(n_res, seed=n_res)
    assert 10 * n_res < s.n < 22 * n_res
    assert abs(s.charge.sum() - round(s.charge.sum())) < 1e-9
    assert S.min_nonbonded_distance(s) > 1.5
    d = np.linalg.norm(s.pos[s.bonds[:, 0]] - s.pos[s.bonds[:, 1]], axis=1)
    assert np.allclose(d, s.bond_l)
    c = s.cov
    for i in range(s.n):
        for k in range(c["off12"][i], c["off12"][i + 1]):
            j = c["idx12"][k]
            assert i in c["idx12"][c["off12"][j]:c["off12"][j + 1]]
