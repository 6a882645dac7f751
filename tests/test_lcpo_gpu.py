"""LCPO SASA on the device against the oracle's restatement of
FF_SASA_LCPO::calcLCPOSasaForces.

Bars: per-atom SASA within 1e-12 relative (both sides FP64 in PD's order);
forces (64-bit fixed point) within 1e-8 of the RMS force; the cavity energy
within 1e-12 relative.  The UpdateSasa schedule (full recalculation every
UpdateSasa steps, stored forces re-applied in between) is checked through the
per-atom SASA the engine keeps.
"""
import numpy as np
import pytest

from oracle import oracle as O
from pd_amd import engine as G
from pd_amd import systems as S
from pd_amd import units as U
from tests.helpers import rel_rms

pytestmark = pytest.mark.gpu

ASP = 0.005  # kcal/mol/A^2, PD's examples (examples/md/gbsa_langevin/run.py:45)


def _all_pairs_list(n):
    counts = np.arange(n, dtype=np.int32)
    packed = np.concatenate([np.arange(i, dtype=np.uint32) | np.uint32(7 << 24) for i in range(n)])
    return counts, packed.astype(np.uint32)


def _oracle(pos, use, rad, P):
    sig = np.full(len(pos), ASP / U.NA / U.J2KCAL)
    c, pk = _all_pairs_list(len(pos))
    return O.lcpo_forces(pos, use, rad, P, sig, c, pk)


def _bare_engine(pos):
    n = len(pos)
    e = G.Engine(n, 0)
    z = np.zeros(n)
    e.set_atom_params(z, z, z, np.full(n, 12.0 * U.AMU))
    e.set_positions(pos)
    return e


def test_lcpo_peptide_matches_oracle():
    s = S.peptide_chain(30, seed=4)
    use, rad, P = S.lcpo_params(s)
    e = G.engine_for(s)
    e.setup_lcpo(G.LCPOParams(GlobalASP=ASP, UpdateSasa=5), rad, P, use)
    f, en = e.compute(G.PDGPU_SASA)
    f_ref, sasa_ref, e_ref = _oracle(s.pos, use, rad, P)
    np.testing.assert_allclose(e.sasa(), sasa_ref, rtol=1e-12, atol=1e-12)
    assert rel_rms(f, f_ref) <= 1e-8
    assert en["epot_surf"] == pytest.approx(e_ref, rel=1e-12)
    assert en["epot"] == pytest.approx(e_ref, rel=1e-12)


def test_lcpo_dense_cluster_truncation_and_clamp():
    """More than 80 overlap partners per atom and buried atoms clamped at zero."""
    rng = np.random.default_rng(5)
    n = 140
    pos = rng.uniform(-5.0, 5.0, (n, 3))
    rad = rng.choice([1.6, 1.65, 1.7, 1.9], n)
    use = (rng.random(n) > 0.1).astype(np.int8)
    P = np.column_stack([rng.uniform(0.05, 0.8, n), rng.uniform(-0.3, -0.02, n),
                         rng.uniform(-2e-3, 1e-3, n), rng.uniform(1e-5, 5e-4, n)])
    e = _bare_engine(pos)
    e.setup_lcpo(G.LCPOParams(GlobalASP=ASP, UpdateSasa=5), rad, P, use)
    f, en = e.compute(G.PDGPU_SASA)
    f_ref, sasa_ref, e_ref = _oracle(pos, use, rad, P)
    assert (sasa_ref[use > 0] == 0).any()
    np.testing.assert_allclose(e.sasa(), sasa_ref, rtol=1e-12, atol=1e-9)
    assert rel_rms(f, f_ref) <= 1e-8
    assert np.all(f[sasa_ref == 0] == 0.0)
    assert en["epot_surf"] == pytest.approx(e_ref, rel=1e-12)


def test_lcpo_update_schedule_in_md():
    """Steps with step % UpdateSasa != 0 re-apply the last full result."""
    s = S.peptide_chain(12, seed=8)
    use, rad, P = S.lcpo_params(s)
    e = G.engine_for(s)
    e.setup_lcpo(G.LCPOParams(GlobalASP=ASP, UpdateSasa=5), rad, P, use)
    e.compute(G.PDGPU_SASA)  # full at the starting positions A
    sasa_a = e.sasa()
    rng = np.random.default_rng(1)
    pos_b = s.pos + rng.normal(scale=0.05, size=s.pos.shape)
    e.set_positions(pos_b)
    md = dict(integrator=G.PDGPU_VELOCITY_VERLET, components=G.PDGPU_SASA, timestep=1e-16, temperature=300.0,
              fric_coeff=0.0, update_nlist=0, total_steps=100, seed=1, noise_stream=0)
    e.md(G.MDParams(**md, step_offset=3), 1)  # step 3: re-apply, SASA unchanged
    np.testing.assert_array_equal(e.sasa(), sasa_a)
    pos_c = e.get_positions()
    e.md(G.MDParams(**md, step_offset=5), 1)  # step 5: full recalculation at the current positions
    _, sasa_c, _ = _oracle(pos_c, use, rad, P)
    np.testing.assert_allclose(e.sasa(), sasa_c, rtol=1e-12, atol=1e-12)


def test_lcpo_batch_equals_single_systems():
    sys_ = [S.peptide_chain(k, seed=20 + k) for k in (6, 15, 9)]
    engines = []
    for s in sys_:
        use, rad, P = S.lcpo_params(s)
        e = G.engine_for(s, bonded=True)
        e.setup_lcpo(G.LCPOParams(GlobalASP=ASP, UpdateSasa=5), rad, P, use)
        engines.append(e)
    b = G.Batch(engines)
    comps = G.PDGPU_SASA | G.PDGPU_BONDED
    fb, eb = b.compute(comps)
    for k, e in enumerate(engines):
        f, en = e.compute(comps)
        assert np.array_equal(fb[k], f), k
        assert eb[k]["epot_surf"] == pytest.approx(en["epot_surf"], rel=1e-12)
        assert eb[k]["epot"] == pytest.approx(en["epot"], rel=1e-12)
