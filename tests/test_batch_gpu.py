"""Batches of uneven GB/SA systems and REMD replicas: a batch must reproduce the
single-system engine exactly, and REMD exchanges must follow PD's criterion."""
import numpy as np
import pytest

from oracle import oracle as O
from pd_amd import engine as G
from pd_amd import systems as S
from pd_amd import units as U

pytestmark = pytest.mark.gpu

GB = dict(Cutoff=999.0, InnerCutoff=999.0, VdwCutoff=999.0, VdwInnerCutoff=999.0, GbsaStillCutoff=999.0,
          DielectricSolvent=78.5, Vdw14Scaling=0.5, Elec14Scaling=0.83333)
COMPS = G.PDGPU_GB | G.PDGPU_BONDED


def gb_engine(s, seed):
    e = G.engine_for(s, None, bonded=True)
    e.setup_gb(G.gb_params(**GB), s.gbrad, s.cov)
    rng = np.random.default_rng(seed)
    e.set_velocities(rng.normal(size=(s.n, 3)) * np.sqrt(U.KB * 300.0 / s.mass)[:, None])
    return e


def chains(lengths, seed0=0):
    return [S.peptide_chain(n, seed=seed0 + k) for k, n in enumerate(lengths)]


def test_batch_compute_equals_single_systems():
    sys_ = chains([7, 20, 3, 12])
    engines = [gb_engine(s, k) for k, s in enumerate(sys_)]
    b = G.Batch(engines)
    fb, eb = b.compute(COMPS)
    for k, e in enumerate(engines):
        f, en = e.compute(COMPS)
        assert np.array_equal(fb[k], f), k
        for key in ("epot_pol", "epot_vdw", "epot_elec", "epot_bond", "epot_angle", "epot_torsion", "epot"):
            assert eb[k][key] == pytest.approx(en[key], rel=1e-12, abs=1e-30), (k, key)


def test_batch_langevin_equals_single_systems():
    sys_ = chains([5, 11, 8], seed0=10)
    engines = [gb_engine(s, 20 + k) for k, s in enumerate(sys_)]
    b = G.Batch(engines)
    b.set_temperatures([300.0, 350.0, 400.0])
    md = G.MDParams(integrator=G.PDGPU_LANGEVIN, components=COMPS, timestep=2e-15, temperature=300.0,
                    fric_coeff=1e12, update_nlist=0, step_offset=0, total_steps=25, seed=99, noise_stream=0)
    ep_b, ek_b = b.md(md, 25)
    for k, e in enumerate(engines):
        m = G.MDParams(integrator=G.PDGPU_LANGEVIN, components=COMPS, timestep=2e-15,
                       temperature=[300.0, 350.0, 400.0][k], fric_coeff=1e12, update_nlist=0, step_offset=0,
                       total_steps=25, seed=99, noise_stream=k)
        ep, ek = e.md(m, 25)
        assert np.array_equal(b.get_positions(k), e.get_positions()), k
        assert np.array_equal(b.get_velocities(k), e.get_velocities()), k
        np.testing.assert_allclose(ep_b[:, k], ep, rtol=1e-12)
        np.testing.assert_allclose(ek_b[:, k], ek, rtol=1e-12)


def test_remd_exchanges_follow_metropolis():
    base = S.peptide_chain(6, seed=3)
    temps = 300.0 * (500.0 / 300.0) ** (np.arange(4) / 3.0)
    engines = [gb_engine(base, 40 + k) for k in range(4)]
    b = G.Batch(engines)
    b.set_temperatures(temps)
    md = G.MDParams(integrator=G.PDGPU_LANGEVIN, components=COMPS, timestep=2e-15, temperature=300.0,
                    fric_coeff=1e12, update_nlist=0, step_offset=0, total_steps=20, seed=5, noise_stream=0)
    # run the segment by hand to know the energies and structures before the exchange
    b2 = G.Batch(engines)
    b2.set_temperatures(temps)
    ep, _ = b2.md(md, 20)
    e_last = ep[-1]
    before = [b2.get_positions(k) for k in range(4)]
    seed = 1234
    acc, att, fe = b.remd(md, 1, seed)
    np.testing.assert_array_equal(fe, e_last)
    for r in (0, 2):  # round 0 tries the even pairs
        assert att[r] == 1
        expect = O.remd_accept(e_last[r], e_last[r + 1], temps[r], temps[r + 1], O.remd_uniform(seed, r, 0))
        assert bool(acc[r]) == expect
        pr, pr1 = b.get_positions(r), b.get_positions(r + 1)
        if expect:
            assert np.array_equal(pr, before[r + 1]) and np.array_equal(pr1, before[r])
        else:
            assert np.array_equal(pr, before[r]) and np.array_equal(pr1, before[r + 1])
    assert att[1] == 0 and att[3] == 0
