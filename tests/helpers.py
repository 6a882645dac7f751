"""Shared helpers for the parity tests: oracle evaluations of a MolSystem in PD units."""
import numpy as np

from oracle import oracle as O
from pd_amd import units as U


def oracle_list(s, pos, cutoff, padding, shadow):
    bo = O.bond_order(s)
    counts, packed = O.build_list(s, pos, cutoff, padding, shadow, bo)
    return counts, packed


def oracle_nb(s, nb_kw, padding=0.5, shadow=False):
    """cleanSpace + list + fast2 forces with the oracle, as PD would do on a rebuild step."""
    pos = O.clean_space(s)
    p = O.nb_params(**nb_kw)
    cut = max(p.Cutoff, p.VdwCutoff)
    counts, packed = oracle_list(s, pos, cut, padding, shadow)
    basis = O.basis_vectors(s.box is not None, s.box, cut)
    f, e = O.nb_forces(s, pos, counts, packed, p, basis)
    e["abs_sum"] = O.nb_abs_energy(s, pos, counts, packed, p, basis)
    return pos, counts, packed, f, e


def energy_close(got, ref, abs_sum, rel=1e-5, cond=1e-6):
    """|got - ref| <= rel*|ref| + cond*sum|E_pair|: the relative bar plus the FP32
    summation allowance for totals that cancel (sum|E_pair| >> |E|)."""
    return abs(got - ref) <= rel * abs(ref) + cond * abs_sum


def rel_rms(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.sqrt(((a - b) ** 2).mean()) / np.sqrt((b ** 2).mean()))


def kcal(e):
    return e * U.J_TO_KCALMOL
