"""Small GeCCo-layout operator sets used by the tests (built with the oracle)."""
import numpy as np

from oracle.melist import Orbitals, MEList
from oracle.contract import Contraction, sym_ab

H, P = 1, 2


def round_robin(n, nsym, start=0):
    return [((start + k) % nsym) + 1 for k in range(n)]


def orbitals(nocc, nvir, nsym):
    return Orbitals(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)


def spin_symmetric(mel, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    x = scale * rng.standard_normal(mel.length)
    sym_ab(mel, x, 0.5, float(mel.absym))
    return x


def ladder(orbs):
    """R(ab,ij) += sum_{c<d} V(ab,cd) T(cd,ij)."""
    T = MEList(orbs, {P: 2}, {H: 2})
    V = MEList(orbs, {P: 2}, {P: 2})
    R = MEList(orbs, {P: 2}, {H: 2})
    return Contraction(V, T, R, {}, {P: 2}), V, T, R


def ring(orbs):
    """R(ab,ij) += sum_{kc} T(ac,ik) W(kb,cj): merged strings on every operand."""
    T = MEList(orbs, {P: 2}, {H: 2})
    W = MEList(orbs, {H: 1, P: 1}, {H: 1, P: 1})
    R = MEList(orbs, {P: 2}, {H: 2})
    return Contraction(T, W, R, {P: 1}, {H: 1}), T, W, R


def t1t1(orbs):
    """R(ab,ij) += T1(a,i) T1(b,j): no contraction, non-injective scatter."""
    T1 = MEList(orbs, {P: 1}, {H: 1})
    R = MEList(orbs, {P: 2}, {H: 2})
    return Contraction(T1, T1, R, {}, {}), T1, T1, R


def hh_ladder(orbs):
    """R(ab,ij) += sum_{k<l} T(ab,kl) W(kl,ij)."""
    T = MEList(orbs, {P: 2}, {H: 2})
    W = MEList(orbs, {H: 2}, {H: 2})
    R = MEList(orbs, {P: 2}, {H: 2})
    return Contraction(T, W, R, {}, {H: 2}), T, W, R


def fock_pp(orbs):
    """R(ab,ij) += sum_c T(ac,ij) F(c,b)-like: op1 C merges, op2 one-body."""
    T = MEList(orbs, {P: 2}, {H: 2})
    F = MEList(orbs, {P: 1}, {P: 1})
    R = MEList(orbs, {P: 2}, {H: 2})
    return Contraction(T, F, R, {P: 1}, {}), T, F, R


def energy_like(orbs):
    """E += sum V(ij,ab) T(ab,ij): full contraction to a scalar list."""
    T = MEList(orbs, {P: 2}, {H: 2})
    V = MEList(orbs, {H: 2}, {P: 2})
    E = MEList(orbs, {}, {})
    return Contraction(V, T, E, {H: 2}, {P: 2}), V, T, E


CASES = {'ladder': ladder, 'ring': ring, 't1t1': t1t1, 'hh_ladder': hh_ladder,
         'fock_pp': fock_pp, 'energy': energy_like}
