"""The spin-projection oracle (oracle/spin_project.py) on the CPU: the
projector onto singlet operators of a list, pinned by what the mathematics
fixes -- a symmetric idempotent map whose image commutes with S+ and S-, that
keeps spin-free operators, with the known number of singlets per group."""
import numpy as np
import pytest

from models import orbitals
from oracle.melist import MEList
from oracle.spin_project import SpinSpace

H, P = 1, 2
LISTS = {
    't1': ({P: 1}, {H: 1}),
    't2': ({P: 2}, {H: 2}),
    'fae': ({P: 1}, {P: 1}),
    'wmbej': ({P: 1, H: 1}, {P: 1, H: 1}),
    't3': ({P: 3}, {H: 3}),
}


def _setup(name, nocc=2, nvir=2, nsym=1):
    orbs = orbitals(nocc, nvir, nsym)
    c, a = LISTS[name]
    return MEList(orbs, c, a), SpinSpace(nocc, nvir)


@pytest.mark.parametrize('name', list(LISTS))
def test_projector_is_the_singlet_projector(name):
    L, sp = _setup(name)
    Pm = sp.projector(L)
    assert np.allclose(Pm, Pm.T, atol=1e-12)
    assert np.allclose(Pm @ Pm, Pm, atol=1e-12)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(L.length)
    X = sp.operator_of(L, Pm @ x)
    for S in (sp.splus, sp.sminus):
        assert np.abs(S @ X - X @ S).max() < 1e-12
    # the part removed is not a singlet: the projection is not the identity
    assert np.abs(Pm @ x - x).max() > 1e-3


def test_spin_free_operators_are_fixed():
    """sum t_ij^ab E_ai E_bj with E_pq = sum_sigma a+_p,sigma a_q,sigma is a
    singlet: its list coefficients are unchanged by the projection."""
    L, sp = _setup('t2')
    rng = np.random.default_rng(5)
    no, nv = 2, 2
    E = lambda p, q: sp.ad[2 * p] @ sp.a[2 * q] + sp.ad[2 * p + 1] @ sp.a[2 * q + 1]
    X = 0
    for i in range(no):
        for j in range(no):
            for a in range(nv):
                for b in range(nv):
                    X = X + rng.standard_normal() * E(no + a, i) @ E(no + b, j)
    x = sp.coefficients(L, X)
    # X lies in the span of the list's operators (double excitations)
    assert np.abs(sp.operator_of(L, x) - X).max() < 1e-12
    assert np.abs(sp.projector(L) @ x - x).max() < 1e-12


def test_singlet_counts():
    """Two creators and two annihilators on four different spatial orbitals:
    the group holds the 6 spin patterns with zero total Ms and 2 singlets
    (two ways of coupling two spin-1/2 pairs to S = 0 overall); one creator
    and one annihilator: 2 patterns, 1 singlet (the spin-free E_ai)."""
    for name, per_group, singlets in (('t2', 6, 2), ('t1', 2, 1)):
        L, sp = _setup(name, nsym=1)
        from oracle.ccsd_route3 import list_spinorbs
        so = list_spinorbs(L)
        nc = L.nc
        groups = {}
        for e in range(L.length):
            key = (tuple(sorted(so[e, :nc] >> 1)), tuple(sorted(so[e, nc:] >> 1)))
            groups.setdefault(key, []).append(e)
        Pm = sp.projector(L)
        distinct = [g for k, g in groups.items() if len(set(k[0])) == len(k[0]) and len(set(k[1])) == len(k[1])]
        assert distinct
        for g in distinct:
            assert len(g) == per_group
            assert round(float(np.trace(Pm[np.ix_(g, g)]))) == singlets
            # the projector does not couple different groups
            others = [e for e in range(L.length) if e not in g]
            assert np.abs(Pm[np.ix_(g, others)]).max(initial=0.0) < 1e-12
