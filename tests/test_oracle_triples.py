"""Oracle pins for the CCSDT triples contractions (T3 <- V.T3 and V.T2) on
three-body ME-lists: each term against the dense antisymmetrised einsum of
the same product on tiny orbital spaces (the result's creator string merges
op1's remaining creators before op2's, its annihilator string op2's before
op1's; a contraction over k equal orbitals carries 1/k!, a merge of n1 and
n2 orbitals of one space 1/(n1! n2!) of the full antisymmetriser), route 2
== route 3, and spin-flip symmetry on/off."""
import numpy as np
import pytest

from oracle.contract import to_dense, from_dense
from models import orbitals, spin_symmetric, T3_CASES
from test_oracle_contract import antisym

C3, A3 = [0, 1, 2], [3, 4, 5]


def dense_t3(name, D1, D2):
    if name == 't3_pp_ladder':
        return antisym(0.5 * np.einsum('abde,decijk->abcijk', D1, D2), [C3]) / 2
    if name == 't3_hh_ladder':
        return antisym(0.5 * np.einsum('lmij,abclmk->abckij', D1, D2), [A3]) / 2
    if name == 't3_ring':
        return antisym(np.einsum('dablij,cldk->abckij', D1, D2), [C3, A3]) / 4
    if name == 't3_vvvo_t2':
        return antisym(np.einsum('bcek,eaij->bcaijk', D1, D2), [C3, A3]) / 4
    if name == 't3_ovoo_t2':
        return antisym(np.einsum('cmjk,abmi->cabijk', D1, D2), [C3, A3]) / 4
    raise KeyError(name)


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 4), (2, 3, 4)])
@pytest.mark.parametrize('name', list(T3_CASES))
@pytest.mark.parametrize('use_tr,route', [(False, 3), (True, 3), (False, 2)])
def test_triples_match_dense_einsum(nsym, nocc, nvir, name, use_tr, route):
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = T3_CASES[name](orbs)
    x1 = spin_symmetric(a, 21)
    x2 = spin_symmetric(b, 22)
    ref = from_dense(r, dense_t3(name, to_dense(a, x1), to_dense(b, x2)))
    assert np.abs(ref).max() > 0.1
    xr = np.zeros(r.length)
    c.run(x1, x2, xr, route=route, use_tr=use_tr)
    assert np.abs(xr - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_triples_dense_check_catches_a_dropped_term():
    """Leaving one kernel call out of the ring term breaks the pin."""
    orbs = orbitals(3, 4, 2)
    c, a, b, r = T3_CASES['t3_ring'](orbs)
    x1 = spin_symmetric(a, 21)
    x2 = spin_symmetric(b, 22)
    ref = from_dense(r, dense_t3('t3_ring', to_dense(a, x1), to_dense(b, x2)))
    blocks = c.blocks(False)
    xr = np.zeros(r.length)
    c.run(x1, x2, xr, use_tr=False, blocks=blocks[1:])
    assert np.abs(xr - ref).max() > 1e-6
