"""Oracle pins for contractions of operators with several vertices
(Contraction.joined, the multi-vertex form of contr_op1op2's arguments):
every case against a dense einsum built from the vertex-level merge maps
alone (each operand component holds its contraction lines before its
external lines, a result creator component EX1's lines before EX2's and an
annihilator component EX2's before EX1's; a contraction over k lines of one
component carries 1/k!, a merge of n1 and n2 lines 1/(n1! n2!) of the full
antisymmetriser), route 2 == route 3, spin-flip symmetry on and off, and a
planted error caught."""
import itertools
import math

import numpy as np
import pytest

from oracle.contract import to_dense, from_dense, _merge_lists
from oracle.melist import HPVXSEQ
from models import JOINED_CASES, orbitals, spin_symmetric
from test_oracle_contract import antisym


def dense_joined(c, D1, D2, ex2_first=False):
    """The dense result of c (a Contraction.joined) from the dense operands,
    using only the per-vertex occupations and the merge maps (ex2_first: a
    planted error, EX2's creators before EX1's)."""
    counter = itertools.count()
    lab = {}

    def labels(part, vocc, v, side, h):
        n = vocc[v - 1][side].get(h, 0)
        key = (part, v, side, h)
        if key not in lab:
            lab[key] = [next(counter) for _ in range(n)]
        return lab[key]

    m1, m2, m12, m21 = c.merge
    op1, op2, res = c.op1, c.op2, c.res

    def axes(mel, fn):
        out = []
        for side, vocc in ((0, mel.vocc_c), (1, mel.vocc_a)):
            for h in HPVXSEQ:
                for v in range(1, mel.njoined + 1):
                    if vocc[v - 1].get(h, 0) > 0:
                        out.append(fn(v, side, h))
        return out

    def op1_comp(v, s, h):
        k1, e1 = _merge_lists(m1, op1.njoined)[v - 1]
        return [x for k in k1 for x in labels('cnt', c.cnt_vocc, k, s, h)] + \
               [x for e in e1 for x in labels('ex1', c.ex1_vocc, e, s, h)]

    def op2_comp(v, s, h):
        k2, e2 = _merge_lists(m2, op2.njoined)[v - 1]
        return [x for k in k2 for x in labels('cnt', c.cnt_vocc, k, 1 - s, h)] + \
               [x for e in e2 for x in labels('ex2', c.ex2_vocc, e, s, h)]

    def res_comp(v, s, h):
        if s == 0:
            a, b = _merge_lists(m12, res.njoined)[v - 1]
            parts = [[x for e in a for x in labels('ex1', c.ex1_vocc, e, 0, h)],
                     [x for e in b for x in labels('ex2', c.ex2_vocc, e, 0, h)]]
            return parts[::-1] if ex2_first else parts
        a, b = _merge_lists(m21, res.njoined)[v - 1]
        return [[x for e in a for x in labels('ex2', c.ex2_vocc, e, 1, h)],
                [x for e in b for x in labels('ex1', c.ex1_vocc, e, 1, h)]]

    l1 = [x for comp in axes(op1, op1_comp) for x in comp]
    l2 = [x for comp in axes(op2, op2_comp) for x in comp]
    rc = axes(res, res_comp)
    lr = [x for comp in rc for part in comp for x in part]
    assert sorted(set(l1) | set(l2)) == sorted(set(lr) | (set(l1) & set(l2)))
    T = np.einsum(D1, l1, D2, l2, lr) if lr else np.array(np.einsum(D1, l1, D2, l2))
    for (part, v, side, h), ls in lab.items():
        if part == 'cnt':
            T = T / math.factorial(len(ls))
    groups, norm, pos = [], 1, 0
    for comp in rc:
        n = sum(len(p) for p in comp)
        if sum(1 for p in comp if p) > 1:
            groups.append(list(range(pos, pos + n)))
            for p in comp:
                norm *= math.factorial(len(p))
        pos += n
    return antisym(T, groups) / norm if groups else T


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 2, 3), (2, 3, 3)])
@pytest.mark.parametrize('name', list(JOINED_CASES))
@pytest.mark.parametrize('use_tr,route', [(False, 3), (True, 3), (False, 2)])
def test_joined_match_dense_einsum(nsym, nocc, nvir, name, use_tr, route):
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = JOINED_CASES[name](orbs)
    x1, x2 = spin_symmetric(a, 31), spin_symmetric(b, 32)
    D = dense_joined(c, to_dense(a, x1), to_dense(b, x2))
    ref = from_dense(r, D) if r.length > 1 or r.oc or r.oa else np.array([float(D)])
    assert np.abs(ref).max() > 1e-3
    xr = np.zeros(r.length)
    c.run(x1, x2, xr, route=route, use_tr=use_tr, screen=False)
    assert np.abs(xr - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_joined_dense_check_catches_a_wrong_merge_order():
    """EX2's creators placed before EX1's in the result's merged creator
    component change its sign: that dense check disagrees with route 3."""
    orbs = orbitals(3, 3, 2)
    c, a, b, r = JOINED_CASES['result_merge'](orbs)
    x1, x2 = spin_symmetric(a, 31), spin_symmetric(b, 32)
    xr = np.zeros(r.length)
    c.run(x1, x2, xr, route=3, use_tr=False, screen=False)
    ref = from_dense(r, dense_joined(c, to_dense(a, x1), to_dense(b, x2), ex2_first=True))
    assert np.abs(xr - ref).max() > 1e-3


def test_one_vertex_cases_agree_with_joined_form():
    """A one-vertex contraction written through Contraction.joined (one CNT
    vertex, trivial merge maps) enumerates the same kernel calls and gives
    the same result as the one-vertex constructor."""
    from oracle.melist import MEList
    from oracle.contract import Contraction
    orbs = orbitals(3, 4, 2)
    P, H = 2, 1
    T = MEList(orbs, {P: 2}, {H: 2})
    W = MEList(orbs, {H: 1, P: 1}, {H: 1, P: 1})
    R = MEList(orbs, {P: 2}, {H: 2})
    c1 = Contraction(T, W, R, {P: 1}, {H: 1})
    c2 = Contraction.joined(T, W, R, ex1=[({P: 1}, {H: 1})], ex2=[({P: 1}, {H: 1})],
                            cnt=[({P: 1}, {H: 1})], merge1=[1, 1, 1, 1], merge2=[1, 1, 1, 1],
                            merge12=[1, 1, 1, 1], merge21=[1, 1, 1, 1])
    assert len(c1.blocks(True)) == len(c2.blocks(True))
    x1, x2 = spin_symmetric(T, 41), spin_symmetric(W, 42)
    y1, y2 = np.zeros(R.length), np.zeros(R.length)
    c1.run(x1, x2, y1)
    c2.run(x1, x2, y2)
    assert np.abs(y1 - y2).max() <= 1e-14 * np.abs(y1).max()


def test_two_components_of_one_source_are_refused():
    """A vertex that contracts the same space with two vertices of the other
    operator would merge two CNT components into one operand component
    ('multi map needed'): refused, as route 3 keeps one string per target
    component and source."""
    from oracle.melist import MEList
    from oracle.contract import Contraction
    orbs = orbitals(2, 3, 1)
    P, H = 2, 1
    V = MEList(orbs, {H: 2}, {P: 2})
    X = MEList(orbs, [{P: 1}, {P: 1}], [{H: 1}, {H: 1}])
    E = MEList(orbs, {}, {})
    with pytest.raises(ValueError, match='multi map'):
        Contraction.joined(V, X, E, ex1=[({}, {})], ex2=[({}, {}), ({}, {})],
                           cnt=[({H: 1}, {P: 1}), ({H: 1}, {P: 1})],
                           merge1=[2, 1, 2, 1, 1], merge2=[1, 1, 1, 1, 1, 2, 1, 1],
                           merge12=[1, 1, 2, 1, 2], merge21=[2, 1, 2, 1, 1])


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 4), (2, 3, 4), (4, 4, 5)])
def test_joined_product_then_reorder_equals_one_vertex_product(nsym, nocc, nvir):
    """The reorder after a contraction (contract/contr_op1op2_wmaps_c.f:
    1962-1995): T1 x T1 into the two-vertex X(ai;bj), then the second
    vertex's lines moved onto the first (creators merge (I0, K), annihilators
    (K, I0)), equals the one-vertex product R(ab,ij) += T1(a,i) T1(b,j)
    (creators EX1 then EX2, annihilators EX2 then EX1); the two-vertex target
    with an empty second vertex has the one-vertex list's layout."""
    from oracle.melist import MEList
    from oracle.reorder import Reorder
    from models import t1t1
    orbs = orbitals(nocc, nvir, nsym)
    c, T1, _, X = JOINED_CASES['t1t1_joined'](orbs)
    P, H = 2, 1
    Xp = MEList(orbs, [{P: 2}, {}], [{H: 2}, {}])
    c1, _, _, R = t1t1(orbs)
    assert Xp.length == R.length
    x = spin_symmetric(T1, 51)
    xx = np.zeros(X.length)
    c.run(x, x, xx, use_tr=False, screen=False)
    y = Reorder(X, Xp, 1, 0, {P: 1}, {H: 1}).run(xx, np.zeros(Xp.length))
    ref = np.zeros(R.length)
    c1.run(x, x, ref, use_tr=False, screen=False)
    assert np.abs(ref).max() > 1e-3
    assert np.abs(y - ref).max() <= 1e-14 * np.abs(ref).max()
    # the emptied second vertex dropped: the one-vertex list itself as REO
    y1 = Reorder(X, R, 1, 0, {P: 1}, {H: 1}).run(xx, np.zeros(R.length))
    assert np.abs(y1 - ref).max() <= 1e-14 * np.abs(ref).max()
