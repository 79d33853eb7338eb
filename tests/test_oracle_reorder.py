"""Pins of the reorder oracle (CPU): dense axis permutation of the
antisymmetric tensors, identity of a move there and back, and the
multi-vertex layout (condense_occ order)."""
import numpy as np
import pytest

from models import orbitals, spin_symmetric
from oracle.melist import MEList, condense_n
from oracle.contract import to_dense
from oracle.reorder import Reorder, dense_reorder

H, P = 1, 2

# (ORI vertices C, ORI vertices A, from, to, K C, K A)
CASES = {
    'p_to_v1': ([{P: 1}, {P: 1}], [{H: 1}, {H: 1}], 0, 1, {P: 1}, {}),
    'h_to_v0': ([{P: 1}, {P: 1}], [{H: 1}, {H: 1}], 1, 0, {}, {H: 1}),
    'ph_merge': ([{P: 1}, {P: 1}], [{H: 1}, {H: 1}], 0, 1, {P: 1}, {H: 1}),
    'mixed': ([{H: 1, P: 1}, {P: 1}], [{P: 1, H: 1}, {H: 1}], 0, 1, {H: 1}, {}),
    'split': ([{P: 2}, {}], [{H: 1}, {H: 1}], 0, 1, {P: 1}, {}),
    'p_to_empty': ([{P: 1}, {}], [{H: 1}, {}], 0, 1, {P: 1}, {}),
}


def _reo_occ(oc, oa, frm, to, kc, ka):
    rc = [dict(x) for x in oc]
    ra = [dict(x) for x in oa]
    for h, n in kc.items():
        rc[frm][h] -= n
        rc[to][h] = rc[to].get(h, 0) + n
    for h, n in ka.items():
        ra[frm][h] -= n
        ra[to][h] = ra[to].get(h, 0) + n
    return rc, ra


def test_odd_parity_rejected():
    with pytest.raises(ValueError):
        MEList(orbitals(2, 3, 1), [{H: 1, P: 1}, {P: 1}], [{P: 1}, {H: 1}])


def test_condense_order_vertex_inside_space():
    oc, hc, vc, oa, ha, va = condense_n([{P: 1, H: 1}, {P: 2}], [{H: 1}, {P: 1, H: 1}])
    assert (oc, hc, vc) == ([1, 2, 1], [P, P, H], [0, 1, 0])
    assert (oa, ha, va) == ([1, 1, 1], [P, H, H], [1, 0, 1])


@pytest.mark.parametrize('nsym', [1, 2, 4])
@pytest.mark.parametrize('name', list(CASES))
def test_reorder_matches_dense(name, nsym):
    oc, oa, frm, to, kc, ka = CASES[name]
    orbs = orbitals(2, 3, nsym)
    ori = MEList(orbs, oc, oa)
    rc, ra = _reo_occ(oc, oa, frm, to, kc, ka)
    reo = MEList(orbs, rc, ra)
    ro = Reorder(ori, reo, frm, to, kc, ka, fact=0.5, sign_reo=-1)
    x = np.random.default_rng(1).standard_normal(ori.length)
    got = ro.run(x, np.zeros(reo.length))
    ref = dense_reorder(to_dense(ori, x), ori, reo, ro)
    np.testing.assert_allclose(to_dense(reo, got), ref, rtol=0, atol=1e-12)
    assert np.abs(got).max() > 0


@pytest.mark.parametrize('name', ['mixed', 'p_to_empty'])
def test_reorder_there_and_back(name):
    """When K meets no string of its space on the to-vertex, the reorder is a
    signed bijection: ORI -> REO -> ORI is the identity (sign_reo^2 = 1)."""
    oc, oa, frm, to, kc, ka = CASES[name]
    orbs = orbitals(2, 3, 2)
    ori = MEList(orbs, oc, oa)
    rc, ra = _reo_occ(oc, oa, frm, to, kc, ka)
    reo = MEList(orbs, rc, ra)
    for h in kc:
        assert reo.vocc_c[to][h] == kc[h]
    for h in ka:
        assert reo.vocc_a[to][h] == ka[h]
    x = spin_symmetric(ori, 5)
    y = Reorder(ori, reo, frm, to, kc, ka, sign_reo=-1).run(x, np.zeros(reo.length))
    back = Reorder(reo, ori, to, frm, kc, ka, sign_reo=-1).run(y, np.zeros(ori.length))
    np.testing.assert_allclose(back, x, rtol=0, atol=1e-13)


@pytest.mark.parametrize('nsym', [1, 4])
@pytest.mark.parametrize('name', list(CASES))
def test_product_reorder_enumeration_counts(name, nsym):
    """The product enumerator (host only) produces as many element terms as
    the oracle restatement."""
    import gecco_hip
    from models import round_robin
    oc, oa, frm, to, kc, ka = CASES[name]
    orbs = orbitals(2, 3, nsym)
    ori = MEList(orbs, oc, oa)
    rc, ra = _reo_occ(oc, oa, frm, to, kc, ka)
    reo = MEList(orbs, rc, ra)
    o = gecco_hip.Orbs(round_robin(2, nsym), round_robin(3, nsym, 1), nsym)
    A = gecco_hip.MEList(o, oc, oa)
    B = gecco_hip.MEList(o, rc, ra)
    assert (A.length, B.length) == (ori.length, reo.length)
    p = gecco_hip.Plan()
    n = p.add_reorder(A, B, frm, to, kc, ka, 1.0, -1)
    assert n == len(Reorder(ori, reo, frm, to, kc, ka).triples())
