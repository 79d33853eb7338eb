"""Pins of the trace_op oracle against the dense partial trace of the
antisymmetric operator tensors (CPU)."""
import numpy as np
import pytest

from models import orbitals, spin_symmetric
from oracle.melist import MEList
from oracle.contract import to_dense
from oracle.trace import Trace, dense_trace

H, P = 1, 2

CASES = {
    # op (C, A), J occupation
    'pp_over_p': (({P: 2}, {P: 2}), {P: 1}),
    'pp_full': (({P: 2}, {P: 2}), {P: 2}),
    'hp_over_h': (({H: 1, P: 1}, {H: 1, P: 1}), {H: 1}),
    'hp_full': (({H: 1, P: 1}, {H: 1, P: 1}), {H: 1, P: 1}),
    'ph_over_p': (({P: 2}, {H: 1, P: 1}), {P: 1}),
    'p_full': (({P: 1}, {P: 1}), {P: 1}),
}


@pytest.mark.parametrize('nsym', [1, 2, 4])
@pytest.mark.parametrize('name', list(CASES))
def test_trace_matches_dense(name, nsym):
    (occ_c, occ_a), cnt = CASES[name]
    orbs = orbitals(3, 4, nsym)
    op = MEList(orbs, occ_c, occ_a)
    tr_c = {h: n - cnt.get(h, 0) for h, n in occ_c.items() if n - cnt.get(h, 0)}
    tr_a = {h: n - cnt.get(h, 0) for h, n in occ_a.items() if n - cnt.get(h, 0)}
    trop = MEList(orbs, tr_c, tr_a)
    x = spin_symmetric(op, 7)
    t = Trace(op, trop, cnt, fact=0.75)
    got = t.run(x, np.zeros(trop.length))
    ref = 0.75 * dense_trace(to_dense(op, x), op, trop, cnt)
    if trop.length == 1 and not trop.oc and not trop.oa:
        np.testing.assert_allclose(got[0], ref, rtol=1e-13, atol=1e-13)
    else:
        np.testing.assert_allclose(to_dense(trop, got), ref, rtol=0, atol=1e-12)
    assert np.abs(got).max() > 0


def test_full_trace_of_one_body_is_sum_of_diagonal():
    """Closed form: the trace of a one-body operator over P is the sum of its
    diagonal elements (both spins)."""
    orbs = orbitals(2, 5, 2)
    op = MEList(orbs, {P: 1}, {P: 1})
    tr = MEList(orbs, {}, {})
    x = spin_symmetric(op, 3)
    D = to_dense(op, x)
    got = Trace(op, tr, {P: 1}).run(x, np.zeros(1))[0]
    assert abs(got - np.trace(D)) < 1e-13


@pytest.mark.parametrize('nsym', [1, 4])
@pytest.mark.parametrize('name', list(CASES))
def test_product_trace_enumeration_counts(name, nsym):
    """The product enumerator (host only, no device work) produces as many
    element terms as the oracle restatement."""
    import gecco_hip
    from models import round_robin
    (occ_c, occ_a), cnt = CASES[name]
    orbs = orbitals(3, 4, nsym)
    op = MEList(orbs, occ_c, occ_a)
    tr_c = {h: n - cnt.get(h, 0) for h, n in occ_c.items() if n - cnt.get(h, 0)}
    tr_a = {h: n - cnt.get(h, 0) for h, n in occ_a.items() if n - cnt.get(h, 0)}
    trop = MEList(orbs, tr_c, tr_a)
    o = gecco_hip.Orbs(round_robin(3, nsym), round_robin(4, nsym, 1), nsym)
    A = gecco_hip.MEList(o, occ_c, occ_a)
    T = gecco_hip.MEList(o, tr_c, tr_a)
    p = gecco_hip.Plan()
    n = p.add_trace(A, T, cnt, 1.0)
    assert n == len(Trace(op, trop, cnt).triples())
