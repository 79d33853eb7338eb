"""CPU emulation of a plan's kernel calls, with the device's address
arithmetic, against the oracle's contr_op1op2 restatement.

Each kernel call R(m,n) += xfac * sum_k A(k,m) B(k,n) is evaluated in numpy
from the plan's resolved index data (gecco_hip_plan_block_index) and operand
slots/offsets (gecco_hip_plan_block_info), reading a separable index vector
that the plan marks affine as 1 + a0 + a1*i, as the kernels do.  Every
address must fall inside its list, and the emulated result must equal the
oracle (two-body CCSD cases and the three-body CCSDT triples terms).
Tolerance: FP64, 1e-12 relative (summation order)."""
import numpy as np
import pytest

from models import CASES, T3_CASES, round_robin, spin_symmetric
from oracle.contract import sym_ab
from oracle.melist import Orbitals


def _operand(plan, i, which, info, length, F, S):
    """ok, address and sign of operand `which` over its F x S index grid."""
    first, second = plan.block_index(i, which)
    if second is None:
        v = first.astype(np.int64)
        return v != 0, np.abs(v) - 1, v < 0
    aff = info[10 + 2 * which], info[11 + 2 * which]

    def vec(x, affine):
        x = x.astype(np.int64)
        if affine:                      # what the kernels compute
            a0, a1 = x[0] - 1, (x[1] - x[0]) if len(x) > 1 else 0
            return np.ones(len(x), bool), a0 + a1 * np.arange(len(x)), np.zeros(len(x), bool)
        return x != 0, np.abs(x) - 1, x < 0
    ok1, a1, n1 = vec(first, aff[0])
    ok2, a2, n2 = vec(second, aff[1])
    return (ok1[:, None] & ok2[None, :], a1[:, None] + a2[None, :], n1[:, None] != n2[None, :])


def emulate(plan, nb, bufs):
    for i in range(nb):
        info = plan.block_info(i)
        M, N, K = info[0], info[1], info[2]
        (ba, oa), (bb, ob), xfac = plan.block_operands(i)
        bc, oc = info[20], info[21]
        vals = []
        for which, (buf, off, F, S) in enumerate(((ba, oa, K, M), (bb, ob, K, N))):
            ok, ad, ng = _operand(plan, i, which, info, len(bufs[buf]), F, S)
            assert ad.shape == (F, S)
            adr = off + ad[ok]
            assert (adr >= 0).all() and (adr < len(bufs[buf])).all(), (i, which)
            x = np.zeros((F, S))
            x[ok] = bufs[buf][adr]
            vals.append(np.where(ng, -x, x))
        ok, ad, ng = _operand(plan, i, 2, info, len(bufs[bc]), M, N)
        adr = oc + ad[ok]
        assert (adr >= 0).all() and (adr < len(bufs[bc])).all(), (i, 'C')
        res = xfac * (vals[0].T @ vals[1])
        np.add.at(bufs[bc], adr, np.where(ng, -res, res)[ok])


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (2, 3, 4), (2, 4, 6), (4, 5, 9)])
@pytest.mark.parametrize('name', list(CASES) + list(T3_CASES))
@pytest.mark.parametrize('use_tr', [True, False])
def test_plan_emulation_matches_oracle(nsym, nocc, nvir, name, use_tr):
    import gecco_hip
    cases = {**CASES, **T3_CASES}
    oo = Orbitals(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    c, a, b, r = cases[name](oo)
    x1, x2, x0 = spin_symmetric(a, 71), spin_symmetric(b, 72), spin_symmetric(r, 73)
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=use_tr, screen=False)
    p = gecco_hip.Plan()
    A, B, R = (gecco_hip.MEList(o, m.occ_c, m.occ_a) for m in (a, b, r))
    nb, _ = p.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, use_tr)
    bufs = [x1.copy(), x2.copy(), x0.copy()]
    emulate(p, nb, bufs)
    if use_tr:
        sym_ab(r, bufs[2], 0.5, float(r.absym))
    assert np.abs(bufs[2] - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (2, 3, 4), (4, 5, 9)])
@pytest.mark.parametrize('name', ['energy_t1', 'ring_t2', 't1_ovoo_t2', 'ladder_t12', 'fock_t12', 'vvvo_t1'])
@pytest.mark.parametrize('use_tr', [True, False])
def test_transposed_operands_emulate_the_oracle(nsym, nocc, nvir, name, use_tr):
    """Operands and results given as adjoint views (gecco_hip_list_adjoint):
    the product's enumerator records the oracle's kernel calls (count and
    FLOP) with the transpose flags, and its resolved addresses -- the
    transposed branch of set_op_ldim_c in the plan -- reproduce the oracle."""
    import gecco_hip
    from models import TRA_CASES
    oo = Orbitals(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    c, stored, tra = TRA_CASES[name](oo)
    xs = [spin_symmetric(m, 81 + k) for k, m in enumerate(stored)]
    ref = xs[2].copy()
    c.run(xs[0], xs[1], ref, route=3, use_tr=use_tr, screen=False)
    lists = [gecco_hip.MEList(o, m.occ_c, m.occ_a) for m in stored]
    used = [L.adjoint() if t else L for L, t in zip(lists, tra)]
    p = gecco_hip.Plan()
    nb, fl = p.add_contraction(used[0], used[1], used[2], c.cnt_c, c.cnt_a, 1.0, use_tr)
    assert nb == len(c.blocks(use_tr)) and fl == c.flops(use_tr)
    bufs = [x.copy() for x in xs]
    emulate(p, nb, bufs)
    if use_tr:
        sym_ab(stored[2], bufs[2], 0.5, float(stored[2].absym))
    assert np.abs(bufs[2] - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def _joined_plan(gecco_hip, o, c, lists, use_tr):
    A, B, R = (gecco_hip.MEList(o, m.vocc_c, m.vocc_a) for m in lists)
    p = gecco_hip.Plan()
    m1, m2, m12, m21 = c.merge
    nb, fl = p.add_contraction_n(A, B, R, c.ex1_vocc, c.ex2_vocc, c.cnt_vocc, m1, m2, m12, m21,
                                 1.0, use_tr)
    return p, nb, fl


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 2, 3), (2, 3, 4), (4, 4, 5)])
@pytest.mark.parametrize('name', ['t1t1_joined', 'vertex_fock', 'pairwise_scalar', 'ring_one_vertex',
                                  'result_merge'])
@pytest.mark.parametrize('use_tr', [True, False])
def test_joined_contractions_emulate_the_oracle(nsym, nocc, nvir, name, use_tr):
    """Operators with several vertices (gecco_hip_plan_add_contraction_n):
    the product records the oracle's kernel calls (count and FLOP), and its
    resolved addresses reproduce the oracle's route 3."""
    import gecco_hip
    from models import JOINED_CASES
    oo = Orbitals(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    c, a, b, r = JOINED_CASES[name](oo)
    x1, x2, x0 = spin_symmetric(a, 91), spin_symmetric(b, 92), spin_symmetric(r, 93)
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=use_tr, screen=False)
    p, nb, fl = _joined_plan(gecco_hip, o, c, (a, b, r), use_tr)
    assert nb == len(c.blocks(use_tr)) and fl == c.flops(use_tr)
    bufs = [x1.copy(), x2.copy(), x0.copy()]
    emulate(p, nb, bufs)
    if use_tr:
        sym_ab(r, bufs[2], 0.5, float(r.absym))
    assert np.abs(bufs[2] - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_joined_contraction_refuses_a_multi_map():
    """A vertex contracting one space with two vertices of the other
    operator needs two CNT components merged into one operand component:
    refused with the reference's 'multi map needed'."""
    import gecco_hip
    o = gecco_hip.Orbs(round_robin(2, 1), round_robin(3, 1, 1), 1)
    H, P = 1, 2
    V = gecco_hip.MEList(o, {H: 2}, {P: 2})
    X = gecco_hip.MEList(o, [{P: 1}, {P: 1}], [{H: 1}, {H: 1}])
    E = gecco_hip.MEList(o, {}, {})
    p = gecco_hip.Plan()
    with pytest.raises(gecco_hip.GeccoHipError, match='multi map'):
        p.add_contraction_n(V, X, E, [({}, {})], [({}, {}), ({}, {})],
                            [({H: 1}, {P: 1}), ({H: 1}, {P: 1})], [2, 1, 2, 1, 1],
                            [1, 1, 1, 1, 1, 2, 1, 1], [1, 1, 2, 1, 2], [2, 1, 2, 1, 1])
    with pytest.raises(gecco_hip.GeccoHipError, match='vertex'):
        p.add_contraction_n(V, X, E, [({}, {})], [({}, {}), ({}, {})],
                            [({H: 1}, {P: 1})], [1, 1, 1, 1],
                            [1, 1, 1, 1, 0, 1, 2], [1, 1, 2, 1, 2], [2, 1, 2, 1, 1])
