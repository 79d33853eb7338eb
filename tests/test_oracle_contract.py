"""Oracle pins for the binary contraction (CPU): dense antisymmetrised
einsum, route 2 == route 3, blocking invariance, spin-flip on/off."""
import itertools

import numpy as np
import pytest

from oracle.contract import to_dense, from_dense, _perm_sign
from models import orbitals, spin_symmetric, CASES, ladder, ring, t1t1


def antisym(Y, groups):
    out = np.zeros_like(Y)
    for combo in itertools.product(*[list(itertools.permutations(g)) for g in groups]):
        perm = list(range(Y.ndim))
        s = 1
        for g, p in zip(groups, combo):
            s *= _perm_sign([g.index(q) for q in p])
            for a, b in zip(g, p):
                perm[a] = b
        out += s * Y.transpose(perm)
    return out


def dense_ref(name, D1, D2):
    if name == 'ladder':
        return 0.5 * np.einsum('abcd,cdij->abij', D1, D2)
    if name == 'ring':
        return antisym(np.einsum('caki,bkcj->abji', D1, D2), [[0, 1], [2, 3]])
    if name == 't1t1':
        return antisym(np.einsum('ai,bj->abji', D1, D2), [[0, 1], [2, 3]])
    raise KeyError(name)


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 4), (2, 3, 4), (4, 4, 5)])
@pytest.mark.parametrize('name', ['ladder', 'ring', 't1t1'])
@pytest.mark.parametrize('use_tr', [False, True])
@pytest.mark.parametrize('route', [3, 2])
def test_matches_dense_einsum(nsym, nocc, nvir, name, use_tr, route):
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = CASES[name](orbs)
    x1 = spin_symmetric(a, 11)
    x2 = spin_symmetric(b, 12)
    ref = from_dense(r, dense_ref(name, to_dense(a, x1), to_dense(b, x2)))
    xr = np.zeros(r.length)
    c.run(x1, x2, xr, route=route, use_tr=use_tr)
    assert np.abs(xr - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def test_dense_check_catches_a_sign_mistake():
    """Flipping one map sign in the oracle's input must break the pin."""
    orbs = orbitals(3, 4, 2)
    c, a, b, r = ring(orbs)
    x1 = spin_symmetric(a, 11)
    x2 = spin_symmetric(b, 12)
    ref = from_dense(r, dense_ref('ring', to_dense(a, x1), to_dense(b, x2)))
    blocks = c.blocks(False)
    for bl in blocks:
        m = bl['map_ex1cntc']
        nz = [i for i, v in enumerate(m) if v]
        if nz:
            m[nz[0]] = -m[nz[0]]
    xr = np.zeros(r.length)
    c.run(x1, x2, xr, use_tr=False, blocks=blocks)
    assert np.abs(xr - ref).max() > 1e-6


@pytest.mark.parametrize('name', list(CASES))
def test_blocking_invariance(name):
    """test/batch-*.inp style: str_block / cnt_block must not change results."""
    orbs = orbitals(4, 6, 2)
    c, a, b, r = CASES[name](orbs)
    x1 = spin_symmetric(a, 3)
    x2 = spin_symmetric(b, 4)
    outs = []
    for lb, lc in [(200, -1), (5, 3), (2, 1), (7, -1)]:
        xr = np.zeros(r.length)
        c.run(x1, x2, xr, lenblock=lb, lenblock_cnt=lc)
        outs.append(xr)
    for o in outs[1:]:
        assert np.abs(o - outs[0]).max() <= 1e-13 * max(1.0, np.abs(outs[0]).max())


@pytest.mark.parametrize('name', list(CASES))
def test_route2_equals_route3_and_use_tr(name):
    orbs = orbitals(4, 5, 4)
    c, a, b, r = CASES[name](orbs)
    x1 = spin_symmetric(a, 5)
    x2 = spin_symmetric(b, 6)
    outs = []
    for route, ut in [(3, True), (2, True), (3, False), (2, False)]:
        xr = np.zeros(r.length)
        c.run(x1, x2, xr, route=route, use_tr=ut)
        outs.append(xr)
    for o in outs[1:]:
        assert np.abs(o - outs[0]).max() <= 1e-13 * max(1.0, np.abs(outs[0]).max())


def test_ladder_flop_count():
    """nocc=8, nvir=32, C1 ladder: 2*(496*28*496) + 2*(1024*64*1024) with
    the spin-flip skip (contract/contr_op1op2_wmaps_c.f:1007-1019)."""
    c, *_ = ladder(orbitals(8, 32, 1))
    assert c.flops(use_tr=True) == 147994624
    assert c.flops(use_tr=False) == 2 * (2 * 496 * 28 * 496 + 1024 * 64 * 1024)


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 4), (2, 3, 5), (4, 4, 5)])
@pytest.mark.parametrize('name', sorted(__import__('models').TRA_CASES))
@pytest.mark.parametrize('use_tr', [False, True])
@pytest.mark.parametrize('route', [3, 2])
def test_transposed_operands_equal_explicit_transposes(nsym, nocc, nvir, name, use_tr, route):
    """A kernel call with tra_op1 / tra_op2 / tra_op1op2 (the reference's
    dagger flags, contract/contr_op1op2_wmaps_c.f:844-855, addressed by the
    transposed branch of contract/set_op_ldim_c.f) gives what the same
    contraction gives on explicitly transposed lists (each stored element
    (a; c) placed at (c; a) of a list of the swapped occupation, through the
    dense antisymmetric form).  The non-transposed contractions are pinned
    by the dense einsums above."""
    from models import TRA_CASES, transposed_data
    from oracle.contract import Contraction
    from oracle.melist import MEList
    orbs = orbitals(nocc, nvir, nsym)
    c, stored, tra = TRA_CASES[name](orbs)
    assert any(b['tra_op1'] or b['tra_op2'] or b['tra_op1op2'] for b in c.blocks(use_tr)) or \
        not c.blocks(use_tr)
    xs = [spin_symmetric(m, 21 + k) for k, m in enumerate(stored)]
    xs[2] = spin_symmetric(stored[2], 31, 0.1)           # results accumulate
    ops, xe = [], []
    for m, x, t in zip(stored, xs, tra):
        if t:
            L, y = transposed_data(m, x)
        else:
            L, y = m, x.copy()
        ops.append(L)
        xe.append(y)
    ce = Contraction(ops[0], ops[1], ops[2], c.cnt_c, c.cnt_a, c.fact)
    got = xs[2].copy()
    c.run(xs[0], xs[1], got, route=route, use_tr=use_tr)
    ce.run(xe[0], xe[1], xe[2], route=route, use_tr=use_tr)
    ref = transposed_data(ops[2], xe[2])[1] if tra[2] else xe[2]
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), np.abs(got - ref).max()
    assert np.abs(ref - xs[2]).max() > 1e-3 * np.abs(ref).max()    # the contraction added something


def test_transposed_layout_differs_from_plain():
    """The transposed leading dimensions matter: reading a stored list through
    its adjoint with the transpose flag cleared gives a different result
    (so the check above would catch a dropped flag)."""
    from models import TRA_CASES
    from oracle.contract import Contraction
    orbs = orbitals(3, 5, 2)
    c, stored, tra = TRA_CASES['ring_t2'](orbs)
    xs = [spin_symmetric(m, 41 + k) for k, m in enumerate(stored)]
    got = np.zeros(stored[2].length)
    c.run(xs[0], xs[1], got, route=3, use_tr=False)
    blocks = [dict(b, tra_op2=0) for b in c.blocks(False)]
    bad = np.zeros(stored[2].length)
    c.run(xs[0], xs[1], bad, route=3, use_tr=False, blocks=blocks)
    assert np.abs(got - bad).max() > 1e-3 * np.abs(got).max()


def test_collect_scatter_hook_is_used():
    """The route-3 test hook (oracle.contract.use_collect_scatter) really
    replaces the collect and scatter steps: with a collect that leaves each
    block at zero, or a scatter that adds nothing, the contraction adds
    nothing; with the hook cleared, it adds the oracle's result."""
    import ctypes
    from models import CASES, orbitals, spin_symmetric
    from oracle.contract import use_collect_scatter
    c, a, b, r = CASES['ladder'](orbitals(4, 6, 2))
    x1, x2 = spin_symmetric(a, 1), spin_symmetric(b, 2)
    P = ctypes.c_void_p
    calls = {'c': 0, 's': 0}

    def coll(xblk, *args):
        calls['c'] += 1                   # xblk keeps its zeros

    def scat(*args):
        calls['s'] += 1
    cf = ctypes.CFUNCTYPE(None, *([P] * 39))(coll)
    sf = ctypes.CFUNCTYPE(None, *([P] * 42))(scat)
    for hook in ((cf, None), (None, sf)):
        xr = np.zeros(r.length)
        try:
            use_collect_scatter(ctypes.cast(hook[0], P).value if hook[0] else None,
                                ctypes.cast(hook[1], P).value if hook[1] else None)
            c.run(x1, x2, xr, route=3, use_tr=False, screen=False)
        finally:
            use_collect_scatter(None, None)
        if hook[0]:
            assert calls['c'] > 0 and not xr.any()   # zero blocks: nothing added
        else:
            assert calls['s'] > 0 and not xr.any()
    xr = np.zeros(r.length)
    c.run(x1, x2, xr, route=3, use_tr=False, screen=False)
    assert np.abs(xr).max() > 0
