"""The CCSD formula table of the package, evaluated with the oracle's
contr_op1op2 restatement, reproduces the dense textbook CCSD residuals and
energy (CPU)."""
import numpy as np
import pytest

import ccsd
from synth import synthetic
from oracle.ccsd_dense import CCSD as DenseCCSD, fci_two_electron
from oracle.melist import Orbitals, MEList
from oracle.contract import Contraction, to_dense, from_dense, sym_ab

H, P = 1, 2


def oracle_lists(ham):
    orbs = Orbitals(ham.irrep_o, ham.irrep_v, ham.nsym)
    return orbs, {n: MEList(orbs, c, a) for n, (c, a) in ccsd.LISTS.items()}


def integral_vectors(L, g, o, v):
    X = {}
    for name in ccsd.INTEGRALS:
        Lst = L[name]
        axes = []
        for h in Lst.hc:
            axes += ['v' if h == P else 'o'] * Lst.occ_c[h]
        for h in Lst.ha:
            axes += ['v' if h == P else 'o'] * Lst.occ_a[h]
        X[name] = from_dense(Lst, g[tuple(v if a == 'v' else o for a in axes)])
    return X


def evaluate_with_oracle(L, X, use_tr=True):
    for n in ccsd.LISTS:
        if n in ccsd.INIT:
            src, f = ccsd.INIT[n]
            X[n] = f * X[src]
        elif n not in ccsd.INTEGRALS and n not in ('T1', 'T2'):
            X[n] = np.zeros(L[n].length)
    for lv, level in zip(ccsd.SCHEDULE, ccsd.LEVELS):
        for it in lv['add']:            # ADD items run before the contractions
            X[it['target']] = X[it['target']] + it['factor'] * X[it['source']]
        written = []
        for res, a, b, cc, ca, fact in level:
            Contraction(L[a], L[b], L[res], cc, ca, fact).run(X[a], X[b], X[res], use_tr=use_tr)
            # run() applies sym_ab after each contraction; applying it once per
            # level (as the engine does) is the same projector
            written.append(res)
    return X


@pytest.mark.parametrize('nocc,nvir,nsym', [(2, 4, 1), (3, 5, 2), (3, 6, 4)])
def test_formula_matches_dense_ccsd(nocc, nvir, nsym):
    ham = synthetic(nocc, nvir, 3 * nvir, nsym=nsym, seed=nocc + 10 * nvir, sigma=0.1)
    cc = DenseCCSD(ham)
    e, t1, t2, _ = cc.solve(1e-9)
    rng = np.random.default_rng(4)
    orbs, L = oracle_lists(ham)
    X = integral_vectors(L, cc.g, cc.o, cc.v)
    X['T1'] = from_dense(L['T1'], (t1 + 0.01 * rng.standard_normal(t1.shape)).T)
    sym_ab(L['T1'], X['T1'], 0.5, 1.0)
    X['T2'] = from_dense(L['T2'], t2.transpose(2, 3, 0, 1))
    t1 = to_dense(L['T1'], X['T1']).T
    r1, r2 = cc.residual(t1, t2)
    evaluate_with_oracle(L, X)
    ref1 = from_dense(L['R1'], r1.T)
    ref2 = from_dense(L['R2'], r2.transpose(2, 3, 0, 1))
    assert np.abs(X['R1'] - ref1).max() <= 1e-12 * max(1, np.abs(ref1).max())
    assert np.abs(X['R2'] - ref2).max() <= 1e-12 * max(1, np.abs(ref2).max())
    assert abs(X['E'][0] - cc.energy(t1, t2)) <= 1e-13


def test_dense_ccsd_is_exact_for_two_electrons():
    for nsym in (1, 2, 4):
        ham = synthetic(1, 6, 12, nsym=nsym, seed=3, sigma=0.15)
        e, *_ = DenseCCSD(ham).solve(1e-12)
        assert abs(e - fci_two_electron(ham)) < 1e-12


def test_dense_ccsd_reproduces_reference_ch2_energy():
    """Known answer of the reference's own test: CH2/DZ (DALTON files of
    test/molecules/CH2-DZ-DALTON, frozen 1s), CCSD total energy
    -38.907618798245 (test/CCPT-energy.e.ref:690, printed to 1e-12)."""
    import json
    import os
    from dalton import hamiltonian_from_dalton
    gold = os.path.join(os.path.dirname(__file__), 'golden')
    ham = hamiltonian_from_dalton(os.path.join(gold, 'ch2_dz_SIRIFC.gz'),
                                  os.path.join(gold, 'ch2_dz_MOTWOINT.gz'))
    assert (ham.nocc, ham.nvir, ham.nsym) == (3, 9, 4)
    assert abs(ham.e_ref - (-38.821908746679)) < 1e-11      # E(SCF) of SIRIFC
    e, *_ = DenseCCSD(ham).solve(1e-11)
    ref = json.load(open(os.path.join(gold, 'ch2_dz.json')))['e_ccsd_ref']
    assert abs(ham.e_ref + e - ref) < 5e-12
