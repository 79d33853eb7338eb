"""Brute-force Fock-space coupled cluster (oracle/cc_fock.py) and the
generated CC formulas (gecco-public_amd/ccgen.py), on the CPU.

The Fock-space residual <mu|e^-T H e^T|0> is pinned by the textbook CCSD
(oracle/ccsd_dense.py: its solution makes it vanish, with the same energy)
and by the hand-factorised CCSD formula table (gecco-public_amd/ccsd.py).
It then pins the generated CCSD and CCSDT formulas term by term at random
amplitudes, and a planted sign error in one generated term is caught.  The
generated CCSDT reproduces the reference's CCSDT known answer for CH2/DZ
(test/CC-energy-I.e.ref:606, frozen (1,0,0,0), deleted (2,1,0,0))."""
import os

import numpy as np
import pytest

import ccgen
import ccsd
import formula as F
from synth import synthetic
from oracle.cc_fock import FockSpace
from oracle.ccsd_dense import CCSD as Dense
from oracle.ccsd_route3 import Route3CCSD, diagonal
from oracle.contract import sym_ab

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _random_amps(r, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    amps = []
    for n in r.tnames:
        x = scale * rng.standard_normal(r.L[n].length)
        sym_ab(r.L[n], x, 0.5, 1.0)
        r.X[n] = x
        amps.append(x)
    return amps


def _compare(r, fs, seed):
    amps = _random_amps(r, seed)
    om = r.evaluate()
    e, Rs = fs.residuals([r.L[n] for n in r.tnames], amps, [r.L[n] for n in r.rnames])
    return float(abs(r.X['E'][0] - e)), float(np.abs(om - np.concatenate(Rs)).max()), \
        float(np.abs(np.concatenate(Rs)).max())


def test_fock_space_residual_vanishes_at_textbook_ccsd():
    ham = synthetic(1, 3, 9, nsym=1, seed=3, sigma=0.2)
    d = Dense(ham)
    e_ref, t1, t2, _ = d.solve(1e-12)
    fs = FockSpace(ham)
    r = Route3CCSD(ham, ccsd.FORMULA, ccsd.SCHEDULE)
    r.solve(tol=1e-12)
    e, Rs = fs.residuals([r.L['T1'], r.L['T2']], [r.X['T1'], r.X['T2']], [r.L['R1'], r.L['R2']])
    assert abs(e - e_ref) < 1e-11
    assert max(np.abs(x).max() for x in Rs) < 1e-10


@pytest.mark.parametrize('nocc,nvir,nsym', [(1, 2, 1), (2, 2, 1), (1, 3, 2)])
def test_hand_and_generated_ccsd_equal_fock_space(nocc, nvir, nsym):
    ham = synthetic(nocc, nvir, 3 * nvir, nsym=nsym, seed=5, sigma=0.2)
    fs = FockSpace(ham)
    for f, sch in ((ccsd.FORMULA, ccsd.SCHEDULE), (None, None)):
        if f is None:
            f = ccgen.generate(2)
            sch = F.schedule(f['items'])
        de, dr, scale = _compare(Route3CCSD(ham, f, sch), fs, 11)
        assert de < 1e-13 and dr < 1e-13 * max(1.0, scale), (de, dr)


@pytest.mark.parametrize('nocc,nvir,nsym', [(2, 2, 1), (2, 2, 2)])
def test_generated_ccsdt_equals_fock_space(nocc, nvir, nsym):
    ham = synthetic(nocc, nvir, 3 * nvir, nsym=nsym, seed=6, sigma=0.2)
    fs = FockSpace(ham)
    f = ccgen.generate(3)
    r = Route3CCSD(ham, f, F.schedule(f['items']))
    assert r.tnames == ['T1', 'T2', 'T3']
    de, dr, scale = _compare(r, fs, 12)
    assert de < 1e-13 and dr < 1e-13 * max(1.0, scale), (de, dr)


def test_planted_sign_error_is_caught():
    ham = synthetic(2, 2, 6, nsym=1, seed=6, sigma=0.2)
    fs = FockSpace(ham)
    f = ccgen.generate(3)
    items = [dict(it) for it in f['items']]
    k = next(i for i, it in enumerate(items) if it['cmd'] == 'CONTR' and it['target'] == 'R3'
             and it['op2'] == 'T3')
    items[k]['factor'] = -items[k]['factor']
    g = dict(f, items=items)
    de, dr, scale = _compare(Route3CCSD(ham, g, F.schedule(items)), fs, 12)
    assert dr > 1e-3 * scale


def test_generated_ccsd_energy_equals_textbook():
    ham = synthetic(3, 6, 18, nsym=2, seed=4, sigma=0.1)
    e_ref, *_ = Dense(ham).solve(1e-11)
    f = ccgen.generate(2)
    e, it, rn = Route3CCSD(ham, f, F.schedule(f['items'])).solve(tol=1e-10)
    assert abs(e - e_ref) < 1e-10


def test_generated_ccsdt_reproduces_the_reference_known_answer():
    """CH2/DZ CCSDT with frozen (1,0,0,0) and deleted (2,1,0,0) orbitals:
    -38.859570692641 Eh (test/CC-energy-I.e.ref:606, converged to 4.5e-11)."""
    import json
    from dalton import hamiltonian_from_dalton
    meta = json.load(open(os.path.join(GOLD, 'ch2_dz.json')))
    ham = hamiltonian_from_dalton(os.path.join(GOLD, 'ch2_dz_SIRIFC.gz'),
                                  os.path.join(GOLD, 'ch2_dz_MOTWOINT.gz'), frozen=(1, 0, 0, 0),
                                  deleted=(2, 1, 0, 0))
    assert (ham.nocc, ham.nvir) == (3, 6)
    f = ccgen.generate(3)
    e, it, rn = Route3CCSD(ham, f, F.schedule(f['items'])).solve(tol=1e-11)
    assert abs(ham.e_ref + e - meta['e_ccsdt_ref']) < 1e-10, ham.e_ref + e
