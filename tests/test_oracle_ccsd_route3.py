"""The oracle's route-3 evaluation of the CCSD formula table
(oracle/ccsd_route3.py: the blocked restatement of contr_op1op2 with a BLAS
dgemm, the engine's formula and schedule) against the dense textbook CCSD
(oracle/ccsd_dense.py) on small synthetic Hamiltonians."""
import pytest

import ccsd
from synth import synthetic
from oracle.ccsd_dense import CCSD as Dense
from oracle.ccsd_route3 import Route3CCSD


@pytest.mark.parametrize('nocc,nvir,nsym,seed', [(2, 5, 1, 3), (3, 6, 2, 4), (4, 7, 4, 5)])
def test_route3_ccsd_equals_dense(nocc, nvir, nsym, seed):
    ham = synthetic(nocc, nvir, 3 * nvir, nsym=nsym, seed=seed, sigma=0.1)
    e_ref, *_ = Dense(ham).solve(1e-11)
    r = Route3CCSD(ham, ccsd.FORMULA, ccsd.SCHEDULE)
    e, it, rn = r.solve(tol=1e-10)
    assert rn < 1e-10
    assert abs(e - e_ref) < 1e-10, (e, e_ref)
    assert r.calls > 0 and r.flops > 0
