"""Pins of the DIIS/preconditioner oracle (CPU), independent of its own
formulas: Jacobi limit, constrained least squares, finite termination on
linear problems, divide threshold, and the list diagonal of the product
against the per-electron sum."""
import numpy as np
import pytest

from oracle.diis import DIIS, diavc, list_diagonal


def test_single_vector_step_is_jacobi():
    rng = np.random.default_rng(0)
    t = rng.standard_normal(50)
    g = rng.standard_normal(50)
    d = rng.uniform(1, 3, 50)
    new = DIIS(8).step(t, g, d)
    np.testing.assert_allclose(new, t - g / d, rtol=0, atol=1e-15)


def test_weights_are_constrained_least_squares():
    """w minimises ||sum w_i r_i|| subject to sum w_i = 1: solve the same
    problem through the null space of the constraint with lstsq."""
    rng = np.random.default_rng(1)
    n, m = 40, 5
    dd = DIIS(8)
    d = rng.uniform(1, 2, n)
    for _ in range(m):
        dd.step(rng.standard_normal(n), rng.standard_normal(n), d)
    R = np.array(dd.r).T                     # n x m
    N = np.zeros((m, m - 1))
    for i in range(m - 1):
        N[i, i] = 1.0
        N[m - 1, i] = -1.0
    e = np.zeros(m)
    e[m - 1] = 1.0
    c, *_ = np.linalg.lstsq(R @ N, -R @ e, rcond=None)
    np.testing.assert_allclose(dd.w, e + N @ c, rtol=0, atol=1e-10)
    assert abs(dd.w.sum() - 1.0) < 1e-13


@pytest.mark.parametrize('n', [3, 6])
def test_linear_problem_terminates(n):
    """g(t) = A t - b with A = D + E: with the full history DIIS reaches the
    solution after at most n + 1 gradients."""
    rng = np.random.default_rng(n)
    d = rng.uniform(2, 4, n)
    E = 0.3 * rng.standard_normal((n, n))
    A = np.diag(d) + E
    b = rng.standard_normal(n)
    t = np.zeros(n)
    dd = DIIS(n + 2)
    for _ in range(n + 1):
        t = dd.step(t, A @ t - b, d)
    np.testing.assert_allclose(A @ t, b, rtol=0, atol=1e-9)


def test_divide_threshold():
    out = diavc(np.array([1.0, 1.0, 1.0, 2.0]), np.array([0.0, 5e-11, -5e-11, 4.0]))
    np.testing.assert_array_equal(out, [1e10, 1e10, 1e10, 0.5])


def test_subspace_drops_oldest():
    rng = np.random.default_rng(3)
    dd = DIIS(3)
    d = np.ones(10)
    rs = []
    for _ in range(5):
        g = rng.standard_normal(10)
        dd.step(rng.standard_normal(10), g, d)
        rs.append(g)
    assert len(dd.r) == 3
    np.testing.assert_array_equal(dd.r[0], rs[2])


def test_product_list_diagonal_matches_electron_sum():
    import ccsd
    import gecco_hip
    from synth import synthetic
    ham = synthetic(3, 7, 21, nsym=2, seed=5, sigma=0.05)
    o = gecco_hip.Orbs(ham.irrep_o, ham.irrep_v, 2)
    for name in ('T1', 'T2', 'Wmbej', 'Hovvv'):
        L = gecco_hip.MEList(o, *ccsd.LISTS[name])
        nc = sum(L.occ_c.values())
        ref = list_diagonal(L.spinorbs(), L.comp_spaces(), nc, ham.eps_o, ham.eps_v)
        np.testing.assert_allclose(L.diagonal(ham.eps_o, ham.eps_v), ref, rtol=0, atol=1e-13)
