"""Oracle: singlet spin projection of a particle-conserving ME-list, by brute
force in the Fock space.

TEST INFRASTRUCTURE ONLY (see oracle/strings.py header).

The reference projects a list onto its singlet (rank-0 spin tensor) part in
contract/spin_prj_blk.f (genealogical spin couplings, s2 = 0).  Here the
projector is found without any coupling scheme: every list element e is the
operator tau_e = a+_c1 ... a+_cn a_an ... a_a1 (oracle/cc_fock.py
conventions), a dense matrix on the Fock space of a few spin orbitals; the
singlet operators X = sum_e x_e tau_e are those with [S+, X] = 0 (total Ms is
zero); the tau_e are orthogonal with equal norms, so the projection is
x -> Q Q^T x with Q an orthonormal basis (SVD) of the null space of the map
x -> [S+, X].  Small spaces only (2^n-dimensional matrices).

Parity status: pinned -- tests/test_oracle_spin_project.py checks that the
projector is symmetric and idempotent, that the projected operator commutes
with S+ and S-, that spin-free (hence singlet) operators are left unchanged,
and the singlet counts of two-electron groups.
"""
import numpy as np

from .cc_fock import ladder_ops
from .ccsd_route3 import list_spinorbs

H, P = 1, 2


class SpinSpace:
    """Fock space of nocc hole and nvir particle spatial orbitals (modes:
    holes first, 2*orb + spin, spin 0 alpha)."""

    def __init__(self, nocc, nvir):
        self.nh = 2 * nocc
        self.n = 2 * (nocc + nvir)
        self.a = ladder_ops(self.n)
        self.ad = [x.T.copy() for x in self.a]
        self.splus = sum(self.ad[2 * p] @ self.a[2 * p + 1] for p in range(nocc + nvir))
        self.sminus = self.splus.T.copy()

    def mode(self, space, so):
        return (self.nh + so) if space == P else so

    def spaces(self, L):
        sp = []
        for occ in (L.occ_c, L.occ_a):
            for h in (4, 2, 3, 1):
                sp += [h] * occ.get(h, 0)
        return sp

    def tau(self, L, so_e, sp):
        nc = L.nc
        m = np.eye(2 ** self.n)
        for q in range(nc):
            m = m @ self.ad[self.mode(sp[q], so_e[q])]
        for q in reversed(range(nc, len(so_e))):
            m = m @ self.a[self.mode(sp[q], so_e[q])]
        return m

    def operators(self, L):
        so = list_spinorbs(L)
        sp = self.spaces(L)
        return [self.tau(L, so[e], sp) for e in range(L.length)]

    def projector(self, L):
        """The singlet projector on the list's coefficient vectors."""
        taus = self.operators(L)
        A = np.stack([(self.splus @ t - t @ self.splus).ravel() for t in taus], axis=1)
        # null space of A: right singular vectors with zero singular value
        _, s, vt = np.linalg.svd(A, full_matrices=False)
        tol = 1e-9 * max(1.0, s.max() if s.size else 1.0)
        rank = int((s > tol).sum())
        Q = vt[rank:].T
        return Q @ Q.T

    def operator_of(self, L, x):
        return sum(xi * t for xi, t in zip(x, self.operators(L)))

    def coefficients(self, L, X):
        """x_e = <tau_e, X> / <tau_e, tau_e> (Frobenius)."""
        taus = self.operators(L)
        return np.array([float(np.sum(t * X)) / float(np.sum(t * t)) for t in taus])
