"""Oracle (TEST INFRASTRUCTURE ONLY): CPU restatement of GeCCo's diagonal
preconditioner and Pulay DIIS for one amplitude vector.

Only tests, the smoke check and the CPU baseline may import this module; the
product path is gecco_hip.DIIS (HIP kernels in gecco-public_amd/csrc/solver.hip).

Follows
  baserout/diavc.f:1-38            vecout = fac*vecin/(diag+shift), |.| <= 1e-10 -> 1e-10
  solve/optc_diis_sbsp_add.f:150-205   r = D^-1 g to the residual subspace,
                                       v = t - r to the vector subspace
  solve/optc_update_redsp1.f       B(i,j) = <r_i|r_j>, oldest dropped when full
  solve/optc_diis_extr.f:66-105    [B -1; -1 0] (w, l) = (0, -1)
  solve/optc_expand_vec.f          t = sum_i w_i v_i
  contract/dia4op_ev.f:298-365     diagonal = sum eps(C strings) - sum eps(A strings)

The reference solves the augmented system with LINPACK dspco/dspsl
(math/linpack.f, symmetric indefinite packed); this restatement uses
numpy.linalg.solve: both return the unique solution.

Parity status: pinned -- tests/test_oracle_diis.py checks a one-vector step
against the Jacobi step, the weights against an independent constrained
least-squares solution, finite termination on linear problems (DIIS with the
full history spans the same Krylov space as GMRES) and the divide threshold.
"""
import numpy as np

THRESH = 1e-10


def diavc(vec, diag, fac=1.0, shift=0.0):
    """baserout/diavc.f:31-35."""
    d = np.asarray(diag, dtype=np.float64) + shift
    d = np.where(np.abs(d) <= THRESH, THRESH, d)
    return fac * np.asarray(vec) / d


def list_diagonal(spinorbs, spaces, nc, eps_h, eps_p):
    """contract/dia4op_ev.f:298-365 (one-electron part) from the per-element
    spin-orbitals (gecco_hip.MEList.spinorbs layout): +eps for the nc C
    electrons, -eps for the A electrons."""
    out = np.zeros(spinorbs.shape[0])
    for q in range(spinorbs.shape[1]):
        eps = np.asarray(eps_h if spaces[q] == 1 else eps_p)
        out += (1.0 if q < nc else -1.0) * eps[spinorbs[:, q] >> 1]
    return out


class DIIS:
    """One amplitude vector, subspace of at most maxsub vectors."""

    def __init__(self, maxsub=8):
        self.maxsub = maxsub
        self.r = []
        self.v = []
        self.w = None

    def step(self, t, g, diag):
        """New amplitudes from t and the gradient g = g(t)."""
        r = diavc(g, diag)
        v = t - r
        if len(self.r) == self.maxsub:
            self.r.pop(0)
            self.v.pop(0)
        self.r.append(r)
        self.v.append(v)
        m = len(self.r)
        a = np.zeros((m + 1, m + 1))
        for i in range(m):
            for j in range(m):
                a[i, j] = self.r[i] @ self.r[j]
        a[:m, m] = -1.0
        a[m, :m] = -1.0
        b = np.zeros(m + 1)
        b[m] = -1.0
        self.w = np.linalg.solve(a, b)[:m]
        return sum(wi * vi for wi, vi in zip(self.w, self.v))
