"""Oracle: textbook spin-orbital CCSD on dense tensors (numpy).

TEST INFRASTRUCTURE ONLY (see oracle/strings.py header).

The reference's CCSD is generated symbolically (cc_spec/, formula/) and its
energy is pinned by test/CCPT-energy.e.ref:690-696 on a molecule whose
integral files this oracle does not read.  This module is an independent
statement of the same method -- the spin-orbital CCSD equations with the
intermediates of J.F. Stanton, J. Gauss, J.D. Watts, R.J. Bartlett,
J. Chem. Phys. 94, 4334 (1991), eqs. 1-13 -- used to pin energies and
amplitudes of the GPU engine on synthetic Hamiltonians.  Small systems
only (dense spin-orbital tensors).

Parity status: pinned by tests/test_oracle_ccsd.py against an independent
closed-shell (spin-adapted) CCSD energy and against second-order
perturbation theory (MP2) as the first iterate.

Spin-orbital index within a space: so = 2*orb + spin (spin 0 alpha,
1 beta), orbitals in the space's IRREP-sorted order -- GeCCo's string order.
"""
import numpy as np


def _einsum(*args):
    """np.einsum with its contraction order chosen for BLAS (results equal
    the left-to-right evaluation up to rounding)."""
    return np.einsum(*args, optimize=True)


def spin_integrals(ham):
    """Antisymmetrised <pq||rs> blocks and orbital energies, spin-orbital."""
    no, nv = ham.nocc, ham.nvir
    n = no + nv
    eri = ham.eri()                                 # (pq|rs) spatial
    ns = 2 * n
    sp = np.arange(ns) // 2                         # spatial of spin-orbital
    sg = np.arange(ns) % 2
    # order spin-orbitals: occ block (2*no) then vir block (2*nv)
    order = np.concatenate([np.arange(2 * no), 2 * no + np.arange(2 * nv)])
    sp = sp[order]
    sg = sg[order]
    phys = eri[sp[:, None, None, None], sp[None, None, :, None],
               sp[None, :, None, None], sp[None, None, None, :]]  # <pq|rs> = (pr|qs)
    phys = phys * ((sg[:, None, None, None] == sg[None, None, :, None]) &
                   (sg[None, :, None, None] == sg[None, None, None, :]))
    g = phys - phys.transpose(0, 1, 3, 2)
    eps = np.concatenate([np.repeat(ham.eps_o, 2), np.repeat(ham.eps_v, 2)])
    return g, eps, 2 * no


class CCSD:
    def __init__(self, ham):
        g, eps, nso = spin_integrals(ham)
        o = slice(0, nso)
        v = slice(nso, g.shape[0])
        self.g = g
        self.o, self.v = o, v
        self.eo, self.ev = eps[o], eps[v]
        self.oovv = g[o, o, v, v]
        self.Dia = self.eo[:, None] - self.ev[None, :]
        self.Dijab = (self.eo[:, None, None, None] + self.eo[None, :, None, None]
                      - self.ev[None, None, :, None] - self.ev[None, None, None, :])

    def energy(self, t1, t2):
        return 0.25 * _einsum('ijab,ijab', self.oovv, t2) + \
            0.5 * _einsum('ijab,ia,jb', self.oovv, t1, t1)

    def residual(self, t1, t2):
        """Stanton-Gauss-Watts-Bartlett 1991 with a diagonal Fock operator.
        Returns (R1, R2) such that the amplitude equations are R = 0, with
        R1 = D_ia t_ia - rhs1 sign-free form  (here: rhs - D*t)."""
        g, o, v = self.g, self.o, self.v
        ooov = g[o, o, o, v]; oovv = self.oovv; ovvv = g[o, v, v, v]
        oooo = g[o, o, o, o]; vvvv = g[v, v, v, v]; ovvo = g[o, v, v, o]
        ovoo = g[o, v, o, o]; vvvo = g[v, v, v, o]; oovo = g[o, o, v, o]
        tau = t2 + _einsum('ia,jb->ijab', t1, t1) - _einsum('ib,ja->ijab', t1, t1)
        taut = t2 + 0.5 * (_einsum('ia,jb->ijab', t1, t1) - _einsum('ib,ja->ijab', t1, t1))
        Fae = _einsum('mf,mafe->ae', t1, ovvv) - 0.5 * _einsum('mnaf,mnef->ae', taut, oovv)
        Fmi = _einsum('ne,mnie->mi', t1, ooov) + 0.5 * _einsum('inef,mnef->mi', taut, oovv)
        Fme = _einsum('nf,mnef->me', t1, oovv)
        Wmnij = oooo + _einsum('je,mnie->mnij', t1, ooov) - _einsum('ie,mnje->mnij', t1, ooov) \
            + 0.25 * _einsum('ijef,mnef->mnij', tau, oovv)
        Wabef = vvvv - _einsum('mb,amef->abef', t1, g[v, o, v, v]) \
            + _einsum('ma,bmef->abef', t1, g[v, o, v, v]) \
            + 0.25 * _einsum('mnab,mnef->abef', tau, oovv)
        Wmbej = ovvo + _einsum('jf,mbef->mbej', t1, ovvv) - _einsum('nb,mnej->mbej', t1, oovo) \
            - _einsum('jnfb,mnef->mbej', 0.5 * t2 + _einsum('jf,nb->jnfb', t1, t1), oovv)
        # T1 equation (eq. 1 with f_ia = 0, off-diagonal f = 0)
        r1 = _einsum('ie,ae->ia', t1, Fae) - _einsum('ma,mi->ia', t1, Fmi) \
            + _einsum('imae,me->ia', t2, Fme) - _einsum('nf,naif->ia', t1, g[o, v, o, v]) \
            - 0.5 * _einsum('imef,maef->ia', t2, ovvv) \
            - 0.5 * _einsum('mnae,nmei->ia', t2, oovo)
        # T2 equation (eq. 2)
        Fbe = Fae - 0.5 * _einsum('mb,me->be', t1, Fme)
        Fmj = Fmi + 0.5 * _einsum('je,me->mj', t1, Fme)
        x = _einsum('ijae,be->ijab', t2, Fbe)
        r2 = g[o, o, v, v].copy()
        r2 += x - x.transpose(0, 1, 3, 2)
        x = _einsum('imab,mj->ijab', t2, Fmj)
        r2 -= x - x.transpose(1, 0, 2, 3)
        r2 += 0.5 * _einsum('mnab,mnij->ijab', tau, Wmnij)
        r2 += 0.5 * _einsum('ijef,abef->ijab', tau, Wabef)
        x = _einsum('imae,mbej->ijab', t2, Wmbej) - _einsum('ie,ma,mbej->ijab', t1, t1, ovvo)
        r2 += x - x.transpose(0, 1, 3, 2) - x.transpose(1, 0, 2, 3) + x.transpose(1, 0, 3, 2)
        x = _einsum('ie,abej->ijab', t1, vvvo)
        r2 += x - x.transpose(1, 0, 2, 3)
        x = _einsum('ma,mbij->ijab', t1, ovoo)
        r2 -= x - x.transpose(0, 1, 3, 2)
        return r1, r2

    def solve(self, tol=1e-10, maxit=200, diis=8):
        """Jacobi update t = rhs/D with Pulay DIIS on the step; returns
        (E_corr, t1, t2, iterations)."""
        t1 = np.zeros_like(self.Dia)
        t2 = np.zeros_like(self.Dijab)
        hist_t, hist_r = [], []
        e_old = 0.0
        for it in range(1, maxit + 1):
            r1, r2 = self.residual(t1, t2)
            new1 = r1 / self.Dia
            new2 = r2 / self.Dijab
            step = np.concatenate([(new1 - t1).ravel(), (new2 - t2).ravel()])
            hist_t.append(np.concatenate([new1.ravel(), new2.ravel()]))
            hist_r.append(step)
            hist_t, hist_r = hist_t[-diis:], hist_r[-diis:]
            m = len(hist_t)
            Bm = -np.ones((m + 1, m + 1))
            Bm[m, m] = 0.0
            for a in range(m):
                for b in range(m):
                    Bm[a, b] = hist_r[a] @ hist_r[b]
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            c = np.linalg.solve(Bm, rhs)[:m]
            vec = sum(ci * ti for ci, ti in zip(c, hist_t))
            t1 = vec[:t1.size].reshape(t1.shape)
            t2 = vec[t1.size:].reshape(t2.shape)
            e = self.energy(t1, t2)
            if np.linalg.norm(step) < tol and abs(e - e_old) < tol:
                return e, t1, t2, it
            e_old = e
        raise RuntimeError('CCSD did not converge')


def mp2_energy(cc):
    return 0.25 * np.sum(cc.oovv ** 2 / cc.Dijab)


def fci_two_electron(ham):
    """Exact correlation energy of a two-electron system (nocc=1): lowest
    eigenvalue of H in the Ms=0 determinant basis minus the HF energy.
    h is recovered from the diagonal Fock operator, h_pq = f_pq - 2(pq|00) + (p0|0q)."""
    assert ham.nocc == 1
    eri = ham.eri()
    n = ham.nocc + ham.nvir
    eps = np.concatenate([ham.eps_o, ham.eps_v])
    h = np.diag(eps) - 2.0 * eri[:, :, 0, 0] + eri[:, 0, 0, :]
    e_hf = 2.0 * h[0, 0] + eri[0, 0, 0, 0]
    H = np.zeros((n * n, n * n))
    for p in range(n):
        for q in range(n):
            for r in range(n):
                for s in range(n):
                    H[p * n + q, r * n + s] = h[p, r] * (q == s) + (p == r) * h[q, s] + eri[p, r, q, s]
    return np.linalg.eigvalsh(H)[0] - e_hf
