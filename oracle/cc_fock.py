"""Oracle: coupled-cluster residuals by brute force in the Fock space.

TEST INFRASTRUCTURE ONLY (see oracle/strings.py header).

For a few spin orbitals the creation and annihilation operators are dense
matrices on the 2^n-dimensional Fock space (Jordan-Wigner), the Hamiltonian
is H = sum_pq h_pq p+ q + 1/4 sum_pqrs <pq||rs> p+ q+ s r with h = f - sum_i
<p i||q i> for the reference's Fock operator f = diag(eps), the cluster
operator T is built from the amplitude lists, and the residual of every
excitation mu is <mu| e^-T H e^T |0> -- no diagrams, no factorisation, no
sign rules: the definition of the CC equations (the method the reference's
cc_spec/ targets generate).  Used to pin generated formulas of any
excitation level (gecco-public_amd/ccgen.py) on tiny systems.

Conventions (those of the ME-lists, oracle/ccsd_route3.py): spin orbital
2*(orb-1)+spin within a space, holes then particles; a list element with C
strings c_1 < ... (components in list order, each string ascending) and A
strings a_1, ... is the coefficient of a+_c1 ... a+_cn a_an ... a_a1 (the
annihilators in reverse order); <pq||rs> is the coefficient of p+ q+ s r.

Parity status: pinned -- the Fock-space CCSD residual vanishes at the dense
textbook CCSD solution (oracle/ccsd_dense.py) and its energy there is the
textbook energy (tests/test_oracle_cc_fock.py).
"""
import numpy as np

from .ccsd_route3 import list_spinorbs


def ladder_ops(n):
    """Annihilators a_p (dense 2^n x 2^n) of n fermionic modes, mode p the
    p-th Jordan-Wigner factor."""
    I = np.eye(2)
    Z = np.diag([1.0, -1.0])
    low = np.array([[0.0, 1.0], [0.0, 0.0]])      # |0><1| on one mode: annihilates
    ops = []
    for p in range(n):
        m = np.array([[1.0]])
        for q in range(n):
            m = np.kron(m, Z if q < p else (low if q == p else I))
        ops.append(m)
    return ops


class FockSpace:
    """Spin orbitals: 2*nocc holes (modes 0..2nocc-1), then 2*nvir particles."""

    def __init__(self, ham):
        self.ham = ham
        no, nv = ham.nocc, ham.nvir
        self.nh, self.n = 2 * no, 2 * (no + nv)
        self.a = ladder_ops(self.n)
        self.ad = [x.T.copy() for x in self.a]
        vac = np.zeros(2 ** self.n)
        vac[0] = 1.0
        ref = vac
        for p in range(self.nh - 1, -1, -1):
            ref = self.ad[p] @ ref
        self.ref = ref
        self.H = self._hamiltonian()

    def mode(self, space_is_p, so):
        return (self.nh + so) if space_is_p else so

    def _hamiltonian(self):
        ham = self.ham
        no = ham.nocc
        n = self.n
        norb = no + ham.nvir
        eps = np.concatenate([ham.eps_o, ham.eps_v])
        eri = ham.eri()                 # (pq|rs) spatial, occupied then virtual
        # spatial orbital and spin of each mode
        sp = np.zeros(n, dtype=int)
        sg = np.zeros(n, dtype=int)
        for m in range(n):
            if m < self.nh:
                sp[m], sg[m] = m // 2, m % 2
            else:
                sp[m], sg[m] = no + (m - self.nh) // 2, (m - self.nh) % 2
        H = np.zeros((2 ** n, 2 ** n))
        # one-electron part h = f - sum_i <p i||q i>: the Fock operator of the
        # reference is the diagonal eps (canonical orbitals)
        for p in range(n):
            for q in range(n):
                if sg[p] != sg[q]:
                    continue
                h = eps[sp[p]] if p == q else 0.0
                for i in range(self.nh):
                    h -= eri[sp[p], sp[q], sp[i], sp[i]]
                    if sg[i] == sg[p]:
                        h += eri[sp[p], sp[i], sp[i], sp[q]]
                if h != 0.0:
                    H += h * (self.ad[p] @ self.a[q])
        # 1/4 sum <pq||rs> p+ q+ s r  =  1/2 sum <pq|rs> p+ q+ s r
        for p in range(n):
            for q in range(n):
                if p == q:
                    continue
                pq = self.ad[p] @ self.ad[q]
                for r in range(n):
                    for s in range(n):
                        if r == s or sg[p] != sg[r] or sg[q] != sg[s]:
                            continue
                        v = eri[sp[p], sp[r], sp[q], sp[s]]      # <pq|rs> = (pr|qs)
                        if v != 0.0:
                            H += 0.5 * v * (pq @ self.a[s] @ self.a[r])
        assert norb == sp.max() + 1
        return H

    def tau(self, L, elem_so):
        """Excitation operator of one list element (its spin orbitals from
        list_spinorbs: C electrons then A electrons)."""
        nc = L.nc
        cs = [(self.mode(True, so)) for so in elem_so[:nc]]
        as_ = [(self.mode(False, so)) for so in elem_so[nc:]]
        m = np.eye(2 ** self.n)
        for c in cs:
            m = m @ self.ad[c]
        for a in reversed(as_):
            m = m @ self.a[a]
        return m

    def cluster(self, lists, amps):
        """T = sum over the amplitude lists (excitation operators: C
        particles, A holes) of element * tau(element)."""
        T = np.zeros((2 ** self.n, 2 ** self.n))
        for L, x in zip(lists, amps):
            so = list_spinorbs(L)
            for e in range(L.length):
                if x[e] != 0.0:
                    T += x[e] * self.tau(L, so[e])
        return T

    def residuals(self, lists, amps, res_lists):
        """(E_corr, [R list per residual list]) with R_mu = <mu|e^-T H e^T|0>
        for the elements of res_lists and E = <0|e^-T H e^T|0> - <0|H|0>."""
        T = self.cluster(lists, amps)
        eT = np.eye(len(T))
        term = np.eye(len(T))
        for k in range(1, 2 * 4 + 2):       # T is nilpotent
            term = term @ T / k
            eT = eT + term
        emT = np.eye(len(T))
        term = np.eye(len(T))
        for k in range(1, 2 * 4 + 2):
            term = -term @ T / k
            emT = emT + term
        hbar0 = emT @ (self.H @ (eT @ self.ref))
        e = float(self.ref @ hbar0 - self.ref @ (self.H @ self.ref))
        out = []
        for L in res_lists:
            so = list_spinorbs(L)
            r = np.zeros(L.length)
            for k in range(L.length):
                mu = self.tau(L, so[k]) @ self.ref
                r[k] = float(mu @ hbar0)
            out.append(r)
        return e, out
