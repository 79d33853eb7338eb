"""Seeded synthetic closed-shell Hamiltonians (inputs for tests and bench).

Recipe: orbital energies eps_occ ~ U[occ_lo, occ_hi], eps_vir ~ U[vir_lo,
vir_hi]; two-electron integrals from a density-fitted factorisation
(pq|rs) = sum_P B(P,pq) B(P,rs) with B ~ N(0, sigma^2) symmetric in pq;
orbitals and auxiliary functions assigned round-robin to nsym IRREPs of a
D2h subgroup (XOR products), B(P,pq) = 0 unless IRREP(p)*IRREP(q) =
IRREP(P), so (pq|rs) is totally symmetric.  The Fock operator is diagonal
(canonical orbitals): f_pq = eps_p delta_pq.

Within each space the orbitals are reordered by IRREP (stable), which is
the orbital order GeCCo's strings assume; irrep_o/irrep_v are returned in
that order.
"""
import numpy as np


def multd2h(a, b):
    return ((a - 1) ^ (b - 1)) + 1


class Hamiltonian:
    def __init__(self, eps_o, eps_v, irrep_o, irrep_v, B, nsym):
        self.eps_o, self.eps_v = eps_o, eps_v
        self.irrep_o, self.irrep_v = irrep_o, irrep_v
        self.B = B            # (naux, n, n), orbitals occ then vir
        self.nsym = nsym
        self.nocc, self.nvir = len(eps_o), len(eps_v)

    def eri(self):
        """(pq|rs) chemists' notation, dense (n^4)."""
        n = self.nocc + self.nvir
        Bm = self.B.reshape(self.B.shape[0], n * n)
        return (Bm.T @ Bm).reshape(n, n, n, n)


class DenseHamiltonian:
    """Canonical closed-shell Hamiltonian given by orbital energies and the
    dense (pq|rs) over the correlated orbitals (occupied then virtual, each
    ordered by IRREP), e.g. imported from a quantum-chemistry program."""

    def __init__(self, eps_o, eps_v, irrep_o, irrep_v, eri, nsym, e_ref=0.0):
        self.eps_o, self.eps_v = np.asarray(eps_o, float), np.asarray(eps_v, float)
        self.irrep_o = [int(x) for x in irrep_o]
        self.irrep_v = [int(x) for x in irrep_v]
        self.eri_full = np.ascontiguousarray(eri, dtype=np.float64)
        self.nsym = int(nsym)
        self.nocc, self.nvir = len(self.eps_o), len(self.eps_v)
        self.e_ref = float(e_ref)       # reference (SCF) energy

    def eri(self):
        return self.eri_full


def synthetic(nocc, nvir, naux, nsym=1, seed=0, sigma=0.03, occ_range=(-2.0, -0.3),
              vir_range=(0.2, 3.0)):
    rng = np.random.default_rng(seed)
    eps_o = rng.uniform(*occ_range, nocc)
    eps_v = rng.uniform(*vir_range, nvir)
    n = nocc + nvir
    irr = np.array([(p % nsym) + 1 for p in range(n)])
    # order by IRREP inside each space (stable)
    po = np.argsort(irr[:nocc], kind='stable')
    pv = nocc + np.argsort(irr[nocc:], kind='stable')
    perm = np.concatenate([po, pv])
    irr = irr[perm]
    eps = np.concatenate([eps_o, eps_v])[perm]
    G = rng.standard_normal((naux, n, n)) * sigma
    G = (G + G.transpose(0, 2, 1)) / np.sqrt(2.0)
    G = G[:, perm][:, :, perm]
    irr_aux = np.array([(P % nsym) + 1 for P in range(naux)])
    gpq = ((irr[:, None] - 1) ^ (irr[None, :] - 1)) + 1
    mask = gpq[None, :, :] == irr_aux[:, None, None]
    B = np.where(mask, G, 0.0)
    return Hamiltonian(eps[:nocc], eps[nocc:], irr[:nocc].tolist(), irr[nocc:].tolist(), B, nsym)


def batch_recipe(count=64, nocc_range=(8, 24), nvir_range=(40, 160), aux_per_vir=3, nsym=4,
                 sigma=0.03):
    """Arguments of synthetic() for a batch of independent instances of
    uneven size: instance k has seed k, nocc and nvir uniform integers in the
    (inclusive) ranges drawn from default_rng([k, 1]), naux = aux_per_vir *
    nvir."""
    out = []
    for k in range(count):
        r = np.random.default_rng([k, 1])
        o = int(r.integers(nocc_range[0], nocc_range[1] + 1))
        v = int(r.integers(nvir_range[0], nvir_range[1] + 1))
        out.append(dict(nocc=o, nvir=v, naux=aux_per_vir * v, nsym=nsym, seed=k, sigma=sigma))
    return out
