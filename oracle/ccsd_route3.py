"""Oracle: a CCSD formula evaluated with the oracle's restatement of GeCCo's
route 3 (collect, BLAS dgemm, scatter per kernel call; oracle/contract.py),
and the Jacobi + Pulay-DIIS solver of the reference (solve/optc_diis_*.f).

TEST INFRASTRUCTURE ONLY (see oracle/strings.py header): used by
tests/golden/make_ccsd_reference.py to write the float64 energy fixture and by
bench.py's cpu_baseline leg, which times one whole formula iteration of it on
one host core -- the reference's own arithmetic path, contract/
contr_blk1blk2_blocked_mm.f with a sequential BLAS dgemm (:465,523).

Restated: the formula schedule and initialisations of the CCSD formula
table (gecco-public_amd/ccsd.py, the role GeCCo's formula file plays for
contract/frm_sched2.f:117-365); each contraction's kernel calls from
oracle.contract.Contraction (contract/contr_op1op2_wmaps_c.f:964-2070);
the spin-flip projection of each written list after its level
(contract/sym_ab_blk.f); lists of several vertices and REORDER items
(oracle.contract.Contraction.joined, oracle/reorder.py); the diagonal
preconditioner (contract/dia4op_ev.f:300-365) and DIIS with 8 vectors
(solve/optc_diis_extr.f:43-108).

Parity status: pinned -- its energy equals the dense textbook CCSD
(oracle/ccsd_dense.py) on small systems (tests/test_oracle_ccsd_route3.py)."""
import time

import numpy as np

from .contract import Contraction, sym_ab_pairs, apply_sym_pairs
from .reorder import Reorder
from .melist import Orbitals, MEList, Adjoint
from .strings import strings_of, to_spinorbs

H, P = 1, 2


def list_spinorbs(L):
    """(length, nel) spin-orbital indices 2*(orb-1)+spin of every element of
    list L, electrons in component order (C components, then A)."""
    nel = L.nc + L.na
    out = np.zeros((L.length, nel), dtype=np.int64)
    for dl in L.blocks.values():
        for d in dl:
            comps = []
            for k in range(len(L.oc)):
                comps.append(strings_of(L.orbs.spaces[L.hc[k]], L.oc[k], d['msc'][k], d['gamc'][k]))
            for k in range(len(L.oa)):
                comps.append(strings_of(L.orbs.spaces[L.ha[k]], L.oa[k], d['msa'][k], d['gama'][k]))
            arrs = [np.array([[2 * (o - 1) + (0 if s == 1 else 1) for o, s in to_spinorbs(x)]
                              for x in c], dtype=np.int64).reshape(len(c), -1) for c in comps]
            ld = d['ldc'] + d['lda']
            grids = np.meshgrid(*[np.arange(len(c)) for c in comps], indexing='ij')
            offs = d['off'] + sum(g * l for g, l in zip(grids, ld))
            cols = [a[g] for a, g in zip(arrs, grids)]
            block = np.concatenate([c[..., None] if c.ndim == len(comps) else c for c in cols], axis=-1)
            out[offs.ravel()] = block.reshape(-1, nel)
    return out


def _electron_spaces(L):
    sp = []
    for side in (L.occ_c, L.occ_a):
        for h in (4, 2, 3, 1):
            sp += [h] * side.get(h, 0)
    return sp


def integral_lists(lists, ham, names):
    """<c1 c2||a1 a2> in every named two-body list, from the Hamiltonian's
    (pq|rs) (density fitted or dense)."""
    n = ham.nocc + ham.nvir
    if getattr(ham, 'eri_full', None) is not None:
        eri = ham.eri_full
    else:
        Bm = ham.B.reshape(ham.B.shape[0], n * n)
        eri = (Bm.T @ Bm).reshape(n, n, n, n)
    out = {}
    for name in names:
        L = lists[name]
        so = list_spinorbs(L)
        sp = _electron_spaces(L)
        g = [(so[:, q] >> 1) + (ham.nocc if sp[q] == P else 0) for q in range(4)]
        s = [so[:, q] & 1 for q in range(4)]
        out[name] = eri[g[0], g[2], g[1], g[3]] * ((s[0] == s[2]) & (s[1] == s[3])) - \
            eri[g[0], g[3], g[1], g[2]] * ((s[0] == s[3]) & (s[1] == s[2]))
    return out


def diagonal(L, ham):
    """dia4op_ev: eps of the C electrons minus eps of the A electrons."""
    eps = np.concatenate([ham.eps_o, ham.eps_v])
    so = list_spinorbs(L)
    sp = _electron_spaces(L)
    d = np.zeros(L.length)
    for q in range(so.shape[1]):
        e = eps[(so[:, q] >> 1) + (ham.nocc if sp[q] == P else 0)]
        d += -e if q < L.nc else e
    return d


class Route3CCSD:
    """formula: {'lists', 'constant', 'items'} with its schedule (levels of
    {'init', 'contr', 'sym'}), as gecco-public_amd/formula.py writes them."""

    def __init__(self, ham, formula, schedule):
        self.ham = ham
        orbs = Orbitals(ham.irrep_o, ham.irrep_v, ham.nsym)
        self.L = {k: MEList(orbs, c, a) for k, (c, a) in formula['lists'].items()}
        self.constant = list(formula['constant'])
        self.X = integral_lists(self.L, ham, self.constant)
        # amplitude lists T1, T2, ... (as many as the formula has)
        self.tnames = [f'T{k}' for k in range(1, 10) if f'T{k}' in self.L]
        self.rnames = [f'R{k}' for k in range(1, len(self.tnames) + 1)]
        self.D = np.concatenate([diagonal(self.L[n], ham) for n in self.tnames])
        self.bounds = np.cumsum([0] + [self.L[n].length for n in self.tnames])
        self.levels = []
        self.flops = 0.0
        self.calls = 0
        for lv in schedule:
            cs = []
            for it in lv['contr']:
                v = it.get('vertices')
                # dagger flags: adjoint views of the flagged one-vertex lists
                L1, L2, Lt = [Adjoint(self.L[n]) if f else self.L[n] for n, f in
                              zip((it['op1'], it['op2'], it['target']), it.get('tra') or (0, 0, 0))]
                if v is None:
                    c = Contraction(L1, L2, Lt, it['cnt_c'], it['cnt_a'], it['factor'])
                else:       # lists of several vertices
                    c = Contraction.joined(L1, L2, Lt, v['ex1'], v['ex2'], v['cnt'], *v['merge'],
                                           fact=it['factor'])
                blocks = c.blocks(True)
                self.flops += sum(2.0 * np.prod(b['lstr_ex1']) * np.prod(b['lstr_ex2']) *
                                  np.prod(b['lstr_cnt']) for b in blocks)
                self.calls += len(blocks)
                cs.append((it['target'], it['op1'], it['op2'], c, blocks))
            reos = [(it['target'], it['source'],
                     Reorder(self.L[it['source']], self.L[it['target']], it['from'], it['to'],
                             it['shift_c'], it['shift_a'], it['factor'], it['sign']))
                    for it in lv.get('reo', [])]
            self.levels.append((lv['init'], lv.get('add', []), cs, reos, lv['sym']))
        self.pairs = {n: sym_ab_pairs(self.L[n], 1.0) for *_, sym in self.levels for n in sym}
        for n in self.tnames:
            self.X[n] = np.zeros(self.L[n].length)

    def evaluate(self):
        """One pass of the formula for the amplitudes in X['T1'], X['T2'],
        ...: returns the residual Omega = R - D t (the function g of the
        solver) over all ranks."""
        X, L = self.X, self.L
        for inits, adds, cs, reos, sym in self.levels:
            for it in inits:
                n = it['target']
                X[n] = it['factor'] * X[it['source']] if 'source' in it else np.zeros(L[n].length)
            for it in adds:                 # ADD items (contract/fs_add_drv.f)
                X[it['target']] = X[it['target']] + it['factor'] * X[it['source']]
            for res, a, b, c, blocks in cs:
                c.run(X[a], X[b], X[res], route=3, use_tr=False, screen=False, blocks=blocks)
            for res, src, ro in reos:       # REORDER items (contract/fs_reo_drv.f)
                X[res] = ro.run(X[src], X[res])
            for n in sym:
                apply_sym_pairs(X[n], self.pairs[n], 0.5)
        t = np.concatenate([X[n] for n in self.tnames])
        return np.concatenate([X[n] for n in self.rnames]) - self.D * t

    def solve(self, tol=1e-10, maxit=200, ndiis=8, log=None):
        X = self.X
        hx, he = [], []
        t0 = time.time()
        for it in range(1, maxit + 1):
            omega = self.evaluate()
            rn = float(np.linalg.norm(omega))
            energy = float(X['E'][0])
            if log:
                log(f'it {it:3d} E {energy:.14f} |R| {rn:.3e} ({time.time() - t0:.0f} s)')
            if rn < tol:
                return energy, it, rn
            t = np.concatenate([X[n] for n in self.tnames])
            step = omega / self.D
            hx.append(t + step)
            he.append(step)
            hx, he = hx[-ndiis:], he[-ndiis:]
            m = len(hx)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for i in range(m):
                for j in range(m):
                    B[i, j] = he[i] @ he[j]
            r = np.zeros(m + 1)
            r[m] = -1.0
            coef = np.linalg.solve(B, r)[:m]
            new = sum(ci * xi for ci, xi in zip(coef, hx))
            for k, n in enumerate(self.tnames):
                X[n] = new[self.bounds[k]:self.bounds[k + 1]].copy()
        return energy, maxit, rn
