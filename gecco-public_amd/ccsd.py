"""Closed-shell CCSD on the MI355X block-contraction engine.

The amplitude equations are the spin-orbital CCSD equations with the
intermediates of Stanton, Gauss, Watts and Bartlett (J. Chem. Phys. 94, 4334
(1991)), written as a list of GeCCo binary contractions over ME-lists in
GeCCo's Ms/IRREP-blocked layout -- the role the factorised formula file
plays for GeCCo's scheduler (contract/frm_sched2.f:117-365): every item is a
[CONTR] of two lists into a third, evaluated by the engine's plan
(contr_op1op2 semantics, spin-flip symmetry on).  Items of one level do not
depend on each other and run as one grouped launch.

Solver (solve/solve_nleq.f:377-787 with solve/optcont.f): diagonal
preconditioner D = sum(eps of A strings) - sum(eps of C strings)
(contract/dia4op_ev.f:300-365, step x = R/D as baserout/diavc.f), Pulay
DIIS on the preconditioned step (solve/optc_diis_extr.f:43-108), energy
and residual norm in FP64.

List conventions (dense reading of a list; C components then A components,
P before H inside each side, as the list layout orders them):
  T1[a;i] = t_i^a, T2[ab;ij] = t_ij^ab, H<C><A>[c..;a..] = <c..||a..>,
  Fae[a;e], Fmi[m;i], Fme[m;e], Wmnij[mn;ij], Wabef[ab;ef], Wmbej[b m;e j].
"""
import time

import numpy as np

H, P = 1, 2

# name: (C occupation, A occupation)
LISTS = {
    'T1': ({P: 1}, {H: 1}), 'T2': ({P: 2}, {H: 2}),
    'R1': ({P: 1}, {H: 1}), 'R2': ({P: 2}, {H: 2}),
    'tau': ({P: 2}, {H: 2}), 'taut': ({P: 2}, {H: 2}),
    'Fae': ({P: 1}, {P: 1}), 'F2ae': ({P: 1}, {P: 1}),
    'Fmi': ({H: 1}, {H: 1}), 'F2mi': ({H: 1}, {H: 1}),
    'Fme': ({H: 1}, {P: 1}),
    'Wmnij': ({H: 2}, {H: 2}), 'Wabef': ({P: 2}, {P: 2}),
    'Wmbej': ({P: 1, H: 1}, {P: 1, H: 1}),
    'Z': ({H: 2}, {P: 1, H: 1}), 'Y': ({P: 2}, {P: 1, H: 1}),
    'E': ({}, {}),
    'Hoovv': ({H: 2}, {P: 2}), 'Hvvoo': ({P: 2}, {H: 2}),
    'Hoooo': ({H: 2}, {H: 2}), 'Hvvvv': ({P: 2}, {P: 2}),
    'Hooov': ({H: 2}, {P: 1, H: 1}), 'Hovvv': ({P: 1, H: 1}, {P: 2}),
    'Hovvo': ({P: 1, H: 1}, {P: 1, H: 1}), 'Hvvvo': ({P: 2}, {P: 1, H: 1}),
    'Hovoo': ({P: 1, H: 1}, {H: 2}),
}
INTEGRALS = [n for n in LISTS if n.startswith('H')]
# lists initialised as factor * source at the start of each iteration
# (others: zero).  Hovvo reads <bm||ej> = -<mb||ej>, hence -1 for Wmbej.
INIT = {'tau': ('T2', 1.0), 'taut': ('T2', 1.0), 'Wmnij': ('Hoooo', 1.0),
        'Wabef': ('Hvvvv', 1.0), 'Wmbej': ('Hovvo', -1.0), 'R2': ('Hvvoo', 1.0)}

# (result, op1, op2, CNT on op1's C side, CNT on op1's A side, factor)
LEVELS = [
    [('tau', 'T1', 'T1', {}, {}, -0.5),
     ('taut', 'T1', 'T1', {}, {}, -0.25),
     ('Fme', 'Hoovv', 'T1', {H: 1}, {P: 1}, 1.0),
     ('Z', 'Hoovv', 'T1', {}, {P: 1}, -1.0),
     ('Fae', 'T1', 'Hovvv', {P: 1}, {H: 1}, -1.0),
     ('F2ae', 'T1', 'Hovvv', {P: 1}, {H: 1}, -1.0),
     ('Fmi', 'T1', 'Hooov', {P: 1}, {H: 1}, 1.0),
     ('F2mi', 'T1', 'Hooov', {P: 1}, {H: 1}, 1.0),
     ('Wmnij', 'Hooov', 'T1', {}, {P: 1}, 1.0),
     ('Wabef', 'Hovvv', 'T1', {H: 1}, {}, -1.0),
     ('Wmbej', 'Hovvv', 'T1', {}, {P: 1}, 1.0),
     ('Wmbej', 'T1', 'Hooov', {}, {H: 1}, 1.0),
     ('Wmbej', 'T2', 'Hoovv', {P: 1}, {H: 1}, 0.5),
     ('Y', 'T1', 'Hovvo', {}, {H: 1}, -1.0),
     ('R1', 'T1', 'Hovvo', {P: 1}, {H: 1}, -1.0),
     ('R1', 'Hovvv', 'T2', {H: 1}, {P: 2}, -1.0),
     ('R1', 'T2', 'Hooov', {P: 1}, {H: 2}, -1.0),
     ('R2', 'Hvvvo', 'T1', {}, {P: 1}, 1.0),
     ('R2', 'T1', 'Hovoo', {}, {H: 1}, 1.0)],
    [('Fae', 'taut', 'Hoovv', {P: 1}, {H: 2}, -1.0),
     ('F2ae', 'taut', 'Hoovv', {P: 1}, {H: 2}, -1.0),
     ('Fmi', 'Hoovv', 'taut', {H: 1}, {P: 2}, 1.0),
     ('F2mi', 'Hoovv', 'taut', {H: 1}, {P: 2}, 1.0),
     ('F2ae', 'T1', 'Fme', {}, {H: 1}, -0.5),
     ('F2mi', 'Fme', 'T1', {}, {P: 1}, 0.5),
     ('Wmnij', 'Hoovv', 'tau', {}, {P: 2}, 0.5),
     ('Wabef', 'tau', 'Hoovv', {}, {H: 2}, 0.5),
     ('Wmbej', 'T1', 'Z', {}, {H: 1}, 1.0),
     ('R1', 'T2', 'Fme', {P: 1}, {H: 1}, 1.0),
     ('R2', 'Y', 'T1', {}, {P: 1}, -1.0),
     ('E', 'Hoovv', 'tau', {H: 2}, {P: 2}, 1.0)],
    [('R1', 'Fae', 'T1', {}, {P: 1}, 1.0),
     ('R1', 'T1', 'Fmi', {}, {H: 1}, -1.0),
     ('R2', 'F2ae', 'T2', {}, {P: 1}, 1.0),
     ('R2', 'T2', 'F2mi', {}, {H: 1}, -1.0),
     ('R2', 'tau', 'Wmnij', {}, {H: 2}, 1.0),
     ('R2', 'Wabef', 'tau', {}, {P: 2}, 1.0),
     ('R2', 'T2', 'Wmbej', {P: 1}, {H: 1}, -1.0)],
]


class CCSD:
    """Device-resident closed-shell CCSD driver.

    ham: synthetic Hamiltonian (synth.Hamiltonian); device: torch device.
    """

    def __init__(self, ham, device='cuda:0', use_tr=True):
        import torch
        import gecco_hip
        self.torch = torch
        self.ham = ham
        self.dev = torch.device(device)
        self.use_tr = use_tr
        self.orbs = gecco_hip.Orbs(ham.irrep_o, ham.irrep_v, ham.nsym)
        self.L = {n: gecco_hip.MEList(self.orbs, c, a) for n, (c, a) in LISTS.items()}
        self.slot = {n: i for i, n in enumerate(LISTS)}
        self.buf = {n: torch.zeros(max(1, self.L[n].length), dtype=torch.float64, device=self.dev)
                    for n in LISTS}
        t0 = time.perf_counter()
        self._integrals()
        self._denominators()
        self.t_setup_integrals = time.perf_counter() - t0
        t0 = time.perf_counter()
        self.plans = []
        self.flops = 0.0
        self.nblocks = 0
        for level in LEVELS:
            p = gecco_hip.Plan()
            written = []
            for res, a, b, cc, ca, fact in level:
                nb, fl = p.add_contraction(self.L[a], self.L[b], self.L[res], cc, ca, fact, use_tr,
                                           buf=(self.slot[a], self.slot[b], self.slot[res]))
                self.flops += fl
                self.nblocks += nb
                if res not in written:
                    written.append(res)
            if use_tr:
                for res in written:
                    p.add_sym_ab(self.L[res], self.slot[res])
            p.finalize()
            self.plans.append(p)
        self.t_setup_plans = time.perf_counter() - t0
        self.bufs = [self.buf[n] for n in LISTS]

    # ------------------------------------------------------------ setup
    def _integrals(self):
        """<pq||rs> in every integral list from the DF factors, on the device."""
        torch = self.torch
        ham = self.ham
        n = ham.nocc + ham.nvir
        B = torch.from_numpy(ham.B.reshape(ham.B.shape[0], n * n)).to(self.dev)
        eri = (B.T @ B).reshape(n, n, n, n)      # (pq|rs)
        del B
        for name in INTEGRALS:
            L = self.L[name]
            so = torch.from_numpy(L.spinorbs().astype(np.int64)).to(self.dev)
            sp = L.comp_spaces()
            g, s = [], []
            for q in range(4):
                orb = so[:, q] >> 1
                g.append(orb if sp[q] == H else orb + ham.nocc)
                s.append(so[:, q] & 1)
            c1, c2, a1, a2 = g
            sc1, sc2, sa1, sa2 = s
            val = eri[c1, a1, c2, a2] * ((sc1 == sa1) & (sc2 == sa2)) - \
                eri[c1, a2, c2, a1] * ((sc1 == sa2) & (sc2 == sa1))
            self.buf[name].copy_(val)
            del so, val
        del eri
        torch.cuda.empty_cache()

    def _denominators(self):
        torch = self.torch
        ham = self.ham
        eps = torch.from_numpy(np.concatenate([ham.eps_o, ham.eps_v])).to(self.dev)
        self.D = {}
        for name in ('T1', 'T2'):
            L = self.L[name]
            so = torch.from_numpy(L.spinorbs().astype(np.int64)).to(self.dev)
            sp = L.comp_spaces()
            nc = sum(L.occ_c.values())
            d = torch.zeros(L.length, dtype=torch.float64, device=self.dev)
            for q in range(L.nel):
                orb = so[:, q] >> 1
                e = eps[orb if sp[q] == H else orb + ham.nocc]
                d += -e if q < nc else e
            self.D[name] = d

    # ---------------------------------------------------------- iteration
    def evaluate(self, stream=None):
        """Energy and residual lists for the amplitudes in T1/T2 (one pass of
        the formula)."""
        buf = self.buf
        torch = self.torch
        for n in LISTS:
            if n in INIT:
                src, f = INIT[n]
                if f == 1.0:
                    buf[n].copy_(buf[src])
                else:
                    torch.mul(buf[src], f, out=buf[n])
            elif n not in INTEGRALS and n not in ('T1', 'T2'):
                buf[n].zero_()
        sh = None if stream is None else stream
        for p in self.plans:
            p.execute(self.bufs, sh)

    def solve(self, tol=1e-8, maxit=100, ndiis=8, verbose=False, on_iter=None, restart=True):
        """Iterate to ||Omega|| < tol (Omega = rhs - D*t over T1 and T2), at
        most maxit iterations.  restart: start from T = 0 with an empty DIIS
        space; otherwise continue the previous solve.  Returns a dict with
        energy, iterations, residual norm and timings."""
        torch = self.torch
        buf, D = self.buf, self.D
        n1 = self.L['T1'].length
        Dcat = torch.cat([D['T1'], D['T2']])
        if restart or not hasattr(self, '_diis'):
            buf['T1'].zero_()
            buf['T2'].zero_()
            self._diis = ([], [], np.zeros((ndiis, ndiis)))
        hist_x, hist_e, bmat = self._diis   # DIIS overlaps <e_i|e_j>, updated by rows
        energy = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it = 0
        rnorm = float('inf')
        for it in range(1, maxit + 1):
            self.evaluate(torch.cuda.current_stream().cuda_stream)
            t = torch.cat([buf['T1'], buf['T2']])
            rhs = torch.cat([buf['R1'], buf['R2']])
            omega = rhs - Dcat * t
            rnorm = float(torch.linalg.vector_norm(omega))
            energy = float(buf['E'][0])
            if on_iter:
                on_iter(it, energy, rnorm)
            if verbose:
                print(f'  it {it:3d}  E = {energy:.12f}  |R| = {rnorm:.3e}', flush=True)
            if rnorm < tol:
                break
            step = omega / Dcat
            if len(hist_x) == ndiis:
                hist_x.pop(0)
                hist_e.pop(0)
                bmat[:-1, :-1] = bmat[1:, 1:].copy()
            hist_x.append(t + step)
            hist_e.append(step)
            m = len(hist_x)
            # new row of the overlap matrix: elementwise products and
            # reductions (a library GEMM with K = vector length and M,N <= 8
            # runs without K-splitting and takes hundreds of ms)
            row = torch.stack([(e * step).sum() for e in hist_e]).cpu().numpy()
            bmat[m - 1, :m] = row
            bmat[:m, m - 1] = row
            Bm = np.zeros((m + 1, m + 1))
            Bm[:m, :m] = bmat[:m, :m]
            Bm[m, :m] = Bm[:m, m] = -1.0
            rhs_d = np.zeros(m + 1)
            rhs_d[m] = -1.0
            c = np.linalg.solve(Bm, rhs_d)[:m]
            new = hist_x[-1] * float(c[-1])
            for ci, xi in zip(c[:-1], hist_x[:-1]):
                new.add_(xi, alpha=float(ci))
            buf['T1'].copy_(new[:n1])
            buf['T2'].copy_(new[n1:])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return dict(energy=energy, iterations=it, residual=rnorm, seconds=wall,
                    iters_per_s=it / wall, flops_per_iter=self.flops, kernel_calls=self.nblocks)
