"""Closed-shell CCSD on the MI355X block-contraction engine.

The amplitude equations are the spin-orbital CCSD equations with the
intermediates of Stanton, Gauss, Watts and Bartlett (J. Chem. Phys. 94, 4334
(1991)), refactored so that no v^4 intermediate is built:
  1/2 tau_ij^ef W_abef = 1/2 tau_ij^ef <ab||ef> - P(ab) t_m^b Zamij
                         + 1/8 tau_mn^ab tau_ij^ef <mn||ef>,
  Zamij = 1/2 tau_ij^ef <am||ef>,
and the last term is folded into W_mnij, whose tau.<mn||ef> term then
carries 1/2 instead of 1/4.  The doubles term -P(ij)P(ab) t_i^e t_m^a
<mb||ej> contracts the virtual index first, t_i^e <mb||ej> (an o^3 v
intermediate instead of the o v^3 one of contracting m first); it has
Zamij's index structure and reaches R2 through the same t_m Zamij
contraction, so it accumulates into Zamij.  The equations are
written as a list of GeCCo binary contractions over ME-lists in
GeCCo's Ms/IRREP-blocked layout -- the role the factorised formula file
plays for GeCCo's scheduler (contract/frm_sched2.f:117-365): every item is a
[CONTR] of two lists into a third, evaluated by the engine's plan
(contr_op1op2 semantics, spin-flip symmetry on).  Items of one level do not
depend on each other and run as one grouped launch.

Solver (solve/solve_nleq.f:377-787 with solve/optcont.f): diagonal
preconditioner D = sum(eps of C strings) - sum(eps of A strings)
(contract/dia4op_ev.f:300-365, step r = g/D as baserout/diavc.f), Pulay
DIIS on the preconditioned step (solve/optc_diis_extr.f:43-108), energy
and residual norm in FP64 -- the library's native DIIS step
(gecco_hip.DIIS).

List conventions (dense reading of a list; C components then A components,
P before H inside each side, as the list layout orders them):
  T1[a;i] = t_i^a, T2[ab;ij] = t_ij^ab, H<C><A>[c..;a..] = <c..||a..>,
  Fae[a;e], Fmi[m;i], Fme[m;e], Wmnij[mn;ij], Zamij[a m;ij], Wmbej[b m;e j].
"""
import math
import time

import numpy as np

import formula as _formula

H, P = 1, 2

# name: (C occupation, A occupation)
LISTS = {
    'T1': ({P: 1}, {H: 1}), 'T2': ({P: 2}, {H: 2}),
    'R1': ({P: 1}, {H: 1}), 'R2': ({P: 2}, {H: 2}),
    'tau': ({P: 2}, {H: 2}), 'taut': ({P: 2}, {H: 2}),
    'Fae': ({P: 1}, {P: 1}), 'F2ae': ({P: 1}, {P: 1}),
    'Fmi': ({H: 1}, {H: 1}), 'F2mi': ({H: 1}, {H: 1}),
    'Fme': ({H: 1}, {P: 1}),
    'Wmnij': ({H: 2}, {H: 2}), 'Zamij': ({P: 1, H: 1}, {H: 2}),
    'Wmbej': ({P: 1, H: 1}, {P: 1, H: 1}),
    'Z': ({H: 2}, {P: 1, H: 1}),
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
        'Wmbej': ('Hovvo', -1.0), 'R2': ('Hvvoo', 1.0)}

# The formula as a flat list in GeCCo's order (formula.py): list
# initialisations, then the binary contractions (result, op1, op2, CNT on
# op1's C side, CNT on op1's A side, factor).  The levels below are derived
# from it by formula.schedule; the nesting here shows them.
_TERM_LEVELS = [
    [('tau', 'T1', 'T1', {}, {}, -0.5),
     ('taut', 'T1', 'T1', {}, {}, -0.25),
     ('Fme', 'Hoovv', 'T1', {H: 1}, {P: 1}, 1.0),
     ('Z', 'Hoovv', 'T1', {}, {P: 1}, -1.0),
     ('Fae', 'T1', 'Hovvv', {P: 1}, {H: 1}, -1.0),
     ('Fmi', 'T1', 'Hooov', {P: 1}, {H: 1}, 1.0),
     ('Wmnij', 'Hooov', 'T1', {}, {P: 1}, 1.0),
     ('Wmbej', 'Hovvv', 'T1', {}, {P: 1}, 1.0),
     ('Wmbej', 'T1', 'Hooov', {}, {H: 1}, 1.0),
     ('Wmbej', 'T2', 'Hoovv', {P: 1}, {H: 1}, 0.5),
     ('Zamij', 'Hovvo', 'T1', {}, {P: 1}, 1.0),
     ('R1', 'T1', 'Hovvo', {P: 1}, {H: 1}, -1.0),
     ('R1', 'Hovvv', 'T2', {H: 1}, {P: 2}, -1.0),
     ('R1', 'T2', 'Hooov', {P: 1}, {H: 2}, -1.0),
     ('R2', 'Hvvvo', 'T1', {}, {P: 1}, 1.0),
     ('R2', 'T1', 'Hovoo', {}, {H: 1}, 1.0)],
    [('Fae', 'taut', 'Hoovv', {P: 1}, {H: 2}, -1.0),
     ('Fmi', 'Hoovv', 'taut', {H: 1}, {P: 2}, 1.0),
     ('F2ae', 'T1', 'Fme', {}, {H: 1}, -0.5),
     ('F2mi', 'Fme', 'T1', {}, {P: 1}, 0.5),
     ('Wmnij', 'Hoovv', 'tau', {}, {P: 2}, 1.0),
     ('R2', 'Hvvvv', 'tau', {}, {P: 2}, 1.0),
     ('Zamij', 'Hovvv', 'tau', {}, {P: 2}, 1.0),
     ('Wmbej', 'T1', 'Z', {}, {H: 1}, 1.0),
     ('R1', 'T2', 'Fme', {P: 1}, {H: 1}, 1.0),
     ('E', 'Hoovv', 'tau', {H: 2}, {P: 2}, 1.0)],
    [('ADD', 'F2ae', 'Fae', 1.0),
     ('ADD', 'F2mi', 'Fmi', 1.0),
     ('R1', 'Fae', 'T1', {}, {P: 1}, 1.0),
     ('R1', 'T1', 'Fmi', {}, {H: 1}, -1.0),
     ('R2', 'F2ae', 'T2', {}, {P: 1}, 1.0),
     ('R2', 'T2', 'F2mi', {}, {H: 1}, -1.0),
     ('R2', 'tau', 'Wmnij', {}, {H: 2}, 1.0),
     ('R2', 'Zamij', 'T1', {H: 1}, {}, -1.0),
     ('R2', 'T2', 'Wmbej', {P: 1}, {H: 1}, -1.0)],
]



TERMS = [t for level in _TERM_LEVELS for t in level if t[0] != 'ADD']
# ADD items (command_add_intm): F2ae = Fae + (-1/2 t_m^a F_me) and F2mi =
# Fmi + 1/2 t_i^e F_me share Fae's and Fmi's T1 and tau~ terms, so those
# are contracted once (into Fae, Fmi) and added to F2ae, F2mi in the
# initialisation launch of the level that reads them.


def record_contraction(p, c, lists, slot, use_tr):
    """Record a CONTR item into plan p: lists maps names to gecco_hip.MEList,
    slot names to buffer slots.  An item with vertex data (lists of several
    vertices) goes through gecco_hip_plan_add_contraction_n."""
    a, b, t = c['op1'], c['op2'], c['target']
    buf = (slot(a), slot(b), slot(t))
    # dagger flags: the flagged lists take part as their adjoint views
    L1, L2, Lt = [lists[n].adjoint() if f else lists[n]
                  for n, f in zip((a, b, t), c.get('tra') or (False, False, False))]
    v = c.get('vertices')
    if v is None:
        return p.add_contraction(L1, L2, Lt, c['cnt_c'], c['cnt_a'], c['factor'], use_tr, buf=buf)
    return p.add_contraction_n(L1, L2, Lt, v['ex1'], v['ex2'], v['cnt'], *v['merge'],
                               fact=c['factor'], use_tr=use_tr, buf=buf)


def record_reorder(p, r, lists, slot):
    """Record an ADD_REO item (REORDER items expand into INIT + ADD_REO)
    into plan p: it runs after the plan's contractions."""
    s, t = r['source'], r['target']
    p.add_reorder(lists[s], lists[t], r['from'], r['to'], r['shift_c'], r['shift_a'], r['factor'],
                  r['sign'], buf=(slot(s), slot(t)))


def _formula_items():
    items = []
    for n in LISTS:
        if n in INIT:
            src, f = INIT[n]
            items.append(_formula.init(n, src, f))
        elif n not in INTEGRALS and n not in ('T1', 'T2'):
            items.append(_formula.init(n))
    items += [_formula.add(*t[1:]) if t[0] == 'ADD' else _formula.contr(*t)
              for level in _TERM_LEVELS for t in level]
    return items


# The CCSD formula ({"lists", "constant", "items"}, see formula.py), its
# schedule, and the contraction tuples and projections per level.
FORMULA = {'lists': LISTS, 'constant': INTEGRALS, 'items': _formula_items()}
SCHEDULE = _formula.schedule(FORMULA['items'])
LEVELS = [[(c['target'], c['op1'], c['op2'], c['cnt_c'], c['cnt_a'], c['factor']) for c in lv['contr']]
          for lv in SCHEDULE]
SYM_AFTER = [lv['sym'] for lv in SCHEDULE]


class CCSD:
    """Device-resident closed-shell coupled-cluster driver.

    ham: Hamiltonian (synth.Hamiltonian / synth.DenseHamiltonian); device:
    torch device; formula: a formula ({"lists", "constant", "items"}, see
    formula.py), by default the CCSD formula above.  The formula's lists
    must include the amplitudes T1, T2, the residuals R1, R2 and the energy
    E; its constant lists are integral lists <c..||a..> named
    H<spaces of C><spaces of A> as above (filled from ham).  plans=False
    sets up the lists, integrals and preconditioner only (for a batched
    driver that builds the plans of several instances together).
    pool=True lets intermediates whose lifetimes do not meet share device
    memory (formula.lifetimes / formula.pool_layout: command_new_intermediate
    and command_del_intermediate); the refined solve, whose linearised
    formula reads every intermediate's final value, then refuses to run.
    spin_project=True projects the residual lists onto their singlet part
    after their last write (gecco_hip_plan_add_spin_project: the reference's
    spin_prj_list for a target with s2 = 0, contract/frm_sched2.f:146-150).
    deterministic=True builds every plan with fixed-order summation
    (gecco_hip_plan_set_deterministic): solves are bit-reproducible.
    """

    def __init__(self, ham, device='cuda:0', use_tr=True, formula=None, plans=True, pool=False,
                 spin_project=False, deterministic=False):
        import torch
        import gecco_hip
        self.torch = torch
        self.deterministic = deterministic
        self.ham = ham
        self.dev = torch.device(device)
        self.use_tr = use_tr
        self.spin_project = spin_project
        self.formula = FORMULA if formula is None else formula
        lists = self.formula['lists']
        for n in ('T1', 'T2', 'R1', 'R2', 'E'):
            if n not in lists:
                raise ValueError(f'formula lacks the list {n}')
        self.constant = list(self.formula['constant'])
        self.orbs = gecco_hip.Orbs(ham.irrep_o, ham.irrep_v, ham.nsym)
        self.L = {n: gecco_hip.MEList(self.orbs, c, a) for n, (c, a) in lists.items()}
        self.slot = {n: i for i, n in enumerate(lists)}
        # T1|T2|... and R1|R2|... are contiguous: the amplitude vector and
        # the residual (without its diagonal part) of the solver; a formula
        # with T3 (CCSDT) adds T3/R3 behind them
        self.tnames = [f'T{k}' for k in range(1, 10) if f'T{k}' in lists]
        self.rnames = [f'R{k}' for k in range(1, len(self.tnames) + 1)]
        for n in self.rnames:
            if n not in lists:
                raise ValueError(f'formula lacks the list {n}')
        self.n1, self.n2 = self.L['T1'].length, self.L['T2'].length
        sizes = [self.L[n].length for n in self.tnames]
        self.seg = {}
        o = 0
        for k, n in enumerate(self.tnames):
            self.seg[n] = self.seg[self.rnames[k]] = (o, o + sizes[k])
            o += sizes[k]
        self.namp = o
        self.amp = torch.zeros(o, dtype=torch.float64, device=self.dev)
        self.res = torch.zeros(o, dtype=torch.float64, device=self.dev)
        # intermediates with a lifetime share one pool when asked to
        self.pool_spans, self.pool_off, self.pool_len = {}, {}, 0
        if pool:
            spans = _formula.lifetimes(self.formula, keep=self.tnames + self.rnames + ['E'])
            self.pool_spans = {n: v for n, v in spans.items() if n not in self.seg}
            self.pool_off, self.pool_len = _formula.pool_layout(
                self.pool_spans, {n: self.L[n].length for n in self.pool_spans})
        self.buf = self._list_buffers(torch.float64, self.amp, self.res)
        t0 = time.perf_counter()
        self._integrals()
        self._denominators()
        self.t_setup_integrals = time.perf_counter() - t0
        t0 = time.perf_counter()
        self.plans = []
        self.flops = 0.0
        self.nblocks = 0
        self.schedule = SCHEDULE if formula is None else _formula.schedule(self.formula['items'])
        if plans:
            self.build_plans()
        self.t_setup_plans = time.perf_counter() - t0
        self.bufs = [self.buf[n] for n in lists]

    def _list_buffers(self, dtype, amp, res, source=None):
        """A buffer per list: amplitude and residual lists as views of amp
        and res, pooled intermediates as views of one pool, the others
        their own zeros (or source[n] converted to dtype: the integrals)."""
        torch = self.torch
        pool = torch.zeros(max(1, self.pool_len), dtype=dtype, device=self.dev) if self.pool_spans else None
        out = {}
        for n in self.formula['lists']:
            ln = max(1, self.L[n].length)
            if n in self.seg:
                v = amp if n[0] == 'T' else res
                out[n] = v[self.seg[n][0]:self.seg[n][1]]
            elif n in self.pool_spans:
                out[n] = pool[self.pool_off[n]:self.pool_off[n] + ln]
            elif source is not None:
                out[n] = source[n].to(dtype)
            else:
                out[n] = torch.zeros(ln, dtype=dtype, device=self.dev)
        return out

    def list_bytes(self):
        """Device bytes of the float64 lists (the pool counted once)."""
        own = sum(max(1, self.L[n].length) for n in self.formula['lists']
                  if n not in self.seg and n not in self.pool_spans)
        return 8 * (own + 2 * self.namp + (max(1, self.pool_len) if self.pool_spans else 0))

    def build_plans(self, tuning=None):
        """One plan per scheduled level of this instance alone (buffer slots
        = list indices); tuning: optional launch settings per level (dicts
        of gecco_hip.Plan.set_tuning arguments, e.g. from tune())."""
        import gecco_hip
        self.plans = []
        self.flops = 0.0
        self.nblocks = 0
        for li, level in enumerate(self.schedule):
            p = gecco_hip.Plan()
            if tuning is not None and tuning[li]:
                p.set_tuning(**tuning[li])
            self.record_level(p, level, 0)
            p.finalize()
            self.plans.append(p)
        self.tuning = tuning
        self._graph = None

    def tune(self, reps=5, axes=None, log=None):
        """Pick each level's launch settings (tile variant, task split,
        contraction chunk) by timing on the live buffers
        (autotune.tune_inplace: the level is recorded once and finalized
        again per candidate); returns the per-level table.  The amplitudes
        are not changed (a level writes intermediates and residuals only)."""
        import gecco_hip
        import autotune
        flops, nblocks = self.flops, self.nblocks
        tuning, report = [], []
        for li, level in enumerate(self.schedule):
            plan = gecco_hip.Plan()
            plan.keep_host(True)
            self.record_level(plan, level, 0)
            plan.finalize()
            # levels before this one leave their outputs in the buffers
            for q in self.plans[:li]:
                q.execute(self.bufs)
            opts, ms, table = autotune.tune_inplace(plan, self.bufs, axes=axes or autotune.AXES,
                                                    reps=reps, log=log)
            plan.keep_host(False)
            self.plans[li] = plan
            tuning.append(opts)
            report.append({'level': li, 'chosen': opts, 'ms': round(ms, 4),
                           'default_ms': round(table[0][1], 4), 'candidates': len(table)})
        self.flops, self.nblocks = flops, nblocks
        self.tuning = tuning
        self._graph = None
        return report

    def record_level(self, p, level, base):
        """Record one scheduled level of this instance into plan p, its
        lists in buffer slots base + list index: constant slots, list
        initialisations, contractions and spin-flip projections."""
        if getattr(self, 'deterministic', False):   # fixed-order sums, bit-reproducible
            p.set_deterministic(True)
        for n in self.constant:       # fixed during the solve
            p.set_constant(base + self.slot[n])
        # list initialisations of the level: from a source or zero
        for it in level['init']:
            n = it['target']
            if 'source' in it:
                p.add_init(base + self.slot[n], self.L[n].length, buf_src=base + self.slot[it['source']],
                           factor=it['factor'])
            else:
                p.add_init(base + self.slot[n], self.L[n].length)
        for it in level.get('add', []):
            n, src = it['target'], it['source']
            if self.L[src].length != self.L[n].length:
                raise ValueError(f'ADD {src} into {n}: lists of different layout')
            p.add_sum(base + self.slot[n], self.L[n].length, base + self.slot[src], factor=it['factor'])
        for c in level['contr']:
            nb, fl = record_contraction(p, c, self.L, lambda n: base + self.slot[n], self.use_tr)
            self.flops += fl
            self.nblocks += nb
        for r in level.get('reo', []):
            record_reorder(p, r, self.L, lambda n: base + self.slot[n])
        if self.use_tr:
            for res in level['sym']:
                p.add_sym_ab(self.L[res], base + self.slot[res])
        if self.spin_project:
            for res in level['sym']:
                if res in self.rnames:
                    p.add_spin_project(self.L[res], base + self.slot[res])

    def solve_mixed(self, tol=1e-8, switch=1e-6, maxit=100, ndiis=8, mode='graph', verbose=False,
                    stall=6, bulk=None):
        """Mixed-precision solve: float32 residuals (solve(precision='f32'))
        until ||g|| < switch (or the float32 residual stalls: `stall`
        iterations without a 10% gain), then float64 iterations from those
        amplitudes with a fresh DIIS subspace until ||g|| < tol -- iterative
        refinement of the amplitude residual, whose converged point is the
        float64 one.  At nocc=20, nvir=100 the float32 residual levels off
        near 1e-6; a fresh subspace then needs fewer float64 iterations than
        continuing the float32 one.
        Returns the float64 phase's result with 'f32_iterations',
        'f32_seconds' and the total 'seconds' / 'iterations'."""
        if bulk is not None and getattr(self, 'bulk', None) != bulk:
            self.build_f32(bulk=bulk)       # bulk: 'f32' or 'bf16x3' float32 plans
            self._graph = None
        r32 = self.solve(tol=switch, maxit=maxit, ndiis=ndiis, mode=mode, precision='f32',
                         verbose=verbose, stall=stall)
        r64 = self.solve(tol=tol, maxit=maxit, ndiis=ndiis, mode=mode, precision='f64',
                         keep_amplitudes=True, verbose=verbose)
        out = dict(r64)
        out.update(f32_iterations=r32['iterations'], f32_seconds=r32['seconds'],
                   f64_iterations=r64['iterations'], f64_seconds=r64['seconds'],
                   iterations=r32['iterations'] + r64['iterations'],
                   seconds=r32['seconds'] + r64['seconds'], precision='mixed', switch=switch)
        return out

    # ------------------------------------------------------------ setup
    def _integrals(self):
        """<pq||rs> in every integral list from the DF factors, on the device."""
        torch = self.torch
        ham = self.ham
        n = ham.nocc + ham.nvir
        if getattr(ham, 'eri_full', None) is not None:   # dense (pq|rs) given
            eri = torch.from_numpy(ham.eri_full).to(self.dev)
        else:                                             # density fitting
            B = torch.from_numpy(ham.B.reshape(ham.B.shape[0], n * n)).to(self.dev)
            eri = (B.T @ B).reshape(n, n, n, n)      # (pq|rs)
            del B
        for name in self.constant:
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
        """Diagonal preconditioner of T1|T2 (dia4op_ev): eps of the C strings
        minus eps of the A strings, built by the library on the host."""
        ham = self.ham
        d = np.concatenate([self.L[n].diagonal(ham.eps_o, ham.eps_v) for n in self.tnames])
        self.diag = self.torch.from_numpy(d).to(self.dev)

    # ---------------------------------------------------------- iteration
    def evaluate(self, stream=None):
        """Energy and residual lists for the amplitudes in T1/T2 (one pass of
        the formula: list initialisation, three grouped contraction levels,
        spin-flip symmetrisation -- all native launches)."""
        for p in self.plans:
            p.execute(self.bufs, stream)

    # ------------------------------------------------------ float32 lists
    def build_f32(self, plans=True, bulk='f32'):
        """Float32 copies of every list (integrals rounded once) and plans
        that run the formula on them: the bulk of a mixed-precision solve
        (plans=False: the lists only, for a batched driver; bulk 'bf16x3':
        the GEMM tiles as split-bfloat16 three-term products)."""
        import gecco_hip
        torch = self.torch
        lists = self.formula['lists']
        self.amp32 = torch.zeros(self.namp, dtype=torch.float32, device=self.dev)
        self.res32 = torch.zeros_like(self.amp32)
        self.buf32 = self._list_buffers(torch.float32, self.amp32, self.res32, source=self.buf)
        self.bufs32 = [self.buf32[n] for n in lists]
        if not plans:
            return
        flops, nblocks = self.flops, self.nblocks
        self.plans32 = []
        for li, level in enumerate(self.schedule):
            p = gecco_hip.Plan()
            p.set_precision(bulk)
            self.record_level(p, level, 0)
            p.finalize()
            self.plans32.append(p)
        self.bulk = bulk
        self.flops, self.nblocks = flops, nblocks

    # ------------------------------------------- refinement of the residual
    def build_linear(self, bulk='f32'):
        """Plans of the formula linearised about the current lists
        (formula.linearize over T1, T2): base lists are the float32 lists
        (build_f32), each list X that depends on the amplitudes gets a float32
        delta list dX, and one pass gives dR1|dR2 = J dT1|dT2, the directional
        derivative of the residual lists along the amplitude step in
        dT1|dT2.  bulk: 'f32' or 'bf16x3' (split-BF16 GEMM tiles)."""
        import gecco_hip
        torch = self.torch
        if self.pool_spans:
            raise ValueError('build_linear: the linearised formula reads every intermediate at the '
                             'current amplitudes; pooled intermediates (pool=True) do not keep them')
        if getattr(self, 'buf32', None) is None:
            self.build_f32(plans=False)
        items, dnames = _formula.linearize(self.formula, self.tnames)
        lists = list(self.formula['lists'])
        nl = len(lists)
        base = {d: d[1:] for d in dnames}
        self.damp32 = torch.zeros(self.namp, dtype=torch.float32, device=self.dev)
        self.dres32 = torch.zeros_like(self.damp32)
        dbuf = {}
        for d in dnames:
            n = base[d]
            if n in self.seg:
                v = self.damp32 if n[0] == 'T' else self.dres32
                dbuf[d] = v[self.seg[n][0]:self.seg[n][1]]
            else:
                dbuf[d] = torch.zeros(max(1, self.L[n].length), dtype=torch.float32, device=self.dev)
        slot = dict(self.slot)
        slot.update({d: nl + k for k, d in enumerate(dnames)})
        self.lin_bufs = self.bufs32 + [dbuf[d] for d in dnames]
        lst = lambda x: self.L[base.get(x, x)]
        flops = 0.0
        self.lin_plans = []
        for lv in _formula.schedule(items):
            p = gecco_hip.Plan()
            p.set_precision(bulk)
            if self.deterministic:
                p.set_deterministic(True)
            for n in self.constant:
                p.set_constant(slot[n])
            for it in lv['init']:
                if 'source' in it:
                    p.add_init(slot[it['target']], lst(it['target']).length, buf_src=slot[it['source']],
                               factor=it['factor'])
                else:
                    p.add_init(slot[it['target']], lst(it['target']).length)
            for it in lv['add']:
                p.add_sum(slot[it['target']], lst(it['target']).length, slot[it['source']],
                          factor=it['factor'])
            lists = {n: lst(n) for n in slot}
            for c in lv['contr']:
                nb, fl = record_contraction(p, c, lists, slot.__getitem__, self.use_tr)
                flops += fl
            for r in lv.get('reo', []):
                record_reorder(p, r, lists, slot.__getitem__)
            if self.use_tr:
                for x in lv['sym']:
                    p.add_sym_ab(lst(x), slot[x])
            p.finalize()
            self.lin_plans.append(p)
        self.lin_flops = flops
        self.lin_bulk = bulk
        self._lin_state = None      # its captured graph ran the previous plans

    def _gradient64(self):
        """One float64 evaluation at the current amplitudes: g = D t + R and
        the energy (a float64 residual evaluation)."""
        self.evaluate()
        g = self.diag * self.amp + self.res
        return g, float(self.buf['E'][0])

    def _solve_linear(self, g, tol, maxit=60, ndiis=8):
        """Solve J d = -g for the amplitude step d (float64 vector; J from
        the linearised plans about the current lists, in the plans' bulk
        precision): DIIS with the diagonal preconditioner on the linear
        function g + J d = D d + (g + dR(d)).  The iteration is captured once
        as a HIP graph (per tolerance) over persistent vectors and replayed;
        each solve resets the subspace and starts from d = 0.  Returns (d,
        iterations, ||g + J d||); d is overwritten by the next solve."""
        import gecco_hip
        torch = self.torch
        st = getattr(self, '_lin_state', None)
        if st is None or st['ndiis'] != ndiis:
            n = self.amp.numel()
            st = dict(ndiis=ndiis, delta=torch.zeros(n, dtype=torch.float64, device=self.dev),
                      rest=torch.zeros(n, dtype=torch.float64, device=self.dev),
                      g=torch.zeros(n, dtype=torch.float64, device=self.dev),
                      diis=gecco_hip.DIIS(n, ndiis), graph=None)
            self._lin_state = st
        delta, rest, gb, diis = st['delta'], st['rest'], st['g'], st['diis']
        gb.copy_(g)
        delta.zero_()
        torch.cuda.synchronize()
        diis.reset()

        def iteration(s):
            self.damp32.copy_(delta)
            for p in self.lin_plans:
                p.execute(self.lin_bufs, s)
            rest.copy_(self.dres32)
            rest.add_(gb)
            diis.step_async(delta, rest, self.diag, tol=tol, stream=s)
        stream = torch.cuda.current_stream()
        graph = st['graph'][1] if st['graph'] is not None and st['graph'][0] == tol else None
        events = []
        rn, it, done = float('inf'), 0, 0
        for it in range(1, maxit + 1):
            if graph is None:
                iteration(stream.cuda_stream)
                stream.synchronize()
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph):
                    iteration(torch.cuda.current_stream().cuda_stream)
                st['graph'] = (tol, graph)
            else:
                graph.replay()
            ev = torch.cuda.Event()
            ev.record()
            events.append(ev)
            stop = False
            while done < it - 1 or (it == maxit and done < it):
                events[done].synchronize()
                rn, _, code = diis.status(done)
                done += 1
                if code < 0:
                    raise gecco_hip.GeccoHipError(f'linear DIIS matrix is singular at iteration {done}')
                if code == 1:
                    stop = True
                    break
            if stop:
                it = done
                break
        torch.cuda.synchronize()
        return delta, it, rn

    def solve_refined(self, tol=1e-8, switch=1e-6, maxit=100, bulk='f32', inner_bulk=None,
                      max_f64=4, stall=6, verbose=False):
        """Mixed-precision solve with float64 kept for the energy, the DIIS
        solves and the last residual evaluations:
          1. float32 (or split-BF16) residuals with float64 DIIS until
             ||g|| < switch or the low-precision residual stalls
             (solve(precision='f32'));
          2. a float64 evaluation g(t) and the energy;
          3. the step d from J d = -g, J the formula linearised about t
             (build_linear), solved in the low precision -- d is small, so
             its rounding error is small in absolute terms -- then t += d;
          4. a float64 evaluation; done when ||g|| < tol, else back to 3
             (at most max_f64 float64 evaluations).
        This is iterative refinement of the amplitude residual: each float64
        evaluation corrects the low-precision Jacobian's step, and from a
        residual near the float32 floor one correction reaches 1e-8 or
        better, so a solve normally makes two float64 evaluations.
        Returns the result of the last evaluation with the counts of each
        phase."""
        torch = self.torch
        inner_bulk = inner_bulk or bulk
        fresh = getattr(self, 'bulk', None) != bulk or getattr(self, 'plans32', None) is None
        if fresh:      # new float32 lists: the linearised plans read them too
            self.build_f32(bulk=bulk)
            self._graph = None
        if fresh or getattr(self, 'lin_bulk', None) != inner_bulk:
            self.build_linear(inner_bulk)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r32 = self.solve(tol=switch, maxit=maxit, precision='f32', stall=stall, verbose=verbose)
        inner = []
        nf64 = 0
        while True:
            g, energy = self._gradient64()
            nf64 += 1
            gn = float(g.norm())
            if verbose:
                print(f'  float64 evaluation {nf64}: E = {energy:.12f}  |g| = {gn:.3e}', flush=True)
            if gn < tol or nf64 >= max_f64:
                break
            # base lists: the float64 lists at t, rounded once
            for n in self.formula['lists']:
                if n not in self.constant and n not in self.seg:
                    self.buf32[n].copy_(self.buf[n])
            self.amp32.copy_(self.amp)
            self.res32.copy_(self.res)
            d, k, rn = self._solve_linear(g, 0.3 * tol, maxit=maxit)
            inner.append(k)
            self.amp.add_(d)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return dict(energy=energy, residual=gn, seconds=wall, f64_evaluations=nf64,
                    f32_iterations=r32['iterations'], linear_iterations=inner,
                    iterations=r32['iterations'] + sum(inner), precision='refined', bulk=bulk,
                    inner_bulk=inner_bulk, switch=switch, converged=gn < tol)

    def evaluate32(self, stream=None):
        """evaluate() on the float32 lists: T1|T2 rounded to float32, the
        residual and energy returned in float64."""
        if getattr(self, 'plans32', None) is None:
            self.build_f32()
        self.amp32.copy_(self.amp)
        for p in self.plans32:
            p.execute(self.bufs32, stream)
        self.res.copy_(self.res32)
        self.buf['E'].copy_(self.buf32['E'])

    def solve(self, tol=1e-8, maxit=100, ndiis=8, verbose=False, on_iter=None, restart=True,
              mode='graph', precond=None, fallback=2.0, precision='f64', keep_amplitudes=False,
              stall=None):
        """Iterate to ||g|| < tol, g = D*t + R over T1|T2 (R: the formula's
        residual lists, D: the diagonal preconditioner), at most maxit
        iterations, with the native DIIS step (gecco_hip.DIIS).  restart:
        start from T = 0 with an empty subspace; otherwise continue.

        mode 'sync': the DIIS step solves its small system on the host, one
        synchronisation per iteration.  'async': the whole step stays on the
        device; the host reads iteration i's norms (from pinned memory) while
        iteration i+1 runs, so the GPU queue never drains and one iteration
        after the converged one is also run (its DIIS step meets the
        tolerance again and leaves T unchanged).  'graph': as 'async', with
        the iteration (three plan executes and the DIIS step) captured once
        as a HIP graph and replayed.  Returns a dict with energy, iterations,
        residual norm and timings.

        precond: a learned.LearnedPreconditioner; each step then uses
        r = s * g / D with s from the network (mode 'graph' runs as
        'async': the iteration index is a kernel argument).  Fallback to the
        plain preconditioner for the rest of the solve once ||g|| exceeds
        `fallback` times its smallest value so far (from the fourth
        iteration on) or is not finite; the result reports the iteration
        ('fallback_at', None if the network stayed on).

        precision 'f32': the residual comes from the float32 lists
        (evaluate32); the amplitudes, the DIIS step and the convergence test
        stay float64.  keep_amplitudes: with restart, start the new subspace
        from the current amplitudes instead of T = 0 (see solve_mixed).
        stall: also stop once ||g|| has not fallen below 0.9 times its best
        value for `stall` iterations (the float32 residual's floor)."""
        import gecco_hip
        torch = self.torch
        if mode not in ('sync', 'async', 'graph'):
            raise ValueError(f'unknown solve mode {mode!r}')
        if precond is not None and mode == 'graph':
            mode = 'async'
        want_async = mode != 'sync'
        d = getattr(self, '_diis', None)
        if precision not in ('f64', 'f32'):
            raise ValueError(f'unknown precision {precision!r}')
        evaluate = self.evaluate32 if precision == 'f32' else self.evaluate
        if precision == 'f32' and getattr(self, 'plans32', None) is None:
            self.build_f32()
        if restart or d is None or d.maxsub != ndiis or getattr(self, '_diis_async', None) != want_async:
            if not keep_amplitudes:
                self.amp.zero_()
            self._diis = gecco_hip.DIIS(self.amp.numel(), ndiis)
            self._diis_async = want_async
            self._diis_steps = 0
            self._graph = None
        diis = self._diis
        stream = torch.cuda.current_stream().cuda_stream
        # learned preconditioner: per-amplitude factors s, refilled before
        # every step
        learn = {'on': precond is not None, 'best': float('inf'), 'at': None}
        if precond is not None:
            if getattr(self, '_scale', None) is None:
                self._scale = torch.empty_like(self.amp)
            diis.set_scale(self._scale)
        else:
            diis.set_scale(None)

        def scale_step(k, s):
            if learn['on']:
                precond.scale(self.diag, self.res, self.amp, k, self._scale, s)

        prog = {'best': float('inf'), 'at': 0, 'stalled': False}

        def stalled(k, rn):
            if rn < 0.9 * prog['best']:
                prog['best'], prog['at'] = rn, k
            prog['stalled'] = stall is not None and k - prog['at'] >= stall
            return prog['stalled']

        def watch(k, rn):
            if not learn['on']:
                return
            if not math.isfinite(rn) or (k >= 4 and rn > fallback * learn['best']):
                learn['on'] = False
                learn['at'] = k
                diis.set_scale(None)
            learn['best'] = min(learn['best'], rn)
        energy = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it = 0
        rnorm = float('inf')
        if mode == 'sync':
            for it in range(1, maxit + 1):
                evaluate(stream)
                scale_step(it - 1, stream)
                rnorm, energy, conv = diis.step(self.amp, self.res, self.diag, energy=self.buf['E'],
                                                tol=tol, stream=stream)
                watch(it, rnorm)
                if stalled(it, rnorm):
                    break
                if on_iter:
                    on_iter(it, energy, rnorm)
                if verbose:
                    print(f'  it {it:3d}  E = {energy:.12f}  |R| = {rnorm:.3e}', flush=True)
                if conv:
                    break
        else:
            def iteration(s, k=0):
                evaluate(s)
                scale_step(k, s)
                diis.step_async(self.amp, self.res, self.diag, energy=self.buf['E'], tol=tol, stream=s)
            graph = None
            if mode == 'graph':
                key = (tol, stream, precision)
                if self._graph is not None and self._graph[0] == key:
                    graph = self._graph[1]
            events = []
            base = self._diis_steps
            done = 0
            for it in range(1, maxit + 1):
                if graph is not None:
                    graph.replay()
                else:
                    iteration(stream, it - 1)
                    if mode == 'graph':
                        # captured after one eager iteration (a plan's first
                        # execute uploads its buffer table and constant
                        # gathers); replayed from the next iteration on
                        torch.cuda.current_stream().synchronize()
                        graph = torch.cuda.CUDAGraph()
                        with torch.cuda.graph(graph):
                            iteration(torch.cuda.current_stream().cuda_stream)
                        self._graph = ((tol, stream, precision), graph)
                ev = torch.cuda.Event()
                ev.record()
                events.append(ev)
                self._diis_steps += 1
                # the previous iteration's record, while this one runs
                conv = False
                while done < it - 1 or (it == maxit and done < it):
                    events[done].synchronize()
                    rnorm, energy, code = diis.status(base + done)
                    done += 1
                    watch(done, rnorm)
                    if stalled(done, rnorm):
                        conv = True
                        break
                    if on_iter:
                        on_iter(done, energy, rnorm)
                    if verbose:
                        print(f'  it {done:3d}  E = {energy:.12f}  |R| = {rnorm:.3e}', flush=True)
                    if code < 0:
                        # the reference stops on a singular subspace matrix
                        # (solve/optc_diis_extr.f:70); so does the sync step
                        torch.cuda.synchronize()
                        raise gecco_hip.GeccoHipError(
                            f'DIIS matrix is singular at iteration {done} (||g|| = {rnorm:.3e})')
                    if code != 0:
                        conv = True
                        break
                if conv:
                    it = done
                    break
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        diis.set_scale(None)
        return dict(energy=energy, iterations=it, residual=rnorm, seconds=wall, precision=precision,
                    stalled=prog['stalled'],
                    iters_per_s=it / wall, flops_per_iter=self.flops, kernel_calls=self.nblocks,
                    preconditioner='learned' if precond is not None else 'diagonal',
                    fallback_at=learn['at'])
