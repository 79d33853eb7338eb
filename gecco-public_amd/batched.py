"""Many independent coupled-cluster instances (different molecules or
orbital spaces, uneven sizes) iterated together: the kernel calls of every
instance's formula level run in one grouped launch, and the whole batched
iteration -- the level launches and every instance's asynchronous DIIS step
-- is captured once as a HIP graph and replayed.

Each instance keeps its own ME-lists, integrals, preconditioner and DIIS
subspace (ccsd.CCSD with plans=False); instance k's lists take the buffer
slots k * nlists + list index of the shared plans and it is instance group
k of every plan.  Each instance's asynchronous DIIS step sets its entry of a
device flag array when it meets the tolerance (leaving the amplitudes
unchanged; all instances' DIIS steps run as one group of three launches,
gecco_hip.DIISGroup), and every launch skips the work of the flagged
instances, so a converged instance costs no arithmetic while the replayed
graph stays the same.  Once the host has seen a quarter of the remaining
instances converge, the plans switch to live-task lists (Plan.compact: the
tasks of the unconverged instances only) and the iteration is captured
again, so that converged instances stop costing workgroups too.  The batch
iterates until every instance has converged or maxit is reached.
"""
import time

import ccsd


class BatchedCCSD:
    def __init__(self, hams, device='cuda:0', use_tr=True, formula=None, on_instance=None,
                 skip_converged=True, compact=True):
        """on_instance(k, instance), if given, is called after instance k's
        lists and integrals are set up.  skip_converged=False keeps
        evaluating converged instances; compact=False keeps launching the
        full task lists (converged instances' workgroups only read their
        flag)."""
        import gecco_hip
        import torch
        self.torch = torch
        self.inst = []
        for k, h in enumerate(hams):
            self.inst.append(ccsd.CCSD(h, device, use_tr, formula, plans=False))
            if on_instance is not None:
                on_instance(k, self.inst[-1])
        nl = len(self.inst[0].formula['lists'])
        self.nlists = nl
        t0 = time.perf_counter()
        # converged flags, one per instance (set by the DIIS steps)
        self.done = torch.zeros(len(self.inst), dtype=torch.int32, device=device)
        self.skip_converged = skip_converged
        self.compact = compact and skip_converged
        self.compactions = 0
        self.plans = []
        for li in range(len(self.inst[0].schedule)):
            p = gecco_hip.Plan()
            for k, c in enumerate(self.inst):
                p.set_group(k)
                c.record_level(p, c.schedule[li], k * nl)
            p.finalize()
            if skip_converged:
                p.set_skip(self.done)
            self.plans.append(p)
        self.t_setup_plans = time.perf_counter() - t0
        self.bufs = [b for c in self.inst for b in c.bufs]
        self.flops = sum(c.flops for c in self.inst)
        self.nblocks = sum(c.nblocks for c in self.inst)
        self._graph = None
        self._diis = None

    def _record(self, precision):
        import gecco_hip
        nl = self.nlists
        plans = []
        for li in range(len(self.inst[0].schedule)):
            p = gecco_hip.Plan()
            p.set_precision(precision)
            for k, c in enumerate(self.inst):
                p.set_group(k)
                c.record_level(p, c.schedule[li], k * nl)
            p.finalize()
            if self.skip_converged:
                p.set_skip(self.done)
            plans.append(p)
        return plans

    def build_f32(self):
        """Float32 lists of every instance and the combined float32 plans
        (same instance groups and skip flags)."""
        saved = [(c.flops, c.nblocks) for c in self.inst]
        for c in self.inst:
            c.build_f32(plans=False)
        self.bufs32 = [b for c in self.inst for b in c.bufs32]
        self.plans32 = self._record('f32')
        for c, (f, nb) in zip(self.inst, saved):
            c.flops, c.nblocks = f, nb

    def evaluate(self, stream=None, precision='f64'):
        if precision == 'f32':
            if getattr(self, 'plans32', None) is None:
                self.build_f32()
            for c in self.inst:
                c.amp32.copy_(c.amp)
            for p in self.plans32:
                p.execute(self.bufs32, stream)
            for c in self.inst:
                c.res.copy_(c.res32)
                c.buf['E'].copy_(c.buf32['E'])
            return
        for p in self.plans:
            p.execute(self.bufs, stream)

    def tune(self, reps=3, log=None):
        """Launch settings of every combined level plan by timing
        (autotune.tune_inplace on a re-recorded plan that keeps its host
        data); returns the per-level choices."""
        import gecco_hip
        import autotune
        nl = self.nlists
        report = []
        for li in range(len(self.plans)):
            p = gecco_hip.Plan()
            p.keep_host(True)
            saved = [(c.flops, c.nblocks) for c in self.inst]
            for k, c in enumerate(self.inst):
                p.set_group(k)
                c.record_level(p, c.schedule[li], k * nl)
            for c, (f, nb) in zip(self.inst, saved):
                c.flops, c.nblocks = f, nb
            p.finalize()
            if self.skip_converged:
                p.set_skip(self.done)
            for q in self.plans[:li]:
                q.execute(self.bufs)
            opts, ms, table = autotune.tune_inplace(p, self.bufs, reps=reps, log=log)
            p.keep_host(False)
            self.plans[li] = p
            report.append({'level': li, 'chosen': opts, 'ms': round(ms, 3),
                           'default_ms': round(table[0][1], 3), 'candidates': len(table)})
        self._graph = None
        return report

    def solve(self, tol=1e-8, maxit=100, ndiis=8, mode='graph', restart=True, timing=False,
              precision='f64', keep_amplitudes=False):
        """Iterate every instance to ||g|| < tol (at most maxit iterations);
        mode 'graph' replays the captured batched iteration, 'async' launches
        it eagerly.  Returns per-instance energies, iterations and residuals,
        and the batch's wall time (timing=True: also the GPU milliseconds of
        every batched iteration, 'iteration_ms').  precision 'f32': float32
        lists (evaluate(precision='f32')), float64 amplitudes and DIIS;
        keep_amplitudes: restart the subspaces from the current amplitudes."""
        import gecco_hip
        torch = self.torch
        if mode not in ('async', 'graph'):
            raise ValueError(f'unknown solve mode {mode!r}')
        if precision == 'f32' and getattr(self, 'plans32', None) is None:
            self.build_f32()
        if restart or self._diis is None:
            for c in self.inst:
                if not keep_amplitudes:
                    c.amp.zero_()
            self._diis = [gecco_hip.DIIS(c.amp.numel(), ndiis) for c in self.inst]
            # full task lists before the flags go back to 0
            for ps in (self.plans, getattr(self, 'plans32', None) or []):
                for p in ps:
                    p.compact(False)
            self.done.zero_()
            for k, d in enumerate(self._diis):
                d.set_flag(self.done[k:k + 1])
            # every instance's DIIS step in three launches
            self._group = gecco_hip.DIISGroup(self._diis, [c.amp for c in self.inst],
                                              [c.res for c in self.inst], [c.diag for c in self.inst],
                                              [c.buf['E'] for c in self.inst])
            self._steps = 0
            self._graph = None
        stream = torch.cuda.current_stream().cuda_stream

        def iteration(s):
            self.evaluate(s, precision)
            self._group.step_async(tol, s)

        n = len(self.inst)
        live_at_compaction = n
        conv_at = [None] * n
        singular_at = [None] * n
        energy = [None] * n
        resid = [None] * n
        events = []
        base = self._steps
        done = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=timing)
        ev0.record()
        graph = self._graph[1] if (mode == 'graph' and self._graph is not None and
                                   self._graph[0] == (tol, precision)) else None
        it = 0
        for it in range(1, maxit + 1):
            if graph is not None:
                graph.replay()
            else:
                iteration(stream)
                if mode == 'graph':
                    torch.cuda.current_stream().synchronize()
                    graph = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(graph):
                        iteration(torch.cuda.current_stream().cuda_stream)
                    self._graph = ((tol, precision), graph)
            ev = torch.cuda.Event(enable_timing=timing)
            ev.record()
            events.append(ev)
            self._steps += 1
            finished = False
            while done < it - 1 or (it == maxit and done < it):
                events[done].synchronize()
                for k, d in enumerate(self._diis):
                    if conv_at[k] is not None:
                        continue
                    g, e, code = d.status(base + done)
                    energy[k], resid[k] = e, g
                    if code < 0 and singular_at[k] is None:
                        # a singular subspace matrix stops this instance as
                        # the reference stops (solve/optc_diis_extr.f:70);
                        # its amplitudes stay, it is not converged
                        singular_at[k] = done + 1
                    if code != 0:
                        conv_at[k] = done + 1
                done += 1
                if all(x is not None for x in conv_at):
                    finished = True
                    break
            if finished:
                break
            live = sum(x is None for x in conv_at)
            if self.compact and live <= 0.75 * live_at_compaction:
                # live-task lists of the unconverged instances (made on the
                # stream after the iterations queued so far), captured anew
                for p in (self.plans32 if precision == 'f32' else self.plans):
                    p.compact(True, stream)
                live_at_compaction = live
                self.compactions += 1
                graph = None
                self._graph = None
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out = dict(energies=energy, iterations=[x if x is not None else it for x in conv_at],
                   residuals=resid, seconds=wall, singular_at=singular_at,
                   converged=[x is not None and s is None for x, s in zip(conv_at, singular_at)],
                   batch_iterations=done, flops_per_iter=self.flops, kernel_calls=self.nblocks)
        if timing:
            out['iteration_ms'] = [a.elapsed_time(b) for a, b in zip([ev0] + events[:-1], events)]
        return out

    def solve_mixed(self, tol=1e-8, switch=1e-5, maxit=100, maxit_f32=60, ndiis=8, mode='graph'):
        """Float32 lists until every instance has ||g|| < switch (at most
        maxit_f32 batched iterations; an instance that meets switch is
        skipped meanwhile), then float64 refinement from those amplitudes
        with fresh subspaces to ||g|| < tol."""
        r32 = self.solve(tol=switch, maxit=maxit_f32, ndiis=ndiis, mode=mode, precision='f32')
        r64 = self.solve(tol=tol, maxit=maxit, ndiis=ndiis, mode=mode, precision='f64',
                         keep_amplitudes=True)
        out = dict(r64)
        out.update(f32_batch_iterations=r32['batch_iterations'], f32_seconds=r32['seconds'],
                   f64_batch_iterations=r64['batch_iterations'], f64_seconds=r64['seconds'],
                   seconds=r32['seconds'] + r64['seconds'], switch=switch,
                   f32_converged=r32['converged'])
        return out
