"""Benchmark of the MI355X block-contraction engine.

Default workload (--workload ccsd): closed-shell CCSD amplitude iterations on
the synthetic Hamiltonian nocc=20, nvir=100, naux=300, 4 IRREPs
(round-robin), eps_occ ~ U[-2,-0.3], eps_vir ~ U[0.2,3.0], B ~ N(0,0.03^2),
GeCCo Ms/IRREP-blocked ME-lists, spin-flip symmetry on, FP64.  One step =
one iteration: the whole CCSD formula (37 binary contractions, 4640 kernel
calls of GeCCo's enumerator, 21.0 GFLOP, three plan executes with list
initialisation, grouped contraction launches and spin-flip
symmetrisation) and the native DIIS step, operands resident in HBM.  After the timed steps a fresh solve runs to
||Omega|| < 1e-8 and its energy is compared with the oracle's FP64 energy
(tests/golden/ccsd_o20_v100_s4.json) when that fixture matches the recipe.

--workload ladder: the particle-particle ladder contraction alone (8 kernel
calls, 5.58 GFLOP) at the same sizes.

--workload triples: the CCSDT triples contractions T3 <- V.T3 (pp ladder,
hh ladder, ring) and T3 <- V.T2 at nocc=12, nvir=60, naux=180, 4 IRREPs
(gecco-public_amd/formulas/ccsdt_triples.json): one step = one evaluation of
R3 (initialisation, one grouped launch for the five terms, spin-flip
symmetrisation) with random spin-symmetric T2 and T3.

--workload learned: the learned per-amplitude preconditioner (trained by
gecco-public_amd/train_precond.py, parameters in --precond) against the
diagonal one on closed-shell CCSD at nocc=30, nvir=150, naux=450, 4 IRREPs:
iterations and time to ||Omega|| < 1e-8 with each, the energy difference,
and the network's inference kernel.  One step = one learned-preconditioner
iteration (graph mode is not used: the iteration index is an argument).

--workload batched: 64 independent CCSD instances of uneven size (nocc ~
U[8,24], nvir ~ U[40,160], naux = 3 nvir, 4 IRREPs, seeds 0..63) in one
grouped launch per formula level and one replayed graph per batched
iteration; also their time to convergence against solving them one after
another.

Prints one JSON line.  --gpus N > 1 is not supported: the engine runs on one
GPU here.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'gecco-public_amd'))
sys.path.insert(0, ROOT)

NOCC, NVIR, NAUX, NSYM, SEED, SIGMA = 20, 100, 300, 4, 0, 0.03
T3_NOCC, T3_NVIR, T3_NAUX = 12, 60, 180
L_NOCC, L_NVIR, L_NAUX = 30, 150, 450
PRECOND = os.path.join(ROOT, 'gecco-public_amd', 'precond', 'ccsd_mlp.npz')
PEAK_FP64 = 78.6        # TFLOP/s, MI355X dense FP64 (vector and MFMA)
PEAK_HBM = 8000.0       # GB/s
METRIC = 'CCSD iters/s & contraction TFLOP/s vs plain FP64 GPU at |dE|<=1e-8 Eh'
# HBM bytes per launch of the ladder kernel at nocc=20 nvir=100 nsym=4, from
# rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 (copy-calibrated factors);
# see profiles/grouped_contract_affine_f64_ccsd_ladder_o20_v100_s4_pmc.md
# HBM bytes per dispatch of the dominant kernel from rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes scaled by a plain-copy calibration (summaries in
# profiles/); for the CCSD iteration it is the average over the three
# formula-level launches of the strided-GEMM kernel
MEASURED_TRAFFIC = {('ladder', 20, 100, 4): 396.5e6, ('ladder', 32, 96, 1): 5.665e9,
                    ('ccsd', 20, 100, 300, 4, 'grouped_contract_affine<64, 64, 16, double, 1, 2, false>'): 658.6e6}
GOLDEN = os.path.join(ROOT, 'tests', 'golden', 'ccsd_o20_v100_s4.json')


def round_robin(n, nsym, start=0):
    return [((start + k) % nsym) + 1 for k in range(n)]


# ------------------------------------------------------------------- ladder
def build_ladder(nocc, nvir, nsym):
    import gecco_hip
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    V = gecco_hip.MEList(o, {2: 2}, {2: 2})
    T = gecco_hip.MEList(o, {2: 2}, {1: 2})
    R = gecco_hip.MEList(o, {2: 2}, {1: 2})
    pc = gecco_hip.Plan()
    nb, fl = pc.add_contraction(V, T, R, {}, {2: 2}, 1.0, True, buf=(0, 1, 2))
    pc.finalize()
    ps = gecco_hip.Plan()
    ps.add_sym_ab(R, 2)
    ps.finalize()
    pv = gecco_hip.Plan()
    pv.add_sym_ab(V, 0)
    pv.finalize()
    pt = gecco_hip.Plan()
    pt.add_sym_ab(T, 1)
    pt.finalize()
    return dict(o=o, V=V, T=T, R=R, pc=pc, ps=ps, pv=pv, pt=pt, nblocks=nb, flops=fl)


def cpu_ladder(nocc, nvir, nsym):
    """Oracle route 3 (collect, dgemm as a plain loop, scatter; one core) on
    one full ladder step."""
    import numpy as np
    from oracle.melist import Orbitals, MEList
    from oracle.contract import Contraction
    orbs = Orbitals(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    T = MEList(orbs, {2: 2}, {1: 2})
    V = MEList(orbs, {2: 2}, {2: 2})
    R = MEList(orbs, {2: 2}, {1: 2})
    c = Contraction(V, T, R, {}, {2: 2})
    blocks = c.blocks(True)
    rng = np.random.default_rng(1)
    xv = rng.standard_normal(V.length)
    xt = rng.standard_normal(T.length)
    xr = np.zeros(R.length)
    fl = sum(2.0 * np.prod(b['lstr_ex1']) * np.prod(b['lstr_ex2']) * np.prod(b['lstr_cnt']) for b in blocks)
    t0 = time.perf_counter()
    c.run(xv, xt, xr, route=3, use_tr=False, screen=False, blocks=blocks)
    dt = time.perf_counter() - t0
    return fl / dt / 1e12, dt, (f'one full step: {len(blocks)} kernel calls ({fl / 1e9:.2f} GFLOP), '
                                'oracle route 3 with a plain-loop dgemm')


def run_ladder(args):
    import torch
    dev = torch.device('cuda:0')
    w = build_ladder(args.nocc, args.nvir, args.nsym)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    xv = torch.randn(w['V'].length, dtype=torch.float64, device=dev, generator=g) * 0.05
    xt = torch.randn(w['T'].length, dtype=torch.float64, device=dev, generator=g) * 0.01
    xr = torch.zeros(w['R'].length, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    w['pv'].execute([xv], sh)
    w['pt'].execute([xv, xt], sh)
    bufs = [xv, xt, xr]
    for _ in range(args.warmup):
        w['pc'].execute(bufs, sh)
        w['ps'].execute(bufs, sh)
    torch.cuda.synchronize()
    w['pc'].set_timing(True)
    w['pc'].timing()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        w['pc'].execute(bufs, sh)
        w['ps'].execute(bufs, sh)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    tm = w['pc'].timing()
    kern_ms = tm['ms'][0] / max(1, tm['launches'][0])
    flops = w['flops']
    achieved = tm['flops'][0] / (kern_ms * 1e-3) / 1e12
    out = {
        'metric': METRIC, 'value': round(flops * args.steps / wall / 1e12, 4), 'unit': 'TFLOP/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(wall * 1e3 / args.steps, 5), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic (random spin-flip-symmetric V and T in GeCCo ME-list layout)',
        'config': {'workload': 'ccsd_pp_ladder_contraction', 'nocc': args.nocc, 'nvir': args.nvir,
                   'nsym': args.nsym, 'use_tr': True, 'kernel_calls': w['nblocks'],
                   'gflop_per_step': round(flops / 1e9, 4), 'tasks': w['pc'].stats()['ntasks']},
        'roofline': {'bound': 'mfma', 'achieved': round(achieved, 4), 'peak': PEAK_FP64,
                     'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_FP64, 4),
                     'traffic': MEASURED_TRAFFIC.get(('ladder', args.nocc, args.nvir, args.nsym)),
                     'kernel': 'grouped_contract_affine<64, 64, 16, double, 1, 2, false>', 'kernel_ms': round(kern_ms, 5)},
    }
    if not args.no_cpu:
        v, dt, sample = cpu_ladder(args.nocc, args.nvir, args.nsym)
        out['cpu_baseline'] = {'value': round(v, 6), 'unit': 'TFLOP/s', 'cores': 1, 'kind': 'port',
                               'sample': sample, 'seconds': round(dt, 2),
                               'cpu': platform.processor() or platform.machine()}
    return out


# --------------------------------------------------------------------- ccsd
def kernel_times(plans, run, steps):
    """Kernel durations for the roofline: run() launches `steps` iterations
    one by one (the same kernels as the replayed graph) with HIP-event timing
    on the launch stream.  Launch classes: 0 strided-GEMM launch (with the
    direct short-contraction tasks), 1 gather kernel, 2 spin-flip
    symmetrisation, 3 operand materialisation, 4 unused.  The dominant
    contraction kernel by time is the strided-GEMM launch (its GEMM tiles and
    the direct tasks that share it: FLOP classes 0 and 2) or the gather
    kernel (FLOP class 1)."""
    import torch
    for p in plans:
        p.set_timing(True)
        p.timing()
    run()
    torch.cuda.synchronize()
    ms = [0.0] * 5
    nl = [0] * 5
    fl = [0.0] * 3
    gbytes = 0.0
    for p in plans:
        tm = p.timing()
        p.set_timing(False)
        gbytes += p.workspace()['gather_bytes']
        for i in range(5):
            ms[i] += tm['ms'][i]
            nl[i] += tm['launches'][i]
        for i in range(3):
            fl[i] += tm['flops'][i]
    cand = [(ms[0], 0, fl[0] + fl[2]), (ms[1], 1, fl[1])]
    _, dom, dom_fl = max(cand)
    dom_ms = ms[dom] / max(1, nl[dom])
    achieved = dom_fl / (ms[dom] / steps * 1e-3) / 1e12 if ms[dom] > 0 else 0.0
    return dict(ms=ms, launches=nl, names=['grouped_contract_affine<64, 64, 16, double, 1, 2, false>', 'grouped_contract<double>'],
                dom=dom, dom_ms=dom_ms, achieved=achieved, gather_bytes=gbytes)


def cpu_ccsd_iteration(args, iters=2):
    """One whole CCSD formula iteration of the oracle's route 3 on one host
    core (oracle/cpu_baseline.py in a child process: the bundled OpenBLAS
    dgemm limited to one thread before numpy loads), best of `iters`; args
    has nocc, nvir, naux, nsym, seed, sigma (attributes or keys)."""
    import subprocess
    get = (lambda k: args[k]) if isinstance(args, dict) else (lambda k: getattr(args, k))
    cmd = [sys.executable, '-m', 'oracle.cpu_baseline'] + \
        [str(get(k)) for k in ('nocc', 'nvir', 'naux', 'nsym', 'seed', 'sigma')] + [str(iters)]
    out = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        raise RuntimeError('cpu baseline failed: ' + out.stderr[-2000:])
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return r


def level_bytes(solver):
    """Algorithmic HBM bytes of the strided-GEMM launch, averaged over the
    formula levels: every list a level's contractions read, read once, and
    every list they write, written once (8 bytes per element)."""
    tot = 0.0
    for lv in solver.schedule:
        reads = {c['op1'] for c in lv['contr']} | {c['op2'] for c in lv['contr']}
        writes = {c['target'] for c in lv['contr']}
        tot += 8.0 * (sum(solver.L[n].length for n in reads) + sum(solver.L[n].length for n in writes))
    return tot / max(1, len(solver.schedule))


def run_ccsd(args):
    import torch
    import ccsd
    from synth import synthetic
    ham = synthetic(args.nocc, args.nvir, args.naux, nsym=args.nsym, seed=args.seed, sigma=args.sigma)
    t0 = time.perf_counter()
    solver = ccsd.CCSD(ham)
    setup_s = time.perf_counter() - t0
    torch.cuda.synchronize()
    step = lambda m: print(f'# {m}', file=sys.stderr, flush=True)
    step(f'setup {setup_s:.1f} s')
    tune_report, tune_s = None, 0.0
    if args.tune:
        t0 = time.perf_counter()
        solver.solve(tol=0.0, maxit=3, mode='async')        # amplitudes of a few iterations
        tune_report = solver.tune(log=lambda m: print('# tune ' + m, file=sys.stderr, flush=True))
        tune_s = time.perf_counter() - t0
        print(f'# tuned in {tune_s:.1f} s: {tune_report}', file=sys.stderr, flush=True)
    # warmup iterations from T = 0 (the first one captures the iteration as
    # a HIP graph), then the timed steps continue the solve with replays
    solver.solve(tol=0.0, maxit=args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed = solver.solve(tol=0.0, maxit=args.steps, restart=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    done = timed['iterations']         # all of them unless the solve stopped early
    step(f'timed steps {wall:.3f} s, {done} iterations')
    # kernel durations (roofline): as many further iterations launched one by
    # one (the same kernels as the replayed graph), timed with HIP events on
    # the launch stream
    kt = kernel_times(solver.plans, lambda: solver.solve(tol=0.0, maxit=args.steps, restart=False,
                                                        mode='async'), args.steps)
    ms, names, dom, dom_ms, achieved, gbytes = (kt[k] for k in ('ms', 'names', 'dom', 'dom_ms',
                                                                'achieved', 'gather_bytes'))
    # time to convergence from T = 0: each scheme runs twice (its first run
    # captures its graphs) and the second run is the one reported
    def twice(f):
        f()
        torch.cuda.synchronize()
        return f()
    res = twice(lambda: solver.solve(tol=args.tol, maxit=200))
    step(f'float64 solve {res["iterations"]} iterations {res["seconds"]:.4f} s')
    t0 = time.perf_counter()
    solver.build_f32(bulk='f32')
    f32_setup = time.perf_counter() - t0
    # float32 residuals until ||g|| < 1e-6, then float64 iterations
    mixed = twice(lambda: solver.solve_mixed(tol=args.tol, switch=1e-6, maxit=200, bulk='f32'))
    step(f'mixed solve {mixed["f32_iterations"]} + {mixed["f64_iterations"]} iterations')
    # float64 only for two residual evaluations: float32 phase, the step of
    # the linearised formula in float32, the final evaluation
    refined = twice(lambda: solver.solve_refined(tol=args.tol, switch=1e-6, bulk='f32'))
    step(f'refined solve {refined["f32_iterations"]} + {refined["linear_iterations"]} iterations, '
         f'{refined["f64_evaluations"]} float64 evaluations')
    solver.solve(tol=0.0, maxit=args.warmup, precision='f32')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.solve(tol=0.0, maxit=args.steps, precision='f32', restart=False)
    torch.cuda.synchronize()
    wall32 = time.perf_counter() - t0
    # the same with split-bfloat16 GEMM tiles in the low-precision phases
    refined_bf = twice(lambda: solver.solve_refined(tol=args.tol, switch=1e-6, bulk='bf16x3'))
    step(f'refined split-bf16 solve {refined_bf["f32_iterations"]} + {refined_bf["linear_iterations"]} '
         'iterations')
    solver.solve(tol=0.0, maxit=args.warmup, precision='f32')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.solve(tol=0.0, maxit=args.steps, precision='f32', restart=False)
    torch.cuda.synchronize()
    wall_bf = time.perf_counter() - t0
    ref = None
    if os.path.exists(GOLDEN):
        g = json.load(open(GOLDEN))
        rec = g['recipe']
        if (rec['nocc'], rec['nvir'], rec['naux'], rec['nsym'], rec['seed'], rec['sigma']) == \
                (args.nocc, args.nvir, args.naux, args.nsym, args.seed, args.sigma):
            ref = g['energy']
    out = {
        'metric': METRIC, 'value': round(done / wall, 3), 'unit': 'iters/s',
        'n_gpus': 1, 'steps': done, 'warmup': args.warmup,
        'ms_per_step': round(wall * 1e3 / done, 4), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': f'synthetic DF Hamiltonian (seed {args.seed}, naux {args.naux}, sigma {args.sigma})',
        'config': {'workload': 'ccsd_iteration', 'nocc': args.nocc, 'nvir': args.nvir,
                   'naux': args.naux, 'nsym': args.nsym, 'use_tr': True,
                   'kernel_calls_per_iter': solver.nblocks,
                   'gflop_per_iter': round(solver.flops / 1e9, 3)},
        'contraction_tflops': round(solver.flops * done / wall / 1e12, 3),
        'kernels_ms_per_iter': {'grouped_contract_affine<64, 64, 16, double, 1, 2, false>': round(ms[0] / args.steps, 4),
                                'grouped_contract<double>': round(ms[1] / args.steps, 4),
                                'sym_ab_batched<double>': round(ms[2] / args.steps, 4),
                                'gather_operands<double>': round(ms[3] / args.steps, 4)},
        'gather_operands_gbps': round(gbytes / (ms[3] / args.steps * 1e-3) / 1e9, 1) if ms[3] > 0 else None,
        'roofline': {'bound': 'mfma', 'achieved': round(achieved, 4), 'peak': PEAK_FP64,
                     'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_FP64, 4),
                     'traffic': MEASURED_TRAFFIC.get(('ccsd', args.nocc, args.nvir, args.naux, args.nsym,
                                                      names[dom])),
                     'traffic_algorithmic': round(level_bytes(solver), 1),
                     'kernel': names[dom], 'kernel_ms': round(dom_ms, 5)},
        'convergence': {'tol': args.tol, 'iterations': res['iterations'],
                        'seconds': round(res['seconds'], 4), 'energy': res['energy'],
                        'residual': res['residual'],
                        'oracle_energy': ref,
                        'abs_dE': None if ref is None else abs(res['energy'] - ref)},
        'time_to_convergence': {
            'tol': args.tol, 'oracle_energy': ref,
            'schemes': {
                'f64': {'seconds': round(res['seconds'], 5), 'iterations': res['iterations'],
                        'f64_evaluations': res['iterations'],
                        'abs_dE_vs_oracle': None if ref is None else abs(res['energy'] - ref)},
                'mixed_f32_then_f64': {'seconds': round(mixed['seconds'], 5),
                                       'f32_iterations': mixed['f32_iterations'],
                                       'f64_evaluations': mixed['f64_iterations'],
                                       'abs_dE_vs_oracle': None if ref is None else abs(mixed['energy'] - ref)},
                'refined_f32': {'seconds': round(refined['seconds'], 5),
                                'f32_iterations': refined['f32_iterations'],
                                'linear_iterations': refined['linear_iterations'],
                                'f64_evaluations': refined['f64_evaluations'],
                                'residual': refined['residual'],
                                'abs_dE_vs_oracle': None if ref is None else abs(refined['energy'] - ref)},
                'refined_split_bf16': {'seconds': round(refined_bf['seconds'], 5),
                                       'f32_iterations': refined_bf['f32_iterations'],
                                       'linear_iterations': refined_bf['linear_iterations'],
                                       'f64_evaluations': refined_bf['f64_evaluations'],
                                       'residual': refined_bf['residual'],
                                       'abs_dE_vs_oracle': None if ref is None else
                                       abs(refined_bf['energy'] - ref)}},
            'speedup_vs_plain_f64': {k: round(res['seconds'] / v['seconds'], 3) for k, v in
                                     (('mixed_f32_then_f64', mixed), ('refined_f32', refined),
                                      ('refined_split_bf16', refined_bf))},
            'f32_iters_per_s': round(args.steps / wall32, 3),
            'split_bf16_iters_per_s': round(args.steps / wall_bf, 3),
            'linearised_gflop_per_iter': round(solver.lin_flops / 1e9, 3),
            'f32_setup_seconds': round(f32_setup, 2)},
        'setup_seconds': round(setup_s, 2),
    }
    if tune_report is not None:
        out['tuning'] = {'seconds': round(tune_s, 1), 'levels': tune_report}
    if not args.no_cpu:
        step('cpu baseline (one core, oracle route 3)')
        r = cpu_ccsd_iteration(args)
        out['cpu_baseline'] = {
            'value': round(1.0 / r['seconds_per_iteration'], 6), 'unit': 'iters/s', 'cores': 1,
            'kind': 'port',
            'sample': f"one whole CCSD formula iteration ({r['kernel_calls']} kernel calls, "
                      f"{r['gflop_per_iteration']:.2f} GFLOP) of the oracle's route 3 (collect, dgemm, "
                      f"scatter per kernel call) with {r['blas']}, best of {r['iterations_timed']}",
            'seconds': round(sum(r['times']), 3), 'setup_seconds': round(r['setup_seconds'], 1),
            'gflops': round(r['gflop_per_iteration'] / r['seconds_per_iteration'], 3),
            'cpu': r['cpu'], 'nproc': r['nproc']}
    return out


# ------------------------------------------------------------------ triples
def cpu_triples_sample(ham, flops_eval):
    """Oracle route 3 (one core) on the T3 <- V(ov,oo).T2 term; returns the
    per-evaluation rate scaled by FLOP, the seconds and a description."""
    import numpy as np
    from oracle.melist import Orbitals, MEList
    from oracle.contract import Contraction
    from oracle import contract
    contract.blas_threads(1)
    dgemm = 'the bundled OpenBLAS ILP64 dgemm on 1 thread' if contract.blas() else 'a plain-loop dgemm'
    H, P = 1, 2
    orbs = Orbitals(ham.irrep_o, ham.irrep_v, ham.nsym)
    V = MEList(orbs, {P: 1, H: 1}, {H: 2})
    T2 = MEList(orbs, {P: 2}, {H: 2})
    R3 = MEList(orbs, {P: 3}, {H: 3})
    c = Contraction(V, T2, R3, {H: 1}, {})
    blocks = c.blocks(True)
    fl = sum(2.0 * np.prod(x['lstr_ex1']) * np.prod(x['lstr_ex2']) * np.prod(x['lstr_cnt'])
             for x in blocks)
    rng = np.random.default_rng(3)
    x1, x2 = rng.standard_normal(V.length), rng.standard_normal(T2.length)
    xr = np.zeros(R3.length)
    t0 = time.perf_counter()
    c.run(x1, x2, xr, route=3, use_tr=False, screen=False, blocks=blocks)
    dt = time.perf_counter() - t0
    per_eval = dt * flops_eval / fl
    return 1.0 / per_eval, dt, (f'the T3 <- V(ov,oo).T2 term ({fl / 1e9:.2f} of {flops_eval / 1e9:.1f} '
                                f'GFLOP per evaluation), oracle route 3 with {dgemm}, '
                                'scaled to one evaluation by FLOP')


def run_triples(args):
    import torch
    import ccsd
    import formula
    import gecco_hip
    from synth import synthetic
    ham = synthetic(args.nocc, args.nvir, args.naux, nsym=args.nsym, seed=args.seed, sigma=args.sigma)
    fpath = os.path.join(ROOT, 'gecco-public_amd', 'formulas', 'ccsdt_triples.json')
    t0 = time.perf_counter()
    solver = ccsd.CCSD(ham, formula=formula.load(fpath))
    setup_s = time.perf_counter() - t0
    print(f'# triples setup {setup_s:.1f} s, {solver.flops / 1e9:.1f} GFLOP, {solver.nblocks} kernel '
          'calls per evaluation', file=sys.stderr, flush=True)
    # random spin-symmetric amplitudes (projected by the library)
    g = torch.Generator(device='cuda:0')
    g.manual_seed(5)
    for n, scale in (('T2', 0.05), ('T3', 0.01)):
        solver.buf[n].copy_(scale * torch.randn(solver.buf[n].numel(), dtype=torch.float64,
                                                device='cuda:0', generator=g))
    proj = gecco_hip.Plan()
    for n in ('T2', 'T3'):
        proj.add_sym_ab(solver.L[n], solver.slot[n])
    proj.finalize()
    proj.execute(solver.bufs)
    tune_report, tune_s = None, 0.0
    if args.tune:
        t1 = time.perf_counter()
        tune_report = solver.tune(log=lambda m: print('# tune ' + m, file=sys.stderr, flush=True))
        tune_s = time.perf_counter() - t1
        print(f'# tuned in {tune_s:.1f} s: {tune_report}', file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        solver.evaluate()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.evaluate()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = kernel_times(solver.plans, lambda: [solver.evaluate() for _ in range(args.steps)], args.steps)
    # float32 lists (f32 MFMA): rate and error against the float64 R3
    solver.build_f32()
    for n in ('T2', 'T3'):
        solver.buf32[n].copy_(solver.buf[n])
    for _ in range(args.warmup):
        for p in solver.plans32:
            p.execute(solver.bufs32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for p in solver.plans32:
            p.execute(solver.bufs32)
    torch.cuda.synchronize()
    wall32 = time.perf_counter() - t0
    r64 = solver.buf['R3']
    r32 = solver.buf32['R3'].double()
    f32_err = {'max_abs_rel': float((r32 - r64).abs().max() / r64.abs().max()),
               'norm_rel': float((r32 - r64).norm() / r64.norm())}
    # error per kernel-call block of R3 is bounded by the list-wide maximum;
    # the worst single element relative to its own magnitude above 1e-3 max
    big = r64.abs() > 1e-3 * r64.abs().max()
    f32_err['worst_element_rel'] = float(((r32 - r64).abs()[big] / r64.abs()[big]).max())
    # split-bfloat16 GEMM tiles
    solver.build_f32(bulk='bf16x3')
    for n in ('T2', 'T3'):
        solver.buf32[n].copy_(solver.buf[n])
    for _ in range(args.warmup):
        for p in solver.plans32:
            p.execute(solver.bufs32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for p in solver.plans32:
            p.execute(solver.bufs32)
    torch.cuda.synchronize()
    wall_bf = time.perf_counter() - t0
    rbf = solver.buf32['R3'].double()
    bf_err = {'max_abs_rel': float((rbf - r64).abs().max() / r64.abs().max()),
              'norm_rel': float((rbf - r64).norm() / r64.norm())}
    # size-independent check: every term is linear in (T2, T3)
    r1 = solver.buf['R3'].clone()
    for n in ('T2', 'T3'):
        solver.buf[n].mul_(2.0)
    solver.evaluate()
    torch.cuda.synchronize()
    lin = float((solver.buf['R3'] - 2.0 * r1).abs().max() / (2.0 * r1.abs().max()))
    ms, names, dom = kt['ms'], kt['names'], kt['dom']
    out = {
        'metric': METRIC, 'value': round(solver.flops * args.steps / wall / 1e12, 4), 'unit': 'TFLOP/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(wall * 1e3 / args.steps, 4), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': f'synthetic DF Hamiltonian (seed {args.seed}, naux {args.naux}, sigma {args.sigma}), '
                'random spin-symmetric T2/T3',
        'config': {'workload': 'ccsdt_triples_contractions', 'nocc': args.nocc, 'nvir': args.nvir,
                   'naux': args.naux, 'nsym': args.nsym, 'use_tr': True,
                   'terms': ['pp ladder V(vv,vv).T3', 'hh ladder V(oo,oo).T3', 'ring T3.V(ov,vo)',
                             'V(vv,vo).T2', 'V(ov,oo).T2'],
                   'kernel_calls': solver.nblocks, 'gflop_per_step': round(solver.flops / 1e9, 3),
                   't3_elements': solver.L['T3'].length},
        'kernels_ms_per_step': {'grouped_contract_affine<64, 64, 16, double, 1, 2, false>': round(ms[0] / args.steps, 4),
                                'grouped_contract<double>': round(ms[1] / args.steps, 4),
                                'sym_ab_batched<double>': round(ms[2] / args.steps, 4),
                                'gather_operands<double>': round(ms[3] / args.steps, 4)},
        'roofline': {'bound': 'mfma', 'achieved': round(kt['achieved'], 4), 'peak': PEAK_FP64,
                     'unit': 'TFLOP/s', 'frac': round(kt['achieved'] / PEAK_FP64, 4),
                     'traffic': MEASURED_TRAFFIC.get(('triples', args.nocc, args.nvir, args.nsym,
                                                      names[dom])),
                     'kernel': names[dom], 'kernel_ms': round(kt['dom_ms'], 5)},
        'linearity_rel_err': lin,
        'f32': {'tflops': round(solver.flops * args.steps / wall32 / 1e12, 4),
                'ms_per_step': round(wall32 * 1e3 / args.steps, 4), 'error_vs_f64': f32_err},
        'split_bf16': {'tflops': round(solver.flops * args.steps / wall_bf / 1e12, 4),
                       'ms_per_step': round(wall_bf * 1e3 / args.steps, 4), 'error_vs_f64': bf_err},
        'setup_seconds': round(setup_s, 2),
    }
    if tune_report is not None:
        out['tuning'] = {'seconds': round(tune_s, 1), 'levels': tune_report}
    if not args.no_cpu:
        v, dt, sample = cpu_triples_sample(ham, solver.flops)
        out['cpu_baseline'] = {'value': round(v * solver.flops / 1e12, 6), 'unit': 'TFLOP/s', 'cores': 1,
                               'kind': 'port', 'sample': sample, 'seconds': round(dt, 2),
                               'cpu': platform.processor() or platform.machine()}
    return out


# ------------------------------------------------------------------ learned
def run_learned(args):
    import torch
    import ccsd
    import gecco_hip
    import learned
    from synth import synthetic
    if not os.path.exists(args.precond):
        raise SystemExit(f'trained preconditioner {args.precond} not found '
                         '(gecco-public_amd/train_precond.py writes it)')
    pre = learned.LearnedPreconditioner.load(args.precond)
    ham = synthetic(args.nocc, args.nvir, args.naux, nsym=args.nsym, seed=args.seed, sigma=args.sigma)
    t0 = time.perf_counter()
    solver = ccsd.CCSD(ham)
    setup_s = time.perf_counter() - t0
    print(f'# learned-preconditioner setup {setup_s:.1f} s, {solver.flops / 1e9:.1f} GFLOP per iteration, '
          f'{solver.amp.numel()} amplitudes', file=sys.stderr, flush=True)
    solver.solve(tol=0.0, maxit=2, mode='async')                  # warm-up
    plain = solver.solve(tol=args.tol, maxit=200, mode='async')
    print(f"# diagonal: {plain['iterations']} iterations {plain['seconds']:.3f} s", file=sys.stderr,
          flush=True)
    lrn = solver.solve(tol=args.tol, maxit=200, mode='async', precond=pre)
    print(f"# learned: {lrn['iterations']} iterations {lrn['seconds']:.3f} s, fallback at "
          f"{lrn['fallback_at']}", file=sys.stderr, flush=True)
    plain_g = solver.solve(tol=args.tol, maxit=200, mode='graph')
    # the converged energies at a tighter residual
    tight_p = solver.solve(tol=1e-10, maxit=300, mode='graph')
    tight_l = solver.solve(tol=1e-10, maxit=300, mode='async', precond=pre)
    # throughput of learned-preconditioner iterations (no convergence test)
    solver.solve(tol=0.0, maxit=args.warmup, mode='async', precond=pre)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.solve(tol=0.0, maxit=args.steps, mode='async', precond=pre)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the inference kernel alone, HIP events on the launch stream
    n = solver.amp.numel()
    scale = torch.empty_like(solver.amp)
    stream = torch.cuda.current_stream()
    for _ in range(3):
        pre.scale(solver.diag, solver.res, solver.amp, 5, scale, stream.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        pre.scale(solver.diag, solver.res, solver.amp, 5, scale, stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    inf_ms = e0.elapsed_time(e1) / reps
    mac = 128 * 4 + 128 * 128 + 128                  # per amplitude
    inf_tflops = 2.0 * mac * n / (inf_ms * 1e-3) / 1e12
    out = {
        'metric': METRIC, 'value': round(args.steps / wall, 3), 'unit': 'iters/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(wall * 1e3 / args.steps, 4), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': f'synthetic DF Hamiltonian (seed {args.seed}, naux {args.naux}, sigma {args.sigma})',
        'config': {'workload': 'ccsd_learned_preconditioner', 'nocc': args.nocc, 'nvir': args.nvir,
                   'naux': args.naux, 'nsym': args.nsym, 'use_tr': True, 'amplitudes': n,
                   'gflop_per_iter': round(solver.flops / 1e9, 3),
                   'network': '4-128-128-1 SiLU, bf16 weights, fp32 accumulation'},
        'convergence': {'tol': args.tol,
                        'diagonal': {'iterations': plain['iterations'], 'seconds': round(plain['seconds'], 4),
                                     'energy': plain['energy'], 'residual': plain['residual']},
                        'diagonal_graph': {'iterations': plain_g['iterations'],
                                           'seconds': round(plain_g['seconds'], 4)},
                        'learned': {'iterations': lrn['iterations'], 'seconds': round(lrn['seconds'], 4),
                                    'energy': lrn['energy'], 'residual': lrn['residual'],
                                    'fallback_at': lrn['fallback_at']},
                        'abs_dE_learned_vs_diagonal': abs(lrn['energy'] - plain['energy']),
                        'tol_1e-10': {'diagonal_iterations': tight_p['iterations'],
                                      'learned_iterations': tight_l['iterations'],
                                      'learned_fallback_at': tight_l['fallback_at'],
                                      'abs_dE': abs(tight_l['energy'] - tight_p['energy'])}},
        'roofline': {'bound': 'mfma', 'achieved': round(inf_tflops, 4), 'peak': 2500.0,
                     'unit': 'TFLOP/s', 'frac': round(inf_tflops / 2500.0, 6), 'traffic': None,
                     'kernel': 'mlp_infer_kernel<0> (+ feature norms)', 'kernel_ms': round(inf_ms, 5)},
        'setup_seconds': round(setup_s, 2),
    }
    summary = os.path.splitext(args.precond)[0] + '.json'
    if os.path.exists(summary):
        out['training'] = json.load(open(summary))
        out['training'].pop('hamiltonians', None)
    return out


# ------------------------------------------------------------------ batched
def run_batched(args):
    """--workload batched: args.instances independent CCSD instances of
    uneven size (synth.batch_recipe: nocc ~ U[8,24], nvir ~ U[40,160],
    naux = 3 nvir, 4 IRREPs, seeds 0..count-1) iterated together, one grouped
    launch per formula level for all instances and one replayed graph per
    batched iteration.  One step = one batched iteration (every instance
    advances by one).  The batch is then solved to convergence from T = 0,
    and the same instances are solved one after another (each with its own
    plans and graph) for the batching's gain."""
    import torch
    from batched import BatchedCCSD
    from synth import synthetic, batch_recipe
    recipe = batch_recipe(args.instances, nsym=args.nsym, sigma=args.sigma)
    t0 = time.perf_counter()
    hams = [synthetic(**r) for r in recipe]

    def progress(k, c):
        if k % 8 == 7 or k == len(hams) - 1:
            print(f'# instance {k + 1}/{len(hams)} set up ({time.perf_counter() - t0:.1f} s)',
                  file=sys.stderr, flush=True)
    b = BatchedCCSD(hams, on_instance=progress)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    print(f'# batched setup: {len(hams)} instances, {b.flops / 1e9:.1f} GFLOP and {b.nblocks} kernel '
          f'calls per batched iteration, {setup_s:.1f} s', file=sys.stderr, flush=True)
    b.solve(tol=0.0, maxit=args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.solve(tol=0.0, maxit=args.steps, restart=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kt = kernel_times(b.plans, lambda: b.solve(tol=0.0, maxit=args.steps, restart=False, mode='async'),
                      args.steps)
    res = b.solve(tol=args.tol, maxit=200, timing=True)
    print(f'# batched solve: {res["batch_iterations"]} batched iterations, {res["seconds"]:.3f} s',
          file=sys.stderr, flush=True)
    # mixed precision: float32 lists to 1e-5, then float64 refinement; and
    # the float32 batched iteration rate
    t0 = time.perf_counter()
    b.build_f32()
    print(f'# float32 lists and plans {time.perf_counter() - t0:.1f} s', file=sys.stderr, flush=True)
    mixed = b.solve_mixed(tol=args.tol, switch=1e-5, maxit=200)
    print(f'# mixed solve {mixed["f32_batch_iterations"]} + {mixed["f64_batch_iterations"]} batched '
          f'iterations, {mixed["seconds"]:.3f} s', file=sys.stderr, flush=True)
    b.solve(tol=0.0, maxit=args.warmup, precision='f32')
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.solve(tol=0.0, maxit=args.steps, precision='f32', restart=False)
    torch.cuda.synchronize()
    wall32 = time.perf_counter() - t0
    # the same instances one after another (per-instance plans and graphs)
    seq_s, seq_iters, seq_dE = 0.0, 0, 0.0
    seq_ms = []             # per-iteration milliseconds of each instance alone
    for k, c in enumerate(b.inst):
        if k % 16 == 15:
            print(f'# sequential solve {k + 1}/{len(b.inst)}', file=sys.stderr, flush=True)
        c.build_plans()
        r = c.solve(tol=args.tol, maxit=200)
        seq_s += r['seconds']
        seq_iters += r['iterations']
        seq_ms.append(1e3 * r['seconds'] / max(1, r['iterations']))
        seq_dE = max(seq_dE, abs(r['energy'] - res['energies'][k]))
        c.plans = []
        c._graph = None
    print(f'# sequential solves: {seq_iters} iterations, {seq_s:.3f} s', file=sys.stderr, flush=True)
    # batched iteration i against the same instances' iterations alone
    # (instance k is active in batched iterations 1..iterations[k])
    tail = []
    for i, ms_i in enumerate(res['iteration_ms']):
        act = [k for k, n in enumerate(res['iterations']) if n > i]
        tail.append((i + 1, len(act), round(ms_i, 3), round(sum(seq_ms[k] for k in act), 3)))
    for row in tail[::8] + ([tail[-1]] if (len(tail) - 1) % 8 else []):
        print('# batched iteration %d: %d active, %.3f ms (alone: %.3f ms)' % row, file=sys.stderr,
              flush=True)
    ms, names, dom = kt['ms'], kt['names'], kt['dom']
    sizes = [(r['nocc'], r['nvir']) for r in recipe]
    out = {
        'metric': METRIC, 'value': round(args.steps / wall, 3), 'unit': 'batched iters/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(wall * 1e3 / args.steps, 4), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': f'synthetic DF Hamiltonians (seeds 0..{args.instances - 1}, naux = 3 nvir, '
                f'sigma {args.sigma})',
        'config': {'workload': 'ccsd_batched_instances', 'instances': args.instances,
                   'nocc_range': [8, 24], 'nvir_range': [40, 160], 'nsym': args.nsym,
                   'use_tr': True, 'sizes': sizes, 'kernel_calls_per_iter': b.nblocks,
                   'gflop_per_iter': round(b.flops / 1e9, 3)},
        'instance_iters_per_s': round(args.instances * args.steps / wall, 2),
        'contraction_tflops': round(b.flops * args.steps / wall / 1e12, 3),
        'kernels_ms_per_iter': {'grouped_contract_affine<64, 64, 16, double, 1, 2, false>': round(ms[0] / args.steps, 4),
                                'grouped_contract<double>': round(ms[1] / args.steps, 4),
                                'sym_ab_batched<double>': round(ms[2] / args.steps, 4),
                                'gather_operands<double>': round(ms[3] / args.steps, 4)},
        'roofline': {'bound': 'mfma', 'achieved': round(kt['achieved'], 4), 'peak': PEAK_FP64,
                     'unit': 'TFLOP/s', 'frac': round(kt['achieved'] / PEAK_FP64, 4),
                     'traffic': MEASURED_TRAFFIC.get(('batched', args.instances, args.nsym, names[dom])),
                     'kernel': names[dom], 'kernel_ms': round(kt['dom_ms'], 5)},
        'convergence': {'tol': args.tol, 'batch_iterations': res['batch_iterations'],
                        'seconds': round(res['seconds'], 4), 'all_converged': all(res['converged']),
                        'iterations': res['iterations'],
                        'sequential_seconds': round(seq_s, 4), 'sequential_iterations': seq_iters,
                        'speedup_vs_sequential': round(seq_s / res['seconds'], 3),
                        'batched_gpu_ms': round(sum(res['iteration_ms']), 2),
                        'max_abs_dE_vs_sequential': seq_dE},
        'mixed_precision': {'switch': 1e-5, 'f32_batch_iterations': mixed['f32_batch_iterations'],
                            'f64_batch_iterations': mixed['f64_batch_iterations'],
                            'seconds': round(mixed['seconds'], 4), 'all_converged': all(mixed['converged']),
                            'max_abs_dE_vs_f64': max(abs(x - y) for x, y in
                                                     zip(mixed['energies'], res['energies'])),
                            'f32_batched_iters_per_s': round(args.steps / wall32, 3),
                            'f32_contraction_tflops': round(b.flops * args.steps / wall32 / 1e12, 3)},
        'setup_seconds': round(setup_s, 2),
    }
    if not args.no_cpu:
        # the instance of median FLOP count, one whole formula iteration of
        # the oracle's route 3 on one core; the batched rate at its GFLOP/s
        mid = sorted(range(len(hams)), key=lambda k: b.inst[k].flops)[len(hams) // 2]
        r = cpu_ccsd_iteration(recipe[mid], iters=1)
        gflops = r['gflop_per_iteration'] / r['seconds_per_iteration']
        out['cpu_baseline'] = {
            'value': round(gflops / (b.flops / 1e9), 6), 'unit': 'batched iters/s', 'cores': 1,
            'kind': 'port',
            'sample': f"instance {mid} (nocc {sizes[mid][0]}, nvir {sizes[mid][1]}): one whole CCSD "
                      f"formula iteration ({r['kernel_calls']} kernel calls, {r['gflop_per_iteration']:.2f} "
                      f"GFLOP) of the oracle's route 3 with {r['blas']}; the batched iteration's "
                      f"{b.flops / 1e9:.1f} GFLOP at that rate",
            'seconds': round(sum(r['times']), 3), 'gflops': round(gflops, 3),
            'cpu': r['cpu'], 'nproc': r['nproc']}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=30)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', choices=['ccsd', 'ladder', 'batched', 'triples', 'learned'],
                    default='ccsd')
    ap.add_argument('--precond', default=PRECOND, help='learned workload: trained parameters (.npz)')
    ap.add_argument('--instances', type=int, default=64, help='batched workload: instance count')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--tune', action='store_true',
                    help='ccsd / triples workloads: autotune the launch settings of every formula level first')
    ap.add_argument('--nocc', type=int, default=None, help=f'default {NOCC} (triples: {T3_NOCC})')
    ap.add_argument('--nvir', type=int, default=None, help=f'default {NVIR} (triples: {T3_NVIR})')
    ap.add_argument('--naux', type=int, default=None, help=f'default {NAUX} (triples: {T3_NAUX})')
    ap.add_argument('--nsym', type=int, default=NSYM)
    ap.add_argument('--seed', type=int, default=SEED)
    ap.add_argument('--sigma', type=float, default=SIGMA)
    ap.add_argument('--tol', type=float, default=1e-8)
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MB',
                    help='only run a plain device copy of MB megabytes (calibrates '
                         'FETCH_SIZE/WRITE_SIZE counters) and exit')
    args = ap.parse_args()
    dflt = {'triples': (T3_NOCC, T3_NVIR, T3_NAUX), 'learned': (L_NOCC, L_NVIR, L_NAUX)}.get(
        args.workload, (NOCC, NVIR, NAUX))
    args.nocc = args.nocc if args.nocc is not None else dflt[0]
    args.nvir = args.nvir if args.nvir is not None else dflt[1]
    args.naux = args.naux if args.naux is not None else dflt[2]
    if args.gpus != 1:
        print(json.dumps({'error': 'this engine runs on one GPU; --gpus > 1 is not supported'}))
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device='cuda:0')
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({'copy_bytes_read': 8 * n, 'copy_bytes_written': 8 * n, 'launches': 5}))
        return
    out = {'ladder': run_ladder, 'ccsd': run_ccsd, 'batched': run_batched,
           'triples': run_triples, 'learned': run_learned}[args.workload](args)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
