"""Benchmark of the MI355X block-contraction engine.

Workload: the particle-particle ladder of closed-shell CCSD,
    R(ab,ij) += sum_{c<d} V(ab,cd) T(cd,ij),
on the CCSD Hamiltonian of nocc=20, nvir=100, 4 IRREPs assigned
round-robin, GeCCo Ms/IRREP-blocked ME-lists, spin-flip symmetry on
(use_tr), FP64.  One step = one grouped launch of every kernel call of the
contraction (8 blocks, 5.58 GFLOP) followed by the spin-flip
symmetrisation of R, operands resident in HBM.

Prints one JSON line (see --help).  --gpus N > 1 is not supported: the
engine runs on one GPU here.
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

NOCC, NVIR, NSYM = 20, 100, 4
PEAK_FP64 = 78.6        # TFLOP/s, MI355X dense FP64 (vector and MFMA)
PEAK_HBM = 8000.0       # GB/s
# HBM bytes per launch of the ladder kernel at nocc=20 nvir=100 nsym=4, from
# rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 (copy-calibrated factors);
# see profiles/grouped_contract_affine_f64_ccsd_ladder_o20_v100_s4_pmc.md
MEASURED_TRAFFIC = {(20, 100, 4): 565.6e6}


def round_robin(n, nsym, start=0):
    return [((start + k) % nsym) + 1 for k in range(n)]


def build(nocc, nvir, nsym):
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
    # spin-flip symmetrisation plans for the synthetic inputs
    pv = gecco_hip.Plan()
    pv.add_sym_ab(V, 0)
    pv.finalize()
    pt = gecco_hip.Plan()
    pt.add_sym_ab(T, 1)
    pt.finalize()
    return dict(o=o, V=V, T=T, R=R, pc=pc, ps=ps, pv=pv, pt=pt, nblocks=nb, flops=fl)


def cpu_baseline(nocc, nvir, nsym):
    """Oracle route 3 (collect, dgemm as a plain loop, scatter; single core)
    on the same contraction; returns (TFLOP/s, seconds, sample description)."""
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
    sample = blocks  # one full step: every kernel call of the contraction
    fl = sum(2.0 * np.prod(b['lstr_ex1']) * np.prod(b['lstr_ex2']) * np.prod(b['lstr_cnt']) for b in sample)
    t0 = time.perf_counter()
    c.run(xv, xt, xr, route=3, use_tr=False, screen=False, blocks=sample)
    dt = time.perf_counter() - t0
    return fl / dt / 1e12, dt, (f'one full step: {len(sample)} of {len(blocks)} kernel calls '
                                f'({fl / 1e9:.2f} GFLOP), oracle route 3 with a plain-loop dgemm')


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--nocc', type=int, default=NOCC)
    ap.add_argument('--nvir', type=int, default=NVIR)
    ap.add_argument('--nsym', type=int, default=NSYM)
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MB',
                    help='only run a plain device copy of MB megabytes (calibrates '
                         'FETCH_SIZE/WRITE_SIZE counters) and exit')
    args = ap.parse_args()
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
    if args.gpus != 1:
        print(json.dumps({'error': 'this engine runs on one GPU; --gpus > 1 is not supported'}))
        sys.exit(2)
    import torch
    dev = torch.device('cuda:0')
    w = build(args.nocc, args.nvir, args.nsym)
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
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in ev:
        e0.record(stream)
        w['pc'].execute(bufs, sh)
        e1.record(stream)
        w['ps'].execute(bufs, sh)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = sum(e0.elapsed_time(e1) for e0, e1 in ev) / len(ev)
    ms_step = wall * 1e3 / args.steps
    flops = w['flops']
    st = w['pc'].stats()
    achieved = flops / (kern_ms * 1e-3) / 1e12
    value = flops * args.steps / wall / 1e12
    out = {
        'metric': 'CCSD iters/s & contraction TFLOP/s vs plain FP64 GPU at |dE|<=1e-8 Eh',
        'value': round(value, 4),
        'unit': 'TFLOP/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_step, 5),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (random spin-flip-symmetric V and T in GeCCo ME-list layout)',
        'config': {'workload': 'ccsd_pp_ladder_contraction', 'nocc': args.nocc,
                   'nvir': args.nvir, 'nsym': args.nsym, 'use_tr': True,
                   'kernel_calls': w['nblocks'], 'gflop_per_step': round(flops / 1e9, 4),
                   'tasks': st['ntasks']},
        'roofline': {'bound': 'mfma', 'achieved': round(achieved, 4), 'peak': PEAK_FP64,
                     'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_FP64, 4),
                     'traffic': MEASURED_TRAFFIC.get((args.nocc, args.nvir, args.nsym)),
                     'kernel': 'grouped_contract_affine_f64',
                     'kernel_ms': round(kern_ms, 5)},
    }
    if not args.no_cpu:
        v, dt, sample = cpu_baseline(args.nocc, args.nvir, args.nsym)
        out['cpu_baseline'] = {'value': round(v, 6), 'unit': 'TFLOP/s', 'cores': 1,
                               'kind': 'port', 'sample': sample, 'seconds': round(dt, 2),
                               'cpu': platform.processor() or platform.machine()}
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
