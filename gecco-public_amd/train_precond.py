"""Train the learned CCSD preconditioner on the GPU and write its parameters.

Training pairs come from plain (diagonal-preconditioner) DIIS solves of
synthetic closed-shell Hamiltonians (synth.synthetic: DF integrals,
naux = 3 nvir, 4 IRREPs) with nocc and nvir drawn uniformly from the given
ranges; each solve contributes `per_iter` random amplitudes per iteration
(learned.collect).  The network is trained with the library's gradient and
Adam kernels and written as a numpy .npz (one float32 array 'params') with
a JSON summary beside it.

    python gecco-public_amd/train_precond.py --out gecco-public_amd/precond/ccsd_mlp.npz
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', default=os.path.join(HERE, 'precond', 'ccsd_mlp.npz'))
    ap.add_argument('--hamiltonians', type=int, default=200)
    ap.add_argument('--nocc', type=int, nargs=2, default=[10, 24])
    ap.add_argument('--nvir', type=int, nargs=2, default=[50, 120])
    ap.add_argument('--per-iter', type=int, default=256)
    ap.add_argument('--steps', type=int, default=4000)
    ap.add_argument('--batch', type=int, default=65536)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--seed', type=int, default=100)
    args = ap.parse_args()
    import numpy as np
    import torch
    import ccsd
    import learned
    from synth import synthetic
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    parts, sizes, iters = [], [], []
    for k in range(args.hamiltonians):
        o = int(rng.integers(args.nocc[0], args.nocc[1] + 1))
        v = int(rng.integers(args.nvir[0], args.nvir[1] + 1))
        h = synthetic(o, v, 3 * v, nsym=4, seed=args.seed + k)
        solver = ccsd.CCSD(h)
        data, nit = learned.collect(solver, maxit=150, tol=1e-9, per_iter=args.per_iter, seed=k)
        parts.append(data)
        sizes.append((o, v))
        iters.append(nit)
        del solver
        torch.cuda.empty_cache()
        print(f'# hamiltonian {k + 1}/{args.hamiltonians}: nocc {o} nvir {v}, {nit} iterations, '
              f'{data[1].numel()} pairs ({time.perf_counter() - t0:.0f} s)', flush=True)
    data = tuple(torch.cat([d[i] for d in parts]).contiguous() for i in range(3))
    del parts
    pre = learned.LearnedPreconditioner(seed=args.seed)
    l0 = pre.loss(data)
    t1 = time.perf_counter()
    losses = []
    for chunk in range(0, args.steps, 500):
        n = min(500, args.steps - chunk)
        losses += pre.train(data, steps=n, lr=args.lr, batch=args.batch, seed=args.seed + chunk)
        print(f'# step {chunk + n}: minibatch loss {losses[-1]:.4e}', flush=True)
    train_s = time.perf_counter() - t1
    l1 = pre.loss(data)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    pre.save(args.out)
    summary = {'pairs': int(data[1].numel()), 'hamiltonians': sizes, 'plain_iterations': iters,
               'nsym': 4, 'naux_per_vir': 3, 'seed': args.seed, 'steps': args.steps,
               'batch': args.batch, 'lr': args.lr, 'loss_initial': l0, 'loss_final': l1,
               'train_seconds': round(train_s, 2), 'collect_seconds': round(t1 - t0, 2)}
    with open(os.path.splitext(args.out)[0] + '.json', 'w') as f:
        json.dump(summary, f, indent=1)
    print(json.dumps(summary))


if __name__ == '__main__':
    main()
