"""CPU baseline: one whole CCSD formula iteration on one host core through
the oracle's route 3 (oracle/ccsd_route3.py: GeCCo's collect / dgemm('t','n')
/ scatter per kernel call, contract/contr_blk1blk2_blocked_mm.f, with the
bundled ILP64 OpenBLAS dgemm limited to one thread, as the reference links a
sequential BLAS).

TEST INFRASTRUCTURE ONLY (see oracle/strings.py header): run by bench.py's
cpu_baseline leg as a child process, so that the BLAS thread limit is set
before numpy loads.  Prints one JSON line.

Usage: python -m oracle.cpu_baseline NOCC NVIR NAUX NSYM SEED SIGMA [ITERS]
"""
import os

os.environ['OPENBLAS_NUM_THREADS'] = '1'
os.environ['OMP_NUM_THREADS'] = '1'

import json  # noqa: E402
import platform  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    root = os.path.dirname(here)
    sys.path[:0] = [root, os.path.join(root, 'gecco-public_amd')]
    nocc, nvir, naux, nsym, seed = (int(x) for x in sys.argv[1:6])
    sigma = float(sys.argv[6])
    iters = int(sys.argv[7]) if len(sys.argv) > 7 else 1
    import numpy as np
    import ccsd
    from synth import synthetic
    from oracle import contract
    from oracle.ccsd_route3 import Route3CCSD
    contract.blas_threads(1)
    t0 = time.perf_counter()
    ham = synthetic(nocc, nvir, naux, nsym=nsym, seed=seed, sigma=sigma)
    r = Route3CCSD(ham, ccsd.FORMULA, ccsd.SCHEDULE)
    setup = time.perf_counter() - t0
    rng = np.random.default_rng(1)
    r.X['T1'][:] = 0.01 * rng.standard_normal(r.X['T1'].size)
    r.X['T2'][:] = 0.01 * rng.standard_normal(r.X['T2'].size)
    times = []
    for _ in range(iters):
        t0 = time.perf_counter()
        r.evaluate()
        times.append(time.perf_counter() - t0)
    best = min(times)
    print(json.dumps({'seconds_per_iteration': best, 'iterations_timed': iters, 'times': times,
                      'gflop_per_iteration': r.flops / 1e9, 'kernel_calls': r.calls,
                      'setup_seconds': setup, 'cores': 1, 'cpu': cpu_model(), 'nproc': os.cpu_count(),
                      'blas': 'bundled OpenBLAS ILP64 dgemm, 1 thread' if contract.blas() else
                              'plain-loop dgemm'}))


if __name__ == '__main__':
    main()
