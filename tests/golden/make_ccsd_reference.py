"""Writes the FP64 oracle CCSD energy of a synthetic Hamiltonian as a golden
fixture: the package's CCSD formula table evaluated with the oracle's
restatement of contr_op1op2 / contr_blk1blk2_blocked_mm (oracle/ccsd_route3.py),
Jacobi steps with Pulay DIIS, converged to ||Omega|| < 1e-10.

Usage: python tests/golden/make_ccsd_reference.py NOCC NVIR NAUX NSYM SEED SIGMA OUT.json
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'gecco-public_amd'))

import ccsd  # noqa: E402
from synth import synthetic  # noqa: E402
from oracle.ccsd_route3 import Route3CCSD  # noqa: E402


def main():
    nocc, nvir, naux, nsym, seed = (int(x) for x in sys.argv[1:6])
    sigma = float(sys.argv[6])
    out_path = sys.argv[7]
    t0 = time.time()
    ham = synthetic(nocc, nvir, naux, nsym=nsym, seed=seed, sigma=sigma)
    r = Route3CCSD(ham, ccsd.FORMULA, ccsd.SCHEDULE)
    print(f'setup {time.time() - t0:.1f} s', flush=True)
    energy, it, rn = r.solve(tol=1e-10, log=lambda m: print(m, flush=True))
    json.dump({'recipe': {'nocc': nocc, 'nvir': nvir, 'naux': naux, 'nsym': nsym, 'seed': seed,
                          'sigma': sigma, 'occ_range': [-2.0, -0.3], 'vir_range': [0.2, 3.0]},
               'energy': energy, 'iterations': it, 'residual': rn,
               'method': 'oracle route 3, spin-flip symmetrisation after each level, Jacobi+DIIS(8)',
               'seconds': round(time.time() - t0, 1)},
              open(out_path, 'w'), indent=1)
    print('wrote', out_path, energy)


if __name__ == '__main__':
    main()
