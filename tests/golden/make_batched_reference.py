"""Writes the dense textbook CCSD energies (oracle/ccsd_dense.py, converged
to 1e-11) of the smallest instances of the 64-instance batch recipe
(synth.batch_recipe(64): nocc ~ U[8,24], nvir ~ U[40,160], naux = 3 nvir,
4 IRREPs, seed = instance index) as a golden fixture.

Usage: python tests/golden/make_batched_reference.py OUT.json K [K ...]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, 'gecco-public_amd')]


def main():
    from synth import batch_recipe, synthetic
    from oracle.ccsd_dense import CCSD as Dense
    out = sys.argv[1]
    ks = [int(x) for x in sys.argv[2:]]
    recipe = batch_recipe(64)
    rows = []
    for k in ks:
        r = recipe[k]
        t0 = time.time()
        ham = synthetic(r['nocc'], r['nvir'], r['naux'], nsym=r['nsym'], seed=r['seed'], sigma=r['sigma'])
        e, *_ = Dense(ham).solve(1e-11)
        rows.append(dict(instance=k, **r, energy=float(e), seconds=round(time.time() - t0, 1)))
        print(rows[-1], flush=True)
    json.dump({'recipe': 'synth.batch_recipe(64)', 'method': 'oracle/ccsd_dense.py, residual 1e-11',
               'instances': rows}, open(out, 'w'), indent=1)


if __name__ == '__main__':
    main()
