"""The Fortran boundary on the GPU: tests/fortran/contr_driver.f, an F77-style
caller built with flang and 8-byte default INTEGER/LOGICAL, links
libgecco_hip.so and calls

  * contr_blk1blk2_blocked_mm / contr_blk1blk2_wmaps_c with the reference's
    argument lists (implicit interfaces, host arrays), every kernel call of a
    contraction (twice: the second reuses the library's kept plans), as
    contract/contr_op1op2_wmaps_c.f:1899,1927 does;
  * gecco_hip_contr_op1op2 through the ISO_C_BINDING module
    (gecco-public_amd/fortran/gecco_hip_f.f90) with contr_op1op2's
    arguments flattened.

Its results are compared with the oracle's route 3 restatement.  Tolerance:
FP64, 1e-12 relative; the printed values carry 17 significant digits."""
import os
import subprocess

import numpy as np
import pytest

from models import orbitals, spin_symmetric, CASES, TRA_CASES, round_robin

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, 'gecco-public_amd', 'fortran', '_build', 'contr_driver')
KEYS = ['hpvxop1c', 'hpvxop1a', 'hpvxop2c', 'hpvxop2a', 'hpvxop1op2c', 'hpvxop1op2a', 'lstrop1',
        'lstrop2', 'lstrop1op2', 'lstr_ex1', 'lstr_ex2', 'lstr_cnt', 'map_info_12c', 'map_info_12a',
        'map_info_1c', 'map_info_1a', 'map_info_2c', 'map_info_2a', 'map_ex1ex2c', 'map_ex1ex2a',
        'map_ex1cntc', 'map_ex1cnta', 'map_ex2cntc', 'map_ex2cnta']
DIMS = ['ncblk_op1', 'nablk_op1', 'ncblk_ex1', 'nablk_ex1', 'ncblk_op2', 'nablk_op2', 'ncblk_ex2',
        'nablk_ex2', 'ncblk_cnt', 'nablk_cnt', 'ncblk_op1op2', 'nablk_op1op2']


def _iocc(c, a):
    return [c.get(h, 0) for h in (1, 2, 3, 4)] + [a.get(h, 0) for h in (1, 2, 3, 4)]


def _fmt(v):
    return ' '.join(repr(float(x)) if isinstance(x, (float, np.floating)) else str(int(x)) for x in v)


@pytest.mark.parametrize('name,nsym,nocc,nvir', [('ladder', 2, 4, 6), ('ring', 4, 5, 8),
                                                 ('t1t1', 1, 3, 5), ('ring_t2', 2, 4, 6),
                                                 ('ladder_t12', 4, 4, 8)])
def test_fortran_caller_matches_oracle(name, nsym, nocc, nvir, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail('tests/fortran/contr_driver.f is not built (__graft_entry__.build())')
    orbs = orbitals(nocc, nvir, nsym)
    if name in CASES:
        c, a, b, r = CASES[name](orbs)
        stored, tra = (a, b, r), (0, 0, 0)
    else:
        c, stored, tra = TRA_CASES[name](orbs)
    xs = [spin_symmetric(m, 61 + k) for k, m in enumerate(stored)]
    blocks = c.blocks(True)
    ref = xs[2].copy()
    c.run(xs[0], xs[1], ref, route=3, use_tr=False, screen=False, blocks=blocks)
    full = xs[2].copy()
    c.run(xs[0], xs[1], full, route=3, use_tr=True, screen=False)
    lines = [_fmt([len(x) for x in xs]), _fmt(xs[0]), _fmt(xs[1]), _fmt(xs[2]), str(len(blocks))]
    for k, bl in enumerate(blocks):
        lines += [f"{3 if k % 2 == 0 else 2} {bl['xfac']!r} {bl['tra_op1']} {bl['tra_op2']} {bl['tra_op1op2']}",
                  _fmt([bl['off1'], bl['off2'], bl['off12']]), _fmt([bl[d] for d in DIMS]),
                  _fmt([len(bl[key]) for key in KEYS])]
        lines += [_fmt(bl[key]) for key in KEYS if len(bl[key])]
    used = (c.op1, c.op2, c.res)
    lines += [f'1.0 1 1 {int(tra[0])} {int(tra[1])} {int(tra[2])}',
              _fmt(_iocc(used[0].occ_c, used[0].occ_a) + _iocc(used[1].occ_c, used[1].occ_a) +
                   _iocc(used[2].occ_c, used[2].occ_a) + _iocc(c.cnt_c, c.cnt_a)),
              f'{nsym} {nocc}', _fmt(round_robin(nocc, nsym)), f'{nvir}', _fmt(round_robin(nvir, nsym, 1))]
    inp = tmp_path / 'in.txt'
    inp.write_text('\n'.join(lines) + '\n')
    print(f'fortran caller {name} nsym={nsym} nocc={nocc} nvir={nvir}: {len(blocks)} kernel calls', flush=True)
    out = subprocess.run([DRIVER], stdin=open(inp), capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    text = out.stdout.split()
    i1, i2 = text.index('kernel'), text.index('contr_op1op2')
    got = np.array([float(x) for x in text[i1 + 2:i2]])
    got2 = np.array([float(x) for x in text[i2 + 1:]])
    # every kernel call ran twice: the contraction added twice
    want = xs[2] + 2.0 * (ref - xs[2])
    assert len(got) == len(want) and len(got2) == len(full)
    scale = max(1e-300, np.abs(want).max())
    assert np.abs(got - want).max() <= 1e-12 * scale
    assert np.abs(got2 - full).max() <= 1e-12 * max(1e-300, np.abs(full).max())
