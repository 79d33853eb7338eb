"""The allocation-free entry point contr_blk1blk2_blocked_mm_noalloc_ against
the allocating one (contr_blk1blk2_blocked_mm_), through the C ABI, on two
calls in a row -- the second with a smaller input -- and its refusal of a
call whose index data exceed its fixed scratch.

Tolerance: the same kernels in the same order on the same inputs, so the
results agree bit for bit except where atomic result additions reorder
(1e-14 relative); both against the oracle to 1e-12."""
import numpy as np
import pytest

from models import orbitals, spin_symmetric, CASES

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _run(entry, blocks, d1, d2, nres, torch):
    dr = torch.zeros(nres, dtype=torch.float64, device='cuda:0')
    for bl in blocks:
        entry(bl, d1[bl['off1']:], d2[bl['off2']:], dr[bl['off12']:])
    torch.cuda.synchronize()
    return dr.cpu().numpy()


def _case(name, nocc, nvir, nsym, seed):
    c, a, b, r = CASES[name](orbitals(nocc, nvir, nsym))
    return c, spin_symmetric(a, seed), spin_symmetric(b, seed + 1), r.length


@pytest.mark.parametrize('first,second', [(('ladder', 8, 32, 1), ('ladder', 3, 5, 2)),
                                          (('ring', 6, 14, 2), ('t1t1', 3, 5, 1)),
                                          (('energy', 5, 12, 4), ('hh_ladder', 3, 4, 1))])
def test_noalloc_matches_allocating_entry(first, second):
    torch = _torch()
    import gecco_hip
    for k, (name, no, nv, ns) in enumerate((first, second)):
        print(f'noalloc call {k}: {name} nocc={no} nvir={nv} nsym={ns}', flush=True)
        c, x1, x2, nres = _case(name, no, nv, ns, 5 + 2 * k)
        blocks = c.blocks(True)
        ref = np.zeros(nres)
        c.run(x1, x2, ref, route=3, use_tr=False, screen=False, blocks=blocks)
        d1 = torch.from_numpy(x1).cuda()
        d2 = torch.from_numpy(x2).cuda()
        got = _run(gecco_hip.contr_blk1blk2_blocked_mm_noalloc, blocks, d1, d2, nres, torch)
        want = _run(gecco_hip.contr_blk1blk2_blocked_mm, blocks, d1, d2, nres, torch)
        scale = max(1e-300, np.abs(want).max())
        assert np.abs(got - want).max() <= 1e-14 * scale
        assert np.abs(got - ref).max() <= 1e-12 * max(1e-300, np.abs(ref).max())


def test_noalloc_refuses_more_than_its_scratch():
    """t1t1 at nocc 32, nvir 160 (C1): the result map of a block is a table
    of more than 2^24 int32 entries, beyond the 64 MiB array."""
    torch = _torch()
    import gecco_hip
    assert gecco_hip.noalloc_scratch_bytes() == 64 << 20
    c, a, b, r = CASES['t1t1'](orbitals(32, 160, 1))
    blocks = c.blocks(False)
    big = max(blocks, key=lambda bl: int(np.prod(bl['lstr_ex1'])) * int(np.prod(bl['lstr_ex2'])))
    assert int(np.prod(big['lstr_ex1'])) * int(np.prod(big['lstr_ex2'])) > (1 << 24)
    d1 = torch.zeros(a.length, dtype=torch.float64, device='cuda:0')
    d2 = torch.zeros(b.length, dtype=torch.float64, device='cuda:0')
    dr = torch.zeros(r.length, dtype=torch.float64, device='cuda:0')
    with pytest.raises(gecco_hip.GeccoHipError, match='scratch'):
        gecco_hip.contr_blk1blk2_blocked_mm_noalloc(big, d1[big['off1']:], d2[big['off2']:],
                                                   dr[big['off12']:])
    # host operands are refused (staging them would need device memory)
    small = CASES['ladder'](orbitals(2, 3, 1))[0].blocks(False)[0]
    with pytest.raises(gecco_hip.GeccoHipError, match='device memory'):
        gecco_hip.contr_blk1blk2_blocked_mm_noalloc(small, np.zeros(100), np.zeros(100), np.zeros(100))
    # the entry still works after a refusal
    c, x1, x2, nres = _case('ladder', 3, 5, 2, 9)
    blocks = c.blocks(True)
    d1, d2 = torch.from_numpy(x1).cuda(), torch.from_numpy(x2).cuda()
    got = _run(gecco_hip.contr_blk1blk2_blocked_mm_noalloc, blocks, d1, d2, nres, torch)
    want = _run(gecco_hip.contr_blk1blk2_blocked_mm, blocks, d1, d2, nres, torch)
    assert np.abs(got - want).max() <= 1e-14 * max(1e-300, np.abs(want).max())
