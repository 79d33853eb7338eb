"""GPU parity of the fused gather/MFMA/scatter kernel against the oracle's
route-3 restatement (contract/contr_blk1blk2_blocked_mm.f).

Tolerance: FP64 with different summation order (MFMA K-steps, split-K
atomics): max |gpu - oracle| <= 1e-12 * max |oracle|.
"""
import numpy as np
import pytest

from models import orbitals, spin_symmetric, CASES, T3_CASES

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _oracle(c, x1, x2, nres, blocks):
    ref = np.zeros(nres)
    c.run(x1, x2, ref, route=3, use_tr=False, screen=False, blocks=blocks)
    return ref


def _relerr(got, ref):
    return float(np.abs(got - ref).max() / max(1e-300, np.abs(ref).max()))


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (2, 4, 6), (4, 5, 9), (1, 8, 32)])
@pytest.mark.parametrize('name', list(CASES))
@pytest.mark.parametrize('materialize', [0, 1, 32])
def test_plan_matches_oracle(nsym, nocc, nvir, name, materialize):
    """materialize = min(M, N) from which gathered operands are written into
    scratch and the call runs on the strided-GEMM kernel (0: never)."""
    torch = _torch()
    import gecco_hip
    print(f'case {name} nsym={nsym} nocc={nocc} nvir={nvir} materialize={materialize}', flush=True)
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = CASES[name](orbs)
    x1 = spin_symmetric(a, 21)
    x2 = spin_symmetric(b, 22)
    for use_tr in (True, False):
        blocks = c.blocks(use_tr)
        ref = _oracle(c, x1, x2, r.length, blocks)
        dev = torch.device('cuda:0')
        d1 = torch.from_numpy(x1).to(dev)
        d2 = torch.from_numpy(x2).to(dev)
        dr = torch.zeros(r.length, dtype=torch.float64, device=dev)
        plan = gecco_hip.Plan()
        plan.set_materialize(materialize)
        for bl in blocks:
            plan.add(bl)
        plan.finalize()
        plan.execute([d1, d2, dr])
        torch.cuda.synchronize()
        assert _relerr(dr.cpu().numpy(), ref) <= TOL
        # results accumulate: a second execute doubles them
        plan.execute([d1, d2, dr])
        torch.cuda.synchronize()
        assert _relerr(dr.cpu().numpy(), 2 * ref) <= TOL


@pytest.mark.parametrize('name', ['ladder', 'ring', 't1t1'])
def test_fortran_entry_points_host_and_device(name):
    """contr_blk1blk2_blocked_mm_ / contr_blk1blk2_wmaps_c_ called block by
    block exactly as contract/contr_op1op2_wmaps_c.f:1899,1927 does."""
    torch = _torch()
    import gecco_hip
    orbs = orbitals(4, 6, 2)
    c, a, b, r = CASES[name](orbs)
    x1 = spin_symmetric(a, 31)
    x2 = spin_symmetric(b, 32)
    blocks = c.blocks(True)
    ref = _oracle(c, x1, x2, r.length, blocks)
    got = np.zeros(r.length)
    dev = torch.device('cuda:0')
    d1 = torch.from_numpy(x1).to(dev)
    d2 = torch.from_numpy(x2).to(dev)
    dr = torch.zeros(r.length, dtype=torch.float64, device=dev)
    for i, bl in enumerate(blocks):
        f = gecco_hip.contr_blk1blk2_blocked_mm if i % 2 == 0 else gecco_hip.contr_blk1blk2_wmaps_c
        f(bl, x1[bl['off1']:], x2[bl['off2']:], got[bl['off12']:])
        f(bl, d1[bl['off1']:], d2[bl['off2']:], dr[bl['off12']:])
    torch.cuda.synchronize()
    assert _relerr(got, ref) <= TOL
    assert _relerr(dr.cpu().numpy(), ref) <= TOL


def test_many_contractions_one_launch():
    """Several contractions with different result lists in one grouped launch."""
    torch = _torch()
    import gecco_hip
    orbs = orbitals(5, 8, 4)
    dev = torch.device('cuda:0')
    plan = gecco_hip.Plan()
    bufs, refs = [], []
    for k, name in enumerate(['ladder', 'ring', 'hh_ladder', 'fock_pp', 't1t1', 'energy']):
        c, a, b, r = CASES[name](orbs)
        x1 = spin_symmetric(a, 40 + k)
        x2 = spin_symmetric(b, 50 + k)
        blocks = c.blocks(True)
        refs.append(_oracle(c, x1, x2, r.length, blocks))
        base = len(bufs)
        bufs += [torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev),
                 torch.zeros(r.length, dtype=torch.float64, device=dev)]
        for bl in blocks:
            plan.add(bl, buf=(base, base + 1, base + 2))
    plan.finalize()
    plan.execute(bufs)
    torch.cuda.synchronize()
    for k, ref in enumerate(refs):
        assert _relerr(bufs[3 * k + 2].cpu().numpy(), ref) <= TOL


def test_instance_groups_skip_flags():
    """Contractions recorded as instance groups of one plan: with skip flags
    set, the flagged group's list initialisation, gathers and kernel calls do
    nothing (its result keeps its prior contents exactly) and the others
    match the oracle; cleared flags or no flags run everything."""
    torch = _torch()
    import gecco_hip
    orbs = orbitals(5, 8, 4)
    dev = torch.device('cuda:0')
    plan = gecco_hip.Plan()
    bufs, refs, prior = [], [], []
    rng = np.random.default_rng(7)
    for k, name in enumerate(['ladder', 'ring', 't1t1']):
        plan.set_group(k)
        c, a, b, r = CASES[name](orbs)
        x1 = spin_symmetric(a, 60 + k)
        x2 = spin_symmetric(b, 70 + k)
        blocks = c.blocks(True)
        refs.append(_oracle(c, x1, x2, r.length, blocks))
        prior.append(rng.standard_normal(r.length))
        base = len(bufs)
        bufs += [torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev),
                 torch.from_numpy(prior[-1]).to(dev)]
        plan.add_init(base + 2, r.length)           # result = 0, then the contraction
        for bl in blocks:
            plan.add(bl, buf=(base, base + 1, base + 2))
    plan.finalize()
    flags = torch.tensor([0, 1, 0], dtype=torch.int32, device=dev)
    with pytest.raises(gecco_hip.GeccoHipError):
        plan.set_skip(flags[:2])                     # groups 0..2 need 3 flags
    plan.set_skip(flags)
    plan.execute(bufs)
    torch.cuda.synchronize()
    for k in (0, 2):
        assert _relerr(bufs[3 * k + 2].cpu().numpy(), refs[k]) <= TOL
    assert np.array_equal(bufs[5].cpu().numpy(), prior[1])
    flags.zero_()
    plan.execute(bufs)
    torch.cuda.synchronize()
    for k in range(3):
        assert _relerr(bufs[3 * k + 2].cpu().numpy(), refs[k]) <= TOL
    flags.fill_(1)
    bufs[2].fill_(3.0)
    plan.set_skip(None)                              # no flags: everything runs
    plan.execute(bufs)
    torch.cuda.synchronize()
    assert _relerr(bufs[2].cpu().numpy(), refs[0]) <= TOL


def test_dense_ladder_blocks_match_oracle():
    """The pp ladder at C1 symmetry, nocc 16, nvir 48: dense blocks of
    2304 x 2304 x 256 (the shape class of the dense-block measurement, at a
    size the oracle's route 3 finishes in seconds), many tiles and long
    contraction ranges per task, against the oracle."""
    torch = _torch()
    import gecco_hip
    print('dense ladder nsym=1 nocc=16 nvir=48', flush=True)
    orbs = orbitals(16, 48, 1)
    c, a, b, r = CASES['ladder'](orbs)
    x1 = spin_symmetric(a, 31)
    x2 = spin_symmetric(b, 32)
    blocks = c.blocks(True)
    ref = _oracle(c, x1, x2, r.length, blocks)
    dev = torch.device('cuda:0')
    bufs = [torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev),
            torch.zeros(r.length, dtype=torch.float64, device=dev)]
    plan = gecco_hip.Plan()
    for bl in blocks:
        plan.add(bl)
    plan.finalize()
    plan.execute(bufs)
    torch.cuda.synchronize()
    assert np.abs(ref).max() > 1e-3
    assert _relerr(bufs[2].cpu().numpy(), ref) <= TOL


def test_live_task_lists_match_full_lists():
    """Plan.compact: live-task lists of the groups whose skip flag is 0 give
    the results of the full lists (the oracle's for live groups; flagged
    groups keep their contents exactly), launch fewer tasks, still read flags
    set after the compaction, work inside a replayed graph, and give way to
    the full lists again before flags are cleared."""
    torch = _torch()
    import gecco_hip
    orbs = orbitals(5, 8, 4)
    dev = torch.device('cuda:0')
    plan = gecco_hip.Plan()
    names = ['ladder', 'ring', 't1t1', 'ladder', 'ring', 't1t1']
    bufs, refs, prior = [], [], []
    rng = np.random.default_rng(11)
    for k, name in enumerate(names):
        plan.set_group(k)
        c, a, b, r = CASES[name](orbs)
        x1 = spin_symmetric(a, 80 + k)
        x2 = spin_symmetric(b, 90 + k)
        blocks = c.blocks(True)
        refs.append(_oracle(c, x1, x2, r.length, blocks))
        prior.append(rng.standard_normal(r.length))
        base = len(bufs)
        bufs += [torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev),
                 torch.from_numpy(prior[-1]).to(dev)]
        plan.add_init(base + 2, r.length)
        for bl in blocks:
            plan.add(bl, buf=(base, base + 1, base + 2))
    plan.finalize()
    with pytest.raises(gecco_hip.GeccoHipError):
        plan.compact(True)                            # no skip flags yet
    flags = torch.tensor([0, 1, 0, 1, 1, 0], dtype=torch.int32, device=dev)
    plan.set_skip(flags)
    full = plan.live_tasks()
    print('live tasks: full', full, flush=True)
    plan.compact(True)
    live = plan.live_tasks()
    print('live tasks: compacted', live, flush=True)
    assert all(x <= y for x, y in zip(live, full)) and sum(live) < sum(full)

    def check(skipped):
        torch.cuda.synchronize()
        for k in range(len(names)):
            got = bufs[3 * k + 2].cpu().numpy()
            if k in skipped:
                assert np.array_equal(got, prior[k]), k
            else:
                assert _relerr(got, refs[k]) <= TOL, k
    plan.execute(bufs)
    check({1, 3, 4})
    # a flag set after the compaction is still read
    flags[0] = 1
    bufs[2].copy_(torch.from_numpy(prior[0]))
    plan.execute(bufs)
    check({0, 1, 3, 4})
    # a replayed graph of the compacted plan
    plan.compact(True)
    assert plan.live_tasks()[0] < live[0]
    for k in (2, 5):
        bufs[3 * k + 2].fill_(7.0)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.execute(bufs, s.cuda_stream)             # buffer table upload outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.execute(bufs, torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        for k in (2, 5):
            bufs[3 * k + 2].fill_(float(rep))
        torch.cuda.synchronize()
        g.replay()
        check({0, 1, 3, 4})
    # the full lists again, then every flag cleared
    plan.compact(False)
    assert plan.live_tasks() == full
    flags.zero_()
    plan.execute(bufs)
    check(set())
    # compaction with no flag set keeps every task
    plan.compact(True)
    assert plan.live_tasks() == full
    plan.execute(bufs)
    check(set())


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (2, 4, 6), (4, 6, 10)])
@pytest.mark.parametrize('name', list(CASES))
@pytest.mark.parametrize('use_tr', [True, False])
def test_product_enumerator_matches_oracle(nsym, nocc, nvir, name, use_tr):
    """Product plan builder (strings, maps, layout, block enumerator, spin-flip
    symmetrisation on the GPU) against the oracle's contr_op1op2 restatement."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    print(f'case {name} nsym={nsym} nocc={nocc} nvir={nvir} use_tr={use_tr}', flush=True)
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = CASES[name](orbs)
    x1 = spin_symmetric(a, 61)
    x2 = spin_symmetric(b, 62)
    x0 = spin_symmetric(r, 63)          # result accumulates onto existing data
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=use_tr, screen=False)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    A = gecco_hip.MEList(o, a.occ_c, a.occ_a)
    B = gecco_hip.MEList(o, b.occ_c, b.occ_a)
    R = gecco_hip.MEList(o, r.occ_c, r.occ_a)
    plan = gecco_hip.Plan()
    plan.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, use_tr)
    if use_tr:
        plan.add_sym_ab(R, 2)
    plan.finalize()
    dev = torch.device('cuda:0')
    d1 = torch.from_numpy(x1).to(dev)
    d2 = torch.from_numpy(x2).to(dev)
    dr = torch.from_numpy(x0).to(dev)
    plan.execute([d1, d2, dr])
    torch.cuda.synchronize()
    assert _relerr(dr.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (2, 3, 4), (2, 4, 6), (4, 6, 10)])
@pytest.mark.parametrize('name', list(T3_CASES))
@pytest.mark.parametrize('use_tr', [True, False])
def test_triples_match_oracle(nsym, nocc, nvir, name, use_tr):
    """CCSDT triples terms (three-body ME-lists) through the product plan
    builder against the oracle's contr_op1op2 restatement."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    print(f'triples {name} nsym={nsym} nocc={nocc} nvir={nvir} use_tr={use_tr}', flush=True)
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = T3_CASES[name](orbs)
    x1 = spin_symmetric(a, 71)
    x2 = spin_symmetric(b, 72)
    x0 = spin_symmetric(r, 73)
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=use_tr, screen=False)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    A = gecco_hip.MEList(o, a.occ_c, a.occ_a)
    B = gecco_hip.MEList(o, b.occ_c, b.occ_a)
    R = gecco_hip.MEList(o, r.occ_c, r.occ_a)
    plan = gecco_hip.Plan()
    plan.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, use_tr)
    if use_tr:
        plan.add_sym_ab(R, 2)
    plan.finalize()
    dev = torch.device('cuda:0')
    dr = torch.from_numpy(x0).to(dev)
    plan.execute([torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev), dr])
    torch.cuda.synchronize()
    assert _relerr(dr.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize('nsym,nocc,nvir', [(2, 4, 6), (4, 6, 10)])
@pytest.mark.parametrize('name', list(CASES) + list(T3_CASES))
@pytest.mark.parametrize('precision', ['f32', 'bf16x3'])
def test_f32_plan_matches_oracle(nsym, nocc, nvir, name, precision):
    """A float32 plan (float32 lists, f32 MFMA; bf16x3: GEMM tiles as
    split-bfloat16 three-term products) against the oracle's float64
    contraction of the same (float32-representable) inputs.  Tolerance:
    f32: float32 rounding of the products and sums, max |err| <= 2e-5 max
    |ref|; bf16x3: each element is hi + lo to 2^-16 relative, max |err| <=
    1e-4 max |ref|."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    print(f'{precision} {name} nsym={nsym} nocc={nocc} nvir={nvir}', flush=True)
    cases = {**CASES, **T3_CASES}
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = cases[name](orbs)
    x1 = spin_symmetric(a, 81).astype(np.float32).astype(np.float64)
    x2 = spin_symmetric(b, 82).astype(np.float32).astype(np.float64)
    x0 = spin_symmetric(r, 83).astype(np.float32).astype(np.float64)
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=True, screen=False)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    A, B, R = (gecco_hip.MEList(o, m.occ_c, m.occ_a) for m in (a, b, r))
    plan = gecco_hip.Plan()
    plan.set_precision(precision)
    plan.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, True)
    plan.add_sym_ab(R, 2)
    plan.finalize()
    dev = torch.device('cuda:0')
    f = lambda v: torch.from_numpy(v.astype(np.float32)).to(dev)
    dr = f(x0)
    with pytest.raises(gecco_hip.GeccoHipError):
        plan.execute([torch.from_numpy(x1).to(dev), f(x2), dr])      # float64 buffer
    plan.execute([f(x1), f(x2), dr])
    torch.cuda.synchronize()
    err = _relerr(dr.cpu().numpy().astype(np.float64), ref)
    print(f'rel err {err:.2e}')
    assert err <= (2e-5 if precision == 'f32' else 1e-4)


def test_split_bf16_tiles_run_and_meet_their_bound():
    """GEMM tiles in split-bfloat16 (x = hi + lo, three BF16 MFMA products):
    on the nocc=8, nvir=32 ladder (GEMM tiles, not thin or direct tasks) the
    error against the float64 oracle is at the split's 2^-16 level -- above
    the float32 plan's, so the split tiles did run -- and below 1e-4."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    nsym, nocc, nvir = 1, 8, 32
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = CASES['ladder'](orbs)
    x1 = spin_symmetric(a, 91).astype(np.float32).astype(np.float64)
    x2 = spin_symmetric(b, 92).astype(np.float32).astype(np.float64)
    ref = np.zeros(r.length)
    c.run(x1, x2, ref, route=3, use_tr=True, screen=False)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    A, B, R = (gecco_hip.MEList(o, m.occ_c, m.occ_a) for m in (a, b, r))
    dev = torch.device('cuda:0')
    f = lambda v: torch.from_numpy(v.astype(np.float32)).to(dev)
    errs = {}
    for precision in ('f32', 'bf16x3'):
        plan = gecco_hip.Plan()
        plan.set_precision(precision)
        nb, _ = plan.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, True)
        plan.add_sym_ab(R, 2)
        plan.finalize()
        tiles = [i for i in range(nb) if plan.block_info(i)[16] and not plan.block_info(i)[19]
                 and not plan.block_info(i)[24]]
        assert tiles, 'no GEMM-tile kernel call in this case'
        dr = torch.zeros(r.length, dtype=torch.float32, device=dev)
        plan.execute([f(x1), f(x2), dr])
        torch.cuda.synchronize()
        errs[precision] = _relerr(dr.cpu().numpy().astype(np.float64), ref)
    print('errors', errs)
    assert errs['f32'] <= 2e-5
    assert errs['f32'] < errs['bf16x3'] <= 1e-4
    with pytest.raises(gecco_hip.GeccoHipError):     # split tiles need float32 lists
        p = gecco_hip.Plan()
        p.set_tuning(variant=8)
        p.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, True)
        p.finalize()


def test_all_triples_terms_one_launch():
    """The five triples terms into one R3 list in one grouped launch (the
    triples bench's plan) against the sum of the oracle's terms."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    nsym, nocc, nvir = 4, 5, 9
    orbs = orbitals(nocc, nvir, nsym)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    dev = torch.device('cuda:0')
    plan = gecco_hip.Plan()
    bufs = [None]
    ref = None
    for k, name in enumerate(T3_CASES):
        c, a, b, r = T3_CASES[name](orbs)
        if ref is None:
            ref = np.zeros(r.length)
            R = gecco_hip.MEList(o, r.occ_c, r.occ_a)
            bufs[0] = torch.zeros(r.length, dtype=torch.float64, device=dev)
        x1 = spin_symmetric(a, 80 + k)
        x2 = spin_symmetric(b, 90 + k)
        c.run(x1, x2, ref, route=3, use_tr=True, screen=False)
        A = gecco_hip.MEList(o, a.occ_c, a.occ_a)
        B = gecco_hip.MEList(o, b.occ_c, b.occ_a)
        s = len(bufs)
        bufs += [torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev)]
        plan.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, True, buf=(s, s + 1, 0))
    plan.add_sym_ab(R, 0)
    plan.finalize()
    plan.execute(bufs)
    torch.cuda.synchronize()
    assert _relerr(bufs[0].cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize('variant', [0, 1, 2, 3, 4, 5, 6, 7, 9])
@pytest.mark.parametrize('name', ['ladder', 'hh_ladder', 'energy', 'ring'])
def test_affine_tile_variants(variant, name, monkeypatch):
    """Every tile variant of the strided-GEMM kernel gives the oracle result."""
    torch = _torch()
    import gecco_hip
    monkeypatch.setenv('GECCO_HIP_AFFINE_TILE', str(variant))
    monkeypatch.setenv('GECCO_HIP_MIN_KCHUNK', '32')
    print(f'variant {variant} case {name}', flush=True)
    orbs = orbitals(6, 21, 2)
    c, a, b, r = CASES[name](orbs)
    x1 = spin_symmetric(a, 71)
    x2 = spin_symmetric(b, 72)
    blocks = c.blocks(True)
    ref = _oracle(c, x1, x2, r.length, blocks)
    dev = torch.device('cuda:0')
    bufs = [torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev),
            torch.zeros(r.length, dtype=torch.float64, device=dev)]
    plan = gecco_hip.Plan()
    for bl in blocks:
        plan.add(bl)
    plan.finalize()
    plan.execute(bufs)
    torch.cuda.synchronize()
    assert _relerr(bufs[2].cpu().numpy(), ref) <= TOL


def test_materialised_gathers_share_jobs():
    """The ring contraction (all three index maps gathered) recorded twice
    into one plan: the second copy reuses the first copy's materialised
    operands, and both accumulate into the same result."""
    torch = _torch()
    import gecco_hip
    orbs = orbitals(5, 12, 2)
    c, a, b, r = CASES['ring'](orbs)
    x1 = spin_symmetric(a, 81)
    x2 = spin_symmetric(b, 82)
    blocks = c.blocks(True)
    ref = _oracle(c, x1, x2, r.length, blocks)
    dev = torch.device('cuda:0')
    bufs = [torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev),
            torch.zeros(r.length, dtype=torch.float64, device=dev)]
    one = gecco_hip.Plan()
    one.set_materialize(1)
    two = gecco_hip.Plan()
    two.set_materialize(1)
    for bl in blocks:
        one.add(bl)
        two.add(bl)
        two.add(bl)
    one.finalize()
    two.finalize()
    infos = [one.block_info(i) for i in range(len(blocks))]
    assert any(x[17] or x[18] for x in infos), 'ring gathers are expected to be materialised'
    assert all(x[16] for x in infos)
    w1, w2 = one.workspace(), two.workspace()
    assert w1['gather_jobs'] > 0
    assert w2['gather_jobs'] == w1['gather_jobs']
    assert w2['scratch_bytes'] == w1['scratch_bytes']
    two.execute(bufs)
    torch.cuda.synchronize()
    assert _relerr(bufs[2].cpu().numpy(), 2 * ref) <= TOL


@pytest.mark.parametrize('materialize,variant,chain', [(0, 1, 1), (1, 1, 1), (1, 0, 1), (1, 4, 1),
                                                    (1, 1, 0), (1, 9, 1)])
def test_ccsd_terms_match_oracle(materialize, variant, chain, monkeypatch):
    """Every binary contraction of the CCSD formula table, as kernel calls of
    the product enumerator, against the oracle's contr_op1op2 restatement.
    Covers table-indexed and non-injective results (antisymmetrised outer
    products) on the LDS-staged write-back of the strided-GEMM kernel, and
    chained kernel calls (one GEMM over the contraction ranges of the calls
    that add into one result block)."""
    torch = _torch()
    import gecco_hip
    import ccsd
    from oracle.melist import MEList as OMEList
    from models import round_robin
    monkeypatch.setenv('GECCO_HIP_AFFINE_TILE', str(variant))
    nocc, nvir, nsym = 4, 9, 2
    orbs = orbitals(nocc, nvir, nsym)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    dev = torch.device('cuda:0')
    staged = linked = skinny = direct = 0
    for li, level in enumerate(ccsd.LEVELS):
        for ti, (res, a, b, cc, ca, fact) in enumerate(level):
            print(f'term L{li}.{ti} {res}<-{a}x{b} materialize={materialize} variant={variant} '
                  f'chain={chain}', flush=True)
            La, Lb, Lr = (OMEList(orbs, *ccsd.LISTS[n]) for n in (a, b, res))
            from oracle.contract import Contraction
            c = Contraction(La, Lb, Lr, cc, ca, fact)
            x1 = spin_symmetric(La, 100 + ti)
            x2 = spin_symmetric(Lb, 200 + ti)
            x0 = spin_symmetric(Lr, 300 + ti)
            ref = x0.copy()
            c.run(x1, x2, ref, route=3, use_tr=True, screen=False)
            A, B, R = (gecco_hip.MEList(o, *ccsd.LISTS[n]) for n in (a, b, res))
            plan = gecco_hip.Plan()
            plan.set_materialize(materialize)
            plan.set_chain(chain)
            nb, _ = plan.add_contraction(A, B, R, cc, ca, fact, True)
            plan.add_sym_ab(R, 2)
            plan.finalize()
            staged += plan.workspace()['staged_tasks']
            infos = [plan.block_info(i) for i in range(nb)]
            linked += sum(x[23] >= 0 for x in infos)
            skinny += sum(x[24] for x in infos)
            direct += sum(x[19] for x in infos)
            bufs = [torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev),
                    torch.from_numpy(x0).to(dev)]
            plan.execute(bufs)
            torch.cuda.synchronize()
            assert _relerr(bufs[2].cpu().numpy(), ref) <= TOL, (res, a, b)
    if materialize:
        assert staged > 0
    # chains link strided-GEMM calls only (gathered operands are on that
    # kernel only when materialised)
    assert (linked > 0) == bool(chain and materialize)
    assert skinny > 0 or not materialize     # 16 x 16 wave-split-K tiles
    assert direct > 0                        # short contractions on the vector ALU


@pytest.mark.parametrize('nsym', [1, 2, 4])
@pytest.mark.parametrize('name', ['pp_over_p', 'pp_full', 'hp_over_h', 'hp_full', 'ph_over_p',
                                  'p_full'])
def test_trace_matches_oracle(name, nsym):
    """Partial trace (trace_op) on the GPU against the oracle restatement,
    accumulating onto existing data, in a plan after a contraction."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    from oracle.melist import MEList as OMEList
    from oracle.trace import Trace
    from test_oracle_trace import CASES as TCASES
    print(f'trace {name} nsym={nsym}', flush=True)
    (occ_c, occ_a), cnt = TCASES[name]
    orbs = orbitals(4, 7, nsym)
    op = OMEList(orbs, occ_c, occ_a)
    tr_c = {h: n - cnt.get(h, 0) for h, n in occ_c.items() if n - cnt.get(h, 0)}
    tr_a = {h: n - cnt.get(h, 0) for h, n in occ_a.items() if n - cnt.get(h, 0)}
    trop = OMEList(orbs, tr_c, tr_a)
    x = spin_symmetric(op, 11)
    x0 = spin_symmetric(trop, 12)
    ref = Trace(op, trop, cnt, fact=-0.5).run(x, x0.copy())
    o = gecco_hip.Orbs(round_robin(4, nsym), round_robin(7, nsym, 1), nsym)
    A = gecco_hip.MEList(o, occ_c, occ_a)
    T = gecco_hip.MEList(o, tr_c, tr_a)
    p = gecco_hip.Plan()
    n = p.add_trace(A, T, cnt, -0.5, buf=(0, 1))
    assert n > 0
    p.finalize()
    dev = torch.device('cuda:0')
    bufs = [torch.from_numpy(x).to(dev), torch.from_numpy(x0).to(dev)]
    p.execute(bufs)
    torch.cuda.synchronize()
    assert _relerr(bufs[1].cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize('nsym', [1, 2, 4])
@pytest.mark.parametrize('name', ['p_to_v1', 'h_to_v0', 'ph_merge', 'mixed', 'split', 'p_to_empty'])
def test_reorder_matches_oracle(name, nsym):
    """Reorder of a two-vertex list (reo_op_wmaps_c) on the GPU against the
    oracle restatement, accumulating onto existing data."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    from oracle.melist import MEList as OMEList
    from oracle.reorder import Reorder
    from test_oracle_reorder import CASES as RCASES, _reo_occ
    print(f'reorder {name} nsym={nsym}', flush=True)
    oc, oa, frm, to, kc, ka = RCASES[name]
    orbs = orbitals(3, 6, nsym)
    ori = OMEList(orbs, oc, oa)
    rc, ra = _reo_occ(oc, oa, frm, to, kc, ka)
    reo = OMEList(orbs, rc, ra)
    x = np.random.default_rng(13).standard_normal(ori.length)
    y0 = np.random.default_rng(14).standard_normal(reo.length)
    ref = Reorder(ori, reo, frm, to, kc, ka, fact=0.25, sign_reo=-1).run(x, y0.copy())
    o = gecco_hip.Orbs(round_robin(3, nsym), round_robin(6, nsym, 1), nsym)
    A = gecco_hip.MEList(o, oc, oa)
    B = gecco_hip.MEList(o, rc, ra)
    p = gecco_hip.Plan()
    assert p.add_reorder(A, B, frm, to, kc, ka, 0.25, -1, buf=(0, 1)) > 0
    p.finalize()
    dev = torch.device('cuda:0')
    bufs = [torch.from_numpy(x).to(dev), torch.from_numpy(y0).to(dev)]
    p.execute(bufs)
    torch.cuda.synchronize()
    assert _relerr(bufs[1].cpu().numpy(), ref) <= TOL
    # the same through the flattened reo_op_wmaps_c entry, host and device
    # buffers, accumulating and overwriting
    irreps = (nsym, round_robin(3, nsym), round_robin(6, nsym, 1))
    vocc = lambda c, a: ([c] if isinstance(c, dict) else c, [a] if isinstance(a, dict) else a)
    y = y0.copy()
    gecco_hip.reo_op(0.25, vocc(oc, oa), vocc(rc, ra), frm, to, (kc, ka), x, y, irreps, sign_reo=-1)
    assert _relerr(y, ref) <= TOL
    y = y0.copy()
    gecco_hip.reo_op(0.25, vocc(oc, oa), vocc(rc, ra), frm, to, (kc, ka), x, y, irreps, update=False,
                     sign_reo=-1)
    assert _relerr(y, ref - y0) <= TOL
    dy = torch.from_numpy(y0).to(dev)
    gecco_hip.reo_op(0.25, vocc(oc, oa), vocc(rc, ra), frm, to, (kc, ka), bufs[0], dy, irreps,
                     sign_reo=-1)
    torch.cuda.synchronize()
    assert _relerr(dy.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (2, 4, 6), (4, 6, 10)])
@pytest.mark.parametrize('name', ['energy_t1', 'ring_t2', 't1_ovoo_t2', 'ladder_t12', 'fock_t12',
                                  'vvvo_t1'])
@pytest.mark.parametrize('use_tr', [True, False])
def test_transposed_operands_match_oracle(nsym, nocc, nvir, name, use_tr):
    """Kernel calls with tra_op1 / tra_op2 / tra_op1op2 set (operands and
    results used as adjoints, contract/contr_op1op2_wmaps_c.f:844-855): the
    oracle's calls recorded one by one into a plan, the product enumerator
    given adjoint views (gecco_hip_list_adjoint), and the Fortran-signature
    entry point contr_blk1blk2_blocked_mm_ on device pointers, each against
    the oracle."""
    torch = _torch()
    import gecco_hip
    from models import TRA_CASES, round_robin
    print(f'transposed case {name} nsym={nsym} nocc={nocc} nvir={nvir} use_tr={use_tr}', flush=True)
    orbs = orbitals(nocc, nvir, nsym)
    c, stored, tra = TRA_CASES[name](orbs)
    xs = [spin_symmetric(m, 91 + k) for k, m in enumerate(stored)]
    blocks = c.blocks(use_tr)
    assert blocks and all(bl['tra_op1'] == tra[0] and bl['tra_op2'] == tra[1] and
                          bl['tra_op1op2'] == tra[2] for bl in blocks)
    dev = torch.device('cuda:0')
    d1, d2 = (torch.from_numpy(x).to(dev) for x in xs[:2])
    # the oracle's kernel calls, one plan
    ref = _oracle(c, xs[0], xs[1], stored[2].length, blocks)
    dr = torch.zeros(stored[2].length, dtype=torch.float64, device=dev)
    plan = gecco_hip.Plan()
    for bl in blocks:
        plan.add(bl)
    plan.finalize()
    plan.execute([d1, d2, dr])
    torch.cuda.synchronize()
    assert _relerr(dr.cpu().numpy(), ref) <= TOL
    # the Fortran-signature kernel entry, block by block on device memory
    dr.zero_()
    for bl in blocks:
        gecco_hip.contr_blk1blk2_blocked_mm(bl, d1[bl['off1']:], d2[bl['off2']:], dr[bl['off12']:])
    torch.cuda.synchronize()
    assert _relerr(dr.cpu().numpy(), ref) <= TOL
    # the product's enumerator on adjoint views, with the projection
    full = xs[2].copy()
    c.run(xs[0], xs[1], full, route=3, use_tr=use_tr, screen=False)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    lists = [gecco_hip.MEList(o, m.occ_c, m.occ_a) for m in stored]
    used = [L.adjoint() if t else L for L, t in zip(lists, tra)]
    plan = gecco_hip.Plan()
    plan.add_contraction(used[0], used[1], used[2], c.cnt_c, c.cnt_a, 1.0, use_tr)
    if use_tr:
        plan.add_sym_ab(lists[2], 2)
    plan.finalize()
    dr = torch.from_numpy(xs[2]).to(dev)
    plan.execute([d1, d2, dr])
    torch.cuda.synchronize()
    assert _relerr(dr.cpu().numpy(), full) <= TOL


def test_fortran_entry_reuses_its_plans():
    """The Fortran-signature entries keep one plan per distinct kernel call:
    repeating the calls (host and device operands) adds the same result
    again without new plans, and another factor is another plan."""
    torch = _torch()
    import gecco_hip
    orbs = orbitals(4, 6, 2)
    c, a, b, r = CASES['ring'](orbs)
    x1 = spin_symmetric(a, 35)
    x2 = spin_symmetric(b, 36)
    blocks = c.blocks(True)
    ref = _oracle(c, x1, x2, r.length, blocks)
    gecco_hip.call_cache(clear=True)
    dev = torch.device('cuda:0')
    d1, d2 = torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev)
    dr = torch.zeros(r.length, dtype=torch.float64, device=dev)
    got = np.zeros(r.length)
    counts = []
    for rep in range(3):
        for bl in blocks:
            gecco_hip.contr_blk1blk2_blocked_mm(bl, d1[bl['off1']:], d2[bl['off2']:], dr[bl['off12']:])
            gecco_hip.contr_blk1blk2_wmaps_c(bl, x1[bl['off1']:], x2[bl['off2']:], got[bl['off12']:])
        counts.append(gecco_hip.call_cache())
    torch.cuda.synchronize()
    assert counts[0] == len(blocks) and counts[1] == counts[0] and counts[2] == counts[0]
    assert _relerr(dr.cpu().numpy(), 3 * ref) <= TOL
    assert _relerr(got, 3 * ref) <= TOL
    for bl in blocks:
        bl2 = dict(bl, xfac=-0.5 * bl['xfac'])
        gecco_hip.contr_blk1blk2_blocked_mm(bl2, d1[bl['off1']:], d2[bl['off2']:], dr[bl['off12']:])
    torch.cuda.synchronize()
    assert gecco_hip.call_cache() == 2 * len(blocks)
    assert _relerr(dr.cpu().numpy(), 2.5 * ref) <= TOL
    assert gecco_hip.call_cache(clear=True) == 0


@pytest.mark.parametrize('name', sorted(CASES) + ['energy_t1', 'ring_t2', 't1_ovoo_t2', 'ladder_t12',
                                                  'fock_t12', 'vvvo_t1'])
@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (4, 6, 10)])
def test_contr_op1op2_entry_matches_oracle(name, nsym, nocc, nvir):
    """gecco_hip_contr_op1op2_ (contr_op1op2's arguments flattened:
    occupations, dagger flags, update, use_tr, orbital IRREPs, whole lists)
    on host arrays and on device tensors, accumulating (update) and
    overwriting the result, against the oracle's contr_op1op2
    restatement."""
    torch = _torch()
    import gecco_hip
    from models import TRA_CASES, round_robin
    print(f'contr_op1op2 {name} nsym={nsym} nocc={nocc} nvir={nvir}', flush=True)
    orbs = orbitals(nocc, nvir, nsym)
    if name in CASES:
        c, a, b, r = CASES[name](orbs)
        stored, tra = (a, b, r), (False, False, False)
    else:
        c, stored, tra = TRA_CASES[name](orbs)
    used = (c.op1, c.op2, c.res)
    irreps = (nsym, round_robin(nocc, nsym), round_robin(nvir, nsym, 1))
    xs = [spin_symmetric(m, 51 + k) for k, m in enumerate(stored)]
    for update in (True, False):
        ref = xs[2].copy() if update else np.zeros(stored[2].length)
        c.run(xs[0], xs[1], ref, route=3, use_tr=True, screen=False)
        args = (0.75, (used[0].occ_c, used[0].occ_a), (used[1].occ_c, used[1].occ_a),
                (used[2].occ_c, used[2].occ_a), (c.cnt_c, c.cnt_a))
        host = xs[2].copy()
        ref75 = xs[2] + 0.75 * (ref - xs[2]) if update else 0.75 * ref
        for rep in range(2):        # the second call reuses the kept plan
            host = xs[2].copy()
            gecco_hip.contr_op1op2(*args, xs[0], xs[1], host, irreps, update=update, tra=tra)
            assert _relerr(host, ref75) <= TOL, (update, rep)
        dev = [torch.from_numpy(x).cuda() for x in xs]
        gecco_hip.contr_op1op2(*args, dev[0], dev[1], dev[2], irreps, update=update, tra=tra)
        torch.cuda.synchronize()
        assert _relerr(dev[2].cpu().numpy(), ref75) <= TOL


@pytest.mark.parametrize('name', ['ladder', 'ring', 't1t1', 'energy'])
@pytest.mark.parametrize('lenblock,lenscr', [(200, None), (7, 900)])
def test_collect_and_scatter_entries_inside_route_3(name, lenblock, lenscr):
    """collect_block_impr_ and scatter_block_ (the reference's Fortran
    argument lists, host arrays): the oracle's route 3 run with its collect
    and scatter steps replaced by the two entries gives the oracle's own
    result.  Small lenblock / lenscr split the strings and the contraction
    range into batches (partial KA/LA ranges, kca_trp for Op2), so every
    branch of the batch bookkeeping is used.  Collect moves data only (exact);
    scatter adds atomically (summation order: 1e-13 relative)."""
    _torch()
    import ctypes
    import gecco_hip
    from oracle.contract import use_collect_scatter
    if name not in CASES:
        pytest.skip(f'no case {name}')
    print(f'collect/scatter {name} lenblock={lenblock}', flush=True)
    orbs = orbitals(4, 6, 2)
    c, a, b, r = CASES[name](orbs)
    x1 = spin_symmetric(a, 41)
    x2 = spin_symmetric(b, 42)
    x0 = spin_symmetric(r, 43)
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=False, screen=False, lenblock=lenblock, lenscr=lenscr)
    got = x0.copy()
    fc = ctypes.cast(gecco_hip._lib.collect_block_impr_, ctypes.c_void_p).value
    fs = ctypes.cast(gecco_hip._lib.scatter_block_, ctypes.c_void_p).value
    try:
        use_collect_scatter(fc, fs)
        c.run(x1, x2, got, route=3, use_tr=False, screen=False, lenblock=lenblock, lenscr=lenscr)
    finally:
        use_collect_scatter(None, None)
    assert gecco_hip._lib.gecco_hip_last_error(None, 0) == 0
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize('name', ['ladder', 'ring', 't1t1', 'fock_pp'])
def test_norm_screening_matches_oracle(name):
    """Plan.set_screen (contract/contr_op1op2_wmaps_c.f:1652-1655,1767-1770):
    with one Op1 distribution block zeroed and one Op2 block scaled to 1e-16
    (squared norm below 1e-28), the screened plan equals the oracle's
    screened route 3 -- also where the skipped calls were the only
    contributions, which stay exactly zero -- and the unscreened plan the
    unscreened oracle."""
    torch = _torch()
    import gecco_hip
    print(f'screening {name}', flush=True)
    orbs = orbitals(4, 6, 2)
    c, a, b, r = CASES[name](orbs)
    x1, x2 = spin_symmetric(a, 51), spin_symmetric(b, 52)
    blocks = c.blocks(False)
    x1[blocks[0]['off1']:blocks[0]['off1'] + blocks[0]['len1']] = 0.0
    x2[blocks[-1]['off2']:blocks[-1]['off2'] + blocks[-1]['len2']] *= 1e-16
    from models import round_robin
    o = gecco_hip.Orbs(round_robin(4, 2), round_robin(6, 2, 1), 2)
    A, B, R = (gecco_hip.MEList(o, m.occ_c, m.occ_a) for m in (a, b, r))
    dev = torch.device('cuda:0')
    d1, d2 = torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev)
    out = {}
    for screen in (True, False):
        ref = np.zeros(r.length)
        c.run(x1, x2, ref, route=3, use_tr=False, screen=screen)
        p = gecco_hip.Plan()
        p.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, False)
        p.set_screen(screen)
        p.finalize()
        dr = torch.zeros(r.length, dtype=torch.float64, device=dev)
        p.execute([d1, d2, dr])
        torch.cuda.synchronize()
        got = dr.cpu().numpy()
        assert _relerr(got, ref) <= TOL, screen
        out[screen] = (got, ref)
    (g1, r1), (g0, r0) = out[True], out[False]
    only = (r1 == 0) & (r0 != 0)          # reached only by screened calls
    assert only.any() or name == 'ring'
    assert np.all(g1[only] == 0.0) and np.all(g0[only] != 0.0)


@pytest.mark.parametrize('name', ['t1', 't2', 'fae', 'wmbej', 't3'])
@pytest.mark.parametrize('nsym', [1, 2])
def test_spin_projection_matches_oracle(name, nsym):
    """gecco_hip_plan_add_spin_project (the singlet projector built from the
    combinatorics of [S+, tau_e] and a Jacobi null space) against the
    oracle's Fock-space projector (dense operators, SVD) on random lists;
    projecting twice changes nothing."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    from oracle.spin_project import SpinSpace
    from test_oracle_spin_project import LISTS
    print(f'spin projection {name} nsym={nsym}', flush=True)
    c, a = LISTS[name]
    orbs = orbitals(2, 2, nsym)
    from oracle.melist import MEList as OList
    Lo = OList(orbs, c, a)
    Pm = SpinSpace(2, 2).projector(Lo)
    o = gecco_hip.Orbs(round_robin(2, nsym), round_robin(2, nsym, 1), nsym)
    L = gecco_hip.MEList(o, c, a)
    assert L.length == Lo.length
    x = np.random.default_rng(7).standard_normal(L.length)
    p = gecco_hip.Plan()
    p.add_spin_project(L, 0)
    p.finalize()
    d = torch.from_numpy(x.copy()).to('cuda:0')
    p.execute([d])
    torch.cuda.synchronize()
    once = d.cpu().numpy()
    assert np.abs(once - Pm @ x).max() <= 1e-12 * max(1.0, np.abs(x).max())
    p.execute([d])
    torch.cuda.synchronize()
    assert np.abs(d.cpu().numpy() - once).max() <= 1e-12


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (2, 4, 6), (4, 6, 10)])
@pytest.mark.parametrize('name', ['t1t1_joined', 'vertex_fock', 'pairwise_scalar', 'ring_one_vertex',
                                  'result_merge'])
@pytest.mark.parametrize('use_tr', [True, False])
def test_joined_contraction_matches_oracle(nsym, nocc, nvir, name, use_tr):
    """Contractions of operators with several vertices
    (gecco_hip_plan_add_contraction_n) on the GPU against the oracle's
    multi-vertex contr_op1op2 restatement, accumulating onto existing data."""
    torch = _torch()
    import gecco_hip
    from models import round_robin, JOINED_CASES
    print(f'joined {name} nsym={nsym} nocc={nocc} nvir={nvir} use_tr={use_tr}', flush=True)
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = JOINED_CASES[name](orbs)
    x1, x2, x0 = spin_symmetric(a, 64), spin_symmetric(b, 65), spin_symmetric(r, 66)
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=use_tr, screen=False)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    A, B, R = (gecco_hip.MEList(o, m.vocc_c, m.vocc_a) for m in (a, b, r))
    plan = gecco_hip.Plan()
    m1, m2, m12, m21 = c.merge
    nb, fl = plan.add_contraction_n(A, B, R, c.ex1_vocc, c.ex2_vocc, c.cnt_vocc, m1, m2, m12, m21,
                                    1.0, use_tr)
    assert nb == len(c.blocks(use_tr))
    if use_tr:
        plan.add_sym_ab(R, 2)
    plan.finalize()
    dev = torch.device('cuda:0')
    bufs = [torch.from_numpy(x).to(dev) for x in (x1, x2, x0)]
    plan.execute(bufs)
    torch.cuda.synchronize()
    assert _relerr(bufs[2].cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize('nsym,nocc,nvir', [(1, 3, 5), (4, 6, 10)])
def test_joined_product_reordered_in_one_plan(nsym, nocc, nvir):
    """The reorder after a contraction in one plan: T1 x T1 into the
    two-vertex X, then X's second vertex moved onto its first (the plan runs
    reorders after its contractions) -- equal to the one-vertex product
    R(ab,ij) += T1(a,i) T1(b,j) of the oracle."""
    torch = _torch()
    import gecco_hip
    from models import round_robin, t1t1
    print(f'joined+reorder nsym={nsym} nocc={nocc} nvir={nvir}', flush=True)
    orbs = orbitals(nocc, nvir, nsym)
    c1, T1o, _, Ro = t1t1(orbs)
    x = spin_symmetric(T1o, 67)
    ref = np.zeros(Ro.length)
    c1.run(x, x, ref, route=3, use_tr=False, screen=False)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    P, H = 2, 1
    T1 = gecco_hip.MEList(o, {P: 1}, {H: 1})
    X = gecco_hip.MEList(o, [{P: 1}, {P: 1}], [{H: 1}, {H: 1}])
    Xp = gecco_hip.MEList(o, [{P: 2}, {}], [{H: 2}, {}])
    assert Xp.length == Ro.length
    plan = gecco_hip.Plan()
    plan.add_init(2, X.length)
    plan.add_contraction_n(T1, T1, X, [({P: 1}, {H: 1})], [({P: 1}, {H: 1})], [], [0, 1, 1],
                           [0, 1, 1], [1, 1, 0, 0, 1, 1], [0, 1, 1, 1, 1, 0], 1.0, False,
                           buf=(0, 0, 2))
    plan.add_reorder(X, Xp, 1, 0, {P: 1}, {H: 1}, buf=(2, 3))
    plan.finalize()
    dev = torch.device('cuda:0')
    bufs = [torch.from_numpy(x).to(dev), torch.zeros(1, dtype=torch.float64, device=dev),
            torch.zeros(X.length, dtype=torch.float64, device=dev),
            torch.zeros(Xp.length, dtype=torch.float64, device=dev)]
    plan.execute(bufs)
    plan.execute(bufs)      # the initialisation restarts X; the reorder accumulates
    torch.cuda.synchronize()
    assert _relerr(bufs[3].cpu().numpy(), 2 * ref) <= TOL


@pytest.mark.parametrize('name', ['t1t1_joined', 'vertex_fock', 'pairwise_scalar', 'ring_one_vertex',
                                  'result_merge'])
def test_flattened_joined_entry_matches_oracle(name):
    """gecco_hip_contr_op1op2_n_ (contr_op1op2's multi-vertex argument list,
    flattened) against the oracle, with host and device buffers, update on
    and off, and with the reorder after the kernel calls (reo_info): T1 x T1
    into the two-vertex op1op2tmp, reordered into the one-vertex R."""
    torch = _torch()
    import gecco_hip
    from models import round_robin, JOINED_CASES, t1t1
    nocc, nvir, nsym = 4, 6, 2
    print(f'flattened joined {name}', flush=True)
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = JOINED_CASES[name](orbs)
    irreps = (nsym, round_robin(nocc, nsym), round_robin(nvir, nsym, 1))
    x1, x2, x0 = spin_symmetric(a, 71), spin_symmetric(b, 72), spin_symmetric(r, 73)
    occ = lambda m: (m.vocc_c, m.vocc_a)
    for use_tr in (True, False):
        ref = x0.copy()
        c.run(x1, x2, ref, route=3, use_tr=use_tr, screen=False)
        y = x0.copy()
        gecco_hip.contr_op1op2_n(1.0, occ(a), occ(b), occ(r), c.ex1_vocc, c.ex2_vocc, c.cnt_vocc, c.merge,
                                 x1, x2, y, irreps, use_tr=use_tr)
        assert _relerr(y, ref) <= TOL
        y = x0.copy()
        gecco_hip.contr_op1op2_n(1.0, occ(a), occ(b), occ(r), c.ex1_vocc, c.ex2_vocc, c.cnt_vocc, c.merge,
                                 x1, x2, y, irreps, update=False, use_tr=use_tr)
        assert _relerr(y, ref - x0) <= TOL
    dev = torch.device('cuda:0')
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=True, screen=False)
    dy = torch.from_numpy(x0).to(dev)
    gecco_hip.contr_op1op2_n(1.0, occ(a), occ(b), occ(r), c.ex1_vocc, c.ex2_vocc, c.cnt_vocc, c.merge,
                             torch.from_numpy(x1).to(dev), torch.from_numpy(x2).to(dev), dy, irreps)
    torch.cuda.synchronize()
    assert _relerr(dy.cpu().numpy(), ref) <= TOL
    if name == 't1t1_joined':          # with the reorder into the one-vertex product
        c1, T1o, _, Ro = t1t1(orbs)
        xt = spin_symmetric(T1o, 74)
        r0 = spin_symmetric(Ro, 75)
        ref = r0.copy()
        c1.run(xt, xt, ref, route=3, use_tr=False, screen=False)
        y = r0.copy()
        P, H = 2, 1
        gecco_hip.contr_op1op2_n(1.0, occ(T1o), occ(T1o), ([{P: 2}, {}], [{H: 2}, {}]), c.ex1_vocc,
                                 c.ex2_vocc, c.cnt_vocc, c.merge, xt, xt, y, irreps, use_tr=False,
                                 op1op2tmp=occ(r), reo=(1, 0, ({P: 1}, {H: 1}), 1))
        assert _relerr(y, ref) <= TOL


@pytest.mark.parametrize('name', list(CASES) + ['t3_ring', 't3_pp_ladder'])
@pytest.mark.parametrize('variant', [1, 9])
def test_fixed_order_plans_are_bit_reproducible(name, variant, monkeypatch):
    """gecco_hip_plan_set_deterministic: the plan matches the oracle, and
    two executes from the same inputs give bit-identical results (every
    result element's contributions summed in task order), with split-K,
    chains, merged-string (table) results and direct tasks in the launch."""
    torch = _torch()
    import gecco_hip
    from models import round_robin
    monkeypatch.setenv('GECCO_HIP_AFFINE_TILE', str(variant))
    print(f'fixed-order {name} variant {variant}', flush=True)
    nocc, nvir, nsym = (4, 9, 2) if name.startswith('t3') else (6, 20, 2)
    orbs = orbitals(nocc, nvir, nsym)
    c, a, b, r = {**CASES, **T3_CASES}[name](orbs)
    x1, x2, x0 = spin_symmetric(a, 81), spin_symmetric(b, 82), spin_symmetric(r, 83)
    ref = x0.copy()
    c.run(x1, x2, ref, route=3, use_tr=True, screen=False)
    o = gecco_hip.Orbs(round_robin(nocc, nsym), round_robin(nvir, nsym, 1), nsym)
    A, B, R = (gecco_hip.MEList(o, m.occ_c, m.occ_a) for m in (a, b, r))
    plan = gecco_hip.Plan()
    plan.set_deterministic(True)
    plan.set_tuning(target_tasks=4096, min_kchunk=32)      # plenty of split-K
    plan.add_contraction(A, B, R, c.cnt_c, c.cnt_a, 1.0, True)
    plan.add_sym_ab(R, 2)
    plan.finalize()
    dev = torch.device('cuda:0')
    outs = []
    for rep in range(3):
        bufs = [torch.from_numpy(x).to(dev) for x in (x1, x2, x0)]
        plan.execute(bufs)
        torch.cuda.synchronize()
        outs.append(bufs[2].cpu().numpy())
    assert _relerr(outs[0], ref) <= TOL
    assert all(np.array_equal(outs[0], y) for y in outs[1:])


def test_fixed_order_ccsd_solves_are_bit_identical():
    """CCSD with every plan in fixed-order mode (CCSD(deterministic=True)):
    two solves give bit-identical energies and amplitudes, equal to the
    default (atomic) solve within 1e-10."""
    torch = _torch()
    import ccsd
    from synth import synthetic
    ham = synthetic(5, 16, 48, nsym=2, seed=23, sigma=0.08)
    print('fixed-order ccsd 5/16/2', flush=True)
    e0 = ccsd.CCSD(ham).solve(tol=1e-10, maxit=80)['energy']
    runs = []
    for rep in range(2):
        c = ccsd.CCSD(ham, deterministic=True)
        r = c.solve(tol=1e-10, maxit=80)
        runs.append((r['energy'], r['iterations'], c.amp.cpu().numpy().copy()))
    assert runs[0][0] == runs[1][0] and runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][2], runs[1][2])
    assert abs(runs[0][0] - e0) < 1e-10
