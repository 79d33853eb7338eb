"""GPU parity of the native amplitude solver (diagonal preconditioner + DIIS,
gecco-public_amd/csrc/solver.hip) against the oracle restatement of
baserout/diavc.f and solve/optc_diis_*.f, and of plan list initialisation.

Tolerance: FP64; the device reductions sum in a different order than numpy:
relative 1e-11 on amplitudes, 1e-12 on norms.
"""
import numpy as np
import pytest

from oracle.diis import DIIS as OracleDIIS

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


@pytest.mark.parametrize('n,maxsub,steps', [(1000, 8, 6), (3000017, 4, 9), (257, 1, 3)])
def test_diis_matches_oracle(n, maxsub, steps):
    """Non-linear vector function g(t) = D t + b + c*t^2 (elementwise) plus a
    coupling term; the native step and the oracle step see the same g."""
    torch = _torch()
    import gecco_hip
    print(f'diis n={n} maxsub={maxsub} steps={steps}', flush=True)
    rng = np.random.default_rng(n)
    d = rng.uniform(0.5, 3.0, n)
    # entries with |D| <= 1e-10 exercise the divide threshold; their
    # gradient is a small constant, so r there is 1e-14 / 1e-10
    tiny = np.zeros(n, dtype=bool)
    tiny[::97] = True
    tiny[1::97] = True
    d[::97] = 0.0
    d[1::97] = -3e-11
    b = 0.1 * rng.standard_normal(n)
    c = 0.05 * rng.standard_normal(n)

    def rest_of(t):                    # g - D t
        r = b + c * t * t + 0.01 * np.roll(t, 1)
        r[tiny] = 1e-14 - d[tiny] * t[tiny]
        return r

    dev = torch.device('cuda:0')
    t_dev = torch.zeros(n, dtype=torch.float64, device=dev)
    d_dev = torch.from_numpy(d).to(dev)
    e_dev = torch.tensor([1.25], dtype=torch.float64, device=dev)
    diis = gecco_hip.DIIS(n, maxsub)
    orc = OracleDIIS(maxsub)
    t = np.zeros(n)
    for it in range(steps):
        rest = rest_of(t)
        g = d * t + rest
        gn, en, conv = diis.step(t_dev, torch.from_numpy(rest).to(dev), d_dev, energy=e_dev)
        assert not conv
        assert en == 1.25
        assert abs(gn - np.linalg.norm(g)) <= 1e-12 * np.linalg.norm(g)
        t = orc.step(t, g, d)
        got = t_dev.cpu().numpy()
        scale = max(1e-300, np.abs(t).max())
        assert np.abs(got - t).max() <= 1e-11 * scale, (it, np.abs(got - t).max() / scale)
        # the subspace matrix holds <r_i|r_j> of the kept vectors; the
        # weights inherit the conditioning of the augmented system
        R = np.array(orc.r)
        B = R @ R.T
        m = len(orc.r)
        aug = np.zeros((m + 1, m + 1))
        aug[:m, :m] = B
        aug[:m, m] = aug[m, :m] = -1.0
        wtol = 1e-15 * max(1.0, np.linalg.cond(aug)) * 10
        np.testing.assert_allclose(diis.weights(), orc.w, rtol=0, atol=max(1e-13, wtol))
        np.testing.assert_allclose(diis.matrix(), B, rtol=0, atol=1e-12 * np.abs(B).max())
        t = got                        # both continue from the same amplitudes


def test_diis_convergence_flag_leaves_t():
    torch = _torch()
    import gecco_hip
    n = 500
    dev = torch.device('cuda:0')
    d = torch.full((n,), 2.0, dtype=torch.float64, device=dev)
    t = torch.full((n,), 0.5, dtype=torch.float64, device=dev)
    rest = torch.full((n,), -1.0, dtype=torch.float64, device=dev)   # g = 2*0.5 - 1 = 0
    diis = gecco_hip.DIIS(n, 4)
    gn, _, conv = diis.step(t, rest, d, tol=1e-12)
    assert conv and gn == 0.0
    assert torch.equal(t, torch.full((n,), 0.5, dtype=torch.float64, device=dev))
    diis.reset()
    assert diis.matrix().shape == (0, 0)


def test_plan_list_initialisation():
    torch = _torch()
    import gecco_hip
    dev = torch.device('cuda:0')
    rng = np.random.default_rng(9)
    src = rng.standard_normal(40000)
    bufs = [torch.from_numpy(src).to(dev), torch.full((50000,), 7.0, dtype=torch.float64, device=dev),
            torch.full((300,), 3.0, dtype=torch.float64, device=dev)]
    p = gecco_hip.Plan()
    p.add_init(1, 30000, buf_src=0, factor=-0.5, off_dst=100, off_src=5000)
    p.add_init(2, 200, off_dst=50)              # zero
    p.finalize()
    p.execute(bufs)
    torch.cuda.synchronize()
    b1 = np.full(50000, 7.0)
    b1[100:30100] = -0.5 * src[5000:35000]
    b2 = np.full(300, 3.0)
    b2[50:250] = 0.0
    np.testing.assert_array_equal(bufs[1].cpu().numpy(), b1)
    np.testing.assert_array_equal(bufs[2].cpu().numpy(), b2)


@pytest.mark.parametrize('n,maxsub,steps', [(1000, 8, 12), (300001, 4, 9)])
def test_async_step_matches_host_step(n, maxsub, steps):
    """The device-resident step (bookkeeping and the augmented solve on the
    GPU, records in pinned memory) against the host-solve step on the same
    vector function."""
    torch = _torch()
    import gecco_hip
    print(f'diis async n={n} maxsub={maxsub}', flush=True)
    rng = np.random.default_rng(n + 1)
    d = rng.uniform(0.5, 3.0, n)
    b = 0.1 * rng.standard_normal(n)
    c = 0.05 * rng.standard_normal(n)
    dev = torch.device('cuda:0')
    d_dev = torch.from_numpy(d).to(dev)
    e_dev = torch.tensor([-2.5], dtype=torch.float64, device=dev)
    sync, asy = gecco_hip.DIIS(n, maxsub), gecco_hip.DIIS(n, maxsub)
    t1 = torch.zeros(n, dtype=torch.float64, device=dev)
    t2 = torch.zeros(n, dtype=torch.float64, device=dev)
    for it in range(steps):
        tt = t1.cpu().numpy()
        rest = torch.from_numpy(b + c * tt * tt + 0.01 * np.roll(tt, 1)).to(dev)
        gn, en, conv = sync.step(t1, rest, d_dev, energy=e_dev)
        asy.step_async(t2, rest, d_dev, energy=e_dev)
        torch.cuda.synchronize()
        gn2, en2, code = asy.status(it)
        assert code == 0 and en2 == en == -2.5
        assert abs(gn2 - gn) <= 1e-12 * gn
        scale = float(t1.abs().max())
        assert float((t1 - t2).abs().max()) <= 1e-11 * scale, it
        np.testing.assert_allclose(asy.matrix(), sync.matrix(), rtol=0, atol=1e-13 * np.abs(sync.matrix()).max())
        t2.copy_(t1)                   # both continue from the same amplitudes
    with pytest.raises(gecco_hip.GeccoHipError):
        asy.status(steps)              # not taken yet
    with pytest.raises(gecco_hip.GeccoHipError):
        sync.step_async(t1, rest, d_dev)   # subspace holds synchronous steps


def test_async_step_converged_leaves_t():
    torch = _torch()
    import gecco_hip
    n = 500
    dev = torch.device('cuda:0')
    d = torch.full((n,), 2.0, dtype=torch.float64, device=dev)
    t = torch.full((n,), 0.5, dtype=torch.float64, device=dev)
    rest = torch.full((n,), -1.0, dtype=torch.float64, device=dev)
    diis = gecco_hip.DIIS(n, 4)
    diis.step_async(t, rest, d, tol=1e-12)
    torch.cuda.synchronize()
    gn, _, code = diis.status(0)
    assert code == 1 and gn == 0.0
    assert torch.equal(t, torch.full((n,), 0.5, dtype=torch.float64, device=dev))
    diis.reset()
    with pytest.raises(gecco_hip.GeccoHipError):
        diis.status(0)


def test_async_flag_freezes_converged_steps():
    """With a convergence flag, the converging asynchronous step sets it;
    later steps leave t and the subspace unchanged whatever residual they
    are given and repeat the converged record until the flag is cleared."""
    torch = _torch()
    import gecco_hip
    n = 700
    dev = torch.device('cuda:0')
    g = torch.Generator(device='cpu').manual_seed(3)
    v = torch.randn(n, generator=g, dtype=torch.float64).to(dev)
    d = (1.0 + torch.rand(n, generator=g, dtype=torch.float64)).to(dev)
    t = torch.zeros(n, dtype=torch.float64, device=dev)
    flag = torch.zeros(1, dtype=torch.int32, device=dev)
    diis = gecco_hip.DIIS(n, 4)
    diis.set_flag(flag)
    # with t = 0 before each step, g = D t + rest = c v and ||g|| = c ||v||
    tol = 1e-3 * float(torch.linalg.norm(v))
    codes = []
    for k, c in enumerate([1.0, 0.3, 0.05, 1e-4]):
        t.zero_()
        diis.step_async(t, c * v, d, tol=tol)
        torch.cuda.synchronize()
        gn, _, code = diis.status(k)
        codes.append(code)
        assert int(flag.item()) == (1 if code == 1 else 0)
    assert codes == [0, 0, 0, 1]
    gn_conv = diis.status(3)[0]
    w = torch.randn(n, generator=g, dtype=torch.float64).to(dev)
    t.copy_(w)
    t_conv = t.clone()
    b_conv = diis.matrix()
    diis.step_async(t, 5.0 * v, d, tol=tol)      # a large gradient: ignored while flagged
    torch.cuda.synchronize()
    gn, _, code = diis.status(4)
    assert code == 1 and gn == gn_conv
    assert torch.equal(t, t_conv)
    assert np.array_equal(diis.matrix(), b_conv)
    # cleared flag and reset: steps run again
    flag.zero_()
    diis.reset()
    diis.step_async(t, 5.0 * v, d, tol=tol)
    torch.cuda.synchronize()
    assert diis.status(0)[2] == 0 and int(flag.item()) == 0
    assert not torch.equal(t, t_conv)
    with pytest.raises(gecco_hip.GeccoHipError):
        diis.set_flag(torch.zeros(2, dtype=torch.int32, device=dev))


def test_diis_group_matches_single_steps():
    """Three subspaces of different lengths stepped as one group give the
    same amplitudes and records as separate asynchronous steps."""
    torch = _torch()
    import gecco_hip
    dev = torch.device('cuda:0')
    ns = [1000, 70001, 313]
    rng = np.random.default_rng(17)
    probs = []
    for n in ns:
        d = rng.uniform(0.5, 3.0, n)
        b = 0.1 * rng.standard_normal(n)
        c = 0.05 * rng.standard_normal(n)
        probs.append((torch.from_numpy(d).to(dev), b, c))
    single = [gecco_hip.DIIS(n, 4) for n in ns]
    grouped = [gecco_hip.DIIS(n, 4) for n in ns]
    t1 = [torch.zeros(n, dtype=torch.float64, device=dev) for n in ns]
    t2 = [torch.zeros(n, dtype=torch.float64, device=dev) for n in ns]
    rest = [torch.zeros(n, dtype=torch.float64, device=dev) for n in ns]
    e = [torch.tensor([-1.0 - k], dtype=torch.float64, device=dev) for k in range(3)]
    group = gecco_hip.DIISGroup(grouped, t2, rest, [p[0] for p in probs], e)
    for it in range(7):
        for k, (d, b, c) in enumerate(probs):
            tt = t1[k].cpu().numpy()
            rest[k].copy_(torch.from_numpy(b + c * tt * tt + 0.01 * np.roll(tt, 1)))
            t2[k].copy_(t1[k])
            single[k].step_async(t1[k], rest[k], d, energy=e[k])
        group.step_async()
        torch.cuda.synchronize()
        for k in range(3):
            g1, e1, c1 = single[k].status(it)
            g2, e2, c2 = grouped[k].status(it)
            assert c1 == c2 == 0 and e1 == e2 == -1.0 - k
            assert abs(g1 - g2) <= 1e-12 * g1
            assert torch.equal(t1[k], t2[k]), (it, k)


def test_singular_subspace_is_reported_and_stops_the_solve():
    """Two bitwise-identical preconditioned steps make the augmented DIIS
    matrix exactly singular.  The reference stops there
    (solve/optc_diis_extr.f:70); the synchronous step raises, the
    asynchronous record reads -1, and ccsd.solve raises in every mode
    instead of reporting convergence."""
    torch = _torch()
    import gecco_hip
    import ccsd
    from synth import synthetic
    dev = torch.device('cuda:0')
    n = 5000
    g = torch.linspace(-1e-2, 1e-2, n, dtype=torch.float64, device=dev) + 1e-3
    d = torch.full((n,), 2.0, dtype=torch.float64, device=dev)
    t0 = torch.zeros(n, dtype=torch.float64, device=dev)
    # synchronous steps: the second sees the same g at the same t
    diis = gecco_hip.DIIS(n, 4)
    t = t0.clone()
    diis.step(t, g, d, tol=0.0)
    t.copy_(t0)
    with pytest.raises(gecco_hip.GeccoHipError, match='singular'):
        diis.step(t, g, d, tol=0.0)
    # asynchronous steps: the record of the second step is -1
    diis = gecco_hip.DIIS(n, 4)
    t = t0.clone()
    diis.step_async(t, g, d, tol=0.0)
    t.copy_(t0)
    diis.step_async(t, g, d, tol=0.0)
    torch.cuda.synchronize()
    assert diis.status(0)[2] == 0 and diis.status(1)[2] == -1
    # the solver: every evaluation returns g at T = 0
    ham = synthetic(2, 5, 15, nsym=1, seed=3, sigma=0.1)
    c = ccsd.CCSD(ham)
    g0 = torch.linspace(1e-4, 2e-3, c.amp.numel(), dtype=torch.float64, device=dev)

    def evaluate(stream=None):
        c.amp.zero_()
        c.res.copy_(g0)
    c.evaluate = evaluate
    for mode in ('sync', 'async', 'graph'):
        print('singular subspace, mode', mode, flush=True)
        with pytest.raises(gecco_hip.GeccoHipError, match='singular'):
            c.solve(tol=1e-14, maxit=6, mode=mode)
