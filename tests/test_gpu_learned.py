"""GPU parity of the learned preconditioner (csrc/mlp.hip) against the
oracle (oracle/mlp.py), and what the trained network gains over the
diagonal preconditioner.

Tolerances: inference and features run in float32 with bfloat16 weights and
activations, the oracle in float64 with the same roundings.  A float32
first-layer sum can land on the other side of a bfloat16 rounding boundary
than the float64 one, so single outputs may differ by a few bfloat16 ulps of
one activation: max |dy| <= 2e-3 (1 + |y|), mean |dy| <= 1e-5 (1 + mean |y|).
Gradients: float32 accumulation over the samples, relative 2e-3 in norm;
loss relative 1e-5.
"""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


def _params(seed):
    rng = np.random.default_rng(seed)
    p = omlp.init_params(seed)
    p[omlp.OFF['b1']:omlp.OFF['b1'] + 128] = rng.uniform(-0.2, 0.2, 128)
    p[omlp.OFF['b2']:omlp.OFF['b2'] + 128] = rng.uniform(-0.2, 0.2, 128)
    p[omlp.OFF['w3']:omlp.OFF['w3'] + 128] = rng.uniform(-0.3, 0.3, 128)
    p[omlp.OFF['b3']] = 0.1
    return p.astype(np.float32)


def _close(got, ref):
    err = np.abs(got - ref)
    return (float(err.max() / (1 + np.abs(ref).max())), float(err.mean() / (1 + np.abs(ref).mean())))


@pytest.mark.parametrize('n', [1000, 64, 7])
def test_mlp_infer_matches_oracle(n):
    torch = _torch()
    import gecco_hip
    print(f'mlp infer n={n}', flush=True)
    p = _params(1)
    x = np.random.default_rng(2).standard_normal((n, 4)).astype(np.float32)
    dev = torch.device('cuda:0')
    y = gecco_hip.mlp_infer(torch.from_numpy(p).to(dev), torch.from_numpy(x).to(dev)).cpu().numpy()
    ref, _ = omlp.forward(p, x)
    mx, mean = _close(y, ref)
    print(f'max {mx:.2e} mean {mean:.2e}')
    assert mx <= 2e-3 and mean <= 1e-5


def test_mlp_scale_matches_oracle():
    """Features from (diag, rest, t) on the device, then the network."""
    torch = _torch()
    import gecco_hip
    rng = np.random.default_rng(5)
    n = 5003
    diag = rng.uniform(0.5, 9.0, n)
    diag[::101] = 1e-12                      # the solver's |D| <= 1e-10 threshold
    t = 0.01 * rng.standard_normal(n)
    rest = 0.05 * rng.standard_normal(n)
    p = _params(3)
    dev = torch.device('cuda:0')
    dd, rr, tt = (torch.from_numpy(v).to(dev) for v in (diag, rest, t))
    s = torch.empty(n, dtype=torch.float64, device=dev)
    work = torch.empty(gecco_hip.MLP_SCALE_WORK, dtype=torch.float64, device=dev)
    gecco_hip.mlp_scale(torch.from_numpy(p).to(dev), dd, rr, tt, 7, s, work)
    x, u, ru = omlp.features(diag, rest, t, 7)
    assert abs(float(work[0]) - ru) <= 1e-12 * ru
    assert abs(float(work[1]) - np.sqrt(np.mean(t * t))) <= 1e-12 * np.sqrt(np.mean(t * t))
    ref = 1.0 + omlp.forward(p, x)[0]
    mx, mean = _close(s.cpu().numpy(), ref)
    print(f'max {mx:.2e} mean {mean:.2e}')
    assert mx <= 2e-3 and mean <= 1e-5


@pytest.mark.parametrize('n', [3000, 50])
def test_mlp_grad_matches_oracle(n):
    torch = _torch()
    import gecco_hip
    print(f'mlp grad n={n}', flush=True)
    rng = np.random.default_rng(n)
    p = _params(4)
    x = rng.standard_normal((n, 4)).astype(np.float32)
    ut = rng.standard_normal(n).astype(np.float32)
    dt = (-ut * rng.uniform(0.5, 1.5, n)).astype(np.float32)
    dev = torch.device('cuda:0')
    grad = torch.zeros(gecco_hip.MLP_P, dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float64, device=dev)
    work = torch.empty(gecco_hip.MLP_GRAD_WORK, dtype=torch.float32, device=dev)
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to(dev)
    gecco_hip.mlp_grad(t(p), t(x), t(ut), t(dt), grad, loss, work)
    lref, gref = omlp.loss_grad(p, x, ut, dt)
    g = grad.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f'grad rel {rel:.2e} loss {float(loss):.6e} ref {lref:.6e}')
    assert rel <= 2e-3
    assert abs(float(loss) - lref) <= 1e-5 * lref
    # a second call: workgroup partials are added in a fixed order, the
    # small parameter groups inside a workgroup with LDS atomics (last bits)
    grad2 = torch.zeros_like(grad)
    gecco_hip.mlp_grad(t(p), t(x), t(ut), t(dt), grad2, None, work)
    assert float((grad - grad2).norm() / grad.norm()) <= 1e-6


def test_mlp_adam_matches_oracle():
    torch = _torch()
    import gecco_hip
    rng = np.random.default_rng(8)
    P = gecco_hip.MLP_P
    p = rng.standard_normal(P).astype(np.float32)
    g1 = rng.standard_normal(P).astype(np.float32)
    g2 = rng.standard_normal(P).astype(np.float32)
    dev = torch.device('cuda:0')
    dp = torch.from_numpy(p).to(dev)
    m = torch.zeros(P, dtype=torch.float32, device=dev)
    v = torch.zeros_like(m)
    gecco_hip.mlp_adam(dp, m, v, torch.from_numpy(g1).to(dev), 1, 1e-2)
    gecco_hip.mlp_adam(dp, m, v, torch.from_numpy(g2).to(dev), 2, 1e-2)
    q, mm, vv = omlp.adam(p.astype(np.float64), np.zeros(P), np.zeros(P), g1, 1, 1e-2)
    q, mm, vv = omlp.adam(q, mm, vv, g2, 2, 1e-2)
    assert np.abs(dp.cpu().numpy() - q).max() <= 1e-5


def test_learned_preconditioner_training_and_gain():
    """Pairs from plain solves of four small synthetic Hamiltonians train
    the network on the GPU; on a held-out Hamiltonian the learned and the
    diagonal preconditioner reach the same energy, and the iteration counts
    are reported."""
    torch = _torch()
    import ccsd
    import learned
    from synth import synthetic
    train_h = [synthetic(4, 12, 36, nsym=2, seed=11, sigma=0.1),
               synthetic(5, 14, 42, nsym=2, seed=12, sigma=0.1),
               synthetic(6, 16, 48, nsym=1, seed=13, sigma=0.1),
               synthetic(5, 15, 45, nsym=4, seed=14, sigma=0.1)]
    parts = []
    for h in train_h:
        data, nit = learned.collect(ccsd.CCSD(h), maxit=60, tol=1e-10, per_iter=2048, seed=1)
        parts.append(data)
        print('collected', data[1].numel(), 'pairs from', nit, 'iterations', flush=True)
    data = tuple(torch.cat([d[i] for d in parts]).contiguous() for i in range(3))
    pre = learned.LearnedPreconditioner(seed=0)
    l0 = pre.loss(data)
    losses = pre.train(data, steps=300, lr=2e-3, batch=32768)
    l1 = pre.loss(data)
    print(f'loss {l0:.4e} -> {l1:.4e}', flush=True)
    assert l1 < 0.9 * l0
    held = synthetic(5, 16, 48, nsym=2, seed=21, sigma=0.1)
    solver = ccsd.CCSD(held)
    plain = solver.solve(tol=1e-10, maxit=150, mode='sync')
    lrn = solver.solve(tol=1e-10, maxit=150, mode='sync', precond=pre)
    print(f"diagonal: {plain['iterations']} iterations; learned: {lrn['iterations']} iterations, "
          f"fallback at {lrn['fallback_at']}", flush=True)
    assert plain['residual'] < 1e-10 and lrn['residual'] < 1e-10
    assert abs(lrn['energy'] - plain['energy']) < 1e-9
    # measured: 29 (diagonal) against 27 (learned) iterations, no fallback
    assert lrn['fallback_at'] is None
    assert lrn['iterations'] <= plain['iterations']
    # the untrained network is the plain preconditioner
    untrained = solver.solve(tol=1e-10, maxit=150, mode='sync',
                             precond=learned.LearnedPreconditioner(seed=0))
    assert untrained['iterations'] == plain['iterations']
    assert abs(untrained['energy'] - plain['energy']) < 1e-12
