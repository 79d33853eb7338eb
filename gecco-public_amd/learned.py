"""Learned per-amplitude preconditioner for the amplitude solver.

The reference preconditions a step with the orbital-energy denominator,
r = g / D (contract/dia4op_ev.f, baserout/diavc.f:31-35, used from
solve/optc_diis_sbsp_add.f:125-140).  Here a small network
(4 -> 128 -> 128 -> 1, SiLU; include/gecco_hip_mlp.h) scales that step per
amplitude, r = s * g / D with s = 1 + y(x), from the features
    x = [D/4 - 1, asinh(u / rms(u)), asinh(t / rms(t)), min(it, 32)/32],  u = g / D.
Inference runs in bfloat16 on MFMA (gecco_hip_mlp_scale), the DIIS step
takes s through gecco_hip_diis_set_scale, and the network is trained on the
GPU by the library's own gradient and Adam kernels on pairs (features,
accepted update) collected from plain-solver runs: the accepted update of a
DIIS step against the plain step -u, both over rms(u), with the loss
mean((1 + y) u~ + dt~)^2.  With the output layer at zero the network is the
plain preconditioner; CCSD.solve falls back to it when the residual norm
grows (see ccsd.CCSD.solve).
"""
import numpy as np

import gecco_hip

P = gecco_hip.MLP_P


def initial_params(seed=0):
    """Uniform(-1/2, 1/2) first layer, uniform(-1, 1)/sqrt(128) second
    layer, zero biases and output layer (y = 0: the plain step)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    p[0:512] = rng.uniform(-0.5, 0.5, 512)
    p[640:17024] = rng.uniform(-1.0, 1.0, 128 * 128) / np.sqrt(128.0)
    return p


class LearnedPreconditioner:
    def __init__(self, params=None, device='cuda:0', seed=0):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        p = initial_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if p.shape != (P,):
            raise ValueError(f'expected {P} parameters, got {p.shape}')
        self.params = torch.from_numpy(p.copy()).to(self.device)
        self.work = torch.empty(gecco_hip.MLP_SCALE_WORK, dtype=torch.float64, device=self.device)

    # -------------------------------------------------------------- files
    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy())

    @classmethod
    def load(cls, path, device='cuda:0'):
        with np.load(path) as f:           # numeric arrays only (no pickle)
            return cls(f['params'], device)

    # ---------------------------------------------------------- inference
    def scale(self, diag, rest, t, it, out, stream=None):
        """out = 1 + y(features of every amplitude) for iteration it."""
        gecco_hip.mlp_scale(self.params, diag, rest, t, it, out, self.work, stream)

    def predict(self, x):
        return gecco_hip.mlp_infer(self.params, x.contiguous())

    # ----------------------------------------------------------- training
    def train(self, data, steps=400, lr=2e-3, batch=65536, seed=0):
        """Adam on minibatches of the pairs data = (x, ut, dt) (device
        float32); returns the loss of every step (full-batch loss of the
        minibatch before the update)."""
        torch = self.torch
        x, ut, dt = data
        n = ut.numel()
        m = torch.zeros(P, dtype=torch.float32, device=self.device)
        v = torch.zeros_like(m)
        grad = torch.zeros_like(m)
        loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        work = torch.empty(gecco_hip.MLP_GRAD_WORK, dtype=torch.float32, device=self.device)
        g = torch.Generator(device='cpu').manual_seed(seed)
        losses = []
        for step in range(1, steps + 1):
            if n > batch:
                idx = torch.randint(0, n, (batch,), generator=g).to(self.device)
                xb, ub, db = x[idx].contiguous(), ut[idx].contiguous(), dt[idx].contiguous()
            else:
                xb, ub, db = x, ut, dt
            gecco_hip.mlp_grad(self.params, xb, ub, db, grad, loss, work)
            gecco_hip.mlp_adam(self.params, m, v, grad, step, lr)
            losses.append(loss)
            loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        return [float(l.item()) for l in losses]

    def loss(self, data):
        torch = self.torch
        x, ut, dt = data
        grad = torch.zeros(P, dtype=torch.float32, device=self.device)
        loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        work = torch.empty(gecco_hip.MLP_GRAD_WORK, dtype=torch.float32, device=self.device)
        gecco_hip.mlp_grad(self.params, x, ut, dt, grad, loss, work)
        return float(loss.item())


def collect(solver, maxit=60, tol=1e-9, per_iter=4096, ndiis=8, seed=0):
    """Training pairs from a plain (diagonal-preconditioner) DIIS solve of a
    ccsd.CCSD instance from T = 0: at every iteration k, per_iter random
    amplitudes with their features (iteration index k) and u~ = u / rms(u),
    dt~ = (t_(k+1) - t_k) / rms(u).  Returns (x, ut, dt) device float32
    tensors and the number of iterations."""
    import torch
    dev = solver.amp.device
    solver.amp.zero_()
    d = gecco_hip.DIIS(solver.amp.numel(), ndiis)
    n = solver.amp.numel()
    gen = torch.Generator(device='cpu').manual_seed(seed)
    xs, us, ds = [], [], []
    diag = solver.diag
    dv = torch.where(diag.abs() <= 1e-10, torch.full_like(diag, 1e-10), diag)
    k = 0
    for k in range(maxit):
        solver.evaluate()
        t = solver.amp.clone()
        u = (diag * t + solver.res) / dv
        ru = torch.sqrt(torch.mean(u * u))
        rt = torch.sqrt(torch.mean(t * t))
        _, _, conv = d.step(solver.amp, solver.res, diag, energy=solver.buf['E'], tol=tol)
        if conv:
            break
        idx = torch.randint(0, n, (min(per_iter, n),), generator=gen).to(dev)
        x = torch.empty((idx.numel(), 4), dtype=torch.float64, device=dev)
        x[:, 0] = diag[idx] / 4.0 - 1.0
        x[:, 1] = torch.asinh(u[idx] / ru)
        x[:, 2] = torch.asinh(t[idx] / rt) if float(rt) > 0 else 0.0
        x[:, 3] = min(k, 32) / 32.0
        xs.append(x.float())
        us.append((u[idx] / ru).float())
        ds.append(((solver.amp[idx] - t[idx]) / ru).float())
    return (torch.cat(xs).contiguous(), torch.cat(us).contiguous(), torch.cat(ds).contiguous()), k + 1
