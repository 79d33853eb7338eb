"""Oracle for the learned per-amplitude preconditioner (test infrastructure:
only tests/ and bench.py's cpu_baseline use it).

It has no counterpart in the reference. The reference preconditions with
the orbital-energy denominator (contract/dia4op_ev.f, baserout/diavc.f:31-35:
r = g / D, |D| <= 1e-10 -> 1e-10). Here a small network scales that step per
amplitude: r = s * g / D with s = 1 + y(x).

Network (float32 parameters, P = 17281, in this order):
    W1[128][4], b1[128], W2[128][128] (row = output), b2[128], w3[128], b3
    z1 = bf(W1) x + b1,  a = bf(silu(z1))
    z2 = bf(W2) a + b2,  h2 = silu(z2)
    y  = bf(w3) . h2 + b3
bf() rounds a float32 to bfloat16 (round to nearest, ties to even): the
weights of the GPU inference are bfloat16, its 128 x 128 layer runs on BF16
MFMA with float32 accumulation, so the first layer's activations are rounded
too. Training uses float32 master weights and passes gradients straight
through bf().

Features of amplitude i (float64, then float32), from the denominator D,
the residual g = D t + rest, the amplitude t and the iteration index it:
    u = g / D' (D' = D, or 1e-10 where |D| <= 1e-10)
    x = [D / 4 - 1, asinh(u / rms(u)), asinh(t / rms(t)) (0 if rms(t) = 0), min(it, 32) / 32]
rms over the whole amplitude vector.

Training pairs come from plain-solver runs: the accepted update dt = t_next
- t of a DIIS step against the plain step -u. With u~ = u / rms(u) and
dt~ = dt / rms(u) per sample, the loss is mean((1 + y) u~ + dt~)^2.

Pins (tests/test_oracle_mlp.py): bf() against hand-worked bfloat16
roundings; the backward pass against central finite differences of the
forward pass without rounding (bf = identity); the features against their
definition on a hand-checked vector.
"""
import numpy as np

NH = 128
NIN = 4
P = NH * NIN + NH + NH * NH + NH + NH + 1
OFF = {'W1': 0, 'b1': 512, 'W2': 640, 'b2': 17024, 'w3': 17152, 'b3': 17280}
DIA_THRESH = 1e-10


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[0:512].reshape(NH, NIN), p[512:640], p[640:17024].reshape(NH, NH), p[17024:17152],
            p[17152:17280], p[17280])


def bf(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases, and
    w3 = 0, so that the untrained network is the plain preconditioner."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P, dtype=np.float32)
    p[0:512] = rng.uniform(-0.5, 0.5, 512)
    p[640:17024] = rng.uniform(-1.0, 1.0, NH * NH) / np.sqrt(NH)
    return p


def forward(p, x, rounding=True):
    """y for features x (n x 4); also the intermediates for backward."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    x = np.asarray(x, dtype=np.float32).astype(np.float64)
    z1 = x @ r(W1).T + b1
    a = r(silu(z1))
    z2 = a @ r(W2).T + b2
    h2 = silu(z2)
    y = h2 @ r(w3) + b3
    return y, (x, z1, a, z2, h2)


def loss_grad(p, x, ut, dt, rounding=True):
    """Loss mean((1 + y) ut + dt)^2 and its gradient over the parameters
    (straight through bf())."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    y, (x, z1, a, z2, h2) = forward(p, x, rounding)
    ut = np.asarray(ut, dtype=np.float32).astype(np.float64)
    dt = np.asarray(dt, dtype=np.float32).astype(np.float64)
    n = len(y)
    e = (1.0 + y) * ut + dt
    loss = float(np.mean(e * e))
    dy = 2.0 * e * ut / n
    g = np.zeros(P)
    g[OFF['b3']] = dy.sum()
    g[OFF['w3']:OFF['w3'] + NH] = dy @ h2
    dz2 = (dy[:, None] * r(w3)[None, :]) * dsilu(z2)
    g[OFF['b2']:OFF['b2'] + NH] = dz2.sum(0)
    g[OFF['W2']:OFF['W2'] + NH * NH] = (dz2.T @ a).ravel()
    da = dz2 @ r(W2)
    dz1 = da * dsilu(z1)
    g[OFF['b1']:OFF['b1'] + NH] = dz1.sum(0)
    g[OFF['W1']:OFF['W1'] + NH * NIN] = (dz1.T @ x).ravel()
    return loss, g


def features(diag, rest, t, it):
    """Feature rows (n x 4, float32) and u, rms(u) of an amplitude vector."""
    diag = np.asarray(diag, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    g = diag * t + np.asarray(rest, dtype=np.float64)
    dv = np.where(np.abs(diag) <= DIA_THRESH, DIA_THRESH, diag)
    u = g / dv
    ru = np.sqrt(np.mean(u * u))
    rt = np.sqrt(np.mean(t * t))
    x = np.empty((len(u), 4))
    x[:, 0] = diag / 4.0 - 1.0
    x[:, 1] = np.arcsinh(u / ru) if ru > 0 else 0.0
    x[:, 2] = np.arcsinh(t / rt) if rt > 0 else 0.0
    x[:, 3] = min(it, 32) / 32.0
    return x.astype(np.float32), u, ru


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (float64 restatement of the device kernel)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
