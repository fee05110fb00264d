"""CPU oracle for the learned Sommerfeld surrogate -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the network the HIP kernels train and run
(momwire_amd/csrc/mw_mlp.hip): the same features, the same BF16 roundings of
weights, activations and back-propagated gradients (``rounding=True``), FP64
sums in place of the GPU's FP32 MFMA accumulation.  With ``rounding=False``
it is the plain float64 network, whose gradients ``fd_check`` pins against
central finite differences of the loss (the mathematics the backward pass
must reproduce).  No reference counterpart exists: the surrogate is new.
"""

import numpy as np

FIN = [64, 256, 256, 256, 256]
FOUT = [256, 256, 256, 256, 64]
N_USED = 8
NF_USED = 12


def n_params():
    return sum(o * i + o for i, o in zip(FIN, FOUT))


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def unpack(params):
    out, o = [], 0
    for i, f in zip(FIN, FOUT):
        W = params[o:o + f * i].reshape(f, i)
        o += f * i
        b = params[o:o + f]
        o += f
        out.append((W, b))
    return out


def features(x):
    """(n, 4) raw inputs (rho, h, eps_r, sigma/(omega eps0)) -> (n, 64), in
    float32 arithmetic like the GPU's."""
    f32 = np.float32
    x = np.asarray(x, dtype=f32)
    rho, h, er, s = x.T
    r1 = np.sqrt(rho * rho + h * h)
    th = np.arctan2(h, rho)
    f = np.zeros((x.shape[0], FIN[0]), dtype=f32)
    f[:, 0] = f32(0.5) * rho - f32(1.0)
    f[:, 1] = f32(0.5) * h - f32(1.0)
    f[:, 2] = (er - f32(17.5)) * f32(0.08)
    f[:, 3] = np.log(np.maximum(s, f32(1e-30)) * f32(0.0745355992)) * f32(0.321810032)
    f[:, 4] = r1 * f32(0.353553391) - f32(1.0)
    f[:, 5] = th * f32(1.27323954) - f32(1.0)
    f[:, 6] = np.exp(f32(-2.0) * r1)
    f[:, 7] = np.sin(th)
    mag = np.sqrt(np.sqrt(er * er + s * s))
    half = f32(0.5) * np.arctan2(-s, er)
    tp = f32(6.28318531)
    k1r, k1i = tp * mag * np.cos(half), tp * mag * np.sin(half)
    damp = np.exp(k1i * r1)
    beat = (k1r - tp) * r1
    f[:, 8] = np.cos(beat) * damp
    f[:, 9] = np.sin(beat) * damp
    f[:, 10] = np.exp(-tp * mag * r1)
    f[:, 11] = np.log(mag) * f32(0.5) - f32(1.0)
    return f


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, rounding=True):
    rnd = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float32 if rounding else np.float64))
    H = [rnd(features(x)).astype(np.float64)]
    Z = []
    for l, (W, b) in enumerate(layers):
        z = H[-1] @ rnd(W).astype(np.float64).T + b.astype(np.float64)
        if rounding:
            z = z.astype(np.float32).astype(np.float64)
        if l < len(layers) - 1:
            Z.append(z)
            H.append(rnd(_silu(z).astype(np.float32) if rounding else _silu(z)).astype(np.float64))
        else:
            out = z
    return dict(H=H, Z=Z, out=out)


def loss_and_grads(params, x, t, rounding=True):
    """Mean-square loss over the 8 used outputs and its gradient (flat)."""
    n = x.shape[0]
    fw = forward(params, x, rounding)
    layers = unpack(np.asarray(params, dtype=np.float64))
    d = fw["out"][:, :N_USED] - np.asarray(t, dtype=np.float64)
    sc = 1.0 / (n * N_USED)
    loss = float((d * d).sum() * sc)
    g = np.zeros((n, FOUT[-1]))
    if rounding:
        d32 = fw["out"][:, :N_USED].astype(np.float32) - np.asarray(t, dtype=np.float32)
        g[:, :N_USED] = bf16(np.float32(2.0) * d32 * np.float32(sc))
    else:
        g[:, :N_USED] = 2.0 * d * sc
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, _b = layers[l]
        Wr = bf16(W.astype(np.float32)).astype(np.float64) if rounding else W
        gW = g.T @ fw["H"][l]
        gb = g.sum(0)
        grads.append((gW, gb))
        if l > 0:
            c = g @ Wr
            if rounding:
                c = c.astype(np.float32).astype(np.float64)
            gz = c * _dsilu(fw["Z"][l - 1])
            g = bf16(gz.astype(np.float32)).astype(np.float64) if rounding else gz
    flat = []
    for gW, gb in reversed(grads):
        flat += [gW.ravel(), gb]
    return loss, np.concatenate(flat)


def fd_check(params, x, t, idx, h=1e-6):
    """Central differences of the float64 loss at parameter indices idx."""
    p = np.asarray(params, dtype=np.float64).copy()
    out = []
    for i in idx:
        p[i] += h
        lp = loss_and_grads(p, x, t, rounding=False)[0]
        p[i] -= 2 * h
        lm = loss_and_grads(p, x, t, rounding=False)[0]
        p[i] += h
        out.append((lp - lm) / (2 * h))
    return np.array(out)


def init_params(seed=0):
    """The surrogate's initialisation: uniform(+-sqrt(6/fan_in)) weights, zero
    biases, features/outputs beyond the used ones zeroed."""
    rng = np.random.default_rng(seed)
    flat = []
    for l, (i, f) in enumerate(zip(FIN, FOUT)):
        lim = np.sqrt(6.0 / (NF_USED if l == 0 else i))
        W = rng.uniform(-lim, lim, (f, i))
        if l == 0:
            W[:, NF_USED:] = 0.0
        if l == len(FIN) - 1:
            W *= 0.1
            W[N_USED:, :] = 0.0
        flat += [W.ravel(), np.zeros(f)]
    return np.concatenate(flat).astype(np.float32)
