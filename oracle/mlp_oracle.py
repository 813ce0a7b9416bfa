"""ORACLE -- test infrastructure, not product code.

CPU (numpy, float64) restatement of the surrogate MLP of vmc_amd/csrc/afq_mlp.hip:
F -> H -> H -> H -> 1 with exact-erf GELU on the hidden layers, standardised
inputs (x - mu)/sd, loss 1/2 mean((y - (t - t_mu)/t_sd)^2).  Every GEMM operand
is rounded to bfloat16 (nearest-even through float32) exactly where the GPU
rounds it when staging in LDS; accumulation here is float64 (the GPU's is
float32), so the two agree to float32-accumulation error, except for the rare
operand that sits on a bf16 rounding boundary.
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


_bf16 = bf16


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(params, dims):
    """flattened [W0 b0 W1 b1 W2 b2 W3 b3 mu sd t_mu t_sd] -> lists"""
    off = 0
    Ws, bs = [], []
    for l in range(4):
        n = dims[l] * dims[l + 1]
        Ws.append(np.asarray(params[off:off + n], dtype=np.float64).reshape(dims[l], dims[l + 1]))
        off += n
        bs.append(np.asarray(params[off:off + dims[l + 1]], dtype=np.float64))
        off += dims[l + 1]
    F = dims[0]
    mu = np.asarray(params[off:off + F], dtype=np.float64)
    sd = np.asarray(params[off + F:off + 2 * F], dtype=np.float64)
    tmu, tsd = float(params[off + 2 * F]), float(params[off + 2 * F + 1])
    return Ws, bs, mu, sd, tmu, tsd


def forward(params, dims, X, rnd=None):
    """rnd: operand rounding of the GEMMs (default bf16, as the GPU; identity for
    the finite-difference check of the algebra)"""
    bf16 = _bf16 if rnd is None else rnd
    Ws, bs, mu, sd, tmu, tsd = unpack(params, dims)
    xin = (np.asarray(X, np.float32) - mu.astype(np.float32)) / sd.astype(np.float32)
    ins, zs = [], []
    h = xin.astype(np.float64)
    for l in range(4):
        ins.append(h)
        z = bf16(h) @ bf16(Ws[l]) + bs[l]
        zs.append(z)
        h = gelu(z) if l < 3 else z
    return h[:, 0], ins, zs


def loss_and_grads(params, dims, X, t, rnd=None):
    bf16 = _bf16 if rnd is None else rnd
    Ws, bs, mu, sd, tmu, tsd = unpack(params, dims)
    y, ins, zs = forward(params, dims, X, rnd)
    B = len(y)
    ts = (np.asarray(t, np.float64) - tmu) / tsd
    d = y - ts
    loss = 0.5 * np.mean(d * d)
    dh = (d / B)[:, None]
    grads = [None] * 8
    for l in range(3, -1, -1):
        dz = dh if l == 3 else dh * gelu_d(zs[l])
        grads[2 * l] = bf16(ins[l]).T @ bf16(dz)
        grads[2 * l + 1] = dz.sum(axis=0)
        if l > 0:
            dh = bf16(dz) @ bf16(Ws[l]).T
    flat = np.concatenate([g.ravel() for g in grads])
    return loss, flat
