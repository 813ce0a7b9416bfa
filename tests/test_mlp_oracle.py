"""The surrogate's CPU oracle (oracle/mlp_oracle.py) pinned without the GPU:
its backward pass against central finite differences of its forward pass (with
operand rounding switched off, so the loss is smooth), GELU and its derivative
against the closed form and finite differences, and bf16 rounding against the
float32 bit pattern rule (round-to-nearest-even on the 16 dropped bits)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def ident(x):
    return np.asarray(x, np.float64)


def test_gelu_and_derivative():
    z = np.linspace(-6, 6, 241)
    assert abs(M.gelu(0.0)) < 1e-15
    assert np.allclose(M.gelu(np.array([8.0, -8.0])), [8.0, 0.0], atol=1e-12)
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_d(z))) < 1e-8


def test_bf16_rounding_rule():
    x = np.array([1.0 + 2 ** -7, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], np.float32)
    r = M.bf16(x)
    # 7 fraction bits: 1 + 2^-7 is exact; 1 + 2^-8 is a tie -> even (1.0);
    # 1 + 3*2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6 -> even (1 + 2^-6); 1 + 2^-9 -> 1.0
    assert r[0] == 1.0 + 2 ** -7 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0
    assert r[4] == -2.5
    rng = np.random.default_rng(0)
    y = rng.standard_normal(1000).astype(np.float32)
    ry = M.bf16(y)
    assert np.all(np.abs(ry - y) <= np.abs(y) * 2 ** -8)
    assert np.all((ry.astype(np.float32).view(np.uint32) & 0xFFFF) == 0)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    F, H, B = 7, 5, 11
    dims = [F, H, H, H, 1]
    nw = sum(dims[l] * dims[l + 1] + dims[l + 1] for l in range(4))
    p = np.concatenate([rng.normal(0, 0.5, nw), rng.normal(0, 0.2, F), rng.uniform(0.5, 2, F), [1.5, 2.0]])
    X = rng.standard_normal((B, F))
    t = rng.normal(1.0, 2.0, B)
    loss, g = M.loss_and_grads(p, dims, X, t, rnd=ident)
    assert g.shape == (nw,)
    h = 1e-6
    for k in rng.choice(nw, 60, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.loss_and_grads(pp, dims, X, t, rnd=ident)[0] - M.loss_and_grads(pm, dims, X, t, rnd=ident)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
