"""CPU FP64 restatement of the learned preconditioner's network (test
infrastructure, like the rest of oracle/: only tests and the benchmark's
checks use it).

Restates, for checking the GPU kernels:
  features  per-point geometry features (the GPU's k_lp_features)
  forward   4-layer ReLU network, BF16 operands rounded where the GPU rounds
            them (weights, the standardized input, each layer's activations)
  backward  the loss and every parameter gradient of one batch, with the GPU's
            roundings (dY and each dZ to BF16) or none (bf16=False)
The backward pass itself is pinned by central finite differences of the loss
(FP64, no rounding) in tests/test_oracle.py.

Parameter layout (one flat FP32 vector, as the GPU stores it): W1 (Fp x H,
row-major), b1 (H), W2 (H x H), b2, W3 (H x H), b3, W4 (H x Op), b4 (Op).
"""
import numpy as np


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def features(X, Y, nx, ny, ds, dx, I0x, I0y, offsets, K):
    """(N, 6(2K+1)+3): per neighbour j along the body's curve (cyclic), the
    offset in grid units in the point's (tangent, normal) frame, sin/cos of
    2 pi times the neighbour's fractional primal-grid coordinates; then nx,
    ny, ds/dx."""
    N = len(X)
    W = 2 * K + 1
    nbr = neighbours(offsets, N, K)
    x, y = np.asarray(X) / dx, np.asarray(Y) / dx
    ddx, ddy = x[nbr] - x[:, None], y[nbr] - y[:, None]
    nxc, nyc = np.asarray(nx)[:, None], np.asarray(ny)[:, None]
    ft = -ddx * nyc + ddy * nxc
    fn = ddx * nxc + ddy * nyc
    fx = x[nbr] - (1.0 - I0x)
    fy = y[nbr] - (1.0 - I0y)
    fx -= np.floor(fx)
    fy -= np.floor(fy)
    two_pi = 2 * np.pi
    F = np.concatenate([ft, fn, np.sin(two_pi * fx), np.cos(two_pi * fx), np.sin(two_pi * fy), np.cos(two_pi * fy),
                        nxc, nyc, (np.asarray(ds) / dx)[:, None]], axis=1)
    assert F.shape == (N, 6 * W + 3)
    return F


def neighbours(offsets, N, K):
    off = list(offsets) if offsets is not None else [0, N]
    W = 2 * K + 1
    nbr = np.zeros((N, W), dtype=np.int64)
    for b in range(len(off) - 1):
        s, m = off[b], off[b + 1] - off[b]
        a = np.arange(m)
        nbr[s:s + m] = s + (a[:, None] + np.arange(-K, K + 1)[None, :]) % m
    return nbr


def unpack(params, dims):
    K, W, F, Fp, H, Op, n = dims
    p = np.asarray(params, dtype=np.float64)
    shapes = [(Fp, H), (H,), (H, H), (H,), (H, H), (H,), (H, Op), (Op,)]
    out, o = [], 0
    for sh in shapes:
        sz = int(np.prod(sh))
        out.append(p[o:o + sz].reshape(sh))
        o += sz
    assert o == n
    return out


def pack(parts):
    return np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in parts])


def loss_grad(params, dims, X, T, bf16_round=True):
    """Loss mean((Y - T)^2 over the 2K+1 real outputs) and the gradient of
    every parameter, for standardized inputs X (rows x Fp) and targets T
    (rows x Op)."""
    K, W, Fp, H, Op = dims[0], dims[1], dims[3], dims[4], dims[5]
    r = (lambda a: bf16(a)) if bf16_round else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params, dims)
    Ws = [r(W1), r(W2), r(W3), r(W4)]
    bs = [b1, b2, b3, b4]
    X = np.asarray(X, dtype=np.float64)
    acts = [r(X)]
    for l in range(3):
        acts.append(r(np.maximum(acts[l] @ Ws[l] + bs[l], 0.0)))
    Y = acts[3] @ Ws[3] + bs[3]
    rows = X.shape[0]
    D = np.zeros_like(Y)
    D[:, :W] = Y[:, :W] - T[:, :W]
    loss = float(np.sum(D[:, :W] ** 2) / (rows * W))
    dZ = r(2.0 * D / (rows * W))
    grads = [None] * 8
    for l in range(3, -1, -1):
        grads[2 * l] = acts[l].T @ dZ
        grads[2 * l + 1] = dZ.sum(axis=0)
        if l > 0:
            dA = dZ @ Ws[l].T
            dZ = r(np.where(acts[l] > 0, dA, 0.0))
    return loss, pack(grads)


def predict(params, dims, Xstd):
    """Standardized prediction (rows x Op) of BF16-rounded inputs."""
    K, W, Fp, H, Op = dims[0], dims[1], dims[3], dims[4], dims[5]
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params, dims)
    a = bf16(Xstd)
    for Wl, bl in ((W1, b1), (W2, b2), (W3, b3)):
        a = bf16(np.maximum(a @ bf16(Wl) + bl, 0.0))
    return a @ bf16(W4) + b4
