"""ORACLE - test infrastructure only.  FP64 restatement of the scheduler's
16-64-64-5 ReLU MLP (forward, masked-MSE loss, backward), used to check the
GPU kernels of x-gminer_amd/csrc/mlp.hip.  The GPU feeds BF16 operands to
the matrix cores; this reference rounds the same tensors to BF16 at the same
points (inputs, weights, post-ReLU activations, upstream gradients) and
computes everything else in FP64, so the remaining difference is the FP32
accumulation of the GPU.  No reference-repo counterpart (the reference has
no learned scheduler)."""
import numpy as np

IN, H, OUT = 16, 64, 5
SIZES = [("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, OUT)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def _ident(x):
    return np.asarray(x, dtype=np.float64)


def forward(p, X, rounding=True):
    r = bf16 if rounding else _ident
    w = unpack(p)
    xb = r(X)
    z1 = xb @ r(w["W1"]) + w["b1"]
    a1 = r(np.maximum(z1, 0))
    z2 = a1 @ r(w["W2"]) + w["b2"]
    a2 = r(np.maximum(z2, 0))
    y = a2 @ r(w["W3"]) + w["b3"]
    return dict(xb=xb, z1=z1, a1=a1, z2=z2, a2=a2, y=y, w=w)


def loss(p, X, T, M, rounding=True):
    d = (forward(p, X, rounding)["y"] - T) * M
    return float(np.sum(d * d) / max(1.0, float(np.sum(M))))


def loss_and_grads(p, X, T, M, rounding=True):
    """loss = sum M*(Y-T)^2 / sum M; returns (loss, flat gradient).
    rounding=False gives the exact FP64 model (for finite differences)."""
    bf16_ = bf16 if rounding else _ident
    f = forward(p, X, rounding)
    w = f["w"]
    msum = max(1.0, float(np.sum(M)))
    d = (f["y"] - T) * M
    loss = float(np.sum(d * d) / msum)
    dy = 2 * d / msum
    dyb = bf16_(dy)
    g = {}
    g["W3"] = f["a2"].T @ dyb
    g["b3"] = dy.sum(0)
    da2 = dyb @ bf16_(w["W3"]).T
    dz2 = np.where(f["z2"] > 0, da2, 0.0)
    dz2b = bf16_(dz2)
    g["W2"] = f["a1"].T @ dz2b
    g["b2"] = dz2.sum(0)
    da1 = dz2b @ bf16_(w["W2"]).T
    dz1 = np.where(f["z1"] > 0, da1, 0.0)
    dz1b = bf16_(dz1)
    g["W1"] = f["xb"].T @ dz1b
    g["b1"] = dz1.sum(0)
    flat = np.concatenate([g[name].reshape(-1) for name, _ in SIZES])
    return loss, flat
