"""FP64 restatement of the warm-start MLP -- TEST INFRASTRUCTURE ONLY.

Only tests/ (and bench.py's CPU leg, if ever) may import this module: it is
the test reference for the library's lats_mlp_* forward and backward kernels, never a
product path.  The network has no counterpart in the reference (whose frames
start from Core::predict_position, src/Solver/xpbd_constraints.h:321-384), so
this file restates the library's own definition: 48 features -> 128 -> 128 ->
128 -> 3, SiLU, loss = mean over the M x 3 outputs of (y - t)^2, parameters
laid out as libatssim_amd.warm_start.unpack_params describes.

With bf16=True the values the kernels round to BF16 are rounded the same way
(round to nearest even): the inputs, every weight, each hidden activation fed
to the next layer, and the output / pre-activation gradients fed to the
matrix-core products (dH = dZ W^T and the weight gradients dW = A^T dZ; the
bias gradients sum the unrounded dZ); everything else is FP64.  With bf16=False it is plain
FP64 calculus, which the finite-difference test pins.
"""
import numpy as np

HIDDEN = 128
NIN = 48
NOUT = 3


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _layers(p):
    o = 0
    out = []
    for fin, fout in ((64, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, 16)):
        W = np.asarray(p[o:o + fin * fout], np.float64).reshape(fin, fout)
        o += fin * fout
        b = np.asarray(p[o:o + fout], np.float64)
        o += fout
        out.append((W, b))
    return out, o


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, bf16_round=True):
    """Outputs (M x 3) and the cache of the backward pass."""
    layers, _ = _layers(p)
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    X = np.asarray(X, np.float64)
    A = np.zeros((len(X), 64))
    A[:, :NIN] = rnd(X) if bf16_round else X
    acts, zs = [A], []
    for li, (W, b) in enumerate(layers):
        Wq = rnd(W) if bf16_round else W
        z = b + acts[-1] @ Wq
        if li == 3:
            return z[:, :NOUT], (acts, zs, layers)
        zs.append(z)
        h = _silu(z)
        acts.append(rnd(h) if bf16_round else h)


def loss(p, X, T, bf16_round=True):
    Y, _ = forward(p, X, bf16_round)
    return float(np.mean((Y - np.asarray(T, np.float64)) ** 2))


def gradient(p, X, T, bf16_round=True):
    """(flat gradient in the parameter layout, loss)."""
    Y, (acts, zs, layers) = forward(p, X, bf16_round)
    T = np.asarray(T, np.float64)
    M = len(Y)
    dY = np.zeros((M, 16))
    dY[:, :NOUT] = 2.0 * (Y - T) / (3.0 * M)
    rnd = bf16 if bf16_round else (lambda a: a)
    grads = [None] * 4
    dz = dY
    for li in range(3, -1, -1):
        W, _ = layers[li]
        grads[li] = (acts[li].T @ rnd(dz), dz.sum(axis=0))
        if li == 0:
            break
        Wq = bf16(W) if bf16_round else W
        dh = rnd(dz) @ Wq.T
        dz = dh * _dsilu(zs[li - 1])
    flat = np.concatenate([np.concatenate([g.ravel(), gb]) for g, gb in grads])
    return flat, float(np.mean((Y - T) ** 2))
