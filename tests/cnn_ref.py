"""CPU reference (numpy, float64) of the learned-guess network of
02_cathode-08_modulerev_amd/csrc/learned.hip: forward, loss and gradients,
with the device's BF16 rounding points emulated (features, weights, hidden
activations and activation gradients are rounded to BF16; every sum is
float64).  Test infrastructure: the checker of the device kernels."""
import numpy as np

CIN = [16, 32, 32, 32, 32, 32]
COUT = [32, 32, 32, 32, 32, 16]
NFEAT = 3
TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def layout():
    off, o = [], 0
    for cin, cout in zip(CIN, COUT):
        w = o
        o += cout * 9 * cin
        off.append((w, o))
        o += cout
    return off, o


def unpack(params):
    """[(W[co][tap][ci], b[co])] float64"""
    off, _ = layout()
    out = []
    for (w, b), cin, cout in zip(off, CIN, COUT):
        out.append((params[w:w + cout * 9 * cin].astype(np.float64).reshape(cout, 9, cin),
                    params[b:b + cout].astype(np.float64)))
    return out


def pack(layers):
    off, n = layout()
    p = np.zeros(n, dtype=np.float64)
    for (w, b), (W, B) in zip(off, layers):
        p[w:w + W.size] = W.ravel()
        p[b:b + B.size] = B
    return p


def bf16(a, on=True):
    """round to the nearest BF16 (ties to even), as float64"""
    if not on:
        return np.asarray(a, dtype=np.float64)
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def conv(X, W, ni, nj):
    """X: (ni+2, nj+2, cin) zero ghosts; W: (cout, 9, cin) -> interior (ni, nj, cout)"""
    out = np.zeros((ni, nj, W.shape[0]))
    for t, (di, dj) in enumerate(TAPS):
        Xs = X[1 + di:ni + 1 + di, 1 + dj:nj + 1 + dj, :]
        out += Xs @ W[:, t, :].T
    return out


def embed(Y, ni, nj):
    Z = np.zeros((ni + 2, nj + 2, Y.shape[2]))
    Z[1:-1, 1:-1] = Y
    return Z


def forward(params, feat, ni, nj, rnd=True):
    """feat: (3, ni+2, nj+2).  Returns (y plane (ni+2, nj+2), activations)."""
    L = unpack(params)
    X = np.zeros((ni + 2, nj + 2, 16))
    X[1:-1, 1:-1, :NFEAT] = np.moveaxis(feat[:, 1:-1, 1:-1], 0, -1)
    acts = [bf16(X, rnd)]
    for l, (W, b) in enumerate(L):
        Wr = bf16(W, rnd)
        pre = conv(acts[-1], Wr, ni, nj) + b
        if l < len(L) - 1:
            acts.append(bf16(embed(np.maximum(pre, 0.0), ni, nj), rnd))
        else:
            y = np.zeros((ni + 2, nj + 2))
            y[1:-1, 1:-1] = pre[:, :, 0]
    return y, acts


def loss_and_grad(params, feat, target, ni, nj, rnd=True):
    """mean squared error over interior cells and its gradient (float64)."""
    L = unpack(params)
    y, acts = forward(params, feat, ni, nj, rnd)
    N = ni * nj
    d = (y - target)[1:-1, 1:-1]
    loss = float((d * d).sum() / N)
    dY = np.zeros((ni, nj, 16))
    dY[:, :, 0] = 2.0 * d / N
    dY = bf16(dY, rnd)
    grads = [None] * len(L)
    for l in range(len(L) - 1, -1, -1):
        W, b = L[l]
        X = acts[l]
        gW = np.zeros_like(W)
        for t, (di, dj) in enumerate(TAPS):
            Xs = X[1 + di:ni + 1 + di, 1 + dj:nj + 1 + dj, :]
            gW[:, t, :] = np.einsum("ijo,ijc->oc", dY, Xs)
        gb = dY.sum(axis=(0, 1))
        grads[l] = (gW, gb)
        if l == 0:
            break
        # gradient into the layer input, through the previous ReLU
        Wr = bf16(W, rnd)
        dYe = embed(dY, ni, nj)
        dX = np.zeros((ni, nj, W.shape[2]))
        for t, (di, dj) in enumerate(TAPS):
            # X[p + off] contributed to Y[p]: dX[q] += dY[q - off] W[:, t, :]
            dX += dYe[1 - di:ni + 1 - di, 1 - dj:nj + 1 - dj, :] @ Wr[:, t, :]
        dY = bf16(dX * (X[1:-1, 1:-1, :] > 0), rnd)
    return loss, pack(grads)
