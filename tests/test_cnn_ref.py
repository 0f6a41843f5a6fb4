"""The CPU reference of the learned-guess network (tests/cnn_ref.py) is
itself pinned: without BF16 rounding its gradient matches central finite
differences of its loss, and BF16 rounding matches numpy's float32 -> bfloat16
round-to-nearest-even on known bit patterns."""
import numpy as np

import cnn_ref as R


def _params(rng, scale=0.3):
    _, n = R.layout()
    p = np.zeros(n)
    layers = R.unpack(p)
    out = []
    for l, (W, b) in enumerate(layers):
        W = rng.standard_normal(W.shape) * scale
        if l == 0:
            W[:, :, R.NFEAT:] = 0.0
        if l == len(layers) - 1:
            W[1:] = 0.0
        out.append((W, rng.standard_normal(b.shape) * 0.1))
    return R.pack(out)


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    ni, nj = 6, 5
    p = _params(rng)
    feat = rng.standard_normal((3, ni + 2, nj + 2))
    tgt = rng.standard_normal((ni + 2, nj + 2))
    loss, g = R.loss_and_grad(p, feat, tgt, ni, nj, rnd=False)
    off, _ = R.layout()
    idx = []
    for (w, b), cin, cout in zip(off, R.CIN, R.COUT):
        idx += [w + 0, w + 9 * cin + 1, b]  # two weights and a bias per layer
    for k in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_and_grad(pp, feat, tgt, ni, nj, rnd=False)[0] -
              R.loss_and_grad(pm, feat, tgt, ni, nj, rnd=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 2e-5 * max(1e-3, abs(fd)), (k, fd, g[k])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-30], dtype=np.float64)
    r = R.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie -> even
    assert r[2] == 1.0 + 2 ** -6                # tie -> even (up)
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-30) / 3.0e-30 < 2 ** -8
