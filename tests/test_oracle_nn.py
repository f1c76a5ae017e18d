"""The warm-start network's CPU oracle (oracle/wl_nn_oracle.cpp), pinned independently of its own
formulas:
- one convolution (no rounding) against scipy.ndimage.correlate with zero padding, per (ci, co);
- the skip input against an explicit nearest upsampling (numpy repeat);
- the bf16 rounding against torch's float32 → bfloat16 conversion (round to nearest even);
- the whole backward pass (no rounding, double) against central finite differences of the loss
  along random directions, one per layer;
- one Adam step against torch.optim.Adam;
- positive homogeneity of the bias-free ReLU network (exact, with bf16 rounding).
"""
import numpy as np
import pytest
import scipy.ndimage as ndi
import torch

import wl_oracle as O


def _rand_w(rng, scale=1.0):
    ws = []
    for ci, co in O.NN_LAYERS:
        ws.append((rng.standard_normal((27, ci, co)) * scale * np.sqrt(2.0 / (27 * ci))).astype(np.float32))
    return ws


def _flat(ws):
    return np.concatenate([w.ravel() for w in ws]).astype(np.float32)


def _to_zyx(a, grid, C):
    nx, ny, nz = grid
    return a.reshape(nz, ny, nx, C)


@pytest.mark.parametrize("ci,co", [(8, 16), (16, 32)])
def test_conv_matches_scipy_correlate(ci, co):
    rng = np.random.default_rng(0)
    grid = (6, 5, 4)
    V = 6 * 5 * 4
    a = rng.standard_normal((V, ci)).astype(np.float32)
    w = rng.standard_normal((27, ci, co)).astype(np.float32)
    out = O.nn_conv(grid, w, a, relu=False, round=False, outf32=True)
    az = _to_zyx(a.astype(np.float64), grid, ci)
    ref = np.zeros((4, 5, 6, co))
    for c_in in range(ci):
        for c_out in range(co):
            k = w[:, c_in, c_out].astype(np.float64).reshape(3, 3, 3)  # [dz][dy][dx]
            ref[..., c_out] += ndi.correlate(az[..., c_in], k, mode="constant", cval=0.0)
    np.testing.assert_allclose(out.reshape(4, 5, 6, co), ref, rtol=1e-5, atol=1e-4)


def test_skip_input_is_nearest_upsampling():
    rng = np.random.default_rng(1)
    grid, C = (4, 4, 2), 16
    V = 32
    a = rng.standard_normal((V, C)).astype(np.float32)
    b = rng.standard_normal((V // 8, C)).astype(np.float32)
    w = np.zeros((27, C, 16), np.float32)
    for c in range(C):
        w[13, c, c] = 1.0  # centre tap, identity
    out = O.nn_conv(grid, w, a, b=b, relu=False, round=False, outf32=True)
    up = _to_zyx(b, (2, 2, 1), C).repeat(2, 0).repeat(2, 1).repeat(2, 2)
    np.testing.assert_allclose(out.reshape(2, 4, 4, 16), _to_zyx(a, grid, C) + up, rtol=1e-6, atol=1e-6)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-8, 8, 2000),
                        np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0])]).astype(np.float32)
    V = x.size // 8 * 8
    a = x[:V].reshape(-1, 8)
    w = np.zeros((27, 8, 16), np.float32)
    for c in range(8):
        w[13, c, c] = 1.0
    out = O.nn_conv((V // 8, 1, 1), w, a, relu=False, round=True)  # bf16 in, identity, bf16 out
    ref = torch.from_numpy(a.copy()).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(out[:, :8], ref)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    grid = (8, 8, 8)
    V = 512
    ws = _rand_w(rng)
    in0 = np.zeros((V, 8), np.float32)
    in0[:, :3] = rng.standard_normal((V, 3))
    tgt = rng.standard_normal(V).astype(np.float32)
    w = _flat(ws)
    L, g = O.nn_gradient(grid, w, in0, tgt, round=False)
    off = 0
    for l, (ci, co) in enumerate(O.NN_LAYERS):
        n = 27 * ci * co
        d = np.zeros(w.size)
        d[off:off + n] = rng.standard_normal(n)
        if l == 0:
            d[off:off + n].reshape(27, ci, co)[:, 3:, :] = 0  # padded input channels carry no signal
        # a small step keeps ReLU kinks from being crossed (at 1e-3 they add ~1% noise on the
        # full-resolution layers); the exact difference of the float32 weights used is the direction
        h = 1e-5 / np.linalg.norm(d[off:off + n])
        wp = (w + h * d).astype(np.float32)
        wm = (w - h * d).astype(np.float32)
        Lp = O.nn_gradient(grid, wp, in0, tgt, round=False, grad=False)
        Lm = O.nn_gradient(grid, wm, in0, tgt, round=False, grad=False)
        fd = (Lp - Lm)
        an = float(np.dot(g, wp.astype(np.float64) - wm.astype(np.float64)))
        rel = abs(fd - an) / max(abs(an), 1e-12)
        assert rel < 1e-5, (l, fd, an)
        off += n
    assert 0.0 < L


def test_adam_matches_torch():
    rng = np.random.default_rng(4)
    n = 1000
    w = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    tw = torch.nn.Parameter(torch.from_numpy(w.copy()))
    opt = torch.optim.Adam([tw], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 6):
        g = rng.standard_normal(n).astype(np.float32)
        O.nn_adam(w, g, m, v, 1e-3, t)
        tw.grad = torch.from_numpy(g.copy())
        opt.step()
        np.testing.assert_allclose(w, tw.detach().numpy(), rtol=1e-6, atol=1e-7)


def test_forward_positively_homogeneous():
    rng = np.random.default_rng(5)
    grid = (16, 8, 8)
    V = 16 * 8 * 8
    w = _flat(_rand_w(rng))
    in0 = np.zeros((V, 8), np.float32)
    in0[:, :3] = rng.standard_normal((V, 3))
    y1 = O.nn_forward(grid, w, in0, round=True)
    y2 = O.nn_forward(grid, w, 2 * in0, round=True)
    np.testing.assert_array_equal(y2, 2 * y1)
    assert np.any(y1 != 0)
