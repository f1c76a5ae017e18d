"""The warm-start network's MFMA kernels through the C ABI against the CPU oracle
(oracle/wl_nn_oracle.cpp, pinned in test_oracle_nn.py) on the same weights and inputs.

Both sides round to bf16 at the same points; they differ only in the order of the FP32 (device
MFMA) versus double (oracle) convolution sums, which can move a stored bf16 value by one unit in
the last place and, rarely, flip a ReLU at zero.  Tolerances: forward output relative L2 ≤ 5e-3;
loss relative 1e-3; weight gradient relative L2 ≤ 3e-2 per layer; one Adam step to 1e-6."""
import sys

import numpy as np
import pytest

import wl_oracle as O
from __graft_entry__ import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wl():
    m = load_package()
    m.load_library()
    return m


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _case(wl, grid, seed):
    rng = np.random.default_rng(seed)
    V = grid[0] * grid[1] * grid[2]
    w = wl.init_weights(seed, out_scale=1.0)
    x = np.zeros((V, 8), np.float32)
    x[:, :3] = rng.standard_normal((V, 3))
    return w, x, rng


# grids below 2^16 voxels run the one-tile-per-wave convolution; (64, 32, 32) the 4-tile one
@pytest.mark.parametrize("grid", [(16, 16, 8), (24, 8, 16), (64, 32, 32)])
def test_forward_matches_oracle(wl, grid):
    say(f"case: network forward, grid {grid}")
    w, x, _ = _case(wl, grid, 1)
    net = wl.WarmNet(grid, replay_pairs=1, weights=w)
    g = net.forward(x)
    o = O.nn_forward(grid, w, x, round=True)
    e = rel(g, o)
    say(f"  rel-L2 {e:.2e}")
    assert e <= 5e-3 and np.linalg.norm(o) > 0


@pytest.mark.parametrize("grid", [(16, 16, 8), (24, 8, 16), (64, 32, 32)])
def test_gradient_matches_oracle(wl, grid):
    say(f"case: network gradient, grid {grid}")
    w, x, rng = _case(wl, grid, 2)
    t = rng.standard_normal(x.shape[0]).astype(np.float32)
    net = wl.WarmNet(grid, replay_pairs=1, weights=w)
    Lg, gg = net.gradient(x, t)
    Lo, go = O.nn_gradient(grid, w, x, t, round=True)
    say(f"  loss {Lg:.6f} oracle {Lo:.6f}")
    assert abs(Lg - Lo) <= 1e-3 * abs(Lo)
    off = 0
    for l, (ci, co) in enumerate(O.NN_LAYERS):
        n = 27 * ci * co
        e = rel(gg[off:off + n], go[off:off + n])
        say(f"  layer {l + 1} ({ci}->{co}) grad rel-L2 {e:.2e}")
        assert e <= 3e-2, l
        off += n


def test_adam_step_matches_oracle(wl):
    say("case: one Adam step")
    grid = (16, 8, 8)
    w, x, rng = _case(wl, grid, 3)
    t = rng.standard_normal(x.shape[0]).astype(np.float32)
    net = wl.WarmNet(grid, replay_pairs=1, weights=w)
    _, g = net.gradient(x, t)
    net.adam(1e-3)
    w1 = net.weights
    wo = w.copy()
    m = np.zeros_like(wo)
    v = np.zeros_like(wo)
    O.nn_adam(wo, g, m, v, 1e-3, 1)
    assert np.max(np.abs(w1 - wo)) <= 1e-6


def test_gradient_descent_fits_a_realisable_target(wl):
    """Adam on one pair whose target is another network's output: the loss falls."""
    say("case: fit one pair")
    grid = (16, 16, 8)
    w, x, _ = _case(wl, grid, 4)
    t = O.nn_forward(grid, wl.init_weights(7, out_scale=1.0), x, round=True)
    net = wl.WarmNet(grid, replay_pairs=1, weights=w)
    L0, _ = net.gradient(x, t)
    for _ in range(60):
        net.gradient(x, t)
        net.adam(3e-3)
    L1, _ = net.gradient(x, t)
    say(f"  loss {L0:.4f} -> {L1:.4f}")
    assert L1 < 0.5 * L0
