"""Learned warm start of the pressure solves (no counterpart in the reference).

The reference starts each pressure solve from the previous pressure, x = p·Δt (mom_project!,
src/Flow.jl:225), and its multigrid loop (src/MultiLevelPoisson.jl:108-127) iterates from there to
Σr² < tol.  `WarmNet` adds a learned guess of the increment before that loop: a bias-free 3D U-Net
(3×3×3 convolutions, 16/32/64/64 channels, bf16 weights and activations, FP32 accumulation on the
MFMA units) on the multigrid's level-1 grid, trained on the device with Adam from pairs the solver
records.  The loop afterwards is unchanged, so every projection still ends at the same tolerance.

    net = WarmNet.for_flow(sim.flow, replay_pairs=256)
    set_warmstart(sim, net, guess=False, record=True)   # harvest pairs while stepping
    ...; net.train(steps=500, lr=1e-3)
    set_warmstart(sim, net, guess=True)                 # use the guess
"""
import ctypes as C

import numpy as np

from . import _abi as A

LAYERS = [(8, 16), (16, 16), (16, 32), (32, 64), (64, 64), (64, 32), (32, 16), (16, 16), (16, 8)]
REAL_IN = 3  # input channels carrying data (the rest is padding)


def net_grid(flow):
    """The network grid of a 3D flow: its interior halved in every direction."""
    if flow.D != 3:
        raise AssertionError("the warm-start network is three-dimensional")
    return tuple(n // 2 for n in flow.N)


def init_weights(seed=0, out_scale=0.1):
    """He-normal weights for the ReLU layers (padding channels zero); the last layer is scaled by
    `out_scale` so the initial guess is small."""
    rng = np.random.default_rng(seed)
    ws = []
    for l, (ci, co) in enumerate(LAYERS):
        cin = REAL_IN if l == 0 else ci
        w = np.zeros((27, ci, co), np.float32)
        w[:, :cin, :] = rng.standard_normal((27, cin, co)) * np.sqrt(2.0 / (27 * cin))
        if l == len(LAYERS) - 1:
            w[:] = 0
            w[:, :, 0] = rng.standard_normal((27, ci)) * np.sqrt(1.0 / (27 * ci)) * out_scale
        ws.append(w.astype(np.float32))
    return np.concatenate([w.ravel() for w in ws])


class WarmNet:
    def __init__(self, grid, replay_pairs=64, weights=None, seed=0):
        self.grid = tuple(int(g) for g in grid)
        self.V = self.grid[0] * self.grid[1] * self.grid[2]
        self.h = C.c_void_p()
        A.check(A.load().wl_net_create(C.byref(self.h), (C.c_int32 * 3)(*self.grid), int(replay_pairs)))
        n = C.c_int64()
        A.check(A.load().wl_net_num_weights(self.h, C.byref(n)))
        self.nweights = n.value
        self.weights = init_weights(seed) if weights is None else weights

    @classmethod
    def for_flow(cls, flow, **kw):
        return cls(net_grid(flow), **kw)

    @property
    def weights(self):
        w = np.zeros(self.nweights, np.float32)
        A.check(A.load().wl_net_get_weights(self.h, w.ctypes.data, self.nweights))
        return w

    @weights.setter
    def weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.nweights:
            raise AssertionError(f"expected {self.nweights} weights")
        A.check(A.load().wl_net_set_weights(self.h, w.ctypes.data, self.nweights))

    def _in(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.shape != (self.V, 8):
            raise AssertionError(f"input must have shape ({self.V}, 8)")
        return x

    def forward(self, x):
        """The network on an input [V][8] (rounded to bf16): the prediction [V]."""
        x = self._in(x)
        out = np.zeros(self.V, np.float32)
        A.check(A.load().wl_net_forward(self.h, x.ctypes.data, out.ctypes.data))
        return out

    def gradient(self, x, target):
        """(loss, canonical weight gradient) of the mean-free relative loss on one pair."""
        x = self._in(x)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.nweights, np.float32)
        L = C.c_double()
        A.check(A.load().wl_net_gradient(self.h, x.ctypes.data, t.ctypes.data, C.byref(L), g.ctypes.data))
        return L.value, g

    def adam(self, lr):
        A.check(A.load().wl_net_adam(self.h, float(lr)))

    def train(self, steps, lr=1e-3, seed=0):
        """`steps` Adam steps on recorded pairs drawn at random; the mean loss before the updates."""
        L = C.c_double()
        A.check(A.load().wl_net_train(self.h, int(steps), float(lr), int(seed), C.byref(L)))
        return L.value

    @property
    def pairs(self):
        n = C.c_int()
        A.check(A.load().wl_net_pairs(self.h, C.byref(n)))
        return n.value

    def pair(self, index):
        x = np.zeros((self.V, 8), np.float32)
        t = np.zeros(self.V, np.float32)
        A.check(A.load().wl_net_get_pair(self.h, int(index), x.ctypes.data, t.ctypes.data))
        return x, t

    def save(self, path):
        np.savez(path, weights=self.weights, grid=np.asarray(self.grid))

    @classmethod
    def load(cls, path, grid=None, **kw):
        """Weights from `save`; the network grid may differ from the one trained on (the network is
        convolutional)."""
        with np.load(path, allow_pickle=False) as z:
            w = z["weights"]
            g = tuple(int(v) for v in z["grid"]) if grid is None else grid
        return cls(g, weights=w, **kw)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                A.load().wl_net_destroy(self.h)
                self.h = None
        except Exception:
            pass


def set_warmstart(sim_or_flow, net, guess=True, record=False, stats=False, predictor_only=False):
    """Attach `net` to a flow (None detaches): guess = add its guess before each solve, record =
    store (input, target) pairs, stats = keep Σr² before/after the guess, predictor_only = only in
    the predictor's projection."""
    flow = getattr(sim_or_flow, "flow", sim_or_flow)
    flags = ((A.WL_WARM_GUESS if guess else 0) | (A.WL_WARM_RECORD if record else 0) |
             (A.WL_WARM_STATS if stats else 0) | (A.WL_WARM_PREDICTOR if predictor_only else 0))
    A.check(A.load().wl_flow_set_warmstart(flow.h, net.h if net is not None else None, flags))
    flow._warmnet = net  # keep the network alive while attached


def warm_stats(sim_or_flow):
    """(n, 2) array of Σr² before and after the guess, one row per projection since attach."""
    flow = getattr(sim_or_flow, "flow", sim_or_flow)
    n = C.c_int()
    A.check(A.load().wl_flow_warm_stats(flow.h, None, 0, C.byref(n)))
    buf = (C.c_double * max(n.value, 1))()
    A.check(A.load().wl_flow_warm_stats(flow.h, buf, n.value, C.byref(n)))
    return np.asarray(buf[:n.value], np.float64).reshape(-1, 2)


# ---------------------------------------------------------------------------------------------
# Solver-generated training data: flows past 1–3 random ellipsoids (semi-axes U[8,24] cells at
# 256×128×128, centres uniform in the middle half of the domain, Re U[500,5000], ν = 32/Re)
# ---------------------------------------------------------------------------------------------
def random_ellipsoid_simulation(seed, n=(256, 128, 128), T=np.float32):
    from .api import Ellipsoid, Simulation
    rng = np.random.default_rng(seed)
    s = n[1] / 128
    body = None
    for _ in range(int(rng.integers(1, 4))):
        axes = tuple(rng.uniform(8, 24, 3) * s)
        c = tuple(rng.uniform(n[d] / 4, 3 * n[d] / 4) for d in range(3))
        b = Ellipsoid(c, axes)
        body = b if body is None else body + b
    Re = rng.uniform(500, 5000)
    return Simulation(n, (1, 0, 0), 32 * s, ν=32 * s / Re, body=body, T=T)


def train_on_flows(net, seeds, n=(256, 128, 128), warmup=20, steps=160, lr=(1e-3, 1e-5), predictor_only=False,
                   log=None):
    """Online training: for each seed, step a random-ellipsoid flow `warmup` steps, then `steps`
    steps recording pairs, then one Adam step per recorded pair drawn from the replay store; the
    learning rate decays geometrically from lr[0] to lr[1] over the seeds.  Returns the mean
    losses per seed."""
    from .api import sim_step
    losses = []
    k = len(seeds)
    for i, seed in enumerate(seeds):
        sim = random_ellipsoid_simulation(seed, n)
        for _ in range(warmup):
            sim_step(sim, remeasure=False)
        set_warmstart(sim, net, guess=False, record=True, predictor_only=predictor_only)
        for _ in range(steps):
            sim_step(sim, remeasure=False)
        set_warmstart(sim, None)
        rate = lr[0] * (lr[1] / lr[0]) ** (i / max(k - 1, 1))
        ntrain = steps * (1 if predictor_only else 2)
        losses.append(net.train(ntrain, lr=rate, seed=seed))
        if log:
            log(f"seed {seed}: {ntrain} pairs, lr {rate:.2e}, mean loss {losses[-1]:.4f}")
    return losses
