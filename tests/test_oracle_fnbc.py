"""Function-valued boundary conditions and forcing on the oracle (F4): BC! with a function
(test/test_core.jl:35-44), the periodic Taylor-Green vortex (test/test_flow.jl:100-109), the
boundary-layer profile (:133-140) and the rotating reference frame (:142-159).  The functions are
evaluated on the host on every cell (tests/fn_fields.py) and handed to the oracle as fields."""
import math

import numpy as np
import pytest

import wl_oracle as O
from fn_fields import field, step_times

L2 = lambda a: float(np.sum(np.square(a[tuple(slice(1, -1) for _ in range(a.ndim))], dtype=np.float64)))


def test_bc_function_kat():
    """test/test_core.jl:35-44: a function uBC equal to a tuple gives the tuple's normal ghosts;
    the exit values survive saveexit."""
    Ng = (10, 8)
    U = (1.0, 0.5)
    rng = np.random.default_rng(0)
    v = np.asfortranarray(rng.random(Ng + (2,)))
    u = np.asfortranarray(rng.random(Ng + (2,)))
    UB = field(lambda i, x, t: np.full(x[0].shape, 1.0 if i == 1 else 0.5), Ng, np.float64, 0.0)
    O.bc_fn(v, UB)
    O.bc(u, U)
    for k in (0, 1, -1):
        assert np.array_equal(v[k, :, 0], u[k, :, 0]) and np.array_equal(v[:, k, 1], u[:, k, 1])
    v[-1, :, 0] = 3
    O.bc_fn(v, UB, saveexit=True)
    assert np.all(v[-1, :, 0] == 3)


@pytest.mark.parametrize("D", [2, 3])
def test_bc_function_reproduces_its_values(D):
    """With the interior equal to an integer-valued uBC field, every ghost equals uBC: Dirichlet
    ghosts directly, tangential ghosts through (uBC(I) + a[I±δ]) − uBC(I±δ) exactly.  A wrong
    neighbour offset or sign breaks it."""
    Ng = (8, 6, 5)[:D]
    UB = field(lambda i, x, t: sum((d + 2 * i) * x[d] for d in range(D)) * 2, Ng, np.float64, 0.0)
    a = np.asfortranarray(np.random.default_rng(1).random(Ng + (D,)))
    inner = tuple(slice(1, -1) for _ in range(D))
    a[inner] = UB[inner]
    O.bc_fn(a, UB)
    assert np.array_equal(a, UB)


def tgv(i, x, t, k, nu):
    """TGVsim's vortex (test/helper.jl:4-16), in the precision of x."""
    X, Y = x[0] * k, x[1] * k
    e = np.exp(-2 * k * k * nu * t)
    return -np.sin(X) * np.cos(Y) * e if i == 1 else np.cos(X) * np.sin(Y) * e


def tgv_dt(i, x, t, k, nu):
    return -2 * k * k * nu * tgv(i, x, t, k, nu)


def test_tgv_periodic():
    """test/test_flow.jl:100-109: the Taylor-Green vortex on 64², periodic in x and y, Re 1e8,
    stepped to sim_time π/100: L₂ of each component's error against the exact decay < 1e-4."""
    T = np.float32
    L = 64
    k = T(2 * math.pi / L)
    nu = T(1 / (k * T(1e8)))
    Ng = (L + 2, L + 2)
    fl = O.OFlow(O.make_desc((L, L), (0, 0), nu=float(nu), dtype=O.WL_F32, perdir=(1, 2)))
    f = lambda i, x, t: tgv(i, x, t, k, nu)
    fd = lambda i, x, t: tgv_dt(i, x, t, k, nu)
    fl.set("u", field(f, Ng, T, 0.0))
    fl.set_field_fn(0, field(f, Ng, T, 0.0))
    fl.apply_ic()
    fl.pois_update()
    while fl.time() / L < math.pi / 100:
        dts = fl.dt
        t0, t1 = step_times(dts, T)
        fl.set_field_fn(0, field(f, Ng, T, t1))
        fl.set_field_fn(1, field(fd, Ng, T, t0))
        fl.set_field_fn(2, field(fd, Ng, T, t1))
        fl.mom_step()
    ue = field(f, Ng, T, T(np.sum(np.asarray(fl.dt[:-1], dtype=T), dtype=T)))
    u = fl.get("u")
    assert L2(u[..., 0] - ue[..., 0]) < 1e-4 and L2(u[..., 1] - ue[..., 1]) < 1e-4


def bl_profile(L):
    return lambda i, x, t: (4.0 * ((x[1] + 0.5) / (2 * L) - ((x[1] + 0.5) / (2 * L)) ** 2)).astype(np.float32) \
        if i == 1 else np.zeros_like(x[1])


def test_boundary_layer_profile():
    """test/test_flow.jl:133-140 (make_bl_flow, test/helper.jl:18-22): a laminar profile set as a
    function uBC on 32², ν = 0.001, stepped to sim_time 10: the exit column equals the inlet column."""
    T = np.float32
    L = 32
    Ng = (L + 2, L + 2)
    fl = O.OFlow(O.make_desc((L, L), (0, 0), nu=0.001, dtype=O.WL_F32))
    f = bl_profile(L)
    UB = field(f, Ng, T, 0.0)
    fl.set("u", UB)
    fl.set_field_fn(0, UB)  # time-independent: ∂uBC/∂t = 0, the same field every step
    fl.apply_ic()
    fl.pois_update()
    while fl.time() / L < 10:
        fl.mom_step()
    u = fl.get("u")
    assert np.allclose(u[0, :, 0], u[-1, :, 0], rtol=math.sqrt(np.finfo(T).eps), atol=0)


def rotating(L, T):
    w = T(1 / L)
    x0 = (T(L), T(L))

    def velocity(i, x, t):
        s, c = np.sin(w * t), np.cos(w * t)
        y1, y2 = w * (x[0] - x0[0]), w * (x[1] - x0[1])
        return s * y1 + c * y2 if i == 1 else -c * y1 + s * y2

    def dvel(i, x, t):
        s, c = np.sin(w * t), np.cos(w * t)
        y1, y2 = w * (x[0] - x0[0]), w * (x[1] - x0[1])
        return w * (c * y1 - s * y2) if i == 1 else w * (s * y1 + c * y2)

    def g(i, x, t):
        cor = 2 * w * velocity(2, x, t) if i == 1 else -2 * w * velocity(1, x, t)
        return cor + w * w * (x[i - 1] - x0[i - 1])

    return velocity, dvel, g


def test_rotating_reference_frame():
    """test/test_flow.jl:142-159: a solid-body rotation seen from the rotating frame (function
    uBC plus Coriolis and centrifugal forcing), Float64 on 8²: the pressure stays ≈ 0 (L₂ < 3e-3)."""
    T = np.float64
    Lh = 4
    N = 2 * Lh
    Ng = (N + 2, N + 2)
    velocity, dvel, g = rotating(Lh, T)
    fl = O.OFlow(O.make_desc((N, N), (0, 0), nu=0.0, dtype=O.WL_F64))
    fl.set("u", field(velocity, Ng, T, 0.0))
    fl.set_field_fn(0, field(velocity, Ng, T, 0.0))
    fl.apply_ic()
    fl.pois_update()
    t0, t1 = step_times(fl.dt, T)
    acc = lambda t: field(lambda i, x, tt: g(i, x, tt) + dvel(i, x, tt), Ng, T, t)
    fl.set_field_fn(0, field(velocity, Ng, T, t1))
    fl.set_field_fn(1, acc(t0))
    fl.set_field_fn(2, acc(t1))
    fl.mom_step()
    assert L2(fl.get("p")) < 3e-3
