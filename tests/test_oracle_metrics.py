"""Oracle force/moment integrals and MeanFlow blend (src/Metrics.jl:116-261) against the
reference's known-answer tests (test/test_metrics.jl:35-104) and closed forms of the divergence
theorem, ∮ φ n dA = ∫ ∇φ dV, on the BDIM-smoothed surface."""
import math

import numpy as np
import pytest

import wl_oracle as O


def loc(i, idx, D):
    """loc(i,I) of src/core.jl:177 for a 0-based index tuple: I − 1.5 − ½eᵢ (1-based I)."""
    return np.array([idx[d] + 1 - 1.5 - (0.5 if d == i - 1 else 0.0) for d in range(D)])


def apply_vec(f, Ng, T=np.float64):
    """apply!((i,x)->f(i,x), u) over every cell including ghosts (src/util.jl apply!)."""
    D = len(Ng)
    u = np.zeros(tuple(Ng) + (D,), dtype=T, order="F")
    for i in range(1, D + 1):
        for idx in np.ndindex(*Ng):
            u[idx + (i - 1,)] = f(i, loc(i, idx, D))
    return u


def apply_scalar(f, Ng, T=np.float64):
    p = np.zeros(tuple(Ng), dtype=T, order="F")
    for idx in np.ndindex(*Ng):
        p[idx] = f(loc(0, idx, len(Ng)))
    return p


N = 32
R = N // 4


def body(D):
    return [{"kind": 1, "center": (N / 2,) * D, "radius": R}]


def vol(D):
    return math.pi * R**2 if D == 2 else 4 / 3 * math.pi * R**3


def test_pressure_force_hydrostatic():
    """test/test_metrics.jl:35-40: p = y inside, ∮ p n = πR²·[0,1] within 2e-3."""
    p = np.zeros((N, N), order="F")
    p[1:-1, 1:-1] = apply_scalar(lambda x: x[1], (N, N))[1:-1, 1:-1]
    F = O.metrics(p, None, 1.0, body(2))[0]
    assert np.sum(np.abs(F / (math.pi * R**2) - [0, 1])) < 2e-3


@pytest.mark.parametrize("D", [2, 3])
def test_strain_kats(D):
    """test/test_metrics.jl:41-53: S of zero, of a uniform gradient and of a shear."""
    Ng = (N,) * D
    I = (N // 2,) * D
    assert np.all(O.strain(np.zeros(Ng + (D,), order="F"), I) == 0)
    u = apply_vec(lambda i, x: x[i - 1], Ng)
    assert np.allclose(2 * O.strain(u, I), 2 * np.eye(D))
    u = apply_vec(lambda i, x: x[i % D], Ng)  # x[i%D+1] in 1-based indexing
    want = np.array([[0, 2], [2, 0]]) if D == 2 else np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]])
    assert np.allclose(2 * O.strain(u, I), want)


@pytest.mark.parametrize("D", [2, 3])
def test_zero_strain_forces(D):
    """test/test_metrics.jl:54-61 (u = 0) and a rigid rotation, whose strain vanishes: no viscous
    force or moment."""
    Ng = (N,) * D
    c = N / 2
    for u in (np.zeros(Ng + (D,), order="F"),
              apply_vec(lambda i, x: -(x[1] - c) if i == 1 else ((x[0] - c) if i == 2 else 0.0), Ng)):
        _, Fv, _, Mv = O.metrics(None, u, 1.0, body(D), x0=(c,) * D)
        assert np.max(np.abs(Fv)) < 1e-10 and np.max(np.abs(Mv)) < 1e-9


@pytest.mark.parametrize("D", [2, 3])
def test_pressure_moment(D):
    """test/test_metrics.jl:62-66 (no moment of a hydrostatic p about the centre), and about
    x₀ = c − a·e₁: M = (c − x₀) × ∫∇p dV = a·V·e_z."""
    Ng = (N,) * D
    p = apply_scalar(lambda x: x[1], Ng)
    c = N / 2
    Mp = O.metrics(p, None, 1.0, body(D), x0=(c,) * D)[2]
    assert np.max(np.abs(Mp)) < 1e-9 * vol(D) * N
    a = 3.0
    Mp = O.metrics(p, None, 1.0, body(D), x0=(c - a,) + (c,) * (D - 1))[2]
    want = a * vol(D)
    got = Mp[0] if D == 2 else Mp[2]
    assert abs(got - want) / want < 5e-3
    if D == 3:
        assert abs(Mp[0]) < 1e-6 * want and abs(Mp[1]) < 1e-6 * want


@pytest.mark.parametrize("D", [2, 3])
def test_viscous_force_closed_form(D):
    """u = (y², 0, ...): S_xy = y, ∇·S = e_x, so F_v = −2ν∫∇·S dV = −2ν·V·e_x; about x₀ = c − b·e_y
    the viscous moment is −2ν(c − x₀)×(V e_x) = 2ν·b·V·e_z."""
    Ng = (N,) * D
    nu = 0.7
    u = apply_vec(lambda i, x: x[1] ** 2 if i == 1 else 0.0, Ng)
    c = N / 2
    b = 2.0
    _, Fv, _, Mv = O.metrics(None, u, nu, body(D), x0=(c, c - b) + (c,) * (D - 2))
    V = vol(D)
    assert abs(Fv[0] + 2 * nu * V) / (2 * nu * V) < 5e-3
    assert np.max(np.abs(Fv[1:])) < 1e-6 * V
    got = Mv[0] if D == 2 else Mv[2]
    assert abs(got - 2 * nu * b * V) / (2 * nu * b * V) < 5e-3


def test_planted_sign_error_fails():
    """The closed form above detects a flipped strain sign (planted in the test's own copy)."""
    Ng = (N, N)
    u = apply_vec(lambda i, x: x[1] ** 2 if i == 1 else 0.0, Ng)
    Fv = O.metrics(None, -u, 1.0, body(2))[1]
    assert abs(Fv[0] + 2 * vol(2)) / (2 * vol(2)) > 1.5


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_meanflow_blend_is_time_average(T):
    """update!(meanflow, flow) (src/Metrics.jl:229-241): after updates at times t₁ < t₂ < …, with
    ε = 1 first, the means are the dt-weighted averages Σ dtₖ fₖ / Σ dtₖ (dt₁ = t₁ − t_init)."""
    rng = np.random.default_rng(0)
    n, D = 50, 2
    P = np.zeros(n, T)
    U = np.zeros(n * D, T)
    UU = np.zeros(n * D * D, T)
    t = [T(0)]
    times = np.cumsum(rng.uniform(0.1, 0.5, size=6)).astype(T)
    fs = [rng.standard_normal(n).astype(T) for _ in times]
    us = [rng.standard_normal(n * D).astype(T) for _ in times]
    for tk, f, u in zip(times, fs, us):
        dt = T(tk) - t[-1]
        e = dt / (dt + (t[-1] - t[0]) + T(np.finfo(T).eps))
        if len(t) == 1:
            e = T(1)
        O.meanflow_update(e, P, f, U, u, UU)
        t.append(T(t[-1] + dt))
    w = np.diff(np.concatenate([[0.0], times.astype(np.float64)]))
    Pw = sum(wk * f.astype(np.float64) for wk, f in zip(w, fs)) / w.sum()
    tol = 1e-4 if T == np.float32 else 1e-12
    assert np.max(np.abs(P - Pw)) < tol
    uu = [np.stack([u.reshape(D, n)[i] * u.reshape(D, n)[j] for j in range(D) for i in range(D)]) for u in us]
    UUw = sum(wk * q.astype(np.float64) for wk, q in zip(w, uu)) / w.sum()
    assert np.max(np.abs(UU.reshape(D * D, n) - UUw)) < 10 * tol
