"""Closed-form pins of the oracle's CFL and BDIM! (CPU only), and the reference's scheme-selection test.

Each check is fixed by the mathematics on inputs whose arithmetic is exact in the element type,
not by a restatement of the oracle's formula.  A planted mistake in the test's own numpy copy of
the function (a wrong viscous term, the lower face read for the upper one, a lost ghost, a
flipped or transposed μ₁ term) is shown to fail the same checks, so they have teeth.
"""
import numpy as np
import pytest

import wl_oracle as O

T32 = np.float32


def flow(n, nu=0.0, dtype=O.WL_F32, **kw):
    return O.OFlow(O.make_desc(n, (0.0,) * len(n), nu=nu, dtype=dtype, **kw))


# ---------------------------------------------------------------------------------------------
# CFL (src/Flow.jl:234-244): Δt = min(10, 1/(maximum(σ) + 5ν)), σ = flux_out(u) on inside, the
# maximum over the whole σ (ghosts keep their previous values).
# ---------------------------------------------------------------------------------------------
def cfl_cases(Ng):
    """(u, σ-ghost value or None, ν, expected Δt) with expected values from exact arithmetic."""
    D = len(Ng)
    cases = []
    # 1. uniform u = U: flux_out = Σ|Uᵢ| (dyadic, exact), 5ν rounded in T as the reference does
    U = (0.5, -0.25, 0.125)[:D]
    u = np.zeros(Ng + (D,), T32, order="F")
    for i in range(D):
        u[..., i] = U[i]
    nu = 0.1
    cases.append((u, None, nu, T32(1) / (T32(sum(abs(v) for v in U)) + T32(5) * T32(nu))))
    # 2. a ramp u₁ = x (cell index) along x: the largest outflow of an inside cell is its upper
    #    face u₁[I+e₁] at the last inside cell, i.e. the value on the top ghost face (Ng₀ − 1)
    u = np.zeros(Ng + (D,), T32, order="F")
    ramp = np.arange(Ng[0], dtype=T32) / T32(8)
    u[..., 0] = ramp.reshape((-1,) + (1,) * (D - 1))
    cases.append((u, None, 0.0, min(T32(10), T32(1) / ramp[Ng[0] - 1])))
    # 3. a reversed ramp (u₁ < 0): the outflow is −u₁[I] at the lowest inside cell (index 1)
    u = np.zeros(Ng + (D,), T32, order="F")
    u[..., 0] = -ramp[::-1].reshape((-1,) + (1,) * (D - 1))
    cases.append((u, None, 0.0, min(T32(10), T32(1) / ramp[::-1][1])))
    # 4. a stale ghost σ larger than every inside flux wins the maximum (hazard: max over all σ)
    u = np.zeros(Ng + (D,), T32, order="F")
    u[..., 0] = T32(0.25)
    cases.append((u, T32(2.0), 0.0, T32(1) / T32(2.0)))
    # 5. tiny velocities: Δt is capped at 10
    u = np.zeros(Ng + (D,), T32, order="F")
    u[..., 0] = T32(2.0 ** -10)
    cases.append((u, None, 0.0, T32(10)))
    return cases


def oracle_cfl(Ng, u, ghost, nu):
    f = flow(tuple(v - 2 for v in Ng), nu=nu)
    f.set("u", u)
    s = np.zeros(Ng, T32, order="F")
    if ghost is not None:
        s[(-1,) + (0,) * (len(Ng) - 1)] = ghost  # a corner ghost of σ
    f.set("sigma", s)
    return T32(f.cfl())


def numpy_cfl(Ng, u, ghost, nu, plant=None):
    """The test's own copy of CFL, with an optional planted mistake."""
    D = len(Ng)
    inner = tuple(slice(1, n - 1) for n in Ng)
    sig = np.zeros(Ng, T32)
    if ghost is not None:
        sig[(-1,) + (0,) * (D - 1)] = ghost
    s = np.zeros(tuple(n - 2 for n in Ng), T32)
    for i in range(D):
        up = tuple(slice(2, n) if d == i else slice(1, n - 1) for d, n in enumerate(Ng))
        hi = u[up + (i,)] if plant != "lower_face" else u[inner + (i,)]
        s = s + (np.maximum(T32(0), hi) + np.maximum(T32(0), -u[inner + (i,)]))
    sig[inner] = s
    m = sig[inner].max() if plant == "inside_only" else sig.max()
    visc = T32(4 if plant == "4nu" else 5) * T32(nu)
    return min(T32(10), T32(1) / (m + visc))


@pytest.mark.parametrize("Ng", [(18, 10), (18, 10, 6)])
def test_cfl_closed_forms(Ng):
    for k, (u, ghost, nu, want) in enumerate(cfl_cases(Ng)):
        got = oracle_cfl(Ng, u, ghost, nu)
        assert got == want, (k, got, want)


@pytest.mark.parametrize("plant", ["4nu", "lower_face", "inside_only"])
def test_cfl_closed_forms_catch_planted_mistakes(plant):
    Ng = (18, 10, 6)
    cases = cfl_cases(Ng)
    assert all(numpy_cfl(Ng, u, g, nu) == w for u, g, nu, w in cases)  # the unplanted copy passes
    assert any(numpy_cfl(Ng, u, g, nu, plant) != w for u, g, nu, w in cases), plant


# ---------------------------------------------------------------------------------------------
# BDIM! (src/Flow.jl:176-180, μddn :20-26): f = u⁰ + Δt·f − V on every cell, then on inside
# u[I,i] += ½Σⱼ μ₁[I,i,j](f[I+eⱼ,i] − f[I−eⱼ,i]) + V[I,i] + μ₀[I,i]·f[I,i].
# With f linear, fᵢ = Σₖ Gᵢₖ xₖ, constant μ₁ = m, μ₀ = a, V = v, u⁰ = b and u = c:
# u = c + Δt·Σⱼ mᵢⱼ·Gᵢⱼ + v + a·(b + Δt·Gᵢ·x − v), all dyadic (exact in Float32).
# ---------------------------------------------------------------------------------------------
G = np.array([[1.0, -2.0, 3.0], [0.5, 1.0, -1.0], [2.0, 0.25, -0.5]])
M = np.array([[0.25, -0.5, 0.125], [0.75, 0.5, -0.25], [-0.125, 1.0, 0.5]])  # not symmetric


def bdim_inputs(Ng):
    D = len(Ng)
    idx = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in Ng], indexing="ij")
    f = np.zeros(Ng + (D,), order="F")
    for i in range(D):
        f[..., i] = sum(G[i, k] * idx[k] for k in range(D))
    mu1 = np.zeros(Ng + (D, D), order="F")
    for i in range(D):
        for j in range(D):
            mu1[..., i, j] = M[i, j]
    return f, mu1, idx


def bdim_expected(Ng, dt, a, b, c, v, plant=None):
    D = len(Ng)
    f, _, idx = bdim_inputs(Ng)
    u = np.full(Ng + (D,), c, order="F")
    for i in range(D):
        fi = b + dt * f[..., i] - v
        s = 0.0
        for j in range(D):
            mij = M[j, i] if plant == "transpose" else M[i, j]
            sg = -1.0 if plant == "sign" else 1.0
            s += sg * mij * dt * G[i, j] * (1.0 if plant == "half" else 2.0)
        inner = tuple(slice(1, n - 1) for n in Ng)
        u[inner + (i,)] = c + s / 2 + v + a * fi[inner]
    return u


@pytest.mark.parametrize("Ng", [(10, 8), (10, 8, 6)])
@pytest.mark.parametrize("dtype", [O.WL_F32, O.WL_F64])
def test_bdim_mu1_closed_form(Ng, dtype):
    D = len(Ng)
    dt, a, b, c, v = 0.5, 0.25, 1.5, -2.0, 0.125
    fl = flow(tuple(n - 2 for n in Ng), dtype=dtype)
    f, mu1, _ = bdim_inputs(Ng)
    fl.set("f", f)
    fl.set("mu1", mu1)
    fl.set("mu0", np.full(Ng + (D,), a))
    fl.set("V", np.full(Ng + (D,), v))
    fl.set("u0", np.full(Ng + (D,), b))
    fl.set("u", np.full(Ng + (D,), c))
    fl.phase(1, dt=dt)
    got = fl.get("u").astype(np.float64)
    want = bdim_expected(Ng, dt, a, b, c, v)
    assert np.array_equal(got, want)
    for plant in ("sign", "transpose", "half"):
        assert not np.array_equal(got, bdim_expected(Ng, dt, a, b, c, v, plant)), plant
    # the first loop runs on every cell, ghosts included: f = u⁰ + Δt·f − V
    assert np.array_equal(fl.get("f").astype(np.float64), b + dt * f - v)


# ---------------------------------------------------------------------------------------------
# test/test_flow.jl:87-98: the scheme λ chosen at construction is the one used.  Periodic 16²,
# Float64, u₁ = sinpi(x₁/8) from the initial-condition keyword; quick and cds differ by > 1e-6
# after one sim_step!.
# ---------------------------------------------------------------------------------------------
def nonuniform_ic(Ng, dtype):
    """apply!((i,x) -> i==1 ? sinpi(x[1]/8) : 0, u) at loc(i,I) = I − 1.5 − ½eᵢ (1-based I)."""
    D = len(Ng)
    u = np.zeros(Ng + (D,), dtype, order="F")
    x1 = np.arange(1, Ng[0] + 1, dtype=np.float64) - 1.5 - 0.5
    u[..., 0] = np.sin(np.pi * x1 / 8).reshape((-1,) + (1,) * (D - 1))
    return u


def scheme_sim(scheme):
    s = O.OSimulation((16, 16), (1.0, 0.0), 16, U=1.0, dtype=O.WL_F64, perdir=(1, 2), scheme=scheme)
    s.flow.set("u", nonuniform_ic((18, 18), np.float64))
    s.flow.apply_ic()
    return s


def test_scheme_selection_quick_vs_cds():
    sq, sc = scheme_sim(O.WL_QUICK), scheme_sim(O.WL_CDS)
    assert np.array_equal(sq.flow.get("u"), sc.flow.get("u"))
    sq.step()
    sc.step()
    assert np.max(np.abs(sq.flow.get("u") - sc.flow.get("u"))) > 1e-6
    sv = scheme_sim(O.WL_VANLEER)
    sv.step()
    assert np.max(np.abs(sv.flow.get("u") - sq.flow.get("u"))) > 1e-6
