"""Known-answer tests that pin the oracle's element-wise functions (CPU only).

Each test restates an assertion of the reference's own test suite (WaterLily.jl v1.8.0,
test/test_core.jl, test/test_flow.jl, test/test_bodies.jl) against the oracle, or checks the
oracle against what the mathematics fixes (a true median, a closed form).
"""
import math

import numpy as np
import pytest

import wl_oracle as O


# ---- test/test_flow.jl:3-41: convective schemes and boundary fluxes ----
def lam(s, u, c, d):
    return O.lib().wlo_scheme(s, u, c, d, O.WL_F64)


def flux(kind, f, I, ua, Ip=-1, s=O.WL_QUICK):
    arr = (O.C.c_double * len(f))(*f)
    return O.lib().wlo_phi_flux(kind, s, arr, I, Ip, ua)


def test_vanleer_cds():
    assert lam(O.WL_VANLEER, 1, 0, 1) == 0 and lam(O.WL_VANLEER, 1, 2, 1) == 2
    assert lam(O.WL_VANLEER, 1, 2, 3) == 2.5 and lam(O.WL_VANLEER, 3, 2, 1) == 1.5
    assert lam(O.WL_CDS, 1, 0, 1) == 0.5 and lam(O.WL_CDS, 1, 2, -1) == 0.5


def test_boundary_fluxes():
    f = [0.0, 0.5, 2.0]
    phi = lambda I: (f[I] + f[I - 1]) / 2
    # ϕuL at the inlet: positive flux -> central difference; negative -> backward QUICK
    assert flux(1, f, 1, 1.0) == phi(1)
    assert flux(1, f, 1, -1.0) == -lam(O.WL_QUICK, 2.0, 0.5, 0.0)
    # ϕuR at the outlet: positive -> standard QUICK; negative -> backward CD
    assert flux(2, f, 2, 1.0) == lam(O.WL_QUICK, 0.0, 0.5, 2.0)
    assert flux(2, f, 2, -1.0) == -phi(2)
    # ϕu equals ϕuP when given the same upstream index
    assert flux(0, f, 2, 1.0) == flux(3, f, 2, 1.0, Ip=0)
    assert flux(0, f, 1, -1.0) == flux(3, f, 1, -1.0, Ip=-1)
    # periodic flux
    g = [1.0, 1.25, 1.5, 1.75, 2.0]
    assert flux(3, g, 2, 1.0, Ip=0) == lam(O.WL_QUICK, g[0], g[1], g[2])
    assert flux(3, g, 2, 1.0, Ip=2) == lam(O.WL_QUICK, g[2], g[1], g[2])


def test_quick_is_limited_median():
    """quick(u,c,d) = median((5c+2d-u)/6, c, median(10c-9u, c, d)) with a true median."""
    rng = np.random.default_rng(0)
    for u, c, d in rng.normal(size=(2000, 3)):
        inner = np.median([10 * c - 9 * u, c, d])
        want = np.median([(5 * c + 2 * d - u) / 6, c, inner])
        assert lam(O.WL_QUICK, u, c, d) == want
    # ties exercise every branch of median
    for u, c, d in [(1, 1, 1), (0, 1, 1), (1, 0, 0), (2, 1, 2), (1, 2, 1)]:
        inner = np.median([10 * c - 9 * u, c, d])
        assert lam(O.WL_QUICK, u, c, d) == np.median([(5 * c + 2 * d - u) / 6, c, inner])
    # on a linear profile QUICK reproduces the linear interpolant (it is 3rd-order accurate)
    assert lam(O.WL_QUICK, 0.0, 1.0, 2.0) == pytest.approx(1.5)


# ---- test/test_bodies.jl:2-5: kernel moments ----
def kern(w, d, e=1.0):
    return O.lib().wlo_kern(w, d, e)


def test_kernel_moments():
    assert kern(3, 3.0, 6.0) == kern(3, 0.5, 1.0)
    assert kern(3, 0.0, 1.0) == 0.5
    assert kern(3, np.finfo(float).eps - 1, 1.0) == 0
    assert kern(4, 0.0, 2.0) == 2 * (1 / 4 - 1 / math.pi ** 2)
    # closed forms: kern0(d) = ∫_{-1}^d kern, kern1(d) = -∫_{-1}^d x·kern(x) dx
    for d in (-0.7, 0.0, 0.3, 1.0):
        x = np.linspace(-1, d, 4001)
        k = np.array([kern(0, v) for v in x])
        assert kern(1, d) == pytest.approx(np.trapezoid(k, x), abs=1e-6)
        assert kern(2, d) == pytest.approx(-np.trapezoid(x * k, x), abs=1e-6)


# ---- test/test_core.jl:19-69: boundary conditions ----
def test_bc_dirichlet_neumann():
    rng = np.random.default_rng(1)
    U = (1.0, 0.5)
    u = np.asfortranarray(rng.random((6, 6, 2)))
    O.bc(u, U)
    assert np.all(u[0, :, 0] == U[0]) and np.all(u[1, :, 0] == U[0]) and np.all(u[-1, :, 0] == U[0])
    assert np.all(u[2:-1, 0, 0] == u[2:-1, 1, 0]) and np.all(u[2:-1, -1, 0] == u[2:-1, -2, 0])
    assert np.all(u[:, 0, 1] == U[1]) and np.all(u[:, 1, 1] == U[1]) and np.all(u[:, -1, 1] == U[1])
    assert np.all(u[0, 2:-1, 1] == u[1, 2:-1, 1]) and np.all(u[-1, 2:-1, 1] == u[-2, 2:-1, 1])
    u[-1, :, 0] = 3
    O.bc(u, U, saveexit=True)
    assert np.all(u[-1, :, 0] == 3)
    O.exitbc(u, u, 0.0)  # conservative exit check
    assert np.all(u[-1, 1:-1, 0] == U[0])
    O.bc(u, U, saveexit=True, perdir=(2,))
    assert np.all(u[:, 0:2, 0] == u[:, -2:, 0])
    s = np.asfortranarray(rng.random((6, 6)))
    O.perbc(s, (1, 2))
    assert np.all(s[0, 1:-1] == s[-2, 1:-1]) and np.all(s[1:-1, 0] == s[1:-1, -2])
    u = np.asfortranarray(rng.random((6, 6, 2)))
    O.bc(u, U, saveexit=True, perdir=(1,))
    assert np.all(u[0:2, :, 0] == u[-2:, :, 0]) and np.all(u[0:2, :, 1] == u[-2:, :, 1])
    assert np.all(u[:, 0, 1] == U[1]) and np.all(u[:, 1, 1] == U[1]) and np.all(u[:, -1, 1] == U[1])


def test_bc_3d_order():
    """Corner ghosts follow the i-outer/j-inner write order (src/core.jl:203): the normal
    Dirichlet value wins wherever the normal index is 1, 2 or N; elsewhere a ghost is the copy of
    the nearest interior cell along every tangential direction."""
    rng = np.random.default_rng(2)
    U = (1.0, -0.25, 0.5)
    u = np.asfortranarray(rng.random((7, 6, 5, 3)))
    orig = u.copy()
    O.bc(u, U)
    N = u.shape[:3]
    for i in range(3):
        for I in np.ndindex(*N):
            if I[i] in (0, 1, N[i] - 1):
                assert u[I + (i,)] == U[i]
            else:
                J = tuple(min(max(I[d], 1), N[d] - 2) if d != i else I[d] for d in range(3))
                assert u[I + (i,)] == orig[J + (i,)]
