"""Host-side mirror of WaterLily.jl's time-step API over the gfx950 C ABI.

Names and argument meanings follow the reference (WaterLily.jl v1.8.0): `Flow`, `mom_step`
(mom_step!), `MultiLevelPoisson`, `Poisson`, `solver` (solver!), `mult` (mult!), `update`
(update!), `Simulation`, `sim_step` (sim_step!), `sim_time`, `measure` (measure!) and
`pressure_force`.  Arrays cross the boundary in Julia's layout: column-major numpy arrays of
shape (Nx, Ny[, Nz][, D[, D]]), ghosts included.  All compute happens in libwaterlily_hip.so.
"""
import ctypes as C
import math

import numpy as np

from . import _abi as A

quick, vanLeer, cds = A.WL_QUICK, A.WL_VANLEER, A.WL_CDS


def _dtype_code(T):
    T = np.dtype(T)
    if T == np.float32:
        return A.WL_F32
    if T == np.float64:
        return A.WL_F64
    raise AssertionError(f"element type {T} is not supported (Float32 or Float64)")


def _perdir_mask(perdir):
    return sum(1 << (int(j) - 1) for j in perdir)


# ---------------------------------------------------------------------------------------------
# Bodies: analytic AutoBody SDFs (src/AutoBody.jl), unions as SetBody(min,...) (src/Body.jl:91-103)
# ---------------------------------------------------------------------------------------------
class AbstractBody:
    def parts(self):
        raise NotImplementedError

    def __add__(self, other):
        return BodyUnion(self.parts() + other.parts())

    __or__ = __add__


class Sphere(AbstractBody):
    """AutoBody((x,t)->√sum(abs2,x.-center)-radius); a circle in 2D."""

    def __init__(self, center, radius):
        self.center, self.radius = tuple(float(c) for c in center), float(radius)

    def parts(self):
        return [self]

    def sdf(self, x):
        return math.sqrt(sum((a - b) ** 2 for a, b in zip(x, self.center))) - self.radius


Circle = Sphere


class Ellipsoid(AbstractBody):
    """AutoBody((x,t)->(√sum(abs2,(x.-center)./axes)-1)*minimum(axes))."""

    def __init__(self, center, axes):
        self.center, self.axes = tuple(float(c) for c in center), tuple(float(a) for a in axes)

    def parts(self):
        return [self]

    def sdf(self, x):
        q = math.sqrt(sum(((a - b) / s) ** 2 for a, b, s in zip(x, self.center, self.axes)))
        return (q - 1) * min(self.axes)


class BodyUnion(AbstractBody):
    def __init__(self, bodies):
        if not 1 <= len(bodies) <= 4:
            raise AssertionError("between 1 and 4 bodies are supported")
        self.bodies = list(bodies)

    def parts(self):
        return list(self.bodies)


class NoBody(AbstractBody):
    def parts(self):
        return []


def _body_array(body):
    parts = body.parts()
    arr = (A.WlBody * max(1, len(parts)))()
    for k, b in enumerate(parts):
        if isinstance(b, Sphere):
            arr[k].kind = A.WL_BODY_SPHERE
            arr[k].radius = b.radius
            for q, c in enumerate(b.center):
                arr[k].center[q] = c
        else:
            arr[k].kind = A.WL_BODY_ELLIPSOID
            for q, (c, a) in enumerate(zip(b.center, b.axes)):
                arr[k].center[q] = c
                arr[k].axes[q] = a
    return arr, len(parts)


# ---------------------------------------------------------------------------------------------
# Flow (src/Flow.jl:114-148)
# ---------------------------------------------------------------------------------------------
def _loc_coords(Ng, i, T):
    """Coordinates of loc(i,I) = I − 1.5 − ½eᵢ (src/core.jl:177, 1-based I) on every cell, in T."""
    idx = np.meshgrid(*[np.arange(1, n + 1, dtype=np.float64) for n in Ng], indexing="ij")
    return [np.asarray(idx[d] - 1.5 - (0.5 if d == i - 1 else 0.0), dtype=T) for d in range(len(Ng))]


def _ic_field(u0, Ng, T):
    """The initial condition on every cell: apply!(u0, u) for a function u0(i, x) (x: the
    coordinate arrays of loc(i,I)), or u0[i] per component for a tuple (src/Flow.jl:140)."""
    D = len(Ng)
    out = np.zeros(tuple(Ng) + (D,), dtype=T, order="F")
    for i in range(1, D + 1):
        v = u0(i, _loc_coords(Ng, i, T)) if callable(u0) else u0[i - 1]
        out[..., i - 1] = np.broadcast_to(np.asarray(v, dtype=T), Ng)
    return out


def ic_kwarg(u0, uλ):
    """Resolve the initial-condition keyword (src/Flow.jl:97-101): `uλ` is the deprecated name of
    `u0`; it warns and is used only when `u0` is not given."""
    if uλ is None:
        return u0
    import warnings
    warnings.warn("The `uλ` keyword argument is deprecated and will be removed in a future release. "
                  "Use `u0` instead.", DeprecationWarning, stacklevel=3)
    return uλ if u0 is None else u0


def _fn_field(fns, Ng, T, t):
    """Σ of the functions fns(i, x, t) at loc(i,I) on every cell, in the layout of u (apply!,
    src/util.jl); each function takes coordinate arrays and returns an array (or scalar) of T."""
    D = len(Ng)
    out = np.zeros(tuple(Ng) + (D,), dtype=T, order="F")
    for i in range(1, D + 1):
        x = _loc_coords(Ng, i, T)
        acc = None
        for fn in fns:
            v = np.broadcast_to(np.asarray(fn(i, x, T(t)), dtype=T), Ng)
            acc = v if acc is None else (acc + v).astype(T)
        out[..., i - 1] = acc
    return out


class Flow:
    """Flow(N, uBC; Δt=0.25, ν=0, g=None, u0=None, λ=quick, T=float32, perdir=(), exitBC=False).

    u0: the initial velocity, a function u0(i, x) or a tuple (default: uBC at t = 0); `uλ` is its
    deprecated alias (src/Flow.jl:133-142).
    uBC: a tuple (constant); a pair (U, dU/dt) for a boundary velocity uniform in space that varies
    linearly in time; or a function uBC(i, x, t) (x: the coordinate arrays of the cells' loc(i,I))
    evaluated on the host every step, with its time derivative `uBC_dt(i, x, t)` when it depends
    on t (∂uBC/∂t is taken as 0 without it).  g: None, a pair (g0, g1) for g(i,x,t) = g0[i] +
    g1[i]*t, or a function g(i, x, t).  The flow owns its MultiLevelPoisson (x≡p, L≡μ₀, z≡σ),
    exposed as `flow.pois`.
    """

    def __init__(self, N, uBC, Δt=0.25, ν=0.0, g=None, λ=quick, T=np.float32, perdir=(), exitBC=False,
                 dt=None, nu=None, uBC_dt=None, u0=None, uλ=None):
        u0 = ic_kwarg(u0, uλ)
        if dt is not None:
            Δt = dt
        if nu is not None:
            ν = nu
        self.D = len(N)
        self.N = tuple(int(v) for v in N)
        self.Ng = tuple(v + 2 for v in self.N)
        self.T = np.dtype(T).type
        d = A.WlFlowDesc()
        d.D = self.D
        for k in range(self.D):
            d.n[k] = self.N[k]
        d.dtype = _dtype_code(T)
        d.scheme = int(λ)
        d.perdir_mask = _perdir_mask(perdir)
        d.exit_bc = int(bool(exitBC))
        self.uBC_fn = uBC if callable(uBC) else None
        self.uBC_dt = uBC_dt
        self.g_fn = g if callable(g) else None
        if self.uBC_fn is not None:
            uBC = (0.0,) * self.D
        if self.g_fn is not None:
            g = None
        if len(uBC) == 2 and hasattr(uBC[0], "__len__"):
            U, rate = uBC
            d.has_uBC_rate = 1
            for k in range(self.D):
                d.uBC_rate[k] = float(rate[k])
        else:
            U = uBC
        for k in range(self.D):
            d.uBC[k] = float(U[k])
        if g is not None:
            d.has_g = 1
            for k in range(self.D):
                d.g0[k], d.g1[k] = float(g[0][k]), float(g[1][k])
        d.nu, d.dt0 = float(ν), float(Δt)
        self.desc = d
        self.uBC = tuple(float(U[k]) for k in range(self.D))
        self.ν = self.T(ν)
        self.perdir = tuple(perdir)
        self.exitBC = bool(exitBC)
        self.Δt = [float(self.T(Δt))]  # time-step history on the host (src/Flow.jl:127)
        self.h = C.c_void_p()
        A.check(A.load().wl_flow_create(C.byref(self.h), C.byref(d)))
        self.pois = FlowPoisson(self)
        if self.uBC_fn is not None or u0 is not None:
            # u = u0 (default uBC(i,x,0)) everywhere, then BC! and exitBC! (src/Flow.jl:138-141)
            ubc0 = _fn_field([self.uBC_fn], self.Ng, self.T, 0.0) if self.uBC_fn is not None else None
            self.set("u", _ic_field(u0, self.Ng, self.T) if u0 is not None else ubc0)
            if ubc0 is not None:
                self.set_field_fn(A.WL_FN_UBC, ubc0)
            A.check(A.load().wl_flow_apply_ic(self.h))

    def set_field_fn(self, which, a):
        """Upload a host-evaluated function field (WL_FN_UBC, WL_FN_ACC0, WL_FN_ACC1); None clears."""
        if a is None:
            A.check(A.load().wl_flow_set_field_fn(self.h, which, None, 0))
            return
        a = np.asfortranarray(a, dtype=self.T)
        A.check(A.load().wl_flow_set_field_fn(self.h, which, a.ctypes.data, a.nbytes))

    def _functions_for_step(self, t0, t1):
        """The function fields of the coming step: uBC at t₁; g + ∂uBC/∂t at t₀ and t₁."""
        if self.uBC_fn is not None:
            self.set_field_fn(A.WL_FN_UBC, _fn_field([self.uBC_fn], self.Ng, self.T, t1))
        fns = [f for f in (self.g_fn, self.uBC_dt if self.uBC_fn is not None else None) if f is not None]
        if fns:
            self.set_field_fn(A.WL_FN_ACC0, _fn_field(fns, self.Ng, self.T, t0))
            self.set_field_fn(A.WL_FN_ACC1, _fn_field(fns, self.Ng, self.T, t1))

    # fields, downloaded in Julia layout
    def _shape(self, name):
        if name in ("p", "sigma"):
            return self.Ng
        if name == "mu1":
            return self.Ng + (self.D, self.D)
        return self.Ng + (self.D,)

    def get(self, name):
        a = np.zeros(self._shape(name), dtype=self.T, order="F")
        A.check(A.load().wl_flow_download(self.h, A.FIELDS[name], a.ctypes.data, a.nbytes))
        return a

    def set(self, name, a):
        a = np.asfortranarray(a, dtype=self.T)
        if a.shape != self._shape(name):
            raise AssertionError(f"{name} must have shape {self._shape(name)}")
        A.check(A.load().wl_flow_upload(self.h, A.FIELDS[name], a.ctypes.data, a.nbytes))

    u = property(lambda s: s.get("u"), lambda s, v: s.set("u", v))
    u0 = property(lambda s: s.get("u0"), lambda s, v: s.set("u0", v))
    f = property(lambda s: s.get("f"), lambda s, v: s.set("f", v))
    p = property(lambda s: s.get("p"), lambda s, v: s.set("p", v))
    sigma = property(lambda s: s.get("sigma"), lambda s, v: s.set("sigma", v))
    V = property(lambda s: s.get("V"), lambda s, v: s.set("V", v))
    mu0 = property(lambda s: s.get("mu0"), lambda s, v: s.set("mu0", v))
    mu1 = property(lambda s: s.get("mu1"), lambda s, v: s.set("mu1", v))

    def autotune(self, reps=5):
        """Choose this flow's tile targets (GS half sweeps; restriction/prolongation) by timing;
        results do not change.  Returns the two targets."""
        t = (C.c_int32 * 2)()
        A.check(A.load().wl_flow_autotune(self.h, int(reps), t))
        return int(t[0]), int(t[1])

    def set_tiles(self, gs, rp):
        """Set the tile targets directly (0: the defaults), as autotune would."""
        A.check(A.load().wl_flow_set_tiles(self.h, int(gs), int(rp)))

    def residuals(self):
        """Σr² after residual! and after every V-cycle of the last mom_step (both projections)."""
        buf = (C.c_double * 512)()
        cnt = C.c_int()
        A.check(A.load().wl_flow_last_residuals(self.h, buf, 512, C.byref(cnt)))
        return [buf[k] for k in range(min(cnt.value, 512))]

    def profile(self, enable=True):
        """Start (clearing) or stop per-kernel HIP-event timing on the flow's stream."""
        A.check(A.load().wl_flow_profile(self.h, int(enable)))

    def profile_read(self):
        """Per-kernel stats since profile(True): dicts of name, level, launches, total_ms,
        bytes_per_launch (algorithmic bytes)."""
        buf = (A.WlKernelStat * 256)()
        cnt = C.c_int()
        A.check(A.load().wl_flow_profile_read(self.h, buf, 256, C.byref(cnt)))
        return [dict(name=buf[k].name.decode(), level=buf[k].level, launches=buf[k].launches,
                     total_ms=buf[k].total_ms, bytes_per_launch=buf[k].bytes_per_launch)
                for k in range(min(cnt.value, 256))]

    def __del__(self):
        try:
            if getattr(self, "h", None):
                A.load().wl_flow_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class FlowPoisson:
    """The MultiLevelPoisson(flow.p, flow.μ₀, flow.σ) attached to a Flow (src/WaterLily.jl:97)."""

    def __init__(self, flow):
        self.flow = flow
        self.n = []  # pois.n (src/MultiLevelPoisson.jl:66)

    def update(self):
        A.check(A.load().wl_flow_pois_update(self.flow.h))

    def set_solver(self, tol=1e-4, itmx=32, mode=A.WL_SOLVER_PLAIN):
        A.check(A.load().wl_flow_set_solver(self.flow.h, mode, tol, itmx))


def time(flow):
    """time(a::AbstractFlow) = sum(Δt[1:end-1]) (src/Flow.jl:174); time(meanflow) =
    t[end] − t[1] (src/Metrics.jl:220)."""
    if isinstance(flow, MeanFlow):
        return flow.T(flow.t[-1] - flow.t[0])
    return float(np.sum(np.asarray(flow.Δt[:-1], dtype=flow.T), dtype=flow.T)) if len(flow.Δt) > 1 else 0.0


class UdfFlow:
    """The flow as a udf! sees it during mom_step!: the fields (u, u0, f, p, sigma, V, mu0, mu1) are
    numpy arrays in Julia layout, downloaded when first read; the ones the function changes are
    uploaded when it returns.  Other attributes (ν, Δt, D, N, T, ...) are the flow's."""
    _FIELDS = ("u", "u0", "f", "p", "sigma", "V", "mu0", "mu1")

    def __init__(self, flow):
        self._flow = flow
        self._cache = {}

    def __getattr__(self, name):
        if name in UdfFlow._FIELDS:
            if name not in self._cache:
                a = self._flow.get(name)
                self._cache[name] = (a, a.tobytes())
            return self._cache[name][0]
        return getattr(self._flow, name)

    def _flush(self):
        for name, (a, orig) in self._cache.items():
            if a.tobytes() != orig:
                self._flow.set(name, a)


def _udf_call(udf, view, u, t, kw):
    """udf!(flow, force!, u, t; kw...) (src/Flow.jl:255-257): force!(flow, u, t; kw...) when that
    form applies, else force!(flow, t; kw...)."""
    import inspect
    try:
        inspect.signature(udf).bind(view, u, t, **kw)
        three = True
    except TypeError:
        three = False
    return udf(view, u, t, **kw) if three else udf(view, t, **kw)


def mom_step(flow, pois=None, udf=None, **kw):
    """mom_step!(a::Flow, b::MultiLevelPoisson; udf, kw...) (src/Flow.jl:156-167): one
    predictor-corrector step on the device, then push!(Δt, CFL).  udf(flow, [u,] t; kw...) runs
    between conv_diff! and accelerate! of both phases (the predictor with u = u⁰ at t₀, the
    corrector with u at t₁) on a UdfFlow view of the fields."""
    pois = flow.pois if pois is None else pois
    dt = flow.Δt[-1]
    t1 = float(np.sum(np.asarray(flow.Δt, dtype=flow.T), dtype=flow.T))
    t0 = float(flow.T(t1) - flow.T(dt))
    if flow.uBC_fn is not None or flow.g_fn is not None:
        flow._functions_for_step(t0, t1)
    out = C.c_double()
    nit = (C.c_int16 * 2)()
    err = []
    if udf is not None:
        def hook(user, h, phase, t):
            try:
                view = UdfFlow(flow)
                _udf_call(udf, view, view.u0 if phase == 0 else view.u, flow.T(t), kw)
                view._flush()
                return 0
            except BaseException as e:  # carried out of the C call and raised below
                err.append(e)
                return -1
        cb = A.UDF_FN(hook)
        A.check(A.load().wl_flow_set_udf(flow.h, cb, None))
        try:
            rc = A.load().wl_mom_step(flow.h, dt, t0, t1, C.byref(out), nit)
        finally:
            A.load().wl_flow_set_udf(flow.h, A.UDF_FN(), None)
        if err:
            raise err[0]
    else:
        rc = A.load().wl_mom_step(flow.h, dt, t0, t1, C.byref(out), nit)
    if rc == A.WL_ENAN:
        raise FloatingPointError(A.load().wl_last_error().decode())
    A.check(rc)
    flow.Δt.append(float(flow.T(out.value)))
    pois.n.extend([int(nit[0]), int(nit[1])])
    return pois


# ---------------------------------------------------------------------------------------------
# Poisson / MultiLevelPoisson on host arrays (src/Poisson.jl:22-39, src/MultiLevelPoisson.jl:61-77)
# ---------------------------------------------------------------------------------------------
class _DevicePoisson:
    multilevel = True

    def __init__(self, x, L, z, perdir=()):
        x = np.asfortranarray(x)
        L = np.asfortranarray(L, dtype=x.dtype)
        z = np.asfortranarray(z, dtype=x.dtype)
        if not (x.shape == z.shape and L.shape[:-1] == x.shape and L.shape[-1] == x.ndim):
            raise AssertionError("axes(x) == axes(z) == front(axes(L)) and last(axes(L)) == 1:ndims(x)")
        self.T = x.dtype.type
        self.nd = x.ndim
        self.Ng = x.shape
        Ng = (C.c_int32 * 3)(*(list(self.Ng) + [1] * (3 - self.nd)))
        self.h = C.c_void_p()
        A.check(A.load().wl_pois_create(C.byref(self.h), self.nd, Ng, _dtype_code(self.T), L.ctypes.data,
                                        int(self.multilevel), _perdir_mask(perdir)))
        self.n = []
        if np.any(x):
            self.set("x", x)
        if np.any(z):
            self.set("z", z)

    @property
    def nlevels(self):
        v = C.c_int()
        A.check(A.load().wl_pois_nlevels(self.h, C.byref(v)))
        return v.value

    def level_dims(self, level=0):
        Ng = (C.c_int32 * 3)()
        A.check(A.load().wl_pois_level_dims(self.h, level, Ng))
        return tuple(Ng[k] for k in range(self.nd))

    def get(self, name, level=0):
        shp = self.level_dims(level) + ((self.nd,) if name == "L" else ())
        a = np.zeros(shp, dtype=self.T, order="F")
        A.check(A.load().wl_pois_download(self.h, level, A.PFIELDS[name], a.ctypes.data, a.nbytes))
        return a

    def set(self, name, a, level=0):
        a = np.asfortranarray(a, dtype=self.T)
        A.check(A.load().wl_pois_upload(self.h, level, A.PFIELDS[name], a.ctypes.data, a.nbytes))

    x = property(lambda s: s.get("x"))
    z = property(lambda s: s.get("z"))
    r = property(lambda s: s.get("r"))
    L = property(lambda s: s.get("L"))
    D = property(lambda s: s.get("D"))
    iD = property(lambda s: s.get("iD"))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                A.load().wl_pois_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class Poisson(_DevicePoisson):
    """Poisson(x, L, z): Jacobi-preconditioned CG solver (src/Poisson.jl:22-39,166-214)."""
    multilevel = False


class MultiLevelPoisson(_DevicePoisson):
    """MultiLevelPoisson(x, L, z): geometric multigrid with GS-RB smoothing
    (src/MultiLevelPoisson.jl:61-127)."""
    multilevel = True

    class _LevelView:
        def __init__(self, ml, level):
            self.ml, self.level = ml, level

        def __getattr__(self, name):
            if name in A.PFIELDS:
                return self.ml.get(name, self.level)
            raise AttributeError(name)

    @property
    def levels(self):
        return [MultiLevelPoisson._LevelView(self, l) for l in range(self.nlevels)]


def mult(pois, x):
    """mult!(p, x): z = A x (ghosts 0), stored in p.z and returned (src/Poisson.jl:63-69)."""
    x = np.asfortranarray(x, dtype=pois.T)
    z = np.zeros_like(x, order="F")
    A.check(A.load().wl_pois_mult(pois.h, x.ctypes.data, z.ctypes.data, x.nbytes))
    return z


def update(pois, flow=None):
    """update!(p) (src/Poisson.jl:47, src/MultiLevelPoisson.jl:79-86); update!(meanflow, flow)
    (src/Metrics.jl:229-241) when given a MeanFlow."""
    if isinstance(pois, MeanFlow):
        return meanflow_update(pois, flow)
    if isinstance(pois, FlowPoisson):
        return pois.update()
    A.check(A.load().wl_pois_update(pois.h))


def solver(pois, tol=1e-4, itmx=None):
    """solver!(p; tol, itmx): itmx defaults to 32 (multigrid) or 1000 (pcg)."""
    if itmx is None:
        itmx = 32 if pois.multilevel else 1000
    n = C.c_int32()
    A.check(A.load().wl_pois_solve(pois.h, float(tol), int(itmx), C.byref(n)))
    pois.n.append(n.value)
    return pois


# ---------------------------------------------------------------------------------------------
# Simulation (src/WaterLily.jl:86-161)
# ---------------------------------------------------------------------------------------------
class Simulation:
    """Simulation(dims, uBC, L; U, Δt, ν, ϵ, g, perdir, u0, exitBC, λ, body, T)."""

    def __init__(self, dims, uBC, L, U=None, Δt=0.25, ν=0.0, ϵ=1, g=None, perdir=(), exitBC=False, λ=quick,
                 body=None, T=np.float32, dt=None, nu=None, eps=None, uBC_dt=None, u0=None, uλ=None):
        u0 = ic_kwarg(u0, uλ)
        if dt is not None:
            Δt = dt
        if nu is not None:
            ν = nu
        if eps is not None:
            ϵ = eps
        funcBC = callable(uBC) or (len(uBC) == 2 and hasattr(uBC[0], "__len__"))
        if U is None:
            if funcBC:
                raise AssertionError("`U` (velocity scale) must be specified if boundary conditions `uBC` is a `Function`")
            U = math.sqrt(sum(float(v) ** 2 for v in uBC))
        self.U, self.L, self.ϵ = U, L, ϵ
        self.body = body if body is not None else NoBody()
        self.flow = Flow(dims, uBC, Δt=Δt, ν=ν, g=g, λ=λ, T=T, perdir=perdir, exitBC=exitBC, uBC_dt=uBC_dt,
                         u0=u0)
        measure_flow(self.flow, self.body, t=0.0, ϵ=ϵ)
        self.pois = self.flow.pois
        self.pois.update()


def measure_flow(flow, body, t=0.0, ϵ=1):
    """measure!(flow, body; t, ϵ) (src/Body.jl:28-51) for analytic bodies; NoBody skips."""
    if isinstance(body, NoBody):
        return
    arr, nb = _body_array(body)
    A.check(A.load().wl_flow_measure(flow.h, arr, nb, float(t), float(ϵ)))


def measure(sim, t=None):
    """measure!(sim, t=sum(Δt)): re-measure the body and update! the Poisson coefficients."""
    t = float(sum(sim.flow.Δt)) if t is None else t
    if isinstance(sim.body, NoBody):
        sim.pois.update()
        return
    measure_flow(sim.flow, sim.body, t=t, ϵ=sim.ϵ)
    sim.pois.update()


def sim_time(sim):
    """tU/L with t = sum(Δt[1:end-1]) (src/WaterLily.jl:117)."""
    return time(sim.flow) * sim.U / sim.L


def sim_step(sim, t_end=None, remeasure=True, max_steps=None, udf=None, **kw):
    """sim_step!(sim[, t_end]; remeasure, max_steps, udf, kw...) (src/WaterLily.jl:128-139)."""
    if t_end is None:
        if remeasure:
            measure(sim)
        mom_step(sim.flow, sim.pois, udf=udf, **kw)
        return sim
    steps0 = len(sim.flow.Δt)
    while sim_time(sim) < t_end and (max_steps is None or len(sim.flow.Δt) - steps0 < max_steps):
        sim_step(sim, remeasure=remeasure, udf=udf, **kw)
    return sim


class Batch:
    """Many Simulations of different sizes stepped together: one kernel launch per stage of the
    time step for all members.  Members keep their own Δt, pois.n and fields; `step()` is one
    sim_step!(sim; remeasure=false) (src/WaterLily.jl:127-140) of every member, with the same
    results as stepping each alone.  Members share D, T and λ, use the plain solver, and have no
    periodic directions and no exitBC."""

    def __init__(self, sims):
        self.sims = list(sims)
        hs = (C.c_void_p * len(self.sims))(*[s.flow.h.value for s in self.sims])
        h = C.c_void_p()
        A.check(A.load().wl_batch_create(C.byref(h), hs, len(self.sims)))
        self.h = h

    def step(self):
        n = len(self.sims)
        dt, t0, t1 = (C.c_double * n)(), (C.c_double * n)(), (C.c_double * n)()
        for k, s in enumerate(self.sims):
            f = s.flow
            dt[k] = f.Δt[-1]
            t1[k] = float(np.sum(np.asarray(f.Δt, dtype=f.T), dtype=f.T))
            t0[k] = float(f.T(t1[k]) - f.T(dt[k]))
        out = (C.c_double * n)()
        nit = (C.c_int16 * (2 * n))()
        rc = A.load().wl_batch_step(self.h, dt, t0, t1, out, nit)
        if rc not in (0, A.WL_ENAN):
            A.check(rc)
        # every member took its step: the healthy ones keep their results, a member with a NaN or
        # Inf in u records a NaN Δt (and the call raises)
        for k, s in enumerate(self.sims):
            s.flow.Δt.append(float(s.flow.T(out[k])))
            s.pois.n.extend([int(nit[2 * k]), int(nit[2 * k + 1])])
        if rc == A.WL_ENAN:
            raise FloatingPointError(A.load().wl_last_error().decode())
        return self

    def __del__(self):
        try:
            if getattr(self, "h", None):
                A.load().wl_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


def pressure_force(sim):
    """pressure_force(sim) (src/Metrics.jl:126-133): Σ p·nds(body) in Float64."""
    F = (C.c_double * 3)()
    arr, nb = _body_array(sim.body)
    if nb == 0:
        return np.zeros(sim.flow.D)
    A.check(A.load().wl_pressure_force(sim.flow.h, arr, nb, time(sim.flow), F))
    return np.array([F[k] for k in range(sim.flow.D)])


def _metrics(sim, x0):
    D = sim.flow.D
    arr, nb = _body_array(sim.body)
    outs = [(C.c_double * 3)() for _ in range(4)]
    if nb == 0:
        return [np.zeros(3) for _ in range(4)]
    x = (C.c_double * 3)(*([float(v) for v in x0] + [0.0] * (3 - len(x0))))
    A.check(A.load().wl_flow_metrics(sim.flow.h, arr, nb, time(sim.flow), x, *outs))
    return [np.array(o[:]) for o in outs]


def _moment(m, D):
    # the 2-vector cross product is a scalar broadcast over df[I,:] (src/Metrics.jl:174,188)
    return np.array([m[0], m[0]]) if D == 2 else np.array(m[:3])


def viscous_force(sim):
    """viscous_force(sim) (src/Metrics.jl:146-155): Σ -2ν·S(I,u)·nds(body) in Float64."""
    return _metrics(sim, (0.0, 0.0, 0.0))[1][:sim.flow.D]


def total_force(sim):
    """total_force(sim) = pressure_force(sim) .+ viscous_force(sim) (src/Metrics.jl:157-162)."""
    o = _metrics(sim, (0.0, 0.0, 0.0))
    return o[0][:sim.flow.D] + o[1][:sim.flow.D]


def pressure_moment(x0, sim):
    """pressure_moment(x₀, sim) (src/Metrics.jl:164-175): Σ p·(x−x₀)×nds(body) in Float64."""
    return _moment(_metrics(sim, x0)[2], sim.flow.D)


def viscous_moment(x0, sim):
    """viscous_moment(x₀, sim) (src/Metrics.jl:177-189): Σ -2ν·(x−x₀)×(S·nds) in Float64."""
    return _moment(_metrics(sim, x0)[3], sim.flow.D)


def total_moment(x0, sim):
    """total_moment(x₀, sim) (src/Metrics.jl:191-196)."""
    o = _metrics(sim, x0)
    return _moment(o[2], sim.flow.D) + _moment(o[3], sim.flow.D)


class MeanFlow:
    """MeanFlow(flow; t_init=time(flow), uu_stats=false) (src/Metrics.jl:199-216): temporal
    averages of p, u (and u⊗u) on the flow's device; the time vector `t` is kept here."""

    def __init__(self, flow, t_init=None, uu_stats=False):
        self.T, self.D = flow.T, flow.D
        self.Ng = tuple(flow._shape("p"))
        self.uu_stats = bool(uu_stats)
        self.t = [self.T(time(flow) if t_init is None else t_init)]
        h = C.c_void_p()
        A.check(A.load().wl_meanflow_create(C.byref(h), flow.h, int(self.uu_stats)))
        self.h = h

    def _get(self, name):
        shp = {"P": self.Ng, "U": self.Ng + (self.D,), "UU": self.Ng + (self.D, self.D),
               "tau": self.Ng + (self.D, self.D)}[name]
        a = np.zeros(shp, dtype=self.T, order="F")
        A.check(A.load().wl_meanflow_download(self.h, A.MFIELDS[name], a.ctypes.data, a.nbytes))
        return a

    def _put(self, name, a):
        shp = {"P": self.Ng, "U": self.Ng + (self.D,), "UU": self.Ng + (self.D, self.D)}[name]
        a = np.asfortranarray(a, dtype=self.T)
        if a.shape != shp:
            raise AssertionError(f"{name} must have shape {shp}")
        A.check(A.load().wl_meanflow_upload(self.h, A.MFIELDS[name], a.ctypes.data, a.nbytes))

    P = property(lambda s: s._get("P"))
    U = property(lambda s: s._get("U"))
    UU = property(lambda s: s._get("UU") if s.uu_stats else None)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                A.load().wl_meanflow_destroy(self.h)
                self.h = None
        except Exception:
            pass


def meanflow_update(meanflow, flow):
    """update!(meanflow, flow) (src/Metrics.jl:229-241): ε = dt/(dt + time(meanflow) + eps(T)),
    1 on the first update, formed in T; the blend runs on the device."""
    T = meanflow.T
    dt = T(time(flow)) - meanflow.t[-1]
    e = dt / (dt + time(meanflow) + T(np.finfo(T).eps))
    if len(meanflow.t) == 1:
        e = T(1)
    A.check(A.load().wl_meanflow_update(meanflow.h, flow.h, float(e)))
    meanflow.t.append(T(meanflow.t[-1] + dt))


def reset(meanflow, t_init=0.0):
    """reset!(meanflow; t_init) (src/Metrics.jl:222-227)."""
    A.check(A.load().wl_meanflow_reset(meanflow.h))
    meanflow.t = [meanflow.T(t_init)]


def uu(meanflow):
    """uu(meanflow) (src/Metrics.jl:246-250): τ = UU − U⊗U."""
    return meanflow._get("tau")


def copy(flow, meanflow):
    """copy!(flow, meanflow) (src/Metrics.jl:254-257): u .= U; p .= P."""
    A.check(A.load().wl_meanflow_copy_to_flow(meanflow.h, flow.h))


def L2(a):
    """L₂(a) = Σ a² over inside(a) (src/Poisson.jl:188), on a host array."""
    a = np.asarray(a)
    return float(np.sum(np.square(a[tuple(slice(1, -1) for _ in range(a.ndim))], dtype=np.float64)))


def coarsen_mask(N):
    """coarsen_mask(N) = map(divisible, N) (src/MultiLevelPoisson.jl:29-31,52)."""
    return tuple(n % 2 == 0 and n > 4 for n in N)
