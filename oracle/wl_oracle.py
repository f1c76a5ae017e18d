"""ORACLE -- test infrastructure only.

ctypes binding of oracle/build/libwl_oracle.so, the CPU restatement of WaterLily.jl v1.8.0's time
step (see wl_oracle.hpp).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
import this module; the product package never does.

Arrays cross the boundary in the reference's Julia layout: column-major (x fastest), ghosts
included, vector component last.  numpy arrays here are therefore Fortran-ordered with shape
(Nx, Ny[, Nz][, D[, D]]).
"""
import ctypes as C
import os
import subprocess
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libwl_oracle.so")

WL_F32, WL_F64 = 0, 1
WL_QUICK, WL_VANLEER, WL_CDS = 0, 1, 2
FIELDS = {"u": 0, "u0": 1, "f": 2, "p": 3, "sigma": 4, "V": 5, "mu0": 6, "mu1": 7}
PFIELDS = {"x": 0, "L": 1, "D": 2, "iD": 3, "eps": 4, "r": 5, "z": 6}


class WlBody(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("center", C.c_double * 3),
                ("radius", C.c_double), ("axes", C.c_double * 3)]


class WlFlowDesc(C.Structure):
    _fields_ = [("D", C.c_int32), ("n", C.c_int32 * 3), ("dtype", C.c_int32), ("scheme", C.c_int32),
                ("perdir_mask", C.c_int32), ("exit_bc", C.c_int32), ("has_g", C.c_int32),
                ("has_uBC_rate", C.c_int32), ("uBC", C.c_double * 3), ("uBC_rate", C.c_double * 3),
                ("g0", C.c_double * 3), ("g1", C.c_double * 3), ("nu", C.c_double), ("dt0", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
        P = C.POINTER
        sig = {
            "wlo_threads": (i, [i]),
            "wlo_scheme": (d, [i, d, d, d, i]),
            "wlo_phi_flux": (d, [i, i, P(d), i, i, d]),
            "wlo_kern": (d, [i, d, d]),
            "wlo_bc": (i, [i, i, P(C.c_int32), vp, P(d), i, i]),
            "wlo_perbc": (i, [i, i, P(C.c_int32), vp, i]),
            "wlo_exitbc": (i, [i, i, P(C.c_int32), vp, vp, d]),
            "wlo_flow_create": (i, [P(vp), P(WlFlowDesc)]),
            "wlo_flow_destroy": (i, [vp]),
            "wlo_flow_upload": (i, [vp, i, vp, sz]),
            "wlo_flow_download": (i, [vp, i, vp, sz]),
            "wlo_flow_measure": (i, [vp, P(WlBody), i, d, d]),
            "wlo_flow_pois_update": (i, [vp]),
            "wlo_flow_set_solver": (i, [vp, i, d, i]),
            "wlo_mom_step": (i, [vp, d, d, d, P(d), P(C.c_int16)]),
            "wlo_pressure_force": (i, [vp, P(WlBody), i, d, P(d)]),
            "wlo_flow_last_residuals": (i, [vp, P(d), i, P(i)]),
            "wlo_flow_phase": (i, [vp, i, d, d, d]),
            "wlo_flow_cfl": (d, [vp]),
            "wlo_flow_set_udf": (i, [vp, UDF_FN, vp]),
            "wlo_pois_create": (i, [P(vp), i, P(C.c_int32), i, vp, i, i]),
            "wlo_pois_destroy": (i, [vp]),
            "wlo_pois_nlevels": (i, [vp, P(i)]),
            "wlo_pois_level_dims": (i, [vp, i, P(C.c_int32)]),
            "wlo_pois_upload": (i, [vp, i, i, vp, sz]),
            "wlo_pois_download": (i, [vp, i, i, vp, sz]),
            "wlo_pois_mult": (i, [vp, vp, vp, sz]),
            "wlo_pois_update": (i, [vp]),
            "wlo_pois_solve": (i, [vp, d, i, P(C.c_int32)]),
            "wlo_pois_op": (i, [vp, i, i, d]),
            "wlo_pois_L2": (d, [vp, i]),
            "wlo_pois_vcycle": (i, [vp, d]),
            "wlo_metrics": (i, [i, i, P(C.c_int32), vp, vp, d, P(WlBody), i, P(d), P(d), P(d), P(d), P(d)]),
            "wlo_strain": (i, [i, i, P(C.c_int32), vp, P(C.c_int32), P(d)]),
            "wlo_meanflow_update": (i, [i, C.c_int64, i, d, vp, vp, vp, vp, vp]),
            "wlo_flow_set_field_fn": (i, [vp, i, vp, sz]),
            "wlo_bc_fn": (i, [i, i, P(C.c_int32), vp, vp, i, i]),
            "wlo_flow_apply_ic": (i, [vp]),
            "wlo_nn_num_weights": (C.c_int64, []),
            "wlo_nn_conv": (None, [P(C.c_int32), i, i, i, i, i, i, vp, vp, vp, vp]),
            "wlo_nn_forward": (None, [P(C.c_int32), vp, vp, i, vp]),
            "wlo_nn_gradient": (d, [P(C.c_int32), vp, vp, vp, i, vp]),
            "wlo_nn_adam": (None, [C.c_int64, vp, vp, vp, vp, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_int64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


UDF_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double)


def _np_dtype(dtype):
    return np.float64 if dtype == WL_F64 else np.float32


def _ck(rc, what):
    if rc != 0:
        raise AssertionError(f"oracle {what} failed with code {rc}")


def make_bodies(bodies):
    arr = (WlBody * max(1, len(bodies)))()
    for k, b in enumerate(bodies):
        arr[k].kind = b["kind"]
        for q in range(3):
            arr[k].center[q] = b.get("center", (0, 0, 0))[q] if q < len(b.get("center", ())) else 0.0
            arr[k].axes[q] = b.get("axes", (1, 1, 1))[q] if q < len(b.get("axes", ())) else 1.0
        arr[k].radius = b.get("radius", 0.0)
    return arr


def make_desc(n, uBC, nu=0.0, dt0=0.25, dtype=WL_F32, scheme=WL_QUICK, perdir=(), exit_bc=False,
              uBC_rate=None, g=None):
    d = WlFlowDesc()
    D = len(n)
    d.D = D
    for k in range(3):
        d.n[k] = n[k] if k < D else 0
        d.uBC[k] = uBC[k] if k < D else 0.0
    d.dtype, d.scheme = dtype, scheme
    d.perdir_mask = sum(1 << (j - 1) for j in perdir)
    d.exit_bc = int(exit_bc)
    if uBC_rate is not None:
        d.has_uBC_rate = 1
        for k in range(D):
            d.uBC_rate[k] = uBC_rate[k]
    if g is not None:
        d.has_g = 1
        for k in range(D):
            d.g0[k], d.g1[k] = g[0][k], g[1][k]
    d.nu, d.dt0 = nu, dt0
    return d


class OFlow:
    """Oracle Flow (src/Flow.jl:114-148) with its MultiLevelPoisson (x≡p, L≡μ₀, z≡σ)."""

    def __init__(self, desc):
        self.desc = desc
        self.D = desc.D
        self.n = tuple(desc.n[k] for k in range(self.D))
        self.Ng = tuple(v + 2 for v in self.n)
        self.dtype = desc.dtype
        self.npdt = _np_dtype(desc.dtype)
        self.h = C.c_void_p()
        _ck(lib().wlo_flow_create(C.byref(self.h), C.byref(desc)), "flow_create")
        self.dt = [desc.dt0]   # Δt history on the host, src/Flow.jl:127

    def shape(self, field):
        if field in ("p", "sigma"):
            return self.Ng
        if field == "mu1":
            return self.Ng + (self.D, self.D)
        return self.Ng + (self.D,)

    def get(self, field):
        a = np.zeros(self.shape(field), dtype=self.npdt, order="F")
        _ck(lib().wlo_flow_download(self.h, FIELDS[field], a.ctypes.data, a.nbytes), "download")
        return a

    def set(self, field, a):
        a = np.asfortranarray(a, dtype=self.npdt)
        assert a.shape == self.shape(field)
        _ck(lib().wlo_flow_upload(self.h, FIELDS[field], a.ctypes.data, a.nbytes), "upload")

    def measure(self, bodies, t=0.0, eps=1.0):
        arr = make_bodies(bodies)
        _ck(lib().wlo_flow_measure(self.h, arr, len(bodies), t, eps), "measure")

    def pois_update(self):
        _ck(lib().wlo_flow_pois_update(self.h), "pois_update")

    def set_solver(self, tol=1e-4, itmx=32):
        _ck(lib().wlo_flow_set_solver(self.h, 0, tol, itmx), "set_solver")

    def mom_step(self, udf=None):
        """mom_step!(flow, pois) with the host-side Δt bookkeeping of src/Flow.jl:157,166.
        udf(oflow, phase, t): called between conv_diff! and accelerate! (phase 0 predictor, 1
        corrector); it reads and writes fields with get/set."""
        dt = self.dt[-1]
        t1 = float(np.sum(np.asarray(self.dt, dtype=self.npdt), dtype=self.npdt))
        t0 = t1 - dt
        out = C.c_double()
        nit = (C.c_int16 * 2)()
        cb = None
        if udf is not None:
            cb = UDF_FN(lambda user, h, phase, t: udf(self, phase, t) or 0)
            lib().wlo_flow_set_udf(self.h, cb, None)
        try:
            _ck(lib().wlo_mom_step(self.h, dt, t0, t1, C.byref(out), nit), "mom_step")
        finally:
            if cb is not None:
                lib().wlo_flow_set_udf(self.h, UDF_FN(), None)
        self.dt.append(float(self.npdt(out.value)))
        return (int(nit[0]), int(nit[1]))

    def time(self):
        return float(sum(self.dt[:-1]))

    def set_field_fn(self, which, a):
        """Function-valued uBC (which 0, at t₁) or forcing g + ∂uBC/∂t (1 at t₀, 2 at t₁) on every
        cell, in the layout of u; None clears."""
        if a is None:
            _ck(lib().wlo_flow_set_field_fn(self.h, which, None, 0), "set_field_fn")
            return
        a = np.asfortranarray(a, dtype=_np_dtype(self.dtype))
        _ck(lib().wlo_flow_set_field_fn(self.h, which, a.ctypes.data, a.nbytes), "set_field_fn")

    def apply_ic(self):
        _ck(lib().wlo_flow_apply_ic(self.h), "apply_ic")

    def pressure_force(self, bodies, t=None):
        arr = make_bodies(bodies)
        F = (C.c_double * 3)()
        _ck(lib().wlo_pressure_force(self.h, arr, len(bodies), self.time() if t is None else t, F),
            "pressure_force")
        return np.array([F[k] for k in range(self.D)])

    def residuals(self):
        buf = (C.c_double * 256)()
        cnt = C.c_int()
        lib().wlo_flow_last_residuals(self.h, buf, 256, C.byref(cnt))
        return [buf[k] for k in range(min(cnt.value, 256))]

    def phase(self, phase, dt=0.0, w=1.0, t=0.0):
        _ck(lib().wlo_flow_phase(self.h, phase, dt, w, t), "phase")

    def cfl(self):
        return lib().wlo_flow_cfl(self.h)

    def __del__(self):
        try:
            if self.h:
                lib().wlo_flow_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class OPois:
    """Oracle Poisson / MultiLevelPoisson built from L (src/Poisson.jl:32-38)."""

    def __init__(self, L, multilevel=True, perdir=(), dtype=WL_F32):
        self.npdt = _np_dtype(dtype)
        L = np.asfortranarray(L, dtype=self.npdt)
        self.D = L.ndim - 1
        self.Ng = L.shape[:-1]
        self.dtype = dtype
        self.multilevel = multilevel
        Ng = (C.c_int32 * 3)(*(list(self.Ng) + [1] * (3 - self.D)))
        self.h = C.c_void_p()
        rc = lib().wlo_pois_create(C.byref(self.h), self.D, Ng, dtype, L.ctypes.data, int(multilevel),
                                   sum(1 << (j - 1) for j in perdir))
        if rc != 0:
            raise AssertionError("MultiLevelPoisson requires size=a2ⁿ, where n>2")
        self.n = []

    @property
    def nlevels(self):
        v = C.c_int()
        lib().wlo_pois_nlevels(self.h, C.byref(v))
        return v.value

    def level_dims(self, level):
        Ng = (C.c_int32 * 3)()
        lib().wlo_pois_level_dims(self.h, level, Ng)
        return tuple(Ng[k] for k in range(self.D))

    def get(self, field, level=0):
        shp = self.level_dims(level) + ((self.D,) if field == "L" else ())
        a = np.zeros(shp, dtype=self.npdt, order="F")
        _ck(lib().wlo_pois_download(self.h, level, PFIELDS[field], a.ctypes.data, a.nbytes), "download")
        return a

    def set(self, field, a, level=0):
        a = np.asfortranarray(a, dtype=self.npdt)
        _ck(lib().wlo_pois_upload(self.h, level, PFIELDS[field], a.ctypes.data, a.nbytes), "upload")

    def mult(self, x):
        x = np.asfortranarray(x, dtype=self.npdt)
        z = np.zeros_like(x, order="F")
        _ck(lib().wlo_pois_mult(self.h, x.ctypes.data, z.ctypes.data, x.nbytes), "mult")
        return z

    def update(self):
        _ck(lib().wlo_pois_update(self.h), "update")

    def solve(self, tol=1e-4, itmx=None):
        if itmx is None:
            itmx = 32 if self.multilevel else 1000
        n = C.c_int32()
        _ck(lib().wlo_pois_solve(self.h, tol, itmx, C.byref(n)), "solve")
        self.n.append(n.value)
        return n.value

    def op(self, op, level=0, w=1.0):
        _ck(lib().wlo_pois_op(self.h, level, op, w), "op")

    def L2(self, level=0):
        return lib().wlo_pois_L2(self.h, level)

    def vcycle(self, w=1.0):
        _ck(lib().wlo_pois_vcycle(self.h, w), "vcycle")

    def __del__(self):
        try:
            if self.h:
                lib().wlo_pois_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


def bc(a, U, saveexit=False, perdir=()):
    """BC!(a,U,saveexit,perdir) on a numpy vector field in place (src/core.jl:200-219)."""
    assert a.flags.f_contiguous
    D = a.ndim - 1
    dtype = WL_F64 if a.dtype == np.float64 else WL_F32
    Ng = (C.c_int32 * 3)(*(list(a.shape[:-1]) + [1] * (3 - D)))
    Ub = (C.c_double * 3)(*(list(U) + [0.0] * (3 - len(U))))
    lib().wlo_bc(dtype, D, Ng, a.ctypes.data, Ub, int(saveexit), sum(1 << (j - 1) for j in perdir))
    return a


def bc_fn(a, UB, saveexit=False, perdir=()):
    """BC!(a, uBC::Function, saveexit, perdir, t) with UB[I,i] = uBC(i, loc(i,I), t) (src/core.jl:201-219)."""
    assert a.flags.f_contiguous
    D = a.ndim - 1
    dtype = WL_F64 if a.dtype == np.float64 else WL_F32
    UB = np.asfortranarray(UB, dtype=a.dtype)
    Ng = (C.c_int32 * 3)(*(list(a.shape[:-1]) + [1] * (3 - D)))
    lib().wlo_bc_fn(dtype, D, Ng, a.ctypes.data, UB.ctypes.data, int(saveexit), sum(1 << (j - 1) for j in perdir))
    return a


def perbc(a, perdir):
    D = a.ndim
    dtype = WL_F64 if a.dtype == np.float64 else WL_F32
    Ng = (C.c_int32 * 3)(*(list(a.shape) + [1] * (3 - D)))
    lib().wlo_perbc(dtype, D, Ng, a.ctypes.data, sum(1 << (j - 1) for j in perdir))
    return a


def exitbc(u, u0, dt):
    D = u.ndim - 1
    dtype = WL_F64 if u.dtype == np.float64 else WL_F32
    Ng = (C.c_int32 * 3)(*(list(u.shape[:-1]) + [1] * (3 - D)))
    lib().wlo_exitbc(dtype, D, Ng, u.ctypes.data, u0.ctypes.data, dt)
    return u


def metrics(p, u, nu, bodies, x0=(0.0, 0.0, 0.0)):
    """(pressure_force, viscous_force, pressure_moment, viscous_moment) of src/Metrics.jl:126-197
    on numpy fields p (Ng) and u (Ng×D); either may be None.  Moments: 1 entry in 2D, 3 in 3D."""
    f = p if p is not None else u
    D = f.ndim if p is not None else f.ndim - 1
    Ngl = list(f.shape[:D])
    dtype = WL_F64 if f.dtype == np.float64 else WL_F32
    Ng = (C.c_int32 * 3)(*(Ngl + [1] * (3 - D)))
    outs = [(C.c_double * 3)() for _ in range(4)]
    x = (C.c_double * 3)(*(list(x0) + [0.0] * (3 - len(x0))))
    pp = None if p is None else np.asfortranarray(p)
    uu = None if u is None else np.asfortranarray(u)
    _ck(lib().wlo_metrics(dtype, D, Ng, None if pp is None else pp.ctypes.data, None if uu is None else uu.ctypes.data,
                          nu, make_bodies(bodies), len(bodies), x, *outs), "metrics")
    nM = 1 if D == 2 else 3
    return (np.array(outs[0][:D]), np.array(outs[1][:D]), np.array(outs[2][:nM]), np.array(outs[3][:nM]))


def strain(u, I):
    """S(I,u) (src/Metrics.jl:140) at the 1-based cell I of a numpy velocity field."""
    D = u.ndim - 1
    uu = np.asfortranarray(u)
    dtype = WL_F64 if u.dtype == np.float64 else WL_F32
    Ng = (C.c_int32 * 3)(*(list(u.shape[:-1]) + [1] * (3 - D)))
    Ic = (C.c_int32 * 3)(*(list(I) + [1] * (3 - D)))
    S = (C.c_double * 9)()
    _ck(lib().wlo_strain(dtype, D, Ng, uu.ctypes.data, Ic, S), "strain")
    return np.array(S[:D * D]).reshape(D, D)


def meanflow_update(e, P, p, U, u, UU=None):
    """The update!(meanflow, flow) blend (src/Metrics.jl:234-240) in place, weight e."""
    D = U.size // P.size
    dtype = WL_F64 if p.dtype == np.float64 else WL_F32
    _ck(lib().wlo_meanflow_update(dtype, p.size, D, e, P.ctypes.data, p.ctypes.data, U.ctypes.data, u.ctypes.data,
                                  None if UU is None else UU.ctypes.data), "meanflow_update")


class OSimulation:
    """Simulation(dims, uBC, L; ...) driven on the oracle (src/WaterLily.jl:86-149).

    Bodies are analytic (dicts with kind/center/radius/axes, see make_bodies)."""

    def __init__(self, dims, uBC, L, U=None, dt=0.25, nu=0.0, eps=1.0, perdir=(), exit_bc=False,
                 scheme=WL_QUICK, bodies=(), dtype=WL_F32, uBC_rate=None, g=None):
        self.U = float(np.sqrt(np.sum(np.square(uBC)))) if U is None else U
        self.L, self.eps, self.bodies = L, eps, list(bodies)
        self.flow = OFlow(make_desc(dims, uBC, nu=nu, dt0=dt, dtype=dtype, scheme=scheme, perdir=perdir,
                                    exit_bc=exit_bc, uBC_rate=uBC_rate, g=g))
        if self.bodies:
            self.flow.measure(self.bodies, t=0.0, eps=eps)
            self.flow.pois_update()
        self.n = []   # pois.n

    def sim_time(self):
        return self.flow.time() * self.U / self.L

    def measure(self):
        if self.bodies:
            t = float(sum(self.flow.dt))
            self.flow.measure(self.bodies, t=t, eps=self.eps)
            self.flow.pois_update()

    def step(self, remeasure=True, udf=None):
        if remeasure:
            self.measure()
        self.n.extend(self.flow.mom_step(udf))

    def sim_step(self, t_end, remeasure=True, max_steps=1 << 62, udf=None):
        steps0 = len(self.flow.dt)
        while self.sim_time() < t_end and len(self.flow.dt) - steps0 < max_steps:
            self.step(remeasure, udf)

    def pressure_force(self):
        return self.flow.pressure_force(self.bodies)


# ---------------------------------------------------------------------------------------------
# The warm-start network (wl_nn_oracle.cpp).  Tensors are C-ordered numpy arrays (V, C), voxel
# index v = (z·ny + y)·nx + x; weights are the canonical float32 vector of all layers
# (each layer w[tap][ci][co]).  round=True rounds at the device's bf16 storage points.
# ---------------------------------------------------------------------------------------------
NN_LAYERS = [(8, 16), (16, 16), (16, 32), (32, 64), (64, 64), (64, 32), (32, 16), (16, 16), (16, 8)]


def _f32c(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def nn_num_weights():
    return int(lib().wlo_nn_num_weights())


def nn_conv(grid, w, a, b=None, relu=True, round=True, outf32=False):
    ci, cos = w.shape[1], w.shape[2]
    a, w = _f32c(a), _f32c(w)
    bb = _f32c(b) if b is not None else None
    out = np.zeros((a.shape[0], cos), np.float32)
    g = (C.c_int32 * 3)(*grid)
    lib().wlo_nn_conv(g, ci, cos, int(b is not None), int(relu), int(round), int(outf32), a.ctypes.data,
                      bb.ctypes.data if bb is not None else None, w.ctypes.data, out.ctypes.data)
    return out


def nn_forward(grid, w, in0, round=True):
    w, in0 = _f32c(w), _f32c(in0)
    out = np.zeros(in0.shape[0], np.float32)
    lib().wlo_nn_forward((C.c_int32 * 3)(*grid), w.ctypes.data, in0.ctypes.data, int(round), out.ctypes.data)
    return out


def nn_gradient(grid, w, in0, target, round=True, grad=True):
    w, in0, target = _f32c(w), _f32c(in0), _f32c(target)
    g = np.zeros(w.size, np.float64) if grad else None
    L = lib().wlo_nn_gradient((C.c_int32 * 3)(*grid), w.ctypes.data, in0.ctypes.data, target.ctypes.data,
                              int(round), g.ctypes.data if grad else None)
    return (L, g) if grad else L


def nn_adam(w, g, m, v, lr, t, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in place on float32 arrays w, m, v with gradient g."""
    g = _f32c(g)
    lib().wlo_nn_adam(w.size, w.ctypes.data, g.ctypes.data, m.ctypes.data, v.ctypes.data, lr, b1, b2, eps, t)
