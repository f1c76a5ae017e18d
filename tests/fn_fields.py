"""Host evaluation of function-valued boundary values and forcing on every cell, for the tests:
f(i, x, t) is called with the component i (1-based), the list of coordinate arrays of loc(i,I)
(src/core.jl:177, in T) and t (in T), and returns an array of T."""
import numpy as np


def loc_coords(Ng, i, T):
    """Coordinates of loc(i,I) = I − 1.5 − ½eᵢ (1-based I) for every cell, as T arrays."""
    D = len(Ng)
    idx = np.meshgrid(*[np.arange(1, n + 1, dtype=np.float64) for n in Ng], indexing="ij")
    return [np.asarray(idx[d] - 1.5 - (0.5 if d == i - 1 else 0.0), dtype=T) for d in range(D)]


def field(fn, Ng, T, t):
    """The field F[I, i] = fn(i, loc(i,I), t) in the layout of u (Fortran order, component last)."""
    D = len(Ng)
    out = np.zeros(tuple(Ng) + (D,), dtype=T, order="F")
    for i in range(1, D + 1):
        out[..., i - 1] = np.broadcast_to(np.asarray(fn(i, loc_coords(Ng, i, T), T(t)), dtype=T), Ng)
    return out


def step_times(dts, T):
    """t₁ = sum(Δt), t₀ = t₁ − Δt[end], in T (src/Flow.jl:157)."""
    t1 = T(np.sum(np.asarray(dts, dtype=T), dtype=T))
    return T(t1 - T(dts[-1])), t1
