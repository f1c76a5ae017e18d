"""MI355X (gfx950) WaterLily time-step path: HIP kernels behind a C ABI (include/wl.h) and a
host-side mirror of the reference's Simulation / Flow / MultiLevelPoisson API.

The package directory is named `waterlily.jl_amd`; load it with `load_package()` from
`__graft_entry__` (it registers the module as `waterlily_amd`).
"""
from ._abi import LIB_PATH, WL_SOLVER_MIXED, WL_SOLVER_PLAIN, WaterLilyError, load as load_library  # noqa: F401
from .api import (  # noqa: F401
    AbstractBody, Batch, BodyUnion, Circle, Ellipsoid, Flow, FlowPoisson, L2, MeanFlow, MultiLevelPoisson, NoBody, Poisson,
    Simulation, Sphere, cds, coarsen_mask, copy, ic_kwarg, measure, measure_flow, meanflow_update, mom_step, mult,
    pressure_force, pressure_moment, quick, reset, sim_step, sim_time, solver, time, total_force, total_moment,
    update, uu, vanLeer, viscous_force, viscous_moment,
)
from .io import (  # noqa: F401
    VTKWriter, close, default_attrib, load_flow, load_meanflow, load_pvd, read_vti, save, save_flow, save_meanflow,
    vtk_writer, write_vti,
)
from .learned import WarmNet, init_weights, set_warmstart, warm_stats  # noqa: F401
