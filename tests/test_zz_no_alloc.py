"""wl_flow_metrics_noalloc (the body integrals with their block partials in a static device array)
against wl_flow_metrics (the same kernels with a buffer the library allocates), through the C ABI:
bit-identical outputs on two calls in a row, the second on a smaller flow, at the largest flow the
static array admits, and a clear WL_EINVAL (no device work) for a larger one."""
import ctypes as C
import sys

import numpy as np
import pytest

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


def both(wl, sim, x0):
    """(no-alloc outputs, allocating outputs): Fp, Fv, Mp, Mv as 3-vectors each."""
    from waterlily_amd import _abi as A
    arr, nb = wl.api._body_array(sim.body)
    x = (C.c_double * 3)(*([float(v) for v in x0] + [0.0] * (3 - len(x0))))
    res = []
    for fn in (A.load().wl_flow_metrics_noalloc, A.load().wl_flow_metrics):
        outs = [(C.c_double * 3)(*([np.nan] * 3)) for _ in range(4)]
        A.check(fn(sim.flow.h, arr, nb, wl.api.time(sim.flow), x, *outs))
        res.append(np.array([o[:] for o in outs]))
    return res


def check_equal(a, b, D):
    nM = 1 if D == 2 else 3
    for k, n in enumerate((D, D, nM, nM)):
        assert np.array_equal(a[k, :n], b[k, :n]), (k, a[k, :n], b[k, :n])
        assert np.all(np.isfinite(a[k, :n]))


def test_noalloc_matches_allocating_variant(wl):
    say("case: metrics no-alloc vs allocating, 3D sphere 96x64x64 then 2D circle 64x48")
    big = wl.Simulation((96, 64, 64), (1, 0, 0), 12, ν=12 / 300, body=wl.Sphere((30, 32, 32), 6), T=np.float32)
    for _ in range(3):
        wl.sim_step(big, remeasure=False)
    a, b = both(wl, big, (30.0, 32.0, 32.0))
    say(f"  3D: Fp {a[0]} Fv {a[1]}")
    check_equal(a, b, 3)
    assert np.any(a[0, :3] != 0) and np.any(a[1, :3] != 0)
    small = wl.Simulation((64, 48), (1, 0), 10, ν=10 / 200, body=wl.Sphere((20, 24), 5), T=np.float32)
    for _ in range(2):
        wl.sim_step(small, remeasure=False)
    a2, b2 = both(wl, small, (20.0, 24.0))
    say(f"  2D: Fp {a2[0]} Mp {a2[2]}")
    check_equal(a2, b2, 2)
    # the larger flow again: the scratch holds the smaller call's partials beyond nothing it reads
    a3, b3 = both(wl, big, (31.0, 30.0, 33.0))
    check_equal(a3, b3, 3)


def test_noalloc_at_the_size_limit(wl):
    """512x256x256 (131072 blocks, the array's capacity) is accepted and bit-identical; one more
    plane of blocks is refused on the host."""
    say("case: metrics no-alloc at 512x256x256 (capacity), then 512x256x257 (refused)")
    rng = np.random.default_rng(3)
    sim = wl.Simulation((512, 256, 256), (1, 0, 0), 64, ν=64 / 1e4, body=wl.Sphere((128, 128, 128), 32),
                        T=np.float32)
    sim.flow.set("p", rng.standard_normal(sim.flow.p.shape).astype(np.float32))
    sim.flow.set("u", rng.standard_normal(sim.flow.u.shape).astype(np.float32))
    a, b = both(wl, sim, (128.0, 128.0, 128.0))
    say(f"  capacity: Fp {a[0]} Fv {a[1]}")
    check_equal(a, b, 3)
    del sim
    import gc
    gc.collect()
    over = wl.Simulation((512, 256, 257), (1, 0, 0), 64, ν=64 / 1e4, body=wl.Sphere((128, 128, 128), 32),
                         T=np.float32)
    from waterlily_amd import _abi as A
    arr, nb = wl.api._body_array(over.body)
    x = (C.c_double * 3)(128.0, 128.0, 128.0)
    outs = [(C.c_double * 3)() for _ in range(4)]
    rc = A.load().wl_flow_metrics_noalloc(over.flow.h, arr, nb, 0.0, x, *outs)
    msg = A.load().wl_last_error().decode()
    say(f"  over capacity: rc {rc}: {msg}")
    assert rc == -1 and "131072" in msg
