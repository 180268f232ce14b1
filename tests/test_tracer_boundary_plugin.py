"""Tracer boundary plug-ins on the GPU (gwm_tracers_ssprk3_callback).

advance_passive_tracers_ssprk3 calls a TracerBoundaryUpdater on each stage
input at the step-relative times 0, dt and dt/2 (src/dycore/dry_transport.cu:
485-508).  The callback path must (1) see exactly those times, (2) leave the
result bit-identical to the plain path when the plug-in changes nothing, and
(3) reproduce the device strip boundary bit for bit when the plug-in writes
the same strips on the host.  The grid-size limits of gwm_grid_create are
checked here too.
"""
import numpy as np
import pytest

import gwm_amd as G

pytestmark = pytest.mark.gpu

NX, NY, NZ = 20, 16, 8


@pytest.fixture(scope="module")
def ctx():
    c = G.Context(0, G.F32)
    yield c
    c.close()


def open_state(ctx, seed=5):
    zi = np.linspace(0.0, 8000.0, NZ + 1)
    grid = G.Grid(ctx, NX, NY, NZ, periodic_x=False, periodic_y=False, dx=1000.0, dy=1000.0,
                  z_interfaces=zi)
    st = G.State(grid, 3)
    st.init_constant(rho=1.1, theta=300.0, u=6.0, v=-4.0, w=0.0)
    rng = np.random.default_rng(seed)
    for t in range(3):
        q = (rng.random((NZ, NY, NX)) * 1e-3).astype(np.float32)
        st.set(G.TRACER0 + t, q * st.get(G.RHO))
    return grid, st


def test_plugin_times_and_identity(ctx):
    _, a = open_state(ctx)
    _, b = open_state(ctx)
    cfg = G.step_config(dt=4.0)
    times = []
    a.tracers_ssprk3_callback(cfg, lambda stage, t: times.append(t))
    b.tracers_ssprk3(cfg)
    assert times == [0.0, 4.0, 2.0]
    for t in range(3):
        assert np.array_equal(a.get(G.TRACER0 + t), b.get(G.TRACER0 + t)), t


def test_plugin_strips_equal_device_strips(ctx):
    _, a = open_state(ctx, seed=9)
    _, b = open_state(ctx, seed=9)
    b.capture_boundary()
    ref = [a.storage(G.TRACER0 + t).copy() for t in range(3)]
    h = 1

    def strips(stage, _t):
        for t in range(3):
            s = stage.storage(G.TRACER0 + t)
            r = ref[t]
            # outermost owned cells on the four open sides
            s[:, h:h + NY, h] = r[:, h:h + NY, h]
            s[:, h:h + NY, h + NX - 1] = r[:, h:h + NY, h + NX - 1]
            s[:, h, h:h + NX] = r[:, h, h:h + NX]
            s[:, h + NY - 1, h:h + NX] = r[:, h + NY - 1, h:h + NX]
            stage.set_storage(G.TRACER0 + t, s)

    cfg = G.step_config(dt=4.0)
    for _ in range(2):
        a.tracers_ssprk3_callback(cfg, strips)
        b.tracers_ssprk3(cfg, G.BOUNDARY_STRIPS)
    for t in range(3):
        assert np.array_equal(a.get(G.TRACER0 + t), b.get(G.TRACER0 + t)), t


def test_plugin_error_fails_the_call(ctx):
    _, a = open_state(ctx)

    def bad(stage, t):
        raise ValueError("plug-in failure")

    with pytest.raises(ValueError):
        a.tracers_ssprk3_callback(G.step_config(dt=4.0), bad)


def test_grid_limits(ctx):
    with pytest.raises(G.GwmError, match="512 levels"):
        G.Grid(ctx, 8, 8, 513)
    with pytest.raises(G.GwmError, match="4 GiB"):
        G.Grid(ctx, 9000, 9000, 16)
