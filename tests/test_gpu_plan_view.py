"""Plan-view products on the GPU (gwm_plan_view, one column pass) against the
oracle's restatement of src/io/plan_view_output.cpp: the reference's known
answers on a constant state, and a moist, terrain-following state after a
few steps, in FP64 (relative 1e-12 per field: the library pow/log/log10 of
device and host) and FP32 (1e-5: their float forms differ by an ulp or two,
which RH and the dewpoint carry through a quotient and a logarithm)."""
import numpy as np
import pytest

import gwm_amd as G
import pyoracle as O

pytestmark = pytest.mark.gpu

NAMES = O.PLAN_VIEW_NAMES
assert NAMES == G.PLAN_VIEW_NAMES


def compare(nf_g, out_g, nf_o, out_o, tol, label):
    assert nf_g == nf_o, (label, nf_g, nf_o)
    for f in range(nf_o):
        a, b = out_g[f], out_o[f]
        assert np.all(np.isfinite(a)), (label, NAMES[f])
        scale = max(np.max(np.abs(b)), 1e-300)
        err = np.max(np.abs(a - b)) / scale
        assert err <= tol, (label, NAMES[f], err)
    assert np.all(out_g[nf_o:] == 0.0)


@pytest.mark.parametrize("prec,tol", [(8, 1e-12), (4, 1e-5)])
def test_constant_state_known_answers(prec, tol):
    ctx = G.Context(0, G.F64 if prec == 8 else G.F32)
    grid = G.Grid(ctx, 6, 4, 3, dx=1000.0, dy=1000.0, z_top=9000.0)
    st = G.State(grid, 3)
    st.init_constant(1.2, 300.0, 10.0, -4.0, 0.5)
    rho = st.get(O.RHO)
    for t, q in enumerate((0.004, 0.001, 0.0005)):
        st.set(O.TRACER0 + t, (rho * q).astype(rho.dtype))
    pr = np.zeros((1, 4, 6))
    pr[0, 0, 0] = 0.25
    st.set(O.PRECIP, pr)
    nf, out = st.plan_view(1, 4, 2.0)
    assert nf == 23
    get = lambda n: out[NAMES.index(n)]  # noqa: E731
    assert np.allclose(get("z_center"), 4500.0)
    assert np.allclose(get("wind_speed"), np.sqrt(116.0), atol=1e-5)
    assert np.allclose(get("column_cloud_water"), 10.8, atol=1e-5)
    assert np.allclose(get("column_rain_water"), 5.4, atol=1e-5)
    assert np.allclose(get("column_rain_fraction"), 1.0 / 3.0, atol=1e-6)
    assert abs(get("mean_surface_precipitation_rate")[0, 0] - 112.5) < 1e-4
    w = O.World(6, 4, 3, dx=1000.0, dy=1000.0, z_top=9000.0, n_tracers=3, precision=prec)
    for f in (O.RHO, O.RHO_THETA, O.MOM_U, O.MOM_V, O.MOM_W, 16, 17, 18, O.PRECIP):
        w.set(f, st.get(f))
    compare(nf, out, *w.plan_view(1, 4, 2.0), tol, f"constant {prec}")
    del st, grid
    ctx.close()


@pytest.mark.parametrize("prec,tol", [(8, 1e-12), (4, 1e-5)])
@pytest.mark.parametrize("terrain", [False, True])
def test_moist_state_after_steps(prec, tol, terrain):
    ctx = G.Context(0, G.F64 if prec == 8 else G.F32)
    nx, ny, nz = 40, 28, 16
    zi = G.stretched_levels(nz, 150.0, 900.0, 8000.0)
    kw = dict(dx=1000.0, dy=1000.0, z_interfaces=zi)
    grid = G.Grid(ctx, nx, ny, nz, **kw)
    st = G.State(grid, 3)
    st.init_convection(1.0e5, 11, 0.5, 1000.0)
    rng = np.random.default_rng(2)
    rho = st.get(O.RHO)
    st.set(17, (rho * 1e-3 * rng.random(rho.shape) * (rng.random(rho.shape) < 0.4)).astype(rho.dtype))
    st.set(18, (rho * 5e-4 * rng.random(rho.shape) * (rng.random(rho.shape) < 0.4)).astype(rho.dtype))
    cfg = G.step_config(dt=1.0, fast_substeps=3)
    wr = G.warm_rain_config(dt=1.0, rain_terminal_velocity=12.0)
    for _ in range(3):
        st.moist_step(cfg, wr)
    if terrain:
        # the same fields on a terrain-following grid: the column paths use
        # each column's own cell depths
        tgrid = G.Grid(ctx, nx, ny, nz, terrain=G.cosine_mountain(nx, ny, 700.0), **kw)
        ts = G.State(tgrid, 3)
        for f in (O.RHO, O.RHO_THETA, O.MOM_U, O.MOM_V, O.MOM_W, 16, 17, 18, O.PRECIP):
            ts.set(f, st.get(f))
        st, grid = ts, tgrid
        kw["terrain"] = G.cosine_mountain(nx, ny, 700.0)
    assert st.summary()["precip_sum"] > 0.0
    w = O.World(nx, ny, nz, n_tracers=3, precision=prec, **kw)
    for f in (O.RHO, O.RHO_THETA, O.MOM_U, O.MOM_V, O.MOM_W, 16, 17, 18, O.PRECIP):
        w.set(f, st.get(f))
    for k in (0, 5, nz - 1):
        nf, out = st.plan_view(k, 3, 1.0)
        compare(nf, out, *w.plan_view(k, 3, 1.0), tol, f"moist {prec} terrain={terrain} k={k}")
    nf, out = st.plan_view(5, 3, 1.0, with_precip=False)
    assert nf == 21
    del st, grid
    ctx.close()


def test_perturbation_form_plan_view_matches_full_values():
    ctx = G.Context(0, G.F32)
    nx, ny, nz = 32, 24, 12
    zi = np.linspace(0.0, 12000.0, nz + 1)
    grid = G.Grid(ctx, nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi)
    st = G.State(grid, 3, perturbation=True)
    st.init_convection(1.0e5, 3, 0.5, 1000.0)
    st.moist_step(G.step_config(dt=1.0, fast_substeps=3),
                  G.warm_rain_config(dt=1.0, rain_terminal_velocity=12.0))
    w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi, n_tracers=3, precision=4)
    for f in (O.RHO, O.RHO_THETA, O.MOM_U, O.MOM_V, O.MOM_W, 16, 17, 18, O.PRECIP):
        w.set(f, st.get(f))
    nf, out = st.plan_view(4, 1, 1.0)
    # the device adds the FP32 departure to the FP32 base; the host rounds
    # the FP64 sum once: an ulp of rho and rho*theta on top of the bar above
    compare(nf, out, *w.plan_view(4, 1, 1.0), 2e-5, "perturbation")
    del st, grid
    ctx.close()
