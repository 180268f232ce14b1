"""gwm_budget_sums_noalloc (the allocation-free budget entry point) against
gwm_budget_sums, through the C-ABI: two calls in a row on one GPU, the
second on a smaller state, each equal to the allocating entry point's slots
(counts, extremes and the non-finite count exactly, sums to 1e-12)."""
import numpy as np
import pytest

import gwm_amd as G

pytestmark = pytest.mark.gpu

EXACT = ("rho_min", "rho_max", "theta_min", "theta_max", "u_min", "u_max", "v_min", "v_max",
         "w_min", "w_max", "precip_max", "precip_cells", "precipitating_cells", "qv_min",
         "qv_max", "qc_min", "qc_max", "qr_min", "qr_max", "nonfinite_values")
SUMS = ("total_dry_mass", "total_rho_theta_m", "total_horizontal_momentum_x",
        "total_horizontal_momentum_y", "total_vertical_momentum", "vapor_path", "cloud_path",
        "rain_path", "precip_sum", "precip_mean")


def moist_state(ctx, nx, ny, nz, seed):
    zi = np.linspace(0.0, 12000.0, nz + 1)
    grid = G.Grid(ctx, nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi)
    st = G.State(grid, 3)
    st.init_convection(1.0e5, seed, 0.5, 1000.0)
    rng = np.random.default_rng(seed)
    rho = st.get(0)
    qv = st.get(16)
    st.set(17, (0.02 * qv * (rng.random(qv.shape) < 0.4)).astype(qv.dtype))
    st.set(18, (0.01 * qv * (rng.random(qv.shape) < 0.3)).astype(qv.dtype))
    u = st.get(2) + rng.standard_normal(st.get(2).shape).astype(qv.dtype)
    st.set(2, u)
    p = (rng.random((ny, nx)) * (rng.random((ny, nx)) < 0.5)).astype(qv.dtype)
    st.set(G.PRECIP, p)
    del rho
    return grid, st


def same(a, b):
    for k in EXACT:
        assert a[k] == b[k] or (np.isnan(a[k]) and np.isnan(b[k])), (k, a[k], b[k])
    for k in SUMS:
        assert abs(a[k] - b[k]) <= 1e-12 * max(1.0, abs(b[k])), (k, a[k], b[k])


@pytest.mark.parametrize("prec", [G.F32, G.F64])
def test_noalloc_budget_equals_budget_on_two_calls(prec):
    ctx = G.Context(0, prec)
    g1, big = moist_state(ctx, 96, 80, 24, 3)
    g2, small = moist_state(ctx, 20, 12, 8, 4)
    w = small.get(4)
    w[2, 3, 4] = np.nan  # a non-finite value in the second, smaller state
    small.set(4, w)
    ref_big, ref_small = big.summary(), small.summary()
    got_big = big.summary(noalloc=True)
    got_small = small.summary(noalloc=True)
    same(got_big, ref_big)
    same(got_small, ref_small)
    assert got_small["nonfinite_values"] == 1 and np.isnan(got_small["w_max"])
    assert got_big["precipitating_cells"] > 0 and got_big["nonfinite_values"] == 0
    # the second call left the shared scratch holding its own partials; a
    # third call on the first state still reads only what it wrote
    same(big.summary(noalloc=True), ref_big)
    del big, small, g1, g2
    ctx.close()
