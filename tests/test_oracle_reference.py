"""Pins the CPU oracle against the reference's own known answers.

Each test restates one of the reference's unit/regression tests (cited by
file:line) on the oracle, with the reference's tolerances.  These run on the
CPU only.
"""
import numpy as np
import pytest

import pyoracle


# ----------------------------------------------------------------- halos ---

def test_scalar_halo_exchange_known_answers(orc):
    # tests/unit/test_halo_exchange.cpp:8-19
    w = orc.World(8, 8, 1, halo=1, ranks_x=2, ranks_y=2, precision=4)
    j, i = np.mgrid[0:8, 0:8]
    w.set(orc.RHO, (100 * j + i)[None].astype(np.float32))
    w.exchange(orc.RHO)
    s = w.rank_storage(0, orc.RHO)  # (nz, py, px), index [k, j+h, i+h]
    h = 1
    assert s[0, 0 + h, -1 + h] == 7.0
    assert s[0, 0 + h, 4 + h] == 4.0
    assert s[0, -1 + h, 0 + h] == 700.0
    assert s[0, 4 + h, 0 + h] == 400.0


def test_face_halo_exchange_known_answers(orc):
    w = orc.World(8, 8, 2, halo=1, ranks_x=2, ranks_y=2, precision=4)
    for f in (orc.MOM_U, orc.MOM_V, orc.MOM_W):
        w.fill(f, -999.0)
        for r in range(4):
            d = w.rank_desc(r)
            s = w.rank_storage(r, f)
            fz, py, px = s.shape
            h = 1
            for k in range(fz):
                for jj in range(py - 2 * h):
                    for ii in range(px - 2 * h):
                        s[k, jj + h, ii + h] = 1000 * k + 100 * (d["j_begin"] + jj) + d["i_begin"] + ii
            w.set_rank_storage(r, f, s)
        w.exchange(f)
    h = 1
    x0 = w.rank_storage(0, orc.MOM_U)
    assert x0[0, 0 + h, -1 + h] == 7.0
    assert x0[0, 0 + h, 5 + h] == 5.0
    assert x0[0, -1 + h, 0 + h] == 700.0
    assert x0[0, -1 + h, -1 + h] == 707.0
    y0 = w.rank_storage(0, orc.MOM_V)
    assert y0[0, 0 + h, -1 + h] == 7.0
    assert y0[0, -1 + h, 0 + h] == 700.0
    assert y0[0, 5 + h, 0 + h] == 500.0
    assert y0[0, -1 + h, -1 + h] == 707.0
    z0 = w.rank_storage(0, orc.MOM_W)
    assert z0[0, 0 + h, -1 + h] == 7.0
    assert z0[1, -1 + h, 0 + h] == 1700.0
    assert z0[1, -1 + h, -1 + h] == 1707.0


def test_scalar_halo_buffers_plan(orc):
    # tests/unit/test_scalar_halo_buffers.cpp:10-30
    w = orc.World(8, 8, 2, halo=1, ranks_x=2, ranks_y=2, precision=4)
    nb = w.neighbors(0)
    assert nb == {"west": 1, "east": 1, "south": 2, "north": 2}
    d = w.rank_desc(0)
    nxl, nyl = d["i_end"] - d["i_begin"], d["j_end"] - d["j_begin"]
    assert 1 * nyl * 2 == 8 and 1 * (nxl + 2) * 2 == 12
    k, j, i = np.mgrid[0:2, 0:8, 0:8]
    w.set(orc.RHO, (1000 * k + 100 * j + i).astype(np.float32))
    s = w.rank_storage(0, orc.RHO)
    assert s[0, 1, 1 + 0] == 0.0 and s[0, 1, 1 + 3] == 3.0  # west/east send
    w.exchange(orc.RHO)
    s = w.rank_storage(0, orc.RHO)
    assert s[0, 1, 0] == 7.0
    assert s[1, 0, 1] == 1700.0


@pytest.mark.parametrize("rx,ry,px,py,halo", [(1, 1, True, True, 1), (2, 2, True, True, 1),
                                             (3, 2, False, True, 1), (1, 1, False, False, 2),
                                             (2, 3, True, False, 2)])
def test_scalar_halo_brute_force(orc, rx, ry, px, py, halo):
    """Every ghost is the periodically wrapped owned value, or 0 across an open
    side; corners follow the y phase (halo_exchange.cpp:563-620)."""
    nx, ny, nz = 7, 6, 2
    w = orc.World(nx, ny, nz, halo=halo, ranks_x=rx, ranks_y=ry, periodic_x=px,
                  periodic_y=py, precision=8)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((nz, ny, nx))
    w.set(orc.RHO, g)
    w.exchange(orc.RHO)
    for r in range(w.num_ranks):
        d = w.rank_desc(r)
        s = w.rank_storage(r, orc.RHO)
        for jj in range(-halo, d["j_end"] - d["j_begin"] + halo):
            for ii in range(-halo, d["i_end"] - d["i_begin"] + halo):
                gi, gj = d["i_begin"] + ii, d["j_begin"] + jj
                if (gi < 0 or gi >= nx) and not px:
                    exp = np.zeros(nz)
                elif (gj < 0 or gj >= ny) and not py:
                    exp = np.zeros(nz)
                else:
                    exp = g[:, gj % ny, gi % nx]
                np.testing.assert_array_equal(s[:, jj + halo, ii + halo], exp)


def test_interface_sync_1x1_periodic_averages_wrap_faces(orc):
    # halo_exchange.cpp:255-267: east == self in a 1x1 periodic layout
    w = orc.World(5, 4, 2, precision=8)
    u = np.random.default_rng(0).standard_normal(w.global_shape(orc.MOM_U))
    w.set(orc.MOM_U, u)
    w.sync_interfaces(orc.MOM_U)
    out = w.get(orc.MOM_U)
    avg = 0.5 * (u[:, :, 5] + u[:, :, 0])
    np.testing.assert_array_equal(out[:, :, 0], avg)
    np.testing.assert_array_equal(out[:, :, 5], avg)
    np.testing.assert_array_equal(out[:, :, 1:5], u[:, :, 1:5])


def test_interface_sync_2x1_periodic_skips_wrap_interface(orc):
    # halo_exchange.cpp:256-257: for ranks_x >= 2 the wrap interface is never averaged
    w = orc.World(6, 4, 1, ranks_x=2, precision=8)
    u = np.random.default_rng(1).standard_normal(w.global_shape(orc.MOM_U))
    before = [w.rank_storage(r, orc.MOM_U).copy() for r in range(2)]
    for r in range(2):
        d = w.rank_desc(r)
        s = before[r]
        s[:, 1:-1, 1:-1] = u[:, :, d["i_begin"]:d["i_end"] + 1]
        w.set_rank_storage(r, orc.MOM_U, s)
    w.sync_interfaces(orc.MOM_U)
    s0, s1 = w.rank_storage(0, orc.MOM_U), w.rank_storage(1, orc.MOM_U)
    avg = 0.5 * (u[:, :, 3] + u[:, :, 3])  # internal interface: both ranks hold face 3
    np.testing.assert_array_equal(s0[:, 1:-1, 1 + 3], avg)
    np.testing.assert_array_equal(s1[:, 1:-1, 1 + 0], avg)
    np.testing.assert_array_equal(s0[:, 1:-1, 1 + 0], u[:, :, 0])  # wrap face untouched
    np.testing.assert_array_equal(s1[:, 1:-1, 1 + 3], u[:, :, 6])


# --------------------------------------------------------- layout/metrics ---

def test_idealized_domain_builder_layout(orc):
    # tests/unit/test_idealized_domain_builder.cpp:8-26
    w = orc.World(48, 24, 20, halo=2, ranks_x=2, ranks_y=2, dx=3000.0, dy=3000.0,
                  z_top=18000.0, precision=4)
    assert w.num_ranks == 4
    d = w.rank_desc(0)
    assert d["halo"] == 2 and d["i_end"] - d["i_begin"] == 24 and d["j_end"] - d["j_begin"] == 12
    zf = w.metric(orc.METRIC_FLAT_Z_INTERFACES)
    assert abs((zf[1] - zf[0]) - 900.0) < 1e-9


def test_layout_remainders_go_to_low_coordinates(orc):
    # virtual_rank_layout.cpp:13-29
    w = orc.World(10, 7, 1, ranks_x=3, ranks_y=2, precision=8)
    xs = [w.rank_desc(r)["i_end"] - w.rank_desc(r)["i_begin"] for r in range(3)]
    ys = [w.rank_desc(r)["j_end"] - w.rank_desc(r)["j_begin"] for r in (0, 3)]
    assert xs == [4, 3, 3] and ys == [4, 3]


def cosine_mountain(nx, ny, height, cx=0.5, cy=0.5, hx=0.15, hy=0.15):
    """idealized_domain_builder.cpp:14-35 (restated in numpy, FP64 then float)."""
    i = np.arange(nx) / max(nx - 1, 1)
    j = np.arange(ny) / max(ny - 1, 1)
    X, Y = np.meshgrid(i, j)
    dx, dy = (X - cx) / hx, (Y - cy) / hy
    out = np.where((np.abs(dx) >= 1) | (np.abs(dy) >= 1), 0.0,
                   height * 0.5 * (1 + np.cos(np.pi * dx)) * 0.5 * (1 + np.cos(np.pi * dy)))
    return out.astype(np.float32)


def test_hybrid_height_metrics_flat_and_mountain(orc):
    # tests/unit/test_hybrid_height_metrics.cpp:8-69
    w = orc.World(12, 10, 8, z_top=16000.0, precision=4)
    zi, zc, idz = (w.metric(orc.METRIC_Z_INTERFACES), w.metric(orc.METRIC_Z_CENTERS),
                   w.metric(orc.METRIC_INV_DZ_CELL))
    fzc = w.metric(orc.METRIC_FLAT_Z_CENTERS)
    assert abs(zi[0, 0, 0]) < 1e-6 and abs(zi[8, 0, 0] - 16000.0) < 1e-4
    assert abs(idz[0, 0, 0] - 1 / 2000.0) < 1e-7
    for k in range(8):
        assert abs(zc[k, 0, 0] - fzc[k]) < 1e-4
        assert zi[k + 1, 0, 0] > zi[k, 0, 0]
    t = cosine_mountain(21, 21, 1200.0, hx=0.25, hy=0.25)
    m = orc.World(21, 21, 12, z_top=18000.0, terrain=t, precision=4)
    zi, zc, idz = (m.metric(orc.METRIC_Z_INTERFACES), m.metric(orc.METRIC_Z_CENTERS),
                   m.metric(orc.METRIC_INV_DZ_CELL))
    ic = jc = 10
    assert abs(zi[0, jc, ic] - t[jc, ic]) < 1e-3 and abs(zi[12, jc, ic] - 18000.0) < 1e-4
    for k in range(12):
        assert zi[k + 1, jc, ic] > zi[k, jc, ic]
        assert zi[k, jc, ic] < zc[k, jc, ic] < zi[k + 1, jc, ic]
        assert idz[k, jc, ic] > 0
    fzc = m.metric(orc.METRIC_FLAT_Z_CENTERS)
    assert zc[0, jc, ic] - fzc[0] > zc[11, jc, ic] - fzc[11]
    assert m.metric(orc.METRIC_WEIGHT_CENTERS)[11] < 5e-2


def test_stretched_levels_reduce_to_uniform(orc):
    """Extension: explicit interfaces equal to k*z_top/nz reproduce the
    reference's uniform metrics to within one FP64 rounding."""
    zt, nz = 10000.0, 10
    t = cosine_mountain(9, 7, 500.0, hx=0.3, hy=0.3)
    a = orc.World(9, 7, nz, z_top=zt, terrain=t, precision=8)
    b = orc.World(9, 7, nz, z_top=zt, z_interfaces=np.arange(nz + 1) * zt / nz, terrain=t,
                  precision=8)
    for m in (orc.METRIC_Z_INTERFACES, orc.METRIC_Z_CENTERS, orc.METRIC_INV_DZ_CELL,
              orc.METRIC_INV_DZ_FACE):
        np.testing.assert_allclose(a.metric(m), b.metric(m), rtol=1e-12, atol=1e-9)


# --------------------------------------------------------------- thermo ---

def test_moist_thermo_closed_forms(orc):
    # tests/unit/test_moist_thermo.cpp:11-29 (the oracle's qsat vs its closed form)
    L = orc.lib()
    T, p = 300.0, 90000.0
    es = L.orc_thermo(4, 3, T, 0, 0)
    eps = L.orc_thermo(4, 8, 0, 0, 0)
    expected = eps * es / (p - (1 - eps) * es)
    assert abs(L.orc_thermo(4, 4, T, p, 0) - expected) < 1e-6
    # Magnus/Bolton form at 0 C is 6.112 hPa exactly.
    assert abs(L.orc_thermo(8, 3, 273.15, 0, 0) - 611.2) < 1e-9


def test_dry_eos_closed_forms(orc):
    L = orc.lib()
    for prec, tol in ((8, 1e-12), (4, 2e-6)):
        theta = 300.0
        rho = 1e5 / (287.05 * theta)          # rho R theta = p0  => p = p0
        assert abs(L.orc_thermo(prec, 1, rho, rho * theta, 0) / 1e5 - 1) < tol
        # rho(p(rho)) round trip: rho = p / (R theta Pi), Pi = (p/p0)^kappa
        for r in (0.3, 0.9, 1.2):
            p = L.orc_thermo(prec, 1, r, r * theta, 0)
            assert abs(L.orc_thermo(prec, 2, p, theta, 0) / r - 1) < 10 * tol
    # c_p = R / kappa
    assert abs(L.orc_thermo(8, 7, 0, 0, 0) - 287.05 / 0.2854) < 1e-9


# ---------------------------------------------------------- diagnostics ---

def test_dry_diagnostics_constant_state(orc):
    # tests/unit/test_dry_diagnostics.cpp:8-38
    w = orc.World(4, 3, 2, precision=4)
    w.init_constant(1.5, 300.0, 2.0, -1.0, 0.25)
    s = w.summary()
    assert abs(s["total_dry_mass"] - 4 * 3 * 2 * 1.5) < 1e-6
    assert abs(s["total_rho_theta_m"] - 4 * 3 * 2 * 1.5 * 300) < 1e-6
    assert abs(s["total_horizontal_momentum_x"] - 5 * 3 * 2 * 1.5 * 2) < 1e-6
    assert abs(s["total_horizontal_momentum_y"] - 4 * 4 * 2 * 1.5 * -1) < 1e-6
    assert abs(s["total_vertical_momentum"] - 4 * 3 * 3 * 1.5 * 0.25) < 1e-6
    assert s["rho_min"] == s["rho_max"] == 1.5
    assert abs(s["theta_min"] - 300) < 1e-6 and abs(s["theta_max"] - 300) < 1e-6


# ------------------------------------------------------------ operators ---

def test_pressure_gradient_known_answers(orc):
    # tests/unit/test_dry_pressure_gradient.cpp:8-55
    w = orc.World(6, 5, 2, periodic_x=False, periodic_y=False, precision=4)
    w.fill(orc.PRESSURE, 100000.0)
    w.pressure_gradient_only()
    tu, tv = w.get(orc.TEND_U), w.get(orc.TEND_V)
    assert abs(tu[0, 2, 1]) < 1e-6 and abs(tv[0, 1, 2]) < 1e-6
    s = w.rank_storage(0, orc.PRESSURE)
    k, jj, ii = np.mgrid[0:2, -1:6, -1:7]
    s[:] = (100000.0 + 10.0 * ii + 20.0 * jj).astype(np.float32)
    w.set_rank_storage(0, orc.PRESSURE, s)
    for f in (orc.TEND_U, orc.TEND_V):
        w.fill(f, 0.0)
    w.pressure_gradient_only()
    tu, tv = w.get(orc.TEND_U), w.get(orc.TEND_V)
    assert abs(tu[0, 2, 0]) < 1e-6
    assert abs(tu[0, 2, 1] - (-10.0 / 1000.0)) < 1e-6
    assert abs(tu[0, 2, 6]) < 1e-6
    assert abs(tv[0, 0, 2]) < 1e-6
    assert abs(tv[0, 1, 2] - (-20.0 / 1000.0)) < 1e-6
    assert abs(tv[0, 5, 2]) < 1e-6


def test_momentum_flux_constant_state_zero(orc):
    # tests/unit/test_dry_momentum_flux.cpp:30-84
    w = orc.World(20, 18, 8, ranks_x=2, ranks_y=2, precision=4)
    w.init_constant(1.05, 300.0, 12.0, -4.0, 0.5)
    for f in (orc.RHO, orc.MOM_U, orc.MOM_V, orc.MOM_W):
        w.exchange(f)
    w.momentum_flux_only()
    for r in range(4):
        for f in (orc.TEND_U, orc.TEND_V, orc.TEND_W):
            s = w.rank_storage(r, f)[:, 1:-1, 1:-1]
            assert abs(float(s.astype(np.float64).sum())) < 1e-5


def test_density_current_step_produces_u(orc):
    # tests/unit/test_dry_momentum_flux.cpp:86-131
    w = orc.World(20, 18, 8, ranks_x=2, ranks_y=2, precision=4)
    w.init_bubble(1.0, 300.0, center=(0.5, 0.5, 0.35), radius=(0.15, 0.15, 0.2),
                  theta_perturbation=3.0, sign=-1)
    w.ssprk3_step(orc.step_cfg(dt=2.0, fast_substeps=3))
    u = w.get(orc.MOM_U)
    assert np.all(np.isfinite(u)) and np.any(np.abs(u) > 1e-4)


def test_fast_modes_invariants(orc):
    # tests/unit/test_dry_fast_modes.cpp:49-192
    w = orc.World(20, 20, 8, ranks_x=2, ranks_y=2, precision=4)
    w.init_constant(1.1, 300.0, 12.0, -6.0, 0.0)
    b = w.summary()
    w.fast_modes(orc.step_cfg(dt=2.0, fast_substeps=4))
    a = w.summary()
    assert abs(b["total_dry_mass"] - a["total_dry_mass"]) < 1e-4
    for k in ("total_horizontal_momentum_x", "total_horizontal_momentum_y"):
        assert abs(a[k] - b[k]) < 1e-4
    for k in ("u_min", "u_max", "v_min", "v_max"):
        assert abs(a[k] - b[k]) < 1e-5

    def pulse_world():
        p = orc.World(20, 20, 8, ranks_x=2, ranks_y=2, precision=4)
        p.init_constant(1.0, 300.0)
        rho = p.get(orc.RHO)
        rho[4, 10, 10] += np.float32(0.02)
        p.set(orc.RHO, rho)
        p.set(orc.RHO_THETA, rho * np.float32(300.0))
        return p

    p = pulse_world()
    m0 = p.summary()["total_dry_mass"]
    for _ in range(3):
        p.fast_modes(orc.step_cfg(dt=1.0, fast_substeps=4))
    assert abs(p.summary()["total_dry_mass"] - m0) < 1e-3

    on, off = pulse_world(), pulse_world()
    on.fast_modes(orc.step_cfg(dt=1.0, fast_substeps=1, horizontal=True, density=False))
    off.fast_modes(orc.step_cfg(dt=1.0, fast_substeps=1, horizontal=False, density=False))
    so, sf = on.summary(), off.summary()
    assert max(abs(so["u_min"]), abs(so["u_max"])) > 1e-5
    assert max(abs(so["v_min"]), abs(so["v_max"])) > 1e-5
    assert max(abs(sf["u_min"]), abs(sf["u_max"])) < 1e-8
    assert max(abs(sf["v_min"]), abs(sf["v_max"])) < 1e-8
    for s in (so, sf):
        assert abs(s["total_dry_mass"] - m0) < 1e-5
        assert abs(s["rho_max"] - 1.02) < 1e-6 and abs(s["rho_min"] - 1.0) < 1e-6

    don, doff = pulse_world(), pulse_world()
    don.fast_modes(orc.step_cfg(dt=1.0, fast_substeps=1, horizontal=False, density=True))
    doff.fast_modes(orc.step_cfg(dt=1.0, fast_substeps=1, horizontal=False, density=False))
    sn, sf = don.summary(), doff.summary()
    assert abs(sn["total_dry_mass"] - m0) < 1e-5 and abs(sf["total_dry_mass"] - m0) < 1e-5
    assert abs(sf["rho_max"] - 1.02) < 1e-6 and abs(sf["rho_min"] - 1.0) < 1e-6
    assert abs(sn["rho_max"] - 1.02) > 1e-5


def test_boundary_strips(orc):
    # tests/unit/test_boundary_conditions.cpp:6-86
    w = orc.World(5, 4, 3, periodic_x=False, periodic_y=False, n_tracers=3, precision=4)
    w.init_constant(1.5, 305.0, 4.0, -3.0, 0.5)
    for t, v in enumerate((0.015, 0.001, 0.002)):
        w.fill(orc.TRACER0 + t, v)
    w.set_boundary_from_current()
    w.init_constant(1.0, 300.0)
    for t, v in enumerate((0.01, 0.0, 0.0)):
        w.fill(orc.TRACER0 + t, v)
    w.apply_boundary(orc.BOUNDARY_STRIPS)
    rho, rt, u = w.get(orc.RHO), w.get(orc.RHO_THETA), w.get(orc.MOM_U)
    v, wf, qv = w.get(orc.MOM_V), w.get(orc.MOM_W), w.get(orc.TRACER0)
    f32 = np.float32
    assert rho[0, 0, 0] == f32(1.5) and abs(rt[0, 0, 0] - f32(1.5) * f32(305)) < 1e-6
    assert u[0, 0, 0] == f32(1.5) * f32(4) and v[0, 0, 0] == f32(1.5) * f32(-3)
    assert wf[0, 0, 0] == f32(1.5) * f32(0.5)
    assert rho[2, 3, 4] == f32(1.5) and u[2, 3, 5] == f32(1.5) * f32(4)
    assert rho[1, 2, 2] == 1.0 and rt[1, 2, 2] == 300.0
    assert u[1, 2, 2] == 0 and v[1, 2, 2] == 0 and wf[1, 2, 2] == 0
    assert abs(qv[0, 0, 0] - 0.015) < 1e-6 and abs(qv[1, 2, 2] - 0.01) < 1e-6
    assert abs(w.get(orc.TRACER0 + 1)[0, 0, 0] - 0.001) < 1e-6
    assert abs(w.get(orc.TRACER0 + 2)[0, 0, 0] - 0.002) < 1e-6


# ------------------------------------------------------------ warm rain ---

def _single_cell(orc, rho, theta, qv, qc=0.0, qr=0.0, nz=1, metrics=None):
    w = orc.World(1, 1, nz, n_tracers=3, precision=4, **(metrics or {}))
    w.init_constant(rho, theta)
    for t, q in enumerate((qv, qc, qr)):
        w.set(orc.TRACER0 + t, np.float32(rho) * np.float32(q))
    return w


def _q(w, t):
    return float(w.get(pyoracle.TRACER0 + t)[0, 0, 0])


def test_warm_rain_saturation_adjustment(orc):
    # tests/unit/test_warm_rain_microphysics.cpp:38-71
    w = _single_cell(orc, 1.0, 300.0, 0.030)
    th0 = float(w.get(orc.RHO_THETA)[0, 0, 0])
    water0 = 0.030
    w.warm_rain(orc.warm_rain_cfg(dt=2.0, cloud_autoconversion_rate=0.0), metric_fallout=False)
    qv, qc, qr = _q(w, 0), _q(w, 1), _q(w, 2)
    assert qv < 0.030 and qc > 0.0 and abs(qr) < 1e-6
    assert abs(qv + qc + qr - water0) < 1e-5
    assert float(w.get(orc.RHO_THETA)[0, 0, 0]) > th0


def test_warm_rain_nominal_fallout_known_answer(orc):
    # tests/unit/test_warm_rain_microphysics.cpp:73-105
    w = orc.World(1, 1, 2, n_tracers=3, precision=4)
    w.init_constant(1.0, 295.0)
    qr = np.zeros((2, 1, 1), np.float32)
    qr[0], qr[1] = 0.004, 0.010
    w.set(orc.TRACER0 + 2, qr)
    w.warm_rain(orc.warm_rain_cfg(dt=1.0, cloud_autoconversion_rate=0.0, rain_evaporation_rate=0.0,
                                  rain_terminal_velocity=250.0, sedimentation_layer_depth_m=1000.0,
                                  enable_latent_heating=False), metric_fallout=False)
    out = w.get(orc.TRACER0 + 2)
    assert abs(out[1, 0, 0] - 0.0075) < 1e-6 and abs(out[0, 0, 0] - 0.0055) < 1e-6
    assert abs(w.get(orc.PRECIP)[0, 0, 0] - 1.0) < 1e-6


def test_warm_rain_autoconversion(orc):
    # tests/unit/test_warm_rain_microphysics.cpp:107-137
    w = _single_cell(orc, 1.0, 295.0, 0.006, 0.002, 0.001)
    w.warm_rain(orc.warm_rain_cfg(dt=2.0, cloud_autoconversion_threshold=1e-4,
                                  cloud_autoconversion_rate=1.0, rain_evaporation_rate=0.0,
                                  enable_latent_heating=False), metric_fallout=False)
    qv, qc, qr = _q(w, 0), _q(w, 1), _q(w, 2)
    assert qc < 0.002 and qr > 0.001
    assert abs(qv + qc + qr - 0.009) < 1e-5


def test_warm_rain_metric_fallout_known_answer(orc):
    # tests/unit/test_warm_rain_microphysics.cpp:139-161
    w = _single_cell(orc, 1.0, 295.0, 0.0, 0.0, 0.010,
                     metrics=dict(dx=1000.0, dy=1000.0, z_top=1000.0))
    w.warm_rain(orc.warm_rain_cfg(dt=1.0, cloud_autoconversion_rate=0.0, rain_evaporation_rate=0.0,
                                  rain_terminal_velocity=1500.0, enable_latent_heating=False),
                metric_fallout=True)
    assert abs(_q(w, 2)) < 1e-6
    assert abs(w.get(orc.PRECIP)[0, 0, 0] - 10.0) < 1e-6


def test_warm_rain_closed_box(orc):
    # tests/regression/test_warm_rain_closed_box.cpp:42-151
    w = orc.World(4, 3, 2, n_tracers=3, precision=4)
    w.init_constant(1.0, 300.0)
    qv = np.full((2, 3, 4), 0.020, np.float32)
    qv[:, 1, :] = 0.030
    w.set(orc.TRACER0, qv)
    water0 = sum(w.get(orc.TRACER0 + t).astype(np.float64).sum() for t in range(3))
    rt0 = w.summary()["total_rho_theta_m"]
    cfg = orc.warm_rain_cfg(dt=2.0, cloud_autoconversion_rate=0.20, rain_evaporation_rate=0.0)
    w.warm_rain(cfg, metric_fallout=False)
    w.warm_rain(cfg, metric_fallout=False)
    water1 = sum(w.get(orc.TRACER0 + t).astype(np.float64).sum() for t in range(3))
    assert abs(water1 - water0) < 1e-4
    assert w.summary()["total_rho_theta_m"] > rt0
    assert w.get(orc.TRACER0 + 1).sum() + w.get(orc.TRACER0 + 2).sum() > 0
    assert w.get(orc.TRACER0 + 2).sum() > 0
    # fallout closure (test_warm_rain_closed_box.cpp:115-149)
    f = orc.World(2, 1, 2, dx=1000.0, dy=1000.0, z_top=2000.0, n_tracers=3, precision=4)
    f.init_constant(1.0, 300.0)
    qr = np.zeros((2, 1, 2), np.float32)
    qr[1], qr[0] = 0.010, 0.004
    f.set(orc.TRACER0 + 2, qr)
    idz = f.metric(orc.METRIC_INV_DZ_CELL).astype(np.float64)

    def column_water():
        return sum((f.get(orc.TRACER0 + t).astype(np.float64) / idz).sum() for t in range(3))

    before = column_water()
    f.warm_rain(orc.warm_rain_cfg(dt=1.0, condensation_relaxation=0.0, cloud_autoconversion_rate=0.0,
                                  rain_evaporation_rate=0.0, rain_terminal_velocity=1000.0,
                                  enable_latent_heating=False), metric_fallout=True)
    surface = f.get(orc.PRECIP).astype(np.float64).sum()
    assert abs(column_water() + surface - before) < 1e-5 and surface > 0


# ----------------------------------------------------------- regressions ---

def test_hydrostatic_rest(orc):
    # tests/regression/test_hydrostatic_rest.cpp:9-53
    w = orc.World(18, 18, 12, ranks_x=2, ranks_y=2, dx=3000.0, dy=3000.0, z_top=18000.0,
                  precision=4)
    w.init_hydrostatic_rest(1.22, 300.0)
    b = w.summary()
    w.ssprk3_step(orc.step_cfg(dt=1.0, fast_substeps=3))
    a = w.summary()
    assert abs(b["total_dry_mass"] - a["total_dry_mass"]) < 1e-3
    assert abs(b["total_rho_theta_m"] - a["total_rho_theta_m"]) < 3e-2
    assert abs(a["total_horizontal_momentum_x"]) < 1e-4
    assert abs(a["total_horizontal_momentum_y"]) < 1e-4


def test_terrain_hydrostatic_rest(orc):
    # tests/regression/test_terrain_hydrostatic_rest.cpp:9-60
    t = cosine_mountain(32, 24, 900.0, hx=0.15, hy=0.30)
    w = orc.World(32, 24, 16, ranks_x=2, ranks_y=2, dx=2000.0, dy=2000.0, z_top=18000.0,
                  terrain=t, precision=4)
    w.init_hydrostatic_rest(1.18, 300.0)
    b = w.summary()
    w.ssprk3_step(orc.step_cfg(dt=0.5, fast_substeps=4))
    a = w.summary()
    assert abs(b["total_dry_mass"] - a["total_dry_mass"]) < 2e-2
    assert abs(b["total_rho_theta_m"] - a["total_rho_theta_m"]) < 1.0
    assert abs(a["total_horizontal_momentum_x"]) < 5e-3
    assert abs(a["total_horizontal_momentum_y"]) < 5e-3
    assert a["w_max"] < 5e-2 and a["w_min"] > -5e-2


def test_acoustic_pulse_symmetry(orc):
    # tests/regression/test_acoustic_pulse.cpp:10-60
    w = orc.World(33, 33, 9, precision=4)
    w.init_constant(1.0, 300.0)
    k, j, i = np.mgrid[0:9, 0:33, 0:33].astype(np.float32)
    r2 = (i - 16) ** 2 + (j - 16) ** 2 + (k - 4) ** 2
    rho = (np.float32(1.0) + np.float32(0.01) * np.exp(-r2 / np.float32(6.0))).astype(np.float32)
    w.set(orc.RHO, rho)
    w.set(orc.RHO_THETA, rho * np.float32(300.0))
    m0 = w.summary()["total_dry_mass"]
    for _ in range(8):
        w.fast_modes(orc.step_cfg(dt=0.5, fast_substeps=6))
    s = w.summary()
    r = w.get(orc.RHO)
    assert abs(s["total_dry_mass"] - m0) < 1e-3
    assert s["u_max"] > 1e-4 and s["v_max"] > 1e-4
    assert abs(r[4, 16, 18] - r[4, 16, 14]) < 5e-4
    assert abs(r[4, 18, 16] - r[4, 14, 16]) < 5e-4
    assert abs(s["total_horizontal_momentum_x"]) < 1e-3
    assert abs(s["total_horizontal_momentum_y"]) < 1e-3


def _bubble_world(orc, sign, precision=4, ranks=2):
    w = orc.World(32, 32, 16, ranks_x=ranks, ranks_y=ranks, precision=precision)
    w.init_bubble(1.0, 300.0, center=(0.5, 0.5, 0.35), radius=(0.12, 0.12, 0.16),
                  theta_perturbation=3.0, sign=sign)
    return w


def test_buoyancy_response(orc):
    # tests/regression/test_buoyancy_response.cpp:11-81
    warm, cold = _bubble_world(orc, +1), _bubble_world(orc, -1)
    for wld in (warm, cold):
        s = wld.summary()
        assert abs(s["w_max"]) < 1e-6 and abs(s["w_min"]) < 1e-6
    warm.ssprk3_step(orc.step_cfg(dt=2.0, fast_substeps=3))
    cold.ssprk3_step(orc.step_cfg(dt=2.0, fast_substeps=3))
    assert warm.summary()["w_max"] > 1e-3
    assert cold.summary()["w_min"] < -1e-3


def test_horizontal_pressure_response(orc):
    # tests/regression/test_horizontal_pressure_response.cpp:9-55
    w = _bubble_world(orc, +1)
    m0 = w.summary()["total_dry_mass"]
    for _ in range(4):
        w.ssprk3_step(orc.step_cfg(dt=2.0, fast_substeps=3))
    s = w.summary()
    assert abs(s["total_dry_mass"] - m0) < 1e-2
    assert s["u_max"] > 1e-3 and s["u_min"] < -1e-3 and s["v_max"] > 1e-3 and s["v_min"] < -1e-3
    assert abs(s["total_horizontal_momentum_x"]) < 1e-2
    assert abs(s["total_horizontal_momentum_y"]) < 1e-2


def test_density_current_evolution(orc):
    # tests/regression/test_density_current_evolution.cpp:13-87
    w = orc.World(32, 32, 16, ranks_x=2, ranks_y=2, precision=4)
    w.init_bubble(1.0, 300.0, center=(0.5, 0.5, 0.30), radius=(0.12, 0.12, 0.14),
                  theta_perturbation=3.0, sign=-1)
    m0 = w.summary()["total_dry_mass"]
    for _ in range(5):
        w.ssprk3_step(orc.step_cfg(dt=2.0, fast_substeps=3))
    s = w.summary()
    assert abs(s["total_dry_mass"] - m0) < 2e-2
    assert s["u_max"] > 1e-3 and s["u_min"] < -1e-3
    assert np.all(np.isfinite(w.get(orc.RHO)))


def test_dry_constant_state_bundle(orc):
    # tests/regression/test_dry_constant_state_bundle.cpp:8-38
    w = orc.World(20, 16, 6, ranks_x=2, ranks_y=2, precision=4)
    w.init_constant(1.1, 301.0, 5.0, -2.0, 0.0)
    w.ssprk3_step(orc.step_cfg(dt=2.0, fast_substeps=2))
    assert np.max(np.abs(w.get(orc.RHO) - 1.1)) < 1e-5
    assert np.max(np.abs(w.get(orc.RHO_THETA) - 1.1 * 301.0)) < 1e-3


def test_mountain_wave_evolution(orc):
    # tests/regression/test_mountain_wave_evolution.cpp:10-86
    worlds = {}
    for name, h in (("flat", 0.0), ("mountain", 1200.0)):
        t = cosine_mountain(96, 24, h, hx=0.10, hy=0.40) if h > 0 else None
        w = orc.World(96, 24, 24, ranks_x=2, ranks_y=2, dx=1000.0, dy=1000.0, z_top=18000.0,
                      terrain=t, precision=4)
        w.init_mountain_wave(1.18, 300.0, 15.0)
        worlds[name] = w
    m0 = {k: w.summary()["total_dry_mass"] for k, w in worlds.items()}
    for _ in range(6):
        for w in worlds.values():
            w.ssprk3_step(orc.step_cfg(dt=1.0, fast_substeps=4))
    sf, sm = worlds["flat"].summary(), worlds["mountain"].summary()
    assert abs(sf["total_dry_mass"] - m0["flat"]) < 5e-2
    assert abs(sm["total_dry_mass"] - m0["mountain"]) < 1.0
    assert sf["w_max"] < 2e-2 and sf["w_min"] > -2e-2
    assert sm["w_max"] > 1e-2 or sm["w_min"] < -1e-2
    assert np.isfinite(sm["total_dry_mass"])


def test_scalar_mass_conservation_tracers(orc):
    # tests/regression/test_scalar_mass_conservation.cpp:42-111
    w = orc.World(32, 20, 4, ranks_x=2, ranks_y=2, n_tracers=1, precision=4)
    w.init_constant(1.1, 300.0, 18.0, -6.0, 0.0)
    k, j, i = np.mgrid[0:4, 0:20, 0:32]
    qv = (np.float32(0.003) + np.float32(0.0002) * ((i + 2 * j + k) % 9).astype(np.float32))
    w.set(orc.TRACER0, np.float32(1.1) * qv.astype(np.float32))
    m0 = w.get(orc.TRACER0).astype(np.float64).sum()
    w.tracers_ssprk3(orc.step_cfg(dt=2.0))
    q = w.get(orc.TRACER0) / w.get(orc.RHO)
    assert abs(w.get(orc.TRACER0).astype(np.float64).sum() - m0) < 1e-3
    assert q.min() >= 0.0 and q.max() <= 0.0047


def test_warm_rain_summary_closure(orc):
    # tests/regression/test_warm_rain_summary_closure.cpp:60-124
    w = orc.World(3, 2, 2, dx=1000.0, dy=1000.0, z_top=2000.0, n_tracers=3, precision=4)
    w.init_constant(1.0, 298.0)
    k, j, i = np.mgrid[0:2, 0:2, 0:3]
    w.set(orc.TRACER0, np.where(i == 1, 0.030, 0.020).astype(np.float32))
    idz = w.metric(orc.METRIC_INV_DZ_CELL).astype(np.float64)
    before_w = sum((w.get(orc.TRACER0 + t).astype(np.float64) / idz).sum() for t in range(3))
    rt0 = w.summary()["total_rho_theta_m"]
    cfg = orc.warm_rain_cfg(dt=2.0, cloud_autoconversion_rate=0.20, rain_evaporation_rate=0.0,
                            rain_terminal_velocity=2.0)
    w.warm_rain(cfg, metric_fallout=True)
    w.warm_rain(cfg, metric_fallout=True)
    s = w.summary()
    after_w = s["vapor_path"] + s["cloud_path"] + s["rain_path"]
    assert abs(after_w + s["precip_sum"] - before_w) < 1e-3
    assert s["total_rho_theta_m"] > rt0
    assert s["cloud_path"] + s["rain_path"] > 0 and s["precip_sum"] > 0
    assert s["qv_min"] >= -1e-8 and s["qc_min"] >= -1e-8 and s["qr_min"] >= -1e-8
    assert s["precip_max"] > 0 and s["precip_mean"] > 0
