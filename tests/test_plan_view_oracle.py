"""The oracle's plan-view products (oracle/gwm_oracle.cpp plan_view), pinned
on the CPU by the reference's own known answers
(tests/unit/test_plan_view_output.cpp:55-210): a constant state on a 6x4x3
grid split over two virtual ranks, 9 km top, slice level 1, 4 steps of 2 s;
and by closed forms of the EOS, the reflectivity and the column paths."""
import math

import numpy as np
import pytest

import pyoracle as O

NAMES = O.PLAN_VIEW_NAMES


def constant_world(n_tracers, precision=4):
    w = O.World(6, 4, 3, ranks_x=2, ranks_y=1, dx=1000.0, dy=1000.0, z_top=9000.0,
                n_tracers=n_tracers, precision=precision)
    w.init_constant(1.2, 300.0, 10.0, -4.0, 0.5)
    return w


def field(out, name):
    return out[NAMES.index(name)]


@pytest.mark.parametrize("precision", [4, 8])
def test_dry_bundle_known_answers(precision):
    w = constant_world(0, precision)
    nf, out = w.plan_view(1, 4, 2.0)
    assert nf == 10
    assert np.all(out[nf:] == 0.0)
    assert np.allclose(field(out, "terrain_height"), 0.0)
    assert np.allclose(field(out, "z_center"), 4500.0, atol=1e-6)
    assert np.allclose(field(out, "rho_d"), 1.2, atol=1e-6)
    assert np.allclose(field(out, "theta_m"), 300.0, atol=1e-4)
    assert np.allclose(field(out, "u_velocity"), 10.0, atol=1e-5)
    assert np.allclose(field(out, "v_velocity"), -4.0, atol=1e-5)
    assert np.allclose(field(out, "w_velocity"), 0.5, atol=1e-6)
    assert np.allclose(field(out, "wind_speed"), math.sqrt(116.0), atol=1e-5)
    # closed forms of dry_thermo.hpp: p = p0 (rho R theta / p0)^(1/(1-kappa)),
    # T = theta (p/p0)^kappa
    p = 1e5 * (1.2 * 287.05 * 300.0 / 1e5) ** (1.0 / (1.0 - 0.2854))
    T = 300.0 * (p / 1e5) ** 0.2854
    tol = 1e-2 if precision == 4 else 1e-8
    assert np.allclose(field(out, "air_pressure"), p, rtol=1e-6, atol=tol)
    assert np.allclose(field(out, "air_temperature"), T, rtol=1e-6, atol=1e-4)


@pytest.mark.parametrize("precision", [4, 8])
def test_moist_bundle_known_answers(precision):
    w = constant_world(3, precision)
    rho = w.get(O.RHO)
    for t, q in enumerate((0.004, 0.001, 0.0005)):
        w.set(O.TRACER0 + t, rho * q)
    pr = np.zeros((1, 4, 6))
    pr[0, 0, 0] = 0.25
    w.set(O.PRECIP, pr)
    nf, out = w.plan_view(1, 4, 2.0, with_precip=True)
    assert nf == 23
    T = field(out, "air_temperature")[0, 0]
    assert abs(field(out, "specific_humidity")[0, 0] - 0.004) < 1e-6
    rh = field(out, "relative_humidity")
    assert np.all(rh > 0.0) and np.all(rh <= 100.0)
    assert np.all(field(out, "dewpoint") < T)
    assert abs(field(out, "cloud_water_mixing_ratio")[0, 0] - 0.001) < 1e-6
    assert abs(field(out, "rain_water_mixing_ratio")[0, 0] - 0.0005) < 1e-6
    assert abs(field(out, "total_condensate")[0, 0] - 0.0015) < 1e-6
    # column paths: rho q times the 9 km column
    assert np.allclose(field(out, "column_cloud_water"), 10.8, atol=1e-5)
    assert np.allclose(field(out, "column_rain_water"), 5.4, atol=1e-5)
    assert np.allclose(field(out, "column_total_condensate"), 16.2, atol=1e-5)
    assert np.allclose(field(out, "column_rain_fraction"), 1.0 / 3.0, atol=1e-6)
    # Z = 3.63e9 (rho qr)^1.75, dBZ = 10 log10 Z
    dbz = 10.0 * math.log10(3.63e9 * (1.2 * 0.0005) ** 1.75)
    assert np.allclose(field(out, "synthetic_reflectivity"), dbz, atol=1e-3)
    assert np.all(field(out, "synthetic_reflectivity") <= 75.0)
    acc = field(out, "accumulated_surface_precipitation")
    assert abs(acc[0, 0] - 0.25) < 1e-6 and np.all(acc.ravel()[1:] == 0.0)
    # 0.25 mm over 4 steps of 2 s: 112.5 mm/h
    assert abs(field(out, "mean_surface_precipitation_rate")[0, 0] - 112.5) < 1e-4
    nf, _ = w.plan_view(1, 4, 2.0, with_precip=False)
    assert nf == 21


def test_vapour_only_bundle_and_slice_clamp():
    w = constant_world(1)
    w.set(O.TRACER0, w.get(O.RHO) * 0.004)
    nf, out = w.plan_view(7, 4, 2.0)
    assert nf == 13
    # slice level clamped to nz - 1 = 2: z = 7500 m
    assert np.allclose(field(out, "z_center"), 7500.0, atol=1e-6)
    assert abs(field(out, "specific_humidity")[0, 0] - 0.004) < 1e-6


def test_dewpoint_and_rh_closed_forms():
    # dewpoint inverts the Magnus form: e_s(Td) = e, with e = p q / (eps + (1 - eps) q)
    w = constant_world(1, precision=8)
    w.set(O.TRACER0, w.get(O.RHO) * 0.004)
    nf, out = w.plan_view(0, 1, 1.0)
    p = field(out, "air_pressure")[0, 0]
    T = field(out, "air_temperature")[0, 0]
    eps = 287.05 / 461.5
    e = p * 0.004 / (eps + (1 - eps) * 0.004)
    td = field(out, "dewpoint")[0, 0] - 273.15
    es_td = 6.112 * math.exp(17.67 * td / (td + 243.5)) * 100.0
    assert abs(es_td - e) < 1e-9 * e
    tc = T - 273.15
    es = 6.112 * math.exp(17.67 * tc / (tc + 243.5)) * 100.0
    qs = eps * es / max(p - (1 - eps) * es, 1000.0)
    assert abs(field(out, "relative_humidity")[0, 0] - 100.0 * 0.004 / qs) < 1e-9
