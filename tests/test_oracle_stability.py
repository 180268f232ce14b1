"""Records a property of the reference scheme: its split-explicit step
diverges on the benchmark configurations.

The oracle restates the reference (and the HIP path matches the oracle to
1e-11 in FP64), so these runs show what the reference itself computes:
finite for the few steps of the reference's own regression tests, divergent
for the 64x64x40 warm bubble at dz = 250 m, dt = 2 s, 3 substeps.
"""
import numpy as np

import pyoracle as O


def c1_world(nx=64, ny=64):
    # 64x64x40, dx = dy = 500 m, dz = 250 m, hydrostatic theta = 300 K with
    # p_sfc = 1000 hPa, +2 K bubble of radius 2 km at (16 km, 16 km, 2 km).
    w = O.World(nx, ny, 40, dx=500.0, dy=500.0, z_top=10000.0, precision=8)
    rho_s = 1e5 / (287.05 * 300.0)
    w.init_bubble(rho_background=-rho_s, theta_background=300.0,
                  center=(16000.0 / ((nx - 1) * 500.0), 16000.0 / ((ny - 1) * 500.0), 0.2),
                  radius=(2000.0 / (nx * 500.0), 2000.0 / (ny * 500.0), 0.2),
                  theta_perturbation=2.0)
    return w


def test_reference_scheme_diverges_on_the_warm_bubble_case():
    w = c1_world()
    finite_steps = 0
    for _ in range(10):
        w.ssprk3_step(O.step_cfg(dt=2.0, fast_substeps=3))
        s = w.summary()
        if not (np.isfinite(s["w_max"]) and np.isfinite(s["w_min"])):
            break
        finite_steps += 1
    assert finite_steps < 10, "the reference scheme stayed finite on C1"


def test_reference_regression_window_stays_finite():
    # The reference's own regime: dz = 1000 m, dt = 2 s, 3 substeps, 5 steps
    # (tests/regression/test_density_current_evolution.cpp:13-49).
    w = O.World(32, 32, 16, ranks_x=2, ranks_y=2, precision=8)
    w.init_bubble(1.0, 300.0, center=(0.5, 0.5, 0.30), radius=(0.12, 0.12, 0.14),
                  theta_perturbation=3.0, sign=-1)
    for _ in range(5):
        w.ssprk3_step(O.step_cfg(dt=2.0, fast_substeps=3))
    s = w.summary()
    assert np.isfinite(s["w_max"]) and abs(s["w_max"]) < 50.0
