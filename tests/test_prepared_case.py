"""The prepared-case path: gwm_prepared_case_driver and its start-up.

CPU:
- the driver's flag errors follow the reference's flag contract
  (gwm_prepared_case_driver.cpp:169-250) and its input errors name the
  document, key and problem (the checks of runtime_case.cpp:151-274);
- the prepared-case oracle (oracle/prepared_case_oracle.py) is pinned by
  closed forms: the EOS at rho = 1 returns the pressure rho*theta was built
  from; every level of a balanced column satisfies the trapezoid
  hydrostatic relation with the level below and the EOS with its theta; the
  cell mass divergence of the momenta is their column-mean horizontal
  divergence, which the barotropic projection of open domains removes, and
  w = 0 at the ground and the lid; the SOR potential solves the 5-point
  Poisson problem (checked against a dense solve).
GPU:
- the driver's start-up balance meets the reference's bounds
  (tests/unit/test_prepared_case_init.cpp), and its initial and final
  summaries agree with the oracle: the balanced start-up of
  prepared_case_oracle, stepped by the C++ oracle with the strip boundary
  of the balanced boundary snapshot (vapour strips only, as the prepared
  tracer updater) and warm rain with metric fallout.
"""
import json
import shutil
import os
import subprocess

import numpy as np
import pytest

import prepared_case_fixture as PF
import prepared_case_oracle as PO
import pyoracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "gpu-wm-next_amd", "host", "bin", "gwm_prepared_case_driver")


def _need_driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (run __graft_entry__.build())")


def _run(args, timeout=120):
    _need_driver()
    return subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=timeout)


# --------------------------------------------------------------------------- CPU


@pytest.mark.parametrize("args,msg", [
    (["--bogus", "1"], "Unknown argument: --bogus"),
    (["--steps"], "Missing value for --steps"),
    (["--boundary-cache", "b.json"], "--analysis-state is required"),
    (["--analysis-state", "a.json"], "--boundary-cache is required"),
    (["--analysis-state", "a", "--boundary-cache", "b", "--enable-warm-rain", "maybe"],
     "Invalid boolean value: maybe"),
    (["--analysis-state", "a", "--boundary-cache", "b", "--steps", "-1"],
     "--steps must be nonnegative"),
])
def test_driver_rejects_bad_flags(args, msg):
    r = _run(args)
    assert r.returncode == 1
    assert "gwm_prepared_case_driver error: " + msg in r.stderr, r.stderr


def _mutate(path, fn):
    with open(path) as fh:
        d = json.load(fh)
    fn(d)
    with open(path, "w") as fh:
        json.dump(d, fh)


@pytest.mark.parametrize("mutation,msg", [
    ("truncate", "JSON parse error"),
    ("schema", "declares schema 'gwm-next-analysis-state/v0', expected "
               "'gwm-next-analysis-state/v1'"),
    ("source", ".source names unknown source 'NOAA'"),
    ("missing", "analysis atmosphere lacks the canonical field 'v_wind' (HRRR VGRD)"),
    ("size", "analysis atmosphere field 'u_wind' holds 119 values; the grid needs 120"),
    ("gfs", "analysis source GFS has no runtime field bindings"),
    ("order", "boundary snapshots must have strictly increasing forecast offsets "
              "(snapshot 1 at 0 s follows 3600 s)"),
    ("mismatch", "analysis grid 6x5x4 and boundary-cache grid 5x5x4 differ"),
])
def test_driver_rejects_bad_inputs(tmp_path, mutation, msg):
    snaps = ((0, 1.0, 1.5, 0.001), (3600, 2.0, 3.0, 0.002))
    pa, pb, _ = PF.write_case(str(tmp_path), grid=dict(PF.GRID, nx=6, ny=5, nz=4),
                              snapshots=snaps)
    if mutation == "truncate":
        with open(pa) as fh:
            text = fh.read()
        with open(pa, "w") as fh:
            fh.write(text[: len(text) // 2])
    elif mutation == "schema":
        _mutate(pa, lambda d: d.update(schema_version="gwm-next-analysis-state/v0"))
    elif mutation == "source":
        _mutate(pa, lambda d: d.update(source="NOAA"))
    elif mutation == "missing":
        _mutate(pa, lambda d: d["atmosphere"].pop("v_wind"))
    elif mutation == "size":
        _mutate(pa, lambda d: d["atmosphere"].update(u_wind=d["atmosphere"]["u_wind"][:-1]))
    elif mutation == "gfs":
        _mutate(pa, lambda d: d.update(source="GFS"))
    elif mutation == "order":
        _mutate(pb, lambda d: d["snapshots"].reverse())
    elif mutation == "mismatch":
        def shrink(d):
            g = d["grid"]
            g["nx"] -= 1
            for s in d["snapshots"]:
                for b in ("atmosphere", "surface"):
                    for k, v in s[b].items():
                        n = len(v) // (g["nx"] + 1) * g["nx"]
                        s[b][k] = v[:n]
        _mutate(pb, shrink)
    r = _run(["--analysis-state", pa, "--boundary-cache", pb, "--steps", "0"])
    assert r.returncode == 1
    assert msg in r.stderr, r.stderr


def _metrics(grid, terrain):
    w = O.World(grid["nx"], grid["ny"], grid["nz"], periodic_x=False, periodic_y=False,
                dx=grid["dx"], dy=grid["dy"], z_top=grid["z_top"], terrain=terrain, precision=4)
    return w.metric(1), w.metric(2)  # z_centers, inv_dz_cell


def test_oracle_rho_theta_reproduces_pressure():
    p = np.linspace(2.0e4, 1.05e5, 400).astype(np.float32)
    rt = PO.rho_theta_from_pressure(p)
    back = PO.pressure_from_rho_theta(np.ones_like(rt), rt)
    assert np.max(np.abs(back - p) / p) < 2e-6


def test_oracle_balanced_columns_are_hydrostatic():
    grid = dict(PF.GRID, nx=9, ny=7, nz=10)
    atm, _, static = PF.fields(grid, seed=3)
    zc, _ = _metrics(grid, static["terrain_height"])
    p, th, rho, rt = PO.balanced_columns(atm["air_pressure"], atm["air_temperature"], zc)
    ps = np.maximum(atm["air_pressure"].astype(np.float32), 1.0)
    th_src = atm["air_temperature"] * (1.0e5 / ps.astype(np.float64)) ** 0.2854
    assert np.allclose(th, th_src, rtol=2e-6)
    assert np.allclose(p[0], ps[0], rtol=0)
    # EOS with the level's theta, and the trapezoid relation with the level below
    assert np.allclose(rho, PO.rho_from_pressure_theta(p, th), rtol=0)
    dz = (zc[1:] - zc[:-1]).astype(np.float64)
    resid = p[1:] - (p[:-1] - 9.81 * 0.5 * (rho[:-1] + rho[1:]) * dz)
    assert np.max(np.abs(resid) / p[1:]) < 2e-6
    assert np.allclose(PO.pressure_from_rho_theta(np.ones_like(rt), rt), p, rtol=2e-6)


def test_oracle_sor_solves_poisson():
    rng = np.random.default_rng(0)
    ny, nx, dx, dy = 5, 6, 2000.0, 1500.0
    rhs = (1e-4 * rng.standard_normal((ny, nx))).astype(np.float32)
    phi = PO.sor_potential(rhs, dx, dy)
    n = nx * ny
    A = np.zeros((n, n))
    for j in range(ny):
        for i in range(nx):
            r = j * nx + i
            A[r, r] = -2.0 / dx ** 2 - 2.0 / dy ** 2
            for (jj, ii, c) in ((j, i - 1, dx), (j, i + 1, dx), (j - 1, i, dy), (j + 1, i, dy)):
                if 0 <= jj < ny and 0 <= ii < nx:
                    A[r, jj * nx + ii] = 1.0 / c ** 2
    exact = np.linalg.solve(A, rhs.astype(np.float64).ravel()).reshape(ny, nx)
    assert np.max(np.abs(phi - exact)) <= 1e-4 * np.max(np.abs(exact))


@pytest.mark.parametrize("periodic", [False, True])
def test_oracle_face_momenta_are_mass_balanced(periodic):
    grid = dict(PF.GRID, nx=10, ny=8, nz=6)
    atm, _, static = PF.fields(grid, seed=5)
    zc, idz = _metrics(grid, static["terrain_height"])
    _, _, rho, _ = PO.balanced_columns(atm["air_pressure"], atm["air_temperature"], zc)
    mu, mv, mw = PO.face_momenta(rho, atm["u_wind"], atm["v_wind"], idz, grid["dx"], grid["dy"],
                                 periodic, periodic)
    assert np.all(mw[0] == 0) and np.all(mw[-1] == 0)
    mean, hdiv, dz = PO.column_mean_divergence(mu, mv, idz, grid["dx"], grid["dy"])
    div = hdiv + (mw[1:] - mw[:-1]) * idz
    scale = np.max(np.abs(hdiv))
    # w takes the divergence less its column mean, so the cell divergence is
    # the column mean; the barotropic projection (open domains) removes that
    assert np.max(np.abs(div - mean[None])) <= 1e-5 * scale
    if not periodic:
        assert np.max(np.abs(mean)) <= 1e-4 * scale
        assert np.max(np.abs(div)) <= 1e-4 * scale
    # the interior faces average the balanced density and the analysis wind
    i = grid["nx"] // 2
    ref = 0.5 * (rho[:, :, i - 1] + rho[:, :, i]) * 0.5 * (atm["u_wind"][:, :, i - 1] +
                                                          atm["u_wind"][:, :, i])
    if periodic:
        assert np.allclose(mu[:, :, i], ref, rtol=2e-6)


# --------------------------------------------------------------------------- GPU

SUM_MAP = [("total_dry_mass", "total_dry_mass"), ("total_rho_theta_m", "total_rho_theta_m"),
           ("total_horizontal_momentum_x", "total_horizontal_momentum_x"),
           ("total_horizontal_momentum_y", "total_horizontal_momentum_y"),
           ("total_vertical_momentum", "total_vertical_momentum"),
           ("rho_min", ("rho_d", "min")), ("rho_max", ("rho_d", "max")),
           ("theta_min", ("theta_m", "min")), ("theta_max", ("theta_m", "max")),
           ("u_min", ("u_face", "min")), ("u_max", ("u_face", "max")),
           ("v_min", ("v_face", "min")), ("v_max", ("v_face", "max")),
           ("vapor_path", ("moisture", "vapor_water_path_sum_kg_m2")),
           ("cloud_path", ("moisture", "cloud_water_path_sum_kg_m2")),
           ("rain_path", ("moisture", "rain_water_path_sum_kg_m2")),
           ("precip_sum", ("moisture", "accumulated_surface_precipitation_sum_mm"))]


def _get(js, key):
    if isinstance(key, tuple):
        return js[key[0]][key[1]]
    return js[key]


def _oracle_run(gen, grid, steps, cfg, wr, boundaries):
    atm, _, static = gen["analysis"]
    w = O.World(grid["nx"], grid["ny"], grid["nz"], periodic_x=False, periodic_y=False,
                dx=grid["dx"], dy=grid["dy"], z_top=grid["z_top"],
                terrain=static["terrain_height"], n_tracers=3, precision=4)
    zc, idz = w.metric(1), w.metric(2)

    def load(state):
        for f, key in ((O.RHO, "rho"), (O.RHO_THETA, "rho_theta"), (O.MOM_U, "mom_u"),
                       (O.MOM_V, "mom_v"), (O.MOM_W, "mom_w"), (O.TRACER0, "qv"),
                       (O.TRACER0 + 1, "qc"), (O.TRACER0 + 2, "qr")):
            w.set(f, state[key])

    if boundaries:
        (_, batm, _), = gen["snapshots"]
        load(PO.prepared_state(batm, zc, idz, grid["dx"], grid["dy"]))
        w.set_boundary_from_current()
        w.set_tracer_strip_mask(0b001)  # the snapshot carries vapour only
    load(PO.prepared_state(atm, zc, idz, grid["dx"], grid["dy"]))
    mode = O.BOUNDARY_STRIPS if boundaries else O.BOUNDARY_NONE
    initial = w.summary()
    for _ in range(steps):
        w.ssprk3_step(cfg, mode)
        w.tracers_ssprk3(cfg, mode)
        w.warm_rain(wr, metric_fallout=True)
    return initial, w.summary(), w


@pytest.mark.gpu
@pytest.mark.parametrize("boundaries", [True, False])
def test_prepared_driver_matches_oracle(tmp_path, boundaries):
    grid = PF.GRID
    pa, pb, gen = PF.write_case(str(tmp_path), grid=grid, snapshots=((0, 1.5, 2.0, 0.001),))
    out = str(tmp_path / "summary.json")
    # plan view beside a copy of the analysis named analysis_state.json, which
    # the writer merges in (surface fields and their 2 m humidity products)
    pv_dir = tmp_path / "pv"
    pv_dir.mkdir()
    shutil.copy(pa, pv_dir / "analysis_state.json")
    pv = str(pv_dir / "plan_view.json")
    steps, dt, ns = 3, 2.0, 3
    r = _run(["--analysis-state", pa, "--boundary-cache", pb, "--steps", str(steps),
              "--dt", str(dt), "--fast-substeps", str(ns),
              "--enable-boundaries", "true" if boundaries else "false",
              "--summary-json", out, "--plan-view-json", pv, "--plan-view-level", "2"],
             timeout=300)
    assert r.returncode == 0, r.stderr
    with open(out) as fh:
        js = json.load(fh)
    assert js["case"] == "prepared_case" and js["source"] == "HRRR"
    assert js["grid"]["nx"] == grid["nx"] and js["surface_runtime"] == {"ntile": 1, "nsoil": 4}
    b = js["startup_balance"]
    # bounds of the reference's start-up check (test_prepared_case_init.cpp)
    assert b["max_rel_eos"] <= 1e-6 and b["max_rel_hydrostatic"] <= 1e-6
    assert b["max_rel_fast_vertical"] <= 1e-6
    assert b["max_rel_mass_divergence"] <= 1e-3
    assert b["max_abs_mom_w_bottom"] == 0.0 and b["max_abs_mom_w_top"] == 0.0
    assert b["max_abs_tracer_closure"] <= 1e-6
    assert "max_abs_z_src_minus_metric" in b

    cfg = O.step_cfg(dt=dt, fast_substeps=ns)
    wr = O.warm_rain_cfg(dt=dt, rain_terminal_velocity=12.0)
    initial, final, world = _oracle_run(gen, grid, steps, cfg, wr, boundaries)
    for phase, ours, ref, rtol in (("initial", js["initial"], initial, 2e-6),
                                   ("final", js["final"], final, 2e-4)):
        for okey, jkey in SUM_MAP:
            a, e = _get(ours, jkey), ref[okey]
            scale = max(abs(e), 1e-3 * max(abs(ref["total_horizontal_momentum_x"]), 1.0)
                        if "momentum" in okey else 1e-12)
            if okey.endswith("_path") or okey == "precip_sum":
                scale = max(abs(e), 1e-9)
            assert abs(a - e) <= rtol * scale + (1e-9 if okey.endswith("_path") else 1e-7), \
                (phase, okey, a, e)
    # plan view: the runtime bundle with the bucket (23 fields) plus the five
    # surface fields of the companion analysis; values against the oracle's
    # products of its own final state (the FP32 trajectories differ at the
    # 2e-4 level, the JSON carries ten digits)
    with open(pv) as fh:
        pj = json.load(fh)
    assert pj["schema_version"] == "gwm-next-plan-view/v1" and pj["case"] == "prepared_case"
    assert pj["steps"] == steps and pj["slice_k"] == 2
    names = [f["name"] for f in pj["fields"]]
    assert names[:23] == O.PLAN_VIEW_NAMES
    assert names[23:] == ["surface_pressure", "air_temperature_2m", "specific_humidity_2m",
                          "relative_humidity_2m", "dewpoint_2m"]
    nf, ref = world.plan_view(2, steps, dt, with_precip=True)
    assert nf == 23
    for f, fj in enumerate(pj["fields"][:23]):
        assert fj["storage"] == "row_major_yx" and fj["nx"] == grid["nx"]
        got = np.asarray(fj["values"]).reshape(grid["ny"], grid["nx"])
        scale = max(np.max(np.abs(ref[f])), 1e-12)
        assert np.max(np.abs(got - ref[f])) <= 2e-3 * scale, (fj["name"], np.max(np.abs(got - ref[f])), scale)
