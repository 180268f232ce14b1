"""The C++ host mirror and the idealized driver.

CPU: the binaries are built and resolve libgwm_hip.so; the driver rejects
bad flags like the reference driver (src/core/gwm_idealized_driver.cpp:70-190).
GPU: the C++ mirror tests (tests/cpp/test_mirror_*.cpp) pass, and the driver's
summary JSON matches the oracle on the same case to the six digits the JSON
carries.
"""
import json
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "gpu-wm-next_amd", "host", "bin")
DRIVER = os.path.join(BIN, "gwm_idealized_driver")
CPP_TESTS = sorted(f[:-4] for f in os.listdir(os.path.join(REPO, "tests", "cpp"))
                   if f.startswith("test_") and f.endswith(".cpp"))


def _need(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")


def test_binaries_link_against_the_hip_library():
    for name in ["gwm_idealized_driver", "gwm_prepared_case_driver"] + CPP_TESTS:
        path = os.path.join(BIN, name)
        _need(path)
        out = subprocess.run(["ldd", path], capture_output=True, text=True, check=True).stdout
        line = [ln for ln in out.splitlines() if "libgwm_hip.so" in ln]
        assert line and "not found" not in line[0], out


@pytest.mark.parametrize("args,msg", [
    (["--bogus", "1"], "Unknown argument: --bogus"),
    (["--nx"], "Missing value for --nx"),
    (["--periodic-x", "maybe"], "Invalid boolean value: maybe"),
    (["--terrain-kind", "witch"], "Unsupported terrain kind: witch"),
])
def test_driver_rejects_bad_flags(args, msg):
    _need(DRIVER)
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert msg in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("name", CPP_TESTS)
def test_cpp_mirror(name):
    path = os.path.join(BIN, name)
    _need(path)
    print(f"running {name}", flush=True)
    r = subprocess.run([path], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stdout + r.stderr


CASES = {
    "warm_bubble": dict(
        args=["--case", "warm_bubble", "--nx", "24", "--ny", "20", "--nz", "10", "--dx", "500",
              "--dy", "500", "--z-top", "10000", "--rho-background", "1.16",
              "--center-z-fraction", "0.2", "--radius-x-fraction", "0.2",
              "--radius-y-fraction", "0.2", "--radius-z-fraction", "0.2", "--dt", "2",
              "--fast-substeps", "3", "--steps", "3"],
        world=dict(nx=24, ny=20, nz=10, dx=500.0, dy=500.0, z_top=10000.0),
        init=lambda w: w.init_bubble(1.16, 300.0, center=(0.5, 0.5, 0.2),
                                     radius=(0.2, 0.2, 0.2), theta_perturbation=2.0),
        step=dict(dt=2.0, fast_substeps=3)),
    "mountain_wave_background": dict(
        args=["--case", "mountain_wave_background", "--nx", "32", "--ny", "8", "--nz", "12",
              "--dx", "1000", "--dy", "1000", "--z-top", "12000", "--terrain-kind",
              "cosine_mountain", "--terrain-height", "400", "--terrain-half-width-x", "0.2",
              "--terrain-half-width-y", "0.5", "--u-background", "10", "--dt", "1",
              "--fast-substeps", "4", "--steps", "3"],
        world=dict(nx=32, ny=8, nz=12, dx=1000.0, dy=1000.0, z_top=12000.0),
        terrain=dict(cx=0.5, cy=0.5, hx=0.2, hy=0.5, height=400.0),
        init=lambda w: w.init_mountain_wave(1.18, 300.0, 10.0, 0.0, 0.0),
        step=dict(dt=1.0, fast_substeps=4)),
}

KEYS = [("total_dry_mass", "total_dry_mass"), ("total_rho_theta_m", "total_rho_theta_m"),
        ("total_horizontal_momentum_x", "total_horizontal_momentum_x"),
        ("total_vertical_momentum", "total_vertical_momentum")]
RANGES = [("rho_d", "rho"), ("theta_m", "theta"), ("u_face", "u"), ("w_face", "w")]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_driver_summary_matches_oracle(orc, case, precision):
    import gwm_amd as G

    _need(DRIVER)
    spec = CASES[case]
    print(f"driver {case} {precision}", flush=True)
    r = subprocess.run([DRIVER, "--precision", precision] + spec["args"], capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    assert got["case"] == case and got["steps"] == 3

    terrain = None
    if "terrain" in spec:
        t = spec["terrain"]
        terrain = G.cosine_mountain(spec["world"]["nx"], spec["world"]["ny"], t["height"],
                                    t["cx"], t["cy"], t["hx"], t["hy"])
    w = orc.World(**spec["world"], terrain=terrain, precision=8 if precision == "fp64" else 4)
    spec["init"](w)
    initial = w.summary()
    for _ in range(3):
        w.ssprk3_step(orc.step_cfg(**spec["step"]))
    final = w.summary()
    # The JSON carries six significant digits: a printed value is within
    # 5e-6 of its own magnitude.  FP32 adds the device/host FP32 trajectory
    # difference (device and host powf differ by an ulp now and then, and the
    # scheme amplifies it): the extremes of w move by up to ~1 % of its range
    # after three steps, while the FP32 parity tests bound the rel-L2 at 2e-4.
    digits = 6e-6
    traj = 0.0 if precision == "fp64" else 1e-2
    bad = []
    for part, ref in (("initial", initial), ("final", final)):
        s = got[part]
        scale = abs(ref["total_dry_mass"])
        for jk, ok in KEYS:
            # Momentum totals can be round-off around zero: allow the momentum
            # a 1e-6 m/s (FP64) or 1e-3 m/s (FP32) mean wind would carry.
            floor = (1e-6 if precision == "fp64" else 1e-3) * scale if "momentum" in jk else 0.0
            tol = digits * abs(ref[ok]) + traj * abs(ref[ok]) + floor
            if abs(s[jk] - ref[ok]) > tol:
                bad.append((part, jk, s[jk], ref[ok]))
        for jk, ok in RANGES:
            spread = max(abs(ref[f"{ok}_max"]), abs(ref[f"{ok}_min"]))
            for b in ("min", "max"):
                v = ref[f"{ok}_{b}"]
                tol = digits * abs(v) + traj * spread + 1e-12
                if abs(s[jk][b] - v) > tol:
                    bad.append((part, jk, b, s[jk][b], v))
    assert not bad, bad
    assert np.isfinite(got["final"]["total_dry_mass"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_driver_plan_view_matches_oracle(orc, tmp_path, case):
    """--plan-view-json: the dry bundle (10 fields) of the final state, against
    the oracle's products of its own final state (FP64; the JSON carries ten
    significant digits)."""
    import gwm_amd as G

    _need(DRIVER)
    spec = CASES[case]
    pv = str(tmp_path / "plan_view.json")
    r = subprocess.run([DRIVER, "--precision", "fp64"] + spec["args"] +
                       ["--plan-view-json", pv, "--plan-view-level", "3"],
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr
    with open(pv) as fh:
        pj = json.load(fh)
    assert pj["schema_version"] == "gwm-next-plan-view/v1" and pj["case"] == case
    assert pj["steps"] == 3 and pj["slice_k"] == 3
    assert [f["name"] for f in pj["fields"]] == orc.PLAN_VIEW_NAMES[:10]
    terrain = None
    if "terrain" in spec:
        t = spec["terrain"]
        terrain = G.cosine_mountain(spec["world"]["nx"], spec["world"]["ny"], t["height"],
                                    t["cx"], t["cy"], t["hx"], t["hy"])
    w = orc.World(**spec["world"], terrain=terrain, precision=8)
    spec["init"](w)
    for _ in range(3):
        w.ssprk3_step(orc.step_cfg(**spec["step"]))
    nf, ref = w.plan_view(3, 3, spec["step"]["dt"], with_precip=False)
    assert nf == 10
    nx, ny = spec["world"]["nx"], spec["world"]["ny"]
    assert abs(pj["grid"]["slice_mean_height_m"] - ref[1].mean()) <= 1e-9 * abs(ref[1].mean())
    for f, fj in enumerate(pj["fields"]):
        got = np.asarray(fj["values"]).reshape(ny, nx)
        scale = max(np.max(np.abs(ref[f])), 1e-300)
        # ten printed digits, plus the FP64 device/oracle difference (1e-11)
        assert np.max(np.abs(got - ref[f])) <= 1e-9 * scale + 1e-12, fj["name"]
