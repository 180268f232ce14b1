"""Synthetic prepared cases: an analysis state and a boundary cache in the
reference's JSON schemas ("gwm-next-analysis-state/v1",
"gwm-next-boundary-cache/v1", src/ingest/runtime_case.cpp:151-200), for the
prepared-case driver tests.  Fields are x-fastest (nz, ny, nx) FP32 lists.

The atmosphere is a stable sounding (theta rising 4 K/km from 298 K, a
near-hydrostatic pressure profile) over a Gaussian hill, with sheared winds
and a moist lower layer; a boundary snapshot is the analysis warmed by
`dtheta`, with stronger wind and moister air."""
import json

import numpy as np

GRID = dict(nx=14, ny=12, nz=8, dx=2000.0, dy=2000.0, z_top=8000.0)


def fields(grid, warm=0.0, wind=0.0, moist=0.0, seed=0):
    nx, ny, nz = grid["nx"], grid["ny"], grid["nz"]
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) * grid["dx"]
    y = (np.arange(ny) + 0.5) * grid["dy"]
    hill = 400.0 * np.exp(-(((x[None, :] - x.mean()) / 6000.0) ** 2 +
                            ((y[:, None] - y.mean()) / 5000.0) ** 2))
    z = ((np.arange(nz) + 0.5) / nz * grid["z_top"])[:, None, None] + 0.0 * hill[None]
    theta = 298.0 + 0.004 * z + warm + 0.2 * rng.standard_normal((nz, ny, nx))
    p = 100000.0 * np.exp(-z / 8000.0) * (1.0 - 0.0005 * hill[None] / 100.0)
    t = theta * (p / 100000.0) ** 0.2854
    u = 6.0 + 0.002 * z + wind + 0.1 * rng.standard_normal((nz, ny, nx))
    v = -2.0 + 0.001 * z - 0.5 * wind + 0.1 * rng.standard_normal((nz, ny, nx))
    w = 0.01 * rng.standard_normal((nz, ny, nx))
    q = (0.012 + moist) * np.exp(-z / 2500.0)
    hgt = z + hill[None] * (1.0 - z / grid["z_top"])
    atm = {"air_pressure": p, "air_temperature": t, "u_wind": u, "v_wind": v, "w_wind": w,
           "specific_humidity": q, "geopotential_height": hgt}
    sfc = {"surface_pressure": p[0], "air_temperature_2m": t[0], "specific_humidity_2m": q[0],
           "u_wind_10m": u[0], "v_wind_10m": v[0], "skin_temperature": t[0] + 1.0}
    static = {"terrain_height": hill, "land_mask": np.ones((ny, nx)),
              "land_use_index": np.full((ny, nx), 7.0)}
    return atm, sfc, static


def _flat(d):
    return {k: np.asarray(v, np.float32).ravel().tolist() for k, v in d.items()}


def write_case(dirpath, grid=GRID, snapshots=((0, 1.0, 1.5, 0.001),), source="HRRR", seed=0):
    """Writes analysis.json and boundary.json; returns their paths and the
    generating fields {"analysis": (atm, sfc, static), "snapshots": [(off, atm, sfc)]}."""
    atm, sfc, static = fields(grid, seed=seed)
    meta = {"status": "synthetic", "generator": "tests/prepared_case_fixture.py"}
    analysis = {"schema_version": "gwm-next-analysis-state/v1", "source": source,
                "grid": dict(grid), "cycle_time_utc": "2026-04-03T00:00:00Z",
                "valid_time_utc": "2026-04-03T00:00:00Z", "forecast_offset_seconds": 0,
                "metadata": meta, "atmosphere": _flat(atm), "surface": _flat(sfc),
                "static_surface": _flat(static)}
    snaps, gen = [], []
    for off, warm, wind, moist in snapshots:
        a, s, _ = fields(grid, warm=warm, wind=wind, moist=moist, seed=seed)
        snaps.append({"forecast_offset_seconds": int(off),
                      "valid_time_utc": f"2026-04-03T{int(off) // 3600:02d}:00:00Z",
                      "atmosphere": _flat(a), "surface": _flat(s)})
        gen.append((int(off), a, s))
    boundary = {"schema_version": "gwm-next-boundary-cache/v1", "source": source,
                "grid": dict(grid), "cycle_time_utc": "2026-04-03T00:00:00Z",
                "boundary_interval_seconds": 3600, "metadata": meta, "snapshots": snaps}
    pa, pb = f"{dirpath}/analysis.json", f"{dirpath}/boundary.json"
    with open(pa, "w") as fh:
        json.dump(analysis, fh)
    with open(pb, "w") as fh:
        json.dump(boundary, fh)
    return pa, pb, {"analysis": (atm, sfc, static), "snapshots": gen}
