"""Benchmark: grid-cell RK3 steps per second of the dry + warm-rain step.

Workload (the 512x512x64 warm-rain convection case): dx = dy = 1 km, 64
levels stretched from 100 m to 500 m up to 16 km, doubly periodic, flat;
theta = 300 K + 43 K (z/12 km)^1.25 below 12 km and isothermal above, RH =
1 - 0.75 (z/12 km)^1.25 capped at qv = 14 g/kg, seeded +/-0.5 K theta noise
below 1 km; dt = 5 s, 6 acoustic substeps, warm rain with 12 m/s fallout.
A step is one full moist step: dry SSPRK3 (3 slow stages, 18 fast
substeps), tracer SSPRK3 for qv/qc/qr, warm rain + fallout.  Inputs are
resident in HBM before timing starts.

Prints one JSON line.  `value` is cell-steps/s of the FP32 step on one GPU.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gpu-wm-next_amd"))


def measured_traffic(nx, ny, nz):
    """HBM bytes per k_fast_column launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known
    size); only for the 512x512x64 workload it was measured on."""
    if (nx, ny, nz) != (512, 512, 64):
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "fast_column_traffic_512x512x64.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        return None
    return None


def copy_ceiling(G, nx, ny, nz, reps=10):
    """Device-to-device copy rate of a moist state (read + write bytes over
    wall time, 8 fields and the surface bucket per copy): the box's
    STREAM-copy ceiling beside the 8 TB/s peak."""
    ctx, grid, a, _ = make_case(G, G.F32, nx, ny, nz)
    b = G.State(grid, 3)
    b.copy_from(a)
    import numpy as np
    elems = sum(int(np.prod(a.storage_shape(f))) for f in range(5))
    elems += 3 * int(np.prod(a.storage_shape(0))) + nx * ny
    nbytes = elems * 4 * 2
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_from(a)
    secs = time.perf_counter() - t0
    return nbytes * reps / secs / 1e9


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--nx", type=int, default=512)
    p.add_argument("--ny", type=int, default=512)
    p.add_argument("--nz", type=int, default=64)
    p.add_argument("--fp64-steps", type=int, default=5,
                   help="timed steps of the FP64 plain path (0 to skip)")
    p.add_argument("--cpu-baseline", type=int, default=1,
                   help="time the FP64 CPU oracle on a bounded sample (0 to skip)")
    p.add_argument("--cpu-threads", type=int, default=16)
    p.add_argument("--dt", type=float, default=5.0)
    p.add_argument("--control-dt", type=float, default=0.25,
                   help="dt of the timed runs, one the reference scheme integrates without "
                        "diverging (0: the configuration's --dt)")
    p.add_argument("--surrogate-pairs", type=float, default=2e7,
                   help="training pairs of the surrogate leg, from seeds 0-7 (0 to skip)")
    p.add_argument("--surrogate-epochs", type=int, default=4)
    p.add_argument("--ensemble-members", type=int, default=32,
                   help="members of the batched-ensemble leg (0 to skip)")
    p.add_argument("--ensemble-steps", type=int, default=20)
    p.add_argument("--perturbation", type=int, default=1,
                   help="time the FP32 perturbation form, plain and with the surrogate (0 to skip)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run one device-to-device state copy (profiler byte calibration)")
    return p.parse_args()


def make_case(G, precision, nx, ny, nz, ctx=None, perturbation=False):
    if ctx is None:
        ctx = G.Context(0, precision)
    zi = G.stretched_levels(nz, 100.0, 500.0, 16000.0)
    grid = G.Grid(ctx, nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi)
    st = G.State(grid, 3, perturbation=perturbation)
    st.init_convection(1.0e5, 42, 0.5, 1000.0)
    return ctx, grid, st, zi


def column_mass(st, zi):
    """Dry mass per unit horizontal area summed over columns, sum rho dz
    (FP64 download), the quantity the flux form conserves on stretched
    levels; the summary's total_dry_mass is the unweighted sum of rho."""
    import numpy as np
    rho = st.get_f64(0)
    dz = np.diff(np.asarray(zi, np.float64))
    return float(np.sum(dz * rho.reshape(rho.shape[0], -1).sum(axis=1)))


def run(G, precision, nx, ny, nz, steps, warmup, timing, dt=5.0, surrogate=None, ctx=None,
        calibrate_copy=False, info=None, perturbation=False, mass=False):
    ctx, grid, st, zi = make_case(G, precision, nx, ny, nz, ctx, perturbation)
    if mass:
        info["column_mass_initial"] = column_mass(st, zi)
    cfg = G.step_config(dt=dt, fast_substeps=6)
    wr = G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0)

    def step():
        if surrogate is None:
            st.moist_step(cfg, wr)
        else:
            st.moist_step_surrogate(cfg, wr, surrogate)

    if calibrate_copy:
        other = G.State(grid, 3)
        other.copy_from(st)
        ctx.sync()
        del other
    for _ in range(warmup):
        step()
    ctx.sync()
    if timing:
        ctx.reset_kernel_stats()
        ctx.set_kernel_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    secs = time.perf_counter() - t0
    stats = {}
    if timing:
        for name in ("fast_column", "slow_stage", "level_partials", "level_finalize",
                     "tracer_stage", "warm_rain", "warm_rain_surrogate", "fallout"):
            n, ms = ctx.kernel_stats(name)
            if n:
                stats[name] = {"launches": n, "total_ms": ms, "avg_ms": ms / n}
        ctx.set_kernel_timing(False)
    summary = st.summary()
    if mass:
        info["column_mass_final"] = column_mass(st, zi)
    if info is not None:
        info["tuning"] = ctx.tuning_report()
    return secs, stats, summary


def cpu_baseline(nx, ny, nz, threads, dt):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import numpy as np
    import pyoracle as O
    import gwm_amd as G

    # Bounded sample of the same workload: a quarter of the horizontal domain
    # (identical columns, levels and physics), one moist step, FP64 oracle.
    nx, ny = max(nx // 2, 8), max(ny // 2, 8)
    zi = G.stretched_levels(nz, 100.0, 500.0, 16000.0)
    ctx, grid, st, _ = make_case(G, G.F64, nx, ny, nz)
    w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi, n_tracers=3, precision=8)
    for f in (O.RHO, O.RHO_THETA, O.MOM_U, O.MOM_V, O.MOM_W, O.TRACER0):
        w.set(f, st.get(f))
    del st, grid, ctx
    steps = 1
    secs = w.time_steps(O.step_cfg(dt=dt, fast_substeps=6),
                        O.warm_rain_cfg(dt=dt, rain_terminal_velocity=12.0), moist=True,
                        steps=steps, threads=threads)
    _ = np.float64(secs)
    return {
        "value": nx * ny * nz * steps / secs,
        "unit": "cell-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} full moist step of a {nx}x{ny}x{nz} quarter of the case at dt {dt} s "
                  f"(finite state; FP64 CPU oracle, {threads} OpenMP threads, {secs:.1f} s)",
    }


def surrogate_leg(G, a, cells, info):
    """Train the warm-rain surrogate on plain-scheme pairs (seeds 0-7 of the
    case), report its error against the plain scheme at seed 42, and time the
    FP32+surrogate step."""
    import numpy as np
    import gwm_amd.surrogate as SG

    ctx = G.Context(0, G.F32)
    seeds = list(range(8))
    per = int(a.surrogate_pairs) // len(seeds)
    log(f"surrogate: sampling {a.surrogate_pairs:.0f} pairs")
    t0 = time.perf_counter()
    X, Y, n = SG.generate(ctx, seeds, per, shape=(a.nx, a.ny, a.nz),
                          wr=G.warm_rain_config(dt=a.dt, rain_terminal_velocity=12.0))
    ctx.sync()
    t_gen = time.perf_counter() - t0
    log("surrogate: training")
    t0 = time.perf_counter()
    sur, norm, losses = SG.train(ctx, X, Y, n, seed=0, epochs=a.surrogate_epochs, batch=4096,
                                 lr=2e-3, lr_final=5e-5)
    ctx.sync()
    t_train = time.perf_counter() - t0
    del X, Y
    log(f"surrogate: epoch losses {list(losses)}; error at seed 42")
    errs = SG.tendency_errors(ctx, sur, seed=42, shape=(min(a.nx, 256), min(a.ny, 256), a.nz),
                              wr=G.warm_rain_config(dt=a.dt, rain_terminal_velocity=0.0))
    # Throughput of the FP32+surrogate step at the finite-state dt.
    dt = a.control_dt or a.dt
    sinfo = {}
    log("surrogate: FP32+surrogate step")
    ss, sstats, ssum = run(G, G.F32, a.nx, a.ny, a.nz, a.steps, a.warmup, False, dt,
                           surrogate=sur, ctx=ctx)
    _, sstats, _ = run(G, G.F32, a.nx, a.ny, a.nz, a.steps, a.warmup, True, dt, surrogate=sur,
                       ctx=ctx, info=sinfo)
    return {
        "_model": (sur.params(), sur.norm()),
        "pairs": n, "seeds": seeds, "epochs": a.surrogate_epochs,
        "generate_s": round(t_gen, 2), "train_s": round(t_train, 2),
        "epoch_loss": [float(v) for v in losses],
        "error_vs_plain_seed42": errs,
        "value": cells * a.steps / ss, "dt": dt,
        "state_finite": bool(np.isfinite(ssum["w_max"])),
        "kernels": sstats,
    }


def perturbation_leg(G, a, cells, sur_src):
    """The FP32 perturbation form (rho' and (rho theta)' against FP64
    level-mean bases; within 1e-4 of the FP64 oracle in the tests) on the
    same case at the finite-state dt: plain warm rain, and with the trained
    surrogate."""
    import numpy as np

    dt = a.control_dt or a.dt
    ctx = G.Context(0, G.F32)
    out = {"dt": dt}
    log("perturbation form: plain")
    info = {}
    s, _, summ = run(G, G.F32, a.nx, a.ny, a.nz, a.steps, a.warmup, False, dt, ctx=ctx,
                     perturbation=True, info=info, mass=True)
    out["value_plain"] = cells * a.steps / s
    out["state_finite"] = bool(np.isfinite(summ["w_max"]))
    m0, m1 = info["column_mass_initial"], info["column_mass_final"]
    out["dry_mass_rel_change"] = abs(m1 - m0) / m0  # sum of rho dz, over all steps run
    out["dry_mass_steps"] = a.steps + a.warmup
    if sur_src is not None:
        params, norm = sur_src
        sur = G.Surrogate(ctx, 0)
        sur.set_norm(**{k: v for k, v in norm.items()})
        sur.set_params(params)
        log("perturbation form: surrogate")
        s, _, summ = run(G, G.F32, a.nx, a.ny, a.nz, a.steps, a.warmup, False, dt, surrogate=sur,
                         ctx=ctx, perturbation=True)
        out["value_surrogate"] = cells * a.steps / s
        out["state_finite_surrogate"] = bool(np.isfinite(summ["w_max"]))
    return out


def ensemble_members(G, n, seed=0, z_top=16000.0):
    """Uneven members: nx, ny uniform in {96, 128, ..., 320}, nz in {40, 48,
    60} (seeded); the convection sounding (moist) with, for half the members,
    a warm bubble (+1 to +3 K, seeded centre, radius 2 km) and for the other
    half a density current (-15 K cold bubble, radius 4 km at z = 3 km), as
    a cos^2 theta perturbation at fixed rho_d.  One context (stream) each."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        nx, ny = (int(v) for v in rng.choice(np.arange(96, 321, 32), 2))
        nz = int(rng.choice([40, 48, 60]))
        ctx = G.Context(0, G.F32)
        zi = np.linspace(0.0, z_top, nz + 1)
        grid = G.Grid(ctx, nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi)
        st = G.State(grid, 3)
        st.init_convection(1.0e5, 100 + m, 0.5, 1000.0)
        warm = m % 2 == 0
        if warm:
            amp, r0 = float(rng.uniform(1.0, 3.0)), 2000.0
            cx, cy, cz = rng.uniform(0.3, 0.7) * nx * 1e3, rng.uniform(0.3, 0.7) * ny * 1e3, 2000.0
        else:
            amp, r0 = -15.0, 4000.0
            cx, cy, cz = 0.5 * nx * 1e3, 0.5 * ny * 1e3, 3000.0
        zc = 0.5 * (zi[1:] + zi[:-1])
        x = (np.arange(nx) + 0.5) * 1e3
        y = (np.arange(ny) + 0.5) * 1e3
        r = np.sqrt(((x[None, None, :] - cx) / r0) ** 2 + ((y[None, :, None] - cy) / r0) ** 2 +
                    ((zc[:, None, None] - cz) / r0) ** 2)
        dth = np.where(r < 1.0, amp * np.cos(0.5 * np.pi * r) ** 2, 0.0)
        rho = st.get(0).astype(np.float64)
        st.set(1, (st.get(1) + rho * dth).astype(np.float32))
        out.append((ctx, grid, st, nx * ny * nz))
    return out


def ensemble_leg(G, a, sur_src):
    """C5: 32 uneven members, FP32 + surrogate, dt 2 s, 4 substeps, advanced
    concurrently from captured graphs (one stream per member), against the
    same members stepped one after the other without graphs."""
    import numpy as np

    members = ensemble_members(G, a.ensemble_members)
    cells = sum(m[3] for m in members)
    cfg = G.step_config(dt=2.0, fast_substeps=4)
    wr = G.warm_rain_config(dt=2.0, rain_terminal_velocity=12.0)
    surs = None
    if sur_src is not None:
        params, norm = sur_src
        surs = []
        for ctx, _, _, _ in members:
            s = G.Surrogate(ctx, 0)
            s.set_norm(**{k: v for k, v in norm.items()})
            s.set_params(params)
            surs.append(s)
    states = [m[2] for m in members]
    # warm-up: captures every member's graphs
    G.ensemble_steps(states, cfg, wr, 3, surs)
    t0 = time.perf_counter()
    G.ensemble_steps(states, cfg, wr, a.ensemble_steps, surs)
    t_graph = time.perf_counter() - t0
    # the same work launched member by member, eager
    for st, (ctx, _, _, _) in zip(states, members):
        ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.ensemble_steps):
        for i, st in enumerate(states):
            if surs is None:
                st.moist_step(cfg, wr)
            else:
                st.moist_step_surrogate(cfg, wr, surs[i])
    for ctx, _, _, _ in members:
        ctx.sync()
    t_eager = time.perf_counter() - t0
    finite = all(np.isfinite(st.summary()["w_max"]) for st in states)
    return {
        "members": len(members), "total_cells": cells,
        "shapes": [list(map(int, (m[2].grid.nx, m[2].grid.ny, m[2].grid.nz))) for m in members],
        "steps": a.ensemble_steps, "surrogate": surs is not None,
        "value_graph_concurrent": cells * a.ensemble_steps / t_graph,
        "value_eager_sequential": cells * a.ensemble_steps / t_eager,
        "state_finite": bool(finite),
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the multi-GPU decomposition is described in DESIGN.md)"}))
        sys.exit(2)
    import gwm_amd as G

    cells = a.nx * a.ny * a.nz
    import math

    # The timed step runs at a dt the reference scheme integrates without
    # diverging (the configuration's 5 s blows up in the reference's
    # split-explicit scheme, and in this one, within a few steps); the per-step
    # work does not depend on dt, and the configuration's dt is timed too.
    main_dt = a.control_dt if a.control_dt > 0 else a.dt
    log(f"FP32 plain step, dt {main_dt}")
    secs, _, summary = run(G, G.F32, a.nx, a.ny, a.nz, a.steps, a.warmup, False, main_dt,
                           calibrate_copy=a.calibrate_copy)
    value = cells * a.steps / secs
    # Per-kernel HIP-event timings come from a second pass, so the events do
    # not sit inside the timed region above.
    info = {}
    log("FP32 plain step, per-kernel timing")
    _, stats, _ = run(G, G.F32, a.nx, a.ny, a.nz, a.steps, a.warmup, True, main_dt, info=info)
    finite = all(math.isfinite(summary[k]) for k in ("w_min", "w_max", "total_dry_mass"))
    config_dt = None
    if main_dt != a.dt:
        log(f"FP32 plain step at the configuration's dt {a.dt}")
        cs, _, csum = run(G, G.F32, a.nx, a.ny, a.nz, a.steps, a.warmup, False, a.dt)
        config_dt = {"dt": a.dt, "value": cells * a.steps / cs,
                     "state_finite": all(math.isfinite(csum[k]) for k in ("w_min", "w_max"))}
    ns = 6
    # Dominant kernel: the fast substep, which carries face departures from
    # the subcycle's entry faces.  Algorithmic elements per cell: the first
    # substep reads rho, theta and writes rho and three departures (6); a
    # middle substep reads rho, theta, three departures and writes rho and
    # three departures (9); the last reads rho, theta, three departures and
    # the three entry faces and writes rho, rho*theta and three faces (13).
    kname = "fast_column"
    fs = stats.get(kname)
    roofline = None
    if fs:
        elems = (6 + 9 * (ns - 2) + 13) / ns
        bytes_per_launch = elems * 4 * cells
        achieved = bytes_per_launch / (fs["avg_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                    "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                    "traffic": measured_traffic(a.nx, a.ny, a.nz),
                    "kernel": kname, "avg_ms": round(fs["avg_ms"], 5),
                    "algorithmic_bytes": int(bytes_per_launch)}
    # Other stencil kernels against their own algorithmic bytes: slow stage
    # (mean of modes: stage 1 reads 5 writes 5, stages 2-3 read 10 write 5),
    # tracer stage (3 tracers: rho, 3 faces, 3 in, 3 q0 except stage 1, 3 out).
    kernel_roofline = {}
    for name, elems in (("slow_stage", (10 + 15 + 15) / 3), ("tracer_stage", (10 + 13 + 13) / 3),
                        ("warm_rain", 9)):
        if name in stats:
            b = elems * 4 * cells
            gbs = b / (stats[name]["avg_ms"] * 1e-3) / 1e9
            kernel_roofline[name] = {"achieved_GBs": round(gbs, 1), "frac": round(gbs / 8000.0, 4),
                                     "algorithmic_bytes": int(b)}
    log("copy ceiling")
    copy_gbs = copy_ceiling(G, a.nx, a.ny, a.nz)
    if roofline:
        roofline["copy_ceiling_GBs"] = round(copy_gbs, 1)
        roofline["frac_of_copy_ceiling"] = round(roofline["achieved"] / copy_gbs, 4)
    fp64 = None
    if a.fp64_steps > 0:
        log("FP64 plain step")
        s64, _, _ = run(G, G.F64, a.nx, a.ny, a.nz, a.fp64_steps, 1, False, main_dt)
        fp64 = cells * a.fp64_steps / s64
    sur = surrogate_leg(G, a, cells, info) if a.surrogate_pairs > 0 else None
    if sur is not None and fp64:
        sur["over_fp64_plain"] = sur["value"] / fp64
    ens = None
    model = sur.pop("_model", None) if sur else None
    pert = perturbation_leg(G, a, cells, model) if a.perturbation else None
    if pert is not None and fp64:
        for k in ("value_plain", "value_surrogate"):
            if k in pert:
                pert[k.replace("value", "over_fp64_plain")] = pert[k] / fp64
    if a.ensemble_members > 0:
        log(f"ensemble: {a.ensemble_members} members")
        ens = ensemble_leg(G, a, model)
    if a.cpu_baseline:
        log("CPU baseline")
    cpu = (cpu_baseline(a.nx, a.ny, a.nz, a.cpu_threads, a.control_dt or a.dt)
           if a.cpu_baseline else None)
    out = {
        "metric": "cell-steps/s per GPU, % HBM roofline; FP32+surrogate vs FP64 plain at set error",
        "value": value,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": secs / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"warm-rain convection {a.nx}x{a.ny}x{a.nz}, dx 1 km, stretched "
                               f"100-500 m to 16 km, dt {main_dt} s, 6 substeps, plain warm rain",
                   "global_batch": 1, "parallelism": "single GPU"},
        "state_finite": finite,
        "at_configuration_dt": config_dt,
        "fp64_plain_value": fp64,
        "fp32_over_fp64": (value / fp64) if fp64 else None,
        "surrogate": sur,
        "perturbation_form": pert,
        "ensemble": ens,
        "dry_mass": summary["total_dry_mass"],
        "kernels": stats,
        "kernel_roofline": kernel_roofline,
        "tuning": info.get("tuning"),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
