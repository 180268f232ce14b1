"""Benchmark: grid-cell RK3 steps per second of the dry + warm-rain step.

Workload (the 512x512x64 warm-rain convection case): dx = dy = 1 km, 64
levels stretched from 100 m to 500 m up to 16 km, doubly periodic, flat;
theta = 300 K + 43 K (z/12 km)^1.25 below 12 km and isothermal above, RH =
1 - 0.75 (z/12 km)^1.25 capped at qv = 14 g/kg, seeded +/-0.5 K theta noise
below 1 km; 6 acoustic substeps, warm rain with 12 m/s fallout.  A step is
one full moist step: dry SSPRK3 (3 slow stages, 18 fast substeps), tracer
SSPRK3 for qv/qc/qr, warm rain + fallout.  Inputs are resident in HBM
before timing starts.

The configuration's dt of 5 s overflows the reference's explicit acoustic
scheme within the first step on the 100 m levels, so the timed runs use dt
0.25 s, where the state stays finite; per-step work does not depend on dt,
and the 5 s run is reported beside it, with no throughput credited to a run
whose state holds a NaN or infinity (counted over every prognostic field).

`value` is cell-steps/s of the FP32 perturbation form (rho and rho*theta
stored as FP32 departures from FP64 level bases), the FP32 path that stays
within 1e-4 relative L2 of FP64; its error against the FP64 step (which the
GPU tests hold to the FP64 oracle: 1e-11 on operators and short runs, 1e-8
over the configurations' ill-conditioned trajectories) is measured here at
full size.
Prints one JSON line.
"""
import argparse
import json
import os
import platform
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "gpu-wm-next_amd"))

FIELD_NAMES = {0: "rho", 1: "rho_theta", 2: "u", 3: "v", 4: "w", 16: "qv", 17: "qc", 18: "qr"}
# physical scale of each field: relative L2 floors at 1e-3 of it
FIELD_SCALE = {0: 1.0, 1: 300.0, 2: 1.0, 3: 1.0, 4: 1.0, 16: 1e-3, 17: 1e-3, 18: 1e-3}


def measured_traffic(nx, ny, nz):
    """HBM bytes per k_fast_column launch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated against a copy of known
    size); only for the 512x512x64 workload it was measured on."""
    if (nx, ny, nz) != (512, 512, 64):
        return None
    path = os.path.join(REPO, "profiles", "fast_column_traffic_512x512x64.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        return None
    return None


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
    p.add_argument("--error-steps", type=int, default=10,
                   help="steps of the error leg against the FP64 step (0 to skip)")
    p.add_argument("--cpu-baseline", type=int, default=1,
                   help="time the FP64 CPU oracle on a bounded sample (0 to skip)")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="OpenMP threads of the CPU baseline (0: the host's share)")
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
    p.add_argument("--ensemble-dt", type=float, default=0.5,
                   help="dt of the timed ensemble run (the configuration's 2 s diverges, and "
                        "1 s on the 60-level members within twenty steps)")
    p.add_argument("--c2-steps", type=int, default=10,
                   help="timed steps of the mountain-wave leg (after 2 warm-up steps)")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="also run one device-to-device state copy (profiler byte calibration)")
    p.add_argument("--legs", default="all",
                   help="comma list of legs beyond the headline: full,graph,config_dt,copy,fp64,"
                        "surrogate,error,ensemble,c2 (all: every leg)")
    return p.parse_args()


def make_case(G, precision, nx, ny, nz, ctx=None, perturbation=False, moist=False):
    """The warm-rain convection case (sounding, seed-42 theta noise); with
    `moist` also the seeded moisture perturbations of the surrogate's
    training states (qv scaled by U(0.85, 1.25), cloud and rain in 30 % of
    the cells each), the same tracer values in every precision and form."""
    if ctx is None:
        ctx = G.Context(0, precision)
    zi = G.stretched_levels(nz, 100.0, 500.0, 16000.0)
    grid = G.Grid(ctx, nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi)
    st = G.State(grid, 3, perturbation=perturbation)
    st.init_convection(1.0e5, 42, 0.5, 1000.0)
    if moist:
        for t, q in enumerate(moist_tracers(G, nx, ny, nz)):
            st.set(16 + t, q)
    return ctx, grid, st, zi


_MOIST = {}


def moist_tracers(G, nx, ny, nz):
    """Tracer masses of the moist case, computed once from an FP32
    full-form state of the case."""
    key = (nx, ny, nz)
    if key not in _MOIST:
        import gwm_amd.surrogate as SG
        ctx = G.Context(0, G.F32)
        _, grid, st, _ = make_case(G, G.F32, nx, ny, nz, ctx)
        _MOIST[key] = SG.moist_tracers(st, 42)
        del st, grid
        ctx.close()
    return _MOIST[key]


def column_mass(st, zi):
    """Dry mass per unit horizontal area summed over columns, sum rho dz
    (FP64 download), the quantity the flux form conserves on stretched
    levels; the summary's total_dry_mass is the unweighted sum of rho."""
    import numpy as np
    rho = st.get_f64(0)
    dz = np.diff(np.asarray(zi, np.float64))
    return float(np.sum(dz * rho.reshape(rho.shape[0], -1).sum(axis=1)))


KERNELS = ("fast_column", "slow_stage", "level_partials", "level_finalize", "tracer_stage",
           "warm_rain", "warm_rain_surrogate", "fallout")


def run(G, precision, a, dt, timing=False, surrogate=None, ctx=None, perturbation=False,
        calibrate_copy=False, info=None, steps=None, moist=False, graph=False):
    """Warm-up, then `steps` timed moist steps; returns a leg record with
    the throughput, or value None when the state ended non-finite.  With a
    surrogate the record also gives the share of cells that went through
    the network at the last step (the others took the plain scheme's
    answer: zero source terms, or outside the training box)."""
    steps = a.steps if steps is None else steps
    ctx, grid, st, zi = make_case(G, precision, a.nx, a.ny, a.nz, ctx, perturbation, moist)
    cfg = G.step_config(dt=dt, fast_substeps=6)
    wr = G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0)

    def step():
        if surrogate is None:
            st.moist_step(cfg, wr)
        else:
            st.moist_step_surrogate(cfg, wr, surrogate)

    def steps_graph(n):
        # replayed from the captured two-step graph (bit-identical steps)
        st.moist_steps_graph(cfg, wr, n, surrogate)

    if calibrate_copy:
        other = G.State(grid, 3, perturbation=perturbation)
        other.copy_from(st)
        ctx.sync()
        del other
    if graph:
        steps_graph(max(a.warmup, 2))
    else:
        for _ in range(a.warmup):
            step()
    ctx.sync()
    if timing:
        ctx.reset_kernel_stats()
        ctx.set_kernel_timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    if graph:
        steps_graph(steps)
    else:
        for _ in range(steps):
            step()
    ctx.sync()
    secs = time.perf_counter() - t0
    stats = {}
    if timing:
        for name in KERNELS:
            n, ms = ctx.kernel_stats(name)
            if n:
                stats[name] = {"launches": n, "total_ms": ms, "avg_ms": ms / n}
        ctx.set_kernel_timing(False)
    summary = st.summary()
    finite = summary["nonfinite_values"] == 0
    if info is not None:
        info["tuning"] = ctx.tuning_report()
    cells = a.nx * a.ny * a.nz
    rec = {"dt": dt, "steps": steps, "state_finite": bool(finite),
           "nonfinite_values": int(summary["nonfinite_values"]),
           "value": cells * steps / secs if finite else None,
           "ms_per_step": secs / steps * 1e3}
    if surrogate is not None:
        fb, nw = surrogate.last_counts()
        rec["network_cell_share"] = nw / cells
        rec["fallback_cell_share"] = fb / cells
    if not finite:
        rec["note"] = "state non-finite after the timed steps: no throughput credited"
    return rec, stats, summary


def host_cpu():
    model = platform.processor() or ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = os.cpu_count() or 1
    return model, os.cpu_count() or share, share


def cpu_baseline(a, dt):
    """FP64 CPU oracle on a bounded sample of the workload, as the baseline
    plan states: OpenMP threads = the host's CPU share (OMP_NUM_THREADS when
    the host sets it), one warm-up step, then the median of three timed
    single steps."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as O
    import gwm_amd as G

    model, nproc, share = host_cpu()
    threads = a.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or share
    # A quarter of the horizontal domain (identical columns, levels and
    # physics), FP64 oracle, plain warm rain.
    nx, ny, nz = max(a.nx // 2, 8), max(a.ny // 2, 8), a.nz
    zi = G.stretched_levels(nz, 100.0, 500.0, 16000.0)
    ctx, grid, st, _ = make_case(G, G.F64, nx, ny, nz)
    w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi, n_tracers=3, precision=8)
    for f in (O.RHO, O.RHO_THETA, O.MOM_U, O.MOM_V, O.MOM_W, O.TRACER0):
        w.set(f, st.get(f))
    del st, grid, ctx
    cfg = O.step_cfg(dt=dt, fast_substeps=6)
    wr = O.warm_rain_cfg(dt=dt, rain_terminal_velocity=12.0)
    warm = w.time_steps(cfg, wr, moist=True, steps=1, threads=threads)
    runs = [w.time_steps(cfg, wr, moist=True, steps=1, threads=threads) for _ in range(3)]
    secs = statistics.median(runs)
    return {
        "value": nx * ny * nz / secs,
        "unit": "cell-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{nx}x{ny}x{nz} quarter of the case at dt {dt} s (finite state), FP64 CPU "
                  f"oracle, {threads} OpenMP threads; one warm-up step ({warm:.2f} s), then the "
                  f"median of 3 single steps ({', '.join(f'{r:.2f}' for r in runs)} s)",
        "host_cpu": model, "host_nproc": nproc, "host_cpu_share": share,
    }


def rel_errors(G, ref, states, fields):
    """Relative L2 (floor 1e-3 of each field's physical scale) of each
    state's fields against the reference state, downloaded in FP64."""
    import numpy as np
    out = {name: {} for name in states}
    for f in fields:
        r = ref.get_f64(f) if ref.perturbation else ref.get(f).astype(np.float64)
        den = max(np.linalg.norm(r), 1e-3 * FIELD_SCALE[f] * np.sqrt(r.size))
        for name, st in states.items():
            x = st.get_f64(f)
            out[name][FIELD_NAMES[f]] = float(np.linalg.norm(x - r) / den)
        del r
    for name in out:
        out[name]["worst"] = max(out[name].values())
    return out


def error_leg(G, a, dt, sur):
    """From the same initial state, the moist case (cloud, rain and
    supersaturated cells, so the source terms are large): `error_steps`
    moist steps of the FP64 plain step, the FP32 full form, the FP32
    perturbation form, both with the surrogate, and the perturbation form
    without microphysics (what the surrogate stands in for); relative L2 of
    each against the FP64 plain step, and the dry-mass change of the
    perturbation form.  Beside them, the conditioning of the trajectory:
    the FP64 plain step started from the same state rounded to FP32 (every
    field), whose distance from the reference is what FP32 storage of the
    initial state alone does to the FP64 scheme; and the FP32 forms'
    errors after 1, 2 and 5 steps."""
    import numpy as np
    ctx64 = G.Context(0, G.F64)
    ctx32 = G.Context(0, G.F32)
    _, _, ref, zi = make_case(G, G.F64, a.nx, a.ny, a.nz, ctx64, moist=True)
    _, _, twin, _ = make_case(G, G.F64, a.nx, a.ny, a.nz, ctx64, moist=True)
    for f in (0, 1, 2, 3, 4, 16, 17, 18):
        twin.set(f, twin.get(f).astype(np.float32).astype(np.float64))
    states = {}
    _, _, states["fp32_full"], _ = make_case(G, G.F32, a.nx, a.ny, a.nz, ctx32, moist=True)
    _, _, states["fp32_perturbation"], _ = make_case(G, G.F32, a.nx, a.ny, a.nz, ctx32,
                                                     perturbation=True, moist=True)
    surs = {}
    if sur is not None:
        s = G.Surrogate(ctx32, 0)
        s.set_norm(**{k: v for k, v in sur[1].items()})
        s.set_params(sur[0])
        for form, pert in (("fp32_full_surrogate", False), ("fp32_perturbation_surrogate", True)):
            _, _, states[form], _ = make_case(G, G.F32, a.nx, a.ny, a.nz, ctx32, perturbation=pert,
                                              moist=True)
            surs[form] = s
        # what the surrogate stands in for: the same run without microphysics
        _, _, states["fp32_perturbation_no_microphysics"], _ = make_case(
            G, G.F32, a.nx, a.ny, a.nz, ctx32, perturbation=True, moist=True)
    cfg = G.step_config(dt=dt, fast_substeps=6)
    wr = G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0)
    m0 = column_mass(states["fp32_perturbation"], zi)
    masses = [m0]
    fields = (0, 1, 2, 3, 4, 16, 17, 18)
    by_step = {}
    for step in range(1, a.error_steps + 1):
        ref.moist_step(cfg, wr)
        twin.moist_step(cfg, wr)
        for name, st in states.items():
            if name in surs:
                st.moist_step_surrogate(cfg, wr, surs[name])
            elif name.endswith("no_microphysics"):
                st.ssprk3_step(cfg)
                st.tracers_ssprk3(cfg)
            else:
                st.moist_step(cfg, wr)
        masses.append(column_mass(states["fp32_perturbation"], zi))
        if step in (1, 2, 5) and step < a.error_steps:
            e = rel_errors(G, ref, {"fp64_from_fp32_rounded_start": twin,
                                    "fp32_perturbation": states["fp32_perturbation"],
                                    "fp32_full": states["fp32_full"]}, fields)
            by_step[step] = {name: v["worst"] for name, v in e.items()}
    finite = {name: st.summary()["nonfinite_values"] == 0 for name, st in states.items()}
    finite["fp64_plain"] = ref.summary()["nonfinite_values"] == 0
    errs = rel_errors(G, ref, dict(states, fp64_from_fp32_rounded_start=twin), fields)
    by_step[a.error_steps] = {name: errs[name]["worst"]
                              for name in ("fp64_from_fp32_rounded_start", "fp32_perturbation",
                                           "fp32_full")}
    drift = [abs(m1 - m0) / m0 for m0, m1 in zip(masses[:-1], masses[1:])]
    return {"dt": dt, "steps": a.error_steps, "reference": "FP64 plain step on the device",
            "rel_l2_floor": "1e-3 of each field's physical scale",
            "rel_l2": errs, "worst_by_step": by_step, "state_finite": finite,
            "perturbation_dry_mass_drift_per_step_max": max(drift) if drift else None,
            "perturbation_dry_mass_rel_change_total": abs(masses[-1] - m0) / m0}


def surrogate_train(G, a):
    """Train the warm-rain surrogate on plain-scheme pairs (seeds 0-7 of the
    case) and report its tendency error against the plain scheme at seed
    42, next to the untrained network."""
    import gwm_amd.surrogate as SG

    ctx = G.Context(0, G.F32)
    seeds = list(range(8))
    per = int(a.surrogate_pairs) // len(seeds)
    dt = a.control_dt or a.dt
    log(f"surrogate: sampling {a.surrogate_pairs:.0f} pairs")
    t0 = time.perf_counter()
    # pairs from four-step plain runs of each seed (the start and the
    # near-saturated states after the first adjustments)
    run = (G.step_config(dt=dt, fast_substeps=6),
           G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0), 4)
    X, Y, n = SG.generate(ctx, seeds, per, shape=(a.nx, a.ny, a.nz),
                          wr=G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0), run=run)
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
    shape = (min(a.nx, 256), min(a.ny, 256), a.nz)
    wr0 = G.warm_rain_config(dt=dt, rain_terminal_velocity=0.0)
    errs = SG.tendency_errors(ctx, sur, seed=42, shape=shape, wr=wr0)
    raw = G.Surrogate(ctx, 0)
    raw.set_norm(**{k: v for k, v in norm.items()})
    errs_untrained = SG.tendency_errors(ctx, raw, seed=42, shape=shape, wr=wr0)
    rec = {"pairs": n, "seeds": seeds, "epochs": a.surrogate_epochs, "dt": dt,
           "generate_s": round(t_gen, 2), "train_s": round(t_train, 2),
           "epoch_loss": [float(v) for v in losses],
           "tendency_error_vs_plain_seed42": {
               k: (v if not isinstance(v, dict) else
                   {"trained": v["surrogate"], "untrained": errs_untrained[k]["surrogate"],
                    "no_source_terms": v["no_source_terms"]})
               for k, v in errs.items()}}
    return ctx, sur, (sur.params(), sur.norm()), rec


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
    """32 uneven members, FP32 + surrogate, 4 substeps, dt 0.5 s (the
    configuration's 2 s overflows within three steps, 1 s within twenty on
    the 60-level members), three ways from fresh members each: one launch per
    kernel of the step over all members (batched, graph-replayed), one
    context (stream) per member with its own graph replays, and the members
    stepped one after the other without graphs."""
    dt = a.ensemble_dt
    cfg = G.step_config(dt=dt, fast_substeps=4)
    wr = G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0)

    def fresh():
        members = ensemble_members(G, a.ensemble_members)
        surs = None
        if sur_src is not None:
            params, norm = sur_src
            surs = []
            for ctx, _, _, _ in members:
                s = G.Surrogate(ctx, 0)
                s.set_norm(**{k: v for k, v in norm.items()})
                s.set_params(params)
                surs.append(s)
        return members, surs

    def finish(members):
        for ctx, _, _, _ in members:
            ctx.sync()
        return sum(int(m[2].summary()["nonfinite_values"]) for m in members)

    out = {}
    for mode in ("batched", "streams", "eager"):
        members, surs = fresh()
        states = [m[2] for m in members]
        cells = sum(m[3] for m in members)
        if mode == "eager":
            def run_steps(n):
                for _ in range(n):
                    for i, st in enumerate(states):
                        if surs is None:
                            st.moist_step(cfg, wr)
                        else:
                            st.moist_step_surrogate(cfg, wr, surs[i])
        else:
            def run_steps(n):
                G.ensemble_steps(states, cfg, wr, n, surs, streams=(mode == "streams"))
        # warm-up: allocation, per-member tuning, and the batched form's
        # autotuner (one timed step pair per level-chunk candidate)
        run_steps(9)
        finish(members)
        t0 = time.perf_counter()
        run_steps(a.ensemble_steps)
        bad = finish(members)
        secs = time.perf_counter() - t0
        out[mode] = {"value": cells * a.ensemble_steps / secs if bad == 0 else None,
                     "state_finite": bad == 0, "nonfinite_values": bad}
        shapes = [list(map(int, (m[2].grid.nx, m[2].grid.ny, m[2].grid.nz))) for m in members]
        del states, members, surs
    finite = all(v["state_finite"] for v in out.values())
    return {
        "members": a.ensemble_members, "total_cells": cells, "dt": dt, "shapes": shapes,
        "steps": a.ensemble_steps, "surrogate": sur_src is not None,
        "batched_autotuner": G.ensemble_tuning_report().strip().splitlines(),
        "value_batched": out["batched"]["value"],
        "value_streams_graphs": out["streams"]["value"],
        "value_eager_sequential": out["eager"]["value"],
        "modes": out, "state_finite": finite,
    }


def c2_leg(G, a):
    """The mountain-wave configuration at its size: 256x256x60 columns,
    dx 2 km, levels stretched 50-600 m to 20 km, an Agnesi ridge (400 m,
    half-width 10 km, centred in x), constant N = 0.01 s-1 with u = 10 m/s,
    relaxation at the open x sides (width 5), 6 substeps; dt 0.25 s (its
    6 s overflows in the first step).  FP64, the FP32 full form (from the
    FP64 initial state rounded once) and the FP32 perturbation form (its
    3-D base the FP64 initial state itself) from the same state:
    throughput of each over the timed steps and each FP32 state's relative
    L2 against the FP64 state at the end."""
    import numpy as np
    nx, ny, nz, dx = 256, 256, 60, 2000.0
    zi = G.stretched_levels(nz, 50.0, 600.0, 20000.0)
    ter = G.agnesi_ridge(nx, ny, dx, 0.5 * nx * dx, 10000.0, 400.0)
    cfg = G.step_config(dt=0.25, fast_substeps=6)
    bc = G.BoundaryConfig(G.BOUNDARY_RELAX, 5.0, G.BOUNDARY_FN(), None)
    steps, warm = a.c2_steps, 2
    out, states = {}, {}
    for name, prec in (("fp64", G.F64), ("fp32_full", G.F32), ("fp32_pert", G.F32)):
        ctx = G.Context(0, prec)
        grid = G.Grid(ctx, nx, ny, nz, dx=dx, dy=dx, z_interfaces=zi, terrain=ter,
                      periodic_x=False)
        st = G.State(grid, 0, perturbation=name == "fp32_pert")
        st.init_constant_n(theta_surface=288.0, brunt_vaisala=0.01, p_surface=1.0e5, u=10.0)
        if name == "fp32_full":  # the FP64 initial state, rounded once
            for f in range(5):
                st.set(f, states["fp64_init"][f].astype(np.float32))
        else:
            states["fp64_init"] = [st.get(f) for f in range(5)]
        st.capture_boundary()
        for _ in range(warm):
            st.ssprk3_step(cfg, bc)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            st.ssprk3_step(cfg, bc)
        ctx.sync()
        secs = time.perf_counter() - t0
        bad = int(st.summary()["nonfinite_values"])
        out[name] = {"value": nx * ny * nz * steps / secs if bad == 0 else None,
                     "ms_per_step": secs / steps * 1e3, "state_finite": bad == 0}
        states[name] = (ctx, grid, st)
    errs = rel_errors(G, states["fp64"][2], {k: states[k][2] for k in ("fp32_full", "fp32_pert")},
                      (0, 1, 2, 3, 4))
    rec = {"workload": "C2 mountain wave 256x256x60, Agnesi ridge 400 m, constant N, "
                       "stretched 50-600 m to 20 km, relaxation in x, dry",
           "dt": 0.25, "steps": steps, "warmup": warm, "fp64": out["fp64"],
           "fp32_full": out["fp32_full"], "fp32_pert": out["fp32_pert"],
           "fp32_full_vs_fp64_rel_l2": errs["fp32_full"],
           "fp32_pert_vs_fp64_rel_l2": errs["fp32_pert"],
           "w_max_fp64": states["fp64"][2].summary()["w_max"]}
    for k in ("fp32_full", "fp32_pert"):
        if out["fp64"]["value"] and out[k]["value"]:
            rec[k + "_over_fp64"] = out[k]["value"] / out["fp64"]["value"]
    return rec


def copy_ceiling(G, a, reps=10):
    """Device-to-device copy rate of a moist state (read + write bytes over
    wall time, 8 fields and the surface bucket per copy): the box's
    STREAM-copy ceiling beside the 8 TB/s peak."""
    import numpy as np
    ctx, grid, src, _ = make_case(G, G.F32, a.nx, a.ny, a.nz)
    dst = G.State(grid, 3)
    dst.copy_from(src)
    elems = sum(int(np.prod(src.storage_shape(f))) for f in range(5))
    elems += 3 * int(np.prod(src.storage_shape(0))) + a.nx * a.ny
    nbytes = elems * 4 * 2
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_from(src)
    ctx.sync()
    secs = time.perf_counter() - t0
    return nbytes * reps / secs / 1e9


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the multi-GPU decomposition is described in DESIGN.md)"}))
        sys.exit(2)
    import gwm_amd as G

    legs = set(("full,graph,config_dt,copy,fp64,surrogate,error,ensemble,c2" if a.legs == "all"
                else a.legs).split(","))
    cells = a.nx * a.ny * a.nz
    dt = a.control_dt if a.control_dt > 0 else a.dt
    ns = 6

    log(f"FP32 perturbation form, plain warm rain, dt {dt}")
    head, _, summary = run(G, G.F32, a, dt, perturbation=True, calibrate_copy=a.calibrate_copy)
    # Per-kernel HIP-event timings come from a second pass, so the events do
    # not sit inside the timed region above.
    info = {}
    log("FP32 perturbation form, per-kernel timing")
    _, stats, _ = run(G, G.F32, a, dt, timing=True, perturbation=True, info=info)

    # Dominant kernel: the fast substep, which carries face departures from
    # the subcycle's entry faces.  Algorithmic elements per cell: the first
    # substep reads rho, theta and writes rho and three departures (6); a
    # middle substep reads rho, theta, three departures and writes rho and
    # three departures (9); the last reads rho, theta, three departures and
    # the three entry faces and writes rho, rho*theta and three faces (13).
    roofline = None
    fs = stats.get("fast_column")
    if fs:
        elems = (6 + 9 * (ns - 2) + 13) / ns
        bytes_per_launch = elems * 4 * cells
        achieved = bytes_per_launch / (fs["avg_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0,
                    "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                    "traffic": measured_traffic(a.nx, a.ny, a.nz),
                    "kernel": "fast_column", "avg_ms": round(fs["avg_ms"], 5),
                    "algorithmic_bytes": int(bytes_per_launch)}
    # Other stencil kernels against their own algorithmic bytes: slow stage
    # (mean of modes: stage 1 reads 5 writes 5, stages 2-3 read 10 write 5),
    # tracer stage (3 tracers: rho, 3 faces, 3 in, 3 q0 except stage 1, 3
    # out), warm rain (read 5, write 4).
    kernel_roofline = {}
    for name, elems in (("slow_stage", (10 + 15 + 15) / 3), ("tracer_stage", (10 + 13 + 13) / 3),
                        ("warm_rain", 9)):
        if name in stats:
            b = elems * 4 * cells
            gbs = b / (stats[name]["avg_ms"] * 1e-3) / 1e9
            kernel_roofline[name] = {"achieved_GBs": round(gbs, 1), "frac": round(gbs / 8000.0, 4),
                                     "algorithmic_bytes": int(b)}

    full = config_dt = fp64 = sur_rec = err = ens = graph = None
    if "graph" in legs:
        log("FP32 perturbation form replayed from the captured two-step graph")
        graph, _, _ = run(G, G.F32, a, dt, perturbation=True, graph=True)
    if "full" in legs:
        log("FP32 full form (the reference's formulation), dt %g" % dt)
        full, _, _ = run(G, G.F32, a, dt)
    if "config_dt" in legs and dt != a.dt:
        log(f"FP32 full form at the configuration's dt {a.dt}")
        config_dt, _, _ = run(G, G.F32, a, a.dt)
    if "copy" in legs:
        log("copy ceiling")
        copy_gbs = copy_ceiling(G, a)
        if roofline:
            roofline["copy_ceiling_GBs"] = round(copy_gbs, 1)
            roofline["frac_of_copy_ceiling"] = round(roofline["achieved"] / copy_gbs, 4)
    if "fp64" in legs and a.fp64_steps > 0:
        log("FP64 plain step")
        fp64, _, _ = run(G, G.F64, a, dt, steps=a.fp64_steps)
    model = None
    if "surrogate" in legs and a.surrogate_pairs > 0:
        sctx, sur, model, sur_rec = surrogate_train(G, a)
        log("FP32 + surrogate: full form, then perturbation form")
        sfull, sstats, _ = run(G, G.F32, a, dt, surrogate=sur, ctx=sctx)
        _, sstats, _ = run(G, G.F32, a, dt, surrogate=sur, ctx=sctx, timing=True)
        spert, _, _ = run(G, G.F32, a, dt, surrogate=sur, ctx=sctx, perturbation=True)
        sur_rec["full_form"] = sfull
        sur_rec["perturbation_form"] = spert
        sur_rec["kernels"] = sstats
        # The case above has no cloud or rain and stays subsaturated over the
        # timed steps, so the network has almost nothing to do there; the
        # moist case (cloud, rain and supersaturation in most cells) is the
        # one where it replaces real source terms.
        log("moist case: FP32 perturbation form plain and with the surrogate, FP64 plain")
        mplain, _, _ = run(G, G.F32, a, dt, perturbation=True, moist=True)
        mpert, _, _ = run(G, G.F32, a, dt, surrogate=sur, ctx=sctx, perturbation=True, moist=True)
        _, mstats, _ = run(G, G.F32, a, dt, surrogate=sur, ctx=sctx, perturbation=True,
                           moist=True, timing=True)
        m64 = None
        if a.fp64_steps > 0:
            m64, _, _ = run(G, G.F64, a, dt, steps=a.fp64_steps, moist=True)
        moist = {"fp32_perturbation_plain": mplain, "fp32_perturbation_surrogate": mpert,
                 "fp64_plain": m64, "kernels": mstats}
        if m64 and m64["value"]:
            moist["over_fp64_plain"] = {
                k: moist[k]["value"] / m64["value"]
                for k in ("fp32_perturbation_plain", "fp32_perturbation_surrogate")
                if moist[k]["value"]}
        sur_rec["moist_case"] = moist
        del sur, sctx
    if "error" in legs and a.error_steps > 0:
        log(f"error leg: {a.error_steps} steps against the FP64 step")
        err = error_leg(G, a, dt, model)

    ratios = {}
    if fp64 and fp64["value"]:
        for name, rec in (("fp32_perturbation", head), ("fp32_full", full),
                          ("fp32_perturbation_graph_replay", graph),
                          ("fp32_full_surrogate", sur_rec and sur_rec["full_form"]),
                          ("fp32_perturbation_surrogate", sur_rec and sur_rec["perturbation_form"])):
            if rec and rec.get("value"):
                ratios[name] = rec["value"] / fp64["value"]
    if "ensemble" in legs and a.ensemble_members > 0:
        log(f"ensemble: {a.ensemble_members} members")
        ens = ensemble_leg(G, a, model)
    c2 = None
    if "c2" in legs:
        log("mountain wave (C2) at 256x256x60: FP64, FP32 full and perturbation forms")
        c2 = c2_leg(G, a)
    cpu = None
    if a.cpu_baseline:
        log("CPU baseline")
        cpu = cpu_baseline(a, dt)

    out = {
        "metric": "cell-steps/s per GPU, % HBM roofline; FP32+surrogate vs FP64 plain at set error",
        "value": head["value"],
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"warm-rain convection {a.nx}x{a.ny}x{a.nz}, dx 1 km, stretched "
                               f"100-500 m to 16 km, dt {dt} s, 6 substeps, plain warm rain, "
                               f"FP32 perturbation form",
                   "global_batch": 1, "parallelism": "single GPU"},
        "state_finite": head["state_finite"],
        "fp32_full_form": full,
        "fp32_perturbation_graph_replay": graph,
        "at_configuration_dt": config_dt,
        "fp64_plain": fp64,
        "over_fp64_plain": ratios,
        "surrogate": sur_rec,
        "error_vs_fp64": err,
        "ensemble": ens,
        "mountain_wave_c2": c2,
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
