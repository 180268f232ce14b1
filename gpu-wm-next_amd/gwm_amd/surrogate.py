"""Warm-rain surrogate pipeline: training data from the plain scheme,
normalisation and fallback box, training on the device, and evaluation
against the plain scheme.  The arithmetic runs in the HIP kernels of
libgwm_hip.so (gwm_surrogate.hip); this module only orchestrates.

No counterpart in the reference (FahrenheitResearch/gpu-wm-next): the
surrogate replaces the source terms of apply_warm_rain_microphysics
(src/physics/warm_rain.cu:36-120), whose plain form generates the data,
handles the cells outside the training box and judges the error.
"""
import numpy as np

from . import (SURROGATE_IN, SURROGATE_OUT, DeviceArray, F32, Grid, State, Surrogate,
               stretched_levels, surrogate_sample, warm_rain_config)

FEATURES = ("T", "p", "rho_d", "qv", "qc", "qr", "dt")
TARGETS = ("dqv", "dqc", "dqr", "dtheta")


def convection_state(ctx, nx, ny, nz, seed, dz_bottom=100.0, dz_top=500.0, z_top=16000.0,
                     perturb=True):
    """The warm-rain convection sounding (gwm_init_convection) with seeded
    theta noise, plus -- with `perturb` -- seeded moisture perturbations so
    that saturated, cloudy and raining cells occur: qv scaled by U(0.85,
    1.25), and cloud and rain water (|N(0,1)| * 1.5 g/kg and * 0.8 g/kg) in
    a random 30 % of the cells each."""
    zi = stretched_levels(nz, dz_bottom, dz_top, z_top)
    grid = Grid(ctx, nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi)
    st = State(grid, 3)
    st.init_convection(1.0e5, int(seed), 0.5, 1000.0)
    if perturb:
        for t, q in enumerate(moist_tracers(st, seed)):
            st.set(16 + t, q)
    return grid, st


def moist_tracers(st, seed):
    """The seeded moisture perturbations of convection_state for a full-form
    state just initialised by gwm_init_convection: the tracer masses
    (rho qv, rho qc, rho qr) as FP32 arrays, to be set on this state or on
    others of the same grid (any precision or form)."""
    rng = np.random.default_rng(10_000 + int(seed))
    rho = st.get(0)
    qv = st.get(16)
    shape = qv.shape
    qv = qv * rng.uniform(0.85, 1.25, shape).astype(np.float32)
    qc = rho * (np.abs(rng.standard_normal(shape)) * 1.5e-3 *
                (rng.random(shape) < 0.3)).astype(np.float32)
    qr = rho * (np.abs(rng.standard_normal(shape)) * 0.8e-3 *
                (rng.random(shape) < 0.3)).astype(np.float32)
    return qv.astype(np.float32), qc.astype(np.float32), qr.astype(np.float32)


def generate(ctx, seeds, pairs_per_seed, shape=(512, 512, 64), wr=None, run=None):
    """Training pairs from the plain scheme on convection states of the
    given seeds: returns device arrays X [n][7], Y [n][4] and n.

    With `run` = (step_cfg, moist_wr, steps) the pairs come from plain-scheme
    runs: each seed's pairs are split evenly over its state at steps 0, 1,
    .., `steps` of moist steps (dynamics, tracers, warm rain with `moist_wr`),
    so the network also sees the near-saturated states the scheme leaves
    behind after its first adjustment, not only the forced start."""
    if ctx.precision != F32:
        raise ValueError("the surrogate works on FP32 contexts")
    wr = wr or warm_rain_config(dt=5.0, rain_terminal_velocity=12.0)
    per = int(pairs_per_seed)
    n = len(seeds) * per
    X = DeviceArray(ctx, (n, SURROGATE_IN))
    Y = DeviceArray(ctx, (n, SURROGATE_OUT))
    snaps = 1 if run is None else int(run[2]) + 1
    if snaps < 1 or per < snaps:
        raise ValueError("generate: need at least one pair per sampled state")
    counts = [per // snaps + (1 if k < per % snaps else 0) for k in range(snaps)]
    for s, seed in enumerate(seeds):
        _, st = convection_state(ctx, *shape, seed)
        row = s * per
        for k, cnt in enumerate(counts):
            if k:
                st.moist_step(run[0], run[1])
            surrogate_sample(st, wr, seed + 1009 * k, cnt, X, Y, row=row)
            row += cnt
        del st
    return X, Y, n


def fit_norm(x, y, box=(1.0, 99.0)):
    """Per-feature / per-target mean and standard deviation, and the
    fallback box [p_lo, p_hi] of each raw feature."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    xs = x.std(axis=0)
    ys = y.std(axis=0)
    xs[xs == 0] = 1.0
    ys[ys == 0] = 1.0
    return dict(x_mean=x.mean(axis=0), x_std=xs, y_mean=y.mean(axis=0), y_std=ys,
                box_lo=np.percentile(x, box[0], axis=0), box_hi=np.percentile(x, box[1], axis=0))


def spread_rows(arr, n, m, blocks=64):
    """About m rows of the first n rows of a device array, taken as `blocks`
    contiguous runs spread evenly over them (the pairs are stored seed by
    seed and state by state, so a prefix would see only the first ones)."""
    if m >= n:
        return arr.download_rows(n)
    blocks = max(1, min(blocks, m))
    run = m // blocks
    starts = np.linspace(0, n - run, blocks).astype(np.int64)
    return np.concatenate([arr.download_range(int(s), run) for s in starts])


def train(ctx, X, Y, n, seed=0, epochs=4, batch=4096, lr=2e-3, lr_final=5e-5, norm_rows=1_000_000):
    """A surrogate with its normalisation fitted on up to `norm_rows` pairs
    spread over all n, and trained on all n pairs."""
    sur = Surrogate(ctx, seed)
    m = min(n, norm_rows)
    norm = fit_norm(spread_rows(X, n, m), spread_rows(Y, n, m))
    sur.set_norm(**norm)
    losses = sur.train(X, Y, n, batch=batch, epochs=epochs, lr=lr, lr_final=lr_final)
    return sur, norm, losses


def tendency_errors(ctx, sur, seed=42, shape=(128, 128, 64), wr=None, run=None):
    """One warm-rain call with the surrogate and with the plain scheme from
    the same convection state: relative L2 error of the surrogate's changes
    of (qv, qc, qr, theta) against the plain changes, the same for the
    method without the network (no source terms at all), and the fallback
    count.  With `run` = (step_cfg, moist_wr, steps) the state is first
    advanced `steps` plain moist steps."""
    wr = wr or warm_rain_config(dt=5.0, rain_terminal_velocity=0.0)
    _, a = convection_state(ctx, *shape, seed)
    _, b = convection_state(ctx, *shape, seed)
    if run is not None:
        for _ in range(int(run[2])):
            a.moist_step(run[0], run[1])
            b.moist_step(run[0], run[1])

    def snap(st):
        rho = st.get(0).astype(np.float64)
        th = st.get(1) / np.maximum(rho, 1e-6)
        return [st.get(16 + t) / np.maximum(rho, 1e-6) for t in range(3)] + [th]

    before = snap(a)
    a.warm_rain(wr, metric_fallout=True)
    nfall = b.warm_rain_surrogate(wr, sur, metric_fallout=True)
    plain = [p - q for p, q in zip(snap(a), before)]
    surr = [p - q for p, q in zip(snap(b), before)]
    out = {}
    for name, dp, ds in zip(TARGETS, plain, surr):
        den = max(np.linalg.norm(dp), 1e-30)
        out[name] = {"surrogate": float(np.linalg.norm(ds - dp) / den),
                     "no_source_terms": float(np.linalg.norm(dp) / den)}
    cells = shape[0] * shape[1] * shape[2]
    out["fallback_fraction"] = nfall / cells
    return out
