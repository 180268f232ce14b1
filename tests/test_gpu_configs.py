"""The five benchmark configurations on the GPU, against the CPU oracle.

Each configuration runs at its own vertical levels, time step and physics,
at a horizontal size the oracle steps in seconds, on states that stay
finite (checked first: the reference's split-explicit scheme diverges at
the configurations' own dt, so those runs use the largest dt of a finite
trajectory, and the configuration's dt runs only as far as it stays finite).

These trajectories are ill-conditioned: the reference's acoustic scheme
amplifies grid-scale noise even where the state stays finite, so a random
one-ulp perturbation of the C1 initial rho*theta moves the FP64 oracle's own
w by 1e-10 after ten steps (a million-fold).  The device's and the host's
pow/exp differ by an ulp now and then at every operation, so the bars of
the single-operator tests (1e-11 FP64, 2e-5 FP32) do not carry over to
multi-step trajectories; the bars here are the conditioning's:
  - FP64 device vs FP64 oracle: relative L2 <= 1e-8 (1e-9 on the first C1
    step at 2 s);
  - FP32 device vs FP32 oracle: 5e-3 on the momenta (the FP32 ulp is 2^29
    times the FP64 one); the thermodynamic fields to 1e-5;
  - FP32 perturbation form (flat, doubly periodic) vs FP64 oracle: 1e-4,
    the north-star bar (its departures carry the pressure differences at
    their own relative precision; on the moist 512x512x64 case the bench
    measures 3.8e-5 after five steps and 3.0e-4 after ten, where the FP64
    step started from the FP32-rounded state is itself 1.5e-4 away);
  - FP32 full form vs FP64 oracle: 5e-3, the reference formulation's own
    error (one ulp of rho*theta moves p by ~10 mPa against ~1 Pa
    horizontal differences);
  - FP32 + surrogate vs FP64 plain oracle: stated per field, with dry mass
    and total water (+ precipitation) budgets.
Relative L2 uses a floor of 1e-3 of each field's physical scale (rho 1 kg
m-3, rho*theta 300 K kg m-3, momenta 1 kg m-2 s-1, water 1e-3 kg m-3), so a
field at rest is judged against a meaningful scale, not its own noise.
"""
import numpy as np
import pytest

import gwm_amd as G
import pyoracle as O

pytestmark = pytest.mark.gpu

FIELDS = (O.RHO, O.RHO_THETA, O.MOM_U, O.MOM_V, O.MOM_W)
WATER = (O.TRACER0, O.TRACER0 + 1, O.TRACER0 + 2)
SCALE = {O.RHO: 1.0, O.RHO_THETA: 300.0, O.MOM_U: 1.0, O.MOM_V: 1.0, O.MOM_W: 1.0,
         O.TRACER0: 1e-3, O.TRACER0 + 1: 1e-3, O.TRACER0 + 2: 1e-3}
NAMES = {O.RHO: "rho", O.RHO_THETA: "rho_theta", O.MOM_U: "u", O.MOM_V: "v", O.MOM_W: "w",
         O.TRACER0: "qv", O.TRACER0 + 1: "qc", O.TRACER0 + 2: "qr"}


@pytest.fixture(scope="module")
def ctx32():
    c = G.Context(0, G.F32)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ctx64():
    c = G.Context(0, G.F64)
    yield c
    c.close()


def rel_l2(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), floor * np.sqrt(b.size), 1e-300)


def errors(get, w, fields, floor_frac=1e-3):
    return {NAMES[f]: rel_l2(get(f), w.get(f), SCALE[f] * floor_frac) for f in fields}


class Checks:
    """Collects every comparison of a test, prints all of them, and fails
    at the end with the ones above their bars."""

    def __init__(self):
        self.bad = []

    def __call__(self, errs, tol, label, fields=None):
        errs = {k: v for k, v in errs.items() if fields is None or k in fields}
        print(label, {k: f"{v:.2e}" for k, v in errs.items()}, f"(bar {tol:g})")
        bad = {k: v for k, v in errs.items() if not v <= tol}
        if bad:
            self.bad.append(f"{label}: rel-L2 above {tol}: {bad}")

    def done(self):
        assert not self.bad, "\n".join(self.bad)


DRY = ("rho", "rho_theta")
MOM = ("u", "v", "w")


def check(errs, tol, label):
    c = Checks()
    c(errs, tol, label)
    c.done()


def finite(st, label):
    s = st.summary()
    assert s["nonfinite_values"] == 0, (label, s["nonfinite_values"])
    return s


def oracle_finite(w, label):
    s = w.summary()
    assert all(np.isfinite(s[k]) for k in ("w_min", "w_max", "u_max", "rho_min")), (label, s)
    return s


def to_oracle(st, w, fields, f64=False):
    for f in fields:
        w.set(f, st.get_f64(f) if f64 else st.get(f))


def to_state(w, st, fields):
    for f in fields:
        st.set(f, w.get(f).astype(st.ctx.dtype))


# -------------------------------------------------------------------- C1 ---
# Dry warm bubble: 64x64x40, dx = dy = 500 m, dz = 250 m (10 km top), flat,
# doubly periodic; theta 300 K hydrostatic with p_sfc = 1000 hPa; +2 K
# cos^2 bubble of radius 2 km at (16 km, 16 km, 2 km); 3 substeps.  At the
# configuration's dt of 2 s the scheme is finite for one step (|w| 14 m/s
# after it, 5e20 after the second), and finite at dt 0.5 s.

C1 = dict(nx=64, ny=64, nz=40, dx=500.0, dy=500.0, z_top=10000.0)
RHO_S = 1.0e5 / (287.05 * 300.0)


def c1_bubble(st, nx=64, ny=64):
    st.init_bubble(rho_background=-RHO_S, theta_background=300.0,
                   center=(16000.0 / ((nx - 1) * 500.0), 16000.0 / ((ny - 1) * 500.0), 0.2),
                   radius=(2000.0 / (nx * 500.0), 2000.0 / (ny * 500.0), 0.2),
                   theta_perturbation=2.0)


def c1_world(prec):
    w = O.World(C1["nx"], C1["ny"], C1["nz"], dx=500.0, dy=500.0, z_top=10000.0, precision=prec)
    w.init_bubble(rho_background=-RHO_S, theta_background=300.0,
                  center=(16000.0 / (63 * 500.0), 16000.0 / (63 * 500.0), 0.2),
                  radius=(2000.0 / (64 * 500.0), 2000.0 / (64 * 500.0), 0.2),
                  theta_perturbation=2.0)
    return w


def c1_grid(ctx):
    return G.Grid(ctx, C1["nx"], C1["ny"], C1["nz"], dx=500.0, dy=500.0, z_top=10000.0)


def test_c1_initial_state_and_first_step_at_configuration_dt(ctx64):
    """The verbatim configuration (dt 2 s, 3 substeps) for the one step the
    scheme stays finite; the device builds the initial state itself."""
    grid = c1_grid(ctx64)
    st = G.State(grid, 0)
    c1_bubble(st)
    w = c1_world(8)
    for f in FIELDS:
        assert np.array_equal(st.get(f), w.get(f)), NAMES[f]
    cfg = dict(dt=2.0, fast_substeps=3)
    st.ssprk3_step(G.step_config(**cfg))
    w.ssprk3_step(O.step_cfg(**cfg))
    finite(st, "C1 dt 2")
    oracle_finite(w, "C1 dt 2")
    # This step already amplifies perturbations violently (|w| reaches 14
    # m/s from rest, and 5e20 a step later), so an ulp of difference between
    # the device's and the host's FP64 pow/exp grows to ~1e-10 in the momenta.
    check(errors(st.get, w, FIELDS, 1e-6), 1e-9, "C1 FP64, dt 2 s, step 1")


def test_c1_ten_steps_fp64_fp32_and_perturbation_form(ctx32, ctx64):
    cfg = dict(dt=0.5, fast_substeps=3)
    steps = 10
    w64 = c1_world(8)
    w32 = c1_world(4)
    s64 = G.State(c1_grid(ctx64), 0)
    s32 = G.State(c1_grid(ctx32), 0)
    sp = G.State(c1_grid(ctx32), 0, perturbation=True)
    c1_bubble(s64)
    c1_bubble(sp)
    to_state(w32, s32, FIELDS)
    for _ in range(steps):
        for st in (s64, s32, sp):
            st.ssprk3_step(G.step_config(**cfg))
        for w in (w64, w32):
            w.ssprk3_step(O.step_cfg(**cfg))
    for st, lab in ((s64, "fp64"), (s32, "fp32"), (sp, "perturbation")):
        finite(st, lab)
    oracle_finite(w64, "C1 oracle")
    assert w64.summary()["w_max"] > 0.3  # the bubble rises
    c = Checks()
    c(errors(s64.get, w64, FIELDS, 1e-6), 1e-8, "C1 FP64 vs FP64 oracle")
    e32 = errors(s32.get, w32, FIELDS)
    c(e32, 1e-5, "C1 FP32 vs FP32 oracle, thermodynamic", DRY)
    c(e32, 5e-3, "C1 FP32 vs FP32 oracle, momenta", MOM)
    c(errors(sp.get_f64, w64, FIELDS), 1e-4, "C1 FP32 perturbation form vs FP64 oracle")
    c(errors(s32.get, w64, FIELDS), 5e-3, "C1 FP32 full form vs FP64 oracle")
    c.done()


# -------------------------------------------------------------------- C2 ---
# Mountain wave: dx = dy = 2 km, 60 hybrid-height levels stretched 50 m ->
# 600 m to a 20 km top, Witch-of-Agnesi ridge h = 400 m / (1 + (x/10 km)^2)
# uniform in y, U = 10 m/s, constant N = 0.01 s^-1 from theta_sfc = 288 K,
# relaxation lateral boundaries in x (y periodic), 6 substeps.  Reduced to
# 64 x 4 columns (the ridge and 27 km either side); at the configuration's
# dt of 6 s the explicit vertical acoustic step on 50 m levels overflows in
# the first step, dt 0.25 s stays finite.

def c2_pair(ctx, prec, nx=64, ny=4, nz=60):
    dx = 2000.0
    zi = G.stretched_levels(nz, 50.0, 600.0, 20000.0)
    ter = G.agnesi_ridge(nx, ny, dx, 0.5 * nx * dx, 10000.0, 400.0)
    kw = dict(dx=dx, dy=dx, z_interfaces=zi, terrain=ter, periodic_x=False)
    grid = G.Grid(ctx, nx, ny, nz, **kw)
    st = G.State(grid, 0)
    w = O.World(nx, ny, nz, precision=prec, **kw)
    return grid, st, w


def test_c2_mountain_wave_relaxation(ctx32, ctx64):
    cfg = dict(dt=0.25, fast_substeps=6)
    steps = 6
    g64, s64, w64 = c2_pair(ctx64, 8)
    g32, s32, w32 = c2_pair(ctx32, 4)
    # the constant-N state on the terrain-following levels, built by the library
    s64.init_constant_n(theta_surface=288.0, brunt_vaisala=0.01, p_surface=1.0e5, u=10.0)
    to_oracle(s64, w64, FIELDS)
    to_oracle(s64, w32, FIELDS)
    to_state(w32, s32, FIELDS)
    for st in (s64, s32):
        st.capture_boundary()
    for w in (w64, w32):
        w.set_boundary_from_current()
    bc = G.BoundaryConfig(G.BOUNDARY_RELAX, 5.0, G.BOUNDARY_FN(), None)
    for _ in range(steps):
        s64.ssprk3_step(G.step_config(**cfg), bc)
        s32.ssprk3_step(G.step_config(**cfg), bc)
        for w in (w64, w32):
            w.ssprk3_step(O.step_cfg(**cfg), boundary=O.BOUNDARY_RELAX, relax_width=5.0)
    finite(s64, "C2 fp64")
    finite(s32, "C2 fp32")
    s = oracle_finite(w64, "C2 oracle")
    assert s["w_max"] > 1e-3  # the ridge forces a wave
    c = Checks()
    c(errors(s64.get, w64, FIELDS, 1e-6), 1e-8, "C2 FP64 vs FP64 oracle")
    e32 = errors(s32.get, w32, FIELDS)
    c(e32, 1e-5, "C2 FP32 vs FP32 oracle, thermodynamic", DRY)
    c(e32, 5e-3, "C2 FP32 vs FP32 oracle, momenta", MOM)
    # The FP32 full form on terrain: the perturbation form is built for flat,
    # doubly periodic grids only, so this is the reference formulation's
    # FP32 error (~3e-3 on w, whose amplitude is ~0.05 m/s after 1.5 s).
    c(errors(s32.get, w64, FIELDS), 1e-2, "C2 FP32 full form vs FP64 oracle")
    c.done()


def test_c2_at_its_size_stays_inside_the_metric_arrays(ctx32, ctx64):
    """C2 at 256x256x60 columns, one step in FP64 and FP32.  Its FP64
    metric arrays end exactly on a 2 MiB boundary (256*256*60*8 B = 15 *
    2 MiB), where a load one element past the end is not mapped: the slow
    stage's closed outer x faces on the open sides load the metric of the
    column beyond the edge before selecting their zero flux, and that
    column is clamped to the edge one.  Both steps stay finite.  FP32
    against FP64: rho, rho*theta and the horizontal momenta to 1e-5; w,
    after one step only the start of the wave, carries the full form's
    pressure rounding (one ulp of rho*theta moves p by ~10 mPa against
    ~1 Pa horizontal differences): 1.4 % measured, bar 5 %."""
    nx, ny, nz, dx = 256, 256, 60, 2000.0
    zi = G.stretched_levels(nz, 50.0, 600.0, 20000.0)
    ter = G.agnesi_ridge(nx, ny, dx, 0.5 * nx * dx, 10000.0, 400.0)
    cfg = G.step_config(dt=0.25, fast_substeps=6)
    bc = G.BoundaryConfig(G.BOUNDARY_RELAX, 5.0, G.BOUNDARY_FN(), None)
    out = {}
    for name, ctx in (("fp64", ctx64), ("fp32", ctx32)):
        grid = G.Grid(ctx, nx, ny, nz, dx=dx, dy=dx, z_interfaces=zi, terrain=ter,
                      periodic_x=False)
        st = G.State(grid, 0)
        if name == "fp64":
            st.init_constant_n(theta_surface=288.0, brunt_vaisala=0.01, p_surface=1.0e5,
                               u=10.0)
            init = [st.get(f) for f in FIELDS]
        else:
            for f, v in zip(FIELDS, init):
                st.set(f, v.astype(np.float32))
        st.capture_boundary()
        st.ssprk3_step(cfg, bc)
        finite(st, f"C2 {name}")
        out[name] = (grid, st)
    e = errors(out["fp32"][1].get, _StateAsWorld(out["fp64"][1]), FIELDS)
    c = Checks()
    c(e, 1e-5, "C2 256x256x60 FP32 full form vs FP64, one step", DRY + ("u", "v"))
    c(e, 5e-2, "C2 256x256x60 FP32 full form vs FP64, one step", ("w",))
    c.done()


class _StateAsWorld:
    """A device state read like an oracle world (for errors())."""

    def __init__(self, st):
        self.st = st

    def get(self, f):
        return self.st.get(f)


# -------------------------------------------------------------------- C3 ---
# Warm-rain convection: dx = 1 km, 64 levels stretched 100 m -> 500 m to
# 16 km, doubly periodic, the analytic sounding with +/-0.5 K theta noise
# below 1 km, qc = qr = 0, warm rain with 12 m/s metric fallout, 6
# substeps.  Reduced to 32 x 24 columns; dt 0.25 s (5 s overflows in the
# first step on 100 m levels).

def c3_levels():
    return G.stretched_levels(64, 100.0, 500.0, 16000.0)


def c3_grid(ctx, nx=32, ny=24):
    return G.Grid(ctx, nx, ny, 64, dx=1000.0, dy=1000.0, z_interfaces=c3_levels())


def c3_world(prec, nx=32, ny=24):
    return O.World(nx, ny, 64, dx=1000.0, dy=1000.0, z_interfaces=c3_levels(), n_tracers=3,
                   precision=prec)


def moist_step_oracle(w, dt, ns):
    w.ssprk3_step(O.step_cfg(dt=dt, fast_substeps=ns))
    w.tracers_ssprk3(O.step_cfg(dt=dt, fast_substeps=ns))
    w.warm_rain(O.warm_rain_cfg(dt=dt, rain_terminal_velocity=12.0), metric_fallout=True)


def test_c3_convection_levels(ctx32, ctx64):
    dt, ns, steps = 0.25, 6, 4
    s64 = G.State(c3_grid(ctx64), 3)
    s32 = G.State(c3_grid(ctx32), 3)
    sp = G.State(c3_grid(ctx32), 3, perturbation=True)
    s64.init_convection(1.0e5, 42, 0.5, 1000.0)
    sp.init_convection(1.0e5, 42, 0.5, 1000.0)
    w64, w32 = c3_world(8), c3_world(4)
    to_oracle(s64, w64, FIELDS + (O.TRACER0,))
    to_oracle(s64, w32, FIELDS + (O.TRACER0,))
    to_state(w32, s32, FIELDS + (O.TRACER0,))
    cfg = G.step_config(dt=dt, fast_substeps=ns)
    wr = G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0)
    for _ in range(steps):
        for st in (s64, s32, sp):
            st.moist_step(cfg, wr)
        for w in (w64, w32):
            moist_step_oracle(w, dt, ns)
    for st, lab in ((s64, "fp64"), (s32, "fp32"), (sp, "perturbation")):
        finite(st, lab)
    oracle_finite(w64, "C3 oracle")
    f = FIELDS + (O.TRACER0,)
    c = Checks()
    c(errors(s64.get, w64, f, 1e-6), 1e-8, "C3 FP64 vs FP64 oracle")
    e32 = errors(s32.get, w32, f)
    c(e32, 1e-5, "C3 FP32 vs FP32 oracle, thermodynamic", DRY + ("qv",))
    c(e32, 5e-3, "C3 FP32 vs FP32 oracle, momenta", MOM)
    c(errors(sp.get_f64, w64, f), 1e-4, "C3 FP32 perturbation form vs FP64 oracle")
    c(errors(s32.get, w64, f), 5e-3, "C3 FP32 full form vs FP64 oracle")
    c.done()


# -------------------------------------------------------------------- C4 ---
# C3 with the warm-rain surrogate (trained here on plain-scheme pairs from
# four-step plain runs of convection states of seeds 0-1), evaluated from a
# seed-42 state with cloudy, raining and supersaturated cells (qv scaled by
# U(0.85, 1.25), cloud and rain in 30 % of the cells), so the source terms
# are large.

C4_DT, C4_NS, C4_STEPS = 0.25, 6, 4


def c4_run():
    return (G.step_config(dt=C4_DT, fast_substeps=C4_NS),
            G.warm_rain_config(dt=C4_DT, rain_terminal_velocity=12.0), C4_STEPS)


@pytest.fixture(scope="module")
def trained(ctx32):
    import gwm_amd.surrogate as SG
    X, Y, n = SG.generate(ctx32, [0, 1], 250_000, shape=(64, 64, 64),
                          wr=G.warm_rain_config(dt=C4_DT, rain_terminal_velocity=0.0),
                          run=c4_run())
    sur, norm, losses = SG.train(ctx32, X, Y, n, seed=0, epochs=6, batch=1024, lr=3e-3,
                                 lr_final=1e-4)
    print("surrogate epoch losses", losses)
    return sur, losses


def column_mass(st, dz):
    """Sum over columns of rho dz, the dry mass the flux form conserves on
    stretched levels (the summary's total_dry_mass is the plain sum of
    rho, which vertical motion moves between thin and thick levels)."""
    rho = st.get(0).astype(np.float64)
    return float(np.sum(dz * rho.reshape(rho.shape[0], -1).sum(axis=1)))


def water(summary):
    return (summary["vapor_path"] + summary["cloud_path"] + summary["rain_path"] +
            summary["precip_sum"])


def test_c4_surrogate_step_against_fp64_plain_oracle(ctx32, ctx64, trained):
    import gwm_amd.surrogate as SG
    sur, losses = trained
    assert losses[-1] < 0.5 * losses[0]
    dt, ns, steps = C4_DT, C4_NS, C4_STEPS
    _, s_sur = SG.convection_state(ctx32, 32, 24, 64, 42)
    _, s_plain = SG.convection_state(ctx32, 32, 24, 64, 42)
    _, s_none = SG.convection_state(ctx32, 32, 24, 64, 42)
    w64 = c3_world(8)
    to_oracle(s_sur, w64, FIELDS + WATER)
    cfg = G.step_config(dt=dt, fast_substeps=ns)
    wr = G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0)
    dz = np.diff(G.stretched_levels(64, 100.0, 500.0, 16000.0))
    b0 = s_sur.summary()
    masses, waters, nfall = [column_mass(s_sur, dz)], [water(b0)], []
    masses_plain = [column_mass(s_plain, dz)]
    for _ in range(steps):
        s_sur.moist_step_surrogate(cfg, wr, sur)
        s_plain.moist_step(cfg, wr)
        s_none.ssprk3_step(cfg)
        s_none.tracers_ssprk3(cfg)
        moist_step_oracle(w64, dt, ns)
        b = finite(s_sur, "surrogate")
        masses.append(column_mass(s_sur, dz))
        masses_plain.append(column_mass(s_plain, dz))
        waters.append(water(b))
    finite(s_plain, "plain")
    finite(s_none, "no microphysics")
    oracle_finite(w64, "C4 oracle")
    f = FIELDS + WATER
    e_sur = errors(s_sur.get, w64, f)
    e_plain = errors(s_plain.get, w64, f)
    e_none = errors(s_none.get, w64, f)
    print("FP32 plain vs FP64 plain oracle", {k: f"{v:.2e}" for k, v in e_plain.items()})
    print("FP32 without microphysics vs FP64 plain oracle",
          {k: f"{v:.2e}" for k, v in e_none.items()})
    print("C4 FP32+surrogate vs FP64 plain oracle", {k: f"{v:.2e}" for k, v in e_sur.items()})
    # The surrogate's error is its source-term error carried by the state:
    # the water fields directly, the dry fields through latent heating and
    # the buoyancy it drives.  The bar is what the surrogate is for: on
    # every field it must remove most (75 %) of the difference that leaving
    # the microphysics out makes after these steps, so each ratio below is
    # at most 0.25; an untrained or misplaced network scores about 1.
    ratio = {k: e_sur[k] / e_none[k] for k in e_sur}
    c = Checks()
    c(ratio, 0.25, "C4 surrogate error / error without microphysics")
    c(e_plain, 1e-4, "C4 FP32 plain vs FP64 plain oracle, thermodynamic", DRY + ("qv", "qr"))
    c.done()
    # Budgets.  The warm rain does not touch rho, so the surrogate leaves
    # the column dry mass to the FP32 full form's own drift (~3e-8 a step
    # here, measured on the plain run beside it; the perturbation form holds
    # it to 1e-12, tests/test_perturbation.py).  The surrogate's conservation
    # layer keeps total water + the precipitation bucket at the FP32 plain
    # scheme's drift (~3e-8 a step).
    dm = np.abs(np.diff(masses)) / masses[0]
    dm_plain = np.abs(np.diff(masses_plain)) / masses_plain[0]
    dw = np.abs(np.diff(waters)) / waters[0]
    print("dry-mass drift per step", dm, "plain", dm_plain, "water drift per step", dw)
    assert np.all(dm <= 1.5 * dm_plain.max() + 1e-10), (dm, dm_plain)
    assert np.all(dw <= 2e-7), dw


def test_c4_training_gains_over_untrained_network(ctx32, trained):
    """What training buys: the trained network's source terms against the
    plain scheme, next to the same network untrained and to no source terms
    at all (error 1 by definition).  The plain scheme itself is the truth."""
    import gwm_amd.surrogate as SG
    sur, _ = trained
    wr = G.warm_rain_config(dt=0.25, rain_terminal_velocity=0.0)
    e_tr = SG.tendency_errors(ctx32, sur, seed=42, shape=(48, 48, 64), wr=wr)
    raw = G.Surrogate(ctx32, seed=0)
    raw.set_norm(**{k: v for k, v in sur.norm().items()})
    e_un = SG.tendency_errors(ctx32, raw, seed=42, shape=(48, 48, 64), wr=wr)
    print("trained", e_tr)
    print("untrained", e_un)
    # the same after two plain steps, where the first adjustment has left
    # most cells near saturation and the source terms are small
    e_late = SG.tendency_errors(ctx32, sur, seed=42, shape=(48, 48, 64), wr=wr,
                                run=c4_run()[:2] + (2,))
    print("trained, after two plain steps", e_late)
    for name in SG.TARGETS:
        assert e_tr[name]["surrogate"] < 0.35, (name, e_tr[name])
        assert e_tr[name]["surrogate"] < 0.2 * e_un[name]["surrogate"], name
    assert 0.0 < e_tr["fallback_fraction"] < 0.5


# -------------------------------------------------------------------- C5 ---
# Uneven ensemble members, dx = 1 km, uniform levels to 16 km: half warm
# bubbles (+1 to +3 K, seeded centre, radius 2 km at 2 km) and half density
# currents (-15 K, radius 4 km at 3 km), on the convection sounding, 4
# substeps.  dt 0.5 s (2 s overflows within three steps, 1 s within twenty
# on 60 levels), members advanced together from their captured graphs, each
# checked against its oracle.

def c5_member(ctx, nx, ny, nz, m, warm, rng):
    zi = np.linspace(0.0, 16000.0, nz + 1)
    grid = G.Grid(ctx, nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi)
    st = G.State(grid, 3)
    st.init_convection(1.0e5, 100 + m, 0.5, 1000.0)
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
    rho = st.get(O.RHO).astype(np.float64)
    st.set(O.RHO_THETA, (st.get(O.RHO_THETA) + rho * dth).astype(st.ctx.dtype))
    w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi, n_tracers=3,
                precision=8 if st.ctx.dtype == np.float64 else 4)
    return grid, st, w


@pytest.mark.parametrize("prec", [8, 4])
def test_c5_uneven_members_against_oracle(prec):
    shapes = [(32, 24, 40), (40, 32, 48), (24, 40, 60), (48, 16, 40)]
    dt, ns, steps = 0.5, 4, 4
    rng = np.random.default_rng(0)
    ctxs = [G.Context(0, G.F64 if prec == 8 else G.F32) for _ in shapes]
    keep, members, worlds = [], [], []
    for m, (ctx, shape) in enumerate(zip(ctxs, shapes)):
        grid, st, w = c5_member(ctx, *shape, m, m % 2 == 0, rng)
        to_oracle(st, w, FIELDS + (O.TRACER0,))
        keep.append(grid)
        members.append(st)
        worlds.append(w)
    G.ensemble_steps(members, G.step_config(dt=dt, fast_substeps=ns),
                     G.warm_rain_config(dt=dt, rain_terminal_velocity=12.0), steps)
    for w in worlds:
        for _ in range(steps):
            moist_step_oracle(w, dt, ns)
    f = FIELDS + WATER
    c = Checks()
    for m, (st, w) in enumerate(zip(members, worlds)):
        finite(st, f"member {m}")
        oracle_finite(w, f"member {m}")
        if prec == 8:
            c(errors(st.get, w, f, 1e-6), 1e-8, f"C5 member {m} {shapes[m]} FP64")
        else:
            e32 = errors(st.get, w, f)
            c(e32, 1e-5, f"C5 member {m} {shapes[m]} FP32 vs FP32 oracle, thermodynamic",
              DRY + ("qv",))
            c(e32, 5e-3, f"C5 member {m} {shapes[m]} FP32 vs FP32 oracle, momenta", MOM)
    c.done()
    del members, keep
    for c in ctxs:
        c.close()
