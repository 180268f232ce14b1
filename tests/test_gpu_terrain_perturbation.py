"""The FP32 perturbation form on terrain, against the FP64 step.

On terrain the perturbation form keeps a 3-D base (the FP64 state it was
built from, cell by cell), stores rho and rho*theta as FP32 departures from
it, and holds the base's own pressure-gradient and gravity forcing per face,
so only departures pass through FP32 pressure arithmetic.  Its K14 acts on
the departures without a reference state: the column reference's terms
cancel between its recurrence and K14 up to the recurrence's residual
(dry_fast_modes.cu:108-158), which the z-face base forcing carries,
computed on the base column; the FP64 step with the column reference is
the reference here.

The cases are the mountain wave (Agnesi ridge 400 m, half-width 10 km,
constant N = 0.01 s-1, u = 10 m/s, levels stretched 50-600 m to 20 km,
dt 0.25 s, 6 substeps) at reduced horizontal size, with relaxation at open
x sides and doubly periodic.  The FP64 step is held to the oracle on the
same configurations by test_gpu_configs.py and test_gpu_parity.py.
"""
import os
import tempfile

import numpy as np
import pytest

import gwm_amd as G

pytestmark = pytest.mark.gpu

FIELDS = (0, 1, 2, 3, 4)
NAMES = ("rho", "rho_theta", "mom_u", "mom_v", "mom_w")
SCALE = {0: 1.0, 1: 300.0, 2: 1.0, 3: 1.0, 4: 1.0}


@pytest.fixture(scope="module")
def ctxs():
    c32, c64 = G.Context(0, G.F32), G.Context(0, G.F64)
    yield c32, c64
    c32.close()
    c64.close()


def mountain(ctx, nx, ny, nz, periodic_x, perturbation=False, steps=0, dt=0.25):
    dx = 2000.0
    zi = G.stretched_levels(nz, 50.0, 600.0, 20000.0)
    ter = G.agnesi_ridge(nx, ny, dx, 0.5 * nx * dx, 10000.0, 400.0)
    grid = G.Grid(ctx, nx, ny, nz, dx=dx, dy=dx, z_interfaces=zi, terrain=ter,
                  periodic_x=periodic_x)
    st = G.State(grid, 0, perturbation=perturbation)
    st.init_constant_n(theta_surface=288.0, brunt_vaisala=0.01, p_surface=1.0e5, u=10.0)
    bc = None
    if not periodic_x:
        st.capture_boundary()
        bc = G.BoundaryConfig(G.BOUNDARY_RELAX, 5.0, G.BOUNDARY_FN(), None)
    cfg = G.step_config(dt=dt, fast_substeps=6)
    for _ in range(steps):
        if bc is None:
            st.ssprk3_step(cfg)
        else:
            st.ssprk3_step(cfg, bc)
    return grid, st


def rel_l2(ref, st):
    out = {}
    for f in FIELDS:
        r = ref.get(f).astype(np.float64)
        x = st.get_f64(f)
        den = max(np.linalg.norm(r), 1e-3 * SCALE[f] * np.sqrt(r.size))
        out[NAMES[f]] = float(np.linalg.norm(x - r) / den)
    return out


def test_terrain_perturbation_starts_on_the_fp64_state(ctxs):
    """The base is the FP64 initial state itself (built on the FP64
    metrics), so the departures start at zero and the FP64 readout equals
    the FP64 context's state bit for bit; the dry-mass sum agrees to FP64
    rounding."""
    c32, c64 = ctxs
    _, ref = mountain(c64, 64, 8, 30, False)
    _, st = mountain(c32, 64, 8, 30, False, perturbation=True)
    assert st.perturbation
    for f in (0, 1):
        assert np.array_equal(st.get_f64(f), ref.get(f)), NAMES[f]
    for f in (2, 3, 4):
        assert np.array_equal(st.get(f), ref.get(f).astype(np.float32)), NAMES[f]
    m32, m64 = st.summary()["total_dry_mass"], ref.summary()["total_dry_mass"]
    assert abs(m32 - m64) <= 1e-13 * abs(m64)


@pytest.mark.parametrize("periodic_x", [False, True], ids=["relax_open_x", "periodic"])
def test_terrain_perturbation_against_fp64(ctxs, periodic_x):
    """Twelve steps (3 s): every field of the perturbation form within
    relative L2 1e-4 of the FP64 step, and w at least a hundred times closer
    than the full FP32 form's (whose one-ulp rho*theta rounding moves p by
    ~10 mPa against the young wave's differences)."""
    c32, c64 = ctxs
    nx, ny, nz, steps = 96, 8, 40, 12
    _, ref = mountain(c64, nx, ny, nz, periodic_x, steps=steps)
    _, tp = mountain(c32, nx, ny, nz, periodic_x, perturbation=True, steps=steps)
    _, full = mountain(c32, nx, ny, nz, periodic_x, steps=steps)
    for s in (ref, tp, full):
        assert s.summary()["nonfinite_values"] == 0
    e_tp, e_full = rel_l2(ref, tp), rel_l2(ref, full)
    print("perturbation form", e_tp)
    print("full form", e_full)
    assert max(e_tp.values()) <= 1e-4, e_tp
    assert e_tp["mom_w"] * 100 <= e_full["mom_w"], (e_tp, e_full)


def test_terrain_perturbation_checkpoint_restart_is_bit_exact(ctxs):
    """The restart dump carries the 3-D FP64 bases: a restored state steps
    on bit for bit, boundary reference included."""
    c32, _ = ctxs
    grid, a = mountain(c32, 64, 8, 30, False, perturbation=True, steps=2)
    cfg = G.step_config(dt=0.25, fast_substeps=6)
    bc = G.BoundaryConfig(G.BOUNDARY_RELAX, 5.0, G.BOUNDARY_FN(), None)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "tp.ckpt")
        a.checkpoint_write(path)
        b = G.State(grid, 0, perturbation=True)
        b.checkpoint_read(path)
    for f in FIELDS:
        assert np.array_equal(a.get_f64(f), b.get_f64(f)), NAMES[f]
    for _ in range(2):
        a.ssprk3_step(cfg, bc)
        b.ssprk3_step(cfg, bc)
    for f in FIELDS:
        assert np.array_equal(a.get_f64(f), b.get_f64(f)), NAMES[f]


def test_terrain_perturbation_refuses_what_it_does_not_carry(ctxs):
    c32, _ = ctxs
    zi = G.stretched_levels(20, 50.0, 600.0, 20000.0)
    ter = G.agnesi_ridge(64, 8, 2000.0, 64000.0, 10000.0, 400.0)
    grid = G.Grid(c32, 64, 8, 20, dx=2000.0, dy=2000.0, z_interfaces=zi, terrain=ter,
                  periodic_x=False)
    with pytest.raises(G.GwmError, match="no tracers"):
        G.State(grid, 3, perturbation=True)
    st = G.State(grid, 0, perturbation=True)
    st.init_constant_n(theta_surface=288.0, brunt_vaisala=0.01, p_surface=1.0e5, u=10.0)
    with pytest.raises(G.GwmError, match="terrain"):
        st.plan_view(0, 0, 0.0)


def test_terrain_perturbation_single_operators(ctxs):
    """From the same (initial) state: the slow tendencies (whose pressure
    gradient is the departures' plus the base forcing of each face) and one
    fast subcycle (theta and rho*theta rebuilt from the cell's base) against
    the FP64 operators, beside the full FP32 form.  The momentum tendencies,
    where pressure enters, within 1e-5 (full form: 5.6e-3 on u); the mass
    and rho*theta tendencies, FP32 flux differences of the stored momenta
    in either form, no worse than the full form's."""
    c32, c64 = ctxs
    cfg = G.step_config(dt=0.25, fast_substeps=6)
    st = {}
    for key, ctx, pert in (("fp64", c64, False), ("tp", c32, True), ("full", c32, False)):
        _, s = mountain(ctx, 64, 8, 30, False, perturbation=pert)
        s.slow_tendencies(cfg)
        st[key] = s
    for f in (G.TEND_RHO, G.TEND_RHO_THETA, G.TEND_U, G.TEND_W):
        r = st["fp64"].get(f).astype(np.float64)
        e = {k: np.linalg.norm(st[k].get(f).astype(np.float64) - r) / max(np.linalg.norm(r), 1e-30)
             for k in ("tp", "full")}
        print("tendency", f, e)
        if f in (G.TEND_U, G.TEND_W):
            assert e["tp"] <= 1e-5, (f, e)
        else:
            assert e["tp"] <= e["full"], (f, e)
    for s in st.values():
        s.fast_modes(cfg)
    e = rel_l2(st["fp64"], st["tp"])
    print("fast subcycle", e)
    assert max(e.values()) <= 1e-6, e


def test_terrain_perturbation_with_boundary_strips(ctxs):
    """The strip boundary overwrites the outermost owned strips from the
    captured reference, which the perturbation form holds as departures
    (a linear map either way): eight steps against the FP64 step."""
    c32, c64 = ctxs
    cfg = G.step_config(dt=0.25, fast_substeps=6)
    bc = G.BoundaryConfig(G.BOUNDARY_STRIPS, 0.0, G.BOUNDARY_FN(), None)
    out = []
    for ctx, pert in ((c64, False), (c32, True)):
        _, st = mountain(ctx, 64, 8, 30, False, perturbation=pert)
        for _ in range(8):
            st.ssprk3_step(cfg, bc)
        assert st.summary()["nonfinite_values"] == 0
        out.append(st)
    e = rel_l2(out[0], out[1])
    print("strips", e)
    assert max(e.values()) <= 1e-4, e


def test_boundary_callback_view_reads_full_values(ctxs):
    """A boundary plug-in handed a stage view of a perturbation-form state
    reads full values through it (the view converts with its owner's
    bases), as from a full-form state: at the step's entry the view's rho
    is the state's, bit for bit, on the owned cells."""
    import ctypes
    c32, _ = ctxs
    grid, st = mountain(c32, 64, 8, 30, True, perturbation=True)
    rho0 = st.get(0)
    shape, h = st.storage_shape(0), grid.halo
    seen = []

    def fn(hd, t, _user):
        buf = np.zeros(int(np.prod(shape)), np.float32)
        rc = G.lib().gwm_state_download(hd, 0, buf.ctypes.data_as(ctypes.c_void_p))
        owned = buf.reshape(shape)[:, h:shape[1] - h, h:shape[2] - h]
        seen.append((t, rc, owned.copy()))
        return 0

    bc = G.BoundaryConfig(G.BOUNDARY_NONE, 0.0, G.BOUNDARY_FN(fn), None)
    st.ssprk3_step(G.step_config(dt=0.25, fast_substeps=6), bc)
    assert [x[0] for x in seen] == [0.0, 0.25, 0.125]
    assert all(x[1] == 0 for x in seen)
    assert np.array_equal(seen[0][2], rho0)
    for _, _, r in seen[1:]:
        assert np.max(np.abs(r - rho0)) <= 1e-4 * np.max(rho0)
