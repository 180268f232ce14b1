"""Warm-rain surrogate on the GPU, through the C-ABI.

- the training-data sampler reproduces the plain scheme (the FP32 oracle's
  warm rain without fallout) at the sampled cells;
- the FP32 training kernels' loss and gradient match the FP64 oracle
  (oracle/surrogate_oracle.py, itself pinned by finite differences);
- the BF16 matrix-core inference matches the oracle's BF16 emulation, which
  also pins the MFMA operand layouts;
- with a box that excludes every cell, the surrogate warm rain equals the
  plain warm rain bit for bit (fallback path and separate fallout);
- training on plain-scheme pairs cuts the loss, and the trained network's
  source terms beat the method without it (no source terms) by a wide margin.
"""
import numpy as np
import pytest

import gwm_amd as G
import gwm_amd.surrogate as SG
import pyoracle as O
import surrogate_oracle as SO

pytestmark = pytest.mark.gpu

WR = dict(dt=5.0, rain_terminal_velocity=0.0)


@pytest.fixture(scope="module")
def ctx():
    c = G.Context(0, G.F32)
    yield c
    c.close()


def small_state(ctx, seed=3, shape=(24, 20, 16)):
    return SG.convection_state(ctx, *shape, seed, dz_bottom=150.0, dz_top=900.0, z_top=8000.0)


def sample(ctx, st, n, seed=7):
    X = G.DeviceArray(ctx, (n, 7))
    Y = G.DeviceArray(ctx, (n, 4))
    C = G.DeviceArray(ctx, (n,), np.int64)
    G.surrogate_sample(st, G.warm_rain_config(**WR), seed, n, X, Y, C)
    return X, Y, C


def test_sampler_reproduces_plain_scheme(ctx):
    grid, st = small_state(ctx)
    nx, ny, nz = 24, 20, 16
    n = 3000
    X, Y, C = sample(ctx, st, n)
    x, y, cells = X.download(), Y.download(), C.download()
    assert cells.min() >= 0 and cells.max() < nx * ny * nz
    zi = G.stretched_levels(nz, 150.0, 900.0, 8000.0)
    w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi, n_tracers=3, precision=4)
    for f in (O.RHO, O.RHO_THETA, O.MOM_U, O.MOM_V, O.MOM_W, 16, 17, 18):
        w.set(f, st.get(f))
    rho = st.get(O.RHO).astype(np.float64).ravel()
    before = [st.get(16 + t).astype(np.float64).ravel() / rho for t in range(3)]
    th0 = st.get(O.RHO_THETA).astype(np.float64).ravel() / rho
    w.warm_rain(O.warm_rain_cfg(**WR), metric_fallout=True)
    after = [w.get(16 + t).astype(np.float64).ravel() / rho for t in range(3)]
    th1 = w.get(O.RHO_THETA).astype(np.float64).ravel() / rho
    ref = np.stack([after[0] - before[0], after[1] - before[1], after[2] - before[2],
                    th1 - th0], axis=1)[cells]
    # features: rho and the mixing ratios are the cell's own values
    assert np.allclose(x[:, 2], rho[cells], rtol=1e-7)
    assert np.allclose(x[:, 3], np.maximum(before[0][cells], 0), rtol=1e-6, atol=1e-12)
    assert np.all(x[:, 6] == 5.0)
    # targets: FP32 device vs FP32 host; ulp-level differences of powf/expf
    # and of the rho*q / rho round trip
    scale = np.abs(ref).max(axis=0)
    err = np.abs(y - ref).max(axis=0)
    assert np.all(err <= 2e-4 * scale + 1e-9), (err, scale)
    assert (np.abs(ref[:, 1]) > 0).mean() > 0.2  # the sample exercises the source terms


def _norm_of(x, y):
    return SG.fit_norm(x, y)


def test_loss_and_gradient_match_oracle(ctx):
    _, st = small_state(ctx, seed=4)
    n = 2048
    X, Y, _ = sample(ctx, st, n, seed=11)
    x, y = X.download(), Y.download()
    norm = _norm_of(x, y)
    sur = G.Surrogate(ctx, seed=5)
    sur.set_norm(**norm)
    p = SO.init_xavier(9).astype(np.float32)
    sur.set_params(p)
    loss, grad = sur.loss_grad(X, Y, n)
    n32 = {k: np.asarray(v, np.float32).astype(np.float64) for k, v in sur.norm().items()}
    # the device normalises features and targets in FP32 before its FP32
    # layers; the oracle does the same rounding and then runs in FP64
    ref_loss, ref_grad = SO.loss_grad(p.astype(np.float64), x, y, n32, fp32_normalise=True)
    print("loss", loss, ref_loss)
    assert abs(loss - ref_loss) <= 1e-5 * ref_loss, (loss, ref_loss)
    rel = np.linalg.norm(grad - ref_grad) / np.linalg.norm(ref_grad)
    assert rel < 1e-4, rel
    # every layer's block individually
    for lo, hi in ((0, 512), (512, 4672), (4672, 8832), (8832, 9092)):
        r = np.linalg.norm(grad[lo:hi] - ref_grad[lo:hi]) / np.linalg.norm(ref_grad[lo:hi])
        assert r < 1e-4, (lo, r)


def test_bf16_inference_matches_oracle(ctx):
    _, st = small_state(ctx, seed=6)
    n = 1000  # not a multiple of the 16-cell MFMA group
    X, Y, _ = sample(ctx, st, n, seed=13)
    x, y = X.download(), Y.download()
    sur = G.Surrogate(ctx, seed=3)
    sur.set_norm(**_norm_of(x, y))
    p = (SO.init_xavier(4) * 1.5).astype(np.float32)
    sur.set_params(p)
    out = G.DeviceArray(ctx, (n, 4))
    sur.infer(X, n, out)
    got = out.download()
    norm = {k: np.asarray(v, np.float64) for k, v in sur.norm().items()}
    emu = SO.forward_bf16(p, x, norm)
    exact = SO.forward(p, SO.normalise_x(x, norm))[2] * norm["y_std"] + norm["y_mean"]
    yn = lambda a: (a - norm["y_mean"]) / norm["y_std"]  # noqa: E731
    # BF16 rounding of an activation can differ by one unit where the FP32
    # device sum and the host sum straddle a rounding boundary.
    print("finite", np.isfinite(got).all(), np.isfinite(emu).all(),
          "max |got|", np.abs(got).max(), "max |emu|", np.abs(emu).max())
    assert np.isfinite(got).all() and np.isfinite(emu).all()
    e_emu = np.linalg.norm(yn(got) - yn(emu)) / np.linalg.norm(yn(emu))
    e_exact = np.linalg.norm(yn(got) - yn(exact)) / np.linalg.norm(yn(exact))
    assert e_emu < 1e-2, e_emu
    assert e_exact < 5e-2, e_exact


def test_fallback_everywhere_equals_plain_warm_rain(ctx):
    wr = G.warm_rain_config(dt=5.0, rain_terminal_velocity=12.0)
    for metric in (True, False):
        _, a = small_state(ctx, seed=8)
        _, b = small_state(ctx, seed=8)
        if not metric:
            wr = G.warm_rain_config(dt=5.0, rain_terminal_velocity=12.0,
                                    sedimentation_layer_depth_m=500.0)
        sur = G.Surrogate(ctx, seed=1)
        # a box no cell can be inside (T in [-1, -1])
        sur.set_norm(x_mean=np.zeros(7), x_std=np.ones(7), y_mean=np.zeros(4), y_std=np.ones(4),
                     box_lo=-np.ones(7), box_hi=-np.ones(7))
        a.warm_rain(wr, metric_fallout=metric)
        nfall = b.warm_rain_surrogate(wr, sur, metric_fallout=metric)
        assert nfall == 24 * 20 * 16
        for f in (O.RHO_THETA, 16, 17, 18, O.PRECIP):
            assert np.array_equal(a.get(f), b.get(f)), (metric, f)


def test_training_cuts_loss_and_beats_no_source_terms(ctx):
    seeds = [0, 1]
    X, Y, n = SG.generate(ctx, seeds, 100_000, shape=(64, 64, 32),
                          wr=G.warm_rain_config(**WR))
    sur, norm, losses = SG.train(ctx, X, Y, n, seed=0, epochs=6, batch=1024, lr=3e-3,
                                 lr_final=1e-4)
    print("epoch losses", losses)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.25 * losses[0]
    errs = SG.tendency_errors(ctx, sur, seed=42, shape=(48, 48, 32),
                              wr=G.warm_rain_config(**WR))
    print("tendency errors", errs)
    # the method without the network (no source terms) has error 1 by
    # definition; the trained network must recover most of the change
    for name in SG.TARGETS:
        assert errs[name]["surrogate"] < 0.35, (name, errs[name])
    assert 0.0 < errs["fallback_fraction"] < 0.5


def test_moist_step_surrogate_keeps_dry_mass(ctx):
    grid, st = small_state(ctx, seed=9)
    sur = G.Surrogate(ctx, seed=2)
    x = np.zeros((1, 7))
    sur.set_norm(x_mean=np.zeros(7), x_std=np.ones(7), y_mean=np.zeros(4), y_std=np.ones(4),
                 box_lo=-np.ones(7), box_hi=-np.ones(7))
    m0 = st.summary()["total_dry_mass"]
    for _ in range(2):
        st.moist_step_surrogate(G.step_config(dt=1.0, fast_substeps=3),
                                G.warm_rain_config(dt=1.0, rain_terminal_velocity=12.0), sur)
    s = st.summary()
    assert np.isfinite(s["w_max"])
    assert abs(s["total_dry_mass"] - m0) <= 1e-6 * m0


def test_quiescent_cells_take_the_plain_answer(ctx):
    """Cells without cloud or rain at or below saturation have zero source
    terms in the plain scheme; the surrogate step gives them exactly the
    plain answer even when the network's output there is far from zero,
    and applies the network (inside the box) everywhere else."""
    wr = G.warm_rain_config(dt=5.0, rain_terminal_velocity=0.0)
    _, a = small_state(ctx, seed=12)
    _, b = small_state(ctx, seed=12)
    qc0, qr0 = a.get(17), a.get(18)
    sur = G.Surrogate(ctx, seed=4)
    big = 1e30
    sur.set_norm(x_mean=np.zeros(7), x_std=np.ones(7) * 1e3, y_mean=np.zeros(4),
                 y_std=np.array([1e-4, 1e-4, 1e-4, 1.0]), box_lo=-big * np.ones(7),
                 box_hi=big * np.ones(7))
    sur.set_params((SO.init_xavier(6) * 3.0 + 0.5).astype(np.float32))
    a.warm_rain(wr, metric_fallout=True)
    nfall = b.warm_rain_surrogate(wr, sur, metric_fallout=True)
    assert nfall == 0
    quiet = (qc0 == 0) & (qr0 == 0) & (a.get(17) == 0)
    assert 0.1 < quiet.mean() < 0.9, quiet.mean()
    for f in (O.RHO_THETA, 16, 17, 18):
        x, y = a.get(f), b.get(f)
        assert np.array_equal(x[quiet], y[quiet]), f
    # elsewhere the network's (large) changes were applied
    assert np.mean(a.get(17)[~quiet] != b.get(17)[~quiet]) > 0.5


def test_packed_network_cells_match_unpacked(ctx):
    """Only the cells that take the network's answer go through it, packed
    into 16-cell groups per tile.  A cell's result must not depend on which
    other cells are packed with it: quieting half the cells (no water at
    all, so every source term is zero) changes the packing of the rest but
    not one bit of their results; the counts report the cells that went
    through the network: those with cloud or rain, or above saturation (the
    plain scheme condenses there)."""
    wr = G.warm_rain_config(dt=5.0, rain_terminal_velocity=0.0)
    _, a = small_state(ctx, seed=14)
    _, b = small_state(ctx, seed=14)
    _, c = small_state(ctx, seed=14)
    qc0, qr0 = a.get(17), a.get(18)
    a.warm_rain(wr, metric_fallout=True)
    quiet = (qc0 == 0) & (qr0 == 0) & (a.get(17) == 0)
    sur = G.Surrogate(ctx, seed=5)
    big = 1e30
    sur.set_norm(x_mean=np.zeros(7), x_std=np.ones(7) * 1e3, y_mean=np.zeros(4),
                 y_std=np.array([1e-4, 1e-4, 1e-4, 1.0]), box_lo=-big * np.ones(7),
                 box_hi=big * np.ones(7))
    sur.set_params((SO.init_xavier(7) * 3.0 + 0.5).astype(np.float32))
    mask = np.random.default_rng(3).random(qc0.shape) < 0.5
    for f in (16, 17, 18):
        q = c.get(f)
        q[mask] = 0.0
        c.set(f, q)
    assert b.warm_rain_surrogate(wr, sur, metric_fallout=True) == 0
    fb_b, nw_b = sur.last_counts()
    assert c.warm_rain_surrogate(wr, sur, metric_fallout=True) == 0
    fb_c, nw_c = sur.last_counts()
    print("network cells", nw_b, nw_c, "of", mask.size, "quiet", int(quiet.sum()))
    assert fb_b == 0 and fb_c == 0
    assert nw_b == int(np.sum(~quiet))
    assert nw_c == int(np.sum(~quiet & ~mask))
    assert 0.2 * mask.size < nw_c < nw_b
    for f in (O.RHO_THETA, 16, 17, 18):
        x, y = b.get(f), c.get(f)
        assert np.array_equal(x[~mask], y[~mask]), f
    # the network's (large) changes were applied in the unmasked cells
    assert np.mean(a.get(17)[~mask & ~quiet] != c.get(17)[~mask & ~quiet]) > 0.5


@pytest.mark.parametrize("metric", [True, False])
def test_split_fallout_equals_whole_column_fallout(ctx, metric):
    """The classify pass carries the rain fallout down each column until the
    column's first network cell (all the way, into the bucket, in a column
    without one); a fallout pass finishes the other columns from that level.
    The result must be the fallout of whole columns applied after the source
    terms: here, the surrogate call without fallout followed by the plain
    warm rain with every source term switched off (its fallout alone).  Half
    the columns get 20 g/kg more rain, above the box's rain bound, so every
    cell there is a fallback cell and the column finishes in the classify
    pass; the others hold network cells at various heights."""
    base = dict(dt=5.0)
    if not metric:
        base["sedimentation_layer_depth_m"] = 500.0
    _, a = small_state(ctx, seed=15)
    _, b = small_state(ctx, seed=15)
    rho = a.get(O.RHO)
    cols = np.random.default_rng(4).random(rho.shape[1:]) < 0.5
    for st in (a, b):
        q = st.get(18)
        q[:, cols] += 0.02 * rho[:, cols]
        st.set(18, q)
    sur = G.Surrogate(ctx, seed=5)
    big = 1e30
    hi = big * np.ones(7)
    hi[5] = 0.01  # rain mixing ratio
    sur.set_norm(x_mean=np.zeros(7), x_std=np.ones(7) * 1e3, y_mean=np.zeros(4),
                 y_std=np.array([1e-4, 1e-4, 1e-4, 1.0]), box_lo=-big * np.ones(7), box_hi=hi)
    sur.set_params((SO.init_xavier(7) * 3.0 + 0.5).astype(np.float32))
    a.warm_rain_surrogate(G.warm_rain_config(rain_terminal_velocity=12.0, **base), sur,
                          metric_fallout=metric)
    fb, nw = sur.last_counts()
    b.warm_rain_surrogate(G.warm_rain_config(rain_terminal_velocity=0.0, **base), sur,
                          metric_fallout=metric)
    assert sur.last_counts() == (fb, nw)
    assert fb == rho.shape[0] * int(cols.sum()) and nw > 0
    for f in (O.RHO_THETA, 16, 17):
        assert np.array_equal(a.get(f), b.get(f)), f
    no_sources = G.warm_rain_config(rain_terminal_velocity=12.0, condensation_relaxation=0.0,
                                    cloud_autoconversion_rate=0.0, rain_evaporation_rate=0.0,
                                    enable_latent_heating=False, **base)
    b.warm_rain(no_sources, metric_fallout=metric)
    qa, qb = a.get(18).astype(np.float64), b.get(18).astype(np.float64)
    # the plain call rounds rho * (rho q_r / rho) once more: an ulp or so
    assert np.allclose(qa, qb, rtol=1e-6, atol=1e-12 * np.max(qb))
    pa = a.get(O.PRECIP).astype(np.float64).reshape(cols.shape)
    pb = b.get(O.PRECIP).astype(np.float64).reshape(cols.shape)
    # rain reaches the ground in columns of both kinds
    assert np.all(pb[cols] > 0) and np.any(pb[~cols] > 0)
    assert np.allclose(pa, pb, rtol=1e-6, atol=1e-12 * np.max(pb))
