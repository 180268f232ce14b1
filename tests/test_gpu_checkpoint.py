"""Restart dump (gwm_state_checkpoint_write / _read) on the GPU: a run split
by a checkpoint and a restore into a fresh state equals the unbroken run bit
for bit (full FP32 with strips and tracers, FP64, the perturbation form);
the manifest satisfies the reference's checkpoint-stub contract
(tools/verify/run_verification.py:729-755 checks schema, restart family and
cold-start groups); corrupted and mismatched dumps are refused with errors
that name the problem."""
import json

import numpy as np
import pytest

import gwm_amd as G

pytestmark = pytest.mark.gpu

FIELDS = (G.RHO, G.RHO_THETA, G.MOM_U, G.MOM_V, G.MOM_W)


def moist(ctx, nx=24, ny=20, nz=12, perturbation=False, periodic=True, seed=5):
    zi = np.linspace(0.0, 12000.0, nz + 1)
    grid = G.Grid(ctx, nx, ny, nz, dx=1000.0, dy=1000.0, z_interfaces=zi, periodic_x=periodic,
                  periodic_y=periodic)
    st = G.State(grid, 3, perturbation=perturbation)
    st.init_convection(1.0e5, seed, 0.5, 1000.0)
    qv = st.get(G.TRACER0)
    st.set(G.TRACER0 + 1, (0.02 * qv).astype(qv.dtype))
    st.set(G.TRACER0 + 2, (0.01 * qv).astype(qv.dtype))
    return grid, st


def run(st, n, bc=None):
    cfg = G.step_config(dt=1.0, fast_substeps=3)
    wr = G.warm_rain_config(dt=1.0, rain_terminal_velocity=12.0)
    for _ in range(n):
        st.moist_step(cfg, wr, bc)


def snapshot(st):
    out = [st.get(f) for f in FIELDS + (G.TRACER0, G.TRACER0 + 1, G.TRACER0 + 2, G.PRECIP)]
    if st.perturbation:
        out += [st.get_f64(G.RHO), st.get_f64(G.RHO_THETA)]
    return out


@pytest.mark.parametrize("prec,perturbation,periodic", [(G.F32, False, False), (G.F64, False, True),
                                                        (G.F32, True, True)])
def test_restart_is_bit_exact(tmp_path, prec, perturbation, periodic):
    ctx = G.Context(0, prec)
    grid, a = moist(ctx, perturbation=perturbation, periodic=periodic)
    _, b = moist(ctx, perturbation=perturbation, periodic=periodic)
    bc = None
    if not periodic:
        for s in (a, b):
            s.capture_boundary()
        bc = G.BoundaryConfig(G.BOUNDARY_STRIPS, 0.0, G.BOUNDARY_FN(), None)
    run(a, 6, bc)
    run(b, 3, bc)
    path = str(tmp_path / "state.ckpt")
    b.checkpoint_write(path, metadata={"step": 3, "note": "split run"})
    # a fresh state on the same grid, seeded differently, restored from the dump
    c = G.State(grid, 3, perturbation=perturbation)
    c.init_convection(1.0e5, 99, 0.5, 1000.0)
    c.checkpoint_read(path)
    for x, y in zip(snapshot(b), snapshot(c)):
        assert np.array_equal(x, y)
    run(c, 3, bc)
    assert c.summary()["nonfinite_values"] == 0
    for x, y in zip(snapshot(a), snapshot(c)):
        assert np.array_equal(x, y)
    man = json.load(open(path + ".json"))
    assert man["schema_version"] == "gwm-next-checkpoint-stub/v1"
    assert man["restart_family"] == "gwm-native-checkpoint"
    assert man["cold_start_contract"]["required_groups"] == ["atmosphere", "surface",
                                                             "static_surface"]
    assert man["checkpoint"]["grid"]["nx"] == 24 and man["checkpoint"]["n_tracers"] == 3
    assert man["checkpoint"]["boundary_reference"] == (not periodic)
    assert man["metadata"] == {"step": 3, "note": "split run"}
    del a, b, c, grid
    ctx.close()


def test_corrupt_and_mismatched_dumps_are_refused(tmp_path):
    ctx = G.Context(0, G.F32)
    grid, a = moist(ctx)
    path = str(tmp_path / "a.ckpt")
    a.checkpoint_write(path)
    raw = bytearray(open(path, "rb").read())
    raw[-5] ^= 0xFF
    bad = str(tmp_path / "bad.ckpt")
    open(bad, "wb").write(raw)
    with pytest.raises(G.GwmError, match="fails its checksum"):
        a.checkpoint_read(bad)
    open(bad, "wb").write(raw[:1000])
    with pytest.raises(G.GwmError, match="truncated|does not match"):
        a.checkpoint_read(bad)
    _, other = moist(ctx, nx=20)
    with pytest.raises(G.GwmError, match="grid shape"):
        other.checkpoint_read(path)
    ctx64 = G.Context(0, G.F64)
    _, s64 = moist(ctx64)
    with pytest.raises(G.GwmError, match="precision"):
        s64.checkpoint_read(path)
    _, p = moist(ctx, perturbation=True)
    with pytest.raises(G.GwmError, match="storage form"):
        p.checkpoint_read(path)
    open(bad, "wb").write(b"not a checkpoint" * 20)
    with pytest.raises(G.GwmError, match="not a gwm checkpoint"):
        a.checkpoint_read(bad)
    del a, other, s64, p, grid
    ctx64.close()
    ctx.close()
