"""Checks that pin the oracle's transport, stepping and terrain operators by
what the mathematics fixes (test infrastructure).

Each pin raises AssertionError when the oracle disagrees with a closed form
or an exact property; none restates the oracle's own formulas:

- tracer_ssp_shift: K23 (dry_transport.cu:49-125) with SSPRK3 (440-518).
  With uniform density and a uniform mass flux at Courant number 1, one
  upwind stage is the one-cell shift S, so the step must return exactly
  (1/3) q + (1/2) S q + (1/6) S^3 q; a downwind flux gives S^-1, other
  stage weights another polynomial.
- dry_ssp_shift: advance_dry_state_ssprk3 (dry_core.cu:541-651) with
  K17/K18 (dry_momentum_flux.cu:79-235).  Uniform rho and rho*theta, a
  uniform cross flow at Courant 1 and a momentum component that varies
  only along that flow: pressure, buoyancy and the fast modes vanish, the
  slow operator is the same shift, so the step is the same polynomial in S.
- theta_flux_shift: K2 (dry_core.cu:62-144).  An Euler step of the rho*theta
  tendency at Courant 1 shifts theta by one cell (x, y, interior z, both
  flow directions).
- w_momentum_shift: K19 (dry_momentum_flux.cu:237-312), x advection of m_w.
- u_self_advection_diffuses: K16/K17 x flux of m_u by its own cell
  velocity.  For a small bump on a uniform flow u0, the tendency's second
  central moment grows at +u0 dx (first-order upwind is diffusive); a
  downwind flux gives -u0 dx, a centred one 0.
- column_reference_second_order: K11 (dry_fast_modes.cu:108-158) against
  the exact constant-theta column Pi(z) = Pi_0 - g (z - z_0) / (c_p theta):
  the trapezoid recurrence is second order, so halving dz divides the
  error by about 4.
- terrain_pgf_cancels: K21/K22 (dry_pressure_gradient.cu:75-123) on a
  sloped hybrid-height grid with p linear in z: the along-level difference
  and the metric correction cancel, so the horizontal gradient is zero.
- taper_smoothstep: the terrain taper (grid_metrics.hpp:126-144) is 1 up to
  eta_t and 0 at the lid, equals 1/2 halfway, and leaves both ends with
  zero slope (1 - w and w grow quadratically away from the ends).
"""
import numpy as np

RHO, RHO_THETA, MOM_U, MOM_V, MOM_W = 0, 1, 2, 3, 4
TEND_RHO, TEND_RHO_THETA, TEND_U, TEND_V, TEND_W = 6, 7, 8, 9, 10
TRACER0 = 16


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _shift(q, axis, sign):
    """S q: the value one cell upstream (sign +1: flow towards +axis)."""
    return np.roll(q, sign, axis=axis)


def _ssp_poly(q, axis, sign):
    s1 = _shift(q, axis, sign)
    s3 = _shift(_shift(s1, axis, sign), axis, sign)
    return q / 3.0 + s1 / 2.0 + s3 / 6.0


def tracer_ssp_shift(O):
    nx, ny, nz = 16, 10, 3
    dt = 10.0
    rng = np.random.default_rng(11)
    for axis, field, sign in ((2, MOM_U, +1), (2, MOM_U, -1), (1, MOM_V, +1), (1, MOM_V, -1)):
        w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_top=3000.0, n_tracers=1, precision=8)
        w.init_constant(1.0, 300.0)
        w.fill(field, sign * 100.0)  # rho u dt / dx = 1
        q = 1.0 + rng.random((nz, ny, nx))
        w.set(TRACER0, q)
        w.tracers_ssprk3(O.step_cfg(dt=dt))
        e = _rel(w.get(TRACER0), _ssp_poly(q, axis, sign))
        assert e <= 1e-13, ("tracer SSPRK3 at Courant 1", axis, sign, e)


def dry_ssp_shift(O):
    nx, ny, nz = 16, 12, 4
    dt = 10.0
    x = np.arange(nx)
    y = np.arange(ny)
    for case, sign in (("v along x", +1), ("v along x", -1), ("u along y", +1),
                       ("u along y", -1)):
        w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_top=4000.0, precision=8)
        w.init_constant(1.0, 300.0)
        if case == "v along x":
            w.fill(MOM_U, sign * 100.0)
            prof = 3.0 + np.sin(2 * np.pi * x / nx) + 0.5 * np.cos(6 * np.pi * x / nx)
            m = np.broadcast_to(prof[None, None, :], (nz, ny + 1, nx)).copy()
            field, axis = MOM_V, 2
        else:
            w.fill(MOM_V, sign * 100.0)
            prof = -2.0 + np.cos(2 * np.pi * y / ny) + 0.3 * np.sin(4 * np.pi * y / ny)
            m = np.broadcast_to(prof[None, :, None], (nz, ny, nx + 1)).copy()
            field, axis = MOM_U, 1
        w.set(field, m)
        w.ssprk3_step(O.step_cfg(dt=dt, fast_substeps=3))
        got = w.get(field)
        want = _ssp_poly(prof, 0, sign)
        want = (np.broadcast_to(want[None, None, :], m.shape) if axis == 2
                else np.broadcast_to(want[None, :, None], m.shape))
        e = _rel(got, want)
        assert e <= 1e-12, ("dry SSPRK3 at Courant 1", case, sign, e)
        # everything else stays put
        assert _rel(w.get(RHO), np.ones((nz, ny, nx))) <= 1e-13
        assert np.max(np.abs(w.get(MOM_W))) <= 1e-9


def theta_flux_shift(O):
    nx, ny, nz = 12, 10, 8
    dt = 10.0
    rng = np.random.default_rng(5)
    for field, axis, sign in ((MOM_U, 2, +1), (MOM_U, 2, -1), (MOM_V, 1, +1), (MOM_V, 1, -1),
                              (MOM_W, 0, +1), (MOM_W, 0, -1)):
        w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_top=8000.0, precision=8)
        w.init_constant(1.0, 300.0)
        rt = 300.0 + 5.0 * rng.random((nz, ny, nx))
        w.set(RHO_THETA, rt)
        if field == MOM_W:
            m = np.full((nz + 1, ny, nx), sign * 100.0)  # dz = 1000 m
            m[0] = m[-1] = 0.0
            w.set(MOM_W, m)
        else:
            w.fill(field, sign * 100.0)
        w.slow_tendencies(O.step_cfg(dt=dt))
        new = rt + dt * w.get(TEND_RHO_THETA)
        want = _shift(rt, axis, sign)
        if axis == 0:  # interior levels only: the lid and ground carry no flux
            new, want = new[1:-1], want[1:-1]
        e = _rel(new, want)
        assert e <= 1e-13, ("theta flux at Courant 1", field, sign, e)


def w_momentum_shift(O):
    nx, ny, nz = 16, 6, 7
    dt = 10.0
    x = np.arange(nx)
    for sign in (+1, -1):
        w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_top=7000.0, precision=8)
        w.init_constant(1.0, 300.0)
        w.fill(MOM_U, sign * 100.0)
        prof = 0.5 + np.sin(2 * np.pi * x / nx)
        m = np.broadcast_to(prof[None, None, :], (nz + 1, ny, nx)).copy()
        m[0] = m[-1] = 0.0
        w.set(MOM_W, m)
        for f in (RHO, MOM_U, MOM_V, MOM_W):
            w.exchange(f)
        for f in (TEND_U, TEND_V, TEND_W):
            w.fill(f, 0.0)
        w.momentum_flux_only()
        new = m + dt * w.get(TEND_W)
        want = np.broadcast_to(_shift(prof, 0, sign)[None, None, :], m.shape)
        # faces 2 .. nz-2: their vertical fluxes cancel (m_w uniform in z)
        e = _rel(new[2:-2], want[2:-2])
        assert e <= 1e-13, ("w momentum at Courant 1", sign, e)


def u_self_advection_diffuses(O):
    nx, ny, nz = 64, 4, 3
    dx = 1000.0
    for sign in (+1, -1):
        u0 = sign * 20.0
        w = O.World(nx, ny, nz, dx=dx, dy=dx, z_top=3000.0, precision=8)
        w.init_constant(1.0, 300.0)
        fx = np.arange(nx + 1)
        eps = 1e-5 * abs(u0)
        bump = eps * np.exp(-0.5 * ((fx - nx / 2) / 2.0) ** 2)
        m = np.broadcast_to((u0 + bump)[None, None, :], (nz, ny, nx + 1)).copy()
        m[..., -1] = m[..., 0]  # the periodic wrap face equals face 0
        w.set(MOM_U, m)
        for f in (RHO, MOM_U, MOM_V, MOM_W):
            w.exchange(f)
        for f in (TEND_U, TEND_V, TEND_W):
            w.fill(f, 0.0)
        w.momentum_flux_only()
        t = w.get(TEND_U)[0, 0, :nx]  # owned faces 0 .. nx-1 (face nx == face 0)
        d = bump[:nx]
        xs = fx[:nx] * dx
        mass = d.sum()
        assert abs(t.sum()) <= 1e-9 * np.abs(t).max()  # conservative
        xbar = (xs * d).sum() / mass
        # d/dt of the bump's variance <x^2> - xbar^2, with d/dt xbar the
        # drift of its first moment
        drift = (xs * t).sum() / mass
        rate = ((xs ** 2) * t).sum() / mass - 2.0 * xbar * drift
        assert abs(rate / (abs(u0) * dx) - 1.0) <= 1e-2, ("m_u self advection", sign, rate)


def column_reference_second_order(O):
    g, th, cp = 9.81, 300.0, 287.05 / 0.2854
    p0, Rd, kap = 1.0e5, 287.05, 0.2854
    ztop = 12000.0

    def exact_p(z, z0, ps):
        pi0 = (ps / p0) ** kap
        return p0 * (pi0 - g * (z - z0) / (cp * th)) ** (1.0 / kap)

    errs = []
    for nz in (30, 60, 120):
        dz = ztop / nz
        zc = (np.arange(nz) + 0.5) * dz
        ps = 98000.0
        p, r = O.column_reference(np.full(nz, th), zc, ps, gravity=g, precision=8)
        pe = exact_p(zc, zc[0], ps)
        re = pe / (Rd * th * (pe / p0) ** kap)
        assert abs(p[0] - ps) == 0.0
        assert np.max(np.abs(r - re) / re) <= 1e-3
        errs.append(np.max(np.abs(p - pe) / pe))
    ratio1, ratio2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.5 <= ratio1 <= 4.5 and 3.5 <= ratio2 <= 4.5, ("K11 convergence", errs)
    assert errs[0] <= 1e-3, errs  # dz 400 m


def terrain_pgf_cancels(O):
    nx, ny, nz = 24, 20, 10
    j, i = np.mgrid[0:ny, 0:nx]
    hill = 900.0 * (0.5 * (1 + np.cos(np.pi * np.clip(np.hypot((i - 12) / 7.0, (j - 9) / 6.0),
                                                        0, 1))))
    w = O.World(nx, ny, nz, dx=2000.0, dy=2000.0, z_top=12000.0, terrain=hill, precision=8)
    zc = w.metric(1)
    a, b = 1.0e5, 5.0  # p = a - b z, positive up to the lid
    p = a - b * zc
    rt = (1.0e5 / 287.05) * (p / 1.0e5) ** (1.0 - 0.2854)  # EOS inverse at any rho
    w.init_constant(1.0, 300.0)
    w.set(RHO_THETA, rt)
    w.exchange(RHO)
    w.exchange(RHO_THETA)
    w.pressure()
    w.exchange(5)
    for f in (TEND_U, TEND_V, TEND_W):
        w.fill(f, 0.0)
    w.pressure_gradient_only()
    slope = max(np.max(np.abs(np.diff(zc, axis=2))) / 2000.0,
                np.max(np.abs(np.diff(zc, axis=1))) / 2000.0)
    scale = b * slope
    assert scale > 1e-3
    for f in (TEND_U, TEND_V):
        e = np.max(np.abs(w.get(f))) / scale
        assert e <= 1e-8, ("terrain PGF with p(z) linear", f, e)


def taper_smoothstep(O):
    nx, ny = 4, 3
    terr = np.full((ny, nx), 600.0)
    for nz, checks in ((8, "mid"), (400, "ends")):
        w = O.World(nx, ny, nz, dx=1000.0, dy=1000.0, z_top=10000.0, terrain=terr,
                    taper_eta=0.25, precision=8)
        zi = w.metric(0)[:, 0, 0]
        eta = np.arange(nz + 1) / nz
        wt = (zi - eta * 10000.0) / 600.0
        if checks == "mid":
            assert abs(wt[2] - 1.0) <= 1e-12  # eta = 0.25
            assert abs(wt[5] - 0.5) <= 1e-12  # eta = 0.625, halfway through the taper
            assert abs(wt[8]) <= 1e-12
        else:
            k0 = 100  # eta = 0.25
            assert np.all(np.abs(wt[:k0 + 1] - 1.0) <= 1e-12)
            r_lo = (1.0 - wt[k0 + 2]) / (1.0 - wt[k0 + 1])
            r_hi = wt[nz - 2] / wt[nz - 1]
            assert 3.9 <= r_lo <= 4.1, ("taper slope at eta_t", r_lo)
            assert 3.9 <= r_hi <= 4.1, ("taper slope at the lid", r_hi)


PINS = {
    "tracer_ssp_shift": tracer_ssp_shift,
    "dry_ssp_shift": dry_ssp_shift,
    "theta_flux_shift": theta_flux_shift,
    "w_momentum_shift": w_momentum_shift,
    "u_self_advection_diffuses": u_self_advection_diffuses,
    "column_reference_second_order": column_reference_second_order,
    "terrain_pgf_cancels": terrain_pgf_cancels,
    "taper_smoothstep": taper_smoothstep,
}
