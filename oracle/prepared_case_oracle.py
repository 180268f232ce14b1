"""CPU oracle of the prepared-case start-up -- TEST INFRASTRUCTURE ONLY.

Imported by tests/ as the checker of the host mirror's prepared-case path
(gpu-wm-next_amd/host/include/gwm/gwm_ingest.hpp) and gwm_prepared_case_driver;
the product never calls it.  It restates, in FP32 numpy like the
reference's `real`, src/ingest/prepared_case_init.cpp of
FahrenheitResearch/gpu-wm-next:

- rho_theta_m_from_pressure (prepared_case_init.cpp:38-50),
- build_balanced_prepared_case_columns (320-384): hydrostatic columns from
  the analysis pressure at the ground, trapezoid recurrence with four
  fixed-point passes,
- build_projected_prepared_case_face_momenta (172-318): face momenta from the
  balanced density and the analysis winds, barotropic projection on fully
  open domains (successive over-relaxation, 126-170), w from the column
  continuity equation,
- make_tracers_from_analysis (923-1007) and the boundary snapshot
  interpolation (runtime_case.cpp:214-240).

Metric heights and layer thicknesses come from the C++ oracle's metrics
(pyoracle.World.metric), which are pinned to the reference's
GridMetrics::make_hybrid_height.

Pinned by (tests/test_prepared_case.py): closed forms of the balanced state
(the EOS at rho = 1 reproduces the pressure the rho*theta was built from;
each level satisfies the trapezoid hydrostatic relation with the level
below; rho*theta / rho is the analysis theta), the projected momenta's
vanishing column-mean divergence and discrete mass divergence, w = 0 at the
ground and lid, and brute-force agreement of the vectorised projection sweep
with a scalar restatement on a tiny grid.
"""
import numpy as np

F = np.float32
P0, RD, KAPPA, G = F(100000.0), F(287.05), F(0.2854), F(9.81)


def _pow(a, b):
    return np.power(F(a) if np.isscalar(a) else a.astype(F), F(b)).astype(F)


def pressure_from_rho_theta(rho, rt):
    """dry_thermo.hpp:26-33 in FP32."""
    rho = np.asarray(rho, F)
    th = (np.asarray(rt, F) / np.maximum(rho, F(1e-6))).astype(F)
    ratio = np.maximum((rho * RD * th / P0).astype(F), F(1e-6))
    return (P0 * _pow(ratio, F(1.0) / (F(1.0) - KAPPA))).astype(F)


def rho_from_pressure_theta(p, th):
    """dry_thermo.hpp:35-40 in FP32."""
    p = np.asarray(p, F)
    ex = _pow(np.maximum((p / P0).astype(F), F(1e-6)), KAPPA)
    return (p / np.maximum((RD * np.asarray(th, F) * ex).astype(F), F(1e-6))).astype(F)


def rho_theta_from_pressure(p):
    p = np.asarray(p, F)
    rt = ((P0 / RD) * _pow(np.maximum((p / P0).astype(F), F(1e-6)), F(1.0) - KAPPA)).astype(F)
    for _ in range(2):
        pe = pressure_from_rho_theta(np.ones_like(rt), rt)
        ratio = np.maximum((p / np.maximum(pe, F(1.0))).astype(F), F(1e-6))
        rt = (rt * _pow(ratio, F(1.0) - KAPPA)).astype(F)
    return rt


def balanced_columns(p_src, t_src, zc):
    """(pressure, theta, rho_d, rho_theta_m), arrays (nz, ny, nx);
    zc: metric level heights (nz, ny, nx)."""
    p_src = np.asarray(p_src, F)
    t_src = np.asarray(t_src, F)
    nz = p_src.shape[0]
    ps = np.maximum(p_src, F(1.0))
    theta = (t_src * _pow((P0 / ps).astype(F), KAPPA)).astype(F)
    p = np.zeros_like(ps)
    rho = np.zeros_like(ps)
    p[0] = ps[0]
    rho[0] = rho_from_pressure_theta(p[0], theta[0])
    for k in range(1, nz):
        dz = (np.asarray(zc[k], F) - np.asarray(zc[k - 1], F)).astype(F)
        pn = np.maximum(F(1.0), (p[k - 1] - (G * rho[k - 1] * dz).astype(F)).astype(F))
        for _ in range(4):
            rn = rho_from_pressure_theta(pn, theta[k])
            pn = (p[k - 1] - (G * F(0.5) * (rho[k - 1] + rn).astype(F) * dz).astype(F)).astype(F)
            pn = np.maximum(pn, F(1.0))
        p[k] = pn
        rho[k] = rho_from_pressure_theta(pn, theta[k])
    return p, theta, rho, rho_theta_from_pressure(p)


def _wrap_or_clamp(i, n, periodic):
    return i % n if periodic else min(max(i, 0), n - 1)


def column_mean_divergence(mu, mv, idz, dx, dy):
    hdiv = ((mu[:, :, 1:] - mu[:, :, :-1]) / F(dx) + (mv[:, 1:, :] - mv[:, :-1, :]) / F(dy)).astype(F)
    dz = (F(1.0) / np.maximum(np.asarray(idz, F), F(1e-6))).astype(F)
    bias = np.zeros(hdiv.shape[1:], F)
    height = np.zeros(hdiv.shape[1:], F)
    for k in range(hdiv.shape[0]):  # sequential FP32 sums, level order
        bias = (bias + (hdiv[k] * dz[k]).astype(F)).astype(F)
        height = (height + dz[k]).astype(F)
    return (bias / np.maximum(height, F(1e-6))).astype(F), hdiv, dz


def sor_potential(rhs, dx, dy):
    """lap(phi) = rhs with phi = 0 outside, Gauss-Seidel SOR in row order
    (omega 1.85, <= 2000 sweeps, stop at max update <= 1e-10)."""
    ny, nx = rhs.shape
    phi = np.zeros((ny, nx), F)
    ix2 = F(1.0) / max(F(dx) * F(dx), F(1e-6))
    iy2 = F(1.0) / max(F(dy) * F(dy), F(1e-6))
    den = F(2.0) * (ix2 + iy2)
    for _ in range(2000):
        big = F(0.0)
        for j in range(ny):
            for i in range(nx):
                w = phi[j, i - 1] if i > 0 else F(0.0)
                e = phi[j, i + 1] if i + 1 < nx else F(0.0)
                s = phi[j - 1, i] if j > 0 else F(0.0)
                n = phi[j + 1, i] if j + 1 < ny else F(0.0)
                star = F((F((w + e) * ix2) + F((s + n) * iy2) - rhs[j, i]) / den)
                upd = F(phi[j, i] + F(F(1.85) * F(star - phi[j, i])))
                big = max(big, F(abs(F(upd - phi[j, i]))))
                phi[j, i] = upd
        if big <= F(1e-10):
            break
    return phi


def face_momenta(rho, u_wind, v_wind, idz, dx, dy, periodic_x, periodic_y):
    """(mom_u (nz, ny, nx+1), mom_v (nz, ny+1, nx), mom_w (nz+1, ny, nx))."""
    rho = np.asarray(rho, F)
    uw = np.asarray(u_wind, F)
    vw = np.asarray(v_wind, F)
    nz, ny, nx = rho.shape
    il = [_wrap_or_clamp(fi - 1, nx, periodic_x) for fi in range(nx + 1)]
    ir = [_wrap_or_clamp(fi, nx, periodic_x) for fi in range(nx + 1)]
    mu = ((F(0.5) * (rho[:, :, il] + rho[:, :, ir])).astype(F) *
          (F(0.5) * (uw[:, :, il] + uw[:, :, ir])).astype(F)).astype(F)
    js = [_wrap_or_clamp(fj - 1, ny, periodic_y) for fj in range(ny + 1)]
    jn = [_wrap_or_clamp(fj, ny, periodic_y) for fj in range(ny + 1)]
    mv = ((F(0.5) * (rho[:, js, :] + rho[:, jn, :])).astype(F) *
          (F(0.5) * (vw[:, js, :] + vw[:, jn, :])).astype(F)).astype(F)
    if not periodic_x and not periodic_y:
        for _ in range(3):
            mean, _, _ = column_mean_divergence(mu, mv, idz, dx, dy)
            if np.max(np.abs(mean)) <= F(1e-8):
                break
            phi = sor_potential(mean, dx, dy)
            pad_x = np.zeros((ny, nx + 2), F)
            pad_x[:, 1:-1] = phi
            corr_u = ((pad_x[:, 1:] - pad_x[:, :-1]) / F(dx)).astype(F)  # (ny, nx+1)
            mu = (mu - corr_u[None]).astype(F)
            pad_y = np.zeros((ny + 2, nx), F)
            pad_y[1:-1] = phi
            corr_v = ((pad_y[1:] - pad_y[:-1]) / F(dy)).astype(F)  # (ny+1, nx)
            mv = (mv - corr_v[None]).astype(F)
    mean, hdiv, dz = column_mean_divergence(mu, mv, idz, dx, dy)
    mw = np.zeros((nz + 1, ny, nx), F)
    for k in range(nz):
        mw[k + 1] = (mw[k] - ((hdiv[k] - mean).astype(F) * dz[k]).astype(F)).astype(F)
    mw[nz] = 0.0
    return mu, mv, mw


def prepared_state(fields, zc, idz, dx, dy, periodic_x=False, periodic_y=False):
    """Balanced dry state and warm-rain tracer masses of an analysis.

    fields: dict of (nz, ny, nx) arrays air_pressure, air_temperature,
    u_wind, v_wind, specific_humidity (optionally cloud / rain mixing ratios).
    Returns dict rho, rho_theta, mom_u, mom_v, mom_w, qv, qc, qr, pressure."""
    p, th, rho, rt = balanced_columns(fields["air_pressure"], fields["air_temperature"], zc)
    mu, mv, mw = face_momenta(rho, fields["u_wind"], fields["v_wind"], idz, dx, dy,
                              periodic_x, periodic_y)
    out = dict(rho=rho, rho_theta=rt, mom_u=mu, mom_v=mv, mom_w=mw, pressure=p, theta=th)
    for key, name in (("qv", "specific_humidity"), ("qc", "cloud_water_mixing_ratio"),
                      ("qr", "rain_water_mixing_ratio")):
        q = fields.get(name)
        out[key] = (rho * np.asarray(q, F)).astype(F) if q is not None else np.zeros_like(rho)
    return out


def interpolate_snapshot(snapshots, t):
    """Fields of the boundary cache at forecast offset t (seconds): the
    bracketing snapshots blended linearly, clamped to the first / last.
    snapshots: list of (offset, {name: array})."""
    lo = hi = None
    for off, f in snapshots:
        if off <= t:
            lo = (off, f)
        if off >= t:
            hi = (off, f)
            break
    lo = lo or snapshots[0]
    hi = hi or snapshots[-1]
    if lo[0] == hi[0]:
        return lo[1]
    a = F(t - lo[0]) / F(hi[0] - lo[0])
    return {k: ((F(1.0) - a) * np.asarray(lo[1][k], F) + a * np.asarray(hi[1][k], F)).astype(F)
            for k in lo[1]}
