// gwm_micro.hpp -- the warm-rain source terms of one cell (K24,
// src/physics/warm_rain.cu:36-120), shared by the plain warm-rain kernel, the
// surrogate's training-data sampler and its fallback path.
//
// micro_prepare derives the thermodynamic state of a cell from the conserved
// variables; micro_apply runs saturation adjustment (condensation or cloud
// evaporation, with latent heating through the pre-adjustment Exner
// function, warm_rain.cu:57-58, 71-72), autoconversion, the accretion
// extension (rate 0 reproduces the reference), then rain evaporation
// against the re-derived saturation ratio, and clips the mixing ratios at 0.
// Arithmetic is the step kernels' (q* in gwm_internal.hpp).
#pragma once

#include "gwm_internal.hpp"

namespace gwm {

template <class R>
struct MicroCfg {
  R dt, relax, thr, krate, erate, kacc;
  int heat;
};

template <class R>
struct MicroCell {
  R rho;   // max(rho_d, 1e-6)
  R th;    // theta
  R p, T;  // pressure and temperature before the adjustment
  R qv, qc, qr;
  R lfac;  // L / (c_p * Exner)
};

template <class R>
__device__ inline MicroCell<R> micro_prepare(R rho_raw, R rt0, R rqv, R rqc, R rqr) {
  MicroCell<R> m;
  m.rho = rmax(rho_raw, R(1.0e-6));
  m.th = qdiv(rt0, m.rho);
  m.p = q_pressure(m.rho, rt0);
  m.T = q_theta(m.rho, rt0) * qpow(rmax(qdiv(m.p, Phys<R>::p0), R(1.0e-6)), Phys<R>::kappa);
  const R ex = rmax(qdiv(m.T, rmax(m.th, R(1.0e-6))), R(1.0e-6));
  m.qv = rmax(qdiv(rqv, m.rho), R(0));
  m.qc = rmax(qdiv(rqc, m.rho), R(0));
  m.qr = rmax(qdiv(rqr, m.rho), R(0));
  m.lfac = qdiv(Phys<R>::Lv, d_cp<R>() * rmax(ex, R(1.0e-6)));
  return m;
}

template <class R>
__device__ inline void micro_apply(MicroCell<R>& m, const MicroCfg<R>& c) {
  const R rho = m.rho;
  R th = m.th, qv = m.qv, qc = m.qc, qr = m.qr;
  const R lfac = m.lfac;
  R qs = q_qsat(m.T, m.p);
  if (qv > qs) {
    const R cnd = rmin((qv - qs) * rmax(c.relax, R(0)), qv);
    qv -= cnd;
    qc += cnd;
    if (c.heat && cnd > R(0)) th += lfac * cnd;
  } else if (qv < qs && qc > R(0)) {
    const R e = rmin((qs - qv) * rmax(c.relax, R(0)), qc);
    qv += e;
    qc -= e;
    if (c.heat && e > R(0)) th -= lfac * e;
  }
  const R exc = rmax(qc - c.thr, R(0));
  if (exc > R(0) && c.krate > R(0)) {
    const R cv = rmin(exc, c.dt * c.krate * exc);
    qc -= cv;
    qr += cv;
  }
  if (c.kacc > R(0) && qc > R(0) && qr > R(0)) {  // accretion extension
    const R ac = rmin(qc, c.dt * c.kacc * qc * qr);
    qc -= ac;
    qr += ac;
  }
  const R p = q_pressure(rho, rho * th);
  const R T = th * qpow(rmax(qdiv(p, Phys<R>::p0), R(1.0e-6)), Phys<R>::kappa);
  qs = q_qsat(T, p);
  if (qr > R(0) && qv < qs && c.erate > R(0)) {
    const R e = rmin(qr, c.dt * c.erate * rmin(qs - qv, qr));
    qr -= e;
    qv += e;
    if (c.heat && e > R(0)) th -= lfac * e;
  }
  m.th = th;
  m.qv = rmax(qv, R(0));
  m.qc = rmax(qc, R(0));
  m.qr = rmax(qr, R(0));
}

}  // namespace gwm
