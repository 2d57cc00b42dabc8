// viscalc.hpp -- the ductile viscosity of a marker, shared by the marker
// kernels and the flow adjoint.
#pragma once
#include "kernels.hpp"

namespace i2e {

// ---------------------------------------------------------------- rheology
// viscalc, markmart.c:1855-2063 (ductile branches A-D and the limits; the
// brittle and Peierls branches need rock constants the configs set to zero)
__device__ inline double viscalc_d(const i2e_params *p, double mtk, double mpb, int mm2, double epsin) {
  const i2e_rock &r = p->rock[mm2];
  double rt = 8.314 * mtk, e1, k1, sduct, nunewt = 0, nupowl = 0, nueff;
  if (mm2 > 20) return r.n0;
  if (r.nu0 != 0) {
    if (r.dh == 0 && r.dv == 0 && (r.ss == 0 || r.mm == 1.0)) nunewt = 1.0 / r.nu0;
    if ((r.dh != 0 || r.dv != 0) && (r.ss == 0 || r.mm == 1.0)) {
      e1 = (r.dh + r.dv * mpb) / rt;
      if (e1 > 150.0) e1 = 150.0;
      if (r.dh < 0 && r.dv > 0) {
        e1 = r.dh * (mtk - r.dv);
        if (e1 < -150.0) e1 = -150.0;
      }
      if (r.dh < 0 && r.dv < 0) {
        e1 = (-r.dh) * (1.0 - (mtk - (-r.dv)) / (-r.dv)) / 8.314 / (-r.dv);
        if (e1 > 150.0) e1 = 150.0;
      }
      nunewt = 1.0 / (r.nu0 * exp(e1));
    }
    if ((r.dh == 0 && r.dv == 0) && r.ss != 0 && r.mm != 1.0) {
      k1 = r.nu0 / pow(r.ss, r.mm - 1.0);
      nunewt = 1.0 / (0.5 * k1);
      sduct = pow(r.nu0 * epsin, 1.0 / r.mm);
      if (epsin > 0) nupowl = 1.0 / (0.5 * sduct / epsin);
    }
    if ((r.dh != 0 || r.dv != 0) && r.ss != 0 && r.mm != 1.0) {
      e1 = (r.dh + r.dv * mpb) / rt;
      if (e1 > 150.0) e1 = 150.0;
      e1 = exp(e1);
      k1 = r.nu0 / pow(r.ss, r.mm - 1.0);
      nunewt = 1.0 / (0.5 * k1 * e1);
      sduct = pow(r.nu0 * e1 * epsin, 1.0 / r.mm);
      if (epsin > 0) nupowl = 1.0 / (0.5 * sduct / epsin);
    }
  }
  nueff = nunewt + nupowl;
  if (nueff <= 0) nueff = 1.0 / r.n1;
  nueff = 1.0 / nueff;
  if (nueff < r.n0) nueff = r.n0;
  if (nueff > r.n1) nueff = r.n1;
  if (nueff < p->nubeg) nueff = p->nubeg;
  if (nueff > p->nuend) nueff = p->nuend;
  return nueff;
}

}  // namespace i2e
