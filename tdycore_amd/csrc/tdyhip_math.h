// Constitutive relations of the Richards path (FP64), device and host.
#pragma once

#include <hip/hip_runtime.h>

namespace tdyhip {

// van Genuchten saturation and its first two pressure derivatives
// (PressureSaturation_VanGenuchten, src/materials/tdycharacteristiccurves.c:692-717).
__host__ __device__ inline void vg_saturation(double m, double alpha, double Sr, double Pc, double &S,
                                              double &dS_dP, double &d2S_dP2) {
  if (Pc <= 0.0) {
    S = 1.0; dS_dP = 0.0; d2S_dP2 = 0.0;
    return;
  }
  const double n = 1.0 / (1.0 - m);
  const double pa = Pc * alpha;
  const double pan = pow(pa, n);
  const double opn = 1.0 + pan;
  const double Se = pow(opn, -m);
  S = (1.0 - Sr) * Se + Sr;
  const double dSe_dPc = -m * n * alpha * pan / (pa * pow(opn, m + 1.0));
  dS_dP = -dSe_dPc * (1.0 - Sr);
  const double d2Se = m * n * pan * pow(opn, -m - 2.0) * ((m * n + 1.0) * pan - n + 1.0) / (Pc * Pc);
  d2S_dP2 = (1.0 - Sr) * d2Se;
}

// Mualem relative permeability; a cubic (coefficients c[], variable
// (Se-x0)/dx) above Se = x1 (RelativePermeability_Mualem and
// _Unsmoothed, tdycharacteristiccurves.c:468-485, 551-562, 607-614).
__host__ __device__ inline void mualem_kr(double m, double x0, double x1, double dx, const double *c,
                                          double Se, double &Kr, double &dKr_dSe) {
  if (Se > x1) {
    const double xt = (Se - x0) / dx;
    Kr = c[0] + xt * (c[1] + xt * (c[2] + xt * c[3]));
    dKr_dSe = (c[1] + xt * (2.0 * c[2] + xt * 3.0 * c[3])) / dx;
    return;
  }
  Kr = 1.0; dKr_dSe = 0.0;
  if (Se >= 1.0) return;
  const double s1m = pow(Se, 1.0 / m);
  const double tmp = pow(1.0 - s1m, m);
  Kr = sqrt(Se) * ((1.0 - tmp) * (1.0 - tmp));
  dKr_dSe = 0.5 * Kr / Se + 2.0 * pow(Se, 1.0 / m - 0.5) * pow(1.0 - s1m, m - 1.0) * (1.0 - pow(1.0 - s1m, m));
}

// Water density (constant or exponential, src/materials/tdyeos.c:5-37).
__host__ __device__ inline void water_density(int type, double p, double &rho, double &drho, double &d2rho) {
  if (type == 0) {
    rho = 998.0; drho = 0.0; d2rho = 0.0;
    return;
  }
  const double rho0 = 997.16, p0 = 101325.0, beta = 1.0e-8;
  if (p < p0) { rho = rho0; drho = 0.0; d2rho = 0.0; return; }
  rho = rho0 * exp(beta * (p - p0));
  drho = beta * rho;
  d2rho = beta * drho;
}

// Constant water viscosity (src/materials/tdyeos.c:87-98).
constexpr double kViscosity = 9.94e-4;

}  // namespace tdyhip
