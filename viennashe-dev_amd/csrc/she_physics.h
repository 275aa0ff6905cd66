// she_physics.h -- FP64 physics of the SHE coefficient fill, callable on host and device.
//
// Silicon electrons/holes with the Modena non-parabolic band (the reference's
// default dispersion, viennashe/config.hpp:138; physics/dispersion.hpp:207-290),
// the SI constants of physics/constants.hpp:60-76 and the phonon parameters of
// she/scattering/config.hpp:50-116.  Expressions keep the reference's operation
// order where it fixes the rounding of a value that decides the unknown mapping
// (density of states > 0, kinetic energy <= 0).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace shephys {

#define SHE_HD __host__ __device__ inline

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr double kQ = 1.602176565e-19;
constexpr double kKB = 1.3806488e-23;
constexpr double kM0 = 9.10938291e-31;
SHE_HD double hbar() { return 6.62606896e-34 / (2.0 * 3.1415926535897932384626433832795); }

// effective masses (materials/all.hpp:49-107); carrier 0 = electrons, 1 = holes
SHE_HD double m_dos(int ct) { return ct == 0 ? 0.32 * kM0 : 0.8 * kM0; }
SHE_HD double m_long(int ct) { return ct == 0 ? 0.98 * kM0 : 0.16 * kM0; }
SHE_HD double m_trans(int ct) { return ct == 0 ? 0.19 * kM0 : 0.16 * kM0; }
SHE_HD double m_cond(int ct) {
  return ct == 0 ? 3.0 / (1.0 / m_long(ct) + 1.0 / m_trans(ct) + 1.0 / m_trans(ct)) : 0.36 * kM0;
}
SHE_HD double band_gap() { return 1.11 * kQ; }
SHE_HD double symmetry_factor(int ct) { return ct == 0 ? 6.0 : 4.0; }

// Modena band: eps (1 + alpha eps) = hbar^2 k^2 / (2 m), alpha = 0.5 / eV
SHE_HD double dos(int ct, double eps) {
  if (eps > 0.0) {
    const double h = hbar(), a = 0.5 / kQ, m = m_dos(ct);
    const double pre = m * sqrt(2.0 * m) / (8.0 * kPi * kPi * kPi * h * h * h);
    return pre * 4.0 * kPi * (1.0 + 2.0 * a * eps) * sqrt(eps * (1.0 + a * eps));
  }
  return 0.0;
}
SHE_HD double velocity(int ct, double eps) {
  if (eps < 0) return 0;
  const double a = 0.5 / kQ;
  const double g = eps * (1.0 + a * eps), dg = 1.0 + 2.0 * a * eps;
  return sqrt(m_cond(ct) / m_dos(ct)) * sqrt(2.0 * g / m_dos(ct)) / dg;
}

// physics/physics.hpp:71-184
SHE_HD double optical_phonon_number(double e, double T) { return 1.0 / (exp(e / (kKB * T)) - 1.0); }
SHE_HD double band_weight(double T, int ct) {
  const double Ncv0 = 2.50933670314359713002e19 * 100 * 100 * 100;
  const double me = ct == 0 ? 0.33 : 0.6;
  return Ncv0 * pow(me * (T / 300.0), 1.5) * 6.0;
}
SHE_HD double band_edge(int ct) { return (ct == 0 ? 1.0 : -1.0) * band_gap() * 0.5; }
SHE_HD double aux_conc(double T, int ct) {
  return band_weight(T, ct) * exp((ct == 0 ? -1.0 : 1.0) * band_edge(ct) / (kKB * T));
}
SHE_HD double built_in_potential(double T, double dn, double dp) {
  if (dn <= 0 && dp <= 0) return 0;
  const double net = dn - dp, na = aux_conc(T, 0), pa = aux_conc(T, 1);
  double x = sqrt(net * net + 4.0 * na * pa);
  if (net >= 0.0) return T * log((x + net) * 0.5 / na) * kKB / kQ;
  return -T * log((x - net) * 0.5 / pa) * kKB / kQ;
}
SHE_HD double contact_density(double T, double dn, double dp, int ct) {
  const double phiB = built_in_potential(T, dn, dp);
  return band_weight(T, ct) * exp((ct == 0 ? -1.0 : 1.0) * (band_edge(ct) - kQ * phiB) / (kKB * T));
}

// scattering rates (acoustic_phonon_scattering.hpp:98-114, optical_phonon_scattering.hpp:165-179)
SHE_HD double acoustic_rate(int ct, double T) {
  const double ul = 9040.0, rho = 2330.0, E1 = ct == 0 ? 8.9 * kQ : 5.12 * kQ;
  const double fit = ct == 0 ? 1.05 : 0.875;
  return fit * (2.0 * kPi * kKB * T * E1 * E1 / (hbar() * ul * ul * rho));
}
// phonon energy 62 meV snapped to the energy grid (optical_phonon_scattering.hpp:62-80)
SHE_HD double optical_phonon_energy(double spacing) {
  double hw = 1.24 * (kQ / 20.0);
  if (spacing > 0) {
    unsigned long long steps = (unsigned long long)(hw / spacing);
    double e0 = hw - steps * spacing, e1 = (steps + 1.0) * spacing - hw;
    if (e1 < e0) ++steps;
    hw = steps * spacing;
  }
  return hw;
}
SHE_HD double optical_rate(double hw) {
  const double rho = 2330.0, DtK = 5e10 * kQ;
  return 0.7 * (kPi * DtK * DtK * hbar() / (2.0 * rho * hw));
}

// Energy-box quantities of an element whose kinetic energies at H-1, H, H+1 are
// em, e0, ep (has_m / has_p false at the ends of the energy grid):
// box_height (she/assemble_common.hpp:411-427),
// averaged_density_of_states (:431-462), integral_vZ (:519-551).
SHE_HD double box_height3(double em, double e0, double ep, bool has_m, bool has_p) {
  double h = 0;
  if (has_m) h = fabs(e0 - em);
  if (has_p) h += fabs(ep - e0);
  return h / 2.0;
}
SHE_HD double avg_dos3(int ct, double em, double e0, double ep, bool has_m, bool has_p) {
  double Z = dos(ct, e0), n = 1;
  if (has_p) { Z += dos(ct, (e0 + ep) / 2.0); ++n; }
  if (has_m) { Z += dos(ct, (e0 + em) / 2.0); ++n; }
  return Z / n;
}
SHE_HD double integral_vZ3(int ct, double em, double e0, double ep, bool has_m, bool has_p) {
  double vZ = velocity(ct, e0) * dos(ct, e0), n = 1;
  if (has_p) { double k = (e0 + ep) / 2.0; vZ += velocity(ct, k) * dos(ct, k); ++n; }
  if (has_m) { double k = (e0 + em) / 2.0; vZ += velocity(ct, k) * dos(ct, k); ++n; }
  vZ /= n;
  return vZ * box_height3(em, e0, ep, has_m, has_p);
}

#undef SHE_HD
}  // namespace shephys
