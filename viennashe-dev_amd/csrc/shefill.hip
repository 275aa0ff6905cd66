// shefill.hip -- SHE coefficient fill on MI355X behind include/shefill.h.
//
// Host: structured mesh, energy grid, expansion orders, unknown mapping and the
// sparsity of the reduced (odd-eliminated) system.  The pattern is value
// independent on structured meshes: the drift coupling b_dir is nonzero only
// where the diffusion coupling a_dir is (checked for L = 1..7 in both
// directions), so every entry the reference creates is an a-coupling product.
// Device: all FP64 values -- scalar coefficients per (cell, H) and (facet, H),
// the eliminated rows -S^eo (S^oo)^-1 S^oe per row, unknown scaling and the
// column/row normalisation -- written straight into the resident CSR.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "coupling_tables.inc"
#include "she_physics.h"
#include "shefill.h"

using namespace shephys;

// Device scratch of shefill_density_update_device_noalloc: kNoAllocParts
// per-workgroup partial sums of the update norm, then the norm itself.  Every
// device the library runs on holds its own copy for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (noalloc_scratch), never through this name.
constexpr int kNoAllocParts = 65536;
static __device__ __align__(256) double scratch[kNoAllocParts + 1];

namespace {

thread_local std::string g_ferr;
void ferr(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_ferr = buf;
}
#define FHIP(expr)                                                                    \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ferr("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return -10;                                                                     \
    }                                                                                 \
  } while (0)

constexpr int kB = 256;
// facet slots of a cell (creation order of a quadrilateral's edges) and the
// direction from the cell to the neighbour across each: +x 0, -x 1, +y 2, -y 3
constexpr int kSlotDir[4] = {3, 1, 0, 2};  // bottom, left, right, top
__host__ __device__ inline int rev_dir(int d) { return d ^ 1; }

int even_count(int L) { int n = 0; for (int l = 0; l <= L; l += 2) n += 2 * l + 1; return n; }
int odd_count(int L) { int n = 0; for (int l = 1; l <= L; l += 2) n += 2 * l + 1; return n; }
std::vector<int> harmonics(int L, bool odd) {
  std::vector<int> h;
  for (int l = odd ? 1 : 0; l <= L; l += 2)
    for (int m = -l; m <= l; ++m) h.push_back(l * l + l + m);
  return h;
}

// --- coupling matrices with the reference's std::map semantics -----------------
struct CMat {
  int K = 0;
  std::vector<char> p;
  std::vector<double> v;
  explicit CMat(int k = 0) : K(k), p((size_t)k * k, 0), v((size_t)k * k, 0.0) {}
};
CMat transpose(const CMat &m) {
  CMat t(m.K);
  for (int i = 0; i < m.K; ++i)
    for (int j = 0; j < m.K; ++j)
      if (m.p[i * m.K + j]) { t.p[j * m.K + i] = 1; t.v[j * m.K + i] = m.v[i * m.K + j]; }
  return t;
}
// sparse_matrix::operator*(factor) (math/linalg_util.hpp:132-143) scales the
// transposed position of every stored entry, inserting it when absent.
CMat scaled(const CMat &m, double f) {
  CMat r = m;
  for (int i = 0; i < m.K; ++i)
    for (int j = 0; j < m.K; ++j)
      if (m.p[i * m.K + j]) { r.p[j * m.K + i] = 1; r.v[j * m.K + i] *= f; }
  return r;
}
void add_into(CMat &d, const CMat &s) {
  for (size_t i = 0; i < d.p.size(); ++i)
    if (s.p[i]) { d.p[i] = 1; d.v[i] += s.v[i]; }
}
// coupling_matrix_in_direction (she/assemble_common.hpp:176-290) for the unit
// direction of code d; 1D uses the x matrix only.
CMat in_direction(const CMat *m, int d, int dim, int ct) {
  const double px = d == 0 ? 1.0 : (d == 1 ? -1.0 : 0.0), py = d == 2 ? 1.0 : (d == 3 ? -1.0 : 0.0);
  const double s = sqrt(m_dos(ct) / m_trans(ct));
  CMat r = scaled(m[0], px * s);
  if (dim > 1) add_into(r, scaled(m[1], py * s));
  return r;
}

// --- device kernels -----------------------------------------------------------
struct DevTables {
  int K, NH, carrier, dim;
  double T, kT, hw, rate_ac, rate_op, Nop, tau, spacing;
  const double *Hval, *shc, *shf, *cvol, *fvol, *flen, *dn, *dp;
  const int *contact, *fc1, *fc2;
  const double *Ae, *Be, *Ao, *Bo;  // [4][K][K]
  const int *ev_h, *od_h;           // harmonic index lists
  const int *ev_off, *od_off;       // offsets per L in the lists (L+1 entries)
};

__device__ inline double kin_cell(const DevTables &t, int c, int H) {
  return t.carrier == 0 ? t.Hval[H] - t.shc[c] : t.shc[c] - t.Hval[H];
}
__device__ inline double kin_facet(const DevTables &t, int f, int H) {
  return t.carrier == 0 ? t.Hval[H] - t.shf[f] : t.shf[f] - t.Hval[H];
}
__device__ inline void kin3(const DevTables &t, bool facet, int e, int H, double &em, double &e0, double &ep,
                            bool &hm, bool &hp) {
  hm = H > 0;
  hp = H < t.NH - 1;
  e0 = facet ? kin_facet(t, e, H) : kin_cell(t, e, H);
  em = hm ? (facet ? kin_facet(t, e, H - 1) : kin_cell(t, e, H - 1)) : 0.0;
  ep = hp ? (facet ? kin_facet(t, e, H + 1) : kin_cell(t, e, H + 1)) : 0.0;
}
__device__ inline double avgZ(const DevTables &t, bool facet, int e, int H) {
  double em, e0, ep;
  bool hm, hp;
  kin3(t, facet, e, H, em, e0, ep, hm, hp);
  return avg_dos3(t.carrier, em, e0, ep, hm, hp);
}
__device__ inline double bheight(const DevTables &t, bool facet, int e, int H) {
  double em, e0, ep;
  bool hm, hp;
  kin3(t, facet, e, H, em, e0, ep, hm, hp);
  return box_height3(em, e0, ep, hm, hp);
}

// contact normalisation (she/boundary_conditions.hpp:74-101) and contact density
__global__ void k_bnd_corr(DevTables t, int ncells, double *corr, double *ncont) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  if (!t.contact[c]) { corr[c] = 1.0; ncont[c] = 0.0; return; }
  double s = 0;
  for (int H = 1; H < t.NH - 1; ++H) {
    double mid = kin_cell(t, c, H);
    double lo = fmax((kin_cell(t, c, H - 1) + mid) / 2.0, 0.0);
    double up = (kin_cell(t, c, H + 1) + mid) / 2.0;
    if (lo < up && up >= 0) s += exp(-mid / t.kT) * avgZ(t, false, c, H) * bheight(t, false, c, H);
  }
  corr[c] = s * symmetry_factor(t.carrier);
  ncont[c] = contact_density(t.T, t.dn[c], t.dp[c], t.carrier);
}

// Per facet node (f, H): diffusion coefficient, drift coefficient seen from each
// cell, odd diagonal (out-scattering; no elastic round-off skip for odd rows).
__global__ void k_fnode(DevTables t, int nfn, const int *fn_f, const int *fn_H, const int *fn_out,
                        double *cdiff, double *cdrift, double *odiag) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfn) return;
  const int f = fn_f[k], H = fn_H[k], c1 = t.fc1[f], c2 = t.fc2[f];
  double em, e0, ep;
  bool hm, hp;
  kin3(t, true, f, H, em, e0, ep, hm, hp);
  const double area = t.fvol[f], len = t.flen[f];
  cdiff[k] = integral_vZ3(t.carrier, em, e0, ep, hm, hp) * area;
  // lower/upper kinetic energies of the facet box (assemble_common.hpp:323-388), interior H
  const double lo = (kin_cell(t, c1, H - 1) + kin_cell(t, c2, H - 1) + kin_cell(t, c1, H) + kin_cell(t, c2, H)) / 4.0;
  const double up = (kin_cell(t, c1, H + 1) + kin_cell(t, c2, H + 1) + kin_cell(t, c1, H) + kin_cell(t, c2, H)) / 4.0;
  const double izhk = 0.5 * (dos(t.carrier, up) * velocity(t.carrier, up) - dos(t.carrier, lo) * velocity(t.carrier, lo));
  // force_on_facet_accessor: (E_kin(cell) - E_kin(other)) / distance
  const double f1 = (kin_cell(t, c1, H) - kin_cell(t, c2, H)) / len;
  const double f2 = (kin_cell(t, c2, H) - kin_cell(t, c1, H)) / len;
  cdrift[2 * k] = izhk * f1 * area * len / 2.0;
  cdrift[2 * k + 1] = izhk * f2 * area * len / 2.0;
  // out-scattering on the facet box, volume |f| * d / dim
  const double V = area * (len / (double)t.dim);
  const double Z0 = avg_dos3(t.carrier, em, e0, ep, hm, hp), h0 = box_height3(em, e0, ep, hm, hp);
  double d = 0;
  if (t.rate_ac > 0) d += 2.0 * kPi * t.rate_ac * Z0 * Z0 * V * h0;
  for (int q = 0; q < 2; ++q) {
    const int Hf = fn_out[2 * k + q];
    if (Hf < 0) continue;
    const double rate = q == 0 ? t.rate_op * (t.Nop + 1.0) : t.rate_op * t.Nop;
    const double Zf = avgZ(t, true, f, Hf), hf = bheight(t, true, f, Hf);
    d += 2.0 * kPi * rate * Z0 * Zf * V * fmin(h0, hf);
  }
  odiag[k] = d;
}

// Per cell node (c, H): diagonal scattering/boundary terms, rhs and the two
// optical in-scattering coefficients (she/assemble_scattering.hpp:116-335,
// she/assemble_boundary.hpp:69-183).
__global__ void k_node(DevTables t, int nn, const int *nd_c, const int *nd_H, const int *nd_out, const int *nd_in,
                       const double *corr, const double *ncont, double *diag0, double *diagr, double *rhs0,
                       double *inco) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nn) return;
  const int c = nd_c[k], H = nd_H[k];
  double em, e0, ep;
  bool hm, hp;
  kin3(t, false, c, H, em, e0, ep, hm, hp);
  const double V = t.cvol[c];
  const double Z0 = avg_dos3(t.carrier, em, e0, ep, hm, hp), h0 = box_height3(em, e0, ep, hm, hp);
  double d0 = 0, dr = 0, r0 = 0;
  if (t.contact[c]) {
    const double coeff = V * h0 / t.tau * Z0;
    double bnd = 0.0;
    if (H >= 1 && H < t.NH - 1) {
      const double lo = fmax((em + e0) / 2.0, 0.0), up = (ep + e0) / 2.0;
      if (lo < up && up >= 0) bnd = ncont[c] * exp(-e0 / t.kT) / corr[c];
    }
    d0 += coeff;
    dr += coeff;
    r0 += bnd * coeff;
  }
  // acoustic: elastic, the (0,0) out/in pair is skipped
  if (t.rate_ac > 0) dr += 2.0 * kPi * t.rate_ac * Z0 * Z0 * V * h0;
  // optical out (emission to H-k with N+1, absorption to H+k with N)
  for (int q = 0; q < 2; ++q) {
    const int Hf = nd_out[2 * k + q];
    if (Hf < 0) continue;
    const double rate = q == 0 ? t.rate_op * (t.Nop + 1.0) : t.rate_op * t.Nop;
    const double Zf = avgZ(t, false, c, Hf), hf = bheight(t, false, c, Hf);
    const double coeff = 2.0 * kPi * rate * Z0 * Zf * V * fmin(h0, hf);
    d0 += coeff;
    dr += coeff;
  }
  // optical in: from H-k (absorption, N) and from H+k (emission, N+1)
  for (int q = 0; q < 2; ++q) {
    const int Hi = nd_in[2 * k + q];
    double v = 0;
    if (Hi >= 0) {
      const double rate = q == 0 ? t.rate_op * t.Nop : t.rate_op * (t.Nop + 1.0);
      const double Zi = avgZ(t, false, c, Hi), hi = bheight(t, false, c, Hi);
      v = -2.0 * kPi * rate * Zi * Z0 * V * fmin(hi, h0);
    }
    inco[2 * k + q] = v;
  }
  diag0[k] = d0;
  diagr[k] = dr;
  rhs0[k] = r0;
}

// ---- solution back end: scale back, odd recovery, carrier density, current
// x_full[i] = x[i] * scaling[i] * col_factor[i] for the even unknowns
// (she/linear_solver.hpp:172-181)
__global__ void k_scale_back(int n, const double *x, const double *scal, const double *colfac, int on, double *xf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  xf[i] = on ? x[i] * (scal[i] * colfac[i]) : x[i];
}

struct RecArgs {
  int n, NH, nfn;
  const int *fn_f, *fn_H, *fn_L, *fn_base, *fdir, *node_of, *nd_base, *nd_L;
  const double *cdiff, *cdrift, *odiag;
};

// Odd unknowns of one facet node from its row of the full system:
// f_o = (b_o - sum_e S^oe f_e) / S^oo_oo with b_o = 0 (she/eliminate.hpp:305-354).
// S^oe couples to the even unknowns of both cells X of the facet:
// diffusion * a^T + drift seen from X * b^T, in the direction from the other
// cell to X -- the same entries the elimination used.
__global__ void k_recover_odd(DevTables t, RecArgs a, double *xf) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.nfn) return;
  const int f = a.fn_f[k], H = a.fn_H[k], Lf = a.fn_L[k], K = t.K;
  const int d12 = a.fdir[f];
  const int ob = t.od_off[Lf], oe = t.od_off[Lf + 1];
  double *out = xf + a.n + a.fn_base[k];
  for (int o = ob; o < oe; ++o) {
    const int ho = t.od_h[o];
    double acc = 0.0;
    for (int side = 0; side < 2; ++side) {
      const int X = side ? t.fc2[f] : t.fc1[f];
      const int dir = side ? d12 : rev_dir(d12);
      const int nX = a.node_of[(size_t)X * a.NH + H];
      const int base = a.nd_base[nX], LX = a.nd_L[nX];
      const double *Ao = t.Ao + (size_t)dir * K * K, *Bo = t.Bo + (size_t)dir * K * K;
      const double cd = a.cdiff[k], cr = a.cdrift[2 * k + side];
      for (int r = t.ev_off[LX]; r < t.ev_off[LX + 1]; ++r) {
        const int hq = t.ev_h[r];
        const double soe = cd * Ao[ho * K + hq] + cr * Bo[ho * K + hq];
        acc += soe * xf[base + (r - t.ev_off[LX])];
      }
    }
    out[o - ob] = -acc / a.odiag[k];
  }
}

// carrier density per cell (she/postproc/carrier_density.hpp:68-131, window [0, 1.0])
__global__ void k_density(DevTables t, int nc, const int *idxc, const double *xf, double *dens) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double n = 0.0;
  for (int H = 1; H < t.NH - 1; ++H) {
    const int u = idxc[(size_t)c * t.NH + H];
    if (u < 0) continue;
    double lo = (kin_cell(t, c, H - 1) + kin_cell(t, c, H)) / 2.0;
    if (lo < 0.0) lo = 0.0;
    double up = (kin_cell(t, c, H + 1) + kin_cell(t, c, H)) / 2.0;
    if (up > 1.0) up = 1.0;
    if (up >= 0 && lo < up) n += xf[u] * avgZ(t, false, c, H) * bheight(t, false, c, H);
  }
  dens[c] = n * symmetry_factor(t.carrier);
}

struct CurArgs {
  int nf, n;
  const int *fnode_of, *fn_base;
  const double *cx, *cy;
  double ax[3], ay[3];
};

// current through each interior facet (she/postproc/current_density.hpp:48-131):
// projection of the l = 1 components on the facet normal
__global__ void k_current(DevTables t, CurArgs a, const double *xf, double *cur) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= a.nf) return;
  const int c1 = t.fc1[f], c2 = t.fc2[f];
  double nx = a.cx[c2] - a.cx[c1], ny = t.dim > 1 ? a.cy[c2] - a.cy[c1] : 0.0;
  const double nn = sqrt(nx * nx + ny * ny);
  nx /= nn;
  ny /= nn;
  const double s = sqrt(m_dos(t.carrier) / m_trans(t.carrier));
  double vn = 0.0;
  for (int H = 1; H < t.NH - 1; ++H) {
    const int k = a.fnode_of[(size_t)f * t.NH + H];
    if (k < 0) continue;
    const double *x = xf + a.n + a.fn_base[k];
    double em, e0, ep;
    bool hm, hp;
    kin3(t, true, f, H, em, e0, ep, hm, hp);
    const double factor = integral_vZ3(t.carrier, em, e0, ep, hm, hp) * symmetry_factor(t.carrier);
    vn += (a.ax[0] * x[0] + a.ax[1] * x[1] + a.ax[2] * x[2]) * factor * s * nx;
    if (t.dim > 1) vn += (a.ay[0] * x[0] + a.ay[1] * x[1] + a.ay[2] * x[2]) * factor * s * ny;
  }
  const double pol = t.carrier == 0 ? -1.0 : 1.0;
  const double Y00 = sqrt(1.0 / (2.0 * kPi)) / sqrt(2.0);
  cur[f] = -pol * kQ * vn / Y00;
}

// Gummel update of the carrier density (simulator.hpp:524-600 update_quantity):
// the damped value, linear cur (1 - a) + new a, or logarithmic cur^(1-a) new^a
// with new limited to >= 1e-10 cur; then the relative update norm of
// simulator.hpp:286-320 between the new value and the updated density,
// (new - upd) / upd (new where upd is 0), squared and summed per block.
__global__ void k_density_update(int nc, const double *dnew, double *dcur, double alpha, int logarithmic,
                                 double *part) {
  __shared__ double red[kB / 64];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double z2 = 0.0;
  if (c < nc) {
    const double nv = dnew[c];
    const double cur = dcur[c];
    double upd;
    if (logarithmic) {
      const double nl = nv < cur * 1e-10 ? cur * 1e-10 : nv;
      upd = pow(cur, 1.0 - alpha) * pow(nl, alpha);
    } else {
      upd = cur * (1.0 - alpha) + nv * alpha;
    }
    dcur[c] = upd;
    const double zw = fabs(upd) > 0 ? (nv - upd) / upd : nv;
    z2 = zw * zw;
  }
  for (int off = 32; off > 0; off >>= 1) z2 += __shfl_xor(z2, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = z2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kB / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

// Terminal current (postproc/terminal_current.hpp:50-107): the facet currents
// summed over the facets with exactly one cell in the terminal, times the
// facet area (the weighted interface area of an orthogonal mesh), with +
// where the terminal cell is the facet's first cell; per-block partial sums.
__global__ void k_terminal_current(int nf, const int *fc1, const int *fc2, const double *fvol,
                                   const unsigned char *mask, const double *jf, double *part) {
  __shared__ double red[kB / 64];
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (f < nf) {
    const bool m1 = mask[fc1[f]] != 0, m2 = mask[fc2[f]] != 0;
    if (m1 != m2) v = m1 ? jf[f] * fvol[f] : -(jf[f] * fvol[f]);
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kB / 64; ++w) t += red[w];
    part[blockIdx.x] = t;
  }
}

// Previous iterate on the unknowns of a new build (she/transfer_to_new_h_space.hpp:58-196,
// the cell part): each new (cell, level) takes the values of the old level of
// the same cell whose kinetic energy is nearest (energy_index_H,
// she/assemble_scattering.hpp:64-105, walked on the old grid from the same
// index), moving up past old levels without unknowns (below the band edge);
// harmonics the old node lacks are zero, as are levels the old grid does not
// reach.  f_new holds the even part of a full solution of the new build.
struct TransferArgs {
  int nc, NH, NH_old, carrier, L_max;
  const double *H, *shc, *H_old, *shc_old;
  const int *idx, *ord, *idx_old, *ord_old;
  const double *x_old;  // full solution of the old build (even unknowns first)
  double *f_new;
};
__device__ inline int even_of(int L) { return L * (L + 1) / 2; }  // even harmonics up to odd L
__global__ void k_transfer_even(TransferArgs a) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= (int64_t)a.nc * a.NH) return;
  const int c = (int)(k / a.NH), H = (int)(k % a.NH);
  const int u = a.idx[k];
  if (u < 0) return;
  const int e_new = even_of(a.ord[k]);
  for (int i = 0; i < e_new; ++i) a.f_new[u + i] = 0.0;
  if (H < 1 || H > a.NH - 2) return;
  const double E = a.carrier == 0 ? a.H[H] - a.shc[c] : a.shc[c] - a.H[H];
  auto kold = [&](long h) { return a.carrier == 0 ? a.H_old[h] - a.shc_old[c] : a.shc_old[c] - a.H_old[h]; };
  long r = H < a.NH_old - 1 ? H : a.NH_old - 1;
  if (E < kold(r) || E > kold(r)) {
    long inc = 1;
    double diff = E - kold(r), diff0 = diff;
    if (diff < 0) inc = -1;
    while (r < (long)(a.NH_old - 1) && r > 0 && diff * diff0 > 0) {
      r += inc;
      diff = E - kold(r);
    }
    if (diff * diff0 > 0) return;  // the old grid does not reach this energy
    double e1 = fabs(diff), e2 = e1;
    if (r - inc >= 0 && r - inc < (long)a.NH_old) e2 = fabs(E - kold(r - inc));
    if (e2 < e1) r -= inc;
  }
  const int64_t ko = (int64_t)c * a.NH_old;
  while (a.ord_old[ko + r] == 0 && r < a.NH_old - 1) ++r;
  const int uo = a.idx_old[ko + r];
  if (a.ord_old[ko + r] <= 0 || uo < 0) return;
  const int e_old = even_of(a.ord_old[ko + r]);
  for (int i = 0; i < e_new && i < e_old; ++i) a.f_new[u + i] = a.x_old[uo + i];
}
// Density rescaling of the transferred values per cell (density_old /
// density_new where both are positive), then the reduced, scaled unknowns
// x = f / (scaling * column factor) of the new system.
__global__ void k_transfer_scale(int nc, int NH, const int *idx, const int *ord, const double *dens_old,
                                 const double *dens_new, const double *f_new, const double *scal,
                                 const double *colfac, int scaled, double *x) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= (int64_t)nc * NH) return;
  const int u = idx[k];
  if (u < 0) return;
  const int c = (int)(k / NH);
  const double s = (dens_new[c] > 0 && dens_old[c] > 0) ? dens_old[c] / dens_new[c] : 1.0;
  const int e = even_of(ord[k]);
  for (int i = 0; i < e; ++i) {
    const double f = f_new[u + i] * s;
    x[u + i] = scaled ? f / (scal[u + i] * colfac[u + i]) : f;
  }
}

// Sums part[0..cnt) in index order in one thread -- the order of the host
// loop of shefill_density_update_device, so the norm is the same bit for bit --
// and stores the square root in *out.
__global__ void k_sum_sqrt_ordered(int cnt, const double *__restrict__ part, double *__restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double t = 0.0;
  for (int i = 0; i < cnt; ++i) t += part[i];
  *out = sqrt(t);
}

struct RowArgs {
  int n;
  const int *rp, *ci, *row_node;
  const int *nd_c, *nd_H, *nd_base, *nd_L, *nd_fn, *nd_nb, *nd_inn;
  const int *fn_f, *fn_L;
  const double *cdiff, *cdrift, *odiag, *diag0, *diagr, *rhs0, *inco;
  double *val, *rhs;
};

// One thread per row: value of every stored entry.  (Lanes per row measured
// slower: every lane repeats the row's facet set-up and the entry loops diverge.)
constexpr int kRowLanes = 8;  // lanes per row of the normalisation
// With K <= kRowsLdsK the four coupling tables (4 directions each) are staged
// in LDS first: their lookups in the elimination sums below go to LDS instead
// of scattered cached loads.  The arithmetic is the same either way.
constexpr int kRowsLdsK = 16;
__global__ __launch_bounds__(kB) void k_rows(DevTables t, RowArgs a) {
  extern __shared__ double tab[];
  const int K = t.K;
  if (K <= kRowsLdsK) {
    const int kk = 4 * K * K;
    for (int i = threadIdx.x; i < kk; i += blockDim.x) {
      tab[i] = t.Ae[i];
      tab[kk + i] = t.Be[i];
      tab[2 * kk + i] = t.Ao[i];
      tab[3 * kk + i] = t.Bo[i];
    }
    __syncthreads();
    t.Ae = tab;
    t.Be = tab + kk;
    t.Ao = tab + 2 * kk;
    t.Bo = tab + 3 * kk;
  }
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.n) return;
  const int nd = a.row_node[row];
  const int c = a.nd_c[nd], base = a.nd_base[nd], Lc = a.nd_L[nd];
  const int r = row - base;
  const int hr = t.ev_h[t.ev_off[Lc] + r];
  // facets with odd unknowns: slot data
  int fnk[4], nbk[4], dirk[4], side_c[4];
  for (int s = 0; s < 4; ++s) {
    fnk[s] = a.nd_fn[4 * nd + s];
    nbk[s] = a.nd_nb[4 * nd + s];
    dirk[s] = kSlotDir[s];
    side_c[s] = 0;
    if (fnk[s] >= 0) side_c[s] = (t.fc1[a.fn_f[fnk[s]]] == c) ? 0 : 1;
  }
  for (int j = a.rp[row]; j < a.rp[row + 1]; ++j) {
    const int col = a.ci[j];
    const int cn = a.row_node[col];
    const int rq = col - a.nd_base[cn];
    const int hq = t.ev_h[t.ev_off[a.nd_L[cn]] + rq];
    double v = 0.0;
    if (cn == nd) {
      if (rq == r) v += (r == 0) ? a.diag0[nd] : a.diagr[nd];
    } else if (cn == a.nd_inn[2 * nd]) {
      v = (r == 0 && rq == 0) ? a.inco[2 * nd] : 0.0;
    } else if (cn == a.nd_inn[2 * nd + 1]) {
      v = (r == 0 && rq == 0) ? a.inco[2 * nd + 1] : 0.0;
    }
    // elimination through every facet whose far side is this column's node
    for (int s = 0; s < 4; ++s) {
      const int fk = fnk[s];
      if (fk < 0) continue;
      const bool self = (cn == nd);
      if (!self && cn != nbk[s]) continue;
      const int d = dirk[s];
      const int dd = self ? rev_dir(d) : d;  // direction other -> column cell
      const int Lf = a.fn_L[fk];
      const double cd = a.cdiff[fk];
      const double crow = a.cdrift[2 * fk + side_c[s]];
      const double ccol = a.cdrift[2 * fk + (self ? side_c[s] : 1 - side_c[s])];
      const double *Ae = t.Ae + (size_t)d * K * K, *Be = t.Be + (size_t)rev_dir(d) * K * K;
      const double *Ao = t.Ao + (size_t)dd * K * K, *Bo = t.Bo + (size_t)dd * K * K;
      double s_acc = 0.0;
      const int ob = t.od_off[Lf], oe = t.od_off[Lf + 1];
      for (int o = ob; o < oe; ++o) {
        const int ho = t.od_h[o];
        const double seo = cd * Ae[hr * K + ho] + crow * Be[hr * K + ho];
        if (seo == 0.0) continue;
        const double soe = cd * Ao[ho * K + hq] + ccol * Bo[ho * K + hq];
        s_acc += seo * soe;
      }
      v -= s_acc / a.odiag[fk];
    }
    a.val[j] = v;
  }
  a.rhs[row] = (r == 0) ? a.rhs0[nd] : 0.0;
}

// unknown scaling of the columns (she/rescaling.hpp:89-107): every entry
// times its column's factor
__global__ void k_scale(int64_t nnz, const int *ci, const double *scal, double *val) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nnz; j += (int64_t)gridDim.x * blockDim.x)
    val[j] *= scal[ci[j]];
}
// log-magnitude sum of the structural entries of each column, in ascending
// row order (the order of the transposed std::map rows)
__global__ void k_collog(int n, const int *cptr, const int *cpos, const double *val, double *logsum) {
  int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  double s = 0.0;
  for (int k = cptr[col]; k < cptr[col + 1]; ++k) s += log(fabs(val[cpos[k]]));
  logsum[col] = s;
}
// Column-ordered positions of the structural entries (the diagonal that the
// normalisation inserts is not one), built on the device from the stored
// pattern: counts per column, then each entry takes a slot in its column, then
// each column's slots are sorted, which orders them by row.
__global__ void k_colcount(int n, const int *rp, const int *ci, const int *diag_pos, const unsigned char *dstruct,
                           int *colcount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = rp[i]; j < rp[i + 1]; ++j)
    if (j != diag_pos[i] || dstruct[i]) atomicAdd(&colcount[ci[j]], 1);
}
__global__ void k_colfill(int n, const int *rp, const int *ci, const int *diag_pos, const unsigned char *dstruct,
                          int *fillp, int *cpos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = rp[i]; j < rp[i + 1]; ++j)
    if (j != diag_pos[i] || dstruct[i]) cpos[atomicAdd(&fillp[ci[j]], 1)] = j;
}
__global__ void k_colsort(int n, const int *cptr, int *cpos) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n) return;
  const int b = cptr[col], e = cptr[col + 1];
  for (int k = b + 1; k < e; ++k) {  // insertion sort: a column holds a few entries
    const int v = cpos[k];
    int m = k - 1;
    while (m >= b && cpos[m] > v) {
      cpos[m + 1] = cpos[m];
      --m;
    }
    cpos[m + 1] = v;
  }
}
__global__ void k_colnorm(int n, const int *colcount, const int *diag_pos, const double *val, double *logsum,
                          double *colfac) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double cn = exp(logsum[i] / (double)colcount[i]);
  if (val[diag_pos[i]] < 0.0) cn *= -1.0;
  logsum[i] = cn;  // reused as the column normaliser
  colfac[i] = 1.0 / cn;
}
// entries divided by their column normaliser, then the row by its 2-norm
// (signed like the diagonal); kRowLanes lanes per row load and scale, the
// squares are summed in column order as linalg_util.hpp:200-282 does
__global__ void k_rownorm(int n, const int *rp, const int *ci, const int *diag_pos, const double *cn, double *val,
                          double *rhs) {
  const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int row = (int)(gt / kRowLanes), lane = (int)(gt % kRowLanes);
  const bool ok = row < n;
  const int b = ok ? rp[row] : 0, len = ok ? rp[row + 1] - b : 0;
  double s = 0.0, vd = 0.0;
  if (ok) vd = val[diag_pos[row]] / cn[row];
  for (int j0 = 0; j0 < len; j0 += kRowLanes) {
    const int j = j0 + lane;
    const double v = j < len ? val[b + j] / cn[ci[b + j]] : 0.0;
    const int m = min(kRowLanes, len - j0);
    for (int q = 0; q < m; ++q) {
      const double w = __shfl(v, q, kRowLanes);
      s += w * w;
    }
  }
  if (!ok) return;
  double rn = sqrt(s);
  if (vd < 0.0) rn *= -1.0;
  for (int j = rp[row] + lane; j < rp[row + 1]; j += kRowLanes) val[j] = (val[j] / cn[ci[j]]) / rn;
  if (lane == 0) rhs[row] /= rn;
}
__global__ void k_node_scaling(int n, const int *row_node, const int *nd_c, const int *nd_H, DevTables t, int on,
                               double *scal) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const int nd = row_node[row];
  scal[row] = on ? exp(-kin_cell(t, nd_c[nd], nd_H[nd]) / t.kT) : 1.0;
}
__global__ void k_rp64(int n, const int *rp, int64_t *rp64) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) rp64[i] = rp[i];
}

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  ~DBuf() { if (p) (void)hipFree(p); }
  int alloc(size_t cnt) {
    if (cnt <= n && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(cnt, 1)) != hipSuccess) {
      ferr("hipMalloc of %zu bytes failed", sizeof(T) * cnt);
      return -11;
    }
    n = cnt;
    return 0;
  }
};

}  // namespace

struct shefill_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double fill_ms = 0;
  shefill_config cfg{};
  int dim = 1, nx = 0, ny = 1, nc = 0, nf = 0, NH = 0, K = 0;
  double T = 300;
  std::vector<double> cx, cy, cvol, dn, dp, psi, fvol, flen, Hval, shc, shf;
  std::vector<int> contact, fc1, fc2, cf, ordc, idxc, ordf, idxf, node_of, fnode_of;
  // nodes
  std::vector<int> nd_c, nd_H, nd_base, nd_L, nd_fn, nd_nb, nd_inn, nd_in, nd_out;
  std::vector<int> fn_f, fn_H, fn_L, fn_out;
  std::vector<int> fdir;     // direction code of c1 -> c2 per facet (0 = +x, 2 = +y)
  std::vector<int> fn_base;  // first odd unknown of each facet node, relative to n
  double a_l01[2][3] = {{0, 0, 0}, {0, 0, 0}};  // raw a_x, a_y entries (0, 1..3)
  int64_t n = 0, n_odd = 0, nnz = 0;
  std::vector<int> rp, ci, row_node, diag_pos;
  std::vector<unsigned char> diag_struct;
  std::vector<int64_t> blocks;
  // device
  DBuf<double> dHval, dshc, dshf, dcvol, dfvol, dflen, ddn, ddp, dAe, dBe, dAo, dBo;
  DBuf<int> dcontact, dfc1, dfc2, dev_h, dod_h, dev_off, dod_off;
  DBuf<int> dnd_c, dnd_H, dnd_base, dnd_L, dnd_fn, dnd_nb, dnd_inn, dnd_in, dnd_out, dfn_f, dfn_H, dfn_L, dfn_out;
  DBuf<int> drp, dci, drow_node, ddiag_pos, dcolcount, dcptr, dcpos;
  DBuf<unsigned char> ddiag_struct;
  DBuf<double> dcorr, dncont, dcdiff, dcdrift, dodiag, ddiag0, ddiagr, drhs0, dinco;
  DBuf<double> dval, drhs, dscal, dcolfac, dlog;
  DBuf<int64_t> drp64;
  DBuf<int> dfdir, dfn_base, didxc, dfnode_of, dnode_of, dordc;
  // the previous build on the same mesh, for shefill_transfer_device: its
  // energies, band-edge shifts and even mapping (prev_nc < 0: none)
  int prev_nc = -1, prev_NH = 0;
  DBuf<double> dprev_H, dprev_shc, dtr_f, dtr_dens;
  DBuf<int> dprev_idx, dprev_ord;
  DBuf<double> dcx, dcy;
  DevTables t{};
  bool built = false;
};

namespace {

template <typename T>
int upload(shefill_ctx *c, DBuf<T> &b, const std::vector<T> &h) {
  if (int e = b.alloc(h.size())) return e;
  if (!h.empty()) FHIP(hipMemcpyAsync(b.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->stream));
  return 0;
}

double kin_c(const shefill_ctx *c, int cell, int H) {
  return c->cfg.carrier == 0 ? c->Hval[H] - c->shc[cell] : c->shc[cell] - c->Hval[H];
}
double kin_f(const shefill_ctx *c, int f, int H) {
  return c->cfg.carrier == 0 ? c->Hval[H] - c->shf[f] : c->shf[f] - c->Hval[H];
}
double avg_dos_host(const shefill_ctx *c, bool facet, int e, int H) {
  auto k = [&](int h) { return facet ? kin_f(c, e, h) : kin_c(c, e, h); };
  const bool hm = H > 0, hp = H < c->NH - 1;
  return avg_dos3(c->cfg.carrier, hm ? k(H - 1) : 0.0, k(H), hp ? k(H + 1) : 0.0, hm, hp);
}
// energy_index_H: she/assemble_scattering.hpp:64-105
long energy_index_host(const shefill_ctx *c, bool facet, int e, long hint, double E) {
  auto k = [&](long h) { return facet ? kin_f(c, e, (int)h) : kin_c(c, e, (int)h); };
  long r = hint;
  if (E < k(r) || E > k(r)) {
    long inc = 1;
    double diff = E - k(r), diff0 = diff;
    if (diff < 0) inc = -1;
    while (r < (long)(c->NH - 1) && r > 0 && diff * diff0 > 0) {
      r += inc;
      diff = E - k(r);
    }
    if (diff * diff0 > 0) return -1;
    double e1 = fabs(diff), e2 = e1;
    if (r - inc >= 0 && r - inc < (long)c->NH) e2 = fabs(E - k(r - inc));
    if (e2 < e1) return r - inc;
  }
  return r;
}

int build_geometry(shefill_ctx *c, const shefill_device *d) {
  c->dim = d->dim;
  c->nx = d->nx;
  c->ny = d->dim > 1 ? d->ny : 1;
  c->nc = c->nx * c->ny;
  c->T = d->temperature;
  const int nc = c->nc;
  c->cx.assign(nc, 0); c->cy.assign(nc, 0); c->cvol.assign(nc, 0);
  c->dn.assign(d->doping_n, d->doping_n + nc);
  c->dp.assign(d->doping_p, d->doping_p + nc);
  c->psi.assign(d->potential, d->potential + nc);
  c->contact.assign(d->contact, d->contact + nc);
  for (int j = 0; j < c->ny; ++j)
    for (int i = 0; i < c->nx; ++i) {
      const int k = j * c->nx + i;
      if (c->dim == 1) {
        c->cx[k] = (d->xv[i] + d->xv[i + 1]) / 2.0;
        c->cvol[k] = fabs(d->xv[i + 1] - d->xv[i]);
      } else {
        c->cx[k] = (d->xv[i] + d->xv[i + 1] + d->xv[i] + d->xv[i + 1]) / 4.0;
        c->cy[k] = (d->yv[j] + d->yv[j] + d->yv[j + 1] + d->yv[j + 1]) / 4.0;
        c->cvol[k] = fabs(d->xv[i + 1] - d->xv[i]) * fabs(d->yv[j + 1] - d->yv[j]);
      }
    }
  c->fc1.clear(); c->fc2.clear(); c->fvol.clear(); c->flen.clear(); c->fdir.clear();
  c->cf.assign((size_t)nc * 4, -1);
  for (int j = 0; j < c->ny; ++j)
    for (int i = 0; i < c->nx; ++i) {
      const int k = j * c->nx + i;
      auto add = [&](int o, double vol, int slot_here, int slot_other) {
        const int f = (int)c->fc1.size();
        c->fc1.push_back(k);
        c->fc2.push_back(o);
        c->fvol.push_back(vol);
        const double dx = c->cx[k] - c->cx[o], dy = c->dim > 1 ? c->cy[k] - c->cy[o] : 0.0;
        c->flen.push_back(sqrt(dx * dx + dy * dy));
        c->cf[(size_t)k * 4 + slot_here] = f;
        c->cf[(size_t)o * 4 + slot_other] = f;
        c->fdir.push_back(kSlotDir[slot_here]);
      };
      if (i + 1 < c->nx) add(k + 1, c->dim == 1 ? 1.0 : fabs(d->yv[j + 1] - d->yv[j]), 2, 1);
      if (c->dim > 1 && j + 1 < c->ny) add(k + c->nx, fabs(d->xv[i + 1] - d->xv[i]), 3, 0);
    }
  c->nf = (int)c->fc1.size();
  return 0;
}

int build_energies_orders_mapping(shefill_ctx *c) {
  const shefill_config &cf = c->cfg;
  // setup_energies (she/setup_energies.hpp:61-153)
  double maxp = -1.7976931348623157e308, minp = 1.7976931348623157e308;
  for (double p : c->psi) {
    p = p + 0.0;
    if (p < minp) minp = p;
    if (p > maxp) maxp = p;
  }
  const double range = kQ * (maxp - minp) + 2.0 * cf.energy_spacing + cf.min_kin_range;
  const double edge = cf.carrier == 0 ? -kQ * maxp - cf.energy_spacing + band_gap() / 2.0
                                      : -kQ * minp + cf.energy_spacing - band_gap() / 2.0;
  const double inc = (cf.carrier == 0 ? 1.0 : -1.0) * cf.energy_spacing;
  const size_t NH = (size_t)(range / cf.energy_spacing) + 1;
  c->NH = (int)NH;
  c->Hval.resize(NH);
  for (size_t h = 0; h < NH; ++h) c->Hval[h] = edge + h * inc;
  c->shc.resize(c->nc);
  for (int k = 0; k < c->nc; ++k)
    c->shc[k] = cf.carrier == 0 ? -kQ * c->psi[k] + band_gap() / 2.0 : -kQ * c->psi[k] - band_gap() / 2.0;
  c->shf.resize(c->nf);
  for (int f = 0; f < c->nf; ++f) c->shf[f] = (c->shc[c->fc1[f]] + c->shc[c->fc2[f]]) / 2.0;
  // expansion orders (she/expansion_order.hpp:262-363), non-adaptive
  const int L = cf.L_max;
  c->ordc.assign((size_t)c->nc * NH, L);
  c->ordf.assign((size_t)c->nf * NH, L);
  for (int k = 0; k < c->nc; ++k) {
    for (int H = 1; H < (int)NH - 1; ++H) {
      if (kin_c(c, k, H) <= 0) { c->ordc[(size_t)k * NH + H] = 1; continue; }
      bool allpos = true;
      for (int s = 0; s < 4; ++s) {
        const int f = c->cf[(size_t)k * 4 + s];
        if (f < 0) continue;
        const int o = c->fc1[f] == k ? c->fc2[f] : c->fc1[f];
        if (kin_c(c, o, H - 1) <= 0) allpos = false;
      }
      if (allpos) break;
      c->ordc[(size_t)k * NH + H] = 1;
      for (int s = 0; s < 4; ++s) {
        const int f = c->cf[(size_t)k * 4 + s];
        if (f < 0) continue;
        const int o = c->fc1[f] == k ? c->fc2[f] : c->fc1[f];
        c->ordc[(size_t)o * NH + H] = 1;
      }
    }
  }
  for (int it = 0; it < 10; ++it) {  // smooth_expansion_order: lowers the first cell of a facet
    bool changed = false;
    for (int f = 0; f < c->nf; ++f)
      for (size_t H = 0; H < NH; ++H) {
        int &o1 = c->ordc[(size_t)c->fc1[f] * NH + H];
        const int o2 = c->ordc[(size_t)c->fc2[f] * NH + H];
        if (o1 > o2 + 2) { changed = true; o1 = o2 + 2; }
      }
    if (!changed) break;
  }
  // mapping (mapping.hpp:97-329): energy-major, then odd unknowns on facets
  c->idxc.assign((size_t)c->nc * NH, -1);
  c->idxf.assign((size_t)c->nf * NH, -1);
  // cells of each energy level in mesh order, or from both ends of a 1D
  // device toward the middle
  std::vector<int> cells(c->nc);
  for (int k = 0; k < c->nc; ++k) cells[k] = k;
  if (cf.ordering == 1 && c->dim == 1)
    for (int k = 0; k < c->nc; ++k) cells[k] = (k % 2 == 0) ? k / 2 : c->nc - 1 - k / 2;
  // which (cell, level) and (facet, level) pairs carry unknowns, evaluated on
  // host threads; the numbering below is sequential
  std::vector<char> livec((size_t)c->nc * NH, 0), livef((size_t)c->nf * NH, 0);
  {
    const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    auto par = [&](int64_t m, auto fn) {
      const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthr, m));
      std::vector<std::thread> th;
      for (int t = 1; t < nt; ++t) th.emplace_back([&, t] { fn(m * t / nt, m * (t + 1) / nt); });
      fn(0, m / nt);
      for (auto &x : th) x.join();
    };
    par(c->nc, [&](int64_t k0, int64_t k1) {
      for (int64_t k = k0; k < k1; ++k)
        for (size_t H = 1; H + 1 < NH; ++H)
          livec[(size_t)k * NH + H] = avg_dos_host(c, false, (int)k, (int)H) > 0 && kin_c(c, (int)k, (int)H) < cf.max_kin_range;
    });
    par(c->nf, [&](int64_t f0, int64_t f1) {
      for (int64_t f = f0; f < f1; ++f)
        for (size_t H = 0; H < NH; ++H) livef[(size_t)f * NH + H] = avg_dos_host(c, true, (int)f, (int)H) > 0;
    });
  }
  int64_t u = 0;
  for (size_t H = 0; H < NH; ++H)
    for (int kk = 0; kk < c->nc; ++kk) {
      const int k = cells[kk];
      const size_t q = (size_t)k * NH + H;
      if (H == 0 || H == NH - 1) { c->idxc[q] = -1; continue; }
      if (livec[q]) {
        c->idxc[q] = (int)u;
        u += even_count(c->ordc[q]);
      } else {
        c->idxc[q] = -1;
        c->ordc[q] = 0;
      }
    }
  if (u > 0x7fffffff) { ferr("more than 2^31 reduced unknowns"); return -3; }
  c->n = u;
  int64_t uo = 0;
  for (size_t H = 0; H < NH; ++H)
    for (int f = 0; f < c->nf; ++f) {
      const size_t q = (size_t)f * NH + H;
      if (c->idxc[(size_t)c->fc1[f] * NH + H] > -1 && c->idxc[(size_t)c->fc2[f] * NH + H] > -1 && livef[q]) {
        c->idxf[q] = 1;
        uo += odd_count(c->ordf[q]);
      } else {
        c->idxf[q] = -1;
        c->ordf[q] = 0;
      }
    }
  c->n_odd = uo;
  return 0;
}

// Pattern of one eliminated block: P[r] = sorted r' with some odd o linking them.
using Pattern = std::vector<std::vector<int>>;

struct Builder {
  shefill_ctx *c;
  std::vector<CMat> Ae, Be, Ao, Bo;  // [4]
  std::vector<std::vector<int>> ev, od;
  std::map<std::tuple<int, int, int, int, int>, Pattern> cache;

  explicit Builder(shefill_ctx *cc) : c(cc) {}

  int tables() {
    const int L = c->cfg.L_max;
    const int K = (L + 1) * (L + 1);
    c->K = K;
    CMat a[3] = {CMat(K), CMat(K), CMat(K)}, b[3] = {CMat(K), CMat(K), CMat(K)};
    bool found = false;
    for (const auto &t : she_tables::tables) {
      if (t.L != L) continue;
      found = true;
      CMat &m = t.which < 3 ? a[t.which] : b[t.which - 3];
      for (int q = 0; q < t.n; ++q) {
        m.p[t.e[q].r * K + t.e[q].c] = 1;
        m.v[t.e[q].r * K + t.e[q].c] = t.e[q].v;
      }
    }
    if (!found) { ferr("expansion order %d has no precomputed coupling table", L); return 2; }
    for (int w = 0; w < 2; ++w)
      for (int q = 1; q <= 3 && q < K; ++q) c->a_l01[w][q - 1] = a[w].v[q];
    CMat aT[3] = {transpose(a[0]), transpose(a[1]), transpose(a[2])};
    CMat bT[3] = {transpose(b[0]), transpose(b[1]), transpose(b[2])};
    Ae.clear(); Be.clear(); Ao.clear(); Bo.clear();
    for (int d = 0; d < 4; ++d) {
      Ae.push_back(in_direction(a, d, c->dim, c->cfg.carrier));
      Be.push_back(in_direction(b, d, c->dim, c->cfg.carrier));
      Ao.push_back(in_direction(aT, d, c->dim, c->cfg.carrier));
      Bo.push_back(in_direction(bT, d, c->dim, c->cfg.carrier));
    }
    ev.assign(L + 1, {});
    od.assign(L + 1, {});
    for (int l = 0; l <= L; ++l) { ev[l] = harmonics(l, false); od[l] = harmonics(l, true); }
    return 0;
  }

  const Pattern &elim(int d, int La, int Lf, int Lb, int dd) {
    auto key = std::make_tuple(d, La, Lf, Lb, dd);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    const int K = c->K;
    Pattern P(ev[La].size());
    for (size_t r = 0; r < ev[La].size(); ++r)
      for (size_t q = 0; q < ev[Lb].size(); ++q) {
        bool any = false;
        for (int ho : od[Lf])
          if (Ae[d].v[ev[La][r] * K + ho] != 0.0 && Ao[dd].v[ho * K + ev[Lb][q]] != 0.0) { any = true; break; }
        if (any) P[r].push_back((int)q);
      }
    return cache.emplace(key, std::move(P)).first->second;
  }
};

// SHESOLVE_VERBOSE: durations of the build phases on stderr
struct FillClock {
  bool on = getenv("SHESOLVE_VERBOSE") != nullptr;
  std::chrono::steady_clock::time_point tp = std::chrono::steady_clock::now();
  void operator()(const char *what) {
    if (!on) return;
    const auto t = std::chrono::steady_clock::now();
    fprintf(stderr, "[shefill] build %s: %.1f ms\n", what, std::chrono::duration<double, std::milli>(t - tp).count());
    tp = t;
  }
};

int build_nodes_and_pattern(shefill_ctx *c, Builder &B) {
  FillClock phase;
  const size_t NH = c->NH;
  const double hw = optical_phonon_energy(c->cfg.energy_spacing);
  if (c->cfg.optical && !(hw > 0)) { ferr("optical phonon energy snaps to zero on this grid"); return 2; }
  // (facet and cell nodes are numbered energy level by energy level, in mesh
  // order inside a level: per-level counts, prefix sums, then every level
  // filled on host threads; the node data after that, node by node in parallel)
  const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto par = [&](int64_t m, auto fn) {
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(nthr, m));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back([&, t] { fn(m * t / nt, m * (t + 1) / nt); });
    fn(0, m / nt);
    for (auto &x : th) x.join();
  };
  // facet nodes
  c->fnode_of.assign((size_t)c->nf * NH, -1);
  std::vector<int> fcount(NH + 1, 0);
  par((int64_t)NH, [&](int64_t h0, int64_t h1) {
    for (int64_t H = h0; H < h1; ++H) {
      int cnt = 0;
      for (int f = 0; f < c->nf; ++f) cnt += c->idxf[(size_t)f * NH + H] >= 0;
      fcount[H + 1] = cnt;
    }
  });
  for (size_t H = 0; H < NH; ++H) fcount[H + 1] += fcount[H];
  const int nfn_all = fcount[NH];
  c->fn_f.assign(nfn_all, 0);
  c->fn_H.assign(nfn_all, 0);
  c->fn_L.assign(nfn_all, 0);
  c->fn_out.assign((size_t)2 * nfn_all, -1);
  par((int64_t)NH, [&](int64_t h0, int64_t h1) {
    for (int64_t H = h0; H < h1; ++H) {
      int w = fcount[H];
      for (int f = 0; f < c->nf; ++f) {
        const size_t q = (size_t)f * NH + H;
        if (c->idxf[q] < 0) continue;
        c->fnode_of[q] = w;
        c->fn_f[w] = f;
        c->fn_H[w] = (int)H;
        c->fn_L[w] = c->ordf[q];
        const double E = kin_f(c, f, (int)H);
        for (int s2 = 0; s2 < 2; ++s2) {
          int Hf = -1;
          if (c->cfg.optical) {
            long h = energy_index_host(c, true, f, (long)H, s2 == 0 ? E - hw : E + hw);
            if (h >= 0 && c->idxf[(size_t)f * NH + h] >= 0) Hf = (int)h;
          }
          c->fn_out[(size_t)2 * w + s2] = Hf;
        }
        ++w;
      }
    }
  });
  c->fn_base.assign(c->fn_f.size(), 0);
  {
    int acc = 0;
    for (size_t k = 0; k < c->fn_f.size(); ++k) {
      c->fn_base[k] = acc;
      acc += odd_count(c->fn_L[k]);
    }
  }
  // cell nodes in unknown order
  c->node_of.assign((size_t)c->nc * NH, -1);
  std::vector<int> ccount(NH + 1, 0);
  par((int64_t)NH, [&](int64_t h0, int64_t h1) {
    for (int64_t H = h0; H < h1; ++H) {
      int cnt = 0;
      for (int k = 0; k < c->nc; ++k) cnt += c->idxc[(size_t)k * NH + H] >= 0;
      ccount[H + 1] = cnt;
    }
  });
  for (size_t H = 0; H < NH; ++H) ccount[H + 1] += ccount[H];
  const int nn = ccount[NH];
  std::vector<int> order(nn);
  par((int64_t)NH, [&](int64_t h0, int64_t h1) {
    for (int64_t H = h0; H < h1; ++H) {
      int w = ccount[H];
      for (int k = 0; k < c->nc; ++k)
        if (c->idxc[(size_t)k * NH + H] >= 0) {
          c->node_of[(size_t)k * NH + H] = w;
          order[w++] = k * (int)NH + (int)H;
        }
    }
  });
  c->nd_c.resize(nn); c->nd_H.resize(nn); c->nd_base.resize(nn); c->nd_L.resize(nn);
  c->nd_fn.assign((size_t)nn * 4, -1); c->nd_nb.assign((size_t)nn * 4, -1);
  c->nd_inn.assign((size_t)nn * 2, -1); c->nd_in.assign((size_t)nn * 2, -1); c->nd_out.assign((size_t)nn * 2, -1);
  par(nn, [&](int64_t k0, int64_t k1) {
    for (int k = (int)k0; k < (int)k1; ++k) {
      const int cell = order[k] / (int)NH, H = order[k] % (int)NH;
      const size_t q = (size_t)cell * NH + H;
      c->nd_c[k] = cell;
      c->nd_H[k] = H;
      c->nd_base[k] = c->idxc[q];
      c->nd_L[k] = c->ordc[q];
      for (int s2 = 0; s2 < 4; ++s2) {
        const int f = c->cf[(size_t)cell * 4 + s2];
        if (f < 0) continue;
        const int fn = c->fnode_of[(size_t)f * NH + H];
        if (fn < 0) continue;
        const int o = c->fc1[f] == cell ? c->fc2[f] : c->fc1[f];
        c->nd_fn[(size_t)k * 4 + s2] = fn;
        c->nd_nb[(size_t)k * 4 + s2] = c->node_of[(size_t)o * NH + H];
      }
      if (c->cfg.optical) {
        const double E = kin_c(c, cell, H);
        // out: to E - hw (slot 0) and E + hw (slot 1); in: from E - hw (slot 0) and E + hw (slot 1)
        for (int s2 = 0; s2 < 2; ++s2) {
          long h = energy_index_host(c, false, cell, H, s2 == 0 ? E - hw : E + hw);
          if (h >= 0 && c->idxc[(size_t)cell * NH + h] >= 0) {
            c->nd_out[2 * k + s2] = (int)h;
            c->nd_in[2 * k + s2] = (int)h;
          }
        }
      }
    }
  });
  // in-scattering partner nodes (need node ids of other energies: second pass)
  par(nn, [&](int64_t k0, int64_t k1) {
    for (int k = (int)k0; k < (int)k1; ++k)
      for (int s2 = 0; s2 < 2; ++s2)
        if (c->nd_in[2 * k + s2] >= 0)
          c->nd_inn[2 * k + s2] = c->node_of[(size_t)c->nd_c[k] * NH + c->nd_in[2 * k + s2]];
  });
  phase("  nodes");
  // rows: pattern per node
  const int64_t n = c->n;
  c->rp.assign(n + 1, 0);
  c->row_node.assign(n, 0);
  c->diag_pos.assign(n, 0);
  c->diag_struct.assign(n, 0);
  // row lists of one node, computed twice (count, fill) in parallel over nodes
  // elimination patterns by direct index once the cache holds every key this
  // device uses (a std::map lookup per facet and node cost as much as the rows)
  const int LM = c->cfg.L_max + 1;
  std::vector<const Pattern *> etab;
  auto elim = [&](int d, int La, int Lf, int Lb, int dd) -> const Pattern & {
    if (!etab.empty() && d >= 0 && d < 4 && dd >= 0 && dd < 4 && La >= 0 && La < LM && Lf >= 0 && Lf < LM &&
        Lb >= 0 && Lb < LM) {
      const Pattern *q = etab[(size_t)(((d * LM + La) * LM + Lf) * LM + Lb) * 4 + dd];
      if (q) return *q;
    }
    return B.elim(d, La, Lf, Lb, dd);
  };
  // (cols keeps its row vectors between nodes, so their storage is reused)
  auto node_rows = [&](int k, std::vector<std::vector<int>> &cols, std::vector<char> &dstruct) -> int {
    const int cell = c->nd_c[k], Lc = c->nd_L[k];
    const int e = (int)B.ev[Lc].size();
    if ((int)cols.size() < e) cols.resize(e);
    for (int r = 0; r < e; ++r) cols[r].clear();
    dstruct.assign(e, 0);
    // blocks in ascending column order: lower partner, bottom, left, self, right, top, upper partner
    struct Blk { int base; int kind; int slot; };
    Blk blks[7];
    int nb = 0;
    if (c->nd_inn[2 * k] >= 0) blks[nb++] = {c->nd_base[c->nd_inn[2 * k]], 1, 0};
    for (int s = 0; s < 2; ++s)
      if (c->nd_fn[(size_t)k * 4 + s] >= 0) blks[nb++] = {c->nd_base[c->nd_nb[(size_t)k * 4 + s]], 2, s};
    blks[nb++] = {c->nd_base[k], 0, -1};
    for (int s = 2; s < 4; ++s)
      if (c->nd_fn[(size_t)k * 4 + s] >= 0) blks[nb++] = {c->nd_base[c->nd_nb[(size_t)k * 4 + s]], 2, s};
    if (c->nd_inn[2 * k + 1] >= 0) blks[nb++] = {c->nd_base[c->nd_inn[2 * k + 1]], 1, 1};
    // (already ascending in mesh-order numbering; the two-sided numbering of
    // 1D cells reorders the left and right neighbours)
    std::sort(blks, blks + nb, [](const Blk &a, const Blk &b) { return a.base < b.base; });
    // self block pattern
    thread_local std::vector<char> selfv;
    selfv.assign((size_t)e * e, 0);
    auto selfm = [&](int r, int q) -> char & { return selfv[(size_t)r * e + q]; };
    const bool optical_out = c->nd_out[2 * k] >= 0 || c->nd_out[2 * k + 1] >= 0;
    for (int r = 0; r < e; ++r) {
      if (c->contact[cell] || optical_out) selfm(r, r) = 1;
      if (c->cfg.acoustic && r > 0) selfm(r, r) = 1;
    }
    for (int s = 0; s < 4; ++s) {
      const int fn = c->nd_fn[(size_t)k * 4 + s];
      if (fn < 0) continue;
      const int d = kSlotDir[s];
      const Pattern &P = elim(d, Lc, c->fn_L[fn], Lc, rev_dir(d));
      for (int r = 0; r < e; ++r)
        for (int q : P[r]) selfm(r, q) = 1;
    }
    for (int bi = 0; bi < nb; ++bi) {
      const Blk &bk = blks[bi];
      if (bk.kind == 1) {
        cols[0].push_back(bk.base);
      } else if (bk.kind == 0) {
        for (int r = 0; r < e; ++r) {
          dstruct[r] = selfm(r, r);
          for (int q = 0; q < e; ++q)
            if (selfm(r, q) || q == r) cols[r].push_back(bk.base + q);
        }
      } else {
        const int s = bk.slot;
        const int fn = c->nd_fn[(size_t)k * 4 + s], nbn = c->nd_nb[(size_t)k * 4 + s];
        const int d = kSlotDir[s];
        const Pattern &P = elim(d, Lc, c->fn_L[fn], c->nd_L[nbn], d);
        for (int r = 0; r < e; ++r)
          for (int q : P[r]) cols[r].push_back(bk.base + q);
      }
    }
    return e;
  };
  // pattern cache is filled serially first (all key combinations of this device)
  {
    std::vector<std::vector<int>> cols;
    std::vector<char> ds;
    for (int k = 0; k < nn; ++k) {
      bool needed = false;
      for (int s = 0; s < 4 && !needed; ++s) {
        const int fn = c->nd_fn[(size_t)k * 4 + s];
        if (fn < 0) continue;
        const int d = kSlotDir[s];
        needed = !B.cache.count(std::make_tuple(d, c->nd_L[k], c->fn_L[fn], c->nd_L[k], rev_dir(d))) ||
                 !B.cache.count(std::make_tuple(d, c->nd_L[k], c->fn_L[fn], c->nd_L[c->nd_nb[(size_t)k * 4 + s]], d));
      }
      if (needed) node_rows(k, cols, ds);
    }
  }
  etab.assign((size_t)4 * LM * LM * LM * 4, nullptr);
  for (const auto &kv : B.cache) {
    const int d = std::get<0>(kv.first), La = std::get<1>(kv.first), Lf = std::get<2>(kv.first),
              Lb = std::get<3>(kv.first), dd = std::get<4>(kv.first);
    if (d >= 0 && d < 4 && dd >= 0 && dd < 4 && La >= 0 && La < LM && Lf >= 0 && Lf < LM && Lb >= 0 && Lb < LM)
      etab[(size_t)(((d * LM + La) * LM + Lf) * LM + Lb) * 4 + dd] = &kv.second;
  }
  const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto parallel = [&](auto fn) {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t)
      th.emplace_back([&, t] {
        std::vector<std::vector<int>> cols;
        std::vector<char> ds;
        for (int k = (int)((int64_t)nn * t / nthreads); k < (int)((int64_t)nn * (t + 1) / nthreads); ++k)
          fn(k, cols, ds);
      });
    for (auto &x : th) x.join();
  };
  parallel([&](int k, std::vector<std::vector<int>> &cols, std::vector<char> &ds) {
    const int e = node_rows(k, cols, ds);
    const int base = c->nd_base[k];
    for (int r = 0; r < e; ++r) {
      c->rp[base + r + 1] = (int)cols[r].size();
      c->row_node[base + r] = k;
      c->diag_struct[base + r] = ds[r];
    }
  });
  for (int64_t i = 0; i < n; ++i) {
    if ((int64_t)c->rp[i] + c->rp[i + 1] > 0x7fffffffLL) { ferr("more than 2^31 nonzeros"); return -3; }
    c->rp[i + 1] += c->rp[i];
  }
  phase("  row lengths");
  c->nnz = c->rp[n];
  c->ci.assign(c->nnz, 0);
  parallel([&](int k, std::vector<std::vector<int>> &cols, std::vector<char> &ds) {
    const int e = node_rows(k, cols, ds);
    const int base = c->nd_base[k];
    for (int r = 0; r < e; ++r) {
      int p = c->rp[base + r];
      for (int col : cols[r]) {
        if (col == base + (int)r) c->diag_pos[base + r] = p;
        c->ci[p++] = col;
      }
    }
  });
  phase("  columns");
  // block ranges (she/linear_solver.hpp:42-83)
  c->blocks.assign((size_t)2 * NH, 0);
  int64_t last = 0;
  for (size_t H = 0; H < NH; ++H) {
    c->blocks[2 * H] = last;
    for (int k = 0; k < c->nc; ++k) {
      const int u = c->idxc[(size_t)k * NH + H];
      if (u > -1) {
        last = u;
        c->blocks[2 * H] = last;
        if (H > 0) c->blocks[2 * (H - 1) + 1] = last;
        break;
      }
    }
    if (c->blocks[2 * H + 1] < c->blocks[2 * H]) c->blocks[2 * H + 1] = c->blocks[2 * H];
  }
  c->blocks[2 * (NH - 1) + 1] = n;
  return 0;
}

int grid(int64_t n) { return (int)std::max<int64_t>(1, (n + kB - 1) / kB); }
int launch_fill(shefill_ctx *c);

// The column-ordered positions of the structural entries on the device
// (k_colcount, k_colfill, k_colsort); the prefix sums of the counts on the host.
int build_column_positions(shefill_ctx *c) {
  const int n = (int)c->n;
  if (c->dcolcount.alloc(n) || c->dcptr.alloc((size_t)n + 1)) return -11;
  FHIP(hipMemsetAsync(c->dcolcount.p, 0, sizeof(int) * std::max(n, 1), c->stream));
  if (n > 0)
    hipLaunchKernelGGL(k_colcount, dim3(grid(n)), dim3(kB), 0, c->stream, n, c->drp.p, c->dci.p, c->ddiag_pos.p,
                       c->ddiag_struct.p, c->dcolcount.p);
  std::vector<int> cnt(n), ptr((size_t)n + 1, 0);
  if (n > 0) FHIP(hipMemcpyAsync(cnt.data(), c->dcolcount.p, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  FHIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + cnt[i];
  DBuf<int> fillp;
  if (c->dcpos.alloc(std::max(ptr[n], 1)) || fillp.alloc((size_t)n + 1)) return -11;
  FHIP(hipMemcpyAsync(c->dcptr.p, ptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
  FHIP(hipMemcpyAsync(fillp.p, ptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
  if (n > 0) {
    hipLaunchKernelGGL(k_colfill, dim3(grid(n)), dim3(kB), 0, c->stream, n, c->drp.p, c->dci.p, c->ddiag_pos.p,
                       c->ddiag_struct.p, fillp.p, c->dcpos.p);
    hipLaunchKernelGGL(k_colsort, dim3(grid(n)), dim3(kB), 0, c->stream, n, c->dcptr.p, c->dcpos.p);
  }
  FHIP(hipGetLastError());
  FHIP(hipStreamSynchronize(c->stream));  // ptr and fillp live until the copies and kernels are done
  return 0;
}

int fill_values(shefill_ctx *c, Builder &B) {
  const int K = c->K;
  std::vector<double> Ae(4 * K * K), Be(4 * K * K), Ao(4 * K * K), Bo(4 * K * K);
  for (int d = 0; d < 4; ++d)
    for (int q = 0; q < K * K; ++q) {
      Ae[d * K * K + q] = B.Ae[d].v[q];
      Be[d * K * K + q] = B.Be[d].v[q];
      Ao[d * K * K + q] = B.Ao[d].v[q];
      Bo[d * K * K + q] = B.Bo[d].v[q];
    }
  std::vector<int> evh, odh, evo, odo;
  for (int l = 0; l <= c->cfg.L_max; ++l) {
    evo.push_back((int)evh.size());
    odo.push_back((int)odh.size());
    for (int h : B.ev[l]) evh.push_back(h);
    for (int h : B.od[l]) odh.push_back(h);
  }
  evo.push_back((int)evh.size());
  odo.push_back((int)odh.size());
  int e = 0;
  e = e ? e : upload(c, c->dHval, c->Hval);
  e = e ? e : upload(c, c->dshc, c->shc);
  e = e ? e : upload(c, c->dshf, c->shf);
  e = e ? e : upload(c, c->dcvol, c->cvol);
  e = e ? e : upload(c, c->dfvol, c->fvol);
  e = e ? e : upload(c, c->dflen, c->flen);
  e = e ? e : upload(c, c->ddn, c->dn);
  e = e ? e : upload(c, c->ddp, c->dp);
  e = e ? e : upload(c, c->dcontact, c->contact);
  e = e ? e : upload(c, c->dfc1, c->fc1);
  e = e ? e : upload(c, c->dfc2, c->fc2);
  e = e ? e : upload(c, c->dAe, Ae);
  e = e ? e : upload(c, c->dBe, Be);
  e = e ? e : upload(c, c->dAo, Ao);
  e = e ? e : upload(c, c->dBo, Bo);
  e = e ? e : upload(c, c->dev_h, evh);
  e = e ? e : upload(c, c->dod_h, odh);
  e = e ? e : upload(c, c->dev_off, evo);
  e = e ? e : upload(c, c->dod_off, odo);
  e = e ? e : upload(c, c->dnd_c, c->nd_c);
  e = e ? e : upload(c, c->dnd_H, c->nd_H);
  e = e ? e : upload(c, c->dnd_base, c->nd_base);
  e = e ? e : upload(c, c->dnd_L, c->nd_L);
  e = e ? e : upload(c, c->dnd_fn, c->nd_fn);
  e = e ? e : upload(c, c->dnd_nb, c->nd_nb);
  e = e ? e : upload(c, c->dnd_inn, c->nd_inn);
  e = e ? e : upload(c, c->dnd_in, c->nd_in);
  e = e ? e : upload(c, c->dnd_out, c->nd_out);
  e = e ? e : upload(c, c->dfn_f, c->fn_f);
  e = e ? e : upload(c, c->dfn_H, c->fn_H);
  e = e ? e : upload(c, c->dfn_L, c->fn_L);
  e = e ? e : upload(c, c->dfn_out, c->fn_out);
  e = e ? e : upload(c, c->dfdir, c->fdir);
  e = e ? e : upload(c, c->dfn_base, c->fn_base);
  e = e ? e : upload(c, c->didxc, c->idxc);
  e = e ? e : upload(c, c->dordc, c->ordc);
  e = e ? e : upload(c, c->dfnode_of, c->fnode_of);
  e = e ? e : upload(c, c->dnode_of, c->node_of);
  e = e ? e : upload(c, c->dcx, c->cx);
  e = e ? e : upload(c, c->dcy, c->cy);
  e = e ? e : upload(c, c->drp, c->rp);
  e = e ? e : upload(c, c->dci, c->ci);
  e = e ? e : upload(c, c->drow_node, c->row_node);
  e = e ? e : upload(c, c->ddiag_pos, c->diag_pos);
  e = e ? e : upload(c, c->ddiag_struct, c->diag_struct);
  e = e ? e : build_column_positions(c);
  const int nn = (int)c->nd_c.size(), nfn = (int)c->fn_f.size();
  const int64_t n = c->n;
  e = e ? e : c->dcorr.alloc(c->nc);
  e = e ? e : c->dncont.alloc(c->nc);
  e = e ? e : c->dcdiff.alloc(nfn);
  e = e ? e : c->dcdrift.alloc(2 * (size_t)nfn);
  e = e ? e : c->dodiag.alloc(nfn);
  e = e ? e : c->ddiag0.alloc(nn);
  e = e ? e : c->ddiagr.alloc(nn);
  e = e ? e : c->drhs0.alloc(nn);
  e = e ? e : c->dinco.alloc(2 * (size_t)nn);
  e = e ? e : c->dval.alloc(c->nnz);
  e = e ? e : c->drhs.alloc(n);
  e = e ? e : c->dscal.alloc(n);
  e = e ? e : c->dcolfac.alloc(n);
  e = e ? e : c->dlog.alloc(n);
  e = e ? e : c->drp64.alloc(n + 1);
  if (e) return e;

  DevTables t{};
  t.K = K;
  t.NH = c->NH;
  t.carrier = c->cfg.carrier;
  t.dim = c->dim;
  t.T = c->T;
  t.kT = kKB * c->T;
  t.hw = optical_phonon_energy(c->cfg.energy_spacing);
  t.rate_ac = c->cfg.acoustic ? acoustic_rate(c->cfg.carrier, c->T) : 0.0;
  t.rate_op = c->cfg.optical ? optical_rate(t.hw) : 0.0;
  t.Nop = optical_phonon_number(t.hw, c->T);
  t.tau = c->cfg.tau;
  t.spacing = c->cfg.energy_spacing;
  t.Hval = c->dHval.p; t.shc = c->dshc.p; t.shf = c->dshf.p; t.cvol = c->dcvol.p; t.fvol = c->dfvol.p;
  t.flen = c->dflen.p; t.dn = c->ddn.p; t.dp = c->ddp.p; t.contact = c->dcontact.p; t.fc1 = c->dfc1.p;
  t.fc2 = c->dfc2.p; t.Ae = c->dAe.p; t.Be = c->dBe.p; t.Ao = c->dAo.p; t.Bo = c->dBo.p;
  t.ev_h = c->dev_h.p; t.od_h = c->dod_h.p; t.ev_off = c->dev_off.p; t.od_off = c->dod_off.p;

  c->t = t;
  c->built = true;
  return launch_fill(c);
}

int launch_fill(shefill_ctx *c) {
  const DevTables &t = c->t;
  const int nn = (int)c->nd_c.size(), nfn = (int)c->fn_f.size();
  const int64_t n = c->n;
  FHIP(hipEventRecord(c->e0, c->stream));
  hipLaunchKernelGGL(k_bnd_corr, dim3(grid(c->nc)), dim3(kB), 0, c->stream, t, c->nc, c->dcorr.p, c->dncont.p);
  if (nfn > 0)
    hipLaunchKernelGGL(k_fnode, dim3(grid(nfn)), dim3(kB), 0, c->stream, t, nfn, c->dfn_f.p, c->dfn_H.p, c->dfn_out.p,
                       c->dcdiff.p, c->dcdrift.p, c->dodiag.p);
  if (nn > 0)
    hipLaunchKernelGGL(k_node, dim3(grid(nn)), dim3(kB), 0, c->stream, t, nn, c->dnd_c.p, c->dnd_H.p, c->dnd_out.p,
                       c->dnd_in.p, c->dcorr.p, c->dncont.p, c->ddiag0.p, c->ddiagr.p, c->drhs0.p, c->dinco.p);
  RowArgs a{};
  a.n = (int)n; a.rp = c->drp.p; a.ci = c->dci.p; a.row_node = c->drow_node.p; a.nd_c = c->dnd_c.p;
  a.nd_H = c->dnd_H.p; a.nd_base = c->dnd_base.p; a.nd_L = c->dnd_L.p; a.nd_fn = c->dnd_fn.p; a.nd_nb = c->dnd_nb.p;
  a.nd_inn = c->dnd_inn.p; a.fn_f = c->dfn_f.p; a.fn_L = c->dfn_L.p; a.cdiff = c->dcdiff.p; a.cdrift = c->dcdrift.p;
  a.odiag = c->dodiag.p; a.diag0 = c->ddiag0.p; a.diagr = c->ddiagr.p; a.rhs0 = c->drhs0.p; a.inco = c->dinco.p;
  a.val = c->dval.p; a.rhs = c->drhs.p;
  if (n > 0) {
    const size_t lds = t.K <= kRowsLdsK ? sizeof(double) * 16 * t.K * t.K : 0;
    hipLaunchKernelGGL(k_rows, dim3(grid(n)), dim3(kB), lds, c->stream, t, a);
    hipLaunchKernelGGL(k_node_scaling, dim3(grid(n)), dim3(kB), 0, c->stream, (int)n, c->drow_node.p, c->dnd_c.p,
                       c->dnd_H.p, t, c->cfg.scale, c->dscal.p);
    hipLaunchKernelGGL(k_scale, dim3(std::min(grid(c->nnz), 32768)), dim3(kB), 0, c->stream, (int64_t)c->nnz,
                       c->dci.p, c->dscal.p, c->dval.p);
    hipLaunchKernelGGL(k_collog, dim3(grid(n)), dim3(kB), 0, c->stream, (int)n, c->dcptr.p, c->dcpos.p, c->dval.p,
                       c->dlog.p);
    hipLaunchKernelGGL(k_colnorm, dim3(grid(n)), dim3(kB), 0, c->stream, (int)n, c->dcolcount.p, c->ddiag_pos.p,
                       c->dval.p, c->dlog.p, c->dcolfac.p);
    hipLaunchKernelGGL(k_rownorm, dim3(grid((int64_t)n * kRowLanes)), dim3(kB), 0, c->stream, (int)n, c->drp.p,
                       c->dci.p, c->ddiag_pos.p, c->dlog.p, c->dval.p, c->drhs.p);
    hipLaunchKernelGGL(k_rp64, dim3(grid(n + 1)), dim3(kB), 0, c->stream, (int)n, c->drp.p, c->drp64.p);
  }
  FHIP(hipGetLastError());
  FHIP(hipEventRecord(c->e1, c->stream));
  FHIP(hipEventSynchronize(c->e1));
  float ms = 0;
  FHIP(hipEventElapsedTime(&ms, c->e0, c->e1));
  c->fill_ms = ms;
  return 0;
}

// Address of this device's scratch, looked up once per device with that
// device current (shefill_create); nullptr until then.
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;

int noalloc_scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return 0;
  double *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    ferr("hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
    return -10;
  }
  g_scratch_ptr[device] = p;
  return 0;
}

double *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

}  // namespace

extern "C" {

int shefill_create(shefill_handle *h, int device) {
  if (!h) { ferr("handle pointer is NULL"); return 1; }
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) { ferr("no HIP device available"); return -2; }
  if (device < 0 || device >= cnt) { ferr("device %d out of range", device); return 2; }
  FHIP(hipSetDevice(device));
  if (int e = noalloc_scratch_setup(device)) return e;
  auto *c = new shefill_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&c->e0) != hipSuccess ||
      hipEventCreate(&c->e1) != hipSuccess) {
    ferr("stream/event creation failed");
    delete c;
    return -10;
  }
  *h = c;
  return 0;
}

int shefill_destroy(shefill_handle h) {
  if (!h) return 1;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  (void)hipEventDestroy(h->e0);
  (void)hipEventDestroy(h->e1);
  (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

const char *shefill_last_error(void) { return g_ferr.c_str(); }

int shefill_build(shefill_handle h, const shefill_device *d, const shefill_config *cfg) {
  if (!h) return 1;
  if (!d || !d->xv || !d->doping_n || !d->doping_p || !d->potential || !d->contact) { ferr("device arrays missing"); return 2; }
  if (!cfg) return 3;
  if (d->dim != 1 && d->dim != 2) { ferr("dim must be 1 or 2"); return 2; }
  if (d->nx < 2 || (d->dim == 2 && (d->ny < 2 || !d->yv))) { ferr("mesh too small"); return 2; }
  if (!(cfg->energy_spacing > 0) || !(cfg->tau > 0)) { ferr("energy spacing and tau must be positive"); return 3; }
  if (cfg->carrier != 0 && cfg->carrier != 1) { ferr("carrier must be 0 or 1"); return 3; }
  if (cfg->ordering != 0 && cfg->ordering != 1) { ferr("ordering must be 0 or 1"); return 3; }
  FHIP(hipSetDevice(h->device));
  // keep what a transfer of this build's solution to the next build needs
  h->prev_nc = -1;
  if (h->built && h->nc == d->nx * (d->dim == 2 ? d->ny : 1)) {
    const size_t m = (size_t)h->nc * h->NH;
    if (h->dprev_H.alloc(h->NH) || h->dprev_shc.alloc(h->nc) || h->dprev_idx.alloc(m) || h->dprev_ord.alloc(m))
      return -11;
    FHIP(hipMemcpyAsync(h->dprev_H.p, h->dHval.p, sizeof(double) * h->NH, hipMemcpyDeviceToDevice, h->stream));
    FHIP(hipMemcpyAsync(h->dprev_shc.p, h->dshc.p, sizeof(double) * h->nc, hipMemcpyDeviceToDevice, h->stream));
    FHIP(hipMemcpyAsync(h->dprev_idx.p, h->didxc.p, sizeof(int) * m, hipMemcpyDeviceToDevice, h->stream));
    FHIP(hipMemcpyAsync(h->dprev_ord.p, h->dordc.p, sizeof(int) * m, hipMemcpyDeviceToDevice, h->stream));
    FHIP(hipStreamSynchronize(h->stream));
    h->prev_nc = h->nc;
    h->prev_NH = h->NH;
  }
  h->built = false;
  h->cfg = *cfg;
  FillClock phase;
  if (int e = build_geometry(h, d)) return e;
  phase("geometry");
  if (int e = build_energies_orders_mapping(h)) return e;
  phase("energies and mapping");
  Builder B(h);
  if (int e = B.tables()) return e == 2 ? 3 : e;
  phase("tables");
  if (int e = build_nodes_and_pattern(h, B)) return e;
  phase("nodes and pattern");
  const int e = fill_values(h, B);
  if (phase.on) hipStreamSynchronize(h->stream);
  phase("uploads and fill");
  return e;
}

int shefill_refill(shefill_handle h) {
  if (!h) return 1;
  if (!h->built) { ferr("nothing built"); return -1; }
  FHIP(hipSetDevice(h->device));
  return launch_fill(h);
}

int shefill_sizes(shefill_handle h, int64_t *o) {
  if (!h) return 1;
  if (!o) return 2;
  o[0] = h->NH; o[1] = h->n; o[2] = h->nnz; o[3] = h->n_odd; o[4] = (int64_t)h->nd_c.size(); o[5] = h->nc; o[6] = h->nf;
  return 0;
}

int shefill_pattern_hash(shefill_handle h, uint64_t *out) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!out) { ferr("out is NULL"); return 2; }
  if (!h->built) { ferr("nothing built"); return -1; }
  // FNV-1a over 32-bit words: n, then the row pointer and the column indices
  // cut into 16 slices hashed in parallel, the slice hashes mixed in order
  constexpr uint64_t kBasis = 1469598103934665603ULL, kPrime = 1099511628211ULL;
  auto words = [&](const std::vector<int> &v, size_t i0, size_t i1) {
    uint64_t x = kBasis;
    for (size_t i = i0; i < i1; ++i) {
      x ^= (uint32_t)v[i];
      x *= kPrime;
    }
    return x;
  };
  constexpr int kSlices = 16;
  uint64_t part[2][kSlices];
  std::vector<std::thread> th;
  for (int t = 0; t < kSlices; ++t)
    th.emplace_back([&, t] {
      part[0][t] = words(h->rp, h->rp.size() * t / kSlices, h->rp.size() * (t + 1) / kSlices);
      part[1][t] = words(h->ci, h->ci.size() * t / kSlices, h->ci.size() * (t + 1) / kSlices);
    });
  for (auto &x : th) x.join();
  uint64_t x = kBasis;
  auto mix = [&](uint64_t v) {
    x ^= v;
    x *= kPrime;
  };
  mix((uint64_t)h->n);
  for (int a = 0; a < 2; ++a)
    for (int t = 0; t < kSlices; ++t) mix(part[a][t]);
  *out = x;
  return 0;
}

int shefill_device_system(shefill_handle h, const int64_t **rp, const int32_t **ci, const double **val,
                          const double **rhs) {
  if (!h) return 1;
  if (!h->dval.p) { ferr("nothing built"); return -1; }
  if (rp) *rp = h->drp64.p;
  if (ci) *ci = h->dci.p;
  if (val) *val = h->dval.p;
  if (rhs) *rhs = h->drhs.p;
  return 0;
}

int shefill_get_system(shefill_handle h, int64_t *rp, int32_t *ci, double *val, double *rhs, double *sc, double *cf) {
  if (!h) return 1;
  if (!h->dval.p) { ferr("nothing built"); return -1; }
  FHIP(hipSetDevice(h->device));
  if (rp) for (int64_t i = 0; i <= h->n; ++i) rp[i] = h->rp[i];
  if (ci) std::memcpy(ci, h->ci.data(), sizeof(int32_t) * h->nnz);
  if (val) FHIP(hipMemcpyAsync(val, h->dval.p, sizeof(double) * h->nnz, hipMemcpyDeviceToHost, h->stream));
  if (rhs) FHIP(hipMemcpyAsync(rhs, h->drhs.p, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  if (sc) FHIP(hipMemcpyAsync(sc, h->dscal.p, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  if (cf) FHIP(hipMemcpyAsync(cf, h->dcolfac.p, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  FHIP(hipStreamSynchronize(h->stream));
  return 0;
}

int shefill_block_ranges(shefill_handle h, int64_t *r) {
  if (!h) return 1;
  if (!r) return 2;
  std::memcpy(r, h->blocks.data(), sizeof(int64_t) * h->blocks.size());
  return 0;
}

int shefill_mapping(shefill_handle h, int64_t *ei, int32_t *eo) {
  if (!h) return 1;
  if (!ei) return 2;
  if (!eo) return 3;
  for (size_t q = 0; q < h->idxc.size(); ++q) {
    ei[q] = h->idxc[q];
    eo[q] = h->ordc[q];
  }
  return 0;
}

int shefill_last_fill_ms(shefill_handle h, double *ms) {
  if (!h) return 1;
  if (!ms) return 2;
  *ms = h->fill_ms;
  return 0;
}

int shefill_recover_device(shefill_handle h, const double *d_x, double *d_x_full) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!d_x) { ferr("x is NULL"); return 2; }
  if (!d_x_full) { ferr("x_full is NULL"); return 3; }
  FHIP(hipSetDevice(h->device));
  const int n = (int)h->n, nfn = (int)h->fn_f.size();
  if (n > 0)
    hipLaunchKernelGGL(k_scale_back, dim3(grid(n)), dim3(kB), 0, h->stream, n, d_x, h->dscal.p, h->dcolfac.p,
                       h->cfg.scale, d_x_full);
  RecArgs a{};
  a.n = n; a.NH = h->NH; a.nfn = nfn;
  a.fn_f = h->dfn_f.p; a.fn_H = h->dfn_H.p; a.fn_L = h->dfn_L.p; a.fn_base = h->dfn_base.p; a.fdir = h->dfdir.p;
  a.node_of = h->dnode_of.p; a.nd_base = h->dnd_base.p; a.nd_L = h->dnd_L.p;
  a.cdiff = h->dcdiff.p; a.cdrift = h->dcdrift.p; a.odiag = h->dodiag.p;
  if (nfn > 0) hipLaunchKernelGGL(k_recover_odd, dim3(grid(nfn)), dim3(kB), 0, h->stream, h->t, a, d_x_full);
  FHIP(hipGetLastError());
  FHIP(hipStreamSynchronize(h->stream));
  return 0;
}

int shefill_postproc_device(shefill_handle h, const double *d_x_full, double *d_density, double *d_current) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!d_x_full) { ferr("x_full is NULL"); return 2; }
  FHIP(hipSetDevice(h->device));
  if (d_density && h->nc > 0)
    hipLaunchKernelGGL(k_density, dim3(grid(h->nc)), dim3(kB), 0, h->stream, h->t, h->nc, h->didxc.p, d_x_full,
                       d_density);
  if (d_current && h->nf > 0) {
    CurArgs a{};
    a.nf = h->nf; a.n = (int)h->n; a.fnode_of = h->dfnode_of.p; a.fn_base = h->dfn_base.p;
    a.cx = h->dcx.p; a.cy = h->dcy.p;
    for (int q = 0; q < 3; ++q) { a.ax[q] = h->a_l01[0][q]; a.ay[q] = h->a_l01[1][q]; }
    hipLaunchKernelGGL(k_current, dim3(grid(h->nf)), dim3(kB), 0, h->stream, h->t, a, d_x_full, d_current);
  }
  FHIP(hipGetLastError());
  FHIP(hipStreamSynchronize(h->stream));
  return 0;
}

int shefill_recover(shefill_handle h, const double *x, double *x_full) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!x) { ferr("x is NULL"); return 2; }
  if (!x_full) { ferr("x_full is NULL"); return 3; }
  FHIP(hipSetDevice(h->device));
  const size_t n = (size_t)h->n, N = n + (size_t)h->n_odd;
  double *dx = nullptr, *dxf = nullptr;
  FHIP(hipMalloc(&dx, sizeof(double) * std::max<size_t>(n, 1)));
  if (hipMalloc(&dxf, sizeof(double) * std::max<size_t>(N, 1)) != hipSuccess) {
    hipFree(dx);
    ferr("out of device memory");
    return -10;
  }
  int e = 0;
  if (hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) { ferr("copy failed"); e = -10; }
  if (!e) e = shefill_recover_device(h, dx, dxf);
  if (!e && hipMemcpy(x_full, dxf, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess) {
    ferr("copy failed");
    e = -10;
  }
  hipFree(dx);
  hipFree(dxf);
  return e;
}

int shefill_postproc(shefill_handle h, const double *x_full, double *density, double *current) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!x_full) { ferr("x_full is NULL"); return 2; }
  FHIP(hipSetDevice(h->device));
  const size_t N = (size_t)h->n + (size_t)h->n_odd;
  double *dxf = nullptr, *dd = nullptr, *dc = nullptr;
  FHIP(hipMalloc(&dxf, sizeof(double) * std::max<size_t>(N, 1)));
  FHIP(hipMalloc(&dd, sizeof(double) * std::max(h->nc, 1)));
  FHIP(hipMalloc(&dc, sizeof(double) * std::max(h->nf, 1)));
  int e = 0;
  if (hipMemcpy(dxf, x_full, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess) { ferr("copy failed"); e = -10; }
  if (!e) e = shefill_postproc_device(h, dxf, density ? dd : nullptr, current ? dc : nullptr);
  if (!e && density && hipMemcpy(density, dd, sizeof(double) * h->nc, hipMemcpyDeviceToHost) != hipSuccess) e = -10;
  if (!e && current && h->nf > 0 && hipMemcpy(current, dc, sizeof(double) * h->nf, hipMemcpyDeviceToHost) != hipSuccess)
    e = -10;
  hipFree(dxf);
  hipFree(dd);
  hipFree(dc);
  return e;
}

int shefill_density_update_device(shefill_handle h, const double *d_new, double *d_density, double alpha,
                                  int logarithmic, double *update_norm) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!d_new) { ferr("new density is NULL"); return 2; }
  if (!d_density) { ferr("density is NULL"); return 3; }
  if (!(alpha > 0.0)) { ferr("damping must be positive"); return 4; }
  FHIP(hipSetDevice(h->device));
  const int nb = grid(h->nc);
  double *part = nullptr;
  FHIP(hipMalloc(&part, sizeof(double) * nb));
  hipLaunchKernelGGL(k_density_update, dim3(nb), dim3(kB), 0, h->stream, h->nc, d_new, d_density, alpha, logarithmic,
                     part);
  std::vector<double> hp(nb);
  int e = 0;
  if (hipMemcpyAsync(hp.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess) {
    ferr("density update failed");
    e = -10;
  }
  hipFree(part);
  if (!e && update_norm) {
    double t = 0.0;
    for (double v : hp) t += v;
    *update_norm = sqrt(t);
  }
  return e;
}

int shefill_transfer_device(shefill_handle h, const double *d_x_full_prev, const double *d_density_prev,
                            double *d_x) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!d_x_full_prev) { ferr("previous solution is NULL"); return 2; }
  if (!d_density_prev) { ferr("previous density is NULL"); return 3; }
  if (!d_x) { ferr("x is NULL"); return 4; }
  if (h->prev_nc != h->nc) { ferr("no previous build on this mesh"); return -1; }
  FHIP(hipSetDevice(h->device));
  const int64_t m = (int64_t)h->nc * h->NH;
  const int64_t N = h->n + h->n_odd;
  if (h->dtr_f.alloc(std::max<int64_t>(N, 1)) || h->dtr_dens.alloc(std::max(h->nc, 1))) return -11;
  TransferArgs a{};
  a.nc = h->nc; a.NH = h->NH; a.NH_old = h->prev_NH; a.carrier = h->cfg.carrier; a.L_max = h->cfg.L_max;
  a.H = h->dHval.p; a.shc = h->dshc.p; a.H_old = h->dprev_H.p; a.shc_old = h->dprev_shc.p;
  a.idx = h->didxc.p; a.ord = h->dordc.p; a.idx_old = h->dprev_idx.p; a.ord_old = h->dprev_ord.p;
  a.x_old = d_x_full_prev; a.f_new = h->dtr_f.p;
  const int g = (int)std::max<int64_t>(1, (m + kB - 1) / kB);
  hipLaunchKernelGGL(k_transfer_even, dim3(g), dim3(kB), 0, h->stream, a);
  // density of the transferred values on the new grid (the even part is all k_density reads)
  hipLaunchKernelGGL(k_density, dim3(grid(h->nc)), dim3(kB), 0, h->stream, h->t, h->nc, h->didxc.p, h->dtr_f.p,
                     h->dtr_dens.p);
  hipLaunchKernelGGL(k_transfer_scale, dim3(g), dim3(kB), 0, h->stream, h->nc, h->NH, h->didxc.p, h->dordc.p,
                     d_density_prev, h->dtr_dens.p, h->dtr_f.p, h->dscal.p, h->dcolfac.p, h->cfg.scale, d_x);
  FHIP(hipGetLastError());
  FHIP(hipStreamSynchronize(h->stream));
  return 0;
}

int shefill_density_update_device_noalloc(shefill_handle h, const double *d_new, double *d_density, double alpha,
                                          int logarithmic, double *update_norm) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!d_new) { ferr("new density is NULL"); return 2; }
  if (!d_density) { ferr("density is NULL"); return 3; }
  if (!(alpha > 0.0)) { ferr("damping must be positive"); return 4; }
  const int nb = grid(h->nc);
  if (nb > kNoAllocParts) {
    ferr("%d cells exceed the %lld of the allocation-free update", h->nc, (long long)kNoAllocParts * kB);
    return 1;
  }
  double *part = noalloc_scratch(h->device);
  if (!part) { ferr("no scratch address for device %d", h->device); return -1; }
  FHIP(hipSetDevice(h->device));
  // every partial the sum reads is written by this launch: nothing to clear
  hipLaunchKernelGGL(k_density_update, dim3(nb), dim3(kB), 0, h->stream, h->nc, d_new, d_density, alpha, logarithmic,
                     part);
  if (update_norm) {
    hipLaunchKernelGGL(k_sum_sqrt_ordered, dim3(1), dim3(64), 0, h->stream, nb, (const double *)part,
                       part + kNoAllocParts);
    FHIP(hipMemcpyAsync(update_norm, part + kNoAllocParts, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  FHIP(hipGetLastError());
  FHIP(hipStreamSynchronize(h->stream));
  return 0;
}

int shefill_density_update(shefill_handle h, const double *new_density, double *density, double alpha,
                           int logarithmic, double *update_norm) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!new_density) { ferr("new density is NULL"); return 2; }
  if (!density) { ferr("density is NULL"); return 3; }
  FHIP(hipSetDevice(h->device));
  double *dn = nullptr, *dc = nullptr;
  FHIP(hipMalloc(&dn, sizeof(double) * std::max(h->nc, 1)));
  FHIP(hipMalloc(&dc, sizeof(double) * std::max(h->nc, 1)));
  int e = 0;
  if (hipMemcpy(dn, new_density, sizeof(double) * h->nc, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dc, density, sizeof(double) * h->nc, hipMemcpyHostToDevice) != hipSuccess) {
    ferr("copy failed");
    e = -10;
  }
  if (!e) e = shefill_density_update_device(h, dn, dc, alpha, logarithmic, update_norm);
  if (!e && hipMemcpy(density, dc, sizeof(double) * h->nc, hipMemcpyDeviceToHost) != hipSuccess) e = -10;
  hipFree(dn);
  hipFree(dc);
  return e;
}

int shefill_terminal_current_device(shefill_handle h, const double *d_facet_current,
                                    const unsigned char *d_terminal_mask, double *current) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!d_facet_current && h->nf > 0) { ferr("facet current is NULL"); return 2; }
  if (!d_terminal_mask) { ferr("terminal mask is NULL"); return 3; }
  if (!current) { ferr("current is NULL"); return 4; }
  FHIP(hipSetDevice(h->device));
  *current = 0.0;
  if (h->nf == 0) return 0;
  const int nb = grid(h->nf);
  double *part = nullptr;
  FHIP(hipMalloc(&part, sizeof(double) * nb));
  hipLaunchKernelGGL(k_terminal_current, dim3(nb), dim3(kB), 0, h->stream, h->nf, h->dfc1.p, h->dfc2.p, h->dfvol.p,
                     d_terminal_mask, d_facet_current, part);
  std::vector<double> hp(nb);
  int e = 0;
  if (hipMemcpyAsync(hp.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess) {
    ferr("terminal current failed");
    e = -10;
  }
  hipFree(part);
  if (!e) {
    double t = 0.0;
    for (double v : hp) t += v;
    *current = t;
  }
  return e;
}

int shefill_terminal_current(shefill_handle h, const double *facet_current, const unsigned char *terminal_mask,
                             double *current) {
  if (!h) { ferr("handle is NULL"); return 1; }
  if (!h->built) { ferr("nothing built"); return -1; }
  if (!facet_current && h->nf > 0) { ferr("facet current is NULL"); return 2; }
  if (!terminal_mask) { ferr("terminal mask is NULL"); return 3; }
  if (!current) { ferr("current is NULL"); return 4; }
  FHIP(hipSetDevice(h->device));
  double *dj = nullptr;
  unsigned char *dm = nullptr;
  FHIP(hipMalloc(&dj, sizeof(double) * std::max(h->nf, 1)));
  FHIP(hipMalloc(&dm, std::max(h->nc, 1)));
  int e = 0;
  if ((h->nf > 0 && hipMemcpy(dj, facet_current, sizeof(double) * h->nf, hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(dm, terminal_mask, h->nc, hipMemcpyHostToDevice) != hipSuccess) {
    ferr("copy failed");
    e = -10;
  }
  if (!e) e = shefill_terminal_current_device(h, dj, dm, current);
  hipFree(dj);
  hipFree(dm);
  return e;
}

}  // extern "C"
