/*
 * she_oracle.c -- CPU FP64 restatement of the ViennaSHE SHE linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 * Every function names the reference file:line it restates.  Reference
 * paths are relative to the ViennaSHE 1.3.0 source tree.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "she_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* constants: physics/constants.hpp:60-76, math/constants.hpp:42       */
/* ------------------------------------------------------------------ */
#define PI_ 3.1415926535897932384626433832795
static const double Q_ = 1.602176565e-19;
static const double KB_ = 1.3806488e-23;
static const double M0_ = 9.10938291e-31;
static double hbar_(void) { return 6.62606896e-34 / (2.0 * 3.1415926535897932384626433832795); }

static char g_err[512];

/* Test-only scale factors of the fill (oracle_set_plant): 1 everywhere except
 * in the physics pins, which switch a term off or plant a factor to show that
 * their check catches it. */
static double g_plant_diff = 1.0, g_plant_drift = 1.0, g_plant_ac = 1.0, g_plant_op = 1.0;
void oracle_set_plant(double diffusion, double drift, double acoustic, double optical) {
  g_plant_diff = diffusion;
  g_plant_drift = drift;
  g_plant_ac = acoustic;
  g_plant_op = optical;
}
const char *oracle_last_error(void) { return g_err; }
static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

/* ------------------------------------------------------------------ */
/* silicon parameters: materials/all.hpp:49-107                        */
/* ------------------------------------------------------------------ */
static double m_dos(int ct) { return ct == 0 ? 0.32 * M0_ : 0.8 * M0_; }
static double m_long(int ct) { return ct == 0 ? 0.98 * M0_ : 0.16 * M0_; }
static double m_trans(int ct) { return ct == 0 ? 0.19 * M0_ : 0.16 * M0_; }
static double m_cond(int ct) {
  if (ct == 0) return 3.0 / (1.0 / m_long(ct) + 1.0 / m_trans(ct) + 1.0 / m_trans(ct));
  return 0.36 * M0_;
}
static double band_gap(void) { return 1.11 * Q_; }

/* Modena dispersion: physics/dispersion.hpp:207-290 (alpha = 0.5/eV) */
static double dos(int ct, double eps) {
  if (eps > 0.0) {
    const double hbar = hbar_();
    const double alpha = 0.5 / Q_;
    const double mstar = m_dos(ct);
    const double pi = PI_;
    const double prefactor = mstar * sqrt(2.0 * mstar) / (8.0 * pi * pi * pi * hbar * hbar * hbar);
    return prefactor * 4.0 * pi * (1.0 + 2.0 * alpha * eps) * sqrt(eps * (1.0 + alpha * eps));
  }
  return 0.0;
}
static double velocity(int ct, double eps) {
  if (eps < 0) return 0;
  const double alpha = 0.5 / Q_;
  double gamma = eps * (1.0 + alpha * eps);
  double dgam = 1.0 + 2.0 * alpha * eps;
  double mc = m_cond(ct), md = m_dos(ct);
  return sqrt(mc / md) * sqrt(2.0 * gamma / m_dos(ct)) / dgam;
}
static double symmetry_factor(int ct) { return ct == 0 ? 6.0 : 4.0; }

/* physics/physics.hpp:71-75, :84-89, :97-100, :125-133, :141-163, :172-184 */
static double optical_phonon_number(double e, double T) { return 1.0 / (exp(e / (KB_ * T)) - 1.0); }
static double band_weight(double T, int ct) {
  const double Ncv0 = 2.50933670314359713002e19 * 100 * 100 * 100;
  const double me = (ct == 0) ? 0.33 : 0.6;
  return Ncv0 * pow(me * (T / 300.0), 1.5) * 6.0;
}
static double band_edge(int ct) { return (ct == 0 ? 1.0 : -1.0) * band_gap() * 0.5; }
static double aux_conc(double T, int ct) {
  const double kbT = KB_ * T;
  const double pol = (ct == 0) ? -1.0 : 1.0;
  return band_weight(T, ct) * exp(pol * band_edge(ct) / kbT);
}
static double built_in_potential(double T, double dn, double dp) {
  if (dn <= 0 && dp <= 0) return 0;
  const double net = dn - dp;
  const double naux = aux_conc(T, 0), paux = aux_conc(T, 1);
  double x = sqrt(net * net + 4.0 * naux * paux);
  if (net >= 0.0) {
    x = (x + net) * 0.5 / naux;
    return T * log(x) * KB_ / Q_;
  }
  x = (x - net) * 0.5 / paux;
  return -T * log(x) * KB_ / Q_;
}
static double contact_carrier_ohm(double T, double dn, double dp, int ct) {
  const double kbT = KB_ * T;
  const double phiB = built_in_potential(T, dn, dp);
  const double pol = (ct == 0) ? -1.0 : 1.0;
  return band_weight(T, ct) * exp(pol * (band_edge(ct) - Q_ * phiB) / kbT);
}

/* ------------------------------------------------------------------ */
/* std::map-row sparse matrix: math/linalg_util.hpp:67-196             */
/* operator() inserts a zero entry when absent; erase removes.         */
/* ------------------------------------------------------------------ */
typedef struct { int64_t *c; double *v; int32_t len, cap; } orow;
typedef struct { int64_t n; orow *r; } omap;

static int omap_init(omap *m, int64_t n) {
  m->n = n;
  m->r = (orow *)calloc((size_t)(n > 0 ? n : 1), sizeof(orow));
  return m->r != NULL;
}
static void omap_free(omap *m) {
  if (!m->r) return;
  for (int64_t i = 0; i < m->n; ++i) { free(m->r[i].c); free(m->r[i].v); }
  free(m->r);
  m->r = NULL;
}
static int32_t orow_lower(const orow *r, int64_t j) {
  int32_t lo = 0, hi = r->len;
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    if (r->c[mid] < j) lo = mid + 1; else hi = mid;
  }
  return lo;
}
static double *omap_ref(omap *m, int64_t i, int64_t j) {
  orow *r = &m->r[i];
  int32_t p = orow_lower(r, j);
  if (p < r->len && r->c[p] == j) return &r->v[p];
  if (r->len == r->cap) {
    int32_t nc = r->cap ? 2 * r->cap : 8;
    r->c = (int64_t *)realloc(r->c, (size_t)nc * sizeof(int64_t));
    r->v = (double *)realloc(r->v, (size_t)nc * sizeof(double));
    r->cap = nc;
  }
  memmove(r->c + p + 1, r->c + p, (size_t)(r->len - p) * sizeof(int64_t));
  memmove(r->v + p + 1, r->v + p, (size_t)(r->len - p) * sizeof(double));
  r->c[p] = j;
  r->v[p] = 0.0;
  r->len++;
  return &r->v[p];
}
static void omap_erase(omap *m, int64_t i, int64_t j) {
  orow *r = &m->r[i];
  int32_t p = orow_lower(r, j);
  if (p < r->len && r->c[p] == j) {
    memmove(r->c + p, r->c + p + 1, (size_t)(r->len - p - 1) * sizeof(int64_t));
    memmove(r->v + p, r->v + p + 1, (size_t)(r->len - p - 1) * sizeof(double));
    r->len--;
  }
}
static int64_t omap_nnz(const omap *m) {
  int64_t s = 0;
  for (int64_t i = 0; i < m->n; ++i) s += m->r[i].len;
  return s;
}
/* trans(): linalg_util.hpp:118-130 */
static void omap_trans(const omap *a, omap *t) {
  omap_init(t, a->n);
  /* count per column then fill in ascending source-row order: equals the
   * map insertion result (keys sorted). */
  for (int64_t i = 0; i < a->n; ++i)
    for (int32_t k = 0; k < a->r[i].len; ++k) t->r[a->r[i].c[k]].cap++;
  for (int64_t i = 0; i < a->n; ++i) {
    orow *r = &t->r[i];
    if (r->cap) {
      r->c = (int64_t *)malloc((size_t)r->cap * sizeof(int64_t));
      r->v = (double *)malloc((size_t)r->cap * sizeof(double));
    }
  }
  for (int64_t i = 0; i < a->n; ++i)
    for (int32_t k = 0; k < a->r[i].len; ++k) {
      orow *r = &t->r[a->r[i].c[k]];
      r->c[r->len] = i;
      r->v[r->len] = a->r[i].v[k];
      r->len++;
    }
}

/* ------------------------------------------------------------------ */
/* coupling matrices with map semantics (present flag + value)          */
/* ------------------------------------------------------------------ */
typedef struct { int K; unsigned char *p; double *v; } cmat;
static void cmat_init(cmat *m, int K) {
  m->K = K;
  m->p = (unsigned char *)calloc((size_t)K * K, 1);
  m->v = (double *)calloc((size_t)K * K, sizeof(double));
}
static void cmat_free(cmat *m) { free(m->p); free(m->v); }
static void cmat_copy(cmat *d, const cmat *s) {
  cmat_init(d, s->K);
  memcpy(d->p, s->p, (size_t)s->K * s->K);
  memcpy(d->v, s->v, (size_t)s->K * s->K * sizeof(double));
}
static double cmat_get(const cmat *m, int i, int j) { return m->p[i * m->K + j] ? m->v[i * m->K + j] : 0.0; }
static void cmat_trans(cmat *d, const cmat *s) {
  cmat_init(d, s->K);
  for (int i = 0; i < s->K; ++i)
    for (int j = 0; j < s->K; ++j)
      if (s->p[i * s->K + j]) { d->p[j * s->K + i] = 1; d->v[j * s->K + i] = s->v[i * s->K + j]; }
}
/* operator*(factor): linalg_util.hpp:132-143.  For every stored (i,j) of the
 * source it scales result(j,i) -- the transposed position -- creating it
 * (with value 0) when absent.  Reproduced as is. */
static void cmat_scaled(cmat *d, const cmat *s, double f) {
  cmat_copy(d, s);
  int K = s->K;
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j)
      if (s->p[i * K + j]) {
        d->p[j * K + i] = 1;
        d->v[j * K + i] *= f;
      }
}
/* operator+=: linalg_util.hpp:145-156 */
static void cmat_add(cmat *d, const cmat *s) {
  int K = s->K;
  for (int i = 0; i < K * K; ++i)
    if (s->p[i]) { d->p[i] = 1; d->v[i] += s->v[i]; }
}

/* ------------------------------------------------------------------ */
/* spherical harmonics iteration: math/spherical_harmonics.hpp:213-252  */
/* ------------------------------------------------------------------ */
static int harmonics_list(int L, int odd, int *out) {
  int n = 0;
  for (int l = odd ? 1 : 0; l <= L; l += 2)
    for (int m = -l; m <= l; ++m) out[n++] = l * l + l + m;
  return n;
}
/* she_quantity.hpp:35-54 */
static int even_unknowns(int L) { int n = 0; for (int l = 0; l <= L; l += 2) n += 2 * l + 1; return n; }
static int odd_unknowns(int L) { int n = 0; for (int l = 1; l <= L; l += 2) n += 2 * l + 1; return n; }

/* ------------------------------------------------------------------ */
/* the system                                                          */
/* ------------------------------------------------------------------ */
struct oracle_system {
  oracle_config cfg;
  int dim, nx, ny, ncells, nfacets, NH;
  double T;
  double *cx, *cy, *cvol;           /* centroids, cell volume */
  int32_t *fc1, *fc2, *fdir;        /* facet cells (coboundary order), direction */
  double *fvol;                     /* facet volume */
  int32_t *cf;                      /* facets on cell: 4 per cell, -1 = boundary */
  double *dn, *dp, *psi;
  int32_t *contact;
  double *Hval, *shift_c, *shift_f;
  int32_t *ord_c, *ord_f;
  int64_t *idx_c, *idx_f;
  double *bnd;                      /* boundary values per (cell, H) */
  int64_t n_even, n_odd, N;
  omap full;                        /* after diagonalisation + healing */
  double *b_full;
  omap unscaled;                    /* compressed system before scaling */
  double *b_unscaled;
  omap red;                         /* scaled + normalised */
  double *b_red, *scaling, *colfac;
  int64_t *blocks;
  cmat a[3], b[3];
  int Lc;                            /* table order actually used */
};

static double kin_c(const oracle_system *s, int c, int H) {
  return s->cfg.carrier == 0 ? s->Hval[H] - s->shift_c[c] : s->shift_c[c] - s->Hval[H];
}
static double kin_f(const oracle_system *s, int f, int H) {
  return s->cfg.carrier == 0 ? s->Hval[H] - s->shift_f[f] : s->shift_f[f] - s->Hval[H];
}
static double kin(const oracle_system *s, int isf, int e, int H) { return isf ? kin_f(s, e, H) : kin_c(s, e, H); }

/* box_height: she/assemble_common.hpp:411-427 */
static double box_height(const oracle_system *s, int isf, int e, int H) {
  double h = 0;
  if (H > 0) h = fabs(kin(s, isf, e, H) - kin(s, isf, e, H - 1));
  if (H < s->NH - 1) h += fabs(kin(s, isf, e, H + 1) - kin(s, isf, e, H));
  return h / 2.0;
}
/* averaged_density_of_states: assemble_common.hpp:431-462 */
static double avg_dos(const oracle_system *s, int isf, int e, int H) {
  int ct = s->cfg.carrier;
  double Z = dos(ct, kin(s, isf, e, H));
  double nc = 1;
  if (H < s->NH - 1) { Z += dos(ct, (kin(s, isf, e, H) + kin(s, isf, e, H + 1)) / 2.0); ++nc; }
  if (H > 0) { Z += dos(ct, (kin(s, isf, e, H) + kin(s, isf, e, H - 1)) / 2.0); ++nc; }
  return Z / nc;
}
/* integral_vZ: assemble_common.hpp:519-551 */
static double integral_vZ(const oracle_system *s, int isf, int e, int H) {
  int ct = s->cfg.carrier;
  double ec = kin(s, isf, e, H);
  double vZ = velocity(ct, ec) * dos(ct, ec);
  double nc = 1;
  if (H < s->NH - 1) {
    double k = (kin(s, isf, e, H) + kin(s, isf, e, H + 1)) / 2.0;
    vZ += velocity(ct, k) * dos(ct, k);
    ++nc;
  }
  if (H > 0) {
    double k = (kin(s, isf, e, H) + kin(s, isf, e, H - 1)) / 2.0;
    vZ += velocity(ct, k) * dos(ct, k);
    ++nc;
  }
  vZ /= nc;
  return vZ * box_height(s, isf, e, H);
}
/* integral_Z_over_hk: assemble_common.hpp:152-164 */
static int integral_Z_over_hk(int ct, double lo, double up, double *out) {
  if (up < lo) { set_err("Invalid interval in integral_Z_over_hk(): [%g, %g]", lo, up); return -1; }
  *out = 0.5 * (dos(ct, up) * velocity(ct, up) - dos(ct, lo) * velocity(ct, lo));
  return 0;
}
/* lower/upper_kinetic_energy_at_facet: assemble_common.hpp:323-388
 * (including the operator precedence of the index_H == 0 / N_H-1 branch) */
static double lower_kin_facet(const oracle_system *s, int f, int H) {
  int c1 = s->fc1[f], c2 = s->fc2[f];
  if (H == 0) return kin_c(s, c1, H) + kin_c(s, c2, H) / 2.0;
  return (kin_c(s, c1, H - 1) + kin_c(s, c2, H - 1) + kin_c(s, c1, H) + kin_c(s, c2, H)) / 4.0;
}
static double upper_kin_facet(const oracle_system *s, int f, int H) {
  int c1 = s->fc1[f], c2 = s->fc2[f];
  if (H == s->NH - 1) return kin_c(s, c1, H) + kin_c(s, c2, H) / 2.0;
  return (kin_c(s, c1, H + 1) + kin_c(s, c2, H + 1) + kin_c(s, c1, H) + kin_c(s, c2, H)) / 4.0;
}
static int64_t idx_of(const oracle_system *s, int isf, int e, int H) {
  return isf ? s->idx_f[(int64_t)e * s->NH + H] : s->idx_c[(int64_t)e * s->NH + H];
}
static int ord_of(const oracle_system *s, int isf, int e, int H) {
  return isf ? s->ord_f[(int64_t)e * s->NH + H] : s->ord_c[(int64_t)e * s->NH + H];
}

/* add_block_matrix: util/block_matrix_writer.hpp:44-81 */
static void add_block_matrix(omap *A, int64_t row_index, int64_t col_index, double prefactor,
                             const cmat *C, int Lr, int odd_r, int Lc, int odd_c, int elastic_prev) {
  int ri[64], ci[64];
  int nr = harmonics_list(Lr, odd_r, ri), nc = harmonics_list(Lc, odd_c, ci);
  for (int a = 0; a < nr; ++a) {
    int64_t row = row_index + a;
    for (int b = 0; b < nc; ++b) {
      int64_t col = col_index + b;
      double entry = cmat_get(C, ri[a], ci[b]);
      if (entry != 0.0) {
        if (elastic_prev && row == row_index && col == col_index) continue;
        *omap_ref(A, row, col) += prefactor * entry;
      }
    }
  }
}

/* coupling_matrix_in_direction: assemble_common.hpp:176-290.
 * p3 = normalised (centroid(c2) - centroid(c1)). */
static void coupling_in_direction(const oracle_system *s, const cmat *m, int c1, int c2, cmat *out) {
  int ct = s->cfg.carrier;
  double px = s->cx[c2] - s->cx[c1], py = s->dim > 1 ? s->cy[c2] - s->cy[c1] : 0.0;
  double nrm = sqrt(px * px + py * py);
  px /= nrm;
  py /= nrm;
  double mt = m_trans(ct), md = m_dos(ct);
  cmat_scaled(out, &m[0], px * sqrt(md / mt));
  if (s->dim > 1) {
    cmat t;
    cmat_scaled(&t, &m[1], py * sqrt(md / mt));
    cmat_add(out, &t);
    cmat_free(&t);
  }
}

/* energy_index_H: she/assemble_scattering.hpp:64-105 */
static long energy_index_H(const oracle_system *s, int isf, int e, long hint, double E) {
  long r = hint;
  if (E < kin(s, isf, e, (int)r) || E > kin(s, isf, e, (int)r)) {
    long inc = 1;
    double diff = E - kin(s, isf, e, (int)r);
    double diff_init = diff;
    if (diff < 0) inc = -1;
    while ((r < (long)(s->NH - 1)) && (r > 0) && (diff * diff_init > 0)) {
      r += inc;
      diff = E - kin(s, isf, e, (int)r);
    }
    if (diff * diff_init > 0) return -1;
    double e1 = fabs(diff), e2 = e1;
    if ((r - inc >= 0) && (r - inc < (long)s->NH)) e2 = fabs(E - kin(s, isf, e, (int)(r - inc)));
    if (e2 < e1) return r - inc;
  }
  return r;
}

/* scattering events: acoustic_phonon_scattering.hpp:83-114,
 * optical_phonon_scattering.hpp:62-80,122-179, scattering/config.hpp:50-116 */
typedef struct { double ei, ef, rate, gen; } sevent;
static int scattering_events(const oracle_system *s, double E, sevent *ev) {
  int n = 0, ct = s->cfg.carrier;
  const double T = s->T, pi = PI_, hbar = hbar_();
  if (s->cfg.acoustic) {
    const double ul = 9040.0, rho = 2330.0;
    const double E1 = ct == 0 ? 8.9 * Q_ : 5.12 * Q_;
    const double rate = 2.0 * pi * KB_ * T * E1 * E1 / (hbar * ul * ul * rho);
    const double fit = ct == 0 ? 1.05 : 0.875;
    ev[n].ei = E; ev[n].ef = E; ev[n].rate = g_plant_ac * fit * rate; ev[n].gen = 0; ++n;
  }
  if (s->cfg.optical) {
    double hw = 1.24 * (Q_ / 20.0);
    double sp = s->cfg.energy_spacing;
    if (sp > 0) {
      size_t steps = (size_t)(hw / sp);
      double err0 = hw - steps * sp, err1 = (steps + 1.0) * sp - hw;
      if (err1 < err0) ++steps;
      hw = steps * sp;
    }
    const double Nop = optical_phonon_number(hw, T);
    const double rho = 2330.0, DtK = 5e10 * Q_;
    const double r = g_plant_op * 0.7 * (pi * DtK * DtK * hbar / (2.0 * rho * hw));
    ev[n].ei = E;      ev[n].ef = E - hw; ev[n].rate = r * (Nop + 1.0); ev[n].gen = 0; ++n;
    ev[n].ei = E + hw; ev[n].ef = E;      ev[n].rate = r * (Nop + 1.0); ev[n].gen = 0; ++n;
    ev[n].ei = E;      ev[n].ef = E + hw; ev[n].rate = r * Nop;         ev[n].gen = 0; ++n;
    ev[n].ei = E - hw; ev[n].ef = E;      ev[n].rate = r * Nop;         ev[n].gen = 0; ++n;
  }
  return n;
}

/* assemble_scattering_operator_on_box: she/assemble_scattering.hpp:116-335 */
static void assemble_scattering(oracle_system *s, int isf, int e, int H, const cmat *sc_in, const cmat *sc_out) {
  const double pi = PI_;
  int64_t row = idx_of(s, isf, e, H);
  if (row < 0) return;
  double E = kin(s, isf, e, H);
  int Lmid = ord_of(s, isf, e, H);
  double V;
  if (isf) {
    int f = e;
    double dx = s->cx[s->fc1[f]] - s->cx[s->fc2[f]];
    double dy = s->dim > 1 ? s->cy[s->fc1[f]] - s->cy[s->fc2[f]] : 0.0;
    V = s->fvol[f] * (sqrt(dx * dx + dy * dy) / (double)s->dim);
  } else {
    V = s->cvol[e];
  }
  sevent ev[8];
  int ne = scattering_events(s, E, ev);
  for (int k = 0; k < ne; ++k) {
    double ei = ev[k].ei, ef = ev[k].ef, rate = ev[k].rate, gen = ev[k].gen;
    if (rate <= 0 && gen <= 0) continue;
    long iHi = energy_index_H(s, isf, e, H, ei), iHf = energy_index_H(s, isf, e, H, ef);
    if (iHi == -1 || iHf == -1) continue;
    double Zi = avg_dos(s, isf, e, (int)iHi), Zf = avg_dos(s, isf, e, (int)iHf);
    double h1 = box_height(s, isf, e, (int)iHi), h2 = box_height(s, isf, e, (int)iHf);
    double height = h1 < h2 ? h1 : h2;
    int elastic_prev = (iHi == iHf) && !isf;
    if (ei <= E && ei >= E) {
      int64_t col = idx_of(s, isf, e, (int)iHf);
      if (col >= 0) {
        double coeff = 2.0 * pi * rate * Zi * Zf * V * height;
        add_block_matrix(&s->full, row, row, coeff, sc_out, Lmid, isf, Lmid, isf, elastic_prev);
      }
    }
    if (ef <= E && ef >= E) {
      int64_t col = idx_of(s, isf, e, (int)iHi);
      if (col >= 0) {
        double coeff = -2.0 * pi * rate * Zi * Zf * V * height - gen * Zi * V * height;
        add_block_matrix(&s->full, row, col, coeff, sc_in, Lmid, isf, Lmid, isf, elastic_prev);
      }
    }
  }
}

/* assemble_boundary_on_box (generation/recombination branch):
 * she/assemble_boundary.hpp:69-183; write_boundary: assemble_common.hpp:300-316 */
static void assemble_boundary(oracle_system *s, int c, int H, const cmat *ident) {
  int64_t row = s->idx_c[(int64_t)c * s->NH + H];
  if (row < 0) return;
  int L = s->ord_c[(int64_t)c * s->NH + H];
  if (!s->contact[c]) return;
  double tau = s->cfg.tau;
  double height = box_height(s, 0, c, H);
  double bnd = s->bnd[(int64_t)c * s->NH + H];
  double V = s->cvol[c];
  double coeff = V * height / tau * avg_dos(s, 0, c, H);
  add_block_matrix(&s->full, row, row, coeff, ident, L, 0, L, 0, 0);
  /* write_boundary over the zeroth-order even harmonic only */
  s->b_full[row] -= (-bnd * coeff) * cmat_get(ident, 0, 0);
}

/* assemble_free_streaming_operator_on_box: she/assemble_streaming.hpp:73-365 */
static int assemble_streaming(oracle_system *s, int c, int f, int H, const cmat *diff, const cmat *drift, int odd) {
  int other = (s->fc1[f] == c) ? s->fc2[f] : s->fc1[f];
  int64_t row = odd ? s->idx_f[(int64_t)f * s->NH + H] : s->idx_c[(int64_t)c * s->NH + H];
  if (row < 0) return 0;
  int64_t col = odd ? s->idx_c[(int64_t)c * s->NH + H] : s->idx_f[(int64_t)f * s->NH + H];
  if (col < 0) return 0;
  double dx = s->cx[c] - s->cx[other], dy = s->dim > 1 ? s->cy[c] - s->cy[other] : 0.0;
  double conn_len = sqrt(dx * dx + dy * dy);
  /* structured meshes: the outer facet normal is parallel to the cell connection */
  double area = s->fvol[f] * 1.0;
  int Lr = odd ? s->ord_f[(int64_t)f * s->NH + H] : s->ord_c[(int64_t)c * s->NH + H];
  int Lcol = odd ? s->ord_c[(int64_t)c * s->NH + H] : s->ord_f[(int64_t)f * s->NH + H];
  double lo = lower_kin_facet(s, f, H), up = upper_kin_facet(s, f, H);
  double cdiff = g_plant_diff * integral_vZ(s, 1, f, H) * area;
  add_block_matrix(&s->full, row, col, cdiff, diff, Lr, odd, Lcol, !odd, 0);
  double izhk;
  if (integral_Z_over_hk(s->cfg.carrier, lo, up, &izhk)) return -1;
  /* force_on_facet_accessor()(quan, other, cell, H): (E_kin(cell) - E_kin(other)) / distance */
  double ddx = s->cx[other] - s->cx[c], ddy = s->dim > 1 ? s->cy[other] - s->cy[c] : 0.0;
  double dist = sqrt(ddx * ddx + ddy * ddy);
  double force = (kin_c(s, c, H) - kin_c(s, other, H)) / dist;
  double cdrift = g_plant_drift * izhk * force * area * conn_len / 2.0;
  add_block_matrix(&s->full, row, col, cdrift, drift, Lr, odd, Lcol, !odd, 0);
  return 0;
}

/* ------------------------------------------------------------------ */
/* setup: device geometry, energies, orders, boundary values, mapping  */
/* ------------------------------------------------------------------ */
static int setup_geometry(oracle_system *s, const oracle_device *d) {
  s->dim = d->dim;
  s->nx = d->nx;
  s->ny = d->dim > 1 ? d->ny : 1;
  s->ncells = s->nx * s->ny;
  s->T = d->temperature;
  int nc = s->ncells;
  s->cx = (double *)calloc(nc, sizeof(double));
  s->cy = (double *)calloc(nc, sizeof(double));
  s->cvol = (double *)calloc(nc, sizeof(double));
  s->dn = (double *)malloc(nc * sizeof(double));
  s->dp = (double *)malloc(nc * sizeof(double));
  s->psi = (double *)malloc(nc * sizeof(double));
  s->contact = (int32_t *)malloc(nc * sizeof(int32_t));
  memcpy(s->dn, d->doping_n, nc * sizeof(double));
  memcpy(s->dp, d->doping_p, nc * sizeof(double));
  memcpy(s->psi, d->potential, nc * sizeof(double));
  memcpy(s->contact, d->contact, nc * sizeof(int32_t));
  /* cell centroids = vertex averages; cell ids y-outer, x-inner
   * (util/generate_device.hpp:249-268 for 1D, :426-450 for 2D) */
  for (int j = 0; j < s->ny; ++j)
    for (int i = 0; i < s->nx; ++i) {
      int c = j * s->nx + i;
      if (s->dim == 1) {
        s->cx[c] = (d->xv[i] + d->xv[i + 1]) / 2.0;
        s->cvol[c] = fabs(d->xv[i + 1] - d->xv[i]);
      } else {
        s->cx[c] = (d->xv[i] + d->xv[i + 1] + d->xv[i] + d->xv[i + 1]) / 4.0;
        s->cy[c] = (d->yv[j] + d->yv[j] + d->yv[j + 1] + d->yv[j + 1]) / 4.0;
        s->cvol[c] = fabs(d->xv[i + 1] - d->xv[i]) * fabs(d->yv[j + 1] - d->yv[j]);
      }
    }
  /* interior facets in creation order: per cell, the right (x) facet then
   * the top (y) facet.  Facets on cell: bottom, left, right, top. */
  int maxf = 2 * nc;
  s->fc1 = (int32_t *)malloc(maxf * sizeof(int32_t));
  s->fc2 = (int32_t *)malloc(maxf * sizeof(int32_t));
  s->fdir = (int32_t *)malloc(maxf * sizeof(int32_t));
  s->fvol = (double *)malloc(maxf * sizeof(double));
  s->cf = (int32_t *)malloc((size_t)nc * 4 * sizeof(int32_t));
  for (int k = 0; k < nc * 4; ++k) s->cf[k] = -1;
  int nf = 0;
  for (int j = 0; j < s->ny; ++j)
    for (int i = 0; i < s->nx; ++i) {
      int c = j * s->nx + i;
      if (i + 1 < s->nx) {
        s->fc1[nf] = c; s->fc2[nf] = c + 1; s->fdir[nf] = 0;
        s->fvol[nf] = s->dim == 1 ? 1.0 : fabs(d->yv[j + 1] - d->yv[j]);
        s->cf[c * 4 + 2] = nf; s->cf[(c + 1) * 4 + 1] = nf;
        ++nf;
      }
      if (s->dim > 1 && j + 1 < s->ny) {
        s->fc1[nf] = c; s->fc2[nf] = c + s->nx; s->fdir[nf] = 1;
        s->fvol[nf] = fabs(d->xv[i + 1] - d->xv[i]);
        s->cf[c * 4 + 3] = nf; s->cf[(c + s->nx) * 4 + 0] = nf;
        ++nf;
      }
    }
  s->nfacets = nf;
  return 0;
}

/* setup_energies: she/setup_energies.hpp:61-153 (H-transform, no quantum correction) */
static int setup_energies(oracle_system *s) {
  double maxp = -1.0 * 1.7976931348623157e308, minp = 1.7976931348623157e308;
  for (int c = 0; c < s->ncells; ++c) {
    double p = s->psi[c] + 0.0;
    if (p < minp) minp = p;
    if (p > maxp) maxp = p;
  }
  const oracle_config *cf = &s->cfg;
  double range = Q_ * (maxp - minp) + 2.0 * cf->energy_spacing + cf->min_kin_range;
  double edge = cf->carrier == 0 ? -Q_ * maxp - cf->energy_spacing + band_gap() / 2.0
                                 : -Q_ * minp + cf->energy_spacing - band_gap() / 2.0;
  double inc = (cf->carrier == 0 ? 1.0 : -1.0) * cf->energy_spacing;
  size_t NH = (size_t)(range / cf->energy_spacing) + 1;
  s->NH = (int)NH;
  s->Hval = (double *)malloc(NH * sizeof(double));
  for (size_t h = 0; h < NH; ++h) s->Hval[h] = edge + h * inc;
  s->shift_c = (double *)malloc(s->ncells * sizeof(double));
  s->shift_f = (double *)malloc((s->nfacets > 0 ? s->nfacets : 1) * sizeof(double));
  for (int c = 0; c < s->ncells; ++c)
    s->shift_c[c] = cf->carrier == 0 ? -Q_ * s->psi[c] + band_gap() / 2.0 : -Q_ * s->psi[c] - band_gap() / 2.0;
  for (int f = 0; f < s->nfacets; ++f) s->shift_f[f] = (s->shift_c[s->fc1[f]] + s->shift_c[s->fc2[f]]) / 2.0;
  return 0;
}

/* distribute_expansion_orders: she/expansion_order.hpp:262-363 (non-adaptive),
 * lower_expansion_order_near_bandedge :60-130, smooth_expansion_order :137-190 */
static void distribute_orders(oracle_system *s) {
  int NH = s->NH, L = s->cfg.L_max;
  s->ord_c = (int32_t *)malloc((size_t)s->ncells * NH * sizeof(int32_t));
  s->ord_f = (int32_t *)malloc((size_t)(s->nfacets > 0 ? s->nfacets : 1) * NH * sizeof(int32_t));
  for (int64_t k = 0; k < (int64_t)s->ncells * NH; ++k) s->ord_c[k] = L;
  for (int64_t k = 0; k < (int64_t)s->nfacets * NH; ++k) s->ord_f[k] = L;
  for (int c = 0; c < s->ncells; ++c) {
    for (int H = 1; H < NH - 1; ++H) {
      if (kin_c(s, c, H) <= 0) { s->ord_c[(int64_t)c * NH + H] = 1; continue; }
      int allpos = 1;
      for (int k = 0; k < 4; ++k) {
        int f = s->cf[c * 4 + k];
        if (f < 0) continue;
        int o = s->fc1[f] == c ? s->fc2[f] : s->fc1[f];
        if (kin_c(s, o, H - 1) <= 0) allpos = 0;
      }
      if (allpos) break;
      s->ord_c[(int64_t)c * NH + H] = 1;
      for (int k = 0; k < 4; ++k) {
        int f = s->cf[c * 4 + k];
        if (f < 0) continue;
        int o = s->fc1[f] == c ? s->fc2[f] : s->fc1[f];
        s->ord_c[(int64_t)o * NH + H] = 1;
      }
    }
  }
  for (int it = 0; it < 10; ++it) {
    int changed = 0;
    for (int f = 0; f < s->nfacets; ++f) {
      int c1 = s->fc1[f], c2 = s->fc2[f];
      for (int H = 0; H < NH; ++H) {
        int32_t *o1 = &s->ord_c[(int64_t)c1 * NH + H], o2 = s->ord_c[(int64_t)c2 * NH + H];
        if (*o1 > o2 + 2) { changed = 1; *o1 = o2 + 2; }
      }
    }
    if (!changed) break;
  }
}

/* write_boundary_conditions: she/boundary_conditions.hpp:57-136 (DF discretisation) */
static void boundary_values(oracle_system *s) {
  int NH = s->NH, ct = s->cfg.carrier;
  s->bnd = (double *)calloc((size_t)s->ncells * NH, sizeof(double));
  for (int c = 0; c < s->ncells; ++c) {
    const double T = s->T;
    double corr = 0;
    for (int H = 1; H < NH - 1; ++H) {
      double mid = kin_c(s, c, H);
      double lo = (kin_c(s, c, H - 1) + mid) / 2.0;
      if (lo < 0.0) lo = 0.0;
      double up = (kin_c(s, c, H + 1) + mid) / 2.0;
      if (lo < up && up >= 0) corr += exp(-mid / (KB_ * T)) * avg_dos(s, 0, c, H) * box_height(s, 0, c, H);
    }
    corr *= symmetry_factor(ct);
    for (int H = 1; H < NH - 1; ++H) {
      double en = kin_c(s, c, H);
      double lo = (kin_c(s, c, H - 1) + en) / 2.0;
      if (lo < 0.0) lo = 0.0;
      double up = (kin_c(s, c, H + 1) + en) / 2.0;
      double v = 0.0;
      if (lo < up && up >= 0)
        v = contact_carrier_ohm(T, s->dn[c], s->dp[c], ct) * exp(-en / (KB_ * T)) / corr;
      s->bnd[(int64_t)c * NH + H] = v;
    }
  }
}

/* create_even_mapping / create_odd_mapping: mapping.hpp:97-329 (single process) */
static void create_mapping(oracle_system *s) {
  int NH = s->NH;
  s->idx_c = (int64_t *)malloc((size_t)s->ncells * NH * sizeof(int64_t));
  s->idx_f = (int64_t *)malloc((size_t)(s->nfacets > 0 ? s->nfacets : 1) * NH * sizeof(int64_t));
  int64_t u = 0;
  for (int H = 0; H < NH; ++H)
    for (int c = 0; c < s->ncells; ++c) {
      int64_t k = (int64_t)c * NH + H;
      if (H == 0 || H == NH - 1) { s->idx_c[k] = -1; continue; }
      if (avg_dos(s, 0, c, H) > 0 && kin_c(s, c, H) < s->cfg.max_kin_range) {
        s->idx_c[k] = u;
        u += even_unknowns(s->ord_c[k]);
      } else {
        s->idx_c[k] = -1;
        s->ord_c[k] = 0;
      }
    }
  s->n_even = u;
  for (int H = 0; H < NH; ++H)
    for (int f = 0; f < s->nfacets; ++f) {
      int64_t k = (int64_t)f * NH + H;
      int64_t d1 = s->idx_c[(int64_t)s->fc1[f] * NH + H], d2 = s->idx_c[(int64_t)s->fc2[f] * NH + H];
      if (d1 > -1 && d2 > -1 && avg_dos(s, 1, f, H) > 0) {
        s->idx_f[k] = u;
        u += odd_unknowns(s->ord_f[k]);
      } else {
        s->idx_f[k] = -1;
        s->ord_f[k] = 0;
      }
    }
  s->n_odd = u - s->n_even;
  s->N = u;
}

/* she::assemble: she/assemble_all.hpp:46-424 */
static int assemble(oracle_system *s) {
  int K = (s->Lc + 1) * (s->Lc + 1);
  cmat ident, sc_in, sc_out, aT[3], bT[3];
  cmat_init(&ident, K);
  cmat_init(&sc_in, K);
  cmat_init(&sc_out, K);
  for (int i = 0; i < K; ++i)
    for (int j = 0; j < K; ++j) { ident.p[i * K + j] = 1; ident.v[i * K + j] = (i == j) ? 1.0 : 0.0; }
  int Ksc = (s->cfg.L_max + 1) * (s->cfg.L_max + 1);
  for (int i = 0; i < Ksc && i < K; ++i) { sc_out.p[i * K + i] = 1; sc_out.v[i * K + i] += 1.0; }
  sc_in.p[0] = 1; sc_in.v[0] += 1.0;
  for (int d = 0; d < 3; ++d) { cmat_trans(&aT[d], &s->a[d]); cmat_trans(&bT[d], &s->b[d]); }
  int NH = s->NH, rc = 0;
  omap_init(&s->full, s->N);
  s->b_full = (double *)calloc((size_t)(s->N > 0 ? s->N : 1), sizeof(double));
  /* step 1: cells (even unknowns) */
  for (int c = 0; c < s->ncells && !rc; ++c) {
    for (int H = 0; H < NH; ++H) {
      assemble_boundary(s, c, H, &ident);
      assemble_scattering(s, 0, c, H, &sc_in, &sc_out);
    }
    static const int order[4] = {0, 1, 2, 3};
    for (int k = 0; k < 4 && !rc; ++k) {
      int f = s->cf[c * 4 + order[k]];
      if (f < 0) continue;
      int other = s->fc1[f] == c ? s->fc2[f] : s->fc1[f];
      cmat cd, cr;
      coupling_in_direction(s, s->a, c, other, &cd);
      coupling_in_direction(s, s->b, other, c, &cr);
      for (int H = 0; H < NH && !rc; ++H) rc = assemble_streaming(s, c, f, H, &cd, &cr, 0);
      cmat_free(&cd);
      cmat_free(&cr);
    }
  }
  /* step 2: facets (odd unknowns) */
  for (int f = 0; f < s->nfacets && !rc; ++f) {
    for (int H = 0; H < NH; ++H) assemble_scattering(s, 1, f, H, &sc_in, &sc_out);
    int cells[2] = {s->fc1[f], s->fc2[f]};
    for (int k = 0; k < 2 && !rc; ++k) {
      int c = cells[k], other = cells[1 - k];
      cmat cd, cr;
      coupling_in_direction(s, aT, other, c, &cd);
      coupling_in_direction(s, bT, other, c, &cr);
      for (int H = 0; H < NH && !rc; ++H) rc = assemble_streaming(s, c, f, H, &cd, &cr, 1);
      cmat_free(&cd);
      cmat_free(&cr);
    }
  }
  cmat_free(&ident); cmat_free(&sc_in); cmat_free(&sc_out);
  for (int d = 0; d < 3; ++d) { cmat_free(&aT[d]); cmat_free(&bT[d]); }
  return rc;
}

/* ------------------------------------------------------------------ */
/* she::solve front end: she/linear_solver.hpp:96-189                  */
/* ------------------------------------------------------------------ */
/* combine_lines: she/eliminate.hpp:40-68 */
static void combine_lines(omap *A, double *rhs, int64_t line1, int64_t lta, int64_t colz) {
  double s = *omap_ref(A, line1, colz) / *omap_ref(A, lta, lta);
  if (s == 0.0) return;
  rhs[line1] = rhs[line1] - rhs[lta] * s;
  orow *r = &A->r[lta];
  for (int32_t k = 0; k < r->len; ++k) {
    double v = r->v[k];
    int64_t c = r->c[k];
    if (v == 0.0) continue;
    *omap_ref(A, line1, c) -= v * s;
  }
  omap_erase(A, line1, colz);
}
static void heal_rows(omap *A, int64_t from) {
  for (int64_t i = from < 0 ? 0 : from; i < A->n; ++i) {
    orow *r = &A->r[i];
    int32_t w = 0;
    for (int32_t k = 0; k < r->len; ++k)
      if (r->v[k] != 0.0) { r->c[w] = r->c[k]; r->v[w] = r->v[k]; ++w; }
    r->len = w;
  }
}
/* diagonalise_odd2odd_coupling_matrix: she/eliminate.hpp:80-214 */
static int diagonalise(oracle_system *s) {
  omap *A = &s->full;
  int64_t N = s->N, ne = s->n_even;
  for (int64_t i = ne + 1; i < N; ++i) {
    orow *r = &A->r[i];
    for (int32_t k = 0; k < r->len; ++k) {
      int64_t c = r->c[k];
      if (c >= i) break;
      if (c < ne) continue;
      combine_lines(A, s->b_full, i, c, c);
      break;
    }
  }
  heal_rows(A, ne - 1);
  for (int64_t i = N - 1; i >= ne && i >= 0; --i) {
    orow *r = &A->r[i];
    for (int32_t k = 0; k < r->len; ++k) {
      int64_t c = r->c[k];
      if (c < i && c >= ne) { set_err("values left of the odd diagonal remain (row %lld)", (long long)i); return -1; }
      if (c < ne || c == i) continue;
      combine_lines(A, s->b_full, i, c, c);
      break;
    }
  }
  heal_rows(A, ne - 1);
  return 0;
}
/* eliminate_odd_unknowns: she/eliminate.hpp:226-293 */
static int eliminate(oracle_system *s) {
  int64_t ne = s->n_even;
  omap *A = &s->full, *C = &s->unscaled;
  omap_init(C, ne);
  s->b_unscaled = (double *)calloc((size_t)(ne > 0 ? ne : 1), sizeof(double));
  for (int64_t i = 0; i < ne; ++i) {
    s->b_unscaled[i] = s->b_full[i];
    orow *r = &A->r[i];
    for (int32_t k = 0; k < r->len; ++k) {
      int64_t col = r->c[k];
      double v = r->v[k];
      if (col < ne) {
        *omap_ref(C, i, col) += v;
      } else {
        double d = -1.0;
        orow *o = &A->r[col];
        for (int32_t q = 0; q < o->len; ++q)
          if (o->c[q] == col) { d = o->v[q]; break; }
        if (d <= 0.0) {
          set_err("eliminate_odd_unknowns(): Diagonal entry not positive or not existing! (%g at %lld)", d, (long long)col);
          return -1;
        }
        for (int32_t q = 0; q < o->len; ++q)
          if (o->c[q] < ne) *omap_ref(C, i, o->c[q]) -= v * o->v[q] / d;
        s->b_unscaled[i] -= v * s->b_full[col] / d;
      }
    }
  }
  return 0;
}
/* setup_unknown_scaling: she/rescaling.hpp:47-82 (kinetic-energy scaling) */
static void unknown_scaling(oracle_system *s) {
  s->scaling = (double *)malloc((size_t)(s->n_even > 0 ? s->n_even : 1) * sizeof(double));
  for (int64_t i = 0; i < s->n_even; ++i) s->scaling[i] = 1.0;
  if (!s->cfg.scale) return;
  for (int c = 0; c < s->ncells; ++c) {
    double T = s->T;
    for (int H = 0; H < s->NH; ++H) {
      int64_t base = s->idx_c[(int64_t)c * s->NH + H];
      if (base < 0) continue;
      int n = even_unknowns(s->ord_c[(int64_t)c * s->NH + H]);
      for (int i = 0; i < n; ++i) s->scaling[base + i] = exp(-kin_c(s, c, H) / (KB_ * T));
    }
  }
}
static void omap_clone(omap *d, const omap *a) {
  omap_init(d, a->n);
  for (int64_t i = 0; i < a->n; ++i) {
    orow *r = &d->r[i];
    r->len = r->cap = a->r[i].len;
    if (r->cap) {
      r->c = (int64_t *)malloc((size_t)r->cap * sizeof(int64_t));
      r->v = (double *)malloc((size_t)r->cap * sizeof(double));
      memcpy(r->c, a->r[i].c, (size_t)r->len * sizeof(int64_t));
      memcpy(r->v, a->r[i].v, (size_t)r->len * sizeof(double));
    }
  }
}
/* rescale_system: she/rescaling.hpp:89-107; row_normalize_system: math/linalg_util.hpp:200-282 */
static void normalise(oracle_system *s) {
  int64_t n = s->n_even;
  omap A;
  omap_clone(&A, &s->unscaled);
  s->b_red = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  memcpy(s->b_red, s->b_unscaled, (size_t)n * sizeof(double));
  if (s->cfg.scale)
    for (int64_t i = 0; i < n; ++i)
      for (int32_t k = 0; k < A.r[i].len; ++k) A.r[i].v[k] *= s->scaling[A.r[i].c[k]];
  s->colfac = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for (int64_t i = 0; i < n; ++i) s->colfac[i] = 1.0;
  omap At;
  omap_trans(&A, &At);
  omap_free(&A);
  for (int64_t i = 0; i < n; ++i) {
    orow *ci = &At.r[i];
    double nrm = 0.0;
    for (int32_t k = 0; k < ci->len; ++k) nrm += log(fabs(ci->v[k]));
    nrm = exp(nrm / (double)ci->len);
    if (*omap_ref(&At, i, i) < 0.0) nrm *= -1.0;
    ci = &At.r[i];
    for (int32_t k = 0; k < ci->len; ++k) ci->v[k] /= nrm;
    s->colfac[i] /= nrm;
  }
  omap_trans(&At, &s->red);
  omap_free(&At);
  for (int64_t i = 0; i < n; ++i) {
    orow *ri = &s->red.r[i];
    double nrm = 0.0;
    for (int32_t k = 0; k < ri->len; ++k) nrm += ri->v[k] * ri->v[k];
    nrm = sqrt(nrm);
    if (*omap_ref(&s->red, i, i) < 0.0) nrm *= -1.0;
    ri = &s->red.r[i];
    for (int32_t k = 0; k < ri->len; ++k) ri->v[k] /= nrm;
    s->b_red[i] /= nrm;
  }
}
/* fill_block_indices: she/linear_solver.hpp:42-83 */
static void block_indices(oracle_system *s) {
  int NH = s->NH;
  s->blocks = (int64_t *)calloc((size_t)2 * NH, sizeof(int64_t));
  int64_t last = 0;
  for (int H = 0; H < NH; ++H) {
    s->blocks[2 * H] = last;
    for (int c = 0; c < s->ncells; ++c) {
      int64_t u = s->idx_c[(int64_t)c * NH + H];
      if (u > -1) {
        last = u;
        s->blocks[2 * H] = last;
        if (H > 0) s->blocks[2 * (H - 1) + 1] = last;
        break;
      }
    }
    if (s->blocks[2 * H + 1] < s->blocks[2 * H]) s->blocks[2 * H + 1] = s->blocks[2 * H];
  }
  s->blocks[2 * (NH - 1) + 1] = s->n_even;
}

static int load_coupling(oracle_system *s, const oracle_coupling *cpl) {
  int L = s->cfg.L_max;
  if (L != 1 && L != 3 && L != 5 && L != 7) {
    set_err("expansion order %d has no precomputed coupling table", L);
    return -1;
  }
  s->Lc = L;
  int K = (L + 1) * (L + 1);
  for (int d = 0; d < 3; ++d) { cmat_init(&s->a[d], K); cmat_init(&s->b[d], K); }
  for (int t = 0; t < 6; ++t) {
    cmat *m = t < 3 ? &s->a[t] : &s->b[t - 3];
    for (int k = 0; k < cpl->nnz[t]; ++k) {
      int i = cpl->row[t][k], j = cpl->col[t][k];
      if (i < 0 || j < 0 || i >= K || j >= K) { set_err("coupling entry out of range"); return -1; }
      m->p[i * K + j] = 1;
      m->v[i * K + j] = cpl->val[t][k];
    }
  }
  return 0;
}

oracle_system *oracle_build(const oracle_device *dev, const oracle_config *cfg, const oracle_coupling *cpl) {
  g_err[0] = 0;
  if (dev->dim != 1 && dev->dim != 2) { set_err("dim must be 1 or 2"); return NULL; }
  if (dev->nx < 2 || (dev->dim == 2 && dev->ny < 2)) { set_err("mesh too small"); return NULL; }
  oracle_system *s = (oracle_system *)calloc(1, sizeof(oracle_system));
  s->cfg = *cfg;
  if (load_coupling(s, cpl) || setup_geometry(s, dev) || setup_energies(s)) { oracle_free(s); return NULL; }
  distribute_orders(s);
  boundary_values(s);
  create_mapping(s);
  if (assemble(s) || diagonalise(s) || eliminate(s)) { oracle_free(s); return NULL; }
  unknown_scaling(s);
  normalise(s);
  block_indices(s);
  return s;
}

void oracle_free(oracle_system *s) {
  if (!s) return;
  free(s->cx); free(s->cy); free(s->cvol); free(s->fc1); free(s->fc2); free(s->fdir); free(s->fvol);
  free(s->cf); free(s->dn); free(s->dp); free(s->psi); free(s->contact); free(s->Hval);
  free(s->shift_c); free(s->shift_f); free(s->ord_c); free(s->ord_f); free(s->idx_c); free(s->idx_f);
  free(s->bnd); free(s->b_full); free(s->b_unscaled); free(s->b_red); free(s->scaling); free(s->colfac);
  free(s->blocks);
  omap_free(&s->full); omap_free(&s->unscaled); omap_free(&s->red);
  if (s->a[0].p) for (int d = 0; d < 3; ++d) { cmat_free(&s->a[d]); cmat_free(&s->b[d]); }
  free(s);
}

void oracle_sizes(const oracle_system *s, int64_t *o) {
  o[0] = s->NH; o[1] = s->n_even; o[2] = s->n_odd; o[3] = omap_nnz(&s->full);
  o[4] = omap_nnz(&s->red); o[5] = s->ncells; o[6] = s->nfacets;
  int64_t z = 0;
  for (int64_t i = 0; i < s->red.n; ++i)
    for (int32_t k = 0; k < s->red.r[i].len; ++k) z += s->red.r[i].v[k] == 0.0;
  o[7] = z;
}
void oracle_mapping(const oracle_system *s, int64_t *ei, int32_t *eo, int64_t *oi, int32_t *oo) {
  size_t nc = (size_t)s->ncells * s->NH, nf = (size_t)s->nfacets * s->NH;
  memcpy(ei, s->idx_c, nc * sizeof(int64_t));
  memcpy(eo, s->ord_c, nc * sizeof(int32_t));
  if (nf) { memcpy(oi, s->idx_f, nf * sizeof(int64_t)); memcpy(oo, s->ord_f, nf * sizeof(int32_t)); }
}
void oracle_facets(const oracle_system *s, int32_t *c1, int32_t *c2, int32_t *dir) {
  memcpy(c1, s->fc1, (size_t)s->nfacets * sizeof(int32_t));
  memcpy(c2, s->fc2, (size_t)s->nfacets * sizeof(int32_t));
  memcpy(dir, s->fdir, (size_t)s->nfacets * sizeof(int32_t));
}
void oracle_energies(const oracle_system *s, double *H, double *sh) {
  memcpy(H, s->Hval, (size_t)s->NH * sizeof(double));
  memcpy(sh, s->shift_c, (size_t)s->ncells * sizeof(double));
}
static void export_csr(const omap *m, int64_t *rp, int32_t *ci, double *v) {
  int64_t p = 0;
  rp[0] = 0;
  for (int64_t i = 0; i < m->n; ++i) {
    for (int32_t k = 0; k < m->r[i].len; ++k) { ci[p] = (int32_t)m->r[i].c[k]; v[p] = m->r[i].v[k]; ++p; }
    rp[i + 1] = p;
  }
}
void oracle_reduced(const oracle_system *s, int64_t *rp, int32_t *ci, double *v, double *rhs, double *sc, double *cf) {
  export_csr(&s->red, rp, ci, v);
  memcpy(rhs, s->b_red, (size_t)s->n_even * sizeof(double));
  memcpy(sc, s->scaling, (size_t)s->n_even * sizeof(double));
  memcpy(cf, s->colfac, (size_t)s->n_even * sizeof(double));
}
int64_t oracle_unscaled_nnz(const oracle_system *s) { return omap_nnz(&s->unscaled); }
void oracle_unscaled(const oracle_system *s, int64_t *rp, int32_t *ci, double *v, double *rhs) {
  export_csr(&s->unscaled, rp, ci, v);
  memcpy(rhs, s->b_unscaled, (size_t)s->n_even * sizeof(double));
}
void oracle_full(const oracle_system *s, int64_t *rp, int32_t *ci, double *v, double *rhs) {
  export_csr(&s->full, rp, ci, v);
  memcpy(rhs, s->b_full, (size_t)s->N * sizeof(double));
}
void oracle_block_ranges(const oracle_system *s, int64_t *r) { memcpy(r, s->blocks, (size_t)2 * s->NH * sizeof(int64_t)); }

/* she::solve tail (linear_solver.hpp:169-187) + recover_odd_unknowns (eliminate.hpp:305-354) */
double oracle_finish(const oracle_system *s, const double *xr, double *xf) {
  int64_t ne = s->n_even, N = s->N;
  for (int64_t i = 0; i < ne; ++i) xf[i] = s->cfg.scale ? xr[i] * (s->scaling[i] * s->colfac[i]) : xr[i];
  for (int64_t i = ne; i < N; ++i) {
    const orow *r = &s->full.r[i];
    double d = 0.0;
    xf[i] = s->b_full[i];
    for (int32_t k = 0; k < r->len; ++k) {
      if (r->c[k] < ne) xf[i] -= r->v[k] * xf[r->c[k]];
      else if (r->c[k] == i) d = r->v[k];
    }
    xf[i] /= d;
  }
  double num = 0.0, den = 0.0;
  for (int64_t i = 0; i < N; ++i) {
    const orow *r = &s->full.r[i];
    double v = 0.0;
    for (int32_t k = 0; k < r->len; ++k) v += r->v[k] * xf[r->c[k]];
    double d = v - s->b_full[i];
    num += d * d;
    den += s->b_full[i] * s->b_full[i];
  }
  return sqrt(num) / sqrt(den);
}

/* carrier_density_wrapper_by_reference::operator() (DF): she/postproc/carrier_density.hpp:68-131,
 * window [0, 1.0]; current_on_facet_by_ref_calculator: she/postproc/current_density.hpp:48-131 */
void oracle_postproc(const oracle_system *s, const double *xf, double *dens, double *cur) {
  int NH = s->NH, ct = s->cfg.carrier;
  for (int c = 0; c < s->ncells; ++c) {
    double n = 0;
    for (int H = 1; H < NH - 1; ++H) {
      int64_t u = s->idx_c[(int64_t)c * NH + H];
      if (u < 0) continue;
      double lo = (kin_c(s, c, H - 1) + kin_c(s, c, H)) / 2.0;
      if (lo < 0.0) lo = 0.0;
      double up = (kin_c(s, c, H + 1) + kin_c(s, c, H)) / 2.0;
      if (up > 1.0) up = 1.0;
      if (up >= 0 && lo < up) n += xf[u] * avg_dos(s, 0, c, H) * box_height(s, 0, c, H);
    }
    dens[c] = n * symmetry_factor(ct);
  }
  /* L = 1 coupling entries a_x(0,3) and a_y(0,1) (the only l=0 -> l=1 entries) */
  const double ax03 = cmat_get(&s->a[0], 0, 3), ay01 = cmat_get(&s->a[1], 0, 1);
  const double ax01 = cmat_get(&s->a[0], 0, 1), ax02 = cmat_get(&s->a[0], 0, 2);
  const double ay02 = cmat_get(&s->a[1], 0, 2), ay03 = cmat_get(&s->a[1], 0, 3);
  double mt = m_trans(ct), md = m_dos(ct);
  double pol = ct == 0 ? -1.0 : 1.0;
  double Y00 = sqrt((2.0 * 0 + 1.0) * 1.0 / (2.0 * PI_ * 1.0)) * 1.0 / sqrt(2.0);
  for (int f = 0; f < s->nfacets; ++f) {
    int c1 = s->fc1[f], c2 = s->fc2[f];
    double nx = s->cx[c2] - s->cx[c1], ny = s->dim > 1 ? s->cy[c2] - s->cy[c1] : 0.0;
    double nn = sqrt(nx * nx + ny * ny);
    nx /= nn; ny /= nn;
    double vn = 0;
    for (int H = 1; H < NH - 1; ++H) {
      int64_t u = s->idx_f[(int64_t)f * NH + H];
      if (u < 0) continue;
      double factor = integral_vZ(s, 1, f, H) * symmetry_factor(ct);
      vn += (ax01 * xf[u] + ax02 * xf[u + 1] + ax03 * xf[u + 2]) * factor * sqrt(md / mt) * nx;
      if (s->dim > 1) vn += (ay01 * xf[u] + ay02 * xf[u + 1] + ay03 * xf[u + 2]) * factor * sqrt(md / mt) * ny;
    }
    cur[f] = -pol * Q_ * vn / Y00;
  }
}

/* ------------------------------------------------------------------ */
/* CSR solvers                                                         */
/* ------------------------------------------------------------------ */
struct oracle_precond {
  int kind;
  int64_t n;
  int64_t *rp;     /* factor pattern */
  int32_t *ci;
  double *v;
  int64_t *diag;   /* position of the diagonal in each row */
  int nseg;        /* block ILU0 whose ranges tile [0, n): the blocks' starts seg[0..nseg), seg[nseg] = n */
  int64_t *seg;
};

void oracle_dispersion(int carrier, double eps, double *Z, double *v) {
  *Z = dos(carrier, eps);
  *v = velocity(carrier, eps);
}

/* Rows are independent: with par the row loop runs on all OpenMP threads (the
 * parallel CPU baseline), otherwise in order on the calling thread. */
static void spmv_par(int par, int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *x,
                     double *y) {
#pragma omp parallel for if (par) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
    y[i] = s;
  }
}

void oracle_spmv(int64_t n, const int64_t *rp, const int32_t *ci, const double *v, const double *x, double *y) {
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[ci[k]];
    y[i] = s;
  }
}

/* ILU(0), row-wise IKJ variant: Saad 2nd ed. Alg. 10.4, on the pattern held in
 * p (A's pattern for ILU0, the level-of-fill pattern for ILU(k)).  drop > 0
 * applies ILUT's dual threshold rule (Saad Alg. 10.6, lines 5 and 10) inside
 * the fixed pattern: a multiplier |l_ik| < drop * ||a_i||_2 is set to zero and
 * not used, and so is an upper entry below the same bound after the row is
 * done; dropped entries stay in the pattern as zeros. */
static int ilu0_factor_drop(oracle_precond *p, double drop);
static int ilu0_factor(oracle_precond *p) { return ilu0_factor_drop(p, 0.0); }
/* Rows [r0, r1) of the factorisation; pos (column -> position in the current
 * row, -1 elsewhere) is left as it was found.  Rows of different blocks of a
 * block ILU0 touch disjoint columns, so blocks may run concurrently on one pos. */
static int ilu0_rows(oracle_precond *p, int64_t r0, int64_t r1, double drop, int64_t *pos) {
  for (int64_t i = r0; i < r1; ++i) {
    double rn = 0.0;
    for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) {
      pos[p->ci[k]] = k;
      rn += p->v[k] * p->v[k];
    }
    const double tau = drop * sqrt(rn);
    p->diag[i] = pos[i];
    if (pos[i] < 0) {
      for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) pos[p->ci[k]] = -1;
      return (int)-1;
    }
    for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) {
      int64_t kc = p->ci[k];
      if (kc >= i) break;
      double piv = p->v[p->diag[kc]];
      p->v[k] /= piv;
      if (drop > 0.0 && fabs(p->v[k]) < tau) { p->v[k] = 0.0; continue; }
      double lik = p->v[k];
      for (int64_t q = p->diag[kc] + 1; q < p->rp[kc + 1]; ++q) {
        int64_t j = pos[p->ci[q]];
        if (j >= 0) p->v[j] -= lik * p->v[q];
      }
    }
    if (drop > 0.0)
      for (int64_t k = p->diag[i] + 1; k < p->rp[i + 1]; ++k)
        if (fabs(p->v[k]) < tau) p->v[k] = 0.0;
    for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) pos[p->ci[k]] = -1;
  }
  return 0;
}

static int64_t *ilu0_begin(oracle_precond *p) {
  const int64_t n = p->n;
  int64_t *pos = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  for (int64_t i = 0; i < n; ++i) pos[i] = -1;
  p->diag = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  return pos;
}

static int ilu0_factor_drop(oracle_precond *p, double drop) {
  int64_t *pos = ilu0_begin(p);
  int e = ilu0_rows(p, 0, p->n, drop, pos);
  free(pos);
  if (e) set_err("ILU0: a row has no diagonal entry");
  return e;
}

/* The same factorisation with the blocks of a block ILU0 (p->seg) on all
 * OpenMP threads; every row computes exactly as in ilu0_factor_drop. */
static int ilu0_factor_blocks_parallel(oracle_precond *p) {
  int64_t *pos = ilu0_begin(p);
  int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
  for (int b = 0; b < p->nseg; ++b) err |= ilu0_rows(p, p->seg[b], p->seg[b + 1], 0.0, pos) != 0;
  free(pos);
  if (err) set_err("ILU0: a row has no diagonal entry");
  return err ? -1 : 0;
}

/* Symbolic ILU(k): level-of-fill pattern (Saad 2nd ed. Alg. 10.5).  A's
 * entries have level 0; eliminating with row j gives column c the level
 * lev(i,j) + lev(j,c) + 1, the minimum over all paths; columns with level <= k
 * are kept.  Rows are processed in order, the lower part of each row in
 * ascending column order.  Output: row pointer and columns (ascending). */
static int iluk_symbolic(int64_t n, const int64_t *rp, const int32_t *ci, int k, int64_t **out_rp,
                         int32_t **out_ci) {
  int *lev = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  int32_t *lst = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  for (int64_t i = 0; i < n; ++i) lev[i] = -1;
  int64_t cap = rp[n] * 2 + 16, used = 0;
  int64_t *orp = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
  int32_t *oci = (int32_t *)malloc((size_t)cap * sizeof(int32_t));
  unsigned char *olev = (unsigned char *)malloc((size_t)cap);
  int64_t *odg = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  orp[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t nl = 0;
    for (int64_t q = rp[i]; q < rp[i + 1]; ++q) {
      lev[ci[q]] = 0;
      lst[nl++] = ci[q];
    }
    /* ascending processing of the lower part: selection of the next smallest */
    int32_t done = -1;
    for (;;) {
      int32_t j = (int32_t)i;
      for (int32_t t = 0; t < nl; ++t)
        if (lst[t] > done && lst[t] < j) j = lst[t];
      if (j >= i) break;
      done = j;
      const int lj = lev[j];
      for (int64_t q = odg[j] + 1; q < orp[j + 1]; ++q) {
        const int32_t c = oci[q];
        const int nlv = lj + olev[q] + 1;
        if (nlv > k) continue;
        if (lev[c] < 0) { lev[c] = nlv; lst[nl++] = c; }
        else if (nlv < lev[c]) lev[c] = nlv;
      }
    }
    /* sort the row's columns */
    for (int32_t a = 1; a < nl; ++a) {
      int32_t v = lst[a], b = a - 1;
      while (b >= 0 && lst[b] > v) { lst[b + 1] = lst[b]; --b; }
      lst[b + 1] = v;
    }
    if (used + nl > cap) {
      cap = (used + nl) * 2;
      oci = (int32_t *)realloc(oci, (size_t)cap * sizeof(int32_t));
      olev = (unsigned char *)realloc(olev, (size_t)cap);
    }
    odg[i] = -1;
    for (int32_t t = 0; t < nl; ++t) {
      if (lst[t] == i) odg[i] = used;
      oci[used] = lst[t];
      olev[used] = (unsigned char)lev[lst[t]];
      ++used;
      lev[lst[t]] = -1;
    }
    if (odg[i] < 0) {
      free(lev); free(lst); free(orp); free(oci); free(olev); free(odg);
      set_err("ILU(k): missing diagonal in row %lld", (long long)i);
      return -1;
    }
    orp[i + 1] = used;
  }
  free(lev); free(lst); free(olev); free(odg);
  *out_rp = orp;
  *out_ci = oci;
  return 0;
}

/* ILUT(p, tau): Saad 2nd ed. Alg. 10.6 (dual dropping), diagonal always kept */
typedef struct { int32_t j; double v; } jv;
static int cmp_abs_desc(const void *a, const void *b) {
  double x = fabs(((const jv *)a)->v), y = fabs(((const jv *)b)->v);
  return x < y ? 1 : (x > y ? -1 : (((const jv *)a)->j - ((const jv *)b)->j));
}
static int cmp_col(const void *a, const void *b) { return ((const jv *)a)->j - ((const jv *)b)->j; }
static int ilut_factor(oracle_precond *p, int64_t n, const int64_t *rp, const int32_t *ci, const double *val,
                       double drop, int fill) {
  double *w = (double *)calloc((size_t)n, sizeof(double));
  char *nz = (char *)calloc((size_t)n, 1);
  int32_t *lst = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  jv *buf = (jv *)malloc((size_t)(2 * n + 2) * sizeof(jv));
  int64_t cap = rp[n] * 2 + n + 16, used = 0;
  p->rp = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
  p->ci = (int32_t *)malloc((size_t)cap * sizeof(int32_t));
  p->v = (double *)malloc((size_t)cap * sizeof(double));
  p->diag = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  p->rp[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t nl = 0;
    double rn = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      int32_t j = ci[k];
      w[j] = val[k];
      rn += val[k] * val[k];
      if (!nz[j]) { nz[j] = 1; lst[nl++] = j; }
    }
    double tau_i = drop * sqrt(rn);
    /* eliminate in ascending column order: repeatedly take the smallest unprocessed k < i */
    int32_t done_upto = -1;
    for (;;) {
      int32_t kmin = (int32_t)i;
      for (int32_t t = 0; t < nl; ++t)
        if (lst[t] < kmin && lst[t] > done_upto && nz[lst[t]]) kmin = lst[t];
      if (kmin >= i) break;
      done_upto = kmin;
      double wk = w[kmin] / p->v[p->diag[kmin]];
      if (fabs(wk) < tau_i) { w[kmin] = 0.0; continue; }
      w[kmin] = wk;
      for (int64_t q = p->diag[kmin] + 1; q < p->rp[kmin + 1]; ++q) {
        int32_t j = p->ci[q];
        if (!nz[j]) { nz[j] = 1; lst[nl++] = j; w[j] = 0.0; }
        w[j] -= wk * p->v[q];
      }
    }
    int32_t nlw = 0, nuw = 0;
    jv *L = buf, *U = buf + nl;
    double dval = w[i];
    for (int32_t t = 0; t < nl; ++t) {
      int32_t j = lst[t];
      if (j == i) continue;
      if (w[j] == 0.0 || fabs(w[j]) < tau_i) continue;
      if (j < i) { L[nlw].j = j; L[nlw].v = w[j]; ++nlw; }
    }
    for (int32_t t = 0; t < nl; ++t) {
      int32_t j = lst[t];
      if (j <= i) continue;
      if (w[j] == 0.0 || fabs(w[j]) < tau_i) continue;
      U[nuw].j = j; U[nuw].v = w[j]; ++nuw;
    }
    if (nlw > fill) { qsort(L, nlw, sizeof(jv), cmp_abs_desc); nlw = fill; }
    if (nuw > fill) { qsort(U, nuw, sizeof(jv), cmp_abs_desc); nuw = fill; }
    qsort(L, nlw, sizeof(jv), cmp_col);
    qsort(U, nuw, sizeof(jv), cmp_col);
    if (used + nlw + nuw + 1 > cap) {
      cap = (used + nlw + nuw + 1) * 2;
      p->ci = (int32_t *)realloc(p->ci, (size_t)cap * sizeof(int32_t));
      p->v = (double *)realloc(p->v, (size_t)cap * sizeof(double));
    }
    for (int32_t t = 0; t < nlw; ++t) { p->ci[used] = L[t].j; p->v[used] = L[t].v; ++used; }
    if (dval == 0.0) dval = tau_i > 0 ? tau_i : 1e-4;
    p->diag[i] = used;
    p->ci[used] = (int32_t)i; p->v[used] = dval; ++used;
    for (int32_t t = 0; t < nuw; ++t) { p->ci[used] = U[t].j; p->v[used] = U[t].v; ++used; }
    p->rp[i + 1] = used;
    for (int32_t t = 0; t < nl; ++t) { nz[lst[t]] = 0; w[lst[t]] = 0.0; }
  }
  free(w); free(nz); free(lst); free(buf);
  return 0;
}

static oracle_precond *precond_create_impl(int parallel, int kind, int64_t n, const int64_t *rp, const int32_t *ci,
                                           const double *val, double drop, int fill, int nblocks,
                                           const int64_t *ranges) {
  oracle_precond *p = (oracle_precond *)calloc(1, sizeof(oracle_precond));
  p->kind = kind;
  p->n = n;
  if (kind == ORACLE_PC_NONE) return p;
  if (kind == ORACLE_PC_ILUT) {
    if (ilut_factor(p, n, rp, ci, val, drop, fill)) { oracle_precond_free(p); return NULL; }
    return p;
  }
  if (kind == ORACLE_PC_ILUK) {
    /* fill = level k; A's values scattered into the level-k pattern */
    if (iluk_symbolic(n, rp, ci, fill, &p->rp, &p->ci)) { oracle_precond_free(p); return NULL; }
    p->v = (double *)calloc((size_t)(p->rp[n] > 0 ? p->rp[n] : 1), sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
      int64_t q = p->rp[i];
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
        while (p->ci[q] < ci[k]) ++q;
        p->v[q] = val[k];
      }
    }
    if (ilu0_factor_drop(p, drop)) { oracle_precond_free(p); return NULL; }
    return p;
  }
  /* ILU0 on A, or on the block-diagonal part of A for block ILU0 */
  int64_t *blk = NULL;
  if (kind == ORACLE_PC_BLOCK_ILU0) {
    blk = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) blk[i] = -1;
    for (int b = 0; b < nblocks; ++b)
      for (int64_t i = ranges[2 * b]; i < ranges[2 * b + 1] && i < n; ++i)
        if (blk[i] < 0) blk[i] = b;
    for (int64_t i = 0; i < n; ++i) if (blk[i] < 0) blk[i] = nblocks; /* uncovered rows form their own block */
  }
  p->rp = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
  p->ci = (int32_t *)malloc((size_t)(rp[n] > 0 ? rp[n] : 1) * sizeof(int32_t));
  p->v = (double *)malloc((size_t)(rp[n] > 0 ? rp[n] : 1) * sizeof(double));
  int64_t q = 0;
  p->rp[0] = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      if (blk && blk[ci[k]] != blk[i]) continue;
      p->ci[q] = ci[k];
      p->v[q] = val[k];
      ++q;
    }
    p->rp[i + 1] = q;
  }
  free(blk);
  if (kind == ORACLE_PC_BLOCK_ILU0 && nblocks > 0) {
    /* nonempty ranges that tile [0, n) in order (empty ones, energy levels
     * without unknowns, are skipped): the blocks apply independently
     * (parallel_linear_solver.hpp:79-93 applies them with OpenMP) */
    int64_t *st = (int64_t *)malloc((size_t)(nblocks + 1) * sizeof(int64_t));
    int m = 0, tiled = 1;
    int64_t at = 0;
    for (int b = 0; b < nblocks && tiled; ++b) {
      const int64_t f = ranges[2 * b], l = ranges[2 * b + 1];
      if (l == f) continue;
      if (f != at || l < f) tiled = 0;
      else { st[m++] = f; at = l; }
    }
    if (tiled && m > 0 && at == n) {
      st[m] = n;
      p->nseg = m;
      p->seg = st;
    } else {
      free(st);
    }
  }
  if (parallel && p->nseg > 0 ? ilu0_factor_blocks_parallel(p) : ilu0_factor(p)) {
    oracle_precond_free(p);
    return NULL;
  }
  return p;
}

/* Forward then backward sweep over rows [r0, r1) of a factor with no entries
 * leaving that range (a block of block ILU0). */
static void precond_apply_range(const oracle_precond *p, int64_t r0, int64_t r1, const double *in, double *out) {
  for (int64_t i = r0; i < r1; ++i) {
    double s = in[i];
    for (int64_t k = p->rp[i]; k < p->diag[i]; ++k) s -= p->v[k] * out[p->ci[k]];
    out[i] = s;
  }
  for (int64_t i = r1 - 1; i >= r0; --i) {
    double s = out[i];
    for (int64_t k = p->diag[i] + 1; k < p->rp[i + 1]; ++k) s -= p->v[k] * out[p->ci[k]];
    out[i] = s / p->v[p->diag[i]];
  }
}

oracle_precond *oracle_precond_create(int kind, int64_t n, const int64_t *rp, const int32_t *ci, const double *val,
                                      double drop, int fill, int nblocks, const int64_t *ranges) {
  return precond_create_impl(0, kind, n, rp, ci, val, drop, fill, nblocks, ranges);
}
oracle_precond *oracle_precond_create_parallel(int kind, int64_t n, const int64_t *rp, const int32_t *ci,
                                               const double *val, double drop, int fill, int nblocks,
                                               const int64_t *ranges) {
  return precond_create_impl(1, kind, n, rp, ci, val, drop, fill, nblocks, ranges);
}

static void precond_apply_n(const oracle_precond *p, int64_t n, const double *in, double *out) {
  if (!p || p->kind == ORACLE_PC_NONE) { memcpy(out, in, (size_t)n * sizeof(double)); return; }
  for (int64_t i = 0; i < n; ++i) {
    double s = in[i];
    for (int64_t k = p->rp[i]; k < p->diag[i]; ++k) s -= p->v[k] * out[p->ci[k]];
    out[i] = s;
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    double s = out[i];
    for (int64_t k = p->diag[i] + 1; k < p->rp[i + 1]; ++k) s -= p->v[k] * out[p->ci[k]];
    out[i] = s / p->v[p->diag[i]];
  }
}
void oracle_precond_apply(const oracle_precond *p, const double *in, double *out) { precond_apply_n(p, p->n, in, out); }
int64_t oracle_precond_nnz(const oracle_precond *p) { return p->rp ? p->rp[p->n] : 0; }
void oracle_precond_factors(const oracle_precond *p, int64_t *rp, int32_t *ci, double *v) {
  if (!p->rp) return;
  memcpy(rp, p->rp, (size_t)(p->n + 1) * sizeof(int64_t));
  memcpy(ci, p->ci, (size_t)p->rp[p->n] * sizeof(int32_t));
  memcpy(v, p->v, (size_t)p->rp[p->n] * sizeof(double));
}
void oracle_precond_free(oracle_precond *p) {
  if (!p) return;
  free(p->rp); free(p->ci); free(p->v); free(p->diag); free(p->seg);
  free(p);
}

static double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
static double dot_par(int par, int64_t n, const double *a, const double *b) {
  if (!par) return dot(n, a, b);
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
/* the preconditioner application of the parallel baseline: the blocks of a
 * block ILU0 in parallel (same values as the sequential sweep, no entry leaves
 * a block); anything else in order */
static void precond_apply_par(int par, const oracle_precond *p, int64_t n, const double *in, double *out) {
  if (par && p && p->kind == ORACLE_PC_BLOCK_ILU0 && p->nseg > 0) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < p->nseg; ++b) precond_apply_range(p, p->seg[b], p->seg[b + 1], in, out);
    return;
  }
  precond_apply_n(p, n, in, out);
}

int oracle_precond_blocks(const oracle_precond *p) { return p ? p->nseg : 0; }

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

/* Right-preconditioned BiCGStab (van der Vorst 1992) as the reference calls it
 * (serial_linear_solver.hpp:129-136).  With par the vector work, the SpMV and
 * the block ILU0 application run on all OpenMP threads (the parallel CPU
 * baseline; the same steps, dot products summed in another order). */
static int bicgstab_impl(int par, int64_t n, const int64_t *rp, const int32_t *ci, const double *val, const double *b,
                         double *x, const oracle_precond *p, double tol, int max_iters, double max_seconds,
                         double *rel_res) {
  double t0 = now_s();
  size_t sz = (size_t)(n > 0 ? n : 1) * sizeof(double);
  double *r = (double *)malloc(sz), *r0 = (double *)malloc(sz), *pp = (double *)calloc(1, sz),
         *v = (double *)calloc(1, sz), *ph = (double *)malloc(sz), *sv = (double *)malloc(sz),
         *sh = (double *)malloc(sz), *t = (double *)malloc(sz);
  double nb = sqrt(dot_par(par, n, b, b));
  int it = 0;
  if (nb == 0.0) {
    /* solve_impl: zero right-hand side gives the zero solution (src/solvers/solve.hpp:61-66) */
    for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
    *rel_res = 0.0;
    goto done;
  }
  spmv_par(par, n, rp, ci, val, x, r);
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
  memcpy(r0, r, sz);
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  double res = sqrt(dot_par(par, n, r, r)) / nb;
  while (res > tol && it < max_iters) {
    if (max_seconds > 0 && now_s() - t0 > max_seconds) break;
    ++it;
    double rho_new = dot_par(par, n, r0, r);
    if (rho_new == 0.0 || omega == 0.0) {
      /* breakdown: restart from the current residual */
      memcpy(r0, r, sz);
      for (int64_t i = 0; i < n; ++i) { pp[i] = 0.0; v[i] = 0.0; }
      rho = alpha = omega = 1.0;
      rho_new = dot_par(par, n, r0, r);
    }
    double beta = (rho_new / rho) * (alpha / omega);
#pragma omp parallel for if (par) schedule(static)
    for (int64_t i = 0; i < n; ++i) pp[i] = r[i] + beta * (pp[i] - omega * v[i]);
    precond_apply_par(par, p, n, pp, ph);
    spmv_par(par, n, rp, ci, val, ph, v);
    alpha = rho_new / dot_par(par, n, r0, v);
#pragma omp parallel for if (par) schedule(static)
    for (int64_t i = 0; i < n; ++i) sv[i] = r[i] - alpha * v[i];
    double sres = sqrt(dot_par(par, n, sv, sv)) / nb;
    if (sres <= tol) {
      for (int64_t i = 0; i < n; ++i) x[i] += alpha * ph[i];
      spmv_par(par, n, rp, ci, val, x, r);
      for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
      res = sqrt(dot_par(par, n, r, r)) / nb;
      if (res <= tol) break;
      memcpy(r0, r, sz);
      for (int64_t i = 0; i < n; ++i) { pp[i] = 0.0; v[i] = 0.0; }
      rho = alpha = omega = 1.0;
      continue;
    }
    precond_apply_par(par, p, n, sv, sh);
    spmv_par(par, n, rp, ci, val, sh, t);
    double tt = dot_par(par, n, t, t);
    omega = tt > 0 ? dot_par(par, n, t, sv) / tt : 0.0;
#pragma omp parallel for if (par) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = sv[i] - omega * t[i];
    }
    rho = rho_new;
    res = sqrt(dot_par(par, n, r, r)) / nb;
    if (res <= tol) {
      /* confirm with the explicit residual */
      spmv_par(par, n, rp, ci, val, x, r);
      for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
      res = sqrt(dot_par(par, n, r, r)) / nb;
      if (res > tol) {
        memcpy(r0, r, sz);
        for (int64_t i = 0; i < n; ++i) { pp[i] = 0.0; v[i] = 0.0; }
        rho = alpha = omega = 1.0;
      }
    }
  }
  spmv_par(par, n, rp, ci, val, x, r);
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
  *rel_res = sqrt(dot_par(par, n, r, r)) / nb;
done:
  free(r); free(r0); free(pp); free(v); free(ph); free(sv); free(sh); free(t);
  return it;
}

int oracle_bicgstab(int64_t n, const int64_t *rp, const int32_t *ci, const double *val, const double *b, double *x,
                    const oracle_precond *p, double tol, int max_iters, double max_seconds, double *rel_res) {
  return bicgstab_impl(0, n, rp, ci, val, b, x, p, tol, max_iters, max_seconds, rel_res);
}

int oracle_bicgstab_parallel(int64_t n, const int64_t *rp, const int32_t *ci, const double *val, const double *b,
                             double *x, const oracle_precond *p, double tol, int max_iters, double max_seconds,
                             double *rel_res) {
  return bicgstab_impl(1, n, rp, ci, val, b, x, p, tol, max_iters, max_seconds, rel_res);
}
