/*
 * she_oracle.h -- CPU FP64 restatement of the ViennaSHE SHE linear-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * implementation: only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark may load it.  The product never links or calls it.
 *
 * It restates, on structured 1D line and 2D quadrilateral meshes:
 *   - energy grid / band-edge shift       she/setup_energies.hpp:61-153
 *   - expansion orders                    she/expansion_order.hpp:60-364
 *   - contact boundary values             she/boundary_conditions.hpp:57-136
 *   - even/odd unknown mapping            mapping.hpp:97-329
 *   - coefficient fill (boundary, scattering, streaming)
 *                                         she/assemble_all.hpp:46-424 and the files it calls
 *   - odd elimination, unknown scaling, row normalisation, block ranges
 *                                         she/linear_solver.hpp:42-189, she/eliminate.hpp,
 *                                         she/rescaling.hpp, math/linalg_util.hpp:200-282
 *   - ILU0 / ILUT / block-ILU0 + BiCGStab (ViennaCL, not in the reference tree;
 *     restated from Saad, "Iterative Methods for Sparse Linear Systems", 2nd ed.,
 *     Alg. 10.4 (ILU0), Alg. 10.6 (ILUT) and van der Vorst 1992 (BiCGStab)),
 *     and the same block-ILU0 + BiCGStab over all OpenMP threads, the CPU
 *     baseline              src/solvers/viennacl/parallel_linear_solver.hpp:64-114
 *   - odd recovery, carrier density, facet current
 *                                         she/eliminate.hpp:305-354,
 *                                         she/postproc/carrier_density.hpp:68-131,
 *                                         she/postproc/current_density.hpp:48-131
 *
 * The sparse matrix of the reference is a vector of std::map rows
 * (math/linalg_util.hpp:67-196); the restatement keeps the same
 * insert-on-access semantics so that the sparsity pattern it produces is the
 * reference's, entry for entry.
 */
#ifndef SHE_ORACLE_H
#define SHE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Coupling table in coordinate form: entries (row, col, value) for one of
 * a_x, a_y, a_z, b_x, b_y, b_z (harmonics_coupling.hpp:281-1046). */
typedef struct {
  int nnz[6];
  const int32_t *row[6];
  const int32_t *col[6];
  const double *val[6];
} oracle_coupling;

typedef struct {
  int dim;                 /* 1 or 2 */
  int nx, ny;              /* cells per direction (ny = 1 in 1D) */
  const double *xv;        /* nx+1 vertex x coordinates [m] */
  const double *yv;        /* ny+1 vertex y coordinates [m] (2D) */
  const double *doping_n;  /* per cell [m^-3], cell id = j*nx + i */
  const double *doping_p;
  const double *potential; /* per cell [V] */
  const int32_t *contact;  /* per cell: 1 if a contact potential is set */
  double temperature;      /* lattice temperature [K] */
} oracle_device;

typedef struct {
  int carrier;             /* 0 electrons, 1 holes */
  int L_max;               /* maximum expansion order */
  double energy_spacing;   /* [J] */
  double min_kin_range;    /* [J] */
  double max_kin_range;    /* [J] */
  double tau;              /* G/R boundary rate parameter [s] */
  int acoustic;            /* acoustic phonon scattering on/off */
  int optical;             /* optical phonon scattering on/off */
  int scale;               /* unknown scaling exp(-E_kin/kT) on/off */
} oracle_config;

typedef struct oracle_system oracle_system;

/* Builds the full (even+odd) system and the reduced, scaled, normalised
 * system exactly as she::assemble + she::solve do before the solver call.
 * Returns NULL on failure; oracle_last_error() then says why. */
oracle_system *oracle_build(const oracle_device *dev, const oracle_config *cfg,
                            const oracle_coupling *cpl);
void oracle_free(oracle_system *s);
const char *oracle_last_error(void);

/* sizes: [0] N_H, [1] n_even, [2] n_odd, [3] nnz_full, [4] nnz_reduced,
 *        [5] n_cells, [6] n_facets, [7] explicit zeros in the reduced matrix */
void oracle_sizes(const oracle_system *s, int64_t *out8);

/* Unknown mapping: even index per (cell, H) [n_cells*N_H], odd index per
 * (facet, H) [n_facets*N_H], orders likewise (-1 where no unknown). */
void oracle_mapping(const oracle_system *s, int64_t *even_idx, int32_t *even_order,
                    int64_t *odd_idx, int32_t *odd_order);
/* Facet table: c1, c2 (cell ids) and direction per interior facet. */
void oracle_facets(const oracle_system *s, int32_t *c1, int32_t *c2, int32_t *dir);
/* Per-cell band-edge shift and H values. */
void oracle_energies(const oracle_system *s, double *H_values, double *cell_shift);

/* Reduced system in CSR (ascending columns): the matrix handed to
 * solvers::solve, its right-hand side, the unknown scaling vector and the
 * column factors returned by row_normalize_system. */
void oracle_reduced(const oracle_system *s, int64_t *row_ptr, int32_t *col, double *val,
                    double *rhs, double *unknown_scaling, double *col_factors);
/* Reduced system before scaling and normalisation (CSR, same pattern rules as
 * eliminate_odd_unknowns) and its right-hand side. */
int64_t oracle_unscaled_nnz(const oracle_system *s);
void oracle_unscaled(const oracle_system *s, int64_t *row_ptr, int32_t *col, double *val,
                     double *rhs);
/* Full system (after diagonalisation + healing) in CSR and its rhs. */
void oracle_full(const oracle_system *s, int64_t *row_ptr, int32_t *col, double *val,
                 double *rhs);
/* Per-energy block ranges from fill_block_indices: 2*N_H values (first, second). */
void oracle_block_ranges(const oracle_system *s, int64_t *ranges);

/* Scale back (x * scaling * col_factor), recover odd unknowns, return the
 * full solution [n_even+n_odd] and the full-system relative residual. */
double oracle_finish(const oracle_system *s, const double *x_reduced, double *x_full);
/* Carrier density per cell [m^-3] and current per interior facet from a full
 * solution (carrier_density.hpp:68-131, current_density.hpp:48-131). */
void oracle_postproc(const oracle_system *s, const double *x_full, double *density,
                     double *facet_current);

/* TEST ONLY: scale factors of the fill's diffusion coefficient (integral_vZ *
 * area, assemble_streaming.hpp:127), drift coefficient (assemble_streaming.hpp:241-251),
 * acoustic rate and optical rate (scattering/config.hpp:50-116); 1 by default.
 * The physics pins use them to switch the drift term off and to plant a factor
 * that their checks must catch.  Applies to every later oracle_build. */
void oracle_set_plant(double diffusion, double drift, double acoustic, double optical);

/* Modena density of states [1/(J m^3)] and velocity [m/s] at kinetic energy eps [J]. */
void oracle_dispersion(int carrier, double eps, double *Z, double *v);

/* ---------------- solvers on CSR (FP64) ---------------- */
enum { ORACLE_PC_NONE = 0, ORACLE_PC_ILU0 = 1, ORACLE_PC_ILUT = 2, ORACLE_PC_BLOCK_ILU0 = 3, ORACLE_PC_ILUK = 5 };
/* ORACLE_PC_ILUK: ILU(k) on the level-of-fill pattern (Saad Alg. 10.5 symbolic,
 * Alg. 10.4 numeric), k = ilut_fill; ilut_drop > 0 adds ILUT's threshold rule
 * inside that pattern (dropped entries kept as zeros). */

typedef struct oracle_precond oracle_precond;
oracle_precond *oracle_precond_create(int kind, int64_t n, const int64_t *row_ptr,
                                      const int32_t *col, const double *val,
                                      double ilut_drop, int ilut_fill,
                                      int nblocks, const int64_t *ranges);
void oracle_precond_apply(const oracle_precond *p, const double *in, double *out);
/* Factor storage (ILU0 / block ILU0): CSR pattern of the factors and values,
 * L strictly lower (unit diagonal implied), U upper with diagonal. */
int64_t oracle_precond_nnz(const oracle_precond *p);
void oracle_precond_factors(const oracle_precond *p, int64_t *row_ptr, int32_t *col, double *val);
void oracle_precond_free(oracle_precond *p);

void oracle_spmv(int64_t n, const int64_t *row_ptr, const int32_t *col, const double *val,
                 const double *x, double *y);

/* Right-preconditioned BiCGStab on A x = b, x holds the initial guess.
 * Stops when ||b - A x||_2 / ||b||_2 <= tol (the recurrence residual is
 * checked each half step and confirmed with an explicit residual).
 * Returns the number of iterations; *rel_res receives the final true
 * relative residual.  max_seconds > 0 bounds the wall time (returns early). */
int oracle_bicgstab(int64_t n, const int64_t *row_ptr, const int32_t *col, const double *val,
                    const double *b, double *x, const oracle_precond *p, double tol,
                    int max_iters, double max_seconds, double *rel_res);
/* The same iteration on all OpenMP threads: SpMV, dot products and vector
 * updates in parallel, the blocks of a block ILU0 (ranges tiling the rows)
 * applied in parallel, as parallel_linear_solver does
 * (src/solvers/viennacl/parallel_linear_solver.hpp:64-114).  The CPU baseline
 * over the host's cores; oracle_threads() is the thread count it uses. */
int oracle_bicgstab_parallel(int64_t n, const int64_t *row_ptr, const int32_t *col, const double *val,
                             const double *b, double *x, const oracle_precond *p, double tol,
                             int max_iters, double max_seconds, double *rel_res);
int oracle_threads(void);
/* oracle_precond_create with the blocks of a block ILU0 (ranges tiling the
 * rows) factorised on all OpenMP threads; the factor is the same bit for bit. */
oracle_precond *oracle_precond_create_parallel(int kind, int64_t n, const int64_t *row_ptr,
                                               const int32_t *col, const double *val,
                                               double ilut_drop, int ilut_fill,
                                               int nblocks, const int64_t *ranges);
/* Blocks a block ILU0 applies in parallel (0: its ranges do not tile the rows). */
int oracle_precond_blocks(const oracle_precond *p);

#ifdef __cplusplus
}
#endif
#endif
