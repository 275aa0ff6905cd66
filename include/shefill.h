/*
 * shefill.h -- C ABI of the MI355X SHE coefficient fill (libshesolve.so).
 *
 * Builds, for a structured 1D line or 2D quadrilateral device, the system that
 * the reference hands to its linear solver:
 *
 *   she::assemble(...)  viennashe/she/assemble_all.hpp:46-424 (streaming,
 *                       scattering, contact boundary terms)
 *   she::solve(...)     viennashe/she/linear_solver.hpp:96-189 up to the
 *                       solvers::solve call: odd elimination (she/eliminate.hpp:80-293),
 *                       unknown scaling (she/rescaling.hpp:47-107), row
 *                       normalisation (math/linalg_util.hpp:200-282) and the
 *                       per-energy block ranges (she/linear_solver.hpp:42-83)
 *
 * with the energy grid, expansion orders, boundary values and unknown mapping
 * of she/setup_energies.hpp:61-153, she/expansion_order.hpp:60-364,
 * she/boundary_conditions.hpp:57-136 and mapping.hpp:97-329.
 *
 * The host computes the integer structure (energies, orders, mapping,
 * sparsity); the device computes every value in FP64 and keeps the reduced
 * CSR system resident, ready for shesolve_set_matrix_device.  Error
 * conventions as in shesolve.h (0 ok, >0 bad argument position, <0 internal).
 */
#ifndef SHEFILL_H
#define SHEFILL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shefill_ctx *shefill_handle;

typedef struct {
  int dim;                 /* 1 or 2 */
  int nx, ny;              /* cells per direction (ny = 1 in 1D); cell id = j*nx + i */
  const double *xv;        /* nx+1 vertex x coordinates [m] */
  const double *yv;        /* ny+1 vertex y coordinates [m] (2D; ignored in 1D) */
  const double *doping_n;  /* donor doping per cell [m^-3] */
  const double *doping_p;  /* acceptor doping per cell [m^-3] */
  const double *potential; /* electrostatic potential per cell [V] */
  const int32_t *contact;  /* 1 where the cell carries a contact potential */
  double temperature;      /* lattice temperature [K] */
} shefill_device;

typedef struct {
  int carrier;             /* 0 electrons, 1 holes */
  int L_max;               /* expansion order: 1, 3, 5 or 7 */
  double energy_spacing;   /* [J] */
  double min_kin_range;    /* [J] */
  double max_kin_range;    /* [J] */
  double tau;              /* generation/recombination contact parameter [s] */
  int acoustic;            /* acoustic phonon scattering on/off */
  int optical;             /* optical phonon scattering on/off */
  int scale;               /* kinetic-energy unknown scaling on/off */
  int ordering;            /* numbering of the even unknowns within each energy level: 0 = cells in mesh
                              order (mapping.hpp:249-271); 1 = 1D devices only, cells from both ends
                              toward the middle (0, N-1, 1, N-2, ...): the same system permuted, with
                              about half the ILU level depth */
} shefill_config;

const char *shefill_last_error(void);
int shefill_create(shefill_handle *h, int device);
int shefill_destroy(shefill_handle h);

/* Builds structure and values.  Re-running with a new potential on the same
 * mesh reuses buffers; the structure is rebuilt from the new energies. */
int shefill_build(shefill_handle h, const shefill_device *dev, const shefill_config *cfg);

/* Recomputes every value of the last build on the device (same structure and
 * inputs): the per-solve coefficient fill of a Gummel step. */
int shefill_refill(shefill_handle h);

/* sizes: [0] N_H, [1] n (reduced unknowns), [2] nnz, [3] n_odd (eliminated),
 *        [4] nodes (cell, H) carrying unknowns, [5] cells, [6] interior facets */
int shefill_sizes(shefill_handle h, int64_t *out7);

/* Device pointers of the resident reduced system (owned by the handle, valid
 * until the next build or destroy): row_ptr (int64, n+1), col (int32), val
 * (FP64), rhs (FP64). */
int shefill_device_system(shefill_handle h, const int64_t **d_row_ptr, const int32_t **d_col,
                          const double **d_val, const double **d_rhs);
/* Host copies of the reduced system, the unknown scaling exp(-E_kin/kT) and
 * the column factors of the normalisation. */
int shefill_get_system(shefill_handle h, int64_t *row_ptr, int32_t *col, double *val, double *rhs,
                       double *unknown_scaling, double *col_factors);
/* Block ranges (2*N_H values), as she::fill_block_indices writes them. */
int shefill_block_ranges(shefill_handle h, int64_t *ranges);
/* Even mapping: first unknown index and expansion order per (cell, H),
 * cell-major [cells * N_H]; -1 / 0 where a node carries no unknown. */
int shefill_mapping(shefill_handle h, int64_t *even_idx, int32_t *even_order);

/* 64-bit hash of the reduced pattern (n, row pointer, columns) of the last
 * build: equal hashes across builds mean the solver can keep its analysis and
 * take new values with shesolve_update_values_device. */
int shefill_pattern_hash(shefill_handle h, uint64_t *out);
/* Device milliseconds of the last value fill (scalars + rows + normalisation). */
int shefill_last_fill_ms(shefill_handle h, double *ms);

/* Full solution from a reduced one: the even unknowns scaled back,
 * x_full[i] = x[i] * scaling[i] * col_factor[i] (she/linear_solver.hpp:172-181),
 * then the odd unknowns recovered from their rows of the full system,
 * f_o = (b_o - S^oe f_e) / S^oo_oo (she/eliminate.hpp:305-354).  x has n
 * entries, x_full n + n_odd (even unknowns first, then the odd ones in the odd
 * mapping's order).  _device: device buffers; otherwise host buffers. */
int shefill_recover(shefill_handle h, const double *x, double *x_full);
int shefill_recover_device(shefill_handle h, const double *d_x, double *d_x_full);
/* Carrier density per cell [m^-3] (she/postproc/carrier_density.hpp:68-131) and
 * current per interior facet (she/postproc/current_density.hpp:48-131, facet
 * order of the build: per cell the right then the top neighbour) from a full
 * solution.  density or current may be NULL.  _device: device buffers. */
int shefill_postproc(shefill_handle h, const double *x_full, double *density, double *facet_current);
int shefill_postproc_device(shefill_handle h, const double *d_x_full, double *d_density,
                            double *d_facet_current);
/* Initial guess for the current build from the solution of the previous
 * build on the same mesh (the step before a potential change), on the device:
 * she::transfer_to_new_h_space (she/transfer_to_new_h_space.hpp:58-196) for
 * the even unknowns -- per cell and level the old level of nearest kinetic
 * energy, values rescaled per cell so that the cell keeps the previous
 * density -- then converted to the reduced, scaled unknowns of the current
 * system.  d_x_full_prev: full solution of the previous build (as
 * shefill_recover_device wrote it), d_density_prev: its density per cell (as
 * shefill_postproc_device wrote it); d_x: n values.  Odd unknowns are
 * eliminated, so the reference's facet part has no counterpart here.  Returns
 * -1 when the previous build was on another mesh or there was none. */
int shefill_transfer_device(shefill_handle h, const double *d_x_full_prev, const double *d_density_prev,
                            double *d_x);
/* Gummel update of the carrier density per cell (simulator.hpp:524-600,
 * update_quantity with damping alpha in (0, 1]): density <- density (1 - alpha)
 * + new alpha, or with logarithmic != 0, density^(1-alpha) *
 * max(new, 1e-10 density)^alpha.  The reference's force_no_damping is a plain
 * copy of new.  *update_norm (may be NULL) receives the relative update norm
 * of simulator.hpp:286-320 between new and the updated density,
 * sqrt(sum ((new - upd) / upd)^2), with new in place of the quotient where
 * upd = 0: the value the reference's Gummel loop reports at its next
 * iteration, where it evaluates the norm before writing the next solution
 * back (simulator.hpp:1081-1086).  _device: device buffers, n_cells values
 * each. */
/* Terminal current (postproc/terminal_current.hpp:50-107, get_terminal_current):
 * the facet currents of shefill_postproc summed over the interior facets with
 * exactly one cell in the terminal (terminal_mask: one byte per cell, nonzero
 * = the cell belongs to the terminal), times the facet area, positive where the
 * terminal cell is the facet's first cell (left / bottom).  _device: device
 * buffers. */
int shefill_terminal_current(shefill_handle h, const double *facet_current, const unsigned char *terminal_mask,
                             double *current);
int shefill_terminal_current_device(shefill_handle h, const double *d_facet_current,
                                    const unsigned char *d_terminal_mask, double *current);
/* The device update above without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and error codes as shefill_density_update_device (the
 * density bit for bit, the update norm bit for bit); its partial sums live in
 * a fixed device array of the library, one per GPU, held until the process
 * ends.  At most SHEFILL_NOALLOC_MAX_CELLS cells (returns 1 beyond).  All calls
 * on one GPU share that array: calls that overlap in time on one GPU (two
 * host threads, two streams) race and must be serialised by the caller. */
#define SHEFILL_NOALLOC_MAX_CELLS (65536 * 256)
int shefill_density_update_device_noalloc(shefill_handle h, const double *d_new, double *d_density, double alpha,
                                          int logarithmic, double *update_norm);
int shefill_density_update(shefill_handle h, const double *new_density, double *density, double alpha,
                           int logarithmic, double *update_norm);
int shefill_density_update_device(shefill_handle h, const double *d_new, double *d_density, double alpha,
                                  int logarithmic, double *update_norm);

#ifdef __cplusplus
}
#endif
#endif
