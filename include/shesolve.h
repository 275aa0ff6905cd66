/*
 * shesolve.h -- C ABI of the MI355X SHE linear solver (libshesolve.so).
 *
 * Drop-in for the solver call of the ViennaSHE SHE linear-solve path:
 *
 *   viennashe::solvers::solve(sparse_matrix<double>& A, std::vector<double> const& b,
 *                             linear_solver_config const& config)
 *     viennashe/solvers/forwards.h:38-41, defined src/solver.cpp:38-42,
 *     dispatched by solve_impl in src/solvers/solve.hpp:47-87
 *
 * The dormant solver id gpu_parallel_linear_solver (viennashe/solvers/config.hpp:46,
 * libviennashe/include/config.h:34-35) selects this library; the C++ shim in
 * viennashe-dev_amd/host/ converts the std::map rows to CSR exactly like
 * solvers::detail::copy (src/solvers/viennacl/serial_linear_solver.hpp:51-89)
 * and calls the functions below.
 *
 * Conventions (mirroring libviennashe, libviennashe/src/viennashe_all.hpp:72-77):
 *   every function returns int: 0 on success, >0 = 1-based position of an
 *   invalid argument, <0 = internal / HIP error.  Non-convergence is not an
 *   error: the solve returns 0 with report.status != 0 (as the reference's
 *   solvers::solve returns its last iterate without throwing).
 *   shesolve_last_error() returns a message for the last failure of the calling thread.
 *   All calls are synchronous at the ABI (results are complete on return).
 *   A handle is not thread-safe; use one handle per host thread.
 *   Host buffers stay owned by the caller; device buffers passed to *_device
 *   functions must be device allocations on the handle's device.
 */
#ifndef SHESOLVE_H
#define SHESOLVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shesolve_ctx *shesolve_handle;

/* Preconditioners.  ILU0 and BLOCK_ILU0 are the reference's
 * ilu0_precond / block_ilu_precond<ilu0> (serial_linear_solver.hpp:125-126,
 * parallel_linear_solver.hpp:79-93); block ranges come from
 * she::fill_block_indices (she/linear_solver.hpp:42-83). */
enum {
  SHESOLVE_PRECOND_NONE = 0,
  SHESOLVE_PRECOND_ILU0 = 1,        /* ILU(0), natural row order, level-scheduled */
  SHESOLVE_PRECOND_BLOCK_ILU0 = 3,  /* ILU(0) of the diagonal blocks (per-energy slabs) */
  SHESOLVE_PRECOND_MC_ILU0 = 4,     /* ILU(0) of the matrix in a greedy multicolour row
                                       ordering: a few dependency levels instead of
                                       hundreds; not in the reference, same stopping rule */
  SHESOLVE_PRECOND_ILUK = 5,        /* ILU(k) on the level-of-fill pattern, k = fill_level,
                                       natural order, level-scheduled; drop_tol > 0 adds
                                       ILUT's threshold rule inside that pattern.  The
                                       GPU form of the ILUT(drop, fill) slot of the
                                       reference config (solvers/config.hpp:166-186) */
  SHESOLVE_PRECOND_BLOCK_ILUK = 7,  /* ILU(k) of the diagonal blocks (per-energy slabs), k =
                                       fill_level (per system in batches); k = 0 is BLOCK_ILU0.
                                       The reference's block_ilu_precond with ILU(k) in place of
                                       its ilu0_tag (parallel_linear_solver.hpp:79-93); the
                                       ranges must be contiguous row ranges */
  SHESOLVE_PRECOND_BMC_ILU0 = 6     /* ILU(0) of the matrix in a block multicolour ordering:
                                       blocks of 32 consecutive rows coloured greedily, a
                                       block's rows kept together in natural order; one
                                       sweep launch per colour, each block solved by half
                                       a wavefront.  Not in the reference, same stopping rule */
};

enum {
  SHESOLVE_FP64 = 0,     /* FP64 BiCGStab */
  SHESOLVE_FP32_IR = 1   /* FP32 BiCGStab inside FP64 iterative refinement */
};

typedef struct {
  double tol;            /* target ||b - A x||_2 / ||b||_2 (FP64), default 1e-10 */
  int max_iters;         /* BiCGStab iteration cap per (inner) solve, default 1000 */
  int precond;           /* SHESOLVE_PRECOND_* */
  int precision;         /* SHESOLVE_FP64 or SHESOLVE_FP32_IR */
  double inner_tol;      /* relative reduction of each FP32 inner solve, default 1e-2 */
  int max_ir_steps;      /* cap on refinement steps, default 30 */
  int lanes_per_row;     /* SpMV lanes per row (0 = pick from nnz/row) */
  int fill_level;        /* ILUK: level of fill k, default 1 */
  double drop_tol;       /* ILUK: relative drop tolerance (0 = plain ILU(k)), default 0 */
  int autotune;          /* 1: time the CSR-stream launch shapes at set-up and keep the fastest
                            (default; a handle re-set with a matrix of the same preconditioner
                            within 10 % of the row count it tuned for keeps those shapes);
                            0: fixed shape (2048-entry row blocks, strided), so that results do
                            not depend on timings */
} shesolve_params;

typedef struct {
  int status;            /* 0 converged, 1 iteration cap reached, 2 breakdown */
  int iterations;        /* total BiCGStab iterations (all refinement steps) */
  int ir_steps;          /* refinement steps (FP32_IR), 0 for FP64 */
  double rel_residual;   /* final FP64 ||b - A x|| / ||b|| */
  double setup_ms;       /* last shesolve_set_matrix* (transfer + analysis + factorisation) */
  double solve_ms;       /* last solve */
} shesolve_report;

/* Fills *p with the defaults listed above. */
void shesolve_default_params(shesolve_params *p);
const char *shesolve_last_error(void);

/* Creates a solver bound to HIP device `device`. */
int shesolve_create(shesolve_handle *h, int device, const shesolve_params *p);
int shesolve_destroy(shesolve_handle h);
int shesolve_set_params(shesolve_handle h, const shesolve_params *p);

/* Uploads an n x n CSR matrix (ascending column indices per row, host
 * buffers), analyses its pattern and factorises the preconditioner.
 * block_ranges: nblocks (first, last) row pairs for BLOCK_ILU0, may be NULL
 * otherwise.  Rows not covered by any range form one extra block. */
int shesolve_set_matrix(shesolve_handle h, int64_t n, int64_t nnz, const int64_t *row_ptr,
                        const int32_t *col, const double *val, int nblocks,
                        const int64_t *block_ranges);
/* Same with the CSR already in device memory (row_ptr int64, col int32, val FP64). */
int shesolve_set_matrix_device(shesolve_handle h, int64_t n, int64_t nnz, const int64_t *d_row_ptr,
                               const int32_t *d_col, const double *d_val, int nblocks,
                               const int64_t *block_ranges);
/* New values on the stored pattern (host / device), refactorises. */
int shesolve_update_values(shesolve_handle h, const double *val);
int shesolve_update_values_device(shesolve_handle h, const double *d_val);

/* Solves A x = b.  x holds the initial guess when warm_start != 0, otherwise
 * it is overwritten from zero.  A zero right-hand side gives x = 0 without
 * iterating (src/solvers/solve.hpp:61-66).  report may be NULL. */
int shesolve_solve(shesolve_handle h, const double *b, double *x, int warm_start, shesolve_report *report);
int shesolve_solve_device(shesolve_handle h, const double *d_b, double *d_x, int warm_start,
                          shesolve_report *report);

/* y = A x in FP64 on the device (host buffers); used to check the SpMV. */
int shesolve_spmv(shesolve_handle h, const double *x, double *y);
/* y = A x in FP32 (the cast copy of A), host buffers in float. */
int shesolve_spmv_f32(shesolve_handle h, const float *x, float *y);
/* z = M^{-1} r with the FP64 preconditioner (host buffers). */
int shesolve_precond_apply(shesolve_handle h, const double *r, double *z);
/* Copies the FP64 factor values (pattern = the stored matrix, or its block-
 * diagonal part for BLOCK_ILU0) and the factor pattern to host buffers sized
 * by shesolve_factor_nnz. */
int64_t shesolve_factor_nnz(shesolve_handle h);
int shesolve_get_factors(shesolve_handle h, int64_t *row_ptr, int32_t *col, double *val);
/* Launch shape the setup autotuner picked: entries per thread of the
 * CSR-stream kernels (row block = 256 * items entries; the sweeps always use
 * them), or -LPR when the vector-CSR SpMV with LPR lanes per row won the SpMV;
 * XCD-contiguous block mapping (0/1); and the nine timings it compared (FP32
 * SpMV, microseconds; items 4/8/16 x mapping strided/contiguous, then vector
 * CSR with 4/8/16 lanes per row; with a multicolour factor the first six are
 * those of the permuted matrix's SpMV with its fused dot product and 4-byte
 * column indices).  Any pointer may be NULL. */
int shesolve_stream_config(shesolve_handle h, int32_t *items, int32_t *xcd, double *tune_us9);
/* Row ordering of the factor (new -> old); the identity unless MC_ILU0. */
int shesolve_get_ordering(shesolve_handle h, int32_t *perm);
/* Level-schedule statistics: [0] forward levels, [1] backward levels of the
 * factor's dependency graph (the factorisation's launches).  The sweeps of
 * MC_ILU0 take one launch per colour on these levels; those of BMC_ILU0 take
 * one launch per colour of blocks, fewer than these levels. */
int shesolve_levels(shesolve_handle h, int64_t *out2);

/* Refinement history of the last solve on this handle.  The reference solves
 * in FP64 only (serial_linear_solver.hpp:129-136) and has no counterpart: this
 * reports the FP32 + FP64-refinement path against an FP64 solution.
 * One record per FP64 residual evaluation of an FP32_IR solve: the initial
 * guess (step 0), then the iterate after each refinement step, then, when
 * the refinement handed over to FP64 BiCGStab, the final iterate (kind 2).
 * An FP64 solve leaves one record of kind 2. */
typedef struct {
  int32_t step;              /* refinement steps applied before this record */
  int32_t inner_iterations;  /* FP32 BiCGStab iterations of that step (FP64 ones for kind 2) */
  int32_t kind;              /* 0 initial guess or accepted FP32 correction, 1 correction that did
                                not reduce the FP64 residual (undone), 2 after the FP64 solve */
  int32_t reserved;
  double rel_residual;       /* FP64 ||b - A x||_2 / ||b||_2 of the iterate */
  double rel_error;          /* ||x - x_ref||_2 / ||x_ref||_2, NaN without a reference solution */
} shesolve_ir_step;

/* Reference solution x_ref (n values, host or device memory as on_device says)
 * against which the history records the solution error; NULL clears it.  It
 * is kept until cleared, a new matrix is set or the handle is destroyed; the
 * error costs one extra vector pass per record, ordinary solves none. */
int shesolve_set_reference_solution(shesolve_handle h, int64_t n, const double *x_ref, int on_device);
/* Copies up to cap records of the last solve to out and their total number to *count. */
int shesolve_ir_history(shesolve_handle h, int32_t cap, shesolve_ir_step *out, int32_t *count);

/* Timing of one kernel class over the solves since profiling was switched on:
 * average device time per launch in microseconds and the launch count, for
 * the FP32 SpMV (kind 0), forward sweep level (1), backward sweep level (2),
 * vector kernels (3), FP64 SpMV (4: FP64 solves and refinement residuals).
 * shesolve_set_profiling(h, mode) before the solves: mode 1 brackets every
 * launch of the classes with events and waits for each (the solve slows
 * down); mode 2 times only the SpMV launches (kinds 0 and 4), with events
 * recorded on the solver stream without waiting, resolved when
 * shesolve_kernel_stats is called; 0 switches profiling off.  Switching
 * resets the statistics. */
int shesolve_set_profiling(shesolve_handle h, int on);
int shesolve_kernel_stats(shesolve_handle h, int kind, double *avg_us, int64_t *launches);
/* Times `reps` back-to-back SpMV launches (FP64 if fp32 == 0, else FP32) with
 * HIP events on the solver stream; returns the average per launch in
 * microseconds. */
int shesolve_time_spmv(shesolve_handle h, int fp32, int reps, double *avg_us);
/* Algorithmic bytes of one SpMV in the format the solver launches (FP32 values
 * and vectors if fp32 != 0, else FP64): the values, the 4-byte column indices
 * or, in the block multicolour solve space, the 16-bit column codes and their
 * per-row-block tables, the row pointer, x once and y.  Measurement helper for
 * the roofline of the SpMV (no counterpart in the reference). */
int shesolve_spmv_bytes(shesolve_handle h, int fp32, int64_t *bytes);

/* ---------------- batched solves ----------------
 * Many independent systems of uneven size (bias points, devices, expansion
 * orders) solved together.  The reference solves them one solvers::solve call
 * at a time (src/solvers/solve.hpp:47-87); here they are stored block-
 * diagonally on one device and every BiCGStab phase is one launch over all
 * systems, with per-system scalars, stopping tests and convergence masks kept
 * on the device; one iteration is captured into a HIP graph and replayed.
 * Each system keeps its own result: the same preconditioner (factored per
 * system: the factor of a block-diagonal matrix is block-diagonal), the same
 * stopping rule on its own ||b_s - A_s x_s|| / ||b_s|| as a single solve.
 * The parameters (precond, precision, tolerances) apply to every system.
 * ILU sweeps run one workgroup per system over its own level schedule. */
typedef struct shesolve_batch_ctx *shesolve_batch;

int shesolve_batch_create(shesolve_batch *h, int device, const shesolve_params *p);
int shesolve_batch_destroy(shesolve_batch h);
int shesolve_batch_set_params(shesolve_batch h, const shesolve_params *p);
/* count systems; system s has n[s] rows, CSR row_ptr[s] (int64, n[s]+1),
 * col[s] (int32, ascending per row), val[s] (FP64); block ranges per system
 * for BLOCK_ILU0 (nblocks and block_ranges may be NULL).  Host buffers, or
 * device buffers for the _device variant.  Total rows and entries must fit
 * 32-bit indices. */
int shesolve_batch_set_systems(shesolve_batch h, int count, const int64_t *n, const int64_t *const *row_ptr,
                               const int32_t *const *col, const double *const *val, const int *nblocks,
                               const int64_t *const *block_ranges);
int shesolve_batch_set_systems_device(shesolve_batch h, int count, const int64_t *n,
                                      const int64_t *const *d_row_ptr, const int32_t *const *d_col,
                                      const double *const *d_val, const int *nblocks,
                                      const int64_t *const *block_ranges);
/* New values for every system on the stored patterns; refactorises. */
int shesolve_batch_update_values(shesolve_batch h, const double *const *val);
int shesolve_batch_update_values_device(shesolve_batch h, const double *const *d_val);
/* Solves every system; reports[s] (may be NULL) receives system s's result. */
int shesolve_batch_solve(shesolve_batch h, const double *const *b, double *const *x, int warm_start,
                         shesolve_report *reports);
int shesolve_batch_solve_device(shesolve_batch h, const double *const *d_b, double *const *d_x, int warm_start,
                                shesolve_report *reports);
/* ILU(k) level per system for SHESOLVE_PRECOND_ILUK (0 = ILU(0) for that
 * system), used by the next shesolve_batch_set_systems*; count must then equal
 * the number of systems.  count = 0 goes back to params.fill_level for all. */
int shesolve_batch_set_fill_levels(shesolve_batch h, int count, const int *levels);
/* use_graph: replay one captured iteration (default 1); poll: iterations
 * between host checks of the convergence masks (default 8). */
int shesolve_batch_options(shesolve_batch h, int use_graph, int poll);
/* [0] systems, [1] total rows, [2] total entries, [3] deepest per-system level
 * schedule, [4] levels of the concatenated schedule, [5] factor entries. */
int shesolve_batch_info(shesolve_batch h, int64_t *out6);
/* The underlying single-system handle of the concatenated matrix (SpMV,
 * factors, timing helpers); owned by the batch. */
shesolve_handle shesolve_batch_solver(shesolve_batch h);

#ifdef __cplusplus
}
#endif
#endif
