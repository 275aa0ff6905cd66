// Test declarations under ViennaSHE's own path and names: linear_solver_ids
// and linear_solver_config of viennashe/solvers/config.hpp:39-195 (ids,
// defaults: tolerance 1e-13, 1000 iterations, ILUT 60 / 1e-4, scaling on,
// block boundaries; the GPU id accepted only with VIENNASHE_HAVE_GPU_SOLVER),
// plus -- unless SHESOLVE_DECLS_UNPATCHED is defined -- the GPU members that
// the configuration patch of the integration adds.  Used only to compile the
// GPU solver overload standalone.
#pragma once

#include <algorithm>
#include <cstddef>
#include <string>
#include <utility>
#include <vector>

#include "viennashe/solvers/exception.hpp"

namespace viennashe {
namespace solvers {

struct linear_solver_ids {
  enum {
    dense_linear_solver,
    serial_linear_solver,
    parallel_linear_solver,
    gpu_parallel_linear_solver,
    petsc_parallel_linear_solver,
    petsc_parallel_AMGX_solver
  };
};

class linear_solver_config : public linear_solver_ids {
 public:
  typedef std::vector<std::pair<std::size_t, std::size_t> > block_preconditioner_boundaries_container;
#ifndef SHESOLVE_DECLS_UNPATCHED
  // the GPU members (values: shesolve.h's SHESOLVE_FP* and SHESOLVE_PRECOND_*)
  enum gpu_precision_type { gpu_fp64 = 0, gpu_fp32_refined = 1 };
  enum gpu_preconditioner_type {
    gpu_ilu0 = 1,
    gpu_block_ilu0 = 3,
    gpu_mc_ilu0 = 4,
    gpu_iluk = 5,
    gpu_bmc_ilu0 = 6,
    gpu_block_iluk = 7
  };
#endif

  linear_solver_config()
      : id_(serial_linear_solver), tol_(1e-13), max_iters_(1000), ilut_entries_(60), ilut_drop_tol_(1e-4),
        do_scale_(true) {}

  void scale(bool value) { do_scale_ = value; }
  bool scale() const { return do_scale_; }
  long id() const { return id_; }

  void set(long solver_id) {
    switch (solver_id) {
      case dense_linear_solver:
      case serial_linear_solver:
      case parallel_linear_solver:
      case petsc_parallel_linear_solver:
      case petsc_parallel_AMGX_solver:
#ifdef VIENNASHE_HAVE_GPU_SOLVER
      case gpu_parallel_linear_solver:
#endif
        id_ = solver_id;
        break;
      default:
        throw invalid_linear_solver_exception();
    }
  }
  void set(std::string solver_name) {
    std::transform(solver_name.begin(), solver_name.end(), solver_name.begin(), ::tolower);
    if (solver_name == "dense_linear_solver") id_ = dense_linear_solver;
    else if (solver_name == "serial_linear_solver") id_ = serial_linear_solver;
    else if (solver_name == "parallel_linear_solver") id_ = parallel_linear_solver;
    else if (solver_name == "petsc_parallel_linear_solver") id_ = petsc_parallel_linear_solver;
#ifdef VIENNASHE_HAVE_GPU_SOLVER
    else if (solver_name == "gpu_parallel_linear_solver") id_ = gpu_parallel_linear_solver;
#endif
    else throw invalid_linear_solver_exception();
  }

  double tolerance() const { return tol_; }
  void tolerance(double tol) { if (tol > 0) tol_ = tol; }
  std::size_t max_iters() const { return max_iters_; }
  void max_iters(std::size_t iters) { max_iters_ = iters; }
  std::size_t ilut_entries() const { return ilut_entries_; }
  void ilut_entries(std::size_t num) { ilut_entries_ = num; }
  double ilut_drop_tolerance() const { return ilut_drop_tol_; }
  void ilut_drop_tolerance(double tol) { if (tol > 0) ilut_drop_tol_ = tol; }
  block_preconditioner_boundaries_container &block_preconditioner_boundaries() { return block_precond_boundaries_; }
  block_preconditioner_boundaries_container const &block_preconditioner_boundaries() const {
    return block_precond_boundaries_;
  }

#ifndef SHESOLVE_DECLS_UNPATCHED
  int gpu_device() const { return gpu_device_; }
  void gpu_device(int d) { gpu_device_ = d; }
  gpu_precision_type gpu_precision() const { return gpu_precision_; }
  void gpu_precision(gpu_precision_type p) { gpu_precision_ = p; }
  gpu_preconditioner_type gpu_preconditioner() const { return gpu_precond_; }
  void gpu_preconditioner(gpu_preconditioner_type p) { gpu_precond_ = p; }
  int gpu_fill_level() const { return gpu_fill_level_; }
  void gpu_fill_level(int k) { if (k >= 0) gpu_fill_level_ = k; }
  double gpu_drop_tolerance() const { return gpu_drop_tol_; }
  void gpu_drop_tolerance(double tol) { if (tol >= 0 && tol < 1) gpu_drop_tol_ = tol; }
#endif

 private:
  long id_;
  double tol_;
  std::size_t max_iters_;
  std::size_t ilut_entries_;
  double ilut_drop_tol_;
  bool do_scale_;
  block_preconditioner_boundaries_container block_precond_boundaries_;
#ifndef SHESOLVE_DECLS_UNPATCHED
  int gpu_device_ = 0;
  gpu_precision_type gpu_precision_ = gpu_fp32_refined;
  gpu_preconditioner_type gpu_precond_ = gpu_ilu0;
  int gpu_fill_level_ = 1;
  double gpu_drop_tol_ = 0.0;
#endif
};

}  // namespace solvers
}  // namespace viennashe
