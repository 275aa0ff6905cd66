// viennashe/solvers/config.hpp -- linear solver selection with the MI355X solver enabled.
//
// Mirrors viennashe/solvers/config.hpp:39-215 (ids, defaults tol 1e-13,
// 1000 iterations, ILUT 60 / 1e-4, scaling on, block boundaries) and turns the
// dormant id gpu_parallel_linear_solver (config.hpp:46, set() cases :93-95,
// :127-131) into the selector of libshesolve.  The extra knobs choose the
// precision and the preconditioner of the GPU solve.
#pragma once

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace viennashe {
namespace solvers {

/** Thrown for an unknown or unavailable solver id (viennashe/solvers/exception.hpp:35-51). */
class invalid_linear_solver_exception : public std::runtime_error {
 public:
  invalid_linear_solver_exception() : std::runtime_error("* ERROR: Invalid linear solver specified!") {}
  explicit invalid_linear_solver_exception(std::string const &s) : std::runtime_error(s) {}
};
/** Thrown when the solver itself fails (util::linear_solver_exception). */
class linear_solver_exception : public std::runtime_error {
 public:
  explicit linear_solver_exception(std::string const &s) : std::runtime_error(s) {}
};
/** Thrown for NaN/Inf entries (util::check_vector_for_valid_entries). */
class invalid_value_exception : public std::runtime_error {
 public:
  explicit invalid_value_exception(std::string const &s) : std::runtime_error(s) {}
};

struct linear_solver_ids {
  enum {
    dense_linear_solver,
    serial_linear_solver,
    parallel_linear_solver,
    gpu_parallel_linear_solver,  // libshesolve on MI355X
    petsc_parallel_linear_solver,
    petsc_parallel_AMGX_solver
  };
};

class linear_solver_config : public linear_solver_ids {
 public:
  typedef std::vector<std::pair<std::size_t, std::size_t> > block_preconditioner_boundaries_container;
  enum gpu_precision_type { gpu_fp64 = 0, gpu_fp32_refined = 1 };
  enum gpu_preconditioner_type { gpu_ilu0 = 1, gpu_block_ilu0 = 3, gpu_mc_ilu0 = 4, gpu_iluk = 5, gpu_bmc_ilu0 = 6 };

  linear_solver_config()
      : id_(serial_linear_solver), tol_(1e-13), max_iters_(1000), ilut_entries_(60), ilut_drop_tol_(1e-4),
        do_scale_(true), gpu_device_(0), gpu_precision_(gpu_fp32_refined), gpu_precond_(gpu_ilu0),
        gpu_fill_level_(1), gpu_drop_tol_(0.0) {}

  void scale(bool value) { do_scale_ = value; }
  bool scale() const { return do_scale_; }
  long id() const { return id_; }

  /** The CPU ids stay the reference's; this library provides the GPU one. */
  void set(long solver_id) {
    switch (solver_id) {
      case dense_linear_solver:
      case serial_linear_solver:
      case parallel_linear_solver:
      case petsc_parallel_linear_solver:
      case petsc_parallel_AMGX_solver:
      case gpu_parallel_linear_solver:
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
    else if (solver_name == "petsc_parallel_amgx_solver") id_ = petsc_parallel_AMGX_solver;
    else if (solver_name == "gpu_parallel_linear_solver") id_ = gpu_parallel_linear_solver;
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

  // GPU solver knobs
  int gpu_device() const { return gpu_device_; }
  void gpu_device(int d) { gpu_device_ = d; }
  gpu_precision_type gpu_precision() const { return gpu_precision_; }
  void gpu_precision(gpu_precision_type p) { gpu_precision_ = p; }
  gpu_preconditioner_type gpu_preconditioner() const { return gpu_precond_; }
  void gpu_preconditioner(gpu_preconditioner_type p) { gpu_precond_ = p; }
  /** gpu_iluk: level of fill k and relative drop tolerance (0 = plain ILU(k)) */
  int gpu_fill_level() const { return gpu_fill_level_; }
  void gpu_fill_level(int k) { if (k >= 0) gpu_fill_level_ = k; }
  double gpu_drop_tolerance() const { return gpu_drop_tol_; }
  void gpu_drop_tolerance(double tol) { if (tol >= 0 && tol < 1) gpu_drop_tol_ = tol; }

 private:
  long id_;
  double tol_;
  std::size_t max_iters_;
  std::size_t ilut_entries_;
  double ilut_drop_tol_;
  bool do_scale_;
  block_preconditioner_boundaries_container block_precond_boundaries_;
  int gpu_device_;
  gpu_precision_type gpu_precision_;
  gpu_preconditioner_type gpu_precond_;
  int gpu_fill_level_;
  double gpu_drop_tol_;
};

}  // namespace solvers
}  // namespace viennashe
