// viennashe/solvers/gpu_linear_solver.hpp -- solvers::solve for gpu_parallel_linear_solver.
//
// The tag overload that solve_impl (src/solvers/solve.hpp:68-86) dispatches to
// for the dormant GPU id.  It copies the std::map rows to CSR exactly like
// solvers::detail::copy (src/solvers/viennacl/serial_linear_solver.hpp:51-89:
// rows in order, columns ascending, every stored entry -- explicit zeros too)
// and calls the C ABI of libshesolve (include/shesolve.h).
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <sstream>
#include <vector>

#include "shesolve.h"
#include "viennashe/math/sparse_matrix.hpp"
#include "viennashe/solvers/config.hpp"

namespace viennashe {
namespace solvers {

struct gpu_parallel_linear_solver_tag {};

namespace detail {

struct csr_matrix {
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col;
  std::vector<double> val;
};

/** std::map rows -> CSR (the order of solvers::detail::copy). */
template <typename NumericT>
csr_matrix copy_to_csr(viennashe::math::sparse_matrix<NumericT> const &A) {
  csr_matrix m;
  const std::size_t n = A.size1();
  m.row_ptr.resize(n + 1);
  m.row_ptr[0] = 0;
  std::size_t nnz = A.nnz();
  m.col.reserve(nnz);
  m.val.reserve(nnz);
  for (std::size_t i = 0; i < n; ++i) {
    typename viennashe::math::sparse_matrix<NumericT>::row_type const &r = A.row(i);
    for (typename viennashe::math::sparse_matrix<NumericT>::const_iterator2 it = r.begin(); it != r.end(); ++it) {
      m.col.push_back(static_cast<int32_t>(it->first));
      m.val.push_back(static_cast<double>(it->second));
    }
    m.row_ptr[i + 1] = static_cast<int64_t>(m.col.size());
  }
  return m;
}

inline void check_shesolve(int rc, const char *what) {
  if (rc != 0) {
    std::ostringstream ss;
    ss << "solve(): " << what << " failed (" << rc << "): " << shesolve_last_error();
    throw linear_solver_exception(ss.str());
  }
}

struct handle_deleter {
  void operator()(shesolve_ctx *h) const { shesolve_destroy(h); }
};

}  // namespace detail

/** Solves A x = b on the GPU.  Non-convergence is not an error at this layer
 * (the reference logs it); the achieved residual is returned through
 * `report` when given. */
template <typename NumericT>
std::vector<double> solve(viennashe::math::sparse_matrix<NumericT> const &A, std::vector<double> const &b,
                          linear_solver_config const &config, gpu_parallel_linear_solver_tag,
                          shesolve_report *report = 0) {
  detail::csr_matrix m = detail::copy_to_csr(A);
  shesolve_params p;
  shesolve_default_params(&p);
  p.tol = config.tolerance();
  p.max_iters = static_cast<int>(config.max_iters());
  p.precision = config.gpu_precision() == linear_solver_config::gpu_fp64 ? SHESOLVE_FP64 : SHESOLVE_FP32_IR;
  std::vector<int64_t> ranges;
  linear_solver_config::block_preconditioner_boundaries_container const &blk = config.block_preconditioner_boundaries();
  if (config.gpu_preconditioner() == linear_solver_config::gpu_block_ilu0 && !blk.empty()) {
    p.precond = SHESOLVE_PRECOND_BLOCK_ILU0;
    for (std::size_t k = 0; k < blk.size(); ++k) {
      ranges.push_back(static_cast<int64_t>(blk[k].first));
      ranges.push_back(static_cast<int64_t>(blk[k].second));
    }
  } else if (config.gpu_preconditioner() == linear_solver_config::gpu_iluk) {
    p.precond = SHESOLVE_PRECOND_ILUK;
    p.fill_level = config.gpu_fill_level();
    p.drop_tol = config.gpu_drop_tolerance();
  } else if (config.gpu_preconditioner() == linear_solver_config::gpu_mc_ilu0) {
    p.precond = SHESOLVE_PRECOND_MC_ILU0;
  } else if (config.gpu_preconditioner() == linear_solver_config::gpu_bmc_ilu0) {
    p.precond = SHESOLVE_PRECOND_BMC_ILU0;
  } else {
    p.precond = SHESOLVE_PRECOND_ILU0;
  }
  shesolve_ctx *raw = 0;
  detail::check_shesolve(shesolve_create(&raw, config.gpu_device(), &p), "shesolve_create");
  std::unique_ptr<shesolve_ctx, detail::handle_deleter> h(raw);
  detail::check_shesolve(shesolve_set_matrix(h.get(), static_cast<int64_t>(b.size()), m.row_ptr.back(),
                                             m.row_ptr.data(), m.col.data(), m.val.data(),
                                             static_cast<int>(ranges.size() / 2), ranges.empty() ? 0 : ranges.data()),
                         "shesolve_set_matrix");
  std::vector<double> x(b.size(), 0.0);
  shesolve_report rep;
  detail::check_shesolve(shesolve_solve(h.get(), b.data(), x.data(), 0, &rep), "shesolve_solve");
  if (report) *report = rep;
  return x;
}

/** solve_impl restricted to the id this library provides: the same checks as
 * src/solvers/solve.hpp:47-87 (NaN/Inf in b throws, zero b gives x = 0), then
 * the GPU solve.  The CPU ids belong to the reference's own libshesolvers. */
template <typename NumericT>
std::vector<double> solve(viennashe::math::sparse_matrix<NumericT> &A, std::vector<double> const &b,
                          linear_solver_config const &config) {
  if (A.size1() != b.size()) throw linear_solver_exception("solve(): size mismatch of matrix and right-hand side");
  // matrix_consistency_check: an empty row is an error (src/solvers/solve.hpp:53-58)
  for (std::size_t i = 0; i < A.size1(); ++i)
    if (A.row(i).empty()) throw linear_solver_exception("solve(): Found empty row in system_matrix.");
  double nrm = 0.0;
  for (std::size_t i = 0; i < b.size(); ++i) {
    if (!std::isfinite(b[i])) {
      std::ostringstream ss;
      ss << "solve(): invalid entry " << b[i] << " at index " << i;
      throw invalid_value_exception(ss.str());
    }
    nrm += b[i] * b[i];
  }
  if (nrm == 0.0) return std::vector<double>(b.size(), 0.0);
  if (config.id() != linear_solver_config::gpu_parallel_linear_solver)
    throw invalid_linear_solver_exception("solve(): this library provides gpu_parallel_linear_solver only");
  return solve(static_cast<viennashe::math::sparse_matrix<NumericT> const &>(A), b, config,
               gpu_parallel_linear_solver_tag());
}

}  // namespace solvers
}  // namespace viennashe
