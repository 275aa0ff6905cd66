// viennashe/solvers/gpu_linear_solver.hpp -- the GPU case of ViennaSHE's linear solver dispatch.
//
// Defines viennashe::solvers::solve(A, b, config, gpu_parallel_linear_solver_tag),
// the overload that solve_impl (src/solvers/solve.hpp:78-80) calls for the id
// gpu_parallel_linear_solver once VIENNASHE_HAVE_GPU_SOLVER is defined.  It is
// written against ViennaSHE's own declarations and includes them by their
// paths in that tree: the tag (viennashe/forwards.h:216), the matrix
// (viennashe/math/linalg_util.hpp:67), the configuration
// (viennashe/solvers/config.hpp:56) and the exception types
// (viennashe/util/exception.hpp:38,52).  solve_impl's own checks (empty rows,
// NaN/Inf in b, b = 0; src/solvers/solve.hpp:53-66) run before it is reached.
//
// The rows are copied to CSR as solvers::detail::copy does
// (src/solvers/viennacl/serial_linear_solver.hpp:51-89: rows in order,
// columns ascending, every stored entry including explicit zeros) and handed
// to libshesolve's C ABI (include/shesolve.h).  A failure of the library is
// thrown as viennashe::util::linear_solver_exception, the type solve_impl
// throws for a failed system; a non-finite matrix value as
// viennashe::util::invalid_value_exception, the type of
// util::check_vector_for_valid_entries (viennashe/util/checks.hpp:110-130).
// Non-convergence is not an error at this layer (the reference logs it); the
// achieved residual is returned through `report` when given.
//
// The GPU choices come from the configuration when it provides them
// (gpu_precision(), gpu_preconditioner(), gpu_device(), gpu_fill_level(),
// gpu_drop_tolerance(); the preconditioner values are the C ABI's
// SHESOLVE_PRECOND_*).  A configuration without them gets FP32 iterations
// with FP64 refinement, per-energy block ILU0 when block boundaries are set
// (the blocks the reference's parallel solver uses,
// src/solvers/viennacl/parallel_linear_solver.hpp:79-93) and ILU0 otherwise,
// on device 0.
#ifndef VIENNASHE_SOLVERS_GPU_LINEAR_SOLVER_HPP
#define VIENNASHE_SOLVERS_GPU_LINEAR_SOLVER_HPP

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <sstream>
#include <type_traits>
#include <utility>
#include <vector>

#include "shesolve.h"
#include "viennashe/forwards.h"
#include "viennashe/math/linalg_util.hpp"
#include "viennashe/solvers/config.hpp"
#include "viennashe/util/exception.hpp"

namespace viennashe {
namespace solvers {
namespace gpu_detail {

struct csr_matrix {
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col;
  std::vector<double> val;
};

/** std::map rows -> CSR (the order of solvers::detail::copy). */
template <typename MatrixT>
csr_matrix copy_to_csr(MatrixT const &A) {
  csr_matrix m;
  const std::size_t n = A.size1();
  m.row_ptr.resize(n + 1);
  m.row_ptr[0] = 0;
  m.col.reserve(A.nnz());
  m.val.reserve(A.nnz());
  for (std::size_t i = 0; i < n; ++i) {
    typename MatrixT::row_type const &r = A.row(i);
    for (typename MatrixT::const_iterator2 it = r.begin(); it != r.end(); ++it) {
      if (!std::isfinite(static_cast<double>(it->second))) {
        std::ostringstream ss;
        ss << "solve(): invalid matrix entry " << it->second << " at (" << i << ", " << it->first << ")";
        throw viennashe::util::invalid_value_exception(ss.str());
      }
      m.col.push_back(static_cast<int32_t>(it->first));
      m.val.push_back(static_cast<double>(it->second));
    }
    m.row_ptr[i + 1] = static_cast<int64_t>(m.col.size());
  }
  return m;
}

inline void check(int rc, const char *what) {
  if (rc != 0) {
    std::ostringstream ss;
    ss << "solve(): " << what << " failed (" << rc << "): " << shesolve_last_error();
    throw viennashe::util::linear_solver_exception(ss.str());
  }
}

struct handle_deleter {
  void operator()(shesolve_ctx *h) const { shesolve_destroy(h); }
};

// Whether the configuration carries the GPU choices (C++11 detection).
template <typename...>
struct voider {
  typedef void type;
};
template <typename C, typename = void>
struct has_gpu_knobs {
  static const bool value = false;
};
template <typename C>
struct has_gpu_knobs<C, typename voider<decltype(std::declval<C const &>().gpu_precision()),
                                        decltype(std::declval<C const &>().gpu_preconditioner()),
                                        decltype(std::declval<C const &>().gpu_device()),
                                        decltype(std::declval<C const &>().gpu_fill_level()),
                                        decltype(std::declval<C const &>().gpu_drop_tolerance())>::type> {
  static const bool value = true;
};

struct choice {
  int device = 0;
  int precision = SHESOLVE_FP32_IR;
  int precond = -1;  // -1: block ILU0 when blocks are set, else ILU0
  int fill_level = 1;
  double drop_tol = 0.0;
};

template <typename Config>
typename std::enable_if<has_gpu_knobs<Config>::value, choice>::type choose(Config const &c) {
  choice ch;
  ch.device = static_cast<int>(c.gpu_device());
  ch.precision = static_cast<int>(c.gpu_precision()) == SHESOLVE_FP64 ? SHESOLVE_FP64 : SHESOLVE_FP32_IR;
  ch.precond = static_cast<int>(c.gpu_preconditioner());
  ch.fill_level = static_cast<int>(c.gpu_fill_level());
  ch.drop_tol = static_cast<double>(c.gpu_drop_tolerance());
  return ch;
}
template <typename Config>
typename std::enable_if<!has_gpu_knobs<Config>::value, choice>::type choose(Config const &) {
  return choice();
}

}  // namespace gpu_detail

/** Solves A x = b on the GPU (the id gpu_parallel_linear_solver of solve_impl). */
template <typename NumericT, typename VectorType>
VectorType solve(viennashe::math::sparse_matrix<NumericT> const &A, VectorType const &b,
                 linear_solver_config const &config, gpu_parallel_linear_solver_tag,
                 shesolve_report *report = 0) {
  gpu_detail::csr_matrix m = gpu_detail::copy_to_csr(A);
  const gpu_detail::choice ch = gpu_detail::choose(config);
  shesolve_params p;
  shesolve_default_params(&p);
  p.tol = config.tolerance();
  p.max_iters = static_cast<int>(config.max_iters());
  p.precision = ch.precision;
  std::vector<int64_t> ranges;
  linear_solver_config::block_preconditioner_boundaries_container const &blk = config.block_preconditioner_boundaries();
  for (std::size_t k = 0; k < blk.size(); ++k) {
    ranges.push_back(static_cast<int64_t>(blk[k].first));
    ranges.push_back(static_cast<int64_t>(blk[k].second));
  }
  p.precond = ch.precond >= 0 ? ch.precond : (ranges.empty() ? SHESOLVE_PRECOND_ILU0 : SHESOLVE_PRECOND_BLOCK_ILU0);
  const bool block_kind = p.precond == SHESOLVE_PRECOND_BLOCK_ILU0 || p.precond == SHESOLVE_PRECOND_BLOCK_ILUK;
  if (block_kind && ranges.empty())  // no blocks given: the whole matrix is one block
    p.precond = p.precond == SHESOLVE_PRECOND_BLOCK_ILUK ? SHESOLVE_PRECOND_ILUK : SHESOLVE_PRECOND_ILU0;
  if (!block_kind) ranges.clear();
  p.fill_level = ch.fill_level;
  p.drop_tol = ch.drop_tol;
  shesolve_ctx *raw = 0;
  gpu_detail::check(shesolve_create(&raw, ch.device, &p), "shesolve_create");
  std::unique_ptr<shesolve_ctx, gpu_detail::handle_deleter> h(raw);
  gpu_detail::check(shesolve_set_matrix(h.get(), static_cast<int64_t>(b.size()), m.row_ptr.back(), m.row_ptr.data(),
                                        m.col.data(), m.val.data(), static_cast<int>(ranges.size() / 2),
                                        ranges.empty() ? 0 : ranges.data()),
                    "shesolve_set_matrix");
  std::vector<double> bb(b.begin(), b.end()), x(b.size(), 0.0);
  shesolve_report rep;
  gpu_detail::check(shesolve_solve(h.get(), bb.data(), x.data(), 0, &rep), "shesolve_solve");
  if (report) *report = rep;
  VectorType result(b.size());
  for (std::size_t i = 0; i < x.size(); ++i) result[i] = x[i];
  return result;
}

}  // namespace solvers
}  // namespace viennashe

#endif
