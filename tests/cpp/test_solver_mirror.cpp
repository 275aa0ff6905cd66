// Exercises the C++ solver front end (viennashe-dev_amd/host) the way
// she::solve calls solvers::solve (she/linear_solver.hpp:165-167).
//   test_solver_mirror cpu             -- config / copy / argument checks, no GPU use
//   test_solver_mirror gpu in.bin out  -- solve a CSR system read from in.bin
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "viennashe/solvers/gpu_linear_solver.hpp"

using viennashe::math::sparse_matrix;
using namespace viennashe::solvers;

#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

static int cpu_checks() {
  linear_solver_config conf;
  CHECK(conf.id() == linear_solver_config::serial_linear_solver);
  CHECK(conf.tolerance() == 1e-13 && conf.max_iters() == 1000 && conf.scale());
  conf.set(linear_solver_ids::gpu_parallel_linear_solver);
  CHECK(conf.id() == 3);
  conf.set(std::string("Serial_Linear_Solver"));
  CHECK(conf.id() == linear_solver_config::serial_linear_solver);
  conf.set(std::string("GPU_PARALLEL_LINEAR_SOLVER"));
  CHECK(conf.id() == linear_solver_config::gpu_parallel_linear_solver);
  bool threw = false;
  try { conf.set(42L); } catch (invalid_linear_solver_exception const &) { threw = true; }
  CHECK(threw);
  threw = false;
  try { conf.set(std::string("cuda_solver")); } catch (invalid_linear_solver_exception const &) { threw = true; }
  CHECK(threw);
  conf.tolerance(-1.0);
  CHECK(conf.tolerance() == 1e-13);
  // std::map rows -> CSR: ascending columns, explicit zeros kept
  sparse_matrix<double> A(3, 3);
  A(0, 2) = 2.0; A(0, 0) = 1.0; A(1, 1) = 0.0; A(2, 0) = -1.0; A(2, 2) = 4.0;
  detail::csr_matrix m = detail::copy_to_csr(A);
  CHECK(m.row_ptr.size() == 4 && m.row_ptr[3] == 5);
  const int32_t cols[5] = {0, 2, 1, 0, 2};
  const double vals[5] = {1.0, 2.0, 0.0, -1.0, 4.0};
  for (int k = 0; k < 5; ++k) CHECK(m.col[k] == cols[k] && m.val[k] == vals[k]);
  // zero right-hand side: x = 0 without a solve (src/solvers/solve.hpp:61-66)
  std::vector<double> b(3, 0.0);
  std::vector<double> x = solve(A, b, conf);
  CHECK(x.size() == 3 && x[0] == 0.0 && x[1] == 0.0 && x[2] == 0.0);
  // NaN in the right-hand side throws (util/checks.hpp:110-130)
  b[1] = std::nan("");
  threw = false;
  try { solve(A, b, conf); } catch (invalid_value_exception const &) { threw = true; }
  CHECK(threw);
  // a CPU id is not served by this library
  b[1] = 1.0;
  conf.set(linear_solver_ids::serial_linear_solver);
  threw = false;
  try { solve(A, b, conf); } catch (invalid_linear_solver_exception const &) { threw = true; }
  CHECK(threw);
  // an empty row is rejected before any solve (src/solvers/solve.hpp:53-58)
  sparse_matrix<double> E(2, 2);
  E(0, 0) = 1.0;
  std::vector<double> b2(2, 1.0);
  threw = false;
  try { solve(E, b2, conf); } catch (linear_solver_exception const &) { threw = true; }
  CHECK(threw);
  // the ILU(k) knobs
  CHECK(conf.gpu_fill_level() == 1 && conf.gpu_drop_tolerance() == 0.0);
  conf.gpu_fill_level(2);
  conf.gpu_drop_tolerance(1e-3);
  conf.gpu_drop_tolerance(2.0);
  CHECK(conf.gpu_fill_level() == 2 && conf.gpu_drop_tolerance() == 1e-3);
  // the preconditioner knob carries the C ABI's values
  CHECK(linear_solver_config::gpu_ilu0 == SHESOLVE_PRECOND_ILU0);
  CHECK(linear_solver_config::gpu_block_ilu0 == SHESOLVE_PRECOND_BLOCK_ILU0);
  CHECK(linear_solver_config::gpu_mc_ilu0 == SHESOLVE_PRECOND_MC_ILU0);
  CHECK(linear_solver_config::gpu_iluk == SHESOLVE_PRECOND_ILUK);
  CHECK(linear_solver_config::gpu_bmc_ilu0 == SHESOLVE_PRECOND_BMC_ILU0);
  CHECK(conf.gpu_preconditioner() == linear_solver_config::gpu_ilu0);
  conf.gpu_preconditioner(linear_solver_config::gpu_bmc_ilu0);
  CHECK(conf.gpu_preconditioner() == linear_solver_config::gpu_bmc_ilu0);
  std::printf("cpu checks passed\n");
  return 0;
}

static int gpu_solve(const char *in, const char *out, const char *precond) {
  FILE *f = std::fopen(in, "rb");
  if (!f) return 2;
  int64_t n = 0, nnz = 0;
  if (std::fread(&n, 8, 1, f) != 1 || std::fread(&nnz, 8, 1, f) != 1) return 2;
  std::vector<int64_t> rp(n + 1);
  std::vector<int32_t> ci(nnz);
  std::vector<double> v(nnz), b(n);
  if (std::fread(rp.data(), 8, n + 1, f) != (size_t)(n + 1) || std::fread(ci.data(), 4, nnz, f) != (size_t)nnz ||
      std::fread(v.data(), 8, nnz, f) != (size_t)nnz || std::fread(b.data(), 8, n, f) != (size_t)n)
    return 2;
  std::fclose(f);
  sparse_matrix<double> A(n, n);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) A(i, ci[k]) = v[k];
  linear_solver_config conf;
  conf.set(std::string("gpu_parallel_linear_solver"));
  conf.tolerance(1e-12);
  conf.max_iters(5000);
  if (precond && std::strcmp(precond, "bmc_ilu0") == 0) conf.gpu_preconditioner(linear_solver_config::gpu_bmc_ilu0);
  std::vector<double> x = solve(A, b, conf);
  FILE *g = std::fopen(out, "wb");
  std::fwrite(x.data(), 8, n, g);
  std::fclose(g);
  std::printf("gpu solve done n=%lld\n", (long long)n);
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && std::strcmp(argv[1], "cpu") == 0) return cpu_checks();
  if (argc >= 4 && std::strcmp(argv[1], "gpu") == 0) return gpu_solve(argv[2], argv[3], argc >= 5 ? argv[4] : nullptr);
  std::fprintf(stderr, "usage: %s cpu | gpu in.bin out.bin\n", argv[0]);
  return 2;
}
