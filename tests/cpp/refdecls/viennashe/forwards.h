// Test declarations under ViennaSHE's own path and names: the solver tags of
// viennashe/forwards.h:200-217 (the reference header also pulls in ViennaGrid,
// absent here).  Used only to compile the GPU solver overload standalone.
#pragma once

namespace viennashe {
namespace solvers {

class dense_linear_solver_tag {};
class serial_linear_solver_tag {};
class parallel_linear_solver_tag {};
class petsc_linear_solver_tag {};
class petsc_amgx_solver_tag {};
class gpu_parallel_linear_solver_tag {};

}  // namespace solvers
}  // namespace viennashe
