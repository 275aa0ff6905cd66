// Test declarations under ViennaSHE's own path and names: the solver-selection
// exception of viennashe/solvers/exception.hpp:35-42.  When the reference tree
// is at hand the test compiles against its header instead.
#pragma once

#include <stdexcept>

namespace viennashe {
namespace solvers {

class invalid_linear_solver_exception : public std::runtime_error {
 public:
  invalid_linear_solver_exception() : std::runtime_error("* Invalid linear solver specified.") {}
};

}  // namespace solvers
}  // namespace viennashe
