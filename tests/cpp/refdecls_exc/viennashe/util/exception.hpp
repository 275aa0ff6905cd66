// Test declarations under ViennaSHE's own path and names: the two exception
// types of viennashe/util/exception.hpp:38-56 that the solver layer throws.
// When the reference tree is at hand the test compiles against its header
// instead.
#pragma once

#include <stdexcept>
#include <string>

namespace viennashe {
namespace util {

class linear_solver_exception : public std::runtime_error {
 public:
  linear_solver_exception(std::string const &str) : std::runtime_error(str) {}
};

class invalid_value_exception : public std::runtime_error {
 public:
  invalid_value_exception(std::string const &str) : std::runtime_error(str) {}
};

}  // namespace util
}  // namespace viennashe
