// Error type and status codes shared by the host library and the C ABI.
#pragma once

#include <stdexcept>
#include <string>

namespace nodals {

// Status codes returned by every C-ABI entry point (include/nodals_hip.h).
enum Status : int {
  kOk = 0,
  kErrInvalid = 1,       // bad argument, malformed input, unsupported option
  kErrHip = 2,           // HIP runtime error
  kErrNonFinite = 3,     // NaN/Inf in a norm or scalar
  kErrBreakdown = 4,     // Krylov breakdown (non-positive pAp, rho, ...)
  kErrNotConverged = 5,  // iteration cap reached
  kErrInternal = 6,
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(const std::string& m, int code = kErrInvalid) { throw Error(code, m); }

inline void require(bool ok, const std::string& m, int code = kErrInvalid) {
  if (!ok) throw Error(code, m);
}

}  // namespace nodals
