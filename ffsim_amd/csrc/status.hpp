// Status handling shared by the library's source files: a failure carries the
// C ABI status code and a message; guarded() turns it into the return code and
// the thread's last-error string.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/ffsim_amd.h"

namespace ffsim_amd {

struct Failure {
  int code;
  std::string msg;
};

#define HIP_CHECK(expr)                                                               \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      throw Failure{FFSIM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
  } while (0)

inline void require(bool ok, int code, const std::string& msg) {
  if (!ok) throw Failure{code, msg};
}

// The string ffsim_amd_last_error() returns on this thread.
void set_last_error(const std::string& msg);

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return FFSIM_OK;
  } catch (const Failure& e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return FFSIM_ERR_VALUE;
  }
}

}  // namespace ffsim_amd
