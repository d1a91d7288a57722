// c_guard.hpp — keeps C++ exceptions inside the C ABI. An entry point whose body allocates runs
// that body under c_guard: an allocation failure returns DNALN_E_NOMEM and any other exception
// DNALN_E_INTERNAL, so a C, ctypes or cgo caller gets a status instead of std::terminate.
#pragma once

#include <new>

#include "dnaln_hip.h"

namespace dnaln_amd {

template <class F>
int c_guard(F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return DNALN_E_NOMEM;
  } catch (...) {
    return DNALN_E_INTERNAL;
  }
}

}  // namespace dnaln_amd
