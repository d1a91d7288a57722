// Exceptions stay inside the C ABI (csrc/host/c_guard.hpp): an allocation failure becomes
// DNALN_E_NOMEM, any other exception DNALN_E_INTERNAL, and a body's own status passes through.
// Prints "ok" when every case holds.
#include <cstdio>
#include <new>
#include <stdexcept>
#include <vector>

#include "host/c_guard.hpp"

int main() {
  using dnaln_amd::c_guard;
  int bad = 0;
  auto expect = [&bad](const char* what, int got, int want) {
    if (got != want) {
      std::printf("%s: %d, expected %d\n", what, got, want);
      ++bad;
    }
  };
  expect("status", c_guard([] { return DNALN_E_CAPACITY; }), DNALN_E_CAPACITY);
  expect("ok", c_guard([] { return DNALN_OK; }), DNALN_OK);
  expect("bad_alloc", c_guard([]() -> int { throw std::bad_alloc(); }), DNALN_E_NOMEM);
  expect("length_error", c_guard([]() -> int {
           std::vector<int> v;
           v.reserve(v.max_size() + 1);
           return DNALN_OK;
         }),
         DNALN_E_INTERNAL);
  expect("runtime_error", c_guard([]() -> int { throw std::runtime_error("x"); }), DNALN_E_INTERNAL);
  expect("int", c_guard([]() -> int { throw 7; }), DNALN_E_INTERNAL);
  if (bad) return 1;
  std::printf("ok\n");
  return 0;
}
