// Exceptions stay inside the C ABI (csrc/host/c_guard.hpp): an allocation failure becomes
// DNALN_E_NOMEM, any other exception DNALN_E_INTERNAL, and a body's own status passes through.
// The same holds for an entry point of the HIP library: dnaln_hip_index_build sizes its host
// contig arrays before any HIP call, so under an address-space limit that allocation fails on
// any machine, GPU or not. Prints "ok" when every case holds.
#include <sys/resource.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "dnaln_hip.h"
#include "host/c_guard.hpp"

// Virtual size of this process, bytes (first field of /proc/self/statm, in pages).
static unsigned long long vm_bytes() {
  unsigned long long pages = 0;
  FILE* f = std::fopen("/proc/self/statm", "r");
  if (f) {
    if (std::fscanf(f, "%llu", &pages) != 1) pages = 0;
    std::fclose(f);
  }
  return pages * 4096ull;
}

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
  {
    rlimit old{};
    getrlimit(RLIMIT_AS, &old);
    rlimit lim = old;
    lim.rlim_cur = vm_bytes() + (256ull << 20);
    setrlimit(RLIMIT_AS, &lim);
    dnaln_contig one{0, 0, 0};
    dnaln_hip_index* ix = nullptr;
    // 2^28 contigs: 3 GiB of host arrays, far beyond the limit; the contigs array is never read.
    const int rc = dnaln_hip_index_build(nullptr, 0, &one, 1 << 28, 15, 10, 0, &ix);
    setrlimit(RLIMIT_AS, &old);
    expect("dnaln_hip_index_build under an address-space limit", rc, DNALN_E_NOMEM);
    if (rc == DNALN_E_NOMEM && std::strcmp(dnaln_hip_last_error(), "host allocation failed") != 0) {
      std::printf("message: %s\n", dnaln_hip_last_error());
      ++bad;
    }
  }
  if (bad) return 1;
  std::printf("ok\n");
  return 0;
}
