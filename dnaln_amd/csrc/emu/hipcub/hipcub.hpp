// hipcub.hpp (CPU emulation) — the few hipCUB device-wide algorithms the library calls, run on
// the host (see hip/hip_runtime.h beside it). Test infrastructure for the emulated build only.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

namespace hipcub {

namespace detail {
template <class F>
inline hipError_t run(hipStream_t st, F&& f) {
  hip_emu_no_capture(st, "a hipCUB call");
  f();
  return hipSuccess;
}
}  // namespace detail

struct DeviceScan {
  template <class I, class O, class N>
  static hipError_t ExclusiveSum(void* tmp, size_t& bytes, I in, O out, N n, hipStream_t st) {
    if (!tmp) {
      bytes = 16;
      return hipSuccess;
    }
    return detail::run(st, [&] {
      typename std::remove_reference<decltype(*out)>::type acc = 0;
      for (N i = 0; i < n; ++i) {
        const auto v = in[i];
        out[i] = acc;
        acc += v;
      }
    });
  }
};

struct DeviceRadixSort {
  template <class K, class V>
  static hipError_t SortPairs(void* tmp, size_t& bytes, const K* kin, K* kout, const V* vin, V* vout, uint32_t n,
                              int b0 = 0, int b1 = 8 * sizeof(K), hipStream_t st = nullptr) {
    return sort(tmp, bytes, kin, kout, vin, vout, n, b0, b1, st, false);
  }
  template <class K, class V>
  static hipError_t SortPairsDescending(void* tmp, size_t& bytes, const K* kin, K* kout, const V* vin, V* vout,
                                        uint32_t n, int b0 = 0, int b1 = 8 * sizeof(K), hipStream_t st = nullptr) {
    return sort(tmp, bytes, kin, kout, vin, vout, n, b0, b1, st, true);
  }
  template <class K>
  static hipError_t SortKeys(void* tmp, size_t& bytes, const K* kin, K* kout, int n, int b0 = 0,
                             int b1 = 8 * sizeof(K), hipStream_t st = nullptr) {
    if (!tmp) {
      bytes = 16;
      return hipSuccess;
    }
    return detail::run(st, [&] {
      std::vector<K> k(kin, kin + n);
      const K mask = b1 >= (int)(8 * sizeof(K)) ? ~K(0) : ((K(1) << b1) - 1);
      std::stable_sort(k.begin(), k.end(), [&](K a, K b) { return ((a & mask) >> b0) < ((b & mask) >> b0); });
      std::copy(k.begin(), k.end(), kout);
    });
  }

 private:
  template <class K, class V>
  static hipError_t sort(void* tmp, size_t& bytes, const K* kin, K* kout, const V* vin, V* vout, uint32_t n, int b0,
                         int b1, hipStream_t st, bool desc) {
    if (!tmp) {
      bytes = 16;
      return hipSuccess;
    }
    return detail::run(st, [&] {
      std::vector<uint32_t> idx(n);
      std::iota(idx.begin(), idx.end(), 0u);
      const K mask = b1 >= (int)(8 * sizeof(K)) ? ~K(0) : ((K(1) << b1) - 1);
      auto key = [&](uint32_t i) { return (kin[i] & mask) >> b0; };
      std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return desc ? key(a) > key(b) : key(a) < key(b); });
      std::vector<K> ko(n);
      std::vector<V> vo(n);
      for (uint32_t i = 0; i < n; ++i) {
        ko[i] = kin[idx[i]];
        vo[i] = vin[idx[i]];
      }
      std::copy(ko.begin(), ko.end(), kout);
      std::copy(vo.begin(), vo.end(), vout);
    });
  }
};

}  // namespace hipcub
