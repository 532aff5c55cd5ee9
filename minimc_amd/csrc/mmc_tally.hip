// Deterministic FP64 tally reduction: see mmc_tally.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>

#include "mmc_tally.hpp"

namespace mmc {

namespace {

constexpr int kReduceBlock = 256;

// s_{h,b}: the sum of the run of records that share one key (bin, history),
// in the run's value-bit order, stored at the run's first record; 0 at every
// other record, so that sums of run_s and of run_s^2 over a bin are the bin's
// S and S2 contributions.
__global__ void KRunSums(const uint64_t* __restrict__ key, const uint64_t* __restrict__ vbits,
                         uint32_t n, double* __restrict__ run_s) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += gridDim.x * blockDim.x) {
    const uint64_t kk = key[k];
    double s = 0;
    if (k == 0 || key[k - 1] != kk) {
      for (uint32_t j = k; j < n && key[j] == kk; j++) s += __longlong_as_double(vbits[j]);
    }
    run_s[k] = s;
  }
}

// First record whose key is >= v (keys ascending).
__device__ __forceinline__ uint32_t LowerBound(const uint64_t* key, uint32_t n, uint64_t v) {
  uint32_t lo = 0, len = n;
  while (len > 0) {
    const uint32_t half = len >> 1;
    if (key[lo + half] < v) {
      lo += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo;
}

// Per bin b: the records of b form one segment of the sorted keys.  Thread t
// adds records lo + t, lo + t + 256, ... in order, then a fixed binary tree
// adds the 256 partial sums.
__global__ void __launch_bounds__(kReduceBlock)
    KBinReduce(const uint64_t* __restrict__ key, const double* __restrict__ run_s, uint32_t n,
               int nb, double* __restrict__ sum, double* __restrict__ sq) {
  __shared__ double s1[kReduceBlock], s2[kReduceBlock];
  __shared__ uint32_t seg[2];
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    if (threadIdx.x < 2)
      seg[threadIdx.x] = LowerBound(key, n, uint64_t(b + threadIdx.x) << 32);
    __syncthreads();
    const uint32_t lo = seg[0], hi = seg[1];
    double a = 0, a2 = 0;
    for (uint32_t k = lo + threadIdx.x; k < hi; k += kReduceBlock) {
      const double v = run_s[k];
      a += v;
      a2 += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = a2;
    __syncthreads();
    for (int off = kReduceBlock / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) {
        s1[threadIdx.x] += s1[threadIdx.x + off];
        s2[threadIdx.x] += s2[threadIdx.x + off];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      sum[b] = s1[0];
      if (sq) sq[b] = s2[0];
    }
    __syncthreads();  // seg and s1/s2 are reused by the next bin
  }
}

template <class T>
T* Alloc(size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
  return static_cast<T*>(p);
}

}  // namespace

void TallyWorkFree(TallyWork& w) {
  void* ps[] = {w.key_b, w.val_a, w.val_b, w.key_c, w.run_s, w.temp};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  w = TallyWork{};
}

int TallyWorkEnsure(TallyWork& w, uint32_t cap) {
  if (w.cap >= cap) return 0;
  TallyWorkFree(w);
  size_t b1 = 0, b2 = 0;
  if (rocprim::radix_sort_pairs(nullptr, b1, static_cast<const uint64_t*>(nullptr),
                                static_cast<uint64_t*>(nullptr),
                                static_cast<const uint64_t*>(nullptr),
                                static_cast<uint64_t*>(nullptr), cap) != hipSuccess ||
      rocprim::radix_sort_pairs(nullptr, b2, static_cast<const uint64_t*>(nullptr),
                                static_cast<uint64_t*>(nullptr),
                                static_cast<const uint64_t*>(nullptr),
                                static_cast<uint64_t*>(nullptr), cap, 0u, 64u) != hipSuccess)
    return 1;
  w.temp_bytes = std::max(b1, b2);
  w.key_b = Alloc<uint64_t>(cap);
  w.val_a = Alloc<uint64_t>(cap);
  w.val_b = Alloc<uint64_t>(cap);
  w.key_c = Alloc<uint64_t>(cap);
  w.run_s = Alloc<double>(cap);
  w.temp = Alloc<unsigned char>(w.temp_bytes);
  if (!w.key_b || !w.val_a || !w.val_b || !w.key_c || !w.run_s || !w.temp) {
    TallyWorkFree(w);
    return 1;
  }
  w.cap = cap;
  return 0;
}

int TallyReduce(TallyWork& w, uint64_t* key, uint64_t* val, uint32_t n, int nb, double* sum,
                double* sq, hipStream_t s) {
  if (nb <= 0) return 0;
  if (n > w.cap) return 1;
  if (n > 0) {
    size_t bytes = w.temp_bytes;
    // 1: by value bits (any fixed order of the values of one run will do)
    if (rocprim::radix_sort_pairs(w.temp, bytes, val, w.val_a, key, w.key_b, n, 0u, 64u, s) !=
        hipSuccess)
      return 1;
    // 2: by (bin, history), stable: each run keeps the value order of pass 1
    unsigned end_bit = 33;
    while (end_bit < 64 && (uint64_t(1) << (end_bit - 32)) < uint64_t(nb)) end_bit++;
    bytes = w.temp_bytes;
    if (rocprim::radix_sort_pairs(w.temp, bytes, w.key_b, w.key_c, w.val_a, w.val_b, n, 0u,
                                  end_bit, s) != hipSuccess)
      return 1;
    const uint32_t g = std::min<uint32_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(KRunSums, dim3(g), dim3(256), 0, s, w.key_c, w.val_b, n, w.run_s);
  }
  hipLaunchKernelGGL(KBinReduce, dim3(std::min(nb, 4096)), dim3(kReduceBlock), 0, s, w.key_c,
                     w.run_s, n, nb, sum, sq);
  return 0;
}

}  // namespace mmc
