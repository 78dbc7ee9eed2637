// Device-only helpers shared by the kernel files.
#pragma once

#include "mpic_internal.hpp"

namespace mpic {

// 12 / 16 consecutive doubles of element e (records are 96 / 128 B, 16-B aligned).
__device__ __forceinline__ void load12(const double* __restrict__ base, int e, double o[12]) {
  const double2* q = reinterpret_cast<const double2*>(base + (size_t)e * 12);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double2 t = q[i];
    o[2 * i] = t.x;
    o[2 * i + 1] = t.y;
  }
}
__device__ __forceinline__ void load16(const double* __restrict__ base, int e, double o[16]) {
  const double2* q = reinterpret_cast<const double2*>(base + (size_t)e * 16);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const double2 t = q[i];
    o[2 * i] = t.x;
    o[2 * i + 1] = t.y;
  }
}

// physToReferenceVector, mesh.hpp:245-250: out_r = ((0 + J_r0 v0) + J_r1 v1) + J_r2 v2.
__device__ __forceinline__ void p2r_vec(const double* Ji, const double v[3], double o[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double acc = 0.0;
    acc += Ji[r * 3 + 0] * v[0];
    acc += Ji[r * 3 + 1] * v[1];
    acc += Ji[r * 3 + 2] * v[2];
    o[r] = acc;
  }
}

// 128-bit atomic add: 64-bit add on the low word, carry folded into the high word.
// Each low-word add sees an exact old value, so the carries sum exactly and the
// final value is exact in any order.
__device__ __forceinline__ void atomic_add_fx(fx128* dst, fx128 v) {
  if (v.lo == 0ull && v.hi == 0ll) return;
  const unsigned long long old =
      __hip_atomic_fetch_add(&dst->lo, v.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long carry = (old + v.lo) < old ? 1ull : 0ull;
  const unsigned long long h = (unsigned long long)v.hi + carry;
  if (h != 0ull)
    __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(&dst->hi), h, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Wavefront-segmented charge deposit.  Lanes holding the same element in a
// contiguous run (particles sorted by element) are summed with shuffles; the
// first lane of each run adds the run's 4 node charges to the element's
// accumulator slots acc[key*4 + i].  key < 0 lanes contribute nothing.
__device__ __forceinline__ void deposit_segmented(int key, fx128 c[4], fx128* __restrict__ acc) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1);
  const bool head = (lane == 0) || (prev != key);
  const unsigned long long heads = __ballot(head);
  if (heads != ~0ull) {
    const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int run = __popcll(heads & le);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int orun = __shfl_down(run, off);
      const bool take = (lane + off < 64) && (orun == run);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const unsigned long long olo = __shfl_down(c[i].lo, off);
        const long long ohi = __shfl_down(c[i].hi, off);
        if (take) c[i] = fx_add(c[i], fx128{olo, ohi});
      }
    }
  }
  if (head && key >= 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) atomic_add_fx(acc + (size_t)key * 4 + i, c[i]);
  }
}

}  // namespace mpic
