// Node charge assembly (the finalize half of PIC::weight_charge, PIC.cpp:139-160)
// and small utility kernels.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

// rhs[i] = background[i] + sum of the element slots touching node i, exact in
// fixed point; boundary rows are zeroed (PIC.cpp:155-158).  Each slot belongs to
// exactly one node, so the slot is cleared here for the next deposit.
__global__ __launch_bounds__(BLOCK) void k_node_charge(DevMesh m, fx128* __restrict__ acc, int S,
                                                       double* __restrict__ rhs) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= m.nn) return;
  fx128 s = m.bgfix ? m.bgfix[i] : fx128{0, 0};
  const int b = m.adj_ptr[i], e = m.adj_ptr[i + 1];
  for (int k = b; k < e; ++k) {
    const int slot = m.adj[k];
    s = fx_add(s, acc[slot]);
    acc[slot] = fx128{0, 0};
  }
  rhs[i] = m.bnd[i] ? 0.0 : from_fixed(s, S);
}

// max |w*q| over live particles (as ordered double bits) and the live count.
__global__ __launch_bounds__(BLOCK) void k_charge_bound(DevParticles P, int64_t n, const double* sp,
                                                        unsigned long long* out2) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  unsigned long long mx = 0, live = 0;
  if (p < n && P.elem[p] >= 0) {
    mx = __double_as_longlong(fabs(P.w[p] * sp[P.type[p]]));
    live = 1;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const unsigned long long o = __shfl_xor(mx, off);
    mx = o > mx ? o : mx;
    live += __shfl_xor(live, off);
  }
  if ((threadIdx.x & 63) == 0) {
    if (mx) atomicMax(out2, mx);
    if (live) atomicAdd(out2 + 1, live);
  }
}

__global__ void k_zero(unsigned long long* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

hipError_t launch_node_charge(const DevMesh& m, fx128* acc_tet, int S, double* rhs, hipStream_t s) {
  hipLaunchKernelGGL(k_node_charge, dim3((m.nn + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, m, acc_tet, S, rhs);
  return hipGetLastError();
}

hipError_t launch_charge_bound(const DevParticles& P, int64_t n, const double* species, unsigned long long* out2,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_charge_bound, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, P, n, species,
                     out2);
  return hipGetLastError();
}

// Zero fill with a kernel (usable inside captured graphs).
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 8;
  if (n == 0) return hipSuccess;
  size_t g = (n + BLOCK - 1) / BLOCK;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)g), dim3(BLOCK), 0, s, (unsigned long long*)p, n);
  return hipGetLastError();
}

}  // namespace mpic
