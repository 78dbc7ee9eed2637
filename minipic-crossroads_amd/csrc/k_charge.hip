// Node charge assembly (the finalize half of PIC::weight_charge, PIC.cpp:139-160)
// and small utility kernels.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

// rhs[i] = background[i] + sum of the element slots touching node i, exact in
// fixed point; boundary rows are zeroed (PIC.cpp:155-158).  Each slot (element
// e, local node k: 4 digit words at acc[(e*4+k)*4]) belongs to exactly one node,
// so the slot is cleared here for the next deposit.
__global__ __launch_bounds__(BLOCK) void k_node_charge(DevMesh m, unsigned long long* __restrict__ acc, int S,
                                                       double* __restrict__ rhs) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= m.nn) return;
  // the node's element slots (slot e*4+node: 4 digit words), NODE_BATCH at a
  // time with all their loads in flight (indices clamped to the last slot,
  // whose extra reads are not added); each slot is cleared for the next deposit
  fx128 s = m.bgfix ? m.bgfix[i] : fx128{0, 0};
  const int b = m.adj_ptr[i], e = m.adj_ptr[i + 1];
  for (int k = b; k < e; k += NODE_BATCH) {
    int a[NODE_BATCH];
#pragma unroll
    for (int u = 0; u < NODE_BATCH; ++u) a[u] = m.adj[min(k + u, e - 1)];
    ulonglong2 w01[NODE_BATCH], w23[NODE_BATCH];
#pragma unroll
    for (int u = 0; u < NODE_BATCH; ++u) {
      const unsigned long long* w = acc + (size_t)a[u] * 4;
      w01[u] = *reinterpret_cast<const ulonglong2*>(w);
      w23[u] = *reinterpret_cast<const ulonglong2*>(w + 2);
    }
#pragma unroll
    for (int u = 0; u < NODE_BATCH; ++u)
      if (k + u < e) {
        const unsigned long long ws[4] = {w01[u].x, w01[u].y, w23[u].x, w23[u].y};
        s = fx_add(s, dep_slot_value(ws));
        unsigned long long* w = acc + (size_t)a[u] * 4;
        *reinterpret_cast<ulonglong2*>(w) = ulonglong2{0, 0};
        *reinterpret_cast<ulonglong2*>(w + 2) = ulonglong2{0, 0};
      }
  }
  rhs[i] = m.bnd[i] ? 0.0 : from_fixed(s, S);
}

__global__ void k_zero(unsigned long long* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

hipError_t launch_node_charge(const DevMesh& m, unsigned long long* acc, int S, double* rhs, hipStream_t s) {
  hipLaunchKernelGGL(k_node_charge, dim3((m.nn + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, m, acc, S, rhs);
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
