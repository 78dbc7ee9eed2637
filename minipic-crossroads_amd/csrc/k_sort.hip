// Counting sort of the particle list: the build's replacement for
// ParticleList::sort (Kokkos::BinSort, particle_list.cpp:90-116) and
// compute_start_particles_by_element (particle_list.cpp:118-146).  The key is the
// owning element, or, for the timestep's own reordering, the spatial bucket of
// the owning element (a few dozen neighbouring elements): most particles stay in
// their bucket from one step to the next, so the reordering keeps their relative
// order and the next push reads them almost in sequence.
//   1. count[key] = particles with that key (key-grouped atomics; the fused push
//      produces bucket counts itself);
//   2. start = exclusive scan of count (start[nkeys] = live particles);
//   3. rank: perm[start[key] + rank] = position of every live particle; dead
//      particles (elem < 0) get no slot, which also compacts the list
//      (DeleteParticles, particle_list.cpp:148-191).
// Only the 4-byte permutation is scattered: the next fused push gathers each
// particle through it and writes the whole state in key order (or
// launch_permute applies it).  Reading in storage order and writing each
// record to its slot instead was measured slower on MI355X (the scattered
// 64/96-B record writes cost more than the scattered reads).  As in the reference, the order inside one key is
// not specified (it follows the order of the ranking's atomics); the
// identity key (id) travels with the particle.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;

// Lanes of a wavefront with the same element are grouped (wave_group_leader);
// the leader adds the group size with one atomic, so every atomic instruction
// touches distinct addresses.
// Sort key of particle p: -1 when dead or out of range.
__device__ __forceinline__ int sort_key(const int32_t* __restrict__ elem, const int32_t* __restrict__ kmap, int64_t p,
                                        int64_t n) {
  const int e = p < n ? elem[p] : -1;
  return (e >= 0 && kmap) ? kmap[e] : e;
}

__global__ __launch_bounds__(BLOCK) void k_count(const int32_t* __restrict__ elem, const int32_t* __restrict__ kmap,
                                                 int64_t n, int32_t* count) {
  const int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int e = sort_key(elem, kmap, p, n);
  unsigned long long grp;
  const int leader = wave_group_leader(e, &grp);
  if (leader == (int)(threadIdx.x & 63)) atomicAdd(count + e, (int)__popcll(grp));
}

// Exclusive scan, pass 1: tile totals.
__global__ __launch_bounds__(BLOCK) void k_scan_tiles(const int32_t* __restrict__ in, int64_t n, int32_t* tile_sum) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) s += in[base + k];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off);
  __shared__ int ws[BLOCK / 64];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Block-wide exclusive scan of one int per thread; returns the block total in *total.
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  __shared__ int ws[BLOCK / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int o = __shfl_up(inc, off);
    if (lane >= off) inc += o;
  }
  __syncthreads();
  if (lane == 63) ws[wave] = inc;
  __syncthreads();
  int wbase = 0;
  for (int w = 0; w < wave; ++w) wbase += ws[w];
  *total = ws[0] + ws[1] + ws[2] + ws[3];
  return wbase + inc - v;
}

// Pass 2: scan of the tile totals in one workgroup (in place, exclusive).
__global__ __launch_bounds__(BLOCK) void k_scan_top(int32_t* tile_sum, int ntiles) {
  const int per = (ntiles + BLOCK - 1) / BLOCK;
  const int b = threadIdx.x * per;
  int s = 0;
  for (int k = 0; k < per; ++k)
    if (b + k < ntiles) s += tile_sum[b + k];
  int total;
  int run = block_excl_scan(s, &total);
  for (int k = 0; k < per; ++k)
    if (b + k < ntiles) {
      const int t = tile_sum[b + k];
      tile_sum[b + k] = run;
      run += t;
    }
}

// Pass 3: exclusive scan inside each tile plus the tile offset.
__global__ __launch_bounds__(BLOCK) void k_scan_apply(const int32_t* __restrict__ in, int64_t n,
                                                      const int32_t* __restrict__ tile_off, int32_t* out,
                                                      int32_t* cursor, int64_t ncursor) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int vals[SCAN_ITEMS];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    vals[k] = (base + k < n) ? in[base + k] : 0;
    s += vals[k];
  }
  int total;
  int run = block_excl_scan(s, &total) + tile_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) {
      out[base + k] = run;
      if (base + k < ncursor) cursor[base + k] = run;
    }
    run += vals[k];
  }
}

// Ranking, one workgroup per RANK_ITEMS * BLOCK consecutive particles: each
// particle takes a local rank among the workgroup's particles with its key from
// an LDS table (key, count), then one thread per key the workgroup holds adds
// the key's count to cursor[key] with one atomic, and every particle's slot is
// the returned base plus its local rank.  Between reorderings a particle moves
// only to neighbouring elements, so a workgroup's 2,048 particles hold a few
// dozen keys: a small table (512 entries, cleared and scanned in two passes of
// the workgroup) holds them, and a particle whose key finds no entry within
// RANK_PROBES slots (a list far from key order: after a fill, or sorted by a
// different key) takes its own slot with one atomic on cursor[key].  Every
// particle's elements are loaded before any key is mapped or ranked.  Per
// ranking at 100M particles, between reorderings: 1,024 particles and a
// 2,048-entry table that never overflows 0.75 ms; this 0.39 ms, and 0.84 ms for
// the first ranking after a fill, whose workgroups hold a few hundred keys (a
// 256-entry table: 0.38 and 1.9 ms).
constexpr int RANK_ITEMS = 8;
constexpr int RANK_BITS = 9;
constexpr int RANK_TABLE = 1 << RANK_BITS;
constexpr int RANK_PROBES = 32;
static_assert(RANK_TABLE % BLOCK == 0 && RANK_PROBES <= RANK_TABLE, "rank table");

__global__ __launch_bounds__(BLOCK) void k_rank(const int32_t* __restrict__ elem, const int32_t* __restrict__ kmap,
                                                int64_t n, int32_t* __restrict__ cursor, int32_t* __restrict__ perm) {
  __shared__ int tkey[RANK_TABLE];
  __shared__ int tcnt[RANK_TABLE];
  for (int k = threadIdx.x; k < RANK_TABLE; k += BLOCK) {
    tkey[k] = -1;
    tcnt[k] = 0;
  }
  const int64_t p0 = (int64_t)blockIdx.x * (RANK_ITEMS * BLOCK) + threadIdx.x;
  int key[RANK_ITEMS], slot[RANK_ITEMS], rank[RANK_ITEMS];
#pragma unroll
  for (int u = 0; u < RANK_ITEMS; ++u) {
    const int64_t p = p0 + u * BLOCK;
    key[u] = p < n ? elem[p] : -1;
  }
  if (kmap) {
#pragma unroll
    for (int u = 0; u < RANK_ITEMS; ++u) key[u] = key[u] >= 0 ? kmap[key[u]] : -1;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < RANK_ITEMS; ++u) {
    const int e = key[u];
    slot[u] = -1;  // dead: no slot
    rank[u] = 0;
    if (e < 0) continue;
    unsigned h = ((unsigned)e * 2654435761u) >> (32 - RANK_BITS);
    int probe = 0;
    for (; probe < RANK_PROBES; ++probe) {
      int k = __hip_atomic_load(&tkey[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (k == -1) {
        const int old = atomicCAS(&tkey[h], -1, e);
        k = old == -1 ? e : old;
      }
      if (k == e) break;
      h = (h + 1) & (RANK_TABLE - 1);
    }
    if (probe == RANK_PROBES) {
      slot[u] = -2;  // no entry: a slot of its own below
      continue;
    }
    slot[u] = (int)h;
    rank[u] = atomicAdd(&tcnt[h], 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < RANK_TABLE; k += BLOCK) {
    const int kk = tkey[k];
    if (kk >= 0) tcnt[k] = atomicAdd(cursor + kk, tcnt[k]);  // the key's base slot
  }
  __syncthreads();
  // every destination first (the atomics of the particles without an entry all in
  // flight together), then the stores back to back: a wavefront's memory
  // operations complete in order, so a store between them would be waited for
#pragma unroll
  for (int u = 0; u < RANK_ITEMS; ++u)
    if (slot[u] == -2) rank[u] = atomicAdd(cursor + key[u], 1);
#pragma unroll
  for (int u = 0; u < RANK_ITEMS; ++u) rank[u] = slot[u] >= 0 ? tcnt[slot[u]] + rank[u] : slot[u] == -2 ? rank[u] : -1;
#pragma unroll
  for (int u = 0; u < RANK_ITEMS; ++u)
    if (rank[u] >= 0) perm[rank[u]] = (int32_t)(p0 + u * BLOCK);
}

__global__ __launch_bounds__(BLOCK) void k_permute(DevParticles src, DevParticles dst, const int32_t* __restrict__ perm,
                                                   int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (q >= n) return;
  const Rec64 r = load_rec(src.r, perm[q]);
  store_rec(dst.r, q, r);
  dst.elem[q] = r.elem;
}

}  // namespace

hipError_t launch_rank(const DevMesh& m, const DevParticles& P, int64_t n, bool by_bucket, bool counts_ready,
                       SortBuffers& sb, int32_t* perm, hipStream_t s) {
  const int32_t* kmap = by_bucket ? m.bucket : nullptr;
  const int64_t nkeys = by_bucket ? m.nbuckets : m.ne;
  const int64_t nk1 = nkeys + 1;
  const unsigned gp = (unsigned)((n + BLOCK - 1) / BLOCK);
  if (!counts_ready) {
    hipError_t err = launch_zero(sb.count, (size_t)((nk1 + 1) / 2) * 8, s);
    if (err != hipSuccess) return err;
    if (n > 0) hipLaunchKernelGGL(k_count, dim3(gp), dim3(BLOCK), 0, s, P.elem, kmap, n, sb.count);
  }
  const int ntiles = (int)((nk1 + SCAN_TILE - 1) / SCAN_TILE);
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(BLOCK), 0, s, sb.count, nk1, sb.blocksum);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(BLOCK), 0, s, sb.blocksum, ntiles);
  hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(BLOCK), 0, s, sb.count, nk1, sb.blocksum, sb.start, sb.cursor,
                     nkeys);
  const unsigned gr = (unsigned)((n + RANK_ITEMS * BLOCK - 1) / (RANK_ITEMS * BLOCK));
  if (n > 0) hipLaunchKernelGGL(k_rank, dim3(gr), dim3(BLOCK), 0, s, P.elem, kmap, n, sb.cursor, perm);
  return hipGetLastError();
}

hipError_t launch_permute(const DevParticles& src, const DevParticles& dst, const int32_t* perm, int64_t n,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_permute, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, src, dst, perm, n);
  return hipGetLastError();
}

}  // namespace mpic
