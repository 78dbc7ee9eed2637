// Device particle fill: ParticleFill / FillFunctor (particle_fill.hpp:59-164)
// with one thread per (element, ipart) over the elements [e0, e0 + ecount)
// (the whole mesh, or one problem of an ensemble).  As the reference's
// FillFunctor, each team (workgroup) builds its particles in LDS and appends
// them with the team-collective ParticleList::push_back (particle_list.hpp:103-139,
// include/minipic/particle_list_view.hpp): one atomic per workgroup.  Identity
// key = offset + (e - e0)*nppe + ipart, independent of the order in which the
// teams append; the seed formula takes the element index and element count of
// that range.
//   seed = elem + (34030+ipart)*Nglob + Nglob*12345*type   (u64 wrap, particle_fill.hpp:147)
//   10 burned draws; reference position by rejection (sum > 1 rejected);
//   x = refToPhysical; velocity by set_vel (particle_fill.hpp:101-124);
//   weight = density*(detJ/6)/nppe.
// Extensions: seed += seed_offset * 0x9E3779B97F4A7C15 (0 = reference seeds);
// weight *= 1 + pert_amp*sin(pert_k*(x0 - lo0)).  Both sines are fill_sin.
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

// The fill's sine (the beam perturbation of set_vel, particle_fill.hpp:106-111,
// and the weight perturbation): a fixed sequence of IEEE operations, so the
// fill gives the same bits wherever it runs.  Cody and Waite's reduction
// x = k pi/2 + r, |r| <= pi/4, with pi/2 split into a 33-bit head and a tail
// (the split of fdlibm's __ieee754_rem_pio2), then the Taylor polynomials of
// sin r (to r^21) and cos r (to r^20) in Horner form in z = r^2; within 1 ulp
// of a correctly rounded sine for |x| below a few thousand.  The file is built
// with -ffp-contract=off, so no product is fused into an addition.
__device__ __forceinline__ double fill_sin(double x) {
  const double k = rint(x * 0x1.45f306dc9c883p-1);  // x * 2/pi
  const double r = (x - k * 0x1.921fb54400000p+0) - k * 0x1.0b4611a626331p-34;
  const double z = r * r;
  double s = 1.9572941063391263e-20;
  s = s * z + -8.22063524662433e-18;
  s = s * z + 2.8114572543455206e-15;
  s = s * z + -7.647163731819816e-13;
  s = s * z + 1.6059043836821613e-10;
  s = s * z + -2.505210838544172e-08;
  s = s * z + 2.7557319223985893e-06;
  s = s * z + -0.0001984126984126984;
  s = s * z + 0.008333333333333333;
  s = s * z + -0.16666666666666666;
  const double sn = r + (r * z) * s;
  double c = 4.110317623312165e-19;
  c = c * z + -1.5619206968586225e-16;
  c = c * z + 4.779477332387385e-14;
  c = c * z + -1.1470745597729725e-11;
  c = c * z + 2.08767569878681e-09;
  c = c * z + -2.755731922398589e-07;
  c = c * z + 2.48015873015873e-05;
  c = c * z + -0.001388888888888889;
  c = c * z + 0.041666666666666664;
  c = c * z + -0.5;
  const double cs = 1.0 + z * c;
  const int q = (int)(long long)k & 3;
  return q == 0 ? sn : q == 1 ? cs : q == 2 ? -sn : -cs;
}

// dRandom, random.hpp:43-87: s = A s + C mod 2^64, drand = (double)s * 2^-64.
struct Lcg {
  unsigned long long s;
  __device__ __forceinline__ unsigned long long next() {
    s = 6364136223846793005ull * s + 1442695040888963407ull;
    return s;
  }
  __device__ __forceinline__ double drand() { return (double)next() * 5.421010862427522170e-20; }
};

__global__ __launch_bounds__(BLOCK) void k_fill(FillArgs a) {
  __shared__ minipic::Particle team_p[BLOCK];
  __shared__ long team_add[BLOCK];
  __shared__ int scratch[MPIC_TEAM_SCRATCH];
  const minipic::member_type dev(scratch);
  const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t total = (int64_t)a.ecount * a.nppe;
  const bool mine = q < total;
  if (mine) {
    const int el = (int)(q / a.nppe);  // element within the filled range
    const int ip = (int)(q - (int64_t)el * a.nppe);
    const int e = a.e0 + el;
    const unsigned long long N = (unsigned long long)a.ecount;
    Lcg r{(unsigned long long)el + (34030ull + (unsigned long long)ip) * N + N * 12345ull * (unsigned long long)a.type};
    r.s += a.seed_offset * 0x9E3779B97F4A7C15ull;
    for (int i = 0; i < 10; ++i) r.next();
    double ref[3];
    do {
      ref[0] = r.drand();
      ref[1] = r.drand();
      ref[2] = r.drand();
    } while (ref[0] + ref[1] + ref[2] > 1.0);
    // refToPhysical, mesh.hpp:183-190
    const double wts[4] = {1 - ref[0] - ref[1] - ref[2], ref[0], ref[1], ref[2]};
    const int4 nd = a.m.conn[e];
    const int nid[4] = {nd.x, nd.y, nd.z, nd.w};
    double ph[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 4; ++i)
      for (int d = 0; d < 3; ++d) ph[d] += wts[i] * a.m.nodes[(size_t)nid[i] * 3 + d];
    double vv[3];
    const double temp = a.temp;
    if (temp == 0) {
      vv[0] = a.vel[0];
      vv[1] = a.vel[1];
      vv[2] = a.vel[2];
    } else if (temp < 0) {
      vv[0] = a.vel[0] * (1. + temp * fill_sin(1. * 3.14159265358979323846 * ph[0]));
      vv[1] = vv[2] = 0.0;
    } else {
      while (true) {
        const double rr = r.drand();
        const double vmax = sqrt(6 * temp / a.mass);
        double u[3];
        u[0] = vmax * 2 * (-.5 + r.drand());
        u[1] = vmax * 2 * (-.5 + r.drand());
        u[2] = vmax * 2 * (-.5 + r.drand());
        const double prob = exp(-(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / temp);
        if (rr < prob) {
          vv[0] = a.vel[0] + u[0];
          vv[1] = a.vel[1] + u[1];
          vv[2] = a.vel[2] + u[2];
          break;
        }
      }
    }
    const double vol = a.m.detj[e] / 6.;
    double wp = a.density * vol / a.nppe;
    if (a.pert_amp != 0.0) wp = wp * (1.0 + a.pert_amp * fill_sin(a.pert_k * (ph[0] - a.lo0)));
    minipic::Particle& o = team_p[threadIdx.x];
    for (int d = 0; d < 3; ++d) {
      o.x[d] = ph[d];
      o.x_ref[d] = ref[d];
      o.v[d] = vv[d];
    }
    o.weight = wp;
    o.ielement = e;
    o.type = a.type;
    o.is_dead = false;
    o.id = (uint32_t)(a.offset + q);
  }
  team_add[threadIdx.x] = mine ? 1 : 0;
  a.view.push_back(dev, team_p, team_add, BLOCK);
}

__global__ void k_view_reset(unsigned long long* counters, int64_t n) {
  if (threadIdx.x < 4) counters[threadIdx.x] = threadIdx.x == 0 ? (unsigned long long)n : 0ull;
}

__global__ __launch_bounds__(BLOCK) void k_elem_copy(DevParticles P, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) P.elem[i] = P.r[i].elem;
}

}  // namespace

hipError_t launch_fill(const FillArgs& a, hipStream_t s) {
  const int64_t total = (int64_t)a.ecount * a.nppe;
  if (total <= 0) return hipSuccess;
  if (a.e0 < 0 || (int64_t)a.e0 + a.ecount > a.m.ne || !a.view.rec || !a.view.counters ||
      a.offset + total > a.view.capacity)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_view_reset, dim3(1), dim3(64), 0, s, a.view.counters, a.offset);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((total + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_view_reset(unsigned long long* counters, int64_t n, hipStream_t s) {
  if (!counters) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_view_reset, dim3(1), dim3(64), 0, s, counters, n);
  return hipGetLastError();
}

hipError_t launch_elem_copy(const DevParticles& P, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_elem_copy, dim3(grid_for(n)), dim3(BLOCK), 0, s, P, n);
  return hipGetLastError();
}

}  // namespace mpic
