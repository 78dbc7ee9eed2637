// Device particle fill: ParticleFill / FillFunctor (particle_fill.hpp:59-164)
// with one thread per (element, ipart) over the elements [e0, e0 + ecount)
// (the whole mesh, or one problem of an ensemble).  Particle index = offset +
// (e - e0)*nppe + ipart; the seed formula takes the element index and element
// count of that range.
//   seed = elem + (34030+ipart)*Nglob + Nglob*12345*type   (u64 wrap, particle_fill.hpp:147)
//   10 burned draws; reference position by rejection (sum > 1 rejected);
//   x = refToPhysical; velocity by set_vel (particle_fill.hpp:101-124);
//   weight = density*(detJ/6)/nppe.
// Extensions: seed += seed_offset * 0x9E3779B97F4A7C15 (0 = reference seeds);
// weight *= 1 + pert_amp*sin(pert_k*(x0 - lo0)).
#include "mpic_internal.hpp"
#include "mpic_device.hpp"

namespace mpic {
namespace {

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
  const int64_t q = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const int64_t total = (int64_t)a.ecount * a.nppe;
  if (q >= total) return;
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
    vv[0] = a.vel[0] * (1. + temp * sin(1. * 3.14159265358979323846 * ph[0]));
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
  if (a.pert_amp != 0.0) wp = wp * (1.0 + a.pert_amp * sin(a.pert_k * (ph[0] - a.lo0)));
  const int64_t p = a.offset + q;
  Rec64 o;
  for (int d = 0; d < 3; ++d) {
    o.x[d] = ph[d];
    o.xr[d] = ref[d];
    o.v[d] = vv[d];
  }
  o.w = wp;
  o.elem = e;
  o.id = (uint32_t)p;
  o.type = (uint32_t)a.type;
  o.pad = 0;
  store_rec(a.P.r, p, o);
  a.P.elem[p] = e;
}

}  // namespace

hipError_t launch_fill(const FillArgs& a, hipStream_t s) {
  const int64_t total = (int64_t)a.ecount * a.nppe;
  if (total <= 0) return hipSuccess;
  if (a.e0 < 0 || (int64_t)a.e0 + a.ecount > a.m.ne) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((total + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

}  // namespace mpic
