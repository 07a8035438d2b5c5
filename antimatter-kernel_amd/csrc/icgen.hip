// Seeded cloud generator on the device for large clouds (no reference counterpart at this size: the
// reference draws its initial conditions serially with libstdc++'s <random>, penning_trap.cc:51-161,
// which the host generator of ic.hip reproduces bit for bit).  Same distributions, counter-based
// draws: positions uniform in the ellipsoid of radii (a, b, c) by rejection from the bounding box,
// velocities Maxwellian per species, N(0, k_B T / m) per component, then, as the reference does,
// each species' mean velocity removed and the set rescaled by sqrt(n / (n - 1)).  Species occupy
// consecutive blocks in enum order.  The draws of particle i depend only on (seed, i), so the cloud
// does not depend on the launch shape; it is not bit-identical to the host generator.
#include <cmath>

#include "am_internal.hpp"

namespace am {

// Uniform in [-1, 1) with 24 bits from one Philox word.
__device__ inline double USym(uint32_t w) { return ((double)(w >> 8) * (1.0 / 8388608.0)) - 1.0; }

// Particle i: position by rejection (each Philox block of four words gives one candidate; the
// acceptance probability is pi/6, so 64 rounds leave a miss with probability 1e-21, after which the
// candidate is scaled onto the ellipsoid's surface), velocity from two Box-Muller pairs.
__global__ void CloudKernel(int64_t n, uint64_t seed, const int64_t* __restrict__ species_end,
                            const double* __restrict__ sigma, double ra, double rb, double rc, double* __restrict__ x,
                            double* __restrict__ y, double* __restrict__ z, double* __restrict__ vx,
                            double* __restrict__ vy, double* __restrict__ vz, uint8_t* __restrict__ species) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s = 0;
  while (s < 5 && i >= species_end[s]) ++s;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  double p[3] = {0.0, 0.0, 0.0};
  for (uint32_t round = 0; round < 64; ++round) {
    uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), round, 3u};
    Philox(c, k0, k1);
    p[0] = USym(c[0]);
    p[1] = USym(c[1]);
    p[2] = USym(c[2]);
    const double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    if (r2 <= 1.0) break;
    if (round == 63) {
      const double inv = 1.0 / sqrt(r2);
      p[0] *= inv;
      p[1] *= inv;
      p[2] *= inv;
    }
  }
  x[i] = p[0] * ra;
  y[i] = p[1] * rb;
  z[i] = p[2] * rc;
  uint32_t c[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0xffffffffu, 4u};
  Philox(c, k0, k1);
  const double u1 = ((double)(c[0] >> 8) + 0.5) * (1.0 / 16777216.0);
  const double u2 = (double)(c[1] >> 8) * (1.0 / 16777216.0);
  const double u3 = ((double)(c[2] >> 8) + 0.5) * (1.0 / 16777216.0);
  const double u4 = (double)(c[3] >> 8) * (1.0 / 16777216.0);
  const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
  const double sg = sigma[s];
  vx[i] = sg * r1 * cos(2.0 * kPi * u2);
  vy[i] = sg * r1 * sin(2.0 * kPi * u2);
  vz[i] = sg * r2 * cos(2.0 * kPi * u4);
  species[i] = (uint8_t)s;
}

// v <- (v - mean) * scale over particles [b, e) of one species (mean and scale per component set).
__global__ void CenterKernel(int64_t b, int64_t e, double mx, double my, double mz, double scale,
                             double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz) {
  const int64_t i = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= e) return;
  vx[i] = (vx[i] - mx) * scale;
  vy[i] = (vy[i] - my) * scale;
  vz[i] = (vz[i] - mz) * scale;
}

__global__ void SpeciesDefaultsKernel(int64_t n, const uint8_t* __restrict__ species, double* __restrict__ q,
                                      double* __restrict__ m, uint8_t* __restrict__ flags, uint32_t* __restrict__ trap,
                                      double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = species[i];
  // GetDefaultSpeciesProperties (src/agents/particle_species.h:27-46)
  const double mp = kProtonMass, me = kElectronMass, mn = kNeutronMass, e = kElementaryCharge;
  const double mass[6] = {mp, me, mp + me, me, mp, 2.0 * (mp + mn)};
  const double charge[6] = {-e, e, 0.0, -e, e, 0.0};
  q[i] = charge[s];
  m[i] = mass[s];
  flags[i] = 0;
  trap[i] = 0;
  fx[i] = 0.0;
  fy[i] = 0.0;
  fz[i] = 0.0;
}

}  // namespace am

using namespace am;

extern "C" int am_store_generate_cloud(am_store* s, const uint64_t counts[6], const double radii[3],
                                       const double temperatures[6], uint64_t seed) {
  if (!s || !counts || !radii || !temperatures) return Fail(AM_ERR_INVALID, "am_store_generate_cloud: NULL argument");
  if (!(radii[0] > 0.0 && radii[1] > 0.0 && radii[2] > 0.0))
    return Fail(AM_ERR_INVALID, "am_store_generate_cloud: every cloud radius must be positive");
  int64_t n = 0, end[6];
  double sigma[6];
  for (int sp = 0; sp < 6; ++sp) {
    if (!(temperatures[sp] >= 0.0)) return Fail(AM_ERR_INVALID, "am_store_generate_cloud: temperature < 0");
    n += (int64_t)counts[sp];
    end[sp] = n;
    double mass, charge;
    SpeciesProperties(sp, &mass, &charge);
    sigma[sp] = std::sqrt(kBoltzmann * temperatures[sp] / mass);  // a cold species is at rest
  }
  if (n <= 0) return Fail(AM_ERR_INVALID, "am_store_generate_cloud: the populations hold no particles");
  if (n > s->capacity) return Fail(AM_ERR_CAPACITY, "am_store_generate_cloud: n exceeds store capacity");
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  AM_HIP(hipSetDevice(ctx->device));
  void* tab;
  int rc = Scratch(ctx, "icgen_tab", 12 * sizeof(double), &tab);
  if (rc) return rc;
  AM_HIP(hipMemcpyAsync(tab, end, 6 * sizeof(int64_t), hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync((double*)tab + 6, sigma, 6 * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(CloudKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, seed, (const int64_t*)tab,
                     (const double*)tab + 6, radii[0], radii[1], radii[2], s->x, s->y, s->z, s->vx, s->vy, s->vz,
                     s->species);
  hipLaunchKernelGGL(SpeciesDefaultsKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, (const uint8_t*)s->species,
                     s->q, s->m, s->flags, s->trap, s->fx, s->fy, s->fz);
  AM_HIP(hipGetLastError());
  // Per species: remove the mean velocity and restore the expected kinetic energy
  // (penning_trap.cc:82-104), with the deterministic FP64 reduction.
  for (int sp = 0; sp < 6; ++sp) {
    const int64_t b = sp == 0 ? 0 : end[sp - 1];
    const int64_t cnt = end[sp] - b;
    if (cnt < 2 || sigma[sp] == 0.0) continue;
    double mean[3];
    double* comp[3] = {s->vx, s->vy, s->vz};
    for (int d = 0; d < 3; ++d) {
      if ((rc = ReduceSum(ctx, comp[d] + b, cnt, &mean[d]))) return rc;
      mean[d] /= (double)cnt;
    }
    const double scale = std::sqrt((double)cnt / (double)(cnt - 1));
    hipLaunchKernelGGL(CenterKernel, dim3(Blocks(cnt, 256)), dim3(256), 0, st, b, end[sp], mean[0], mean[1], mean[2],
                       scale, s->vx, s->vy, s->vz);
    AM_HIP(hipGetLastError());
  }
  s->n = n;
  s->generation += 1;
  s->perm_n = -1;
  if ((rc = RefreshSource(s))) return rc;
  AM_HIP(hipStreamSynchronize(st));
  return AM_OK;
}
