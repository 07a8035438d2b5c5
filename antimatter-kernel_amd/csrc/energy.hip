// ComputeTotalEnergy (src/analysis/energy_analysis.cc:84-145) on the device: FP64 per-particle
// terms reduced in a fixed two-stage tree (deterministic for a given n), time-centring on the
// host exactly as the reference pairs consecutive calls.
#include <algorithm>
#include <cmath>

#include "am_internal.hpp"

namespace am {

int StandardEnergyTerms(am_store* s, double rc, double* out);
int AllPairsPotentialTerms(am_store* s, int precision, double* out, double* si_scale);
int BarnesHutPotentialTerms(am_store* s, double theta, int precision, double* out, double* si_scale);
int FastMultipolePotentialEnergy(am_store* s, int order, double theta, double* pe);

constexpr int kReduceThreads = 256;
constexpr int kReduceBlocks = 512;

__global__ __launch_bounds__(kReduceThreads) void ReducePartialKernel(const double* __restrict__ in, int64_t n,
                                                                      double* __restrict__ partial) {
  __shared__ double sh[kReduceThreads];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kReduceThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kReduceThreads)
    s += in[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kReduceThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kReduceThreads) void ReduceFinalKernel(const double* __restrict__ partial, int count,
                                                                    double* __restrict__ out) {
  __shared__ double sh[kReduceThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += kReduceThreads) s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kReduceThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// Deterministic FP64 sum of n device doubles into *host_out (synchronises the stream).
int ReduceSum(am_ctx* ctx, const double* in, int64_t n, double* host_out) {
  if (n <= 0) { *host_out = 0.0; return AM_OK; }
  void* part;
  int rc = Scratch(ctx, "reduce_partial", (kReduceBlocks + 1) * sizeof(double), &part);
  if (rc) return rc;
  const int blocks = (int)std::min<int64_t>(kReduceBlocks, (n + kReduceThreads - 1) / kReduceThreads);
  double* p = (double*)part;
  hipLaunchKernelGGL(ReducePartialKernel, dim3(blocks), dim3(kReduceThreads), 0, ctx->stream, in, n, p);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(ReduceFinalKernel, dim3(1), dim3(kReduceThreads), 0, ctx->stream, (const double*)p, blocks,
                     p + kReduceBlocks);
  AM_HIP(hipGetLastError());
  AM_HIP(hipMemcpyAsync(host_out, p + kReduceBlocks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return AM_OK;
}

// KE and trap-potential terms (energy_analysis.cc:137-146): ke_i = 0.5 m v.v,
// pt_i = q phi(x) with phi = (U0 / 4 d^2)(2 z^2 - rho^2) (penning_trap.cc:35-43).
__global__ void KineticTrapKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                  const double* __restrict__ z, const double* __restrict__ vx,
                                  const double* __restrict__ vy, const double* __restrict__ vz,
                                  const double* __restrict__ q, const double* __restrict__ m,
                                  const uint8_t* __restrict__ flags, double u0, double d, double* __restrict__ ke,
                                  double* __restrict__ pt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flags[i] & kFlagAnnihilated) {
    ke[i] = 0.0;
    pt[i] = 0.0;
    return;
  }
  const double a = vx[i], b = vy[i], c = vz[i];
  ke[i] = 0.5 * m[i] * Dot3(a, b, c, a, b, c);
  const double px = x[i], py = y[i], pz = z[i];
  const double rho2 = px * px + py * py;
  const double d2 = d * d;
  pt[i] = q[i] * ((u0 / (4.0 * d2)) * (2.0 * pz * pz - rho2));
}

}  // namespace am

using namespace am;

extern "C" int am_compute_total_energy_ex(am_store* s, const am_trap_field* field,
                                          const am_energy_settings* settings, am_energy_history* history,
                                          int precision, am_energy_breakdown* out) {
  if (!s || !field || !settings || !out) return Fail(AM_ERR_INVALID, "am_compute_total_energy: NULL argument");
  if (precision != AM_FP32 && precision != AM_FP64) return Fail(AM_ERR_INVALID, "unknown precision");
  AM_HIP(hipSetDevice(s->ctx->device));
  am_ctx* ctx = s->ctx;
  const int64_t n = s->n;
  am_energy_breakdown r{0.0, 0.0, 0.0, 0};
  int rc = AM_OK;
  if (n > 0) {
    r.num_active_particles = CountActive(s, &rc);
    if (rc) return rc;
    void* buf;
    if ((rc = Scratch(ctx, "energy_terms", (size_t)n * 3 * sizeof(double), &buf))) return rc;
    double* ke = (double*)buf;
    double* pt = ke + n;
    double* pc = pt + n;
    hipLaunchKernelGGL(KineticTrapKernel, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, n, s->x, s->y, s->z,
                       s->vx, s->vy, s->vz, s->q, s->m, s->flags, field->trap_voltage, field->d, ke, pt);
    AM_HIP(hipGetLastError());
    if ((rc = ReduceSum(ctx, ke, n, &r.kinetic))) return rc;
    if ((rc = ReduceSum(ctx, pt, n, &r.trap_potential))) return rc;
    double sum = 0.0;
    switch (settings->method) {
      case AM_METHOD_FAST_MULTIPOLE: {
        double pe = 0.0;
        const int order = std::min(std::max((int)settings->fast_multipole_order, 1), 8);
        if ((rc = FastMultipolePotentialEnergy(s, order, settings->fast_multipole_theta, &pe))) return rc;
        r.energy_potential = pe;
        break;
      }
      case AM_METHOD_BARNES_HUT:
        if (settings->barnes_hut_theta > 0.0) {
          double scale = 1.0;
          if ((rc = BarnesHutPotentialTerms(s, settings->barnes_hut_theta, precision, pc, &scale))) return rc;
          if ((rc = ReduceSum(ctx, pc, n, &sum))) return rc;
          r.energy_potential = 0.5 * kCoulomb * scale * sum;
          break;
        }
        [[fallthrough]];  // theta = 0 opens every node: the exact all-pairs potential
      case AM_METHOD_ALLPAIRS: {
        double scale = 1.0;
        if ((rc = AllPairsPotentialTerms(s, precision, pc, &scale))) return rc;
        if ((rc = ReduceSum(ctx, pc, n, &sum))) return rc;
        r.energy_potential = 0.5 * kCoulomb * scale * sum;
        break;
      }
      case AM_METHOD_STANDARD:
      default:
        if ((rc = StandardEnergyTerms(s, settings->cutoff_radius, pc))) return rc;
        if ((rc = ReduceSum(ctx, pc, n, &sum))) return rc;
        r.energy_potential = 0.5 * sum;
        break;
    }
  }
  if (history) {  // energy_analysis.cc:126-142
    const double raw_t = r.trap_potential, raw_c = r.energy_potential;
    if (history->has_previous) {
      r.trap_potential = 0.5 * (history->prev_trap_potential + raw_t);
      r.energy_potential = 0.5 * (history->prev_coulomb_potential + raw_c);
    }
    history->prev_trap_potential = raw_t;
    history->prev_coulomb_potential = raw_c;
    history->has_previous = 1;
  }
  *out = r;
  return AM_OK;
}

extern "C" int am_compute_total_energy(am_store* s, const am_trap_field* field, const am_energy_settings* settings,
                                       am_energy_history* history, am_energy_breakdown* out) {
  return am_compute_total_energy_ex(s, field, settings, history, AM_FP64, out);
}
