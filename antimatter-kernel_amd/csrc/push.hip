// Fused per-particle push: LorentzForce (src/forces/lorentz.cc:18-71) then CyclotronCooling
// (src/forces/cooling.cc:16-65), in the per-agent behaviour order of the reference driver
// (src/antimatter-kernel.cc:93-94), plus the refresh of the FP32 hi/lo source mirror used by the
// pair kernels.  All arithmetic is FP64 and uncontracted (the library is built with
// -ffp-contract=off), in the same operation order as the CPU restatement.
//
// HBM traffic per particle-step (algorithmic): read x,v,F (72 B), q,m (16 B), flags (1 B);
// write x,v (48 B), F = 0 (24 B), flags (1 B), source mirror (32 B).
#include <cmath>

#include "physics.hpp"

namespace am {

__global__ __launch_bounds__(256) void PushKernel(int64_t n, double* __restrict__ x, double* __restrict__ y,
                                                   double* __restrict__ z, double* __restrict__ vx,
                                                   double* __restrict__ vy, double* __restrict__ vz,
                                                   double* __restrict__ fx, double* __restrict__ fy,
                                                   double* __restrict__ fz, const double* __restrict__ q,
                                                   const double* __restrict__ m, uint8_t* __restrict__ flags,
                                                   float4* __restrict__ src, PushParams p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  PushParticle(i, x, y, z, vx, vy, vz, fx, fy, fz, q, m, flags, src, p, (uint32_t)i);
}

static int LaunchPush(am_store* s, const PushParams& p) {
  if (s->n == 0) return AM_OK;
  AM_HIP(hipSetDevice(s->ctx->device));
  int slot;
  TimerBegin(s->ctx, "push", &slot);
  hipLaunchKernelGGL(PushKernel, dim3(Blocks(s->n, 256)), dim3(256), 0, s->ctx->stream, s->n, s->x, s->y, s->z,
                     s->vx, s->vy, s->vz, s->fx, s->fy, s->fz, s->q, s->m, s->flags, s->src, p);
  TimerEnd(s->ctx, slot);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

}  // namespace am

using namespace am;

extern "C" {

int am_lorentz_force(am_store* s, const am_trap_field* f, double dt) {
  if (!s || !f) return Fail(AM_ERR_INVALID, "am_lorentz_force: NULL argument");
  PushParams p{f->b0, f->trap_voltage, f->d, dt, 1, 0.0, 0.0, 0, 0};
  return LaunchPush(s, p);
}

int am_cyclotron_cooling(am_store* s, const am_trap_field* f, double tau_scale, double bath_temp_K, double dt,
                         uint64_t seed, uint64_t step) {
  if (!s || !f) return Fail(AM_ERR_INVALID, "am_cyclotron_cooling: NULL argument");
  if (tau_scale <= 0.0) return AM_OK;
  PushParams p{f->b0, f->trap_voltage, f->d, dt, 0, tau_scale, bath_temp_K, seed, step};
  return LaunchPush(s, p);
}

int am_push(am_store* s, const am_trap_field* f, double dt, double tau_scale, double bath_temp_K, uint64_t seed,
            uint64_t step) {
  if (!s || !f) return Fail(AM_ERR_INVALID, "am_push: NULL argument");
  PushParams p{f->b0, f->trap_voltage, f->d, dt, 1, tau_scale, bath_temp_K, seed, step};
  return LaunchPush(s, p);
}

}  // extern "C"
