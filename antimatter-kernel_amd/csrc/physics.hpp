// Per-particle push physics shared by the single-store and ensemble push kernels:
// LorentzForce (src/forces/lorentz.cc:18-71) then CyclotronCooling (src/forces/cooling.cc:16-65),
// FP64 and uncontracted (the library is built with -ffp-contract=off), in the operation order of
// the reference (and of the CPU restatement).
#pragma once

#include <cmath>

#include "am_internal.hpp"

namespace am {

struct PushParams {
  double b0, u0, d;  // PenningTrapField
  double dt;
  int lorentz;        // run LorentzForce
  double tau_scale;   // CyclotronCooling (<= 0: off)
  double bath_temp;
  uint64_t seed, step;
};

struct D3 {
  double x, y, z;
};
__device__ inline D3 Add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 Sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 Scale(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline double Dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 Cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// BorisVelocity (lorentz.cc:18-34).
__device__ inline D3 Boris(D3 v, D3 f, D3 b, double q, double m, double dt) {
  const double h = 0.5 * dt;
  const D3 kick = Scale(f, h / m);
  const D3 vm = Add(v, kick);
  const D3 t = Scale(b, q * h / m);
  const D3 vp = Add(vm, Cross(vm, t));
  const double t2 = Dot(t, t);
  const D3 s = Scale(t, 2.0 / (1.0 + t2));
  return Add(Add(vm, Cross(vp, s)), kick);
}

// Three standard normals (Box-Muller on Philox words), purpose word 2.
__device__ inline void CoolingNormals(uint64_t seed, uint64_t step, uint32_t i, double g[3]) {
  uint32_t c[4] = {(uint32_t)step, i, 0u, 2u};
  Philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u1 = ((double)(c[0] >> 8) + 0.5) * (1.0 / 16777216.0);
  const double u2 = (double)(c[1] >> 8) * (1.0 / 16777216.0);
  const double u3 = ((double)(c[2] >> 8) + 0.5) * (1.0 / 16777216.0);
  const double u4 = (double)(c[3] >> 8) * (1.0 / 16777216.0);
  const double r1 = sqrt(-2.0 * log(u1));
  const double r2 = sqrt(-2.0 * log(u3));
  g[0] = r1 * cos(2.0 * kPi * u2);
  g[1] = r1 * sin(2.0 * kPi * u2);
  g[2] = r2 * cos(2.0 * kPi * u4);
}

__device__ inline double CoolingTimeDev(double q, double m, double b) {
  const double q2 = q * q;
  const double denom = q2 * q2 * b * b;
  if (denom <= 0.0) return INFINITY;
  return (0.75 / kCoulomb) * m * m * m * kLight * kLight * kLight / denom;
}

// One particle: Lorentz push (optional) then cooling (optional); updates x, v, F, flags and the
// FP32 source mirror.  Annihilated particles are skipped (lorentz.cc:42, cooling.cc:30).
// rng_index keys the particle's cooling-noise draws.
__device__ inline void PushParticle(int64_t i, double* __restrict__ x, double* __restrict__ y,
                                    double* __restrict__ z, double* __restrict__ vx, double* __restrict__ vy,
                                    double* __restrict__ vz, double* __restrict__ fx, double* __restrict__ fy,
                                    double* __restrict__ fz, const double* __restrict__ q,
                                    const double* __restrict__ m, uint8_t* __restrict__ flags,
                                    float4* __restrict__ src, const PushParams& p, uint32_t rng_index) {
  uint8_t fl = flags[i];
  if (fl & kFlagAnnihilated) return;
  const double qi = q[i], mi = m[i];
  D3 pos{x[i], y[i], z[i]};
  D3 vel{vx[i], vy[i], vz[i]};
  const D3 B{0.0, 0.0, p.b0};
  bool moved = false;
  if (p.lorentz && !(mi <= 0.0 || p.dt <= 0.0)) {
    // PenningTrapField::GetElectricField (penning_trap.cc:22-33).
    const double d2 = p.d * p.d;
    const double radial = p.u0 / (2.0 * d2);
    const double axial = p.u0 / d2;
    const D3 E{radial * pos.x, radial * pos.y, -axial * pos.z};
    const D3 F = Add(Scale(E, qi), D3{fx[i], fy[i], fz[i]});
    if (!(fl & kFlagStaggered)) {  // first call: v(0) -> v(-dt/2), lorentz.cc:56-60
      vel = Boris(vel, F, B, qi, mi, -0.5 * p.dt);
      fl |= kFlagStaggered;
    }
    vel = Boris(vel, F, B, qi, mi, p.dt);
    pos = Add(pos, Scale(vel, p.dt));
    fx[i] = 0.0;
    fy[i] = 0.0;
    fz[i] = 0.0;
    moved = true;
  }
  if (p.tau_scale > 0.0 && !(qi == 0.0 || mi <= 0.0 || p.dt <= 0.0)) {  // cooling.cc:35
    const double bmag = sqrt(Dot(B, B));
    if (bmag > 0.0) {
      const double tau = p.tau_scale * CoolingTimeDev(fabs(qi), mi, bmag);
      if (isfinite(tau)) {
        const D3 bh = Scale(B, 1.0 / bmag);
        const double vpar = Dot(vel, bh);
        const D3 vparv = Scale(bh, vpar);
        D3 vperp = Sub(vel, vparv);
        const double decay = exp(-p.dt / (2.0 * tau));
        vperp = Scale(vperp, decay);
        if (p.bath_temp > 0.0) {
          const double sigma = sqrt(kBoltzmann * p.bath_temp / mi * (1.0 - decay * decay));
          double g[3];
          CoolingNormals(p.seed, p.step, rng_index, g);
          D3 noise{sigma * g[0], sigma * g[1], sigma * g[2]};
          noise = Sub(noise, Scale(bh, Dot(noise, bh)));
          vperp = Add(vperp, noise);
        }
        vel = Add(vparv, vperp);
      }
    }
  }
  vx[i] = vel.x;
  vy[i] = vel.y;
  vz[i] = vel.z;
  flags[i] = fl;
  if (moved) {
    x[i] = pos.x;
    y[i] = pos.y;
    z[i] = pos.z;
    WriteMirror(src, i, pos.x, pos.y, pos.z, qi, false);
  }
}

}  // namespace am
