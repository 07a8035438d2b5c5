// Learned far-field correction for Barnes-Hut (no reference counterpart; the method it corrects is
// the dual-monopole walk of src/interaction/barnes-hut.cc:183-221).
//
// For an accepted node and a charge sign s, the dual monopole places the node's sign-s charge Q_s
// at its charge centre c_s.  Its error is led by the quadrupole of those charges, so the residual
// is learned in the dimensionless form
//     y = dF_s * R^4 / (sigma_s Q_s h^2),   dF_s = sum_{j in s} q_j (t - x_j)/|t - x_j|^3 - sigma_s Q_s r/R^3
// (r = t - c_s, R = |r|, h = node half width, sigma_s = +-1; units e and 2^-20 m), from 13 features:
//     u = r/R (3), a = h/R (1), (c_s - c_geo)/h (3), S_s/(Q_s h^2) (6: xx yy zz xy xz yz)
// padded to 16.  Network 16 -> 64 -> 64 -> 3, SiLU, inputs standardised, outputs scaled.  The
// network predicts y by itself (default), or, in the quadrupole mode, what the analytic quadrupole
// term of the same features leaves (octupole and beyond, which the features determine only in part).
//
// Inference: BF16 weights and activations on MFMA (v_mfma_f32_32x32x16_bf16), FP32 accumulation,
// 32 rows per batch, SiLU from an interpolated table in LDS.  Training: the same BF16 rounding of
// weights and activations in the forward pass, exact SiLU, FP32 gradients (straight through the
// rounding), Adam; a fused per-32-row-tile kernel writes per-tile gradient partials that a second
// kernel reduces in a fixed order.
// Fallback: a row whose prediction exceeds the bound |y| <= ybound, or whose h/R exceeds
// amax, is corrected with the exact residual summed over the node's sign-s particles.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bh.hpp"

namespace am {

constexpr int kIn = 16, kH = 64, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kOut, kParams = kOffB3 + kOut;  // 5443
constexpr int kTrainRows = 32;

__host__ __device__ inline float Bf16Round(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}
__host__ __device__ inline uint16_t Bf16Bits(float f) {
  const float r = Bf16Round(f);
  uint32_t u;
  memcpy(&u, &r, 4);
  return (uint16_t)(u >> 16);
}
// SiLU and its derivative, shared by the training and inference kernels.
__device__ inline float SigmoidFast(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ inline float Silu(float z) { return z * SigmoidFast(z); }

// The inference kernels take SiLU from a table in LDS: sigma at z_k = -16 + k/32 (k = 0..1024),
// linearly interpolated, so the quarter-rate exp and reciprocal per hidden value become a few
// full-rate operations and one LDS read.  Interpolation error: h^2/8 max|sigma''| = 1.2e-5 in sigma
// (h = 1/32), far below the bf16 rounding of the activation that follows; outside [-16, 16] sigma
// is clamped to its end values (within 1.2e-7 of 0 and 1).  Entry k holds (sigma_k,
// sigma_{k+1} - sigma_k).
constexpr int kSiluTab = 1024;

__device__ inline float SiluTab(const float2* __restrict__ tab, float z) {
  const float t = fminf(fmaxf(fmaf(z, 32.f, 512.f), 0.f), 1023.99f);  // NaN z gives t = 0
  const int i = (int)t;
  const float2 e = tab[i];
  return z * fmaf(t - (float)i, e.y, e.x);
}
__device__ inline float SiluGrad(float z) {
  const float s = SigmoidFast(z);
  return s * (1.f + z * (1.f - s));
}

struct MlpNorm {
  float in_mean[kIn], in_std[kIn], in_istd[kIn], out_scale[kOut];  // in_istd = 1 / in_std
};

// ---- training ------------------------------------------------------------------------------------
// One workgroup per tile of 32 rows: forward (BF16-rounded operands), loss, backward, and the
// tile's gradient into part[tile][kParams].  X: normalised features [rows][16]; Y: scaled labels.
__global__ __launch_bounds__(256) void MlpTrainTileKernel(int rows, const float* __restrict__ X,
                                                          const float* __restrict__ Y, const float* __restrict__ w,
                                                          float inv_count, float* __restrict__ part,
                                                          float* __restrict__ loss_part) {
  __shared__ float xs[kTrainRows][kIn];
  __shared__ float z1[kTrainRows][kH], a1[kTrainRows][kH], z2[kTrainRows][kH], a2[kTrainRows][kH];
  __shared__ float dp[kTrainRows][kOut];
  __shared__ float dz2[kTrainRows][kH], dz1[kTrainRows][kH];
  __shared__ float lsum[256];
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * kTrainRows;
  float* g = part + (size_t)blockIdx.x * kParams;
  for (int e = t; e < kTrainRows * kIn; e += 256) {
    const int r = e / kIn, k = e % kIn;
    xs[r][k] = r0 + r < rows ? Bf16Round(X[(size_t)(r0 + r) * kIn + k]) : 0.f;
  }
  __syncthreads();
  for (int e = t; e < kTrainRows * kH; e += 256) {
    const int r = e / kH, j = e % kH;
    float acc = w[kOffB1 + j];
    for (int k = 0; k < kIn; ++k) acc = fmaf(xs[r][k], Bf16Round(w[kOffW1 + k * kH + j]), acc);
    z1[r][j] = acc;
    a1[r][j] = Bf16Round(Silu(acc));
  }
  __syncthreads();
  for (int e = t; e < kTrainRows * kH; e += 256) {
    const int r = e / kH, j = e % kH;
    float acc = w[kOffB2 + j];
    for (int k = 0; k < kH; ++k) acc = fmaf(a1[r][k], Bf16Round(w[kOffW2 + k * kH + j]), acc);
    z2[r][j] = acc;
    a2[r][j] = Bf16Round(Silu(acc));
  }
  __syncthreads();
  float l = 0.f;
  if (t < kTrainRows * kOut) {
    const int r = t / kOut, o = t % kOut;
    float acc = w[kOffB3 + o];
    for (int k = 0; k < kH; ++k) acc = fmaf(a2[r][k], Bf16Round(w[kOffW3 + k * kOut + o]), acc);
    const bool valid = r0 + r < rows;
    const float d = valid ? acc - Y[(size_t)(r0 + r) * kOut + o] : 0.f;
    l = d * d;
    dp[r][o] = 2.f * d * inv_count;
  }
  lsum[t] = l;
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if (t < s2) lsum[t] += lsum[t + s2];
    __syncthreads();
  }
  if (t == 0) loss_part[blockIdx.x] = lsum[0];
  // layer 3 gradients
  if (t < kH * kOut) {
    const int j = t / kOut, o = t % kOut;
    float acc = 0.f;
    for (int r = 0; r < kTrainRows; ++r) acc = fmaf(a2[r][j], dp[r][o], acc);
    g[kOffW3 + j * kOut + o] = acc;
  }
  if (t < kOut) {
    float acc = 0.f;
    for (int r = 0; r < kTrainRows; ++r) acc += dp[r][t];
    g[kOffB3 + t] = acc;
  }
  for (int e = t; e < kTrainRows * kH; e += 256) {
    const int r = e / kH, j = e % kH;
    float acc = 0.f;
    for (int o = 0; o < kOut; ++o) acc = fmaf(dp[r][o], Bf16Round(w[kOffW3 + j * kOut + o]), acc);
    dz2[r][j] = acc * SiluGrad(z2[r][j]);
  }
  __syncthreads();
  for (int e = t; e < kH * kH; e += 256) {
    const int i = e / kH, j = e % kH;
    float acc = 0.f;
    for (int r = 0; r < kTrainRows; ++r) acc = fmaf(a1[r][i], dz2[r][j], acc);
    g[kOffW2 + i * kH + j] = acc;
  }
  if (t < kH) {
    float acc = 0.f;
    for (int r = 0; r < kTrainRows; ++r) acc += dz2[r][t];
    g[kOffB2 + t] = acc;
  }
  for (int e = t; e < kTrainRows * kH; e += 256) {
    const int r = e / kH, i = e % kH;
    float acc = 0.f;
    for (int j = 0; j < kH; ++j) acc = fmaf(dz2[r][j], Bf16Round(w[kOffW2 + i * kH + j]), acc);
    dz1[r][i] = acc * SiluGrad(z1[r][i]);
  }
  __syncthreads();
  for (int e = t; e < kIn * kH; e += 256) {
    const int k = e / kH, j = e % kH;
    float acc = 0.f;
    for (int r = 0; r < kTrainRows; ++r) acc = fmaf(xs[r][k], dz1[r][j], acc);
    g[kOffW1 + k * kH + j] = acc;
  }
  if (t < kH) {
    float acc = 0.f;
    for (int r = 0; r < kTrainRows; ++r) acc += dz1[r][t];
    g[kOffB1 + t] = acc;
  }
}

// Sums the tiles' gradients in tile order; the batch's summed squared error goes to *loss.
__global__ void GradReduceKernel(int tiles, const float* __restrict__ part, const float* __restrict__ loss_part,
                                 float* __restrict__ grad, float* __restrict__ loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kParams) {
    float acc = 0.f;
    for (int b = 0; b < tiles; ++b) acc += part[(size_t)b * kParams + p];
    grad[p] = acc;
  }
  if (p == kParams) {
    float acc = 0.f;
    for (int b = 0; b < tiles; ++b) acc += loss_part[b];
    *loss = acc;
  }
}

// The training loop's update in one launch: each thread sums one parameter's tile gradients in tile
// order (as GradReduceKernel does) and applies the Adam step to it.
__global__ void GradReduceAdamKernel(int tiles, const float* __restrict__ part, const float* __restrict__ loss_part,
                                     double* __restrict__ epoch_loss, float inv_count, float* __restrict__ w,
                                     float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                                     float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kParams) {
    float g = 0.f;
    for (int b = 0; b < tiles; ++b) g += part[(size_t)b * kParams + p];
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    w[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
  }
  if (p == kParams) {
    float acc = 0.f;
    for (int b = 0; b < tiles; ++b) acc += loss_part[b];
    *epoch_loss += (double)(acc * inv_count);
  }
}

// Epoch shuffle on the device: a Philox draw per row (counter (row, epoch), key = seed), then a
// stable sort of (draw, row), so ties keep row order and the order depends only on the seed.
__global__ void ShuffleKeysKernel(int64_t rows, uint64_t seed, int ep, uint64_t* __restrict__ keys,
                                  uint32_t* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  uint32_t c[4] = {(uint32_t)r, (uint32_t)((uint64_t)r >> 32), (uint32_t)ep, 11u};
  Philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  keys[r] = c[0];
  vals[r] = (uint32_t)r;
}

__global__ void ZeroEpochLossKernel(int n, double* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// Rows of the normalised dataset in the epoch's order.
__global__ void GatherRowsKernel(int64_t rows, const uint32_t* __restrict__ perm, const float* __restrict__ xn,
                                 const float* __restrict__ yn, float* __restrict__ bx, float* __restrict__ by) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int64_t q = perm[r];
  for (int k = 0; k < kIn; ++k) bx[r * kIn + k] = xn[q * kIn + k];
  for (int k = 0; k < kOut; ++k) by[r * kOut + k] = yn[q * kOut + k];
}

// ---- MFMA inference --------------------------------------------------------------------------------
// v_mfma_f32_32x32x16_bf16 (gfx950): A 32x16, B 16x32, C/D 32x32 FP32.  Lane l (n = l%32, h = l/32)
// holds A[n][8h+j] and B[8h+j][n] (j = 0..7) and D[8(v/4) + 4h + v%4][n] (v = 0..15).
// The network runs transposed, H^T = W^T X^T, with the 32 rows of a batch as the N dimension: the
// weights are the A operand, and a layer's output fragment is the next layer's B operand once the
// hidden units are taken in the order in which D holds them,
//     U(ks, h, j) = 16 ks + 8 (j/4) + 4 h + j%4    (k-step ks = 0..3, B position 8h + j),
// i.e. B_ks[j] = D_{ks/2}[8 (ks%2) + j].  Activations never leave registers.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kFragL1 = 2, kFragL2 = 8, kFragL3 = 4, kFrags = kFragL1 + kFragL2 + kFragL3;  // 14

__host__ __device__ inline int HiddenOrder(int ks, int h, int j) { return 16 * ks + 8 * (j / 4) + 4 * h + (j % 4); }

// A-operand fragment f for lane l (8 bf16): layer 1 (f = mb): W1[8h+j][32 mb + n]; layer 2
// (f = 2 + 4 mb + ks): W2[U(ks,h,j)][32 mb + n]; layer 3 (f = 10 + ks): W3[U(ks,h,j)][n], n < 3.
__global__ void MlpFragKernel(const float* __restrict__ w, uint16_t* __restrict__ frag) {
  const int f = blockIdx.x, l = threadIdx.x;  // 64 lanes
  const int n = l % 32, h = l / 32;
  for (int j = 0; j < 8; ++j) {
    float v = 0.f;
    if (f < kFragL1) {
      v = w[kOffW1 + (8 * h + j) * kH + 32 * f + n];
    } else if (f < kFragL1 + kFragL2) {
      const int mb = (f - kFragL1) / 4, ks = (f - kFragL1) % 4;
      v = w[kOffW2 + HiddenOrder(ks, h, j) * kH + 32 * mb + n];
    } else {
      const int ks = f - kFragL1 - kFragL2;
      v = n < kOut ? w[kOffW3 + HiddenOrder(ks, h, j) * kOut + n] : 0.f;
    }
    frag[((size_t)f * 64 + l) * 8 + j] = Bf16Bits(v);
  }
}

struct MlpDev {
  const uint16_t* frag;  // kFrags x 64 lanes x 8 bf16
  const float* w;        // FP32 master (biases are used in FP32)
  MlpNorm norm;
  int baseline;          // add the analytic quadrupole term to the network output
};

// Quadrupole term of the label from the raw features (u = f[0..2], S/(Q h^2) = f[7..12]):
// the dipole of one sign vanishes about its charge centre, so the leading residual is
// -grad of r.Qt.r / (2 R^5) with Qt = 3 S - tr(S) I, i.e. in label units
//     y_q = -(Qt u - 5/2 (u.Qt.u) u).
__host__ __device__ inline void QuadBaseline(const float* f, float y[3]) {
  const float ux = f[0], uy = f[1], uz = f[2];
  const float tr = f[7] + f[8] + f[9];
  const float qxx = 3.f * f[7] - tr, qyy = 3.f * f[8] - tr, qzz = 3.f * f[9] - tr;
  const float qxy = 3.f * f[10], qxz = 3.f * f[11], qyz = 3.f * f[12];
  const float qux = qxx * ux + qxy * uy + qxz * uz;
  const float quy = qxy * ux + qyy * uy + qyz * uz;
  const float quz = qxz * ux + qyz * uy + qzz * uz;
  const float uqu = ux * qux + uy * quy + uz * quz;
  y[0] = 2.5f * uqu * ux - qux;
  y[1] = 2.5f * uqu * uy - quy;
  y[2] = 2.5f * uqu * uz - quz;
}

// Per-wave network state: weight fragments in registers; FP32 biases of the hidden layers and the
// SiLU table in LDS (one wave per workgroup).
struct MlpWave {
  bf16x8 wf[kFrags];
  float b3[kOut], oscale[kOut];
};

struct MlpLds {
  float4 bias4[2 * kH / 4];
  float2 silu[kSiluTab];
  __device__ const float* bias() const { return reinterpret_cast<const float*>(bias4); }
};

__device__ inline void MlpWaveLoad(const MlpDev& net, MlpLds& lds, MlpWave& s) {
  const int l = threadIdx.x & 63;
  float* bias = reinterpret_cast<float*>(lds.bias4);
  bias[l] = net.w[kOffB1 + l];
  bias[kH + l] = net.w[kOffB2 + l];
  for (int k = l; k < kSiluTab; k += 64) {
    const float z = -16.f + (float)k * (1.f / 32.f);
    const float s0 = SigmoidFast(z);
    lds.silu[k] = make_float2(s0, SigmoidFast(z + 1.f / 32.f) - s0);
  }
  const bf16x8* p = reinterpret_cast<const bf16x8*>(net.frag);
  for (int f = 0; f < kFrags; ++f) s.wf[f] = p[f * 64 + l];
  for (int o = 0; o < kOut; ++o) {
    s.b3[o] = net.w[kOffB3 + o];
    s.oscale[o] = net.norm.out_scale[o];
  }
  __syncthreads();
}

// This lane's B operand of layer 1: raw features f (16) of row n, standardised, features 8h..8h+7.
__device__ inline bf16x8 MlpInput(const MlpNorm& nm, const float f[kIn], bool ok) {
  const int h = (threadIdx.x & 63) >> 5;
  bf16x8 xb;
  for (int j = 0; j < 8; ++j) {
    const float fa = h ? f[8 + j] : f[j];
    const float mu = h ? nm.in_mean[8 + j] : nm.in_mean[j];
    const float is = h ? nm.in_istd[8 + j] : nm.in_istd[j];
    xb[j] = (__bf16)(ok ? (fa - mu) * is : 0.f);
  }
  return xb;
}

// Both halves of a row's standardised features (features 0..7 and 8..15), for a lane that owns a
// whole row.
__device__ inline void MlpInputHalves(const MlpNorm& nm, const float f[kIn], bool ok, bf16x8& lo, bf16x8& hi) {
  for (int j = 0; j < 8; ++j) {
    lo[j] = (__bf16)(ok ? (f[j] - nm.in_mean[j]) * nm.in_istd[j] : 0.f);
    hi[j] = (__bf16)(ok ? (f[8 + j] - nm.in_mean[8 + j]) * nm.in_istd[8 + j] : 0.f);
  }
}

// v from lane l ^ 32 (the other half of the wave).
__device__ inline bf16x8 OtherHalf(const bf16x8 v) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  u32x4 u = __builtin_bit_cast(u32x4, v);
  for (int k = 0; k < 4; ++k) u[k] = (uint32_t)__shfl_xor((int)u[k], 32);
  return __builtin_bit_cast(bf16x8, u);
}

__device__ inline void HiddenFromBias(const float* bias, f32x16& acc) {
  const int h = (threadIdx.x & 63) >> 5;
  for (int g = 0; g < 4; ++g) {
    const float4 b = *reinterpret_cast<const float4*>(bias + 8 * g + 4 * h);
    acc[4 * g + 0] = b.x;
    acc[4 * g + 1] = b.y;
    acc[4 * g + 2] = b.z;
    acc[4 * g + 3] = b.w;
  }
}

__device__ inline void SiluPack(const float2* __restrict__ tab, const f32x16 acc[2], bf16x8 hb[4]) {
  for (int ks = 0; ks < 4; ++ks)
    for (int j = 0; j < 8; ++j) hb[ks][j] = (__bf16)SiluTab(tab, acc[ks >> 1][8 * (ks & 1) + j]);
}

// One batch of 32 rows (features 0..7 of row n in lane n, 8..15 in lane n + 32).  Returns the
// de-normalised outputs of row n in lane n (lanes 32..63 return unspecified values).
__device__ inline void MlpForwardRegs(const MlpWave& s, const MlpLds& lds, const bf16x8 xb, float y[kOut]) {
  const float* bias_lds = lds.bias();
  f32x16 acc[2];
  for (int mb = 0; mb < 2; ++mb) {
    HiddenFromBias(bias_lds + 32 * mb, acc[mb]);
    acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(s.wf[mb], xb, acc[mb], 0, 0, 0);
  }
  bf16x8 hb[4];
  SiluPack(lds.silu, acc, hb);
  for (int mb = 0; mb < 2; ++mb) {
    HiddenFromBias(bias_lds + kH + 32 * mb, acc[mb]);
    for (int ks = 0; ks < 4; ++ks)
      acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(s.wf[kFragL1 + 4 * mb + ks], hb[ks], acc[mb], 0, 0, 0);
  }
  SiluPack(lds.silu, acc, hb);
  f32x16 o;
  for (int v = 0; v < 16; ++v) o[v] = 0.f;
  for (int ks = 0; ks < 4; ++ks)
    o = __builtin_amdgcn_mfma_f32_32x32x16_bf16(s.wf[kFragL1 + kFragL2 + ks], hb[ks], o, 0, 0, 0);
  for (int k = 0; k < kOut; ++k) y[k] = (o[k] + s.b3[k]) * s.oscale[k];
}

// Batched inference (raw features in, de-normalised predictions out), one wave per 32 rows.
__global__ __launch_bounds__(64) void MlpInferKernel(int rows, const float* __restrict__ X, MlpDev net,
                                                     float* __restrict__ Yout) {
  __shared__ MlpLds lds;
  MlpWave ws;
  MlpWaveLoad(net, lds, ws);
  const int l = threadIdx.x, n = l & 31;
  const int r = blockIdx.x * 32 + n;
  float f[kIn];
  for (int k = 0; k < kIn; ++k) f[k] = r < rows ? X[(size_t)r * kIn + k] : 0.f;
  float yb[3] = {0.f, 0.f, 0.f};
  if (net.baseline) QuadBaseline(f, yb);
  float y[kOut];
  MlpForwardRegs(ws, lds, MlpInput(net.norm, f, true), y);
  if (l < 32 && r < rows)
    for (int o = 0; o < kOut; ++o) Yout[(size_t)r * kOut + o] = y[o] + yb[o];
}

// ---- features, labels and the correction ------------------------------------------------------------
// Raw features of (target t, node c, sign s) and the monopole geometry; returns false if the node
// has no sign-s charge.  Units: 2^-20 m and e.
__device__ inline bool RowFeatures(const float4 th, const float4 tl, const WalkNodeAux& nf, const FeatNode& fe, int sgn,
                                   float feat[kIn], float& R, float& Qs_signed, float& h) {
  const float* f = sgn == 0 ? fe.fp : fe.fn;
  const float Q = f[9];
  if (!(Q > 0.f)) return false;
  const float4 ch = sgn == 0 ? nf.p_hi : nf.n_hi, cl = sgn == 0 ? nf.p_lo : nf.n_lo;
  const float rx = (th.x - ch.x) + (tl.x - cl.x), ry = (th.y - ch.y) + (tl.y - cl.y), rz = (th.z - ch.z) + (tl.z - cl.z);
  R = sqrtf(rx * rx + ry * ry + rz * rz);
  h = fe.h;
  const float ir = 1.f / R;
  feat[0] = rx * ir;
  feat[1] = ry * ir;
  feat[2] = rz * ir;
  feat[3] = h * ir;
  for (int k = 0; k < 9; ++k) feat[4 + k] = f[k];
  feat[13] = feat[14] = feat[15] = 0.f;
  Qs_signed = sgn == 0 ? Q : -Q;
  return true;
}

// Exact residual (e/unit^2) of node c's sign-s particles on target i, FP32 pair math.
__device__ inline void ExactResidual(uint32_t i, const float4 th, const float4 tl, const WalkNode& nd,
                                     const WalkNodeAux& nf, int sgn, const uint32_t* __restrict__ order,
                                     const float4* __restrict__ src, float out[3]) {
  float fx = 0.f, fy = 0.f, fz = 0.f;
  for (int k = nd.b; k < nd.e; ++k) {
    const uint32_t j = order[k];
    if (j == i) continue;
    const float4 sh = src[2 * j];
    if (sgn == 0 ? !(sh.w > 0.f) : !(sh.w < 0.f)) continue;
    PairF32(th, tl, sh, src[2 * j + 1], fx, fy, fz);
  }
  float mx = 0.f, my = 0.f, mz = 0.f;
  PairF32(th, tl, sgn == 0 ? nf.p_hi : nf.n_hi, sgn == 0 ? nf.p_lo : nf.n_lo, mx, my, mz);
  out[0] = fx - mx;
  out[1] = fy - my;
  out[2] = fz - mz;
}

// The same residual summed by the whole wave, lane l taking particles b + l, b + l + 64, ...; the
// monopole is subtracted in lane 0 only, so the wave's lanes add up to the residual.
__device__ inline void ExactResidualWave(uint32_t i, const float4 th, const float4 tl, const WalkNode& nd,
                                         const WalkNodeAux& nf, int sgn, const uint32_t* __restrict__ order,
                                         const float4* __restrict__ src, int lane, float out[3]) {
  float fx = 0.f, fy = 0.f, fz = 0.f;
  for (int k = nd.b + lane; k < nd.e; k += 64) {
    const uint32_t j = order[k];
    if (j == i) continue;
    const float4 sh = src[2 * j];
    if (sgn == 0 ? !(sh.w > 0.f) : !(sh.w < 0.f)) continue;
    PairF32(th, tl, sh, src[2 * j + 1], fx, fy, fz);
  }
  if (lane == 0) {
    float mx = 0.f, my = 0.f, mz = 0.f;
    PairF32(th, tl, sgn == 0 ? nf.p_hi : nf.n_hi, sgn == 0 ? nf.p_lo : nf.n_lo, mx, my, mz);
    fx -= mx;
    fy -= my;
    fz -= mz;
  }
  out[0] = fx;
  out[1] = fy;
  out[2] = fz;
}

// Fallback rows over nodes of more than this many particles are summed by the whole wave: in one
// lane, a node of S particles holds the wave for S pair iterations.
constexpr int kFallbackLaneMax = 64;

// Training rows: for sampled target k (sorted position tsel[k]) and its recorded node slot,
// both signs: raw features [16] and the normalised label y (FP64 exact residual).
__global__ void MlpDataKernel(int n_targets, const uint32_t* __restrict__ tsel, const uint32_t* __restrict__ rec,
                              int rec_cap, const uint32_t* __restrict__ rec_cnt, const uint32_t* __restrict__ row_off,
                              const WalkNode* __restrict__ nodes, const WalkNodeAux* __restrict__ nodes32,
                              const FeatNode* __restrict__ feat, const uint32_t* __restrict__ order,
                              const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                              const double* __restrict__ q, const float4* __restrict__ src, float* __restrict__ X,
                              float* __restrict__ Y, uint8_t* __restrict__ valid, int32_t* __restrict__ tgt,
                              int32_t* __restrict__ nid, float* __restrict__ scl) {
  const int k = blockIdx.y;
  if (k >= n_targets) return;
  const int cnt = min((int)rec_cnt[k], rec_cap);
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= cnt) return;
  const uint32_t i = order[tsel[k]];
  const int c = (int)rec[(size_t)k * rec_cap + slot];
  const WalkNode nd = nodes[c];
  const WalkNodeAux nf = nodes32[c];
  const FeatNode fe = feat[c];
  const float4 th = src[2 * i], tl = src[2 * i + 1];
  const double tx = x[i] * kUnitScale, ty = y[i] * kUnitScale, tz = z[i] * kUnitScale;
  for (int sgn = 0; sgn < 2; ++sgn) {
    const size_t row = (size_t)row_off[k] * 2 + 2 * slot + sgn;
    float f[kIn];
    float R, Qs, h;
    const bool ok = RowFeatures(th, tl, nf, fe, sgn, f, R, Qs, h);
    valid[row] = ok ? 1 : 0;
    for (int a = 0; a < kIn; ++a) X[row * kIn + a] = ok ? f[a] : 0.f;
    double dx = 0.0, dy = 0.0, dz = 0.0, lscale = 0.0;
    if (ok) {
      const double cs[3] = {(sgn == 0 ? nf.pqx : nf.nqx) * kUnitScale, (sgn == 0 ? nf.pqy : nf.nqy) * kUnitScale,
                            (sgn == 0 ? nf.pqz : nf.nqz) * kUnitScale};
      for (int kk = nd.b; kk < nd.e; ++kk) {
        const uint32_t j = order[kk];
        const double qj = q[j] / kElementaryCharge;
        if (j == i || (sgn == 0 ? !(qj > 0.0) : !(qj < 0.0))) continue;
        const double rx = tx - x[j] * kUnitScale, ry = ty - y[j] * kUnitScale, rz = tz - z[j] * kUnitScale;
        const double r2 = rx * rx + ry * ry + rz * rz;
        const double ir3 = 1.0 / (r2 * sqrt(r2));
        dx += qj * rx * ir3;
        dy += qj * ry * ir3;
        dz += qj * rz * ir3;
      }
      const double rx = tx - cs[0], ry = ty - cs[1], rz = tz - cs[2];
      const double r2 = rx * rx + ry * ry + rz * rz;
      const double ir3 = 1.0 / (r2 * sqrt(r2));
      const double Qd = (sgn == 0 ? nd.qp : -nd.qn) / kElementaryCharge;
      const double hd = nd.half * kUnitScale;
      dx -= Qd * rx * ir3;
      dy -= Qd * ry * ir3;
      dz -= Qd * rz * ir3;
      const double norm = r2 * r2 / (Qd * hd * hd);
      lscale = 1.0 / norm;
      dx *= norm;
      dy *= norm;
      dz *= norm;
    }
    tgt[row] = (int32_t)i;
    nid[row] = c;
    scl[row] = (float)lscale;
    Y[row * kOut + 0] = (float)dx;
    Y[row * kOut + 1] = (float)dy;
    Y[row * kOut + 2] = (float)dz;
  }
}

// Correction pass: one wave per target (sorted position k); rows = (recorded node, sign), one
// per lane, 64 per pass as two MLP batches of 32.  Lane l computes the features of row base + l
// once; the halves that the other batch's B operand needs cross the wave (lane l ^ 32), and batch
// 2's outputs come back to the lanes of their rows the same way.  Adds the predicted residual (or,
// past the bound, the exact one) to the force.  Each lane sums its rows' corrections in FP64 and
// the wave adds the 64 partial sums once per target in a fixed order: deterministic.
constexpr int kCorrectTake = 8;  // targets per queue take
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void MlpCorrectKernel(int m, const uint32_t* __restrict__ rec, int rec_cap,
                                                       const uint32_t* __restrict__ rec_cnt,
                                                       const WalkNode* __restrict__ nodes,
                                                       const WalkNodeAux* __restrict__ nodes32,
                                                       const FeatNode* __restrict__ feat,
                                                       const uint32_t* __restrict__ order,
                                                       const double* __restrict__ q, const float4* __restrict__ src,
                                                       MlpDev net, float ybound, float amax,
                                                       double* __restrict__ fx, double* __restrict__ fy,
                                                       double* __restrict__ fz, unsigned long long* __restrict__ stats,
                                                       unsigned* __restrict__ queue) {
  __shared__ MlpLds lds;
  MlpWave ws;
  MlpWaveLoad(net, lds, ws);
  const int l = threadIdx.x;
  const bool upper = l >= 32;
  unsigned long long n_model = 0, n_fallback = 0, n_over = 0;
  // Persistent over targets: the weight fragments stay in registers for the whole launch, and each wave
  // takes kCorrectTake targets at a time from a queue, so the waves that run ahead (a SIMD issues from
  // its oldest waves first) take more; the grid is the number of waves that fit at once.
  for (;;) {
    unsigned k0 = 0;
    if (l == 0) k0 = atomicAdd(&queue[0], (unsigned)kCorrectTake);
    k0 = __builtin_amdgcn_readfirstlane(k0);
    if ((int)k0 >= m) break;
    const int k1 = min(m, (int)k0 + kCorrectTake);
    for (int k = (int)k0; k < k1; ++k) {
      const uint32_t i = order[k];
      const double qi = q[i];
      const float4 th = src[2 * i], tl = src[2 * i + 1];
      const int cnt = min((int)rec_cnt[k], rec_cap);
      n_over += (int)rec_cnt[k] > rec_cap ? 1 : 0;
      const int nrows = 2 * cnt;
      double cx = 0.0, cy = 0.0, cz = 0.0;
      for (int base = 0; base < nrows; base += 64) {
        const int row = base + l;
        bool ok = false;
        float R = 1.f, Qs = 0.f, hh = 0.f;
        float yb[3] = {0.f, 0.f, 0.f};
        float f[kIn];
        for (int a = 0; a < kIn; ++a) f[a] = 0.f;
        int c = 0;
        if (row < nrows) {
          c = (int)rec[(size_t)k * rec_cap + row / 2];
          ok = RowFeatures(th, tl, nodes32[c], feat[c], row & 1, f, R, Qs, hh);
          if (ok && net.baseline) QuadBaseline(f, yb);
        }
        bf16x8 lo, hi;
        MlpInputHalves(net.norm, f, ok, lo, hi);
        // batch 1 = rows base .. base+31, batch 2 = rows base+32 .. base+63: lane n < 32 feeds
        // features 0..7 of row n of its batch, lane n + 32 features 8..15 of row n
        const bf16x8 got = OtherHalf(upper ? lo : hi);
        float y1[kOut], y2[kOut] = {0.f, 0.f, 0.f};
        MlpForwardRegs(ws, lds, upper ? got : lo, y1);
        if (base + 32 < nrows) MlpForwardRegs(ws, lds, upper ? hi : got, y2);
        float y[kOut];
        for (int o = 0; o < kOut; ++o) {
          const float moved = __shfl_xor(y2[o], 32);  // batch 2's row n + 32 is output in lane n
          y[o] = upper ? moved : y1[o];
        }
        bool wave_fallback = false;
        if (ok) {
          float corr[3] = {0.f, 0.f, 0.f};
          const float p0 = y[0] + yb[0], p1 = y[1] + yb[1], p2 = y[2] + yb[2];
          const float a = hh / R;
          const float mag = sqrtf(p0 * p0 + p1 * p1 + p2 * p2);
          if (mag <= ybound && a <= amax && mag == mag) {
            const float sc = Qs * hh * hh / (R * R * R * R);
            corr[0] = p0 * sc;
            corr[1] = p1 * sc;
            corr[2] = p2 * sc;
            ++n_model;
          } else if (nodes[c].e - nodes[c].b <= kFallbackLaneMax) {
            ExactResidual(i, th, tl, nodes[c], nodes32[c], row & 1, order, src, corr);
            ++n_fallback;
          } else {
            wave_fallback = true;
          }
          cx += corr[0];
          cy += corr[1];
          cz += corr[2];
        }
        // the large-node fallback rows of this batch, one at a time over the whole wave
        for (unsigned long long mask = __ballot(wave_fallback); mask; mask &= mask - 1) {
          const int owner = __builtin_ctzll(mask);
          const int fc = __builtin_amdgcn_readfirstlane(__shfl(c, owner));
          const int fs = __builtin_amdgcn_readfirstlane(__shfl(row & 1, owner));
          float corr[3];
          ExactResidualWave(i, th, tl, nodes[fc], nodes32[fc], fs, order, src, l, corr);
          cx += corr[0];
          cy += corr[1];
          cz += corr[2];
          if (l == owner) ++n_fallback;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        cx += __shfl_xor(cx, off);
        cy += __shfl_xor(cy, off);
        cz += __shfl_xor(cz, off);
      }
      if (l == 0 && qi != 0.0) {
        const double scale = kCoulomb * qi * kElementaryCharge * kInvUnit2;
        fx[i] += scale * cx;
        fy[i] += scale * cy;
        fz[i] += scale * cz;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    n_model += __shfl_xor(n_model, off);
    n_fallback += __shfl_xor(n_fallback, off);
  }
  if (l == 0 && stats) {
    atomicAdd(&stats[0], n_model);
    atomicAdd(&stats[1], n_fallback);
    atomicAdd(&stats[2], n_over);
  }
  if (l == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {  // one wave per workgroup
      queue[0] = 0u;
      queue[1] = 0u;
      __threadfence();
    }
  }
}

}  // namespace am

using namespace am;

struct am_mlp {
  am_ctx* ctx = nullptr;
  float* w = nullptr;  // kParams FP32 master weights
  float* grad = nullptr;
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  uint16_t* frag = nullptr;
  float* loss = nullptr;
  int64_t adam_step = 0;
  MlpNorm norm{};
  int baseline = 0;
  bool frag_valid = false;
};

namespace am {

static int MlpRefreshFrag(am_mlp* net) {
  if (net->frag_valid) return AM_OK;
  hipLaunchKernelGGL(MlpFragKernel, dim3(kFrags), dim3(64), 0, net->ctx->stream, (const float*)net->w, net->frag);
  AM_HIP(hipGetLastError());
  net->frag_valid = true;
  return AM_OK;
}

static MlpDev MlpView(am_mlp* net) { return MlpDev{net->frag, net->w, net->norm, net->baseline}; }

// Host -> device copy ordered on the context stream (kernels queued there may still read dst);
// returns once the copy is done, so the host buffer can be reused.
static int UploadSync(am_ctx* ctx, void* dst, const void* src, size_t bytes) {
  AM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return AM_OK;
}

}  // namespace am

extern "C" {

int am_mlp_create(am_ctx* ctx, uint64_t seed, am_mlp** out) {
  if (!ctx || !out) return Fail(AM_ERR_INVALID, "am_mlp_create: NULL argument");
  AM_HIP(hipSetDevice(ctx->device));
  am_mlp* net = new am_mlp();
  net->ctx = ctx;
  hipError_t e = hipMalloc((void**)&net->w, kParams * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&net->grad, kParams * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&net->adam_m, kParams * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&net->adam_v, kParams * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&net->frag, (size_t)kFrags * 64 * 8 * sizeof(uint16_t));
  if (e == hipSuccess) e = hipMalloc((void**)&net->loss, 16);
  if (e != hipSuccess) {
    am_mlp_destroy(net);
    return HipFail(e, "am_mlp_create", __FILE__, __LINE__);
  }
  // He-style initialisation from a counter-based generator (deterministic in seed).
  std::vector<float> w(kParams, 0.f), zero(kParams, 0.f);
  auto init = [&](int off, int fan_in, int count) {
    const double sd = std::sqrt(2.0 / fan_in);
    for (int k = 0; k < count; ++k) {
      uint32_t c[4] = {(uint32_t)(off + k), 0u, 7u, 3u};
      Philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
      const double u1 = ((double)(c[0] >> 8) + 0.5) / 16777216.0, u2 = (double)(c[1] >> 8) / 16777216.0;
      w[off + k] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2));
    }
  };
  init(kOffW1, kIn, kIn * kH);
  init(kOffW2, kH, kH * kH);
  init(kOffW3, kH, kH * kOut);
  for (int k = 0; k < kOut; ++k) w[kOffW3 + k] *= 0.1f;
  for (int k = 0; k < kIn; ++k) {
    net->norm.in_mean[k] = 0.f;
    net->norm.in_std[k] = 1.f;
    net->norm.in_istd[k] = 1.f;
  }
  for (int k = 0; k < kOut; ++k) net->norm.out_scale[k] = 1.f;
  int rc;
  if ((rc = UploadSync(ctx, net->w, w.data(), kParams * 4)) || (rc = UploadSync(ctx, net->adam_m, zero.data(), kParams * 4)) ||
      (rc = UploadSync(ctx, net->adam_v, zero.data(), kParams * 4))) {
    am_mlp_destroy(net);
    return rc;
  }
  *out = net;
  return AM_OK;
}

int am_mlp_destroy(am_mlp* net) {
  if (!net) return AM_OK;
  void* p[] = {net->w, net->grad, net->adam_m, net->adam_v, net->frag, net->loss};
  for (void* a : p)
    if (a) (void)hipFree(a);
  delete net;
  return AM_OK;
}

int am_mlp_num_params(void) { return kParams; }

int am_mlp_set_baseline(am_mlp* net, int quadrupole) {
  if (!net) return Fail(AM_ERR_INVALID, "am_mlp_set_baseline: NULL network");
  net->baseline = quadrupole ? 1 : 0;
  return AM_OK;
}

int am_mlp_quadrupole_baseline(const float* X, int64_t rows, float* Y) {
  if (!X || !Y || rows < 0) return Fail(AM_ERR_INVALID, "am_mlp_quadrupole_baseline: bad arguments");
  for (int64_t r = 0; r < rows; ++r) QuadBaseline(X + r * kIn, Y + r * kOut);
  return AM_OK;
}

int am_mlp_set_params(am_mlp* net, const float* w) {
  if (!net || !w) return Fail(AM_ERR_INVALID, "am_mlp_set_params: NULL argument");
  int rc;
  if ((rc = UploadSync(net->ctx, net->w, w, kParams * 4))) return rc;
  net->frag_valid = false;
  return AM_OK;
}

int am_mlp_get_params(am_mlp* net, float* w) {
  if (!net || !w) return Fail(AM_ERR_INVALID, "am_mlp_get_params: NULL argument");
  AM_HIP(hipMemcpyAsync(w, net->w, kParams * 4, hipMemcpyDeviceToHost, net->ctx->stream));
  AM_HIP(hipStreamSynchronize(net->ctx->stream));
  return AM_OK;
}

// Input standardisation (mean, std per 16 features) and output scale (3).
int am_mlp_set_normalization(am_mlp* net, const float* in_mean, const float* in_std, const float* out_scale) {
  if (!net || !in_mean || !in_std || !out_scale) return Fail(AM_ERR_INVALID, "am_mlp_set_normalization: NULL");
  for (int k = 0; k < kIn; ++k) {
    if (!(in_std[k] > 0.f)) return Fail(AM_ERR_INVALID, "am_mlp_set_normalization: std must be > 0");
    net->norm.in_mean[k] = in_mean[k];
    net->norm.in_std[k] = in_std[k];
    net->norm.in_istd[k] = 1.f / in_std[k];
  }
  for (int k = 0; k < kOut; ++k) net->norm.out_scale[k] = out_scale[k];
  return AM_OK;
}

// Forward + backward on a device batch (X normalised [rows][16], Y scaled labels [rows][3]):
// gradient of the mean squared error into grad (host, kParams) and the loss.  No update.
static int MlpGradients(am_mlp* net, const float* dX, const float* dY, int rows, float* loss_out) {
  am_ctx* ctx = net->ctx;
  const int tiles = (rows + kTrainRows - 1) / kTrainRows;
  void *part, *lpart;
  int rc;
  if ((rc = Scratch(ctx, "mlp_part", (size_t)tiles * kParams * 4, &part))) return rc;
  if ((rc = Scratch(ctx, "mlp_lpart", (size_t)tiles * 4, &lpart))) return rc;
  const float inv_count = 1.f / ((float)rows * kOut);
  hipLaunchKernelGGL(MlpTrainTileKernel, dim3(tiles), dim3(256), 0, ctx->stream, rows, dX, dY, (const float*)net->w,
                     inv_count, (float*)part, (float*)lpart);
  hipLaunchKernelGGL(GradReduceKernel, dim3(Blocks(kParams + 1, 256)), dim3(256), 0, ctx->stream, tiles,
                     (const float*)part, (const float*)lpart, net->grad, net->loss);
  AM_HIP(hipGetLastError());
  if (loss_out) {
    float l = 0.f;
    AM_HIP(hipMemcpyAsync(&l, net->loss, 4, hipMemcpyDeviceToHost, ctx->stream));
    AM_HIP(hipStreamSynchronize(ctx->stream));
    *loss_out = l * inv_count;
  }
  return AM_OK;
}

// Gradient check entry point: host batch in, host gradient out.
int am_mlp_gradients(am_mlp* net, const float* X, const float* Y, int rows, float* grad, float* loss) {
  if (!net || !X || !Y || rows <= 0) return Fail(AM_ERR_INVALID, "am_mlp_gradients: bad arguments");
  am_ctx* ctx = net->ctx;
  void *dX, *dY;
  int rc;
  if ((rc = Scratch(ctx, "mlp_gx", (size_t)rows * kIn * 4, &dX))) return rc;
  if ((rc = Scratch(ctx, "mlp_gy", (size_t)rows * kOut * 4, &dY))) return rc;
  AM_HIP(hipMemcpyAsync(dX, X, (size_t)rows * kIn * 4, hipMemcpyHostToDevice, ctx->stream));
  AM_HIP(hipMemcpyAsync(dY, Y, (size_t)rows * kOut * 4, hipMemcpyHostToDevice, ctx->stream));
  if ((rc = MlpGradients(net, (const float*)dX, (const float*)dY, rows, loss))) return rc;
  if (grad) {
    AM_HIP(hipMemcpyAsync(grad, net->grad, kParams * 4, hipMemcpyDeviceToHost, ctx->stream));
    AM_HIP(hipStreamSynchronize(ctx->stream));
  }
  return AM_OK;
}

// Adam training on a dataset (raw features [rows][16], labels [rows][3], host): features are
// standardised and labels divided by the output scale with the network's normalisation (once, on
// the host), the normalised set is uploaded once, and each of the `epochs` passes runs on the
// device without host round trips: a seeded shuffle (ShuffleKeysKernel and a radix sort), the
// gather of the shuffled rows, and per minibatch the gradient tiles, their reduction and the Adam
// update.  losses (optional, epochs): mean minibatch loss per epoch, summed on the device.
int am_mlp_train(am_mlp* net, const float* X, const float* Y, int64_t rows, int epochs, int batch, float lr,
                 uint64_t seed, float* losses) {
  if (!net || !X || !Y || rows <= 0 || epochs <= 0 || batch <= 0)
    return Fail(AM_ERR_INVALID, "am_mlp_train: bad arguments");
  if (rows > (int64_t)UINT32_MAX) return Fail(AM_ERR_INVALID, "am_mlp_train: more than 2^32 rows");
  am_ctx* ctx = net->ctx;
  hipStream_t st = ctx->stream;
  std::vector<float> xn((size_t)rows * kIn), yn((size_t)rows * kOut);
  for (int64_t r = 0; r < rows; ++r) {
    for (int k = 0; k < kIn; ++k)
      xn[r * kIn + k] = (X[r * kIn + k] - net->norm.in_mean[k]) * net->norm.in_istd[k];  // as on the device
    float yb[3] = {0.f, 0.f, 0.f};
    if (net->baseline) QuadBaseline(X + r * kIn, yb);  // the network learns what the quadrupole leaves
    for (int k = 0; k < kOut; ++k) yn[r * kOut + k] = (Y[r * kOut + k] - yb[k]) / net->norm.out_scale[k];
  }
  void *dxn, *dyn, *bX, *bY, *keys, *perm, *eloss;
  int rc;
  if ((rc = Scratch(ctx, "mlp_xn", (size_t)rows * kIn * 4, &dxn))) return rc;
  if ((rc = Scratch(ctx, "mlp_yn", (size_t)rows * kOut * 4, &dyn))) return rc;
  if ((rc = Scratch(ctx, "mlp_bx", (size_t)rows * kIn * 4, &bX))) return rc;
  if ((rc = Scratch(ctx, "mlp_by", (size_t)rows * kOut * 4, &bY))) return rc;
  if ((rc = Scratch(ctx, "mlp_shuffle_keys", (size_t)rows * 8, &keys))) return rc;
  if ((rc = Scratch(ctx, "mlp_shuffle_perm", (size_t)rows * 4, &perm))) return rc;
  if ((rc = Scratch(ctx, "mlp_epoch_loss", (size_t)epochs * 8, &eloss))) return rc;
  if ((rc = UploadSync(ctx, dxn, xn.data(), (size_t)rows * kIn * 4))) return rc;
  if ((rc = UploadSync(ctx, dyn, yn.data(), (size_t)rows * kOut * 4))) return rc;
  hipLaunchKernelGGL(ZeroEpochLossKernel, dim3(Blocks(epochs, 256)), dim3(256), 0, st, epochs, (double*)eloss);
  AM_HIP(hipGetLastError());
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  std::vector<int> nbatches(epochs, 0);
  for (int ep = 0; ep < epochs; ++ep) {
    hipLaunchKernelGGL(ShuffleKeysKernel, dim3(Blocks(rows, 256)), dim3(256), 0, st, rows, seed, ep, (uint64_t*)keys,
                       (uint32_t*)perm);
    AM_HIP(hipGetLastError());
    if ((rc = RadixSortPairs(ctx, (uint64_t*)keys, (uint32_t*)perm, rows, 32))) return rc;
    hipLaunchKernelGGL(GatherRowsKernel, dim3(Blocks(rows, 256)), dim3(256), 0, st, rows, (const uint32_t*)perm,
                       (const float*)dxn, (const float*)dyn, (float*)bX, (float*)bY);
    AM_HIP(hipGetLastError());
    for (int64_t r0 = 0; r0 < rows; r0 += batch) {
      const int nrow = (int)std::min<int64_t>(batch, rows - r0);
      const int tiles = (nrow + kTrainRows - 1) / kTrainRows;
      void *part, *lpart;
      if ((rc = Scratch(ctx, "mlp_part", (size_t)tiles * kParams * 4, &part))) return rc;
      if ((rc = Scratch(ctx, "mlp_lpart", (size_t)tiles * 4, &lpart))) return rc;
      const float inv_count = 1.f / ((float)nrow * kOut);
      hipLaunchKernelGGL(MlpTrainTileKernel, dim3(tiles), dim3(256), 0, st, nrow, (const float*)bX + r0 * kIn,
                         (const float*)bY + r0 * kOut, (const float*)net->w, inv_count, (float*)part, (float*)lpart);
      ++nbatches[ep];
      ++net->adam_step;
      const float c1 = 1.f / (1.f - std::pow(b1, (float)net->adam_step));
      const float c2 = 1.f / (1.f - std::pow(b2, (float)net->adam_step));
      hipLaunchKernelGGL(GradReduceAdamKernel, dim3(Blocks(kParams + 1, 256)), dim3(256), 0, st, tiles,
                         (const float*)part, (const float*)lpart, (double*)eloss + ep, inv_count, net->w,
                         net->adam_m, net->adam_v, lr, b1, b2, eps, c1, c2);
      AM_HIP(hipGetLastError());
    }
  }
  std::vector<double> el(epochs);
  AM_HIP(hipMemcpyAsync(el.data(), eloss, (size_t)epochs * 8, hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  // the shuffles' sorts report a look-back give-up here (their permutation would be wrong)
  if ((rc = CheckSortError(ctx))) return rc;
  if (losses)
    for (int ep = 0; ep < epochs; ++ep) losses[ep] = (float)(el[ep] / std::max(nbatches[ep], 1));
  net->frag_valid = false;
  return AM_OK;
}

// Batched inference on the MFMA path: raw features in, predictions (de-normalised) out (host).
int am_mlp_infer(am_mlp* net, const float* X, int rows, float* Y) {
  if (!net || !X || !Y || rows <= 0) return Fail(AM_ERR_INVALID, "am_mlp_infer: bad arguments");
  am_ctx* ctx = net->ctx;
  int rc;
  if ((rc = MlpRefreshFrag(net))) return rc;
  void *dX, *dY;
  if ((rc = Scratch(ctx, "mlp_ix", (size_t)rows * kIn * 4, &dX))) return rc;
  if ((rc = Scratch(ctx, "mlp_iy", (size_t)rows * kOut * 4, &dY))) return rc;
  AM_HIP(hipMemcpyAsync(dX, X, (size_t)rows * kIn * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(MlpInferKernel, dim3((rows + 31) / 32), dim3(64), 0, ctx->stream, rows, (const float*)dX,
                     MlpView(net), (float*)dY);
  AM_HIP(hipGetLastError());
  AM_HIP(hipMemcpyAsync(Y, dY, (size_t)rows * kOut * 4, hipMemcpyDeviceToHost, ctx->stream));
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return AM_OK;
}

// Training data from the store's cloud: n_targets sampled sorted positions (every stride-th of the
// tree order, from offset), Barnes-Hut walk at theta recording up to rec_cap accepted nodes per
// target, then two rows (signs) per recorded node: raw features [16] and normalised exact residual
// [3], with (optional) the target particle index, the node (preorder position) and the label
// scale sigma_s Q_s h^2 / R^4 of each row.  Rows of a sign absent from the node are dropped.  Returns the number of rows in *rows_out;
// the outputs must hold 2 * n_targets * rec_cap rows.
int am_bh_learned_training_data(am_store* s, double theta, int n_targets, int stride, int offset, int rec_cap,
                                float* X, float* Y, int32_t* target, int32_t* node, float* scale,
                                int64_t* rows_out) {
  if (!s || !X || !Y || !rows_out || n_targets <= 0 || stride <= 0 || offset < 0 || rec_cap <= 0)
    return Fail(AM_ERR_INVALID, "am_bh_learned_training_data: bad arguments");
  if (n_targets > 65535) return Fail(AM_ERR_CAPACITY, "am_bh_learned_training_data: at most 65535 targets per call");
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  AM_HIP(hipSetDevice(ctx->device));
  BHTree t;
  int rc;
  if ((rc = BuildBHTree(s, 8, &t, true))) return rc;
  *rows_out = 0;
  if (t.m == 0) return AM_OK;
  std::vector<uint32_t> sel;
  for (int k = 0; k < n_targets; ++k) {
    const int64_t p = (int64_t)offset + (int64_t)k * stride;
    if (p >= t.m) break;
    sel.push_back((uint32_t)p);
  }
  const int nt = (int)sel.size();
  if (nt == 0) return AM_OK;
  void *dsel, *rec, *cnt, *off, *dX, *dY, *dv, *tmpf, *dt, *dn, *ds;
  if ((rc = Scratch(ctx, "mlpd_sel", (size_t)nt * 4, &dsel))) return rc;
  if ((rc = Scratch(ctx, "mlpd_rec", (size_t)nt * rec_cap * 4, &rec))) return rc;
  if ((rc = Scratch(ctx, "mlpd_cnt", (size_t)nt * 4, &cnt))) return rc;
  if ((rc = Scratch(ctx, "mlpd_off", (size_t)(nt + 1) * 4, &off))) return rc;
  if ((rc = Scratch(ctx, "mlpd_force", (size_t)s->n * 3 * 8, &tmpf))) return rc;
  AM_HIP(hipMemcpyAsync(dsel, sel.data(), (size_t)nt * 4, hipMemcpyHostToDevice, st));
  double* tf = (double*)tmpf;
  if ((rc = LaunchBHWalkF32(s, t, theta, (const uint32_t*)dsel, nt, tf, tf + s->n, tf + 2 * s->n, (uint32_t*)rec,
                            rec_cap, (uint32_t*)cnt)))
    return rc;
  std::vector<uint32_t> hc(nt), ho(nt + 1, 0);
  AM_HIP(hipMemcpyAsync(hc.data(), cnt, (size_t)nt * 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  int maxc = 0;
  for (int k = 0; k < nt; ++k) {
    const int c = std::min<int>(hc[k], rec_cap);
    ho[k + 1] = ho[k] + c;
    maxc = std::max(maxc, c);
  }
  const int64_t slots = ho[nt];
  if (slots == 0) return AM_OK;
  if ((rc = Scratch(ctx, "mlpd_X", (size_t)slots * 2 * kIn * 4, &dX))) return rc;
  if ((rc = Scratch(ctx, "mlpd_Y", (size_t)slots * 2 * kOut * 4, &dY))) return rc;
  if ((rc = Scratch(ctx, "mlpd_valid", (size_t)slots * 2, &dv))) return rc;
  if ((rc = Scratch(ctx, "mlpd_tgt", (size_t)slots * 2 * 4, &dt))) return rc;
  if ((rc = Scratch(ctx, "mlpd_scl", (size_t)slots * 2 * 4, &ds))) return rc;
  if ((rc = Scratch(ctx, "mlpd_nid", (size_t)slots * 2 * 4, &dn))) return rc;
  AM_HIP(hipMemcpyAsync(off, ho.data(), (size_t)(nt + 1) * 4, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(MlpDataKernel, dim3(Blocks(maxc, 64), nt), dim3(64), 0, st, nt, (const uint32_t*)dsel,
                     (const uint32_t*)rec, rec_cap, (const uint32_t*)cnt, (const uint32_t*)off,
                     (const WalkNode*)t.nodes, (const WalkNodeAux*)t.nodes32, (const FeatNode*)t.feat,
                     (const uint32_t*)t.order, s->x, s->y, s->z, s->q, (const float4*)s->src, (float*)dX,
                     (float*)dY, (uint8_t*)dv, (int32_t*)dt, (int32_t*)dn,
                     (float*)ds);
  AM_HIP(hipGetLastError());
  std::vector<float> hx((size_t)slots * 2 * kIn), hy((size_t)slots * 2 * kOut);
  std::vector<uint8_t> hv((size_t)slots * 2);
  std::vector<int32_t> htg((size_t)slots * 2), hn((size_t)slots * 2);
  AM_HIP(hipMemcpyAsync(hn.data(), dn, hn.size() * 4, hipMemcpyDeviceToHost, st));
  std::vector<float> hs((size_t)slots * 2);
  AM_HIP(hipMemcpyAsync(htg.data(), dt, htg.size() * 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipMemcpyAsync(hs.data(), ds, hs.size() * 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipMemcpyAsync(hx.data(), dX, hx.size() * 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipMemcpyAsync(hy.data(), dY, hy.size() * 4, hipMemcpyDeviceToHost, st));
  AM_HIP(hipMemcpyAsync(hv.data(), dv, hv.size(), hipMemcpyDeviceToHost, st));
  AM_HIP(hipStreamSynchronize(st));
  int64_t w = 0;
  for (int64_t r = 0; r < slots * 2; ++r) {
    if (!hv[r]) continue;
    std::memcpy(X + w * kIn, &hx[r * kIn], kIn * 4);
    std::memcpy(Y + w * kOut, &hy[r * kOut], kOut * 4);
    if (target) target[w] = htg[r];
    if (node) node[w] = hn[r];
    if (scale) scale[w] = hs[r];
    ++w;
  }
  *rows_out = w;
  return AM_OK;
}

// Barnes-Hut with the learned far-field correction (FP32 walk).  rec_cap: accepted nodes recorded
// per particle (nodes past it keep the plain monopole and are counted); ybound / amax: fallback
// bounds on |y| and h/R.  stats (optional, 3): rows corrected by the model, rows that fell back to
// the exact residual, particles whose node list overflowed rec_cap.
int am_compute_barnes_hut_learned_forces(am_store* s, double theta, am_mlp* net, int rec_cap, float ybound,
                                         float amax, int64_t* stats) {
  if (!s || !net || rec_cap <= 0) return Fail(AM_ERR_INVALID, "am_compute_barnes_hut_learned_forces: bad arguments");
  if (!(theta >= 0.0)) return Fail(AM_ERR_INVALID, "barnes_hut_theta must be >= 0");
  if (net->ctx != s->ctx) return Fail(AM_ERR_INVALID, "am_compute_barnes_hut_learned_forces: network on another context");
  if (s->n == 0) return AM_OK;
  if (s->n > (int64_t)1 << 30) return Fail(AM_ERR_CAPACITY, "Barnes-Hut: n too large");
  if ((int64_t)s->n * rec_cap > ((int64_t)1 << 34))
    return Fail(AM_ERR_CAPACITY, "am_compute_barnes_hut_learned_forces: n * rec_cap exceeds 2^34 node records");
  am_ctx* ctx = s->ctx;
  hipStream_t st = ctx->stream;
  AM_HIP(hipSetDevice(ctx->device));
  int rc;
  if ((rc = MlpRefreshFrag(net))) return rc;
  BHTree t;
  if ((rc = BuildBHTree(s, 8, &t, true))) return rc;
  if ((rc = ZeroNonFiniteForces(s))) return rc;  // they are in no tree
  if (t.m == 0) return AM_OK;
  void *rec, *cnt, *dstats;
  if ((rc = Scratch(ctx, "bhl_rec", (size_t)t.m * rec_cap * 4, &rec))) return rc;
  if ((rc = Scratch(ctx, "bhl_cnt", (size_t)t.m * 4, &cnt))) return rc;
  if ((rc = Scratch(ctx, "bhl_stats", 32, &dstats))) return rc;
  AM_HIP(hipMemsetAsync(dstats, 0, 32, st));
  if ((rc = LaunchBHWalkF32(s, t, theta, nullptr, t.m, s->fx, s->fy, s->fz, (uint32_t*)rec, rec_cap,
                            (uint32_t*)cnt)))
    return rc;
  int slot;
  TimerBegin(ctx, "bh_learned_correct", &slot);
  unsigned* queue;
  if ((rc = QueueCounters(ctx, "bhl_queue", &queue))) return rc;
  // 2 waves per SIMD fit (the weight fragments take about 230 VGPRs): 8 one-wave workgroups per CU.
  const int grid = std::min((t.m + kCorrectTake - 1) / kCorrectTake, std::max(1, ctx->num_cus) * 8);
  hipLaunchKernelGGL(MlpCorrectKernel, dim3(grid), dim3(64), 0, st, t.m, (const uint32_t*)rec, rec_cap,
                     (const uint32_t*)cnt, (const WalkNode*)t.nodes, (const WalkNodeAux*)t.nodes32,
                     (const FeatNode*)t.feat, (const uint32_t*)t.order, s->q, (const float4*)s->src, MlpView(net),
                     ybound, amax, s->fx, s->fy, s->fz, (unsigned long long*)dstats, queue);
  TimerEnd(ctx, slot);
  AM_HIP(hipGetLastError());
  if (stats) {
    unsigned long long h[3];
    AM_HIP(hipMemcpyAsync(h, dstats, 24, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 3; ++k) stats[k] = (int64_t)h[k];
  }
  return AM_OK;
}

}  // extern "C"
