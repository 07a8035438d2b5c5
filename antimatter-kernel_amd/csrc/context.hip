// Context, particle store, errors and timing.
#include <cmath>
#include <cstdio>

#include "am_internal.hpp"

namespace am {

static thread_local std::string g_error;

void SetError(const std::string& msg) { g_error = msg; }

int Fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

int HipFail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file,
                line);
  g_error = buf;
  return AM_ERR_HIP;
}

int Scratch(am_ctx* ctx, const char* name, size_t bytes, void** out) {
  Buffer& b = ctx->scratch[name];
  if (b.bytes < bytes) {
    if (b.ptr) AM_HIP(hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
    size_t want = bytes + bytes / 4 + 256;
    AM_HIP(hipMalloc(&b.ptr, want));
    b.bytes = want;
  }
  *out = b.ptr;
  return AM_OK;
}

static hipEvent_t TakeEvent(am_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void TimerBegin(am_ctx* ctx, const char* name, int* slot) {
  *slot = -1;
  if (!ctx->timing) return;
  am_ctx::Pending p;
  p.name = name;
  p.start = TakeEvent(ctx);
  p.stop = TakeEvent(ctx);
  if (!p.start || !p.stop || hipEventRecord(p.start, ctx->stream) != hipSuccess) {  // timing is best effort
    if (p.start) ctx->event_pool.push_back(p.start);
    if (p.stop) ctx->event_pool.push_back(p.stop);
    return;
  }
  ctx->pending.push_back(p);
  *slot = (int)ctx->pending.size() - 1;
}

void TimerEnd(am_ctx* ctx, int slot) {
  if (slot < 0 || slot >= (int)ctx->pending.size()) return;
  (void)hipEventRecord(ctx->pending[slot].stop, ctx->stream);  // timing is best effort
}

static int Drain(am_ctx* ctx) {
  if (ctx->pending.empty()) return AM_OK;
  AM_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->pending) {
    float ms = 0.f;
    AM_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
    auto& t = ctx->timers[p.name];
    t.first += ms;
    t.second += 1;
    ctx->event_pool.push_back(p.start);
    ctx->event_pool.push_back(p.stop);
  }
  ctx->pending.clear();
  return AM_OK;
}

// ---- source mirror ----------------------------------------------------------------------------
__global__ void RefreshSourceKernel(int64_t n, int64_t n_pad, const double* __restrict__ x,
                                    const double* __restrict__ y, const double* __restrict__ z,
                                    const double* __restrict__ q, const uint8_t* __restrict__ flags,
                                    float4* __restrict__ src) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  if (i >= n) {
    src[2 * i] = make_float4(0.f, 0.f, 0.f, 0.f);
    src[2 * i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  WriteMirror(src, i, x[i], y[i], z[i], q[i], (flags[i] & kFlagAnnihilated) != 0);
}

__global__ void ZeroNonFiniteForceKernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                         const double* __restrict__ z, const uint8_t* __restrict__ flags,
                                         double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (flags[i] & kFlagAnnihilated) || Finite3(x[i], y[i], z[i])) return;
  fx[i] = 0.0;
  fy[i] = 0.0;
  fz[i] = 0.0;
}

int ZeroNonFiniteForces(am_store* s) {
  if (s->n == 0) return AM_OK;
  hipLaunchKernelGGL(ZeroNonFiniteForceKernel, dim3(Blocks(s->n, 256)), dim3(256), 0, s->ctx->stream, s->n, s->x,
                     s->y, s->z, s->flags, s->fx, s->fy, s->fz);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

int RefreshSource(am_store* s) {
  if (s->n == 0) return AM_OK;
  const int64_t n_pad = PadN(s->n);
  hipLaunchKernelGGL(RefreshSourceKernel, dim3(Blocks(n_pad, 256)), dim3(256), 0, s->ctx->stream, s->n, n_pad, s->x,
                     s->y, s->z, s->q, s->flags, s->src);
  AM_HIP(hipGetLastError());
  return AM_OK;
}

__global__ void CountActiveKernel(int64_t n, const uint8_t* __restrict__ flags, unsigned long long* out) {
  __shared__ unsigned int part[4];
  unsigned int c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += (flags[i] & kFlagAnnihilated) ? 0u : 1u;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

__global__ void ZeroU64Kernel(unsigned long long* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

int QueueCounters(am_ctx* ctx, const char* name, unsigned** out) {
  auto it = ctx->scratch.find(name);
  const bool fresh = it == ctx->scratch.end() || it->second.ptr == nullptr;
  void* q;
  int rc = Scratch(ctx, name, sizeof(unsigned long long), &q);
  if (rc) return rc;
  if (fresh) {  // a kernel, not a memset, so no memset node enters a captured graph
    hipLaunchKernelGGL(ZeroU64Kernel, dim3(1), dim3(64), 0, ctx->stream, (unsigned long long*)q, (int64_t)1);
    AM_HIP(hipGetLastError());
  }
  *out = (unsigned*)q;
  return AM_OK;
}

int64_t CountActive(am_store* s, int* status) {
  *status = AM_OK;
  if (s->n == 0) return 0;
  void* p = nullptr;
  int rc = Scratch(s->ctx, "count_active", sizeof(unsigned long long), &p);
  if (rc) { *status = rc; return -1; }
  unsigned long long h = 0;
  hipLaunchKernelGGL(ZeroU64Kernel, dim3(1), dim3(64), 0, s->ctx->stream, (unsigned long long*)p, (int64_t)1);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    const int blocks = std::min(Blocks(s->n, 256), 1024);
    hipLaunchKernelGGL(CountActiveKernel, dim3(blocks), dim3(256), 0, s->ctx->stream, s->n, s->flags,
                       (unsigned long long*)p);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&h, p, sizeof(h), hipMemcpyDeviceToHost, s->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
  if (e != hipSuccess) { *status = HipFail(e, "CountActive", __FILE__, __LINE__); return -1; }
  return (int64_t)h;
}

__global__ void UnpackStateKernel(int64_t n, const double* __restrict__ pos, const double* __restrict__ vel,
                                  double* x, double* y, double* z, double* vx, double* vy, double* vz, double* fx,
                                  double* fy, double* fz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = pos[3 * i]; y[i] = pos[3 * i + 1]; z[i] = pos[3 * i + 2];
  vx[i] = vel[3 * i]; vy[i] = vel[3 * i + 1]; vz[i] = vel[3 * i + 2];
  fx[i] = 0.0; fy[i] = 0.0; fz[i] = 0.0;
}

__global__ void PackVec3Kernel(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                               const double* __restrict__ c, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[3 * i] = a[i]; out[3 * i + 1] = b[i]; out[3 * i + 2] = c[i];
}

__global__ void UnpackVec3Kernel(int64_t n, const double* __restrict__ in, double* a, double* b, double* c) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[i] = in[3 * i]; b[i] = in[3 * i + 1]; c[i] = in[3 * i + 2];
}

}  // namespace am

using namespace am;

extern "C" {

const char* am_last_error(void) { return g_error.c_str(); }
int am_version(void) { return 1; }

int am_trap_make_from_geometry(double b0, double trap_voltage, double z0, double rho0, am_trap_field* out) {
  if (!out) return Fail(AM_ERR_INVALID, "am_trap_make_from_geometry: out is NULL");
  const double d = std::sqrt((z0 * z0) / 2.0 + (rho0 * rho0) / 4.0);
  if (!(d > 0.0)) return Fail(AM_ERR_INVALID, "characteristic dimension must be > 0");  // penning_trap.cc:12
  out->b0 = b0;
  out->trap_voltage = trap_voltage;
  out->d = d;
  return AM_OK;
}

double am_cooling_time(double charge, double mass, double b_mag) {
  const double q2 = charge * charge;
  const double denom = q2 * q2 * b_mag * b_mag;
  if (denom <= 0.0) return INFINITY;
  return (0.75 / kCoulomb) * mass * mass * mass * kLight * kLight * kLight / denom;
}

int am_species_properties(int species, double* mass, double* charge) {
  if (!mass || !charge) return Fail(AM_ERR_INVALID, "am_species_properties: NULL output");
  SpeciesProperties(species, mass, charge);
  return AM_OK;
}

int am_is_antiparticle_pair(int a, int b) { return IsPair(a, b) ? 1 : 0; }

int am_ctx_create(int device, am_ctx** out) {
  if (!out) return Fail(AM_ERR_INVALID, "am_ctx_create: out is NULL");
  *out = nullptr;
  int count = 0;
  AM_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return Fail(AM_ERR_INVALID, "am_ctx_create: no such device");
  AM_HIP(hipSetDevice(device));
  am_ctx* c = new am_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return HipFail(e, "hipStreamCreate", __FILE__, __LINE__); }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  *out = c;
  return AM_OK;
}

int am_ctx_destroy(am_ctx* ctx) {
  if (!ctx) return AM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->scratch)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  for (auto& p : ctx->pending) { (void)hipEventDestroy(p.start); (void)hipEventDestroy(p.stop); }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return AM_OK;
}

int am_ctx_synchronize(am_ctx* ctx) {
  if (!ctx) return Fail(AM_ERR_INVALID, "NULL context");
  AM_HIP(hipStreamSynchronize(ctx->stream));
  return AM_OK;
}

int am_ctx_enable_timing(am_ctx* ctx, int enable) {
  if (!ctx) return Fail(AM_ERR_INVALID, "NULL context");
  ctx->timing = enable != 0;
  return AM_OK;
}

int am_ctx_kernel_time(am_ctx* ctx, const char* name, double* total_ms, int64_t* launches) {
  if (!ctx || !name) return Fail(AM_ERR_INVALID, "NULL argument");
  int rc = Drain(ctx);
  if (rc) return rc;
  auto it = ctx->timers.find(name);
  if (total_ms) *total_ms = it == ctx->timers.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == ctx->timers.end() ? 0 : it->second.second;
  return AM_OK;
}

int am_ctx_reset_timing(am_ctx* ctx) {
  if (!ctx) return Fail(AM_ERR_INVALID, "NULL context");
  int rc = Drain(ctx);
  ctx->timers.clear();
  return rc;
}

int am_store_create(am_ctx* ctx, int64_t capacity, am_store** out) {
  if (!ctx || !out) return Fail(AM_ERR_INVALID, "am_store_create: NULL argument");
  if (capacity <= 0 || capacity > (int64_t)1 << 30) return Fail(AM_ERR_INVALID, "am_store_create: capacity out of range");
  *out = nullptr;
  AM_HIP(hipSetDevice(ctx->device));
  am_store* s = new am_store();
  s->ctx = ctx;
  s->capacity = capacity;
  const size_t c = (size_t)capacity;
  double** dbl[] = {&s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz, &s->fx, &s->fy, &s->fz, &s->q, &s->m};
  hipError_t e = hipSuccess;
  for (double** p : dbl)
    if (e == hipSuccess) e = hipMalloc((void**)p, c * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void**)&s->species, c);
  if (e == hipSuccess) e = hipMalloc((void**)&s->flags, c);
  if (e == hipSuccess) e = hipMalloc((void**)&s->trap, c * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&s->src, 2 * (size_t)PadN(capacity) * sizeof(float4));
  if (e != hipSuccess) {
    am_store_destroy(s);
    return HipFail(e, "am_store_create hipMalloc", __FILE__, __LINE__);
  }
  *out = s;
  return AM_OK;
}

int am_store_destroy(am_store* s) {
  if (!s) return AM_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  void* ptrs[] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->fx, s->fy, s->fz, s->q, s->m,
                  s->species, s->flags, s->trap, s->src, s->perm};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete s;
  return AM_OK;
}

int am_store_upload(am_store* s, int64_t n, const double* pos, const double* vel, const uint8_t* species,
                    const double* charge, const double* mass, const uint8_t* annihilated, const uint32_t* trap) {
  if (!s) return Fail(AM_ERR_INVALID, "am_store_upload: NULL store");
  if (n < 0 || n > s->capacity) return Fail(AM_ERR_CAPACITY, "am_store_upload: n exceeds store capacity");
  if (n > 0 && (!pos || !vel || !species)) return Fail(AM_ERR_INVALID, "am_store_upload: pos/vel/species required");
  hipStream_t st = s->ctx->stream;
  AM_HIP(hipSetDevice(s->ctx->device));
  s->n = n;
  s->generation += 1;
  s->perm_n = -1;  // spatial order must be rebuilt
  if (n == 0) return AM_OK;
  std::vector<double> hq(n), hm(n);
  std::vector<uint8_t> hf(n);
  for (int64_t i = 0; i < n; ++i) {
    if (species[i] > AM_ANTIHELIUM) return Fail(AM_ERR_INVALID, "am_store_upload: unknown species");
    double mm, qq;
    SpeciesProperties(species[i], &mm, &qq);
    hq[i] = charge ? charge[i] : qq;
    hm[i] = mass ? mass[i] : mm;
    hf[i] = (annihilated && annihilated[i]) ? kFlagAnnihilated : 0;
  }
  std::vector<uint32_t> ht;
  if (!trap) ht.assign(n, 0u);
  void* staging = nullptr;
  {
    const int rc0 = Scratch(s->ctx, "upload", (size_t)n * 6 * sizeof(double), &staging);
    if (rc0) return rc0;
  }
  double* dpos = (double*)staging;
  double* dvel = dpos + 3 * n;
  AM_HIP(hipMemcpyAsync(dpos, pos, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync(dvel, vel, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync(s->q, hq.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync(s->m, hm.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync(s->species, species, n, hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync(s->flags, hf.data(), n, hipMemcpyHostToDevice, st));
  AM_HIP(hipMemcpyAsync(s->trap, trap ? trap : ht.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(UnpackStateKernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, dpos, dvel, s->x, s->y, s->z,
                     s->vx, s->vy, s->vz, s->fx, s->fy, s->fz);
  AM_HIP(hipGetLastError());
  int rc = RefreshSource(s);
  if (rc) return rc;
  AM_HIP(hipStreamSynchronize(st));  // host staging vectors go out of scope
  return AM_OK;
}

int am_store_download(am_store* s, double* pos, double* vel, double* force, uint8_t* annihilated,
                      uint8_t* staggered) {
  if (!s) return Fail(AM_ERR_INVALID, "am_store_download: NULL store");
  const int64_t n = s->n;
  if (n == 0) return AM_OK;
  hipStream_t st = s->ctx->stream;
  AM_HIP(hipSetDevice(s->ctx->device));
  void* tmp = nullptr;
  int rc = Scratch(s->ctx, "download", (size_t)n * 3 * sizeof(double), &tmp);
  if (rc) return rc;
  struct Item { double* out; double *a, *b, *c; } items[] = {{pos, s->x, s->y, s->z}, {vel, s->vx, s->vy, s->vz},
                                                            {force, s->fx, s->fy, s->fz}};
  for (auto& it : items) {
    if (!it.out) continue;
    hipLaunchKernelGGL(PackVec3Kernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, it.a, it.b, it.c, (double*)tmp);
    AM_HIP(hipGetLastError());
    AM_HIP(hipMemcpyAsync(it.out, tmp, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
  }
  if (annihilated || staggered) {
    std::vector<uint8_t> f(n);
    AM_HIP(hipMemcpyAsync(f.data(), s->flags, n, hipMemcpyDeviceToHost, st));
    AM_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i) {
      if (annihilated) annihilated[i] = (f[i] & kFlagAnnihilated) ? 1 : 0;
      if (staggered) staggered[i] = (f[i] & kFlagStaggered) ? 1 : 0;
    }
  }
  return AM_OK;
}

int am_store_set_coulomb_force(am_store* s, const double* force) {
  if (!s || (!force && s->n > 0)) return Fail(AM_ERR_INVALID, "am_store_set_coulomb_force: NULL argument");
  const int64_t n = s->n;
  if (n == 0) return AM_OK;
  hipStream_t st = s->ctx->stream;
  void* tmp = nullptr;
  int rc = Scratch(s->ctx, "upload", (size_t)n * 3 * sizeof(double), &tmp);
  if (rc) return rc;
  AM_HIP(hipMemcpyAsync(tmp, force, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(UnpackVec3Kernel, dim3(Blocks(n, 256)), dim3(256), 0, st, n, (const double*)tmp, s->fx, s->fy,
                     s->fz);
  AM_HIP(hipGetLastError());
  AM_HIP(hipStreamSynchronize(st));
  return AM_OK;
}

int am_store_download_species(am_store* s, uint8_t* species) {
  if (!s || (!species && s->n > 0)) return Fail(AM_ERR_INVALID, "am_store_download_species: NULL argument");
  if (s->n == 0) return AM_OK;
  AM_HIP(hipMemcpyAsync(species, s->species, (size_t)s->n, hipMemcpyDeviceToHost, s->ctx->stream));
  AM_HIP(hipStreamSynchronize(s->ctx->stream));
  return AM_OK;
}

int am_store_count(am_store* s, int64_t* n, int64_t* n_active) {
  if (!s) return Fail(AM_ERR_INVALID, "am_store_count: NULL store");
  if (n) *n = s->n;
  if (n_active) {
    int rc = AM_OK;
    *n_active = CountActive(s, &rc);
    if (rc) return rc;
  }
  return AM_OK;
}

}  // extern "C"
