// Error reporting and device query of the C ABI.
#include <hip/hip_runtime.h>
#include <mutex>
#include <set>
#include <string>
#include <vector>
#include <stdint.h>
#include "mpcx_device.h"
#include "mpcx_host.h"
#include "../../include/mpcx.h"

namespace mpcx {
static thread_local std::string g_last_error;
int set_error(int code, const char* msg) {
  g_last_error = msg ? msg : "";
  return code;
}
static std::mutex g_handles_mu;
static std::set<void*> g_handles[H_KINDS];
void handle_add(int kind, void* p) {
  std::lock_guard<std::mutex> lock(g_handles_mu);
  g_handles[kind].insert(p);
}
void handle_del(int kind, void* p) {
  std::lock_guard<std::mutex> lock(g_handles_mu);
  g_handles[kind].erase(p);
}
}  // namespace mpcx

// Every handle still live, destroyed in dependency order (batches before the
// engines they run), then the device drained: the streams, events, graphs and
// buffers are gone before the HIP runtime (or a profiler attached to it) tears
// down at process exit.
extern "C" int mpcx_shutdown(void) {
  std::vector<void*> live[mpcx::H_KINDS];
  {
    std::lock_guard<std::mutex> lock(mpcx::g_handles_mu);
    for (int k = 0; k < mpcx::H_KINDS; k++) live[k].assign(mpcx::g_handles[k].begin(), mpcx::g_handles[k].end());
  }
  for (void* p : live[mpcx::H_COB]) (void)mpcx_cob_destroy((mpcx_cob*)p);
  for (void* p : live[mpcx::H_CO]) (void)mpcx_co_destroy((mpcx_co*)p);
  for (void* p : live[mpcx::H_SK]) (void)mpcx_sk_destroy((mpcx_sk*)p);
  for (void* p : live[mpcx::H_MLP]) (void)mpcx_mlp_destroy((mpcx_mlp*)p);
  int n = 0;
  if (hipGetDeviceCount(&n) == hipSuccess && n > 0) (void)hipDeviceSynchronize();
  return 0;
}

extern "C" const char* mpcx_last_error(void) { return mpcx::g_last_error.c_str(); }

extern "C" int mpcx_device_synchronize(void) {
  hipError_t r = hipDeviceSynchronize();
  if (r != hipSuccess) return mpcx::set_error(MPCX_E_HIP, hipGetErrorString(r));
  return 0;
}

extern "C" int mpcx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// Plain device-to-device copy of a known size: the reference point that turns
// the hardware's fetch/write counters into bytes.
__global__ __launch_bounds__(256) void probe_copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                         long long n) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    dst[i] = src[i];
}

// the throughput mode's collision draws, philox_uniform2(seed, stream, step, idx0 + i), i < n
__global__ void philox_fill_kernel(uint64_t seed, uint32_t stream, uint64_t step, uint64_t idx0, long long n,
                                   double* out) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double a, b;
    mpcx::philox_uniform2(seed, stream, step, idx0 + (uint64_t)i, a, b);
    out[2 * i] = a;
    out[2 * i + 1] = b;
  }
}

extern "C" int mpcx_philox_uniform2(uint64_t seed, uint32_t stream, uint64_t step, uint64_t idx0, int64_t n,
                                    double* out) {
  if (n < 0 || n > (1ll << 28) || (n > 0 && !out) || stream > 0xffffu)
    return mpcx::set_error(MPCX_E_ARG, "bad arguments (n <= 2^28, stream < 65536)");
  if (n == 0) return 0;
  double* d = nullptr;
  if (hipMalloc(&d, (size_t)n * 16) != hipSuccess) return mpcx::set_error(MPCX_E_HIP, "hipMalloc failed");
  hipLaunchKernelGGL(philox_fill_kernel, dim3(1024), dim3(256), 0, 0, seed, stream, step, idx0, (long long)n, d);
  hipError_t r = hipGetLastError();
  if (r == hipSuccess) r = hipMemcpy(out, d, (size_t)n * 16, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return r == hipSuccess ? 0 : mpcx::set_error(MPCX_E_HIP, hipGetErrorString(r));
}

extern "C" int mpcx_probe_copy(int64_t nbytes, int32_t reps, double* ms_per_copy) {
  if (nbytes < 16 || reps < 1 || !ms_per_copy) return mpcx::set_error(MPCX_E_ARG, "bad arguments");
  const long long n = nbytes / 16;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    rc = mpcx::set_error(MPCX_E_HIP, "hipMalloc failed");
  } else {
    (void)hipMemset(a, 0, n * 16);
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int grid = 8192;
    hipLaunchKernelGGL(probe_copy_kernel, dim3(grid), dim3(256), 0, 0, (const double2*)a, (double2*)b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++)
      hipLaunchKernelGGL(probe_copy_kernel, dim3(grid), dim3(256), 0, 0, (const double2*)a, (double2*)b, n);
    (void)hipEventRecord(e1, 0);
    hipError_t r = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_copy = ms / reps;
    if (r != hipSuccess) rc = mpcx::set_error(MPCX_E_HIP, hipGetErrorString(r));
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

// ---------------------------------------------------------------------------
// Host Mersenne Twister streams in the two flavours the simulators use.  They
// produce the uniforms the parity mode uploads, in the reference's call order.
// ---------------------------------------------------------------------------
namespace {
void mt_twist(uint32_t* mt) {
  const uint32_t UM = 0x80000000u, LM = 0x7fffffffu, A = 0x9908b0dfu;
  int kk;
  uint32_t y;
  for (kk = 0; kk < 624 - 397; kk++) {
    y = (mt[kk] & UM) | (mt[kk + 1] & LM);
    mt[kk] = mt[kk + 397] ^ (y >> 1) ^ ((y & 1u) ? A : 0u);
  }
  for (; kk < 623; kk++) {
    y = (mt[kk] & UM) | (mt[kk + 1] & LM);
    mt[kk] = mt[kk + (397 - 624)] ^ (y >> 1) ^ ((y & 1u) ? A : 0u);
  }
  y = (mt[623] & UM) | (mt[0] & LM);
  mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? A : 0u);
}
inline uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}
}  // namespace

// sgrnd (collective_code/mtmod.f90:72-85)
extern "C" int mpcx_mtmod_seed(int32_t seed, uint32_t* mt, int32_t* mti) {
  if (!mt || !mti) return mpcx::set_error(MPCX_E_ARG, "null argument");
  mt[0] = (uint32_t)seed;
  for (int i = 1; i < 624; i++) mt[i] = 69069u * mt[i - 1];
  *mti = 624;
  return 0;
}
// n successive grnd() (mtmod.f90:88-144): u / (2^32 - 1)
extern "C" int mpcx_mtmod_fill(uint32_t* mt, int32_t* mti, double* out, int64_t n) {
  if (!mt || !mti || (!out && n > 0)) return mpcx::set_error(MPCX_E_ARG, "null argument");
  int32_t i = *mti;
  for (int64_t k = 0; k < n; k++) {
    if (i >= 624) { mt_twist(mt); i = 0; }
    out[k] = (double)mt_temper(mt[i++]) / 4294967295.0;
  }
  *mti = i;
  return 0;
}
// init_by_scalar (selkov_code/mt95.f90:378-404); cnt is 1-based as in mt95
extern "C" int mpcx_mt95_seed(int32_t seed, uint32_t* val, int32_t* cnt) {
  if (!val || !cnt) return mpcx::set_error(MPCX_E_ARG, "null argument");
  val[0] = (uint32_t)seed;
  for (int i = 1; i < 624; i++) val[i] = 1812433253u * (val[i - 1] ^ (val[i - 1] >> 30)) + (uint32_t)i;
  *cnt = 625;
  return 0;
}
// n successive genrand_real1 (mt95.f90:828-841): real(a)*(1/(2^32-1)) + 2^31/(2^32-1)
extern "C" int mpcx_mt95_fill(uint32_t* val, int32_t* cnt, double* out, int64_t n) {
  if (!val || !cnt || (!out && n > 0)) return mpcx::set_error(MPCX_E_ARG, "null argument");
  const double pi232_1 = 1.0 / 4294967295.0, p231d232_1 = 2147483648.0 / 4294967295.0;
  int32_t c = *cnt;
  for (int64_t k = 0; k < n; k++) {
    if (c > 624) { mt_twist(val); c = 1; }
    int32_t a = (int32_t)mt_temper(val[c - 1]);
    c++;
    volatile double m = (double)a * pi232_1;   // two roundings, as the reference
    out[k] = m + p231d232_1;
  }
  *cnt = c;
  return 0;
}
