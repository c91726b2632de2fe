// MLP surrogate of the MPC runs: FP32 master weights, BF16 operands on the
// MFMA units (v_mfma_f32_16x16x32_bf16, FP32 accumulation), hand-written
// forward and backward passes and Adam, all on the device.  It maps the
// parameters of a run (radii, bond length, interaction energies, reaction
// probability, rate constants) to the time-averaged body-frame density field
// and the propulsion velocity that full MPC runs produce.  No reference
// counterpart: the reference has no learned part.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "mpcx_device.h"
#include "mpcx_host.h"
#include "../../include/mpcx.h"

namespace mpcx {
namespace ml {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 256;                 // GEMM block: 4 waves
constexpr int BM = 64, BN = 64, BK = 32;

__device__ __forceinline__ __bf16 to_bf16(float f) { return (__bf16)f; }   // round to nearest even

// C[m][n] (ldc) = act(sum_k A[m*am + k*ak] * B[k*bk + n*bn] + bias[n]) (+ C if accumulate);
// A, B in BF16, accumulation in FP32.  Each wave owns a 32x32 quarter of the
// 64x64 block tile: 2x2 MFMA 16x16x32 tiles.  Z (optional) keeps the
// pre-activation for the backward pass.
__global__ __launch_bounds__(GT) void gemm_bf16(int M, int N, int K, const __bf16* __restrict__ A, long long am,
                                                long long ak, const __bf16* __restrict__ B, long long bk, long long bn,
                                                float* __restrict__ C, long long ldc, const float* __restrict__ bias,
                                                int relu, float* __restrict__ Z, __bf16* __restrict__ Cb,
                                                int accumulate) {
  __shared__ __bf16 As[BM][BK + 8];   // +8: 16-byte row padding against LDS bank conflicts
  __shared__ __bf16 Bs[BN][BK + 8];   // B tile stored transposed: Bs[n][k]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += BK) {
    // stage the tiles (zero beyond the matrix edges): 64x32 elements each, 8 per thread
    for (int e = tid; e < BM * BK; e += GT) {
      const int r = e / BK, c = e % BK;
      const int gm = m0 + r, gk = k0 + c;
      As[r][c] = (gm < M && gk < K) ? A[gm * am + gk * ak] : to_bf16(0.f);
      const int gn = n0 + r;
      Bs[r][c] = (gn < N && gk < K) ? B[gk * bk + gn * bn] : to_bf16(0.f);
    }
    __syncthreads();
    // fragments: lane l holds row/column l%16 and k = 8*(l/16) .. +7
    const int rr = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
#pragma unroll
      for (int t = 0; t < 8; t++) {
        a[i][t] = As[wm + i * 16 + rr][kq + t];
        b[i][t] = Bs[wn + i * 16 + rr][kq + t];
      }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // result of tile (i,j): lane l holds column l%16, rows 4*(l/16) .. +3
  const int cc = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int gm = m0 + wm + i * 16 + rq + t, gn = n0 + wn + j * 16 + cc;
        if (gm >= M || gn >= N) continue;
        float v = acc[i][j][t];
        if (bias) v = v + bias[gn];
        if (Z) Z[gm * ldc + gn] = v;
        if (relu) v = v > 0.f ? v : 0.f;
        if (accumulate) v = v + C[gm * ldc + gn];
        C[gm * ldc + gn] = v;
        if (Cb) Cb[gm * ldc + gn] = to_bf16(v);
      }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = to_bf16(x[i]);
}

// rows of x gathered by idx (a minibatch), in FP32 and BF16
__global__ void gather_rows(const float* __restrict__ src, const int* __restrict__ idx, int rows, int cols,
                            float* __restrict__ dst, __bf16* __restrict__ dstb) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)rows * cols;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e / cols), c = (int)(e % cols);
    const float v = src[(long long)idx[r] * cols + c];
    dst[e] = v;
    if (dstb) dstb[e] = to_bf16(v);
  }
}

// weighted squared error: loss partials per block, dY = 2 w (Y - T) / batch
__global__ __launch_bounds__(256) void mse_grad(const float* __restrict__ Y, const float* __restrict__ T,
                                                const float* __restrict__ wout, int rows, int cols,
                                                float* __restrict__ dY, __bf16* __restrict__ dYb,
                                                double* __restrict__ part) {
  double acc = 0.0;
  for (long long e = blockIdx.x * 256ll + threadIdx.x; e < (long long)rows * cols; e += (long long)gridDim.x * 256) {
    const int c = (int)(e % cols);
    const float d = Y[e] - T[e];
    const float wv = wout[c];
    acc += (double)wv * (double)d * (double)d;
    const float g = 2.f * wv * d / (float)rows;
    dY[e] = g;
    dYb[e] = to_bf16(g);
  }
  __shared__ double red[4];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// dZ = dH * (Z > 0), in FP32 and BF16
__global__ void relu_back(const float* __restrict__ dH, const float* __restrict__ Z, long long n,
                          float* __restrict__ dZ, __bf16* __restrict__ dZb) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float g = Z[i] > 0.f ? dH[i] : 0.f;
    dZ[i] = g;
    dZb[i] = to_bf16(g);
  }
}

// db[n] = sum over rows of dZ[r][n] (rows in order)
__global__ void col_sum(const float* __restrict__ dZ, int rows, int cols, float* __restrict__ db) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; r++) s += dZ[(long long)r * cols + c];
  db[c] = s;
}

// Adam on the FP32 master parameters
__global__ void adam_step(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                          float c2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

}  // namespace ml
}  // namespace mpcx

using namespace mpcx;
using namespace mpcx::ml;

struct mpcx_mlp {
  int n_in, n_hid, n_layers, n_out;   // n_layers hidden layers
  std::vector<int> dims;              // dims[l] -> dims[l+1]
  std::vector<long long> woff, boff;  // offsets of W_l and b_l in the flat parameter vector
  long long nparam = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  __bf16* pb = nullptr;               // BF16 copy of the parameters (GEMM operands)
  int cap = 0;                        // batch capacity of the activation buffers
  std::vector<float*> H, Z, dH;       // activations per layer (H[0] = input)
  std::vector<__bf16*> Hb, dZb;
  float* dZ = nullptr;
  double* part = nullptr;
  int* idx = nullptr;
  float *X = nullptr, *T = nullptr, *W = nullptr;   // dataset on the device
  long long ndata = 0;
  hipStream_t st = nullptr;
  long long t = 0;                    // Adam step counter
  double gemm_ms = 0;
  long long gemm_flops = 0;
};

#define MLCHK(x)                                                              \
  do {                                                                        \
    hipError_t _r = (x);                                                      \
    if (_r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(_r)); \
  } while (0)

static void gemm(hipStream_t st, int M, int N, int K, const __bf16* A, long long am, long long ak, const __bf16* B,
                 long long bk, long long bn, float* C, long long ldc, const float* bias, int relu, float* Z,
                 __bf16* Cb, int acc) {
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  hipLaunchKernelGGL(gemm_bf16, grid, dim3(GT), 0, st, M, N, K, A, am, ak, B, bk, bn, C, ldc, bias, relu, Z, Cb, acc);
}

static int mlp_alloc_batch(mpcx_mlp* e, int batch) {
  if (batch <= e->cap) return 0;
  auto fr = [](void* q) { if (q) (void)hipFree(q); };
  for (auto q : e->H) fr(q);
  for (auto q : e->Z) fr(q);
  for (auto q : e->dH) fr(q);
  for (auto q : e->Hb) fr(q);
  for (auto q : e->dZb) fr(q);
  fr(e->dZ); fr(e->part); fr(e->idx);
  const int L = (int)e->dims.size();
  e->H.assign(L, nullptr); e->Z.assign(L, nullptr); e->dH.assign(L, nullptr);
  e->Hb.assign(L, nullptr); e->dZb.assign(L, nullptr);
  int maxd = 0;
  for (int l = 0; l < L; l++) {
    const size_t n = (size_t)batch * e->dims[l];
    maxd = std::max(maxd, e->dims[l]);
    MLCHK(hipMalloc(&e->H[l], n * 4)); MLCHK(hipMalloc(&e->Z[l], n * 4)); MLCHK(hipMalloc(&e->dH[l], n * 4));
    MLCHK(hipMalloc(&e->Hb[l], n * 2)); MLCHK(hipMalloc(&e->dZb[l], n * 2));
  }
  MLCHK(hipMalloc(&e->dZ, (size_t)batch * maxd * 4));
  MLCHK(hipMalloc(&e->part, 1024 * 8));
  MLCHK(hipMalloc(&e->idx, (size_t)batch * 4));
  e->cap = batch;
  return 0;
}

// forward pass of `rows` rows already in H[0]/Hb[0]; output in H[L-1]
static void mlp_forward(mpcx_mlp* e, int rows) {
  const int L = (int)e->dims.size();
  for (int l = 0; l + 1 < L; l++) {
    const int K = e->dims[l], N = e->dims[l + 1];
    const bool last = (l + 2 == L);
    // Z = H W^T + b: W_l is (N x K) row-major, so B[k][n] = W[n*K + k]
    gemm(e->st, rows, N, K, e->Hb[l], K, 1, e->pb + e->woff[l], 1, K, e->H[l + 1], N, e->p + e->boff[l],
         last ? 0 : 1, e->Z[l + 1], e->Hb[l + 1], 0);
    e->gemm_flops += 2ll * rows * N * K;
  }
}

// backward pass from dY (dZ of the output layer, in dZb[L-1] and dZ); gradients into g
static void mlp_backward(mpcx_mlp* e, int rows) {
  const int L = (int)e->dims.size();
  for (int l = L - 2; l >= 0; l--) {
    const int K = e->dims[l], N = e->dims[l + 1];
    // dW_l = dZ^T H_l : (N x rows)(rows x K)
    gemm(e->st, N, K, rows, e->dZb[l + 1], 1, N, e->Hb[l], K, 1, e->g + e->woff[l], K, nullptr, 0, nullptr, nullptr, 0);
    hipLaunchKernelGGL(col_sum, dim3((N + 255) / 256), dim3(256), 0, e->st, (const float*)e->dH[l + 1], rows, N,
                       e->g + e->boff[l]);
    e->gemm_flops += 2ll * rows * N * K;
    if (l == 0) break;
    // dH_l = dZ W_l : (rows x N)(N x K), then dZ_l = dH_l * relu'(Z_l)
    gemm(e->st, rows, K, N, e->dZb[l + 1], N, 1, e->pb + e->woff[l], K, 1, e->dZ, K, nullptr, 0, nullptr, nullptr, 0);
    e->gemm_flops += 2ll * rows * N * K;
    const long long n = (long long)rows * K;
    hipLaunchKernelGGL(relu_back, dim3((unsigned)std::min<long long>((n + 255) / 256, 4096)), dim3(256), 0, e->st,
                       (const float*)e->dZ, (const float*)e->Z[l], n, e->dH[l], e->dZb[l]);
  }
}

static void refresh_bf16(mpcx_mlp* e) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)std::min<long long>((e->nparam + 255) / 256, 4096)), dim3(256), 0,
                     e->st, (const float*)e->p, e->pb, e->nparam);
}

extern "C" int mpcx_mlp_destroy(mpcx_mlp* e) {
  if (!e) return 0;
  mpcx::handle_del(mpcx::H_MLP, e);
  if (e->st) (void)hipStreamSynchronize(e->st);
  auto fr = [](void* q) { if (q) (void)hipFree(q); };
  fr(e->p); fr(e->g); fr(e->m); fr(e->v); fr(e->pb);
  for (auto q : e->H) fr(q);
  for (auto q : e->Z) fr(q);
  for (auto q : e->dH) fr(q);
  for (auto q : e->Hb) fr(q);
  for (auto q : e->dZb) fr(q);
  fr(e->dZ); fr(e->part); fr(e->idx); fr(e->X); fr(e->T); fr(e->W);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return 0;
}

extern "C" int mpcx_mlp_create(mpcx_mlp** out, int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out,
                               uint64_t seed) {
  if (!out || n_in < 1 || n_hidden < 1 || n_layers < 1 || n_out < 1) return set_error(MPCX_E_ARG, "bad sizes");
  mpcx_mlp* e = new mpcx_mlp();
  e->n_in = n_in; e->n_hid = n_hidden; e->n_layers = n_layers; e->n_out = n_out;
  e->dims.push_back(n_in);
  for (int l = 0; l < n_layers; l++) e->dims.push_back(n_hidden);
  e->dims.push_back(n_out);
  long long o = 0;
  for (size_t l = 0; l + 1 < e->dims.size(); l++) {
    e->woff.push_back(o); o += (long long)e->dims[l] * e->dims[l + 1];
    e->boff.push_back(o); o += e->dims[l + 1];
  }
  e->nparam = o;
  // He-uniform initial weights from a 64-bit LCG (Knuth MMIX constants), zero biases
  std::vector<float> p(o, 0.f);
  unsigned long long s = seed * 6364136223846793005ull + 1442695040888963407ull;
  for (size_t l = 0; l + 1 < e->dims.size(); l++) {
    const double lim = sqrt(6.0 / e->dims[l]);
    for (long long i = 0; i < (long long)e->dims[l] * e->dims[l + 1]; i++) {
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      const double u = (double)(s >> 11) * (1.0 / 9007199254740992.0);
      p[e->woff[l] + i] = (float)((2.0 * u - 1.0) * lim);
    }
  }
  if (hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&e->p, o * 4) != hipSuccess || hipMalloc(&e->g, o * 4) != hipSuccess ||
      hipMalloc(&e->m, o * 4) != hipSuccess || hipMalloc(&e->v, o * 4) != hipSuccess ||
      hipMalloc(&e->pb, o * 2) != hipSuccess) {
    mpcx_mlp_destroy(e);
    return set_error(MPCX_E_HIP, "device allocation failed");
  }
  MLCHK(hipMemcpy(e->p, p.data(), o * 4, hipMemcpyHostToDevice));
  MLCHK(hipMemset(e->m, 0, o * 4));
  MLCHK(hipMemset(e->v, 0, o * 4));
  MLCHK(hipMemset(e->g, 0, o * 4));
  refresh_bf16(e);
  MLCHK(hipStreamSynchronize(e->st));
  *out = e;
  mpcx::handle_add(mpcx::H_MLP, e);
  return 0;
}

extern "C" int64_t mpcx_mlp_num_params(mpcx_mlp* e) { return e ? e->nparam : 0; }

extern "C" int mpcx_mlp_get_params(mpcx_mlp* e, float* flat) {
  if (!e || !flat) return set_error(MPCX_E_ARG, "null argument");
  MLCHK(hipStreamSynchronize(e->st));
  MLCHK(hipMemcpy(flat, e->p, e->nparam * 4, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int mpcx_mlp_set_params(mpcx_mlp* e, const float* flat) {
  if (!e || !flat) return set_error(MPCX_E_ARG, "null argument");
  MLCHK(hipStreamSynchronize(e->st));
  MLCHK(hipMemcpy(e->p, flat, e->nparam * 4, hipMemcpyHostToDevice));
  refresh_bf16(e);
  MLCHK(hipStreamSynchronize(e->st));
  return 0;
}

// rows -> H[0] (FP32 and BF16) from a host array
static int load_rows(mpcx_mlp* e, int rows, const float* x) {
  const size_t n = (size_t)rows * e->n_in;
  MLCHK(hipMemcpyAsync(e->H[0], x, n * 4, hipMemcpyHostToDevice, e->st));
  hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, e->st,
                     (const float*)e->H[0], e->Hb[0], (long long)n);
  return 0;
}

extern "C" int mpcx_mlp_forward(mpcx_mlp* e, int32_t rows, const float* x, float* y) {
  if (!e || rows < 1 || !x || !y) return set_error(MPCX_E_ARG, "bad arguments");
  int rc = mlp_alloc_batch(e, rows);
  if (rc) return rc;
  if ((rc = load_rows(e, rows, x))) return rc;
  mlp_forward(e, rows);
  const int L = (int)e->dims.size();
  MLCHK(hipMemcpyAsync(y, e->H[L - 1], (size_t)rows * e->n_out * 4, hipMemcpyDeviceToHost, e->st));
  MLCHK(hipStreamSynchronize(e->st));
  return 0;
}

// loss and gradient of one batch held in H[0] with targets T (device)
static int loss_grad(mpcx_mlp* e, int rows, const float* Td, const float* wd, double* loss) {
  const int L = (int)e->dims.size();
  mlp_forward(e, rows);
  const long long n = (long long)rows * e->n_out;
  const int nb = (int)std::min<long long>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(mse_grad, dim3(nb), dim3(256), 0, e->st, (const float*)e->H[L - 1], Td, wd, rows, e->n_out,
                     e->dH[L - 1], e->dZb[L - 1], e->part);
  mlp_backward(e, rows);
  std::vector<double> part(nb);
  MLCHK(hipMemcpyAsync(part.data(), e->part, nb * 8, hipMemcpyDeviceToHost, e->st));
  MLCHK(hipStreamSynchronize(e->st));
  double s = 0.0;
  for (int b = 0; b < nb; b++) s += part[b];
  *loss = s / rows;
  return 0;
}

extern "C" int mpcx_mlp_loss_grad(mpcx_mlp* e, int32_t rows, const float* x, const float* target, const float* wout,
                                  double* loss, float* grad) {
  if (!e || rows < 1 || !x || !target || !wout || !loss) return set_error(MPCX_E_ARG, "bad arguments");
  int rc = mlp_alloc_batch(e, rows);
  if (rc) return rc;
  float *Td = nullptr, *wd = nullptr;
  MLCHK(hipMalloc(&Td, (size_t)rows * e->n_out * 4));
  MLCHK(hipMalloc(&wd, (size_t)e->n_out * 4));
  MLCHK(hipMemcpyAsync(Td, target, (size_t)rows * e->n_out * 4, hipMemcpyHostToDevice, e->st));
  MLCHK(hipMemcpyAsync(wd, wout, (size_t)e->n_out * 4, hipMemcpyHostToDevice, e->st));
  if ((rc = load_rows(e, rows, x)) == 0) rc = loss_grad(e, rows, Td, wd, loss);
  if (rc == 0 && grad) {
    if (hipMemcpy(grad, e->g, e->nparam * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = set_error(MPCX_E_HIP, "hipMemcpy failed");
  }
  (void)hipFree(Td);
  (void)hipFree(wd);
  return rc;
}

// Minibatch Adam over a dataset held on the device.  Each epoch visits the n
// rows in an order shuffled by a 64-bit LCG seeded with `seed` and the epoch.
extern "C" int mpcx_mlp_train(mpcx_mlp* e, int64_t n, const float* x, const float* target, const float* wout,
                              int32_t batch, int32_t epochs, float lr, uint64_t seed, double* epoch_loss) {
  if (!e || n < 1 || !x || !target || !wout || batch < 1 || epochs < 0) return set_error(MPCX_E_ARG, "bad arguments");
  int rc = mlp_alloc_batch(e, batch);
  if (rc) return rc;
  auto fr = [](void* q) { if (q) (void)hipFree(q); };
  fr(e->X); fr(e->T); fr(e->W);
  e->X = e->T = e->W = nullptr;
  MLCHK(hipMalloc(&e->X, (size_t)n * e->n_in * 4));
  MLCHK(hipMalloc(&e->T, (size_t)n * e->n_out * 4));
  MLCHK(hipMalloc(&e->W, (size_t)e->n_out * 4));
  MLCHK(hipMemcpy(e->X, x, (size_t)n * e->n_in * 4, hipMemcpyHostToDevice));
  MLCHK(hipMemcpy(e->T, target, (size_t)n * e->n_out * 4, hipMemcpyHostToDevice));
  MLCHK(hipMemcpy(e->W, wout, (size_t)e->n_out * 4, hipMemcpyHostToDevice));
  e->ndata = n;
  float* Tb = nullptr;
  MLCHK(hipMalloc(&Tb, (size_t)batch * e->n_out * 4));
  std::vector<int> order(n);
  hipEvent_t a0, a1;
  (void)hipEventCreate(&a0);
  (void)hipEventCreate(&a1);
  (void)hipEventRecord(a0, e->st);
  e->gemm_flops = 0;
  for (int ep = 0; ep < epochs; ep++) {
    for (long long i = 0; i < n; i++) order[i] = (int)i;
    unsigned long long s = (seed + 1) * 6364136223846793005ull + (unsigned long long)ep * 1442695040888963407ull;
    for (long long i = n - 1; i > 0; i--) {   // Fisher-Yates
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      const long long j = (long long)((s >> 33) % (unsigned long long)(i + 1));
      std::swap(order[i], order[j]);
    }
    double tot = 0.0;
    long long seen = 0;
    for (long long b0 = 0; b0 < n; b0 += batch) {
      const int rows = (int)std::min<long long>(batch, n - b0);
      MLCHK(hipMemcpyAsync(e->idx, order.data() + b0, rows * 4, hipMemcpyHostToDevice, e->st));
      hipLaunchKernelGGL(gather_rows, dim3(64), dim3(256), 0, e->st, (const float*)e->X, (const int*)e->idx, rows,
                         e->n_in, e->H[0], e->Hb[0]);
      hipLaunchKernelGGL(gather_rows, dim3(256), dim3(256), 0, e->st, (const float*)e->T, (const int*)e->idx, rows,
                         e->n_out, Tb, (__bf16*)nullptr);
      double l = 0.0;
      if ((rc = loss_grad(e, rows, Tb, e->W, &l))) break;
      tot += l * rows;
      seen += rows;
      e->t += 1;
      const float c1 = 1.f - powf(0.9f, (float)e->t), c2 = 1.f - powf(0.999f, (float)e->t);
      hipLaunchKernelGGL(adam_step, dim3((unsigned)std::min<long long>((e->nparam + 255) / 256, 4096)), dim3(256), 0,
                         e->st, e->p, (const float*)e->g, e->m, e->v, e->nparam, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
      refresh_bf16(e);
    }
    if (rc) break;
    if (epoch_loss) epoch_loss[ep] = tot / (double)seen;
  }
  (void)hipEventRecord(a1, e->st);
  (void)hipEventSynchronize(a1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, a0, a1);
  e->gemm_ms = ms;
  (void)hipEventDestroy(a0);
  (void)hipEventDestroy(a1);
  (void)hipFree(Tb);
  if (rc) return rc;
  MLCHK(hipStreamSynchronize(e->st));
  return 0;
}

// wall time (ms) and GEMM flops of the last mpcx_mlp_train
extern "C" int mpcx_mlp_train_stats(mpcx_mlp* e, double* ms, double* gemm_flops) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  if (ms) *ms = e->gemm_ms;
  if (gemm_flops) *gemm_flops = (double)e->gemm_flops;
  return 0;
}

// C (M x N, FP32) = A (M x K) B (K x N) with A, B rounded to BF16 and FP32
// accumulation on the MFMA units; row-major host arrays.  For tests and timing.
extern "C" int mpcx_gemm_bf16(int32_t M, int32_t N, int32_t K, const float* A, const float* B, float* C, int32_t reps,
                              double* ms) {
  if (M < 1 || N < 1 || K < 1 || !A || !B || !C || reps < 1) return set_error(MPCX_E_ARG, "bad arguments");
  float *Af = nullptr, *Bf = nullptr, *Cd = nullptr;
  __bf16 *Ab = nullptr, *Bb = nullptr;
  MLCHK(hipMalloc(&Af, (size_t)M * K * 4)); MLCHK(hipMalloc(&Bf, (size_t)K * N * 4));
  MLCHK(hipMalloc(&Ab, (size_t)M * K * 2)); MLCHK(hipMalloc(&Bb, (size_t)K * N * 2));
  MLCHK(hipMalloc(&Cd, (size_t)M * N * 4));
  MLCHK(hipMemcpy(Af, A, (size_t)M * K * 4, hipMemcpyHostToDevice));
  MLCHK(hipMemcpy(Bf, B, (size_t)K * N * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(to_bf16_kernel, dim3(1024), dim3(256), 0, 0, (const float*)Af, Ab, (long long)M * K);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(1024), dim3(256), 0, 0, (const float*)Bf, Bb, (long long)K * N);
  hipEvent_t a0, a1;
  (void)hipEventCreate(&a0);
  (void)hipEventCreate(&a1);
  gemm(0, M, N, K, Ab, K, 1, Bb, N, 1, Cd, N, nullptr, 0, nullptr, nullptr, 0);
  (void)hipEventRecord(a0, 0);
  for (int r = 0; r < reps; r++) gemm(0, M, N, K, Ab, K, 1, Bb, N, 1, Cd, N, nullptr, 0, nullptr, nullptr, 0);
  (void)hipEventRecord(a1, 0);
  hipError_t rr = hipEventSynchronize(a1);
  float t = 0;
  (void)hipEventElapsedTime(&t, a0, a1);
  if (ms) *ms = t / reps;
  if (rr == hipSuccess) rr = hipMemcpy(C, Cd, (size_t)M * N * 4, hipMemcpyDeviceToHost);
  (void)hipEventDestroy(a0);
  (void)hipEventDestroy(a1);
  (void)hipFree(Af); (void)hipFree(Bf); (void)hipFree(Ab); (void)hipFree(Bb); (void)hipFree(Cd);
  if (rr != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(rr));
  return 0;
}
