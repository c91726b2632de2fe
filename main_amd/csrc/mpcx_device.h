// Device helpers shared by the collective and Selkov kernels (gfx950, wave64).
//
// All FP64 arithmetic that has a counterpart in the reference keeps the
// reference's operand order; the library is compiled with -ffp-contract=off so
// no multiply/add pair is fused behind our back.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MPCX_WAVE 64

namespace mpcx {

// ----------------------------------------------------------------- errors
enum DevError : int {
  ERR_NONE = 0,
  ERR_CELL_RANGE = 1,      // a particle maps outside the cell arrays
  ERR_MEMBERS = 2,         // more members in a cell than the reference arrays hold
  ERR_INSIDE_CAP = 3,      // inside set larger than the near engine's register slots
  ERR_NEAR_GUARD = 4,      // a sphere moved further than the near/far margin allows
  ERR_MEMBERSHIPS = 5,     // a particle in more neighbour lists than supported
  ERR_BUCKET = 6,          // more particles in one collision cell than the sort buffer holds
  ERR_NEAR_CAP = 7,        // near set larger than its buffer
  ERR_LIST_CAP = 8,        // a neighbour list longer than max_neighbour
  ERR_EVENTS = 9,          // more rebuild events in one interval than recorded
  ERR_RATTLE = 10,         // the bond constraint did not converge
  ERR_INSIDE_LOST = 11,    // a list member missing from the next interval's near set
  ERR_DRAWS = 12,          // the host draw buffer ran out before the run ended
  ERR_CLUSTER = 13,        // more than four dimers whose near regions overlap
  ERR_BARRIER = 14,        // the clusters of one system fell out of step
  ERR_CLUSTER_MEMB = 15,   // a list member carried over to a cluster without its dimer
  ERR_INIT = 16,           // no place outside the spheres found for a solvent particle
  ERR_FLIGHT = 17,         // the rebuild events of an interval do not add up to the interval
  ERR_WALL_CAP = 18,       // more far particles reach a wall in one interval than the deferred list holds
  ERR_ZONE = 19,           // a near or deferred particle ended in a cell collided before the near engine finished
  ERR_ZONE_CAP = 20,       // more late collision cells than their list holds
};

__device__ __forceinline__ void raise_error(int* err, int code) { atomicCAS(err, 0, code); }

// value of v in lane `lane` of the wave, for a wave-uniform lane: one
// v_readlane instead of a shuffle through the LDS crossbar (__shfl), which the
// grouping loops would otherwise wait on once per distinct key
__device__ __forceinline__ int wave_bcast(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// --------------------------------------------------------- geometry helpers
// "if (d > 0.5d0*L) d = d - L; else if (d < -0.5d0*L) d = d + L"
__device__ __forceinline__ double pbc_half(double d, double L) {
  if (d > 0.5 * L) return d - L;
  if (d < -0.5 * L) return d + L;
  return d;
}
__device__ __forceinline__ double pbc_h(double d, double H, double L) {
  if (d > H) return d - L;
  if (d < -H) return d + L;
  return d;
}
// floor() of the reference returns a default (32-bit) integer
__device__ __forceinline__ double ffloor_i(double v) { return (double)(int)floor(v); }

// int(floor(x) + box1*floor(y) + box1*box2*floor(z) + 1) - 1 (0-based)
__device__ __forceinline__ long long cell_unshifted0(double x, double y, double z, double b0, double b1) {
  double c = ffloor_i(x) + b0 * ffloor_i(y);
  c = c + (b0 * b1) * ffloor_i(z);
  c = c + 1.0;
  return (long long)c - 1;
}

// ----------------------------------------------------------- LJ (WCA) terms
// forces_mod.f95:82-88
__device__ __forceinline__ void lj_terms(double dRsq, double R6, double R12, double r12R12, double r6R6,
                                         double& Fb, double& Ub) {
  double IdRsq = 1.0 / dRsq;
  double IdR6 = (IdRsq * IdRsq) * IdRsq;
  double IdR8 = IdR6 * IdRsq;
  double IdR12 = IdR6 * IdR6;
  double IdR14 = IdR12 * IdRsq;
  Fb = 4.0 * (r12R12 * IdR14 - r6R6 * IdR8);
  Ub = 4.0 * (R12 * IdR12 - R6 * IdR6) + 1.0;
}

// ------------------------------------------------------- wave/block reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, MPCX_WAVE);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, MPCX_WAVE));
  return v;
}
// The same reductions on the VALU through DPP lane moves (no LDS round trip
// per step): sums over pairs, quads, half-rows and rows of 16 by quad
// permutes and row mirrors, then the four row totals, read from lanes 15, 31,
// 47 and 63, added in row order.  Each step adds a lane's value to its
// partner's and the partner does the same (commutative), so every lane of a
// row holds the same bits and the result is one wave-uniform value.  All 64
// lanes must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_mov_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double lane_read_d(double v, int lane) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int DPP_QUAD_1032 = 0xb1, DPP_QUAD_2301 = 0x4e, DPP_ROW_MIRROR = 0x140, DPP_ROW_HALF_MIRROR = 0x141;
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v = v + dpp_mov_d<DPP_QUAD_1032>(v);
  v = v + dpp_mov_d<DPP_QUAD_2301>(v);
  v = v + dpp_mov_d<DPP_ROW_HALF_MIRROR>(v);
  v = v + dpp_mov_d<DPP_ROW_MIRROR>(v);
  return ((lane_read_d(v, 15) + lane_read_d(v, 31)) + lane_read_d(v, 47)) + lane_read_d(v, 63);
}
__device__ __forceinline__ double wave_max_dpp(double v) {
  v = fmax(v, dpp_mov_d<DPP_QUAD_1032>(v));
  v = fmax(v, dpp_mov_d<DPP_QUAD_2301>(v));
  v = fmax(v, dpp_mov_d<DPP_ROW_HALF_MIRROR>(v));
  v = fmax(v, dpp_mov_d<DPP_ROW_MIRROR>(v));
  return fmax(fmax(fmax(lane_read_d(v, 15), lane_read_d(v, 31)), lane_read_d(v, 47)), lane_read_d(v, 63));
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, MPCX_WAVE);
  return v;
}
// inclusive prefix sum over the wave, in lane order
__device__ __forceinline__ int wave_incl_scan(int v) {
  int lane = threadIdx.x & (MPCX_WAVE - 1);
#pragma unroll
  for (int o = 1; o < MPCX_WAVE; o <<= 1) {
    int t = __shfl_up(v, o, MPCX_WAVE);
    if (lane >= o) v += t;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_incl_scan64(unsigned long long v) {
  int lane = threadIdx.x & (MPCX_WAVE - 1);
#pragma unroll
  for (int o = 1; o < MPCX_WAVE; o <<= 1) {
    unsigned long long t = __shfl_up(v, o, MPCX_WAVE);
    if (lane >= o) v += t;
  }
  return v;
}
// rank of this lane among lanes (of lower id) whose predicate holds
__device__ __forceinline__ int wave_rank(bool pred, unsigned long long& mask) {
  mask = __ballot(pred);
  int lane = threadIdx.x & (MPCX_WAVE - 1);
  unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  return __popcll(mask & lt);
}

// Block exclusive scan of a 64-bit value over blockDim.x threads (fixed order,
// so results are deterministic).  scratch: >= blockDim.x/64 + 1 entries.
__device__ __forceinline__ unsigned long long block_excl_scan64(unsigned long long v,
                                                                unsigned long long* scratch,
                                                                unsigned long long& total) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long inc = wave_incl_scan64(v);
  if (lane == 63) scratch[w] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long run = 0;
    for (int i = 0; i < nw; i++) { unsigned long long t = scratch[i]; scratch[i] = run; run += t; }
    scratch[nw] = run;
  }
  __syncthreads();
  unsigned long long r = scratch[w] + inc - v;
  total = scratch[nw];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based generator
// used in throughput mode for the collision draws.
struct u32x4 { uint32_t v[4]; };
__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t hi0 = __umulhi(M0, c.v[0]), lo0 = M0 * c.v[0];
    uint32_t hi1 = __umulhi(M1, c.v[2]), lo1 = M1 * c.v[2];
    u32x4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}
// two doubles in [0,1) with 53 random bits each
__device__ __forceinline__ void philox_uniform2(uint64_t seed, uint32_t stream, uint64_t step, uint64_t idx,
                                                double& a, double& b) {
  u32x4 c;
  c.v[0] = (uint32_t)idx; c.v[1] = (uint32_t)(idx >> 32); c.v[2] = (uint32_t)step;
  c.v[3] = ((uint32_t)(step >> 32) & 0xffffu) | (stream << 16);
  u32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double k = 1.0 / 9007199254740992.0;  // 2^-53
  a = (double)(((uint64_t)r.v[0] << 21) ^ (r.v[1] >> 11)) * k;
  b = (double)(((uint64_t)r.v[2] << 21) ^ (r.v[3] >> 11)) * k;
}

}  // namespace mpcx
