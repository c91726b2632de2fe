// Kernels and launch helpers shared by the engines: ordered scans, set-bit
// lists, fills, wave-grouped counting and per-bucket ordering.  Internal
// linkage: every translation unit that includes this holds its own copy.
#pragma once
#include <hip/hip_runtime.h>
#include "mpcx_device.h"

namespace mpcx {
namespace {

// Adds the number of active lanes holding `key` to counter[key], once per
// distinct key of the wave: a ballot loop finds the groups (no memory access),
// then every group leader issues its atomic in one instruction.  Returns the
// lane's rank among equal keys plus the counter's old value.
__device__ __forceinline__ int wave_atomic_rank(int* counter, int key, bool active) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  unsigned long long todo = __ballot(active);
  int leader_of = 0, rank = 0, gsize = 0;
  while (todo) {
    int leader = __ffsll((unsigned long long)todo) - 1;
    int k = wave_bcast(key, leader);
    bool mine = active && key == k;
    unsigned long long grp = __ballot(mine);
    if (mine) { leader_of = leader; rank = __popcll(grp & lt); }
    if (lane == leader) gsize = __popcll(grp);
    todo &= ~grp;
  }
  int base = 0;
  if (active && lane == leader_of) base = atomicAdd(&counter[key], gsize);
  base = __shfl(base, leader_of);
  return base + rank;
}

// exclusive scan of cnt[0..n) into off[0..n] (off[n] = total), three passes
__global__ __launch_bounds__(1024) void scan_partial(const int* cnt, long long n, int* partial) {
  __shared__ unsigned long long s[17];
  long long base = (long long)blockIdx.x * 4096;
  unsigned long long v = 0;
  for (int j = 0; j < 4; j++) {
    long long i = base + threadIdx.x * 4 + j;
    if (i < n) v += cnt[i];
  }
  unsigned long long tot;
  block_excl_scan64(v, s, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = (int)tot;
}
__global__ __launch_bounds__(1024) void scan_top(int* partial, int nb) {
  __shared__ unsigned long long s[17];
  __shared__ long long run;
  if (threadIdx.x == 0) run = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    int i = b0 + threadIdx.x;
    unsigned long long v = i < nb ? (unsigned)partial[i] : 0;
    unsigned long long tot;
    unsigned long long e = block_excl_scan64(v, s, tot);
    if (i < nb) partial[i] = (int)(run + (long long)e);
    __syncthreads();
    if (threadIdx.x == 0) run += (long long)tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[nb] = (int)run;
}
__global__ __launch_bounds__(1024) void scan_final(const int* cnt, long long n, const int* partial, int* off) {
  __shared__ unsigned long long s[17];
  long long base = (long long)blockIdx.x * 4096;
  int v[4];
  unsigned long long sum = 0;
  for (int j = 0; j < 4; j++) {
    long long i = base + threadIdx.x * 4 + j;
    v[j] = i < n ? cnt[i] : 0;
    sum += v[j];
  }
  unsigned long long tot;
  unsigned long long e = block_excl_scan64(sum, s, tot);
  long long run = partial[blockIdx.x] + (long long)e;
  for (int j = 0; j < 4; j++) {
    long long i = base + threadIdx.x * 4 + j;
    if (i < n) off[i] = (int)run;
    run += v[j];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) off[n] = partial[gridDim.x];
}

// ordered list of the set bits: counts per 1024-word block, then positions
__global__ __launch_bounds__(1024) void bits_count(const unsigned* bits, long long nwords, int* partial) {
  __shared__ unsigned long long s[17];
  long long w = (long long)blockIdx.x * 1024 + threadIdx.x;
  unsigned long long v = w < nwords ? __popc(bits[w]) : 0;
  unsigned long long tot;
  block_excl_scan64(v, s, tot);
  if (threadIdx.x == 0) partial[blockIdx.x] = (int)tot;
}
__global__ __launch_bounds__(1024) void bits_emit(const unsigned* bits, long long nwords, const int* partial,
                                                  int* list, int cap, int NT, int* out_n, int* err, int nblocks) {
  __shared__ unsigned long long s[17];
  long long w = (long long)blockIdx.x * 1024 + threadIdx.x;
  unsigned word = w < nwords ? bits[w] : 0u;
  unsigned long long tot;
  unsigned long long e = block_excl_scan64(__popc(word), s, tot);
  int pos = partial[blockIdx.x] + (int)e;
  while (word) {
    int b = __ffs(word) - 1;
    word &= word - 1;
    int idx = (int)(w * 32 + b);
    if (idx < NT) {
      if (pos < cap) list[pos] = idx;
      pos++;
    }
  }
  if (blockIdx.x == nblocks - 1 && threadIdx.x == 0) {
    int n = partial[nblocks];
    *out_n = n < cap ? n : cap;
    if (n > cap) raise_error(err, ERR_NEAR_CAP);
  }
}

template <typename T>
__global__ void fill_kernel(T* p, long long n, T v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
// members of each bucket in ascending order (insertion sort; buckets hold at
// most `cap` entries, longer ones are cut and reported by the caller)
// The workgroup's buckets t0..t0+255 are one stretch of sb (soff has ncell+1
// entries): sorted in LDS when the stretch fits, in place in sb otherwise.
constexpr int SO_CAP = 8192;
__global__ __launch_bounds__(256) void sort_order(long long ncell, const int* scnt, const int* soff, int* sb, int cap) {
  __shared__ int stage[SO_CAP];
  __shared__ int srange[2];
  const long long t0 = (long long)blockIdx.x * blockDim.x, t = t0 + threadIdx.x;
  if (threadIdx.x == 0) {
    const long long te = t0 + blockDim.x < ncell ? t0 + blockDim.x : ncell;
    srange[0] = soff[t0];
    srange[1] = soff[te];
  }
  __syncthreads();
  const int s0 = srange[0], total = srange[1] - s0;
  const bool staged = total >= 0 && total <= SO_CAP;
  if (staged)
    for (int j = threadIdx.x; j < total; j += blockDim.x) stage[j] = sb[s0 + j];
  __syncthreads();
  if (t < ncell) {
    const int n = min(scnt[t], cap), o = soff[t];
    int* a = staged ? stage + (o - s0) : sb + o;
    for (int j = 1; j < n; j++) {
      int v = a[j], k = j;
      while (k > 0 && a[k - 1] > v) { a[k] = a[k - 1]; k--; }
      a[k] = v;
    }
  }
  __syncthreads();
  if (staged)
    for (int j = threadIdx.x; j < total; j += blockDim.x) sb[s0 + j] = stage[j];
}

// ------------------------------------------------------------ host helpers
static int grid_for(long long n, int threads, int cap = 4096) {
  long long g = (n + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// scan_final with the scan of the block sums folded in: block b adds the sums
// of blocks 0..b-1 itself (cheap while the blocks are few)
__global__ __launch_bounds__(1024) void scan_final_sums(const int* cnt, long long n, const int* sums, int* off) {
  __shared__ unsigned long long s[17];
  unsigned long long pre = 0;
  for (int j = threadIdx.x; j < (int)blockIdx.x; j += blockDim.x) pre += (unsigned long long)sums[j];
  unsigned long long before;
  block_excl_scan64(pre, s, before);
  long long base = (long long)blockIdx.x * 4096;
  int v[4];
  unsigned long long sum = 0;
  for (int j = 0; j < 4; j++) {
    long long i = base + threadIdx.x * 4 + j;
    v[j] = i < n ? cnt[i] : 0;
    sum += v[j];
  }
  unsigned long long tot;
  unsigned long long e = block_excl_scan64(sum, s, tot);
  long long run = (long long)before + (long long)e;
  for (int j = 0; j < 4; j++) {
    long long i = base + threadIdx.x * 4 + j;
    if (i < n) off[i] = (int)run;
    run += v[j];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) off[n] = (int)((long long)before + (long long)tot);
}

static void launch_scan(hipStream_t st, const int* cnt, long long n, int* partial, int* off) {
  int nb = (int)((n + 4095) / 4096);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(scan_partial, dim3(nb), dim3(1024), 0, st, cnt, n, partial);
  if (nb <= 4096) {   // up to 1.7e7 entries: one launch less
    hipLaunchKernelGGL(scan_final_sums, dim3(nb), dim3(1024), 0, st, cnt, n, (const int*)partial, off);
    return;
  }
  hipLaunchKernelGGL(scan_top, dim3(1), dim3(1024), 0, st, partial, nb);
  hipLaunchKernelGGL(scan_final, dim3(nb), dim3(1024), 0, st, cnt, n, (const int*)partial, off);
}


// ordered list of the set bits of a bitmap (ascending index), its length in *out_n
// bits_emit with the scan of the block counts folded in: block b sums the
// counts of blocks 0..b-1 itself (cheap while the blocks are few)
__global__ __launch_bounds__(1024) void bits_emit_scan(const unsigned* bits, long long nwords, const int* counts,
                                                       int* list, int cap, int NT, int* out_n, int* err, int nblocks) {
  __shared__ unsigned long long s[17];
  unsigned long long pre = 0;
  for (int j = threadIdx.x; j < (int)blockIdx.x; j += blockDim.x) pre += (unsigned long long)counts[j];
  unsigned long long before;
  block_excl_scan64(pre, s, before);
  long long w = (long long)blockIdx.x * 1024 + threadIdx.x;
  unsigned word = w < nwords ? bits[w] : 0u;
  unsigned long long tot;
  unsigned long long e = block_excl_scan64(__popc(word), s, tot);
  int pos = (int)(before + e);
  while (word) {
    int b = __ffs(word) - 1;
    word &= word - 1;
    int idx = (int)(w * 32 + b);
    if (idx < NT) {
      if (pos < cap) list[pos] = idx;
      pos++;
    }
  }
  if (blockIdx.x == nblocks - 1 && threadIdx.x == 0) {
    const long long n = (long long)before + (long long)tot;
    *out_n = n < cap ? (int)n : cap;
    if (n > cap) raise_error(err, ERR_NEAR_CAP);
  }
}

static inline void launch_bits_list(hipStream_t st, const unsigned* bits, long long nwords, int* partial, int* list,
                                    int cap, int NT, int* out_n, int* err) {
  int nb = (int)((nwords + 1023) / 1024);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(bits_count, dim3(nb), dim3(1024), 0, st, bits, nwords, partial);
  if (nb <= 8192) {   // up to 2.7e8 particles: one launch less
    hipLaunchKernelGGL(bits_emit_scan, dim3(nb), dim3(1024), 0, st, bits, nwords, (const int*)partial, list, cap, NT,
                       out_n, err, nb);
    return;
  }
  hipLaunchKernelGGL(scan_top, dim3(1), dim3(1024), 0, st, partial, nb);
  hipLaunchKernelGGL(bits_emit, dim3(nb), dim3(1024), 0, st, bits, nwords, (const int*)partial, list, cap, NT, out_n,
                     err, nb);
}

}  // namespace
}  // namespace mpcx
