// Reference-shaped entry points: one call = one Fortran module procedure of
// collective_code, with its host arrays and argument list.  Each call stages
// its arrays through the device, runs exact kernels and copies the outputs
// back (the device-resident engine avoids these copies for the time loop).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>
#include "mpcx_device.h"
#include "mpcx_host.h"
#include "../../include/mpcx.h"

#pragma clang fp contract(off)

namespace mpcx {
namespace dropin {

constexpr int BT = 256;
constexpr int CB = 1024;   // elements per block of the compaction scans

// -------------------------------------------------------------- device buffers
struct Buf {
  void* p = nullptr;
  size_t n = 0;
  ~Buf() { if (p) (void)hipFree(p); }
  bool alloc(size_t bytes) {
    n = bytes;
    return hipMalloc(&p, bytes ? bytes : 8) == hipSuccess;
  }
  template <typename T> T* as() { return (T*)p; }
};

#define DCHK(x)                                                              \
  do {                                                                       \
    hipError_t _r = (x);                                                     \
    if (_r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(_r)); \
  } while (0)

static int up(Buf& b, const void* h, size_t bytes) {
  if (!b.alloc(bytes)) return set_error(MPCX_E_HIP, "hipMalloc failed");
  if (h && bytes) DCHK(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
  return 0;
}
static int down(void* h, Buf& b, size_t bytes) {
  if (h && bytes) DCHK(hipMemcpy(h, b.p, bytes, hipMemcpyDeviceToHost));
  return 0;
}

// exclusive scan of counts (int) into offsets (int, n+1 entries) on the host
// side of small arrays is enough here: counts come back, offsets go up.
static void host_scan(std::vector<int>& cnt, std::vector<int>& off) {
  off.assign(cnt.size() + 1, 0);
  for (size_t i = 0; i < cnt.size(); i++) off[i + 1] = off[i] + cnt[i];
}

// ------------------------------------------------------------ mpc_rotation
// mpc_rotation_mod.f95:116-145
__global__ void k_cells(int NT, double b0, double b1, double xs, double ys, double zs, const double* x,
                        const double* y, const double* z, long long numcell, long long* ncn, int* cnt, int* rank,
                        int* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NT) return;
  double X = x[i] + xs;
  X = X - b0 * floor(X / b0);
  double Y = y[i] + ys;
  Y = Y - b1 * floor(Y / b1);
  double Z = z[i] + zs;
  double fz = zs < 0.0 ? floor(Z + 1.0) : floor(Z);
  double c = ffloor_i(X) + b0 * ffloor_i(Y);
  c = c + (b0 * b1) * (double)(int)fz;
  c = c + 1.0;
  long long cn = (long long)c;
  ncn[i] = cn;
  if (cn < 1 || cn > numcell) { raise_error(err, ERR_CELL_RANGE); rank[i] = -1; return; }
  rank[i] = atomicAdd(&cnt[cn - 1], 1);
}
__global__ void k_scatter(int NT, const long long* ncn, const int* rank, const int* off, int* bucket) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NT || rank[i] < 0) return;
  bucket[off[ncn[i] - 1] + rank[i]] = i;
}
// mpc_rotation_mod.f95:153-193: sum of each cell's velocities in ascending
// particle order, division by the count, axis from two draws
__global__ void k_cell_cm(long long numcell, const int* cnt, const int* off, int* bucket, const double* vx,
                          const double* vy, const double* vz, const double* draws, double twopi, double* vxcm,
                          double* vycm, double* vzcm, double* tcs, double* uvx, double* uvy, double* uvz) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= numcell) return;
  int n = cnt[c];
  int* b = bucket + off[c];
  for (int j = 1; j < n; j++) {
    int v = b[j], k = j;
    while (k > 0 && b[k - 1] > v) { b[k] = b[k - 1]; k--; }
    b[k] = v;
  }
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int j = 0; j < n; j++) { sx = sx + vx[b[j]]; sy = sy + vy[b[j]]; sz = sz + vz[b[j]]; }
  double t = (double)n;
  if (n != 0) { sx = sx / t; sy = sy / t; sz = sz / t; }
  vxcm[c] = sx; vycm[c] = sy; vzcm[c] = sz; tcs[c] = t;
  double phi = twopi * draws[3 + 2 * c];
  double th = 2.0 * draws[4 + 2 * c] - 1.0;
  double st = sqrt(1.0 - th * th);
  double sp, cp;
  sincos(phi, &sp, &cp);   // one reduction; the same values as sin() and cos()
  uvx[c] = st * cp; uvy[c] = st * sp; uvz[c] = th;
}
// mpc_rotation_mod.f95:205-230
__global__ void k_rotate(int NT, const long long* ncn, const double* vxcm, const double* vycm, const double* vzcm,
                         const double* uvx, const double* uvy, const double* uvz, double cr, double sr,
                         double* vx, double* vy, double* vz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NT) return;
  long long c = ncn[i] - 1;
  double ux = uvx[c], uy = uvy[c], uz = uvz[c], cx = vxcm[c], cy = vycm[c], cz = vzcm[c];
  double omc = 1.0 - cr;
  double dx = vx[i] - cx, dy = vy[i] - cy, dz = vz[i] - cz;
  double rx = (dx * (cr + (ux * ux) * omc) + dy * ((ux * uy) * omc - uz * sr)) + dz * ((ux * uz) * omc + uy * sr);
  double ry = (dx * ((ux * uy) * omc + uz * sr) + dy * (cr + (uy * uy) * omc)) + dz * ((uy * uz) * omc - ux * sr);
  double rz = (dx * ((ux * uz) * omc - uy * sr) + dy * ((uy * uz) * omc + ux * sr)) + dz * (cr + (uz * uz) * omc);
  vx[i] = cx + rx; vy[i] = cy + ry; vz[i] = cz + rz;
}

// reaction_mod.f95:55-60: cell i turns its first B member into A when
// 0 < r <= prob1(tBCellN(i)); several cells never name one particle twice
// (each particle is in one cell), and each writes the same value.
__global__ void k_react(long long numcell, int NT, const long long* tB, const long long* sB, const double* draws,
                        const double* tab, int* flag, int* err) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numcell) return;
  double r = draws[i];
  if (0.0 < r && r <= tab[tB[i]]) {
    long long q = sB[i] - 1;
    if (q < 0 || q >= NT) { raise_error(err, ERR_CELL_RANGE); return; }
    flag[q] = 0;
  }
}

// ----------------------------------------------------------- ordered compaction
// predicate per particle (neighbour_list_mod.f95:38-57, :129-167), then the
// indices of the hits in ascending order: counts per block, host scan, write.
__device__ __forceinline__ double pbc(double d, double L) {
  if (d > 0.5 * L) return d - L;
  if (d < -0.5 * L) return d + L;
  return d;
}
// mode 0: within sqrt(r2a) of sphere A; mode 1: within r2a of any C sphere or
// r2b of any N sphere (inside); mode 2: the complement (outside)
__device__ __forceinline__ bool pred(int mode, int i, const double* x, const double* y, const double* z, double b0,
                                     double b1, const double* sx, const double* sy, const double* sz, int nd,
                                     double r2a, double r2b) {
  if (mode == 0) {
    double d1 = pbc(x[i] - sx[0], b0), d2 = pbc(y[i] - sy[0], b1), d3 = z[i] - sz[0];
    return d1 * d1 + d2 * d2 + d3 * d3 <= r2a;
  }
  bool in = false;
  for (int m = 0; m < nd; m++)
    for (int j = 0; j < 2; j++) {
      int s = m + j * nd;
      double d1 = pbc(x[i] - sx[s], b0), d2 = pbc(y[i] - sy[s], b1), d3 = z[i] - sz[s];
      if (d1 * d1 + d2 * d2 + d3 * d3 <= (j == 0 ? r2a : r2b)) in = true;
    }
  return mode == 1 ? in : !in;
}
__global__ __launch_bounds__(CB) void k_compact_count(int mode, int NT, const double* x, const double* y,
                                                      const double* z, double b0, double b1, const double* sx,
                                                      const double* sy, const double* sz, int nd, double r2a,
                                                      double r2b, int* bcount) {
  __shared__ unsigned long long s[CB / 64 + 1];
  int i = blockIdx.x * CB + threadIdx.x;
  unsigned long long v = (i < NT && pred(mode, i, x, y, z, b0, b1, sx, sy, sz, nd, r2a, r2b)) ? 1 : 0;
  unsigned long long tot;
  block_excl_scan64(v, s, tot);
  if (threadIdx.x == 0) bcount[blockIdx.x] = (int)tot;
}
__global__ __launch_bounds__(CB) void k_compact_write(int mode, int NT, const double* x, const double* y,
                                                      const double* z, double b0, double b1, const double* sx,
                                                      const double* sy, const double* sz, int nd, double r2a,
                                                      double r2b, const int* boff, long long* out) {
  __shared__ unsigned long long s[CB / 64 + 1];
  int i = blockIdx.x * CB + threadIdx.x;
  bool p = i < NT && pred(mode, i, x, y, z, b0, b1, sx, sy, sz, nd, r2a, r2b);
  unsigned long long tot;
  unsigned long long e = block_excl_scan64(p ? 1 : 0, s, tot);
  if (p) out[boff[blockIdx.x] + (long long)e] = i + 1;
}

// exclusive scan of the nb block counts into off[0..nb] (one block, in block order)
__global__ __launch_bounds__(CB) void k_scan_blocks(int nb, const int* cnt, int* off) {
  __shared__ unsigned long long s[CB / 64 + 1];
  __shared__ long long run;
  if (threadIdx.x == 0) run = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += CB) {
    const int b = b0 + threadIdx.x;
    const unsigned long long v = b < nb ? (unsigned long long)cnt[b] : 0ull;
    unsigned long long tot;
    const unsigned long long e = block_excl_scan64(v, s, tot);
    if (b < nb) off[b] = (int)(run + (long long)e);
    __syncthreads();
    if (threadIdx.x == 0) run += (long long)tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) off[nb] = (int)run;
}

static int compact(int mode, int NT, Buf& x, Buf& y, Buf& z, const double* box, Buf& sph, int nd, double r2a,
                   double r2b, std::vector<long long>& result) {
  int nb = (NT + CB - 1) / CB;
  Buf bc, bo, out;
  if (!bc.alloc(sizeof(int) * nb) || !bo.alloc(sizeof(int) * (nb + 1)) || !out.alloc(sizeof(long long) * NT))
    return set_error(MPCX_E_HIP, "hipMalloc failed");
  const double* s = sph.as<double>();
  hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(CB), 0, 0, mode, NT, x.as<double>(), y.as<double>(),
                     z.as<double>(), box[0], box[1], s, s + 2 * nd, s + 4 * nd, nd, r2a, r2b, bc.as<int>());
  std::vector<int> cnt(nb), off;
  DCHK(hipMemcpy(cnt.data(), bc.p, sizeof(int) * nb, hipMemcpyDeviceToHost));
  host_scan(cnt, off);
  DCHK(hipMemcpy(bo.p, off.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_compact_write, dim3(nb), dim3(CB), 0, 0, mode, NT, x.as<double>(), y.as<double>(),
                     z.as<double>(), box[0], box[1], s, s + 2 * nd, s + 4 * nd, nd, r2a, r2b, bo.as<int>(),
                     out.as<long long>());
  result.resize((size_t)off[nb]);
  if (off[nb]) DCHK(hipMemcpy(result.data(), out.p, sizeof(long long) * off[nb], hipMemcpyDeviceToHost));
  return 0;
}

// ------------------------------------------------------------------- census
// distribute_mod.f95:95-183: every listed particle (outside list first, then
// the inside ones outside every cutoff) keyed by its place in that order;
// per (species, cell): count and the keys in order.
__global__ void k_census(long long n_out, long long n_in, const long long* ol, const long long* il, const int* flag,
                         const double* x, const double* y, const double* z, double b0, double b1, const double* sx,
                         const double* sy, const double* sz, int nd, double c2, double n2, long long vol,
                         int* cnt, int* rank, int* cellk, int* err) {
  long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_out + n_in) return;
  cellk[j] = -1;
  long long i = (j < n_out ? ol[j] : il[j - n_out]) - 1;
  if (j >= n_out) {
    for (int m = 0; m < nd; m++) {
      double a1 = pbc(x[i] - sx[m], b0), a2 = pbc(y[i] - sy[m], b1), a3 = z[i] - sz[m];
      double e1 = pbc(x[i] - sx[m + nd], b0), e2 = pbc(y[i] - sy[m + nd], b1), e3 = z[i] - sz[m + nd];
      if (a1 * a1 + a2 * a2 + a3 * a3 <= c2 || e1 * e1 + e2 * e2 + e3 * e3 <= n2) return;
    }
  }
  int f = flag[i];
  int sp = f == 0 ? 0 : f == 1 ? 1 : f == 2 ? 2 : -1;
  if (sp < 0) return;
  long long u = cell_unshifted0(x[i], y[i], z[i], b0, b1);
  if (u < 0 || u >= vol) { raise_error(err, ERR_CELL_RANGE); return; }
  long long k = sp * vol + u;
  cellk[j] = (int)k;
  rank[j] = atomicAdd(&cnt[k], 1);
}
__global__ void k_census_scatter(long long n, const int* cellk, const int* rank, const int* off, int* bucket) {
  long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || cellk[j] < 0) return;
  bucket[off[cellk[j]] + rank[j]] = (int)j;
}
__global__ void k_sort_buckets(long long ncell, const int* cnt, const int* off, int* bucket) {
  long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  int n = cnt[c];
  int* b = bucket + off[c];
  for (int j = 1; j < n; j++) {
    int v = b[j], k = j;
    while (k > 0 && b[k - 1] > v) { b[k] = b[k - 1]; k--; }
    b[k] = v;
  }
}

// ------------------------------------------------------------------ sorting
__global__ void k_sort_bin(int NT, const double* x, const double* y, const double* z, double b0, double b1,
                           long long vol, const int* tourpos, int* cnt, int* key, int* rank, int* err) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NT) return;
  long long u = cell_unshifted0(x[i], y[i], z[i], b0, b1);
  if (u < 0 || u >= vol || tourpos[u] < 0) { raise_error(err, ERR_CELL_RANGE); key[i] = -1; return; }
  key[i] = tourpos[u];
  rank[i] = atomicAdd(&cnt[key[i]], 1);
}
__global__ void k_sort_scatter(int NT, const int* key, const int* rank, const int* off, int* order) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NT || key[i] < 0) return;
  order[off[key[i]] + rank[i]] = i;
}
template <typename T>
__global__ void k_gather(long long n, const int* order, const T* src, T* dst) {
  long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) dst[j] = src[order[j]];
}

static int read_err(Buf& e) {
  int v = 0;
  DCHK(hipMemcpy(&v, e.p, sizeof(int), hipMemcpyDeviceToHost));
  if (v) return set_error(MPCX_E_DEVICE, "a particle maps outside the cell arrays");
  return 0;
}

}  // namespace dropin
}  // namespace mpcx

using namespace mpcx;
using namespace mpcx::dropin;

static inline unsigned nblk(long long n, int t = BT) { return (unsigned)std::max<long long>(1, (n + t - 1) / t); }

extern "C" int mpcx_mpc_rotation(int32_t NT, double rotang, int64_t numcell, double a0, const double* box,
                                 const double* Xsol, const double* Ysol, const double* Zsol, double* VXsol,
                                 double* VYsol, double* VZsol, double* VXCM, double* VYCM, double* VZCM,
                                 double* tCellShift, int64_t* NewCellN, double* UVX, double* UVY, double* UVZ,
                                 const double* draws) {
  if (NT < 0 || numcell < 1 || !box || !Xsol || !draws) return set_error(MPCX_E_ARG, "bad arguments");
  const size_t n = (size_t)NT, mc = (size_t)numcell;
  Buf x, y, z, vx, vy, vz, dr, ncn, cnt, rank, bucket, offb, cm[7], err;
  int r;
  if ((r = up(x, Xsol, n * 8)) || (r = up(y, Ysol, n * 8)) || (r = up(z, Zsol, n * 8)) ||
      (r = up(vx, VXsol, n * 8)) || (r = up(vy, VYsol, n * 8)) || (r = up(vz, VZsol, n * 8)) ||
      (r = up(dr, draws, (3 + 2 * mc) * 8)) || (r = up(ncn, nullptr, n * 8)) || (r = up(cnt, nullptr, mc * 4)) ||
      (r = up(rank, nullptr, n * 4)) || (r = up(bucket, nullptr, n * 4)) || (r = up(offb, nullptr, (mc + 1) * 4)) ||
      (r = up(err, nullptr, 4)))
    return r;
  for (auto& b : cm)
    if ((r = up(b, nullptr, mc * 8))) return r;
  DCHK(hipMemset(cnt.p, 0, mc * 4));
  DCHK(hipMemset(err.p, 0, 4));
  double xs = a0 * (draws[0] - 0.5), ys = a0 * (draws[1] - 0.5), zs = a0 * (draws[2] - 0.5);
  hipLaunchKernelGGL(k_cells, dim3(nblk(n)), dim3(BT), 0, 0, NT, box[0], box[1], xs, ys, zs, x.as<double>(),
                     y.as<double>(), z.as<double>(), (long long)numcell, ncn.as<long long>(), cnt.as<int>(),
                     rank.as<int>(), err.as<int>());
  if ((r = read_err(err))) return r;
  std::vector<int> c(mc), off;
  DCHK(hipMemcpy(c.data(), cnt.p, mc * 4, hipMemcpyDeviceToHost));
  host_scan(c, off);
  DCHK(hipMemcpy(offb.p, off.data(), (mc + 1) * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_scatter, dim3(nblk(n)), dim3(BT), 0, 0, NT, ncn.as<long long>(), rank.as<int>(),
                     offb.as<int>(), bucket.as<int>());
  const double twopi = 2.0 * (4.0 * atan(1.0));
  hipLaunchKernelGGL(k_cell_cm, dim3(nblk(mc)), dim3(BT), 0, 0, (long long)numcell, cnt.as<int>(), offb.as<int>(),
                     bucket.as<int>(), vx.as<double>(), vy.as<double>(), vz.as<double>(), dr.as<double>(), twopi,
                     cm[0].as<double>(), cm[1].as<double>(), cm[2].as<double>(), cm[3].as<double>(),
                     cm[4].as<double>(), cm[5].as<double>(), cm[6].as<double>());
  hipLaunchKernelGGL(k_rotate, dim3(nblk(n)), dim3(BT), 0, 0, NT, ncn.as<long long>(), cm[0].as<double>(),
                     cm[1].as<double>(), cm[2].as<double>(), cm[4].as<double>(), cm[5].as<double>(),
                     cm[6].as<double>(), cos(rotang), sin(rotang), vx.as<double>(), vy.as<double>(),
                     vz.as<double>());
  if ((r = down(VXsol, vx, n * 8)) || (r = down(VYsol, vy, n * 8)) || (r = down(VZsol, vz, n * 8)) ||
      (r = down(VXCM, cm[0], mc * 8)) || (r = down(VYCM, cm[1], mc * 8)) || (r = down(VZCM, cm[2], mc * 8)) ||
      (r = down(tCellShift, cm[3], mc * 8)) || (r = down(UVX, cm[4], mc * 8)) || (r = down(UVY, cm[5], mc * 8)) ||
      (r = down(UVZ, cm[6], mc * 8)) || (r = down(NewCellN, ncn, n * 8)))
    return r;
  return 0;
}

extern "C" int mpcx_reaction_simple(double MPCt, int64_t numcell, double k1, int32_t NT, int32_t* flag,
                                    const int64_t* tBCellN, const int64_t* sBCellN, int64_t vol, const double* draws) {
  if (NT < 0 || !flag || !tBCellN || !sBCellN || !draws || numcell < 0 || numcell > vol)
    return set_error(MPCX_E_ARG, "bad arguments");
  if (numcell == 0) return 0;
  // prob1 = 1 - dexp(-a1*MPCt) with a1 = k1*tBCellN(i) (reaction_mod.f95:45-54),
  // tabulated on the host per count so the device decision uses the host libm.
  int64_t nmax = 0;
  for (int64_t i = 0; i < numcell; i++) {
    if (tBCellN[i] < 0) return set_error(MPCX_E_ARG, "negative tBCellN");
    nmax = std::max(nmax, tBCellN[i]);
  }
  std::vector<double> tab((size_t)nmax + 1);
  for (int64_t n = 0; n <= nmax; n++) {
    double a1 = n == 0 ? 0.0 : k1 * (double)n;
    tab[n] = 1.0 - exp(-(a1 * MPCt));
  }
  Buf f, tb, sb, dr, tp, err;
  int r;
  const size_t nc = (size_t)numcell;
  if ((r = up(f, flag, (size_t)NT * 4)) || (r = up(tb, tBCellN, nc * 8)) || (r = up(sb, sBCellN, nc * 8)) ||
      (r = up(dr, draws, nc * 8)) || (r = up(tp, tab.data(), tab.size() * 8)) || (r = up(err, nullptr, 4)))
    return r;
  DCHK(hipMemset(err.p, 0, 4));
  hipLaunchKernelGGL(k_react, dim3(nblk(numcell)), dim3(BT), 0, 0, (long long)numcell, NT, tb.as<long long>(),
                     sb.as<long long>(), dr.as<double>(), tp.as<double>(), f.as<int>(), err.as<int>());
  if ((r = read_err(err))) return r;
  return down(flag, f, (size_t)NT * 4);
}

extern "C" int mpcx_neighbour_list(int32_t NT, const double* box, double cutoff, double skinrad, const double* Xsol,
                                   const double* Ysol, const double* Zsol, double Xdim, double Ydim, double Zdim,
                                   int64_t* neighbour_num, int64_t* neighbour, int64_t stride, int64_t max_neighbour) {
  if (!box || !Xsol || !neighbour_num || !neighbour || stride < 1) return set_error(MPCX_E_ARG, "bad arguments");
  Buf x, y, z, sph;
  int r;
  double s[3] = {Xdim, Ydim, Zdim};
  if ((r = up(x, Xsol, NT * 8ull)) || (r = up(y, Ysol, NT * 8ull)) || (r = up(z, Zsol, NT * 8ull)) ||
      (r = up(sph, nullptr, 6 * 8)))
    return r;
  double sv[6] = {s[0], 0, s[1], 0, s[2], 0};   // layout (x[2*nd], y[2*nd], z[2*nd]) with nd = 1
  DCHK(hipMemcpy(sph.p, sv, sizeof sv, hipMemcpyHostToDevice));
  double full = cutoff + skinrad;
  std::vector<long long> out;
  if ((r = compact(0, NT, x, y, z, box, sph, 1, full * full, 0.0, out))) return r;
  if (*neighbour_num + (int64_t)out.size() > max_neighbour)
    return set_error(MPCX_E_ARG, "neighbour list longer than max_neighbour");
  for (size_t j = 0; j < out.size(); j++) neighbour[(*neighbour_num + (int64_t)j) * stride] = out[j];
  *neighbour_num += (int64_t)out.size();
  return 0;
}

extern "C" int mpcx_inout_particle_list(int32_t NT, const double* box, double cutoffC, double cutoffN,
                                        double skinrad, const double* xsol, const double* ysol, const double* zsol,
                                        const double* xdim, const double* ydim, const double* zdim,
                                        int64_t* outside_particle_num, int64_t* outside_particle,
                                        int64_t* inside_particle_num, int64_t* inside_particle,
                                        int64_t inside_capacity, int32_t numdims) {
  if (!box || !xsol || !outside_particle_num || !inside_particle_num || numdims < 1)
    return set_error(MPCX_E_ARG, "bad arguments");
  Buf x, y, z, sph;
  int r;
  const int nd = numdims;
  std::vector<double> sv(6 * nd);
  for (int s = 0; s < 2 * nd; s++) { sv[s] = xdim[s]; sv[2 * nd + s] = ydim[s]; sv[4 * nd + s] = zdim[s]; }
  if ((r = up(x, xsol, NT * 8ull)) || (r = up(y, ysol, NT * 8ull)) || (r = up(z, zsol, NT * 8ull)) ||
      (r = up(sph, sv.data(), sv.size() * 8)))
    return r;
  double fc = cutoffC + skinrad, fn = cutoffN + skinrad;
  std::vector<long long> in, out;
  if ((r = compact(1, NT, x, y, z, box, sph, nd, fc * fc, fn * fn, in))) return r;
  if ((r = compact(2, NT, x, y, z, box, sph, nd, fc * fc, fn * fn, out))) return r;
  if (*inside_particle_num + (int64_t)in.size() > inside_capacity)
    return set_error(MPCX_E_ARG, "inside list longer than its array");
  memcpy(inside_particle + *inside_particle_num, in.data(), in.size() * 8);
  memcpy(outside_particle + *outside_particle_num, out.data(), out.size() * 8);
  *inside_particle_num += (int64_t)in.size();
  *outside_particle_num += (int64_t)out.size();
  return 0;
}

// ------------------------------------------ inout_particle_list, no allocation
// The scratch of mpcx_inout_particle_list_noalloc: the compaction's block
// counts and offsets and the sphere coordinates.  Every GPU the library runs
// on holds one copy until the process ends; host code reaches it only through
// the address hipGetSymbolAddress returns on that device (the array's name in
// host code is a stand-in, not a device address).
namespace mpcx {
namespace dropin {
constexpr long long NA_MAX_NT = 1ll << 26;              // particles per call
constexpr int NA_MAX_BLOCKS = (int)(NA_MAX_NT / CB);    // compaction blocks
constexpr int NA_MAX_DIMERS = 64;
constexpr size_t NA_OFF_COUNT = 0;                                          // int[NA_MAX_BLOCKS]
constexpr size_t NA_OFF_OFFS = NA_OFF_COUNT + 4 * (size_t)NA_MAX_BLOCKS;    // int[NA_MAX_BLOCKS + 1]
constexpr size_t NA_OFF_SPH = (NA_OFF_OFFS + 4 * ((size_t)NA_MAX_BLOCKS + 1) + 255) / 256 * 256;   // double[6 * 64]
constexpr size_t NA_SCRATCH_BYTES = NA_OFF_SPH + 8 * 6 * (size_t)NA_MAX_DIMERS;   // 527,616 bytes
static __device__ __align__(256) unsigned char scratch[NA_SCRATCH_BYTES];

constexpr int NA_MAX_DEVICES = 64;
static void* na_scratch_ptr[NA_MAX_DEVICES];
static std::mutex na_mutex;

// the current device's copy of scratch, resolved on the first call there
static int na_scratch(unsigned char** out) {
  int dev = -1;
  DCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEVICES) return set_error(MPCX_E_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(na_mutex);
  if (!na_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return set_error(MPCX_E_HIP, "hipGetSymbolAddress of the scratch array failed");
    na_scratch_ptr[dev] = p;
  }
  *out = (unsigned char*)na_scratch_ptr[dev];
  return 0;
}

// one list (mode 1 inside, 2 outside) of the particles in device memory,
// appended at out (room for cap entries); *added = its length
static int na_compact(int mode, int NT, const double* x, const double* y, const double* z, const double* box,
                      unsigned char* sc, int nd, double r2a, double r2b, long long* out, long long cap,
                      long long* added) {
  const int nb = (NT + CB - 1) / CB;
  int* cnt = (int*)(sc + NA_OFF_COUNT);
  int* off = (int*)(sc + NA_OFF_OFFS);
  const double* s = (const double*)(sc + NA_OFF_SPH);
  hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(CB), 0, 0, mode, NT, x, y, z, box[0], box[1], s, s + 2 * nd,
                     s + 4 * nd, nd, r2a, r2b, cnt);
  DCHK(hipGetLastError());
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(CB), 0, 0, nb, (const int*)cnt, off);
  DCHK(hipGetLastError());
  int total = 0;
  DCHK(hipMemcpy(&total, off + nb, sizeof(int), hipMemcpyDeviceToHost));
  if (total > cap) return set_error(MPCX_E_ARG, "particle list longer than its array");
  hipLaunchKernelGGL(k_compact_write, dim3(nb), dim3(CB), 0, 0, mode, NT, x, y, z, box[0], box[1], s, s + 2 * nd,
                     s + 4 * nd, nd, r2a, r2b, (const int*)off, out);
  DCHK(hipGetLastError());
  *added = total;
  return 0;
}
}  // namespace dropin
}  // namespace mpcx

extern "C" int mpcx_inout_particle_list_noalloc(int32_t NT, const double* box, double cutoffC, double cutoffN,
                                                double skinrad, const double* xsol, const double* ysol,
                                                const double* zsol, const double* xdim, const double* ydim,
                                                const double* zdim, int64_t* outside_particle_num,
                                                int64_t* outside_particle, int64_t outside_capacity,
                                                int64_t* inside_particle_num, int64_t* inside_particle,
                                                int64_t inside_capacity, int32_t numdims) {
  if (NT < 0 || (long long)NT > NA_MAX_NT)
    return set_error(MPCX_E_ARG, "NT beyond the no-allocation limit (67,108,864 particles)");
  if (numdims < 1 || numdims > NA_MAX_DIMERS)
    return set_error(MPCX_E_ARG, "numdims beyond the no-allocation limit (64 dimers)");
  if (!box || !xdim || !ydim || !zdim || !outside_particle_num || !inside_particle_num ||
      (NT > 0 && (!xsol || !ysol || !zsol || !outside_particle || !inside_particle)))
    return set_error(MPCX_E_ARG, "bad arguments");
  const long long on = *outside_particle_num, inn = *inside_particle_num;
  if (on < 0 || inn < 0 || on > outside_capacity || inn > inside_capacity)
    return set_error(MPCX_E_ARG, "list counts outside their arrays");
  if (NT == 0) return 0;
  unsigned char* sc = nullptr;
  int r = na_scratch(&sc);
  if (r) return r;
  const int nd = numdims;
  double* sph = (double*)(sc + NA_OFF_SPH);   // (x[2nd], y[2nd], z[2nd]), sphere s = k + i*nd
  DCHK(hipMemcpy(sph, xdim, 2 * nd * sizeof(double), hipMemcpyHostToDevice));
  DCHK(hipMemcpy(sph + 2 * nd, ydim, 2 * nd * sizeof(double), hipMemcpyHostToDevice));
  DCHK(hipMemcpy(sph + 4 * nd, zdim, 2 * nd * sizeof(double), hipMemcpyHostToDevice));
  const double fc = cutoffC + skinrad, fn = cutoffN + skinrad;
  long long n_in = 0, n_out = 0;
  if ((r = na_compact(1, NT, xsol, ysol, zsol, box, sc, nd, fc * fc, fn * fn, (long long*)inside_particle + inn,
                      inside_capacity - inn, &n_in)))
    return r;
  if ((r = na_compact(2, NT, xsol, ysol, zsol, box, sc, nd, fc * fc, fn * fn, (long long*)outside_particle + on,
                      outside_capacity - on, &n_out)))
    return r;
  DCHK(hipStreamSynchronize(nullptr));
  *inside_particle_num = inn + n_in;
  *outside_particle_num = on + n_out;
  return 0;
}

extern "C" int mpcx_distribute(int32_t NT, int32_t numdims, const int32_t* flag, const double* Xsol,
                               const double* Ysol, const double* Zsol, const double* Xdim, const double* Ydim,
                               const double* Zdim, const double* box, int64_t vol, int64_t* tACellN,
                               int64_t* tBCellN, int64_t* tSCellN, int64_t* sACellN, int64_t* sBCellN,
                               int64_t* sSCellN, int64_t maxs, double cutoffC2, double cutoffN2,
                               int64_t outside_particle_num, const int64_t* outside_particle,
                               int64_t inside_particle_num, const int64_t* inside_particle) {
  if (!flag || !Xsol || !box || vol < 1 || numdims < 1) return set_error(MPCX_E_ARG, "bad arguments");
  const int nd = numdims;
  const long long no = outside_particle_num, ni = inside_particle_num, nt = no + ni, nk = 3 * vol;
  if (no < 0 || ni < 0 || (no && !outside_particle) || (ni && !inside_particle))
    return set_error(MPCX_E_ARG, "bad particle lists");
  for (long long j = 0; j < no; j++)
    if (outside_particle[j] < 1 || outside_particle[j] > NT) return set_error(MPCX_E_ARG, "outside_particle entry outside 1..NT");
  for (long long j = 0; j < ni; j++)
    if (inside_particle[j] < 1 || inside_particle[j] > NT) return set_error(MPCX_E_ARG, "inside_particle entry outside 1..NT");
  std::vector<double> sv(6 * nd);
  for (int s = 0; s < 2 * nd; s++) { sv[s] = Xdim[s]; sv[2 * nd + s] = Ydim[s]; sv[4 * nd + s] = Zdim[s]; }
  Buf x, y, z, f, ol, il, sph, cnt, rank, ck, bucket, offb, err;
  int r;
  if ((r = up(x, Xsol, NT * 8ull)) || (r = up(y, Ysol, NT * 8ull)) || (r = up(z, Zsol, NT * 8ull)) ||
      (r = up(f, flag, NT * 4ull)) || (r = up(ol, outside_particle, no * 8)) || (r = up(il, inside_particle, ni * 8)) ||
      (r = up(sph, sv.data(), sv.size() * 8)) || (r = up(cnt, nullptr, nk * 4)) || (r = up(rank, nullptr, nt * 4)) ||
      (r = up(ck, nullptr, nt * 4)) || (r = up(bucket, nullptr, nt * 4)) || (r = up(offb, nullptr, (nk + 1) * 4)) ||
      (r = up(err, nullptr, 4)))
    return r;
  DCHK(hipMemset(cnt.p, 0, nk * 4));
  DCHK(hipMemset(err.p, 0, 4));
  const double* s = sph.as<double>();
  hipLaunchKernelGGL(k_census, dim3(nblk(nt)), dim3(BT), 0, 0, no, ni, ol.as<long long>(), il.as<long long>(),
                     f.as<int>(), x.as<double>(), y.as<double>(), z.as<double>(), box[0], box[1], s, s + 2 * nd,
                     s + 4 * nd, nd, cutoffC2, cutoffN2, (long long)vol, cnt.as<int>(), rank.as<int>(), ck.as<int>(),
                     err.as<int>());
  if ((r = read_err(err))) return r;
  std::vector<int> c(nk), off;
  DCHK(hipMemcpy(c.data(), cnt.p, nk * 4, hipMemcpyDeviceToHost));
  for (long long k = 0; k < nk; k++)
    if (c[k] > maxs) return set_error(MPCX_E_DEVICE, "a cell holds more members than the list arrays");
  host_scan(c, off);
  DCHK(hipMemcpy(offb.p, off.data(), (nk + 1) * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_census_scatter, dim3(nblk(nt)), dim3(BT), 0, 0, nt, ck.as<int>(), rank.as<int>(),
                     offb.as<int>(), bucket.as<int>());
  hipLaunchKernelGGL(k_sort_buckets, dim3(nblk(nk)), dim3(BT), 0, 0, nk, cnt.as<int>(), offb.as<int>(),
                     bucket.as<int>());
  std::vector<int> b((size_t)std::max<long long>(nt, 1));
  DCHK(hipMemcpy(b.data(), bucket.p, nt * 4, hipMemcpyDeviceToHost));
  int64_t* T[3] = {tACellN, tBCellN, tSCellN};
  int64_t* L[3] = {sACellN, sBCellN, sSCellN};
  for (int sp = 0; sp < 3; sp++)
    for (long long u = 0; u < vol; u++) {
      long long k = sp * vol + u;
      T[sp][u] = c[k];
      for (int j = 0; j < c[k]; j++) {   // list order -> particle number
        long long pos = b[off[k] + j];
        L[sp][u + (long long)j * vol] = pos < no ? outside_particle[pos] : inside_particle[pos - no];
      }
    }
  return 0;
}

extern "C" int mpcx_sort_hilbert(int64_t vol, int32_t NT, const double* box, const int64_t* tour, double* xsol,
                                 double* ysol, double* zsol, int32_t* flag, double* vxsol, double* vysol,
                                 double* vzsol) {
  if (!box || !tour || !xsol || !flag || vol < 1) return set_error(MPCX_E_ARG, "bad arguments");
  std::vector<int> tp((size_t)vol, -1);
  for (long long t = 0; t < vol; t++) {
    long long c = tour[t] - 1;
    if (c < 0 || c >= vol || tp[c] >= 0) return set_error(MPCX_E_ARG, "the tour must visit every cell exactly once");
    tp[c] = (int)t;
  }
  const size_t n = (size_t)NT;
  Buf a[6], fl, tpb, cnt, key, rank, order, offb, err, tmp;
  int r;
  double* h[6] = {xsol, ysol, zsol, vxsol, vysol, vzsol};
  for (int j = 0; j < 6; j++)
    if ((r = up(a[j], h[j], n * 8))) return r;
  if ((r = up(fl, flag, n * 4)) || (r = up(tpb, tp.data(), vol * 4)) || (r = up(cnt, nullptr, vol * 4)) ||
      (r = up(key, nullptr, n * 4)) || (r = up(rank, nullptr, n * 4)) || (r = up(order, nullptr, n * 4)) ||
      (r = up(offb, nullptr, (vol + 1) * 4)) || (r = up(err, nullptr, 4)) || (r = up(tmp, nullptr, n * 8)))
    return r;
  DCHK(hipMemset(cnt.p, 0, vol * 4));
  DCHK(hipMemset(err.p, 0, 4));
  hipLaunchKernelGGL(k_sort_bin, dim3(nblk(n)), dim3(BT), 0, 0, NT, a[0].as<double>(), a[1].as<double>(),
                     a[2].as<double>(), box[0], box[1], (long long)vol, tpb.as<int>(), cnt.as<int>(), key.as<int>(),
                     rank.as<int>(), err.as<int>());
  if ((r = read_err(err))) return r;
  std::vector<int> c((size_t)vol), off;
  DCHK(hipMemcpy(c.data(), cnt.p, vol * 4, hipMemcpyDeviceToHost));
  host_scan(c, off);
  DCHK(hipMemcpy(offb.p, off.data(), (vol + 1) * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_sort_scatter, dim3(nblk(n)), dim3(BT), 0, 0, NT, key.as<int>(), rank.as<int>(),
                     offb.as<int>(), order.as<int>());
  // stable: the particles of one cell keep ascending original index (sort_hilbert_mod.f95:50-54)
  hipLaunchKernelGGL(k_sort_buckets, dim3(nblk(vol)), dim3(BT), 0, 0, (long long)vol, cnt.as<int>(), offb.as<int>(),
                     order.as<int>());
  for (int j = 0; j < 6; j++) {
    hipLaunchKernelGGL(k_gather<double>, dim3(nblk(n)), dim3(BT), 0, 0, (long long)n, order.as<int>(),
                       a[j].as<double>(), tmp.as<double>());
    if ((r = down(h[j], tmp, n * 8))) return r;
  }
  hipLaunchKernelGGL(k_gather<int>, dim3(nblk(n)), dim3(BT), 0, 0, (long long)n, order.as<int>(), fl.as<int>(),
                     tmp.as<int>());
  return down(flag, tmp, n * 4);
}
