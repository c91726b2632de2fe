// Diagnostics on the device-resident state of either engine: the system
// observables, concentration slabs and catalytic shells the reference writes
// at its output cadence, and the body-frame (r,z) density around a dimer.
//   collective_code/output_mod.f95:681-805   momentum, kinetic energy, temperature
//   collective_code/output_mod.f95:486-678   A/B concentration slabs
//   collective_code/output_mod.f95:807-886   bond unit vector
//   selkov_code/Output_SelkovDimer.f90:144-184   propulsion velocity DotVZ
//   selkov_code/Output_Concentration_Catalytic.f90:7-191   shells around the C sphere
// Counts are integers (exact); sums are wave/block tree sums added on the host
// in block order (deterministic; the reference adds in particle order).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <vector>
#include "mpcx_device.h"
#include "mpcx_host.h"
#include "../../include/mpcx.h"

#pragma clang fp contract(off)

namespace mpcx {
namespace dg {

constexpr int DT = 256;
constexpr int DB = 1024;   // blocks of the reductions

struct ViewK {   // the kernel-side part of a DevView
  const double *x, *y, *z, *vx, *vy, *vz;
  const int* flag;
  long long n;
  double box[3];
  double mass[4];
  int species[4];
};

__device__ __forceinline__ int sp_of(const ViewK& v, int f) { return (f >= 0 && f < 4) ? v.species[f] : -1; }

// per block: counts of the three species columns, momentum, kinetic energy
__global__ __launch_bounds__(DT) void dg_sums(ViewK v, double* part) {
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (long long i = blockIdx.x * (long long)DT + threadIdx.x; i < v.n; i += (long long)gridDim.x * DT) {
    const int f = v.flag[i];
    const int sp = sp_of(v, f);
    if (sp < 0) continue;
    const double m = v.mass[f];
    const double vx = v.vx[i], vy = v.vy[i], vz = v.vz[i];
    const double px = m * vx, py = m * vy, pz = m * vz;
    acc[sp] += 1.0;
    acc[3] += px; acc[4] += py; acc[5] += pz;
    acc[6] += 0.5 * (vx * px + vy * py + vz * pz);
  }
  __shared__ double red[DT / 64][7];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 7; k++) {
    double r = wave_sum(acc[k]);
    if (lane == 0) red[w][k] = r;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    double a = red[0][threadIdx.x];
    for (int j = 1; j < DT / 64; j++) a = a + red[j][threadIdx.x];
    part[blockIdx.x * 7 + threadIdx.x] = a;
  }
}

// A/B counts in the slabs of output_mod.f95:503-568: bins of ceiling(x/width)
struct Slabs {
  double xs, ys, zs, half, xw, yw, zw;
  int nx, ny, nz;
};
__global__ __launch_bounds__(DT) void dg_slabs(ViewK v, Slabs S, int* out) {
  const int nx = S.nx, ny = S.ny, nz = S.nz;
  int* A1 = out;
  int* B1 = A1 + nx;
  int* Axy = B1 + nx;
  int* Bxy = Axy + nx * ny;
  int* Axz = Bxy + nx * ny;
  int* Bxz = Axz + nx * nz;
  int* Ayz = Bxz + nx * nz;
  int* Byz = Ayz + ny * nz;
  for (long long i = blockIdx.x * (long long)DT + threadIdx.x; i < v.n; i += (long long)gridDim.x * DT) {
    const int f = v.flag[i];
    if (f != 0 && f != 1) continue;
    const double x = v.x[i], y = v.y[i], z = v.z[i];
    // ceiling(x/width), 1-based; a particle exactly on the lower edge has no bin
    const int jx = (int)ceil(x / S.xw) - 1, jy = (int)ceil(y / S.yw) - 1, jz = (int)ceil(z / S.zw) - 1;
    const bool bx = jx >= 0 && jx < nx, by = jy >= 0 && jy < ny, bz = jz >= 0 && jz < nz;
    if (z <= S.zs + S.half && z >= S.zs - S.half) {
      if (bx) atomicAdd(f == 0 ? &A1[jx] : &B1[jx], 1);
      if (bx && by) atomicAdd(f == 0 ? &Axy[jx + nx * jy] : &Bxy[jx + nx * jy], 1);
    }
    if (y <= S.ys + S.half && y >= S.ys - S.half && bx && bz)
      atomicAdd(f == 0 ? &Axz[jx + nx * jz] : &Bxz[jx + nx * jz], 1);
    if (x <= S.xs + S.half && x >= S.xs - S.half && by && bz)
      atomicAdd(f == 0 ? &Ayz[jy + ny * jz] : &Byz[jy + ny * jz], 1);
  }
}

// Output_Concentration_Catalytic.f90:50-156: species counts in spherical shells
// around the C sphere (R^2 bounds 4, 9, 16, 36, 64, 100, 144, 225, 400)
__global__ __launch_bounds__(DT) void dg_shells(ViewK v, double cx, double cy, double cz, long long* out) {
  __shared__ int cnt[3][8];
  if (threadIdx.x < 24) cnt[threadIdx.x / 8][threadIdx.x % 8] = 0;
  __syncthreads();
  const double hb[3] = {0.5 * v.box[0], 0.5 * v.box[1], 0.5 * v.box[2]};
  const double mx = hb[0] - cx, my = hb[1] - cy, mz = hb[2] - cz;
  const double lim[9] = {4.0, 9.0, 16.0, 36.0, 64.0, 100.0, 144.0, 225.0, 400.0};
  for (long long i = blockIdx.x * (long long)DT + threadIdx.x; i < v.n; i += (long long)gridDim.x * DT) {
    const int sp = sp_of(v, v.flag[i]);
    if (sp < 0) continue;
    double X = v.x[i] + mx, Y = v.y[i] + my, Z = v.z[i] + mz;
    if (X > v.box[0]) X = X - v.box[0]; else if (X < 0.0) X = X + v.box[0];
    if (Y > v.box[1]) Y = Y - v.box[1]; else if (Y < 0.0) Y = Y + v.box[1];
    if (Z > v.box[2]) Z = Z - v.box[2]; else if (Z < 0.0) Z = Z + v.box[2];
    const double rx = X - hb[0], ry = Y - hb[1], rz = Z - hb[2];
    const double R2 = rx * rx + ry * ry + rz * rz;
    if (R2 < lim[0] || !(R2 < lim[8])) continue;
    int b = 0;
    while (!(R2 < lim[b + 1])) b++;
    atomicAdd(&cnt[sp][b], 1);
  }
  __syncthreads();
  if (threadIdx.x < 24) atomicAdd((unsigned long long*)&out[threadIdx.x], (unsigned long long)cnt[threadIdx.x / 8][threadIdx.x % 8]);
}

// counts per (species, z bin, r bin) in the frame of a dimer: origin at its C
// sphere, axis u from the N to the C sphere, z = (r - rC).u, r = |(r - rC) - z u|
struct Body {
  double ox, oy, oz, ux, uy, uz, rmax, zmin, zmax;
  int nr, nz, periodic_z;
};
__global__ __launch_bounds__(DT) void dg_body(ViewK v, Body B, unsigned* out) {
  const double dr = B.rmax / B.nr, dz = (B.zmax - B.zmin) / B.nz;
  for (long long i = blockIdx.x * (long long)DT + threadIdx.x; i < v.n; i += (long long)gridDim.x * DT) {
    const int sp = sp_of(v, v.flag[i]);
    if (sp < 0) continue;
    const double ax = pbc_half(v.x[i] - B.ox, v.box[0]);
    const double ay = pbc_half(v.y[i] - B.oy, v.box[1]);
    const double az = B.periodic_z ? pbc_half(v.z[i] - B.oz, v.box[2]) : v.z[i] - B.oz;
    const double zr = (ax * B.ux + ay * B.uy) + az * B.uz;
    const double px = ax - zr * B.ux, py = ay - zr * B.uy, pz = az - zr * B.uz;
    const double rr = sqrt((px * px + py * py) + pz * pz);
    if (!(zr >= B.zmin && zr < B.zmax && rr < B.rmax)) continue;
    const int iz = min((int)floor((zr - B.zmin) / dz), B.nz - 1), ir = min((int)floor(rr / dr), B.nr - 1);
    atomicAdd(&out[((long long)sp * B.nz + iz) * B.nr + ir], 1u);
  }
}

static ViewK kview(const DevView& v) {
  ViewK k;
  k.x = v.x; k.y = v.y; k.z = v.z; k.vx = v.vx; k.vy = v.vy; k.vz = v.vz; k.flag = v.flag; k.n = v.n;
  for (int i = 0; i < 3; i++) k.box[i] = v.box[i];
  for (int i = 0; i < 4; i++) { k.mass[i] = v.mass[i]; k.species[i] = v.species[i]; }
  return k;
}

static int grid_n(long long n) {
  long long g = (n + DT - 1) / DT;
  if (g > DB) g = DB;
  return (int)(g < 1 ? 1 : g);
}

#define DGCHK(x)                                                              \
  do {                                                                        \
    hipError_t _r = (x);                                                      \
    if (_r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(_r)); \
  } while (0)

// The observables of collective_code/output_mod.f95:681-756 (selkov = false)
// or selkov_code/Output_SelkovDimer.f90:147-300 (selkov = true).  The solvent
// sums are tree sums (the reference adds in particle order); the dimer terms
// and the temperatures follow the reference's expressions and constants: the
// temperatures multiply by the default-real literal 2./3. (0.6666666865...,
// the source files declare no -fdefault-real-8); in the Selkov form the dimer
// enters through its centre-of-mass velocity (momentum (Mc+Mn)*Vdimav,
// kinetic energy of the centre of mass), in the collective form sphere by sphere.
static int observables(const DevView& v, double* out, bool selkov) {
  hipStream_t st = (hipStream_t)v.stream;
  const int g = grid_n(v.n);
  double* part = nullptr;
  DGCHK(hipMalloc(&part, (size_t)g * 7 * 8));
  hipLaunchKernelGGL(dg_sums, dim3(g), dim3(DT), 0, st, kview(v), part);
  std::vector<double> h((size_t)g * 7);
  hipError_t r = hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, st);
  if (r == hipSuccess) r = hipStreamSynchronize(st);
  (void)hipFree(part);
  if (r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(r));
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < g; b++)
    for (int k = 0; k < 7; k++) s[k] = s[k] + h[(size_t)b * 7 + k];
  for (int i = 0; i < 16; i++) out[i] = 0.0;
  const double two_thirds = (double)(2.0f / 3.0f);
  const double n = s[0] + s[1] + s[2];
  out[0] = n; out[1] = s[0]; out[2] = s[1]; out[3] = s[2];
  out[4] = s[3]; out[5] = s[4]; out[6] = s[5];
  // DotVZ of dimer 0 (Output_SelkovDimer.f90:147-173): centre-of-mass velocity along the N -> C unit vector
  const int c = 0, nn = v.nd;
  double d[3];
  for (int a = 0; a < 3; a++) {
    d[a] = v.dim[a][c] - v.dim[a][nn];
    if (a < 2 || v.periodic_z) {
      if (d[a] > 0.5 * v.box[a]) d[a] = d[a] - v.box[a]; else if (d[a] < -0.5 * v.box[a]) d[a] = d[a] + v.box[a];
    }
  }
  const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double mc = v.msph[0], mn = v.msph[1];
  double vav[3];
  for (int a = 0; a < 3; a++) vav[a] = (mc * v.dim[3 + a][c] + mn * v.dim[3 + a][nn]) / (mn + mc);
  out[15] = (vav[0] * (d[0] / len) + vav[1] * (d[1] / len)) + vav[2] * (d[2] / len);
  double pt[3] = {s[3], s[4], s[5]}, ked = 0.0;
  if (selkov) {   // Output_SelkovDimer.f90:270-285
    double pd[3];
    for (int a = 0; a < 3; a++) {
      pd[a] = (mc + mn) * vav[a];
      pt[a] = pt[a] + pd[a];
    }
    ked = 0.0 + (0.5 * (1.0 / (mc + mn))) * ((pd[0] * pd[0] + pd[1] * pd[1]) + pd[2] * pd[2]);
  } else {        // output_mod.f95:719-732
    for (int k = 0; k < v.nd; k++) {
      const int sc = k, sn = k + v.nd;
      const double pc[3] = {mc * v.dim[3][sc], mc * v.dim[4][sc], mc * v.dim[5][sc]};
      const double pn[3] = {mn * v.dim[3][sn], mn * v.dim[4][sn], mn * v.dim[5][sn]};
      for (int a = 0; a < 3; a++) pt[a] = (pt[a] + pc[a]) + pn[a];
      ked = (ked + (0.5 * (1.0 / mc)) * ((pc[0] * pc[0] + pc[1] * pc[1]) + pc[2] * pc[2])) +
            (0.5 * (1.0 / mn)) * ((pn[0] * pn[0] + pn[1] * pn[1]) + pn[2] * pn[2]);
    }
  }
  out[7] = pt[0]; out[8] = pt[1]; out[9] = pt[2];
  out[10] = s[6]; out[11] = ked;
  out[12] = n > 0 ? (s[6] / n) * two_thirds : 0.0;
  out[13] = ked * two_thirds;
  out[14] = (ked + s[6]) + v.tpotential;
  return 0;
}

static int body_density(const DevView& v, int dimer, int nr, int nz, double rmax, double zmin, double zmax,
                        uint32_t* out) {
  if (dimer < 0 || dimer >= v.nd || nr < 1 || nz < 1 || !(rmax > 0) || !(zmax > zmin) || !out)
    return set_error(MPCX_E_ARG, "bad arguments");
  hipStream_t st = (hipStream_t)v.stream;
  Body B;
  const int c = dimer, nn = dimer + v.nd;
  B.ox = v.dim[0][c]; B.oy = v.dim[1][c]; B.oz = v.dim[2][c];
  double d[3];
  for (int a = 0; a < 3; a++) {
    d[a] = v.dim[a][c] - v.dim[a][nn];
    if (a < 2 || v.periodic_z) {
      if (d[a] > 0.5 * v.box[a]) d[a] = d[a] - v.box[a]; else if (d[a] < -0.5 * v.box[a]) d[a] = d[a] + v.box[a];
    }
  }
  const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  B.ux = d[0] / len; B.uy = d[1] / len; B.uz = d[2] / len;
  B.rmax = rmax; B.zmin = zmin; B.zmax = zmax; B.nr = nr; B.nz = nz; B.periodic_z = v.periodic_z;
  const size_t bins = (size_t)3 * nz * nr;
  unsigned* h = nullptr;
  DGCHK(hipMalloc(&h, bins * 4));
  hipError_t r = hipMemsetAsync(h, 0, bins * 4, st);
  if (r == hipSuccess) {
    hipLaunchKernelGGL(dg_body, dim3(grid_n(v.n)), dim3(DT), 0, st, kview(v), B, h);
    r = hipMemcpyAsync(out, h, bins * 4, hipMemcpyDeviceToHost, st);
  }
  if (r == hipSuccess) r = hipStreamSynchronize(st);
  (void)hipFree(h);
  if (r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(r));
  return 0;
}

}  // namespace dg
}  // namespace mpcx

using namespace mpcx;
using namespace mpcx::dg;

extern "C" int mpcx_co_observables(mpcx_co* e, double* out16) {
  if (!out16) return set_error(MPCX_E_ARG, "null argument");
  DevView v;
  int rc = co_view(e, &v);
  return rc ? rc : observables(v, out16, false);
}

extern "C" int mpcx_sk_observables(mpcx_sk* e, double* out16) {
  if (!out16) return set_error(MPCX_E_ARG, "null argument");
  DevView v;
  int rc = sk_view(e, &v);
  return rc ? rc : observables(v, out16, true);
}

extern "C" int mpcx_co_conc_slabs(mpcx_co* e, double x_slab, double y_slab, double z_slab, double half_width,
                                  double x_width, double y_width, double z_width, int32_t* counts) {
  if (!counts || !(x_width > 0) || !(y_width > 0) || !(z_width > 0)) return set_error(MPCX_E_ARG, "bad arguments");
  DevView v;
  int rc = co_view(e, &v);
  if (rc) return rc;
  Slabs S;
  S.xs = x_slab; S.ys = y_slab; S.zs = z_slab; S.half = half_width; S.xw = x_width; S.yw = y_width; S.zw = z_width;
  S.nx = (int)(v.box[0] / x_width); S.ny = (int)(v.box[1] / y_width); S.nz = (int)(v.box[2] / z_width);
  const size_t n = 2 * ((size_t)S.nx + (size_t)S.nx * S.ny + (size_t)S.nx * S.nz + (size_t)S.ny * S.nz);
  hipStream_t st = (hipStream_t)v.stream;
  int* d = nullptr;
  DGCHK(hipMalloc(&d, n * 4));
  hipError_t r = hipMemsetAsync(d, 0, n * 4, st);
  if (r == hipSuccess) {
    hipLaunchKernelGGL(dg_slabs, dim3(grid_n(v.n)), dim3(DT), 0, st, kview(v), S, d);
    r = hipMemcpyAsync(counts, d, n * 4, hipMemcpyDeviceToHost, st);
  }
  if (r == hipSuccess) r = hipStreamSynchronize(st);
  (void)hipFree(d);
  if (r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(r));
  return 0;
}

extern "C" int mpcx_sk_shells(mpcx_sk* e, int64_t* counts24) {
  if (!counts24) return set_error(MPCX_E_ARG, "null argument");
  DevView v;
  int rc = sk_view(e, &v);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)v.stream;
  long long* d = nullptr;
  DGCHK(hipMalloc(&d, 24 * 8));
  hipError_t r = hipMemsetAsync(d, 0, 24 * 8, st);
  if (r == hipSuccess) {
    hipLaunchKernelGGL(dg_shells, dim3(grid_n(v.n)), dim3(DT), 0, st, kview(v), v.dim[0][0], v.dim[1][0], v.dim[2][0], d);
    r = hipMemcpyAsync(counts24, d, 24 * 8, hipMemcpyDeviceToHost, st);
  }
  if (r == hipSuccess) r = hipStreamSynchronize(st);
  (void)hipFree(d);
  if (r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(r));
  return 0;
}

extern "C" int mpcx_co_body_density(mpcx_co* e, int32_t dimer, int32_t nr, int32_t nz, double rmax, double zmin,
                                    double zmax, uint32_t* counts) {
  DevView v;
  int rc = co_view(e, &v);
  return rc ? rc : body_density(v, dimer, nr, nz, rmax, zmin, zmax, counts);
}

extern "C" int mpcx_sk_body_density(mpcx_sk* e, int32_t nr, int32_t nz, double rmax, double zmin, double zmax,
                                    uint32_t* counts) {
  DevView v;
  int rc = sk_view(e, &v);
  return rc ? rc : body_density(v, 0, nr, nz, rmax, zmin, zmax, counts);
}
