// Selkov engine: the MD + reactive-MPC time loop of selkov_code on one MI355X
// (Main_SelkovDimer.f90:125-149).
//
// Work split per collision interval (MPCtstep MD steps, the last of which
// ends in the reactive collision):
//  * near engine (one persistent 512-thread workgroup): the particles that can
//    come within a sphere's cutoff during the interval ("near" set), plus the
//    dimer.  Every MD step it moves them (Motion_SelkovDimer.f90:41-64),
//    computes their sphere forces with the C-sphere conversions in particle
//    order (ForcesMD_SelkovDimer.f90:36-246; in parity mode the k-th converting
//    particle of a step takes the k-th host draw after those already used),
//    kicks them (:251-270) and integrates the dimer with RATTLE
//    (Motion_SelkovDimer.f90:66-159, ForcesMD_SelkovDimer.f90:272-303).
//  * bulk kernels over all particles: a far particle never feels a force, so
//    its MPCtstep moves "x + v*dt, wrapped" are replayed in registers in one
//    pass, fused with the census (Distribute_SelkovDimer.f90) and the
//    shifted-cell binning; then per cell the ordered centre-of-mass sums,
//    axes and reactions (RMPC_SelkovDimer.f90:15-324), and the rotation.
// A far particle is one whose distance beyond every cutoff exceeds
// |v|*tau + guard and that carries no force; the near engine checks every MD
// step that no sphere moved beyond the guard.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <string>
#include <vector>
#include "mpcx_device.h"
#include "mpcx_host.h"
#include "mpcx_kernels.h"
#include "../../include/mpcx.h"

#pragma clang fp contract(off)

namespace mpcx {
namespace sk {

constexpr int SK_T = 512;              // near-engine threads
constexpr int SK_W = SK_T / 64;
constexpr int SK_EMAX = 64;            // near entries per lane
constexpr int SK_NEAR = SK_T * SK_EMAX;
constexpr int SLCAP = 1792;            // near slots held in LDS (the rest in HBM)
constexpr int SK_MAXS = 200;           // MaxF/MaxG/MaxI, Global_SelkovDimer.f90:129-131
constexpr int SK_BMAX = 64;            // particles per collision cell in the ordered sum
constexpr int CT = 256;
constexpr long long RATTLE_CAP = 100000;
enum : int { SM_STOP = 2, SM_RESUME = 4 };
enum : unsigned char { CF_FORCE = 0x80 };   // cflag bit: the particle carries a force
enum : int { SX = 0, SY, SZ, SVX, SVY, SVZ, SFX, SFY, SFZ, SQX, SQY, SQZ };

struct SkScalars {
  long long cstep;      // last completed MD step
  long long cursor;     // host draws used so far (parity mode)
  int near_num, err, NT, pad;
  double tpotential;
  double bond[3];       // dX, dY, dZ left by the dimer motion for the velocity RATTLE
  double x0[2], y0[2], z0[2];
  unsigned long long prof[8];
  double xs0[2], ys0[2], zs0[2];   // spheres at the interval start (kept while the near engine moves them)
};

struct SkP {
  int NT, space, S, sort_step, rng_mode, Lx, Ly, Lz;
  long long vol;
  unsigned long long seed;
  double box[3], hbox[3];
  double MDt, IMDt, Dcn, hc, hn, IMc, IMn, IIMcIMn;
  double h2[4], hk[4];            // by flag 0 (F), 1 (G), 3 (I): move and kick factors
  double cutC2, cutN2, cutC, cutN;
  double R6[2], R12[2], r6R6[2], r12R12[2];   // [0] C sphere, [1] N sphere
  double EC[4], EN[4];
  double RPfg, RPgf, A0, k[6], MPCt, cosrot, sinrot, twopi, tau, guard;
};

struct SkD {
  double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz;
  int* flag;
  unsigned char* cflag;   // census flag of near particles at the collision step, | CF_FORCE
  double* dim;            // [12][2]
  SkScalars* sc;
  unsigned* near_bits;
  int* near_list;
  int near_cap;
  double *ns_x, *ns_y, *ns_z, *ns_vx, *ns_vy, *ns_vz, *ns_fx, *ns_fy, *ns_fz;
  int* ns_flag;
  unsigned char* ns_cflag;
  int *cnt, *off, *bucket, *partial, *cell_of;
  unsigned short* rank_of;
  double* cellrec;        // [vol][8]: vcm x y z, axis x y z, n, 0
  unsigned char* dq;
  int* tcount;            // [3][vol]: F, G, I census counts
  unsigned* first;        // [3][vol]: first member (lowest index) of each species
  int *unaff, *uoff;      // unaffected cells and their prefix (reaction draw slots)
  const double* U;        // host draws (rng_mode 0)
  long long nU;
  const double* ptab;     // [2][201][201] reaction probability by (tI > 0, tF, tG)
  unsigned long long coll;
};

__device__ __forceinline__ int species(int f) { return f == 0 ? 0 : f == 1 ? 1 : f == 3 ? 2 : -1; }
__device__ __forceinline__ bool active(int f) { return f == 0 || f == 1 || f == 3; }
__device__ __forceinline__ double wrap(double x, double L) { return x - L * floor(x / L); }

// INT(FLOOR(X)+Box(1)*FLOOR(Y)+Box(1)*Box(2)*FLOOR(Z)+1) - 1
__device__ __forceinline__ long long sk_cell0(const SkP& P, double x, double y, double z) {
  double c = ffloor_i(x) + P.box[0] * ffloor_i(y);
  c = c + (P.box[0] * P.box[1]) * ffloor_i(z);
  return (long long)(c + 1.0) - 1;
}

// the host draw at position idx of the parity stream, bounds-checked
__device__ __forceinline__ double host_draw(const SkD& D, long long idx) {
  if (idx < 0 || idx >= D.nU) { raise_error(&D.sc->err, ERR_DRAWS); return 0.5; }
  return D.U[idx];
}

// ------------------------------------------------------------- dimer (thread 0)
struct Dim2 { double v[12][2]; double bond[3]; };

// Motion_SelkovDimer.f90:66-159: advance the spheres and RATTLE the bond length
__device__ __forceinline__ void sk_dimer_motion(const SkP& P, Dim2& d, int* err) {
  const double MDt = P.MDt;
  double prX[2], prY[2], prZ[2], prVX[2], prVY[2], prVZ[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    prX[i] = d.v[SX][i]; prY[i] = d.v[SY][i]; prZ[i] = d.v[SZ][i];
    prVX[i] = d.v[SVX][i]; prVY[i] = d.v[SVY][i]; prVZ[i] = d.v[SVZ][i];
  }
  const double dXp = pbc_h(prX[0] - prX[1], P.hbox[0], P.box[0]);
  const double dYp = pbc_h(prY[0] - prY[1], P.hbox[1], P.box[1]);
  const double dZp = pbc_h(prZ[0] - prZ[1], P.hbox[2], P.box[2]);
  d.v[SQX][0] = prVX[0] + P.hc * d.v[SFX][0]; d.v[SQY][0] = prVY[0] + P.hc * d.v[SFY][0];
  d.v[SQZ][0] = prVZ[0] + P.hc * d.v[SFZ][0];
  d.v[SQX][1] = prVX[1] + P.hn * d.v[SFX][1]; d.v[SQY][1] = prVY[1] + P.hn * d.v[SFY][1];
  d.v[SQZ][1] = prVZ[1] + P.hn * d.v[SFZ][1];
#pragma unroll
  for (int i = 0; i < 2; i++) {
    d.v[SX][i] = wrap(prX[i] + MDt * d.v[SQX][i], P.box[0]);
    d.v[SY][i] = wrap(prY[i] + MDt * d.v[SQY][i], P.box[1]);
    d.v[SZ][i] = wrap(prZ[i] + MDt * d.v[SQZ][i], P.box[2]);
  }
  for (long long acc = 0;; acc++) {
    const double dx = pbc_h(d.v[SX][0] - d.v[SX][1], P.hbox[0], P.box[0]);
    const double dy = pbc_h(d.v[SY][0] - d.v[SY][1], P.hbox[1], P.box[1]);
    const double dz = pbc_h(d.v[SZ][0] - d.v[SZ][1], P.hbox[2], P.box[2]);
    d.bond[0] = dx; d.bond[1] = dy; d.bond[2] = dz;
    const double c1 = (dx * dx + dy * dy + dz * dz) - P.Dcn * P.Dcn;
    if (fabs(c1) < 1e-6) break;
    if (acc == RATTLE_CAP) { raise_error(err, ERR_RATTLE); break; }
    const double G = (((0.5 * c1) * P.IMDt) * P.IIMcIMn) / (dx * dXp + dy * dYp + dz * dZp);
    d.v[SQX][0] = d.v[SQX][0] - (G * P.IMc) * dXp; d.v[SQY][0] = d.v[SQY][0] - (G * P.IMc) * dYp;
    d.v[SQZ][0] = d.v[SQZ][0] - (G * P.IMc) * dZp;
    d.v[SQX][1] = d.v[SQX][1] + (G * P.IMn) * dXp; d.v[SQY][1] = d.v[SQY][1] + (G * P.IMn) * dYp;
    d.v[SQZ][1] = d.v[SQZ][1] + (G * P.IMn) * dZp;
#pragma unroll
    for (int i = 0; i < 2; i++) {
      d.v[SX][i] = wrap(prX[i] + MDt * d.v[SQX][i], P.box[0]);
      d.v[SY][i] = wrap(prY[i] + MDt * d.v[SQY][i], P.box[1]);
      d.v[SZ][i] = wrap(prZ[i] + MDt * d.v[SQZ][i], P.box[2]);
    }
  }
}

// ForcesMD_SelkovDimer.f90:237-303: sphere forces from the (tree-reduced)
// solvent sums, velocity update and velocity RATTLE.  sum[0..2]: sum of the C
// pair forces, sum[3..5]: N, sum[6]: potential.
__device__ __forceinline__ void sk_dimer_forces(const SkP& P, Dim2& d, const double* sum, double& tpot, int* err) {
  d.v[SFX][0] = 0.0 - sum[0]; d.v[SFY][0] = 0.0 - sum[1]; d.v[SFZ][0] = 0.0 - sum[2];
  d.v[SFX][1] = 0.0 - sum[3]; d.v[SFY][1] = 0.0 - sum[4]; d.v[SFZ][1] = 0.0 - sum[5];
  tpot = 0.0 + sum[6];
  d.v[SVX][0] = d.v[SQX][0] + P.hc * d.v[SFX][0]; d.v[SVY][0] = d.v[SQY][0] + P.hc * d.v[SFY][0];
  d.v[SVZ][0] = d.v[SQZ][0] + P.hc * d.v[SFZ][0];
  d.v[SVX][1] = d.v[SQX][1] + P.hn * d.v[SFX][1]; d.v[SVY][1] = d.v[SQY][1] + P.hn * d.v[SFY][1];
  d.v[SVZ][1] = d.v[SQZ][1] + P.hn * d.v[SFZ][1];
  const double dx = d.bond[0], dy = d.bond[1], dz = d.bond[2];
  for (long long acc = 0;; acc++) {
    const double dvx = d.v[SVX][0] - d.v[SVX][1], dvy = d.v[SVY][0] - d.v[SVY][1], dvz = d.v[SVZ][0] - d.v[SVZ][1];
    const double c2 = dvx * dx + dvy * dy + dvz * dz;
    if (fabs(c2) < 1e-8) break;
    if (acc == RATTLE_CAP) { raise_error(err, ERR_RATTLE); break; }
    const double O = (P.IIMcIMn * c2) / (P.Dcn * P.Dcn);
    d.v[SVX][0] = d.v[SVX][0] - (O * dx) * P.IMc; d.v[SVY][0] = d.v[SVY][0] - (O * dy) * P.IMc;
    d.v[SVZ][0] = d.v[SVZ][0] - (O * dz) * P.IMc;
    d.v[SVX][1] = d.v[SVX][1] + (O * dx) * P.IMn; d.v[SVY][1] = d.v[SVY][1] + (O * dy) * P.IMn;
    d.v[SVZ][1] = d.v[SVZ][1] + (O * dz) * P.IMn;
  }
}

// ------------------------------------------------------------- near engine
struct SkShared {
  double lx[SLCAP], ly[SLCAP], lz[SLCAP], lvx[SLCAP], lvy[SLCAP], lvz[SLCAP], lfx[SLCAP], lfy[SLCAP], lfz[SLCAP];
  int lflag[SLCAP];
  unsigned char lcflag[SLCAP];
  double sd[12][2];
  double bond[3];
  double red[SK_W][8];
  double fin[8];
  int wsum[SK_W + 1];
  long long cursor, cstep;
  int nn;
  SkP P;
  SkD D;
};
__shared__ SkShared g_sk;

// slot q: LDS below SLCAP, the HBM near arrays above
struct SkSlot {
  double x, y, z, vx, vy, vz, fx, fy, fz;
  int flag;
};
__device__ __forceinline__ SkSlot slot_load(const SkD& D, int q) {
  SkSlot s;
  if (q < SLCAP) {
    s.x = g_sk.lx[q]; s.y = g_sk.ly[q]; s.z = g_sk.lz[q]; s.vx = g_sk.lvx[q]; s.vy = g_sk.lvy[q]; s.vz = g_sk.lvz[q];
    s.fx = g_sk.lfx[q]; s.fy = g_sk.lfy[q]; s.fz = g_sk.lfz[q]; s.flag = g_sk.lflag[q];
  } else {
    s.x = D.ns_x[q]; s.y = D.ns_y[q]; s.z = D.ns_z[q]; s.vx = D.ns_vx[q]; s.vy = D.ns_vy[q]; s.vz = D.ns_vz[q];
    s.fx = D.ns_fx[q]; s.fy = D.ns_fy[q]; s.fz = D.ns_fz[q]; s.flag = D.ns_flag[q];
  }
  return s;
}
__device__ __forceinline__ void slot_store(const SkD& D, int q, const SkSlot& s) {
  if (q < SLCAP) {
    g_sk.lx[q] = s.x; g_sk.ly[q] = s.y; g_sk.lz[q] = s.z; g_sk.lvx[q] = s.vx; g_sk.lvy[q] = s.vy; g_sk.lvz[q] = s.vz;
    g_sk.lfx[q] = s.fx; g_sk.lfy[q] = s.fy; g_sk.lfz[q] = s.fz; g_sk.lflag[q] = s.flag;
  } else {
    D.ns_x[q] = s.x; D.ns_y[q] = s.y; D.ns_z[q] = s.z; D.ns_vx[q] = s.vx; D.ns_vy[q] = s.vy; D.ns_vz[q] = s.vz;
    D.ns_fx[q] = s.fx; D.ns_fy[q] = s.fy; D.ns_fz[q] = s.fz; D.ns_flag[q] = s.flag;
  }
}

// Motion_SelkovDimer.f90:41-64 for one particle (pf = the force of the last step)
__device__ __forceinline__ void sk_move(const SkP& P, SkSlot& s) {
  if (active(s.flag)) {
    const double h = P.h2[s.flag];
    s.x = (s.x + s.vx * P.MDt) + s.fx * h;
    s.y = (s.y + s.vy * P.MDt) + s.fy * h;
    s.z = (s.z + s.vz * P.MDt) + s.fz * h;
  }
  s.x = wrap(s.x, P.box[0]);
  s.y = wrap(s.y, P.box[1]);
  s.z = wrap(s.z, P.box[2]);
}

// distance of the slot's particle to sphere sph, squared, with the components
__device__ __forceinline__ double sk_d2(const SkP& P, const SkSlot& s, int sph, double& d1, double& d2, double& d3) {
  d1 = pbc_h(s.x - g_sk.sd[SX][sph], P.hbox[0], P.box[0]);
  d2 = pbc_h(s.y - g_sk.sd[SY][sph], P.hbox[1], P.box[1]);
  d3 = pbc_h(s.z - g_sk.sd[SZ][sph], P.hbox[2], P.box[2]);
  return d1 * d1 + d2 * d2 + d3 * d3;
}

// does the slot's particle take a conversion draw this step (F or G within the C cutoff)?
__device__ __forceinline__ bool sk_draws(const SkP& P, const SkSlot& s) {
  if (s.flag != 0 && s.flag != 1) return false;
  double a, b, c;
  return sk_d2(P, s, 0, a, b, c) <= P.cutC2;
}

// the lane's near indices: wave w owns [w*64*E, (w+1)*64*E), lane-interleaved
__device__ __forceinline__ int sk_E(int nn) { return (nn + SK_T - 1) / SK_T; }

// exclusive scan over the waves of g_sk.wsum (counts) -> offsets, total in wsum[SK_W]
__device__ __forceinline__ void sk_wave_scan() {
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < SK_W; w++) { int t = g_sk.wsum[w]; g_sk.wsum[w] = run; run += t; }
    g_sk.wsum[SK_W] = run;
  }
  __syncthreads();
}

__device__ __attribute__((noinline)) void sk_begin_motion() {
  // thread 0: the dimer part of Motion_SelkovDimer for the next step, and the guard check
  const SkP& P = g_sk.P;
  const SkD& D = g_sk.D;
  Dim2 d;
#pragma unroll
  for (int f = 0; f < 12; f++) { d.v[f][0] = g_sk.sd[f][0]; d.v[f][1] = g_sk.sd[f][1]; }
  sk_dimer_motion(P, d, &D.sc->err);
#pragma unroll
  for (int f = 0; f < 12; f++) { g_sk.sd[f][0] = d.v[f][0]; g_sk.sd[f][1] = d.v[f][1]; }
  g_sk.bond[0] = d.bond[0]; g_sk.bond[1] = d.bond[1]; g_sk.bond[2] = d.bond[2];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    double a = pbc_h(d.v[SX][s] - D.sc->x0[s], P.hbox[0], P.box[0]);
    double b = pbc_h(d.v[SY][s] - D.sc->y0[s], P.hbox[1], P.box[1]);
    double c = pbc_h(d.v[SZ][s] - D.sc->z0[s], P.hbox[2], P.box[2]);
    if (a * a + b * b + c * c > P.guard * P.guard) raise_error(&D.sc->err, ERR_NEAR_GUARD);
  }
}

__device__ __attribute__((noinline)) void sk_end_forces() {
  // thread 0: the dimer part of ForcesMD_SelkovDimer
  const SkP& P = g_sk.P;
  const SkD& D = g_sk.D;
  Dim2 d;
#pragma unroll
  for (int f = 0; f < 12; f++) { d.v[f][0] = g_sk.sd[f][0]; d.v[f][1] = g_sk.sd[f][1]; }
  d.bond[0] = g_sk.bond[0]; d.bond[1] = g_sk.bond[1]; d.bond[2] = g_sk.bond[2];
  double sum[7];
#pragma unroll
  for (int v = 0; v < 7; v++) sum[v] = g_sk.fin[v];
  double tpot;
  sk_dimer_forces(P, d, sum, tpot, &D.sc->err);
  D.sc->tpotential = tpot;
#pragma unroll
  for (int f = 0; f < 12; f++) { g_sk.sd[f][0] = d.v[f][0]; g_sk.sd[f][1] = d.v[f][1]; }
}

// Parity mode: the number of conversion draws each wave takes this step, and
// their exclusive offsets (sk_wave_scan).  Returns nothing; offsets in wsum.
__device__ __attribute__((noinline)) void sk_count_draws() {
  const SkP& P = g_sk.P;
  const SkD& D = g_sk.D;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nn = g_sk.nn, E = sk_E(nn), qb = wv * 64 * E + lane;
  int cnt = 0;
  for (int j = 0; j < E; j++) {
    const int q = qb + j * 64;
    bool need = false;
    if (q < nn) need = sk_draws(P, slot_load(D, q));
    cnt += __popcll(__ballot(need));
  }
  if (lane == 0) g_sk.wsum[wv] = cnt;
  sk_wave_scan();
}

// One MD step of the near set after its motion: forces with conversions
// (ForcesMD_SelkovDimer.f90:42-234), kick (:255-269), and the motion of the
// next step unless `last` (Motion_SelkovDimer.f90:41-64).  At `census` the
// flag before this step's conversions is kept for the census.
__device__ __attribute__((noinline)) void sk_step(bool last, bool census) {
  const SkP& P = g_sk.P;
  const SkD& D = g_sk.D;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nn = g_sk.nn, E = sk_E(nn), qb = wv * 64 * E + lane;
  const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const long long step = g_sk.cstep;
  long long base = g_sk.cursor + (P.rng_mode == 0 ? g_sk.wsum[wv] : 0);
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < E; j++) {
    const int q = qb + j * 64;
    const bool valid = q < nn;
    SkSlot s;
    if (valid) s = slot_load(D, q);
    const bool need = valid && sk_draws(P, s);
    const unsigned long long bal = __ballot(need);
    if (valid) {
      const int f0 = s.flag;
      if (census) D.ns_cflag[q] = (unsigned char)f0;
      const double pfx = s.fx, pfy = s.fy, pfz = s.fz;
      double fx = 0.0, fy = 0.0, fz = 0.0, pot = 0.0;
      double nx = 0.0, ny = 0.0, nz = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
      double d1, d2, d3;
      // N sphere (:42-113)
      double r2 = sk_d2(P, s, 1, d1, d2, d3);
      if (r2 <= P.cutN2 && active(s.flag)) {
        const double E_ = P.EN[s.flag];
        double F, U;
        lj_terms(r2, P.R6[1], P.R12[1], P.r12R12[1], P.r6R6[1], F, U);
        const double EF = E_ * F;
        nx = EF * d1; ny = EF * d2; nz = EF * d3;
        pot = E_ * U;
        fx = nx; fy = ny; fz = nz;
      }
      // C sphere (:123-234), conversions in particle order
      r2 = sk_d2(P, s, 0, d1, d2, d3);
      if (r2 <= P.cutC2 && active(s.flag)) {
        double E_;
        if (s.flag == 0 || s.flag == 1) {
          double r;
          if (P.rng_mode == 0) {
            r = host_draw(D, base + __popcll(bal & lt));
          } else {
            double t;
            philox_uniform2(P.seed, 3, (unsigned long long)step, (unsigned long long)D.near_list[q], r, t);
          }
          if (s.flag == 0) { E_ = P.EC[0]; if (P.RPfg > r) s.flag = 1; }
          else { E_ = P.EC[1]; if (P.RPgf > r) s.flag = 0; }
        } else {
          E_ = P.EC[3];
        }
        double F, U;
        lj_terms(r2, P.R6[0], P.R12[0], P.r12R12[0], P.r6R6[0], F, U);
        const double EF = E_ * F;
        cx = EF * d1; cy = EF * d2; cz = EF * d3;
        pot = E_ * U;
        fx = cx; fy = cy; fz = cz;
      }
      if (active(s.flag)) {
        acc[0] += cx; acc[1] += cy; acc[2] += cz;
        acc[3] += nx; acc[4] += ny; acc[5] += nz;
        acc[6] += pot;
        // kick (:255-269) with the flag after the conversions
        const double h = P.hk[s.flag];
        s.vx = s.vx + (pfx + fx) * h;
        s.vy = s.vy + (pfy + fy) * h;
        s.vz = s.vz + (pfz + fz) * h;
      }
      s.fx = fx; s.fy = fy; s.fz = fz;
      if (!last) sk_move(P, s);
      slot_store(D, q, s);
    }
    base += __popcll(bal);
  }
#pragma unroll
  for (int v = 0; v < 7; v++) {
    double r = wave_sum(acc[v]);
    if (lane == 0) g_sk.red[wv][v] = r;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    double a = g_sk.red[0][threadIdx.x];
    for (int w = 1; w < SK_W; w++) a = a + g_sk.red[w][threadIdx.x];
    g_sk.fin[threadIdx.x] = a;
  }
  __syncthreads();
}

// gather the near set from the particle arrays, and scatter it back
__device__ __forceinline__ void sk_gather(const SkD& D, int nn) {
  for (int q = threadIdx.x; q < nn; q += SK_T) {
    const int p = D.near_list[q];
    SkSlot s;
    s.x = D.x[p]; s.y = D.y[p]; s.z = D.z[p]; s.vx = D.vx[p]; s.vy = D.vy[p]; s.vz = D.vz[p];
    s.fx = D.fx[p]; s.fy = D.fy[p]; s.fz = D.fz[p]; s.flag = D.flag[p];
    slot_store(D, q, s);
    D.ns_cflag[q] = (unsigned char)s.flag;
  }
}
__device__ __forceinline__ void sk_scatter(const SkD& D, int nn) {
  for (int q = threadIdx.x; q < nn; q += SK_T) {
    const int p = D.near_list[q];
    SkSlot s = slot_load(D, q);
    D.x[p] = s.x; D.y[p] = s.y; D.z[p] = s.z; D.vx[p] = s.vx; D.vy[p] = s.vy; D.vz[p] = s.vz;
    D.fx[p] = s.fx; D.fy[p] = s.fy; D.fz[p] = s.fz; D.flag[p] = s.flag;
    const bool force = s.fx != 0.0 || s.fy != 0.0 || s.fz != 0.0;
    D.cflag[p] = (unsigned char)(D.ns_cflag[q] | (force ? CF_FORCE : 0));
  }
}

// the motion of the first step of a launch over the near set
__device__ __forceinline__ void sk_first_motion(const SkP& P, const SkD& D, int nn) {
  for (int q = threadIdx.x; q < nn; q += SK_T) {
    SkSlot s = slot_load(D, q);
    sk_move(P, s);
    slot_store(D, q, s);
  }
}

// mode 0: steps 1..S of an interval (motion of step 1 first);
// SM_STOP: steps 1..S-1, without the motion of step S (a sort comes first);
// SM_RESUME: the motion and the step S.
__global__ __launch_bounds__(SK_T) void sk_near_engine(SkP P, SkD D, int mode) {
  const int tid = threadIdx.x;
  SkScalars* sc = D.sc;
  if (tid == 0) {
    g_sk.P = P;
    g_sk.D = D;
    g_sk.nn = sc->near_num;
    g_sk.cursor = sc->cursor;
    g_sk.cstep = sc->cstep;
    for (int f = 0; f < 12; f++) { g_sk.sd[f][0] = D.dim[f * 2]; g_sk.sd[f][1] = D.dim[f * 2 + 1]; }
    if (!(mode & SM_RESUME))
      for (int s = 0; s < 2; s++) { sc->x0[s] = D.dim[SX * 2 + s]; sc->y0[s] = D.dim[SY * 2 + s]; sc->z0[s] = D.dim[SZ * 2 + s]; }
  }
  __syncthreads();
  const int nn = g_sk.nn;
  if (nn > SK_NEAR) {
    if (tid == 0) raise_error(&sc->err, ERR_NEAR_CAP);
    return;
  }
  unsigned long long T0 = __builtin_amdgcn_s_memrealtime();
  sk_gather(D, nn);
  __threadfence_block();
  __syncthreads();
  // motion of the launch's first step, solvent and dimer
  sk_first_motion(P, D, nn);
  if (tid == 0) {
    g_sk.cstep += 1;
    sk_begin_motion();
  }
  __threadfence_block();
  __syncthreads();
  const int t0 = (mode & SM_RESUME) ? P.S : 1;
  const int t1 = (mode & SM_STOP) ? P.S - 1 : P.S;
  for (int t = t0; t <= t1; t++) {
    const bool last = (t == t1);
    if (P.rng_mode == 0) sk_count_draws();
    sk_step(last, t == P.S);
    if (tid == 0) {
      if (P.rng_mode == 0) g_sk.cursor += g_sk.wsum[SK_W];
      sk_end_forces();
      if (!last) {
        g_sk.cstep += 1;
        sk_begin_motion();
      }
    }
    __threadfence_block();
    __syncthreads();
  }
  sk_scatter(D, nn);
  if (tid == 0) {
    sc->cstep = g_sk.cstep;
    sc->cursor = g_sk.cursor;
    sc->bond[0] = g_sk.bond[0]; sc->bond[1] = g_sk.bond[1]; sc->bond[2] = g_sk.bond[2];
    for (int f = 0; f < 12; f++) { D.dim[f * 2] = g_sk.sd[f][0]; D.dim[f * 2 + 1] = g_sk.sd[f][1]; }
    sc->prof[0] += __builtin_amdgcn_s_memrealtime() - T0;
    sc->prof[5] = nn;
    sc->prof[7] += 1;
  }
}

// ------------------------------------------------------------- bulk kernels
// Far particles: nmove motions "x + v*dt (+ 0*h), wrapped" (they carry no
// force); then, with do_bin, the census of Distribute_SelkovDimer.f90:12-35
// (counts and lowest index per species and cell, with the flag each particle
// had before the step's conversions), the shifted cell of
// RMPC_SelkovDimer.f90:22-46 counted by cell and rank, and the distance byte
// for the next near set.
__device__ __forceinline__ void sk_bin_one(const SkP& P, const SkD& D, int i, bool valid, bool near, int f,
                                           double x, double y, double z, double xs, double ys, double zs,
                                           const double (*sph)[2], bool asc);
// far: the far particles only, beside the near engine (sk_bin_near bins the
// near set afterwards), distances to the spheres at the interval's start
__global__ __launch_bounds__(256) void sk_bulk(SkP P, SkD D, int nmove, int do_bin, int far) {
  const int NT = P.space;
  double xs = 0, ys = 0, zs = 0;
  if (do_bin) {
    double u0, u1, u2;
    if (P.rng_mode == 0) {
      const long long c = D.sc->cursor;
      u0 = host_draw(D, c); u1 = host_draw(D, c + 1); u2 = host_draw(D, c + 2);
    } else {
      double t;
      philox_uniform2(P.seed, 4, D.coll, 0, u0, u1);
      philox_uniform2(P.seed, 4, D.coll, 1, u2, t);
    }
    xs = P.A0 * (u0 - 0.5); ys = P.A0 * (u1 - 0.5); zs = P.A0 * (u2 - 0.5);
  }
  __shared__ double sph[3][2];
  if (threadIdx.x < 2) {
    sph[0][threadIdx.x] = far ? D.sc->xs0[threadIdx.x] : D.dim[SX * 2 + threadIdx.x];
    sph[1][threadIdx.x] = far ? D.sc->ys0[threadIdx.x] : D.dim[SY * 2 + threadIdx.x];
    sph[2][threadIdx.x] = far ? D.sc->zs0[threadIdx.x] : D.dim[SZ * 2 + threadIdx.x];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < NT; b0 += stride) {
    const int i = (int)(b0 + threadIdx.x);
    const bool valid = i < NT;
    bool near = false;
    int f = 2;
    double x = 0, y = 0, z = 0;
    if (valid) {
      near = (D.near_bits[i >> 5] >> (i & 31)) & 1u;
      f = D.flag[i];
      x = D.x[i]; y = D.y[i]; z = D.z[i];
      if (!near && active(f) && nmove > 0) {
        const double vx = D.vx[i], vy = D.vy[i], vz = D.vz[i];
        const double h = P.h2[f];
        for (int m = 0; m < nmove; m++) {
          x = wrap((x + vx * P.MDt) + 0.0 * h, P.box[0]);
          y = wrap((y + vy * P.MDt) + 0.0 * h, P.box[1]);
          z = wrap((z + vz * P.MDt) + 0.0 * h, P.box[2]);
        }
        D.x[i] = x; D.y[i] = y; D.z[i] = z;
      }
    }
    if (!do_bin) continue;
    sk_bin_one(P, D, i, valid && !(far && near), near, f, x, y, z, xs, ys, zs, sph, true);
  }
}
// distance beyond the nearest cutoff in 1/32 units, rounded down, capped at 255
__device__ __forceinline__ unsigned char sk_qdist(const SkP& P, double x, double y, double z,
                                                  const double (*sph)[2]) {
  double dmin = 1e30;
  for (int s = 0; s < 2; s++) {
    double a = pbc_h(x - sph[0][s], P.hbox[0], P.box[0]);
    double b = pbc_h(y - sph[1][s], P.hbox[1], P.box[1]);
    double cz = pbc_h(z - sph[2][s], P.hbox[2], P.box[2]);
    dmin = fmin(dmin, sqrt(a * a + b * b + cz * cz) - (s == 0 ? P.cutC : P.cutN));
  }
  const double qd = floor(fmax(dmin, 0.0) * 32.0 * (1.0 - 1e-12));
  return (unsigned char)fmin(qd, 255.0);
}
// Binning of one particle per lane (whole waves call it together; `valid`
// false for idle lanes): shifted cell and rank, census with the flag before
// the collision step's conversions, distance byte.  asc: the wave's lanes
// hold ascending indices, so a group's leader holds its lowest index.
__device__ __forceinline__ void sk_bin_one(const SkP& P, const SkD& D, int i, bool valid, bool near, int f,
                                           double x, double y, double z, double xs, double ys, double zs,
                                           const double (*sph)[2], bool asc) {
  const int lane = threadIdx.x & 63;
  {
    const bool act = valid && active(f);
    // shifted cell (all three directions periodic)
    bool ok = act;
    int c = 0;
    if (act) {
      double X = wrap(x + xs, P.box[0]), Y = wrap(y + ys, P.box[1]), Z = wrap(z + zs, P.box[2]);
      long long cc = sk_cell0(P, X, Y, Z);
      if (cc < 0 || cc >= P.vol) { raise_error(&D.sc->err, ERR_CELL_RANGE); ok = false; } else c = (int)cc;
    }
    int r = wave_atomic_rank(D.cnt, c, ok);
    if (valid) {
      if (ok && r >= SK_BMAX) raise_error(&D.sc->err, ERR_BUCKET);
      D.cell_of[i] = ok ? c : -1;
      D.rank_of[i] = ok ? (unsigned short)min(r, 65535) : (unsigned short)0xffff;
    }
    // census with the flag before the collision step's conversions
    unsigned char cf = (unsigned char)f;
    if (valid && near) cf = D.cflag[i];
    const int sp = valid ? species(cf & 0x7f) : -1;
    bool cen = false;
    int ck = 0;
    if (sp >= 0) {
      long long u = sk_cell0(P, x, y, z);
      if (u < 0 || u >= P.vol) raise_error(&D.sc->err, ERR_CELL_RANGE);
      else { cen = true; ck = (int)(sp * P.vol + u); }
    }
    {
      unsigned long long todo = __ballot(cen);
      int gsize = 0;
      unsigned gmin = 0;
      bool leader = false;
      while (todo) {
        const int ld = __ffsll((unsigned long long)todo) - 1;
        const int k = wave_bcast(ck, ld);
        const unsigned long long grp = __ballot(cen && ck == k);
        if (lane == ld) { leader = true; gsize = __popcll(grp); gmin = (unsigned)i; }
        todo &= ~grp;
      }
      if (leader) {
        int old = atomicAdd(&D.tcount[ck], gsize);
        if (old + gsize > SK_MAXS) raise_error(&D.sc->err, ERR_MEMBERS);
        if (asc) atomicMin(&D.first[ck], gmin);
      }
      if (!asc && cen) atomicMin(&D.first[ck], (unsigned)i);
    }
    // distance beyond the nearest cutoff, 1/32 units (0 for particles carrying a force)
    if (valid && act) D.dq[i] = (near && (D.cflag[i] & CF_FORCE)) ? (unsigned char)0 : sk_qdist(P, x, y, z, sph);
  }
}
// the near set once the near engine is done (beside it, sk_bulk binned the
// far particles): binning with distances to the spheres where they ended
__global__ __launch_bounds__(256) void sk_bin_near(SkP P, SkD D) {
  double u0, u1, u2, t;
  philox_uniform2(P.seed, 4, D.coll, 0, u0, u1);   // the shift of sk_bulk (device draws)
  philox_uniform2(P.seed, 4, D.coll, 1, u2, t);
  const double xs = P.A0 * (u0 - 0.5), ys = P.A0 * (u1 - 0.5), zs = P.A0 * (u2 - 0.5);
  __shared__ double sph[3][2];
  if (threadIdx.x < 2) {
    sph[0][threadIdx.x] = D.dim[SX * 2 + threadIdx.x];
    sph[1][threadIdx.x] = D.dim[SY * 2 + threadIdx.x];
    sph[2][threadIdx.x] = D.dim[SZ * 2 + threadIdx.x];
  }
  __syncthreads();
  const long long nn = min(D.sc->near_num, D.near_cap);
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < nn; b0 += (long long)gridDim.x * blockDim.x) {
    const long long q = b0 + threadIdx.x;
    const bool valid = q < nn;
    int i = 0, f = 2;
    double x = 0, y = 0, z = 0;
    if (valid) {
      i = D.near_list[q];
      f = D.flag[i];
      x = D.x[i]; y = D.y[i]; z = D.z[i];
    }
    sk_bin_one(P, D, i, valid, true, f, x, y, z, xs, ys, zs, sph, false);
  }
}

__global__ __launch_bounds__(256) void sk_scatter_cells(SkP P, SkD D) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.space; i += gridDim.x * blockDim.x) {
    const int r = D.rank_of[i];
    if (r < SK_BMAX) D.bucket[D.off[D.cell_of[i]] + r] = i;
  }
}

// Affected_Cell.f90:5-99: the cells around each sphere take no reaction
__global__ __launch_bounds__(256) void sk_affected(SkP P, SkD D) {
  for (int sph = 0; sph < 2; sph++) {
    const double cut = sph == 0 ? P.cutC : P.cutN, cut2 = sph == 0 ? P.cutC2 : P.cutN2;
    const double X = D.dim[SX * 2 + sph], Y = D.dim[SY * 2 + sph], Z = D.dim[SZ * 2 + sph];
    const long long lx = (long long)ceil(X - (cut + 1.0)), ux = (long long)floor(X + (cut + 1.0));
    const long long ly = (long long)ceil(Y - (cut + 1.0)), uy = (long long)floor(Y + (cut + 1.0));
    const long long lz = (long long)ceil(Z - (cut + 1.0)), uz = (long long)floor(Z + (cut + 1.0));
    const long long nx = ux - lx + 1, ny = uy - ly + 1, nz = uz - lz + 1;
    for (long long t = threadIdx.x; t < nx * ny * nz; t += blockDim.x) {
      const long long i = lx + t % nx, j = ly + (t / nx) % ny, k = lz + t / (nx * ny);
      // REAL(i) is single precision; the difference and "- 1." are double
      const double a = fabs((double)(float)i - X) - 1.0;
      const double b = fabs((double)(float)j - Y) - 1.0;
      const double c = fabs((double)(float)k - Z) - 1.0;
      if (a * a + b * b + c * c < cut2) {
        long long cx = i, cy = j, cz = k;
        if (cx < 0) cx += P.Lx; else if (cx >= P.Lx) cx -= P.Lx;
        if (cy < 0) cy += P.Ly; else if (cy >= P.Ly) cy -= P.Ly;
        if (cz < 0) cz += P.Lz; else if (cz >= P.Lz) cz -= P.Lz;
        const long long cn = cx + (long long)P.Lx * cy + (long long)P.Lx * P.Ly * cz;
        if (cn >= 0 && cn < P.vol) D.unaff[cn] = 0;
        else raise_error(&D.sc->err, ERR_CELL_RANGE);
      }
    }
  }
}

// Per cell c: members of shifted cell c in ascending index, centre-of-mass
// velocity (RMPC_SelkovDimer.f90:31-62), random axis (:67-75); and the
// reaction of unshifted cell c unless it is affected (:127-324).
__global__ __launch_bounds__(CT) void sk_collide(SkP P, SkD D) {
  constexpr int LB = 32;
  __shared__ int sidx[LB * CT];
  const long long c = (long long)blockIdx.x * CT + threadIdx.x;
  if (c >= P.vol) return;
  const long long base = D.sc->cursor;
  int n = D.cnt[c];
  if (n > SK_BMAX) n = SK_BMAX;   // ERR_BUCKET already raised
  const int* b = D.bucket + D.off[c];
  int* idx;
  int istride;
  if (n <= LB) {
    idx = sidx + threadIdx.x;
    istride = CT;
    for (int j = 0; j < n; j++) {
      int v = b[j], k = j;
      while (k > 0 && idx[(k - 1) * CT] > v) { idx[k * CT] = idx[(k - 1) * CT]; k--; }
      idx[k * CT] = v;
    }
  } else {
    idx = (int*)b;
    istride = 1;
    for (int j = 1; j < n; j++) {
      int v = idx[j], k = j;
      while (k > 0 && idx[k - 1] > v) { idx[k] = idx[k - 1]; k--; }
      idx[k] = v;
    }
  }
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int j = 0; j < n; j++) {
    const int i = idx[j * istride];
    sx = sx + D.vx[i]; sy = sy + D.vy[i]; sz = sz + D.vz[i];
  }
  if (n != 0) {
    const double it = 1.0 / (double)n;
    sx = sx * it; sy = sy * it; sz = sz * it;
  }
  double u1, u2;
  if (P.rng_mode == 0) {
    u1 = host_draw(D, base + 3 + 2 * c);
    u2 = host_draw(D, base + 4 + 2 * c);
  } else {
    philox_uniform2(P.seed, 5, D.coll, (unsigned long long)c, u1, u2);
  }
  const double phi = P.twopi * u1, th = 2.0 * u2 - 1.0, st = sqrt(1.0 - th * th);
  double2* rec = (double2*)(D.cellrec + c * 8);
  rec[0] = make_double2(sx, sy);
  double sp, cp;
  sincos(phi, &sp, &cp);   // one reduction; the same values as sin() and cos()
  rec[1] = make_double2(sz, st * cp);
  rec[2] = make_double2(st * sp, th);
  rec[3] = make_double2((double)n, 0.0);

  if (!D.unaff[c]) return;
  const long long tF = D.tcount[c], tG = D.tcount[P.vol + c], tI = D.tcount[2 * P.vol + c];
  double a1 = P.k[0], a2 = P.k[1] * (double)tF, a5 = P.k[4] * (double)tG, a6 = P.k[5], a3, a4;
  if (tG == 0) a3 = 0.0; else a3 = ((P.k[2] * (double)tF) * (double)tG) * (double)(tG - 1);
  if (tG <= 1) a4 = 0.0; else a4 = ((P.k[3] * (double)tG) * (double)(tG - 1)) * (double)(tG - 2);
  if (tI == 0) { a1 = 0.0; a6 = 0.0; }
  const double a00 = ((((a1 + a2) + a3) + a4) + a5) + a6;
  const double Pr = D.ptab[((tI > 0 ? 1 : 0) * 201 + min(tF, (long long)SK_MAXS)) * 201 + min(tG, (long long)SK_MAXS)];
  double x1, x2;
  if (P.rng_mode == 0) {
    const long long r0 = base + 3 + 2 * P.vol + 2 * (long long)D.uoff[c];
    x1 = host_draw(D, r0);
    x2 = host_draw(D, r0 + 1);
  } else {
    philox_uniform2(P.seed, 6, D.coll, (unsigned long long)c, x1, x2);
  }
  const double ar = a00 * x1;
  if (!(Pr > x2)) return;
  const unsigned fF = D.first[c], fG = D.first[P.vol + c], fI = D.first[2 * P.vol + c];
  unsigned tgt = 0xffffffffu;
  int val = 0;
  // the decision tree of RMPC_SelkovDimer.f90:140-322: (count case, propensity interval) -> (member, new flag)
  if (tI > 0) {
    if (tF == 0) {
      if (tG == 0) {
        if (ar < a1) { tgt = fI; val = 0; }
        else if (ar >= a1 && ar < a00) { tgt = fI; val = 1; }
      } else if (tG <= 2) {
        if (ar < a1) { tgt = fI; val = 0; }
        else if (ar >= a1 && ar < a1 + a5) { tgt = fG; val = 3; }
        else if (ar >= a1 + a5 && ar < a00) { tgt = fI; val = 1; }
      } else {
        if (ar < a1) { tgt = fI; val = 0; }
        else if (ar >= a1 && ar < a1 + a4) { tgt = fG; val = 0; }
        else if (ar >= a1 + a4 && ar < (a1 + a4) + a5) { tgt = fG; val = 3; }
        else if (ar >= (a1 + a4) + a5 && ar < a00) { tgt = fI; val = 1; }
      }
    } else {
      const double s12 = a1 + a2, s123 = s12 + a3;
      if (tG == 0) {
        if (ar < a1) { tgt = fI; val = 0; }
        else if (ar >= a1 && ar < s12) { tgt = fF; val = 3; }
        else if (ar >= s12 && ar < a00) { tgt = fI; val = 1; }
      } else if (tG == 1) {
        if (ar < a1) { tgt = fI; val = 0; }
        else if (ar >= a1 && ar < s12) { tgt = fF; val = 3; }
        else if (ar >= s12 && ar < s12 + a5) { tgt = fG; val = 3; }
        else if (ar >= s12 + a5 && ar < a00) { tgt = fI; val = 1; }
      } else if (tG == 2) {
        if (ar < a1) { tgt = fI; val = 0; }
        else if (ar >= a1 && ar < s12) { tgt = fF; val = 3; }
        else if (ar >= s12 && ar < s123) { tgt = fF; val = 1; }
        else if (ar >= s123 && ar < s123 + a5) { tgt = fG; val = 3; }
        else if (ar >= s123 + a5 && ar < a00) { tgt = fI; val = 1; }
      } else {
        const double s1234 = s123 + a4;
        if (ar < a1) { tgt = fI; val = 0; }
        else if (ar >= a1 && ar < s12) { tgt = fF; val = 3; }
        else if (ar >= s12 && ar < s123) { tgt = fF; val = 1; }
        else if (ar >= s123 && ar < s1234) { tgt = fG; val = 0; }
        else if (ar >= s1234 && ar < s1234 + a5) { tgt = fG; val = 3; }
        else if (ar >= s1234 + a5 && ar < a00) { tgt = fI; val = 1; }
      }
    }
  } else {
    if (tF == 0) {
      if (tG == 0) {
      } else if (tG <= 2) {
        if (ar >= a1 && ar < a00) { tgt = fG; val = 3; }
      } else {
        if (ar >= a1 && ar < a1 + a4) { tgt = fG; val = 0; }
        else if (ar >= a1 + a4 && ar < a00) { tgt = fG; val = 3; }
      }
    } else {
      const double s12 = a1 + a2, s123 = s12 + a3;
      if (tG == 0) {
        if (ar >= a1 && ar < a00) { tgt = fF; val = 3; }
      } else if (tG == 1) {
        if (ar >= a1 && ar < s12) { tgt = fF; val = 3; }
        else if (ar >= s12 && ar < a00) { tgt = fG; val = 3; }
      } else if (tG == 2) {
        if (ar >= a1 && ar < s12) { tgt = fF; val = 3; }
        else if (ar >= s12 && ar < s123) { tgt = fF; val = 1; }
        else if (ar >= s123 && ar < a00) { tgt = fG; val = 3; }
      } else {
        if (ar >= a1 && ar < s12) { tgt = fF; val = 3; }
        else if (ar >= s12 && ar < s123) { tgt = fF; val = 1; }
        else if (ar >= s123 && ar < s123 + a4) { tgt = fG; val = 0; }
        else if (ar >= s123 + a4 && ar < a00) { tgt = fG; val = 3; }
      }
    }
  }
  if (tgt < (unsigned)P.space) D.flag[tgt] = val;
  else if (tgt != 0xffffffffu) raise_error(&D.sc->err, ERR_CELL_RANGE);
}

// draws of the collision: 3 shift + 2 per cell + 2 per unaffected cell
// before the rotation: the draw cursor advanced past this collision's draws,
// and the near bitmap (nwords words) cleared for the next near set
__global__ __launch_bounds__(256) void sk_rotate_begin(SkP P, SkD D, long long nwords) {
  const long long i0 = (long long)blockIdx.x * 256 + threadIdx.x, st = (long long)gridDim.x * 256;
  for (long long i = i0; i < nwords; i += st) D.near_bits[i] = 0u;
  if (i0 == 0 && P.rng_mode == 0) {
    const long long used = 3 + 2 * P.vol + 2 * (long long)D.uoff[P.vol];
    if (D.sc->cursor + used > D.nU) raise_error(&D.sc->err, ERR_DRAWS);
    D.sc->cursor += used;
  }
}
// interval start: cleared cell counts, census and first keys, every cell
// unaffected until sk_affected marks the ones near a sphere
__global__ __launch_bounds__(256) void sk_interval_begin(SkD D, long long V) {
  const long long i0 = (long long)blockIdx.x * 256 + threadIdx.x, st = (long long)gridDim.x * 256;
  if (i0 < 2) {   // the spheres' start positions, for the far particles' distances
    D.sc->xs0[i0] = D.dim[SX * 2 + i0]; D.sc->ys0[i0] = D.dim[SY * 2 + i0]; D.sc->zs0[i0] = D.dim[SZ * 2 + i0];
  }
  for (long long i = i0; i < V; i += st) { D.cnt[i] = 0; D.unaff[i] = 1; }
  for (long long i = i0; i < 3 * V; i += st) { D.tcount[i] = 0; D.first[i] = 0xffffffffu; }
}

// Rotation of RMPC_SelkovDimer.f90:79-121 (projection, cos-rotated
// orthogonal part, sin-weighted cross product) and the next near set: within
// |v| tau + guard of a cutoff by the distance byte, or carrying a force.
// overlap: the far particles' distance bytes were measured to the spheres at
// the interval's start (the near set's to where they ended); a far particle
// within reach + the spheres' largest displacement + 1/16 of that measure is
// measured again to where the spheres ended, so the near set is the one of
// the one-stream order.  The near bitmap still holds this interval's set; each
// wave reads its 64 bits and writes them whole.
__global__ __launch_bounds__(256) void sk_rotate(SkP P, SkD D, int overlap) {
  const double tau = P.tau * (1.0 + 1e-9);
  __shared__ double sph[3][2];
  __shared__ double shift_s;
  if (overlap) {
    if (threadIdx.x < 2) {
      sph[0][threadIdx.x] = D.dim[SX * 2 + threadIdx.x];
      sph[1][threadIdx.x] = D.dim[SY * 2 + threadIdx.x];
      sph[2][threadIdx.x] = D.dim[SZ * 2 + threadIdx.x];
    }
    if (threadIdx.x == 0) {
      double sh = 0.0;
      for (int s = 0; s < 2; s++) {
        const double a = pbc_h(D.dim[SX * 2 + s] - D.sc->xs0[s], P.hbox[0], P.box[0]);
        const double b = pbc_h(D.dim[SY * 2 + s] - D.sc->ys0[s], P.hbox[1], P.box[1]);
        const double c = pbc_h(D.dim[SZ * 2 + s] - D.sc->zs0[s], P.hbox[2], P.box[2]);
        sh = fmax(sh, sqrt(a * a + b * b + c * c));
      }
      shift_s = sh;
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < P.space; b0 += stride) {
    const int i = (int)(b0 + threadIdx.x);
    const int c = i < P.space ? D.cell_of[i] : -1;
    bool nb = false;
    if (c >= 0) {
      const double2* rec = (const double2*)(D.cellrec + (long long)c * 8);
      const double2 r0 = rec[0], r1 = rec[1], r2 = rec[2];
      const double cx = r0.x, cy = r0.y, cz = r1.x, ux = r1.y, uy = r2.x, uz = r2.y;
      const double dx = D.vx[i] - cx, dy = D.vy[i] - cy, dz = D.vz[i] - cz;
      const double pr = (dx * ux + dy * uy) + dz * uz;
      const double px = pr * ux, py = pr * uy, pz = pr * uz;
      const double ox = dx - px, oy = dy - py, oz = dz - pz;
      const double o2x = (oy * uz) - (oz * uy), o2y = (oz * ux) - (ox * uz), o2z = (ox * uy) - (oy * ux);
      const double vx = cx + ((px + ox * P.cosrot) + o2x * P.sinrot);
      const double vy = cy + ((py + oy * P.cosrot) + o2y * P.sinrot);
      const double vz = cz + ((pz + oz * P.cosrot) + o2z * P.sinrot);
      D.vx[i] = vx; D.vy[i] = vy; D.vz[i] = vz;
      const double reach = sqrt((vx * vx + vy * vy) + vz * vz) * tau + P.guard + 1e-6;
      unsigned char q = D.dq[i];
      if (overlap) {
        const bool was = (D.near_bits[i >> 5] >> (i & 31)) & 1u;
        if (!was && (double)q * (1.0 / 32.0) <= reach + shift_s + 0.0625) q = sk_qdist(P, D.x[i], D.y[i], D.z[i], sph);
      }
      nb = (double)q * (1.0 / 32.0) <= reach;
      if (!overlap && nb) atomicOr(&D.near_bits[i >> 5], 1u << (i & 31));
    }
    if (overlap) {   // lanes 0 and 32 write the wave's two words; words past the last particle
                     // ((space + 31) / 32 of them hold indices below space) are left alone
      const unsigned long long m = __ballot(nb);
      if ((lane == 0 || lane == 32) && (long long)(i >> 5) < (P.space + 31) / 32)
        D.near_bits[i >> 5] = (unsigned)(m >> lane);
    }
  }
}

// near set from positions (set-up): within |v| tau + guard of a cutoff, or carrying a force
__global__ __launch_bounds__(256) void sk_classify_all(SkP P, SkD D) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.space; i += gridDim.x * blockDim.x) {
    if (!active(D.flag[i])) continue;
    const double x = D.x[i], y = D.y[i], z = D.z[i], vx = D.vx[i], vy = D.vy[i], vz = D.vz[i];
    const double reach = sqrt((vx * vx + vy * vy) + vz * vz) * P.tau * (1.0 + 1e-9) + P.guard + 1e-6;
    bool near = D.fx[i] != 0.0 || D.fy[i] != 0.0 || D.fz[i] != 0.0;
    for (int s = 0; s < 2; s++) {
      double a = pbc_h(x - D.dim[SX * 2 + s], P.hbox[0], P.box[0]);
      double b = pbc_h(y - D.dim[SY * 2 + s], P.hbox[1], P.box[1]);
      double c = pbc_h(z - D.dim[SZ * 2 + s], P.hbox[2], P.box[2]);
      double r = (s == 0 ? P.cutC : P.cutN) + reach;
      if (a * a + b * b + c * c <= r * r) near = true;
    }
    if (near) atomicOr(&D.near_bits[i >> 5], 1u << (i & 31));
  }
}

// ------------------------------------------------------------------ Sort.f90
// key = 3*cell + species (F, G, I), cells ascending; inactive particles drop out
__global__ __launch_bounds__(256) void sk_sort_bin(SkP P, SkD D, int* scnt, int* key, unsigned short* rk) {
  const int lane = threadIdx.x & 63;
  (void)lane;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < P.space; b0 += stride) {
    const int i = (int)(b0 + threadIdx.x);
    const bool valid = i < P.space;
    int k = 0;
    bool ok = false;
    if (valid) {
      const int sp = species(D.flag[i]);
      if (sp >= 0) {
        long long u = sk_cell0(P, D.x[i], D.y[i], D.z[i]);
        if (u < 0 || u >= P.vol) raise_error(&D.sc->err, ERR_CELL_RANGE);
        else { ok = true; k = (int)(3 * u + sp); }
      }
    }
    const int r = wave_atomic_rank(scnt, k, ok);
    if (valid) {
      if (ok && r >= SK_MAXS) raise_error(&D.sc->err, ERR_MEMBERS);
      key[i] = ok ? k : -1;
      rk[i] = ok ? (unsigned short)min(r, 65535) : (unsigned short)0xffff;
    }
  }
}
__global__ __launch_bounds__(256) void sk_sort_scatter(int n, const int* key, const unsigned short* rk,
                                                       const int* soff, int* sb) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (key[i] >= 0 && rk[i] < SK_MAXS) sb[soff[key[i]] + rk[i]] = i;
}
// new slot j < NTnew takes particle sb[j]; slots past NTnew are cleared to FLAG 2
__global__ __launch_bounds__(256) void sk_sort_gather(SkP P, SkD D, const int* sb, const int* soff, long long nkeys,
                                                      double* const* alt, int* aflag, unsigned char* acflag,
                                                      unsigned* abits) {
  const int lane = threadIdx.x & 63;
  const int NTn = soff[nkeys];
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long b0 = (long long)blockIdx.x * blockDim.x; b0 < P.space; b0 += stride) {
    const long long j = b0 + threadIdx.x;
    bool bit = false;
    if (j < P.space) {
      if (j < NTn) {
        const int i = sb[j];
        alt[0][j] = D.x[i]; alt[1][j] = D.y[i]; alt[2][j] = D.z[i];
        alt[3][j] = D.vx[i]; alt[4][j] = D.vy[i]; alt[5][j] = D.vz[i];
        alt[6][j] = D.fx[i]; alt[7][j] = D.fy[i]; alt[8][j] = D.fz[i];
        aflag[j] = D.flag[i];
        acflag[j] = D.cflag[i];
        bit = (D.near_bits[i >> 5] >> (i & 31)) & 1u;
      } else {
        for (int a = 0; a < 9; a++) alt[a][j] = 0.0;
        aflag[j] = 2;
        acflag[j] = 2;
      }
    }
    const unsigned long long m = __ballot(bit);
    const long long w0 = (b0 + (threadIdx.x & ~63)) >> 5, nw = (P.space + 31) / 32;   // words holding slots
    if (lane == 0 && w0 < nw) abits[w0] = (unsigned)m;
    if (lane == 32 && w0 + 1 < nw) abits[w0 + 1] = (unsigned)(m >> 32);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) D.sc->NT = NTn;
}

}  // namespace sk
}  // namespace mpcx

// =====================================================================================
// Host side
// =====================================================================================
using namespace mpcx;
using namespace mpcx::sk;

struct mpcx_sk {
  mpcx_sk_config cfg;
  SkP P;
  SkD D;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // with device draws, sk_bulk streams and bins the far particles on st2
  // beside the near engine (st2 waits for fork, st for join)
  hipStream_t st2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  bool overlap = true;
  // near engine, sk_bulk, sk_collide, sk_rotate: every TSTRIDE-th interval is
  // timed (each event costs the stream a few us), at most TMAX of them
  static constexpr int NKT = 4, TMAX = 256, TSTRIDE = 8;
  std::vector<hipEvent_t> tev;
  std::vector<char> tev_on;
  double kt_ms[NKT] = {0};
  double run_ms = 0;
  long long nwords = 0;
  double* d_U = nullptr;
  long long nU_cap = 0;
  int *d_scnt = nullptr, *d_soff = nullptr, *d_sb = nullptr, *d_skey = nullptr;
  unsigned short* d_srk = nullptr;
  double* alt[9] = {nullptr};
  double** d_alt = nullptr;
  int* alt_flag = nullptr;
  unsigned char* alt_cflag = nullptr;
  unsigned* alt_bits = nullptr;
  std::vector<void*> allocs;
};

#define SKCHK(x)                                                              \
  do {                                                                        \
    hipError_t _r = (x);                                                      \
    if (_r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(_r)); \
  } while (0)

template <typename T>
static T* skalloc(mpcx_sk* e, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, n * sizeof(T) + 64) != hipSuccess) return nullptr;
  e->allocs.push_back(p);
  return (T*)p;
}

static int sk_check_error(mpcx_sk* e) {
  int err = 0;
  SKCHK(hipMemcpy(&err, &e->D.sc->err, sizeof(int), hipMemcpyDeviceToHost));
  if (!err) return 0;
  static const char* msg[] = {"none", "a particle maps outside the cell arrays",
                              "a cell holds more members of one species than the reference arrays (200)", "",
                              "a sphere moved further than the near/far guard within one interval", "",
                              "a collision cell holds more than 64 particles", "near set larger than its buffer",
                              "", "", "the dimer bond constraint did not converge", "",
                              "the host draw buffer ran out"};
  std::string m = std::string("device check failed: ") + (err > 0 && err <= 12 ? msg[err] : "unknown");
  int zero = 0;
  (void)hipMemcpy(&e->D.sc->err, &zero, sizeof(int), hipMemcpyHostToDevice);
  return set_error(MPCX_E_DEVICE, m.c_str());
}

extern "C" int mpcx_sk_destroy(mpcx_sk* e) {
  if (!e) return 0;
  handle_del(H_SK, e);
  if (e->st) (void)hipStreamSynchronize(e->st);
  if (e->st2) (void)hipStreamSynchronize(e->st2);
  for (void* p : e->allocs)
    if (p) (void)hipFree(p);
  if (e->st) (void)hipStreamDestroy(e->st);
  if (e->st2) (void)hipStreamDestroy(e->st2);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->join) (void)hipEventDestroy(e->join);
  if (e->e0) (void)hipEventDestroy(e->e0);
  if (e->e1) (void)hipEventDestroy(e->e1);
  for (hipEvent_t ev : e->tev)
    if (ev) (void)hipEventDestroy(ev);
  delete e;
  return 0;
}

extern "C" int mpcx_sk_create(mpcx_sk** out, const mpcx_sk_config* cfg) {
  if (!out || !cfg) return set_error(MPCX_E_ARG, "null argument");
  *out = nullptr;
  if (cfg->space < 1 || cfg->space > 0x7fffff00LL || cfg->NT < 0 || cfg->NT > cfg->space)
    return set_error(MPCX_E_ARG, "bad particle counts");
  if (cfg->MPCtstep < 1 || cfg->sort_step < 1 || cfg->sort_step % cfg->MPCtstep != 0)
    return set_error(MPCX_E_UNSUPPORTED, "sort_step must be a positive multiple of MPCtstep");
  for (int i = 0; i < 3; i++)
    if (cfg->box[i] < 4.0 || cfg->box[i] != floor(cfg->box[i])) return set_error(MPCX_E_ARG, "box edges must be whole cells");
  if (cfg->A0 != 1.0) return set_error(MPCX_E_UNSUPPORTED, "the reference's cell index assumes A0 = 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(MPCX_E_HIP, "no HIP device");
  mpcx_sk* e = new mpcx_sk();
  e->cfg = *cfg;
  SkP& P = e->P;
  memset(&P, 0, sizeof P);
  P.NT = (int)cfg->NT; P.space = (int)cfg->space; P.S = cfg->MPCtstep; P.sort_step = cfg->sort_step;
  P.rng_mode = cfg->rng_mode; P.seed = cfg->seed;
  P.Lx = (int)cfg->box[0]; P.Ly = (int)cfg->box[1]; P.Lz = (int)cfg->box[2];
  P.vol = (long long)P.Lx * P.Ly * P.Lz;
  for (int i = 0; i < 3; i++) { P.box[i] = cfg->box[i]; P.hbox[i] = 0.5 * cfg->box[i]; }
  const double MDt = cfg->MDt;
  P.MDt = MDt; P.IMDt = 1.0 / MDt; P.Dcn = cfg->Dcn;
  P.hc = (0.5 * MDt) / cfg->Mc; P.hn = (0.5 * MDt) / cfg->Mn;
  P.IMc = 1.0 / cfg->Mc; P.IMn = 1.0 / cfg->Mn; P.IIMcIMn = 1.0 / (P.IMc + P.IMn);
  const double M[4] = {cfg->Mf, cfg->Mg, 1.0, cfg->Mi};
  for (int f = 0; f < 4; f++) { P.h2[f] = ((0.5 * MDt) * MDt) / M[f]; P.hk[f] = (0.5 * MDt) / M[f]; }
  const double cut = pow(2.0, 1.0 / 6.0);
  P.cutC = cfg->Rc * cut; P.cutN = cfg->Rn * cut;
  P.cutC2 = P.cutC * P.cutC; P.cutN2 = P.cutN * P.cutN;
  const double R[2] = {cfg->Rc, cfg->Rn};
  for (int i = 0; i < 2; i++) {
    P.R6[i] = pow(R[i], 6); P.R12[i] = pow(R[i], 12);
    P.r6R6[i] = 6.0 * P.R6[i]; P.r12R12[i] = 12.0 * P.R12[i];
  }
  P.EC[0] = cfg->EfC; P.EC[1] = cfg->EgC; P.EC[3] = cfg->EiC;
  P.EN[0] = cfg->EfN; P.EN[1] = cfg->EgN; P.EN[3] = cfg->EiN;
  P.RPfg = cfg->RPfg; P.RPgf = cfg->RPgf; P.A0 = cfg->A0; P.MPCt = cfg->MPCt;
  for (int i = 0; i < 6; i++) P.k[i] = cfg->k[i];
  P.cosrot = cfg->cosrot; P.sinrot = cfg->sinrot;
  P.twopi = 2.0 * (4.0 * atan(1.0));
  P.tau = (double)cfg->MPCtstep * MDt;
  P.guard = cfg->sphere_guard > 0 ? cfg->sphere_guard : 1.0;

  const long long NS = cfg->space, V = P.vol;
  SkD& D = e->D;
  memset(&D, 0, sizeof D);
  D.x = skalloc<double>(e, NS); D.y = skalloc<double>(e, NS); D.z = skalloc<double>(e, NS);
  D.vx = skalloc<double>(e, NS); D.vy = skalloc<double>(e, NS); D.vz = skalloc<double>(e, NS);
  D.fx = skalloc<double>(e, NS); D.fy = skalloc<double>(e, NS); D.fz = skalloc<double>(e, NS);
  D.flag = skalloc<int>(e, NS);
  D.cflag = skalloc<unsigned char>(e, NS);
  D.dim = skalloc<double>(e, 24);
  D.sc = skalloc<SkScalars>(e, 1);
  e->nwords = (NS + 63) / 64 * 2;
  D.near_bits = skalloc<unsigned>(e, e->nwords + 2);
  D.near_cap = (int)std::min<long long>(NS, SK_NEAR);
  D.near_list = skalloc<int>(e, D.near_cap);
  D.ns_x = skalloc<double>(e, D.near_cap); D.ns_y = skalloc<double>(e, D.near_cap); D.ns_z = skalloc<double>(e, D.near_cap);
  D.ns_vx = skalloc<double>(e, D.near_cap); D.ns_vy = skalloc<double>(e, D.near_cap); D.ns_vz = skalloc<double>(e, D.near_cap);
  D.ns_fx = skalloc<double>(e, D.near_cap); D.ns_fy = skalloc<double>(e, D.near_cap); D.ns_fz = skalloc<double>(e, D.near_cap);
  D.ns_flag = skalloc<int>(e, D.near_cap);
  D.ns_cflag = skalloc<unsigned char>(e, D.near_cap);
  D.cnt = skalloc<int>(e, V); D.off = skalloc<int>(e, V + 1); D.bucket = skalloc<int>(e, NS);
  long long npart = std::max((3 * V + 4095) / 4096, (e->nwords + 1023) / 1024) + 2;
  D.partial = skalloc<int>(e, npart);
  D.cell_of = skalloc<int>(e, NS); D.rank_of = skalloc<unsigned short>(e, NS);
  D.cellrec = skalloc<double>(e, V * 8); D.dq = skalloc<unsigned char>(e, NS);
  D.tcount = skalloc<int>(e, 3 * V); D.first = skalloc<unsigned>(e, 3 * V);
  D.unaff = skalloc<int>(e, V); D.uoff = skalloc<int>(e, V + 1);
  double* ptab = skalloc<double>(e, 2 * 201 * 201);
  D.ptab = ptab;
  e->d_scnt = skalloc<int>(e, 3 * V); e->d_soff = skalloc<int>(e, 3 * V + 1);
  e->d_sb = skalloc<int>(e, NS); e->d_skey = skalloc<int>(e, NS); e->d_srk = skalloc<unsigned short>(e, NS);
  for (int a = 0; a < 9; a++) e->alt[a] = skalloc<double>(e, NS);
  e->d_alt = skalloc<double*>(e, 9);
  e->alt_flag = skalloc<int>(e, NS);
  e->alt_cflag = skalloc<unsigned char>(e, NS);
  e->alt_bits = skalloc<unsigned>(e, e->nwords + 2);
  if (std::find(e->allocs.begin(), e->allocs.end(), nullptr) != e->allocs.end()) {
    mpcx_sk_destroy(e);
    return set_error(MPCX_E_HIP, "device allocation failed");
  }
  SKCHK(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking));
  SKCHK(hipStreamCreateWithFlags(&e->st2, hipStreamNonBlocking));
  SKCHK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
  SKCHK(hipEventCreateWithFlags(&e->join, hipEventDisableTiming));
  SKCHK(hipEventCreate(&e->e0)); SKCHK(hipEventCreate(&e->e1));
  e->tev.assign((size_t)mpcx_sk::TMAX * mpcx_sk::NKT * 2, nullptr);
  e->tev_on.assign((size_t)mpcx_sk::TMAX * mpcx_sk::NKT, 0);
  for (auto& ev : e->tev) SKCHK(hipEventCreate(&ev));
  // reaction probabilities 1 - dexp(-a00*MPCt) (RMPC_SelkovDimer.f90:136-137) on the host libm
  std::vector<double> tab(2 * 201 * 201);
  for (int ti = 0; ti < 2; ti++)
    for (int tF = 0; tF <= 200; tF++)
      for (int tG = 0; tG <= 200; tG++) {
        double a1 = P.k[0], a2 = P.k[1] * (double)tF, a5 = P.k[4] * (double)tG, a6 = P.k[5], a3, a4;
        if (tG == 0) a3 = 0.0; else a3 = ((P.k[2] * (double)tF) * (double)tG) * (double)(tG - 1);
        if (tG <= 1) a4 = 0.0; else a4 = ((P.k[3] * (double)tG) * (double)(tG - 1)) * (double)(tG - 2);
        if (ti == 0) { a1 = 0.0; a6 = 0.0; }
        const double a00 = ((((a1 + a2) + a3) + a4) + a5) + a6;
        tab[(ti * 201 + tF) * 201 + tG] = 1.0 - exp(-(a00 * P.MPCt));
      }
  SKCHK(hipMemcpy(ptab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
  SKCHK(hipMemcpy(e->d_alt, e->alt, sizeof(e->alt), hipMemcpyHostToDevice));
  SKCHK(hipMemset(D.sc, 0, sizeof(SkScalars)));
  SKCHK(hipMemset(D.near_bits, 0, (e->nwords + 2) * 4));
  SKCHK(hipMemset(e->alt_bits, 0, (e->nwords + 2) * 4));   // words past the last particle stay zero in both
  SKCHK(hipMemset(D.cflag, 0, NS));
  *out = e;
  handle_add(H_SK, e);
  return 0;
}

extern "C" int mpcx_sk_set_state(mpcx_sk* e, int64_t cstep, int64_t NT, const double* x, const double* y,
                                 const double* z, const double* vx, const double* vy, const double* vz,
                                 const double* fx, const double* fy, const double* fz, const int64_t* flag,
                                 const double* dm) {
  if (!e || !x || !y || !z || !vx || !vy || !vz || !flag || !dm) return set_error(MPCX_E_ARG, "null argument");
  if (NT < 0 || NT > e->P.space) return set_error(MPCX_E_ARG, "NT outside 0..space");
  if (cstep % e->P.S != 0) return set_error(MPCX_E_ARG, "the engine starts at a collision step");
  const size_t n = (size_t)e->P.space;
  SkD& D = e->D;
  SKCHK(hipMemcpy(D.x, x, n * 8, hipMemcpyHostToDevice)); SKCHK(hipMemcpy(D.y, y, n * 8, hipMemcpyHostToDevice));
  SKCHK(hipMemcpy(D.z, z, n * 8, hipMemcpyHostToDevice)); SKCHK(hipMemcpy(D.vx, vx, n * 8, hipMemcpyHostToDevice));
  SKCHK(hipMemcpy(D.vy, vy, n * 8, hipMemcpyHostToDevice)); SKCHK(hipMemcpy(D.vz, vz, n * 8, hipMemcpyHostToDevice));
  std::vector<double> zero(n, 0.0);
  SKCHK(hipMemcpy(D.fx, fx ? fx : zero.data(), n * 8, hipMemcpyHostToDevice));
  SKCHK(hipMemcpy(D.fy, fy ? fy : zero.data(), n * 8, hipMemcpyHostToDevice));
  SKCHK(hipMemcpy(D.fz, fz ? fz : zero.data(), n * 8, hipMemcpyHostToDevice));
  std::vector<int> f32(n);
  for (size_t i = 0; i < n; i++) {
    if (flag[i] < 0 || flag[i] > 3) return set_error(MPCX_E_ARG, "FLAG must be 0, 1, 2 or 3");
    f32[i] = (int)flag[i];
  }
  SKCHK(hipMemcpy(D.flag, f32.data(), n * 4, hipMemcpyHostToDevice));
  // dm is (2,12) column-major: field f of sphere s at dm[f*2 + s]
  std::vector<double> dim(24, 0.0);
  for (int f = 0; f < 9; f++)
    for (int s = 0; s < 2; s++) dim[f * 2 + s] = dm[f * 2 + s];
  SKCHK(hipMemcpy(D.dim, dim.data(), 24 * 8, hipMemcpyHostToDevice));
  SkScalars sc;
  memset(&sc, 0, sizeof sc);
  sc.cstep = cstep;
  sc.NT = (int)NT;
  SKCHK(hipMemcpy(D.sc, &sc, sizeof sc, hipMemcpyHostToDevice));
  // the near set of the first interval
  hipStream_t st = e->st;
  hipLaunchKernelGGL(fill_kernel<unsigned>, dim3(grid_for(e->nwords + 2, 256)), dim3(256), 0, st, D.near_bits,
                     e->nwords + 2, 0u);
  hipLaunchKernelGGL(sk_classify_all, dim3(grid_for(n, 256, 2048)), dim3(256), 0, st, e->P, D);
  launch_bits_list(st, D.near_bits, e->nwords, D.partial, D.near_list, D.near_cap, e->P.space, &D.sc->near_num,
                   &D.sc->err);
  SKCHK(hipStreamSynchronize(st));
  return sk_check_error(e);
}

static void sk_near_list(mpcx_sk* e) {
  launch_bits_list(e->st, e->D.near_bits, e->nwords, e->D.partial, e->D.near_list, e->D.near_cap, e->P.space,
                   &e->D.sc->near_num, &e->D.sc->err);
}

static void sk_sort(mpcx_sk* e) {
  SkD& D = e->D;
  SkP& P = e->P;
  const long long NS = P.space, K = 3 * P.vol;
  hipStream_t st = e->st;
  hipLaunchKernelGGL(fill_kernel<int>, dim3(grid_for(K, 256)), dim3(256), 0, st, e->d_scnt, K, 0);
  hipLaunchKernelGGL(sk_sort_bin, dim3(grid_for(NS, 256, 2048)), dim3(256), 0, st, P, D, e->d_scnt, e->d_skey,
                     e->d_srk);
  launch_scan(st, e->d_scnt, K, D.partial, e->d_soff);
  hipLaunchKernelGGL(fill_kernel<int>, dim3(grid_for(NS, 256)), dim3(256), 0, st, e->d_sb, NS, 0);
  hipLaunchKernelGGL(sk_sort_scatter, dim3(grid_for(NS, 256)), dim3(256), 0, st, (int)NS, (const int*)e->d_skey,
                     (const unsigned short*)e->d_srk, (const int*)e->d_soff, e->d_sb);
  hipLaunchKernelGGL(sort_order, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, st, K, (const int*)e->d_scnt,
                     (const int*)e->d_soff, e->d_sb, SK_MAXS);
  hipLaunchKernelGGL(sk_sort_gather, dim3(grid_for(NS, 256)), dim3(256), 0, st, P, D, (const int*)e->d_sb,
                     (const int*)e->d_soff, K, (double* const*)e->d_alt, e->alt_flag, e->alt_cflag, e->alt_bits);
  double** f[9] = {&D.x, &D.y, &D.z, &D.vx, &D.vy, &D.vz, &D.fx, &D.fy, &D.fz};
  for (int a = 0; a < 9; a++) std::swap(*f[a], e->alt[a]);
  std::swap(D.flag, e->alt_flag);
  std::swap(D.cflag, e->alt_cflag);
  std::swap(D.near_bits, e->alt_bits);
  (void)hipMemcpyAsync(e->d_alt, e->alt, sizeof(e->alt), hipMemcpyHostToDevice, st);
}

extern "C" int mpcx_sk_run(mpcx_sk* e, int64_t ncoll, const double* draws, int64_t ndraws, int64_t* used) {
  if (!e || ncoll < 0) return set_error(MPCX_E_ARG, "bad arguments");
  SkD& D = e->D;
  SkP& P = e->P;
  hipStream_t st = e->st;
  if (P.rng_mode == 0) {
    if (!draws || ndraws < 1) return set_error(MPCX_E_ARG, "rng_mode 0 needs host draws");
    if (ndraws > e->nU_cap) {
      void* p = nullptr;
      if (hipMalloc(&p, ndraws * 8) != hipSuccess) return set_error(MPCX_E_HIP, "hipMalloc failed");
      e->allocs.push_back(p);
      e->d_U = (double*)p;
      e->nU_cap = ndraws;
    }
    SKCHK(hipMemcpyAsync(e->d_U, draws, ndraws * 8, hipMemcpyHostToDevice, st));
    D.U = e->d_U;
    D.nU = ndraws;
  }
  SkScalars sc;
  SKCHK(hipMemcpy(&sc, D.sc, sizeof sc, hipMemcpyDeviceToHost));
  long long cstep = sc.cstep;
  const long long cursor0 = 0;
  sc.cursor = cursor0;
  SKCHK(hipMemcpy(&D.sc->cursor, &sc.cursor, sizeof(long long), hipMemcpyHostToDevice));
  const long long NS = P.space, V = P.vol;
  const int gN = grid_for(NS, 256, 2048);
  SKCHK(hipEventRecord(e->e0, st));
  for (int64_t c = 0; c < ncoll; c++) {
    const long long cnext = cstep + P.S;
    const bool sort_due = cnext % P.sort_step == 0;
    D.coll = (unsigned long long)(cnext / P.S);
    const long long slot = c / mpcx_sk::TSTRIDE;
    const bool timed = c % mpcx_sk::TSTRIDE == 0 && slot < mpcx_sk::TMAX;
    auto tmark = [&](int k, int edge) -> hipError_t {
      if (!timed) return hipSuccess;
      e->tev_on[(size_t)slot * mpcx_sk::NKT + k] = 1;
      return hipEventRecord(e->tev[((size_t)slot * mpcx_sk::NKT + k) * 2 + edge], st);
    };
    if (timed)
      for (int k = 0; k < mpcx_sk::NKT; k++) e->tev_on[(size_t)slot * mpcx_sk::NKT + k] = 0;
    hipLaunchKernelGGL(sk_interval_begin, dim3(grid_for(3 * V, 256, 2048)), dim3(256), 0, st, D, (long long)V);
    // device draws: the far particles stream and bin beside the near engine
    // (host draws: the shift is read at the draw cursor the near engine moves)
    const bool ov = e->overlap && P.rng_mode == 1 && !sort_due;
    if (ov) SKCHK(hipEventRecord(e->fork, st));
    SKCHK(tmark(0, 0));
    hipLaunchKernelGGL(sk_near_engine, dim3(1), dim3(SK_T), 0, st, P, D, sort_due ? SM_STOP : 0);
    SKCHK(tmark(0, 1));
    if (ov) {   // enqueued after the near engine, so that it is dispatched first
      SKCHK(hipStreamWaitEvent(e->st2, e->fork, 0));
      if (timed) {
        e->tev_on[(size_t)slot * mpcx_sk::NKT + 1] = 1;
        SKCHK(hipEventRecord(e->tev[((size_t)slot * mpcx_sk::NKT + 1) * 2], e->st2));
      }
      hipLaunchKernelGGL(sk_bulk, dim3(gN), dim3(256), 0, e->st2, P, D, P.S, 1, 1);
      if (timed) SKCHK(hipEventRecord(e->tev[((size_t)slot * mpcx_sk::NKT + 1) * 2 + 1], e->st2));
      SKCHK(hipEventRecord(e->join, e->st2));
    }
    if (ov) {
      SKCHK(hipStreamWaitEvent(st, e->join, 0));
      hipLaunchKernelGGL(sk_bin_near, dim3(64), dim3(256), 0, st, P, D);
    } else if (sort_due) {
      hipLaunchKernelGGL(sk_bulk, dim3(gN), dim3(256), 0, st, P, D, P.S - 1, 0, 0);
      sk_sort(e);
      sk_near_list(e);
      hipLaunchKernelGGL(sk_near_engine, dim3(1), dim3(SK_T), 0, st, P, D, SM_RESUME);
      hipLaunchKernelGGL(sk_bulk, dim3(gN), dim3(256), 0, st, P, D, 1, 1, 0);
    } else {
      SKCHK(tmark(1, 0));
      hipLaunchKernelGGL(sk_bulk, dim3(gN), dim3(256), 0, st, P, D, P.S, 1, 0);
      SKCHK(tmark(1, 1));
    }
    hipLaunchKernelGGL(sk_affected, dim3(1), dim3(256), 0, st, P, D);
    launch_scan(st, D.unaff, V, D.partial, D.uoff);
    launch_scan(st, D.cnt, V, D.partial, D.off);
    hipLaunchKernelGGL(sk_scatter_cells, dim3(gN), dim3(256), 0, st, P, D);
    SKCHK(tmark(2, 0));
    hipLaunchKernelGGL(sk_collide, dim3((unsigned)((V + CT - 1) / CT)), dim3(CT), 0, st, P, D);
    SKCHK(tmark(2, 1));
    hipLaunchKernelGGL(sk_rotate_begin, dim3(grid_for(e->nwords + 2, 256, 2048)), dim3(256), 0, st, P, D,
                       ov ? 0LL : (long long)(e->nwords + 2));
    SKCHK(tmark(3, 0));
    hipLaunchKernelGGL(sk_rotate, dim3(gN), dim3(256), 0, st, P, D, ov ? 1 : 0);
    SKCHK(tmark(3, 1));
    sk_near_list(e);
    cstep = cnext;
  }
  SKCHK(hipEventRecord(e->e1, st));
  SKCHK(hipEventSynchronize(e->e1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e->e0, e->e1);
  e->run_ms = ms;
  const long long kt_n = std::min<long long>((ncoll + mpcx_sk::TSTRIDE - 1) / mpcx_sk::TSTRIDE, mpcx_sk::TMAX);
  for (int k = 0; k < mpcx_sk::NKT; k++) {
    double sum = 0;
    long long n = 0;
    for (long long c = 0; c < kt_n; c++) {
      if (!e->tev_on[(size_t)c * mpcx_sk::NKT + k]) continue;
      float t = 0;
      (void)hipEventElapsedTime(&t, e->tev[((size_t)c * mpcx_sk::NKT + k) * 2], e->tev[((size_t)c * mpcx_sk::NKT + k) * 2 + 1]);
      sum += t;
      n++;
    }
    e->kt_ms[k] = n ? sum / (double)n : 0.0;
  }
  int rc = sk_check_error(e);
  SKCHK(hipMemcpy(&sc, D.sc, sizeof sc, hipMemcpyDeviceToHost));
  if (used) *used = sc.cursor;
  return rc;
}

extern "C" int mpcx_sk_get_state(mpcx_sk* e, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                                 double* fx, double* fy, double* fz, int64_t* flag, double* dm, int64_t* scalars,
                                 double* tpotential) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  SKCHK(hipStreamSynchronize(e->st));
  const size_t n = (size_t)e->P.space;
  SkD& D = e->D;
  double* dst[9] = {x, y, z, vx, vy, vz, fx, fy, fz};
  double* src[9] = {D.x, D.y, D.z, D.vx, D.vy, D.vz, D.fx, D.fy, D.fz};
  for (int a = 0; a < 9; a++)
    if (dst[a]) SKCHK(hipMemcpy(dst[a], src[a], n * 8, hipMemcpyDeviceToHost));
  if (flag) {
    std::vector<int> f32(n);
    SKCHK(hipMemcpy(f32.data(), D.flag, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) flag[i] = f32[i];
  }
  if (dm) {
    std::vector<double> dim(24);
    SKCHK(hipMemcpy(dim.data(), D.dim, 24 * 8, hipMemcpyDeviceToHost));
    for (int f = 0; f < 12; f++)
      for (int s = 0; s < 2; s++) dm[f * 2 + s] = dim[f * 2 + s];
  }
  SkScalars sc;
  SKCHK(hipMemcpy(&sc, D.sc, sizeof sc, hipMemcpyDeviceToHost));
  if (scalars) { scalars[0] = sc.cstep; scalars[1] = sc.NT; scalars[2] = sc.cursor; scalars[3] = sc.near_num; }
  if (tpotential) *tpotential = sc.tpotential;
  return 0;
}

extern "C" int mpcx_sk_get_cells(mpcx_sk* e, double* rec8, int32_t* cell_of, int32_t* census, uint8_t* unaffected) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  SKCHK(hipStreamSynchronize(e->st));
  const long long V = e->P.vol, n = e->P.space;
  if (rec8) SKCHK(hipMemcpy(rec8, e->D.cellrec, V * 64, hipMemcpyDeviceToHost));
  if (cell_of) SKCHK(hipMemcpy(cell_of, e->D.cell_of, n * 4, hipMemcpyDeviceToHost));
  if (census) SKCHK(hipMemcpy(census, e->D.tcount, 3 * V * 4, hipMemcpyDeviceToHost));
  if (unaffected) {
    std::vector<int> u(V);
    SKCHK(hipMemcpy(u.data(), e->D.unaff, V * 4, hipMemcpyDeviceToHost));
    for (long long c = 0; c < V; c++) unaffected[c] = (uint8_t)u[c];
  }
  return 0;
}

namespace mpcx {
int sk_view(mpcx_sk* e, DevView* v) {
  if (!e || !v) return set_error(MPCX_E_ARG, "null argument");
  if (hipStreamSynchronize(e->st) != hipSuccess) return set_error(MPCX_E_HIP, "stream synchronize failed");
  memset(v, 0, sizeof *v);
  const SkD& D = e->D;
  v->x = D.x; v->y = D.y; v->z = D.z; v->vx = D.vx; v->vy = D.vy; v->vz = D.vz; v->flag = D.flag;
  v->n = e->P.space;
  for (int i = 0; i < 3; i++) v->box[i] = e->P.box[i];
  v->mass[0] = e->cfg.Mf; v->mass[1] = e->cfg.Mg; v->mass[3] = e->cfg.Mi;
  v->species[0] = 0; v->species[1] = 1; v->species[2] = -1; v->species[3] = 2;
  v->periodic_z = 1;
  v->nd = 1;
  std::vector<double> dim(24);
  if (hipMemcpy(dim.data(), D.dim, 24 * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(MPCX_E_HIP, "hipMemcpy failed");
  for (int f = 0; f < 12; f++)
    for (int s = 0; s < 2; s++) v->dim[f][s] = dim[f * 2 + s];
  v->msph[0] = e->cfg.Mc; v->msph[1] = e->cfg.Mn;
  SkScalars sc;
  if (hipMemcpy(&sc, D.sc, sizeof sc, hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(MPCX_E_HIP, "hipMemcpy failed");
  v->tpotential = sc.tpotential;
  v->stream = e->st;
  return 0;
}
}  // namespace mpcx

extern "C" int mpcx_sk_set_overlap(mpcx_sk* e, int32_t on) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  e->overlap = on != 0;
  return 0;
}

extern "C" int mpcx_sk_kernel_ms(mpcx_sk* e, double* ms4, double* run_ms) {
  if (!e || !ms4) return set_error(MPCX_E_ARG, "null argument");
  for (int k = 0; k < mpcx_sk::NKT; k++) ms4[k] = e->kt_ms[k];
  if (run_ms) *run_ms = e->run_ms;
  return 0;
}
