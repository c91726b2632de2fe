// Collective engine: the MD + MPC time loop of collective_code on one MI355X.
//
// Work split per collision interval (MPCtstep MD steps + one collision):
//  * near engine (one persistent 1024-thread workgroup per system): the
//    particles that can enter a neighbour list during the interval ("near"
//    set) are advanced exactly like evolve_mod.f95:6-732 -- inside particles
//    live in registers across all MD steps, the dimer lives in LDS and is
//    integrated by thread 0 (RATTLE), neighbour lists are rebuilt by ordered
//    block-wide compaction of the near set.
//  * bulk kernels over all particles: the "far" particles are only ever
//    touched by the lazy streaming of evolve_mod.f95:467-529; the rebuild
//    schedule recorded by the near engine is replayed in registers in one
//    pass, fused with the shifted-cell binning of the collision.  The
//    collision (mpc_rotation_mod.f95:83-232) sums each cell's velocities in
//    ascending particle order (the reference's order at one thread), rotates,
//    and the same pass selects the near set of the next interval.
// The near/far split is exact: a far particle is one whose cell lies further
// from every sphere than cutoff+skin+|v|*tau+guard, and the near engine checks
// at every rebuild that no sphere moved more than the guard.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>
#include "mpcx_device.h"
#include "mpcx_host.h"
#include "mpcx_kernels.h"
#include "../../include/mpcx.h"

#pragma clang fp contract(off)

namespace mpcx {

constexpr int NE_T = 768;    // near-engine threads (one workgroup, 256 VGPRs per lane)
constexpr int NE_CAP = 16384;  // inside-list capacity of the near engine
constexpr int MAXD = 4;      // dimers one near-engine workgroup handles (a cluster)
constexpr int MAXS = 2 * MAXD;
constexpr int MAXM = MAXS;   // force slots kept per inside entry (one per list of its cluster)
constexpr int MAXDG = 120;   // dimers of one system
constexpr int MAXSG = 2 * MAXDG;   // stride of the sphere fields in CoD::dim
constexpr int MAXEV = 64;    // rebuilds recorded per interval
constexpr int BMAX = 64;     // particles per collision cell in the ordered sum
constexpr int CENSUS_MAX = 100;  // MaxA/MaxB/MaxS, global_mod.f95:114-116
constexpr int CT = 256;      // threads of the per-cell kernel
constexpr int NXCD = 8;      // XCDs (accelerator complex dies) of an MI355X, one L2 each
// The reference caps its RATTLE iterations at 2e8 (evolve_mod.f95:104,115); on
// the GPU a non-converging constraint stops at this cap and is reported.
constexpr long long RATTLE_CAP = 100000;

enum : uint8_t { ST_INSIDE = 1, ST_EXCL = 2 };
enum : int { M_SETUP = 1, M_STOP = 2, M_RESUME = 4, M_CLUSTER = 8 };
enum : int { DX = 0, DY, DZ, DVX, DVY, DVZ, DFX, DFY, DFZ, DQX, DQY, DQZ };

struct CoScalars {
  int cstep, time_count, time_penetrate, in_num;
  int near_num, n_ev, err, err_site;   // err_site: the SITE_* of a split kernel's err, or 0
  int events[MAXEV];
  unsigned long long far_maxv2;    // bits of a non-negative double
  double near_vmax2;
  double tpotential;
  double x0[MAXSG], y0[MAXSG], z0[MAXSG];   // spheres at the interval start
  unsigned long long prof[10];  // near engine: total, rebuild, particle phase, serial (100 MHz ticks), rebuilds,
                                // near set, inside list, launches, set-up + gather, end phase (ticks)
  unsigned long long coll;      // collision counter of the current interval (Philox counter)
  double xs0[MAXSG], ys0[MAXSG], zs0[MAXSG];   // copy of x0..z0 made at the interval start (the near
                                               // engine overwrites x0..z0 while co_far still reads them)
  int n_wall, pad_w;            // far particles co_far left to co_bin_rest (they reach a wall)
  int n_late, err_idx;          // late collision cells listed (zone around the spheres, deferred particles);
                                // the particle of a split kernel's err (-1: none named)
  unsigned long long next_fmax2;   // far max |v|^2 of the next interval while far_maxv2 is still in use
  unsigned xmask;                  // split storage: XCDs whose accumulator copies hold this interval's census
  unsigned sums_copied;            // split storage: some cell sums of this interval went to the copies
};
// where a split-storage kernel raised the run's first error (named in its report)
enum : int { SITE_BIN = 1, SITE_BIN_CENSUS, SITE_STREAM, SITE_BULK_CELL, SITE_NEAR_STORE, SITE_INBOX, SITE_ROTATE,
             SITE_SORT_BIN, SITE_PATCH, SITE_COUNT };
__device__ __forceinline__ void raise_error_at(CoScalars* sc, int code, int site, int idx = -1) {
  if (atomicCAS(&sc->err, 0, code) == 0) {
    __hip_atomic_store(&sc->err_site, site, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&sc->err_idx, idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Dimers whose near regions can overlap within an interval form a cluster;
// each cluster is advanced by its own near-engine workgroup, and the
// workgroups of one system meet at a grid barrier once per MD step (and per
// list rebuild) to take the system-wide rebuild decisions of the reference.
struct CluInfo {
  int nclu, abort;
  int cnt[MAXDG];            // dimers per cluster
  int dims[MAXDG][MAXD];     // their global indices, ascending
  int clu_of[MAXDG];         // cluster of each dimer
  int loc[MAXDG];            // position of each dimer within its cluster
  unsigned count, gen;       // barrier arrivals and generation
  unsigned long long pub[2][MAXDG];   // exchanged values (double bits), by generation parity
  unsigned adj[MAXDG][(MAXDG + 31) / 32];   // dimers linked through a near particle
};

struct CoP {
  int NT, nd, MPCtstep, sort_step, react_dir, Lx, Ly, Lz, nrb, rng_mode;
  long long vol, mpcnumcell, rxnnumcell;
  unsigned long long seed;
  double box[3];
  double MDt, hA, hB, hS, hc, hn, IIMcIMn, Mc, Mn, Dcn2;
  double cut2[2], full2[2], full[2];
  double R6[2], R12[2], r6R6[2], r12R12[2];
  double Ea[2], Eb[2], Es[2];
  double Em, mm_cut2[3], mm_R6[3], mm_R12[3];
  double Ew, wR3[2], wR9[2], wA[2], wB[2], wcut2[2], pref;
  double hs2, a0, cosrot, sinrot, omc, twopi, tau, guard, skinrad;
};

struct SpD;
struct CoD {
  double *x, *y, *z, *vx, *vy, *vz;
  int* flag;
  double* dim;            // [12][MAXSG], sphere s = k + i*nd
  CoScalars* sc;
  const int* react_box;
  int* near_list;         // near set of the interval, ascending particle index
  double *ns_x, *ns_y, *ns_z, *ns_vx, *ns_vy, *ns_vz;   // near slots, by near index
  int* ns_flag;
  unsigned char* ns_stat;
  unsigned* near_bits;
  unsigned char* status;
  int* in_list;           // inside list as near indices
  int* in_plist;          // the same entries as particle indices (written at launch end)
  unsigned* in_memb;
  double* in_fp;          // [(m*3+c)*in_cap + q]
  double* in_pot;         // [m*in_cap + q]
  double* in_msd;         // max_solvent_dist of inside entries (slot state below, by entry)
  double *sl_x, *sl_y, *sl_z, *sl_vx, *sl_vy, *sl_vz;
  int* sl_flag;
  int in_cap, near_cap;
  int* cell_of;
  unsigned short* rank_of;
  int* cnt;
  int* off;
  int* bucket;
  int* partial;
  double* cellrec;        // [mpcnumcell][8] by tour position: vcm x y z, uv x y z, n, 0
  const int* mpc_pos;     // tour position of each collision cell
  const int* mpc_cell;    // collision cell at each tour position
  unsigned char* dq;      // distance to the nearest neighbour-list sphere, 1/32 units, capped
  int* tcount;            // [3][vol]
  unsigned* firstB;       // [vol]
  const double* U;        // host draws of this collision (rng_mode 0)
  const double* ptab;     // reaction probabilities, 101 entries
  // clusters (many dimers): the system's cluster table, this workgroup's
  // cluster, and per-particle records of list memberships carried between
  // intervals (up to four global list ids 2k+i as bytes, their pair forces
  // and potentials, max_solvent_dist)
  CluInfo* clu;
  int cid, pad_c;
  unsigned* rec_ids;
  double* rec_msd;
  double* rec_fp;         // [p][4][3]
  double* rec_pot;        // [p][4]
  int* wall_list;         // [wall_cap] far particles whose flight reaches a wall (co_far -> co_bin_rest)
  int wall_cap, pad_w;
  // cells collided in two passes (one beside the near engine, one after it):
  int* slab;              // [mpcnumcell][BMAX] members of each collision cell (by tour position), fixed slots
  int* stamp;             // [mpcnumcell] epoch at which the cell was listed late (collided after the near engine)
  int* late;              // [late_cap] late cells of the interval (tour positions)
  unsigned* next_bits;    // the next interval's near bitmap while near_bits is still in use
  int late_cap, epoch;
  // split storage: per collision cell, by tour position (mpc_pos), the
  // velocity sums in 64-bit fixed point, three words, the member count in the
  // low 12 bits of the z word; and the cell records written from them (vcm x y
  // z, count; axis); cell_of holds a particle's collision cell the same way
  unsigned long long* sacc;
  // the part of each tour cell's sums binned by the particle tile that owns
  // it (the tile holding the cell's first particle at the last re-sort):
  // written whole, by plain stores, by that tile in every interval
  unsigned long long* sown;
  const SpD* spd;         // a split batch's copy of the descriptor: the split arrays the near engine gathers from
  const int* soff;        // [vol + 1] first particle of each tour position at the last re-sort
  float4* srec;
  int* xtc;               // split storage B census, one copy per XCD: [8][vol], by tour position
  unsigned* xfb;          // split storage first-B keys, one copy per XCD: [8][vol], by tour position
};

// ------------------------------------------------------------- streaming
// One streaming update of evolve_mod.f95:259-312 (tc = 1: inside particles,
// whose "x + v*MDt" equals "(v*MDt)*1"), or of the lazy update :474-527.
// the wall part of a streaming update, after the free flight
__device__ __forceinline__ bool stream_walls(double& x, double& y, double& z, double& vx, double& vy, double& vz,
                                             int& flag, const CoP& P, const int* react_box, int* err) {
  const double L0 = P.box[0], L1 = P.box[1], L2 = P.box[2];
  if (z < 0.0) {
    double t = z / vz;
    // reactive_grid_mod.f95:128-149
    double ax = x - vx * t, ay = y - vy * t;
    if (ax >= L0) ax = ax - L0; else if (ax < 0.0) ax = ax + L0;
    if (ay >= L1) ay = ay - L1; else if (ay < 0.0) ay = ay + L1;
    int hit = (int)(ffloor_i(ax) + L0 * ffloor_i(ay) + 1.0);
    if (hit < 1 || hit > P.nrb) raise_error(err, ERR_CELL_RANGE);
    else if (react_box[hit - 1] == 1) {
      if (P.react_dir == 0 && flag == 0) flag = 1;
      else if (P.react_dir == 1 && flag == 1) flag = 0;
    }
    vx = -vx; vy = -vy; vz = -vz;
    x = x + (2.0 * vx) * t; y = y + (2.0 * vy) * t; z = z + (2.0 * vz) * t;
    if (x >= L0) x = x - L0; else if (x < 0.0) x = x + L0;
    if (y >= L1) y = y - L1; else if (y < 0.0) y = y + L1;
    return true;
  } else if (z > L2) {
    double t = (z - L2) / vz;
    vx = -vx; vy = -vy; vz = -vz;
    x = x + (2.0 * vx) * t; y = y + (2.0 * vy) * t; z = z + (2.0 * vz) * t;
    if (x >= L0) x = x - L0; else if (x < 0.0) x = x + L0;
    if (y >= L1) y = y - L1; else if (y < 0.0) y = y + L1;
    return true;
  }
  return false;
}
// free flight of tc steps with the step displacements d = v*MDt (the
// reference's (v*MDt)*tc) and periodic wrap in x and y
__device__ __forceinline__ void stream_flight(double& x, double& y, double& z, double dx, double dy, double dz,
                                              double tc, const CoP& P) {
  const double L0 = P.box[0], L1 = P.box[1];
  x = x + dx * tc;
  y = y + dy * tc;
  z = z + dz * tc;
  // the reference's "if (x >= L) x = x - L; else if (x < 0) x = x + L" as selects
  const double xm = x - L0, xp = x + L0, ym = y - L1, yp = y + L1;
  x = x >= L0 ? xm : (x < 0.0 ? xp : x);
  y = y >= L1 ? ym : (y < 0.0 ? yp : y);
}
__device__ __forceinline__ bool stream_particle(double& x, double& y, double& z, double& vx, double& vy,
                                                double& vz, int& flag, double tc, const CoP& P,
                                                const int* react_box, int* err) {
  const double MDt = P.MDt;
  stream_flight(x, y, z, vx * MDt, vy * MDt, vz * MDt, tc, P);
  return stream_walls(x, y, z, vx, vy, vz, flag, P, react_box, err);
}

__device__ __forceinline__ double kick_factor(const CoP& P, int flag, bool& ok) {
  ok = true;
  if (flag == 0) return P.hA;
  if (flag == 1) return P.hB;
  if (flag == 2) return P.hS;
  ok = false;
  return 0.0;
}

// neighbour-list membership of a position (bit 2k+i: sphere i of dimer k)
template <int ND>
__device__ __forceinline__ unsigned memberships(double x, double y, double z, const CoP& P,
                                                const double (*sd)[MAXS]) {
  unsigned m = 0;
#pragma unroll
  for (int k = 0; k < ND; k++)
#pragma unroll
    for (int i = 0; i < 2; i++) {
      int s = k + i * ND;
      double d1 = pbc_half(x - sd[DX][s], P.box[0]);
      double d2 = pbc_half(y - sd[DY][s], P.box[1]);
      double d3 = z - sd[DZ][s];
      if (d1 * d1 + d2 * d2 + d3 * d3 <= P.full2[i]) m |= 1u << (2 * k + i);
    }
  return m;
}

// X - L*floor(X/L) of the reference without the division: for -L <= X < 2L a
// correctly rounded X/L lies in [-1,0), [0,1) or [1,2) exactly when X does, so
// floor(X/L) is -1, 0 or 1 by comparison and the subtraction is the same
// IEEE operation (X - (-L) is X + L; X - 0*L keeps X, -0.0 included)
__device__ __forceinline__ double wrap_periodic(double X, double L) {
  return X < 0.0 ? X + L : (X >= L ? X - L : X);
}
// the same for any X: the division only outside [-L, 2L)
__device__ __forceinline__ double wrap_periodic_any(double X, double L) {
  if (__builtin_expect(X < -L || X >= 2.0 * L, 0)) return X - L * floor(X / L);
  return wrap_periodic(X, L);
}

// ------------------------------------------------------------- dimer (thread 0)
// The dimer state of one system, held in registers while thread 0 integrates
// it (all indices are compile-time after unrolling).  Sphere s = k + i*ND.
template <int ND>
struct Dimer {
  double v[12][2 * ND];
};

template <int ND>
__device__ __forceinline__ void dimer_load(Dimer<ND>& d, const double (*sd)[MAXS]) {
#pragma unroll
  for (int f = 0; f < 12; f++)
#pragma unroll
    for (int s = 0; s < 2 * ND; s++) d.v[f][s] = sd[f][s];
}
template <int ND>
__device__ __forceinline__ void dimer_store(const Dimer<ND>& d, double (*sd)[MAXS]) {
#pragma unroll
  for (int f = 0; f < 12; f++)
#pragma unroll
    for (int s = 0; s < 2 * ND; s++) sd[f][s] = d.v[f][s];
}

// evolve_mod.f95:325-415: Q = V + h F, advance, RATTLE on the bond length.
template <int ND>
__device__ __forceinline__ void rattle_positions(const CoP& P, Dimer<ND>& d, int* err) {
  const double MDt = P.MDt, L0 = P.box[0], L1 = P.box[1];
#pragma unroll
  for (int k = 0; k < ND; k++) {
    const int s0 = k, s1 = k + ND;
    double prX[2] = {d.v[DX][s0], d.v[DX][s1]}, prY[2] = {d.v[DY][s0], d.v[DY][s1]},
           prZ[2] = {d.v[DZ][s0], d.v[DZ][s1]};
    double dP0 = pbc_half(d.v[DX][s0] - d.v[DX][s1], L0);
    double dP1 = pbc_half(d.v[DY][s0] - d.v[DY][s1], L1);
    double dP2 = d.v[DZ][s0] - d.v[DZ][s1];
    d.v[DQX][s0] = d.v[DVX][s0] + P.hc * d.v[DFX][s0];
    d.v[DQY][s0] = d.v[DVY][s0] + P.hc * d.v[DFY][s0];
    d.v[DQZ][s0] = d.v[DVZ][s0] + P.hc * d.v[DFZ][s0];
    d.v[DQX][s1] = d.v[DVX][s1] + P.hn * d.v[DFX][s1];
    d.v[DQY][s1] = d.v[DVY][s1] + P.hn * d.v[DFY][s1];
    d.v[DQZ][s1] = d.v[DVZ][s1] + P.hn * d.v[DFZ][s1];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int s = k + i * ND;
      double x = d.v[DX][s] + MDt * d.v[DQX][s];
      double y = d.v[DY][s] + MDt * d.v[DQY][s];
      d.v[DZ][s] = d.v[DZ][s] + MDt * d.v[DQZ][s];
      d.v[DX][s] = wrap_periodic_any(x, L0);
      d.v[DY][s] = wrap_periodic_any(y, L1);
    }
    for (long long acc = 0;; acc++) {
      if (acc == RATTLE_CAP) { raise_error(err, ERR_RATTLE); break; }
      double d0 = pbc_half(d.v[DX][s0] - d.v[DX][s1], L0);
      double d1 = pbc_half(d.v[DY][s0] - d.v[DY][s1], L1);
      double d2 = d.v[DZ][s0] - d.v[DZ][s1];
      double c1 = (d0 * d0 + d1 * d1 + d2 * d2) - P.Dcn2;
      if (fabs(c1) < 1e-6) break;
      double G = -(((0.25 * c1) * P.IIMcIMn) / (d0 * dP0 + d1 * dP1 + d2 * dP2));
      double g2 = 2.0 * G;
      double x0 = (prX[0] + d.v[DQX][s0] * MDt) + (g2 * dP0) / P.Mc;
      double y0 = (prY[0] + d.v[DQY][s0] * MDt) + (g2 * dP1) / P.Mc;
      double z0 = (prZ[0] + d.v[DQZ][s0] * MDt) + (g2 * dP2) / P.Mc;
      double x1 = (prX[1] + d.v[DQX][s1] * MDt) - (g2 * dP0) / P.Mn;
      double y1 = (prY[1] + d.v[DQY][s1] * MDt) - (g2 * dP1) / P.Mn;
      double z1 = (prZ[1] + d.v[DQZ][s1] * MDt) - (g2 * dP2) / P.Mn;
      d.v[DX][s0] = wrap_periodic_any(x0, L0); d.v[DY][s0] = wrap_periodic_any(y0, L1); d.v[DZ][s0] = z0;
      d.v[DX][s1] = wrap_periodic_any(x1, L0); d.v[DY][s1] = wrap_periodic_any(y1, L1); d.v[DZ][s1] = z1;
      d.v[DQX][s0] = d.v[DQX][s0] + (g2 * dP0) / (P.Mc * MDt);
      d.v[DQY][s0] = d.v[DQY][s0] + (g2 * dP1) / (P.Mc * MDt);
      d.v[DQZ][s0] = d.v[DQZ][s0] + (g2 * dP2) / (P.Mc * MDt);
      d.v[DQX][s1] = d.v[DQX][s1] - (g2 * dP0) / (P.Mn * MDt);
      d.v[DQY][s1] = d.v[DQY][s1] - (g2 * dP1) / (P.Mn * MDt);
      d.v[DQZ][s1] = d.v[DQZ][s1] - (g2 * dP2) / (P.Mn * MDt);
    }
  }
}

// forces_mod.f95:183-208
__device__ __forceinline__ void lj_mot_mot(const CoP& P, double x1, double y1, double z1, double x2, double y2,
                                           double z2, int kind, double& fx, double& fy, double& fz, double& fp) {
  double d1 = pbc_half(x1 - x2, P.box[0]);
  double d2 = pbc_half(y1 - y2, P.box[1]);
  double d3 = z1 - z2;
  double dRsq = d1 * d1 + d2 * d2 + d3 * d3;
  fx = fy = fz = fp = 0.0;
  if (dRsq <= P.mm_cut2[kind]) {
    double F, U;
    lj_terms(dRsq, P.mm_R6[kind], P.mm_R12[kind], 12.0 * P.mm_R12[kind], 6.0 * P.mm_R6[kind], F, U);
    double EF = P.Em * F;
    fx = EF * d1; fy = EF * d2; fz = EF * d3;
    fp = P.Em * U;
  }
}

// forces_mod.f95:303-333 with the dummy order of the call at evolve_mod.f95:597-600
__device__ __forceinline__ void lj_mot_wall(const CoP& P, double zd, int i, double& fw, double& pot) {
  const double R3 = P.wR3[i], R9 = P.wR9[i], r9R9 = P.wA[i], r3R3 = P.wB[i];
  const double pe = P.pref * P.Ew;
  fw = 0.0;
  pot = 0.0;
  double dR = P.box[2] - zd;
  if (dR * dR < P.wcut2[i]) {
    double IdRsq = 1.0 / (dR * dR);
    double IdR3 = pow(IdRsq, 1.5);
    double IdR9 = (IdR3 * IdR3) * IdR3;
    double IdR5 = IdR3 * IdRsq;
    double IdR11 = IdR9 * IdRsq;
    fw = (pe * (r9R9 * IdR11 - r3R3 * IdR5)) * dR;
    pot = pe * (R9 * IdR9 - R3 * IdR3) + 1.0;
  }
  dR = 0.0 - zd;
  if (dR * dR < P.wcut2[i]) {
    double IdRsq = 1.0 / (dR * dR);
    double IdR3 = pow(IdRsq, 1.5);
    double IdR9 = (IdR3 * IdR3) * IdR3;
    double IdR5 = IdR3 * IdRsq;
    double IdR11 = IdR9 * IdRsq;
    fw = fw + (pe * (r9R9 * IdR11 - r3R3 * IdR5)) * dR;
    pot = (pot + pe * (R9 * IdR9 - R3 * IdR3)) + 1.0;
  }
}

// evolve_mod.f95:605-636 (solvent sums arrive as tree-reduced totals) and
// :679-724 (velocity RATTLE).  sumF[4*s + 0..2] = sum of Fpair of sphere s,
// sumF[4*s + 3] = sum of its Potential.
template <int ND>
__device__ __forceinline__ void dimer_forces_and_rattle_v(const CoP& P, Dimer<ND>& d, const double* sumF,
                                                          double& tpot, int* err) {
  double tpp[2 * ND];
#pragma unroll
  for (int k = 0; k < ND; k++)
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int s = k + i * ND;
      double fw, wp;
      lj_mot_wall(P, d.v[DZ][s], i, fw, wp);
      double fx = 0.0 - sumF[4 * s + 0], fy = 0.0 - sumF[4 * s + 1], fz = 0.0 - sumF[4 * s + 2];
      double tp = 0.0 + sumF[4 * s + 3];
#pragma unroll
      for (int j = 0; j < ND; j++)
#pragma unroll
        for (int l = 0; l < 2; l++) {
          if (j == k) continue;   // Fmot(k,i,k,l) stays zero (evolve_mod.f95:580)
          const int b = j + l * ND;
          const int kind = (i == 0 && l == 0) ? 0 : (i == 1 && l == 1) ? 2 : 1;
          double mx, my, mz, mp;
          lj_mot_mot(P, d.v[DX][s], d.v[DY][s], d.v[DZ][s], d.v[DX][b], d.v[DY][b], d.v[DZ][b], kind, mx, my, mz,
                     mp);
          fx = fx + mx; fy = fy + my; fz = fz + mz;
          if (j < k) tp = tp + mp;
        }
      fz = fz + fw;
      tp = tp + wp;
      d.v[DFX][s] = fx; d.v[DFY][s] = fy; d.v[DFZ][s] = fz;
      tpp[s] = tp;
    }
  tpot = 0.0;
#pragma unroll
  for (int k = 0; k < ND; k++)
#pragma unroll
    for (int i = 0; i < 2; i++) tpot = tpot + tpp[k + i * ND];
#pragma unroll
  for (int k = 0; k < ND; k++) {
    const int s0 = k, s1 = k + ND;
    d.v[DVX][s0] = d.v[DQX][s0] + P.hc * d.v[DFX][s0];
    d.v[DVY][s0] = d.v[DQY][s0] + P.hc * d.v[DFY][s0];
    d.v[DVZ][s0] = d.v[DQZ][s0] + P.hc * d.v[DFZ][s0];
    d.v[DVX][s1] = d.v[DQX][s1] + P.hn * d.v[DFX][s1];
    d.v[DVY][s1] = d.v[DQY][s1] + P.hn * d.v[DFY][s1];
    d.v[DVZ][s1] = d.v[DQZ][s1] + P.hn * d.v[DFZ][s1];
    for (long long acc = 0;; acc++) {
      if (acc == RATTLE_CAP) { raise_error(err, ERR_RATTLE); break; }
      double dvx = d.v[DVX][s0] - d.v[DVX][s1], dvy = d.v[DVY][s0] - d.v[DVY][s1], dvz = d.v[DVZ][s0] - d.v[DVZ][s1];
      double d0 = pbc_half(d.v[DX][s0] - d.v[DX][s1], P.box[0]);
      double d1 = pbc_half(d.v[DY][s0] - d.v[DY][s1], P.box[1]);
      double d2 = d.v[DZ][s0] - d.v[DZ][s1];
      double c2 = 2.0 * (dvx * d0 + dvy * d1 + dvz * d2);
      if (fabs(c2) < 1e-8) break;
      double O = -((0.25 * (P.IIMcIMn * c2)) / P.Dcn2);
      double o2 = 2.0 * O;
      d.v[DVX][s0] = d.v[DVX][s0] + (o2 * d0) / P.Mc;
      d.v[DVY][s0] = d.v[DVY][s0] + (o2 * d1) / P.Mc;
      d.v[DVZ][s0] = d.v[DVZ][s0] + (o2 * d2) / P.Mc;
      d.v[DVX][s1] = d.v[DVX][s1] - (o2 * d0) / P.Mn;
      d.v[DVY][s1] = d.v[DVY][s1] - (o2 * d1) / P.Mn;
      d.v[DVZ][s1] = d.v[DVZ][s1] - (o2 * d2) / P.Mn;
    }
  }
}

// Wave forms of the one-dimer step: every lane of the wave holds the dimer
// and computes the same scalars; in each RATTLE iteration lane c (0..5)
// updates component c of the update -- sphere c / 3, axis c % 3 -- with the
// expression of the one-thread form above, and the six results are read back
// from lanes 0..5 into every lane.  The results are bitwise those of the
// one-thread form, and an iteration's chain holds two or three divisions
// instead of twelve (a wave issues every instruction whatever lanes are
// active).  Lane 0 alone reports errors.
__device__ __forceinline__ double pick6(int c, double a0, double a1, double a2, double b0, double b1, double b2) {
  return c < 3 ? (c == 0 ? a0 : c == 1 ? a1 : a2) : (c == 3 ? b0 : c == 4 ? b1 : b2);
}
__device__ __forceinline__ void bcast6(double v, double& a0, double& a1, double& a2, double& b0, double& b1,
                                       double& b2) {
  a0 = lane_read_d(v, 0); a1 = lane_read_d(v, 1); a2 = lane_read_d(v, 2);
  b0 = lane_read_d(v, 3); b1 = lane_read_d(v, 4); b2 = lane_read_d(v, 5);
}
__device__ __forceinline__ void rattle_positions_wave(const CoP& P, Dimer<1>& d, int* err, int lane) {
  const double MDt = P.MDt, L0 = P.box[0], L1 = P.box[1];
  const double prX[2] = {d.v[DX][0], d.v[DX][1]}, prY[2] = {d.v[DY][0], d.v[DY][1]},
               prZ[2] = {d.v[DZ][0], d.v[DZ][1]};
  const double dP0 = pbc_half(d.v[DX][0] - d.v[DX][1], L0);
  const double dP1 = pbc_half(d.v[DY][0] - d.v[DY][1], L1);
  const double dP2 = d.v[DZ][0] - d.v[DZ][1];
  d.v[DQX][0] = d.v[DVX][0] + P.hc * d.v[DFX][0];
  d.v[DQY][0] = d.v[DVY][0] + P.hc * d.v[DFY][0];
  d.v[DQZ][0] = d.v[DVZ][0] + P.hc * d.v[DFZ][0];
  d.v[DQX][1] = d.v[DVX][1] + P.hn * d.v[DFX][1];
  d.v[DQY][1] = d.v[DVY][1] + P.hn * d.v[DFY][1];
  d.v[DQZ][1] = d.v[DVZ][1] + P.hn * d.v[DFZ][1];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    double x = d.v[DX][s] + MDt * d.v[DQX][s];
    double y = d.v[DY][s] + MDt * d.v[DQY][s];
    d.v[DZ][s] = d.v[DZ][s] + MDt * d.v[DQZ][s];
    d.v[DX][s] = wrap_periodic_any(x, L0);
    d.v[DY][s] = wrap_periodic_any(y, L1);
  }
  const int c = lane < 6 ? lane : 0, ax = c % 3;
  const bool first = c < 3;   // sphere 0 (C, mass Mc) or 1 (N, mass Mn)
  const double pr = pick6(c, prX[0], prY[0], prZ[0], prX[1], prY[1], prZ[1]);
  const double dPa = ax == 0 ? dP0 : (ax == 1 ? dP1 : dP2);
  const double M = first ? P.Mc : P.Mn;
  double q = pick6(c, d.v[DQX][0], d.v[DQY][0], d.v[DQZ][0], d.v[DQX][1], d.v[DQY][1], d.v[DQZ][1]);
  for (long long acc = 0;; acc++) {
    if (acc == RATTLE_CAP) { if (lane == 0) raise_error(err, ERR_RATTLE); break; }
    const double d0 = pbc_half(d.v[DX][0] - d.v[DX][1], L0);
    const double d1 = pbc_half(d.v[DY][0] - d.v[DY][1], L1);
    const double d2 = d.v[DZ][0] - d.v[DZ][1];
    const double c1 = (d0 * d0 + d1 * d1 + d2 * d2) - P.Dcn2;
    if (fabs(c1) < 1e-6) break;
    const double G = -(((0.25 * c1) * P.IIMcIMn) / (d0 * dP0 + d1 * dP1 + d2 * dP2));
    const double g2 = 2.0 * G;
    const double base = pr + q * MDt, dl = (g2 * dPa) / M;
    double x = first ? base + dl : base - dl;
    if (ax == 0) x = wrap_periodic_any(x, L0);
    else if (ax == 1) x = wrap_periodic_any(x, L1);
    const double dq = (g2 * dPa) / (M * MDt);
    q = first ? q + dq : q - dq;
    bcast6(x, d.v[DX][0], d.v[DY][0], d.v[DZ][0], d.v[DX][1], d.v[DY][1], d.v[DZ][1]);
  }
  bcast6(q, d.v[DQX][0], d.v[DQY][0], d.v[DQZ][0], d.v[DQX][1], d.v[DQY][1], d.v[DQZ][1]);
}
__device__ __forceinline__ void dimer_forces_and_rattle_v_wave(const CoP& P, Dimer<1>& d, const double* sumF,
                                                               double& tpot, int* err, int lane) {
  double tpp[2];
#pragma unroll
  for (int i = 0; i < 2; i++) {   // one dimer: no sphere-sphere terms (evolve_mod.f95:580)
    double fw, wp;
    lj_mot_wall(P, d.v[DZ][i], i, fw, wp);
    double fx = 0.0 - sumF[4 * i + 0], fy = 0.0 - sumF[4 * i + 1], fz = 0.0 - sumF[4 * i + 2];
    double tp = 0.0 + sumF[4 * i + 3];
    fz = fz + fw;
    tp = tp + wp;
    d.v[DFX][i] = fx; d.v[DFY][i] = fy; d.v[DFZ][i] = fz;
    tpp[i] = tp;
  }
  tpot = 0.0;
  tpot = tpot + tpp[0];
  tpot = tpot + tpp[1];
  d.v[DVX][0] = d.v[DQX][0] + P.hc * d.v[DFX][0];
  d.v[DVY][0] = d.v[DQY][0] + P.hc * d.v[DFY][0];
  d.v[DVZ][0] = d.v[DQZ][0] + P.hc * d.v[DFZ][0];
  d.v[DVX][1] = d.v[DQX][1] + P.hn * d.v[DFX][1];
  d.v[DVY][1] = d.v[DQY][1] + P.hn * d.v[DFY][1];
  d.v[DVZ][1] = d.v[DQZ][1] + P.hn * d.v[DFZ][1];
  const int c = lane < 6 ? lane : 0, ax = c % 3;
  const bool first = c < 3;
  const double M = first ? P.Mc : P.Mn;
  double v = pick6(c, d.v[DVX][0], d.v[DVY][0], d.v[DVZ][0], d.v[DVX][1], d.v[DVY][1], d.v[DVZ][1]);
  for (long long acc = 0;; acc++) {
    if (acc == RATTLE_CAP) { if (lane == 0) raise_error(err, ERR_RATTLE); break; }
    const double dvx = d.v[DVX][0] - d.v[DVX][1], dvy = d.v[DVY][0] - d.v[DVY][1], dvz = d.v[DVZ][0] - d.v[DVZ][1];
    const double d0 = pbc_half(d.v[DX][0] - d.v[DX][1], P.box[0]);
    const double d1 = pbc_half(d.v[DY][0] - d.v[DY][1], P.box[1]);
    const double d2 = d.v[DZ][0] - d.v[DZ][1];
    const double c2 = 2.0 * (dvx * d0 + dvy * d1 + dvz * d2);
    if (fabs(c2) < 1e-8) break;
    const double O = -((0.25 * (P.IIMcIMn * c2)) / P.Dcn2);
    const double o2 = 2.0 * O;
    const double da = ax == 0 ? d0 : (ax == 1 ? d1 : d2);
    const double dl = (o2 * da) / M;
    v = first ? v + dl : v - dl;
    bcast6(v, d.v[DVX][0], d.v[DVY][0], d.v[DVZ][0], d.v[DVX][1], d.v[DVY][1], d.v[DVZ][1]);
  }
}

// ------------------------------------------------------------- near engine
// One workgroup advances the near set of one system through a whole collision
// interval.  The near particles are gathered once per interval into
// contiguous slot arrays (ns_*, by near index q, ascending particle index;
// they stay in L2), so list rebuilds read coalesced data.  The inside list
// holds near indices; the state of its entries lives in LDS (entry k < LCAP)
// or in the HBM slot arrays sl_* (k >= LCAP) for all MD steps between two
// rebuilds.  Rebuild partition: wave w owns near indices
// [w*64*E, (w+1)*64*E), lane-interleaved (q = w*64*E + j*64 + lane), so the
// order of a compaction follows ballots within a wave and one scan over the
// waves.
constexpr int NE_EMAX = 64;    // near entries per lane: near set <= NE_T * NE_EMAX
constexpr int NE_NEAR = NE_T * NE_EMAX;
constexpr int NE_W = NE_T / 64;

// memberships of a lane's near entries, B bits each (B = 4 for up to two
// dimers, 8 for up to four): entry j in field j
template <int B>
struct Memb {
  static constexpr int PER = 64 / B;
  static constexpr unsigned MASK = (1u << B) - 1u;
  unsigned long long w[NE_EMAX / PER];
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int i = 0; i < NE_EMAX / PER; i++) w[i] = 0ull;
  }
  __device__ __forceinline__ void set(int j, unsigned m) {
#pragma unroll
    for (int i = 0; i < NE_EMAX / PER; i++)
      if (j / PER == i) w[i] |= (unsigned long long)(m & MASK) << (B * (j % PER));
  }
  __device__ __forceinline__ unsigned get(int j) const {
    unsigned r = 0;
#pragma unroll
    for (int i = 0; i < NE_EMAX / PER; i++)
      if (j / PER == i) r = (unsigned)(w[i] >> (B * (j % PER))) & MASK;
    return r;
  }
};
template <int ND>
using MembN = Memb<(ND <= 2 ? 4 : 8)>;

__device__ __forceinline__ unsigned long long lanes_below() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Grid barrier of the clusters of one system (thread 0 of each cluster's
// workgroup): publishes v, waits for every cluster and returns the maximum.
// Values are double-buffered by the generation's parity, so a fast cluster
// entering the next barrier cannot overwrite what a slow one still reads.  A
// bounded wait: a cluster that stopped early sets `abort`, and every wait
// gives up after SPIN_CAP polls, so all workgroups always finish.
constexpr long long SPIN_CAP = 1ll << 24;
__device__ __attribute__((noinline)) double clu_exchange_max(CluInfo* ci, int cid, double v, int* err) {
  const int n = ci->nclu;
  if (n <= 1) return v;
  // The clusters exchange only these values, as agent-scope atomics (each
  // keeps its dimers and particles for the whole launch), so no L2 write-back
  // or invalidate: a store completes (s_waitcnt 0) before the next atomic
  // that others wait on.  (Acquire loads in the poll invalidated L2 at every
  // poll; 8 clustered dimers in 64x64x32: 2.32 -> 2.30 ms per interval.)
  const unsigned gen = __hip_atomic_load(&ci->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  unsigned long long* pub = ci->pub[gen & 1u];
  __hip_atomic_store(&pub[cid], (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_s_waitcnt(0);
  const unsigned arrived = atomicAdd(&ci->count, 1u);
  if (arrived == (unsigned)n - 1u) {
    atomicExch(&ci->count, 0u);
    __builtin_amdgcn_s_waitcnt(0);
    atomicAdd(&ci->gen, 1u);
  } else {
    long long spins = 0;
    while (__hip_atomic_load(&ci->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
      if (__hip_atomic_load(&ci->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > SPIN_CAP) {
        raise_error(err, ERR_BARRIER);
        atomicExch(&ci->abort, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  double m = v;
  for (int c = 0; c < n; c++)
    m = fmax(m, __longlong_as_double((long long)__hip_atomic_load(&pub[c], __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_AGENT)));
  return m;
}


// Split-precision storage of the solvent (the kernels further below): per
// particle the packed unshifted cell, FP32 offsets in it, FP32 velocities and
// the species.  The batch near engines gather their near sets from it.
struct SpD {
  int* cell;
  float *ox, *oy, *oz, *vx, *vy, *vz;
  signed char* flag;
};

// The cell of a split-storage particle is held as its packed coordinates
// ix | iy << 11 | iz << 22 (boxes up to 2048 x 2048 x 1024 cells), so no
// kernel divides by the box edges to find them.
__device__ __forceinline__ int sp_pack(int ix, int iy, int iz) { return ix | (iy << 11) | (iz << 22); }
__device__ __forceinline__ void sp_cell_xyz(const CoP& P, int cell, int& ix, int& iy, int& iz) {
  (void)P;
  ix = cell & 2047;
  iy = (cell >> 11) & 2047;
  iz = (int)((unsigned)cell >> 22);
}

// The near engine (its LDS state and kernels): up to LCAP inside entries are
// staged in LDS (150 KB with the rest of its state: one engine per CU); the
// entries beyond them live in the HBM slot arrays (the same arithmetic).
constexpr int LCAP = 2304;
#include "mpcx_near_engine.inc"


// ------------------------------------------------------------- clusters
// Segment bases of the per-cluster near engines (cluster c owns entries
// [c*cap, (c+1)*cap) of each array) and their descriptors.
struct CluBase {
  CoP* Ps;
  CoD* Ds;
  CoScalars* scs;
  double* dimc;                 // [c][12*MAXSG]
  int* near_list;
  double *ns_x, *ns_y, *ns_z, *ns_vx, *ns_vy, *ns_vz;
  int* ns_flag;
  unsigned char* ns_stat;
  int *in_list, *in_plist;
  unsigned* in_memb;
  double *in_fp, *in_pot, *in_msd;
  double *sl_x, *sl_y, *sl_z, *sl_vx, *sl_vy, *sl_vz;
  int* sl_flag;
  unsigned char* assign;        // dimer of the nearest sphere of each entry of the system's near list
};

__device__ __forceinline__ void sphere_xyz(const CoD& D, int s, double& x, double& y, double& z) {
  x = D.dim[DX * MAXSG + s]; y = D.dim[DY * MAXSG + s]; z = D.dim[DZ * MAXSG + s];
}

// Links between dimers, from the near set: a near particle within reach of
// spheres of two dimers (|v| tau + guard, plus the skin for list members
// carried over and 1/32 for the quantized near test) links them, so two
// clusters never share a particle that could enter a list of either within
// the interval.  Also the dimer of each entry's nearest sphere (beyond its
// list radius, the measure of the near/far split), which picks its cluster.
__global__ __launch_bounds__(256) void clu_links(CoP P, CoD G, unsigned char* near_dimer) {
  __shared__ double sx[MAXSG], sy[MAXSG], sz[MAXSG];
  const int nd = P.nd;
  for (int s = threadIdx.x; s < 2 * nd; s += 256) sphere_xyz(G, s, sx[s], sy[s], sz[s]);
  __syncthreads();
  const double tau = P.tau * (1.0 + 1e-9), extra = P.guard + 1e-6 + P.skinrad + 1.0 / 32.0;
  const int nn = min(G.sc->near_num, G.near_cap);
  for (int q = blockIdx.x * 256 + threadIdx.x; q < nn; q += gridDim.x * 256) {
    const int p = G.near_list[q];
    const double x = G.x[p], y = G.y[p], z = G.z[p], vx = G.vx[p], vy = G.vy[p], vz = G.vz[p];
    const double reach = sqrt((vx * vx + vy * vy) + vz * vz) * tau + extra;
    double best = 1e300;
    int bs = 0, k0 = -1;
    for (int s = 0; s < 2 * nd; s++) {
      const double a = pbc_half(x - sx[s], P.box[0]), b = pbc_half(y - sy[s], P.box[1]), c = z - sz[s];
      const double d = sqrt(a * a + b * b + c * c) - P.full[s >= nd];
      if (d < best) { best = d; bs = s; }
      if (d <= reach) {
        const int k = s < nd ? s : s - nd;
        if (k0 < 0) k0 = k;
        else if (k != k0) {
          atomicOr(&G.clu->adj[k0][k >> 5], 1u << (k & 31));
          atomicOr(&G.clu->adj[k][k0 >> 5], 1u << (k0 & 31));
        }
      }
    }
    near_dimer[q] = (unsigned char)(bs < nd ? bs : bs - nd);
  }
}

__global__ void clu_clear(CluInfo* ci) {
  for (int j = threadIdx.x; j < MAXDG * ((MAXDG + 31) / 32); j += blockDim.x) ci->adj[j / ((MAXDG + 31) / 32)][j % ((MAXDG + 31) / 32)] = 0u;
}

// Clusters: the connected groups of the near-particle links and of spheres
// within the sphere-sphere cutoff of each other plus twice the guard (each may
// move that far in the interval), numbered by their smallest dimer index.
// Also writes each cluster's parameters, descriptor, scalars and local
// sphere fields.
__global__ __launch_bounds__(256) void clu_build(CoP P, CoD G, CluBase B) {
  CluInfo* ci = G.clu;
  const int nd = P.nd, tid = threadIdx.x;
  __shared__ unsigned adj[MAXDG][(MAXDG + 31) / 32];
  __shared__ int lab[MAXDG], nlab[MAXDG];
  __shared__ int changed;
  for (int j = tid; j < MAXDG * ((MAXDG + 31) / 32); j += 256)
    adj[j / ((MAXDG + 31) / 32)][j % ((MAXDG + 31) / 32)] = ci->adj[j / ((MAXDG + 31) / 32)][j % ((MAXDG + 31) / 32)];
  for (int k = tid; k < nd; k += 256) lab[k] = k;
  __syncthreads();
  for (int pr = tid; pr < nd * nd; pr += 256) {
    const int k = pr / nd, l = pr % nd;
    if (k >= l) continue;
    bool link = false;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) {
        double ax, ay, az, bx, by, bz;
        sphere_xyz(G, k + i * nd, ax, ay, az);
        sphere_xyz(G, l + j * nd, bx, by, bz);
        const double dx = pbc_half(ax - bx, P.box[0]), dy = pbc_half(ay - by, P.box[1]), dz = az - bz;
        const int kind = (i == 0 && j == 0) ? 0 : (i == 1 && j == 1) ? 2 : 1;
        const double lim = sqrt(P.mm_cut2[kind]) + 2.0 * P.guard;
        if (dx * dx + dy * dy + dz * dz <= lim * lim) link = true;
      }
    if (link) {
      atomicOr(&adj[k][l >> 5], 1u << (l & 31));
      atomicOr(&adj[l][k >> 5], 1u << (k & 31));
    }
  }
  __syncthreads();
  for (int it = 0; it <= nd; it++) {   // minimum-label propagation over the links
    if (tid == 0) changed = 0;
    __syncthreads();
    for (int k = tid; k < nd; k += 256) {
      int m = lab[k];
      for (int l = 0; l < nd; l++)
        if ((adj[k][l >> 5] >> (l & 31)) & 1u) m = min(m, lab[l]);
      nlab[k] = m;
      if (m != lab[k]) changed = 1;
    }
    __syncthreads();
    for (int k = tid; k < nd; k += 256) lab[k] = nlab[k];
    __syncthreads();
    if (!changed) break;
  }
  if (tid == 0) {
    int n = 0;
    for (int k = 0; k < nd; k++)
      if (lab[k] == k) { nlab[k] = n; ci->cnt[n] = 0; n++; }
    for (int k = 0; k < nd; k++) {
      const int c = nlab[lab[k]];
      ci->clu_of[k] = c;
      const int j = ci->cnt[c]++;
      ci->loc[k] = j;
      if (j < MAXD) ci->dims[c][j] = k; else raise_error(&G.sc->err, ERR_CLUSTER);
    }
    ci->nclu = n;
    ci->count = 0u;
    ci->abort = 0;
  }
  __syncthreads();
  const int nclu = ci->nclu;
  for (int c = tid; c < nclu; c += 256) {
    const int nl = min(ci->cnt[c], MAXD);
    CoP Pc = P;
    Pc.nd = nl;
    B.Ps[c] = Pc;
    CoD D = G;
    const size_t nq = (size_t)NE_NEAR * c, ni = (size_t)NE_CAP * c;
    D.sc = B.scs + c;
    D.dim = B.dimc + (size_t)c * 12 * MAXSG;
    D.near_list = B.near_list + nq;
    D.ns_x = B.ns_x + nq; D.ns_y = B.ns_y + nq; D.ns_z = B.ns_z + nq;
    D.ns_vx = B.ns_vx + nq; D.ns_vy = B.ns_vy + nq; D.ns_vz = B.ns_vz + nq;
    D.ns_flag = B.ns_flag + nq; D.ns_stat = B.ns_stat + nq;
    D.in_list = B.in_list + ni; D.in_plist = B.in_plist + ni; D.in_memb = B.in_memb + ni;
    D.in_fp = B.in_fp + ni * MAXM * 3; D.in_pot = B.in_pot + ni * MAXM; D.in_msd = B.in_msd + ni;
    D.sl_x = B.sl_x + ni; D.sl_y = B.sl_y + ni; D.sl_z = B.sl_z + ni;
    D.sl_vx = B.sl_vx + ni; D.sl_vy = B.sl_vy + ni; D.sl_vz = B.sl_vz + ni; D.sl_flag = B.sl_flag + ni;
    D.in_cap = NE_CAP;
    D.near_cap = NE_NEAR;
    D.cid = c;
    B.Ds[c] = D;
    // scalars: the system's, with this cluster's spheres at the interval start
    CoScalars* cs = B.scs + c;
    const CoScalars* gs = G.sc;
    cs->cstep = gs->cstep; cs->time_count = gs->time_count; cs->time_penetrate = gs->time_penetrate;
    cs->in_num = 0; cs->near_num = 0; cs->n_ev = gs->n_ev; cs->err = 0;
    for (int e = 0; e < MAXEV; e++) cs->events[e] = gs->events[e];
    cs->far_maxv2 = gs->far_maxv2; cs->near_vmax2 = gs->near_vmax2; cs->tpotential = 0.0;
    cs->coll = gs->coll;
    for (int e = 0; e < 8; e++) cs->prof[e] = gs->prof[e];
    for (int j = 0; j < nl; j++)
      for (int i = 0; i < 2; i++) {
        const int sg = ci->dims[c][j] + i * nd, sl = j + i * nl;
        cs->x0[sl] = gs->x0[sg]; cs->y0[sl] = gs->y0[sg]; cs->z0[sl] = gs->z0[sg];
      }
  }
  // local sphere fields
  for (int w = tid; w < nclu * 12 * MAXS; w += 256) {
    const int c = w / (12 * MAXS), f = (w / MAXS) % 12, sl = w % MAXS;
    const int nl = min(ci->cnt[c], MAXD);
    if (sl >= 2 * nl) continue;
    const int j = sl % nl, i = sl / nl;
    B.dimc[(size_t)c * 12 * MAXSG + f * MAXSG + sl] = G.dim[f * MAXSG + ci->dims[c][j] + i * nd];
  }
}

// per cluster (one workgroup each): its entries of the near list in ascending
// particle order, and the number of them that are list members
__global__ __launch_bounds__(1024) void clu_lists(CoD G, CluBase B) {
  const int c = blockIdx.x;
  if (c >= G.clu->nclu) return;
  __shared__ int wcnt[16], wins[16];
  __shared__ int base;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x == 0) base = 0;
  int ins = 0;
  const int nn = min(G.sc->near_num, G.near_cap);
  int* out = B.near_list + (size_t)NE_NEAR * c;
  __syncthreads();
  for (int q0 = 0; q0 < nn; q0 += 1024) {
    const int q = q0 + threadIdx.x;
    const bool mine = q < nn && G.clu->clu_of[B.assign[q]] == c;
    int p = mine ? G.near_list[q] : 0;
    const unsigned long long bl = __ballot(mine);
    if (lane == 0) wcnt[wv] = __popcll(bl);
    if (mine && (G.status[p] & ST_INSIDE)) ins++;
    __syncthreads();
    int off = base;
    for (int w = 0; w < wv; w++) off += wcnt[w];
    const int r = off + __popcll(bl & lanes_below());
    if (mine && r < NE_NEAR) out[r] = p;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int w = 0; w < 16; w++) t += wcnt[w];
      base += t;
    }
    __syncthreads();
  }
  ins = wave_sum_i(ins);
  if (lane == 0) wins[wv] = ins;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < 16; w++) t += wins[w];
    B.scs[c].near_num = base;
    B.scs[c].in_num = t;
  }
}

// ------------------------------------------------------------- global near engine
// Many dimers at any density: the near set of one system is split into
// contiguous chunks (ascending particle index) over G cooperative
// workgroups.  Every workgroup knows every sphere; a particle's list
// memberships are up to four global list ids 2k+i (ascending, 0xff = none),
// so the number of dimers is a run-time value.  Per MD step: each workgroup
// advances its own inside entries and sums their pair forces per sphere in a
// fixed order (per wave, then over waves); at a grid barrier workgroup 0 adds
// the workgroups' sums in workgroup order, integrates every dimer (one thread
// per sphere for the forces, one per dimer for RATTLE) and takes the rebuild
// decision; a second barrier hands the new sphere positions and the decision
// to all.  A rebuild compacts the inside list with a prefix over the
// workgroups' counts, so lists keep the reference's ascending order.
constexpr int GE_T = 512;
constexpr int GE_W = GE_T / 64;
constexpr int GE_LCAP = 768;    // inside entries of a workgroup staged in LDS
constexpr int GE_GMAX = 128;    // workgroups of one launch
constexpr int SG_CELLS = 1024;  // coarse cells of the sphere grid
constexpr int SG_CAP = 8;       // spheres listed per coarse cell (more: the cell scans all spheres)
constexpr int MM_CAP = 8;       // sphere-sphere candidates per sphere (more: all spheres)
struct GeSync {
  unsigned count, gen;   // gen: the last barrier epoch released (set to 0 before every launch)
  int abort, pad;
  unsigned arrive[GE_GMAX];   // the last barrier epoch each workgroup reached
  int cnt[GE_GMAX];
  unsigned long long vmax2[GE_GMAX];
  // decision block written by workgroup 0
  int rebuild, tc, tp, cstep, nev, pad2;
};
struct GeD {
  GeSync* sy;
  unsigned* memb_q;      // [near_cap] memberships of the near entries at the last rebuild
  unsigned* in_memb4;    // [in_cap] memberships of the inside entries
  double* in_fp4;        // [in_cap][4][3] pair forces by membership slot
  double* in_pot4;       // [in_cap][4]
  double* part;          // [2][GE_GMAX][4*MAXSG+1] per-workgroup sphere sums, max displacement (by step parity)
  unsigned* touch;       // [2][MAXSG][GE_GMAX/32] workgroups holding members of each sphere's list
  int in_cap, pad;
};
struct GeShared {
  double lx[GE_LCAP], ly[GE_LCAP], lz[GE_LCAP], lvx[GE_LCAP], lvy[GE_LCAP], lvz[GE_LCAP], lmsd[GE_LCAP];
  int lflag[GE_LCAP];
  unsigned lmemb[GE_LCAP];
  double dim[12][MAXSG];
  double wacc[GE_W][4 * MAXSG];
  double fin[4 * MAXSG];
  double tpp[MAXSG];
  int wsum[GE_W + 1];
  int cstep, tc, tp, in_num, nev, rebuild;
  int qb, qe, ib, ie;
  unsigned ep;   // grid barriers passed in this launch
  int rc[GE_GMAX];         // every workgroup's published count and max (ge_read_counts)
  double rv[GE_GMAX];
  double vmax2, red[GE_W];
  unsigned long long tm[8];   // real-time ticks by phase (workgroup 0 reports them)
  unsigned touched[2][(MAXSG + 31) / 32];   // spheres this workgroup published, by step parity
  // coarse grid of the spheres (positions at the launch start, radius
  // full + guard) and each sphere's sphere-sphere candidates
  unsigned char sg_ids[SG_CELLS][SG_CAP];
  unsigned char sg_n[SG_CELLS];
  unsigned char mm_ids[MAXSG][MM_CAP];
  unsigned char mm_n[MAXSG];
  int sg_nx, sg_ny, sg_nz;
  double sg_h;
};
__shared__ GeShared g_ge;
// phase timers: 1 rebuild, 2 particle phase, 3 dimers, 5 barrier waits, 6 totals over workgroups
#define GE_T0() const unsigned long long t0_ = __builtin_amdgcn_s_memrealtime()
#define GE_T1(k) if (threadIdx.x == 0) g_ge.tm[k] += __builtin_amdgcn_s_memrealtime() - t0_

__device__ __forceinline__ int wave_min_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, MPCX_WAVE));
  return v;
}

// Grid barrier of the G workgroups, run by wave 0 of each: every workgroup
// stores the barrier's epoch in its own arrival word, wave 0 of workgroup 0
// polls all of them (two words a lane) and then releases the epoch through
// gen, which the others poll.  The data the workgroups exchange across it
// across the barriers of the MD steps (sphere sums, touch bits) are
// agent-scope atomics (written through L2, read past it), and between list
// rebuilds every entry and slot keeps its owner workgroup; so a step's
// barrier is light: every wave waits for its stores to complete (s_waitcnt 0,
// the wait an agent-scope release makes after its L2 write-back; a
// workgroup-scope release fence emits none on gfx950) before the arrival, and
// no L2 write-back or invalidate follows.  The barriers of a rebuild and of
// the interval start, where entries change owners, keep the agent-scope
// fences: without the write-back, a line the old owner left dirty in its
// XCD's L2 could be evicted over the new owner's data.  With
// agent-scope fences (L2 write-back and invalidate on every XCD) a barrier
// of 128 workgroups measured 13 us, with one counter or with these words;
// with workgroup-scope fences, 4 us.  Bounded like clu_exchange_max.
__device__ __attribute__((noinline)) void ge_barrier_wave(GeSync* sy, int G, unsigned ep, int* err) {
  const int lane = threadIdx.x & 63;
  if (G <= 1) return;
  if (lane == 0) __hip_atomic_store(&sy->arrive[blockIdx.x], ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (blockIdx.x == 0) {
    long long spins = 0;
    while (true) {
      bool ok = true;
      for (int g = lane; g < G; g += 64)
        ok = ok && (int)(__hip_atomic_load(&sy->arrive[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - ep) >= 0;
      if (__all(ok)) break;
      if (__hip_atomic_load(&sy->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > SPIN_CAP) {
        if (lane == 0) { raise_error(err, ERR_BARRIER); atomicExch(&sy->abort, 1); }
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (lane == 0) __hip_atomic_store(&sy->gen, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (lane == 0) {
    long long spins = 0;
    while ((int)(__hip_atomic_load(&sy->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - ep) < 0) {
      if (__hip_atomic_load(&sy->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > SPIN_CAP) {
        raise_error(err, ERR_BARRIER);
        atomicExch(&sy->abort, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
}
__device__ __forceinline__ void ge_sync(const GeD& E, int* err, bool light = false) {
  if (light) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // keeps the compiler's order
    __builtin_amdgcn_s_waitcnt(0);                           // this wave's stores have completed
  } else {
    __threadfence();
  }
  __syncthreads();
  const unsigned ep = g_ge.ep + 1u;
  GE_T0();
  if (threadIdx.x < 64) ge_barrier_wave(E.sy, (int)gridDim.x, ep, err);
  GE_T1(5);
  __syncthreads();
  if (light) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  else __threadfence();
  if (threadIdx.x == 0) g_ge.ep = ep;   // read again only after the next barrier's first __syncthreads
}

__device__ __forceinline__ int sg_cell(double x, double y, double z) {
  const double h = g_ge.sg_h;
  const int ix = min(max((int)floor(x / h), 0), g_ge.sg_nx - 1);
  const int iy = min(max((int)floor(y / h), 0), g_ge.sg_ny - 1);
  const int iz = min(max((int)floor(z / h), 0), g_ge.sg_nz - 1);
  return ix + g_ge.sg_nx * (iy + g_ge.sg_ny * iz);
}
// list b = 2k+i holds the particle (neighbour_list_mod.f95:38-57)
__device__ __forceinline__ bool ge_in_list(double x, double y, double z, const CoP& P, int b) {
  const int k = b >> 1, i = b & 1, s = k + i * P.nd;
  double d1 = pbc_half(x - g_ge.dim[DX][s], P.box[0]);
  double d2 = pbc_half(y - g_ge.dim[DY][s], P.box[1]);
  double d3 = z - g_ge.dim[DZ][s];
  return d1 * d1 + d2 * d2 + d3 * d3 <= P.full2[i];
}
// the lists holding a particle, ascending b, from its coarse cell's candidates
__device__ __forceinline__ unsigned ge_memberships(double x, double y, double z, const CoP& P, int* err) {
  unsigned ids = 0xffffffffu;
  int m = 0;
  const int c = sg_cell(x, y, z);
  const int n = g_ge.sg_n[c];
  const int nb = n <= SG_CAP ? n : 2 * P.nd;
  for (int j = 0; j < nb; j++) {
    const int b = n <= SG_CAP ? (int)g_ge.sg_ids[c][j] : j;
    if (ge_in_list(x, y, z, P, b)) {
      if (m == 4) { raise_error(err, ERR_MEMBERSHIPS); return ids; }
      ids = (ids & ~(0xffu << (8 * m))) | ((unsigned)b << (8 * m));
      m++;
    }
  }
  return ids;
}
// per launch, in every workgroup: the coarse grid and the sphere-sphere
// candidates; spheres move at most the guard in an interval
__device__ __forceinline__ void ge_build_grids(const CoP& P) {
  const int nd = P.nd, tid = threadIdx.x;
  if (tid == 0) {
    double h = 8.0;
    int nx, ny, nz;
    for (;;) {
      nx = (int)ceil(P.box[0] / h); ny = (int)ceil(P.box[1] / h); nz = (int)ceil(P.box[2] / h);
      if ((long long)nx * ny * nz <= SG_CELLS) break;
      h *= 1.25;
    }
    g_ge.sg_h = h; g_ge.sg_nx = nx; g_ge.sg_ny = ny; g_ge.sg_nz = nz;
  }
  __syncthreads();
  const double h = g_ge.sg_h;
  const int ncell = g_ge.sg_nx * g_ge.sg_ny * g_ge.sg_nz;
  for (int c = tid; c < ncell; c += GE_T) {
    const int ix = c % g_ge.sg_nx, iy = (c / g_ge.sg_nx) % g_ge.sg_ny, iz = c / (g_ge.sg_nx * g_ge.sg_ny);
    const double cx = (ix + 0.5) * h, cy = (iy + 0.5) * h, cz = (iz + 0.5) * h;
    int n = 0;
    for (int b = 0; b < 2 * nd; b++) {
      const int k = b >> 1, i = b & 1, s = k + i * nd;
      const double r = P.full[i] + P.guard + 1e-6;
      const double ax = fmax(fabs(pbc_half(g_ge.dim[DX][s] - cx, P.box[0])) - 0.5 * h, 0.0);
      const double ay = fmax(fabs(pbc_half(g_ge.dim[DY][s] - cy, P.box[1])) - 0.5 * h, 0.0);
      const double az = fmax(fabs(g_ge.dim[DZ][s] - cz) - 0.5 * h, 0.0);
      if (ax * ax + ay * ay + az * az <= r * r) {
        if (n < SG_CAP) g_ge.sg_ids[c][n] = (unsigned char)b;
        n++;
      }
    }
    g_ge.sg_n[c] = (unsigned char)min(n, 255);
  }
  for (int s = tid; s < 2 * nd; s += GE_T) {   // every workgroup integrates the dimers
      const int k = s < nd ? s : s - nd, i = s < nd ? 0 : 1;
      int n = 0;
      for (int j = 0; j < nd; j++)
        for (int l = 0; l < 2; l++) {
          if (j == k) continue;
          const int b = j + l * nd;
          const int kind = (i == 0 && l == 0) ? 0 : (i == 1 && l == 1) ? 2 : 1;
          const double r = sqrt(P.mm_cut2[kind]) + 2.0 * P.guard + 1e-6;
          const double dx = pbc_half(g_ge.dim[DX][s] - g_ge.dim[DX][b], P.box[0]);
          const double dy = pbc_half(g_ge.dim[DY][s] - g_ge.dim[DY][b], P.box[1]);
          const double dz = g_ge.dim[DZ][s] - g_ge.dim[DZ][b];
          if (dx * dx + dy * dy + dz * dz <= r * r) {
            if (n < MM_CAP) g_ge.mm_ids[s][n] = (unsigned char)(2 * j + l);
            n++;
          }
        }
      g_ge.mm_n[s] = (unsigned char)min(n, 255);
    }
  __syncthreads();
}

__device__ __forceinline__ void ge_slot_load(const CoD& D, int k, double& x, double& y, double& z, double& vx,
                                             double& vy, double& vz, int& fl, double& ms) {
  const int kl = k - g_ge.ib;
  if (kl < GE_LCAP) {
    x = g_ge.lx[kl]; y = g_ge.ly[kl]; z = g_ge.lz[kl]; vx = g_ge.lvx[kl]; vy = g_ge.lvy[kl]; vz = g_ge.lvz[kl];
    fl = g_ge.lflag[kl]; ms = g_ge.lmsd[kl];
  } else {
    x = D.sl_x[k]; y = D.sl_y[k]; z = D.sl_z[k]; vx = D.sl_vx[k]; vy = D.sl_vy[k]; vz = D.sl_vz[k];
    fl = D.sl_flag[k]; ms = D.in_msd[k];
  }
}
__device__ __forceinline__ void ge_slot_store(const CoD& D, int k, double x, double y, double z, double vx,
                                              double vy, double vz, int fl, double ms) {
  const int kl = k - g_ge.ib;
  if (kl < GE_LCAP) {
    g_ge.lx[kl] = x; g_ge.ly[kl] = y; g_ge.lz[kl] = z; g_ge.lvx[kl] = vx; g_ge.lvy[kl] = vy; g_ge.lvz[kl] = vz;
    g_ge.lflag[kl] = fl; g_ge.lmsd[kl] = ms;
  } else {
    D.sl_x[k] = x; D.sl_y[k] = y; D.sl_z[k] = z; D.sl_vx[k] = vx; D.sl_vy[k] = vy; D.sl_vz[k] = vz;
    D.sl_flag[k] = fl; D.in_msd[k] = ms;
  }
}
// near slot q -> entry k of this workgroup
__device__ __forceinline__ void ge_stage(const CoD& D, int k, int q, double ms) {
  ge_slot_store(D, k, D.ns_x[q], D.ns_y[q], D.ns_z[q], D.ns_vx[q], D.ns_vy[q], D.ns_vz[q], D.ns_flag[q], ms);
}
// this workgroup's entries -> their near slots
__device__ __forceinline__ void ge_unstage(const CoD& D) {
  for (int k = g_ge.ib + threadIdx.x; k < g_ge.ie; k += GE_T) {
    double x, y, z, vx, vy, vz, ms;
    int fl;
    ge_slot_load(D, k, x, y, z, vx, vy, vz, fl, ms);
    const int q = D.in_list[k];
    D.ns_x[q] = x; D.ns_y[q] = y; D.ns_z[q] = z; D.ns_vx[q] = vx; D.ns_vy[q] = vy; D.ns_vz[q] = vz;
    D.ns_flag[q] = fl;
    if (k - g_ge.ib < GE_LCAP) D.in_msd[k] = ms;
  }
}

// exclusive prefix of a per-workgroup count over the grid (after a barrier)
__device__ __forceinline__ void ge_publish_count(const GeD& E, int c, double vm) {
  if (threadIdx.x == 0) {
    __hip_atomic_store(&E.sy->cnt[blockIdx.x], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&E.sy->vmax2[blockIdx.x], (unsigned long long)__double_as_longlong(vm), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}
// (every thread together: one load per workgroup's words, one thread each,
// where every thread loading all of them in turn waited on each)
__device__ __forceinline__ void ge_read_counts(const GeD& E, int& before, int& total, double& vmax) {
  const int G = (int)gridDim.x;
  if ((int)threadIdx.x < G) {
    g_ge.rc[threadIdx.x] = __hip_atomic_load(&E.sy->cnt[threadIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    g_ge.rv[threadIdx.x] = __longlong_as_double((long long)__hip_atomic_load(&E.sy->vmax2[threadIdx.x],
                                                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  }
  __syncthreads();
  before = 0; total = 0; vmax = 0.0;
  for (int w = 0; w < G; w++) {
    const int c = g_ge.rc[w];
    if (w < (int)blockIdx.x) before += c;
    total += c;
    vmax = fmax(vmax, g_ge.rv[w]);
  }
  __syncthreads();   // the words are read again by the next call
}

// ordered compaction of this workgroup's near entries with pred(q) into
// entries [base, ...): wave w owns [qb + w*64*E, ...), lanes interleaved
template <typename Pred, typename Emit>
__device__ __forceinline__ void ge_compact_chunk(int base, Pred pred, Emit emit) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = g_ge.qe - g_ge.qb;
  const int E = (n + GE_T - 1) / GE_T;
  const int q0 = g_ge.qb + wv * 64 * E + lane;
  int cnt = 0;
  for (int j = 0; j < E; j++) {
    const int q = q0 + j * 64;
    cnt += __popcll(__ballot(q < g_ge.qe && pred(q)));
  }
  if (lane == 0) g_ge.wsum[wv] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < GE_W; w++) { int t = g_ge.wsum[w]; g_ge.wsum[w] = run; run += t; }
    g_ge.wsum[GE_W] = run;
  }
  __syncthreads();
  int b = base + g_ge.wsum[wv];
  const unsigned long long lt = lanes_below();
  for (int j = 0; j < E; j++) {
    const int q = q0 + j * 64;
    const bool in = q < g_ge.qe && pred(q);
    const unsigned long long bl = __ballot(in);
    const int k = b + __popcll(bl & lt);
    b += __popcll(bl);
    if (in) emit(q, k);
  }
  __syncthreads();
}
__device__ __forceinline__ int ge_count_chunk_pred_total() { return g_ge.wsum[GE_W]; }

// Rebuild over this workgroup's chunk: lazy streaming of its near particles
// outside the lists (when do_stream), memberships, then the ordered
// compaction of all workgroups' members into the inside list.
__device__ __attribute__((noinline)) void ge_rebuild(const CoP& P, const CoD& D, const GeD& E, bool store_first,
                                                     bool do_stream, bool compact) {
  if (store_first) {
    ge_unstage(D);
    __threadfence_block();
    __syncthreads();
  }
  const double tc = (double)g_ge.tc;
  double vm = 0.0;
  int* err = &D.sc->err;
  for (int q = g_ge.qb + threadIdx.x; q < g_ge.qe; q += GE_T) {
    double x = D.ns_x[q], y = D.ns_y[q], z = D.ns_z[q], vx = D.ns_vx[q], vy = D.ns_vy[q], vz = D.ns_vz[q];
    int f = D.ns_flag[q];
    if (do_stream && !(D.ns_stat[q] & ST_INSIDE)) {
      const int f0 = f;
      bool b = stream_particle(x, y, z, vx, vy, vz, f, tc, P, D.react_box, err);
      D.ns_x[q] = x; D.ns_y[q] = y; D.ns_z[q] = z;
      if (b) { D.ns_vx[q] = vx; D.ns_vy[q] = vy; D.ns_vz[q] = vz; }
      if (f != f0) D.ns_flag[q] = f;
    }
    vm = fmax(vm, (vx * vx + vy * vy) + vz * vz);
    if (compact) E.memb_q[q] = ge_memberships(x, y, z, P, err);
  }
  vm = wave_max(vm);
  if ((threadIdx.x & 63) == 0) g_ge.red[threadIdx.x >> 6] = vm;
  __threadfence_block();
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < GE_W; w++) m = fmax(m, g_ge.red[w]);
    if (do_stream) g_ge.vmax2 = m;
  }
  if (!compact) {   // stopping for a sort: the system-wide max v^2 is kept for the resumed rebuild
    __syncthreads();
    if (threadIdx.x == 0) ge_publish_count(E, 0, g_ge.vmax2);
    ge_sync(E, err);
    int before, total;
    double vmax;
    ge_read_counts(E, before, total, vmax);
    if (threadIdx.x == 0) g_ge.vmax2 = vmax;
    __syncthreads();
    return;
  }
  // members of this chunk, then their place in the system's inside list
  const auto member = [&](int q) { return E.memb_q[q] != 0xffffffffu; };
  int cnt = 0;
  for (int q = g_ge.qb + threadIdx.x; q < g_ge.qe; q += GE_T) cnt += member(q) ? 1 : 0;
  cnt = wave_sum_i(cnt);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) g_ge.wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < GE_W; w++) t += g_ge.wsum[w];
    ge_publish_count(E, t, g_ge.vmax2);
  }
  ge_sync(E, err);
  int before, total;
  double vmax;
  ge_read_counts(E, before, total, vmax);
  if (total > E.in_cap) {
    if (threadIdx.x == 0) raise_error(err, ERR_INSIDE_CAP);
    total = E.in_cap;
  }
  if (threadIdx.x == 0) {
    g_ge.ib = min(before, E.in_cap);
    g_ge.in_num = total;
    g_ge.vmax2 = vmax;
    double fm = __longlong_as_double((long long)D.sc->far_maxv2);
    double maxt = fmax(fm, vmax);
    g_ge.tp = (int)floor(((0.5 * P.skinrad) / sqrt(maxt)) / P.MDt);
    g_ge.tc = 0;
  }
  __syncthreads();
  ge_compact_chunk(g_ge.ib, member, [&](int q, int k) {
    if (k < E.in_cap) {
      D.in_list[k] = q;
      E.in_memb4[k] = E.memb_q[q];
      if (k - g_ge.ib < GE_LCAP) g_ge.lmemb[k - g_ge.ib] = E.memb_q[q];
      ge_stage(D, k, q, 0.0);
    }
  });
  for (int q = g_ge.qb + threadIdx.x; q < g_ge.qe; q += GE_T) D.ns_stat[q] = member(q) ? ST_INSIDE : 0;
  if (threadIdx.x == 0) g_ge.ie = min(g_ge.ib + g_ge.wsum[GE_W], E.in_cap);
  __threadfence_block();
  __syncthreads();
}

// Interval start: the inside list (unchanged since the last rebuild) as near
// indices of this interval's set; entry k keeps its memberships, forces and
// max_solvent_dist in the system's entry arrays.
__device__ __attribute__((noinline)) void ge_map_inside(const CoD& D, const GeD& E) {
  const auto ins = [&](int q) { return (D.ns_stat[q] & ST_INSIDE) != 0; };
  int cnt = 0;
  for (int q = g_ge.qb + threadIdx.x; q < g_ge.qe; q += GE_T) cnt += ins(q) ? 1 : 0;
  cnt = wave_sum_i(cnt);
  if ((threadIdx.x & 63) == 0) g_ge.wsum[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < GE_W; w++) t += g_ge.wsum[w];
    ge_publish_count(E, t, 0.0);
  }
  ge_sync(E, &D.sc->err);
  int before, total;
  double vm;
  ge_read_counts(E, before, total, vm);
  if (threadIdx.x == 0) {
    if (total != g_ge.in_num) raise_error(&D.sc->err, ERR_INSIDE_LOST);
    g_ge.ib = min(before, E.in_cap);
  }
  __syncthreads();
  ge_compact_chunk(g_ge.ib, ins, [&](int q, int k) {
    if (k < E.in_cap) {
      D.in_list[k] = q;
      const int kl = k - g_ge.ib;
      ge_stage(D, k, q, kl < GE_LCAP ? D.in_msd[k] : D.in_msd[k]);
      if (kl < GE_LCAP) g_ge.lmemb[kl] = E.in_memb4[k];
    }
  });
  if (threadIdx.x == 0) g_ge.ie = min(g_ge.ib + g_ge.wsum[GE_W], E.in_cap);
  __syncthreads();
}

__device__ __forceinline__ unsigned ge_memb_of(const GeD& E, int k) {
  const int kl = k - g_ge.ib;
  return kl < GE_LCAP ? g_ge.lmemb[kl] : E.in_memb4[k];
}

// kick with up to four pair forces (list order)
__device__ __forceinline__ void ge_kick(const CoP& P, int fl, unsigned ids, const double* fx, const double* fy,
                                        const double* fz, double& vx, double& vy, double& vz) {
  bool ok;
  const double h = kick_factor(P, fl, ok);
  if (!ok) return;
#pragma unroll
  for (int m = 0; m < 4; m++) {
    if (((ids >> (8 * m)) & 0xffu) == 0xffu) break;
    vx = vx + h * fx[m];
    vy = vy + h * fy[m];
    vz = vz + h * fz[m];
  }
}

// Particle phase of this workgroup's entries (forces_mod.f95:62-136, kick,
// and unless the last step the next kick and streaming); per-sphere sums
// into wacc by wave in a fixed order.
__device__ __attribute__((noinline)) double ge_particles(const CoP& P, const CoD& D, const GeD& E, bool last) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nd = P.nd;
  double mx = 0.0;
  const int ib = g_ge.ib, ie = g_ge.ie;
  for (int k0 = ib + wv * 64; k0 < ie; k0 += GE_T) {
    const int k = k0 + lane;
    const bool valid = k < ie;
    double fx[4] = {0, 0, 0, 0}, fy[4] = {0, 0, 0, 0}, fz[4] = {0, 0, 0, 0}, pot[4] = {0, 0, 0, 0};
    unsigned ids = 0xffffffffu;
    if (valid) {
      double x, y, z, vx, vy, vz, ms;
      int fl;
      ge_slot_load(D, k, x, y, z, vx, vy, vz, fl, ms);
      ids = ge_memb_of(E, k);
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const unsigned b = (ids >> (8 * m)) & 0xffu;
        if (b == 0xffu) break;
        const int kk = (int)(b >> 1), ii = (int)(b & 1u), s = kk + ii * nd;
        double d1 = pbc_half(x - g_ge.dim[DX][s], P.box[0]);
        double d2 = pbc_half(y - g_ge.dim[DY][s], P.box[1]);
        double d3 = z - g_ge.dim[DZ][s];
        double dRsq = d1 * d1 + d2 * d2 + d3 * d3;
        if (dRsq <= P.cut2[ii]) {
          double En = 0.0;
          bool on = false;
          if (fl == 0) {
            if (ii == 0) { fl = 1; En = P.Eb[0]; } else En = P.Ea[1];
            on = true;
          } else if (fl == 1) { En = P.Eb[ii]; on = true; }
          else if (fl == 2) { En = P.Es[ii]; on = true; }
          if (on) {
            double F, Ub;
            lj_terms(dRsq, P.R6[ii], P.R12[ii], P.r12R12[ii], P.r6R6[ii], F, Ub);
            double EF = En * F;
            fx[m] = EF * d1; fy[m] = EF * d2; fz[m] = EF * d3; pot[m] = En * Ub;
          }
        }
      }
      ge_kick(P, fl, ids, fx, fy, fz, vx, vy, vz);
      if (!last) {
        ge_kick(P, fl, ids, fx, fy, fz, vx, vy, vz);
        stream_particle(x, y, z, vx, vy, vz, fl, 1.0, P, D.react_box, &D.sc->err);
        double v2 = (vx * vx + vy * vy) + vz * vz;
        ms = ms + v2 * (P.MDt * P.MDt);
        mx = fmax(mx, ms);
      } else {
        for (int m = 0; m < 4; m++) {
          E.in_fp4[(size_t)k * 12 + m * 3 + 0] = fx[m];
          E.in_fp4[(size_t)k * 12 + m * 3 + 1] = fy[m];
          E.in_fp4[(size_t)k * 12 + m * 3 + 2] = fz[m];
          E.in_pot4[(size_t)k * 4 + m] = pot[m];
        }
      }
      ge_slot_store(D, k, x, y, z, vx, vy, vz, fl, ms);
    }
    // per-sphere sums of this wave's 64 entries, lists in ascending order
    int ptr = 0;
    for (;;) {
      const unsigned b = ptr < 4 ? (ids >> (8 * ptr)) & 0xffu : 0xffu;
      const int key = b == 0xffu ? 0x7fffffff : (int)b;
      const int bmin = wave_min_i(key);
      if (bmin == 0x7fffffff) break;
      const int smin = (bmin >> 1) + (bmin & 1) * nd;
      const bool mine = key == bmin;
      double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0;
      if (mine) {
#pragma unroll
        for (int m = 0; m < 4; m++)
          if (m == ptr) { c0 = fx[m]; c1 = fy[m]; c2 = fz[m]; c3 = pot[m]; }
      }
      c0 = wave_sum(c0); c1 = wave_sum(c1); c2 = wave_sum(c2); c3 = wave_sum(c3);
      if (lane == 0) {
        double* a = &g_ge.wacc[wv][4 * smin];
        a[0] = a[0] + c0; a[1] = a[1] + c1; a[2] = a[2] + c2; a[3] = a[3] + c3;
      }
      if (mine) ptr++;
    }
  }
  return mx;
}

// thread s < 2 nd of workgroup 0: forces on sphere s (evolve_mod.f95:605-636,
// the order of dimer_forces_and_rattle_v)
__device__ __forceinline__ void ge_sphere_force(const CoP& P, int s) {
  const int nd = P.nd, k = s < nd ? s : s - nd, i = s < nd ? 0 : 1;
  double fw, wp;
  lj_mot_wall(P, g_ge.dim[DZ][s], i, fw, wp);
  double fx = 0.0 - g_ge.fin[4 * s + 0], fy = 0.0 - g_ge.fin[4 * s + 1], fz = 0.0 - g_ge.fin[4 * s + 2];
  double tp = 0.0 + g_ge.fin[4 * s + 3];
  // partners beyond the sphere-sphere cutoff add exact zeros; only the
  // candidates within it (plus the guard each may move) are visited, in order
  const int nc = g_ge.mm_n[s];
  const int nj = nc <= MM_CAP ? nc : 2 * nd;
  for (int q = 0; q < nj; q++) {
      const int jl = nc <= MM_CAP ? (int)g_ge.mm_ids[s][q] : q;
      const int j = jl >> 1, l = jl & 1;
      if (j == k) continue;
      const int b = j + l * nd;
      const int kind = (i == 0 && l == 0) ? 0 : (i == 1 && l == 1) ? 2 : 1;
      double mx, my, mz, mp;
      lj_mot_mot(P, g_ge.dim[DX][s], g_ge.dim[DY][s], g_ge.dim[DZ][s], g_ge.dim[DX][b], g_ge.dim[DY][b],
                 g_ge.dim[DZ][b], kind, mx, my, mz, mp);
      fx = fx + mx; fy = fy + my; fz = fz + mz;
      if (j < k) tp = tp + mp;
    }
  fz = fz + fw;
  tp = tp + wp;
  g_ge.dim[DFX][s] = fx; g_ge.dim[DFY][s] = fy; g_ge.dim[DFZ][s] = fz;
  g_ge.tpp[s] = tp;
}
// thread k < nd of workgroup 0: one dimer through rattle_positions / the
// velocity RATTLE of dimer_forces_and_rattle_v (the same arithmetic)
__device__ __forceinline__ void ge_dimer_load(Dimer<1>& d, int k, int nd) {
  for (int f = 0; f < 12; f++) { d.v[f][0] = g_ge.dim[f][k]; d.v[f][1] = g_ge.dim[f][k + nd]; }
}
__device__ __forceinline__ void ge_dimer_store(const Dimer<1>& d, int k, int nd) {
  for (int f = 0; f < 12; f++) { g_ge.dim[f][k] = d.v[f][0]; g_ge.dim[f][k + nd] = d.v[f][1]; }
}
__device__ __forceinline__ void ge_rattle_v(const CoP& P, Dimer<1>& d, int* err) {
  const int s0 = 0, s1 = 1;
  d.v[DVX][s0] = d.v[DQX][s0] + P.hc * d.v[DFX][s0];
  d.v[DVY][s0] = d.v[DQY][s0] + P.hc * d.v[DFY][s0];
  d.v[DVZ][s0] = d.v[DQZ][s0] + P.hc * d.v[DFZ][s0];
  d.v[DVX][s1] = d.v[DQX][s1] + P.hn * d.v[DFX][s1];
  d.v[DVY][s1] = d.v[DQY][s1] + P.hn * d.v[DFY][s1];
  d.v[DVZ][s1] = d.v[DQZ][s1] + P.hn * d.v[DFZ][s1];
  for (long long acc = 0;; acc++) {
    if (acc == RATTLE_CAP) { raise_error(err, ERR_RATTLE); break; }
    double dvx = d.v[DVX][s0] - d.v[DVX][s1], dvy = d.v[DVY][s0] - d.v[DVY][s1], dvz = d.v[DVZ][s0] - d.v[DVZ][s1];
    double d0 = pbc_half(d.v[DX][s0] - d.v[DX][s1], P.box[0]);
    double d1 = pbc_half(d.v[DY][s0] - d.v[DY][s1], P.box[1]);
    double d2 = d.v[DZ][s0] - d.v[DZ][s1];
    double c2 = 2.0 * (dvx * d0 + dvy * d1 + dvz * d2);
    if (fabs(c2) < 1e-8) break;
    double O = -((0.25 * (P.IIMcIMn * c2)) / P.Dcn2);
    double o2 = 2.0 * O;
    d.v[DVX][s0] = d.v[DVX][s0] + (o2 * d0) / P.Mc;
    d.v[DVY][s0] = d.v[DVY][s0] + (o2 * d1) / P.Mc;
    d.v[DVZ][s0] = d.v[DVZ][s0] + (o2 * d2) / P.Mc;
    d.v[DVX][s1] = d.v[DVX][s1] - (o2 * d0) / P.Mn;
    d.v[DVY][s1] = d.v[DVY][s1] - (o2 * d1) / P.Mn;
    d.v[DVZ][s1] = d.v[DVZ][s1] - (o2 * d2) / P.Mn;
  }
}

// workgroup 0: the dimers at the end of step t (forces, velocity RATTLE) and,
// unless last, the start of step t+1 (positions RATTLE, rebuild decision)
__device__ __attribute__((noinline)) void ge_dimers(const CoP& P, const CoD& D, const GeD& E, bool end_step,
                                                    bool begin_step, double maxv) {
  const int nd = P.nd, tid = threadIdx.x;
  int* err = &D.sc->err;
  if (end_step) {
    for (int s = tid; s < 2 * nd; s += GE_T) ge_sphere_force(P, s);
    __syncthreads();
    if (tid == 0 && blockIdx.x == 0) {
      double tpot = 0.0;
      for (int k = 0; k < nd; k++)
        for (int i = 0; i < 2; i++) tpot = tpot + g_ge.tpp[k + i * nd];
      D.sc->tpotential = tpot;
    }
    for (int k = tid; k < nd; k += GE_T) {
      Dimer<1> d;
      ge_dimer_load(d, k, nd);
      ge_rattle_v(P, d, err);
      ge_dimer_store(d, k, nd);
    }
    __syncthreads();
  }
  if (!begin_step) return;
  if (tid == 0) { g_ge.cstep += 1; g_ge.tc += 1; }
  for (int k = tid; k < nd; k += GE_T) {
    Dimer<1> d;
    ge_dimer_load(d, k, nd);
    rattle_positions<1>(P, d, err);
    ge_dimer_store(d, k, nd);
  }
  __syncthreads();
  if (tid == 0) {
    const int rb = (maxv > P.hs2 || g_ge.tc == g_ge.tp || g_ge.cstep % P.MPCtstep == 0);
    g_ge.rebuild = rb;
    if (rb) {
      if (g_ge.nev >= MAXEV) raise_error(err, ERR_EVENTS);
      else if (blockIdx.x == 0) D.sc->events[g_ge.nev] = g_ge.tc;
      g_ge.nev += 1;
    }
  }
  __syncthreads();
  if (g_ge.rebuild)
    for (int s = tid; s < 2 * nd; s += GE_T) {
      double a = pbc_half(g_ge.dim[DX][s] - D.sc->x0[s], P.box[0]);
      double b = pbc_half(g_ge.dim[DY][s] - D.sc->y0[s], P.box[1]);
      double c = g_ge.dim[DZ][s] - D.sc->z0[s];
      if (a * a + b * b + c * c > P.guard * P.guard) raise_error(err, ERR_NEAR_GUARD);
    }
}

// Every workgroup integrates every dimer (the same arithmetic on the same
// totals, so all hold the same dimer state and decisions) after one barrier
// per MD step.  A workgroup publishes the per-sphere sums of its entries only
// for the spheres it holds members of, with a bit in that sphere's word of
// workgroups; the totals add the published sums in workgroup order.  Sums,
// bits and the max displacement are double-buffered by step parity: a buffer
// is rewritten two barriers after it was read, and a workgroup clears its
// bits of a buffer after the barrier that follows their readers.
__device__ __forceinline__ void ge_post(const CoP& P, const GeD& E, double mx, int par) {
  const int nd = P.nd, G = (int)gridDim.x;
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) g_ge.red[threadIdx.x >> 6] = mx;
  __syncthreads();
  double* part = E.part + ((size_t)par * GE_GMAX + blockIdx.x) * (4 * MAXSG + 1);
  unsigned* tb = E.touch + (size_t)par * MAXSG * (GE_GMAX / 32);
  for (int sph = threadIdx.x; sph < 2 * nd; sph += GE_T) {
    double a[4];
    bool any = false;
    for (int c = 0; c < 4; c++) {
      const int v = 4 * sph + c;
      double t = g_ge.wacc[0][v];
      g_ge.wacc[0][v] = 0.0;
      for (int w = 1; w < GE_W; w++) { t = t + g_ge.wacc[w][v]; g_ge.wacc[w][v] = 0.0; }
      a[c] = t;
      any = any || t != 0.0;
    }
    if (any) {   // read by other workgroups after the barrier: agent-scope stores
      for (int c = 0; c < 4; c++)
        __hip_atomic_store((unsigned long long*)&part[4 * sph + c], (unsigned long long)__double_as_longlong(a[c]),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      atomicOr(&tb[sph * (GE_GMAX / 32) + (blockIdx.x >> 5)], 1u << (blockIdx.x & 31));
      atomicOr(&g_ge.touched[par][sph >> 5], 1u << (sph & 31));
    }
  }
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < GE_W; w++) m = fmax(m, g_ge.red[w]);
    __hip_atomic_store((unsigned long long*)&part[4 * MAXSG], (unsigned long long)__double_as_longlong(m),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  (void)G;
}
// after the barrier: totals of every sphere into fin, this workgroup's bits of
// the other parity cleared; returns the system's max displacement
__device__ __forceinline__ double ge_totals(const CoP& P, const GeD& E, int par) {
  const int nd = P.nd, nv = 4 * 2 * nd, G = (int)gridDim.x;
  const double* part = E.part + (size_t)par * GE_GMAX * (4 * MAXSG + 1);
  const unsigned* tb = E.touch + (size_t)par * MAXSG * (GE_GMAX / 32);
  // every word of the sphere's bits, then up to 8 of its workgroups' sums at
  // a time, loaded before any is added (each load crosses L2 to memory), and
  // added in workgroup order as before
  constexpr int NW = GE_GMAX / 32, NB = 8;
  for (int v = threadIdx.x; v < nv; v += GE_T) {
    const int sph = v >> 2;
    unsigned bw[NW];
#pragma unroll
    for (int wd = 0; wd < NW; wd++)
      bw[wd] = __hip_atomic_load(&tb[sph * NW + wd], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double a = 0.0;
#pragma unroll
    for (int wd = 0; wd < NW; wd++) {
      unsigned bits = bw[wd];
      while (bits) {
        double val[NB];
        int cnt = 0;
#pragma unroll
        for (int j = 0; j < NB; j++) {
          val[j] = 0.0;
          if (bits) {
            const int w = wd * 32 + __ffs(bits) - 1;
            bits &= bits - 1u;
            val[j] = __longlong_as_double((long long)__hip_atomic_load(
                (unsigned long long*)&part[(size_t)w * (4 * MAXSG + 1) + v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            cnt = j + 1;
          }
        }
#pragma unroll
        for (int j = 0; j < NB; j++)
          if (j < cnt) a = a + val[j];
      }
    }
    g_ge.fin[v] = a;
  }
  double m = 0.0;
  for (int w = threadIdx.x; w < G; w += GE_T)
    m = fmax(m, __longlong_as_double((long long)__hip_atomic_load(
                    (unsigned long long*)&part[(size_t)w * (4 * MAXSG + 1) + 4 * MAXSG], __ATOMIC_RELAXED,
                    __HIP_MEMORY_SCOPE_AGENT)));
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) g_ge.red[threadIdx.x >> 6] = m;
  // this workgroup's bits of the other buffer (read before this barrier) are cleared
  unsigned* to = E.touch + (size_t)(par ^ 1) * MAXSG * (GE_GMAX / 32);
  for (int sph = threadIdx.x; sph < 2 * nd; sph += GE_T)
    if ((g_ge.touched[par ^ 1][sph >> 5] >> (sph & 31)) & 1u)
      atomicAnd(&to[sph * (GE_GMAX / 32) + (blockIdx.x >> 5)], ~(1u << (blockIdx.x & 31)));
  __syncthreads();
  if (threadIdx.x < (MAXSG + 31) / 32) g_ge.touched[par ^ 1][threadIdx.x] = 0u;
  double r = 0.0;
  for (int w = 0; w < GE_W; w++) r = fmax(r, g_ge.red[w]);
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(GE_T) void co_ge_engine(CoP P, CoD D, GeD E, int nsteps, int mode) {
  const int tid = threadIdx.x, nd = P.nd, G = (int)gridDim.x;
  CoScalars* sc = D.sc;
  int* err = &sc->err;
  for (int j = tid; j < 12 * 2 * nd; j += GE_T) {
    const int f = j / (2 * nd), s = j % (2 * nd);
    g_ge.dim[f][s] = D.dim[f * MAXSG + s];
  }
  for (int j = tid; j < GE_W * 4 * MAXSG; j += GE_T) g_ge.wacc[j / (4 * MAXSG)][j % (4 * MAXSG)] = 0.0;
  if (tid < 8) g_ge.tm[tid] = 0ull;
  if (tid < 2 * ((MAXSG + 31) / 32)) g_ge.touched[tid / ((MAXSG + 31) / 32)][tid % ((MAXSG + 31) / 32)] = 0u;
  int par = 0;
  const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
  __syncthreads();
  ge_build_grids(P);
  const int nn = min(sc->near_num, D.near_cap);
  const int chunk = (nn + G - 1) / G;
  if (tid == 0) {
    g_ge.ep = 0u;   // ge_reset cleared the barrier words before this launch
    g_ge.cstep = sc->cstep; g_ge.tc = sc->time_count; g_ge.tp = sc->time_penetrate; g_ge.in_num = sc->in_num;
    g_ge.nev = (mode & M_RESUME) ? sc->n_ev : 0;
    g_ge.vmax2 = sc->near_vmax2;
    g_ge.rebuild = 0;
    g_ge.qb = min(nn, (int)blockIdx.x * chunk);
    g_ge.qe = min(nn, g_ge.qb + chunk);
    g_ge.ib = 0; g_ge.ie = 0;
  }
  __syncthreads();
  // near particles of this chunk into their slots
  for (int q = g_ge.qb + tid; q < g_ge.qe; q += GE_T) {
    const int p = D.near_list[q];
    D.ns_x[q] = D.x[p]; D.ns_y[q] = D.y[p]; D.ns_z[q] = D.z[p];
    D.ns_vx[q] = D.vx[p]; D.ns_vy[q] = D.vy[p]; D.ns_vz[q] = D.vz[p];
    D.ns_flag[q] = D.flag[p]; D.ns_stat[q] = D.status[p];
  }
  __threadfence_block();
  __syncthreads();
  const auto scatter_chunk = [&]() {
    for (int q = g_ge.qb + tid; q < g_ge.qe; q += GE_T) {
      const int p = D.near_list[q];
      D.x[p] = D.ns_x[q]; D.y[p] = D.ns_y[q]; D.z[p] = D.ns_z[q];
      D.vx[p] = D.ns_vx[q]; D.vy[p] = D.ns_vy[q]; D.vz[p] = D.ns_vz[q];
      D.flag[p] = D.ns_flag[q]; D.status[p] = D.ns_stat[q];
    }
  };
  const auto save_scalars = [&](bool with_x0) {
    if (blockIdx.x == 0 && tid == 0) {
      sc->cstep = g_ge.cstep; sc->time_count = g_ge.tc; sc->time_penetrate = g_ge.tp; sc->in_num = g_ge.in_num;
      sc->n_ev = g_ge.nev; sc->near_vmax2 = g_ge.vmax2;
    }
    if (blockIdx.x == 0) {
      for (int j = tid; j < 12 * 2 * nd; j += GE_T) {
        const int f = j / (2 * nd), s = j % (2 * nd);
        D.dim[f * MAXSG + s] = g_ge.dim[f][s];
      }
      if (with_x0)
        for (int s = tid; s < 2 * nd; s += GE_T) { sc->x0[s] = g_ge.dim[DX][s]; sc->y0[s] = g_ge.dim[DY][s]; sc->z0[s] = g_ge.dim[DZ][s]; }
    }
  };

  if (mode & M_SETUP) {
    ge_rebuild(P, D, E, false, false, true);
    for (int k = g_ge.ib + tid; k < g_ge.ie; k += GE_T) {
      D.in_plist[k] = D.near_list[D.in_list[k]];
      for (int j = 0; j < 12; j++) E.in_fp4[(size_t)k * 12 + j] = 0.0;
      for (int j = 0; j < 4; j++) E.in_pot4[(size_t)k * 4 + j] = 0.0;
      D.in_msd[k] = 0.0;
    }
    __threadfence_block();
    __syncthreads();
    scatter_chunk();
    if (tid == 0 && blockIdx.x == 0) sc->n_ev = 0;
    save_scalars(true);
    if (blockIdx.x == 0 && tid == 0) sc->n_ev = 0;
    return;
  }

  const bool resume = (mode & M_RESUME) != 0;
  if (!resume) {
    ge_map_inside(D, E);
    double mx = 0.0;
    for (int k = g_ge.ib + tid; k < g_ge.ie; k += GE_T) {
      double x, y, z, vx, vy, vz, ms;
      int fl;
      ge_slot_load(D, k, x, y, z, vx, vy, vz, fl, ms);
      const unsigned ids = ge_memb_of(E, k);
      double fx[4], fy[4], fz[4];
      for (int m = 0; m < 4; m++) {
        fx[m] = E.in_fp4[(size_t)k * 12 + m * 3 + 0];
        fy[m] = E.in_fp4[(size_t)k * 12 + m * 3 + 1];
        fz[m] = E.in_fp4[(size_t)k * 12 + m * 3 + 2];
      }
      ge_kick(P, fl, ids, fx, fy, fz, vx, vy, vz);
      stream_particle(x, y, z, vx, vy, vz, fl, 1.0, P, D.react_box, err);
      double v2 = (vx * vx + vy * vy) + vz * vz;
      ms = ms + v2 * (P.MDt * P.MDt);
      mx = fmax(mx, ms);
      ge_slot_store(D, k, x, y, z, vx, vy, vz, fl, ms);
    }
    ge_post(P, E, mx, par);
    ge_sync(E, err, true);
    const double m = ge_totals(P, E, par);
    par ^= 1;
    ge_dimers(P, D, E, false, true, m);
  } else {
    if (tid == 0) g_ge.rebuild = 1;
    __syncthreads();
  }

  for (int t = 1; t <= nsteps; t++) {
    const bool last = (t == nsteps);
    if (g_ge.rebuild) {
      if (tid == 0) g_ge.tm[4] += 1;
      GE_T0();
      if (!(resume && last)) {
        const bool stop = (mode & M_STOP) && last;
        ge_rebuild(P, D, E, true, true, !stop);
        if (stop) {
          __threadfence_block();
          __syncthreads();
          scatter_chunk();
          save_scalars(false);
          return;
        }
      } else {
        ge_rebuild(P, D, E, false, false, true);
      }
      GE_T1(1);
    }
    {
      GE_T0();
      const double mx = ge_particles(P, D, E, last);
      __syncthreads();
      ge_post(P, E, mx, par);
      GE_T1(2);
    }
    ge_sync(E, err, true);
    double m;
    {
      GE_T0();
      m = ge_totals(P, E, par);
      par ^= 1;
      GE_T1(6);
    }
    {
      GE_T0();
      ge_dimers(P, D, E, true, !last, m);
      GE_T1(3);
    }
  }

  // end of the interval: census exclusion, inside list as particle indices,
  // entries back into their slots and the particle arrays
  for (int k = g_ge.ib + tid; k < g_ge.ie; k += GE_T) {
    double x, y, z, vx, vy, vz, ms;
    int fl;
    ge_slot_load(D, k, x, y, z, vx, vy, vz, fl, ms);
    const unsigned ids = ge_memb_of(E, k);
    bool ex = false;
    for (int m = 0; m < 4; m++) {
      const unsigned b = (ids >> (8 * m)) & 0xffu;
      if (b == 0xffu) break;
      const int kk = (int)(b >> 1), ii = (int)(b & 1u), s = kk + ii * nd;
      double d1 = pbc_half(x - g_ge.dim[DX][s], P.box[0]);
      double d2 = pbc_half(y - g_ge.dim[DY][s], P.box[1]);
      double d3 = z - g_ge.dim[DZ][s];
      if (d1 * d1 + d2 * d2 + d3 * d3 <= P.cut2[ii]) ex = true;
    }
    const int q = D.in_list[k];
    D.ns_stat[q] = ST_INSIDE | (ex ? ST_EXCL : 0);
    D.in_plist[k] = D.near_list[q];
  }
  ge_unstage(D);
  __threadfence_block();
  __syncthreads();
  scatter_chunk();
  save_scalars(true);
  if (blockIdx.x == 0 && tid == 0) {
    sc->prof[0] += __builtin_amdgcn_s_memrealtime() - t_start;
    sc->prof[1] += g_ge.tm[1]; sc->prof[2] += g_ge.tm[2]; sc->prof[3] += g_ge.tm[3]; sc->prof[4] += g_ge.tm[4];
    sc->prof[5] += g_ge.tm[5]; sc->prof[6] += g_ge.tm[6]; sc->prof[7] += 1;
  }
}

// after the clusters' launch: their sphere fields and scalars back into the
// system's (all clusters agree on the step counters and the rebuild events)
__global__ __launch_bounds__(256) void clu_finish(CoP P, CoD G, CluBase B, int mode) {
  CluInfo* ci = G.clu;
  const int nclu = ci->nclu, nd = P.nd;
  for (int w = threadIdx.x; w < nclu * 12 * MAXS; w += 256) {
    const int c = w / (12 * MAXS), f = (w / MAXS) % 12, sl = w % MAXS;
    const int nl = min(ci->cnt[c], MAXD);
    if (sl >= 2 * nl) continue;
    const int j = sl % nl, i = sl / nl;
    G.dim[f * MAXSG + ci->dims[c][j] + i * nd] = B.dimc[(size_t)c * 12 * MAXSG + f * MAXSG + sl];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    CoScalars* gs = G.sc;
    const CoScalars* s0 = B.scs;
    int err = 0, in_num = 0, near_num = 0;
    double tp = 0.0, nv = 0.0;
    unsigned long long pf[5] = {0, 0, 0, 0, 0};
    for (int c = 0; c < nclu; c++) {
      const CoScalars* cs = B.scs + c;
      if (!err && cs->err) err = cs->err;
      if (cs->cstep != s0->cstep || cs->time_count != s0->time_count || cs->time_penetrate != s0->time_penetrate ||
          cs->n_ev != s0->n_ev)
        if (!err) err = ERR_BARRIER;
      in_num += cs->in_num;
      near_num += cs->near_num;
      tp = tp + cs->tpotential;
      nv = fmax(nv, cs->near_vmax2);
      for (int e = 0; e < 5; e++) pf[e] = max(pf[e], cs->prof[e]);
    }
    if (!err && ci->abort) err = ERR_BARRIER;
    if (err) raise_error(&gs->err, err);
    gs->cstep = s0->cstep; gs->time_count = s0->time_count; gs->time_penetrate = s0->time_penetrate;
    gs->n_ev = s0->n_ev;
    for (int e = 0; e < MAXEV; e++) gs->events[e] = s0->events[e];
    gs->in_num = in_num;
    gs->near_vmax2 = nv;
    gs->tpotential = tp;
    for (int e = 0; e < 5; e++) gs->prof[e] = pf[e];
    gs->prof[5] = near_num; gs->prof[6] = in_num; gs->prof[7] = s0->prof[7];
  }
  if (!(mode & M_STOP))
    for (int s = threadIdx.x; s < 2 * nd; s += 256) sphere_xyz(G, s, G.sc->x0[s], G.sc->y0[s], G.sc->z0[s]);
}

// ------------------------------------------------------------- bulk kernels
// A block of a batched launch works on elements [begin, end) of replica r.
struct Tile {
  int r, pad;
  long long begin, end;
};

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b to XCD b % 8).
// The logical block of workgroup b makes each XCD's workgroups cover one
// contiguous stretch of the arrays, so neighbouring data shares an L2.
__device__ __forceinline__ int xcd_block(int b, int G) {
  const int q = G >> 3, r = G & 7, x = b & 7, k = b >> 3;
  return x * q + min(x, r) + k;
}
// contiguous range [first, end) of logical block lb when n items are split over
// G blocks in multiples of 256
__device__ __forceinline__ void block_range(long long n, int G, int lb, long long& first, long long& end) {
  const long long chunk = (((n + G - 1) / G) + 255) / 256 * 256;
  first = (long long)lb * chunk;
  end = min(first + chunk, n);
}

__device__ __forceinline__ void collision_shift(const CoP& P, const CoD& D, double& xs, double& ys, double& zs) {
  double u0, u1, u2;
  if (P.rng_mode == 0) {
    u0 = D.U[0]; u1 = D.U[1]; u2 = D.U[2];
  } else {
    double t;
    philox_uniform2(P.seed, 0, D.sc->coll, 0, u0, u1);
    philox_uniform2(P.seed, 0, D.sc->coll, 1, u2, t);
  }
  xs = P.a0 * (u0 - 0.5);
  ys = P.a0 * (u1 - 0.5);
  zs = P.a0 * (u2 - 0.5);
}


// mpc_rotation_mod.f95:116-145: cell of the shifted grid, 0-based
__device__ __forceinline__ long long shifted_cell(const CoP& P, double x, double y, double z, double xs, double ys,
                                                  double zs) {
  double X = x + xs;
  X = wrap_periodic(X, P.box[0]);
  double Y = y + ys;
  Y = wrap_periodic(Y, P.box[1]);
  double Z = z + zs;
  double fz = zs < 0.0 ? floor(Z + 1.0) : floor(Z);
  double c = ffloor_i(X) + P.box[0] * ffloor_i(Y);
  c = c + (P.box[0] * P.box[1]) * (double)(int)fz;
  c = c + 1.0;
  return (long long)c - 1;
}

// how far beyond the larger list radius a near particle may end the interval:
// its reach (|v| tau + guard, within 1.5 for |v| below 2.4) plus its own
// travel, with margin, plus half a cell diagonal for the test on cell centres
constexpr double ZONE_PAD = 5.9;
// a deferred particle's cell is taken from its single flight; faces within
// this distance of it add the neighbouring cells (the replay differs by rounding)
constexpr double LATE_MARGIN = 1e-6;

// lists tour position t as late, once per interval
__device__ __forceinline__ void late_cell(const CoD& D, int t) {
  if (atomicMax(&D.stamp[t], D.epoch) < D.epoch) {
    const int k = atomicAdd(&D.sc->n_late, 1);
    if (k < D.late_cap) D.late[k] = t;
    else raise_error(&D.sc->err, ERR_ZONE_CAP);
  }
}
// A deferred (wall-reaching) far particle: its single flight with the bounce
// locates its end cell, listed late with the neighbours across faces within
// LATE_MARGIN; returns that cell's tour position (or -1 out of range)
__device__ __forceinline__ int late_cells_of(const CoP& P, const CoD& D, double x, double y, double z, double vx,
                                             double vy, double vz, double xs, double ys, double zs) {
  int f = 0;
  stream_flight(x, y, z, vx * P.MDt, vy * P.MDt, vz * P.MDt, (double)P.MPCtstep, P);
  (void)stream_walls(x, y, z, vx, vy, vz, f, P, D.react_box, &D.sc->err);
  const double X = wrap_periodic_any(x + xs, P.box[0]), Y = wrap_periodic_any(y + ys, P.box[1]);
  const double Zc = z + zs + (zs < 0.0 ? 1.0 : 0.0);
  const double fx = floor(X), fy = floor(Y), fz = floor(Zc);
  int lo[3], hi[3];
  const double fr[3] = {X - fx, Y - fy, Zc - fz};
  const int base[3] = {(int)fx, (int)fy, (int)fz};
  for (int d = 0; d < 3; d++) {
    lo[d] = base[d] - (fr[d] < LATE_MARGIN ? 1 : 0);
    hi[d] = base[d] + (fr[d] > 1.0 - LATE_MARGIN ? 1 : 0);
  }
  int first = -1;
  for (int kz = lo[2]; kz <= hi[2]; kz++)
    for (int ky = lo[1]; ky <= hi[1]; ky++)
      for (int kx = lo[0]; kx <= hi[0]; kx++) {
        if (kz < 0 || kz > P.Lz) continue;
        const int ix = ((kx % P.Lx) + P.Lx) % P.Lx, iy = ((ky % P.Ly) + P.Ly) % P.Ly;
        const int t = D.mpc_pos[(long long)ix + (long long)P.Lx * iy + (long long)P.Lx * P.Ly * kz];
        late_cell(D, t);
        if (kx == base[0] && ky == base[1] && kz == base[2]) first = t;
      }
  return first;
}

// Grouping of a block's binning (every thread of the block calls it
// together): each lane's rank among the block's lanes with the same cell t,
// and the census -- count and minimum key per (species, cell) ck.  The keys go
// into hash tables in LDS with one LDS atomic per lane, then one global atomic
// per distinct key of the block: no loop over a wave's distinct keys.
constexpr int GT = 512;   // hash slots per table, twice the block
struct GroupTables {
  int rk[GT], rc[GT], rb[GT];   // cell key, lanes, global base
  int ck[GT], cc[GT];           // census key, lanes
  unsigned cm[GT];              // census minimum key
};
__device__ __forceinline__ int gt_slot(int* keys, int key) {
  int j = (int)(((unsigned)key * 2654435761u) >> 23);   // 9 bits
  for (;;) {   // fewer keys than slots: an empty or equal slot is always found
    const int old = atomicCAS(&keys[j], -1, key);
    if (old == -1 || old == key) return j;
    j = (j + 1) & (GT - 1);
  }
}
__device__ __forceinline__ int block_groups(const CoP& P, const CoD& D, GroupTables& G, int t, bool ok, int ck,
                                            bool cen, unsigned key) {
  for (int j = threadIdx.x; j < GT; j += blockDim.x) {
    G.rk[j] = -1; G.rc[j] = 0; G.ck[j] = -1; G.cc[j] = 0; G.cm[j] = 0xffffffffu;
  }
  __syncthreads();
  int rs = 0, local = 0;
  if (ok) { rs = gt_slot(G.rk, t); local = atomicAdd(&G.rc[rs], 1); }
  if (cen) {
    const int cs = gt_slot(G.ck, ck);
    atomicAdd(&G.cc[cs], 1);
    atomicMin(&G.cm[cs], key);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < GT; j += blockDim.x) {
    if (G.rc[j]) G.rb[j] = atomicAdd(&D.cnt[G.rk[j]], G.rc[j]);
    if (G.cc[j]) {
      const int k = G.ck[j];
      atomicAdd(&D.tcount[k], G.cc[j]);   // no return: the collision kernel checks the totals against CENSUS_MAX
      if (k >= P.vol && k < 2 * P.vol) atomicMin(&D.firstB[k - P.vol], G.cm[j]);
    }
  }
  __syncthreads();
  return ok ? G.rb[rs] + local : 0;
}

// Binning of one particle per lane; every thread of the block calls it
// together (`valid` false for idle lanes).  The shifted cell, counted by tour
// position with the particle's rank in it; the distance to the nearest sphere
// of `sph` beyond its list radius (selection of the next interval's near set);
// the census of distribute_mod.f95:95-183: outside particles keyed by index,
// inside ones after all outside ones, inside ones within a cutoff skipped.
// slab 0: cell and rank for co_scatter; 1: the particle's slot in its cell's
// slab written directly; 2: the same, and the cell must be listed late (the
// particle was binned after the first collision pass).
__device__ __forceinline__ void bin_particle(const CoP& P, const CoD& D, int i, bool valid, bool near,
                                             unsigned char st, double x, double y, double z, int f, double xs,
                                             double ys, double zs, const double (*sph)[MAXSG], int slab = 0) {
  __shared__ GroupTables G;
  const int NT = P.NT;
  bool ok = valid;
  int t = 0;
  if (valid) {
    long long c = shifted_cell(P, x, y, z, xs, ys, zs);
    if (c < 0 || c >= P.mpcnumcell) { raise_error(&D.sc->err, ERR_CELL_RANGE); ok = false; }
    else t = D.mpc_pos[c];
  }
  unsigned key = (unsigned)i;
  bool cen = false;
  int ck = 0;
  if (valid && ok) {
    bool skip = false;
    if (near && (st & ST_INSIDE)) { key = (unsigned)NT + (unsigned)i; skip = (st & ST_EXCL) != 0; }
    int sp = f == 0 ? 0 : f == 1 ? 1 : f == 2 ? 2 : -1;
    if (!skip && sp >= 0) {
      long long u = cell_unshifted0(x, y, z, P.box[0], P.box[1]);
      if (u < 0 || u >= P.vol) raise_error(&D.sc->err, ERR_CELL_RANGE);
      else { cen = true; ck = (int)(sp * P.vol + u); }
    }
  }
  const int r = block_groups(P, D, G, t, ok, ck, cen, key);
  if (valid) {
    if (ok && r >= BMAX) raise_error(&D.sc->err, ERR_BUCKET);
    D.cell_of[i] = t;
    if (!slab) D.rank_of[i] = ok ? (unsigned short)min(r, 65535) : (unsigned short)0xffff;
    else if (ok && r < BMAX) {
      D.slab[(long long)t * BMAX + r] = i;
      if (slab == 2 && D.stamp[t] != D.epoch) raise_error(&D.sc->err, ERR_ZONE);
    }
    // distance beyond the list radius; 8 or more beyond every radius gives the
    // cap (255 at 7.97) without a sqrt
    double dmin = 1e30;
    bool close = false;
    for (int s = 0; s < 2 * P.nd; s++) {
      double a = pbc_half(x - sph[0][s], P.box[0]);
      double b = pbc_half(y - sph[1][s], P.box[1]);
      double c = z - sph[2][s];
      const double d2 = a * a + b * b + c * c, far = P.full[s >= P.nd] + 8.0;
      close |= d2 < far * far;
    }
    if (__ballot(close))   // most waves lie far from every sphere: no square roots at all
      for (int s = 0; s < 2 * P.nd; s++) {
        double a = pbc_half(x - sph[0][s], P.box[0]);
        double b = pbc_half(y - sph[1][s], P.box[1]);
        double c = z - sph[2][s];
        const double d2 = a * a + b * b + c * c, far = P.full[s >= P.nd] + 8.0;
        if (d2 < far * far) dmin = fmin(dmin, sqrt(d2) - P.full[s >= P.nd]);
      }
    double qd = floor(fmax(dmin, 0.0) * 32.0 * (1.0 - 1e-12));
    if (st & ST_INSIDE) qd = 0.0;   // list members stay near until the next rebuild
    D.dq[i] = (unsigned char)fmin(qd, 255.0);
  }
}

// The lazy-streaming events of the interval (evolve_mod.f95:474-527) for one
// far particle, with the reference's rounding: one free flight per rebuild
// event, the walls checked after each; position, velocity (at a bounce) and
// species (at a reactive wall) written back.
__device__ __forceinline__ void replay_events(const CoP& P, const CoD& D, int i, double& x, double& y, double& z,
                                              int& f, const int* ev, int nev) {
  double vx = D.vx[i], vy = D.vy[i], vz = D.vz[i];
  // step displacements v*MDt, negated with v at a bounce ((-v)*MDt is -(v*MDt) exactly)
  double dx = vx * P.MDt, dy = vy * P.MDt, dz = vz * P.MDt;
  const int f0 = f;
  bool bounced = false;
  for (int e = 0; e < nev; e++) {
    stream_flight(x, y, z, dx, dy, dz, (double)ev[e], P);
    if (stream_walls(x, y, z, vx, vy, vz, f, P, D.react_box, &D.sc->err)) {
      bounced = true;
      dx = -dx; dy = -dy; dz = -dz;
    }
  }
  D.x[i] = x; D.y[i] = y; D.z[i] = z;
  if (bounced) { D.vx[i] = vx; D.vy[i] = vy; D.vz[i] = vz; }
  if (f != f0) D.flag[i] = f;
}

// A far particle whose single flight ends this close to a wall, or beyond it,
// is left to the event replay: far beyond the flights' rounding difference
constexpr double WALL_MARGIN = 1e-9;

// appends the lanes' particles with `wall` set to the deferred list, one
// atomic per wave (whole waves call it together)
__device__ __forceinline__ void wall_defer(const CoD& D, bool wall, int i) {
  const unsigned long long m = __ballot(wall);
  if (!m) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll(m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(&D.sc->n_wall, __popcll(m));
  base = wave_bcast(base, leader);
  if (wall) {
    const int k = base + __popcll(m & ((1ull << lane) - 1ull));
    if (k < D.wall_cap) D.wall_list[k] = i;
    else raise_error(&D.sc->err, ERR_WALL_CAP);
  }
}

// Streaming of the far particles, then binning.
// FAR = false: the interval's rebuild events replayed (replay_events), every
// particle binned, distances to the spheres at the interval's end (x0..z0).
// FAR = true (co_far, concurrent with the near engine): far particles only,
// one free flight of the whole interval -- the replay's sum of flights up to
// a few ulps, since a far particle meets no list -- and distances to the
// spheres at the interval's start (xs0..zs0).  A flight that reaches a wall
// is not taken: the reference's bounce time (z - Lz)/vz magnifies the ulps by
// |v|/|vz|, so such particles go to the deferred list and replay the events
// in co_bin_rest, which also bins the near set once the near engine is done.
// slab (FAR only): bin into the cells' slabs (first collision pass beside the
// near engine); deferred particles then list their possible end cells late.
template <bool FAR>
__device__ __forceinline__ void bulk_body(const CoP& P, const CoD& D, int do_stream, int do_bin, long long first,
                                          long long end, long long step, int slab = 0) {
  const int nev = FAR ? 0 : min(D.sc->n_ev, MAXEV);
  __shared__ int ev[MAXEV];
  __shared__ double sph[3][MAXSG];
  if (!FAR)
    for (int j = threadIdx.x; j < nev; j += blockDim.x) ev[j] = D.sc->events[j];
  const double* sx = FAR ? D.sc->xs0 : D.sc->x0;
  const double* sy = FAR ? D.sc->ys0 : D.sc->y0;
  const double* sz = FAR ? D.sc->zs0 : D.sc->z0;
  for (int j = threadIdx.x; j < 2 * P.nd; j += blockDim.x) {
    sph[0][j] = sx[j];
    sph[1][j] = sy[j];
    sph[2][j] = sz[j];
  }
  __syncthreads();
  double xs = 0, ys = 0, zs = 0;
  if (do_bin) collision_shift(P, D, xs, ys, zs);
  const double flight = (double)P.MPCtstep;
  const double ztop = P.box[2] - WALL_MARGIN;
  for (long long base = first; base < end; base += step) {
    const int i = (int)(base + threadIdx.x);
    const bool valid = i < end;
    bool near = false, wall = false;
    double x = 0, y = 0, z = 0;
    int f = -1;
    unsigned char st = 0;
    if (valid) near = (D.near_bits[i >> 5] >> (i & 31)) & 1u;
    const bool mine = valid && !(FAR && near);
    if (mine) {
      st = D.status[i];
      if (!near && st != 0) { D.status[i] = 0; st = 0; }   // left the near set earlier
      if (FAR) {   // streamed past the caches: the near engine's slots stay in L2 and the
                   // binning's outputs in the infinity cache for co_scatter (1.48 -> 1.45 ms per interval)
        x = __builtin_nontemporal_load(&D.x[i]);
        y = __builtin_nontemporal_load(&D.y[i]);
        z = __builtin_nontemporal_load(&D.z[i]);
        f = D.flag[i];
        double vx = __builtin_nontemporal_load(&D.vx[i]), vy = __builtin_nontemporal_load(&D.vy[i]);
        double vz = __builtin_nontemporal_load(&D.vz[i]);
        const double dx = vx * P.MDt, dy = vy * P.MDt, dz = vz * P.MDt;
        const double zt = z + dz * flight;   // the flight's z, as stream_flight computes it
        // The bounce time z/vz (or (z - Lz)/vz) turns the flights' rounding
        // difference dz in z into 2 |v_xy| dz / |vz| in x and y.  At z = 0 the
        // ulps are below 1e-16, so a bounce is taken here unless it is
        // shallower than |vz| = |v_xy| / 100 (then below 1e-13); at the top
        // wall the ulps of Lz are not, so every crossing there is deferred.
        // A flight ending within WALL_MARGIN of z = 0 is deferred as well: the
        // replay might end on the other side.
        const bool top = !(zt < ztop), clear = zt > WALL_MARGIN;
        const bool bottom = zt <= -WALL_MARGIN && fabs(vz) >= 0.01 * fmax(fabs(vx), fabs(vy));
        wall = top || !(clear || bottom);
        if (!wall) {
          stream_flight(x, y, z, dx, dy, dz, flight, P);
          if (bottom) {
            const int f0 = f;
            (void)stream_walls(x, y, z, vx, vy, vz, f, P, D.react_box, &D.sc->err);
            D.vx[i] = vx; D.vy[i] = vy; D.vz[i] = vz;
            if (f != f0) D.flag[i] = f;
          }
          __builtin_nontemporal_store(x, &D.x[i]);
          __builtin_nontemporal_store(y, &D.y[i]);
          __builtin_nontemporal_store(z, &D.z[i]);
        } else if (slab && do_bin) {   // the first pass must leave its possible end cells alone
          const int t = late_cells_of(P, D, x, y, z, vx, vy, vz, xs, ys, zs);
          D.cell_of[i] = t >= 0 ? t : 0;   // co_rotate_far skips it there (a late cell)
          if (t < 0) raise_error(&D.sc->err, ERR_CELL_RANGE);
        }
      } else {
        x = D.x[i]; y = D.y[i]; z = D.z[i];
        f = D.flag[i];
      }
      if (!FAR && do_stream && !near && nev > 0) {
        replay_events(P, D, i, x, y, z, f, ev, nev);
      }
    }
    if (FAR) wall_defer(D, wall, i);
    if (!do_bin) continue;
    bin_particle(P, D, i, mine && !wall, near, st, x, y, z, f, xs, ys, zs, sph, slab);
  }
}
__global__ __launch_bounds__(256) void co_bulk(CoP P, CoD D, int do_stream, int do_bin) {
  bulk_body<false>(P, D, do_stream, do_bin, (long long)blockIdx.x * blockDim.x, P.NT,
                   (long long)gridDim.x * blockDim.x);
}
// six waves per SIMD (at most 80 VGPRs, four of them spilled) where the
// register count alone gave five: 0.376 -> 0.358 ms at 100^3 (seven waves
// spilled sixteen and gained less)
__global__ __launch_bounds__(256, 6) void co_far(CoP P, CoD D, int do_bin, int slab) {
  bulk_body<true>(P, D, 1, do_bin, (long long)blockIdx.x * blockDim.x, P.NT, (long long)gridDim.x * blockDim.x,
                  slab);
}
// co_far's single flight covers the whole interval: the rebuild events
// recorded by the near engine must add up to it
__device__ __forceinline__ void check_flight(const CoP& P, const CoD& D) {
  int sum = 0;
  const int nev = D.sc->n_ev;
  for (int e = 0; e < nev && e < MAXEV; e++) sum += D.sc->events[e];
  if (nev > MAXEV || sum != P.MPCtstep) raise_error(&D.sc->err, ERR_FLIGHT);
}
// After the near engine and co_far: the deferred far particles replay the
// interval's events; with do_bin they and the near set are binned (distances
// to the spheres at the interval's start, as co_far).  Then the flight check.
// slab: bin into the cells' slabs, each cell listed late (second pass)
__global__ __launch_bounds__(256) void co_bin_rest(CoP P, CoD D, int do_bin, int slab) {
  __shared__ int ev[MAXEV];
  __shared__ double sph[3][MAXSG];
  const int nev = min(D.sc->n_ev, MAXEV);
  for (int j = threadIdx.x; j < nev; j += blockDim.x) ev[j] = D.sc->events[j];
  for (int j = threadIdx.x; j < 2 * P.nd; j += blockDim.x) {
    sph[0][j] = D.sc->xs0[j];
    sph[1][j] = D.sc->ys0[j];
    sph[2][j] = D.sc->zs0[j];
  }
  __syncthreads();
  double xs = 0, ys = 0, zs = 0;
  if (do_bin) collision_shift(P, D, xs, ys, zs);
  // an overflowing near list or deferred list was reported when it was built
  const long long nn = do_bin ? min(D.sc->near_num, D.near_cap) : 0;
  const long long total = nn + min(D.sc->n_wall, D.wall_cap);
  for (long long base = (long long)blockIdx.x * blockDim.x; base < total; base += (long long)gridDim.x * blockDim.x) {
    const long long q = base + threadIdx.x;
    const bool valid = q < total, near = q < nn;
    int i = 0, f = -1;
    double x = 0, y = 0, z = 0;
    unsigned char st = 0;
    if (valid) {
      i = near ? D.near_list[q] : D.wall_list[q - nn];
      x = D.x[i]; y = D.y[i]; z = D.z[i];
      f = D.flag[i];
      st = D.status[i];
      if (!near) replay_events(P, D, i, x, y, z, f, ev, nev);
    }
    if (do_bin) bin_particle(P, D, i, valid, near, st, x, y, z, f, xs, ys, zs, sph, slab ? 2 : 0);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) check_flight(P, D);
}
__global__ __launch_bounds__(256) void co_bulk_b(const CoP* Ps, const CoD* Ds, const Tile* tiles, int do_stream,
                                                 int do_bin) {
  const Tile t = tiles[blockIdx.x];
  const CoP P = Ps[t.r];
  const CoD D = Ds[t.r];
  bulk_body<false>(P, D, do_stream, do_bin, t.begin, t.end, 256);
}

__device__ __forceinline__ void scatter_body(const CoD& D, long long first, long long end, long long step) {
  for (long long i = first + threadIdx.x; i < end; i += step) {
    int r = D.rank_of[i];
    if (r < BMAX) D.bucket[D.off[D.cell_of[i]] + r] = (int)i;
  }
}
__global__ __launch_bounds__(256) void co_scatter(CoP P, CoD D) {
  scatter_body(D, (long long)blockIdx.x * blockDim.x, P.NT, (long long)gridDim.x * blockDim.x);
}
__global__ __launch_bounds__(256) void co_scatter_b(const CoD* Ds, const Tile* tiles) {
  const Tile t = tiles[blockIdx.x];
  const CoD D = Ds[t.r];
  scatter_body(D, t.begin, t.end, 256);
}

// Rotation of mpc_rotation_mod.f95:205-230: v <- c + R(axis u) (v - c)
__device__ __forceinline__ void rotate_velocity(const CoP& P, double cx, double cy, double cz, double ux, double uy,
                                                double uz, double& vx, double& vy, double& vz) {
  double dx = vx - cx, dy = vy - cy, dz = vz - cz;
  const double cr = P.cosrot, sr = P.sinrot, omc = P.omc;
  double rx = (dx * (cr + (ux * ux) * omc) + dy * ((ux * uy) * omc - uz * sr)) + dz * ((ux * uz) * omc + uy * sr);
  double ry = (dx * ((ux * uy) * omc + uz * sr) + dy * (cr + (uy * uy) * omc)) + dz * ((uy * uz) * omc - ux * sr);
  double rz = (dx * ((ux * uz) * omc - uy * sr) + dy * ((uy * uz) * omc + ux * sr)) + dz * (cr + (uz * uz) * omc);
  vx = cx + rx; vy = cy + ry; vz = cz + rz;
}
// the next interval's near set: near when the distance beyond a list radius
// (recorded by the binning, rounded down) is within |v| tau + guard, widened
// by `shift` (sphere motion since the distance was measured); far ones feed
// the far max |v|^2
__device__ __forceinline__ void classify_next(const CoP& P, unsigned* bits, const unsigned char* dq, int i,
                                              double vx, double vy, double vz, double tau, double shift,
                                              double& fmax2) {
  double v2 = (vx * vx + vy * vy) + vz * vz;
  double reach = sqrt(v2) * tau + P.guard + 1e-6 + shift;
  if ((double)dq[i] * (1.0 / 32.0) <= reach) atomicOr(&bits[i >> 5], 1u << (i & 31));
  else fmax2 = fmax(fmax2, v2);
}
// a block's far max |v|^2 into *dst (every thread calls it)
__device__ __forceinline__ void far_max_publish(unsigned long long* dst, double fmax2) {
  fmax2 = wave_max(fmax2);
  __shared__ double red[16];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = fmax2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
    atomicMax(dst, (unsigned long long)__double_as_longlong(m));
  }
}

// the centre velocity from the ordered sums, the random axis and the record
// (written when `write`) of the cell at tour position t
__device__ __forceinline__ void cell_finish(const CoP& P, const CoD& D, long long t, int c, int n, double sx,
                                            double sy, double sz, double* r6, bool write) {
  if (n != 0) {
    double tn = (double)n;
    sx = sx / tn; sy = sy / tn; sz = sz / tn;
  }
  double u1, u2;
  if (P.rng_mode == 0) {
    u1 = D.U[3 + 2 * (long long)c];
    u2 = D.U[4 + 2 * (long long)c];
  } else {
    philox_uniform2(P.seed, 1, D.sc->coll, (unsigned long long)c, u1, u2);
  }
  double phi = P.twopi * u1;
  double th = 2.0 * u2 - 1.0;
  double st = sqrt(1.0 - th * th);
  double sp, cp;
  sincos(phi, &sp, &cp);   // one reduction; the same values as sin() and cos()
  if (write) {
    double2* rec = (double2*)(D.cellrec + t * 8);
    rec[0] = make_double2(sx, sy);
    rec[1] = make_double2(sz, st * cp);
    rec[2] = make_double2(st * sp, th);
    rec[3] = make_double2((double)n, 0.0);
  }
  r6[0] = sx; r6[1] = sy; r6[2] = sz; r6[3] = st * cp; r6[4] = st * sp; r6[5] = th;
}
// The cell's centre-of-mass velocity from its n members idx[0], idx[istride],
// ... (ascending particle index) summed in that order (mpc_rotation_mod.f95:
// 153-182), its random axis (:184-193), and its record at tour position t;
// r6 receives the velocity and the axis.
__device__ __forceinline__ void cell_record(const CoP& P, const CoD& D, long long t, int c, const int* idx,
                                            int istride, int n, double* r6) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  int j = 0;
  for (; j + 4 <= n; j += 4) {   // loads of four members in flight, sums in order
    int i0 = idx[j * istride], i1 = idx[(j + 1) * istride], i2 = idx[(j + 2) * istride], i3 = idx[(j + 3) * istride];
    double ax = D.vx[i0], bx = D.vx[i1], cx = D.vx[i2], dx = D.vx[i3];
    double ay = D.vy[i0], by = D.vy[i1], cy = D.vy[i2], dy = D.vy[i3];
    double az = D.vz[i0], bz = D.vz[i1], cz = D.vz[i2], dz = D.vz[i3];
    sx = sx + ax; sx = sx + bx; sx = sx + cx; sx = sx + dx;
    sy = sy + ay; sy = sy + by; sy = sy + cy; sy = sy + dy;
    sz = sz + az; sz = sz + bz; sz = sz + cz; sz = sz + dz;
  }
  for (; j < n; j++) {
    int i = idx[j * istride];
    sx = sx + D.vx[i];
    sy = sy + D.vy[i];
    sz = sz + D.vz[i];
  }
  cell_finish(P, D, t, c, n, sx, sy, sz, r6, true);
}
// the census limit and the bulk B->A reaction of unshifted cell t
// (reaction_mod.f95:41-62)
__device__ __forceinline__ void cell_react(const CoP& P, const CoD& D, long long t) {
  if (t < P.vol)   // the census totals of the binning kernel against the reference's MaxA/MaxB/MaxS
    if (D.tcount[t] > CENSUS_MAX || D.tcount[P.vol + t] > CENSUS_MAX || D.tcount[2 * P.vol + t] > CENSUS_MAX)
      raise_error(&D.sc->err, ERR_MEMBERS);
  if (t < P.rxnnumcell) {
    int nb = D.tcount[1 * P.vol + t];
    double prob = D.ptab[min(nb, CENSUS_MAX)];
    double r;
    if (P.rng_mode == 0) r = D.U[3 + 2 * P.mpcnumcell + t];
    else { double q; philox_uniform2(P.seed, 2, D.sc->coll, (unsigned long long)t, r, q); }
    if (0.0 < r && r <= prob) {
      unsigned key = D.firstB[t];
      int i = key >= (unsigned)P.NT ? (int)(key - (unsigned)P.NT) : (int)key;
      D.flag[i] = 0;
    }
  }
}

// Per collision cell, in tour order (thread t <-> t-th cell of the tour, so a
// wave touches one compact patch of the Hilbert-sorted particle arrays):
// members in ascending particle order, centre-of-mass velocity
// (mpc_rotation_mod.f95:153-182), random axis (:184-193).  Thread t also does
// the B->A reaction of unshifted cell t (reaction_mod.f95:41-62).
constexpr int LB = 32;   // members sorted in LDS; larger cells sort in place in the bucket
// stage: the workgroup's whole stretch of the bucket copied to LDS (its cells'
// members are contiguous there, from bucket entry sbase), or null
__device__ __forceinline__ void collide_body(const CoP& P, const CoD& D, long long t, int* sidx,
                                             int* stage = nullptr, int sbase = 0) {
  if (t < P.mpcnumcell) {
    const int c = D.mpc_cell[t];
    int n = D.cnt[t];
    if (n > BMAX) n = BMAX;   // ERR_BUCKET already raised
    const int* b = D.bucket + D.off[t];
    int* idx;
    int istride;
    if (stage) {   // insertion sort in place in the staged stretch: ascending particle index
      idx = stage + (D.off[t] - sbase);
      istride = 1;
      for (int j = 1; j < n; j++) {
        int v = idx[j], k = j;
        while (k > 0 && idx[k - 1] > v) { idx[k] = idx[k - 1]; k--; }
        idx[k] = v;
      }
    } else if (n <= LB) {
      idx = sidx + threadIdx.x;
      istride = CT;
      for (int j = 0; j < n; j++) {   // insertion: ascending particle index
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
    double r6[6];
    cell_record(P, D, t, c, idx, istride, n, r6);
  }
  cell_react(P, D, t);
}

__global__ __launch_bounds__(CT) void co_collide(CoP P, CoD D) {
  __shared__ int sidx[LB * CT];
  __shared__ int srange[2];
  // the workgroup's cells t0 .. t0+CT-1 own one contiguous stretch of the
  // bucket: read it once, coalesced, instead of member by member per thread
  const long long t0 = (long long)xcd_block(blockIdx.x, gridDim.x) * CT;
  if (threadIdx.x == 0) {
    const long long te = min(t0 + CT, P.mpcnumcell);
    srange[0] = t0 < te ? D.off[t0] : 0;
    srange[1] = t0 < te ? D.off[te] : 0;
  }
  __syncthreads();
  const int sb = srange[0], total = srange[1] - sb;
  const bool staged = total >= 0 && total <= LB * CT;
  if (staged)
    for (int j = threadIdx.x; j < total; j += CT) sidx[j] = D.bucket[sb + j];
  __syncthreads();
  collide_body(P, D, t0 + threadIdx.x, sidx, staged ? sidx : nullptr, sb);
}
__global__ __launch_bounds__(CT) void co_collide_b(const CoP* Ps, const CoD* Ds, const Tile* tiles) {
  __shared__ int sidx[LB * CT];
  const Tile t = tiles[blockIdx.x];
  const CoP P = Ps[t.r];
  const CoD D = Ds[t.r];
  const long long c = t.begin + threadIdx.x;
  if (c < t.end) collide_body(P, D, c, sidx);
}

// Rotation of mpc_rotation_mod.f95:205-230 and selection of the next
// interval's near set: a particle is near when its distance beyond a list
// radius (recorded by co_bulk, rounded down) is within |v| tau + guard.
// the largest sphere displacement since the interval's start (xs0 -> x0)
__device__ __forceinline__ double start_shift(const CoP& P, const CoD& D) {
  double shift = 0.0;
  for (int s = 0; s < 2 * P.nd; s++) {
    const double a = pbc_half(D.sc->x0[s] - D.sc->xs0[s], P.box[0]);
    const double b = pbc_half(D.sc->y0[s] - D.sc->ys0[s], P.box[1]);
    const double c = D.sc->z0[s] - D.sc->zs0[s];
    shift = fmax(shift, sqrt(a * a + b * b + c * c));
  }
  return shift;
}
// classify 0: rotation only; 1: distances measured to the spheres at the
// interval's end; 2: measured to the spheres at its start (co_far,
// co_bin_rest), so the reach widens by the largest sphere displacement since.
__device__ __forceinline__ void rotate_body(const CoP& P, const CoD& D, int classify, long long first, long long end,
                                            long long step) {
  double fmax2 = 0.0;
  const double tau = P.tau * (1.0 + 1e-9);
  const double shift = classify == 2 ? start_shift(P, D) : 0.0;
  for (long long i = first + threadIdx.x; i < end; i += step) {
    int t = D.cell_of[i];
    const double2* rec = (const double2*)(D.cellrec + (long long)t * 8);
    double2 r0 = rec[0], r1 = rec[1], r2 = rec[2];
    double vx = D.vx[i], vy = D.vy[i], vz = D.vz[i];
    rotate_velocity(P, r0.x, r0.y, r1.x, r1.y, r2.x, r2.y, vx, vy, vz);
    D.vx[i] = vx; D.vy[i] = vy; D.vz[i] = vz;
    if (classify) classify_next(P, D.near_bits, D.dq, (int)i, vx, vy, vz, tau, shift, fmax2);
  }
  if (classify) far_max_publish(&D.sc->far_maxv2, fmax2);
}
__global__ __launch_bounds__(256) void co_rotate(CoP P, CoD D, int classify) {
  rotate_body(P, D, classify, (long long)blockIdx.x * blockDim.x, P.NT, (long long)gridDim.x * blockDim.x);
}
__global__ __launch_bounds__(256) void co_rotate_b(const CoP* Ps, const CoD* Ds, const Tile* tiles) {
  const Tile t = tiles[blockIdx.x];
  const CoP P = Ps[t.r];
  const CoD D = Ds[t.r];
  rotate_body(P, D, 1, t.begin, t.end, 256);
}

// ------------------------------------------------- collisions in two passes
// In an interval without a re-sort the cells are collided in two passes.
// Beside the near engine: every cell no near particle can end in (co_far has
// binned all its members by then), and the rotation of their particles.
// After it: the late cells -- those within reach of the near set (a ball
// around each sphere's start position) and those a wall-reaching far particle
// may end in -- collided and rotated together.  Every cell keeps one slab of
// BMAX slots (no prefix scan), and a cell is late when its stamp holds the
// interval's epoch.  A near or deferred particle outside the late cells is
// reported (ERR_ZONE) rather than collided twice.

// the late cells within ZONE_PAD beyond the larger list radius of sphere
// blockIdx.y (its start position), by their centres in particle coordinates;
// the candidate cube is dealt over the blockIdx.x workgroups
__global__ __launch_bounds__(256) void co_zone_mark(CoP P, CoD D) {
  const int s = blockIdx.y;
  if (s >= 2 * P.nd) return;
  double xs, ys, zs;
  collision_shift(P, D, xs, ys, zs);
  const double R = fmax(P.full[0], P.full[1]) + ZONE_PAD;
  const int r = (int)ceil(R) + 1, w = 2 * r + 1;
  const double sx = D.sc->xs0[s], sy = D.sc->ys0[s], sz = D.sc->zs0[s];
  const double zlow = zs < 0.0 ? 1.0 : 0.0;   // cell iz holds Z + zlow in [iz, iz + 1)
  const int cx0 = (int)floor(sx + xs), cy0 = (int)floor(sy + ys), cz0 = (int)floor(sz + zs + zlow);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < w * w * w; k += gridDim.x * blockDim.x) {
    const int cx = cx0 + k % w - r, cy = cy0 + (k / w) % w - r, cz = cz0 + k / (w * w) - r;
    if (cz < 0 || cz > P.Lz) continue;
    const double a = pbc_half((cx + 0.5 - xs) - sx, P.box[0]);
    const double b = pbc_half((cy + 0.5 - ys) - sy, P.box[1]);
    const double c = (cz + 0.5 - zlow - zs) - sz;
    if (a * a + b * b + c * c > R * R) continue;
    const int ix = ((cx % P.Lx) + P.Lx) % P.Lx, iy = ((cy % P.Ly) + P.Ly) % P.Ly;
    late_cell(D, D.mpc_pos[(long long)ix + (long long)P.Lx * iy + (long long)P.Lx * P.Ly * cz]);
  }
}
// the members of slab cell t sorted by particle index (in LDS, strided per
// thread, or in place in the slab)
__device__ __forceinline__ const int* slab_members(const CoD& D, int t, int* sidx, int& istride, int& n) {
  n = min(D.cnt[t], BMAX);   // ERR_BUCKET already raised beyond
  int* b = D.slab + (long long)t * BMAX;
  int* idx;
  if (n <= LB) {
    idx = sidx + threadIdx.x;
    istride = blockDim.x;
    for (int j = 0; j < n; j++) {
      int v = b[j], k = j;
      while (k > 0 && idx[(k - 1) * istride] > v) { idx[k * istride] = idx[(k - 1) * istride]; k--; }
      idx[k * istride] = v;
    }
  } else {
    idx = b;
    istride = 1;
    for (int j = 1; j < n; j++) {
      int v = idx[j], k = j;
      while (k > 0 && idx[k - 1] > v) { idx[k] = idx[k - 1]; k--; }
      idx[k] = v;
    }
  }
  return idx;
}
// one cell of the slab layout: sorted members, then cell_record
__device__ __forceinline__ const int* slab_cell(const CoP& P, const CoD& D, int t, int* sidx, int& istride, int& n,
                                                double* r6) {
  const int* idx = slab_members(D, t, sidx, istride, n);
  cell_record(P, D, t, D.mpc_cell[t], idx, istride, n, r6);
  return idx;
}
// first pass, beside the near engine: the cells that are not late
__global__ __launch_bounds__(CT) void co_collide_far(CoP P, CoD D) {
  __shared__ int sidx[LB * CT];
  const long long t = (long long)xcd_block(blockIdx.x, gridDim.x) * CT + threadIdx.x;
  if (t < P.mpcnumcell && D.stamp[t] != D.epoch) {
    int istride, n;
    double r6[6];
    (void)slab_cell(P, D, (int)t, sidx, istride, n, r6);
  }
}
// first pass: the particles of those cells rotated (coalesced), classified
// into the next bitmap with the largest displacement the guard allows
__global__ __launch_bounds__(256) void co_rotate_far(CoP P, CoD D) {
  double fmax2 = 0.0;
  const double tau = P.tau * (1.0 + 1e-9);
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < P.NT; i += (long long)gridDim.x * 256) {
    if ((D.near_bits[i >> 5] >> (i & 31)) & 1u) continue;   // the near set is binned after the near engine
    const int t = D.cell_of[i];
    if (D.stamp[t] == D.epoch) continue;   // late cell: the second pass
    const double2* rec = (const double2*)(D.cellrec + (long long)t * 8);
    double2 r0 = rec[0], r1 = rec[1], r2 = rec[2];
    double vx = D.vx[i], vy = D.vy[i], vz = D.vz[i];
    rotate_velocity(P, r0.x, r0.y, r1.x, r1.y, r2.x, r2.y, vx, vy, vz);
    D.vx[i] = vx; D.vy[i] = vy; D.vz[i] = vz;
    classify_next(P, D.next_bits, D.dq, (int)i, vx, vy, vz, tau, P.guard, fmax2);
  }
  far_max_publish(&D.sc->next_fmax2, fmax2);
}
// second pass, after the near engine: the late cells collided and their
// members rotated and classified (the spheres' actual displacement); the
// workgroups from `nlate_wg` on run the census limits and bulk reactions of
// every unshifted cell instead (the census is complete once the near set is
// binned, and the collisions read no species, so the two are independent)
__global__ __launch_bounds__(CT) void co_collide_late(CoP P, CoD D, int nlate_wg) {
  if ((int)blockIdx.x >= nlate_wg) {
    const long long nt = max(P.vol, P.rxnnumcell);
    const long long g = (long long)(gridDim.x - nlate_wg) * CT;
    for (long long t = (long long)(blockIdx.x - nlate_wg) * CT + threadIdx.x; t < nt; t += g) cell_react(P, D, t);
    return;
  }
  const double tau = P.tau * (1.0 + 1e-9), shift = start_shift(P, D);
  const long long nl = min(D.sc->n_late, D.late_cap);
  double fmax2 = 0.0;
  // one wave per late cell: lane j loads the cell's slab slot j and its
  // velocity; the members' velocities go to LDS in ascending particle order,
  // where every lane sums them in that order (the sums of cell_record); each
  // lane rotates and classifies its own member
  __shared__ double sv[CT / 64][3][BMAX];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (long long k = ((long long)blockIdx.x * CT + threadIdx.x) >> 6; k < nl; k += (long long)nlate_wg * (CT / 64)) {
    const int t = D.late[k];
    const int n = min(D.cnt[t], BMAX);   // ERR_BUCKET already raised beyond
    const bool mine = lane < n;
    const int v = mine ? D.slab[(long long)t * BMAX + lane] : 0x7fffffff;
    int rank = 0;
    for (int q = 0; q < n; q++) rank += wave_bcast(v, q) < v;
    double vx = 0.0, vy = 0.0, vz = 0.0;
    if (mine) {
      vx = D.vx[v]; vy = D.vy[v]; vz = D.vz[v];
      sv[w][0][rank] = vx; sv[w][1][rank] = vy; sv[w][2][rank] = vz;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j = 0; j < n; j++) { sx = sx + sv[w][0][j]; sy = sy + sv[w][1][j]; sz = sz + sv[w][2][j]; }
    double r6[6];
    cell_finish(P, D, t, D.mpc_cell[t], n, sx, sy, sz, r6, lane == 0);
    if (mine) {
      rotate_velocity(P, r6[0], r6[1], r6[2], r6[3], r6[4], r6[5], vx, vy, vz);
      D.vx[v] = vx; D.vy[v] = vy; D.vz[v] = vz;
      classify_next(P, D.next_bits, D.dq, v, vx, vy, vz, tau, shift, fmax2);
    }
    __builtin_amdgcn_wave_barrier();   // the sums are read before the next cell's stores
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
  far_max_publish(&D.sc->next_fmax2, fmax2);
}
// the far max of a two-pass interval handed over (after both passes)
__global__ void co_far_handoff(CoD D) {
  if (threadIdx.x == 0) {
    D.sc->far_maxv2 = D.sc->next_fmax2;
    D.sc->next_fmax2 = 0ull;
  }
}

// Near set from particle positions (set-up, before any collision).
__global__ __launch_bounds__(256) void co_classify_all(CoP P, CoD D) {
  const int nd = P.nd;
  double fmax2 = 0.0, nmax2 = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.NT; i += gridDim.x * blockDim.x) {
    double x = D.x[i], y = D.y[i], z = D.z[i], vx = D.vx[i], vy = D.vy[i], vz = D.vz[i];
    double v2 = (vx * vx + vy * vy) + vz * vz;
    double reach = sqrt(v2) * P.tau * (1.0 + 1e-9) + P.guard + 1e-6;
    bool near = false;
    for (int s = 0; s < 2 * nd; s++) {
      int ii = s >= nd;
      double a = pbc_half(x - D.sc->x0[s], P.box[0]);
      double b = pbc_half(y - D.sc->y0[s], P.box[1]);
      double c = z - D.sc->z0[s];
      double r = P.full[ii] + reach;
      if (a * a + b * b + c * c <= r * r) near = true;
    }
    if (near) { atomicOr(&D.near_bits[i >> 5], 1u << (i & 31)); nmax2 = fmax(nmax2, v2); }
    else fmax2 = fmax(fmax2, v2);
  }
  fmax2 = wave_max(fmax2);
  nmax2 = wave_max(nmax2);
  __shared__ double red[8];
  if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6] = fmax2; red[4 + (threadIdx.x >> 6)] = nmax2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    double n = fmax(fmax(red[4], red[5]), fmax(red[6], red[7]));
    atomicMax(&D.sc->far_maxv2, (unsigned long long)__double_as_longlong(m));
    atomicMax((unsigned long long*)&D.sc->near_vmax2, (unsigned long long)__double_as_longlong(n));
  }
}

// interval start: collision counter, cleared cell counts, census and first-B
// keys, the spheres' start positions kept for co_far
// (clear, nclear: the next interval's bitmap of the two-pass collisions;
// handoff: the previous interval's far max, held in next_fmax2)
__global__ __launch_bounds__(256) void interval_begin(CoP P, CoD D, unsigned long long coll, unsigned* clear,
                                                      long long nclear, int handoff) {
  const long long mc = P.mpcnumcell, vol = P.vol;
  const long long i0 = (long long)blockIdx.x * 256 + threadIdx.x, st = (long long)gridDim.x * 256;
  for (long long i = i0; i < mc; i += st) D.cnt[i] = 0;
  for (long long i = i0; i < 3 * vol; i += st) D.tcount[i] = 0;
  for (long long i = i0; i < vol; i += st) D.firstB[i] = 0xffffffffu;
  if (clear)
    for (long long i = i0; i < nclear; i += st) clear[i] = 0u;
  if (i0 == 0) {
    D.sc->coll = coll; D.sc->n_wall = 0; D.sc->n_late = 0;
    if (handoff) { D.sc->far_maxv2 = D.sc->next_fmax2; D.sc->next_fmax2 = 0ull; }
  }
  if (i0 < 2 * P.nd) { D.sc->xs0[i0] = D.sc->x0[i0]; D.sc->ys0[i0] = D.sc->y0[i0]; D.sc->zs0[i0] = D.sc->z0[i0]; }
}
// before the rotation: cleared near bitmap (nwords + 1 words) and far max v^2
__global__ __launch_bounds__(256) void rotate_begin(CoD D, long long nwords) {
  const long long i0 = (long long)blockIdx.x * 256 + threadIdx.x, st = (long long)gridDim.x * 256;
  for (long long i = i0; i <= nwords; i += st) D.near_bits[i] = 0u;
  if (i0 == 0) D.sc->far_maxv2 = 0ull;
}

// --------------------------------------------------------------- Hilbert sort
// sort_hilbert_mod.f95:50-92: stable bin by tour position of the unshifted cell
// first and step are multiples of the wave size: every wave runs the loop as a
// whole, one counter atomic per distinct cell of the wave
__device__ __forceinline__ void sort_bin_body(const CoP& P, const CoD& D, const int* tourpos, int* scnt, int* key,
                                              unsigned short* rk, long long first, long long end, long long step) {
  for (long long base = first; base < end; base += step) {
    const long long i = base + threadIdx.x;
    int t = 0;
    bool ok = false;
    if (i < end) {
      const long long u = cell_unshifted0(D.x[i], D.y[i], D.z[i], P.box[0], P.box[1]);
      if (u < 0 || u >= P.vol || tourpos[u] < 0) { raise_error(&D.sc->err, ERR_CELL_RANGE); key[i] = 0; rk[i] = 0; }
      else { t = tourpos[u]; ok = true; }
    }
    const int r = wave_atomic_rank(scnt, t, ok);
    if (ok) {
      if (r >= BMAX) raise_error(&D.sc->err, ERR_BUCKET);
      key[i] = t;
      rk[i] = (unsigned short)min(r, 65535);
    }
  }
}
__global__ __launch_bounds__(256) void sort_bin(CoP P, CoD D, const int* tourpos, int* scnt, int* key,
                                                unsigned short* rk) {
  sort_bin_body(P, D, tourpos, scnt, key, rk, (long long)blockIdx.x * blockDim.x, P.NT,
                (long long)gridDim.x * blockDim.x);
}
__device__ __forceinline__ void sort_scatter_body(const int* key, const unsigned short* rk, const int* soff, int* sb,
                                                  long long first, long long end, long long step) {
  for (long long i = first + threadIdx.x; i < end; i += step)
    if (rk[i] < BMAX) sb[soff[key[i]] + rk[i]] = (int)i;
}
__global__ __launch_bounds__(256) void sort_scatter(int NT, const int* key, const unsigned short* rk, const int* soff,
                                                    int* sb) {
  sort_scatter_body(key, rk, soff, sb, (long long)blockIdx.x * blockDim.x, NT, (long long)gridDim.x * blockDim.x);
}
__device__ __forceinline__ void sort_gather_body(const CoD& D, const int* sb, double* nx, double* ny, double* nz,
                                                 double* nvx, double* nvy, double* nvz, int* nf, unsigned char* nst,
                                                 unsigned* nbits, long long first, long long end, long long step) {
  const int lane = threadIdx.x & 63;
  for (long long base = first; base < end; base += step) {
    const long long j = base + threadIdx.x;
    bool bit = false;
    if (j < end) {
      int i = sb[j];
      nx[j] = D.x[i]; ny[j] = D.y[i]; nz[j] = D.z[i];
      nvx[j] = D.vx[i]; nvy[j] = D.vy[i]; nvz[j] = D.vz[i];
      nf[j] = D.flag[i];
      nst[j] = D.status[i];
      bit = (D.near_bits[i >> 5] >> (i & 31)) & 1u;
    }
    unsigned long long m = __ballot(bit);   // the wave covers 64 consecutive j: two words
    const long long w0 = (base + (threadIdx.x & ~63)) >> 5, nw = (end + 31) >> 5;   // words below end
    if (lane == 0 && w0 < nw) nbits[w0] = (unsigned)m;
    if (lane == 32 && w0 + 1 < nw) nbits[w0 + 1] = (unsigned)(m >> 32);
  }
}

// one array of the permutation at a time: the working set of each pass (one
// source array) stays in the last-level cache, where the eight interleaved
// arrays of sort_gather did not
// the six FP64 state arrays in one pass: the source index read once, six
// loads in flight per element
struct Six { const double* s[6]; double* d[6]; };
__global__ __launch_bounds__(256) void sort_gather_six(long long n, const int* __restrict__ sb, Six a) {
  long long first, end;
  block_range(n, gridDim.x, xcd_block(blockIdx.x, gridDim.x), first, end);
  for (long long j = first + threadIdx.x; j < end; j += 256) {
    const int i = sb[j];
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = a.s[k][i];
#pragma unroll
    for (int k = 0; k < 6; k++) a.d[k][j] = v[k];
  }
}
template <typename T>
__global__ __launch_bounds__(256) void sort_gather_one(long long n, const int* __restrict__ sb,
                                                       const T* __restrict__ src, T* __restrict__ dst) {
  long long first, end;
  block_range(n, gridDim.x, xcd_block(blockIdx.x, gridDim.x), first, end);
  for (long long j = first + threadIdx.x; j < end; j += 256) dst[j] = src[sb[j]];
}
__global__ __launch_bounds__(256) void sort_gather_bits(CoP P, CoD D, const int* sb, unsigned char* nst,
                                                        unsigned* nbits) {
  long long first, end;
  block_range(P.NT, gridDim.x, xcd_block(blockIdx.x, gridDim.x), first, end);
  const int lane = threadIdx.x & 63;
  for (long long base = first; base < end; base += 256) {
    const long long j = base + threadIdx.x;
    bool bit = false;
    if (j < end) {
      const int i = sb[j];
      nst[j] = D.status[i];
      bit = (D.near_bits[i >> 5] >> (i & 31)) & 1u;
    }
    unsigned long long m = __ballot(bit);   // the wave covers 64 consecutive j: two words
    const long long w0 = (base + (threadIdx.x & ~63)) >> 5, nw = (end + 31) >> 5;   // words below end
    if (lane == 0 && w0 < nw) nbits[w0] = (unsigned)m;
    if (lane == 32 && w0 + 1 < nw) nbits[w0 + 1] = (unsigned)(m >> 32);
  }
}

}  // namespace mpcx

namespace mpcx {
// ------------------------------------------------------------- split precision
// Solvent stored as the unshifted cell with FP32 offsets in [0,1) inside it,
// FP32 velocities and an 8-bit species: 29 bytes read and written per
// particle and interval instead of 52.  Far particles stream, bounce, bin and
// rotate in FP32 cell coordinates (no large coordinates ever formed, walls
// handled relative to the wall's cell); the per-cell velocity sums are exact
// 64-bit fixed point (the split-precision intervals below).  Near particles
// and the dimers keep the FP64 near engines: their state is expanded into FP64
// before each near-engine launch and stored back after it.
__device__ __forceinline__ void sp_split(double x, double y, double z, const CoP& P, int& cell, float& ox, float& oy,
                                         float& oz) {
  int ix = (int)floor(x), iy = (int)floor(y), iz = (int)floor(z);
  float fx = (float)(x - (double)ix), fy = (float)(y - (double)iy), fz = (float)(z - (double)iz);
  if (fx >= 1.0f) { fx = 0.0f; ix += 1; }
  if (fy >= 1.0f) { fy = 0.0f; iy += 1; }
  if (fz >= 1.0f) { fz = 0.0f; iz += 1; }
  if (iz >= P.Lz) { iz = P.Lz - 1; fz = 0x1.fffffep-1f; }   // within half an FP32 ulp of the top wall
  if (ix >= P.Lx) ix -= P.Lx; else if (ix < 0) ix += P.Lx;
  if (iy >= P.Ly) iy -= P.Ly; else if (iy < 0) iy += P.Ly;
  cell = sp_pack(ix, iy, iz);
  ox = fx; oy = fy; oz = fz;
}

// a split particle whose cell lies in the box and whose offsets and
// velocity are finite (anything else -- say a velocity so large that its
// flight overflows the cell coordinates -- is reported, never binned)
__device__ __forceinline__ bool sp_sane(const CoP& P, int ix, int iy, int iz, float ox, float oy, float oz, float vx,
                                        float vy, float vz) {
  return ix >= 0 && ix < P.Lx && iy >= 0 && iy < P.Ly && iz >= 0 && iz < P.Lz && ox >= 0.0f && ox < 1.0f &&
         oy >= 0.0f && oy < 1.0f && oz >= 0.0f && oz < 1.0f && __builtin_isfinite(vx) && __builtin_isfinite(vy) &&
         __builtin_isfinite(vz);
}
// an FP64 position inside the box and a finite velocity.  x = Lx (y = Ly)
// is the periodic image of 0, which sp_split stores as cell 0, offset 0: the
// reference's wrap "if (x < 0) x = x + L" yields it when a flight ends a
// rounding error below 0, and a near particle whose FP32 start and velocity
// put its flight's end on the boundary exactly (x + v tau = 0) gets there
// with either sign of its FP64 steps' rounding.  z walls: 0 <= z <= Lz.
__device__ __forceinline__ bool fp64_sane(const CoP& P, double x, double y, double z, double vx, double vy, double vz) {
  return x >= 0.0 && x <= P.box[0] && y >= 0.0 && y <= P.box[1] && z >= 0.0 && z <= P.box[2] &&
         __builtin_isfinite(vx) && __builtin_isfinite(vy) && __builtin_isfinite(vz);
}

// FP64 staging <-> split storage, for all particles or the near list
__global__ __launch_bounds__(256) void sp_store(CoP P, CoD D, SpD S, int near_only) {
  const long long n = near_only ? (long long)min(D.sc->near_num, D.near_cap) : (long long)P.NT;
  for (long long j = (long long)blockIdx.x * 256 + threadIdx.x; j < n; j += (long long)gridDim.x * 256) {
    const int i = near_only ? D.near_list[j] : (int)j;
    if (!fp64_sane(P, D.x[i], D.y[i], D.z[i], D.vx[i], D.vy[i], D.vz[i])) {
      raise_error(&D.sc->err, ERR_CELL_RANGE);
      continue;
    }
    int c;
    float ox, oy, oz;
    sp_split(D.x[i], D.y[i], D.z[i], P, c, ox, oy, oz);
    S.cell[i] = c; S.ox[i] = ox; S.oy[i] = oy; S.oz[i] = oz;
    S.vx[i] = (float)D.vx[i]; S.vy[i] = (float)D.vy[i]; S.vz[i] = (float)D.vz[i];
    S.flag[i] = (signed char)D.flag[i];
  }
}
__global__ __launch_bounds__(256) void sp_load(CoP P, CoD D, SpD S, int near_only) {
  const long long n = near_only ? (long long)min(D.sc->near_num, D.near_cap) : (long long)P.NT;
  for (long long j = (long long)blockIdx.x * 256 + threadIdx.x; j < n; j += (long long)gridDim.x * 256) {
    const int i = near_only ? D.near_list[j] : (int)j;
    int ix, iy, iz;
    sp_cell_xyz(P, S.cell[i], ix, iy, iz);
    D.x[i] = (double)ix + (double)S.ox[i]; D.y[i] = (double)iy + (double)S.oy[i]; D.z[i] = (double)iz + (double)S.oz[i];
    D.vx[i] = (double)S.vx[i]; D.vy[i] = (double)S.vy[i]; D.vz[i] = (double)S.vz[i];
    D.flag[i] = (int)S.flag[i];
  }
}

__device__ __forceinline__ int sp_wrap(int i, int L) { return i >= L ? i - L : (i < 0 ? i + L : i); }
__device__ __forceinline__ void sp_carry(int& i, float& o, int L) {
  const float k = floorf(o);
  o = o - k;
  if (o >= 1.0f) { o = 0.0f; i += 1; }
  i = sp_wrap(i + (int)k, L);
}
// A position that rounds to z = Lz in FP32 (a flight ending within half an
// ulp below the top wall, a reflection within half an ulp of it, or z = Lz
// exactly, which the reference leaves unreflected) has no cell in [0, Lz):
// it goes to the top cell's last FP32 offset, 2^-24 below the wall.
__device__ __forceinline__ void sp_below_top(int& iz, float& oz, int Lz) {
  if (iz >= Lz) { iz = Lz - 1; oz = 0x1.fffffep-1f; }
}
// evolve_mod.f95:259-312 / :474-527 in cell coordinates (FP32): tc MD steps
// of free flight, bounce-back at the z walls, A<->B on a reactive patch of z = 0
__device__ __forceinline__ bool sp_stream(int& ix, int& iy, int& iz, float& ox, float& oy, float& oz, float& vx,
                                          float& vy, float& vz, int& flag, float dtc, const CoP& P,
                                          const int* react_box, CoScalars* sc) {
  ox = ox + vx * dtc;
  oy = oy + vy * dtc;
  oz = oz + vz * dtc;
  sp_carry(ix, ox, P.Lx);
  sp_carry(iy, oy, P.Ly);
  const float kz = floorf(oz);
  const int nz = iz + (int)kz;
  if (nz < 0) {   // below z = 0: z = iz + oz < 0
    const float z = (float)iz + oz;
    const float t = z / vz;
    float ax = ox - vx * t, ay = oy - vy * t;
    int hx = ix, hy = iy;
    sp_carry(hx, ax, P.Lx);
    sp_carry(hy, ay, P.Ly);
    const int hit = hx + P.Lx * hy + 1;
    if (hit < 1 || hit > P.nrb) raise_error_at(sc, ERR_CELL_RANGE, SITE_PATCH);
    else if (react_box[hit - 1] == 1) {
      if (P.react_dir == 0 && flag == 0) flag = 1;
      else if (P.react_dir == 1 && flag == 1) flag = 0;
    }
    vx = -vx; vy = -vy; vz = -vz;
    ox = ox + (2.0f * vx) * t; oy = oy + (2.0f * vy) * t;
    float zn = z + (2.0f * vz) * t;
    sp_carry(ix, ox, P.Lx);
    sp_carry(iy, oy, P.Ly);
    const float kk = floorf(zn);
    iz = (int)kk; oz = zn - kk;
    if (oz >= 1.0f) { oz = 0.0f; iz += 1; }
    return true;
  }
  const float over = (float)(nz - P.Lz) + (oz - kz);   // z - Lz
  if (nz >= P.Lz && over > 0.0f) {
    const float t = over / vz;
    vx = -vx; vy = -vy; vz = -vz;
    ox = ox + (2.0f * vx) * t; oy = oy + (2.0f * vy) * t;
    const float zr = over + (2.0f * vz) * t;   // z - Lz after the bounce
    sp_carry(ix, ox, P.Lx);
    sp_carry(iy, oy, P.Ly);
    const float kk = floorf(zr);
    iz = P.Lz + (int)kk; oz = zr - kk;
    if (oz >= 1.0f) { oz = 0.0f; iz += 1; }
    sp_below_top(iz, oz, P.Lz);
    return true;
  }
  iz = nz;
  oz = oz - kz;
  if (oz >= 1.0f) { oz = 0.0f; iz += 1; }
  sp_below_top(iz, oz, P.Lz);
  return false;
}

// co_rotate in split storage (FP32 rotation) and the next near set
// split storage: the rotation of co_rotate in FP32 with the record's mean and axis
__device__ __forceinline__ void sp_rotate_velocity(const CoP& P, double rcx, double rcy, double rcz, double rux,
                                                   double ruy, double ruz, float& vx, float& vy, float& vz) {
  const float cr = (float)P.cosrot, sr = (float)P.sinrot, omc = (float)P.omc;
  const float cx = (float)rcx, cy = (float)rcy, cz = (float)rcz;
  const float ux = (float)rux, uy = (float)ruy, uz = (float)ruz;
  const float dx = vx - cx, dy = vy - cy, dz = vz - cz;
  const float rx = (dx * (cr + (ux * ux) * omc) + dy * ((ux * uy) * omc - uz * sr)) + dz * ((ux * uz) * omc + uy * sr);
  const float ry = (dx * ((ux * uy) * omc + uz * sr) + dy * (cr + (uy * uy) * omc)) + dz * ((uy * uz) * omc - ux * sr);
  const float rz = (dx * ((ux * uz) * omc - uy * sr) + dy * ((uy * uz) * omc + ux * sr)) + dz * (cr + (uz * uz) * omc);
  vx = cx + rx; vy = cy + ry; vz = cz + rz;
}
// FP32 velocities: the far maximum is raised by a relative 1e-6 so time_penetrate stays conservative
__device__ __forceinline__ void sp_far_max_publish(unsigned long long* dst, float fmax2) {
  far_max_publish(dst, (double)fmax2 * (1.0 + 1e-6));
}
}  // namespace mpcx

// =====================================================================================
// Host side
// =====================================================================================
using namespace mpcx;

struct mpcx_co {
  mpcx_co_config cfg;
  CoP P;
  CoD D;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr, b0 = nullptr, b1 = nullptr, n0 = nullptr, n1 = nullptr;
  // co_far runs on st2 beside the near engine: st2 waits for fork (cleared
  // counters, draws), st waits for join before binning the near set
  hipStream_t st2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, far_done = nullptr;   // far_done: co_far finished (two passes)
  bool overlap = true;
  int epoch = 0;   // intervals collided in two passes (the cells' late stamps)
  // after a two-pass interval: the far max waits in next_fmax2 and the near
  // list is built at the start of the next interval (beside its co_far)
  bool handoff = false, list_due = false;
  double run_ms = 0, bulk_ms = 0, near_ms = 0;
  // timed kernels, timed intervals, and the sampling: every TSTRIDE-th interval
  // carries the kernel-timing events (each event costs the stream a few us)
  static constexpr int NKT = 5, TMAX = 256, TSTRIDE = 8;
  std::vector<hipEvent_t> tev;                // [TMAX][NKT][2]
  std::vector<char> tev_on;                   // [TMAX][NKT]: pair recorded
  double kt_ms[NKT] = {0};
  long long kt_n = 0;
  long long ndraw = 0;
  double* d_U = nullptr;
  int* d_tourpos = nullptr;
  bool has_tour = false;
  // sort buffers
  int *d_skey = nullptr, *d_scnt = nullptr, *d_soff = nullptr, *d_sb = nullptr, *d_spart = nullptr;
  unsigned short* d_srk = nullptr;
  double *alt[6] = {nullptr};
  int* alt_flag = nullptr;
  unsigned char* alt_status = nullptr;
  unsigned* alt_bits = nullptr;
  std::vector<void*> allocs;
  long long nwords = 0;
  int nd = 1;
  std::vector<int> mpc_pos, mpc_cell;   // host copies of the collision-cell tour
  // near engine with two or more dimers: 0 one workgroup (up to four
  // dimers), 1 one workgroup per cluster of dimers, 2 the system's near set
  // over G cooperative workgroups
  int near_mode = 0;
  bool clustered = false, can_cluster = false, has_ge = false;
  CluBase cb;
  CoD Dg;          // the system's descriptor with the global engine's entry arrays
  GeD E;
  int geG = 1;
  // split precision: solvent in cell + FP32 offsets / FP32 velocities / 8-bit species
  bool split = false;
  SpD S, Salt;
  bool split_dirty = false;   // the accumulator copies may hold partial sums of a run that failed
  mpcx_cob* solo = nullptr;   // split runs: this engine as a batch of one (made on the first run)
  double* d_rsum = nullptr;   // [4][MAXSG] per-sphere totals of the restart evaluation
  int g_bulk = 0, g_spbulk = 0;   // co_bulk / sp_bulk workgroups (0: not chosen yet)
  bool cells_valid = false;   // a collision has binned the particles since the last upload
};

template <typename T>
static T* dalloc(mpcx_co* e, size_t n) {
  void* p = nullptr;
  if (hipMalloc(&p, n * sizeof(T) + 64) != hipSuccess) return nullptr;
  e->allocs.push_back(p);
  return (T*)p;
}

static int check_device_error(mpcx_co* e, int replica = -1) {
  int ev[2] = {0, 0}, idx = -1;   // err, err_site; err_idx
  if (hipMemcpy(ev, &e->D.sc->err, sizeof(ev), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&idx, &e->D.sc->err_idx, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(MPCX_E_HIP, "hipMemcpy of the error flag failed");
  const int err = ev[0], site = ev[1];
  if (err) {
    static const char* msg[] = {"none",
                                "a particle maps outside the cell arrays",
                                "a cell holds more members than the reference arrays (100)",
                                "inside set larger than the near engine capacity (16384)",
                                "a sphere moved further than the near/far guard within one interval",
                                "a particle belongs to more than 4 neighbour lists",
                                "a collision cell holds more than 64 particles",
                                "near set larger than its buffer",
                                "a neighbour list is longer than max_neighbour",
                                "more than 64 list rebuilds in one interval",
                                "the dimer bond constraint did not converge",
                                "a neighbour-list member is missing from the near set",
                                "the host draw buffer ran out",
                                "more than four dimers lie within reach of each other (one cluster)",
                                "the clusters of the system fell out of step",
                                "a neighbour-list member was carried to a cluster without its dimer",
                                "no place outside the spheres found for a solvent particle",
                                "the rebuild events of an interval do not add up to the interval",
                                "more far particles reach a wall in one interval than the deferred list holds",
                                "a near or deferred particle ended in a cell collided before the near engine finished",
                                "more late collision cells than their list holds"};
    static const char* where[SITE_COUNT] = {"",
                                            "sb_bin (shifted cell)",
                                            "sb_bin (census cell)",
                                            "sb_bulk (free flight)",
                                            "sb_bulk (state after a re-sort)",
                                            "sb_near (near set back into split storage)",
                                            "sb_cells_own (inbox entry of a cell the tile does not own)",
                                            "sb_rotate (collision cell)",
                                            "sb_sort_bin (unshifted cell)",
                                            "sp_stream (reactive patch)"};
    std::string m = std::string("device check failed: ") + (err > 0 && err < 21 ? msg[err] : "unknown");
    if (site > 0 && site < SITE_COUNT) {
      m += std::string(" [") + where[site];
      if (idx >= 0) m += ", particle " + std::to_string(idx) + " (0-based)";
      m += "]";
    }
    if (replica >= 0) m += " [replica " + std::to_string(replica) + "]";
    const int zero[2] = {0, 0};
    (void)hipMemcpy(&e->D.sc->err, zero, sizeof(zero), hipMemcpyHostToDevice);
    return set_error(MPCX_E_DEVICE, m.c_str());
  }
  return 0;
}

#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t _r = (x);                                                 \
    if (_r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(_r)); \
  } while (0)

// The global near engine's arrays (entries of the whole system): made at
// creation above two dimers, or when the mode is selected.
static int ge_alloc(mpcx_co* e) {
  if (e->has_ge) return 0;
  CoD& D = e->D;
  const long long NT = e->P.NT;
  CoD& G = e->Dg;
  G = D;   // particle arrays, scalars and cell data are shared (refreshed at each launch)
  const size_t nq = (size_t)D.near_cap;
  const int icap = (int)std::min<long long>(NT, (long long)NE_CAP + 8192LL * e->nd);
  G.ns_x = dalloc<double>(e, nq); G.ns_y = dalloc<double>(e, nq); G.ns_z = dalloc<double>(e, nq);
  G.ns_vx = dalloc<double>(e, nq); G.ns_vy = dalloc<double>(e, nq); G.ns_vz = dalloc<double>(e, nq);
  G.ns_flag = dalloc<int>(e, nq); G.ns_stat = dalloc<unsigned char>(e, nq);
  G.in_cap = icap;
  G.in_list = dalloc<int>(e, icap); G.in_plist = dalloc<int>(e, icap); G.in_msd = dalloc<double>(e, icap);
  G.sl_x = dalloc<double>(e, icap); G.sl_y = dalloc<double>(e, icap); G.sl_z = dalloc<double>(e, icap);
  G.sl_vx = dalloc<double>(e, icap); G.sl_vy = dalloc<double>(e, icap); G.sl_vz = dalloc<double>(e, icap);
  G.sl_flag = dalloc<int>(e, icap);
  GeD& E = e->E;
  E.sy = dalloc<GeSync>(e, 1);
  E.memb_q = dalloc<unsigned>(e, nq);
  E.in_memb4 = dalloc<unsigned>(e, icap);
  E.in_fp4 = dalloc<double>(e, (size_t)icap * 12);
  E.in_pot4 = dalloc<double>(e, (size_t)icap * 4);
  E.part = dalloc<double>(e, (size_t)2 * GE_GMAX * (4 * MAXSG + 1));
  E.touch = dalloc<unsigned>(e, (size_t)2 * MAXSG * (GE_GMAX / 32));
  E.in_cap = icap;
  if (!G.ns_x || !G.sl_flag || !E.sy || !E.touch || !E.in_fp4) return set_error(MPCX_E_HIP, "device allocation failed");
  e->geG = std::min(GE_GMAX, std::max(2, 2 * e->nd));
  HIPCHK(hipMemset(E.sy, 0, sizeof(GeSync)));
  HIPCHK(hipMemset(E.touch, 0, (size_t)2 * MAXSG * (GE_GMAX / 32) * 4));
  HIPCHK(hipMemset(E.in_fp4, 0, (size_t)icap * 12 * 8));
  HIPCHK(hipMemset(E.in_pot4, 0, (size_t)icap * 4 * 8));
  HIPCHK(hipMemset(G.in_msd, 0, (size_t)icap * 8));
  e->has_ge = true;
  return 0;
}

extern "C" int mpcx_co_create(mpcx_co** out, const mpcx_co_config* cfg) {
  if (!out || !cfg) return set_error(MPCX_E_ARG, "null argument");
  *out = nullptr;
  if (cfg->numdims < 1 || cfg->numdims > MAXDG)
    return set_error(MPCX_E_UNSUPPORTED, "the device-resident collective engine handles 1 to 120 dimers");
  if (cfg->NT < 1 || cfg->NT > 0x7fffff00LL || cfg->vol < 1) return set_error(MPCX_E_ARG, "bad sizes");
  if (cfg->sort_step % cfg->MPCtstep != 0)
    return set_error(MPCX_E_UNSUPPORTED, "sort_step must be a multiple of MPCtstep for the device-resident engine");
  if (cfg->a0 != 1.0) return set_error(MPCX_E_UNSUPPORTED, "the reference's cell index assumes a0 = 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(MPCX_E_HIP, "no HIP device");
  mpcx_co* e = new mpcx_co();
  e->cfg = *cfg;
  e->nd = (int)cfg->numdims;
  CoP& P = e->P;
  memset(&P, 0, sizeof P);
  P.NT = (int)cfg->NT; P.nd = (int)cfg->numdims; P.MPCtstep = cfg->MPCtstep; P.sort_step = cfg->sort_step;
  P.react_dir = cfg->react_dir; P.rng_mode = cfg->rng_mode; P.seed = cfg->seed;
  P.Lx = (int)cfg->box[0]; P.Ly = (int)cfg->box[1]; P.Lz = (int)cfg->box[2];
  P.nrb = P.Lx * P.Ly;
  P.vol = cfg->vol; P.mpcnumcell = cfg->mpcnumcell; P.rxnnumcell = cfg->rxnnumcell;
  for (int i = 0; i < 3; i++) P.box[i] = cfg->box[i];
  const double MDt = cfg->MDt;
  P.MDt = MDt;
  P.hA = (0.5 * MDt) / cfg->Ma; P.hB = (0.5 * MDt) / cfg->Mb; P.hS = (0.5 * MDt) / cfg->Ms;
  P.hc = (0.5 * MDt) / cfg->Mc; P.hn = (0.5 * MDt) / cfg->Mn;
  P.IIMcIMn = 1.0 / ((1.0 / cfg->Mc) + (1.0 / cfg->Mn));
  P.Mc = cfg->Mc; P.Mn = cfg->Mn; P.Dcn2 = cfg->Dcn * cfg->Dcn;
  const double R[2] = {cfg->Rc, cfg->Rn};
  for (int i = 0; i < 2; i++) {
    double cut = R[i] * cfg->cutoff;
    P.cut2[i] = cut * cut;
    P.full[i] = cut + cfg->skinrad;
    P.full2[i] = P.full[i] * P.full[i];
    P.R6[i] = ipow(R[i], 6); P.R12[i] = ipow(R[i], 12);
    P.r6R6[i] = 6.0 * P.R6[i]; P.r12R12[i] = 12.0 * P.R12[i];
  }
  P.Ea[0] = cfg->EaC; P.Eb[0] = cfg->EbC; P.Es[0] = cfg->EsC;
  P.Ea[1] = cfg->EaN; P.Eb[1] = cfg->EbN; P.Es[1] = cfg->EsN;
  P.Em = cfg->Em;
  const double dd[3] = {cfg->Rc + cfg->Rc + cfg->Rm_delta, cfg->Rc + cfg->Rn + cfg->Rm_delta,
                        cfg->Rn + cfg->Rn + cfg->Rm_delta};
  for (int j = 0; j < 3; j++) {
    P.mm_R6[j] = ipow(dd[j], 6); P.mm_R12[j] = ipow(dd[j], 12);
    double c = dd[j] * cfg->cutoff;
    P.mm_cut2[j] = c * c;
  }
  P.Ew = cfg->Ew;
  for (int i = 0; i < 2; i++) {
    double w = cfg->Rw + R[i];
    P.wR3[i] = ipow(w, 3); P.wR9[i] = ipow(w, 9);
    P.wA[i] = 3.0 * P.wR3[i];   // received as r9R9 (evolve_mod.f95:597-600)
    P.wB[i] = 9.0 * P.wR9[i];   // received as r3R3
    double c = w * cfg->wall_w6;
    P.wcut2[i] = c * c;
  }
  P.pref = 1.5 * sqrt(3.0);
  P.hs2 = (0.5 * cfg->skinrad) * (0.5 * cfg->skinrad);
  P.a0 = cfg->a0; P.cosrot = cfg->cosrot; P.sinrot = cfg->sinrot; P.omc = 1.0 - cfg->cosrot;
  P.twopi = 2.0 * (4.0 * atan(1.0));
  P.tau = (double)cfg->MPCtstep * MDt;
  P.guard = cfg->sphere_guard > 0 ? cfg->sphere_guard : 0.3;
  P.skinrad = cfg->skinrad;

  const long long NT = cfg->NT, vol = cfg->vol, mc = cfg->mpcnumcell;
  CoD& D = e->D;
  memset(&D, 0, sizeof D);
  D.x = dalloc<double>(e, NT); D.y = dalloc<double>(e, NT); D.z = dalloc<double>(e, NT);
  D.vx = dalloc<double>(e, NT); D.vy = dalloc<double>(e, NT); D.vz = dalloc<double>(e, NT);
  D.flag = dalloc<int>(e, NT);
  D.dim = dalloc<double>(e, 12 * MAXSG);
  D.sc = dalloc<CoScalars>(e, 1);
  D.react_box = dalloc<int>(e, std::max(1, P.nrb));
  D.in_cap = NE_CAP;
  e->can_cluster = e->nd >= 2;
  e->near_mode = e->nd > 2 ? 2 : 0;
  e->clustered = false;
  // the system's near list holds every cluster's near set
  const long long nsq = std::min<long long>(NT, NE_NEAR);
  D.near_cap = (int)std::min<long long>(NT, (long long)NE_NEAR * (e->can_cluster ? e->nd : 1));
  D.near_list = dalloc<int>(e, D.near_cap);
  // far particles reaching a wall within one interval: about <|vz|> tau / Lz of
  // them (0.2 % at 100^3); the list holds 1/16 of the particles, at least 65536
  D.wall_cap = (int)std::min<long long>(NT, std::max<long long>(65536, NT / 16));
  D.wall_list = dalloc<int>(e, D.wall_cap);
  D.ns_x = dalloc<double>(e, nsq); D.ns_y = dalloc<double>(e, nsq); D.ns_z = dalloc<double>(e, nsq);
  D.ns_vx = dalloc<double>(e, nsq); D.ns_vy = dalloc<double>(e, nsq);
  D.ns_vz = dalloc<double>(e, nsq);
  D.ns_flag = dalloc<int>(e, nsq);
  D.ns_stat = dalloc<unsigned char>(e, nsq);
  if (e->can_cluster) {
    const size_t nc = (size_t)e->nd, nq = nc * NE_NEAR, ni = nc * NE_CAP;
    CluBase& B = e->cb;
    memset(&B, 0, sizeof B);
    D.clu = dalloc<CluInfo>(e, 1);
    B.Ps = dalloc<CoP>(e, nc); B.Ds = dalloc<CoD>(e, nc); B.scs = dalloc<CoScalars>(e, nc);
    B.dimc = dalloc<double>(e, nc * 12 * MAXSG);
    B.near_list = dalloc<int>(e, nq);
    B.ns_x = dalloc<double>(e, nq); B.ns_y = dalloc<double>(e, nq); B.ns_z = dalloc<double>(e, nq);
    B.ns_vx = dalloc<double>(e, nq); B.ns_vy = dalloc<double>(e, nq); B.ns_vz = dalloc<double>(e, nq);
    B.ns_flag = dalloc<int>(e, nq); B.ns_stat = dalloc<unsigned char>(e, nq);
    B.in_list = dalloc<int>(e, ni); B.in_plist = dalloc<int>(e, ni); B.in_memb = dalloc<unsigned>(e, ni);
    B.in_fp = dalloc<double>(e, ni * MAXM * 3); B.in_pot = dalloc<double>(e, ni * MAXM);
    B.in_msd = dalloc<double>(e, ni);
    B.sl_x = dalloc<double>(e, ni); B.sl_y = dalloc<double>(e, ni); B.sl_z = dalloc<double>(e, ni);
    B.sl_vx = dalloc<double>(e, ni); B.sl_vy = dalloc<double>(e, ni); B.sl_vz = dalloc<double>(e, ni);
    B.sl_flag = dalloc<int>(e, ni);
    B.assign = dalloc<unsigned char>(e, (size_t)D.near_cap);
    D.rec_ids = dalloc<unsigned>(e, NT);
    D.rec_msd = dalloc<double>(e, NT);
    D.rec_fp = dalloc<double>(e, (size_t)NT * 12);
    D.rec_pot = dalloc<double>(e, (size_t)NT * 4);
  }
  e->nwords = (NT + 31) / 32;
  D.near_bits = dalloc<unsigned>(e, e->nwords + 1);
  D.next_bits = dalloc<unsigned>(e, e->nwords + 1);
  // two-pass collisions: a slab of BMAX slots per collision cell, the cells'
  // late stamps and the list of late cells (each cell at most once)
  D.slab = dalloc<int>(e, (size_t)P.mpcnumcell * BMAX);
  D.stamp = dalloc<int>(e, (size_t)P.mpcnumcell);
  D.late = dalloc<int>(e, (size_t)P.mpcnumcell);
  D.late_cap = (int)P.mpcnumcell;
  D.status = dalloc<unsigned char>(e, NT);
  D.in_list = dalloc<int>(e, D.in_cap);
  D.in_plist = dalloc<int>(e, D.in_cap);
  D.in_memb = dalloc<unsigned>(e, D.in_cap);
  D.in_fp = dalloc<double>(e, (size_t)MAXM * 3 * D.in_cap);
  D.in_pot = dalloc<double>(e, (size_t)MAXM * D.in_cap);
  D.in_msd = dalloc<double>(e, D.in_cap);
  D.sl_x = dalloc<double>(e, D.in_cap); D.sl_y = dalloc<double>(e, D.in_cap); D.sl_z = dalloc<double>(e, D.in_cap);
  D.sl_vx = dalloc<double>(e, D.in_cap); D.sl_vy = dalloc<double>(e, D.in_cap); D.sl_vz = dalloc<double>(e, D.in_cap);
  D.sl_flag = dalloc<int>(e, D.in_cap);
  if (e->nd >= 2) {
    int rc = ge_alloc(e);
    if (rc) { mpcx_co_destroy(e); return rc; }
  }
  D.cell_of = dalloc<int>(e, NT);
  D.rank_of = dalloc<unsigned short>(e, NT);
  D.cnt = dalloc<int>(e, mc);
  D.off = dalloc<int>(e, mc + 1);
  D.bucket = dalloc<int>(e, NT);
  long long npart = std::max((mc + 4095) / 4096, std::max((vol + 4095) / 4096, (e->nwords + 1023) / 1024)) + 2;
  D.partial = dalloc<int>(e, npart);
  D.cellrec = dalloc<double>(e, mc * 8);
  D.dq = dalloc<unsigned char>(e, NT);
  int* mpos = dalloc<int>(e, mc);
  int* mcell = dalloc<int>(e, mc);
  D.mpc_pos = mpos;
  D.mpc_cell = mcell;
  D.tcount = dalloc<int>(e, 3 * vol);
  D.firstB = dalloc<unsigned>(e, vol);
  double* ptab = dalloc<double>(e, 101);
  D.ptab = ptab;
  e->ndraw = 3 + 2 * mc + cfg->rxnnumcell;
  e->d_U = dalloc<double>(e, cfg->rng_mode == 0 ? e->ndraw : 4);
  e->d_rsum = dalloc<double>(e, 4 * MAXSG);
  D.U = e->d_U;
  e->d_tourpos = dalloc<int>(e, vol);
  e->d_skey = dalloc<int>(e, NT); e->d_srk = dalloc<unsigned short>(e, NT);
  e->d_scnt = dalloc<int>(e, vol); e->d_soff = dalloc<int>(e, vol + 1); e->d_sb = dalloc<int>(e, NT);
  for (int j = 0; j < 6; j++) e->alt[j] = dalloc<double>(e, NT);
  e->alt_flag = dalloc<int>(e, NT);
  e->alt_status = dalloc<unsigned char>(e, NT);
  e->alt_bits = dalloc<unsigned>(e, e->nwords + 1);
  for (void* p : e->allocs)
    if (!p) { mpcx_co_destroy(e); return set_error(MPCX_E_HIP, "device allocation failed"); }
  if (std::find(e->allocs.begin(), e->allocs.end(), nullptr) != e->allocs.end() || !D.x || !e->alt_bits) {
    mpcx_co_destroy(e);
    return set_error(MPCX_E_HIP, "device allocation failed");
  }
  HIPCHK(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&e->st2, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&e->join, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&e->far_done, hipEventDisableTiming));
  HIPCHK(hipEventCreate(&e->e0)); HIPCHK(hipEventCreate(&e->e1));
  HIPCHK(hipEventCreate(&e->b0)); HIPCHK(hipEventCreate(&e->b1));
  HIPCHK(hipEventCreate(&e->n0)); HIPCHK(hipEventCreate(&e->n1));
  e->tev_on.assign((size_t)mpcx_co::TMAX * mpcx_co::NKT, 0);
  HIPCHK(hipMemcpy(ptab, cfg->reaction_p, 101 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(D.sc, 0, sizeof(CoScalars)));
  HIPCHK(hipMemset(D.status, 0, NT));
  HIPCHK(hipMemset(D.cell_of, 0, NT * 4));
  HIPCHK(hipMemset(D.near_bits, 0, (e->nwords + 1) * 4));
  HIPCHK(hipMemset(D.next_bits, 0, (e->nwords + 1) * 4));
  HIPCHK(hipMemset(D.stamp, 0, (size_t)P.mpcnumcell * 4));
  HIPCHK(hipMemset(e->d_soff, 0, ((size_t)vol + 1) * 4));   // all zero: a valid cell ownership before any sort
  D.soff = e->d_soff;
  {   // collision-cell tour maps: the identity until a state with its tour is uploaded
    e->mpc_pos.resize((size_t)mc);
    e->mpc_cell.resize((size_t)mc);
    for (long long c = 0; c < mc; c++) { e->mpc_pos[c] = (int)c; e->mpc_cell[c] = (int)c; }
    HIPCHK(hipMemcpy((void*)D.mpc_pos, e->mpc_pos.data(), mc * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void*)D.mpc_cell, e->mpc_cell.data(), mc * 4, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemset(D.in_fp, 0, (size_t)MAXM * 3 * D.in_cap * 8));
  HIPCHK(hipMemset(D.in_pot, 0, (size_t)MAXM * D.in_cap * 8));
  HIPCHK(hipMemset(D.in_msd, 0, (size_t)D.in_cap * 8));
  if (e->can_cluster) {
    HIPCHK(hipMemset(D.clu, 0, sizeof(CluInfo)));
    HIPCHK(hipMemset(D.rec_ids, 0xff, (size_t)NT * 4));
    HIPCHK(hipMemset(D.rec_msd, 0, (size_t)NT * 8));
  }
  *out = e;
  handle_add(H_CO, e);
  return 0;
}

// 0: one workgroup for the whole system (up to four dimers); 1: one
// workgroup per cluster of dimers (clusters of up to four); 2: the system's
// near set over cooperative workgroups (any number of dimers)
// 1: split-precision solvent storage (device Philox draws only); call before
// mpcx_co_setup, which converts the uploaded FP64 state
extern "C" int mpcx_co_set_precision(mpcx_co* e, int32_t split) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  if (split && e->P.rng_mode != 1) return set_error(MPCX_E_UNSUPPORTED, "split precision runs with device Philox draws");
  if (split && (e->P.Lx > 2048 || e->P.Ly > 2048 || e->P.Lz > 1024))
    return set_error(MPCX_E_UNSUPPORTED, "split precision holds boxes up to 2048 x 2048 x 1024 cells");
  if (split && !e->S.cell) {
    const size_t NT = (size_t)e->P.NT;
    for (SpD* q : {&e->S, &e->Salt}) {
      q->cell = dalloc<int>(e, NT);
      q->ox = dalloc<float>(e, NT); q->oy = dalloc<float>(e, NT); q->oz = dalloc<float>(e, NT);
      q->vx = dalloc<float>(e, NT); q->vy = dalloc<float>(e, NT); q->vz = dalloc<float>(e, NT);
      q->flag = dalloc<signed char>(e, NT);
      if (!q->cell || !q->ox || !q->oy || !q->oz || !q->vx || !q->vy || !q->vz || !q->flag)
        return set_error(MPCX_E_HIP, "device allocation failed");
    }
    const size_t mc = (size_t)e->P.mpcnumcell, vol = (size_t)e->P.vol;
    e->D.sacc = dalloc<unsigned long long>(e, NXCD * 3 * mc);
    e->D.sown = dalloc<unsigned long long>(e, 3 * mc);
    e->D.srec = dalloc<float4>(e, 2 * mc);
    e->D.xtc = dalloc<int>(e, NXCD * vol);
    e->D.xfb = dalloc<unsigned>(e, NXCD * vol);
    if (!e->D.sacc || !e->D.sown || !e->D.srec || !e->D.xtc || !e->D.xfb)
      return set_error(MPCX_E_HIP, "device allocation failed");
    HIPCHK(hipMemset(e->D.sacc, 0, NXCD * 3 * mc * 8));
    HIPCHK(hipMemset(e->D.sown, 0, 3 * mc * 8));
    HIPCHK(hipMemset(e->D.srec, 0, 2 * mc * 16));
    HIPCHK(hipMemset(e->D.xtc, 0, NXCD * vol * 4));
    HIPCHK(hipMemset(e->D.xfb, 0xff, NXCD * vol * 4));
  }
  e->split = split != 0;
  return 0;
}

// workgroups of the global near engine (mode 2): 1..128
extern "C" int mpcx_co_set_near_workgroups(mpcx_co* e, int32_t G) {
  if (!e || G < 1 || G > GE_GMAX) return set_error(MPCX_E_ARG, "1 to 128 workgroups");
  e->geG = G;
  return 0;
}

extern "C" int mpcx_co_set_near_mode(mpcx_co* e, int32_t mode) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  if (mode < 0 || mode > 2) return set_error(MPCX_E_ARG, "near mode is 0, 1 or 2");
  if (mode == 1 && !e->can_cluster) return set_error(MPCX_E_UNSUPPORTED, "clusters need two or more dimers");
  if (!mode && e->nd > MAXD) return set_error(MPCX_E_UNSUPPORTED, "one workgroup handles up to four dimers");
  if (mode == 2) {
    int rc = ge_alloc(e);
    if (rc) return rc;
  }
  e->near_mode = mode;
  e->clustered = mode == 1;
  return 0;
}

extern "C" int mpcx_co_destroy(mpcx_co* e) {
  if (!e) return 0;
  handle_del(H_CO, e);
  if (e->solo) { (void)mpcx_cob_destroy(e->solo); e->solo = nullptr; }
  if (e->st) (void)hipStreamSynchronize(e->st);
  if (e->st2) (void)hipStreamSynchronize(e->st2);
  for (void* p : e->allocs)
    if (p) (void)hipFree(p);
  if (e->st) (void)hipStreamDestroy(e->st);
  if (e->st2) (void)hipStreamDestroy(e->st2);
  for (hipEvent_t ev : {e->e0, e->e1, e->b0, e->b1, e->n0, e->n1, e->fork, e->join, e->far_done})
    if (ev) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : e->tev)
    if (ev) (void)hipEventDestroy(ev);
  delete e;
  return 0;
}

extern "C" int mpcx_co_set_state(mpcx_co* e, int32_t cstep, const double* x, const double* y, const double* z,
                                 const double* vx, const double* vy, const double* vz, const int32_t* flag,
                                 const double* xd, const double* yd, const double* zd, const double* vxd,
                                 const double* vyd, const double* vzd, const double* fxd, const double* fyd,
                                 const double* fzd, const int32_t* react_box, const int64_t* tour) {
  // the solvent arrays may all be NULL: mpcx_co_init_solvent then fills them on the device
  const bool solvent = x || y || z || vx || vy || vz || flag;
  if ((solvent && (!x || !y || !z || !vx || !vy || !vz || !flag)) || !xd || !yd || !zd)
    return set_error(MPCX_E_ARG, "null argument");
  const size_t NT = (size_t)e->P.NT, nd = (size_t)e->nd;
  CoD& D = e->D;
  if (solvent) {
    HIPCHK(hipMemcpy(D.x, x, NT * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.y, y, NT * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.z, z, NT * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.vx, vx, NT * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.vy, vy, NT * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.vz, vz, NT * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.flag, flag, NT * 4, hipMemcpyHostToDevice));
  }
  std::vector<double> dim(12 * MAXSG, 0.0);
  const double* src[9] = {xd, yd, zd, vxd, vyd, vzd, fxd, fyd, fzd};
  for (int f = 0; f < 9; f++)
    for (size_t s = 0; s < 2 * nd; s++) dim[f * MAXSG + s] = src[f] ? src[f][s] : 0.0;
  HIPCHK(hipMemcpy(D.dim, dim.data(), dim.size() * 8, hipMemcpyHostToDevice));
  std::vector<int> rb((size_t)std::max(1, e->P.nrb), 0);
  if (react_box)
    for (int i = 0; i < e->P.nrb; i++) rb[i] = react_box[i];
  HIPCHK(hipMemcpy((void*)D.react_box, rb.data(), rb.size() * 4, hipMemcpyHostToDevice));
  e->has_tour = false;
  // collision cells are visited in the tour order of the unshifted cells they
  // share an index with (x + Lx*y + Lx*Ly*z, z < Lz), then the extra z layer
  {
    const long long mc = e->P.mpcnumcell, vol = e->P.vol;
    e->mpc_cell.assign((size_t)mc, 0);
    e->mpc_pos.assign((size_t)mc, -1);
    long long t = 0;
    if (tour)
      for (long long j = 0; j < vol; j++) {
        long long c = tour[j] - 1;
        if (c >= 0 && c < vol && c < mc && e->mpc_pos[c] < 0) { e->mpc_pos[c] = (int)t; e->mpc_cell[t] = (int)c; t++; }
      }
    for (long long c = 0; c < mc; c++)
      if (e->mpc_pos[c] < 0) { e->mpc_pos[c] = (int)t; e->mpc_cell[t] = (int)c; t++; }
    HIPCHK(hipMemcpy((void*)D.mpc_pos, e->mpc_pos.data(), mc * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void*)D.mpc_cell, e->mpc_cell.data(), mc * 4, hipMemcpyHostToDevice));
  }
  if (tour) {
    std::vector<int> tp((size_t)e->P.vol, -1);
    for (long long t = 0; t < e->P.vol; t++) {
      long long c = tour[t] - 1;
      if (c < 0 || c >= e->P.vol || tp[c] >= 0)
        return set_error(MPCX_E_ARG, "the tour must visit every cell exactly once");
      tp[c] = (int)t;
    }
    HIPCHK(hipMemcpy(e->d_tourpos, tp.data(), tp.size() * 4, hipMemcpyHostToDevice));
    e->has_tour = true;
  }
  CoScalars sc;
  memset(&sc, 0, sizeof sc);
  sc.cstep = cstep;
  HIPCHK(hipMemcpy(D.sc, &sc, sizeof sc, hipMemcpyHostToDevice));
  e->cells_valid = false;
  return 0;
}

static int co_sort(mpcx_co* e, hipStream_t st) {
  CoD& D = e->D;
  CoP& P = e->P;
  const long long NT = P.NT, vol = P.vol;
  if (!e->has_tour) return set_error(MPCX_E_ARG, "a sort step needs the cell tour (pass it to mpcx_co_set_state)");
  hipLaunchKernelGGL(fill_kernel<int>, dim3(grid_for(vol, 256)), dim3(256), 0, st, e->d_scnt, vol, 0);
  hipLaunchKernelGGL(sort_bin, dim3(grid_for(NT, 256)), dim3(256), 0, st, P, D, (const int*)e->d_tourpos, e->d_scnt,
                     e->d_skey, e->d_srk);
  launch_scan(st, e->d_scnt, vol, D.partial, e->d_soff);
  hipLaunchKernelGGL(fill_kernel<int>, dim3(grid_for(NT, 256)), dim3(256), 0, st, e->d_sb, NT, 0);
  hipLaunchKernelGGL(sort_scatter, dim3(grid_for(NT, 256)), dim3(256), 0, st, (int)NT, (const int*)e->d_skey,
                     (const unsigned short*)e->d_srk, (const int*)e->d_soff, e->d_sb);
  hipLaunchKernelGGL(sort_order, dim3((unsigned)((vol + 255) / 256)), dim3(256), 0, st, vol, (const int*)e->d_scnt,
                     (const int*)e->d_soff, e->d_sb, BMAX);
  hipLaunchKernelGGL(fill_kernel<unsigned>, dim3(grid_for(e->nwords + 1, 256)), dim3(256), 0, st, e->alt_bits,
                     e->nwords + 1, 0u);
  {
    const int g = grid_for(NT, 256);
    const int* sb = e->d_sb;
    Six six = {{D.x, D.y, D.z, D.vx, D.vy, D.vz}, {e->alt[0], e->alt[1], e->alt[2], e->alt[3], e->alt[4], e->alt[5]}};
    hipLaunchKernelGGL(sort_gather_six, dim3(g), dim3(256), 0, st, NT, sb, six);
    hipLaunchKernelGGL(sort_gather_one<int>, dim3(g), dim3(256), 0, st, NT, sb, (const int*)D.flag, e->alt_flag);
    hipLaunchKernelGGL(sort_gather_bits, dim3(g), dim3(256), 0, st, P, D, sb, e->alt_status, e->alt_bits);
  }
  std::swap(D.x, e->alt[0]); std::swap(D.y, e->alt[1]); std::swap(D.z, e->alt[2]);
  std::swap(D.vx, e->alt[3]); std::swap(D.vy, e->alt[4]); std::swap(D.vz, e->alt[5]);
  std::swap(D.flag, e->alt_flag); std::swap(D.status, e->alt_status); std::swap(D.near_bits, e->alt_bits);
  return 0;
}

// split precision: the near set (or every particle) between the split arrays and the FP64 staging
static void sp_convert(mpcx_co* e, bool to_fp64, bool near_only) {
  const int g = near_only ? 256 : grid_for(e->P.NT, 256, 2048);
  if (to_fp64) hipLaunchKernelGGL(sp_load, dim3(g), dim3(256), 0, e->st, e->P, e->D, e->S, near_only ? 1 : 0);
  else hipLaunchKernelGGL(sp_store, dim3(g), dim3(256), 0, e->st, e->P, e->D, e->S, near_only ? 1 : 0);
}

static void co_near_list(mpcx_co* e) {
  launch_bits_list(e->st, (const unsigned*)e->D.near_bits, e->nwords, e->D.partial, e->D.near_list, e->D.near_cap,
                   e->P.NT, &e->D.sc->near_num, &e->D.sc->err);
}

// before every launch of the global engine: barrier state and the
// workgroup bits of the sphere sums (the last step's bits stay set)
__global__ void ge_reset(GeSync* sy, unsigned* touch) {
  if (threadIdx.x == 0) { sy->count = 0u; sy->abort = 0; sy->gen = 0u; }
  for (int g = threadIdx.x; g < GE_GMAX; g += blockDim.x) sy->arrive[g] = 0u;
  for (int j = threadIdx.x; j < 2 * MAXSG * (GE_GMAX / 32); j += blockDim.x) touch[j] = 0u;
}

static hipError_t launch_near(mpcx_co* e, int nsteps, int mode) {
  if (e->near_mode == 2) {
    hipLaunchKernelGGL(ge_reset, dim3(1), dim3(256), 0, e->st, e->E.sy, e->E.touch);
    const CoP P = e->P;
    CoD G = e->Dg;
    {   // the particle arrays the system holds now (a sort swaps them with its buffers)
      const CoD& D = e->D;
      G.x = D.x; G.y = D.y; G.z = D.z; G.vx = D.vx; G.vy = D.vy; G.vz = D.vz; G.flag = D.flag;
      G.status = D.status; G.near_bits = D.near_bits; G.near_list = D.near_list; G.dim = D.dim; G.sc = D.sc;
    }
    const GeD E = e->E;
    void* args[] = {(void*)&P, (void*)&G, (void*)&E, (void*)&nsteps, (void*)&mode};
    hipError_t r = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(co_ge_engine), dim3(e->geG), dim3(GE_T),
                                              args, 0, e->st);
    if (r != hipSuccess) return r;
    return hipGetLastError();
  }
  if (e->clustered) {
    CoD& D = e->D;
    CluBase& B = e->cb;
    hipLaunchKernelGGL(clu_clear, dim3(1), dim3(256), 0, e->st, D.clu);
    hipLaunchKernelGGL(clu_links, dim3(grid_for(D.near_cap, 256, 1024)), dim3(256), 0, e->st, e->P, D, B.assign);
    hipLaunchKernelGGL(clu_build, dim3(1), dim3(256), 0, e->st, e->P, D, B);
    hipLaunchKernelGGL(clu_lists, dim3(e->nd), dim3(1024), 0, e->st, D, B);
    const CoP* Ps = B.Ps;
    const CoD* Ds = B.Ds;
    const CluInfo* ci = D.clu;
    void* args[] = {(void*)&Ps, (void*)&Ds, (void*)&ci, (void*)&nsteps, (void*)&mode};
    hipError_t r = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(co_near_engine_c), dim3(e->nd),
                                              dim3(NE_T), args, 0, e->st);
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(clu_finish, dim3(1), dim3(256), 0, e->st, e->P, D, B, mode);
    return hipGetLastError();
  }
  switch (e->nd) {
    case 1: hipLaunchKernelGGL(co_near_engine<1>, dim3(1), dim3(NE_T), 0, e->st, e->P, e->D, nsteps, mode); break;
    case 2: hipLaunchKernelGGL(co_near_engine<2>, dim3(1), dim3(NE_T), 0, e->st, e->P, e->D, nsteps, mode); break;
    case 3: hipLaunchKernelGGL(co_near_engine<3>, dim3(1), dim3(NE_T), 0, e->st, e->P, e->D, nsteps, mode); break;
    default: hipLaunchKernelGGL(co_near_engine<4>, dim3(1), dim3(NE_T), 0, e->st, e->P, e->D, nsteps, mode); break;
  }
  return hipGetLastError();
}

// ------------------------------------------------------ restart evaluation
// restart_mod.f95:429-495 after the lists are built: the pair force of every
// inside entry with each sphere whose list holds it (the forces_mod.f95:62-136
// arithmetic; an A particle inside a catalytic cutoff turns B), the
// per-sphere totals, then the dimer forces and the potential with the
// motor-motor and wall terms.  The run then starts from these forces instead
// of the zero forces of a fresh start (initial_mod.f95:348-353).

// pair force of one particle with sphere s (list id 2k+i, ii = i)
__device__ __forceinline__ void restart_pair(const CoP& P, const CoD& D, double x, double y, double z, int& fl,
                                             int s, int ii, double& fx, double& fy, double& fz, double& pot) {
  fx = fy = fz = pot = 0.0;
  double d1 = pbc_half(x - D.dim[DX * MAXSG + s], P.box[0]);
  double d2 = pbc_half(y - D.dim[DY * MAXSG + s], P.box[1]);
  double d3 = z - D.dim[DZ * MAXSG + s];
  double dRsq = d1 * d1 + d2 * d2 + d3 * d3;
  if (!(dRsq <= P.cut2[ii])) return;
  double E;
  if (fl == 0) {
    if (ii == 0) { fl = 1; E = P.Eb[0]; } else E = P.Ea[1];
  } else if (fl == 1) E = P.Eb[ii];
  else if (fl == 2) E = P.Es[ii];
  else return;
  double F, Ub;
  lj_terms(dRsq, P.R6[ii], P.R12[ii], P.r12R12[ii], P.r6R6[ii], F, Ub);
  double EF = E * F;
  fx = EF * d1; fy = EF * d2; fz = EF * d3; pot = E * Ub;
}

// one-workgroup engine: entry k holds membership bits b = 2k'+i (sphere k' + i*nd)
__global__ __launch_bounds__(256) void co_restart_pairs(CoP P, CoD D) {
  const int cap = D.in_cap, n = min(D.sc->in_num, cap), nd = P.nd;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int p = D.in_plist[k];
    const unsigned mb = D.in_memb[k];
    const double x = D.x[p], y = D.y[p], z = D.z[p];
    const int f0 = D.flag[p];
    int fl = f0;
    for (int b = 0; b < 2 * nd; b++) {
      double fx = 0.0, fy = 0.0, fz = 0.0, pot = 0.0;
      if ((mb >> b) & 1u) restart_pair(P, D, x, y, z, fl, (b >> 1) + (b & 1) * nd, b & 1, fx, fy, fz, pot);
      D.in_fp[(b * 3 + 0) * cap + k] = fx;
      D.in_fp[(b * 3 + 1) * cap + k] = fy;
      D.in_fp[(b * 3 + 2) * cap + k] = fz;
      D.in_pot[b * cap + k] = pot;
    }
    if (fl != f0) D.flag[p] = fl;
  }
}

// global engine: entry k holds up to four list ids (bytes, ascending, 0xff ends)
__global__ __launch_bounds__(256) void co_restart_pairs_ge(CoP P, CoD D, GeD E) {
  const int n = min(D.sc->in_num, E.in_cap), nd = P.nd;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int p = D.in_plist[k];
    const unsigned ids = E.in_memb4[k];
    const double x = D.x[p], y = D.y[p], z = D.z[p];
    const int f0 = D.flag[p];
    int fl = f0;
    for (int m = 0; m < 4; m++) {
      const unsigned b = (ids >> (8 * m)) & 0xffu;
      double fx = 0.0, fy = 0.0, fz = 0.0, pot = 0.0;
      if (b != 0xffu) restart_pair(P, D, x, y, z, fl, (int)(b >> 1) + (int)(b & 1u) * nd, (int)(b & 1u), fx, fy, fz, pot);
      E.in_fp4[(size_t)k * 12 + m * 3 + 0] = fx;
      E.in_fp4[(size_t)k * 12 + m * 3 + 1] = fy;
      E.in_fp4[(size_t)k * 12 + m * 3 + 2] = fz;
      E.in_pot4[(size_t)k * 4 + m] = pot;
    }
    if (fl != f0) D.flag[p] = fl;
  }
}

// per-sphere totals of the pair forces and potential (one workgroup per
// sphere, fixed summation order)
__global__ __launch_bounds__(256) void co_restart_sums(CoP P, CoD D, GeD E, int ge, double* out) {
  __shared__ double red[4][4];
  const int s = blockIdx.x, nd = P.nd, tid = threadIdx.x;
  const unsigned b = 2u * (unsigned)(s % nd) + (unsigned)(s / nd);
  const int cap = D.in_cap, n = min(D.sc->in_num, ge ? E.in_cap : cap);
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = tid; k < n; k += blockDim.x) {
    if (!ge) {
      if ((D.in_memb[k] >> b) & 1u) {
        for (int c = 0; c < 3; c++) a[c] += D.in_fp[(b * 3 + c) * cap + k];
        a[3] += D.in_pot[b * cap + k];
      }
    } else {
      const unsigned ids = E.in_memb4[k];
      for (int m = 0; m < 4; m++)
        if (((ids >> (8 * m)) & 0xffu) == b) {
          for (int c = 0; c < 3; c++) a[c] += E.in_fp4[(size_t)k * 12 + m * 3 + c];
          a[3] += E.in_pot4[(size_t)k * 4 + m];
        }
    }
  }
  for (int c = 0; c < 4; c++) a[c] = wave_sum(a[c]);
  if ((tid & 63) == 0)
    for (int c = 0; c < 4; c++) red[tid >> 6][c] = a[c];
  __syncthreads();
  if (tid < 4) out[4 * s + tid] = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
}

// dimer forces and potential from the totals (the force part of
// dimer_forces_and_rattle_v, velocities untouched)
__global__ __launch_bounds__(256) void co_restart_dimers(CoP P, CoD D, const double* sums) {
  __shared__ double tp[MAXSG];
  const int nd = P.nd, s = threadIdx.x;
  if (s < 2 * nd) {
    const int k = s % nd, i = s / nd;
    const double* dm = D.dim;
    double fw, wp;
    lj_mot_wall(P, dm[DZ * MAXSG + s], i, fw, wp);
    double fx = 0.0 - sums[4 * s + 0], fy = 0.0 - sums[4 * s + 1], fz = 0.0 - sums[4 * s + 2];
    double t = 0.0 + sums[4 * s + 3];
    for (int j = 0; j < nd; j++)
      for (int l = 0; l < 2; l++) {
        if (j == k) continue;
        const int b = j + l * nd;
        const int kind = (i == 0 && l == 0) ? 0 : (i == 1 && l == 1) ? 2 : 1;
        double mx, my, mz, mp;
        lj_mot_mot(P, dm[DX * MAXSG + s], dm[DY * MAXSG + s], dm[DZ * MAXSG + s], dm[DX * MAXSG + b],
                   dm[DY * MAXSG + b], dm[DZ * MAXSG + b], kind, mx, my, mz, mp);
        fx = fx + mx; fy = fy + my; fz = fz + mz;
        if (j < k) t = t + mp;
      }
    fz = fz + fw;
    t = t + wp;
    D.dim[DFX * MAXSG + s] = fx; D.dim[DFY * MAXSG + s] = fy; D.dim[DFZ * MAXSG + s] = fz;
    tp[s] = t;
  }
  __syncthreads();
  if (s == 0) {
    double T = 0.0;
    for (int k = 0; k < nd; k++)
      for (int i = 0; i < 2; i++) T = T + tp[k + i * nd];
    D.sc->tpotential = T;
  }
}

static hipError_t launch_restart(mpcx_co* e) {
  const CoP& P = e->P;
  hipStream_t st = e->st;
  const int ge = e->near_mode == 2 ? 1 : 0;
  CoD D = ge ? e->Dg : e->D;
  {   // particle arrays of the system as they are now
    const CoD& S = e->D;
    D.x = S.x; D.y = S.y; D.z = S.z; D.flag = S.flag; D.dim = S.dim; D.sc = S.sc;
  }
  const int cap = ge ? e->E.in_cap : D.in_cap;
  const int grid = std::max(1, std::min(1024, (cap + 255) / 256));
  if (ge)
    hipLaunchKernelGGL(co_restart_pairs_ge, dim3(grid), dim3(256), 0, st, P, D, e->E);
  else
    hipLaunchKernelGGL(co_restart_pairs, dim3(grid), dim3(256), 0, st, P, D);
  hipLaunchKernelGGL(co_restart_sums, dim3(2 * e->nd), dim3(256), 0, st, P, D, e->E, ge, e->d_rsum);
  hipLaunchKernelGGL(co_restart_dimers, dim3(1), dim3(256), 0, st, P, D, (const double*)e->d_rsum);
  return hipGetLastError();
}

static int co_setup(mpcx_co* e, int32_t do_sort, bool restart) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  CoD& D = e->D;
  const long long NT = e->P.NT;
  hipStream_t st = e->st;
  // sphere positions of this interval start
  std::vector<double> dim(12 * MAXSG);
  HIPCHK(hipMemcpy(dim.data(), D.dim, dim.size() * 8, hipMemcpyDeviceToHost));
  CoScalars sc;
  HIPCHK(hipMemcpy(&sc, D.sc, sizeof sc, hipMemcpyDeviceToHost));
  for (int s = 0; s < 2 * e->nd; s++) { sc.x0[s] = dim[DX * MAXSG + s]; sc.y0[s] = dim[DY * MAXSG + s]; sc.z0[s] = dim[DZ * MAXSG + s]; }
  sc.far_maxv2 = 0; sc.near_vmax2 = 0.0; sc.n_ev = 0; sc.err = 0; sc.next_fmax2 = 0;
  HIPCHK(hipMemcpy(D.sc, &sc, sizeof sc, hipMemcpyHostToDevice));
  e->handoff = e->list_due = false;
  hipLaunchKernelGGL(fill_kernel<unsigned>, dim3(grid_for(e->nwords + 1, 256)), dim3(256), 0, st, D.near_bits,
                     e->nwords + 1, 0u);
  if (do_sort) {
    int r = co_sort(e, st);
    if (r) return r;
  }
  hipLaunchKernelGGL(fill_kernel<unsigned>, dim3(grid_for(e->nwords + 1, 256)), dim3(256), 0, st, D.near_bits,
                     e->nwords + 1, 0u);
  hipLaunchKernelGGL(co_classify_all, dim3(grid_for(NT, 256, 2048)), dim3(256), 0, st, e->P, D);
  co_near_list(e);
  hipLaunchKernelGGL(fill_kernel<unsigned char>, dim3(grid_for(NT, 256)), dim3(256), 0, st, D.status, NT,
                     (unsigned char)0);
  HIPCHK(launch_near(e, 0, M_SETUP));
  if (restart) HIPCHK(launch_restart(e));
  if (e->split) sp_convert(e, false, false);
  HIPCHK(hipStreamSynchronize(st));
  return check_device_error(e);
}

extern "C" int mpcx_co_setup(mpcx_co* e, int32_t do_sort) { return co_setup(e, do_sort, false); }

extern "C" int mpcx_co_setup_restart(mpcx_co* e, int32_t do_sort) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  if (e->clustered)
    return set_error(MPCX_E_UNSUPPORTED, "the restart evaluation runs with near mode 0 or 2");
  return co_setup(e, do_sort, true);
}

// Workgroups of co_bulk / sp_bulk: the per-particle cost varies (far particles
// replay every streaming event, near ones none; cells near a sphere take square
// roots), so many short workgroups that the hardware hands out as others finish
// balance better than one resident round: up to 8192 (measured at 100^3:
// 0.50 ms per launch against 0.55 with 2048 and 0.65 with one resident round).
static int bulk_grid(long long n) { return grid_for(n, 256, 8192); }

extern "C" int mpcx_co_set_bulk_grid(mpcx_co* e, int32_t grid) {
  if (!e || grid < 0 || grid > 65536) return set_error(MPCX_E_ARG, "grid must be 0 (automatic) to 65536");
  e->g_bulk = e->g_spbulk = grid ? grid : bulk_grid(e->P.NT);
  return 0;
}

extern "C" int mpcx_co_set_overlap(mpcx_co* e, int32_t on) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  e->overlap = on != 0;
  return 0;
}

static int co_run_split(mpcx_co* e, int64_t ncoll);

extern "C" int mpcx_co_run(mpcx_co* e, int64_t ncoll, const double* draws) {
  if (!e || ncoll < 0) return set_error(MPCX_E_ARG, "bad arguments");
  if (e->split) return co_run_split(e, ncoll);   // the split-precision interval (one-system batch)
  {
    int cs = 0;
    HIPCHK(hipMemcpy(&cs, &e->D.sc->cstep, sizeof(int), hipMemcpyDeviceToHost));
    if (cs % e->P.MPCtstep != 0) return set_error(MPCX_E_ARG, "mpcx_co_run starts at a collision step");
  }
  if (e->P.rng_mode == 0 && !draws && ncoll > 0) return set_error(MPCX_E_ARG, "rng_mode 0 needs host draws");
  CoD& D = e->D;
  CoP& P = e->P;
  const long long NT = P.NT, mc = P.mpcnumcell, vol = P.vol;
  hipStream_t st = e->st;
  int cstep = 0;
  HIPCHK(hipMemcpy(&cstep, &D.sc->cstep, sizeof(int), hipMemcpyDeviceToHost));
  e->bulk_ms = 0; e->near_ms = 0;
  if (e->tev.empty()) {   // per-kernel timing events, made on the first run
    e->tev.assign((size_t)mpcx_co::TMAX * mpcx_co::NKT * 2, nullptr);
    for (auto& ev : e->tev) HIPCHK(hipEventCreate(&ev));
  }
  HIPCHK(hipEventRecord(e->e0, st));
  const int gN = grid_for(NT, 256, 2048);
  if (!e->g_bulk) e->g_bulk = e->g_spbulk = bulk_grid(NT);
  for (int64_t c = 0; c < ncoll; c++) {
    const int cnext = cstep + P.MPCtstep;
    const bool sort_due = (cnext % P.sort_step) == 0;
    // FP64 with the far flight beside the one-workgroup near engine: cells
    // collided in two passes (a re-sort interval keeps one pass after the sort)
    const bool two = e->overlap && e->near_mode == 0 && !sort_due;
    if (two) D.epoch = ++e->epoch;
    hipLaunchKernelGGL(interval_begin, dim3(grid_for(3 * vol, 256, 2048)), dim3(256), 0, st, P, D,
                       (unsigned long long)(cnext / P.MPCtstep), two ? D.next_bits : nullptr,
                       (long long)e->nwords + 1, e->handoff ? 1 : 0);
    e->handoff = false;
    if (P.rng_mode == 0)
      HIPCHK(hipMemcpyAsync(e->d_U, draws + c * e->ndraw, e->ndraw * 8, hipMemcpyHostToDevice, st));
    const long long slot = c / mpcx_co::TSTRIDE;
    const bool timed = c % mpcx_co::TSTRIDE == 0 && slot < mpcx_co::TMAX;
    const bool last = c == ncoll - 1;   // the group events (mpcx_co_timing) of the last interval
    auto tmark = [&](int k, int edge, hipStream_t s) -> hipError_t {
      if (!timed) return hipSuccess;
      e->tev_on[(size_t)slot * mpcx_co::NKT + k] = 1;
      return hipEventRecord(e->tev[((size_t)slot * mpcx_co::NKT + k) * 2 + edge], s);
    };
    if (timed)
      for (int k = 0; k < mpcx_co::NKT; k++) e->tev_on[(size_t)slot * mpcx_co::NKT + k] = 0;
    // The near list (the set classified by the last two-pass interval) comes
    // before the fork: the near engine must reach a CU before co_far's
    // workgroups fill them all, since it needs nearly a whole CU's LDS.
    if (e->list_due) {
      co_near_list(e);
      e->list_due = false;
    }
    // the far particles' flight and binning beside the near engine
    const bool ov = e->overlap && e->near_mode == 0;
    if (ov) {
      hipStream_t s2 = e->st2;
      HIPCHK(hipEventRecord(e->fork, st));
      HIPCHK(hipStreamWaitEvent(s2, e->fork, 0));
      if (!sort_due) HIPCHK(tmark(1, 0, s2));
      hipLaunchKernelGGL(co_far, dim3(e->g_bulk), dim3(256), 0, s2, P, D, sort_due ? 0 : 1, two ? 1 : 0);
      if (!sort_due) HIPCHK(tmark(1, 1, s2));
      if (two) {   // the late cells around the spheres; the first collision pass
        hipLaunchKernelGGL(co_zone_mark, dim3(64, 2 * e->nd), dim3(256), 0, s2, P, D);
        HIPCHK(hipEventRecord(e->far_done, s2));
        HIPCHK(tmark(2, 0, s2));
        hipLaunchKernelGGL(co_collide_far, dim3((unsigned)((mc + CT - 1) / CT)), dim3(CT), 0, s2, P, D);
        HIPCHK(tmark(2, 1, s2));
        HIPCHK(tmark(3, 0, s2));
        hipLaunchKernelGGL(co_rotate_far, dim3(gN), dim3(256), 0, s2, P, D);
        HIPCHK(tmark(3, 1, s2));
      }
      HIPCHK(hipEventRecord(e->join, s2));
    }
    if (last) HIPCHK(hipEventRecord(e->n0, st));
    HIPCHK(tmark(0, 0, st));
    HIPCHK(launch_near(e, P.MPCtstep, sort_due ? M_STOP : 0));
    HIPCHK(tmark(0, 1, st));
    if (last) HIPCHK(hipEventRecord(e->n1, st));
    if (two) {   // the second pass needs co_far's binning only, not the first pass
      if (last) HIPCHK(hipEventRecord(e->b0, st));
      HIPCHK(hipStreamWaitEvent(st, e->far_done, 0));
      hipLaunchKernelGGL(co_bin_rest, dim3(128), dim3(256), 0, st, P, D, 1, 1);
      hipLaunchKernelGGL(co_collide_late, dim3(512 + 1024), dim3(CT), 0, st, P, D, 512);   // + reactions
    }
    if (ov) HIPCHK(hipStreamWaitEvent(st, e->join, 0));
    auto bulk = [&](int stream, int bin) {
      hipLaunchKernelGGL(co_bulk, dim3(e->g_bulk), dim3(256), 0, st, P, D, stream, bin);
    };
    if (sort_due) {
      if (ov) hipLaunchKernelGGL(co_bin_rest, dim3(128), dim3(256), 0, st, P, D, 0, 0);   // the flight check
      else bulk(1, 0);
      int r = co_sort(e, st);
      if (r) return r;
      co_near_list(e);
      HIPCHK(launch_near(e, 1, M_RESUME));
      if (last) HIPCHK(hipEventRecord(e->b0, st));
      bulk(0, 1);
    } else if (two) {   // after both passes: the next interval's near set (its list and
                        // far max follow at the next interval's start)
      if (last) HIPCHK(hipEventRecord(e->b1, st));
      std::swap(D.near_bits, D.next_bits);
      e->list_due = e->handoff = true;
      cstep = cnext;
      continue;
    } else if (ov) {
      if (last) HIPCHK(hipEventRecord(e->b0, st));
      hipLaunchKernelGGL(co_bin_rest, dim3(128), dim3(256), 0, st, P, D, 1, 0);
    } else {
      if (last) HIPCHK(hipEventRecord(e->b0, st));
      HIPCHK(tmark(1, 0, st));
      bulk(1, 1);
      HIPCHK(tmark(1, 1, st));
    }
    // with two-pass intervals the timed collide and rotate are their first
    // passes; the re-sort intervals' single pass is left out of the means
    auto tmark1 = [&](int k, int edge) -> hipError_t {
      return (e->overlap && e->near_mode == 0) ? hipSuccess : tmark(k, edge, st);
    };
    launch_scan(st, D.cnt, mc, D.partial, D.off);
    HIPCHK(tmark1(4, 0));
    hipLaunchKernelGGL(co_scatter, dim3(gN), dim3(256), 0, st, P, D);
    HIPCHK(tmark1(4, 1));
    HIPCHK(tmark1(2, 0));
    hipLaunchKernelGGL(co_collide, dim3((unsigned)((mc + CT - 1) / CT)), dim3(CT), 0, st, P, D);
    HIPCHK(tmark1(2, 1));
    hipLaunchKernelGGL(rotate_begin, dim3(grid_for(e->nwords + 1, 256, 2048)), dim3(256), 0, st, D, e->nwords);
    HIPCHK(tmark1(3, 0));
    hipLaunchKernelGGL(co_rotate, dim3(gN), dim3(256), 0, st, P, D, ov && !sort_due ? 2 : 1);
    HIPCHK(tmark1(3, 1));
    if (last) HIPCHK(hipEventRecord(e->b1, st));
    co_near_list(e);
    cstep = cnext;
  }
  if (e->list_due) {   // leave the state complete for the other entry points
    co_near_list(e);
    e->list_due = false;
  }
  if (e->handoff) {
    hipLaunchKernelGGL(co_far_handoff, dim3(1), dim3(64), 0, st, D);
    e->handoff = false;
  }
  if (ncoll > 0) {   // kernel-group times of the last interval
    HIPCHK(hipEventSynchronize(e->b1));
    float nm = 0, bm = 0;
    (void)hipEventElapsedTime(&nm, e->n0, e->n1);
    (void)hipEventElapsedTime(&bm, e->b0, e->b1);
    e->near_ms = nm;
    e->bulk_ms = bm;
  }
  HIPCHK(hipEventRecord(e->e1, st));
  if (ncoll > 0) e->cells_valid = true;
  HIPCHK(hipEventSynchronize(e->e1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e->e0, e->e1);
  e->run_ms = ms;
  e->kt_n = std::min<long long>((ncoll + mpcx_co::TSTRIDE - 1) / mpcx_co::TSTRIDE, mpcx_co::TMAX);
  for (int k = 0; k < mpcx_co::NKT; k++) {
    double sum = 0;
    long long n = 0;
    for (long long c = 0; c < e->kt_n; c++) {
      if (!e->tev_on[(size_t)c * mpcx_co::NKT + k]) continue;
      float t = 0;
      (void)hipEventElapsedTime(&t, e->tev[((size_t)c * mpcx_co::NKT + k) * 2], e->tev[((size_t)c * mpcx_co::NKT + k) * 2 + 1]);
      sum += t;
      n++;
    }
    e->kt_ms[k] = n ? sum / (double)n : 0.0;
  }
  return check_device_error(e);
}

// ------------------------------------------------------------- initial solvent
// initial_mod.f95:122-345 on the device: species by index (A below Na0, then
// B below Na0+Nb0, then S), positions uniform in the box and redrawn while
// within the cutoff of any sphere, Gaussian velocities with variance T/m of
// the species (Box-Muller), then the mean velocity subtracted.  Draws come
// from Philox streams 8 (positions, one counter per attempt) and 9
// (velocities), keyed by the seed and the particle index.
namespace mpcx {
constexpr int INIT_TRIES = 100000;
__global__ __launch_bounds__(256) void init_solvent(CoP P, CoD D, long long na, long long nb, double sa, double sb,
                                                    double ss, unsigned long long seed, double* part) {
  __shared__ double sx[MAXSG], sy[MAXSG], sz[MAXSG];
  __shared__ double red[4][3];
  const int nd = P.nd;
  for (int s = threadIdx.x; s < 2 * nd; s += 256) {
    sx[s] = D.dim[DX * MAXSG + s]; sy[s] = D.dim[DY * MAXSG + s]; sz[s] = D.dim[DZ * MAXSG + s];
  }
  __syncthreads();
  const double twopi = 2.0 * (4.0 * atan(1.0));
  double acc[3] = {0.0, 0.0, 0.0};
  const long long per = (P.NT + gridDim.x - 1) / gridDim.x;
  const long long i0 = (long long)blockIdx.x * per, i1 = min((long long)P.NT, i0 + per);
  for (long long i = i0 + threadIdx.x; i < i1; i += 256) {
    double x = 0, y = 0, z = 0;
    bool ok = false;
    for (int a = 0; a < INIT_TRIES && !ok; a++) {
      double u, v, w, unused;
      philox_uniform2(seed, 8, 2ull * a, (uint64_t)i, u, v);
      philox_uniform2(seed, 8, 2ull * a + 1, (uint64_t)i, w, unused);
      x = P.box[0] * u; y = P.box[1] * v; z = P.box[2] * w;
      ok = true;
      for (int s = 0; s < 2 * nd && ok; s++) {
        const double a1 = pbc_half(x - sx[s], P.box[0]), a2 = pbc_half(y - sy[s], P.box[1]), a3 = z - sz[s];
        if (a1 * a1 + a2 * a2 + a3 * a3 < P.cut2[s >= nd]) ok = false;
      }
    }
    if (!ok) raise_error(&D.sc->err, ERR_INIT);
    const int f = i < na ? 0 : (i < na + nb ? 1 : 2);
    const double sd = f == 0 ? sa : (f == 1 ? sb : ss);
    double u1, u2, u3, u4;
    philox_uniform2(seed, 9, 0, (uint64_t)i, u1, u2);
    philox_uniform2(seed, 9, 1, (uint64_t)i, u3, u4);
    const double r1 = sqrt(-2.0 * log(1.0 - u1)), r2 = sqrt(-2.0 * log(1.0 - u3));
    const double vx = sd * (r1 * cos(twopi * u2)), vy = sd * (r1 * sin(twopi * u2)), vz = sd * (r2 * cos(twopi * u4));
    D.x[i] = x; D.y[i] = y; D.z[i] = z;
    D.vx[i] = vx; D.vy[i] = vy; D.vz[i] = vz;
    D.flag[i] = f;
    acc[0] += vx; acc[1] += vy; acc[2] += vz;
  }
  for (int c = 0; c < 3; c++) {
    const double r = wave_sum(acc[c]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = r;
  }
  __syncthreads();
  if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] =
      ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}
__global__ __launch_bounds__(256) void init_zero_momentum(CoP P, CoD D, const double* part, int nblk) {
  __shared__ double mean[3];
  if (threadIdx.x < 3) {
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s = s + part[b * 3 + threadIdx.x];
    mean[threadIdx.x] = s / (double)P.NT;
  }
  __syncthreads();
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < P.NT; i += (long long)gridDim.x * 256) {
    D.vx[i] = D.vx[i] - mean[0];
    D.vy[i] = D.vy[i] - mean[1];
    D.vz[i] = D.vz[i] - mean[2];
  }
}
}  // namespace mpcx

extern "C" int mpcx_co_init_solvent(mpcx_co* e, int64_t Na0, int64_t Nb0, double T, uint64_t seed) {
  if (!e || Na0 < 0 || Nb0 < 0 || Na0 + Nb0 > e->P.NT || !(T >= 0)) return set_error(MPCX_E_ARG, "bad arguments");
  const int nblk = 2048;
  double* part = nullptr;
  HIPCHK(hipMalloc(&part, (size_t)nblk * 3 * 8));
  hipLaunchKernelGGL(init_solvent, dim3(nblk), dim3(256), 0, e->st, e->P, e->D, (long long)Na0, (long long)Nb0,
                     sqrt(T / e->cfg.Ma), sqrt(T / e->cfg.Mb), sqrt(T / e->cfg.Ms), (unsigned long long)seed, part);
  hipLaunchKernelGGL(init_zero_momentum, dim3(1024), dim3(256), 0, e->st, e->P, e->D, (const double*)part, nblk);
  hipError_t r = hipStreamSynchronize(e->st);
  (void)hipFree(part);
  if (r != hipSuccess) return set_error(MPCX_E_HIP, hipGetErrorString(r));
  return check_device_error(e);
}

extern "C" int mpcx_co_get_state(mpcx_co* e, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                                 int32_t* flag, double* dimer12, int32_t* scalars) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  if (e->split) sp_convert(e, true, false);   // the split arrays hold the state
  HIPCHK(hipStreamSynchronize(e->st));
  const size_t NT = (size_t)e->P.NT;
  CoD& D = e->D;
  if (x) HIPCHK(hipMemcpy(x, D.x, NT * 8, hipMemcpyDeviceToHost));
  if (y) HIPCHK(hipMemcpy(y, D.y, NT * 8, hipMemcpyDeviceToHost));
  if (z) HIPCHK(hipMemcpy(z, D.z, NT * 8, hipMemcpyDeviceToHost));
  if (vx) HIPCHK(hipMemcpy(vx, D.vx, NT * 8, hipMemcpyDeviceToHost));
  if (vy) HIPCHK(hipMemcpy(vy, D.vy, NT * 8, hipMemcpyDeviceToHost));
  if (vz) HIPCHK(hipMemcpy(vz, D.vz, NT * 8, hipMemcpyDeviceToHost));
  if (flag) HIPCHK(hipMemcpy(flag, D.flag, NT * 4, hipMemcpyDeviceToHost));
  if (dimer12) {
    std::vector<double> dim(12 * MAXSG);
    HIPCHK(hipMemcpy(dim.data(), D.dim, dim.size() * 8, hipMemcpyDeviceToHost));
    const int ns = 2 * e->nd;
    for (int f = 0; f < 12; f++)
      for (int s = 0; s < ns; s++) dimer12[f * ns + s] = dim[f * MAXSG + s];
  }
  if (scalars) {
    CoScalars sc;
    HIPCHK(hipMemcpy(&sc, D.sc, sizeof sc, hipMemcpyDeviceToHost));
    scalars[0] = sc.cstep; scalars[1] = sc.time_count; scalars[2] = sc.time_penetrate; scalars[3] = sc.in_num;
  }
  return 0;
}

// ---------------------------------------------------------------- frames
// One frame of the system, written at a collision step and read back before
// mpcx_co_setup_restart; it carries what restart_mod.f95:362-406 reads from
// the last data.h5 frame (solvent positions, velocities and species, dimer
// positions and velocities).  Layout (host byte order, little-endian here):
//   0  char[8] "MPCXFRM1"   8 int32 NT   12 int32 numdims   16 int32 cstep
//  20  int32 0             24 double box[3]                 48 uint64 payload bytes
//  56  uint64 FNV-1a hash of the payload
//  64  x y z vx vy vz (NT doubles each), flag (NT int32),
//      xd yd zd vxd vyd vzd (2*numdims doubles each, sphere s = k + i*numdims)
// The file is written under a temporary name and renamed into place.
namespace {
struct FrameHeader {
  char magic[8];
  int32_t NT, nd, cstep, zero;
  double box[3];
  uint64_t bytes, hash;
};
static_assert(sizeof(FrameHeader) == 64, "frame header is 64 bytes");
uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = (const unsigned char*)p;
  for (size_t i = 0; i < n; i++) { h ^= c[i]; h *= 1099511628211ull; }
  return h;
}
}  // namespace

extern "C" int mpcx_co_save_frame(mpcx_co* e, const char* path) {
  if (!e || !path) return set_error(MPCX_E_ARG, "null argument");
  const size_t NT = (size_t)e->P.NT, ns = 2 * (size_t)e->nd;
  std::vector<double> a(6 * NT);
  std::vector<int32_t> flag(NT);
  std::vector<double> dim(12 * ns);
  int32_t sc[4];
  int r = mpcx_co_get_state(e, a.data(), a.data() + NT, a.data() + 2 * NT, a.data() + 3 * NT, a.data() + 4 * NT,
                            a.data() + 5 * NT, flag.data(), dim.data(), sc);
  if (r) return r;
  if (sc[0] % e->P.MPCtstep != 0) return set_error(MPCX_E_ARG, "a frame is written at a collision step");
  FrameHeader h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, "MPCXFRM1", 8);
  h.NT = (int32_t)NT; h.nd = e->nd; h.cstep = sc[0];
  for (int i = 0; i < 3; i++) h.box[i] = e->P.box[i];
  h.bytes = 6 * NT * 8 + NT * 4 + 6 * ns * 8;
  uint64_t hs = 1469598103934665603ull;
  hs = fnv1a(hs, a.data(), a.size() * 8);
  hs = fnv1a(hs, flag.data(), NT * 4);
  hs = fnv1a(hs, dim.data(), 6 * ns * 8);   // positions and velocities: fields 0-5
  h.hash = hs;
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return set_error(MPCX_E_ARG, "frame: cannot open for writing");
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(a.data(), 8, a.size(), f) == a.size() &&
            fwrite(flag.data(), 4, NT, f) == NT && fwrite(dim.data(), 8, 6 * ns, f) == 6 * ns;
  ok = (fclose(f) == 0) && ok;
  if (!ok || rename(tmp.c_str(), path) != 0) {
    remove(tmp.c_str());
    return set_error(MPCX_E_ARG, "frame: write failed");
  }
  return 0;
}

extern "C" int mpcx_co_load_frame(mpcx_co* e, const char* path, const int32_t* react_box, const int64_t* tour,
                                  int32_t* cstep) {
  if (!e || !path) return set_error(MPCX_E_ARG, "null argument");
  const size_t NT = (size_t)e->P.NT, ns = 2 * (size_t)e->nd;
  FILE* f = fopen(path, "rb");
  if (!f) return set_error(MPCX_E_ARG, "frame: cannot open");
  FrameHeader h;
  std::vector<double> a(6 * NT), dim(6 * ns);
  std::vector<int32_t> flag(NT);
  bool ok = fread(&h, sizeof h, 1, f) == 1;
  if (ok && (memcmp(h.magic, "MPCXFRM1", 8) != 0 || h.NT != (int32_t)NT || h.nd != e->nd ||
             h.bytes != 6 * NT * 8 + NT * 4 + 6 * ns * 8 || h.box[0] != e->P.box[0] || h.box[1] != e->P.box[1] ||
             h.box[2] != e->P.box[2])) {
    fclose(f);
    return set_error(MPCX_E_ARG, "frame: written for another system (particle count, dimers or box)");
  }
  ok = ok && fread(a.data(), 8, a.size(), f) == a.size() && fread(flag.data(), 4, NT, f) == NT &&
       fread(dim.data(), 8, 6 * ns, f) == 6 * ns;
  fclose(f);
  if (!ok) return set_error(MPCX_E_ARG, "frame: truncated");
  uint64_t hs = 1469598103934665603ull;
  hs = fnv1a(hs, a.data(), a.size() * 8);
  hs = fnv1a(hs, flag.data(), NT * 4);
  hs = fnv1a(hs, dim.data(), 6 * ns * 8);
  if (hs != h.hash) return set_error(MPCX_E_ARG, "frame: checksum mismatch");
  if (h.cstep % e->P.MPCtstep != 0) return set_error(MPCX_E_ARG, "frame: not at a collision step");
  const double* d = dim.data();
  int r = mpcx_co_set_state(e, h.cstep, a.data(), a.data() + NT, a.data() + 2 * NT, a.data() + 3 * NT,
                            a.data() + 4 * NT, a.data() + 5 * NT, flag.data(), d, d + ns, d + 2 * ns, d + 3 * ns,
                            d + 4 * ns, d + 5 * ns, nullptr, nullptr, nullptr, react_box, tour);
  if (r) return r;
  if (cstep) *cstep = h.cstep;
  return 0;
}

extern "C" int mpcx_co_get_inside(mpcx_co* e, int32_t* list, uint32_t* memb, double* fp, double* pot, int64_t cap,
                                  int64_t* n) {
  if (!e || !n) return set_error(MPCX_E_ARG, "null argument");
  if (e->near_mode != 0) return set_error(MPCX_E_UNSUPPORTED, "in this near mode the lists are read with mpcx_co_get_lists");
  HIPCHK(hipStreamSynchronize(e->st));
  CoScalars sc;
  HIPCHK(hipMemcpy(&sc, e->D.sc, sizeof sc, hipMemcpyDeviceToHost));
  *n = sc.in_num;
  if (cap < sc.in_num) return set_error(MPCX_E_ARG, "buffer smaller than the inside list");
  const size_t m = (size_t)sc.in_num, ic = (size_t)e->D.in_cap;
  if (list) HIPCHK(hipMemcpy(list, e->D.in_plist, m * 4, hipMemcpyDeviceToHost));
  if (memb) HIPCHK(hipMemcpy(memb, e->D.in_memb, m * 4, hipMemcpyDeviceToHost));
  if (fp)
    for (int j = 0; j < MAXM * 3; j++)
      HIPCHK(hipMemcpy(fp + j * cap, e->D.in_fp + j * ic, m * 8, hipMemcpyDeviceToHost));
  if (pot)
    for (int j = 0; j < MAXM; j++)
      HIPCHK(hipMemcpy(pot + j * cap, e->D.in_pot + j * ic, m * 8, hipMemcpyDeviceToHost));
  return 0;
}

// Neighbour lists in the reference's layout (neighbour_list_mod.f95:38-57):
// nb/fpx/fpy/fpz/pot as (numdims, 2, max_nb) column-major with 1-based
// particle indices in ascending order, nbnum as (numdims, 2); the inside list
// (0-based, ascending) in in_list.  Works with and without clusters.
extern "C" int mpcx_co_get_lists(mpcx_co* e, int64_t max_nb, int64_t* nb, int64_t* nbnum, double* fpx, double* fpy,
                                 double* fpz, double* pot, int32_t* in_list, int64_t in_cap, int64_t* in_num) {
  if (!e || !nbnum || !in_num || max_nb < 0) return set_error(MPCX_E_ARG, "bad arguments");
  HIPCHK(hipStreamSynchronize(e->st));
  const int nd = e->nd;
  std::vector<std::vector<int>> lp(2 * nd);
  std::vector<std::vector<double>> lf(2 * nd);   // fx fy fz pot per member
  std::vector<int> all;
  auto take = [&](const CoD& D, int n, const int* dims) -> int {
    const size_t cap = (size_t)D.in_cap;
    std::vector<int> pl(n);
    std::vector<unsigned> mb(n);
    std::vector<double> fp((size_t)MAXM * 3 * n), pt((size_t)MAXM * n);
    if (n) {
      HIPCHK(hipMemcpy(pl.data(), D.in_plist, (size_t)n * 4, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(mb.data(), D.in_memb, (size_t)n * 4, hipMemcpyDeviceToHost));
      for (int j = 0; j < MAXM * 3; j++)
        HIPCHK(hipMemcpy(fp.data() + (size_t)j * n, D.in_fp + j * cap, (size_t)n * 8, hipMemcpyDeviceToHost));
      for (int j = 0; j < MAXM; j++)
        HIPCHK(hipMemcpy(pt.data() + (size_t)j * n, D.in_pot + j * cap, (size_t)n * 8, hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < n; k++) {
      all.push_back(pl[k]);
      for (int b = 0; b < MAXM; b++) {
        if (!((mb[k] >> b) & 1u)) continue;
        const int kg = dims ? dims[b >> 1] : (b >> 1), i = b & 1;
        if (kg < 0 || kg >= nd) return set_error(MPCX_E_DEVICE, "list membership outside the system's dimers");
        const int L = kg + nd * i;
        lp[L].push_back(pl[k]);
        for (int c = 0; c < 3; c++) lf[L].push_back(fp[(size_t)(b * 3 + c) * n + k]);
        lf[L].push_back(pt[(size_t)b * n + k]);
      }
    }
    return 0;
  };
  if (e->near_mode == 2) {
    CoScalars sc;
    HIPCHK(hipMemcpy(&sc, e->D.sc, sizeof sc, hipMemcpyDeviceToHost));
    const int n = std::min(sc.in_num, e->E.in_cap);
    std::vector<int> pl(n);
    std::vector<unsigned> ids(n);
    std::vector<double> fp((size_t)n * 12), pt((size_t)n * 4);
    if (n) {
      HIPCHK(hipMemcpy(pl.data(), e->Dg.in_plist, (size_t)n * 4, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(ids.data(), e->E.in_memb4, (size_t)n * 4, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(fp.data(), e->E.in_fp4, (size_t)n * 96, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(pt.data(), e->E.in_pot4, (size_t)n * 32, hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < n; k++) {
      all.push_back(pl[k]);
      for (int m = 0; m < 4; m++) {
        const unsigned b = (ids[k] >> (8 * m)) & 0xffu;
        if (b == 0xffu) break;
        const int kg = (int)(b >> 1), i = (int)(b & 1u);
        if (kg >= nd) return set_error(MPCX_E_DEVICE, "list membership outside the system's dimers");
        const int L = kg + nd * i;
        lp[L].push_back(pl[k]);
        for (int c = 0; c < 3; c++) lf[L].push_back(fp[(size_t)k * 12 + m * 3 + c]);
        lf[L].push_back(pt[(size_t)k * 4 + m]);
      }
    }
  } else if (!e->clustered) {
    CoScalars sc;
    HIPCHK(hipMemcpy(&sc, e->D.sc, sizeof sc, hipMemcpyDeviceToHost));
    int rc = take(e->D, sc.in_num, nullptr);
    if (rc) return rc;
  } else {
    CluInfo ci;
    HIPCHK(hipMemcpy(&ci, e->D.clu, sizeof ci, hipMemcpyDeviceToHost));
    std::vector<CoD> Ds(ci.nclu);
    std::vector<CoScalars> scs(ci.nclu);
    HIPCHK(hipMemcpy(Ds.data(), e->cb.Ds, sizeof(CoD) * ci.nclu, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(scs.data(), e->cb.scs, sizeof(CoScalars) * ci.nclu, hipMemcpyDeviceToHost));
    for (int c = 0; c < ci.nclu; c++) {
      int rc = take(Ds[c], scs[c].in_num, ci.dims[c]);
      if (rc) return rc;
    }
  }
  std::sort(all.begin(), all.end());
  *in_num = (int64_t)all.size();
  if (in_list) {
    if ((int64_t)all.size() > in_cap) return set_error(MPCX_E_ARG, "buffer smaller than the inside list");
    for (size_t j = 0; j < all.size(); j++) in_list[j] = all[j];
  }
  for (int L = 0; L < 2 * nd; L++) {
    const int64_t cnt = (int64_t)lp[L].size();
    nbnum[L] = cnt;
    if (cnt > max_nb) return set_error(MPCX_E_DEVICE, "a neighbour list is longer than max_nb");
    for (int64_t j = 0; j < cnt; j++) {
      const size_t o = (size_t)L + (size_t)2 * nd * (size_t)j;
      if (nb) nb[o] = (int64_t)lp[L][j] + 1;
      if (fpx) fpx[o] = lf[L][4 * j + 0];
      if (fpy) fpy[o] = lf[L][4 * j + 1];
      if (fpz) fpz[o] = lf[L][4 * j + 2];
      if (pot) pot[o] = lf[L][4 * j + 3];
    }
  }
  return 0;
}

extern "C" int mpcx_co_get_cells(mpcx_co* e, double* vcm_uv, int32_t* cell_of, int32_t* census, double* tpotential) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  HIPCHK(hipStreamSynchronize(e->st));
  if (!e->cells_valid && (vcm_uv || cell_of || census))
    return set_error(MPCX_E_ARG, "no collision has run since the state was uploaded: the cell arrays are empty");
  CoD& D = e->D;
  const size_t mc = (size_t)e->P.mpcnumcell;
  if (vcm_uv && e->split) {   // FP32 records (vcm, count; axis) by tour position; returned by cell
    std::vector<float> rec(mc * 8);
    HIPCHK(hipMemcpy(rec.data(), D.srec, mc * 32, hipMemcpyDeviceToHost));
    for (size_t t = 0; t < mc; t++) {
      double* o = vcm_uv + (size_t)e->mpc_cell[t] * 8;
      const float* r = &rec[t * 8];
      o[0] = r[0]; o[1] = r[1]; o[2] = r[2]; o[3] = r[4]; o[4] = r[5]; o[5] = r[6]; o[6] = r[3]; o[7] = 0.0;
    }
  } else if (vcm_uv) {   // stored by tour position; returned by cell
    std::vector<double> rec(mc * 8);
    HIPCHK(hipMemcpy(rec.data(), D.cellrec, mc * 64, hipMemcpyDeviceToHost));
    for (size_t t = 0; t < mc; t++) memcpy(vcm_uv + (size_t)e->mpc_cell[t] * 8, &rec[t * 8], 64);
  }
  if (cell_of) {
    HIPCHK(hipMemcpy(cell_of, D.cell_of, (size_t)e->P.NT * 4, hipMemcpyDeviceToHost));
    for (long long i = 0; i < e->P.NT; i++) {
      if (cell_of[i] < 0 || (size_t)cell_of[i] >= mc) return set_error(MPCX_E_DEVICE, "cell index out of range");
      cell_of[i] = e->mpc_cell[cell_of[i]];   // both storages bin by tour position
    }
  }
  if (census && e->split) return set_error(MPCX_E_UNSUPPORTED, "split storage clears its census after each collision");
  if (census) HIPCHK(hipMemcpy(census, D.tcount, (size_t)e->P.vol * 12, hipMemcpyDeviceToHost));
  if (tpotential) HIPCHK(hipMemcpy(tpotential, &D.sc->tpotential, 8, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int mpcx_co_near_profile(mpcx_co* e, uint64_t* prof8) {
  if (!e || !prof8) return set_error(MPCX_E_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(e->st));
  HIPCHK(hipMemcpy(prof8, e->D.sc->prof, 10 * sizeof(uint64_t), hipMemcpyDeviceToHost));
  return 0;
}

namespace mpcx {
int co_view(mpcx_co* e, DevView* v) {
  if (!e || !v) return set_error(MPCX_E_ARG, "null argument");
  if (e->split) sp_convert(e, true, false);
  if (hipStreamSynchronize(e->st) != hipSuccess) return set_error(MPCX_E_HIP, "stream synchronize failed");
  memset(v, 0, sizeof *v);
  const CoD& D = e->D;
  v->x = D.x; v->y = D.y; v->z = D.z; v->vx = D.vx; v->vy = D.vy; v->vz = D.vz; v->flag = D.flag;
  v->n = e->P.NT;
  for (int i = 0; i < 3; i++) v->box[i] = e->P.box[i];
  v->mass[0] = e->cfg.Ma; v->mass[1] = e->cfg.Mb; v->mass[2] = e->cfg.Ms;
  v->species[0] = 0; v->species[1] = 1; v->species[2] = 2; v->species[3] = -1;
  v->periodic_z = 0;
  v->nd = e->nd;
  std::vector<double> dim(12 * MAXSG);
  if (hipMemcpy(dim.data(), D.dim, dim.size() * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(MPCX_E_HIP, "hipMemcpy failed");
  for (int f = 0; f < 12; f++)
    for (int s = 0; s < 2 * e->nd; s++) v->dim[f][s] = dim[f * MAXSG + s];
  v->msph[0] = e->cfg.Mc; v->msph[1] = e->cfg.Mn;
  if (hipMemcpy(&v->tpotential, &D.sc->tpotential, 8, hipMemcpyDeviceToHost) != hipSuccess)
    return set_error(MPCX_E_HIP, "hipMemcpy failed");
  v->stream = e->st;
  return 0;
}
}  // namespace mpcx

extern "C" int mpcx_co_kernel_ms(mpcx_co* e, double* ms5, int64_t* count) {
  if (!e || !ms5) return set_error(MPCX_E_ARG, "null argument");
  for (int k = 0; k < mpcx_co::NKT; k++) ms5[k] = e->kt_ms[k];
  if (count) *count = e->kt_n;
  return 0;
}

extern "C" int mpcx_co_timing(mpcx_co* e, double* run_ms, double* bulk_ms, double* near_ms) {
  if (!e) return set_error(MPCX_E_ARG, "null engine");
  if (run_ms) *run_ms = e->run_ms;
  if (bulk_ms) *bulk_ms = e->bulk_ms;
  if (near_ms) *near_ms = e->near_ms;
  return 0;
}

// =====================================================================================
// Batches: many independent collective systems ("replicas") advanced together.
// Each replica is an ordinary engine (its own device arrays, parameters and
// Philox key); the batch launches every kernel of an interval once for all of
// them: the near engine with one workgroup per replica, the particle and cell
// kernels over a table of tiles (replica, first, last).  A replica's results
// are bit-identical to running it alone.
// =====================================================================================
namespace mpcx {

// per replica: exclusive scan of the collision-cell counts (one workgroup each)
__global__ __launch_bounds__(1024) void scan_cells_b(const CoP* Ps, const CoD* Ds) {
  const CoD D = Ds[blockIdx.x];
  const long long n = Ps[blockIdx.x].mpcnumcell;
  __shared__ unsigned long long s[17];
  __shared__ long long run;
  if (threadIdx.x == 0) run = 0;
  __syncthreads();
  for (long long b0 = 0; b0 < n; b0 += 4096) {
    int v[4];
    unsigned long long sum = 0;
    for (int j = 0; j < 4; j++) {
      const long long i = b0 + threadIdx.x * 4 + j;
      v[j] = i < n ? D.cnt[i] : 0;
      sum += v[j];
    }
    unsigned long long tot;
    const unsigned long long e = block_excl_scan64(sum, s, tot);
    long long r = run + (long long)e;
    for (int j = 0; j < 4; j++) {
      const long long i = b0 + threadIdx.x * 4 + j;
      if (i < n) D.off[i] = (int)r;
      r += v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) run += (long long)tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) D.off[n] = (int)run;
}

// per replica: the ordered near list from its bitmap (one workgroup each)
__global__ __launch_bounds__(1024) void near_list_b(const CoP* Ps, const CoD* Ds) {
  const CoD D = Ds[blockIdx.x];
  const int NT = Ps[blockIdx.x].NT;
  const long long nwords = (NT + 31) / 32;
  __shared__ unsigned long long s[17];
  __shared__ long long run;
  if (threadIdx.x == 0) run = 0;
  __syncthreads();
  for (long long w0 = 0; w0 < nwords; w0 += 1024) {
    const long long w = w0 + threadIdx.x;
    unsigned word = w < nwords ? D.near_bits[w] : 0u;
    unsigned long long tot;
    const unsigned long long e = block_excl_scan64(__popc(word), s, tot);
    long long pos = run + (long long)e;
    while (word) {
      const int b = __ffs(word) - 1;
      word &= word - 1;
      const long long idx = w * 32 + b;
      if (idx < NT) {
        if (pos < D.near_cap) D.near_list[pos] = (int)idx;
        pos++;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) run += (long long)tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    D.sc->near_num = run < D.near_cap ? (int)run : D.near_cap;
    if (run > D.near_cap) raise_error(&D.sc->err, ERR_NEAR_CAP);
  }
}

// per replica (blockIdx.y): clear the collision-cell and census counters, next collision
__global__ __launch_bounds__(256) void fills_begin_b(const CoP* Ps, const CoD* Ds) {
  const CoD D = Ds[blockIdx.y];
  const CoP& P = Ps[blockIdx.y];
  const long long mc = P.mpcnumcell, vol = P.vol;
  const long long i0 = blockIdx.x * 256ll + threadIdx.x, st = (long long)gridDim.x * 256;
  for (long long i = i0; i < mc; i += st) D.cnt[i] = 0;
  for (long long i = i0; i < 3 * vol; i += st) D.tcount[i] = 0;
  for (long long i = i0; i < vol; i += st) D.firstB[i] = 0xffffffffu;
  if (blockIdx.x == 0 && threadIdx.x == 0) D.sc->coll += 1;
}
// per replica: clear the near bitmap and the far speed maximum before the rotation
__global__ __launch_bounds__(256) void fills_rotate_b(const CoP* Ps, const CoD* Ds) {
  const CoD D = Ds[blockIdx.y];
  const long long nw = (Ps[blockIdx.y].NT + 31) / 32 + 1;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < nw; i += (long long)gridDim.x * 256) D.near_bits[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) D.sc->far_maxv2 = 0ull;
}

// sort buffers of one replica (the engine's own)
struct CoSortD {
  const int* tourpos;
  int *scnt, *soff, *sb, *key;
  unsigned short* rk;
  double* alt[6];
  int* aflag;
  unsigned char* astat;
  unsigned* abits;
};

__global__ __launch_bounds__(256) void sort_fills_b(const CoP* Ps, const CoSortD* Ss) {
  const CoSortD S = Ss[blockIdx.y];
  const CoP& P = Ps[blockIdx.y];
  const long long nw = (P.NT + 31) / 32 + 1;
  const long long i0 = blockIdx.x * 256ll + threadIdx.x, st = (long long)gridDim.x * 256;
  for (long long i = i0; i < P.vol; i += st) S.scnt[i] = 0;
  for (long long i = i0; i < P.NT; i += st) S.sb[i] = 0;
  for (long long i = i0; i < nw; i += st) S.abits[i] = 0u;
}
__global__ __launch_bounds__(256) void sort_bin_b(const CoP* Ps, const CoD* Ds, const CoSortD* Ss, const Tile* tiles) {
  const Tile t = tiles[blockIdx.x];
  const CoP P = Ps[t.r];
  const CoD D = Ds[t.r];
  const CoSortD S = Ss[t.r];
  sort_bin_body(P, D, S.tourpos, S.scnt, S.key, S.rk, t.begin, t.end, 256);
}
// per replica: exclusive scan of the sort counts over the vol cells
__global__ __launch_bounds__(1024) void sort_scan_b(const CoP* Ps, const CoSortD* Ss) {
  const CoSortD S = Ss[blockIdx.x];
  const long long n = Ps[blockIdx.x].vol;
  __shared__ unsigned long long sh[17];
  __shared__ long long run;
  if (threadIdx.x == 0) run = 0;
  __syncthreads();
  for (long long b0 = 0; b0 < n; b0 += 4096) {
    int v[4];
    unsigned long long sum = 0;
    for (int j = 0; j < 4; j++) {
      const long long i = b0 + threadIdx.x * 4 + j;
      v[j] = i < n ? S.scnt[i] : 0;
      sum += v[j];
    }
    unsigned long long tot;
    const unsigned long long e = block_excl_scan64(sum, sh, tot);
    long long r = run + (long long)e;
    for (int j = 0; j < 4; j++) {
      const long long i = b0 + threadIdx.x * 4 + j;
      if (i < n) S.soff[i] = (int)r;
      r += v[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) run += (long long)tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) S.soff[n] = (int)run;
}
__global__ __launch_bounds__(256) void sort_scatter_b(const CoSortD* Ss, const Tile* tiles) {
  const Tile t = tiles[blockIdx.x];
  const CoSortD S = Ss[t.r];
  sort_scatter_body(S.key, S.rk, S.soff, S.sb, t.begin, t.end, 256);
}
// per tour cell of a tile (256 cells): its members in ascending original
// index.  The tile's members are one contiguous run of sb; a run of at most
// ORD_CAP entries (10 per cell on average: 2,560) is sorted in LDS, read and
// written back coalesced, a longer one in place in global memory.
constexpr int ORD_CAP = 6144;
__global__ __launch_bounds__(256) void sort_order_b(const CoSortD* Ss, const Tile* tiles) {
  __shared__ int s[ORD_CAP];
  const Tile t = tiles[blockIdx.x];
  const CoSortD S = Ss[t.r];
  const long long c = t.begin + threadIdx.x;
  const int o0 = S.soff[t.begin], o1 = S.soff[t.end];   // soff holds vol + 1 entries
  const int m = o1 - o0;
  const bool lds = m <= ORD_CAP;   // the same for the whole block
  int* seg = nullptr;
  int n = 0;
  if (c < t.end) n = min(S.scnt[c], BMAX);
  if (lds) {
    for (int j = threadIdx.x; j < m; j += 256) s[j] = S.sb[o0 + j];
    __syncthreads();
    if (c < t.end) seg = s + (S.soff[c] - o0);
  } else if (c < t.end) {
    seg = S.sb + S.soff[c];
  }
  for (int j = 1; j < n; j++) {
    int v = seg[j], k = j;
    while (k > 0 && seg[k - 1] > v) { seg[k] = seg[k - 1]; k--; }
    seg[k] = v;
  }
  if (lds) {
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += 256) S.sb[o0 + j] = s[j];
  }
}
__global__ __launch_bounds__(256) void sort_gather_b(const CoD* Ds, const CoSortD* Ss, const Tile* tiles) {
  const Tile t = tiles[blockIdx.x];
  const CoD D = Ds[t.r];
  const CoSortD S = Ss[t.r];
  sort_gather_body(D, S.sb, S.alt[0], S.alt[1], S.alt[2], S.alt[3], S.alt[4], S.alt[5], S.aflag, S.astat, S.abits,
                   t.begin, t.end, 256);
}


// =====================================================================================
// Split-precision intervals (one system, or a batch of systems)
//
// The far particles take one free flight of the interval and are binned on a
// second stream while the near engines run; the near set is binned when they
// are done; every collision cell is then collided and every particle rotated.
// The per-cell velocity sums are exact: each FP32 velocity is scaled by 2^40
// (an integer for |v| >= 2^-17; below that the rounding is under 2^-41) and
// added in 64-bit integer arithmetic, grouped per particle tile in an LDS hash
// table.  Cells are kept by tour position; after each re-sort every particle
// tile owns the tour cells whose first particle it holds: it stores their sums
// whole, and sends the sums of other tiles' cells to their owners' inboxes
// (sb_flush_owned); the cell pass of each owner adds them.  Integer sums do
// not depend on the order the members arrive in, so the far particles need
// not wait for the near set, no cell is collided twice, and a system's results
// do not depend on which workgroups ran it or on the particle order.  The
// census (unshifted cells, distribute_mod.f95:95-183) is grouped per tile the
// same way, and the bulk B->A reaction (reaction_mod.f95:28-66) runs per
// unshifted cell from it.  Philox streams: 0 the grid shift, 1 the axes
// (counter: collision cell), 2 the reactions (counter: unshifted cell), keyed
// by the system's seed.
// =====================================================================================
constexpr double SB_SCALE = 0x1p40;   // fixed point of the cell sums
// What is left to atomics -- the near set's sums (sb_near), cells no tile
// owns or whose owner's inbox is full, and the census -- is added into one
// copy of the accumulators per XCD, named by the hardware XCC_ID register, so
// no two L2s hold the same accumulator line; sb_cells_own adds the copies
// (integer sums: the result does not depend on how the work was spread) after
// the kernels that fill them have ended.  The atomics are workgroup-scoped,
// but the L2 does not execute them: the counters show every atomic request
// going to the memory fabric (TCC_EA0_ATOMIC equals TCC_ATOMIC), at a rate
// far below the L2's, hence the plain stores above for most sums.
__device__ __forceinline__ int hw_xcc() {
  unsigned v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return (int)(v & (unsigned)(NXCD - 1));
}
// Global-memory views of one system's arrays for the split kernels.  The
// descriptors are read from device memory, so the compiler cannot tell their
// pointers are global and would emit FLAT accesses, which also count against
// the LDS wait counter: every LDS wait of the binning would then wait for all
// outstanding memory operations.  Address-space 1 pointers give GLOBAL ones.
template <typename T>
using gp = __attribute__((address_space(1))) T*;
typedef float f4v __attribute__((ext_vector_type(4)));   // a cell record half (builtin vector: no operator= member)
template <typename T>
__device__ __forceinline__ gp<T> g1(T* p) { return (gp<T>)p; }
struct SbV {
  gp<int> cell;
  gp<float> ox, oy, oz, vx, vy, vz;
  gp<signed char> flag;
  gp<unsigned> near_bits;
  gp<unsigned char> status, dq;
  gp<int> cell_of;
  gp<unsigned long long> acc;   // this workgroup's XCD copy: [3 * mpcnumcell], by collision cell
  gp<int> tc;                   // B census: [vol]
  gp<unsigned> fb;              // first-B keys: [vol]
  gp<unsigned> copied;          // set when a cell sum goes to the copy (the cell pass then reads them)
};
__device__ __forceinline__ SbV sb_view(const CoP& P, const CoD& D, const SpD& S, int x) {
  SbV v;
  v.cell = g1(S.cell); v.ox = g1(S.ox); v.oy = g1(S.oy); v.oz = g1(S.oz);
  v.vx = g1(S.vx); v.vy = g1(S.vy); v.vz = g1(S.vz); v.flag = g1(S.flag);
  v.near_bits = g1(D.near_bits); v.status = g1(D.status); v.dq = g1(D.dq); v.cell_of = g1(D.cell_of);
  v.acc = g1(D.sacc + (long long)x * 3 * P.mpcnumcell);
  v.tc = g1(D.xtc + (long long)x * P.vol);
  v.fb = g1(D.xfb + (long long)x * P.vol);
  v.copied = g1(&D.sc->sums_copied);
  return v;
}
template <typename T>
__device__ __forceinline__ void l2_add(gp<T> p, T v) {
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void l2_min(gp<unsigned> p, unsigned v) {
  (void)__hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

constexpr int SB_NBITS = 12;          // count bits below the z sum (at most 4095 members)
__device__ __forceinline__ long long sb_fix(float v) { return __double2ll_rn((double)v * SB_SCALE); }

// Per-workgroup hash tables of the binning, filled over the workgroup's whole
// tile and added to the global sums once at its end (one atomic per word and
// distinct cell or census key of the tile): collision-cell sums (GT slots; the
// count in the low bits of the z word) and the B census (SB_CT slots; count
// and minimum key per unshifted cell).  A key that finds no free slot within
// SB_PROBE probes goes to the global sums directly.  The distinct collision
// cells of a tile grow as the particles drift from their sorted order (a
// 4096-particle tile of 10 per cell: about 580 just after a re-sort, about
// 1020 nine collisions later), so the tile size picks GT to keep the table
// at most about half full.  The census holds only the unshifted cells of B
// particles, about a tenth of a tile's particles' worth of keys: the particle
// tiles take CT = GT / 2 slots (with the 4-sphere table of a batch,
// sb_bulk<.., 1024> fits four workgroups in a CU's LDS, eight waves per SIMD).
constexpr int SB_CT = 1024;    // census slots of the near-set binning
constexpr int SB_PROBE = 64;
constexpr int SB_DEST = 64;    // owner tiles a tile sends foreign sums to (beyond: the XCD copies)
template <int GT, int CT>
struct SbTables {
  static_assert((GT & (GT - 1)) == 0 && (CT & (CT - 1)) == 0 && GT >= CT, "slots: powers of two, GT >= CT");
  static constexpr int SHIFT = 32 - __builtin_ctz(GT);
  int rk[GT];
  unsigned long long rs[3][GT];
  int ck[CT], cc[CT];
  unsigned cm[CT];
  int dk[SB_DEST], dn[SB_DEST], db[SB_DEST];   // owner tile, entries for it, first inbox slot claimed
};
template <int GT, int CT>
__device__ __forceinline__ void sb_tables_init(SbTables<GT, CT>& G) {
  for (int j = threadIdx.x; j < GT; j += blockDim.x) {
    G.rk[j] = -1; G.rs[0][j] = 0ull; G.rs[1][j] = 0ull; G.rs[2][j] = 0ull;
    if (j < CT) { G.ck[j] = -1; G.cc[j] = 0; G.cm[j] = 0xffffffffu; }
    if (j < SB_DEST) { G.dk[j] = -1; G.dn[j] = 0; }
  }
  __syncthreads();
}
// slot of key in keys[N] (inserted if absent), or -1 after SB_PROBE full probes
template <int N>
__device__ __forceinline__ int sb_slot(int* keys, int key) {
  int j = (int)(((unsigned)key * 2654435761u) >> (32 - __builtin_ctz(N)));
#pragma unroll 1
  for (int p = 0; p < SB_PROBE; p++) {
    const int old = atomicCAS(&keys[j], -1, key);
    if (old == -1 || old == key) return j;
    j = (j + 1) & (N - 1);
  }
  return -1;
}
// one lane's velocity into cell t (tour position; ok) and census entry ck with key (cen); no barrier
template <int GT, int CT>
__device__ __forceinline__ void sb_insert(const CoP& P, const SbV& A, SbTables<GT, CT>& G, bool ok, int t, float vx, float vy,
                                          float vz, bool cen, int ck, unsigned key) {
  if (ok) {
    const unsigned long long w0 = (unsigned long long)sb_fix(vx), w1 = (unsigned long long)sb_fix(vy);
    const unsigned long long w2 = ((unsigned long long)sb_fix(vz) << SB_NBITS) + 1ull;
    const int j = sb_slot<GT>(G.rk, t);
    if (j >= 0) {
      atomicAdd(&G.rs[0][j], w0);
      atomicAdd(&G.rs[1][j], w1);
      atomicAdd(&G.rs[2][j], w2);
    } else {
      const gp<unsigned long long> a = A.acc + 3ll * t;
      l2_add<unsigned long long>(a, w0);
      l2_add<unsigned long long>(a + 1, w1);
      l2_add<unsigned long long>(a + 2, w2);
      *A.copied = 1u;
    }
  }
  if (cen) {
    const int j = sb_slot<CT>(G.ck, ck);
    if (j >= 0) {
      atomicAdd(&G.cc[j], 1);
      atomicMin(&G.cm[j], key);
    } else {
      l2_add<int>(A.tc + ck, 1);
      l2_min(A.fb + ck, key);
    }
  }
}
// slot of key in keys[N] (the table complete), or -1
template <int N>
__device__ __forceinline__ int sb_find(const int* keys, int key) {
  int j = (int)(((unsigned)key * 2654435761u) >> (32 - __builtin_ctz(N)));
#pragma unroll 1
  for (int p = 0; p < SB_PROBE; p++) {
    const int k = keys[j];
    if (k == key) return j;
    if (k == -1) return -1;
    j = (j + 1) & (N - 1);
  }
  return -1;
}
// The sums a tile binned into tour cells that another tile owns go to that
// owner's inbox: one slot claimed per entry with a counter atomic, the entry
// written with plain stores; the owner's cell pass merges them (integer sums:
// the order of arrival does not matter).  Every atomic on device memory runs
// in the memory fabric here, not in L2, so three per cell were the binning's
// largest cost.
struct FxEntry {
  int p, pad;
  unsigned long long w[3];
};
struct SbInbox {
  const int* tb;   // [R] index of each system's first particle tile among all tiles
  int* cnt;        // [tiles] entries appended this interval
  FxEntry* box;    // [tiles][cap]
  int2* rng;       // [tiles] owned tour positions [x, y) (set after every re-sort)
  int T, cap;      // particles per tile, entries per inbox
};
// Cell ownership of a particle tile [b, e) of a system of NT particles: the
// tour positions [own_bound(b), own_bound(e)) whose first particle at the last
// re-sort (soff, non-decreasing) lies in the tile, the first tile's range from
// 0 and the last one's to vol.  Consecutive tiles get adjacent ranges, so each
// tour cell has one owner whatever soff holds.
__device__ __forceinline__ int own_bound(const int* soff, long long vol, long long i, long long NT) {
  if (i <= 0) return 0;
  if (i >= NT) return (int)vol;
  long long lo = 0, hi = vol;   // first p < vol with soff[p] >= i, else vol
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (soff[mid] < i) lo = mid + 1;
    else hi = mid;
  }
  return (int)lo;
}
// the owned tour positions of every particle tile of the list (after a
// re-sort, and at the start of a run)
__global__ __launch_bounds__(256) void sb_own_ranges(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds,
                                                     const Tile* __restrict__ tiles, int n, SbInbox X) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const Tile tl = tiles[i];
  if (tl.begin >= tl.end) return;
  const long long vol = Ps[tl.r].vol, NT = Ps[tl.r].NT;
  const int* soff = Ds[tl.r].soff;
  X.rng[X.tb[tl.r] + (int)(tl.begin / X.T)] = make_int2(own_bound(soff, vol, tl.begin, NT), own_bound(soff, vol, tl.end, NT));
}
// The extra z layer of collision cells (tour positions [vol, mpcnumcell))
// has no first particle to own it by: it is dealt to a system's particle
// tiles in equal runs of ext_run(P, T) positions, tile k the k-th run.
__device__ __forceinline__ long long ext_run(const CoP& P, int T) {
  const long long nt = (P.NT + T - 1) / T;
  return max((P.mpcnumcell - P.vol + nt - 1) / nt, 1ll);
}
// the owner tile (index among all tiles) of tour position t of system r
__device__ __forceinline__ int owner_tile(const CoP& P, const CoD& D, const SbInbox& X, int r, int t) {
  const int nt = (int)((P.NT + X.T - 1) / X.T);
  const long long o = t < P.vol ? D.soff[t] / X.T : (t - P.vol) / ext_run(P, X.T);
  return X.tb[r] + (int)min(o, (long long)nt - 1);
}
// every thread of the block together, after the tile: the sums of the cells
// the tile owns (tour positions [plo, phi), one contiguous run of sown, and
// its run [e0, e1) of the extra layer) stored whole -- zero for an owned cell
// it binned nothing into -- and the other table entries sent to their
// owners' inboxes (the near-set tiles of sb_near own nothing)
template <int GT, int CT>
__device__ __forceinline__ void sb_flush_owned(const CoP& P, const CoD& D, const SbV& A, SbTables<GT, CT>& G, int plo,
                                               int phi, int e0, int e1, const SbInbox& X, int r) {
  __syncthreads();
  const gp<unsigned long long> own = g1(D.sown);
  const int n0 = phi - plo, nown = n0 + (e1 - e0);
  for (int i = (int)threadIdx.x; i < nown; i += blockDim.x) {
    const int p = i < n0 ? plo + i : e0 + (i - n0);
    const int j = sb_find<GT>(G.rk, p);
    unsigned long long w0 = 0ull, w1 = 0ull, w2 = 0ull;
    if (j >= 0) { w0 = G.rs[0][j]; w1 = G.rs[1][j]; w2 = G.rs[2][j]; }
    own[3ll * p] = w0; own[3ll * p + 1] = w1; own[3ll * p + 2] = w2;
  }
  auto owned = [&](int t) { return (t >= plo && t < phi) || (t >= e0 && t < e1); };
  // the foreign entries grouped by owner tile: ranks within each group in
  // LDS, one inbox claim per group, then the entries written at their slots
  constexpr int K = GT / 256;   // table slots per thread (at least 256 threads)
  int ds[K], rk[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int j = threadIdx.x + k * blockDim.x;
    ds[k] = -1; rk[k] = 0;
    if (j < GT) {
      const int t = G.rk[j];
      if (t >= 0 && !owned(t)) {
        const int d = sb_slot<SB_DEST>(G.dk, owner_tile(P, D, X, r, t));
        if (d >= 0) { ds[k] = d; rk[k] = atomicAdd(&G.dn[d], 1); }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < SB_DEST && G.dk[threadIdx.x] >= 0)
    G.db[threadIdx.x] = atomicAdd(&X.cnt[G.dk[threadIdx.x]], G.dn[threadIdx.x]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int j = threadIdx.x + k * blockDim.x;
    if (j >= GT) continue;
    const int t = G.rk[j];
    if (t >= 0 && !owned(t)) {
      bool sent = false;
      if (ds[k] >= 0) {
        const int idx = G.db[ds[k]] + rk[k];
        if (idx < X.cap) {
          FxEntry* e = X.box + (long long)G.dk[ds[k]] * X.cap + idx;
          e->p = t;
          e->w[0] = G.rs[0][j]; e->w[1] = G.rs[1][j]; e->w[2] = G.rs[2][j];
          sent = true;
        }
      }
      if (!sent) {   // a full inbox or too many owners: the XCD copy
        const gp<unsigned long long> a = A.acc + 3ll * t;
        l2_add<unsigned long long>(a, G.rs[0][j]);
        l2_add<unsigned long long>(a + 1, G.rs[1][j]);
        l2_add<unsigned long long>(a + 2, G.rs[2][j]);
        *A.copied = 1u;
      }
    }
    const int c = j < CT ? G.ck[j] : -1;
    if (c >= 0) {
      l2_add<int>(A.tc + c, G.cc[j]);   // checked against CENSUS_MAX by sb_cells_own
      l2_min(A.fb + c, G.cm[j]);
    }
  }
}

// The spheres of the distance bytes: each coordinate as its cell (integer)
// and the offset in it (FP32)
template <int NS>
struct SbSpheres {   // NS: sphere capacity (4 in a batch of one- or two-dimer replicas)
  int ic[3][NS];
  float fr[3][NS];
};
__device__ __forceinline__ float sp_pbc(float d, float L) {
  return d > 0.5f * L ? d - L : (d < -0.5f * L ? d + L : d);
}

// Binning of one split-storage particle per lane (every thread of the block
// together, `valid` false for idle lanes): its shifted cell (mpc_rotation_mod.f95:
// 116-145 in cell coordinates), its velocity into that cell's sums, its census
// entry (outside-list particles keyed by index, inside ones after them, inside
// ones within an LJ cutoff left out), and the distance byte to the spheres sph
// (the selection of the next near set).
template <int GT, int CT, int NS>
__device__ __forceinline__ void sb_bin(const CoP& P, const CoD& D, const SbV& A, SbTables<GT, CT>& G, int i, bool valid, bool inside,
                                       bool excl, int ix, int iy, int iz, float ox, float oy, float oz, float vx,
                                       float vy, float vz, int f, float fxs, float fys, float fzs, double zs,
                                       const SbSpheres<NS>& sph) {
  bool ok = valid;
  int t = 0;
  if (valid) {
    const int cx = sp_wrap(ix + (int)floorf(ox + fxs), P.Lx);
    const int cy = sp_wrap(iy + (int)floorf(oy + fys), P.Ly);
    const int cz = iz + (int)floorf(oz + fzs) + (zs < 0.0 ? 1 : 0);
    const long long c = (long long)cx + (long long)P.Lx * cy + (long long)P.Lx * P.Ly * cz;
    if (cx < 0 || cx >= P.Lx || cy < 0 || cy >= P.Ly || cz < 0 || c >= P.mpcnumcell) {
      raise_error_at(D.sc, ERR_CELL_RANGE, SITE_BIN);
      ok = false;
    } else {
      t = g1(D.mpc_pos)[c];   // cells are kept by tour position: a tile's cells lie in a few runs
    }
  }
  unsigned key = (unsigned)i;
  bool cen = false;
  int ck = 0;
  if (ok && f == 1 && !(inside && excl)) {   // the B census (all the bulk reaction reads)
    if (inside) key = (unsigned)P.NT + (unsigned)i;
    const long long u = (long long)ix + (long long)P.Lx * iy + (long long)P.Lx * P.Ly * iz;
    if (u < 0 || u >= P.vol) raise_error_at(D.sc, ERR_CELL_RANGE, SITE_BIN_CENSUS);
    else { cen = true; ck = g1(D.mpc_pos)[u]; }   // by tour position, as the cell sums
  }
  sb_insert(P, A, G, ok, t, vx, vy, vz, cen, ck, key);
  if (valid) {
    A.cell_of[i] = t;
    // distance beyond each list radius in FP32, from differences of cells
    // (exact integers) and of offsets, so no large coordinate is rounded; the
    // rounding (under 1e-5 at these distances) is covered by the 1e-4 taken off
    float dmin = 1e30f;
    bool close = false;
    auto d2_of = [&](int s) {
      const float a = sp_pbc((float)(ix - sph.ic[0][s]) + (ox - sph.fr[0][s]), (float)P.Lx);
      const float b = sp_pbc((float)(iy - sph.ic[1][s]) + (oy - sph.fr[1][s]), (float)P.Ly);
      const float c = (float)(iz - sph.ic[2][s]) + (oz - sph.fr[2][s]);
      return (a * a + b * b) + c * c;
    };
    for (int s = 0; s < 2 * P.nd; s++) {
      const float far = (float)P.full[s >= P.nd] + 8.0f;
      close |= d2_of(s) < far * far;
    }
    if (__ballot(close))   // most waves lie far from every sphere: no square roots at all
      for (int s = 0; s < 2 * P.nd; s++) {
        const float far = (float)P.full[s >= P.nd] + 8.0f, d2 = d2_of(s);
        if (d2 < far * far) dmin = fminf(dmin, sqrtf(d2) - (float)P.full[s >= P.nd]);
      }
    float qd = floorf(fmaxf(dmin - 1e-4f, 0.0f) * 32.0f);
    if (inside) qd = 0.0f;   // list members stay near until the next rebuild
    A.dq[i] = (unsigned char)fminf(qd, 255.0f);
  }
}

template <int NS>
__device__ __forceinline__ void sb_load_spheres(const CoP& P, const double* x, const double* y, const double* z,
                                                SbSpheres<NS>& sph) {
  for (int j = threadIdx.x; j < 2 * P.nd; j += blockDim.x) {
    const double c[3] = {x[j], y[j], z[j]};
    for (int a = 0; a < 3; a++) {
      const double f = floor(c[a]);
      sph.ic[a][j] = (int)f;
      sph.fr[a][j] = (float)(c[a] - f);
    }
  }
}

// Tile of particles of one system.  ALL = false (beside the near engine): the
// particles outside the interval's near set take one free flight of MPCtstep
// steps (sp_stream: bounce-back at the z walls, the reactive patch of z = 0)
// and, with do_bin, are binned with distances to the spheres' start
// positions.  ALL = true (a re-sort interval, after the sort and the near
// engine): every particle binned where it is, distances to the spheres' end
// positions.
// A split particle's state, loaded one round ahead of its use
struct SbPart {
  int cell, f;
  float ox, oy, oz, vx, vy, vz;
  unsigned char st;
  bool near;
};
__device__ __forceinline__ void sb_fetch(const SbV& V, long long i, long long end, SbPart& p) {
  p.near = false; p.cell = 0; p.f = -1; p.st = 0;
  p.ox = p.oy = p.oz = p.vx = p.vy = p.vz = 0.0f;
  if (i < end) {
    p.near = (V.near_bits[i >> 5] >> (i & 31)) & 1u;
    p.st = V.status[i];
    p.cell = V.cell[i];
    p.ox = V.ox[i]; p.oy = V.oy[i]; p.oz = V.oz[i];
    p.vx = V.vx[i]; p.vy = V.vy[i]; p.vz = V.vz[i];
    p.f = V.flag[i];
  }
}

// Tile of particles of one system (an empty tile, begin == end, pads the
// placement of tiles on XCDs).  ALL = false (beside the near engine): the
// particles outside the interval's near set take one free flight of MPCtstep
// steps (sp_stream: bounce-back at the z walls, the reactive patch of z = 0)
// and, with do_bin, are binned with distances to the spheres' start
// positions.  ALL = true (a re-sort interval, after the sort and the near
// engine): every particle binned where it is, distances to the spheres' end
// positions.  The next round's loads are issued before a round is binned,
// but gfx9 counts loads and stores in one counter and the round's stores make
// the compiler wait for all of them there: the latency is hidden by
// occupancy (NS = 4 spheres and GT / 2 census slots: four 512-thread
// workgroups per CU at GT = 1024), not by the prefetch.
template <bool ALL, int T, int GT, int NS>
__global__ __launch_bounds__(T, (NS == 4 && GT == 1024) ? 8 : 1) void sb_bulk(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds,
                                             const SpD* __restrict__ Ss, const Tile* __restrict__ tiles, int do_bin,
                                             SbInbox X) {
  __shared__ SbTables<GT, GT / 2> G;
  __shared__ SbSpheres<NS> sph;
  const Tile tl = tiles[blockIdx.x];
  if (tl.begin >= tl.end) return;
  const CoP P = Ps[tl.r];
  const CoD D = Ds[tl.r];
  const SpD S = Ss[tl.r];
  const int xcc = hw_xcc();
  const SbV V = sb_view(P, D, S, xcc);
  if (do_bin && threadIdx.x == 0) atomicOr(&D.sc->xmask, 1u << xcc);
  if (ALL) sb_load_spheres(P, D.sc->x0, D.sc->y0, D.sc->z0, sph);
  else sb_load_spheres(P, D.sc->xs0, D.sc->ys0, D.sc->zs0, sph);
  if (do_bin) sb_tables_init(G);
  else __syncthreads();
  double xs = 0, ys = 0, zs = 0;
  if (do_bin) collision_shift(P, D, xs, ys, zs);
  const float fxs = (float)xs, fys = (float)ys, fzs = (float)zs;
  const float dtc = (float)P.MDt * (float)P.MPCtstep;
  SbPart nxt;
  sb_fetch(V, tl.begin + threadIdx.x, tl.end, nxt);
  for (long long base = tl.begin; base < tl.end; base += T) {
    const SbPart cur = nxt;
    sb_fetch(V, base + T + threadIdx.x, tl.end, nxt);
    const int i = (int)(base + threadIdx.x);
    const bool valid = i < tl.end;
    bool mine = valid && (ALL || !cur.near);
    int ix = 0, iy = 0, iz = 0, f = cur.f;
    float ox = cur.ox, oy = cur.oy, oz = cur.oz, vx = cur.vx, vy = cur.vy, vz = cur.vz;
    bool inside = false, excl = false;
    if (mine) {
      if (cur.near) { inside = (cur.st & ST_INSIDE) != 0; excl = (cur.st & ST_EXCL) != 0; }
      else if (cur.st) V.status[i] = 0;
      sp_cell_xyz(P, cur.cell, ix, iy, iz);
      if (!ALL) {
        const int f0 = f;
        const bool bounced = sp_stream(ix, iy, iz, ox, oy, oz, vx, vy, vz, f, dtc, P, D.react_box, D.sc);
        if (sp_sane(P, ix, iy, iz, ox, oy, oz, vx, vy, vz)) {
          V.cell[i] = sp_pack(ix, iy, iz);
          V.ox[i] = ox; V.oy[i] = oy; V.oz[i] = oz;
          if (bounced) { V.vx[i] = vx; V.vy[i] = vy; V.vz[i] = vz; }
          if (f != f0) V.flag[i] = (signed char)f;
        } else {
          raise_error_at(D.sc, ERR_CELL_RANGE, SITE_STREAM);
          mine = false;
        }
      } else if (!sp_sane(P, ix, iy, iz, ox, oy, oz, vx, vy, vz)) {
        raise_error_at(D.sc, ERR_CELL_RANGE, SITE_BULK_CELL);
        mine = false;
      }
    }
    if (do_bin)
      sb_bin(P, D, V, G, i, mine, inside, excl, ix, iy, iz, ox, oy, oz, vx, vy, vz, f, fxs, fys, fzs, zs, sph);
  }
  if (do_bin) {
    const int2 own = X.rng[X.tb[tl.r] + (int)(tl.begin / X.T)];
    const long long ex = ext_run(P, X.T), k = tl.begin / X.T;
    const int e0 = (int)min(P.vol + k * ex, P.mpcnumcell), e1 = (int)min(P.vol + (k + 1) * ex, P.mpcnumcell);
    sb_flush_owned(P, D, V, G, own.x, own.y, e0, e1, X, tl.r);
  }
}

// Per system (blockIdx.y): the interval's start -- the collision counter and
// the spheres' start positions (start = 1) -- and, with load, the near set
// from split storage into the FP64 arrays (for a solo engine's near engine;
// a batch's near engines gather from split storage themselves).
__global__ __launch_bounds__(256) void sb_load_near(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds,
                                                    const SpD* __restrict__ Ss, int start, int load) {
  const int r = blockIdx.y;
  const CoP P = Ps[r];
  const CoD D = Ds[r];
  const SpD S = Ss[r];
  if (start && blockIdx.x == 0) {
    if (threadIdx.x == 0) D.sc->coll += 1;
    for (int j = threadIdx.x; j < 2 * P.nd; j += blockDim.x) {
      D.sc->xs0[j] = D.sc->x0[j]; D.sc->ys0[j] = D.sc->y0[j]; D.sc->zs0[j] = D.sc->z0[j];
    }
  }
  const long long nn = load ? min(D.sc->near_num, D.near_cap) : 0;
  for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < nn; q += (long long)gridDim.x * 256) {
    const int i = D.near_list[q];
    int ix, iy, iz;
    sp_cell_xyz(P, S.cell[i], ix, iy, iz);
    D.x[i] = (double)ix + (double)S.ox[i]; D.y[i] = (double)iy + (double)S.oy[i]; D.z[i] = (double)iz + (double)S.oz[i];
    D.vx[i] = (double)S.vx[i]; D.vy[i] = (double)S.vy[i]; D.vz[i] = (double)S.vz[i];
    D.flag[i] = (int)S.flag[i];
  }
}

// After the near engines: near-set tile (replica r, chunk begin of end
// chunks; begin == end pads the placement on XCDs) of the near list back into
// split storage and, with do_bin, binned (its sums to the cells' owner tiles;
// distances to the spheres' start
// positions, as sb_bulk<false>); with check, the flight check (the near
// engine's rebuild events must add up to the interval sb_bulk flew).
// NS = 4 (a batch of one- or two-dimer replicas) takes CT = SB_CT / 2 census
// slots: 35.7 KB of LDS, four workgroups per CU, so a 256-replica batch's
// 1024 near-set tiles run at once.
template <int NS>
__global__ __launch_bounds__(256) void sb_near(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds,
                                               const SpD* __restrict__ Ss, const Tile* __restrict__ tiles, int do_bin,
                                               int check, SbInbox X) {
  __shared__ SbTables<SB_CT, NS == 4 ? SB_CT / 2 : SB_CT> G;
  __shared__ SbSpheres<NS> sph;
  const Tile tl = tiles[blockIdx.x];
  if (tl.begin >= tl.end) return;
  const CoP P = Ps[tl.r];
  const CoD D = Ds[tl.r];
  const SpD S = Ss[tl.r];
  const int xcc = hw_xcc();
  const SbV V = sb_view(P, D, S, xcc);
  if (do_bin && threadIdx.x == 0) atomicOr(&D.sc->xmask, 1u << xcc);
  sb_load_spheres(P, D.sc->xs0, D.sc->ys0, D.sc->zs0, sph);
  if (do_bin) sb_tables_init(G);
  else __syncthreads();
  double xs = 0, ys = 0, zs = 0;
  if (do_bin) collision_shift(P, D, xs, ys, zs);
  const float fxs = (float)xs, fys = (float)ys, fzs = (float)zs;
  const long long nn = min(D.sc->near_num, D.near_cap);
  for (long long base = tl.begin * 256; base < nn; base += tl.end * 256) {
    const long long q = base + threadIdx.x;
    bool valid = q < nn;
    int i = 0, ix = 0, iy = 0, iz = 0, f = -1;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, vx = 0.0f, vy = 0.0f, vz = 0.0f;
    bool inside = false, excl = false;
    if (valid) {
      i = D.near_list[q];
      const double x = D.x[i], y = D.y[i], z = D.z[i], dvx = D.vx[i], dvy = D.vy[i], dvz = D.vz[i];
      if (!fp64_sane(P, x, y, z, dvx, dvy, dvz)) {   // reported; the particle keeps its split state
        raise_error_at(D.sc, ERR_CELL_RANGE, SITE_NEAR_STORE, i);
        valid = false;
      }
    }
    if (valid) {
      int c;
      sp_split(D.x[i], D.y[i], D.z[i], P, c, ox, oy, oz);
      vx = (float)D.vx[i]; vy = (float)D.vy[i]; vz = (float)D.vz[i];
      f = D.flag[i];
      S.cell[i] = c; S.ox[i] = ox; S.oy[i] = oy; S.oz[i] = oz;
      S.vx[i] = vx; S.vy[i] = vy; S.vz[i] = vz; S.flag[i] = (signed char)f;
      sp_cell_xyz(P, c, ix, iy, iz);
      const unsigned char st = D.status[i];
      inside = (st & ST_INSIDE) != 0; excl = (st & ST_EXCL) != 0;
    }
    if (do_bin)
      sb_bin(P, D, V, G, i, valid, inside, excl, ix, iy, iz, ox, oy, oz, vx, vy, vz, f, fxs, fys, fzs, zs, sph);
  }
  if (do_bin) sb_flush_owned(P, D, V, G, 0, 0, 0, 0, X, tl.r);   // owns nothing: every sum to its owner
  if (check && tl.begin == 0 && threadIdx.x == 0) check_flight(P, D);
}

// The cells of one system's particle tile, once every particle is binned:
// those it owns -- tour positions [plo, phi) and its run of the extra z
// layer -- OWNCAP at a time, each with its inbox added in LDS.  Per cell:
// the record of the collision cell (mean velocity from the exact sums of its
// owner, of the owner's inbox and, when some sums fell back to them, of the
// XCD copies; random axis of mpc_rotation_mod.f95:184-194); and, for an
// unshifted cell, the bulk B->A reaction (reaction_mod.f95:28-66: p from the
// host table of 1 - exp(-k1 nB tau), one draw per cell, the first B member in
// census order turned A) with its B census cleared.  Split storage keeps the
// census of B only: nothing else reads the A and S counts.
constexpr int OWNCAP = 1024;
__device__ __forceinline__ void sb_cell(const CoP& P, const CoD& D, const SpD& S, long long p, unsigned long long x0,
                                        unsigned long long x1, unsigned long long x2, bool copies);
__global__ __launch_bounds__(256) void sb_cells_own(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds,
                                                    const SpD* __restrict__ Ss, const Tile* __restrict__ tiles,
                                                    SbInbox X) {
  __shared__ unsigned long long acc[3][OWNCAP];
  const Tile tl = tiles[blockIdx.x];
  if (tl.begin >= tl.end) return;
  const CoP P = Ps[tl.r];
  const CoD D = Ds[tl.r];
  const SpD S = Ss[tl.r];
  const int g = X.tb[tl.r] + (int)(tl.begin / X.T);
  const int2 own = X.rng[g];
  const long long ex = ext_run(P, X.T), k = tl.begin / X.T;
  const long long e0 = min(P.vol + k * ex, P.mpcnumcell), e1 = min(P.vol + (k + 1) * ex, P.mpcnumcell);
  const long long lo[2] = {own.x, e0}, hi[2] = {own.y, e1};
  const int n = min(X.cnt[g], X.cap);
  const FxEntry* box = X.box + (long long)g * X.cap;
  const bool copies = D.sc->sums_copied != 0u;
  if (own.x == 0 && threadIdx.x == 0) D.sc->far_maxv2 = 0ull;   // sb_rotate publishes the next interval's far maximum
  for (int e = threadIdx.x; e < n; e += 256) {   // every entry is for a cell the tile owns
    const long long q = box[e].p;
    if (!((q >= lo[0] && q < hi[0]) || (q >= lo[1] && q < hi[1]))) raise_error_at(D.sc, ERR_CELL_RANGE, SITE_INBOX);
  }
  for (int part = 0; part < 2; part++)
    for (long long c0 = lo[part]; c0 < hi[part]; c0 += OWNCAP) {
      const long long c1 = min(c0 + (long long)OWNCAP, hi[part]);
      for (int j = threadIdx.x; j < c1 - c0; j += 256) { acc[0][j] = 0ull; acc[1][j] = 0ull; acc[2][j] = 0ull; }
      __syncthreads();
      for (int e = threadIdx.x; e < n; e += 256) {
        const long long q = box[e].p;
        if (q >= c0 && q < c1) {
          atomicAdd(&acc[0][q - c0], box[e].w[0]);
          atomicAdd(&acc[1][q - c0], box[e].w[1]);
          atomicAdd(&acc[2][q - c0], box[e].w[2]);
        }
      }
      __syncthreads();
      for (long long q = c0 + threadIdx.x; q < c1; q += 256)
        sb_cell(P, D, S, q, acc[0][q - c0], acc[1][q - c0], acc[2][q - c0], copies);
      __syncthreads();
    }
  if (threadIdx.x == 0) X.cnt[g] = 0;   // the inbox is read: empty for the next interval
}
__device__ __forceinline__ void sb_cell(const CoP& P, const CoD& D, const SpD& S, long long p, unsigned long long x0,
                                        unsigned long long x1, unsigned long long x2, bool copies) {
  const unsigned xm = D.sc->xmask;   // copies filled this interval
  if (p < P.mpcnumcell) {
    const gp<const unsigned long long> own = g1((const unsigned long long*)D.sown) + 3 * p;
    long long sx = (long long)(own[0] + x0), sy = (long long)(own[1] + x1), sz = (long long)(own[2] + x2);
    for (unsigned m = copies ? xm : 0u; m; m &= m - 1) {
      const gp<unsigned long long> a = g1(D.sacc + (long long)(__ffs(m) - 1) * 3 * P.mpcnumcell + 3 * p);
      sx += (long long)a[0]; sy += (long long)a[1]; sz += (long long)a[2];
      a[0] = 0ull; a[1] = 0ull; a[2] = 0ull;
    }
    const int n = (int)(sz & ((1ll << SB_NBITS) - 1));
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (n) {
      const double dn = (double)n;
      cx = (float)(((double)sx / SB_SCALE) / dn);
      cy = (float)(((double)sy / SB_SCALE) / dn);
      cz = (float)(((double)((sz - n) >> SB_NBITS) / SB_SCALE) / dn);
    }
    double u1, u2;   // the draw is keyed by the collision cell itself
    philox_uniform2(P.seed, 1, D.sc->coll, (unsigned long long)g1(D.mpc_cell)[p], u1, u2);
    const double phi = P.twopi * u1, th = 2.0 * u2 - 1.0, st = sqrt(1.0 - th * th);
    double sp, cp;
    sincos(phi, &sp, &cp);
    const gp<f4v> rec = g1((f4v*)D.srec);
    rec[2 * p] = f4v{cx, cy, cz, (float)n};
    rec[2 * p + 1] = f4v{(float)(st * cp), (float)(st * sp), (float)th, 0.0f};
  }
  if (p < P.vol) {   // the unshifted cell at the same tour position (census by tour position)
    const long long t = g1(D.mpc_cell)[p];
    int nb = 0;
    unsigned first = 0xffffffffu;
    for (unsigned m = xm; m; m &= m - 1) {
      const int x = __ffs(m) - 1;
      const gp<int> tc = g1(D.xtc + (long long)x * P.vol);
      const gp<unsigned> fb = g1(D.xfb + (long long)x * P.vol);
      nb += tc[p];
      first = min(first, fb[p]);
      tc[p] = 0;
      fb[p] = 0xffffffffu;
    }
    if (nb > CENSUS_MAX) raise_error(&D.sc->err, ERR_MEMBERS);   // the reference's MaxB
    if (t < P.rxnnumcell) {
      const double prob = g1(D.ptab)[min(nb, CENSUS_MAX)];
      double u, w;
      philox_uniform2(P.seed, 2, D.sc->coll, (unsigned long long)t, u, w);
      if (0.0 < u && u <= prob) {
        const int i = first >= (unsigned)P.NT ? (int)(first - (unsigned)P.NT) : (int)first;
        g1(S.flag)[i] = 0;
      }
    }
  }
}

// Tile of particles of one system: the rotation in FP32 with its cell's
// record (mpc_rotation_mod.f95:198-230) and the next interval's near set: a
// particle is near when its distance byte (beyond a list radius) is within
// |v| tau + guard, widened by the spheres' largest displacement since the
// interval's start when the distances were measured there (classify 2).
// Whole words of the near bitmap are written (tiles are multiples of 256).
__global__ __launch_bounds__(256) void sb_rotate(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds,
                                                 const SpD* __restrict__ Ss, const Tile* __restrict__ tiles,
                                                 int classify) {
  const Tile tl = tiles[blockIdx.x];
  const CoP P = Ps[tl.r];
  const CoD D = Ds[tl.r];
  const SpD S = Ss[tl.r];
  const double tau = P.tau * (1.0 + 1e-9);
  const double reach0 = P.guard + 1e-6 + (classify == 2 ? start_shift(P, D) : 0.0);
  const int lane = threadIdx.x & 63;
  if (tl.begin == 0 && threadIdx.x == 0) { D.sc->xmask = 0u; D.sc->sums_copied = 0u; }   // sb_cells_own merged the copies
  const gp<const int> cell_of = g1((const int*)D.cell_of);
  const gp<const f4v> rec = g1((const f4v*)D.srec);
  const gp<const unsigned char> dq = g1((const unsigned char*)D.dq);
  const gp<unsigned> bits = g1(D.near_bits);
  const gp<float> svx = g1(S.vx), svy = g1(S.vy), svz = g1(S.vz);
  float fmax2 = 0.0f;
  for (long long base = tl.begin; base < tl.end; base += 256) {
    const long long i = base + threadIdx.x;
    bool nb = false;
    if (i < tl.end) {
      const int t = cell_of[i];
      float vx = svx[i], vy = svy[i], vz = svz[i];
      if (t >= 0 && t < P.mpcnumcell) {
        const f4v r0 = rec[2 * (long long)t], r1 = rec[2 * (long long)t + 1];
        sp_rotate_velocity(P, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, vx, vy, vz);
        svx[i] = vx; svy[i] = vy; svz[i] = vz;
      } else {
        raise_error_at(D.sc, ERR_CELL_RANGE, SITE_ROTATE);
      }
      const float v2 = (vx * vx + vy * vy) + vz * vz;
      nb = (double)dq[i] * (1.0 / 32.0) <= sqrt((double)v2) * tau + reach0;
      if (!nb) fmax2 = fmaxf(fmax2, v2);
    }
    const unsigned long long m = __ballot(nb);
    const long long w0 = (base + (threadIdx.x & ~63)) >> 5, nw = (tl.end + 31) >> 5;   // words below the tile's end
    if (lane == 0 && w0 < nw) bits[w0] = (unsigned)m;
    if (lane == 32 && w0 + 1 < nw) bits[w0 + 1] = (unsigned)(m >> 32);
  }
  sp_far_max_publish(&D.sc->far_maxv2, fmax2);
}

// Per system (blockIdx.y): cleared cell sums, census and first-B keys of
// every XCD copy (every split run starts from them; each interval's sb_cells_own
// leaves the copies it merged cleared)
__global__ __launch_bounds__(256) void sb_clear(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds) {
  const int r = blockIdx.y;
  const CoP P = Ps[r];
  const CoD D = Ds[r];
  const long long i0 = (long long)blockIdx.x * 256 + threadIdx.x, st = (long long)gridDim.x * 256;
  for (long long i = i0; i < NXCD * 3 * P.mpcnumcell; i += st) D.sacc[i] = 0ull;
  for (long long i = i0; i < NXCD * P.vol; i += st) D.xtc[i] = 0;
  for (long long i = i0; i < NXCD * P.vol; i += st) D.xfb[i] = 0xffffffffu;
  if (i0 == 0) { D.sc->xmask = 0u; D.sc->sums_copied = 0u; }
}

// Hilbert re-sort of split storage (sort_hilbert_mod.f95:50-92), per tile:
// bin by the tour position of the unshifted cell; gather into the second buffers
__global__ __launch_bounds__(256) void sb_sort_bin(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds,
                                                   const SpD* __restrict__ Ss, const CoSortD* __restrict__ Xs,
                                                   const Tile* __restrict__ tiles) {
  const Tile tl = tiles[blockIdx.x];
  const CoP P = Ps[tl.r];
  const CoD D = Ds[tl.r];
  const SpD S = Ss[tl.r];
  const CoSortD X = Xs[tl.r];
  // whole waves run the loop (tiles are multiples of 256): one counter atomic
  // per distinct cell of the wave (the particles are in tour order)
  for (long long base = tl.begin; base < tl.end; base += 256) {
    const long long i = base + threadIdx.x;
    int t = 0;
    bool ok = false;
    if (i < tl.end) {
      int ix, iy, iz;
      sp_cell_xyz(P, S.cell[i], ix, iy, iz);
      const long long u = (long long)ix + (long long)P.Lx * (iy + (long long)P.Ly * iz);
      if (ix >= P.Lx || iy >= P.Ly || iz >= P.Lz || X.tourpos[u] < 0) {
        raise_error_at(D.sc, ERR_CELL_RANGE, SITE_SORT_BIN);
        X.key[i] = 0; X.rk[i] = 0;
      } else {
        t = X.tourpos[u];
        ok = true;
      }
    }
    const int r = wave_atomic_rank(X.scnt, t, ok);
    if (ok) {
      if (r >= BMAX) raise_error(&D.sc->err, ERR_BUCKET);
      X.key[i] = t;
      X.rk[i] = (unsigned short)min(r, 65535);
    }
  }
}
// The gather of the re-sort, one field per pass (pass 0 the cell with the
// status byte and the near bit, 1 the species, 2-7 the offsets and
// velocities): each pass's random reads fall in one array of the replicas a
// XCD is working through, a few MB that its L2 holds, where the fields all
// at once overflowed it and fetched a line per element read.
template <int PASS>
__global__ __launch_bounds__(256) void sb_sort_gather(const CoD* __restrict__ Ds, const SpD* __restrict__ Ss,
                                                      const SpD* __restrict__ Ns, const CoSortD* __restrict__ Xs,
                                                      const Tile* __restrict__ tiles) {
  const Tile tl = tiles[blockIdx.x];
  const CoSortD X = Xs[tl.r];
  const SpD S = Ss[tl.r], N = Ns[tl.r];
  const gp<const int> sb = g1((const int*)X.sb);
  if constexpr (PASS == 0) {
    const CoD D = Ds[tl.r];
    const int lane = threadIdx.x & 63;
    for (long long base = tl.begin; base < tl.end; base += 256) {
      const long long j = base + threadIdx.x;
      bool bit = false;
      if (j < tl.end) {
        const int i = sb[j];
        g1(N.cell)[j] = g1((const int*)S.cell)[i];
        g1(X.astat)[j] = g1((const unsigned char*)D.status)[i];
        bit = (g1((const unsigned*)D.near_bits)[i >> 5] >> (i & 31)) & 1u;
      }
      const unsigned long long m = __ballot(bit);
      const long long w0 = (base + (threadIdx.x & ~63)) >> 5, nw = (tl.end + 31) >> 5;
      if (lane == 0 && w0 < nw) g1(X.abits)[w0] = (unsigned)m;
      if (lane == 32 && w0 + 1 < nw) g1(X.abits)[w0 + 1] = (unsigned)(m >> 32);
    }
  } else if constexpr (PASS == 1) {
    for (long long j = tl.begin + threadIdx.x; j < tl.end; j += 256)
      g1(N.flag)[j] = g1((const signed char*)S.flag)[sb[j]];
  } else {
    const float* src = PASS == 2 ? S.ox : PASS == 3 ? S.oy : PASS == 4 ? S.oz : PASS == 5 ? S.vx : PASS == 6 ? S.vy : S.vz;
    float* dst = PASS == 2 ? N.ox : PASS == 3 ? N.oy : PASS == 4 ? N.oz : PASS == 5 ? N.vx : PASS == 6 ? N.vy : N.vz;
    const gp<const float> s1 = g1(src);
    const gp<float> d1 = g1(dst);
    for (long long j = tl.begin + threadIdx.x; j < tl.end; j += 256) d1[j] = s1[sb[j]];
  }
}

}  // namespace mpcx

struct mpcx_cob {
  int R = 0, nd = 1, S = 0, sort_step = 0;
  std::vector<mpcx_co*> eng;
  CoP* d_P = nullptr;
  CoD* d_D = nullptr;
  std::vector<CoD> h_D;
  Tile *d_pt = nullptr, *d_ct = nullptr, *d_vt = nullptr, *d_gt = nullptr;
  int n_pt = 0, n_ct = 0, n_vt = 0, tile = 2048;
  int bulk_slots = 0;   // cell-table slots of sb_bulk: 1024, 2048, or 0 (2048 for tiles of 4096 and more)
  // split replicas: particle tiles of sb_bulk and near-set tiles of sb_near
  // placed on the XCDs (workgroup b runs on XCD b % 8): each replica on one
  // XCD when there are enough replicas to balance them, so its sums fill one
  // accumulator copy; a solo system spreads over all eight
  Tile *d_xt = nullptr, *d_nt = nullptr;
  int n_xt = 0, n_nt = 0;
  std::vector<int> xcd_of;   // XCD of each replica (-1: spread)
  // split replicas: each particle tile's inbox of sums binned by other tiles
  // into the cells it owns (tile g of replica r: tb[r] + first particle / tile)
  int* d_tb = nullptr;
  int* d_icnt = nullptr;
  FxEntry* d_ibox = nullptr;
  int2* d_orng = nullptr;
  int n_tiles = 0, icap = 0;
  int icap_set = 0;   // entries per inbox when set (else max(tile / 2, 256))
  CoSortD* d_X = nullptr;   // sort buffers of each replica
  std::vector<CoSortD> h_X;
  // split storage: the replicas' split arrays and their second (sort) buffers
  bool split = false;
  SpD *d_S = nullptr, *d_N = nullptr;
  std::vector<SpD> h_S, h_N;
  // solo: one engine run through the split interval on its own streams (its
  // near engine in whatever mode the engine uses); otherwise one
  // workgroup per replica.  overlap: the far particles stream beside the near
  // engines (one-workgroup engines only: a cooperative near engine needs
  // every workgroup resident)
  bool solo = false, overlap = true;
  hipStream_t st = nullptr, st2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double run_ms = 0;
  long long particles = 0, cells = 0;
  int fill_blocks = 64;
  // kernel timing of split runs without a graph: every TSTRIDE-th interval,
  // [0] near engines, [1] sb_bulk (far flight and binning), [2] sb_cells_own,
  // [3] sb_rotate, [4] sb_near (near set binning)
  static constexpr int NKT = 5, TMAX = 256, TSTRIDE = 8;
  std::vector<hipEvent_t> tev;
  std::vector<char> tev_on;
  double kt_ms[NKT] = {0};
  long long kt_n = 0;
  // each replica's step counter (run start) or error flag (run end), read in one copy
  int* d_io = nullptr;
  std::vector<int> h_io;
};

// every replica's step counter into io[r], its collision counter set from it
// (the host checks the steps before any interval runs)
__global__ void cob_begin_k(const CoD* __restrict__ D, int R, int S, int* __restrict__ io) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  CoScalars* sc = D[r].sc;
  const int cs = sc->cstep;
  io[r] = cs;
  sc->coll = (unsigned long long)(cs / S);
}

// every replica's device error flag into io[r]
__global__ void cob_errs_k(const CoD* __restrict__ D, int R, int* __restrict__ io) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < R) io[r] = D[r].sc->err;
}

static int cob_tiles(mpcx_cob* b) {
  std::vector<Tile> pt, ct, vt;
  for (int r = 0; r < b->R; r++) {
    const CoP& P = b->eng[r]->P;
    for (long long i = 0; i < P.NT; i += b->tile) pt.push_back(Tile{r, 0, i, std::min<long long>(i + b->tile, P.NT)});
    for (long long t = 0; t < P.mpcnumcell; t += CT)
      ct.push_back(Tile{r, 0, t, std::min<long long>(t + CT, P.mpcnumcell)});
    for (long long t = 0; t < P.vol; t += 256) vt.push_back(Tile{r, 0, t, std::min<long long>(t + 256, P.vol)});
  }
  // workgroup b runs on XCD b % 8: lay the tiles out so that each XCD's
  // workgroups take one contiguous run of tiles
  auto xcd_order = [](std::vector<Tile>& v) {
    const size_t n = v.size(), q = n / 8, r = n % 8;
    std::vector<Tile> o(n);
    for (size_t b = 0; b < n; b++) {
      const size_t x = b % 8, k = b / 8;
      o[b] = v[x * q + std::min(x, r) + k];
    }
    v.swap(o);
  };
  std::vector<Tile> gt = pt;   // gather kernels (sort gather) take the XCD-contiguous order
  xcd_order(gt);
  xcd_order(ct);
  // split: replicas onto XCDs, largest first onto the least loaded (when each
  // XCD gets at least two and no replica exceeds 1/16 of the particles)
  std::vector<Tile> xt, nt;
  {
    long long total = 0, big = 0;
    for (int r = 0; r < b->R; r++) { total += b->eng[r]->P.NT; big = std::max<long long>(big, b->eng[r]->P.NT); }
    const bool place = b->R >= 2 * NXCD && big * 16 <= total;
    b->xcd_of.assign(b->R, -1);
    std::vector<std::vector<Tile>> xq(NXCD), nq(NXCD);
    const int K = b->solo ? grid_for(b->eng[0]->D.near_cap, 256, 1024) : 4;   // near-set chunks per replica
    if (place) {
      std::vector<int> order(b->R);
      for (int r = 0; r < b->R; r++) order[r] = r;
      std::stable_sort(order.begin(), order.end(),
                       [&](int a, int c) { return b->eng[a]->P.NT > b->eng[c]->P.NT; });
      long long load[NXCD] = {0};
      for (int r : order) {
        const int x = (int)(std::min_element(load, load + NXCD) - load);
        b->xcd_of[r] = x;
        load[x] += b->eng[r]->P.NT;
      }
      for (const Tile& t : pt) xq[b->xcd_of[t.r]].push_back(t);
      for (int r = 0; r < b->R; r++)
        for (int k = 0; k < K; k++) nq[b->xcd_of[r]].push_back(Tile{r, 0, k, K});
      auto interleave = [](std::vector<std::vector<Tile>>& q, std::vector<Tile>& out) {
        size_t m = 0;
        for (auto& v : q) m = std::max(m, v.size());
        out.assign(m * NXCD, Tile{0, 0, 0, 0});   // begin == end: an empty tile
        for (int x = 0; x < NXCD; x++)
          for (size_t k = 0; k < q[x].size(); k++) out[k * NXCD + x] = q[x][k];
      };
      interleave(xq, xt);
      interleave(nq, nt);
    } else {
      xt = pt;
      for (int r = 0; r < b->R; r++)
        for (int k = 0; k < K; k++) nt.push_back(Tile{r, 0, k, K});
    }
  }
  for (Tile** d : {&b->d_pt, &b->d_ct, &b->d_vt, &b->d_gt, &b->d_xt, &b->d_nt})
    if (*d) { (void)hipFree(*d); *d = nullptr; }
  HIPCHK(hipMalloc(&b->d_xt, xt.size() * sizeof(Tile)));
  HIPCHK(hipMalloc(&b->d_nt, nt.size() * sizeof(Tile)));
  HIPCHK(hipMemcpy(b->d_xt, xt.data(), xt.size() * sizeof(Tile), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_nt, nt.data(), nt.size() * sizeof(Tile), hipMemcpyHostToDevice));
  b->n_xt = (int)xt.size();
  b->n_nt = (int)nt.size();
  HIPCHK(hipMalloc(&b->d_pt, pt.size() * sizeof(Tile)));
  HIPCHK(hipMalloc(&b->d_gt, gt.size() * sizeof(Tile)));
  HIPCHK(hipMalloc(&b->d_ct, ct.size() * sizeof(Tile)));
  HIPCHK(hipMalloc(&b->d_vt, vt.size() * sizeof(Tile)));
  HIPCHK(hipMemcpy(b->d_gt, gt.data(), gt.size() * sizeof(Tile), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_pt, pt.data(), pt.size() * sizeof(Tile), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_ct, ct.data(), ct.size() * sizeof(Tile), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_vt, vt.data(), vt.size() * sizeof(Tile), hipMemcpyHostToDevice));
  b->n_pt = (int)pt.size();
  b->n_ct = (int)ct.size();
  b->n_vt = (int)vt.size();
  if (b->split) {   // the inboxes: max(tile / 2, 256) entries per particle tile
    for (void** d : {(void**)&b->d_tb, (void**)&b->d_icnt, (void**)&b->d_ibox, (void**)&b->d_orng})
      if (*d) { (void)hipFree(*d); *d = nullptr; }
    std::vector<int> tb(b->R);
    long long n = 0;
    for (int r = 0; r < b->R; r++) {
      tb[r] = (int)n;
      n += (b->eng[r]->P.NT + b->tile - 1) / b->tile;
    }
    b->icap = b->icap_set > 0 ? b->icap_set : std::max(b->tile / 2, 256);
    if (n * (long long)b->icap > (1ll << 31)) return set_error(MPCX_E_ARG, "too many particle tiles for the inboxes");
    b->n_tiles = (int)n;
    HIPCHK(hipMalloc(&b->d_tb, b->R * sizeof(int)));
    HIPCHK(hipMalloc(&b->d_icnt, n * sizeof(int)));
    HIPCHK(hipMalloc(&b->d_ibox, n * (size_t)b->icap * sizeof(FxEntry)));
    HIPCHK(hipMalloc(&b->d_orng, n * sizeof(int2)));
    HIPCHK(hipMemcpy(b->d_tb, tb.data(), b->R * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(b->d_icnt, 0, n * sizeof(int)));
  }
  return 0;
}
static SbInbox cob_inbox(const mpcx_cob* b) {
  SbInbox X;
  X.tb = b->d_tb; X.cnt = b->d_icnt; X.box = b->d_ibox; X.rng = b->d_orng; X.T = b->tile; X.cap = b->icap;
  return X;
}

static int cob_upload_D(mpcx_cob* b) {
  for (int r = 0; r < b->R; r++) {
    mpcx_co* e = b->eng[r];
    b->h_D[r] = e->D;
    b->h_D[r].spd = (b->split && !b->solo) ? b->d_S + r : nullptr;   // batch near engines gather from split storage
    CoSortD& X = b->h_X[r];
    X.tourpos = e->d_tourpos; X.scnt = e->d_scnt; X.soff = e->d_soff; X.sb = e->d_sb; X.key = e->d_skey;
    X.rk = e->d_srk;
    for (int a = 0; a < 6; a++) X.alt[a] = e->alt[a];
    X.aflag = e->alt_flag; X.astat = e->alt_status; X.abits = e->alt_bits;
    if (b->split) {
      b->h_S[r] = e->S;
      b->h_N[r] = e->Salt;
    }
  }
  HIPCHK(hipMemcpyAsync(b->d_D, b->h_D.data(), b->R * sizeof(CoD), hipMemcpyHostToDevice, b->st));
  HIPCHK(hipMemcpyAsync(b->d_X, b->h_X.data(), b->R * sizeof(CoSortD), hipMemcpyHostToDevice, b->st));
  if (b->split) {
    HIPCHK(hipMemcpyAsync(b->d_S, b->h_S.data(), b->R * sizeof(SpD), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(b->d_N, b->h_N.data(), b->R * sizeof(SpD), hipMemcpyHostToDevice, b->st));
  }
  HIPCHK(hipStreamSynchronize(b->st));
  return 0;
}

// the Hilbert re-sort of every replica (sort_hilbert_mod.f95:50-92), then the
// double buffers of each engine swap roles
static int cob_sort_all(mpcx_cob* b) {
  for (int r = 0; r < b->R; r++)
    if (!b->eng[r]->has_tour) return set_error(MPCX_E_ARG, "a sort step needs the cell tour of every replica");
  const dim3 fg(b->fill_blocks, b->R);
  hipLaunchKernelGGL(sort_fills_b, fg, dim3(256), 0, b->st, b->d_P, b->d_X);
  if (b->split)
    hipLaunchKernelGGL(sb_sort_bin, dim3(b->n_pt), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_S, b->d_X, b->d_pt);
  else hipLaunchKernelGGL(sort_bin_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_X, b->d_pt);
  hipLaunchKernelGGL(sort_scan_b, dim3(b->R), dim3(1024), 0, b->st, b->d_P, b->d_X);
  hipLaunchKernelGGL(sort_scatter_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_X, b->d_pt);
  hipLaunchKernelGGL(sort_order_b, dim3(b->n_vt), dim3(256), 0, b->st, b->d_X, b->d_vt);
  if (b->split) {   // one field per pass
    const dim3 g(b->n_pt), t(256);
    hipLaunchKernelGGL(sb_sort_gather<0>, g, t, 0, b->st, b->d_D, b->d_S, b->d_N, b->d_X, b->d_gt);
    hipLaunchKernelGGL(sb_sort_gather<1>, g, t, 0, b->st, b->d_D, b->d_S, b->d_N, b->d_X, b->d_gt);
    hipLaunchKernelGGL(sb_sort_gather<2>, g, t, 0, b->st, b->d_D, b->d_S, b->d_N, b->d_X, b->d_gt);
    hipLaunchKernelGGL(sb_sort_gather<3>, g, t, 0, b->st, b->d_D, b->d_S, b->d_N, b->d_X, b->d_gt);
    hipLaunchKernelGGL(sb_sort_gather<4>, g, t, 0, b->st, b->d_D, b->d_S, b->d_N, b->d_X, b->d_gt);
    hipLaunchKernelGGL(sb_sort_gather<5>, g, t, 0, b->st, b->d_D, b->d_S, b->d_N, b->d_X, b->d_gt);
    hipLaunchKernelGGL(sb_sort_gather<6>, g, t, 0, b->st, b->d_D, b->d_S, b->d_N, b->d_X, b->d_gt);
    hipLaunchKernelGGL(sb_sort_gather<7>, g, t, 0, b->st, b->d_D, b->d_S, b->d_N, b->d_X, b->d_gt);
  }
  else hipLaunchKernelGGL(sort_gather_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_D, b->d_X, b->d_gt);
  for (int r = 0; r < b->R; r++) {
    mpcx_co* e = b->eng[r];
    CoD& D = e->D;
    if (b->split) {
      std::swap(e->S, e->Salt);
    } else {
      std::swap(D.x, e->alt[0]); std::swap(D.y, e->alt[1]); std::swap(D.z, e->alt[2]);
      std::swap(D.vx, e->alt[3]); std::swap(D.vy, e->alt[4]); std::swap(D.vz, e->alt[5]);
      std::swap(D.flag, e->alt_flag);
    }
    std::swap(D.status, e->alt_status); std::swap(D.near_bits, e->alt_bits);
  }
  const int rc = cob_upload_D(b);
  if (rc == 0 && b->split)   // the cells each particle tile owns in the new order
    hipLaunchKernelGGL(sb_own_ranges, dim3((b->n_xt + 255) / 256), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_xt,
                       b->n_xt, cob_inbox(b));
  return rc;
}

static hipError_t cob_launch_near(mpcx_cob* b, int nsteps, int mode) {
  if (b->solo) return launch_near(b->eng[0], nsteps, mode);   // the engine's own stream is b->st
  if (b->nd == 1)
    hipLaunchKernelGGL(co_near_engine_b<1>, dim3(b->R), dim3(NE_T), 0, b->st, b->d_P, b->d_D, nsteps, mode);
  else
    hipLaunchKernelGGL(co_near_engine_b<2>, dim3(b->R), dim3(NE_T), 0, b->st, b->d_P, b->d_D, nsteps, mode);
  return hipGetLastError();
}

// the near list of every replica from its bitmap
static void cob_near_list(mpcx_cob* b) {
  if (b->solo) co_near_list(b->eng[0]);   // many workgroups for one large system
  else hipLaunchKernelGGL(near_list_b, dim3(b->R), dim3(1024), 0, b->st, b->d_P, b->d_D);
}

// the launches of one interval without a re-sort (captured into the graph)
static void cob_interval(mpcx_cob* b) {
  const dim3 fg(b->fill_blocks, b->R);
  (void)cob_launch_near(b, b->S, 0);
  hipLaunchKernelGGL(fills_begin_b, fg, dim3(256), 0, b->st, b->d_P, b->d_D);
  hipLaunchKernelGGL(co_bulk_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_pt, 1, 1);
  hipLaunchKernelGGL(scan_cells_b, dim3(b->R), dim3(1024), 0, b->st, b->d_P, b->d_D);
  hipLaunchKernelGGL(co_scatter_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_D, b->d_pt);
  hipLaunchKernelGGL(co_collide_b, dim3(b->n_ct), dim3(CT), 0, b->st, b->d_P, b->d_D, b->d_ct);
  hipLaunchKernelGGL(fills_rotate_b, fg, dim3(256), 0, b->st, b->d_P, b->d_D);
  hipLaunchKernelGGL(co_rotate_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_pt);
  hipLaunchKernelGGL(near_list_b, dim3(b->R), dim3(1024), 0, b->st, b->d_P, b->d_D);
}

// an interval that ends in a re-sort: the near engines stop before the last
// step's list rebuild, every replica is sorted, then they resume
static int cob_sort_interval(mpcx_cob* b) {
  const dim3 fg(b->fill_blocks, b->R);
  HIPCHK(cob_launch_near(b, b->S, M_STOP));
  hipLaunchKernelGGL(fills_begin_b, fg, dim3(256), 0, b->st, b->d_P, b->d_D);
  hipLaunchKernelGGL(co_bulk_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_pt, 1, 0);
  int rc = cob_sort_all(b);
  if (rc) return rc;
  hipLaunchKernelGGL(near_list_b, dim3(b->R), dim3(1024), 0, b->st, b->d_P, b->d_D);
  HIPCHK(cob_launch_near(b, 1, M_RESUME));
  hipLaunchKernelGGL(co_bulk_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_pt, 0, 1);
  hipLaunchKernelGGL(scan_cells_b, dim3(b->R), dim3(1024), 0, b->st, b->d_P, b->d_D);
  hipLaunchKernelGGL(co_scatter_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_D, b->d_pt);
  hipLaunchKernelGGL(co_collide_b, dim3(b->n_ct), dim3(CT), 0, b->st, b->d_P, b->d_D, b->d_ct);
  hipLaunchKernelGGL(fills_rotate_b, fg, dim3(256), 0, b->st, b->d_P, b->d_D);
  hipLaunchKernelGGL(co_rotate_b, dim3(b->n_pt), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_pt);
  hipLaunchKernelGGL(near_list_b, dim3(b->R), dim3(1024), 0, b->st, b->d_P, b->d_D);
  return 0;
}

// sb_bulk over the XCD-placed particle tiles, 512 threads per workgroup; the
// replicas of a batch hold at most two dimers (checked at its creation), so
// their sphere table has 4 entries
template <int GT, int NS>
static void sb_launch_bulk_ns(mpcx_cob* b, bool all, hipStream_t s, int bin) {
  const dim3 g(b->n_xt);
  const SbInbox X = cob_inbox(b);
  if (all) hipLaunchKernelGGL((sb_bulk<true, 512, GT, NS>), g, dim3(512), 0, s, b->d_P, b->d_D, b->d_S, b->d_xt, bin, X);
  else hipLaunchKernelGGL((sb_bulk<false, 512, GT, NS>), g, dim3(512), 0, s, b->d_P, b->d_D, b->d_S, b->d_xt, bin, X);
}
// sb_near over the near-set tiles (the 4-sphere instance for a batch, as sb_bulk)
static void sb_launch_near(mpcx_cob* b, hipStream_t s, int bin, int check) {
  if (!b->solo && 2 * b->nd <= 4)
    hipLaunchKernelGGL(sb_near<4>, dim3(b->n_nt), dim3(256), 0, s, b->d_P, b->d_D, b->d_S, b->d_nt, bin, check, cob_inbox(b));
  else
    hipLaunchKernelGGL(sb_near<MAXSG>, dim3(b->n_nt), dim3(256), 0, s, b->d_P, b->d_D, b->d_S, b->d_nt, bin, check,
                       cob_inbox(b));
}
template <int GT>
static void sb_launch_bulk_t(mpcx_cob* b, bool all, hipStream_t s, int bin) {
  if (!b->solo && 2 * b->nd <= 4) sb_launch_bulk_ns<GT, 4>(b, all, s, bin);
  else sb_launch_bulk_ns<GT, MAXSG>(b, all, s, bin);
}
static void sb_launch_bulk(mpcx_cob* b, bool all, hipStream_t s, int bin) {
  const int slots = b->bulk_slots ? b->bulk_slots : (b->tile >= 4096 ? 2048 : 1024);
  if (slots == 2048) sb_launch_bulk_t<2048>(b, all, s, bin);
  else sb_launch_bulk_t<1024>(b, all, s, bin);
}

// One split-precision interval of every replica (the kernels above).  With
// overlap (a solo engine, whose one near-engine workgroup leaves the other
// CUs free) the far particles' flight and binning (st2) run beside the near
// engine (st); the near set is binned on st when it is done; every cell is
// collided and every particle rotated once both are.  A batch runs them in
// order: its near engines hold a CU each (LDS and VGPRs), so a far flight
// launched beside them only waits for CUs, in dispatch order, and started
// beside them or beside the near sets' binning it measured 3.78 and 3.81 ms
// per interval of the 256-replica ensemble against 3.77 in order.  A re-sort interval
// stops the near engines before the last step's list rebuild (the far flight
// is not binned), re-sorts, lets them resume, then bins every particle.
// slot >= 0: the interval's kernels are timed (events into that slot).
static int sb_interval(mpcx_cob* b, bool sort_due, long long slot) {
  hipStream_t st = b->st, s2 = b->st2;
  const dim3 ng(32, b->R);   // sb_load_near: blockIdx.y = replica
  const bool ov = b->overlap;
  auto tmark = [&](int k, int edge, hipStream_t s) -> hipError_t {
    if (slot < 0) return hipSuccess;
    b->tev_on[(size_t)slot * mpcx_cob::NKT + k] = 1;
    return hipEventRecord(b->tev[((size_t)slot * mpcx_cob::NKT + k) * 2 + edge], s);
  };
  if (slot >= 0)
    for (int k = 0; k < mpcx_cob::NKT; k++) b->tev_on[(size_t)slot * mpcx_cob::NKT + k] = 0;
  // the interval's start; a solo engine's near engine also needs its near set in the FP64 arrays
  if (b->solo) hipLaunchKernelGGL(sb_load_near, ng, dim3(256), 0, st, b->d_P, b->d_D, b->d_S, 1, 1);
  else hipLaunchKernelGGL(sb_load_near, dim3(1, b->R), dim3(256), 0, st, b->d_P, b->d_D, b->d_S, 1, 0);
  const int bin = sort_due ? 0 : 1;
  if (ov) {
    HIPCHK(hipEventRecord(b->fork, st));
    HIPCHK(hipStreamWaitEvent(s2, b->fork, 0));
  }
  // the near engines are submitted before the far flight: each needs most of
  // a CU's LDS, which the far flight's workgroups would otherwise take first
  HIPCHK(tmark(0, 0, st));
  HIPCHK(cob_launch_near(b, b->S, sort_due ? M_STOP : 0));
  HIPCHK(tmark(0, 1, st));
  if (ov) {
    HIPCHK(tmark(1, 0, s2));
    sb_launch_bulk(b, false, s2, bin);
    HIPCHK(tmark(1, 1, s2));
    HIPCHK(hipEventRecord(b->join, s2));
  }
  if (!ov) {
    HIPCHK(tmark(1, 0, st));
    sb_launch_bulk(b, false, st, bin);
    HIPCHK(tmark(1, 1, st));
  }
  HIPCHK(tmark(4, 0, st));
  sb_launch_near(b, st, bin, 1);
  HIPCHK(tmark(4, 1, st));
  if (ov) HIPCHK(hipStreamWaitEvent(st, b->join, 0));
  if (sort_due) {
    int rc = cob_sort_all(b);
    if (rc) return rc;
    cob_near_list(b);
    if (b->solo) hipLaunchKernelGGL(sb_load_near, ng, dim3(256), 0, st, b->d_P, b->d_D, b->d_S, 0, 1);
    HIPCHK(cob_launch_near(b, 1, M_RESUME));
    sb_launch_near(b, st, 0, 0);
    sb_launch_bulk(b, true, st, 1);
  }
  HIPCHK(tmark(2, 0, st));
  hipLaunchKernelGGL(sb_cells_own, dim3(b->n_xt), dim3(256), 0, st, b->d_P, b->d_D, b->d_S, b->d_xt, cob_inbox(b));
  HIPCHK(tmark(2, 1, st));
  HIPCHK(tmark(3, 0, st));
  hipLaunchKernelGGL(sb_rotate, dim3(b->n_pt), dim3(256), 0, st, b->d_P, b->d_D, b->d_S, b->d_pt, sort_due ? 1 : 2);
  HIPCHK(tmark(3, 1, st));
  cob_near_list(b);
  return 0;
}

static void cob_free_graph(mpcx_cob* b) {
  if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
}

extern "C" int mpcx_cob_destroy(mpcx_cob* b) {
  if (!b) return 0;
  handle_del(H_COB, b);
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->st2) (void)hipStreamSynchronize(b->st2);
  cob_free_graph(b);
  for (void* p : {(void*)b->d_P, (void*)b->d_D, (void*)b->d_pt, (void*)b->d_ct, (void*)b->d_vt, (void*)b->d_gt,
                  (void*)b->d_xt, (void*)b->d_nt, (void*)b->d_X, (void*)b->d_S, (void*)b->d_N, (void*)b->d_io,
                  (void*)b->d_tb, (void*)b->d_icnt, (void*)b->d_ibox, (void*)b->d_orng})
    if (p) (void)hipFree(p);
  for (hipEvent_t ev : {b->e0, b->e1, b->fork, b->join})
    if (ev) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : b->tev)
    if (ev) (void)hipEventDestroy(ev);
  if (!b->solo) {   // a solo batch runs on its engine's streams
    if (b->st) (void)hipStreamDestroy(b->st);
    if (b->st2) (void)hipStreamDestroy(b->st2);
  }
  delete b;
  return 0;
}

// solo: one split engine on its own streams (any near mode)
static int cob_make(mpcx_cob** out, int32_t R, mpcx_co* const* engines, int32_t tile, bool solo) {
  if (!out || R < 1 || !engines) return set_error(MPCX_E_ARG, "bad arguments");
  *out = nullptr;
  const mpcx_co* e0 = engines[0];
  for (int r = 0; r < R; r++) {
    const mpcx_co* e = engines[r];
    if (!e) return set_error(MPCX_E_ARG, "null engine in the batch");
    if (e->P.rng_mode != 1) return set_error(MPCX_E_UNSUPPORTED, "a batch runs in rng_mode 1 (device Philox)");
    if (e->nd != e0->nd || e->P.MPCtstep != e0->P.MPCtstep || e->P.sort_step != e0->P.sort_step)
      return set_error(MPCX_E_UNSUPPORTED, "replicas of a batch share numdims, MPCtstep and sort_step");
    if (e->split != e0->split) return set_error(MPCX_E_UNSUPPORTED, "replicas of a batch share their storage precision");
    if (!solo && (e->nd > 2 || e->near_mode != 0))
      return set_error(MPCX_E_UNSUPPORTED, "batched replicas hold one or two dimers (one-workgroup near engines)");
  }
  if (tile < 256 || tile % 256 != 0) return set_error(MPCX_E_ARG, "tile must be a positive multiple of 256");
  mpcx_cob* b = new mpcx_cob();
  b->R = R;
  b->nd = e0->nd;
  b->S = e0->P.MPCtstep;
  b->sort_step = e0->P.sort_step;
  b->tile = tile;
  b->split = e0->split;
  b->solo = solo;
  b->overlap = solo && e0->near_mode == 0 && e0->overlap;   // a batch runs its interval in order
  b->eng.assign(engines, engines + R);
  b->h_D.resize(R);
  b->h_X.resize(R);
  b->h_S.resize(R);
  b->h_N.resize(R);
  b->h_io.resize(R);
  std::vector<CoP> hp(R);
  for (int r = 0; r < R; r++) {
    hp[r] = engines[r]->P;
    b->particles += engines[r]->P.NT;
    b->cells += engines[r]->P.mpcnumcell;
  }
  int rc = 0;
  if (solo) {
    b->st = e0->st;
    b->st2 = e0->st2;
  } else if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
             hipStreamCreateWithFlags(&b->st2, hipStreamNonBlocking) != hipSuccess) {
    rc = set_error(MPCX_E_HIP, "batch stream creation failed");
  }
  if (!rc && (hipMalloc(&b->d_P, R * sizeof(CoP)) != hipSuccess || hipMalloc(&b->d_D, R * sizeof(CoD)) != hipSuccess ||
              hipMalloc(&b->d_X, R * sizeof(CoSortD)) != hipSuccess || hipMalloc(&b->d_S, R * sizeof(SpD)) != hipSuccess ||
              hipMalloc(&b->d_N, R * sizeof(SpD)) != hipSuccess || hipMalloc(&b->d_io, R * sizeof(int)) != hipSuccess ||
              hipEventCreate(&b->e0) != hipSuccess ||
              hipEventCreate(&b->e1) != hipSuccess ||
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess))
    rc = set_error(MPCX_E_HIP, "batch allocation failed");
  if (!rc && hipMemcpy(b->d_P, hp.data(), R * sizeof(CoP), hipMemcpyHostToDevice) != hipSuccess)
    rc = set_error(MPCX_E_HIP, "hipMemcpy failed");
  if (!rc) rc = cob_tiles(b);
  if (rc) {
    mpcx_cob_destroy(b);
    return rc;
  }
  *out = b;
  return 0;
}

extern "C" int mpcx_cob_create(mpcx_cob** out, int32_t R, mpcx_co* const* engines, int32_t tile) {
  const int rc = cob_make(out, R, engines, tile, false);
  if (!rc) handle_add(H_COB, *out);
  return rc;
}

// checks that every replica sits at the same collision step, sets the
// collision counters and uploads the replica descriptors
static int cob_begin(mpcx_cob* b, int* cstep_out) {
  if (!b->solo)
    for (int r = 0; r < b->R; r++) HIPCHK(hipStreamSynchronize(b->eng[r]->st));
  int rc = cob_upload_D(b);
  if (rc) return rc;
  hipLaunchKernelGGL(cob_begin_k, dim3((b->R + 255) / 256), dim3(256), 0, b->st, b->d_D, b->R, b->S, b->d_io);
  HIPCHK(hipMemcpyAsync(b->h_io.data(), b->d_io, b->R * sizeof(int), hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  const int cstep = b->h_io[0];
  for (int r = 0; r < b->R; r++) {
    if (b->h_io[r] % b->S != 0) return set_error(MPCX_E_ARG, "a batch runs from a collision step");
    if (b->h_io[r] != cstep) return set_error(MPCX_E_ARG, "replicas of a batch must be at the same step");
  }
  *cstep_out = cstep;
  if (b->split) {
    // a run that ends without an error leaves every accumulator copy cleared
    // (the cell pass clears what it reads); after a failed one, clear them all
    long long big = 0;
    bool dirty = false;
    for (int r = 0; r < b->R; r++) {
      big = std::max(big, NXCD * 3 * std::max(b->eng[r]->P.mpcnumcell, b->eng[r]->P.vol));
      dirty = dirty || b->eng[r]->split_dirty;
    }
    if (dirty) {
      hipLaunchKernelGGL(sb_clear, dim3(grid_for(big, 256, 1024), b->R), dim3(256), 0, b->st, b->d_P, b->d_D);
      for (int r = 0; r < b->R; r++) b->eng[r]->split_dirty = false;
    }
    hipLaunchKernelGGL(fill_kernel<int>, dim3(grid_for(b->n_tiles, 256, 1024)), dim3(256), 0, b->st, b->d_icnt,
                       (long long)b->n_tiles, 0);   // inboxes of an interval that ended in an error
    hipLaunchKernelGGL(sb_own_ranges, dim3((b->n_xt + 255) / 256), dim3(256), 0, b->st, b->d_P, b->d_D, b->d_xt,
                       b->n_xt, cob_inbox(b));
  }
  return 0;
}

// launches one interval (graph replay when possible); slot >= 0 times it
static int cob_step(mpcx_cob* b, int* cstep, int use_graph, long long slot = -1) {
  const int cnext = *cstep + b->S;
  const bool sort_due = cnext % b->sort_step == 0;
  int rc = 0;
  if (b->split) {
    if (sort_due || !use_graph || slot >= 0) {
      rc = sb_interval(b, sort_due, slot);
    } else {
      if (!b->gexec) {
        HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        rc = sb_interval(b, false, -1);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->st, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        HIPCHK(ce);
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
      }
      HIPCHK(hipGraphLaunch(b->gexec, b->st));
    }
  } else if (sort_due) {
    rc = cob_sort_interval(b);
  } else if (use_graph) {
    if (!b->gexec) {
      HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
      cob_interval(b);
      HIPCHK(hipStreamEndCapture(b->st, &b->graph));
      HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->gexec, b->st));
  } else {
    cob_interval(b);
  }
  for (int r = 0; r < b->R; r++) b->eng[r]->cells_valid = true;
  *cstep = cnext;
  return rc;
}

static int cob_end(mpcx_cob* b) {
  hipLaunchKernelGGL(cob_errs_k, dim3((b->R + 255) / 256), dim3(256), 0, b->st, b->d_D, b->R, b->d_io);
  HIPCHK(hipMemcpyAsync(b->h_io.data(), b->d_io, b->R * sizeof(int), hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  for (int r = 0; r < b->R; r++) {
    if (!b->h_io[r]) continue;
    int rc = check_device_error(b->eng[r], r);   // reports and clears it
    if (rc) return rc;
  }
  return 0;
}

static int cob_run_body(mpcx_cob* b, int64_t ncoll, int32_t use_graph);
static void cob_mark_dirty(mpcx_cob* b) {
  for (int r = 0; r < b->R; r++) b->eng[r]->split_dirty = true;
}
static int cob_run(mpcx_cob* b, int64_t ncoll, int32_t use_graph) {
  const int rc0 = cob_run_body(b, ncoll, use_graph);
  if (rc0) cob_mark_dirty(b);
  return rc0;
}
static int cob_run_body(mpcx_cob* b, int64_t ncoll, int32_t use_graph) {
  int cstep = 0;
  int rc = cob_begin(b, &cstep);
  if (rc) return rc;
  const bool timed_run = b->split && !use_graph;
  if (timed_run && b->tev.empty()) {
    b->tev.assign((size_t)mpcx_cob::TMAX * mpcx_cob::NKT * 2, nullptr);
    b->tev_on.assign((size_t)mpcx_cob::TMAX * mpcx_cob::NKT, 0);
    for (auto& ev : b->tev) HIPCHK(hipEventCreate(&ev));
  }
  HIPCHK(hipEventRecord(b->e0, b->st));
  for (int64_t c = 0; c < ncoll && !rc; c++) {
    const long long slot = c / mpcx_cob::TSTRIDE;
    const bool timed = timed_run && c % mpcx_cob::TSTRIDE == 0 && slot < mpcx_cob::TMAX;
    rc = cob_step(b, &cstep, use_graph, timed ? slot : -1);
  }
  if (rc) return rc;
  HIPCHK(hipEventRecord(b->e1, b->st));
  HIPCHK(hipEventSynchronize(b->e1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, b->e0, b->e1);
  b->run_ms = ms;
  if (timed_run) {
    b->kt_n = std::min<long long>((ncoll + mpcx_cob::TSTRIDE - 1) / mpcx_cob::TSTRIDE, mpcx_cob::TMAX);
    for (int k = 0; k < mpcx_cob::NKT; k++) {
      double sum = 0;
      long long n = 0;
      for (long long c = 0; c < b->kt_n; c++) {
        if (!b->tev_on[(size_t)c * mpcx_cob::NKT + k]) continue;
        float t = 0;
        (void)hipEventElapsedTime(&t, b->tev[((size_t)c * mpcx_cob::NKT + k) * 2],
                                  b->tev[((size_t)c * mpcx_cob::NKT + k) * 2 + 1]);
        sum += t;
        n++;
      }
      b->kt_ms[k] = n ? sum / (double)n : 0.0;
    }
  }
  return cob_end(b);
}

extern "C" int mpcx_cob_run(mpcx_cob* b, int64_t ncoll, int32_t use_graph) {
  if (!b || ncoll < 0) return set_error(MPCX_E_ARG, "bad arguments");
  return cob_run(b, ncoll, use_graph);
}

// mean kernel durations (ms) of the last split run without a graph
extern "C" int mpcx_cob_kernel_ms(mpcx_cob* b, double* ms5, int64_t* count) {
  if (!b || !ms5) return set_error(MPCX_E_ARG, "bad arguments");
  for (int k = 0; k < mpcx_cob::NKT; k++) ms5[k] = b->kt_ms[k];
  if (count) *count = b->kt_n;
  return 0;
}

// particles and collision cells of the batch, and the wall time of the last run (ms)
extern "C" int mpcx_cob_info(mpcx_cob* b, int64_t* particles, int64_t* cells, double* run_ms) {
  if (!b) return set_error(MPCX_E_ARG, "null batch");
  if (particles) *particles = b->particles;
  if (cells) *cells = b->cells;
  if (run_ms) *run_ms = b->run_ms;
  return 0;
}

// cell-table slots of the split binning (1024, 2048, or 0: by the tile); drops a captured graph
extern "C" int mpcx_cob_set_bulk_slots(mpcx_cob* b, int32_t slots) {
  if (!b || (slots != 0 && slots != 1024 && slots != 2048))
    return set_error(MPCX_E_ARG, "bulk slots must be 0, 1024 or 2048");
  HIPCHK(hipStreamSynchronize(b->st));
  cob_free_graph(b);
  b->bulk_slots = slots;
  return 0;
}

// entries per inbox of a split batch (0: max(tile / 2, 256)); the sums that
// do not fit go through the atomics into the XCD copies, so the results do
// not depend on it; drops a captured graph
extern "C" int mpcx_cob_set_inbox_cap(mpcx_cob* b, int32_t cap) {
  if (!b || cap < 0 || cap > (1 << 20)) return set_error(MPCX_E_ARG, "inbox capacity must be 0..1048576");
  if (!b->split) return set_error(MPCX_E_UNSUPPORTED, "FP64 batches have no inboxes");
  HIPCHK(hipStreamSynchronize(b->st));
  cob_free_graph(b);
  b->icap_set = cap;
  return cob_tiles(b);
}

// re-sort period of a split batch in MD steps (a positive multiple of
// MPCtstep): split storage sums cells in exact integers and keys every draw
// by cell, so the particle order only sets the memory locality (and the order
// of the near engines' force sums); FP64 batches keep the engines' period,
// the reference's summation order follows it
extern "C" int mpcx_cob_set_sort_step(mpcx_cob* b, int32_t steps) {
  if (!b || steps <= 0 || steps % b->S != 0) return set_error(MPCX_E_ARG, "the re-sort period must be a multiple of MPCtstep");
  if (!b->split) return set_error(MPCX_E_UNSUPPORTED, "FP64 batches re-sort at their engines' sort_step");
  b->sort_step = steps;
  return 0;
}

// re-tile the particle kernels (the autotuner's knob); drops a captured graph
extern "C" int mpcx_cob_set_tile(mpcx_cob* b, int32_t tile) {
  if (!b || tile < 256 || tile % 256 != 0) return set_error(MPCX_E_ARG, "tile must be a positive multiple of 256");
  HIPCHK(hipStreamSynchronize(b->st));
  cob_free_graph(b);
  b->tile = tile;
  return cob_tiles(b);
}

// A split-precision engine's run: a batch of one on the engine's streams
static int co_run_split(mpcx_co* e, int64_t ncoll) {
  if (!e->solo) {
    mpcx_cob* b = nullptr;
    int rc = cob_make(&b, 1, &e, 2048, true);
    if (rc) return rc;
    e->solo = b;
  }
  mpcx_cob* b = e->solo;
  b->overlap = e->near_mode == 0 && !e->clustered && e->overlap;
  int rc = cob_run(b, ncoll, 0);
  e->run_ms = b->run_ms;
  e->kt_n = b->kt_n;
  for (int k = 0; k < mpcx_co::NKT && k < mpcx_cob::NKT; k++) e->kt_ms[k] = b->kt_ms[k];
  e->near_ms = b->kt_ms[0];
  e->bulk_ms = b->kt_ms[1] + b->kt_ms[2] + b->kt_ms[3];
  return rc;
}

// ---------------------------------------------------------------- batch diagnostics
// Per replica (blockIdx.y): counts of B (flag 1) particles per (z, r) bin in
// the body frame of its dimer 0 (origin at the C sphere, z along the N->C unit
// vector), added to acc[r][nz][nr]; and its dimer's centre-of-mass velocity
// along that axis, added to vacc[r] (one thread).
namespace mpcx {
struct BodyGrid {
  int nr, nz;
  double rmax, zmin, zmax;
};
// Ss (split storage) or null (FP64 storage)
__global__ __launch_bounds__(256) void body_acc_b(const CoP* __restrict__ Ps, const CoD* __restrict__ Ds,
                                                  const SpD* __restrict__ Ss, BodyGrid G, unsigned* acc,
                                                  double* vacc) {
  const int r = blockIdx.y;
  const CoP P = Ps[r];
  const CoD D = Ds[r];
  const int nd = P.nd;
  const double cx = D.dim[DX * MAXSG + 0], cy = D.dim[DY * MAXSG + 0], cz = D.dim[DZ * MAXSG + 0];
  double ux = cx - D.dim[DX * MAXSG + nd], uy = cy - D.dim[DY * MAXSG + nd], uz = cz - D.dim[DZ * MAXSG + nd];
  ux = pbc_half(ux, P.box[0]);
  uy = pbc_half(uy, P.box[1]);
  const double len = sqrt(ux * ux + uy * uy + uz * uz);
  ux = ux / len; uy = uy / len; uz = uz / len;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double M = P.Mc + P.Mn;
    const double vx = (P.Mc * D.dim[DVX * MAXSG + 0] + P.Mn * D.dim[DVX * MAXSG + nd]) / M;
    const double vy = (P.Mc * D.dim[DVY * MAXSG + 0] + P.Mn * D.dim[DVY * MAXSG + nd]) / M;
    const double vz = (P.Mc * D.dim[DVZ * MAXSG + 0] + P.Mn * D.dim[DVZ * MAXSG + nd]) / M;
    vacc[r] += (vx * ux + vy * uy) + vz * uz;
  }
  const double dr = G.rmax / G.nr, dz = (G.zmax - G.zmin) / G.nz;
  unsigned* a = acc + (size_t)r * G.nz * G.nr;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < P.NT; i += (long long)gridDim.x * 256) {
    double x, y, z;
    if (Ss) {
      const SpD S = Ss[r];
      if (S.flag[i] != 1) continue;
      int ix, iy, iz;
      sp_cell_xyz(P, S.cell[i], ix, iy, iz);
      x = (double)ix + (double)S.ox[i]; y = (double)iy + (double)S.oy[i]; z = (double)iz + (double)S.oz[i];
    } else {
      if (D.flag[i] != 1) continue;
      x = D.x[i]; y = D.y[i]; z = D.z[i];
    }
    const double ax = pbc_half(x - cx, P.box[0]), ay = pbc_half(y - cy, P.box[1]), az = z - cz;
    const double zr = (ax * ux + ay * uy) + az * uz;
    const double px = ax - zr * ux, py = ay - zr * uy, pz = az - zr * uz;
    const double rr = sqrt((px * px + py * py) + pz * pz);
    if (!(zr >= G.zmin && zr < G.zmax && rr < G.rmax)) continue;
    const int iz = min((int)floor((zr - G.zmin) / dz), G.nz - 1), ir = min((int)floor(rr / dr), G.nr - 1);
    atomicAdd(&a[iz * G.nr + ir], 1u);
  }
}
}  // namespace mpcx

// Accumulate the body-frame B density and propulsion velocity of every
// replica over n intervals (one sample after each): counts (R, nz, nr) and
// velocity sums (R) are added to the host arrays.
extern "C" int mpcx_cob_body_average(mpcx_cob* b, int64_t n, int32_t nr, int32_t nz, double rmax, double zmin,
                                     double zmax, uint32_t* counts, double* vsum) {
  if (!b || n < 0 || nr < 1 || nz < 1 || !(rmax > 0) || !(zmax > zmin) || !counts || !vsum)
    return set_error(MPCX_E_ARG, "bad arguments");
  const size_t bins = (size_t)b->R * nz * nr;
  unsigned* acc = nullptr;
  double* vacc = nullptr;
  int rc = 0;
  if (hipMalloc(&acc, bins * 4) != hipSuccess || hipMalloc(&vacc, b->R * 8) != hipSuccess ||
      hipMemcpy(acc, counts, bins * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(vacc, vsum, b->R * 8, hipMemcpyHostToDevice) != hipSuccess)
    rc = set_error(MPCX_E_HIP, "body accumulator allocation failed");
  BodyGrid G{nr, nz, rmax, zmin, zmax};
  int cstep = 0;
  if (!rc) rc = cob_begin(b, &cstep);
  for (int64_t i = 0; i < n && !rc; i++) {
    rc = cob_step(b, &cstep, 1);
    if (!rc)
      hipLaunchKernelGGL(body_acc_b, dim3(64, b->R), dim3(256), 0, b->st, b->d_P, b->d_D,
                         (const SpD*)(b->split ? b->d_S : nullptr), G, acc, vacc);
  }
  if (!rc) rc = cob_end(b);
  if (rc) cob_mark_dirty(b);
  if (!rc) {
    if (hipMemcpy(counts, acc, bins * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(vsum, vacc, b->R * 8, hipMemcpyDeviceToHost) != hipSuccess)
      rc = set_error(MPCX_E_HIP, "hipMemcpy failed");
  }
  (void)hipFree(acc);
  (void)hipFree(vacc);
  return rc;
}
