/*
 * CPU restatement of the reference MD + MPC hot path -- TEST INFRASTRUCTURE ONLY.
 * See mpc_oracle.h.  Compiled with -O2 -ffp-contract=off; Fortran 1-based
 * indices appear as "i - 1" offsets and as stored values.
 */
#include "mpc_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IX2(k, i, nd) ((int64_t)(k) + (int64_t)(i) * (nd))                      /* (k,i), 0-based */
#define IX3(k, i, j, nd) ((int64_t)(k) + (int64_t)(i) * (nd) + (int64_t)(j) * 2 * (nd))

static const double PI = 3.141592653589793238462643383279502884; /* 4*atan(1) */

/* ===================== MT19937, mtmod.f90 ===================== */
void or_mtmod_sgrnd(or_mtmod_t *s, int32_t seed) {
  /* mtmod.f90:80-83: mt(0)=seed; mt(i)=69069*mt(i-1) in 32-bit wrap-around */
  s->mt[0] = (uint32_t)seed;
  for (int i = 1; i < 624; i++) s->mt[i] = 69069u * s->mt[i - 1];
  s->mti = 624; /* value the do-loop variable holds on exit */
}

static void mt_twist(uint32_t *mt) {
  /* mtmod.f90:117-126 (identical recurrence in mt95.f90:466-481) */
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

static inline uint32_t mt_temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

double or_mtmod_grnd(or_mtmod_t *s) {
  /* mtmod.f90:115-141; one rounding: u/(2^32-1) */
  if (s->mti >= 624) { mt_twist(s->mt); s->mti = 0; }
  uint32_t y = mt_temper(s->mt[s->mti++]);
  return (double)y / 4294967295.0;
}

void or_mtmod_fill(or_mtmod_t *s, double *out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = or_mtmod_grnd(s);
}

/* ===================== MT19937, mt95.f90 ===================== */
void or_mt95_init(or_mt95_t *s, int32_t seed) {
  /* mt95.f90:378-404 (init_genrand) */
  s->val[0] = (uint32_t)seed;
  for (int i = 1; i < 624; i++)
    s->val[i] = 1812433253u * (s->val[i - 1] ^ (s->val[i - 1] >> 30)) + (uint32_t)i;
  s->cnt = 625;
}

uint32_t or_mt95_int32(or_mt95_t *s) {
  /* mt95.f90:583-602: next_state when cnt > n, then temper val(cnt) */
  if (s->cnt > 624) { mt_twist(s->val); s->cnt = 1; }
  return mt_temper(s->val[s->cnt++ - 1]);
}

double or_mt95_real1(or_mt95_t *s) {
  /* mt95.f90:836-837: real(a)*pi232_1 + p231d232_1 with a the signed word,
   * two roundings (mul, then add). */
  const double pi232_1 = 1.0 / 4294967295.0;
  const double p231d232_1 = 2147483648.0 / 4294967295.0;
  int32_t a = (int32_t)or_mt95_int32(s);
  double m = (double)a * pi232_1;
  return m + p231d232_1;
}

void or_mt95_fill(or_mt95_t *s, double *out, int64_t n) {
  for (int64_t i = 0; i < n; i++) out[i] = or_mt95_real1(s);
}

/* ===================== collective_code ===================== */
static inline double pbc_half(double d, double L) {
  /* "if (d > 0.5d0*L) d = d - L; else if (d < -0.5d0*L) d = d + L" */
  if (d > 0.5 * L) return d - L;
  if (d < -0.5 * L) return d + L;
  return d;
}

void or_co_neighbour_list(int32_t NT, const double *box, double cutoff, double skin,
                          const double *x, const double *y, const double *z,
                          double xd, double yd, double zd,
                          int64_t *num, int64_t *nb, int64_t stride) {
  /* neighbour_list_mod.f95:34-57; caller zeroes *num (evolve_mod.f95:536) */
  double full = cutoff + skin;
  double full2 = full * full;
  for (int64_t i = 0; i < NT; i++) {
    double d1 = pbc_half(x[i] - xd, box[0]);
    double d2 = pbc_half(y[i] - yd, box[1]);
    double d3 = z[i] - zd;
    if (d1 * d1 + d2 * d2 + d3 * d3 <= full2) {
      *num += 1;
      nb[(*num - 1) * stride] = i + 1;
    }
  }
}

void or_co_inout(int32_t NT, const double *box, double cutC, double cutN, double skin,
                 const double *x, const double *y, const double *z,
                 const double *xd, const double *yd, const double *zd, int32_t numdims,
                 int64_t *out_num, int64_t *out_list, int64_t *in_num, int64_t *in_list) {
  /* neighbour_list_mod.f95:105-177 at one OpenMP thread: lists in ascending
   * particle order, appended to the caller's (zeroed) counts. */
  double fc = cutC + skin, fn = cutN + skin;
  double fc2 = fc * fc, fn2 = fn * fn;
  for (int64_t i = 0; i < NT; i++) {
    int bulk = 0;
    for (int m = 0; m < numdims; m++) {
      double c1 = pbc_half(x[i] - xd[IX2(m, 0, numdims)], box[0]);
      double c2 = pbc_half(y[i] - yd[IX2(m, 0, numdims)], box[1]);
      double c3 = z[i] - zd[IX2(m, 0, numdims)];
      double n1 = pbc_half(x[i] - xd[IX2(m, 1, numdims)], box[0]);
      double n2 = pbc_half(y[i] - yd[IX2(m, 1, numdims)], box[1]);
      double n3 = z[i] - zd[IX2(m, 1, numdims)];
      if (c1 * c1 + c2 * c2 + c3 * c3 <= fc2 || n1 * n1 + n2 * n2 + n3 * n3 <= fn2) bulk = 1;
    }
    if (bulk == 0) out_list[(*out_num)++] = i + 1;
    else in_list[(*in_num)++] = i + 1;
  }
}

static inline void lj_terms(double dRsq, double R6, double R12, double r12R12, double r6R6,
                            double *Fbrack, double *Ubrack) {
  /* forces_mod.f95:82-88 */
  double IdRsq = 1.0 / dRsq;
  double IdR6 = (IdRsq * IdRsq) * IdRsq;
  double IdR8 = IdR6 * IdRsq;
  double IdR12 = IdR6 * IdR6;
  double IdR14 = IdR12 * IdRsq;
  *Fbrack = 4.0 * (r12R12 * IdR14 - r6R6 * IdR8);
  *Ubrack = 4.0 * (R12 * IdR12 - R6 * IdR6) + 1.0;
}

void or_co_lj_mot_sol(int32_t sphere, const double *box, int32_t *flag,
                      const double *x, const double *y, const double *z,
                      double xd, double yd, double zd,
                      double *fpx, double *fpy, double *fpz, double *pot,
                      double RPab, double RPba, double cut2, double Ea, double Eb, double Es,
                      double R6, double R12, double r12R12, double r6R6,
                      int64_t num, const int64_t *nb, int64_t stride) {
  (void)RPab; (void)RPba; /* forces_mod.f95:30-31: received, not used */
  for (int64_t i = 0; i < num; i++) {
    int64_t p = nb[i * stride] - 1;
    double d1 = pbc_half(x[p] - xd, box[0]);
    double d2 = pbc_half(y[p] - yd, box[1]);
    double d3 = z[p] - zd;
    double dRsq = d1 * d1 + d2 * d2 + d3 * d3;
    if (!(dRsq <= cut2)) continue;
    double E;
    if (flag[p] == 0) {
      if (sphere == 1) { flag[p] = 1; E = Eb; }          /* A -> B at the catalytic sphere */
      else if (sphere == 2) E = Ea;
      else continue;
    } else if (flag[p] == 1) E = Eb;
    else if (flag[p] == 2) E = Es;
    else continue;
    double F, U;
    lj_terms(dRsq, R6, R12, r12R12, r6R6, &F, &U);
    double EF = E * F;
    fpx[i * stride] = EF * d1;
    fpy[i * stride] = EF * d2;
    fpz[i * stride] = EF * d3;
    pot[i * stride] = pot[i * stride] + E * U;
  }
}

static void lj_mot_mot(const double *box, double x1, double y1, double z1,
                       double x2, double y2, double z2, double *fx, double *fy, double *fz,
                       double *fp, double cut2, double Em, double R6, double R12,
                       double r6R6, double r12R12) {
  /* forces_mod.f95:183-208 */
  double d1 = pbc_half(x1 - x2, box[0]);
  double d2 = pbc_half(y1 - y2, box[1]);
  double d3 = z1 - z2;
  double dRsq = d1 * d1 + d2 * d2 + d3 * d3;
  if (dRsq <= cut2) {
    double F, U;
    lj_terms(dRsq, R6, R12, r12R12, r6R6, &F, &U);
    double EF = Em * F;
    *fx = EF * d1; *fy = EF * d2; *fz = EF * d3;
    *fp = Em * U;
  }
}

static void lj_mot_wall(const double *box, double zd, double *fwallz, double Ew, double R3,
                        double R9, double r9R9, double r3R3, double cut2, double *potential) {
  /* forces_mod.f95:303-333 */
  double pref = 1.5 * sqrt(3.0);           /* 0.5*3. is a single-precision 1.5 */
  double dR = box[2] - zd;
  if (dR * dR < cut2) {
    double IdRsq = 1.0 / (dR * dR);
    double IdR3 = pow(IdRsq, 1.5);
    double IdR9 = (IdR3 * IdR3) * IdR3;
    double IdR5 = IdR3 * IdRsq;
    double IdR11 = IdR9 * IdRsq;
    double Fb = (pref * Ew) * (r9R9 * IdR11 - r3R3 * IdR5);
    *fwallz = Fb * dR;
    *potential = (pref * Ew) * (R9 * IdR9 - R3 * IdR3) + 1.0;
  }
  dR = 0.0 - zd;
  if (dR * dR < cut2) {
    double IdRsq = 1.0 / (dR * dR);
    double IdR3 = pow(IdRsq, 1.5);
    double IdR9 = (IdR3 * IdR3) * IdR3;
    double IdR5 = IdR3 * IdRsq;
    double IdR11 = IdR9 * IdRsq;
    double Fb = (pref * Ew) * (r9R9 * IdR11 - r3R3 * IdR5);
    *fwallz = *fwallz + Fb * dR;
    *potential = *potential + (pref * Ew) * (R9 * IdR9 - R3 * IdR3) + 1.0;
  }
}

void or_co_reactive_grid_setup(const double *box, int32_t *rb, int32_t type) {
  /* reactive_grid_mod.f95:25-97; literal single-precision constants kept */
  int32_t n = (int32_t)trunc(box[0] * box[1]);
  int32_t ib1 = (int32_t)box[0];
  memset(rb, 0, sizeof(int32_t) * (size_t)n);
  for (int32_t i = 1; i <= n; i++) {
    int on = 0;
    if (type == 1) {
      on = (i % ib1) > 0 && (i % ib1) <= (int32_t)(0.5 * box[0]);
    } else if (type == 2) {
      int32_t f = (int32_t)floor((double)i / box[1]);
      on = ((i + f) % 2) == 0;
    } else if (type == 3) {
      double a = 0.5 * box[0] - (double)((i - 1) % ib1) + 0.5;
      double b = 0.5 * box[1] - (double)(int32_t)floor((double)i / box[0]) + 0.5;
      on = a * a + b * b < 400.0;
    } else if (type == 4) {
      const double s6 = (double)(1.0f / 6.0f), f6 = (double)(5.0f / 6.0f);
      double xm = (double)((i - 1) % ib1), ym = (double)(int32_t)floor((double)i / box[0]);
      double a1 = s6 * box[0] - xm + 0.5, a2 = 0.5 * box[0] - xm + 0.5, a3 = f6 * box[0] - xm + 0.5;
      double b1 = 0.25 * box[1] - ym + 0.5, b2 = 0.75 * box[1] - ym + 0.5;
      on = a1 * a1 + b1 * b1 < 100.0 && a1 * a1 + b2 * b2 < 100.0 && a2 * a2 + b1 * b1 < 100.0 &&
           a2 * a2 + b2 * b2 < 100.0 && a3 * a3 + b1 * b1 < 100.0 && a3 * a3 + b2 * b2 < 100.0;
    } else if (type == 5) {
      int32_t lo = (int32_t)floor(0.5 * box[0] - 4.0), hi = (int32_t)floor(0.5 * box[0] + 4.0);
      on = (i % ib1) > lo && (i % ib1) <= hi;
    } else if (type == 6) {
      on = 1;
    }
    if (on) rb[i - 1] = 1;
  }
}

void or_co_reactive_grid(const int32_t *rb, int32_t react_dir, const double *box,
                         double x, double y, double vx, double vy, int32_t *flag, double t) {
  /* reactive_grid_mod.f95:128-149 */
  double ax = x - vx * t, ay = y - vy * t;
  if (ax >= box[0]) ax = ax - box[0];
  else if (ax < 0.0) ax = ax + box[0];
  if (ay >= box[1]) ay = ay - box[1];
  else if (ay < 0.0) ay = ay + box[1];
  int32_t hit = (int32_t)((double)(int32_t)floor(ax) + box[0] * (double)(int32_t)floor(ay) + 1.0);
  if (rb[hit - 1] == 1) {
    if (react_dir == 0 && *flag == 0) *flag = 1;
    else if (react_dir == 1 && *flag == 1) *flag = 0;
  }
}

static inline int64_t cell_unshifted(double x, double y, double z, const double *box) {
  /* int(floor(x) + box1*floor(y) + box1*box2*floor(z) + 1), as in
   * sort_hilbert_mod.f95:51 and distribute_mod.f95:100 */
  double c = (double)(int32_t)floor(x) + box[0] * (double)(int32_t)floor(y);
  c = c + (box[0] * box[1]) * (double)(int32_t)floor(z);
  c = c + 1.0;
  return (int64_t)c;
}

void or_co_sort_hilbert(int64_t vol, int32_t NT, const double *box, const int64_t *tour,
                        double *x, double *y, double *z, int32_t *flag,
                        double *vx, double *vy, double *vz) {
  /* sort_hilbert_mod.f95:47-92: bin by unshifted cell (ascending index within
   * a cell), then emit cells in tour order.  The tour is the list of 1-based
   * cell numbers the reference builds from hilbert.txt (:57-62). */
  int64_t *cnt = calloc((size_t)vol + 1, sizeof(int64_t));
  int64_t *cell = malloc(sizeof(int64_t) * (size_t)NT);
  for (int64_t i = 0; i < NT; i++) {
    cell[i] = cell_unshifted(x[i], y[i], z[i], box);
    if (cell[i] < 1 || cell[i] > vol) cell[i] = 0; /* out of the box: never emitted */
    cnt[cell[i]]++;
  }
  double *nx = malloc(sizeof(double) * (size_t)NT * 6);
  int32_t *nf = malloc(sizeof(int32_t) * (size_t)NT);
  /* slots the tour does not fill keep their old contents (the reference copies
   * uninitialised memory there; a tour covering every cell once fills all). */
  for (int64_t i = 0; i < NT; i++) {
    nx[i] = x[i]; nx[NT + i] = y[i]; nx[2 * NT + i] = z[i];
    nx[3 * NT + i] = vx[i]; nx[4 * NT + i] = vy[i]; nx[5 * NT + i] = vz[i]; nf[i] = flag[i];
  }
  /* replay the reference loop literally (a cell visited twice is emitted twice) */
  int64_t **members = calloc((size_t)vol + 1, sizeof(int64_t *));
  int64_t *fill = calloc((size_t)vol + 1, sizeof(int64_t));
  for (int64_t i = 0; i < NT; i++) {
    int64_t c = cell[i];
    if (!members[c]) members[c] = malloc(sizeof(int64_t) * (size_t)cnt[c]);
    members[c][fill[c]++] = i;
  }
  int64_t nw = 0;
  for (int64_t t = 0; t < vol; t++) {
    int64_t c = tour[t];
    if (c < 1 || c > vol) continue;
    for (int64_t j = 0; j < cnt[c]; j++) {
      int64_t src = members[c][j];
      nx[nw] = x[src]; nx[NT + nw] = y[src]; nx[2 * NT + nw] = z[src];
      nx[3 * NT + nw] = vx[src]; nx[4 * NT + nw] = vy[src]; nx[5 * NT + nw] = vz[src];
      nf[nw] = flag[src];
      nw++;
    }
  }
  for (int64_t i = 0; i < NT; i++) {
    x[i] = nx[i]; y[i] = nx[NT + i]; z[i] = nx[2 * NT + i];
    vx[i] = nx[3 * NT + i]; vy[i] = nx[4 * NT + i]; vz[i] = nx[5 * NT + i]; flag[i] = nf[i];
  }
  for (int64_t c = 0; c <= vol; c++) free(members[c]);
  free(members); free(fill); free(nx); free(nf); free(cell); free(cnt);
}

/* lazy / inside streaming of one particle, evolve_mod.f95:259-312 (dtc = MDt
 * for inside particles; the outside form multiplies (v*MDt)*time_count) */
static inline void stream_one(const or_co_params *p, or_co_state *s, int64_t n, int outside,
                              double tc) {
  const double *box = p->box;
  double MDt = p->MDt;
  double x = s->x[n], y = s->y[n], z = s->z[n];
  double vx = s->vx[n], vy = s->vy[n], vz = s->vz[n];
  if (outside) x = x + (vx * MDt) * tc; else x = x + vx * MDt;
  if (x >= box[0]) x = x - box[0]; else if (x < 0.0) x = x + box[0];
  if (outside) y = y + (vy * MDt) * tc; else y = y + vy * MDt;
  if (y >= box[1]) y = y - box[1]; else if (y < 0.0) y = y + box[1];
  if (outside) z = z + (vz * MDt) * tc; else z = z + vz * MDt;
  if (z < 0.0) {
    double t = z / vz;
    or_co_reactive_grid(s->react_box, p->react_dir, box, x, y, vx, vy, &s->flag[n], t);
    vx = -vx; vy = -vy; vz = -vz;
    x = x + (2.0 * vx) * t; y = y + (2.0 * vy) * t; z = z + (2.0 * vz) * t;
    if (x >= box[0]) x = x - box[0]; else if (x < 0.0) x = x + box[0];
    if (y >= box[1]) y = y - box[1]; else if (y < 0.0) y = y + box[1];
  } else if (z > box[2]) {
    double t = (z - box[2]) / vz;
    vx = -vx; vy = -vy; vz = -vz;
    x = x + (2.0 * vx) * t; y = y + (2.0 * vy) * t; z = z + (2.0 * vz) * t;
    if (x >= box[0]) x = x - box[0]; else if (x < 0.0) x = x + box[0];
    if (y >= box[1]) y = y - box[1]; else if (y < 0.0) y = y + box[1];
  }
  s->x[n] = x; s->y[n] = y; s->z[n] = z;
  s->vx[n] = vx; s->vy[n] = vy; s->vz[n] = vz;
}

static void kick_lists(const or_co_params *p, or_co_state *s) {
  /* evolve_mod.f95:225-250 (and :645-670) */
  int nd = p->numdims;
  double hA = (0.5 * p->MDt) / p->Ma, hB = (0.5 * p->MDt) / p->Mb, hS = (0.5 * p->MDt) / p->Ms;
  for (int k = 0; k < nd; k++)
    for (int i = 0; i < 2; i++) {
      int64_t num = s->nb_num[IX2(k, i, nd)];
      for (int64_t j = 0; j < num; j++) {
        int64_t q = s->nb[IX3(k, i, j, nd)] - 1;
        double h;
        if (s->flag[q] == 0) h = hA;
        else if (s->flag[q] == 1) h = hB;
        else if (s->flag[q] == 2) h = hS;
        else continue;
        s->vx[q] = s->vx[q] + h * s->fpx[IX3(k, i, j, nd)];
        s->vy[q] = s->vy[q] + h * s->fpy[IX3(k, i, j, nd)];
        s->vz[q] = s->vz[q] + h * s->fpz[IX3(k, i, j, nd)];
      }
    }
}

static void rebuild_lists(const or_co_params *p, or_co_state *s) {
  /* evolve_mod.f95:536-553 at one thread */
  int nd = p->numdims;
  double cutC = p->Rc * p->cutoff, cutN = p->Rn * p->cutoff;
  for (int k = 0; k < nd; k++)
    for (int i = 0; i < 2; i++) s->nb_num[IX2(k, i, nd)] = 0;
  s->out_num = 0;
  s->in_num = 0;
  for (int k = 0; k < nd; k++) {
    or_co_neighbour_list(p->NT, p->box, cutC, p->skinrad, s->x, s->y, s->z,
                         s->xd[IX2(k, 0, nd)], s->yd[IX2(k, 0, nd)], s->zd[IX2(k, 0, nd)],
                         &s->nb_num[IX2(k, 0, nd)], &s->nb[IX3(k, 0, 0, nd)], 2 * nd);
    or_co_neighbour_list(p->NT, p->box, cutN, p->skinrad, s->x, s->y, s->z,
                         s->xd[IX2(k, 1, nd)], s->yd[IX2(k, 1, nd)], s->zd[IX2(k, 1, nd)],
                         &s->nb_num[IX2(k, 1, nd)], &s->nb[IX3(k, 1, 0, nd)], 2 * nd);
  }
  or_co_inout(p->NT, p->box, cutC, cutN, p->skinrad, s->x, s->y, s->z, s->xd, s->yd, s->zd,
              nd, &s->out_num, s->out_list, &s->in_num, s->in_list);
}

static int32_t penetrate_steps(const or_co_params *p, const or_co_state *s) {
  /* evolve_mod.f95:555-563 */
  double maxt = s->vx[0] * s->vx[0] + s->vy[0] * s->vy[0] + s->vz[0] * s->vz[0];
  for (int64_t i = 0; i < p->NT; i++) {
    double v = s->vx[i] * s->vx[i] + s->vy[i] * s->vy[i] + s->vz[i] * s->vz[i];
    if (v > maxt) maxt = v;
  }
  return (int32_t)floor(((0.5 * p->skinrad) / sqrt(maxt)) / p->MDt);
}

void or_co_setup_lists(const or_co_params *p, or_co_state *s, int do_sort) {
  /* main_pattern.f95:170-188 */
  if (do_sort)
    or_co_sort_hilbert(p->vol, p->NT, p->box, s->tour, s->x, s->y, s->z, s->flag,
                       s->vx, s->vy, s->vz);
  rebuild_lists(p, s);
  for (int64_t i = 0; i < p->NT; i++) s->msd[i] = 0.0;
  s->time_penetrate = penetrate_steps(p, s);
  s->time_count = 0;
}

/* forces_mod.f95 calls of evolve_mod.f95:567-602 and restart_mod.f95:467-475 (their
 * order and arguments are the same): solvent pair forces over the lists, the
 * motor-motor terms into Fm (FmotX,Y,Z,pot (k,i,j,l)) and the wall terms. */
static void co_pair_terms(const or_co_params *p, or_co_state *s, double *Fm, double *fwz, double *wpot) {
  const int nd = p->numdims;
  const double cutC = p->Rc * p->cutoff, cutN = p->Rn * p->cutoff;
  const double cutC2 = cutC * cutC, cutN2 = cutN * cutN;
  const double Rn6 = pow(p->Rn, 6), Rn12 = pow(p->Rn, 12);
  const double Rc6 = pow(p->Rc, 6), Rc12 = pow(p->Rc, 12);
  const double r6Rn6 = 6.0 * Rn6, r12Rn12 = 12.0 * Rn12, r6Rc6 = 6.0 * Rc6, r12Rc12 = 12.0 * Rc12;
  const double dcc = p->Rc + p->Rc + p->Rm_delta, dcn = p->Rc + p->Rn + p->Rm_delta,
               dnn = p->Rn + p->Rn + p->Rm_delta;
  const double Rcc6 = pow(dcc, 6), Rcc12 = pow(dcc, 12), Rcn6 = pow(dcn, 6), Rcn12 = pow(dcn, 12);
  const double Rnn6 = pow(dnn, 6), Rnn12 = pow(dnn, 12);
  const double mcc2 = (dcc * p->cutoff) * (dcc * p->cutoff), mcn2 = (dcn * p->cutoff) * (dcn * p->cutoff),
               mnn2 = (dnn * p->cutoff) * (dnn * p->cutoff);
  const double wc = p->Rw + p->Rc, wn = p->Rw + p->Rn;
  const double Rwc3 = pow(wc, 3), Rwc9 = pow(wc, 9), Rwn3 = pow(wn, 3), Rwn9 = pow(wn, 9);
  const double r3Rwc3 = 3.0 * Rwc3, r9Rwc9 = 9.0 * Rwc9, r3Rwn3 = 3.0 * Rwn3, r9Rwn9 = 9.0 * Rwn9;
  const double w6 = pow(3.0, (double)(1.0f / 6.0f));
  const double wcC2 = (wc * w6) * (wc * w6), wcN2 = (wn * w6) * (wn * w6);
  const double *box = p->box;
  int64_t nd4 = (int64_t)nd * 2 * nd * 2;
#define FM(c, k, i, j, l) Fm[(int64_t)(c) * nd4 + (k) + (int64_t)(i) * nd + (int64_t)(j) * 2 * nd + (int64_t)(l) * 2 * nd * nd]
  for (int k = 0; k < nd; k++) {                                        /* :567-602 */
    int64_t o0 = IX2(k, 0, nd), o1 = IX2(k, 1, nd);
    or_co_lj_mot_sol(1, box, s->flag, s->x, s->y, s->z, s->xd[o0], s->yd[o0], s->zd[o0],
                     &s->fpx[IX3(k, 0, 0, nd)], &s->fpy[IX3(k, 0, 0, nd)], &s->fpz[IX3(k, 0, 0, nd)],
                     &s->pot[IX3(k, 0, 0, nd)], p->RPab, p->RPba, cutC2, p->EaC, p->EbC, p->EsC,
                     Rc6, Rc12, r12Rc12, r6Rc6, s->nb_num[o0], &s->nb[IX3(k, 0, 0, nd)], 2 * nd);
    or_co_lj_mot_sol(2, box, s->flag, s->x, s->y, s->z, s->xd[o1], s->yd[o1], s->zd[o1],
                     &s->fpx[IX3(k, 1, 0, nd)], &s->fpy[IX3(k, 1, 0, nd)], &s->fpz[IX3(k, 1, 0, nd)],
                     &s->pot[IX3(k, 1, 0, nd)], p->RPab, p->RPba, cutN2, p->EaN, p->EbN, p->EsN,
                     Rn6, Rn12, r12Rn12, r6Rn6, s->nb_num[o1], &s->nb[IX3(k, 1, 0, nd)], 2 * nd);
    for (int j = 0; j < nd; j++) {
      if (j == k) continue;
      int64_t j0 = IX2(j, 0, nd), j1 = IX2(j, 1, nd);
      lj_mot_mot(box, s->xd[o0], s->yd[o0], s->zd[o0], s->xd[j0], s->yd[j0], s->zd[j0],
                 &FM(0, k, 0, j, 0), &FM(1, k, 0, j, 0), &FM(2, k, 0, j, 0), &FM(3, k, 0, j, 0),
                 mcc2, p->Em, Rcc6, Rcc12, 6.0 * Rcc6, 12.0 * Rcc12);
      lj_mot_mot(box, s->xd[o0], s->yd[o0], s->zd[o0], s->xd[j1], s->yd[j1], s->zd[j1],
                 &FM(0, k, 0, j, 1), &FM(1, k, 0, j, 1), &FM(2, k, 0, j, 1), &FM(3, k, 0, j, 1),
                 mcn2, p->Em, Rcn6, Rcn12, 6.0 * Rcn6, 12.0 * Rcn12);
      lj_mot_mot(box, s->xd[o1], s->yd[o1], s->zd[o1], s->xd[j0], s->yd[j0], s->zd[j0],
                 &FM(0, k, 1, j, 0), &FM(1, k, 1, j, 0), &FM(2, k, 1, j, 0), &FM(3, k, 1, j, 0),
                 mcn2, p->Em, Rcn6, Rcn12, 6.0 * Rcn6, 12.0 * Rcn12);
      lj_mot_mot(box, s->xd[o1], s->yd[o1], s->zd[o1], s->xd[j1], s->yd[j1], s->zd[j1],
                 &FM(0, k, 1, j, 1), &FM(1, k, 1, j, 1), &FM(2, k, 1, j, 1), &FM(3, k, 1, j, 1),
                 mnn2, p->Em, Rnn6, Rnn12, 6.0 * Rnn6, 12.0 * Rnn12);
    }
    /* evolve_mod.f95:597-600 passes (R3,R9) then (3R^3, 9R^9) into the dummies
     * named (r9R9, r3R3): the swap is part of the reference arithmetic. */
    lj_mot_wall(box, s->zd[o0], &fwz[o0], p->Ew, Rwc3, Rwc9, r3Rwc3, r9Rwc9, wcC2, &wpot[o0]);
    lj_mot_wall(box, s->zd[o1], &fwz[o1], p->Ew, Rwn3, Rwn9, r3Rwn3, r9Rwn9, wcN2, &wpot[o1]);
  }
#undef FM
}

void or_co_restart_forces(const or_co_params *p, or_co_state *s) {
  /* restart_mod.f95:429-495, after the lists of or_co_setup_lists: every
   * force, pair force and potential is zeroed and evaluated again from the
   * restored positions (A particles inside a catalytic cutoff turn B on the
   * way, as in any force evaluation); TPotential is summed in the restart's
   * running order */
  const int nd = p->numdims;
  double fwz[2 * 64], wpot[2 * 64];
  if (nd > 64) abort();
  double *Fm = calloc((size_t)(4 * nd * 2 * nd * 2), sizeof(double));
  int64_t nd4 = (int64_t)nd * 2 * nd * 2;
#define FM(c, k, i, j, l) Fm[(int64_t)(c) * nd4 + (k) + (int64_t)(i) * nd + (int64_t)(j) * 2 * nd + (int64_t)(l) * 2 * nd * nd]
  for (int k = 0; k < nd; k++)
    for (int i = 0; i < 2; i++) {
      int64_t o = IX2(k, i, nd);
      s->fxd[o] = 0.0; s->fyd[o] = 0.0; s->fzd[o] = 0.0;
      fwz[o] = 0.0; wpot[o] = 0.0;
      for (int64_t j = 0; j < s->max_nb; j++) {
        int64_t q = IX3(k, i, j, nd);
        s->fpx[q] = 0.0; s->fpy[q] = 0.0; s->fpz[q] = 0.0; s->pot[q] = 0.0;
      }
    }
  s->tpotential = 0.0;
  co_pair_terms(p, s, Fm, fwz, wpot);
  for (int k = 0; k < nd; k++)
    for (int i = 0; i < 2; i++) {
      int64_t o = IX2(k, i, nd);
      for (int64_t j = 0; j < s->nb_num[o]; j++) {
        int64_t q = IX3(k, i, j, nd);
        s->tpotential = s->tpotential + s->pot[q];
        s->fxd[o] = s->fxd[o] - s->fpx[q];
        s->fyd[o] = s->fyd[o] - s->fpy[q];
        s->fzd[o] = s->fzd[o] - s->fpz[q];
      }
      for (int j = 0; j < nd; j++)
        for (int l = 0; l < 2; l++) {
          s->fxd[o] = s->fxd[o] + FM(0, k, i, j, l);
          s->fyd[o] = s->fyd[o] + FM(1, k, i, j, l);
          s->fzd[o] = s->fzd[o] + FM(2, k, i, j, l);
          if (j < k) s->tpotential = s->tpotential + FM(3, k, i, j, l);
        }
      s->tpotential = s->tpotential + wpot[o];
      s->fzd[o] = s->fzd[o] + fwz[o];
    }
#undef FM
  free(Fm);
}

void or_co_evolve(const or_co_params *p, or_co_state *s) {
  const int nd = p->numdims;
  const double MDt = p->MDt;
  const double hc = (0.5 * MDt) / p->Mc, hn = (0.5 * MDt) / p->Mn;
  const double IIMcIMn = 1.0 / ((1.0 / p->Mc) + (1.0 / p->Mn));
  const double *box = p->box;
  double QX[2 * 64], QY[2 * 64], QZ[2 * 64];
  if (nd > 64) abort();

  kick_lists(p, s);                                                     /* :225-250 */
  for (int64_t i = 0; i < s->in_num; i++) {                             /* :252-316 */
    int64_t n = s->in_list[i] - 1;
    stream_one(p, s, n, 0, 0.0);
    double v2 = (s->vx[n] * s->vx[n] + s->vy[n] * s->vy[n]) + s->vz[n] * s->vz[n];
    s->msd[n] = s->msd[n] + v2 * (MDt * MDt);
  }
  for (int k = 0; k < nd; k++) {                                        /* :325-415 */
    double prX[2], prY[2], prZ[2], dP[3], d[3];
    for (int i = 0; i < 2; i++) {
      prX[i] = s->xd[IX2(k, i, nd)]; prY[i] = s->yd[IX2(k, i, nd)]; prZ[i] = s->zd[IX2(k, i, nd)];
    }
    dP[0] = pbc_half(s->xd[IX2(k, 0, nd)] - s->xd[IX2(k, 1, nd)], box[0]);
    dP[1] = pbc_half(s->yd[IX2(k, 0, nd)] - s->yd[IX2(k, 1, nd)], box[1]);
    dP[2] = s->zd[IX2(k, 0, nd)] - s->zd[IX2(k, 1, nd)];
    QX[IX2(k, 0, nd)] = s->vxd[IX2(k, 0, nd)] + hc * s->fxd[IX2(k, 0, nd)];
    QY[IX2(k, 0, nd)] = s->vyd[IX2(k, 0, nd)] + hc * s->fyd[IX2(k, 0, nd)];
    QZ[IX2(k, 0, nd)] = s->vzd[IX2(k, 0, nd)] + hc * s->fzd[IX2(k, 0, nd)];
    QX[IX2(k, 1, nd)] = s->vxd[IX2(k, 1, nd)] + hn * s->fxd[IX2(k, 1, nd)];
    QY[IX2(k, 1, nd)] = s->vyd[IX2(k, 1, nd)] + hn * s->fyd[IX2(k, 1, nd)];
    QZ[IX2(k, 1, nd)] = s->vzd[IX2(k, 1, nd)] + hn * s->fzd[IX2(k, 1, nd)];
    for (int i = 0; i < 2; i++) {
      int64_t o = IX2(k, i, nd);
      s->xd[o] = s->xd[o] + MDt * QX[o];
      s->yd[o] = s->yd[o] + MDt * QY[o];
      s->zd[o] = s->zd[o] + MDt * QZ[o];
      s->xd[o] = s->xd[o] - box[0] * floor(s->xd[o] / box[0]);
      s->yd[o] = s->yd[o] - box[1] * floor(s->yd[o] / box[1]);
    }
    int64_t acc = 0;
    for (;;) {
      d[0] = pbc_half(s->xd[IX2(k, 0, nd)] - s->xd[IX2(k, 1, nd)], box[0]);
      d[1] = pbc_half(s->yd[IX2(k, 0, nd)] - s->yd[IX2(k, 1, nd)], box[1]);
      d[2] = s->zd[IX2(k, 0, nd)] - s->zd[IX2(k, 1, nd)];
      double c1 = (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - p->Dcn * p->Dcn;
      if (fabs(c1) < 1e-6 || acc == 200000000) break;
      double G = -(((0.25 * c1) * IIMcIMn) / (d[0] * dP[0] + d[1] * dP[1] + d[2] * dP[2]));
      int64_t o0 = IX2(k, 0, nd), o1 = IX2(k, 1, nd);
      double g2 = 2.0 * G;
      s->xd[o0] = (prX[0] + QX[o0] * MDt) + (g2 * dP[0]) / p->Mc;
      s->yd[o0] = (prY[0] + QY[o0] * MDt) + (g2 * dP[1]) / p->Mc;
      s->zd[o0] = (prZ[0] + QZ[o0] * MDt) + (g2 * dP[2]) / p->Mc;
      s->xd[o1] = (prX[1] + QX[o1] * MDt) - (g2 * dP[0]) / p->Mn;
      s->yd[o1] = (prY[1] + QY[o1] * MDt) - (g2 * dP[1]) / p->Mn;
      s->zd[o1] = (prZ[1] + QZ[o1] * MDt) - (g2 * dP[2]) / p->Mn;
      for (int i = 0; i < 2; i++) {
        int64_t o = IX2(k, i, nd);
        s->xd[o] = s->xd[o] - box[0] * floor(s->xd[o] / box[0]);
        s->yd[o] = s->yd[o] - box[1] * floor(s->yd[o] / box[1]);
      }
      QX[o0] = QX[o0] + (g2 * dP[0]) / (p->Mc * MDt);
      QY[o0] = QY[o0] + (g2 * dP[1]) / (p->Mc * MDt);
      QZ[o0] = QZ[o0] + (g2 * dP[2]) / (p->Mc * MDt);
      QX[o1] = QX[o1] - (g2 * dP[0]) / (p->Mn * MDt);
      QY[o1] = QY[o1] - (g2 * dP[1]) / (p->Mn * MDt);
      QZ[o1] = QZ[o1] - (g2 * dP[2]) / (p->Mn * MDt);
      acc++;
    }
  }
  /* :424-455 zero forces over the current lists */
  double fwz[2 * 64], wpot[2 * 64], tpp[2 * 64];
  double *Fm = calloc((size_t)(4 * nd * 2 * nd * 2), sizeof(double)); /* FmotX,Y,Z,pot (k,i,j,l) */
  int64_t nd4 = (int64_t)nd * 2 * nd * 2;
#define FM(c, k, i, j, l) Fm[(int64_t)(c) * nd4 + (k) + (int64_t)(i) * nd + (int64_t)(j) * 2 * nd + (int64_t)(l) * 2 * nd * nd]
  for (int k = 0; k < nd; k++)
    for (int i = 0; i < 2; i++) {
      int64_t o = IX2(k, i, nd);
      s->fxd[o] = 0.0; s->fyd[o] = 0.0; s->fzd[o] = 0.0;
      fwz[o] = 0.0; wpot[o] = 0.0; tpp[o] = 0.0;
      for (int64_t j = 0; j < s->nb_num[o]; j++) {
        int64_t q = IX3(k, i, j, nd);
        s->fpx[q] = 0.0; s->fpy[q] = 0.0; s->fpz[q] = 0.0; s->pot[q] = 0.0;
      }
    }
  s->tpotential = 0.0;
  double maxv = s->msd[0];                                              /* :457-463 */
  for (int64_t i = 0; i < p->NT; i++) if (s->msd[i] > maxv) maxv = s->msd[i];
  double hs = 0.5 * p->skinrad;
  if (maxv > hs * hs || s->time_count == s->time_penetrate || s->cstep % p->MPCtstep == 0) {
    double tc = (double)s->time_count;
    for (int64_t i = 0; i < s->out_num; i++) stream_one(p, s, s->out_list[i] - 1, 1, tc); /* :467-529 */
    if (s->cstep % p->sort_step == 0)                                   /* :531-533 */
      or_co_sort_hilbert(p->vol, p->NT, box, s->tour, s->x, s->y, s->z, s->flag, s->vx, s->vy, s->vz);
    rebuild_lists(p, s);                                                /* :536-553 */
    for (int64_t i = 0; i < p->NT; i++) s->msd[i] = 0.0;
    s->time_penetrate = penetrate_steps(p, s);                          /* :555-563 */
    s->time_count = 0;
  }
  co_pair_terms(p, s, Fm, fwz, wpot);
  for (int k = 0; k < nd; k++)                                          /* :609-636 */
    for (int i = 0; i < 2; i++) {
      int64_t o = IX2(k, i, nd);
      for (int64_t j = 0; j < s->nb_num[o]; j++) {
        int64_t q = IX3(k, i, j, nd);
        tpp[o] = tpp[o] + s->pot[q];
        s->fxd[o] = s->fxd[o] - s->fpx[q];
        s->fyd[o] = s->fyd[o] - s->fpy[q];
        s->fzd[o] = s->fzd[o] - s->fpz[q];
      }
      for (int j = 0; j < nd; j++)
        for (int l = 0; l < 2; l++) {
          s->fxd[o] = s->fxd[o] + FM(0, k, i, j, l);
          s->fyd[o] = s->fyd[o] + FM(1, k, i, j, l);
          s->fzd[o] = s->fzd[o] + FM(2, k, i, j, l);
          if (j < k) tpp[o] = tpp[o] + FM(3, k, i, j, l);
        }
      s->fzd[o] = s->fzd[o] + fwz[o];
      tpp[o] = tpp[o] + wpot[o];
    }
  for (int k = 0; k < nd; k++)
    for (int i = 0; i < 2; i++) s->tpotential = s->tpotential + tpp[IX2(k, i, nd)];
#undef FM
  free(Fm);
  kick_lists(p, s);                                                     /* :645-670 */
  for (int k = 0; k < nd; k++) {                                        /* :679-724 */
    int64_t o0 = IX2(k, 0, nd), o1 = IX2(k, 1, nd);
    s->vxd[o0] = QX[o0] + hc * s->fxd[o0];
    s->vyd[o0] = QY[o0] + hc * s->fyd[o0];
    s->vzd[o0] = QZ[o0] + hc * s->fzd[o0];
    s->vxd[o1] = QX[o1] + hn * s->fxd[o1];
    s->vyd[o1] = QY[o1] + hn * s->fyd[o1];
    s->vzd[o1] = QZ[o1] + hn * s->fzd[o1];
    int64_t acc = 0;
    for (;;) {
      double dvx = s->vxd[o0] - s->vxd[o1], dvy = s->vyd[o0] - s->vyd[o1], dvz = s->vzd[o0] - s->vzd[o1];
      double d0 = pbc_half(s->xd[o0] - s->xd[o1], box[0]);
      double d1 = pbc_half(s->yd[o0] - s->yd[o1], box[1]);
      double d2 = s->zd[o0] - s->zd[o1];
      double c2 = 2.0 * (dvx * d0 + dvy * d1 + dvz * d2);
      if (fabs(c2) < 1e-8 || acc == 200000000) break;
      double O = -((0.25 * (IIMcIMn * c2)) / (p->Dcn * p->Dcn));
      double o2 = 2.0 * O;
      s->vxd[o0] = s->vxd[o0] + (o2 * d0) / p->Mc;
      s->vyd[o0] = s->vyd[o0] + (o2 * d1) / p->Mc;
      s->vzd[o0] = s->vzd[o0] + (o2 * d2) / p->Mc;
      s->vxd[o1] = s->vxd[o1] - (o2 * d0) / p->Mn;
      s->vyd[o1] = s->vyd[o1] - (o2 * d1) / p->Mn;
      s->vzd[o1] = s->vzd[o1] - (o2 * d2) / p->Mn;
      acc++;
    }
  }
}

int or_co_distribute(const or_co_params *p, const or_co_state *s,
                     int64_t *tA, int64_t *tB, int64_t *tS,
                     int64_t *sA, int64_t *sB, int64_t *sS, int64_t maxs,
                     double cutC2, double cutN2) {
  /* distribute_mod.f95:81-201 at one thread: outside list first, then the
   * inside particles that lie outside every sphere's LJ cutoff.  Returns -1
   * when a cell index or a member count leaves the arrays. */
  int64_t vol = p->vol;
  int nd = p->numdims;
  for (int64_t c = 0; c < vol; c++) { tA[c] = 0; tB[c] = 0; tS[c] = 0; }
  for (int pass = 0; pass < 2; pass++) {
    int64_t n = pass == 0 ? s->out_num : s->in_num;
    const int64_t *lst = pass == 0 ? s->out_list : s->in_list;
    for (int64_t i = 0; i < n; i++) {
      int64_t q = lst[i] - 1;
      if (pass == 1) {
        int bulk = 1;
        for (int k = 0; k < nd; k++) {
          double c1 = pbc_half(s->x[q] - s->xd[IX2(k, 0, nd)], p->box[0]);
          double c2 = pbc_half(s->y[q] - s->yd[IX2(k, 0, nd)], p->box[1]);
          double c3 = s->z[q] - s->zd[IX2(k, 0, nd)];
          double n1 = pbc_half(s->x[q] - s->xd[IX2(k, 1, nd)], p->box[0]);
          double n2 = pbc_half(s->y[q] - s->yd[IX2(k, 1, nd)], p->box[1]);
          double n3 = s->z[q] - s->zd[IX2(k, 1, nd)];
          if (c1 * c1 + c2 * c2 + c3 * c3 <= cutC2 || n1 * n1 + n2 * n2 + n3 * n3 <= cutN2) bulk = 0;
        }
        if (!bulk) continue;
      }
      int64_t *t, *sl;
      if (s->flag[q] == 0) { t = tA; sl = sA; }
      else if (s->flag[q] == 1) { t = tB; sl = sB; }
      else if (s->flag[q] == 2) { t = tS; sl = sS; }
      else continue;
      int64_t c = cell_unshifted(s->x[q], s->y[q], s->z[q], p->box);
      if (c < 1 || c > vol) return -1;
      t[c - 1] += 1;
      if (t[c - 1] > maxs) return -1;
      sl[(c - 1) + (t[c - 1] - 1) * vol] = q + 1;
    }
  }
  return 0;
}

int or_co_mpc_rotation(int32_t NT, double rotang, int64_t numcell, double a0, const double *box,
                       const double *x, const double *y, const double *z,
                       double *vx, double *vy, double *vz,
                       double *vxcm, double *vycm, double *vzcm, double *tcs, int64_t *ncn,
                       double *uvx, double *uvy, double *uvz, or_mtmod_t *rng) {
  /* mpc_rotation_mod.f95:83-232 at one thread */
  double sinrot = sin(rotang), cosrot = cos(rotang);
  for (int64_t c = 0; c < numcell; c++) { vxcm[c] = 0.0; vycm[c] = 0.0; vzcm[c] = 0.0; tcs[c] = 0.0; }
  double xs = or_mtmod_grnd(rng); xs = a0 * (xs - 0.5);
  double ys = or_mtmod_grnd(rng); ys = a0 * (ys - 0.5);
  double zs = or_mtmod_grnd(rng); zs = a0 * (zs - 0.5);
  for (int64_t i = 0; i < NT; i++) {
    double X = x[i] + xs;
    X = X - box[0] * floor(X / box[0]);
    double Y = y[i] + ys;
    Y = Y - box[1] * floor(Y / box[1]);
    double Z = z[i] + zs;
    double fz = zs < 0.0 ? floor(Z + 1.0) : floor(Z);
    double c = (double)(int32_t)floor(X) + box[0] * (double)(int32_t)floor(Y);
    c = c + (box[0] * box[1]) * (double)(int32_t)fz;
    c = c + 1.0;
    ncn[i] = (int64_t)c;
    if (ncn[i] < 1 || ncn[i] > numcell) return -1;
  }
  for (int64_t i = 0; i < NT; i++) {
    int64_t c = ncn[i] - 1;
    tcs[c] = tcs[c] + 1.0;
    vxcm[c] = vxcm[c] + vx[i];
    vycm[c] = vycm[c] + vy[i];
    vzcm[c] = vzcm[c] + vz[i];
  }
  for (int64_t c = 0; c < numcell; c++) {
    if (tcs[c] != 0.0) { vxcm[c] = vxcm[c] / tcs[c]; vycm[c] = vycm[c] / tcs[c]; vzcm[c] = vzcm[c] / tcs[c]; }
    double r = or_mtmod_grnd(rng);
    double phi = (2.0 * PI) * r;
    r = or_mtmod_grnd(rng);
    double th = 2.0 * r - 1.0;
    double st = sqrt(1.0 - th * th);
    uvx[c] = st * cos(phi);
    uvy[c] = st * sin(phi);
    uvz[c] = th;
  }
  double omc = 1.0 - cosrot;
  for (int64_t i = 0; i < NT; i++) {
    int64_t c = ncn[i] - 1;
    double ux = uvx[c], uy = uvy[c], uz = uvz[c];
    double dx = vx[i] - vxcm[c], dy = vy[i] - vycm[c], dz = vz[i] - vzcm[c];
    double rx = (dx * (cosrot + (ux * ux) * omc) + dy * ((ux * uy) * omc - uz * sinrot)) +
                dz * ((ux * uz) * omc + uy * sinrot);
    double ry = (dx * ((ux * uy) * omc + uz * sinrot) + dy * (cosrot + (uy * uy) * omc)) +
                dz * ((uy * uz) * omc - ux * sinrot);
    double rz = (dx * ((ux * uz) * omc - uy * sinrot) + dy * ((uy * uz) * omc + ux * sinrot)) +
                dz * (cosrot + (uz * uz) * omc);
    vx[i] = vxcm[c] + rx;
    vy[i] = vycm[c] + ry;
    vz[i] = vzcm[c] + rz;
  }
  return 0;
}

int or_co_reaction_simple(double MPCt, int64_t numcell, double k1, int32_t *flag,
                          const int64_t *tB, const int64_t *sB, or_mtmod_t *rng) {
  /* reaction_mod.f95:41-62 at one thread */
  for (int64_t i = 0; i < numcell; i++) {
    double a1 = tB[i] == 0 ? 0.0 : k1 * (double)tB[i];
    double prob1 = 1.0 - exp(-(a1 * MPCt));
    double r = or_mtmod_grnd(rng);
    if (0.0 < r && r <= prob1) flag[sB[i] - 1] = 0;
  }
  return 0;
}

int or_co_time_step(const or_co_params *p, or_co_state *s, or_co_mpc *m, or_mtmod_t *rng) {
  /* main_pattern.f95:203-224 */
  s->cstep += 1;
  s->time_count += 1;
  or_co_evolve(p, s);
  if (s->cstep % p->MPCtstep != 0) return 0;
  double cutC = p->Rc * p->cutoff, cutN = p->Rn * p->cutoff;
  if (or_co_distribute(p, s, m->tA, m->tB, m->tS, m->sA, m->sB, m->sS, m->maxs, cutC * cutC,
                       cutN * cutN))
    return -1;
  if (or_co_mpc_rotation(p->NT, m->rotang, m->mpcnumcell, m->a0, p->box, s->x, s->y, s->z, s->vx,
                         s->vy, s->vz, m->vxcm, m->vycm, m->vzcm, m->tcs, m->ncn, m->uvx, m->uvy,
                         m->uvz, rng))
    return -1;
  return or_co_reaction_simple(m->MPCt, m->rxnnumcell, m->k1, s->flag, m->tB, m->sB, rng);
}

/* ===================== selkov_code ===================== */
static inline double pbc_h(double d, double H, double L) {
  if (d > H) return d - L;
  if (d < -H) return d + L;
  return d;
}

void or_sk_motion(const or_sk_params *p, or_sk_state *s) {
  /* Motion_SelkovDimer.f90:8-159 */
  const double MDt = p->MDt;
  const double hf2 = ((0.5 * MDt) * MDt) / p->Mf, hg2 = ((0.5 * MDt) * MDt) / p->Mg,
               hi2 = ((0.5 * MDt) * MDt) / p->Mi;
  const double hc = (0.5 * MDt) / p->Mc, hn = (0.5 * MDt) / p->Mn;
  const double IMc = 1.0 / p->Mc, IMn = 1.0 / p->Mn;
  const double IIMcIMn = 1.0 / (IMc + IMn);
  const double IMDt = 1.0 / MDt;
  for (int64_t i = 0; i < p->NT; i++) { s->pfx[i] = s->fx[i]; s->pfy[i] = s->fy[i]; s->pfz[i] = s->fz[i]; }
  double prX[2], prY[2], prZ[2], prVX[2], prVY[2], prVZ[2];
  for (int i = 0; i < 2; i++) {
    s->pfxd[i] = s->fxd[i]; s->pfyd[i] = s->fyd[i]; s->pfzd[i] = s->fzd[i];
    prX[i] = s->xd[i]; prY[i] = s->yd[i]; prZ[i] = s->zd[i];
    prVX[i] = s->vxd[i]; prVY[i] = s->vyd[i]; prVZ[i] = s->vzd[i];
  }
  for (int64_t i = 0; i < p->NT; i++) {
    double h;
    int64_t f = s->flag[i];
    int mv = 1;
    if (f == 0) h = hf2; else if (f == 1) h = hg2; else if (f == 3) h = hi2; else { h = 0.0; mv = 0; }
    if (mv) {
      s->x[i] = (s->x[i] + s->vx[i] * MDt) + s->pfx[i] * h;
      s->y[i] = (s->y[i] + s->vy[i] * MDt) + s->pfy[i] * h;
      s->z[i] = (s->z[i] + s->vz[i] * MDt) + s->pfz[i] * h;
    }
    s->x[i] = s->x[i] - p->box[0] * floor(s->x[i] / p->box[0]);
    s->y[i] = s->y[i] - p->box[1] * floor(s->y[i] / p->box[1]);
    s->z[i] = s->z[i] - p->box[2] * floor(s->z[i] / p->box[2]);
  }
  double dXp = pbc_h(prX[0] - prX[1], p->hbox[0], p->box[0]);
  double dYp = pbc_h(prY[0] - prY[1], p->hbox[1], p->box[1]);
  double dZp = pbc_h(prZ[0] - prZ[1], p->hbox[2], p->box[2]);
  s->qx[0] = prVX[0] + hc * s->pfxd[0]; s->qy[0] = prVY[0] + hc * s->pfyd[0]; s->qz[0] = prVZ[0] + hc * s->pfzd[0];
  s->qx[1] = prVX[1] + hn * s->pfxd[1]; s->qy[1] = prVY[1] + hn * s->pfyd[1]; s->qz[1] = prVZ[1] + hn * s->pfzd[1];
  for (int i = 0; i < 2; i++) {
    s->xd[i] = prX[i] + MDt * s->qx[i];
    s->yd[i] = prY[i] + MDt * s->qy[i];
    s->zd[i] = prZ[i] + MDt * s->qz[i];
    s->xd[i] = s->xd[i] - p->box[0] * floor(s->xd[i] / p->box[0]);
    s->yd[i] = s->yd[i] - p->box[1] * floor(s->yd[i] / p->box[1]);
    s->zd[i] = s->zd[i] - p->box[2] * floor(s->zd[i] / p->box[2]);
  }
  int64_t acc = 0;
  for (;;) {
    s->dx = pbc_h(s->xd[0] - s->xd[1], p->hbox[0], p->box[0]);
    s->dy = pbc_h(s->yd[0] - s->yd[1], p->hbox[1], p->box[1]);
    s->dz = pbc_h(s->zd[0] - s->zd[1], p->hbox[2], p->box[2]);
    double d2 = s->dx * s->dx + s->dy * s->dy + s->dz * s->dz;
    double c1 = d2 - p->Dcn * p->Dcn;
    if (fabs(c1) < 1e-6 || acc == 200000000) break;
    double G = (((0.5 * c1) * IMDt) * IIMcIMn) / (s->dx * dXp + s->dy * dYp + s->dz * dZp);
    s->qx[0] = s->qx[0] - (G * IMc) * dXp; s->qy[0] = s->qy[0] - (G * IMc) * dYp; s->qz[0] = s->qz[0] - (G * IMc) * dZp;
    s->qx[1] = s->qx[1] + (G * IMn) * dXp; s->qy[1] = s->qy[1] + (G * IMn) * dYp; s->qz[1] = s->qz[1] + (G * IMn) * dZp;
    for (int i = 0; i < 2; i++) {
      s->xd[i] = prX[i] + MDt * s->qx[i];
      s->yd[i] = prY[i] + MDt * s->qy[i];
      s->zd[i] = prZ[i] + MDt * s->qz[i];
      s->xd[i] = s->xd[i] - p->box[0] * floor(s->xd[i] / p->box[0]);
      s->yd[i] = s->yd[i] - p->box[1] * floor(s->yd[i] / p->box[1]);
      s->zd[i] = s->zd[i] - p->box[2] * floor(s->zd[i] / p->box[2]);
    }
    acc++;
  }
}

static inline int64_t sk_cell(const or_sk_params *p, double x, double y, double z) {
  /* INT(FLOOR(X)+Box(1)*FLOOR(Y)+Box(1)*Box(2)*FLOOR(Z)+1) */
  double c = (double)(int32_t)floor(x) + p->box[0] * (double)(int32_t)floor(y);
  c = c + (p->box[0] * p->box[1]) * (double)(int32_t)floor(z);
  return (int64_t)(c + 1.0);
}

int or_sk_distribute(const or_sk_params *p, or_sk_state *s) {
  /* Distribute_SelkovDimer.f90:12-35 */
  for (int64_t c = 0; c < p->vol; c++) { s->tF[c] = 0; s->tG[c] = 0; s->tI[c] = 0; }
  for (int64_t i = 0; i < p->NT; i++) {
    int64_t *t, *sl;
    if (s->flag[i] == 0) { t = s->tF; sl = s->sF; }
    else if (s->flag[i] == 1) { t = s->tG; sl = s->sG; }
    else if (s->flag[i] == 3) { t = s->tI; sl = s->sI; }
    else continue;
    int64_t c = sk_cell(p, s->x[i], s->y[i], s->z[i]);
    if (c < 1 || c > p->vol) return -1;
    t[c - 1] += 1;
    if (t[c - 1] > s->maxs) return -1;
    sl[(c - 1) + (t[c - 1] - 1) * p->vol] = i + 1;
  }
  return 0;
}

void or_sk_forces(const or_sk_params *p, or_sk_state *s, or_mt95_t *rng) {
  /* ForcesMD_SelkovDimer.f90:9-303 at one thread (draws in particle order) */
  const double Rn6 = pow(p->Rn, 6), Rn12 = pow(p->Rn, 12), Rc6 = pow(p->Rc, 6), Rc12 = pow(p->Rc, 12);
  const double r6Rn6 = 6.0 * Rn6, r12Rn12 = 12.0 * Rn12, r6Rc6 = 6.0 * Rc6, r12Rc12 = 12.0 * Rc12;
  const double hf = (0.5 * p->MDt) / p->Mf, hg = (0.5 * p->MDt) / p->Mg, hi = (0.5 * p->MDt) / p->Mi;
  const double hc = (0.5 * p->MDt) / p->Mc, hn = (0.5 * p->MDt) / p->Mn;
  const double IMc = 1.0 / p->Mc, IMn = 1.0 / p->Mn;
  const double IIMcIMn = 1.0 / (IMc + IMn);
  const int64_t NT = p->NT, L = s->lead; /* FpairX(1:space,1:2), column stride = space */
  s->tpotential = 0.0;
  for (int64_t i = 0; i < L; i++) {
    s->fx[i] = 0.0; s->fy[i] = 0.0; s->fz[i] = 0.0; s->pot[i] = 0.0;
    s->fpx[i] = 0.0; s->fpy[i] = 0.0; s->fpz[i] = 0.0;
    s->fpx[L + i] = 0.0; s->fpy[L + i] = 0.0; s->fpz[L + i] = 0.0;
  }
  for (int j = 0; j < 2; j++) { s->fxd[j] = 0.0; s->fyd[j] = 0.0; s->fzd[j] = 0.0; }
  for (int64_t i = 0; i < NT; i++) {                                /* :42-113 N sphere */
    double d1 = pbc_h(s->x[i] - s->xd[1], p->hbox[0], p->box[0]);
    double d2 = pbc_h(s->y[i] - s->yd[1], p->hbox[1], p->box[1]);
    double d3 = pbc_h(s->z[i] - s->zd[1], p->hbox[2], p->box[2]);
    double dRsq = d1 * d1 + d2 * d2 + d3 * d3;
    if (!(dRsq <= p->cutoffN2)) continue;
    double E;
    if (s->flag[i] == 0) E = p->EfN; else if (s->flag[i] == 1) E = p->EgN;
    else if (s->flag[i] == 3) E = p->EiN; else continue;
    double F, U;
    lj_terms(dRsq, Rn6, Rn12, r12Rn12, r6Rn6, &F, &U);
    double EF = E * F;
    s->fpx[L + i] = EF * d1; s->fpy[L + i] = EF * d2; s->fpz[L + i] = EF * d3;
    s->pot[i] = E * U;
    s->fx[i] = s->fpx[L + i]; s->fy[i] = s->fpy[L + i]; s->fz[i] = s->fpz[L + i];
  }
  for (int64_t i = 0; i < NT; i++) {                                /* :123-234 C sphere */
    double d1 = pbc_h(s->x[i] - s->xd[0], p->hbox[0], p->box[0]);
    double d2 = pbc_h(s->y[i] - s->yd[0], p->hbox[1], p->box[1]);
    double d3 = pbc_h(s->z[i] - s->zd[0], p->hbox[2], p->box[2]);
    double dRsq = d1 * d1 + d2 * d2 + d3 * d3;
    if (!(dRsq <= p->cutoffC2)) continue;
    double E;
    int64_t f = s->flag[i];
    if (f == 0) {
      double r = or_mt95_real1(rng);
      E = p->EfC;
      if (p->RPfg > r) s->flag[i] = 1;
    } else if (f == 1) {
      double r = or_mt95_real1(rng);
      E = p->EgC;
      if (p->RPgf > r) s->flag[i] = 0;
    } else if (f == 3) E = p->EiC;
    else continue;
    double F, U;
    lj_terms(dRsq, Rc6, Rc12, r12Rc12, r6Rc6, &F, &U);
    double EF = E * F;
    s->fpx[i] = EF * d1; s->fpy[i] = EF * d2; s->fpz[i] = EF * d3;
    s->pot[i] = E * U;
    s->fx[i] = s->fpx[i]; s->fy[i] = s->fpy[i]; s->fz[i] = s->fpz[i];
  }
  for (int64_t i = 0; i < NT; i++) {                                /* :237-246 */
    int64_t f = s->flag[i];
    if (f == 0 || f == 1 || f == 3) {
      s->tpotential = s->tpotential + s->pot[i];
      for (int j = 0; j < 2; j++) {
        s->fxd[j] = s->fxd[j] - s->fpx[j * L + i];
        s->fyd[j] = s->fyd[j] - s->fpy[j * L + i];
        s->fzd[j] = s->fzd[j] - s->fpz[j * L + i];
      }
    }
  }
  for (int64_t i = 0; i < NT; i++) {                                /* :255-269 */
    double h;
    int64_t f = s->flag[i];
    if (f == 0) h = hf; else if (f == 1) h = hg; else if (f == 3) h = hi; else continue;
    s->vx[i] = s->vx[i] + (s->pfx[i] + s->fx[i]) * h;
    s->vy[i] = s->vy[i] + (s->pfy[i] + s->fy[i]) * h;
    s->vz[i] = s->vz[i] + (s->pfz[i] + s->fz[i]) * h;
  }
  s->vxd[0] = s->qx[0] + hc * s->fxd[0]; s->vyd[0] = s->qy[0] + hc * s->fyd[0]; s->vzd[0] = s->qz[0] + hc * s->fzd[0];
  s->vxd[1] = s->qx[1] + hn * s->fxd[1]; s->vyd[1] = s->qy[1] + hn * s->fyd[1]; s->vzd[1] = s->qz[1] + hn * s->fzd[1];
  int64_t acc = 0;
  for (;;) {                                                        /* :283-303 */
    double dvx = s->vxd[0] - s->vxd[1], dvy = s->vyd[0] - s->vyd[1], dvz = s->vzd[0] - s->vzd[1];
    double c2 = dvx * s->dx + dvy * s->dy + dvz * s->dz;
    if (fabs(c2) < 1e-8 || acc == 200000000) break;
    double O = (IIMcIMn * c2) / (p->Dcn * p->Dcn);
    s->vxd[0] = s->vxd[0] - (O * s->dx) * IMc; s->vyd[0] = s->vyd[0] - (O * s->dy) * IMc; s->vzd[0] = s->vzd[0] - (O * s->dz) * IMc;
    s->vxd[1] = s->vxd[1] + (O * s->dx) * IMn; s->vyd[1] = s->vyd[1] + (O * s->dy) * IMn; s->vzd[1] = s->vzd[1] + (O * s->dz) * IMn;
    acc++;
  }
}

void or_sk_affected_cell(const or_sk_params *p, or_sk_state *s) {
  /* Affected_Cell.f90:5-99 */
  for (int64_t c = 0; c < p->vol; c++) s->affected[c] = 0;
  for (int sph = 0; sph < 2; sph++) {
    double cut = sph == 0 ? p->cutoffC : p->cutoffN;
    double cut2 = sph == 0 ? p->cutoffC2 : p->cutoffN2;
    double X = s->xd[sph], Y = s->yd[sph], Z = s->zd[sph];
    int64_t lx = (int64_t)ceil(X - (cut + 1.0)), ux = (int64_t)floor(X + (cut + 1.0));
    int64_t ly = (int64_t)ceil(Y - (cut + 1.0)), uy = (int64_t)floor(Y + (cut + 1.0));
    int64_t lz = (int64_t)ceil(Z - (cut + 1.0)), uz = (int64_t)floor(Z + (cut + 1.0));
    for (int64_t i = lx; i <= ux; i++)
      for (int64_t j = ly; j <= uy; j++)
        for (int64_t k = lz; k <= uz; k++) {
          /* REAL(i) is single precision; the difference and "-1." promote to double */
          double a = fabs((double)(float)i - X) - 1.0;
          double b = fabs((double)(float)j - Y) - 1.0;
          double c = fabs((double)(float)k - Z) - 1.0;
          if (a * a + b * b + c * c < cut2) {
            int64_t cx = i, cy = j, cz = k;
            if (cx < 0) cx += p->ibox[0]; else if (cx >= p->ibox[0]) cx -= p->ibox[0];
            if (cy < 0) cy += p->ibox[1]; else if (cy >= p->ibox[1]) cy -= p->ibox[1];
            if (cz < 0) cz += p->ibox[2]; else if (cz >= p->ibox[2]) cz -= p->ibox[2];
            int64_t cn = cx + p->ibox[0] * cy + p->ibox[0] * p->ibox[1] * cz + 1;
            s->affected[cn - 1] = 1;
          }
        }
  }
}

int or_sk_rmpc(const or_sk_params *p, or_sk_state *s, or_mt95_t *rng) {
  /* RMPC_SelkovDimer.f90:15-324 */
  const int64_t V = p->vol;
  for (int64_t c = 0; c < V; c++) { s->vxcm[c] = 0.0; s->vycm[c] = 0.0; s->vzcm[c] = 0.0; s->tcs[c] = 0; }
  double xs = or_mt95_real1(rng); xs = p->A0 * (xs - 0.5);
  double ys = or_mt95_real1(rng); ys = p->A0 * (ys - 0.5);
  double zs = or_mt95_real1(rng); zs = p->A0 * (zs - 0.5);
  for (int64_t i = 0; i < p->NT; i++) {
    int64_t f = s->flag[i];
    if (!(f == 0 || f == 1 || f == 3)) continue;
    double X = s->x[i] + xs; X = X - p->box[0] * floor(X / p->box[0]);
    double Y = s->y[i] + ys; Y = Y - p->box[1] * floor(Y / p->box[1]);
    double Z = s->z[i] + zs; Z = Z - p->box[2] * floor(Z / p->box[2]);
    int64_t c = sk_cell(p, X, Y, Z);
    if (c < 1 || c > V) return -1;
    s->ncn[i] = c;
    s->tcs[c - 1] += 1;
    s->vxcm[c - 1] = s->vxcm[c - 1] + s->vx[i];
    s->vycm[c - 1] = s->vycm[c - 1] + s->vy[i];
    s->vzcm[c - 1] = s->vzcm[c - 1] + s->vz[i];
  }
  for (int64_t c = 0; c < V; c++) {
    if (s->tcs[c] != 0) {
      double it = 1.0 / (double)s->tcs[c];
      s->vxcm[c] = s->vxcm[c] * it; s->vycm[c] = s->vycm[c] * it; s->vzcm[c] = s->vzcm[c] * it;
    }
  }
  for (int64_t c = 0; c < V; c++) {
    double x = or_mt95_real1(rng), y = or_mt95_real1(rng);
    double phi = (2.0 * PI) * x;   /* 2*pi folds to the double 2*pi */
    double th = 2.0 * y - 1.0;
    double st = sqrt(1.0 - th * th);
    s->uvx[c] = st * cos(phi); s->uvy[c] = st * sin(phi); s->uvz[c] = th;
  }
  for (int64_t i = 0; i < p->NT; i++) {
    int64_t f = s->flag[i];
    if (!(f == 0 || f == 1 || f == 3)) continue;
    int64_t c = s->ncn[i] - 1;
    double ux = s->uvx[c], uy = s->uvy[c], uz = s->uvz[c];
    double dx = s->vx[i] - s->vxcm[c], dy = s->vy[i] - s->vycm[c], dz = s->vz[i] - s->vzcm[c];
    double pr = (dx * ux + dy * uy) + dz * uz;
    double px = pr * ux, py = pr * uy, pz = pr * uz;
    double ox = dx - px, oy = dy - py, oz = dz - pz;
    double o2x = (oy * uz) - (oz * uy), o2y = (oz * ux) - (ox * uz), o2z = (ox * uy) - (oy * ux);
    double rx = (px + ox * p->cosrot) + o2x * p->sinrot;
    double ry = (py + oy * p->cosrot) + o2y * p->sinrot;
    double rz = (pz + oz * p->cosrot) + o2z * p->sinrot;
    s->vx[i] = s->vxcm[c] + rx; s->vy[i] = s->vycm[c] + ry; s->vz[i] = s->vzcm[c] + rz;
  }
#define SF1(c) (s->sF[(c)] - 1)
#define SG1(c) (s->sG[(c)] - 1)
#define SI1(c) (s->sI[(c)] - 1)
  for (int64_t c = 0; c < V; c++) {                                 /* :127-324 */
    if (s->affected[c] != 0) continue;
    int64_t tF = s->tF[c], tG = s->tG[c], tI = s->tI[c];
    double a1 = p->k1, a2 = p->k2 * (double)tF, a5 = p->k5 * (double)tG, a6 = p->k6, a3, a4;
    if (tG == 0) a3 = 0.0; else a3 = ((p->k3 * (double)tF) * (double)tG) * (double)(tG - 1);
    if (tG <= 1) a4 = 0.0; else a4 = ((p->k4 * (double)tG) * (double)(tG - 1)) * (double)(tG - 2);
    if (tI == 0) { a1 = 0.0; a6 = 0.0; }
    double a00 = ((((a1 + a2) + a3) + a4) + a5) + a6;
    double P = 1.0 - exp(-(a00 * p->MPCt));
    double x = or_mt95_real1(rng);
    double ar = a00 * x;
    x = or_mt95_real1(rng);
    if (!(P > x)) continue;
    int64_t tgt = -1, val = 0;
    if (tI > 0) {
      if (tF == 0) {
        if (tG == 0) {
          if (ar < a1) { tgt = SI1(c); val = 0; }
          else if (ar >= a1 && ar < a00) { tgt = SI1(c); val = 1; }
        } else if (tG > 0 && tG <= 2) {
          if (ar < a1) { tgt = SI1(c); val = 0; }
          else if (ar >= a1 && ar < a1 + a5) { tgt = SG1(c); val = 3; }
          else if (ar >= a1 + a5 && ar < a00) { tgt = SI1(c); val = 1; }
        } else if (tG > 2) {
          if (ar < a1) { tgt = SI1(c); val = 0; }
          else if (ar >= a1 && ar < a1 + a4) { tgt = SG1(c); val = 0; }
          else if (ar >= a1 + a4 && ar < (a1 + a4) + a5) { tgt = SG1(c); val = 3; }
          else if (ar >= (a1 + a4) + a5 && ar < a00) { tgt = SI1(c); val = 1; }
        }
      } else if (tF > 0) {
        double s12 = a1 + a2, s123 = s12 + a3;
        if (tG == 0) {
          if (ar < a1) { tgt = SI1(c); val = 0; }
          else if (ar >= a1 && ar < s12) { tgt = SF1(c); val = 3; }
          else if (ar >= s12 && ar < a00) { tgt = SI1(c); val = 1; }
        } else if (tG == 1) {
          if (ar < a1) { tgt = SI1(c); val = 0; }
          else if (ar >= a1 && ar < s12) { tgt = SF1(c); val = 3; }
          else if (ar >= s12 && ar < s12 + a5) { tgt = SG1(c); val = 3; }
          else if (ar >= s12 + a5 && ar < a00) { tgt = SI1(c); val = 1; }
        } else if (tG == 2) {
          if (ar < a1) { tgt = SI1(c); val = 0; }
          else if (ar >= a1 && ar < s12) { tgt = SF1(c); val = 3; }
          else if (ar >= s12 && ar < s123) { tgt = SF1(c); val = 1; }
          else if (ar >= s123 && ar < s123 + a5) { tgt = SG1(c); val = 3; }
          else if (ar >= s123 + a5 && ar < a00) { tgt = SI1(c); val = 1; }
        } else if (tG > 2) {
          double s1234 = s123 + a4;
          if (ar < a1) { tgt = SI1(c); val = 0; }
          else if (ar >= a1 && ar < s12) { tgt = SF1(c); val = 3; }
          else if (ar >= s12 && ar < s123) { tgt = SF1(c); val = 1; }
          else if (ar >= s123 && ar < s1234) { tgt = SG1(c); val = 0; }
          else if (ar >= s1234 && ar < s1234 + a5) { tgt = SG1(c); val = 3; }
          else if (ar >= s1234 + a5 && ar < a00) { tgt = SI1(c); val = 1; }
        }
      }
    } else if (tI == 0) {
      if (tF == 0) {
        if (tG == 0) {
        } else if (tG > 0 && tG <= 2) {
          if (ar >= a1 && ar < a00) { tgt = SG1(c); val = 3; }
        } else if (tG > 2) {
          if (ar >= a1 && ar < a1 + a4) { tgt = SG1(c); val = 0; }
          else if (ar >= a1 + a4 && ar < a00) { tgt = SG1(c); val = 3; }
        }
      } else if (tF > 0) {
        double s12 = a1 + a2, s123 = s12 + a3;
        if (tG == 0) {
          if (ar >= a1 && ar < a00) { tgt = SF1(c); val = 3; }
        } else if (tG == 1) {
          if (ar >= a1 && ar < s12) { tgt = SF1(c); val = 3; }
          else if (ar >= s12 && ar < a00) { tgt = SG1(c); val = 3; }
        } else if (tG == 2) {
          if (ar >= a1 && ar < s12) { tgt = SF1(c); val = 3; }
          else if (ar >= s12 && ar < s123) { tgt = SF1(c); val = 1; }
          else if (ar >= s123 && ar < a00) { tgt = SG1(c); val = 3; }
        } else if (tG > 2) {
          if (ar >= a1 && ar < s12) { tgt = SF1(c); val = 3; }
          else if (ar >= s12 && ar < s123) { tgt = SF1(c); val = 1; }
          else if (ar >= s123 && ar < s123 + a4) { tgt = SG1(c); val = 0; }
          else if (ar >= s123 + a4 && ar < a00) { tgt = SG1(c); val = 3; }
        }
      }
    }
    if (tgt >= 0) s->flag[tgt] = val;
  }
#undef SF1
#undef SG1
#undef SI1
  return 0;
}

int or_sk_sort(or_sk_params *p, or_sk_state *s) {
  /* Sort.f90:13-183: per cell, F members then G then I, cells ascending;
   * slots past the new count are cleared to FLAG=2 and NT shrinks. */
  int64_t V = p->vol, NT = p->NT;
  if (or_sk_distribute(p, s)) return -1;
  double *b = malloc(sizeof(double) * (size_t)NT * 12);
  int64_t *nf = malloc(sizeof(int64_t) * (size_t)NT);
  int64_t nw = 0;
  for (int64_t c = 0; c < V; c++)
    for (int sp = 0; sp < 3; sp++) {
      int64_t *t = sp == 0 ? s->tF : sp == 1 ? s->tG : s->tI;
      int64_t *sl = sp == 0 ? s->sF : sp == 1 ? s->sG : s->sI;
      for (int64_t j = 0; j < t[c]; j++) {
        int64_t q = sl[c + j * V] - 1;
        double *o = b + nw * 12;
        o[0] = s->x[q]; o[1] = s->y[q]; o[2] = s->z[q]; o[3] = s->vx[q]; o[4] = s->vy[q]; o[5] = s->vz[q];
        o[6] = s->fx[q]; o[7] = s->fy[q]; o[8] = s->fz[q]; o[9] = s->pfx[q]; o[10] = s->pfy[q]; o[11] = s->pfz[q];
        nf[nw] = s->flag[q];
        nw++;
      }
    }
  for (int64_t m = 0; m < NT; m++) {
    if (m < nw) {
      double *o = b + m * 12;
      s->x[m] = o[0]; s->y[m] = o[1]; s->z[m] = o[2]; s->vx[m] = o[3]; s->vy[m] = o[4]; s->vz[m] = o[5];
      s->fx[m] = o[6]; s->fy[m] = o[7]; s->fz[m] = o[8]; s->pfx[m] = o[9]; s->pfy[m] = o[10]; s->pfz[m] = o[11];
      s->flag[m] = nf[m];
    } else {
      s->x[m] = 0; s->y[m] = 0; s->z[m] = 0; s->vx[m] = 0; s->vy[m] = 0; s->vz[m] = 0;
      s->fx[m] = 0; s->fy[m] = 0; s->fz[m] = 0; s->pfx[m] = 0; s->pfy[m] = 0; s->pfz[m] = 0;
      s->flag[m] = 2;
    }
  }
  p->NT = nw;
  free(b); free(nf);
  return 0;
}

int or_sk_time_step(or_sk_params *p, or_sk_state *s, or_mt95_t *rng, int64_t cstep) {
  /* Main_SelkovDimer.f90:126-138 */
  if (cstep % p->MPCtstep != 0) {
    or_sk_motion(p, s);
    or_sk_forces(p, s, rng);
    return 0;
  }
  if (cstep % (p->sort_step > 0 ? p->sort_step : 10000) == 0 && or_sk_sort(p, s)) return -1;
  or_sk_motion(p, s);
  if (or_sk_distribute(p, s)) return -1;
  or_sk_forces(p, s, rng);
  or_sk_affected_cell(p, s);
  return or_sk_rmpc(p, s, rng);
}

/* ===================== flat drivers (same argument lists as oracle/ref_*_harness.f90) ===================== */
int or_co_run(int64_t *ip, double *rp, int64_t nsteps, int32_t do_setup, int32_t *flag,
              double *x, double *y, double *z, double *vx, double *vy, double *vz,
              double *xd, double *yd, double *zd, double *vxd, double *vyd, double *vzd,
              double *fxd, double *fyd, double *fzd, double *fpx, double *fpy, double *fpz, double *pot,
              int64_t *nb, int64_t *nbnum, double *msd, int64_t *ol, int64_t *il, int64_t nrb,
              int32_t *react_box, int64_t *tA, int64_t *tB, int64_t *tS, int64_t *sA, int64_t *sB,
              int64_t *sS, double *vxcm, double *vycm, double *vzcm, double *tcs, int64_t *ncn,
              double *uvx, double *uvy, double *uvz, uint32_t *mt, int32_t *mti, const int64_t *tour) {
  (void)nrb;
  or_co_params p;
  memset(&p, 0, sizeof p);
  p.NT = (int32_t)ip[0]; p.numdims = (int32_t)ip[1]; p.MPCtstep = (int32_t)ip[2];
  p.sort_step = (int32_t)ip[3]; p.react_dir = (int32_t)ip[4]; p.vol = ip[8];
  for (int i = 0; i < 3; i++) p.box[i] = rp[i];
  p.MDt = rp[3]; p.Ma = rp[4]; p.Mb = rp[5]; p.Ms = rp[6]; p.Mc = rp[7]; p.Mn = rp[8];
  p.Dcn = rp[9]; p.cutoff = rp[10]; p.Rc = rp[11]; p.Rn = rp[12];
  p.EaC = rp[13]; p.EbC = rp[14]; p.EsC = rp[15]; p.EaN = rp[16]; p.EbN = rp[17]; p.EsN = rp[18];
  p.Rw = rp[19]; p.Ew = rp[20]; p.Rm_delta = rp[21]; p.Em = rp[22]; p.RPab = rp[23]; p.RPba = rp[24];
  p.skinrad = rp[25];
  or_co_state s;
  memset(&s, 0, sizeof s);
  s.flag = flag; s.x = x; s.y = y; s.z = z; s.vx = vx; s.vy = vy; s.vz = vz;
  s.xd = xd; s.yd = yd; s.zd = zd; s.vxd = vxd; s.vyd = vyd; s.vzd = vzd;
  s.fxd = fxd; s.fyd = fyd; s.fzd = fzd; s.fpx = fpx; s.fpy = fpy; s.fpz = fpz; s.pot = pot;
  s.nb = nb; s.nb_num = nbnum; s.msd = msd; s.out_list = ol; s.in_list = il;
  s.out_num = ip[10]; s.in_num = ip[11]; s.cstep = (int32_t)ip[5];
  s.time_count = (int32_t)ip[6]; s.time_penetrate = (int32_t)ip[7];
  s.react_box = react_box; s.tour = tour; s.max_nb = ip[9];
  or_co_mpc m;
  memset(&m, 0, sizeof m);
  m.rotang = rp[26]; m.a0 = rp[27]; m.MPCt = rp[28]; m.k1 = rp[29];
  m.mpcnumcell = ip[12]; m.rxnnumcell = ip[13]; m.maxs = 100;
  m.vxcm = vxcm; m.vycm = vycm; m.vzcm = vzcm; m.tcs = tcs; m.uvx = uvx; m.uvy = uvy; m.uvz = uvz;
  m.ncn = ncn; m.tA = tA; m.tB = tB; m.tS = tS; m.sA = sA; m.sB = sB; m.sS = sS;
  or_mtmod_t rng;
  memcpy(rng.mt, mt, sizeof rng.mt);
  rng.mti = *mti;
  int rc = 0;
  if (do_setup) or_co_setup_lists(&p, &s, do_setup == 2 || do_setup == 4);
  if (do_setup >= 3) or_co_restart_forces(&p, &s);
  for (int64_t it = 0; it < nsteps && rc == 0; it++) rc = or_co_time_step(&p, &s, &m, &rng);
  memcpy(mt, rng.mt, sizeof rng.mt);
  *mti = rng.mti;
  ip[5] = s.cstep; ip[6] = s.time_count; ip[7] = s.time_penetrate; ip[10] = s.out_num; ip[11] = s.in_num;
  rp[30] = s.tpotential;
  return rc;
}

int or_sk_run(int64_t *ip, double *rp, int32_t seed, int64_t cstep0, int64_t nsteps, int64_t *flag,
              double *x, double *y, double *z, double *vx, double *vy, double *vz,
              double *fx, double *fy, double *fz, double *pfx, double *pfy, double *pfz, double *dm) {
  /* mirrors ref_sk_run in oracle/ref_selkov_harness.f90; dm is (2,12) column-major */
  or_sk_params p;
  memset(&p, 0, sizeof p);
  const int64_t space = ip[1];
  p.NT = ip[0];
  p.sort_step = ip[2];
  for (int i = 0; i < 3; i++) {
    p.box[i] = rp[i]; p.hbox[i] = 0.5 * rp[i]; p.ibox[i] = (int64_t)rp[i];
  }
  p.vol = p.ibox[0] * p.ibox[1] * p.ibox[2];
  p.Mf = rp[3]; p.Mg = rp[4]; p.Mi = rp[5]; p.Rc = rp[6]; p.Rn = rp[7]; p.Dcn = rp[8];
  p.EfC = rp[9]; p.EgC = rp[10]; p.EiC = rp[11]; p.EfN = rp[12]; p.EgN = rp[13]; p.EiN = rp[14];
  p.RPfg = rp[15]; p.RPgf = rp[16]; p.A0 = rp[17]; p.MDt = rp[18]; p.MPCt = rp[19];
  p.k1 = rp[20]; p.k2 = rp[21]; p.k3 = rp[22]; p.k4 = rp[23]; p.k5 = rp[24]; p.k6 = rp[25];
  p.Mc = rp[26]; p.Mn = rp[27];
  p.MPCtstep = (int64_t)(p.MPCt / p.MDt);
  double cut = pow(2.0, 1.0 / 6.0);
  p.cutoffC = p.Rc * cut; p.cutoffN = p.Rn * cut;
  p.cutoffC2 = p.cutoffC * p.cutoffC; p.cutoffN2 = p.cutoffN * p.cutoffN;
  p.sinrot = sin(PI / 2.0); p.cosrot = cos(PI / 2.0);
  or_sk_state s;
  memset(&s, 0, sizeof s);
  int64_t V = p.vol, MS = 200;
  s.flag = flag; s.x = x; s.y = y; s.z = z; s.vx = vx; s.vy = vy; s.vz = vz;
  s.fx = fx; s.fy = fy; s.fz = fz; s.pfx = pfx; s.pfy = pfy; s.pfz = pfz;
  s.lead = space;
  s.fpx = calloc((size_t)space * 2, sizeof(double));
  s.fpy = calloc((size_t)space * 2, sizeof(double));
  s.fpz = calloc((size_t)space * 2, sizeof(double));
  s.pot = calloc((size_t)space, sizeof(double));
  for (int i = 0; i < 2; i++) {
    s.xd[i] = dm[i]; s.yd[i] = dm[2 + i]; s.zd[i] = dm[4 + i];
    s.vxd[i] = dm[6 + i]; s.vyd[i] = dm[8 + i]; s.vzd[i] = dm[10 + i];
    s.fxd[i] = dm[12 + i]; s.fyd[i] = dm[14 + i]; s.fzd[i] = dm[16 + i];
  }
  s.maxs = MS;
  s.tF = calloc((size_t)V, 8); s.tG = calloc((size_t)V, 8); s.tI = calloc((size_t)V, 8);
  s.sF = calloc((size_t)(V * MS), 8); s.sG = calloc((size_t)(V * MS), 8); s.sI = calloc((size_t)(V * MS), 8);
  s.affected = calloc((size_t)V, 8);
  s.vxcm = calloc((size_t)V, 8); s.vycm = calloc((size_t)V, 8); s.vzcm = calloc((size_t)V, 8);
  s.uvx = calloc((size_t)V, 8); s.uvy = calloc((size_t)V, 8); s.uvz = calloc((size_t)V, 8);
  s.tcs = calloc((size_t)V, 8); s.ncn = calloc((size_t)space, 8);
  or_mt95_t rng;
  or_mt95_init(&rng, seed);
  int rc = 0;
  for (int64_t it = 0; it < nsteps && rc == 0; it++) rc = or_sk_time_step(&p, &s, &rng, cstep0 + it);
  ip[0] = p.NT;
  for (int i = 0; i < 2; i++) {
    dm[i] = s.xd[i]; dm[2 + i] = s.yd[i]; dm[4 + i] = s.zd[i];
    dm[6 + i] = s.vxd[i]; dm[8 + i] = s.vyd[i]; dm[10 + i] = s.vzd[i];
    dm[12 + i] = s.fxd[i]; dm[14 + i] = s.fyd[i]; dm[16 + i] = s.fzd[i];
  }
  rp[28] = s.tpotential;
  free(s.fpx); free(s.fpy); free(s.fpz); free(s.pot); free(s.tF); free(s.tG); free(s.tI);
  free(s.sF); free(s.sG); free(s.sI); free(s.affected); free(s.vxcm); free(s.vycm); free(s.vzcm);
  free(s.uvx); free(s.uvy); free(s.uvz); free(s.tcs); free(s.ncn);
  return rc;
}

int or_co_distribute_flat(int32_t NT, int32_t numdims, const int32_t *flag,
                          const double *x, const double *y, const double *z,
                          const double *xd, const double *yd, const double *zd, const double *box,
                          int64_t vol, int64_t *tA, int64_t *tB, int64_t *tS,
                          int64_t *sA, int64_t *sB, int64_t *sS, int64_t maxs,
                          double cutC2, double cutN2,
                          int64_t out_num, const int64_t *out_list, int64_t in_num, const int64_t *in_list) {
  /* or_co_distribute with the flat argument list of distribute_mod.f95:6-40 */
  or_co_params p;
  or_co_state s;
  memset(&p, 0, sizeof p);
  memset(&s, 0, sizeof s);
  p.NT = NT; p.numdims = numdims; p.vol = vol;
  for (int d = 0; d < 3; d++) p.box[d] = box[d];
  s.flag = (int32_t *)flag;
  s.x = (double *)x; s.y = (double *)y; s.z = (double *)z;
  s.xd = (double *)xd; s.yd = (double *)yd; s.zd = (double *)zd;
  s.out_list = (int64_t *)out_list; s.in_list = (int64_t *)in_list;
  s.out_num = out_num; s.in_num = in_num;
  return or_co_distribute(&p, &s, tA, tB, tS, sA, sB, sS, maxs, cutC2, cutN2);
}
