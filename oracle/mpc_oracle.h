/*
 * CPU restatement of the reference MD + MPC hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the HIP path.  Only tests/, the smoke entry
 * and bench.py's CPU-baseline leg may load it; the product path never does.
 *
 * Every function states the reference routine (file:line) it restates.  Array
 * layouts are the Fortran ones (column-major, 1-based particle numbers stored
 * as values), so results compare element by element with the flang-built
 * reference modules in oracle/_ref.
 *
 * Build with -ffp-contract=off: every expression keeps the reference's
 * evaluation order and two-rounding multiply/add.
 */
#ifndef MPC_ORACLE_H
#define MPC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Mersenne Twister, collective_code/mtmod.f90 flavour (1998 seeding) ---- */
typedef struct { uint32_t mt[624]; int32_t mti; } or_mtmod_t;
void   or_mtmod_sgrnd(or_mtmod_t *s, int32_t seed);          /* mtmod.f90:72-85  */
double or_mtmod_grnd(or_mtmod_t *s);                          /* mtmod.f90:88-144 */
void   or_mtmod_fill(or_mtmod_t *s, double *out, int64_t n);  /* n successive grnd() */

/* ---- Mersenne Twister, selkov_code/mt95.f90 flavour (2002 seeding) ---- */
typedef struct { uint32_t val[624]; int32_t cnt; } or_mt95_t;  /* cnt is 1-based like mt95 */
void     or_mt95_init(or_mt95_t *s, int32_t seed);            /* mt95.f90:378-404 */
uint32_t or_mt95_int32(or_mt95_t *s);                         /* mt95.f90:457-485,583-602 */
double   or_mt95_real1(or_mt95_t *s);                         /* mt95.f90:828-841 */
void     or_mt95_fill(or_mt95_t *s, double *out, int64_t n);

/* ---- collective_code parameters (argument lists of evolve_mod.f95:6-87) ---- */
typedef struct {
  int32_t NT, numdims, MPCtstep, sort_step, react_dir, pad0;
  int64_t vol;                 /* Lx*Ly*Lz unshifted cells                    */
  double box[3];
  double MDt, Ma, Mb, Ms, Mc, Mn, Dcn, cutoff, Rc, Rn;
  double EaC, EbC, EsC, EaN, EbN, EsN, Rw, Ew, Rm_delta, Em, RPab, RPba;
  double skinrad;
} or_co_params;

/* Mutable state of one collective system, Fortran layouts:
 *   dimer arrays (numdims,2); neighbour/Fpair/Potential (numdims,2,max_nb);
 *   neighbour_num (numdims,2).                                               */
typedef struct {
  int32_t *flag;
  double *x, *y, *z, *vx, *vy, *vz;
  double *xd, *yd, *zd, *vxd, *vyd, *vzd, *fxd, *fyd, *fzd;
  double *fpx, *fpy, *fpz, *pot;
  int64_t *nb, *nb_num;
  double *msd;                                   /* max_solvent_dist(NT)      */
  int64_t *out_list, *in_list;                   /* outside (NT), inside      */
  int64_t out_num, in_num;
  int32_t time_count, time_penetrate;
  int32_t cstep, pad1;
  double tpotential;
  int32_t *react_box;                            /* (Lx*Ly)                   */
  const int64_t *tour;                           /* hilbert tour, vol entries */
  int64_t max_nb;
} or_co_state;

void or_co_neighbour_list(int32_t NT, const double *box, double cutoff, double skin,
                          const double *x, const double *y, const double *z,
                          double xd, double yd, double zd,
                          int64_t *num, int64_t *nb, int64_t stride);      /* neighbour_list_mod.f95:6-59 */
void or_co_inout(int32_t NT, const double *box, double cutC, double cutN, double skin,
                 const double *x, const double *y, const double *z,
                 const double *xd, const double *yd, const double *zd, int32_t numdims,
                 int64_t *out_num, int64_t *out_list, int64_t *in_num, int64_t *in_list); /* :63-184 */
void or_co_lj_mot_sol(int32_t sphere, const double *box, int32_t *flag,
                      const double *x, const double *y, const double *z,
                      double xd, double yd, double zd,
                      double *fpx, double *fpy, double *fpz, double *pot,
                      double RPab, double RPba, double cut2, double Ea, double Eb, double Es,
                      double R6, double R12, double r12R12, double r6R6,
                      int64_t num, const int64_t *nb, int64_t stride);      /* forces_mod.f95:6-140 */
void or_co_reactive_grid_setup(const double *box, int32_t *react_box, int32_t type); /* reactive_grid_mod.f95:8-99 */
void or_co_reactive_grid(const int32_t *react_box, int32_t react_dir, const double *box,
                         double x, double y, double vx, double vy, int32_t *flag, double t); /* :103-151 */
void or_co_sort_hilbert(int64_t vol, int32_t NT, const double *box, const int64_t *tour,
                        double *x, double *y, double *z, int32_t *flag,
                        double *vx, double *vy, double *vz);               /* sort_hilbert_mod.f95:6-98 */
void or_co_evolve(const or_co_params *p, or_co_state *s);                 /* evolve_mod.f95:6-732 */
int  or_co_distribute(const or_co_params *p, const or_co_state *s,
                      int64_t *tA, int64_t *tB, int64_t *tS,
                      int64_t *sA, int64_t *sB, int64_t *sS, int64_t maxs,
                      double cutC2, double cutN2);                          /* distribute_mod.f95:6-205 */
int  or_co_distribute_flat(int32_t NT, int32_t numdims, const int32_t *flag,
                           const double *x, const double *y, const double *z,
                           const double *xd, const double *yd, const double *zd, const double *box,
                           int64_t vol, int64_t *tA, int64_t *tB, int64_t *tS,
                           int64_t *sA, int64_t *sB, int64_t *sS, int64_t maxs,
                           double cutC2, double cutN2,
                           int64_t out_num, const int64_t *out_list, int64_t in_num, const int64_t *in_list);
int  or_co_mpc_rotation(int32_t NT, double rotang, int64_t numcell, double a0, const double *box,
                        const double *x, const double *y, const double *z,
                        double *vx, double *vy, double *vz,
                        double *vxcm, double *vycm, double *vzcm, double *tcs, int64_t *ncn,
                        double *uvx, double *uvy, double *uvz, or_mtmod_t *rng); /* mpc_rotation_mod.f95:6-239 */
int  or_co_reaction_simple(double MPCt, int64_t numcell, double k1, int32_t *flag,
                           const int64_t *tB, const int64_t *sB, or_mtmod_t *rng); /* reaction_mod.f95:8-66 */
/* one pass of the main_pattern.f95:202-225 time loop (cstep+=1, time_count+=1,
 * evolve, and at collision steps distribute/mpc_rotation/reaction_simple).   */
typedef struct {
  double rotang, a0, MPCt, k1;
  int64_t mpcnumcell, rxnnumcell, maxs;
  double *vxcm, *vycm, *vzcm, *tcs, *uvx, *uvy, *uvz;
  int64_t *ncn, *tA, *tB, *tS, *sA, *sB, *sS;
} or_co_mpc;
int or_co_time_step(const or_co_params *p, or_co_state *s, or_co_mpc *m, or_mtmod_t *rng);
/* main_pattern.f95:170-188: sort, neighbour lists, in/out lists, time_penetrate */
void or_co_setup_lists(const or_co_params *p, or_co_state *s, int do_sort);
/* restart_mod.f95:425-495: forces, pair forces and potential evaluated again
 * from the restored positions over the lists or_co_setup_lists built */
void or_co_restart_forces(const or_co_params *p, or_co_state *s);

/* ---- selkov_code (module globals become explicit arrays) ---- */
typedef struct {
  int64_t NT, vol, MPCtstep;
  int64_t ibox[3];
  double box[3], hbox[3];
  double MDt, MPCt, Mf, Mg, Mi, Mc, Mn, Dcn, Rc, Rn;
  double EfC, EgC, EiC, EfN, EgN, EiN, RPfg, RPgf;
  double cutoffC, cutoffN, cutoffC2, cutoffN2, A0;
  double k1, k2, k3, k4, k5, k6;
  double sinrot, cosrot;
  int64_t sort_step;   /* Main_SelkovDimer.f90:131 sorts every 10000 steps; other values only in tests */
} or_sk_params;
typedef struct {
  int64_t *flag;
  double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz, *pfx, *pfy, *pfz;
  double *fpx, *fpy, *fpz;        /* (NT,2) column-major */
  double *pot;
  double xd[2], yd[2], zd[2], vxd[2], vyd[2], vzd[2], fxd[2], fyd[2], fzd[2];
  double pfxd[2], pfyd[2], pfzd[2];
  double qx[2], qy[2], qz[2];
  double dx, dy, dz;             /* bond vector left by Motion for ForcesMD */
  double tpotential;
  int64_t *tF, *tG, *tI, *sF, *sG, *sI, maxs;
  int64_t *affected;
  double *vxcm, *vycm, *vzcm, *uvx, *uvy, *uvz;
  int64_t *tcs, *ncn;
  int64_t lead;                  /* allocated length "space" of the particle arrays */
} or_sk_state;
void or_sk_motion(const or_sk_params *p, or_sk_state *s);                  /* Motion_SelkovDimer.f90:1-161 */
int  or_sk_distribute(const or_sk_params *p, or_sk_state *s);              /* Distribute_SelkovDimer.f90:1-37 */
void or_sk_forces(const or_sk_params *p, or_sk_state *s, or_mt95_t *rng);  /* ForcesMD_SelkovDimer.f90:1-305 */
void or_sk_affected_cell(const or_sk_params *p, or_sk_state *s);           /* Affected_Cell.f90:1-101 */
int  or_sk_rmpc(const or_sk_params *p, or_sk_state *s, or_mt95_t *rng);    /* RMPC_SelkovDimer.f90:1-326 */
int  or_sk_sort(or_sk_params *p, or_sk_state *s);                          /* Sort.f90:1-185 */
/* one pass of Main_SelkovDimer.f90:125-148 for step cstep */
int  or_sk_time_step(or_sk_params *p, or_sk_state *s, or_mt95_t *rng, int64_t cstep);

#ifdef __cplusplus
}
#endif
#endif
