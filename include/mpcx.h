/*
 * mpcx -- MI355X (gfx950) engine for the MD + multiparticle-collision step of
 * the sphere-dimer motor simulators (collective_code and selkov_code).
 *
 * C ABI: plain pointers and sizes, no C++ or torch types.  Every function
 * returns 0 on success or a negative MPCX_E_* code; mpcx_last_error() gives
 * the message of the last failure on the calling thread.
 *
 * Two ways in:
 *   1. Reference-shaped entry points (mpcx_mpc_rotation, mpcx_distribute, ...)
 *      take the host arrays and argument lists of the Fortran module
 *      procedures they replace, so a Fortran caller binds them one-to-one
 *      through ISO_C_BINDING.  Each call stages its arrays through the device.
 *   2. A device-resident engine (mpcx_co_* / mpcx_sk_*) keeps the solvent and
 *      dimer state in HBM across many collision intervals and replaces the
 *      whole time loop of main_pattern.f95:202-225 / Main_SelkovDimer.f90:125-149.
 *
 * Index conventions follow the reference: particle and cell numbers stored in
 * integer arrays are 1-based; 2-D/3-D arrays are column-major with the
 * reference's shapes.
 */
#ifndef MPCX_H
#define MPCX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPCX_OK 0
#define MPCX_E_ARG -1        /* invalid argument or size                        */
#define MPCX_E_HIP -2        /* HIP runtime failure                             */
#define MPCX_E_DEVICE -3     /* a kernel reported an out-of-range state         */
#define MPCX_E_UNSUPPORTED -4/* configuration outside what this engine handles  */

const char* mpcx_last_error(void);
/* number of visible HIP devices (>0 when a GPU is usable) */
int mpcx_device_count(void);
/* hipDeviceSynchronize on the current device */
int mpcx_device_synchronize(void);
/* No counterpart in the reference: destroys every engine, batch and network
 * still live (batches first) and drains the device, so no stream, event,
 * graph or buffer outlives the HIP runtime at process exit.  Handles are
 * invalid afterwards.  The Python front end calls it at interpreter exit. */
int mpcx_shutdown(void);
/* Measurement aid with no counterpart in the reference: copies nbytes between
 * two device buffers reps times (after one warm-up copy) and returns the mean
 * time of one copy.  Its hardware fetch/write counter readings calibrate the
 * HBM traffic figures of the engine's kernels. */
int mpcx_probe_copy(int64_t nbytes, int32_t reps, double* ms_per_copy);

/* The throughput mode's device draws, with no counterpart in the reference
 * (which draws from Mersenne Twister on the host): Philox4x32-10 (Salmon,
 * Moraes, Dror, Shaw, SC'11) with key = seed, counter = (idx, step, stream),
 * returned as out[2i], out[2i+1] = the two 53-bit uniforms in [0,1) of index
 * idx0 + i; stream 0 is the grid shift, 1 the collision axes, 2 the bulk
 * reaction.  For tests and for callers that want the engine's own stream.
 * n <= 2^28, stream < 65536. */
int mpcx_philox_uniform2(uint64_t seed, uint32_t stream, uint64_t step, uint64_t idx0, int64_t n, double* out);

/* Host Mersenne Twister streams (parity mode draws).
 * mpcx_mtmod_*: sgrnd/grnd of collective_code/mtmod.f90:72-144 (state mt[624], mti).
 * mpcx_mt95_*:  init_by_scalar/genrand_real1 of selkov_code/mt95.f90:378-404,828-841
 *               (state val[624], 1-based cnt). */
int mpcx_mtmod_seed(int32_t seed, uint32_t* mt, int32_t* mti);
int mpcx_mtmod_fill(uint32_t* mt, int32_t* mti, double* out, int64_t n);
int mpcx_mt95_seed(int32_t seed, uint32_t* val, int32_t* cnt);
int mpcx_mt95_fill(uint32_t* val, int32_t* cnt, double* out, int64_t n);

/* ------------------------------------------------------------------------
 * Reference-shaped entry points (collective_code)
 * ---------------------------------------------------------------------- */

/* Replaces mpc_rotation (collective_code/mpc_rotation_mod.f95:6-36).
 * draws: the 3 + 2*numcell uniforms the reference takes from grnd() at one
 * OpenMP thread, in call order (shift x,y,z; then two per cell). */
int mpcx_mpc_rotation(int32_t NT, double rotang, int64_t numcell, double a0, const double* box,
                      const double* Xsol, const double* Ysol, const double* Zsol,
                      double* VXsol, double* VYsol, double* VZsol,
                      double* VXCM, double* VYCM, double* VZCM, double* tCellShift, int64_t* NewCellN,
                      double* UVX, double* UVY, double* UVZ, const double* draws);

/* Replaces reaction_simple (collective_code/reaction_mod.f95:8-27).
 * vol is the leading dimension of sBCellN(vol,100); draws holds numcell uniforms. */
int mpcx_reaction_simple(double MPCt, int64_t numcell, double k1, int32_t NT, int32_t* flag,
                         const int64_t* tBCellN, const int64_t* sBCellN, int64_t vol, const double* draws);

/* Replaces neighbour_list (collective_code/neighbour_list_mod.f95:6-26).  The
 * list is written with the given element stride (2*numdims for a section
 * neighbour(k,i,:) of the reference array); *neighbour_num is incremented as
 * in the reference. */
int mpcx_neighbour_list(int32_t NT, const double* box, double cutoff, double skinrad,
                        const double* Xsol, const double* Ysol, const double* Zsol,
                        double Xdim, double Ydim, double Zdim, int64_t* neighbour_num,
                        int64_t* neighbour, int64_t stride, int64_t max_neighbour);

/* Replaces inout_particle_list (collective_code/neighbour_list_mod.f95:63-87). */
int mpcx_inout_particle_list(int32_t NT, const double* box, double cutoffC, double cutoffN, double skinrad,
                             const double* xsol, const double* ysol, const double* zsol,
                             const double* xdim, const double* ydim, const double* zdim,
                             int64_t* outside_particle_num, int64_t* outside_particle,
                             int64_t* inside_particle_num, int64_t* inside_particle,
                             int64_t inside_capacity, int32_t numdims);

/* No counterpart in the reference: mpcx_inout_particle_list for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same lists, appended the same
 * way, but xsol/ysol/zsol and the two lists are DEVICE arrays the caller owns
 * (outside_capacity/inside_capacity entries); box, the sphere arrays and the
 * counts stay on the host.  It calls no hipMalloc or hipFree: its temporaries
 * live in a fixed 527,616-byte array that each GPU holds until the process
 * ends.  Limits: NT <= 67,108,864 and numdims <= 64 (rejected with an error
 * beyond them).  Runs on the null stream and returns when the lists are
 * written.  Calls that overlap on one GPU share that array and race. */
int mpcx_inout_particle_list_noalloc(int32_t NT, const double* box, double cutoffC, double cutoffN, double skinrad,
                                     const double* xsol, const double* ysol, const double* zsol,
                                     const double* xdim, const double* ydim, const double* zdim,
                                     int64_t* outside_particle_num, int64_t* outside_particle,
                                     int64_t outside_capacity, int64_t* inside_particle_num,
                                     int64_t* inside_particle, int64_t inside_capacity, int32_t numdims);

/* Replaces distribute (collective_code/distribute_mod.f95:6-40); the member
 * lists s?CellN are (vol, maxs) column-major. */
int mpcx_distribute(int32_t NT, int32_t numdims, const int32_t* flag,
                    const double* Xsol, const double* Ysol, const double* Zsol,
                    const double* Xdim, const double* Ydim, const double* Zdim, const double* box,
                    int64_t vol, int64_t* tACellN, int64_t* tBCellN, int64_t* tSCellN,
                    int64_t* sACellN, int64_t* sBCellN, int64_t* sSCellN, int64_t maxs,
                    double cutoffC2, double cutoffN2,
                    int64_t outside_particle_num, const int64_t* outside_particle,
                    int64_t inside_particle_num, const int64_t* inside_particle);

/* Replaces sort_hilbert (collective_code/sort_hilbert_mod.f95:6-24).  tour holds
 * the vol 1-based cell numbers the reference builds from hilbert.txt (:57-62). */
int mpcx_sort_hilbert(int64_t vol, int32_t NT, const double* box, const int64_t* tour,
                      double* xsol, double* ysol, double* zsol, int32_t* flag,
                      double* vxsol, double* vysol, double* vzsol);

/* ------------------------------------------------------------------------
 * Device-resident collective engine
 * ---------------------------------------------------------------------- */
typedef struct mpcx_co mpcx_co;

/* Parameters of one collective_code system: the input.txt quantities
 * (main_pattern.f95:19-64) and the derived ones (:66-76, :117-127). */
typedef struct {
  int64_t NT, numdims, vol, mpcnumcell, rxnnumcell, max_neighbour;
  int32_t MPCtstep, sort_step, react_dir, rng_mode; /* rng_mode 0: host draws, 1: device Philox */
  uint64_t seed;                                    /* Philox key in rng_mode 1              */
  double box[3];
  double MDt, Ma, Mb, Ms, Mc, Mn, Dcn, cutoff, Rc, Rn;
  double EaC, EbC, EsC, EaN, EbN, EsN, Rw, Ew, Rm_delta, Em, RPab, RPba;
  double skinrad, rotang, a0, MPCt, k1;
  double sinrot, cosrot;   /* dsin/dcos(rotang) evaluated on the host (mpc_rotation_mod.f95:83-84) */
  double wall_w6;          /* 3.d0**(1./6.) as the reference evaluates it (evolve_mod.f95:218)    */
  double reaction_p[101];  /* 1 - dexp(-k1*n*MPCt) for n = 0..100 (reaction_mod.f95:45-54)        */
  double sphere_guard;     /* bound on a sphere's displacement per interval (near/far split)      */
} mpcx_co_config;

int mpcx_co_create(mpcx_co** out, const mpcx_co_config* cfg);
int mpcx_co_destroy(mpcx_co* e);

/* Upload solvent and dimer state; dimer arrays are (numdims,2) column-major.
 * react_box is (Lx*Ly) from reactive_grid_setup; tour as in mpcx_sort_hilbert.
 * The seven solvent arrays may all be NULL (then mpcx_co_init_solvent). */
int mpcx_co_set_state(mpcx_co* e, int32_t cstep, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz, const int32_t* flag,
                      const double* xd, const double* yd, const double* zd,
                      const double* vxd, const double* vyd, const double* vzd,
                      const double* fxd, const double* fyd, const double* fzd,
                      const int32_t* react_box, const int64_t* tour);

/* Set-up of main_pattern.f95:170-188: optional sort, neighbour and in/out
 * lists, time_penetrate; prepares the near/far split of the first interval. */
int mpcx_co_setup(mpcx_co* e, int32_t do_sort);
/* Set-up of a restart (restart_mod.f95:416-495): the same sort, lists and
 * time_penetrate, then the forces, pair forces and potential evaluated from
 * the uploaded positions (an A particle inside a catalytic cutoff turns B, as
 * in any force evaluation) instead of the zero forces of a fresh start.
 * do_sort = (cstep % sort_step == 0) as restart_mod.f95:423.  Near modes 0
 * and 2. */
int mpcx_co_setup_restart(mpcx_co* e, int32_t do_sort);

/* Frame of the system at a collision step: solvent positions, velocities and
 * species, dimer positions and velocities -- what restart_mod.f95:362-406
 * reads from the last data.h5 frame -- in a binary file (64-byte header with
 * sizes, box and a checksum; written under a temporary name, then renamed).
 * mpcx_co_load_frame checks the file against the engine's system and uploads
 * it as mpcx_co_set_state does (dimer forces zero until
 * mpcx_co_setup_restart); *cstep (may be NULL) receives the frame's step. */
int mpcx_co_save_frame(mpcx_co* e, const char* path);
int mpcx_co_load_frame(mpcx_co* e, const char* path, const int32_t* react_box, const int64_t* tour,
                       int32_t* cstep);

/* Advance ncoll collision intervals (MPCtstep MD steps plus one collision
 * each).  In rng_mode 0, draws holds ncoll blocks of
 * 3 + 2*mpcnumcell + rxnnumcell uniforms (grnd() order of one collision). */
int mpcx_co_run(mpcx_co* e, int64_t ncoll, const double* draws);

/* Initial solvent on the device (initial_mod.f95:122-345 with Philox draws
 * from `seed` instead of the Mersenne Twister): flags A below Na0, B below
 * Na0+Nb0, S above; positions uniform in the box outside every sphere's LJ
 * cutoff (the dimers set by mpcx_co_set_state, which then takes NULL solvent
 * arrays); Gaussian velocities of variance T/m; total momentum removed. */
int mpcx_co_init_solvent(mpcx_co* e, int64_t Na0, int64_t Nb0, double T, uint64_t seed);

/* Download state.  Arrays may be NULL to skip them. */
int mpcx_co_get_state(mpcx_co* e, double* x, double* y, double* z, double* vx, double* vy, double* vz,
                      int32_t* flag, double* dimer12 /* (numdims*2, 12): x y z vx vy vz fx fy fz qx qy qz */,
                      int32_t* scalars /* cstep, time_count, time_penetrate, inside_num */);
/* Inside list (ascending 0-based particle indices), the neighbour-list
 * membership bit mask of each entry (bit 2k: C sphere of dimer k, 2k+1: N),
 * and the force on each entry from each sphere it belongs to (up to 8,
 * by bit: fp[b*3*cap + c*cap + q]).  Returns the count in *n.  One-workgroup
 * mode only (up to four dimers); mpcx_co_get_lists works in both modes. */
int mpcx_co_get_inside(mpcx_co* e, int32_t* list, uint32_t* memb, double* fp, double* pot, int64_t cap,
                       int64_t* n);
/* Neighbour lists as neighbour_list_mod.f95:38-57 leaves them (neighbour,
 * neighbour_num) with the forces_mod.f95:55-136 pair forces and potentials
 * of the last MD step: nb/fpx/fpy/fpz/pot are (numdims, 2, max_nb)
 * column-major with 1-based particle indices, ascending; nbnum (numdims, 2);
 * in_list the inside list (0-based, ascending, inout_particle_list), *in_num
 * its length.  Any pointer but nbnum and in_num may be NULL. */
int mpcx_co_get_lists(mpcx_co* e, int64_t max_nb, int64_t* nb, int64_t* nbnum, double* fpx, double* fpy,
                      double* fpz, double* pot, int32_t* in_list, int64_t in_cap, int64_t* in_num);
/* Near-engine layout for two or more dimers (call before mpcx_co_setup).
 *  mode 0: one workgroup for the whole system (up to four dimers; the default
 *          for one or two dimers);
 *  mode 1 (two or more dimers): dimers whose near regions can meet within an interval are grouped
 *          into clusters of at most four, one workgroup each;
 *  mode 2: the system's near set split over cooperative workgroups, any number
 *          of dimers at any density (the default above two dimers).
 * In modes 1 and 2 the workgroups meet at a grid barrier every MD step to take
 * the system-wide list-rebuild decision of evolve_mod.f95:457-465.  No
 * counterpart in the reference (its OpenMP loops over dimers and particles in
 * evolve_mod.f95 run in one address space). */
int mpcx_co_set_near_mode(mpcx_co* e, int32_t mode);
/* Workgroups of mode 2 (1..128; default 2 per dimer, at least 2).  Mode 2 also
 * runs one-dimer systems (its near set split over the workgroups). */
int mpcx_co_set_near_workgroups(mpcx_co* e, int32_t G);
/* split = 1: split-precision solvent (the 0-based unshifted cell and FP32
 * offsets in it, FP32 velocities, 8-bit species; far particles stream in one
 * flight beside the near engine and bin, every cell's velocity sum is exact
 * (FP32 velocities in 64-bit fixed point, 2^-40 units, integer addition: no
 * dependence on the order of the members), particles rotate in FP32; near
 * particles and the dimers stay FP64 in the near engine).  Device Philox
 * draws only; call before mpcx_co_setup.  mpcx_co_run then runs the system
 * as a batch of one (mpcx_cob_*), in any near mode.  No counterpart in the
 * reference (FP64 throughout). */
int mpcx_co_set_precision(mpcx_co* e, int32_t split);
/* Workgroups of the interval's streaming/binning kernel (co_bulk, or sp_bulk
 * in split storage), a grid-stride loop over the particles: 0 (the default)
 * takes one per 256 particles up to 8192; a positive value is used as given
 * (at most 65536), for tuning.  Results do not depend on it.  No counterpart
 * in the reference. */
int mpcx_co_set_bulk_grid(mpcx_co* e, int32_t grid);
/* on = 1 (the default): with the one-workgroup near engine (mode 0), the far
 * particles stream in one free flight of the whole interval and bin on a
 * second stream while the near engine runs (co_far; sp_far in split
 * storage); the near set is binned afterwards (co_bin_rest, sp_bin_near).
 * In FP64 storage far positions then agree with the reference's
 * event-by-event lazy streaming (evolve_mod.f95:474-527) to a few ulps rather
 * than bitwise, and a particle whose flight reaches a z wall -- where the
 * bounce time (z - Lz)/vz would magnify those ulps -- replays the events
 * exactly after the near engine; integer state is unaffected unless a
 * particle lies within that rounding of a cell face.  on = 0: every rebuild
 * event is replayed after the near engine, bit for bit.  No counterpart in
 * the reference. */
int mpcx_co_set_overlap(mpcx_co* e, int32_t on);
/* Collision-cell data of the last collision: (mpcnumcell) arrays, 0-based
 * cell of each particle, and unshifted census counts (3, vol). */
int mpcx_co_get_cells(mpcx_co* e, double* vcm_uv /* (mpcnumcell, 8) */, int32_t* cell_of,
                      int32_t* census /* (3, vol) */, double* tpotential);
/* Wall time of the HIP events bracketing the last mpcx_co_run (ms), and of
 * the last interval's bulk kernels (co_bulk .. co_rotate) and near engine (ms). */
int mpcx_co_timing(mpcx_co* e, double* run_ms, double* bulk_ms, double* near_ms);
/* Near-engine counters accumulated since creation (10 values): real-time
 * ticks (100 MHz) of [0] launches from the first MD step to the end, [1] list
 * rebuilds, [2] particle phases, [3] dimer (serial) sections; [4] rebuild
 * count; [5] last near-set size; [6] last inside-list size; [7] launches;
 * ticks of [8] set-up and gather before the first step, [9] the end phase
 * (census exclusion, write-back).  The one-workgroup engine fills [8], [9]. */
int mpcx_co_near_profile(mpcx_co* e, uint64_t* prof10);
/* Mean duration (ms) of each kernel of the collision interval over the last
 * mpcx_co_run, from HIP events on the stream each runs on; every 8th interval
 * is timed (at most 256 of them), since each event costs the stream a few
 * microseconds: [0] near engine, [1] co_far (or co_bulk without overlap;
 * at a re-sort these run as launches around the sort, which are not timed),
 * [2] co_collide, [3] co_rotate, [4] co_scatter; count = timed intervals.
 * Split storage: the kernels of mpcx_cob_kernel_ms. */
int mpcx_co_kernel_ms(mpcx_co* e, double* ms5, int64_t* count);

/* ------------------------------------------------------------------------
 * Batches of collective systems (no reference counterpart: the reference
 * runs one system per process; collective_code/main_pattern.f95:202-225 is
 * the loop each replica runs).  A batch advances R engines created and set
 * up as above (rng_mode 1, same numdims (1 or 2, near mode 0), MPCtstep,
 * sort_step and storage precision; boxes, parameters and keys may differ)
 * with one launch per kernel of an interval for all of them: the near engine
 * with one workgroup per replica, the particle and cell kernels over tiles
 * of `tile` particles / 256 cells.  A split batch runs its interval in
 * order (the near engines hold a CU each; a far flight beside them would
 * only wait for CUs); a one-engine split run streams its far particles on a
 * second stream beside its near engine.  Each
 * replica ends bit-identical to running it alone (a split engine's
 * mpcx_co_run is a batch of one); its state is read with mpcx_co_get_state.
 * With use_graph, the launches of an interval without a re-sort are
 * captured once into a HIP graph and replayed.
 * ---------------------------------------------------------------------- */
typedef struct mpcx_cob mpcx_cob;
int mpcx_cob_create(mpcx_cob** out, int32_t R, mpcx_co* const* engines, int32_t tile);
int mpcx_cob_destroy(mpcx_cob* b);
int mpcx_cob_run(mpcx_cob* b, int64_t ncoll, int32_t use_graph);
int mpcx_cob_set_tile(mpcx_cob* b, int32_t tile);
/* Collision-cell slots of the per-workgroup hash table of the split-precision
 * binning: 1024, 2048, or 0 (default: 2048 for tiles of 4096 particles and
 * more, else 1024).  Results do not depend on it. */
int mpcx_cob_set_bulk_slots(mpcx_cob* b, int32_t slots);
/* Re-sort period of a split-precision batch in MD steps (a positive multiple
 * of MPCtstep; default: the engines' sort_step).  Split storage sums cells in
 * exact integers and keys every draw by cell, so the period changes only the
 * memory locality and the order of the near engines' force sums; a replica
 * matches its engine run alone at the same period.  FP64 batches refuse it. */
int mpcx_cob_set_sort_step(mpcx_cob* b, int32_t steps);
/* Entries per inbox of a split-precision batch (0: max(tile / 2, 256)).  An
 * inbox receives the cell sums other particle tiles binned into the cells
 * its tile owns; what does not fit goes through atomics into the
 * accumulator copies instead, so results do not depend on it. */
int mpcx_cob_set_inbox_cap(mpcx_cob* b, int32_t cap);
int mpcx_cob_info(mpcx_cob* b, int64_t* particles, int64_t* cells, double* run_ms);
/* Split batches run without the graph: mean kernel durations (ms) of the last
 * run from HIP events on the stream each runs on (every 8th interval, at most
 * 256 of them): [0] near engines, [1] sb_bulk (far flight and binning), [2]
 * sb_cells_own (cell records, reactions), [3] sb_rotate, [4] sb_near (near set
 * binning); *count = timed intervals. */
int mpcx_cob_kernel_ms(mpcx_cob* b, double* ms5, int64_t* count);
/* Advance n intervals and, after each, add every replica's body-frame B
 * counts (R, nz, nr) (frame of its dimer 0, as mpcx_co_body_density) to
 * `counts` and its propulsion velocity to vsum[r]: the time averages that
 * train the surrogate. */
int mpcx_cob_body_average(mpcx_cob* b, int64_t n, int32_t nr, int32_t nz, double rmax, double zmin, double zmax,
                          uint32_t* counts, double* vsum);

/* ------------------------------------------------------------------------
 * Device-resident Selkov engine (selkov_code)
 * ---------------------------------------------------------------------- */
typedef struct mpcx_sk mpcx_sk;

/* Parameters of one selkov_code system: Input_SelkovDimer.txt quantities and
 * the derived ones of Main_SelkovDimer.f90:60-104 (T already divided by 5 is
 * not needed here: velocities come with the state). */
typedef struct {
  int64_t NT, space;                 /* active particles, allocated slots (FLAG 2 beyond NT)        */
  int32_t MPCtstep, sort_step;       /* MD steps per collision; re-sort period (reference: 10000) */
  int32_t rng_mode, pad;             /* 0: host mt95 draws, 1: device Philox                      */
  uint64_t seed;                     /* Philox key in rng_mode 1                                  */
  double box[3];
  double MDt, MPCt, Mf, Mg, Mi, Mc, Mn, Dcn, Rc, Rn;
  double EfC, EgC, EiC, EfN, EgN, EiN, RPfg, RPgf, A0;
  double k[6];
  double sinrot, cosrot;             /* sin/cos of pi/2 as the reference evaluates them          */
  double sphere_guard;               /* bound on a sphere's displacement per interval             */
} mpcx_sk_config;

int mpcx_sk_create(mpcx_sk** out, const mpcx_sk_config* cfg);
int mpcx_sk_destroy(mpcx_sk* e);
/* Upload the state at a collision step (cstep a multiple of MPCtstep); arrays
 * of `space` entries as the module arrays of Global_SelkovDimer.f90 (FLAG
 * int64; FXsol.. the forces of the last step); dm is the (2,12) column-major
 * dimer block x y z vx vy vz fx fy fz (+3 unused) of sphere C (row 1) and N. */
int mpcx_sk_set_state(mpcx_sk* e, int64_t cstep, int64_t NT, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz, const double* fx, const double* fy,
                      const double* fz, const int64_t* flag, const double* dm);
/* Advance ncoll collision intervals of Main_SelkovDimer.f90:125-149 (Motion,
 * ForcesMD every step; Sort, Distribute, Affected_Cell, RMPC at collision
 * steps).  In rng_mode 0 `draws` holds ndraws successive genrand_real1()
 * values of the mt95 stream in the reference's call order; *used returns how
 * many the run took (the caller advances its stream by that many). */
int mpcx_sk_run(mpcx_sk* e, int64_t ncoll, const double* draws, int64_t ndraws, int64_t* used);
/* Download state (NULL skips an array); scalars: cstep, NT, draws used, near-set size. */
int mpcx_sk_get_state(mpcx_sk* e, double* x, double* y, double* z, double* vx, double* vy, double* vz, double* fx,
                      double* fy, double* fz, int64_t* flag, double* dm, int64_t* scalars, double* tpotential);
/* Cells of the last collision: (vol, 8) records vcm xyz, axis xyz, count; the
 * collision cell of each particle (-1: inactive); census counts (3, vol) F, G,
 * I; unaffected mask (vol). */
int mpcx_sk_get_cells(mpcx_sk* e, double* rec8, int32_t* cell_of, int32_t* census, uint8_t* unaffected);
/* Mean kernel durations (ms) over the last run (every 8th interval timed, at
 * most 256 of them): near engine, sk_bulk, sk_collide, sk_rotate; and the
 * run's wall time. */
int mpcx_sk_kernel_ms(mpcx_sk* e, double* ms4, double* run_ms);
/* on = 1 (the default): with device Philox draws, the far particles stream
 * and bin on a second stream while the near engine runs, and the near set is
 * binned after it; the results are bitwise those of on = 0 (the far
 * particles' per-step streaming does not depend on the near engine).  Host
 * draws keep the order of the reference's draw stream and never overlap.  No
 * counterpart in the reference. */
int mpcx_sk_set_overlap(mpcx_sk* e, int32_t on);

/* ------------------------------------------------------------------------
 * Diagnostics on the device-resident state (either engine)
 * ---------------------------------------------------------------------- */
/* out16: [0] particles counted, [1..3] per species (A B S / F G I), [4..6]
 * solvent momentum, [7..9] total momentum with the dimers, [10] solvent and
 * [11] dimer kinetic energy, [12] solvent temperature (KE/N)*(2./3.), [13]
 * dimer KE*(2./3.), [14] total energy with TPotential, [15] propulsion
 * velocity of dimer 0: centre-of-mass velocity along the N->C unit vector
 * (selkov_code/Output_SelkovDimer.f90:147-173).  2./3. is the reference's
 * default-real literal (0.6666666865...).  Collective: collective_code/
 * output_mod.f95:681-756, the dimers sphere by sphere.  Selkov:
 * Output_SelkovDimer.f90:243-300, the dimer as its centre of mass (momentum
 * (Mc+Mn)*Vdimav, kinetic energy of the centre of mass).  Solvent sums are
 * tree sums (the reference adds in particle order). */
int mpcx_co_observables(mpcx_co* e, double* out16);
int mpcx_sk_observables(mpcx_sk* e, double* out16);
/* A/B counts in the concentration slabs of collective_code/output_mod.f95:
 * 486-568 (bins ceiling(x/width); divide by the slab volume for densities):
 * A1d, B1d (nx), Axy, Bxy (nx,ny), Axz, Bxz (nx,nz), Ayz, Byz (ny,nz), each
 * column-major, with n? = int(box/width). */
int mpcx_co_conc_slabs(mpcx_co* e, double x_slab, double y_slab, double z_slab, double half_width,
                       double x_width, double y_width, double z_width, int32_t* counts);
/* F, G, I counts (3, 8) in the shells around the C sphere of
 * selkov_code/Output_Concentration_Catalytic.f90:50-156 (R^2 bounds 4, 9, 16,
 * 36, 64, 100, 144, 225, 400). */
int mpcx_sk_shells(mpcx_sk* e, int64_t* counts24);
/* Counts per (species, z bin, r bin), (3, nz, nr) row-major, in the body frame
 * of a dimer: origin at its C sphere, z along the N->C unit vector, r the
 * distance from that axis; z in [zmin, zmax), r in [0, rmax).  No reference
 * counterpart (its analogues are the shells above); the time average of these
 * fields trains the surrogate. */
int mpcx_co_body_density(mpcx_co* e, int32_t dimer, int32_t nr, int32_t nz, double rmax, double zmin, double zmax,
                         uint32_t* counts);
int mpcx_sk_body_density(mpcx_sk* e, int32_t nr, int32_t nz, double rmax, double zmin, double zmax,
                         uint32_t* counts);

/* ------------------------------------------------------------------------
 * MLP surrogate (no reference counterpart): parameters of a run -> the
 * time-averaged body-frame density field and propulsion velocity.  FP32
 * master weights; every GEMM runs on the MFMA units with BF16 operands and
 * FP32 accumulation (forward, both backward products); ReLU hidden layers,
 * linear output, weighted squared error, Adam.
 * ---------------------------------------------------------------------- */
typedef struct mpcx_mlp mpcx_mlp;
int mpcx_mlp_create(mpcx_mlp** out, int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out, uint64_t seed);
int mpcx_mlp_destroy(mpcx_mlp* e);
/* parameters, flat FP32: for each layer W (n_next x n_prev, row-major) then b */
int64_t mpcx_mlp_num_params(mpcx_mlp* e);
int mpcx_mlp_get_params(mpcx_mlp* e, float* flat);
int mpcx_mlp_set_params(mpcx_mlp* e, const float* flat);
int mpcx_mlp_forward(mpcx_mlp* e, int32_t rows, const float* x, float* y);
/* loss = sum_c wout[c] (y - target)^2 averaged over rows; grad (flat, may be NULL) */
int mpcx_mlp_loss_grad(mpcx_mlp* e, int32_t rows, const float* x, const float* target, const float* wout,
                       double* loss, float* grad);
int mpcx_mlp_train(mpcx_mlp* e, int64_t n, const float* x, const float* target, const float* wout, int32_t batch,
                   int32_t epochs, float lr, uint64_t seed, double* epoch_loss);
int mpcx_mlp_train_stats(mpcx_mlp* e, double* ms, double* gemm_flops);
/* C = A B (row-major, M x K and K x N) with BF16 operands on the MFMA units */
int mpcx_gemm_bf16(int32_t M, int32_t N, int32_t K, const float* A, const float* B, float* C, int32_t reps,
                   double* ms);

/* ------------------------------------------------------ restart files (host)
 * The reference's own restart files, byte for byte where the format is fixed:
 *  timestep.dat: main_pattern.f95:130-136 (read) / :243-249 (write): cstep,
 *    iTotA, iTotB, iTotS, one list-directed item per record;
 *  mt_state.dat: mtmod.f90:152-229 (mtsavef/mtgetf, form 'u'): sequential
 *    unformatted records mti(T) and mt(624,T), 4-byte record markers;
 *  motor_positions.dat: initial_mod.f95:79,118 / restart_mod.f95:410-414:
 *    motorCM(numdims,3) (column-major), one dimer per record.
 * List-directed output follows gfortran's layout; the readers take any
 * list-directed form.  mpcx_mt_state_read with mt or mti NULL only returns
 * the stream count. */
int mpcx_timestep_write(const char* path, int32_t cstep, int64_t n_a, int64_t n_b, int64_t n_s);
int mpcx_timestep_read(const char* path, int32_t* cstep, int64_t* n_a, int64_t* n_b, int64_t* n_s);
int mpcx_mt_state_write(const char* path, const uint32_t* mt, const int32_t* mti, int32_t nstreams);
int mpcx_mt_state_read(const char* path, uint32_t* mt, int32_t* mti, int32_t max_streams, int32_t* nstreams);
int mpcx_motor_positions_write(const char* path, const double* cm, int32_t nd);
int mpcx_motor_positions_read(const char* path, double* cm, int32_t nd);
/* One record of fixed-form output with I (Iw), E (Ew.d), S (ESw.d) and X (nX)
 * edit descriptors, e.g. "I8 X1 E23.16 X1 E23.16 X1 E23.16 X1" for the
 * selkov_code dimer records (Output_SelkovDimer.f90:26) or "S23.16 X1 ... I8"
 * for its solvent records (:139); values in order, integers as whole doubles.
 * Writes a NUL-terminated line (no newline) into out[cap]. */
int mpcx_format_record(const char* spec, const double* values, int32_t nvalues, char* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
