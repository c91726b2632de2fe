/* data.h5 of collective_code (output_mod.f95 / restart_mod.f95) through the
 * HDF5 C API: main_amd/libmpcx_h5.so, a host library of its own (it links the
 * HDF5 1.10 library under /opt/conda), so libmpcx.so keeps no HDF5 dependency.
 * Arrays are the reference's: solvent arrays of NT entries, flag 32-bit,
 * dimer arrays (numdims, 2) column-major (sphere C in column 1, N in 2).
 * Every call opens and closes the file; a nonzero return is MPCX_H5_E with
 * the message in mpcx_h5_last_error(). */
#ifndef MPCX_H5_H
#define MPCX_H5_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define MPCX_H5_E 7

const char* mpcx_h5_last_error(void);

/* Replaces output_hdf5_initiate (collective_code/output_mod.f95:96-264):
 * creates (truncates) path with the file attributes, the groups Solvent,
 * Concentration, System and (numdims > 0) Dimer, their io_frequency
 * attributes and their extendible chunked datasets. */
int mpcx_h5_initiate(const char* path, int32_t Freq1, int32_t Freq2, int32_t Freq3, int32_t Freq4, const double* box,
                     double T, double MDt, double MPCt, int32_t NT, int32_t numdims, double x_width, double y_width,
                     double z_width);

/* Replaces output_hdf5_write (output_mod.f95:324-898): the frames due at
 * cstep (solvent every Freq1 steps, concentrations every Freq2, system sums
 * every Freq3, dimers every Freq4), computed from the arrays in the
 * reference's loop order. */
int mpcx_h5_write(const char* path, int32_t cstep, int64_t Vol, int32_t Freq1, int32_t Freq2, int32_t Freq3,
                  int32_t Freq4, int32_t NT, int32_t numdims, double Ma, double Mb, double Ms, double Mc, double Mn,
                  double x_conc_slab, double y_conc_slab, double z_conc_slab, const double* box, double x_width,
                  double y_width, double z_width, const double* xsol, const double* ysol, const double* zsol,
                  const int32_t* flag, const double* vxsol, const double* vysol, const double* vzsol,
                  double cutoffC, double Dcn, const double* xdim, const double* ydim, const double* zdim,
                  const double* vxdim, const double* vydim, const double* vzdim, double tpotential);

/* Replaces the data.h5 reads of restart_mod.f95:359-402: the solvent of
 * frame cstep/Freq1 and the dimers of frame cstep/Freq4. */
int mpcx_h5_read_restart(const char* path, int32_t cstep, int32_t Freq1, int32_t Freq4, int32_t NT, int32_t numdims,
                         double* xsol, double* ysol, double* zsol, int32_t* flag, double* vxsol, double* vysol,
                         double* vzsol, double* xdim, double* ydim, double* zdim, double* vxdim, double* vydim,
                         double* vzdim);

#ifdef __cplusplus
}
#endif
#endif
