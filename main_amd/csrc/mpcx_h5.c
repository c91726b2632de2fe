/* data.h5 of collective_code through the HDF5 C API (host only; built into
 * its own library, libmpcx_h5.so, so the engine library does not depend on
 * HDF5).
 *
 *   mpcx_h5_initiate      output_mod.f95:96-264  (output_hdf5_initiate)
 *   mpcx_h5_write         output_mod.f95:324-898 (output_hdf5_write)
 *   mpcx_h5_read_restart  restart_mod.f95:359-402 (the data.h5 frames a restart reads)
 *
 * The reference writes through HDF5's Fortran interface, which names
 * dataspace dimensions in column-major order; the C interface names them in
 * row-major order, so every shape below is the Fortran shape reversed (a
 * Fortran (NT, frames) dataset is C {frames, NT}) and the bytes in the file are
 * the same.  Names lose the Fortran strings' trailing blanks, as the Fortran
 * interface trims them.  Values are computed from the caller's host arrays in
 * the reference's loop order, so sums round as the reference's do.
 *
 * One quirk is kept on purpose: output_hdf5_initiate sizes the 2-D
 * concentration datasets int(box*width) (box/width elsewhere), so frame 0 of
 * each holds the leading int(Lx*w)*int(Ly*w) values of the field in memory
 * order, and later frames, after the extent grows to box/width, the whole
 * field (output_mod.f95:123-126 against :592-595).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hdf5.h>

#include "../../include/mpcx_h5.h"

static char g_err[256];

static int fail(const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return MPCX_H5_E;
}

const char* mpcx_h5_last_error(void) { return g_err; }

/* an attribute of rank 1 or 2 written from memory of type mtype */
static int put_attr(hid_t loc, const char* name, hid_t ftype, hid_t mtype, int rank, const hsize_t* dims,
                    const void* buf) {
  hid_t sp = H5Screate_simple(rank, dims, NULL);
  hid_t a = sp < 0 ? -1 : H5Acreate2(loc, name, ftype, sp, H5P_DEFAULT, H5P_DEFAULT);
  herr_t w = a < 0 ? -1 : H5Awrite(a, mtype, buf);
  if (a >= 0) H5Aclose(a);
  if (sp >= 0) H5Sclose(sp);
  return (sp < 0 || a < 0 || w < 0) ? fail("attribute creation failed") : 0;
}

/* an extendible chunked dataset of the given initial dims */
static int make_dset(hid_t loc, const char* name, hid_t type, int rank, const hsize_t* dims, const hsize_t* chunk) {
  hsize_t maxd[3] = {H5S_UNLIMITED, H5S_UNLIMITED, H5S_UNLIMITED};
  hid_t sp = H5Screate_simple(rank, dims, maxd);
  hid_t pl = H5Pcreate(H5P_DATASET_CREATE);
  herr_t c = (sp < 0 || pl < 0) ? -1 : H5Pset_chunk(pl, rank, chunk);
  hid_t d = c < 0 ? -1 : H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, pl, H5P_DEFAULT);
  if (d >= 0) H5Dclose(d);
  if (pl >= 0) H5Pclose(pl);
  if (sp >= 0) H5Sclose(sp);
  return d < 0 ? fail("dataset creation failed") : 0;
}

int mpcx_h5_initiate(const char* path, int32_t Freq1, int32_t Freq2, int32_t Freq3, int32_t Freq4, const double* box,
                     double T, double MDt, double MPCt, int32_t NT, int32_t numdims, double x_width, double y_width,
                     double z_width) {
  if (!path || !box || NT < 1 || numdims < 0 || Freq1 < 1 || Freq2 < 1 || Freq3 < 1 || Freq4 < 1)
    return fail("bad arguments");
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) return fail("cannot create the file");
  int rc = 0;
  const hsize_t one = 1, box_dims[2] = {1, 3};
  rc = rc ? rc : put_attr(f, "simulation_box_length", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, 2, box_dims, box);
  rc = rc ? rc : put_attr(f, "solvent_temperature", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, 1, &one, &T);
  rc = rc ? rc : put_attr(f, "MD_time", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, 1, &one, &MDt);
  rc = rc ? rc : put_attr(f, "MPC_time", H5T_NATIVE_DOUBLE, H5T_NATIVE_DOUBLE, 1, &one, &MPCt);
  /* created as doubles, written from integers (output_mod.f95:142-152) */
  rc = rc ? rc : put_attr(f, "total_#_particles", H5T_NATIVE_DOUBLE, H5T_NATIVE_INT, 1, &one, &NT);
  rc = rc ? rc : put_attr(f, "number_of_dimers", H5T_NATIVE_DOUBLE, H5T_NATIVE_INT, 1, &one, &numdims);
  static const char* groups[4] = {"Solvent", "Concentration", "System", "Dimer"};
  hid_t g[4] = {-1, -1, -1, -1};
  for (int i = 0; i < (numdims > 0 ? 4 : 3) && !rc; i++) {
    g[i] = H5Gcreate2(f, groups[i], H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (g[i] < 0) rc = fail("group creation failed");
  }
  /* System */
  if (!rc) rc = put_attr(g[2], "io_frequency", H5T_NATIVE_INT, H5T_NATIVE_INT, 1, &one, &Freq3);
  static const char* sys[6] = {"pxsol", "pysol", "pzsol", "total_energy", "solvent_temperature",
                               "dimer_temperature"};
  const hsize_t c75k = 75000;
  for (int i = 0; i < 6 && !rc; i++) rc = make_dset(g[2], sys[i], H5T_NATIVE_DOUBLE, 1, &one, &c75k);
  /* Concentration */
  if (!rc) rc = put_attr(g[1], "io_frequency", H5T_NATIVE_INT, H5T_NATIVE_INT, 1, &one, &Freq2);
  static const char* conc[12] = {"total_A", "total_B", "concentration_A", "concentration_B",
                                 "1D_concentration_A", "1D_concentration_B", "2D_concentration_A_xy",
                                 "2D_concentration_B_xy", "2D_concentration_A_xz", "2D_concentration_B_xz",
                                 "2D_concentration_A_yz", "2D_concentration_B_yz"};
  const hsize_t n1d = (hsize_t)(int)(box[0] * 2.0f);
  const hsize_t qx = (hsize_t)(int)(box[0] * x_width), qy = (hsize_t)(int)(box[1] * y_width),
                qz = (hsize_t)(int)(box[2] * z_width);
  const hsize_t d1[2] = {1, n1d}, ch1[2] = {800, 120}, ch2[3] = {7, 120, 120};
  const hsize_t dxy[3] = {1, qy, qx}, dxz[3] = {1, qz, qx}, dyz[3] = {1, qz, qy};
  for (int i = 0; i < 12 && !rc; i++) {
    if (i < 4) rc = make_dset(g[1], conc[i], H5T_NATIVE_DOUBLE, 1, &one, &c75k);
    else if (i < 6) rc = make_dset(g[1], conc[i], H5T_NATIVE_DOUBLE, 2, d1, ch1);
    else if (i < 8) rc = make_dset(g[1], conc[i], H5T_NATIVE_DOUBLE, 3, dxy, ch2);
    else if (i < 10) rc = make_dset(g[1], conc[i], H5T_NATIVE_DOUBLE, 3, dxz, ch2);
    else rc = make_dset(g[1], conc[i], H5T_NATIVE_DOUBLE, 3, dyz, ch2);
  }
  /* Solvent */
  if (!rc) rc = put_attr(g[0], "io_frequency", H5T_NATIVE_INT, H5T_NATIVE_INT, 1, &one, &Freq1);
  static const char* solv[7] = {"position_x", "position_y", "position_z", "flag", "velocity_x", "velocity_y",
                                "velocity_z"};
  const hsize_t ds[2] = {1, (hsize_t)NT}, cs[2] = {1, 100000};
  for (int i = 0; i < 7 && !rc; i++)
    rc = make_dset(g[0], solv[i], i == 3 ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE, 2, ds, cs);
  /* Dimer */
  if (!rc && numdims > 0) {
    rc = put_attr(g[3], "io_frequency", H5T_NATIVE_INT, H5T_NATIVE_INT, 1, &one, &Freq4);
    static const char* dim[7] = {"position_x", "position_y", "position_z", "velocity_x", "velocity_y",
                                 "velocity_z", "nucvec"};
    const hsize_t nd = (hsize_t)numdims;
    const hsize_t dd[3] = {1, 2, nd}, dn[3] = {1, 3, nd};
    const hsize_t cd[3] = {(hsize_t)floor(100000.0f / (float)(numdims * 2)), 2, nd};
    const hsize_t cn[3] = {(hsize_t)floor(100000.0f / (float)(3 * numdims)), 3, nd};
    for (int i = 0; i < 7 && !rc; i++)
      rc = make_dset(g[3], dim[i], H5T_NATIVE_DOUBLE, 3, i < 6 ? dd : dn, i < 6 ? cd : cn);
  }
  for (int i = 0; i < 4; i++)
    if (g[i] >= 0) H5Gclose(g[i]);
  if (H5Fclose(f) < 0 && !rc) rc = fail("closing the file failed");
  return rc;
}

/* frame `frame` of dataset loc/name: at frame 0 the whole (initial) extent
 * from buf; later the extent grows to `ext` (frames first) and the hyperslab
 * {frame, 0..} of `count` receives buf */
static int put_frame(hid_t loc, const char* name, hid_t mtype, int rank, const hsize_t* ext, const hsize_t* count,
                     int64_t frame, const void* buf) {
  hid_t d = H5Dopen2(loc, name, H5P_DEFAULT);
  if (d < 0) return fail("dataset missing");
  herr_t w;
  if (frame == 0) {
    w = H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
  } else {
    hsize_t off[3] = {(hsize_t)frame, 0, 0};
    w = H5Dset_extent(d, ext);
    hid_t fs = w < 0 ? -1 : H5Dget_space(d);
    hid_t ms = fs < 0 ? -1 : H5Screate_simple(rank, count, NULL);
    w = ms < 0 ? -1 : H5Sselect_hyperslab(fs, H5S_SELECT_SET, off, NULL, count, NULL);
    if (w >= 0) w = H5Dwrite(d, mtype, ms, fs, H5P_DEFAULT, buf);
    if (ms >= 0) H5Sclose(ms);
    if (fs >= 0) H5Sclose(fs);
  }
  H5Dclose(d);
  return w < 0 ? fail("dataset write failed") : 0;
}

static double pbc_half(double d, double L) {
  if (d > 0.5 * L) return d - L;
  if (d < -0.5 * L) return d + L;
  return d;
}

int mpcx_h5_write(const char* path, int32_t cstep, int64_t Vol, int32_t Freq1, int32_t Freq2, int32_t Freq3,
                  int32_t Freq4, int32_t NT, int32_t numdims, double Ma, double Mb, double Ms, double Mc, double Mn,
                  double x_conc_slab, double y_conc_slab, double z_conc_slab, const double* box, double x_width,
                  double y_width, double z_width, const double* xsol, const double* ysol, const double* zsol,
                  const int32_t* flag, const double* vxsol, const double* vysol, const double* vzsol,
                  double cutoffC, double Dcn, const double* xdim, const double* ydim, const double* zdim,
                  const double* vxdim, const double* vydim, const double* vzdim, double tpotential) {
  if (!path || !box || !xsol || !ysol || !zsol || !flag || !vxsol || !vysol || !vzsol || NT < 1 || numdims < 0 ||
      cstep < 0 || Freq1 < 1 || Freq2 < 1 || Freq3 < 1 || Freq4 < 1 ||
      (numdims > 0 && (!xdim || !ydim || !zdim || !vxdim || !vydim || !vzdim)))
    return fail("bad arguments");
  hid_t f = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) return fail("cannot open the file (mpcx_h5_initiate creates it)");
  int rc = 0;
  const int nt = NT, nd = numdims;
  /* Solvent: output_mod.f95:455-480 */
  if (cstep % Freq1 == 0) {
    hid_t g = H5Gopen2(f, "Solvent", H5P_DEFAULT);
    const int64_t fr = cstep / Freq1;
    const hsize_t ext[2] = {(hsize_t)fr + 1, (hsize_t)nt}, cnt[2] = {1, (hsize_t)nt};
    static const char* names[7] = {"position_x", "position_y", "position_z", "flag", "velocity_x", "velocity_y",
                                   "velocity_z"};
    const void* bufs[7] = {xsol, ysol, zsol, flag, vxsol, vysol, vzsol};
    if (g < 0) rc = fail("group Solvent missing");
    for (int i = 0; i < 7 && !rc; i++)
      rc = put_frame(g, names[i], i == 3 ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE, 2, ext, cnt, fr, bufs[i]);
    if (g >= 0) H5Gclose(g);
  }
  /* Concentration: output_mod.f95:485-678 */
  if (!rc && cstep % Freq2 == 0) {
    double totA = 0.0, totB = 0.0;
    for (int i = 0; i < nt; i++) {
      if (flag[i] == 0) totA = totA + 1.0;
      else if (flag[i] == 1) totB = totB + 1.0;
    }
    const double concA = totA / (double)Vol, concB = totB / (double)Vol;
    const int nx = (int)(box[0] / x_width), ny = (int)(box[1] / y_width), nz = (int)(box[2] / z_width);
    double* a1 = calloc((size_t)nx, sizeof(double));
    double* b1 = calloc((size_t)nx, sizeof(double));
    double* axy = calloc((size_t)nx * ny, sizeof(double));
    double* bxy = calloc((size_t)nx * ny, sizeof(double));
    double* axz = calloc((size_t)nx * nz, sizeof(double));
    double* bxz = calloc((size_t)nx * nz, sizeof(double));
    double* ayz = calloc((size_t)ny * nz, sizeof(double));
    double* byz = calloc((size_t)ny * nz, sizeof(double));
    if (!a1 || !b1 || !axy || !bxy || !axz || !bxz || !ayz || !byz) rc = fail("out of host memory");
    const double c2 = 2.0 * cutoffC;
    /* bins j = ceiling(x/width), 1-based; a particle at a bin edge of 0 has
     * none (the reference would index element 0) and is left out */
#define BIN(v, w, n) ((int)ceil((v) / (w)))
    for (int i = 0; i < nt && !rc; i++) {
      if (flag[i] != 0 && flag[i] != 1) continue;
      double* A1 = flag[i] == 0 ? a1 : b1;
      double* Axy = flag[i] == 0 ? axy : bxy;
      double* Axz = flag[i] == 0 ? axz : bxz;
      double* Ayz = flag[i] == 0 ? ayz : byz;
      if (zsol[i] <= z_conc_slab + cutoffC && zsol[i] >= z_conc_slab - cutoffC) {
        const int j = BIN(xsol[i], x_width, nx), k = BIN(ysol[i], y_width, ny);
        if (j >= 1 && j <= nx) A1[j - 1] = A1[j - 1] + 1.0;
        if (j >= 1 && j <= nx && k >= 1 && k <= ny) Axy[(j - 1) + (size_t)nx * (k - 1)] += 1.0;
      }
      if (ysol[i] <= y_conc_slab + cutoffC && ysol[i] >= y_conc_slab - cutoffC) {
        const int j = BIN(xsol[i], x_width, nx), k = BIN(zsol[i], z_width, nz);
        if (j >= 1 && j <= nx && k >= 1 && k <= nz) Axz[(j - 1) + (size_t)nx * (k - 1)] += 1.0;
      }
      if (xsol[i] <= x_conc_slab + cutoffC && xsol[i] >= x_conc_slab - cutoffC) {
        const int j = BIN(ysol[i], y_width, ny), k = BIN(zsol[i], z_width, nz);
        if (j >= 1 && j <= ny && k >= 1 && k <= nz) Ayz[(j - 1) + (size_t)ny * (k - 1)] += 1.0;
      }
    }
#undef BIN
    if (!rc) {
      const double v1 = (x_width * box[1]) * c2, vxy = (x_width * y_width) * c2, vxz = (x_width * z_width) * c2,
                   vyz = (y_width * z_width) * c2;
      for (int j = 0; j < nx; j++) { a1[j] = a1[j] / v1; b1[j] = b1[j] / v1; }
      for (size_t j = 0; j < (size_t)nx * ny; j++) { axy[j] = axy[j] / vxy; bxy[j] = bxy[j] / vxy; }
      for (size_t j = 0; j < (size_t)nx * nz; j++) { axz[j] = axz[j] / vxz; bxz[j] = bxz[j] / vxz; }
      for (size_t j = 0; j < (size_t)ny * nz; j++) { ayz[j] = ayz[j] / vyz; byz[j] = byz[j] / vyz; }
      hid_t g = H5Gopen2(f, "Concentration", H5P_DEFAULT);
      const int64_t fr = cstep / Freq2;
      const hsize_t e0[1] = {(hsize_t)fr + 1}, c0[1] = {1};
      const hsize_t e1[2] = {(hsize_t)fr + 1, (hsize_t)nx}, cc1[2] = {1, (hsize_t)nx};
      const hsize_t exy[3] = {(hsize_t)fr + 1, (hsize_t)ny, (hsize_t)nx}, cxy[3] = {1, (hsize_t)ny, (hsize_t)nx};
      const hsize_t exz[3] = {(hsize_t)fr + 1, (hsize_t)nz, (hsize_t)nx}, cxz[3] = {1, (hsize_t)nz, (hsize_t)nx};
      const hsize_t eyz[3] = {(hsize_t)fr + 1, (hsize_t)nz, (hsize_t)ny}, cyz[3] = {1, (hsize_t)nz, (hsize_t)ny};
      if (g < 0) rc = fail("group Concentration missing");
      rc = rc ? rc : put_frame(g, "total_A", H5T_NATIVE_DOUBLE, 1, e0, c0, fr, &totA);
      rc = rc ? rc : put_frame(g, "total_B", H5T_NATIVE_DOUBLE, 1, e0, c0, fr, &totB);
      rc = rc ? rc : put_frame(g, "concentration_A", H5T_NATIVE_DOUBLE, 1, e0, c0, fr, &concA);
      rc = rc ? rc : put_frame(g, "concentration_B", H5T_NATIVE_DOUBLE, 1, e0, c0, fr, &concB);
      rc = rc ? rc : put_frame(g, "1D_concentration_A", H5T_NATIVE_DOUBLE, 2, e1, cc1, fr, a1);
      rc = rc ? rc : put_frame(g, "1D_concentration_B", H5T_NATIVE_DOUBLE, 2, e1, cc1, fr, b1);
      rc = rc ? rc : put_frame(g, "2D_concentration_A_xy", H5T_NATIVE_DOUBLE, 3, exy, cxy, fr, axy);
      rc = rc ? rc : put_frame(g, "2D_concentration_B_xy", H5T_NATIVE_DOUBLE, 3, exy, cxy, fr, bxy);
      rc = rc ? rc : put_frame(g, "2D_concentration_A_xz", H5T_NATIVE_DOUBLE, 3, exz, cxz, fr, axz);
      rc = rc ? rc : put_frame(g, "2D_concentration_B_xz", H5T_NATIVE_DOUBLE, 3, exz, cxz, fr, bxz);
      rc = rc ? rc : put_frame(g, "2D_concentration_A_yz", H5T_NATIVE_DOUBLE, 3, eyz, cyz, fr, ayz);
      rc = rc ? rc : put_frame(g, "2D_concentration_B_yz", H5T_NATIVE_DOUBLE, 3, eyz, cyz, fr, byz);
      if (g >= 0) H5Gclose(g);
    }
    free(a1); free(b1); free(axy); free(bxy); free(axz); free(bxz); free(ayz); free(byz);
  }
  /* System: output_mod.f95:680-795 */
  if (!rc && cstep % Freq3 == 0) {
    double sumpx = 0.0, sumpy = 0.0, sumpz = 0.0, kedim = 0.0, kesol = 0.0, totA = 0.0, totB = 0.0, totS = 0.0;
    for (int i = 0; i < nt; i++) {
      double m = 0.0;
      if (flag[i] == 0) { totA = totA + 1.0; m = Ma; }
      else if (flag[i] == 1) { totB = totB + 1.0; m = Mb; }
      else if (flag[i] == 2) { totS = totS + 1.0; m = Ms; }
      else continue;   /* PXsol keeps its old value in the reference; no other flag occurs */
      sumpx = sumpx + m * vxsol[i];
      sumpy = sumpy + m * vysol[i];
      sumpz = sumpz + m * vzsol[i];
    }
    for (int k = 0; k < nd; k++) {
      const double px1 = Mc * vxdim[k], py1 = Mc * vydim[k], pz1 = Mc * vzdim[k];
      const double px2 = Mn * vxdim[k + nd], py2 = Mn * vydim[k + nd], pz2 = Mn * vzdim[k + nd];
      sumpx = (sumpx + px1) + px2;
      sumpy = (sumpy + py1) + py2;
      sumpz = (sumpz + pz1) + pz2;
      kedim = (kedim + (0.5 * (1.0 / Mc)) * ((px1 * px1 + py1 * py1) + pz1 * pz1)) +
              (0.5 * (1.0 / Mn)) * ((px2 * px2 + py2 * py2) + pz2 * pz2);
    }
    for (int i = 0; i < nt; i++) {
      double m = flag[i] == 0 ? Ma : flag[i] == 1 ? Mb : flag[i] == 2 ? Ms : 0.0;
      kesol = kesol + 0.5 * ((vxsol[i] * (m * vxsol[i]) + vysol[i] * (m * vysol[i])) + vzsol[i] * (m * vzsol[i]));
    }
    const double two_thirds = (double)(2.0f / 3.0f);   /* the reference's default-real 2./3. */
    const double ekin = nd > 0 ? kedim + kesol : kesol;
    const double tot_energy = ekin + tpotential;
    const double tempsol = (kesol / ((totA + totB) + totS)) * two_thirds;
    const double tempdim = nd > 0 ? kedim * two_thirds : 0.0;
    hid_t g = H5Gopen2(f, "System", H5P_DEFAULT);
    const int64_t fr = cstep / Freq3;
    const hsize_t e0[1] = {(hsize_t)fr + 1}, c0[1] = {1};
    const double* vals[6] = {&sumpx, &sumpy, &sumpz, &tot_energy, &tempsol, &tempdim};
    static const char* names[6] = {"pxsol", "pysol", "pzsol", "total_energy", "solvent_temperature",
                                   "dimer_temperature"};
    if (g < 0) rc = fail("group System missing");
    for (int i = 0; i < 6 && !rc; i++) rc = put_frame(g, names[i], H5T_NATIVE_DOUBLE, 1, e0, c0, fr, vals[i]);
    if (g >= 0) H5Gclose(g);
  }
  /* Dimer: output_mod.f95:797-896 */
  if (!rc && nd > 0 && cstep % Freq4 == 0) {
    double* nuc = malloc(sizeof(double) * 3 * (size_t)nd);   /* (numdims, 3) column-major */
    if (!nuc) rc = fail("out of host memory");
    for (int j = 0; j < nd && !rc; j++) {
      const double d0 = pbc_half(xdim[j] - xdim[j + nd], box[0]);
      const double d1 = pbc_half(ydim[j] - ydim[j + nd], box[1]);
      const double d2 = zdim[j] - zdim[j + nd];
      nuc[j] = d0 / Dcn; nuc[j + nd] = d1 / Dcn; nuc[j + 2 * nd] = d2 / Dcn;
    }
    hid_t g = rc ? -1 : H5Gopen2(f, "Dimer", H5P_DEFAULT);
    const int64_t fr = cstep / Freq4;
    const hsize_t ed[3] = {(hsize_t)fr + 1, 2, (hsize_t)nd}, cd[3] = {1, 2, (hsize_t)nd};
    const hsize_t en[3] = {(hsize_t)fr + 1, 3, (hsize_t)nd}, cn[3] = {1, 3, (hsize_t)nd};
    static const char* names[6] = {"position_x", "position_y", "position_z", "velocity_x", "velocity_y",
                                   "velocity_z"};
    const double* bufs[6] = {xdim, ydim, zdim, vxdim, vydim, vzdim};
    if (!rc && g < 0) rc = fail("group Dimer missing");
    for (int i = 0; i < 6 && !rc; i++) rc = put_frame(g, names[i], H5T_NATIVE_DOUBLE, 3, ed, cd, fr, bufs[i]);
    if (!rc) rc = put_frame(g, "nucvec", H5T_NATIVE_DOUBLE, 3, en, cn, fr, nuc);
    if (g >= 0) H5Gclose(g);
    free(nuc);
  }
  if (H5Fclose(f) < 0 && !rc) rc = fail("closing the file failed");
  return rc;
}

/* hyperslab {frame, 0..} of `count` from loc/name into buf */
static int get_frame(hid_t loc, const char* name, hid_t mtype, int rank, const hsize_t* count, int64_t frame,
                     void* buf) {
  hid_t d = H5Dopen2(loc, name, H5P_DEFAULT);
  if (d < 0) return fail("dataset missing");
  hid_t fs = H5Dget_space(d);
  hsize_t dims[3] = {0, 0, 0};
  int r = fs < 0 ? -1 : H5Sget_simple_extent_dims(fs, dims, NULL);
  int rc = 0;
  if (r != rank || (hsize_t)frame >= dims[0]) rc = fail("the frame is not in the file");
  for (int i = 1; i < rank && !rc; i++)
    if (dims[i] != count[i]) rc = fail("the file's dataset shape differs from the system's");
  if (!rc) {
    const hsize_t off[3] = {(hsize_t)frame, 0, 0};
    hid_t ms = H5Screate_simple(rank, count, NULL);
    herr_t w = ms < 0 ? -1 : H5Sselect_hyperslab(fs, H5S_SELECT_SET, off, NULL, count, NULL);
    if (w >= 0) w = H5Dread(d, mtype, ms, fs, H5P_DEFAULT, buf);
    if (ms >= 0) H5Sclose(ms);
    if (w < 0) rc = fail("dataset read failed");
  }
  if (fs >= 0) H5Sclose(fs);
  H5Dclose(d);
  return rc;
}

int mpcx_h5_read_restart(const char* path, int32_t cstep, int32_t Freq1, int32_t Freq4, int32_t NT, int32_t numdims,
                         double* xsol, double* ysol, double* zsol, int32_t* flag, double* vxsol, double* vysol,
                         double* vzsol, double* xdim, double* ydim, double* zdim, double* vxdim, double* vydim,
                         double* vzdim) {
  if (!path || NT < 1 || numdims < 0 || cstep < 0 || Freq1 < 1 || Freq4 < 1 || !xsol || !ysol || !zsol || !flag ||
      !vxsol || !vysol || !vzsol || (numdims > 0 && (!xdim || !ydim || !zdim || !vxdim || !vydim || !vzdim)))
    return fail("bad arguments");
  hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return fail("cannot open the file");
  int rc = 0;
  hid_t g = H5Gopen2(f, "Solvent", H5P_DEFAULT);
  const hsize_t cnt[2] = {1, (hsize_t)NT};
  static const char* sn[7] = {"position_x", "position_y", "position_z", "flag", "velocity_x", "velocity_y",
                              "velocity_z"};
  void* sb[7] = {xsol, ysol, zsol, flag, vxsol, vysol, vzsol};
  if (g < 0) rc = fail("group Solvent missing");
  for (int i = 0; i < 7 && !rc; i++)
    rc = get_frame(g, sn[i], i == 3 ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE, 2, cnt, cstep / Freq1, sb[i]);
  if (g >= 0) H5Gclose(g);
  if (!rc && numdims > 0) {
    g = H5Gopen2(f, "Dimer", H5P_DEFAULT);
    const hsize_t cd[3] = {1, 2, (hsize_t)numdims};
    static const char* dn[6] = {"position_x", "position_y", "position_z", "velocity_x", "velocity_y", "velocity_z"};
    void* db[6] = {xdim, ydim, zdim, vxdim, vydim, vzdim};
    if (g < 0) rc = fail("group Dimer missing");
    for (int i = 0; i < 6 && !rc; i++) rc = get_frame(g, dn[i], H5T_NATIVE_DOUBLE, 3, cd, cstep / Freq4, db[i]);
    if (g >= 0) H5Gclose(g);
  }
  H5Fclose(f);
  return rc;
}
