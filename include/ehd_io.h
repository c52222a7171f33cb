/*
 * ehd_io.h -- host-side output and restart files of the EHD driver.
 *
 * The reference writes these files from its host wrappers after copying the
 * device fields back; here the copy is ehd_read_fields (ehd.h) and the
 * formatting is a separate host-only layer, so that the file formats can be
 * checked without a GPU.  All functions take host arrays in the reference
 * layout ([z][y][x], x fastest); fields9 is the 9*S block rho, q, phi, u, v,
 * w, Ex, Ey, Ez (the order ehd_init takes).  The wall rows are linearly
 * extrapolated (2*v[1] - v[2], 2*v[nz-2] - v[nz-3]) on a private copy exactly
 * where the reference does it, the caller's arrays are not modified.
 * Return 0 or EHD_EINVAL (ehd.h) with ehd_last_error() set.
 */
#ifndef EHD_IO_H
#define EHD_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* One Tecplot zone, save_data_tecplot (LBM.cu:1427-1503): with first != 0
 * the VARIABLES line, then a blank line, the ZONE header and one
 * "%g x8 %10.6f x4" row per node (x, y, z, u, v, w, p=rho, charge, phi, Ex,
 * Ey, Ez).  rho, charge, u, v, w are extrapolated at both walls.  append = 0
 * truncates the file ("wb+" of main.cu:167), 1 appends (the open stream). */
int ehd_io_tecplot(const char* path, int append, int first, double t, int nx, int ny, int nz,
                   double dx, double dy, double dz, const double* fields9);

/* data_end.dat, save_data_end (LBM.cu:1505-1566): one row of ten %10.6f per
 * node: t, u, v, w, rho, charge, phi, Ex, Ey, Ez (same wall extrapolation). */
int ehd_io_data_end(const char* path, double t, int nx, int ny, int nz, const double* fields9);

/* read_data (LBM.cu:1596-1643): parses data_end.dat back into fields9 and the
 * time of its last row.  The reference reads unchecked; a short or malformed
 * file is an error here. */
int ehd_io_read_end(const char* path, int nx, int ny, int nz, double* fields9, double* t);

/* "<t>.dat", save_data_dmd (LBM.cu:1568-1594): "%10.6f %10.6f" rows of t and
 * w with w extrapolated at both walls.  The file name is "%g.dat" of t
 * (main.cu:214). */
int ehd_io_dmd(const char* path, double t, int nx, int ny, int nz, const double* uz);

/* One "%10.6f %10.6f" row of umax.dat, record_umax (LBM.cu:1673-1713): the
 * maximum of w over the box (starting from 0) after extrapolating w at the
 * top wall only.  append as in ehd_io_tecplot. */
int ehd_io_umax(const char* path, int append, double t, int nx, int ny, int nz, const double* uz);

/* Host collector current, current() (LBM.cu:1646-1671): q extrapolated at the
 * top wall, I = K dy^2 sum_xy q(x,y,nz-1) Ez(x,y,nz-1), summed y-major in
 * FP64 as the reference loops.  The device reduction of ehd_diagnostics gives
 * the same value up to summation order. */
double ehd_io_current(int nx, int ny, int nz, double K, double dy, const double* q, const double* ez);

#ifdef __cplusplus
}
#endif
#endif
