/* gamera_io.h -- C-ABI of libgamera_io.so: the reference's HDF5 file formats
 * for a GAMERA block, read and written on the host.
 *
 *   corner grid:  datasets X, Y, Z of the (Ni+1, Nj+1, Nk+1) ghost-inclusive
 *                 corners; the active cell counts are each corner count
 *                 minus 2*nG+1 (reference src/gamera/gioH5.F90:36-124
 *                 readH5Grid, src/base/ioH5.F90:166-200 GridSizeH5)
 *   restart:      RunID.gam.Res.#####.h5 (src/gamera/output.F90:147-173):
 *                 root attributes nOut, nRes, ts (int), dt, dt0, t, ot (double),
 *                 datasets X, Y, Z, Gas, magFlux, Bxyz, oGas, omagFlux, oBxyz
 *                 (src/gamera/gioH5.F90:650-704 writeH5Res, :707-872
 *                 readH5Restart)
 *   bcwind:       the solar-wind time series (src/gamera/apps/wind.F90:617-760
 *                 readWind), below
 *
 * Arrays are Fortran order (i fastest) with the ghost cells, the layout of
 * gam_put_state/gam_get_state: xyz (Ni+1, Nj+1, Nk+1, 3), Gas (Ni, Nj, Nk, 5),
 * magFlux (Ni+1, Nj+1, Nk+1, 3), Bxyz (Ni, Nj, Nk, 3), FP64.  A file's
 * datasets have the reference's Fortran shapes, so in HDF5's C view their
 * dimensions are reversed (Gas: 1 x 5 x Nk x Nj x Ni, the leading 1 the
 * species axis BLK:nSpc with nSpc = 0).
 *
 * Every call returns 0 on success, < 0 on an error, with a message in
 * gam_io_last_error() (per thread).  Nothing here touches a GPU.
 */
#ifndef GAMERA_IO_H
#define GAMERA_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* Model%IO counters and the clock a restart carries */
typedef struct {
    int nOut, nRes, ts;
    double t;      /* Model%t = State%time */
    double dt;     /* Model%dt */
    double dt0;    /* Model%dt0 (0 when below 10*TINY) */
    double ot;     /* oState%time */
} gam_io_clock;

/* Active cells (Nip, Njp, Nkp) of the grid in file path: the dimensions of
 * its X dataset minus 2*nG+1 (gioH5.F90:44-52).  A rank-2 X (a 2-D grid) has
 * no k corners: dims[2] is then -(2*nG+1)+1, which the caller rejects. */
int gam_io_grid_dims(const char *path, int nG, int *dims);

/* The corners X, Y, Z into xyz (gioH5.F90:75-123, the whole grid): the file's
 * active dims must be (Nip, Njp, Nkp). */
int gam_io_read_grid(const char *path, int nG, int Nip, int Njp, int Nkp, double *xyz);

/* A grid file holding X, Y, Z: what readH5Grid ingests. */
int gam_io_write_grid(const char *path, int nG, int Nip, int Njp, int Nkp, const double *xyz);

/* writeH5Res (gioH5.F90:650-704): the attributes of c (dt0 written as 0 when
 * below 10*TINY), the corners, State (Gas, magFlux, Bxyz) and oState (oGas,
 * omagFlux, oBxyz; t and ot from c).  doMHD = 0 writes no magFlux/Bxyz. */
int gam_io_write_restart(const char *path, int nG, int Nip, int Njp, int Nkp, int doMHD,
                         const gam_io_clock *c, const double *xyz,
                         const double *Gas, const double *magFlux, const double *Bxyz,
                         const double *oGas, const double *omagFlux, const double *oBxyz);

/* readH5Restart (gioH5.F90:707-872): State and oState Gas/magFlux, and the
 * clock as the reference sets it: nRes one past the file's, dt0 = 0 when the
 * file has none or it is below 10*TINY.  Refuses a file with no oGas ("too
 * old"), with species (Gas's last Fortran dimension above 1), or of another
 * block size; a Gas without ghosts (Ni-2nG wide, older files) fills the
 * active cells only and leaves the rest of the arrays as passed.  Bxyz is
 * not read (the caller derives it from magFlux).  doMHD = 0 reads no
 * magFlux. */
int gam_io_read_restart(const char *path, int nG, int Nip, int Njp, int Nkp, int doMHD, gam_io_clock *c,
                        double *Gas, double *magFlux, double *oGas, double *omagFlux);

/* bcwind: the solar-wind time series readWind ingests (reference
 * src/gamera/apps/wind.F90:617-760), in the file's units (t [s], D [#/cc],
 * V [m/s], Temp [K], B [nT]): datasets T, D, Vx, Vy, Vz, Temp, Bx, By, Bz of
 * nt samples each; the front coefficients ByC, BzC, Bx0 (Bx0 in nT), each a
 * root attribute or a dataset, absent ones 0; the marker xBS, yBS, zBS
 * (datasets, in planet radii), absent: the origin.  V, B and xyzBS are (nt, 3)
 * Fortran-ordered (each component's series contiguous).  A file without Temp
 * is refused, as readWind refuses it (:641-646). */
typedef struct {
    double ByC, BzC, Bx0;   /* Bx0 in nT, as the file holds it */
    int hasCoef[3];         /* ByC, BzC, Bx0 present */
    int hasBS;              /* xBS present */
} gam_io_wind_info;

/* Samples in the file (the length of T). */
int gam_io_wind_len(const char *path, int *nt);
/* The file's series; nt must be gam_io_wind_len's. */
int gam_io_read_wind(const char *path, int nt, double *T, double *D, double *V, double *Temp, double *B,
                     double *xyzBS, gam_io_wind_info *info);
/* A bcwind file of those series; coef (ByC, BzC, Bx0) and xyzBS may be NULL
 * (not written).  The coefficients become root attributes. */
int gam_io_write_wind(const char *path, int nt, const double *T, const double *D, const double *V,
                      const double *Temp, const double *B, const double *xyzBS, const double *coef);

const char *gam_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
