/*
 * gamera_oracle.h -- CPU restatement of the GAMERA finite-volume MHD step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP step in
 * kaiju_amd/; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  The product path never calls into it.
 *
 * Arrays use the reference layout: Fortran column-major, i fastest, one
 * contiguous slab per component, ghost-inclusive index ranges
 *   cell arrays  (isg:ieg, jsg:jeg, ksg:keg, ncomp)        "G"
 *   face arrays  (isg:ieg+1, jsg:jeg+1, ksg:keg+1, ncomp)   "F"
 * with isg = 1-nG, ieg = Nip+nG (reference base/types/gamtypes.F90:163-245).
 *
 * Every part has an independent check in tests/ (listed in the docstring of
 * tests/test_oracle.py).
 */
#ifndef GAMERA_ORACLE_H
#define GAMERA_ORACLE_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_8CENT = 0, ORC_7UP = 1 };

/* Model_T subset (reference gamtypes.F90:71-146) */
typedef struct {
    double gamma, CFL, pdmb, Vd0, cHogH, cHogM, Ca, dFloor, pFloor;
    double t, tFin, dt, dt0, limDT0, limCPR;
    int ts, rmeth, doLFMLim, doBoris, doMHD, do25D, doCPR, fixedTimestep;
    int status;   /* CalcDT outcome: 0 ok, 1 bad (dt drop), 2 disaster */
    int doBackground;   /* background field B0 split (gamtypes.F90:88) */
    int bc[6];          /* boundary type per face INI, OUTI, INJ, OUTJ, INK, OUTK: ORC_BC_* */
    int ncpr;           /* CalcDT CPR repairs done (step.F90:140-142) */
    /* solar wind of the WindBC (apps/wind.F90), constant in time: density,
     * pressure, velocity, field (code units) */
    double wind_rho, wind_p, wind_v[3], wind_b[3];
    /* or a time series (readWind, apps/wind.F90:617-760, after its unit
     * conversion) when wind_nt >= 2: wind_nt rows of ORC_WIND_COLS doubles
     * [t, rho, p, vx, vy, vz, Bx, By, Bz, xBS, yBS, zBS]; tMin, tMax, dtW
     * and the front coefficients ByC, BzC, Bx0 */
    int wind_nt;
    const double *wind_ser;
    double wind_tMin, wind_tMax, wind_dtW, wind_ByC, wind_BzC, wind_Bx0;
} orc_model;
#define ORC_WIND_COLS 12

/* boundary conditions (bcs.F90): periodic, zero gradient, zero extension,
 * Cartesian reflecting, spherical radial (SphIn on INI, SphOut on OUTI),
 * through-the-axis LFM pole (lfmIn on INJ, lfmOut on OUTJ), solar wind (WindBC on
 * OUTI, apps/wind.F90:206-374, with a wind constant in time) */
enum { ORC_BC_PERIODIC = 0, ORC_BC_ZEROGRAD = 1, ORC_BC_ZEROEXTEN = 2, ORC_BC_REFLECT = 3, ORC_BC_SPHERE = 4,
       ORC_BC_LFM = 5, ORC_BC_WIND = 6, ORC_BC_CYLPOLE = 7, ORC_BC_LAZY = 8, ORC_BC_ION = 9 };
#define ORC_MAX_RINGS 32

/* Grid_T subset (reference gamtypes.F90:163-245) */
typedef struct {
    int Nip, Njp, Nkp, nG;
    /* derived by orc_grid_derive */
    int Ni, Nj, Nk;
    int is, ie, js, je, ks, ke;
    int isg, ieg, jsg, jeg, ksg, keg;
    int isDT, ieDT, jsDT, jeDT, ksDT, keDT;
    size_t nc, nf;
    /* caller-owned arrays */
    double *xyz;     /* F x 3   corners */
    double *xyzcc;   /* G x 3   cell barycentres */
    double *xfc;     /* F x 3 x 3 face centres (xyz, face dir) */
    double *face;    /* F x 3   face areas */
    double *Tf;      /* F x 9 x 3 face triads */
    double *edge;    /* F x 3   edge lengths */
    double *Te;      /* F x 9 x 3 edge triads */
    double *Teb;     /* F x 4 x 3 edge B transforms */
    double *volume, *di, *dj, *dk;  /* G */
    /* background field (gamtypes.F90:231-238), G-sized, filled by orc_add_b0 */
    double *fcB0;    /* G x 3 x 3 face-averaged B0 (xyz, face dir) */
    double *edgB0;   /* G x 2 x 3 edge-averaged B0 on the edge's (T1,T2) (slot, edge dir) */
    double *B0;      /* G x 3 cell-averaged B0 */
    double *bFlux0;  /* G x 3 B0 flux through each face */
    double *dpB0;    /* G x 3 -div(B0 Maxwell stress)/volume (active cells) */
    /* ring average about the LFM-grid axis (Model%doRing, gid "lfm"; set before
     * orc_corners2grid): the j = js and j = je+1 faces lie on the axis, k goes
     * once around it (Nkp = Np cells), ring r (1-based) of each pole is split
     * into nch[r-1] chunks */
    int ring, nring;
    int nch[ORC_MAX_RINGS];
} orc_grid;

/* State_T subset (reference gamtypes.F90:248-264) */
typedef struct {
    double time;
    double *Gas;      /* G x 5 */
    double *magFlux;  /* F x 3 */
    double *Bxyz;     /* G x 3 */
    double *Efld;     /* G x 3 */
} orc_state;

typedef struct orc_solver orc_solver;

void   orc_grid_derive(orc_grid *g);

/* setup producers */
void   orc_gen_corners_cart(const orc_grid *g, const double bds[6], int do25D);
void   orc_gen_corners_cyl(const orc_grid *g, const double bds[6]);
void   orc_gen_corners_warp3(const orc_grid *g, const double bds[6], double amp);
void   orc_gen_corners_refwarp(const orc_grid *g, const double bds[6], double w0, double R, int Nw, int do25D);
void   orc_corners2grid(const orc_grid *g);
void   orc_ic_ot(const orc_model *m, const orc_grid *g, orc_state *s,
                 double rho0, double p0, double b0, double kscale);
void   orc_ic_bw(const orc_model *m, const orc_grid *g, orc_state *s,
                 double x0, double y0, double z0, double D0, double P0,
                 double pRat, double dRat, double rC, double B0);
/* background.F90:20-167 AddB0 for B0(x) = bu + dipole(x - xc) with moment
 * (0,0,M0) (dipolehelper.F90:19-30, with r^3 and r^5 as products) */
void   orc_add_b0(const orc_model *m, const orc_grid *g, double M0, const double xc[3], const double bu[3]);

/* solver scratch */
orc_solver *orc_solver_new(const orc_grid *g);
void        orc_solver_free(orc_solver *sv);
double     *orc_solver_array(orc_solver *sv, int which);

/* hot path (reference mhdgroup.F90, stress.F90, fields.F90, step.F90, gamapp.F90) */
void   orc_hatch_finish(orc_model *m, const orc_grid *g, orc_state *s, orc_state *o);
void   orc_advance_mhd(orc_model *m, const orc_grid *g, orc_state *s, orc_state *o,
                       orc_solver *sv, double dt);
double orc_calc_dt(orc_model *m, const orc_grid *g, orc_state *s);
void   orc_enforce_bcs(const orc_model *m, const orc_grid *g, orc_state *s);
void   orc_bflux2fld(const orc_grid *g, const double *magFlux, double *Bxyz);
void   orc_step(orc_model *m, const orc_grid *g, orc_state *s, orc_state *o, orc_solver *sv);
int    orc_run_steps(orc_model *m, const orc_grid *g, orc_state *s, orc_state *o,
                     orc_solver *sv, int nsteps, int dt0_after);
double orc_divb(const orc_grid *g, const orc_state *s);
int orc_omp_threads(void);
long orc_nuke_count(int reset);

/* stage entry points used by stage-level parity tests */
void   orc_predictor(const orc_model *m, const orc_grid *g, const orc_state *s,
                     const orc_state *o, orc_state *p, double pdt);
void   orc_calc_efield(const orc_model *m, const orc_grid *g, const orc_state *n,
                       const orc_state *hf, double *Vf, double *E);
void   orc_calc_stress(const orc_model *m, const orc_grid *g, const orc_state *n,
                       const orc_state *hf, double *gFlx, double *mFlx,
                       double *dGasH, double *dGasM);

/* ring-average pieces exposed for known-answer tests (ringav.F90, ringrecon.F90) */
void   orc_ring_average(const orc_model *m, const orc_grid *g, orc_state *s, double *SmoothE);
void   orc_reconstruct_ring(double *rW, int Nc, int Np);
void   orc_ppm_intlr(double fm2, double fm1, double f, double fp1, double fp2, double *vm, double *vp);
void   orc_lfm_map(const orc_grid *g, int what, int i, int j, int k, int *ip, int *jp, int *kp);
double orc_wgt_wind(const orc_model *m, const double *xyz);
void   orc_wind_at(const orc_model *m, const double *xyz, double t, double *q);

/* cell-level helpers exposed for known-answer tests */
void   orc_cell_c2p(const orc_model *m, const double *con, double *prim);
void   orc_cell_p2c(const orc_model *m, const double *prim, double *con);
double orc_pdm(double pdmb, double q0, double q1, double q2, double qI);
void   orc_pdmlr(double pdmb, const double *nQ, double Qi, double *Ql, double *Qr);
double orc_up7(const double *q);
double orc_central8(const double *q);
void   orc_gaskinflux(const orc_model *m, int q, const double *prim, const double *Tf9,
                      double *gFlx5, double *dW5);
void   orc_magkinflux(const orc_model *m, int q, const double *prim, const double *mag,
                      double Bn, const double *Tf9, double *mFlx3, double *bbD, double *VaD);
/* MagKinFlux with the background-field terms (B0 at the face, B0n its normal) */
void   orc_magkinflux_b0(const orc_model *m, int q, const double *prim, const double *mag,
                         double Bn, const double *Tf9, const double *B0, double B0n,
                         double *mFlx3, double *bbD, double *VaD);
void   orc_cell_boris(const orc_model *m, double *U2, const double *U0, const double *B2,
                      const double *B0, const double *dGasH, const double *dGasM,
                      double Ca, double dt);
void   orc_cell_maxwell(const orc_model *m, double *U, const double *dGasM, double dt);
double orc_cell_dt(const orc_model *m, const orc_grid *g, const orc_state *s, int i, int j, int k);

#ifdef __cplusplus
}
#endif
#endif
