/*
 * gamera_hip.h -- C-ABI of the MI355X GAMERA MHD step (libgamera_hip.so).
 *
 * Plain C types only: dimensions, double pointers, sizes.  Every host array is
 * a contiguous column-major (Fortran) array with the reference's index bounds:
 *   cell arrays  (isg:ieg,   jsg:jeg,   ksg:keg,   ncomp)   "G"
 *   face arrays  (isg:ieg+1, jsg:jeg+1, ksg:keg+1, ncomp)   "F"
 * i.e. exactly what c_loc() of the reference's allocatables gives
 * (reference src/base/types/gamtypes.F90:163-264).
 *
 * Threading: one host thread per handle; all calls on a handle are ordered on
 * the handle's HIP stream.  Calls that return host data synchronise.
 * Errors: every entry point returns 0 on success, > 0 for a recoverable
 * numerical condition (the reference's "bad timestep" path), < 0 for a fatal
 * error (bad argument, HIP failure).  gam_last_error() has the text.  The
 * library never exits the process.
 */
#ifndef GAMERA_HIP_H
#define GAMERA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gam_handle gam_handle;

/* Grid_T sizes: active cells and ghost depth (reference gamtypes.F90:163-180,
 * init.F90:327 nG = 4). */
typedef struct {
    int32_t Nip, Njp, Nkp, nG;
} gam_dims;

/* Model_T numerical knobs (reference gamtypes.F90:71-146, defaults read in
 * init.F90:312-492 initModel). */
typedef struct {
    double gamma, CFL, pdmb, Vd0, cHogH, cHogM, Ca, dFloor, pFloor, tFin, limDT0, limCPR;
    int32_t rmeth;          /* 0 = 8CENT (Cen8LRs), 1 = 7UP (Up7LRs), init.F90:450-462 */
    int32_t doLFMLim;       /* limit on state n (gamtypes.F90:123)                    */
    int32_t doBoris;        /* Boris correction with speed Ca                         */
    int32_t doMHD;
    int32_t do25D;          /* Nkp == 1 (init.F90:612-616)                             */
    int32_t doCPR;
    int32_t fixedTimestep;
    int32_t precision;      /* 0 = FP64 everywhere; 1 = FP32 bulk, FP64 CT/energy/dt   */
} gam_params;
/* precision = 1 (FP32 bulk): Gas, Bxyz, the predictor state, reconstruction,
 * fluxes, EMFs and the per-face/per-cell geometry they read are FP32.  FP64 is
 * kept for magFlux and the CT face update (E2Flux), for Flux2Deltas with the
 * Reynolds/Maxwell/Boris cell update and its energy-to-pressure recovery with
 * floors, for bFlux2Fld (positions stay FP64) and for the dt reduction.  The
 * C-ABI still takes and returns FP64 host arrays: gam_put_state rounds Gas and
 * Bxyz to FP32, gam_get_state widens them. */

/* Model_T / State_T time bookkeeping that the step reads and writes. */
typedef struct {
    double t;           /* Model%t                                   */
    double dt;          /* Model%dt                                  */
    double dt0;         /* Model%dt0 (0 until captured)              */
    double stateTime;   /* State%time                                */
    double oStateTime;  /* oState%time                               */
    int32_t ts;         /* Model%ts                                  */
    int32_t status;     /* last CalcDT outcome: 0 ok, 1 bad, 2 disaster */
    int32_t dt0_after;  /* capture dt0 at the first step with ts >= dt0_after (0: never) */
    int32_t ncpr;       /* CPR repairs done (params doCPR: a bad, not disastrous CalcDT
                           runs StateCPR, step.F90:140-142, and the step goes on) */
} gam_clock;

/* --- lifetime ------------------------------------------------------------ */
/* Replaces the reference's allocation in Hatch/initSolver (init.F90:38-59,
 * 494-516): device State, oState, predictor state and solver scratch. */
int  gam_create(const gam_dims *dims, const gam_params *params, gam_handle **out);
int  gam_destroy(gam_handle *h);
const char *gam_last_error(const gam_handle *h);
int  gam_sync(gam_handle *h);
void *gam_stream(gam_handle *h);         /* the handle's hipStream_t */

/* --- grid ----------------------------------------------------------------- */
/* Upload Grid_T geometry computed by the caller (the arrays of
 * gamtypes.F90:183-221).  Sizes: xyz F*3, xyzcc G*3, xfc F*9, face F*3, Tf F*27,
 * edge F*3, Te F*27, Teb F*12, volume/di/dj/dk G. */
int  gam_set_grid(gam_handle *h, const double *xyz, const double *xyzcc, const double *xfc,
                  const double *face, const double *Tf, const double *edge, const double *Te,
                  const double *Teb, const double *volume, const double *di, const double *dj,
                  const double *dk);
/* Upload the corners only and derive the rest on the GPU
 * (replaces Corners2Grid + ebGeom, reference init.F90:801-1145). */
int  gam_set_corners(gam_handle *h, const double *xyz);
/* Download the geometry (same sizes as gam_set_grid; NULL skips an array). */
int  gam_get_grid(gam_handle *h, double *xyz, double *xyzcc, double *xfc, double *face,
                  double *Tf, double *edge, double *Te, double *Teb, double *volume,
                  double *di, double *dj, double *dk);

/* --- background field (Model%doBackground) ------------------------------- */
/* The split B = B0 + dB of the reference (background.F90:20-167 AddB0 and its
 * uses in stress.F90:185-188,302-305,550-596, fields.F90:417-420,
 * gridutils.F90:82-86, mhdgroup.F90:180-183).  Arrays are Grid_T's
 * (gamtypes.F90:231-238), G-sized: fcB0 G*3*3 (xyz, face dir), edgB0 G*2*3
 * (T1/T2 slot, edge dir), B0 G*3, bFlux0 G*3, dpB0 G*3.  Setting a background
 * turns the split on for every later step; State's magFlux and Bxyz then hold
 * dB only.  gam_set_background uploads arrays the caller computed (AddB0 for
 * any Model%B0); gam_background_dipole runs AddB0 on the GPU, on the current
 * geometry, for B0(x) = bu + dipole(x - xc) with moment (0,0,M0)
 * (dipolehelper.F90:19-30 MagsphereDipole, its rad**3/rad**5 as products). */
int  gam_set_background(gam_handle *h, const double *fcB0, const double *edgB0, const double *B0,
                        const double *bFlux0, const double *dpB0);
int  gam_get_background(gam_handle *h, double *fcB0, double *edgB0, double *B0, double *bFlux0,
                        double *dpB0);
int  gam_background_dipole(gam_handle *h, double M0, const double *xc, const double *bu);

/* --- device buffers (no counterpart in the reference) -------------------- */
/* Plain device memory for callers without a GPU allocator of their own (a
 * correction field, a gradient buffer): gam_memcpy copies in any direction
 * and returns when the copy is done. */
int  gam_device_alloc(size_t bytes, void **p);
int  gam_device_free(void *p);
int  gam_memcpy(void *dst, const void *src, size_t bytes);

/* --- learned coarse-grid correction ------------------------------------- */
/* No counterpart in the reference.  A correction field o (device FP32, [G][8]:
 * 5 gas-flux and 3 edge-EMF channels per cell, e.g. the output of
 * libgamera_net.so) set with gam_set_correction is applied by every following
 * AdvanceMHD: scale*face*o[v] is added to gas flux v of each active interface
 * at the cell's lower faces and scale*edge*o[5+d] to each active edge EMF d,
 * before E2Flux and the update, so the update stays conservative and div.B
 * exact.  o is read at launch time: refresh it between steps (o = NULL turns the
 * correction off).  gam_state_device returns the device arrays of State (0) or
 * oState (1), bulk-typed by the precision.  Training: gam_correction_target
 * uploads a target change (Gas G*5, magFlux F*3) and gam_correction_grad
 * evaluates, for the change o makes to one step of length dt linearised
 * (Gas += dt*divergence, magFlux -= dt*curl), the loss
 *   sum_active_cells sum_v w_v (dGas_v - target_v)^2 + w3 sum_CT_faces (dPhi - target)^2
 * with w4 = {w_rho, w_momentum, w_energy, w_magFlux}, and its gradient dL/do
 * into dout (device FP32 [G][8]). */
int  gam_state_device(gam_handle *h, int which, void **Gas, double **magFlux, void **Bxyz);
int  gam_set_correction(gam_handle *h, const float *o, double scale);
/* A hipEvent_t the step waits for (hipStreamWaitEvent on its stream) right
 * before its first kernel that reads o, e.g. the event gam_net_infer records
 * on the net's own stream: the network's inference then overlaps the step's
 * predictor, bFlux2Fld and Mom2Face.  NULL: no wait.  Graphs refuse handles
 * with a wait set. */
int  gam_set_correction_wait(gam_handle *h, void *event);
int  gam_correction_target(gam_handle *h, const double *dGas, const double *dMag);
int  gam_correction_target_device(gam_handle *h, const double *dGas, const double *dMag);  /* device FP64 */
int  gam_correction_grad(gam_handle *h, const float *o, double scale, double dt, const double *w4,
                         float *dout, double *loss);

/* --- boundary conditions ------------------------------------------------ */
/* Boundary type per face, in the order INI, OUTI, INJ, OUTJ, INK, OUTK (the
 * order EnforceBCs applies them, step.F90:17-55): 0 periodic (bcs.F90:279-450),
 * 1 zero gradient (zeroGrad_*, :799-930), 2 zero extension (zeroExten_*,
 * :454-603), 3 Cartesian reflecting (cartReflect_*, :607-795), 4 spherical
 * radial (SphIn on INI, SphOut on OUTI, :934-982; i faces only), 5 LFM axis
 * (lfmIn on INJ, lfmOut on OUTJ, :1020-1106; both j faces, k periodic over an
 * even Nkp >= 16: the ghosts across the axis from the cells half a turn away),
 * 6 solar wind (WindBC on OUTI, apps/wind.F90:206-374, with the wind of
 * gam_set_wind or gam_set_wind_series; k periodic; the inner i face keeps its
 * own non-periodic BC),
 * 7 cylindrical pole (cylpole on INI, bcs.F90:987-1017: j goes once around the
 * pole, periodic over an even Njp >= 8; ghosts from the cells half a turn away),
 * 8 lazy (lazyBC, bcs.F90:268-275: the face's ghosts keep what they hold, the
 * initial condition's values when nothing else writes them: a fixed boundary),
 * 9 ionospheric inner (IonInner on INI, voltron/ICs/earthcmi.F90:378-490, with
 * no ionosphere attached: the coupling E field zero, so the ghosts take the
 * density and pressure of the mirrored cell at rest and the ghost faces zero
 * flux, the field there being the background dipole; k periodic). 
 * An axis is periodic on both faces or on neither; 2.5D blocks keep k periodic.
 * Default: all periodic.  The reference's BC objects (Grid%externalBCs) map
 * onto these. */
int  gam_set_bcs(gam_handle *h, const int *type);
int  gam_get_bcs(gam_handle *h, int *type);
/* The solar wind (density, pressure, velocity, field; code units) of the
 * GAM_BC_WIND boundary, constant in time: the reference's WindBC reading a
 * constant time series (GetWindAt, apps/wind.F90:160-177).  Before gam_set_bcs. */
int  gam_set_wind(gam_handle *h, double rho, double p, const double *v, const double *b);
/* The solar wind of the GAM_BC_WIND boundary as a time series, replacing
 * readWind (apps/wind.F90:617-760) after its unit conversion: nt >= 2 samples
 * at times t (code units, increasing, uniform to 1e-6 as readWind demands) of
 * density rho, pressure p, velocity v and field b (v, b, xyzBS: (nt, 3), each
 * component's series contiguous, as readWind's Q(:,VELX:VELZ), B and xyzBS
 * hold them); b's x component is replaced by
 * Bx0 + ByC By + BzC Bz (the tilted front, wind.F90:718-721); xyzBS, the
 * series' marker position (the bow shock, xBS/yBS/zBS), may be NULL (zero).
 * The boundary then takes, at the clock's time t, each point's wind at t - DelT
 * (GetWindAt, :160-177), DelT the travel time of the front (RefreshWind,
 * :116-158) from the marker, from the sample whose interval holds that time
 * (discrete walls, InterpWgts :819-868), with a null field at or before the
 * first sample.  Refuses a wind at or above the Boris speed (doBoris), as
 * readWind does.  Before gam_set_bcs; gam_set_wind switches back to a
 * constant wind. */
int  gam_set_wind_series(gam_handle *h, int nt, const double *t, const double *rho, const double *p,
                         const double *v, const double *b, const double *xyzBS, double ByC, double BzC,
                         double Bx0);

/* --- ring average about the LFM axis (Model%doRing, ring/gid = "lfm") ----- */
/* Replaces InitRings/RingGridFix (ringutils.F90:37-390) and turns on, inside
 * every step, RingFlux (stress.F90:137-139), PoleE (mhdgroup.F90:138-142),
 * RingFlux2FieldFix in bFlux2Fld (gridutils.F90:298-300) and RingAverage after
 * the update (mhdgroup.F90:69-74, ringav.F90:25-73; PPM ring reconstruction,
 * one Fourier mode removed from the ring fluxes, fast loop cleaning, no random
 * shift).  Ring r = 1..nring of each pole (cell rows js+r-1 and je-r+1) is
 * averaged in nch[r-1] chunks along k.  Call before gam_set_corners, which then
 * applies RingGridFix to the geometry it derives (gam_set_grid takes geometry
 * as given).  Needs an even Nkp in 16..512, Njp >= 2*nring + 4, chunk counts
 * dividing Nkp and at most Nkp/2.  nring = 0 turns it off.  Allocates device
 * memory held until gam_destroy: the smoothing EMF (3 G doubles) and a ring
 * scratch of about 5 nring Nkp Nip doubles (grown, never shrunk, when called
 * again with more rings). */
int  gam_set_ring(gam_handle *h, int nring, const int *nch);
int  gam_get_ring(gam_handle *h, int *nring, int *nch);

/* --- state ---------------------------------------------------------------- */
/* which: 0 = State, 1 = oState, 2 = predictor (half-step) state (read only).
 * Gas G*5, magFlux F*3, Bxyz G*3; NULL skips an array, except that a NULL
 * Bxyz with a magFlux given is recomputed from it on the device (bFlux2Fld). */
int  gam_put_state(gam_handle *h, int which, double time, const double *Gas,
                   const double *magFlux, const double *Bxyz);
int  gam_get_state(gam_handle *h, int which, double *time, double *Gas,
                   double *magFlux, double *Bxyz);
/* State/oState from FP64 device arrays (Gas G*5, magFlux F*3); Bxyz is
 * recomputed from magFlux.  No counterpart in the reference (a host keeping
 * training samples resident on the GPU). */
int  gam_put_state_device(gam_handle *h, int which, double time, const double *Gas, const double *magFlux);
int  gam_set_clock(gam_handle *h, const gam_clock *c);
int  gam_get_clock(gam_handle *h, gam_clock *c);

/* --- initial conditions on the GPU (setup producers) ----------------------- */
/* prob.F90:74-119 initOT2D with wavenumber factor kscale. */
int  gam_ic_orszag_tang(gam_handle *h, double rho0, double p0, double b0, double kscale);
/* prob.F90:190-298 initBW via GasIC2State and VectorPot2Flux (gridutils.F90:144-237). */
int  gam_ic_blast_wave(gam_handle *h, double x0, double y0, double z0, double D0, double P0,
                       double pRat, double dRat, double rC, double B0);
/* init.F90:204-244: first Bxyz, oState = State, initial BCs, first dt. */
int  gam_hatch(gam_handle *h);

/* --- the step ------------------------------------------------------------- */
/* AdvanceMHD(Model,Grid,State,oState,ooState,Solver,Model%dt), mhdgroup.F90:16-76.
 * Leaves State (pre-BC) = new state, oState = old state. */
int  gam_advance_mhd(gam_handle *h);
/* CalcDT (step.F90:58-168) after UpdateStateData's ts += 1, t += dt
 * (gamapp.F90:112-114).  Writes Model%dt into the clock; *status as gam_clock. */
int  gam_calc_dt(gam_handle *h, double *dt, int *status);
/* EnforceBCs + bFlux2Fld (step.F90:17-55, gamapp.F90:86-90), periodic BCs. */
int  gam_enforce_bcs(gam_handle *h);
/* nsteps x stepGamera (gamapp.F90:73-94), fully on the GPU with dt kept in
 * device memory; returns the first non-zero CalcDT status. */
int  gam_step(gam_handle *h, int nsteps);
/* DivB total (gridutils.F90:717-770). */
int  gam_divb(gam_handle *h, double *tot);
/* The same total, bit for bit, from an entry point that allocates nothing: no
 * device or host allocation on any call, for callers that cannot allocate
 * (tight monitoring loops, harnesses that forbid it).  No counterpart in the
 * reference.  Its block sums live in a file-scope device array of
 * GAM_NOALLOC_BLOCKS + 1 doubles (1 MiB) that each GPU holds until the process
 * ends; blocks with more than GAM_NOALLOC_BLOCKS * 256 active cells are refused
 * (returns -1, gam_last_error says why).  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (handles on different streams,
 * host threads) race and must be serialised by the caller. */
#define GAM_NOALLOC_BLOCKS 131072
int  gam_divb_noalloc(gam_handle *h, double *tot);

/* --- diagnostics ---------------------------------------------------------- */
/* Copy a solver work array of the last step to the host, for stage-level
 * checks (the reference keeps them in gamSolver_T, gamtypes.F90:267-286):
 * which 0 = interface fluxes, G*24 (8 per direction: 5 gas then 3 magnetic,
 * already times face area); 1 = edge EMFs times edge length, G*3; 2 = face
 * velocities of Mom2Face, G*9 (3 per sweep direction I, J, K). */
int  gam_get_work(gam_handle *h, int which, double *out);

/* --- instrumentation ------------------------------------------------------ */
/* Per-stage HIP-event timers, like the reference's Tic/Toc clocks
 * (base/clocks.F90:84-182).  Stage ids are listed by gam_stage_name. */
int  gam_set_timing(gam_handle *h, int enable);
int  gam_get_timing(gam_handle *h, int stage, double *ms_total, int64_t *launches);
const char *gam_stage_name(int stage);
int  gam_num_stages(void);
/* Bytes each stage's kernels must move per launch by the algorithm (inputs
 * read once, outputs written once); used for roofline reporting. */
int  gam_stage_bytes(gam_handle *h, int stage, double *bytes);

/* --- launch tuning --------------------------------------------------------- */
/* No counterpart in the reference (its loops are OpenMP over k,j with i
 * vectorised).  The stencil stages (bflux_half, bflux_post, mom2face, emf,
 * flux_i, flux_j, flux_k, update) launch 256-thread blocks that each cover
 * a ti x tj x tk brick of their index box; 'fast' (0 i, 1 j, 2 k) is the axis
 * along which consecutive bricks of one XCD follow each other, and fast = 3
 * gives each block 256 consecutive points of the box in i-fastest order.
 * Adding 4 to fast keeps the hardware's block order (block b to XCD b % 8)
 * instead of giving each XCD one contiguous run of bricks.  Results do not
 * depend on the tile; only speed does.  gam_set_tile returns -1 for other
 * stages, ti*tj*tk != 256 or fast outside 0..7.  tile4 receives {ti, tj, tk, fast}. */
int  gam_set_tile(gam_handle *h, int stage, int ti, int tj, int tk, int fast);
int  gam_get_tile(gam_handle *h, int stage, int *tile4);

/* --- ensembles of blocks under one captured graph -------------------------- */
/* No counterpart in the reference, which steps one block per MPI rank
 * (gamera/mpi/gamapp_mpi.F90); independent blocks of different sizes are
 * separate runs there.  gam_graph_create captures nsteps (even, >= 2) fused
 * steps of each of the n handles -- the same device sequence as gam_step --
 * as independent branches of one hipGraph; gam_graph_launch replays it
 * `replays` times (nsteps * replays steps per block, each block with its own
 * dt); gam_graph_sync waits and returns the largest CalcDT status (a block
 * whose status is nonzero stops where gam_step would have stopped).  Between
 * gam_graph_launch and gam_graph_sync the handles belong to the graph: make no
 * other call on them.  Stage timers must be off while capturing.  The graph
 * holds each handle's buffers and launch parameters as they were at capture:
 * until gam_graph_destroy, gam_set_tile, gam_set_ring, gam_set_bcs,
 * gam_set_wind, gam_set_wind_series, gam_set_correction, gam_set_background,
 * gam_background_dipole and gam_destroy return -1 on a held handle, and
 * gam_graph_launch returns -1 if a handle that has not halted was stepped an
 * odd number of times outside the graph (its State and oState buffers are then
 * swapped against the capture). */
typedef struct gam_graph gam_graph;
int  gam_graph_create(gam_handle *const *hs, int n, int nsteps, gam_graph **out);
int  gam_graph_launch(gam_graph *g, int replays);
int  gam_graph_sync(gam_graph *g);
int  gam_graph_destroy(gam_graph *g);
const char *gam_graph_error(const gam_graph *g);
/* The same steps with one launch per kernel for all the handles (no
 * counterpart in the reference: its ensemble of independent runs is one
 * process per run).  Each handle's launches are recorded instead of run; the
 * launches of one kernel by several handles become one launch whose blocks are
 * split among them (each handle's blocks start at a multiple of 8, keeping
 * their XCD placement), every block running the kernel with its own handle's
 * arguments, read from a device table the graph owns.  Handles whose launch
 * sequences differ (2.5D and 3D blocks) are merged in order: a launch packs
 * every handle that runs that kernel next.  Periodic MHD handles without ring
 * average or learned correction; others are refused (-1, message on the
 * handle).  gam_graph_launch/sync/destroy as for gam_graph_create. */
int  gam_graph_create_packed(gam_handle *const *hs, int n, int nsteps, gam_graph **out);
/* Nodes (kernel launches and copies) of the captured graph. */
int  gam_graph_nodes(gam_graph *g, int *nodes);

#ifdef __cplusplus
}
#endif
#endif
