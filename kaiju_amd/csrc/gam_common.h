// gam_common.h -- shared device definitions for the MI355X GAMERA step.
//
// Layout on the device is the reference's own (Fortran column-major, i fastest,
// one slab per component, ghost-inclusive), so uploads/downloads are straight
// copies.  Kernels use 0-based local indices (i0 = i - isg), the active block is
// [nG, nG+Nip) in each axis.
#pragma once
#ifdef GAM_EMU
// CPU emulation build used by the tests to check indexing and logic
#include "gam_emu.h"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "gam_erfc.h"

#define GAM_TINY 1.0e-12    // base/kdefs.F90:20 TINY
#define GAM_HUGE 1.0e+12    // base/kdefs.F90:20 HUGE
#define GAM_PI 3.14159265358979323846
#define GAM_SQRTPI 1.77245385090551602730   // sqrt(pi)

enum { DEN = 0, MOMX, MOMY, MOMZ, ENERGY };                                      // kdefs.F90:98-112
enum { NORMX = 0, NORMY, NORMZ, TAN1X, TAN1Y, TAN1Z, TAN2X, TAN2Y, TAN2Z };      // gdefs.F90:27-32
enum { XNQI = 0, YNQI, XNQJ, YNQJ };                                             // gdefs.F90:35-37

// Model_T knobs used inside kernels (passed by value).
struct MP {
    double gamma, CFL, pdmb, Vd0, cHogH, cHogM, Ca, dFloor, pFloor;
    int rmeth, doLFMLim, doBoris, doMHD, do25D;
};

// Device clock: Model%t, Model%dt, dt0, ts and the two State times.  Kernels read
// dt from here so a whole step can run without the host.
struct DClock {
    double t, dt, dt0, tFin, stime, otime, pdt, odt;
    double limDT0, limCPR;
    int ts, status, dt0_after, fixedTimestep, doCPR, halt;
    int cpr_pending, ncpr;   // CalcDT asked for a CPR repair this step; repairs so far
};

// Per-cell / per-face quantities the stencil kernels read, computed once per
// step in the predictor pass.  Each is exactly the expression the reference
// evaluates per stencil point, so moving it changes no bits:
//   Wh, Wn  BlockCon2Prim of the predictor / limiter state  (recon.F90:66-92)
//   VdV     vol*M/max(D,dFloor) of the predictor state      (fields.F90:265-271)
//   bh, bn  magFlux/max(face,TINY) of predictor / limiter   (fields.F90:331-337)
// R is the bulk real type: double for the FP64 step, float for the FP32-bulk
// step (same kernels; every literal goes through R(), so the double
// instantiation is the FP64 arithmetic unchanged).
template <typename R> struct WKT {
    R *Wh, *Wn, *VdV;   // 5G, 5G, 3G
    R *bh, *bn;         // 3F, 3F
};
typedef WKT<double> WK;

// Grid view: sizes + geometry pointers (device).  Positions (xyz, xyzcc, xfc)
// and the FP64 volume dvol stay double in both precisions: bFlux2Fld's lever
// arms are differences of nearby positions (FP32 bulk reads them from b2fc,
// formed in FP64 and rounded once) and Flux2Deltas divides by dvol.
template <typename R> struct GVT {
    int Ni, Nj, Nk, nG, Nip, Njp, Nkp;
    int nc, nf;           // number of cells (G) and face points (F)
    const double *xyz, *xyzcc, *xfc, *dvol;
    const R *face, *Tf, *edge, *Te, *Teb, *vol, *di, *dj, *dk;
    // FP32 bulk: bFlux2Fld's lever arms per cell, 18 slabs [(3*face + x)*nc + c]
    // of +-(face centre - cell centre)/volume for its lower/upper I, J, K faces
    // (differences formed in FP64, then rounded); null in FP64
    const float *b2fc;
};
typedef GVT<double> GV;

// Background field B0 (reference Grid_T fcB0/edgB0/B0/bFlux0/dpB0,
// gamtypes.F90:231-238), G-sized FP64 arrays; on == 0: no background.
//   fcB0  [nc*(xyz + 3*faceDir)]  face-averaged B0 of the face at the lower corner
//   edgB0 [nc*(slot + 2*edgeDir)] edge-averaged B0 on the edge's (T1, T2)
//   B0    [nc*xyz]                cell-averaged B0
//   bFlux0[nc*faceDir]            B0 flux through the face
//   dpB0  [nc*xyz]                -div(B0 Maxwell stress)/volume
struct BGV {
    const double *fcB0, *edgB0, *B0, *bFlux0, *dpB0;
    int on;
};

template <typename R> __host__ __device__ inline int ci(const GVT<R> &g, int i, int j, int k) { return i + g.Ni * (j + g.Nj * k); }
template <typename R> __host__ __device__ inline int fi(const GVT<R> &g, int i, int j, int k) { return i + (g.Ni + 1) * (j + (g.Nj + 1) * k); }
template <typename R> __host__ __device__ inline int cstride(const GVT<R> &g, int d) { return d == 0 ? 1 : (d == 1 ? g.Ni : g.Ni * g.Nj); }
template <typename R> __host__ __device__ inline int fstride(const GVT<R> &g, int d) { return d == 0 ? 1 : (d == 1 ? g.Ni + 1 : (g.Ni + 1) * (g.Nj + 1)); }

template <typename T> __device__ __forceinline__ T sq(T x) { return x * x; }   // Fortran x**2.0

// math in the bulk precision
__device__ __forceinline__ double rmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float rmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double rmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float rmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double rabs(double a) { return fabs(a); }
__device__ __forceinline__ float rabs(float a) { return fabsf(a); }
__device__ __forceinline__ double rsgn(double a) { return copysign(1.0, a); }     // SIGN(1.0, a)
__device__ __forceinline__ float rsgn(float a) { return copysignf(1.0f, a); }
__device__ __forceinline__ double rroot(double a) { return sqrt(a); }
__device__ __forceinline__ float rroot(float a) { return sqrtf(a); }
__device__ __forceinline__ double rexp(double a) { return exp(a); }
__device__ __forceinline__ float rexp(float a) { return expf(a); }

// erfc(x) from e = exp(-x*x), which the flux callers compute anyway:
// erfc(u) = e * P((u - 2)/(u + 2)) for u = |x| (u clamped to 6 inside P;
// erfc(6) < 2.2e-17), erfc(-u) = 2 - erfc(u).  P is this project's own
// Chebyshev fit of erfc(u)exp(u^2) on [0, 6] (gam_erfc.h, made by
// erfc_fit.py): |error| <= 3.4e-16 in FP64 and 2.3e-7 in FP32, absolute,
// which is what the fluxes need (erfc weights O(1) flux terms).
__device__ __forceinline__ double rerfc_e(double x, double e) {
    const double u = fmin(fabs(x), 6.0);
    const double t = (u - 2.0) / (u + 2.0);
    double p = kErfcD[0];
#pragma unroll
    for (int q = 1; q < 22; q++) p = fma(p, t, kErfcD[q]);
    const double r = e * p;
    return x < 0.0 ? 2.0 - r : r;
}
__device__ __forceinline__ float rerfc_e(float x, float e) {
    const float u = fminf(fabsf(x), 6.0f);
    const float t = (u - 2.0f) / (u + 2.0f);
    float p = kErfcF[0];
#pragma unroll
    for (int q = 1; q < 11; q++) p = fmaf(p, t, kErfcF[q]);
    const float r = e * p;
    return x < 0.0f ? 2.0f - r : r;
}

// recon.F90:22-30 weights (integer numerators / denominator in double; the
// FP32 tables are those doubles rounded to float)
__device__ __constant__ static const double cW8[8] = {
    -3.0 / 840.0, 29.0 / 840.0, -139.0 / 840.0, 533.0 / 840.0, 533.0 / 840.0, -139.0 / 840.0, 29.0 / 840.0, -3.0 / 840.0};
__device__ __constant__ static const double cW6[8] = {0.0 / 60.0, 1.0 / 60.0, -8.0 / 60.0, 37.0 / 60.0, 37.0 / 60.0, -8.0 / 60.0, 1.0 / 60.0, 0.0 / 60.0};
__device__ __constant__ static const double cU7[7] = {
    -3.0 / 420.0, 25.0 / 420.0, -101.0 / 420.0, 319.0 / 420.0, 214.0 / 420.0, -38.0 / 420.0, 4.0 / 420.0};
__device__ __constant__ static const float fW8[8] = {
    (float)(-3.0 / 840.0), (float)(29.0 / 840.0), (float)(-139.0 / 840.0), (float)(533.0 / 840.0),
    (float)(533.0 / 840.0), (float)(-139.0 / 840.0), (float)(29.0 / 840.0), (float)(-3.0 / 840.0)};
__device__ __constant__ static const float fU7[7] = {
    (float)(-3.0 / 420.0), (float)(25.0 / 420.0), (float)(-101.0 / 420.0), (float)(319.0 / 420.0),
    (float)(214.0 / 420.0), (float)(-38.0 / 420.0), (float)(4.0 / 420.0)};
template <typename R> __device__ __forceinline__ R w8(int n);
template <> __device__ __forceinline__ double w8<double>(int n) { return cW8[n]; }
template <> __device__ __forceinline__ float w8<float>(int n) { return fW8[n]; }
template <typename R> __device__ __forceinline__ R u7(int n);
template <> __device__ __forceinline__ double u7<double>(int n) { return cU7[n]; }
template <> __device__ __forceinline__ float u7<float>(int n) { return fU7[n]; }

// recon.F90:202-214 Central8 (sequential dot product)
template <typename R> __device__ __forceinline__ R c8(const R *q) {
    R s = R(0.0);
#pragma unroll
    for (int n = 0; n < 8; n++) s += w8<R>(n) * q[n];
    return s;
}

// recon.F90:303-310 Up7
template <typename R>
__device__ __forceinline__ R up7(R q0, R q1, R q2, R q3, R q4, R q5, R q6) {
    return q0 * u7<R>(0) + q1 * u7<R>(1) + q2 * u7<R>(2) + q3 * u7<R>(3) + q4 * u7<R>(4) + q5 * u7<R>(5) + q6 * u7<R>(6);
}

// recon.F90:275-301 PDM (one-sided)
template <typename R> __device__ __forceinline__ R pdm1(R pdmb, R q0, R q1, R q2, R qI) {
    R maxQ = rmax(q1, q2), minQ = rmin(q1, q2);
    R qN = rmax(minQ, rmin(qI, maxQ));
    R dq0 = pdmb * (q1 - q0);
    R dq1 = pdmb * (q2 - q1);
    R s0 = rsgn(dq0), s1 = rsgn(dq1);
    R s01 = rabs(s0 + s1);
    R dqL = qN - q1;
    return qN - s1 * rmax(R(0.0), rabs(dqL) - s01 * rabs(dq0));
}

// recon.F90:231-272 pdmLR
template <typename R>
__device__ __forceinline__ void pdmlr(R pdmb, R v0, R v1, R v2, R v3, R Qi, R &Ql, R &Qr) {
    R maxV = rmax(v1, v2), minV = rmin(v1, v2);
    R vN = rmax(minV, rmin(Qi, maxV));
    R dv0 = pdmb * (v1 - v0), dv1 = pdmb * (v2 - v1), dv2 = pdmb * (v3 - v2);
    R s0 = rsgn(dv0), s1 = rsgn(dv1), s2 = rsgn(dv2);
    R q0 = rabs(s0 + s1), q1 = rabs(s1 + s2);
    R dvL = vN - v1, dvR = v2 - vN;
    Ql = vN - s1 * rmax(R(0.0), rabs(dvL) - q0 * rabs(dv0));
    Qr = vN + s1 * rmax(R(0.0), rabs(dvR) - q1 * rabs(dv2));
}

// GetLRs for one variable: recon.F90:125-158 (Cen8LRs) / 162-198 (Up7LRs).
// rVi = 1/Vi, formed once per interface by the caller: the reference divides
// each interpolant by Vi; multiplying by the reciprocal differs by at most
// an ulp and saves two divisions per variable.
template <int RM, typename R>
__device__ __forceinline__ void get_lrs(R pdmb, const R *dV, const R *nQ, const R *Q, R rVi, R &Ql, R &Qr) {
    R QdV[8];
#pragma unroll
    for (int n = 0; n < 8; n++) QdV[n] = dV[n] * Q[n];
    if (RM == 1) {
        R l = up7(QdV[0], QdV[1], QdV[2], QdV[3], QdV[4], QdV[5], QdV[6]) * rVi;
        R r = up7(QdV[7], QdV[6], QdV[5], QdV[4], QdV[3], QdV[2], QdV[1]) * rVi;
        Ql = pdm1(pdmb, nQ[0], nQ[1], nQ[2], l);
        Qr = pdm1(pdmb, nQ[3], nQ[2], nQ[1], r);
    } else {
        R Qi = c8(QdV) * rVi;
        pdmlr(pdmb, nQ[0], nQ[1], nQ[2], nQ[3], Qi, Ql, Qr);
    }
}

// gamutils.F90:301-321 CellC2P
template <typename R> __device__ __forceinline__ void cell_c2p(const MP &m, const R *c, R *w) {
    R rho = c[DEN];
    R Vx = c[MOMX] / rho, Vy = c[MOMY] / rho, Vz = c[MOMZ] / rho;
    R KinE = R(0.5) * rho * (sq(Vx) + sq(Vy) + sq(Vz));
    w[0] = rmax(rho, R(m.dFloor));
    w[1] = Vx; w[2] = Vy; w[3] = Vz;
    w[4] = rmax(R(m.gamma - 1) * (c[ENERGY] - KinE), R(m.pFloor));
}

// gamutils.F90:362-383 CellP2C
template <typename R> __device__ __forceinline__ void cell_p2c(const MP &m, const R *w, R *c) {
    R rho = rmax(w[0], R(m.dFloor));
    R Vx = w[1], Vy = w[2], Vz = w[3];
    R IntE = rmax(w[4], R(m.pFloor)) / R(m.gamma - 1);
    R KinE = R(0.5) * rho * (sq(Vx) + sq(Vy) + sq(Vz));
    c[0] = rho; c[1] = rho * Vx; c[2] = rho * Vy; c[3] = rho * Vz; c[4] = IntE + KinE;
}

// recon.F90:66-92 BlockCon2Prim
template <typename R> __device__ __forceinline__ void block_c2p(const MP &m, const R *c, R *w) {
    R D = rmax(c[DEN], R(m.dFloor));
    R KinE = R(0.5) * (sq(c[MOMX]) + sq(c[MOMY]) + sq(c[MOMZ])) / D;
    R P = rmax(R(m.gamma - 1) * (c[ENERGY] - KinE), R(m.pFloor));
    w[0] = D; w[1] = c[MOMX] / D; w[2] = c[MOMY] / D; w[3] = c[MOMZ] / D; w[4] = P;
}

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// Launch tile: a block of ti*tj*tk threads (== blockDim.x) covers a brick of
// an nx*ny*nz index box, so a stencil along j or k is served from the bytes the
// block's own waves fetch instead of from 8 separate passes over HBM.  Tiles are
// numbered with axis (fast & 3) varying fastest; (fast & 3) == TILE_LINEAR
// instead gives each block 256 consecutive points of the box in i-fastest order
// (no ragged bricks when nx is not a multiple of ti).  The hardware dispatches
// block b to XCD b % 8; unless fast & TILE_NOXCD, block b takes tile
// (b % 8) * per + b / 8, so each XCD walks one contiguous run of tiles and
// neighbouring tiles meet in the same L2; with TILE_NOXCD block b takes tile b.
struct Tile { int ti, tj, tk, fast; };
#define TILE_LINEAR 3
#define TILE_NOXCD 4

__device__ __forceinline__ int tile_of_block(const Tile &t, int nt, const unsigned vb) {
    if (t.fast & TILE_NOXCD) return (int)vb;
    const int per = cdiv(nt, 8);
    return (int)(vb & 7) * per + (int)(vb >> 3);
}

__host__ __device__ inline int tile_count(const Tile &t, int nx, int ny, int nz) {
    if ((t.fast & 3) == TILE_LINEAR) return cdiv(nx * ny * nz, t.ti * t.tj * t.tk);
    return cdiv(nx, t.ti) * cdiv(ny, t.tj) * cdiv(nz, t.tk);
}
__host__ __device__ inline int tile_blocks(const Tile &t, int nx, int ny, int nz) {
    const int nt = tile_count(t, nx, ny, nz);
    return (t.fast & TILE_NOXCD) ? nt : 8 * cdiv(nt, 8);
}

// (a, b, c) of this thread of block vb in the box; false if outside
__device__ __forceinline__ bool tile_pos(const unsigned vb, const Tile &t, int nx, int ny, int nz, int &a, int &b, int &c) {
    if ((t.fast & 3) == TILE_LINEAR) {
        const int n = nx * ny * nz, bs = t.ti * t.tj * t.tk, nt = cdiv(n, bs);
        const int q = tile_of_block(t, nt, vb);
        const int idx = q * bs + (int)threadIdx.x;
        if (q >= nt || idx >= n) return false;
        a = idx % nx;
        const int r = idx / nx;
        b = r % ny;
        c = r / ny;
        return true;
    }
    const int ntx = cdiv(nx, t.ti), nty = cdiv(ny, t.tj), ntz = cdiv(nz, t.tk);
    const int q = tile_of_block(t, ntx * nty * ntz, vb);
    if (q >= ntx * nty * ntz) return false;
    int tx, ty, tz;
    const int f = t.fast & 3;
    if (f == 0) { tx = q % ntx; int r = q / ntx; ty = r % nty; tz = r / nty; }
    else if (f == 1) { ty = q % nty; int r = q / nty; tx = r % ntx; tz = r / ntx; }
    else { tz = q % ntz; int r = q / ntz; tx = r % ntx; ty = r / ntx; }
    const int l = threadIdx.x;
    a = tx * t.ti + l % t.ti;
    b = ty * t.tj + (l / t.ti) % t.tj;
    c = tz * t.tk + l / (t.ti * t.tj);
    return a < nx && b < ny && c < nz;
}
