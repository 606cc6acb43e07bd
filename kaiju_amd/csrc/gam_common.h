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

#define GAM_TINY 1.0e-12    // base/kdefs.F90:20 TINY
#define GAM_HUGE 1.0e+12    // base/kdefs.F90:20 HUGE
#define GAM_PI 3.14159265358979323846

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
};

// Per-cell / per-face quantities the stencil kernels read, computed once per
// step in the predictor pass.  Each is exactly the expression the reference
// evaluates per stencil point, so moving it changes no bits:
//   Wh, Wn  BlockCon2Prim of the predictor / limiter state  (recon.F90:66-92)
//   VdV     vol*M/max(D,dFloor) of the predictor state      (fields.F90:265-271)
//   bh, bn  magFlux/max(face,TINY) of predictor / limiter   (fields.F90:331-337)
struct WK {
    double *Wh, *Wn, *VdV;   // 5G, 5G, 3G
    double *bh, *bn;         // 3F, 3F
};

// Grid view: sizes + geometry pointers (device).
struct GV {
    int Ni, Nj, Nk, nG, Nip, Njp, Nkp;
    int nc, nf;           // number of cells (G) and face points (F)
    const double *xyz, *xyzcc, *xfc, *face, *Tf, *edge, *Te, *Teb, *vol, *di, *dj, *dk;
};

__host__ __device__ inline int ci(const GV &g, int i, int j, int k) { return i + g.Ni * (j + g.Nj * k); }
__host__ __device__ inline int fi(const GV &g, int i, int j, int k) { return i + (g.Ni + 1) * (j + (g.Nj + 1) * k); }
__host__ __device__ inline int cstride(const GV &g, int d) { return d == 0 ? 1 : (d == 1 ? g.Ni : g.Ni * g.Nj); }
__host__ __device__ inline int fstride(const GV &g, int d) { return d == 0 ? 1 : (d == 1 ? g.Ni + 1 : (g.Ni + 1) * (g.Nj + 1)); }

// component access: slab c of an array with n points per slab
template <typename T> __device__ __forceinline__ T ld(const T *p, int n, int c, int idx) { return p[(size_t)n * c + idx]; }

__device__ __forceinline__ double sq(double x) { return x * x; }   // Fortran x**2.0

// recon.F90:22-30 weights (integer numerators / denominator in double)
__device__ __constant__ static const double cW8[8] = {
    -3.0 / 840.0, 29.0 / 840.0, -139.0 / 840.0, 533.0 / 840.0, 533.0 / 840.0, -139.0 / 840.0, 29.0 / 840.0, -3.0 / 840.0};
__device__ __constant__ static const double cW6[8] = {0.0 / 60.0, 1.0 / 60.0, -8.0 / 60.0, 37.0 / 60.0, 37.0 / 60.0, -8.0 / 60.0, 1.0 / 60.0, 0.0 / 60.0};
__device__ __constant__ static const double cU7[7] = {
    -3.0 / 420.0, 25.0 / 420.0, -101.0 / 420.0, 319.0 / 420.0, 214.0 / 420.0, -38.0 / 420.0, 4.0 / 420.0};

// recon.F90:202-214 Central8 (sequential dot product)
__device__ __forceinline__ double c8(const double *q) {
    double s = 0.0;
#pragma unroll
    for (int n = 0; n < 8; n++) s += cW8[n] * q[n];
    return s;
}

// recon.F90:303-310 Up7
__device__ __forceinline__ double up7(double q0, double q1, double q2, double q3, double q4, double q5, double q6) {
    return q0 * cU7[0] + q1 * cU7[1] + q2 * cU7[2] + q3 * cU7[3] + q4 * cU7[4] + q5 * cU7[5] + q6 * cU7[6];
}

// recon.F90:275-301 PDM (one-sided)
__device__ __forceinline__ double pdm1(double pdmb, double q0, double q1, double q2, double qI) {
    double maxQ = fmax(q1, q2), minQ = fmin(q1, q2);
    double qN = fmax(minQ, fmin(qI, maxQ));
    double dq0 = pdmb * (q1 - q0);
    double dq1 = pdmb * (q2 - q1);
    double s0 = copysign(1.0, dq0), s1 = copysign(1.0, dq1);
    double s01 = fabs(s0 + s1);
    double dqL = qN - q1;
    return qN - s1 * fmax(0.0, fabs(dqL) - s01 * fabs(dq0));
}

// recon.F90:231-272 pdmLR
__device__ __forceinline__ void pdmlr(double pdmb, double v0, double v1, double v2, double v3, double Qi,
                                      double &Ql, double &Qr) {
    double maxV = fmax(v1, v2), minV = fmin(v1, v2);
    double vN = fmax(minV, fmin(Qi, maxV));
    double dv0 = pdmb * (v1 - v0), dv1 = pdmb * (v2 - v1), dv2 = pdmb * (v3 - v2);
    double s0 = copysign(1.0, dv0), s1 = copysign(1.0, dv1), s2 = copysign(1.0, dv2);
    double q0 = fabs(s0 + s1), q1 = fabs(s1 + s2);
    double dvL = vN - v1, dvR = v2 - vN;
    Ql = vN - s1 * fmax(0.0, fabs(dvL) - q0 * fabs(dv0));
    Qr = vN + s1 * fmax(0.0, fabs(dvR) - q1 * fabs(dv2));
}

// GetLRs for one variable: recon.F90:125-158 (Cen8LRs) / 162-198 (Up7LRs)
template <int RM>
__device__ __forceinline__ void get_lrs(double pdmb, const double *dV, const double *nQ, const double *Q, double Vi,
                                        double &Ql, double &Qr) {
    double QdV[8];
#pragma unroll
    for (int n = 0; n < 8; n++) QdV[n] = dV[n] * Q[n];
    if (RM == 1) {
        double l = up7(QdV[0], QdV[1], QdV[2], QdV[3], QdV[4], QdV[5], QdV[6]) / Vi;
        double r = up7(QdV[7], QdV[6], QdV[5], QdV[4], QdV[3], QdV[2], QdV[1]) / Vi;
        Ql = pdm1(pdmb, nQ[0], nQ[1], nQ[2], l);
        Qr = pdm1(pdmb, nQ[3], nQ[2], nQ[1], r);
    } else {
        double Qi = c8(QdV) / Vi;
        pdmlr(pdmb, nQ[0], nQ[1], nQ[2], nQ[3], Qi, Ql, Qr);
    }
}

// gamutils.F90:301-321 CellC2P
__device__ __forceinline__ void cell_c2p(const MP &m, const double *c, double *w) {
    double rho = c[DEN];
    double Vx = c[MOMX] / rho, Vy = c[MOMY] / rho, Vz = c[MOMZ] / rho;
    double KinE = 0.5 * rho * (sq(Vx) + sq(Vy) + sq(Vz));
    w[0] = fmax(rho, m.dFloor);
    w[1] = Vx; w[2] = Vy; w[3] = Vz;
    w[4] = fmax((m.gamma - 1) * (c[ENERGY] - KinE), m.pFloor);
}

// gamutils.F90:362-383 CellP2C
__device__ __forceinline__ void cell_p2c(const MP &m, const double *w, double *c) {
    double rho = fmax(w[0], m.dFloor);
    double Vx = w[1], Vy = w[2], Vz = w[3];
    double IntE = fmax(w[4], m.pFloor) / (m.gamma - 1);
    double KinE = 0.5 * rho * (sq(Vx) + sq(Vy) + sq(Vz));
    c[0] = rho; c[1] = rho * Vx; c[2] = rho * Vy; c[3] = rho * Vz; c[4] = IntE + KinE;
}

// recon.F90:66-92 BlockCon2Prim
__device__ __forceinline__ void block_c2p(const MP &m, const double *c, double *w) {
    double D = fmax(c[DEN], m.dFloor);
    double KinE = 0.5 * (sq(c[MOMX]) + sq(c[MOMY]) + sq(c[MOMZ])) / D;
    double P = fmax((m.gamma - 1) * (c[ENERGY] - KinE), m.pFloor);
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

__device__ __forceinline__ int tile_of_block(const Tile &t, int nt) {
    if (t.fast & TILE_NOXCD) return (int)blockIdx.x;
    const int per = cdiv(nt, 8);
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

__host__ __device__ inline int tile_count(const Tile &t, int nx, int ny, int nz) {
    if ((t.fast & 3) == TILE_LINEAR) return cdiv(nx * ny * nz, t.ti * t.tj * t.tk);
    return cdiv(nx, t.ti) * cdiv(ny, t.tj) * cdiv(nz, t.tk);
}
__host__ __device__ inline int tile_blocks(const Tile &t, int nx, int ny, int nz) {
    const int nt = tile_count(t, nx, ny, nz);
    return (t.fast & TILE_NOXCD) ? nt : 8 * cdiv(nt, 8);
}

// (a, b, c) of this thread in the box; false if outside
__device__ __forceinline__ bool tile_pos(const Tile &t, int nx, int ny, int nz, int &a, int &b, int &c) {
    if ((t.fast & 3) == TILE_LINEAR) {
        const int n = nx * ny * nz, bs = t.ti * t.tj * t.tk, nt = cdiv(n, bs);
        const int q = tile_of_block(t, nt);
        const int idx = q * bs + (int)threadIdx.x;
        if (q >= nt || idx >= n) return false;
        a = idx % nx;
        const int r = idx / nx;
        b = r % ny;
        c = r / ny;
        return true;
    }
    const int ntx = cdiv(nx, t.ti), nty = cdiv(ny, t.tj), ntz = cdiv(nz, t.tk);
    const int q = tile_of_block(t, ntx * nty * ntz);
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
