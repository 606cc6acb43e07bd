// gam_step.hip -- GAMERA MHD step kernels for gfx950, FP64 and FP32-bulk.
//
// One step (reference gamapp.F90:73-116 stepGamera) is, with S = State,
// O = oState, H = predictor (half-step) state:
//   k_clock_begin     time trap + pdt/odt            mhdgroup.F90:25-31,459
//   k_predict   (F)   AB2 magFlux, PredictCell       mhdgroup.F90:445-616
//   k_b2f       (G)   H.Bxyz = bFlux2Fld(H.magFlux)  gridutils.F90:281-301,648-680
//   k_vf        (G)   Mom2Face for the 3 sweeps      fields.F90:110-179,235-285
//   k_emf<d>    (E)   edge EMFs, one direction each  fields.F90:186-226,288-517
//   k_flux<d>   (I)   interface fluxes, 3 launches   stress.F90:84-128,204-599
//   k_ct        (F)   O.magFlux = E2Flux(S.magFlux)  fields.F90:18-51
//   k_update    (G)   Flux2Deltas + Reynolds + Bxyz + Maxwell/Boris + CellDT min
//                                                    stress.F90:152-194, mhdgroup.F90:79-442,
//                                                    gridutils.F90:49-106
//   k_dtfinal   (1)   time update + CalcDT checks    gamapp.F90:112-114, step.F90:58-168
//   k_bc        (F)   periodic ghost fill            bcs.F90:279-450
//   k_b2f       (G)   O.Bxyz = bFlux2Fld(O.magFlux)  gamapp.F90:89
// after which the host swaps the S and O buffer sets (CopyState becomes a
// pointer swap, mhdgroup.F90:202-257).
//
// Precision: the kernels are templates on the bulk real type R.  R = double is
// the FP64 step.  R = float is the FP32-bulk step: Gas, Bxyz, the predictor
// state, the stencil inputs and the fluxes/EMFs are stored and computed in
// FP32, while magFlux, the CT face update (E2Flux), Flux2Deltas with the
// Reynolds/Maxwell/Boris update and its pressure recovery, bFlux2Fld and the
// CFL bound stay FP64 (values are promoted on load).
#include "gam_common.h"

// ----------------------------------------------------------------- helpers
// Streaming (non-temporal) store, used where it measured faster: the Mom2Face
// and update outputs, which later stages read back from HBM anyway.
template <typename T> __device__ __forceinline__ void st_nt(T *p, T v) {
#ifdef GAM_EMU
    *p = v;
#else
    __builtin_nontemporal_store(v, p);
#endif
}

template <typename R> __device__ __forceinline__ void decode_f(const GVT<R> &g, int idx, int &i, int &j, int &k) {
    int nx = g.Ni + 1, ny = g.Nj + 1;
    i = idx % nx; int r = idx / nx; j = r % ny; k = r / ny;
}
template <typename R> __device__ __forceinline__ void decode_c(const GVT<R> &g, int idx, int &i, int &j, int &k) {
    i = idx % g.Ni; int r = idx / g.Ni; j = r % g.Nj; k = r / g.Nj;
}

// mhdgroup.F90:710-718 ExtrapAB2
template <typename T> __device__ __forceinline__ T ab2(T dt, T Q, T oQ, T odt) {
    T dQ = (Q - oQ) / odt;
    return Q + dt * dQ;
}

// gridutils.F90:648-680 gamCellBxyz, in FP64 from a magFlux of type TM.
// FP32 bulk: the same sum over the six faces with the cell's lever arms from
// the table g.b2fc (18 floats per cell instead of 13 FP64 positions and the
// volume; the stored Bxyz is FP32 either way).
template <typename R, typename TM>
__device__ __forceinline__ void cell_bxyz(const GVT<R> &g, const TM *mag, int i, int j, int k, double *B) {
    const int nf = g.nf;
    int c = ci(g, i, j, k), f0 = fi(g, i, j, k);
    int fi1 = f0 + 1, fj1 = f0 + (g.Ni + 1), fk1 = f0 + (g.Ni + 1) * (g.Nj + 1);
    if constexpr (sizeof(R) == sizeof(float)) {
        if (g.b2fc) {
            const size_t nc = g.nc;
            const double fl[6] = {(double)mag[f0], (double)mag[fi1], (double)mag[nf + f0], (double)mag[nf + fj1],
                                  (double)mag[2 * (size_t)nf + f0], (double)mag[2 * (size_t)nf + fk1]};
#pragma unroll
            for (int x = 0; x < 3; x++) {
                double b = fl[0] * (double)g.b2fc[(size_t)x * nc + c];
#pragma unroll
                for (int q = 1; q < 6; q++) b = b + fl[q] * (double)g.b2fc[(size_t)(3 * q + x) * nc + c];
                B[x] = b;
            }
            return;
        }
    }
    double dV = g.dvol[c];
    double fI = mag[f0], fJ = mag[nf + f0], fK = mag[2 * nf + f0];
    double fIp = mag[fi1], fJp = mag[nf + fj1], fKp = mag[2 * (size_t)nf + fk1];
    double Div = fIp - fI + fJp - fJ + fKp - fK;
#pragma unroll
    for (int x = 0; x < 3; x++) {
        const double *xf = g.xfc;
        double xfI = xf[(size_t)nf * (x + 0) + f0], xfIp = xf[(size_t)nf * (x + 0) + fi1];
        double xfJ = xf[(size_t)nf * (x + 3) + f0], xfJp = xf[(size_t)nf * (x + 3) + fj1];
        double xfK = xf[(size_t)nf * (x + 6) + f0], xfKp = xf[(size_t)nf * (x + 6) + fk1];
        double b = (fIp * xfIp + fJp * xfJp + fKp * xfKp - fI * xfI - fJ * xfJ - fK * xfK) / dV;
        B[x] = b - Div * g.xyzcc[(size_t)g.nc * x + c] / dV;
    }
}

// FP32 bulk: the lever arms of every cell for cell_bxyz, +-(face centre - cell
// centre)/volume of its lower (-) and upper (+) I, J, K faces, formed in FP64
__global__ void __launch_bounds__(256) k_b2f_coef(GVT<float> g, const double *__restrict__ xf, const double *__restrict__ xc,
                                                  const double *__restrict__ dvol, float *__restrict__ out) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= g.nc) return;
    int i, j, k;
    decode_c(g, c, i, j, k);
    const size_t nc = g.nc, nf = g.nf;
    const int f0 = fi(g, i, j, k);
    const int fu[3] = {f0 + 1, f0 + (g.Ni + 1), f0 + (g.Ni + 1) * (g.Nj + 1)};
    const double dV = dvol[c];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int x = 0; x < 3; x++) {
            const double cx = xc[nc * x + c];
            const double lo = xf[nf * (x + 3 * d) + f0], up = xf[nf * (x + 3 * d) + fu[d]];
            out[(size_t)(3 * (2 * d) + x) * nc + c] = (float)(-(lo - cx) / dV);
            out[(size_t)(3 * (2 * d + 1) + x) * nc + c] = (float)((up - cx) / dV);
        }
}

// The same, with the lower K face's flux and centre passed in (kc.fK, kc.xfK)
// and the upper K face's returned in their place: a thread walking up a
// column of cells reads each K face once.
struct KCarry { double fK, xfK[3]; };
template <typename R, typename TM>
__device__ __forceinline__ void cell_bxyz_kc(const GVT<R> &g, const TM *mag, int i, int j, int k, double *B,
                                             KCarry &kc, bool have) {
    const int nf = g.nf;
    int c = ci(g, i, j, k), f0 = fi(g, i, j, k);
    int fi1 = f0 + 1, fj1 = f0 + (g.Ni + 1), fk1 = f0 + (g.Ni + 1) * (g.Nj + 1);
    const double *xf = g.xfc;
    if (!have) {
        kc.fK = mag[2 * nf + f0];
#pragma unroll
        for (int x = 0; x < 3; x++) kc.xfK[x] = xf[(size_t)nf * (x + 6) + f0];
    }
    double dV = g.dvol[c];
    double fI = mag[f0], fJ = mag[nf + f0], fK = kc.fK;
    double fIp = mag[fi1], fJp = mag[nf + fj1], fKp = mag[2 * (size_t)nf + fk1];
    double Div = fIp - fI + fJp - fJ + fKp - fK;
#pragma unroll
    for (int x = 0; x < 3; x++) {
        double xfI = xf[(size_t)nf * (x + 0) + f0], xfIp = xf[(size_t)nf * (x + 0) + fi1];
        double xfJ = xf[(size_t)nf * (x + 3) + f0], xfJp = xf[(size_t)nf * (x + 3) + fj1];
        double xfK = kc.xfK[x], xfKp = xf[(size_t)nf * (x + 6) + fk1];
        double b = (fIp * xfIp + fJp * xfJp + fKp * xfKp - fI * xfI - fJ * xfJ - fK * xfK) / dV;
        B[x] = b - Div * g.xyzcc[(size_t)g.nc * x + c] / dV;
        kc.xfK[x] = xfKp;
    }
    kc.fK = fKp;
}

// cells per k_update_col thread along k (a column walked bottom up)
#define UPDATE_KC 4

// ----------------------------------------------------------------- clock
__device__ __forceinline__ void b_clock_begin(const unsigned vb, DClock *clk) {
    // mhdgroup.F90:25-28 trap and :31,459 pdt/odt
    if (threadIdx.x == 0 && vb == 0 && !clk->halt) {
        if (fabs(clk->otime - clk->stime) < GAM_TINY) clk->otime = clk->stime - clk->dt;
        clk->odt = clk->stime - clk->otime;
        clk->pdt = 0.5 * clk->dt;
    }
}
__global__ void k_clock_begin(DClock *clk) {
    b_clock_begin(blockIdx.x, clk);
}

// ----------------------------------------------------------------- predictor
// magFlux AB2 (FP64 in, bulk out) on every face point; PredictCell on every
// cell plus the per-point stencil inputs (WKT).
template <typename R>
__device__ __forceinline__ void b_predict(const unsigned vb, GVT<R> g, MP m, const DClock *__restrict__ clk,
                                                 const R *__restrict__ Gas, const R *__restrict__ oGas,
                                                 const double *__restrict__ mag, const double *__restrict__ omag,
                                                 R *__restrict__ hGas, R *__restrict__ hmag, WKT<R> w) {
    int idx = vb * blockDim.x + threadIdx.x;
    if (idx >= g.nf || clk->halt) return;
    const double pdt = clk->pdt, odt = clk->odt;
    const int nf = g.nf, nc = g.nc;
    if (m.doMHD) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double mn = mag[(size_t)nf * d + idx];
            double hm = ab2(pdt, mn, omag[(size_t)nf * d + idx], odt);
            const R hmR = R(hm);
            hmag[(size_t)nf * d + idx] = hmR;
            R fa = rmax(g.face[(size_t)nf * d + idx], R(GAM_TINY));
            w.bh[(size_t)nf * d + idx] = hmR / fa;
            w.bn[(size_t)nf * d + idx] = (m.doLFMLim ? R(mn) : hmR) / fa;
        }
    }
    int i, j, k;
    decode_f(g, idx, i, j, k);
    if (i >= g.Ni || j >= g.Nj || k >= g.Nk) return;
    int c = ci(g, i, j, k);
    R Q[5], oQ[5], W[5], oW[5], pW[5], pQ[5];
#pragma unroll
    for (int v = 0; v < 5; v++) { Q[v] = Gas[(size_t)nc * v + c]; oQ[v] = oGas[(size_t)nc * v + c]; }
    cell_c2p(m, oQ, oW);
    cell_c2p(m, Q, W);
#pragma unroll
    for (int v = 0; v < 5; v++) pW[v] = ab2(R(pdt), W[v], oW[v], R(odt));
    pW[0] = rmax(pW[0], R(m.dFloor));
    pW[4] = rmax(pW[4], R(m.pFloor));
    cell_p2c(m, pW, pQ);
#pragma unroll
    for (int v = 0; v < 5; v++) hGas[(size_t)nc * v + c] = pQ[v];
    // stencil-point quantities of this cell
    R Wh[5], Wn[5];
    block_c2p(m, pQ, Wh);
    if (m.doLFMLim) block_c2p(m, Q, Wn);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        w.Wh[(size_t)nc * v + c] = Wh[v];
        w.Wn[(size_t)nc * v + c] = m.doLFMLim ? Wn[v] : Wh[v];
    }
    const R vol = g.vol[c];
    const R D = rmax(pQ[0], R(m.dFloor));
#pragma unroll
    for (int x = 0; x < 3; x++) w.VdV[(size_t)nc * x + c] = vol * pQ[1 + x] / D;
}
template <typename R>
__global__ void __launch_bounds__(256) k_predict(GVT<R> g, MP m, const DClock *__restrict__ clk,
                                                 const R *__restrict__ Gas, const R *__restrict__ oGas,
                                                 const double *__restrict__ mag, const double *__restrict__ omag,
                                                 R *__restrict__ hGas, R *__restrict__ hmag, WKT<R> w) {
    b_predict<R>(blockIdx.x, g, m, clk, Gas, oGas, mag, omag, hGas, hmag, w);
}

// bFlux2Fld over all cells, FP64 arithmetic, stored in the bulk type
template <typename R, typename TM>
__device__ __forceinline__ void b_b2f(const unsigned vb, GVT<R> g, Tile t, const DClock *__restrict__ clk, const TM *__restrict__ mag,
                                             R *__restrict__ B) {
    int i, j, k;
    if (!tile_pos(vb, t, g.Ni, g.Nj, g.Nk, i, j, k) || clk->halt) return;
    const int idx = ci(g, i, j, k);
    double b[3];
    cell_bxyz(g, mag, i, j, k, b);
#pragma unroll
    for (int x = 0; x < 3; x++) B[(size_t)g.nc * x + idx] = R(b[x]);
}
template <typename R, typename TM>
__global__ void __launch_bounds__(256) k_b2f(GVT<R> g, Tile t, const DClock *__restrict__ clk, const TM *__restrict__ mag,
                                             R *__restrict__ B) {
    b_b2f<R, TM>(blockIdx.x, g, t, clk, mag, B);
}

// bFlux2Fld on the ghost cells only.  Used after the BCs of a fused 3D step:
// every BC type rewrites ghost faces only (an active cell's six faces are
// owned faces; a normal component's first ghost face is past the owned upper
// face), so the active cells' Bxyz from k_update (or, on the ring rows, from
// the ring average), computed by the same cell_bxyz from the same faces, is
// already what a full pass would write.
// Shell order: the k ghost planes, then the j ghost rows of the active planes,
// then the i ghosts of the active rows.
__host__ __device__ __forceinline__ long b2f_shell_count(int Ni, int Nj, int Nk, int Nip, int Njp, int Nkp) {
    return (long)(Nk - Nkp) * Ni * Nj + (long)Nkp * (Nj - Njp) * Ni + (long)Nkp * Njp * (Ni - Nip);
}
// (i, j, k) of ghost cell idx in that order; false past the shell
template <typename R>
__device__ __forceinline__ bool shell_cell(const GVT<R> &g, long idx, int &i, int &j, int &k) {
    const int Ni = g.Ni, Nj = g.Nj, a0 = g.nG;
    const long nplane = (long)(g.Nk - g.Nkp) * Ni * Nj;
    const long nrow = (long)g.Nkp * (Nj - g.Njp) * Ni;
    const long nig = (long)g.Nkp * g.Njp * (Ni - g.Nip);
    if (idx >= nplane + nrow + nig) return false;
    if (idx < nplane) {
        const int p = (int)(idx / ((long)Ni * Nj)), r = (int)(idx % ((long)Ni * Nj));
        k = p < a0 ? p : p + g.Nkp;
        j = r / Ni;
        i = r % Ni;
    } else if (idx < nplane + nrow) {
        const long q = idx - nplane;
        const int row = (int)(q / Ni), gj = Nj - g.Njp;
        i = (int)(q % Ni);
        k = a0 + row / gj;
        const int t = row % gj;
        j = t < a0 ? t : t + g.Njp;
    } else {
        const long q = idx - nplane - nrow;
        const int gi = Ni - g.Nip;
        const int t = (int)(q % gi), r = (int)(q / gi);
        i = t < a0 ? t : t + g.Nip;
        j = a0 + r % g.Njp;
        k = a0 + r / g.Njp;
    }
    return true;
}

template <typename R>
__device__ __forceinline__ void b_b2f_shell(const unsigned vb, GVT<R> g, const DClock *__restrict__ clk, const double *__restrict__ mag,
                                                   R *__restrict__ B) {
    const long idx = (long)vb * blockDim.x + threadIdx.x;
    int i, j, k;
    if (clk->halt || !shell_cell(g, idx, i, j, k)) return;
    const int c = ci(g, i, j, k);
    double b[3];
    cell_bxyz(g, mag, i, j, k, b);
#pragma unroll
    for (int x = 0; x < 3; x++) B[(size_t)g.nc * x + c] = R(b[x]);
}
template <typename R>
__global__ void __launch_bounds__(256) k_b2f_shell(GVT<R> g, const DClock *__restrict__ clk, const double *__restrict__ mag,
                                                   R *__restrict__ B) {
    b_b2f_shell<R>(blockIdx.x, g, clk, mag, B);
}

// ----------------------------------------------------------------- E field
// fields.F90:235-285 Mom2Face at (i,j,k) along dT, c = cell index; the
// per-cell summand vol*M/max(D,dFloor) comes precomputed in VdV
template <typename R>
__device__ __forceinline__ void mom2face(const GVT<R> &g, const R *__restrict__ VdV, int c, int dT, R *Vf) {
    const int st = cstride(g, dT), nc = g.nc;
    R VolB[8], Vx[8], Vy[8], Vz[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        int cc = c + (n - 4) * st;
        VolB[n] = g.vol[cc];
        Vx[n] = VdV[cc];
        Vy[n] = VdV[(size_t)nc + cc];
        Vz[n] = VdV[2 * (size_t)nc + cc];
    }
    R dV = rmax(c8(VolB), R(GAM_TINY));
    Vf[0] = c8(Vx) / dV;
    Vf[1] = c8(Vy) / dV;
    Vf[2] = c8(Vz) / dV;
}

// Vf arrays: VfA[dT1][xyz] stored as slab (3*dT1 + xyz).  Each thread walks
// VF_KC cells up a k column; the k-direction stencil (8 points along k) is a
// window kept in registers, shifted by one point per cell.
#define VF_KC 8
template <typename R>
__device__ __forceinline__ void b_vf(const unsigned vb, GVT<R> g, Tile t, const DClock *__restrict__ clk, const R *__restrict__ VdV,
                                            R *__restrict__ Vf) {
    int i, j, kb;
    if (!tile_pos(vb, t, g.Ni, g.Nj, cdiv(g.Nk, VF_KC), i, j, kb) || clk->halt) return;
    const int a0 = g.nG;
    const size_t nc = g.nc;
    const int sk = cstride(g, 2);
    R wVol[8], wX[8], wY[8], wZ[8];   // k window of the K-face Mom2Face: points k-4 .. k+3
    bool haveW = false;
#pragma unroll 1
    for (int kk = 0; kk < VF_KC; kk++) {
        const int k = kb * VF_KC + kk;
        if (k >= g.Nk) break;
        const int idx = ci(g, i, j, k);
        R v[3];
        // dT1 = I (used by K edges): i in is..ie+1   (fields.F90:161-175)
        if (i >= a0 && i <= a0 + g.Nip) {
            mom2face(g, VdV, idx, 0, v);
            st_nt(&Vf[nc * 0 + idx], v[0]); st_nt(&Vf[nc * 1 + idx], v[1]); st_nt(&Vf[nc * 2 + idx], v[2]);
        }
        // dT1 = J (used by I edges): j in js..je+1   (fields.F90:119-133)
        if (j >= a0 && j <= a0 + g.Njp) {
            mom2face(g, VdV, idx, 1, v);
            st_nt(&Vf[nc * 3 + idx], v[0]); st_nt(&Vf[nc * 4 + idx], v[1]); st_nt(&Vf[nc * 5 + idx], v[2]);
        }
        // dT1 = K (used by J edges): k in ks..ke+1   (fields.F90:140-154)
        if (k >= a0 && k <= a0 + g.Nkp) {
            if (!haveW) {
#pragma unroll
                for (int n = 0; n < 8; n++) {
                    const int cc = idx + (n - 4) * sk;
                    wVol[n] = g.vol[cc]; wX[n] = VdV[cc]; wY[n] = VdV[nc + cc]; wZ[n] = VdV[2 * nc + cc];
                }
                haveW = true;
            } else {
#pragma unroll
                for (int n = 0; n < 7; n++) { wVol[n] = wVol[n + 1]; wX[n] = wX[n + 1]; wY[n] = wY[n + 1]; wZ[n] = wZ[n + 1]; }
                const int cc = idx + 3 * sk;
                wVol[7] = g.vol[cc]; wX[7] = VdV[cc]; wY[7] = VdV[nc + cc]; wZ[7] = VdV[2 * nc + cc];
            }
            const R dV = rmax(c8(wVol), R(GAM_TINY));
            st_nt(&Vf[nc * 6 + idx], c8(wX) / dV); st_nt(&Vf[nc * 7 + idx], c8(wY) / dV); st_nt(&Vf[nc * 8 + idx], c8(wZ) / dV);
        }
    }
}
template <typename R>
__global__ void __launch_bounds__(256) k_vf(GVT<R> g, Tile t, const DClock *__restrict__ clk, const R *__restrict__ VdV,
                                            R *__restrict__ Vf) {
    b_vf<R>(blockIdx.x, g, t, clk, VdV, Vf);
}

// fields.F90:288-362 EdgeLRs: dT1 faces reconstructed along dT2 at face point f;
// the face-normal field bh/bn = flux/max(area,TINY) comes precomputed
template <int RM, typename R>
__device__ __forceinline__ void edge_lrs(const GVT<R> &g, const MP &m, const R *__restrict__ bn,
                                         const R *__restrict__ bh, int f, int dT1, int dT2,
                                         R &bL, R &bR, R &fA) {
    const int st = fstride(g, dT2);
    const size_t off = (size_t)g.nf * dT1;
    R AreaB[8], MagB[8], nMagB[4];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        int ff = f + (n - 4) * st;
        AreaB[n] = g.face[off + ff];
        MagB[n] = bh[off + ff];
    }
#pragma unroll
    for (int n = 0; n < 4; n++) nMagB[n] = bn[off + f + (n - 2) * st];
    R Vi = c8(AreaB);
    R l, r;
    get_lrs<RM>(R(m.pdmb), AreaB, nMagB, MagB, R(1) / Vi, l, r);
    bL = l * Vi; bR = r * Vi; fA = Vi;
}

// The same from stencils already loaded (AreaB, MagB: 8 points; nMagB: 4)
template <int RM, typename R>
__device__ __forceinline__ void edge_lrs_w(const MP &m, const R *AreaB, const R *MagB, const R *nMagB,
                                           R &bL, R &bR, R &fA) {
    R Vi = c8(AreaB);
    R l, r;
    get_lrs<RM>(R(m.pdmb), AreaB, nMagB, MagB, R(1) / Vi, l, r);
    bL = l * Vi; bR = r * Vi; fA = Vi;
}

// One edge EMF (fields.F90:196-221) for edge direction eD at corner (i,j,k)
// wA/wB/wN: the dT2 faces' stencils along dT1 (face, bh: 8 points; bn: 4)
// already loaded (a k-column walker's register window), or null.
template <int RM, int BORIS, typename R>
__device__ __forceinline__ R emf_edge(const GVT<R> &g, const MP &m, const BGV &bg, double dt64, const R *__restrict__ Vf,
                                      const R *__restrict__ hGas, const R *__restrict__ bn,
                                      const R *__restrict__ bh, int eD, int i, int j, int k,
                                      const R *wA = nullptr, const R *wB = nullptr, const R *wN = nullptr) {
    const R dt = R(dt64);
    const int dT1 = (eD + 1) % 3, dT2 = (eD + 2) % 3;
    const int f = fi(g, i, j, k), c = ci(g, i, j, k);
    const size_t nf = g.nf, nc = g.nc;
    // GetCornerV fields.F90:457-517 (Vf of the dT1 faces interpolated along dT2)
    const int fst = fstride(g, dT2), cst = cstride(g, dT2);
    const R *VfA = Vf + nc * 3 * dT1;
    R AreaB[8], Vx[8], Vy[8], Vz[8];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        AreaB[n] = g.face[nf * dT1 + f + (n - 4) * fst];
        int cc = c + (n - 4) * cst;
        Vx[n] = AreaB[n] * VfA[cc];
        Vy[n] = AreaB[n] * VfA[nc + cc];
        Vz[n] = AreaB[n] * VfA[2 * nc + cc];
    }
    R dA = rmax(c8(AreaB), R(GAM_TINY));
    R Vcx = c8(Vx) / dA, Vcy = c8(Vy) / dA, Vcz = c8(Vz) / dA;
    const R *Te = g.Te + nf * 9 * eD;
    R v1 = Vcx * Te[nf * TAN1X + f] + Vcy * Te[nf * TAN1Y + f] + Vcz * Te[nf * TAN1Z + f];
    R v2 = Vcx * Te[nf * TAN2X + f] + Vcy * Te[nf * TAN2Y + f] + Vcz * Te[nf * TAN2Z + f];
    R vDiff = rroot(sq(v1) + sq(v2));
    // GetCornerD fields.F90:427-454
    // normal plane basis (metric.F90:151-167): I -> (j,k), J -> (i,k), K -> (i,j)
    const int s1 = cstride(g, eD == 0 ? 1 : 0), s2 = cstride(g, eD == 2 ? 1 : 2);
    R Dc = R(0.25) * (hGas[c] + hGas[c - s1] + hGas[c - s2] + hGas[c - s1 - s2]);
    // GetCornerB fields.F90:365-424
    R bT1L, bT1R, fA1, bT2L, bT2R, fA2;
    edge_lrs<RM>(g, m, bn, bh, f, dT1, dT2, bT1L, bT1R, fA1);
    if (wA) edge_lrs_w<RM>(m, wA, wB, wN, bT2L, bT2R, fA2);
    else edge_lrs<RM>(g, m, bn, bh, f, dT2, dT1, bT2L, bT2R, fA2);
    const R *Teb = g.Teb + nf * 4 * eD;
    R xni = Teb[nf * XNQI + f], yni = Teb[nf * YNQI + f];
    R xnj = Teb[nf * XNQJ + f], ynj = Teb[nf * YNQJ + f];
    R detT = R(1.0) / (xni * ynj - xnj * yni);
    R bT1 = R(0.5) * (bT1L + bT1R) / fA1;
    R bT2 = R(0.5) * (bT2L + bT2R) / fA2;
    R db1 = (bT1R - bT1L) / fA1;
    R db2 = (bT2R - bT2L) / fA2;
    R Jd = db2 - db1;
    R b1 = detT * (ynj * bT1 - yni * bT2);
    R b2 = detT * (-xnj * bT1 + xni * bT2);
    if (bg.on) {   // fields.F90:417-420
        b1 = b1 + R(bg.edgB0[nc * (0 + 2 * eD) + c]);
        b2 = b2 + R(bg.edgB0[nc * (1 + 2 * eD) + c]);
    }
    R vA = rroot(sq(b1) + sq(b2)) / rroot(Dc);
    const R Ca = R(m.Ca);
    if (BORIS) vA = Ca * vA / rroot(sq(Ca) + sq(vA));
    vDiff = vDiff + vA;
    R ed = g.edge[nf * eD + f];
    vDiff = rmin(vDiff, R(m.CFL) * ed / dt);
    R e = -(v1 * b2 - v2 * b1) + R(m.Vd0) * vDiff * Jd;
    return e * ed;
}

// The learned coarse-grid correction, added where an EMF or a flux is stored
// (periodic blocks only): scale * edge * o[c][5 + d] on the EMF of every active
// edge of direction d, scale * face * o[c][v] on the gas flux v (0..4) of every
// interface of the flux box, c the cell at the edge's lower corner or below the
// interface, o read at c's periodic image (so the two copies of a boundary face
// or edge get one correction).  The sums are FP64, rounded once to R.
__device__ __forceinline__ int pwrap(int x, int a0, int n) { return a0 + ((x - a0) % n + n) % n; }
template <typename R> __device__ __forceinline__ size_t pimg(const GVT<R> &g, int i, int j, int k) {
    return (size_t)ci(g, pwrap(i, g.nG, g.Nip), pwrap(j, g.nG, g.Njp), pwrap(k, g.nG, g.Nkp));
}
template <typename R> __device__ __forceinline__ R corr_add(R x, double s, double len, float o) {
    return R((double)x + s * len * (double)o);
}

// Edges of direction ED over their active box (fields.F90:110-225): I edges
// [is,ie]x[js,je+1]x[ks,ke+1], J edges [is,ie+1]x[js,je]x[ks,ke+1], K edges
// [is,ie+1]x[js,je+1]x[ks,ke]; E of edge ED at its lower corner (i,j,k).  E
// slabs are written only on active edges; k_ct treats the rest as zero
// (CalcElecField zeroes E, fields.F90:89).  One kernel per direction, so each
// gets a launch brick shaped for its two stencil axes (dT1 and dT2).
template <int ED, int RM, int BORIS, typename R>
__device__ __forceinline__ void b_emf(const unsigned vb, GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk, const R *__restrict__ Vf,
                                             const R *__restrict__ hGas, const R *__restrict__ bn,
                                             const R *__restrict__ bh, R *__restrict__ E,
                                             const float *__restrict__ corr, double cs) {
    int a, b, cz;
    if (!tile_pos(vb, t, g.Nip + (ED != 0), g.Njp + (ED != 1), g.Nkp + (ED != 2), a, b, cz) || clk->halt) return;
    const int i = a + g.nG, j = b + g.nG, k = cz + g.nG;
    R e = emf_edge<RM, BORIS>(g, m, bg, clk->dt, Vf, hGas, bn, bh, ED, i, j, k);
    if (corr) e = corr_add(e, cs, (double)g.edge[(size_t)g.nf * ED + fi(g, i, j, k)], corr[pimg(g, i, j, k) * 8 + 5 + ED]);
    E[(size_t)g.nc * ED + ci(g, i, j, k)] = e;
}
template <int ED, int RM, int BORIS, typename R>
__global__ void __launch_bounds__(256) k_emf(GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk, const R *__restrict__ Vf,
                                             const R *__restrict__ hGas, const R *__restrict__ bn,
                                             const R *__restrict__ bh, R *__restrict__ E,
                                             const float *__restrict__ corr, double cs) {
    b_emf<ED, RM, BORIS, R>(blockIdx.x, g, t, m, bg, clk, Vf, hGas, bn, bh, E, corr, cs);
}

// J edges walking EMF_KC edges up a k column: the I faces' stencils along k
// (the one k-direction stencil of a J edge) are a register window shifted by
// one point per edge.  Box [is,ie+1]x[js,je]x[ks,ke+1].
#define EMF_KC 8
template <int RM, int BORIS, typename R>
__device__ __forceinline__ void b_emf_jcol(const unsigned vb, GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                                  const R *__restrict__ Vf, const R *__restrict__ hGas,
                                                  const R *__restrict__ bn, const R *__restrict__ bh, R *__restrict__ E,
                                                  const float *__restrict__ corr, double cs) {
    const int nz = g.Nkp + 1;
    int a, b, cb;
    if (!tile_pos(vb, t, g.Nip + 1, g.Njp, cdiv(nz, EMF_KC), a, b, cb) || clk->halt) return;
    const int i = a + g.nG, j = b + g.nG;
    const int st = fstride(g, 2);
    const R *fI = g.face, *bhI = bh, *bnI = bn;   // I-face slabs (direction 0)
    R wA[8], wB[8], wN[4];
#pragma unroll 1
    for (int kk = 0; kk < EMF_KC; kk++) {
        const int cz = cb * EMF_KC + kk;
        if (cz >= nz) break;
        const int k = cz + g.nG;
        const int f = fi(g, i, j, k);
        if (kk == 0) {
#pragma unroll
            for (int n = 0; n < 8; n++) { wA[n] = fI[f + (n - 4) * st]; wB[n] = bhI[f + (n - 4) * st]; }
#pragma unroll
            for (int n = 0; n < 4; n++) wN[n] = bnI[f + (n - 2) * st];
        } else {
#pragma unroll
            for (int n = 0; n < 7; n++) { wA[n] = wA[n + 1]; wB[n] = wB[n + 1]; }
#pragma unroll
            for (int n = 0; n < 3; n++) wN[n] = wN[n + 1];
            wA[7] = fI[f + 3 * st]; wB[7] = bhI[f + 3 * st]; wN[3] = bnI[f + st];
        }
        R e = emf_edge<RM, BORIS>(g, m, bg, clk->dt, Vf, hGas, bn, bh, 1, i, j, k, wA, wB, wN);
        if (corr) e = corr_add(e, cs, (double)g.edge[(size_t)g.nf * 1 + f], corr[pimg(g, i, j, k) * 8 + 6]);
        E[(size_t)g.nc * 1 + ci(g, i, j, k)] = e;
    }
}
template <int RM, int BORIS, typename R>
__global__ void __launch_bounds__(256) k_emf_jcol(GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                                  const R *__restrict__ Vf, const R *__restrict__ hGas,
                                                  const R *__restrict__ bn, const R *__restrict__ bh, R *__restrict__ E,
                                                  const float *__restrict__ corr, double cs) {
    b_emf_jcol<RM, BORIS, R>(blockIdx.x, g, t, m, bg, clk, Vf, hGas, bn, bh, E, corr, cs);
}

// ----------------------------------------------------------------- fluxes
// stress.F90:513-599 MagKinFlux (no background field), q = -1/+1
template <int BORIS, typename R>
__device__ __forceinline__ void magkin(const MP &m, R sg, const R *P5, const R *B, R Bn,
                                       const R *T, bool bgon, const R *B0, R B0n, R *mF, R &bbD, R &VaD) {
    R D = P5[0], P = P5[4];
    R Bx = B[0], By = B[1], Bz = B[2];
    R dPb = R(0.5) * (sq(Bx) + sq(By) + sq(Bz));
    R Va2;
    if (bgon) {   // stress.F90:550-552 total field
        Va2 = (sq(Bx + B0[0]) + sq(By + B0[1]) + sq(Bz + B0[2])) / D;
        bbD = bbD + R(0.5) * (sq(Bx + B0[0]) + sq(By + B0[1]) + sq(Bz + B0[2]));
    } else {
        Va2 = 2 * dPb / D;
        bbD = bbD + R(0.5) * (sq(Bx) + sq(By) + sq(Bz));
    }
    R Vac2 = BORIS ? Va2 / (1 + (Va2 / sq(R(m.Ca)))) : Va2;
    R Nx = T[NORMX], Ny = T[NORMY], Nz = T[NORMZ];
    R Vn = Nx * P5[1] + Ny * P5[2] + Nz * P5[3];
    // lambda = 1/(2P/D + Vac2); sqrt(lambda)*Vn formed as Vn/sqrt(1/lambda) (one division, same value to an ulp)
    const R xv = Vn / rroot(2 * P / D + Vac2);
    R Vn0 = R(0.5) * rerfc_e(sg * xv, rexp(-xv * xv));
    VaD = VaD + R(0.5) * rroot(Vac2);
    mF[0] = mF[0] + Vn0 * (dPb * Nx - Bx * Bn);
    mF[1] = mF[1] + Vn0 * (dPb * Ny - By * Bn);
    mF[2] = mF[2] + Vn0 * (dPb * Nz - Bz * Bn);
    if (bgon) {   // stress.F90:589-596 cross terms
        R BdB0 = B0[0] * Bx + B0[1] * By + B0[2] * Bz;
        mF[0] = mF[0] + Vn0 * (-B0[0] * Bn - Bx * B0n + BdB0 * Nx);
        mF[1] = mF[1] + Vn0 * (-B0[1] * Bn - By * B0n + BdB0 * Ny);
        mF[2] = mF[2] + Vn0 * (-B0[2] * Bn - Bz * B0n + BdB0 * Nz);
    }
}

// stress.F90:432-511 GasKinFlux
template <typename R>
__device__ __forceinline__ void gaskin(const MP &m, R sg, const R *P5, const R *T, R *gF, R *dW) {
    R D = P5[0], P = P5[4], Vx = P5[1], Vy = P5[2], Vz = P5[3];
    R E = R(0.5) * D * (sq(Vx) + sq(Vy) + sq(Vz)) + P / R(m.gamma - 1);
    R lambda = D / (2 * P);
    R Vn = T[NORMX] * Vx + T[NORMY] * Vy + T[NORMZ] * Vz;
    R Vt1 = T[TAN1X] * Vx + T[TAN1Y] * Vy + T[TAN1Z] * Vz;
    R Vt2 = T[TAN2X] * Vx + T[TAN2Y] * Vy + T[TAN2Z] * Vz;
    const R e = rexp(-lambda * Vn * Vn);
    const R sl = rroot(lambda);   // sqrt(pi*lambda) = sqrt(pi)*sqrt(lambda), to an ulp
    R Vn0 = R(0.5) * rerfc_e(sg * sl * Vn, e);
    R Vn1 = Vn * Vn0 - R(0.5) * sg * e / (R(GAM_SQRTPI) * sl);
    gF[0] = gF[0] + D * Vn1;
    gF[4] = gF[4] + (E + R(0.5) * P) * Vn1 + R(0.5) * P * Vn0 * Vn;
    R fMn = D * Vn * Vn1 + P * Vn0;
    R fMt1 = D * Vt1 * Vn1;
    R fMt2 = D * Vt2 * Vn1;
    gF[1] = gF[1] + (T[NORMX] * fMn + T[TAN1X] * fMt1 + T[TAN2X] * fMt2);
    gF[2] = gF[2] + (T[NORMY] * fMn + T[TAN1Y] * fMt1 + T[TAN2Y] * fMt2);
    gF[3] = gF[3] + (T[NORMZ] * fMn + T[TAN1Z] * fMt1 + T[TAN2Z] * fMt2);
    dW[0] = dW[0] + sg * D;
    dW[1] = dW[1] + sg * D * Vx;
    dW[2] = dW[2] + sg * D * Vy;
    dW[3] = dW[3] + sg * D * Vz;
    dW[4] = dW[4] + sg * E;
}

// One interface of direction DN (stress.F90:204-429) at its cell c and face
// point f: out[0..4] the gas fluxes, out[5..7] the magnetic ones (MHD), each
// times the face area.  BG: the background-field terms are compiled in (a
// separate instantiation keeps the plain kernel's register count).
// The flux of one interface from its reconstructed states (stress.F90
// Fluxes after BlockStateLRs): kinetic gas and magnetic fluxes, HOGS, the
// Boris nuke; out[0..4] gas and out[5..7] magnetic fluxes times the face area.
template <int DN, int BORIS, int MHD, int BG, typename R>
__device__ __forceinline__ void lr_flux(const GVT<R> &g, const MP &m, const BGV &bg, const R *__restrict__ hmag,
                                        const R *__restrict__ hGas, int c, int f, const R *PL, const R *PR,
                                        const R *MagL, const R *MagR, R *out) {
    const size_t nc = g.nc, nf = g.nf;
    const int st = cstride(g, DN);
    R T[9];
#pragma unroll
    for (int q = 0; q < 9; q++) T[q] = g.Tf[nf * (q + 9 * DN) + f];
    const R faB = g.face[nf * DN + f];
    const R Bn = MHD ? hmag[nf * DN + f] / faB : R(0.0);
    R B0f[3] = {0, 0, 0}, B0n = R(0.0);
    const bool bgon = MHD && BG;
    if (bgon) {   // stress.F90:302-305
#pragma unroll
        for (int x = 0; x < 3; x++) B0f[x] = R(bg.fcB0[nc * (x + 3 * DN) + c]);
        B0n = R(bg.bFlux0[nc * DN + c]) / faB;
    }
    const R Ca = R(m.Ca);
    R VaD = R(0.0);
    if (MHD) {
        R mF[3] = {0, 0, 0}, bbD = R(0.0);
        magkin<BORIS>(m, R(-1.0), PL, MagL, Bn, T, bgon, B0f, B0n, mF, bbD, VaD);
        magkin<BORIS>(m, R(1.0), PR, MagR, Bn, T, bgon, B0f, B0n, mF, bbD, VaD);
        if (BORIS) {   // stress.F90:335-351
#pragma unroll
            for (int d = 0; d < 3; d++) {
                R deeV = PR[1 + d] - PL[1 + d];
                R deeP = bbD * deeV / sq(Ca);
                mF[d] = mF[d] - R(m.cHogM) * VaD * deeP;
            }
        }
#pragma unroll
        for (int d = 0; d < 3; d++) out[5 + d] = faB * mF[d];
    }
    R gF[5] = {0, 0, 0, 0, 0}, dW[5] = {0, 0, 0, 0, 0};
    gaskin(m, R(-1.0), PL, T, gF, dW);
    gaskin(m, R(1.0), PR, T, gF, dW);
#pragma unroll
    for (int v = 0; v < 5; v++) gF[v] = gF[v] - R(1.0) * R(m.cHogH) * VaD * dW[v];
    if (BORIS) {   // stress.F90:399-417
        R vL = rroot(sq(PL[1]) + sq(PL[2]) + sq(PL[3]));
        R vR = rroot(sq(PR[1]) + sq(PR[2]) + sq(PR[3]));
        R Vn = rroot(sq(vL) + sq(vR)) / R((double)sqrtf(2.0f));
        if (Vn >= R(1.5) * Ca) {
#pragma unroll
            for (int v = 0; v < 5; v++) {
                R dCon = hGas[nc * v + c] - hGas[nc * v + c - st];
                gF[v] = gF[v] - Ca * dCon;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++) out[v] = faB * gF[v];
}

template <int DN, int RM, int BORIS, int MHD, int BG, typename R>
__device__ __forceinline__ void face_flux(const GVT<R> &g, const MP &m, const BGV &bg,
                                          const R *__restrict__ Wn, const R *__restrict__ nB,
                                          const R *__restrict__ Wh, const R *__restrict__ hB,
                                          const R *__restrict__ hmag, const R *__restrict__ hGas,
                                          int c, int f, R *out) {
    const size_t nc = g.nc;
    const int st = cstride(g, DN);
    const R pdmb = R(m.pdmb);
    R VolB[8];
#pragma unroll
    for (int n = 0; n < 8; n++) VolB[n] = g.vol[c + (n - 4) * st];
    const R rVi = R(1) / c8(VolB);
    R PL[5], PR[5];
    R MagL[3] = {0, 0, 0}, MagR[3] = {0, 0, 0};
    if (MHD) {
        const R *limB = m.doLFMLim ? nB : hB;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            R Q[8], nQ[4];
#pragma unroll
            for (int n = 0; n < 8; n++) Q[n] = hB[nc * d + c + (n - 4) * st];
#pragma unroll
            for (int n = 0; n < 4; n++) nQ[n] = limB[nc * d + c + (n - 2) * st];
            get_lrs<RM>(pdmb, VolB, nQ, Q, rVi, MagL[d], MagR[d]);
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {   // BlockStateLRs on the prepass primitives
        R Q[8], nQ[4];
#pragma unroll
        for (int n = 0; n < 8; n++) Q[n] = Wh[nc * v + c + (n - 4) * st];
#pragma unroll
        for (int n = 0; n < 4; n++) nQ[n] = Wn[nc * v + c + (n - 2) * st];
        get_lrs<RM>(pdmb, VolB, nQ, Q, rVi, PL[v], PR[v]);
    }
    lr_flux<DN, BORIS, MHD, BG>(g, m, bg, hmag, hGas, c, f, PL, PR, MagL, MagR, out);
}

// The same for the two interfaces at cells c and c + 1 (adjacent along i),
// for a j or k stencil: every stencil point is one 2-wide load.  B's results
// are formed whether or not the caller keeps them (c + 1 is always a cell of
// the array: at most the first ghost past the active range).
#ifdef GAM_EMU
template <typename R> __device__ __forceinline__ void ld2(const R *p, R &a, R &b) { a = p[0]; b = p[1]; }
#else
template <typename R> struct P2 { typedef R T __attribute__((ext_vector_type(2), aligned(sizeof(R)))); };
template <typename R> __device__ __forceinline__ void ld2(const R *p, R &a, R &b) {
    const typename P2<R>::T v = *(const typename P2<R>::T *)p;
    a = v.x;
    b = v.y;
}
#endif
template <int DN, int RM, int BORIS, int MHD, int BG, typename R>
__device__ __forceinline__ void face_flux2(const GVT<R> &g, const MP &m, const BGV &bg,
                                           const R *__restrict__ Wn, const R *__restrict__ nB,
                                           const R *__restrict__ Wh, const R *__restrict__ hB,
                                           const R *__restrict__ hmag, const R *__restrict__ hGas,
                                           int c, int f, R *outA, R *outB) {
    const size_t nc = g.nc;
    const int st = cstride(g, DN);
    const R pdmb = R(m.pdmb);
    R VA[8], VB[8];
#pragma unroll
    for (int n = 0; n < 8; n++) ld2(g.vol + c + (n - 4) * st, VA[n], VB[n]);
    const R rVA = R(1) / c8(VA), rVB = R(1) / c8(VB);
    R PLA[5], PRA[5], PLB[5], PRB[5];
    R MLA[3] = {0, 0, 0}, MRA[3] = {0, 0, 0}, MLB[3] = {0, 0, 0}, MRB[3] = {0, 0, 0};
    if (MHD) {
        const R *limB = m.doLFMLim ? nB : hB;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            R QA[8], QB[8], nA[4], nBq[4];
#pragma unroll
            for (int n = 0; n < 8; n++) ld2(hB + nc * d + c + (n - 4) * st, QA[n], QB[n]);
#pragma unroll
            for (int n = 0; n < 4; n++) ld2(limB + nc * d + c + (n - 2) * st, nA[n], nBq[n]);
            get_lrs<RM>(pdmb, VA, nA, QA, rVA, MLA[d], MRA[d]);
            get_lrs<RM>(pdmb, VB, nBq, QB, rVB, MLB[d], MRB[d]);
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++) {
        R QA[8], QB[8], nA[4], nBq[4];
#pragma unroll
        for (int n = 0; n < 8; n++) ld2(Wh + nc * v + c + (n - 4) * st, QA[n], QB[n]);
#pragma unroll
        for (int n = 0; n < 4; n++) ld2(Wn + nc * v + c + (n - 2) * st, nA[n], nBq[n]);
        get_lrs<RM>(pdmb, VA, nA, QA, rVA, PLA[v], PRA[v]);
        get_lrs<RM>(pdmb, VB, nBq, QB, rVB, PLB[v], PRB[v]);
    }
    lr_flux<DN, BORIS, MHD, BG>(g, m, bg, hmag, hGas, c, f, PLA, PRA, MLA, MRA, outA);
    lr_flux<DN, BORIS, MHD, BG>(g, m, bg, hmag, hGas, c + 1, f + 1, PLB, PRB, MLB, MRB, outB);
}

// Interface fluxes in direction DN for the interface box
// DN=I: [is,ie+1]x[js,je]x[ks,ke]  (k range a single slice in 2.5D)
// Output slabs Flx[(8*DN + v)*nc + c], v = 0..4 gas, 5..7 magnetic.  CORR = 1:
// with the learned correction (corr, cs) added to the gas fluxes (corr_add); a
// separate instantiation, so the plain kernels keep their registers.
template <int DN, int RM, int BORIS, int MHD, int BG, typename R, int CORR = 0>
__device__ __forceinline__ void b_flux(const unsigned vb, GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                              const R *__restrict__ Wn, const R *__restrict__ nB,
                                              const R *__restrict__ Wh, const R *__restrict__ hB,
                                              const R *__restrict__ hmag, const R *__restrict__ hGas,
                                              R *__restrict__ Flx, const float *__restrict__ corr, double cs) {
    const int nx = g.Nip + (DN == 0), ny = g.Njp + (DN == 1), nz = (m.do25D ? 1 : g.Nkp) + (DN == 2);
    int a, b, cz;
    if (!tile_pos(vb, t, nx, ny, nz, a, b, cz) || clk->halt) return;
    const int i = a + g.nG, j = b + g.nG, k = cz + g.nG;
    const size_t nc = g.nc;
    const int c = ci(g, i, j, k);
    R o[8];
    face_flux<DN, RM, BORIS, MHD, BG>(g, m, bg, Wn, nB, Wh, hB, hmag, hGas, c, fi(g, i, j, k), o);
    if (CORR && !(DN == 2 && m.do25D)) {   // CORR instantiations run with a correction set only
        const double A = (double)g.face[(size_t)g.nf * DN + fi(g, i, j, k)];
        const float *oc = corr + pimg(g, i, j, k) * 8;
#pragma unroll
        for (int v = 0; v < 5; v++) o[v] = corr_add(o[v], cs, A, oc[v]);
    }
#pragma unroll
    for (int v = 0; v < (MHD ? 8 : 5); v++) Flx[nc * (8 * DN + v) + c] = o[v];
}
template <int DN, int RM, int BORIS, int MHD, int BG, typename R, int CORR = 0>
__global__ void __launch_bounds__(256) k_flux(GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                              const R *__restrict__ Wn, const R *__restrict__ nB,
                                              const R *__restrict__ Wh, const R *__restrict__ hB,
                                              const R *__restrict__ hmag, const R *__restrict__ hGas,
                                              R *__restrict__ Flx, const float *__restrict__ corr, double cs) {
    b_flux<DN, RM, BORIS, MHD, BG, R, CORR>(blockIdx.x, g, t, m, bg, clk, Wn, nB, Wh, hB, hmag, hGas, Flx, corr, cs);
}

// j or k interfaces two at a time along i (face_flux2): thread (a, b, c) of
// the box ceil(nx/2) x ny x nz owns interfaces i = 2a and 2a + 1.
template <int DN, int RM, int BORIS, int MHD, int BG, typename R, int CORR = 0>
__device__ __forceinline__ void b_flux2(const unsigned vb, GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                        const R *__restrict__ Wn, const R *__restrict__ nB,
                                        const R *__restrict__ Wh, const R *__restrict__ hB,
                                        const R *__restrict__ hmag, const R *__restrict__ hGas,
                                        R *__restrict__ Flx, const float *__restrict__ corr, double cs) {
    static_assert(DN != 0, "pairs along i serve the j and k interfaces");
    const int nx = g.Nip, ny = g.Njp + (DN == 1), nz = (m.do25D ? 1 : g.Nkp) + (DN == 2);
    int a, b, cz;
    if (!tile_pos(vb, t, (nx + 1) / 2, ny, nz, a, b, cz) || clk->halt) return;
    const int i = 2 * a + g.nG, j = b + g.nG, k = cz + g.nG;
    const size_t nc = g.nc;
    const int c = ci(g, i, j, k);
    R oA[8], oB[8];
    face_flux2<DN, RM, BORIS, MHD, BG>(g, m, bg, Wn, nB, Wh, hB, hmag, hGas, c, fi(g, i, j, k), oA, oB);
    const bool two = 2 * a + 1 < nx;
    if (CORR && !(DN == 2 && m.do25D)) {   // CORR instantiations run with a correction set only
        const int f = fi(g, i, j, k);
        const double A = (double)g.face[(size_t)g.nf * DN + f];
        const float *oc = corr + pimg(g, i, j, k) * 8;
#pragma unroll
        for (int v = 0; v < 5; v++) oA[v] = corr_add(oA[v], cs, A, oc[v]);
        if (two) {
            const double AB = (double)g.face[(size_t)g.nf * DN + f + 1];
            const float *ob = corr + pimg(g, i + 1, j, k) * 8;
#pragma unroll
            for (int v = 0; v < 5; v++) oB[v] = corr_add(oB[v], cs, AB, ob[v]);
        }
    }
#pragma unroll
    for (int v = 0; v < (MHD ? 8 : 5); v++) {
        Flx[nc * (8 * DN + v) + c] = oA[v];
        if (two) Flx[nc * (8 * DN + v) + c + 1] = oB[v];
    }
}
template <int DN, int RM, int BORIS, int MHD, int BG, typename R, int CORR = 0>
__global__ void __launch_bounds__(256) k_flux2(GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                               const R *__restrict__ Wn, const R *__restrict__ nB,
                                               const R *__restrict__ Wh, const R *__restrict__ hB,
                                               const R *__restrict__ hmag, const R *__restrict__ hGas,
                                               R *__restrict__ Flx, const float *__restrict__ corr, double cs) {
    b_flux2<DN, RM, BORIS, MHD, BG, R, CORR>(blockIdx.x, g, t, m, bg, clk, Wn, nB, Wh, hB, hmag, hGas, Flx, corr, cs);
}
// The correction of the gas fluxes in direction DN on their own (corr_add on the
// stored fluxes of the flux box), for the one flux kernel whose correcting
// instantiation the compiler cannot fit in its plain registers (the FP64 j pairs).
template <int DN, typename R>
__global__ void __launch_bounds__(256) k_corr_flux(GVT<R> g, MP m, const DClock *__restrict__ clk, R *__restrict__ Flx,
                                                   const float *__restrict__ corr, double cs) {
    const int nx = g.Nip + (DN == 0), ny = g.Njp + (DN == 1), nz = (m.do25D ? 1 : g.Nkp) + (DN == 2);
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= (long)nx * ny * nz || clk->halt || (DN == 2 && m.do25D)) return;
    const int i = (int)(q % nx) + g.nG, j = (int)((q / nx) % ny) + g.nG, k = (int)(q / ((long)nx * ny)) + g.nG;
    const size_t nc = g.nc;
    const int c = ci(g, i, j, k);
    const double A = (double)g.face[(size_t)g.nf * DN + fi(g, i, j, k)];
    const float *oc = corr + pimg(g, i, j, k) * 8;
#pragma unroll
    for (int v = 0; v < 5; v++) Flx[nc * (8 * DN + v) + c] = corr_add(Flx[nc * (8 * DN + v) + c], cs, A, oc[v]);
}

// ----------------------------------------------------------------- CT update
// fields.F90:18-51 E2Flux into the new buffer, in FP64: faces in [is,ie+1]^3
// box get the curl update, all others are copied.  E outside the active edge
// ranges is 0.
template <typename R>
__device__ __forceinline__ double Eval(const GVT<R> &g, const R *E, int d, int i, int j, int k) {
    const int a0 = g.nG;
    int ai = i - a0, aj = j - a0, ak = k - a0;
    int mi = g.Nip - (d == 0), mj = g.Njp - (d == 1), mk = g.Nkp - (d == 2);
    if (ai < 0 || aj < 0 || ak < 0 || ai > mi || aj > mj || ak > mk) return 0.0;
    return (double)E[(size_t)g.nc * d + ci(g, i, j, k)];
}

template <typename R>
__device__ __forceinline__ void b_ct(const unsigned vb, GVT<R> g, const DClock *__restrict__ clk, const double *__restrict__ magN,
                                            const R *__restrict__ E, double *__restrict__ magO) {
    int idx = vb * blockDim.x + threadIdx.x;
    if (idx >= g.nf || clk->halt) return;
    const size_t nf = g.nf;
    int i, j, k;
    decode_f(g, idx, i, j, k);
    const int a0 = g.nG;
    double m0 = magN[idx], m1 = magN[nf + idx], m2 = magN[2 * nf + idx];
    if (i >= a0 && i <= a0 + g.Nip && j >= a0 && j <= a0 + g.Njp && k >= a0 && k <= a0 + g.Nkp) {
        const double dt = clk->dt;
        double Ei = Eval(g, E, 0, i, j, k), Ej = Eval(g, E, 1, i, j, k), Ek = Eval(g, E, 2, i, j, k);
        double dfi = Eval(g, E, 2, i, j + 1, k) - Ek - Eval(g, E, 1, i, j, k + 1) + Ej;
        double dfj = Eval(g, E, 0, i, j, k + 1) - Ei - Eval(g, E, 2, i + 1, j, k) + Ek;
        double dfk = Eval(g, E, 1, i + 1, j, k) - Ej - Eval(g, E, 0, i, j + 1, k) + Ei;
        m0 = m0 - dt * dfi;
        m1 = m1 - dt * dfj;
        m2 = m2 - dt * dfk;
    }
    magO[idx] = m0; magO[nf + idx] = m1; magO[2 * nf + idx] = m2;
}
template <typename R>
__global__ void __launch_bounds__(256) k_ct(GVT<R> g, const DClock *__restrict__ clk, const double *__restrict__ magN,
                                            const R *__restrict__ E, double *__restrict__ magO) {
    b_ct<R>(blockIdx.x, g, clk, magN, E, magO);
}

// ----------------------------------------------------------------- update (FP64)
// mhdgroup.F90:300-330 CellMaxwell
// in the compute type T of the cell update; the total energy to pressure
// recovery with its floors is FP64 whatever T is
template <typename T>
__device__ __forceinline__ void cell_maxwell(const MP &m, T *U, const T *dGM, T dt) {
    const T D = rmax(U[0], T(m.dFloor));
    const double Dd = fmax((double)U[0], m.dFloor);
    const double KinE0 = 0.5 * (sq((double)U[1]) + sq((double)U[2]) + sq((double)U[3])) / Dd;
    const double P = fmax((m.gamma - 1) * ((double)U[4] - KinE0), m.pFloor);
    T Mx = U[1], My = U[2], Mz = U[3];
    Mx = Mx + dt * dGM[0]; My = My + dt * dGM[1]; Mz = Mz + dt * dGM[2];
    const T KinE = T(0.5) * (sq(Mx) + sq(My) + sq(Mz)) / D;
    U[0] = D; U[1] = Mx; U[2] = My; U[3] = Mz;
    U[4] = T(KinE + P / (m.gamma - 1));
}

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void normvec(const double *a, double *o) {   // math.F90:332-343
    double n = sqrt(dot3(a, a));
    if (n <= GAM_TINY) { o[0] = o[1] = o[2] = 0.0; }
    else { o[0] = a[0] / n; o[1] = a[1] / n; o[2] = a[2] / n; }
}

template <typename T> __device__ __forceinline__ T dot3t(const T *a, const T *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <typename T> __device__ __forceinline__ void normvec_t(const T *a, T *o) {   // math.F90:332-343
    T n = rroot(dot3t(a, a));
    if (n <= T(GAM_TINY)) { o[0] = o[1] = o[2] = T(0.0); }
    else { o[0] = a[0] / n; o[1] = a[1] / n; o[2] = a[2] / n; }
}

// mhdgroup.F90:337-389 CellBoris (bulk fluid), in the compute type T of the
// cell update; the total energy to pressure recovery with its floors is FP64
template <typename T>
__device__ __forceinline__ void cell_boris(const MP &m, T *U2, const T *U0, const T *B2, const T *B0,
                                           const T *dGH, const T *dGM, T Ca, T dt) {
    double U2d[5], W2[5];
#pragma unroll
    for (int v = 0; v < 5; v++) U2d[v] = U2[v];
    cell_c2p(m, U2d, W2);
    const T D2 = T(fmax(W2[0], m.dFloor)), P2 = T(fmax(W2[4], m.pFloor));
    T D0 = rmax(U0[0], T(m.dFloor));
    T V0[3] = {U0[1] / D0, U0[2] / D0, U0[3] / D0};
    T B1[3] = {T(0.5) * (B2[0] + B0[0]), T(0.5) * (B2[1] + B0[1]), T(0.5) * (B2[2] + B0[2])};
    T alpha = dot3t(B1, B1) / (D2 * Ca * Ca);
    T dRho = dt * dGH[0];
    T dPGas[3] = {dt * dGH[1], dt * dGH[2], dt * dGH[3]};
    T dPMag[3] = {dt * dGM[0], dt * dGM[1], dt * dGM[2]};
    T pCon = T(1.0) / (T(1.0) + alpha);
    T bh1[3];
    normvec_t(B1, bh1);
    T dMom[3];
    if (alpha > T(GAM_TINY)) {
        T dp = dot3t(dPGas, bh1);
        T sum[3] = {dPGas[0] + dPMag[0], dPGas[1] + dPMag[1], dPGas[2] + dPMag[2]};
        T ds = dot3t(sum, bh1), dv = dot3t(V0, bh1);
        T s = alpha * pCon * dRho;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            T a = dp * bh1[d];
            T b = sum[d] - ds * bh1[d];
            T cc = V0[d] - dv * bh1[d];
            dMom[d] = a + pCon * b + s * cc;
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) dMom[d] = dPGas[d] + dPMag[d];
    }
    T W[5];
    W[0] = D2;
#pragma unroll
    for (int d = 0; d < 3; d++) W[1 + d] = (D0 * V0[d] + dMom[d]) / D2;
    W[4] = P2;
    cell_p2c(m, W, U2);
}

// gridutils.F90:49-106 CellDT (single fluid, no resistivity), FP64
template <int BORIS, typename R>
__device__ __forceinline__ double cell_dt(const GVT<R> &g, const MP &m, const double *U, const double *B, int c) {
    double rho = U[0];
    double Vx = U[1] / rho, Vy = U[2] / rho, Vz = U[3] / rho;
    double MagV = sqrt(sq(Vx) + sq(Vy) + sq(Vz));
    double ke = 0.5 * rho * sq(MagV);
    double e = U[4] - ke;
    double P = (m.gamma - 1) * e;
    double Cs = sqrt(m.gamma * P / rho);
    double Valf = 0.0;
    if (m.doMHD) {
        double MagB = sqrt(sq(B[0]) + sq(B[1]) + sq(B[2]));
        Valf = MagB / sqrt(rho);
        if (BORIS) Valf = m.Ca * Valf / sqrt(m.Ca * m.Ca + Valf * Valf);
    }
    double vCFL = MagV + sqrt(sq(Cs) + sq(Valf)) + 0.0;
    double dl = fmin(fmin((double)g.di[c], (double)g.dj[c]), (double)g.dk[c]);
    return m.CFL * dl / vCFL;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// block min over 256 threads -> partial[vb] (vb: the block's index)
__device__ __forceinline__ void block_min_store(double v, double *partial, const unsigned vb) {
#ifdef GAM_EMU
    // the CPU emulation runs the threads of a block in order 0..blockDim-1
    partial[vb] = threadIdx.x == 0 ? v : fmin(partial[vb], v);
    return;
#endif
    __shared__ double red[4];
    v = wave_min(v);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) r = fmin(r, red[q]);
        partial[vb] = r;
    }
}

// block sum over 256 threads -> partial[vb]
__device__ __forceinline__ void block_sum_store(double v, double *partial, const unsigned vb) {
#ifdef GAM_EMU
    partial[vb] = threadIdx.x == 0 ? v : partial[vb] + v;
    return;
#endif
    __shared__ double reds[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) reds[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = reds[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) r += reds[q];
        partial[vb] = r;
    }
}

template <typename R> __device__ __forceinline__ bool in_dt_box(const GVT<R> &g, int i, int j, int k) {
    // SetDomain init.F90:519-538: active +-1 layer (single block owns all BCs)
    const int a0 = g.nG - 1;
    return i >= a0 && i <= g.nG + g.Nip && j >= a0 && j <= g.nG + g.Njp && k >= a0 && k <= g.nG + g.Nkp;
}

// applyUpdate of one active cell (mhdgroup.F90:79-199) from its deltas: the
// Reynolds update of U (State n in, new state out), then Maxwell or Boris with
// Bn, the cell field of the new face fluxes (total fields with a background).
// Reynolds then Maxwell or Boris (mhdgroup.F90:153-196) in the compute type T
// of the cell update (the bulk type); the fields' totals with B0 are formed in
// FP64 and rounded to T
template <int BORIS, int MHD, typename R, typename T>
__device__ __forceinline__ void cell_apply(const MP &m, const BGV &bg, size_t nc, int idx, T dt, const R *__restrict__ nB,
                                           T *U, const double *Bn, const T *dGH, const T *dGM) {
    T U0[5];
#pragma unroll
    for (int v = 0; v < 5; v++) { U0[v] = U[v]; U[v] = U[v] + dt * dGH[v]; }   // CellReynolds, Model%dt
    if (MHD) {
        if (BORIS) {
            double oBv[3] = {(double)nB[idx], (double)nB[nc + idx], (double)nB[2 * nc + idx]};
            double Bt[3] = {Bn[0], Bn[1], Bn[2]};
            if (bg.on) {   // mhdgroup.F90:180-183 total fields
#pragma unroll
                for (int x = 0; x < 3; x++) { Bt[x] = Bt[x] + bg.B0[nc * x + idx]; oBv[x] = oBv[x] + bg.B0[nc * x + idx]; }
            }
            const T BtT[3] = {T(Bt[0]), T(Bt[1]), T(Bt[2])}, oBT[3] = {T(oBv[0]), T(oBv[1]), T(oBv[2])};
            cell_boris(m, U, U0, BtT, oBT, dGH, dGM, T(m.Ca), dt);
        } else {
            cell_maxwell(m, U, dGM, dt);
        }
    }
}

// CellDT of the stored new state (rounded to the bulk type first) with the
// total field (gridutils.F90:82-86)
template <int BORIS, typename R>
__device__ __forceinline__ double cell_dt_stored(const GVT<R> &g, const MP &m, const BGV &bg, int idx, const double *Us,
                                                 const double *Bs) {
    double U[5], Bn[3];
#pragma unroll
    for (int v = 0; v < 5; v++) U[v] = (double)R(Us[v]);
#pragma unroll
    for (int x = 0; x < 3; x++) {
        Bn[x] = (double)R(Bs[x]);
        if (bg.on) Bn[x] = Bn[x] + bg.B0[(size_t)g.nc * x + idx];
    }
    return cell_dt<BORIS>(g, m, U, Bn, idx);
}

// Flux2Deltas + applyUpdate (Reynolds, Bxyz of the new fluxes, Maxwell/Boris)
// + CellDT min.  Flux2Deltas, Reynolds and Maxwell/Boris compute in the bulk
// type (FP32 bulk: FP32); the energy to pressure recovery with its floors,
// bFlux2Fld of the FP64 magFlux and CellDT stay FP64.  Writes the new Gas/Bxyz
// into the O buffers; ghosts copy Gas n.
template <int BORIS, int MHD, typename R>
__device__ __forceinline__ void b_update(const unsigned vb, GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                                const R *__restrict__ nGas, const R *__restrict__ nB,
                                                const R *__restrict__ Flx, const double *__restrict__ magNew,
                                                R *__restrict__ oGas, R *__restrict__ oB,
                                                double *__restrict__ partial, int skip_nb) {
    if (clk->halt) return;
    double dtl = GAM_HUGE;
    int i, j, k;
    if (tile_pos(vb, t, g.Ni, g.Nj, g.Nk, i, j, k)) {
        const int idx = ci(g, i, j, k);
        const size_t nc = g.nc;
        R U[5];   // the update computes in the bulk type (FP64 step: double)
        double Bn[3] = {0, 0, 0};
#pragma unroll
        for (int v = 0; v < 5; v++) U[v] = nGas[nc * v + idx];
        cell_bxyz(g, magNew, i, j, k, Bn);
        const int a0 = g.nG;
        bool active = i >= a0 && i < a0 + g.Nip && j >= a0 && j < a0 + g.Njp && k >= a0 && k < a0 + g.Nkp;
        if (active) {
            const R dt = R(clk->dt);
            const int ci1 = idx + 1, cj1 = idx + g.Ni, ck1 = idx + g.Ni * g.Nj;
            const R dV = g.vol[idx];   // FP32 bulk: the FP32 copy of dvol (the same rounding, half the bytes)
            auto F = [&](int slab, int c) { return Flx[nc * slab + c]; };
            R dGH[5], dGM[3] = {0, 0, 0};
#pragma unroll
            for (int v = 0; v < 5; v++) {
                dGH[v] = (F(0 + v, idx) - F(0 + v, ci1) + F(8 + v, idx) - F(8 + v, cj1)
                        + F(16 + v, idx) - F(16 + v, ck1)) / dV;
            }
            if (MHD) {
#pragma unroll
                for (int v = 0; v < 3; v++) {
                    dGM[v] = (F(5 + v, idx) - F(5 + v, ci1) + F(13 + v, idx) - F(13 + v, cj1)
                            + F(21 + v, idx) - F(21 + v, ck1)) / dV;
                    if (bg.on) dGM[v] = dGM[v] + R(bg.dpB0[nc * v + idx]);   // stress.F90:185-188
                }
            }
            cell_apply<BORIS, MHD>(m, bg, nc, idx, dt, nB, U, Bn, dGH, dGM);
        }
#pragma unroll
        for (int v = 0; v < 5; v++) st_nt(&oGas[nc * v + idx], U[v]);
#pragma unroll
        for (int x = 0; x < 3; x++) st_nt(&oB[nc * x + idx], R(Bn[x]));
        // skip_nb > 0: the ring rows' CFL bound is taken after the ring average (k_celldt_rows)
        const bool band = skip_nb > 0 && (j < g.nG - 1 + skip_nb || j > g.nG + g.Njp - skip_nb);
        if (in_dt_box(g, i, j, k) && !band) {   // CFL bound of the stored state
            const double Ud[5] = {U[0], U[1], U[2], U[3], U[4]};
            dtl = cell_dt_stored<BORIS>(g, m, bg, idx, Ud, Bn);
        }
    }
    block_min_store(dtl, partial, vb);
}
template <int BORIS, int MHD, typename R>
__global__ void __launch_bounds__(256) k_update(GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                                const R *__restrict__ nGas, const R *__restrict__ nB,
                                                const R *__restrict__ Flx, const double *__restrict__ magNew,
                                                R *__restrict__ oGas, R *__restrict__ oB,
                                                double *__restrict__ partial, int skip_nb) {
    b_update<BORIS, MHD, R>(blockIdx.x, g, t, m, bg, clk, nGas, nB, Flx, magNew, oGas, oB, partial, skip_nb);
}

// k_update walking UPDATE_KC cells up a k column per thread, reading each K
// face's fluxes, magFlux and centre once (the FP64 step's update: less HBM
// traffic there; in FP32 bulk it measured slower and k_update is used).
template <int BORIS, int MHD, typename R>
__device__ __forceinline__ void b_update_col(const unsigned vb, GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                                const R *__restrict__ nGas, const R *__restrict__ nB,
                                                const R *__restrict__ Flx, const double *__restrict__ magNew,
                                                R *__restrict__ oGas, R *__restrict__ oB,
                                                double *__restrict__ partial, int skip_nb) {
    if (clk->halt) return;
    double dtl = GAM_HUGE;
    constexpr int KC = UPDATE_KC;
    int i, j, kb;
    if (tile_pos(vb, t, g.Ni, g.Nj, cdiv(g.Nk, KC), i, j, kb)) {
        const size_t nc = g.nc;
        const int a0 = g.nG;
        const bool act_ij = i >= a0 && i < a0 + g.Nip && j >= a0 && j < a0 + g.Njp;
        auto F = [&](int slab, int c) { return (double)Flx[nc * slab + c]; };
        KCarry kc;
        double cF[8];          // K fluxes of the lower K face, carried from the cell below
        bool haveF = false;
#pragma unroll 1
        for (int kk = 0; kk < KC; kk++) {
            const int k = kb * KC + kk;
            if (k >= g.Nk) break;
            const int idx = ci(g, i, j, k);
            double U[5], Bn[3] = {0, 0, 0};
#pragma unroll
            for (int v = 0; v < 5; v++) U[v] = (double)nGas[nc * v + idx];
            cell_bxyz_kc(g, magNew, i, j, k, Bn, kc, kk > 0);
            const bool active = act_ij && k >= a0 && k < a0 + g.Nkp;
            if (active) {
                const double dt = clk->dt;
                const int ci1 = idx + 1, cj1 = idx + g.Ni, ck1 = idx + g.Ni * g.Nj;
                const double dV = g.dvol[idx];
                if (!haveF) {
#pragma unroll
                    for (int v = 0; v < 8; v++) cF[v] = (v < 5 || MHD) ? F(16 + v, idx) : 0.0;
                }
                double dGH[5], dGM[3] = {0, 0, 0};
#pragma unroll
                for (int v = 0; v < 5; v++) {
                    const double fk1 = F(16 + v, ck1);
                    dGH[v] = (F(0 + v, idx) - F(0 + v, ci1) + F(8 + v, idx) - F(8 + v, cj1)
                            + cF[v] - fk1) / dV;
                    cF[v] = fk1;
                }
                if (MHD) {
#pragma unroll
                    for (int v = 0; v < 3; v++) {
                        const double fk1 = F(21 + v, ck1);
                        dGM[v] = (F(5 + v, idx) - F(5 + v, ci1) + F(13 + v, idx) - F(13 + v, cj1)
                                + cF[5 + v] - fk1) / dV;
                        cF[5 + v] = fk1;
                        if (bg.on) dGM[v] = dGM[v] + bg.dpB0[nc * v + idx];   // stress.F90:185-188
                    }
                }
                haveF = true;
                cell_apply<BORIS, MHD>(m, bg, nc, idx, dt, nB, U, Bn, dGH, dGM);
            } else {
                haveF = false;
            }
#pragma unroll
            for (int v = 0; v < 5; v++) st_nt(&oGas[nc * v + idx], R(U[v]));
#pragma unroll
            for (int x = 0; x < 3; x++) st_nt(&oB[nc * x + idx], R(Bn[x]));
            // skip_nb > 0: the ring rows' CFL bound is taken after the ring average (k_celldt_rows)
            const bool band = skip_nb > 0 && (j < g.nG - 1 + skip_nb || j > g.nG + g.Njp - skip_nb);
            if (in_dt_box(g, i, j, k) && !band) dtl = fmin(dtl, cell_dt_stored<BORIS>(g, m, bg, idx, U, Bn));   // CFL bound of the stored state
        }
    }
    block_min_store(dtl, partial, vb);
}
template <int BORIS, int MHD, typename R>
__global__ void __launch_bounds__(256) k_update_col(GVT<R> g, Tile t, MP m, BGV bg, const DClock *__restrict__ clk,
                                                const R *__restrict__ nGas, const R *__restrict__ nB,
                                                const R *__restrict__ Flx, const double *__restrict__ magNew,
                                                R *__restrict__ oGas, R *__restrict__ oB,
                                                double *__restrict__ partial, int skip_nb) {
    b_update_col<BORIS, MHD, R>(blockIdx.x, g, t, m, bg, clk, nGas, nB, Flx, magNew, oGas, oB, partial, skip_nb);
}

// Stand-alone CellDT min over the DT box (CalcDT without the fused update)
template <int BORIS, typename R>
__global__ void __launch_bounds__(256) k_celldt(GVT<R> g, MP m, BGV bg, const DClock *__restrict__ clk, const R *__restrict__ Gas,
                                                const R *__restrict__ B, double *__restrict__ partial) {
    if (clk->halt) return;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    double dtl = GAM_HUGE;
    if (idx < g.nc) {
        int i, j, k;
        decode_c(g, idx, i, j, k);
        if (in_dt_box(g, i, j, k)) {
            const size_t nc = g.nc;
            double U[5], Bv[3];
#pragma unroll
            for (int v = 0; v < 5; v++) U[v] = (double)Gas[nc * v + idx];
#pragma unroll
            for (int x = 0; x < 3; x++) Bv[x] = (double)B[nc * x + idx];
            if (bg.on) {   // gridutils.F90:82-86
#pragma unroll
                for (int x = 0; x < 3; x++) Bv[x] = Bv[x] + bg.B0[nc * x + idx];
            }
            dtl = cell_dt<BORIS>(g, m, U, Bv, idx);
        }
    }
    block_min_store(dtl, partial, blockIdx.x);
}

// UpdateStateData time update (gamapp.F90:112-114, mhdgroup.F90:104-105,255)
// when advance != 0, then CalcDT's checks and clamps (step.F90:103-166).
__device__ __forceinline__ void b_dtfinal(const unsigned vb, DClock *clk, const double *__restrict__ partial, int np, int advance, int calc) {
    if (clk->halt) return;
    if (threadIdx.x == 0) clk->cpr_pending = 0;
#ifdef GAM_EMU
    if (threadIdx.x != 0) return;
    double dtMin = GAM_HUGE;
    for (int q = 0; q < np; q++) dtMin = fmin(dtMin, partial[q]);
#else
    __shared__ double red[16];
    double v = GAM_HUGE;
    for (int q = threadIdx.x; q < np; q += blockDim.x) v = fmin(v, partial[q]);
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double dtMin = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) dtMin = fmin(dtMin, red[q]);
#endif
    if (advance) {
        clk->otime = clk->stime;
        clk->stime = clk->stime + clk->dt;
        clk->ts = clk->ts + 1;
        clk->t = clk->t + clk->dt;
    }
    if (!calc) return;
    if (clk->fixedTimestep) { clk->status = 0; return; }
    double dtOld = clk->dt;
    int disaster = (dtMin == GAM_HUGE), bad = 0;
    double dt = dtMin;
    int status = 0;
    if (clk->ts > 0) {
        if (dtOld / dtMin >= 10) bad = 1;
        if (clk->dt0 > GAM_TINY && dtMin / clk->dt0 <= clk->limDT0) bad = 1;
        if (clk->doCPR && clk->dt0 > GAM_TINY && dtMin / clk->dt0 <= clk->limCPR) bad = 1;
        if (dtMin <= GAM_TINY) disaster = 1;
        status = disaster ? 2 : (bad ? 1 : 0);
        if (clk->doCPR && bad && !disaster) {   // step.F90:140-142: StateCPR (k_cpr) and go on
            status = 0;
            clk->cpr_pending = 1;
            clk->ncpr = clk->ncpr + 1;
        }
    }
    if ((clk->t + dt) > clk->tFin) dt = fmax(clk->tFin - clk->t, GAM_TINY);
    if (clk->dt0 > GAM_TINY && dt > (10.0 * clk->dt0)) dt = 10.0 * clk->dt0;
    clk->dt = dt;
    clk->status = status;
    if (status) clk->halt = 1;   // the reference stops here (step.F90:138-154); the host decides
    // console-output dt0 capture (output.F90:57-59) at the configured cadence
    if (clk->dt0_after > 0 && clk->ts >= clk->dt0_after && clk->dt0 < GAM_TINY) clk->dt0 = clk->dt;
}
__global__ void k_dtfinal(DClock *clk, const double *__restrict__ partial, int np, int advance, int calc) {
    b_dtfinal(blockIdx.x, clk, partial, np, advance, calc);
}

// step.F90:279-302 StateCPR after a bad (not disastrous) CalcDT with doCPR:
// CellCPR (step.F90:305-430, single fluid) on every active cell whose CFL step
// is at most limCPR of dt0: the fastest of the flow and sound speeds is slowed
// by alpha = 0.9 (velocity) or alpha^2 (pressure); a floored cell loses its
// velocity; the Alfven branch is off in the reference.  FP64 on the stored state.
template <int BORIS, typename R>
__device__ __forceinline__ void b_cpr(const unsigned vb, GVT<R> g, MP m, BGV bg, const DClock *__restrict__ clk, R *__restrict__ Gas,
                                             const R *__restrict__ B) {
    if (!clk->cpr_pending || clk->halt) return;
    const int n = g.Nip * g.Njp * g.Nkp;
    const int q = vb * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int i = g.nG + q % g.Nip, r = q / g.Nip, j = g.nG + r % g.Njp, k = g.nG + r / g.Njp;
    const int c = ci(g, i, j, k);
    const size_t nc = g.nc;
    double U[5], Bt[3];
#pragma unroll
    for (int v = 0; v < 5; v++) U[v] = (double)Gas[nc * v + c];
#pragma unroll
    for (int x = 0; x < 3; x++) Bt[x] = (double)B[nc * x + c] + (bg.on ? bg.B0[nc * x + c] : 0.0);
    if (!(cell_dt<BORIS>(g, m, U, Bt, c) / clk->dt0 <= clk->limCPR)) return;
    const double alpha = 0.9;
    double pW[5];
    cell_c2p(m, U, pW);
    const double Vf = sqrt(sq(pW[1]) + sq(pW[2]) + sq(pW[3]));
    const double rhoF = fmax(U[0], m.dFloor);   // gamutils.F90:324-341 CellPress2Cs
    const double vx = U[1] / rhoF, vy = U[2] / rhoF, vz = U[3] / rhoF;
    const double KinE = 0.5 * rhoF * (sq(vx) + sq(vy) + sq(vz));
    const double P = fmax((m.gamma - 1) * (U[4] - KinE), m.pFloor);
    const double Cs = sqrt(m.gamma * P / rhoF);
    double Va = 0.0;
    if (m.doMHD) {
        Va = sqrt(sq(Bt[0]) + sq(Bt[1]) + sq(Bt[2])) / sqrt(U[0]);
        if (m.doBoris) Va = m.Ca * Va / sqrt(m.Ca * m.Ca + Va * Va);
    }
    const bool doVf = Vf >= Cs && Vf >= Va;
    const bool doCs = !doVf && Cs >= Va && Cs >= Vf;
    const bool floored = (pW[0] < m.dFloor + GAM_TINY) || (pW[4] < m.pFloor + GAM_TINY);
    if (doVf) { pW[1] = alpha * pW[1]; pW[2] = alpha * pW[2]; pW[3] = alpha * pW[3]; }
    if (doCs) pW[4] = alpha * alpha * pW[4];
    if (floored) pW[1] = pW[2] = pW[3] = 0.0;
    double Un[5];
    cell_p2c(m, pW, Un);
#pragma unroll
    for (int v = 0; v < 5; v++) Gas[nc * v + c] = R(Un[v]);
}
template <int BORIS, typename R>
__global__ void __launch_bounds__(256) k_cpr(GVT<R> g, MP m, BGV bg, const DClock *__restrict__ clk, R *__restrict__ Gas,
                                             const R *__restrict__ B) {
    b_cpr<BORIS, R>(blockIdx.x, g, m, bg, clk, Gas, B);
}

// ----------------------------------------------------------------- BCs
// bcs.F90:279-450 periodic BCs, applied in the order INI,OUTI,INJ,OUTJ,INK,OUTK.
// The sequential copies compose into one per-axis index map (every source index
// is an owned index along the axis being mapped), so one in-place pass over the
// ghost points reproduces them.  Cell-centred map:
//   i < is -> i + Nip ;  i > ie -> i - Nip
// Face-normal component along the axis:  i < is -> i + Nip ; i > ie+1 -> i - Nip
// (face ie+1 keeps its own CT-evolved value); both repeated while needed.  2.5D: every k except ks -> ks,
// for all components, k up to keg (bcs.F90:387-394,423-429).
// When the owned extent n is smaller than the ghost depth the reference's
// loops (n = 1..nG, in order) copy from ghosts they have just filled, so the
// map is applied until it lands on an owned index.
__device__ __forceinline__ int pmap_c(int i, int a0, int n) {
    while (i < a0) i += n;
    while (i >= a0 + n) i -= n;
    return i;
}
__device__ __forceinline__ int pmap_f(int i, int a0, int n) {   // owned faces a0..a0+n
    while (i < a0) i += n;
    while (i > a0 + n) i -= n;
    return i;
}

// Both maps are the identity on the owned cells a0..a0+n-1 of a mapped axis
// (2.5D k: on ks alone) and on every index of an axis not mapped, so only the
// points of the (Ni+1)(Nj+1)(Nk+1) face-point box outside that inner box can
// change anything: the pass runs over this shell.  Shell order: the k planes
// outside, then the j rows outside of the inner planes, then the i points
// outside of the inner rows (i fastest in each part).
struct BcShell {
    int lo[3], in[3], np[3];   // per axis: first inner point, inner count, point count
};
__host__ __device__ __forceinline__ BcShell bc_shell(int Ni, int Nj, int Nk, int Nip, int Njp, int Nkp, int a0, int mask,
                                                     int do25D) {
    BcShell s;
    const int N[3] = {Ni, Nj, Nk}, Np[3] = {Nip, Njp, Nkp};
    for (int x = 0; x < 3; x++) {
        s.np[x] = N[x] + 1;
        const bool m = (mask >> x) & 1;
        s.lo[x] = m ? a0 : 0;
        s.in[x] = m ? (x == 2 && do25D ? 1 : Np[x]) : N[x] + 1;
    }
    return s;
}
__host__ __device__ __forceinline__ long bc_shell_count(const BcShell &s) {
    return (long)s.np[0] * s.np[1] * s.np[2] - (long)s.in[0] * s.in[1] * s.in[2];
}

// mask: the periodic axes (bit 0 i, 1 j, 2 k) mapped by this pass; the
// others are left for their own BCs.
template <typename R>
__device__ __forceinline__ void b_bc(const unsigned vb, GVT<R> g, const DClock *__restrict__ clk, int do25D, int mask, R *__restrict__ Gas,
                                            double *__restrict__ mag, R *__restrict__ B) {
    const BcShell sh = bc_shell(g.Ni, g.Nj, g.Nk, g.Nip, g.Njp, g.Nkp, g.nG, mask, do25D);
    const long t = (long)vb * blockDim.x + threadIdx.x;
    const long nplane = (long)(sh.np[2] - sh.in[2]) * sh.np[0] * sh.np[1];
    const long nrow = (long)sh.in[2] * (sh.np[1] - sh.in[1]) * sh.np[0];
    if (t >= bc_shell_count(sh) || clk->halt) return;
    int i, j, k;
    if (t < nplane) {
        const long pl = (long)sh.np[0] * sh.np[1];
        const int p = (int)(t / pl), r = (int)(t % pl);
        k = p < sh.lo[2] ? p : p + sh.in[2];
        j = r / sh.np[0];
        i = r % sh.np[0];
    } else if (t < nplane + nrow) {
        const long q = t - nplane;
        const int gj = sh.np[1] - sh.in[1], row = (int)(q / sh.np[0]);
        i = (int)(q % sh.np[0]);
        k = sh.lo[2] + row / gj;
        const int r = row % gj;
        j = r < sh.lo[1] ? r : r + sh.in[1];
    } else {
        const long q = t - nplane - nrow;
        const int gi = sh.np[0] - sh.in[0];
        const int r = (int)(q % gi), rr = (int)(q / gi);
        i = r < sh.lo[0] ? r : r + sh.in[0];
        j = sh.lo[1] + rr % sh.in[1];
        k = sh.lo[2] + rr / sh.in[1];
    }
    const int idx = i + sh.np[0] * (j + sh.np[1] * k);
    const int a0 = g.nG;
    const size_t nc = g.nc, nf = g.nf;
    const int Nk_eff = g.Nkp;
    const bool mi = mask & 1, mj = mask & 2, mk = mask & 4;
    auto kmapc = [&](int kk) { return !mk ? kk : do25D ? ((kk <= g.Nk - 1) ? a0 : kk) : pmap_c(kk, a0, Nk_eff); };
    // cells
    if (i < g.Ni && j < g.Nj && k < g.Nk) {
        int si = mi ? pmap_c(i, a0, g.Nip) : i, sj = mj ? pmap_c(j, a0, g.Njp) : j, sk = kmapc(k);
        if (si != i || sj != j || sk != k) {
            int s = ci(g, si, sj, sk), d = ci(g, i, j, k);
#pragma unroll
            for (int v = 0; v < 5; v++) Gas[nc * v + d] = Gas[nc * v + s];
            if (B) {   // Bxyz ghosts are copied only where no bFlux2Fld follows (init.F90:230)
#pragma unroll
                for (int v = 0; v < 3; v++) B[nc * v + d] = B[nc * v + s];
            }
        }
    }
    // faces, per component
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int si = !mi ? i : (d == 0) ? pmap_f(i, a0, g.Nip) : pmap_c(i, a0, g.Nip);
        int sj = !mj ? j : (d == 1) ? pmap_f(j, a0, g.Njp) : pmap_c(j, a0, g.Njp);
        int sk;
        if (!mk) sk = k;
        else if (do25D) sk = (k <= g.Nk - 1) ? a0 : k;
        else sk = (d == 2) ? pmap_f(k, a0, g.Nkp) : pmap_c(k, a0, g.Nkp);
        // the +1 slot of a non-normal axis is never written by the reference's BCs
        if (d != 0 && i == g.Ni) si = i;
        if (d != 1 && j == g.Nj) sj = j;
        if (!do25D && d != 2 && k == g.Nk) sk = k;
        if (si != i || sj != j || sk != k) mag[nf * d + idx] = mag[nf * d + fi(g, si, sj, sk)];
    }
}
template <typename R>
__global__ void __launch_bounds__(256) k_bc(GVT<R> g, const DClock *__restrict__ clk, int do25D, int mask, R *__restrict__ Gas,
                                            double *__restrict__ mag, R *__restrict__ B) {
    b_bc<R>(blockIdx.x, g, clk, do25D, mask, Gas, mag, B);
}

// One face's non-periodic BC, restating bcs.F90: zeroGrad_{i,o}bc{I,J,K}
// (:799-930), zeroExten_{i,o}bc{I,J,K} (:454-603), cartReflect_{i,o}bc{I,J,K}
// (:607-795).  face: 0 INI, 1 OUTI, 2 INJ, 3 OUTJ, 4 INK, 5 OUTK.  One thread
// per (n, a, b): ghost layer n = 1..nG and the ghost-inclusive cell ranges of
// the other two axes (the reference's loops).  Every thread writes its own
// ghost cell and faces and reads only points no thread of the launch writes.
// withB: also the Bxyz ghosts (hatch; a step's bFlux2Fld recomputes them).
// face64/Tf64: the FP64 face areas and triads (zeroGrad_obcI's ghost fluxes).
// GAM_BC_SPHERE: SphIn (INI, the same assignments as zeroExten_ibcI) and SphOut
// (OUTI: every ghost face and cell from the last cell ie), bcs.F90:934-982.
enum { GAM_BC_PERIODIC = 0, GAM_BC_ZEROGRAD = 1, GAM_BC_ZEROEXTEN = 2, GAM_BC_REFLECT = 3, GAM_BC_SPHERE = 4, GAM_BC_LFM = 5,
       GAM_BC_WIND = 6, GAM_BC_CYLPOLE = 7, GAM_BC_LAZY = 8, GAM_BC_ION = 9 };
template <typename R>
__global__ void __launch_bounds__(256) k_bc_face(GVT<R> g, const DClock *__restrict__ clk, int face, int type, int withB,
                                                 const double *__restrict__ face64, const double *__restrict__ Tf64,
                                                 R *__restrict__ Gas, double *__restrict__ mag, R *__restrict__ B) {
    const int ax = face >> 1, out = face & 1;
    const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    const int ext[3] = {g.Ni, g.Nj, g.Nk};
    const int n1 = ext[t1], n2 = ext[t2];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.nG * n1 * n2 || clk->halt) return;
    int a, n, b;
    if (ax == 0) {   // i ghosts: the layer fastest, so that neighbouring lanes share cache lines
        n = 1 + (int)(idx % g.nG);
        const long r = idx / g.nG;
        a = (int)(r % n1);
        b = (int)(r / n1);
    } else {
        a = (int)(idx % n1);
        const long r = idx / n1;
        n = 1 + (int)(r % g.nG);
        b = (int)(r / g.nG);
    }
    const int act[3] = {g.Nip, g.Njp, g.Nkp};
    const int lo = g.nG, hi = g.nG + act[ax] - 1;
    const size_t nc = g.nc, nf = g.nf;
    int p[3], q[3];
    p[t1] = q[t1] = a;
    p[t2] = q[t2] = b;
    const int gi = out ? hi + n : lo - n;
    p[ax] = gi;
    q[ax] = type == GAM_BC_REFLECT ? (out ? hi - n + 1 : lo + n - 1) : (out ? hi : lo);
    const int cg = ci(g, p[0], p[1], p[2]), cs = ci(g, q[0], q[1], q[2]);
#pragma unroll
    for (int v = 0; v < 5; v++) {
        R x = Gas[nc * v + cs];
        if (type == GAM_BC_REFLECT && v == 1 + ax) x = R(-1.0) * x;
        Gas[nc * v + cg] = x;
    }
    if (withB && (type == GAM_BC_ZEROEXTEN || type == GAM_BC_SPHERE || (type == GAM_BC_REFLECT && !(ax == 0 && out)))) {
#pragma unroll
        for (int v = 0; v < 3; v++) B[nc * v + cg] = B[nc * v + cs];
    }
    const int fg = fi(g, p[0], p[1], p[2]), fs = fi(g, q[0], q[1], q[2]);
    int pn[3] = {p[0], p[1], p[2]}, qn[3] = {q[0], q[1], q[2]};
    if (!out) {
        if (type == GAM_BC_REFLECT) qn[ax] = lo + n;
        for (int d = 0; d < 3; d++) {
            if (d == ax) mag[nf * d + fi(g, pn[0], pn[1], pn[2])] = mag[nf * d + fi(g, qn[0], qn[1], qn[2])];
            else mag[nf * d + fg] = mag[nf * d + fs];
        }
    } else if (type == GAM_BC_ZEROGRAD && ax == 0) {
        // zeroGrad_obcI: ghost fluxes of the last active cell's Cartesian field
        double Bc[3];
        cell_bxyz(g, mag, hi, a, b, Bc);
        const int f1 = fi(g, gi + 1, a, b);
        for (int d = 0; d < 3; d++) {
            const int f = d == 0 ? f1 : fg;
            const double dotn = Bc[0] * Tf64[nf * (NORMX + 9 * d) + f] + Bc[1] * Tf64[nf * (NORMY + 9 * d) + f] +
                                Bc[2] * Tf64[nf * (NORMZ + 9 * d) + f];
            mag[nf * d + f] = face64[nf * d + f] * dotn;
        }
    } else {
        pn[ax] = gi + 1;
        qn[ax] = type == GAM_BC_REFLECT ? hi - n + 1 : (type == GAM_BC_SPHERE ? hi : hi + 1);
        // zeroGrad_obcK reads the j component at ke+1, which the routine's own
        // n = 1 pass has just set from ke: the value is magFlux(ke, j) for all n
        const bool zgk = type == GAM_BC_ZEROGRAD && ax == 2;
        if (zgk) qn[2] = hi;
        for (int d = 0; d < 3; d++) {
            if (d == ax) {
                mag[nf * d + fi(g, pn[0], pn[1], pn[2])] = mag[nf * (zgk ? 1 : d) + fi(g, qn[0], qn[1], qn[2])];
            } else if (type == GAM_BC_ZEROEXTEN && ax == 0) {
                mag[nf * d + fg] = (double)B[nc * d + ci(g, hi, a, b)];   // zeroExten_obcI: Bxyz of the last cell
            } else {
                mag[nf * d + fg] = mag[nf * d + fs];
            }
        }
    }
}

// ----------------------------------------------------------------- conversions
template <typename TI, typename TO>
__global__ void __launch_bounds__(256) k_cvt(const TI *__restrict__ in, TO *__restrict__ out, size_t n) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) out[idx] = TO(in[idx]);
}

// ----------------------------------------------------------------- diagnostics
// bcs.F90:987-1017 cylpole, the inner i face of a cylindrical grid whose j
// index goes once around the pole (Np = Njp cells, even): ghost layer n of
// column j takes the first cells half a turn away, jp = j + Np/2 wrapped past
// Np, the I and J face fluxes negated (their normals turn over across the
// pole), the K flux and the cell values as they are.  Every (j, k) of the cell
// box, as the reference loops jsg..jeg and ksg..keg (the J faces at jeg+1 and
// the K faces at keg+1 are left alone); sources are owned cells, so one pass.
template <typename R>
__global__ void __launch_bounds__(256) k_bc_cylpole(GVT<R> g, const DClock *__restrict__ clk, int withB,
                                                    R *__restrict__ Gas, double *__restrict__ mag, R *__restrict__ B) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.nG * g.Nj * g.Nk || clk->halt) return;
    const int nl = 1 + (int)(idx % g.nG);   // ghost layer fastest
    const long r = idx / g.nG;
    const int j = (int)(r % g.Nj), k = (int)(r / g.Nj);
    const int Np = g.Njp;
    int jp = j + Np / 2;                    // the reference's jp = j + Np2, 1-based active j
    if (jp - g.nG + 1 > Np) jp -= Np;
    const int ig = g.nG - nl, is = g.nG + nl - 1;
    const size_t nc = g.nc, nf = g.nf;
    const int c = ci(g, ig, j, k), cs = ci(g, is, jp, k);
#pragma unroll
    for (int v = 0; v < 5; v++) Gas[nc * v + c] = Gas[nc * v + cs];
    if (withB) {
#pragma unroll
        for (int x = 0; x < 3; x++) B[nc * x + c] = B[nc * x + cs];
    }
    const int f = fi(g, ig, j, k);
    mag[f] = -1.0 * mag[fi(g, is + 1, jp, k)];
    mag[nf + f] = -1.0 * mag[nf + fi(g, is, jp, k)];
    mag[2 * nf + f] = mag[2 * nf + fi(g, is, jp, k)];
}

__global__ void k_divb(GV g, const double *__restrict__ mag, double *__restrict__ partial) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    const int n = g.Nip * g.Njp * g.Nkp;
    if (idx < n) {
        int a = idx % g.Nip, r = idx / g.Nip, b = r % g.Njp, cz = r / g.Njp;
        int i = a + g.nG, j = b + g.nG, k = cz + g.nG;
        const size_t nf = g.nf;
        int f = fi(g, i, j, k);
        double dFi = mag[f] - mag[f + 1];
        double dFj = mag[nf + f] - mag[nf + f + (g.Ni + 1)];
        double dFk = mag[2 * nf + f] - mag[2 * nf + f + (g.Ni + 1) * (g.Nj + 1)];
        s = fabs(dFi + dFj + dFk);
    }
#ifdef GAM_EMU
    partial[blockIdx.x] = threadIdx.x == 0 ? s : partial[blockIdx.x] + s;
    return;
#endif
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Sum of np values in index order (the order of gam_divb's host loop, so the
// totals agree bit for bit): each round stages 256 values in LDS and thread 0
// adds them in turn.  One block.
__global__ void __launch_bounds__(256) k_sum_ordered(const double *__restrict__ part, int np, double *__restrict__ out) {
#ifdef GAM_EMU
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int q = 0; q < np; q++) s += part[q];
    *out = s;
#else
    __shared__ double v[256];
    double s = 0.0;
    for (int q0 = 0; q0 < np; q0 += 256) {
        const int q = q0 + (int)threadIdx.x;
        v[threadIdx.x] = q < np ? part[q] : 0.0;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = np - q0 < 256 ? np - q0 : 256;
            for (int r = 0; r < m; r++) s += v[r];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s;
#endif
}

// ----------------------------------------------------------------- learned correction
// A correction field o[c][8] (5 gas-flux and 3 edge-EMF channels per cell,
// FP32) adds scale*face_d(c)*o[c][v] to the gas flux v of every active
// interface d at cell c and scale*edge_d(c)*o[c][5+d] to every active edge EMF
// d at cell c (the edge at the cell's lower corner), before CT and the update:
// the update stays conservative and E2Flux keeps div.B exact.  A ghost cell
// uses the o of its periodic image (the step's boundaries are periodic), so
// the two copies of a boundary face or edge get one correction.  The step
// adds it where k_emf/k_emf_jcol and k_flux/k_flux2 store (corr_add); the
// training kernels below use the same faces (flux_box) and edges (edge_active).
template <typename R> __device__ __forceinline__ bool flux_box(const GVT<R> &g, int do25D, int d, int i, int j, int k) {
    const int a0 = g.nG;
    const int nk = do25D ? 1 : g.Nkp;
    return i >= a0 && i < a0 + g.Nip + (d == 0) && j >= a0 && j < a0 + g.Njp + (d == 1) && k >= a0 &&
           k < a0 + nk + (d == 2) && !(do25D && d == 2);
}
template <typename R> __device__ __forceinline__ bool edge_active(const GVT<R> &g, int d, int i, int j, int k) {
    const int a0 = g.nG;
    return i >= a0 && i <= a0 + g.Nip - (d == 0) && j >= a0 && j <= a0 + g.Njp - (d == 1) && k >= a0 &&
           k <= a0 + g.Nkp - (d == 2);
}

// Training: the change the correction o makes to one step, linearised
// (U += dt*dGH for the gas, magFlux -= dt*curl(E)), against a target change.
// loss = sum_cells sum_v wG[v] (dU_v - tG_v)^2 + sum_faces wM (dPhi - tM)^2 over
// the active cells and the active faces (face index along its normal in
// [is, ie+1], the others in [is, ie]).  Pass 1: cell residuals rG = 2 wG (dU - tG) / dV.
template <typename R>
__global__ void __launch_bounds__(256) k_corr_cells(GVT<R> g, int do25D, const float *__restrict__ o, double s, double dt,
                                                    const double *__restrict__ tG, double w0, double w1, double w4,
                                                    double *__restrict__ rG, double *__restrict__ partial) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (idx < g.nc) {
        int i, j, k;
        decode_c(g, idx, i, j, k);
        const size_t nc = g.nc, nf = g.nf;
        const int a0 = g.nG;
        const bool active = i >= a0 && i < a0 + g.Nip && j >= a0 && j < a0 + g.Njp && k >= a0 && k < a0 + g.Nkp;
        const double wv[5] = {w0, w1, w1, w1, w4};
        for (int v = 0; v < 5; v++) {
            double r = 0.0;
            if (active) {
                const double V = g.dvol[idx];
                double sum = 0.0;
                for (int d = 0; d < 3; d++) {
                    if (do25D && d == 2) continue;
                    const int f = fi(g, i, j, k), fs = fstride(g, d);
                    const size_t up = pimg(g, i + (d == 0), j + (d == 1), k + (d == 2));
                    sum += s * (double)g.face[nf * d + f] * (double)o[(size_t)idx * 8 + v]
                         - s * (double)g.face[nf * d + f + fs] * (double)o[up * 8 + v];
                }
                const double dU = dt * sum / V;
                const double e = dU - tG[nc * v + idx];
                l += wv[v] * e * e;
                r = 2.0 * wv[v] * e / V;
            }
            rG[nc * v + idx] = r;
        }
    }
    block_sum_store(l, partial, blockIdx.x);
}

// Pass 2: face residuals rM = 2 wM (dPhi - tM) on the active faces
template <typename R>
__global__ void __launch_bounds__(256) k_corr_faces(GVT<R> g, const float *__restrict__ o, double s, double dt,
                                                    const double *__restrict__ tM, double wM, double *__restrict__ rM,
                                                    double *__restrict__ partial) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (idx < g.nf) {
        int i, j, k;
        decode_f(g, idx, i, j, k);
        const size_t nf = g.nf;
        const int a0 = g.nG;
        auto dE = [&](int d, int a, int b, int c) -> double {   // correction EMF on an active edge
            if (!edge_active(g, d, a, b, c)) return 0.0;
            return s * (double)g.edge[nf * d + fi(g, a, b, c)] * (double)o[pimg(g, a, b, c) * 8 + 5 + d];
        };
        for (int d = 0; d < 3; d++) {
            double r = 0.0;
            if (i >= a0 && i < a0 + g.Nip + (d == 0) && j >= a0 && j < a0 + g.Njp + (d == 1) && k >= a0 &&
                k < a0 + g.Nkp + (d == 2)) {
                double df;
                if (d == 0) df = dE(2, i, j + 1, k) - dE(2, i, j, k) - dE(1, i, j, k + 1) + dE(1, i, j, k);
                else if (d == 1) df = dE(0, i, j, k + 1) - dE(0, i, j, k) - dE(2, i + 1, j, k) + dE(2, i, j, k);
                else df = dE(1, i + 1, j, k) - dE(1, i, j, k) - dE(0, i, j + 1, k) + dE(0, i, j, k);
                const double e = -dt * df - tM[nf * d + idx];
                l += wM * e * e;
                r = 2.0 * wM * e;
            }
            rM[nf * d + idx] = r;
        }
    }
    block_sum_store(l, partial, blockIdx.x);
}

// Pass 3: dL/do per cell (adjoint of passes 1-2).  An active cell gathers the
// uses of itself and of its periodic images; a ghost cell's o is never read.
template <typename R>
__global__ void __launch_bounds__(256) k_corr_adjoint(GVT<R> g, int do25D, double s, double dt,
                                                      const double *__restrict__ rG, const double *__restrict__ rM,
                                                      float *__restrict__ dout) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i0, j0, k0;
    decode_c(g, idx, i0, j0, k0);
    const size_t nc = g.nc, nf = g.nf;
    const int a0 = g.nG;
    double go[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const bool active = i0 >= a0 && i0 < a0 + g.Nip && j0 >= a0 && j0 < a0 + g.Njp && k0 >= a0 && k0 < a0 + g.Nkp;
    const int fI = fstride(g, 0), fJ = fstride(g, 1), fK = fstride(g, 2);
    for (int q = 0; active && q < 8; q++) {
        const int i = i0 + (q & 1) * g.Nip, j = j0 + ((q >> 1) & 1) * g.Njp, k = k0 + (q >> 2) * g.Nkp;
        if (i >= g.Ni || j >= g.Nj || k >= g.Nk) continue;
        const int c = ci(g, i, j, k), f = fi(g, i, j, k);
        for (int d = 0; d < 3; d++) {
            if (!flux_box(g, do25D, d, i, j, k)) continue;
            const int st = cstride(g, d);
            const double A = (double)g.face[nf * d + f];
            // the face is the lower face of cell c and the upper face of c - st
            for (int v = 0; v < 5; v++) go[v] += s * A * dt * (rG[nc * v + c] - rG[nc * v + c - st]);
        }
        for (int d = 0; d < 3; d++) {
            if (!edge_active(g, d, i, j, k)) continue;
            const double *R0 = rM, *R1 = rM + nf, *R2 = rM + 2 * nf;
            double gE;
            if (d == 0) gE = -dt * (R1[f - fK] - R1[f] - R2[f - fJ] + R2[f]);
            else if (d == 1) gE = -dt * (-R0[f - fK] + R0[f] + R2[f - fI] - R2[f]);
            else gE = -dt * (R0[f - fJ] - R0[f] - R1[f - fI] + R1[f]);
            go[5 + d] += s * (double)g.edge[nf * d + f] * gE;
        }
    }
    for (int c = 0; c < 8; c++) dout[(size_t)idx * 8 + c] = (float)go[c];
}
