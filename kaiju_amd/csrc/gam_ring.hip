// gam_ring.hip -- the LFM-grid axis: through-the-axis ghost geometry and BCs,
// the axis fixes of the fluxes and EMFs, and the ring average (reference
// src/gamera/ringutils.F90, ringav.F90, ringrecon.F90, bcs.F90:1020-1106,
// gamdebug.F90:91-175).  Included by gam_api.hip.
//
// Grid: j runs from one axis (the j = js faces) to the other (j = je+1), k goes
// once around the axis in Np = Nkp cells.  A ghost cell j = js-n sits across
// the axis from the active cell js+n-1 half a turn away in k; the face and
// edge maps below follow the reference's index rules.  Ring r of a pole is
// the cell row js+r-1 (or je-r+1), averaged in nch[r-1] chunks along k.
//
// Indices here are 0-based and ghost-inclusive (active cells nG .. nG+N-1);
// the k numbering of the reference's WrapK is 1-based from ks.

#define GAM_RING_MAX 32
#define GAM_RING_NPMAX 512

// Block-cooperative ring kernels stride over a ring with the block's threads
// between __syncthreads() phases; the CPU emulation (threads of a block run
// one after another) does every phase on thread 0 instead.
#ifdef GAM_EMU
#define RING_T0 0
#define RING_NT 1
#define RING_ONE_THREAD() do { if (threadIdx.x != 0) return; } while (0)
#else
#define RING_T0 ((int)threadIdx.x)
#define RING_NT ((int)blockDim.x)
#define RING_ONE_THREAD() do { } while (0)
#endif

struct RingP {
    int on, nring;
    int nch[GAM_RING_MAX];
};

// ringutils.F90:631-639 WrapK on a ghost-inclusive k
__host__ __device__ __forceinline__ int ring_wrapk(int k, int nG, int Np) {
    int kr = k - nG + 1 + Np / 2;
    if (kr > Np) kr = kr - Np;
    return kr + nG - 1;
}

// ringutils.F90:483-520 lfmIJKcc
__host__ __device__ __forceinline__ void lfm_cc(int nG, int Njp, int Nkp, int j, int k, int &jp, int &kp) {
    const int js = nG, je = nG + Njp - 1, ks = nG, ke = nG + Nkp - 1;
    int kk = k < ks ? k + Nkp : (k > ke ? k - Nkp : k), jj = j;
    if (j < js) { jj = js + (js - j) - 1; kk = ring_wrapk(kk, nG, Nkp); }
    if (j > je) { jj = je - (j - je) + 1; kk = ring_wrapk(kk, nG, Nkp); }
    jp = jj; kp = kk;
}

// ringutils.F90:523-577 lfmIJKfc, face direction d
__host__ __device__ __forceinline__ void lfm_fc(int nG, int Njp, int Nkp, int d, int j, int k, int &jp, int &kp) {
    if (d == 0) { lfm_cc(nG, Njp, Nkp, j, k, jp, kp); return; }
    const int js = nG, je = nG + Njp - 1, ks = nG, ke = nG + Nkp - 1;
    int kk = k, jj = j;
    if (k < ks) kk = k + Nkp;
    if (d == 1 && k > ke) kk = k - Nkp;
    if (d == 2 && k > ke + 1) kk = k - Nkp;
    if (d == 1) {
        if (j < js) { jj = js + (js - j); kk = ring_wrapk(k, nG, Nkp); }
        if (j > je + 1) { jj = je + 1 - (j - je) + 1; kk = ring_wrapk(k, nG, Nkp); }
    } else {
        if (j < js) { jj = js + (js - j) - 1; kk = ring_wrapk(k, nG, Nkp); }
        if (j > je) { jj = je - (j - je) + 1; kk = ring_wrapk(k, nG, Nkp); }
    }
    jp = jj; kp = kk;
}

// ----------------------------------------------------------------- grid (RingGridFix)
// ringutils.F90:262-280: the CFL length along k of ring r's cells times the
// chunk width Np/nch[r-1] (all i and k, both poles)
__global__ void __launch_bounds__(256) k_ring_dk(GVw g, RingP rp) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)g.Ni * g.Nk * rp.nring * 2;
    if (idx >= n) return;
    const int i = (int)(idx % g.Ni);
    long r = idx / g.Ni;
    const int k = (int)(r % g.Nk);
    r /= g.Nk;
    const int ring = (int)(r % rp.nring), pole = (int)(r / rp.nring);
    const int j = pole == 0 ? g.nG + ring : g.nG + g.Njp - 1 - ring;
    const int dJ = g.Nkp / rp.nch[ring];
    const int c = ciw(g, i, j, k);
    g.dk[c] = dJ * g.dk[c];
}

// ringutils.F90:282-389: ghost volume, di/dj/dk, face areas and face normals
// across the axis (J and K normals flip); thread = (i, k, ghost layer n, pole)
// over i <= Ni, k <= Nk.  Sources are active rows, which no thread writes.
__global__ void __launch_bounds__(256) k_ring_ghost_geom(GVw g) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n = (long)(g.Ni + 1) * (g.Nk + 1) * g.nG * 2;
    if (idx >= n) return;
    const int i = (int)(idx % (g.Ni + 1));
    long r = idx / (g.Ni + 1);
    const int k = (int)(r % (g.Nk + 1));
    r /= (g.Nk + 1);
    const int nl = 1 + (int)(r % g.nG), pole = (int)(r / g.nG);
    const int js = g.nG, je = g.nG + g.Njp - 1;
    const int jg = pole == 0 ? js - nl : je + nl;
    int jp, kp;
    if (k < g.Nk && i < g.Ni) {
        lfm_cc(g.nG, g.Njp, g.Nkp, jg, k, jp, kp);
        const int c = ciw(g, i, jg, k), cs = ciw(g, i, jp, kp);
        g.vol[c] = g.vol[cs];
        g.di[c] = g.di[cs]; g.dj[c] = g.dj[cs]; g.dk[c] = g.dk[cs];
    }
    const size_t nf = g.nf;
    for (int d = 0; d < 3; d++) {
        const int jj = (d == 1 && pole == 1) ? jg + 1 : jg;   // past je+1 the first J ghost face is je+2
        lfm_fc(g.nG, g.Njp, g.Nkp, d, jj, k, jp, kp);
        const int f = fiw(g, i, jj, k), fs = fiw(g, i, jp, kp);
        for (int x = 0; x < 3; x++) {
            const double t = g.Tf[nf * (NORMX + x + 9 * d) + fs];
            g.Tf[nf * (NORMX + x + 9 * d) + f] = d == 0 ? t : -t;
        }
        g.face[nf * d + f] = g.face[nf * d + fs];
    }
}

// ----------------------------------------------------------------- state fixes
// bcs.F90:1020-1106 lfmIn (out = 0) / lfmOut (out = 1); thread = (i, k, n) over
// i <= Ni, k <= Nk.  Cells and Bxyz from the conjugate cell, I faces as they
// are, J and K faces with the sign flipped.
template <typename R>
__global__ void __launch_bounds__(256) k_bc_lfm(GVT<R> g, const DClock *__restrict__ clk, int out, R *__restrict__ Gas,
                                                double *__restrict__ mag, R *__restrict__ B) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)(g.Ni + 1) * (g.Nk + 1) * g.nG || clk->halt) return;
    const int i = (int)(idx % (g.Ni + 1));
    const long r = idx / (g.Ni + 1);
    const int k = (int)(r % (g.Nk + 1)), nl = 1 + (int)(r / (g.Nk + 1));
    const int jg = out ? g.nG + g.Njp - 1 + nl : g.nG - nl;
    const size_t nc = g.nc, nf = g.nf;
    int jp, kp;
    if (k < g.Nk && i < g.Ni) {
        lfm_cc(g.nG, g.Njp, g.Nkp, jg, k, jp, kp);
        const int c = ci(g, i, jg, k), cs = ci(g, i, jp, kp);
#pragma unroll
        for (int v = 0; v < 5; v++) Gas[nc * v + c] = Gas[nc * v + cs];
#pragma unroll
        for (int v = 0; v < 3; v++) B[nc * v + c] = B[nc * v + cs];
    }
    lfm_fc(g.nG, g.Njp, g.Nkp, 0, jg, k, jp, kp);
    mag[fi(g, i, jg, k)] = mag[fi(g, i, jp, kp)];
    const int jf = out ? jg + 1 : jg;
    lfm_fc(g.nG, g.Njp, g.Nkp, 1, jf, k, jp, kp);
    mag[nf + fi(g, i, jf, k)] = -mag[nf + fi(g, i, jp, kp)];
    lfm_fc(g.nG, g.Njp, g.Nkp, 2, jg, k, jp, kp);
    mag[2 * nf + fi(g, i, jg, k)] = -mag[2 * nf + fi(g, i, jp, kp)];
}

// ringutils.F90:393-444 RingFlux2FieldFix (after every bFlux2Fld): Bxyz of the
// j ghost cells from their conjugates; thread = (i, k, n, pole)
template <typename R>
__global__ void __launch_bounds__(256) k_ring_bfix(GVT<R> g, const DClock *__restrict__ clk, R *__restrict__ B) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.Ni * g.Nk * g.nG * 2 || clk->halt) return;
    const int i = (int)(idx % g.Ni);
    long r = idx / g.Ni;
    const int k = (int)(r % g.Nk);
    r /= g.Nk;
    const int nl = 1 + (int)(r % g.nG), pole = (int)(r / g.nG);
    const int jg = pole == 0 ? g.nG - nl : g.nG + g.Njp - 1 + nl;
    int jp, kp;
    lfm_cc(g.nG, g.Njp, g.Nkp, jg, k, jp, kp);
    const size_t nc = g.nc;
    const int c = ci(g, i, jg, k), cs = ci(g, i, jp, kp);
#pragma unroll
    for (int x = 0; x < 3; x++) B[nc * x + c] = B[nc * x + cs];
}

// ringutils.F90:664-679 RingFlux: no gas or magnetic flux through the axis
// faces (J faces js and je+1, all i and k); with MHD, gamdebug.F90:91-124
// FixGasFluxLFM: one k flux on the periodic seam (faces ks and ke+1 averaged).
// Flux slabs Flx[(8*d + v)*nc + c], v = 0..4 gas, 5..7 magnetic.
template <typename R>
__global__ void __launch_bounds__(256) k_ring_flux(GVT<R> g, const DClock *__restrict__ clk, int mhd, R *__restrict__ Flx) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nA = (long)g.Ni * g.Nk * 2, nB = (long)g.Nip * g.Njp;
    if (idx >= nA + nB || clk->halt) return;
    const size_t nc = g.nc;
    const int nv = mhd ? 8 : 5;
    if (idx < nA) {
        const int i = (int)(idx % g.Ni);
        const long r = idx / g.Ni;
        const int k = (int)(r % g.Nk), pole = (int)(r / g.Nk);
        const int j = pole == 0 ? g.nG : g.nG + g.Njp;
        const int c = ci(g, i, j, k);
        for (int v = 0; v < nv; v++) Flx[nc * (8 + v) + c] = R(0);
        return;
    }
    if (!mhd) return;
    const long q = idx - nA;
    const int i = g.nG + (int)(q % g.Nip), j = g.nG + (int)(q / g.Nip);
    const int a = ci(g, i, j, g.nG), b = ci(g, i, j, g.nG + g.Nkp);
    for (int v = 0; v < 8; v++) {
        const double m = 0.5 * ((double)Flx[nc * (16 + v) + a] + (double)Flx[nc * (16 + v) + b]);
        Flx[nc * (16 + v) + a] = R(m);
        Flx[nc * (16 + v) + b] = R(m);
    }
}

// ringav.F90:402-431 PoleE ("lfm") on the axis rows js and je+1: no k EMF, the
// i EMF is its ring mean (summed in k order); one block per (slice i in
// is..ie+1, axis row), the column staged in LDS.
template <typename R>
__global__ void __launch_bounds__(256) k_pole_e(GVT<R> g, const DClock *__restrict__ clk, R *__restrict__ E) {
    __shared__ double col[GAM_RING_NPMAX];
    __shared__ double mean;
    if (clk->halt) return;
    RING_ONE_THREAD();
    const int i = g.nG + (int)(blockIdx.x % (g.Nip + 1)), pole = (int)(blockIdx.x / (g.Nip + 1));
    const int j = pole ? g.nG + g.Njp : g.nG, ks = g.nG, Np = g.Nkp;
    const size_t nc = g.nc;
    R *Ei = E, *Ek = E + 2 * nc;
    for (int n = RING_T0; n < Np; n += RING_NT) col[n] = (double)Ei[ci(g, i, j, ks + n)];
    __syncthreads();
    if (RING_T0 == 0) {
        double sum = 0.0;
#pragma unroll 8
        for (int n = 0; n < Np; n++) sum = sum + col[n];
        mean = sum / Np;
    }
    __syncthreads();
    const R avg = R(mean);
    for (int k = RING_T0; k < g.Nk; k += RING_NT) {
        Ek[ci(g, i, j, k)] = R(0);
        Ei[ci(g, i, j, k)] = avg;
    }
}

// gamdebug.F90:149-175 FixEFieldLFM (after PoleE): one i and one j EMF on the
// periodic seam, ks and ke+1 averaged; thread = (i, j) over i in is..ie+1,
// j in js..je+1
template <typename R>
__global__ void __launch_bounds__(256) k_fix_e(GVT<R> g, const DClock *__restrict__ clk, R *__restrict__ E) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)(g.Nip + 1) * (g.Njp + 1) || clk->halt) return;
    const int i = g.nG + (int)(idx % (g.Nip + 1)), j = g.nG + (int)(idx / (g.Nip + 1));
    const size_t nc = g.nc;
    R *Ei = E, *Ej = E + nc;
    const int a = ci(g, i, j, g.nG), b = ci(g, i, j, g.nG + g.Nkp);
    const R ei = R(0.5 * ((double)Ei[a] + (double)Ei[b])), ej = R(0.5 * ((double)Ej[a] + (double)Ej[b]));
    Ei[a] = ei; Ei[b] = ei; Ej[a] = ej; Ej[b] = ej;
}

// ----------------------------------------------------------------- ring average
// math.F90:192-197 minmod and ringrecon.F90 PPM_IntLR
__device__ __forceinline__ double ring_minmod(double x, double y) {
    return 0.5 * (copysign(1.0, x) + copysign(1.0, y)) * fmin(fabs(x), fabs(y));
}
__device__ __forceinline__ void ring_ppm(double fm2, double fm1, double f, double fp1, double fp2, double &vm, double &vp) {
    const double dvm2 = fm1 - fm2, dvm1 = f - fm1, dvp1 = fp1 - f, dvp2 = fp2 - fp1;
    double dvc = 0.5 * (dvm2 + dvm1), SM = 2.0 * ring_minmod(dvm2, dvm1);
    const double Sm1 = ring_minmod(dvc, SM);
    dvc = 0.5 * (dvm1 + dvp1); SM = 2.0 * ring_minmod(dvm1, dvp1);
    const double Sp1 = ring_minmod(dvc, SM);
    dvc = 0.5 * (dvp2 + dvp1); SM = 2.0 * ring_minmod(dvp2, dvp1);
    const double Sp2 = ring_minmod(dvc, SM);
    const double p = 0.5 * (f + fp1) - (Sp2 - Sp1) / 6.0;
    const double q = 0.5 * (f + fm1) - (Sp1 - Sm1) / 6.0;
    double ap = p - f, am = q - f;
    if (ap * am >= 0.0) { ap = 0.0; am = 0.0; }
    else {
        if (fabs(ap) >= 2.0 * fabs(am)) ap = -2.0 * am;
        if (fabs(am) >= 2.0 * fabs(ap)) am = -2.0 * ap;
    }
    vp = f + ap;
    vm = f + am;
}

// ringrecon.F90:33-87 ReconstructRing for cell n of a ring whose Nc chunk
// means are mean[0], mean[stride], ... (periodic in the chunk index): the
// cell mean of its chunk's PPM parabola
struct RingParab {
    double a, b, cc;
};
__device__ __forceinline__ RingParab ring_recon_chunk(const double *mean, int Nc, int c, int stride) {
    auto ch = [&](int q) { return mean[(size_t)(((q % Nc) + Nc) % Nc) * stride]; };
    double fL, fR;
    const double f = ch(c);
    ring_ppm(ch(c - 2), ch(c - 1), f, ch(c + 1), ch(c + 2), fL, fR);
    return RingParab{3 * (fL + fR - 2 * f), 2 * (3 * f - fR - 2 * fL), fL};
}
// cell m = 1..dJ of the chunk
__device__ __forceinline__ double ring_recon_eval(const RingParab &p, int m, int dJ) {
    return (p.a / 3.0) * (3 * m * m - 3 * m + 1) / (dJ * dJ) + 0.5 * p.b * (2 * m - 1) / dJ + p.cc;
}
__device__ __forceinline__ double ring_recon_cell(const double *mean, int Nc, int Np, int n, int stride = 1) {
    const int dJ = Np / Nc, c = n / dJ, m = n - c * dJ + 1;
    return ring_recon_eval(ring_recon_chunk(mean, Nc, c, stride), m, dJ);
}

// ringav.F90:206-329 CalcSmoothE, in three passes over threads with the slice
// i in is..ie+1 fastest (a wave reads whole cache lines of the face fluxes).
// Per (slice, ring, pole) the reference pulls the ring's J face fluxes (d = 0,
// their EMF goes to the I edges) and I face fluxes (d = 1, to the J edges),
// removes their m = 0, 1 Fourier modes (CleanRing, NFTMAG = 1), chunk-averages
// and PPM-reconstructs the rest, adds the modes back (DirtyRing) and turns
// the flux change into an EMF summed along each chunk.  SE (FP64, G-sized,
// zero outside the ring rows) gets that EMF on the ring rows; the PoleE and
// FixEFieldLFM the reference applies to SE next leave it unchanged (the axis
// rows stay zero, the seam values are written equal), so they are not
// launched.  tab: cos, sin of phi_n.  Scratch (gam_set_ring sizes it):
// chunk means [pole][ring][d][Np/2][Nip+1], then the Fourier coefficients
// [pole][ring][d][4][Nip+1].
struct SmoothEIdx {
    int s, d, j;   // slice, flux kind, row of the fluxes and of their EMF
};
__device__ __forceinline__ SmoothEIdx smoothe_idx(const GV &g, int s, int d, int nR, int pole) {
    const int js = g.nG, je = g.nG + g.Njp - 1;
    // J faces -> I edges on row js+nR (je-nR+1); I faces -> J edges on js+nR-1 (je-nR+1)
    const int j = pole ? je - nR + 1 : (d == 0 ? js + nR : js + nR - 1);
    return SmoothEIdx{s, d, j};
}
__device__ __forceinline__ double smoothe_flux(const GV &g, const SmoothEIdx &x, const double *__restrict__ mag, int n) {
    const size_t f = fi(g, g.nG + x.s, x.j, g.nG + n);
    return x.d == 0 ? mag[g.nf + f] : mag[f];
}

// CleanRing coefficients, each a sequential sum over the ring; thread = (slice,
// d, ring, pole)
__global__ void __launch_bounds__(256) k_ring_se_coef(GV g, RingP rp, const DClock *__restrict__ clk,
                                                     const double *__restrict__ tab, const double *__restrict__ mag,
                                                     double *__restrict__ scr) {
    const int Np = g.Nkp, ns = g.Nip + 1;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ns * 2 * rp.nring * 2 || clk->halt) return;
    const int s = (int)(idx % ns);
    long r = idx / ns;
    const int d = (int)(r % 2);
    r /= 2;
    const int nR = 1 + (int)(r % rp.nring), pole = (int)(r / rp.nring);
    const SmoothEIdx x = smoothe_idx(g, s, d, nR, pole);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    constexpr int NB = 16;   // loads issued ahead of the (sequential) sums
    for (int n0 = 0; n0 < Np; n0 += NB) {
        double B[NB], C[NB], S[NB];
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int n = n0 + u < Np ? n0 + u : Np - 1;
            B[u] = smoothe_flux(g, x, mag, n);
            C[u] = tab[n];
            S[u] = tab[Np + n];
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            if (n0 + u < Np) {
                s0 = s0 + B[u] * 1.0;
                s1 = s1 + B[u] * 0.0;
                s2 = s2 + B[u] * C[u];
                s3 = s3 + B[u] * S[u];
            }
        }
    }
    const double aScl = 2.0 / Np;
    const size_t hc = Np / 2;
    double *cf = scr + (size_t)2 * rp.nring * 2 * hc * ns + ((size_t)((pole * rp.nring + nR - 1) * 2 + d) * 4) * ns + s;
    cf[0] = 0.5 * aScl * s0;
    cf[ns] = 0.5 * aScl * s1;
    cf[2 * (size_t)ns] = 1.0 * aScl * s2;
    cf[3 * (size_t)ns] = 1.0 * aScl * s3;
}

// the cleaned flux of cell n
__device__ __forceinline__ double smoothe_clean(const double *cf, size_t ns, double B, double c, double sn) {
    const double q = B - cf[0] * 1.0 - cf[ns] * 0.0;
    return q - cf[2 * ns] * c - cf[3 * ns] * sn;
}

// chunk means of the cleaned fluxes (each summed in order of k); thread =
// (slice, chunk, d, ring, pole)
__global__ void __launch_bounds__(256) k_ring_se_mean(GV g, RingP rp, const DClock *__restrict__ clk,
                                                     const double *__restrict__ tab, const double *__restrict__ mag,
                                                     double *__restrict__ scr) {
    const int Np = g.Nkp, ns = g.Nip + 1, hc = Np / 2;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ns * hc * 2 * rp.nring * 2 || clk->halt) return;
    const int s = (int)(idx % ns);
    long r = idx / ns;
    const int c = (int)(r % hc);
    r /= hc;
    const int d = (int)(r % 2);
    r /= 2;
    const int nR = 1 + (int)(r % rp.nring), pole = (int)(r / rp.nring);
    const int NumCh = rp.nch[nR - 1];
    if (c >= NumCh) return;
    const int dC = Np / NumCh;
    const SmoothEIdx x = smoothe_idx(g, s, d, nR, pole);
    const int pr = (pole * rp.nring + nR - 1) * 2 + d;
    const double *cf = scr + (size_t)2 * rp.nring * 2 * hc * ns + (size_t)pr * 4 * ns + s;
    double sum = 0.0;
    for (int m = 0; m < dC; m++) {
        const int n = c * dC + m;
        sum = sum + smoothe_clean(cf, ns, smoothe_flux(g, x, mag, n), tab[n], tab[Np + n]);
    }
    scr[((size_t)pr * hc + c) * ns + s] = sum / dC;
}

// reconstruct, DirtyRing, the flux change, and its EMF summed along the chunk
// (zero at the chunk's first edge) into SE; thread = (slice, chunk, d, ring,
// pole).  The k = ke+1 edges repeat k = ks (a chunk's first edge: zero).
__global__ void __launch_bounds__(256) k_ring_se_emf(GV g, RingP rp, const DClock *__restrict__ clk,
                                                    const double *__restrict__ tab, const double *__restrict__ mag,
                                                    const double *__restrict__ scr, double *__restrict__ SE) {
    const int Np = g.Nkp, ns = g.Nip + 1, hc = Np / 2;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ns * hc * 2 * rp.nring * 2 || clk->halt) return;
    const int s = (int)(idx % ns);
    long r = idx / ns;
    const int c = (int)(r % hc);
    r /= hc;
    const int d = (int)(r % 2);
    r /= 2;
    const int nR = 1 + (int)(r % rp.nring), pole = (int)(r / rp.nring);
    const int NumCh = rp.nch[nR - 1];
    if (c >= NumCh) return;
    const int dC = Np / NumCh;
    const SmoothEIdx x = smoothe_idx(g, s, d, nR, pole);
    const int pr = (pole * rp.nring + nR - 1) * 2 + d;
    const double *cf = scr + (size_t)2 * rp.nring * 2 * hc * ns + (size_t)pr * 4 * ns + s;
    const double *mean = scr + (size_t)pr * hc * ns + s;
    const double eS = d == 0 ? 1.0 : -1.0;
    double *E = SE + (d == 0 ? 0 : g.nc);
    const int lRs = c * dC;
    const RingParab pb = ring_recon_chunk(mean, NumCh, c, ns);
    const double c0 = cf[0], c1 = cf[ns], c2 = cf[2 * (size_t)ns], c3 = cf[3 * (size_t)ns];
    double cum = 0.0;
    constexpr int NB = 8;   // loads issued ahead of the (sequential) chunk sum
    for (int m0 = 1; m0 <= dC; m0 += NB) {
        double B[NB];
#pragma unroll
        for (int u = 0; u < NB; u++) B[u] = smoothe_flux(g, x, mag, lRs + (m0 + u <= dC ? m0 + u : dC) - 1);
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int m = m0 + u;
            if (m > dC) break;
            const int n = lRs + m - 1;
            double q = ring_recon_eval(pb, m, dC) + c0 * 1.0 + c1 * 0.0;
            q = q + c2 * tab[n] + c3 * tab[Np + n];
            const double dF = q - B[u];
            const int e = ci(g, g.nG + s, x.j, g.nG + n);
            if (m == 1) {
                E[e] = 0.0;
                cum = -dF;
            } else {
                E[e] = eS * cum;
                cum = cum - dF;
            }
        }
    }
    if (c == 0) E[ci(g, g.nG + s, x.j, g.nG + Np)] = 0.0;
}

// fields.F90:18-51 E2Flux with dt = 1 and E = SE, in place, on the faces whose
// curl can see a ring row: j in [js, js+nring+1] and [je-nring, je+1] (the
// library keeps these bands apart); thread = (i, k, band row) over i in
// is..ie+1, k in ks..ke+1
__global__ void __launch_bounds__(256) k_ring_ct(GV g, RingP rp, const DClock *__restrict__ clk, const double *__restrict__ SE,
                                                 double *__restrict__ mag) {
    const int nb = rp.nring + 2;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)(g.Nip + 1) * (g.Nkp + 1) * 2 * nb || clk->halt) return;
    const int i = g.nG + (int)(idx % (g.Nip + 1));
    long r = idx / (g.Nip + 1);
    const int k = g.nG + (int)(r % (g.Nkp + 1));
    const int row = (int)(r / (g.Nkp + 1));
    const int j = row < nb ? g.nG + row : g.nG + g.Njp - (row - nb);
    const size_t nc = g.nc, nf = g.nf;
    const double *Ei = SE, *Ej = SE + nc, *Ek = SE + 2 * nc;
    const int c = ci(g, i, j, k);
    const double dfi = Ek[ci(g, i, j + 1, k)] - Ek[c] - Ej[ci(g, i, j, k + 1)] + Ej[c];
    const double dfj = Ei[ci(g, i, j, k + 1)] - Ei[c] - Ek[ci(g, i + 1, j, k)] + Ek[c];
    const double dfk = Ej[ci(g, i + 1, j, k)] - Ej[c] - Ei[ci(g, i, j + 1, k)] + Ei[c];
    const int f = fi(g, i, j, k);
    mag[f] = mag[f] - 1.0 * dfi;
    mag[nf + f] = mag[nf + f] - 1.0 * dfj;
    mag[2 * nf + f] = mag[2 * nf + f] - 1.0 * dfk;
}

// ringav.F90:349-388 CleanLoops (doFastLoop): the ring mean of the K fluxes of
// the cells next to the axis (rows js and je) is removed; one block per (slice
// i in is..ie+1, pole), the column staged in LDS
__global__ void __launch_bounds__(256) k_ring_loops(GV g, const DClock *__restrict__ clk, double *__restrict__ mag) {
    __shared__ double col[GAM_RING_NPMAX];
    __shared__ double mean;
    if (clk->halt) return;
    RING_ONE_THREAD();
    const int i = g.nG + (int)(blockIdx.x % (g.Nip + 1)), pole = (int)(blockIdx.x / (g.Nip + 1));
    const int j = pole == 0 ? g.nG : g.nG + g.Njp - 1, ks = g.nG, Np = g.Nkp;
    const size_t nf = g.nf;
    for (int n = RING_T0; n < Np; n += RING_NT) col[n] = mag[2 * nf + fi(g, i, j, ks + n)];
    __syncthreads();
    if (RING_T0 == 0) {
        double s = 0.0;
#pragma unroll 8
        for (int n = 0; n < Np; n++) s = s + col[n];
        mean = s / Np;
    }
    __syncthreads();
    const double avg = mean;
    for (int n = RING_T0; n <= Np; n += RING_NT) {
        const int f = 2 * nf + fi(g, i, j, ks + n);
        mag[f] = (n < Np ? col[n] : mag[f]) - 1.0 * avg;
    }
}

// ringav.F90:76-200 SmoothGas (single fluid, doMassRA = F, no shift, no
// velocity cleaning), in two passes.  First the chunk means of the ring
// variables rho, momentum, internal energy (Gas2Ring); thread = (i in is..ie,
// chunk, ring, pole), i fastest so that a wave reads whole cache lines, the
// chunk's cells summed in order of k.  mean[((pole*nring + ring)*5 + v)*(Np/2)
// + chunk][i - is] (G-independent scratch sized by gam_set_ring).
template <typename R>
__global__ void __launch_bounds__(256) k_ring_gas_mean(GVT<R> g, MP mp, RingP rp, const DClock *__restrict__ clk,
                                                       const R *__restrict__ Gas, double *__restrict__ mean) {
    const int Np = g.Nkp, hc = Np / 2;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.Nip * hc * rp.nring * 2 || clk->halt) return;
    const int i = (int)(idx % g.Nip);
    long r = idx / g.Nip;
    const int c = (int)(r % hc);
    r /= hc;
    const int nR = 1 + (int)(r % rp.nring), pole = (int)(r / rp.nring);
    const int NumCh = rp.nch[nR - 1];
    if (c >= NumCh) return;
    const int j = pole ? g.nG + g.Njp - 1 - (nR - 1) : g.nG + nR - 1;
    const int dC = Np / NumCh;
    const size_t nc = g.nc;
    const double gm1 = mp.gamma - 1;
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int m = 0; m < dC; m++) {   // PullRingCC + Gas2Ring
        const int cc = ci(g, g.nG + i, j, g.nG + c * dC + m);
        const double D = fmax((double)Gas[cc], mp.dFloor);
        const double Mx = Gas[nc + cc], My = Gas[2 * nc + cc], Mz = Gas[3 * nc + cc], E = Gas[4 * nc + cc];
        const double KinE = 0.5 * (Mx * Mx + My * My + Mz * Mz) / D;
        const double P = fmax(gm1 * (E - KinE), mp.pFloor);
        s[0] = s[0] + D; s[1] = s[1] + Mx; s[2] = s[2] + My; s[3] = s[3] + Mz;
        s[4] = s[4] + P / gm1;
    }
    const size_t sv = (size_t)hc * g.Nip;
    double *o = mean + (size_t)((pole * rp.nring + nR - 1) * 5) * sv + (size_t)c * g.Nip + i;
#pragma unroll
    for (int v = 0; v < 5; v++) o[v * sv] = s[v] / dC;
}

// Then each ring cell from the PPM reconstruction of its chunk's neighbourhood
// of means, turned back into conserved variables (Ring2Gas + PushRingCC);
// thread = (i in is..ie, k, ring, pole)
template <typename R>
__global__ void __launch_bounds__(256) k_ring_gas_set(GVT<R> g, MP mp, RingP rp, const DClock *__restrict__ clk,
                                                      const double *__restrict__ mean, R *__restrict__ Gas) {
    const int Np = g.Nkp, hc = Np / 2;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.Nip * Np * rp.nring * 2 || clk->halt) return;
    const int i = (int)(idx % g.Nip);
    long r = idx / g.Nip;
    const int n = (int)(r % Np);
    r /= Np;
    const int nR = 1 + (int)(r % rp.nring), pole = (int)(r / rp.nring);
    const int NumCh = rp.nch[nR - 1];
    const int j = pole ? g.nG + g.Njp - 1 - (nR - 1) : g.nG + nR - 1;
    const size_t nc = g.nc, sv = (size_t)hc * g.Nip;
    const double *mb = mean + (size_t)((pole * rp.nring + nR - 1) * 5) * sv + i;
    const double gm1 = mp.gamma - 1;
    double q[5];
#pragma unroll
    for (int v = 0; v < 5; v++) q[v] = ring_recon_cell(mb + v * sv, NumCh, Np, n, g.Nip);
    const double D = fmax(q[0], mp.dFloor);
    double P = gm1 * q[4];
    P = fmax(P, mp.pFloor);
    const double KinE = 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / D;
    const int cidx = ci(g, g.nG + i, j, g.nG + n);
    Gas[cidx] = R(D);
    Gas[nc + cidx] = R(q[1]); Gas[2 * nc + cidx] = R(q[2]); Gas[3 * nc + cidx] = R(q[3]);
    Gas[4 * nc + cidx] = R(KinE + P / gm1);
}

// The ring band: cell rows j in [js-1, js+nring+1] and [je-nring-1, je+1]
// (nb = nring + 3 rows each), the cells whose faces or Gas the ring average
// can change, with the axis ghost row.  Row r of 2*nb -> j.
__host__ __device__ __forceinline__ int ring_band_row(int nG, int Njp, int nb, int r) {
    return r < nb ? nG - 1 + r : nG + Njp - (r - nb);
}

// bFlux2Fld on the ring band (every other cell's faces are unchanged since
// k_update wrote its Bxyz); thread = (i, k, band row) over all i, k
template <typename R>
__global__ void __launch_bounds__(256) k_b2f_rows(GVT<R> g, const DClock *__restrict__ clk, int nb,
                                                  const double *__restrict__ mag, R *__restrict__ B) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.Ni * g.Nk * 2 * nb || clk->halt) return;
    const int i = (int)(idx % g.Ni);
    const long r = idx / g.Ni;
    const int k = (int)(r % g.Nk), j = ring_band_row(g.nG, g.Njp, nb, (int)(r / g.Nk));
    double b[3];
    cell_bxyz(g, mag, i, j, k, b);
    const int c = ci(g, i, j, k);
#pragma unroll
    for (int x = 0; x < 3; x++) B[(size_t)g.nc * x + c] = R(b[x]);
}

// CellDT block minima over the DT box cells of the ring band, after the ring
// average (k_update leaves these rows out of its minima); thread = (i, k, band
// row) over i in is-1..ie+1, k in ks-1..ke+1
template <int BORIS, typename R>
__global__ void __launch_bounds__(256) k_celldt_rows(GVT<R> g, MP m, BGV bg, const DClock *__restrict__ clk, int nb,
                                                     const R *__restrict__ Gas, const R *__restrict__ B,
                                                     double *__restrict__ partial) {
    if (clk->halt) return;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double dtl = GAM_HUGE;
    if (idx < (long)(g.Nip + 2) * (g.Nkp + 2) * 2 * nb) {
        const int i = g.nG - 1 + (int)(idx % (g.Nip + 2));
        const long r = idx / (g.Nip + 2);
        const int k = g.nG - 1 + (int)(r % (g.Nkp + 2)), j = ring_band_row(g.nG, g.Njp, nb, (int)(r / (g.Nkp + 2)));
        const int c = ci(g, i, j, k);
        const size_t nc = g.nc;
        double U[5], Bv[3];
#pragma unroll
        for (int v = 0; v < 5; v++) U[v] = (double)Gas[nc * v + c];
#pragma unroll
        for (int x = 0; x < 3; x++) Bv[x] = (double)B[nc * x + c];
        if (bg.on) {
#pragma unroll
            for (int x = 0; x < 3; x++) Bv[x] = Bv[x] + bg.B0[nc * x + c];
        }
        dtl = cell_dt<BORIS>(g, m, U, Bv, c);
    }
    block_min_store(dtl, partial, blockIdx.x);
}

// ----------------------------------------------------------------- solar-wind BC
// apps/wind.F90:206-374 WindBC on the outer i face, single fluid.  The wind is
// constant in time, or a time series as readWind (wind.F90:617-760) leaves it:
// nt rows of WIND_COLS doubles [t, rho, p, vx, vy, vz, bx, by, bz, xBS, yBS,
// zBS] in code units, uniform in t, bx already Bx0 + ByC by + BzC bz.
// jmap: the ghost Bxyz maps j across the axis (ring grids).
#define WIND_COLS 12
struct WindP {
    double rho, p, v[3], b[3];   // the constant wind (nt == 0)
    int on, nt;
    const double *ser;           // the series, nt x WIND_COLS (device memory)
    double tMin, tMax, dtW, ByC, BzC;
};

// InterpWgts (wind.F90:819-868) with discrete walls (doWindInterp false): the
// row, 0-based, whose values hold at time t
__device__ __forceinline__ int wind_row(const WindP &w, double t) {
    if (t >= w.tMax) return w.nt - 1;
    if (t <= w.tMin) return 0;
    const int i0 = (int)floor((t - w.tMin) / w.dtW);
    return i0 + 2 <= w.nt ? i0 : w.nt - 2;
}

// InterpWind (wind.F90:765-806) with discrete walls: q = rho, p, v[3], B[3] at
// time t; at or before the first sample the field is null (no spin-up sequence)
__device__ __forceinline__ void wind_get(const WindP &w, double t, double *q) {
    if (w.nt == 0) {
        q[0] = w.rho; q[1] = w.p;
        for (int x = 0; x < 3; x++) { q[2 + x] = w.v[x]; q[5 + x] = w.b[x]; }
        return;
    }
    const double *r = w.ser + (size_t)WIND_COLS * wind_row(w, t);
    for (int x = 0; x < 8; x++) q[x] = r[1 + x];
    if (t <= w.tMin) q[5] = q[6] = q[7] = 0.0;
}

// RefreshWind + PlaceWind (wind.F90:116-158, 808-817) at time t: the series'
// marker xyzW and the front velocity, |v(t)| against the flow, along x or
// tilted by the front coefficients ByC, BzC
struct WindFront {
    double xyzW[3], vFr[3];
};
__device__ __forceinline__ void wind_front(const WindP &w, double t, WindFront &fr) {
    const double *r = w.ser + (size_t)WIND_COLS * wind_row(w, t);
    for (int x = 0; x < 3; x++) fr.xyzW[x] = r[9 + x];
    double q[8];
    wind_get(w, t, q);
    const double vSpd = -sqrt(q[2] * q[2] + q[3] * q[3] + q[4] * q[4]);
    const double bcc = sqrt(w.ByC * w.ByC + w.BzC * w.BzC), btt = sqrt(1 + bcc * bcc);
    if (bcc > GAM_TINY) {
        const double CosBp = w.ByC / bcc, SinBp = w.BzC / bcc, CosF = 1.0 / btt, SinF = bcc / btt;
        fr.vFr[0] = vSpd * (CosF * CosF);
        fr.vFr[1] = -(vSpd * CosF * SinF * CosBp);
        fr.vFr[2] = -(vSpd * CosF * SinF * SinBp);
    } else {
        fr.vFr[0] = vSpd; fr.vFr[1] = 0.0; fr.vFr[2] = 0.0;
    }
}

// GetWindAt (wind.F90:160-176): the wind at point xyz and time t, taken at
// t - DelT, DelT the front's travel time from the marker to xyz
__device__ __forceinline__ void wind_at(const WindP &w, const WindFront &fr, const double *xyz, double t, double *q) {
    if (w.nt == 0) { wind_get(w, t, q); return; }
    const double vMag = sqrt(dot3(fr.vFr, fr.vFr));
    const double d[3] = {xyz[0] - fr.xyzW[0], xyz[1] - fr.xyzW[1], xyz[2] - fr.xyzW[2]};
    const double DelT = dot3(d, fr.vFr) / (vMag * vMag);
    wind_get(w, t - DelT, q);
}

// math.F90:97-111 RampDown
__device__ __forceinline__ double wind_ramp_down(double r, double rC, double lC) {
    if (r <= rC) return 1.0;
    if (r >= rC + lC) return 0.0;
    const double rScl = (r - rC) / lC;
    const double eArg = -1.0 / (1.0 - rScl * rScl);
    return exp(eArg + 1.0);
}

// apps/wind.F90:180-203 wgtWind, given the wind velocity v there
__device__ __forceinline__ double wind_wgt(const double *v, const double *xyz) {
    double nH[3], vH[3];
    normvec(xyz, nH);
    normvec(v, vH);
    const double d = -vH[0] * nH[0] + -vH[1] * nH[1] + -vH[2] * nH[2];
    const double dTh = acos(d) * 180.0 / GAM_PI;
    return wind_ramp_down(dTh, 90.0, 90.0);
}

// gamutils.F90:324-360 CellPress2Cs + Con2Fast
__device__ __forceinline__ double wind_con2fast(const MP &m, const double *U, const double *B) {
    const double rho = fmax(U[0], m.dFloor);
    const double Vx = U[1] / rho, Vy = U[2] / rho, Vz = U[3] / rho;
    const double KinE = 0.5 * rho * (sq(Vx) + sq(Vy) + sq(Vz));
    const double P = fmax((m.gamma - 1) * (U[4] - KinE), m.pFloor);
    const double Cs = sqrt(m.gamma * P / rho);
    const double MagB = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    double Va = MagB / sqrt(U[0]);
    if (m.doBoris) Va = m.Ca * Va / sqrt(sq(m.Ca) + sq(Va));
    return sqrt(sq(Cs) + sq(Va));
}

// Ghost cells (weight at the cell centre: the wind there (GetWindAt), or the
// last cell's state pushed outward at >= MinMach = 1 fast speed) and ghost face
// fluxes (the flux of the wind's field at the face centre mixed with the last
// cell's); the wind is taken at the clock's time; thread = (n, j, k) over j <= Nj,
// k <= Nk.  Faces that bound no cell are left alone.  FP64 arithmetic.
template <typename R>
__global__ void __launch_bounds__(256) k_bc_wind(GVT<R> g, MP m, WindP w, const DClock *__restrict__ clk,
                                                 const double *__restrict__ xyzcc, const double *__restrict__ xfc,
                                                 const double *__restrict__ face64, const double *__restrict__ Tf64,
                                                 R *__restrict__ Gas, double *__restrict__ mag, const R *__restrict__ B) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)(g.Nj + 1) * (g.Nk + 1) * g.nG || clk->halt) return;
    const int nl = 1 + (int)(idx % g.nG);   // ghost layer fastest: neighbouring lanes share cache lines
    const long r = idx / g.nG;
    const int j = (int)(r % (g.Nj + 1)), k = (int)(r / (g.Nj + 1));
    const int ip = g.nG + g.Nip - 1, ig = ip + nl;
    const size_t nc = g.nc, nf = g.nf;
    const double MinMach = 1.0;
    const double t = clk->t;
    WindFront fr = {};
    if (w.nt > 0) wind_front(w, t, fr);   // RefreshWind at Model%t
    if (k < g.Nk && j < g.Nj) {
        const int c = ci(g, ig, j, k), cp = ci(g, ip, j, k);
        const double xcc[3] = {xyzcc[c], xyzcc[nc + c], xyzcc[2 * nc + c]};
        double q[8];
        wind_at(w, fr, xcc, t, q);
        const double wSW = wind_wgt(q + 2, xcc), wMHD = 1.0 - wSW;
        double gSW[5] = {0, 0, 0, 0, 0}, gIN[5] = {0, 0, 0, 0, 0};
        if (wSW > GAM_TINY) {
            gSW[0] = q[0]; gSW[4] = q[1];
            for (int x = 0; x < 3; x++) gSW[1 + x] = q[2 + x];
        }
        if (wMHD > GAM_TINY) {
            double U[5], Bin[3];
            for (int v = 0; v < 5; v++) U[v] = (double)Gas[nc * v + cp];
            for (int x = 0; x < 3; x++) Bin[x] = (double)B[nc * x + cp];
            cell_c2p(m, U, gIN);
            const double Cx = wind_con2fast(m, U, Bin);
            const int fn = fi(g, ip + 1, j, k);
            const double nH[3] = {Tf64[nf * NORMX + fn], Tf64[nf * NORMY + fn], Tf64[nf * NORMZ + fn]};
            double V[3] = {gIN[1], gIN[2], gIN[3]};
            const double vn = V[0] * nH[0] + V[1] * nH[1] + V[2] * nH[2];
            if (vn < MinMach * Cx)
                for (int x = 0; x < 3; x++) V[x] = V[x] - vn * nH[x] + MinMach * Cx * nH[x];
            for (int x = 0; x < 3; x++) gIN[1 + x] = V[x];
        }
        double gW[5], U[5];
        for (int v = 0; v < 5; v++) gW[v] = wSW * gSW[v] + wMHD * gIN[v];
        cell_p2c(m, gW, U);
        for (int v = 0; v < 5; v++) Gas[nc * v + c] = R(U[v]);
    }
    for (int d = 0; d < 3; d++) {
        if ((d != 1 && j >= g.Nj) || (d != 2 && k >= g.Nk)) continue;
        const int f = fi(g, d == 0 ? ig + 1 : ig, j, k), fin = fi(g, d == 0 ? ip + 1 : ip, j, k);
        const double x[3] = {xfc[nf * (0 + 3 * d) + f], xfc[nf * (1 + 3 * d) + f], xfc[nf * (2 + 3 * d) + f]};
        double q[8];
        wind_at(w, fr, x, t, q);
        const double wSW = wind_wgt(q + 2, x);
        const double bn = Tf64[nf * (NORMX + 9 * d) + f] * q[5] + Tf64[nf * (NORMY + 9 * d) + f] * q[6] +
                          Tf64[nf * (NORMZ + 9 * d) + f] * q[7];
        const double swFlx = face64[nf * d + f] * bn;
        mag[nf * d + f] = wSW * swFlx + (1 - wSW) * mag[nf * d + fin];
    }
}

// ------------------------------------------------------- ionospheric inner BC
// voltron/ICs/earthcmi.F90:378-490 IonInner with its coupling field inExyz
// zero (no ionosphere model attached): the E x B velocity is then zero, and
// CellPlasmaBC (apps/msphingest.F90:268-372; single fluid, no source terms)
// gives each ghost cell the density and pressure of the cell mirrored across
// the inner face (i = is + n - 1; j, k through lfmIJKcc) at rest; every ghost
// face of the inner layers that bounds a cell carries zero flux (the field
// there is the background dipole alone, isGoodFaceIJK).  The ghost Bxyz follow
// from k_wind_bxyz / k_wind_bfix with inner set.  thread = (n, j, k) over
// j <= Nj, k <= Nk.  FP64 arithmetic.
template <typename R>
__global__ void __launch_bounds__(256) k_bc_ion(GVT<R> g, MP m, const DClock *__restrict__ clk, int jmap,
                                                R *__restrict__ Gas, double *__restrict__ mag) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)(g.Nj + 1) * (g.Nk + 1) * g.nG || clk->halt) return;
    const int nl = 1 + (int)(idx % g.nG);
    const long r = idx / g.nG;
    const int j = (int)(r % (g.Nj + 1)), k = (int)(r / (g.Nj + 1));
    const int ig = g.nG - nl, ip = g.nG + nl - 1;
    const size_t nc = g.nc, nf = g.nf;
    if (k < g.Nk && j < g.Nj) {
        int jp = j, kp;
        if (jmap) lfm_cc(g.nG, g.Njp, g.Nkp, j, k, jp, kp);
        else kp = k < g.nG ? k + g.Nkp : (k > g.nG + g.Nkp - 1 ? k - g.Nkp : k);
        const int c = ci(g, ig, j, k), cp = ci(g, ip, jp, kp);
        double U[5], W[5];
#pragma unroll
        for (int v = 0; v < 5; v++) U[v] = (double)Gas[nc * v + cp];
        cell_c2p(m, U, W);
        const double gW[5] = {W[0], 0.0, 0.0, 0.0, W[4]};
        cell_p2c(m, gW, U);
#pragma unroll
        for (int v = 0; v < 5; v++) Gas[nc * v + c] = R(U[v]);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if ((d != 1 && j >= g.Nj) || (d != 2 && k >= g.Nk)) continue;
        mag[nf * d + fi(g, ig, j, k)] = 0.0;
    }
}

// WindBC's (IonInner's when inner) ghost Bxyz: bFlux2Fld of the cell the map
// names (k periodic; j across the axis when jmap); thread = (i, j, k) over the
// ghost cells i > ie (i < is)
template <typename R>
__global__ void __launch_bounds__(256) k_wind_bxyz(GVT<R> g, const DClock *__restrict__ clk, int jmap, int inner,
                                                   const double *__restrict__ mag, R *__restrict__ B) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.Nj * g.Nk * g.nG || clk->halt) return;
    const int ig = inner ? g.nG - 1 - (int)(idx % g.nG) : g.nG + g.Nip + (int)(idx % g.nG);
    const long r = idx / g.nG;
    const int j = (int)(r % g.Nj), k = (int)(r / g.Nj);
    int jp = j, kp;
    if (jmap) lfm_cc(g.nG, g.Njp, g.Nkp, j, k, jp, kp);
    else kp = k < g.nG ? k + g.Nkp : (k > g.nG + g.Nkp - 1 ? k - g.Nkp : k);
    double b[3];
    cell_bxyz(g, mag, ig, jp, kp, b);
    const int c = ci(g, ig, j, k);
#pragma unroll
    for (int x = 0; x < 3; x++) B[(size_t)g.nc * x + c] = R(b[x]);
}

// then, on ring grids, the j ghosts of those i ghost cells from their
// conjugates; thread = (i > ie (i < is when inner), k, n, pole)
template <typename R>
__global__ void __launch_bounds__(256) k_wind_bfix(GVT<R> g, const DClock *__restrict__ clk, int inner, R *__restrict__ B) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)g.nG * g.Nk * g.nG * 2 || clk->halt) return;
    const int i = inner ? (int)(idx % g.nG) : g.nG + g.Nip + (int)(idx % g.nG);
    long r = idx / g.nG;
    const int k = (int)(r % g.Nk);
    r /= g.Nk;
    const int nl = 1 + (int)(r % g.nG), pole = (int)(r / g.nG);
    const int jg = pole == 0 ? g.nG - nl : g.nG + g.Njp - 1 + nl;
    int jp, kp;
    lfm_cc(g.nG, g.Njp, g.Nkp, jg, k, jp, kp);
    const int c = ci(g, i, jg, k), cs = ci(g, i, jp, kp);
#pragma unroll
    for (int x = 0; x < 3; x++) B[(size_t)g.nc * x + c] = B[(size_t)g.nc * x + cs];
}
