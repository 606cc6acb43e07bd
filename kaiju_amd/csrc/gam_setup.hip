// gam_setup.hip -- setup producers on the GPU: grid metric from corners and the
// initial conditions of the benchmark problems.  These run once per run; the
// quadratures (12-point Gauss per axis) are what make a 256^3 setup take
// seconds here instead of tens of minutes on the host.
//
// Reference: init.F90:801-1145 (Corners2Grid, ebGeom), quadrature.F90:42-389,
// metric.F90:79-148, prob.F90:74-119 (initOT2D), 190-298 (initBW),
// gridutils.F90:144-237 (GasIC2State, VectorPot2Flux).
#include "gam_common.h"

// quadrature.F90:11-26
__device__ __constant__ static const double gA[12] = {
    0.1252334085114689154724, 0.3678314989981801937527, 0.5873179542866174472967,
    0.7699026741943046870369, 0.9041172563704748566785, 0.9815606342467192506906,
    -0.1252334085114689154724, -0.3678314989981801937527, -0.5873179542866174472967,
    -0.7699026741943046870369, -0.9041172563704748566785, -0.9815606342467192506906};
__device__ __constant__ static const double gW[12] = {
    0.2491470458134027850006, 0.233492536538354808761, 0.203167426723065921749,
    0.160078328543346226335, 0.1069393259953184309603, 0.0471753363865118271946,
    0.2491470458134027850006, 0.233492536538354808761, 0.203167426723065921749,
    0.160078328543346226335, 0.1069393259953184309603, 0.0471753363865118271946};

struct GVw {   // writable grid view for the metric kernels
    int Ni, Nj, Nk, nG, Nip, Njp, Nkp, nc, nf;
    const double *xyz;
    double *xyzcc, *xfc, *face, *Tf, *edge, *Te, *Teb, *vol, *di, *dj, *dk;
};

__device__ __forceinline__ int fiw(const GVw &g, int i, int j, int k) { return i + (g.Ni + 1) * (j + (g.Nj + 1) * k); }
__device__ __forceinline__ int ciw(const GVw &g, int i, int j, int k) { return i + g.Ni * (j + g.Nj * k); }
__device__ __forceinline__ void xyzv(const GVw &g, int i, int j, int k, double *p) {
    int f = fiw(g, i, j, k);
    p[0] = g.xyz[f]; p[1] = g.xyz[(size_t)g.nf + f]; p[2] = g.xyz[2 * (size_t)g.nf + f];
}
__device__ __forceinline__ double nrm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {   // math.F90:265-274
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__global__ void k_fill(double *p, size_t n, double v) {
    size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (idx < n) p[idx] = v;
}

// Faces of direction d: area + centre (quadrature.F90:159-221 with A=1 and A=r)
// and the face triad (init.F90:873-906, doQuadFT = F).
__global__ void __launch_bounds__(256) k_faces(GVw g, int d) {
    const int nx = g.Ni + (d == 0), ny = g.Nj + (d == 1), nz = g.Nk + (d == 2);
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny * nz) return;
    int i = idx % nx, r = idx / nx, j = r % ny, k = r / ny;
    double f0[3], f1[3], f2[3], f3[3];   // metric.F90:79-108 faceCoords
    xyzv(g, i, j, k, f0);
    if (d == 0) { xyzv(g, i, j + 1, k, f1); xyzv(g, i, j, k + 1, f2); xyzv(g, i, j + 1, k + 1, f3); }
    else if (d == 1) { xyzv(g, i, j, k + 1, f1); xyzv(g, i + 1, j, k, f2); xyzv(g, i + 1, j, k + 1, f3); }
    else { xyzv(g, i + 1, j, k, f1); xyzv(g, i, j + 1, k, f2); xyzv(g, i + 1, j + 1, k, f3); }
    double dx[3], dy[3], dz[3];   // quadrature.F90:373-389
    dx[0] = f1[0] - f0[0]; dx[1] = f2[0] - f0[0]; dx[2] = f3[0] + f0[0] - f2[0] - f1[0];
    dy[0] = f1[1] - f0[1]; dy[1] = f2[1] - f0[1]; dy[2] = f3[1] + f0[1] - f2[1] - f1[1];
    dz[0] = f1[2] - f0[2]; dz[1] = f2[2] - f0[2]; dz[2] = f3[2] + f0[2] - f2[2] - f1[2];
    double A = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int a = 0; a < 12; a++)
        for (int b = 0; b < 12; b++) {
            double eta = 0.5 * (1 + gA[a]), psi = 0.5 * (1 + gA[b]);
            double w = gW[a] * gW[b];
            double xp = f0[0] + dx[0] * eta + dx[1] * psi + dx[2] * eta * psi;
            double yp = f0[1] + dy[0] * eta + dy[1] * psi + dy[2] * eta * psi;
            double zp = f0[2] + dz[0] * eta + dz[1] * psi + dz[2] * eta * psi;
            double eV[3] = {dx[0] + dx[2] * psi, dy[0] + dy[2] * psi, dz[0] + dz[2] * psi};
            double pV[3] = {dx[1] + dx[2] * eta, dy[1] + dy[2] * eta, dz[1] + dz[2] * eta};
            double cr[3];
            cross3(eV, pV, cr);
            double area = 0.25 * nrm3(cr) * w;
            A = A + area * 1.0;
            cx = cx + area * xp; cy = cy + area * yp; cz = cz + area * zp;
        }
    const size_t nf = g.nf;
    int f = fiw(g, i, j, k);
    g.xfc[nf * (0 + 3 * d) + f] = cx / A;
    g.xfc[nf * (1 + 3 * d) + f] = cy / A;
    g.xfc[nf * (2 + 3 * d) + f] = cz / A;
    g.face[nf * d + f] = A;
    double a4[3], b4[3], c4[3], e4[3], dT1[3], dT2[3];
    if (d == 0) {
        xyzv(g, i, j + 1, k + 1, a4); xyzv(g, i, j, k + 1, b4); xyzv(g, i, j + 1, k, c4); xyzv(g, i, j, k, e4);
        for (int x = 0; x < 3; x++) dT1[x] = 0.5 * (a4[x] - b4[x]) + 0.5 * (c4[x] - e4[x]);
        xyzv(g, i, j + 1, k + 1, a4); xyzv(g, i, j + 1, k, b4); xyzv(g, i, j, k + 1, c4);
        for (int x = 0; x < 3; x++) dT2[x] = 0.5 * (a4[x] - b4[x]) + 0.5 * (c4[x] - e4[x]);
    } else if (d == 1) {
        xyzv(g, i + 1, j, k + 1, a4); xyzv(g, i + 1, j, k, b4); xyzv(g, i, j, k + 1, c4); xyzv(g, i, j, k, e4);
        for (int x = 0; x < 3; x++) dT1[x] = 0.5 * (a4[x] - b4[x]) + 0.5 * (c4[x] - e4[x]);
        xyzv(g, i + 1, j, k + 1, a4); xyzv(g, i, j, k + 1, b4); xyzv(g, i + 1, j, k, c4);
        for (int x = 0; x < 3; x++) dT2[x] = 0.5 * (a4[x] - b4[x]) + 0.5 * (c4[x] - e4[x]);
    } else {
        xyzv(g, i + 1, j + 1, k, a4); xyzv(g, i, j + 1, k, b4); xyzv(g, i + 1, j, k, c4); xyzv(g, i, j, k, e4);
        for (int x = 0; x < 3; x++) dT1[x] = 0.5 * (a4[x] - b4[x]) + 0.5 * (c4[x] - e4[x]);
        xyzv(g, i + 1, j + 1, k, a4); xyzv(g, i + 1, j, k, b4); xyzv(g, i, j + 1, k, c4);
        for (int x = 0; x < 3; x++) dT2[x] = 0.5 * (a4[x] - b4[x]) + 0.5 * (c4[x] - e4[x]);
    }
    double cr[3], eN[3], eT1[3], eT2[3];
    cross3(dT1, dT2, cr);
    double n0 = nrm3(cr);
    for (int x = 0; x < 3; x++) eN[x] = cr[x] / n0;
    double n2 = nrm3(dT2);
    for (int x = 0; x < 3; x++) eT2[x] = dT2[x] / n2;
    cross3(eT2, eN, cr);
    double n1 = nrm3(cr);
    for (int x = 0; x < 3; x++) eT1[x] = cr[x] / n1;
    for (int x = 0; x < 3; x++) {
        g.Tf[nf * (NORMX + x + 9 * d) + f] = eN[x];
        g.Tf[nf * (TAN1X + x + 9 * d) + f] = eT1[x];
        g.Tf[nf * (TAN2X + x + 9 * d) + f] = eT2[x];
    }
}

// quadrature.F90:42-100 GaussianVolumeIntegral; W selects the integrand:
// 0 = CellDV (metric.F90:298-307), 1 = blast-wave gas (prob.F90:280-297)
struct BWp { double x0, y0, z0, D0, P0, pRat, dRat, rC; };

// quadrature.F90:42-100 GaussianVolumeIntegral of the integrand fn(x, y, z, W5)
template <class Fn>
__device__ void gauss_volume_fn(const double C[8][3], const Fn &fn, double *out5) {
    const int iS[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const int jS[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const int kS[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int v = 0; v < 5; v++) out5[v] = 0.0;
    for (int a = 0; a < 12; a++)
        for (int b = 0; b < 12; b++)
            for (int c = 0; c < 12; c++) {
                double wijk = gW[a] * gW[b] * gW[c];
                double xp[3], J[3][3];
#pragma unroll
                for (int x = 0; x < 3; x++) {
                    double s = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                    for (int n = 0; n < 8; n++) {
                        double iScl = iS[n] < 0 ? 1 - gA[a] : 1 + gA[a];
                        double jScl = jS[n] < 0 ? 1 - gA[b] : 1 + gA[b];
                        double kScl = kS[n] < 0 ? 1 - gA[c] : 1 + gA[c];
                        s += C[n][x] * (iScl * jScl * kScl);
                        s1 += C[n][x] * ((double)iS[n] * jScl * kScl);
                        s2 += C[n][x] * (iScl * (double)jS[n] * kScl);
                        s3 += C[n][x] * (iScl * jScl * (double)kS[n]);
                    }
                    xp[x] = 0.125 * s; J[x][0] = 0.125 * s1; J[x][1] = 0.125 * s2; J[x][2] = 0.125 * s3;
                }
                double dJ = fabs(J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1])
                               - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                               + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]));
                double Wp[5];
                fn(xp[0], xp[1], xp[2], Wp);
#pragma unroll
                for (int v = 0; v < 5; v++) out5[v] = out5[v] + dJ * wijk * Wp[v];
            }
}

// W selects the integrand: 0 = CellDV (metric.F90:298-307), 1 = blast-wave gas
// (prob.F90:280-297)
template <int W>
__device__ void gauss_volume(const double C[8][3], const BWp &bw, double *out5) {
    gauss_volume_fn(C, [&](double x, double y, double z, double *Wp) {
        if (W == 0) { Wp[0] = 1; Wp[1] = x; Wp[2] = y; Wp[3] = z; Wp[4] = 0; }
        else {
            double R = sqrt(sq(x - bw.x0) + sq(y - bw.y0) + sq(z - bw.z0));
            double D = bw.D0, P = bw.P0;
            if (R <= bw.rC) { D = bw.dRat * D; P = bw.pRat * P; }
            Wp[0] = D; Wp[1] = 0.0; Wp[2] = 0.0; Wp[3] = 0.0; Wp[4] = P;
        }
    }, out5);
}

__device__ __forceinline__ void cell_coords(const GVw &g, int i, int j, int k, double C[8][3]) {   // metric.F90:133-148
    xyzv(g, i, j, k, C[0]); xyzv(g, i + 1, j, k, C[1]); xyzv(g, i + 1, j + 1, k, C[2]); xyzv(g, i, j + 1, k, C[3]);
    xyzv(g, i, j, k + 1, C[4]); xyzv(g, i + 1, j, k + 1, C[5]); xyzv(g, i + 1, j + 1, k + 1, C[6]); xyzv(g, i, j + 1, k + 1, C[7]);
}

// init.F90:922-965 volume, barycentre, di/dj/dk (needs faces first)
__global__ void __launch_bounds__(256) k_volume(GVw g) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i = idx % g.Ni, r = idx / g.Ni, j = r % g.Nj, k = r / g.Nj;
    double C[8][3], W[5];
    cell_coords(g, i, j, k, C);
    BWp dummy = {};
    gauss_volume<0>(C, dummy, W);
    const size_t nc = g.nc, nf = g.nf;
    g.vol[idx] = W[0];
    g.xyzcc[idx] = W[1] / W[0];
    g.xyzcc[nc + idx] = W[2] / W[0];
    g.xyzcc[2 * nc + idx] = W[3] / W[0];
    int f = fiw(g, i, j, k);
    double fA = fmax(g.face[fiw(g, i + 1, j, k)], g.face[f]);
    g.di[idx] = W[0] / fA;
    fA = fmax(g.face[nf + fiw(g, i, j + 1, k)], g.face[nf + f]);
    g.dj[idx] = W[0] / fA;
    fA = fmax(g.face[2 * nf + fiw(g, i, j, k + 1)], g.face[2 * nf + f]);
    g.dk[idx] = W[0] / fA;
}

// init.F90:978-1048 edge lengths and edge triads for direction d
__global__ void __launch_bounds__(256) k_edges(GVw g, int d) {
    int is = 0, ie = g.Ni, js = 0, je = g.Nj, ks = 0, ke = g.Nk;   // inclusive, F extents
    if (d == 0) { ie = g.Ni - 1; ks = 1; ke = g.Nk - 1; }
    else if (d == 1) { je = g.Nj - 1; is = 1; ie = g.Ni - 1; }
    else { ke = g.Nk - 1; js = 1; je = g.Nj - 1; }
    const int nx = ie - is + 1, ny = je - js + 1, nz = ke - ks + 1;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny * nz) return;
    int i = is + idx % nx, r = idx / nx, j = js + r % ny, k = ks + r / ny;
    double p0[3], p1[3], a[3], b[3], c[3], e[3];
    xyzv(g, i, j, k, p0);
    if (d == 0) {
        xyzv(g, i + 1, j, k, p1);
        xyzv(g, i, j, k + 1, a); xyzv(g, i, j, k - 1, b); xyzv(g, i + 1, j, k + 1, c); xyzv(g, i + 1, j, k - 1, e);
    } else if (d == 1) {
        xyzv(g, i, j + 1, k, p1);
        xyzv(g, i + 1, j, k, a); xyzv(g, i - 1, j, k, b); xyzv(g, i + 1, j + 1, k, c); xyzv(g, i - 1, j + 1, k, e);
    } else {
        xyzv(g, i, j, k + 1, p1);
        xyzv(g, i, j + 1, k, a); xyzv(g, i, j - 1, k, b); xyzv(g, i, j + 1, k + 1, c); xyzv(g, i, j - 1, k + 1, e);
    }
    double dE[3], eAvg[3];
    for (int x = 0; x < 3; x++) { dE[x] = p1[x] - p0[x]; eAvg[x] = 0.5 * (a[x] - b[x] + c[x] - e[x]); }
    double len = nrm3(dE), eN[3], eT1[3], eT2[3], cr[3];
    for (int x = 0; x < 3; x++) eN[x] = dE[x] / len;
    cross3(eAvg, eN, cr);
    double n1 = nrm3(cr);
    for (int x = 0; x < 3; x++) eT1[x] = cr[x] / n1;
    cross3(eN, eT1, cr);
    double n2 = nrm3(cr);
    for (int x = 0; x < 3; x++) eT2[x] = cr[x] / n2;
    const size_t nf = g.nf;
    int f = fiw(g, i, j, k);
    g.edge[nf * d + f] = len;
    for (int x = 0; x < 3; x++) {
        g.Te[nf * (NORMX + x + 9 * d) + f] = eN[x];
        g.Te[nf * (TAN1X + x + 9 * d) + f] = eT1[x];
        g.Te[nf * (TAN2X + x + 9 * d) + f] = eT2[x];
    }
}

// init.F90:1083-1145 ebGeom for edge dir dN with (dT1, dT2) into slots slot, slot+1
__global__ void __launch_bounds__(256) k_teb(GVw g, int dN, int dT1, int dT2, int slot) {
    const int nx = g.Nip + 1, ny = g.Njp + 1, nz = g.Nkp + 1;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny * nz) return;
    int i = g.nG + idx % nx, r = idx / nx, j = g.nG + r % ny, k = g.nG + r / ny;
    const size_t nf = g.nf;
    const int st = dT2 == 0 ? 1 : (dT2 == 1 ? g.Ni + 1 : (g.Ni + 1) * (g.Nj + 1));
    int f = fiw(g, i, j, k);
    double fArB[8], Nx[3][8];
    for (int q = 0; q < 8; q++) {
        int ff = f + (q - 4) * st;
        fArB[q] = g.face[nf * dT1 + ff];
        for (int x = 0; x < 3; x++) Nx[x][q] = g.Tf[nf * (NORMX + x + 9 * dT1) + ff];
    }
    double fAi = 0.0;
    for (int q = 0; q < 8; q++) fAi += cW6[q] * fArB[q];
    double Nec[3];
    for (int x = 0; x < 3; x++) {
        double s = 0.0;
        for (int q = 0; q < 8; q++) s += (Nx[x][q] * fArB[q]) * cW6[q];
        Nec[x] = s / fAi;
    }
    double t1[3], t2[3];
    for (int x = 0; x < 3; x++) {
        t1[x] = g.Te[nf * (TAN1X + x + 9 * dN) + f];
        t2[x] = g.Te[nf * (TAN2X + x + 9 * dN) + f];
    }
    double dxT = Nec[0] * t1[0] + Nec[1] * t1[1] + Nec[2] * t1[2];
    double dyT = Nec[0] * t2[0] + Nec[1] * t2[1] + Nec[2] * t2[2];
    double dScl = sqrt(sq(dxT) + sq(dyT));
    if (dScl <= GAM_TINY) dScl = 1.0;
    g.Teb[nf * (slot + 4 * dN) + f] = dxT / dScl;
    g.Teb[nf * (slot + 1 + 4 * dN) + f] = dyT / dScl;
}

// ------------------------------------------------------------ initial conditions
// prob.F90:74-119 initOT2D over is..ie+1 (cells and faces)
__global__ void k_ic_ot(GV g, double gamma, double rho0, double p0, double b0, double kscale, double *Gas, double *mag) {
    const int nx = g.Nip + 1, ny = g.Njp + 1, nz = g.Nkp + 1;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny * nz) return;
    int i = g.nG + idx % nx, r = idx / nx, j = g.nG + r % ny, k = g.nG + r / ny;
    const size_t nc = g.nc, nf = g.nf;
    int c = ci(g, i, j, k), f = fi(g, i, j, k);
    double xc = g.xyzcc[c], yc = g.xyzcc[nc + c];
    Gas[c] = rho0;
    double Vx = -sin(kscale * yc), Vy = sin(kscale * xc), Vz = 0.0;
    double KinE = 0.5 * rho0 * (sq(Vx) + sq(Vy) + sq(Vz));
    Gas[nc * 1 + c] = Vx * rho0;
    Gas[nc * 2 + c] = Vy * rho0;
    Gas[nc * 3 + c] = Vz * rho0;
    Gas[nc * 4 + c] = KinE + p0 / (gamma - 1);
    mag[f] = (-b0 * sin(kscale * yc)) * g.face[f];
    mag[nf + f] = (b0 * sin((2 * kscale) * xc)) * g.face[nf + f];
    mag[2 * nf + f] = 0.0;
}

// gridutils.F90:144-189 GasIC2State with the blast-wave integrand (active cells)
__global__ void __launch_bounds__(256) k_ic_bw_gas(GVw g, double gamma, BWp bw, double *Gas) {
    const int nx = g.Nip, ny = g.Njp, nz = g.Nkp;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny * nz) return;
    int i = g.nG + idx % nx, r = idx / nx, j = g.nG + r % ny, k = g.nG + r / ny;
    double C[8][3], W[5];
    cell_coords(g, i, j, k, C);
    int c = ciw(g, i, j, k);
    double dV = g.vol[c];
    gauss_volume<1>(C, bw, W);
    for (int v = 0; v < 5; v++) W[v] = W[v] / dV;
    double D = W[0], Mx = D * W[1], My = D * W[2], Mz = D * W[3], P = W[4];
    double KinE = 0.5 * (sq(Mx) + sq(My) + sq(Mz)) / D;
    double IntE = P / (gamma - 1.0);
    const size_t nc = g.nc;
    Gas[c] = D; Gas[nc + c] = Mx; Gas[2 * nc + c] = My; Gas[3 * nc + c] = Mz;
    Gas[4 * nc + c] = IntE + KinE;
}

// gridutils.F90:205-223 edge integrals of A (quadrature.F90:345-370), over isg..ieg-1
__global__ void __launch_bounds__(256) k_ic_bw_la(GVw g, double bScl, double bSclz, double *lA) {
    const int nx = g.Ni - 1, ny = g.Nj - 1, nz = g.Nk - 1;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny * nz) return;
    int i = idx % nx, r = idx / nx, j = r % ny, k = r / ny;
    double e0[3], eX[3][3];
    xyzv(g, i, j, k, e0);
    xyzv(g, i + 1, j, k, eX[0]); xyzv(g, i, j + 1, k, eX[1]); xyzv(g, i, j, k + 1, eX[2]);
    int c = ciw(g, i, j, k);
    for (int d = 0; d < 3; d++) {
        const double *e1 = eX[d], *e2 = e0;
        double dR2[3], xBar[3], eInt[3] = {0, 0, 0};
        for (int x = 0; x < 3; x++) { dR2[x] = 0.5 * (e2[x] - e1[x]); xBar[x] = 0.5 * (e2[x] + e1[x]); }
        for (int q = 0; q < 12; q++) {
            double xp[3];
            for (int x = 0; x < 3; x++) xp[x] = xBar[x] + gA[q] * dR2[x];
            double Ap[3] = {0.0, bSclz * xp[0], bScl * (xp[1] - xp[0])};   // prob.F90:271-278
            for (int x = 0; x < 3; x++) eInt[x] = eInt[x] + 0.5 * gW[q] * Ap[x];
        }
        double dl[3] = {e1[0] - e2[0], e1[1] - e2[1], e1[2] - e2[2]};
        lA[(size_t)g.nc * d + c] = eInt[0] * dl[0] + eInt[1] * dl[1] + eInt[2] * dl[2];
    }
}

// gridutils.F90:226-235 curl of the edge integrals
__global__ void __launch_bounds__(256) k_ic_bw_flux(GVw g, const double *lA, double *mag) {
    const int nx = g.Ni - 1, ny = g.Nj - 1, nz = g.Nk - 1;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny * nz) return;
    int i = idx % nx, r = idx / nx, j = r % ny, k = r / ny;
    const size_t nc = g.nc, nf = g.nf;
    const double *LI = lA, *LJ = lA + nc, *LK = lA + 2 * nc;
    int c = ciw(g, i, j, k), f = fiw(g, i, j, k);
    mag[f] = LJ[c] - LJ[ciw(g, i, j, k + 1)] + LK[ciw(g, i, j + 1, k)] - LK[c];
    mag[nf + f] = -(LI[c] - LI[ciw(g, i, j, k + 1)] + LK[ciw(g, i + 1, j, k)] - LK[c]);
    mag[2 * nf + f] = LI[c] - LI[ciw(g, i, j + 1, k)] + LJ[ciw(g, i + 1, j, k)] - LJ[c];
}

// ----------------------------------------------------------------- background field
// background.F90:20-167 AddB0 for B0(x) = bu + dipole(x - xc), moment (0,0,M0)
// (dipolehelper.F90:19-30 MagsphereDipole; rad**5 and rad**3 as products).
struct B0P { double M0, xc[3], bu[3]; };

__device__ __forceinline__ void b0_eval(const B0P &p, double x, double y, double z, double *B) {
    double r[3] = {x - p.xc[0], y - p.xc[1], z - p.xc[2]};
    B[0] = p.bu[0]; B[1] = p.bu[1]; B[2] = p.bu[2];
    if (p.M0 != 0.0) {
        double rad = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        double r3 = rad * rad * rad, r5 = r3 * rad * rad;
        double mdr = p.M0 * r[2];
        double m3[3] = {0.0, 0.0, p.M0};
        for (int x = 0; x < 3; x++) B[x] = B[x] + (3 * mdr * r[x] / r5 - m3[x] / r3);
    }
}

// Per cell of the ghost-inclusive box, the three faces at its lower corner:
// B0 flux (GaussianFaceFlux), face-averaged B0 (GaussianFaceIntegral / Face)
// and the face Maxwell stress of B0 (GaussianFaceStress, into fS), one sweep
// of the 12x12 Gauss points (quadrature.F90:105-280).
__global__ void __launch_bounds__(256) k_b0_faces(GVw g, B0P p, double *fcB0, double *bFlux0, double *fS) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i = idx % g.Ni, r = idx / g.Ni, j = r % g.Nj, k = r / g.Nj;
    const size_t nc = g.nc, nf = g.nf;
    const int f = fiw(g, i, j, k);
    for (int d = 0; d < 3; d++) {
        double f0[3], f1[3], f2[3], f3[3];   // metric.F90:79-108 faceCoords
        xyzv(g, i, j, k, f0);
        if (d == 0) { xyzv(g, i, j + 1, k, f1); xyzv(g, i, j, k + 1, f2); xyzv(g, i, j + 1, k + 1, f3); }
        else if (d == 1) { xyzv(g, i, j, k + 1, f1); xyzv(g, i + 1, j, k, f2); xyzv(g, i + 1, j, k + 1, f3); }
        else { xyzv(g, i + 1, j, k, f1); xyzv(g, i, j + 1, k, f2); xyzv(g, i + 1, j + 1, k, f3); }
        double dx[3], dy[3], dz[3];
        dx[0] = f1[0] - f0[0]; dx[1] = f2[0] - f0[0]; dx[2] = f3[0] + f0[0] - f2[0] - f1[0];
        dy[0] = f1[1] - f0[1]; dy[1] = f2[1] - f0[1]; dy[2] = f3[1] + f0[1] - f2[1] - f1[1];
        dz[0] = f1[2] - f0[2]; dz[1] = f2[2] - f0[2]; dz[2] = f3[2] + f0[2] - f2[2] - f1[2];
        double flx = 0.0, fInt[3] = {0, 0, 0}, Mbb[3] = {0, 0, 0};
        for (int a = 0; a < 12; a++)
            for (int b = 0; b < 12; b++) {
                double eta = 0.5 * (1 + gA[a]), psi = 0.5 * (1 + gA[b]);
                double w = gW[a] * gW[b];
                double xp = f0[0] + dx[0] * eta + dx[1] * psi + dx[2] * eta * psi;
                double yp = f0[1] + dy[0] * eta + dy[1] * psi + dy[2] * eta * psi;
                double zp = f0[2] + dz[0] * eta + dz[1] * psi + dz[2] * eta * psi;
                double A[3];
                b0_eval(p, xp, yp, zp, A);
                double eV[3] = {dx[0] + dx[2] * psi, dy[0] + dy[2] * psi, dz[0] + dz[2] * psi};
                double pV[3] = {dx[1] + dx[2] * eta, dy[1] + dy[2] * eta, dz[1] + dz[2] * eta};
                double cr[3];
                cross3(eV, pV, cr);
                double N[3] = {0.25 * w * cr[0], 0.25 * w * cr[1], 0.25 * w * cr[2]};
                flx = flx + (N[0] * A[0] + N[1] * A[1] + N[2] * A[2]);
                double area = 0.25 * nrm3(cr) * w;
                for (int x = 0; x < 3; x++) fInt[x] = fInt[x] + area * A[x];
                double BB = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
                double BdS = A[0] * N[0] + A[1] * N[1] + A[2] * N[2];
                for (int x = 0; x < 3; x++) Mbb[x] = Mbb[x] + 0.5 * BB * N[x] - A[x] * BdS;
            }
        bFlux0[nc * d + idx] = flx;
        for (int x = 0; x < 3; x++) fcB0[nc * (x + 3 * d) + idx] = fInt[x] / g.face[nf * d + f];
        for (int x = 0; x < 3; x++) fS[nc * (d + 3 * x) + idx] = Mbb[x];
    }
}

// Per cell, the three edges at its lower corner: the mean of B0 along the edge
// (GaussianEdgeIntegral, quadrature.F90:345-370) on the edge's (T1, T2)
__global__ void __launch_bounds__(256) k_b0_edges(GVw g, B0P p, double *edgB0) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i = idx % g.Ni, r = idx / g.Ni, j = r % g.Nj, k = r / g.Nj;
    const size_t nc = g.nc, nf = g.nf;
    const int f = fiw(g, i, j, k);
    for (int d = 0; d < 3; d++) {
        double e1[3], e2[3];   // metric.F90:111-130 edgeCoords
        xyzv(g, i, j, k, e1);
        if (d == 0) xyzv(g, i + 1, j, k, e2);
        else if (d == 1) xyzv(g, i, j + 1, k, e2);
        else xyzv(g, i, j, k + 1, e2);
        double dR2[3], xBar[3], eInt[3];
        for (int x = 0; x < 3; x++) { dR2[x] = 0.5 * (e2[x] - e1[x]); xBar[x] = 0.5 * (e2[x] + e1[x]); eInt[x] = 0.0; }
        for (int a = 0; a < 12; a++) {
            double A[3];
            b0_eval(p, xBar[0] + gA[a] * dR2[0], xBar[1] + gA[a] * dR2[1], xBar[2] + gA[a] * dR2[2], A);
            for (int x = 0; x < 3; x++) eInt[x] = eInt[x] + 0.5 * gW[a] * A[x];
        }
        const double *Te = g.Te + nf * 9 * d;
        edgB0[nc * (0 + 2 * d) + idx] = eInt[0] * Te[nf * TAN1X + f] + eInt[1] * Te[nf * TAN1Y + f] + eInt[2] * Te[nf * TAN1Z + f];
        edgB0[nc * (1 + 2 * d) + idx] = eInt[0] * Te[nf * TAN2X + f] + eInt[1] * Te[nf * TAN2Y + f] + eInt[2] * Te[nf * TAN2Z + f];
    }
}

// Cell-averaged B0 (GaussianVolumeIntegral / volume)
__global__ void __launch_bounds__(256) k_b0_cells(GVw g, B0P p, double *B0) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nc) return;
    int i = idx % g.Ni, r = idx / g.Ni, j = r % g.Nj, k = r / g.Nj;
    double C[8][3], W[5];
    cell_coords(g, i, j, k, C);
    gauss_volume_fn(C, [&](double x, double y, double z, double *Wp) {   // background.F90:157-165 B0Wrap
        double B[3];
        b0_eval(p, x, y, z, B);
        Wp[0] = B[0]; Wp[1] = B[1]; Wp[2] = B[2]; Wp[3] = 0.0; Wp[4] = 0.0;
    }, W);
    for (int x = 0; x < 3; x++) B0[(size_t)g.nc * x + idx] = W[x] / g.vol[idx];
}

// background.F90:137-150: -div(face stress)/volume on the active cells
__global__ void __launch_bounds__(256) k_b0_dp(GVw g, const double *fS, double *dpB0) {
    const int nx = g.Nip, ny = g.Njp, nz = g.Nkp;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nx * ny * nz) return;
    int i = g.nG + idx % nx, r = idx / nx, j = g.nG + r % ny, k = g.nG + r / ny;
    const size_t nc = g.nc;
    const int c = ciw(g, i, j, k), ci1 = ciw(g, i + 1, j, k), cj1 = ciw(g, i, j + 1, k), ck1 = ciw(g, i, j, k + 1);
    for (int x = 0; x < 3; x++)
        dpB0[nc * x + c] = -(fS[nc * (0 + 3 * x) + ci1] - fS[nc * (0 + 3 * x) + c]
                           + fS[nc * (1 + 3 * x) + cj1] - fS[nc * (1 + 3 * x) + c]
                           + fS[nc * (2 + 3 * x) + ck1] - fS[nc * (2 + 3 * x) + c]) / g.vol[c];
}
