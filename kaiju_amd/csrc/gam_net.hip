// gam_net.hip -- the learned coarse-grid correction: a 3-layer 3x3x3 conv net
// (8 -> 32 -> 32 -> 8 channels, leaky ReLU between) on the ghost-inclusive cell grid,
// with BF16 weights and activations and FP32 accumulation on gfx950's
// v_mfma_f32_16x16x32_bf16.  Forward, backward (data and weights) and Adam are
// hand-written here; libgamera_net.so exports them (include/gamera_net.h).
//
// Layouts (nc = cells of the ghost-inclusive box, p = linear cell index, i fastest):
//   activations  act[p][32]   BF16, channels last (A operand of the convolutions)
//                actT[c][p]   BF16, channel-major copy (operand of the weight gradients),
//                             rows of ncs >= nc elements (a multiple of 8, with a gap),
//                             NET_PAD elements of zeros before row 0 and after row 31
//   weights      W[l][tap][co][ci] BF16 (ci = 32 padded), tap = dz*9 + dy*3 + dx;
//                the first layer (8 input channels) packs 4 taps per K = 32:
//                W0[g][co][4*8] with taps 4g..4g+3 (zero past tap 26)
//   output       out[p][8]    FP32 (the correction field, channels last)
// A convolution is an implicit GEMM: M = 16 consecutive cells per MFMA, N = 16
// output channels per tile, K = 32 input channels of one tap per MFMA (of four
// taps in the first layer).  Output
// cells cover [lo, lo + 16*ntile) with lo = Ni*Nj + Ni + 1, so every tap's
// neighbour p + off stays inside the array; cells near the box faces read
// wrapped neighbours and are never used (the correction is read on the active
// box plus one layer only, 4 ghost layers deep).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <cstring>
#include <cmath>

#include "gamera_net.h"

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;

#define NET_C 32      // padded channel count of every activation
#define NET_TAPS 27
#define NET_G0 7      // K steps of the first layer: ceil(27 taps / 4)
#define NET_PAD 64    // zero elements before and after a channel-major array
#define NET_MT 4      // 16-cell tiles per wave in the convolutions

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__host__ __device__ inline u16 f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (u16)(r >> 16);
}
__host__ __device__ inline float bf2f(u16 b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

struct NetDims { int Ni, Nj, Nk, nc, ncs, lo, ntile; int off[NET_TAPS]; };

__device__ __forceinline__ bf16x8 ld8(const u16 *p) {   // 16-byte aligned
    return *reinterpret_cast<const bf16x8 *>(p);
}

#define NET_LEAK 0.125f

// ----------------------------------------------------------------- convolution
// Hidden activation: leaky ReLU, x > 0 ? x : x/8 (a slope that keeps every
// unit trainable; 1/8 is exact in BF16 and FP32).
// MODE 0: activation, write act_out (channels last) and actT_out (channel-major), NCO = 32
// MODE 1: linear output, write out[p][0..7] FP32 (NCO = 16, channels 8..15 padding)
// MODE 2: backward data: multiply by the activation's slope at the forward
//         output (mask act > 0: 1, else 1/8), write both BF16 layouts, NCO = 32
// TPK taps per K step: 1 (K = 32 channels of one tap) or 4 (the first layer:
// K = 8 channels of 4 taps, lane group kg reading tap 4s + kg).
// One wave = NET_MT tiles of 16 cells x NCO channels (each K step's weights
// loaded once for all its tiles); blockDim 256 = 4 waves.
template <int NCO, int MODE, int TPK>
__global__ void __launch_bounds__(256) k_conv(NetDims d, const u16 *__restrict__ in, const u16 *__restrict__ W,
                                              const float *__restrict__ bias, const u16 *__restrict__ mask,
                                              u16 *__restrict__ act_out, u16 *__restrict__ actT_out,
                                              float *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int tile0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * NET_MT;   // this wave's NET_MT consecutive tiles
    if (tile0 >= d.ntile) return;                      // whole waves exit together
    const int row = lane & 15, kg = lane >> 4;         // A row (cell), K group of 8 channels
    f32x4 acc[NET_MT][NCO / 16];
#pragma unroll
    for (int m = 0; m < NET_MT; m++)
#pragma unroll
        for (int t = 0; t < NCO / 16; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int NSTEP = (NET_TAPS + TPK - 1) / TPK;
    for (int st = 0; st < NSTEP; st++) {
        const int tap = TPK == 1 ? st : TPK * st + kg;
        bf16x8 b[NCO / 16];                            // this K step's weights, used by every tile of the wave
#pragma unroll
        for (int t = 0; t < NCO / 16; t++) b[t] = ld8(W + ((size_t)st * NCO + 16 * t + row) * NET_C + 8 * kg);
        const bool live = TPK == 1 || tap < NET_TAPS;  // K slots past the last tap: weights are zero there too
        const int q0 = d.lo + 16 * tile0 + row + (live ? d.off[tap] : 0);
#pragma unroll
        for (int m = 0; m < NET_MT; m++) {
            if (tile0 + m >= d.ntile) break;           // uniform in the wave
            const bf16x8 a = live ? ld8(in + (size_t)(q0 + 16 * m) * NET_C + (TPK == 1 ? 8 * kg : 0)) : bf16x8{};
#pragma unroll
            for (int t = 0; t < NCO / 16; t++) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[t], acc[m][t], 0, 0, 0);
        }
    }
    // D layout: lane holds rows 4*kg + r (cells), column row (channel)
#pragma unroll
    for (int m = 0; m < NET_MT; m++) {
        if (tile0 + m >= d.ntile) break;
        const int p0 = d.lo + 16 * (tile0 + m);
#pragma unroll
        for (int t = 0; t < NCO / 16; t++) {
            const int co = 16 * t + row;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int p = p0 + 4 * kg + r;
                float v = acc[m][t][r];
                if (MODE == 0) {
                    v = v + bias[co];
                    v = v > 0.f ? v : NET_LEAK * v;
                    const u16 bb = f2bf(v);
                    act_out[(size_t)p * NET_C + co] = bb;
                    actT_out[(size_t)co * d.ncs + p] = bb;
                } else if (MODE == 1) {
                    if (co < 8) out[(size_t)p * 8 + co] = v + bias[co];
                } else {
                    if (bf2f(mask[(size_t)p * NET_C + co]) <= 0.f) v = NET_LEAK * v;   // slope at the forward output
                    const u16 bb = f2bf(v);
                    act_out[(size_t)p * NET_C + co] = bb;
                    actT_out[(size_t)co * d.ncs + p] = bb;
                }
            }
        }
    }
}

// 8 BF16 at element offset S (0..7) of the 16 aligned ones in w[0..7] (dwords)
template <int S> __device__ __forceinline__ bf16x8 take8(const uint32_t *w) {
    uint32_t r[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
        r[j] = (S % 2 == 0) ? w[j + S / 2] : __builtin_amdgcn_alignbyte(w[j + S / 2 + 1], w[j + S / 2], 2);
    bf16x8 v;
    memcpy(&v, r, 16);
    return v;
}
__device__ __forceinline__ void ld16x2(const u16 *p, uint32_t *w) {   // 32 bytes from a 16-byte aligned address
    const uint4 a = *reinterpret_cast<const uint4 *>(p), b = *reinterpret_cast<const uint4 *>(p + 8);
    w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w; w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
}

// Weight gradient of one layer, dW[tap][co][ci] += sum_p g[p][co] * a[p + off][ci]:
// per tap a GEMM with M = output channels, N = input channels and K = cells.
// Block = 4 waves on 4 consecutive chunks of cells of one tap (blocks ordered
// tap fastest, so the 27 taps of a chunk run together and share its gradient
// rows in L2).  Per 32 cells a wave loads its A tiles (gradient, channel-major,
// 16 aligned bytes per lane: chunks start at multiples of 8 cells) and its B
// tiles (activation at the tap-shifted cells: 32 aligned bytes per lane, then
// the 8 wanted values taken at the wave-uniform misalignment S of the tap
// offset).  Cells outside the computed range carry a zero gradient (never
// written, zero from allocation), the arrays' zero padding covers the reads
// past either end.  The 4 waves' sums meet in LDS; one FP32 atomic per weight
// per block goes to the unpadded parameter layout [tap][nco][nci].
template <int NCO, int NT, int S>
__device__ __forceinline__ void wgrad_chunk(const NetDims &d, const u16 *__restrict__ gT, const u16 *__restrict__ aT,
                                            int c0, int cend, int off, f32x4 (&acc)[NCO / 16][NT]) {
    const int lane = threadIdx.x & 63, row = lane & 15, kg = lane >> 4;
    for (int c = c0; c < cend; c += 32) {
        const int pk = c + 8 * kg;                             // 8 cells of this lane group
        bf16x8 bv[NT];
#pragma unroll
        for (int n = 0; n < NT; n++) {                         // B[k = cell][n = ci] at cell pk + off
            uint32_t w[8];
            ld16x2(aT + (size_t)(16 * n + row) * d.ncs + pk + off - S, w);
            bv[n] = take8<S>(w);
        }
#pragma unroll
        for (int m = 0; m < NCO / 16; m++) {                   // A[m = co][k = cell]
            const bf16x8 av = ld8(gT + (size_t)(16 * m + row) * d.ncs + pk);
#pragma unroll
            for (int n = 0; n < NT; n++) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv[n], acc[m][n], 0, 0, 0);
        }
    }
}

template <int NCO, int NT>
__global__ void __launch_bounds__(256) k_wgrad(NetDims d, const u16 *__restrict__ gT, const u16 *__restrict__ aT,
                                               int chunk, int nchunk, int nco, int nci, float *__restrict__ dW) {
    __shared__ float red[4][NCO][16 * NT];
    const int tap = blockIdx.x % NET_TAPS;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, row = lane & 15, kg = lane >> 4;
    const int q = (blockIdx.x / NET_TAPS) * 4 + wave;          // this wave's chunk
    const int A0 = d.lo & ~7, A1 = d.lo + 16 * d.ntile;
    const int off = d.off[tap];
    f32x4 acc[NCO / 16][NT];
#pragma unroll
    for (int m = 0; m < NCO / 16; m++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (q < nchunk) {
        const int c0 = A0 + q * chunk, cend = min(c0 + chunk, A1);
        switch (off & 7) {   // the misalignment S of cell pk + off, pk a multiple of 8
        case 0: wgrad_chunk<NCO, NT, 0>(d, gT, aT, c0, cend, off, acc); break;
        case 1: wgrad_chunk<NCO, NT, 1>(d, gT, aT, c0, cend, off, acc); break;
        case 2: wgrad_chunk<NCO, NT, 2>(d, gT, aT, c0, cend, off, acc); break;
        case 3: wgrad_chunk<NCO, NT, 3>(d, gT, aT, c0, cend, off, acc); break;
        case 4: wgrad_chunk<NCO, NT, 4>(d, gT, aT, c0, cend, off, acc); break;
        case 5: wgrad_chunk<NCO, NT, 5>(d, gT, aT, c0, cend, off, acc); break;
        case 6: wgrad_chunk<NCO, NT, 6>(d, gT, aT, c0, cend, off, acc); break;
        default: wgrad_chunk<NCO, NT, 7>(d, gT, aT, c0, cend, off, acc); break;
        }
    }
    // D[i = co][j = ci]: lane holds co = 16m + 4kg + r, ci = 16n + row
#pragma unroll
    for (int m = 0; m < NCO / 16; m++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int r = 0; r < 4; r++) red[wave][16 * m + 4 * kg + r][16 * n + row] = acc[m][n][r];
    __syncthreads();
    for (int e = threadIdx.x; e < NCO * 16 * NT; e += 256) {
        const int co = e / (16 * NT), ci = e % (16 * NT);
        if (co < nco && ci < nci)
            atomicAdd(&dW[((size_t)tap * nco + co) * nci + ci], red[0][co][ci] + red[1][co][ci] + red[2][co][ci] + red[3][co][ci]);
    }
}

// bias gradient: db[co] += sum over the tile range of g[p][co]
template <int NCO>
__global__ void __launch_bounds__(256) k_bgrad(NetDims d, const u16 *__restrict__ gT, float *__restrict__ db) {
    __shared__ float red[256];
    const int co = blockIdx.y;
    const int n = 16 * d.ntile;
    float s = 0.f;
    for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += gridDim.x * 256) s += bf2f(gT[(size_t)co * d.ncs + d.lo + q]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&db[co], red[0]);
}

// dL/dout (FP32 [nc][8]) -> BF16 gradient of the last layer, both layouts (16
// padded channels); zero outside the computed cells, where the output is not defined
__global__ void __launch_bounds__(256) k_dout(NetDims d, const float *__restrict__ dout, u16 *__restrict__ g, u16 *__restrict__ gT) {
    const int nc = d.nc;
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nc) return;
    const bool in = p >= d.lo && p < d.lo + 16 * d.ntile;
    for (int c = 0; c < NET_C; c++) {
        const u16 b = (c < 8 && in) ? f2bf(dout[(size_t)p * 8 + c]) : (u16)0;
        g[(size_t)p * NET_C + c] = b;
        if (c < 16) gT[(size_t)c * d.ncs + p] = b;
    }
}

// state -> input features ((rho - rho0)/s_rho, v/s_v, (p - p0)/s_p, B/s_B) of cell p
// from the cell's (rho, rho v, E, B) rounded to FP32 in q[8]
__device__ __forceinline__ void features_calc(const float (&q)[8], float gamma, float4 sc, float2 off, float (&f)[8]) {
    float rho = q[0], mx = q[1], my = q[2], mz = q[3];
    float E = q[4];
    float r = fmaxf(rho, 1e-12f);
    float vx = mx / r, vy = my / r, vz = mz / r;
    float P = (gamma - 1.f) * (E - 0.5f * r * (vx * vx + vy * vy + vz * vz));
    f[0] = (rho - off.x) / sc.x; f[1] = vx / sc.y; f[2] = vy / sc.y; f[3] = vz / sc.y; f[4] = (P - off.y) / sc.z;
    f[5] = q[5] / sc.w; f[6] = q[6] / sc.w; f[7] = q[7] / sc.w;
}
template <typename R>
__device__ __forceinline__ void state_of(size_t p, size_t n, const R *__restrict__ Gas, const R *__restrict__ B, float (&q)[8]) {
#pragma unroll
    for (int v = 0; v < 5; v++) q[v] = (float)Gas[v * n + p];
#pragma unroll
    for (int v = 0; v < 3; v++) q[5 + v] = (float)B[v * n + p];
}
template <typename R>
__device__ __forceinline__ void features_of(size_t p, size_t n, const R *__restrict__ Gas, const R *__restrict__ B, float gamma,
                                            float4 sc, float2 off, float (&f)[8]) {
    float q[8];
    state_of<R>(p, n, Gas, B, q);
    features_calc(q, gamma, sc, off, f);
}

// the features, BF16, both layouts
template <typename R>
__global__ void __launch_bounds__(256) k_features(int nc, int ncs, const R *__restrict__ Gas, const R *__restrict__ B, float gamma,
                                                  float4 sc, float2 off, u16 *__restrict__ act, u16 *__restrict__ actT) {
    int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= nc) return;
    float f[8];
    features_of<R>(p, nc, Gas, B, gamma, sc, off, f);
    for (int c = 0; c < NET_C; c++) {
        const u16 b = c < 8 ? f2bf(f[c]) : (u16)0;
        act[(size_t)p * NET_C + c] = b;
        actT[(size_t)c * ncs + p] = b;
    }
}

// ----------------------------------------------------------------- inference
// One layer over a brick of NB_I x NB_J x NB_K output cells with its input (the
// brick and one cell around it) staged once in LDS: every input cell comes from
// L2/HBM about 2.5 times instead of once per tap, and the MFMA A operands are
// LDS reads.  The layer's outputs are stored on a box [lo, lo + n) per axis only
// (the last layer's box is the cells at least `halo` from the grid faces, each
// earlier layer's box one cell wider on every side, so every stored output reads
// stored inputs); staged cells off the grid are zeros and feed only outputs
// that are not stored.  The first layer (CIN = 8) computes the features of its
// staged cells from Gas and B itself (features_of, as k_features).  Per output
// cell the operands, their K order, the bias and the activation are k_conv's,
// so the stored values equal gam_net_forward's bit for bit.
// Wave w of the block computes the brick's k row w: NB_J tiles of 16 cells (one
// per j row).  The results go through LDS to leave as whole cells (64 bytes of
// BF16 activations, or 32 bytes of FP32 output).
#define NB_I 16
#ifndef NB_J
#define NB_J 4
#endif
#define NB_K 4
#define NB_LI (NB_I + 2)
#define NB_LJ (NB_J + 2)
#define NB_LK (NB_K + 2)
#define NB_LC (NB_LI * NB_LJ * NB_LK)   // staged cells
#define NB_OC (NB_I * NB_J * NB_K)      // output cells

struct InferBox { int lo[3], n[3], nb[3]; };

template <int CIN, int NCO, int MODE, typename R>
__global__ void __launch_bounds__(256) k_infer(NetDims d, InferBox bx, const u16 *__restrict__ in, const R *__restrict__ Gas,
                                               const R *__restrict__ Bc, float gamma, float4 sc, float2 offs,
                                               const u16 *__restrict__ W, const float *__restrict__ bias,
                                               u16 *__restrict__ act_out, float *__restrict__ out) {
    static_assert(CIN == 8 || CIN == 32, "8 or 32 input channels");
    static_assert(NB_I == 16 && NB_K == 4, "one 16-cell tile per row, one k row per wave");
    constexpr int TPK = CIN == 8 ? 4 : 1;
    constexpr int NSTEP = (NET_TAPS + TPK - 1) / TPK;
    constexpr int OB = MODE == 0 ? 64 : 32;                          // bytes per output cell
    // a staged cell of 32 channels takes 80 bytes: consecutive cells of an A
    // operand then start 5 x 16 bytes apart, so the lanes of a 16-byte LDS read
    // fall on distinct banks (64 bytes apart, every fourth lane would share them)
    constexpr int CS = CIN == 32 ? 40 : 8;                           // BF16 elements per staged cell
    constexpr int SB = NB_LC * CS * 2 > NB_OC * OB ? NB_LC * CS * 2 : NB_OC * OB;
    __shared__ uint4 smem[SB / 16];
    u16 *sA = reinterpret_cast<u16 *>(smem);
    const int b = blockIdx.x;
    const int bi = b % bx.nb[0], bj = (b / bx.nb[0]) % bx.nb[1], bk = b / (bx.nb[0] * bx.nb[1]);
    const int oi = bx.lo[0] + NB_I * bi, oj = bx.lo[1] + NB_J * bj, ok = bx.lo[2] + NB_K * bk;   // brick origin
    const size_t nc = d.nc;
    // stage the input: cell c of the staged box is (oi - 1, oj - 1, ok - 1) + (li, lj, lk).
    // Every load of the thread is issued before the first LDS store.
    auto cell_at = [&](int c, size_t &p) -> bool {
        const int li = c % NB_LI, lj = (c / NB_LI) % NB_LJ, lk = c / (NB_LI * NB_LJ);
        const int i = oi - 1 + li, j = oj - 1 + lj, k = ok - 1 + lk;
        p = (size_t)i + (size_t)d.Ni * ((size_t)j + (size_t)d.Nj * k);
        return i >= 0 && i < d.Ni && j >= 0 && j < d.Nj && k >= 0 && k < d.Nk;
    };
    if (CIN == 8) {
        constexpr int NIT = (NB_LC + 255) / 256;
        float q[NIT][8];
        bool ok_[NIT];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int c = threadIdx.x + 256 * it;
            size_t p = 0;
            ok_[it] = c < NB_LC && cell_at(c, p);
            if (ok_[it]) state_of<R>(p, nc, Gas, Bc, q[it]);
        }
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int c = threadIdx.x + 256 * it;
            if (c >= NB_LC) break;
            uint4 v = {0u, 0u, 0u, 0u};
            if (ok_[it]) {
                float f[8];
                features_calc(q[it], gamma, sc, offs, f);
                v.x = (uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
                v.y = (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
                v.z = (uint32_t)f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16);
                v.w = (uint32_t)f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16);
            }
            smem[c] = v;
        }
    } else {
        constexpr int NIT = (NB_LC * 4 + 255) / 256;   // 4 x 16 bytes per cell, consecutive threads along i
        uint4 v[NIT];
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int e = threadIdx.x + 256 * it;
            size_t p = 0;
            v[it] = uint4{0u, 0u, 0u, 0u};
            if (e < NB_LC * 4 && cell_at(e >> 2, p)) v[it] = *reinterpret_cast<const uint4 *>(in + p * NET_C + 8 * (e & 3));
        }
#pragma unroll
        for (int it = 0; it < NIT; it++) {
            const int e = threadIdx.x + 256 * it;
            if (e < NB_LC * 4) smem[(e >> 2) * (CS / 8) + (e & 3)] = v[it];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, row = lane & 15, kg = lane >> 4;
    f32x4 acc[NB_J][NCO / 16];
#pragma unroll
    for (int m = 0; m < NB_J; m++)
#pragma unroll
        for (int t = 0; t < NCO / 16; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // the K steps' weights (the same for every block, from L2) are loaded two
    // steps ahead of their MFMAs, so their latency overlaps the two steps before
    const u16 *wl = W + ((size_t)row) * NET_C + 8 * kg;
    bf16x8 wn[2][NCO / 16];
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int t = 0; t < NCO / 16; t++) wn[q][t] = ld8(wl + ((size_t)q * NCO + 16 * t) * NET_C);
    for (int st = 0; st < NSTEP; st++) {
        const int tap = TPK == 1 ? st : TPK * st + kg;
        const bool live = TPK == 1 || tap < NET_TAPS;   // K slots past the last tap: zero A (their weights are zero too)
        const int tp = live ? tap : 13;                 // a staged cell for the dead slots
        bf16x8 bw[NCO / 16];
        const int sn = st + 2 < NSTEP ? st + 2 : NSTEP - 1;   // the step whose weights are fetched now
#pragma unroll
        for (int t = 0; t < NCO / 16; t++) {
            bw[t] = wn[0][t];
            wn[0][t] = wn[1][t];
            wn[1][t] = ld8(wl + ((size_t)sn * NCO + 16 * t) * NET_C);
        }
        const int dx = tp % 3 - 1, dy = (tp / 3) % 3 - 1, dz = tp / 9 - 1;
        const int c0 = (wave + 1 + dz) * (NB_LJ * NB_LI) + (1 + dy) * NB_LI + 1 + dx + row;
#pragma unroll
        for (int m = 0; m < NB_J; m++) {
            bf16x8 a = *reinterpret_cast<const bf16x8 *>(sA + (size_t)(c0 + m * NB_LI) * CS + (TPK == 1 ? 8 * kg : 0));
            if (!live) a = bf16x8{};
#pragma unroll
            for (int t = 0; t < NCO / 16; t++) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[t], acc[m][t], 0, 0, 0);
        }
    }
    __syncthreads();   // every wave is done with the staged input: the LDS takes the outputs
    // D layout: lane holds cells 4*kg + r of the tile, channel 16*t + row
#pragma unroll
    for (int m = 0; m < NB_J; m++)
#pragma unroll
        for (int t = 0; t < NCO / 16; t++) {
            const int co = 16 * t + row;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int oc = (wave * NB_J + m) * NB_I + 4 * kg + r;
                const float v = acc[m][t][r] + bias[co < (MODE == 0 ? NCO : 8) ? co : 0];
                if (MODE == 0) {
                    sA[oc * NET_C + co] = f2bf(v > 0.f ? v : NET_LEAK * v);
                } else if (co < 8) {
                    reinterpret_cast<float *>(smem)[oc * 8 + co] = v;
                }
            }
        }
    __syncthreads();
    constexpr int CH = OB / 16;   // 16-byte chunks per output cell
    for (int e = threadIdx.x; e < NB_OC * CH; e += 256) {
        const int oc = e / CH, part = e % CH;
        const int ii = oc % NB_I, jj = (oc / NB_I) % NB_J, kk = oc / (NB_I * NB_J);
        if (NB_I * bi + ii >= bx.n[0] || NB_J * bj + jj >= bx.n[1] || NB_K * bk + kk >= bx.n[2]) continue;
        const size_t p = (size_t)(oi + ii) + (size_t)d.Ni * ((size_t)(oj + jj) + (size_t)d.Nj * (ok + kk));
        if (MODE == 0) *reinterpret_cast<uint4 *>(act_out + p * NET_C + 8 * part) = smem[e];
        else *reinterpret_cast<uint4 *>(out + p * 8 + 4 * part) = smem[e];
    }
}

// Adam on the FP32 master parameters (Kingma & Ba 2015, bias-corrected)
__global__ void __launch_bounds__(256) k_adam(int n, float *__restrict__ w, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                                              float eps, float c1, float c2) {
    int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const float gq = g[q];
    const float mq = b1 * m[q] + (1.f - b1) * gq;
    const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    w[q] = w[q] - lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

// master FP32 parameters -> BF16 forward and transposed/flipped backward weights
// layer l forward: Wf[tap][co][ci]; backward data: Wb[tap'][ci][co] with tap' the mirrored tap
__global__ void __launch_bounds__(256) k_pack(const float *__restrict__ w, int off, int nco_real, int nci_real, int nco_pad,
                                              u16 *__restrict__ Wf, u16 *__restrict__ Wb) {
    int q = blockIdx.x * 256 + threadIdx.x;
    const int total = NET_TAPS * nco_pad * NET_C;
    if (q >= total) return;
    const int ci = q % NET_C, co = (q / NET_C) % nco_pad, tap = q / (NET_C * nco_pad);
    const float val = (co < nco_real && ci < nci_real) ? w[off + (tap * nco_real + co) * nci_real + ci] : 0.f;
    Wf[q] = f2bf(val);
    // backward data kernel has NCO = 32 outputs (= ci) and 32 inputs (= co, padded)
    const int mt = NET_TAPS - 1 - tap;
    if (ci < NET_C && co < NET_C) Wb[((size_t)mt * NET_C + ci) * NET_C + co] = f2bf(val);
}

// the first layer's weights, 4 taps per K step: W0[g][co][8q + ci] = w[tap = 4g + q][co][ci]
__global__ void __launch_bounds__(256) k_pack0(const float *__restrict__ w, u16 *__restrict__ W0) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= NET_G0 * NET_C * NET_C) return;
    const int k = q % NET_C, co = (q / NET_C) % NET_C, g = q / (NET_C * NET_C);
    const int tap = 4 * g + k / 8, ci = k % 8;
    W0[q] = tap < NET_TAPS ? f2bf(w[(tap * NET_C + co) * 8 + ci]) : (u16)0;
}

// ----------------------------------------------------------------- host side
struct gam_net {
    NetDims d;
    hipStream_t stream = nullptr;        // where the net's work goes: its own stream or the caller's
    hipStream_t own = nullptr;           // the stream the net created
    hipEvent_t done = nullptr;           // recorded after each gam_net_infer
    // parameters: FP32 master, grads, Adam moments
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    u16 *Wf[3] = {}, *Wb[3] = {};
    u16 *act[4] = {}, *actT[4] = {};     // 0 features, 1..2 hidden, 3 scratch gradient
    u16 *gr[2] = {}, *grT[2] = {};       // gradient ping-pong
    u16 *actT0[4] = {}, *grT0[2] = {};   // allocations of the channel-major arrays (NET_PAD before row 0)
    float *out = nullptr;
    float *dbuf = nullptr;               // upstream gradient uploaded from the host
    float4 scales = {1.f, 1.f, 1.f, 1.f};
    float gamma = 5.f / 3.f;
    float2 offs = {0.f, 0.f};
    std::string err;
};

// parameter layout (FP32): layer 0 W[27][32][8], b[32]; layer 1 W[27][32][32], b[32]; layer 2 W[27][8][32], b[8]
static const int kCo[3] = {32, 32, 8}, kCi[3] = {8, 32, 32}, kCoPad[3] = {32, 32, 16};
static int woff(int l) { int o = 0; for (int q = 0; q < l; q++) o += NET_TAPS * kCo[q] * kCi[q] + kCo[q]; return o; }
static int boff(int l) { return woff(l) + NET_TAPS * kCo[l] * kCi[l]; }

extern "C" int gam_net_num_params(void) { return woff(3); }

static int nfail(gam_net *n, int code, const char *msg, hipError_t e) {
    if (n) n->err = std::string(msg) + ": " + hipGetErrorString(e);
    return code;
}
#define NCHK(n, x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return nfail(n, -2, #x, e_); } while (0)

extern "C" int gam_net_destroy(gam_net *n) {
    if (!n) return 0;
    if (n->own) (void)hipStreamSynchronize(n->own);   // work on a caller's stream: hipFree below waits for the device
    void *bufs[] = {n->w, n->g, n->m, n->v, n->out, n->dbuf};
    for (void *b : bufs) if (b) (void)hipFree(b);
    for (int l = 0; l < 3; l++) { (void)hipFree(n->Wf[l]); (void)hipFree(n->Wb[l]); }
    for (int q = 0; q < 4; q++) { (void)hipFree(n->act[q]); (void)hipFree(n->actT0[q]); }
    for (int q = 0; q < 2; q++) { (void)hipFree(n->gr[q]); (void)hipFree(n->grT0[q]); }
    if (n->done) (void)hipEventDestroy(n->done);
    if (n->own) (void)hipStreamDestroy(n->own);
    delete n;
    return 0;
}

extern "C" int gam_net_create(int Ni, int Nj, int Nk, gam_net **out) {
    if (!out || Ni < 8 || Nj < 8 || Nk < 8) return -1;   // the 3-layer stencil needs 3 cells around each used cell
    *out = nullptr;
    long nc = (long)Ni * Nj * Nk;
    if ((nc + 24) * NET_C + 2 * NET_PAD > 0x7fffffffL) return -1;
    gam_net *n = new gam_net();
    NetDims &d = n->d;
    d.Ni = Ni; d.Nj = Nj; d.Nk = Nk; d.nc = (int)nc;
    d.ncs = (int)(((nc + 7) / 8) * 8 + 16);   // channel-major row stride: 16-byte aligned rows, a gap between them
    const int O = Ni * Nj + Ni + 1;
    d.lo = O;
    d.ntile = (int)((nc - 2L * O) / 16);
    for (int t = 0; t < NET_TAPS; t++) {
        int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
        d.off[t] = dx + Ni * (dy + Nj * dz);
    }
    auto cleanup = [&](int c) { gam_net_destroy(n); return c; };
    if (hipStreamCreateWithFlags(&n->own, hipStreamNonBlocking) != hipSuccess) return cleanup(-2);
    n->stream = n->own;
    if (hipEventCreateWithFlags(&n->done, hipEventDisableTiming) != hipSuccess) return cleanup(-2);
    const int np = gam_net_num_params();
    for (float **p : {&n->w, &n->g, &n->m, &n->v}) {
        if (hipMalloc(p, np * sizeof(float)) != hipSuccess) return cleanup(-3);
        if (hipMemset(*p, 0, np * sizeof(float)) != hipSuccess) return cleanup(-2);
    }
    for (int l = 0; l < 3; l++) {
        if (hipMalloc(&n->Wf[l], (size_t)NET_TAPS * kCoPad[l] * NET_C * 2) != hipSuccess) return cleanup(-3);
        if (hipMalloc(&n->Wb[l], (size_t)NET_TAPS * NET_C * NET_C * 2) != hipSuccess) return cleanup(-3);
        if (hipMemset(n->Wb[l], 0, (size_t)NET_TAPS * NET_C * NET_C * 2) != hipSuccess) return cleanup(-2);
    }
    const size_t ab = (size_t)nc * NET_C * 2, abT = ((size_t)d.ncs * NET_C + 2 * NET_PAD) * 2;
    for (int q = 0; q < 4; q++) {
        if (hipMalloc(&n->act[q], ab) != hipSuccess || hipMalloc(&n->actT0[q], abT) != hipSuccess) return cleanup(-3);
        if (hipMemset(n->act[q], 0, ab) != hipSuccess || hipMemset(n->actT0[q], 0, abT) != hipSuccess) return cleanup(-2);
        n->actT[q] = n->actT0[q] + NET_PAD;
    }
    for (int q = 0; q < 2; q++) {
        if (hipMalloc(&n->gr[q], ab) != hipSuccess || hipMalloc(&n->grT0[q], abT) != hipSuccess) return cleanup(-3);
        if (hipMemset(n->gr[q], 0, ab) != hipSuccess || hipMemset(n->grT0[q], 0, abT) != hipSuccess) return cleanup(-2);
        n->grT[q] = n->grT0[q] + NET_PAD;
    }
    if (hipMalloc(&n->out, (size_t)nc * 8 * sizeof(float)) != hipSuccess) return cleanup(-3);
    if (hipMalloc(&n->dbuf, (size_t)nc * 8 * sizeof(float)) != hipSuccess) return cleanup(-3);
    if (hipMemset(n->out, 0, (size_t)nc * 8 * sizeof(float)) != hipSuccess) return cleanup(-2);
    *out = n;
    return 0;
}

extern "C" const char *gam_net_error(const gam_net *n) { return n ? n->err.c_str() : "null net"; }
extern "C" void *gam_net_stream(gam_net *n) { return n ? (void *)n->stream : nullptr; }

extern "C" int gam_net_set_stream(gam_net *n, void *stream) {
    if (!n) return -1;
    hipStream_t to = stream ? (hipStream_t)stream : n->own;
    if (to == n->stream) return 0;
    NCHK(n, hipStreamSynchronize(n->stream));   // the work already queued ends before any on the new stream
    n->stream = to;
    return 0;
}
extern "C" const float *gam_net_output(gam_net *n) { return n ? n->out : nullptr; }
extern "C" float *gam_net_grad_buffer(gam_net *n) { return n ? n->dbuf : nullptr; }

static int pack(gam_net *n) {
    hipLaunchKernelGGL(k_pack0, dim3((NET_G0 * NET_C * NET_C + 255) / 256), 256, 0, n->stream, (const float *)n->w + woff(0),
                       n->Wf[0]);
    for (int l = 1; l < 3; l++) {
        const int total = NET_TAPS * kCoPad[l] * NET_C;
        hipLaunchKernelGGL(k_pack, dim3((total + 255) / 256), 256, 0, n->stream, (const float *)n->w, woff(l), kCo[l], kCi[l],
                           kCoPad[l], n->Wf[l], n->Wb[l]);
    }
    NCHK(n, hipGetLastError());
    return 0;
}

extern "C" int gam_net_set_params(gam_net *n, const float *w) {
    if (!n || !w) return -1;
    NCHK(n, hipMemcpyAsync(n->w, w, gam_net_num_params() * sizeof(float), hipMemcpyHostToDevice, n->stream));
    int rc = pack(n);
    if (rc) return rc;
    NCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

extern "C" int gam_net_get_params(gam_net *n, float *w, float *grad) {
    if (!n) return -1;
    if (w) NCHK(n, hipMemcpyAsync(w, n->w, gam_net_num_params() * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    if (grad) NCHK(n, hipMemcpyAsync(grad, n->g, gam_net_num_params() * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

extern "C" int gam_net_set_scales(gam_net *n, float rho, float v, float p, float b, float gamma) {
    if (!n || !(rho > 0) || !(v > 0) || !(p > 0) || !(b > 0)) return -1;
    n->scales = make_float4(rho, v, p, b);
    n->gamma = gamma;
    return 0;
}

extern "C" int gam_net_set_offsets(gam_net *n, float rho0, float p0) {
    if (!n) return -1;
    n->offs = make_float2(rho0, p0);
    return 0;
}

extern "C" int gam_net_features(gam_net *n, const void *Gas, const void *B, int precision) {
    if (!n || !Gas || !B) return -1;
    const int nb = (n->d.nc + 255) / 256;
    if (precision) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_features<float>), dim3(nb), 256, 0, n->stream, n->d.nc, n->d.ncs, (const float *)Gas,
                                      (const float *)B, n->gamma, n->scales, n->offs, n->act[0], n->actT[0]);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_features<double>), dim3(nb), 256, 0, n->stream, n->d.nc, n->d.ncs, (const double *)Gas,
                            (const double *)B, n->gamma, n->scales, n->offs, n->act[0], n->actT[0]);
    NCHK(n, hipGetLastError());
    return 0;
}

extern "C" int gam_net_set_features(gam_net *n, const float *feat) {
    // host features [nc][8] FP32 (tests): rounded to BF16 like k_features
    if (!n || !feat) return -1;
    const size_t nc = n->d.nc, ncs = n->d.ncs;
    std::vector<u16> a(nc * NET_C, 0), aT(ncs * NET_C, 0);
    for (size_t p = 0; p < nc; p++)
        for (int c = 0; c < 8; c++) { u16 b = f2bf(feat[p * 8 + c]); a[p * NET_C + c] = b; aT[c * ncs + p] = b; }
    NCHK(n, hipMemcpyAsync(n->act[0], a.data(), a.size() * 2, hipMemcpyHostToDevice, n->stream));
    NCHK(n, hipMemcpyAsync(n->actT[0], aT.data(), aT.size() * 2, hipMemcpyHostToDevice, n->stream));
    NCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

static dim3 conv_grid(const gam_net *n) { return dim3((unsigned)((n->d.ntile + 4 * NET_MT - 1) / (4 * NET_MT))); }

extern "C" int gam_net_forward(gam_net *n) {
    if (!n) return -1;
    const NetDims &d = n->d;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_conv<32, 0, 4>), conv_grid(n), 256, 0, n->stream, d, (const u16 *)n->act[0], (const u16 *)n->Wf[0],
                       (const float *)(n->w + boff(0)), (const u16 *)nullptr, n->act[1], n->actT[1], (float *)nullptr);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_conv<32, 0, 1>), conv_grid(n), 256, 0, n->stream, d, (const u16 *)n->act[1], (const u16 *)n->Wf[1],
                       (const float *)(n->w + boff(1)), (const u16 *)nullptr, n->act[2], n->actT[2], (float *)nullptr);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_conv<16, 1, 1>), conv_grid(n), 256, 0, n->stream, d, (const u16 *)n->act[2], (const u16 *)n->Wf[2],
                       (const float *)(n->w + boff(2)), (const u16 *)nullptr, (u16 *)nullptr, (u16 *)nullptr, n->out);
    NCHK(n, hipGetLastError());
    return 0;
}

extern "C" int gam_net_infer(gam_net *n, const void *Gas, const void *B, int precision, int halo) {
    if (!n || !Gas || !B) return -1;
    const NetDims &d = n->d;
    const int N[3] = {d.Ni, d.Nj, d.Nk};
    if (halo < 3) { n->err = "gam_net_infer: halo must be >= 3 (three layers of 3-point stencils)"; return -1; }
    for (int a = 0; a < 3; a++)
        if (N[a] < 2 * halo + 1) { n->err = "gam_net_infer: the grid is thinner than 2*halo + 1 cells"; return -1; }
    InferBox bx[3];
    for (int l = 0; l < 3; l++) {   // layer l stores the cells at least halo - (2 - l) from the faces
        const int lo = halo - (2 - l);
        const int nb[3] = {NB_I, NB_J, NB_K};
        for (int a = 0; a < 3; a++) {
            bx[l].lo[a] = lo;
            bx[l].n[a] = N[a] - 2 * lo;
            bx[l].nb[a] = (bx[l].n[a] + nb[a] - 1) / nb[a];
        }
    }
    auto grid = [](const InferBox &b) { return dim3((unsigned)(b.nb[0] * b.nb[1] * b.nb[2])); };
    if (precision)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_infer<8, 32, 0, float>), grid(bx[0]), 256, 0, n->stream, d, bx[0], (const u16 *)nullptr,
                           (const float *)Gas, (const float *)B, n->gamma, n->scales, n->offs, (const u16 *)n->Wf[0],
                           (const float *)(n->w + boff(0)), n->act[1], (float *)nullptr);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_infer<8, 32, 0, double>), grid(bx[0]), 256, 0, n->stream, d, bx[0], (const u16 *)nullptr,
                           (const double *)Gas, (const double *)B, n->gamma, n->scales, n->offs, (const u16 *)n->Wf[0],
                           (const float *)(n->w + boff(0)), n->act[1], (float *)nullptr);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_infer<32, 32, 0, float>), grid(bx[1]), 256, 0, n->stream, d, bx[1], (const u16 *)n->act[1],
                       (const float *)nullptr, (const float *)nullptr, n->gamma, n->scales, n->offs, (const u16 *)n->Wf[1],
                       (const float *)(n->w + boff(1)), n->act[2], (float *)nullptr);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_infer<32, 16, 1, float>), grid(bx[2]), 256, 0, n->stream, d, bx[2], (const u16 *)n->act[2],
                       (const float *)nullptr, (const float *)nullptr, n->gamma, n->scales, n->offs, (const u16 *)n->Wf[2],
                       (const float *)(n->w + boff(2)), (u16 *)nullptr, n->out);
    NCHK(n, hipGetLastError());
    NCHK(n, hipEventRecord(n->done, n->stream));
    return 0;
}

extern "C" void *gam_net_infer_event(gam_net *n) { return n ? (void *)n->done : nullptr; }

// dout: device FP32 [nc][8] = dL/d(out).  Accumulates into the gradient buffer
// (gam_net_zero_grad clears it).  Uses the activations of the last forward.
extern "C" int gam_net_backward(gam_net *n, const float *dout) {
    if (!n || !dout) return -1;
    const NetDims &d = n->d;
    const int chunk = 8192;   // cells per wave, a multiple of 32
    const int nchunk = (d.lo + 16 * d.ntile - (d.lo & ~7) + chunk - 1) / chunk;
    const dim3 wg((unsigned)(NET_TAPS * ((nchunk + 3) / 4)));
    hipLaunchKernelGGL(k_dout, dim3((d.nc + 255) / 256), 256, 0, n->stream, d, dout, n->gr[0], n->grT[0]);
    // layer 2 (linear, 8 outputs padded to 16)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wgrad<16, 2>), wg, 256, 0, n->stream, d, (const u16 *)n->grT[0], (const u16 *)n->actT[2], chunk,
                       nchunk, kCo[2], kCi[2], n->g + woff(2));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgrad<16>), dim3(64, 8), 256, 0, n->stream, d, (const u16 *)n->grT[0], n->g + boff(2));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_conv<32, 2, 1>), conv_grid(n), 256, 0, n->stream, d, (const u16 *)n->gr[0], (const u16 *)n->Wb[2],
                       (const float *)nullptr, (const u16 *)n->act[2], n->gr[1], n->grT[1], (float *)nullptr);
    // layer 1
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wgrad<32, 2>), wg, 256, 0, n->stream, d, (const u16 *)n->grT[1], (const u16 *)n->actT[1], chunk,
                       nchunk, kCo[1], kCi[1], n->g + woff(1));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgrad<32>), dim3(64, 32), 256, 0, n->stream, d, (const u16 *)n->grT[1], n->g + boff(1));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_conv<32, 2, 1>), conv_grid(n), 256, 0, n->stream, d, (const u16 *)n->gr[1], (const u16 *)n->Wb[1],
                       (const float *)nullptr, (const u16 *)n->act[1], n->gr[0], n->grT[0], (float *)nullptr);
    // layer 0 (input gradient not needed)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wgrad<32, 1>), wg, 256, 0, n->stream, d, (const u16 *)n->grT[0], (const u16 *)n->actT[0], chunk,
                       nchunk, kCo[0], kCi[0], n->g + woff(0));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bgrad<32>), dim3(64, 32), 256, 0, n->stream, d, (const u16 *)n->grT[0], n->g + boff(0));
    NCHK(n, hipGetLastError());
    return 0;
}

extern "C" int gam_net_zero_grad(gam_net *n) {
    if (!n) return -1;
    NCHK(n, hipMemsetAsync(n->g, 0, gam_net_num_params() * sizeof(float), n->stream));
    return 0;
}

extern "C" int gam_net_adam(gam_net *n, float lr, float b1, float b2, float eps, int t) {
    if (!n || t < 1) return -1;
    const int np = gam_net_num_params();
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((np + 255) / 256), 256, 0, n->stream, np, n->w, (const float *)n->g, n->m, n->v, lr, b1, b2,
                       eps, c1, c2);
    NCHK(n, hipGetLastError());
    return pack(n);
}

extern "C" int gam_net_sync(gam_net *n) {
    if (!n) return -1;
    NCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

extern "C" int gam_net_get_output(gam_net *n, float *out) {
    if (!n || !out) return -1;
    NCHK(n, hipMemcpyAsync(out, n->out, (size_t)n->d.nc * 8 * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    NCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

// channels-last BF16 activations of the last forward (layer 0 = features), for tests
extern "C" int gam_net_get_act(gam_net *n, int layer, unsigned short *out) {
    if (!n || !out || layer < 0 || layer > 2) return -1;
    NCHK(n, hipMemcpyAsync(out, n->act[layer], (size_t)n->d.nc * NET_C * 2, hipMemcpyDeviceToHost, n->stream));
    NCHK(n, hipStreamSynchronize(n->stream));
    return 0;
}

extern "C" int gam_net_backward_host(gam_net *n, const float *dout) {
    if (!n || !dout) return -1;
    NCHK(n, hipMemcpyAsync(n->dbuf, dout, (size_t)n->d.nc * 8 * sizeof(float), hipMemcpyHostToDevice, n->stream));
    return gam_net_backward(n, n->dbuf);
}
