// gam_api.hip -- the C-ABI of libgamera_hip.so: handle, device memory, stream,
// launch sequencing and per-stage timers.  Declared in include/gamera_hip.h.
#include "gamera_hip.h"
#include "gam_common.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

// The kernels are part of this translation unit.
#include "gam_step.hip"
#include "gam_setup.hip"
#include "gam_ring.hip"

// ----------------------------------------------------------------- stages
enum Stage {
    ST_CLOCK = 0, ST_PREDICT, ST_BHALF, ST_MOM2FACE, ST_EMF_I, ST_EMF_J, ST_EMF_K, ST_FLUX_I, ST_FLUX_J,
    ST_FLUX_K, ST_CT, ST_UPDATE, ST_DTFINAL, ST_BC, ST_BPOST, ST_CELLDT, ST_RING, ST_N
};
static const char *kStageNames[ST_N] = {
    "clock", "predictor", "bflux_half", "mom2face", "emf_i", "emf_j", "emf_k", "flux_i", "flux_j",
    "flux_k", "ct_update", "update", "dt_final", "bcs", "bflux_post", "celldt", "ring"};

enum { G_XYZ = 0, G_XYZCC, G_XFC, G_FACE, G_TF, G_EDGE, G_TE, G_TEB, G_VOL, G_DI, G_DJ, G_DK, G_N };
// geometry the bulk kernels read, kept as an FP32 copy in precision 1
static const int kG32[9] = {G_FACE, G_TF, G_EDGE, G_TE, G_TEB, G_VOL, G_DI, G_DJ, G_DK};
enum { W_WH = 0, W_WN, W_VDV, W_BH, W_BN, W_N };

struct PendingEv { int stage; hipEvent_t a, b; };

struct gam_handle {
    gam_dims dims;
    gam_params par;
    GV gv;                      // FP64 view (setup, FP64 step)
    GVT<float> gv32;            // FP32 view of the bulk geometry (precision 1)
    MP mp;
    int prec = 0;               // 0 FP64, 1 FP32 bulk
    size_t es = 8;              // bytes per bulk element
    hipStream_t stream = nullptr;
    double *grid[G_N] = {};
    float *grid32[9] = {};
    float *b2fc = nullptr;               // FP32 bulk: bFlux2Fld lever arms (18 x G)
    // bulk-typed arrays (double or float by prec); magFlux is always FP64
    void *Gas[2] = {}, *B[2] = {};
    double *mag[2] = {};
    int s = 0;                 // index of State in the {Gas,mag,B} pairs
    void *hGas = nullptr, *hmag = nullptr, *hB = nullptr;
    void *Vf = nullptr, *E = nullptr, *Flx = nullptr;
    void *wk[W_N] = {};        // predictor-pass stencil quantities
    double *bg[5] = {};        // background field fcB0, edgB0, B0, bFlux0, dpB0 (G-sized)
    const float *corr = nullptr; // learned correction field [G][8] (device), or none
    double corr_scale = 0.0;
    hipEvent_t corr_wait = nullptr;    // the step waits for it before the first kernel that reads corr
    int bct[6] = {0, 0, 0, 0, 0, 0};   // boundary type per face INI, OUTI, INJ, OUTJ, INK, OUTK (GAM_BC_*)
    int docpr = 0;                     // params doCPR: StateCPR after a bad CalcDT
    RingP ring = {};                   // ring average about the LFM axis (gam_set_ring)
    double *ringSE = nullptr;          // its smoothing EMF (G x 3, zero off the ring rows)
    double *ringTab = nullptr;         // cos, sin of the ring angles phi_n
    double *ringMean = nullptr;        // ring-average scratch (chunk means, Fourier coefficients)
    size_t ringMean_n = 0;             // its capacity in doubles
    int fused_np = 0;                  // block minima the fused step's CalcDT reduces
    WindP wind = {};                   // the solar wind of a GAM_BC_WIND outer i face (gam_set_wind*)
    double *windSer = nullptr;         // its time series (nt x WIND_COLS), wind.ser points here
    bool corners_set = false;          // geometry derived by gam_set_corners (RingGridFix applied then)
    double *ctG = nullptr, *ctM = nullptr, *crG = nullptr, *crM = nullptr;   // correction training buffers
    BGV bgv = {};              // device view of bg; bgv.on = 0 until a background is set
    double *partial = nullptr;
    Tile tile[ST_N] = {};       // launch tiles of the tiled stages
    int npart = 0;             // blocks of the 1-D CellDT launch
    int pcap = 0;              // capacity of partial (block minima / sums)
    DClock *clk = nullptr;      // device
    DClock *clk_host = nullptr; // pinned
    std::string err;
    bool timing = false;
    std::vector<PendingEv> pending;
    std::vector<hipEvent_t> evpool;
    double ms[ST_N] = {};
    int64_t launches[ST_N] = {};
    int graph_refs = 0;        // captured graphs holding this handle's buffers and launch parameters
    int device = 0;            // the GPU current at gam_create (its stream's device)
    double *scratch = nullptr; // that GPU's copy of the file-scope scratch array (gam_divb_noalloc)
    std::vector<std::unique_ptr<struct PackRec>> *rec = nullptr;   // set: launches are recorded, not run
};

// Calls that would change what a captured graph launches (kernels, their
// by-value parameters, or buffers it baked in) are refused while one holds h.
static int refuse_if_captured(gam_handle *h, const char *what) {
    if (h && h->graph_refs > 0) {
        h->err = std::string(what) + ": a captured graph holds this handle (gam_graph_destroy it first)";
        return -1;
    }
    return 0;
}

static int fail(gam_handle *h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

#define HIPCHK(h, x)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return fail(h, -2, "%s: %s", #x, hipGetErrorString(e_));      \
    } while (0)

static size_t gsz(const gam_handle *h) { return (size_t)h->gv.nc; }
static size_t fsz(const gam_handle *h) { return (size_t)h->gv.nf; }

template <typename R> static R *P(void *p) { return static_cast<R *>(p); }
template <typename R> static const GVT<R> &view(const gam_handle *h);
template <> const GVT<double> &view<double>(const gam_handle *h) { return h->gv; }
template <> const GVT<float> &view<float>(const gam_handle *h) { return h->gv32; }
template <typename R> static WKT<R> wkt(gam_handle *h) {
    WKT<R> w;
    w.Wh = P<R>(h->wk[W_WH]); w.Wn = P<R>(h->wk[W_WN]); w.VdV = P<R>(h->wk[W_VDV]);
    w.bh = P<R>(h->wk[W_BH]); w.bn = P<R>(h->wk[W_BN]);
    return w;
}

// ----------------------------------------------------------------- timers
static hipEvent_t ev_get(gam_handle *h) {
    if (!h->evpool.empty()) { hipEvent_t e = h->evpool.back(); h->evpool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) { h->timing = false; return nullptr; }
    return e;
}

static void collect_timers(gam_handle *h) {
    for (auto &p : h->pending) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) { h->ms[p.stage] += t; h->launches[p.stage]++; }
        h->evpool.push_back(p.a);
        h->evpool.push_back(p.b);
    }
    h->pending.clear();
}

struct StageTimer {
    gam_handle *h; int st; hipEvent_t a = nullptr, b = nullptr;
    StageTimer(gam_handle *h_, int st_) : h(h_), st(st_) {
        if (h->timing) {
            a = ev_get(h);
            if (!a || hipEventRecord(a, h->stream) != hipSuccess) { h->timing = false; a = nullptr; }
        }
    }
    ~StageTimer() {
        if (h->timing && a) {
            b = ev_get(h);
            if (b && hipEventRecord(b, h->stream) == hipSuccess) h->pending.push_back({st, a, b});
            else h->timing = false;
        }
    }
};

static inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

// ----------------------------------------------------------------- packed launches
// One launch of a kernel for several handles (the blocks of an ensemble): the
// blocks of member m are blocks first[m] .. first[m] + count[m] - 1 of the
// packed grid, first[m] a multiple of 8 so that each member's blocks keep
// their XCD classes; each runs the kernel's body (b_*) with the member's own
// arguments, read from a device table.  Pointers read from the table are
// re-typed as global memory, so the bodies' loads stay global loads as they
// are when the pointers arrive as kernel arguments.
template <typename... A> struct PkArgs;
template <> struct PkArgs<> {};
template <typename A0, typename... A> struct PkArgs<A0, A...> {
    A0 a;
    PkArgs<A...> r;
};
template <size_t I, typename P> __host__ __device__ constexpr const auto &pk_get(const P &p) {
    if constexpr (I == 0) return p.a;
    else return pk_get<I - 1>(p.r);
}
template <typename... A> struct PkMake;
template <> struct PkMake<> {
    static PkArgs<> make() { return {}; }
};
template <typename A0, typename... A> struct PkMake<A0, A...> {
    template <typename B0, typename... B> static PkArgs<A0, A...> make(B0 &&b0, B &&...b) {
        return {static_cast<A0>(b0), PkMake<A...>::make(std::forward<B>(b)...)};
    }
};
// the argument list of a body void b_*(unsigned vb, A...)
template <typename F> struct PkSig;
template <typename... A> struct PkSig<void (*)(unsigned, A...)> {
    using Args = PkArgs<A...>;
    using Make = PkMake<A...>;
    static constexpr size_t N = sizeof...(A);
};

template <typename T> __device__ __forceinline__ T pk_glob(const T &v) { return v; }
#ifdef GAM_EMU
template <typename T> __device__ __forceinline__ T *pk_glob(T *const &p) { return p; }
#else
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wincompatible-pointer-types-discards-qualifiers"
template <typename T> __device__ __forceinline__ T *pk_glob(T *const &p) {
    typedef __attribute__((address_space(1))) T GT;
    return (T *)*(GT *const *)&p;
}
#pragma clang diagnostic pop
#endif
template <typename R> __device__ __forceinline__ GVT<R> pk_glob(const GVT<R> &g) {
    GVT<R> o = g;
    o.xyz = pk_glob(g.xyz); o.xyzcc = pk_glob(g.xyzcc); o.xfc = pk_glob(g.xfc); o.dvol = pk_glob(g.dvol);
    o.face = pk_glob(g.face); o.Tf = pk_glob(g.Tf); o.edge = pk_glob(g.edge); o.Te = pk_glob(g.Te);
    o.Teb = pk_glob(g.Teb); o.vol = pk_glob(g.vol); o.di = pk_glob(g.di); o.dj = pk_glob(g.dj); o.dk = pk_glob(g.dk);
    o.b2fc = pk_glob(g.b2fc);
    return o;
}
template <typename R> __device__ __forceinline__ WKT<R> pk_glob(const WKT<R> &w) {
    WKT<R> o;
    o.Wh = pk_glob(w.Wh); o.Wn = pk_glob(w.Wn); o.VdV = pk_glob(w.VdV); o.bh = pk_glob(w.bh); o.bn = pk_glob(w.bn);
    return o;
}
__device__ __forceinline__ BGV pk_glob(const BGV &b) {
    BGV o = b;
    o.fcB0 = pk_glob(b.fcB0); o.edgB0 = pk_glob(b.edgB0); o.B0 = pk_glob(b.B0); o.bFlux0 = pk_glob(b.bFlux0);
    o.dpB0 = pk_glob(b.dpB0);
    return o;
}

template <auto Body, typename P, size_t... I>
__device__ __forceinline__ void pk_call(const unsigned vb, const P &p, std::index_sequence<I...>) {
    Body(vb, pk_glob(pk_get<I>(p))...);
}

template <int NT, auto Body>
__global__ void __launch_bounds__(NT) k_packed(const typename PkSig<decltype(Body)>::Args *__restrict__ args,
                                               const unsigned *__restrict__ first, const unsigned *__restrict__ count,
                                               int n) {
    int lo = 0, hi = n - 1;   // the last member whose first block is at or below this one
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const unsigned vb = blockIdx.x - first[lo];
    if (vb >= count[lo]) return;
    pk_call<Body>(vb, args[lo], std::make_index_sequence<PkSig<decltype(Body)>::N>{});
}

// one recorded launch of a handle (gam_handle::rec set): grid, block size and
// the body's arguments
struct PackRec {
    virtual ~PackRec() {}
    dim3 grid;
    int nt = 256;
    virtual const void *key() const = 0;   // the packed kernel's instantiation
    virtual size_t arg_bytes() const = 0;
    virtual void put_args(void *dst) const = 0;
    virtual void launch(const void *args, const unsigned *first, const unsigned *count, int n, unsigned total,
                        hipStream_t s) const = 0;
};
template <int NT, auto Body> struct PackRecT : PackRec {
    using Args = typename PkSig<decltype(Body)>::Args;
    Args args;
    PackRecT(dim3 g, const Args &a) : args(a) { grid = g; nt = NT; }
    const void *key() const override { return reinterpret_cast<const void *>(&k_packed<NT, Body>); }
    size_t arg_bytes() const override { return sizeof(Args); }
    void put_args(void *dst) const override { std::memcpy(dst, &args, sizeof(Args)); }
    void launch(const void *a, const unsigned *first, const unsigned *count, int n, unsigned total,
                hipStream_t s) const override {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_packed<NT, Body>), dim3(total), dim3(NT), 0, s, (const Args *)a, first,
                           count, n);
    }
};

// launches kernel KERN, or records it while h->rec is set (packed graphs);
// KERN and BODY (the kernel's b_* body) are parenthesized
template <auto Body, int NT, typename... B>
static void pk_record(std::vector<std::unique_ptr<PackRec>> *rec, dim3 grid, B &&...b) {
    using Sig = PkSig<decltype(Body)>;
    static_assert(sizeof...(B) == Sig::N, "argument count");
    rec->push_back(std::make_unique<PackRecT<NT, Body>>(grid, Sig::Make::make(std::forward<B>(b)...)));
}
#define PLAUNCH(h, KERN, BODY, grid, nt, ...)                                              \
    do {                                                                                   \
        if ((h)->rec) pk_record<BODY, nt>((h)->rec, (grid), __VA_ARGS__);                  \
        else hipLaunchKernelGGL(HIP_KERNEL_NAME KERN, (grid), (nt), 0, (h)->stream, __VA_ARGS__); \
    } while (0)

// stages whose kernels take a Tile, and the default shapes (256 threads each)
static bool tiled_stage(int st) {
    return st == ST_BHALF || st == ST_BPOST || st == ST_MOM2FACE || st == ST_EMF_I || st == ST_EMF_J ||
           st == ST_EMF_K || st == ST_FLUX_I || st == ST_FLUX_J || st == ST_FLUX_K || st == ST_UPDATE;
}
static void default_tiles(gam_handle *h) {
    // measured best on MI355X for the 256^3 blast wave (ms per launch, stage
    // timers): flux_j/flux_k/emf bricks halve their HBM traffic against
    // row-shaped blocks; the full-box kernels keep plain row-major blocks
    h->tile[ST_BHALF] = {256, 1, 1, TILE_LINEAR | TILE_NOXCD};
    h->tile[ST_BPOST] = {256, 1, 1, TILE_LINEAR | TILE_NOXCD};
    h->tile[ST_MOM2FACE] = {256, 1, 1, TILE_LINEAR};
    h->tile[ST_EMF_I] = {64, 1, 4, TILE_NOXCD};
    h->tile[ST_EMF_J] = {256, 1, 1, TILE_LINEAR | TILE_NOXCD};
    h->tile[ST_EMF_K] = {32, 8, 1, 1};
    h->tile[ST_UPDATE] = {256, 1, 1, TILE_LINEAR | TILE_NOXCD};
    h->tile[ST_FLUX_I] = {256, 1, 1, TILE_LINEAR | TILE_NOXCD};
    h->tile[ST_FLUX_J] = {32, 8, 1, 0};
    h->tile[ST_FLUX_K] = {64, 1, 4, 2};
}
static inline dim3 tgrid(const gam_handle *h, int st, int nx, int ny, int nz) {
    return dim3((unsigned)tile_blocks(h->tile[st], nx, ny, nz));
}

// FP32 copies of the bulk geometry (precision 1), refreshed after every change
static int sync_grid32(gam_handle *h) {
    if (!h->prec) return 0;
    const size_t G = gsz(h), F = fsz(h);
    const size_t n[G_N] = {3 * F, 3 * G, 9 * F, 3 * F, 27 * F, 3 * F, 27 * F, 12 * F, G, G, G, G};
    for (int q = 0; q < 9; q++) {
        const size_t m = n[kG32[q]];
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cvt<double, float>), dim3((unsigned)((m + 255) / 256)), 256, 0, h->stream,
                           (const double *)h->grid[kG32[q]], h->grid32[q], m);
    }
    hipLaunchKernelGGL(k_b2f_coef, grid1((long)G), 256, 0, h->stream, h->gv32, (const double *)h->grid[G_XFC],
                       (const double *)h->grid[G_XYZCC], (const double *)h->grid[G_VOL], h->b2fc);
    HIPCHK(h, hipGetLastError());
    return 0;
}

// ----------------------------------------------------------------- scratch
// Device memory of gam_divb_noalloc, the entry point that allocates nothing:
// a file-scope array of GAM_NOALLOC_BLOCKS block sums plus their total.  Every
// GPU the library runs on holds its own copy until the process ends.  In host
// code the name `scratch` is no device address: each GPU's copy is reached
// through the address hipGetSymbolAddress returns with that GPU current, which
// gam_create looks up once per device and keeps in g_scratch_dev.
static __device__ __align__(256) double scratch[GAM_NOALLOC_BLOCKS + 1];
#define GAM_MAX_DEVICES 64
static double *g_scratch_dev[GAM_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

static int scratch_lookup(gam_handle *h) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (h->device < 0 || h->device >= GAM_MAX_DEVICES) return fail(h, -1, "device %d: at most %d GPUs", h->device, GAM_MAX_DEVICES);
    if (!g_scratch_dev[h->device]) {
        void *p = nullptr;
        HIPCHK(h, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(h, -2, "hipGetSymbolAddress(scratch) returned no address");
        g_scratch_dev[h->device] = static_cast<double *>(p);
    }
    h->scratch = g_scratch_dev[h->device];
    return 0;
}

// ----------------------------------------------------------------- lifetime
extern "C" int gam_create(const gam_dims *dims, const gam_params *par, gam_handle **out) {
    if (!dims || !par || !out) return -1;
    *out = nullptr;
    if (dims->Nip < 1 || dims->Njp < 1 || dims->Nkp < 1 || dims->nG != 4)
        return -1;   // the stencils assume 4 ghost layers (init.F90:327)
    if (par->precision != 0 && par->precision != 1) return -1;
    if (par->doMHD != 0 && par->doMHD != 1) return -1;
    long Ni = dims->Nip + 2L * dims->nG, Nj = dims->Njp + 2L * dims->nG, Nk = dims->Nkp + 2L * dims->nG;
    long nf = (Ni + 1) * (Nj + 1) * (Nk + 1);
    if (nf * 27 > 0x7fffffffL * 8L) return -1;
    if (nf > 0x7fffffffL) return -1;
    gam_handle *h = new gam_handle();
    h->dims = *dims;
    h->par = *par;
    h->prec = par->precision;
    h->es = h->prec ? sizeof(float) : sizeof(double);
    GV &g = h->gv;
    g.Ni = (int)Ni; g.Nj = (int)Nj; g.Nk = (int)Nk; g.nG = dims->nG;
    g.Nip = dims->Nip; g.Njp = dims->Njp; g.Nkp = dims->Nkp;
    g.nc = (int)(Ni * Nj * Nk); g.nf = (int)nf;
    MP &m = h->mp;
    m.gamma = par->gamma; m.CFL = par->CFL; m.pdmb = par->pdmb; m.Vd0 = par->Vd0;
    m.cHogH = par->cHogH; m.cHogM = par->cHogM; m.Ca = par->Ca; m.dFloor = par->dFloor; m.pFloor = par->pFloor;
    m.rmeth = par->rmeth; m.doLFMLim = par->doLFMLim; m.doBoris = par->doBoris; m.doMHD = par->doMHD;
    m.do25D = par->do25D;
    auto cleanup = [&](int code) { gam_destroy(h); return code; };
    if (hipGetDevice(&h->device) != hipSuccess) return cleanup(-2);
    if (scratch_lookup(h)) return cleanup(-2);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(-2);
    const size_t G = gsz(h), F = fsz(h), es = h->es;
    const size_t gcomp[G_N] = {3 * F, 3 * G, 9 * F, 3 * F, 27 * F, 3 * F, 27 * F, 12 * F, G, G, G, G};
    for (int q = 0; q < G_N; q++)
        if (hipMalloc(&h->grid[q], gcomp[q] * sizeof(double)) != hipSuccess) return cleanup(-3);
    if (h->prec) {
        for (int q = 0; q < 9; q++)
            if (hipMalloc(&h->grid32[q], gcomp[kG32[q]] * sizeof(float)) != hipSuccess) return cleanup(-3);
        if (hipMalloc(&h->b2fc, 18 * G * sizeof(float)) != hipSuccess) return cleanup(-3);
    }
    for (int q = 0; q < 2; q++) {
        if (hipMalloc(&h->Gas[q], 5 * G * es) != hipSuccess) return cleanup(-3);
        if (hipMalloc(&h->mag[q], 3 * F * sizeof(double)) != hipSuccess) return cleanup(-3);
        if (hipMalloc(&h->B[q], 3 * G * es) != hipSuccess) return cleanup(-3);
        if (hipMemset(h->Gas[q], 0, 5 * G * es) != hipSuccess) return cleanup(-2);
        if (hipMemset(h->mag[q], 0, 3 * F * sizeof(double)) != hipSuccess) return cleanup(-2);
        if (hipMemset(h->B[q], 0, 3 * G * es) != hipSuccess) return cleanup(-2);
    }
    {
        void **z[11] = {&h->hGas, &h->hmag, &h->hB, &h->Vf, &h->E, &h->Flx, &h->wk[W_WH], &h->wk[W_WN],
                        &h->wk[W_VDV], &h->wk[W_BH], &h->wk[W_BN]};
        // K fluxes stay 0 in 2.5D (init.F90:500-501 zero the flux work arrays)
        const size_t zn[11] = {5 * G, 3 * F, 3 * G, 9 * G, 3 * G, 24 * G, 5 * G, 5 * G, 3 * G, 3 * F, 3 * F};
        for (int q = 0; q < 11; q++) {
            if (hipMalloc(z[q], zn[q] * es) != hipSuccess) return cleanup(-3);
            if (hipMemset(*z[q], 0, zn[q] * es) != hipSuccess) return cleanup(-2);
        }
    }
    default_tiles(h);
    h->npart = (int)((G + 255) / 256);
    h->pcap = std::max(h->npart, tile_blocks(h->tile[ST_UPDATE], g.Ni, g.Nj, g.Nk));
    if (hipMalloc(&h->partial, (size_t)h->pcap * sizeof(double)) != hipSuccess) return cleanup(-3);
    if (hipMalloc(&h->clk, sizeof(DClock)) != hipSuccess) return cleanup(-3);
    if (hipHostMalloc(&h->clk_host, sizeof(DClock), hipHostMallocDefault) != hipSuccess) return cleanup(-3);
    DClock c = {};
    c.tFin = par->tFin; c.limDT0 = par->limDT0; c.limCPR = par->limCPR;
    c.fixedTimestep = par->fixedTimestep; c.doCPR = par->doCPR;
    h->docpr = par->doCPR;
    *h->clk_host = c;
    if (hipMemcpy(h->clk, &c, sizeof c, hipMemcpyHostToDevice) != hipSuccess) return cleanup(-2);
    // geometry defaults (init.F90:846-850,973-976,1098)
    double *gp[4] = {h->grid[G_FACE], h->grid[G_EDGE], h->grid[G_TEB], nullptr};
    size_t gn[3] = {3 * F, 3 * F, 12 * F};
    for (int q = 0; q < 3; q++) hipLaunchKernelGGL(k_fill, grid1((long)gn[q]), 256, 0, h->stream, gp[q], gn[q], 1.0);
    if (hipMemsetAsync(h->grid[G_XFC], 0, 9 * F * sizeof(double), h->stream) != hipSuccess ||
        hipMemsetAsync(h->grid[G_TF], 0, 27 * F * sizeof(double), h->stream) != hipSuccess ||
        hipMemsetAsync(h->grid[G_TE], 0, 27 * F * sizeof(double), h->stream) != hipSuccess)
        return cleanup(-2);
    // device views
    g.xyz = h->grid[G_XYZ]; g.xyzcc = h->grid[G_XYZCC]; g.xfc = h->grid[G_XFC]; g.face = h->grid[G_FACE];
    g.Tf = h->grid[G_TF]; g.edge = h->grid[G_EDGE]; g.Te = h->grid[G_TE]; g.Teb = h->grid[G_TEB];
    g.vol = h->grid[G_VOL]; g.di = h->grid[G_DI]; g.dj = h->grid[G_DJ]; g.dk = h->grid[G_DK];
    g.dvol = h->grid[G_VOL];
    GVT<float> &g32 = h->gv32;
    g32.Ni = g.Ni; g32.Nj = g.Nj; g32.Nk = g.Nk; g32.nG = g.nG; g32.Nip = g.Nip; g32.Njp = g.Njp; g32.Nkp = g.Nkp;
    g32.nc = g.nc; g32.nf = g.nf;
    g32.xyz = g.xyz; g32.xyzcc = g.xyzcc; g32.xfc = g.xfc; g32.dvol = g.dvol;
    g32.face = h->grid32[0]; g32.Tf = h->grid32[1]; g32.edge = h->grid32[2]; g32.Te = h->grid32[3];
    g32.Teb = h->grid32[4]; g32.vol = h->grid32[5]; g32.di = h->grid32[6]; g32.dj = h->grid32[7]; g32.dk = h->grid32[8];
    g32.b2fc = h->b2fc;
    g.b2fc = nullptr;
    if (sync_grid32(h)) return cleanup(-2);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup(-2);
    *out = h;
    return 0;
}

extern "C" int gam_destroy(gam_handle *h) {
    if (!h) return 0;
    if (refuse_if_captured(h, "gam_destroy")) return -1;
    // teardown: failures here cannot be acted on, the handle is gone either way
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    collect_timers(h);
    for (auto e : h->evpool) (void)hipEventDestroy(e);
    for (int q = 0; q < G_N; q++) if (h->grid[q]) (void)hipFree(h->grid[q]);
    for (int q = 0; q < 9; q++) if (h->grid32[q]) (void)hipFree(h->grid32[q]);
    if (h->b2fc) (void)hipFree(h->b2fc);
    for (int q = 0; q < 5; q++) if (h->bg[q]) (void)hipFree(h->bg[q]);
    for (double *b : {h->ctG, h->ctM, h->crG, h->crM, h->ringSE, h->ringTab, h->ringMean, h->windSer}) if (b) (void)hipFree(b);
    for (int q = 0; q < 2; q++) { (void)hipFree(h->Gas[q]); (void)hipFree(h->mag[q]); (void)hipFree(h->B[q]); }
    void *bufs[13] = {h->hGas, h->hmag, h->hB, h->Vf, h->E, h->Flx, h->partial, h->clk,
                      h->wk[0], h->wk[1], h->wk[2], h->wk[3], h->wk[4]};
    for (void *b : bufs) if (b) (void)hipFree(b);
    if (h->clk_host) (void)hipHostFree(h->clk_host);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

extern "C" const char *gam_last_error(const gam_handle *h) { return h ? h->err.c_str() : "null handle"; }

extern "C" int gam_sync(gam_handle *h) {
    if (!h) return -1;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    collect_timers(h);
    return 0;
}

extern "C" void *gam_stream(gam_handle *h) { return h ? (void *)h->stream : nullptr; }

// ----------------------------------------------------------------- copies
static int h2d(gam_handle *h, double *dst, const double *src, size_t n) {
    if (!src) return 0;
    HIPCHK(h, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    return 0;
}
static int d2h(gam_handle *h, double *dst, const double *src, size_t n) {
    if (!dst) return 0;
    HIPCHK(h, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    return 0;
}
// host FP64 <-> device bulk array of n elements (converted on the host in FP32 mode)
static int h2d_bulk(gam_handle *h, void *dst, const double *src, size_t n) {
    if (!src) return 0;
    if (!h->prec) return h2d(h, (double *)dst, src, n);
    std::vector<float> tmp(n);
    for (size_t q = 0; q < n; q++) tmp[q] = (float)src[q];
    HIPCHK(h, hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));   // tmp goes out of scope
    return 0;
}
static int d2h_bulk(gam_handle *h, double *dst, const void *src, size_t n, bool bulk = true) {
    if (!dst) return 0;
    if (!h->prec || !bulk) return d2h(h, dst, (const double *)src, n);
    std::vector<float> tmp(n);
    HIPCHK(h, hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (size_t q = 0; q < n; q++) dst[q] = (double)tmp[q];
    return 0;
}

static GVw gvw(gam_handle *h) {
    const GV &g = h->gv;
    GVw w;
    w.Ni = g.Ni; w.Nj = g.Nj; w.Nk = g.Nk; w.nG = g.nG; w.Nip = g.Nip; w.Njp = g.Njp; w.Nkp = g.Nkp;
    w.nc = g.nc; w.nf = g.nf;
    w.xyz = h->grid[G_XYZ]; w.xyzcc = h->grid[G_XYZCC]; w.xfc = h->grid[G_XFC]; w.face = h->grid[G_FACE];
    w.Tf = h->grid[G_TF]; w.edge = h->grid[G_EDGE]; w.Te = h->grid[G_TE]; w.Teb = h->grid[G_TEB];
    w.vol = h->grid[G_VOL]; w.di = h->grid[G_DI]; w.dj = h->grid[G_DJ]; w.dk = h->grid[G_DK];
    return w;
}

extern "C" int gam_set_grid(gam_handle *h, const double *xyz, const double *xyzcc, const double *xfc,
                            const double *face, const double *Tf, const double *edge, const double *Te,
                            const double *Teb, const double *volume, const double *di, const double *dj,
                            const double *dk) {
    if (!h) return -1;
    const size_t G = gsz(h), F = fsz(h);
    const double *src[G_N] = {xyz, xyzcc, xfc, face, Tf, edge, Te, Teb, volume, di, dj, dk};
    const size_t n[G_N] = {3 * F, 3 * G, 9 * F, 3 * F, 27 * F, 3 * F, 27 * F, 12 * F, G, G, G, G};
    for (int q = 0; q < G_N; q++) { int rc = h2d(h, h->grid[q], src[q], n[q]); if (rc) return rc; }
    int rc = sync_grid32(h);
    if (rc) return rc;
    return gam_sync(h);
}

extern "C" int gam_get_grid(gam_handle *h, double *xyz, double *xyzcc, double *xfc, double *face, double *Tf,
                            double *edge, double *Te, double *Teb, double *volume, double *di, double *dj, double *dk) {
    if (!h) return -1;
    const size_t G = gsz(h), F = fsz(h);
    double *dst[G_N] = {xyz, xyzcc, xfc, face, Tf, edge, Te, Teb, volume, di, dj, dk};
    const size_t n[G_N] = {3 * F, 3 * G, 9 * F, 3 * F, 27 * F, 3 * F, 27 * F, 12 * F, G, G, G, G};
    for (int q = 0; q < G_N; q++) { int rc = d2h(h, dst[q], h->grid[q], n[q]); if (rc) return rc; }
    return gam_sync(h);
}

extern "C" int gam_set_corners(gam_handle *h, const double *xyz) {
    // init.F90:801-1080 Corners2Grid + ebGeom on the GPU
    if (!h || !xyz) return -1;
    const GV &g = h->gv;
    if (g.Nip < 8) return fail(h, -1, "gam_set_corners: Nip must be >= 8 for the 8-point stencils");
    int rc = h2d(h, h->grid[G_XYZ], xyz, 3 * fsz(h));
    if (rc) return rc;
    GVw w = gvw(h);
    for (int d = 0; d < 3; d++) {
        long n = (long)(g.Ni + (d == 0)) * (g.Nj + (d == 1)) * (g.Nk + (d == 2));
        hipLaunchKernelGGL(k_faces, grid1(n), 256, 0, h->stream, w, d);
    }
    hipLaunchKernelGGL(k_volume, grid1(g.nc), 256, 0, h->stream, w);
    if (h->ring.on) {   // RingGridFix (init.F90:968), before the edge systems read the face normals
        hipLaunchKernelGGL(k_ring_dk, grid1((long)g.Ni * g.Nk * h->ring.nring * 2), 256, 0, h->stream, w, h->ring);
        hipLaunchKernelGGL(k_ring_ghost_geom, grid1((long)(g.Ni + 1) * (g.Nk + 1) * g.nG * 2), 256, 0, h->stream, w);
    }
    for (int d = 0; d < 3; d++) {
        long nx = (d == 1) ? g.Ni - 1 : (d == 0 ? g.Ni : g.Ni + 1);
        long ny = (d == 0) ? g.Nj + 1 : (d == 1 ? g.Nj : g.Nj - 1);
        long nz = (d == 0) ? g.Nk - 1 : (d == 1 ? g.Nk + 1 : g.Nk);
        hipLaunchKernelGGL(k_edges, grid1(nx * ny * nz), 256, 0, h->stream, w, d);
    }
    long ne = (long)(g.Nip + 1) * (g.Njp + 1) * (g.Nkp + 1);
    for (int dN = 0; dN < 3; dN++) {
        int T1 = (dN + 1) % 3, T2 = (dN + 2) % 3;
        hipLaunchKernelGGL(k_teb, grid1(ne), 256, 0, h->stream, w, dN, T1, T2, XNQI);
        hipLaunchKernelGGL(k_teb, grid1(ne), 256, 0, h->stream, w, dN, T2, T1, XNQJ);
    }
    HIPCHK(h, hipGetLastError());
    if ((rc = sync_grid32(h))) return rc;
    h->corners_set = true;
    return gam_sync(h);
}

// which: 0 State, 1 oState, 2 predictor state (its magFlux is bulk-typed)
static void state_ptrs(gam_handle *h, int which, void **Gas, void **mag, void **B) {
    if (which == 2) { *Gas = h->hGas; *mag = h->hmag; *B = h->hB; return; }
    int q = which == 0 ? h->s : 1 - h->s;
    *Gas = h->Gas[q]; *mag = h->mag[q]; *B = h->B[q];
}

template <typename R, typename TM> static void launch_b2f(gam_handle *h, int st, const TM *mag, void *B);

extern "C" int gam_put_state(gam_handle *h, int which, double time, const double *Gas, const double *magFlux,
                             const double *Bxyz) {
    if (!h || which < 0 || which > 1) return -1;
    void *g, *m, *b;
    state_ptrs(h, which, &g, &m, &b);
    int rc;
    if ((rc = h2d_bulk(h, g, Gas, 5 * gsz(h)))) return rc;
    if ((rc = h2d(h, (double *)m, magFlux, 3 * fsz(h)))) return rc;
    if (Bxyz) {
        if ((rc = h2d_bulk(h, b, Bxyz, 3 * gsz(h)))) return rc;
    } else if (magFlux && h->prec) {   // bFlux2Fld of the uploaded magFlux
        launch_b2f<float, double>(h, ST_BPOST, (const double *)m, b);
    } else if (magFlux) {
        launch_b2f<double, double>(h, ST_BPOST, (const double *)m, b);
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(h->clk_host, h->clk, sizeof(DClock), hipMemcpyDeviceToHost));
    if (which == 0) h->clk_host->stime = time; else h->clk_host->otime = time;
    HIPCHK(h, hipMemcpy(h->clk, h->clk_host, sizeof(DClock), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int gam_put_state_device(gam_handle *h, int which, double time, const double *Gas, const double *magFlux) {
    // FP64 device sources; Bxyz = bFlux2Fld(magFlux) on the device
    if (!h || which < 0 || which > 1 || !Gas || !magFlux) return -1;
    void *g, *m, *b;
    state_ptrs(h, which, &g, &m, &b);
    const size_t nG5 = 5 * gsz(h);
    if (h->prec)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cvt<double, float>), dim3((unsigned)((nG5 + 255) / 256)), 256, 0, h->stream,
                           Gas, (float *)g, nG5);
    else
        HIPCHK(h, hipMemcpyAsync(g, Gas, nG5 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(m, magFlux, 3 * fsz(h) * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    if (h->prec) launch_b2f<float, double>(h, ST_BPOST, (const double *)m, b);
    else launch_b2f<double, double>(h, ST_BPOST, (const double *)m, b);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(h->clk_host, h->clk, sizeof(DClock), hipMemcpyDeviceToHost));
    if (which == 0) h->clk_host->stime = time; else h->clk_host->otime = time;
    HIPCHK(h, hipMemcpy(h->clk, h->clk_host, sizeof(DClock), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int gam_get_state(gam_handle *h, int which, double *time, double *Gas, double *magFlux, double *Bxyz) {
    if (!h || which < 0 || which > 2) return -1;
    void *g, *m, *b;
    state_ptrs(h, which, &g, &m, &b);
    int rc;
    if ((rc = d2h_bulk(h, Gas, g, 5 * gsz(h)))) return rc;
    if ((rc = d2h_bulk(h, magFlux, m, 3 * fsz(h), which == 2))) return rc;
    if ((rc = d2h_bulk(h, Bxyz, b, 3 * gsz(h)))) return rc;
    HIPCHK(h, hipMemcpyAsync(h->clk_host, h->clk, sizeof(DClock), hipMemcpyDeviceToHost, h->stream));
    if ((rc = gam_sync(h))) return rc;
    if (time) *time = which == 0 ? h->clk_host->stime : (which == 1 ? h->clk_host->otime : h->clk_host->stime + h->clk_host->pdt);
    return 0;
}

extern "C" int gam_set_clock(gam_handle *h, const gam_clock *c) {
    if (!h || !c) return -1;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(h->clk_host, h->clk, sizeof(DClock), hipMemcpyDeviceToHost));
    DClock &d = *h->clk_host;
    d.t = c->t; d.dt = c->dt; d.dt0 = c->dt0; d.stime = c->stateTime; d.otime = c->oStateTime;
    d.ts = c->ts; d.status = c->status; d.dt0_after = c->dt0_after; d.ncpr = c->ncpr; d.cpr_pending = 0;
    d.halt = 0;
    HIPCHK(h, hipMemcpy(h->clk, h->clk_host, sizeof(DClock), hipMemcpyHostToDevice));
    return 0;
}

extern "C" int gam_get_clock(gam_handle *h, gam_clock *c) {
    if (!h || !c) return -1;
    HIPCHK(h, hipMemcpyAsync(h->clk_host, h->clk, sizeof(DClock), hipMemcpyDeviceToHost, h->stream));
    int rc = gam_sync(h);
    if (rc) return rc;
    const DClock &d = *h->clk_host;
    c->t = d.t; c->dt = d.dt; c->dt0 = d.dt0; c->stateTime = d.stime; c->oStateTime = d.otime;
    c->ts = d.ts; c->status = d.status; c->dt0_after = d.dt0_after; c->ncpr = d.ncpr;
    return 0;
}

// ----------------------------------------------------------------- ICs
// The initial conditions are computed in FP64.  In FP32 mode they go to an
// FP64 scratch first and are rounded into the FP32 State.
struct IcScratch {
    double *buf = nullptr;
    ~IcScratch() { if (buf) (void)hipFree(buf); }
};

static int ic_gas_target(gam_handle *h, IcScratch &sc, double **gas, double **lA) {
    const size_t G = gsz(h);
    if (!h->prec) { *gas = (double *)h->Gas[h->s]; *lA = (double *)h->E; return 0; }
    HIPCHK(h, hipMalloc(&sc.buf, 8 * G * sizeof(double)));
    *gas = sc.buf;
    *lA = sc.buf + 5 * G;
    // the FP32 State keeps whatever the scratch holds outside the IC's cells
    const size_t n = 5 * G;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cvt<float, double>), dim3((unsigned)((n + 255) / 256)), 256, 0, h->stream,
                       (const float *)h->Gas[h->s], *gas, n);
    HIPCHK(h, hipGetLastError());
    return 0;
}

static int ic_gas_finish(gam_handle *h, IcScratch &sc) {
    if (!h->prec) return 0;
    const size_t n = 5 * gsz(h);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cvt<double, float>), dim3((unsigned)((n + 255) / 256)), 256, 0, h->stream,
                       (const double *)sc.buf, (float *)h->Gas[h->s], n);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));   // before the scratch is freed
    return 0;
}

extern "C" int gam_ic_orszag_tang(gam_handle *h, double rho0, double p0, double b0, double kscale) {
    if (!h) return -1;
    const GV &g = h->gv;
    IcScratch sc;
    double *gas, *lA;
    int rc = ic_gas_target(h, sc, &gas, &lA);
    if (rc) return rc;
    long n = (long)(g.Nip + 1) * (g.Njp + 1) * (g.Nkp + 1);
    hipLaunchKernelGGL(k_ic_ot, grid1(n), 256, 0, h->stream, g, h->mp.gamma, rho0, p0, b0, kscale, gas, h->mag[h->s]);
    HIPCHK(h, hipGetLastError());
    if ((rc = ic_gas_finish(h, sc))) return rc;
    return gam_sync(h);
}

extern "C" int gam_ic_blast_wave(gam_handle *h, double x0, double y0, double z0, double D0, double P0,
                                 double pRat, double dRat, double rC, double B0) {
    if (!h) return -1;
    const GV &g = h->gv;
    GVw w = gvw(h);
    IcScratch sc;
    double *gas, *lA;
    int rc = ic_gas_target(h, sc, &gas, &lA);
    if (rc) return rc;
    BWp bw = {x0, y0, z0, D0, P0, pRat, dRat, rC};
    long n = (long)g.Nip * g.Njp * g.Nkp;
    hipLaunchKernelGGL(k_ic_bw_gas, grid1(n), 256, 0, h->stream, w, h->mp.gamma, bw, gas);
    if (h->mp.doMHD) {
        // prob.F90:234 bScl = B0/sqrt(2.0) with a default-real sqrt(2.0)
        double bScl = B0 / (double)sqrtf(2.0f), bSclz = 0.0;
        // scratch, zeroed like the reference's allocGridVec
        HIPCHK(h, hipMemsetAsync(lA, 0, 3 * gsz(h) * sizeof(double), h->stream));
        long m = (long)(g.Ni - 1) * (g.Nj - 1) * (g.Nk - 1);
        hipLaunchKernelGGL(k_ic_bw_la, grid1(m), 256, 0, h->stream, w, bScl, bSclz, lA);
        hipLaunchKernelGGL(k_ic_bw_flux, grid1(m), 256, 0, h->stream, w, (const double *)lA, h->mag[h->s]);
    }
    HIPCHK(h, hipGetLastError());
    if ((rc = ic_gas_finish(h, sc))) return rc;
    return gam_sync(h);
}

// ----------------------------------------------------------------- background field
// Grid_T fcB0/edgB0/B0/bFlux0/dpB0 (gamtypes.F90:231-238), G-sized, FP64.
static const int kBgComp[5] = {9, 6, 3, 3, 3};

static int bg_alloc(gam_handle *h) {
    const size_t G = gsz(h);
    for (int q = 0; q < 5; q++)
        if (!h->bg[q]) HIPCHK(h, hipMalloc(&h->bg[q], kBgComp[q] * G * sizeof(double)));
    h->bgv.fcB0 = h->bg[0]; h->bgv.edgB0 = h->bg[1]; h->bgv.B0 = h->bg[2];
    h->bgv.bFlux0 = h->bg[3]; h->bgv.dpB0 = h->bg[4];
    return 0;
}

extern "C" int gam_set_background(gam_handle *h, const double *fcB0, const double *edgB0, const double *B0,
                                  const double *bFlux0, const double *dpB0) {
    if (!h || !fcB0 || !edgB0 || !B0 || !bFlux0 || !dpB0) return -1;
    if (refuse_if_captured(h, "gam_set_background")) return -1;
    int rc = bg_alloc(h);
    if (rc) return rc;
    const double *src[5] = {fcB0, edgB0, B0, bFlux0, dpB0};
    for (int q = 0; q < 5; q++)
        if ((rc = h2d(h, h->bg[q], src[q], kBgComp[q] * gsz(h)))) return rc;
    if ((rc = gam_sync(h))) return rc;
    h->bgv.on = 1;
    return 0;
}

extern "C" int gam_get_background(gam_handle *h, double *fcB0, double *edgB0, double *B0, double *bFlux0,
                                  double *dpB0) {
    if (!h) return -1;
    if (!h->bgv.on) return fail(h, -1, "gam_get_background: no background field set");
    double *dst[5] = {fcB0, edgB0, B0, bFlux0, dpB0};
    int rc;
    for (int q = 0; q < 5; q++)
        if ((rc = d2h(h, dst[q], h->bg[q], kBgComp[q] * gsz(h)))) return rc;
    return gam_sync(h);
}

extern "C" int gam_background_dipole(gam_handle *h, double M0, const double *xc, const double *bu) {
    // background.F90:20-167 AddB0 on the GPU for B0 = bu + dipole(x - xc)
    if (!h || !xc || !bu) return -1;
    if (refuse_if_captured(h, "gam_background_dipole")) return -1;
    int rc = bg_alloc(h);
    if (rc) return rc;
    const GV &g = h->gv;
    const size_t G = gsz(h);
    double *fS = nullptr;
    HIPCHK(h, hipMalloc(&fS, 9 * G * sizeof(double)));
    B0P p = {M0, {xc[0], xc[1], xc[2]}, {bu[0], bu[1], bu[2]}};
    GVw w = gvw(h);
    hipError_t e = hipSuccess;
    for (int q = 0; q < 5 && e == hipSuccess; q++) e = hipMemsetAsync(h->bg[q], 0, kBgComp[q] * G * sizeof(double), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(fS, 0, 9 * G * sizeof(double), h->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_b0_faces, grid1(g.nc), 256, 0, h->stream, w, p, h->bg[0], h->bg[3], fS);
        hipLaunchKernelGGL(k_b0_edges, grid1(g.nc), 256, 0, h->stream, w, p, h->bg[1]);
        hipLaunchKernelGGL(k_b0_cells, grid1(g.nc), 256, 0, h->stream, w, p, h->bg[2]);
        hipLaunchKernelGGL(k_b0_dp, grid1((long)g.Nip * g.Njp * g.Nkp), 256, 0, h->stream, w, (const double *)fS, h->bg[4]);
        e = hipGetLastError();
    }
    hipError_t e2 = hipStreamSynchronize(h->stream);
    (void)hipFree(fS);
    if (e != hipSuccess) return fail(h, -2, "gam_background_dipole: %s", hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(h, -2, "gam_background_dipole: %s", hipGetErrorString(e2));
    h->bgv.on = 1;
    return 0;
}

static bool all_periodic(const gam_handle *h);

// ----------------------------------------------------------------- device buffers
extern "C" int gam_device_alloc(size_t bytes, void **p) {
    if (!p) return -1;
    *p = nullptr;
    if (bytes == 0) return -1;
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return -3;
    }
    return 0;
}

extern "C" int gam_device_free(void *p) {
    if (!p) return 0;
    return hipFree(p) == hipSuccess ? 0 : -3;
}

extern "C" int gam_memcpy(void *dst, const void *src, size_t bytes) {
    if (!dst || !src) return -1;
    return hipMemcpy(dst, src, bytes, hipMemcpyDefault) == hipSuccess ? 0 : -3;
}

// ----------------------------------------------------------------- learned correction
extern "C" int gam_state_device(gam_handle *h, int which, void **Gas, double **magFlux, void **Bxyz) {
    if (!h || which < 0 || which > 1) return -1;
    void *g, *m, *b;
    state_ptrs(h, which, &g, &m, &b);
    if (Gas) *Gas = g;
    if (magFlux) *magFlux = (double *)m;
    if (Bxyz) *Bxyz = b;
    return 0;
}

extern "C" int gam_set_correction_wait(gam_handle *h, void *event) {
    if (!h) return -1;
    if (refuse_if_captured(h, "gam_set_correction_wait")) return -1;
    int rc = gam_sync(h);
    if (rc) return rc;
    h->corr_wait = (hipEvent_t)event;
    return 0;
}

extern "C" int gam_set_correction(gam_handle *h, const float *o, double scale) {
    if (!h) return -1;
    if (refuse_if_captured(h, "gam_set_correction")) return -1;
    if (o && !all_periodic(h)) return fail(h, -1, "gam_set_correction: periodic boundaries only");
    int rc = gam_sync(h);
    if (rc) return rc;
    h->corr = o;
    h->corr_scale = scale;
    return 0;
}

static int corr_target(gam_handle *h, const double *dGas, const double *dMag, bool dev) {
    if (!h || !dGas || !dMag) return -1;
    const size_t G = gsz(h), F = fsz(h);
    if (!h->ctG) {
        HIPCHK(h, hipMalloc(&h->ctG, 5 * G * sizeof(double)));
        HIPCHK(h, hipMalloc(&h->ctM, 3 * F * sizeof(double)));
        HIPCHK(h, hipMalloc(&h->crG, 5 * G * sizeof(double)));
        HIPCHK(h, hipMalloc(&h->crM, 3 * F * sizeof(double)));
    }
    if (dev) {
        HIPCHK(h, hipMemcpyAsync(h->ctG, dGas, 5 * G * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->ctM, dMag, 3 * F * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    } else {
        int rc;
        if ((rc = h2d(h, h->ctG, dGas, 5 * G))) return rc;
        if ((rc = h2d(h, h->ctM, dMag, 3 * F))) return rc;
    }
    return gam_sync(h);
}

extern "C" int gam_correction_target(gam_handle *h, const double *dGas, const double *dMag) {
    return corr_target(h, dGas, dMag, false);
}

extern "C" int gam_correction_target_device(gam_handle *h, const double *dGas, const double *dMag) {
    return corr_target(h, dGas, dMag, true);
}

template <typename R>
static int corr_grad_t(gam_handle *h, const float *o, double scale, double dt, const double *w, float *dout, double *loss) {
    const GVT<R> &g = view<R>(h);
    const int nbc = (int)grid1(g.nc).x, nbf = (int)grid1(g.nf).x;
    if (nbc > h->pcap || nbf > h->pcap) {   // the block sums reuse h->partial
        double *p = nullptr;
        HIPCHK(h, hipMalloc(&p, (size_t)std::max(nbc, nbf) * sizeof(double)));
        HIPCHK(h, hipFree(h->partial));
        h->partial = p;
        h->pcap = std::max(nbc, nbf);
    }
    std::vector<double> hp(std::max(nbc, nbf));
    double L = 0.0;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_corr_cells<R>), dim3(nbc), 256, 0, h->stream, g, h->mp.do25D, o, scale, dt,
                       (const double *)h->ctG, w[0], w[1], w[2], h->crG, h->partial);
    HIPCHK(h, hipMemcpyAsync(hp.data(), h->partial, nbc * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int q = 0; q < nbc; q++) L += hp[q];
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_corr_faces<R>), dim3(nbf), 256, 0, h->stream, g, o, scale, dt,
                       (const double *)h->ctM, w[3], h->crM, h->partial);
    HIPCHK(h, hipMemcpyAsync(hp.data(), h->partial, nbf * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int q = 0; q < nbf; q++) L += hp[q];
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_corr_adjoint<R>), dim3(nbc), 256, 0, h->stream, g, h->mp.do25D, scale, dt,
                       (const double *)h->crG, (const double *)h->crM, dout);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (loss) *loss = L;
    return 0;
}

extern "C" int gam_correction_grad(gam_handle *h, const float *o, double scale, double dt, const double *w4,
                                   float *dout, double *loss) {
    if (!h || !o || !w4 || !dout) return -1;
    if (!h->ctG) return fail(h, -1, "gam_correction_grad: no target set (gam_correction_target)");
    return h->prec ? corr_grad_t<float>(h, o, scale, dt, w4, dout, loss) : corr_grad_t<double>(h, o, scale, dt, w4, dout, loss);
}

// ----------------------------------------------------------------- step pieces
// CellDT min over the DT box of (Gas, B) into h->partial (CalcDT's loop).
template <typename R>
static int launch_celldt(gam_handle *h, const void *Gas, const void *B) {
    if (!Gas || !B) return fail(h, -1, "launch_celldt: null state");
    StageTimer t(h, ST_CELLDT);
    const GVT<R> &g = view<R>(h);
    if (h->mp.doBoris) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_celldt<1, R>), dim3(h->npart), 256, 0, h->stream, g, h->mp, h->bgv, h->clk, (const R *)Gas, (const R *)B, h->partial);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_celldt<0, R>), dim3(h->npart), 256, 0, h->stream, g, h->mp, h->bgv, h->clk, (const R *)Gas, (const R *)B, h->partial);
    HIPCHK(h, hipGetLastError());
    return 0;
}
static int celldt(gam_handle *h, const void *Gas, const void *B) {
    return h->prec ? launch_celldt<float>(h, Gas, B) : launch_celldt<double>(h, Gas, B);
}

// k_update's launch box: (i, j, column of UpdateKC cells in k)
static int update_kc(const gam_handle *h) { return h->prec ? 1 : UPDATE_KC; }
static int update_blocks(const gam_handle *h) {
    return tile_blocks(h->tile[ST_UPDATE], h->gv.Ni, h->gv.Nj, cdiv(h->gv.Nk, update_kc(h)));
}

// blocks of k_celldt_rows (DT box cells of the ring band), 0 without rings
static int ring_dt_blocks(const gam_handle *h) {
    if (!h->ring.on) return 0;
    const long n = (long)(h->gv.Nip + 2) * (h->gv.Nkp + 2) * 2 * (h->ring.nring + 3);
    return (int)((n + 255) / 256);
}

// room for n block minima in h->partial
static int ensure_partial(gam_handle *h, int n) {
    if (n <= h->pcap) return 0;
    double *p = nullptr;
    HIPCHK(h, hipMalloc(&p, (size_t)n * sizeof(double)));
    HIPCHK(h, hipFree(h->partial));
    h->partial = p;
    h->pcap = n;
    return 0;
}

// np = number of block minima in h->partial (calc != 0)
static void launch_dtfinal(gam_handle *h, int advance, int calc, int np) {
    StageTimer t(h, ST_DTFINAL);
    PLAUNCH(h, (k_dtfinal), (b_dtfinal), dim3(1), 1024, h->clk, h->partial, np, advance, calc);
}

// StateCPR on State when the CalcDT just done asked for it (doCPR only; the
// kernel returns at once unless the device clock says so)
template <typename R>
static void launch_cpr_t(gam_handle *h) {
    const GVT<R> &g = view<R>(h);
    const long n = (long)g.Nip * g.Njp * g.Nkp;
    if (h->mp.doBoris)
        PLAUNCH(h, (k_cpr<1, R>), (b_cpr<1, R>), grid1(n), 256, g, h->mp, h->bgv, h->clk,
                           P<R>(h->Gas[h->s]), P<R>(h->B[h->s]));
    else
        PLAUNCH(h, (k_cpr<0, R>), (b_cpr<0, R>), grid1(n), 256, g, h->mp, h->bgv, h->clk,
                           P<R>(h->Gas[h->s]), P<R>(h->B[h->s]));
}
static void launch_cpr(gam_handle *h) {
    if (!h->docpr) return;
    if (h->prec) launch_cpr_t<float>(h); else launch_cpr_t<double>(h);
}

// the flux launches with the learned correction added at the store (a block
// with a correction is periodic, no ring average, so never packed)
template <int RM, int BO, int MH, int BG, typename R>
static void launch_flux_corr(gam_handle *h, int d, dim3 gr, dim3 gr2, Tile tl, const WKT<R> &w, const R *nB, const R *hB,
                             const R *hmag, const R *hGas, R *Flx) {
    const GVT<R> &g = view<R>(h);
#define FLUX_CORR(KERN, GRID) hipLaunchKernelGGL(HIP_KERNEL_NAME KERN, (GRID), 256, 0, h->stream, g, tl, h->mp, h->bgv, h->clk, \
                                                 w.Wn, nB, w.Wh, hB, hmag, hGas, Flx, h->corr, h->corr_scale)
    if (d == 0) FLUX_CORR((k_flux<0, RM, BO, MH, BG, R, 1>), gr);
    else if (sizeof(R) == sizeof(float)) {
        if (d == 1) FLUX_CORR((k_flux<1, RM, BO, MH, BG, R, 1>), gr);
        else FLUX_CORR((k_flux<2, RM, BO, MH, BG, R, 1>), gr);
    }
    else if (d == 1) {   // the plain pairs, then the correction on its own
        FLUX_CORR((k_flux2<1, RM, BO, MH, BG, R, 0>), gr2);
        const long n = (long)g.Nip * (g.Njp + 1) * (h->mp.do25D ? 1 : g.Nkp);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_corr_flux<1, R>), grid1(n), 256, 0, h->stream, g, h->mp, h->clk, Flx, h->corr,
                           h->corr_scale);
    }
    else FLUX_CORR((k_flux2<2, RM, BO, MH, BG, R, 1>), gr2);
#undef FLUX_CORR
}

template <int RM, int BO, int MH, int BG, typename R>
static void launch_flux_bg(gam_handle *h, int d) {
    const GVT<R> &g = view<R>(h);
    const int S = h->s;
    long nx = g.Nip + (d == 0), ny = g.Njp + (d == 1), nz = (h->mp.do25D ? 1 : g.Nkp) + (d == 2);
    const Tile tl = h->tile[ST_FLUX_I + d];
    dim3 gr = tgrid(h, ST_FLUX_I + d, (int)nx, (int)ny, (int)nz);
    const dim3 gr2 = tgrid(h, ST_FLUX_I + d, (int)(nx + 1) / 2, (int)ny, (int)nz);   // j, k: interface pairs along i
    const WKT<R> w = wkt<R>(h);
    const R *nB = P<R>(h->B[S]), *hB = P<R>(h->hB), *hmag = P<R>(h->hmag), *hGas = P<R>(h->hGas);
    R *Flx = P<R>(h->Flx);
    StageTimer t(h, ST_FLUX_I + d);
    if (h->corr) launch_flux_corr<RM, BO, MH, BG, R>(h, d, gr, gr2, tl, w, nB, hB, hmag, hGas, Flx);
    else if (d == 0) PLAUNCH(h, (k_flux<0, RM, BO, MH, BG, R>), (b_flux<0, RM, BO, MH, BG, R>), gr, 256, g, tl, h->mp, h->bgv, h->clk, w.Wn, nB, w.Wh, hB, hmag, hGas, Flx, h->corr, h->corr_scale);
    else if (sizeof(R) == sizeof(float)) {   // FP32 bulk: one interface per thread measured as fast as pairs
        if (d == 1) PLAUNCH(h, (k_flux<1, RM, BO, MH, BG, R>), (b_flux<1, RM, BO, MH, BG, R>), gr, 256, g, tl, h->mp, h->bgv, h->clk, w.Wn, nB, w.Wh, hB, hmag, hGas, Flx, h->corr, h->corr_scale);
        else PLAUNCH(h, (k_flux<2, RM, BO, MH, BG, R>), (b_flux<2, RM, BO, MH, BG, R>), gr, 256, g, tl, h->mp, h->bgv, h->clk, w.Wn, nB, w.Wh, hB, hmag, hGas, Flx, h->corr, h->corr_scale);
    }
    else if (d == 1) PLAUNCH(h, (k_flux2<1, RM, BO, MH, BG, R>), (b_flux2<1, RM, BO, MH, BG, R>), gr2, 256, g, tl, h->mp, h->bgv, h->clk, w.Wn, nB, w.Wh, hB, hmag, hGas, Flx, h->corr, h->corr_scale);
    else PLAUNCH(h, (k_flux2<2, RM, BO, MH, BG, R>), (b_flux2<2, RM, BO, MH, BG, R>), gr2, 256, g, tl, h->mp, h->bgv, h->clk, w.Wn, nB, w.Wh, hB, hmag, hGas, Flx, h->corr, h->corr_scale);
}

template <int RM, int BO, int MH, typename R>
static void launch_flux(gam_handle *h, int d) {
    if (MH && h->bgv.on) launch_flux_bg<RM, BO, MH, MH, R>(h, d);   // background terms only with MHD
    else launch_flux_bg<RM, BO, MH, 0, R>(h, d);
}

template <int ED, int RM, int BO, typename R>
static void launch_emf(gam_handle *h) {
    const GVT<R> &g = view<R>(h);
    const int st = ST_EMF_I + ED;
    const WKT<R> w = wkt<R>(h);
    StageTimer t(h, st);
    if (ED == 1)   // J edges: k columns with the k stencil in registers
        PLAUNCH(h, (k_emf_jcol<RM, BO, R>), (b_emf_jcol<RM, BO, R>), tgrid(h, st, g.Nip + 1, g.Njp, cdiv(g.Nkp + 1, EMF_KC)),
                256, g, h->tile[st], h->mp, h->bgv, h->clk, P<R>(h->Vf), P<R>(h->hGas), w.bn, w.bh, P<R>(h->E), h->corr, h->corr_scale);
    else
        PLAUNCH(h, (k_emf<ED, RM, BO, R>), (b_emf<ED, RM, BO, R>), tgrid(h, st, g.Nip + (ED != 0), g.Njp + (ED != 1), g.Nkp + (ED != 2)),
                256, g, h->tile[st], h->mp, h->bgv, h->clk, P<R>(h->Vf), P<R>(h->hGas), w.bn, w.bh, P<R>(h->E), h->corr, h->corr_scale);
}

template <int RM, int BO, typename R>
static void launch_emfs(gam_handle *h) {
    launch_emf<0, RM, BO, R>(h);
    launch_emf<1, RM, BO, R>(h);
    launch_emf<2, RM, BO, R>(h);
}

template <int RM, int BO, int MH, typename R>
static void launch_fluxes(gam_handle *h) {
    launch_flux<RM, BO, MH, R>(h, 0);
    launch_flux<RM, BO, MH, R>(h, 1);
    if (!h->mp.do25D) launch_flux<RM, BO, MH, R>(h, 2);
}

// bFlux2Fld of magFlux (type TM) into the bulk array B
template <typename R, typename TM>
static void launch_b2f(gam_handle *h, int st, const TM *mag, void *B) {
    const GVT<R> &g = view<R>(h);
    PLAUNCH(h, (k_b2f<R, TM>), (b_b2f<R, TM>), tgrid(h, st, g.Ni, g.Nj, g.Nk), 256, g, h->tile[st],
            h->clk, mag, P<R>(B));
    if (h->ring.on)   // RingFlux2FieldFix (gridutils.F90:298-300)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ring_bfix<R>), grid1((long)g.Ni * g.Nk * g.nG * 2), 256, 0, h->stream, g,
                           h->clk, P<R>(B));
}

// RingAverage (ringav.F90:25-73) on State, after the update: smoothing EMF,
// its E2Flux, CleanLoops, bFlux2Fld, SmoothGas
template <typename R>
static void ring_average(gam_handle *h) {
    const GVT<R> &g = view<R>(h);
    const GV &g64 = h->gv;
    const int S = h->s;
    const RingP &rp = h->ring;
    StageTimer t(h, ST_RING);
    const long nsr = (long)(g.Nip + 1) * 2 * rp.nring * 2;   // (slice, d, ring, pole)
    hipLaunchKernelGGL(k_ring_se_coef, grid1(nsr), 256, 0, h->stream, g64, rp, h->clk, (const double *)h->ringTab,
                       (const double *)h->mag[S], h->ringMean);
    hipLaunchKernelGGL(k_ring_se_mean, grid1(nsr * (g.Nkp / 2)), 256, 0, h->stream, g64, rp, h->clk,
                       (const double *)h->ringTab, (const double *)h->mag[S], h->ringMean);
    hipLaunchKernelGGL(k_ring_se_emf, grid1(nsr * (g.Nkp / 2)), 256, 0, h->stream, g64, rp, h->clk,
                       (const double *)h->ringTab, (const double *)h->mag[S], (const double *)h->ringMean, h->ringSE);
    hipLaunchKernelGGL(k_ring_ct, grid1((long)(g.Nip + 1) * (g.Nkp + 1) * 2 * (rp.nring + 2)), 256, 0, h->stream, g64, rp,
                       h->clk, h->ringSE, h->mag[S]);
    hipLaunchKernelGGL(k_ring_loops, dim3((unsigned)((g.Nip + 1) * 2)), 256, 0, h->stream, g64, h->clk, h->mag[S]);
    const int nb = rp.nring + 3;   // bFlux2Fld where faces changed, then RingFlux2FieldFix
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_b2f_rows<R>), grid1((long)g.Ni * g.Nk * 2 * nb), 256, 0, h->stream, g, h->clk, nb,
                       (const double *)h->mag[S], P<R>(h->B[S]));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ring_bfix<R>), grid1((long)g.Ni * g.Nk * g.nG * 2), 256, 0, h->stream, g, h->clk,
                       P<R>(h->B[S]));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ring_gas_mean<R>), grid1((long)g.Nip * (g.Nkp / 2) * rp.nring * 2), 256, 0,
                       h->stream, g, h->mp, rp, h->clk, (const R *)P<R>(h->Gas[S]), h->ringMean);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ring_gas_set<R>), grid1((long)g.Nip * g.Nkp * rp.nring * 2), 256, 0, h->stream, g,
                       h->mp, rp, h->clk, (const double *)h->ringMean, P<R>(h->Gas[S]));
}

// AdvanceMHD + UpdateStateData time update; the dt min of the fused path is
// left in h->partial when fuse_dt != 0.
template <typename R>
static int advance_t(gam_handle *h, int fuse_dt) {
    const GVT<R> &g = view<R>(h);
    const MP &m = h->mp;
    const int S = h->s, O = 1 - h->s;
    { StageTimer t(h, ST_CLOCK); PLAUNCH(h, (k_clock_begin), (b_clock_begin), dim3(1), 64, h->clk); }
    { StageTimer t(h, ST_PREDICT);
      PLAUNCH(h, (k_predict<R>), (b_predict<R>), grid1(g.nf), 256, g, m, h->clk, P<R>(h->Gas[S]), P<R>(h->Gas[O]),
              h->mag[S], h->mag[O], P<R>(h->hGas), P<R>(h->hmag), wkt<R>(h)); }
    { StageTimer t(h, ST_BHALF); launch_b2f<R, R>(h, ST_BHALF, P<R>(h->hmag), h->hB); }
    if (m.doMHD) {
        { StageTimer t(h, ST_MOM2FACE);
          PLAUNCH(h, (k_vf<R>), (b_vf<R>), tgrid(h, ST_MOM2FACE, g.Ni, g.Nj, cdiv(g.Nk, VF_KC)), 256, g,
                  h->tile[ST_MOM2FACE], h->clk, P<R>(h->wk[W_VDV]), P<R>(h->Vf)); }
        if (h->corr && h->corr_wait) HIPCHK(h, hipStreamWaitEvent(h->stream, h->corr_wait, 0));
        if (m.rmeth == 1) { if (m.doBoris) launch_emfs<1, 1, R>(h); else launch_emfs<1, 0, R>(h); }
        else { if (m.doBoris) launch_emfs<0, 1, R>(h); else launch_emfs<0, 0, R>(h); }
    } else if (h->corr && h->corr_wait) HIPCHK(h, hipStreamWaitEvent(h->stream, h->corr_wait, 0));
    if (m.doMHD) {
        if (m.rmeth == 1) { if (m.doBoris) launch_fluxes<1, 1, 1, R>(h); else launch_fluxes<1, 0, 1, R>(h); }
        else { if (m.doBoris) launch_fluxes<0, 1, 1, R>(h); else launch_fluxes<0, 0, 1, R>(h); }
    } else {
        if (m.rmeth == 1) launch_fluxes<1, 0, 0, R>(h); else launch_fluxes<0, 0, 0, R>(h);
    }
    if (h->ring.on)   // RingFlux (stress.F90:137-139)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ring_flux<R>), grid1((long)g.Ni * g.Nk * 2 + (long)g.Nip * g.Njp), 256, 0,
                           h->stream, g, h->clk, m.doMHD, P<R>(h->Flx));
    if (h->ring.on && m.doMHD) {   // PoleE + FixEFieldLFM (mhdgroup.F90:138-142)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pole_e<R>), dim3((unsigned)((g.Nip + 1) * 2)), 256, 0, h->stream, g, h->clk,
                           P<R>(h->E));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fix_e<R>), grid1((long)(g.Nip + 1) * (g.Njp + 1)), 256, 0, h->stream, g,
                           h->clk, P<R>(h->E));
    }
    {
        StageTimer t(h, ST_CT);
        if (m.doMHD) PLAUNCH(h, (k_ct<R>), (b_ct<R>), grid1(g.nf), 256, g, h->clk, h->mag[S], P<R>(h->E), h->mag[O]);
        else HIPCHK(h, hipMemcpyAsync(h->mag[O], h->mag[S], 3 * fsz(h) * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    }
    {
        StageTimer t(h, ST_UPDATE);
        dim3 gr(update_blocks(h));
        const Tile tu = h->tile[ST_UPDATE];
        const R *nGas = P<R>(h->Gas[S]), *nB = P<R>(h->B[S]), *Flx = P<R>(h->Flx);
        R *oGas = P<R>(h->Gas[O]), *oB = P<R>(h->B[O]);
        const int skip = h->ring.on ? h->ring.nring + 3 : 0;
        // FP64: the k-column kernel; FP32 bulk: one cell per thread
        if constexpr (sizeof(R) == sizeof(double)) {
            if (m.doMHD) {
                if (m.doBoris) PLAUNCH(h, (k_update_col<1, 1, R>), (b_update_col<1, 1, R>), gr, 256, g, tu, m, h->bgv, h->clk, nGas, nB, Flx, h->mag[O], oGas, oB, h->partial, skip);
                else PLAUNCH(h, (k_update_col<0, 1, R>), (b_update_col<0, 1, R>), gr, 256, g, tu, m, h->bgv, h->clk, nGas, nB, Flx, h->mag[O], oGas, oB, h->partial, skip);
            } else {
                PLAUNCH(h, (k_update_col<0, 0, R>), (b_update_col<0, 0, R>), gr, 256, g, tu, m, h->bgv, h->clk, nGas, nB, Flx, h->mag[O], oGas, oB, h->partial, skip);
            }
        } else {
          if (m.doMHD) {
            if (m.doBoris) PLAUNCH(h, (k_update<1, 1, R>), (b_update<1, 1, R>), gr, 256, g, tu, m, h->bgv, h->clk, nGas, nB, Flx, h->mag[O], oGas, oB, h->partial, skip);
            else PLAUNCH(h, (k_update<0, 1, R>), (b_update<0, 1, R>), gr, 256, g, tu, m, h->bgv, h->clk, nGas, nB, Flx, h->mag[O], oGas, oB, h->partial, skip);
        } else {
            PLAUNCH(h, (k_update<0, 0, R>), (b_update<0, 0, R>), gr, 256, g, tu, m, h->bgv, h->clk, nGas, nB, Flx, h->mag[O], oGas, oB, h->partial, skip);
          }
        }
    }
    // CopyState State->oState is the buffer swap
    h->s = O;
    h->fused_np = update_blocks(h);
    if (h->ring.on) {   // mhdgroup.F90:69-74; CalcDT then sees the averaged rings
        ring_average<R>(h);
        if (fuse_dt) {   // the ring band's CFL minima after k_update's
            const int nb = h->ring.nring + 3;
            StageTimer t(h, ST_CELLDT);
            const dim3 gr(ring_dt_blocks(h));
            const R *Gs = P<R>(h->Gas[h->s]), *Bs = P<R>(h->B[h->s]);
            double *pt = h->partial + h->fused_np;
            if (h->mp.doBoris)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_celldt_rows<1, R>), gr, 256, 0, h->stream, g, m, h->bgv, h->clk, nb, Gs, Bs, pt);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_celldt_rows<0, R>), gr, 256, 0, h->stream, g, m, h->bgv, h->clk, nb, Gs, Bs, pt);
            h->fused_np += ring_dt_blocks(h);
        }
    }
    if (!fuse_dt) launch_dtfinal(h, /*advance*/ 1, /*calc*/ 0, 0);
    HIPCHK(h, hipGetLastError());
    return 0;
}
static int advance(gam_handle *h, int fuse_dt) {
    return h->prec ? advance_t<float>(h, fuse_dt) : advance_t<double>(h, fuse_dt);
}

static bool all_periodic(const gam_handle *h) {
    for (int f = 0; f < 6; f++) if (h->bct[f] != GAM_BC_PERIODIC) return false;
    return true;
}

// BCs on State, in the reference's order (INI, OUTI, INJ, OUTJ, INK, OUTK);
// when withB, Bxyz ghosts are written too (hatch).  All-periodic: one pass
// with the composed per-axis maps; otherwise per axis, a periodic pass or the
// two faces' own BCs one after the other.
template <typename R>
static void launch_bc(gam_handle *h, bool withB) {
    const GVT<R> &g = view<R>(h);
    const int S = h->s;
    R *Bw = withB ? P<R>(h->B[S]) : (R *)nullptr;
    if (all_periodic(h)) {
        PLAUNCH(h, (k_bc<R>), (b_bc<R>), grid1(bc_shell_count(bc_shell(g.Ni, g.Nj, g.Nk, g.Nip, g.Njp, g.Nkp, g.nG, 7, h->mp.do25D))),
                256, g, h->clk, h->mp.do25D, 7, P<R>(h->Gas[S]), h->mag[S], Bw);
        return;
    }
    const int ext[3] = {g.Ni, g.Nj, g.Nk};
    // one face's own (non-periodic) BC
    auto face_bc = [&](int ax, int side) {
        const int t = h->bct[2 * ax + side];
        const long n = (long)g.nG * ext[ax == 0 ? 1 : 0] * ext[ax == 2 ? 1 : 2];
        if (t == GAM_BC_LAZY) return;   // lazyBC (bcs.F90:268-275): nothing to do
        if (t == GAM_BC_CYLPOLE) {       // cylpole, the inner i face (bcs.F90:987-1017)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bc_cylpole<R>), grid1(n), 256, 0, h->stream, g, h->clk, (int)withB,
                               P<R>(h->Gas[S]), h->mag[S], P<R>(h->B[S]));
            return;
        }
        if (t == GAM_BC_ION) {           // IonInner, uncoupled (voltron/ICs/earthcmi.F90:378-490)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bc_ion<R>), grid1((long)(g.Nj + 1) * (g.Nk + 1) * g.nG), 256, 0,
                               h->stream, g, h->mp, h->clk, h->ring.on, P<R>(h->Gas[S]), h->mag[S]);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wind_bxyz<R>), grid1((long)g.Nj * g.Nk * g.nG), 256, 0, h->stream, g,
                               h->clk, h->ring.on, 1, (const double *)h->mag[S], P<R>(h->B[S]));
            if (h->ring.on)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wind_bfix<R>), grid1((long)g.nG * g.Nk * g.nG * 2), 256, 0,
                                   h->stream, g, h->clk, 1, P<R>(h->B[S]));
            return;
        }
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bc_face<R>), grid1(n), 256, 0, h->stream, g, h->clk, 2 * ax + side, t,
                           (int)withB, (const double *)h->grid[G_FACE], (const double *)h->grid[G_TF],
                           P<R>(h->Gas[S]), h->mag[S], P<R>(h->B[S]));
    };
    for (int ax = 0; ax < 3; ax++) {
        if (h->bct[2 * ax] == GAM_BC_PERIODIC) {
            const long nsh = bc_shell_count(bc_shell(g.Ni, g.Nj, g.Nk, g.Nip, g.Njp, g.Nkp, g.nG, 1 << ax, h->mp.do25D));
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bc<R>), grid1(nsh), 256, 0, h->stream, g, h->clk, h->mp.do25D, 1 << ax,
                               P<R>(h->Gas[S]), h->mag[S], Bw);
            continue;
        }
        if (ax == 0 && h->bct[1] == GAM_BC_WIND) {   // the inner face's BC, then WindBC (apps/wind.F90:206-374)
            face_bc(0, 0);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bc_wind<R>), grid1((long)(g.Nj + 1) * (g.Nk + 1) * g.nG), 256, 0, h->stream, g,
                               h->mp, h->wind, h->clk, (const double *)h->grid[G_XYZCC], (const double *)h->grid[G_XFC],
                               (const double *)h->grid[G_FACE], (const double *)h->grid[G_TF], P<R>(h->Gas[S]), h->mag[S],
                               P<R>(h->B[S]));
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wind_bxyz<R>), grid1((long)g.Nj * g.Nk * g.nG), 256, 0, h->stream, g, h->clk,
                               h->ring.on, 0, (const double *)h->mag[S], P<R>(h->B[S]));
            if (h->ring.on)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_wind_bfix<R>), grid1((long)g.nG * g.Nk * g.nG * 2), 256, 0, h->stream, g,
                                   h->clk, 0, P<R>(h->B[S]));
            continue;
        }
        if (h->bct[2 * ax] == GAM_BC_LFM) {   // lfmIn then lfmOut (bcs.F90:1020-1106)
            for (int side = 0; side < 2; side++)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bc_lfm<R>), grid1((long)(g.Ni + 1) * (g.Nk + 1) * g.nG), 256, 0,
                                   h->stream, g, h->clk, side, P<R>(h->Gas[S]), h->mag[S], P<R>(h->B[S]));
            continue;
        }
        for (int side = 0; side < 2; side++) face_bc(ax, side);
    }
}

extern "C" int gam_set_bcs(gam_handle *h, const int *type) {
    if (!h || !type) return -1;
    if (refuse_if_captured(h, "gam_set_bcs")) return -1;
    for (int ax = 0; ax < 3; ax++) {
        const int a = type[2 * ax], b = type[2 * ax + 1];
        if (a < 0 || a > GAM_BC_ION || b < 0 || b > GAM_BC_ION)
            return fail(h, -1, "gam_set_bcs: unknown boundary type");
        if ((a == GAM_BC_ION && ax != 0) || b == GAM_BC_ION)
            return fail(h, -1, "gam_set_bcs: the ionospheric BC is the inner i face's BC");
        if ((a == GAM_BC_CYLPOLE && ax != 0) || b == GAM_BC_CYLPOLE)
            return fail(h, -1, "gam_set_bcs: the cylindrical pole BC is the inner i face's BC");
        if (a == GAM_BC_WIND || (b == GAM_BC_WIND && ax != 0))
            return fail(h, -1, "gam_set_bcs: the solar-wind BC is the outer i face's BC");
        if ((a == GAM_BC_LFM || b == GAM_BC_LFM) && (ax != 1 || a != b))
            return fail(h, -1, "gam_set_bcs: the LFM axis BCs are the two j faces' BC");
        if (ax > 0 && (a == GAM_BC_SPHERE || b == GAM_BC_SPHERE))
            return fail(h, -1, "gam_set_bcs: the spherical BCs are i-face BCs");
        if ((a == GAM_BC_PERIODIC) != (b == GAM_BC_PERIODIC))
            return fail(h, -1, "gam_set_bcs: an axis is periodic on both faces or on neither");
    }
    if (h->mp.do25D && type[4] != GAM_BC_PERIODIC)
        return fail(h, -1, "gam_set_bcs: 2.5D blocks keep the k boundaries periodic");
    if (type[2] == GAM_BC_LFM && (type[4] != GAM_BC_PERIODIC || h->gv.Nkp < 16 || (h->gv.Nkp & 1)))
        return fail(h, -1, "gam_set_bcs: the LFM axis BCs need k periodic over an even Nkp >= 16");
    if (type[0] == GAM_BC_CYLPOLE && (type[2] != GAM_BC_PERIODIC || h->gv.Njp < 8 || (h->gv.Njp & 1) || type[1] == GAM_BC_WIND))
        return fail(h, -1, "gam_set_bcs: the cylindrical pole BC needs j periodic over an even Njp >= 8 (no wind)");
    if (type[0] == GAM_BC_ION && type[4] != GAM_BC_PERIODIC)
        return fail(h, -1, "gam_set_bcs: the ionospheric BC needs k periodic");
    if (type[1] == GAM_BC_WIND && (!h->wind.on || type[4] != GAM_BC_PERIODIC || type[0] == GAM_BC_PERIODIC))
        return fail(h, -1, "gam_set_bcs: the solar-wind BC needs gam_set_wind first, k periodic and a non-periodic i axis");
    if (h->corr && !(type[0] == 0 && type[2] == 0 && type[4] == 0))
        return fail(h, -1, "gam_set_bcs: a learned correction is set (periodic boundaries only)");
    int rc = gam_sync(h);
    if (rc) return rc;
    for (int f = 0; f < 6; f++) h->bct[f] = type[f];
    return 0;
}

extern "C" int gam_get_bcs(gam_handle *h, int *type) {
    if (!h || !type) return -1;
    for (int f = 0; f < 6; f++) type[f] = h->bct[f];
    return 0;
}

extern "C" int gam_set_wind(gam_handle *h, double rho, double p, const double *v, const double *b) {
    if (!h || !v || !b) return -1;
    if (refuse_if_captured(h, "gam_set_wind")) return -1;
    if (!(rho > 0) || !(p > 0) || !(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 0))
        return fail(h, -1, "gam_set_wind: density, pressure and speed must be positive");
    int rc = gam_sync(h);
    if (rc) return rc;
    if (h->windSer) HIPCHK(h, hipFree(h->windSer));
    h->windSer = nullptr;
    h->wind = WindP{};
    h->wind.rho = rho; h->wind.p = p;
    for (int x = 0; x < 3; x++) { h->wind.v[x] = v[x]; h->wind.b[x] = b[x]; }
    h->wind.on = 1;
    return 0;
}

// readWind (wind.F90:617-760) after its unit conversion, and InitWind's
// discrete series (wind.F90:69-111): the checks it makes (uniform spacing to
// 1e-6, the wind slower than the Boris speed), Bx replaced by
// Bx0 + ByC By + BzC Bz, the bow-shock marker xyzBS zero when not given.
extern "C" int gam_set_wind_series(gam_handle *h, int nt, const double *t, const double *rho, const double *p,
                                   const double *v, const double *b, const double *xyzBS, double ByC, double BzC,
                                   double Bx0) {
    if (!h || !t || !rho || !p || !v || !b) return -1;
    if (refuse_if_captured(h, "gam_set_wind_series")) return -1;
    if (nt < 2) return fail(h, -1, "gam_set_wind_series: the series needs at least two samples");
    double dtMin = t[1] - t[0], dtMax = dtMin, tMin = t[0], tMax = t[0], vMax = 0.0;
    for (int n = 0; n < nt; n++) {
        if (n > 0) { dtMin = fmin(dtMin, t[n] - t[n - 1]); dtMax = fmax(dtMax, t[n] - t[n - 1]); }
        tMin = fmin(tMin, t[n]); tMax = fmax(tMax, t[n]);
        const double vx = v[n], vy = v[nt + n], vz = v[2 * nt + n];
        vMax = fmax(vMax, sqrt(vx * vx + vy * vy + vz * vz));
        if (!(rho[n] > 0) || !(p[n] > 0) || !(vx * vx + vy * vy + vz * vz > 0))
            return fail(h, -1, "gam_set_wind_series: density, pressure and speed must be positive");
    }
    if (!(dtMin > 0)) return fail(h, -1, "gam_set_wind_series: the sample times must increase");
    if (fabs(dtMin - dtMax) > 1.0e-6 * sqrt(dtMin * dtMin + dtMax * dtMax))
        return fail(h, -1, "gam_set_wind_series: the series requires uniform spacing");
    if (h->mp.doBoris && vMax >= h->mp.Ca)
        return fail(h, -1, "gam_set_wind_series: the solar wind is faster than the Boris speed of light");
    std::vector<double> tab((size_t)nt * WIND_COLS);
    for (int n = 0; n < nt; n++) {
        double *r = &tab[(size_t)WIND_COLS * n];
        r[0] = t[n]; r[1] = rho[n]; r[2] = p[n];
        for (int x = 0; x < 3; x++) {
            const size_t q = (size_t)nt * x + n;
            r[3 + x] = v[q]; r[6 + x] = b[q]; r[9 + x] = xyzBS ? xyzBS[q] : 0.0;
        }
        r[6] = Bx0 + ByC * r[7] + BzC * r[8];
    }
    int rc = gam_sync(h);
    if (rc) return rc;
    if (h->windSer) HIPCHK(h, hipFree(h->windSer));
    h->windSer = nullptr;
    h->wind = WindP{};   // off until the new table is in place
    HIPCHK(h, hipMalloc(&h->windSer, tab.size() * sizeof(double)));
    HIPCHK(h, hipMemcpy(h->windSer, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    WindP w = {};
    w.rho = rho[0]; w.p = p[0];
    for (int x = 0; x < 3; x++) { w.v[x] = tab[3 + x]; w.b[x] = tab[6 + x]; }
    w.on = 1; w.nt = nt; w.ser = h->windSer;
    w.tMin = tMin; w.tMax = tMax; w.dtW = dtMin; w.ByC = ByC; w.BzC = BzC;
    h->wind = w;
    return 0;
}

extern "C" int gam_set_ring(gam_handle *h, int nring, const int *nch) {
    if (!h) return -1;
    if (refuse_if_captured(h, "gam_set_ring")) return -1;
    const GV &g = h->gv;
    int rc = gam_sync(h);
    if (rc) return rc;
    if (h->corners_set) {   // RingGridFix scaled the geometry by these rings and chunks
        bool same = nring == (h->ring.on ? h->ring.nring : 0);
        for (int r = 0; same && r < nring; r++) same = nch && nch[r] == h->ring.nch[r];
        if (!same)
            return fail(h, -1, "gam_set_ring: call it before gam_set_corners (RingGridFix is part of the geometry)");
    }
    if (nring == 0) { h->ring.on = 0; return 0; }
    if (nring < 0 || nring > GAM_RING_MAX || !nch) return fail(h, -1, "gam_set_ring: 1 to %d rings", GAM_RING_MAX);
    if (h->mp.do25D || g.Nkp < 16 || g.Nkp > GAM_RING_NPMAX || (g.Nkp & 1))
        return fail(h, -1, "gam_set_ring: Nkp must be even, 16 .. %d", GAM_RING_NPMAX);
    if (g.Njp < 2 * nring + 4) return fail(h, -1, "gam_set_ring: Njp must be >= 2*nring + 4");
    for (int r = 0; r < nring; r++)
        if (nch[r] < 1 || nch[r] > g.Nkp / 2 || g.Nkp % nch[r])
            return fail(h, -1, "gam_set_ring: ring %d: chunks must divide Nkp, at most Nkp/2", r + 1);
    if (h->corr) return fail(h, -1, "gam_set_ring: a learned correction is set (periodic boundaries only)");
    const size_t nse = 3 * gsz(h);
    if (!h->ringSE) {
        HIPCHK(h, hipMalloc(&h->ringSE, nse * sizeof(double)));
        HIPCHK(h, hipMalloc(&h->ringTab, 2 * (size_t)GAM_RING_NPMAX * sizeof(double)));
    }
    HIPCHK(h, hipMemset(h->ringSE, 0, nse * sizeof(double)));   // never written off the ring rows
    // ring scratch: SmoothGas chunk means, or CalcSmoothE chunk means + Fourier coefficients
    const size_t hc = g.Nkp / 2;
    const size_t nmean = std::max((size_t)2 * nring * 5 * hc * g.Nip, (size_t)2 * nring * 2 * (hc + 4) * (g.Nip + 1));
    if (nmean > h->ringMean_n) {
        if (h->ringMean) HIPCHK(h, hipFree(h->ringMean));
        h->ringMean = nullptr;
        h->ringMean_n = 0;
        HIPCHK(h, hipMalloc(&h->ringMean, nmean * sizeof(double)));
        h->ringMean_n = nmean;
    }
    // ringav.F90:447-451: phi_n = dp*n - dp/2, n = 1..Np, dp = 2 pi/Np
    std::vector<double> tab(2 * (size_t)g.Nkp);
    const double dp = 2.0 * M_PI / g.Nkp;
    for (int n = 1; n <= g.Nkp; n++) {
        const double phi = dp * n - 0.5 * dp;
        tab[n - 1] = cos(1.0 * phi);
        tab[g.Nkp + n - 1] = sin(1.0 * phi);
    }
    HIPCHK(h, hipMemcpy(h->ringTab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    h->ring.on = 1;
    h->ring.nring = nring;
    for (int r = 0; r < GAM_RING_MAX; r++) h->ring.nch[r] = r < nring ? nch[r] : 0;
    return ensure_partial(h, update_blocks(h) + ring_dt_blocks(h));
}

extern "C" int gam_get_ring(gam_handle *h, int *nring, int *nch) {
    if (!h || !nring) return -1;
    *nring = h->ring.on ? h->ring.nring : 0;
    if (nch) for (int r = 0; r < *nring; r++) nch[r] = h->ring.nch[r];
    return 0;
}

// shell: after a fused step's update only the ghost cells' Bxyz can change
// (k_b2f_shell); the stand-alone EnforceBCs recomputes every cell.
template <typename R>
static int bcs_t(gam_handle *h, bool shell) {
    const int S = h->s;
    { StageTimer t(h, ST_BC); launch_bc<R>(h, false); }
    {
        StageTimer t(h, ST_BPOST);
        const GVT<R> &g = view<R>(h);
        if (shell && !h->mp.do25D) {
            const long n = b2f_shell_count(g.Ni, g.Nj, g.Nk, g.Nip, g.Njp, g.Nkp);
            PLAUNCH(h, (k_b2f_shell<R>), (b_b2f_shell<R>), dim3((unsigned)((n + 255) / 256)), 256, g, h->clk,
                    (const double *)h->mag[S], P<R>(h->B[S]));
            if (h->ring.on)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_ring_bfix<R>), grid1((long)g.Ni * g.Nk * g.nG * 2), 256, 0, h->stream, g,
                                   h->clk, P<R>(h->B[S]));
        } else {
            launch_b2f<R, double>(h, ST_BPOST, h->mag[S], h->B[S]);
        }
    }
    HIPCHK(h, hipGetLastError());
    return 0;
}
static int bcs(gam_handle *h, bool shell = false) { return h->prec ? bcs_t<float>(h, shell) : bcs_t<double>(h, shell); }

extern "C" int gam_advance_mhd(gam_handle *h) {
    if (!h) return -1;
    int rc = advance(h, 0);
    if (rc) return rc;
    return gam_sync(h);
}

extern "C" int gam_calc_dt(gam_handle *h, double *dt, int *status) {
    if (!h) return -1;
    int rc = celldt(h, h->Gas[h->s], h->B[h->s]);
    if (rc) return rc;
    launch_dtfinal(h, 0, 1, h->npart);
    launch_cpr(h);
    gam_clock c;
    rc = gam_get_clock(h, &c);
    if (rc) return rc;
    if (dt) *dt = c.dt;
    if (status) *status = c.status;
    return c.status;
}

extern "C" int gam_enforce_bcs(gam_handle *h) {
    if (!h) return -1;
    int rc = bcs(h);
    if (rc) return rc;
    return gam_sync(h);
}

extern "C" int gam_step(gam_handle *h, int nsteps) {
    if (!h || nsteps < 0) return -1;
    HIPCHK(h, hipMemcpyAsync(h->clk_host, h->clk, sizeof(DClock), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int ts0 = h->clk_host->ts;
    const int s0 = h->s;
    for (int n = 0; n < nsteps; n++) {
        int rc = advance(h, 1);
        if (rc) return rc;
        launch_dtfinal(h, 1, 1, h->fused_np);   // block minima left in h->partial by k_update (or CellDT)
        launch_cpr(h);
        if ((rc = bcs(h, true))) return rc;
    }
    gam_clock c;
    int rc = gam_get_clock(h, &c);
    if (rc) return rc;
    // steps after a halt launched but did nothing: undo their buffer swaps
    int done = c.ts - ts0;
    if (((s0 + done) & 1) != h->s) h->s = 1 - h->s;
    return c.status;
}

extern "C" int gam_hatch(gam_handle *h) {
    // init.F90:204-244: Bxyz from fluxes, initial BCs (Gas, Bxyz and magFlux
    // ghosts copied; no Bxyz recompute after them), oState = State, first dt
    // (ts = 0), oState%time = State%time - dt.
    if (!h) return -1;
    const int S = h->s, O = 1 - h->s;
    gam_clock c0;
    int rc = gam_get_clock(h, &c0);
    if (rc) return rc;
    c0.dt0 = 0.0; c0.ts = 0; c0.status = 0;
    if ((rc = gam_set_clock(h, &c0))) return rc;
    if (h->prec) { launch_b2f<float, double>(h, ST_BPOST, h->mag[S], h->B[S]); launch_bc<float>(h, true); }
    else { launch_b2f<double, double>(h, ST_BPOST, h->mag[S], h->B[S]); launch_bc<double>(h, true); }
    HIPCHK(h, hipMemcpyAsync(h->Gas[O], h->Gas[S], 5 * gsz(h) * h->es, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->mag[O], h->mag[S], 3 * fsz(h) * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->B[O], h->B[S], 3 * gsz(h) * h->es, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipGetLastError());
    if ((rc = celldt(h, h->Gas[S], h->B[S]))) return rc;
    launch_dtfinal(h, 0, 1, h->npart);
    gam_clock c;
    if ((rc = gam_get_clock(h, &c))) return rc;
    c.oStateTime = c.stateTime - c.dt;
    int st = c.status;
    if ((rc = gam_set_clock(h, &c))) return rc;
    return st;
}

extern "C" int gam_divb(gam_handle *h, double *tot) {
    if (!h || !tot) return -1;
    const GV &g = h->gv;
    long n = (long)g.Nip * g.Njp * g.Nkp;
    int nb = (int)((n + 255) / 256);
    if (nb > h->pcap) return fail(h, -1, "gam_divb: partial buffer too small");
    hipLaunchKernelGGL(k_divb, nb, 256, 0, h->stream, g, h->mag[h->s], h->partial);
    HIPCHK(h, hipGetLastError());
    std::vector<double> hp(nb);
    HIPCHK(h, hipMemcpyAsync(hp.data(), h->partial, nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    int rc = gam_sync(h);
    if (rc) return rc;
    double s = 0.0;
    for (double v : hp) s += v;
    *tot = s;
    return 0;
}

extern "C" int gam_divb_noalloc(gam_handle *h, double *tot) {
    if (!h || !tot) return -1;
    const GV &g = h->gv;
    const long n = (long)g.Nip * g.Njp * g.Nkp;
    const long nb = (n + 255) / 256;
    if (nb > GAM_NOALLOC_BLOCKS)
        return fail(h, -1, "gam_divb_noalloc: %ld active cells need %ld block sums, more than the %d of its scratch array "
                    "(at most %ld cells; use gam_divb)", n, nb, GAM_NOALLOC_BLOCKS, (long)GAM_NOALLOC_BLOCKS * 256);
    double *sc = h->scratch;   // this GPU's copy of `scratch`
    hipLaunchKernelGGL(k_divb, dim3((unsigned)nb), 256, 0, h->stream, g, (const double *)h->mag[h->s], sc);
    hipLaunchKernelGGL(k_sum_ordered, 1, 256, 0, h->stream, (const double *)sc, (int)nb, sc + GAM_NOALLOC_BLOCKS);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(tot, sc + GAM_NOALLOC_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    return gam_sync(h);
}

extern "C" int gam_get_work(gam_handle *h, int which, double *out) {
    if (!h || !out) return -1;
    const size_t G = gsz(h);
    int rc;
    if (which == 0) rc = d2h_bulk(h, out, h->Flx, 24 * G);
    else if (which == 1) rc = d2h_bulk(h, out, h->E, 3 * G);
    else if (which == 2) rc = d2h_bulk(h, out, h->Vf, 9 * G);
    else return fail(h, -1, "gam_get_work: which must be 0, 1 or 2");
    if (rc) return rc;
    return gam_sync(h);
}

// ----------------------------------------------------------------- ensembles
// Several handles stepped by one captured graph.  Each handle's step sequence
// (advance + dt + BCs, all device-resident) is captured on the handle's own
// stream, forked from and joined back to a capture stream, so the blocks are
// independent branches of one graph.  nsteps is even: the State/oState buffer
// pointers baked into the graph are then the same at every replay.
struct gam_graph {
    std::vector<gam_handle *> hs;
    std::vector<int> ts0, s0;
    std::vector<int> s_cap;              // State index of each handle when captured
    std::vector<double *> partial_cap;   // its block-minima buffer baked into the graph
    int nsteps = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> evs;
    void *pk_arena = nullptr;   // packed graphs: the members' argument tables (device)
    std::string err;
};

static int gfail(gam_graph *g, int code, const char *msg, hipError_t e) {
    if (g) g->err = std::string(msg) + ": " + hipGetErrorString(e);
    return code;
}

extern "C" int gam_graph_destroy(gam_graph *g) {
    if (!g) return 0;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->exec)
        for (auto h : g->hs) h->graph_refs--;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (auto e : g->evs) (void)hipEventDestroy(e);
    if (g->pk_arena) (void)hipFree(g->pk_arena);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return 0;
}

extern "C" int gam_graph_create(gam_handle *const *hs, int n, int nsteps, gam_graph **out) {
    if (!hs || n < 1 || !out || nsteps < 2 || (nsteps & 1)) return -1;
    *out = nullptr;
    for (int q = 0; q < n; q++) {
        if (!hs[q]) return -1;
        for (int r = 0; r < q; r++) if (hs[r] == hs[q]) return -1;   // one branch per handle
        int rc = gam_sync(hs[q]);
        if (rc) return rc;
        if (hs[q]->timing) return fail(hs[q], -1, "gam_graph_create: stage timers must be off");
        if (hs[q]->corr_wait)
            return fail(hs[q], -1, "gam_graph_create: the correction waits for an event of another stream (gam_set_correction_wait)");
    }
    gam_graph *g = new gam_graph();
    g->hs.assign(hs, hs + n);
    g->nsteps = nsteps;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking)) != hipSuccess) {
        gam_graph_destroy(g);
        return -2;
    }
    g->evs.resize(2 * n + 1, nullptr);
    for (auto &ev : g->evs)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) { gam_graph_destroy(g); return -2; }
    if ((e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        gam_graph_destroy(g);
        return -2;
    }
    int rc = 0;
    if ((e = hipEventRecord(g->evs[0], g->stream)) != hipSuccess) rc = -2;
    for (int q = 0; q < n && !rc; q++) {
        gam_handle *h = hs[q];
        if ((e = hipStreamWaitEvent(h->stream, g->evs[0], 0)) != hipSuccess) { rc = -2; break; }
        for (int k = 0; k < nsteps && !rc; k++) {
            rc = advance(h, 1);
            if (!rc) launch_dtfinal(h, 1, 1, h->fused_np);
            if (!rc) launch_cpr(h);
            if (!rc) rc = bcs(h, true);
        }
        if (!rc && (e = hipEventRecord(g->evs[1 + q], h->stream)) != hipSuccess) rc = -2;
        if (!rc && (e = hipStreamWaitEvent(g->stream, g->evs[1 + q], 0)) != hipSuccess) rc = -2;
    }
    hipGraph_t graph = nullptr;
    hipError_t ee = hipStreamEndCapture(g->stream, &graph);
    g->graph = graph;
    if (rc || ee != hipSuccess) {
        if (!rc) e = ee;
        gam_graph_destroy(g);
        for (int q = 0; q < n; q++) (void)hipGetLastError();
        return rc ? rc : -2;
    }
    if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess) {
        gam_graph_destroy(g);
        return -2;
    }
    g->ts0.assign(n, 0);
    g->s0.assign(n, 0);
    for (int q = 0; q < n; q++) {   // nsteps is even: every handle is back at its capture-time State index
        g->s_cap.push_back(hs[q]->s);
        g->partial_cap.push_back(hs[q]->partial);
        hs[q]->graph_refs++;
    }
    *out = g;
    return 0;
}

// Every handle's nsteps fused steps with one launch per kernel for all of
// them: each handle's launches are recorded (not run), and launches of one
// kernel by several handles become one packed launch (k_packed) reading each
// handle's arguments from a device table.  Handles with periodic boundaries
// and MHD, without ring average or correction.
extern "C" int gam_graph_create_packed(gam_handle *const *hs, int n, int nsteps, gam_graph **out) {
    if (!hs || n < 1 || !out || nsteps < 2 || (nsteps & 1)) return -1;
    *out = nullptr;
    for (int q = 0; q < n; q++) {
        gam_handle *h = hs[q];
        if (!h) return -1;
        for (int r = 0; r < q; r++) if (hs[r] == h) return -1;
        int rc = gam_sync(h);
        if (rc) return rc;
        if (h->timing) return fail(h, -1, "gam_graph_create_packed: stage timers must be off");
        if (!all_periodic(h) || h->ring.on || h->corr || !h->mp.doMHD)
            return fail(h, -1, "gam_graph_create_packed: periodic MHD blocks only (no ring average, no correction)");
    }
    std::vector<std::vector<std::unique_ptr<PackRec>>> recs(n);
    for (int q = 0; q < n; q++) {
        gam_handle *h = hs[q];
        h->rec = &recs[q];
        int rc = 0;
        for (int k = 0; k < nsteps && !rc; k++) {   // gam_step's sequence; the State swaps return with nsteps even
            rc = advance(h, 1);
            if (!rc) launch_dtfinal(h, 1, 1, h->fused_np);
            if (!rc) launch_cpr(h);
            if (!rc) rc = bcs(h, true);
        }
        h->rec = nullptr;
        if (rc) return rc;
    }
    // Merge the handles' launch sequences: each packed launch takes, of the
    // kernel that most handles launch next, the next launch of every handle
    // that launches it next; every handle's launches keep their order, and
    // handles with the same sequence (the same kind of block) share every
    // launch.
    struct Group { std::vector<const PackRec *> r; };
    std::vector<Group> plan;
    std::vector<size_t> pos(n, 0);
    for (;;) {
        int best = -1, bestc = 0;
        for (int q = 0; q < n; q++) {
            if (pos[q] >= recs[q].size()) continue;
            const PackRec *c = recs[q][pos[q]].get();
            int cnt = 0;
            for (int r = 0; r < n; r++)
                if (pos[r] < recs[r].size() && recs[r][pos[r]]->key() == c->key() && recs[r][pos[r]]->nt == c->nt) cnt++;
            if (cnt > bestc) { best = q; bestc = cnt; }
        }
        if (best < 0) break;
        const PackRec *c = recs[best][pos[best]].get();
        Group grp;
        for (int r = 0; r < n; r++)
            if (pos[r] < recs[r].size() && recs[r][pos[r]]->key() == c->key() && recs[r][pos[r]]->nt == c->nt)
                grp.r.push_back(recs[r][pos[r]++].get());
        plan.push_back(std::move(grp));
    }
    const size_t L = plan.size();
    // per packed launch: the argument table, then first[] and count[], each 256-byte aligned
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    std::vector<size_t> off(L + 1, 0);
    std::vector<unsigned> tot(L, 0);
    for (size_t p = 0; p < L; p++) {
        const size_t m = plan[p].r.size(), ab = plan[p].r[0]->arg_bytes();
        off[p + 1] = off[p] + al(ab * m) + 2 * al(sizeof(unsigned) * m);
        for (const PackRec *r : plan[p].r)
            if (r->grid.y != 1 || r->grid.z != 1) return fail(hs[0], -1, "gam_graph_create_packed: a launch with a 2-D grid");
    }
    std::vector<unsigned char> host(off[L]);
    for (size_t p = 0; p < L; p++) {
        const size_t m = plan[p].r.size(), ab = plan[p].r[0]->arg_bytes();
        unsigned char *base = host.data() + off[p];
        unsigned *first = (unsigned *)(base + al(ab * m)), *count = (unsigned *)(base + al(ab * m) + al(sizeof(unsigned) * m));
        unsigned long long t = 0;
        for (size_t q = 0; q < m; q++) {
            plan[p].r[q]->put_args(base + ab * q);
            first[q] = (unsigned)t;
            count[q] = plan[p].r[q]->grid.x;
            t += (count[q] + 7ull) / 8 * 8;
        }
        if (t > 0xffffffffull) return fail(hs[0], -1, "gam_graph_create_packed: more than 2^32 blocks in one launch");
        tot[p] = first[m - 1] + count[m - 1];
    }
    gam_graph *g = new gam_graph();
    g->hs.assign(hs, hs + n);
    g->nsteps = nsteps;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMalloc(&g->pk_arena, host.size())) != hipSuccess ||
        (e = hipMemcpy(g->pk_arena, host.data(), host.size(), hipMemcpyHostToDevice)) != hipSuccess) {
        gam_graph_destroy(g);
        return -2;
    }
    if ((e = hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) {
        gam_graph_destroy(g);
        return -2;
    }
    for (size_t p = 0; p < L; p++) {
        if (!tot[p]) continue;
        const size_t m = plan[p].r.size(), ab = plan[p].r[0]->arg_bytes();
        const unsigned char *base = (const unsigned char *)g->pk_arena + off[p];
        plan[p].r[0]->launch(base, (const unsigned *)(base + al(ab * m)),
                             (const unsigned *)(base + al(ab * m) + al(sizeof(unsigned) * m)), (int)m, tot[p], g->stream);
    }
    hipGraph_t graph = nullptr;
    hipError_t ee = hipStreamEndCapture(g->stream, &graph);
    g->graph = graph;
    if (ee != hipSuccess || (e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess) {
        gam_graph_destroy(g);
        (void)hipGetLastError();
        return -2;
    }
    g->ts0.assign(n, 0);
    g->s0.assign(n, 0);
    for (int q = 0; q < n; q++) {
        g->s_cap.push_back(hs[q]->s);
        g->partial_cap.push_back(hs[q]->partial);
        hs[q]->graph_refs++;
    }
    *out = g;
    return 0;
}

// Replays the graph; every handle's clock and buffers belong to the graph's
// stream until gam_graph_sync returns.
extern "C" int gam_graph_launch(gam_graph *g, int replays) {
    if (!g || replays < 0) return -1;
    for (size_t q = 0; q < g->hs.size(); q++) {
        gam_handle *h = g->hs[q];
        gam_clock c;
        int rc = gam_get_clock(h, &c);   // also waits for the handle's own stream
        if (rc) return rc;
        // the graph's nodes hold the buffers of the capture-time State index; a
        // halted handle's kernels return at once, so only a running one must match
        if (h->partial != g->partial_cap[q] || (!h->clk_host->halt && h->s != g->s_cap[q])) {
            g->err = "gam_graph_launch: handle " + std::to_string(q) +
                     " was stepped an odd number of times (or re-tiled) since the capture";
            return -1;
        }
        g->ts0[q] = c.ts;
        g->s0[q] = h->s;
    }
    hipError_t e;
    for (int r = 0; r < replays; r++)
        if ((e = hipGraphLaunch(g->exec, g->stream)) != hipSuccess) return gfail(g, -2, "hipGraphLaunch", e);
    return 0;
}

// Waits for the replays; returns the largest CalcDT status over the handles
// (a halted handle skipped its remaining steps; its State is where it stopped).
extern "C" int gam_graph_sync(gam_graph *g) {
    if (!g) return -1;
    hipError_t e;
    if ((e = hipStreamSynchronize(g->stream)) != hipSuccess) return gfail(g, -2, "hipStreamSynchronize", e);
    int worst = 0;
    for (size_t q = 0; q < g->hs.size(); q++) {
        gam_handle *h = g->hs[q];
        gam_clock c;
        int rc = gam_get_clock(h, &c);
        if (rc) return rc;
        const int done = c.ts - g->ts0[q];
        h->s = (g->s0[q] + done) & 1;   // each step swaps State and oState
        g->ts0[q] = c.ts;
        g->s0[q] = h->s;
        worst = std::max(worst, c.status);
    }
    return worst;
}

extern "C" const char *gam_graph_error(const gam_graph *g) { return g ? g->err.c_str() : "null graph"; }

extern "C" int gam_graph_nodes(gam_graph *g, int *nodes) {
    if (!g || !nodes || !g->graph) return -1;
    size_t k = 0;
    hipError_t e;
    if ((e = hipGraphGetNodes(g->graph, nullptr, &k)) != hipSuccess) return gfail(g, -2, "hipGraphGetNodes", e);
    *nodes = (int)k;
    return 0;
}

// ----------------------------------------------------------------- timers API
extern "C" int gam_set_timing(gam_handle *h, int enable) {
    if (!h) return -1;
    int rc = gam_sync(h);
    if (rc) return rc;
    h->timing = enable != 0;
    for (int q = 0; q < ST_N; q++) { h->ms[q] = 0; h->launches[q] = 0; }
    return 0;
}

extern "C" int gam_get_timing(gam_handle *h, int stage, double *ms_total, int64_t *launches) {
    if (!h || stage < 0 || stage >= ST_N) return -1;
    int rc = gam_sync(h);
    if (rc) return rc;
    if (ms_total) *ms_total = h->ms[stage];
    if (launches) *launches = h->launches[stage];
    return 0;
}

extern "C" int gam_set_tile(gam_handle *h, int stage, int ti, int tj, int tk, int fast) {
    if (!h) return -1;
    if (refuse_if_captured(h, "gam_set_tile")) return -1;
    if (stage < 0 || stage >= ST_N || !tiled_stage(stage))
        return fail(h, -1, "gam_set_tile: stage %d takes no tile", stage);
    if (ti < 1 || tj < 1 || tk < 1 || ti * tj * tk != 256 || fast < 0 || fast > (TILE_LINEAR | TILE_NOXCD))
        return fail(h, -1, "gam_set_tile: need ti*tj*tk == 256 and fast in 0..7 (got %d %d %d %d)", ti, tj, tk, fast);
    int rc = gam_sync(h);   // launches already queued keep the tile they were given
    if (rc) return rc;
    const Tile t = {ti, tj, tk, fast};
    if (stage == ST_UPDATE) {   // one block minimum per update block, then the ring band's
        rc = ensure_partial(h, tile_blocks(t, h->gv.Ni, h->gv.Nj, cdiv(h->gv.Nk, update_kc(h))) + ring_dt_blocks(h));
        if (rc) return rc;
    }
    h->tile[stage] = t;
    return 0;
}

extern "C" int gam_get_tile(gam_handle *h, int stage, int *tile4) {
    if (!h || !tile4 || stage < 0 || stage >= ST_N || !tiled_stage(stage)) return -1;
    tile4[0] = h->tile[stage].ti; tile4[1] = h->tile[stage].tj; tile4[2] = h->tile[stage].tk;
    tile4[3] = h->tile[stage].fast;
    return 0;
}

extern "C" const char *gam_stage_name(int stage) { return (stage >= 0 && stage < ST_N) ? kStageNames[stage] : ""; }
extern "C" int gam_num_stages(void) { return ST_N; }

extern "C" int gam_stage_bytes(gam_handle *h, int stage, double *bytes) {
    // Unique bytes each stage must move: every input array read once over the
    // stage's domain, every output written once, each at its stored precision
    // (b = bulk elements, 8 or 4 bytes; d = FP64 elements).
    if (!h || !bytes || stage < 0 || stage >= ST_N) return -1;
    const GV &g = h->gv;
    const double G = g.nc, F = g.nf;
    const double nI = (double)(g.Nip + 1) * g.Njp * g.Nkp, nJ = (double)g.Nip * (g.Njp + 1) * g.Nkp,
                 nK = (double)g.Nip * g.Njp * (g.Nkp + 1), nA = (double)g.Nip * g.Njp * g.Nkp;
    const double nE = (double)(g.Nip + 1) * (g.Njp + 1) * (g.Nkp + 1);
    double b = 0, d = 0;
    switch (stage) {
    case ST_PREDICT:   // mag n,o (d) ; face -> hmag, bh, bn ; Gas n,o, vol -> hGas, Wh, Wn, VdV
        d = F * 6; b = F * (3 + 3 + 3 + 3) + G * (5 + 5 + 1 + 5 + 5 + 5 + 3); break;
    case ST_BHALF:   // hmag (b), xfc, xyzcc, vol (d) -> B (b)
        b = F * 3 + G * 3; d = F * 9 + G * (3 + 1); break;
    case ST_BPOST: {   // mag, xfc, xyzcc, vol (d) -> B (b); the fused 3D step recomputes the ghost shell only
        const GV &g = h->gv;
        const double S = h->mp.do25D ? G : (double)b2f_shell_count(g.Ni, g.Nj, g.Nk, g.Nip, g.Njp, g.Nkp);
        b = S * 3; d = S * (3 + 9 + 3 + 1); break;
    }
    case ST_MOM2FACE: b = G * (3 + 1 + 9); break;                                 // VdV, vol -> Vf (9)
    case ST_EMF_I: case ST_EMF_J: case ST_EMF_K: {   // Vf (dT1), rho ; face, bn, bh (dT1, dT2), Te, Teb, edge -> E
        const int e = stage - ST_EMF_I;
        const double nEd = (double)(g.Nip + (e != 0)) * (g.Njp + (e != 1)) * (g.Nkp + (e != 2));
        b = G * (3 + 1) + nEd * (2 + 2 + 2 + 6 + 4 + 1 + 1); break;
    }
    case ST_FLUX_I: b = (G * (5 + 5 + 3 + 3 + 1)) + nI * (9 + 1 + 1 + 8); d = h->bgv.on ? nI * 4 : 0; break;
    case ST_FLUX_J: b = (G * (5 + 5 + 3 + 3 + 1)) + nJ * (9 + 1 + 1 + 8); d = h->bgv.on ? nJ * 4 : 0; break;
    case ST_FLUX_K: b = (G * (5 + 5 + 3 + 3 + 1)) + nK * (9 + 1 + 1 + 8); d = h->bgv.on ? nK * 4 : 0; break;
    case ST_CT: d = F * (3 + 3); b = G * 3; break;
    case ST_UPDATE:   // Gas n, fluxes (8 per direction), di/dj/dk -> Gas, Bxyz (b) ; magFlux new, xfc, xyzcc, vol (d)
        b = G * (5 + 24 + 3 + 5 + 3 + (h->mp.doBoris ? 3 : 0)); d = F * (3 + 9) + G * (3 + 1);
        if (h->bgv.on) d += G * (3 + 3); break;                                  // dpB0, B0
    case ST_BC: b = 2.0 * (G - nA) * 5; d = 2.0 * (F - nE) * 3; break;
    case ST_CELLDT: b = G * (5 + 3 + 3); break;
    default: b = 0;
    }
    *bytes = (double)h->es * b + 8.0 * d;
    return 0;
}
