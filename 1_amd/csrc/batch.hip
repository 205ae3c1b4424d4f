// Batched engine: many independent simulations (same Runge-Kutta tableau, any
// sizes and materials) advanced together, one launch per pass and length
// class per RK stage for the whole batch.
//
// Each simulation keeps its own device state exactly as in the single-simulation
// engine (sim.hip): m, k_i, spectra B1/B2, kernel spectrum, EngineState, trial
// log.  A batched kernel covers every simulation whose padded length matches its
// template: blocks are numbered per simulation and prefix tables map a block to
// (simulation, local block), after which the kernel runs the same device body
// as the single-simulation kernel (passes.h, rk_dev.h).  So a simulation gets
// bit-identical arithmetic alone and in a batch; only its scheduling differs.
//
// Step control stays per simulation on the device (own dt, error, accept /
// reject, FSAL state, stop time), so a batched trial is a fixed launch
// sequence, captured once into a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/mumax_engine.h"
#include "engine.h"
#include "passes.h"
#include "rk_dev.h"
#include "sim.h"

namespace gk {

// one simulation of a batch, as the batched kernels see it (device table)
struct BSim {
    int Nx, Ny, Nz, Q, PY, PX;
    long N;
    float* m;
    float* mi;  // stage input stored by the Z forward pass (BStage::use_mi)
    float* k[7];
    float* mtmp;
    float2* B1;
    float2* B2;
    const float* Kc;
    EngineState* st;
    float* part;  // 3 x npart: |k0|, |k_last|, error
    int npart;
    int nfin;     // blocks of the RK final (error partials)
    TorqueArgs t; // c and part are set per stage
    CtrlParams cp;
    // twiddle tables of the simulation's lengths (the length-dispatch kernels)
    const float2* twM;  // Z half length M
    const float2* twP;  // Z padded length 2M
    const float2* twY;  // padded Y
    const float2* twX;  // padded X
};

// one RK stage, the same for every simulation of the batch
struct BStage {
    int i;
    int fsal0;  // stage 0 of an FSAL tableau: simulations holding an FSAL k skip the convolution
    int nk;
    int kj[6];
    float a[6];
    int nu;
    int uj[6];
    float ua[6], ub[6];
    float bs;
    int fin;             // epilogue mode (zinv_epi_tile's FIN)
    int use_mi;          // the Z forward pass stores m_i, the epilogue reads it (nu = 0)
    double ce[6], ce_self;  // fin == 2: b2_j - b_j of the slots and of this stage
    float c, cmax;
    int partslot;  // -1: none, 0: |k0|, 1: |k_last|
};

// largest j with start[j] <= b (start[0] = 0, increasing)
__device__ __forceinline__ int bfind(const int* __restrict__ start, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// the simulation of this block and the block's index within it
struct BBlock {
    const BSim* d;
    int lb;
};
__device__ __forceinline__ BBlock bblock(const BSim* sims, const int* idx, const int* start, int n) {
    const int j = bfind(start, n, blockIdx.x);
    return BBlock{sims + idx[j], (int)blockIdx.x - start[j]};
}
__device__ __forceinline__ bool bskip(const BSim& d, const BStage& sg) {
    return !d.st->trial_active || (sg.fsal0 && d.st->fsal_init);
}

// ------------------------------------------------------------------ passes
// The bodies of the batched passes: block lb of simulation d.  The per-length
// kernels (kb_*) and the length-dispatch kernels (kb_*_any, one launch for
// many length classes) both run them, with the same block shapes.
template <int M, int TR, int NK>
__device__ __forceinline__ void bz_fwd(const BSim& d, int lb, const BStage& sg, const float2* __restrict__ twM,
                                       const float2* __restrict__ twP, float2* lds) {
    const int nyb = (d.Ny + TR - 1) / TR;
    const int yb = lb % nyb, r = lb / nyb, x = r % d.Nx, c = r / d.Nx;
    StageIn si{};
    si.m = d.m;
    si.st = d.st;
    si.nk = sg.nk;
    si.mout = sg.use_mi ? d.mi : nullptr;
    for (int j = 0; j < 6; j++) {
        si.k[j] = j < sg.nk ? d.k[sg.kj[j]] : d.m;
        si.a[j] = sg.a[j];
    }
    zfwd_tile<M, TR, NK>(si, d.B1, twM, twP, d.Nx, d.Ny, d.Nz, c, x, yb * TR, lds);
}

template <int M, int TR, int NK>
__global__ __launch_bounds__(TR * RegFFT<M>::G) void kb_zfwd(const BSim* sims, const int* idx, const int* start,
                                                               int n, BStage sg, const float2* __restrict__ twM,
                                                               const float2* __restrict__ twP) {
    __shared__ float2 lds[zfwd_lds<M, TR>()];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    bz_fwd<M, TR, NK>(d, bb.lb, sg, twM, twP, lds);
}

// 3-D Y passes: lds holds the exchange buffers and then the twiddle copy
template <int PY> constexpr int ypass_lds() {
    return (kNT / RegFFT<PY>::G) * RegFFT<PY>::LDS + (PY <= 1024 ? PY : 1);
}
template <int PY, bool FWD>
__device__ __forceinline__ void b_ypass(const BSim& d, int lb, const float2* __restrict__ tw, float2* lds) {
    using F = RegFFT<PY>;
    constexpr int E = F::E, G = F::G, LPB = kNT / G;
    constexpr bool TWL = PY <= 1024;
    float2* twl = lds + LPB * F::LDS;
    if constexpr (TWL)
        for (int i = threadIdx.x; i < PY; i += kNT) twl[i] = tw[i];
    const int li = threadIdx.x / G, g = threadIdx.x % G;
    const long nlines = 3L * d.Nx * d.Q;
    const long line = (long)lb * LPB + li;
    float2 v[E];
    ypass_load<PY, FWD>(v, FWD ? d.B1 : d.B2, line, line < nlines, d.Nx, d.Ny, d.Q, g);
    __syncthreads();  // twiddle table
    reg_fft<float, PY, FWD ? -1 : +1>(v, g, lds + li * F::LDS, TWL ? twl : tw);
    if (line < nlines) ypass_store<PY, FWD>(v, FWD ? d.B2 : d.B1, line, d.Nx, d.Ny, d.Q, g);
}

template <int PY, bool FWD>
__global__ __launch_bounds__(kNT) void kb_ypass(const BSim* sims, const int* idx, const int* start, int n, BStage sg,
                                                const float2* __restrict__ tw) {
    __shared__ float2 lds[ypass_lds<PY>()];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    b_ypass<PY, FWD>(d, bb.lb, tw, lds);
}

// 3-D X . K . X: lds holds the exchange buffers and then the twiddle copy
template <int PX> constexpr int xmul_lds() { return (kNT / RegFFT<PX>::G) * RegFFT<PX>::LDS + PX; }
template <int PX>
__device__ __forceinline__ void b_xmul(const BSim& d, int lb, const float2* __restrict__ tw, float2* lds) {
    using F = RegFFT<PX>;
    constexpr int E = F::E, G = F::G, LPB = kNT / G;
    float2* twl = lds + LPB * F::LDS;
    for (int i = threadIdx.x; i < PX; i += kNT) twl[i] = tw[i];
    const int li = threadIdx.x / G, g = threadIdx.x % G;
    const int lpr = LPB < d.PY ? LPB : d.PY;
    const int rows = (d.PY + lpr - 1) / lpr;
    int kz, ky;
    xmul_column(lb, (long)d.Q * rows, rows, lpr, li, d.PY, kz, ky);
    float2 v[3][E];
    xmul_load<PX>(v, d.B2, kz, ky, d.Nx, d.PY, d.Q, g);
    __syncthreads();  // twiddle table
#pragma unroll
    for (int c = 0; c < 3; c++) reg_fft<float, PX, -1>(v[c], g, lds + li * F::LDS, twl);
    if (ky >= 0) xmul_kernel<PX>(v, d.Kc, kz, ky, d.PY, g);
#pragma unroll
    for (int c = 0; c < 3; c++) reg_fft<float, PX, +1>(v[c], g, lds + li * F::LDS, twl);
    if (ky >= 0) xmul_store<PX>(v, d.B2, kz, ky, d.Nx, d.PY, d.Q, g);
}

template <int PX>
__global__ __launch_bounds__(kNT) void kb_xmul(const BSim* sims, const int* idx, const int* start, int n, BStage sg,
                                               const float2* __restrict__ tw) {
    __shared__ float2 lds[xmul_lds<PX>()];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    b_xmul<PX>(d, bb.lb, tw, lds);
}

template <int PY, int PX>
__device__ __forceinline__ void b_yxy(const BSim& d, int lb, const float2* __restrict__ twY,
                                      const float2* __restrict__ twX, float2* lds) {
    using Y = Yxy<PY, PX>;
    float2 wx[PX];
#pragma unroll
    for (int j = 0; j < PX; j++) wx[j] = twX[j];
    float2 v[Y::RMAX][Y::E];
    yxy_load<PY, PX>(v, d.B1, d.Nx, d.Ny, d.Q, lb);
    yxy_plane<PY, PX>(v, d.B1, d.Kc, twY, wx, d.Nx, d.Ny, d.Q, lb, lds);
}

template <int PY, int PX>
__global__ __launch_bounds__(kYxyThreads) void kb_yxy(const BSim* sims, const int* idx, const int* start, int n,
                                                       BStage sg, const float2* __restrict__ twY,
                                                       const float2* __restrict__ twX) {
    __shared__ float2 lds[yxy_lds<PY, PX>()];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    b_yxy<PY, PX>(d, bb.lb, twY, twX, lds);
}

// 2-D: lds holds the exchange buffers (LPB RegFFT<PY>::LDS) and then the
// twiddle copy (PY, for PY <= 1024)
template <int PY> constexpr int y2r_lds() {
    return (kNT / RegFFT<PY>::G) * RegFFT<PY>::LDS + (PY <= 1024 ? PY : 1);
}
template <int PY>
__device__ __forceinline__ void b_y2r(const BSim& d, int lb, const float2* __restrict__ tw, float2* lds) {
    using F = RegFFT<PY>;
    constexpr int E = F::E, G = F::G, LPB = kNT / G;
    constexpr bool TWL = PY <= 1024;
    float2* twl = lds + LPB * F::LDS;
    if constexpr (TWL)
        for (int i = threadIdx.x; i < PY; i += kNT) twl[i] = tw[i];
    const int li = threadIdx.x / G, g = threadIdx.x % G;
    const long kz = (long)lb * LPB + li;
    __syncthreads();  // twiddle table
    if constexpr (E > 16) {
        y2r_column_split<PY>(d.B1, d.Kc, kz, kz < d.Q, d.Ny, d.Q, g, lds + li * F::LDS, TWL ? twl : tw);
    } else {
        float2 v[3][E];
        y2r_load<PY>(v, d.B1, kz, kz < d.Q, d.Ny, d.Q, g);
        y2r_column<PY>(v, d.Kc, kz, kz < d.Q, d.Q, g, lds + li * F::LDS, TWL ? twl : tw);
        if (kz < d.Q) y2r_store<PY>(v, d.B1, kz, d.Ny, d.Q, g);
    }
}

template <int PY>
__global__ __launch_bounds__(kNT) void kb_y2r(const BSim* sims, const int* idx, const int* start, int n, BStage sg,
                                              const float2* __restrict__ tw) {
    __shared__ float2 lds[y2r_lds<PY>()];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    b_y2r<PY>(d, bb.lb, tw, lds);
}

template <int M, int TR, int NU, int FIN>
__device__ __forceinline__ void bz_inv(const BSim& d, int lb, const BStage& sg, const float2* __restrict__ twM,
                                       const float2* __restrict__ twP, float2* lds, float* hext_s, float* red) {
    const int nyb = (d.Ny + TR - 1) / TR;
    const int yb = lb % nyb, x = lb / nyb;
    EpiArgs ea{};
    ea.m = sg.use_mi ? d.mi : d.m;
    ea.out = FIN == 1 ? d.m : d.k[sg.i];
    for (int j = 0; j < 6; j++) {
        const bool u = j < sg.nu;
        ea.k[j] = u ? d.k[sg.uj[j]] : d.m;  // unused slots: m with zero coefficients
        ea.a[j] = u ? sg.ua[j] : 0.f;
        ea.b[j] = u ? sg.ub[j] : 0.f;
        ea.ce[j] = u ? sg.ce[j] : 0.0;
    }
    ea.ce_self = sg.ce_self;
    ea.mhigh = d.mtmp;
    ea.part_err = d.part + 2L * d.npart;
    ea.bs = sg.bs;
    ea.nu = sg.nu;
    ea.fin = sg.fin;
    ea.t = d.t;
    ea.t.c = sg.c;
    ea.t.cmax = sg.cmax;
    ea.t.part = sg.partslot < 0 ? nullptr : d.part + (long)sg.partslot * d.npart;
    zinv_epi_tile<M, TR, NU, FIN>(d.B1, ea, twM, twP, d.st, d.Nx, d.Ny, d.Nz, x, yb * TR, lb, lds, hext_s, red);
}

template <int M, int TR, int NU, int FIN>
__global__ __launch_bounds__(3 * TR * RegFFT<M>::G) void kb_zinv_epi(const BSim* sims, const int* idx,
                                                                       const int* start, int n, BStage sg,
                                                                       const float2* __restrict__ twM,
                                                                       const float2* __restrict__ twP) {
    __shared__ float2 lds[zinv_epi_lds<M, TR>()];
    __shared__ float hext_s[3];
    __shared__ float red[3 * TR * RegFFT<M>::G];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    bz_inv<M, TR, NU, FIN>(d, bb.lb, sg, twM, twP, lds, hext_s, red);
}

// ------------------------------------------------------------------ RK pointwise and control
// FSAL: k_0 <- k_{S-1} for simulations that hold one (rk.go:311-315)
__global__ __launch_bounds__(256) void kb_fsal_copy(const BSim* sims, const int* idx, const int* start, int n,
                                                    int last) {
    const int j = bfind(start, n, blockIdx.x);
    const BSim& d = sims[idx[j]];
    if (!d.st->trial_active || !d.st->fsal_init) return;
    copy_body(d.k[0], d.k[last], 3 * d.N, (int)blockIdx.x - start[j], start[j + 1] - start[j]);
}

__global__ __launch_bounds__(256) void kb_rk_final(const BSim* sims, const int* idx, const int* start, int n,
                                                   FinalArgs fa) {
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (!d.st->trial_active) return;
    KPtrs kp;
    for (int s = 0; s < 7; s++) kp.k[s] = d.k[s] ? d.k[s] : d.k[0];
    FinalArgs a = fa;
    a.part_err = d.part + 2L * d.npart;
    a.normMap = d.t.normMap;
    rk_final_body(d.m, kp, d.mtmp, d.st, a, (int)d.N, bb.lb, d.nfin);
}

__global__ __launch_bounds__(256) void kb_commit(const BSim* sims, const int* idx, const int* start, int n) {
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    commit_body(d.m, d.mtmp, d.st, d.t.normMap, (int)d.N, bb.lb, d.nfin);
}

// one thread per simulation
__global__ void kb_ctrl_begin(const BSim* sims, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) ctrl_begin_body(sims[j].st);
}
// one wave per simulation
__global__ __launch_bounds__(64) void kb_ctrl_after_k0(const BSim* sims, int fsal) {
    const BSim& d = sims[blockIdx.x];
    if (!d.st->trial_active || d.cp.fold_k0) return;
    ctrl_after_k0_body(d.st, fsal && d.st->fsal_init, d.cp);
}
__global__ __launch_bounds__(64) void kb_ctrl_end(const BSim* sims) {
    const BSim& d = sims[blockIdx.x];
    ctrl_end_body(d.st, d.cp);
}
// status flags (0 running, 1 done, 2 stopped on a NaN dt or a non-finite torque or
// error, i.e. failed); with rem, each running simulation's remaining trials at its
// current dt, (stop - time) / dt, or its remaining steps under a step target
__global__ void kb_done(const BSim* sims, int n, int* done, float* rem) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const EngineState* st = sims[j].st;
    done[j] = st->done ? (st->nan_dt || st->nonfinite ? 2 : 1) : 0;
    if (rem) {
        double r = 0.0;
        if (!st->done)
            r = st->stop < INFINITY ? (st->stop - st->time) / fmax((double)st->dt, 1e-30)
                                    : (double)(st->steps_target - st->steps);
        rem[j] = (float)fmin(r, 1e30);
    }
}

}  // namespace gk

using namespace gk;

// ------------------------------------------------------------------ host
namespace {

// blocks of one pass for one simulation, per padded length class
struct LaunchSet {
    int L = 0;       // template length of the class
    int cnt = 0;     // simulations
    int blocks = 0;  // total blocks
    int* idx = nullptr;
    int* start = nullptr;
    size_t idx_off = 0, start_off = 0;  // offsets in the batch's table block (set-up)
    std::vector<int> hidx, hst;         // host copies of idx and start (set-up)
    int any = 0;                        // 1: small length classes merged, launched by a dispatch kernel
    int nt = 0, lds = 0;                // dispatch launch: threads per block and dynamic LDS bytes
};

// Host image of a batch's device tables, uploaded with one copy on the batch's
// stream: a synchronous hipMemcpy per table would wait for other batches'
// queued work.
struct TableImage {
    std::vector<char> bytes;
    size_t put(const void* src, size_t n) {
        const size_t off = (bytes.size() + 255) & ~(size_t)255;
        bytes.resize(off + n);
        memcpy(bytes.data() + off, src, n);
        return off;
    }
};

int epi_rows_host(int M);
int z_rows_host(int M);

}  // namespace

struct mmBatch {
    std::vector<mmSim*> sims;
    Tableau tb;
    hipStream_t s = nullptr;
    BSim* dsims = nullptr;
    int* ddone = nullptr;
    float* drem = nullptr;      // remaining trials per simulation (kb_done)
    std::vector<void*> owned;  // device allocations of the tables
    TableImage image;          // set-up only: the tables before upload
    // per pass: one LaunchSet per length class
    std::vector<LaunchSet> zf, yf, xm, yi, y2, yx, zi, fin, all256;
    int useGraph = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long long trials = 0;
    int* hdone = nullptr;       // pinned host copy of the done flags (mm_batch_enqueue)
    float* hrem = nullptr;      // pinned host copy of drem
    hipEvent_t ev = nullptr;    // end of the last enqueued work
};

namespace {

// rows per Z block (same choice as the single-simulation launchers)
template <int M> constexpr int z_rows_c() {
    constexpr int G = RegFFT<M>::G;
    int tr = 1024 / G;
    if (tr > 16) tr = 16;
    while (tr > 1 && tr * (M + 1) * 8 > 72 * 1024) tr /= 2;
    return tr;
}
template <int M> constexpr int epi_rows_c() {
    constexpr int G = RegFFT<M>::G;
    int tr = 1024 / (3 * G);
    if (tr > 16) tr = 16;
    int p = 1;
    while (p * 2 <= tr) p *= 2;
    tr = p;
    while (tr > 1 && zinv_epi_lds_elems(M, tr) * 8 > 80 * 1024) tr /= 2;
    return tr;
}


int z_rows_host(int M) {
    switch (M) {
#define GK_CASE(n) \
    case n: return z_rows_c<n>();
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    }
    abort();
}
int epi_rows_host(int M) {
    switch (M) {
#define GK_CASE(n) \
    case n: return epi_rows_c<n>();
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    }
    abort();
}
int lanes_host(int L) { return reg_fft_G(L); }  // RegFFT<L>::G

// ---------------------------------------------------------------- length-dispatch kernels
// One launch for the small length classes of a pass: each block reads its
// simulation's lengths and runs the body of that length, with the block shape
// of the per-length kernel (threads beyond it leave at once) and its LDS carved
// from dynamic shared memory sized for the launch's largest class.  A batch of
// many small simulations of different sizes then issues a few launches per
// stage instead of one per length class and pass.
constexpr int kAnyLds = 80 * 1024;
#define GK_ANY_M(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
#define GK_ANY_Y(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)
bool any_m(int M) {
    switch (M) {
#define GK_CASE(n) case n:
        GK_ANY_M(GK_CASE)
#undef GK_CASE
        return true;
    }
    return false;
}
bool any_y(int PY) {
    switch (PY) {
#define GK_CASE(n) case n:
        GK_ANY_Y(GK_CASE)
#undef GK_CASE
        return true;
    }
    return false;
}
bool any_yx(int PY, int PX) { return any_y(PY) && (PX == 4 || PX == 8) && yxy_fits(PX, PY); }

template <int M> constexpr int zinv_any_bytes() {
    constexpr int TR = epi_rows_c<M>(), NR = TR * RegFFT<M>::G;
    return zinv_epi_lds<M, TR>() * 8 + 16 + 3 * NR * 4;
}

template <int NK>
__global__ __launch_bounds__(1024) void kb_zfwd_any(const BSim* sims, const int* idx, const int* start, int n,
                                                    BStage sg) {
    extern __shared__ float2 dyn[];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    switch (d.Q - 1) {
#define GK_CASE(m)                                                              case m: {                                                                       constexpr int TR = z_rows_c<m>();                                           if (threadIdx.x >= TR * RegFFT<m>::G) return;                               bz_fwd<m, TR, NK>(d, bb.lb, sg, d.twM, d.twP, dyn);                         break;                                                                  }
        GK_ANY_M(GK_CASE)
#undef GK_CASE
    }
}

template <int NU, int FIN>
__global__ __launch_bounds__(768) void kb_zinv_any(const BSim* sims, const int* idx, const int* start, int n,
                                                   BStage sg) {
    extern __shared__ float2 dyn[];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    switch (d.Q - 1) {
#define GK_CASE(m)                                                                              case m: {                                                                                       constexpr int TR = epi_rows_c<m>(), NT = 3 * TR * RegFFT<m>::G;                             if (threadIdx.x >= NT) return;                                                              float* hext_s = reinterpret_cast<float*>(dyn + zinv_epi_lds<m, TR>());                      bz_inv<m, TR, NU, FIN>(d, bb.lb, sg, d.twM, d.twP, dyn, hext_s, hext_s + 4);                break;                                                                                  }
        GK_ANY_M(GK_CASE)
#undef GK_CASE
    }
}

__global__ __launch_bounds__(kYxyThreads) void kb_yxy_any(const BSim* sims, const int* idx, const int* start, int n,
                                                          BStage sg) {
    extern __shared__ float2 dyn[];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    if (d.PX == 4) {
        switch (d.PY) {
#define GK_CASE(py)     case py: if constexpr (yxy_fits(4, py)) b_yxy<py, 4>(d, bb.lb, d.twY, d.twX, dyn); break;
            GK_ANY_Y(GK_CASE)
#undef GK_CASE
        }
    } else {
        switch (d.PY) {
#define GK_CASE(py)     case py: if constexpr (yxy_fits(8, py)) b_yxy<py, 8>(d, bb.lb, d.twY, d.twX, dyn); break;
            GK_ANY_Y(GK_CASE)
#undef GK_CASE
        }
    }
}

__global__ __launch_bounds__(kNT) void kb_y2r_any(const BSim* sims, const int* idx, const int* start, int n,
                                                  BStage sg) {
    extern __shared__ float2 dyn[];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    switch (d.PY) {
#define GK_CASE(py)     case py: b_y2r<py>(d, bb.lb, d.twY, dyn); break;
        GK_ANY_Y(GK_CASE)
#undef GK_CASE
    }
}

#define GK_ANY_X(X) X(4) X(8) X(16) X(32)
bool any_x(int PX) { return PX == 4 || PX == 8 || PX == 16 || PX == 32; }

template <bool FWD>
__global__ __launch_bounds__(kNT) void kb_ypass_any(const BSim* sims, const int* idx, const int* start, int n,
                                                    BStage sg) {
    extern __shared__ float2 dyn[];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    switch (d.PY) {
#define GK_CASE(py) \
    case py: b_ypass<py, FWD>(d, bb.lb, d.twY, dyn); break;
        GK_ANY_Y(GK_CASE)
#undef GK_CASE
    }
}

__global__ __launch_bounds__(kNT) void kb_xmul_any(const BSim* sims, const int* idx, const int* start, int n,
                                                   BStage sg) {
    extern __shared__ float2 dyn[];
    const BBlock bb = bblock(sims, idx, start, n);
    const BSim& d = *bb.d;
    if (bskip(d, sg)) return;
    switch (d.PX) {
#define GK_CASE(px) \
    case px: b_xmul<px>(d, bb.lb, d.twX, dyn); break;
        GK_ANY_X(GK_CASE)
#undef GK_CASE
    }
}

// per length: threads and dynamic LDS bytes of a block in the dispatch kernels
struct AnyShape {
    int nt, lds;
};
AnyShape any_shape(int pass, int L) {
    switch (pass) {
    case 0:  // Z forward, L = M
        switch (L) {
#define GK_CASE(n) case n: return {z_rows_c<n>() * RegFFT<n>::G, zfwd_lds<n, z_rows_c<n>()>() * 8};
            GK_ANY_M(GK_CASE)
#undef GK_CASE
        }
        break;
    case 1:  // Z inverse + epilogue, L = M
        switch (L) {
#define GK_CASE(n) case n: return {3 * epi_rows_c<n>() * RegFFT<n>::G, zinv_any_bytes<n>()};
            GK_ANY_M(GK_CASE)
#undef GK_CASE
        }
        break;
    case 2: {  // thin plane, L = PY * 64 + PX
        const int PY = L / 64, PX = L % 64;
        switch (PY) {
#define GK_CASE(n)     case n: return {kYxyThreads, PX == 4 ? yxy_lds_elems(4, n) * 8 : yxy_lds_elems(8, n) * 8};
            GK_ANY_Y(GK_CASE)
#undef GK_CASE
        }
        break;
    }
    case 3:  // 2-D, L = PY
        switch (L) {
#define GK_CASE(n) case n: return {kNT, y2r_lds<n>() * 8};
            GK_ANY_Y(GK_CASE)
#undef GK_CASE
        }
        break;
    case 4:  // 3-D Y passes, L = PY
        switch (L) {
#define GK_CASE(n) case n: return {kNT, ypass_lds<n>() * 8};
            GK_ANY_Y(GK_CASE)
#undef GK_CASE
        }
        break;
    case 5:  // 3-D X . K . X, L = PX
        switch (L) {
#define GK_CASE(n) case n: return {kNT, xmul_lds<n>() * 8};
            GK_ANY_X(GK_CASE)
#undef GK_CASE
        }
        break;
    }
    abort();
}

void any_kernel_attributes() {
    static std::once_flag once;
    std::call_once(once, [] {
        auto set = [](const void* f) {
            GK_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kAnyLds));
        };
        set((const void*)kb_zfwd_any<0>);
        set((const void*)kb_zfwd_any<1>);
        set((const void*)kb_zfwd_any<2>);
        set((const void*)kb_zfwd_any<3>);
        set((const void*)kb_zfwd_any<5>);
        set((const void*)kb_zinv_any<0, 0>);
        set((const void*)kb_zinv_any<1, 0>);
        set((const void*)kb_zinv_any<3, 0>);
        set((const void*)kb_zinv_any<6, 0>);
        set((const void*)kb_zinv_any<1, 1>);
        set((const void*)kb_zinv_any<3, 1>);
        set((const void*)kb_zinv_any<3, 2>);
        set((const void*)kb_zinv_any<6, 2>);
        set((const void*)kb_yxy_any);
        set((const void*)kb_y2r_any);
        set((const void*)kb_ypass_any<true>);
        set((const void*)kb_ypass_any<false>);
        set((const void*)kb_xmul_any);
    });
}

void* balloc(mmBatch* b, size_t bytes) {
    void* p = cached_malloc(bytes);
    b->owned.push_back(p);
    return p;
}

// group simulations by a length key; blocks(sim) gives each simulation's block count
template <typename Key, typename Blocks>
std::vector<LaunchSet> make_sets(mmBatch* b, const std::vector<int>& members, Key key, Blocks blocks) {
    std::map<int, std::vector<int>> groups;
    for (int j : members) groups[key(b->sims[j])].push_back(j);
    std::vector<LaunchSet> out;
    for (auto& kv : groups) {
        LaunchSet ls;
        ls.L = kv.first;
        ls.cnt = (int)kv.second.size();
        std::vector<int> st(ls.cnt + 1, 0);
        for (int i = 0; i < ls.cnt; i++) st[i + 1] = st[i] + blocks(b->sims[kv.second[i]]);
        ls.blocks = st[ls.cnt];
        ls.hidx = kv.second;
        ls.hst = st;
        if (ls.blocks > 0) out.push_back(ls);
    }
    return out;
}

// Length classes of a pass (0 Z forward, 1 Z inverse, 2 thin plane, 3 2-D, 4 3-D
// Y forward or inverse, 5 3-D X . K . X) with
// fewer blocks than this are merged into one dispatch launch when two or more
// of them are small and have a body in the dispatch kernel.
std::atomic<int> g_merge_blocks{1024};
std::vector<LaunchSet> merge_small(std::vector<LaunchSet> sets, int pass) {
    std::vector<LaunchSet> out, small;
    const int limit = g_merge_blocks.load();
    for (auto& ls : sets) {
        const bool has = pass <= 1   ? any_m(ls.L)
                         : pass == 2 ? any_yx(ls.L / 64, ls.L % 64)
                         : pass == 5 ? any_x(ls.L)
                                     : any_y(ls.L);
        (has && ls.blocks < limit ? small : out).push_back(std::move(ls));
    }
    if (small.size() < 2) {
        for (auto& ls : small) out.push_back(std::move(ls));
        return out;
    }
    LaunchSet m;
    m.any = 1;
    m.L = -1;
    m.hst.push_back(0);
    for (auto& ls : small) {
        const AnyShape a = any_shape(pass, ls.L);
        m.nt = std::max(m.nt, a.nt);
        m.lds = std::max(m.lds, a.lds);
        for (int i = 0; i < ls.cnt; i++) {
            m.hidx.push_back(ls.hidx[i]);
            m.hst.push_back(m.hst.back() + ls.hst[i + 1] - ls.hst[i]);
        }
        m.cnt += ls.cnt;
    }
    m.blocks = m.hst.back();
    GK_ASSERT(m.lds <= kAnyLds && m.nt <= 1024, "dispatch launch: %d threads, %d bytes of LDS", m.nt, m.lds);
    out.push_back(std::move(m));
    return out;
}

// the launch sets' simulation lists and block starts into the table image
void stage_sets(mmBatch* b, std::vector<LaunchSet>& sets) {
    for (auto& ls : sets) {
        ls.idx_off = b->image.put(ls.hidx.data(), ls.cnt * sizeof(int));
        ls.start_off = b->image.put(ls.hst.data(), (ls.cnt + 1) * sizeof(int));
    }
}

int fin_blocks(const mmSim* s) { return std::max(1, std::min(kRedBlocks, div_up(s->N, 256))); }

// ---------------------------------------------------------------- launchers (one length class each)
void bl_zfwd(mmBatch* b, const LaunchSet& ls, const BStage& sg) {
    if (ls.any) {
        switch (zfwd_nk(sg.nk)) {
        case 0: kb_zfwd_any<0><<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg); break;
        case 1: kb_zfwd_any<1><<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg); break;
        case 2: kb_zfwd_any<2><<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg); break;
        case 3: kb_zfwd_any<3><<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg); break;
        default: kb_zfwd_any<5><<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg); break;
        }
        GK_LAUNCH_CHECK();
        return;
    }
    const float2* twM = twiddle_f32(ls.L);
    const float2* twP = twiddle_f32(2 * ls.L);
    switch (ls.L) {
#define GK_ZF(n, K) \
    kb_zfwd<n, z_rows_c<n>(), K><<<ls.blocks, z_rows_c<n>() * RegFFT<n>::G, 0, b->s>>>(b->dsims, ls.idx, ls.start, \
                                                                                     ls.cnt, sg, twM, twP)
#define GK_CASE(n)                                                                                  \
    case n:                                                                                         \
        switch (zfwd_nk(sg.nk)) {                                                                   \
        case 0: GK_ZF(n, 0); break;                                                                 \
        case 1: GK_ZF(n, 1); break;                                                                 \
        case 2: GK_ZF(n, 2); break;                                                                 \
        case 3: GK_ZF(n, 3); break;                                                                 \
        default: GK_ZF(n, 5); break;                                                                \
        }                                                                                           \
        break;
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    #undef GK_ZF
    GK_LAUNCH_CHECK();
}
void bl_ypass(mmBatch* b, const LaunchSet& ls, const BStage& sg, bool fwd) {
    if (ls.any) {
        if (fwd) kb_ypass_any<true><<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg);
        else kb_ypass_any<false><<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg);
        GK_LAUNCH_CHECK();
        return;
    }
    const float2* tw = twiddle_f32(ls.L);
    switch (ls.L) {
#define GK_CASE(n)                                                                                       \
    case n:                                                                                              \
        if (fwd) kb_ypass<n, true><<<ls.blocks, kNT, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, tw);  \
        else kb_ypass<n, false><<<ls.blocks, kNT, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, tw);     \
        break;
        GK_LEN_Y(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}
void bl_xmul(mmBatch* b, const LaunchSet& ls, const BStage& sg) {
    if (ls.any) {
        kb_xmul_any<<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg);
        GK_LAUNCH_CHECK();
        return;
    }
    const float2* tw = twiddle_f32(ls.L);
    switch (ls.L) {
#define GK_CASE(n)                                                                            \
    case n: kb_xmul<n><<<ls.blocks, kNT, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, tw); break;
        GK_LEN_X(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}
// thin 3-D plane pass; the class key is PY * 64 + PX
void bl_yxy(mmBatch* b, const LaunchSet& ls, const BStage& sg) {
    if (ls.any) {
        kb_yxy_any<<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg);
        GK_LAUNCH_CHECK();
        return;
    }
    const int PY = ls.L / 64, PX = ls.L % 64;
    const float2* twY = twiddle_f32(PY);
    const float2* twX = twiddle_f32(PX);
    switch (PY) {
#define GK_YX(n, px)                                                                                              \
    case px:                                                                                                      \
        if constexpr (yxy_fits(px, n))                                                                            \
            kb_yxy<n, px><<<ls.blocks, kYxyThreads, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twY, twX); \
        else                                                                                                      \
            abort();                                                                                              \
        break;
#define GK_CASE(n)                                                                                                \
    case n:                                                                                                       \
        switch (PX) {                                                                                             \
            GK_YX(n, 4) GK_YX(n, 6) GK_YX(n, 8) GK_YX(n, 12) GK_YX(n, 16)                                         \
        default: abort();                                                                                         \
        }                                                                                                         \
        break;
        GK_LEN_Y(GK_CASE)
#undef GK_CASE
#undef GK_YX
    default: abort();
    }
    GK_LAUNCH_CHECK();
}
void bl_y2d(mmBatch* b, const LaunchSet& ls, const BStage& sg) {
    if (ls.any) {
        kb_y2r_any<<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg);
        GK_LAUNCH_CHECK();
        return;
    }
    const float2* tw = twiddle_f32(ls.L);
    switch (ls.L) {
#define GK_CASE(n)                                                                                           \
    case n: kb_y2r<n><<<ls.blocks, kNT, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, tw); \
        break;
        GK_LEN_Y(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}
void bl_zinv_epi(mmBatch* b, const LaunchSet& ls, const BStage& sg) {
    if (ls.any) {
#define GK_ZA(nu, fin) kb_zinv_any<nu, fin><<<ls.blocks, ls.nt, ls.lds, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg)
        if (sg.fin == 1) {
            if (sg.nu <= 1) GK_ZA(1, 1);
            else GK_ZA(3, 1);
        } else if (sg.fin == 2) {
            if (sg.nu <= 3) GK_ZA(3, 2);
            else GK_ZA(6, 2);
        } else {
            if (sg.nu == 0) GK_ZA(0, 0);
            else if (sg.nu == 1) GK_ZA(1, 0);
            else if (sg.nu <= 3) GK_ZA(3, 0);
            else GK_ZA(6, 0);
        }
#undef GK_ZA
        GK_LAUNCH_CHECK();
        return;
    }
    const float2* twM = twiddle_f32(ls.L);
    const float2* twP = twiddle_f32(2 * ls.L);
    switch (ls.L) {
#define GK_CASE(n)                                                                                       \
    case n: {                                                                                            \
        constexpr int TR = epi_rows_c<n>();                                                              \
        const int th = 3 * TR * RegFFT<n>::G;                                                            \
        if (sg.fin == 1) {                                                                               \
            if (sg.nu <= 1) kb_zinv_epi<n, TR, 1, 1><<<ls.blocks, th, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twM, twP); \
            else kb_zinv_epi<n, TR, 3, 1><<<ls.blocks, th, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twM, twP);         \
        } else if (sg.fin == 2) {                                                                        \
            if (sg.nu <= 3) kb_zinv_epi<n, TR, 3, 2><<<ls.blocks, th, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twM, twP); \
            else kb_zinv_epi<n, TR, 6, 2><<<ls.blocks, th, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twM, twP);         \
        } else {                                                                                         \
            if (sg.nu == 0) kb_zinv_epi<n, TR, 0, 0><<<ls.blocks, th, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twM, twP); \
            else if (sg.nu == 1) kb_zinv_epi<n, TR, 1, 0><<<ls.blocks, th, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twM, twP); \
            else if (sg.nu <= 3) kb_zinv_epi<n, TR, 3, 0><<<ls.blocks, th, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twM, twP); \
            else kb_zinv_epi<n, TR, 6, 0><<<ls.blocks, th, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, sg, twM, twP);            \
        }                                                                                                \
        break;                                                                                           \
    }
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- one batched trial
void enqueue_btrial(mmBatch* b) {
    const Tableau& t = b->tb;
    const int n = (int)b->sims.size();
    kb_ctrl_begin<<<div_up(n, 64), 64, 0, b->s>>>(b->dsims, n);
    GK_LAUNCH_CHECK();
    for (int i = 0; i < t.S; i++) {
        BStage sg{};
        sg.i = i;
        sg.fsal0 = (i == 0 && t.fsal) ? 1 : 0;
        if (sg.fsal0)
            for (auto& ls : b->all256) {
                kb_fsal_copy<<<ls.blocks, 256, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, t.S - 1);
                GK_LAUNCH_CHECK();
            }
        for (int j = 0; j < t.S; j++)
            if (t.a[i][j] != 0.f) {
                sg.kj[sg.nk] = j;
                sg.a[sg.nk] = t.a[i][j];
                sg.nk++;
            }
        const bool fin = !t.adaptive && i == t.S - 1;
        const bool efin = t.fsal_final() && i == t.S - 1;
        for (int j = 0; j < i; j++)
            if (t.a[i][j] != 0.f || (fin && t.b[j] != 0.f) || (efin && t.b2[j] != t.b[j])) {
                sg.uj[sg.nu] = j;
                sg.ua[sg.nu] = t.a[i][j];
                sg.ub[sg.nu] = fin ? t.b[j] : 0.f;
                sg.ce[sg.nu] = efin ? (double)t.b2[j] - (double)t.b[j] : 0.0;
                sg.nu++;
            }
        sg.fin = fin ? 1 : efin ? 2 : 0;
        sg.ce_self = efin ? (double)t.b2[i] - (double)t.b[i] : 0.0;
        sg.use_mi = (!fin && !efin && sg.nu >= 2) ? 1 : 0;
        if (sg.use_mi) sg.nu = 0;
        sg.bs = fin ? t.b[i] : 0.f;
        sg.c = t.c[i];
        sg.cmax = t.cmax(i);
        sg.partslot = i == 0 ? 0 : (t.fsal && i == t.S - 1) ? 1 : -1;
        for (auto& ls : b->zf) bl_zfwd(b, ls, sg);
        for (auto& ls : b->yf) bl_ypass(b, ls, sg, true);
        for (auto& ls : b->xm) bl_xmul(b, ls, sg);
        for (auto& ls : b->yi) bl_ypass(b, ls, sg, false);
        for (auto& ls : b->y2) bl_y2d(b, ls, sg);
        for (auto& ls : b->yx) bl_yxy(b, ls, sg);
        for (auto& ls : b->zi) bl_zinv_epi(b, ls, sg);
        if (i == 0) {
            kb_ctrl_after_k0<<<n, 64, 0, b->s>>>(b->dsims, t.fsal);
            GK_LAUNCH_CHECK();
        }
    }
    if (t.adaptive && !t.fsal_final()) {
        FinalArgs fa{};
        fa.S = t.S;
        for (int i = 0; i < 7; i++) {
            fa.b[i] = t.b[i];
            fa.b2[i] = t.b2[i];
        }
        fa.adaptive = 1;
        for (auto& ls : b->fin) {
            kb_rk_final<<<ls.blocks, 256, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt, fa);
            GK_LAUNCH_CHECK();
        }
    }
    kb_ctrl_end<<<n, 64, 0, b->s>>>(b->dsims);
    GK_LAUNCH_CHECK();
    if (t.adaptive)
        for (auto& ls : b->fin) {
            kb_commit<<<ls.blocks, 256, 0, b->s>>>(b->dsims, ls.idx, ls.start, ls.cnt);
            GK_LAUNCH_CHECK();
        }
}

}  // namespace

extern "C" {

mmBatch* mm_batch_new(mmSim* const* sims, int n) {
    GK_ASSERT(n > 0 && sims, "mm_batch_new: empty batch");
    mmBatch* b = new mmBatch();
    b->sims.assign(sims, sims + n);
    b->tb = sims[0]->tb;
    for (int j = 0; j < n; j++) {
        const mmSim* s = sims[j];
        GK_ASSERT(s->mode == MM_MODE_FUSED && s->conv, "mm_batch_new: simulation %d is not a fused-mode engine", j);
        GK_ASSERT(s->p.solver == sims[0]->p.solver, "mm_batch_new: simulations must share the solver");
    }
    b->s = cached_stream();
    // simulation table
    std::vector<BSim> h(n);
    for (int j = 0; j < n; j++) {
        mmSim* s = sims[j];
        BSim& d = h[j];
        d = BSim{};
        d.Nx = s->cp.N[0];
        d.Ny = s->cp.N[1];
        d.Nz = s->cp.N[2];
        d.Q = s->cp.Q;
        d.PY = s->cp.P[1];
        d.PX = s->cp.P[0];
        d.N = s->N;
        d.m = s->m;
        d.mi = s->mi;
        for (int i = 0; i < 7; i++) d.k[i] = s->k[i];
        d.mtmp = s->mtmp;
        d.B1 = s->cp.B1;
        d.B2 = s->cp.B2;
        d.Kc = s->cp.Kc;
        d.st = s->st;
        d.part = s->part;
        d.npart = s->npart;
        d.nfin = fin_blocks(s);
        d.t = torque_args(s, 0.f, nullptr);
        if (s->field.n >= 2) {
            // the batch keeps its own copy of the pointwise field table, so the
            // simulation can replace its table in place (mm_sim_set_field)
            const int nf = s->field.n;
            double* ft = (double*)balloc(b, nf * sizeof(double));
            float* fbv = (float*)balloc(b, 3 * nf * sizeof(float));
            GK_CHECK(hipMemcpyAsync(ft, s->field.t, nf * sizeof(double), hipMemcpyDeviceToDevice, b->s));
            GK_CHECK(hipMemcpyAsync(fbv, s->field.b, 3 * nf * sizeof(float), hipMemcpyDeviceToDevice, b->s));
            d.t.field.t = ft;
            d.t.field.b = fbv;
        }
        d.twM = twiddle_f32(s->cp.M);
        d.twP = twiddle_f32(s->cp.P[2]);
        d.twY = s->cp.P[1] > 1 ? twiddle_f32(s->cp.P[1]) : nullptr;
        d.twX = s->cp.P[0] > 1 ? twiddle_f32(s->cp.P[0]) : nullptr;
        d.cp = ctrl_params(s);
        d.cp.n_err = d.nfin;
        d.cp.part_e = nullptr;  // the batched epilogue computes no energy
    }
    const size_t sims_off = b->image.put(h.data(), n * sizeof(BSim));
    b->ddone = (int*)balloc(b, n * sizeof(int));
    b->drem = (float*)balloc(b, n * sizeof(float));
    // launch sets per pass and length class
    std::vector<int> all(n), d3, d2, dt;
    for (int j = 0; j < n; j++) {
        all[j] = j;
        (sims[j]->cp.P[0] == 1 ? d2 : sims[j]->cp.yxy ? dt : d3).push_back(j);
    }
    auto M = [](const mmSim* s) { return s->cp.M; };
    auto PY = [](const mmSim* s) { return s->cp.P[1]; };
    auto PX = [](const mmSim* s) { return s->cp.P[0]; };
    b->zf = make_sets(b, all, M, [](const mmSim* s) { return div_up(s->cp.N[1], z_rows_host(s->cp.M)) * s->cp.N[0] * 3; });
    b->zi = make_sets(b, all, M, [](const mmSim* s) { return div_up(s->cp.N[1], epi_rows_host(s->cp.M)) * s->cp.N[0]; });
    auto ylines = [](const mmSim* s) { return div_up(3L * s->cp.N[0] * s->cp.Q, 256 / lanes_host(s->cp.P[1])); };
    b->yf = make_sets(b, d3, PY, ylines);
    b->yi = make_sets(b, d3, PY, ylines);
    b->xm = make_sets(b, d3, PX, [](const mmSim* s) {
        const int lpb = 256 / lanes_host(s->cp.P[0]);
        const int lpr = std::min(lpb, s->cp.P[1]);
        return s->cp.Q * div_up(s->cp.P[1], lpr);
    });
    b->y2 = make_sets(b, d2, PY, [](const mmSim* s) { return div_up(s->cp.Q, 256 / lanes_host(s->cp.P[1])); });
    b->yx = make_sets(b, dt, [](const mmSim* s) { return s->cp.P[1] * 64 + s->cp.P[0]; },
                      [](const mmSim* s) { return s->cp.Q; });
    b->fin = make_sets(b, all, [](const mmSim*) { return 0; }, [](const mmSim* s) { return fin_blocks(s); });
    b->all256 = make_sets(b, all, [](const mmSim*) { return 0; },
                          [](const mmSim* s) { return std::max(1, std::min(1024, div_up(3L * s->N, 256 * 4))); });
    // small length classes of each pass: one dispatch launch per pass
    any_kernel_attributes();
    b->zf = merge_small(std::move(b->zf), 0);
    b->zi = merge_small(std::move(b->zi), 1);
    b->yx = merge_small(std::move(b->yx), 2);
    b->y2 = merge_small(std::move(b->y2), 3);
    b->yf = merge_small(std::move(b->yf), 4);
    b->yi = merge_small(std::move(b->yi), 4);
    b->xm = merge_small(std::move(b->xm), 5);
    for (auto* sets : {&b->zf, &b->zi, &b->yf, &b->yi, &b->xm, &b->y2, &b->yx, &b->fin, &b->all256})
        stage_sets(b, *sets);
    // upload the tables: one block, one copy on the batch's stream from pinned memory
    {
        const size_t nb = b->image.bytes.size();
        char* dev = (char*)balloc(b, nb);
        char* host = (char*)cached_host_malloc(nb);
        memcpy(host, b->image.bytes.data(), nb);
        GK_CHECK(hipMemcpyAsync(dev, host, nb, hipMemcpyHostToDevice, b->s));
        GK_CHECK(hipStreamSynchronize(b->s));
        cached_host_free(host);
        b->dsims = (BSim*)(dev + sims_off);
        for (auto* sets : {&b->zf, &b->zi, &b->yf, &b->yi, &b->xm, &b->y2, &b->yx, &b->fin, &b->all256})
            for (LaunchSet& ls : *sets) {
                ls.idx = (int*)(dev + ls.idx_off);
                ls.start = (int*)(dev + ls.start_off);
            }
        b->image = TableImage{};
    }
    // twiddle tables are created on first use: make them now, outside any graph capture
    for (int j = 0; j < n; j++) {
        const ConvPlan& cp = sims[j]->cp;
        (void)twiddle_f32(cp.M);
        (void)twiddle_f32(cp.P[2]);
        if (cp.P[1] > 1) (void)twiddle_f32(cp.P[1]);
        if (cp.P[0] > 1) (void)twiddle_f32(cp.P[0]);
    }
    // tables above were copied synchronously; other batches' streams keep running
    GK_CHECK(hipStreamSynchronize(b->s));
    return b;
}

void mm_batch_free(mmBatch* b) {
    if (!b) return;
    if (b->s) (void)hipStreamSynchronize(b->s);
    cached_host_free(b->hdone);
    cached_host_free(b->hrem);
    if (b->ev) (void)hipEventDestroy(b->ev);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void* p : b->owned) cached_free(p);
    cached_stream_release(b->s);
    delete b;
}

void mm_batch_set_graph(mmBatch* b, int on) { b->useGraph = on; }

int mm_batch_set_merge_blocks(int blocks) {
    GK_ASSERT(blocks >= 0, "mm_batch_set_merge_blocks: %d", blocks);
    return g_merge_blocks.exchange(blocks);
}

int mm_batch_size(mmBatch* b) { return (int)b->sims.size(); }

// Run(): every simulation j gets the stop time time_j + seconds[j] and is
// normalised first (simrun.go:19-24); Steps() (seconds null): steps[j] steps,
// no normalisation (simrun.go:85-92).
void mm_batch_start(mmBatch* b, const double* seconds, const long long* steps) {
    for (size_t j = 0; j < b->sims.size(); j++) {
        mmSim* s = b->sims[j];
        GK_CHECK(hipStreamSynchronize(s->s));
        EngineState h = read_state(s);
        const double stop = seconds ? h.time + seconds[j] / s->unitTime : INFINITY;
        if (seconds) launch_normalize_active(s->m, nullptr, s->normMap, s->N, b->s);
        launch_set_target(s->st, steps ? steps[j] : -1, stop, b->s);
    }
    GK_CHECK(hipStreamSynchronize(b->s));
}

// enqueue n batched trials and wait; returns the number of simulations not done
void mm_batch_enqueue(mmBatch* b, int n) {
    if (b->useGraph) {
        if (!b->exec) {
            GK_CHECK(hipStreamBeginCapture(b->s, hipStreamCaptureModeRelaxed));
            enqueue_btrial(b);
            GK_CHECK(hipStreamEndCapture(b->s, &b->graph));
            GK_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        for (int i = 0; i < n; i++) GK_CHECK(hipGraphLaunch(b->exec, b->s));
    } else {
        for (int i = 0; i < n; i++) enqueue_btrial(b);
    }
    b->trials += n;
    const int ns = (int)b->sims.size();
    kb_done<<<div_up(ns, 64), 64, 0, b->s>>>(b->dsims, ns, b->ddone, b->drem);
    GK_LAUNCH_CHECK();
    if (!b->hdone) {
        b->hdone = (int*)cached_host_malloc(ns * sizeof(int));
        b->hrem = (float*)cached_host_malloc(ns * sizeof(float));
        GK_CHECK(hipEventCreateWithFlags(&b->ev, hipEventDisableTiming));
    }
    GK_CHECK(hipMemcpyAsync(b->hdone, b->ddone, ns * sizeof(int), hipMemcpyDeviceToHost, b->s));
    GK_CHECK(hipMemcpyAsync(b->hrem, b->drem, ns * sizeof(float), hipMemcpyDeviceToHost, b->s));
    GK_CHECK(hipEventRecord(b->ev, b->s));
}

int mm_batch_ready(mmBatch* b) {
    if (!b->ev) return 1;
    const hipError_t e = hipEventQuery(b->ev);
    if (e == hipErrorNotReady) return 0;
    GK_CHECK(e);
    return 1;
}

int mm_batch_flags(mmBatch* b, int* done) {
    GK_ASSERT(b->ev != nullptr, "mm_batch_flags before mm_batch_enqueue");
    GK_CHECK(hipEventSynchronize(b->ev));
    const int ns = (int)b->sims.size();
    int active = 0;
    for (int i = 0; i < ns; i++) {
        if (done) done[i] = b->hdone[i];
        active += !b->hdone[i];
    }
    return active;
}

int mm_batch_remaining(mmBatch* b, float* rem) {
    GK_ASSERT(b->ev != nullptr, "mm_batch_remaining before mm_batch_enqueue");
    GK_CHECK(hipEventSynchronize(b->ev));
    const int ns = (int)b->sims.size();
    for (int i = 0; i < ns; i++) rem[i] = b->hrem[i];
    return ns;
}

int mm_batch_trials(mmBatch* b, int n) {
    mm_batch_enqueue(b, n);
    return mm_batch_flags(b, nullptr);
}

// status of every simulation (0 running, 1 reached its stop time / step target, 2 failed)
void mm_batch_done(mmBatch* b, int* done) {
    const int ns = (int)b->sims.size();
    kb_done<<<div_up(ns, 64), 64, 0, b->s>>>(b->dsims, ns, b->ddone, nullptr);
    GK_LAUNCH_CHECK();
    GK_CHECK(hipMemcpyAsync(done, b->ddone, ns * sizeof(int), hipMemcpyDeviceToHost, b->s));
    GK_CHECK(hipStreamSynchronize(b->s));
}

// run until every simulation is done, polling every `poll` trials; returns batched trials
long long mm_batch_run(mmBatch* b, const double* seconds, int poll) {
    mm_batch_start(b, seconds, nullptr);
    long long n = 0;
    if (poll < 1) poll = 16;
    while (true) {
        const int active = mm_batch_trials(b, poll);
        n += poll;
        if (!active) break;
    }
    return n;
}

}  // extern "C"
