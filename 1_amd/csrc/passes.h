// Device code of the convolution passes, shared by the single-simulation
// kernels (conv.hip) and the batched kernels (batch.hip): each pass is a
// __device__ body over one tile / line / column set, wrapped by kernels that
// only decide which tile a block works on.
#pragma once
#include <hip/hip_runtime.h>

#include "engine.h"
#include "fft_device.h"

// Lengths the engine's register-FFT passes are built for (2^a and 3 * 2^a):
// M = half the padded Z length, PY the padded Y length, PX the padded X length.
#define GK_LEN_M(X)                                                                                              \
    X(1) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) \
        X(768) X(1024) X(1536) X(2048)
#define GK_LEN_Y(X)                                                                                              \
    X(2) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)    \
        X(1024) X(1536) X(2048) X(3072) X(4096)
#define GK_LEN_X(X)                                                                                              \
    X(2) X(4) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768)    \
        X(1024)

namespace gk {

// =============================================================== passes
constexpr int kNT = 256;

// Workgroups are dispatched to the 8 XCDs round-robin (block b runs on XCD b % 8),
// and each XCD has its own L2.  Renumber so that every XCD gets one contiguous
// range of logical blocks: neighbouring blocks, which share cache lines (and
// mirrored kernel rows), then meet in the same L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    constexpr int kXcd = 8;
    if (nb % kXcd) return b;
    return (b % kXcd) * (nb / kXcd) + b / kXcd;
}

template <int L> constexpr int rows_per_block() {
    int n = 71680 / ((L + 1) * 8);
    n = n < 1 ? 1 : (n > 16 ? 16 : n);
    int p = 1;
    while (p * 2 <= n) p *= 2;
    return p;
}

// shuffle of a complex value within groups of W lanes
template <int W> __device__ __forceinline__ float2 shfl2(float2 v, int src) {
    return make_float2(__shfl(v.x, src, W), __shfl(v.y, src, W));
}

// Z forward.  TR rows (y0..y0+TR) of one (c, x) plane; each row is a group of G
// lanes holding the packed complex row z[n] = x[2n] + i x[2n+1] (RK stage build
// m + sum dt a_ij k_j applied while loading), transformed in registers; the
// r2c split takes Z[M-k] from the partner lane by a shuffle, and the spectrum is
// transposed through a padded LDS tile into B1[c][x][kz][y] (TR consecutive y).
template <int M, int TR> constexpr int zfwd_lds() {
    constexpr int FFT_LDS = TR * RegFFT<M>::LDS, TILE_LDS = (M + 1) * (TR + 1);
    return FFT_LDS > TILE_LDS ? FFT_LDS : TILE_LDS;
}

// body: component c, plane x, rows y0..y0+TR; lds holds zfwd_lds<M,TR>() complex.
// NK earlier stages enter the stage build (a compile-time count, so every load
// of a lane is issued before the first is used; unused slots hold m with a
// zero coefficient).
constexpr int zfwd_nk(int nk) { return nk <= 3 ? nk : 5; }
template <int M, int TR, int NK>
__device__ __forceinline__ void zfwd_tile(const StageIn& si, float2* __restrict__ B1, const float2* __restrict__ twM,
                                          const float2* __restrict__ twP, int Nx, int Ny, int Nz, int c, int x,
                                          int y0, float2* lds) {
    using F = RegFFT<M>;
    constexpr int E = F::E, G = F::G, NT = TR * G, Q = M + 1, TS = TR + 1;
    const int row = threadIdx.x / G, g = threadIdx.x % G;
    const int nrows = min(TR, Ny - y0);
    const bool ok = row < nrows;
    const long N = (long)Nx * Ny * Nz;
    const long rbase = c * N + ((long)x * Ny + y0 + row) * Nz;
    const float dt = si.st ? si.st->dt_used : 0.f;
    float coef[6];
#pragma unroll
    for (int j = 0; j < 6; j++) coef[j] = dt * si.a[j];
    float2 v[E];
    if ((Nz & 1) == 0) {
        // pairs (2n, 2n+1) as float2: rows start at an even float offset
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int n = g + G * e;
            float2 z = make_float2(0.f, 0.f);
            if (!pad_half<M>(e) && ok && 2 * n < Nz) {
                z = *reinterpret_cast<const float2*>(si.m + rbase + 2 * n);
#pragma unroll
                for (int j = 0; j < NK; j++) {
                    const float2 kv = *reinterpret_cast<const float2*>(si.k[j] + rbase + 2 * n);
                    z.x = fmaf(coef[j], kv.x, z.x);
                    z.y = fmaf(coef[j], kv.y, z.y);
                }
                if (si.mout) *reinterpret_cast<float2*>(si.mout + rbase + 2 * n) = z;
            }
            v[e] = z;
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int n = g + G * e;
            float z2[2] = {0.f, 0.f};
            for (int h = 0; h < 2; h++) {
                const int i = 2 * n + h;
                if (!pad_half<M>(e) && ok && i < Nz) {
                    float a = si.m[rbase + i];
#pragma unroll
                    for (int j = 0; j < NK; j++) a = fmaf(coef[j], si.k[j][rbase + i], a);
                    if (si.mout) si.mout[rbase + i] = a;
                    z2[h] = a;
                }
            }
            v[e] = make_float2(z2[0], z2[1]);
        }
    }
    reg_fft<float, M, -1>(v, g, lds + row * F::LDS, twM);
    // r2c split: lane g holds Z[k], k = g + G e; Z[M-k] is in lane (G-g) % G at e' = E-1-e
    // (or, for g = 0, in this lane at e' = (E-e) % E)
    const int partner = (G - g) % G;
    float2 X[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const float2 t = shfl2<G>(v[E - 1 - e], partner);
        const float2 zmk = g == 0 ? v[(E - e) % E] : t;
        X[e] = r2c_split<float>(v[e], zmk, twP[g + G * e]);
    }
    const float2 nyq = make_float2(v[0].x - v[0].y, 0.f);  // X[M], valid in lane 0
    __syncthreads();  // the FFT exchange buffer becomes the transpose tile
#pragma unroll
    for (int e = 0; e < E; e++) lds[(g + G * e) * TS + row] = X[e];
    if (g == 0) lds[M * TS + row] = nyq;
    __syncthreads();
    float2* dst = B1 + ((long)c * Nx + x) * Q * Ny + y0;
    for (int t = threadIdx.x; t < Q * TR; t += NT) {
        const int r = t % TR, kz = t / TR;
        if (r < nrows) dst[(long)kz * Ny + r] = lds[kz * TS + r];
    }
}

template <int M, int TR, int NK>
__global__ __launch_bounds__(TR * RegFFT<M>::G) void k_zfwd(StageIn si, float2* __restrict__ B1,
                                                              const float2* __restrict__ twM,
                                                              const float2* __restrict__ twP, int Nx, int Ny, int Nz) {
    __shared__ float2 lds[zfwd_lds<M, TR>()];
    if (si.st && !si.st->trial_active) return;
    zfwd_tile<M, TR, NK>(si, B1, twM, twP, Nx, Ny, Nz, blockIdx.z, blockIdx.y, blockIdx.x * TR, lds);
}

// Z inverse: B1[c][x][kz][y0..y0+TR) -> rows of h.  The spectrum is read through
// a padded LDS tile (TR consecutive y per kz), merged (c2r) in registers and
// transformed back; each lane writes its (2n, 2n+1) pairs as float2.
template <int M, int TR>
__global__ __launch_bounds__(TR * RegFFT<M>::G) void k_zinv(const float2* __restrict__ B1, float* __restrict__ h,
                                                              const float2* __restrict__ twM,
                                                              const float2* __restrict__ twP, const EngineState* st,
                                                              int Nx, int Ny, int Nz) {
    using F = RegFFT<M>;
    constexpr int E = F::E, G = F::G, NT = TR * G, Q = M + 1, TS = TR + 1;
    __shared__ float2 lds[zfwd_lds<M, TR>()];
    if (st && !st->trial_active) return;
    const int row = threadIdx.x / G, g = threadIdx.x % G;
    const int c = blockIdx.z, x = blockIdx.y, y0 = blockIdx.x * TR;
    const int nrows = min(TR, Ny - y0);
    const float2* src = B1 + ((long)c * Nx + x) * Q * Ny + y0;
    batched_copy<float2, NT>(
        Q * TR,
        [&](int t) {
            const int r = t % TR, kz = t / TR;
            return r < nrows ? src[(long)kz * Ny + r] : make_float2(0.f, 0.f);
        },
        [&](int t, float2 val) { lds[(t / TR) * TS + t % TR] = val; });
    __syncthreads();
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int k = g + G * e;
        v[e] = c2r_merge<float>(lds[k * TS + row], lds[(M - k) * TS + row], twP[k]);
    }
    __syncthreads();  // the tile becomes the FFT exchange buffer
    reg_fft<float, M, +1>(v, g, lds + row * F::LDS, twM);
    if (row < nrows) {
        const long N = (long)Nx * Ny * Nz;
        float* dst = h + c * N + ((long)x * Ny + y0 + row) * Nz;
        if ((Nz & 1) == 0) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int n = g + G * e;
                if (!pad_half<M>(e) && 2 * n < Nz) *reinterpret_cast<float2*>(dst + 2 * n) = v[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int n = g + G * e;
                if (!pad_half<M>(e) && 2 * n < Nz) dst[2 * n] = v[e].x;
                if (!pad_half<M>(e) && 2 * n + 1 < Nz) dst[2 * n + 1] = v[e].y;
            }
        }
    }
}

// Fused Z inverse + RK stage epilogue.  One block: TR rows (y0..y0+TR) of one x
// plane for all 3 components (3 x TR lane groups).  The 3 spectra are read
// through transposed LDS tiles, merged and transformed back in registers, and
// the fields are exchanged through LDS so that every thread sees h of all 3
// components for its cells.  The epilogue (EpiArgs) rebuilds the stage input
// m_i from m and the earlier k_j (the same fmaf sequence as the forward pass's
// stage build), evaluates the torque, and stores k_i (with the per-block max
// |k_i|) or, for the last stage of a fixed-step tableau, the normalised new m.
// h itself never leaves the chip.  NU bounds the number of earlier stages read
// (the launcher points unused slots at m with zero coefficients, which leaves
// every fmaf chain unchanged), so all loads are unconditional.
__host__ __device__ constexpr int zinv_epi_lds_elems(int M, int TR) {
    const int Q = M + 1, TS = TR + 1, HS = M / 2 > 0 ? M / 2 : 1;
    const int TILE = 3 * Q * TS, FFTB = 3 * TR * (M + M / 8), HB = 3 * TR * HS;  // RegFFT<M>::LDS = M + M/8
    return TILE > FFTB ? (TILE > HB ? TILE : HB) : (FFTB > HB ? FFTB : HB);
}
template <int M, int TR> constexpr int zinv_epi_lds() {
    static_assert(RegFFT<M>::LDS == M + M / 8, "exchange buffer size");
    return zinv_epi_lds_elems(M, TR);
}

// body: plane x, rows y0..y0+TR, all 3 components; the per-block max |k| goes to
// ea.t.part[pidx]; lds / hext_s / red: zinv_epi_lds<M,TR>() complex, 3 and
// 3 TR G floats
// FIN: 0 torque only; 1 last stage of a fixed-step tableau (m <- normalise(m +
// dt sum b_j k_j)); 2 last stage of an adaptive FSAL tableau, whose stage input
// is the high-order solution (a_Sj = b_j): it is stored to ea.mhigh and the
// error dt sum (b2_j - b_j) k_j, this stage's k included, is reduced per block
// into ea.part_err (rk.go:381-403), in place of a separate final pass.
template <int M, int TR, int NU, int FIN>
__device__ __forceinline__ void zinv_epi_tile(const float2* __restrict__ B1, const EpiArgs& ea,
                                              const float2* __restrict__ twM, const float2* __restrict__ twP,
                                              const EngineState* st, int Nx, int Ny, int Nz, int x, int y0, int pidx,
                                              float2* lds, float* hext_s, float* red) {
    using F = RegFFT<M>;
    constexpr int E = F::E, G = F::G, NR = TR * G, NT = 3 * NR, Q = M + 1, TS = TR + 1;
    constexpr int HS = M / 2 > 0 ? M / 2 : 1;  // float2 per h row (z pairs); Nz <= M
    const int c = threadIdx.x / NR, rem = threadIdx.x - c * NR, row = rem / G, g = rem % G;
    const int nrows = min(TR, Ny - y0);
    if (threadIdx.x == 0) stage_field(ea.t, st, hext_s);
    const long cst = (long)Nx * Q * Ny;
    const float2* src = B1 + (long)x * Q * Ny + y0;
    batched_copy<float2, NT>(
        3 * Q * TR,
        [&](int t) {
            const int cc = t / (Q * TR), u = t - cc * (Q * TR), kz = u / TR, r = u - kz * TR;
            return r < nrows ? src[cc * cst + (long)kz * Ny + r] : make_float2(0.f, 0.f);
        },
        [&](int t, float2 val) {
            const int cc = t / (Q * TR), u = t - cc * (Q * TR), kz = u / TR, r = u - kz * TR;
            lds[(cc * Q + kz) * TS + r] = val;
        });
    __syncthreads();
    float2 v[E];
    {
        const float2* tile = lds + c * Q * TS;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = g + G * e;
            v[e] = c2r_merge<float>(tile[k * TS + row], tile[(M - k) * TS + row], twP[k]);
        }
    }
    __syncthreads();  // the tiles become the FFT exchange buffers
    reg_fft<float, M, +1>(v, g, lds + (c * TR + row) * F::LDS, twM);
    __syncthreads();  // ... and then the h rows
    const int nh = (Nz + 1) >> 1;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int n = g + G * e;
        if (!pad_half<M>(e) && n < nh) lds[(c * TR + row) * HS + n] = v[e];
    }
    __syncthreads();

    const float hext[3] = {hext_s[0], hext_s[1], hext_s[2]};
    const float dt = st->dt_used;
    float ca[NU > 0 ? NU : 1], cb[NU > 0 ? NU : 1];
#pragma unroll
    for (int j = 0; j < NU; j++) {
        ca[j] = dt * ea.a[j];
        cb[j] = dt * ea.b[j];
    }
    const float cbs = dt * ea.bs;
    double ce[NU > 0 ? NU : 1];
#pragma unroll
    for (int j = 0; j < NU; j++) ce[j] = FIN == 2 ? ea.ce[j] * (double)dt : 0.0;
    const double ce_self = FIN == 2 ? ea.ce_self * (double)dt : 0.0;
    const long N = (long)Nx * Ny * Nz;
    float lmax = 0.f, emax = 0.f;
    double esum = 0.0;
    const bool want_e = ea.epart != nullptr;
    // one cell from the accumulated stage input mi (and final sum hi, or the
    // error of the earlier stages er) and the field h
    auto finish = [&](long ci, const float mi[3], const float hi[3], const double er[3], const float h[3],
                      float o[3]) {
        if (want_e) {
            // phi = -1/2 m.H_conv - m.B_ext - 1/2 k (m.u)^2 (field.go:64-71, gpu_local_contr3.cu:238-256)
            const double mx = mi[0], my = mi[1], mz = mi[2];
            double phi = -0.5 * (mx * h[0] + my * h[1] + mz * h[2]);
            phi -= mx * hext[0] + my * hext[1] + mz * hext[2];
            if (ea.t.anisType == 1) {
                const double mu = mx * ea.t.u[0] + my * ea.t.u[1] + mz * ea.t.u[2];
                phi -= 0.5 * ea.t.kAnis * mu * mu;
            }
            esum += phi;
        }
        float k[3], ht[3] = {h[0], h[1], h[2]};
        add_thermal(ea.t, st, ci, ht);  // not in the energy density (field.go:64-80)
        llg_torque_cell(ea.t, cell_alpha(ea.t, ci), hext, mi[0], mi[1], mi[2], ht[0], ht[1], ht[2], k);
        lmax = fmaxf(lmax, fmaxf(abs_keep_nan(k[0]), fmaxf(abs_keep_nan(k[1]), abs_keep_nan(k[2]))));
        if constexpr (FIN == 2) {
#pragma unroll
            for (int q = 0; q < 3; q++) emax = fmaxf(emax, abs_keep_nan((float)fma(ce_self, (double)k[q], er[q])));
        }
        if constexpr (FIN == 1) {
            float f[3];
#pragma unroll
            for (int q = 0; q < 3; q++) f[q] = fmaf(cbs, k[q], hi[q]);
            const float sc = norm_scale(f[0], f[1], f[2], ea.t.normMap, ci);
#pragma unroll
            for (int q = 0; q < 3; q++) o[q] = f[q] * sc;
        } else {
#pragma unroll
            for (int q = 0; q < 3; q++) o[q] = k[q];
        }
    };
    const long rowbase = ((long)x * Ny + y0) * Nz;
    if ((Nz & 1) == 0) {
        // cell pairs (2n, 2n+1) as float2
        constexpr int IT = (TR * HS + NT - 1) / NT;
        const int nzh = Nz >> 1, P = nrows * nzh;
        float2 ai[IT][3], af[IT][3];
        double ex[FIN == 2 ? IT : 1][3], ey[FIN == 2 ? IT : 1][3];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int p = threadIdx.x + it * NT;
            if (p < P) {
                const int r = p / nzh, n = p - r * nzh;
                const long gi = rowbase + (long)r * Nz + 2 * n;
#pragma unroll
                for (int q = 0; q < 3; q++) ai[it][q] = af[it][q] = *reinterpret_cast<const float2*>(ea.m + q * N + gi);
                if constexpr (FIN == 2)
#pragma unroll
                    for (int q = 0; q < 3; q++) ex[it][q] = ey[it][q] = 0.0;
#pragma unroll
                for (int j = 0; j < NU; j++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const float2 kv = *reinterpret_cast<const float2*>(ea.k[j] + q * N + gi);
                        ai[it][q].x = fmaf(ca[j], kv.x, ai[it][q].x);
                        ai[it][q].y = fmaf(ca[j], kv.y, ai[it][q].y);
                        if constexpr (FIN == 1) {
                            af[it][q].x = fmaf(cb[j], kv.x, af[it][q].x);
                            af[it][q].y = fmaf(cb[j], kv.y, af[it][q].y);
                        }
                        if constexpr (FIN == 2) {
                            ex[it][q] = fma(ce[j], (double)kv.x, ex[it][q]);
                            ey[it][q] = fma(ce[j], (double)kv.y, ey[it][q]);
                        }
                    }
            }
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int p = threadIdx.x + it * NT;
            if (p < P) {
                const int r = p / nzh, n = p - r * nzh;
                const long gi = rowbase + (long)r * Nz + 2 * n;
                float2 hv[3];
#pragma unroll
                for (int q = 0; q < 3; q++) hv[q] = lds[(q * TR + r) * HS + n];
                float mi[3], hi[3], h[3], o0[3], o1[3];
                double er[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    mi[q] = ai[it][q].x;
                    hi[q] = af[it][q].x;
                    h[q] = hv[q].x;
                    if constexpr (FIN == 2) er[q] = ex[it][q];
                }
                finish(gi, mi, hi, er, h, o0);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    mi[q] = ai[it][q].y;
                    hi[q] = af[it][q].y;
                    h[q] = hv[q].y;
                    if constexpr (FIN == 2) er[q] = ey[it][q];
                }
                finish(gi + 1, mi, hi, er, h, o1);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    *reinterpret_cast<float2*>(ea.out + q * N + gi) = make_float2(o0[q], o1[q]);
                    if constexpr (FIN == 2)
                        *reinterpret_cast<float2*>(ea.mhigh + q * N + gi) = ai[it][q];
                }
            }
        }
    } else {
        const float* hf = reinterpret_cast<const float*>(lds);
        const int P = nrows * Nz;
        for (int p = threadIdx.x; p < P; p += NT) {
            const int r = p / Nz, z = p - r * Nz;
            const long gi = rowbase + (long)r * Nz + z;
            float mi[3], hi[3], h[3], o[3];
            double er[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 3; q++) {
                mi[q] = hi[q] = ea.m[q * N + gi];
                h[q] = hf[(q * TR + r) * 2 * HS + z];
#pragma unroll
                for (int j = 0; j < NU; j++) {
                    const float kv = ea.k[j][q * N + gi];
                    mi[q] = fmaf(ca[j], kv, mi[q]);
                    hi[q] = fmaf(cb[j], kv, hi[q]);
                    if constexpr (FIN == 2) er[q] = fma(ce[j], (double)kv, er[q]);
                }
            }
            finish(gi, mi, hi, er, h, o);
#pragma unroll
            for (int q = 0; q < 3; q++) {
                ea.out[q * N + gi] = o[q];
                if constexpr (FIN == 2) ea.mhigh[q * N + gi] = mi[q];
            }
        }
    }
    if (want_e) {
        // one slot per thread in the tile buffer (3 (M+1)(TR+1) >= NT complex; the h rows are done)
        __syncthreads();
        double* dred = reinterpret_cast<double*>(lds);
        dred[threadIdx.x] = esum;
        __syncthreads();
        if (threadIdx.x == 0) {
            double e = 0.0;
            for (int i = 0; i < NT; i++) e += dred[i];
            ea.epart[pidx] = e;
        }
    }
    // per-block maximum of v (>= 0) into out[pidx]
    auto block_max = [&](float v, float* out) {
        __syncthreads();
        red[threadIdx.x] = v;
        __syncthreads();
        if constexpr (NT >= 64) {
            if (threadIdx.x < 64) {
                float vm = 0.f;
                for (int i = threadIdx.x; i < NT; i += 64) vm = fmaxf(vm, red[i]);
                for (int off = 32; off > 0; off >>= 1) vm = fmaxf(vm, __shfl_xor(vm, off, 64));
                if (threadIdx.x == 0) out[pidx] = vm;
            }
        } else {
            if (threadIdx.x == 0) {
                float vm = 0.f;
                for (int i = 0; i < NT; i++) vm = fmaxf(vm, red[i]);
                out[pidx] = vm;
            }
        }
    };
    if (ea.t.part) block_max(lmax, ea.t.part);
    if constexpr (FIN == 2) block_max(emax, ea.part_err);
}

template <int M, int TR, int NU, int FIN>
__global__ __launch_bounds__(3 * TR * RegFFT<M>::G) void k_zinv_epi(const float2* __restrict__ B1, EpiArgs ea,
                                                                      const float2* __restrict__ twM,
                                                                      const float2* __restrict__ twP,
                                                                      const EngineState* st, int Nx, int Ny, int Nz) {
    __shared__ float2 lds[zinv_epi_lds<M, TR>()];
    __shared__ float hext_s[3];
    __shared__ float red[3 * TR * RegFFT<M>::G];
    if (!st->trial_active) return;
    const int lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    zinv_epi_tile<M, TR, NU, FIN>(B1, ea, twM, twP, st, Nx, Ny, Nz, lb / gridDim.x, (lb % gridDim.x) * TR,
                                  blockIdx.y * gridDim.x + blockIdx.x, lds, hext_s, red);
}

// 3-D Y passes, software-pipelined.  One line per lane group and iteration:
//   forward  B1[c][x][kz][0..Ny) -> B2[c][kz][x][0..Py)   (zero padded)
//   inverse  B2[c][kz][x][0..Py) -> B1[c][x][kz][0..Ny)   (rows >= Ny dropped)
// Lines are numbered with x fastest (a block touches consecutive B2 lines).
// Each lane group walks `nit` lines strided by the grid and loads line i+1
// before transforming line i, so the memory of one line overlaps the FFT of
// the previous one; twiddles come from LDS, so no global load sits between a
// prefetch and its use.
// B1 line (c, x, kz) and B2 line (c, kz, x) of a Y-pass line number (x fastest)
__device__ __forceinline__ long ypass_b1(long ln, int Nx, int Ny, int Q) {
    const int x = (int)(ln % Nx);
    const long r = ln / Nx;
    const int kz = (int)(r % Q), c = (int)(r / Q);
    return (((long)c * Nx + x) * Q + kz) * Ny;
}
__device__ __forceinline__ long ypass_b2(long ln, int Nx, int Q, int PY) {
    const int x = (int)(ln % Nx);
    const long r = ln / Nx;
    const int kz = (int)(r % Q), c = (int)(r / Q);
    return (((long)c * Q + kz) * Nx + x) * PY;
}
// one line into the lane-strided registers of lane g (zeros when !ok)
template <int PY, bool FWD>
__device__ __forceinline__ void ypass_load(float2* v, const float2* __restrict__ src, long ln, bool ok, int Nx,
                                           int Ny, int Q, int g) {
    using F = RegFFT<PY>;
    if constexpr (FWD) {
        const float2* p = src + ypass_b1(ok ? ln : 0, Nx, Ny, Q);
#pragma unroll
        for (int e = 0; e < F::E; e++) {
            const int i = g + F::G * e;
            v[e] = (!pad_half<PY>(e) && ok && i < Ny) ? p[i] : make_float2(0.f, 0.f);
        }
    } else {
        const float2* p = src + ypass_b2(ok ? ln : 0, Nx, Q, PY);
#pragma unroll
        for (int e = 0; e < F::E; e++) v[e] = ok ? p[g + F::G * e] : make_float2(0.f, 0.f);
    }
}
template <int PY, bool FWD>
__device__ __forceinline__ void ypass_store(const float2* v, float2* __restrict__ dst, long ln, int Nx, int Ny,
                                            int Q, int g) {
    using F = RegFFT<PY>;
    if constexpr (FWD) {
        float2* p = dst + ypass_b2(ln, Nx, Q, PY);
#pragma unroll
        for (int e = 0; e < F::E; e++) p[g + F::G * e] = v[e];
    } else {
        float2* p = dst + ypass_b1(ln, Nx, Ny, Q);
#pragma unroll
        for (int e = 0; e < F::E; e++) {
            const int i = g + F::G * e;
            if (!pad_half<PY>(e) && i < Ny) p[i] = v[e];
        }
    }
}

template <int PY, bool FWD>
__global__ __launch_bounds__(kNT) void k_ypass(const float2* __restrict__ src, float2* __restrict__ dst,
                                               const float2* __restrict__ tw, const EngineState* st, int Nx, int Ny,
                                               int Q, int nit) {
    using F = RegFFT<PY>;
    constexpr int E = F::E, G = F::G, LPB = kNT / G;
    constexpr bool TWL = PY <= 1024;  // twiddle table in LDS (longer lines keep it in global memory)
    __shared__ float2 lds[LPB * F::LDS];
    __shared__ float2 twl[TWL ? PY : 1];
    if (st && !st->trial_active) return;
    if constexpr (TWL)
        for (int i = threadIdx.x; i < PY; i += kNT) twl[i] = tw[i];
    const int li = threadIdx.x / G, g = threadIdx.x % G;
    const long nlines = 3L * Nx * Q, stride = (long)gridDim.x * LPB;
    long line = (long)blockIdx.x * LPB + li;
    constexpr bool PIPE = E <= 16;  // registers for a second line (longer lines load in place)
    float2 nxt[PIPE ? E : 1];
    if constexpr (PIPE) ypass_load<PY, FWD>(nxt, src, line, line < nlines, Nx, Ny, Q, g);
    __syncthreads();  // twiddle table
    for (int it = 0; it < nit; it++, line += stride) {
        float2 v[E];
        if constexpr (PIPE) {
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = nxt[e];
            if (it + 1 < nit) ypass_load<PY, FWD>(nxt, src, line + stride, line + stride < nlines, Nx, Ny, Q, g);
        } else {
            ypass_load<PY, FWD>(v, src, line, line < nlines, Nx, Ny, Q, g);
        }
        reg_fft<float, PY, FWD ? -1 : +1>(v, g, lds + li * F::LDS, TWL ? twl : tw);
        if (line < nlines) ypass_store<PY, FWD>(v, dst, line, Nx, Ny, Q, g);
    }
}

// 3-D X forward . kernel . X inverse, software-pipelined.  One (kz, ky) column
// per group of G lanes, all 3 components in registers (lane-strided x = g + G e):
// the loads and stores touch G rows of consecutive ky per instruction, the
// 6-component symmetric multiply is lane-local, and each transform needs one
// LDS exchange per radix pass.  A block item is one kz and a run of lpr
// consecutive ky; a block walks `nit` consecutive items, loading the columns of
// the next item before transforming the current one.  Items are ordered so that
// the runs ky and Py - ky (which read the same kernel rows) follow each other.
// the 3 components of column (kz, ky) into lane g's registers (ky < 0: zeros)
template <int PX>
__device__ __forceinline__ void xmul_load(float2 (*v)[RegFFT<PX>::E], const float2* __restrict__ B2, int kz, int ky,
                                          int Nx, int PY, int Q, int g) {
    using F = RegFFT<PX>;
    const long cstride = (long)Q * Nx * PY;
    const float2* base = B2 + (long)kz * Nx * PY + (ky < 0 ? 0 : ky);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int e = 0; e < F::E; e++) {
            const int x = g + F::G * e;
            v[c][e] = (!pad_half<PX>(e) && ky >= 0 && x < Nx) ? base[c * cstride + (long)x * PY] : make_float2(0.f, 0.f);
        }
}
template <int PX>
__device__ __forceinline__ void xmul_store(float2 (*v)[RegFFT<PX>::E], float2* __restrict__ B2, int kz, int ky,
                                           int Nx, int PY, int Q, int g) {
    using F = RegFFT<PX>;
    const long cstride = (long)Q * Nx * PY;
    float2* base = B2 + (long)kz * Nx * PY + ky;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int e = 0; e < F::E; e++) {
            const int x = g + F::G * e;
            if (!pad_half<PX>(e) && x < Nx) base[c * cstride + (long)x * PY] = v[c][e];
        }
}
// H = K M on the spectrum of column (kz, ky): the symmetric 6-component kernel,
// stored for kx <= PX/2, ky <= PY/2, with the parity signs restored
template <int PX>
__device__ __forceinline__ void xmul_kernel(float2 (*v)[RegFFT<PX>::E], const float* __restrict__ Kc, int kz, int ky,
                                            int PY, int g) {
    GK_FFT_FP
    using F = RegFFT<PX>;
    const int hx = PX / 2 + 1, hy = PY / 2 + 1;
    const bool fy = ky > PY / 2;
    const int kyp = fy ? PY - ky : ky;
    const float* kb = Kc + ((long)kz * hy + kyp) * hx * 6;
#pragma unroll
    for (int e = 0; e < F::E; e++) {
        const int kx = g + F::G * e;
        const bool fx = kx > PX / 2;
        const int kxp = fx ? PX - kx : kx;
        const float2* kk = reinterpret_cast<const float2*>(kb + kxp * 6);
        const float2 k01 = kk[0], k23 = kk[1], k45 = kk[2];
        const float kxx = k01.x, kyy = k01.y, kzz = k23.x;
        float kyz = k23.y, kxz = k45.x, kxy = k45.y;
        if (fy) kyz = -kyz;
        if (fx) kxz = -kxz;
        if (fx != fy) kxy = -kxy;
        const float2 a = v[0][e], b = v[1][e], c = v[2][e];
        v[0][e] = make_float2(a.x * kxx + b.x * kxy + c.x * kxz, a.y * kxx + b.y * kxy + c.y * kxz);
        v[1][e] = make_float2(a.x * kxy + b.x * kyy + c.x * kyz, a.y * kxy + b.y * kyy + c.y * kyz);
        v[2][e] = make_float2(a.x * kxz + b.x * kyz + c.x * kzz, a.y * kxz + b.y * kyz + c.y * kzz);
    }
}
// column of item t for lane group li: items are (kz, run of lpr ky), with the
// runs ky and PY - ky adjacent; ky < 0 when the lane group has no column
__device__ __forceinline__ void xmul_column(long t, long items, int rows, int lpr, int li, int PY, int& kz,
                                            int& ky) {
    kz = (int)(t / rows);
    const int rr = (int)(t % rows);
    const int r = (rr & 1) ? rows - 1 - (rr >> 1) : (rr >> 1);
    ky = (t < items && li < lpr && r * lpr + li < PY) ? r * lpr + li : -1;
}

// Launch variants of the X . K . X pass (TuneParams::xv): EE elements per lane
// (the default RegFFT<PX>::E, or half of it over twice the lanes: fewer live
// registers, one more LDS exchange per transform) and PIPE, the next item's
// columns loaded before the current one is transformed (twice the data
// registers) or after.  Lane g of a line holds x = g + (PX / EE) e.
template <int PX, int EE>
__device__ __forceinline__ void xmul_load_e(float2 (*v)[EE], const float2* __restrict__ B2, int kz, int ky, int Nx,
                                            int PY, int Q, int g) {
    constexpr int G = PX / EE;
    const long cstride = (long)Q * Nx * PY;
    const float2* base = B2 + (long)kz * Nx * PY + (ky < 0 ? 0 : ky);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int e = 0; e < EE; e++) {
            const int x = g + G * e;
            const bool pad = EE >= 2 && EE % 2 == 0 && e >= EE / 2;  // x >= PX / 2 >= Nx
            v[c][e] = (!pad && ky >= 0 && x < Nx) ? base[c * cstride + (long)x * PY] : make_float2(0.f, 0.f);
        }
}
template <int PX, int EE>
__device__ __forceinline__ void xmul_store_e(float2 (*v)[EE], float2* __restrict__ B2, int kz, int ky, int Nx, int PY,
                                             int Q, int g) {
    constexpr int G = PX / EE;
    const long cstride = (long)Q * Nx * PY;
    float2* base = B2 + (long)kz * Nx * PY + ky;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int e = 0; e < EE; e++) {
            const int x = g + G * e;
            const bool pad = EE >= 2 && EE % 2 == 0 && e >= EE / 2;
            if (!pad && x < Nx) base[c * cstride + (long)x * PY] = v[c][e];
        }
}
template <int PX, int EE>
__device__ __forceinline__ void xmul_kernel_e(float2 (*v)[EE], const float* __restrict__ Kc, int kz, int ky, int PY,
                                              int g) {
    GK_FFT_FP
    constexpr int G = PX / EE;
    const int hx = PX / 2 + 1, hy = PY / 2 + 1;
    const bool fy = ky > PY / 2;
    const int kyp = fy ? PY - ky : ky;
    const float* kb = Kc + ((long)kz * hy + kyp) * hx * 6;
#pragma unroll
    for (int e = 0; e < EE; e++) {
        const int kx = g + G * e;
        const bool fx = kx > PX / 2;
        const int kxp = fx ? PX - kx : kx;
        const float2* kk = reinterpret_cast<const float2*>(kb + kxp * 6);
        const float2 k01 = kk[0], k23 = kk[1], k45 = kk[2];
        const float kxx = k01.x, kyy = k01.y, kzz = k23.x;
        float kyz = k23.y, kxz = k45.x, kxy = k45.y;
        if (fy) kyz = -kyz;
        if (fx) kxz = -kxz;
        if (fx != fy) kxy = -kxy;
        const float2 a = v[0][e], b = v[1][e], c = v[2][e];
        v[0][e] = make_float2(a.x * kxx + b.x * kxy + c.x * kxz, a.y * kxx + b.y * kxy + c.y * kxz);
        v[1][e] = make_float2(a.x * kxy + b.x * kyy + c.x * kyz, a.y * kxy + b.y * kyy + c.y * kyz);
        v[2][e] = make_float2(a.x * kxz + b.x * kyz + c.x * kzz, a.y * kxz + b.y * kyz + c.y * kzz);
    }
}
template <int PX, int EE, bool PIPE>
__global__ __launch_bounds__(kNT) void k_xmul_v(float2* __restrict__ B2, const float* __restrict__ Kc,
                                                const float2* __restrict__ tw, const EngineState* st, int Nx, int PY,
                                                int Q, int nit) {
    constexpr int E = EE, G = PX / EE, LPB = kNT / G, LDS = RegFFT<PX>::LDS;
    __shared__ float2 lds[LPB * LDS];
    __shared__ float2 twl[PX];
    if (st && !st->trial_active) return;
    for (int i = threadIdx.x; i < PX; i += kNT) twl[i] = tw[i];
    const int li = threadIdx.x / G, g = threadIdx.x % G;
    const int lpr = LPB < PY ? LPB : PY;
    const int rows = (PY + lpr - 1) / lpr;
    const long items = (long)Q * rows;
    const long t0 = (long)blockIdx.x * nit;
    float2 nxt[3][E];
    if constexpr (PIPE) {
        int kz, ky;
        xmul_column(t0, items, rows, lpr, li, PY, kz, ky);
        xmul_load_e<PX, E>(nxt, B2, kz, ky, Nx, PY, Q, g);
    }
    __syncthreads();  // twiddle table
    for (int it = 0; it < nit; it++) {
        const long t = t0 + it;
        int kz, ky;
        xmul_column(t, items, rows, lpr, li, PY, kz, ky);
        float2 v[3][E];
        if constexpr (PIPE) {
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int e = 0; e < E; e++) v[c][e] = nxt[c][e];
            if (it + 1 < nit) {
                int kz1, ky1;
                xmul_column(t + 1, items, rows, lpr, li, PY, kz1, ky1);
                xmul_load_e<PX, E>(nxt, B2, kz1, ky1, Nx, PY, Q, g);
            }
        } else {
            xmul_load_e<PX, E>(v, B2, kz, ky, Nx, PY, Q, g);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) reg_fft<float, PX, -1, 1, E>(v[c], g, lds + li * LDS, twl);
        if (ky >= 0) xmul_kernel_e<PX, E>(v, Kc, kz, ky, PY, g);
#pragma unroll
        for (int c = 0; c < 3; c++) reg_fft<float, PX, +1, 1, E>(v[c], g, lds + li * LDS, twl);
        if (ky >= 0) xmul_store_e<PX, E>(v, B2, kz, ky, Nx, PY, Q, g);
    }
}

template <int PX>
__global__ __launch_bounds__(kNT) void k_xmul(float2* __restrict__ B2, const float* __restrict__ Kc,
                                              const float2* __restrict__ tw, const EngineState* st, int Nx, int PY,
                                              int Q, int nit) {
    using F = RegFFT<PX>;
    constexpr int E = F::E, G = F::G, LPB = kNT / G;
    __shared__ float2 lds[LPB * F::LDS];
    __shared__ float2 twl[PX];
    if (st && !st->trial_active) return;
    for (int i = threadIdx.x; i < PX; i += kNT) twl[i] = tw[i];
    const int li = threadIdx.x / G, g = threadIdx.x % G;
    const int lpr = LPB < PY ? LPB : PY;  // columns per item
    const int rows = (PY + lpr - 1) / lpr;  // items per kz plane (the last may be partial)
    const long items = (long)Q * rows;
    const long t0 = (long)blockIdx.x * nit;
    float2 nxt[3][E];
    {
        int kz, ky;
        xmul_column(t0, items, rows, lpr, li, PY, kz, ky);
        xmul_load<PX>(nxt, B2, kz, ky, Nx, PY, Q, g);
    }
    __syncthreads();  // twiddle table
    for (int it = 0; it < nit; it++) {
        const long t = t0 + it;
        float2 v[3][E];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int e = 0; e < E; e++) v[c][e] = nxt[c][e];
        if (it + 1 < nit) {
            int kz1, ky1;
            xmul_column(t + 1, items, rows, lpr, li, PY, kz1, ky1);
            xmul_load<PX>(nxt, B2, kz1, ky1, Nx, PY, Q, g);
        }
        int kz, ky;
        xmul_column(t, items, rows, lpr, li, PY, kz, ky);
#pragma unroll
        for (int c = 0; c < 3; c++) reg_fft<float, PX, -1>(v[c], g, lds + li * F::LDS, twl);
        if (ky >= 0) xmul_kernel<PX>(v, Kc, kz, ky, PY, g);
#pragma unroll
        for (int c = 0; c < 3; c++) reg_fft<float, PX, +1>(v[c], g, lds + li * F::LDS, twl);
        if (ky >= 0) xmul_store<PX>(v, B2, kz, ky, Nx, PY, Q, g);
    }
}

// 3-D Y forward . X.K.X . Y inverse of one kz plane in one pass, for padded
// Y = 512 and padded X = 16 or 32 (Nx <= PX/2, Ny <= 256): per field
// evaluation B1 is read and written once and the Y-padded B2 is not used
// (48 A1 + kernel bytes instead of 48 A1 + 96 A2 + kernel for the three
// separate passes).
//
// The length-512 Y transform of a line of 256 inputs splits by the parity p of
// ky: Y(2k + p) = FFT_256(y[n] w^(p n))[k], w = exp(-2 pi i / 512), and the
// inverse, keeping n < 256, is h[n] = sum_p w^(-p n) IFFT_256(H_p)[n].  The
// two parities are independent through the X.K.X stage (a column ky only
// meets the kernel at its own ky), so a block does them one after the other
// and adds the two inverse halves.  Wave x (x < Nx) holds the 3 component
// lines of x (a 256-line is one wave, 4 elements per lane, n or k = g + 64 e);
// each round's spectra go to an LDS staging tile [k][c][x], whose columns the
// block takes through X forward, the kernel multiply and X inverse 128 at a
// time (8 lanes per column, 4 elements per lane), and come back.
//
// Block 0 packs the two real planes kz = 0 and kz = M (= Q - 1; their Z
// spectra are real) as P = A + i B.  After the Y transform, A = (P(k) +
// conj P(-k)) / 2 and B = (P(k) - conj P(-k)) / 2i (-k within the same
// parity round: a lane shuffle), and only the columns k <= its Hermitian
// partner are needed: K_A of A (kernel plane 0) and K_A of B (kernel plane M).
// Coming back, H_P = H_A + i H_B, with the partner's conjugate for the other
// half, and the real and imaginary parts go to planes 0 and M.  So Q planes
// take M = Q - 1 blocks: 256 for the 256x256x16 cuboid, one per CU.
constexpr int kPlaneThreads = 1024;
constexpr int kPlanePY = 512;
__host__ __device__ constexpr bool plane3_fits(int PX, int PY, int Nx) {
    return PY == kPlanePY && (PX == 16 || PX == 32) && Nx >= 1 && 2 * Nx <= PX;
}
template <int PX> struct Plane3 {
    static constexpr int L = kPlanePY / 2;             // half-length Y transform (256)
    static constexpr int YE = 4;                       // elements per lane of a 256-line (64 lanes)
    static constexpr int XE = 4, XG = PX / XE;         // X layout: elements / lanes per column
    static constexpr int SUB = kPlaneThreads / XG;     // columns in the X layout at once
    static constexpr int HX = PX / 2;                  // x slots of a column (Nx <= HX)
    static constexpr int SQ = 3 * HX + 1;              // staging stride per column, padded
    static constexpr int SLOTS = L + 2;                // columns of a round (block 0, even: 2 x 129)
    static constexpr int STAGE = SLOTS * SQ;
    static constexpr int XEXCH = SUB * (PX + PX / 8);
    static constexpr int YEXCH = (kPlaneThreads / 64) * (L + L / 8);
    static constexpr int EXCH = XEXCH > YEXCH ? XEXCH : YEXCH;
    static constexpr int LDS = STAGE + EXCH + L + kPlanePY + PX;  // + twiddle tables
};

template <int PX>
__global__ __launch_bounds__(kPlaneThreads) void k_plane3(float2* __restrict__ B1, const float* __restrict__ Kc,
                                                          const float2* __restrict__ twLg,
                                                          const float2* __restrict__ twPg,
                                                          const float2* __restrict__ twXg, const EngineState* st,
                                                          int Nx, int Ny, int Q) {
    GK_FFT_FP
    using P3 = Plane3<PX>;
    constexpr int L = P3::L, YE = P3::YE, XE = P3::XE, XG = P3::XG, SQ = P3::SQ, HX = P3::HX, SUB = P3::SUB;
    constexpr int PY = kPlanePY;
    __shared__ float2 lds[P3::LDS];
    if (st && !st->trial_active) return;
    float2* stage = lds;                      // [SLOTS][SQ]
    float2* exch = lds + P3::STAGE;           // Y (per wave) or X (per column) exchange buffers
    float2* twL = exch + P3::EXCH;            // exp(-2 pi i j / 256)
    float2* twP = twL + L;                    // exp(-2 pi i j / 512)
    float2* twX = twP + PY;                   // exp(-2 pi i j / PX)
    for (int i = threadIdx.x; i < L; i += kPlaneThreads) twL[i] = twLg[i];
    for (int i = threadIdx.x; i < PY; i += kPlaneThreads) twP[i] = twPg[i];
    if (threadIdx.x < PX) twX[threadIdx.x] = twXg[threadIdx.x];
    const int M = Q - 1;
    const bool packed = blockIdx.x == 0;
    const int kz = blockIdx.x;  // plane of an unpacked block
    const int x = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), g = threadIdx.x % 64;  // wave x, lane g
    const bool xok = x < Nx;
    const long cs = (long)Nx * Q * Ny;
    const long row = ((long)(xok ? x : 0) * Q + kz) * Ny;
    const long rowM = ((long)(xok ? x : 0) * Q + M) * Ny;
    __syncthreads();  // twiddle tables
    const int q = threadIdx.x / XG, gx = threadIdx.x % XG;  // X layout: column q of a sub-round, lane gx
    const int hx = PX / 2 + 1, hy = PY / 2 + 1;
    float2 half0[3][YE];  // the even round's inverse half of the lines
#pragma unroll
    for (int p = 0; p < 2; p++) {
        // columns of block 0: A for k < KA, B for KA <= slot < 2 KA (KA = 129 even, 128 odd)
        const int KA = p ? L / 2 : L / 2 + 1;
        const int nslots = packed ? 2 * KA : L;
        // ---- the plane's lines (read again for the odd round rather than held in
        // registers through the even one), then Y forward of this parity:
        // FFT_256(y[n] w^(p n)); loads from clamped indices and a select, so no
        // divergent branch sits around them
        float2 u[3][YE];
        if (packed) {
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int e = 0; e < YE; e++) {
                    const int n = g + 64 * e, nn = n < Ny ? n : Ny - 1;
                    const float2 a = B1[c * cs + row + nn], b = B1[c * cs + rowM + nn];
                    u[c][e] = (xok && n < Ny) ? make_float2(a.x, b.x) : make_float2(0.f, 0.f);
                }
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int e = 0; e < YE; e++) {
                    const int n = g + 64 * e, nn = n < Ny ? n : Ny - 1;
                    const float2 a = B1[c * cs + row + nn];
                    u[c][e] = (xok && n < Ny) ? a : make_float2(0.f, 0.f);
                }
        }
        if (p && xok) {
            // the even round's half of h parks in B1 (its inputs are read): registers stay free for the odd round
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int e = 0; e < YE; e++) {
                    const int n = g + 64 * e;
                    if (n < Ny) B1[c * cs + row + n] = half0[c][e];
                }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (p)
#pragma unroll
                for (int e = 0; e < YE; e++) u[c][e] = cmul<float>(u[c][e], twP[g + 64 * e]);
            reg_fft<float, L, -1, 1, YE, true>(u[c], g, exch + x * (L + L / 8), twL);
        }
        // ---- into the staging tile
        if (xok) {
            if (!packed) {
#pragma unroll
                for (int e = 0; e < YE; e++)
#pragma unroll
                    for (int c = 0; c < 3; c++) stage[(g + 64 * e) * SQ + c * HX + x] = u[c][e];
            } else {
                // partner -k of k = g + 64 e: lane (64 - g) % 64 (even) or 63 - g (odd), element 3 - e;
                // for g = 0 of the even round, this lane's element (4 - e) % 4
                const int src = p ? 63 - g : (64 - g) & 63;
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int e = 0; e < YE; e++) {
                        const float2 ta = shfl2<64>(u[c][YE - 1 - e], src);
                        const float2 pm = (p == 0 && g == 0) ? u[c][(YE - e) % YE] : ta;
                        const int k = g + 64 * e;
                        if (k < KA) {
                            const float2 a = u[c][e];
                            stage[k * SQ + c * HX + x] = make_float2((a.x + pm.x) * 0.5f, (a.y - pm.y) * 0.5f);
                            stage[(KA + k) * SQ + c * HX + x] =
                                make_float2((a.y + pm.y) * 0.5f, (pm.x - a.x) * 0.5f);
                        }
                    }
            }
        }
        __syncthreads();  // the Y exchange buffers become the X ones
        // ---- X . K . X of every column, SUB at a time
        for (int j = 0; j * SUB < nslots; j++) {
            const int slot = j * SUB + q;
            const bool ok = slot < nslots;
            float2* col = stage + (ok ? slot : 0) * SQ;
            // kernel plane and ky of the column
            const bool bcol = packed && slot >= KA;
            const int kzk = packed ? (bcol ? M : 0) : kz;
            const int ky = 2 * (bcol ? slot - KA : slot) + p;
            float2 xv[3][XE];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int e = 0; e < XE; e++) {
                    const int xx = gx + XG * e;
                    xv[c][e] = (ok && 2 * e < XE && xx < Nx) ? col[c * HX + xx] : make_float2(0.f, 0.f);
                }
#pragma unroll
            for (int c = 0; c < 3; c++) reg_fft<float, PX, -1, 1, XE, true>(xv[c], gx, exch + q * (PX + PX / 8), twX);
            if (ok) {
                // H = K M with the compressed kernel's parity signs restored (as xmul_kernel)
                const bool fy = ky > PY / 2;
                const float* kb = Kc + ((long)kzk * hy + (fy ? PY - ky : ky)) * hx * 6;
#pragma unroll
                for (int e = 0; e < XE; e++) {
                    const int kx = gx + XG * e;
                    const bool fx = kx > PX / 2;
                    const float2* kk = reinterpret_cast<const float2*>(kb + (fx ? PX - kx : kx) * 6);
                    const float2 k01 = kk[0], k23 = kk[1], k45 = kk[2];
                    const float kxx = k01.x, kyy = k01.y, kzz = k23.x;
                    const float kyz = fy ? -k23.y : k23.y, kxz = fx ? -k45.x : k45.x;
                    const float kxy = (fx != fy) ? -k45.y : k45.y;
                    const float2 a = xv[0][e], b = xv[1][e], cc = xv[2][e];
                    xv[0][e] = make_float2(a.x * kxx + b.x * kxy + cc.x * kxz, a.y * kxx + b.y * kxy + cc.y * kxz);
                    xv[1][e] = make_float2(a.x * kxy + b.x * kyy + cc.x * kyz, a.y * kxy + b.y * kyy + cc.y * kyz);
                    xv[2][e] = make_float2(a.x * kxz + b.x * kyz + cc.x * kzz, a.y * kxz + b.y * kyz + cc.y * kzz);
                }
            }
#pragma unroll
            for (int c = 0; c < 3; c++) reg_fft<float, PX, +1, 1, XE, true>(xv[c], gx, exch + q * (PX + PX / 8), twX);
            if (ok)
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int e = 0; e < XE; e++) {
                        const int xx = gx + XG * e;
                        if (2 * e < XE && xx < Nx) col[c * HX + xx] = xv[c][e];
                    }
        }
        __syncthreads();  // the staging tile is complete; the X exchange buffers become the Y ones
        // ---- back into the lines and Y inverse of this parity
#pragma unroll
        for (int e = 0; e < YE; e++) {
            const int k = g + 64 * e;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float2 h = make_float2(0.f, 0.f);
                if (xok) {
                    if (!packed) {
                        h = stage[k * SQ + c * HX + x];
                    } else {
                        // H_P = H_A + i H_B; beyond KA from the Hermitian partner, conjugated
                        const bool low = k < KA;
                        const int kk = low ? k : (p ? L - 1 - k : L - k);
                        float2 ha = stage[kk * SQ + c * HX + x], hb = stage[(KA + kk) * SQ + c * HX + x];
                        if (!low) {
                            ha.y = -ha.y;
                            hb.y = -hb.y;
                        }
                        h = make_float2(ha.x - hb.y, ha.y + hb.x);
                    }
                }
                u[c][e] = h;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            reg_fft<float, L, +1, 1, YE, true>(u[c], g, exch + x * (L + L / 8), twL);
            if (p) {
#pragma unroll
                for (int e = 0; e < YE; e++) {
                    float2 w = twP[g + 64 * e];
                    w.y = -w.y;
                    u[c][e] = cmul<float>(u[c][e], w);
                }
            }
        }
        if (p == 0) {
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int e = 0; e < YE; e++) half0[c][e] = u[c][e];
        } else if (xok) {
            // h = even half (parked in B1 by this lane) + odd half
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int e = 0; e < YE; e++) {
                    const int n = g + 64 * e;
                    if (n < Ny) {
                        const float2 h0 = B1[c * cs + row + n];
                        const float2 h = make_float2(h0.x + u[c][e].x, h0.y + u[c][e].y);
                        if (packed) {
                            B1[c * cs + row + n] = make_float2(h.x, 0.f);
                            B1[c * cs + rowM + n] = make_float2(h.y, 0.f);
                        } else {
                            B1[c * cs + row + n] = h;
                        }
                    }
                }
        }
        __syncthreads();  // the staging tile and exchange buffers are free for the next round
    }
}

// Thin 3-D grids (padded X of 4 to 16): Y forward, X . K . X and Y inverse of
// one kz plane in one pass.  The plane of Y spectra [3][Nx][PY] stays in LDS,
// so per field evaluation the pass reads and writes B1 once, where the three
// separate passes move B1 twice and the Y-padded B2 four times.  The X
// transforms of length PX (Nx <= PX/2 nonzero inputs, Nx outputs kept) are
// direct sums with the table exp(-2 pi i j / PX).
constexpr int kYxyThreads = 256;
constexpr int kYxyLdsBytes = 80 * 1024;
// Which plans take the plane pass: a plane of 3 (PX/2) lines plus separate
// FFT exchange buffers within 80 KB (the sizes its registers were sized for).
__host__ __device__ constexpr bool yxy_fits(int PX, int PY) {
    return PX >= 4 && PX <= 16 && reg_fft_length(PY) &&
           (3 * (PX / 2) * PY + (kYxyThreads / reg_fft_G(PY)) * (PY + PY / 8)) * 8 <= kYxyLdsBytes;
}
// LDS of the pass: the plane's rows l = r LPB + li at l PY; the FFT exchange
// buffers overlay the rows of the last round of lines (r = RMAX - 1), which are
// written only after every lane group's last forward FFT and read back before
// any inverse FFT (yxy_plane), so the two share memory: ~50 KB instead of
// ~66 KB for PY = 512, PX = 8, three blocks per CU instead of two.
__host__ __device__ constexpr int yxy_lds_elems(int PX, int PY) {
    return (3 * (PX / 2) * PY) >
                   ((3 * (PX / 2) + kYxyThreads / reg_fft_G(PY) - 1) / (kYxyThreads / reg_fft_G(PY)) - 1) *
                           (kYxyThreads / reg_fft_G(PY)) * PY +
                       (kYxyThreads / reg_fft_G(PY)) * (PY + PY / 8)
               ? 3 * (PX / 2) * PY
               : ((3 * (PX / 2) + kYxyThreads / reg_fft_G(PY) - 1) / (kYxyThreads / reg_fft_G(PY)) - 1) *
                         (kYxyThreads / reg_fft_G(PY)) * PY +
                     (kYxyThreads / reg_fft_G(PY)) * (PY + PY / 8);
}
template <int PY, int PX> constexpr int yxy_lds() { return yxy_lds_elems(PX, PY); }
template <int PY, int PX> struct Yxy {
    using F = RegFFT<PY>;
    static constexpr int E = F::E, G = F::G, LPB = kYxyThreads / G;
    static constexpr int RMAX = (3 * (PX / 2) + LPB - 1) / LPB;  // rounds of lines per lane group, at most
};

// the 3 Nx lines (l = c Nx + x) of plane kz into the lane group's registers
template <int PY, int PX>
__device__ __forceinline__ void yxy_load(float2 (*v)[RegFFT<PY>::E], const float2* __restrict__ B1, int Nx, int Ny,
                                         int Q, int kz) {
    using Y = Yxy<PY, PX>;
    const int li = threadIdx.x / Y::G, g = threadIdx.x % Y::G;
#pragma unroll
    for (int r = 0; r < Y::RMAX; r++) {
        const int l = r * Y::LPB + li;
        const bool ok = l < 3 * Nx;
        const float2* src = B1 + ((long)(ok ? l : 0) * Q + kz) * Ny;  // B1[c][x][kz][y], (c, x) = line l
#pragma unroll
        for (int e = 0; e < Y::E; e++) {
            const int i = g + Y::G * e;
            v[r][e] = (!pad_half<PY>(e) && ok && i < Ny) ? src[i] : make_float2(0.f, 0.f);
        }
    }
}

// the plane pass on loaded lines (called by every thread of the block)
template <int PY, int PX>
__device__ __forceinline__ void yxy_plane(float2 (*v)[RegFFT<PY>::E], float2* __restrict__ B1,
                                          const float* __restrict__ Kc, const float2* twY, const float2* wx, int Nx,
                                          int Ny, int Q, int kz, float2* lds) {
    using Y = Yxy<PY, PX>;
    using F = RegFFT<PY>;
    constexpr int E = Y::E, G = Y::G, NT = kYxyThreads, LPB = Y::LPB, RMAX = Y::RMAX;
    constexpr int hx = PX / 2 + 1, hy = PY / 2 + 1;
    float2* plane = lds;                                  // [3 Nx][PY]
    float2* exch = lds + (long)(RMAX - 1) * LPB * PY;     // LPB exchange buffers over the last round's rows
    static_assert((RMAX - 1) * LPB * PY + LPB * F::LDS <= yxy_lds<PY, PX>(), "LDS layout");
    const int li = threadIdx.x / G, g = threadIdx.x % G;
    const int nl = 3 * Nx;
    // Y forward
#pragma unroll
    for (int r = 0; r < RMAX; r++) {
        const int l = r * LPB + li;
        if (r * LPB < nl) {  // uniform: some lane group has a line in this round
            reg_fft<float, PY, -1>(v[r], g, exch + li * F::LDS, twY);
            // the last round's rows hold the exchange buffers until every lane group is through
            if (r == RMAX - 1) __syncthreads();
            if (l < nl)
#pragma unroll
                for (int e = 0; e < E; e++) plane[(long)l * PY + g + G * e] = v[r][e];
        }
    }
    __syncthreads();
    // X . K . X for every ky of the plane
    for (int ky = threadIdx.x; ky < PY; ky += NT) {
        GK_FFT_FP
        constexpr bool BF = PX == 4 || PX == 8;  // radix-4 / radix-8 butterflies, else direct sums
        float2 a[3][PX];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int kx = 0; kx < PX; kx++) a[c][kx] = make_float2(0.f, 0.f);
        if constexpr (BF) {
#pragma unroll
            for (int x = 0; x < PX / 2; x++)
                if (x < Nx)
#pragma unroll
                    for (int c = 0; c < 3; c++) a[c][x] = plane[(long)(c * Nx + x) * PY + ky];
#pragma unroll
            for (int c = 0; c < 3; c++) dft<float, PX, -1>(a[c]);
        } else {
#pragma unroll
            for (int x = 0; x < PX / 2; x++) {
                if (x < Nx) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        const float2 u = plane[(long)(c * Nx + x) * PY + ky];
#pragma unroll
                        for (int kx = 0; kx < PX; kx++)
                            a[c][kx] = cadd<float>(a[c][kx], cmul<float>(u, wx[(kx * x) % PX]));
                    }
                }
            }
        }
        // kernel layout of plane-pass plans: [kz][component][kx <= PX/2][ky <= PY/2], so that
        // consecutive ky (threads) read consecutive floats
        const bool fy = ky > PY / 2;
        const int kyp = fy ? PY - ky : ky;
        const float* kb = Kc + (long)kz * 6 * hx * hy + kyp;
#pragma unroll
        for (int kx = 0; kx < PX; kx++) {
            const bool fx = kx > PX / 2;
            const float* kk = kb + (fx ? PX - kx : kx) * hy;
            constexpr int cs = hx * hy;
            const float kxx = kk[0], kyy = kk[cs], kzz = kk[2 * cs];
            const float kyz = fy ? -kk[3 * cs] : kk[3 * cs], kxz = fx ? -kk[4 * cs] : kk[4 * cs];
            const float kxy = (fx != fy) ? -kk[5 * cs] : kk[5 * cs];
            const float2 p = a[0][kx], q = a[1][kx], w = a[2][kx];
            a[0][kx] = make_float2(p.x * kxx + q.x * kxy + w.x * kxz, p.y * kxx + q.y * kxy + w.y * kxz);
            a[1][kx] = make_float2(p.x * kxy + q.x * kyy + w.x * kyz, p.y * kxy + q.y * kyy + w.y * kyz);
            a[2][kx] = make_float2(p.x * kxz + q.x * kyz + w.x * kzz, p.y * kxz + q.y * kyz + w.y * kzz);
        }
        if constexpr (BF) {
#pragma unroll
            for (int c = 0; c < 3; c++) dft<float, PX, +1>(a[c]);
#pragma unroll
            for (int x = 0; x < PX / 2; x++)
                if (x < Nx)
#pragma unroll
                    for (int c = 0; c < 3; c++) plane[(long)(c * Nx + x) * PY + ky] = a[c][x];
        } else {
#pragma unroll
            for (int x = 0; x < PX / 2; x++) {
                if (x < Nx) {
#pragma unroll
                    for (int c = 0; c < 3; c++) {
                        float2 t = make_float2(0.f, 0.f);
#pragma unroll
                        for (int kx = 0; kx < PX; kx++)
                            t = cadd<float>(t, cmul<float>(a[c][kx], cconj<float>(wx[(kx * x) % PX])));
                        plane[(long)(c * Nx + x) * PY + ky] = t;
                    }
                }
            }
        }
    }
    __syncthreads();
    // Y inverse, first Ny kept; the last round first, whose rows the exchange
    // buffers take over once every lane group has read them
#pragma unroll
    for (int r = RMAX - 1; r >= 0; r--) {
        if (r * LPB >= nl) continue;  // uniform
        const int l = r * LPB + li;
        const bool ok = l < nl;
        float2 w[E];
#pragma unroll
        for (int e = 0; e < E; e++) w[e] = ok ? plane[(long)l * PY + g + G * e] : make_float2(0.f, 0.f);
        if (r == RMAX - 1) __syncthreads();
        reg_fft<float, PY, +1>(w, g, exch + li * F::LDS, twY);
        if (ok) {
            float2* dst = B1 + ((long)l * Q + kz) * Ny;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int i = g + G * e;
                if (!pad_half<PY>(e) && i < Ny) dst[i] = w[e];
            }
        }
    }
}

// single simulation: one block per kz plane
template <int PY, int PX>
__global__ __launch_bounds__(kYxyThreads) void k_yxy(float2* __restrict__ B1, const float* __restrict__ Kc,
                                                      const float2* __restrict__ twY, const float2* __restrict__ twX,
                                                      const EngineState* st, int Nx, int Ny, int Q) {
    using Y = Yxy<PY, PX>;
    __shared__ float2 lds[yxy_lds<PY, PX>()];
    if (st && !st->trial_active) return;
    float2 wx[PX];
#pragma unroll
    for (int j = 0; j < PX; j++) wx[j] = twX[j];
    float2 v[Y::RMAX][Y::E];
    yxy_load<PY, PX>(v, B1, Nx, Ny, Q, blockIdx.x);
    yxy_plane<PY, PX>(v, B1, Kc, twY, wx, Nx, Ny, Q, blockIdx.x, lds);
}

// 2-D: Y forward . kernel . Y inverse of one kz column (3 components) in
// registers, in place in B1[c][0][kz][y].  Lane group li of a block takes
// column kz; the 4 stored kernel components (xx, yy, zz, yz) for ky <= PY/2
// are read with the yz sign restored for ky > PY/2.  Components c0..c0+NC.
template <int PY, int NC = 3>
__device__ __forceinline__ void y2r_load(float2 (*v)[RegFFT<PY>::E], const float2* __restrict__ B1, long kz, bool ok,
                                         int Ny, int Q, int g, int c0 = 0) {
    using F = RegFFT<PY>;
    const long cs = (long)Q * Ny;
    const float2* p = B1 + (ok ? kz : 0) * Ny;
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int e = 0; e < F::E; e++) {
            const int i = g + F::G * e;
            v[c][e] = (!pad_half<PY>(e) && ok && i < Ny) ? p[(c0 + c) * cs + i] : make_float2(0.f, 0.f);
        }
}
template <int PY, int NC = 3>
__device__ __forceinline__ void y2r_store(float2 (*v)[RegFFT<PY>::E], float2* __restrict__ B1, long kz, int Ny, int Q,
                                          int g, int c0 = 0) {
    using F = RegFFT<PY>;
    const long cs = (long)Q * Ny;
    float2* p = B1 + kz * Ny;
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int e = 0; e < F::E; e++) {
            const int i = g + F::G * e;
            if (!pad_half<PY>(e) && i < Ny) p[(c0 + c) * cs + i] = v[c][e];
        }
}
// kernel multiply; C0 = 0: all 3 components in v[0..2]; C0 = 1: only x in v[0];
// C0 = 2: only y, z in v[0..1]
template <int PY, int C0 = 0>
__device__ __forceinline__ void y2r_kernel(float2 (*v)[RegFFT<PY>::E], const float* __restrict__ Kc, long kz, int Q,
                                           int g) {
    GK_FFT_FP
    using F = RegFFT<PY>;
    constexpr int hy = PY / 2 + 1;
    const long cs = (long)Q * hy;
#pragma unroll
    for (int e = 0; e < F::E; e++) {
        const int ky = g + F::G * e;
        const bool fy = ky > PY / 2;
        const long b = kz * hy + (fy ? PY - ky : ky);
        if constexpr (C0 <= 1) {
            const float kxx = Kc[b];
            const float2 a = v[0][e];
            v[0][e] = make_float2(a.x * kxx, a.y * kxx);
        }
        if constexpr (C0 != 1) {
            constexpr int o = C0 == 0 ? 1 : 0;
            const float kyy = Kc[cs + b], kzz = Kc[2 * cs + b];
            const float kyz = fy ? -Kc[3 * cs + b] : Kc[3 * cs + b];
            const float2 bb = v[o][e], cc = v[o + 1][e];
            v[o][e] = make_float2(bb.x * kyy + cc.x * kyz, bb.y * kyy + cc.y * kyz);
            v[o + 1][e] = make_float2(bb.x * kyz + cc.x * kzz, bb.y * kyz + cc.y * kzz);
        }
    }
}
// transforms of one loaded column (every lane group of the block calls it: the
// exchanges synchronise the block)
template <int PY, int NC = 3, int C0 = 0>
__device__ __forceinline__ void y2r_column(float2 (*v)[RegFFT<PY>::E], const float* __restrict__ Kc, long kz, bool ok,
                                           int Q, int g, float2* lds, const float2* tw) {
#pragma unroll
    for (int c = 0; c < NC; c++) reg_fft<float, PY, -1>(v[c], g, lds, tw);
    if (ok) y2r_kernel<PY, C0>(v, Kc, kz, Q, g);
#pragma unroll
    for (int c = 0; c < NC; c++) reg_fft<float, PY, +1>(v[c], g, lds, tw);
}
// long lines (more than 16 elements per lane): x alone, then the y-z pair,
// to bound the registers of one column
template <int PY>
__device__ __forceinline__ void y2r_column_split(float2* __restrict__ B1, const float* __restrict__ Kc, long kz,
                                                 bool ok, int Ny, int Q, int g, float2* lds, const float2* tw) {
    constexpr int E = RegFFT<PY>::E;
    {
        float2 v[1][E];
        y2r_load<PY, 1>(v, B1, kz, ok, Ny, Q, g, 0);
        y2r_column<PY, 1, 1>(v, Kc, kz, ok, Q, g, lds, tw);
        if (ok) y2r_store<PY, 1>(v, B1, kz, Ny, Q, g, 0);
    }
    {
        float2 v[2][E];
        y2r_load<PY, 2>(v, B1, kz, ok, Ny, Q, g, 1);
        y2r_column<PY, 2, 2>(v, Kc, kz, ok, Q, g, lds, tw);
        if (ok) y2r_store<PY, 2>(v, B1, kz, Ny, Q, g, 1);
    }
}

// single simulation: persistent over kz columns, the next column loaded while
// the current one is transformed (when its registers fit)
template <int PY>
__global__ __launch_bounds__(kNT) void k_y2r(float2* __restrict__ B1, const float* __restrict__ Kc,
                                             const float2* __restrict__ tw, const EngineState* st, int Ny, int Q,
                                             int nit) {
    using F = RegFFT<PY>;
    constexpr int E = F::E, G = F::G, LPB = kNT / G;
    constexpr bool TWL = PY <= 1024;
    __shared__ float2 lds[LPB * F::LDS];
    __shared__ float2 twl[TWL ? PY : 1];
    if (st && !st->trial_active) return;
    if constexpr (TWL)
        for (int i = threadIdx.x; i < PY; i += kNT) twl[i] = tw[i];
    const int li = threadIdx.x / G, g = threadIdx.x % G;
    const long stride = (long)gridDim.x * LPB;
    long kz = (long)blockIdx.x * LPB + li;
    constexpr bool PIPE = E <= 8;
    float2 nxt[PIPE ? 3 : 1][PIPE ? E : 1];
    if constexpr (PIPE) y2r_load<PY>(nxt, B1, kz, kz < Q, Ny, Q, g);
    __syncthreads();  // twiddle table
    if constexpr (E > 16) {
        for (int it = 0; it < nit; it++, kz += stride)
            y2r_column_split<PY>(B1, Kc, kz, kz < Q, Ny, Q, g, lds + li * F::LDS, TWL ? twl : tw);
    } else {
        for (int it = 0; it < nit; it++, kz += stride) {
            float2 v[3][E];
            if constexpr (PIPE) {
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int e = 0; e < E; e++) v[c][e] = nxt[c][e];
                if (it + 1 < nit) y2r_load<PY>(nxt, B1, kz + stride, kz + stride < Q, Ny, Q, g);
            } else {
                y2r_load<PY>(v, B1, kz, kz < Q, Ny, Q, g);
            }
            y2r_column<PY>(v, Kc, kz, kz < Q, Q, g, lds + li * F::LDS, TWL ? twl : tw);
            if (kz < Q) y2r_store<PY>(v, B1, kz, Ny, Q, g);
        }
    }
}

}  // namespace gk
