// Demag(+exchange) convolution for the device-resident engine.
//
// Reference behaviour restated: conv.go:57-100 (three padded forward FFTs,
// k-space multiply with the symmetric kernel, three inverse FFTs and unpad),
// conv.go:110-167 (kernel spectrum = Re FFT(K) / (Px Py Pz)), demag.go:26-84,142-183
// (FaceKernel6), exchange.go:44-69 + init.go:256-267 (exchange folded in).
//
// MI355X design.  Intermediate spectra are kept in transposed layouts so that
// every FFT pass reads and writes whole contiguous lines (coalesced 64-lane
// accesses) and no separate transpose or pad/unpad pass exists:
//   m[c][x][y][z]  --Z r2c (stage build fused)-->  B1[c][x][kz][y]   (y fastest)
//   3-D: B1 --Y fwd--> B2[c][kz][x][ky] --X fwd . K . X inv (in place)--> B2
//        --Y inv--> B1 --Z c2r--> h[c][x][y][z]
//   2-D: B1 --Y fwd . K . Y inv (in place)--> B1 --Z c2r--> h
// Only data lines that can be non-zero are transformed (rows y >= Ny and
// planes x >= Nx are zero padding), and only the kept half of each inverse
// transform is stored, so B1/B2 hold Nx (not Px) planes.
//
// The kernel spectrum is real and each component is even or odd along every
// axis (Kxy odd in x and y, Kxz odd in x, Kyz odd in y, diagonal even), so
// only kx <= Px/2, ky <= Py/2 is stored (Kc), a quarter of the reference's
// 6 x Px x Py x (Pz/2+1) floats; signs are restored on the fly.  3-D: Kc is
// [kz][ky][kx][6] (the 6 components of one frequency are 24 contiguous bytes,
// read as 3 float2 by the lane that needs them); 2-D: [4][kz][ky].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "fft_device.h"
#include "gk_fft.h"

namespace gk {

// =============================================================== kernel setup
__constant__ int c_kern_idx[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};  // xx yy zz yz xz xy

// faceIntegral (demag.go:142-183) in FP64
__device__ void face_integral_f64(double B[3], const double R[3], const double cell[3], int s, int n) {
    const int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
    const double charge = cell[v] * cell[w];
    const double pu = cell[u] * 0.5;
    const double inv4pi = 1.0 / (4.0 * 3.14159265358979323846);
    B[0] = B[1] = B[2] = 0.0;
    for (int i = 0; i < n; i++) {
        const double pv = -0.5 * cell[v] + cell[v] / (2.0 * n) + i * (cell[v] / n);
        for (int j = 0; j < n; j++) {
            const double pw = -0.5 * cell[w] + cell[w] / (2.0 * n) + j * (cell[w] / n);
            for (int pole = 0; pole < 2; pole++) {
                double pos[3];
                pos[u] = pole == 0 ? pu : -pu;
                pos[v] = pv;
                pos[w] = pw;
                const double d0 = R[0] - pos[0], d1 = R[1] - pos[1], d2 = R[2] - pos[2];
                const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                const double r = sqrt(r2);
                const double q = (pole == 0 ? charge : -charge) * inv4pi / (r2 * r);
                B[0] += d0 * q;
                B[1] += d1 * q;
                B[2] += d2 * q;
            }
        }
    }
    const double inv = 1.0 / (double)(n * n);
    B[0] *= inv;
    B[1] *= inv;
    B[2] *= inv;
}

// wrapped index -> displacement within FaceKernel6's loop range [-(P-1)/2, P/2-1]
// (demag.go:35-53); false for the index no displacement maps to.
__device__ __forceinline__ bool unwrap(int w, int P, int& x) {
    int x1 = -((P - 1) / 2), x2 = P / 2 - 1;
    if (P == 1) x1 = x2 = 0;
    if (w <= x2) {
        x = w;
        return true;
    }
    x = w - P;
    return x >= x1;
}

__global__ void k_facekernel6(double* __restrict__ K, int P0, int P1, int P2, double c0, double c1, double c2,
                              int acc) {
    const long NP = (long)P0 * P1 * P2;
    const double cell[3] = {c0, c1, c2};
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < NP; idx += (long)gridDim.x * blockDim.x) {
        const int zw = (int)(idx % P2);
        const long r = idx / P2;
        const int yw = (int)(r % P1), xw = (int)(r / P1);
        int x, y, z;
        double kk[6] = {0, 0, 0, 0, 0, 0};
        if (unwrap(xw, P0, x) && unwrap(yw, P1, y) && unwrap(zw, P2, z)) {
            const double R[3] = {x * c0, y * c1, z * c2};
            for (int s = 0; s < 3; s++) {
                double B[3];
                face_integral_f64(B, R, cell, s, acc);
                for (int d = s; d < 3; d++) kk[c_kern_idx[s][d]] = B[d];
            }
        }
        for (int k = 0; k < 6; k++) K[k * NP + idx] = kk[k];
    }
}

struct ExchPoints {
    long idx[7];
    double val[7];
    int n;
};
__global__ void k_add_exch_points(double* K, long NP, ExchPoints p) {
    const int t = threadIdx.x;
    if (t < 3 * p.n) {
        const int comp = t / p.n, i = t % p.n;  // comps 0,1,2 = xx, yy, zz
        K[comp * NP + p.idx[i]] += p.val[i];
    }
}

__global__ void k_real_to_complex(const double* __restrict__ in, double2* __restrict__ out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = make_double2(in[i], 0.0);
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* a, double v) {
    atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// spectrum -> compressed kernel (and optionally the reference's full layout)
__global__ void k_extract_kernel(const double2* __restrict__ W, int P0, int P1, int P2, int Q, double norm,
                                 float* __restrict__ Kc, int comp, int nK, float* __restrict__ full_comp,
                                 unsigned long long* maxre, unsigned long long* maxim) {
    const long total = (long)P0 * P1 * Q;
    const int hx = P0 / 2 + 1, hy = P1 / 2 + 1;
    double lre = 0, lim = 0;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int kz = (int)(t % Q);
        const long r = t / Q;
        const int ky = (int)(r % P1), kx = (int)(r / P1);
        const double2 w = W[((long)kx * P1 + ky) * P2 + kz];
        const double re = w.x * norm, im = w.y * norm;
        lre = fmax(lre, fabs(re));
        lim = fmax(lim, fabs(im));
        if (full_comp) full_comp[t] = (float)re;
        if (kx < hx && ky < hy) {
            if (P0 == 1) Kc[((long)comp * Q + kz) * hy + ky] = (float)re;
            else Kc[(((long)kz * hy + ky) * hx + kx) * nK + comp] = (float)re;
        }
    }
    atomic_max_nonneg(maxre, lre);
    atomic_max_nonneg(maxim, lim);
}

// =============================================================== plan
static int pad_len(int n) {
    // smallest supported FFT length >= 2n-1 (exact linear convolution, any P >= 2N-1)
    if (n == 1) return 1;
    int p = 1;
    while (p < 2 * n - 1) p *= 2;
    return p;
}

void conv_plan_init(ConvPlan& cp, const int N[3]) {
    GK_ASSERT(N[2] >= 2, "the Z (last user x) size must be > 1, got %d (init.go:127-129)", N[2]);
    for (int i = 0; i < 3; i++) {
        cp.N[i] = N[i];
        cp.P[i] = pad_len(N[i]);
    }
    if (cp.P[2] < 2) cp.P[2] = 2;
    cp.M = cp.P[2] / 2;
    cp.Q = cp.M + 1;
    cp.nK = cp.P[0] == 1 ? 4 : 6;
    GK_ASSERT(cp.M <= 2048 && cp.P[1] <= 4096 && cp.P[0] <= 1024, "grid too large for this build (padded %d %d %d)",
              cp.P[0], cp.P[1], cp.P[2]);
    const long hx = cp.P[0] / 2 + 1, hy = cp.P[1] / 2 + 1;
    const long kc = cp.P[0] == 1 ? (long)cp.nK * cp.Q * hy : (long)cp.nK * cp.Q * hy * hx;
    const long b1 = 3L * N[0] * cp.Q * N[1];
    const long b2 = cp.P[0] == 1 ? 0 : 3L * cp.Q * cp.P[1] * N[0];
    GK_CHECK(hipMalloc(&cp.Kc, kc * sizeof(float)));
    GK_CHECK(hipMalloc(&cp.B1, b1 * sizeof(float2)));
    cp.B2 = nullptr;  // layout [3][Q][Nx][Py]
    if (b2) GK_CHECK(hipMalloc(&cp.B2, b2 * sizeof(float2)));
    cp.bytes = kc * sizeof(float) + (b1 + b2) * sizeof(float2);
}

void conv_plan_free(ConvPlan& cp) {
    if (cp.Kc) GK_CHECK(hipFree(cp.Kc));
    if (cp.B1) GK_CHECK(hipFree(cp.B1));
    if (cp.B2) GK_CHECK(hipFree(cp.B2));
    cp.Kc = nullptr;
    cp.B1 = cp.B2 = nullptr;
}

static int wrapi(int number, int max) {
    while (number < 0) number += max;
    while (number >= max) number -= max;
    return number;
}

// real-space FaceKernel6 (+ Exch6) over the padded grid P, 6 x P doubles on the device
static double* realspace_kernel_f64(const int P[3], const double cell[3], int accuracy, bool exch, hipStream_t s) {
    const int P0 = P[0], P1 = P[1], P2 = P[2];
    const long NP = (long)P0 * P1 * P2;
    double* K = nullptr;
    GK_CHECK(hipMalloc(&K, 6 * NP * sizeof(double)));
    k_facekernel6<<<grid1d(NP), kBlock, 0, s>>>(K, P0, P1, P2, cell[0], cell[1], cell[2], accuracy);
    GK_LAUNCH_CHECK();
    if (exch) {
        // Exch6NgbrKernel written in the reference's order: with a padded size of 1,
        // wrap(+-1,1) = 0 overwrites the centre (exchange.go:58-65)
        std::vector<std::pair<long, double>> pts;
        auto put = [&](int a, int b, int c, double v) {
            long id = ((long)wrapi(a, P0) * P1 + wrapi(b, P1)) * P2 + wrapi(c, P2);
            for (auto& pr : pts)
                if (pr.first == id) {
                    pr.second = v;
                    return;
                }
            pts.push_back({id, v});
        };
        const double hx = cell[0] * cell[0], hy = cell[1] * cell[1], hz = cell[2] * cell[2];
        put(0, 0, 0, -2 / hx - 2 / hy - 2 / hz);
        put(1, 0, 0, 1 / hx);
        put(-1, 0, 0, 1 / hx);
        put(0, 1, 0, 1 / hy);
        put(0, -1, 0, 1 / hy);
        put(0, 0, 1, 1 / hz);
        put(0, 0, -1, 1 / hz);
        ExchPoints ep{};
        ep.n = (int)pts.size();
        for (int i = 0; i < ep.n; i++) {
            ep.idx[i] = pts[i].first;
            ep.val[i] = pts[i].second;
        }
        k_add_exch_points<<<1, 64, 0, s>>>(K, NP, ep);
        GK_LAUNCH_CHECK();
    }
    return K;
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

void conv_realspace_kernel(const int P[3], const double cell[3], int accuracy, bool exch, float* out_host) {
    hipStream_t s = stream();
    const long NP = (long)P[0] * P[1] * P[2];
    double* K = realspace_kernel_f64(P, cell, accuracy, exch, s);
    float* f = nullptr;
    GK_CHECK(hipMalloc(&f, 6 * NP * sizeof(float)));
    k_to_float<<<grid1d(6 * NP), kBlock, 0, s>>>(K, f, 6 * NP);
    GK_LAUNCH_CHECK();
    GK_CHECK(hipMemcpyAsync(out_host, f, 6 * NP * sizeof(float), hipMemcpyDeviceToHost, s));
    sync();
    GK_CHECK(hipFree(K));
    GK_CHECK(hipFree(f));
}

double conv_build_kernel(ConvPlan& cp, const double cell[3], int accuracy, bool exch, float* full_host,
                         float* full_dev) {
    hipStream_t s = stream();
    const int P0 = cp.P[0], P1 = cp.P[1], P2 = cp.P[2];
    const long NP = (long)P0 * P1 * P2;
    double* K = realspace_kernel_f64(cp.P, cell, accuracy, exch, s);
    double2* W = nullptr;
    unsigned long long* mx = nullptr;
    GK_CHECK(hipMalloc(&W, NP * sizeof(double2)));
    GK_CHECK(hipMalloc(&mx, 2 * sizeof(unsigned long long)));
    GK_CHECK(hipMemsetAsync(mx, 0, 2 * sizeof(unsigned long long), s));
    const int hx = P0 / 2 + 1, hy = P1 / 2 + 1;
    const long fulls = (long)P0 * P1 * cp.Q;
    const double norm = 1.0 / (double)NP;
    // components stored: 3-D xx yy zz yz xz xy; 2-D xx yy zz yz
    for (int k = 0; k < cp.nK; k++) {
        k_real_to_complex<<<grid1d(NP), kBlock, 0, s>>>(K + k * NP, W, NP);
        fft3_c2c_f64(W, cp.P, -1, s);
        k_extract_kernel<<<grid1d(fulls), kBlock, 0, s>>>(W, P0, P1, P2, cp.Q, norm, cp.Kc, k, cp.nK,
                                                            full_dev ? full_dev + k * fulls : nullptr, mx, mx + 1);
        GK_LAUNCH_CHECK();
    }
    unsigned long long h[2];
    GK_CHECK(hipMemcpyAsync(h, mx, sizeof(h), hipMemcpyDeviceToHost, s));
    if (full_host && full_dev)
        GK_CHECK(hipMemcpyAsync(full_host, full_dev, cp.nK * fulls * sizeof(float), hipMemcpyDeviceToHost, s));
    sync();
    GK_CHECK(hipFree(K));
    GK_CHECK(hipFree(W));
    GK_CHECK(hipFree(mx));
    double re, im;
    memcpy(&re, &h[0], 8);
    memcpy(&im, &h[1], 8);
    return re > 0 ? im / re : 0.0;
}

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
template <int M, int TR>
__global__ __launch_bounds__(TR * RegFFT<M>::G) void k_zfwd(StageIn si, float2* __restrict__ B1,
                                                              const float2* __restrict__ twM,
                                                              const float2* __restrict__ twP, int Nx, int Ny, int Nz) {
    using F = RegFFT<M>;
    constexpr int E = F::E, G = F::G, NT = TR * G, Q = M + 1, TS = TR + 1;
    constexpr int FFT_LDS = TR * F::LDS, TILE_LDS = Q * TS;
    __shared__ float2 lds[FFT_LDS > TILE_LDS ? FFT_LDS : TILE_LDS];
    if (si.st && !si.st->trial_active) return;
    const int row = threadIdx.x / G, g = threadIdx.x % G;
    const int c = blockIdx.z, x = blockIdx.y, y0 = blockIdx.x * TR;
    const int nrows = min(TR, Ny - y0);
    const bool ok = row < nrows;
    const long N = (long)Nx * Ny * Nz;
    const long rbase = c * N + ((long)x * Ny + y0 + row) * Nz;
    const float dt = si.st ? si.st->dt_used : 0.f;
    float coef[6];
#pragma unroll
    for (int j = 0; j < 6; j++) coef[j] = dt * si.a[j];
    const int nk = si.nk;
    float2 v[E];
    if ((Nz & 1) == 0) {
        // pairs (2n, 2n+1) as float2: rows start at an even float offset
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int n = g + G * e;
            float2 z = make_float2(0.f, 0.f);
            if (!pad_half<M>(e) && ok && 2 * n < Nz) {
                z = *reinterpret_cast<const float2*>(si.m + rbase + 2 * n);
                for (int j = 0; j < nk; j++) {
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
                    for (int j = 0; j < nk; j++) a = fmaf(coef[j], si.k[j][rbase + i], a);
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
    constexpr int FFT_LDS = TR * F::LDS, TILE_LDS = Q * TS;
    __shared__ float2 lds[FFT_LDS > TILE_LDS ? FFT_LDS : TILE_LDS];
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
template <int M, int TR, int NU, bool FIN>
__global__ __launch_bounds__(3 * TR * RegFFT<M>::G) void k_zinv_epi(const float2* __restrict__ B1, EpiArgs ea,
                                                                      const float2* __restrict__ twM,
                                                                      const float2* __restrict__ twP,
                                                                      const EngineState* st, int Nx, int Ny, int Nz) {
    using F = RegFFT<M>;
    constexpr int E = F::E, G = F::G, NR = TR * G, NT = 3 * NR, Q = M + 1, TS = TR + 1;
    constexpr int HS = M / 2 > 0 ? M / 2 : 1;  // float2 per h row (z pairs); Nz <= M
    constexpr int TILE = 3 * Q * TS, FFTB = 3 * TR * F::LDS, HB = 3 * TR * HS;
    constexpr int LN = TILE > FFTB ? (TILE > HB ? TILE : HB) : (FFTB > HB ? FFTB : HB);
    __shared__ float2 lds[LN];
    __shared__ float hext_s[3];
    __shared__ float red[NT];
    if (!st->trial_active) return;
    const int c = threadIdx.x / NR, rem = threadIdx.x - c * NR, row = rem / G, g = rem % G;
    const int lb = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int x = lb / gridDim.x, y0 = (lb % gridDim.x) * TR;
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
    const long N = (long)Nx * Ny * Nz;
    float lmax = 0.f;
    // one cell from the accumulated stage input mi (and final sum hi) and the field h
    auto finish = [&](const float mi[3], const float hi[3], const float h[3], float o[3]) {
        float k[3];
        llg_torque_cell(ea.t, hext, mi[0], mi[1], mi[2], h[0], h[1], h[2], k);
        lmax = fmaxf(lmax, fmaxf(fabsf(k[0]), fmaxf(fabsf(k[1]), fabsf(k[2]))));
        if constexpr (FIN) {
            float f[3];
#pragma unroll
            for (int q = 0; q < 3; q++) f[q] = fmaf(cbs, k[q], hi[q]);
            const float sc = rsqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
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
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int p = threadIdx.x + it * NT;
            if (p < P) {
                const int r = p / nzh, n = p - r * nzh;
                const long gi = rowbase + (long)r * Nz + 2 * n;
#pragma unroll
                for (int q = 0; q < 3; q++) ai[it][q] = af[it][q] = *reinterpret_cast<const float2*>(ea.m + q * N + gi);
#pragma unroll
                for (int j = 0; j < NU; j++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const float2 kv = *reinterpret_cast<const float2*>(ea.k[j] + q * N + gi);
                        ai[it][q].x = fmaf(ca[j], kv.x, ai[it][q].x);
                        ai[it][q].y = fmaf(ca[j], kv.y, ai[it][q].y);
                        if constexpr (FIN) {
                            af[it][q].x = fmaf(cb[j], kv.x, af[it][q].x);
                            af[it][q].y = fmaf(cb[j], kv.y, af[it][q].y);
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
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    mi[q] = ai[it][q].x;
                    hi[q] = af[it][q].x;
                    h[q] = hv[q].x;
                }
                finish(mi, hi, h, o0);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    mi[q] = ai[it][q].y;
                    hi[q] = af[it][q].y;
                    h[q] = hv[q].y;
                }
                finish(mi, hi, h, o1);
#pragma unroll
                for (int q = 0; q < 3; q++)
                    *reinterpret_cast<float2*>(ea.out + q * N + gi) = make_float2(o0[q], o1[q]);
            }
        }
    } else {
        const float* hf = reinterpret_cast<const float*>(lds);
        const int P = nrows * Nz;
        for (int p = threadIdx.x; p < P; p += NT) {
            const int r = p / Nz, z = p - r * Nz;
            const long gi = rowbase + (long)r * Nz + z;
            float mi[3], hi[3], h[3], o[3];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                mi[q] = hi[q] = ea.m[q * N + gi];
                h[q] = hf[(q * TR + r) * 2 * HS + z];
#pragma unroll
                for (int j = 0; j < NU; j++) {
                    const float kv = ea.k[j][q * N + gi];
                    mi[q] = fmaf(ca[j], kv, mi[q]);
                    hi[q] = fmaf(cb[j], kv, hi[q]);
                }
            }
            finish(mi, hi, h, o);
#pragma unroll
            for (int q = 0; q < 3; q++) ea.out[q * N + gi] = o[q];
        }
    }
    if (ea.t.part) {
        red[threadIdx.x] = lmax;
        __syncthreads();
        if constexpr (NT >= 64) {
            if (threadIdx.x < 64) {
                float vm = 0.f;
                for (int i = threadIdx.x; i < NT; i += 64) vm = fmaxf(vm, red[i]);
                for (int off = 32; off > 0; off >>= 1) vm = fmaxf(vm, __shfl_xor(vm, off, 64));
                if (threadIdx.x == 0) ea.t.part[blockIdx.y * gridDim.x + blockIdx.x] = vm;
            }
        } else {
            if (threadIdx.x == 0) {
                float vm = 0.f;
                for (int i = 0; i < NT; i++) vm = fmaxf(vm, red[i]);
                ea.t.part[blockIdx.y * gridDim.x + blockIdx.x] = vm;
            }
        }
    }
}

// 3-D Y passes, software-pipelined.  One line per lane group and iteration:
//   forward  B1[c][x][kz][0..Ny) -> B2[c][kz][x][0..Py)   (zero padded)
//   inverse  B2[c][kz][x][0..Py) -> B1[c][x][kz][0..Ny)   (rows >= Ny dropped)
// Lines are numbered with x fastest (a block touches consecutive B2 lines).
// Each lane group walks `nit` lines strided by the grid and loads line i+1
// before transforming line i, so the memory of one line overlaps the FFT of
// the previous one; twiddles come from LDS, so no global load sits between a
// prefetch and its use.
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
    // B1 line (c, x, kz) and B2 line (c, kz, x) of a line number (x fastest)
    auto b1 = [&](long ln) {
        const int x = (int)(ln % Nx);
        const long r = ln / Nx;
        const int kz = (int)(r % Q), c = (int)(r / Q);
        return (((long)c * Nx + x) * Q + kz) * Ny;
    };
    auto b2 = [&](long ln) {
        const int x = (int)(ln % Nx);
        const long r = ln / Nx;
        const int kz = (int)(r % Q), c = (int)(r / Q);
        return (((long)c * Q + kz) * Nx + x) * PY;
    };
    auto load = [&](float2* v, long ln) {
        const bool ok = ln < nlines;
        if constexpr (FWD) {
            const float2* p = src + b1(ok ? ln : 0);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int i = g + G * e;
                v[e] = (!pad_half<PY>(e) && ok && i < Ny) ? p[i] : make_float2(0.f, 0.f);
            }
        } else {
            const float2* p = src + b2(ok ? ln : 0);
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = ok ? p[g + G * e] : make_float2(0.f, 0.f);
        }
    };
    constexpr bool PIPE = E <= 16;  // registers for a second line (longer lines load in place)
    float2 nxt[PIPE ? E : 1];
    if constexpr (PIPE) load(nxt, line);
    __syncthreads();  // twiddle table
    for (int it = 0; it < nit; it++, line += stride) {
        float2 v[E];
        if constexpr (PIPE) {
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = nxt[e];
            if (it + 1 < nit) load(nxt, line + stride);
        } else {
            load(v, line);
        }
        reg_fft<float, PY, FWD ? -1 : +1>(v, g, lds + li * F::LDS, TWL ? twl : tw);
        if (line < nlines) {
            if constexpr (FWD) {
                float2* p = dst + b2(line);
#pragma unroll
                for (int e = 0; e < E; e++) p[g + G * e] = v[e];
            } else {
                float2* p = dst + b1(line);
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int i = g + G * e;
                    if (!pad_half<PY>(e) && i < Ny) p[i] = v[e];
                }
            }
        }
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
    const int rows = PY / lpr;            // items per kz plane
    const long items = (long)Q * rows;
    const long cstride = (long)Q * Nx * PY;
    const int hx = PX / 2 + 1, hy = PY / 2 + 1;
    // column of item t for this lane group (ky < 0: no column)
    auto column = [&](long t, int& kz, int& ky) {
        kz = (int)(t / rows);
        const int rr = (int)(t % rows);
        const int r = (rr & 1) ? rows - 1 - (rr >> 1) : (rr >> 1);
        ky = (t < items && li < lpr) ? r * lpr + li : -1;
    };
    auto load = [&](float2 (*v)[E], long t) {
        int kz, ky;
        column(t, kz, ky);
        const float2* base = B2 + (long)kz * Nx * PY + (ky < 0 ? 0 : ky);
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int x = g + G * e;
                v[c][e] = (!pad_half<PX>(e) && ky >= 0 && x < Nx) ? base[c * cstride + (long)x * PY]
                                                                   : make_float2(0.f, 0.f);
            }
    };
    const long t0 = (long)blockIdx.x * nit;
    float2 nxt[3][E];
    load(nxt, t0);
    __syncthreads();  // twiddle table
    for (int it = 0; it < nit; it++) {
        const long t = t0 + it;
        float2 v[3][E];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int e = 0; e < E; e++) v[c][e] = nxt[c][e];
        if (it + 1 < nit) load(nxt, t + 1);
        int kz, ky;
        column(t, kz, ky);
#pragma unroll
        for (int c = 0; c < 3; c++) reg_fft<float, PX, -1>(v[c], g, lds + li * F::LDS, twl);
        if (ky >= 0) {
            const bool fy = ky > PY / 2;
            const int kyp = fy ? PY - ky : ky;
            const float* kb = Kc + ((long)kz * hy + kyp) * hx * 6;
#pragma unroll
            for (int e = 0; e < E; e++) {
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
#pragma unroll
        for (int c = 0; c < 3; c++) reg_fft<float, PX, +1>(v[c], g, lds + li * F::LDS, twl);
        if (ky >= 0) {
            float2* base = B2 + (long)kz * Nx * PY + ky;
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int x = g + G * e;
                    if (!pad_half<PX>(e) && x < Nx) base[c * cstride + (long)x * PY] = v[c][e];
                }
        }
    }
}

// 2-D: Y forward . kernel . Y inverse on NKZ consecutive kz rows, all 3 components, in place in B1
template <int PY, int NKZ>
__global__ __launch_bounds__(kNT) void k_y2d(float2* __restrict__ B1, const float* __restrict__ Kc,
                                             const float2* __restrict__ tw, const EngineState* st, int Ny, int Q) {
    constexpr int LS = PY + 1;
    __shared__ float2 lds[3 * NKZ * LS];
    if (st && !st->trial_active) return;
    const int kz0 = blockIdx.x * NKZ;
    const int nkz = min(NKZ, Q - kz0);
    const long cstride = (long)Q * Ny;
    const int per = nkz * Ny;
    batched_copy<float2, kNT>(
        3 * per,
        [&](int t) {
            const int c = t / per, r = t - c * per;
            return B1[c * cstride + (long)kz0 * Ny + r];
        },
        [&](int t, float2 v) {
            const int c = t / per, r = t - c * per;
            const int j = r / Ny, y = r - j * Ny;
            lds[(c * NKZ + j) * LS + y] = v;
        });
    __syncthreads();
    fft_lds<float, PY, -1, 3 * NKZ, kNT>(lds, LS, tw, Ny);
    const int hy = PY / 2 + 1;
    const long cs = (long)Q * hy;
    for (int t = threadIdx.x; t < nkz * PY; t += kNT) {
        const int j = t / PY, ky = t - j * PY;
        const int kz = kz0 + j;
        const bool fy = ky > PY / 2;
        const int kyp = fy ? PY - ky : ky;
        const long b = (long)kz * hy + kyp;
        const float kxx = Kc[b], kyy = Kc[cs + b], kzz = Kc[2 * cs + b];
        const float kyz = fy ? -Kc[3 * cs + b] : Kc[3 * cs + b];
        float2* px = &lds[(0 * NKZ + j) * LS + ky];
        float2* py = &lds[(1 * NKZ + j) * LS + ky];
        float2* pz = &lds[(2 * NKZ + j) * LS + ky];
        const float2 a = *px, bb = *py, cc = *pz;
        *px = make_float2(a.x * kxx, a.y * kxx);
        *py = make_float2(bb.x * kyy + cc.x * kyz, bb.y * kyy + cc.y * kyz);
        *pz = make_float2(bb.x * kyz + cc.x * kzz, bb.y * kyz + cc.y * kzz);
    }
    __syncthreads();
    fft_lds<float, PY, +1, 3 * NKZ, kNT>(lds, LS, tw, PY);
    for (int c = 0; c < 3; c++) {
        float2* dst = B1 + c * cstride + (long)kz0 * Ny;
        for (int t = threadIdx.x; t < nkz * Ny; t += kNT) {
            const int j = t / Ny, y = t - j * Ny;
            dst[t] = lds[(c * NKZ + j) * LS + y];
        }
    }
}

// =============================================================== launchers
#define GK_POW2_1_2048(X) X(1) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)
#define GK_POW2_2_4096(X) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#define GK_POW2_2_1024(X) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)

template <int L> constexpr int groups_per_block(int budget_lines_bytes) {
    // how many 3-line groups of length L fit in `budget` bytes of LDS (>= 1, <= 32, power of two)
    int n = budget_lines_bytes / (3 * (L + 1) * 8);
    n = n < 1 ? 1 : (n > 32 ? 32 : n);
    int p = 1;
    while (p * 2 <= n) p *= 2;
    return p;
}

// rows per Z block: enough consecutive y for 64-128 B transposed segments, at most 1024 threads
template <int M> constexpr int z_rows() {
    constexpr int G = RegFFT<M>::G;
    int tr = 1024 / G;
    if (tr > 16) tr = 16;
    while (tr > 1 && tr * (M + 1) * 8 > 72 * 1024) tr /= 2;
    return tr;
}

static void launch_zfwd(const ConvPlan& cp, const StageIn& si, hipStream_t s) {
    const int M = cp.M, Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2];
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    switch (M) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int TR = z_rows<n>();                                                         \
        dim3 g(div_up(Ny, TR), Nx, 3);                                                          \
        k_zfwd<n, TR><<<g, TR * RegFFT<n>::G, 0, s>>>(si, cp.B1, twM, twP, Nx, Ny, Nz);         \
        break;                                                                                  \
    }
        GK_POW2_1_2048(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

static void launch_zinv(const ConvPlan& cp, float* h, const EngineState* st, hipStream_t s) {
    const int M = cp.M, Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2];
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    switch (M) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int TR = z_rows<n>();                                                         \
        dim3 g(div_up(Ny, TR), Nx, 3);                                                          \
        k_zinv<n, TR><<<g, TR * RegFFT<n>::G, 0, s>>>(cp.B1, h, twM, twP, st, Nx, Ny, Nz);      \
        break;                                                                                  \
    }
        GK_POW2_1_2048(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

// rows per block of the fused Z inverse: 3 x TR lane groups <= 1024 threads, the
// 3 transposed tiles of (M+1) x (TR+1) complex within 96 KB of LDS
template <int M> constexpr int epi_rows() {
    constexpr int G = RegFFT<M>::G;
    int tr = 1024 / (3 * G);
    if (tr > 16) tr = 16;
    int p = 1;
    while (p * 2 <= tr) p *= 2;
    tr = p;
    while (tr > 1 && 3 * (M + 1) * (tr + 1) * 8 > 96 * 1024) tr /= 2;
    return tr;
}

static int epi_rows_rt(int M) {
    switch (M) {
#define GK_CASE(n) \
    case n: return epi_rows<n>();
        GK_POW2_1_2048(GK_CASE)
#undef GK_CASE
    default: abort();
    }
}

int conv_epi_blocks(const ConvPlan& cp) { return div_up(cp.N[1], epi_rows_rt(cp.M)) * cp.N[0]; }

static void launch_zinv_epi(const ConvPlan& cp, const EpiArgs& ea, const EngineState* st, hipStream_t s) {
    const int M = cp.M, Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2];
    GK_ASSERT(st != nullptr && ea.nu >= 0 && ea.nu <= 6 && (!ea.fin || ea.nu <= 3), "fused Z inverse: bad epilogue");
    // pad the list of earlier stages to the kernel's bound with (m, 0, 0): fmaf(0, m, z) == z
    const int nu = ea.nu;
    EpiArgs e = ea;
    for (int j = nu; j < 6; j++) {
        e.k[j] = ea.m;
        e.a[j] = 0.f;
        e.b[j] = 0.f;
    }
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    switch (M) {
#define GK_CASE(n)                                                                               \
    case n: {                                                                                    \
        constexpr int TR = epi_rows<n>();                                                        \
        const dim3 g(div_up(Ny, TR), Nx), b(3 * TR * RegFFT<n>::G);                              \
        if (e.fin) {                                                                             \
            if (nu <= 1) k_zinv_epi<n, TR, 1, true><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else k_zinv_epi<n, TR, 3, true><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);         \
        } else {                                                                                 \
            if (nu == 0) k_zinv_epi<n, TR, 0, false><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else if (nu == 1) k_zinv_epi<n, TR, 1, false><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else if (nu <= 3) k_zinv_epi<n, TR, 3, false><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else k_zinv_epi<n, TR, 6, false><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);            \
        }                                                                                        \
        break;                                                                                   \
    }
        GK_POW2_1_2048(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

// Blocks per CU of the pipelined Y passes (each walks several lines).
#ifndef GK_Y_BLOCKS_PER_CU
#define GK_Y_BLOCKS_PER_CU 8
#endif
constexpr int kYBlocksPerCU = GK_Y_BLOCKS_PER_CU;
// ... and of the pipelined X . K . X pass
#ifndef GK_X_BLOCKS_PER_CU
#define GK_X_BLOCKS_PER_CU 4
#endif
constexpr int kXBlocksPerCU = GK_X_BLOCKS_PER_CU;

static int cu_count() {
    static int n = 0;
    if (!n) {
        int dev = 0;
        GK_CHECK(hipGetDevice(&dev));
        GK_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    }
    return n;
}

// Split `items` units of work (lines per block row) over at most `blocks` blocks
// so that every block runs the same number of iterations: returns the iteration
// count and sets `grid`.
static int pipeline_split(long items, long blocks, int* grid) {
    const long nit = std::max(1L, (items + blocks - 1) / blocks);
    *grid = (int)((items + nit - 1) / nit);
    return (int)nit;
}

static void launch_y3d(const ConvPlan& cp, const EngineState* st, bool fwd, hipStream_t s) {
    const int PY = cp.P[1], Nx = cp.N[0], Ny = cp.N[1], Q = cp.Q;
    const float2* tw = twiddle_f32(PY);
    const long nlines = 3L * Nx * Q;
    switch (PY) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int LPB = kNT / RegFFT<n>::G;                                                 \
        int g = 0;                                                                              \
        const int nit = pipeline_split(div_up(nlines, LPB), (long)kYBlocksPerCU * cu_count(), &g); \
        if (fwd) k_ypass<n, true><<<g, kNT, 0, s>>>(cp.B1, cp.B2, tw, st, Nx, Ny, Q, nit);      \
        else k_ypass<n, false><<<g, kNT, 0, s>>>(cp.B2, cp.B1, tw, st, Nx, Ny, Q, nit);          \
        break;                                                                                  \
    }
        GK_POW2_2_4096(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

static void launch_xmul(const ConvPlan& cp, const EngineState* st, hipStream_t s) {
    const int PX = cp.P[0], PY = cp.P[1], Nx = cp.N[0], Q = cp.Q;
    const float2* tw = twiddle_f32(PX);
    switch (PX) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int LPB = kNT / RegFFT<n>::G;                                                 \
        const int lpr = LPB < PY ? LPB : PY;                                                    \
        int g = 0;                                                                              \
        const int nit = pipeline_split((long)Q * (PY / lpr), (long)kXBlocksPerCU * cu_count(), &g); \
        k_xmul<n><<<g, kNT, 0, s>>>(cp.B2, cp.Kc, tw, st, Nx, PY, Q, nit);                      \
        break;                                                                                  \
    }
        GK_POW2_2_1024(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

static void launch_y2d(const ConvPlan& cp, const EngineState* st, hipStream_t s) {
    const int PY = cp.P[1], Ny = cp.N[1], Q = cp.Q;
    const float2* tw = twiddle_f32(PY);
    switch (PY) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int NKZ = groups_per_block<n>(49152);                                         \
        k_y2d<n, NKZ><<<div_up(Q, NKZ), kNT, 0, s>>>(cp.B1, cp.Kc, tw, st, Ny, Q);              \
        break;                                                                                  \
    }
        GK_POW2_2_4096(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

void Prof::begin(int cls, hipStream_t s) {
    if (!on) return;
    while ((int)pool.size() < used + 2) {
        hipEvent_t e;
        GK_CHECK(hipEventCreate(&e));
        pool.push_back(e);
    }
    marks.push_back({cls, used});
    GK_CHECK(hipEventRecord(pool[used], s));
    used += 2;
}
void Prof::end(hipStream_t s) {
    if (!on) return;
    GK_CHECK(hipEventRecord(pool[marks.back().second + 1], s));
}
void Prof::collect(double ms[kClasses], int counts[kClasses]) {
    for (int c = 0; c < kClasses; c++) {
        ms[c] = 0;
        counts[c] = 0;
    }
    for (auto& mk : marks) {
        GK_CHECK(hipEventSynchronize(pool[mk.second + 1]));
        float t = 0.f;
        GK_CHECK(hipEventElapsedTime(&t, pool[mk.second], pool[mk.second + 1]));
        ms[mk.first] += t;
        counts[mk.first] += 1;
    }
}
Prof::~Prof() {
    for (auto e : pool) (void)hipEventDestroy(e);
}

void conv_run_pass(const ConvPlan& cp, int cls, const StageIn& in, float* h, const EpiArgs* epi, hipStream_t s) {
    const bool d3 = cp.P[0] > 1;
    switch (cls) {
    case 0: launch_zfwd(cp, in, s); break;
    case 1: GK_ASSERT(d3, "no Y forward pass in 2-D"); launch_y3d(cp, in.st, true, s); break;
    case 2: if (d3) launch_xmul(cp, in.st, s); else launch_y2d(cp, in.st, s); break;
    case 3: GK_ASSERT(d3, "no Y inverse pass in 2-D"); launch_y3d(cp, in.st, false, s); break;
    case 4: if (epi) launch_zinv_epi(cp, *epi, in.st, s); else launch_zinv(cp, h, in.st, s); break;
    default: GK_ASSERT(false, "pass class %d", cls);
    }
}

void conv_apply(const ConvPlan& cp, const StageIn& in, float* h, hipStream_t s, Prof* prof, const EpiArgs* epi) {
    Prof dummy;
    Prof& p = prof ? *prof : dummy;
    p.begin(0, s);
    launch_zfwd(cp, in, s);
    p.end(s);
    if (cp.P[0] == 1) {
        p.begin(2, s);
        launch_y2d(cp, in.st, s);
        p.end(s);
    } else {
        p.begin(1, s);
        launch_y3d(cp, in.st, true, s);
        p.end(s);
        p.begin(2, s);
        launch_xmul(cp, in.st, s);
        p.end(s);
        p.begin(3, s);
        launch_y3d(cp, in.st, false, s);
        p.end(s);
    }
    p.begin(4, s);
    if (epi)
        launch_zinv_epi(cp, *epi, in.st, s);
    else
        launch_zinv(cp, h, in.st, s);
    p.end(s);
}

}  // namespace gk
