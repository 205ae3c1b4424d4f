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
//   3-D: B1 --Y fwd--> B2[c][kz][ky][x] --X fwd . K . X inv (in place)--> B2
//        --Y inv--> B1 --Z c2r--> h[c][x][y][z]
//   2-D: B1 --Y fwd . K . Y inv (in place)--> B1 --Z c2r--> h
// Only data lines that can be non-zero are transformed (rows y >= Ny and
// planes x >= Nx are zero padding), and only the kept half of each inverse
// transform is stored, so B1/B2 hold Nx (not Px) planes.
//
// The kernel spectrum is real and each component is even or odd along every
// axis (Kxy odd in x and y, Kxz odd in x, Kyz odd in y, diagonal even), so
// only kx <= Px/2, ky <= Py/2 is stored (Kc), a quarter of the reference's
// 6 x Px x Py x (Pz/2+1) floats; signs are restored on the fly.
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
                                 float* __restrict__ Kc_comp, float* __restrict__ full_comp,
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
            if (P0 == 1) Kc_comp[(long)kz * hy + ky] = (float)re;
            else Kc_comp[((long)kz * hy + ky) * hx + kx] = (float)re;
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
    cp.B2 = nullptr;
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
    const long kcs = P0 == 1 ? (long)cp.Q * hy : (long)cp.Q * hy * hx;
    const long fulls = (long)P0 * P1 * cp.Q;
    const double norm = 1.0 / (double)NP;
    // components stored: 3-D xx yy zz yz xz xy; 2-D xx yy zz yz
    for (int k = 0; k < cp.nK; k++) {
        k_real_to_complex<<<grid1d(NP), kBlock, 0, s>>>(K + k * NP, W, NP);
        fft3_c2c_f64(W, cp.P, -1, s);
        k_extract_kernel<<<grid1d(fulls), kBlock, 0, s>>>(W, P0, P1, P2, cp.Q, norm, cp.Kc + k * kcs,
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

template <int L> constexpr int rows_per_block() {
    int n = 71680 / ((L + 1) * 8);
    n = n < 1 ? 1 : (n > 16 ? 16 : n);
    int p = 1;
    while (p * 2 <= n) p *= 2;
    return p;
}

// Z forward: rows of m (stage build on the fly) -> B1[c][x][kz][y]
template <int M, int TR>
__global__ __launch_bounds__(kNT) void k_zfwd(StageIn si, float2* __restrict__ B1, const float2* __restrict__ twM,
                                              const float2* __restrict__ twP, int Nx, int Ny, int Nz) {
    constexpr int LS = M + 1, Q = M + 1;
    __shared__ float2 lds[TR * LS];
    if (si.st && !si.st->trial_active) return;
    const int c = blockIdx.z, x = blockIdx.y, y0 = blockIdx.x * TR;
    const int nrows = min(TR, Ny - y0);
    const long N = (long)Nx * Ny * Nz;
    const long base = c * N + ((long)x * Ny + y0) * Nz;
    const float dt = si.st ? si.st->dt_used : 0.f;
    float coef[6];
#pragma unroll
    for (int j = 0; j < 6; j++) coef[j] = dt * si.a[j];
    const int cnt = nrows * Nz;
    const int nk = si.nk;
    float* mout = si.mout;
    batched_copy<float, kNT>(
        cnt,
        [&](int e) {
            float v = si.m[base + e];
            for (int j = 0; j < nk; j++) v = fmaf(coef[j], si.k[j][base + e], v);
            return v;
        },
        [&](int e, float v) {
            if (mout) mout[base + e] = v;
            const int row = e / Nz, col = e - row * Nz;
            reinterpret_cast<float*>(lds + row * LS)[col] = v;
        });
    if (Nz & 1)
        for (int row = threadIdx.x; row < nrows; row += kNT) reinterpret_cast<float*>(lds + row * LS)[Nz] = 0.f;
    __syncthreads();
    fft_lds<float, M, -1, TR, kNT>(lds, LS, twM, (Nz + 1) / 2);
    float2* dst = B1 + ((long)c * Nx + x) * Q * Ny + y0;
    for (int t = threadIdx.x; t < Q * TR; t += kNT) {
        const int row = t % TR, kz = t / TR;
        if (row < nrows) {
            float2 zk = lds[row * LS + (kz % M)], zmk = lds[row * LS + ((M - kz) % M)];
            dst[(long)kz * Ny + row] = r2c_split<float>(zk, zmk, twP[kz]);
        }
    }
}

// Z inverse: B1[c][x][kz][y] -> rows of h
template <int M, int TR>
__global__ __launch_bounds__(kNT) void k_zinv(const float2* __restrict__ B1, float* __restrict__ h,
                                              const float2* __restrict__ twM, const float2* __restrict__ twP,
                                              const EngineState* st, int Nx, int Ny, int Nz) {
    constexpr int LS = M + 1, Q = M + 1;
    constexpr int NY = (TR * M + kNT - 1) / kNT;
    __shared__ float2 lds[TR * LS];
    if (st && !st->trial_active) return;
    const int c = blockIdx.z, x = blockIdx.y, y0 = blockIdx.x * TR;
    const int nrows = min(TR, Ny - y0);
    const float2* src = B1 + ((long)c * Nx + x) * Q * Ny + y0;
    batched_copy<float2, kNT>(
        Q * TR,
        [&](int t) {
            const int row = t % TR, kz = t / TR;
            return row < nrows ? src[(long)kz * Ny + row] : make_float2(0.f, 0.f);
        },
        [&](int t, float2 v) { lds[(t % TR) * LS + t / TR] = v; });
    __syncthreads();
    float2 y[NY];
#pragma unroll
    for (int q = 0; q < NY; q++) {
        const int t = threadIdx.x + q * kNT;
        if (t < TR * M) {
            const int row = t / M, k = t % M;
            y[q] = c2r_merge<float>(lds[row * LS + k], lds[row * LS + (M - k)], twP[k]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NY; q++) {
        const int t = threadIdx.x + q * kNT;
        if (t < TR * M) lds[(t / M) * LS + (t % M)] = y[q];
    }
    __syncthreads();
    fft_lds<float, M, +1, TR, kNT>(lds, LS, twM, M);
    const long N = (long)Nx * Ny * Nz;
    float* dst = h + c * N + ((long)x * Ny + y0) * Nz;
    const int cnt = nrows * Nz;
    for (int e = threadIdx.x; e < cnt; e += kNT) {
        const int row = e / Nz, col = e - row * Nz;
        dst[e] = reinterpret_cast<const float*>(lds + row * LS)[col];
    }
}

// 3-D Y forward: B1[c][x][kz][0..Ny) -> B2[c][kz][ky][x]
template <int PY, int XC>
__global__ __launch_bounds__(kNT) void k_yfwd(const float2* __restrict__ B1, float2* __restrict__ B2,
                                              const float2* __restrict__ tw, const EngineState* st, int Nx, int Ny,
                                              int Q) {
    constexpr int LS = PY + 1;
    __shared__ float2 lds[XC * LS];
    if (st && !st->trial_active) return;
    const int c = blockIdx.z, x0 = blockIdx.y * XC, kz = blockIdx.x;
    const int nx = min(XC, Nx - x0);
    batched_copy<float2, kNT>(
        nx * Ny,
        [&](int t) {
            const int xi = t / Ny, y = t - xi * Ny;
            return B1[(((long)c * Nx + x0 + xi) * Q + kz) * Ny + y];
        },
        [&](int t, float2 v) {
            const int xi = t / Ny, y = t - xi * Ny;
            lds[xi * LS + y] = v;
        });
    __syncthreads();
    fft_lds<float, PY, -1, XC, kNT>(lds, LS, tw, Ny);
    float2* dst = B2 + (((long)c * Q + kz) * PY) * Nx + x0;
    for (int t = threadIdx.x; t < PY * nx; t += kNT) {
        const int xi = t % nx, ky = t / nx;
        dst[(long)ky * Nx + xi] = lds[xi * LS + ky];
    }
}

// 3-D Y inverse: B2[c][kz][ky][x] -> B1[c][x][kz][0..Ny)
template <int PY, int XC>
__global__ __launch_bounds__(kNT) void k_yinv(const float2* __restrict__ B2, float2* __restrict__ B1,
                                              const float2* __restrict__ tw, const EngineState* st, int Nx, int Ny,
                                              int Q) {
    constexpr int LS = PY + 1;
    __shared__ float2 lds[XC * LS];
    if (st && !st->trial_active) return;
    const int c = blockIdx.z, x0 = blockIdx.y * XC, kz = blockIdx.x;
    const int nx = min(XC, Nx - x0);
    const float2* src = B2 + (((long)c * Q + kz) * PY) * Nx + x0;
    batched_copy<float2, kNT>(
        PY * nx, [&](int t) { return src[(long)(t / nx) * Nx + (t % nx)]; },
        [&](int t, float2 v) { lds[(t % nx) * LS + t / nx] = v; });
    __syncthreads();
    fft_lds<float, PY, +1, XC, kNT>(lds, LS, tw, PY);
    for (int t = threadIdx.x; t < nx * Ny; t += kNT) {
        const int xi = t / Ny, y = t - xi * Ny;
        B1[(((long)c * Nx + x0 + xi) * Q + kz) * Ny + y] = lds[xi * LS + y];
    }
}

// 3-D X forward . kernel . X inverse on NP consecutive (kz,ky) pairs, all 3 components
template <int PX, int NP>
__global__ __launch_bounds__(kNT) void k_xmul(float2* __restrict__ B2, const float* __restrict__ Kc,
                                              const float2* __restrict__ tw, const EngineState* st, int Nx, int PY,
                                              int Q) {
    constexpr int LS = PX + 1;
    __shared__ float2 lds[3 * NP * LS];
    if (st && !st->trial_active) return;
    const long npairs = (long)Q * PY;
    const long p0 = (long)blockIdx.x * NP;
    const int np = (int)min((long)NP, npairs - p0);
    const long cstride = npairs * Nx;
    const int per = np * Nx;
    batched_copy<float2, kNT>(
        3 * per,
        [&](int t) {
            const int c = t / per, r = t - c * per;
            return B2[c * cstride + p0 * Nx + r];
        },
        [&](int t, float2 v) {
            const int c = t / per, r = t - c * per;
            const int pi = r / Nx, x = r - pi * Nx;
            lds[(c * NP + pi) * LS + x] = v;
        });
    __syncthreads();
    fft_lds<float, PX, -1, 3 * NP, kNT>(lds, LS, tw, Nx);
    const int hx = PX / 2 + 1, hy = PY / 2 + 1;
    const long cs = (long)Q * hy * hx;
    for (int t = threadIdx.x; t < np * PX; t += kNT) {
        const int pi = t / PX, kx = t - pi * PX;
        const long p = p0 + pi;
        const int kz = (int)(p / PY), ky = (int)(p - (long)kz * PY);
        const bool fx = kx > PX / 2, fy = ky > PY / 2;
        const int kxp = fx ? PX - kx : kx, kyp = fy ? PY - ky : ky;
        const long b = ((long)kz * hy + kyp) * hx + kxp;
        const float kxx = Kc[b], kyy = Kc[cs + b], kzz = Kc[2 * cs + b];
        float kyz = Kc[3 * cs + b], kxz = Kc[4 * cs + b], kxy = Kc[5 * cs + b];
        if (fy) kyz = -kyz;
        if (fx) kxz = -kxz;
        if (fx != fy) kxy = -kxy;
        float2* px = &lds[(0 * NP + pi) * LS + kx];
        float2* py = &lds[(1 * NP + pi) * LS + kx];
        float2* pz = &lds[(2 * NP + pi) * LS + kx];
        const float2 a = *px, bb = *py, cc = *pz;
        *px = make_float2(a.x * kxx + bb.x * kxy + cc.x * kxz, a.y * kxx + bb.y * kxy + cc.y * kxz);
        *py = make_float2(a.x * kxy + bb.x * kyy + cc.x * kyz, a.y * kxy + bb.y * kyy + cc.y * kyz);
        *pz = make_float2(a.x * kxz + bb.x * kyz + cc.x * kzz, a.y * kxz + bb.y * kyz + cc.y * kzz);
    }
    __syncthreads();
    fft_lds<float, PX, +1, 3 * NP, kNT>(lds, LS, tw, PX);
    for (int c = 0; c < 3; c++) {
        float2* dst = B2 + c * cstride + p0 * Nx;
        for (int t = threadIdx.x; t < np * Nx; t += kNT) {
            const int pi = t / Nx, x = t - pi * Nx;
            dst[t] = lds[(c * NP + pi) * LS + x];
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

static void launch_zfwd(const ConvPlan& cp, const StageIn& si, hipStream_t s) {
    const int M = cp.M, Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2];
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    switch (M) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int TR = rows_per_block<n>();                                                 \
        dim3 g(div_up(Ny, TR), Nx, 3);                                                          \
        k_zfwd<n, TR><<<g, kNT, 0, s>>>(si, cp.B1, twM, twP, Nx, Ny, Nz);                       \
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
        constexpr int TR = rows_per_block<n>();                                                 \
        dim3 g(div_up(Ny, TR), Nx, 3);                                                          \
        k_zinv<n, TR><<<g, kNT, 0, s>>>(cp.B1, h, twM, twP, st, Nx, Ny, Nz);                    \
        break;                                                                                  \
    }
        GK_POW2_1_2048(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

static void launch_y3d(const ConvPlan& cp, const EngineState* st, bool fwd, hipStream_t s) {
    const int PY = cp.P[1], Nx = cp.N[0], Ny = cp.N[1], Q = cp.Q;
    const float2* tw = twiddle_f32(PY);
    switch (PY) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int XC = rows_per_block<n>();                                                 \
        dim3 g(Q, div_up(Nx, XC), 3);                                                           \
        if (fwd) k_yfwd<n, XC><<<g, kNT, 0, s>>>(cp.B1, cp.B2, tw, st, Nx, Ny, Q);              \
        else k_yinv<n, XC><<<g, kNT, 0, s>>>(cp.B2, cp.B1, tw, st, Nx, Ny, Q);                  \
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
    const long npairs = (long)Q * PY;
    switch (PX) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int NP = groups_per_block<n>(49152);                                          \
        k_xmul<n, NP><<<div_up(npairs, NP), kNT, 0, s>>>(cp.B2, cp.Kc, tw, st, Nx, PY, Q);      \
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

void conv_apply(const ConvPlan& cp, const StageIn& in, float* h, hipStream_t s, Prof* prof) {
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
    launch_zinv(cp, h, in.st, s);
    p.end(s);
}

}  // namespace gk
