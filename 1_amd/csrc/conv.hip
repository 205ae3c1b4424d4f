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
#include "passes.h"

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
// Compressed layouts: 2-D [comp][kz][ky]; 3-D [kz][ky][kx][comp] (the 6 values
// of one frequency together, for the X . K . X pass); thin 3-D plane-pass
// plans [kz][comp][kx][ky] (consecutive ky together, one thread per ky).
__global__ void k_extract_kernel(const double2* __restrict__ W, int P0, int P1, int P2, int Q, double norm,
                                 float* __restrict__ Kc, int comp, int nK, int plane_layout,
                                 float* __restrict__ full_comp, unsigned long long* maxre,
                                 unsigned long long* maxim) {
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
            else if (plane_layout) Kc[(((long)kz * nK + comp) * hx + kx) * hy + ky] = (float)re;
            else Kc[(((long)kz * hy + ky) * hx + kx) * nK + comp] = (float)re;
        }
    }
    atomic_max_nonneg(maxre, lre);
    atomic_max_nonneg(maxim, lim);
}

// =============================================================== plan
static int pad_len(int n) {
    // smallest register-FFT length 2^a or 3 * 2^a that is >= 2n (any P >= 2n - 1
    // gives the same linear convolution; the reference pads to exactly 2n)
    if (n == 1) return 1;
    int p = 2;
    while (p < 2 * n) p *= 2;
    return (p / 4) * 3 >= 2 * n ? (p / 4) * 3 : p;
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
    cp.yxy = cp.P[0] > 1 && yxy_fits(cp.P[0], cp.P[1]);
    GK_ASSERT(cp.M <= 2048 && cp.P[1] <= 4096 && cp.P[0] <= 1024, "grid too large for this build (padded %d %d %d)",
              cp.P[0], cp.P[1], cp.P[2]);
    const long hx = cp.P[0] / 2 + 1, hy = cp.P[1] / 2 + 1;
    const long kc = cp.P[0] == 1 ? (long)cp.nK * cp.Q * hy : (long)cp.nK * cp.Q * hy * hx;
    const long b1 = 3L * N[0] * cp.Q * N[1];
    const long b2 = (cp.P[0] == 1 || cp.yxy) ? 0 : 3L * cp.Q * cp.P[1] * N[0];
    GK_CHECK(hipMalloc(&cp.Kc, kc * sizeof(float)));
    GK_CHECK(hipMalloc(&cp.B1, b1 * sizeof(float2)));
    cp.B2 = nullptr;  // layout [3][Q][Nx][Py]
    if (b2) GK_CHECK(hipMalloc(&cp.B2, b2 * sizeof(float2)));
    cp.bytes = kc * sizeof(float) + (b1 + b2) * sizeof(float2);
}

static int epi_rows_rt(int M);

size_t conv_plan_footprint(const int N[3], int* epi_blocks) {
    int P[3];
    for (int i = 0; i < 3; i++) P[i] = pad_len(N[i]);
    if (P[2] < 2) P[2] = 2;
    const int M = P[2] / 2, Q = M + 1, nK = P[0] == 1 ? 4 : 6;
    const long hx = P[0] / 2 + 1, hy = P[1] / 2 + 1;
    const long kc = P[0] == 1 ? (long)nK * Q * hy : (long)nK * Q * hy * hx;
    const long b1 = 3L * N[0] * Q * N[1];
    const long b2 = (P[0] == 1 || yxy_fits(P[0], P[1])) ? 0 : 3L * Q * P[1] * N[0];
    if (epi_blocks) *epi_blocks = div_up(N[1], epi_rows_rt(M)) * N[0];
    return kc * sizeof(float) + (b1 + b2) * sizeof(float2);
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
        k_extract_kernel<<<grid1d(fulls), kBlock, 0, s>>>(W, P0, P1, P2, cp.Q, norm, cp.Kc, k, cp.nK, cp.yxy,
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

// =============================================================== launchers


// rows per Z block: enough consecutive y for 64-128 B transposed segments, at most 1024 threads
template <int M> constexpr int z_rows() {
    constexpr int G = RegFFT<M>::G;
    int tr = 1024 / G;
    if (tr > 16) tr = 16;
    while (tr > 1 && tr * (M + 1) * 8 > 72 * 1024) tr /= 2;
    return tr;
}

static void launch_zfwd(const ConvPlan& cp, const StageIn& in, hipStream_t s) {
    const int M = cp.M, Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2];
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    GK_ASSERT(in.nk >= 0 && in.nk <= 5, "stage build with %d earlier stages", in.nk);
    // pad the stage list to the kernel's count with (m, 0): fmaf(0, m, z) == z
    StageIn si = in;
    const int NKp = zfwd_nk(in.nk);
    for (int j = in.nk; j < 6; j++) {
        si.k[j] = in.m;
        si.a[j] = 0.f;
    }
    si.nk = NKp;
    switch (M) {
#define GK_ZF(n, K) k_zfwd<n, TR, K><<<g, TR * RegFFT<n>::G, 0, s>>>(si, cp.B1, twM, twP, Nx, Ny, Nz)
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int TR = z_rows<n>();                                                         \
        dim3 g(div_up(Ny, TR), Nx, 3);                                                          \
        switch (NKp) {                                                                          \
        case 0: GK_ZF(n, 0); break;                                                             \
        case 1: GK_ZF(n, 1); break;                                                             \
        case 2: GK_ZF(n, 2); break;                                                             \
        case 3: GK_ZF(n, 3); break;                                                             \
        default: GK_ZF(n, 5); break;                                                            \
        }                                                                                       \
        break;                                                                                  \
    }
        GK_LEN_M(GK_CASE)
#undef GK_CASE
#undef GK_ZF
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
        GK_LEN_M(GK_CASE)
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
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    default: abort();
    }
}

int conv_epi_blocks(const ConvPlan& cp) { return div_up(cp.N[1], epi_rows_rt(cp.M)) * cp.N[0]; }

static void launch_zinv_epi(const ConvPlan& cp, const EpiArgs& ea, const EngineState* st, hipStream_t s) {
    const int M = cp.M, Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2];
    GK_ASSERT(st != nullptr && ea.nu >= 0 && ea.nu <= 6 && (ea.fin != 1 || ea.nu <= 3) &&
                  (ea.fin != 2 || (ea.mhigh && ea.part_err)),
              "fused Z inverse: bad epilogue");
    // pad the list of earlier stages to the kernel's bound with (m, 0, 0): fmaf(0, m, z) == z
    const int nu = ea.nu;
    EpiArgs e = ea;
    for (int j = nu; j < 6; j++) {
        e.k[j] = ea.m;
        e.a[j] = 0.f;
        e.b[j] = 0.f;
        e.ce[j] = 0.0;
    }
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    switch (M) {
#define GK_CASE(n)                                                                               \
    case n: {                                                                                    \
        constexpr int TR = epi_rows<n>();                                                        \
        const dim3 g(div_up(Ny, TR), Nx), b(3 * TR * RegFFT<n>::G);                              \
        if (e.fin == 1) {                                                                        \
            if (nu <= 1) k_zinv_epi<n, TR, 1, 1><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else k_zinv_epi<n, TR, 3, 1><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);         \
        } else if (e.fin == 2) {                                                                 \
            if (nu <= 3) k_zinv_epi<n, TR, 3, 2><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else k_zinv_epi<n, TR, 6, 2><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);         \
        } else {                                                                                 \
            if (nu == 0) k_zinv_epi<n, TR, 0, 0><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else if (nu == 1) k_zinv_epi<n, TR, 1, 0><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else if (nu <= 3) k_zinv_epi<n, TR, 3, 0><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz); \
            else k_zinv_epi<n, TR, 6, 0><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);            \
        }                                                                                        \
        break;                                                                                   \
    }
        GK_LEN_M(GK_CASE)
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
        GK_LEN_Y(GK_CASE)
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
        const int nit = pipeline_split((long)Q * div_up(PY, lpr), (long)kXBlocksPerCU * cu_count(), &g); \
        k_xmul<n><<<g, kNT, 0, s>>>(cp.B2, cp.Kc, tw, st, Nx, PY, Q, nit);                      \
        break;                                                                                  \
    }
        GK_LEN_X(GK_CASE)
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
        constexpr int LPB = kNT / RegFFT<n>::G;                                                 \
        int g = 0;                                                                              \
        const int nit = pipeline_split(div_up(Q, LPB), (long)kYBlocksPerCU * cu_count(), &g);  \
        k_y2r<n><<<g, kNT, 0, s>>>(cp.B1, cp.Kc, tw, st, Ny, Q, nit);                           \
        break;                                                                                  \
    }
        GK_LEN_Y(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

static void launch_yxy(const ConvPlan& cp, const EngineState* st, hipStream_t s) {
    const int PX = cp.P[0], PY = cp.P[1], Nx = cp.N[0], Ny = cp.N[1], Q = cp.Q;
    const float2* twY = twiddle_f32(PY);
    const float2* twX = twiddle_f32(PX);
    switch (PY) {
#define GK_YX(n, px)                                                                                       \
    case px:                                                                                               \
        if constexpr (yxy_fits(px, n))                                                                     \
            k_yxy<n, px><<<Q, kYxyThreads, 0, s>>>(cp.B1, cp.Kc, twY, twX, st, Nx, Ny, Q);                 \
        else                                                                                               \
            abort();                                                                                       \
        break;
#define GK_CASE(n)                                                                                         \
    case n:                                                                                                \
        switch (PX) {                                                                                      \
            GK_YX(n, 4) GK_YX(n, 6) GK_YX(n, 8) GK_YX(n, 12) GK_YX(n, 16)                                  \
        default: abort();                                                                                  \
        }                                                                                                  \
        break;
        GK_LEN_Y(GK_CASE)
#undef GK_CASE
#undef GK_YX
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
    case 1: GK_ASSERT(d3 && !cp.yxy, "no separate Y forward pass"); launch_y3d(cp, in.st, true, s); break;
    case 2:
        if (cp.yxy) launch_yxy(cp, in.st, s);
        else if (d3) launch_xmul(cp, in.st, s);
        else launch_y2d(cp, in.st, s);
        break;
    case 3: GK_ASSERT(d3 && !cp.yxy, "no separate Y inverse pass"); launch_y3d(cp, in.st, false, s); break;
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
    } else if (cp.yxy) {
        p.begin(2, s);
        launch_yxy(cp, in.st, s);
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
