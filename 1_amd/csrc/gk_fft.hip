// Padded real FFT plan (drop-in for gpu_fft0.cu:30-112) and the generic
// line-transform kernels it is built from.
//
// The reference runs cuFFT over the whole zero-padded volume after zeroing it
// (gpu_fft0.cu:96-101).  Here the transform is split per axis and only lines
// that can be non-zero are transformed:
//   forward: Z real->complex on the Nx*Ny input rows (padding done while
//            loading), Y on the Nx planes (rows >= Ny are known zeros),
//            X on every (ky,kz) column (planes >= Nx are known zeros);
//   inverse: X on every column, Y on the Nx planes only, Z complex->real on
//            the Nx*Ny output rows only (the unpad is the store).
// No scratch zeroing and no separate pad/unpad launches are needed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "../../include/gpukern.h"
#include "fft_device.h"
#include "gk_fft.h"

namespace gk {

// ----------------------------------------------------------------------------
// generic line transform: line l starts at (l / n_inner)*outer + (l % n_inner)*inner,
// elements at stride `estride`.  A block handles NL consecutive lines.  Only the
// first `nvalid` inputs are read (the rest are zero padding) and only the first
// `nkeep` outputs are stored.
template <typename T, int L, int DIR, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_lines(C2<T>* __restrict__ data, const C2<T>* __restrict__ tw,
                                                  long nlines, long n_inner, long inner, long outer, long estride,
                                                  int nvalid, int nkeep) {
    constexpr int LS = L + 1;
    __shared__ C2<T> lds[NL * LS];
    const long l0 = (long)blockIdx.x * NL;
    const bool col = estride != 1;
    auto pos = [&](int t, int n, int& i, int& e) {
        if (col) {
            i = t % NL;
            e = t / NL;
        } else {
            i = t / n;
            e = t % n;
        }
    };
    batched_copy<C2<T>, NT>(
        NL * nvalid,
        [&](int t) {
            int i, e;
            pos(t, nvalid, i, e);
            const long l = l0 + i;
            return l < nlines ? data[(l / n_inner) * outer + (l % n_inner) * inner + e * estride] : mk<T>(0, 0);
        },
        [&](int t, C2<T> v) {
            int i, e;
            pos(t, nvalid, i, e);
            lds[i * LS + e] = v;
        });
    __syncthreads();
    fft_lds<T, L, DIR, NL, NT>(lds, LS, tw, nvalid);
    for (int t = threadIdx.x; t < NL * nkeep; t += NT) {
        int i, e;
        if (col) {
            i = t % NL;
            e = t / NL;
        } else {
            i = t / nkeep;
            e = t % nkeep;
        }
        const long l = l0 + i;
        if (l < nlines) data[(l / n_inner) * outer + (l % n_inner) * inner + e * estride] = lds[i * LS + e];
    }
}

// rows of n_real floats -> M+1 complex (real transform of length P = 2M, zero padded)
template <int M, int NL, int NT>
__global__ __launch_bounds__(NT) void k_rows_r2c(const float* __restrict__ in, float2* __restrict__ out,
                                                 const float2* __restrict__ twM, const float2* __restrict__ twP,
                                                 long nrows, long n_inner, long in_inner, long in_outer,
                                                 long out_inner, long out_outer, int n_real) {
    constexpr int LS = M + 1;
    __shared__ float2 lds[NL * LS];
    const long r0 = (long)blockIdx.x * NL;
    const int npair = (n_real + 1) / 2;
    batched_copy<float2, NT>(
        NL * npair,
        [&](int t) {
            const int i = t / npair, e = t % npair;
            const long r = r0 + i;
            float2 v = make_float2(0.f, 0.f);
            if (r < nrows) {
                const float* src = in + (r / n_inner) * in_outer + (r % n_inner) * in_inner;
                v.x = src[2 * e];
                v.y = (2 * e + 1 < n_real) ? src[2 * e + 1] : 0.f;
            }
            return v;
        },
        [&](int t, float2 v) { lds[(t / npair) * LS + t % npair] = v; });
    __syncthreads();
    fft_lds<float, M, -1, NL, NT>(lds, LS, twM, npair);
    for (int t = threadIdx.x; t < NL * (M + 1); t += NT) {
        const int i = t / (M + 1), k = t % (M + 1);
        const long r = r0 + i;
        if (r < nrows) {
            float2 zk = lds[i * LS + (k % M)], zmk = lds[i * LS + ((M - k) % M)];
            out[(r / n_inner) * out_outer + (r % n_inner) * out_inner + k] = r2c_split<float>(zk, zmk, twP[k]);
        }
    }
}

// M+1 complex per row -> first n_real floats of the length-2M real inverse transform
template <int M, int NL, int NT>
__global__ __launch_bounds__(NT) void k_rows_c2r(const float2* __restrict__ in, float* __restrict__ out,
                                                 const float2* __restrict__ twM, const float2* __restrict__ twP,
                                                 long nrows, long n_inner, long in_inner, long in_outer,
                                                 long out_inner, long out_outer, int n_real) {
    constexpr int LS = M + 1;
    constexpr int NY = (NL * M + NT - 1) / NT;
    __shared__ float2 lds[NL * LS];
    const long r0 = (long)blockIdx.x * NL;
    batched_copy<float2, NT>(
        NL * (M + 1),
        [&](int t) {
            const int i = t / (M + 1), k = t % (M + 1);
            const long r = r0 + i;
            return r < nrows ? in[(r / n_inner) * in_outer + (r % n_inner) * in_inner + k] : make_float2(0.f, 0.f);
        },
        [&](int t, float2 v) { lds[(t / (M + 1)) * LS + t % (M + 1)] = v; });
    __syncthreads();
    float2 y[NY];
#pragma unroll
    for (int q = 0; q < NY; q++) {
        const int t = threadIdx.x + q * NT;
        if (t < NL * M) {
            const int i = t / M, k = t % M;
            y[q] = c2r_merge<float>(lds[i * LS + k], lds[i * LS + (M - k)], twP[k]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NY; q++) {
        const int t = threadIdx.x + q * NT;
        if (t < NL * M) lds[(t / M) * LS + (t % M)] = y[q];
    }
    __syncthreads();
    fft_lds<float, M, +1, NL, NT>(lds, LS, twM, M);
    for (int t = threadIdx.x; t < NL * n_real; t += NT) {
        const int i = t / n_real, e = t % n_real;
        const long r = r0 + i;
        if (r < nrows) {
            float2 z = lds[i * LS + (e >> 1)];
            out[(r / n_inner) * out_outer + (r % n_inner) * out_inner + e] = (e & 1) ? z.y : z.x;
        }
    }
}

// ----------------------------------------------------------------------------
// dispatch over the supported compile-time lengths
constexpr int kFFTThreads = 256;
constexpr int pow2floor(int v) {
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}
template <typename T> constexpr int lines_per_block(int L) {
    int n = 71680 / ((L + 1) * (int)sizeof(C2<T>));
    n = n < 1 ? 1 : (n > 16 ? 16 : n);
    return pow2floor(n);
}

#define GK_FOR_EACH_POW2(X) X(1) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

bool fft_length_supported(int L) {
    switch (L) {
#define GK_CASE(n) case n:
        GK_FOR_EACH_POW2(GK_CASE)
#undef GK_CASE
        return true;
    default: return false;
    }
}

template <typename T, int DIR>
void fft_lines(C2<T>* data, int L, long nlines, long n_inner, long inner, long outer, long estride, int nvalid,
               int nkeep, hipStream_t s) {
    if (nlines <= 0) return;
    const C2<T>* tw;
    if constexpr (sizeof(T) == 4) tw = twiddle_f32(L);
    else tw = twiddle_f64(L);
    switch (L) {
#define GK_CASE(n)                                                                                              \
    case n: {                                                                                                   \
        constexpr int NL = lines_per_block<T>(n);                                                               \
        k_fft_lines<T, n, DIR, NL, kFFTThreads><<<div_up(nlines, NL), kFFTThreads, 0, s>>>(                     \
            data, tw, nlines, n_inner, inner, outer, estride, nvalid, nkeep);                                   \
        break;                                                                                                  \
    }
        GK_FOR_EACH_POW2(GK_CASE)
#undef GK_CASE
    default: fprintf(stderr, "libgpukern: FFT length %d is not supported\n", L); abort();
    }
    GK_LAUNCH_CHECK();
}

template void fft_lines<float, -1>(float2*, int, long, long, long, long, long, int, int, hipStream_t);
template void fft_lines<float, +1>(float2*, int, long, long, long, long, long, int, int, hipStream_t);
template void fft_lines<double, -1>(double2*, int, long, long, long, long, long, int, int, hipStream_t);
template void fft_lines<double, +1>(double2*, int, long, long, long, long, long, int, int, hipStream_t);

void fft_rows_r2c(const float* in, float2* out, int P, long nrows, long n_inner, long in_inner, long in_outer,
                  long out_inner, long out_outer, int n_real, hipStream_t s) {
    if (nrows <= 0) return;
    GK_ASSERT(P % 2 == 0 && n_real <= P, "real FFT length %d (data %d)", P, n_real);
    const int M = P / 2;
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(P);
    switch (M) {
#define GK_CASE(n)                                                                                              \
    case n: {                                                                                                   \
        constexpr int NL = lines_per_block<float>(n);                                                           \
        k_rows_r2c<n, NL, kFFTThreads><<<div_up(nrows, NL), kFFTThreads, 0, s>>>(                               \
            in, out, twM, twP, nrows, n_inner, in_inner, in_outer, out_inner, out_outer, n_real);               \
        break;                                                                                                  \
    }
        GK_FOR_EACH_POW2(GK_CASE)
#undef GK_CASE
    default: fprintf(stderr, "libgpukern: real FFT length %d is not supported\n", P); abort();
    }
    GK_LAUNCH_CHECK();
}

void fft_rows_c2r(const float2* in, float* out, int P, long nrows, long n_inner, long in_inner, long in_outer,
                  long out_inner, long out_outer, int n_real, hipStream_t s) {
    if (nrows <= 0) return;
    GK_ASSERT(P % 2 == 0 && n_real <= P, "real FFT length %d (data %d)", P, n_real);
    const int M = P / 2;
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(P);
    switch (M) {
#define GK_CASE(n)                                                                                              \
    case n: {                                                                                                   \
        constexpr int NL = lines_per_block<float>(n);                                                           \
        k_rows_c2r<n, NL, kFFTThreads><<<div_up(nrows, NL), kFFTThreads, 0, s>>>(                               \
            in, out, twM, twP, nrows, n_inner, in_inner, in_outer, out_inner, out_outer, n_real);               \
        break;                                                                                                  \
    }
        GK_FOR_EACH_POW2(GK_CASE)
#undef GK_CASE
    default: fprintf(stderr, "libgpukern: real FFT length %d is not supported\n", P); abort();
    }
    GK_LAUNCH_CHECK();
}

// 3-D complex FFT of [n0][n1][n2] in FP64 (kernel setup), all lines.
void fft3_c2c_f64(double2* d, const int n[3], int dir, hipStream_t s) {
    const long n12 = (long)n[1] * n[2];
    if (dir < 0) {
        if (n[2] > 1) fft_lines_any<double, -1>(d, n[2], (long)n[0] * n[1], (long)n[0] * n[1], n[2], 0, 1, n[2], n[2], s);
        if (n[1] > 1) fft_lines_any<double, -1>(d, n[1], (long)n[0] * n[2], n[2], 1, n12, n[2], n[1], n[1], s);
        if (n[0] > 1) fft_lines_any<double, -1>(d, n[0], n12, n12, 1, 0, n12, n[0], n[0], s);
    } else {
        if (n[2] > 1) fft_lines_any<double, +1>(d, n[2], (long)n[0] * n[1], (long)n[0] * n[1], n[2], 0, 1, n[2], n[2], s);
        if (n[1] > 1) fft_lines_any<double, +1>(d, n[1], (long)n[0] * n[2], n[2], 1, n12, n[2], n[1], n[1], s);
        if (n[0] > 1) fft_lines_any<double, +1>(d, n[0], n12, n12, 1, 0, n12, n[0], n[0], s);
    }
}

}  // namespace gk

// ----------------------------------------------------------------------------
// drop-in plan
using namespace gk;

// Scratch of gpuFFT3dPlan_forward_noalloc: the packed rows and Bluestein buffers
// of the any-length transforms, 4 Mi float2 (32 MiB) per device, held by every
// device the library runs on until the process ends.  Host code never passes
// `scratch` itself to a kernel: it takes the device address that
// hipGetSymbolAddress gives on each device (scratch_device_ptr).
constexpr size_t kFFTScratchElems = size_t(4) << 20;
static __device__ __align__(256) float2 scratch[kFFTScratchElems];

static float2* scratch_device_ptr() {
    static std::mutex mu;
    static std::map<int, float2*> ptr;  // device -> address of its copy of `scratch`
    int dev = 0;
    GK_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptr.find(dev);
    if (it != ptr.end()) return it->second;
    void* a = nullptr;
    GK_CHECK(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
    GK_ASSERT(a != nullptr, "no device address for the FFT scratch array");
    ptr[dev] = static_cast<float2*>(a);
    return ptr[dev];
}

struct gpuFFT3dPlan {
    int size[3];
    int padded[3];
    int pzc;
};

extern "C" {

gpuFFT3dPlan* new_gpuFFT3dPlan_padded(int* size, int* paddedSize) {
    GK_ASSERT(paddedSize[0] > 0 && paddedSize[1] > 0 && paddedSize[2] > 1,
              "new_gpuFFT3dPlan_padded: padded size %d x %d x %d", paddedSize[0], paddedSize[1], paddedSize[2]);
    for (int i = 0; i < 3; i++)
        GK_ASSERT(size[i] > 0 && size[i] <= paddedSize[i], "new_gpuFFT3dPlan_padded: size[%d]=%d padded=%d", i,
                  size[i], paddedSize[i]);
    gpuFFT3dPlan* p = new gpuFFT3dPlan;
    for (int i = 0; i < 3; i++) {
        p->size[i] = size[i];
        p->padded[i] = paddedSize[i];
    }
    p->pzc = paddedSize[2] / 2 + 1;
    // the twiddle and chirp tables of both directions are built here, once, so
    // that no transform call has to allocate them (without a device they are
    // built by the first transform instead)
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) {
        for (int dir = -1; dir <= 1; dir += 2) {
            fft_rows_any_tables(paddedSize[2], dir);
            if (paddedSize[1] > 1) fft_lines_any_tables(paddedSize[1], dir);
            if (paddedSize[0] > 1) fft_lines_any_tables(paddedSize[0], dir);
        }
    }
    return p;
}

void delete_gpuFFT3dPlan(gpuFFT3dPlan* plan) { delete plan; }

int gpuFFT3dPlan_normalization(gpuFFT3dPlan* p) { return p->padded[0] * p->padded[1] * p->padded[2]; }

void gpuFFT3dPlan_forward(gpuFFT3dPlan* p, float* input, float* output) {
    const int N0 = p->size[0], N1 = p->size[1], N2 = p->size[2];
    const int P0 = p->padded[0], P1 = p->padded[1], P2 = p->padded[2], Q = p->pzc;
    float2* out = (float2*)output;
    hipStream_t s = stream();
    // Z: real rows (x<N0, y<N1) -> out[x][y][0..Q)
    fft_rows_r2c_any(input, out, P2, (long)N0 * N1, N1, N2, (long)N1 * N2, Q, (long)P1 * Q, N2, s);
    // Y: columns (x<N0, kz) over y; rows y>=N1 are zero padding and not read
    if (P1 > 1) fft_lines_any<float, -1>(out, P1, (long)N0 * Q, Q, 1, (long)P1 * Q, Q, N1, P1, s);
    // X: every (ky,kz) column; planes x>=N0 are zero padding and not read
    if (P0 > 1) fft_lines_any<float, -1>(out, P0, (long)P1 * Q, (long)P1 * Q, 1, 0, (long)P1 * Q, N0, P0, s);
    sync();
}

// float2 elements of scratch the forward transform of plan p needs: its three
// passes run one after the other on one stream, so they share it
static size_t forward_scratch_elems(const gpuFFT3dPlan* p) {
    const long N0 = p->size[0], N1 = p->size[1];
    const int P0 = p->padded[0], P1 = p->padded[1], P2 = p->padded[2], Q = p->pzc;
    size_t need = fft_rows_r2c_any_scratch(P2, N0 * N1);
    if (P1 > 1) need = std::max(need, fft_lines_any_scratch(P1, N0 * Q));
    if (P0 > 1) need = std::max(need, fft_lines_any_scratch(P0, (long)P1 * Q));
    return need;
}

void gpuFFT3dPlan_forward_noalloc(gpuFFT3dPlan* p, float* input, float* output) {
    const size_t need = forward_scratch_elems(p);
    GK_ASSERT(need <= kFFTScratchElems,
              "gpuFFT3dPlan_forward_noalloc: padded size %d x %d x %d needs %zu scratch elements (8 bytes each); "
              "the maximum is %zu (use gpuFFT3dPlan_forward)",
              p->padded[0], p->padded[1], p->padded[2], need, kFFTScratchElems);
    const int N0 = p->size[0], N1 = p->size[1], N2 = p->size[2];
    const int P0 = p->padded[0], P1 = p->padded[1], P2 = p->padded[2], Q = p->pzc;
    const FFTScratch sc{need ? scratch_device_ptr() : nullptr, kFFTScratchElems};
    float2* out = (float2*)output;
    hipStream_t s = stream();
    tables_frozen(true);
    fft_rows_r2c_any(input, out, P2, (long)N0 * N1, N1, N2, (long)N1 * N2, Q, (long)P1 * Q, N2, s, &sc);
    if (P1 > 1) fft_lines_any<float, -1>(out, P1, (long)N0 * Q, Q, 1, (long)P1 * Q, Q, N1, P1, s, &sc);
    if (P0 > 1) fft_lines_any<float, -1>(out, P0, (long)P1 * Q, (long)P1 * Q, 1, 0, (long)P1 * Q, N0, P0, s, &sc);
    tables_frozen(false);
    sync();
}

void gpuFFT3dPlan_inverse(gpuFFT3dPlan* p, float* input, float* output) {
    const int N0 = p->size[0], N1 = p->size[1], N2 = p->size[2];
    const int P0 = p->padded[0], P1 = p->padded[1], P2 = p->padded[2], Q = p->pzc;
    float2* in = (float2*)input;
    hipStream_t s = stream();
    // X: all columns, only planes x<N0 are kept
    if (P0 > 1) fft_lines_any<float, +1>(in, P0, (long)P1 * Q, (long)P1 * Q, 1, 0, (long)P1 * Q, P0, N0, s);
    // Y: columns of the kept planes, only rows y<N1 are kept
    if (P1 > 1) fft_lines_any<float, +1>(in, P1, (long)N0 * Q, Q, 1, (long)P1 * Q, Q, P1, N1, s);
    // Z: complex rows (x<N0, y<N1) -> first N2 reals
    fft_rows_c2r_any(in, output, P2, (long)N0 * N1, N1, Q, (long)P1 * Q, N2, (long)N1 * N2, N2, s);
    sync();
}

}  // extern "C"
