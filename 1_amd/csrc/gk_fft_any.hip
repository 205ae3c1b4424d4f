// Transforms of any length for the drop-in plan (gpu_fft0.cu:30-112 accepts
// every size cuFFT does; the Go driver pads each axis to 2N, kernel.go:80-90,
// so lengths such as 2^a 3^b 5^c 7^d and 2p for a prime p all occur).
//
// Lengths whose factors are 2, 3, 5 and 7 run as Stockham passes in LDS with
// a runtime list of radices (8, 4, 2, 3, 5, 7), one line per block.  Any other
// length uses Bluestein's chirp-z identity nk = (n^2 + k^2 - (k-n)^2)/2:
//   X[k] = w[k] sum_n (x[n] w[n]) conj(w)[k-n],   w[n] = exp(DIR pi i n^2 / L),
// i.e. a circular convolution of length Lp >= 2L-1 (a power of two) done with
// the power-of-two line transforms; the chirp and the transformed filter are
// computed on the host in long double.  Real rows of even length P use the
// half-length complex transform (same split / merge as the power-of-two path);
// odd P (a periodic axis) uses the full complex transform of length P.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <map>
#include <mutex>
#include <vector>

#include "fft_device.h"
#include "gk_common.h"
#include "gk_fft.h"

namespace gk {

struct RadixList {
    int n;
    int r[32];
};

static bool factor_radices(int L, RadixList& rl) {
    rl.n = 0;
    int rem = L;
    while (rem > 1) {
        const int R = rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : rem % 5 == 0 ? 5
                                                                                                 : rem % 7 == 0 ? 7
                                                                                                                : 0;
        if (!R || rl.n == 32) return false;
        rl.r[rl.n++] = R;
        rem /= R;
    }
    return true;
}

// one Stockham pass of radix RR over a line of L held in LDS (src -> dst)
template <typename T, int DIR, int RR>
__device__ __forceinline__ void stockham_pass_rt(const C2<T>* src, C2<T>* dst, int L, int NS,
                                                 const C2<T>* __restrict__ tw) {
    const int nb = L / RR;
    for (int j = threadIdx.x; j < nb; j += blockDim.x) {
        const int k = j % NS;
        C2<T> v[RR];
#pragma unroll
        for (int r = 0; r < RR; r++) v[r] = src[j + r * nb];
        if (NS > 1) {
            const int step = L / (NS * RR);
#pragma unroll
            for (int r = 1; r < RR; r++) v[r] = cmul<T>(v[r], twid<T, DIR>(tw, r * k * step));
        }
        dft<T, RR, DIR>(v);
        const int base = (j / NS) * NS * RR + k;
#pragma unroll
        for (int r = 0; r < RR; r++) dst[base + r * NS] = v[r];
    }
}

// lines as in k_fft_lines: line l starts at (l / n_inner) outer + (l % n_inner) inner,
// elements at stride estride; the first nvalid inputs are read, the first nkeep stored
template <typename T, int DIR>
__global__ __launch_bounds__(256) void k_fft_lines_rt(C2<T>* __restrict__ data, const C2<T>* __restrict__ tw, int L,
                                                      RadixList rl, long nlines, long n_inner, long inner, long outer,
                                                      long estride, int nvalid, int nkeep) {
    extern __shared__ __align__(16) unsigned char smem_rt[];
    C2<T>* a = reinterpret_cast<C2<T>*>(smem_rt);
    C2<T>* b = a + L;
    const long l = blockIdx.x;
    if (l >= nlines) return;
    C2<T>* line = data + (l / n_inner) * outer + (l % n_inner) * inner;
    for (int i = threadIdx.x; i < L; i += blockDim.x) a[i] = i < nvalid ? line[(long)i * estride] : mk<T>(0, 0);
    __syncthreads();
    int NS = 1;
    for (int p = 0; p < rl.n; p++) {
        const int R = rl.r[p];
        switch (R) {
        case 8: stockham_pass_rt<T, DIR, 8>(a, b, L, NS, tw); break;
        case 4: stockham_pass_rt<T, DIR, 4>(a, b, L, NS, tw); break;
        case 2: stockham_pass_rt<T, DIR, 2>(a, b, L, NS, tw); break;
        case 3: stockham_pass_rt<T, DIR, 3>(a, b, L, NS, tw); break;
        case 5: stockham_pass_rt<T, DIR, 5>(a, b, L, NS, tw); break;
        default: stockham_pass_rt<T, DIR, 7>(a, b, L, NS, tw); break;
        }
        __syncthreads();
        C2<T>* t = a;
        a = b;
        b = t;
        NS *= R;
    }
    for (int i = threadIdx.x; i < nkeep; i += blockDim.x) line[(long)i * estride] = a[i];
}

// ---------------------------------------------------------------- Bluestein
struct BlueTables {
    int L = 0, Lp = 0;
    float2* w = nullptr;  // chirp, L entries
    float2* B = nullptr;  // Lp-point transform of the conjugate chirp filter, divided by Lp
};

static void host_fft_pow2(std::vector<std::complex<long double>>& a, int sign) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; k++) {
                const long double ang = sign * 2 * pi * (long double)k / (long double)len;
                const std::complex<long double> w(cosl(ang), sinl(ang));
                const std::complex<long double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

static const BlueTables& blue_tables(int L, int dir) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, BlueTables> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({L, dir});
    if (it != cache.end()) return it->second;
    GK_ASSERT(!tables_frozen(), "chirp table of length %d is missing in a call that may not allocate", L);
    BlueTables t;
    t.L = L;
    t.Lp = 1;
    while (t.Lp < 2 * L - 1) t.Lp *= 2;
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<std::complex<long double>> w(L), b(t.Lp, 0);
    for (int n = 0; n < L; n++) {
        const long long n2 = ((long long)n * n) % (2LL * L);  // exp(i pi n^2 / L) has period 2L in n^2
        const long double ang = dir * pi * (long double)n2 / (long double)L;
        w[n] = std::complex<long double>(cosl(ang), sinl(ang));
    }
    for (int m = 0; m < L; m++) {
        b[m] = std::conj(w[m]);
        if (m) b[t.Lp - m] = std::conj(w[m]);
    }
    host_fft_pow2(b, -1);
    std::vector<float2> hw(L), hB(t.Lp);
    for (int n = 0; n < L; n++) hw[n] = make_float2((float)w[n].real(), (float)w[n].imag());
    for (int k = 0; k < t.Lp; k++)
        hB[k] = make_float2((float)(b[k].real() / t.Lp), (float)(b[k].imag() / t.Lp));
    GK_CHECK(hipMalloc(&t.w, L * sizeof(float2)));
    GK_CHECK(hipMalloc(&t.B, t.Lp * sizeof(float2)));
    GK_CHECK(hipMemcpy(t.w, hw.data(), L * sizeof(float2), hipMemcpyHostToDevice));
    GK_CHECK(hipMemcpy(t.B, hB.data(), t.Lp * sizeof(float2), hipMemcpyHostToDevice));
    return cache.emplace(std::make_pair(L, dir), t).first->second;
}

// scratch[l][n] = x_l[n] w[n] (n < nvalid), zero up to Lp
__global__ void k_blue_pre(const float2* __restrict__ data, float2* __restrict__ scr, const float2* __restrict__ w,
                           int L, int Lp, long nlines, long n_inner, long inner, long outer, long estride, int nvalid) {
    const long tot = nlines * Lp;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long l = t / Lp;
        const int n = (int)(t - l * Lp);
        float2 v = make_float2(0.f, 0.f);
        if (n < nvalid && n < L) {
            const float2 x = data[(l / n_inner) * outer + (l % n_inner) * inner + (long)n * estride];
            v = cmul<float>(x, w[n]);
        }
        scr[t] = v;
    }
}
__global__ void k_blue_mul(float2* __restrict__ scr, const float2* __restrict__ B, int Lp, long nlines) {
    const long tot = nlines * Lp;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x)
        scr[t] = cmul<float>(scr[t], B[t % Lp]);
}
__global__ void k_blue_post(float2* __restrict__ data, const float2* __restrict__ scr, const float2* __restrict__ w,
                            int Lp, long nlines, long n_inner, long inner, long outer, long estride, int nkeep) {
    const long tot = nlines * nkeep;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long l = t / nkeep;
        const int k = (int)(t - l * nkeep);
        data[(l / n_inner) * outer + (l % n_inner) * inner + (long)k * estride] = cmul<float>(scr[l * Lp + k], w[k]);
    }
}

static int grid_for(long n) {
    long g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (int)(g < 1 ? 1 : g);
}

static int blue_length(int L) {
    int Lp = 1;
    while (Lp < 2 * L - 1) Lp *= 2;
    return Lp;
}

size_t fft_lines_any_scratch(int L, long nlines) {
    if (nlines <= 0 || fft_length_supported(L)) return 0;
    RadixList rl;
    if (factor_radices(L, rl)) return 0;
    return (size_t)nlines * blue_length(L);
}

void fft_lines_any_tables(int L, int dir) {
    RadixList rl;
    if (fft_length_supported(L) || factor_radices(L, rl)) {
        twiddle_f32(L);
        return;
    }
    blue_tables(L, dir);
    twiddle_f32(blue_length(L));
}

template <typename T, int DIR>
void fft_lines_any(C2<T>* data, int L, long nlines, long n_inner, long inner, long outer, long estride, int nvalid,
                   int nkeep, hipStream_t s, const FFTScratch* sc) {
    if (nlines <= 0) return;
    if (fft_length_supported(L)) {
        fft_lines<T, DIR>(data, L, nlines, n_inner, inner, outer, estride, nvalid, nkeep, s);
        return;
    }
    RadixList rl;
    if (factor_radices(L, rl)) {
        GK_ASSERT(2L * L * (long)sizeof(C2<T>) <= 65536, "FFT length %d is too long for one block", L);
        const C2<T>* tw;
        if constexpr (sizeof(T) == 4) tw = twiddle_f32(L);
        else tw = twiddle_f64(L);
        const int nt = L <= 512 ? 64 : 256;
        k_fft_lines_rt<T, DIR><<<(unsigned)nlines, nt, 2 * L * sizeof(C2<T>), s>>>(data, tw, L, rl, nlines, n_inner,
                                                                                   inner, outer, estride, nvalid, nkeep);
        GK_LAUNCH_CHECK();
        return;
    }
    if constexpr (sizeof(T) == 4) {
        const BlueTables& bt = blue_tables(L, DIR);
        GK_ASSERT(bt.Lp <= 4096 * 2 && fft_length_supported(bt.Lp), "FFT length %d is not supported", L);
        float2* scr = nullptr;
        if (sc) {
            GK_ASSERT((size_t)nlines * bt.Lp <= sc->cap, "FFT scratch of %zu elements, %ld needed", sc->cap,
                      nlines * (long)bt.Lp);
            scr = sc->base;
        } else {
            GK_CHECK(hipMalloc(&scr, nlines * bt.Lp * sizeof(float2)));
        }
        k_blue_pre<<<grid_for(nlines * bt.Lp), 256, 0, s>>>(data, scr, bt.w, L, bt.Lp, nlines, n_inner, inner, outer,
                                                          estride, nvalid);
        GK_LAUNCH_CHECK();
        fft_lines<float, -1>(scr, bt.Lp, nlines, nlines, bt.Lp, 0, 1, bt.Lp, bt.Lp, s);
        k_blue_mul<<<grid_for(nlines * bt.Lp), 256, 0, s>>>(scr, bt.B, bt.Lp, nlines);
        GK_LAUNCH_CHECK();
        fft_lines<float, +1>(scr, bt.Lp, nlines, nlines, bt.Lp, 0, 1, bt.Lp, bt.Lp, s);
        k_blue_post<<<grid_for(nlines * nkeep), 256, 0, s>>>(data, scr, bt.w, bt.Lp, nlines, n_inner, inner, outer,
                                                           estride, nkeep);
        GK_LAUNCH_CHECK();
        if (!sc) {
            GK_CHECK(hipStreamSynchronize(s));
            GK_CHECK(hipFree(scr));
        }
    } else {
        GK_ASSERT(false, "FP64 FFT length %d is not supported", L);
    }
}

template void fft_lines_any<float, -1>(float2*, int, long, long, long, long, long, int, int, hipStream_t,
                                       const FFTScratch*);
template void fft_lines_any<float, +1>(float2*, int, long, long, long, long, long, int, int, hipStream_t,
                                       const FFTScratch*);
template void fft_lines_any<double, -1>(double2*, int, long, long, long, long, long, int, int, hipStream_t,
                                        const FFTScratch*);
template void fft_lines_any<double, +1>(double2*, int, long, long, long, long, long, int, int, hipStream_t,
                                        const FFTScratch*);

// ---------------------------------------------------------------- real rows of any length
// row r: input at (r / n_inner) in_outer + (r % n_inner) in_inner (n_real floats),
// output at (r / n_inner) out_outer + (r % n_inner) out_inner (P/2+1 complex)
__global__ void k_rows_pack(const float* __restrict__ in, float2* __restrict__ z, int H, long nrows, long n_inner,
                            long in_inner, long in_outer, int n_real, int interleave) {
    const long tot = nrows * H;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long r = t / H;
        const int n = (int)(t - r * H);
        const float* row = in + (r / n_inner) * in_outer + (r % n_inner) * in_inner;
        float a = 0.f, b = 0.f;
        if (interleave) {  // z[n] = x[2n] + i x[2n+1]
            if (2 * n < n_real) a = row[2 * n];
            if (2 * n + 1 < n_real) b = row[2 * n + 1];
        } else if (n < n_real) {
            a = row[n];
        }
        z[t] = make_float2(a, b);
    }
}
__global__ void k_rows_split(const float2* __restrict__ Z, float2* __restrict__ out, const float2* __restrict__ twP,
                             int M, long nrows, long n_inner, long out_inner, long out_outer) {
    const long tot = nrows * (M + 1);
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long r = t / (M + 1);
        const int k = (int)(t - r * (M + 1));
        const float2* z = Z + r * M;
        float2 X;
        if (k == M) X = make_float2(z[0].x - z[0].y, 0.f);
        else X = r2c_split<float>(z[k], z[(M - k) % M], twP[k]);
        out[(r / n_inner) * out_outer + (r % n_inner) * out_inner + k] = X;
    }
}
__global__ void k_rows_keep(const float2* __restrict__ Z, float2* __restrict__ out, int P, int Q, long nrows,
                            long n_inner, long out_inner, long out_outer) {
    const long tot = nrows * Q;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long r = t / Q;
        const int k = (int)(t - r * Q);
        out[(r / n_inner) * out_outer + (r % n_inner) * out_inner + k] = Z[r * P + k];
    }
}
// inverse: even P -> merged half-length spectrum; odd P -> full Hermitian spectrum
__global__ void k_rows_merge(const float2* __restrict__ in, float2* __restrict__ Y, const float2* __restrict__ twP,
                             int P, long nrows, long n_inner, long in_inner, long in_outer) {
    const int even = (P & 1) == 0, H = even ? P / 2 : P;
    const long tot = nrows * H;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long r = t / H;
        const int k = (int)(t - r * H);
        const float2* X = in + (r / n_inner) * in_outer + (r % n_inner) * in_inner;
        if (even) Y[t] = c2r_merge<float>(X[k], X[H - k], twP[k]);
        else Y[t] = k <= P / 2 ? X[k] : cconj<float>(X[P - k]);
    }
}
__global__ void k_rows_unpack(const float2* __restrict__ Y, float* __restrict__ out, int H, long nrows, long n_inner,
                              long out_inner, long out_outer, int n_real, int interleave) {
    const long tot = nrows * (long)n_real;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot; t += (long)gridDim.x * blockDim.x) {
        const long r = t / n_real;
        const int i = (int)(t - r * n_real);
        const float2 y = interleave ? Y[r * H + i / 2] : Y[r * H + i];
        out[(r / n_inner) * out_outer + (r % n_inner) * out_inner + i] = interleave ? ((i & 1) ? y.y : y.x) : y.x;
    }
}

size_t fft_rows_r2c_any_scratch(int P, long nrows) {
    if (nrows <= 0 || (P % 2 == 0 && fft_length_supported(P / 2))) return 0;
    const int H = P % 2 == 0 ? P / 2 : P;
    return (size_t)nrows * H + fft_lines_any_scratch(H, nrows);
}

void fft_rows_any_tables(int P, int dir) {
    if (P % 2 == 0) {
        twiddle_f32(P);
        if (fft_length_supported(P / 2)) {
            twiddle_f32(P / 2);
            return;
        }
        fft_lines_any_tables(P / 2, dir);
    } else {
        fft_lines_any_tables(P, dir);
    }
}

void fft_rows_r2c_any(const float* in, float2* out, int P, long nrows, long n_inner, long in_inner, long in_outer,
                      long out_inner, long out_outer, int n_real, hipStream_t s, const FFTScratch* sc) {
    if (nrows <= 0) return;
    if (P % 2 == 0 && fft_length_supported(P / 2)) {
        fft_rows_r2c(in, out, P, nrows, n_inner, in_inner, in_outer, out_inner, out_outer, n_real, s);
        return;
    }
    const bool even = P % 2 == 0;
    const int H = even ? P / 2 : P;
    float2* z = nullptr;
    FFTScratch rest;  // what the packed rows leave for the line transform
    if (sc) {
        GK_ASSERT(fft_rows_r2c_any_scratch(P, nrows) <= sc->cap, "FFT scratch of %zu elements, %zu needed", sc->cap,
                  fft_rows_r2c_any_scratch(P, nrows));
        z = sc->base;
        rest.base = sc->base + (size_t)nrows * H;
        rest.cap = sc->cap - (size_t)nrows * H;
    } else {
        GK_CHECK(hipMalloc(&z, nrows * H * sizeof(float2)));
    }
    k_rows_pack<<<grid_for(nrows * H), 256, 0, s>>>(in, z, H, nrows, n_inner, in_inner, in_outer, n_real, even);
    GK_LAUNCH_CHECK();
    fft_lines_any<float, -1>(z, H, nrows, nrows, H, 0, 1, H, H, s, sc ? &rest : nullptr);
    if (even) {
        k_rows_split<<<grid_for(nrows * (H + 1)), 256, 0, s>>>(z, out, twiddle_f32(P), H, nrows, n_inner, out_inner,
                                                              out_outer);
    } else {
        k_rows_keep<<<grid_for(nrows * (P / 2 + 1)), 256, 0, s>>>(z, out, P, P / 2 + 1, nrows, n_inner, out_inner,
                                                                 out_outer);
    }
    GK_LAUNCH_CHECK();
    if (!sc) {
        GK_CHECK(hipStreamSynchronize(s));
        GK_CHECK(hipFree(z));
    }
}

void fft_rows_c2r_any(const float2* in, float* out, int P, long nrows, long n_inner, long in_inner, long in_outer,
                      long out_inner, long out_outer, int n_real, hipStream_t s) {
    if (nrows <= 0) return;
    if (P % 2 == 0 && fft_length_supported(P / 2)) {
        fft_rows_c2r(in, out, P, nrows, n_inner, in_inner, in_outer, out_inner, out_outer, n_real, s);
        return;
    }
    const bool even = P % 2 == 0;
    const int H = even ? P / 2 : P;
    float2* y = nullptr;
    GK_CHECK(hipMalloc(&y, nrows * H * sizeof(float2)));
    k_rows_merge<<<grid_for(nrows * H), 256, 0, s>>>(in, y, twiddle_f32(P), P, nrows, n_inner, in_inner, in_outer);
    GK_LAUNCH_CHECK();
    fft_lines_any<float, +1>(y, H, nrows, nrows, H, 0, 1, H, H, s);
    k_rows_unpack<<<grid_for(nrows * (long)n_real), 256, 0, s>>>(y, out, H, nrows, n_inner, out_inner, out_outer,
                                                                n_real, even);
    GK_LAUNCH_CHECK();
    GK_CHECK(hipStreamSynchronize(s));
    GK_CHECK(hipFree(y));
}

}  // namespace gk
