// Device-side FFT building blocks for gfx950.
//
// Lines of complex data live in LDS (one line per row, row stride `ls`
// complex elements).  A block of NT threads transforms NL lines of length L
// (L = 2^a 3^b 5^c 7^d) with Stockham auto-sort passes of radix 8/4/2/3/5/7:
// every pass loads its butterfly inputs into registers, barriers, applies the
// twiddles and the radix-R DFT, and stores in natural order, so the output
// needs no bit reversal.  L, NL and NT are compile-time, so every loop is
// unrolled and the working set stays in VGPRs.
//
// Conventions follow the reference's cuFFT usage (gpu_fft0.cu:57-109):
// forward = exp(-i...), inverse = exp(+i...), unnormalised.
#pragma once
#include <hip/hip_runtime.h>

#include "gk_common.h"

// Multiply-adds in the butterflies, twiddles and real-transform splits are
// fused only where one expression holds both (fp contract "on"), never across
// statements: the result of a transform is then the same in every kernel that
// inlines it, whatever the surrounding code (the batched and single-simulation
// passes give bit-identical fields).
#define GK_FFT_FP _Pragma("clang fp contract(on)")

namespace gk {

template <typename T> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };
template <typename T> using C2 = typename V2<T>::type;

template <typename T> __device__ __forceinline__ C2<T> mk(T a, T b) {
    C2<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
    GK_FFT_FP
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ C2<T> cconj(C2<T> a) { return mk<T>(a.x, -a.y); }
template <typename T> __device__ __forceinline__ C2<T> cscale(C2<T> a, T s) {
    GK_FFT_FP
    return mk<T>(a.x * s, a.y * s);
}

// multiply by -i (DIR<0) or +i (DIR>0)
template <typename T, int DIR> __device__ __forceinline__ C2<T> rot90(C2<T> v) {
    return DIR < 0 ? mk<T>(v.y, -v.x) : mk<T>(-v.y, v.x);
}

// twiddle exp(DIR * 2 pi i * idx / L) from the forward table of length L
template <typename T, int DIR> __device__ __forceinline__ C2<T> twid(const C2<T>* tw, int idx) {
    C2<T> w = tw[idx];
    if (DIR > 0) w.y = -w.y;
    return w;
}

template <typename T, int DIR> __device__ __forceinline__ void dft2(C2<T>* v) {
    GK_FFT_FP
    C2<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
}

template <typename T, int DIR> __device__ __forceinline__ void dft4(C2<T>& v0, C2<T>& v1, C2<T>& v2, C2<T>& v3) {
    GK_FFT_FP
    C2<T> t0 = cadd<T>(v0, v2), t1 = csub<T>(v0, v2);
    C2<T> t2 = cadd<T>(v1, v3), t3 = rot90<T, DIR>(csub<T>(v1, v3));
    v0 = cadd<T>(t0, t2);
    v1 = cadd<T>(t1, t3);
    v2 = csub<T>(t0, t2);
    v3 = csub<T>(t1, t3);
}

// multiply by exp(DIR * 2 pi i * k / 8), k = 1, 2, 3
template <typename T, int DIR, int K> __device__ __forceinline__ C2<T> w8(C2<T> v) {
    GK_FFT_FP
    const T h = (T)0.70710678118654752440;
    if (K == 2) return rot90<T, DIR>(v);
    if (DIR < 0) {
        if (K == 1) return mk<T>((v.x + v.y) * h, (v.y - v.x) * h);
        return mk<T>((v.y - v.x) * h, -(v.x + v.y) * h);  // K == 3
    } else {
        if (K == 1) return mk<T>((v.x - v.y) * h, (v.x + v.y) * h);
        return mk<T>(-(v.x + v.y) * h, (v.x - v.y) * h);  // K == 3
    }
}

template <typename T, int DIR> __device__ __forceinline__ void dft8(C2<T>* v) {
    GK_FFT_FP
    C2<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C2<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<T, DIR>(e0, e1, e2, e3);
    dft4<T, DIR>(o0, o1, o2, o3);
    o1 = w8<T, DIR, 1>(o1);
    o2 = w8<T, DIR, 2>(o2);
    o3 = w8<T, DIR, 3>(o3);
    v[0] = cadd<T>(e0, o0);
    v[4] = csub<T>(e0, o0);
    v[1] = cadd<T>(e1, o1);
    v[5] = csub<T>(e1, o1);
    v[2] = cadd<T>(e2, o2);
    v[6] = csub<T>(e2, o2);
    v[3] = cadd<T>(e3, o3);
    v[7] = csub<T>(e3, o3);
}

template <typename T, int DIR> __device__ __forceinline__ void dft3(C2<T>* v) {
    GK_FFT_FP
    const T s = (T)0.86602540378443864676;  // sin(2 pi / 3)
    C2<T> t1 = cadd<T>(v[1], v[2]);
    C2<T> t2 = mk<T>(v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y);
    C2<T> t3 = cscale<T>(rot90<T, DIR>(csub<T>(v[1], v[2])), s);
    v[0] = cadd<T>(v[0], t1);
    v[1] = cadd<T>(t2, t3);
    v[2] = csub<T>(t2, t3);
}

template <typename T, int DIR> __device__ __forceinline__ void dft5(C2<T>* v) {
    GK_FFT_FP
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    C2<T> t1 = cadd<T>(v[1], v[4]), t2 = cadd<T>(v[2], v[3]);
    C2<T> t3 = csub<T>(v[1], v[4]), t4 = csub<T>(v[2], v[3]);
    C2<T> a1 = mk<T>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    C2<T> a2 = mk<T>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    C2<T> b1 = rot90<T, DIR>(mk<T>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
    C2<T> b2 = rot90<T, DIR>(mk<T>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
    v[0] = cadd<T>(v[0], cadd<T>(t1, t2));
    v[1] = cadd<T>(a1, b1);
    v[4] = csub<T>(a1, b1);
    v[2] = cadd<T>(a2, b2);
    v[3] = csub<T>(a2, b2);
}

template <typename T, int DIR> __device__ __forceinline__ void dft7(C2<T>* v) {
    GK_FFT_FP
    // cos / sin of 2 pi k / 7, k = 1, 2, 3
    const T c[4] = {(T)1, (T)0.62348980185873353053, (T)-0.22252093395631440429, (T)-0.90096886790241912624};
    const T sn[4] = {(T)0, (T)0.78183148246802980871, (T)0.97492791218182360702, (T)0.43388373911755812048};
    C2<T> tp[4], tm[4];
#pragma unroll
    for (int k = 1; k <= 3; k++) {
        tp[k] = cadd<T>(v[k], v[7 - k]);
        tm[k] = csub<T>(v[k], v[7 - k]);
    }
    C2<T> out[7];
    out[0] = cadd<T>(v[0], cadd<T>(tp[1], cadd<T>(tp[2], tp[3])));
#pragma unroll
    for (int j = 1; j <= 3; j++) {
        C2<T> a = v[0], b = mk<T>(0, 0);
#pragma unroll
        for (int k = 1; k <= 3; k++) {
            const int jk = (j * k) % 7;
            const T cc = jk <= 3 ? c[jk] : c[7 - jk];
            const T ss = jk <= 3 ? sn[jk] : -sn[7 - jk];
            a = mk<T>(a.x + cc * tp[k].x, a.y + cc * tp[k].y);
            b = mk<T>(b.x + ss * tm[k].x, b.y + ss * tm[k].y);
        }
        b = rot90<T, DIR>(b);
        out[j] = cadd<T>(a, b);
        out[7 - j] = csub<T>(a, b);
    }
#pragma unroll
    for (int r = 0; r < 7; r++) v[r] = out[r];
}

template <typename T, int R, int DIR> __device__ __forceinline__ void dft(C2<T>* v) {
    if constexpr (R == 2) dft2<T, DIR>(v);
    else if constexpr (R == 4) dft4<T, DIR>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8<T, DIR>(v);
    else if constexpr (R == 3) dft3<T, DIR>(v);
    else if constexpr (R == 5) dft5<T, DIR>(v);
    else dft7<T, DIR>(v);
}

// radix chosen for the remaining length `rem` (product of the radices still to do)
constexpr int next_radix(int rem) {
    return rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : rem % 5 == 0 ? 5 : 7;
}
constexpr bool fft_size_ok(int L) {
    while (L % 2 == 0) L /= 2;
    while (L % 3 == 0) L /= 3;
    while (L % 5 == 0) L /= 5;
    while (L % 7 == 0) L /= 7;
    return L == 1;
}

// One Stockham DIT pass (radix R, Ns = product of the previous radices) over
// NL lines of length L in LDS.  On the first pass (NS==1) elements at index
// >= nvalid are known zeros (zero padding) and are not read.
template <typename T, int L, int R, int NS, int DIR, int NL, int NT>
__device__ __forceinline__ void stockham_pass(C2<T>* lds, int ls, const C2<T>* tw, int nvalid) {
    constexpr int NBF = L / R;
    constexpr int TOT = NL * NBF;
    constexpr int NB = (TOT + NT - 1) / NT;
    C2<T> v[NB][R];
    const int tid = threadIdx.x;
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = tid + q * NT;
        if ((TOT % NT) == 0 || b < TOT) {
            const int line = b / NBF, j = b % NBF;
            const C2<T>* src = lds + line * ls;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int idx = j + r * NBF;
                if (NS == 1) v[q][r] = idx < nvalid ? src[idx] : mk<T>(0, 0);
                else v[q][r] = src[idx];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int b = tid + q * NT;
        if ((TOT % NT) == 0 || b < TOT) {
            const int line = b / NBF, j = b % NBF;
            const int k = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[q][r] = cmul<T>(v[q][r], twid<T, DIR>(tw, r * k * (L / (NS * R))));
            }
            dft<T, R, DIR>(v[q]);
            C2<T>* dst = lds + line * ls;
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) dst[base + r * NS] = v[q][r];
        }
    }
    __syncthreads();
}

// Full FFT of NL lines of length L in LDS (in place, natural order in and out).
// `tw` is the table exp(-2 pi i k / L), k < L.
// The caller must __syncthreads() between filling LDS and calling this.
template <typename T, int L, int DIR, int NL, int NT, int NS = 1>
__device__ __forceinline__ void fft_lds(C2<T>* lds, int ls, const C2<T>* tw, int nvalid) {
    if constexpr (NS < L) {
        constexpr int rem = L / NS;
        constexpr int R = next_radix(rem);
        static_assert(fft_size_ok(L), "FFT length must factor over 2, 3, 5, 7");
        stockham_pass<T, L, R, NS, DIR, NL, NT>(lds, ls, tw, nvalid);
        fft_lds<T, L, DIR, NL, NT, NS * R>(lds, ls, tw, nvalid);
    }
}

// ---------------------------------------------------------------------------
// Register FFT: one line of length L is held by G consecutive lanes, E = L/G
// elements per lane in the lane-strided layout v[e] = x[g + G e] (g = lane
// within its group).  In that layout every butterfly of a Stockham pass finds
// all its inputs in one lane, and the last pass writes its outputs back into
// the same layout, so a whole transform needs only (passes - 1) exchanges
// through LDS, and loading / storing the line from / to memory is coalesced.
// `lds` is this line's exchange buffer (RegFFT<L>::LDS complex), `tw` the table of L twiddles.
// Register-FFT lengths are 2^a and 3 * 2^a.  Elements per lane: 8 for 2^a up
// to 512 (radix-8 passes), 12 for 3 * 2^a up to 768 (radix 4, 2 and one radix
// 3), L/64 beyond (64 lanes per line); G = L/E lanes per line is a power of two.
constexpr int reg_fft_E(int L) {
    return L % 3 == 0 ? (L <= 12 ? L : (L / 12 <= 64 ? 12 : L / 64)) : (L <= 512 ? (L < 8 ? L : 8) : L / 64);
}
constexpr int reg_fft_G(int L) { return L / reg_fft_E(L); }
constexpr bool reg_fft_length(int L) {
    int m = L % 3 == 0 ? L / 3 : L;
    while (m % 2 == 0) m /= 2;
    return L >= 1 && m == 1;
}
// radix of the next register pass: the largest of 8, 4, 2 dividing both the
// remaining length and the elements per lane, else 3
constexpr int reg_radix(int rem, int E) {
    return (rem % 8 == 0 && E % 8 == 0) ? 8
         : (rem % 4 == 0 && E % 4 == 0) ? 4
         : (rem % 2 == 0 && E % 2 == 0) ? 2
         : (rem % 3 == 0 && E % 3 == 0) ? 3 : 0;
}

template <int L> struct RegFFT {
    static_assert(reg_fft_length(L), "register FFT lengths are 2^a and 3 * 2^a");
    static constexpr int E = reg_fft_E(L);  // elements per lane
    static constexpr int G = L / E;         // lanes per line
    // exchange buffer per line: one float2 of padding per 8, so the stride-8
    // scatter of a radix-8 pass and neighbouring lines spread over all 64 banks
    static constexpr int LDS = L + L / 8;
};
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 3); }

// A padded length L >= 2 satisfies L >= 2n - 1 and is even, so only indices
// < L/2 can hold data (forward inputs) or be kept (inverse outputs).  In the
// lane-strided layout (index g + G e) these are exactly the elements e < E/2;
// the others are compile-time zeros / dead, and the compiler prunes the first
// and last passes accordingly.
template <int L> __device__ constexpr bool pad_half(int e) {
    return RegFFT<L>::E >= 2 && RegFFT<L>::E % 2 == 0 && e >= RegFFT<L>::E / 2;
}

// The lanes of one line (G = L / E <= 64) are always in one wave, and the
// exchange buffer belongs to that line, so a wave-level barrier orders its LDS
// traffic (the pattern of rocPRIM's wave_barrier: release fence, wave barrier,
// acquire fence), and waves do not wait for each other between radix passes.
// Every caller's exchange buffer is private to its line during the transform;
// where the buffers overlay other data, the callers separate the uses with
// block barriers of their own.  WS = false keeps block barriers.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// EE: elements per lane (default RegFFT<L>::E; a smaller power-of-two EE spreads
// a line over more lanes, L / EE, with the same exchange buffer); WS: wave-level
// instead of block-level barriers around the exchanges (see wave_sync)
template <typename T, int L, int DIR, int NS = 1, int EE = reg_fft_E(L), bool WS = true>
__device__ __forceinline__ void reg_fft(C2<T>* v, int g, C2<T>* lds, const C2<T>* tw) {
    if constexpr (NS < L) {
        constexpr int E = EE, G = L / EE;
        static_assert(L % EE == 0, "elements per lane must divide the length");
        constexpr int R = reg_radix(L / NS, E);
        static_assert(R != 0 && E % R == 0, "radix must divide the elements per lane");
        constexpr int NB = E / R;  // butterflies per lane
        constexpr bool last = NS * R == L;
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const int j = g + G * q;
            const int k = j % NS;
            C2<T> b[R];
#pragma unroll
            for (int r = 0; r < R; r++) b[r] = v[q + r * NB];
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) b[r] = cmul<T>(b[r], twid<T, DIR>(tw, r * k * (L / (NS * R))));
            }
            dft<T, R, DIR>(b);
            if (last) {
                // output index j + r L/R = g + G (q + r NB): back in the lane-strided slots
#pragma unroll
                for (int r = 0; r < R; r++) v[q + r * NB] = b[r];
            } else {
                const int base = (j / NS) * NS * R + k;
#pragma unroll
                for (int r = 0; r < R; r++) lds[lds_pad(base + r * NS)] = b[r];
            }
        }
        if constexpr (!last) {
            if constexpr (WS) wave_sync();
            else __syncthreads();
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = lds[lds_pad(g + G * e)];
            if constexpr (WS) wave_sync();
            else __syncthreads();
            reg_fft<T, L, DIR, NS * R, EE, WS>(v, g, lds, tw);
        }
    }
}

// Load phases issue kUnroll independent global loads per lane before any of
// them is consumed, so a block keeps up to 8 x NT loads in flight instead of
// waiting one memory latency per element (loop trip counts are runtime values,
// so the compiler would not batch them by itself).
constexpr int kUnroll = 8;
template <typename T, int NT, typename Ld, typename St>
__device__ __forceinline__ void batched_copy(int cnt, Ld ld, St st) {
    for (int base = threadIdx.x; base < cnt; base += NT * kUnroll) {
        T v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const int i = base + u * NT;
            if (i < cnt) v[u] = ld(i);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const int i = base + u * NT;
            if (i < cnt) st(i, v[u]);
        }
    }
}

// Real-to-complex split: given Z = FFT_M(x[2n] + i x[2n+1]) (M = P/2), returns
// X[k] of the length-P real transform, k in [0, M]:  X = E + W^k O with
// E = (Z[k] + conj Z[M-k]) / 2,  O = (Z[k] - conj Z[M-k]) / 2i,  W = exp(-2 pi i / P).
template <typename T>
__device__ __forceinline__ C2<T> r2c_split(C2<T> zk, C2<T> zmk, C2<T> wk /* exp(-2 pi i k/P) */) {
    GK_FFT_FP
    C2<T> e = mk<T>((zk.x + zmk.x) * (T)0.5, (zk.y - zmk.y) * (T)0.5);
    // D = zk - conj(zmk);  O = D / 2i = (D.y/2, -D.x/2)
    T dx = zk.x - zmk.x, dy = zk.y + zmk.y;
    C2<T> o = mk<T>(dy * (T)0.5, -dx * (T)0.5);
    return cadd<T>(e, cmul<T>(wk, o));
}

// Complex-to-real merge (inverse of r2c_split, unnormalised like a C2R):
// Y[k] = (X[k] + conj X[M-k]) + i W^-k (X[k] - conj X[M-k]),  k in [0, M);
// then IFFT_M(Y) = x[2n] + i x[2n+1].
template <typename T>
__device__ __forceinline__ C2<T> c2r_merge(C2<T> xk, C2<T> xmk, C2<T> wk /* exp(-2 pi i k/P) */) {
    GK_FFT_FP
    C2<T> s = mk<T>(xk.x + xmk.x, xk.y - xmk.y);
    C2<T> d = mk<T>(xk.x - xmk.x, xk.y + xmk.y);
    C2<T> t = cmul<T>(cconj<T>(wk), d);  // W^-k * d
    return mk<T>(s.x - t.y, s.y + t.x);   // s + i t
}

}  // namespace gk
