// Host-side launchers of the FFT kernels (gk_fft.hip), shared by the
// drop-in plan and the engine.
#pragma once
#include <hip/hip_runtime.h>

#include "fft_device.h"

namespace gk {

bool fft_length_supported(int L);

// Lines of length L: line l starts at (l / n_inner)*outer + (l % n_inner)*inner,
// elements at stride estride (in complex elements).  Reads the first nvalid
// elements (the rest are zero padding), stores the first nkeep.
template <typename T, int DIR>
void fft_lines(C2<T>* data, int L, long nlines,
               long n_inner, long inner, long outer, long estride, int nvalid, int nkeep, hipStream_t s);

// Rows of n_real floats (row r at (r / n_inner)*in_outer + (r % n_inner)*in_inner)
// -> P/2+1 complex (same indexing with out_inner/out_outer, complex units).
void fft_rows_r2c(const float* in, float2* out, int P, long nrows, long n_inner, long in_inner, long in_outer,
                  long out_inner, long out_outer, int n_real, hipStream_t s);
// inverse: P/2+1 complex per row -> first n_real floats of the unnormalised real transform
void fft_rows_c2r(const float2* in, float* out, int P, long nrows, long n_inner, long in_inner, long in_outer,
                  long out_inner, long out_outer, int n_real, hipStream_t s);

// Device memory for the temporaries of the any-length transforms below: when
// `scr` is given, their packed rows and Bluestein buffers are carved from it
// (stream-ordered reuse, nothing allocated or freed); else they are allocated
// per call.  `cap` counts float2 elements.
struct FFTScratch {
    float2* base = nullptr;
    size_t cap = 0;
};

// The same for any length (gk_fft_any.hip): power-of-two lengths go to the
// kernels above, lengths with factors 2, 3, 5, 7 to LDS Stockham passes with a
// runtime radix list, others through Bluestein's chirp-z convolution.
template <typename T, int DIR>
void fft_lines_any(C2<T>* data, int L, long nlines, long n_inner, long inner, long outer, long estride, int nvalid,
                   int nkeep, hipStream_t s, const FFTScratch* scr = nullptr);
void fft_rows_r2c_any(const float* in, float2* out, int P, long nrows, long n_inner, long in_inner, long in_outer,
                      long out_inner, long out_outer, int n_real, hipStream_t s, const FFTScratch* scr = nullptr);
// float2 elements of scratch the FP32 calls above need for these sizes (0: none)
size_t fft_lines_any_scratch(int L, long nlines);
size_t fft_rows_r2c_any_scratch(int P, long nrows);
// Build (once, cached) every twiddle / chirp table the FP32 transforms of this
// length and direction read, so that a later call finds them all.
void fft_lines_any_tables(int L, int dir);
void fft_rows_any_tables(int P, int dir);
void fft_rows_c2r_any(const float2* in, float* out, int P, long nrows, long n_inner, long in_inner, long in_outer,
                      long out_inner, long out_outer, int n_real, hipStream_t s);

// in-place 3-D complex FFT in FP64 of [n0][n1][n2] (dir -1 forward, +1 inverse)
void fft3_c2c_f64(double2* d, const int n[3], int dir, hipStream_t s);

}  // namespace gk
