// Common host/device helpers for libgpukern (MI355X / gfx950).
//
// Error behaviour follows the reference library: every failure prints a
// message to stderr and aborts the process (gpu_safe.h:36-57); there are no
// return codes on the reference entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define GK_CHECK(call)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            fprintf(stderr, "libgpukern: HIP error %s (%d) at %s:%d: %s\n", hipGetErrorString(e_), \
                    (int)e_, __FILE__, __LINE__, #call);                                           \
            abort();                                                                               \
        }                                                                                          \
    } while (0)

#define GK_ASSERT(cond, ...)                                                        \
    do {                                                                            \
        if (!(cond)) {                                                              \
            fprintf(stderr, "libgpukern: assertion failed at %s:%d: %s: ", __FILE__, \
                    __LINE__, #cond);                                               \
            fprintf(stderr, __VA_ARGS__);                                           \
            fprintf(stderr, "\n");                                                  \
            abort();                                                                \
        }                                                                           \
    } while (0)

#define GK_LAUNCH_CHECK() GK_CHECK(hipGetLastError())

namespace gk {

// Stream every libgpukern launch goes to.  The reference entry points are
// synchronous (gpu_safe.h:41); the engine entry points enqueue on this stream
// and synchronise only where their contract says so.
hipStream_t stream();
void sync();  // hipStreamSynchronize(stream()) + error check

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// 1-D launch geometry for N elements, 256 threads (4 wave64s) per block.
constexpr int kBlock = 256;
inline dim3 grid1d(long n, int block = kBlock) {
    long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 2147483647L) g = 2147483647L;
    return dim3((unsigned)g);
}

// Twiddle tables exp(-2 pi i k / L), k < L, in float2 and double2, computed
// on the host in long double and uploaded once per length (cached).
const float2* twiddle_f32(int L);
const double2* twiddle_f64(int L);

// While set on the calling thread, a table lookup that would have to build
// (allocate) a missing table aborts with a message instead: the no-allocation
// entry points run with it set.
void tables_frozen(bool on);
bool tables_frozen();

// Device memory, pinned host memory and streams of the engine (simulations,
// convolution plans, kernel spectra and their set-up, batches): freed blocks
// go to a per-device cache and are reused by later requests of the same size
// class.  hipFree and hipHostFree wait for every stream of the device, so a
// simulation or batch freed while others run on their own streams (a sweep
// rebuilding a batch, the serving loop retiring a request) would stall all of
// them until their queued work drained.  The caller frees only blocks whose
// work has finished (the engine synchronises the owner's stream first).
// Classes: powers of two from 512 B to 1 MiB, then whole MiB.  A failed
// hipMalloc releases the device's cache and retries; the cache holds at most
// 1/16 of the device's memory (beyond that a free goes to hipFree).
void* cached_malloc(size_t bytes);
void cached_free(void* p);
void* cached_host_malloc(size_t bytes);
void cached_host_free(void* p);
hipStream_t cached_stream();             // a non-blocking stream of the current device
void cached_stream_release(hipStream_t s);  // after its work has finished
size_t cached_trim();                    // hipFree the current device's cached blocks (and destroy its idle
                                         // streams); bytes released
size_t cached_bytes();                   // bytes the current device's cache holds

}  // namespace gk
