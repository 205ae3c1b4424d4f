// Reference-named entry points of libgpukern: initialisation, memory, linear
// algebra, reductions, normalisation, torque, local fields, exchange stencil,
// k-space multiply, padding.  Each function cites the reference function it
// replaces; the kernels are written for wave64 CDNA4 (256-thread blocks,
// grid-stride loops, shuffle reductions), not translated from the CUDA code.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/gpukern.h"
#include "gk_common.h"

namespace gk {

static std::once_flag g_init_flag;
static hipStream_t g_stream = nullptr;
static int g_maxthreads = 0;
static int g_stride_cache = -1;
static std::atomic<unsigned long long> g_usedmem{0};
static std::mutex g_mem_mutex;
static std::map<void*, size_t> g_alloc_sizes;
static std::mutex g_tw_mutex;
static std::map<int, float2*> g_tw32;
static std::map<int, double2*> g_tw64;

static void init_once() { GK_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking)); }

static void ensure_init() { std::call_once(g_init_flag, init_once); }

hipStream_t stream() {
    ensure_init();
    return g_stream;
}
void sync() {
    GK_CHECK(hipGetLastError());
    GK_CHECK(hipStreamSynchronize(stream()));
}
// ---------------------------------------------------------------- block cache
namespace {
struct BlockCache {
    std::mutex mu;
    std::map<void*, std::pair<int, size_t>> live;           // device block -> (device, class)
    std::map<std::pair<int, size_t>, std::vector<void*>> free_dev;
    std::map<int, size_t> held, cap;                         // cached bytes and their limit per device
    std::map<void*, size_t> live_host;
    std::map<size_t, std::vector<void*>> free_host;
    std::map<int, std::vector<hipStream_t>> streams;  // idle, last freed on top
};
BlockCache& block_cache() {
    static BlockCache* c = new BlockCache;  // never destroyed: blocks may be freed during exit
    return *c;
}
size_t size_class(size_t bytes) {
    if (bytes <= (1u << 20)) {
        size_t c = 512;
        while (c < bytes) c <<= 1;
        return c;
    }
    return (bytes + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
}
int current_device() {
    int d = 0;
    GK_CHECK(hipGetDevice(&d));
    return d;
}
size_t trim_locked(BlockCache& c, int dev) {
    size_t n = 0;
    for (auto it = c.free_dev.begin(); it != c.free_dev.end();) {
        if (it->first.first != dev) {
            ++it;
            continue;
        }
        for (void* p : it->second) {
            GK_CHECK(hipFree(p));
            n += it->first.second;
        }
        it = c.free_dev.erase(it);
    }
    c.held[dev] = 0;
    return n;
}
}  // namespace

void* cached_malloc(size_t bytes) {
    BlockCache& c = block_cache();
    const int dev = current_device();
    const size_t cls = size_class(bytes > 0 ? bytes : 1);
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.free_dev.find({dev, cls});
    void* p = nullptr;
    if (it != c.free_dev.end() && !it->second.empty()) {
        p = it->second.back();
        it->second.pop_back();
        c.held[dev] -= cls;
    } else {
        hipError_t e = hipMalloc(&p, cls);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            trim_locked(c, dev);
            e = hipMalloc(&p, cls);
        }
        GK_CHECK(e);
    }
    c.live[p] = {dev, cls};
    return p;
}

void cached_free(void* p) {
    if (!p) return;
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.live.find(p);
    GK_ASSERT(it != c.live.end(), "cached_free of a block the cache did not give out");
    const int dev = it->second.first;
    const size_t cls = it->second.second;
    c.live.erase(it);
    if (!c.cap.count(dev)) {
        size_t fr = 0, tot = 0;
        GK_CHECK(hipMemGetInfo(&fr, &tot));
        c.cap[dev] = tot / 16;
    }
    if (c.held[dev] + cls > c.cap[dev]) {
        GK_CHECK(hipFree(p));
        return;
    }
    c.free_dev[{dev, cls}].push_back(p);
    c.held[dev] += cls;
}

void* cached_host_malloc(size_t bytes) {
    BlockCache& c = block_cache();
    const size_t cls = size_class(bytes > 0 ? bytes : 1);
    std::lock_guard<std::mutex> lk(c.mu);
    void* p = nullptr;
    auto& fl = c.free_host[cls];
    if (!fl.empty()) {
        p = fl.back();
        fl.pop_back();
    } else {
        GK_CHECK(hipHostMalloc(&p, cls, hipHostMallocDefault));
    }
    c.live_host[p] = cls;
    return p;
}

void cached_host_free(void* p) {
    if (!p) return;
    BlockCache& c = block_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto it = c.live_host.find(p);
    GK_ASSERT(it != c.live_host.end(), "cached_host_free of a block the cache did not give out");
    c.free_host[it->second].push_back(p);
    c.live_host.erase(it);
}

hipStream_t cached_stream() {
    BlockCache& c = block_cache();
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(c.mu);
    auto& v = c.streams[dev];
    if (!v.empty()) {
        hipStream_t s = v.back();
        v.pop_back();
        return s;
    }
    hipStream_t s = nullptr;
    GK_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

void cached_stream_release(hipStream_t s) {
    if (!s) return;
    BlockCache& c = block_cache();
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(c.mu);
    c.streams[dev].push_back(s);
}

size_t cached_trim() {
    BlockCache& c = block_cache();
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(c.mu);
    // idle streams too: the next batches create new ones.  A stream keeps the
    // hardware queue it got at creation (4 per process here), so which of a new
    // workload's concurrent batches share a queue depends on the streams it
    // inherits; starting from new ones gives it the assignment of a fresh process.
    for (hipStream_t st : c.streams[dev]) GK_CHECK(hipStreamDestroy(st));
    c.streams[dev].clear();
    return trim_locked(c, dev);
}

size_t cached_bytes() {
    BlockCache& c = block_cache();
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(c.mu);
    return c.held[dev];
}

static thread_local bool g_tables_frozen = false;
void tables_frozen(bool on) { g_tables_frozen = on; }
bool tables_frozen() { return g_tables_frozen; }

template <typename C, typename R>
static C* make_table(std::map<int, C*>& cache, int L) {
    std::lock_guard<std::mutex> lk(g_tw_mutex);
    auto it = cache.find(L);
    if (it != cache.end()) return it->second;
    GK_ASSERT(!g_tables_frozen, "twiddle table of length %d is missing in a call that may not allocate", L);
    std::vector<C> t(L);
    for (int k = 0; k < L; k++) {
        long double a = -2.0L * 3.141592653589793238462643383279L * (long double)k / (long double)L;
        t[k].x = (R)cosl(a);
        t[k].y = (R)sinl(a);
    }
    C* d = nullptr;
    GK_CHECK(hipMalloc(&d, (size_t)L * sizeof(C)));
    GK_CHECK(hipMemcpy(d, t.data(), (size_t)L * sizeof(C), hipMemcpyHostToDevice));
    cache[L] = d;
    return d;
}
const float2* twiddle_f32(int L) {
    ensure_init();
    return make_table<float2, float>(g_tw32, L);
}
const double2* twiddle_f64(int L) {
    ensure_init();
    return make_table<double2, double>(g_tw64, L);
}

// ---------------------------------------------------------------- kernels

__global__ void k_fill(float* __restrict__ a, float v, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_add(float* __restrict__ a, const float* __restrict__ b, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] += b[i];
}
__global__ void k_madd(float* __restrict__ a, float c, const float* __restrict__ b, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] += c * b[i];
}
__global__ void k_madd2(float* __restrict__ a, const float* __restrict__ b, const float* __restrict__ c, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] += b[i] * c[i];
}
__global__ void k_add_constant(float* __restrict__ a, float c, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] += c;
}
__global__ void k_lincomb(float* __restrict__ a, const float* __restrict__ b, float wa, float wb, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = wa * a[i] + wb * b[i];
}

struct LinCombMany {
    const float* v[16];
    float w[16];
};
__global__ void k_lincomb_many(float* __restrict__ r, LinCombMany lc, int nv, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float acc = r[i];
        for (int j = 0; j < nv; j++) acc += lc.w[j] * lc.v[j][i];
        r[i] = acc;
    }
}

__global__ void k_scale_dot(float* __restrict__ r, const float* __restrict__ a, const float* __restrict__ b, float s,
                            int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        r[i] = s * (a[i] * b[i] + a[n + i] * b[n + i] + a[2 * n + i] * b[2 * n + i]);
}

__global__ void k_normalize(float* __restrict__ m, const float* __restrict__ map, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float x = m[i], y = m[n + i], z = m[2 * n + i];
        float s = rsqrtf(x * x + y * y + z * z);
        if (map) s = map[i] == 0.f ? 0.f : s * map[i];
        m[i] = x * s;
        m[n + i] = y * s;
        m[2 * n + i] = z * s;
    }
}

// h <- dtg (m x h - alpha m x (m x h)), gpu_torque.cu:20-50
__global__ void k_deltaM(const float* __restrict__ m, float* __restrict__ h, float alpha_mul,
                         const float* __restrict__ alpha_map, float dtg, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float alpha = alpha_map ? alpha_mul * alpha_map[i] : alpha_mul;
        float Mx = m[i], My = m[n + i], Mz = m[2 * n + i];
        float Hx = h[i], Hy = h[n + i], Hz = h[2 * n + i];
        float cx = My * Hz - Mz * Hy, cy = Mz * Hx - Mx * Hz, cz = Mx * Hy - My * Hx;
        float dx = My * cz - Mz * cy, dy = Mz * cx - Mx * cz, dz = Mx * cy - My * cx;
        h[i] = dtg * (cx - alpha * dx);
        h[n + i] = dtg * (cy - alpha * dy);
        h[2 * n + i] = dtg * (cz - alpha * dz);
    }
}

struct LocalParams {
    float hext[3];
    int anisType;
    float k0, k1;
    float u[9];
};

// gpu_local_contr3.cu:21-229 (H) and 238-466 (H and phi), one kernel for all variants
template <bool PHI>
__global__ void k_local_fields(const float* __restrict__ m, float* __restrict__ h, float* __restrict__ phi,
                               const float* __restrict__ hmap, LocalParams p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float mx = m[i], my = m[n + i], mz = m[2 * n + i];
        float ex = p.hext[0], ey = p.hext[1], ez = p.hext[2];
        if (hmap) {
            ex *= hmap[i];
            ey *= hmap[n + i];
            ez *= hmap[2 * n + i];
        }
        float ax = 0.f, ay = 0.f, az = 0.f, e_anis = 0.f;
        if (p.anisType == 1) {
            float mu = mx * p.u[0] + my * p.u[1] + mz * p.u[2];
            ax = p.k0 * mu * p.u[0];
            ay = p.k0 * mu * p.u[1];
            az = p.k0 * mu * p.u[2];
            e_anis = 0.5f * p.k0 * mu * mu;
        } else if (p.anisType == 2) {
            float a0 = mx * p.u[0] + my * p.u[1] + mz * p.u[2];
            float a1 = mx * p.u[3] + my * p.u[4] + mz * p.u[5];
            float a2 = mx * p.u[6] + my * p.u[7] + mz * p.u[8];
            float a00 = a0 * a0, a11 = a1 * a1, a22 = a2 * a2;
            float d0 = p.k0 * (a11 + a22) * a0 + p.k1 * a0 * a11 * a22;
            float d1 = p.k0 * (a00 + a22) * a1 + p.k1 * a00 * a1 * a22;
            float d2 = p.k0 * (a00 + a11) * a2 + p.k1 * a00 * a11 * a2;
            ax = -d0 * p.u[0] - d1 * p.u[3] - d2 * p.u[6];
            ay = -d0 * p.u[1] - d1 * p.u[4] - d2 * p.u[7];
            az = -d0 * p.u[2] - d1 * p.u[5] - d2 * p.u[8];
            // phi -= m.Hext - m.cub/4 with H_anis = -cub (gpu_local_contr3.cu:308-315)
            e_anis = 0.25f * (mx * ax + my * ay + mz * az);
        }
        h[i] += ex + ax;
        h[n + i] += ey + ay;
        h[2 * n + i] += ez + az;
        if (PHI) phi[i] -= mx * ex + my * ey + mz * ez + e_anis;
    }
}

__global__ void k_lin_anis(float* hx, float* hy, float* hz, const float* mx, const float* my, const float* mz,
                           const float* kxx, const float* kyy, const float* kzz, const float* kyz, const float* kxz,
                           const float* kxy, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float x = mx[i], y = my[i], z = mz[i];
        hx[i] += kxx[i] * x + kxy[i] * y + kxz[i] * z;
        hy[i] += kxy[i] * x + kyy[i] * y + kyz[i] * z;
        hz[i] += kxz[i] * x + kyz[i] * y + kzz[i] * z;
    }
}

// ---------------------------------------------------------------- reductions
// Each block reduces one contiguous chunk; lanes first accumulate privately,
// then combine with wave64 xor-shuffles and a 4-entry LDS exchange.
template <int OP> struct RedOp;
template <> struct RedOp<REDUCE_ADD> {
    using acc_t = double;
    __device__ static acc_t init() { return 0.0; }
    __device__ static acc_t load(float x) { return (double)x; }
    __device__ static acc_t comb(acc_t a, acc_t b) { return a + b; }
};
template <> struct RedOp<REDUCE_SUMABS> {
    using acc_t = double;
    __device__ static acc_t init() { return 0.0; }
    __device__ static acc_t load(float x) { return (double)fabsf(x); }
    __device__ static acc_t comb(acc_t a, acc_t b) { return a + b; }
};
template <> struct RedOp<REDUCE_MAX> {
    using acc_t = float;
    __device__ static acc_t init() { return -INFINITY; }
    __device__ static acc_t load(float x) { return x; }
    __device__ static acc_t comb(acc_t a, acc_t b) { return fmaxf(a, b); }
};
template <> struct RedOp<REDUCE_MIN> {
    using acc_t = float;
    __device__ static acc_t init() { return INFINITY; }
    __device__ static acc_t load(float x) { return x; }
    __device__ static acc_t comb(acc_t a, acc_t b) { return fminf(a, b); }
};
template <> struct RedOp<REDUCE_MAXABS> {
    using acc_t = float;
    __device__ static acc_t init() { return 0.f; }
    __device__ static acc_t load(float x) { return fabsf(x); }
    __device__ static acc_t comb(acc_t a, acc_t b) { return fmaxf(a, b); }
};

template <int OP>
__global__ __launch_bounds__(256) void k_reduce_partial(const float* __restrict__ in, float* __restrict__ out, int n,
                                                        int chunk) {
    using R = RedOp<OP>;
    using A = typename R::acc_t;
    __shared__ A wave_acc[4];
    const long lo = (long)blockIdx.x * chunk;
    long hi = lo + chunk;
    if (hi > n) hi = n;
    A acc = R::init();
    for (long i = lo + threadIdx.x; i < hi; i += blockDim.x) acc = R::comb(acc, R::load(in[i]));
    for (int off = 32; off > 0; off >>= 1) acc = R::comb(acc, __shfl_xor(acc, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_acc[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        A r = wave_acc[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = R::comb(r, wave_acc[w]);
        out[blockIdx.x] = (float)r;
    }
}

// ---------------------------------------------------------------- exchange stencil
// 6-neighbour Laplacian (gpu_exch1.cu:21-301): h += c0 m + sum_d c_d m(nb),
// missing neighbours count as zero (free boundary) unless periodic.  Blocks
// are 64 (Z) x 8 (Y) cells; the current X-plane tile with its one-cell Y/Z
// halo is staged in LDS, the X-1 / X+1 values ride along in registers while
// the block marches through X.  Every tile slot (interior or halo) is filled
// through the same wrap-or-zero loader, so periodic wrap also works at a
// partially filled edge block.
constexpr int EX_TZ = 64, EX_TY = 8;

__device__ __forceinline__ int wrap_or_neg(int i, int n, int periodic) {
    if (i >= 0 && i < n) return i;
    if (!periodic) return -1;
    return i < 0 ? i + n : i - n;
}

__global__ __launch_bounds__(EX_TZ* EX_TY) void k_exch6(const float* __restrict__ m, float* __restrict__ h, int n0,
                                                         int n1, int n2, int p0, int p1, int p2, float cx, float cy,
                                                         float cz, float c0) {
    __shared__ float tile[EX_TY + 2][EX_TZ + 2];
    const int tz = threadIdx.x, ty = threadIdx.y;
    const int z0 = blockIdx.x * EX_TZ, y0 = blockIdx.y * EX_TY;
    const int z = z0 + tz, y = y0 + ty;
    const bool inside = z < n2 && y < n1;
    const long plane = (long)n1 * n2;
    auto ld = [&](int x, int yy, int zz) -> float {
        int xw = wrap_or_neg(x, n0, p0), yw = wrap_or_neg(yy, n1, p1), zw = wrap_or_neg(zz, n2, p2);
        return (xw >= 0 && yw >= 0 && zw >= 0) ? m[(long)xw * plane + (long)yw * n2 + zw] : 0.f;
    };
    float below = (n0 > 1) ? ld(-1, y, z) : 0.f;
    float cur = ld(0, y, z);
    for (int x = 0; x < n0; x++) {
        float above = (n0 > 1) ? ld(x + 1, y, z) : 0.f;
        tile[ty + 1][tz + 1] = cur;
        if (tz == 0) tile[ty + 1][0] = ld(x, y, z0 - 1);
        if (tz == 1) tile[ty + 1][EX_TZ + 1] = ld(x, y, z0 + EX_TZ);
        if (ty == 0) tile[0][tz + 1] = ld(x, y0 - 1, z);
        if (ty == 1) tile[EX_TY + 1][tz + 1] = ld(x, y0 + EX_TY, z);
        __syncthreads();
        if (inside) {
            float r = c0 * cur + cx * (below + above) + cy * (tile[ty][tz + 1] + tile[ty + 2][tz + 1]) +
                      cz * (tile[ty + 1][tz] + tile[ty + 1][tz + 2]);
            h[(long)x * plane + (long)y * n2 + z] += r;
        }
        __syncthreads();
        below = cur;
        cur = above;
    }
}

// ---------------------------------------------------------------- k-space multiply
__global__ void k_kernelmul6(float2* __restrict__ mx, float2* __restrict__ my, float2* __restrict__ mz,
                             const float* __restrict__ kxx, const float* __restrict__ kyy,
                             const float* __restrict__ kzz, const float* __restrict__ kyz,
                             const float* __restrict__ kxz, const float* __restrict__ kxy, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float2 a = mx[i], b = my[i], c = mz[i];
        float xx = kxx[i], yy = kyy[i], zz = kzz[i], yz = kyz[i], xz = kxz[i], xy = kxy[i];
        mx[i] = make_float2(a.x * xx + b.x * xy + c.x * xz, a.y * xx + b.y * xy + c.y * xz);
        my[i] = make_float2(a.x * xy + b.x * yy + c.x * yz, a.y * xy + b.y * yy + c.y * yz);
        mz[i] = make_float2(a.x * xz + b.x * yz + c.x * zz, a.y * xz + b.y * yz + c.y * zz);
    }
}
__global__ void k_kernelmul4(float2* __restrict__ mx, float2* __restrict__ my, float2* __restrict__ mz,
                             const float* __restrict__ kxx, const float* __restrict__ kyy,
                             const float* __restrict__ kzz, const float* __restrict__ kyz, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float2 a = mx[i], b = my[i], c = mz[i];
        float xx = kxx[i], yy = kyy[i], zz = kzz[i], yz = kyz[i];
        mx[i] = make_float2(a.x * xx, a.y * xx);
        my[i] = make_float2(b.x * yy + c.x * yz, b.y * yy + c.y * yz);
        mz[i] = make_float2(b.x * yz + c.x * zz, b.y * yz + c.y * zz);
    }
}

// ---------------------------------------------------------------- padding
// Copies the overlap [0,min(S0,D0)) x [0,min(S1,D1)) x [0,min(S2,D2)) in one launch.
__global__ void k_copy_region(const float* __restrict__ src, float* __restrict__ dst, int S1, int S2, int D1, int D2,
                              int c0, int c1, int c2) {
    long total = (long)c0 * c1 * c2;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        int k = (int)(t % c2);
        long r = t / c2;
        int j = (int)(r % c1);
        int i = (int)(r / c1);
        dst[((long)i * D1 + j) * D2 + k] = src[((long)i * S1 + j) * S2 + k];
    }
}

// ---------------------------------------------------------------- spin-transfer torque
// Zhang-Li torque (gpu_spintorque.cu:24-180): Hspin = (u . grad) m by central
// differences, one-sided at the box edges (the edge cell stands in for its
// missing neighbour on every side); then
// h <- dtg (m x (H + eps Hspin) - m x (m x (alpha H + beta Hspin))).
__global__ void k_spintorque(const float* __restrict__ m, float* __restrict__ h, float alpha, float beta, float eps,
                             float ux, float uy, float uz, const float* __restrict__ jmap, float dtg, int n0, int n1,
                             int n2) {
    const long N = (long)n0 * n1 * n2;
    for (long I = blockIdx.x * (long)blockDim.x + threadIdx.x; I < N; I += (long)gridDim.x * blockDim.x) {
        const int k = (int)(I % n2);
        const long r = I / n2;
        const int j = (int)(r % n1), i = (int)(r / n1);
        float u0 = ux, u1 = uy, u2 = uz;
        if (jmap) {
            u0 *= jmap[I];
            u1 *= jmap[N + I];
            u2 *= jmap[2 * N + I];
        }
        const long im = i > 0 ? I - (long)n1 * n2 : I, ip = i + 1 < n0 ? I + (long)n1 * n2 : I;
        const long jm = j > 0 ? I - n2 : I, jp = j + 1 < n1 ? I + n2 : I;
        const long km = k > 0 ? I - 1 : I, kp = k + 1 < n2 ? I + 1 : I;
        float d[3];
        for (int c = 0; c < 3; c++) {
            const float* mc = m + c * N;
            d[c] = u0 * (mc[ip] - mc[im]) + u1 * (mc[jp] - mc[jm]) + u2 * (mc[kp] - mc[km]);
        }
        const float Mx = m[I], My = m[N + I], Mz = m[2 * N + I];
        const float Hx = h[I], Hy = h[N + I], Hz = h[2 * N + I];
        const float px = Hx + eps * d[0], py = Hy + eps * d[1], pz = Hz + eps * d[2];
        const float qx = alpha * Hx + beta * d[0], qy = alpha * Hy + beta * d[1], qz = alpha * Hz + beta * d[2];
        const float ax = My * pz - Mz * py, ay = Mz * px - Mx * pz, az = Mx * py - My * px;  // m x Hp
        const float cx = My * qz - Mz * qy, cy = Mz * qx - Mx * qz, cz = Mx * qy - My * qx;  // m x Hd
        const float ex = My * cz - Mz * cy, ey = Mz * cx - Mx * cz, ez = Mx * cy - My * cx;  // m x (m x Hd)
        h[I] = dtg * (ax - ex);
        h[N + I] = dtg * (ay - ey);
        h[2 * N + I] = dtg * (az - ez);
    }
}

// ---------------------------------------------------------------- semi-analytical step
// Exact solution of the LL equation for a constant field over dt
// (gpu_anal.cu:23-84): rotate m into a frame whose first axis is h, advance
// the polar angle (damping) and the azimuth (precession) in closed form,
// rotate back.  Cells with m = 0 are left untouched.
__global__ void k_anal_fw(const float* __restrict__ min, float* __restrict__ mout, const float* __restrict__ h,
                          float dt, float alpha, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float mx = min[i], my = min[n + i], mz = min[2 * n + i];
        if (mx == 0.f && my == 0.f && mz == 0.f) continue;
        const float hx = h[i], hy = h[n + i], hz = h[2 * n + i];
        // orthonormal frame e1 = h/|h|, e2, e3 (rows of R)
        float r0, r1, r2, r3, r4, r5, r6, r8, inv_h;
        if (hx == 0.f && hy == 0.f) {
            r0 = 0.f; r1 = 0.f; r2 = -1.f;
            r3 = 0.f; r4 = 1.f; r5 = 0.f;
            r6 = 1.f; r8 = 0.f;
            inv_h = 1.f / hz;
        } else {
            const float t = hx * hx + hy * hy;
            const float inv_xy = rsqrtf(t);
            inv_h = rsqrtf(t + hz * hz);
            r0 = hx * inv_h;
            r1 = -hy * inv_xy;
            r2 = -r0 * hz * inv_xy;
            r3 = hy * inv_h;
            r4 = hx * inv_xy;
            r5 = r1 * hz * inv_h;
            r6 = hz * inv_h;
            r8 = inv_h / inv_xy;
        }
        const float a = mx * r0 + my * r3 + mz * r6;
        const float b = mx * r1 + my * r4;
        const float c = mx * r2 + my * r5 + mz * r8;
        const float qt = dt / (1.f + alpha * alpha);
        const float e = expf(alpha * qt / inv_h);
        float sn, cs;
        sincosf(qt / inv_h, &sn, &cs);
        const float den = e * (1.f + a) + (1.f - a) / e;
        const float a2 = (e * (1.f + a) - (1.f - a) / e) / den;
        const float b2 = 2.f * (b * cs - c * sn) / den;
        const float c2 = 2.f * (b * sn + c * cs) / den;
        mout[i] = a2 * r0 + b2 * r1 + c2 * r2;
        mout[n + i] = a2 * r3 + b2 * r4 + c2 * r5;
        mout[2 * n + i] = a2 * r6 + c2 * r8;
    }
}

// ---------------------------------------------------------------- noise
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__global__ void k_gaussian(float* __restrict__ d, float mean, float sd, int n, uint64_t seed) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t a = splitmix64(seed ^ (2ull * i)), b = splitmix64(seed ^ (2ull * i + 1));
        double u1 = ((a >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        double u2 = (b >> 11) * (1.0 / 9007199254740992.0);
        double g = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
        d[i] = (float)(mean + sd * g);
    }
}

}  // namespace gk

using namespace gk;

static dim3 g1(long n) {
    long g = (n + kBlock - 1) / kBlock;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return dim3((unsigned)g);
}

extern "C" {

// ---------------------------------------------------------------- init (gpu_init.cu, gpu_conf.cu)
void gpu_init(int threads, int options) {
    (void)options;
    if (threads > 0) g_maxthreads = threads;
    ensure_init();
}

void gpu_set_device(int devid) {
    int n = 0;
    GK_CHECK(hipGetDeviceCount(&n));
    GK_ASSERT(devid >= 0 && devid < n, "device %d out of range (%d devices)", devid, n);
    GK_CHECK(hipSetDevice(devid));
}

int gpu_maxthreads(void) {
    if (g_maxthreads <= 0) {
        hipDeviceProp_t prop;
        int dev = 0;
        GK_CHECK(hipGetDevice(&dev));
        GK_CHECK(hipGetDeviceProperties(&prop, dev));
        g_maxthreads = prop.maxThreadsPerBlock;
    }
    return g_maxthreads;
}

void gpu_print_properties_stdout(void) {
    hipDeviceProp_t p;
    int dev = 0;
    GK_CHECK(hipGetDevice(&dev));
    GK_CHECK(hipGetDeviceProperties(&p, dev));
    printf("Device number: %d\n", dev);
    printf("Device name: %s (%s)\n", p.name, p.gcnArchName);
    printf("Global Memory: %zu MB\n", (size_t)(p.totalGlobalMem >> 20));
    printf("Shared Memory: %zu kB/block\n", (size_t)(p.sharedMemPerBlock >> 10));
    printf("Registers: %d per block\n", p.regsPerBlock);
    printf("Warp size: %d threads\n", p.warpSize);
    printf("Max threads: %d per block\n", p.maxThreadsPerBlock);
    printf("Max block size: %d x %d x %d threads\n", p.maxThreadsDim[0], p.maxThreadsDim[1], p.maxThreadsDim[2]);
    printf("Max grid size: %d x %d x %d blocks\n", p.maxGridSize[0], p.maxGridSize[1], p.maxGridSize[2]);
    printf("Clock rate: %d MHz\n", p.clockRate / 1000);
    printf("Multiprocessors: %d\n", p.multiProcessorCount);
    fflush(stdout);
}

// ---------------------------------------------------------------- memory (gpu_mem.cu:28-216)
float* new_gpu_array(int size) {
    GK_ASSERT(size > 0, "new_gpu_array: size %d", size);
    ensure_init();
    float* p = nullptr;
    GK_CHECK(hipMalloc((void**)&p, (size_t)size * sizeof(float)));
    GK_ASSERT(p != nullptr, "hipMalloc returned NULL");
    GK_CHECK(hipMemsetAsync(p, 0, (size_t)size * sizeof(float), stream()));
    sync();
    {
        std::lock_guard<std::mutex> lk(g_mem_mutex);
        g_alloc_sizes[p] = (size_t)size * sizeof(float);
    }
    g_usedmem += (unsigned long long)size * sizeof(float);
    return p;
}

void free_gpu_array(float* ptr) {
    if (!ptr) return;
    {
        std::lock_guard<std::mutex> lk(g_mem_mutex);
        auto it = g_alloc_sizes.find(ptr);
        if (it != g_alloc_sizes.end()) {
            g_usedmem -= it->second;
            g_alloc_sizes.erase(it);
        }
    }
    sync();
    GK_CHECK(hipFree(ptr));
}

unsigned long long int gpu_usedmem(void) { return g_usedmem.load(); }

int gpu_stride_float(void) {
    if (g_stride_cache == -1) {
        // hipMallocPitch of a 1x1 array gives the device's preferred row pitch
        void* p = nullptr;
        size_t pitch = 0;
        GK_CHECK(hipMallocPitch(&p, &pitch, sizeof(float), 1));
        GK_CHECK(hipFree(p));
        g_stride_cache = (int)(pitch / sizeof(float));
    }
    return g_stride_cache;
}

void gpu_override_stride(int nFloats) {
    GK_ASSERT(nFloats > -2, "gpu_override_stride(%d)", nFloats);
    g_stride_cache = nFloats;
}

void memcpy_to_gpu(float* source, float* dest, int n) {
    GK_ASSERT(n > 0, "memcpy_to_gpu: n=%d", n);
    GK_CHECK(hipMemcpyAsync(dest, source, (size_t)n * sizeof(float), hipMemcpyHostToDevice, stream()));
    sync();
}
void memcpy_from_gpu(float* source, float* dest, int n) {
    GK_ASSERT(n > 0, "memcpy_from_gpu: n=%d", n);
    GK_CHECK(hipMemcpyAsync(dest, source, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, stream()));
    sync();
}
void memcpy_on_gpu(float* source, float* dest, int n) {
    GK_ASSERT(n > 0, "memcpy_on_gpu: n=%d", n);
    GK_CHECK(hipMemcpyAsync(dest, source, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, stream()));
    sync();
}
void memcpy_gpu_dir(float* source, float* dest, int n, int direction) {
    switch (direction) {
    case DIR_TO: memcpy_to_gpu(source, dest, n); break;
    case DIR_ON: memcpy_on_gpu(source, dest, n); break;
    case DIR_FROM: memcpy_from_gpu(source, dest, n); break;
    default: fprintf(stderr, "illegal memcpy direction: %d\n", direction); abort();
    }
}
float gpu_array_get(float* p, int index) {
    float r = 666.0f;
    memcpy_from_gpu(p + index, &r, 1);
    return r;
}
void gpu_array_set(float* p, int index, float value) { memcpy_to_gpu(&value, p + index, 1); }

void gpu_zero(float* data, int n) {
    GK_ASSERT(n >= 0, "gpu_zero: n=%d", n);
    if (n == 0) return;
    k_fill<<<g1(n), kBlock, 0, stream()>>>(data, 0.f, n);
    sync();
}

// ---------------------------------------------------------------- linalg (gpu_linalg.cu:20-140)
void gpu_add(float* a, float* b, int N) {
    k_add<<<g1(N), kBlock, 0, stream()>>>(a, b, N);
    sync();
}
void gpu_madd(float* a, float c, float* b, int N) {
    k_madd<<<g1(N), kBlock, 0, stream()>>>(a, c, b, N);
    sync();
}
void gpu_madd2(float* a, float* b, float* c, int N) {
    k_madd2<<<g1(N), kBlock, 0, stream()>>>(a, b, c, N);
    sync();
}
void gpu_add_constant(float* a, float c, int N) {
    k_add_constant<<<g1(N), kBlock, 0, stream()>>>(a, c, N);
    sync();
}
void gpu_linear_combination(float* a, float* b, float wa, float wb, int N) {
    k_lincomb<<<g1(N), kBlock, 0, stream()>>>(a, b, wa, wb, N);
    sync();
}
void gpu_linear_combination_many(float* result, float** vectors, float* weights, int nv, int N) {
    for (int base = 0; base < nv; base += 16) {
        LinCombMany lc;
        int cnt = nv - base < 16 ? nv - base : 16;
        for (int j = 0; j < cnt; j++) {
            lc.v[j] = vectors[base + j];
            lc.w[j] = weights[base + j];
        }
        k_lincomb_many<<<g1(N), kBlock, 0, stream()>>>(result, lc, cnt, N);
    }
    sync();
}
void gpu_scale_dot_product(float* r, float* a, float* b, float s, int N) {
    k_scale_dot<<<g1(N), kBlock, 0, stream()>>>(r, a, b, s, N);
    sync();
}

// ---------------------------------------------------------------- reduce (gpu_reduction.cu:532-630)
float gpu_reduce(int op, float* input, float* devbuf, float* hostbuf, int blocks, int threads, int N) {
    GK_ASSERT(blocks > 0 && N > 0, "gpu_reduce: blocks=%d N=%d", blocks, N);
    GK_ASSERT((long)blocks * threads * 2 >= N, "gpu_reduce: blocks*threads*2=%ld < N=%d", (long)blocks * threads * 2,
              N);
    int chunk = div_up(N, blocks);
    switch (op) {
    case REDUCE_ADD: k_reduce_partial<REDUCE_ADD><<<blocks, 256, 0, stream()>>>(input, devbuf, N, chunk); break;
    case REDUCE_MAX: k_reduce_partial<REDUCE_MAX><<<blocks, 256, 0, stream()>>>(input, devbuf, N, chunk); break;
    case REDUCE_MAXABS: k_reduce_partial<REDUCE_MAXABS><<<blocks, 256, 0, stream()>>>(input, devbuf, N, chunk); break;
    case REDUCE_MIN: k_reduce_partial<REDUCE_MIN><<<blocks, 256, 0, stream()>>>(input, devbuf, N, chunk); break;
    case REDUCE_SUMABS: k_reduce_partial<REDUCE_SUMABS><<<blocks, 256, 0, stream()>>>(input, devbuf, N, chunk); break;
    default: fprintf(stderr, "gpu_reduce: illegal operation %d\n", op); abort();
    }
    GK_LAUNCH_CHECK();
    memcpy_from_gpu(devbuf, hostbuf, blocks);
    switch (op) {
    case REDUCE_ADD:
    case REDUCE_SUMABS: {
        double s = 0;
        for (int i = 0; i < blocks; i++) s += hostbuf[i];
        return (float)s;
    }
    case REDUCE_MAX: {
        float r = hostbuf[0];
        for (int i = 1; i < blocks; i++) r = fmaxf(r, hostbuf[i]);
        return r;
    }
    case REDUCE_MIN: {
        float r = hostbuf[0];
        for (int i = 1; i < blocks; i++) r = fminf(r, hostbuf[i]);
        return r;
    }
    default: {
        float r = 0.f;
        for (int i = 0; i < blocks; i++) r = fmaxf(r, hostbuf[i]);
        return r;
    }
    }
}

// ---------------------------------------------------------------- normalize / torque
void gpu_normalize_uniform(float* m, int N) {
    k_normalize<<<g1(N), kBlock, 0, stream()>>>(m, nullptr, N);
    sync();
}
void gpu_normalize_map(float* m, float* map, int N) {
    k_normalize<<<g1(N), kBlock, 0, stream()>>>(m, map, N);
    sync();
}
void gpu_deltaM(float* m, float* h, float alpha_mul, float* alpha_map, float dt_gilb, int N) {
    k_deltaM<<<g1(N), kBlock, 0, stream()>>>(m, h, alpha_mul, alpha_map, dt_gilb, N);
    sync();
}

// ---------------------------------------------------------------- local fields
static LocalParams local_params(const float* Hext, int anisType, const float* anisK, const float* anisAxes) {
    LocalParams p{};
    for (int i = 0; i < 3; i++) p.hext[i] = Hext[i];
    p.anisType = anisType;
    GK_ASSERT(anisType >= 0 && anisType <= 2, "gpu_add_local_fields: anisType %d", anisType);
    if (anisType == 1) {
        p.k0 = anisK[0];
        for (int i = 0; i < 3; i++) p.u[i] = anisAxes[i];
    } else if (anisType == 2) {
        p.k0 = anisK[0];
        p.k1 = anisK[1];
        for (int i = 0; i < 9; i++) p.u[i] = anisAxes[i];
    }
    return p;
}
void gpu_add_local_fields(float* m, float* h, int N, float* Hext, float* hmap, int anisType, float* anisK,
                          float* anisAxes) {
    LocalParams p = local_params(Hext, anisType, anisK, anisAxes);
    k_local_fields<false><<<g1(N), kBlock, 0, stream()>>>(m, h, nullptr, hmap, p, N);
    sync();
}
void gpu_add_local_fields_H_and_phi(float* m, float* h, float* phi, int N, float* Hext, float* hmap, int anisType,
                                    float* anisK, float* anisAxes) {
    LocalParams p = local_params(Hext, anisType, anisK, anisAxes);
    k_local_fields<true><<<g1(N), kBlock, 0, stream()>>>(m, h, phi, hmap, p, N);
    sync();
}
void gpu_add_lin_anisotropy(float* hx, float* hy, float* hz, float* mx, float* my, float* mz, float* kxx, float* kyy,
                            float* kzz, float* kyz, float* kxz, float* kxy, int N) {
    k_lin_anis<<<g1(N), kBlock, 0, stream()>>>(hx, hy, hz, mx, my, mz, kxx, kyy, kzz, kyz, kxz, kxy, N);
    sync();
}

// ---------------------------------------------------------------- exchange (gpu_exch1.cu:20-55,186-205,282-301)
void gpu_add_exch(float* m, float* h, int* size, int* periodic, int* exchInConv, float* cellSize, int type) {
    if (exchInConv[0] && exchInConv[1] && exchInConv[2]) return;
    if (type != EXCH_6NGBR) {
        fprintf(stderr, "libgpukern: gpu_add_exch: exchange type %d is not supported by this build (only %d)\n", type,
                EXCH_6NGBR);
        abort();
    }
    int n0 = size[0], n1 = size[1], n2 = size[2];
    float cx = 1.0f / cellSize[0] / cellSize[0];
    float cy = 1.0f / cellSize[1] / cellSize[1];
    float cz = 1.0f / cellSize[2] / cellSize[2];
    float c0 = n0 == 1 ? -2.0f / cellSize[1] / cellSize[1] - 2.0f / cellSize[2] / cellSize[2]
                       : -2.0f / cellSize[0] / cellSize[0] - 2.0f / cellSize[1] / cellSize[1] -
                             2.0f / cellSize[2] / cellSize[2];
    long N = (long)n0 * n1 * n2;
    dim3 grid(div_up(n2, EX_TZ), div_up(n1, EX_TY));
    dim3 block(EX_TZ, EX_TY);
    for (int c = 0; c < 3; c++) {
        if (exchInConv[c]) continue;
        k_exch6<<<grid, block, 0, stream()>>>(m + c * N, h + c * N, n0, n1, n2, periodic[0], periodic[1],
                                              periodic[2], cx, cy, cz, c0);
    }
    sync();
}

// ---------------------------------------------------------------- kernel multiply (gpu_kernmul.cu:25-141)
void gpu_kernelmul6(float* mx, float* my, float* mz, float* kxx, float* kyy, float* kzz, float* kyz, float* kxz,
                    float* kxy, int nReal) {
    GK_ASSERT(nReal > 0 && nReal % 2 == 0, "gpu_kernelmul6: nRealNumbers=%d", nReal);
    int n = nReal / 2;
    k_kernelmul6<<<g1(n), kBlock, 0, stream()>>>((float2*)mx, (float2*)my, (float2*)mz, kxx, kyy, kzz, kyz, kxz, kxy,
                                                 n);
    sync();
}
void gpu_kernelmul4(float* mx, float* my, float* mz, float* kxx, float* kyy, float* kzz, float* kyz, int nReal) {
    GK_ASSERT(nReal > 0 && nReal % 2 == 0, "gpu_kernelmul4: nRealNumbers=%d", nReal);
    int n = nReal / 2;
    k_kernelmul4<<<g1(n), kBlock, 0, stream()>>>((float2*)mx, (float2*)my, (float2*)mz, kxx, kyy, kzz, kyz, n);
    sync();
}

// ---------------------------------------------------------------- padding (gpu_zeropad.cu:83-107)
// The reference's unpad loops over the source planes and writes past the end of
// the destination when S0 > D0; here only the overlap is copied.
void gpu_copy_pad(float* src, float* dst, int S0, int S1, int S2, int D0, int D1, int D2) {
    GK_ASSERT(S0 <= D0 && S1 <= D1 && S2 <= D2, "gpu_copy_pad: source larger than destination");
    k_copy_region<<<g1((long)S0 * S1 * S2), kBlock, 0, stream()>>>(src, dst, S1, S2, D1, D2, S0, S1, S2);
    sync();
}
void gpu_copy_unpad(float* src, float* dst, int S0, int S1, int S2, int D0, int D1, int D2) {
    GK_ASSERT(S0 >= D0 && S1 >= D1 && S2 >= D2, "gpu_copy_unpad: destination larger than source");
    k_copy_region<<<g1((long)D0 * D1 * D2), kBlock, 0, stream()>>>(src, dst, S1, S2, D1, D2, D0, D1, D2);
    sync();
}

// ---------------------------------------------------------------- STT / semi-analytical (gpu_spintorque.cu:188-211, gpu_anal.cu:87-98)
void gpu_spintorque_deltaM(float* m, float* h, float alpha, float beta, float epsillon, float* u, float* jmap,
                           float dt_gilb, int N0, int N1, int N2) {
    long N = (long)N0 * N1 * N2;
    k_spintorque<<<g1(N), kBlock, 0, stream()>>>(m, h, alpha, beta, epsillon, u[0], u[1], u[2], jmap, dt_gilb, N0, N1,
                                                 N2);
    sync();
}

void gpu_anal_fw_step(float dt, float alpha, int N, float* m_in, float* m_out, float* h) {
    k_anal_fw<<<g1(N), kBlock, 0, stream()>>>(m_in, m_out, h, dt, alpha, N);
    sync();
}

// ---------------------------------------------------------------- noise (gpu_temperature.cu:31-41)
void gpu_gaussian_noise(float* d, float mean, float sd, int n) {
    static uint64_t counter = 0;
    uint64_t seed = 0x1234567ull + (counter++) * 0x9E3779B97F4A7C15ull;
    k_gaussian<<<g1(n), kBlock, 0, stream()>>>(d, mean, sd, n, seed);
    sync();
}

}  // extern "C"
