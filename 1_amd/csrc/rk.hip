// Runge-Kutta stage kernels and device-side step control for the engine.
//
// Restates rk.go:279-445 (stage loop, dt control at stage 0, error estimate,
// step factor (maxErr/err)^(1/order) clamped to [0.01, 2], reject when
// err > 2 maxErr, at most 10 trials, normalise at the end of the step),
// field.go:50-61 (applied field at the stage time), gpu_local_contr3.cu
// (Zeeman + uniaxial anisotropy) and gpu_torque.cu:20-50 (LL torque).
//
// All scalar control lives in EngineState on the device and is updated by
// single-thread kernels, so a run is a stream of launches with no host
// round trip; reductions are wave64 shuffles + one atomicMax per block on the
// float bit pattern (non-negative floats order like unsigned ints).
#include <hip/hip_runtime.h>

#include <cmath>

#include "engine.h"
#include "rk_dev.h"

namespace gk {

// k = dtg (m x h - alpha m x (m x h)) with h = H_conv + B_ext(t_i) + anisotropy (+ exchange)
__global__ __launch_bounds__(256) void k_torque(const float* __restrict__ m, float* __restrict__ hk,
                                                EngineState* st, TorqueArgs a, int N) {
    __shared__ float hext_s[3];
    if (!st->trial_active) return;
    if (threadIdx.x == 0) stage_field(a, st, hext_s);
    __syncthreads();
    const float hext[3] = {hext_s[0], hext_s[1], hext_s[2]};
    float lmax = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        float k[3], h[3] = {hk[i], hk[N + i], hk[2 * N + i]};
        add_thermal(a, st, i, h);
        llg_torque_cell(a, cell_alpha(a, i), hext, m[i], m[N + i], m[2 * N + i], h[0], h[1], h[2], k);
        hk[i] = k[0];
        hk[N + i] = k[1];
        hk[2 * N + i] = k[2];
        lmax = fmaxf(lmax, fmaxf(abs_keep_nan(k[0]), fmaxf(abs_keep_nan(k[1]), abs_keep_nan(k[2]))));
    }
    if (a.part) {
        const float bm = block_max_nonneg(lmax);
        if (threadIdx.x == 0) a.part[blockIdx.x] = bm;
    }
}

// h += the thermal field of the current step (calcHeff's addThermalField, field.go:74-80)
__global__ void k_add_thermal(float* __restrict__ h, const EngineState* st, TorqueArgs a, int N) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        float t[3] = {0.f, 0.f, 0.f};
        add_thermal(a, st, i, t);
        h[i] += t[0];
        h[N + i] += t[1];
        h[2 * N + i] += t[2];
    }
}

// m_high and error partials (adaptive) or m = normalise(m_high) (fixed step)
__global__ __launch_bounds__(256) void k_rk_final(float* __restrict__ m, KPtrs kp, float* __restrict__ mtmp,
                                                  EngineState* st, FinalArgs a, int N) {
    if (!st->trial_active) return;
    rk_final_body(m, kp, mtmp, st, a, N, blockIdx.x, gridDim.x);
}

__global__ void k_copy_if_active(float* __restrict__ dst, const float* __restrict__ src, long n,
                                 const EngineState* st) {
    if (!st->trial_active) return;
    copy_body(dst, src, n, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(256) void k_commit(float* __restrict__ m, const float* __restrict__ mtmp,
                                                const EngineState* st, const float* __restrict__ nmap, int N) {
    commit_body(m, mtmp, st, nmap, N, blockIdx.x, gridDim.x);
}

__global__ void k_normalize_active(float* __restrict__ m, const EngineState* st, const float* __restrict__ nmap,
                                   int N) {
    if (st && !st->trial_active) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const float x = m[i], y = m[N + i], z = m[2 * N + i];
        const float s = norm_scale(x, y, z, nmap, i);
        m[i] = x * s;
        m[N + i] = y * s;
        m[2 * N + i] = z * s;
    }
}

// ---------------------------------------------------------------- control (one thread / one wave)
__global__ void k_ctrl_begin(EngineState* st) { ctrl_begin_body(st); }
__global__ void k_ctrl_after_k0(EngineState* st, int fsal_copy, CtrlParams p) { ctrl_after_k0_body(st, fsal_copy, p); }
__global__ void k_ctrl_end(EngineState* st, CtrlParams p) { ctrl_end_body(st, p); }

// ---------------------------------------------------------------- FP64 output reductions
__global__ __launch_bounds__(256) void k_sum_comp(const float* __restrict__ v, int N, int ncomp,
                                                  double* __restrict__ out) {
    __shared__ double wsum[4];
    for (int c = 0; c < ncomp; c++) {
        double acc = 0.0;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) acc += v[c * (long)N + i];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) atomicAdd(&out[c], wsum[0] + wsum[1] + wsum[2] + wsum[3]);
        __syncthreads();
    }
}

static dim3 gridN(long n) {
    long g = (n + 255) / 256;
    if (g > kRedBlocks) g = kRedBlocks;
    if (g < 1) g = 1;
    return dim3((unsigned)g);
}

// energy density phi = -1/2 m.H_conv - m.B_ext - 1/2 k (m.u)^2 summed in FP64
// (field.go:17-20,64-71; gpu_local_contr3.cu:238-256)
__global__ __launch_bounds__(256) void k_energy(const float* __restrict__ m, const float* __restrict__ h, float ex,
                                                float ey, float ez, int anisType, float kAnis, float u0, float u1,
                                                float u2, int N, double* __restrict__ out) {
    __shared__ double wsum[4];
    double acc = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double mx = m[i], my = m[N + i], mz = m[2 * N + i];
        double phi = -0.5 * (mx * h[i] + my * h[N + i] + mz * h[2 * N + i]);
        phi -= mx * ex + my * ey + mz * ez;
        if (anisType == 1) {
            const double mu = mx * u0 + my * u1 + mz * u2;
            phi -= 0.5 * kAnis * mu * mu;
        }
        acc += phi;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

__global__ void k_set_target(EngineState* st, long long nsteps, double stop) {
    st->steps_target = nsteps < 0 ? 0x7fffffffffffffffLL : st->steps + nsteps;
    st->stop = stop;
    st->done = (st->time >= st->stop || st->steps >= st->steps_target) ? 1 : 0;
    st->nan_dt = 0;
}

// ---------------------------------------------------------------- launchers
void launch_energy(const float* m, const float* h, const float hext[3], int anisType, float kAnis, const float u[3],
                   int N, double* out, hipStream_t s) {
    k_energy<<<gridN(N), 256, 0, s>>>(m, h, hext[0], hext[1], hext[2], anisType, kAnis, u[0], u[1], u[2], N, out);
    GK_LAUNCH_CHECK();
}
__global__ void k_set_active(EngineState* st) { st->trial_active = 1; }
__global__ void k_set_relax(EngineState* st, int on) {
    st->relax = on;
    st->relax_prev = INFINITY;
}
void launch_set_relax(EngineState* st, int on, hipStream_t s) {
    k_set_relax<<<1, 1, 0, s>>>(st, on);
    GK_LAUNCH_CHECK();
}
void launch_set_active(EngineState* st, hipStream_t s) {
    k_set_active<<<1, 1, 0, s>>>(st);
    GK_LAUNCH_CHECK();
}
void launch_set_target(EngineState* st, long long nsteps, double stop, hipStream_t s) {
    k_set_target<<<1, 1, 0, s>>>(st, nsteps, stop);
    GK_LAUNCH_CHECK();
}


void launch_add_thermal(float* h, const EngineState* st, const TorqueArgs& a, int N, hipStream_t s) {
    if (a.thermalK == 0.f) return;
    k_add_thermal<<<gridN(N), 256, 0, s>>>(h, st, a, N);
    GK_LAUNCH_CHECK();
}

void launch_torque(const float* mi, float* hk, EngineState* st, const TorqueArgs& a, int N, hipStream_t s) {
    // a partial buffer has kRedBlocks entries: launch exactly that many blocks when reducing
    k_torque<<<a.part ? dim3(kRedBlocks) : gridN(N), 256, 0, s>>>(mi, hk, st, a, N);
    GK_LAUNCH_CHECK();
}

void launch_rk_final(float* m, const float* const k[7], float* mtmp, EngineState* st, const FinalArgs& a, int N,
                     hipStream_t s) {
    KPtrs kp;
    for (int i = 0; i < 7; i++) kp.k[i] = k[i];
    k_rk_final<<<a.adaptive ? dim3(kRedBlocks) : gridN(N), 256, 0, s>>>(m, kp, mtmp, st, a, N);
    GK_LAUNCH_CHECK();
}

void launch_copy_if_active(float* dst, const float* src, long n, const EngineState* st, hipStream_t s) {
    k_copy_if_active<<<gridN(n), 256, 0, s>>>(dst, src, n, st);
    GK_LAUNCH_CHECK();
}

void launch_ctrl_begin(EngineState* st, hipStream_t s) {
    k_ctrl_begin<<<1, 1, 0, s>>>(st);
    GK_LAUNCH_CHECK();
}
void launch_ctrl_after_k0(EngineState* st, int fsal_copy, const CtrlParams& p, hipStream_t s) {
    k_ctrl_after_k0<<<1, 64, 0, s>>>(st, fsal_copy, p);
    GK_LAUNCH_CHECK();
}
void launch_ctrl_end(EngineState* st, const CtrlParams& p, hipStream_t s) {
    k_ctrl_end<<<1, 64, 0, s>>>(st, p);
    GK_LAUNCH_CHECK();
}
void launch_commit(float* m, const float* mtmp, EngineState* st, const float* nmap, int N, hipStream_t s) {
    k_commit<<<gridN(N), 256, 0, s>>>(m, mtmp, st, nmap, N);
    GK_LAUNCH_CHECK();
}
void launch_normalize_active(float* m, const EngineState* st, const float* nmap, int N, hipStream_t s) {
    k_normalize_active<<<gridN(N), 256, 0, s>>>(m, st, nmap, N);
    GK_LAUNCH_CHECK();
}
void launch_sum3(const float* v, int N, double* out, hipStream_t s) {
    k_sum_comp<<<gridN(N), 256, 0, s>>>(v, N, 3, out);
    GK_LAUNCH_CHECK();
}
void launch_sum1(const float* v, int N, double* out, hipStream_t s) {
    k_sum_comp<<<gridN(N), 256, 0, s>>>(v, N, 1, out);
    GK_LAUNCH_CHECK();
}

}  // namespace gk
