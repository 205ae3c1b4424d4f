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

namespace gk {

__device__ __forceinline__ float block_max_nonneg(float v) {
    __shared__ float wmax[16];
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wmax[wave] = v;
    __syncthreads();
    float r = 0.f;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) r = fmaxf(r, wmax[w]);
    return r;  // valid in thread 0
}

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
        float k[3];
        llg_torque_cell(a, hext, m[i], m[N + i], m[2 * N + i], hk[i], hk[N + i], hk[2 * N + i], k);
        hk[i] = k[0];
        hk[N + i] = k[1];
        hk[2 * N + i] = k[2];
        lmax = fmaxf(lmax, fmaxf(fabsf(k[0]), fmaxf(fabsf(k[1]), fabsf(k[2]))));
    }
    if (a.part) {
        const float bm = block_max_nonneg(lmax);
        if (threadIdx.x == 0) a.part[blockIdx.x] = bm;
    }
}

struct KPtrs {
    const float* k[7];
};

// m_high = m + sum b_i dt k_i, m_low = m + sum b2_i dt k_i (rk.go:381-403)
__global__ __launch_bounds__(256) void k_rk_final(float* __restrict__ m, KPtrs kp, float* __restrict__ mtmp,
                                                  EngineState* st, FinalArgs a, int N) {
    if (!st->trial_active) return;
    const float dt = st->dt_used;
    float cb[7], cb2[7];
#pragma unroll
    for (int s = 0; s < 7; s++) {
        cb[s] = a.b[s] * dt;
        cb2[s] = a.b2[s] * dt;
    }
    float lmax = 0.f;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        float hi[3], lo[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float m0 = m[c * N + i];
            float h = m0, l = m0;
            for (int s = 0; s < a.S; s++) {
                const float kv = (a.b[s] != 0.f || a.b2[s] != 0.f) ? kp.k[s][c * N + i] : 0.f;
                if (a.b[s] != 0.f) h = fmaf(cb[s], kv, h);
                if (a.adaptive && a.b2[s] != 0.f) l = fmaf(cb2[s], kv, l);
            }
            hi[c] = h;
            lo[c] = l;
        }
        if (a.adaptive) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                mtmp[c * N + i] = hi[c];
                lmax = fmaxf(lmax, fabsf(lo[c] - hi[c]));
            }
        } else {
            const float s = rsqrtf(hi[0] * hi[0] + hi[1] * hi[1] + hi[2] * hi[2]);
#pragma unroll
            for (int c = 0; c < 3; c++) m[c * N + i] = hi[c] * s;
        }
    }
    if (a.adaptive) {
        const float bm = block_max_nonneg(lmax);
        if (threadIdx.x == 0) a.part_err[blockIdx.x] = bm;
    }
}

__global__ void k_copy_if_active(float* __restrict__ dst, const float* __restrict__ src, long n,
                                 const EngineState* st) {
    if (!st->trial_active) return;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// accepted: m = normalize(m_high); step ended after the last rejected trial:
// m (restored, i.e. untouched) is normalised (rk.go:426-444)
__global__ __launch_bounds__(256) void k_commit(float* __restrict__ m, const float* __restrict__ mtmp,
                                                const EngineState* st, int N) {
    if (!st->trial_active || !st->step_end) return;
    const float* src = st->accept ? mtmp : m;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const float x = src[i], y = src[N + i], z = src[2 * N + i];
        const float s = rsqrtf(x * x + y * y + z * z);
        m[i] = x * s;
        m[N + i] = y * s;
        m[2 * N + i] = z * s;
    }
}

__global__ void k_normalize_active(float* __restrict__ m, const EngineState* st, int N) {
    if (st && !st->trial_active) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const float x = m[i], y = m[N + i], z = m[2 * N + i];
        const float s = rsqrtf(x * x + y * y + z * z);
        m[i] = x * s;
        m[N + i] = y * s;
        m[2 * N + i] = z * s;
    }
}

// ---------------------------------------------------------------- control (single thread)
__global__ void k_ctrl_begin(EngineState* st) {
    st->trial_active = !st->done;
    if (!st->trial_active) return;
    if (st->new_step) {
        st->time0 = st->time;
        st->trials = 0;
        st->new_step = 0;
    }
    st->dt_used = st->dt;
}

// one wave reduces n per-block maxima; result valid in every lane
__device__ __forceinline__ float wave_max_partials(const float* part, int n) {
    float v = 0.f;
    for (int i = threadIdx.x; i < n; i += 64) v = fmaxf(v, part[i]);
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// after k0 is known (rk.go:348-373); launched with one wave
__global__ void k_ctrl_after_k0(EngineState* st, int fsal_copy, CtrlParams p) {
    if (!st->trial_active) return;
    const float maxTorque = wave_max_partials(fsal_copy ? p.part_k6 : p.part_k0, p.n_k);
    if (threadIdx.x != 0) return;
    if (!p.adaptive) st->dt = p.inputDt;
    st->max_torque = maxTorque;
    float dt = st->dt;
    const float dm = dt * maxTorque;
    if (dm < p.minDm) dt = p.minDm / maxTorque;
    if (dt < p.minDt) dt = p.minDt;
    if (p.maxDm != 0.f && dm > p.maxDm) dt = p.maxDm / maxTorque;
    if (p.maxDt != 0.f && dt > p.maxDt) dt = p.maxDt;
    if (isnan(dt)) st->nan_dt = 1;
    st->dt = dt;
    st->dt_used = dt;
}

// after the high/low solutions (rk.go:376-440)
__global__ void k_ctrl_end(EngineState* st, CtrlParams p) {
    if (!st->trial_active) return;
    const float err_red = p.adaptive ? wave_max_partials(p.part_err, p.n_err) : 0.f;
    const float k0_red = p.fold_k0 ? wave_max_partials(p.part_k0, p.n_k) : 0.f;
    if (threadIdx.x != 0) return;
    if (p.fold_k0) {
        st->max_torque = k0_red;
        if (!p.adaptive) st->dt = p.inputDt;
    }
    const float dtu = st->dt_used;
    st->time = st->time0 + (double)dtu;
    st->trials++;
    st->trials_total++;
    st->fsal_init = 1;
    int bad = 0;
    if (p.adaptive) {
        const float err = err_red;
        st->err = err;
        float factor = (float)pow((double)(p.maxError / err), 1.0 / (double)p.errorOrder);
        if (factor > 2.f) factor = 2.f;
        if (factor < 0.01f) factor = 0.01f;
        float dt = dtu * factor;
        if (isnan(dt)) st->nan_dt = 1;
        bad = err > 2.f * p.maxError;
        if (!bad) {
            if (dt < p.minDt) dt = p.minDt;
            if (p.maxDt != 0.f && dt > p.maxDt) dt = p.maxDt;
        }
        st->dt = dt;
    }
    if (bad) st->rejects_total++;
    st->accept = !bad;
    st->step_end = !bad || st->trials >= kMaxStepTrials;
    if (st->log_n < p.log_cap) {
        p.log_dt[st->log_n] = (double)dtu;
        p.log_accept[st->log_n] = !bad;
    }
    st->log_n++;
    if (st->step_end) {
        st->steps++;
        st->new_step = 1;
        if (st->time >= st->stop || st->steps >= st->steps_target || st->nan_dt) st->done = 1;
    }
}

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
void launch_set_active(EngineState* st, hipStream_t s) {
    k_set_active<<<1, 1, 0, s>>>(st);
    GK_LAUNCH_CHECK();
}
void launch_set_target(EngineState* st, long long nsteps, double stop, hipStream_t s) {
    k_set_target<<<1, 1, 0, s>>>(st, nsteps, stop);
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
void launch_commit(float* m, const float* mtmp, EngineState* st, int N, hipStream_t s) {
    k_commit<<<gridN(N), 256, 0, s>>>(m, mtmp, st, N);
    GK_LAUNCH_CHECK();
}
void launch_normalize_active(float* m, const EngineState* st, int N, hipStream_t s) {
    k_normalize_active<<<gridN(N), 256, 0, s>>>(m, st, N);
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
