// Device bodies of the Runge-Kutta pointwise and control kernels, shared by
// the single-simulation kernels (rk.hip) and the batched kernels (batch.hip).
// Block-strided bodies take their block index and block count explicitly.
#pragma once
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

struct KPtrs {
    const float* k[7];
};

// m_high = m + sum b_i dt k_i (rk.go:381-390); the error m_low - m_high =
// dt sum (b2_i - b_i) k_i (rk.go:391-403) is accumulated in FP64 from the k_i
// directly, so it carries no cancellation of two float32 solutions.
__device__ __forceinline__ void rk_final_body(float* __restrict__ m, const KPtrs& kp, float* __restrict__ mtmp,
                                              const EngineState* st, const FinalArgs& a, int N, int blk, int nblk) {
    const float dt = st->dt_used;
    float cb[7];
    double ce[7];
#pragma unroll
    for (int s = 0; s < 7; s++) {
        cb[s] = a.b[s] * dt;
        ce[s] = ((double)a.b2[s] - (double)a.b[s]) * (double)dt;
    }
    float lmax = 0.f;
    for (int i = blk * blockDim.x + threadIdx.x; i < N; i += nblk * blockDim.x) {
        float hi[3];
        double er[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float m0 = m[c * N + i];
            float h = m0;
            double e = 0.0;
            for (int s = 0; s < a.S; s++) {
                const bool used = a.b[s] != 0.f || (a.adaptive && a.b2[s] != 0.f);
                const float kv = used ? kp.k[s][c * N + i] : 0.f;
                if (a.b[s] != 0.f) h = fmaf(cb[s], kv, h);
                if (a.adaptive) e = fma(ce[s], (double)kv, e);
            }
            hi[c] = h;
            er[c] = e;
        }
        if (a.adaptive) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                mtmp[c * N + i] = hi[c];
                lmax = fmaxf(lmax, abs_keep_nan((float)er[c]));
            }
        } else {
            const float s = norm_scale(hi[0], hi[1], hi[2], a.normMap, i);
#pragma unroll
            for (int c = 0; c < 3; c++) m[c * N + i] = hi[c] * s;
        }
    }
    if (a.adaptive) {
        const float bm = block_max_nonneg(lmax);
        if (threadIdx.x == 0) a.part_err[blk] = bm;
    }
}

__device__ __forceinline__ void copy_body(float* __restrict__ dst, const float* __restrict__ src, long n, int blk,
                                          int nblk) {
    for (long i = blk * (long)blockDim.x + threadIdx.x; i < n; i += (long)nblk * blockDim.x) dst[i] = src[i];
}

// accepted: m = normalize(m_high); step ended after the last rejected trial:
// m (restored, i.e. untouched) is normalised (rk.go:426-444); with a norm map
// each cell to its own norm (sim.go:230-239)
__device__ __forceinline__ void commit_body(float* __restrict__ m, const float* __restrict__ mtmp,
                                            const EngineState* st, const float* __restrict__ nmap, int N, int blk,
                                            int nblk) {
    if (!st->trial_active || !st->step_end) return;
    const float* src = st->accept ? mtmp : m;
    for (int i = blk * blockDim.x + threadIdx.x; i < N; i += nblk * blockDim.x) {
        const float x = src[i], y = src[N + i], z = src[2 * N + i];
        const float s = norm_scale(x, y, z, nmap, i);
        m[i] = x * s;
        m[N + i] = y * s;
        m[2 * N + i] = z * s;
    }
}

// ---------------------------------------------------------------- control (single thread)
__device__ __forceinline__ void ctrl_begin_body(EngineState* st) {
    st->trial_active = !st->done;
    if (!st->trial_active) return;
    if (st->new_step) {
        st->time0 = st->time;
        st->trials = 0;
        st->new_step = 0;
        st->prev_energy = st->energy == 0 ? 9999999999.0 : st->energy;  // rk.go:297-298
        // the thermal field of this step: drawn from dt at its start (rk.go:289-292)
        st->dt_noise = st->dt;
        st->noise_step = st->steps;
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

// after k0 is known (rk.go:348-373); run by one wave
__device__ __forceinline__ void ctrl_after_k0_body(EngineState* st, int fsal_copy, const CtrlParams& p) {
    if (!st->trial_active) return;
    const float maxTorque = wave_max_partials(fsal_copy ? p.part_k6 : p.part_k0, p.n_k);
    if (threadIdx.x != 0) return;
    if (!isfinite(maxTorque)) st->nonfinite = 1;
    if (p.thermal) return;  // rk.go:348: the dt control after k0 only runs at T = 0
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

// after the high/low solutions (rk.go:376-440); run by one wave
__device__ __forceinline__ void ctrl_end_body(EngineState* st, const CtrlParams& p) {
    if (!st->trial_active) return;
    const float err_red = p.adaptive ? wave_max_partials(p.part_err, p.n_err) : 0.f;
    const float k0_red = p.fold_k0 ? wave_max_partials(p.part_k0, p.n_k) : 0.f;
    double e_red = 0.0;
    if (p.part_e) {
        for (int i = threadIdx.x; i < p.n_k; i += 64) e_red += p.part_e[i];
        for (int off = 32; off > 0; off >>= 1) e_red += __shfl_xor(e_red, off, 64);
    }
    if (threadIdx.x != 0) return;
    if (p.part_e) st->energy = e_red * p.cell_volume;
    if (p.fold_k0) {
        st->max_torque = k0_red;
        if (!isfinite(k0_red)) st->nonfinite = 1;
        if (!p.adaptive && !p.thermal) st->dt = p.inputDt;
    }
    const float dtu = st->dt_used;
    st->time = st->time0 + (double)dtu;
    st->field_tmax = fmax(st->field_tmax, st->time0 + (double)(p.cmax * dtu));
    st->trials++;
    st->trials_total++;
    st->fsal_init = 1;
    int bad = 0;
    if (p.adaptive) {
        const float err = err_red;
        st->err = err;
        if (!isfinite(err)) st->nonfinite = 1;
        float factor = (float)pow((double)(p.maxError / err), 1.0 / (double)p.errorOrder);
        if (factor > 2.f) factor = 2.f;
        if (factor < 0.01f) factor = 0.01f;
        float dt = dtu * factor;
        if (isnan(dt)) st->nan_dt = 1;
        // wantEnergy: the energy must not rise within a step (rk.go:424-426)
        bad = err > 2.f * p.maxError || (p.part_e && st->energy > st->prev_energy);
        if (!bad) {
            if (dt < p.minDt) dt = p.minDt;
            if (p.maxDt != 0.f && dt > p.maxDt) dt = p.maxDt;
        }
        st->dt = dt;
    }
    if (bad) st->rejects_total++;
    st->accept = !bad;
    st->step_end = !bad || st->trials >= kMaxStepTrials || st->nonfinite;
    if (st->log_n < p.log_cap) {
        p.log_dt[st->log_n] = (double)dtu;
        p.log_accept[st->log_n] = !bad;
        p.log_err[st->log_n] = p.adaptive ? st->err : 0.f;
    }
    st->log_n++;
    if (st->step_end) {
        st->steps++;
        st->new_step = 1;
        if (st->relax) {  // relax.go:14-25: stop once a step raised the energy
            if (st->energy > st->relax_prev) st->done = 1;
            st->relax_prev = st->energy;
        }
        if (st->time >= st->stop || st->steps >= st->steps_target || st->nan_dt || st->nonfinite) st->done = 1;
    }
}

}  // namespace gk
