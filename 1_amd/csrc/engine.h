// Device-resident LLG engine: shared declarations between the convolution
// kernels (conv.hip), the Runge-Kutta kernels (rk.hip) and the host-side
// simulation (sim.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/mumax_engine.h"
#include "gk_common.h"

namespace gk {

// ---------------------------------------------------------------- device state
// One per simulation, in device memory.  Control kernels (one thread) own
// the scalar fields; every stage kernel reads dt/time from here, so a whole
// run can be enqueued (or graph-replayed) without host round trips.
struct EngineState {
    double time;        // internal units
    double time0;       // start of the current step
    double stop;        // Run(): done when time >= stop
    long long steps;
    long long steps_target;  // Steps(n): done when steps >= steps_target
    long long trials_total, rejects_total;
    float dt;           // dt for the next trial (rk.dt)
    float dt_used;      // dt of the current trial
    float max_torque;   // maxabs(k0) of the current step (rk.go:353)
    float err;          // error estimate of the last trial (rk.go:403)
    int trials;         // trials of the current step (MAX_STEP_TRIALS = 10)
    int accept;         // last trial accepted
    int step_end;       // last trial ended the step (accepted or out of trials)
    int fsal_init;      // an FSAL k is available (rk.go:311)
    int done;           // no further trials: kernels become no-ops
    int trial_active;   // latched at the start of each trial
    int new_step;       // next trial starts a new step
    int log_n;          // trials logged
    int nan_dt;         // dt became NaN (rk.go:448-452)
    double energy;      // energy of the last evaluated trial (wantEnergy), internal units
    double prev_energy; // energy at the start of the current step (rk.go:297-298)
    int relax;          // Relax(): done once an accepted step raised the energy (relax.go:14-25)
    double relax_prev;  // energy after the previous step of the relaxation
    double field_tmax;  // latest stage time (internal) the applied field was evaluated at
    int nonfinite;      // max |torque| or the error estimate was NaN or infinite: the run stopped
    float dt_noise;     // dt at the start of the current step: the thermal field's (rk.go:289-292)
    long long noise_step;  // step the thermal field was drawn for (its counter)
};

// |x| for a per-block maximum that must not lose a NaN: fmaxf returns the other
// operand of a NaN, so a NaN maps to +inf, which every later fmaxf keeps.
__device__ __forceinline__ float abs_keep_nan(float x) { return x == x ? fabsf(x) : INFINITY; }

constexpr int kMaxStepTrials = 10;  // rk.go:277

// Applied field: 0 none, 1 static, >= 2 points interpolated linearly in time
// (simfield.go:101-153).  Times in seconds, fields in tesla, internal order.
struct FieldTable {
    int n;
    const double* t;  // device
    const float* b;   // device, 3n
    float b0[3];      // static value (n == 1)
    int fn;           // MM_FIELD_* function of time instead of the table (0: table)
    double fp[4];     // its parameters
    float fb[3];      // its amplitude (tesla, internal X, Y, Z)
};

// Applied-field functions of time, t in seconds (simfield.go:155-260): pulse
// (fp: duration, rise time; raised-cosine edges), RF (freq), RF ramp (freq,
// ramp time), sawtooth (freq), rotating (freq, phases of X, Y, Z).  The scale
// is computed in double and rounded to float32, as the reference does.
__host__ __device__ inline void field_function(int fn, const double fp[4], const float fb[3], double t,
                                               float out[3]) {
    const double pi = 3.14159265358979323846;
    float s[3] = {0.f, 0.f, 0.f};
    if (fn == MM_FIELD_PULSE) {
        const double d = fp[0], r = fp[1];
        double sc = 0.0;
        if (t > 0 && t < r) sc = 0.5 - 0.5 * cos(t * pi / r);
        else if (t >= r && t < d + r) sc = 1.0;
        else if (t >= d + r && t < d + 2.0 * r) sc = 0.5 + 0.5 * cos((t - d - r) * pi / r);
        s[0] = s[1] = s[2] = (float)sc;
    } else if (fn == MM_FIELD_RF) {
        s[0] = s[1] = s[2] = (float)sin(fp[0] * 2 * pi * t);
    } else if (fn == MM_FIELD_RF_RAMP) {
        const double fac = t < fp[1] ? t / fp[1] : 1.0;
        s[0] = s[1] = s[2] = (float)(sin(fp[0] * 2 * pi * t) * fac);
    } else if (fn == MM_FIELD_SAWTOOTH) {
        s[0] = s[1] = s[2] = (float)(asin(sin(fp[0] * 2 * pi * t)) / (pi / 2));
    } else if (fn == MM_FIELD_ROTATING) {
        for (int c = 0; c < 3; c++) s[c] = (float)sin(fp[0] * 2 * pi * t + fp[1 + c]);
    }
    for (int c = 0; c < 3; c++) out[c] = s[c] * fb[c];
}

// The reference's pointwise field keeps its search index and only moves it
// forward (simfield.go:126-133): the interval used at time t is the one ending
// at the first point at or past the latest time evaluated so far (t_search =
// max of all evaluated times, t included), so a trial that steps back in time
// after a rejection keeps the later interval (and a zero field past the end).
__host__ __device__ inline void applied_field_at(int n, const double* T, const float* B, const float* b0, double t_si,
                                                 double t_search_si, float out[3]) {
    out[0] = out[1] = out[2] = 0.f;
    if (n == 0) return;
    if (n == 1) {
        out[0] = b0[0];
        out[1] = b0[1];
        out[2] = b0[2];
        return;
    }
    int i = 0;
    while (i < n && T[i] < t_search_si) i++;  // first point at or past t_search
    if (i - 1 < 0 || i >= n) return;          // out of range: zero field
    const double dt = T[i] - T[i - 1];
    const float s = (float)((t_si - T[i - 1]) / dt);
    for (int c = 0; c < 3; c++) out[c] = B[3 * (i - 1) + c] + s * (B[3 * i + c] - B[3 * (i - 1) + c]);
}

__device__ __forceinline__ void applied_field(const FieldTable& f, double t_si, double t_search_si, float out[3]) {
    if (f.fn) {
        field_function(f.fn, f.fp, f.fb, t_si, out);
        return;
    }
    applied_field_at(f.n, f.t, f.b, f.b0, t_si, t_search_si, out);
}

// ---------------------------------------------------------------- stage build
// m_i = m + sum_j (dt a_ij) k_j, evaluated while loading the forward Z pass
// (rk.go:316-322).  Coefficients a_ij are float32 like the reference; dt is
// read from the device state.
struct StageIn {
    const float* m;
    const float* k[6];
    float a[6];
    int nk;
    float* mout;             // where m_i is stored for the torque pass (may be null)
    const EngineState* st;   // null: plain input (no stage build)
};

// ---------------------------------------------------------------- convolution plan
// What the convolution kernel is made of: cell size (internal units), the demag
// kernel (MM_KERNEL_FACE6 with `accuracy` charges per face edge, or
// MM_KERNEL_NEWELL) and whether the 6-neighbour exchange is folded in.
struct KernelSpec {
    double cell[3];
    int accuracy;
    int exch;
    int type;
    // periodic boundaries (Periodic, simgeom.go:84-88): per axis the period in
    // cells (the data size) and the number of images, 0 and 0 for none
    int period[3];
    int images[3];
};

// Launch parameters of a plan's passes (tile shapes, grid sizes).  The
// defaults are hand-picked; the autotuner (mm_sim_tune) times the candidates
// on the device and a table of tuned values (per data size) feeds new plans.
struct TuneParams {
    int zr = 0;   // Z forward rows per block: variant 0 = z_rows<M>(), 1 = half, 2 = double
    int er = 0;   // fused Z inverse rows per block: variant 0 = epi_rows<M>(), 1 = half
    int yb = 8;   // blocks per CU of the pipelined 3-D Y passes
    int xb = 4;   // blocks per CU of the pipelined X.K.X pass
    int y2b = 8;  // blocks per CU of the 2-D Y.K.Y pass
    int pp = 0;   // 3-D plans that fit it: 1 = Y.X K X.Y in one plane pass (k_plane3), 0 = separate passes
    int xv = 0;   // X.K.X variant (padded X 16-64): 0 default, 1 no prefetch, 2 half the elements per lane, 3 both
};

struct ConvPlan {
    int N[3];   // data size (internal X,Y,Z)
    int P[3];   // padded FFT size
    int M, Q;   // Pz/2, Pz/2+1
    int nK;     // 6 (3-D) or 4 (2-D)
    float* Kc;  // compressed kernel spectrum (see conv.hip), shared between plans of one spec
    long kc_elems;
    float2* B1; // [3][Nx][Q][Ny]
    float2* B2; // 3-D: [3][Q][Py][Nx] (not used by the thin plane pass)
    int yxy;    // thin 3-D: Y . X K X . Y inverse in one plane pass (no B2)
    int plane3; // 3-D plan the register-resident plane pass fits (used when tp.pp is set)
    size_t bytes;
    TuneParams tp;  // from the tuning table (conv_tune_lookup) when it holds N, else defaults
};

// Tuning table: data size N -> launch parameters (process-wide; conv_plan_init
// reads it).  Text file: one line "Nx Ny Nz zr er yb xb y2b pp" per size (a
// line without pp reads as pp = 0).
bool conv_tune_lookup(const int N[3], TuneParams* tp);
void conv_tune_store(const int N[3], const TuneParams& tp);
int conv_tune_save(const char* path);  // entries written, or -1
int conv_tune_load(const char* path);  // entries read, or -1
void conv_tune_clear();
// rows per block of the fused Z inverse for Z half-length M and variant v, and
// whether variant v of the Z forward / fused Z inverse rows is a distinct shape
int conv_epi_rows(int M, int v);
bool conv_zr_valid(int M, int v);
bool conv_er_valid(int M, int v);
bool conv_xv_valid(int PX, int v);

// Choose padded sizes, allocate buffers.  Kc is set by conv_acquire_kernel.
void conv_plan_init(ConvPlan& cp, const int N[3]);
// Frees the buffers and releases the plan's kernel.
void conv_plan_free(ConvPlan& cp);
// Points cp.Kc at the kernel spectrum of spec `ks` for cp's padded size on the
// current device, building it (conv_build_kernel) if no live plan holds it:
// plans of the same padded size and spec share one spectrum, freed with the
// last of them.  Returns the spectrum's max|Im|/max|Re|.
double conv_acquire_kernel(ConvPlan& cp, const KernelSpec& ks);
// number of distinct kernel spectra held on the current device
int conv_kernel_cache_size();
// bytes conv_plan_init allocates for data size N, and the blocks of the fused Z inverse pass
size_t conv_plan_footprint(const int N[3], int* epi_blocks);
// FaceKernel6 (accuracy x accuracy charges per face) or the Newell tensor in
// FP64 on the device, optionally plus the 6-neighbour exchange kernel, FFT in
// FP64, stored compressed in cp.Kc (cp.kc_elems floats, allocated by the caller).  If full != null, also writes nK x [Px][Py][Q] floats
// in the reference's layout (conv.go:110-167).  Returns max|Im|/max|Re| of the
// spectrum (the reference warns above 1e-4, conv.go:153).
double conv_build_kernel(ConvPlan& cp, const KernelSpec& ks, float* full_host, float* full_dev);
// real-space demag (+exchange) kernel over the padded grid P, 6 x P floats / doubles to the host
void conv_realspace_kernel(const int P[3], const KernelSpec& ks, float* out_host);
void conv_realspace_kernel_f64(const int P[3], const KernelSpec& ks, double* out_host);
// Optional per-kernel-class event timing (bench/profiling).  Classes:
// 0 Z fwd, 1 Y fwd, 2 X.K.X (3-D) or Y.K.Y (2-D), 3 Y inv, 4 Z inv (+ torque epilogue),
// 5 separate torque, 6 separate RK final (adaptive tableaux), 7 control/copy.
struct Prof {
    static constexpr int kClasses = 8;
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, int>> marks;  // (class, index of start event; end = start+1)
    int used = 0;
    void begin(int cls, hipStream_t s);
    void end(hipStream_t s);
    void collect(double ms[kClasses], int counts[kClasses]);
    void reset() { marks.clear(); used = 0; }
    ~Prof();
};
struct EpiArgs;
// h = K * m (planar 3N each); in = stage build description, out = 3N floats.
// With epi != null the Z inverse pass ends in the torque / final-combination
// epilogue described by *epi instead of writing h.
void conv_apply(const ConvPlan& cp, const StageIn& in, float* h, hipStream_t s, Prof* prof = nullptr,
                const EpiArgs* epi = nullptr);
// number of blocks (= per-block partial maxima) of the fused Z inverse + torque pass
int conv_epi_blocks(const ConvPlan& cp);  // blocks of the fused Z inverse at cp.tp.er
int conv_epi_blocks_max(const int N[3]);   // ... the most over every row variant
// one pass of class cls (0 Z fwd, 1 Y fwd, 2 X.K.X / Y.K.Y, 3 Y inv, 4 Z inv), for timing
void conv_run_pass(const ConvPlan& cp, int cls, const StageIn& in, float* h, const EpiArgs* epi, hipStream_t s);
// st->trial_active = 1 (stage kernels do their work until the next trial start)
void launch_set_active(EngineState* st, hipStream_t s);
// st->relax = on (relax_prev = +inf): Relax() mode of the step control
void launch_set_relax(EngineState* st, int on, hipStream_t s);

// ---------------------------------------------------------------- RK kernels
struct TorqueArgs {
    float alpha, dtg;       // damping, 1/(1+alpha^2)
    float thermalK;         // thermal field sigma * sqrt(dt) (internal units); 0: T = 0
    unsigned long long noiseSeed;
    const float* alphaMap;  // per-cell damping multiplier (null: none); dtg keeps the global alpha (torque.go:49-51)
    const float* normMap;   // per-cell |m| after each normalisation (null: 1; sim.go:230-239)
    int anisType;
    float kAnis;            // 2K1/(mu0 Ms^2) (init.go:103-108)
    float u[3];
    FieldTable field;
    float unitField;        // tesla per internal field unit
    double unitTime;        // seconds per internal time unit
    float c;                // stage node c_i
    float cmax;             // max c_j over the stages evaluated so far in this trial
    float* part;            // per-block max |k| (null: no reduction at this stage)
    int exch;               // add the 6-neighbour exchange stencil here
    int n[3];
    float cexch[4];         // cx, cy, cz, c0
};
// k = torque(m_i, H + local fields (+ exchange)); in place over hk.
void launch_torque(const float* mi, float* hk, EngineState* st, const TorqueArgs& a, int N, hipStream_t s);
// h += the thermal field of the current step (no-op at T = 0)
void launch_add_thermal(float* h, const EngineState* st, const TorqueArgs& a, int N, hipStream_t s);

// Applied field in internal units at the stage node: t = time0 + float64(c_i dt)
// (rk.go:316), B(t) in tesla / B_unit (field.go:50-61).
__device__ __forceinline__ void stage_field(const TorqueArgs& a, const EngineState* st, float hext[3]) {
    const double t = st->time0 + (double)(a.c * st->dt_used);
    const double ts = fmax(st->field_tmax, st->time0 + (double)(a.cmax * st->dt_used));
    float b[3];
    applied_field(a.field, t * a.unitTime, ts * a.unitTime, b);
    for (int i = 0; i < 3; i++) hext[i] = b[i] / a.unitField;
}

// Thermal field (temperature.go:29-36): per step and cell, three standard
// normals from a counter-based stream -- SplitMix64's output function
// (Steele, Lea & Flood 2014) over (seed, step, cell), 24-bit uniforms, the
// Box-Muller transform -- times sigma = K / sqrt(dt_step).  The same draw in
// every stage and trial of a step, in every launch shape and in batches.
__host__ __device__ inline unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline unsigned long long thermal_key(unsigned long long seed, long long step, long cell, int j) {
    const unsigned long long base = mix64(seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(step + 1));
    return mix64(base + 0x9E3779B97F4A7C15ull * (unsigned long long)(4 * cell + j + 1));
}
// uniform in (0, 1) (half-offset) and in [0, 1) from the top 24 bits
__host__ __device__ inline float thermal_u_open(unsigned long long r) { return ((float)(r >> 40) + 0.5f) * 0x1p-24f; }
__host__ __device__ inline float thermal_u(unsigned long long r) { return (float)(r >> 40) * 0x1p-24f; }
__device__ __forceinline__ void add_thermal(const TorqueArgs& a, const EngineState* st, long cell, float h[3]) {
    if (a.thermalK == 0.f) return;
    const float sigma = a.thermalK * rsqrtf(st->dt_noise);
    const float u1 = thermal_u_open(thermal_key(a.noiseSeed, st->noise_step, cell, 0));
    const float u2 = thermal_u(thermal_key(a.noiseSeed, st->noise_step, cell, 1));
    const float u3 = thermal_u_open(thermal_key(a.noiseSeed, st->noise_step, cell, 2));
    const float u4 = thermal_u(thermal_key(a.noiseSeed, st->noise_step, cell, 3));
    const float r1 = sqrtf(-2.f * logf(u1)), r2 = sqrtf(-2.f * logf(u3));
    float s1, c1, s2, c2;
    sincospif(2.f * u2, &s1, &c1);
    sincospif(2.f * u4, &s2, &c2);
    h[0] += sigma * (r1 * c1);
    h[1] += sigma * (r1 * s1);
    h[2] += sigma * (r2 * c2);
}

// LL torque of one cell: h = H + B_ext + uniaxial anisotropy of m
// (gpu_local_contr3.cu), k = dtg (m x h - alpha m x (m x h)) (gpu_torque.cu:20-50).
// damping of cell i: alpha * alphaMap[i] (gpu_torque.cu's alpha_map)
__device__ __forceinline__ float cell_alpha(const TorqueArgs& a, long i) {
    return a.alphaMap ? a.alpha * a.alphaMap[i] : a.alpha;
}
// scale that normalises (x, y, z): 1/|m|, times the cell's norm when a norm map
// is given, 0 where the map is 0 (gpu_normalize.cu:20-30 uniform, :43-71 map)
__device__ __forceinline__ float norm_scale(float x, float y, float z, const float* nmap, long i) {
    const float s = rsqrtf(x * x + y * y + z * z);
    if (!nmap) return s;
    const float w = nmap[i];
    return w == 0.f ? 0.f : s * w;
}
__device__ __forceinline__ void llg_torque_cell(const TorqueArgs& a, float alpha, const float hext[3], float Mx,
                                                float My, float Mz, float Hx, float Hy, float Hz, float k[3]) {
    Hx += hext[0];
    Hy += hext[1];
    Hz += hext[2];
    if (a.anisType == 1) {
        const float mu = Mx * a.u[0] + My * a.u[1] + Mz * a.u[2];
        Hx += a.kAnis * mu * a.u[0];
        Hy += a.kAnis * mu * a.u[1];
        Hz += a.kAnis * mu * a.u[2];
    }
    const float cx = My * Hz - Mz * Hy, cy = Mz * Hx - Mx * Hz, cz = Mx * Hy - My * Hx;
    const float dx = My * cz - Mz * cy, dy = Mz * cx - Mx * cz, dz = Mx * cy - My * cx;
    k[0] = a.dtg * (cx - alpha * dx);
    k[1] = a.dtg * (cy - alpha * dy);
    k[2] = a.dtg * (cz - alpha * dz);
}

// Epilogue of the fused Z inverse pass (one RK stage):
//   m_i = m + dt sum_j a_j k_j               (the stage input, rebuilt from m and the k_j)
//   k_i = torque(m_i, h)                     -> out (with per-block max |k_i| if t.part)
// or, when fin (last stage of a fixed-step tableau, rk.go:381-390 + the final normalise),
//   m <- normalise(m + dt (sum_j b_j k_j + bs k_i))   (out == m, k_i is not stored)
struct EpiArgs {
    const float* m;
    float* out;
    const float* k[6];
    float a[6], b[6];
    float bs;
    int nu;
    int fin;          // 0 torque; 1 fixed-step final; 2 adaptive FSAL final (see zinv_epi_tile)
    TorqueArgs t;
    double* epart;    // non-null: per-block sums of the energy density of m_i (FP64)
    // fin == 2: b2_j - b_j of the k slots and of this stage (times dt on the
    // device), where m_high goes, and the per-block maximum error
    double ce[6];
    double ce_self;
    float* mhigh;
    float* part_err;
};

struct FinalArgs {
    int S;
    float b[7], b2[7];
    int adaptive;
    float* part_err;        // per-block max |m_low - m_high| (adaptive)
    const float* normMap;   // fixed step: the norm map of the final normalisation (null: 1)
};

// Pointwise kernels run a fixed grid of kRedBlocks blocks (grid-stride), so
// every per-block partial buffer has exactly kRedBlocks entries.
constexpr int kRedBlocks = 1024;
// high/low order solutions; adaptive: mtmp = m_high, part_err = per-block max|m_low-m_high|;
// fixed step: m = normalize(m_high) in place.
void launch_rk_final(float* m, const float* const k[7], float* mtmp, EngineState* st, const FinalArgs& a, int N,
                     hipStream_t s);
void launch_copy_if_active(float* dst, const float* src, long n, const EngineState* st, hipStream_t s);

struct CtrlParams {
    int adaptive;
    float maxError, errorOrder;
    float inputDt;          // internal
    float minDm, maxDm, minDt, maxDt;  // minDt/maxDt in internal units
    double* log_dt;
    int* log_accept;
    float* log_err;         // per trial: the error estimate (adaptive) or 0
    int log_cap;
    const float* part_k0;   // per-block max |k0| (stage 0 torque)
    const float* part_k6;   // per-block max |k_last| (FSAL tableau)
    const float* part_err;  // per-block error (adaptive)
    int n_k, n_err;         // entries of part_k0/part_k6 and of part_err
    int fold_k0;            // ctrl_end also records max |k0| (no separate after-k0 control)
    const double* part_e;   // non-null (wantEnergy): n_k per-block energy sums of this trial
    double cell_volume;     // internal units
    float cmax;             // largest stage node of the tableau (advances field_tmax)
    int thermal;            // T > 0: no dt control after the first torque (rk.go:348)
};
void launch_ctrl_begin(EngineState* st, hipStream_t s);
void launch_ctrl_after_k0(EngineState* st, int fsal_copy, const CtrlParams& p, hipStream_t s);
void launch_ctrl_end(EngineState* st, const CtrlParams& p, hipStream_t s);
void launch_commit(float* m, const float* mtmp, EngineState* st, const float* nmap, int N, hipStream_t s);
void launch_normalize_active(float* m, const EngineState* st, const float* nmap, int N, hipStream_t s);

void launch_energy(const float* m, const float* h, const float hext[3], int anisType, float kAnis, const float u[3],
                   int N, double* out, hipStream_t s);
// steps_target = steps + nsteps (nsteps < 0: unlimited), stop time; recomputes `done`
void launch_set_target(EngineState* st, long long nsteps, double stop, hipStream_t s);
// FP64 reductions for outputs: sums[0..2] = sum of each component; energy
void launch_sum3(const float* v, int N, double* out_dev, hipStream_t s);
void launch_sum1(const float* v, int N, double* out_dev, hipStream_t s);

}  // namespace gk
