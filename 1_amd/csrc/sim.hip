// Host side of the LLG engine: the reference's Sim/RK driver (sim.go,
// init.go:80-287, rk.go:41-445, simrun.go:19-92, field.go, torque.go)
// restated in C++ on top of libgpukern.
//
// MM_MODE_FUSED enqueues each Runge-Kutta trial as a fixed sequence of fused
// kernels whose scalars (dt, time, accept) live on the device, so Run()
// only synchronises once per batch of trials.  MM_MODE_ABI replays the
// reference's own call sequence through the drop-in gpukern.h functions, one
// synchronous call at a time, as the Go driver would.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpukern.h"
#include "../../include/mumax_engine.h"
#include "engine.h"
#include "sim.h"

using namespace gk;

namespace {

// rk.go:62-212, coefficients rounded to float32 as in the reference
Tableau make_tableau(int solver) {
    Tableau t;
    // the semi-analytic solvers keep RK2's buffers: k_0 holds h, k_1 the corrector's m3
    if (solver == MM_SEMIANAL1 || solver == MM_SEMIANAL2) solver = MM_RK2;
    switch (solver) {
    case MM_RK2:
    case MM_RK12:
        t.S = 2;
        t.c[1] = 1.f;
        t.a[1][0] = 1.f;
        t.b[0] = .5f;
        t.b[1] = .5f;
        if (solver == MM_RK12) {
            t.adaptive = 1;
            t.order = 1.f;
            t.b2[0] = 1.f;
        }
        break;
    case MM_RK3:
        t.S = 3;
        t.c[1] = 1.f / 2.f;
        t.c[2] = 1.f;
        t.a[1][0] = 1.f / 2.f;
        t.a[2][0] = -1.f;
        t.a[2][1] = 2.f;
        t.b[0] = 1.f / 6.f;
        t.b[1] = 2.f / 3.f;
        t.b[2] = 1.f / 6.f;
        break;
    case MM_RK23:
        t.S = 4;
        t.fsal = 1;
        t.c[1] = 1.f / 2.f;
        t.c[2] = 3.f / 4.f;
        t.c[3] = 1.f;
        t.a[1][0] = 1.f / 2.f;
        t.a[2][1] = 3.f / 4.f;
        t.a[3][0] = 2.f / 9.f;
        t.a[3][1] = 1.f / 3.f;
        t.a[3][2] = 4.f / 9.f;
        t.b[0] = 2.f / 9.f;
        t.b[1] = 1.f / 3.f;
        t.b[2] = 4.f / 9.f;
        t.adaptive = 1;
        t.order = 2.f;
        t.b2[0] = 7.f / 24.f;
        t.b2[1] = 1.f / 4.f;
        t.b2[2] = 1.f / 3.f;
        t.b2[3] = 1.f / 8.f;
        break;
    case MM_RK4:
        t.S = 4;
        t.c[1] = 1.f / 2.f;
        t.c[2] = 1.f / 2.f;
        t.c[3] = 1.f;
        t.a[1][0] = 1.f / 2.f;
        t.a[2][1] = 1.f / 2.f;
        t.a[3][2] = 1.f;
        t.b[0] = 1.f / 6.f;
        t.b[1] = 1.f / 3.f;
        t.b[2] = 1.f / 3.f;
        t.b[3] = 1.f / 6.f;
        break;
    case MM_RKCK:
        t.S = 6;
        t.c[1] = 1.f / 5.f;
        t.c[2] = 3.f / 10.f;
        t.c[3] = 3.f / 5.f;
        t.c[4] = 1.f;
        t.c[5] = 7.f / 8.f;
        t.a[1][0] = 1.f / 5.f;
        t.a[2][0] = 3.f / 40.f;
        t.a[2][1] = 9.f / 40.f;
        t.a[3][0] = 3.f / 10.f;
        t.a[3][1] = -9.f / 10.f;
        t.a[3][2] = 6.f / 5.f;
        t.a[4][0] = -11.f / 54.f;
        t.a[4][1] = 5.f / 2.f;
        t.a[4][2] = -70.f / 27.f;
        t.a[4][3] = 35.f / 27.f;
        t.a[5][0] = 1631.f / 55296.f;
        t.a[5][1] = 175.f / 512.f;
        t.a[5][2] = 575.f / 13824.f;
        t.a[5][3] = 44275.f / 110592.f;
        t.a[5][4] = 253.f / 4096.f;
        t.b[0] = 37.f / 378.f;
        t.b[2] = 250.f / 621.f;
        t.b[3] = 125.f / 594.f;
        t.b[5] = 512.f / 1771.f;
        t.adaptive = 1;
        t.order = 4.f;
        t.b2[0] = 2825.f / 27648.f;
        t.b2[2] = 18575.f / 48384.f;
        t.b2[3] = 13525.f / 55296.f;
        t.b2[4] = 277.f / 14336.f;
        t.b2[5] = 1.f / 4.f;
        break;
    case MM_RKDP:
        t.S = 7;
        t.fsal = 1;
        t.c[1] = 1.f / 5.f;
        t.c[2] = 3.f / 10.f;
        t.c[3] = 4.f / 5.f;
        t.c[4] = 8.f / 9.f;
        t.c[5] = 1.f;
        t.c[6] = 1.f;
        t.a[1][0] = 1.f / 5.f;
        t.a[2][0] = 3.f / 40.f;
        t.a[2][1] = 9.f / 40.f;
        t.a[3][0] = 44.f / 45.f;
        t.a[3][1] = -56.f / 15.f;
        t.a[3][2] = 32.f / 9.f;
        t.a[4][0] = 19372.f / 6561.f;
        t.a[4][1] = -25360.f / 2187.f;
        t.a[4][2] = 64448.f / 6561.f;
        t.a[4][3] = -212.f / 729.f;
        t.a[5][0] = 9017.f / 3168.f;
        t.a[5][1] = -355.f / 33.f;
        t.a[5][2] = 46732.f / 5247.f;
        t.a[5][3] = 49.f / 176.f;
        t.a[5][4] = -5103.f / 18656.f;
        t.a[6][0] = 35.f / 384.f;
        t.a[6][2] = 500.f / 1113.f;
        t.a[6][3] = 125.f / 192.f;
        t.a[6][4] = -2187.f / 6784.f;
        t.a[6][5] = 11.f / 84.f;
        t.b[0] = 35.f / 384.f;
        t.b[2] = 500.f / 1113.f;
        t.b[3] = 125.f / 192.f;
        t.b[4] = -2187.f / 6784.f;
        t.b[5] = 11.f / 84.f;
        t.adaptive = 1;
        t.order = 4.f;
        t.b2[0] = 5179.f / 57600.f;
        t.b2[2] = 7571.f / 16695.f;
        t.b2[3] = 393.f / 640.f;
        t.b2[4] = -92097.f / 339200.f;
        t.b2[5] = 187.f / 2100.f;
        t.b2[6] = 1.f / 40.f;
        break;
    default: fprintf(stderr, "libgpukern: unknown solver type %d\n", solver); abort();
    }
    return t;
}

const double kMu0 = 4.0e-7 * 3.14159265358979323846;
const double kGamma0 = 2.211e5;

}  // namespace

// device copy of the applied field table (static value in the table itself)
// The pointwise field table lives in one buffer per simulation, replaced in
// place (batches keep copies of their own), grown by doubling when a longer
// table arrives.
static void upload_field(mmSim* s) {
    const int n = s->p.nField;
    s->field = FieldTable{};
    s->field.n = n;
    if (n == 1)
        for (int c = 0; c < 3; c++) s->field.b0[c] = s->fieldB[c];
    if (n >= 2) {
        if (s->s) GK_CHECK(hipStreamSynchronize(s->s));  // queued trials of this simulation read the old table
        if (n > s->fieldCap) {
            if (s->fT) {
                cached_free(s->fT);
                cached_free(s->fB);
                s->bytes -= s->fieldCap * (sizeof(double) + 3 * sizeof(float));
            }
            const int cap = std::max(n, 2 * s->fieldCap);
            s->fT = (double*)cached_malloc(cap * sizeof(double));
            s->fB = (float*)cached_malloc(3 * cap * sizeof(float));
            s->fieldCap = cap;
            s->bytes += cap * (sizeof(double) + 3 * sizeof(float));
        }
        GK_CHECK(hipMemcpyAsync(s->fT, s->fieldT.data(), n * sizeof(double), hipMemcpyHostToDevice, s->s));
        GK_CHECK(hipMemcpyAsync(s->fB, s->fieldB.data(), 3 * n * sizeof(float), hipMemcpyHostToDevice, s->s));
        GK_CHECK(hipStreamSynchronize(s->s));
        s->field.t = s->fT;
        s->field.b = s->fB;
    }
}

void* gk::dmalloc(mmSim* s, size_t bytes) {
    void* p = cached_malloc(bytes);
    s->bytes += bytes;
    return p;
}

// applied field at internal time t; t_search (internal) is the latest time the
// field was evaluated at, t included (the forward-only search of simfield.go:126-133)
static void applied_field_host(const mmSim* s, double t, double t_search, float out[3]) {
    if (s->field.fn) {
        field_function(s->field.fn, s->field.fp, s->field.fb, t * s->unitTime, out);
        return;
    }
    const float* b = s->fieldB.data();
    applied_field_at(s->p.nField, s->fieldT.data(), b, b, t * s->unitTime, t_search * s->unitTime, out);
}

EngineState gk::read_state(mmSim* s) {
    EngineState h;
    GK_CHECK(hipMemcpyAsync(&h, s->st, sizeof(h), hipMemcpyDeviceToHost, s->s));
    GK_CHECK(hipStreamSynchronize(s->s));
    return h;
}

bool gk::needs_after_k0(const mmSim* s) {
    return s->tb.fsal || s->p.minDm != 0 || s->p.maxDm != 0 || s->p.minDt != 0 || s->p.maxDt != 0;
}

CtrlParams gk::ctrl_params(const mmSim* s) {
    CtrlParams c{};
    c.part_k0 = s->part;
    c.part_k6 = s->part + s->npart;
    c.part_err = s->part + 2 * s->npart;
    c.n_k = s->nk_part;
    // per-block errors: from the last stage's epilogue (adaptive FSAL) or the final pass
    c.n_err = s->tb.fsal_final() ? s->nk_part : kRedBlocks;
    c.part_e = s->wantEnergy ? s->epart : nullptr;
    c.cmax = s->tb.cmax(s->tb.S - 1);
    c.cell_volume = s->cell[0] * s->cell[1] * s->cell[2];
    c.fold_k0 = !needs_after_k0(s);
    c.adaptive = s->tb.adaptive;
    c.maxError = (float)s->p.maxError;
    c.errorOrder = s->tb.order;
    c.inputDt = (float)(s->p.dt / s->unitTime);
    c.minDm = (float)s->p.minDm;
    c.maxDm = (float)s->p.maxDm;
    c.minDt = (float)(s->p.minDt / s->unitTime);
    c.maxDt = (float)(s->p.maxDt / s->unitTime);
    c.log_dt = s->log_dt;
    c.log_accept = s->log_acc;
    c.log_err = s->log_err;
    c.log_cap = s->log_cap;
    c.thermal = s->p.temperature > 0;
    return c;
}

TorqueArgs gk::torque_args(const mmSim* s, float c, float* part) {
    TorqueArgs a{};
    a.alpha = (float)s->p.alpha;
    a.dtg = (float)(1.0 / (1.0 + s->p.alpha * s->p.alpha));
    a.alphaMap = s->alphaMap;
    a.normMap = s->normMap;
    a.anisType = s->p.anisType;
    a.kAnis = s->kAnis;
    for (int i = 0; i < 3; i++) a.u[i] = s->u[i];
    a.field = s->field;
    a.unitField = (float)s->unitField;
    a.unitTime = s->unitTime;
    a.c = c;
    a.part = part;
    a.thermalK = s->thermalK;
    a.noiseSeed = s->p.seed;
    return a;
}

// ---------------------------------------------------------------- fused mode
static void enqueue_trial(mmSim* s) {
    const Tableau& t = s->tb;
    const CtrlParams cp = ctrl_params(s);
    Prof& pf = s->prof;
    pf.begin(7, s->s);
    launch_ctrl_begin(s->st, s->s);
    pf.end(s->s);
    for (int i = 0; i < t.S; i++) {
        if (i == 0 && t.fsal && s->fsalHost) {
            pf.begin(7, s->s);
            launch_copy_if_active(s->k[0], s->k[t.S - 1], s->N3, s->st, s->s);  // rk.go:311-315
            launch_ctrl_after_k0(s->st, 1, cp, s->s);
            pf.end(s->s);
            continue;
        }
        StageIn si{};
        si.m = s->m;
        si.st = s->st;
        for (int j = 0; j < t.S; j++)
            if (t.a[i][j] != 0.f) {
                si.k[si.nk] = s->k[j];
                si.a[si.nk] = t.a[i][j];
                si.nk++;
            }
        // the torque runs in the epilogue of the Z inverse pass, which rebuilds m_i
        // from m and the k_j; the last stage of a fixed-step tableau also forms
        // and normalises the new m there (rk.go:381-390)
        const bool fin = !t.adaptive && i == t.S - 1;
        // adaptive FSAL: the last stage's epilogue also forms m_high and the error (rk.go:381-403)
        const bool efin = t.fsal_final() && i == t.S - 1;
        EpiArgs ea{};
        ea.m = s->m;
        for (int j = 0; j < i; j++)
            if (t.a[i][j] != 0.f || (fin && t.b[j] != 0.f) || (efin && t.b2[j] != t.b[j])) {
                ea.k[ea.nu] = s->k[j];
                ea.a[ea.nu] = t.a[i][j];
                ea.b[ea.nu] = fin ? t.b[j] : 0.f;
                ea.ce[ea.nu] = efin ? (double)t.b2[j] - (double)t.b[j] : 0.0;
                ea.nu++;
            }
        ea.fin = fin ? 1 : efin ? 2 : 0;
        if (!fin && !efin && ea.nu >= 2) {
            // the Z forward pass stores m_i, which the epilogue reads instead of
            // rebuilding it from m and nu earlier stages (same fmaf chain, same bits)
            si.mout = s->mi;
            ea.m = s->mi;
            ea.nu = 0;
        }
        if (efin) {
            ea.ce_self = (double)t.b2[i] - (double)t.b[i];
            ea.mhigh = s->mtmp;
            ea.part_err = s->part + 2 * s->npart;
        }
        ea.bs = fin ? t.b[i] : 0.f;
        ea.out = fin ? s->m : s->k[i];
        // max|k| is needed for k0 (rk.go:353) and, with FSAL, for the last stage (next k0)
        float* part = i == 0 ? s->part : (t.fsal && i == t.S - 1) ? s->part + s->npart : nullptr;
        ea.t = torque_args(s, t.c[i], part);
        ea.t.cmax = t.cmax(i);
        // wantEnergy: the energy of the new state (FSAL, last stage) or of the step's start
        // (rk.go:324-333)
        if (s->wantEnergy && i == (t.fsal ? t.S - 1 : 0)) ea.epart = s->epart;
        conv_apply(s->cp, si, nullptr, s->s, &pf, &ea);
        if (i == 0 && needs_after_k0(s)) {
            pf.begin(7, s->s);
            launch_ctrl_after_k0(s->st, 0, cp, s->s);
            pf.end(s->s);
        }
    }
    FinalArgs fa{};
    fa.S = t.S;
    for (int i = 0; i < 7; i++) {
        fa.b[i] = t.b[i];
        fa.b2[i] = t.b2[i];
    }
    fa.adaptive = t.adaptive;
    fa.part_err = s->part + 2 * s->npart;
    fa.normMap = s->normMap;
    const float* kk[7];
    for (int i = 0; i < 7; i++) kk[i] = s->k[i] ? s->k[i] : s->k[0];
    if (t.adaptive && !t.fsal_final()) {
        pf.begin(6, s->s);
        launch_rk_final(s->m, kk, s->mtmp, s->st, fa, s->N, s->s);
        pf.end(s->s);
    }
    pf.begin(7, s->s);
    launch_ctrl_end(s->st, cp, s->s);
    if (t.adaptive) launch_commit(s->m, s->mtmp, s->st, s->normMap, s->N, s->s);
    pf.end(s->s);
    if (t.fsal) s->fsalHost = true;
}

static long fused_run(mmSim* s, long long nsteps, double stop, bool wait) {
    EngineState h = read_state(s);
    s->fsalHost = h.fsal_init != 0;
    const long long steps0 = h.steps;
    launch_set_target(s->st, nsteps, stop, s->s);
    if (!wait) {
        // enqueue exactly nsteps trials (fixed-step tableaux never reject)
        for (long long i = 0; i < nsteps; i++) enqueue_trial(s);
        return nsteps;
    }
    int batch = 4;
    while (true) {
        for (int i = 0; i < batch; i++) enqueue_trial(s);
        h = read_state(s);
        if (h.nan_dt) {
            fprintf(stderr, "libgpukern: dt = NaN (rk.go:448-452)\n");
            abort();
        }
        if (h.nonfinite) {
            fprintf(stderr, "libgpukern: the torque or the error estimate is NaN or infinite at step %lld\n", h.steps);
            abort();
        }
        if (h.done) break;
        if (batch < 64) batch *= 2;
    }
    return (long)(h.steps - steps0);
}

// ---------------------------------------------------------------- ABI mode (reference call sequence)
static void abi_calc_heff(mmSim* s, float* m, float* h) {
    // conv.go:57-100
    const int N = s->N;
    for (int c = 0; c < 3; c++) gpuFFT3dPlan_forward(s->plan, m + c * N, s->abuf[c]);
    const int nReal = s->Pabi[0] * s->Pabi[1] * (s->Pabi[2] + 2);
    if (s->Pabi[0] > 1)
        gpu_kernelmul6(s->abuf[0], s->abuf[1], s->abuf[2], s->akern[0], s->akern[1], s->akern[2], s->akern[3],
                       s->akern[4], s->akern[5], nReal);
    else
        gpu_kernelmul4(s->abuf[0], s->abuf[1], s->abuf[2], s->akern[0], s->akern[1], s->akern[2], s->akern[3],
                       nReal);
    for (int c = 0; c < 3; c++) gpuFFT3dPlan_inverse(s->plan, s->abuf[c], h + c * N);
    // addLocalFields (field.go:36-61)
    float b[3];
    s->fieldTmaxHost = std::max(s->fieldTmaxHost, s->htime);
    applied_field_host(s, s->htime, s->fieldTmaxHost, b);
    float hext[3];
    for (int c = 0; c < 3; c++) hext[c] = b[c] / (float)s->unitField;
    float kk = s->kAnis;
    float ax[3] = {s->u[0], s->u[1], s->u[2]};
    gpu_add_local_fields(m, h, N, hext, nullptr, s->p.anisType, &kk, ax);
}

static bool has_current(const mmParams& p) { return p.current[0] != 0 || p.current[1] != 0 || p.current[2] != 0; }
static bool has_current(const mmSim* s) { return s->curOn || has_current(s->p); }

// s.input.j for this torque: the applied current at the stage time when there is
// one (torque.go:42-44, simfield.go:87-260), else CurrentDensity's value
static void current_at(mmSim* s, float j[3]) {
    if (!s->curOn) {
        for (int i = 0; i < 3; i++) j[i] = (float)s->p.current[i];
        return;
    }
    s->curTmaxHost = std::max(s->curTmaxHost, s->htime);
    const double t = s->htime * s->unitTime;
    if (s->curFn) {
        field_function(s->curFn, s->curFp, s->curFb, t, j);
        return;
    }
    const float* v = s->curV.data();
    applied_field_at(s->curN, s->curT.data(), v, v, t, s->curTmaxHost * s->unitTime, j);
}

// SpintorqueDeltaM (torque.go:64-91) with dt = 1, in the reference's float32:
// b = P muB / (e (1 + xi^2)), beta = b (1 + alpha xi), eps = b (xi - alpha),
// u = j / (2 cell), muB, e and j in internal units (material.go:83-110)
static void abi_spintorque(mmSim* s, float* m, float* h) {
    const float Ms = (float)s->p.msat, L = (float)s->unitLength, T = (float)s->unitTime;
    const float alpha = (float)s->p.alpha;
    const float dtGilb = 1.f / (1.f + alpha * alpha);
    const float muB = 9.2740091523E-24f / (Ms * (L * L * L));  // init.go:84 / UnitMoment
    const float e = 1.60217646E-19f / ((Ms * L) * T);          // init.go:85 / UnitCharge
    const float P = (float)s->p.spinPol, xi = (float)s->p.xi;
    const float b = P * muB / (e * 1 * (1 + xi * xi));
    const float beta = b * (1 + alpha * xi), eps = b * (xi - alpha);
    float j[3], u[3];
    current_at(s, j);
    for (int i = 0; i < 3; i++) u[i] = 0.5f * (j[i] / (Ms / L)) / (float)s->cell[i];
    gpu_spintorque_deltaM(m, h, alpha, beta, eps, u, s->jMask, dtGilb, s->p.size[0], s->p.size[1], s->p.size[2]);
}

static void abi_torque(mmSim* s, float* m, float* h) {
    // torque.go:33-53 with dt = 1; the spin-transfer form with a current (:40-49)
    if (has_current(s)) {
        abi_spintorque(s, m, h);
        return;
    }
    const float alpha = (float)s->p.alpha;
    gpu_deltaM(m, h, alpha, s->alphaMap, 1.f / (1.f + alpha * alpha), s->N);
}

// Sim.Normalize (sim.go:230-239): to the norm map when there is one
static void normalize_abi(mmSim* s, float* m) {
    if (s->normMap)
        gpu_normalize_map(m, s->normMap, s->N);
    else
        gpu_normalize_uniform(m, s->N);
}

static bool semianal(const mmSim* s) { return s->p.solver == MM_SEMIANAL1 || s->p.solver == MM_SEMIANAL2; }

// SemiAnal1.step / SemiAnal2.step (semianal.go:23-58, 87-140) with Step's counter
// (simrun.go:76-77): m2 (am1) carries over between steps; the time is not advanced
static void abi_semianal_step(mmSim* s) {
    const int N = s->N, n3 = (int)s->N3;
    const float dt = (float)(s->p.dt / s->unitTime), alpha = (float)s->p.alpha;
    float *m1 = s->m, *m2 = s->am1, *m3 = s->k[1], *h = s->k[0];
    if (s->hsteps == 0) {
        abi_calc_heff(s, m1, h);
        gpu_anal_fw_step(dt / 2.f, alpha, N, m1, m2, h);
        abi_calc_heff(s, m2, h);
        gpu_anal_fw_step(dt, alpha, N, m1, m1, h);
    } else if (s->p.solver == MM_SEMIANAL1) {
        abi_calc_heff(s, m1, h);
        gpu_anal_fw_step(dt, alpha, N, m2, m2, h);
        abi_calc_heff(s, m2, h);
        gpu_anal_fw_step(dt, alpha, N, m1, m1, h);
    } else {
        abi_calc_heff(s, m1, h);
        gpu_anal_fw_step(dt, alpha, N, m2, m3, h);
        gpu_linear_combination(m3, m2, 0.5f, 0.5f, n3);
        abi_calc_heff(s, m3, h);
        gpu_anal_fw_step(dt, alpha, N, m2, m2, h);
        abi_calc_heff(s, m2, h);
        gpu_anal_fw_step(dt, alpha, N, m1, m3, h);
        gpu_linear_combination(m3, m1, 0.5f, 0.5f, n3);
        abi_calc_heff(s, m3, h);
        gpu_anal_fw_step(dt, alpha, N, m1, m1, h);
    }
    if (s->hsteps % 100 == 0) {
        normalize_abi(s, m1);
        normalize_abi(s, m2);
    }
    s->hsteps++;
}

static void abi_step(mmSim* s) {
    if (semianal(s)) {
        abi_semianal_step(s);
        return;
    }
    // rk.go:279-445
    const Tableau& t = s->tb;
    const int order = t.S;
    const int n3 = (int)s->N3;
    float* m = s->m;
    float* m1 = s->am1;
    const double time0 = s->htime;
    memcpy_on_gpu(m, s->abackup, n3);
    bool good = false;
    int trials = 0;
    while (!good && trials < kMaxStepTrials) {
        trials++;
        s->htrials++;
        for (int i = 0; i < order; i++) {
            if (t.fsal && i == 0 && s->hfsal) {
                memcpy_on_gpu(s->k[order - 1], s->k[0], n3);
            } else {
                s->htime = time0 + (double)(t.c[i] * s->hdt);
                memcpy_on_gpu(m, m1, n3);
                for (int j = 0; j < order; j++)
                    if (t.a[i][j] != 0.f) gpu_madd(m1, s->hdt * t.a[i][j], s->k[j], n3);
                abi_calc_heff(s, m1, s->k[i]);
                abi_torque(s, m1, s->k[i]);
                s->hfsal = true;
            }
            if (i == 0) {
                if (!t.adaptive) s->hdt = (float)(s->p.dt / s->unitTime);
                float maxTorque = gpu_reduce(REDUCE_MAXABS, s->k[0], s->redDev, s->redHost.data(), s->redBlocks,
                                             s->redThreads, n3);
                s->htorque = maxTorque;
                float dm = s->hdt * maxTorque;
                if (dm < (float)s->p.minDm) s->hdt = (float)s->p.minDm / maxTorque;
                if (s->hdt * s->unitTime < s->p.minDt) s->hdt = (float)(s->p.minDt / s->unitTime);
                if (s->p.maxDm != 0 && dm > (float)s->p.maxDm) s->hdt = (float)s->p.maxDm / maxTorque;
                if (s->p.maxDt != 0 && s->hdt * s->unitTime > s->p.maxDt) s->hdt = (float)(s->p.maxDt / s->unitTime);
                GK_ASSERT(!std::isnan(s->hdt), "dt = NaN");
            }
        }
        const float dtu = s->hdt;
        s->htime = time0 + (double)s->hdt;
        if (t.adaptive) {
            memcpy_on_gpu(m, m1, n3);
            for (int i = 0; i < order; i++)
                if (t.b2[i] != 0.f) gpu_madd(m1, t.b2[i] * s->hdt, s->k[i], n3);
        }
        for (int i = 0; i < order; i++)
            if (t.b[i] != 0.f) gpu_madd(m, t.b[i] * s->hdt, s->k[i], n3);
        bool bad = false;
        if (t.adaptive) {
            gpu_madd(m1, -1.f, m, n3);
            float err = gpu_reduce(REDUCE_MAXABS, m1, s->redDev, s->redHost.data(), s->redBlocks, s->redThreads, n3);
            s->herr = err;
            float factor = (float)pow((double)((float)s->p.maxError / err), 1.0 / (double)t.order);
            if (factor > 2.f) factor = 2.f;
            if (factor < 0.01f) factor = 0.01f;
            s->hdt = s->hdt * factor;
            GK_ASSERT(!std::isnan(s->hdt), "dt = NaN");
            if (err > 2.f * (float)s->p.maxError) {
                memcpy_on_gpu(s->abackup, m, n3);
                bad = true;
                s->hrejects++;
            } else {
                if (s->hdt * s->unitTime < s->p.minDt) s->hdt = (float)(s->p.minDt / s->unitTime);
                if (s->p.maxDt != 0 && s->hdt * s->unitTime > s->p.maxDt) s->hdt = (float)(s->p.maxDt / s->unitTime);
            }
        }
        good = !bad;
        s->hlog_dt.push_back(dtu);
        s->hlog_acc.push_back(!bad);
        s->hlog_err.push_back(t.adaptive ? s->herr : 0.f);
    }
    normalize_abi(s, m);  // rk.go:444
    s->hsteps++;
}

// the convolution kernel of this simulation for a cell size in internal units
static KernelSpec kernel_spec(const mmSim* s, const double cell[3]) {
    KernelSpec ks{};
    for (int i = 0; i < 3; i++) ks.cell[i] = cell[i];
    ks.accuracy = s->p.demagAccuracy > 0 ? s->p.demagAccuracy : 8;
    ks.exch = s->p.exchType == 6 && s->p.exchInConv;
    ks.type = s->p.demagKernel;
    for (int i = 0; i < 3; i++) {
        ks.period[i] = s->p.periodic[i] ? s->p.size[i] : 0;
        ks.images[i] = s->p.periodic[i];
    }
    return ks;
}

static void abi_init(mmSim* s, const double cellInt[3]) {
    // initConv (init.go:234-274) + NewConv/loadKernel6 (conv.go:35-167) through the drop-in API
    int size[3] = {s->p.size[0], s->p.size[1], s->p.size[2]};
    for (int i = 0; i < 3; i++) s->Pabi[i] = size[i] > 1 && !s->p.periodic[i] ? 2 * size[i] : size[i];
    s->plan = new_gpuFFT3dPlan_padded(size, s->Pabi);
    const long phys = (long)s->Pabi[0] * s->Pabi[1] * (s->Pabi[2] + 2);
    for (int c = 0; c < 3; c++) s->abuf[c] = new_gpu_array((int)phys);
    // real-space kernel in FP64 on the device (FaceKernel6 + Exch6), rounded to float32;
    // the spectrum of the real-space kernel is taken with the drop-in plan, as loadKernel6 does
    const int nK = s->Pabi[0] == 1 ? 4 : 6;
    const long NP = (long)s->Pabi[0] * s->Pabi[1] * s->Pabi[2];
    std::vector<float> Kr((size_t)6 * NP);
    conv_realspace_kernel(s->Pabi, kernel_spec(s, cellInt), Kr.data());
    gpuFFT3dPlan* kplan = new_gpuFFT3dPlan_padded(s->Pabi, s->Pabi);
    float* devIn = new_gpu_array((int)NP);
    float* devOut = new_gpu_array((int)phys);
    std::vector<float> hostOut(phys);
    const float norm = 1.0f / (float)gpuFFT3dPlan_normalization(kplan);
    const int Q = s->Pabi[2] / 2 + 1;
    const long ksize = (long)s->Pabi[0] * s->Pabi[1] * Q;
    for (int k = 0; k < nK; k++) {
        memcpy_to_gpu(Kr.data() + k * NP, devIn, (int)NP);
        gpuFFT3dPlan_forward(kplan, devIn, devOut);
        memcpy_from_gpu(devOut, hostOut.data(), (int)phys);
        for (long j = 0; j < ksize; j++) hostOut[j] = hostOut[2 * j] * norm;
        s->akern[k] = new_gpu_array((int)ksize);
        memcpy_to_gpu(hostOut.data(), s->akern[k], (int)ksize);
    }
    free_gpu_array(devIn);
    free_gpu_array(devOut);
    delete_gpuFFT3dPlan(kplan);
    s->am1 = new_gpu_array((int)s->N3);
    s->abackup = new_gpu_array((int)s->N3);
    // Reductor init (redux.go:103-121)
    s->redThreads = gpu_maxthreads() / 2;
    if (s->redThreads == 0) s->redThreads = 1;
    while (s->N3 <= s->redThreads) s->redThreads /= 2;
    s->redBlocks = (int)((s->N3 + s->redThreads * 2 - 1) / (s->redThreads * 2));
    s->redDev = new_gpu_array(s->redBlocks);
    s->redHost.resize(s->redBlocks);
}

// ---------------------------------------------------------------- C ABI
extern "C" {

// thermal field: sigma(dt) = sqrt(2 alpha kB T / (gamma0 mu0 Ms V dt_SI)) / Ms (temperature.go:29-36),
// dt_SI = dt UnitTime, so sigma = thermalK / sqrt(dt) in internal units
static void set_thermal(mmSim* s, double T) {
    GK_ASSERT(T >= 0, "temperature %g K", T);
    s->p.temperature = T;
    s->thermalK = 0.f;
    if (T > 0) {
        const double kB = 1.380650424e-23, gamma0 = 2.211e5, mu0 = 4e-7 * M_PI;  // material.go:27, init.go:82-83
        const double V = s->p.cellsize[0] * s->p.cellsize[1] * s->p.cellsize[2];
        s->thermalK = (float)(std::sqrt(2 * s->p.alpha * kB * T / (gamma0 * mu0 * s->p.msat * V * s->unitTime)) /
                              s->p.msat);
    }
}

mmSim* mm_sim_new(const mmParams* p, const float* m0, int mode) {
    mmSim* s = new mmSim;
    s->p = *p;
    s->mode = mode;
    s->s = stream();
    GK_ASSERT(p->size[0] > 0 && p->size[1] > 0 && p->size[2] > 1, "bad grid size %d %d %d", p->size[0], p->size[1],
              p->size[2]);
    for (int i = 0; i < 3; i++) {
        GK_ASSERT(p->periodic[i] >= 0, "periodic images must be >= 0, got %d", p->periodic[i]);
        // the fused engine stores a quarter of the kernel spectrum and restores the rest from
        // the kernel's parity along each axis; an odd period has an asymmetric image range
        // (demag.go:42-53), whose kernel has no such parity
        GK_ASSERT(!p->periodic[i] || mode != MM_MODE_FUSED || p->size[i] % 2 == 0 || p->size[i] == 1,
                  "periodic axis %d of odd size %d: use MM_MODE_ABI", i, p->size[i]);
        GK_ASSERT(!p->periodic[i] || p->demagKernel == MM_KERNEL_FACE6,
                  "periodic boundaries use the FaceKernel6 demag kernel (demag.go:26-84)");
    }
    GK_ASSERT(p->msat > 0 && p->aexch > 0, "msat and aexch must be set (init.go:87-97)");
    GK_ASSERT(p->wantDemag, "the engine needs demag on");
    GK_ASSERT(p->demagKernel == MM_KERNEL_FACE6 || p->demagKernel == MM_KERNEL_NEWELL, "unknown demag kernel %d",
              p->demagKernel);
    GK_ASSERT(p->exchType == 0 || (p->exchType == 6 && p->exchInConv),
              "exchange must be 0 or 6-neighbour folded into the convolution");
    s->N = p->size[0] * p->size[1] * p->size[2];
    s->N3 = 3L * s->N;
    // units (material.go:36-110)
    s->unitLength = sqrt(2.0 * p->aexch / (kMu0 * p->msat * p->msat));
    s->unitTime = 1.0 / (kGamma0 * p->msat);
    s->unitField = kMu0 * p->msat;
    s->unitEnergy = p->aexch * s->unitLength;
    s->unitEdens = s->unitEnergy / (s->unitLength * s->unitLength * s->unitLength);
    for (int i = 0; i < 3; i++) s->cell[i] = p->cellsize[i] / s->unitLength;
    s->kAnis = (float)(p->k1 / s->unitEdens);
    GK_ASSERT(p->temperature == 0 || mode == MM_MODE_FUSED, "a finite temperature needs the fused mode");
    if (has_current(*p)) {
        GK_ASSERT(mode == MM_MODE_ABI, "spin-transfer torque runs in MM_MODE_ABI");
        GK_ASSERT(p->currentMask, "Spin-transfer torque requires a currentMask to be set (torque.go:68-71)");
        GK_ASSERT(!p->alphaMap, "spin-transfer torque with a space-dependent alpha (torque.go:66)");
    }
    set_thermal(s, p->temperature);
    for (int i = 0; i < 3; i++) s->u[i] = (float)p->anisAxis[i];
    s->tb = make_tableau(p->solver);
    GK_ASSERT(!(p->solver == MM_SEMIANAL1 || p->solver == MM_SEMIANAL2) || mode == MM_MODE_ABI,
              "the semi-analytic solvers run in MM_MODE_ABI");
    GK_ASSERT(!s->tb.adaptive || p->maxError > 0, "maxError must be > 0 for an adaptive solver");
    GK_ASSERT(s->tb.adaptive || p->dt > 0, "a fixed-step solver needs dt > 0");
    if (p->nField > 0) {
        s->fieldT.assign(p->fieldT, p->fieldT + p->nField);
        s->fieldB.resize(3 * p->nField);
        for (int i = 0; i < 3 * p->nField; i++) s->fieldB[i] = (float)p->fieldB[i];
    }
    s->p.fieldT = nullptr;
    s->p.fieldB = nullptr;
    // device buffers
    s->m = (float*)dmalloc(s, s->N3 * sizeof(float));
    // set-up copies go to this simulation's stream: synchronous copies would wait for
    // the work other simulations and batches have queued on their own streams
    GK_CHECK(hipMemcpyAsync(s->m, m0, s->N3 * sizeof(float), hipMemcpyHostToDevice, s->s));
    for (int i = 0; i < s->tb.S; i++) {
        s->k[i] = (float*)dmalloc(s, s->N3 * sizeof(float));
        GK_CHECK(hipMemsetAsync(s->k[i], 0, s->N3 * sizeof(float), s->s));
    }
    s->mi = (float*)dmalloc(s, s->N3 * sizeof(float));
    s->mtmp = (float*)dmalloc(s, s->N3 * sizeof(float));
    s->st = (EngineState*)dmalloc(s, sizeof(EngineState));
    s->log_dt = (double*)dmalloc(s, s->log_cap * sizeof(double));
    s->log_acc = (int*)dmalloc(s, s->log_cap * sizeof(int));
    s->log_err = (float*)dmalloc(s, s->log_cap * sizeof(float));
    s->dsum = (double*)dmalloc(s, 4 * sizeof(double));
    EngineState h{};
    h.dt = (float)(p->dt / s->unitTime);  // init.go:280-284
    if (h.dt == 0.f) h.dt = 1e-4f;
    h.new_step = 1;
    h.dt_noise = h.dt;
    h.noise_step = 0;
    h.stop = 0;
    h.done = 1;
    h.field_tmax = -INFINITY;
    GK_CHECK(hipMemcpyAsync(s->st, &h, sizeof(h), hipMemcpyHostToDevice, s->s));
    GK_CHECK(hipStreamSynchronize(s->s));  // h lives on this frame
    s->hdt = h.dt;
    upload_field(s);
    if (p->alphaMap) {
        s->alphaMap = (float*)dmalloc(s, (size_t)s->N * sizeof(float));
        GK_CHECK(hipMemcpyAsync(s->alphaMap, p->alphaMap, (size_t)s->N * sizeof(float), hipMemcpyHostToDevice, s->s));
    }
    s->p.alphaMap = nullptr;  // the caller's array is not kept
    if (p->normMap) {
        s->normMap = (float*)dmalloc(s, (size_t)s->N * sizeof(float));
        GK_CHECK(hipMemcpyAsync(s->normMap, p->normMap, (size_t)s->N * sizeof(float), hipMemcpyHostToDevice, s->s));
    }
    s->p.normMap = nullptr;
    if (p->currentMask) {
        s->jMask = (float*)dmalloc(s, (size_t)s->N3 * sizeof(float));
        GK_CHECK(hipMemcpyAsync(s->jMask, p->currentMask, (size_t)s->N3 * sizeof(float), hipMemcpyHostToDevice, s->s));
    }
    s->p.currentMask = nullptr;
    // convolution (init.go:234-274)
    conv_plan_init(s->cp, p->size);
    s->bytes += s->cp.bytes;
    s->kernelImagRatio = conv_acquire_kernel(s->cp, kernel_spec(s, s->cell));
    if (s->kernelImagRatio > 1e-4)
        fprintf(stderr, "Warning: FFT Kernel max imaginary part ratio = %g\n", s->kernelImagRatio);
    s->conv = true;
    s->nk_part = conv_epi_blocks(s->cp);
    // room for the partials of every row shape the tuner may choose
    s->npart = std::max(kRedBlocks, conv_epi_blocks_max(s->cp.N));
    s->part = (float*)dmalloc(s, 3 * (size_t)s->npart * sizeof(float));
    GK_CHECK(hipMemsetAsync(s->part, 0, 3 * (size_t)s->npart * sizeof(float), s->s));
    s->epart = (double*)dmalloc(s, (size_t)s->npart * sizeof(double));
    GK_CHECK(hipMemsetAsync(s->epart, 0, (size_t)s->npart * sizeof(double), s->s));
    s->wantEnergy = p->wantEnergy != 0;
    if (mode == MM_MODE_ABI) abi_init(s, s->cell);
    // set-up work ran on the library stream and this simulation's stream (the
    // others, e.g. running batches, are not waited for)
    GK_CHECK(hipStreamSynchronize(stream()));
    GK_CHECK(hipStreamSynchronize(s->s));
    return s;
}

void mm_sim_free(mmSim* s) {
    if (!s) return;
    GK_CHECK(hipStreamSynchronize(s->s));
    for (int i = 0; i < 7; i++)
        cached_free(s->k[i]);
    for (float* p : {s->m, s->mi, s->mtmp, s->fB, s->alphaMap, s->normMap, s->jMask}) cached_free(p);
    for (void* p : {(void*)s->st, (void*)s->log_dt, (void*)s->log_acc, (void*)s->log_err, (void*)s->dsum, (void*)s->fT,
                    (void*)s->part, (void*)s->epart})
        cached_free(p);
    conv_plan_free(s->cp);
    if (s->plan) delete_gpuFFT3dPlan(s->plan);
    for (int c = 0; c < 3; c++)
        if (s->abuf[c]) free_gpu_array(s->abuf[c]);
    for (int k = 0; k < 6; k++)
        if (s->akern[k]) free_gpu_array(s->akern[k]);
    for (float* p : {s->am1, s->abackup, s->redDev})
        if (p) free_gpu_array(p);
    delete s;
}

long mm_sim_run(mmSim* s, double seconds) {
    GK_ASSERT(!semianal(s), "the semi-analytic solvers do not advance the time (semianal.go): use Steps");
    if (s->mode == MM_MODE_ABI) {
        const double stop = s->htime + seconds / s->unitTime;
        normalize_abi(s, s->m);  // simrun.go:24
        long n = 0;
        while (s->htime < stop) {
            abi_step(s);
            n++;
        }
        return n;
    }
    EngineState h = read_state(s);
    const double stop = h.time + seconds / s->unitTime;
    launch_normalize_active(s->m, nullptr, s->normMap, s->N, s->s);  // simrun.go:24
    return fused_run(s, -1, stop, true);
}

long mm_sim_run_until(mmSim* s, double stop) {
    GK_ASSERT(!semianal(s), "the semi-analytic solvers do not advance the time (semianal.go): use Steps");
    if (s->mode == MM_MODE_ABI) {
        long n = 0;
        while (s->htime < stop) {
            abi_step(s);
            n++;
        }
        return n;
    }
    return fused_run(s, -1, stop, true);
}

void mm_sim_steps(mmSim* s, int n) {
    if (s->mode == MM_MODE_ABI) {
        for (int i = 0; i < n; i++) abi_step(s);
        return;
    }
    fused_run(s, n, INFINITY, true);
}

void mm_sim_steps_async(mmSim* s, int n) {
    if (s->mode == MM_MODE_ABI) {
        mm_sim_steps(s, n);
        return;
    }
    GK_ASSERT(!s->tb.adaptive, "mm_sim_steps_async needs a fixed-step solver (no rejected trials)");
    fused_run(s, n, INFINITY, false);
}

void mm_sim_sync(mmSim* s) { GK_CHECK(hipStreamSynchronize(s->s)); }

void mm_sim_get_m(mmSim* s, float* out) {
    GK_CHECK(hipMemcpyAsync(out, s->m, s->N3 * sizeof(float), hipMemcpyDeviceToHost, s->s));
    GK_CHECK(hipStreamSynchronize(s->s));
}

void mm_sim_set_m(mmSim* s, const float* in) {
    GK_CHECK(hipMemcpyAsync(s->m, in, s->N3 * sizeof(float), hipMemcpyHostToDevice, s->s));
    GK_CHECK(hipStreamSynchronize(s->s));
}

float* mm_sim_device_m(mmSim* s) { return s->m; }

void mm_sim_state(mmSim* s, double* st) {
    if (s->mode == MM_MODE_ABI) {
        st[0] = s->htime;
        st[1] = s->hdt;
        st[2] = (double)s->hsteps;
        st[3] = s->herr;
        st[4] = s->htorque;
        st[5] = (double)s->htrials;
        st[6] = (double)s->hrejects;
        st[7] = 0;
    } else {
        EngineState h = read_state(s);
        st[0] = h.time;
        st[1] = h.dt;
        st[2] = (double)h.steps;
        st[3] = h.err;
        st[4] = h.max_torque;
        st[5] = (double)h.trials_total;
        st[6] = (double)h.rejects_total;
        st[7] = h.energy;
    }
    st[8] = s->unitTime;
    st[9] = s->unitField;
    for (int i = 0; i < 3; i++) st[10 + i] = s->mode == MM_MODE_ABI ? s->Pabi[i] : s->cp.P[i];
}

int mm_sim_trial_log(mmSim* s, double* dt, int* accept, int max) {
    if (s->mode == MM_MODE_ABI) {
        int n = (int)std::min<size_t>(max, s->hlog_dt.size());
        size_t off = s->hlog_dt.size() - n;
        for (int i = 0; i < n; i++) {
            dt[i] = s->hlog_dt[off + i];
            accept[i] = s->hlog_acc[off + i];
        }
        return n;
    }
    EngineState h = read_state(s);
    int n = std::min(h.log_n, s->log_cap);
    n = std::min(n, max);
    if (n <= 0) return 0;
    GK_CHECK(hipMemcpy(dt, s->log_dt, n * sizeof(double), hipMemcpyDeviceToHost));
    GK_CHECK(hipMemcpy(accept, s->log_acc, n * sizeof(int), hipMemcpyDeviceToHost));
    return n;
}

int mm_sim_trial_errors(mmSim* s, float* err, int max) {
    if (s->mode == MM_MODE_ABI) {
        int n = (int)std::min<size_t>(max, s->hlog_err.size());
        size_t off = s->hlog_err.size() - n;
        for (int i = 0; i < n; i++) err[i] = s->hlog_err[off + i];
        return n;
    }
    EngineState h = read_state(s);
    int n = std::min(std::min(h.log_n, s->log_cap), max);
    if (n <= 0) return 0;
    GK_CHECK(hipMemcpy(err, s->log_err, n * sizeof(float), hipMemcpyDeviceToHost));
    return n;
}

static void eval_heff(mmSim* s, const float* m_host, float* h_dev) {
    GK_CHECK(hipMemcpyAsync(s->mi, m_host, s->N3 * sizeof(float), hipMemcpyHostToDevice, s->s));
    if (s->mode == MM_MODE_ABI) {
        GK_CHECK(hipStreamSynchronize(s->s));
        abi_calc_heff(s, s->mi, h_dev);
        return;
    }
    StageIn si{};
    si.m = s->mi;
    conv_apply(s->cp, si, h_dev, s->s);
    EngineState h = read_state(s);
    float b[3], hext[3];
    applied_field_host(s, h.time, std::max(h.field_tmax, h.time), b);
    for (int c = 0; c < 3; c++) hext[c] = b[c] / (float)s->unitField;
    float kk = s->kAnis;
    float ax[3] = {s->u[0], s->u[1], s->u[2]};
    gpu_add_local_fields(s->mi, h_dev, s->N, hext, nullptr, s->p.anisType, &kk, ax);
    // the thermal field of the current step (field.go:74-80)
    launch_add_thermal(h_dev, s->st, torque_args(s, 0.f, nullptr), s->N, s->s);
    GK_CHECK(hipStreamSynchronize(s->s));
}

void mm_sim_heff(mmSim* s, const float* m, float* h) {
    eval_heff(s, m, s->mtmp);
    GK_CHECK(hipMemcpy(h, s->mtmp, s->N3 * sizeof(float), hipMemcpyDeviceToHost));
}

void mm_sim_torque(mmSim* s, const float* m, float* k) {
    eval_heff(s, m, s->mtmp);
    if (has_current(s)) {
        abi_spintorque(s, s->mi, s->mtmp);  // torque.go:40-49
    } else {
        const float alpha = (float)s->p.alpha;
        gpu_deltaM(s->mi, s->mtmp, alpha, s->alphaMap, (float)(1.0 / (1.0 + s->p.alpha * s->p.alpha)), s->N);
    }
    GK_CHECK(hipMemcpy(k, s->mtmp, s->N3 * sizeof(float), hipMemcpyDeviceToHost));
}

double mm_sim_energy(mmSim* s) {
    StageIn si{};
    si.m = s->m;
    conv_apply(s->cp, si, s->mtmp, s->s);
    double t, ts;
    if (s->mode == MM_MODE_ABI) {
        t = s->htime;
        ts = std::max(s->fieldTmaxHost, t);
    } else {
        const EngineState h = read_state(s);
        t = h.time;
        ts = std::max(h.field_tmax, t);
    }
    float b[3], hext[3];
    applied_field_host(s, t, ts, b);
    for (int c = 0; c < 3; c++) hext[c] = b[c] / (float)s->unitField;
    GK_CHECK(hipMemsetAsync(s->dsum, 0, sizeof(double), s->s));
    launch_energy(s->m, s->mtmp, hext, s->p.anisType, s->kAnis, s->u, s->N, s->dsum, s->s);
    double e = 0;
    GK_CHECK(hipMemcpyAsync(&e, s->dsum, sizeof(double), hipMemcpyDeviceToHost, s->s));
    GK_CHECK(hipStreamSynchronize(s->s));
    return e * s->cell[0] * s->cell[1] * s->cell[2];
}

void mm_sim_average_m(mmSim* s, double* avg) {
    GK_CHECK(hipMemsetAsync(s->dsum, 0, 3 * sizeof(double), s->s));
    launch_sum3(s->m, s->N, s->dsum, s->s);
    GK_CHECK(hipMemcpyAsync(avg, s->dsum, 3 * sizeof(double), hipMemcpyDeviceToHost, s->s));
    GK_CHECK(hipStreamSynchronize(s->s));
    for (int c = 0; c < 3; c++) avg[c] /= (double)s->N;
}

long mm_sim_kernel_spectrum(mmSim* s, float* out) {
    const int Q = s->cp.Q;
    const long fulls = (long)s->cp.P[0] * s->cp.P[1] * Q;
    const long total = s->cp.nK * fulls;
    if (!out) return total;
    float* dev = nullptr;
    dev = (float*)cached_malloc(total * sizeof(float));
    ConvPlan tmp = s->cp;  // a private spectrum: the engine's may be shared
    tmp.Kc = (float*)cached_malloc(tmp.kc_elems * sizeof(float));
    conv_build_kernel(tmp, kernel_spec(s, s->cell), out, dev);
    cached_free(tmp.Kc);
    cached_free(dev);
    return total;
}

void mm_demag_kernel(const int P[3], const double cellsize[3], int kernel, int accuracy, double* out) {
    GK_ASSERT(P[0] > 0 && P[1] > 0 && P[2] > 0, "bad padded size %d %d %d", P[0], P[1], P[2]);
    GK_ASSERT(kernel == MM_KERNEL_FACE6 || kernel == MM_KERNEL_NEWELL, "unknown demag kernel %d", kernel);
    KernelSpec ks{};
    for (int i = 0; i < 3; i++) ks.cell[i] = cellsize[i];
    ks.accuracy = accuracy > 0 ? accuracy : 8;
    ks.exch = 0;
    ks.type = kernel;
    conv_realspace_kernel_f64(P, ks, out);
}

unsigned long long mm_memory_cached(void) { return cached_bytes(); }
unsigned long long mm_memory_trim(void) { return cached_trim(); }

int mm_kernel_cache_size(void) {
    return conv_kernel_cache_size();
}

unsigned long long mm_sim_device_bytes(mmSim* s) { return s->bytes; }

double mm_sim_profile_steps(mmSim* s, int n, double* ms, int* counts) {
    GK_ASSERT(s->mode == MM_MODE_FUSED, "profiling needs the fused mode");
    GK_ASSERT(!s->tb.adaptive, "profiling needs a fixed-step solver");
    hipEvent_t e0, e1;
    GK_CHECK(hipEventCreate(&e0));
    GK_CHECK(hipEventCreate(&e1));
    s->prof.reset();
    s->prof.on = true;
    GK_CHECK(hipEventRecord(e0, s->s));
    fused_run(s, n, INFINITY, false);
    GK_CHECK(hipEventRecord(e1, s->s));
    GK_CHECK(hipEventSynchronize(e1));
    float total = 0.f;
    GK_CHECK(hipEventElapsedTime(&total, e0, e1));
    s->prof.collect(ms, counts);
    s->prof.on = false;
    s->prof.reset();
    GK_CHECK(hipEventDestroy(e0));
    GK_CHECK(hipEventDestroy(e1));
    return total;
}

unsigned long long mm_sim_footprint(const mmParams* p) {
    const long N3 = 3L * p->size[0] * p->size[1] * p->size[2];
    const Tableau t = make_tableau(p->solver);
    int epi = 0;
    size_t b = conv_plan_footprint(p->size, &epi);
    const int npart = std::max(kRedBlocks, epi);
    b += N3 * sizeof(float) * (size_t)(1 + t.S + 2);  // m, k_0..k_{S-1}, m_i, m_tmp
    b += sizeof(EngineState) + (size_t)(1 << 16) * (sizeof(double) + sizeof(int) + sizeof(float)) + 4 * sizeof(double);
    b += 3 * (size_t)npart * sizeof(float) + (size_t)npart * sizeof(double);  // part, epart
    if (p->nField >= 2) b += (size_t)p->nField * (sizeof(double) + 3 * sizeof(float));
    if (p->alphaMap) b += (size_t)(N3 / 3) * sizeof(float);
    if (p->normMap) b += (size_t)(N3 / 3) * sizeof(float);
    if (p->currentMask) b += (size_t)N3 * sizeof(float);
    return b;
}

void mm_sim_set_field(mmSim* s, int n, const double* t, const double* b) {
    GK_ASSERT(n >= 0 && (n <= 1 || t), "mm_sim_set_field: %d points", n);
    GK_CHECK(hipStreamSynchronize(s->s));
    if (n >= 2 && s->p.nField < 2) {
        // a new pointwise field starts its search at the first point; one that
        // grows an existing table keeps its position (simfield.go:101-110)
        const double t0 = -INFINITY;
        GK_CHECK(hipMemcpy(&s->st->field_tmax, &t0, sizeof(double), hipMemcpyHostToDevice));
        s->fieldTmaxHost = t0;
    }
    s->p.nField = n;
    s->fieldT.assign(t ? t : nullptr, t ? t + n : nullptr);
    s->fieldB.resize(3 * n);
    for (int i = 0; i < 3 * n; i++) s->fieldB[i] = (float)b[i];
    upload_field(s);
}

void mm_sim_set_field_function(mmSim* s, int kind, const double* b, const double* params) {
    GK_ASSERT(kind >= MM_FIELD_PULSE && kind <= MM_FIELD_ROTATING, "unknown field function %d", kind);
    GK_CHECK(hipStreamSynchronize(s->s));
    s->p.nField = 0;
    s->fieldT.clear();
    s->fieldB.clear();
    upload_field(s);
    s->field.fn = kind;
    for (int i = 0; i < 4; i++) s->field.fp[i] = params[i];
    for (int c = 0; c < 3; c++) s->field.fb[c] = (float)b[c];
}

long mm_sim_relax(mmSim* s, long max_steps) {
    GK_ASSERT(s->mode == MM_MODE_FUSED, "mm_sim_relax needs the fused mode");
    s->wantEnergy = true;  // relax.go:16
    launch_set_relax(s->st, 1, s->s);
    const long n = fused_run(s, max_steps > 0 ? max_steps : -1, INFINITY, true);
    launch_set_relax(s->st, 0, s->s);
    GK_CHECK(hipStreamSynchronize(s->s));
    return n;
}

void mm_sim_begin_run(mmSim* s, double seconds) {
    GK_ASSERT(s->mode == MM_MODE_FUSED, "mm_sim_begin_run needs the fused mode");
    EngineState h = read_state(s);
    const double stop = h.time + seconds / s->unitTime;
    launch_normalize_active(s->m, nullptr, s->normMap, s->N, s->s);  // simrun.go:24
    launch_set_target(s->st, -1, stop, s->s);
    GK_CHECK(hipStreamSynchronize(s->s));
}

// n back-to-back launches of pass class cls; the Z forward's stage build reads
// nk earlier stages (the first stage of a step reads none, most others one)
static double time_pass(mmSim* s, int cls, int n, int nk) {
    GK_ASSERT(s->mode == MM_MODE_FUSED && n > 0 && cls >= 0 && cls <= 4, "time_pass: fused mode, class 0..4");
    GK_ASSERT(nk >= 0 && nk < s->tb.S, "time_pass: %d earlier stages", nk);
    StageIn si{};
    si.m = s->m;
    si.st = s->st;
    si.nk = nk;
    for (int j = 0; j < nk; j++) {
        si.k[j] = s->k[1 + j];
        si.a[j] = 0.5f;
    }
    EpiArgs ea{};
    ea.m = s->m;
    ea.out = s->k[0];
    ea.t = torque_args(s, s->tb.c[0], s->part);
    launch_set_active(s->st, s->s);
    hipEvent_t e0, e1;
    GK_CHECK(hipEventCreate(&e0));
    GK_CHECK(hipEventCreate(&e1));
    GK_CHECK(hipEventRecord(e0, s->s));
    for (int i = 0; i < n; i++) conv_run_pass(s->cp, cls, si, nullptr, &ea, s->s);
    GK_CHECK(hipEventRecord(e1, s->s));
    GK_CHECK(hipEventSynchronize(e1));
    float total = 0.f;
    GK_CHECK(hipEventElapsedTime(&total, e0, e1));
    GK_CHECK(hipEventDestroy(e0));
    GK_CHECK(hipEventDestroy(e1));
    return total;
}

double mm_sim_time_pass(mmSim* s, int cls, int n) { return time_pass(s, cls, n, 0); }

int mm_sim_tune(mmSim* s, int reps, int params[7], double ms_default[8], double ms_tuned[8]) {
    GK_ASSERT(s->mode == MM_MODE_FUSED && s->conv, "tune: an initialised fused-mode simulation");
    if (reps < 1) reps = 1;
    ConvPlan& cp = s->cp;
    const bool d3 = cp.P[0] > 1, sep = d3 && !cp.yxy;
    const TuneParams dflt{};
    // ms per launch of classes cls_a (+ cls_b): the least of 3 runs of `reps`
    // back-to-back launches, after 2 untimed launches
    // (the Z forward with one earlier stage in its stage build, as in most stages)
    const int nk0 = s->tb.S >= 2 ? 1 : 0;
    auto timed = [&](int cls_a, int cls_b) {
        double t = 0;
        for (int cls : {cls_a, cls_b}) {
            if (cls < 0) continue;
            const int nk = cls == 0 ? nk0 : 0;
            time_pass(s, cls, 2, nk);
            double best = INFINITY;
            for (int r = 0; r < 3; r++) best = std::min(best, time_pass(s, cls, reps, nk) / reps);
            t += best;
        }
        return t;
    };
    // one tunable: try its candidates, keep the fastest (ties keep the earlier)
    auto tune_one = [&](int slot, int* field, int dval, const std::vector<int>& cands, int cls_a, int cls_b) {
        double t_d = -1, t_best = INFINITY;
        int v_best = dval;
        for (int v : cands) {
            *field = v;
            if (field == &cp.tp.er) s->nk_part = conv_epi_blocks(cp);
            const double t = timed(cls_a, cls_b);
            if (v == dval) t_d = t;
            if (t < t_best) {
                t_best = t;
                v_best = v;
            }
        }
        *field = v_best;
        if (field == &cp.tp.er) s->nk_part = conv_epi_blocks(cp);
        ms_default[slot] = t_d;
        ms_tuned[slot] = t_best;
    };
    for (int i = 0; i < 8; i++) ms_default[i] = ms_tuned[i] = -1;
    std::vector<int> zr{0}, er{0};
    for (int v = 1; v <= 2; v++)
        if (conv_zr_valid(cp.M, v)) zr.push_back(v);
    if (conv_er_valid(cp.M, 1)) er.push_back(1);
    const std::vector<int> grid{2, 3, 4, 6, 8, 12, 16};
    tune_one(0, &cp.tp.zr, dflt.zr, zr, 0, -1);
    tune_one(1, &cp.tp.er, dflt.er, er, 4, -1);
    if (sep) {
        // separate Y and X.K.X passes first (the plane pass off), then the choice between them
        const int pp0 = cp.tp.pp;
        cp.tp.pp = 0;
        tune_one(2, &cp.tp.yb, dflt.yb, grid, 1, 3);
        tune_one(3, &cp.tp.xb, dflt.xb, grid, 2, -1);
        std::vector<int> xv{0};
        for (int v = 1; v <= 3; v++)
            if (conv_xv_valid(cp.P[0], v)) xv.push_back(v);
        if (xv.size() > 1) tune_one(6, &cp.tp.xv, dflt.xv, xv, 2, -1);
        cp.tp.pp = pp0;
        if (cp.plane3) {
            double t[2];
            for (int v = 0; v < 2; v++) {
                cp.tp.pp = v;
                t[v] = v ? timed(2, -1) : timed(1, 3) + timed(2, -1);
            }
            cp.tp.pp = t[1] < t[0] ? 1 : 0;
            ms_default[5] = t[dflt.pp];
            ms_tuned[5] = t[cp.tp.pp];
            ms_default[7] = t[0];  // the three passes
            ms_tuned[7] = t[1];    // the plane pass
        }
    }
    if (!d3) tune_one(4, &cp.tp.y2b, dflt.y2b, grid, 2, -1);
    GK_CHECK(hipStreamSynchronize(s->s));
    conv_tune_store(cp.N, cp.tp);
    if (params) {
        params[0] = cp.tp.zr;
        params[1] = cp.tp.er;
        params[2] = cp.tp.yb;
        params[3] = cp.tp.xb;
        params[4] = cp.tp.y2b;
        params[5] = cp.tp.pp;
        params[6] = cp.tp.xv;
    }
    return 0;
}

void mm_sim_set_temperature(mmSim* s, double temperature) {
    GK_ASSERT(temperature == 0 || s->mode == MM_MODE_FUSED, "a finite temperature needs the fused mode");
    GK_CHECK(hipStreamSynchronize(s->s));
    set_thermal(s, temperature);
}

void mm_sim_set_spintorque(mmSim* s, const double j[3], double spinPol, double xi) {
    const bool on = j[0] != 0 || j[1] != 0 || j[2] != 0;
    GK_ASSERT(!on || s->mode == MM_MODE_ABI, "spin-transfer torque runs in MM_MODE_ABI");
    GK_ASSERT(!on || s->jMask, "Spin-transfer torque requires a currentMask to be set (torque.go:68-71)");
    GK_ASSERT(!on || !s->alphaMap, "spin-transfer torque with a space-dependent alpha (torque.go:66)");
    for (int i = 0; i < 3; i++) s->p.current[i] = j[i];
    s->p.spinPol = spinPol;
    s->p.xi = xi;
}

void mm_sim_set_current(mmSim* s, int n, const double* t, const double* j, int kind, const double* params) {
    const bool on = kind > 0 || n > 0;
    GK_ASSERT(!on || s->mode == MM_MODE_ABI, "spin-transfer torque runs in MM_MODE_ABI");
    GK_ASSERT(!on || s->jMask, "Spin-transfer torque requires a currentMask to be set (torque.go:68-71)");
    GK_ASSERT(!on || !s->alphaMap, "spin-transfer torque with a space-dependent alpha (torque.go:66)");
    GK_ASSERT(kind >= 0 && kind <= MM_FIELD_ROTATING && n >= 0 && (kind || n <= 1 || t), "mm_sim_set_current: %d %d",
              kind, n);
    s->curOn = on;
    s->curFn = kind;
    s->curN = kind ? 0 : n;
    s->curT.assign(t && !kind && n > 1 ? t : nullptr, t && !kind && n > 1 ? t + n : nullptr);
    s->curV.resize(kind ? 0 : 3 * (size_t)n);
    for (size_t i = 0; i < s->curV.size(); i++) s->curV[i] = (float)j[i];
    if (kind) {
        for (int i = 0; i < 4; i++) s->curFp[i] = params[i];
        for (int c = 0; c < 3; c++) s->curFb[c] = (float)j[c];
    }
    s->curTmaxHost = -INFINITY;
}

void mm_sim_get_tune(mmSim* s, int params[7]) {
    params[0] = s->cp.tp.zr;
    params[1] = s->cp.tp.er;
    params[2] = s->cp.tp.yb;
    params[3] = s->cp.tp.xb;
    params[4] = s->cp.tp.y2b;
    params[5] = s->cp.tp.pp;
    params[6] = s->cp.tp.xv;
}

int mm_tune_save(const char* path) { return conv_tune_save(path); }
int mm_tune_load(const char* path) { return conv_tune_load(path); }
void mm_tune_clear(void) { conv_tune_clear(); }

double mm_sim_time_steps(mmSim* s, int n) {
    GK_ASSERT(s->mode == MM_MODE_FUSED && !s->tb.adaptive, "timing needs the fused mode and a fixed-step solver");
    hipEvent_t e0, e1;
    GK_CHECK(hipEventCreate(&e0));
    GK_CHECK(hipEventCreate(&e1));
    GK_CHECK(hipEventRecord(e0, s->s));
    fused_run(s, n, INFINITY, false);
    GK_CHECK(hipEventRecord(e1, s->s));
    GK_CHECK(hipEventSynchronize(e1));
    float total = 0.f;
    GK_CHECK(hipEventElapsedTime(&total, e0, e1));
    GK_CHECK(hipEventDestroy(e0));
    GK_CHECK(hipEventDestroy(e1));
    return total;
}

}  // extern "C"
