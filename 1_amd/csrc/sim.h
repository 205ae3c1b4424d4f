// Host-side simulation state of the engine, shared by the single-simulation
// driver (sim.hip) and the batched driver (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/gpukern.h"
#include "../../include/mumax_engine.h"
#include "engine.h"

namespace gk {

struct Tableau {
    int S = 0, fsal = 0, adaptive = 0;
    float order = 0;
    float a[7][7] = {}, b[7] = {}, b2[7] = {}, c[7] = {};
    // adaptive FSAL tableau whose last stage input is the high-order solution
    // (a_{S-1,j} == b_j): its final combination and error run in that stage's
    // epilogue
    bool fsal_final() const {
        if (!(adaptive && fsal)) return false;
        for (int j = 0; j < S; j++)
            if (a[S - 1][j] != b[j]) return false;
        return true;
    }
    // largest node among stages 0..i (the applied field's search time)
    float cmax(int i) const {
        float m = 0.f;
        for (int j = 0; j <= i; j++) m = c[j] > m ? c[j] : m;
        return m;
    }
};

}  // namespace gk

using gk::ConvPlan;
using gk::EngineState;
using gk::FieldTable;
using gk::Prof;
using gk::Tableau;

struct mmSim {
    mmParams p;
    int mode = MM_MODE_FUSED;
    int N = 0;
    long N3 = 0;
    double unitLength = 0, unitTime = 0, unitField = 0, unitEnergy = 0, unitEdens = 0;
    double cell[3] = {};  // internal units
    float kAnis = 0;
    float thermalK = 0;  // thermal field sigma * sqrt(dt), internal units (0: T = 0)
    float u[3] = {};
    Tableau tb;
    std::vector<double> fieldT;
    std::vector<float> fieldB;
    hipStream_t s = nullptr;
    // device
    float* m = nullptr;
    float* k[7] = {};
    float* mi = nullptr;
    float* mtmp = nullptr;
    EngineState* st = nullptr;
    double* log_dt = nullptr;
    int* log_acc = nullptr;
    float* log_err = nullptr;  // per trial: the error estimate (adaptive) or 0
    int log_cap = 1 << 16;
    double* dsum = nullptr;
    float* part = nullptr;  // 3 x npart: |k0|, |k_last|, error
    int npart = 0;          // max(kRedBlocks, blocks of the fused Z inverse + torque pass)
    int nk_part = 0;        // partial maxima written per torque evaluation
    double* epart = nullptr;  // per-block energy sums (npart), wantEnergy
    bool wantEnergy = false;
    int fieldCap = 0;                // points in fT / fB
    double fieldTmaxHost = -INFINITY;  // ABI mode: latest time the field was evaluated at
    // ABI mode: an applied current density (Apply*("j", ...)), evaluated on the host
    // at each torque's stage time with its own forward-only search
    bool curOn = false;
    int curN = 0, curFn = 0;
    std::vector<double> curT;
    std::vector<float> curV;  // 3 per point, A/m^2, internal order
    double curFp[4] = {0, 0, 0, 0};
    float curFb[3] = {0, 0, 0};
    double curTmaxHost = -INFINITY;
    double* fT = nullptr;
    float* fB = nullptr;
    float* alphaMap = nullptr;       // device copy of mmParams.alphaMap (N floats)
    float* normMap = nullptr;        // device copy of mmParams.normMap (N floats)
    float* jMask = nullptr;          // device copy of mmParams.currentMask (3N floats)
    FieldTable field{};
    ConvPlan cp{};
    bool conv = false;
    double kernelImagRatio = 0;
    bool fsalHost = false;
    unsigned long long bytes = 0;
    Prof prof;
    // ABI mode
    gpuFFT3dPlan* plan = nullptr;
    int Pabi[3] = {};
    float* abuf[3] = {};
    float* akern[6] = {};
    float* am1 = nullptr;
    float* abackup = nullptr;
    float* redDev = nullptr;
    std::vector<float> redHost;
    int redBlocks = 0, redThreads = 0;
    // host mirror of the reference's RK state for ABI mode
    double htime = 0;
    float hdt = 0;
    long long hsteps = 0, htrials = 0, hrejects = 0;
    float herr = 0, htorque = 0;
    bool hfsal = false;
    std::vector<double> hlog_dt;
    std::vector<int> hlog_acc;
    std::vector<float> hlog_err;
};


namespace gk {
EngineState read_state(mmSim* s);
// the dt clamps of rk.go:356-371 need max|k0| before stage 1; FSAL tableaux take
// max|k0| from the previous trial's last stage
bool needs_after_k0(const mmSim* s);
CtrlParams ctrl_params(const mmSim* s);
TorqueArgs torque_args(const mmSim* s, float c, float* part);
void* dmalloc(mmSim* s, size_t bytes);
}  // namespace gk
