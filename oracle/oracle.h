/*
 * oracle.h -- CPU restatement of the MuMax 1.1 (barnex/1) LLG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (libgpukern.so) never links, loads or calls it.
 *
 * Precision: everything is `or_real`, double by default (liboracle.so).  The
 * same source compiled with -DOR_REAL_FLOAT gives liboracle_f32.so, an FP32
 * twin used to measure the FP32-vs-FP64 floor of the algorithm itself.
 *
 * Layout conventions follow the reference (internal X = thickness, Z fastest; planar components):
 *   - internal axes (X,Y,Z) = user (z,y,x); X is the film thickness;
 *   - vector fields are planar: [3][X][Y][Z], Z fastest;
 *   - kernels are 6 real arrays (xx,yy,zz,yz,xz,xy) over the padded grid,
 *     displacements wrapped (kernel.go:14-24).
 *
 * Pinning of each function is listed at the top of oracle.c.
 */
#ifndef MUMAX_ORACLE_H
#define MUMAX_ORACLE_H

#ifdef OR_REAL_FLOAT
typedef float or_real;
#else
typedef double or_real;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* kernel component indices (kernel.go:95-101) */
enum { OR_XX = 0, OR_YY = 1, OR_ZZ = 2, OR_YZ = 3, OR_XZ = 4, OR_XY = 5 };
/* reduction op codes (gpu_reduction.h:24-38) */
enum { OR_REDUCE_ADD = 1, OR_REDUCE_MAX = 2, OR_REDUCE_MAXABS = 3, OR_REDUCE_MIN = 4, OR_REDUCE_SUMABS = 5 };
/* anisotropy types (simanis.go:76-80) */
enum { OR_ANIS_NONE = 0, OR_ANIS_UNIAXIAL = 1 };
/* solvers (solver.go:16-53; rk.go tableaux) */
enum { OR_RK2 = 2, OR_RK12 = 12, OR_RK3 = 3, OR_RK23 = 23, OR_RK4 = 4, OR_RKCK = 45, OR_RKDP = 54 };
enum { OR_SEMIANAL1 = 101, OR_SEMIANAL2 = 102 }; /* semianal.go: fixed dt, the time is not advanced */

int or_real_size(void);

/* ---------------- units (material.go:36-110) ---------------- */
/* out[0]=length (m), out[1]=time (s), out[2]=field (T), out[3]=energy (J), out[4]=energy density (J/m^3) */
void or_units(double msat, double aexch, double out[5]);

/* ---------------- kernels ---------------- */
/* padded size: 2N where N>1 and not periodic (kernel.go:79-90) */
void or_pad_size(const int size[3], const int periodic[3], int padded[3]);
/* FaceKernel6 (demag.go:26-84,142-183): K is 6*P0*P1*P2, overwritten. */
void or_facekernel6(const int P[3], const or_real cell[3], int accuracy, const int periodic[3], or_real* K);
/* float32 transcription of faceIntegral for bit-level comparison (tensor.Vector is [3]float32). */
void or_facekernel6_f32(const int P[3], const float cell[3], int accuracy, const int periodic[3], float* K);
/* Exch6NgbrKernel (exchange.go:44-69), ADDED into K (init.go:256-267), including the
 * wrap(+-1,1)=0 overwrite of the centre when a padded size is 1 (exchange.go:58-65). */
void or_add_exch6_kernel(const int P[3], const or_real cell[3], or_real* K);

/* ---------------- FFT ---------------- */
/* in-place complex FFT of `howmany` contiguous lines of length n (interleaved re,im);
 * sign=-1 forward, +1 inverse (unnormalised).  Any n>=1 (mixed radix, generic odd primes). */
void or_fft_c2c(or_real* data, int n, int howmany, int sign);
/* in-place 3-D complex FFT of [n0][n1][n2] */
void or_fft3_c2c(or_real* data, const int n[3], int sign);

/* ---------------- convolution (conv.go:35-198, fft.go, gpu_fft0.cu) ---------------- */
typedef struct or_conv or_conv;
/* N = data size, P = padded size, K6 = 6*P real-space kernel (demag [+exchange]) */
or_conv* or_conv_new(const int N[3], const int P[3], const or_real* K6);
void or_conv_free(or_conv* c);
/* h = K * m (planar 3*N each) */
void or_conv_apply(or_conv* c, const or_real* m, or_real* h);
/* normalised real kernel spectrum, 6 * P0*P1*(P2/2+1) (conv.go:126-160) */
void or_conv_kernel_spectrum(const or_conv* c, or_real* out);
/* padded forward transform of one component, as gpuFFT3dPlan_forward: out is
 * [P0][P1][P2/2+1] complex (interleaved) */
void or_conv_forward1(or_conv* c, const or_real* in, or_real* out);
/* brute-force O(N^2) linear convolution with the wrapped kernel, for pinning */
void or_conv_direct(const int N[3], const int P[3], const or_real* K6, const or_real* m, or_real* h);

/* ---------------- pointwise ---------------- */
void or_local_fields(const or_real* m, or_real* h, int N, const or_real hext[3], int anisType,
                     or_real k, const or_real axis[3]);
void or_local_fields_phi(const or_real* m, or_real* h, or_real* phi, int N, const or_real hext[3],
                         int anisType, or_real k, const or_real axis[3]);
void or_deltaM(const or_real* m, or_real* h, or_real alpha, or_real dtgilb, int N);
void or_deltaM_map(const or_real* m, or_real* h, or_real alpha_mul, const or_real* amap, or_real dtgilb, int N);
void or_normalize(or_real* m, int N);
/* gpu_anal.cu:23-84: m advanced over dt in a constant field h in closed form (in place allowed) */
void or_anal_step(const or_real* min, or_real* mout, const or_real* h, double dt, double alpha, int N);
void or_normalize_map(or_real* m, const or_real* nmap, int N);
void or_scale_dot(or_real* r, const or_real* a, const or_real* b, or_real s, int N);
double or_reduce(int op, const or_real* x, int n);
/* 6-neighbour exchange stencil with free (zero-halo) or periodic boundaries (gpu_exch1.cu:21-301) */
void or_add_exch6_stencil(const or_real* m, or_real* h, const int size[3], const int periodic[3],
                          const or_real cell[3]);

/* ---------------- simulation (sim.go, init.go, rk.go, simrun.go, field.go, torque.go) ---------------- */
typedef struct {
    int size[3];            /* internal X,Y,Z cells */
    int periodic[3];
    double cellsize[3];     /* SI metres, internal order */
    double msat, aexch, alpha;
    int anisType;           /* OR_ANIS_* */
    double k1;              /* J/m^3 */
    double anisAxis[3];     /* internal order, normalised by caller */
    int exchType;           /* 0 or 6 */
    int exchInConv;         /* 1 = exchange folded into the convolution kernel (default) */
    int wantDemag;          /* 1 = demag on */
    int demagAccuracy;      /* 8 */
    int solver;             /* OR_RK* */
    double maxError, dt;    /* dt in seconds; 0 = DEFAULT_DT_INTERNAL (1e-4 internal) */
    double minDm, maxDm, minDt, maxDt;
    int nField;             /* 0: no field; 1: static; >=2: pointwise (simfield.go:101-153) */
    const double* fieldT;   /* nField times (s) */
    const double* fieldB;   /* nField*3 tesla, internal order */
    int wantEnergy;         /* compute energy at the FSAL stage (rk.go:323-337) */
    double temperature;     /* K; thermal field drawn once per step (temperature.go:29-36, rk.go:289-292) */
    unsigned long long seed; /* thermal stream seed */
} or_params;

/* The thermal field's standard normals of cell `cell` at step `step` (3 values):
 * SplitMix64's output function over (seed, step, cell), the top 24 bits as uniforms
 * ((r + 1/2) 2^-24 and r 2^-24), Box-Muller in double.  The engine draws the same
 * uniforms; the reference's cuRAND stream is not reproduced (its seed and
 * generator are cuRAND's, gpu_temperature.cu:31-39). */
void or_thermal_normals(unsigned long long seed, long long step, long cell, double out[3]);

typedef struct or_sim or_sim;
or_sim* or_sim_new(const or_params* p, const float* m0 /* 3*N, internal planar, any norm */);
void or_sim_free(or_sim* s);
/* one reference step (rk.go:279-445).  If dt_trial != NULL and n_trial > 0 the step replays a
 * schedule: trial t uses dt_trial[t] and is accepted iff accept[t]!=0 (the GPU's decisions). */
int or_sim_step(or_sim* s);
int or_sim_step_replay(or_sim* s, const double* dt_trial, const int* accept, int n_trial);
/* Run(duration): loop until time >= stop (simrun.go:19-57) */
int or_sim_run(or_sim* s, double seconds);
void or_sim_get_m(const or_sim* s, double* m);        /* 3*N */
void or_sim_set_m(or_sim* s, const double* m);
/* state[0]=time(internal) [1]=dt(internal) [2]=steps [3]=last error [4]=maxTorque
 * [5]=trials total [6]=rejects total [7]=energy (internal) [8]=unit time (s) */
void or_sim_state(const or_sim* s, double* st);
/* effective field and torque of the current m (calcHeff + Torque) */
void or_sim_heff(or_sim* s, const double* m, double* h);
void or_sim_torque(or_sim* s, const double* m, double* k);
/* replace the applied-field table (n points, t[n] s, b[3n] T internal order) */
void or_sim_set_field(or_sim* s, int n, const double* t, const double* b);
/* 1: round the RK coefficients to float32 as the reference does (default 0: exact) */
void or_set_tableau_f32(int on);
/* per-trial log: dt used and accepted flag of every trial so far (first cap copied) */
long or_sim_trial_log(const or_sim* s, double* dt, int* acc, long cap);
/* energy-controlled stepping on/off (relax.go:16, rk.go:424-426) */
void or_sim_set_want_energy(or_sim* s, int on);
/* applied field as a function of time (1 pulse, 2 RF, 3 RF ramp, 4 sawtooth,
 * 5 rotating; params as mm_sim_set_field_function), replacing the table */
void or_sim_set_field_function(or_sim* s, int kind, const double* b, const double* params);
/* the applied field (tesla, internal order) at t seconds */
void or_sim_field_at(or_sim* s, double t_si, double* out);
/* per-cell damping multiplier, N floats (NULL: none); copied */
void or_sim_set_alpha_map(or_sim* s, const float* amap);
/* spin-transfer torque (CurrentDensity, SpinPolarization, Xi, CurrentMask;
 * simspintorque.go): current density j (A/m^2, internal order), 3N mask floats
 * (NULL: 1); copied.  A nonzero j selects the STT torque (torque.go:40-49) */
void or_sim_set_spintorque(or_sim* s, const double j[3], double spinPol, double xi, const float* jmask);
/* an applied current density (Apply*("j", ...), simfield.go:61-70): n points
 * (t s, j A/m^2 internal order; n = 1 static, n = 0 none) or, kind > 0, a function
 * of time with amplitude j[3] (kinds and params as or_sim_set_field_function);
 * evaluated at each stage's time, it replaces the current density (torque.go:42-44) */
void or_sim_set_applied_current(or_sim* s, int n, const double* t, const double* j, int kind, const double* params);
void or_spintorque_deltaM(const or_real* m, or_real* h, double alpha, double beta, double eps, const double u[3],
                          const or_real* jmap, double dtg, const int n[3]);
/* per-cell norm of m after each normalisation, N floats (NULL: 1); copied */
void or_sim_set_norm_map(or_sim* s, const float* nmap);
/* total energy (internal units * cell volume), energy.go:38-52 */
double or_sim_energy(or_sim* s);

#ifdef __cplusplus
}
#endif
#endif
