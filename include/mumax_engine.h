/*
 * mumax_engine.h -- additions to libgpukern.so: the device-resident LLG engine.
 *
 * None of these entry points exists in the reference.  They are the host
 * side of the reference's solver (the Go packages sim/rk.go, sim/conv.go,
 * sim/field.go, sim/torque.go, sim/simrun.go) restated in C++ on top of the
 * kernels, so a caller gets a whole time step per call instead of ~150
 * synchronous kernel calls.  Two step implementations are selectable:
 *
 *   MM_MODE_FUSED  MI355X-native: zero-skipping transposed FFT convolution
 *                  with the k-space multiply fused into the outermost axis,
 *                  Runge-Kutta stage build fused into the forward Z pass,
 *                  local fields + torque + step-control reductions fused into
 *                  one pass, dt control on the device (no host round trip
 *                  per step).
 *   MM_MODE_ABI    the reference's call sequence (rk.go:279-445 ->
 *                  conv.go:57-100 -> field.go:74-80 -> torque.go:33-53) issued
 *                  through the drop-in gpukern.h functions, one synchronous
 *                  call at a time, as gpu.go would.
 *
 * Arrays crossing this boundary are host float/double arrays in the
 * reference's internal layout: planar [3][X][Y][Z], Z fastest, X = thickness.
 * Errors print to stderr and abort(), like the rest of libgpukern.
 */
#ifndef MUMAX_ENGINE_H
#define MUMAX_ENGINE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { MM_MODE_FUSED = 0, MM_MODE_ABI = 1 };
enum { MM_ANIS_NONE = 0, MM_ANIS_UNIAXIAL = 1 };
/* demag kernel: the reference's FaceKernel6 (demag.go:26-84) or the Newell
 * tensor (cell-averaged field of a uniformly magnetised cell, FP64) */
enum { MM_KERNEL_FACE6 = 0, MM_KERNEL_NEWELL = 1 };
/* applied-field functions of time (simfield.go:155-260) */
enum { MM_FIELD_PULSE = 1, MM_FIELD_RF = 2, MM_FIELD_RF_RAMP = 3, MM_FIELD_SAWTOOTH = 4, MM_FIELD_ROTATING = 5 };
/* solver codes: rk.go tableaux (solver.go:16-53) */
enum { MM_RK2 = 2, MM_RK12 = 12, MM_RK3 = 3, MM_RK23 = 23, MM_RK4 = 4, MM_RKCK = 45, MM_RKDP = 54 };
/* The semi-analytic solvers (SemiAnal1 / SemiAnal2, semianal.go; solver.go:46-49):
 * MM_MODE_ABI, fixed dt, mm_sim_steps only -- their steps never advance the time,
 * so a Run would not end (the reference's do not either). */
enum { MM_SEMIANAL1 = 101, MM_SEMIANAL2 = 102 };

/* Simulation parameters: the reference's input commands (sim.go:50-170),
 * already in internal axis order. */
typedef struct {
    int size[3];            /* cells, internal X,Y,Z (gridsize z y x) */
    int periodic[3];        /* periodic images per axis (Periodic, simgeom.go:84-88); 0: free.
                               A periodic axis keeps its size in the reference's kernel grid
                               (padSize, kernel.go:79-90) and the kernel adds the images
                               (demag.go:48-79). Fused mode: even sizes (or 1) only */
    double cellsize[3];     /* metres, internal order (cellsize z y x) */
    double msat, aexch, alpha;
    int anisType;           /* MM_ANIS_* */
    double k1;              /* J/m^3 */
    double anisAxis[3];     /* internal order, unit vector */
    int exchType;           /* 0 or 6 */
    int exchInConv;         /* 1: exchange folded into the demag kernel (sim.go:203) */
    int wantDemag;
    int demagAccuracy;      /* face-charge points per face edge (8 in the reference) */
    int solver;             /* MM_RK* */
    double maxError, dt;    /* dt in seconds, 0 = default 1e-4 internal units */
    double minDm, maxDm, minDt, maxDt;
    int nField;             /* 0 none, 1 static, >=2 pointwise linear (simfield.go:87-153) */
    const double* fieldT;   /* nField times (s) */
    const double* fieldB;   /* nField*3 tesla, internal order */
    int wantEnergy;         /* energy at the FSAL stage (rk.go:323-337) */
    int demagKernel;        /* MM_KERNEL_* (0: FaceKernel6, as the reference) */
    const float* alphaMap;  /* N host floats: per-cell damping multiplier (SetAlpha /
                               SetAlphaRange, simgeom.go:226-264); null: uniform alpha.
                               The torque's 1/(1+alpha^2) keeps the global alpha
                               (torque.go:49-51). Copied at mm_sim_new. */
    double temperature;     /* K (Temperature, temperature.go:29-36, sim.go:78); 0: none.
                               Fused mode: a thermal field of standard deviation
                               sqrt(2 alpha kB T / (gamma0 mu0 Ms V dt)) / Ms per component,
                               drawn once per step from dt at the step's start and kept
                               through its trials (rk.go:289-292) */
    unsigned long long seed;  /* thermal noise stream: counter-based (SplitMix64 of seed,
                               step and cell, Box-Muller), so it does not depend on launch
                               shapes or batching */
    const float* normMap;   /* N host floats: per-cell norm of m, the "normalized
                               saturation magnetization" of SetMsat / SetMsatRange /
                               SetMsatEllips (simgeom.go:266-340); every normalisation
                               (the run's start, each step's end) scales a cell to its
                               value, 0 leaving it empty (Sim.Normalize, sim.go:230-239;
                               gpu_normalize.cu:43-71). Null: 1 everywhere. <m> still
                               divides by the cell count (avgNorm, init.go:135).
                               Copied at mm_sim_new. */
    double current[3];      /* current density, A/m^2, internal order (CurrentDensity,
                               simspintorque.go:30-36). Nonzero: every torque is the
                               spin-transfer (Zhang-Li) form (torque.go:40-91,
                               gpu_spintorque.cu). MM_MODE_ABI only */
    double spinPol, xi;     /* SpinPolarization, Xi (simspintorque.go:19-28) */
    const float* currentMask; /* 3N host floats (CurrentMask): each component of the
                               current scaled per cell; required with a current, as in
                               the reference (torque.go:68-71). Copied at mm_sim_new. */
} mmParams;

typedef struct mmSim mmSim;

/* m0: host 3*N floats, any norm (normalised by the first run/steps call) */
mmSim* mm_sim_new(const mmParams* p, const float* m0, int mode);
void mm_sim_free(mmSim* s);

/* Run(duration): steps until time >= stop (simrun.go:19-57); returns accepted steps */
long mm_sim_run(mmSim* s, double seconds);
/* The step loop of Run() (simrun.go:29-49) up to an absolute stop time in
 * internal units (state[0] and state[8] give the time and the unit), without
 * Run()'s normalisation at the start: a Run() interrupted at output times is
 * mm_sim_run(s, 0) (the normalisation) followed by mm_sim_run_until for each
 * stretch between outputs.  Returns the number of accepted steps. */
long mm_sim_run_until(mmSim* s, double stop);
/* Steps(n) (simrun.go:88-92) */
void mm_sim_steps(mmSim* s, int n);
/* Enqueue n steps without waiting (fused mode; for timing); mm_sim_sync waits. */
void mm_sim_steps_async(mmSim* s, int n);
void mm_sim_sync(mmSim* s);

void mm_sim_get_m(mmSim* s, float* m);          /* 3N host floats */
void mm_sim_set_m(mmSim* s, const float* m);
float* mm_sim_device_m(mmSim* s);               /* device pointer to planar m (3N) */

/* st[0]=time (internal) [1]=dt (internal) [2]=steps [3]=last error [4]=max torque
 * [5]=trials [6]=rejects [7]=energy (internal) [8]=unit time (s) [9]=unit field (T)
 * [10..12] padded size used by the convolution */
void mm_sim_state(mmSim* s, double* st);

/* per-trial schedule (dt internal, accepted flag) of the last `max` trials since
 * creation, oldest first; returns the number written */
int mm_sim_trial_log(mmSim* s, double* dt, int* accept, int max);
/* the error estimate of the same trials (0 for fixed-step tableaux): the max-abs
 * of m_low - m_high from which each trial's next dt was chosen (rk.go:403-420) */
int mm_sim_trial_errors(mmSim* s, float* err, int max);

/* one evaluation of the effective field / torque of m (host 3N floats) */
void mm_sim_heff(mmSim* s, const float* m, float* h);
void mm_sim_torque(mmSim* s, const float* m, float* k);

/* energy of the current m, internal units x cell volume (energy.go:38-52), FP64 reductions */
double mm_sim_energy(mmSim* s);
/* <m> per component, FP64 reduction (simoutput.go:234-238), internal order */
void mm_sim_average_m(mmSim* s, double* avg);

/* Demag(+exchange) kernel spectrum used by the engine, as the reference stores it:
 * nK x [Px][Py][Pz/2+1] floats, Re(FFT(K))/(Px Py Pz), nK = 6 (3-D) or 4 (2-D,
 * xx yy zz yz).  out may be NULL to query; returns the number of floats. */
long mm_sim_kernel_spectrum(mmSim* s, float* out);

/* Real-space demag kernel in FP64 over a padded grid P (internal X,Y,Z), cell
 * in metres: 6 x [P0][P1][P2] doubles (xx yy zz yz xz xy) at wrapped
 * displacements, zero where FaceKernel6's loop leaves them (demag.go:35-53);
 * kernel = MM_KERNEL_*, accuracy = face-charge points per edge (FaceKernel6).
 * The values are in the internal unit system (the cell is scaled by the
 * exchange length of msat/aexch, which a dimensionless kernel does not see). */
void mm_demag_kernel(const int P[3], const double cellsize[3], int kernel, int accuracy, double* out);

/* bytes of device memory held by the simulation (a kernel spectrum shared with
 * other simulations of the same padded size, cell and kernel is counted by each) */
unsigned long long mm_sim_device_bytes(mmSim* s);
/* distinct kernel spectra held on the current device by live simulations */
int mm_kernel_cache_size(void);
/* Freed engine memory (simulations, plans, spectra, batches) is kept per
 * device and reused by later simulations and batches, so that freeing does not
 * wait for the device's other streams (hipFree does).  Bytes the current
 * device's cache holds, and a release of all of them to the device (this
 * waits for the device, as hipFree does) together with the idle batch streams,
 * so that the next batches get new streams (and the hardware queues new
 * streams get). */
unsigned long long mm_memory_cached(void);
unsigned long long mm_memory_trim(void);

/* Fixed-step fused mode only.  Enqueue n steps bracketed by HIP events on the
 * engine's stream; returns the elapsed ms.  The _profile variant also brackets
 * every kernel class (0 Z fwd, 1 Y fwd, 2 X.K.X or Y.K.Y, 3 Y inv, 4 Z inv,
 * 5 torque, 6 RK final, 7 control) and returns per-class total ms and launch
 * counts in ms[8], counts[8]. */
double mm_sim_time_steps(mmSim* s, int n);
double mm_sim_profile_steps(mmSim* s, int n, double* ms, int* counts);
/* Fused mode.  n back-to-back launches of one convolution pass class (0..4 as
 * above; the Z inverse with the stage-1 torque epilogue) on the engine's
 * stream between two HIP events; returns the elapsed ms.  The passes work on
 * the engine's scratch spectra and its first stage buffer, so the state m is
 * unchanged; the next step recomputes everything they touch. */
double mm_sim_time_pass(mmSim* s, int cls, int n);

/* Autotuner of the launch shapes.  The reference fixes them by one rule for
 * every kernel (make1dconf, gpu_conf.cu:68-97, with the size filter of
 * cufftsize.go:14-52); here each pass of a fused-mode simulation has its own
 * tunables, timed on the device: params[6] = {zr: Z forward rows per block
 * (variant 0 default, 1 half, 2 double), er: fused Z inverse rows per block
 * (0 default, 1 half), yb: blocks per CU of the 3-D Y passes, xb: blocks per CU
 * of the X.K.X pass, y2b: blocks per CU of the 2-D Y.K.Y pass, pp: for 3-D
 * plans with padded Y 512 and padded X 16 or 32, 1 = Y forward, X.K.X and Y
 * inverse as one register-resident plane pass, 0 = three passes, xv: X.K.X
 * variant for padded X 16-64, 0 default, 1 without the next item's prefetch,
 * 2 half the elements per lane over twice the lanes, 3 both}.  Each tunable's
 * candidates run `reps` back-to-back launches (least of 3 runs); the fastest is
 * kept in s and stored in the process's tuning table under s's data size.
 * ms_default / ms_tuned [0..6]: ms per launch of the affected pass(es) at the
 * default and the chosen value (-1 where the plan has no such pass); [7]: the
 * three separate passes and the plane pass (-1 where it does not fit).  m is
 * unchanged.
 * Row counts, blocks per CU and the prefetch change no result bit, except that
 * the fused Z inverse's row count sets the summation order of the FP64
 * per-block energy sums; the plane pass and xv 2, 3 compute the same
 * transforms in another order (FP32 rounding apart), and the batched engine
 * keeps the default shapes. */
int mm_sim_tune(mmSim* s, int reps, int params[7], double ms_default[8], double ms_tuned[8]);
void mm_sim_get_tune(mmSim* s, int params[7]);
/* Tuning table: new simulations of a listed data size take its parameters.
 * Text, one line "Nx Ny Nz zr er yb xb y2b pp xv" per size (internal order;
 * lines of 8 or 9 columns read pp and xv as 0).  save
 * and load return the number of entries, or -1 if the file cannot be used. */
int mm_tune_save(const char* path);
int mm_tune_load(const char* path);
void mm_tune_clear(void);

/* device bytes mm_sim_new(p, ..., MM_MODE_FUSED) allocates (admission by HBM
 * footprint); the kernel set-up needs transient FP64 buffers beyond this */
unsigned long long mm_sim_footprint(const mmParams* p);
/* Run()'s set-up without stepping (fused mode): m normalised, stop time =
 * time + seconds (simrun.go:19-24); the trials are then driven by a batch
 * (mm_batch_trials), which is how a running batch takes in a new request. */
void mm_sim_begin_run(mmSim* s, double seconds);
/* ApplyStatic / ApplyPointwise on a live simulation (simfield.go:87-153): n
 * points (t[n] seconds, b[3n] tesla, internal order; n = 1: static, n = 0:
 * none).  Takes effect at the next trial; a batch built before keeps the
 * copy of the table it was built with.  The table is replaced in place (the
 * call waits for the simulation's queued trials), in a buffer that grows by
 * doubling. */
void mm_sim_set_field(mmSim* s, int n, const double* t, const double* b);
/* Temperature(T) on a live simulation (temperature.go:47-56): the thermal field's
 * standard deviation from the next trial on (fused mode); a batch built before
 * keeps the temperature it was built with. */
void mm_sim_set_temperature(mmSim* s, double temperature);
/* CurrentDensity / SpinPolarization / Xi on a live simulation (ABI mode; the
 * mask stays the one given at creation): j[3] in A/m^2, internal order. */
void mm_sim_set_spintorque(mmSim* s, const double j[3], double spinPol, double xi);
/* ApplyStatic / ApplyPointwise / ApplyRf ... ("j", ...) on an ABI-mode
 * simulation (simfield.go:61-70): the current density as n points (t[n]
 * seconds, j[3n] A/m^2 internal order; n = 1 static; n = 0 and kind = 0: off,
 * back to CurrentDensity's value) or, kind = MM_FIELD_*, a function of time of
 * amplitude j[3] with params[4] as mm_sim_set_field_function.  Evaluated at each
 * torque's stage time with its own forward-only search, it replaces the current
 * density there; any applied current selects the spin-transfer torque
 * (torque.go:40-45).  Needs the current mask. */
void mm_sim_set_current(mmSim* s, int n, const double* t, const double* j, int kind, const double* params);
/* Replace the applied field by a function of time (MM_FIELD_*), amplitude b[3]
 * in tesla (internal X, Y, Z), params[4]: PULSE (duration s, rise time s),
 * RF (frequency Hz), RF_RAMP (frequency, ramp time), SAWTOOTH (frequency),
 * ROTATING (frequency, phase X, Y, Z rad) -- ApplyPulse, ApplyRf, ApplyRfRamp,
 * ApplySawtooth, ApplyRotating (simfield.go:155-260). */
void mm_sim_set_field_function(mmSim* s, int kind, const double* b, const double* params);
/* Relax() (relax.go:14-25): wantEnergy on (trials that raise the energy are
 * rejected, rk.go:424-426), steps until a step raises the energy, or until
 * max_steps steps when max_steps > 0 (the reference has no bound); returns
 * the steps taken.  Fused mode.  wantEnergy stays on afterwards, as in the
 * reference. */
long mm_sim_relax(mmSim* s, long max_steps);

/* ---- batched engine ----------------------------------------------------
 * n fused-mode simulations sharing the solver (any sizes, materials, fields)
 * advanced together: every RK stage is one launch per pass and padded length
 * for the whole batch, step control stays per simulation on the device, and a
 * batched trial is captured once into a HIP graph and replayed.  The batch
 * references the simulations (they stay owned by the caller and usable alone
 * after mm_batch_free); a simulation's arithmetic is the same alone and in a
 * batch. */
typedef struct mmBatch mmBatch;
mmBatch* mm_batch_new(mmSim* const* sims, int n);
void mm_batch_free(mmBatch* b);
/* 1 (default): replay a captured graph per trial; 0: enqueue the launches */
void mm_batch_set_graph(mmBatch* b, int on);
/* Batches created from now on launch the length classes of each pass (Z,
 * thin-plane, 2-D, 3-D Y and X.K.X) that have fewer than `blocks` blocks
 * together, in one launch
 * of a kernel that runs each simulation's per-length body (default 1024; 0:
 * one launch per length class).  Results do not depend on it.  Returns the
 * previous value. */
int mm_batch_set_merge_blocks(int blocks);
int mm_batch_size(mmBatch* b);
/* Run() set-up: simulation j stops at time_j + seconds[j], m normalised first
 * (simrun.go:19-24); or, seconds null, Steps(): steps[j] more steps, no
 * normalisation (simrun.go:85-92).  steps null: no step target. */
void mm_batch_start(mmBatch* b, const double* seconds, const long long* steps);
/* enqueue n batched trials and wait; returns the simulations not yet done */
int mm_batch_trials(mmBatch* b, int n);
/* The same without waiting: enqueue n batched trials and the copy of the done
 * flags on the batch's own stream.  mm_batch_ready is 1 once that work has
 * finished; mm_batch_flags then fills done[n_sims] (may be null) with each
 * simulation's status -- 0 running, 1 done (stop time or step target reached),
 * 2 failed (stopped because dt became NaN, rk.go:448-452, or its torque or error
 * estimate was NaN or infinite) -- and returns the simulations still running
 * (it waits if the work is still running).
 * Several batches on their own streams can run concurrently this way. */
void mm_batch_enqueue(mmBatch* b, int n);
int mm_batch_ready(mmBatch* b);
int mm_batch_flags(mmBatch* b, int* done);
/* After mm_batch_enqueue: each simulation's remaining trials at the dt it will
 * try next, (stop - time) / dt (0 once done; remaining steps under a step
 * target), as of the end of that work, into rem[n_sims]; waits like
 * mm_batch_flags.  For schedulers that give long-running members their own
 * batch.  Returns n_sims. */
int mm_batch_remaining(mmBatch* b, float* rem);
/* done[j]: status of simulation j as in mm_batch_flags (0 running, 1 done, 2 failed) */
void mm_batch_done(mmBatch* b, int* done);
/* mm_batch_start(seconds) then trials until all are done (polling every
 * `poll` trials); returns the batched trials enqueued */
long long mm_batch_run(mmBatch* b, const double* seconds, int poll);

#ifdef __cplusplus
}
#endif
#endif
