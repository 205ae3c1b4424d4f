/*
 * oracle.c -- CPU restatement of the MuMax 1.1 (barnex/1) LLG hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libgpukern.so.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference root).  How each part is pinned (tests/test_oracle_*.py):
 *   or_units             closed form: lex(Py)=5.686 nm, t_u=5.654 ps (Ms 8e5 A/m, A 1.3e-11 J/m)
 *   or_facekernel6       self term: trace -> -1, cube -> -1/3 each; thin film Nxx -> -1;
 *                        far field -> point dipole; symmetry K(-r)=+-K(r); float32 twin
 *                        bit-compared with or_facekernel6_f32 (the literal transcription)
 *   or_add_exch6_kernel  spin-wave eigenvalue -(2/h^2)sum(1-cos kh); conv exchange == stencil (3-D)
 *   or_fft_*             brute-force DFT on small n; round trip; Parseval; impulse -> constant;
 *                        numpy.fft (pocketfft) on random data
 *   or_conv_*            brute-force direct convolution; unit kernel -> identity; zero kernel -> 0;
 *                        padded transform == manually padded transform (fft_test.go:79-165)
 *   or_local_fields      closed forms (Zeeman constant, uniaxial k(m.u)u)
 *   or_deltaM            m.k = 0; |k| = |m x h| sqrt(1+a^2)/(1+a^2); Larmor precession
 *   or_reduce            redux_test.go:13-102 known answers
 *   or_sim (RK)          tableau order on dm/dt=linear; analytic macrospin precession+damping;
 *                        forced reject path with a tiny maxError; SP4 <m>(t) vs solution3.txt (loose);
 *                        float32-rounded coefficients (the reference's literals) vs exact rationals
 *   energy (calcHeffEnergy) closed forms: Zeeman -m.B, anisotropy and demag -m.H/2, two-cell
 *                        exchange vol (4 - cos t)/h^2 with the stencil's zero boundary
 *   or_sim_set_field     identical to the same field given at creation
 *   Relax (Python loop)  energy never rises over accepted steps, the stopping step raises it
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define X 0
#define Y 1
#define Z 2

int or_real_size(void) { return (int)sizeof(or_real); }

static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu x %zu)\n", n, sz); abort(); }
    return p;
}

/* ====================================================================== */
/* units: material.go:36-110, init.go:80-113                               */
/* ====================================================================== */
static const double MU0 = 4.0e-7 * 3.14159265358979323846;
static const double GAMMA0 = 2.211e5;

void or_units(double msat, double aexch, double out[5]) {
    double L = sqrt(2.0 * aexch / (MU0 * msat * msat)); /* UnitLength, material.go:253-256 */
    double T = 1.0 / (GAMMA0 * msat);                   /* UnitTime */
    double B = MU0 * msat;                              /* UnitField */
    double E = aexch * L;                               /* UnitEnergy */
    double Ed = E / (L * L * L);                        /* UnitEnergyDensity */
    out[0] = L; out[1] = T; out[2] = B; out[3] = E; out[4] = Ed;
}

/* ====================================================================== */
/* kernels: kernel.go:62-98, demag.go:26-84,142-183, exchange.go:44-69     */
/* ====================================================================== */
static int wrapi(int number, int max) { /* kernel.go:62-72 */
    while (number < 0) number += max;
    while (number >= max) number -= max;
    return number;
}

void or_pad_size(const int size[3], const int periodic[3], int padded[3]) { /* kernel.go:79-90 */
    for (int i = 0; i < 3; i++)
        padded[i] = (size[i] > 1 && periodic[i] == 0) ? 2 * size[i] : size[i];
}

static const int KERN_IDX[3][3] = {{OR_XX, OR_XY, OR_XZ}, {OR_XY, OR_YY, OR_YZ}, {OR_XZ, OR_YZ, OR_ZZ}};

/* faceIntegral (demag.go:142-183) in or_real: field at R of a cell uniformly
 * magnetised along s, from n x n point charges on each of the two faces normal to s. */
static void face_integral(or_real B[3], const or_real R[3], const or_real cell[3], int s, int n) {
    int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
    or_real charge = cell[v] * cell[w];
    or_real pu1 = cell[u] / 2, pu2 = -pu1;
    const or_real fourpi = (or_real)(4.0 * 3.14159265358979323846);
    B[0] = B[1] = B[2] = 0;
    for (int i = 0; i < n; i++) {
        or_real pv = -(cell[v] / 2) + cell[v] / (or_real)(2 * n) + (or_real)i * (cell[v] / (or_real)n);
        for (int j = 0; j < n; j++) {
            or_real pw = -(cell[w] / 2) + cell[w] / (or_real)(2 * n) + (or_real)j * (cell[w] / (or_real)n);
            for (int pole = 0; pole < 2; pole++) {
                or_real pos[3];
                pos[u] = pole == 0 ? pu1 : pu2;
                pos[v] = pv;
                pos[w] = pw;
                or_real d[3] = {R[0] - pos[0], R[1] - pos[1], R[2] - pos[2]};
                or_real r = (or_real)sqrt((double)(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
                or_real q = (pole == 0 ? charge : -charge) / (fourpi * r * r);
                /* normalised direction times q */
                B[0] += d[0] / r * q;
                B[1] += d[1] / r * q;
                B[2] += d[2] / r * q;
            }
        }
    }
    or_real inv = (or_real)1 / (or_real)(n * n);
    B[0] *= inv; B[1] *= inv; B[2] *= inv;
}

void or_facekernel6(const int P[3], const or_real cell[3], int accuracy, const int periodic[3], or_real* K) {
    size_t NP = (size_t)P[0] * P[1] * P[2];
    memset(K, 0, 6 * NP * sizeof(or_real));
    int x1 = -(P[X] - 1) / 2, x2 = P[X] / 2 - 1;
    if (P[X] == 1 && periodic[X] == 0) x2 = 0; /* demag.go:37-40 */
    int y1 = -(P[Y] - 1) / 2, y2 = P[Y] / 2 - 1;
    int z1 = -(P[Z] - 1) / 2, z2 = P[Z] / 2 - 1;
    x1 *= periodic[X] + 1; x2 *= periodic[X] + 1;
    y1 *= periodic[Y] + 1; y2 *= periodic[Y] + 1;
    z1 *= periodic[Z] + 1; z2 *= periodic[Z] + 1;
    for (int s = 0; s < 3; s++) {
        /* periodic images add into the same wrapped cell: x stays serial and each
         * thread owns one wrapped y, walking its images in ascending order, so every
         * entry is summed in the serial loop's order (x, then y, then z ascending) */
        for (int x = x1; x <= x2; x++) {
            int xw = wrapi(x, P[X]);
#pragma omp parallel for schedule(static)
            for (int yw = 0; yw < P[Y]; yw++) {
                if (y1 > y2) continue;
                for (int y = y1 + wrapi(yw - y1, P[Y]); y <= y2; y += P[Y])
                for (int z = z1; z <= z2; z++) {
                    int zw = wrapi(z, P[Z]);
                    or_real R[3] = {(or_real)x * cell[X], (or_real)y * cell[Y], (or_real)z * cell[Z]};
                    or_real B[3];
                    face_integral(B, R, cell, s, accuracy);
                    size_t idx = ((size_t)xw * P[Y] + yw) * P[Z] + zw;
                    for (int d = s; d < 3; d++) K[KERN_IDX[s][d] * NP + idx] += B[d];
                }
            }
        }
    }
}

/* literal float32 transcription of faceIntegral for bit-level checks:
 * tensor.Vector is [3]float32, Norm() is float32(math.Sqrt(float64(...))),
 * Normalize() multiplies by 1/norm, Scale multiplies (demag.go:142-183, vector.go:62-71). */
static void face_integral_f32(float B[3], const float R[3], const float cell[3], int s, int n) {
    int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
    float surface = cell[v] * cell[w];
    float charge = surface;
    float pu1 = cell[u] / 2.f, pu2 = -pu1;
    const float fourpi = (float)(4.0 * 3.14159265358979323846);
    B[0] = B[1] = B[2] = 0.f;
    for (int i = 0; i < n; i++) {
        volatile float pv = -(cell[v] / 2.f) + cell[v] / (float)(2 * n) + (float)i * (cell[v] / (float)n);
        for (int j = 0; j < n; j++) {
            volatile float pw = -(cell[w] / 2.f) + cell[w] / (float)(2 * n) + (float)j * (cell[w] / (float)n);
            for (int pole = 0; pole < 2; pole++) {
                float pos[3];
                pos[u] = pole == 0 ? pu1 : pu2;
                pos[v] = pv;
                pos[w] = pw;
                volatile float R2[3] = {R[0] - pos[0], R[1] - pos[1], R[2] - pos[2]};
                volatile float ss = R2[0] * R2[0];
                ss = ss + R2[1] * R2[1];
                ss = ss + R2[2] * R2[2];
                float r = (float)sqrt((double)ss);
                volatile float inv = 1.f / r;
                for (int c = 0; c < 3; c++) R2[c] = R2[c] * inv;
                volatile float den = fourpi * r;
                den = den * r;
                volatile float sc = (pole == 0 ? charge : -charge) / den;
                for (int c = 0; c < 3; c++) R2[c] = R2[c] * sc;
                for (int c = 0; c < 3; c++) B[c] = B[c] + R2[c];
            }
        }
    }
    float inv = 1.f / (float)(n * n);
    for (int c = 0; c < 3; c++) B[c] = B[c] * inv;
}

void or_facekernel6_f32(const int P[3], const float cell[3], int accuracy, const int periodic[3], float* K) {
    size_t NP = (size_t)P[0] * P[1] * P[2];
    memset(K, 0, 6 * NP * sizeof(float));
    int x1 = -(P[X] - 1) / 2, x2 = P[X] / 2 - 1;
    if (P[X] == 1 && periodic[X] == 0) x2 = 0;
    int y1 = -(P[Y] - 1) / 2, y2 = P[Y] / 2 - 1;
    int z1 = -(P[Z] - 1) / 2, z2 = P[Z] / 2 - 1;
    x1 *= periodic[X] + 1; x2 *= periodic[X] + 1;
    y1 *= periodic[Y] + 1; y2 *= periodic[Y] + 1;
    z1 *= periodic[Z] + 1; z2 *= periodic[Z] + 1;
    for (int s = 0; s < 3; s++)
        for (int x = x1; x <= x2; x++) {
            int xw = wrapi(x, P[X]);
#pragma omp parallel for schedule(static)
            for (int yw = 0; yw < P[Y]; yw++) {
                if (y1 > y2) continue;
                for (int y = y1 + wrapi(yw - y1, P[Y]); y <= y2; y += P[Y])
                for (int z = z1; z <= z2; z++) {
                    int zw = wrapi(z, P[Z]);
                    float R[3] = {(float)x * cell[X], (float)y * cell[Y], (float)z * cell[Z]};
                    float B[3];
                    face_integral_f32(B, R, cell, s, accuracy);
                    size_t idx = ((size_t)xw * P[Y] + yw) * P[Z] + zw;
                    for (int d = s; d < 3; d++) K[KERN_IDX[s][d] * NP + idx] += B[d];
                }
            }
        }
}

void or_add_exch6_kernel(const int P[3], const or_real cell[3], or_real* K) {
    /* exchange.go:44-69 builds a fresh 6NGBR kernel and init.go:256-267 adds it
     * element-wise.  Writes happen in the reference's order, so a padded size of 1
     * maps wrap(+-1,1)=0 onto the centre and the last write wins (exchange.go:58-65). */
    size_t NP = (size_t)P[0] * P[1] * P[2];
    or_real* E = (or_real*)xcalloc(NP, sizeof(or_real));
    or_real hx = cell[X] * cell[X], hy = cell[Y] * cell[Y], hz = cell[Z] * cell[Z];
#define EI(a, b, c) (((size_t)wrapi(a, P[X]) * P[Y] + wrapi(b, P[Y])) * P[Z] + wrapi(c, P[Z]))
    E[EI(0, 0, 0)] = -2 / hx - 2 / hy - 2 / hz;
    E[EI(1, 0, 0)] = 1 / hx;
    E[EI(-1, 0, 0)] = 1 / hx;
    E[EI(0, 1, 0)] = 1 / hy;
    E[EI(0, -1, 0)] = 1 / hy;
    E[EI(0, 0, 1)] = 1 / hz;
    E[EI(0, 0, -1)] = 1 / hz;
#undef EI
    for (int s = 0; s < 3; s++) {
        or_real* k = K + KERN_IDX[s][s] * NP;
        for (size_t i = 0; i < NP; i++) k[i] += E[i];
    }
    free(E);
}

/* ====================================================================== */
/* FFT: mixed-radix recursive decimation in time, double twiddles          */
/* ====================================================================== */
typedef struct { or_real re, im; } cpx;

typedef struct {
    int n, nf;
    int f[64];
    cpx* tw; /* tw[k] = exp(-2 pi i k / n) */
} fftplan;

#define MAXPLANS 64
static fftplan g_plans[MAXPLANS];
static int g_nplans = 0;

static const fftplan* get_plan(int n) {
    const fftplan* found = NULL;
#pragma omp critical(or_fft_plan)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i].n == n) { found = &g_plans[i]; break; }
        if (!found) {
            if (g_nplans >= MAXPLANS) { fprintf(stderr, "oracle: too many FFT plans\n"); abort(); }
            fftplan* p = &g_plans[g_nplans];
            p->n = n;
            p->nf = 0;
            int m = n;
            while (m % 4 == 0 && m > 4) { p->f[p->nf++] = 4; m /= 4; }
            while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
            for (int q = 3; m > 1; q += 2)
                while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
            p->tw = (cpx*)xcalloc((size_t)n, sizeof(cpx));
            for (int k = 0; k < n; k++) {
                /* exact reduction of the angle: k/n in [0,1) */
                double a = -2.0 * 3.14159265358979323846 * (double)k / (double)n;
                p->tw[k].re = (or_real)cos(a);
                p->tw[k].im = (or_real)sin(a);
            }
            g_nplans++;
            found = p;
        }
    }
    return found;
}

static inline cpx cmul(cpx a, cpx b) {
    cpx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

/* out[0..n) = DFT of in[0], in[is], ...; N = plan length, tws = N/n */
static void fft_rec(const cpx* in, cpx* out, int n, size_t is, const int* f, const cpx* tw, int N, int tws,
                    int sign) {
    if (n == 1) { out[0] = in[0]; return; }
    int p = f[0], m = n / p;
    for (int q = 0; q < p; q++) fft_rec(in + q * is, out + (size_t)q * m, m, is * p, f + 1, tw, N, tws * p, sign);
    if (p == 2) {
        for (int k = 0; k < m; k++) {
            cpx w = tw[(size_t)k * tws];
            if (sign > 0) w.im = -w.im;
            cpx a = out[k], b = cmul(out[k + m], w);
            out[k].re = a.re + b.re; out[k].im = a.im + b.im;
            out[k + m].re = a.re - b.re; out[k + m].im = a.im - b.im;
        }
        return;
    }
    if (p == 4) {
        for (int k = 0; k < m; k++) {
            cpx y[4];
            y[0] = out[k];
            for (int r = 1; r < 4; r++) {
                cpx w = tw[((size_t)r * k * tws) % N];
                if (sign > 0) w.im = -w.im;
                y[r] = cmul(out[k + r * m], w);
            }
            cpx t0 = {y[0].re + y[2].re, y[0].im + y[2].im};
            cpx t1 = {y[0].re - y[2].re, y[0].im - y[2].im};
            cpx t2 = {y[1].re + y[3].re, y[1].im + y[3].im};
            cpx t3 = {y[1].re - y[3].re, y[1].im - y[3].im};
            /* multiply t3 by -i (forward) or +i (inverse) */
            cpx t3r = sign < 0 ? (cpx){t3.im, -t3.re} : (cpx){-t3.im, t3.re};
            out[k].re = t0.re + t2.re; out[k].im = t0.im + t2.im;
            out[k + m].re = t1.re + t3r.re; out[k + m].im = t1.im + t3r.im;
            out[k + 2 * m].re = t0.re - t2.re; out[k + 2 * m].im = t0.im - t2.im;
            out[k + 3 * m].re = t1.re - t3r.re; out[k + 3 * m].im = t1.im - t3r.im;
        }
        return;
    }
    cpx ystack[64];
    cpx* y = p <= 64 ? ystack : (cpx*)xcalloc((size_t)p, sizeof(cpx));
    cpx zstack[64];
    cpx* zz = p <= 64 ? zstack : (cpx*)xcalloc((size_t)p, sizeof(cpx));
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) {
            cpx w = tw[((size_t)r * k * tws) % N];
            if (sign > 0) w.im = -w.im;
            y[r] = cmul(out[k + (size_t)r * m], w);
        }
        for (int s = 0; s < p; s++) {
            cpx acc = {0, 0};
            for (int r = 0; r < p; r++) {
                cpx w = tw[((size_t)r * s * m * tws) % N];
                if (sign > 0) w.im = -w.im;
                cpx t = cmul(y[r], w);
                acc.re += t.re; acc.im += t.im;
            }
            zz[s] = acc;
        }
        for (int s = 0; s < p; s++) out[k + (size_t)s * m] = zz[s];
    }
    if (p > 64) { free(y); free(zz); }
}

static void fft_line(cpx* line, cpx* tmp, const fftplan* pl, int sign) {
    memcpy(tmp, line, (size_t)pl->n * sizeof(cpx));
    fft_rec(tmp, line, pl->n, 1, pl->f, pl->tw, pl->n, 1, sign);
}

void or_fft_c2c(or_real* data, int n, int howmany, int sign) {
    const fftplan* pl = get_plan(n);
    cpx* d = (cpx*)data;
#pragma omp parallel
    {
        cpx* tmp = (cpx*)xcalloc((size_t)n, sizeof(cpx));
#pragma omp for schedule(static)
        for (int l = 0; l < howmany; l++) fft_line(d + (size_t)l * n, tmp, pl, sign);
        free(tmp);
    }
}

/* FFT along one axis of a [n0][n1][n2] complex array, restricted to lines whose
 * other two indices lie below lim (zero-skipping, gpu_fft6.cu:54-256 idea). */
static void fft3_axis(cpx* d, const int n[3], int axis, const int lim[3], int sign) {
    const fftplan* pl = get_plan(n[axis]);
    int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
    if (a1 > a2) { int t = a1; a1 = a2; a2 = t; }
    size_t strides[3] = {(size_t)n[1] * n[2], (size_t)n[2], 1};
    long nl = (long)lim[a1] * lim[a2];
    int L = n[axis];
#pragma omp parallel
    {
        cpx* line = (cpx*)xcalloc((size_t)L, sizeof(cpx));
        cpx* tmp = (cpx*)xcalloc((size_t)L, sizeof(cpx));
#pragma omp for schedule(static)
        for (long q = 0; q < nl; q++) {
            int i1 = (int)(q / lim[a2]), i2 = (int)(q % lim[a2]);
            cpx* base = d + i1 * strides[a1] + i2 * strides[a2];
            size_t st = strides[axis];
            for (int i = 0; i < L; i++) line[i] = base[i * st];
            fft_line(line, tmp, pl, sign);
            for (int i = 0; i < L; i++) base[i * st] = line[i];
        }
        free(line);
        free(tmp);
    }
}

void or_fft3_c2c(or_real* data, const int n[3], int sign) {
    for (int ax = 2; ax >= 0; ax--) fft3_axis((cpx*)data, n, ax, n, sign);
}

/* ====================================================================== */
/* convolution: conv.go:57-167, gpu_fft0.cu:96-109, gpu_kernmul.cu:25-141  */
/* ====================================================================== */
struct or_conv {
    int N[3], P[3];
    size_t NP;
    or_real* Kr; /* 6 * NP, real parts of FFT(K)/(P0 P1 P2) (conv.go:141-158) */
    cpx* buf[3];
};

or_conv* or_conv_new(const int N[3], const int P[3], const or_real* K6) {
    or_conv* c = (or_conv*)xcalloc(1, sizeof(or_conv));
    for (int i = 0; i < 3; i++) { c->N[i] = N[i]; c->P[i] = P[i]; }
    c->NP = (size_t)P[0] * P[1] * P[2];
    c->Kr = (or_real*)xcalloc(6 * c->NP, sizeof(or_real));
    for (int i = 0; i < 3; i++) c->buf[i] = (cpx*)xcalloc(c->NP, sizeof(cpx));
    double norm = 1.0 / (double)c->NP;
    for (int k = 0; k < 6; k++) {
        cpx* b = c->buf[0];
        for (size_t i = 0; i < c->NP; i++) { b[i].re = K6[k * c->NP + i]; b[i].im = 0; }
        or_fft3_c2c((or_real*)b, P, -1);
        for (size_t i = 0; i < c->NP; i++) c->Kr[k * c->NP + i] = (or_real)(b[i].re * norm);
    }
    return c;
}

void or_conv_free(or_conv* c) {
    if (!c) return;
    free(c->Kr);
    for (int i = 0; i < 3; i++) free(c->buf[i]);
    free(c);
}

void or_conv_kernel_spectrum(const or_conv* c, or_real* out) {
    int Pzc = c->P[2] / 2 + 1;
    for (int k = 0; k < 6; k++)
        for (int x = 0; x < c->P[0]; x++)
            for (int y = 0; y < c->P[1]; y++)
                for (int z = 0; z < Pzc; z++)
                    out[(((size_t)k * c->P[0] + x) * c->P[1] + y) * Pzc + z] =
                        c->Kr[k * c->NP + ((size_t)x * c->P[1] + y) * c->P[2] + z];
}

static void pad_into(const or_conv* c, const or_real* in, cpx* b) {
    memset(b, 0, c->NP * sizeof(cpx));
    for (int x = 0; x < c->N[0]; x++)
        for (int y = 0; y < c->N[1]; y++)
            for (int z = 0; z < c->N[2]; z++)
                b[((size_t)x * c->P[1] + y) * c->P[2] + z].re = in[((size_t)x * c->N[1] + y) * c->N[2] + z];
}

static void forward_padded(or_conv* c, cpx* b) {
    /* Z lines: only x<Nx, y<Ny are non-zero; Y lines: only x<Nx; X lines: all */
    int limZ[3] = {c->N[0], c->N[1], c->P[2]};
    int limY[3] = {c->N[0], c->P[1], c->P[2]};
    fft3_axis(b, c->P, 2, limZ, -1);
    fft3_axis(b, c->P, 1, limY, -1);
    fft3_axis(b, c->P, 0, c->P, -1);
}

static void inverse_padded(or_conv* c, cpx* b) {
    int limY[3] = {c->N[0], c->P[1], c->P[2]};
    int limZ[3] = {c->N[0], c->N[1], c->P[2]};
    fft3_axis(b, c->P, 0, c->P, +1);
    fft3_axis(b, c->P, 1, limY, +1);
    fft3_axis(b, c->P, 2, limZ, +1);
}

void or_conv_forward1(or_conv* c, const or_real* in, or_real* out) {
    cpx* b = c->buf[0];
    pad_into(c, in, b);
    forward_padded(c, b);
    int Pzc = c->P[2] / 2 + 1;
    for (int x = 0; x < c->P[0]; x++)
        for (int y = 0; y < c->P[1]; y++)
            for (int z = 0; z < Pzc; z++) {
                cpx v = b[((size_t)x * c->P[1] + y) * c->P[2] + z];
                size_t o = (((size_t)x * c->P[1] + y) * Pzc + z) * 2;
                out[o] = v.re;
                out[o + 1] = v.im;
            }
}

void or_conv_apply(or_conv* c, const or_real* m, or_real* h) {
    size_t N = (size_t)c->N[0] * c->N[1] * c->N[2];
    for (int i = 0; i < 3; i++) {
        pad_into(c, m + i * N, c->buf[i]);
        forward_padded(c, c->buf[i]);
    }
    const or_real *Kxx = c->Kr, *Kyy = c->Kr + c->NP, *Kzz = c->Kr + 2 * c->NP;
    const or_real *Kyz = c->Kr + 3 * c->NP, *Kxz = c->Kr + 4 * c->NP, *Kxy = c->Kr + 5 * c->NP;
    cpx *bx = c->buf[0], *by = c->buf[1], *bz = c->buf[2];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)c->NP; i++) {
        cpx mx = bx[i], my = by[i], mz = bz[i];
        bx[i].re = mx.re * Kxx[i] + my.re * Kxy[i] + mz.re * Kxz[i];
        bx[i].im = mx.im * Kxx[i] + my.im * Kxy[i] + mz.im * Kxz[i];
        by[i].re = mx.re * Kxy[i] + my.re * Kyy[i] + mz.re * Kyz[i];
        by[i].im = mx.im * Kxy[i] + my.im * Kyy[i] + mz.im * Kyz[i];
        bz[i].re = mx.re * Kxz[i] + my.re * Kyz[i] + mz.re * Kzz[i];
        bz[i].im = mx.im * Kxz[i] + my.im * Kyz[i] + mz.im * Kzz[i];
    }
    for (int i = 0; i < 3; i++) {
        inverse_padded(c, c->buf[i]);
        for (int x = 0; x < c->N[0]; x++)
            for (int y = 0; y < c->N[1]; y++)
                for (int z = 0; z < c->N[2]; z++)
                    h[i * N + ((size_t)x * c->N[1] + y) * c->N[2] + z] =
                        c->buf[i][((size_t)x * c->P[1] + y) * c->P[2] + z].re;
    }
}

void or_conv_direct(const int N[3], const int P[3], const or_real* K6, const or_real* m, or_real* h) {
    size_t n = (size_t)N[0] * N[1] * N[2], NP = (size_t)P[0] * P[1] * P[2];
    memset(h, 0, 3 * n * sizeof(or_real));
    for (int x = 0; x < N[0]; x++)
        for (int y = 0; y < N[1]; y++)
            for (int z = 0; z < N[2]; z++) {
                size_t o = ((size_t)x * N[1] + y) * N[2] + z;
                for (int xs = 0; xs < N[0]; xs++)
                    for (int ys = 0; ys < N[1]; ys++)
                        for (int zs = 0; zs < N[2]; zs++) {
                            size_t src = ((size_t)xs * N[1] + ys) * N[2] + zs;
                            size_t kid = ((size_t)wrapi(x - xs, P[0]) * P[1] + wrapi(y - ys, P[1])) * P[2] +
                                         wrapi(z - zs, P[2]);
                            for (int d = 0; d < 3; d++)
                                for (int s = 0; s < 3; s++)
                                    h[d * n + o] += K6[KERN_IDX[s][d] * NP + kid] * m[s * n + src];
                        }
            }
}

/* ====================================================================== */
/* pointwise: gpu_local_contr3.cu:21-50,238-256; gpu_torque.cu:20-50;      */
/* gpu_normalize.cu:20-30; gpu_linalg.cu:121-132; gpu_reduction.cu:576-630 */
/* ====================================================================== */
void or_local_fields(const or_real* m, or_real* h, int N, const or_real hext[3], int anisType, or_real k,
                     const or_real u[3]) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        or_real mu = 0;
        if (anisType == OR_ANIS_UNIAXIAL) mu = m[i] * u[0] + m[N + i] * u[1] + m[2 * N + i] * u[2];
        for (int c = 0; c < 3; c++) {
            or_real a = anisType == OR_ANIS_UNIAXIAL ? k * mu * u[c] : 0;
            h[c * N + i] += hext[c] + a;
        }
    }
}

void or_local_fields_phi(const or_real* m, or_real* h, or_real* phi, int N, const or_real hext[3], int anisType,
                         or_real k, const or_real u[3]) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        or_real mx = m[i], my = m[N + i], mz = m[2 * N + i];
        or_real mu = 0;
        if (anisType == OR_ANIS_UNIAXIAL) mu = mx * u[0] + my * u[1] + mz * u[2];
        for (int c = 0; c < 3; c++) {
            or_real a = anisType == OR_ANIS_UNIAXIAL ? k * mu * u[c] : 0;
            h[c * N + i] += hext[c] + a;
        }
        phi[i] -= mx * hext[0] + my * hext[1] + mz * hext[2] + (anisType == OR_ANIS_UNIAXIAL ? 0.5 * k * mu * mu : 0);
    }
}

void or_deltaM(const or_real* m, or_real* h, or_real alpha, or_real dtgilb, int N) {
    or_deltaM_map(m, h, alpha, NULL, dtgilb, N);
}

/* gpu_torque.cu with alpha_map: the damping of cell i is alpha * amap[i];
 * dtgilb stays that of the global alpha (torque.go:49-51) */
void or_deltaM_map(const or_real* m, or_real* h, or_real alpha_mul, const or_real* amap, or_real dtgilb, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        const or_real alpha = amap ? alpha_mul * amap[i] : alpha_mul;
        or_real Mx = m[i], My = m[N + i], Mz = m[2 * N + i];
        or_real Hx = h[i], Hy = h[N + i], Hz = h[2 * N + i];
        or_real cx = My * Hz - Mz * Hy, cy = Mz * Hx - Mx * Hz, cz = Mx * Hy - My * Hx; /* m x H */
        or_real dx = My * cz - Mz * cy, dy = Mz * cx - Mx * cz, dz = Mx * cy - My * cx; /* m x (m x H) */
        h[i] = dtgilb * (cx - alpha * dx);
        h[N + i] = dtgilb * (cy - alpha * dy);
        h[2 * N + i] = dtgilb * (cz - alpha * dz);
    }
}

void or_normalize(or_real* m, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double n2 = (double)m[i] * m[i] + (double)m[N + i] * m[N + i] + (double)m[2 * N + i] * m[2 * N + i];
        double inv = 1.0 / sqrt(n2);
        m[i] = (or_real)(m[i] * inv);
        m[N + i] = (or_real)(m[N + i] * inv);
        m[2 * N + i] = (or_real)(m[2 * N + i] * inv);
    }
}

/* gpu_anal.cu:23-84: the exact solution of the LL equation in a constant field over dt.
 * m is rotated into a frame whose first axis is h (the fixed frame for h along z),
 * the polar component advanced by the damping, the azimuth by the precession,
 * and rotated back; a cell with m = 0 is left as it is. */
void or_anal_step(const or_real* min, or_real* mout, const or_real* h, double dt, double alpha, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        const double mx = min[i], my = min[N + i], mz = min[2 * N + i];
        if (mx == 0 && my == 0 && mz == 0) continue;
        const double hx = h[i], hy = h[N + i], hz = h[2 * N + i];
        double r0, r1, r2, r3, r4, r5, r6, r8, hr; /* rows of the rotation; hr = 1/|h| */
        if (hx == 0 && hy == 0) {
            r0 = 0; r1 = 0; r2 = -1; r3 = 0; r4 = 1; r5 = 0; r6 = 1; r8 = 0;
            hr = 1.0 / hz;
        } else {
            const double t = hx * hx + hy * hy, hxy = 1.0 / sqrt(t);
            hr = 1.0 / sqrt(t + hz * hz);
            r0 = hx * hr; r1 = -hy * hxy; r2 = -r0 * hz * hxy;
            r3 = hy * hr; r4 = hx * hxy; r5 = r1 * hz * hr;
            r6 = hz * hr; r8 = hr / hxy;
        }
        const double a = mx * r0 + my * r3 + mz * r6, b = mx * r1 + my * r4, c = mx * r2 + my * r5 + mz * r8;
        const double qt = dt / (1 + alpha * alpha), ex = exp(alpha * qt / hr), sn = sin(qt / hr), cs = cos(qt / hr);
        const double den = ex * (1 + a) + (1 - a) / ex;
        const double an = (ex * (1 + a) - (1 - a) / ex) / den, bn = 2 * (b * cs - c * sn) / den,
                     cn = 2 * (b * sn + c * cs) / den;
        mout[i] = (or_real)(an * r0 + bn * r1 + cn * r2);
        mout[N + i] = (or_real)(an * r3 + bn * r4 + cn * r5);
        mout[2 * N + i] = (or_real)(an * r6 + cn * r8);
    }
}

/* gpu_spintorque.cu:24-180 (Zhang-Li torque): Hspin = sum_d u_d (m(+1) - m(-1))
 * along each axis, the cell itself standing in for a missing neighbour at the
 * box edges (:58-83), u_d scaled by the mask's component d of the cell (NULL: 1);
 * then h <- dtg (m x (H + eps Hspin) - m x (m x (alpha H + beta Hspin))) */
void or_spintorque_deltaM(const or_real* m, or_real* h, double alpha, double beta, double eps, const double u[3],
                          const or_real* jmap, double dtg, const int n[3]) {
    const long N = (long)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (long I = 0; I < N; I++) {
        const int k = (int)(I % n[2]);
        const long r = I / n[2];
        const int j = (int)(r % n[1]), i = (int)(r / n[1]);
        double ud[3] = {u[0], u[1], u[2]};
        if (jmap)
            for (int d = 0; d < 3; d++) ud[d] *= jmap[d * N + I];
        const long st[3] = {(long)n[1] * n[2], n[2], 1};
        const int at[3] = {i, j, k};
        double hs[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++) {
            const long lo = at[d] > 0 ? I - st[d] : I, hi = at[d] + 1 < n[d] ? I + st[d] : I;
            for (int c = 0; c < 3; c++) hs[c] += ud[d] * ((double)m[c * N + hi] - (double)m[c * N + lo]);
        }
        const double M[3] = {m[I], m[N + I], m[2 * N + I]}, H[3] = {h[I], h[N + I], h[2 * N + I]};
        double p[3], q[3], a[3], c[3], e[3];
        for (int x = 0; x < 3; x++) {
            p[x] = H[x] + eps * hs[x];
            q[x] = alpha * H[x] + beta * hs[x];
        }
        a[0] = M[1] * p[2] - M[2] * p[1]; a[1] = M[2] * p[0] - M[0] * p[2]; a[2] = M[0] * p[1] - M[1] * p[0];
        c[0] = M[1] * q[2] - M[2] * q[1]; c[1] = M[2] * q[0] - M[0] * q[2]; c[2] = M[0] * q[1] - M[1] * q[0];
        e[0] = M[1] * c[2] - M[2] * c[1]; e[1] = M[2] * c[0] - M[0] * c[2]; e[2] = M[0] * c[1] - M[1] * c[0];
        for (int x = 0; x < 3; x++) h[x * N + I] = (or_real)(dtg * (a[x] - e[x]));
    }
}

/* gpu_normalize.cu:43-71 (_gpu_normalize_map): each cell to the norm nmap[i];
 * a cell whose norm is 0 becomes 0 (the kernel's special case, so m = 0 there
 * gives 0 rather than 0/0) */
void or_normalize_map(or_real* m, const or_real* nmap, int N) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double n2 = (double)m[i] * m[i] + (double)m[N + i] * m[N + i] + (double)m[2 * N + i] * m[2 * N + i];
        double sc = nmap[i] == 0 ? 0.0 : (double)nmap[i] / sqrt(n2);
        m[i] = (or_real)(m[i] * sc);
        m[N + i] = (or_real)(m[N + i] * sc);
        m[2 * N + i] = (or_real)(m[2 * N + i] * sc);
    }
}

void or_scale_dot(or_real* r, const or_real* a, const or_real* b, or_real s, int N) {
    for (int i = 0; i < N; i++) r[i] = s * (a[i] * b[i] + a[N + i] * b[N + i] + a[2 * N + i] * b[2 * N + i]);
}

double or_reduce(int op, const or_real* x, int n) {
    double acc;
    switch (op) {
    case OR_REDUCE_ADD:
        acc = 0;
        for (int i = 0; i < n; i++) acc += x[i];
        return acc;
    case OR_REDUCE_SUMABS:
        acc = 0;
        for (int i = 0; i < n; i++) acc += fabs((double)x[i]);
        return acc;
    case OR_REDUCE_MAX:
        acc = x[0];
        for (int i = 0; i < n; i++) if (x[i] > acc) acc = x[i];
        return acc;
    case OR_REDUCE_MIN:
        acc = x[0];
        for (int i = 0; i < n; i++) if (x[i] < acc) acc = x[i];
        return acc;
    case OR_REDUCE_MAXABS:
        acc = 0;
        for (int i = 0; i < n; i++) if (fabs((double)x[i]) > acc) acc = fabs((double)x[i]);
        return acc;
    }
    fprintf(stderr, "oracle: bad reduce op %d\n", op);
    abort();
}

void or_add_exch6_stencil(const or_real* m, or_real* h, const int size[3], const int periodic[3],
                          const or_real cell[3]) {
    /* gpu_exch1.cu:21-301: per component h += c0 m + sum_d c_d m(neighbour), a missing
     * neighbour counts as zero (free boundary) unless the direction is periodic. */
    int n0 = size[0], n1 = size[1], n2 = size[2];
    size_t N = (size_t)n0 * n1 * n2;
    or_real cx = 1 / (cell[X] * cell[X]), cy = 1 / (cell[Y] * cell[Y]), cz = 1 / (cell[Z] * cell[Z]);
    or_real c0 = n0 == 1 ? -2 * cy - 2 * cz : -2 * cx - 2 * cy - 2 * cz;
    for (int c = 0; c < 3; c++) {
        const or_real* mc = m + c * N;
        or_real* hc = h + c * N;
#pragma omp parallel for schedule(static)
        for (int x = 0; x < n0; x++)
            for (int y = 0; y < n1; y++)
                for (int z = 0; z < n2; z++) {
                    size_t i = ((size_t)x * n1 + y) * n2 + z;
                    or_real r = c0 * mc[i];
#define NB(xx, yy, zz, coef)                                                                            \
    do {                                                                                                \
        int a_ = (xx), b_ = (yy), d_ = (zz);                                                            \
        if (periodic[0]) a_ = (a_ + n0) % n0;                                                           \
        if (periodic[1]) b_ = (b_ + n1) % n1;                                                           \
        if (periodic[2]) d_ = (d_ + n2) % n2;                                                           \
        if (a_ >= 0 && a_ < n0 && b_ >= 0 && b_ < n1 && d_ >= 0 && d_ < n2)                             \
            r += (coef) * mc[((size_t)a_ * n1 + b_) * n2 + d_];                                         \
    } while (0)
                    if (n0 > 1) { NB(x - 1, y, z, cx); NB(x + 1, y, z, cx); }
                    NB(x, y - 1, z, cy);
                    NB(x, y + 1, z, cy);
                    NB(x, y, z - 1, cz);
                    NB(x, y, z + 1, cz);
#undef NB
                    hc[i] += r;
                }
    }
}

/* ====================================================================== */
/* simulation: sim.go, init.go:80-287, field.go, torque.go, rk.go, simrun  */
/* ====================================================================== */
#define MAX_STAGES 7
#define MAX_STEP_TRIALS 10 /* rk.go:277 */

static unsigned long long or_mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static unsigned long long thermal_key(unsigned long long seed, long long step, long cell, int j) {
    unsigned long long base = or_mix64(seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(step + 1));
    return or_mix64(base + 0x9E3779B97F4A7C15ull * (unsigned long long)(4 * cell + j + 1));
}
void or_thermal_normals(unsigned long long seed, long long step, long cell, double out[3]) {
    const double u1 = ((double)(thermal_key(seed, step, cell, 0) >> 40) + 0.5) * 0x1p-24;
    const double u2 = (double)(thermal_key(seed, step, cell, 1) >> 40) * 0x1p-24;
    const double u3 = ((double)(thermal_key(seed, step, cell, 2) >> 40) + 0.5) * 0x1p-24;
    const double u4 = (double)(thermal_key(seed, step, cell, 3) >> 40) * 0x1p-24;
    const double pi = 3.14159265358979323846;
    const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    out[0] = r1 * cos(2.0 * pi * u2);
    out[1] = r1 * sin(2.0 * pi * u2);
    out[2] = r2 * cos(2.0 * pi * u4);
}

struct or_sim {
    or_params p;
    double* fieldT;
    double* fieldB;
    int fieldLastIdx; /* pointwiseField.lastIdx: only moves forward (simfield.go:126-133) */
    int N;
    double units[5];
    or_real cell[3];
    or_real kInt, axis[3];
    or_conv* conv;
    or_real *m, *m1, *mbackup, *k[MAX_STAGES], *phi, *tmp;
    int stages, fsal, fsal_initiated, has_b2;
    double errororder;
    double a[MAX_STAGES][MAX_STAGES], b[MAX_STAGES], c[MAX_STAGES], b2[MAX_STAGES];
    double time, dt;
    long steps, trials_total, rejects_total;
    double lastErr, maxTorque, energy, hextSI[3];
    double* log_dt; /* per trial: dt used and accepted flag (the GPU engine's trial log) */
    int* log_acc;
    long log_n, log_cap;
    or_real* amap; /* per-cell damping multiplier (SetAlpha, simgeom.go:246-264), NULL: none */
    or_real* nmap; /* per-cell norm of m (SetMsat, simgeom.go:266-340), NULL: 1 */
    double j[3];   /* current density, A/m^2 (CurrentDensity, simspintorque.go:30-36) */
    double spinPol, xi;
    or_real* jmask; /* 3N current mask (CurrentMask), NULL: 1 */
    or_real *sa_m2, *sa_m3, *sa_h; /* semi-analytic solvers' m2, m3 and field (allocated at their first step) */
    int japplied;   /* an applied current (Apply*("j", ...), simfield.go:61-70) replaces j at each torque */
    int jn, jfn, jLastIdx;
    double *jT, *jV, jfp[4], jfb[3];
    int fn;        /* applied-field function of time (1 pulse, 2 RF, 3 RF ramp, 4 sawtooth, 5 rotating), 0: table */
    double fp[4], fb[3];
    or_real* noise; /* thermal field of the current step, 3N (NULL at T = 0) */
};

/* Sim.Normalize (sim.go:230-239): to the norm map when there is one */
static void sim_normalize(or_sim* s, or_real* m) {
    if (s->nmap)
        or_normalize_map(m, s->nmap, s->N);
    else
        or_normalize(m, s->N);
}

/* updateTempNoise (temperature.go:29-36): the thermal field for this step from
 * dt (internal) at its start, stddev = sqrt(2 alpha kB T / (gamma0 mu0 Ms V dt_SI)) / Ms */
static void update_temp_noise(or_sim* s, double dt) {
    const double kB = 1.380650424e-23, gamma0 = 2.211e5, mu0 = 4e-7 * 3.14159265358979323846;
    const double V = s->p.cellsize[0] * s->p.cellsize[1] * s->p.cellsize[2];
    const double sd = sqrt(2 * s->p.alpha * kB * s->p.temperature / (gamma0 * mu0 * s->p.msat * V * dt * s->units[1])) /
                      s->p.msat;
    const int N = s->N;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; i++) {
        double n[3];
        or_thermal_normals(s->p.seed, s->steps, i, n);
        for (int c = 0; c < 3; c++) s->noise[c * (size_t)N + i] = (or_real)(sd * n[c]);
    }
}

static void log_trial(or_sim* s, double dt, int acc) {
    if (s->log_n == s->log_cap) {
        s->log_cap = s->log_cap ? 2 * s->log_cap : 1024;
        s->log_dt = (double*)realloc(s->log_dt, (size_t)s->log_cap * sizeof(double));
        s->log_acc = (int*)realloc(s->log_acc, (size_t)s->log_cap * sizeof(int));
        if (!s->log_dt || !s->log_acc) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    }
    s->log_dt[s->log_n] = dt;
    s->log_acc[s->log_n] = acc;
    s->log_n++;
}

static int g_tableau_f32 = 0;
void or_set_tableau_f32(int on) { g_tableau_f32 = on; }

static void set_tableau(or_sim* s) {
    /* coefficients as rationals (rk.go:62-212); the reference rounds them to float32 */
    memset(s->a, 0, sizeof(s->a));
    memset(s->b, 0, sizeof(s->b));
    memset(s->c, 0, sizeof(s->c));
    memset(s->b2, 0, sizeof(s->b2));
    s->fsal = 0;
    s->has_b2 = 0;
    switch (s->p.solver == OR_SEMIANAL1 || s->p.solver == OR_SEMIANAL2 ? OR_RK2 : s->p.solver) {
    case OR_RK2:
    case OR_RK12:
        s->stages = 2;
        s->c[1] = 1;
        s->a[1][0] = 1;
        s->b[0] = .5; s->b[1] = .5;
        if (s->p.solver == OR_RK12) { s->has_b2 = 1; s->errororder = 1; s->b2[0] = 1; s->b2[1] = 0; }
        break;
    case OR_RK3:
        s->stages = 3;
        s->c[1] = .5; s->c[2] = 1;
        s->a[1][0] = .5; s->a[2][0] = -1; s->a[2][1] = 2;
        s->b[0] = 1. / 6; s->b[1] = 2. / 3; s->b[2] = 1. / 6;
        break;
    case OR_RK23:
        s->stages = 4;
        s->fsal = 1;
        s->c[1] = .5; s->c[2] = .75; s->c[3] = 1;
        s->a[1][0] = .5;
        s->a[2][1] = .75;
        s->a[3][0] = 2. / 9; s->a[3][1] = 1. / 3; s->a[3][2] = 4. / 9;
        s->b[0] = 2. / 9; s->b[1] = 1. / 3; s->b[2] = 4. / 9;
        s->has_b2 = 1; s->errororder = 2;
        s->b2[0] = 7. / 24; s->b2[1] = .25; s->b2[2] = 1. / 3; s->b2[3] = 1. / 8;
        break;
    case OR_RK4:
        s->stages = 4;
        s->c[1] = .5; s->c[2] = .5; s->c[3] = 1;
        s->a[1][0] = .5; s->a[2][1] = .5; s->a[3][2] = 1;
        s->b[0] = 1. / 6; s->b[1] = 1. / 3; s->b[2] = 1. / 3; s->b[3] = 1. / 6;
        break;
    case OR_RKCK:
        s->stages = 6;
        s->c[1] = .2; s->c[2] = .3; s->c[3] = .6; s->c[4] = 1; s->c[5] = 7. / 8;
        s->a[1][0] = .2;
        s->a[2][0] = 3. / 40; s->a[2][1] = 9. / 40;
        s->a[3][0] = .3; s->a[3][1] = -.9; s->a[3][2] = 1.2;
        s->a[4][0] = -11. / 54; s->a[4][1] = 2.5; s->a[4][2] = -70. / 27; s->a[4][3] = 35. / 27;
        s->a[5][0] = 1631. / 55296; s->a[5][1] = 175. / 512; s->a[5][2] = 575. / 13824;
        s->a[5][3] = 44275. / 110592; s->a[5][4] = 253. / 4096;
        s->b[0] = 37. / 378; s->b[2] = 250. / 621; s->b[3] = 125. / 594; s->b[5] = 512. / 1771;
        s->has_b2 = 1; s->errororder = 4;
        s->b2[0] = 2825. / 27648; s->b2[2] = 18575. / 48384; s->b2[3] = 13525. / 55296;
        s->b2[4] = 277. / 14336; s->b2[5] = .25;
        break;
    case OR_RKDP:
        s->stages = 7;
        s->fsal = 1;
        s->c[1] = .2; s->c[2] = .3; s->c[3] = .8; s->c[4] = 8. / 9; s->c[5] = 1; s->c[6] = 1;
        s->a[1][0] = .2;
        s->a[2][0] = 3. / 40; s->a[2][1] = 9. / 40;
        s->a[3][0] = 44. / 45; s->a[3][1] = -56. / 15; s->a[3][2] = 32. / 9;
        s->a[4][0] = 19372. / 6561; s->a[4][1] = -25360. / 2187; s->a[4][2] = 64448. / 6561; s->a[4][3] = -212. / 729;
        s->a[5][0] = 9017. / 3168; s->a[5][1] = -355. / 33; s->a[5][2] = 46732. / 5247; s->a[5][3] = 49. / 176;
        s->a[5][4] = -5103. / 18656;
        s->a[6][0] = 35. / 384; s->a[6][2] = 500. / 1113; s->a[6][3] = 125. / 192; s->a[6][4] = -2187. / 6784;
        s->a[6][5] = 11. / 84;
        s->b[0] = 35. / 384; s->b[2] = 500. / 1113; s->b[3] = 125. / 192; s->b[4] = -2187. / 6784; s->b[5] = 11. / 84;
        s->has_b2 = 1; s->errororder = 4;
        s->b2[0] = 5179. / 57600; s->b2[2] = 7571. / 16695; s->b2[3] = 393. / 640; s->b2[4] = -92097. / 339200;
        s->b2[5] = 187. / 2100; s->b2[6] = 1. / 40;
        break;
    default:
        fprintf(stderr, "oracle: unknown solver %d\n", s->p.solver);
        abort();
    }
    if (g_tableau_f32) { /* the reference's float32 coefficients (rk.go:62-212 are float32 literals) */
        for (int r = 0; r < MAX_STAGES; r++) {
            for (int q = 0; q < MAX_STAGES; q++) s->a[r][q] = (float)s->a[r][q];
            s->b[r] = (float)s->b[r];
            s->b2[r] = (float)s->b2[r];
            s->c[r] = (float)s->c[r];
        }
    }
}

/* applied field in tesla at SI time t (simfield.go:87-153) */
/* pulsedField, rfField, rfRamp, sawtooth, rotatingField .GetAppliedField
 * (simfield.go:155-260): scale in double, rounded to float32, times the
 * float32 amplitude */
static void field_function(int fn, const double fp[4], const double fb[3], double t, double B[3]) {
    const double pi = 3.14159265358979323846, f = fp[0];
    double sc[3] = {0, 0, 0};
    switch (fn) {
    case 1: {
        const double d = fp[0], r = fp[1];
        double v = 0;
        if (t > 0 && t < r) v = 0.5 - 0.5 * cos(t * pi / r);
        else if (t >= r && t < d + r) v = 1;
        else if (t >= d + r && t < d + 2 * r) v = 0.5 + 0.5 * cos((t - d - r) * pi / r);
        sc[0] = sc[1] = sc[2] = v;
        break;
    }
    case 2: sc[0] = sc[1] = sc[2] = sin(f * 2 * pi * t); break;
    case 3: sc[0] = sc[1] = sc[2] = sin(f * 2 * pi * t) * (t < fp[1] ? t / fp[1] : 1.0); break;
    case 4: sc[0] = sc[1] = sc[2] = asin(sin(f * 2 * pi * t)) / (pi / 2); break;
    case 5: for (int c = 0; c < 3; c++) sc[c] = sin(f * 2 * pi * t + fp[1 + c]); break;
    }
    for (int c = 0; c < 3; c++) B[c] = (double)((float)sc[c] * (float)fb[c]);
}

/* an AppliedField (simfield.go:87-260): a function of time, a static value (n = 1)
 * or n points interpolated linearly with the forward-only search index *last */
static void applied_eval(int n, const double* T, const double* V, int* last, int fn, const double fp[4],
                         const double fb[3], double t, double B[3]) {
    B[0] = B[1] = B[2] = 0;
    if (fn) { field_function(fn, fp, fb, t, B); return; }
    if (n == 0) return;
    if (n == 1) { for (int i = 0; i < 3; i++) B[i] = V[i]; return; }
    for (; *last < n; (*last)++)
        if (T[*last] >= t) break;
    int i = *last;
    if (i - 1 < 0 || i >= n) return;
    double dt = T[i] - T[i - 1];
    double f = (t - T[i - 1]) / dt;
    for (int c = 0; c < 3; c++) B[c] = V[3 * (i - 1) + c] + f * (V[3 * i + c] - V[3 * (i - 1) + c]);
}

static void applied_field(or_sim* s, double t, double B[3]) {
    applied_eval(s->p.nField, s->fieldT, s->fieldB, &s->fieldLastIdx, s->fn, s->fp, s->fb, t, B);
}

or_sim* or_sim_new(const or_params* p, const float* m0) {
    or_sim* s = (or_sim*)xcalloc(1, sizeof(or_sim));
    s->p = *p;
    if (p->nField > 0) {
        s->fieldT = (double*)xcalloc((size_t)p->nField, sizeof(double));
        s->fieldB = (double*)xcalloc((size_t)p->nField * 3, sizeof(double));
        memcpy(s->fieldT, p->fieldT, (size_t)p->nField * sizeof(double));
        memcpy(s->fieldB, p->fieldB, (size_t)p->nField * 3 * sizeof(double));
    }
    s->p.fieldT = s->fieldT;
    s->p.fieldB = s->fieldB;
    s->N = p->size[0] * p->size[1] * p->size[2];
    or_units(p->msat, p->aexch, s->units);
    for (int i = 0; i < 3; i++) s->cell[i] = (or_real)(p->cellsize[i] / s->units[0]); /* init.go:139-148 */
    s->kInt = (or_real)(p->k1 / s->units[4]);                                          /* init.go:103-108 */
    for (int i = 0; i < 3; i++) s->axis[i] = (or_real)p->anisAxis[i];
    size_t n3 = 3 * (size_t)s->N;
    s->m = (or_real*)xcalloc(n3, sizeof(or_real));
    s->m1 = (or_real*)xcalloc(n3, sizeof(or_real));
    s->mbackup = (or_real*)xcalloc(n3, sizeof(or_real));
    s->tmp = (or_real*)xcalloc(n3, sizeof(or_real));
    s->phi = (or_real*)xcalloc((size_t)s->N, sizeof(or_real));
    for (int i = 0; i < MAX_STAGES; i++) s->k[i] = (or_real*)xcalloc(n3, sizeof(or_real));
    for (size_t i = 0; i < n3; i++) s->m[i] = m0[i];
    /* initConv (init.go:234-274) */
    if (p->wantDemag) {
        int P[3];
        or_pad_size(p->size, p->periodic, P);
        size_t NP = (size_t)P[0] * P[1] * P[2];
        or_real* K = (or_real*)xcalloc(6 * NP, sizeof(or_real));
        or_facekernel6(P, s->cell, p->demagAccuracy > 0 ? p->demagAccuracy : 8, p->periodic, K);
        if (p->exchInConv && p->exchType == 6) or_add_exch6_kernel(P, s->cell, K);
        s->conv = or_conv_new(p->size, P, K);
        free(K);
    }
    /* initSolver (init.go:277-287) */
    s->dt = p->dt / s->units[1];
    if (s->dt == 0) s->dt = 1e-4; /* DEFAULT_DT_INTERNAL, sim.go:30 */
    set_tableau(s);
    if (p->temperature > 0) {
        s->noise = (or_real*)xcalloc(n3, sizeof(or_real));
        update_temp_noise(s, s->dt);
    }
    return s;
}

void or_sim_free(or_sim* s) {
    if (!s) return;
    or_conv_free(s->conv);
    free(s->m); free(s->m1); free(s->mbackup); free(s->tmp); free(s->phi);
    for (int i = 0; i < MAX_STAGES; i++) free(s->k[i]);
    free(s->fieldT); free(s->fieldB);
    free(s->log_dt); free(s->log_acc);
    free(s->amap);
    free(s->nmap);
    free(s->jmask);
    free(s->jT);
    free(s->jV);
    free(s->sa_m2);
    free(s->sa_m3);
    free(s->sa_h);
    free(s->noise);
    free(s);
}

static void madd(or_real* a, double c, const or_real* b, size_t n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; i++) a[i] += (or_real)(c * b[i]);
}

static void update_hext(or_sim* s, or_real hint[3]) { /* field.go:50-61 */
    applied_field(s, s->time * s->units[1], s->hextSI);
    for (int i = 0; i < 3; i++) hint[i] = (or_real)(s->hextSI[i] / s->units[2]);
}

/* calcHeff (field.go:74-80) / calcHeffEnergy (field.go:83-87) */
static void calc_heff(or_sim* s, const or_real* m, or_real* h, int energy) {
    int N = s->N;
    if (s->p.wantDemag) or_conv_apply(s->conv, m, h);
    else memset(h, 0, 3 * (size_t)N * sizeof(or_real));
    if (!s->p.wantDemag || !s->p.exchInConv)
        if (s->p.exchType == 6) or_add_exch6_stencil(m, h, s->p.size, s->p.periodic, s->cell);
    or_real hint[3];
    update_hext(s, hint);
    if (energy) {
        or_scale_dot(s->phi, m, h, -0.5, N);
        or_local_fields_phi(m, h, s->phi, N, hint, s->p.anisType, s->kInt, s->axis);
        double vol = (double)s->cell[0] * s->cell[1] * s->cell[2];
        s->energy = or_reduce(OR_REDUCE_ADD, s->phi, N) * vol;
    } else {
        or_local_fields(m, h, N, hint, s->p.anisType, s->kInt, s->axis);
    }
    if (s->noise) madd(h, 1.0, s->noise, 3 * (size_t)N); /* addThermalField (temperature.go:20-23) */
}

static void torque(or_sim* s, const or_real* m, or_real* h) { /* torque.go:33-53,98-102 */
    double a = s->p.alpha;
    double j[3] = {s->j[0], s->j[1], s->j[2]};
    if (s->japplied) /* s.input.j from the applied current at the stage time (torque.go:42-44) */
        applied_eval(s->jn, s->jT, s->jV, &s->jLastIdx, s->jfn, s->jfp, s->jfb, s->time * s->units[1], j);
    if (s->japplied || j[0] != 0 || j[1] != 0 || j[2] != 0) { /* torque.go:40-49: the STT form */
        /* SpintorqueDeltaM (torque.go:64-91) with dt = 1; units material.go:83-110 */
        const double L = s->units[0], T = s->units[1], Ms = s->p.msat;
        const double muB = 9.2740091523E-24 / (Ms * L * L * L); /* init.go:84, / UnitMoment */
        const double e = 1.60217646E-19 / (Ms * L * T);         /* init.go:85, / UnitCharge */
        const double xi = s->xi;
        const double b = s->spinPol * muB / (e * (1 + xi * xi));
        double u[3];
        for (int i = 0; i < 3; i++) u[i] = 0.5 * (j[i] / (Ms / L)) / (double)s->cell[i];
        or_spintorque_deltaM(m, h, a, b * (1 + a * xi), b * (xi - a), u, s->jmask, 1.0 / (1.0 + a * a), s->p.size);
        return;
    }
    or_deltaM_map(m, h, (or_real)a, s->amap, (or_real)(1.0 / (1.0 + a * a)), s->N);
}


/* one RK step, rk.go:279-445; replay != NULL overrides dt and acceptance per trial */
static int rk_step(or_sim* s, const double* dt_trial, const int* accept, int n_trial) {
    int order = s->stages;
    size_t n3 = 3 * (size_t)s->N;
    double time0 = s->time;
    or_real *m = s->m, *m1 = s->m1;
    memcpy(s->mbackup, m, n3 * sizeof(or_real));
    int goodstep = 0, trials = 0;
    double prevEnergy = s->energy == 0 ? 9999999999.0 : s->energy;
    /* the thermal field is constant during the step (rk.go:289-292); a replayed
     * step draws it from the dt its first trial used */
    if (s->noise) update_temp_noise(s, (dt_trial && n_trial > 0) ? dt_trial[0] : s->dt);
    while (!goodstep && trials < MAX_STEP_TRIALS) {
        int t = trials;
        trials++;
        s->trials_total++;
        for (int i = 0; i < order; i++) {
            if (s->fsal && i == 0 && s->fsal_initiated) {
                memcpy(s->k[0], s->k[order - 1], n3 * sizeof(or_real));
            } else {
                s->time = time0 + s->c[i] * s->dt;
                memcpy(m1, m, n3 * sizeof(or_real));
                for (int j = 0; j < order; j++)
                    if (s->a[i][j] != 0) madd(m1, s->dt * s->a[i][j], s->k[j], n3);
                int wantE = s->p.wantEnergy && (s->fsal ? i == order - 1 : i == 0);
                calc_heff(s, m1, s->k[i], wantE);
                torque(s, m1, s->k[i]);
                s->fsal_initiated = 1;
            }
            if (i == 0) {
                if (s->p.temperature == 0) { /* rk.go:348-373: only at T = 0 */
                    if (!s->has_b2) s->dt = s->p.dt / s->units[1];
                    double maxTorque = or_reduce(OR_REDUCE_MAXABS, s->k[0], (int)n3);
                    s->maxTorque = maxTorque;
                    double dm = s->dt * maxTorque;
                    if (dm < s->p.minDm) s->dt = s->p.minDm / maxTorque;
                    if (s->dt * s->units[1] < s->p.minDt) s->dt = s->p.minDt / s->units[1];
                    if (s->p.maxDm != 0 && dm > s->p.maxDm) s->dt = s->p.maxDm / maxTorque;
                    if (s->p.maxDt != 0 && s->dt * s->units[1] > s->p.maxDt) s->dt = s->p.maxDt / s->units[1];
                }
                if (dt_trial && t < n_trial) s->dt = dt_trial[t];
                if (isnan(s->dt)) { fprintf(stderr, "oracle: dt = NaN\n"); return -1; }
            }
        }
        s->time = time0 + s->dt; /* rk.go:376 */
        const double dt_used = s->dt;
        if (s->has_b2) {
            memcpy(m1, m, n3 * sizeof(or_real));
            for (int i = 0; i < order; i++)
                if (s->b2[i] != 0) madd(m1, s->b2[i] * s->dt, s->k[i], n3);
        }
        for (int i = 0; i < order; i++)
            if (s->b[i] != 0) madd(m, s->b[i] * s->dt, s->k[i], n3);
        if (s->has_b2) {
            madd(m1, -1, m, n3);
            double err = or_reduce(OR_REDUCE_MAXABS, m1, (int)n3);
            s->lastErr = err;
            double factor = pow(s->p.maxError / err, 1.0 / s->errororder);
            if (factor > 2) factor = 2;
            if (factor < 0.01) factor = 0.01;
            s->dt = s->dt * factor;
            if (isnan(s->dt)) { fprintf(stderr, "oracle: dt = NaN\n"); return -1; }
            int bad = err > 2 * s->p.maxError || (s->p.wantEnergy && s->energy > prevEnergy);
            if (dt_trial && t < n_trial) bad = !accept[t];
            log_trial(s, dt_used, !bad);
            if (bad) {
                memcpy(m, s->mbackup, n3 * sizeof(or_real));
                goodstep = 0;
                s->rejects_total++;
            } else {
                goodstep = 1;
                if (s->dt * s->units[1] < s->p.minDt) s->dt = s->p.minDt / s->units[1];
                if (s->p.maxDt != 0 && s->dt * s->units[1] > s->p.maxDt) s->dt = s->p.maxDt / s->units[1];
            }
        } else {
            goodstep = 1;
            log_trial(s, dt_used, 1);
        }
    }
    sim_normalize(s, m); /* rk.go:444 */
    return trials;
}

/* SemiAnal1.step / SemiAnal2.step (semianal.go:23-58, 87-140): m2 carries over
 * between steps, the time is not advanced; m, m2 normalised every 100 steps */
static void semianal_step(or_sim* s) {
    const int N = s->N;
    const size_t n3 = 3 * (size_t)N;
    if (!s->sa_m2) {
        s->sa_m2 = (or_real*)xcalloc(n3, sizeof(or_real));
        s->sa_m3 = (or_real*)xcalloc(n3, sizeof(or_real));
        s->sa_h = (or_real*)xcalloc(n3, sizeof(or_real));
    }
    or_real *m1 = s->m, *m2 = s->sa_m2, *m3 = s->sa_m3, *h = s->sa_h;
    const double dt = s->p.dt / s->units[1], a = s->p.alpha;
    if (s->steps == 0) {
        calc_heff(s, m1, h, 0);
        or_anal_step(m1, m2, h, dt / 2, a, N);
        calc_heff(s, m2, h, 0);
        or_anal_step(m1, m1, h, dt, a, N);
    } else if (s->p.solver == OR_SEMIANAL1) {
        calc_heff(s, m1, h, 0);
        or_anal_step(m2, m2, h, dt, a, N);
        calc_heff(s, m2, h, 0);
        or_anal_step(m1, m1, h, dt, a, N);
    } else {
        calc_heff(s, m1, h, 0);
        or_anal_step(m2, m3, h, dt, a, N);
        for (size_t i = 0; i < n3; i++) m3[i] = (or_real)(0.5 * m3[i] + 0.5 * m2[i]); /* LinearCombination */
        calc_heff(s, m3, h, 0);
        or_anal_step(m2, m2, h, dt, a, N);
        calc_heff(s, m2, h, 0);
        or_anal_step(m1, m3, h, dt, a, N);
        for (size_t i = 0; i < n3; i++) m3[i] = (or_real)(0.5 * m3[i] + 0.5 * m1[i]);
        calc_heff(s, m3, h, 0);
        or_anal_step(m1, m1, h, dt, a, N);
    }
    if (s->steps % 100 == 0) {
        sim_normalize(s, m1);
        sim_normalize(s, m2);
    }
}

int or_sim_step(or_sim* s) {
    if (s->p.solver == OR_SEMIANAL1 || s->p.solver == OR_SEMIANAL2) {
        semianal_step(s);
        s->steps++;
        return 1;
    }
    int r = rk_step(s, NULL, NULL, 0);
    s->steps++;
    return r;
}

int or_sim_step_replay(or_sim* s, const double* dt_trial, const int* accept, int n_trial) {
    int r = rk_step(s, dt_trial, accept, n_trial);
    s->steps++;
    return r;
}

int or_sim_run(or_sim* s, double seconds) { /* simrun.go:19-57 */
    double stop = s->time + seconds / s->units[1];
    sim_normalize(s, s->m); /* simrun.go:24 */
    int n = 0;
    while (s->time < stop) {
        if (or_sim_step(s) < 0) return -1;
        n++;
        if (isnan(s->time) || isinf(s->time)) return -1;
    }
    return n;
}

void or_sim_get_m(const or_sim* s, double* m) {
    for (size_t i = 0; i < 3 * (size_t)s->N; i++) m[i] = s->m[i];
}

void or_sim_set_m(or_sim* s, const double* m) {
    for (size_t i = 0; i < 3 * (size_t)s->N; i++) s->m[i] = (or_real)m[i];
}

void or_sim_state(const or_sim* s, double* st) {
    st[0] = s->time; st[1] = s->dt; st[2] = (double)s->steps; st[3] = s->lastErr; st[4] = s->maxTorque;
    st[5] = (double)s->trials_total; st[6] = (double)s->rejects_total; st[7] = s->energy; st[8] = s->units[1];
}

void or_sim_heff(or_sim* s, const double* m, double* h) {
    size_t n3 = 3 * (size_t)s->N;
    for (size_t i = 0; i < n3; i++) s->tmp[i] = (or_real)m[i];
    calc_heff(s, s->tmp, s->m1, 0);
    for (size_t i = 0; i < n3; i++) h[i] = s->m1[i];
}

void or_sim_torque(or_sim* s, const double* m, double* k) {
    size_t n3 = 3 * (size_t)s->N;
    for (size_t i = 0; i < n3; i++) s->tmp[i] = (or_real)m[i];
    calc_heff(s, s->tmp, s->m1, 0);
    torque(s, s->tmp, s->m1);
    for (size_t i = 0; i < n3; i++) k[i] = s->m1[i];
}

/* ApplyStatic / ApplyPointwise on a running simulation (simfield.go:87-110):
 * replaces the field table; the pointwise search restarts at its first point */
void or_sim_set_field(or_sim* s, int n, const double* t, const double* b) {
    free(s->fieldT);
    free(s->fieldB);
    s->fieldT = s->fieldB = NULL;
    if (n > 0) {
        s->fieldT = (double*)xcalloc((size_t)n, sizeof(double));
        s->fieldB = (double*)xcalloc((size_t)n * 3, sizeof(double));
        if (t) memcpy(s->fieldT, t, (size_t)n * sizeof(double));
        memcpy(s->fieldB, b, (size_t)n * 3 * sizeof(double));
    }
    s->p.nField = n;
    s->p.fieldT = s->fieldT;
    s->p.fieldB = s->fieldB;
    s->fieldLastIdx = 0;
}

/* the first cap entries of the per-trial log (dt used, accepted); returns its length */
long or_sim_trial_log(const or_sim* s, double* dt, int* acc, long cap) {
    for (long i = 0; i < s->log_n && i < cap; i++) {
        dt[i] = s->log_dt[i];
        acc[i] = s->log_acc[i];
    }
    return s->log_n;
}

/* Sim.wantEnergy (relax.go:16): energy of each step, steps may not raise it */
void or_sim_set_want_energy(or_sim* s, int on) { s->p.wantEnergy = on; }

void or_sim_set_field_function(or_sim* s, int kind, const double* b, const double* params) {
    s->fn = kind;
    for (int i = 0; i < 4; i++) s->fp[i] = params[i];
    for (int c = 0; c < 3; c++) s->fb[c] = b[c];
}

void or_sim_field_at(or_sim* s, double t_si, double* out) { applied_field(s, t_si, out); }

void or_sim_set_alpha_map(or_sim* s, const float* amap) {
    free(s->amap);
    s->amap = NULL;
    if (!amap) return;
    s->amap = (or_real*)malloc((size_t)s->N * sizeof(or_real));
    if (!s->amap) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    for (int i = 0; i < s->N; i++) s->amap[i] = (or_real)amap[i];
}

void or_sim_set_spintorque(or_sim* s, const double j[3], double spinPol, double xi, const float* jmask) {
    for (int i = 0; i < 3; i++) s->j[i] = j[i];
    s->spinPol = spinPol;
    s->xi = xi;
    free(s->jmask);
    s->jmask = NULL;
    if (!jmask) return;
    s->jmask = (or_real*)malloc(3 * (size_t)s->N * sizeof(or_real));
    if (!s->jmask) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    for (int i = 0; i < 3 * s->N; i++) s->jmask[i] = (or_real)jmask[i];
}

void or_sim_set_applied_current(or_sim* s, int n, const double* t, const double* j, int kind, const double* params) {
    free(s->jT);
    free(s->jV);
    s->jT = s->jV = NULL;
    s->jn = kind ? 0 : n;
    s->jfn = kind;
    s->jLastIdx = 0;
    s->japplied = kind || n > 0;
    if (kind) {
        for (int i = 0; i < 4; i++) s->jfp[i] = params[i];
        for (int c = 0; c < 3; c++) s->jfb[c] = j[c];
        return;
    }
    if (n > 0) {
        s->jT = (double*)xcalloc((size_t)n, sizeof(double));
        s->jV = (double*)xcalloc(3 * (size_t)n, sizeof(double));
        memcpy(s->jT, t, (size_t)n * sizeof(double));
        memcpy(s->jV, j, 3 * (size_t)n * sizeof(double));
    }
}

void or_sim_set_norm_map(or_sim* s, const float* nmap) {
    free(s->nmap);
    s->nmap = NULL;
    if (!nmap) return;
    s->nmap = (or_real*)malloc((size_t)s->N * sizeof(or_real));
    if (!s->nmap) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    for (int i = 0; i < s->N; i++) s->nmap[i] = (or_real)nmap[i];
}

double or_sim_energy(or_sim* s) {
    calc_heff(s, s->m, s->m1, 1);
    return s->energy;
}
