// Demag(+exchange) convolution for the device-resident engine.
//
// Reference behaviour restated: conv.go:57-100 (three padded forward FFTs,
// k-space multiply with the symmetric kernel, three inverse FFTs and unpad),
// conv.go:110-167 (kernel spectrum = Re FFT(K) / (Px Py Pz)), demag.go:26-84,142-183
// (FaceKernel6), exchange.go:44-69 + init.go:256-267 (exchange folded in).
//
// MI355X design.  Intermediate spectra are kept in transposed layouts so that
// every FFT pass reads and writes whole contiguous lines (coalesced 64-lane
// accesses) and no separate transpose or pad/unpad pass exists:
//   m[c][x][y][z]  --Z r2c (stage build fused)-->  B1[c][x][kz][y]   (y fastest)
//   3-D: B1 --Y fwd--> B2[c][kz][x][ky] --X fwd . K . X inv (in place)--> B2
//        --Y inv--> B1 --Z c2r--> h[c][x][y][z]
//   2-D: B1 --Y fwd . K . Y inv (in place)--> B1 --Z c2r--> h
// Only data lines that can be non-zero are transformed (rows y >= Ny and
// planes x >= Nx are zero padding), and only the kept half of each inverse
// transform is stored, so B1/B2 hold Nx (not Px) planes.
//
// The kernel spectrum is real and each component is even or odd along every
// axis (Kxy odd in x and y, Kxz odd in x, Kyz odd in y, diagonal even), so
// only kx <= Px/2, ky <= Py/2 is stored (Kc), a quarter of the reference's
// 6 x Px x Py x (Pz/2+1) floats; signs are restored on the fly.  3-D: Kc is
// [kz][ky][kx][6] (the 6 components of one frequency are 24 contiguous bytes,
// read as 3 float2 by the lane that needs them); 2-D: [4][kz][ky].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "engine.h"
#include "fft_device.h"
#include "gk_fft.h"
#include "passes.h"

namespace gk {

// =============================================================== kernel setup
__constant__ int c_kern_idx[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};  // xx yy zz yz xz xy

// faceIntegral (demag.go:142-183) in FP64
__device__ void face_integral_f64(double B[3], const double R[3], const double cell[3], int s, int n) {
    const int u = s, v = (s + 1) % 3, w = (s + 2) % 3;
    const double charge = cell[v] * cell[w];
    const double pu = cell[u] * 0.5;
    const double inv4pi = 1.0 / (4.0 * 3.14159265358979323846);
    B[0] = B[1] = B[2] = 0.0;
    for (int i = 0; i < n; i++) {
        const double pv = -0.5 * cell[v] + cell[v] / (2.0 * n) + i * (cell[v] / n);
        for (int j = 0; j < n; j++) {
            const double pw = -0.5 * cell[w] + cell[w] / (2.0 * n) + j * (cell[w] / n);
            for (int pole = 0; pole < 2; pole++) {
                double pos[3];
                pos[u] = pole == 0 ? pu : -pu;
                pos[v] = pv;
                pos[w] = pw;
                const double d0 = R[0] - pos[0], d1 = R[1] - pos[1], d2 = R[2] - pos[2];
                const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                const double r = sqrt(r2);
                const double q = (pole == 0 ? charge : -charge) * inv4pi / (r2 * r);
                B[0] += d0 * q;
                B[1] += d1 * q;
                B[2] += d2 * q;
            }
        }
    }
    const double inv = 1.0 / (double)(n * n);
    B[0] *= inv;
    B[1] *= inv;
    B[2] *= inv;
}

// wrapped index -> displacement within FaceKernel6's loop range [-(P-1)/2, P/2-1]
// (demag.go:35-53); false for the index no displacement maps to.
__device__ __forceinline__ bool unwrap(int w, int P, int& x) {
    int x1 = -((P - 1) / 2), x2 = P / 2 - 1;
    if (P == 1) x1 = x2 = 0;
    if (w <= x2) {
        x = w;
        return true;
    }
    x = w - P;
    return x >= x1;
}

__device__ __forceinline__ int wrap_dev(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// The displacements (in cells) along one axis whose face integrals add into
// wrapped index w of padded length P: first, first + step, ... (count of them).
// Without periodic images (period == 0) the one displacement unwrap(w); a
// length of 1 keeps displacement 0 along X only, as FaceKernel6's loop range
// does (demag.go:35-46: along Y and Z the range of a length 1 is empty).
// With `images` periodic images of an axis of `period` cells, the reference
// builds the kernel over the period itself (padSize, kernel.go:79-90) and adds
// every displacement of [-((period-1)/2)(images+1), (period/2-1)(images+1)]
// into its wrapped index (demag.go:48-79); on the padded length P >= 2 period
// the linear convolution reproduces that circular one when index w holds the
// sum for the residue of unwrap(w) modulo the period.
struct AxisImages {
    int first, step, count;
};
__device__ __forceinline__ AxisImages axis_images(int w, int P, int period, int images, bool axisX) {
    AxisImages a{0, 1, 0};
    if (period == 0) {
        if (P == 1) {
            a.count = axisX ? 1 : 0;
            return a;
        }
        a.count = unwrap(w, P, a.first) ? 1 : 0;
        return a;
    }
    // the residue: w itself on the reference's own grid (P == period), else that
    // of the displacement of w (the index P/2 of an even P >= 2 period is never read)
    int rsd = w;
    if (P != period) {
        if (!unwrap(w, P, rsd)) return a;
        rsd = wrap_dev(rsd, period);
    }
    const int lo = -((period - 1) / 2) * (images + 1), hi = (period / 2 - 1) * (images + 1);
    a.step = period;
    a.first = lo + wrap_dev(rsd - lo, period);
    a.count = a.first > hi ? 0 : (hi - a.first) / period + 1;
    return a;
}

struct KernelImages {
    int period[3], images[3];
};

__global__ void k_facekernel6(double* __restrict__ K, int P0, int P1, int P2, double c0, double c1, double c2,
                              int acc, KernelImages im) {
    const long NP = (long)P0 * P1 * P2;
    const double cell[3] = {c0, c1, c2};
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < NP; idx += (long)gridDim.x * blockDim.x) {
        const int zw = (int)(idx % P2);
        const long r = idx / P2;
        const int yw = (int)(r % P1), xw = (int)(r / P1);
        const AxisImages ax = axis_images(xw, P0, im.period[0], im.images[0], true);
        const AxisImages ay = axis_images(yw, P1, im.period[1], im.images[1], false);
        const AxisImages az = axis_images(zw, P2, im.period[2], im.images[2], false);
        double kk[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < ax.count; i++)
            for (int j = 0; j < ay.count; j++)
                for (int l = 0; l < az.count; l++) {
                    const double R[3] = {(ax.first + i * ax.step) * c0, (ay.first + j * ay.step) * c1,
                                         (az.first + l * az.step) * c2};
                    for (int s = 0; s < 3; s++) {
                        double B[3];
                        face_integral_f64(B, R, cell, s, acc);
                        for (int d = s; d < 3; d++) kk[c_kern_idx[s][d]] += B[d];
                    }
                }
        for (int k = 0; k < 6; k++) K[k * NP + idx] = kk[k];
    }
}

// ---------------------------------------------------------------- Newell tensor
// The demagnetising tensor N between two equal rectangular cells, averaged over
// the target cell (Newell, Williams & Dunlop, J. Geophys. Res. 98, 9551 (1993)):
// N_xx = -(1/(4 pi V)) d2x d2y d2z f(X, Y, Z) and N_xy the same of g, where d2
// is the second difference [1, -2, 1] with the cell size as step.  The kernel
// is K = -N (field per unit magnetisation, the sign of FaceKernel6).
// Cancellation in the 27-term sums grows like (|R|/d)^6; from |R| >= 12 d_max
// on the same cell average is taken by 3-point Gauss-Legendre quadrature in
// each coordinate of both cells (error ~1e-8 relative there).
__device__ double newell_f(double x, double y, double z) {
    x = fabs(x), y = fabs(y), z = fabs(z);
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double R = sqrt(x2 + y2 + z2);
    if (R == 0.0) return 0.0;
    double t = (2.0 * x2 - y2 - z2) * R / 6.0;
    const double rxz = sqrt(x2 + z2), rxy = sqrt(x2 + y2);
    if (y > 0.0 && rxz > 0.0) t += 0.5 * y * (z2 - x2) * asinh(y / rxz);
    if (z > 0.0 && rxy > 0.0) t += 0.5 * z * (y2 - x2) * asinh(z / rxy);
    if (x > 0.0 && y > 0.0 && z > 0.0) t -= x * y * z * atan(y * z / (x * R));
    return t;
}

__device__ double newell_g(double x, double y, double z) {
    const double sgn = ((x < 0.0) != (y < 0.0)) ? -1.0 : 1.0;
    x = fabs(x), y = fabs(y), z = fabs(z);
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double R = sqrt(x2 + y2 + z2);
    if (R == 0.0) return 0.0;
    const double rxy = sqrt(x2 + y2), ryz = sqrt(y2 + z2), rxz = sqrt(x2 + z2);
    double t = -x * y * R / 3.0;
    if (z > 0.0 && rxy > 0.0) t += x * y * z * asinh(z / rxy);
    if (x > 0.0 && ryz > 0.0) t += y / 6.0 * (3.0 * z2 - y2) * asinh(x / ryz);
    if (y > 0.0 && rxz > 0.0) t += x / 6.0 * (3.0 * z2 - x2) * asinh(y / rxz);
    if (z > 0.0) t -= z2 * z / 6.0 * atan(x * y / (z * R));
    if (y > 0.0) t -= z * y2 / 2.0 * atan(x * z / (y * R));
    if (x > 0.0) t -= z * x2 / 2.0 * atan(y * z / (x * R));
    return sgn * t;
}

// sum over the 27 neighbours with weights (-1, 2, -1)^3 of fn(X+i dx, Y+j dy, Z+k dz)
template <bool G>
__device__ double newell_sum(double X, double Y, double Z, double dx, double dy, double dz) {
    const double w[3] = {-1.0, 2.0, -1.0};
    double s = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double sk = 0.0;
            for (int k = 0; k < 3; k++) {
                const double a = X + (i - 1) * dx, b = Y + (j - 1) * dy, c = Z + (k - 1) * dz;
                sk += w[k] * (G ? newell_g(a, b, c) : newell_f(a, b, c));
            }
            s += w[i] * w[j] * sk;
        }
    return s / (4.0 * 3.14159265358979323846 * dx * dy * dz);
}

// K = -N at displacement R (order xx yy zz yz xz xy)
__device__ void newell_kernel(double K[6], const double R[3], const double d[3]) {
    const double r = sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
    const double dmax = fmax(d[0], fmax(d[1], d[2]));
    if (r < 12.0 * dmax) {
        K[0] = -newell_sum<false>(R[0], R[1], R[2], d[0], d[1], d[2]);
        K[1] = -newell_sum<false>(R[1], R[2], R[0], d[1], d[2], d[0]);
        K[2] = -newell_sum<false>(R[2], R[0], R[1], d[2], d[0], d[1]);
        K[3] = -newell_sum<true>(R[1], R[2], R[0], d[1], d[2], d[0]);
        K[4] = -newell_sum<true>(R[0], R[2], R[1], d[0], d[2], d[1]);
        K[5] = -newell_sum<true>(R[0], R[1], R[2], d[0], d[1], d[2]);
        return;
    }
    // V * E[D(R + a - b)], a and b uniform in the cell, D the point-dipole field
    const double gx[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
    const double gw[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
    double u[3][9], wu[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            for (int a = 0; a < 3; a++) u[a][3 * i + j] = 0.5 * (gx[i] - gx[j]) * d[a];
            wu[3 * i + j] = gw[i] * gw[j];
        }
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 9; i++)
        for (int j = 0; j < 9; j++) {
            const double x = R[0] + u[0][i], y = R[1] + u[1][j];
            const double wij = wu[i] * wu[j];
            for (int k = 0; k < 9; k++) {
                const double z = R[2] + u[2][k];
                const double r2 = x * x + y * y + z * z;
                const double ir2 = 1.0 / r2;
                const double ir3 = ir2 / sqrt(r2);
                const double w = wij * wu[k] * ir3;
                const double w3 = 3.0 * w * ir2;
                acc[0] += w3 * x * x - w;
                acc[1] += w3 * y * y - w;
                acc[2] += w3 * z * z - w;
                acc[3] += w3 * y * z;
                acc[4] += w3 * x * z;
                acc[5] += w3 * x * y;
            }
        }
    const double s = d[0] * d[1] * d[2] / (4.0 * 3.14159265358979323846);
    for (int c = 0; c < 6; c++) K[c] = s * acc[c];
}

// one octant (displacements >= 0) of the Newell kernel: O[6][H0][H1][H2]
__global__ void k_newell_octant(double* __restrict__ O, int H0, int H1, int H2, double c0, double c1, double c2) {
    const long NH = (long)H0 * H1 * H2;
    const double d[3] = {c0, c1, c2};
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < NH; idx += (long)gridDim.x * blockDim.x) {
        const int z = (int)(idx % H2);
        const long r = idx / H2;
        const int y = (int)(r % H1), x = (int)(r / H1);
        const double R[3] = {x * c0, y * c1, z * c2};
        double K[6];
        newell_kernel(K, R, d);
        for (int k = 0; k < 6; k++) O[k * NH + idx] = K[k];
    }
}

// the wrapped grid from the octant: diagonal even, yz odd in y and z, xz odd
// in x and z, xy odd in x and y; FaceKernel6's displacement range (unwrap)
__global__ void k_newell_expand(const double* __restrict__ O, int H0, int H1, int H2, double* __restrict__ K,
                                int P0, int P1, int P2) {
    const long NP = (long)P0 * P1 * P2, NH = (long)H0 * H1 * H2;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < NP; idx += (long)gridDim.x * blockDim.x) {
        const int zw = (int)(idx % P2);
        const long r = idx / P2;
        const int yw = (int)(r % P1), xw = (int)(r / P1);
        int x, y, z;
        double kk[6] = {0, 0, 0, 0, 0, 0};
        if (unwrap(xw, P0, x) && unwrap(yw, P1, y) && unwrap(zw, P2, z)) {
            const long o = ((long)abs(x) * H1 + abs(y)) * H2 + abs(z);
            const double sx = x < 0 ? -1.0 : 1.0, sy = y < 0 ? -1.0 : 1.0, sz = z < 0 ? -1.0 : 1.0;
            kk[0] = O[0 * NH + o];
            kk[1] = O[1 * NH + o];
            kk[2] = O[2 * NH + o];
            kk[3] = sy * sz * O[3 * NH + o];
            kk[4] = sx * sz * O[4 * NH + o];
            kk[5] = sx * sy * O[5 * NH + o];
        }
        for (int k = 0; k < 6; k++) K[k * NP + idx] = kk[k];
    }
}

// Exch6NgbrKernel's points as indices of the reference's kernel grid R (the
// padded length, or the period along a periodic axis), added to the diagonal
// components at every index w of the padded grid P whose residue maps to them
// (the identity along axes without periodic images).
struct ExchPoints {
    long idx[7];
    double val[7];
    int n;
    int R[3], period[3];
};
__global__ void k_add_exch_points(double* __restrict__ K, int P0, int P1, int P2, ExchPoints p) {
    const long NP = (long)P0 * P1 * P2;
    for (long w = blockIdx.x * (long)blockDim.x + threadIdx.x; w < NP; w += (long)gridDim.x * blockDim.x) {
        const int wz = (int)(w % P2);
        const long r = w / P2;
        const int wy = (int)(r % P1), wx = (int)(r / P1);
        const int Pw[3] = {P0, P1, P2}, iw[3] = {wx, wy, wz};
        int rho[3];
        bool ok = true;
        for (int a = 0; a < 3; a++) {
            rho[a] = iw[a];
            if (p.period[a] && Pw[a] != p.period[a]) {
                int d = 0;
                ok = ok && unwrap(iw[a], Pw[a], d);
                rho[a] = wrap_dev(d, p.period[a]);
            }
        }
        if (!ok) continue;
        const long id = ((long)rho[0] * p.R[1] + rho[1]) * p.R[2] + rho[2];
        for (int i = 0; i < p.n; i++)
            if (p.idx[i] == id)
                for (int comp = 0; comp < 3; comp++) K[comp * NP + w] += p.val[i];  // xx, yy, zz
    }
}

__global__ void k_real_to_complex(const double* __restrict__ in, double2* __restrict__ out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = make_double2(in[i], 0.0);
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* a, double v) {
    atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// spectrum -> compressed kernel (and optionally the reference's full layout)
// Compressed layouts: 2-D [comp][kz][ky]; 3-D [kz][ky][kx][comp] (the 6 values
// of one frequency together, for the X . K . X pass); thin 3-D plane-pass
// plans [kz][comp][kx][ky] (consecutive ky together, one thread per ky).
__global__ void k_extract_kernel(const double2* __restrict__ W, int P0, int P1, int P2, int Q, double norm,
                                 float* __restrict__ Kc, int comp, int nK, int plane_layout,
                                 float* __restrict__ full_comp, unsigned long long* maxre,
                                 unsigned long long* maxim) {
    const long total = (long)P0 * P1 * Q;
    const int hx = P0 / 2 + 1, hy = P1 / 2 + 1;
    double lre = 0, lim = 0;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int kz = (int)(t % Q);
        const long r = t / Q;
        const int ky = (int)(r % P1), kx = (int)(r / P1);
        const double2 w = W[((long)kx * P1 + ky) * P2 + kz];
        const double re = w.x * norm, im = w.y * norm;
        lre = fmax(lre, fabs(re));
        lim = fmax(lim, fabs(im));
        if (full_comp) full_comp[t] = (float)re;
        if (kx < hx && ky < hy) {
            if (P0 == 1) Kc[((long)comp * Q + kz) * hy + ky] = (float)re;
            else if (plane_layout) Kc[(((long)kz * nK + comp) * hx + kx) * hy + ky] = (float)re;
            else Kc[(((long)kz * hy + ky) * hx + kx) * nK + comp] = (float)re;
        }
    }
    atomic_max_nonneg(maxre, lre);
    atomic_max_nonneg(maxim, lim);
}

// =============================================================== plan
static int pad_len(int n) {
    // smallest register-FFT length 2^a or 3 * 2^a that is >= 2n (any P >= 2n - 1
    // gives the same linear convolution; the reference pads to exactly 2n)
    if (n == 1) return 1;
    int p = 2;
    while (p < 2 * n) p *= 2;
    return (p / 4) * 3 >= 2 * n ? (p / 4) * 3 : p;
}

void conv_plan_init(ConvPlan& cp, const int N[3]) {
    GK_ASSERT(N[2] >= 2, "the Z (last user x) size must be > 1, got %d (init.go:127-129)", N[2]);
    for (int i = 0; i < 3; i++) {
        cp.N[i] = N[i];
        cp.P[i] = pad_len(N[i]);
    }
    if (cp.P[2] < 2) cp.P[2] = 2;
    cp.M = cp.P[2] / 2;
    cp.Q = cp.M + 1;
    cp.nK = cp.P[0] == 1 ? 4 : 6;
    cp.yxy = cp.P[0] > 1 && yxy_fits(cp.P[0], cp.P[1]);
    cp.plane3 = cp.P[0] > 1 && !cp.yxy && plane3_fits(cp.P[0], cp.P[1], N[0]);
    GK_ASSERT(cp.M <= 2048 && cp.P[1] <= 4096 && cp.P[0] <= 1024, "grid too large for this build (padded %d %d %d)",
              cp.P[0], cp.P[1], cp.P[2]);
    const long hx = cp.P[0] / 2 + 1, hy = cp.P[1] / 2 + 1;
    const long kc = cp.P[0] == 1 ? (long)cp.nK * cp.Q * hy : (long)cp.nK * cp.Q * hy * hx;
    const long b1 = 3L * N[0] * cp.Q * N[1];
    const long b2 = (cp.P[0] == 1 || cp.yxy) ? 0 : 3L * cp.Q * cp.P[1] * N[0];
    cp.Kc = nullptr;  // conv_acquire_kernel
    cp.kc_elems = kc;
    cp.B1 = (float2*)cached_malloc(b1 * sizeof(float2));
    cp.B2 = nullptr;  // layout [3][Q][Nx][Py]
    if (b2) cp.B2 = (float2*)cached_malloc(b2 * sizeof(float2));
    cp.bytes = kc * sizeof(float) + (b1 + b2) * sizeof(float2);
    cp.tp = TuneParams{};
    conv_tune_lookup(N, &cp.tp);
}

int conv_epi_blocks_max(const int N[3]) {
    int P2 = pad_len(N[2]);
    if (P2 < 2) P2 = 2;
    const int M = P2 / 2;
    return div_up(N[1], std::min(conv_epi_rows(M, 0), conv_epi_rows(M, 1))) * N[0];
}

// ---------------------------------------------------------------- tuning table
namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 3>, TuneParams> g_tune;
}  // namespace

bool conv_tune_lookup(const int N[3], TuneParams* tp) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find({N[0], N[1], N[2]});
    if (it == g_tune.end()) return false;
    *tp = it->second;
    return true;
}

void conv_tune_store(const int N[3], const TuneParams& tp) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[{N[0], N[1], N[2]}] = tp;
}

void conv_tune_clear() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
}

int conv_tune_save(const char* path) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "# launch parameters per data size (internal Nx Ny Nz): zr er yb xb y2b pp xv\n");
    for (auto& kv : g_tune) {
        const TuneParams& t = kv.second;
        fprintf(f, "%d %d %d %d %d %d %d %d %d %d\n", kv.first[0], kv.first[1], kv.first[2], t.zr, t.er, t.yb, t.xb,
                t.y2b, t.pp, t.xv);
    }
    const int n = (int)g_tune.size();
    return fclose(f) == 0 ? n : -1;
}

int conv_tune_load(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    char line[256];
    int n = 0;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        int N[3];
        TuneParams t;
        const int got = sscanf(line, "%d %d %d %d %d %d %d %d %d %d", &N[0], &N[1], &N[2], &t.zr, &t.er, &t.yb,
                               &t.xb, &t.y2b, &t.pp, &t.xv);
        if (got < 8) continue;
        if (got < 9) t.pp = 0;  // older tables: 8 or 9 columns
        if (got < 10) t.xv = 0;
        if (N[0] < 1 || N[1] < 1 || N[2] < 2 || t.yb < 1 || t.xb < 1 || t.y2b < 1 || t.pp < 0 || t.pp > 1) continue;
        if (!conv_xv_valid(pad_len(N[0]), t.xv)) t.xv = 0;
        int P2 = pad_len(N[2]);
        if (P2 < 2) P2 = 2;
        if (!conv_zr_valid(P2 / 2, t.zr)) t.zr = 0;
        if (!conv_er_valid(P2 / 2, t.er)) t.er = 0;
        if (t.pp && !plane3_fits(pad_len(N[0]), pad_len(N[1]), N[0])) t.pp = 0;
        conv_tune_store(N, t);
        n++;
    }
    fclose(f);
    return n;
}

size_t conv_plan_footprint(const int N[3], int* epi_blocks) {
    int P[3];
    for (int i = 0; i < 3; i++) P[i] = pad_len(N[i]);
    if (P[2] < 2) P[2] = 2;
    const int M = P[2] / 2, Q = M + 1, nK = P[0] == 1 ? 4 : 6;
    const long hx = P[0] / 2 + 1, hy = P[1] / 2 + 1;
    const long kc = P[0] == 1 ? (long)nK * Q * hy : (long)nK * Q * hy * hx;
    const long b1 = 3L * N[0] * Q * N[1];
    const long b2 = (P[0] == 1 || yxy_fits(P[0], P[1])) ? 0 : 3L * Q * P[1] * N[0];
    if (epi_blocks) *epi_blocks = conv_epi_blocks_max(N);
    return kc * sizeof(float) + (b1 + b2) * sizeof(float2);
}

// ---------------------------------------------------------------- shared kernel spectra
// One spectrum per (device, padded size, cell, kernel, accuracy, exchange):
// simulations of one shape -- a parameter sweep, a served batch -- build it
// once and share it (the reference recomputes or reads a disk cache per
// simulation, kernelpipe.go:20-65, wisdom.go:30-133).
namespace {
struct KernelKey {
    int dev, P0, P1, P2, type, accuracy, exch;
    std::array<int, 6> images;  // period and image count per axis
    uint64_t c0, c1, c2;
    auto tie() const { return std::tie(dev, P0, P1, P2, type, accuracy, exch, images, c0, c1, c2); }
    bool operator<(const KernelKey& o) const { return tie() < o.tie(); }
};
struct KernelEntry {
    float* Kc;
    int refs;
    double imag;
};
std::mutex g_kmu;
std::map<KernelKey, KernelEntry> g_kernels;

uint64_t bits(double v) {
    uint64_t u;
    memcpy(&u, &v, 8);
    return u;
}
}  // namespace

double conv_acquire_kernel(ConvPlan& cp, const KernelSpec& ks) {
    KernelKey key{};
    GK_CHECK(hipGetDevice(&key.dev));
    key.P0 = cp.P[0], key.P1 = cp.P[1], key.P2 = cp.P[2];
    key.type = ks.type;
    key.accuracy = ks.type == MM_KERNEL_FACE6 ? ks.accuracy : 0;
    key.exch = ks.exch ? 1 : 0;
    for (int a = 0; a < 3; a++) {
        key.images[a] = ks.period[a];
        key.images[3 + a] = ks.images[a];
    }
    key.c0 = bits(ks.cell[0]), key.c1 = bits(ks.cell[1]), key.c2 = bits(ks.cell[2]);
    std::lock_guard<std::mutex> lock(g_kmu);
    auto it = g_kernels.find(key);
    if (it != g_kernels.end()) {
        it->second.refs++;
        cp.Kc = it->second.Kc;
        return it->second.imag;
    }
    cp.Kc = (float*)cached_malloc(cp.kc_elems * sizeof(float));
    const double imag = conv_build_kernel(cp, ks, nullptr, nullptr);
    g_kernels[key] = KernelEntry{cp.Kc, 1, imag};
    return imag;
}

static void release_kernel(float* Kc) {
    std::lock_guard<std::mutex> lock(g_kmu);
    for (auto it = g_kernels.begin(); it != g_kernels.end(); ++it)
        if (it->second.Kc == Kc) {
            if (--it->second.refs == 0) {
                cached_free(Kc);
                g_kernels.erase(it);
            }
            return;
        }
    cached_free(Kc);
}

int conv_kernel_cache_size() {
    int dev = 0;
    GK_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_kmu);
    int n = 0;
    for (auto& kv : g_kernels) n += kv.first.dev == dev;
    return n;
}

void conv_plan_free(ConvPlan& cp) {
    if (cp.Kc) release_kernel(cp.Kc);
    cached_free(cp.B1);
    cached_free(cp.B2);
    cp.Kc = nullptr;
    cp.B1 = cp.B2 = nullptr;
}

static int wrapi(int number, int max) {
    while (number < 0) number += max;
    while (number >= max) number -= max;
    return number;
}

// real-space demag kernel (FaceKernel6 or Newell) + Exch6 over the padded grid
// P, 6 x P doubles on the device
static double* realspace_kernel_f64(const int P[3], const KernelSpec& ks, hipStream_t s) {
    const int P0 = P[0], P1 = P[1], P2 = P[2];
    const double* cell = ks.cell;
    const long NP = (long)P0 * P1 * P2;
    double* K = nullptr;
    K = (double*)cached_malloc(6 * NP * sizeof(double));
    for (int a = 0; a < 3; a++)
        GK_ASSERT(ks.period[a] == 0 || (ks.images[a] > 0 && (ks.period[a] == P[a] || P[a] >= 2 * ks.period[a])),
                  "periodic axis %d: period %d, %d images, padded length %d", a, ks.period[a], ks.images[a], P[a]);
    if (ks.type == MM_KERNEL_NEWELL) {
        GK_ASSERT(!ks.period[0] && !ks.period[1] && !ks.period[2],
                  "periodic boundaries use the FaceKernel6 demag kernel (demag.go:26-84)");
        // displacements 0 .. P/2 per axis cover unwrap's range [-(P-1)/2, P/2-1] in magnitude
        const int H0 = P0 / 2 + 1, H1 = P1 / 2 + 1, H2 = P2 / 2 + 1;
        const long NH = (long)H0 * H1 * H2;
        double* O = nullptr;
        O = (double*)cached_malloc(6 * NH * sizeof(double));
        k_newell_octant<<<grid1d(NH, 64), 64, 0, s>>>(O, H0, H1, H2, cell[0], cell[1], cell[2]);
        GK_LAUNCH_CHECK();
        k_newell_expand<<<grid1d(NP), kBlock, 0, s>>>(O, H0, H1, H2, K, P0, P1, P2);
        GK_LAUNCH_CHECK();
        GK_CHECK(hipStreamSynchronize(s));
        cached_free(O);
    } else {
        GK_ASSERT(ks.type == MM_KERNEL_FACE6, "unknown demag kernel type %d", ks.type);
        KernelImages im{};
        for (int a = 0; a < 3; a++) {
            im.period[a] = ks.period[a];
            im.images[a] = ks.images[a];
        }
        k_facekernel6<<<grid1d(NP), kBlock, 0, s>>>(K, P0, P1, P2, cell[0], cell[1], cell[2], ks.accuracy, im);
        GK_LAUNCH_CHECK();
    }
    if (ks.exch) {
        // Exch6NgbrKernel written in the reference's order on the reference's grid R:
        // with a length of 1, wrap(+-1,1) = 0 overwrites the centre; with a period
        // of 2, +1 and -1 are one index (exchange.go:58-65)
        int R[3];
        for (int a = 0; a < 3; a++) R[a] = ks.period[a] ? ks.period[a] : P[a];
        std::vector<std::pair<long, double>> pts;
        auto put = [&](int a, int b, int c, double v) {
            long id = ((long)wrapi(a, R[0]) * R[1] + wrapi(b, R[1])) * R[2] + wrapi(c, R[2]);
            for (auto& pr : pts)
                if (pr.first == id) {
                    pr.second = v;
                    return;
                }
            pts.push_back({id, v});
        };
        const double hx = cell[0] * cell[0], hy = cell[1] * cell[1], hz = cell[2] * cell[2];
        put(0, 0, 0, -2 / hx - 2 / hy - 2 / hz);
        put(1, 0, 0, 1 / hx);
        put(-1, 0, 0, 1 / hx);
        put(0, 1, 0, 1 / hy);
        put(0, -1, 0, 1 / hy);
        put(0, 0, 1, 1 / hz);
        put(0, 0, -1, 1 / hz);
        ExchPoints ep{};
        ep.n = (int)pts.size();
        for (int i = 0; i < ep.n; i++) {
            ep.idx[i] = pts[i].first;
            ep.val[i] = pts[i].second;
        }
        for (int a = 0; a < 3; a++) {
            ep.R[a] = R[a];
            ep.period[a] = ks.period[a];
        }
        k_add_exch_points<<<grid1d(NP), kBlock, 0, s>>>(K, P0, P1, P2, ep);
        GK_LAUNCH_CHECK();
    }
    return K;
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

void conv_realspace_kernel_f64(const int P[3], const KernelSpec& ks, double* out_host) {
    hipStream_t s = stream();
    const long NP = (long)P[0] * P[1] * P[2];
    double* K = realspace_kernel_f64(P, ks, s);
    GK_CHECK(hipMemcpyAsync(out_host, K, 6 * NP * sizeof(double), hipMemcpyDeviceToHost, s));
    sync();
    cached_free(K);
}

void conv_realspace_kernel(const int P[3], const KernelSpec& ks, float* out_host) {
    hipStream_t s = stream();
    const long NP = (long)P[0] * P[1] * P[2];
    double* K = realspace_kernel_f64(P, ks, s);
    float* f = nullptr;
    f = (float*)cached_malloc(6 * NP * sizeof(float));
    k_to_float<<<grid1d(6 * NP), kBlock, 0, s>>>(K, f, 6 * NP);
    GK_LAUNCH_CHECK();
    GK_CHECK(hipMemcpyAsync(out_host, f, 6 * NP * sizeof(float), hipMemcpyDeviceToHost, s));
    sync();
    cached_free(K);
    cached_free(f);
}

double conv_build_kernel(ConvPlan& cp, const KernelSpec& ks, float* full_host, float* full_dev) {
    hipStream_t s = stream();
    const int P0 = cp.P[0], P1 = cp.P[1], P2 = cp.P[2];
    const long NP = (long)P0 * P1 * P2;
    double* K = realspace_kernel_f64(cp.P, ks, s);
    double2* W = nullptr;
    unsigned long long* mx = nullptr;
    W = (double2*)cached_malloc(NP * sizeof(double2));
    mx = (unsigned long long*)cached_malloc(2 * sizeof(unsigned long long));
    GK_CHECK(hipMemsetAsync(mx, 0, 2 * sizeof(unsigned long long), s));
    const int hx = P0 / 2 + 1, hy = P1 / 2 + 1;
    const long fulls = (long)P0 * P1 * cp.Q;
    const double norm = 1.0 / (double)NP;
    // components stored: 3-D xx yy zz yz xz xy; 2-D xx yy zz yz
    for (int k = 0; k < cp.nK; k++) {
        k_real_to_complex<<<grid1d(NP), kBlock, 0, s>>>(K + k * NP, W, NP);
        fft3_c2c_f64(W, cp.P, -1, s);
        k_extract_kernel<<<grid1d(fulls), kBlock, 0, s>>>(W, P0, P1, P2, cp.Q, norm, cp.Kc, k, cp.nK, cp.yxy,
                                                            full_dev ? full_dev + k * fulls : nullptr, mx, mx + 1);
        GK_LAUNCH_CHECK();
    }
    unsigned long long h[2];
    GK_CHECK(hipMemcpyAsync(h, mx, sizeof(h), hipMemcpyDeviceToHost, s));
    if (full_host && full_dev)
        GK_CHECK(hipMemcpyAsync(full_host, full_dev, cp.nK * fulls * sizeof(float), hipMemcpyDeviceToHost, s));
    sync();
    cached_free(K);
    cached_free(W);
    cached_free(mx);
    double re, im;
    memcpy(&re, &h[0], 8);
    memcpy(&im, &h[1], 8);
    return re > 0 ? im / re : 0.0;
}

// =============================================================== launchers


// rows per Z block: enough consecutive y for 64-128 B transposed segments, at most 1024 threads
template <int M> constexpr int z_rows() {
    constexpr int G = RegFFT<M>::G;
    int tr = 1024 / G;
    if (tr > 16) tr = 16;
    while (tr > 1 && tr * (M + 1) * 8 > 72 * 1024) tr /= 2;
    return tr;
}

// rows per Z forward block of shape variant v: 0 the default, 1 half of it, 2
// double (while the block stays within 1024 threads and 80 KB of LDS).  Short
// rows (M < 32) have one shape; a variant that does not exist gives the default.
template <int M> constexpr int z_rows_v(int v) {
    constexpr int d = z_rows<M>();
    if (M < 32) return d;
    if (v == 1) return d > 1 ? d / 2 : d;
    if (v == 2) {
        const int t = 2 * d;
        const int fft = t * RegFFT<M>::LDS, tile = (M + 1) * (t + 1);
        return (t * RegFFT<M>::G <= 1024 && (fft > tile ? fft : tile) * 8 <= 80 * 1024) ? t : d;
    }
    return d;
}

template <int M, int TR>
static void zfwd_launch(const ConvPlan& cp, const StageIn& si, int NKp, const float2* twM, const float2* twP,
                        hipStream_t s) {
    const int Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2];
    const dim3 g(div_up(Ny, TR), Nx, 3);
    constexpr int nt = TR * RegFFT<M>::G;
    switch (NKp) {
    case 0: k_zfwd<M, TR, 0><<<g, nt, 0, s>>>(si, cp.B1, twM, twP, Nx, Ny, Nz); break;
    case 1: k_zfwd<M, TR, 1><<<g, nt, 0, s>>>(si, cp.B1, twM, twP, Nx, Ny, Nz); break;
    case 2: k_zfwd<M, TR, 2><<<g, nt, 0, s>>>(si, cp.B1, twM, twP, Nx, Ny, Nz); break;
    case 3: k_zfwd<M, TR, 3><<<g, nt, 0, s>>>(si, cp.B1, twM, twP, Nx, Ny, Nz); break;
    default: k_zfwd<M, TR, 5><<<g, nt, 0, s>>>(si, cp.B1, twM, twP, Nx, Ny, Nz); break;
    }
}

static void launch_zfwd(const ConvPlan& cp, const StageIn& in, hipStream_t s) {
    const int M = cp.M;
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    GK_ASSERT(in.nk >= 0 && in.nk <= 5, "stage build with %d earlier stages", in.nk);
    // pad the stage list to the kernel's count with (m, 0): fmaf(0, m, z) == z
    StageIn si = in;
    const int NKp = zfwd_nk(in.nk);
    for (int j = in.nk; j < 6; j++) {
        si.k[j] = in.m;
        si.a[j] = 0.f;
    }
    si.nk = NKp;
    switch (M) {
#define GK_CASE(n)                                                                               \
    case n:                                                                                      \
        if (cp.tp.zr == 1) zfwd_launch<n, z_rows_v<n>(1)>(cp, si, NKp, twM, twP, s);             \
        else if (cp.tp.zr == 2) zfwd_launch<n, z_rows_v<n>(2)>(cp, si, NKp, twM, twP, s);        \
        else zfwd_launch<n, z_rows_v<n>(0)>(cp, si, NKp, twM, twP, s);                           \
        break;
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

bool conv_zr_valid(int M, int v) {
    if (v == 0) return true;
    if (v < 0 || v > 2) return false;
    switch (M) {
#define GK_CASE(n) \
    case n: return z_rows_v<n>(v) != z_rows_v<n>(0);
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    default: return false;
    }
}

static void launch_zinv(const ConvPlan& cp, float* h, const EngineState* st, hipStream_t s) {
    const int M = cp.M, Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2];
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    switch (M) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int TR = z_rows<n>();                                                         \
        dim3 g(div_up(Ny, TR), Nx, 3);                                                          \
        k_zinv<n, TR><<<g, TR * RegFFT<n>::G, 0, s>>>(cp.B1, h, twM, twP, st, Nx, Ny, Nz);      \
        break;                                                                                  \
    }
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

// rows per block of the fused Z inverse: 3 x TR lane groups <= 1024 threads, its
// LDS (transposed tiles, FFT exchange buffers) within 80 KB (two blocks per CU)
template <int M> constexpr int epi_rows() {
    constexpr int G = RegFFT<M>::G;
    int tr = 1024 / (3 * G);
    if (tr > 16) tr = 16;
    int p = 1;
    while (p * 2 <= tr) p *= 2;
    tr = p;
    while (tr > 1 && zinv_epi_lds_elems(M, tr) * 8 > 80 * 1024) tr /= 2;
    return tr;
}

// rows per fused Z inverse block of shape variant v: 0 the default, 1 half of it
template <int M> constexpr int epi_rows_v(int v) {
    constexpr int d = epi_rows<M>();
    return (v == 1 && M >= 32 && d > 1) ? d / 2 : d;
}

int conv_epi_rows(int M, int v) {
    switch (M) {
#define GK_CASE(n) \
    case n: return epi_rows_v<n>(v);
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    default: abort();
    }
}

bool conv_er_valid(int M, int v) { return v == 0 || (v == 1 && conv_epi_rows(M, 1) != conv_epi_rows(M, 0)); }

int conv_epi_blocks(const ConvPlan& cp) { return div_up(cp.N[1], conv_epi_rows(cp.M, cp.tp.er)) * cp.N[0]; }

template <int M, int TR>
static void zinv_epi_launch(const ConvPlan& cp, const EpiArgs& e, const EngineState* st, const float2* twM,
                            const float2* twP, hipStream_t s) {
    const int Nx = cp.N[0], Ny = cp.N[1], Nz = cp.N[2], nu = e.nu;
    const dim3 g(div_up(Ny, TR), Nx), b(3 * TR * RegFFT<M>::G);
    if (e.fin == 1) {
        if (nu <= 1) k_zinv_epi<M, TR, 1, 1><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);
        else k_zinv_epi<M, TR, 3, 1><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);
    } else if (e.fin == 2) {
        if (nu <= 3) k_zinv_epi<M, TR, 3, 2><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);
        else k_zinv_epi<M, TR, 6, 2><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);
    } else {
        if (nu == 0) k_zinv_epi<M, TR, 0, 0><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);
        else if (nu == 1) k_zinv_epi<M, TR, 1, 0><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);
        else if (nu <= 3) k_zinv_epi<M, TR, 3, 0><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);
        else k_zinv_epi<M, TR, 6, 0><<<g, b, 0, s>>>(cp.B1, e, twM, twP, st, Nx, Ny, Nz);
    }
}

static void launch_zinv_epi(const ConvPlan& cp, const EpiArgs& ea, const EngineState* st, hipStream_t s) {
    const int M = cp.M;
    GK_ASSERT(st != nullptr && ea.nu >= 0 && ea.nu <= 6 && (ea.fin != 1 || ea.nu <= 3) &&
                  (ea.fin != 2 || (ea.mhigh && ea.part_err)),
              "fused Z inverse: bad epilogue");
    // pad the list of earlier stages to the kernel's bound with (m, 0, 0): fmaf(0, m, z) == z
    const int nu = ea.nu;
    EpiArgs e = ea;
    for (int j = nu; j < 6; j++) {
        e.k[j] = ea.m;
        e.a[j] = 0.f;
        e.b[j] = 0.f;
        e.ce[j] = 0.0;
    }
    const float2* twM = twiddle_f32(M);
    const float2* twP = twiddle_f32(cp.P[2]);
    switch (M) {
#define GK_CASE(n)                                                                               \
    case n:                                                                                      \
        if (cp.tp.er == 1) zinv_epi_launch<n, epi_rows_v<n>(1)>(cp, e, st, twM, twP, s);         \
        else zinv_epi_launch<n, epi_rows_v<n>(0)>(cp, e, st, twM, twP, s);                       \
        break;
        GK_LEN_M(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

static int cu_count() {
    static int n = 0;
    if (!n) {
        int dev = 0;
        GK_CHECK(hipGetDevice(&dev));
        GK_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    }
    return n;
}

// Split `items` units of work (lines per block row) over at most `blocks` blocks
// so that every block runs the same number of iterations: returns the iteration
// count and sets `grid`.
static int pipeline_split(long items, long blocks, int* grid) {
    const long nit = std::max(1L, (items + blocks - 1) / blocks);
    *grid = (int)((items + nit - 1) / nit);
    return (int)nit;
}

static void launch_y3d(const ConvPlan& cp, const EngineState* st, bool fwd, hipStream_t s) {
    const int PY = cp.P[1], Nx = cp.N[0], Ny = cp.N[1], Q = cp.Q;
    const float2* tw = twiddle_f32(PY);
    const long nlines = 3L * Nx * Q;
    switch (PY) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int LPB = kNT / RegFFT<n>::G;                                                 \
        int g = 0;                                                                              \
        const int nit = pipeline_split(div_up(nlines, LPB), (long)cp.tp.yb * cu_count(), &g); \
        if (fwd) k_ypass<n, true><<<g, kNT, 0, s>>>(cp.B1, cp.B2, tw, st, Nx, Ny, Q, nit);      \
        else k_ypass<n, false><<<g, kNT, 0, s>>>(cp.B2, cp.B1, tw, st, Nx, Ny, Q, nit);          \
        break;                                                                                  \
    }
        GK_LEN_Y(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

// X . K . X launch variants (TuneParams::xv) for padded X 16, 32, 64: 1 = no
// prefetch of the next item, 2 = half the elements per lane over twice the
// lanes, 3 = both; 0 and other lengths: k_xmul
template <int PX, int EE, bool PIPE>
static void xmul_v_launch(const ConvPlan& cp, const EngineState* st, const float2* tw, hipStream_t s) {
    constexpr int LPB = kNT / (PX / EE);
    const int PY = cp.P[1];
    const int lpr = LPB < PY ? LPB : PY;
    int g = 0;
    const int nit = pipeline_split((long)cp.Q * div_up(PY, lpr), (long)cp.tp.xb * cu_count(), &g);
    k_xmul_v<PX, EE, PIPE><<<g, kNT, 0, s>>>(cp.B2, cp.Kc, tw, st, cp.N[0], PY, cp.Q, nit);
}
bool conv_xv_valid(int PX, int v) { return v == 0 || ((PX == 16 || PX == 32 || PX == 64) && v >= 1 && v <= 3); }

static void launch_xmul(const ConvPlan& cp, const EngineState* st, hipStream_t s) {
    const int PX = cp.P[0], PY = cp.P[1], Nx = cp.N[0], Q = cp.Q;
    const float2* tw = twiddle_f32(PX);
    if (cp.tp.xv && conv_xv_valid(PX, cp.tp.xv)) {
        switch (PX * 4 + cp.tp.xv) {
#define GK_XV(n)                                                                            \
    case n * 4 + 1: xmul_v_launch<n, RegFFT<n>::E, false>(cp, st, tw, s); break;            \
    case n * 4 + 2: xmul_v_launch<n, RegFFT<n>::E / 2, true>(cp, st, tw, s); break;        \
    case n * 4 + 3: xmul_v_launch<n, RegFFT<n>::E / 2, false>(cp, st, tw, s); break;
            GK_XV(16) GK_XV(32) GK_XV(64)
#undef GK_XV
        default: abort();
        }
        GK_LAUNCH_CHECK();
        return;
    }
    switch (PX) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int LPB = kNT / RegFFT<n>::G;                                                 \
        const int lpr = LPB < PY ? LPB : PY;                                                    \
        int g = 0;                                                                              \
        const int nit = pipeline_split((long)Q * div_up(PY, lpr), (long)cp.tp.xb * cu_count(), &g); \
        k_xmul<n><<<g, kNT, 0, s>>>(cp.B2, cp.Kc, tw, st, Nx, PY, Q, nit);                      \
        break;                                                                                  \
    }
        GK_LEN_X(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

static void launch_y2d(const ConvPlan& cp, const EngineState* st, hipStream_t s) {
    const int PY = cp.P[1], Ny = cp.N[1], Q = cp.Q;
    const float2* tw = twiddle_f32(PY);
    switch (PY) {
#define GK_CASE(n)                                                                              \
    case n: {                                                                                   \
        constexpr int LPB = kNT / RegFFT<n>::G;                                                 \
        int g = 0;                                                                              \
        const int nit = pipeline_split(div_up(Q, LPB), (long)cp.tp.y2b * cu_count(), &g);  \
        k_y2r<n><<<g, kNT, 0, s>>>(cp.B1, cp.Kc, tw, st, Ny, Q, nit);                           \
        break;                                                                                  \
    }
        GK_LEN_Y(GK_CASE)
#undef GK_CASE
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

static void launch_plane3(const ConvPlan& cp, const EngineState* st, hipStream_t s) {
    const int PX = cp.P[0], Nx = cp.N[0], Ny = cp.N[1], Q = cp.Q;
    GK_ASSERT(cp.plane3 && cp.P[1] == kPlanePY, "the plane pass does not fit this plan");
    GK_ASSERT(Q >= 2 && 2 * Ny <= kPlanePY, "plane pass: Q = %d, Ny = %d", Q, Ny);
    const float2* twL = twiddle_f32(kPlanePY / 2);
    const float2* twP = twiddle_f32(kPlanePY);
    const float2* twX = twiddle_f32(PX);
    // Q - 1 blocks: block 0 takes the two real planes kz = 0 and kz = Q - 1 together
    if (PX == 16) k_plane3<16><<<Q - 1, kPlaneThreads, 0, s>>>(cp.B1, cp.Kc, twL, twP, twX, st, Nx, Ny, Q);
    else k_plane3<32><<<Q - 1, kPlaneThreads, 0, s>>>(cp.B1, cp.Kc, twL, twP, twX, st, Nx, Ny, Q);
    GK_LAUNCH_CHECK();
}

static void launch_yxy(const ConvPlan& cp, const EngineState* st, hipStream_t s) {
    const int PX = cp.P[0], PY = cp.P[1], Nx = cp.N[0], Ny = cp.N[1], Q = cp.Q;
    const float2* twY = twiddle_f32(PY);
    const float2* twX = twiddle_f32(PX);
    switch (PY) {
#define GK_YX(n, px)                                                                                       \
    case px:                                                                                               \
        if constexpr (yxy_fits(px, n))                                                                     \
            k_yxy<n, px><<<Q, kYxyThreads, 0, s>>>(cp.B1, cp.Kc, twY, twX, st, Nx, Ny, Q);                 \
        else                                                                                               \
            abort();                                                                                       \
        break;
#define GK_CASE(n)                                                                                         \
    case n:                                                                                                \
        switch (PX) {                                                                                      \
            GK_YX(n, 4) GK_YX(n, 6) GK_YX(n, 8) GK_YX(n, 12) GK_YX(n, 16)                                  \
        default: abort();                                                                                  \
        }                                                                                                  \
        break;
        GK_LEN_Y(GK_CASE)
#undef GK_CASE
#undef GK_YX
    default: abort();
    }
    GK_LAUNCH_CHECK();
}

void Prof::begin(int cls, hipStream_t s) {
    if (!on) return;
    while ((int)pool.size() < used + 2) {
        hipEvent_t e;
        GK_CHECK(hipEventCreate(&e));
        pool.push_back(e);
    }
    marks.push_back({cls, used});
    GK_CHECK(hipEventRecord(pool[used], s));
    used += 2;
}
void Prof::end(hipStream_t s) {
    if (!on) return;
    GK_CHECK(hipEventRecord(pool[marks.back().second + 1], s));
}
void Prof::collect(double ms[kClasses], int counts[kClasses]) {
    for (int c = 0; c < kClasses; c++) {
        ms[c] = 0;
        counts[c] = 0;
    }
    for (auto& mk : marks) {
        GK_CHECK(hipEventSynchronize(pool[mk.second + 1]));
        float t = 0.f;
        GK_CHECK(hipEventElapsedTime(&t, pool[mk.second], pool[mk.second + 1]));
        ms[mk.first] += t;
        counts[mk.first] += 1;
    }
}
Prof::~Prof() {
    for (auto e : pool) (void)hipEventDestroy(e);
}

void conv_run_pass(const ConvPlan& cp, int cls, const StageIn& in, float* h, const EpiArgs* epi, hipStream_t s) {
    const bool d3 = cp.P[0] > 1;
    const bool pp = cp.plane3 && cp.tp.pp;
    switch (cls) {
    case 0: launch_zfwd(cp, in, s); break;
    case 1: GK_ASSERT(d3 && !cp.yxy && !pp, "no separate Y forward pass"); launch_y3d(cp, in.st, true, s); break;
    case 2:
        if (cp.yxy) launch_yxy(cp, in.st, s);
        else if (pp) launch_plane3(cp, in.st, s);
        else if (d3) launch_xmul(cp, in.st, s);
        else launch_y2d(cp, in.st, s);
        break;
    case 3: GK_ASSERT(d3 && !cp.yxy && !pp, "no separate Y inverse pass"); launch_y3d(cp, in.st, false, s); break;
    case 4: if (epi) launch_zinv_epi(cp, *epi, in.st, s); else launch_zinv(cp, h, in.st, s); break;
    default: GK_ASSERT(false, "pass class %d", cls);
    }
}

void conv_apply(const ConvPlan& cp, const StageIn& in, float* h, hipStream_t s, Prof* prof, const EpiArgs* epi) {
    Prof dummy;
    Prof& p = prof ? *prof : dummy;
    p.begin(0, s);
    launch_zfwd(cp, in, s);
    p.end(s);
    if (cp.P[0] == 1) {
        p.begin(2, s);
        launch_y2d(cp, in.st, s);
        p.end(s);
    } else if (cp.yxy) {
        p.begin(2, s);
        launch_yxy(cp, in.st, s);
        p.end(s);
    } else if (cp.plane3 && cp.tp.pp) {
        p.begin(2, s);
        launch_plane3(cp, in.st, s);
        p.end(s);
    } else {
        p.begin(1, s);
        launch_y3d(cp, in.st, true, s);
        p.end(s);
        p.begin(2, s);
        launch_xmul(cp, in.st, s);
        p.end(s);
        p.begin(3, s);
        launch_y3d(cp, in.st, false, s);
        p.end(s);
    }
    p.begin(4, s);
    if (epi)
        launch_zinv_epi(cp, *epi, in.st, s);
    else
        launch_zinv(cp, h, in.st, s);
    p.end(s);
}

}  // namespace gk
