// Batched fixed-step RKF78 propagation kernel for gfx950.
//
// Work decomposition: a workgroup owns one tile of 64 objects (one object per
// lane) and runs W waves (W = 1, 2, 4, 8). Every wave evaluates one block of
// spherical-harmonic orders for all 64 objects; because all lanes of a wave
// walk the same (n, m) sequence, the coefficient stream is read with scalar
// (uniform) loads straight into SGPR operands of FP64 FMAs. The per-wave
// partial sums meet in LDS and wave 0 reduces them in fixed wave order
// (deterministic). Wave 0 also owns the RKF78 state: the 13 stage
// derivatives live in LDS, the state in registers; the other forces are
// spread over the last waves (drag, then Earth corrections and SRP) so the
// waves reach the per-stage barrier together. Several integration steps run
// per launch; epoch-only inputs come from the per-epoch records.
//
// Reference semantics (src/common/engine/ unless noted):
//   stage recursion / final sum   third_party/parm/parm/integrate/rk/host/Step.h:33-124
//   RKF78 tableau                 third_party/parm/parm/integrate/rk/coefs/tableau.h:19-51
//   step bookkeeping              paraHPOP/propagators/FixedStep.h:30-37
//   force terms                   paraHPOP/model/accelerations/PrivateAccelerations.h:29-209
//   SH evaluation                 paraHPOP/model/accelerations/sphericalharmonics/Evaluation.h:111-159,282-296
//   point gravity / SRP / drag    PointGravity.h:16-81, SRP.h:62-69, Drag.h:67-73
//   shadow                        paraHPOP/model/environment/eclipse/Shadow.h:84-142
//   tide gradient, relativity     paraHPOP/model/accelerations/hpop_compat/HpopCorrections.h:339-376
//   NRLMSISE-00 inputs            paraHPOP/model/environment/atmosphere/Nrlmsise00.cu:243-327
//
// The SH inner loop differs from the reference's per-term form in three
// exact algebraic rearrangements (rounding-level differences only):
//   * the longitude phasor cos(m l), sin(m l) is applied once per order to
//     sums over degree instead of once per term;
//   * alpha^n is folded into the Legendre recursion (scaled functions);
//   * for m >= 1 the latitude derivative uses
//       cos(phi) dP_nm/dphi = -n sin(phi) P_nm + e_nm P_{n-1,m},
//     e_nm = sqrt((2n+1)(n^2-m^2)/(2n-1)), so the dP recursion and the
//     per-term division by cos(phi) disappear (one division per wave).
// Order 0 keeps the dP recursion; the literal pole (cos(phi) == 0) takes the
// reference form with the analytic m = 1 limit.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "phx_common.h"
#include "phx_kernels.h"
#include "phx_epoch_device.h"
#include "phx_nrlmsise.h"

namespace phx {
namespace {

__constant__ double kRkA[12][12] = {
    {2. / 27, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {1. / 36, 1. / 12, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {1. / 24, 0, 1. / 8, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {5. / 12, 0, -25. / 16, 25. / 16, 0, 0, 0, 0, 0, 0, 0, 0},
    {1. / 20, 0, 0, 1. / 4, 1. / 5, 0, 0, 0, 0, 0, 0, 0},
    {-25. / 108, 0, 0, 125. / 108, -65. / 27, 125. / 54, 0, 0, 0, 0, 0, 0},
    {31. / 300, 0, 0, 0, 61. / 225, -2. / 9, 13. / 900, 0, 0, 0, 0, 0},
    {2.0, 0, 0, -53. / 6, 704. / 45, -107. / 9, 67. / 90, 3.0, 0, 0, 0, 0},
    {-91. / 108, 0, 0, 23. / 108, -976. / 135, 311. / 54, -19. / 60, 17. / 6, -1. / 12, 0, 0, 0},
    {2383. / 4100, 0, 0, -341. / 164, 4496. / 1025, -301. / 82, 2133. / 4100, 45. / 82, 45. / 164, 18. / 41, 0, 0},
    {3. / 205, 0, 0, 0, 0, -6. / 41, -3. / 205, -3. / 41, 3. / 41, 6. / 41, 0, 0},
    {-1777. / 4100, 0, 0, -341. / 164, 4496. / 1025, -289. / 82, 2193. / 4100, 51. / 82, 33. / 164, 12. / 41, 0, 1.0}};
__constant__ double kRkB[13] = {0, 0, 0, 0, 0, 34. / 105, 9. / 35, 9. / 35, 9. / 280, 9. / 280, 0, 41. / 840, 41. / 840};
// stages with nonzero b_s (5-9, 11, 12): the only k-stages the step combination reads
constexpr unsigned kRkBNonZero = 0x1BE0u;

enum : uint8_t { ST_TERMINATED = 1, ST_EOS = 2, ST_PAD = 4 };

// Epoch-record accessors: one record per stage epoch shared by the tile
// (uniform loads), or four grid rows around the object's own stage epoch
// combined with cubic Lagrange weights (uneven epochs).
struct ExactRec {
    const double* __restrict__ p;
    __device__ double operator()(int f) const { return ((const __attribute__((address_space(4))) double*)p)[f]; }
};
struct GridRec {
    const double* __restrict__ p;  // row k-1
    double w0, w1, w2, w3;
    __device__ double operator()(int f) const {
        return w0 * p[f] + w1 * p[REC_SIZE + f] + w2 * p[2 * REC_SIZE + f] + w3 * p[3 * REC_SIZE + f];
    }
};
// Grid rows with the object's leap-second rotation: rows hold R3(-theta) R
// (smooth); the first two rows of R are turned back by the object's theta
// (c = cos theta, s = sin theta).
struct GridRecL {
    GridRec g;
    double c, s;
    __device__ double operator()(int f) const {
        if (f < 3) return c * g(f) + s * g(f + 3);
        if (f < 6) return c * g(f) - s * g(f - 3);
        return g(f);
    }
};
__device__ inline GridRec gridRec(const double* table, int count, double t0, double invH, double ts) {
    const double x = (ts - t0) * invH;
    int k = int(floor(x));
    k = k < 1 ? 1 : (k > count - 3 ? count - 3 : k);
    const double f = x - double(k);
    GridRec r;
    r.p = table + size_t(k - 1) * REC_SIZE;
    r.w0 = -f * (f - 1.0) * (f - 2.0) * (1.0 / 6.0);
    r.w1 = (f + 1.0) * (f - 1.0) * (f - 2.0) * 0.5;
    r.w2 = -(f + 1.0) * f * (f - 2.0) * 0.5;
    r.w3 = (f + 1.0) * f * (f - 1.0) * (1.0 / 6.0);
    return r;
}

// Per-lane records whose fields below the NRLMSISE-00 block (rotation, Earth
// rate, indirect term, bodies, tide coefficients) were interpolated once per
// stage into LDS by the waves idle in the stage-state phase; the rest are
// interpolated on use.
constexpr int kRecStaged = REC_NRL;
struct GridRecF {
    GridRec g;
    const double (*r)[kTile];
    int lane;
    __device__ double operator()(int f) const { return f < kRecStaged ? r[f][lane] : g(f); }
};

// Per-lane geometry of the body-fixed position used by the SH evaluation.
struct ShGeom {
    double sinPhi, cosPhi, sinL, cosL, alpha, r;
    double bx, by, bz;   // body-fixed position
};

template <class Rec>
__device__ inline ShGeom shGeometry(const Rec& rec, double px, double py, double pz, double radius) {
    ShGeom g;
    g.bx = rec(REC_R + 0) * px + rec(REC_R + 1) * py + rec(REC_R + 2) * pz;
    g.by = rec(REC_R + 3) * px + rec(REC_R + 4) * py + rec(REC_R + 5) * pz;
    g.bz = rec(REC_R + 6) * px + rec(REC_R + 7) * py + rec(REC_R + 8) * pz;
    g.r = sqrt(g.bx * g.bx + g.by * g.by + g.bz * g.bz);
    const double rho = sqrt(g.bx * g.bx + g.by * g.by);
    g.sinPhi = g.bz / g.r;
    g.cosPhi = rho / g.r;
    g.sinL = rho > 1e-14 ? g.by / rho : 0.0;
    g.cosL = rho > 1e-14 ? g.bx / rho : 1.0;
    g.alpha = radius / g.r;
    return g;
}

// Read-only tables in the constant address space: uniform indices become
// scalar (SMEM) loads into SGPR operands of the FMAs.
typedef const __attribute__((address_space(4))) double* ConstD;
typedef const __attribute__((address_space(4))) float* ConstF;
typedef const __attribute__((address_space(4))) int* ConstI;

// One run of consecutive degrees of order m >= 1 (column recursion
// P~_{n+1} = a u P~_n - b w P~_{n-1}); T is the arithmetic type of the run, the
// six partial sums are returned in T.
template <class T, class CP>
__device__ __attribute__((always_inline)) inline void shRun(CP t, int count, T u, T w, T& P0, T& Pm1, T& A1, T& A2, T& A5,
                                                            T& A6, T& B1, T& B2) {
    // Four terms per trip: their uniform coefficients are requested together
    // so one scalar-load latency covers the FMAs of four terms.
#pragma unroll 4
    for (int i = 0; i < count; ++i, t += SHT_STRIDE) {
        A1 = fma(P0, t[SHT_C1], A1);
        A2 = fma(P0, t[SHT_S1], A2);
        A5 = fma(P0, t[SHT_C], A5);
        A6 = fma(P0, t[SHT_S], A6);
        B1 = fma(Pm1, t[SHT_EC], B1);
        B2 = fma(Pm1, t[SHT_ES], B2);
        const T Pn = fma(t[SHT_A] * u, P0, -(t[SHT_B] * w) * Pm1);
        Pm1 = P0;
        P0 = Pn;
    }
}

// Order 0: P~ and its latitude derivative by the reference's dP recursion.
template <class T, class CP>
__device__ __attribute__((always_inline)) inline void shRun0(CP t, int count, T u, T v, T w, T& P0, T& Pm1, T& dP0, T& dPm1,
                                                             T& X1, T& D1) {
#pragma unroll 4
    for (int i = 0; i < count; ++i, t += SHT_STRIDE) {
        X1 = fma(P0, t[SHT_C1], X1);
        D1 = fma(dP0, t[SHT_C], D1);
        const T au = t[SHT_A] * u, bw = t[SHT_B] * w;
        const T Pn = fma(au, P0, -bw * Pm1);
        const T dPn = fma(t[SHT_A], fma(v, P0, u * dP0), -bw * dPm1);
        Pm1 = P0;
        P0 = Pn;
        dPm1 = dP0;
        dP0 = dPn;
    }
}

// Order block [m0, m1) of the SH sum for one lane; returns the unscaled
// spherical components (radial, latitude, longitude) of sum_nm alpha^n (...).
// All control flow over (n, m) is uniform across the wave; only the literal
// pole (x = y = 0 in the body frame) takes a divergent fix-up after the loop.
// Mixed precision (fp64Degree >= 0): the orders m <= fp64Order and the degrees
// n <= fp64Degree run in FP64; the other terms run in FP32 (recursion and
// per-order partial sums, FP32 coefficient stream), and each order's FP32
// partial sums are accumulated into the FP64 sums; the diagonal seeds and the
// longitude phasor stay FP64. The low orders stay FP64 because their latitude
// and longitude sums are divided by cos(phi) after a cancellation of terms
// about n times larger than the result near the poles, which an FP32
// recursion does not survive.
__device__ void shOrderBlock(const ShGeom& g, int m0, int m1, int L, const double* termsG, const float* terms32G,
                             int fp64Degree, int fp64Order, const int* orderStartG, const double* diagG, bool normalized,
                             double& outR, double& outPhi, double& outLam) {
    ConstD terms = (ConstD)termsG;
    ConstF terms32 = (ConstF)terms32G;
    ConstI orderStart = (ConstI)orderStartG;
    ConstD diag = (ConstD)diagG;
    m0 = __builtin_amdgcn_readfirstlane(m0);
    m1 = __builtin_amdgcn_readfirstlane(m1);
    L = __builtin_amdgcn_readfirstlane(L);
    // first degree evaluated in FP32 (L + 1: none)
    const int nSplit = __builtin_amdgcn_readfirstlane(fp64Degree < 0 ? L + 1 : min(fp64Degree + 1, L + 1));
    // orders [m0, mFlat) run entirely in FP64 in the flat loop below
    const int mFlat = __builtin_amdgcn_readfirstlane(nSplit > L ? m1 : min(m1, fp64Order + 1));
    const double u = g.alpha * g.sinPhi;   // alpha sin(phi)
    const double v = g.alpha * g.cosPhi;   // alpha cos(phi)
    const double w = g.alpha * g.alpha;
    const float uf = float(u), wf = float(w);
    double ar = 0.0, aphiDirect = 0.0, aphiNum = 0.0, alamNum = 0.0;
    // Seed P~_{m0,m0} = (prod_{k <= m0} c_k) (alpha cos phi)^m0 and
    // cos/sin(m0 lambda): the power and the phasor by binary powering.
    double Pmm = 1.0, cosML = 1.0, sinML = 0.0;
    if (m0 > 0) {
        double vp = 1.0, bv = v, bc = g.cosL, bs = g.sinL;
        for (int e = m0; e > 0; e >>= 1) {
            if (e & 1) {
                vp *= bv;
                const double c = cosML * bc - sinML * bs;
                const double s = sinML * bc + cosML * bs;
                cosML = c;
                sinML = s;
            }
            bv *= bv;
            const double c2 = bc * bc - bs * bs;
            const double s2 = 2.0 * bc * bs;
            bc = c2;
            bs = s2;
        }
        Pmm = diag[L + 1 + m0] * vp;
    }
    if (m0 < mFlat) {
        // FP64 stream: order 0 by its own recursion, orders >= 1 as one flat loop
        // over four-term batches (each order's run is padded with zero terms to a
        // multiple of four), so an order change costs no extra round trip for its
        // start, its remainder or its next diagonal factor
        int m = m0;
        if (m == 0) {
            double P0 = Pmm, Pm1 = 0.0, dP0 = 0.0, dPm1 = 0.0;
            double X1 = 0.0, D1 = 0.0;
            // the padded run (zero terms leave the sums unchanged)
            shRun0<double>(terms + SHT_STRIDE * orderStart[0], (L + 4) & ~3, u, v, w, P0, Pm1, dP0, dPm1, X1, D1);
            ar -= X1;
            aphiDirect += D1;
            if (1 <= L) Pmm = diag[1] * v * Pmm;
            const double s = g.cosL * sinML + g.sinL * cosML;
            const double c = g.cosL * cosML - g.sinL * sinML;
            sinML = s;
            cosML = c;
            m = 1;
        }
        if (m < mFlat) {
            ConstD t = terms + SHT_STRIDE * orderStart[m];
            int left = (L - m + 4) >> 2;
            // the next diagonal factor is prefetched an order ahead
            double dNext = diag[min(m + 1, L)];
            double P0 = Pmm, Pm1 = 0.0;
            double A1 = 0.0, A2 = 0.0, A5 = 0.0, A6 = 0.0, B1 = 0.0, B2 = 0.0;
            while (true) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    ConstD q = t + j * SHT_STRIDE;
                    A1 = fma(P0, q[SHT_C1], A1);
                    A2 = fma(P0, q[SHT_S1], A2);
                    A5 = fma(P0, q[SHT_C], A5);
                    A6 = fma(P0, q[SHT_S], A6);
                    B1 = fma(Pm1, q[SHT_EC], B1);
                    B2 = fma(Pm1, q[SHT_ES], B2);
                    const double Pn = fma(q[SHT_A] * u, P0, -(q[SHT_B] * w) * Pm1);
                    Pm1 = P0;
                    P0 = Pn;
                }
                t += 4 * SHT_STRIDE;
                if (--left == 0) {
                    const double X1 = cosML * A1 + sinML * A2;
                    const double X5 = cosML * A5 + sinML * A6;
                    const double XE = cosML * B1 + sinML * B2;
                    const double X6 = sinML * A5 - cosML * A6;
                    ar -= X1;
                    aphiNum += g.alpha * XE - g.sinPhi * (X1 - X5);
                    alamNum -= double(m) * X6;
                    // next diagonal and phasor (also after the last order: the
                    // mixed orders below start from them)
                    if (m + 1 <= L) Pmm = dNext * v * Pmm;
                    const double s = g.cosL * sinML + g.sinL * cosML;
                    const double c = g.cosL * cosML - g.sinL * sinML;
                    sinML = s;
                    cosML = c;
                    if (++m >= mFlat) break;
                    dNext = diag[min(m + 1, L)];
                    left = (L - m + 4) >> 2;
                    P0 = Pmm;
                    Pm1 = 0.0;
                    A1 = A2 = A5 = A6 = B1 = B2 = 0.0;
                }
            }
        }
    }
    // mixed orders (m >= 1: order 0 is always in the FP64 orders) with degrees
    // in FP64 and in FP32, one order at a time
    const int mFp32 = __builtin_amdgcn_readfirstlane(min(m1, max(max(m0, mFlat), nSplit)));
    for (int m = max(m0, mFlat); m < mFp32; ++m) {
        const int first = orderStart[m];
        const int count = L - m + 1;
        const int c64 = max(0, min(count, nSplit - m));   // degrees m .. m + c64 - 1 in FP64
        const int c32 = count - c64;
        {
            double P0 = Pmm, Pm1 = 0.0;
            double A1 = 0.0, A2 = 0.0, A5 = 0.0, A6 = 0.0, B1 = 0.0, B2 = 0.0;
            shRun<double>(terms + SHT_STRIDE * first, c64, u, w, P0, Pm1, A1, A2, A5, A6, B1, B2);
            if (c32 > 0) {
                float p0 = float(P0), pm1 = float(Pm1);
                float a1 = 0.0f, a2 = 0.0f, a5 = 0.0f, a6 = 0.0f, b1 = 0.0f, b2 = 0.0f;
                shRun<float>(terms32 + SHT_STRIDE * (first + c64), c32, uf, wf, p0, pm1, a1, a2, a5, a6, b1, b2);
                A1 += double(a1);
                A2 += double(a2);
                A5 += double(a5);
                A6 += double(a6);
                B1 += double(b1);
                B2 += double(b2);
            }
            const double X1 = cosML * A1 + sinML * A2;
            const double X5 = cosML * A5 + sinML * A6;
            const double XE = cosML * B1 + sinML * B2;
            const double X6 = sinML * A5 - cosML * A6;
            ar -= X1;
            aphiNum += g.alpha * XE - g.sinPhi * (X1 - X5);
            alamNum -= double(m) * X6;
        }
        // next diagonal and phasor
        if (m + 1 <= L) Pmm = diag[m + 1] * v * Pmm;
        const double s = g.cosL * sinML + g.sinL * cosML;
        const double c = g.cosL * cosML - g.sinL * sinML;
        sinML = s;
        cosML = c;
    }
    // orders whose degrees are all in FP32 (m >= nSplit): one flat loop over the
    // four-term batches of the padded FP32 stream, as the FP64 loop above; each
    // order's FP32 sums are those of the per-order form (the padding terms add
    // exact zeros)
    if (mFp32 < m1) {
        int m = mFp32;
        ConstF t = terms32 + SHT_STRIDE * orderStart[m];
        int left = (L - m + 4) >> 2;
        float p0 = float(Pmm), pm1 = 0.0f;
        float a1 = 0.0f, a2 = 0.0f, a5 = 0.0f, a6 = 0.0f, b1 = 0.0f, b2 = 0.0f;
        while (true) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                ConstF q = t + j * SHT_STRIDE;
                a1 = fma(p0, q[SHT_C1], a1);
                a2 = fma(p0, q[SHT_S1], a2);
                a5 = fma(p0, q[SHT_C], a5);
                a6 = fma(p0, q[SHT_S], a6);
                b1 = fma(pm1, q[SHT_EC], b1);
                b2 = fma(pm1, q[SHT_ES], b2);
                const float pn = fma(q[SHT_A] * uf, p0, -(q[SHT_B] * wf) * pm1);
                pm1 = p0;
                p0 = pn;
            }
            t += 4 * SHT_STRIDE;
            if (--left == 0) {
                double A1 = 0.0, A2 = 0.0, A5 = 0.0, A6 = 0.0, B1 = 0.0, B2 = 0.0;
                A1 += double(a1);
                A2 += double(a2);
                A5 += double(a5);
                A6 += double(a6);
                B1 += double(b1);
                B2 += double(b2);
                const double X1 = cosML * A1 + sinML * A2;
                const double X5 = cosML * A5 + sinML * A6;
                const double XE = cosML * B1 + sinML * B2;
                const double X6 = sinML * A5 - cosML * A6;
                ar -= X1;
                aphiNum += g.alpha * XE - g.sinPhi * (X1 - X5);
                alamNum -= double(m) * X6;
                if (m + 1 <= L) Pmm = diag[m + 1] * v * Pmm;
                const double s = g.cosL * sinML + g.sinL * cosML;
                const double c = g.cosL * cosML - g.sinL * sinML;
                sinML = s;
                cosML = c;
                if (++m >= m1) break;
                left = (L - m + 4) >> 2;
                p0 = float(Pmm);
                pm1 = 0.0f;
                a1 = a2 = a5 = a6 = b1 = b2 = 0.0f;
            }
        }
    }
    outR = ar;
    if (g.cosPhi != 0.0) {
        const double inv = 1.0 / g.cosPhi;
        outPhi = aphiDirect + aphiNum * inv;
        outLam = alamNum * inv;
        return;
    }
    // Literal pole: every P~_nm with m >= 1 vanishes; order 1 keeps a finite
    // latitude derivative (dP recursion from dP~_11 = -c_1 alpha sin(phi)) and
    // the analytic limit of P_n1 / cos(phi) (Coefficients.h:123-132). FP64.
    outPhi = aphiDirect;
    outLam = 0.0;
    if (m0 <= 1 && 1 < m1 && L >= 1) {
        ConstD t = terms + SHT_STRIDE * orderStart[1];
        double P0 = 0.0, Pm1 = 0.0, dP0 = -diag[1] * g.alpha * g.sinPhi, dPm1 = 0.0;
        double D1 = 0.0, D2 = 0.0, Lc = 0.0, Ls = 0.0, an = g.alpha;
        for (int n = 1; n <= L; ++n, t += SHT_STRIDE) {
            D1 = fma(dP0, t[SHT_C], D1);
            D2 = fma(dP0, t[SHT_S], D2);
            const double sign = g.sinPhi > 0.0 ? 1.0 : ((n & 1) ? 1.0 : -1.0);
            double lim = 0.5 * double(n) * (n + 1.0);
            if (normalized) lim *= sqrt(2.0 * (2.0 * n + 1.0) / (double(n) * (n + 1.0)));
            Lc = fma(sign * lim * an, t[SHT_C], Lc);
            Ls = fma(sign * lim * an, t[SHT_S], Ls);
            const double a = t[SHT_A], b = t[SHT_B];
            const double Pn = a * u * P0 - b * w * Pm1;
            const double dPn = a * (u * dP0 + v * P0) - b * w * dPm1;
            Pm1 = P0;
            P0 = Pn;
            dPm1 = dP0;
            dP0 = dPn;
            an *= g.alpha;
        }
        // order-1 phasor: cos(l), sin(l)
        outPhi += g.cosL * D1 + g.sinL * D2;
        outLam = -(g.sinL * Lc - g.cosL * Ls);
    }
}

template <class Rec>
__device__ inline void shToInertial(const Rec& rec, const ShGeom& g, double ar, double aphi, double alam,
                                    double gm, double& ax, double& ay, double& az) {
    const double s = gm / (g.r * g.r);
    ar *= s;
    aphi *= s;
    alam *= s;
    const double bx = g.cosPhi * g.cosL * ar - g.sinPhi * g.cosL * aphi - g.sinL * alam;
    const double by = g.cosPhi * g.sinL * ar - g.sinPhi * g.sinL * aphi + g.cosL * alam;
    const double bz = g.sinPhi * ar + g.cosPhi * aphi;
    ax = rec(REC_R + 0) * bx + rec(REC_R + 3) * by + rec(REC_R + 6) * bz;
    ay = rec(REC_R + 1) * bx + rec(REC_R + 4) * by + rec(REC_R + 7) * bz;
    az = rec(REC_R + 2) * bx + rec(REC_R + 5) * by + rec(REC_R + 8) * bz;
}

__device__ inline double clamp1(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

// Dual-cone occultation factor (Shadow.h:84-142).
__device__ double shadowFactor(double sx, double sy, double sz, double bx, double by, double bz, double RSun, double RBody) {
    const double s2 = sx * sx + sy * sy + sz * sz, b2 = bx * bx + by * by + bz * bz, sb = sx * bx + sy * by + sz * bz;
    if (b2 <= RBody * RBody) return 0.0;
    const double rInvS = 1.0 / sqrt(s2), rInvB = 1.0 / sqrt(b2);
    const double a = asin(clamp1(RSun * rInvS, 0.0, 1.0));
    const double b = asin(clamp1(RBody * rInvB, 0.0, 1.0));
    const double c = acos(clamp1(sb * rInvS * rInvB, -1.0, 1.0));
    if (c >= a + b) return 1.0;
    if (c <= fabs(b - a)) return b >= a ? 0.0 : 1.0 - (b * b) / (a * a);
    const double a2 = a * a, bb = b * b, c2 = c * c;
    const double x = (c2 + a2 - bb) / (2.0 * c);
    const double y = sqrt(fmax(a2 - x * x, 0.0));
    const double al = acos(clamp1(x / a, -1.0, 1.0));
    const double be = acos(clamp1((c - x) / b, -1.0, 1.0));
    return 1.0 - (a2 * al + bb * be - c * y) / (a2 * 3.14159265358979323846);
}

template <class Rec>
__device__ void srpAccel(const ForceModel& fm, const Rec& rec, double px, double py, double pz, double ams, double& ax,
                         double& ay, double& az) {
    const int sb = REC_BODY + 3 * fm.srpBody;
    const double dx = px - rec(sb), dy = py - rec(sb + 1), dz = pz - rec(sb + 2);
    double occ = 1.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k >= fm.nOccluders) break;
        const int ob = REC_BODY + 3 * fm.occBody[k];
        occ *= shadowFactor(dx, dy, dz, px - rec(ob), py - rec(ob + 1), pz - rec(ob + 2), fm.sunRadius, fm.occRadius[k]);
    }
    const double s2 = dx * dx + dy * dy + dz * dz;
    const double k = occ * (fm.srpFactor * ams) / (s2 * sqrt(s2));
    ax = k * dx;
    ay = k * dy;
    az = k * dz;
}

// Geodetic latitude/longitude/altitude of an Earth-fixed position (Nrlmsise00.cu:243-266).
__device__ bool geodetic(double x, double y, double z, double radius, double flat, double& lat, double& lon, double& alt) {
    const double rho2 = x * x + y * y;
    if (!(rho2 + z * z > 0.0)) return false;
    const double e2 = flat * (2.0 - flat);
    double dz = e2 * z, n = radius;
    for (int it = 0; it < 20; ++it) {
        const double zdz = z + dz;
        const double nh = sqrt(rho2 + zdz * zdz);
        const double sp = zdz / nh;
        n = radius / sqrt(1.0 - e2 * sp * sp);
        const double next = n * e2 * sp;
        if (fabs(dz - next) < 1.0e-12 * radius) {
            dz = next;
            break;
        }
        dz = next;
    }
    const double zdz = z + dz;
    const double nh = sqrt(rho2 + zdz * zdz);
    lon = atan2(y, x);
    lat = atan2(zdz, sqrt(rho2));
    alt = nh - n;
    return isfinite(lat) && isfinite(lon) && isfinite(alt);
}

// NRLMSISE-00 inputs of one lane: Earth-fixed position -> geodetic latitude,
// longitude, altitude and the model input record (Nrlmsise00.cu:270-327).
struct NrlLane {
    nrlmsise::Input in;
    int state;   // 0 above hCutoff (rho = 0), 1 model (alt >= 72.5 km), 2 full model below 72.5 km, 3 invalid
};

template <class Rec>
__device__ inline NrlLane nrlLane(const ForceModel& fm, const Rec& rec, const double* nrl, double px, double py, double pz) {
    NrlLane L;
    const double fx = rec(REC_R + 0) * px + rec(REC_R + 1) * py + rec(REC_R + 2) * pz;
    const double fy = rec(REC_R + 3) * px + rec(REC_R + 4) * py + rec(REC_R + 5) * pz;
    const double fz = rec(REC_R + 6) * px + rec(REC_R + 7) * py + rec(REC_R + 8) * pz;
    double lat, lon, alt;
    const double* r = nrl - REC_NRL;
    L.state = 0;
    if (!geodetic(fx, fy, fz, fm.earthRadius, fm.flattening, lat, lon, alt)) {
        // invalid position (its density is NaN): stored at 500 km, where the
        // species waves that evaluate the whole tile take their shortest branches
        L.state = 3;
        alt = 500.0;
        lat = lon = 0.0;
    } else if (alt > fm.hCutoff) {
        L.state = 0;
    } else if (r[REC_NRL_VALID] == 0.0) {
        L.state = 3;
    } else {
        L.state = alt >= 72.5 ? 1 : 2;
    }
    const double rad2deg = 57.2957795130823208768;
    L.in.doy = int(r[REC_NRL_DOY]);
    L.in.sec = r[REC_NRL_SEC];
    L.in.alt = alt;
    L.in.g_lat = lat * rad2deg;
    L.in.g_long = lon * rad2deg;
    double lst = fmod(r[REC_NRL_GAST] + lon * rad2deg / 15.0, 24.0);
    L.in.lst = lst < 0.0 ? lst + 24.0 : lst;
    L.in.f107A = r[REC_NRL_F107A];
    L.in.f107 = r[REC_NRL_F107];
    L.in.ap = r[REC_NRL_AP];
    for (int k = 0; k < 7; ++k) L.in.apa[k] = r[REC_NRL_AP + k];
    return L;
}

// The lane inputs nrlLane derives from the stage position (altitude, geodetic
// latitude and longitude in degrees, local solar time, state), kept in HBM by
// the split SH kernel for the species and finish kernels of the stage.
__device__ inline void nrlLaneStore(const NrlLane& L, double* lane, int nPad, int obj) {
    lane[obj] = L.in.alt;
    lane[size_t(nPad) + obj] = L.in.g_lat;
    lane[size_t(2) * nPad + obj] = L.in.g_long;
    lane[size_t(3) * nPad + obj] = L.in.lst;
    lane[size_t(4) * nPad + obj] = double(L.state);
}

// nrlLaneStore for the one-workgroup SH kernel: the state row always (the drag
// kernels leave a tile whose lanes are all in state 0 after reading it), the
// other rows only when a lane of the tile is below the cutoff. Returns whether
// one is (uniform across the wave).
__device__ inline bool nrlLaneStoreTile(const NrlLane& L, double* lane, int nPad, int obj) {
    const bool any = __ballot(L.state != 0) != 0ull;
    if (any) nrlLaneStore(L, lane, nPad, obj);
    else lane[size_t(4) * nPad + obj] = 0.0;
    return any;
}

// The globe basis trigonometry of a lane's inputs, kept beside them for the
// drag tile kernel (rows of `trig`, BasisTrig order).
__device__ inline void nrlTrigStore(const NrlLane& L, double* trig, int nPad, int obj) {
    const nrlmsise::BasisTrig t = nrlmsise::basisTrig(L.in);
    trig[obj] = t.c;
    trig[size_t(nPad) + obj] = t.s;
    trig[size_t(2) * nPad + obj] = t.sl;
    trig[size_t(3) * nPad + obj] = t.cl;
    trig[size_t(4) * nPad + obj] = t.cg;
    trig[size_t(5) * nPad + obj] = t.sg;
}
__device__ inline nrlmsise::BasisTrig nrlTrigLoad(const double* trig, int nPad, int obj) {
    nrlmsise::BasisTrig t;
    t.c = trig[obj];
    t.s = trig[size_t(nPad) + obj];
    t.sl = trig[size_t(2) * nPad + obj];
    t.cl = trig[size_t(3) * nPad + obj];
    t.cg = trig[size_t(4) * nPad + obj];
    t.sg = trig[size_t(5) * nPad + obj];
    return t;
}

// Whether a lane of the tile needs the drag kernels at this stage, from the
// state row the SH kernel leaves for every tile (uniform across the workgroup).
__device__ inline bool nrlTileActive(const double* lane, int nPad, int obj) {
    return __ballot(lane[size_t(4) * nPad + obj] != 0.0) != 0ull;
}

// The five lane input rows of one lane (altitude, latitude, longitude, local
// time, state), loaded together.
struct LaneRows {
    double v[5];
};
__device__ inline LaneRows laneRows(const double* lane, int nPad, int obj) {
    LaneRows r;
#pragma unroll
    for (int i = 0; i < 5; ++i) r.v[i] = lane[size_t(i) * nPad + obj];
    return r;
}

// nrlLane's result from the kept lane inputs and the stage's epoch record.
__device__ inline NrlLane nrlLaneFrom(const LaneRows& rows, const double* nrl) {
    const double* r = nrl - REC_NRL;
    NrlLane L;
    L.state = int(rows.v[4]);
    L.in.doy = int(r[REC_NRL_DOY]);
    L.in.sec = r[REC_NRL_SEC];
    L.in.alt = rows.v[0];
    L.in.g_lat = rows.v[1];
    L.in.g_long = rows.v[2];
    L.in.lst = rows.v[3];
    L.in.f107A = r[REC_NRL_F107A];
    L.in.f107 = r[REC_NRL_F107];
    L.in.ap = r[REC_NRL_AP];
    for (int k = 0; k < 7; ++k) L.in.apa[k] = r[REC_NRL_AP + k];
    return L;
}

__device__ inline NrlLane nrlLaneLoad(const double* lane, int nPad, int obj, const double* nrl) {
    return nrlLaneFrom(laneRows(lane, nPad, obj), nrl);
}

// This lane's NRLMSISE-00 time inputs (in the epoch record's REC_NRL block
// layout) and weather-slot row at its stage epoch ts, for per-lane epochs (the
// fused kernel's statements); returns the row index, sec is the UT second.
__device__ inline int laneNrlInputs(const StepArgs& a, double ts, double (&nrl)[13], double& sec) {
    EpochArgs wa{};
    wa.weather = a.weather;
    wa.wDays = a.wDays;
    wa.wStart = a.wStart;
    const double mjd = epochdev::etToMjdUtcNrl(ts);
    int doy, record, wslot;
    epochdev::doySecOf(mjd, doy, sec);
    nrl[REC_NRL_DOY - REC_NRL] = double(doy);
    nrl[REC_NRL_SEC - REC_NRL] = sec;
    nrl[REC_NRL_GAST - REC_NRL] = epochdev::gastHours(mjd);
    const bool ok = epochdev::weatherSlotOf(wa, mjd, record, wslot) && a.nrlSlots;
    const int row = ok ? record * 8 + wslot : 0;
    const double* r = a.nrlSlots + size_t(row) * nrlmsise::SLOT_SIZE;
    for (int k = 0; k < 9; ++k) nrl[REC_NRL_F107 - REC_NRL + k] = ok ? r[nrlmsise::SLOT_F107 + k] : nan("");
    nrl[REC_NRL_VALID - REC_NRL] = ok ? 1.0 : 0.0;
    return row;
}

// Per-lane epochs: the lane's UT second, day of year and weather-slot row after
// the five nrlLaneStore fields, and nrlLane's result rebuilt from them.
__device__ inline void nrlLaneStoreGrid(double sec, int doy, int row, double* lane, int nPad, int obj) {
    lane[size_t(5) * nPad + obj] = sec;
    lane[size_t(6) * nPad + obj] = double(doy);
    lane[size_t(7) * nPad + obj] = double(row);
}

// Per-lane epochs: nrlLane's result from the five lane input rows (loaded
// together), the lane's UT second, day of year and weather-slot row.
__device__ inline NrlLane nrlLaneFromGrid(const LaneRows& rows, const double* lane, int nPad, int obj, const double* slots,
                                          int& row, double (&nrl)[13]) {
    row = int(lane[size_t(7) * nPad + obj]);
    nrl[REC_NRL_DOY - REC_NRL] = lane[size_t(6) * nPad + obj];
    nrl[REC_NRL_SEC - REC_NRL] = lane[size_t(5) * nPad + obj];
    for (int k = 0; k < 9; ++k)
        nrl[REC_NRL_F107 - REC_NRL + k] = slots ? slots[size_t(row) * nrlmsise::SLOT_SIZE + nrlmsise::SLOT_F107 + k] : nan("");
    return nrlLaneFrom(rows, nrl);
}

__device__ inline NrlLane nrlLaneLoadGrid(const double* lane, int nPad, int obj, const double* slots, int& row,
                                          double (&nrl)[13]) {
    return nrlLaneFromGrid(laneRows(lane, nPad, obj), lane, nPad, obj, slots, row, nrl);
}

__device__ inline nrlmsise::Flags nrlFlags() {
    nrlmsise::Flags fl;
    for (int k = 0; k < 24; ++k) fl.switches[k] = 1;
    fl.switches[9] = -1;
    nrlmsise::tselec(fl);
    return fl;
}

// Species order for the wave split (wave w evaluates entries w, w + W, ...):
// with four waves each wave gets at most one of O, H and N, the species whose
// mixing corrections run at 300-450 km, and the wave that also adds tides and
// SRP (the third) gets Ar and anomalous O, which never mix there.
__constant__ int kSpeciesOrder[8] = {0, 1, 4, 3, 2, 5, 7, 6};

// Constant phases of the eleven globe7 coefficient tables (nrlmsise::GT_*).
__constant__ double kGlobeTrig[nrlmsise::GE_TABLES * nrlmsise::GT_COUNT];

// NRLMSISE-00 epoch terms of a lane: MODE 1 reads the globe block of the
// lane's epoch record (shared-epoch batches), MODE 2 the weather-slot row of
// its epoch (uneven epochs), MODE 0 evaluates globe7 as published.
struct NrlSlot {
    const double* row;   // weather-slot row (MODE 2)
    double cs, ss;       // cos/sin(sr * seconds of day)
};

// Species k0, k0 + stride, ... of the thermospheric model for this lane
// (cgs number densities before the metric factor) into spec[k][lane].
template <int MODE>
__device__ __attribute__((always_inline)) inline void nrlSpecies(const NrlLane& L, const double* nrl, const NrlSlot& slot,
                                                                 const double* __restrict__ coeff, int k0, int stride,
                                                                 double (*spec)[kTile], int lane) {
    nrlmsise::Input in = L.in;
    const nrlmsise::Flags fl = nrlFlags();
    nrlmsise::State st;
    const nrlmsise::NrlTables T = nrlmsise::nrlTables(coeff);
    nrlmsise::Pre pre;
    // gtd7's gravity/radius at this latitude (switch 2 on), Nrlmsise00Model.inl:976-980
    nrlmsise::glatf(in.g_lat, st.gsurf, st.re);
    if (MODE == 1) {
        const nrlmsise::FastGlobe G{nrl + (REC_NRL_GLOBE - REC_NRL), kGlobeTrig};
        nrlmsise::gts7Prefix(st, T, in, fl, G, pre);
        nrlmsise::DensuShared h;
        nrlmsise::gts7SpeciesShare(h, kSpeciesOrder[k0] == 0, st, T, in, fl, G, pre);
        for (int i = k0; i < 8; i += stride) {
            const int k = kSpeciesOrder[i];
            spec[k][lane] = nrlmsise::gts7Species(k, h, st, T, in, fl, G, pre);
        }
    } else if (MODE == 2) {
        const nrlmsise::SlotGlobe G{slot.row, kGlobeTrig, slot.cs, slot.ss};
        nrlmsise::gts7Prefix(st, T, in, fl, G, pre);
        nrlmsise::DensuShared h;
        nrlmsise::gts7SpeciesShare(h, kSpeciesOrder[k0] == 0, st, T, in, fl, G, pre);
        for (int i = k0; i < 8; i += stride) {
            const int k = kSpeciesOrder[i];
            spec[k][lane] = nrlmsise::gts7Species(k, h, st, T, in, fl, G, pre);
        }
    } else {
        const nrlmsise::GenericGlobe G{};
        nrlmsise::gts7Prefix(st, T, in, fl, G, pre);
        nrlmsise::DensuShared h;
        nrlmsise::gts7SpeciesShare(h, kSpeciesOrder[k0] == 0, st, T, in, fl, G, pre);
        for (int i = k0; i < 8; i += stride) {
            const int k = kSpeciesOrder[i];
            spec[k][lane] = nrlmsise::gts7Species(k, h, st, T, in, fl, G, pre);
        }
    }
}

// Full model for one lane (serial gtd7d), kg/km^3. Takes the density scale
// rather than the ForceModel: a reference to a kernel argument passed to a call
// makes every caller copy the whole argument to scratch memory.
template <int MODE>
__device__ __noinline__ double nrlDensity(double densityScale, const double* nrl, const double* __restrict__ coeff,
                                          nrlmsise::Input in) {
    nrlmsise::Flags fl = nrlFlags();
    for (int k = 0; k < 24; ++k) fl.switches[k] = 1;
    fl.switches[9] = -1;
    nrlmsise::Output out;
    nrlmsise::State st;
    nrlmsise::NrlTables T = nrlmsise::nrlTables(coeff);
    if (MODE == 1) nrlmsise::gtd7d(st, T, in, fl, out, nrlmsise::FastGlobe{nrl + (REC_NRL_GLOBE - REC_NRL), kGlobeTrig});
    else nrlmsise::gtd7d(st, T, in, fl, out, nrlmsise::GenericGlobe{});
    if (!isfinite(out.d[5]) || out.d[5] < 0.0) return nan("");
    return (out.d[5] * 1.0e9) * densityScale;
}

// Full model for one lane with a given globe evaluator (per-lane epochs: the
// lane's weather-slot row), kg/km^3.
template <class Globe>
__device__ __noinline__ double nrlDensityGlobe(double densityScale, const double* __restrict__ coeff, nrlmsise::Input in,
                                               Globe F) {
    nrlmsise::Flags fl = nrlFlags();
    nrlmsise::Output out;
    nrlmsise::State st;
    nrlmsise::NrlTables T = nrlmsise::nrlTables(coeff);
    nrlmsise::gtd7d(st, T, in, fl, out, F);
    if (!isfinite(out.d[5]) || out.d[5] < 0.0) return nan("");
    return (out.d[5] * 1.0e9) * densityScale;
}

// Total density (kg/km^3) from the species: gtd7d's sum and units.
__device__ inline double nrlRhoFromSpecies(const ForceModel& fm, double (*spec)[kTile], int lane) {
    // d[i] = 1e6 n_i; d5 = 1.66e-24 (4 He + 16 O + 28 N2 + 32 O2 + 40 Ar + H + 14 N + 16 anomalous O) / 1000
    const double He = 1.0E6 * spec[1][lane], O = 1.0E6 * spec[2][lane], N2 = 1.0E6 * spec[0][lane], O2 = 1.0E6 * spec[3][lane];
    const double Ar = 1.0E6 * spec[4][lane], H = 1.0E6 * spec[5][lane], N = 1.0E6 * spec[6][lane], AO = 1.0E6 * spec[7][lane];
    double d5 = 1.66E-24 * (4.0 * He + 16.0 * O + 28.0 * N2 + 32.0 * O2 + 40.0 * Ar + H + 14.0 * N + 16.0 * AO);
    d5 = d5 / 1000;
    if (!isfinite(d5) || d5 < 0.0) return nan("");
    return (d5 * 1.0e9) * fm.densityScale;
}

// Drag of one lane with the serial full model (diagnostic probe).
template <int MODE, class Rec>
__device__ void dragAccel(const ForceModel& fm, const Rec& rec, const double* nrl, const double* __restrict__ coeff, double px,
                          double py, double pz, double vx, double vy, double vz, double bcd, double& ax, double& ay,
                          double& az) {
    const NrlLane L = nrlLane(fm, rec, nrl, px, py, pz);
    double rho = 0.0;
    if (L.state == 1 || L.state == 2) rho = nrlDensity<MODE>(fm.densityScale, nrl, coeff, L.in);
    else if (L.state == 3) rho = nan("");
    const double om0 = rec(REC_OMEGA), om1 = rec(REC_OMEGA + 1), om2 = rec(REC_OMEGA + 2);
    const double rx = vx - (om1 * pz - om2 * py);
    const double ry = vy - (om2 * px - om0 * pz);
    const double rz = vz - (om0 * py - om1 * px);
    const double f = -0.5 * rho * bcd * sqrt(rx * rx + ry * ry + rz * rz);
    ax = f * rx;
    ay = f * ry;
    az = f * rz;
}

// Fully normalized Legendre recursion factors for degree <= 6, the same
// formulas as the reference's legendre() (HpopCorrections.h:121-145):
// kTideK = sqrt((2n+1)/((n-m)(n+m))), kTideA = sqrt(2n-1),
// kTideB = sqrt((n+m-1)(n-m-1)/(2n-3)) at [n][m] for n >= m+2;
// kTideD[n] = sqrt((2n+1)/(2n)) (sectoral), kTideS[m] = sqrt(2m+3) (first sub-diagonal).
__constant__ double kTideK[7][7], kTideA[7][7], kTideB[7][7], kTideD[7], kTideS[7];

// Tide gradient (degree <= 6, combined solid + ocean coefficients) and
// Schwarzschild term, SI inside, km/s^2 out (HpopCorrections.h:339-376,
// EarthCorrections.h:57-69). The Legendre values are generated column by
// column (order m outer, degree n inner) and consumed at once, so nothing
// but the two previous degrees of one column is kept in registers.
template <class Rec>
__device__ void correctionsAccel(const ForceModel& fm, const Rec& rec, double px, double py, double pz, double vx, double vy,
                                 double vz, double& ax, double& ay, double& az) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const double mu = fm.muEarthSI;
    if (fm.tides) {
        double E[9];
        for (int k = 0; k < 9; ++k) E[k] = rec(REC_R + k);
        const double x = 1000.0 * px, y = 1000.0 * py, z = 1000.0 * pz;
        const double r0 = E[0] * x + E[1] * y + E[2] * z;
        const double r1 = E[3] * x + E[4] * y + E[5] * z;
        const double r2 = E[6] * x + E[7] * y + E[8] * z;
        const double d = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
        const double xy2 = r0 * r0 + r1 * r1, xy = sqrt(xy2);
        const double s = r2 / d, c = xy / d;
        const double cl = r0 / xy, sl = r1 / xy;
        const double ratio = fm.radiusSI / d;
        // b1_n = (-mu/d^2) (R/d)^n (n+1), b2_n = (mu/d) (R/d)^n
        double pw[7];
        pw[0] = 1.0;
        for (int n = 1; n <= 6; ++n) pw[n] = pw[n - 1] * ratio;
        double dUdr = 0.0, dUdlat = 0.0, dUdlon = 0.0;
        double pmm = 1.0, dpmm = 0.0, cm = 1.0, sm = 0.0;
        for (int m = 0; m <= 6; ++m) {
            if (m == 1) {
                pmm = sqrt(3.0) * c;
                dpmm = -sqrt(3.0) * s;
            } else if (m >= 2) {
                const double k = kTideD[m];
                const double np = k * c * pmm;
                dpmm = k * (c * dpmm - s * pmm);
                pmm = np;
            }
            if (m >= 1) {
                const double nc = cm * cl - sm * sl;
                sm = sm * cl + cm * sl;
                cm = nc;
            }
            double p2 = 0.0, dp2 = 0.0, p1 = pmm, dp1 = dpmm;
            for (int n = m; n <= 6; ++n) {
                double pn, dpn;
                if (n == m) {
                    pn = pmm;
                    dpn = dpmm;
                } else if (n == m + 1) {
                    const double k = kTideS[m];
                    pn = k * s * pmm;
                    dpn = k * (c * pmm + s * dpmm);
                } else {
                    const double k = kTideK[n][m], a = kTideA[n][m], b = kTideB[n][m];
                    pn = k * (a * s * p1 - b * p2);
                    dpn = k * (a * s * dp1 + a * c * p1 - b * dp2);
                }
                if (n > m) {
                    p2 = p1;
                    dp2 = dp1;
                    p1 = pn;
                    dp1 = dpn;
                }
                const double Cnm = rec(REC_TIDE_C + n * (n + 1) / 2 + m), Snm = rec(REC_TIDE_S + n * (n + 1) / 2 + m);
                const double value = Cnm * cm + Snm * sm;
                const double b1 = (-mu / (d * d)) * pw[n] * (n + 1);
                const double b2 = (mu / d) * pw[n];
                dUdr += pn * value * b1;
                dUdlat += dpn * value * b2;
                dUdlon += m * pn * (Snm * cm - Cnm * sm) * b2;
            }
        }
        const double q = (1.0 / d * dUdr - r2 / (d * d * xy) * dUdlat);
        const double bf0 = q * r0 - (1.0 / xy2 * dUdlon) * r1;
        const double bf1 = q * r1 + (1.0 / xy2 * dUdlon) * r0;
        const double bf2 = 1.0 / d * dUdr * r2 + xy / (d * d) * dUdlat;
        sx = E[0] * bf0 + E[3] * bf1 + E[6] * bf2;
        sy = E[1] * bf0 + E[4] * bf1 + E[7] * bf2;
        sz = E[2] * bf0 + E[5] * bf1 + E[8] * bf2;
    }
    if (fm.relativity) {
        const double x = 1000.0 * px, y = 1000.0 * py, z = 1000.0 * pz;
        const double u = 1000.0 * vx, v = 1000.0 * vy, w = 1000.0 * vz;
        const double r = sqrt(x * x + y * y + z * z);
        const double vv = u * u + v * v + w * w;
        const double rv = x * u + y * v + z * w;
        const double f = mu / (fm.clightSI * fm.clightSI * r * r * r);
        const double g = 4.0 * mu / r - vv;
        sx += f * (g * x + 4.0 * rv * u);
        sy += f * (g * y + 4.0 * rv * v);
        sz += f * (g * z + 4.0 * rv * w);
    }
    ax = 1.0e-3 * sx;
    ay = 1.0e-3 * sy;
    az = 1.0e-3 * sz;
}

// Body slots [0, N) of the point-gravity sum, added in config order; the
// unconfigured ones are evaluated and discarded: no branch between the bodies,
// so their square roots and divisions overlap. The fused multiply-adds are
// written out, so every kernel that inlines this forms the same values.
template <int N, class Rec>
__device__ inline void bodySlots(const ForceModel& fm, const Rec& rec, double px, double py, double pz, double& ax, double& ay,
                                 double& az) {
#pragma unroll
    for (int b = 0; b < N; ++b) {
        const bool on = b < fm.nBodies && fm.bodyPointGravity[b] && fm.bodyNaif[b] != 399;
        const int q = REC_BODY + 3 * b;
        const double dx = px - rec(q), dy = py - rec(q + 1), dz = pz - rec(q + 2);
        const double s2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double rc = 1.0 / (s2 * sqrt(s2));
        const double g = fm.bodyGm[b];
        const double tx = fma(-g * dx, rc, ax), ty = fma(-g * dy, rc, ay), tz = fma(-g * dz, rc, az);
        ax = on ? tx : ax;
        ay = on ? ty : ay;
        az = on ? tz : az;
    }
}

// Record view whose rotation rows (REC_R ..) are the axes of body field k.
template <class Rec>
struct BodyAxesRec {
    const Rec& r;
    int off;
    __device__ double operator()(int f) const { return r(off + (f - REC_R)); }
};

// One body field of ForceModel::rot*, selected by a uniform index with
// constant-index reads (a dynamic index into the by-value kernel argument
// would copy it to scratch memory).
struct RotField {
    int body, sh, degree, order, normalized;
    double gm, radius, soi, j2;
    const double* terms;
    const int* orderStart;
    const double* diag;
};

template <class T>
__device__ inline T pick3(int k, const T (&v)[kMaxRotFields]) {
    static_assert(kMaxRotFields == 3, "one select per field");
    return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
}

__device__ inline RotField rotField(const ForceModel& fm, int k) {
    RotField f;
    f.body = pick3(k, fm.rotBody);
    f.sh = pick3(k, fm.rotSh);
    f.degree = pick3(k, fm.rotDegree);
    f.order = pick3(k, fm.rotOrder);
    f.normalized = pick3(k, fm.rotNormalized);
    f.gm = pick3(k, fm.rotGm);
    f.radius = pick3(k, fm.rotRadius);
    f.soi = pick3(k, fm.rotSoi);
    f.j2 = pick3(k, fm.rotJ2);
    f.terms = pick3(k, fm.rotTerms);
    f.orderStart = pick3(k, fm.rotOrderStart);
    f.diag = pick3(k, fm.rotDiag);
    return f;
}

// Field k of the bodies other than the Earth (SphericalHarmonics.cu:90-144,
// J2.h:75-90): evaluated at the position relative to the body in its own axes
// (record rows REC_BROT + 9 k) and rotated back; zero beyond the body's sphere
// of influence. Spherical harmonics run all orders on the calling wave.
template <class Rec>
__device__ inline void rotFieldAccel(const RotField f, int k, const Rec& rec, double px, double py, double pz, double& ax,
                                     double& ay, double& az) {
    {
        const int q = REC_BODY + 3 * f.body;
        const double dx = px - rec(q), dy = py - rec(q + 1), dz = pz - rec(q + 2);
        if (!(sqrt(dx * dx + dy * dy + dz * dz) <= f.soi)) return;
        const BodyAxesRec<Rec> br{rec, REC_BROT + 9 * k};
        double bx, by, bz;
        if (f.sh) {
            const ShGeom g = shGeometry(br, dx, dy, dz, f.radius);
            double r, p, l;
            shOrderBlock(g, 0, f.order + 1, f.degree, f.terms, nullptr, -1, 0, f.orderStart, f.diag, f.normalized != 0, r,
                         p, l);
            shToInertial(br, g, r, p, l, f.gm, bx, by, bz);
        } else {
            // J2 (Wakker eq. 20.6) in the body's axes
            double E[9];
#pragma unroll
            for (int j = 0; j < 9; ++j) E[j] = br(REC_R + j);
            const double fx = E[0] * dx + E[1] * dy + E[2] * dz;
            const double fy = E[3] * dx + E[4] * dy + E[5] * dz;
            const double fz = E[6] * dx + E[7] * dy + E[8] * dz;
            const double invr = 1.0 / sqrt(fx * fx + fy * fy + fz * fz);
            const double invr2 = invr * invr;
            const double invr5 = invr2 * invr2 * invr;
            const double v = 1.0 - 5.0 * (fz * fz) * invr2;
            const double c = -1.5 * f.gm * f.j2 * invr5 * f.radius * f.radius;
            const double cx = c * fx * v, cy = c * fy * v, cz = c * fz * (v + 2.0);
            bx = E[0] * cx + E[3] * cy + E[6] * cz;
            by = E[1] * cx + E[4] * cy + E[7] * cz;
            bz = E[2] * cx + E[5] * cy + E[8] * cz;
        }
        ax += bx;
        ay += by;
        az += bz;
    }
}

// Central and third-body point gravity (PointGravity.h:16-81), then (ROT) the
// fields of the bodies other than the Earth; the indirect third-body term is an
// epoch-only sum taken from the record. ROT = false leaves the field code out
// of the kernel (the launchers choose it when no such field is configured).
template <bool ROT = true, class Rec>
__device__ void pointGravity(const ForceModel& fm, const Rec& rec, double px, double py, double pz, double& ax, double& ay,
                             double& az) {
    ax = ay = az = 0.0;
    if (fm.earthPointGravity) {
        const double s2 = fma(pz, pz, fma(py, py, px * px));
        const double rc = 1.0 / (s2 * sqrt(s2));
        ax = (-fm.gmEarth * px) * rc;
        ay = (-fm.gmEarth * py) * rc;
        az = (-fm.gmEarth * pz) * rc;
    }
    // a few bodies (Earth, Sun, Moon, one more) or every slot the record holds
    if (fm.nBodies <= 4) bodySlots<4>(fm, rec, px, py, pz, ax, ay, az);
    else bodySlots<kMaxBodies>(fm, rec, px, py, pz, ax, ay, az);
    if constexpr (ROT)
        for (int k = 0; k < fm.nRot; ++k) rotFieldAccel(rotField(fm, k), k, rec, px, py, pz, ax, ay, az);
    ax += rec(REC_INDIRECT + 0);
    ay += rec(REC_INDIRECT + 1);
    az += rec(REC_INDIRECT + 2);
}

// J2-only gravity (J2.h:17-92, Wakker eq. 20.6): evaluated in ITRF axes and
// rotated back; zero beyond the sphere of influence.
template <class Rec>
__device__ inline void j2Accel(const ForceModel& fm, const Rec& rec, double px, double py, double pz, double& ax, double& ay,
                               double& az) {
    double E[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) E[k] = rec(REC_R + k);
    const double fx = E[0] * px + E[1] * py + E[2] * pz;
    const double fy = E[3] * px + E[4] * py + E[5] * pz;
    const double fz = E[6] * px + E[7] * py + E[8] * pz;
    const double invr = 1.0 / sqrt(fx * fx + fy * fy + fz * fz);
    const double invr2 = invr * invr;
    const double invr5 = invr2 * invr2 * invr;
    const double v = 1.0 - 5.0 * (fz * fz) * invr2;
    const double c = -1.5 * fm.j2Gm * fm.j2Coef * invr5 * fm.j2Radius * fm.j2Radius;
    const double bx = c * fx * v, by = c * fy * v, bz = c * fz * (v + 2.0);
    const bool in = !(sqrt(px * px + py * py + pz * pz) > fm.j2Soi);
    ax = in ? E[0] * bx + E[3] * by + E[6] * bz : 0.0;
    ay = in ? E[1] * bx + E[4] * by + E[7] * bz : 0.0;
    az = in ? E[2] * bx + E[5] * by + E[8] * bz : 0.0;
}

// Geodetic altitude above the oblate Earth (Geodetic.h:40-72): fixed-point
// iteration on the polar coordinate, at most kGeodeticIter steps (it converges
// in 2-4); ||r|| - R_eq when the flattening is 0.
constexpr int kGeodeticIter = 32;

__device__ inline double geodeticAltitude(double px, double py, double pz, double k0, double k1, double k2, double req,
                                          double f) {
    if (f == 0.0) return sqrt(px * px + py * py + pz * pz) - req;
    const double r2 = px * px + py * py + pz * pz;
    const double z = px * k0 + py * k1 + pz * k2;
    const double xy = r2 - z * z;
    const double e2 = (2.0 - f) * f;
    double d0 = 0.0, d = 1e10, sfq = 0.0, zd = z;
    for (int it = 0; it < kGeodeticIter && fabs(d - d0) > 1e-9; ++it) {
        d0 = d;
        const double z2 = zd * zd;
        sfq = z2 / (xy + z2);
        d = req * e2 * sqrt(sfq / (1.0 - e2 * sfq));
        zd = z + copysign(d, z);
    }
    return sqrt(xy + zd * zd) - req / sqrt(1.0 - e2 * sfq);
}

// Drag in a piecewise exponential atmosphere (PrivateAccelerations.h:167-202,
// PiecewiseExponentialAtmosphere.h:100-170): the band is the last one whose
// base h0 lies at or below the geodetic altitude (band 0 below them all), the
// density rho0 exp(-(h - h0)/H) up to the band's cutoff, zero above it.
template <class Rec>
__device__ inline void expDragAccel(const ForceModel& fm, const Rec& rec, double px, double py, double pz, double vx,
                                    double vy, double vz, double bcd, double& dx, double& dy, double& dz) {
    const double h = geodeticAltitude(px, py, pz, rec(REC_R + 6), rec(REC_R + 7), rec(REC_R + 8), fm.earthRadius,
                                      fm.flattening);
    // bands are sorted by h0, so the last band with h0 <= h is the count of them
    double rho0 = fm.expSeg[0][0], h0 = fm.expSeg[0][1], sh = fm.expSeg[0][2], cut = fm.expSeg[0][3];
#pragma unroll
    for (int s0 = 0; s0 < kMaxExpSegments; s0 += 8) {
        if (s0 >= fm.nExpSeg) break;
#pragma unroll
        for (int s = (s0 == 0 ? 1 : s0); s < s0 + 8; ++s) {
            if (s < fm.nExpSeg && fm.expSeg[s][1] <= h) {
                rho0 = fm.expSeg[s][0];
                h0 = fm.expSeg[s][1];
                sh = fm.expSeg[s][2];
                cut = fm.expSeg[s][3];
            }
        }
    }
    dx = dy = dz = 0.0;
    if (h <= cut) {
        const double rho = rho0 * exp(-(h - h0) / sh);
        const double om0 = rec(REC_OMEGA), om1 = rec(REC_OMEGA + 1), om2 = rec(REC_OMEGA + 2);
        const double rx = vx - (om1 * pz - om2 * py);
        const double ry = vy - (om2 * px - om0 * pz);
        const double rz = vz - (om0 * py - om1 * px);
        const double f = -0.5 * rho * bcd * sqrt(rx * rx + ry * ry + rz * rz);
        dx = f * rx;
        dy = f * ry;
        dz = f * rz;
    }
}

// Which wave carries which extra force family for a W-wave tile.
__device__ constexpr int corrWave(int W) { return W >= 3 ? W - 2 : W - 1; }

__constant__ double kRkC[12] = {2. / 27, 1. / 9, 1. / 6, 5. / 12, 1. / 2, 5. / 6, 1. / 6, 2. / 3, 1. / 3, 1.0, 0, 1.0};

// One stage's force evaluation for the tile: SH order block of this wave,
// the extra families assigned to this wave, then (wave 0) the sum.
template <int W, int MODE, bool DRAG, class Rec>
__device__ inline void stageForces(const StepArgs& a, const ForceModel& fm, const Rec& rec, const double* nrl,
                                   const NrlSlot& slot, int wave,
                                   int lane, int m0, int m1, double ams, double bcd, double (*xch)[kTile],
                                   double (*part)[3][kTile], double (*extra)[3][kTile], double (*spec)[kTile],
                                   double out[3]) {
    const double px = xch[0][lane], py = xch[1][lane], pz = xch[2][lane];
    ShGeom g;
    if (fm.hasSh) {
        double sr = 0.0, sp = 0.0, sl = 0.0;
        g = shGeometry(rec, px, py, pz, fm.shRadius);
        if (m0 < m1)
            shOrderBlock(g, m0, m1, fm.shDegree, a.shTerms, a.shTerms32, a.fp64Degree, a.fp64Order, a.shOrderStart, a.shDiag, fm.shNormalized, sr,
                         sp, sl);
        part[wave][0][lane] = sr;
        part[wave][1][lane] = sp;
        part[wave][2][lane] = sl;
    }
    // NRLMSISE-00: every wave evaluates its share of the species for the
    // lanes below the cutoff; wave 0 assembles the density after the barrier.
    NrlLane nl;
    nl.state = 0;
    if (DRAG && fm.drag) {
        nl = nrlLane(fm, rec, nrl, px, py, pz);
        if (__ballot(nl.state == 1) != 0ull) nrlSpecies<MODE>(nl, nrl, slot, a.nrlCoeff, wave, W, spec, lane);
    }
    // Without a field the families take waves of their own before the barrier
    // (point masses and J2 on the last wave, exponential drag on wave 1 from
    // four waves on), so wave 0 only sums; the values are those of one wave.
    const bool spread = W >= 2 && !fm.hasSh;
    if (spread && wave == W - 1) {
        double gx, gy, gz;
        pointGravity(fm, rec, px, py, pz, gx, gy, gz);
        if (fm.hasJ2) {
            double jx, jy, jz;
            j2Accel(fm, rec, px, py, pz, jx, jy, jz);
            gx += jx;
            gy += jy;
            gz += jz;
        }
        part[0][0][lane] = gx;
        part[0][1][lane] = gy;
        part[0][2][lane] = gz;
    }
    if (spread && W >= 4 && fm.expDrag && wave == 1) {
        double dx, dy, dz;
        expDragAccel(fm, rec, px, py, pz, xch[3][lane], xch[4][lane], xch[5][lane], bcd, dx, dy, dz);
        extra[0][0][lane] = dx;
        extra[0][1][lane] = dy;
        extra[0][2][lane] = dz;
    }
    if (wave == corrWave(W)) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        if (fm.tides || fm.relativity) correctionsAccel(fm, rec, px, py, pz, xch[3][lane], xch[4][lane], xch[5][lane], ax, ay, az);
        if (fm.srp) {
            double bx, by, bz;
            srpAccel(fm, rec, px, py, pz, ams, bx, by, bz);
            ax += bx;
            ay += by;
            az += bz;
        }
        extra[1][0][lane] = ax;
        extra[1][1][lane] = ay;
        extra[1][2][lane] = az;
    }
    __syncthreads();
    if (wave == 0) {
        double ax, ay, az;
        if (spread) {
            ax = part[0][0][lane];
            ay = part[0][1][lane];
            az = part[0][2][lane];
        } else {
            pointGravity(fm, rec, px, py, pz, ax, ay, az);
        }
        if (fm.hasSh) {
            double r = 0.0, p = 0.0, l = 0.0;
            for (int w = 0; w < W; ++w) {
                r += part[w][0][lane];
                p += part[w][1][lane];
                l += part[w][2][lane];
            }
            double hx, hy, hz;
            shToInertial(rec, g, r, p, l, fm.shGm, hx, hy, hz);
            ax += hx;
            ay += hy;
            az += hz;
        } else if (fm.hasJ2 && !spread) {
            double hx, hy, hz;
            j2Accel(fm, rec, px, py, pz, hx, hy, hz);
            ax += hx;
            ay += hy;
            az += hz;
        }
        double dx = 0.0, dy = 0.0, dz = 0.0;
        if (fm.expDrag) {
            if (spread && W >= 4) {
                dx = extra[0][0][lane];
                dy = extra[0][1][lane];
                dz = extra[0][2][lane];
            } else {
                expDragAccel(fm, rec, px, py, pz, xch[3][lane], xch[4][lane], xch[5][lane], bcd, dx, dy, dz);
            }
        }
        if (DRAG && fm.drag && nl.state != 0) {
            double rho;
            if (nl.state == 1) rho = nrlRhoFromSpecies(fm, spec, lane);
            else if (nl.state == 2) rho = nrlDensity<MODE>(fm.densityScale, nrl, a.nrlCoeff, nl.in);
            else rho = nan("");
            // Drag.h:67-73 with v_rel = v - omega x r (Earth is the centre)
            const double vx = xch[3][lane], vy = xch[4][lane], vz = xch[5][lane];
            const double om0 = rec(REC_OMEGA), om1 = rec(REC_OMEGA + 1), om2 = rec(REC_OMEGA + 2);
            const double rx = vx - (om1 * pz - om2 * py);
            const double ry = vy - (om2 * px - om0 * pz);
            const double rz = vz - (om0 * py - om1 * px);
            const double f = -0.5 * rho * bcd * sqrt(rx * rx + ry * ry + rz * rz);
            dx = f * rx;
            dy = f * ry;
            dz = f * rz;
        }
        out[0] = ax + (extra[1][0][lane] + dx);
        out[1] = ay + (extra[1][1][lane] + dy);
        out[2] = az + (extra[1][2][lane] + dz);
    }
}

// DRAG = false: the kernel for force models without NRLMSISE-00 drag, which
// carries none of the drag code and so keeps a small register budget (more
// waves per SIMD, up to 16 waves per tile).
// Drag-free stage forces of the split path: acc = ax + (extra + 0) as in
// stageForces, and the parts ax (point masses + SH) for the drag kernel; the
// corrections wave leaves its sum (SRP, tides, relativity) in extra[1].
//
// FINAL = false (SH over several workgroups): only the first workgroup
// (`extras`) evaluates the other families, and wave 0 returns the raw sums:
// the workgroup's SH components (radial, latitude, longitude) in rpl, point
// gravity in ax and the corrections + SRP sum in ex.
template <int W, bool FINAL = true, bool ROT = true, class Rec>
__device__ inline void stageForcesParts(const StepArgs& a, const ForceModel& fm, const Rec& rec, int wave, int lane, int m0,
                                        int m1, double ams, double (*xch)[kTile], double (*part)[3][kTile],
                                        double (*extra)[3][kTile], double (*pg)[kTile], double acc[3], double ax[3],
                                        double ex[3], bool extras = true, double rpl[3] = nullptr) {
    static_assert(W >= 4, "the split kernel spreads the extra families over waves 1 .. 3");
    const double px = xch[0][lane], py = xch[1][lane], pz = xch[2][lane];
    // the families besides SH, each on a wave of its own before the barrier:
    // point gravity on wave W - 1, tides and relativity on W - 2, SRP on W - 3
    if (!extras) {
    } else if (wave == W - 1) {
        double gx, gy, gz;
        pointGravity<ROT>(fm, rec, px, py, pz, gx, gy, gz);
        pg[0][lane] = gx;
        pg[1][lane] = gy;
        pg[2][lane] = gz;
    } else if (wave == W - 2) {
        double cx = 0.0, cy = 0.0, cz = 0.0;
        if (fm.tides || fm.relativity) correctionsAccel(fm, rec, px, py, pz, xch[3][lane], xch[4][lane], xch[5][lane], cx, cy, cz);
        extra[1][0][lane] = cx;
        extra[1][1][lane] = cy;
        extra[1][2][lane] = cz;
    } else if (wave == W - 3) {
        double bx = 0.0, by = 0.0, bz = 0.0;
        if (fm.srp) srpAccel(fm, rec, px, py, pz, ams, bx, by, bz);
        extra[0][0][lane] = bx;
        extra[0][1][lane] = by;
        extra[0][2][lane] = bz;
    }
    ShGeom g;
    if (fm.hasSh) {
        double sr = 0.0, sp = 0.0, sl = 0.0;
        g = shGeometry(rec, px, py, pz, fm.shRadius);
        if (m0 < m1)
            shOrderBlock(g, m0, m1, fm.shDegree, a.shTerms, a.shTerms32, a.fp64Degree, a.fp64Order, a.shOrderStart, a.shDiag, fm.shNormalized, sr,
                         sp, sl);
        part[wave][0][lane] = sr;
        part[wave][1][lane] = sp;
        part[wave][2][lane] = sl;
    }
    __syncthreads();
    if (!FINAL) {
        if (wave == 0) {
            double r = 0.0, p = 0.0, l = 0.0;
            if (fm.hasSh)
                for (int w = 0; w < W; ++w) {
                    r += part[w][0][lane];
                    p += part[w][1][lane];
                    l += part[w][2][lane];
                }
            rpl[0] = r;
            rpl[1] = p;
            rpl[2] = l;
            if (extras) {
                ax[0] = pg[0][lane];
                ax[1] = pg[1][lane];
                ax[2] = pg[2][lane];
                double cx = extra[1][0][lane], cy = extra[1][1][lane], cz = extra[1][2][lane];
                if (fm.srp) {
                    cx += extra[0][0][lane];
                    cy += extra[0][1][lane];
                    cz += extra[0][2][lane];
                }
                ex[0] = cx;
                ex[1] = cy;
                ex[2] = cz;
            }
        }
        return;
    }
    if (wave == 0) {
        // the fused kernel's sums in its order: ax = point gravity + SH,
        // acc = ax + ((tides + relativity) + SRP)
        double hx = 0.0, hy = 0.0, hz = 0.0;
        double gx = pg[0][lane], gy = pg[1][lane], gz = pg[2][lane];
        if (fm.hasSh) {
            double r = 0.0, p = 0.0, l = 0.0;
            for (int w = 0; w < W; ++w) {
                r += part[w][0][lane];
                p += part[w][1][lane];
                l += part[w][2][lane];
            }
            shToInertial(rec, g, r, p, l, fm.shGm, hx, hy, hz);
            gx += hx;
            gy += hy;
            gz += hz;
        }
        double cx = extra[1][0][lane], cy = extra[1][1][lane], cz = extra[1][2][lane];
        if (fm.srp) {
            cx += extra[0][0][lane];
            cy += extra[0][1][lane];
            cz += extra[0][2][lane];
        }
        ax[0] = gx;
        ax[1] = gy;
        ax[2] = gz;
        ex[0] = cx;
        ex[1] = cy;
        ex[2] = cz;
        acc[0] = gx + (cx + 0.0);
        acc[1] = gy + (cy + 0.0);
        acc[2] = gz + (cz + 0.0);
    }
}

template <int W, bool GRID, bool DRAG>
__global__ void __launch_bounds__(64 * W) rkf78Kernel(StepArgs a, ForceModel fm) {
    __shared__ double kst[kStages][6][kTile];
    __shared__ double xch[7][kTile];          // staged position, velocity, stage epoch
    __shared__ double part[W][3][kTile];
    __shared__ double extra[2][3][kTile];
    __shared__ double spec[8][kTile];          // NRLMSISE-00 species densities
    __shared__ int anyActive;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = __builtin_amdgcn_readfirstlane(a.tileList ? a.tileList[blockIdx.x] : int(blockIdx.x));
    const int obj = tile * kTile + lane;
    const int nPad = a.nPad;

    double y[6];
    double dt = 0.0, cur = 0.0;
    uint8_t st = ST_PAD;
    if (wave == 0) {
        for (int d = 0; d < 6; ++d) y[d] = a.y[d * nPad + obj];
        st = a.status[obj];
        dt = a.nextDt[obj];
        cur = a.cur[obj];
    } else {
        for (int d = 0; d < 6; ++d) y[d] = 0.0;
    }
    const double ams = a.ams[obj];
    const double bcd = a.bcd[obj];
    long long completed = 0;
    const long long recBase = GRID ? 0 : a.tileRecord[tile];
    const double* gridBase = nullptr;
    double gridT0 = 0.0;
    int gridRows = 0;
    if (GRID) {
        gridBase = a.records + a.slotRecord[obj] * REC_SIZE;
        gridT0 = a.slotT0[obj];
        gridRows = a.slotRows[obj];
    }
    const int m0 = a.blockStart[wave], m1 = a.blockStart[wave + 1];
    EpochArgs wa{};
    wa.weather = a.weather;
    wa.wDays = a.wDays;
    wa.wStart = a.wStart;

    for (int step = 0; step < a.steps; ++step) {
        if (wave == 0) {
            const bool active = (st & (ST_TERMINATED | ST_PAD)) == 0;
            const unsigned long long ballot = __ballot(active);
            if (lane == 0) anyActive = ballot != 0ull;
        }
        __syncthreads();
        if (!anyActive) break;
        const double* recStep = GRID ? a.records : a.records + (recBase + (long long)(a.stepBase + step) * kStages) * REC_SIZE;
        for (int s = 0; s < kStages; ++s) {
            if (wave == 0) {
                double ys[6];
                for (int d = 0; d < 6; ++d) ys[d] = y[d];
                if (s >= 1) {
                    for (int j = 0; j < s; ++j) {
                        const double aij = kRkA[s - 1][j];
                        if (aij != 0.0) {
                            const double coef = dt * aij;
                            for (int d = 0; d < 6; ++d) ys[d] += coef * kst[j][d][lane];
                        }
                    }
                }
                for (int d = 0; d < 6; ++d) xch[d][lane] = ys[d];
                double ts = cur;
                if (s >= 1) ts += dt * kRkC[s - 1];
                xch[6][lane] = ts;
            }
            __syncthreads();
            double acc[3];
            if (GRID) {
                const double ts = xch[6][lane];
                const double th = a.leapCount > 0 ? epochdev::leapTheta(a.leaps, a.leapCount, ts) : 0.0;
                const GridRecL rec{gridRec(gridBase, gridRows, gridT0, a.gridInvH, ts), cos(th), sin(th)};
                double nrl[13];
                NrlSlot slot{a.nrlSlots, 0.0, 0.0};
                if (DRAG && fm.drag) {
                    // this lane's NRLMSISE-00 time inputs and weather-slot row
                    const double mjd = epochdev::etToMjdUtcNrl(ts);
                    int doy, record, wslot;
                    double sec;
                    epochdev::doySecOf(mjd, doy, sec);
                    nrl[REC_NRL_DOY - REC_NRL] = double(doy);
                    nrl[REC_NRL_SEC - REC_NRL] = sec;
                    nrl[REC_NRL_GAST - REC_NRL] = epochdev::gastHours(mjd);
                    const bool ok = epochdev::weatherSlotOf(wa, mjd, record, wslot) && a.nrlSlots;
                    const double* row = a.nrlSlots + (ok ? size_t(record * 8 + wslot) * nrlmsise::SLOT_SIZE : 0);
                    for (int k = 0; k < 9; ++k) nrl[REC_NRL_F107 - REC_NRL + k] = ok ? row[nrlmsise::SLOT_F107 + k] : nan("");
                    nrl[REC_NRL_VALID - REC_NRL] = ok ? 1.0 : 0.0;
                    const double sr = 7.2722E-5;
                    slot.row = row;
                    slot.cs = cos(sr * sec);
                    slot.ss = sin(sr * sec);
                }
                stageForces<W, 2, DRAG>(a, fm, rec, nrl, slot, wave, lane, m0, m1, ams, bcd, xch, part, extra, spec, acc);
            } else {
                const ExactRec rec{recStep + s * REC_SIZE};
                const NrlSlot noSlot{nullptr, 0.0, 0.0};
                stageForces<W, 1, DRAG>(a, fm, rec, recStep + s * REC_SIZE + REC_NRL, noSlot, wave, lane, m0, m1, ams, bcd, xch, part,
                                  extra, spec, acc);
            }
            if (wave == 0) {
                kst[s][0][lane] = xch[3][lane];
                kst[s][1][lane] = xch[4][lane];
                kst[s][2][lane] = xch[5][lane];
                kst[s][3][lane] = acc[0];
                kst[s][4][lane] = acc[1];
                kst[s][5][lane] = acc[2];
            }
        }
        if (wave == 0) {
            const bool active = (st & (ST_TERMINATED | ST_PAD)) == 0;
            if (active) {
                for (int s = 0; s < kStages; ++s) {
                    const double bs = kRkB[s];
                    if (bs != 0.0) {
                        const double cb = dt * bs;
                        for (int d = 0; d < 6; ++d) y[d] += cb * kst[s][d][lane];
                    }
                }
                // FixedStep.h:30-37
                cur = cur + dt;
                const double remaining = a.end[obj] - cur;
                const bool finished = remaining <= 0.0;
                st = finished ? (ST_TERMINATED | ST_EOS) : 0;
                dt = finished ? 0.0 : fmin(a.step, remaining);
                ++completed;
            }
        }
    }
    if (wave == 0) {
        for (int d = 0; d < 6; ++d) a.y[d * nPad + obj] = y[d];
        a.status[obj] = st;
        a.nextDt[obj] = dt;
        a.cur[obj] = cur;
        // object-steps completed in this launch, one atomic per wave
        unsigned long long c = (unsigned long long)completed;
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
        if (lane == 0 && a.counters) atomicAdd((unsigned long long*)a.counters, c);
    }
}

// ---------------------------------------------------------------------------
// Split stages (shared-epoch batches with drag). The fused kernel above holds
// the NRLMSISE-00 path, which needs ~500 registers per lane, so every tile runs
// one wave per SIMD and the SH stream's scalar loads go unhidden. Here each
// stage is two launches: a drag-free kernel (SH, point masses, SRP, tides,
// relativity; two waves per SIMD at eight waves per tile) and a drag kernel for
// the tiles with lanes below the cutoff, with the k-stages in HBM; a third
// kernel completes the step. The arithmetic and its order are those of the
// fused kernel, so both give the same bits.
namespace {

__device__ inline size_t ksAt(int s, int d, int nPad, int obj) { return (size_t(s) * 6 + d) * size_t(nPad) + obj; }

// Step index of the exact records of the current stage (uniform).
__device__ inline long long splitStep(const SplitArgs& sp) {
    return sp.stepDev ? (long long)__builtin_amdgcn_readfirstlane(*sp.stepDev) : (long long)sp.step;
}

// Component d of the stage state y_s = y + sum_j (dt a_sj) k_j, in the fused
// kernel's arithmetic and order (so every kernel that forms it gets the same
// bits). Only the k-stages of nonzero a_sj with j < s are read (55 of the 78
// lower-triangle slots over a step); the loads are issued before the sums.
__device__ __attribute__((always_inline)) inline double stageComponent(const StepArgs& a, const SplitArgs& sp, int s, int obj,
                                                                       int d, double dt) {
    const int nPad = a.nPad;
    const double* arow = kRkA[s >= 1 ? s - 1 : 0];
    // the row's coefficients first (one batch of scalar loads), so that the
    // k-stage loads below issue back to back instead of one per scalar round trip
    double ar[kStages - 1];
#pragma unroll
    for (int j = 0; j < kStages - 1; ++j) ar[j] = arow[j];
    double k[kStages - 1];
#pragma unroll
    for (int j = 0; j < kStages - 1; ++j) {
        k[j] = 0.0;
        if (j < s && ar[j] != 0.0) k[j] = sp.ks[ksAt(j, d, nPad, obj)];
    }
    double y = a.y[d * nPad + obj];
#pragma unroll
    for (int j = 0; j < kStages - 1; ++j) {
        const double aij = ar[j];
        const double coef = dt * aij;
        const double v = y + coef * k[j];
        y = (j < s && aij != 0.0) ? v : y;
    }
    return y;
}

// Stage s of the split path for one tile: the stage state, then the forces
// without drag with the SH order blocks over W waves; k_s holds them until the
// finish kernel adds the drag.
// Stage state of a split SH workgroup: y_s = y + sum_j (dt a_sj) k_j and
// t_s = t + dt c_s (as the fused kernel), one component per wave, into xch;
// with `write` also into the velocity k-stage slots and with `writeYs` into ys
// (the drag kernels' copy of the stage state). Per-lane epochs:
// the waves from 6 on (wave 6 % W when there are fewer) interpolate every
// record field below the NRLMSISE-00 block at the lane's stage epoch into recF.
template <int W, bool GRID>
__device__ inline void splitStageState(const StepArgs& a, const SplitArgs& sp, int wave, int lane, int obj, bool write,
                                       bool writeYs, double (*xch)[kTile], double (*recF)[kTile]) {
    const int nPad = a.nPad;
    const int s = sp.stage;
    for (int d = wave; d < 7; d += W) {
        const double dt = a.nextDt[obj];
        double v;
        if (d < 6) {
            v = stageComponent(a, sp, s, obj, d, dt);
            if (d >= 3 && write) sp.ks[ksAt(s, d - 3, nPad, obj)] = v;
        } else {
            v = a.cur[obj];
            if (s >= 1) v += dt * kRkC[s - 1];
        }
        xch[d][lane] = v;
        if (writeYs) sp.ys[size_t(d) * nPad + obj] = v;
    }
    if (GRID) {
        // the idle waves' threads walk (lane, field) items field-fastest, so one
        // load instruction reads consecutive fields of one lane's four rows (the
        // lanes' rows lie apart when their epochs differ)
        const int first = W > 6 ? 6 : 6 % W, stagers = W > 6 ? W - 6 : 1;
        if (wave >= first && (W > 6 || wave == first)) {
            // each thread forms its own lane's rows and weights; the items take
            // them from lane l by cross-lane reads
            double ts = a.cur[obj];
            if (s >= 1) ts += a.nextDt[obj] * kRkC[s - 1];
            const GridRec own = gridRec(a.records + a.slotRecord[obj] * REC_SIZE, a.slotRows[obj], a.slotT0[obj], a.gridInvH, ts);
            const long long ownOff = (long long)(own.p - a.records);
            // leap seconds: the rows' rotation is smooth; this lane's is R3(theta) of it
            double lc = 1.0, ls = 0.0;
            if (a.leapCount > 0) {
                const double th = epochdev::leapTheta(a.leaps, a.leapCount, ts);
                lc = cos(th);
                ls = sin(th);
            }
            for (int item = (wave - first) * kTile + lane; item < kTile * kRecStaged; item += stagers * kTile) {
                const int l = item / kRecStaged, f = item - l * kRecStaged;
                GridRec g;
                g.p = a.records + __shfl(ownOff, l);
                g.w0 = __shfl(own.w0, l);
                g.w1 = __shfl(own.w1, l);
                g.w2 = __shfl(own.w2, l);
                g.w3 = __shfl(own.w3, l);
                double v = g(f);
                if (a.leapCount > 0) {
                    const double c = __shfl(lc, l), sn = __shfl(ls, l);
                    if (f < 6) {
                        const double q = g(f < 3 ? f + 3 : f - 3);
                        v = f < 3 ? c * v + sn * q : c * v - sn * q;
                    }
                }
                recF[GRID ? f : 0][l] = v;
            }
        }
    }
}

template <int W, bool GRID, bool ROT = true>
__global__ void __launch_bounds__(64 * W) splitShKernel(StepArgs a, ForceModel fm, SplitArgs sp) {
    __shared__ double xch[7][kTile];
    __shared__ double recF[GRID ? kRecStaged : 1][kTile];   // per-lane epochs: staged record fields
    __shared__ double part[W][3][kTile];
    __shared__ double extra[2][3][kTile];
    __shared__ double pg[3][kTile];
    __shared__ int dragTile;   // a lane of the tile needs the drag kernel at this stage
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = __builtin_amdgcn_readfirstlane(a.tileList ? a.tileList[blockIdx.x] : int(blockIdx.x));
    const int obj = tile * kTile + lane;
    const int nPad = a.nPad;
    const int s = sp.stage;
    if (threadIdx.x == 0) dragTile = 0;
    splitStageState<W, GRID>(a, sp, wave, lane, obj, true, false, xch, recF);
    __syncthreads();
    const int m0 = a.blockStart[wave], m1 = a.blockStart[wave + 1];
    double acc[3], ax[3], ex[3];
    const double px = xch[0][lane], py = xch[1][lane], pz = xch[2][lane];
    if (GRID) {
        // per-lane epochs: records interpolated at the lane's stage epoch
        const double ts = xch[6][lane];
        const GridRecF rec{gridRec(a.records + a.slotRecord[obj] * REC_SIZE, a.slotRows[obj], a.slotT0[obj], a.gridInvH, ts), recF,
                           lane};
        if (fm.drag && wave == W - 4) {
            // the NRLMSISE-00 lane inputs of this stage for the drag kernel
            double nrl[13], sec;
            const int row = laneNrlInputs(a, ts, nrl, sec);
            const NrlLane L = nrlLane(fm, rec, nrl, px, py, pz);
            if (nrlLaneStoreTile(L, sp.lane, nPad, obj)) {
                nrlLaneStoreGrid(sec, L.in.doy, row, sp.lane, nPad, obj);
                if (sp.trig) nrlTrigStore(L, sp.trig, nPad, obj);
                if (lane == 0) dragTile = 1;
            }
        }
        stageForcesParts<W, true, ROT>(a, fm, rec, wave, lane, m0, m1, a.ams[obj], xch, part, extra, pg, acc, ax, ex);
    } else {
        const double* recp = a.records + (a.tileRecord[tile] + splitStep(sp) * kStages + s) * REC_SIZE;
        const ExactRec rec{recp};
        if (fm.drag && wave == W - 4) {
            const NrlLane L = nrlLane(fm, rec, recp + REC_NRL, px, py, pz);
            if (nrlLaneStoreTile(L, sp.lane, nPad, obj)) {
                if (sp.trig) nrlTrigStore(L, sp.trig, nPad, obj);
                if (lane == 0) dragTile = 1;
            }
        }
        stageForcesParts<W, true, ROT>(a, fm, rec, wave, lane, m0, m1, a.ams[obj], xch, part, extra, pg, acc, ax, ex);
    }
    if (wave == 0) {
        // k_s without drag; a tile with a lane below the cutoff also leaves the
        // stage state and the parts for the drag kernel (which adds the drag to
        // k_s), a tile without one leaves nothing else
        for (int d = 0; d < 3; ++d) sp.ks[ksAt(s, 3 + d, nPad, obj)] = acc[d];
        if (dragTile) {
            for (int d = 0; d < 7; ++d) sp.ys[size_t(d) * nPad + obj] = xch[d][lane];
            for (int d = 0; d < 3; ++d) {
                sp.parts[size_t(d) * nPad + obj] = ax[d];
                sp.parts[size_t(3 + d) * nPad + obj] = ex[d];
            }
        }
    }
}

// Wave 0 of an SH workgroup of a grouped tile: writes the group's partial sums
// (and, for group 0, point gravity and corrections + SRP); the last group to
// arrive at the tile's counter sums the partials in group order and writes
// k_s and the parts as the one-workgroup kernel does, then resets the counter.
// The fences order the partial stores before the count (release) and the count
// before the partial loads (acquire) across the chip's L2 caches.
template <class Rec>
__device__ inline void splitGroupFinish(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, const Rec& rec, int g,
                                        int tile, int obj, int lane, double px, double py, double pz, const double rpl[3],
                                        const double ax[3], const double ex[3]) {
    const int nPad = a.nPad, G = sp.groups;
    for (int c = 0; c < 3; ++c) sp.shPart[(size_t(g) * 3 + c) * nPad + obj] = rpl[c];
    if (g == 0)
        for (int c = 0; c < 3; ++c) {
            sp.famPart[size_t(c) * nPad + obj] = ax[c];
            sp.famPart[size_t(3 + c) * nPad + obj] = ex[c];
        }
    __threadfence();
    unsigned prev = 0;
    if (lane == 0) prev = atomicAdd(sp.tileCount + tile, 1u);
    prev = __shfl(prev, 0);
    if (prev != unsigned(G - 1)) return;
    __threadfence();
    double r = 0.0, p = 0.0, l = 0.0;
    for (int q = 0; q < G; ++q) {
        r += sp.shPart[(size_t(q) * 3 + 0) * nPad + obj];
        p += sp.shPart[(size_t(q) * 3 + 1) * nPad + obj];
        l += sp.shPart[(size_t(q) * 3 + 2) * nPad + obj];
    }
    double gx = sp.famPart[obj], gy = sp.famPart[size_t(nPad) + obj], gz = sp.famPart[2 * size_t(nPad) + obj];
    const double cx = sp.famPart[3 * size_t(nPad) + obj], cy = sp.famPart[4 * size_t(nPad) + obj],
                 cz = sp.famPart[5 * size_t(nPad) + obj];
    if (fm.hasSh) {
        double hx, hy, hz;
        const ShGeom geo = shGeometry(rec, px, py, pz, fm.shRadius);
        shToInertial(rec, geo, r, p, l, fm.shGm, hx, hy, hz);
        gx += hx;
        gy += hy;
        gz += hz;
    }
    const int s = sp.stage;
    const double acc[3] = {gx + (cx + 0.0), gy + (cy + 0.0), gz + (cz + 0.0)};
    const double axs[3] = {gx, gy, gz}, exs[3] = {cx, cy, cz};
    for (int d = 0; d < 3; ++d) {
        sp.ks[ksAt(s, 3 + d, nPad, obj)] = acc[d];
        sp.parts[size_t(d) * nPad + obj] = axs[d];
        sp.parts[size_t(3 + d) * nPad + obj] = exs[d];
    }
    if (lane == 0) sp.tileCount[tile] = 0u;
}

// SH of a tile over sp.groups workgroups when few tiles are active: block
// t * G + g is group g of tile t; its waves run the order blocks
// groupBlockStart[g W + w]; group 0 also carries the other families and the
// NRLMSISE-00 lane inputs; splitGroupFinish assembles k_s.
template <int W, bool GRID>
__global__ void __launch_bounds__(64 * W) splitShGroupKernel(StepArgs a, ForceModel fm, SplitArgs sp) {
    __shared__ double xch[7][kTile];
    __shared__ double recF[GRID ? kRecStaged : 1][kTile];
    __shared__ double part[W][3][kTile];
    __shared__ double extra[2][3][kTile];
    __shared__ double pg[3][kTile];
    const int G = sp.groups;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = __builtin_amdgcn_readfirstlane(int(blockIdx.x) % G);
    const int ti = __builtin_amdgcn_readfirstlane(int(blockIdx.x) / G);
    const int tile = __builtin_amdgcn_readfirstlane(a.tileList ? a.tileList[ti] : ti);
    const int obj = tile * kTile + lane;
    const int nPad = a.nPad;
    const int s = sp.stage;
    splitStageState<W, GRID>(a, sp, wave, lane, obj, g == 0, g == 0, xch, recF);
    __syncthreads();
    const int m0 = sp.groupBlockStart[g * W + wave], m1 = sp.groupBlockStart[g * W + wave + 1];
    double acc[3], ax[3] = {0.0, 0.0, 0.0}, ex[3] = {0.0, 0.0, 0.0}, rpl[3] = {0.0, 0.0, 0.0};
    const double px = xch[0][lane], py = xch[1][lane], pz = xch[2][lane];
    if (GRID) {
        const double ts = xch[6][lane];
        const GridRecF rec{gridRec(a.records + a.slotRecord[obj] * REC_SIZE, a.slotRows[obj], a.slotT0[obj], a.gridInvH, ts), recF,
                           lane};
        if (g == 0 && fm.drag && wave == W - 4) {
            double nrl[13], sec;
            const int row = laneNrlInputs(a, ts, nrl, sec);
            const NrlLane L = nrlLane(fm, rec, nrl, px, py, pz);
            nrlLaneStore(L, sp.lane, nPad, obj);
            nrlLaneStoreGrid(sec, L.in.doy, row, sp.lane, nPad, obj);
            if (sp.trig) nrlTrigStore(L, sp.trig, nPad, obj);
        }
        stageForcesParts<W, false>(a, fm, rec, wave, lane, m0, m1, a.ams[obj], xch, part, extra, pg, acc, ax, ex, g == 0, rpl);
        if (wave == 0) splitGroupFinish(a, fm, sp, rec, g, tile, obj, lane, px, py, pz, rpl, ax, ex);
    } else {
        const double* recp = a.records + (a.tileRecord[tile] + splitStep(sp) * kStages + s) * REC_SIZE;
        const ExactRec rec{recp};
        if (g == 0 && fm.drag && wave == W - 4) {
            const NrlLane L = nrlLane(fm, rec, recp + REC_NRL, px, py, pz);
            nrlLaneStore(L, sp.lane, nPad, obj);
            if (sp.trig) nrlTrigStore(L, sp.trig, nPad, obj);
        }
        stageForcesParts<W, false>(a, fm, rec, wave, lane, m0, m1, a.ams[obj], xch, part, extra, pg, acc, ax, ex, g == 0, rpl);
        if (wave == 0) splitGroupFinish(a, fm, sp, rec, g, tile, obj, lane, px, py, pz, rpl, ax, ex);
    }
}

// Drag of one stage for the tiles with lanes below the cutoff: the species split
// over the tile's four waves, then k_s = ax + (extra + drag) as the fused kernel.
__global__ void __launch_bounds__(256) splitDragKernel(StepArgs a, ForceModel fm, SplitArgs sp) {
    constexpr int W = 4;
    __shared__ double spec[8][kTile];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = int(blockIdx.x);
    const int obj = tile * kTile + lane;
    const int nPad = a.nPad;
    const int s = sp.stage;
    // the stage state is kept for tiles with a lane below the cutoff only
    if (!nrlTileActive(sp.lane, nPad, obj)) return;
    const double px = sp.ys[obj], py = sp.ys[size_t(nPad) + obj], pz = sp.ys[size_t(2) * nPad + obj];
    const double* recp = a.records + (a.tileRecord[tile] + splitStep(sp) * kStages + s) * REC_SIZE;
    const ExactRec rec{recp};
    const double* nrl = recp + REC_NRL;
    const NrlLane nl = nrlLane(fm, rec, nrl, px, py, pz);
    // the same ballot in every wave: the whole workgroup leaves together
    if (__ballot(nl.state != 0) == 0ull) return;
    const NrlSlot noSlot{nullptr, 0.0, 0.0};
    if (__ballot(nl.state == 1) != 0ull) nrlSpecies<1>(nl, nrl, noSlot, a.nrlCoeff, wave, W, spec, lane);
    __syncthreads();
    if (wave == 0 && nl.state != 0) {
        double rho;
        if (nl.state == 1) rho = nrlRhoFromSpecies(fm, spec, lane);
        else if (nl.state == 2) rho = nrlDensity<1>(fm.densityScale, nrl, a.nrlCoeff, nl.in);
        else rho = nan("");
        const double vx = sp.ys[size_t(3) * nPad + obj], vy = sp.ys[size_t(4) * nPad + obj], vz = sp.ys[size_t(5) * nPad + obj];
        const double om0 = rec(REC_OMEGA), om1 = rec(REC_OMEGA + 1), om2 = rec(REC_OMEGA + 2);
        const double rx = vx - (om1 * pz - om2 * py);
        const double ry = vy - (om2 * px - om0 * pz);
        const double rz = vz - (om0 * py - om1 * px);
        const double f = -0.5 * rho * a.bcd[obj] * sqrt(rx * rx + ry * ry + rz * rz);
        const double dd[3] = {f * rx, f * ry, f * rz};
        for (int d = 0; d < 3; ++d)
            sp.ks[ksAt(s, 3 + d, nPad, obj)] = sp.parts[size_t(d) * nPad + obj] + (sp.parts[size_t(3 + d) * nPad + obj] + dd[d]);
    }
}

// Drag spread over more CUs: workgroup (tile, group) of SPW waves evaluates
// SPW of the eight species for the tile's lanes below the cutoff and stores
// them in HBM; splitDragFinishKernel assembles the density and the drag.
template <int SPW>
__global__ void __launch_bounds__(64 * SPW) splitSpeciesKernel(StepArgs a, ForceModel fm, SplitArgs sp) {
    constexpr int G = 8 / SPW;
    __shared__ double spec[8][kTile];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = int(blockIdx.x) / G;
    const int group = int(blockIdx.x) % G;
    const int obj = tile * kTile + lane;
    const int nPad = a.nPad;
    const int s = sp.stage;
    const double* recp = a.records + (a.tileRecord[tile] + splitStep(sp) * kStages + s) * REC_SIZE;
    const double* nrl = recp + REC_NRL;
    const NrlLane nl = nrlLaneLoad(sp.lane, nPad, obj, nrl);
    if (__ballot(nl.state == 1) == 0ull) return;
    const NrlSlot noSlot{nullptr, 0.0, 0.0};
    const int i = group * SPW + wave;   // entry of kSpeciesOrder
    nrlSpecies<1>(nl, nrl, noSlot, a.nrlCoeff, i, 8, spec, lane);
    const int k = kSpeciesOrder[i];
    sp.spec[size_t(k) * nPad + obj] = spec[k][lane];
}

// Calls f on every field of the densu quantities a lane's species share.
template <class H, class F>
__device__ __attribute__((always_inline)) inline void densuSharedFields(H& h, F&& f) {
    f(h.re), f(h.za), f(h.zlb), f(h.s2), f(h.tinf), f(h.tlb), f(h.glb), f(h.xmm);
    f(h.zLane), f(h.zg2L), f(h.ttL), f(h.lnL), f(h.laneUpper);
    f(h.zg2A), f(h.ttA), f(h.lnA), f(h.z1), f(h.z2), f(h.t1), f(h.zgdif), f(h.glb1);
    for (int k = 0; k < 5; ++k) f(h.xs[k]);
    for (int k = 0; k < 5; ++k) f(h.ys[k]);
    for (int k = 0; k < 5; ++k) f(h.y2out[k]);
    f(h.pmm), f(h.emm), f(h.db28), f(h.b28);
}
constexpr int kDensuSharedFields = 40;
static_assert(sizeof(nrlmsise::DensuShared) == kDensuSharedFields * sizeof(double), "every densu field is exchanged");
// Fields of densuShareLane (zg2L, ttL, lnL, pmm, emm) in densuSharedFields order.
__device__ constexpr bool densuLaneField(int i) { return i == 9 || i == 10 || i == 11 || i == 36 || i == 37; }

// Globe policy serving precomputed values: globe7 of table j from row j
// (with the Ap function of pt, ps and pd[3] in rows 11-13 for the apt[0] that
// the serial calls leave) and glob7s of lower-thermosphere table j from row
// 14 + j of the workgroup's exchange array.
struct SharedGlobe {
    double (*v)[kTile];
    int lane;
    __device__ __attribute__((always_inline)) void basis(nrlmsise::State&, const nrlmsise::Input&, const nrlmsise::Flags&) const {}
    __device__ __attribute__((always_inline)) double operator()(nrlmsise::State& st, int j, const double* p,
                                                               const nrlmsise::Input&, const nrlmsise::Flags&) const {
        if (((nrlmsise::NrlCD)p)[51] != 0) {
            if (j == 0) st.apt[0] = v[11][lane];
            else if (j == 1) st.apt[0] = v[12][lane];
            else if (j == 5) st.apt[0] = v[13][lane];
        }
        return v[j][lane];
    }
    __device__ __attribute__((always_inline)) double lower(nrlmsise::State&, int j, const double*, const nrlmsise::Input&,
                                                          const nrlmsise::Flags&) const {
        return v[14 + j][lane];
    }
};

// Species of each wave of the drag tile kernel: N2 on wave 0, which forms the
// lower-thermosphere part of the shared densu quantities while wave 5 forms
// their lane part; the four species whose mixing corrections are longest
// (O, H, N, O2) on four different SIMDs (waves w and w + 4 share a SIMD);
// anomalous O beside O2.
__constant__ int kTileSpecies[8] = {0, 1, 4, 3, 2, 5, 6, 7};

// Drag of one tile and stage in one eight-wave workgroup. The waves first
// evaluate the globe7 tables of gts7's prefix and, when a lane is below
// 300 km, the five glob7s node tables (two tables per wave at most, from the
// basis trigonometry the SH kernel left); every wave forms the prefix; wave 0
// forms the lower-thermosphere densu quantities every species shares and wave
// 5 the lane ones while the others evaluate the remaining species tables;
// after the hand-over each wave runs its species (kTileSpecies[wave]) up to the
// factor that needs b28, which wave 0 forms and flags, then its tail; wave 0
// assembles the density and the drag, k_s = ax + (extra + drag), as the fused
// kernel. Every value is the one the serial evaluation forms.
template <bool GRID>
__global__ void __launch_bounds__(512) splitDragTileKernel(StepArgs a, ForceModel fm, SplitArgs sp) {
    __shared__ double gv[19][kTile];
    __shared__ double spec[8][kTile];
    __shared__ double hx[kDensuSharedFields + 2][kTile];   // the shared densu quantities and lower-boundary node
    __shared__ double hb[kTile];                            // b28 of wave 0
    __shared__ int b28Ready;                                // set by wave 0 once hb holds b28
    __shared__ double hot[nrlmsise::HOT_COUNT];             // the chains' coefficients, read from LDS
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile = __builtin_amdgcn_readfirstlane(a.tileList ? a.tileList[blockIdx.x] : int(blockIdx.x));
    const int obj = tile * kTile + lane;
    const int nPad = a.nPad;
    const int s = sp.stage;
    // the lane input rows in one batch of loads; a tile without a lane below the
    // cutoff has k_s complete (every wave leaves on the state row)
    const LaneRows rows = laneRows(sp.lane, nPad, obj);
    // the basis trigonometry the SH kernel formed beside the lane inputs
    nrlmsise::BasisTrig trig{};
    if (sp.trig) trig = nrlTrigLoad(sp.trig, nPad, obj);
    if (__ballot(rows.v[4] != 0.0) == 0ull) return;
    if (threadIdx.x == 0) b28Ready = 0;   // ordered before wave 0 sets it by the barriers below
    // the copy completes before the first barrier below, long before the prefix
    if (threadIdx.x < nrlmsise::HOT_COUNT) hot[threadIdx.x] = a.nrlCoeff[nrlmsise::COEFF_COUNT + threadIdx.x];
    // shared epochs: the tile's record; per-lane epochs: the lane's time inputs
    // and weather-slot row (the globe terms of its slot)
    const double* recp = GRID ? nullptr : a.records + (a.tileRecord[tile] + splitStep(sp) * kStages + s) * REC_SIZE;
    double nrlLocal[13];
    int row = 0;
    const double* nrl = GRID ? nrlLocal : recp + REC_NRL;
    const NrlLane nl = GRID ? nrlLaneFromGrid(rows, sp.lane, nPad, obj, a.nrlSlots, row, nrlLocal) : nrlLaneFrom(rows, nrl);
    using Globe = typename std::conditional<GRID, nrlmsise::SlotGlobe, nrlmsise::FastGlobe>::type;
    // the same lanes in every wave: the branch is uniform across the workgroup
    if (__ballot(nl.state == 1) != 0ull) {
        const nrlmsise::Input in0 = nl.in;
        const nrlmsise::Flags fl = nrlFlags();
        const nrlmsise::NrlTables T{a.nrlCoeff, hot};
        Globe F;
        if constexpr (GRID) {
            const double sr = 7.2722E-5;
            const double sec = nrl[REC_NRL_SEC - REC_NRL];
            F = nrlmsise::SlotGlobe{a.nrlSlots + size_t(row) * nrlmsise::SLOT_SIZE, kGlobeTrig, cos(sr * sec), sin(sr * sec)};
        } else {
            F = nrlmsise::FastGlobe{nrl + (REC_NRL_GLOBE - REC_NRL), kGlobeTrig};
        }
        const int k = kTileSpecies[wave];
        // globe7 table of this wave's species
        const int jk = k == 7 ? 10 : nrlmsise::kSpeciesPar[k].globe;
        // a lane below 300 km: the glob7s node tables (the same in every wave)
        const bool low = __ballot(in0.alt < 300.0) != 0ull;
        {
            // globe7 of the prefix tables pt, ps and pd[3] on waves 0-2; on waves
            // 3-7, with a lane below 300 km, first one of the five glob7s node
            // tables (each wave forms the apt[0] that the globe7 calls of pt, ps
            // and pd[3] leave, their Ap functions, itself), then the globe7 table
            // of pd[2] (N2) or of the wave's species (O, H, N, anomalous O)
            nrlmsise::State st;
            if (sp.trig) F.basisFrom(st, trig);
            else F.basis(st, in0, fl);
            if (wave < 3) {
                const int j = wave == 2 ? 5 : wave;
                double apt = 0.0;
                gv[j][lane] = F.value(st, j, nrlmsise::globeTable(T, j), in0, apt);
                gv[11 + wave][lane] = apt;
            } else {
                if (low) {
                    const double za = ((nrlmsise::NrlCD)(T.base + nrlmsise::OFF_PDL + 25 * 1))[15];
                    const double* t0 = nrlmsise::globeTable(T, 0);
                    const double* t1 = nrlmsise::globeTable(T, 1);
                    const double* t5 = nrlmsise::globeTable(T, 5);
                    double apt = 0.0;
                    if (in0.alt > za && ((nrlmsise::NrlCD)t0)[51] != 0) apt = F.apt(0, t0, in0);
                    if (in0.alt > 72.5 && ((nrlmsise::NrlCD)t1)[51] != 0) apt = F.apt(1, t1, in0);
                    if (((nrlmsise::NrlCD)t5)[51] != 0) apt = F.apt(5, t5, in0);
                    st.apt[0] = apt;
                    const int j = wave - 3;
                    gv[14 + j][lane] = F.lower(st, j, nrlmsise::lowerTable(T, j), in0, fl);
                }
                const int j = wave == 3 ? 4 : jk;
                double apt = 0.0;
                gv[j][lane] = F.value(st, j, nrlmsise::globeTable(T, j), in0, apt);
            }
            __syncthreads();
        }
        nrlmsise::Input in = in0;
        nrlmsise::State st;
        nrlmsise::glatf(in.g_lat, st.gsurf, st.re);
        const SharedGlobe G{gv, lane};
        nrlmsise::Pre pre;
        nrlmsise::gts7Prefix(st, T, in, fl, G, pre);
        nrlmsise::DensuShared h;
        if (wave == 0 || wave == 5) {
            // the densu quantities every species shares, formed once and handed
            // to every wave: wave 0 the lower-thermosphere part, wave 5 the lane
            // part, at once (db28 and b28 follow the hand-over)
            const bool lanePart = wave == 5;
            if (lanePart) {
                nrlmsise::gts7ShareLane(h, st, T, in, pre);
            } else {
                nrlmsise::gts7SpeciesShare<SharedGlobe, false, false, false>(h, true, st, T, in, fl, G, pre);
                hx[kDensuSharedFields][lane] = st.meso_tn1[0];
                hx[kDensuSharedFields + 1][lane] = st.meso_tgn1[0];
            }
            int i = 0;
            densuSharedFields(h, [&](auto& x) {
                if (densuLaneField(i) == lanePart) hx[i][lane] = double(x);
                ++i;
            });
        } else if (wave < 4) {
            // meanwhile waves 1-3 evaluate their species' globe7 tables
            nrlmsise::State sb;
            if (sp.trig) F.basisFrom(sb, trig);
            else F.basis(sb, in, fl);
            double apt = 0.0;
            gv[jk][lane] = F.value(sb, jk, nrlmsise::globeTable(T, jk), in, apt);
        }
        // each species' correction factors of the lane altitude alone, in the
        // time its wave waits for the shared quantities
        nrlmsise::SpeciesPre q{};
        if (k != 7) q = nrlmsise::gts7SpeciesPre(k, T, in, fl);
        __syncthreads();
        // every species up to its b28 factor while wave 0 forms b28 (before its
        // own N2 head), then the tails from b28: a wave with a tail waits for
        // wave 0's flag, not for the other species' heads
        nrlmsise::SpeciesHead head{};
        if (k != 7) {
            int i = 0;
            densuSharedFields(h, [&](auto& x) {
                using T = typename std::remove_reference<decltype(x)>::type;
                x = T(hx[i++][lane]);
            });
            st.meso_tn1[0] = hx[kDensuSharedFields][lane];
            st.meso_tgn1[0] = hx[kDensuSharedFields + 1][lane];
            if (wave == 0) {
                nrlmsise::gts7ShareDb28(h, true, st, T, in, fl, G, pre);
                nrlmsise::gts7ShareB28(h, st, T, in, pre);
                hb[lane] = h.b28;
                // the release orders the stores of hb before the flag
                __hip_atomic_store(&b28Ready, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            head = nrlmsise::gts7SpeciesHead(k, h, st, T, in, fl, G, pre, q);
            if (wave != 0 && __ballot(head.tail != 0) != 0ull) {
                // wave 0 of this workgroup sets the flag unconditionally above;
                // the wait is bounded all the same (about 8M cycles)
                for (int spin = 0; spin < (1 << 16) &&
                                   __hip_atomic_load(&b28Ready, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == 0;
                     ++spin)
                    __builtin_amdgcn_s_sleep(2);
            }
            h.b28 = hb[lane];
            spec[k][lane] = nrlmsise::gts7SpeciesTail(k, head, h, st, T, pre);
        } else {
            // the anomalous oxygen (its own densu profile) needs nothing else
            spec[k][lane] = nrlmsise::gts7Species(k, h, st, T, in, fl, G, pre);
        }
    }
    __syncthreads();
    if (wave != 0 || nl.state == 0) return;
    double rho;
    if (nl.state == 1) rho = nrlRhoFromSpecies(fm, spec, lane);
    else if (nl.state == 2) {
        // below 72.5 km: the serial model, with the lane's weather-slot globe terms
        // for per-lane epochs
        if constexpr (GRID) {
            const double sec = nrl[REC_NRL_SEC - REC_NRL];
            rho = nrlDensityGlobe(fm.densityScale, a.nrlCoeff, nl.in,
                                  nrlmsise::SlotGlobe{a.nrlSlots + size_t(row) * nrlmsise::SLOT_SIZE, kGlobeTrig,
                                                      cos(7.2722E-5 * sec), sin(7.2722E-5 * sec)});
        } else {
            rho = nrlDensity<1>(fm.densityScale, nrl, a.nrlCoeff, nl.in);
        }
    } else rho = nan("");
    const double px = sp.ys[obj], py = sp.ys[size_t(nPad) + obj], pz = sp.ys[size_t(2) * nPad + obj];
    const double vx = sp.ys[size_t(3) * nPad + obj], vy = sp.ys[size_t(4) * nPad + obj], vz = sp.ys[size_t(5) * nPad + obj];
    double om0, om1, om2;
    if constexpr (GRID) {
        const GridRec rec = gridRec(a.records + a.slotRecord[obj] * REC_SIZE, a.slotRows[obj], a.slotT0[obj], a.gridInvH,
                                    sp.ys[size_t(6) * nPad + obj]);
        om0 = rec(REC_OMEGA), om1 = rec(REC_OMEGA + 1), om2 = rec(REC_OMEGA + 2);
    } else {
        const ExactRec rec{recp};
        om0 = rec(REC_OMEGA), om1 = rec(REC_OMEGA + 1), om2 = rec(REC_OMEGA + 2);
    }
    const double rx = vx - (om1 * pz - om2 * py);
    const double ry = vy - (om2 * px - om0 * pz);
    const double rz = vz - (om0 * py - om1 * px);
    const double f = -0.5 * rho * a.bcd[obj] * sqrt(rx * rx + ry * ry + rz * rz);
    const double dd[3] = {f * rx, f * ry, f * rz};
    for (int d = 0; d < 3; ++d)
        sp.ks[ksAt(s, 3 + d, nPad, obj)] = sp.parts[size_t(d) * nPad + obj] + (sp.parts[size_t(3 + d) * nPad + obj] + dd[d]);
}

__global__ void __launch_bounds__(64) splitDragFinishKernel(StepArgs a, ForceModel fm, SplitArgs sp) {
    __shared__ double spec[8][kTile];
    const int lane = threadIdx.x & 63;
    const int tile = int(blockIdx.x);
    const int obj = tile * kTile + lane;
    const int nPad = a.nPad;
    const int s = sp.stage;
    const double* recp = a.records + (a.tileRecord[tile] + splitStep(sp) * kStages + s) * REC_SIZE;
    const ExactRec rec{recp};
    const double* nrl = recp + REC_NRL;
    const NrlLane nl = nrlLaneLoad(sp.lane, nPad, obj, nrl);
    if (nl.state == 0) return;
    const double px = sp.ys[obj], py = sp.ys[size_t(nPad) + obj], pz = sp.ys[size_t(2) * nPad + obj];
    double rho;
    if (nl.state == 1) {
        for (int k = 0; k < 8; ++k) spec[k][lane] = sp.spec[size_t(k) * nPad + obj];
        rho = nrlRhoFromSpecies(fm, spec, lane);
    } else if (nl.state == 2) {
        rho = nrlDensity<1>(fm.densityScale, nrl, a.nrlCoeff, nl.in);
    } else {
        rho = nan("");
    }
    const double vx = sp.ys[size_t(3) * nPad + obj], vy = sp.ys[size_t(4) * nPad + obj], vz = sp.ys[size_t(5) * nPad + obj];
    const double om0 = rec(REC_OMEGA), om1 = rec(REC_OMEGA + 1), om2 = rec(REC_OMEGA + 2);
    const double rx = vx - (om1 * pz - om2 * py);
    const double ry = vy - (om2 * px - om0 * pz);
    const double rz = vz - (om0 * py - om1 * px);
    const double f = -0.5 * rho * a.bcd[obj] * sqrt(rx * rx + ry * ry + rz * rz);
    const double dd[3] = {f * rx, f * ry, f * rz};
    for (int d = 0; d < 3; ++d)
        sp.ks[ksAt(s, 3 + d, nPad, obj)] = sp.parts[size_t(d) * nPad + obj] + (sp.parts[size_t(3 + d) * nPad + obj] + dd[d]);
}

// End of step: y += sum (dt b_s) k_s and the bookkeeping of FixedStep.h:30-37.
__global__ void __launch_bounds__(64) splitCombineKernel(StepArgs a, SplitArgs sp) {
    const int lane = threadIdx.x & 63;
    const int tile = __builtin_amdgcn_readfirstlane(a.tileList ? a.tileList[blockIdx.x] : int(blockIdx.x));
    const int obj = tile * kTile + lane;
    const int nPad = a.nPad;
    uint8_t st = a.status[obj];
    unsigned long long completed = 0;
    if ((st & (ST_TERMINATED | ST_PAD)) == 0) {
        double dt = a.nextDt[obj], cur = a.cur[obj];
        // y += sum_s (dt b_s) k_s over the nonzero b_s, in stage order; the
        // k-stage loads (of the nonzero b_s only) are issued before the sums
        double y[6];
        for (int d = 0; d < 6; ++d) {
            double k[kStages];
#pragma unroll
            for (int s = 0; s < kStages; ++s) {
                k[s] = 0.0;
                if ((kRkBNonZero >> s) & 1u) k[s] = sp.ks[ksAt(s, d, nPad, obj)];
            }
            double v = a.y[d * nPad + obj];
#pragma unroll
            for (int s = 0; s < kStages; ++s) {
                const double bs = kRkB[s];
                const double cb = dt * bs;
                const double w = v + cb * k[s];
                v = bs != 0.0 ? w : v;
            }
            y[d] = v;
        }
        cur = cur + dt;
        const double remaining = a.end[obj] - cur;
        const bool finished = remaining <= 0.0;
        st = finished ? (ST_TERMINATED | ST_EOS) : 0;
        dt = finished ? 0.0 : fmin(a.step, remaining);
        for (int d = 0; d < 6; ++d) a.y[d * nPad + obj] = y[d];
        a.status[obj] = st;
        a.nextDt[obj] = dt;
        a.cur[obj] = cur;
        completed = 1;
    }
    for (int off = 32; off > 0; off >>= 1) completed += __shfl_xor(completed, off);
    if (lane == 0 && a.counters && completed) atomicAdd((unsigned long long*)a.counters, completed);
    // the next step's records (no kernel of this step reads the index any more)
    if (sp.stepDev && blockIdx.x == 0 && threadIdx.x == 0) *sp.stepDev = *sp.stepDev + 1;
}

}  // namespace

__global__ void __launch_bounds__(64) forceProbeKernel(int n, const double* pos, const double* vel, const double* ams,
                                                       const double* bcd, const double* recp, int which, ForceModel fm,
                                                       const double* shTerms, const int* shOrderStart, const double* shDiag,
                                                       const double* nrlCoeff, double* acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = pos[i], py = pos[n + i], pz = pos[2 * n + i];
    const double vx = vel[i], vy = vel[n + i], vz = vel[2 * n + i];
    const ExactRec rec{recp};
    double ax = 0.0, ay = 0.0, az = 0.0, bx, by, bz;
    if (which < 0 || which == 0 || which == 2) {
        ForceModel f = fm;
        if (which == 0) f.nBodies = f.nRot = 0;   // central term only
        if (which == 2) f.earthPointGravity = 0;  // third bodies only
        pointGravity(f, rec, px, py, pz, bx, by, bz);
        if (which == 0) {
            bx -= rec(REC_INDIRECT);
            by -= rec(REC_INDIRECT + 1);
            bz -= rec(REC_INDIRECT + 2);
        }
        ax += bx, ay += by, az += bz;
    }
    if ((which < 0 || which == 1) && fm.hasSh) {
        const ShGeom g = shGeometry(rec, px, py, pz, fm.shRadius);
        double r, p, l;
        shOrderBlock(g, 0, fm.shOrder + 1, fm.shDegree, shTerms, nullptr, -1, 0, shOrderStart, shDiag, fm.shNormalized, r, p, l);
        shToInertial(rec, g, r, p, l, fm.shGm, bx, by, bz);
        ax += bx, ay += by, az += bz;
    } else if ((which < 0 || which == 1) && fm.hasJ2) {
        j2Accel(fm, rec, px, py, pz, bx, by, bz);
        ax += bx, ay += by, az += bz;
    }
    if ((which < 0 || which == 3) && fm.srp) {
        srpAccel(fm, rec, px, py, pz, ams[i], bx, by, bz);
        ax += bx, ay += by, az += bz;
    }
    if ((which < 0 || which == 4) && fm.drag) {
        dragAccel<1>(fm, rec, recp + REC_NRL, nrlCoeff, px, py, pz, vx, vy, vz, bcd[i], bx, by, bz);
        ax += bx, ay += by, az += bz;
    }
    if ((which < 0 || which == 4) && fm.expDrag) {
        expDragAccel(fm, rec, px, py, pz, vx, vy, vz, bcd[i], bx, by, bz);
        ax += bx, ay += by, az += bz;
    }
    if ((which < 0 || which == 5) && (fm.tides || fm.relativity)) {
        correctionsAccel(fm, rec, px, py, pz, vx, vy, vz, bx, by, bz);
        ax += bx, ay += by, az += bz;
    }
    acc[i] = ax;
    acc[n + i] = ay;
    acc[2 * n + i] = az;
}

__global__ void __launch_bounds__(64) nrlProbeKernel(int n, const double* in, int useApArray, const double* coeff, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* r = in + 16 * i;
    nrlmsise::Input x;
    x.doy = int(r[0]);
    x.sec = r[1];
    x.alt = r[2];
    x.g_lat = r[3];
    x.g_long = r[4];
    x.lst = r[5];
    x.f107A = r[6];
    x.f107 = r[7];
    x.ap = r[8];
    for (int k = 0; k < 7; ++k) x.apa[k] = r[9 + k];
    nrlmsise::Flags fl;
    for (int k = 0; k < 24; ++k) fl.switches[k] = 1;
    fl.switches[0] = 0;
    if (useApArray) fl.switches[9] = -1;
    nrlmsise::Output o;
    nrlmsise::State st;
    nrlmsise::NrlTables T = nrlmsise::nrlTables(coeff);
    nrlmsise::gtd7d(st, T, x, fl, o, nrlmsise::GenericGlobe{});
    for (int k = 0; k < 9; ++k) out[11 * i + k] = o.d[k];
    out[11 * i + 9] = o.t[0];
    out[11 * i + 10] = o.t[1];
}

}  // namespace

int uploadTideFactors() {
    double K[7][7] = {}, A[7][7] = {}, B[7][7] = {}, D[7] = {}, S[7] = {};
    for (int m = 0; m <= 6; ++m)
        for (int n = m + 2; n <= 6; ++n) {
            K[n][m] = sqrt((2.0 * n + 1) / ((n - m) * double(n + m)));
            A[n][m] = sqrt(2.0 * n - 1);
            B[n][m] = sqrt(((n + m - 1) * double(n - m - 1)) / (2.0 * n - 3));
        }
    for (int n = 2; n <= 6; ++n) D[n] = sqrt((2.0 * n + 1) / (2.0 * n));
    for (int m = 0; m <= 6; ++m) S[m] = sqrt(2.0 * m + 3);
    if (hipMemcpyToSymbol(HIP_SYMBOL(kTideK), K, sizeof K) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(kTideA), A, sizeof A) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(kTideB), B, sizeof B) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(kTideD), D, sizeof D) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(kTideS), S, sizeof S) != hipSuccess) return -1;
    return 0;
}

int uploadGlobeTrig(const double* nrlCoeffHost) {
    const double hr = 0.2618, dgtr = 1.74533E-2;
    double t[nrlmsise::GE_TABLES * nrlmsise::GT_COUNT];
    for (int j = 0; j < nrlmsise::GE_TABLES; ++j) {
        const double* p = nrlCoeffHost + (j == 0 ? nrlmsise::OFF_PT : (j == 1 ? nrlmsise::OFF_PS : nrlmsise::OFF_PD + 150 * (j - 2)));
        double* o = t + nrlmsise::GT_COUNT * j;
        o[nrlmsise::GT_C131] = cos(hr * p[131]);
        o[nrlmsise::GT_S131] = sin(hr * p[131]);
        o[nrlmsise::GT_C97] = cos(dgtr * p[97]);
        o[nrlmsise::GT_S97] = sin(dgtr * p[97]);
        o[nrlmsise::GT_C136] = cos(dgtr * p[136]);
        o[nrlmsise::GT_S136] = sin(dgtr * p[136]);
    }
    return hipMemcpyToSymbol(HIP_SYMBOL(kGlobeTrig), t, sizeof t) == hipSuccess ? 0 : -1;
}

namespace {
template <bool GRID, bool DRAG>
int launchShape(const StepArgs& a, const ForceModel& fm, int waves, hipStream_t stream) {
    const dim3 grid(a.nTiles);
    switch (waves) {
        case 1: hipLaunchKernelGGL((rkf78Kernel<1, GRID, DRAG>), grid, dim3(64), 0, stream, a, fm); break;
        case 2: hipLaunchKernelGGL((rkf78Kernel<2, GRID, DRAG>), grid, dim3(128), 0, stream, a, fm); break;
        case 4: hipLaunchKernelGGL((rkf78Kernel<4, GRID, DRAG>), grid, dim3(256), 0, stream, a, fm); break;
        case 8:
            if (DRAG) return -2;
            hipLaunchKernelGGL((rkf78Kernel<8, GRID, false>), grid, dim3(512), 0, stream, a, fm);
            break;
        case 16:
            if (DRAG) return -2;
            hipLaunchKernelGGL((rkf78Kernel<16, GRID, false>), grid, dim3(1024), 0, stream, a, fm);
            break;
        default: return -2;
    }
    return 0;
}
}  // namespace

static int launchSplitSh(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, int shWaves, hipStream_t stream);
static int launchSplitSpecies(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, hipStream_t stream);
static int launchSplitFinish(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, hipStream_t stream);

// One stage of the split path: the SH kernel, then the species kernel (one
// workgroup per tile and species group) and the finish kernel, or the
// four-wave drag kernel.
int launchSplitStage(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, int shWaves, hipStream_t stream) {
    int rc = launchSplitSh(a, fm, sp, shWaves, stream);
    if (rc != 0 || a.nTiles <= 0 || !fm.drag) return rc;
    if (sp.speciesPerGroup == 8) {
        if (a.grid) hipLaunchKernelGGL((splitDragTileKernel<true>), dim3(a.nTiles), dim3(512), 0, stream, a, fm, sp);
        else hipLaunchKernelGGL((splitDragTileKernel<false>), dim3(a.nTiles), dim3(512), 0, stream, a, fm, sp);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    // the per-species and four-wave drag kernels read shared-epoch records of
    // contiguous tiles
    if (a.grid || a.tileList) return -2;
    if (sp.speciesPerGroup != 0) rc = launchSplitSpecies(a, fm, sp, stream);
    if (rc == 0) rc = launchSplitFinish(a, fm, sp, stream);
    return rc;
}

static int launchSplitSh(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, int shWaves, hipStream_t stream) {
    if (a.nTiles <= 0) return 0;
    if (sp.groups > 1) {
        // several SH workgroups per tile (eight or sixteen waves each), counters
        // and partial buffers supplied by the caller
        if (sp.groups > kMaxGroups || !sp.shPart || !sp.famPart || !sp.tileCount) return -2;
        const dim3 grid(a.nTiles * sp.groups);
        switch (shWaves) {
            case 8:
                if (a.grid) hipLaunchKernelGGL((splitShGroupKernel<8, true>), grid, dim3(512), 0, stream, a, fm, sp);
                else hipLaunchKernelGGL((splitShGroupKernel<8, false>), grid, dim3(512), 0, stream, a, fm, sp);
                break;
            case 16:
                if (a.grid) hipLaunchKernelGGL((splitShGroupKernel<16, true>), grid, dim3(1024), 0, stream, a, fm, sp);
                else hipLaunchKernelGGL((splitShGroupKernel<16, false>), grid, dim3(1024), 0, stream, a, fm, sp);
                break;
            default: return -2;
        }
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    const dim3 grid(a.nTiles);
    if (a.grid) {
        switch (shWaves) {
            case 4: hipLaunchKernelGGL((splitShKernel<4, true>), grid, dim3(256), 0, stream, a, fm, sp); break;
            case 8: hipLaunchKernelGGL((splitShKernel<8, true>), grid, dim3(512), 0, stream, a, fm, sp); break;
            case 16:
                if (fm.nRot == 0) hipLaunchKernelGGL((splitShKernel<16, true, false>), grid, dim3(1024), 0, stream, a, fm, sp);
                else hipLaunchKernelGGL((splitShKernel<16, true>), grid, dim3(1024), 0, stream, a, fm, sp);
                break;
            default: return -2;
        }
    } else {
        switch (shWaves) {
            case 4: hipLaunchKernelGGL((splitShKernel<4, false>), grid, dim3(256), 0, stream, a, fm, sp); break;
            case 8: hipLaunchKernelGGL((splitShKernel<8, false>), grid, dim3(512), 0, stream, a, fm, sp); break;
            case 16:
                if (fm.nRot == 0) hipLaunchKernelGGL((splitShKernel<16, false, false>), grid, dim3(1024), 0, stream, a, fm, sp);
                else hipLaunchKernelGGL((splitShKernel<16, false>), grid, dim3(1024), 0, stream, a, fm, sp);
                break;
            default: return -2;
        }
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static int launchSplitSpecies(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, hipStream_t stream) {
    if (a.nTiles <= 0 || !fm.drag) return 0;
    switch (sp.speciesPerGroup) {
        case 1: hipLaunchKernelGGL((splitSpeciesKernel<1>), dim3(a.nTiles * 8), dim3(64), 0, stream, a, fm, sp); break;
        case 2: hipLaunchKernelGGL((splitSpeciesKernel<2>), dim3(a.nTiles * 4), dim3(128), 0, stream, a, fm, sp); break;
        case 4: hipLaunchKernelGGL((splitSpeciesKernel<4>), dim3(a.nTiles * 2), dim3(256), 0, stream, a, fm, sp); break;
        default: return -2;   // the four-wave drag kernel needs the SH kernel's stage state
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static int launchSplitFinish(const StepArgs& a, const ForceModel& fm, const SplitArgs& sp, hipStream_t stream) {
    if (a.nTiles <= 0 || !fm.drag) return 0;
    if (sp.speciesPerGroup == 0) hipLaunchKernelGGL(splitDragKernel, dim3(a.nTiles), dim3(256), 0, stream, a, fm, sp);
    else hipLaunchKernelGGL(splitDragFinishKernel, dim3(a.nTiles), dim3(64), 0, stream, a, fm, sp);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launchSplitCombine(const StepArgs& a, const SplitArgs& sp, hipStream_t stream) {
    if (a.nTiles <= 0) return 0;
    hipLaunchKernelGGL(splitCombineKernel, dim3(a.nTiles), dim3(64), 0, stream, a, sp);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launchStep(const StepArgs& a, const ForceModel& fm, int waves, hipStream_t stream) {
    if (a.nTiles <= 0) return 0;
    int rc;
    if (a.grid) rc = fm.drag ? launchShape<true, true>(a, fm, waves, stream) : launchShape<true, false>(a, fm, waves, stream);
    else rc = fm.drag ? launchShape<false, true>(a, fm, waves, stream) : launchShape<false, false>(a, fm, waves, stream);
    if (rc != 0) return rc;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launchForceProbe(int n, const double* pos, const double* vel, const double* ams, const double* bcd, const double* record,
                     int which, const ForceModel& fm, const double* shTerms, const int* shOrderStart, const double* shDiag,
                     const double* nrlCoeff, double* acc, hipStream_t stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(forceProbeKernel, dim3((n + 63) / 64), dim3(64), 0, stream, n, pos, vel, ams, bcd, record, which, fm,
                       shTerms, shOrderStart, shDiag, nrlCoeff, acc);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launchNrlmsiseProbe(int n, const double* in, int useApArray, const double* nrlCoeff, double* out, hipStream_t stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(nrlProbeKernel, dim3((n + 63) / 64), dim3(64), 0, stream, n, in, useApArray, nrlCoeff, out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace phx
