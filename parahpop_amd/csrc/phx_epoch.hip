// Epoch-record kernel: one lane per epoch computes every quantity of the force
// model that depends on the epoch alone (Earth orientation, Sun/Moon/planet
// positions, the indirect third-body term, the solid+ocean tide Stokes
// corrections, NRLMSISE-00 time and weather inputs). The step kernel reads one
// record per stage with scalar loads (shared-epoch batches) or interpolates a
// grid of records (uneven epochs), so this work is done once per epoch instead
// of once per object and stage.
#include <hip/hip_runtime.h>

#include "phx_common.h"
#include "phx_epoch_device.h"
#include "phx_kernels.h"

namespace phx {
namespace {
using namespace epochdev;

// Position (km) of one ephemeris slot relative to the Earth at `et`: the chain's
// body units, each by backward Clenshaw over the coefficients of the interval
// holding `et` (BodyInterpolator.h:55-100, chebyshev/Interpolator.h:50-90),
// the same operations as the host evaluation in brie.h.
__device__ D3 ephPosition(const double* __restrict__ eph, const double* __restrict__ row, double et) {
    double out[3] = {0.0, 0.0, 0.0};
    const int steps = int(row[0]);
    for (int k = 0; k < steps && k < kEphSteps; ++k) {
        const double* st = row + 1 + 4 * k;
        const long long n = (long long)st[1];
        const int pdeg = int(st[2]);
        const double sign = st[3];
        const double* d = eph + (long long)st[0];
        const double r = d[0];
        long long idx = (long long)((et - d[1]) / (2.0 * r));
        idx = idx < 0 ? 0 : (idx > n - 1 ? n - 1 : idx);
        const double s = (et - d[1 + idx]) / r - 1.0;
        const double s2 = 2.0 * s;
        for (int c = 0; c < 3; ++c) {
            const double* q = d + 1 + n + (long long)c * n * (pdeg + 1) + idx;
            double w = 0.0, x = 0.0;
            for (int i = pdeg; i >= 1; --i) {
                const double tmp = q[i * n] - w;
                w = x;
                x = fma(s2, x, tmp);
            }
            out[c] += sign * fma(s, x, q[0] - w);
        }
    }
    return D3{out[0], out[1], out[2]};
}

// Orientation of a body with a field of its own: ICRF -> body axes as the
// passive 3-1-3 rotation R3(psi) R1(theta) R3(phi) of the IAU angles
// (frames/Rotation.h:158-210, raDecPMToQuat: phi = pi/2 + RA, theta = pi/2 - Dec,
// psi = W), from an IAU body unit (orientations/RefEphUnit.h:190-330: RA, Dec in
// centuries and W in days past J2000 TDB, degrees, plus the nutation-precession
// sums of core/Polynomial.h:40-80, each angle reduced modulo 360 and converted
// to radians, frames/Frames.h:250-270) or a binary frame unit (Chebyshev Euler
// angles, RefEphUnit.h:400-420; ECLIPJ2000-based units, INERT_FRAME 17, compose
// the constant bias R1(eps), frames/Frames.h:40-75).
__device__ inline double polyAscending(const double* c, int n, double t) {
    double r = 0.0;
    for (int i = n - 1; i >= 0; --i) r = fma(r, t, c[i]);
    return r;
}

// sum_i amp[i] f(angle_i(t) deg), angle_i the i-th angle polynomial (ascending
// powers of t) of the barycenter unit; f = sin or cos
__device__ inline double npSum(const double* amp, int n, const double* ang, int angDeg, double t, bool useSin) {
    const double deg2rad = 3.14159265358979323846 / 180.0;
    double r = 0.0;
    for (int i = 0; i < n; ++i) {
        const double* q = ang + (long long)i * (angDeg + 1);
        double v = 0.0, tp = 1.0;
        for (int j = 0; j <= angDeg; ++j) {
            v = fma(q[j], tp, v);
            tp *= t;
        }
        r = fma(amp[i], useSin ? sin(v * deg2rad) : cos(v * deg2rad), r);
    }
    return r;
}

__device__ inline void passive313(double phi, double theta, double psi, double R[9]) {
    const double c1 = cos(phi), s1 = sin(phi), c2 = cos(theta), s2 = sin(theta), c3 = cos(psi), s3 = sin(psi);
    // R3(psi) R1(theta) R3(phi), passive (frame) rotations
    R[0] = c3 * c1 - s3 * c2 * s1;
    R[1] = c3 * s1 + s3 * c2 * c1;
    R[2] = s3 * s2;
    R[3] = -s3 * c1 - c3 * c2 * s1;
    R[4] = -s3 * s1 + c3 * c2 * c1;
    R[5] = c3 * s2;
    R[6] = s2 * s1;
    R[7] = -s2 * c1;
    R[8] = c2;
}

__device__ void bodyAxes(const EpochArgs& a, const RotUnit& u, double et, double R[9]) {
    const double halfPi = 1.57079632679489661923;
    const double* d = a.brot + u.off;
    double ra, dec, pm;
    if (u.type == 1) {
        const double tc = et / 3155760000.0, td = et / 86400.0;
        ra = polyAscending(d, u.degRa + 1, tc);
        dec = polyAscending(d + u.degRa + 1, u.degDec + 1, tc);
        pm = polyAscending(d + u.degRa + 1 + u.degDec + 1, u.degPm + 1, td);
        if (u.bary >= 0) {
            const double* amp = d + (u.degRa + 1) + (u.degDec + 1) + (u.degPm + 1);
            const double* ang = a.brot + u.bary;
            ra += npSum(amp, u.npRa, ang, u.baryDeg, tc, true);
            dec += npSum(amp + u.npRa, u.npDec, ang, u.baryDeg, tc, false);
            pm += npSum(amp + u.npRa + u.npDec, u.npPm, ang, u.baryDeg, tc, true);
        }
        const double deg2rad = 3.14159265358979323846 / 180.0;
        ra = fmod(ra, 360.0) * deg2rad;
        dec = fmod(dec, 360.0) * deg2rad;
        pm = fmod(pm, 360.0) * deg2rad;
    } else {
        // Chebyshev Euler angles (phi, theta, psi) of the interval holding et
        const double r = d[0];
        const long long n = u.nIntervals;
        long long idx = (long long)((et - d[1]) / (2.0 * r));
        idx = idx < 0 ? 0 : (idx > n - 1 ? n - 1 : idx);
        const double s = (et - d[1 + idx]) / r - 1.0, s2 = 2.0 * s;
        double e[3];
        for (int k = 0; k < 3; ++k) {
            const double* c = d + 1 + n + (long long)k * n * (u.pdeg + 1) + idx;
            double w = 0.0, x = 0.0;
            for (int i = u.pdeg; i >= 1; --i) {
                const double tmp = c[i * n] - w;
                w = x;
                x = fma(s2, x, tmp);
            }
            e[k] = fma(s, x, c[0] - w);
        }
        ra = e[0] - halfPi;
        dec = halfPi - e[1];
        pm = fmod(e[2], 2.0 * 3.14159265358979323846);
    }
    passive313(halfPi + ra, halfPi - dec, pm, R);
    if (u.type == 3 && u.inertFrame == 17) {
        // R(ICRF -> body) = R(ECLIPJ2000 -> body) R1(eps)
        const double eps = 84381.448 / 3600.0 * (3.14159265358979323846 / 180.0);
        const double ce = cos(eps), se = sin(eps);
        for (int row = 0; row < 3; ++row) {
            const double y = R[3 * row + 1], z = R[3 * row + 2];
            R[3 * row + 1] = y * ce - z * se;
            R[3 * row + 2] = y * se + z * ce;
        }
    }
}

__global__ void __launch_bounds__(64) epochRecordKernel(EpochArgs a, ForceModel fm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const double et = a.epochs[i];
    double* rec = a.out + size_t(i) * REC_SIZE;
    for (int k = 0; k < REC_SIZE; ++k) rec[k] = 0.0;
    double R[9], w[3];
    if (a.iers) iersOrientation(a, et, R, w);
    else orientation(a, et, R, w);
    for (int k = 0; k < 9; ++k) rec[REC_R + k] = R[k];
    if (a.gridRows && a.leapCount > 0 && !a.iers) {
        // interpolation rows: R3(-theta) R, smooth across the UT1 steps (the
        // step kernels apply R3(theta) of each object's own epoch)
        const double th = leapTheta(a.leaps, a.leapCount, et);
        if (th != 0.0) {
            const double c = cos(th), s = sin(th);
            for (int k = 0; k < 3; ++k) {
                rec[REC_R + k] = c * R[k] - s * R[3 + k];
                rec[REC_R + 3 + k] = s * R[k] + c * R[3 + k];
            }
        }
    }
    for (int k = 0; k < 3; ++k) rec[REC_OMEGA + k] = w[k];
    double ind[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < fm.nBodies; ++b) {
        const D3 p = a.eph ? ephPosition(a.eph, a.ephChain + b * kEphRow, et) : bodyPos(fm.bodyNaif[b], et);
        rec[REC_BODY + 3 * b + 0] = p.x;
        rec[REC_BODY + 3 * b + 1] = p.y;
        rec[REC_BODY + 3 * b + 2] = p.z;
        if (fm.bodyPointGravity[b] && fm.bodyNaif[b] != 399) {
            const double s2 = p.x * p.x + p.y * p.y + p.z * p.z;
            const double rc = 1.0 / (s2 * sqrt(s2));
            const double g = fm.bodyGm[b];
            ind[0] += (-g * p.x) * rc;
            ind[1] += (-g * p.y) * rc;
            ind[2] += (-g * p.z) * rc;
        }
    }
    for (int k = 0; k < 3; ++k) rec[REC_INDIRECT + k] = ind[k];
    // constant indices into the by-value argument (no copy to scratch memory)
    static_assert(kMaxRotFields == 3, "one call per body field");
    double B[9];
    if (a.nRot > 0) {
        bodyAxes(a, a.rot[0], et, B);
        for (int j = 0; j < 9; ++j) rec[REC_BROT + j] = B[j];
    }
    if (a.nRot > 1) {
        bodyAxes(a, a.rot[1], et, B);
        for (int j = 0; j < 9; ++j) rec[REC_BROT + 9 + j] = B[j];
    }
    if (a.nRot > 2) {
        bodyAxes(a, a.rot[2], et, B);
        for (int j = 0; j < 9; ++j) rec[REC_BROT + 18 + j] = B[j];
    }
    if (a.solid || a.ocean) {
        const Times tm = eopTimes(a, et);
        double C[7][7], S[7][7];
        for (int n = 0; n < 7; ++n) for (int m = 0; m < 7; ++m) C[n][m] = S[n][m] = 0.0;
        if (!tm.valid) {
            for (int n = 0; n < 7; ++n) for (int m = 0; m < 7; ++m) C[n][m] = S[n][m] = nan("");
        } else {
            if (a.solid) {
                const double T = et / (86400.0 * 36525.0);
                const D3 sun = a.eph ? ephPosition(a.eph, a.ephChain + kMaxBodies * kEphRow, et) : sunPos(T);
                const D3 moon = a.eph ? ephPosition(a.eph, a.ephChain + (kMaxBodies + 1) * kEphRow, et) : moonPos(T);
                solidTide(a, tm, R, sun, moon, C, S);
            }
            if (a.ocean) oceanTide(a, tm, C, S);
        }
        for (int n = 0; n <= 6; ++n)
            for (int m = 0; m <= n; ++m) {
                rec[REC_TIDE_C + n * (n + 1) / 2 + m] = C[n][m];
                rec[REC_TIDE_S + n * (n + 1) / 2 + m] = S[n][m];
            }
    }
    if (fm.drag) nrlEpoch(a, et, rec + REC_NRL, rec + REC_NRL_GLOBE);
}

__global__ void __launch_bounds__(64) nrlSlotKernel(EpochArgs a, double* out, int rows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    nrlSlotRow(a, i / 8, i % 8, out + size_t(i) * nrlmsise::SLOT_SIZE);
}

}  // namespace

int launchNrlSlots(const EpochArgs& a, double* out, int rows, hipStream_t stream) {
    if (rows <= 0) return 0;
    hipLaunchKernelGGL(nrlSlotKernel, dim3((rows + 63) / 64), dim3(64), 0, stream, a, out, rows);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launchEpochRecords(const EpochArgs& a, const ForceModel& fm, hipStream_t stream) {
    if (a.count <= 0) return 0;
    const int block = 64;
    const int grid = (a.count + block - 1) / block;
    hipLaunchKernelGGL(epochRecordKernel, dim3(grid), dim3(block), 0, stream, a, fm);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace phx
