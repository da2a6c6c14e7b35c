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
