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

__global__ void __launch_bounds__(64) epochRecordKernel(EpochArgs a, ForceModel fm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    const double et = a.epochs[i];
    double* rec = a.out + size_t(i) * REC_SIZE;
    for (int k = 0; k < REC_SIZE; ++k) rec[k] = 0.0;
    double R[9], w[3];
    orientation(a, et, R, w);
    for (int k = 0; k < 9; ++k) rec[REC_R + k] = R[k];
    for (int k = 0; k < 3; ++k) rec[REC_OMEGA + k] = w[k];
    double ind[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < fm.nBodies; ++b) {
        const D3 p = bodyPos(fm.bodyNaif[b], et);
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
                solidTide(a, tm, R, sunPos(T), moonPos(T), C, S);
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
