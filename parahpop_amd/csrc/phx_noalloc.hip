// phx_propagate_noalloc: phx_propagate on caller-owned device buffers with
// every working array in one fixed device array of the library, so the call
// allocates no device memory (for callers that cannot: graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The call runs the per-lane epoch path of device batches (epoch records on a
// uniform grid, per-lane interpolation; the split stages when the force model
// has NRLMSISE-00 drag and a field, else the fused kernel), so its results are
// those of phx_propagate with the option force_grid = 1: the per-object set-up
// below forms the epochs, first step and area ratios with the same roundings
// as the host code (driver.cpp:117-120, propagators/FixedStep.h:7-29), with no
// contracted multiply-adds.
//
// Limits: at most kNaMaxObjects objects and an epoch span of kNaMaxRows - 5
// grid steps. Every call on one GPU uses that GPU's array and leaves its
// contents; the set-up kernels clear what must start at zero. Calls that
// overlap on one GPU race on the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "phx_internal.h"

namespace phx {
namespace {

constexpr int kNaMaxObjects = 16384;   // padded object slots
constexpr int kNaMaxRows = 2048;       // epoch-record grid rows
constexpr int kNaMaxDevices = 64;

constexpr size_t alignUp(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t S = kNaMaxObjects;
// byte offsets of the arrays in the scratch array
constexpr size_t kOffY = 0;                                          // [6][S] state
constexpr size_t kOffCur = alignUp(kOffY + 6 * S * 8);               // [S] epoch (ET s)
constexpr size_t kOffEnd = alignUp(kOffCur + S * 8);                 // [S]
constexpr size_t kOffDt = alignUp(kOffEnd + S * 8);                  // [S]
constexpr size_t kOffAms = alignUp(kOffDt + S * 8);                  // [S]
constexpr size_t kOffBcd = alignUp(kOffAms + S * 8);                 // [S]
constexpr size_t kOffSlotT0 = alignUp(kOffBcd + S * 8);              // [S]
constexpr size_t kOffSlotRec = alignUp(kOffSlotT0 + S * 8);          // [S] long long
constexpr size_t kOffSlotRows = alignUp(kOffSlotRec + S * 8);        // [S] int
constexpr size_t kOffStatus = alignUp(kOffSlotRows + S * 4);         // [S] uint8
constexpr size_t kOffKs = alignUp(kOffStatus + S);                   // [13][6][S] k-stages
constexpr size_t kOffYs = alignUp(kOffKs + size_t(kStages) * 6 * S * 8);   // [7][S] stage state
constexpr size_t kOffParts = alignUp(kOffYs + 7 * S * 8);            // [6][S]
constexpr size_t kOffLane = alignUp(kOffParts + 6 * S * 8);          // [8][S] NRLMSISE-00 lane inputs
constexpr size_t kOffRec = alignUp(kOffLane + 8 * S * 8);            // [rows][REC_SIZE] epoch records
constexpr size_t kOffRowEt = alignUp(kOffRec + size_t(kNaMaxRows) * REC_SIZE * 8);   // [rows] grid epochs
constexpr size_t kOffCounters = alignUp(kOffRowEt + size_t(kNaMaxRows) * 8);          // [4] long long
constexpr size_t kOffSummary = alignUp(kOffCounters + 4 * 8);       // NaSummary
constexpr size_t kNaBytes = alignUp(kOffSummary + 64);

// Set-up results read back by the host.
struct NaSummary {
    double lo, hi;       // earliest start, latest end (ET s)
    double count;        // max over objects of ceil(span / step) + 1
    int flags;           // 1 nonfinite state, 2 invalid epoch or parameter, 4 duration, 8 not at end
    int pad;
};
static_assert(sizeof(NaSummary) <= 64, "summary block");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

void* g_scratch[kNaMaxDevices] = {};
std::mutex g_scratchMu;

template <class T>
__device__ __host__ inline T* at(unsigned char* base, size_t off) {
    return reinterpret_cast<T*>(base + off);
}

// Per-object set-up over all padded slots (one workgroup): epochs, first step,
// area ratios, status, grid segment; validation and the span reductions.
__global__ void __launch_bounds__(1024) naSetupKernel(unsigned char* base, int n, int nPad, const double* epoch,
                                                      const double* state6, const double* mass, const double* drag,
                                                      const double* srp, const double* cd, const double* cr, double duration,
                                                      double step) {
    __shared__ double sLo[1024], sHi[1024], sCount[1024];
    __shared__ int sFlags;
    if (threadIdx.x == 0) sFlags = 0;
    __syncthreads();
    double lo = INFINITY, hi = -INFINITY, count = 0.0;
    int flags = 0;
    for (int s = threadIdx.x; s < nPad; s += blockDim.x) {
        const bool pad = s >= n;
        const int i = pad ? 0 : s;   // padding lanes copy object 0
        for (int d = 0; d < 6; ++d) {
            const double v = state6[size_t(d) * n + i];
            at<double>(base, kOffY)[size_t(d) * S + s] = v;
            if (!isfinite(v)) flags |= 1;
        }
        const double e = epoch[i], m = mass[i], a = drag[i], r = srp[i], c = cd[i], q = cr[i];
        if (!isfinite(e) || !isfinite(m) || !isfinite(a) || !isfinite(r) || !isfinite(c) || !isfinite(q) || m <= 0 || a <= 0 ||
            r < 0 || c < 0 || q < 0)
            flags |= 2;
        // FixedStep.h:7-29 and driver.cpp:117-120, rounded as the host code
        const double start = __dadd_rn(__dmul_rn(e, 86400.0), -43200.0);
        const double end = __dadd_rn(__dmul_rn(__dadd_rn(e, __ddiv_rn(duration, 86400.0)), 86400.0), -43200.0);
        const double span = __dsub_rn(end, start);
        if (!(span > 0) || !isfinite(span) || ceil(__ddiv_rn(span, step)) >= 2147483647.0) flags |= 4;
        const double area = __dmul_rn(a, 1e-6);
        const double crEff = __ddiv_rn(__dmul_rn(q, r), a);
        at<double>(base, kOffCur)[s] = start;
        at<double>(base, kOffEnd)[s] = end;
        at<double>(base, kOffDt)[s] = fmin(step, span);
        at<double>(base, kOffAms)[s] = __ddiv_rn(__dmul_rn(crEff, area), m);
        at<double>(base, kOffBcd)[s] = __ddiv_rn(__dmul_rn(c, area), m);
        at<uint8_t>(base, kOffStatus)[s] = pad ? 4 : 0;
        if (!pad) {
            lo = fmin(lo, start);
            hi = fmax(hi, end);
            count = fmax(count, __dadd_rn(ceil(__ddiv_rn(span, step)), 1.0));
        }
    }
    sLo[threadIdx.x] = lo;
    sHi[threadIdx.x] = hi;
    sCount[threadIdx.x] = count;
    if (flags) atomicOr(&sFlags, flags);
    __syncthreads();
    for (int k = blockDim.x / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) {
            sLo[threadIdx.x] = fmin(sLo[threadIdx.x], sLo[threadIdx.x + k]);
            sHi[threadIdx.x] = fmax(sHi[threadIdx.x], sHi[threadIdx.x + k]);
            sCount[threadIdx.x] = fmax(sCount[threadIdx.x], sCount[threadIdx.x + k]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        NaSummary* sum = at<NaSummary>(base, kOffSummary);
        sum->lo = sLo[0];
        sum->hi = sHi[0];
        sum->count = sCount[0];
        sum->flags = sFlags;
        long long* ctr = at<long long>(base, kOffCounters);
        for (int k = 0; k < 4; ++k) ctr[k] = 0;
    }
}

// The grid segment every slot reads (row 0 of the record array) and the grid
// row epochs gridT0 + k h, rounded as the host code.
__global__ void __launch_bounds__(256) naGridKernel(unsigned char* base, int nPad, int rows, double gridT0, double h) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nPad) {
        at<double>(base, kOffSlotT0)[t] = gridT0;
        at<long long>(base, kOffSlotRec)[t] = 0;
        at<int>(base, kOffSlotRows)[t] = rows;
    }
    if (t < rows) at<double>(base, kOffRowEt)[t] = __dadd_rn(gridT0, __dmul_rn(double(t), h));
}

// Final epochs (MJD2000 TDB, TimeConversion.h:36-40) and states into the
// caller's buffers; flags a target that did not reach its end.
__global__ void __launch_bounds__(256) naFinishKernel(unsigned char* base, int n, double* epoch, double* state6) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    epoch[i] = __dadd_rn(__dmul_rn(at<double>(base, kOffCur)[i], 1.0 / 86400.0), 0.5);
    for (int d = 0; d < 6; ++d) state6[size_t(d) * n + i] = at<double>(base, kOffY)[size_t(d) * S + i];
    if (!(at<uint8_t>(base, kOffStatus)[i] & 2)) atomicOr(&at<NaSummary>(base, kOffSummary)->flags, 8);
}

}  // namespace

// Device address of this device's scratch array (set-up: phx_create calls it
// once per device with that device current).
int noallocInit() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (g_scratch[dev]) return 0;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    g_scratch[dev] = p;
    return 0;
}

}  // namespace phx

using namespace phx;

extern "C" int phx_propagate_noalloc(phx_ctx* c, size_t n, double* d_epoch, double* d_state6_soa, const double* d_mass,
                                     const double* d_drag_area, const double* d_srp_area, const double* d_cd,
                                     const double* d_cr, double duration_s, phx_stats* stats) {
    if (!c) return PHX_EINVAL;
    try {
        if (n == 0) throw std::invalid_argument("targets must be a nonempty array.");
        if (n > size_t(kNaMaxObjects))
            throw std::invalid_argument("phx_propagate_noalloc holds at most " + std::to_string(kNaMaxObjects) + " objects (got " +
                                        std::to_string(n) + ")");
        if (!d_epoch || !d_state6_soa || !d_mass || !d_drag_area || !d_srp_area || !d_cd || !d_cr)
            throw std::invalid_argument("phx_propagate_noalloc: a device buffer is null.");
        if (!(duration_s > 0) || !std::isfinite(duration_s)) throw std::invalid_argument("duration_seconds must be finite and positive.");
        int dev = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        unsigned char* base = dev >= 0 && dev < kNaMaxDevices ? static_cast<unsigned char*>(g_scratch[dev]) : nullptr;
        if (!base) throw HipError("phx_propagate_noalloc: the device scratch array was not set up on this device");
        const auto w0 = std::chrono::steady_clock::now();
        hipStream_t s = c->stream;
        const int nn = int(n);
        const int nPad = (nn + kTile - 1) / kTile * kTile;
        const double step = c->run.stepSeconds, h = c->gridStep;
        hipLaunchKernelGGL(naSetupKernel, dim3(1), dim3(1024), 0, s, base, nn, nPad, (const double*)d_epoch,
                           (const double*)d_state6_soa, d_mass, d_drag_area, d_srp_area, d_cd, d_cr, duration_s, step);
        check(hipGetLastError(), "set-up kernel launch");
        NaSummary sum{};
        check(hipMemcpyAsync(&sum, base + kOffSummary, sizeof sum, hipMemcpyDeviceToHost, s), "set-up read-back");
        check(hipStreamSynchronize(s), "set-up");
        if (sum.flags & 1) throw std::invalid_argument("State contains a nonfinite value.");
        if (sum.flags & 2) throw std::invalid_argument("Invalid target epoch or physical parameter.");
        if (sum.flags & 4) throw std::invalid_argument("Duration/step cannot be represented by the integration time/index types.");
        if (needsCoverageCheck(*c)) checkEphemerisCoverage(*c, sum.lo, sum.hi);
        // the grid of the batch path: two rows of margin each side
        const double gridT0 = sum.lo - 2.0 * h;
        const double rowsD = std::ceil((sum.hi - gridT0) / h) + 3;
        if (!(rowsD <= double(kNaMaxRows)))
            throw std::invalid_argument("phx_propagate_noalloc: the epoch span needs " + std::to_string((long long)rowsD) +
                                        " record rows; at most " + std::to_string(kNaMaxRows) + " fit");
        const int rows = int(rowsD);
        const long long launches = std::min((long long)sum.count, c->run.maxNumSteps);
        hipLaunchKernelGGL(naGridKernel, dim3((std::max(nPad, rows) + 255) / 256), dim3(256), 0, s, base, nPad, rows, gridT0, h);
        check(hipGetLastError(), "grid kernel launch");
        EpochArgs ea = epochArgs(*c);
        ea.gridRows = 1;
        ea.count = rows;
        ea.epochs = at<double>(base, kOffRowEt);
        ea.out = at<double>(base, kOffRec);
        if (launchEpochRecords(ea, c->fm, s) != 0) throw HipError("epoch-record kernel launch failed");
        int W = 1;
        StepArgs a = contextStepArgs(*c, W);
        a.nTiles = nPad / kTile;
        a.nPad = int(S);
        a.y = at<double>(base, kOffY);
        a.cur = at<double>(base, kOffCur);
        a.end = at<double>(base, kOffEnd);
        a.nextDt = at<double>(base, kOffDt);
        a.status = at<uint8_t>(base, kOffStatus);
        a.ams = at<double>(base, kOffAms);
        a.bcd = at<double>(base, kOffBcd);
        a.records = at<double>(base, kOffRec);
        a.tileRecord = nullptr;
        a.tileList = nullptr;
        a.counters = at<long long>(base, kOffCounters);
        a.grid = 1;
        a.gridInvH = 1.0 / h;
        a.slotRecord = at<long long>(base, kOffSlotRec);
        a.slotT0 = at<double>(base, kOffSlotT0);
        a.slotRows = at<int>(base, kOffSlotRows);
        const bool split = c->split && c->fm.drag && c->fm.hasSh && c->splitSpecies == 8;
        long long kernels = 0;
        if (split) {
            W = c->splitWaves;
            orderBlocks(*c, W, a.blockStart, true);
            SplitArgs sp{};
            sp.ks = at<double>(base, kOffKs);
            sp.ys = at<double>(base, kOffYs);
            sp.parts = at<double>(base, kOffParts);
            sp.lane = at<double>(base, kOffLane);
            sp.speciesPerGroup = 8;
            sp.groups = 1;
            for (long long k = 0; k < launches; ++k) {
                sp.step = int(k);
                for (int st = 0; st < kStages; ++st) {
                    sp.stage = st;
                    if (launchSplitStage(a, c->fm, sp, W, s) != 0) throw HipError("split stage launch failed");
                }
                if (launchSplitCombine(a, sp, s) != 0) throw HipError("split combine launch failed");
                kernels += 2 * kStages + 1;
            }
        } else {
            const int K = std::max(1, c->stepsPerLaunch);
            for (long long k = 0; k < launches; k += K) {
                a.stepBase = int(k);
                a.steps = int(std::min<long long>(K, launches - k));
                if (launchStep(a, c->fm, W, s) != 0) throw HipError("step kernel launch failed");
                ++kernels;
            }
        }
        hipLaunchKernelGGL(naFinishKernel, dim3((nn + 255) / 256), dim3(256), 0, s, base, nn, d_epoch, d_state6_soa);
        check(hipGetLastError(), "finish kernel launch");
        long long steps = 0;
        check(hipMemcpyAsync(&sum, base + kOffSummary, sizeof sum, hipMemcpyDeviceToHost, s), "read-back");
        check(hipMemcpyAsync(&steps, base + kOffCounters, sizeof steps, hipMemcpyDeviceToHost, s), "read-back");
        check(hipStreamSynchronize(s), "propagation");
        if (stats) {
            *stats = phx_stats{};
            stats->object_steps = double(steps);
            stats->epoch_records = rows;
            stats->launches = double(kernels);
            stats->waves = W;
            stats->steps_per_launch = split ? 1 : c->stepsPerLaunch;
            stats->kernels_per_step = split ? 2 * kStages + 1 : 0;
            stats->sh_groups_min = stats->sh_groups_max = split ? 1 : 0;
            stats->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
        }
        if (sum.flags & 8) {
            c->setError("A target stopped without reaching the requested end of simulation.");
            return PHX_ENOTDONE;
        }
        return PHX_OK;
    } catch (const HipError& e) {
        c->setError(e.what());
        return PHX_EDEVICE;
    } catch (const std::exception& e) {
        c->setError(e.what());
        return PHX_EINVAL;
    }
}
