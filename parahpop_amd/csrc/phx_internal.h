// Host-side internals shared by the C-ABI (phx_api.cpp) and the request
// server (phx_serve.cpp): device buffers, the context and batch objects and the
// helpers that fill the step-kernel arguments.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/phx.h"
#include "model.h"
#include "phx_common.h"
#include "phx_kernels.h"

namespace phx {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// Owning device array (hipMalloc/hipFree).
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
        n = count;
    }
    void upload(const T* host, size_t count, hipStream_t s) {
        alloc(count);
        if (count) check(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    }
};

class Server;

// RKF78 node fractions c_2..c_13 (parm/integrate/rk/coefs/tableau.h:19-31).
extern const double kRkC[12];

}  // namespace phx

struct phx_ctx {
    phx::RunConfig run;
    phx::Environment env;
    phx::ForceModel fm{};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // last error message: phx_submit / phx_poll may run on any thread, so it is
    // written and copied under errMu
    std::string error;
    mutable std::mutex errMu;
    void setError(const std::string& m) {
        std::lock_guard<std::mutex> lk(errMu);
        error = m;
    }
    std::string lastError() const {
        std::lock_guard<std::mutex> lk(errMu);
        return error;
    }
    std::string dataDir;
    // device tables
    phx::DBuf<double> shTerms, shDiag, nrl, freq, waves, terms, eop, weather, nrlSlots;
    phx::DBuf<double> eph, ephChain;   // Chebyshev ephemeris data and per-slot chains (empty: analytic)
    phx::DBuf<double> iers;            // IERS2000 orientation payload (empty: analytic model)
    // UT1 steps of the EOP time model at leap seconds: [leapCount] ET of each
    // step, then [leapCount] Earth rotation angle offsets after it
    phx::DBuf<double> leaps;
    int leapCount = 0;
    std::vector<std::vector<phx::BrieStep>> ephChains;   // host copy, for coverage checks
    phx::DBuf<float> shTerms32;
    phx::DBuf<int> shOrderStart;
    // fields of the bodies other than the Earth: term streams and the
    // orientation table their axes come from
    phx::DBuf<double> rotTerms[phx::kMaxRotFields], rotDiag[phx::kMaxRotFields], brotBody;
    phx::DBuf<int> rotStart[phx::kMaxRotFields];
    int freqCount[3] = {0, 0, 0};
    int waveCount = 0;
    // options
    int numCU = 256;
    int nWaves = 4;
    int stepsPerLaunch = 32;
    double share0 = 0.85, shareCorr = 0.85;   // SH order-block shares of wave 0 and of the corrections wave
    int sortAltitude = 1;
    int exactGroupsMax = 16;     // more distinct (start, end) groups than this -> grid epochs
    double gridStep = 60.0;      // s, spacing of the epoch grid
    int graphSteps = 8;          // split stages: steps per replayed graph (0: launch every kernel)
    unsigned long long optGen = 0;   // advanced by every option change (captured step graphs expire)
    int forceGrid = 0;
    int split = 1;               // shared-epoch batches with drag: split stages
    int splitWaves = 16;         // waves per tile of the drag-free split kernel
    int splitGroups = 1;         // at most this many SH workgroups per tile while the active tiles leave CUs idle
                                 // (1: one workgroup, whose sums do not depend on the rest of the batch)
    int splitSpecies = 8;        // 8: one eight-wave drag workgroup per tile (species and drag); 1, 2, 4:
                                 // species per workgroup, then a finish kernel; 0: one four-wave workgroup per tile
    // split SH kernel (8 and 16 waves): costs of the waves' extra families and of
    // an order, in FP64 terms of one wave (balancing of the order blocks)
    double splitCostPg = 105.0, splitCostCorr = 246.0, splitCostSrp = 83.0, splitCostLane = 104.0;
    double splitCostOrder = 2.0, splitCostOrder0 = 1.8;
    // serving options
    int serveStepsPerTick = 16;
    long long serveSlots = 262144;      // object slots of the device pool
    double serveRecordMiB = 8192.0;     // epoch-record arena
    // request server, created by the first phx_submit; serverMu guards its
    // creation and every read of the pointer from the serving calls
    std::unique_ptr<phx::Server> server;
    mutable std::mutex serverMu;
};

struct phx_batch {
    phx_ctx* ctx = nullptr;
    size_t n = 0;
    int nTiles = 0;
    int nPad = 0;
    std::vector<int> slotOf;     // original index -> padded slot
    int totalSteps = 0;          // steps of the longest group
    std::vector<int> tileSteps;  // per tile: steps until its last object ends (upper bound)
    int stepsDone = 0;
    phx::DBuf<double> y, cur, end, nextDt, ams, bcd, records, slotT0;
    phx::DBuf<double> ks, ysBuf, parts, spec, lane, trig;   // split stages (allocated on first use)
    phx::DBuf<double> shPart, famPart;               // grouped SH workgroups: partial sums
    phx::DBuf<unsigned> tileCount;                   // grouped SH workgroups: arrivals per tile (zeroed)
    phx::DBuf<uint8_t> status;
    phx::DBuf<long long> tileRecord, counters, slotRecord;
    phx::DBuf<int> slotRows;
    std::vector<double> y0, cur0, dt0;
    std::vector<uint8_t> st0;
    size_t epochRecords = 0;
    double epochSeconds = 0.0;
    bool grid = false;
    double gridInvH = 0.0;
    // split stages: the device step index of exact records, and the captured
    // graphs of `steps` split steps for `active` tiles and `groups` SH
    // workgroups per tile, valid while the context's option generation is `gen`
    // (kept across calls; the oldest is dropped beyond kMaxStepGraphs)
    phx::DBuf<int> stepDev;
    int stepHost = 0;
    struct StepGraph {
        int active, groups, steps;
        unsigned long long gen;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    static constexpr size_t kMaxStepGraphs = 16;
    std::vector<StepGraph> graphs;
    long long graphReplays = 0, graphCaptures = 0;
    phx_batch() = default;
    phx_batch(const phx_batch&) = delete;
    phx_batch& operator=(const phx_batch&) = delete;
    void dropGraph(size_t i) {
        (void)hipGraphExecDestroy(graphs[i].exec);
        (void)hipGraphDestroy(graphs[i].graph);
        graphs.erase(graphs.begin() + long(i));
    }
    ~phx_batch() {
        while (!graphs.empty()) dropGraph(graphs.size() - 1);
    }
};

namespace phx {

EpochArgs epochArgs(const phx_ctx& c);
// Order blocks per wave, balanced on term counts with relative shares.
// SH order blocks [out[w], out[w + 1]) of the W waves of a tile, balanced by
// terms with the per-wave shares of the fused (split = false) or the split SH
// kernel.
void orderBlocks(const phx_ctx& c, int W, int out[kMaxWaves + 1], bool split);
// Split SH order blocks over G workgroups of W waves (out: G W + 1 entries);
// the other families sit on waves W - 1 .. W - 4 of the first workgroup.
void splitOrderBlocks(const phx_ctx& c, int W, int G, int* out);
// Target validation of driver.cpp:99-112 (throws std::invalid_argument).
// Throws unless the Chebyshev ephemeris and the IERS2000 orientation (when
// configured) cover ET [lo, hi].
void checkEphemerisCoverage(const phx_ctx& c, double lo, double hi);
bool needsCoverageCheck(const phx_ctx& c);

void validateTargets(size_t n, const int64_t* id, const double* epoch, const double* state6, const double* mass,
                     const double* drag, const double* srp, const double* cd, const double* cr);
// Context-wide step-kernel arguments (tables, weather, wave split); the caller
// fills the per-batch buffers.
StepArgs contextStepArgs(const phx_ctx& c, int& waves);
// Looks up this device's address of the fixed scratch array that
// phx_propagate_noalloc works in (once per device, that device current).
int noallocInit();
// Steps an object needs to reach its end (propagators/FixedStep.h:7-37).
int stepsToEnd(double startEt, double endEt, double step);

}  // namespace phx
