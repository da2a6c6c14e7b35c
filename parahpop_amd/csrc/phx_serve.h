// Continuous-batching request server over a device-resident slot pool.
//
// Requests (phx_submit) are validated and converted on the caller's thread
// and queued. One worker thread owns a HIP stream and repeats a tick:
//   1. admit queued requests in arrival order while their footprint fits:
//      a contiguous range of object slots in the pool and a segment of rows in
//      the epoch-record arena (grid records at the context's grid step,
//      covering the request's epochs); upload the states and compute the
//      request's records on the device;
//   2. launch the step kernel for `steps_per_tick` steps over every tile that
//      holds a running object (a captured hipGraph, re-captured only when the
//      set of active tiles changes);
//   3. download the requests that reached their last step, release their
//      slots and rows, and publish the results for phx_poll.
// Objects of different requests share tiles; every lane reads only its own
// state and its own request's records, so a result does not depend on what
// else runs beside it.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "phx_internal.h"

namespace phx {

// First-fit allocator over [0, capacity) (units: slots or record rows).
class RangeAllocator {
public:
    explicit RangeAllocator(long long capacity = 0);
    long long alloc(long long len);              // start, or -1 when nothing fits
    void release(long long start, long long len);
    long long capacity() const { return capacity_; }
    long long used() const { return used_; }

private:
    std::map<long long, long long> free_;        // start -> length
    long long capacity_ = 0, used_ = 0;
};

class Server {
public:
    explicit Server(phx_ctx& c);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Validates and queues a request; throws std::invalid_argument (bad input)
    // or std::length_error (larger than the whole pool).
    uint64_t submit(const phx_request& r);
    // PHX_OK and the result, PHX_EAGAIN while queued or running, or the
    // request's error code (message in `err`). A delivered result is dropped.
    int poll(uint64_t ticket, phx_result* out, std::string& err);
    double stat(const std::string& key) const;

private:
    struct Request {
        uint64_t ticket = 0;
        size_t n = 0;
        std::vector<double> y, start, end, dt0, ams, bcd;   // SoA inputs (ET s, km, km/s)
        std::vector<int> perm;         // slot order: slot0 + j holds object perm[j]
        int steps = 0;                 // steps until every object has ended (capped by maxNumSteps)
        double gridT0 = 0.0;
        long long rows = 0;            // epoch-record rows of the request's segment
        long long slot0 = -1, rec0 = -1, padEnd = 0;
        int stepsRun = 0;
        int rc = PHX_OK;
        std::string error;
        std::vector<double> outEpoch, outState;
        std::vector<double> dlY, dlCur;
        std::vector<uint8_t> dlSt;
        double submitT = 0.0, admitT = 0.0, finishT = 0.0;
    };

    void loop();
    void init();
    bool admitOne(Request& r);
    void rebuildTiles();
    void tick();
    void failAll(int rc, const std::string& msg);

    phx_ctx& c_;
    // step shape read once from the context's options when the server is
    // created, so later option changes (phx_set_option, phx_batch_autotune on
    // the caller's thread) never race the worker
    StepArgs base_{};                  // context-wide step arguments and order blocks
    int waves_ = 1;                    // waves per tile (split SH kernel or fused kernel)
    bool split_ = false;               // split stages (SH kernel, drag kernel, combine)
    int stepsPerTick_ = 1;
    const double h_;                   // grid step (s)
    const long long slots_;            // pool slots (multiple of 64)
    const long long recRows_;          // arena rows
    RangeAllocator slotAlloc_, rowAlloc_;
    long long highWater_ = 0;          // slots below this were written at least once

    // device pool (SoA over slots_)
    DBuf<double> y_, cur_, end_, nextDt_, ams_, bcd_, slotT0_, records_, epochStage_;
    DBuf<uint8_t> status_;
    DBuf<long long> slotRecord_, counters_;
    DBuf<int> slotRows_, tileList_;
    std::vector<int> tileListHost_;
    int activeTiles_ = 0;
    bool tilesDirty_ = true;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    int graphTiles_ = -1, graphWaves_ = -1, graphSteps_ = -1, graphSplit_ = -1;
    // split-stage buffers over the pool (allocated on first use)
    DBuf<double> ks_, ys_, parts_, lane_, trig_;

    // shared with callers
    mutable std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::unique_ptr<Request>> queue_;
    std::map<uint64_t, std::unique_ptr<Request>> done_;
    std::map<uint64_t, int> inFlight_;   // ticket -> 0 queued, 1 running
    uint64_t nextTicket_ = 1;
    bool stop_ = false;
    bool broken_ = false;
    std::string brokenMsg_;
    std::thread worker_;
    bool started_ = false;

    // worker-only
    std::vector<std::unique_ptr<Request>> running_;
    long long lastCounter_ = 0;
    // statistics (read under mu_)
    double objectSteps_ = 0.0, kernelSeconds_ = 0.0;
    long long ticks_ = 0, captures_ = 0, completed_ = 0, activeObjects_ = 0;
};

double monotonicSeconds();

}  // namespace phx
