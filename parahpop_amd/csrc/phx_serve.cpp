// Continuous-batching request server (see phx_serve.h for the tick structure).
// Compiled with -ffp-contract=off like the rest of the host code: the epoch
// and unit conversions round exactly as the batch path's.
#include "phx_serve.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <set>

namespace phx {

double monotonicSeconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

RangeAllocator::RangeAllocator(long long capacity) : capacity_(capacity) {
    if (capacity > 0) free_[0] = capacity;
}

long long RangeAllocator::alloc(long long len) {
    if (len <= 0) return -1;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < len) continue;
        const long long start = it->first, rest = it->second - len;
        free_.erase(it);
        if (rest > 0) free_[start + len] = rest;
        used_ += len;
        return start;
    }
    return -1;
}

void RangeAllocator::release(long long start, long long len) {
    if (len <= 0) return;
    used_ -= len;
    auto it = free_.emplace(start, len).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
        it->second += next->second;
        free_.erase(next);
    }
    if (it != free_.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            free_.erase(it);
        }
    }
}

namespace {
enum : uint8_t { ST_TERMINATED = 1, ST_EOS = 2, ST_PAD = 4 };

long long roundUp(long long x, long long m) { return (x + m - 1) / m * m; }
}  // namespace

Server::Server(phx_ctx& c)
    : c_(c),
      h_(c.gridStep),
      slots_(roundUp(std::max<long long>(c.serveSlots, kTile), kTile)),
      recRows_(std::max<long long>(8, (long long)(c.serveRecordMiB * 1048576.0 / (REC_SIZE * sizeof(double))))),
      slotAlloc_(slots_),
      rowAlloc_(recRows_) {
    base_ = contextStepArgs(c, waves_);
    stepsPerTick_ = std::max(1, c.serveStepsPerTick);
    // split stages (per-lane epochs with the eight-wave drag kernel) when the
    // context selects them, as for device batches
    split_ = c.split && c.fm.drag && c.fm.hasSh && c.splitSpecies == 8;
    if (split_) {
        waves_ = c.splitWaves;
        orderBlocks(c, waves_, base_.blockStart, true);
    }
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

// Device pool and arena (first submit, on the caller's thread).
void Server::init() {
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreate(&ev0_), "hipEventCreate");
    check(hipEventCreate(&ev1_), "hipEventCreate");
    const size_t S = size_t(slots_);
    y_.alloc(6 * S);
    cur_.alloc(S);
    end_.alloc(S);
    nextDt_.alloc(S);
    ams_.alloc(S);
    bcd_.alloc(S);
    slotT0_.alloc(S);
    status_.alloc(S);
    slotRecord_.alloc(S);
    slotRows_.alloc(S);
    tileList_.alloc(S / kTile);
    counters_.alloc(4);
    records_.alloc(size_t(recRows_) * REC_SIZE);
    check(hipMemsetAsync(counters_.p, 0, 4 * sizeof(long long), stream_), "hipMemsetAsync");
    check(hipStreamSynchronize(stream_), "serve pool");
}

uint64_t Server::submit(const phx_request& q) {
    if (!q.id || !q.epoch_mjd2000_tdb || !q.state6_soa || !q.mass_kg || !q.drag_area_m2 || !q.srp_area_m2 || !q.cd || !q.cr)
        throw std::invalid_argument("phx_request has a null array.");
    validateTargets(q.n, q.id, q.epoch_mjd2000_tdb, q.state6_soa, q.mass_kg, q.drag_area_m2, q.srp_area_m2, q.cd, q.cr);
    const double duration = q.duration_s;
    if (!(duration > 0) || !std::isfinite(duration)) throw std::invalid_argument("duration_seconds must be finite and positive.");
    const double step = c_.run.stepSeconds;
    auto r = std::make_unique<Request>();
    const size_t n = q.n;
    r->n = n;
    r->y.assign(q.state6_soa, q.state6_soa + 6 * n);
    r->start.resize(n);
    r->end.resize(n);
    r->dt0.resize(n);
    r->ams.resize(n);
    r->bcd.resize(n);
    int steps = 0;
    double lo = 0.0, hi = 0.0;
    for (size_t i = 0; i < n; ++i) {
        // FixedStep.h:7-29 and driver.cpp:117-120, as in the batch path
        const double st = q.epoch_mjd2000_tdb[i] * 86400.0 + -43200.0;
        const double en = (q.epoch_mjd2000_tdb[i] + duration / 86400.0) * 86400.0 + -43200.0;
        const double span = en - st;
        if (!(span > 0) || !std::isfinite(span) || std::ceil(span / step) >= 2147483647.0)
            throw std::invalid_argument("Duration/step cannot be represented by the integration time/index types.");
        r->start[i] = st;
        r->end[i] = en;
        r->dt0[i] = std::min(step, span);
        const double area = q.drag_area_m2[i] * 1e-6;
        const double crEff = q.cr[i] * q.srp_area_m2[i] / q.drag_area_m2[i];
        r->ams[i] = crEff * area / q.mass_kg[i];
        r->bcd[i] = q.cd[i] * area / q.mass_kg[i];
        steps = std::max(steps, stepsToEnd(st, en, step));
        lo = i == 0 ? st : std::min(lo, st);
        hi = i == 0 ? en : std::max(hi, en);
    }
    checkEphemerisCoverage(c_, lo, hi);
    r->steps = int(std::min<long long>(steps, c_.run.maxNumSteps));
    r->gridT0 = lo - 2.0 * h_;
    r->rows = (long long)std::ceil((hi - r->gridT0) / h_) + 3;
    if ((long long)n > slots_ || r->rows > recRows_)
        throw std::length_error("Request exceeds the serving pool (" + std::to_string(n) + " objects, " +
                                std::to_string(r->rows) + " epoch rows; pool " + std::to_string(slots_) + " slots, " +
                                std::to_string(recRows_) + " rows).");
    r->submitT = monotonicSeconds();
    std::unique_lock<std::mutex> lk(mu_);
    if (broken_) throw HipError(brokenMsg_);
    if (!started_) {
        lk.unlock();
        init();
        lk.lock();
        if (!started_) {
            started_ = true;
            worker_ = std::thread([this] { loop(); });
        }
    }
    r->ticket = nextTicket_++;
    const uint64_t t = r->ticket;
    inFlight_[t] = 0;
    queue_.push_back(std::move(r));
    lk.unlock();
    cv_.notify_all();
    return t;
}

int Server::poll(uint64_t ticket, phx_result* out, std::string& err) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = done_.find(ticket);
    if (it == done_.end()) {
        if (inFlight_.count(ticket)) return PHX_EAGAIN;
        err = "unknown or already delivered ticket " + std::to_string(ticket);
        return PHX_EINVAL;
    }
    Request& r = *it->second;
    const int rc = r.rc;
    if (rc != PHX_OK) {
        err = r.error;
    } else if (out) {
        out->n = r.n;
        if (out->out_epoch) std::memcpy(out->out_epoch, r.outEpoch.data(), r.n * sizeof(double));
        if (out->out_state6_soa) std::memcpy(out->out_state6_soa, r.outState.data(), 6 * r.n * sizeof(double));
    }
    if (out) {
        out->submit_time_s = r.submitT;
        out->finish_time_s = r.finishT;
    }
    done_.erase(it);
    return rc;
}

double Server::stat(const std::string& k) const {
    std::lock_guard<std::mutex> lk(mu_);
    if (k == "serve_object_steps") return objectSteps_;
    if (k == "serve_kernel_seconds") return kernelSeconds_;
    if (k == "serve_ticks") return double(ticks_);
    if (k == "serve_graph_captures") return double(captures_);
    if (k == "serve_completed") return double(completed_);
    if (k == "serve_active_objects") return double(activeObjects_);
    if (k == "serve_queued") return double(queue_.size());
    if (k == "serve_pool_slots") return double(slots_);
    if (k == "serve_record_rows") return double(recRows_);
    return std::nan("");
}

// Places one request in the pool; false when its footprint does not fit now.
bool Server::admitOne(Request& r) {
    const long long n = (long long)r.n;
    const long long s0 = slotAlloc_.alloc(n);
    if (s0 < 0) return false;
    const long long rec0 = rowAlloc_.alloc(r.rows);
    if (rec0 < 0) {
        slotAlloc_.release(s0, n);
        return false;
    }
    r.slot0 = s0;
    r.rec0 = rec0;
    // Slots of the request's last tile that were never written get a copy of
    // the request's first object, marked as padding, so every lane of an
    // active tile holds a valid state and record segment.
    const long long tileEnd = roundUp(s0 + n, kTile);
    r.padEnd = s0 + n;
    if (tileEnd > highWater_) {
        r.padEnd = tileEnd;
        highWater_ = tileEnd;
    }
    const size_t m = size_t(r.padEnd - s0);
    const size_t nn = r.n;
    // objects by radius, so those below the drag cutoff share tiles
    r.perm.resize(nn);
    {
        std::vector<double> rad(nn);
        for (size_t i = 0; i < nn; ++i) {
            const double x = r.y[i], y = r.y[nn + i], z = r.y[2 * nn + i];
            rad[i] = std::sqrt(x * x + y * y + z * z);
            r.perm[i] = int(i);
        }
        std::stable_sort(r.perm.begin(), r.perm.end(), [&](int a, int b) { return rad[size_t(a)] < rad[size_t(b)]; });
    }
    std::vector<double> y(6 * m), cur(m), end(m), dt(m), ams(m), bcd(m), t0(m, r.gridT0);
    std::vector<uint8_t> st(m);
    std::vector<long long> recIdx(m, rec0);
    std::vector<int> rows(m, int(r.rows));
    for (size_t j = 0; j < m; ++j) {
        const size_t i = size_t(r.perm[j < nn ? j : 0]);
        for (int d = 0; d < 6; ++d) y[size_t(d) * m + j] = r.y[size_t(d) * nn + i];
        cur[j] = r.start[i];
        end[j] = r.end[i];
        dt[j] = r.dt0[i];
        ams[j] = r.ams[i];
        bcd[j] = r.bcd[i];
        st[j] = j < nn ? 0 : ST_PAD;
    }
    const size_t S = size_t(slots_);
    const size_t o = size_t(s0);
    auto h2d = [&](void* dst, const void* src, size_t bytes) {
        check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_), "serve upload");
    };
    for (int d = 0; d < 6; ++d) h2d(y_.p + size_t(d) * S + o, y.data() + size_t(d) * m, m * sizeof(double));
    h2d(cur_.p + o, cur.data(), m * sizeof(double));
    h2d(end_.p + o, end.data(), m * sizeof(double));
    h2d(nextDt_.p + o, dt.data(), m * sizeof(double));
    h2d(ams_.p + o, ams.data(), m * sizeof(double));
    h2d(bcd_.p + o, bcd.data(), m * sizeof(double));
    h2d(slotT0_.p + o, t0.data(), m * sizeof(double));
    h2d(status_.p + o, st.data(), m);
    h2d(slotRecord_.p + o, recIdx.data(), m * sizeof(long long));
    h2d(slotRows_.p + o, rows.data(), m * sizeof(int));
    // the request's grid records
    std::vector<double> epochs(size_t(r.rows));
    for (long long k = 0; k < r.rows; ++k) epochs[size_t(k)] = r.gridT0 + double(k) * h_;
    if (epochStage_.n < epochs.size()) {
        check(hipStreamSynchronize(stream_), "serve staging");
        epochStage_.alloc(std::max(epochs.size(), size_t(8192)));
    }
    h2d(epochStage_.p, epochs.data(), epochs.size() * sizeof(double));
    EpochArgs ea = epochArgs(c_);
    ea.gridRows = 1;
    ea.count = int(r.rows);
    ea.epochs = epochStage_.p;
    ea.out = records_.p + size_t(rec0) * REC_SIZE;
    if (launchEpochRecords(ea, c_.fm, stream_) != 0) throw HipError("epoch-record kernel launch failed");
    // the host copies above may be pageable: finish them before the vectors go
    check(hipStreamSynchronize(stream_), "serve admission");
    r.admitT = monotonicSeconds();
    return true;
}

void Server::rebuildTiles() {
    std::set<int> tiles;
    for (const auto& r : running_) {
        const long long first = r->slot0 / kTile, last = (r->slot0 + (long long)r->n - 1) / kTile;
        for (long long t = first; t <= last; ++t) tiles.insert(int(t));
    }
    tileListHost_.assign(tiles.begin(), tiles.end());
    activeTiles_ = int(tileListHost_.size());
    if (activeTiles_)
        check(hipMemcpyAsync(tileList_.p, tileListHost_.data(), tileListHost_.size() * sizeof(int), hipMemcpyHostToDevice,
                             stream_),
              "tile list");
    tilesDirty_ = false;
}

void Server::tick() {
    const int W = waves_;
    StepArgs a = base_;
    const int K = stepsPerTick_;
    const bool split = split_;
    if (split) {
        if (!ks_.p) {
            const size_t S = size_t(slots_);
            ks_.alloc(size_t(kStages) * 6 * S);
            ys_.alloc(7 * S);
            parts_.alloc(6 * S);
            lane_.alloc(8 * S);
            trig_.alloc(6 * S);
        }
    }
    a.nTiles = activeTiles_;
    a.nPad = int(slots_);
    a.y = y_.p;
    a.cur = cur_.p;
    a.end = end_.p;
    a.nextDt = nextDt_.p;
    a.status = status_.p;
    a.ams = ams_.p;
    a.bcd = bcd_.p;
    a.records = records_.p;
    a.tileRecord = nullptr;
    a.tileList = tileList_.p;
    a.counters = counters_.p;
    a.grid = 1;
    a.gridInvH = 1.0 / h_;
    a.slotRecord = slotRecord_.p;
    a.slotT0 = slotT0_.p;
    a.slotRows = slotRows_.p;
    a.stepBase = 0;
    a.steps = K;
    if (graphTiles_ != activeTiles_ || graphWaves_ != W || graphSteps_ != K || graphSplit_ != int(split) || !exec_) {
        if (exec_) check(hipGraphExecDestroy(exec_), "hipGraphExecDestroy");
        if (graph_) check(hipGraphDestroy(graph_), "hipGraphDestroy");
        exec_ = nullptr;
        graph_ = nullptr;
        check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        int rc = 0;
        if (split) {
            // K steps of 13 stages (SH kernel, drag kernel) and the end-of-step kernel
            SplitArgs sp{ks_.p, ys_.p, parts_.p, nullptr, lane_.p, 0, 0, 8};
            sp.trig = trig_.p;
            for (int k = 0; k < K && rc == 0; ++k) {
                sp.step = k;
                for (int s = 0; s < kStages && rc == 0; ++s) {
                    sp.stage = s;
                    rc = launchSplitStage(a, c_.fm, sp, W, stream_);
                }
                if (rc == 0) rc = launchSplitCombine(a, sp, stream_);
            }
        } else {
            rc = launchStep(a, c_.fm, W, stream_);
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(stream_, &g);
        if (rc != 0 || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            throw HipError(std::string("step graph capture failed: ") + hipGetErrorString(e));
        }
        graph_ = g;
        check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
        graphTiles_ = activeTiles_;
        graphWaves_ = W;
        graphSteps_ = K;
        graphSplit_ = int(split);
        std::lock_guard<std::mutex> lk(mu_);
        ++captures_;
    }
    check(hipEventRecord(ev0_, stream_), "hipEventRecord");
    check(hipGraphLaunch(exec_, stream_), "hipGraphLaunch");
    check(hipEventRecord(ev1_, stream_), "hipEventRecord");
    long long counter = 0;
    check(hipMemcpyAsync(&counter, counters_.p, sizeof counter, hipMemcpyDeviceToHost, stream_), "counter read");
    // requests that end within this tick: download after the tick
    std::vector<Request*> finishing;
    for (auto& r : running_) {
        r->stepsRun += K;
        if (r->stepsRun >= r->steps) {
            Request& q = *r;
            const size_t m = q.n, S = size_t(slots_), o = size_t(q.slot0);
            q.dlY.resize(6 * m);
            q.dlCur.resize(m);
            q.dlSt.resize(m);
            for (int d = 0; d < 6; ++d)
                check(hipMemcpyAsync(q.dlY.data() + size_t(d) * m, y_.p + size_t(d) * S + o, m * sizeof(double),
                                     hipMemcpyDeviceToHost, stream_),
                      "serve download");
            check(hipMemcpyAsync(q.dlCur.data(), cur_.p + o, m * sizeof(double), hipMemcpyDeviceToHost, stream_), "serve download");
            check(hipMemcpyAsync(q.dlSt.data(), status_.p + o, m, hipMemcpyDeviceToHost, stream_), "serve download");
            finishing.push_back(r.get());
        }
    }
    check(hipStreamSynchronize(stream_), "serve tick");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, ev0_, ev1_), "hipEventElapsedTime");
    const double now = monotonicSeconds();
    for (Request* q : finishing) {
        q->finishT = now;
        q->outEpoch.resize(q->n);
        q->outState.resize(6 * q->n);
        bool ok = true;
        const size_t nn = q->n;
        for (size_t j = 0; j < nn; ++j) {
            const size_t i = size_t(q->perm[j]);   // slot j holds object i
            if (!(q->dlSt[j] & ST_EOS)) ok = false;
            q->outEpoch[i] = q->dlCur[j] * (1.0 / 86400.0) + 0.5;   // TimeConversion.h:36-40
            for (int d = 0; d < 6; ++d) q->outState[size_t(d) * nn + i] = q->dlY[size_t(d) * nn + j];
        }
        if (!ok) {
            q->rc = PHX_ENOTDONE;
            q->error = "A target stopped without reaching the requested end of simulation.";
        }
        slotAlloc_.release(q->slot0, (long long)q->n);
        rowAlloc_.release(q->rec0, q->rows);
    }
    std::lock_guard<std::mutex> lk(mu_);
    objectSteps_ += double(counter - lastCounter_);
    lastCounter_ = counter;
    kernelSeconds_ += ms * 1e-3;
    ++ticks_;
    if (!finishing.empty()) {
        for (auto it = running_.begin(); it != running_.end();) {
            if ((*it)->stepsRun >= (*it)->steps) {
                const uint64_t t = (*it)->ticket;
                inFlight_.erase(t);
                activeObjects_ -= (long long)(*it)->n;
                ++completed_;
                done_[t] = std::move(*it);
                it = running_.erase(it);
            } else {
                ++it;
            }
        }
        tilesDirty_ = true;
    }
}

void Server::failAll(int rc, const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu_);
    broken_ = true;
    brokenMsg_ = msg;
    auto fin = [&](std::unique_ptr<Request>& r) {
        r->rc = rc;
        r->error = msg;
        r->finishT = monotonicSeconds();
        inFlight_.erase(r->ticket);
        const uint64_t t = r->ticket;
        done_[t] = std::move(r);
    };
    for (auto& r : running_) fin(r);
    running_.clear();
    for (auto& r : queue_) fin(r);
    queue_.clear();
    activeObjects_ = 0;
}

void Server::loop() {
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !queue_.empty() || !running_.empty(); });
            if (stop_) return;
        }
        try {
            // admission in arrival order while the footprint fits
            for (;;) {
                Request* head = nullptr;
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    if (queue_.empty()) break;
                    head = queue_.front().get();
                }
                if (!admitOne(*head)) break;
                std::lock_guard<std::mutex> lk(mu_);
                std::unique_ptr<Request> r = std::move(queue_.front());
                queue_.pop_front();
                inFlight_[r->ticket] = 1;
                activeObjects_ += (long long)r->n;
                running_.push_back(std::move(r));
                tilesDirty_ = true;
            }
            if (running_.empty()) continue;
            if (tilesDirty_) rebuildTiles();
            tick();
        } catch (const std::exception& e) {
            failAll(PHX_EDEVICE, e.what());
            return;
        }
    }
}

}  // namespace phx
