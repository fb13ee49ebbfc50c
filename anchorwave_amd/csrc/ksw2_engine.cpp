// ksw2_engine.cpp -- continuous batching on one device.
//
// Tasks (pairs) wait in a host queue and are admitted by their device
// footprint (direction bytes, state, CIGAR slots) into a pool carved out of
// one hipMalloc, and by waves: the admitted tasks never need more waves than
// the device holds at once, so every launch starts at once and no task waits
// inside the device behind another launch's pending workgroups.  Admitted
// tasks are grouped into plans ("flights") of one size class and one team
// size (one launch) that run concurrently on a few streams.  A task's team
// follows the load: the per-wave step budget T is the work queued and in
// flight over the device's waves, and a pair gets more waves while its steps
// per wave exceed 2T, so at light load a pair runs on up to one wave per
// pass (lowest latency) and at heavy load on one wave (highest throughput).
// Each stripe-kernel pair is delivered the moment its completion flag shows;
// a flight's memory returns when its last kernel completes.  When the task at
// the head of the queue does not fit, later tasks that do are admitted
// instead (backfill), so the device keeps working while memory drains.
//
// Two users: ksw_extd2_hip_batch for batches larger than the memory budget
// (queue ordered longest first), and the serving API (first come, first
// served).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ksw2_hip.h"
#include "ksw2_carve.h"
#include "ksw2_common.h"
#include "ksw2_engine.h"
#include "ksw2_kernels.h"
#include "ksw2_plan.h"

namespace kswhip {

namespace {

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The device block of a batch engine, kept for the process's next batch
// call: allocating and releasing hundreds of GB costs seconds per call,
// and AnchorWave calls the batch aligner many times.
struct CachedBlock {
    std::mutex mu;
    uint8_t *base = nullptr;
    uint64_t bytes = 0;
    bool busy = false;
};
CachedBlock g_cached;

// First-fit allocator over one device allocation (RangeCarver's bookkeeping
// under a lock): blocks are 256-aligned, a block may give back parts early,
// and freed runs coalesce.
class FirstFitPool : public DevicePool {
public:
    int init(uint64_t n, bool cached)
    {
        bytes = round_up(n, 256);
        if (cached) {
            std::lock_guard<std::mutex> lk(g_cached.mu);
            if (!g_cached.busy && g_cached.base && g_cached.bytes < bytes) {
                (void)hipFree(g_cached.base);  // too small for this budget
                g_cached.base = nullptr, g_cached.bytes = 0;
            }
            if (!g_cached.busy && !g_cached.base) {
                if (hipMalloc(&g_cached.base, bytes) != hipSuccess) g_cached.base = nullptr;
                else g_cached.bytes = bytes;
            }
            if (!g_cached.busy && g_cached.base) {
                g_cached.busy = true, cached_ = true;
                base = g_cached.base, bytes = g_cached.bytes;
                carve_.reset(bytes);
                return KSW2_HIP_OK;
            }
        }
        if (hipMalloc(&base, bytes) != hipSuccess) {
            base = nullptr;
            return KSW2_HIP_ENOMEM;
        }
        carve_.reset(bytes);
        return KSW2_HIP_OK;
    }
    ~FirstFitPool() override
    {
        if (cached_) {
            std::lock_guard<std::mutex> lk(g_cached.mu);
            g_cached.busy = false;
        } else if (base) {
            (void)hipFree(base);
        }
    }
    bool alloc(uint64_t n, uint64_t *off) override
    {
        std::lock_guard<std::mutex> lk(mu_);
        return carve_.alloc(n, off);
    }
    void free(uint64_t off) override
    {
        std::lock_guard<std::mutex> lk(mu_);
        carve_.free(off);
    }
    void release(uint64_t off, uint64_t n) override
    {
        std::lock_guard<std::mutex> lk(mu_);
        (void)carve_.release(off, n);
    }
    uint64_t largest_free()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return carve_.largest_free();
    }

private:
    bool cached_ = false;
    std::mutex mu_;
    RangeCarver carve_;
};

}  // namespace

std::unique_ptr<DevicePool> cached_block_pool(uint64_t bytes, uint64_t need)
{
    auto pool = std::make_unique<FirstFitPool>();
    if (pool->init(std::max(bytes, need), true) != KSW2_HIP_OK || pool->bytes < need) return nullptr;
    return pool;
}

// Size classes keep a flight's members of similar duration, so short tasks
// are not held back by long ones that share their plan.
static int size_class(double cells) { return cells < 3e7 ? 0 : cells < 5e8 ? 1 : 2; }

struct Engine::Task {
    uint64_t ticket = 0;
    ksw2_pair_t pair{};
    std::vector<uint8_t> qcopy, tcopy;  // owned copies (serving) or empty
    ksw_extz_t *ez = nullptr;           // caller's output
    uint64_t footprint = 0;
    double cells = 0;
    double steps = 0;  // stripe kernel: passes x steps per pass (one wave)
    // the shortest time the task can take, in steps of one wave: a stripe
    // pair on its widest team (one wave per pass, at most kMaxWide) runs
    // its rounds of passes plus the start lag of the team's last wave; a
    // general pair its anti-diagonals
    double span = 0;
    int passes = 1;    // stripe kernel: most useful team size
    int waves = 1;     // general kernel: waves of its workgroup
    int held = 0;      // waves held while admitted
    int cls = 0;
    int status = 0;    // KSW2_HIP_OK, or why the device path did not run it
    bool general = false;
    bool resident = false;  // sequences in the engine's device arena at seq_off
    uint64_t seq_off = 0, seq_pos = 0;  // and the end of their range in the submission stream
    double t_submit = 0, t_start = 0, t_done = 0;
};

struct Engine::Flight {
    Plan *plan = nullptr;
    std::vector<Task *> tasks;
    std::vector<ksw2_pair_t> pairs;
    // per plan slot; a delivered task may be freed by poll at once, so its
    // entry in tasks is cleared and never read again
    std::vector<char> delivered;
    int undelivered_fast = 0;
    int stream = 0;
    double cells = 0, t_start = 0;  // for the statistics
    uint64_t bytes = 0;             // footprints admitted, less direction bytes given back at delivery
};

Engine::Engine(int8_t m, const int8_t *mat, const KswScore &sc, const ScoreClass &cls, uint64_t budget)
    : m_(m), cls_(cls), budget_(budget)
{
    make_gen_score(m, mat, sc, cls, &gsc_);
    const char *ev = getenv("KSW2_HIP_ENGINE_STATS");
    stats_ = ev ? atoi(ev) : 0;
    if (const char *mt = getenv("KSW2_HIP_MAX_TEAM")) max_team_ = std::max(1, std::min(kMaxWide, atoi(mt)));
    if (const char *bf = getenv("KSW2_HIP_BACKFILL")) backfill_ = std::max(0, atoi(bf));
    if (const char *sp = getenv("KSW2_HIP_SERVE_SPREAD")) serve_spread_ = atoi(sp);
    if (const char *bs = getenv("KSW2_HIP_BATCH_SPREAD")) batch_spread_ = atoi(bs) != 0;
    if (const char *ff = getenv("KSW2_HIP_FAIL_FLIGHTS")) fail_flights_ = std::max(0, atoi(ff));
    if (const char *pr = getenv("KSW2_HIP_PAIR_RELEASE")) pair_release_ = atoi(pr) != 0;
    if (const char *wb = getenv("KSW2_HIP_WAVE_BYTES")) wave_bytes_scale_ = atof(wb);
    if (const char *et = getenv("KSW2_HIP_ENGINE_T")) engine_team_T_ = atoi(et) != 0;
    const char *hq = getenv("GPU_MAX_HW_QUEUES");
    nstreams_ = std::max(1, std::min(kMaxStreams, hq ? atoi(hq) : 4));
    if (const char *ns = getenv("KSW2_HIP_STREAMS")) nstreams_ = std::max(1, std::min(kMaxStreams, atoi(ns)));
    for (int s = 0; s < kMaxStreams; ++s) stream_class_[s] = kStreamClass[s % 4];
    if (const char *sc = getenv("KSW2_HIP_STREAM_CLASSES"))  // e.g. "2210" (tuning)
        for (int s = 0; s < nstreams_ && sc[s]; ++s)
            if (sc[s] >= '0' && sc[s] < '0' + kClasses) stream_class_[s] = sc[s] - '0';
}

uint64_t Engine::cached_block_bytes()
{
    std::lock_guard<std::mutex> lk(g_cached.mu);
    return g_cached.base && !g_cached.busy ? g_cached.bytes : 0;
}

void Engine::release_cached_block()
{
    std::lock_guard<std::mutex> lk(g_cached.mu);
    if (g_cached.base && !g_cached.busy) {
        (void)hipFree(g_cached.base);
        g_cached.base = nullptr, g_cached.bytes = 0;
    }
}

int Engine::start(bool cached, bool resident_seq)
{
    int st = device_init(-1);
    if (st != KSW2_HIP_OK) return st;
    if (hipGetDevice(&device_) != hipSuccess) return KSW2_HIP_EDEVICE;
    auto *pool = new FirstFitPool();
    pool_.reset(pool);
    st = pool->init(budget_, cached);
    budget_ = pool->bytes;  // a cached block may be larger than asked
    if (st != KSW2_HIP_OK) return st;
    if (resident_seq) {
        // a sixteenth of the pool, at most 512 MiB (about 50k tasks of 5 kbp);
        // a pool under 16 MiB keeps none (flights then upload their own)
        SeqArena &A = arena_;
        A.bytes = std::min<uint64_t>(budget_ / 16, 512ull << 20) / 256 * 256;
        if (A.bytes >= (1u << 20)) {
            if (!pool->alloc(A.bytes, &A.pool_off)) return KSW2_HIP_ENOMEM;
            if (hipHostMalloc((void **)&A.ring, A.bytes, hipHostMallocDefault) != hipSuccess) return KSW2_HIP_ENOMEM;
            if (hipStreamCreateWithFlags(&A.up, hipStreamNonBlocking) != hipSuccess) return KSW2_HIP_EDEVICE;
            for (auto &ev : A.ev)
                if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return KSW2_HIP_EDEVICE;
            A.dev = pool->base + A.pool_off;
            budget_ -= A.bytes;
        } else {
            A.bytes = 0;
        }
    }
    // the upload stream (created first) takes a hardware queue of its own:
    // flights wait for its copies, so it must not queue behind a flight's
    // kernels on a shared one
    if (arena_.up && nstreams_ > 1) --nstreams_;
    for (int s = 0; s < nstreams_; ++s)
        if (hipStreamCreateWithFlags(&streams_[s], hipStreamNonBlocking) != hipSuccess) return KSW2_HIP_EDEVICE;
    resident_waves_ = std::max(1, stripe_resident_waves(kGroupLeft));
    if (stats_) fprintf(stderr, "[ksw2 engine] streams %d resident waves %d\n", nstreams_, resident_waves_);
    if (hipHostMalloc((void **)&h_stop_, sizeof(int32_t), hipHostMallocMapped) != hipSuccess) return KSW2_HIP_ENOMEM;
    *h_stop_ = 0;
    if (hipHostGetDevicePointer((void **)&d_stop_, h_stop_, 0) != hipSuccess) return KSW2_HIP_EDEVICE;
    worker_ = std::thread([this] { loop(); });
    return KSW2_HIP_OK;
}

Engine::~Engine()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        quit_ = true;
    }
    cv_work_.notify_all();
    if (worker_.joinable()) worker_.join();
    for (auto &f : flights_) {
        (void)plan_wait(f->plan);
        plan_destroy(f->plan);
    }
    for (auto &s : streams_)
        if (s) wait_slots_forget(s), (void)hipStreamDestroy(s);
    for (auto *b : staging_)
        if (b) (void)hipHostFree(b);
    for (auto &m : mirror_) m.release();
    if (arena_.up) (void)hipStreamSynchronize(arena_.up), (void)hipStreamDestroy(arena_.up);
    for (auto &ev : arena_.ev)
        if (ev) (void)hipEventDestroy(ev);
    if (arena_.ring) (void)hipHostFree(arena_.ring);
    if (h_stop_) (void)hipHostFree(h_stop_);
    for (auto &kv : tasks_) delete kv.second;
    if (stats_)
        fprintf(stderr,
                "[ksw2 engine] flights %lld tasks %lld | host s: create %.3f launch %.3f fetch+deliver %.3f | "
                "device s (sum over flights): %.3f\n",
                (long long)st_flights_, (long long)st_tasks_, st_create_, st_launch_, st_fetch_, st_device_);
    if (stats_ && !st_wait_.empty()) {
        auto med = [](std::vector<float> v) {
            std::nth_element(v.begin(), v.begin() + (long)(v.size() / 2), v.end());
            return v[v.size() / 2] * 1e3;
        };
        fprintf(stderr, "[ksw2 engine] per task p50 ms: queued %.2f, launch to delivery %.2f (%zu tasks)\n",
                med(st_wait_), med(st_serv_), st_wait_.size());
    }
}

bool Engine::queues_empty() const
{
    for (const auto &q : queue_)
        if (!q.empty()) return false;
    return true;
}

// A task leaves the pending work (its result was delivered or it failed).
// Sequence bytes of a task in the arena: the query, zero padding to a
// 16-byte boundary plus 16, then the target likewise.
static uint64_t seq_span(int qlen, int tlen) { return (uint64_t)(qlen + 31) / 16 * 16 + (uint64_t)(tlen + 31) / 16 * 16; }

bool Engine::stage_seq_locked(Task *t)
{
    SeqArena &A = arena_;
    const int ql = t->pair.qlen, tl = t->pair.tlen;
    const uint64_t n = seq_span(ql, tl), tq = (uint64_t)(ql + 31) / 16 * 16;
    uint64_t p = A.tail;
    if (p % A.bytes + n > A.bytes) p += A.bytes - p % A.bytes;  // never straddle the end
    if (p + n - A.head > A.bytes) return false;                  // the ring is full
    uint8_t *h = A.ring + p % A.bytes;
    memcpy(h, t->pair.query, (size_t)ql);
    memset(h + ql, 0, (size_t)(tq - ql));
    memcpy(h + tq, t->pair.target, (size_t)tl);
    memset(h + tq + tl, 0, (size_t)(n - tq - tl));
    A.tail = p + n;
    A.live.emplace_back(p, p + n);
    ++resident_tasks_;
    t->seq_off = p % A.bytes, t->seq_pos = p + n, t->resident = true;
    return true;
}

uint64_t Engine::upload_seq()
{
    std::lock_guard<std::mutex> lk(seq_mu_);
    SeqArena &A = arena_;
    if (!A.dev) return 0;
    // completed uploads, oldest first
    while (!A.pending.empty() && hipEventQuery(A.ev[A.pending.front().second]) == hipSuccess)
        A.done = A.pending.front().first, A.pending.pop_front();
    if (A.flushed < A.tail && (int)A.pending.size() < kUpEvents) {
        // what was written since the last upload: one copy, two when it wraps
        bool ok = true;
        uint64_t from = A.flushed;
        while (ok && from < A.tail) {
            const uint64_t o = from % A.bytes, len = std::min(A.tail - from, A.bytes - o);
            ok = hipMemcpyAsync(A.dev + o, A.ring + o, len, hipMemcpyHostToDevice, A.up) == hipSuccess;
            from += len;
        }
        const int e = A.next_ev;
        if (ok && hipEventRecord(A.ev[e], A.up) == hipSuccess) {
            A.pending.emplace_back(A.tail, e);
            A.next_ev = (e + 1) % kUpEvents;
            A.flushed = A.tail;
        }
    }
    return A.done;
}

void Engine::free_seq(Task *t)
{
    if (!t->resident) return;
    std::lock_guard<std::mutex> lk(seq_mu_);
    SeqArena &A = arena_;
    for (auto &r : A.live)
        if (r.second == t->seq_pos) {
            r.second = 0;
            break;
        }
    // the head passes the freed ranges at the front; it never passes a
    // range whose upload may still read the host ring
    while (!A.live.empty() && A.live.front().second == 0) {
        A.live.pop_front();
        A.head = A.live.empty() ? std::min(A.tail, A.done) : A.live.front().first;
    }
    t->resident = false;
}

void Engine::delivered_locked(const Task *t)
{
    free_seq(const_cast<Task *>(t));
    pending_cells_ -= t->cells, since_reap_cells_ += t->cells;
    if (pending_cells_ < 0) pending_cells_ = 0;
}

Engine::Load Engine::load()
{
    std::lock_guard<std::mutex> lk(mu_);
    Load l;
    l.budget = budget_, l.held_bytes = held_bytes_, l.queued_bytes = queued_bytes_;
    l.pending = pending_, l.pending_cells = pending_cells_, l.cells_per_s = cells_per_s_;
    l.staged_flights = staged_flights_;
    {
        std::lock_guard<std::mutex> sl(seq_mu_);
        l.seq_arena_bytes = arena_.bytes, l.resident_tasks = resident_tasks_;
    }
    return l;
}

int Engine::submit(const ksw2_pair_t &p, ksw_extz_t *ez, bool copy, uint64_t *ticket, int *defer_ms)
{
    if (defer_ms && p.qlen > 0 && p.tlen > 0 && p.query && p.target && p.qlen <= 0x7fffffff - p.tlen + 1) {
        const uint64_t fp = pair_footprint(p, takes_general(p, m_, cls_));
        std::lock_guard<std::mutex> lk(mu_);
        if (fp + kPlanSlack <= budget_ && held_bytes_ + queued_bytes_ + fp + kPlanSlack > budget_) {
            // the GPU's memory is committed: the caller parks the task (as
            // AlignmentTaskDeferred does) for about half the time the work
            // ahead needs, 250 ms to 5 s (alignSlidingWindow.cpp:2500-2511)
            const double wait_ms = cells_per_s_ > 0 ? pending_cells_ / cells_per_s_ * 1e3 : 4000.0;
            *defer_ms = (int)std::max(250.0, std::min(5000.0, 0.5 * wait_ms));
            return KSW2_HIP_EAGAIN;
        }
    }
    auto *t = new Task();
    t->pair = p;
    t->ez = ez;
    t->t_submit = now_s();
    reset_ez(ez);
    // calls that finish at submission, as the reference finishes them
    bool finished = true;
    int why = KSW2_HIP_OK;
    if (p.qlen <= 0 || p.tlen <= 0) {
        // reference returns early
    } else if (!p.query || !p.target) {
        ez->zdropped = 1, why = KSW2_HIP_EINVAL;
    } else if (p.qlen > 0x7fffffff - p.tlen + 1) {
        ez->zdropped = 1;                                   // ksw2_extd2_avx.c:138-141
    } else {
        t->general = takes_general(p, m_, cls_);
        t->footprint = pair_footprint(p, t->general);
        if (t->footprint + kPlanSlack > budget_) ez->zdropped = 1, why = KSW2_HIP_ENOMEM;  // never fits
        else finished = false;
    }
    if (!finished) {
        if (copy) {
            t->qcopy.assign(p.query, p.query + p.qlen);
            t->tcopy.assign(p.target, p.target + p.tlen);
            t->pair.query = t->qcopy.data(), t->pair.target = t->tcopy.data();
            if (arena_.dev) {
                std::lock_guard<std::mutex> lk(seq_mu_);
                (void)stage_seq_locked(t);  // else the flight uploads it
            }
        }
        t->cells = (double)p.qlen * p.tlen;
        t->cls = size_class(t->cells);
        if (t->general) {
            t->waves = gen_threads(p) / kWave;
            t->span = (double)p.qlen + p.tlen - 1;
        } else {
            const StripeGeom g = stripe_geom(p);
            t->passes = pass_count(g.kt);
            t->steps = (double)t->passes * (double)pass_steps(g.kq);
            const int team = std::min(t->passes, kMaxWide);
            t->span = (double)((t->passes + team - 1) / team) * (double)pass_steps(g.kq) + (team - 1) * 2.0 * kWave;
        }
    }
    std::lock_guard<std::mutex> lk(mu_);
    t->ticket = ++next_ticket_;
    tasks_[t->ticket] = t;
    if (finished) {
        t->t_done = t->t_submit;
        t->status = why;
        done_.push_back(t->ticket);
        if (why != KSW2_HIP_OK && status_ == KSW2_HIP_OK) status_ = why;
        cv_done_.notify_all();
    } else {
        queue_[t->cls].push_back(t);
        queued_steps_ += t->steps;
        queued_bytes_ += t->footprint, pending_cells_ += t->cells;
        ++pending_;
    }
    if (ticket) *ticket = t->ticket;
    cv_work_.notify_all();
    return KSW2_HIP_OK;
}

void Engine::hold(bool on)
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        hold_ = on;
    }
    cv_work_.notify_all();
}

void Engine::order_longest_first()
{
    std::lock_guard<std::mutex> lk(mu_);
    batch_ = true;
    // longest critical path first (then most cells): a pair whose chain is
    // long must start early, whatever its cell count
    for (auto &q : queue_)
        std::stable_sort(q.begin(), q.end(), [](const Task *a, const Task *b) {
            return a->span != b->span ? a->span > b->span : a->cells > b->cells;
        });
}

int Engine::status()
{
    std::lock_guard<std::mutex> lk(mu_);
    return status_;
}

int Engine::poll(uint64_t *tickets, double *latency, int max, int wait_ms, int32_t *status)
{
    std::unique_lock<std::mutex> lk(mu_);
    if (done_.empty() && wait_ms != 0) {
        auto pred = [this] { return !done_.empty(); };
        if (wait_ms < 0) cv_done_.wait(lk, pred);
        else cv_done_.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
    }
    int n = 0;
    while (n < max && !done_.empty()) {
        const uint64_t tk = done_.front();
        done_.pop_front();
        if (tickets) tickets[n] = tk;
        ++n;
        auto it = tasks_.find(tk);
        if (it != tasks_.end()) {
            if (latency) latency[n - 1] = it->second->t_done - it->second->t_submit;
            if (stats_ && it->second->t_start > 0) {
                st_wait_.push_back((float)(it->second->t_start - it->second->t_submit));
                st_serv_.push_back((float)(it->second->t_done - it->second->t_start));
            }
            if (status) status[n - 1] = it->second->status;
            delete it->second;
            tasks_.erase(it);
        } else {
            if (latency) latency[n - 1] = 0;
            if (status) status[n - 1] = KSW2_HIP_OK;
        }
    }
    return n;
}

int Engine::drain(const Progress *prog)
{
    int calls = 0;
    bool stopped = false;
    std::unique_lock<std::mutex> lk(mu_);
    while (pending_ > 0) {
        if (prog && prog->cb && !stopped) {
            lk.unlock();
            const bool go = progress_continue(prog, calls, 1);
            lk.lock();
            if (!go) {
                stopped = true;
                stop_all_locked();
            }
        }
        cv_done_.wait_for(lk, std::chrono::microseconds(500));
    }
    // forget completed tickets (batch mode has no poller)
    for (uint64_t tk : done_) {
        auto it = tasks_.find(tk);
        if (it != tasks_.end()) delete it->second, tasks_.erase(it);
    }
    done_.clear();
    return status_;
}

// The progress callback refused: queued tasks end as the reference's stop at
// anti-diagonal 0; running plans get their stop flag raised.
void Engine::stop_all_locked()
{
    for (auto &q : queue_) {
        for (Task *t : q) {
            t->ez->stopped = 1;
            t->t_done = now_s();
            done_.push_back(t->ticket);
            --pending_;
            delivered_locked(t);
        }
        q.clear();
    }
    queued_steps_ = 0, queued_bytes_ = 0;
    for (auto &f : flights_) f->plan->stop_requested = true;
    if (h_stop_) __atomic_store_n(h_stop_, 1, __ATOMIC_SEQ_CST);
    cv_done_.notify_all();
}

void Engine::loop()
{
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        if (quit_ && queues_empty() && flights_.empty()) break;
        const bool reaped = reap_locked(lk);
        const bool admitted = admit_locked(lk);
        if (reaped || admitted) continue;
        if (flights_.empty()) {
            if (quit_) break;
            cv_work_.wait_for(lk, std::chrono::milliseconds(5));
        } else {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(100));
            lk.lock();
        }
    }
}

// Completed flights: results to the callers' ez, memory back to the pool.
bool Engine::reap_locked(std::unique_lock<std::mutex> &lk)
{
    bool any = false;
    for (size_t i = 0; i < flights_.size();) {
        Flight *f = flights_[i].get();
        Plan *p = f->plan;
        // pairs the stripe kernel has finished (completion flag in the host
        // mirror) are delivered now, while the rest of the flight runs
        if (p->h_mflag && f->undelivered_fast > 0) {
            const double t = now_s();
            bool got = false;
            for (int k = 0; k < p->n_fast; ++k) {
                if (f->delivered[k] || !__atomic_load_n(p->h_mflag + k, __ATOMIC_ACQUIRE)) continue;
                Task *tk = f->tasks[p->order[k]];
                // the pair's kernels are done with its direction bytes: give
                // them back now rather than when the whole flight ends
                uint64_t doff, dlen;
                if (pair_release_ && plan_dir_span(p, k, &doff, &dlen)) {
                    pool_->release(doff, dlen);
                    dlen = dlen / 256 * 256;
                    f->bytes = f->bytes > dlen ? f->bytes - dlen : 0;
                    held_bytes_ = held_bytes_ > dlen ? held_bytes_ - dlen : 0;
                }
                const int st = deliver_slot(p, k, p->h_mout[k], p->h_mcig, tk->ez, nullptr);
                if (st != KSW2_HIP_OK) tk->ez->zdropped = 1, tk->status = st, status_ = st;
                f->delivered[k] = 1, --f->undelivered_fast;
                f->tasks[p->order[k]] = nullptr;
                busy_waves_ -= tk->held, inflight_steps_ -= tk->steps;
                delivered_locked(tk);
                tk->t_done = t;
                done_.push_back(tk->ticket);
                --pending_;
                got = true;
            }
            if (got) cv_done_.notify_all(), any = true;
        }
        if (!plan_done(p)) {
            ++i;
            continue;
        }
        std::unique_ptr<Flight> own = std::move(flights_[i]);
        flights_.erase(flights_.begin() + (long)i);
        busy_[f->stream] = false;
        lk.unlock();
        const double t_r0 = now_s();
        int st = plan_wait(p);
        std::vector<ksw_extz_t> ez(f->tasks.size());
        for (size_t k = 0; k < f->tasks.size(); ++k)
            if (f->tasks[k]) ez[k] = *f->tasks[k]->ez;
        // the rest (general-kernel pairs, or every pair without a mirror)
        if (st == KSW2_HIP_OK) st = plan_fetch(p, ez.data(), nullptr, &f->delivered);
        const float dev_ms = p->last_dp_ms + p->last_bt_ms;
        const double t = now_s();
        lk.lock();
        if (stats_) {
            st_fetch_ += t - t_r0, st_device_ += dev_ms * 1e-3, ++st_flights_, st_tasks_ += f->tasks.size();
            if (stats_ > 1) {
                int maxteam = 0;
                for (const auto &l : p->launches) maxteam = std::max(maxteam, l.team);
                fprintf(stderr, "[ksw2 engine] flight stream %d tasks %zu cells %.3g dp %.1f ms bt %.1f ms "
                        "fast %d gen %d maxteam %d rows %d wide %d wall %.1f ms\n", f->stream, f->tasks.size(),
                        f->cells, p->last_dp_ms, p->last_bt_ms, p->n_fast, p->n_gen, maxteam, p->R, p->n_wide,
                        (t - f->t_start) * 1e3);
            }
        }
        for (int k = 0; k < p->n; ++k) {
            if (f->delivered[k]) continue;
            Task *tk = f->tasks[p->order[k]];
            ksw_extz_t &z = ez[p->order[k]];
            if (st == KSW2_HIP_OK) {
                *tk->ez = z;
            } else {
                // keep the (possibly reallocated) CIGAR buffer with the caller
                tk->ez->cigar = z.cigar, tk->ez->m_cigar = z.m_cigar;
                tk->ez->zdropped = 1, tk->status = st, status_ = st;
            }
            tk->t_done = t;
            done_.push_back(tk->ticket);
            --pending_;
            busy_waves_ -= tk->held, inflight_steps_ -= tk->steps;
            delivered_locked(tk);
        }
        held_bytes_ = held_bytes_ > f->bytes ? held_bytes_ - f->bytes : 0;
        // throughput under load: cells delivered since the last completion
        // over the time since it
        if (last_reap_t_ > 0 && t > last_reap_t_ && since_reap_cells_ > 0) {
            const double rate = since_reap_cells_ / (t - last_reap_t_);
            cells_per_s_ = cells_per_s_ > 0 ? 0.8 * cells_per_s_ + 0.2 * rate : rate;
        }
        last_reap_t_ = pending_ > 0 ? t : 0, since_reap_cells_ = 0;
        if (busy_waves_ < 0 || flights_.empty()) busy_waves_ = 0;
        if (inflight_steps_ < 0 || flights_.empty()) inflight_steps_ = 0;
        plan_destroy(p);
        cv_done_.notify_all();
        any = true;
    }
    return any;
}

// Team size of a queued task under the per-wave step budget T, at most
// `cap` waves and one wave per pass; teams above kMaxTeam are wide (whole
// workgroups of kWideGroup waves); 0 for a general-kernel task (those form
// flights of their own).
static int team_for(int passes, double steps, bool general, double T, int cap)
{
    if (general) return 0;
    int team = 1;
    while (team < kMaxWide && team < passes && 2 * team <= cap && steps / team > 2 * T) team <<= 1;
    team = std::min(team, passes);
    return team > kMaxTeam ? (team + kWideGroup - 1) / kWideGroup * kWideGroup : team;
}

// Forms and launches flights while streams, waves and memory allow: each free
// stream takes the waiting tasks of its size class, in queue order, as far as
// pool memory and the waves allow; tasks that do not fit are passed over for
// later ones that do (backfill).
//  * Serving: the flights in the device together hold at most the device's
//    waves, and a flight's tasks share the head's team size (one launch) and
//    run within a factor of two of the head's steps per wave, so a stream is
//    not held long after most of its flight is done.
//  * Batch (longest first, latency irrelevant): each flight may hold a
//    device-full of waves, with the plan choosing each pair's team, so the
//    device always has workgroups waiting to fill the tail of a flight.

bool Engine::admit_locked(std::unique_lock<std::mutex> &lk)
{
    bool any = false;
    if (hold_) return false;
    for (int s = 0; s < nstreams_; ++s) {
        if (busy_[s]) continue;
        const int64_t free_waves = batch_ ? resident_waves_ : resident_waves_ - busy_waves_;
        if (free_waves <= 0) break;  // the device is full: wait for pairs to complete
        // a stream serves its own class first; the long-task streams also take
        // shorter tasks when no long ones wait, the others never take long ones
        static const int pref[kClasses][kClasses] = {{0, 1, -1}, {1, 0, -1}, {2, 1, 0}};
        const int own = batch_ ? kBatchStreamClass[s % 4] : stream_class_[s];
        int cls = -1;
        for (int k = 0; k < kClasses && pref[own][k] >= 0; ++k)
            if (!queue_[pref[own][k]].empty()) {
                cls = pref[own][k];
                break;
            }
        if (cls < 0) continue;
        std::deque<Task *> &q = queue_[cls];
        auto *pool = static_cast<FirstFitPool *>(pool_.get());
        const uint64_t share = (uint64_t)(kFlightShare[cls] * (double)budget_);
        const uint64_t free_block = pool->largest_free();
        const uint64_t avail = std::min(free_block, share);
        // per-wave step budget: work queued and in flight over the device's waves
        const double T = std::max(1.0, (queued_steps_ + inflight_steps_) / resident_waves_);
        const int cap = (int)std::min<int64_t>(free_waves, max_team_);
        const Task *head = q.front();
        const int key = batch_ ? (head->general ? 0 : -1) : team_for(head->passes, head->steps, head->general, T, cap);
        const double span = key > 0 ? head->steps / key : 0;  // the head's steps per wave
        std::vector<Task *> pick;
        uint64_t bytes = kPlanSlack, seq = 0, cigw = 0;
        int64_t waves = 0;
        double steps = 0;
        int skipped = 0;
        for (auto it = q.begin(); it != q.end() && (int)pick.size() < kMaxFlightPairs;) {
            Task *t = *it;
            // serving: one team size (one launch) and similar durations per
            // flight; batch: general and stripe pairs in flights of their own
            const bool same = batch_ ? (key == 0) == t->general
                                     : team_for(t->passes, t->steps, t->general, T, cap) == key &&
                                           (key == 0 || (t->steps / key <= 2 * span && 2 * t->steps / key >= span));
            if (!same) {
                if (++skipped > backfill_) break;  // backfill window
                ++it;
                continue;
            }
            const int w = key > 0 ? key : key < 0 ? (t->cls == 2 ? std::min(kMaxWide, t->passes) : 1) : t->waves;
            if (!pick.empty() && waves + w > free_waves) break;
            // a flight holds at most its class's share of the pool, except that
            // its first task may use any free block (a task larger than the share)
            if (bytes + t->footprint > (pick.empty() ? free_block : avail)) {
                if (++skipped > backfill_) break;
                ++it;
                continue;
            }
            pick.push_back(t);
            t->held = w;
            bytes += t->footprint, waves += w, steps += t->steps, seq += (uint64_t)t->pair.qlen + t->pair.tlen + 32;
            cigw += ((uint64_t)t->pair.qlen + t->pair.tlen + 2 + 15) / 16 * 16;
            it = q.erase(it);
        }
        if (pick.empty()) {  // wait for memory
            if (stats_ > 1)
                fprintf(stderr, "[ksw2 engine] t %.3f stream %d class %d: nothing fits (free block %.2f GB, head %.2f GB, skipped %d)\n",
                        now_s(), s, cls, free_block / 1e9, q.front()->footprint / 1e9, skipped);
            continue;
        }
        // Serving flights whose teams are capped by the pairs' pass count
        // (the device has waves to spare) run four rows per lane: twice the
        // passes, so twice the waves share each pair's sweep, while the team
        // stays within one workgroup (kMaxTeam waves) and the flight within a
        // quarter of the free waves.  Team keys and passes are in eight-row
        // passes, so a flight with a forced team runs eight rows otherwise.
        const bool light = queues_empty() && 8 * (busy_waves_ + waves) <= resident_waves_;
        int rows = key > 0 ? 8 : 0, team = std::max(key, 0);
        if (!batch_ && key > 0 && key == head->passes && 2 * key <= kMaxTeam && 4 * waves <= free_waves) {
            rows = 4, team = 2 * key, waves *= 2;
            for (Task *t : pick) t->held *= 2;
        }
        // pinned upload buffer of this stream, grown to the flight's sequences
        if (staging_bytes_[s] < seq) {
            if (staging_[s]) (void)hipHostFree(staging_[s]);
            staging_[s] = nullptr, staging_bytes_[s] = 0;
            const uint64_t want = std::max<uint64_t>(seq * 2, 1 << 20);
            if (hipHostMalloc((void **)&staging_[s], want, hipHostMallocDefault) == hipSuccess) staging_bytes_[s] = want;
        }
        // host mirror of this stream: result slots, CIGAR words, flags
        Mirror &mr = mirror_[s];
        if (mr.slots < pick.size() || mr.words < cigw) {
            mr.release();
            const uint64_t slots = std::max<uint64_t>(pick.size() * 2, 4096);
            const uint64_t words = std::max<uint64_t>(cigw * 2, 1 << 20);
            const unsigned fl = hipHostMallocMapped | hipHostMallocCoherent;
            if (hipHostMalloc((void **)&mr.out, slots * sizeof(KswOut), fl) == hipSuccess &&
                hipHostMalloc((void **)&mr.cig, words * 4, fl) == hipSuccess &&
                hipHostMalloc((void **)&mr.flag, slots * 4, fl) == hipSuccess &&
                hipHostGetDevicePointer((void **)&mr.out_dev, mr.out, 0) == hipSuccess &&
                hipHostGetDevicePointer((void **)&mr.cig_dev, mr.cig, 0) == hipSuccess &&
                hipHostGetDevicePointer((void **)&mr.flag_dev, mr.flag, 0) == hipSuccess) {
                mr.slots = slots, mr.words = words;
            } else {
                mr.release();  // no mirror: the flight completes as a whole
            }
        }
        auto f = std::make_unique<Flight>();
        f->tasks = pick;
        f->stream = s;
        for (const Task *tk : pick) f->cells += tk->cells;
        for (Task *t : pick) f->pairs.push_back(t->pair);
        busy_[s] = true;
        lk.unlock();
        std::vector<int> idx(f->pairs.size());
        for (size_t k = 0; k < idx.size(); ++k) idx[k] = (int)k;
        PlanEnv env;
        env.pool = pool_.get();
        env.stream = streams_[s];
        env.staging = staging_[s], env.staging_bytes = staging_bytes_[s];
        env.h_stop = h_stop_, env.d_stop = d_stop_;
        env.poll_stop = true;  // stop_all_locked raises it
        env.team = team, env.rows = rows;
        if (batch_) {
            // teams by memory: each wave holds at most its share of the pool
            // (the step budget stays the flight's own unless asked otherwise)
            if (engine_team_T_) env.team_T = T;
            if (wave_bytes_scale_ > 0) env.wave_bytes = (uint64_t)(wave_bytes_scale_ * (double)budget_ / resident_waves_);
        }
        // a flight's teams may spread over CUs (tuning switch; off by
        // default: with sixteen streams a spread team gains nothing at 20 %
        // load, and under more load it holds waiting-launch slots that the
        // flights needing wide teams then queue behind); 1 spreads only while
        // the device is lightly loaded (nothing left queued, the flights in
        // the device holding under an eighth of its waves)
        env.spread = batch_ ? batch_spread_ : serve_spread_ == 2 || (serve_spread_ == 1 && light);
        if (mr.slots) {
            env.mirror_out = mr.out, env.mirror_out_dev = mr.out_dev;
            env.mirror_cig = mr.cig, env.mirror_cig_dev = mr.cig_dev;
            env.mirror_flag = mr.flag, env.mirror_flag_dev = mr.flag_dev;
        }
        // every task's sequences in the arena and uploaded: the plan reads
        // them there (else it uploads its own)
        std::vector<uint64_t> seq_offs;
        bool resident = arena_.dev != nullptr;
        uint64_t last = 0;
        for (const Task *tk : pick) resident = resident && tk->resident, last = std::max(last, tk->seq_pos);
        if (resident && last > upload_seq()) {
            // the uploads normally land within microseconds: wait a little,
            // never long (the upload stream may share a busy hardware queue)
            for (int spin = 0; spin < 20 && last > upload_seq(); ++spin)
                std::this_thread::sleep_for(std::chrono::microseconds(10));
        }
        if (resident && last <= upload_seq()) {
            seq_offs.resize(2 * pick.size());
            for (size_t k = 0; k < pick.size(); ++k) {
                seq_offs[2 * k] = pick[k]->seq_off;
                seq_offs[2 * k + 1] = pick[k]->seq_off + (uint64_t)(pick[k]->pair.qlen + 31) / 16 * 16;
            }
            env.seq_dev = arena_.dev, env.seq_offs = seq_offs.data();
        }
        const bool staged = !env.seq_dev;
        const double t_c0 = now_s();
        int st = plan_create(&f->plan, f->pairs.data(), idx.data(), (int)idx.size(), m_, cls_, gsc_, &env);
        if (st == KSW2_HIP_OK) {
            f->delivered.assign(f->plan->n, 0);
            f->undelivered_fast = f->plan->h_mflag ? f->plan->n_fast : 0;
        }
        const double t_c1 = now_s();
        if (st == KSW2_HIP_OK && fail_flights_ > 0) --fail_flights_, st = KSW2_HIP_EDEVICE;  // tests
        if (st == KSW2_HIP_OK) st = plan_launch(f->plan);
        const double t0 = now_s();
        lk.lock();
        if (stats_) st_create_ += t_c1 - t_c0, st_launch_ += t0 - t_c1;
        if (st != KSW2_HIP_OK) {
            busy_[s] = false;
            if (f->plan) plan_destroy(f->plan);
            if (st == KSW2_HIP_ENOMEM && !flights_.empty()) {
                // the estimate was short of the pool's fragmentation: requeue
                for (auto it = pick.rbegin(); it != pick.rend(); ++it) q.push_front(*it);
                continue;
            }
            queued_steps_ -= steps;
            queued_bytes_ = queued_bytes_ > bytes ? queued_bytes_ - bytes : 0;
            for (Task *t : pick) {
                t->ez->zdropped = 1, t->status = st;
                t->t_done = now_s();
                done_.push_back(t->ticket);
                --pending_;
                delivered_locked(t);
            }
            status_ = st;
            cv_done_.notify_all();
            continue;
        }
        for (Task *t : pick) t->t_start = t0;
        if (staged) ++staged_flights_;
        f->t_start = t0;
        f->bytes = bytes;
        if (stats_ > 1) {
            auto *pl = static_cast<FirstFitPool *>(pool_.get());
            fprintf(stderr, "[ksw2 engine] t %.3f admit stream %d class %d tasks %zu bytes %.2f GB free %.2f GB held %.2f GB q %zu/%zu/%zu\n",
                    t0, s, cls, pick.size(), bytes / 1e9, pl->largest_free() / 1e9, held_bytes_ / 1e9,
                    queue_[0].size(), queue_[1].size(), queue_[2].size());
        }
        held_bytes_ += bytes, queued_bytes_ = queued_bytes_ > bytes ? queued_bytes_ - bytes : 0;
        if (last_reap_t_ == 0) last_reap_t_ = t0;
        busy_waves_ += waves, inflight_steps_ += steps, queued_steps_ -= steps;
        if (queued_steps_ < 0 || queues_empty()) queued_steps_ = 0;
        flights_.push_back(std::move(f));
        any = true;
    }
    return any;
}

}  // namespace kswhip
