// ksw2_engine.h -- continuous-batching engine (see ksw2_engine.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ksw2_hip.h"
#include "ksw2_carve.h"
#include "ksw2_common.h"
#include "ksw2_plan.h"

namespace kswhip {

class Engine {
public:
    // Flights in flight at once: one stream each, as many as the process's
    // hardware queues (HIP's GPU_MAX_HW_QUEUES, 4 by default; streams beyond
    // them would share a queue and run one after another), at most
    // kMaxStreams; KSW2_HIP_STREAMS overrides (tuning).
    static constexpr int kMaxStreams = 16;
    static constexpr int kClasses = 3;             // task size classes
    // stream s -> size class it serves first, kStreamClass[s % 4]: per four
    // streams one for the long tasks, one for the medium and two for the
    // short ones (most tasks), so short tasks never queue behind long ones
    static constexpr int kStreamClass[4] = {2, 1, 0, 0};
    // batch mode: the long tasks hold most of the work, so two streams in four
    static constexpr int kBatchStreamClass[4] = {2, 2, 1, 0};
    // share of the pool one flight of a class may hold: long pairs hold
    // memory longest per byte (their speed is capped by one workgroup), so
    // they never take the whole pool and shorter pairs keep the device full
    static constexpr double kFlightShare[kClasses] = {0.25, 0.35, 0.3};
    static constexpr int kMaxFlightPairs = 8192;
    static constexpr uint64_t kPlanSlack = 1 << 16;

    Engine(int8_t m, const int8_t *mat, const KswScore &sc, const ScoreClass &cls, uint64_t budget);
    ~Engine();
    // cached: take the process's cached device block when it is free and
    // large enough (else allocate one and leave it cached), and give it back
    // to the cache at destruction instead of freeing it
    // resident_seq: submissions copy their sequences into a device arena at
    // once (serving), so flights upload none
    int start(bool cached = false, bool resident_seq = false);
    // Bytes of the cached device block (0: none); releases it when idle.
    static uint64_t cached_block_bytes();
    static void release_cached_block();
    // Queues one pair; ez receives the result (reset now, filled when done).
    // copy: keep private copies of the sequences (the caller may free them).
    // defer_ms non-null: when the pair's footprint does not fit beside the
    // memory already held and queued, queue nothing, return KSW2_HIP_EAGAIN
    // and set *defer_ms to the suggested retry delay.
    int submit(const ksw2_pair_t &p, ksw_extz_t *ez, bool copy, uint64_t *ticket, int *defer_ms = nullptr);
    struct Load {
        uint64_t budget, held_bytes, queued_bytes;
        int64_t pending;
        double pending_cells, cells_per_s;
        uint64_t seq_arena_bytes;
        int64_t resident_tasks, staged_flights;
    };
    Load load();
    // Batch mode: hold admission while a batch is queued, then sort the
    // queue longest first and release it.
    void hold(bool on);
    void order_longest_first();
    // Completed tickets (up to max), their submit-to-done latencies and
    // statuses (KSW2_HIP_OK, or why the device path did not run the pair;
    // latency and status may be null); waits up to wait_ms (<0: until one).
    int poll(uint64_t *tickets, double *latency, int max, int wait_ms, int32_t *status = nullptr);
    // The first failure any task of this engine has met (KSW2_HIP_OK: none).
    int status();
    // Waits until every submitted task is done (batch mode); polls the
    // progress callback meanwhile.  Returns the first error seen, else OK.
    int drain(const Progress *prog);

private:
    int max_team_ = 64;         // largest team a flight is keyed on (kMaxWide; KSW2_HIP_MAX_TEAM, tuning)
    int backfill_ = 1024;       // tasks a flight may pass over in the queue (KSW2_HIP_BACKFILL, tuning)
    int serve_spread_ = 0;      // serving flights spread 5-16 wave teams: 0 never, 1 under light load, 2 always (KSW2_HIP_SERVE_SPREAD, tuning)
    bool batch_spread_ = false; // batch flights likewise (KSW2_HIP_BATCH_SPREAD, tuning)
    int nstreams_ = 4;          // streams in use (hardware queues; KSW2_HIP_STREAMS, tuning)
    int stream_class_[kMaxStreams] = {};  // serving: each stream's own size class (KSW2_HIP_STREAM_CLASSES, tuning)
    int fail_flights_ = 0;      // tests: the first this many flights fail before launch (KSW2_HIP_FAIL_FLIGHTS)
    // Tuning switches, off by default (uneven batch in one process, 2.07-2.28 s
    // with neither): a delivered pair's direction bytes go back to the pool at
    // once (KSW2_HIP_PAIR_RELEASE=1: 2.22-2.50 s); batch teams keep each
    // wave's direction bytes under this times budget / resident waves
    // (KSW2_HIP_WAVE_BYTES=1: 2.07 s); batch teams from the engine's per-wave
    // step budget instead of each flight's own (KSW2_HIP_ENGINE_T=1: 2.6 s
    // and slower)
    bool pair_release_ = false;
    double wave_bytes_scale_ = 0.0;
    bool engine_team_T_ = false;
    struct Task;
    struct Flight;
    void loop();
    bool reap_locked(std::unique_lock<std::mutex> &lk);
    bool admit_locked(std::unique_lock<std::mutex> &lk);
    bool queues_empty() const;
    void stop_all_locked();

    int8_t m_;
    ScoreClass cls_;
    GenScore gsc_{};
    uint64_t budget_;
    int device_ = 0;
    int resident_waves_ = 4096;
    std::unique_ptr<DevicePool> pool_;
    hipStream_t streams_[kMaxStreams] = {};
    bool busy_[kMaxStreams] = {};
    // per-stream host mirror (pinned, mapped, coherent) of flight results
    struct Mirror {
        KswOut *out = nullptr, *out_dev = nullptr;
        uint32_t *cig = nullptr, *cig_dev = nullptr;
        int32_t *flag = nullptr, *flag_dev = nullptr;
        uint64_t slots = 0, words = 0;
        void release()
        {
            if (out) (void)hipHostFree(out);
            if (cig) (void)hipHostFree(cig);
            if (flag) (void)hipHostFree(flag);
            *this = Mirror();
        }
    };
    Mirror mirror_[kMaxStreams];
    uint8_t *staging_[kMaxStreams] = {};           // pinned upload buffers, per stream
    uint64_t staging_bytes_[kMaxStreams] = {};
    int32_t *h_stop_ = nullptr, *d_stop_ = nullptr;  // stop flag shared by all flights
    // Device-resident sequences (serving): a device arena carved from the
    // pool and a pinned host ring of the same size, both used in submission
    // order (a task's range never straddles the end).  A submission copies
    // its pair into the host ring (no HIP call); the worker uploads what was
    // written since its last pass with one or two copies on its upload
    // stream and records an event; a flight whose tasks' uploads have all
    // completed reads the arena with no device-side wait (which could stall
    // behind other work sharing the upload stream's hardware queue),
    // otherwise it uploads its own sequences.  Ranges are freed at delivery;
    // the ring's head moves past the oldest freed ones.
    static constexpr int kUpEvents = 64;
    struct SeqArena {
        uint8_t *dev = nullptr;       // device arena (inside the pool)
        uint8_t *ring = nullptr;      // pinned host ring, same layout
        uint64_t pool_off = 0, bytes = 0;
        // positions in the unbounded submission stream (ring offset =
        // position % bytes): live ranges from head, written up to tail,
        // uploaded up to flushed, known complete up to done
        uint64_t head = 0, tail = 0, flushed = 0, done = 0;
        std::deque<std::pair<uint64_t, uint64_t>> live;  // [start, end) in order; end 0: freed
        hipStream_t up = nullptr;
        hipEvent_t ev[kUpEvents] = {};
        std::deque<std::pair<uint64_t, int>> pending;     // (end position, event) of uploads in flight
        int next_ev = 0;
    };
    SeqArena arena_;
    std::mutex seq_mu_;
    int64_t resident_tasks_ = 0;   // under seq_mu_
    int64_t staged_flights_ = 0;   // under mu_
    bool stage_seq_locked(Task *t);  // copies into the host ring (holds seq_mu_)
    uint64_t upload_seq();           // issues the pending uploads; returns the completed position
    void free_seq(Task *t);
    std::thread worker_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Task *> queue_[kClasses];
    std::vector<std::unique_ptr<Flight>> flights_;
    std::deque<uint64_t> done_;
    std::unordered_map<uint64_t, Task *> tasks_;
    uint64_t next_ticket_ = 0;
    int64_t pending_ = 0;
    // admission by waves: waves held by admitted, undelivered tasks, and the
    // stripe-kernel steps queued and in flight (for the per-wave step budget)
    int64_t busy_waves_ = 0;
    double queued_steps_ = 0, inflight_steps_ = 0;
    // memory and work accounting for deferral: footprints of admitted
    // flights and of queued tasks; cells of every pending task; delivered
    // throughput (moving average over completions under load)
    uint64_t held_bytes_ = 0, queued_bytes_ = 0;
    double pending_cells_ = 0, cells_per_s_ = 0, last_reap_t_ = 0, since_reap_cells_ = 0;
    void delivered_locked(const Task *t);
    int status_ = 0;
    bool quit_ = false;
    bool hold_ = false;
    bool batch_ = false;  // longest-first batch: throughput over latency
    // KSW2_HIP_ENGINE_STATS: per-phase host time, printed when the engine ends
    int stats_ = 0;  // 1: totals, 2: also one line per flight
    double st_create_ = 0, st_launch_ = 0, st_fetch_ = 0, st_device_ = 0;
    std::vector<float> st_wait_, st_serv_;  // per delivered task: queued, launched-to-delivered (s)
    int64_t st_flights_ = 0, st_tasks_ = 0;
};

}  // namespace kswhip
