// serve.h -- request loop with continuous batching over the batched hot path.
//
// A request is a CKKS ciphertext and a chain of k links; link i is
//   ct <- rotate(rescale(ct * ct relinearised), steps[i]).
// Evaluation keys stay resident for the server's lifetime; request state
// (the ciphertext) stays resident until its result is taken.
//
// Every tick (step()) admits pending requests in arrival order while the
// memory they reserve (their ciphertext plus the per-item scratch of their
// level) fits the budget and a slot is free, then issues the next operation
// of every active request: the multiplications, the rescales and the rotations
// of one context each go as one transient batch plan. Requests admitted now
// share the launches of requests already in flight. Work of consecutive ticks
// is queued on one stream, so the host runs ahead of the device by at most
// `depth` ticks; a request is complete when the event of the tick that ran its
// last operation has fired.
#pragma once

#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "batch.h"

namespace phx {

struct ServerStats {
    u64 submitted = 0, admitted = 0, completed = 0;
    u64 ticks = 0, ops = 0;
    double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
    double span_s = 0;             // first arrival to last completion
    double req_per_s = 0, ops_per_s = 0;
    u64 peak_reserved = 0, budget = 0;
    u64 peak_active = 0;
};

class Server {
public:
    struct Lane {   // one ring degree / parameter set
        const Context *ctx;
        const KSwitchKey *rlk;
        const GaloisKeys *gk;
    };
    Server(std::vector<Lane> lanes, size_t mem_budget, u32 max_active, u32 depth, hipStream_t s);
    ~Server();

    double now() const;   // seconds on the server clock
    // copies ct; arrival < 0 means now
    u64 submit(u32 lane, const Ciphertext &ct, const std::vector<int> &steps, double arrival);
    // one scheduling tick; returns the operations issued
    u32 step();
    void drain();
    // 0 pending, 1 active, 2 done (result waiting), 3 taken / unknown
    int status(u64 id) const;
    // moves the result out and frees the request's reservation
    void take(u64 id, Ciphertext &out, double *latency_s);
    ServerStats stats() const;

private:
    struct Request {
        u64 id = 0;
        u32 lane = 0;
        Ciphertext ct;
        std::vector<int> steps;
        u32 next_op = 0;   // 3 * link + {0 multiply, 1 rescale, 2 rotate}
        double arrival = 0, admitted = 0, done = 0;
        size_t state_bytes = 0, scratch_bytes = 0;   // reserved while active (state until taken)
        int state = 0;
    };
    // pinned host buffer the work lists of one plan are copied from
    struct Staging {
        uint8_t *p = nullptr;
        size_t bytes = 0;
    };
    struct Tick {
        hipEvent_t ev = nullptr;
        std::vector<std::unique_ptr<BatchPlan>> plans;
        std::vector<Staging> staging;   // returned to the free list when the tick is harvested
        // ciphertexts the tick's rescales replaced: freed once the tick has run, so
        // the pool cannot hand their memory to another stream (the caller's uploads)
        // while this tick's kernels may still read it
        std::vector<Ciphertext> retired;
        std::vector<u64> finishing;
    };
    void harvest(bool wait_oldest);
    // runs a transient plan of `t` in the server's work memory (grown when too small)
    void run_plan(Tick &t, BatchPlan &p);

    std::vector<Lane> lanes_;
    size_t budget_;
    u32 max_active_, depth_;
    hipStream_t s_;
    std::chrono::steady_clock::time_point t0_;
    u64 next_id_ = 1;
    std::map<u64, std::unique_ptr<Request>> reqs_;
    std::deque<u64> pending_;
    std::vector<u64> active_;
    std::deque<Tick> inflight_;
    size_t reserved_ = 0;
    // work memory of the plans: ticks run one after another on one stream, so
    // every plan uses the same arena and table buffer, the table copies ordered
    // on the stream from pinned staging buffers that are recycled per tick
    u64 *arena_ = nullptr;
    size_t arena_bytes_ = 0;
    uint8_t *tables_ = nullptr;
    size_t tables_bytes_ = 0;
    std::vector<Staging> staging_free_;
    ServerStats st_;
    std::vector<double> latencies_;
    double first_arrival_ = -1, last_done_ = 0;
};

}  // namespace phx
