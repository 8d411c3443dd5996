// serve.cpp -- continuous-batching request loop (see serve.h).
#include "serve.h"

#include <algorithm>

namespace phx {

Server::Server(std::vector<Lane> lanes, size_t mem_budget, u32 max_active, u32 depth, hipStream_t s)
    : lanes_(std::move(lanes)), budget_(mem_budget), max_active_(std::max(1u, max_active)),
      depth_(std::max(1u, depth)), s_(s), t0_(std::chrono::steady_clock::now()) {
    require(!lanes_.empty(), "server needs at least one parameter set");
    for (const auto &l : lanes_) {
        require(l.ctx && l.rlk && l.gk, "server lane needs a context, a relinearisation key and Galois keys");
        require(l.ctx->params().scheme == kSchemeCKKS, "the serving loop runs CKKS requests");
    }
    st_.budget = budget_;
}

Server::~Server() {
    try {
        drain();
    } catch (...) {
    }
    for (auto &kv : reqs_) kv.second->ct.release(s_);
    (void)hipStreamSynchronize(s_);
    for (auto &t : inflight_)
        for (auto &c : t.retired) c.release(s_);
    (void)hipFree(arena_);
    (void)hipFree(tables_);
    for (auto &t : inflight_)
        for (auto &b : t.staging) (void)hipHostFree(b.p);
    for (auto &b : staging_free_) (void)hipHostFree(b.p);
}

void Server::run_plan(Tick &t, BatchPlan &p) {
    // growth waits for the work queued on the stream, then replaces the buffer
    if (p.scratch_bytes() > arena_bytes_ || p.table_bytes() > tables_bytes_) {
        PHX_HIP_CHECK(hipStreamSynchronize(s_));
        if (p.scratch_bytes() > arena_bytes_) {
            PHX_HIP_CHECK(hipFree(arena_));
            arena_ = nullptr;
            arena_bytes_ = 0;
            const size_t want = p.scratch_bytes() + p.scratch_bytes() / 4;
            PHX_HIP_CHECK(hipMalloc(&arena_, want));
            arena_bytes_ = want;
        }
        if (p.table_bytes() > tables_bytes_) {
            PHX_HIP_CHECK(hipFree(tables_));
            tables_ = nullptr;
            tables_bytes_ = 0;
            const size_t want = p.table_bytes() * 2;
            PHX_HIP_CHECK(hipMalloc(&tables_, want));
            tables_bytes_ = want;
        }
    }
    // a staging buffer of the free list that is large enough, else a new one
    Staging st;
    for (size_t i = 0; i < staging_free_.size(); ++i)
        if (staging_free_[i].bytes >= p.table_bytes()) {
            st = staging_free_[i];
            staging_free_.erase(staging_free_.begin() + (long)i);
            break;
        }
    if (!st.p) {
        st.bytes = p.table_bytes() * 2;
        void *h = nullptr;
        PHX_HIP_CHECK(hipHostMalloc(&h, st.bytes, hipHostMallocDefault));
        st.p = static_cast<uint8_t *>(h);
    }
    t.staging.push_back(st);
    p.run_with(s_, arena_, tables_, st.p);
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

u64 Server::submit(u32 lane, const Ciphertext &ct, const std::vector<int> &steps, double arrival) {
    require(lane < lanes_.size(), "unknown server lane");
    const Context &ctx = *lanes_[lane].ctx;
    require(ct.data && ct.size == 2 && ct.is_ntt_form, "request ciphertext must be size 2 in NTT form");
    require(ct.poly_degree == ctx.N(), "request ciphertext does not match the lane's ring degree");
    require(!steps.empty(), "request chain is empty");
    require(ct.chain_index + steps.size() <= ctx.size_Q(), "request chain runs past the end of the modulus chain");
    for (int st : steps) {
        const u32 elt = galois_elt_from_step(st, ctx.N());
        const auto &e = lanes_[lane].gk->elts;
        require(std::find(e.begin(), e.end(), elt) != e.end(), "Galois elt not present");
    }
    auto r = std::make_unique<Request>();
    r->id = next_id_++;
    r->lane = lane;
    r->steps = steps;
    r->arrival = arrival >= 0 ? arrival : now();
    // the copy runs on the calling thread's stream, where the caller's buffer
    // is produced and freed; the server stream waits for it
    hipStream_t cs = hipStreamPerThread;
    r->ct.resize(ctx, ct.chain_index, 2, cs);
    r->ct.scale = ct.scale;
    r->ct.is_ntt_form = true;
    PHX_HIP_CHECK(hipMemcpyAsync(r->ct.data, ct.data, (size_t)2 * ct.coeff_modulus_size * ctx.N() * 8,
                                 hipMemcpyDeviceToDevice, cs));
    hipEvent_t ev;
    PHX_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    PHX_HIP_CHECK(hipEventRecord(ev, cs));
    PHX_HIP_CHECK(hipStreamWaitEvent(s_, ev, 0));
    PHX_HIP_CHECK(hipEventDestroy(ev));
    // the ciphertext at its start level (it only shrinks) plus the batch scratch
    // of one item there (rotation scratch is the larger of the two kinds)
    r->state_bytes = (size_t)2 * ctx.limbs_at(ct.chain_index) * ctx.N() * 8;
    r->scratch_bytes = BatchPlan::item_scratch_bytes(ctx, ct.chain_index, BatchPlan::kRotate);
    require(r->state_bytes + r->scratch_bytes <= budget_, "request needs more memory than the server budget");
    if (first_arrival_ < 0 || r->arrival < first_arrival_) first_arrival_ = r->arrival;
    const u64 id = r->id;
    pending_.push_back(id);
    reqs_[id] = std::move(r);
    ++st_.submitted;
    return id;
}

void Server::harvest(bool wait_oldest) {
    while (!inflight_.empty()) {
        Tick &t = inflight_.front();
        if (wait_oldest) {
            PHX_HIP_CHECK(hipEventSynchronize(t.ev));
            wait_oldest = false;
        } else {
            const hipError_t q = hipEventQuery(t.ev);
            if (q == hipErrorNotReady) break;
            PHX_HIP_CHECK(q);
        }
        const double tn = now();
        for (u64 id : t.finishing) {
            Request &r = *reqs_.at(id);
            r.state = 2;
            r.done = tn;
            latencies_.push_back(tn - r.arrival);
            last_done_ = std::max(last_done_, tn);
            ++st_.completed;
        }
        (void)hipEventDestroy(t.ev);
        for (auto &b : t.staging) staging_free_.push_back(b);
        for (auto &c : t.retired) c.release(s_);
        inflight_.pop_front();
    }
}

u32 Server::step() {
    harvest(inflight_.size() >= depth_);
    // admission in arrival order
    while (!pending_.empty() && active_.size() < max_active_) {
        Request &r = *reqs_.at(pending_.front());
        if (reserved_ + r.state_bytes + r.scratch_bytes > budget_) break;
        reserved_ += r.state_bytes + r.scratch_bytes;
        r.state = 1;
        r.admitted = now();
        active_.push_back(r.id);
        pending_.pop_front();
        ++st_.admitted;
    }
    st_.peak_reserved = std::max<u64>(st_.peak_reserved, reserved_);
    st_.peak_active = std::max<u64>(st_.peak_active, active_.size());
    if (active_.empty()) return 0;

    Tick tick;
    u32 issued = 0;
    for (u32 lane = 0; lane < lanes_.size(); ++lane) {
        const Lane &ln = lanes_[lane];
        std::vector<Request *> mul, res, rot;
        for (u64 id : active_) {
            Request &r = *reqs_.at(id);
            if (r.lane != lane) continue;
            const u32 op = r.next_op % 3;
            (op == 0 ? mul : op == 1 ? res : rot).push_back(&r);
        }
        if (!mul.empty()) {
            std::vector<Ciphertext *> a;
            std::vector<const Ciphertext *> b;
            for (Request *r : mul) {
                a.push_back(&r->ct);
                b.push_back(&r->ct);   // squaring
            }
            tick.plans.emplace_back(BatchPlan::mul_relin(*ln.ctx, a, b, *ln.rlk, true));
            run_plan(tick, *tick.plans.back());
        }
        if (!res.empty()) {
            // every rescale of the lane in one batched launch set: outputs at the next
            // level from the pool; the inputs are freed when the tick is harvested
            // (its event has completed), not in stream order on s_: the caller's
            // thread allocates request copies from the same pool on its own stream
            std::vector<Ciphertext> outs(res.size());
            std::vector<const Ciphertext *> ins;
            std::vector<Ciphertext *> outp;
            try {
                for (size_t i = 0; i < res.size(); ++i) {
                    outs[i].resize(*ln.ctx, res[i]->ct.chain_index + 1, 2, s_);
                    ins.push_back(&res[i]->ct);
                    outp.push_back(&outs[i]);
                }
                tick.plans.emplace_back(BatchPlan::rescale(*ln.ctx, ins, outp, true));
                run_plan(tick, *tick.plans.back());
            } catch (...) {
                for (auto &o : outs) o.release(s_);
                throw;
            }
            for (size_t i = 0; i < res.size(); ++i) {
                tick.retired.push_back(res[i]->ct);
                res[i]->ct = outs[i];
            }
        }
        if (!rot.empty()) {
            std::vector<Ciphertext *> a;
            std::vector<int> steps;
            for (Request *r : rot) {
                a.push_back(&r->ct);
                steps.push_back(r->steps[r->next_op / 3]);
            }
            tick.plans.emplace_back(BatchPlan::rotate(*ln.ctx, a, steps, *ln.gk, true));
            run_plan(tick, *tick.plans.back());
        }
        for (auto *v : {&mul, &res, &rot})
            for (Request *r : *v) {
                ++r->next_op;
                ++issued;
            }
    }
    // requests whose chain is now fully issued complete with this tick
    std::vector<u64> still;
    for (u64 id : active_) {
        Request &r = *reqs_.at(id);
        if (r.next_op == 3 * r.steps.size()) {
            tick.finishing.push_back(id);
            reserved_ -= r.scratch_bytes;   // the state stays reserved until taken
            r.scratch_bytes = 0;
        } else {
            still.push_back(id);
        }
    }
    active_.swap(still);
    PHX_HIP_CHECK(hipEventCreateWithFlags(&tick.ev, hipEventDisableTiming));
    PHX_HIP_CHECK(hipEventRecord(tick.ev, s_));
    inflight_.push_back(std::move(tick));
    ++st_.ticks;
    st_.ops += issued;
    return issued;
}

void Server::drain() {
    while (!pending_.empty() || !active_.empty()) {
        if (step() == 0 && !pending_.empty() && active_.empty())
            fail(kLogicError, "a pending request cannot be admitted");
    }
    while (!inflight_.empty()) harvest(true);
}

int Server::status(u64 id) const {
    auto it = reqs_.find(id);
    return it == reqs_.end() ? 3 : it->second->state;
}

void Server::take(u64 id, Ciphertext &out, double *latency_s) {
    auto it = reqs_.find(id);
    require(it != reqs_.end(), "unknown request id");
    Request &r = *it->second;
    require(r.state == 2, "request is not complete");
    if (latency_s) *latency_s = r.done - r.arrival;
    out.release(s_);
    out = r.ct;
    r.ct.data = nullptr;
    reserved_ -= r.state_bytes;
    reqs_.erase(it);
}

ServerStats Server::stats() const {
    ServerStats s = st_;
    if (!latencies_.empty()) {
        std::vector<double> l = latencies_;
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) {
            const size_t i = std::min(l.size() - 1, (size_t)(p * (double)(l.size() - 1) + 0.5));
            return l[i] * 1e3;
        };
        s.p50_ms = pct(0.50);
        s.p99_ms = pct(0.99);
        s.max_ms = l.back() * 1e3;
        double sum = 0;
        for (double v : l) sum += v;
        s.mean_ms = sum / (double)l.size() * 1e3;
        s.span_s = last_done_ - first_arrival_;
        if (s.span_s > 0) {
            s.req_per_s = (double)s.completed / s.span_s;
            s.ops_per_s = (double)s.ops / s.span_s;
        }
    }
    return s;
}

}  // namespace phx
