// batch.h -- batched hot path: many ciphertexts per launch, hipGraph replay.
//
// A BatchPlan binds a fixed list of ciphertexts (and keys) to scratch and
// device pointer tables. Ciphertexts are grouped by level; each pipeline
// stage of each group is one launch over the whole group, so the small
// per-ciphertext launches of the plain path become a few large ones, and the
// key inner product reads a shared key once per coefficient for the whole
// group. The groups run one after another on one stream, so they share one
// scratch arena sized for the largest group.
//
// Persistent plans allocate the arena and tables once at creation: run()
// launches eagerly; capture() records run() into a hipGraph that replay()
// launches with one call (nothing is allocated, set or synchronised inside
// run()). Transient plans (the serving loop) take the arena from the
// stream-ordered pool inside run() and return it at the end of the same
// stream's work; the plan object must outlive that work.
#pragma once

#include <vector>

#include "evaluator.h"

namespace phx {

class BatchPlan {
public:
    enum Kind { kMulRelin = 1, kRotate = 2, kRelin = 3 };

    // ct1[i] *= ct2[i], relinearised with rlk (CKKS, all in NTT form)
    static BatchPlan *mul_relin(const Context &ctx, const std::vector<Ciphertext *> &ct1,
                                const std::vector<const Ciphertext *> &ct2, const KSwitchKey &rlk,
                                bool transient = false);
    // cts[i] <- rotate(cts[i], steps[i]); every step's Galois element must have a key in gk
    static BatchPlan *rotate(const Context &ctx, const std::vector<Ciphertext *> &cts, const std::vector<int> &steps,
                             const GaloisKeys &gk, bool transient = false);
    // scratch bytes one item at chain index `chain` needs (admission budgets)
    static size_t item_scratch_bytes(const Context &ctx, u32 chain, Kind k);
    ~BatchPlan();

    void run(hipStream_t s);
    void capture(hipStream_t s);
    void replay(hipStream_t s);
    size_t size() const { return items_; }
    size_t scratch_bytes() const { return scratch_bytes_; }

private:
    static constexpr u32 kTables = 5;   // a, b, out, keys, perms
    struct Group {
        u32 chain = 0, Ql = 0, QlP = 0, beta = 0, B = 0;
        std::vector<u32> items;
        size_t table0 = 0;             // first entry of this group's tables
        // device pointer tables [B]
        const u64 **a = nullptr;       // operand 1 / ciphertext data
        const u64 **b = nullptr;       // operand 2 (mul-relin)
        u64 **out = nullptr;           // output ciphertext data
        const u64 **keys = nullptr;    // key data per item
        const u32 **perms = nullptr;   // Galois permutation per item
        // scratch (views into the shared arena)
        u64 *c2 = nullptr, *t_cks = nullptr, *t_mod_up = nullptr, *cx = nullptr, *delta = nullptr,
            *base0 = nullptr;
    };
    BatchPlan(const Context &ctx, Kind k, bool transient) : ctx_(ctx), kind_(k), transient_(transient) {}
    void build_groups(const std::vector<u32> &chains);
    // host tables for group g: item i's a, b, out, key, perm
    void set_tables(Group &g, const std::vector<const void *> &a, const std::vector<const void *> &b,
                    const std::vector<const void *> &out, const std::vector<const void *> &keys,
                    const std::vector<const void *> &perms);
    size_t group_scratch_words(const Group &g) const;
    void layout();                               // table offsets, arena size
    void bind(u64 *arena, const void **tables);  // group views into device memory
    void run_group(const Group &g, hipStream_t s);

    const Context &ctx_;
    Kind kind_;
    bool transient_ = false;
    size_t items_ = 0;
    size_t scratch_bytes_ = 0;       // arena bytes
    std::vector<Group> groups_;
    std::vector<const void *> host_tables_;   // [group tables][kTables][B]
    u64 *arena_ = nullptr;                    // persistent plans
    const void **tables_ = nullptr;
    std::vector<Ciphertext *> cts_;
    std::vector<double> scale_mul_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

}  // namespace phx
