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
//
// BFV plans (multiply + relinearise at the top level) stage both operands of
// every item contiguously, run each multiply stage once over all items, key-
// switch them as above and write the results back. A joined plan runs plans of
// different contexts (ring degrees) one after another in one graph, sharing one
// scratch arena.
#pragma once

#include <memory>
#include <vector>

#include "evaluator.h"

namespace phx {

class BatchPlan {
public:
    enum Kind { kMulRelin = 1, kRotate = 2, kRelin = 3, kBfvMulRelin = 4, kJoin = 5 };

    // ct1[i] *= ct2[i], relinearised with rlk. CKKS: all in NTT form, any level.
    // BFV (persistent plans only): coefficient form at the top level, multiplied
    // with the context's technique (BEHZ or HPS).
    static BatchPlan *mul_relin(const Context &ctx, const std::vector<Ciphertext *> &ct1,
                                const std::vector<const Ciphertext *> &ct2, const KSwitchKey &rlk,
                                bool transient = false);
    // one plan running `parts` (persistent, any contexts) in order; takes ownership
    static BatchPlan *join(const std::vector<BatchPlan *> &parts);
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
    static constexpr u32 kTables = 6;   // a, b, out, keys, perms, key-switch outputs (BFV)
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
        u64 **ks_out = nullptr;        // BFV: the key switch adds into these (NTT form)
        // scratch (views into the shared arena)
        u64 *c2 = nullptr, *t_cks = nullptr, *t_mod_up = nullptr, *cx = nullptr, *delta = nullptr,
            *base0 = nullptr;
        u64 c2_stride = 0;             // words between items' c2
        // BFV: staged operands [B][2][Ql][N], multiply work, product [B][3][Ql][N]
        u64 *in1 = nullptr, *in2 = nullptr, *work = nullptr, *c3 = nullptr;
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
    // persistent plans: bind to `arena`, point arena-relative table entries at it and
    // upload the tables
    void rebind(u64 *arena);
    void run_group(const Group &g, hipStream_t s);
    void keyswitch_group(const Group &g, hipStream_t s);
    void record(hipStream_t s);                  // every launch of the plan
    void after_launch();                         // host-side metadata of the outputs

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
    std::vector<std::unique_ptr<BatchPlan>> parts_;   // joined plans
};

}  // namespace phx
