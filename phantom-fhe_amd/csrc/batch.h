// batch.h -- batched hot path: many ciphertexts per launch, hipGraph replay.
//
// A BatchPlan binds a fixed list of ciphertexts (and keys) to scratch and to
// device work lists. Every pipeline stage is ONE launch over all items of the
// plan, whatever their levels: the NTT and FastBConv launches take a table of
// (polynomial, limb) entries with their own buffers, table rows and level
// constants; the tensor / permutation and key inner product take an item table
// and skip the limbs an item's level does not have. So a plan over ciphertexts
// of thirty different levels issues the same dozen launches as a plan over
// one level, and the key inner product reads a key shared by items of any level
// once per coefficient.
//
// Persistent plans allocate the arena (one scratch region per item) and the
// device tables at creation: run() launches eagerly; capture() records run()
// into a hipGraph that replay() launches with one call (nothing is allocated,
// set or synchronised inside run()). Transient plans (the serving loop) run in
// memory the caller hands over (run_with) or take it from the stream-ordered
// pool inside run(); they upload their tables on the same stream from pinned
// host memory, and the plan object must outlive that work.
//
// BFV plans (multiply + relinearise at the top level) stage both operands of
// every item contiguously, run each multiply stage once over all items, key-
// switch them as above and write the results back.
//
// Concurrent branches: a joined plan runs its parts (plans of any contexts /
// ring degrees) on up to `batch_branches` branches: branch 0 on the caller's
// stream, every other branch on a plan-owned stream forked from it by an event
// and joined back into it by another, so one graph holds parallel chains of
// launches whose ramp-up and tail overlap. Parts of one branch run one after
// another in one arena of the largest part's size. A persistent plan of at least
// 2 x batch_branches items is built as such a joined plan of contiguous runs of
// items of about equal work, each with its own work lists.
#pragma once

#include <memory>
#include <vector>

#include "evaluator.h"
#include "ntt.h"

namespace phx {

class BatchPlan {
public:
    enum Kind { kMulRelin = 1, kRotate = 2, kRelin = 3, kBfvMulRelin = 4, kJoin = 5, kRescale = 6 };

    // ct1[i] *= ct2[i], relinearised with rlk. CKKS: all in NTT form, any level.
    // BFV (persistent plans only): coefficient form at the top level, multiplied
    // with the context's technique (BEHZ or HPS).
    static BatchPlan *mul_relin(const Context &ctx, const std::vector<Ciphertext *> &ct1,
                                const std::vector<const Ciphertext *> &ct2, const KSwitchKey &rlk,
                                bool transient = false);
    // one plan running `parts` (persistent, any contexts; joined plans are
    // flattened into their parts) on concurrent branches; takes ownership
    static BatchPlan *join(const std::vector<BatchPlan *> &parts);
    // cts[i] <- rotate(cts[i], steps[i]); every step's Galois element must have a key in gk
    static BatchPlan *rotate(const Context &ctx, const std::vector<Ciphertext *> &cts, const std::vector<int> &steps,
                             const GaloisKeys &gk, bool transient = false);
    // outs[i] <- rescale_to_next(ins[i]) (CKKS, NTT form, any levels); outs[i] must
    // already be sized for chain index ins[i].chain_index + 1 and differ from ins[i]
    static BatchPlan *rescale(const Context &ctx, const std::vector<const Ciphertext *> &ins,
                              const std::vector<Ciphertext *> &outs, bool transient = false);
    // scratch bytes one item at chain index `chain` needs (admission budgets)
    static size_t item_scratch_bytes(const Context &ctx, u32 chain, Kind k);
    ~BatchPlan();

    void run(hipStream_t s);
    // transient plans: run in caller memory of at least scratch_bytes() and
    // table_bytes() (device), free of other work on s from here on; the work
    // lists go through `staging`, pinned host memory of table_bytes() that must
    // stay untouched until the stream has passed this call's work (the copy
    // from it is ordered on s after the work queued before it)
    void run_with(hipStream_t s, u64 *arena, uint8_t *tables, uint8_t *staging);
    size_t table_bytes() const;
    void capture(hipStream_t s);
    void replay(hipStream_t s);
    size_t size() const { return items_.size() + joined_items_; }
    size_t scratch_bytes() const { return scratch_bytes_; }
    // kernel launches one run issues (the graph's kernel nodes)
    size_t launches() const;
    // concurrent branches (1 unless joined)
    size_t branches() const { return kind_ == kJoin ? nbranch_ : 1; }

private:
    struct Item {
        u32 chain = 0, Ql = 0, QlP = 0, beta = 0;
        const u64 *a = nullptr;      // operand 1 / the ciphertext (device)
        const u64 *b = nullptr;      // operand 2 (mul-relin)
        u64 *out = nullptr;          // output ciphertext data
        const u64 *key = nullptr;    // key-switching key
        const u32 *perm = nullptr;   // Galois permutation (rotate)
        size_t off = 0;              // first word of the item's scratch in the arena
    };
    // BFV staging (one group at the top level): operands, multiply work, product
    struct BfvStage {
        u64 *in1 = nullptr, *in2 = nullptr, *work = nullptr, *c3 = nullptr;
    };

    BatchPlan(const Context &ctx, Kind k, bool transient) : ctx_(ctx), kind_(k), transient_(transient) {}
    // a plan with its items set: laid out, or split into concurrent parts
    static BatchPlan *finish(std::unique_ptr<BatchPlan> p);
    // the joined plan of `parts` (persistent, unjoined) on nb branches; it owns
    // them when it returns, and nothing has changed when it throws
    static BatchPlan *join_parts(const std::vector<BatchPlan *> &parts, u32 nb);
    size_t item_words(const Item &it) const;
    void layout();                           // scratch offsets, arena size; persistent: allocate and bind
    void bind(u64 *arena);                   // build the host work lists against an arena
    void upload_tables(hipStream_t s, uint8_t *staging);
    void record(hipStream_t s);              // every launch of the plan
    void record_keyswitch(hipStream_t s);
    void after_launch();                     // host-side metadata of the outputs

    const Context &ctx_;
    Kind kind_;
    bool transient_ = false;
    std::vector<Item> items_;
    size_t joined_items_ = 0;
    size_t scratch_bytes_ = 0;       // arena bytes
    size_t bfv_words_ = 0;           // BFV staging words at the start of the arena
    BfvStage bfv_;
    u64 *arena_ = nullptr;           // persistent plans

    // host work lists and their device copy (one buffer, sections at offsets)
    struct PreItem {
        const u64 *a, *b;
        u64 *out;
        u64 *c2, *base0;
        const u32 *perm;
        u32 Ql, pad;
    };
    struct IpItem {
        const u64 *t, *c2;
        u64 *cx;
        const u64 *key;
        u32 Ql, QlP, beta, pad;
    };
    std::vector<PreItem> pre_;
    std::vector<IpItem> ip_;
    std::vector<NttJob> mu_inv_, mu_fwd_, md_inv_, md_fwd_;
    std::vector<BconvRow> mu_bc_, md_bc_;
    std::vector<BconvDesc> mu_desc_host_, md_desc_host_;   // every descriptor the rows use
    std::vector<BconvDesc> mu_fdesc_host_, md_fdesc_host_; // and their fused forms
    u32 max_Ql_ = 0, max_beta_ = 0;
    std::vector<uint8_t> host_tables_;
    uint8_t *tables_ = nullptr;      // device copy (persistent plans)
    uint8_t *tables_bound_ = nullptr;
    size_t off_pre_ = 0, off_ip_ = 0, off_mu_inv_ = 0, off_mu_fwd_ = 0, off_md_inv_ = 0, off_md_fwd_ = 0,
           off_mu_bc_ = 0, off_md_bc_ = 0, off_bfv_sq_ = 0;
    std::vector<u64> bfv_square_;    // BFV: per item 1 when both operands are one ciphertext

    std::vector<Ciphertext *> cts_;
    std::vector<const Ciphertext *> ins_;   // rescale inputs
    std::vector<double> scale_mul_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    std::vector<std::unique_ptr<BatchPlan>> parts_;   // joined plans
    // joined plans: part i runs on branch branch_[i]; branch 0 on the caller's
    // stream, branch b > 0 on aux_[b - 1]
    std::vector<u32> branch_;
    u32 nbranch_ = 1;
    std::vector<u64 *> branch_arenas_;   // shared arenas of branches with several parts
    std::vector<hipStream_t> aux_;
    hipEvent_t fork_ = nullptr;
    std::vector<hipEvent_t> joins_;
};

}  // namespace phx
