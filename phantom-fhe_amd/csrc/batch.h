// batch.h -- batched hot path: many ciphertexts per launch, hipGraph replay.
//
// A BatchPlan binds a fixed list of ciphertexts (and keys) to preallocated
// scratch and device pointer tables. Ciphertexts are grouped by level; each
// pipeline stage of each group is one launch over the whole group, so the
// small per-ciphertext launches of the plain path become a few large ones,
// and the key inner product reads a shared key once per coefficient for the
// whole group. run() launches eagerly; capture() records run() into a
// hipGraph that replay() launches with one call (nothing is allocated, set or
// synchronised inside run()).
#pragma once

#include <vector>

#include "evaluator.h"

namespace phx {

class BatchPlan {
public:
    enum Kind { kMulRelin = 1, kRotate = 2, kRelin = 3 };

    // ct1[i] *= ct2[i], relinearised with rlk (CKKS, all in NTT form)
    static BatchPlan *mul_relin(const Context &ctx, const std::vector<Ciphertext *> &ct1,
                                const std::vector<const Ciphertext *> &ct2, const KSwitchKey &rlk);
    // cts[i] <- rotate(cts[i], steps[i]); every step's Galois element must have a key in gk
    static BatchPlan *rotate(const Context &ctx, const std::vector<Ciphertext *> &cts, const std::vector<int> &steps,
                             const GaloisKeys &gk);
    ~BatchPlan();

    void run(hipStream_t s);
    void capture(hipStream_t s);
    void replay(hipStream_t s);
    size_t size() const { return items_; }
    size_t scratch_bytes() const { return scratch_bytes_; }

private:
    struct Group {
        u32 chain = 0, Ql = 0, QlP = 0, beta = 0, B = 0;
        std::vector<u32> items;
        // device pointer tables [B]
        const u64 **a = nullptr;       // operand 1 / ciphertext data
        const u64 **b = nullptr;       // operand 2 (mul-relin)
        u64 **out = nullptr;           // output ciphertext data
        const u64 **keys = nullptr;    // key data per item
        const u32 **perms = nullptr;   // Galois permutation per item
        // scratch
        u64 *c2 = nullptr, *t_cks = nullptr, *t_mod_up = nullptr, *cx = nullptr, *delta = nullptr,
            *base0 = nullptr;
    };
    BatchPlan(const Context &ctx, Kind k) : ctx_(ctx), kind_(k) {}
    void build_groups(const std::vector<u32> &chains);
    void alloc_group(Group &g);
    void run_group(const Group &g, hipStream_t s);

    const Context &ctx_;
    Kind kind_;
    size_t items_ = 0;
    size_t scratch_bytes_ = 0;
    std::vector<Group> groups_;
    std::vector<void *> allocations_;
    std::vector<Ciphertext *> cts_;
    std::vector<double> scale_mul_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

}  // namespace phx
