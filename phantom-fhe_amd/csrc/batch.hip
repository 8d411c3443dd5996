// batch.hip -- batched multiply-and-relinearise and rotation (see batch.h).
#include <algorithm>
#include <map>
#include <memory>

#include "batch.h"
#include "ntt_arith.h"
#include "bfv.h"
#include "tune.h"

namespace phx {

namespace {
constexpr int kThreads = 256;
inline dim3 grid3(u64 N, u32 y, u32 z) { return dim3((u32)((N + kThreads - 1) / kThreads), y, z); }
}  // namespace

// tensor of item z: (d0, d1) into out[z], d2 into c2[z]
// (tensor_prod_2x2_rns_poly, src/polymath.cu:463-498)
__global__ __launch_bounds__(kThreads) void k_tensor_batch(const u64 *const *a, const u64 *const *b, u64 *const *out,
                                                           u64 *__restrict__ c2, const DevModulus *__restrict__ mod,
                                                           u32 N, u32 Ql) {
    const u32 limb = blockIdx.y, item = blockIdx.z;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 LN = (u64)Ql * N, o = (u64)limb * N + k;
    const u64 *pa = a[item], *pb = b[item];
    const DevModulus m = mod[limb];
    const u64 q = m.value;
    const u64 a0 = pa[o], a1 = pa[LN + o], b0 = pb[o], b1 = pb[LN + o];
    const u64 r0 = mul_mod(a0, b0, m);
    const u64 r2 = mul_mod(a1, b1, m);
    u64 r1 = mul_mod(add_mod(a0, a1, q), add_mod(b0, b1, q), m);
    r1 = sub_mod(sub_mod(r1, r0, q), r2, q);
    u64 *po = out[item];
    po[o] = r0;
    po[LN + o] = r1;
    c2[(u64)item * LN + o] = r2;
}

// Galois permutation of item z (apply_galois_ntt_permutation, src/galois.cu:11-18):
// base0[z] = pi(c0), c2[z] = pi(c1)
__global__ __launch_bounds__(kThreads) void k_permute_batch(const u64 *const *ct, const u32 *const *perm,
                                                            u64 *__restrict__ base0, u64 *__restrict__ c2, u32 N,
                                                            u32 Ql) {
    const u32 limb = blockIdx.y, item = blockIdx.z;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 LN = (u64)Ql * N, o = (u64)limb * N;
    const u64 *src = ct[item];
    const u32 pk = perm[item][k];
    base0[(u64)item * LN + o + k] = src[o + pk];
    c2[(u64)item * LN + o + k] = src[LN + o + pk];
}

// key inner product over a group (key_switch_inner_prod_c2_and_evk,
// src/eval_key_switch.cu:14-69): items sharing a key read it once per coefficient;
// digit j's own limbs (pos / alpha == j) are read from c2 (NTT form) directly.
// MB = the digit capacity (registers caching a shared key's digits)
template <u32 MB>
__global__ __launch_bounds__(kThreads) void k_inner_product_batch(const u64 *__restrict__ t, const u64 *__restrict__ c2,
                                                                  const u64 *const *keys, u64 *__restrict__ cx,
                                                                  const DevModulus *__restrict__ mod, u32 N, u32 Ql,
                                                                  u32 QlP, u32 size_Q, u32 QP, u32 beta, u32 alpha,
                                                                  u32 B, u64 c2_stride) {
    const u32 pos = blockIdx.y;
    const u32 own = pos < Ql ? pos / alpha : ~0u;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u32 row = pos < Ql ? pos : size_Q + (pos - Ql);
    const u64 QPN = (u64)QP * N;
    const DevModulus m = mod[row];
    constexpr u32 kMaxBeta = MB;
    if (m.pad) {
        // q < 2^50: exact products on the FP64 pipe (FpArith::mulmod, |r| < q), the
        // partial sums reduced every 4 digits (|sum| < 4.5 q < 2^53), then canonical
#pragma clang fp contract(off)
        const FpArith fa{nullptr, (double)m.value, __longlong_as_double((long long)m.pad)};
        double kb[kMaxBeta], ka[kMaxBeta];
        const u64 *cur = nullptr;
        for (u32 item = 0; item < B; ++item) {
            const u64 *key = keys[item];
            if (key != cur) {
                cur = key;
#pragma unroll
                for (u32 j = 0; j < kMaxBeta; ++j)
                    if (j < beta) {
                        const u64 *p = key + (u64)j * 2 * QPN + (u64)row * N + k;
                        kb[j] = FpArith::from_canon(p[0]);
                        ka[j] = FpArith::from_canon(p[QPN]);
                    }
            }
            const u64 *ti = t + (u64)item * beta * QlP * N + (u64)pos * N + k;
            const u64 *ci = c2 + (u64)item * c2_stride + (u64)pos * N + k;
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (u32 j = 0; j < kMaxBeta; ++j)
                if (j < beta) {
                    if (j && (j & 3) == 0) {
                        s0 = fa.red(s0);
                        s1 = fa.red(s1);
                    }
                    const double x = FpArith::from_canon(j == own ? *ci : ti[(u64)j * QlP * N]);
                    s0 += fa.mulmod(x, kb[j]);
                    s1 += fa.mulmod(x, ka[j]);
                }
            u64 *co = cx + (u64)item * 2 * QlP * N + (u64)pos * N + k;
            co[0] = fa.canon(s0);
            co[(u64)QlP * N] = fa.canon(s1);
        }
        return;
    }
    u64 kb[kMaxBeta], ka[kMaxBeta];
    const u64 *cur = nullptr;
    for (u32 item = 0; item < B; ++item) {
        const u64 *key = keys[item];
        if (key != cur) {
            cur = key;
#pragma unroll
            for (u32 j = 0; j < kMaxBeta; ++j)
                if (j < beta) {
                    const u64 *p = key + (u64)j * 2 * QPN + (u64)row * N + k;
                    kb[j] = p[0];
                    ka[j] = p[QPN];
                }
        }
        const u64 *ti = t + (u64)item * beta * QlP * N + (u64)pos * N + k;
        const u64 *ci = c2 + (u64)item * c2_stride + (u64)pos * N + k;
        u128 a0 = 0, a1 = 0;
#pragma unroll
        for (u32 j = 0; j < kMaxBeta; ++j)
            if (j < beta) {
                const u64 x = j == own ? *ci : ti[(u64)j * QlP * N];
                a0 += (u128)x * kb[j];
                a1 += (u128)x * ka[j];
            }
        u64 *co = cx + (u64)item * 2 * QlP * N + (u64)pos * N + k;
        co[0] = barrett128(a0, m.value, m.ratio_lo, m.ratio_hi);
        co[(u64)QlP * N] = barrett128(a1, m.value, m.ratio_lo, m.ratio_hi);
    }
}

// dst[z][0 .. words) = src_z[0 .. words) for item z = blockIdx.y
__global__ __launch_bounds__(kThreads) void k_gather_items(const u64 *const *src, u64 *__restrict__ dst, u64 words,
                                                           u64 dst_stride) {
    const u64 *a = src[blockIdx.y];
    u64 *d = dst + blockIdx.y * dst_stride;
    for (u64 i = (u64)blockIdx.x * kThreads + threadIdx.x; i < words; i += (u64)gridDim.x * kThreads) d[i] = a[i];
}

// dst_z[0 .. words) = src[z][0 .. words)
__global__ __launch_bounds__(kThreads) void k_scatter_items(u64 *const *dst, const u64 *__restrict__ src, u64 words,
                                                            u64 src_stride) {
    u64 *d = dst[blockIdx.y];
    const u64 *a = src + blockIdx.y * src_stride;
    for (u64 i = (u64)blockIdx.x * kThreads + threadIdx.x; i < words; i += (u64)gridDim.x * kThreads) d[i] = a[i];
}

namespace {
dim3 items_grid(u64 words, u32 items) {
    return dim3((u32)std::min<u64>((words + kThreads - 1) / kThreads, 1024), items);
}
}  // namespace

// ---------------------------------------------------------------------------
BatchPlan::~BatchPlan() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (arena_ || tables_) {
        (void)hipDeviceSynchronize();
        (void)hipFree(arena_);
        (void)hipFree(const_cast<void **>(tables_));
    }
}

void BatchPlan::build_groups(const std::vector<u32> &chains) {
    std::map<u32, std::vector<u32>> by;
    for (u32 i = 0; i < chains.size(); ++i) by[chains[i]].push_back(i);
    for (auto &kv : by) {
        Group g;
        g.chain = kv.first;
        const LevelConsts &L = ctx_.level(kv.first);
        g.Ql = L.size_Ql;
        g.QlP = L.size_QlP;
        g.beta = L.beta;
        g.items = kv.second;
        g.B = (u32)g.items.size();
        groups_.push_back(g);
    }
}

namespace {
// scratch words of one item: c2, t_cks, t_mod_up, cx, delta (+ base0 for rotations)
size_t item_words(u64 N, u32 Ql, u32 QlP, u32 beta, bool rotate) {
    return (u64)N * ((u64)Ql * (2 + (rotate ? 1 : 0)) + (u64)beta * QlP + 2ull * QlP + 2ull * Ql);
}
}  // namespace

size_t BatchPlan::item_scratch_bytes(const Context &ctx, u32 chain, Kind k) {
    const LevelConsts &L = ctx.level(chain);
    return 8 * item_words(ctx.N(), L.size_Ql, L.size_QlP, L.beta, k == kRotate);
}

size_t BatchPlan::group_scratch_words(const Group &g) const {
    size_t w = (size_t)g.B * item_words(ctx_.N(), g.Ql, g.QlP, g.beta, kind_ == kRotate);
    if (kind_ == kBfvMulRelin)   // staged operands, product, multiply work
        w += (size_t)g.B * (7ull * g.Ql * ctx_.N() + bfv_batch_work_words(ctx_));
    return w;
}

void BatchPlan::set_tables(Group &g, const std::vector<const void *> &a, const std::vector<const void *> &b,
                           const std::vector<const void *> &out, const std::vector<const void *> &keys,
                           const std::vector<const void *> &perms) {
    g.table0 = host_tables_.size();
    for (const auto *v : {&a, &b, &out, &keys, &perms})
        for (u32 i = 0; i < g.B; ++i) host_tables_.push_back(v->empty() ? nullptr : (*v)[i]);
    for (u32 i = 0; i < g.B; ++i) host_tables_.push_back(nullptr);   // key-switch outputs, set by rebind
}

void BatchPlan::layout() {
    size_t words = 0;
    for (const auto &g : groups_) words = std::max(words, group_scratch_words(g));
    scratch_bytes_ = words * 8;
    if (!transient_) {
        PHX_HIP_CHECK(hipMalloc(&arena_, std::max<size_t>(scratch_bytes_, 256)));
        void *t = nullptr;
        PHX_HIP_CHECK(hipMalloc(&t, std::max<size_t>(host_tables_.size() * sizeof(void *), 256)));
        tables_ = static_cast<const void **>(t);
        rebind(arena_);
    }
}

void BatchPlan::rebind(u64 *arena) {
    bind(arena, tables_);
    if (kind_ == kBfvMulRelin) {
        // outputs of the key switch: polynomials 0 and 1 of each item's product
        for (const auto &g : groups_) {
            const u64 LN = (u64)g.Ql * ctx_.N();
            for (u32 i = 0; i < g.B; ++i) host_tables_[g.table0 + 5 * g.B + i] = g.c3 + (u64)i * 3 * LN;
        }
    }
    PHX_HIP_CHECK(hipMemcpy(const_cast<void **>(tables_), host_tables_.data(), host_tables_.size() * sizeof(void *),
                            hipMemcpyHostToDevice));
}

void BatchPlan::bind(u64 *arena, const void **tables) {
    const u64 N = ctx_.N();
    for (auto &g : groups_) {
        const u64 LN = (u64)g.Ql * N, LPN = (u64)g.QlP * N;
        u64 *p = arena;
        auto take = [&](u64 words) {
            u64 *r = p;
            p += words;
            return r;
        };
        g.c2_stride = LN;
        if (kind_ == kBfvMulRelin) {
            g.in1 = take((u64)g.B * 2 * LN);
            g.in2 = take((u64)g.B * 2 * LN);
            g.c3 = take((u64)g.B * 3 * LN);
            g.work = take((u64)g.B * bfv_batch_work_words(ctx_));
        }
        g.c2 = take(g.B * LN);
        if (kind_ == kBfvMulRelin) {
            // the key switch reads c2 where the product holds it (NTT form)
            g.c2 = g.c3 + 2 * LN;
            g.c2_stride = 3 * LN;
        }
        g.t_cks = take(g.B * LN);
        g.t_mod_up = take((u64)g.B * g.beta * LPN);
        g.cx = take((u64)g.B * 2 * LPN);
        g.delta = take((u64)g.B * 2 * LN);
        g.base0 = kind_ == kRotate ? take(g.B * LN) : nullptr;
        const void **t = tables + g.table0;
        g.a = reinterpret_cast<const u64 **>(t);
        g.b = reinterpret_cast<const u64 **>(t + g.B);
        g.out = reinterpret_cast<u64 **>(const_cast<void **>(t + 2 * g.B));
        g.keys = reinterpret_cast<const u64 **>(t + 3 * g.B);
        g.perms = reinterpret_cast<const u32 **>(t + 4 * g.B);
        g.ks_out = reinterpret_cast<u64 **>(const_cast<void **>(t + 5 * g.B));
    }
}

BatchPlan *BatchPlan::mul_relin(const Context &ctx, const std::vector<Ciphertext *> &ct1,
                                const std::vector<const Ciphertext *> &ct2, const KSwitchKey &rlk, bool transient) {
    const bool bfv = ctx.params().scheme == kSchemeBFV;
    require(bfv || ctx.params().scheme == kSchemeCKKS, "batched multiply_and_relin: unsupported scheme");
    require(!(bfv && transient), "BFV batches are persistent plans");
    require(ct1.size() == ct2.size() && !ct1.empty(), "batch operands must be non-empty and of equal count");
    require(rlk.data != nullptr, "relinearisation key is empty");
    if (bfv) bfv_batch_work_words(ctx);   // rejects a context without its multiply tool
    std::unique_ptr<BatchPlan> p(new BatchPlan(ctx, bfv ? kBfvMulRelin : kMulRelin, transient));
    std::vector<u32> chains;
    for (size_t i = 0; i < ct1.size(); ++i) {
        require(ct1[i] && ct2[i] && ct1[i]->data && ct2[i]->data, "batch ciphertext is empty");
        require(ct1[i]->chain_index == ct2[i]->chain_index, "encrypted1 and encrypted2 parameter mismatch");
        if (bfv) {
            require(!ct1[i]->is_ntt_form && !ct2[i]->is_ntt_form, "encrypted1 or encrypted2 cannot be in NTT form");
            require(ct1[i]->chain_index == 1, "BFV multiplication runs at the top data level");
        } else {
            require(ct1[i]->is_ntt_form && ct2[i]->is_ntt_form, "encrypted1 and encrypted2 must be in NTT form");
        }
        require(ct1[i]->size == 2 && ct2[i]->size == 2, "poly number mismatch");
        require(rlk.dnum >= ctx.level(ct1[i]->chain_index).beta, "key-switching key has too few digits");
        for (size_t j = 0; j < ct1.size(); ++j) {
            if (j < i) require(ct1[j] != ct1[i], "a ciphertext appears twice as a batch output");
            // an output may only alias its own second operand (squaring)
            if (j != i) require(ct2[j] != ct1[i], "a batch output is another item's operand");
        }
        chains.push_back(ct1[i]->chain_index);
        p->cts_.push_back(ct1[i]);
        p->scale_mul_.push_back(bfv ? 1.0 : ct2[i]->scale);
    }
    p->items_ = ct1.size();
    p->build_groups(chains);
    for (auto &g : p->groups_) {
        std::vector<const void *> a, b, out, keys, square;
        for (u32 i : g.items) {
            a.push_back(ct1[i]->data);
            b.push_back(ct2[i]->data);
            out.push_back(ct1[i]->data);
            keys.push_back(rlk.data);
            // BFV: squaring flag (the same ciphertext as both operands)
            square.push_back(ct1[i]->data == ct2[i]->data ? reinterpret_cast<const void *>(uintptr_t{1}) : nullptr);
        }
        p->set_tables(g, a, b, out, keys, bfv ? square : std::vector<const void *>{});
    }
    p->layout();
    return p.release();
}

BatchPlan *BatchPlan::join(const std::vector<BatchPlan *> &parts) {
    require(!parts.empty(), "nothing to join");
    for (size_t i = 0; i < parts.size(); ++i) {
        require(parts[i] != nullptr, "batch is null");
        require(!parts[i]->transient_ && parts[i]->kind_ != kJoin, "only persistent, unjoined batches can be joined");
        for (size_t j = 0; j < i; ++j) require(parts[j] != parts[i], "a batch appears twice in a join");
    }
    std::unique_ptr<BatchPlan> p(new BatchPlan(parts[0]->ctx_, kJoin, false));
    size_t bytes = 0;
    for (auto *q : parts) {
        bytes = std::max(bytes, q->scratch_bytes_);
        p->items_ += q->items_;
    }
    // the parts run one after another on one stream: one arena of the largest size
    PHX_HIP_CHECK(hipDeviceSynchronize());
    PHX_HIP_CHECK(hipMalloc(&p->arena_, std::max<size_t>(bytes, 256)));
    p->scratch_bytes_ = bytes;
    for (auto *q : parts) {
        PHX_HIP_CHECK(hipFree(q->arena_));
        q->arena_ = nullptr;
        q->rebind(p->arena_);
        if (q->exec_) (void)hipGraphExecDestroy(q->exec_);
        if (q->graph_) (void)hipGraphDestroy(q->graph_);
        q->exec_ = nullptr;
        q->graph_ = nullptr;
        p->parts_.emplace_back(q);
    }
    return p.release();
}

BatchPlan *BatchPlan::rotate(const Context &ctx, const std::vector<Ciphertext *> &cts, const std::vector<int> &steps,
                             const GaloisKeys &gk, bool transient) {
    require(cts.size() == steps.size() && !cts.empty(), "batch operands must be non-empty and of equal count");
    std::unique_ptr<BatchPlan> p(new BatchPlan(ctx, kRotate, transient));
    std::vector<u32> chains;
    std::vector<const u64 *> item_key(cts.size());
    std::vector<const u32 *> item_perm(cts.size());
    for (size_t i = 0; i < cts.size(); ++i) {
        require(cts[i] && cts[i]->data, "batch ciphertext is empty");
        require(cts[i]->size == 2 && cts[i]->is_ntt_form, "encrypted size must be 2 (NTT form)");
        for (size_t j = 0; j < i; ++j) require(cts[j] != cts[i], "a ciphertext appears twice in a batch");
        const u32 elt = galois_elt_from_step(steps[i], ctx.N());
        auto it = std::find(gk.elts.begin(), gk.elts.end(), elt);
        require(it != gk.elts.end(), "Galois elt not present");
        const KSwitchKey &key = gk.keys[it - gk.elts.begin()];
        require(key.dnum >= ctx.level(cts[i]->chain_index).beta, "key-switching key has too few digits");
        item_key[i] = key.data;
        item_perm[i] = ctx.galois_perm(elt);
        chains.push_back(cts[i]->chain_index);
        p->cts_.push_back(cts[i]);
    }
    p->items_ = cts.size();
    p->build_groups(chains);
    for (auto &g : p->groups_) {
        // items sharing a key are placed next to each other so the inner product reuses it
        std::stable_sort(g.items.begin(), g.items.end(),
                         [&](u32 x, u32 y) { return item_key[x] < item_key[y]; });
        std::vector<const void *> a, out, keys, perms;
        for (u32 i : g.items) {
            a.push_back(cts[i]->data);
            out.push_back(cts[i]->data);
            keys.push_back(item_key[i]);
            perms.push_back(item_perm[i]);
        }
        p->set_tables(g, a, {}, out, keys, perms);
    }
    p->layout();
    return p.release();
}

// The same pipeline as keyswitch() in evaluator.hip, each stage over the group.
void BatchPlan::run_group(const Group &g, hipStream_t s) {
    const Context &ctx = ctx_;
    const u64 N = ctx.N();
    const u32 Ql = g.Ql, B = g.B;
    const u64 LN = (u64)Ql * N;
    if (kind_ == kBfvMulRelin) {
        // BFV (evaluate.cu: bfv_multiply_behz / bfv_multiply_hps, then relinearize_inplace):
        // product of every item, its NTT, the key switch of c2 added to (c0, c1) in the
        // NTT domain, and back: INTT(NTT(c) + ks) = c + INTT(ks) exactly
        hipLaunchKernelGGL(k_gather_items, items_grid(2 * LN, B), dim3(kThreads), 0, s, g.a, g.in1, 2 * LN, 2 * LN);
        PHX_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_gather_items, items_grid(2 * LN, B), dim3(kThreads), 0, s, g.b, g.in2, 2 * LN, 2 * LN);
        PHX_HIP_CHECK(hipGetLastError());
        // the perms table of a BFV plan holds each item's squaring flag
        bfv_multiply_batch(ctx, g.in1, g.in2, B, reinterpret_cast<const u64 *>(g.perms), g.work, g.c3, s);
        ntt_forward(ctx.tables(), limb_map(g.c3, 0, Ql, LN), Ql, 3 * B, s);
        keyswitch_group(g, s);
        ntt_inverse(ctx.tables(), limb_map(g.c3, 0, Ql, 3 * LN), Ql, B, s);
        ntt_inverse(ctx.tables(), limb_map(g.c3 + LN, 0, Ql, 3 * LN), Ql, B, s);
        hipLaunchKernelGGL(k_scatter_items, items_grid(2 * LN, B), dim3(kThreads), 0, s, g.out, (const u64 *)g.c3,
                           2 * LN, 3 * LN);
        PHX_HIP_CHECK(hipGetLastError());
        return;
    }
    if (kind_ == kMulRelin) {
        hipLaunchKernelGGL(k_tensor_batch, grid3(N, Ql, B), dim3(kThreads), 0, s, g.a, g.b, g.out, g.c2,
                           ctx.dev_mod(), (u32)N, Ql);
    } else {
        hipLaunchKernelGGL(k_permute_batch, grid3(N, Ql, B), dim3(kThreads), 0, s, g.a, g.perms, g.base0, g.c2,
                           (u32)N, Ql);
    }
    PHX_HIP_CHECK(hipGetLastError());
    keyswitch_group(g, s);
}

// out[item] (+)= keyswitch(c2[item]) in NTT form: the same pipeline as keyswitch()
// in evaluator.hip, each stage over the group
void BatchPlan::keyswitch_group(const Group &g, hipStream_t s) {
    const Context &ctx = ctx_;
    const LevelConsts &L = ctx.level(g.chain);
    const u64 N = ctx.N();
    const u32 Ql = g.Ql, QlP = g.QlP, beta = g.beta, B = g.B, P = ctx.size_P(), alpha = ctx.size_P();
    const u64 LN = (u64)Ql * N, LPN = (u64)QlP * N;
    // modup: INTT(c2) * partQlHatInv for every item
    LimbMap mi = limb_map(g.t_cks, 0, Ql, LN);
    mi.src = g.c2;
    mi.src_poly_stride = g.c2_stride;
    mi.scale = L.modup_scale;
    ntt_inverse(ctx.tables(), mi, Ql, B, s);
    // FastBConv of every digit of every item
    BconvArgs ba{};
    ba.cbuf = ctx.const_buf();
    ba.descs = L.modup_bconv;
    ba.desc_stride = 1;
    ba.jobs_per_batch = beta;
    ba.in = g.t_cks;
    ba.in_job_stride = 0;
    ba.in_batch_stride = LN;
    ba.out = g.t_mod_up;
    ba.out_job_stride = LPN;
    ba.out_batch_stride = (u64)beta * LPN;
    // the inner product reads the digits' own limbs from c2 directly: no copy
    const bool fused_ip = tune_value("ks_fused_ip") != 0;
    ba.copy_src = nullptr;
    ba.copy_job_stride = 0;
    ba.copy_batch_stride = g.c2_stride;
    ba.mod = ctx.dev_mod();
    ba.N = (u32)N;
    bconv_run(ctx, L.h_modup_bconv, ba, B * beta, s);
    // NTT of the complement limbs of every (item, digit)
    LimbMap mf = limb_map(g.t_mod_up, 0, QlP, LPN);
    mf.split = Ql;
    mf.special_row = ctx.size_Q();
    mf.skip_alpha = alpha;
    mf.skip_cap = Ql;
    mf.skip_mod = beta;
    mf.nlimbs = QlP;
    const u32 last_sz = Ql - alpha * (beta - 1);
    require(beta <= 32, "batched key switching supports at most 32 digits");
    require(!fused_ip || beta <= 8, "the fused inner product supports at most 8 digits");
    if (fused_ip) {
        // pass 1 here; pass 2 fused with the inner product
        ntt_forward_cols(ctx.tables(), mf, QlP - last_sz, B * beta, s);
        RowsIpArgs ip{};
        ip.t = g.t_mod_up;
        ip.c2 = g.c2;
        ip.c2_stride = g.c2_stride;
        ip.keys = g.keys;
        ip.cx = g.cx;
        ip.Ql = Ql;
        ip.QlP = QlP;
        ip.size_Q = ctx.size_Q();
        ip.QP = ctx.size_QP();
        ip.beta = beta;
        ip.alpha = alpha;
        ntt_forward_rows_ip(ctx.tables(), ip, B, s);
    } else {
        ntt_forward(ctx.tables(), mf, QlP - last_sz, B * beta, s);
        // inner product (a shared key is read once per coefficient for the group)
        auto ip = beta <= 8 ? k_inner_product_batch<8> : beta <= 16 ? k_inner_product_batch<16>
                                                                    : k_inner_product_batch<32>;
        hipLaunchKernelGGL(ip, dim3((u32)(N / kThreads), QlP), dim3(kThreads), 0, s, g.t_mod_up, g.c2, g.keys, g.cx,
                           ctx.dev_mod(), (u32)N, Ql, QlP, ctx.size_Q(), ctx.size_QP(), beta, alpha, B, g.c2_stride);
        PHX_HIP_CHECK(hipGetLastError());
    }
    // moddown of both halves of every item
    LimbMap md = limb_map(g.cx + LN, ctx.size_Q(), P, LPN);
    md.scale = L.moddown_scale;
    ntt_inverse(ctx.tables(), md, P, 2 * B, s);
    BconvArgs bd{};
    bd.cbuf = ctx.const_buf();
    bd.descs = L.moddown_bconv;
    bd.desc_stride = 0;
    bd.jobs_per_batch = 1;
    bd.in = g.cx;
    bd.in_batch_stride = LPN;
    bd.out = g.delta;
    bd.out_batch_stride = LN;
    bd.mod = ctx.dev_mod();
    bd.N = (u32)N;
    bconv_run(ctx, {L.h_moddown_bconv}, bd, 2 * B, s);
    StoreModDownBatch st{};
    st.out = kind_ == kBfvMulRelin ? g.ks_out : g.out;
    st.base0 = kind_ == kRotate ? g.base0 : nullptr;
    st.cx = g.cx;
    st.pinv = L.pinv;
    st.cx_stride = LPN;
    st.N = (u32)N;
    st.Ql = Ql;
    ntt_forward_store_batch(ctx.tables(), limb_map(g.delta, 0, Ql, LN), Ql, 2 * B, st, s);
}

void BatchPlan::record(hipStream_t s) {
    for (const auto &q : parts_) q->record(s);
    for (const auto &g : groups_) run_group(g, s);
}

void BatchPlan::after_launch() {
    for (const auto &q : parts_) q->after_launch();
    if (kind_ == kMulRelin)
        for (size_t i = 0; i < cts_.size(); ++i) cts_[i]->scale *= scale_mul_[i];
}

void BatchPlan::run(hipStream_t s) {
    if (transient_) {
        Scratch arena(std::max<size_t>(scratch_bytes_, 256), s);
        Scratch t(host_tables_.size() * sizeof(void *), s);
        PHX_HIP_CHECK(hipMemcpyAsync(t.get(), host_tables_.data(), host_tables_.size() * sizeof(void *),
                                     hipMemcpyHostToDevice, s));
        bind(arena.get(), reinterpret_cast<const void **>(t.get()));
        for (const auto &g : groups_) run_group(g, s);
    } else {
        record(s);
    }
    after_launch();
}

void BatchPlan::capture(hipStream_t s) {
    require(!transient_, "a transient batch cannot be captured");
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    PHX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        record(s);
    } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(s, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
    }
    PHX_HIP_CHECK(hipStreamEndCapture(s, &graph_));
    PHX_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

void BatchPlan::replay(hipStream_t s) {
    require(exec_ != nullptr, "batch has not been captured");
    PHX_HIP_CHECK(hipGraphLaunch(exec_, s));
    after_launch();
}

}  // namespace phx
