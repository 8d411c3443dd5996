// batch.hip -- batched multiply-and-relinearise, rotation and rescale over work
// lists (see batch.h).
#include <algorithm>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>

#include "batch.h"
#include "ntt_arith.h"
#include "bfv.h"
#include "tune.h"

namespace phx {

namespace {
constexpr int kThreads = 256;
inline dim3 grid3(u64 N, u32 y, u32 z) { return dim3((u32)((N + kThreads - 1) / kThreads), y, z); }
constexpr u32 kMaxGridY = 65535;
}  // namespace

// Galois permutation of item z (apply_galois_ntt_permutation, src/galois.cu:11-18):
// base0 = pi(c0), c2 = pi(c1)
template <class PreItem>
__global__ __launch_bounds__(kThreads) void k_permute_items(const PreItem *items, u32 N) {
    const PreItem it = items[blockIdx.z];
    const u32 limb = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (limb >= it.Ql || k >= N) return;
    const u64 LN = (u64)it.Ql * N, o = (u64)limb * N;
    const u32 pk = gmem(it.perm)[k];
    const auto a = gmem(it.a);
    gmem(it.base0)[o + k] = a[o + pk];
    gmem(it.c2)[o + k] = a[LN + o + pk];
}

// key inner product over the items of a plan (key_switch_inner_prod_c2_and_evk,
// src/eval_key_switch.cu:14-69), grid row = table row r of the QP chain: item i
// has r as data limb pos = r (r < its Ql) or special limb pos = Ql + (r - size_Q);
// cx[pos] = sum_j t[j][pos] * key_j.b[r], cx[QlP + pos] likewise with key_j.a.
// Consecutive items sharing a key read it once per coefficient (the first item of
// a run of equal keys has the most digits). Digit j's own limbs (pos / alpha == j)
// are read from c2 (NTT form). MB = the digit capacity of the key cache.
template <u32 MB, class IpItem>
__global__ __launch_bounds__(kThreads) void k_inner_product_items(const IpItem *items, u32 nitems,
                                                                  const DevModulus *__restrict__ mod, u32 N,
                                                                  u32 size_Q, u32 QP, u32 alpha) {
#pragma clang fp contract(off)
    const u32 r = blockIdx.y;
    const u32 k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= N) return;
    const u64 QPN = (u64)QP * N;
    const DevModulus m = mod[r];
    const bool fp = m.pad != 0;
    const FpArith fa{nullptr, (double)m.value, __longlong_as_double((long long)m.pad)};
    u64 kb[MB], ka[MB];
    const u64 *cur = nullptr;
    for (u32 i = 0; i < nitems; ++i) {
        const IpItem it = items[i];
        if (r < size_Q && r >= it.Ql) continue;
        const u32 pos = r < size_Q ? r : it.Ql + (r - size_Q);
        const u32 own = pos < it.Ql ? pos / alpha : ~0u;
        if (it.key != cur) {
            cur = it.key;
#pragma unroll
            for (u32 j = 0; j < MB; ++j)
                if (j < it.beta) {
                    const auto p = gmem(it.key + (u64)j * 2 * QPN + (u64)r * N + k);
                    kb[j] = p[0];
                    ka[j] = p[QPN];
                }
        }
        const auto ti = gmem(it.t + (u64)pos * N + k);
        const u64 LPN = (u64)it.QlP * N;
        const u64 xo = own != ~0u ? gmem(it.c2)[(u64)pos * N + k] : 0;
        const auto co = gmem(it.cx + (u64)pos * N + k);
        if (fp) {
            // q < 2^50: exact products on the FP64 pipe (FpArith::mulmod, |r| < q), the
            // partial sums reduced every 4 digits (|sum| < 4.5 q < 2^53), then canonical
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (u32 j = 0; j < MB; ++j)
                if (j < it.beta) {
                    if (j && (j & 3) == 0) {
                        s0 = fa.red(s0);
                        s1 = fa.red(s1);
                    }
                    const double x = FpArith::from_canon(j == own ? xo : ti[(u64)j * LPN]);
                    s0 += fa.mulmod(x, FpArith::from_canon(kb[j]));
                    s1 += fa.mulmod(x, FpArith::from_canon(ka[j]));
                }
            co[0] = fa.canon(s0);
            co[LPN] = fa.canon(s1);
        } else {
            u128 a0 = 0, a1 = 0;
#pragma unroll
            for (u32 j = 0; j < MB; ++j)
                if (j < it.beta) {
                    const u64 x = j == own ? xo : ti[(u64)j * LPN];
                    a0 += (u128)x * kb[j];
                    a1 += (u128)x * ka[j];
                }
            co[0] = barrett128(a0, m.value, m.ratio_lo, m.ratio_hi);
            co[LPN] = barrett128(a1, m.value, m.ratio_lo, m.ratio_hi);
        }
    }
}

// BFV staging: dst[z][0 .. words) = (which ? items[z].b : items[z].a)[0 .. words)
template <class PreItem>
__global__ __launch_bounds__(kThreads) void k_gather_items(const PreItem *items, int which, u64 *__restrict__ dst,
                                                           u64 words, u64 dst_stride) {
    const PreItem it = items[blockIdx.y];
    const auto a = gmem(which ? it.b : it.a);
    u64 *d = dst + blockIdx.y * dst_stride;
    for (u64 i = (u64)blockIdx.x * kThreads + threadIdx.x; i < words; i += (u64)gridDim.x * kThreads) d[i] = a[i];
}

// items[z].out[0 .. words) = src[z][0 .. words)
template <class PreItem>
__global__ __launch_bounds__(kThreads) void k_scatter_items(const PreItem *items, const u64 *__restrict__ src,
                                                            u64 words, u64 src_stride) {
    const auto d = gmem(items[blockIdx.y].out);
    const u64 *a = src + blockIdx.y * src_stride;
    for (u64 i = (u64)blockIdx.x * kThreads + threadIdx.x; i < words; i += (u64)gridDim.x * kThreads) d[i] = a[i];
}

namespace {
dim3 items_grid(u64 words, u32 items) {
    return dim3((u32)std::min<u64>((words + kThreads - 1) / kThreads, 1024), items);
}

// scratch words of one item: c2 (CKKS), t_cks, t_mod_up, cx, delta (+ base0 for rotations)
size_t ks_item_words(u64 N, u32 Ql, u32 QlP, u32 beta, bool own_c2, bool rotate) {
    return (size_t)N * ((u64)Ql * ((own_c2 ? 1 : 0) + 1 + (rotate ? 1 : 0)) + (u64)beta * QlP + 2ull * QlP +
                        2ull * Ql);
}

template <class T>
size_t append(std::vector<uint8_t> &buf, const std::vector<T> &v) {
    const size_t off = (buf.size() + 255) & ~size_t{255};
    buf.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// stable order by table row: entries of one prime run next to each other, so
// the prime's twiddles stay in L2
void sort_by_row(std::vector<NttJob> &v) {
    std::stable_sort(v.begin(), v.end(), [](const NttJob &x, const NttJob &y) { return x.row < y.row; });
}
}  // namespace

// ---------------------------------------------------------------------------
BatchPlan::~BatchPlan() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (arena_ || tables_ || !branch_arenas_.empty() || !aux_.empty()) {
        (void)hipDeviceSynchronize();
        (void)hipFree(arena_);
        (void)hipFree(tables_);
        for (u64 *a : branch_arenas_) (void)hipFree(a);
    }
    for (hipStream_t st : aux_) (void)hipStreamDestroy(st);
    for (hipEvent_t e : joins_) (void)hipEventDestroy(e);
    if (fork_) (void)hipEventDestroy(fork_);
}

size_t BatchPlan::item_scratch_bytes(const Context &ctx, u32 chain, Kind k) {
    const LevelConsts &L = ctx.level(chain);
    if (k == kRescale) return 8 * 2 * (size_t)ctx.N();
    return 8 * ks_item_words(ctx.N(), L.size_Ql, L.size_QlP, L.beta, k != kBfvMulRelin, k == kRotate);
}

size_t BatchPlan::item_words(const Item &it) const {
    if (kind_ == kRescale) return 2 * (size_t)ctx_.N();
    return ks_item_words(ctx_.N(), it.Ql, it.QlP, it.beta, kind_ != kBfvMulRelin, kind_ == kRotate);
}

void BatchPlan::layout() {
    const u64 N = ctx_.N();
    bfv_words_ = 0;
    if (kind_ == kBfvMulRelin) {
        const u64 LN = (u64)items_[0].Ql * N;
        bfv_words_ = items_.size() * (7 * LN + bfv_batch_work_words(ctx_));
    }
    size_t w = bfv_words_;
    for (auto &it : items_) {
        it.off = w;
        w += item_words(it);
    }
    scratch_bytes_ = w * 8;
    if (!transient_) {
        PHX_HIP_CHECK(hipMalloc(&arena_, std::max<size_t>(scratch_bytes_, 256)));
        bind(arena_);
        PHX_HIP_CHECK(hipMalloc(&tables_, std::max<size_t>(host_tables_.size(), 256)));
        upload_tables(nullptr, nullptr);
    }
}

void BatchPlan::bind(u64 *arena) {
    const Context &ctx = ctx_;
    const u64 N = ctx.N();
    const u32 size_Q = ctx.size_Q(), P = ctx.size_P(), alpha = ctx.size_P();
    pre_.clear();
    ip_.clear();
    mu_inv_.clear();
    mu_fwd_.clear();
    md_inv_.clear();
    md_fwd_.clear();
    mu_bc_.clear();
    md_bc_.clear();
    mu_desc_host_.clear();
    md_desc_host_.clear();
    mu_fdesc_host_.clear();
    md_fdesc_host_.clear();
    max_Ql_ = max_beta_ = 0;
    if (kind_ == kBfvMulRelin) {
        const u64 B = items_.size(), LN = (u64)items_[0].Ql * N;
        bfv_.in1 = arena;
        bfv_.in2 = arena + B * 2 * LN;
        bfv_.c3 = arena + B * 4 * LN;
        bfv_.work = arena + B * 7 * LN;
    }
    std::vector<u32> seen_levels;
    for (size_t i = 0; i < items_.size(); ++i) {
        const Item &it = items_[i];
        const LevelConsts &L = ctx.level(it.chain);
        const u32 Ql = it.Ql, QlP = it.QlP, beta = it.beta;
        const u64 LN = (u64)Ql * N, LPN = (u64)QlP * N;
        max_Ql_ = std::max(max_Ql_, Ql);
        max_beta_ = std::max(max_beta_, beta);
        u64 *p = arena + it.off;
        if (kind_ == kRescale) {
            // INTT of the last limb of both polynomials, then the forward NTT of
            // [c_last]_{q_j} with the floor-division tail
            // (DRNSTool::divide_and_round_q_last_ntt, src/rns.cu:1128-1184)
            u64 *last = p;
            u64 *out = const_cast<u64 *>(it.b);   // the output ciphertext's buffer
            for (u32 h = 0; h < 2; ++h) {
                mu_inv_.push_back(NttJob{last + h * N, it.a + (h * Ql + Ql - 1) * N, nullptr, nullptr, nullptr,
                                         Ql - 1, 0});
                for (u32 l = 0; l + 1 < Ql; ++l) {
                    u64 *d = out + ((u64)h * (Ql - 1) + l) * N;
                    md_fwd_.push_back(NttJob{d, it.a + ((u64)h * Ql + l) * N, d, last + h * N,
                                             L.rescale_qinv + 2 * l, l, 0});
                }
            }
            continue;
        }
        u64 *c2;
        if (kind_ == kBfvMulRelin) {
            c2 = bfv_.c3 + i * 3 * LN + 2 * LN;
        } else {
            c2 = p;
            p += LN;
        }
        u64 *t_cks = p;
        p += LN;
        u64 *t_mod_up = p;
        p += (u64)beta * LPN;
        u64 *cx = p;
        p += 2 * LPN;
        u64 *delta = p;
        p += 2 * LN;
        u64 *base0 = kind_ == kRotate ? p : nullptr;
        u64 *dst = kind_ == kBfvMulRelin ? bfv_.c3 + i * 3 * LN : it.out;
        pre_.push_back(PreItem{it.a, it.b, it.out, c2, base0, it.perm, Ql, 0});
        // modup: INTT(c2) * partQlHatInv, FastBConv of every digit, NTT of the complements
        // modup INTT: for a multiplication its first pass loads the tensor product
        // (writing d0, d1 to the output and d2 to c2), else it reads c2
        for (u32 l = 0; l < Ql; ++l) {
            if (kind_ == kMulRelin)
                mu_inv_.push_back(NttJob{t_cks + l * N, it.a + l * N, it.out + l * N, it.b + l * N,
                                         L.modup_scale + 4 * l, l, Ql});
            else
                mu_inv_.push_back(NttJob{t_cks + l * N, c2 + l * N, nullptr, nullptr, L.modup_scale + 4 * l, l, 0});
        }
        for (u32 j = 0; j < beta; ++j) {
            mu_bc_.push_back(BconvRow{L.modup_bconv + j, t_cks, t_mod_up + (u64)j * LPN, L.modup_fbconv + j});
            const u32 lo = alpha * j, hi = std::min(alpha * (j + 1), Ql);
            for (u32 pos = 0; pos < QlP; ++pos) {
                if (pos >= lo && pos < hi) continue;
                const u32 row = pos < Ql ? pos : size_Q + (pos - Ql);
                mu_fwd_.push_back(NttJob{t_mod_up + (u64)j * LPN + (u64)pos * N, nullptr, nullptr, nullptr, nullptr,
                                         row, 0});
            }
        }
        if (std::find(seen_levels.begin(), seen_levels.end(), it.chain) == seen_levels.end()) {
            seen_levels.push_back(it.chain);
            mu_desc_host_.insert(mu_desc_host_.end(), L.h_modup_bconv.begin(), L.h_modup_bconv.end());
            md_desc_host_.push_back(L.h_moddown_bconv);
            mu_fdesc_host_.insert(mu_fdesc_host_.end(), L.h_modup_fbconv.begin(), L.h_modup_fbconv.end());
            md_fdesc_host_.push_back(L.h_moddown_fbconv);
        }
        ip_.push_back(IpItem{t_mod_up, c2, cx, it.key, Ql, QlP, beta, it.chain});
        // moddown of both halves: INTT of the P limbs * PHatInv, FastBConv P -> Ql,
        // NTT of delta with the tail dst = base + (cx - NTT(delta)) * P^-1
        for (u32 h = 0; h < 2; ++h) {
            for (u32 pj = 0; pj < P; ++pj)
                md_inv_.push_back(NttJob{cx + (u64)h * LPN + (u64)(Ql + pj) * N, nullptr, nullptr, nullptr,
                                         L.moddown_scale + 4 * pj, size_Q + pj, 0});
            md_bc_.push_back(BconvRow{L.moddown_bconv, cx + (u64)h * LPN, delta + (u64)h * LN, L.moddown_fbconv});
            for (u32 l = 0; l < Ql; ++l) {
                const u64 *base = kind_ == kRotate && h == 0 ? base0 + (u64)l * N : nullptr;
                md_fwd_.push_back(NttJob{delta + (u64)h * LN + (u64)l * N, cx + (u64)h * LPN + (u64)l * N,
                                         dst + (u64)h * LN + (u64)l * N, base, L.pinv + 2 * l, l,
                                         kind_ == kRotate ? 0u : 1u});
            }
        }
    }
    // items sharing a key next to each other, the most digits first (IpItem.pad holds the chain)
    std::stable_sort(ip_.begin(), ip_.end(), [](const IpItem &x, const IpItem &y) {
        return x.key != y.key ? std::less<const u64 *>{}(x.key, y.key) : x.pad < y.pad;
    });
    sort_by_row(mu_inv_);
    sort_by_row(mu_fwd_);
    sort_by_row(md_inv_);
    sort_by_row(md_fwd_);
    host_tables_.clear();
    off_pre_ = append(host_tables_, pre_);
    off_ip_ = append(host_tables_, ip_);
    off_mu_inv_ = append(host_tables_, mu_inv_);
    off_mu_fwd_ = append(host_tables_, mu_fwd_);
    off_md_inv_ = append(host_tables_, md_inv_);
    off_md_fwd_ = append(host_tables_, md_fwd_);
    off_mu_bc_ = append(host_tables_, mu_bc_);
    off_md_bc_ = append(host_tables_, md_bc_);
    off_bfv_sq_ = append(host_tables_, bfv_square_);
}

// staging null: a blocking copy into the plan's own tables (persistent plans);
// else a copy from the pinned staging buffer ordered on s (transient plans)
void BatchPlan::upload_tables(hipStream_t s, uint8_t *staging) {
    if (staging) {
        std::memcpy(staging, host_tables_.data(), host_tables_.size());
        PHX_HIP_CHECK(hipMemcpyAsync(tables_bound_, staging, host_tables_.size(), hipMemcpyHostToDevice, s));
    } else {
        PHX_HIP_CHECK(hipMemcpy(tables_, host_tables_.data(), host_tables_.size(), hipMemcpyHostToDevice));
        tables_bound_ = tables_;
    }
}

BatchPlan *BatchPlan::mul_relin(const Context &ctx, const std::vector<Ciphertext *> &ct1,
                                const std::vector<const Ciphertext *> &ct2, const KSwitchKey &rlk, bool transient) {
    const bool bfv = ctx.params().scheme == kSchemeBFV;
    require(bfv || ctx.params().scheme == kSchemeCKKS, "batched multiply_and_relin: unsupported scheme");
    require(!(bfv && transient), "BFV batches are persistent plans");
    require(ct1.size() == ct2.size() && !ct1.empty(), "batch operands must be non-empty and of equal count");
    require(ct1.size() <= kMaxGridY, "too many ciphertexts in one batch");
    require(rlk.data != nullptr, "relinearisation key is empty");
    if (bfv) bfv_batch_work_words(ctx);   // rejects a context without its multiply tool
    std::unique_ptr<BatchPlan> p(new BatchPlan(ctx, bfv ? kBfvMulRelin : kMulRelin, transient));
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
        const LevelConsts &L = ctx.level(ct1[i]->chain_index);
        require(L.beta <= 32, "batched key switching supports at most 32 digits");
        Item it;
        it.chain = ct1[i]->chain_index;
        it.Ql = L.size_Ql;
        it.QlP = L.size_QlP;
        it.beta = L.beta;
        it.a = ct1[i]->data;
        it.b = ct2[i]->data;
        it.out = ct1[i]->data;
        it.key = rlk.data;
        p->items_.push_back(it);
        p->cts_.push_back(ct1[i]);
        p->scale_mul_.push_back(bfv ? 1.0 : ct2[i]->scale);
        // BFV: squaring flag (the same ciphertext as both operands)
        if (bfv) p->bfv_square_.push_back(ct1[i]->data == ct2[i]->data ? 1 : 0);
    }
    return finish(std::move(p));
}

namespace {
u32 branches_wanted() { return (u32)std::clamp(tune_value("batch_branches"), 1, 8); }
}  // namespace

BatchPlan *BatchPlan::finish(std::unique_ptr<BatchPlan> p) {
    const size_t n = p->items_.size();
    const u32 nb = p->transient_ ? 1u : (u32)std::min<size_t>(branches_wanted(), n / 2);
    if (nb <= 1) {
        p->layout();
        return p.release();
    }
    // items sharing a key-switching key go to the same run where possible, so each
    // branch's key inner product reads its keys once (rotations by many steps);
    // the items' order within the plan does not matter to their results
    {
        std::vector<size_t> order(n);
        std::iota(order.begin(), order.end(), size_t{0});
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t x, size_t y) { return std::less<const u64 *>{}(p->items_[x].key, p->items_[y].key); });
        auto permute = [&](auto &v) {
            if (v.empty()) return;
            auto copy = v;
            for (size_t k = 0; k < n; ++k) v[k] = copy[order[k]];
        };
        permute(p->items_);
        permute(p->cts_);
        permute(p->ins_);
        permute(p->scale_mul_);
        permute(p->bfv_square_);
    }
    // contiguous runs of items of about equal work (scratch words as its measure):
    // an item joins the run when its midpoint lies below the run's share of the
    // total; every run keeps at least one item
    std::vector<size_t> w(n);
    for (size_t i = 0; i < n; ++i) w[i] = p->item_words(p->items_[i]);
    const double total = (double)std::accumulate(w.begin(), w.end(), size_t{0});
    std::vector<size_t> cut(nb + 1, 0);
    cut[nb] = n;
    double acc = 0;
    size_t i = 0;
    for (u32 b = 1; b < nb; ++b) {
        const double target = total * b / nb;
        while (i < n - (nb - b) && (i == cut[b - 1] || acc + 0.5 * (double)w[i] <= target)) acc += (double)w[i++];
        cut[b] = i;
    }
    std::vector<std::unique_ptr<BatchPlan>> parts;
    for (u32 b = 0; b < nb; ++b) {
        std::unique_ptr<BatchPlan> q(new BatchPlan(p->ctx_, p->kind_, false));
        const size_t lo = cut[b], hi = cut[b + 1];
        auto slice = [&](const auto &from, auto &to) {
            if (!from.empty()) to.assign(from.begin() + lo, from.begin() + hi);
        };
        slice(p->items_, q->items_);
        slice(p->cts_, q->cts_);
        slice(p->ins_, q->ins_);
        slice(p->scale_mul_, q->scale_mul_);
        slice(p->bfv_square_, q->bfv_square_);
        q->layout();
        parts.push_back(std::move(q));
    }
    std::vector<BatchPlan *> raw;
    for (auto &q : parts) raw.push_back(q.get());
    BatchPlan *joined = join_parts(raw, nb);
    for (auto &q : parts) (void)q.release();   // now owned by the joined plan
    return joined;
}

BatchPlan *BatchPlan::join(const std::vector<BatchPlan *> &parts) {
    require(!parts.empty(), "nothing to join");
    for (size_t i = 0; i < parts.size(); ++i) {
        require(parts[i] != nullptr, "batch is null");
        require(!parts[i]->transient_, "only persistent batches can be joined");
        for (size_t j = 0; j < i; ++j) require(parts[j] != parts[i], "a batch appears twice in a join");
    }
    // joined plans are flattened into their parts
    std::vector<BatchPlan *> flat;
    for (auto *q : parts) {
        if (q->kind_ != kJoin) flat.push_back(q);
        else
            for (auto &x : q->parts_) flat.push_back(x.get());
    }
    BatchPlan *joined = join_parts(flat, (u32)std::min<size_t>(branches_wanted(), flat.size()));
    // the joined plan owns every part now; flattened plans are emptied and freed
    for (auto *q : parts) {
        if (q->kind_ != kJoin) continue;
        for (auto &x : q->parts_) (void)x.release();
        q->parts_.clear();
        delete q;
    }
    return joined;
}

BatchPlan *BatchPlan::join_parts(const std::vector<BatchPlan *> &parts, u32 nb) {
    std::unique_ptr<BatchPlan> p(new BatchPlan(parts[0]->ctx_, kJoin, false));
    nb = std::max<u32>(1, std::min<u32>(nb, (u32)parts.size()));
    // the largest part first onto the least loaded branch (scratch bytes as the
    // measure of work)
    std::vector<size_t> order(parts.size());
    std::iota(order.begin(), order.end(), size_t{0});
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return parts[x]->scratch_bytes_ > parts[y]->scratch_bytes_; });
    std::vector<size_t> load(nb, 0);
    std::vector<u32> br(parts.size(), 0);
    for (size_t i : order) {
        const u32 b = (u32)(std::min_element(load.begin(), load.end()) - load.begin());
        br[i] = b;
        load[b] += parts[i]->scratch_bytes_;
    }
    // a branch of one part keeps that part's own arena; the parts of a branch of
    // several share one arena of the largest part's size. Every allocation comes
    // first, so a failure leaves the parts as they were.
    std::vector<std::vector<BatchPlan *>> members(nb);
    for (size_t i = 0; i < parts.size(); ++i) members[br[i]].push_back(parts[i]);
    std::vector<u64 *> fresh(nb, nullptr);
    size_t total = 0;
    try {
        for (u32 b = 0; b < nb; ++b) {
            size_t bytes = 0;
            for (auto *q : members[b]) bytes = std::max(bytes, q->scratch_bytes_);
            total += bytes;
            if (members[b].size() == 1 && members[b][0]->arena_) continue;
            PHX_HIP_CHECK(hipMalloc(&fresh[b], std::max<size_t>(bytes, 256)));
        }
        if (nb > 1) {
            PHX_HIP_CHECK(hipEventCreateWithFlags(&p->fork_, hipEventDisableTiming));
            for (u32 b = 1; b < nb; ++b) {
                hipStream_t st = nullptr;
                PHX_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
                p->aux_.push_back(st);
                hipEvent_t e = nullptr;
                PHX_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                p->joins_.push_back(e);
            }
        }
    } catch (...) {
        for (u64 *a : fresh) (void)hipFree(a);
        throw;
    }
    PHX_HIP_CHECK(hipDeviceSynchronize());
    for (u32 b = 0; b < nb; ++b) {
        if (!fresh[b]) continue;
        if (members[b].size() == 1) {
            members[b][0]->arena_ = fresh[b];
        } else {
            p->branch_arenas_.push_back(fresh[b]);
            for (auto *q : members[b]) {
                (void)hipFree(q->arena_);
                q->arena_ = nullptr;
            }
        }
        for (auto *q : members[b]) {
            q->bind(fresh[b]);
            q->upload_tables(nullptr, nullptr);
        }
    }
    for (size_t i = 0; i < parts.size(); ++i) {
        BatchPlan *q = parts[i];
        if (q->exec_) (void)hipGraphExecDestroy(q->exec_);
        if (q->graph_) (void)hipGraphDestroy(q->graph_);
        q->exec_ = nullptr;
        q->graph_ = nullptr;
        p->joined_items_ += q->size();
        p->branch_.push_back(br[i]);
        p->parts_.emplace_back(q);
    }
    p->nbranch_ = nb;
    p->scratch_bytes_ = total;
    return p.release();
}

BatchPlan *BatchPlan::rotate(const Context &ctx, const std::vector<Ciphertext *> &cts, const std::vector<int> &steps,
                             const GaloisKeys &gk, bool transient) {
    require(cts.size() == steps.size() && !cts.empty(), "batch operands must be non-empty and of equal count");
    require(cts.size() <= kMaxGridY, "too many ciphertexts in one batch");
    require(ctx.params().scheme == kSchemeCKKS, "batched rotation runs CKKS ciphertexts");
    std::unique_ptr<BatchPlan> p(new BatchPlan(ctx, kRotate, transient));
    for (size_t i = 0; i < cts.size(); ++i) {
        require(cts[i] && cts[i]->data, "batch ciphertext is empty");
        require(cts[i]->size == 2 && cts[i]->is_ntt_form, "encrypted size must be 2 (NTT form)");
        for (size_t j = 0; j < i; ++j) require(cts[j] != cts[i], "a ciphertext appears twice in a batch");
        const u32 elt = galois_elt_from_step(steps[i], ctx.N());
        auto itk = std::find(gk.elts.begin(), gk.elts.end(), elt);
        require(itk != gk.elts.end(), "Galois elt not present");
        const KSwitchKey &key = gk.keys[itk - gk.elts.begin()];
        const LevelConsts &L = ctx.level(cts[i]->chain_index);
        require(key.dnum >= L.beta, "key-switching key has too few digits");
        require(L.beta <= 32, "batched key switching supports at most 32 digits");
        Item it;
        it.chain = cts[i]->chain_index;
        it.Ql = L.size_Ql;
        it.QlP = L.size_QlP;
        it.beta = L.beta;
        it.a = cts[i]->data;
        it.out = cts[i]->data;
        it.key = key.data;
        it.perm = ctx.galois_perm(elt);
        p->items_.push_back(it);
        p->cts_.push_back(cts[i]);
    }
    return finish(std::move(p));
}

BatchPlan *BatchPlan::rescale(const Context &ctx, const std::vector<const Ciphertext *> &ins,
                              const std::vector<Ciphertext *> &outs, bool transient) {
    require(ins.size() == outs.size() && !ins.empty(), "batch operands must be non-empty and of equal count");
    require(ins.size() <= kMaxGridY, "too many ciphertexts in one batch");
    require(ctx.params().scheme == kSchemeCKKS, "rescale_to_next: unsupported scheme");
    std::unique_ptr<BatchPlan> p(new BatchPlan(ctx, kRescale, transient));
    for (size_t i = 0; i < ins.size(); ++i) {
        const Ciphertext *in = ins[i];
        Ciphertext *out = outs[i];
        require(in && out && in->data && out->data && in != out, "rescale needs distinct input and output");
        require(in->size == 2 && in->is_ntt_form, "encrypted size must be 2 (NTT form)");
        require(in->chain_index >= 1 && in->chain_index < ctx.size_Q(), "end of modulus switching chain reached");
        require(out->chain_index == in->chain_index + 1 && out->size == 2 &&
                    out->coeff_modulus_size == in->coeff_modulus_size - 1,
                "rescale output must be sized for the next level");
        Item it;
        it.chain = in->chain_index;
        it.Ql = in->coeff_modulus_size;
        it.a = in->data;
        it.b = out->data;
        p->items_.push_back(it);
        p->ins_.push_back(in);
        p->cts_.push_back(out);
    }
    return finish(std::move(p));
}

size_t BatchPlan::launches() const {
    size_t n = 0;
    for (const auto &q : parts_) n += q->launches();
    if (kind_ == kJoin) return n;
    auto ntt = [](size_t jobs) { return 2 * ((jobs + kMaxGridY - 1) / kMaxGridY); };
    if (kind_ == kRescale) return ntt(mu_inv_.size()) + ntt(md_fwd_.size());
    n += (kind_ == kRotate ? 1 : 0) + ntt(mu_inv_.size()) + 1 + ntt(mu_fwd_.size()) + 1 + ntt(md_inv_.size()) + 1 +
         ntt(md_fwd_.size());
    return n;
}

// out[item] (+)= keyswitch(c2[item]) in NTT form for every item: the pipeline of
// keyswitch() in evaluator.hip, each stage one launch over the work lists
void BatchPlan::record_keyswitch(hipStream_t s) {
    const Context &ctx = ctx_;
    const u64 N = ctx.N();
    const NttTables &T = ctx.tables();
    auto tab = [&](size_t off) { return tables_bound_ + off; };
    // modup (for a multiplication the tensor product is the INTT's first load)
    // the fused FastBConv takes its inputs byte-flipped from the INTT's store
    const bool mu_fused = bconv_fused_ok(ctx, mu_fdesc_host_), md_fused = bconv_fused_ok(ctx, md_fdesc_host_);
    if (kind_ == kMulRelin)
        ntt_inverse_jobs_tensor(T, reinterpret_cast<const NttJob *>(tab(off_mu_inv_)), (u32)mu_inv_.size(), s,
                                mu_fused);
    else
        ntt_inverse_jobs(T, reinterpret_cast<const NttJob *>(tab(off_mu_inv_)), (u32)mu_inv_.size(), s, mu_fused);
    require(mu_bc_.size() <= kMaxGridY && md_bc_.size() <= kMaxGridY, "too many base conversions in one batch");
    BconvArgs ba{};
    ba.rows = reinterpret_cast<const BconvRow *>(tab(off_mu_bc_));
    ba.cbuf = ctx.const_buf();
    ba.mod = ctx.dev_mod();
    ba.N = (u32)N;
    // FastBConv and the forward NTT of its outputs: fused with the column pass, or separate
    if (mu_fused) {
        bconv_fwd_cols_run(ctx, mu_fdesc_host_, ba, (u32)mu_bc_.size(), s);
        ntt_forward_rows_jobs(T, reinterpret_cast<const NttJob *>(tab(off_mu_fwd_)), (u32)mu_fwd_.size(), s);
    } else {
        bconv_run(ctx, mu_desc_host_, ba, (u32)mu_bc_.size(), s);
        ntt_forward_jobs(T, reinterpret_cast<const NttJob *>(tab(off_mu_fwd_)), (u32)mu_fwd_.size(), s);
    }
    // key inner product
    auto ip = max_beta_ <= 8 ? k_inner_product_items<8, IpItem>
                             : max_beta_ <= 16 ? k_inner_product_items<16, IpItem> : k_inner_product_items<32, IpItem>;
    hipLaunchKernelGGL(ip, dim3((u32)(N / kThreads), ctx.size_QP()), dim3(kThreads), 0, s,
                       reinterpret_cast<const IpItem *>(tab(off_ip_)), (u32)ip_.size(), ctx.dev_mod(), (u32)N,
                       ctx.size_Q(), ctx.size_QP(), ctx.size_P());
    PHX_HIP_CHECK(hipGetLastError());
    // moddown
    ntt_inverse_jobs(T, reinterpret_cast<const NttJob *>(tab(off_md_inv_)), (u32)md_inv_.size(), s, md_fused);
    BconvArgs bd{};
    bd.rows = reinterpret_cast<const BconvRow *>(tab(off_md_bc_));
    bd.cbuf = ctx.const_buf();
    bd.mod = ctx.dev_mod();
    bd.N = (u32)N;
    if (md_fused) {
        bconv_fwd_cols_run(ctx, md_fdesc_host_, bd, (u32)md_bc_.size(), s);
        ntt_forward_rows_jobs_moddown(T, reinterpret_cast<const NttJob *>(tab(off_md_fwd_)), (u32)md_fwd_.size(), s);
    } else {
        bconv_run(ctx, md_desc_host_, bd, (u32)md_bc_.size(), s);
        ntt_forward_jobs_moddown(T, reinterpret_cast<const NttJob *>(tab(off_md_fwd_)), (u32)md_fwd_.size(), s);
    }
}

void BatchPlan::record(hipStream_t s) {
    if (kind_ == kJoin) {
        if (nbranch_ <= 1) {
            for (const auto &q : parts_) q->record(s);
            return;
        }
        // fork: every branch stream waits for the work queued on s so far
        PHX_HIP_CHECK(hipEventRecord(fork_, s));
        for (hipStream_t st : aux_) PHX_HIP_CHECK(hipStreamWaitEvent(st, fork_, 0));
        for (size_t i = 0; i < parts_.size(); ++i) parts_[i]->record(branch_[i] == 0 ? s : aux_[branch_[i] - 1]);
        // join: s waits for every branch
        for (size_t b = 0; b < aux_.size(); ++b) {
            PHX_HIP_CHECK(hipEventRecord(joins_[b], aux_[b]));
            PHX_HIP_CHECK(hipStreamWaitEvent(s, joins_[b], 0));
        }
        return;
    }
    const Context &ctx = ctx_;
    const u64 N = ctx.N();
    const NttTables &T = ctx.tables();
    const PreItem *pre = reinterpret_cast<const PreItem *>(tables_bound_ + off_pre_);
    const u32 B = (u32)items_.size();
    if (kind_ == kRescale) {
        ntt_inverse_jobs(T, reinterpret_cast<const NttJob *>(tables_bound_ + off_mu_inv_), (u32)mu_inv_.size(), s);
        ntt_forward_jobs_rescale(T, reinterpret_cast<const NttJob *>(tables_bound_ + off_md_fwd_),
                                 (u32)md_fwd_.size(), s);
        return;
    }
    if (kind_ == kBfvMulRelin) {
        // BFV (evaluate.cu: bfv_multiply_behz / bfv_multiply_hps, then relinearize_inplace):
        // product of every item, its NTT, the key switch of c2 added to (c0, c1) in the
        // NTT domain, and back: INTT(NTT(c) + ks) = c + INTT(ks) exactly
        const u32 Ql = items_[0].Ql;
        const u64 LN = (u64)Ql * N;
        hipLaunchKernelGGL(k_gather_items<PreItem>, items_grid(2 * LN, B), dim3(kThreads), 0, s, pre, 0, bfv_.in1,
                           2 * LN, 2 * LN);
        PHX_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_gather_items<PreItem>, items_grid(2 * LN, B), dim3(kThreads), 0, s, pre, 1, bfv_.in2,
                           2 * LN, 2 * LN);
        PHX_HIP_CHECK(hipGetLastError());
        bfv_multiply_batch(ctx, bfv_.in1, bfv_.in2, B, reinterpret_cast<const u64 *>(tables_bound_ + off_bfv_sq_),
                           bfv_.work, bfv_.c3, s);
        ntt_forward(T, limb_map(bfv_.c3, 0, Ql, LN), Ql, 3 * B, s);
        record_keyswitch(s);
        ntt_inverse(T, limb_map(bfv_.c3, 0, Ql, 3 * LN), Ql, B, s);
        ntt_inverse(T, limb_map(bfv_.c3 + LN, 0, Ql, 3 * LN), Ql, B, s);
        hipLaunchKernelGGL(k_scatter_items<PreItem>, items_grid(2 * LN, B), dim3(kThreads), 0, s, pre,
                           (const u64 *)bfv_.c3, 2 * LN, 3 * LN);
        PHX_HIP_CHECK(hipGetLastError());
        return;
    }
    if (kind_ == kRotate) {
        hipLaunchKernelGGL(k_permute_items<PreItem>, grid3(N, max_Ql_, B), dim3(kThreads), 0, s, pre, (u32)N);
        PHX_HIP_CHECK(hipGetLastError());
    }
    record_keyswitch(s);
}

void BatchPlan::after_launch() {
    for (const auto &q : parts_) q->after_launch();
    if (kind_ == kMulRelin)
        for (size_t i = 0; i < cts_.size(); ++i) cts_[i]->scale *= scale_mul_[i];
    if (kind_ == kRescale)
        for (size_t i = 0; i < cts_.size(); ++i) {
            const Ciphertext &in = *ins_[i];
            Ciphertext &out = *cts_[i];
            out.scale = in.scale / (double)ctx_.moduli()[in.coeff_modulus_size - 1];
            out.is_ntt_form = true;
            out.correction_factor = in.correction_factor;
            out.noise_scale_deg = in.noise_scale_deg;
            out.is_asymmetric = in.is_asymmetric;
            out.has_seed = false;
        }
}

size_t BatchPlan::table_bytes() const {
    // the sections bind() writes, each aligned to 256 bytes
    size_t pre = 0, ip = 0, mu_inv = 0, mu_fwd = 0, md_inv = 0, md_fwd = 0, mu_bc = 0, md_bc = 0;
    const size_t P = ctx_.size_P();
    for (const Item &it : items_) {
        if (kind_ == kRescale) {
            mu_inv += 2;
            md_fwd += 2 * (it.Ql - 1);
            continue;
        }
        pre += 1;
        ip += 1;
        mu_inv += it.Ql;
        mu_fwd += (size_t)it.beta * it.QlP - it.Ql;
        md_inv += 2 * P;
        md_fwd += 2 * it.Ql;
        mu_bc += it.beta;
        md_bc += 2;
    }
    const size_t sizes[9] = {pre * sizeof(PreItem), ip * sizeof(IpItem), mu_inv * sizeof(NttJob),
                             mu_fwd * sizeof(NttJob), md_inv * sizeof(NttJob), md_fwd * sizeof(NttJob),
                             mu_bc * sizeof(BconvRow), md_bc * sizeof(BconvRow), bfv_square_.size() * sizeof(u64)};
    size_t off = 0;
    for (size_t v : sizes) off = ((off + 255) & ~size_t{255}) + v;
    return std::max<size_t>(off, 256);
}

void BatchPlan::run_with(hipStream_t s, u64 *arena, uint8_t *tables, uint8_t *staging) {
    require(transient_, "only transient plans run in caller memory");
    require(staging != nullptr, "a transient plan needs pinned staging for its work lists");
    bind(arena);
    require(host_tables_.size() <= table_bytes(), "batch work lists larger than their buffer");
    tables_bound_ = tables;
    upload_tables(s, staging);
    record(s);
    tables_bound_ = nullptr;
    after_launch();
}

void BatchPlan::run(hipStream_t s) {
    if (transient_) {
        // a one-off run: device memory from the pool, the staging copy completed
        // before the staging buffer is freed
        Scratch arena(std::max<size_t>(scratch_bytes_, 256), s);
        Scratch t(table_bytes(), s);
        void *staging = nullptr;
        PHX_HIP_CHECK(hipHostMalloc(&staging, table_bytes(), hipHostMallocDefault));
        try {
            run_with(s, arena.get(), reinterpret_cast<uint8_t *>(t.get()), static_cast<uint8_t *>(staging));
            PHX_HIP_CHECK(hipStreamSynchronize(s));
        } catch (...) {
            (void)hipStreamSynchronize(s);
            (void)hipHostFree(staging);
            throw;
        }
        PHX_HIP_CHECK(hipHostFree(staging));
        return;
    }
    record(s);
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
