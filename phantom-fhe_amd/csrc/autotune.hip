// autotune.hip -- picks the kernel variants of one context by timing them.
//
// Searched, in this order (coordinate descent from the process-wide switches):
//   1. the NTT launch form: two launches per transform with pass-2 launches
//      sliced to {all, 128, 256, 512} limbs, or one workgroup of G x 256
//      threads per limb (G in {1, 2, 4} where N / 4096 is a multiple of G) for
//      launches of at least {0, 64, 256} limbs;
//   2. the tile shape of the two passes: row-pass workgroups of {128, 256, 512}
//      threads (8, 16 or 32 rows of 256 coefficients), then column-pass
//      workgroups of {256, 512} threads (16 or 32 columns at N = 2^16); each
//      thread holds 16 coefficients, so the tile is 16 x threads coefficients;
//   3. the FastBConv kernel: matrix cores or VALU;
//   4. for the matrix-core FastBConv, its workgroups per CU: {1, 2, as many as
//      the LDS holds}.
// The radix split of a transform follows from the tile: the column pass runs
// the first log2(N / 256) stages, the row pass the last 8, each as radix-16
// register rounds.
// The objective is the time of one HMult+relin at the top level plus the
// per-pair time of an 8-pair batch plan of the same (its launches are the
// large ones: a candidate that only changes launches above some size must be
// measured on them), on synthetic residues resident in HBM (ciphertexts and a
// relinearisation key of uniform residues), timed with HIP events: medians
// of three runs after a warm-up. A candidate replaces the incumbent only if it
// is at least 2% faster, so timing noise cannot pick a variant. The winner is
// applied to the context and, when a path is given, stored in a JSON cache
// keyed by (scheme, N, size_Q, size_P), which context creation reads when
// PHX_TUNE_CACHE names it.
#include "autotune.h"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>

#include "batch.h"
#include "evaluator.h"
#include "tune.h"

namespace phx {

namespace {
constexpr int kFillThreads = 256;

// d[p][l][n] = pseudo-random residue mod mod[l], for `polys` polynomials of L limbs
__global__ void k_fill_residues(u64 *d, const DevModulus *mod, u32 N, u32 L, u32 polys, u64 seed) {
    const u64 total = (u64)N * L * polys;
    for (u64 i = (u64)blockIdx.x * kFillThreads + threadIdx.x; i < total; i += (u64)gridDim.x * kFillThreads) {
        const u32 limb = (u32)((i / N) % L);
        u64 x = (i ^ seed) * 0x9e3779b97f4a7c15ull;
        x ^= x >> 29;
        x *= 0xbf58476d1ce4e5b9ull;
        x ^= x >> 32;
        d[i] = x % mod[limb].value;
    }
}

void fill(u64 *d, const Context &ctx, u32 L, u32 polys, u64 seed, hipStream_t s) {
    const u64 total = ctx.N() * L * polys;
    const u32 blocks = (u32)std::min<u64>((total + kFillThreads - 1) / kFillThreads, 4096);
    hipLaunchKernelGGL(k_fill_residues, dim3(blocks), dim3(kFillThreads), 0, s, d, ctx.dev_mod(), (u32)ctx.N(), L,
                       polys, seed);
    PHX_HIP_CHECK(hipGetLastError());
}

class EventTimer {
public:
    EventTimer() {
        PHX_HIP_CHECK(hipEventCreate(&a_));
        PHX_HIP_CHECK(hipEventCreate(&b_));
    }
    ~EventTimer() {
        (void)hipEventDestroy(a_);
        (void)hipEventDestroy(b_);
    }
    // median over 3 runs of `reps` calls, after one warm-up call; microseconds per call
    template <class F>
    double median_us(hipStream_t s, int reps, F &&f) {
        f();
        double r[3];
        for (double &x : r) {
            PHX_HIP_CHECK(hipEventRecord(a_, s));
            for (int i = 0; i < reps; ++i) f();
            PHX_HIP_CHECK(hipEventRecord(b_, s));
            PHX_HIP_CHECK(hipEventSynchronize(b_));
            float ms = 0;
            PHX_HIP_CHECK(hipEventElapsedTime(&ms, a_, b_));
            x = 1e3 * ms / reps;
        }
        std::sort(r, r + 3);
        return r[1];
    }

private:
    hipEvent_t a_, b_;
};

std::string describe(const TuneChoice &c) {
    std::ostringstream o;
    o << "ntt_fused=" << c.ntt_fused << ",ntt_fused_min_limbs=" << c.ntt_fused_min_limbs
      << ",ntt_chunk_limbs=" << c.ntt_chunk_limbs << ",bconv_mfma=" << c.bconv_mfma
      << ",ntt_cols_threads=" << c.ntt_cols_threads << ",ntt_rows_threads=" << c.ntt_rows_threads
      << ",bconv_wgs_per_cu=" << c.bconv_wgs_per_cu << ",bconv_fused=" << c.bconv_fused;
    return o.str();
}
}  // namespace

TuneResult autotune(Context &ctx, hipStream_t s) {
    constexpr u32 kBatch = 8;
    constexpr double kMargin = 0.98;   // a candidate must be 2% faster to be adopted
    const u64 N = ctx.N();
    const u32 Ql = ctx.size_Q(), QP = ctx.size_QP(), dnum = ctx.dnum();
    const bool bfv = ctx.params().scheme == kSchemeBFV;
    const TuneChoice saved = ctx.tuning();
    TuneResult res;

    // one plain pair (c[0], c[1]) and kBatch pairs for the batch plan
    std::vector<Ciphertext> c(2 + 2 * kBatch);
    for (size_t i = 0; i < c.size(); ++i) {
        c[i].resize(ctx, 1, 2, s);
        c[i].is_ntt_form = !bfv;
        fill(c[i].data, ctx, Ql, 2, 11 + i, s);
    }
    KSwitchKey rlk;
    rlk.dnum = dnum;
    rlk.poly_degree = N;
    rlk.size_QP = QP;
    rlk.data = static_cast<u64 *>(dev_alloc((size_t)dnum * 2 * QP * N * 8, s));
    fill(rlk.data, ctx, QP, 2 * dnum, 7, s);
    auto release = [&] {
        for (auto &x : c) x.release(s);
        dev_free(rlk.data, s);
    };
    std::unique_ptr<BatchPlan> plan;

    EventTimer timer;
    auto time_choice = [&](const TuneChoice &ch) {
        ctx.set_tuning(ch);
        const double plain = timer.median_us(s, 8, [&] { multiply_and_relin(ctx, c[0], c[1], rlk, s); });
        const double batch = plan ? timer.median_us(s, 2, [&] { plan->run(s); }) / kBatch : 0.0;
        res.trials.emplace_back(describe(ch), plain + batch);
        return plain + batch;
    };

    try {
        PHX_HIP_CHECK(hipStreamSynchronize(s));
        std::vector<Ciphertext *> b1;
        std::vector<const Ciphertext *> b2;
        for (u32 i = 0; i < kBatch; ++i) {
            b1.push_back(&c[2 + 2 * i]);
            b2.push_back(&c[3 + 2 * i]);
        }
        // hps_overq_leveled has no batch plans (its level depends on each ciphertext)
        if (!bfv || ctx.params().mul_tech != kMulHPSOverQLeveled) plan.reset(BatchPlan::mul_relin(ctx, b1, b2, rlk));
        res.default_us = time_choice(TuneChoice{});
        // 1. NTT launch form, FastBConv left at the process-wide switch
        TuneChoice best;
        best.ntt_fused = 0;
        best.ntt_fused_min_limbs = 0;
        best.ntt_chunk_limbs = 0;
        double best_us = time_choice(best);
        std::vector<TuneChoice> cands;
        for (int chunk : {128, 256, 512}) {
            TuneChoice cand = best;
            cand.ntt_chunk_limbs = chunk;
            cands.push_back(cand);
        }
        const u64 tiles = N / kNttTile;
        for (int g : {1, 2, 4}) {
            if (tiles < (u64)g || tiles % (u64)g) continue;
            for (int min_limbs : {0, 64, 256}) {
                TuneChoice cand;
                cand.ntt_fused = g;
                cand.ntt_fused_min_limbs = min_limbs;
                cand.ntt_chunk_limbs = 0;
                cands.push_back(cand);
            }
        }
        for (const TuneChoice &cand : cands) {
            const double us = time_choice(cand);
            if (us < kMargin * best_us) best_us = us, best = cand;
        }
        // 2. tile shape: the row passes, then the column passes (shapes whose tile
        //    exceeds a limb are left out)
        std::vector<int> rows, cols;
        for (int t : {128, 256, 512})
            if ((u64)t * kNttRegs <= N) rows.push_back(t);
        // column passes: 0 = by launch size (the default), or a fixed shape
        cols.push_back(0);
        for (int t : {256, 512})
            if ((u64)t * kNttRegs <= N) cols.push_back(t);
        auto axis = [&](int TuneChoice::*field, const std::vector<int> &v, int inc) {
            if (v.size() < 2) {
                best.*field = inc;
                return;
            }
            TuneChoice inc_c = best;
            inc_c.*field = inc;
            double inc_us = time_choice(inc_c);
            for (int x : v) {
                if (x == inc) continue;
                TuneChoice cand = best;
                cand.*field = x;
                const double us = time_choice(cand);
                if (us < kMargin * inc_us) inc_us = us, inc_c = cand;
            }
            best = inc_c;
            best_us = inc_us;
        };
        axis(&TuneChoice::ntt_rows_threads, rows, 256);
        axis(&TuneChoice::ntt_cols_threads, cols, 0);
        // 3. FastBConv kernel: the matrix-core form stays unless the VALU form wins
        TuneChoice mf = best;
        mf.bconv_mfma = 1;
        const double mf_us = time_choice(mf);
        TuneChoice va = best;
        va.bconv_mfma = 0;
        const double va_us = time_choice(va);
        best = va_us < kMargin * mf_us ? va : mf;
        best_us = va_us < kMargin * mf_us ? va_us : mf_us;
        // 4. the matrix-core FastBConv's workgroups per CU
        if (best.bconv_mfma == 1) axis(&TuneChoice::bconv_wgs_per_cu, {0, 1, 2}, 0);
        // 5. the matrix-core FastBConv fused with the forward NTT's column pass, or separate
        if (best.bconv_mfma == 1) axis(&TuneChoice::bconv_fused, {1, 0}, 1);
        res.choice = best;
        res.best_us = best_us;
    } catch (...) {
        ctx.set_tuning(saved);
        plan.reset();
        release();
        throw;
    }
    ctx.set_tuning(res.choice);
    PHX_HIP_CHECK(hipStreamSynchronize(s));
    plan.reset();
    release();
    return res;
}

// ---- cache: {"entries": [ {...}, ... ]} with one entry per line ----

std::string tune_cache_key(u32 scheme, u64 N, u32 size_Q, u32 size_P) {
    std::ostringstream k;
    k << "scheme=" << (scheme == kSchemeBFV ? "bfv" : "ckks") << ",N=" << N << ",Q=" << size_Q << ",P=" << size_P;
    return k.str();
}

namespace {
bool entry_field(const std::string &line, const char *name, int *v) {
    const std::string f = std::string("\"") + name + "\": ";
    const size_t p = line.find(f);
    if (p == std::string::npos) return false;
    char *end = nullptr;
    const long x = std::strtol(line.c_str() + p + f.size(), &end, 10);
    if (end == line.c_str() + p + f.size()) return false;
    *v = (int)x;
    return true;
}
}  // namespace

void tune_cache_store(const std::string &path, const std::string &key, const TuneChoice &c, double us) {
    require(!path.empty(), "empty tuning cache path");
    const std::string tag = "\"key\": \"" + key + "\"";
    std::vector<std::string> kept;
    {
        std::ifstream in(path);
        std::string line;
        while (std::getline(in, line)) {
            if (line.find("\"key\": \"") == std::string::npos || line.find(tag) != std::string::npos) continue;
            while (!line.empty() && (line.back() == ',' || line.back() == ' ')) line.pop_back();
            kept.push_back(line);
        }
    }
    std::ostringstream e;
    e << "  {" << tag << ", \"ntt_fused\": " << c.ntt_fused << ", \"ntt_fused_min_limbs\": " << c.ntt_fused_min_limbs
      << ", \"ntt_chunk_limbs\": " << c.ntt_chunk_limbs << ", \"bconv_mfma\": " << c.bconv_mfma
      << ", \"ntt_cols_threads\": " << c.ntt_cols_threads << ", \"ntt_rows_threads\": " << c.ntt_rows_threads
      << ", \"bconv_wgs_per_cu\": " << c.bconv_wgs_per_cu << ", \"bconv_fused\": " << c.bconv_fused
      << ", \"hmult_relin_us\": " << us << "}";
    kept.push_back(e.str());
    std::ofstream out(path, std::ios::trunc);
    require((bool)out, "cannot write the tuning cache " + path);
    out << "{\"entries\": [\n";
    for (size_t i = 0; i < kept.size(); ++i) out << kept[i] << (i + 1 < kept.size() ? ",\n" : "\n");
    out << "]}\n";
    require((bool)out, "cannot write the tuning cache " + path);
}

bool tune_cache_lookup(const std::string &path, const std::string &key, TuneChoice *out) {
    std::ifstream in(path);
    if (!in) return false;
    const std::string tag = "\"key\": \"" + key + "\"";
    std::string line;
    while (std::getline(in, line)) {
        if (line.find(tag) == std::string::npos) continue;
        TuneChoice c;
        if (!entry_field(line, "ntt_fused", &c.ntt_fused) ||
            !entry_field(line, "ntt_fused_min_limbs", &c.ntt_fused_min_limbs) ||
            !entry_field(line, "ntt_chunk_limbs", &c.ntt_chunk_limbs) ||
            !entry_field(line, "bconv_mfma", &c.bconv_mfma))
            fail(kInvalidArgument, "malformed tuning cache entry for " + key + " in " + path);
        // axes added later are absent from older entries: they follow the process-wide switches
        entry_field(line, "ntt_cols_threads", &c.ntt_cols_threads);
        entry_field(line, "ntt_rows_threads", &c.ntt_rows_threads);
        entry_field(line, "bconv_wgs_per_cu", &c.bconv_wgs_per_cu);
        entry_field(line, "bconv_fused", &c.bconv_fused);
        *out = c;
        return true;
    }
    return false;
}

bool tune_cache_load(const std::string &path, const Context &ctx, TuneChoice *out) {
    return tune_cache_lookup(path, tune_cache_key((u32)ctx.params().scheme, ctx.N(), ctx.size_Q(), ctx.size_P()),
                             out);
}

}  // namespace phx
