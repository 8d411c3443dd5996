// ntt_kernels.h -- two-pass negacyclic NTT / INTT for gfx950.
//
// A limb of N = R * 256 coefficients is viewed as R rows x 256 columns
// (index i = r*256 + c). The forward transform (Cooley-Tukey, natural order
// in, bit-reversed order out -- the reference's output order,
// src/ntt/fntt_2d.cu:9-198) runs its first log2(R) stages on columns (pass
// 1) and the last 8 stages on rows (pass 2). The inverse (Gentleman-Sande,
// src/ntt/intt_2d.cu:9-207) mirrors it: rows first, then columns.
//
// Each workgroup owns a 4096-coefficient tile (256 threads x 16 registers).
// Within a pass, stages are applied in register "rounds" of up to 4 stages
// (radix-16) with one LDS exchange between rounds; each pass streams its
// tile from HBM once and back once.
//
// Twiddles: forward table tw[row][k] = (psi^brv(k), Shoup), inverse
// itw[row][k] = (psi^-brv(k), Shoup); stage with m groups and group g uses
// entry m + g, as in the reference tables (src/host/ntt.cu:11-56), except
// that entries [N/16, N) are stored transposed (tw_slot): only the last
// register round of the row passes (distances 8 .. 1) reads them, a thread
// there needs 8 >> jj consecutive entries per stage, and transposed the 16
// threads of a row read 16 consecutive entries per load, as for the data.
#pragma once

#include "common.h"
#include "ntt_arith.h"

namespace phx {

// Every NTT pass streams its tiles through HBM once (the working set of a
// launch is far above the L2), so data and intermediate accesses are
// nontemporal: they do not displace the twiddle tables from L2.
// (An A/B with cached intermediates, so that a launch slice's intermediate could
// stay in the last-level cache, measured no difference.)
__device__ __forceinline__ u64 ld_stream(const u64 *p) { return __builtin_nontemporal_load(gmem(p)); }
__device__ __forceinline__ void st_stream(u64 *p, u64 v) { __builtin_nontemporal_store(v, gmem(p)); }
// the intermediate between the two passes
__device__ __forceinline__ u64 ld_mid(const u64 *p) { return ld_stream(p); }
__device__ __forceinline__ void st_mid(u64 *p, u64 v) { st_stream(p, v); }

constexpr int kNttThreads = 256;
constexpr int kNttRegs = 16;
constexpr int kNttTile = kNttThreads * kNttRegs;  // 4096

// Per-launch limb mapping: blockIdx.z selects a limb of the call, blockIdx.y
// + poly0 a polynomial (stride poly_stride elements). Limbs [skip_lo,
// skip_hi) are not launched (the digit's own limbs in modup, already in NTT
// form).
struct LimbMap {
    u64 *data;          // limb 0 of poly 0
    const u64 *src;     // out-of-place source for the first inverse pass (or null)
    const u64 *scale;   // per-limb [s0, s0_shoup, s1, s1_shoup] for the final inverse stage (or null)
    u64 poly_stride;    // elements between polynomials (blockIdx.z)
    u64 src_poly_stride;
    u32 first_row;      // table row of limb 0
    u32 split;          // limbs >= split use special rows
    u32 special_row;    // table row of limb `split`
    u32 skip_lo, skip_hi;
    u32 skip_alpha;     // nonzero: poly z skips limbs [alpha*d, min(alpha*(d+1), skip_cap)),
    u32 skip_cap;       //   digit d = z (skip_mod == 0) or z % skip_mod
    u32 skip_mod;
    u32 nlimbs;         // limbs >= nlimbs are not processed (ragged launches)
    u32 poly0;          // polynomial of blockIdx.y == 0 (launches split over polynomials)
    u32 pad_;

    __device__ __forceinline__ u32 limb(u32 y, u32 z) const {
        u32 lo = skip_lo, hi = skip_hi;
        if (skip_alpha) {
            lo = skip_alpha * (skip_mod ? z % skip_mod : z);
            hi = lo + skip_alpha < skip_cap ? lo + skip_alpha : skip_cap;
        }
        return y < lo ? y : y + (hi - lo);
    }
    __device__ __forceinline__ u32 row(u32 l) const {
        return l < split ? first_row + l : special_row + (l - split);
    }
};

// Work-list form of a launch: one entry per (polynomial, limb) to transform,
// each with its own buffer and table row, so that one launch covers items of
// different levels (ragged limb counts, separate buffers). Grid row y runs
// entry y. The aux pointers and flags carry the load/store policy's operands
// (out-of-place source of an inverse first pass; the moddown or rescale tail's
// inputs and output); scale: the final inverse stage's [s0, s0', s1, s1']
// (null = the row's n^-1 constants) or a policy's constant pair.
struct NttJob {
    u64 *data;          // the limb transformed in place
    const u64 *aux0;
    u64 *aux1;
    const u64 *aux2;
    const u64 *scale;
    u32 row;            // table row of the limb's prime
    u32 flags;
};

// DevModulus.pad of a row holds the bits of fl(1/q) when the row takes the
// FP64 arithmetic (q < 2^50), 0 otherwise.
struct NttTables {
    const TwPair *tw;        // [rows][N]
    const TwPair *itw;       // [rows][N]
    const double *twd;       // [rows][N] forward twiddles as doubles (FP64 rows)
    const double *itwd;      // [rows][N]
    const DevModulus *mod;   // [rows]
    const u64 *inv_consts;   // [rows][4]: n^-1, shoup, n^-1 * itw[1], shoup
    const u64 *int_consts;   // [rows][4]: -q, -2q, -3q, 3q + 1 mod 2^64 (the integer butterflies)
    u32 logN;
    u32 N;
    // launch shape chosen for this context by the autotuner; -1 = follow the
    // process-wide switch of the same name (ntt_fused, ntt_fused_min_limbs,
    // ntt_chunk_limbs)
    int fused_g;
    int fused_min;
    int chunk_limbs;
    // tile shape: threads of the column / row pass workgroups (128, 256, 512; -1 =
    // the process-wide switches ntt_cols_threads / ntt_rows_threads)
    int cols_nt;
    int rows_nt;
};

// Storage slot of table entry k (N >= 2048): entries below N/16 in place; the
// block [m, 2m) of the stage at distance 2^jj (m = N >> (jj+1), jj < 4) holds
// per row r of 256 coefficients a 16 x c matrix (c = 8 >> jj; entry
// m + r*16c + h*c + i for thread column h and i < c) stored as c x 16.
__host__ __device__ inline u64 tw_slot(u64 k, u64 N) {
    if (k < N / 16) return k;
    u64 m = N / 2;
    int jj = 0;
    while (k < m) {
        m >>= 1;
        ++jj;
    }
    const u64 c = 8 >> jj, off = k - m, r = off / (16 * c), w = off % (16 * c);
    return m + r * 16 * c + (w % c) * 16 + w / c;
}

// ---- register rounds ------------------------------------------------------
// v holds (16 >> NB) blocks of 2^NB values; value e of block b has global
// index base[b] + (e << P). A round applies the NB stages whose pair distance
// is 2^(P+NB-1) ... 2^P. Stage with distance t = 2^(P+jj) has
// m = N / (2t) groups; the group of index i is i >> (P+jj+1), and its twiddle
// is entry m + group of the bit-reversed table. P = 0 is the last round of a
// row pass (base = row * 256 + 16 h): its entries are read from tw_slot's
// transposed storage.
template <int P, int NB>
__device__ __forceinline__ u32 tw_index(u32 m, u32 base, int jj, int e) {
    const u32 k = (u32)e >> (jj + 1);
    if constexpr (P == 0) {
        static_assert(NB == 4, "the transposed twiddle layout serves radix-16 rounds");
        return m + ((base >> 8) << (7 - jj)) + (k << 4) + ((base >> 4) & 15);
    } else {
        return m + (base >> (P + jj + 1)) + k;
    }
}

// Top = the first round of a column pass: every base is below 2^P, so the
// group offset base >> (P+jj+1) is 0 and the twiddle index m + k is the same
// for every thread (a uniform address the compiler can load once per wave).
template <int NB, int P, class A, bool Top = false>
__device__ __forceinline__ void fwd_round(typename A::T (&v)[16], const u32 (&base)[16 >> NB], u32 logN,
                                          const A &a) {
#pragma unroll
    for (int jj = NB - 1; jj >= 0; --jj) {
        const u32 m = 1u << (logN - (P + jj + 1));
#pragma unroll
        for (int b = 0; b < (16 >> NB); ++b) {
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e) {
                if (e & (1 << jj)) continue;
                const u32 idx = Top ? m + ((u32)e >> (jj + 1)) : tw_index<P, NB>(m, base[b], jj, e);
                // odd stages skip the FP64 arithmetic's reduction of x; stage 0,
                // the round's last, always reduces (FpArith's bounds)
                if (jj & 1)
                    a.ct_lazy(v[b * (1 << NB) + e], v[b * (1 << NB) + e + (1 << jj)], idx);
                else
                    a.ct(v[b * (1 << NB) + e], v[b * (1 << NB) + e + (1 << jj)], idx);
            }
        }
    }
}

// The 15 twiddles of a radix-16 forward round (NB = 4, one base per thread) in the
// order fwd_round_w takes them: stage jj's 8 >> jj entries k = 0 .. at slot
// (8 >> jj) - 1 + k. A pass loads them while its values wait in LDS across a
// barrier, so the round after the barrier waits on no global load.
template <int P, class A>
__device__ __forceinline__ void load_round_tw(const A &a, u32 base, u32 logN, typename A::W (&w)[15]) {
#pragma unroll
    for (int jj = 3; jj >= 0; --jj) {
        const u32 m = 1u << (logN - (P + jj + 1));
#pragma unroll
        for (int k = 0; k < (8 >> jj); ++k) w[(8 >> jj) - 1 + k] = a.twiddle(tw_index<P, 4>(m, base, jj, k << (jj + 1)));
    }
}

// fwd_round<4, P> with its twiddles already in registers (load_round_tw)
template <class A>
__device__ __forceinline__ void fwd_round_w(typename A::T (&v)[16], const typename A::W *w, const A &a) {
#pragma unroll
    for (int jj = 3; jj >= 0; --jj) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            if (e & (1 << jj)) continue;
            const auto &tw = w[(8 >> jj) - 1 + (e >> (jj + 1))];
            if (jj & 1)
                a.ct_lazy_w(v[e], v[e + (1 << jj)], tw);
            else
                a.ct_w(v[e], v[e + (1 << jj)], tw);
        }
    }
}

// inv_round<4, P> with its twiddles already in registers (load_round_tw over the
// inverse table: the same entries and slots)
template <class A>
__device__ __forceinline__ void inv_round_w(typename A::T (&v)[16], const typename A::W *w, const A &a) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            if (e & (1 << jj)) continue;
            a.gs_w(v[e], v[e + (1 << jj)], w[(8 >> jj) - 1 + (e >> (jj + 1))]);
        }
    }
}

template <int NB, int P, class A>
__device__ __forceinline__ void inv_round(typename A::T (&v)[16], const u32 (&base)[16 >> NB], u32 logN,
                                          const A &a) {
#pragma unroll
    for (int jj = 0; jj < NB; ++jj) {
        const u32 m = 1u << (logN - (P + jj + 1));
#pragma unroll
        for (int b = 0; b < (16 >> NB); ++b) {
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e) {
                if (e & (1 << jj)) continue;
                a.gs(v[b * (1 << NB) + e], v[b * (1 << NB) + e + (1 << jj)], tw_index<P, NB>(m, base[b], jj, e));
            }
        }
    }
}

// Last inverse round: as inv_round, but the top stage (distance N/2, m = 1)
// multiplies its outputs by s0 (x output: n^-1 * scale) and s1 (y output:
// n^-1 * psi^-brv(1) * scale), leaving them ready for canonical form.
// (the last round of a column pass: every base is below 2^p, so the group
// offset base >> sh is 0 and the twiddle indices are uniform)
template <int NB, class A>
__device__ __forceinline__ void inv_round_last(typename A::T (&v)[16], const u32 (&base)[16 >> NB], int p,
                                               u32 logN, const A &a, u64 s0, u64 s0s, u64 s1, u64 s1s) {
#pragma unroll
    for (int jj = 0; jj < NB - 1; ++jj) {
        const int sh = p + jj + 1;
        const u32 m = 1u << (logN - sh);
#pragma unroll
        for (int b = 0; b < (16 >> NB); ++b) {
            const u32 g0 = m;
#pragma unroll
            for (int e = 0; e < (1 << NB); ++e) {
                if (e & (1 << jj)) continue;
                a.gs(v[b * (1 << NB) + e], v[b * (1 << NB) + e + (1 << jj)], g0 + (e >> (jj + 1)));
            }
        }
    }
    constexpr int jj = NB - 1;
#pragma unroll
    for (int b = 0; b < (16 >> NB); ++b) {
#pragma unroll
        for (int e = 0; e < (1 << NB); ++e) {
            if (e & (1 << jj)) continue;
            a.gs_last(v[b * (1 << NB) + e], v[b * (1 << NB) + e + (1 << jj)], s0, s0s, s1, s1s);
        }
    }
}

}  // namespace phx
