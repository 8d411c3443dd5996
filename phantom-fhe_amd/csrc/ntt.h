// ntt.h -- NTT launch interface and the load/store policies that fuse
// neighbouring element-wise steps into the first/last NTT pass.
#pragma once

#include "ntt_kernels.h"

namespace phx {

struct LoadPlain {
    __device__ __forceinline__ u64 operator()(const u64 *d, u32 i, u32, u32, const DevModulus &) const {
        return ld_stream(d + i);
    }
};

struct StorePlain {
    __device__ __forceinline__ void operator()(u64 *d, u32 i, u64 v, u32, u32, const DevModulus &) const {
        st_stream(d + i, v);
    }
};

// The inverse transforms whose outputs only the FastBConv fused with the forward
// column pass reads (bconv_ntt.hip) store them with bit 7 of every byte flipped:
// the matrix cores take those bytes as signed int8 (byte - 128) as they are.
constexpr u64 kByteFlip = 0x8080808080808080ull;
struct StoreFlipped {
    __device__ __forceinline__ void operator()(u64 *d, u32 i, u64 v, u32, u32, const DevModulus &) const {
        st_stream(d + i, v ^ kByteFlip);
    }
};

// Moddown tail fused into the forward NTT of delta (CKKS):
//   ct[poly][limb] += (cx[poly][limb] - NTT(delta)) * P^-1   (mod q_limb)
// (reference: inplace_fnwt_radix8_phase2_fuse_moddown, src/ntt/ntt_moddown.cu:106-216,
//  followed by add_to_ct_kernel, src/rns_bconv.cu:763-769)
struct StoreModDown {
    u64 *ct;             // [poly][Ql][N] output
    const u64 *base0;    // added to poly 0 (may alias ct; null = 0)
    const u64 *base1;    // added to poly 1 (null = 0)
    const u64 *cx;       // [poly][QlP][N]
    const u64 *pinv;     // [Ql][2] (P^-1 mod q_i, Shoup)
    u64 ct_stride, cx_stride;
    u32 N;
    u32 pad;
    __device__ __forceinline__ void operator()(u64 *d, u32 i, u64 v, u32 limb, u32 poly,
                                               const DevModulus &m) const {
        apply(d, i, v, prefetch(i, limb, poly), limb, poly, m);
    }
    // the tail's operands, loaded by the row pass while its outputs wait in LDS
    struct Pre {
        u64 cx, base;
    };
    __device__ __forceinline__ Pre prefetch(u32 i, u32 limb, u32 poly) const {
        const u64 off = (u64)limb * N + i;
        const u64 *base = poly == 0 ? base0 : base1;
        return Pre{gmem(cx)[poly * cx_stride + off], base ? gmem(base)[off] : 0};
    }
    __device__ __forceinline__ void apply(u64 *, u32 i, u64 v, const Pre &pf, u32 limb, u32 poly,
                                          const DevModulus &m) const {
        const u64 q = m.value;
        const u64 off = (u64)limb * N + i;
        const u64 r = mul_shoup(sub_mod(pf.cx, v, q), pinv[2 * limb], pinv[2 * limb + 1], q);
        gmem(ct)[poly * ct_stride + off] = add_mod(pf.base, r, q);
    }
};

// Tensor product as the load of an inverse first pass over one ciphertext pair
// (the plain multiply's modup): writes d0, d1 to out, d2 to c2 and returns d2
// (tensor_prod_2x2_rns_poly, src/polymath.cu:463-498); out may be a, b may be a
struct LoadTensor {
    const u64 *a, *b;
    u64 *out, *c2;
    u64 LN;
    u32 N;
    __device__ __forceinline__ u64 operator()(const u64 *, u32 i, u32 limb, u32, const DevModulus &m) const {
        const u64 o = (u64)limb * N + i;
        const auto ga = gmem(a), gb = gmem(b);
        u64 r0, r1, r2;
        tensor_coeff(ga[o], ga[LN + o], gb[o], gb[LN + o], m, r0, r1, r2);
        gmem(out)[o] = r0;
        gmem(out)[LN + o] = r1;
        gmem(c2)[o] = r2;
        return r2;
    }
};

// Rescale (CKKS, floor): the NTT input for limb j is [c_last]_{q_j} taken from
// the coefficient-form last limb; the output is (c_j - NTT(.)) * q_last^-1.
// (reference: DRNSTool::divide_and_round_q_last_ntt, src/rns.cu:1128-1184)
struct LoadRescale {
    const u64 *last;     // [poly][N] coefficient form, canonical mod q_last
    u64 last_stride;
    __device__ __forceinline__ u64 operator()(const u64 *, u32 i, u32, u32 poly, const DevModulus &m) const {
        return barrett64(gmem(last)[poly * last_stride + i], m.value, m.ratio_hi);
    }
};
struct StoreRescale {
    const u64 *in;       // [poly][Ql][N] input ciphertext (NTT form)
    u64 *out;            // [poly][Ql-1][N]
    const u64 *qinv;     // [Ql-1][2] q_last^-1 mod q_j, Shoup
    u64 in_stride, out_stride;
    u32 N;
    __device__ __forceinline__ void operator()(u64 *d, u32 i, u64 v, u32 limb, u32 poly,
                                               const DevModulus &m) const {
        apply(d, i, v, prefetch(i, limb, poly), limb, poly, m);
    }
    // the input limb, loaded by the row pass while its outputs wait in LDS
    struct Pre {
        u64 x;
    };
    __device__ __forceinline__ Pre prefetch(u32 i, u32 limb, u32 poly) const {
        return Pre{gmem(in)[poly * in_stride + (u64)limb * N + i]};
    }
    __device__ __forceinline__ void apply(u64 *, u32 i, u64 v, const Pre &pf, u32 limb, u32 poly,
                                          const DevModulus &m) const {
        const u64 q = m.value;
        const u64 off = (u64)limb * N + i;
        gmem(out)[poly * out_stride + off] = mul_shoup(sub_mod(pf.x, v, q), qinv[2 * limb], qinv[2 * limb + 1], q);
    }
};

// ---- policies of work-list launches: bound to their entry's operands ----
// Moddown tail: out[i] = base + (cx[i] - NTT(delta)[i]) * P^-1 with aux0 = the cx
// limb, aux1 = the output limb, aux2 = the base limb (null: the output's own value
// when flags & 1, else zero), scale = (P^-1 mod q, Shoup)
struct StoreModDownJob {
    static constexpr bool kBind = true;
    struct Bound {
        const u64 *cx;
        u64 *out;
        const u64 *base;
        u64 p, ps;
        u32 add_out;
        __device__ __forceinline__ void operator()(u64 *d, u32 i, u64 v, u32 limb, u32 z,
                                                   const DevModulus &m) const {
            apply(d, i, v, prefetch(i, limb, z), limb, z, m);
        }
        // the tail's operands, loaded by the row pass while its outputs wait in LDS
        struct Pre {
            u64 cx, base;
        };
        __device__ __forceinline__ Pre prefetch(u32 i, u32, u32) const {
            return Pre{gmem(cx)[i], base ? gmem(base)[i] : (add_out ? gmem(out)[i] : 0)};
        }
        __device__ __forceinline__ void apply(u64 *, u32 i, u64 v, const Pre &pf, u32, u32,
                                              const DevModulus &m) const {
            const u64 q = m.value;
            gmem(out)[i] = add_mod(pf.base, mul_shoup(sub_mod(pf.cx, v, q), p, ps, q), q);
        }
    };
    __device__ __forceinline__ Bound bind(const NttJob &J) const {
        return Bound{J.aux0, J.aux1, J.aux2, gmem(J.scale)[0], gmem(J.scale)[1], J.flags & 1u};
    }
};
// Rescale: the NTT input is [c_last]_q from aux2 (the coefficient-form last limb of
// the polynomial); the output aux1[i] = (aux0[i] - NTT(.)[i]) * q_last^-1, scale =
// (q_last^-1 mod q, Shoup)
struct LoadRescaleJob {
    static constexpr bool kBind = true;
    struct Bound {
        const u64 *last;
        __device__ __forceinline__ u64 operator()(const u64 *, u32 i, u32, u32, const DevModulus &m) const {
            return barrett64(gmem(last)[i], m.value, m.ratio_hi);
        }
    };
    __device__ __forceinline__ Bound bind(const NttJob &J) const { return Bound{J.aux2}; }
};
struct StoreRescaleJob {
    static constexpr bool kBind = true;
    struct Bound {
        const u64 *in;
        u64 *out;
        u64 p, ps;
        __device__ __forceinline__ void operator()(u64 *d, u32 i, u64 v, u32 limb, u32 z,
                                                   const DevModulus &m) const {
            apply(d, i, v, prefetch(i, limb, z), limb, z, m);
        }
        // the input limb, loaded by the row pass while its outputs wait in LDS
        struct Pre {
            u64 x;
        };
        __device__ __forceinline__ Pre prefetch(u32 i, u32, u32) const { return Pre{gmem(in)[i]}; }
        __device__ __forceinline__ void apply(u64 *, u32 i, u64 v, const Pre &pf, u32, u32,
                                              const DevModulus &m) const {
            const u64 q = m.value;
            gmem(out)[i] = mul_shoup(sub_mod(pf.x, v, q), p, ps, q);
        }
    };
    __device__ __forceinline__ Bound bind(const NttJob &J) const {
        return Bound{J.aux0, J.aux1, gmem(J.scale)[0], gmem(J.scale)[1]};
    }
};

// Tensor product as the load of the modup INTT's first pass (tensor_prod_2x2_rns_poly,
// src/polymath.cu:463-498): for limb l of an item, aux0 = a's limb l (c1 at + Ql N),
// aux2 = b's limb l, aux1 = the output's limb l, flags = Ql. The load writes d0, d1 to
// the output and d2 (NTT form) to data - Ql N (the item's c2, which the key inner
// product reads), and returns d2 for the inverse transform. The output may be a
// (in place) and b may be a (squaring): each coefficient is read before it is written.
struct LoadTensorJob {
    static constexpr bool kBind = true;
    struct Bound {
        const u64 *a, *b;
        u64 *out, *c2;
        u64 LN;
        __device__ __forceinline__ u64 operator()(const u64 *, u32 i, u32, u32, const DevModulus &m) const {
            const auto ga = gmem(a), gb = gmem(b);
            u64 r0, r1, r2;
            tensor_coeff(ga[i], ga[LN + i], gb[i], gb[LN + i], m, r0, r1, r2);
            gmem(out)[i] = r0;
            gmem(out)[LN + i] = r1;
            gmem(c2)[i] = r2;
            return r2;
        }
    };
    u64 N;
    __device__ __forceinline__ Bound bind(const NttJob &J) const {
        const u64 LN = (u64)J.flags * N;
        return Bound{J.aux0, J.aux2, J.aux1, J.data - LN, LN};
    }
};

// work-list launches over njobs entries of a device table (two passes each)
void ntt_forward_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s);
void ntt_forward_jobs_moddown(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s);
void ntt_forward_jobs_rescale(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s);
// the row pass alone, over entries whose column pass ran fused with FastBConv
void ntt_forward_rows_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s);
void ntt_forward_rows_jobs_moddown(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s);
// inverse: entries with aux0 read their first pass from aux0 (out of place);
// flip: outputs stored byte-flipped (StoreFlipped)
void ntt_inverse_jobs(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s, bool flip = false);
// inverse over entries whose first pass loads the tensor product (LoadTensorJob)
void ntt_inverse_jobs_tensor(const NttTables &T, const NttJob *jobs, u32 njobs, hipStream_t s, bool flip = false);

void ntt_forward(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s);
void ntt_inverse(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s, bool flip = false);
void ntt_forward_store(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const StoreModDown &st,
                       hipStream_t s);
// the row pass alone, over limbs whose column pass ran fused with FastBConv
void ntt_forward_rows(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, hipStream_t s);
void ntt_forward_rows_store(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const StoreModDown &st,
                            hipStream_t s);
void ntt_forward_rescale(const NttTables &T, const LimbMap &M, u32 nlimbs, u32 npolys, const LoadRescale &ld,
                         const StoreRescale &st, hipStream_t s);
// inverse of one polynomial whose first pass loads the tensor product (always two passes)
void ntt_inverse_tensor(const NttTables &T, const LimbMap &M, u32 nlimbs, const LoadTensor &ld, hipStream_t s,
                        bool flip = false);

// contiguous limbs [0, L) of one polynomial with rows first_row + i
inline LimbMap limb_map(u64 *data, u32 first_row, u32 L, u64 poly_stride = 0) {
    LimbMap m{};
    m.data = data;
    m.src = nullptr;
    m.scale = nullptr;
    m.poly_stride = poly_stride;
    m.src_poly_stride = 0;
    m.first_row = first_row;
    m.split = L;
    m.special_row = 0;
    m.skip_lo = m.skip_hi = L;
    m.skip_alpha = 0;
    m.skip_cap = L;
    m.skip_mod = 0;
    m.nlimbs = L;
    return m;
}

}  // namespace phx
