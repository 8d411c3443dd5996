// bconv_ntt.hip -- FastBConv fused with the first (column) pass of the forward
// NTT of its outputs.
//
// Modup converts each digit's INTT'd limbs to the complement limbs and then
// transforms them; moddown converts the INTT'd special limbs to delta over Ql and
// transforms that (reference: DRNSTool::modup, src/rns_bconv.cu:530-628, and
// moddown_from_NTT, 776-828). Run separately, the conversion writes every output
// limb to HBM and the column pass reads it back. Here each wave computes the
// conversion of 16 rows x 16 columns for four outputs on the matrix cores (the
// exact int8 form of k_bconv_mfma) and runs the column pass's first register
// round on the results in registers, so the outputs reach HBM once, as the
// intermediate; the row pass (fwd_rows_core_rb) runs the column pass's remaining
// stages before its own.
//
// Lane (g, l) of a wave: output 4T + g of the workgroup's group T, column
// c0 + l, row "low" of the wave. MFMA step e puts row low + e R/16 of the 16
// columns in the B operand's columns, so after 16 steps the lane holds rows
// low, low + R/16, ..., low + 15 R/16 of its output's column: the 16 values of
// a column-pass thread's first register round. The four outputs of a workgroup
// share one butterfly arithmetic: the fused descriptors list the 64-bit-integer
// outputs first and the FP64 ones after (padded to a multiple of four with
// outputs that are computed but not stored); a group that holds both runs the
// integer butterflies, exact for any prime below 2^61, and stores its FP64
// outputs in the FP64 form their row pass reads.
//
// Workgroups over the same (conversion, column tile, rows) read the same inputs;
// they are numbered so that they run on one XCD next to each other, and the
// inputs come from that XCD's L2 after the first read.
#include "ntt_body.h"
#include "evaluator.h"

namespace phx {

namespace {
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr u32 kFcCols = 16;   // columns of a tile (the MFMA's 16 coefficient columns)

template <int SR>
struct FcShape {
    static constexpr int R = 1 << SR;      // rows of a limb (N = 256 R)
    static constexpr int LOWS = R / 16;    // first-round threads per column
    static constexpr int WAVES = LOWS < 4 ? LOWS : 4;   // lows (waves) per workgroup
    static constexpr int NT = WAVES * 64;
    static constexpr int QUADS = LOWS / WAVES;          // workgroups per column tile and group
};

// the column pass's first register round on one lane's 16 values (rows
// low + e R/16 of one column, residues below 2q), stored as the intermediate.
// fp_form: the integer round ran for a prime whose row pass takes the FP64
// arithmetic, so its outputs are stored canonical as doubles.
template <int SR, class Ar>
__device__ __forceinline__ void fc_round_a(const Ar &a, const u64 (&raw)[16], u64 *outp, bool store, u32 logN,
                                           u32 base, bool fp_form = false) {
    using Tv = typename Ar::T;
    constexpr int NB = SR - 4;
    Tv v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = Ar::from_canon(raw[e]);
    const u32 baseA[1] = {base};
    fwd_round<4, 8 + NB, Ar, true>(v, baseA, logN, a);
    if (store) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            u64 w = Ar::to_bits(v[e]);
            if constexpr (std::is_same<Ar, IntArith>::value)
                if (fp_form) w = FpArith::to_bits(FpArith::from_canon(a.fwd_canon(v[e])));
            st_mid(outp + base + ((u32)e << (8 + NB)), w);
        }
    }
}
}  // namespace

// Grid: x = column tiles x quads of lows x groups (XCD-aware order), y =
// conversion (descriptor, input and output limb sets, as bconv_row); groups_max =
// the largest group count. Every wave is independent (no LDS, no barrier).
template <int SR>
__global__ __launch_bounds__(FcShape<SR>::NT, 4) void k_bconv_fwd_cols(BconvArgs A, NttTables T, u32 groups_max) {
    using S = FcShape<SR>;
    BconvDesc d;
    const u64 *inb;
    u64 *outb;
    if (A.rows) {
        const BconvRow r = A.rows[blockIdx.y];
        d = *r.fdesc;
        inb = r.in;
        outb = r.out;
    } else {
        const u32 j = blockIdx.y % A.jobs_per_batch, bt = blockIdx.y / A.jobs_per_batch;
        d = A.descs[j * A.desc_stride];
        inb = A.in + bt * A.in_batch_stride + j * A.in_job_stride;
        outb = A.out + bt * A.out_batch_stride + j * A.out_job_stride;
    }
    // the groups of one (column tile, quad) sit on one XCD next to each other
    // (dispatch round-robins XCDs), so their shared inputs come from its L2
    u32 L = blockIdx.x;
    if ((gridDim.x & 7) == 0) L = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const u32 grp = L % groups_max, tq = L / groups_max;
    if (2 * grp >= d.rt) return;   // this conversion has fewer groups
    const u32 c0 = (tq / S::QUADS) * kFcCols, quad = tq % S::QUADS;

    const u32 N = T.N, ni = d.ni;
    const u32 tid = threadIdx.x, w = tid >> 6, lane = tid & 63, lg = lane >> 4, ln = lane & 15;
    const u32 low = quad * S::WAVES + w;
    const u32 o = 4 * grp + lg;
    const u32 row = (reinterpret_cast<const u32 *>(A.cbuf) + d.row_off)[o];
    const DevModulus md = T.mod[row];
    const u64 p = md.value;
    const double pinv = reinterpret_cast<const double *>(A.cbuf + d.pinv_off)[o];
    const u64 *bias = A.cbuf + d.a_off + (u64)d.rt * 256;
    const u64 bias_lo = bias[2 * o];
    const double biasd = __longlong_as_double((long long)bias[2 * o + 1]);
    // the group's A operand tiles (2 tiles x 2 halves of K)
    const uint4 *Ag = reinterpret_cast<const uint4 *>(A.cbuf + d.a_off);
    v4i a[2][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint4 t = Ag[((2 * grp + s) * 2 + h) * 64 + lane];
            a[s][h] = v4i{(int)t.x, (int)t.y, (int)t.z, (int)t.w};
        }
    // lane (lg, ln) feeds K = 64h + 16 lg + j: the bytes of inputs 8h + 2 lg + r
    // (r < 2) at column c0 + ln, row low + e R/16 in step e. An input beyond ni
    // meets zero operand digits (its products vanish whatever its bytes), so such
    // a lane reads input ni - 1 instead. Per lane the input offsets are fixed and
    // per step the base advances by R/16 rows (a uniform address, so the step
    // costs no vector instructions).
    // the first column round's twiddles of the group's four outputs (entries
    // 1 .. 15 of their table rows: with uniform indices the round reads only
    // those) are staged in LDS, and the output position read, before the loop, so
    // that the round after it waits on no global load
    // (every wave writes the same 64 entries: no branch on the lane index, whose
    // divergence would cost the loop loads below their uniform-base form)
    __shared__ __align__(16) u64 s_tw[4][32];
    const u32 *crow = reinterpret_cast<const u32 *>(A.cbuf) + d.row_off;
    const bool grp_fp = T.mod[crow[4 * grp]].pad != 0;
    {
        const u32 g = (tid >> 4) & 3, k = tid & 15;
        const u64 rg = crow[4 * grp + g];
        if (grp_fp) {
            s_tw[g][k] = (u64)__double_as_longlong(gmem(T.twd)[rg * N + k]);
        } else {
            const auto w = gmem(reinterpret_cast<const u64 *>(T.tw + rg * N + k));
            s_tw[g][2 * k] = w[0];
            s_tw[g][2 * k + 1] = w[1];
        }
    }
    const u32 opos = (reinterpret_cast<const u32 *>(A.cbuf) + d.pos_off)[o];
    __syncthreads();
    const u32 base = low * 256 + c0 + ln;
    const u64 *in0 = inb + (u64)d.in_first * N;
    u32 voff[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const u32 i = 8 * h + 2 * lg + r;
            voff[h][r] = ((i < ni ? i : ni - 1) * N + base) * 8;   // bytes, below 2^32
        }
    u64 xb[2][2][2];   // inputs of two steps in flight
    auto load_x = [&](u64(&x)[2][2], int e) {
        const auto sb = reinterpret_cast<const PHX_GLOBAL uint8_t *>(gmem(in0 + (u64)e * (S::LOWS * 256)));
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int r = 0; r < 2; ++r) x[h][r] = *reinterpret_cast<const PHX_GLOBAL u64 *>(sb + voff[h][r]);
    };
    // the inputs arrive with bit 7 of every byte flipped (StoreFlipped): their bytes
    // are the signed int8 operands s = byte - 128 as they are
    auto products = [&](const u64(&x)[2][2], v4i(&acc)[2]) {
        v4i b[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
            b[h] = v4i{(int)(u32)x[h][0], (int)(u32)(x[h][0] >> 32), (int)(u32)x[h][1], (int)(u32)(x[h][1] >> 32)};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            v4i c = {0, 0, 0, 0};
            c = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[s][0], b[0], c, 0, 0, 0);
            acc[s] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[s][1], b[1], c, 0, 0, 0);
        }
    };
    // T = sum_b 256^b Z_b + C_o in [0, 2^15 p) -> r = T - floor(T/p - 2^-30) p in
    // [0, 2p), as k_bconv_mfma's reduction without its final subtraction: the
    // butterflies take inputs below 2p (the integer ones below 6p, the FP64 ones
    // below 2^52, and 2p < 2^51 there)
    auto reduce = [&](const v4i(&acc)[2]) -> u64 {
        const int y0 = acc[0][0] + acc[0][1] * 256, y1 = acc[0][2] + acc[0][3] * 256;
        const int y2 = acc[1][0] + acc[1][1] * 256, y3 = acc[1][2] + acc[1][3] * 256;
        const u64 lo = bias_lo + (u64)((long long)y0 + (long long)y1 * (1ll << 16)) +
                       (u64)(long long)y2 * (1ull << 32) + (u64)(long long)y3 * (1ull << 48);
        double v = __builtin_fma((double)y3, 0x1p48, biasd);
        v = __builtin_fma((double)y2, 0x1p32, v);
        v = __builtin_fma((double)y1, 0x1p16, v);
        v += (double)y0;
        const u32 kq = (u32)__builtin_fmax(__builtin_floor(__builtin_fma(v, pinv, -0x1p-30)), 0.0);
        return lo - (u64)kq * p;
    };
    // step e + 1's products are issued before step e's reduction, and the inputs
    // of step e + 2 are loaded then
    u64 raw[16];
    v4i acc[2][2];
    load_x(xb[0], 0);
    load_x(xb[1], 1);
    products(xb[0], acc[0]);
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        if (e + 1 < 16) products(xb[(e + 1) & 1], acc[(e + 1) & 1]);
        if (e + 2 < 16) load_x(xb[e & 1], e + 2);
        raw[e] = reduce(acc[e & 1]);
    }

    u64 *outp = outb + (u64)(opos == 0xffffffffu ? 0 : opos) * N;
    const bool store = opos != 0xffffffffu;
    // the group's four outputs share one arithmetic (a uniform branch): the integer
    // one when the group's first output needs it (integer outputs come first)
    if (grp_fp) {
        FpArith fa = fp_arith(T.twd + (u64)row * N, md);
        fa.tw = reinterpret_cast<const double *>(&s_tw[lg][0]);
        fc_round_a<SR>(fa, raw, outp, store, T.logN, base);
    } else {
        IntArith ia = int_fwd(T, row, p);
        ia.tw = reinterpret_cast<const TwPair *>(&s_tw[lg][0]);
        fc_round_a<SR>(ia, raw, outp, store, T.logN, base, md.pad != 0);
    }
}

// true when the fused launch serves these conversions: ring degree 2^12 .. 2^16,
// every descriptor with a fused matrix-core form, the bconv_fused switch on
bool bconv_fused_ok(const Context &ctx, const std::vector<BconvDesc> &fdescs) {
    const u32 logN = ctx.logN();
    if (logN < 12 || logN > 16) return false;
    const int sw = ctx.tuning().bconv_fused >= 0 ? ctx.tuning().bconv_fused : tune_value("bconv_fused");
    if (sw == 0) return false;
    const int mf = ctx.tuning().bconv_mfma >= 0 ? ctx.tuning().bconv_mfma : tune_value("bconv_mfma");
    if (mf == 0) return false;
    for (const auto &d : fdescs)
        if (d.rt == 0) return false;
    return !fdescs.empty();
}

void bconv_fwd_cols_run(const Context &ctx, const std::vector<BconvDesc> &fdescs, const BconvArgs &args, u32 jobs,
                        hipStream_t s) {
    u32 groups = 0;
    for (const auto &d : fdescs) groups = std::max(groups, d.rt / 2);
    require(groups > 0 && jobs > 0 && jobs <= 65535, "fused FastBConv: bad launch");
    const NttTables &T = ctx.tables();
    const u32 tiles = 256 / kFcCols;
    auto launch = [&](auto sr) {
        constexpr int SR = decltype(sr)::value;
        using S = FcShape<SR>;
        hipLaunchKernelGGL(k_bconv_fwd_cols<SR>, dim3(tiles * S::QUADS * groups, jobs), dim3(S::NT), 0, s, args, T,
                           groups);
    };
    switch (ctx.logN()) {
        case 12: launch(std::integral_constant<int, 4>{}); break;
        case 13: launch(std::integral_constant<int, 5>{}); break;
        case 14: launch(std::integral_constant<int, 6>{}); break;
        case 15: launch(std::integral_constant<int, 7>{}); break;
        case 16: launch(std::integral_constant<int, 8>{}); break;
        default: fail(kInvalidArgument, "fused FastBConv: unsupported ring degree");
    }
    PHX_HIP_CHECK(hipGetLastError());
}

}  // namespace phx
