// line_fast.hip -- the fused per-direction pass specialised at compile time for
// the (VEC, M) pairs of the hot configurations.
//
// Same mathematics and layouts as line_kernel in fft.hip (one N-point DFT per
// line, N = VEC * M, loaded from the packed real layout and stored in the
// reciprocal layout of zmodel/transforms/fft.cu:162-181 +
// fft_device_plans.h:64-66), organised for gfx950:
//   * every extent is a compile-time power of two, so all index arithmetic is
//     shifts and masks and LDS addresses fold into immediate offsets;
//   * a workgroup loops over lines; while it transforms line i in LDS the
//     global loads of line i+1 are already in flight into registers;
//   * twiddles come from LDS tables, all generated in FP64 on the GPU and laid
//     out so that the lanes of a wave read consecutive entries: per-pass
//     tables for the sub-transform twiddles, and a per-thread table (or the
//     full W_N, or a two-level W_N = A[e & 63] * B[e >> 6]) for the
//     inter-stage twiddle;
//   * the last radix pass writes straight to HBM with lanes running along the
//     contiguous output index.
#include "common.hpp"
#include <type_traits>
#include "kernels.hpp"

namespace zm {

namespace fast {

template <int N>
struct Log2 {
    static constexpr int value = 1 + Log2<N / 2>::value;
};
template <>
struct Log2<1> {
    static constexpr int value = 0;
};

// radix of pass `p` when a length-2^LN transform is split greedily into
// radix-16 passes followed by one smaller pass
template <int LN, int P>
struct Radix {
    static constexpr int done = 4 * P;
    static constexpr int left = LN - done;
    static constexpr int value = left >= 4 ? 16 : (1 << (left > 0 ? left : 0));
};
template <int LN>
struct NumPasses {
    static constexpr int value = (LN + 3) / 4;
};
// SF: the smaller radix first (then radix 16), else radix 16 first; NS is the
// Stockham stride of pass P (the product of the radices before it)
template <int LN, int P, bool SF>
struct RadixS {
    static constexpr int first = (LN % 4) ? (1 << (LN % 4)) : 16;
    static constexpr int value = SF ? (P == 0 ? first : 16) : Radix<LN, P>::value;
};
template <int LN, int P, bool SF>
struct StrideS {
    static constexpr int value = P == 0 ? 1 : SF ? RadixS<LN, 0, true>::value << (4 * (P - 1)) : 1 << (4 * P);
};

}  // namespace fast

// LDS padding of the line group: one element after every 2^kPadShift, chosen
// by measurement: lines of N >= 4096 pad every 16 (FP32) or 64 (FP64)
// elements, shorter lines every 32 (lowest pass times at N = 8192 and 512).
template <typename T, int LN>
constexpr int kPadShift = LN >= 12 ? (sizeof(T) == 8 ? 6 : 4) : 5;
template <typename T, int LN>
__device__ __forceinline__ int fidx(int i) {
    return i + (i >> kPadShift<T, LN>);
}
template <typename T, int LN>
__host__ __device__ constexpr int fast_padded(int e) {
    return e + (e >> kPadShift<T, LN>) + 1;
}

// Per-thread PRE twiddle table (see fast_pass): vector-ALU path, full W_N
// table in LDS and one line per workgroup, where the per-thread table has the
// same N entries as W_N.
template <bool MF, bool FULLTAB, int LN, int LE>
constexpr bool kPerThreadPre = !MF && FULLTAB && LE == LN;

// Exponent e of the PRE twiddle W_N^e of element r of work item u in the
// first pass of a stage (the index arithmetic of fast_pass).
template <int LN_T, int R, bool CONTIG, int LVEC, int LN>
__device__ __forceinline__ int pre_exp(int u, int r) {
    constexpr int n = 1 << LN_T, nbf = n / R, LNBF = fast::Log2<nbf>::value;
    constexpr int N = 1 << LN, VEC = 1 << LVEC;
    int j, kk;
    if (CONTIG) {
        j = u & (nbf - 1);
        kk = (u >> LNBF) & ((N >> LN_T) - 1);
    } else {
        j = (u >> LVEC) & (nbf - 1);
        kk = u & (VEC - 1);
    }
    return ((j + r * nbf) * kk) & (N - 1);
}

// One Stockham radix-R pass over the line-group held in LDS.
//   CONTIG: transforms are contiguous runs of length n (the VEC-point DFTs,
//           batch index bt = (line, r)); lanes run along j.
//   !CONTIG: element i of transform (line, k1) sits at line*N + k1 + i*VEC
//           (the M-point DFTs); lanes run along k1.
// PRE: multiply element i of transform bt by W_N^(i * kk(bt)) first.
// TOGLOBAL: instead of writing LDS, hand each output to `sink`.
// FROMREG: the inputs are this thread's registers rin[k*R + r] (loaded from
// HBM straight into this pass's order) instead of LDS; the writes need no
// barrier before them, the previous group's last reads of LDS having ended
// at the caller's barrier.
// OOP: read `src`, write `dst` (a different LDS buffer) as soon as each
// transform is done, one barrier at the end; otherwise src == dst and all
// reads finish (barrier) before the writes start.
// Twiddle tables (LDS): PTW: `ta` holds this thread's PRE twiddles at
// (k*R + r)*NT + tid (built by the kernel with pre_exp); otherwise the
// whole W_N (FULLTAB) or the two-level A/B pair.  `ts` is this stage's
// per-pass table, W_{NS R}^(jm r) at r*NS + jm (jm runs along the lanes, so
// both reads are free of bank conflicts).
template <typename T, int LN_T, int NS, int R, int LVEC, int LN, int LE, int NT, bool CONTIG, bool FWD, bool PRE,
          bool TOGLOBAL, bool FULLTAB, bool OOP, bool PTW, bool FROMREG = false, typename Sink>
__device__ __forceinline__ void fast_pass(const Cx<T>* src, Cx<T>* dst, const Cx<T>* ts, const Cx<T>* ta,
                                          const Cx<T>* tb, int tid, Sink&& sink, const Cx<T>* rin = nullptr) {
    using C = Cx<T>;
    constexpr int n = 1 << LN_T;
    constexpr int nbf = n / R;
    constexpr int LNBF = fast::Log2<nbf>::value;
    constexpr int VEC = 1 << LVEC;
    constexpr int N = 1 << LN;
    constexpr int E = 1 << LE;
    constexpr int BPT = EPT / R;
    constexpr int ES = CONTIG ? 1 : VEC;
    static_assert(BPT * NT * R == E, "work split");
    static_assert(NS >= 1 && (NS & (NS - 1)) == 0 && NS <= 16, "per-pass tables cover the second pass of a stage");
    C v[BPT][R];
    int obase[BPT];
#pragma unroll
    for (int k = 0; k < BPT; k++) {
        const int u = tid + k * NT;
        int base, j, kk;
        if (CONTIG) {
            j = u & (nbf - 1);
            const int bt = u >> LNBF;  // (line, r) when n == VEC
            base = bt << LN_T;
            kk = bt & ((N >> LN_T) - 1);  // r for the VEC transforms
        } else {
            const int k1 = u & (VEC - 1);
            const int rest = u >> LVEC;
            j = rest & (nbf - 1);
            const int line = rest >> LNBF;
            base = (line << LN) + k1;
            kk = k1;
        }
        if constexpr (FROMREG) {
#pragma unroll
            for (int r = 0; r < R; r++) v[k][r] = rin[k * R + r];
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[k][r] = src[fidx<T, LN>(base + (j + r * nbf) * ES)];
        }
        if (PRE) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int e = ((j + r * nbf) * kk) & (N - 1);
                // FULLTAB: ta is the whole W_N table; otherwise W_N^e = A[e & 63] * B[e >> 6]
                if constexpr (PTW) v[k][r] = ctw<FWD>(v[k][r], ta[(k * R + r) * NT + tid]);
                else if constexpr (FULLTAB) v[k][r] = ctw<FWD>(v[k][r], ta[e]);
                else v[k][r] = ctw<FWD>(v[k][r], cmul(ta[e & 63], tb[e >> 6]));
            }
        }
        if (NS > 1) {
            const int jm = j & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; r++) v[k][r] = ctw<FWD>(v[k][r], ts[r * NS + jm]);
        }
        dft<R, FWD>(v[k]);
        const int jm = j & (NS - 1);
        const int oidx = (j - jm) * R + jm;  // output index within the transform, before + r*NS
        obase[k] = TOGLOBAL ? oidx : base + oidx * ES;
        if constexpr (OOP && !TOGLOBAL) {
#pragma unroll
            for (int r = 0; r < R; r++) dst[fidx<T, LN>(obase[k] + r * NS * ES)] = v[k][r];
        }
    }
    if constexpr (OOP && !TOGLOBAL) {
        __syncthreads();
    } else if constexpr (!TOGLOBAL) {
        if constexpr (!FROMREG) __syncthreads();
#pragma unroll
        for (int k = 0; k < BPT; k++) {
#pragma unroll
            for (int r = 0; r < R; r++) dst[fidx<T, LN>(obase[k] + r * NS * ES)] = v[k][r];
        }
        __syncthreads();
    } else {
#pragma unroll
        for (int k = 0; k < BPT; k++) {
            const int u = tid + k * NT;
#pragma unroll
            for (int r = 0; r < R; r++) sink(u, obase[k] + r * NS, v[k][r]);
        }
    }
}

// ---------------------------------------------------------------------------
// Split-precision MFMA radix-16 pass (ZM_FP32_MFMA).
//
// A radix-16 DFT of 16 transforms is one real GEMM Y = W X with K = 32
// (k = 2r + c: input r, c = re/im) and 32 output rows (2s + c'), done as two
// 16x16 output tiles of v_mfma_f32_16x16x32_f16.  Operand layouts of that
// instruction (wave64): lane l holds A[l & 15][8 (l >> 4) + t], B[8 (l >> 4) + t]
// [l & 15] and D[4 (l >> 4) + t][l & 15].  So lane l works on transform
// column l & 15 with inputs r = 4g..4g+3 and outputs s = 2g, 2g+1 (tile 0) and
// 8+2g, 9+2g (tile 1), g = l >> 4.
// Data are split x = hi + 2^-12 lo into two FP16 values (round to nearest):
// hi = fp16(x), lo = fp16(2^12 (x - hi)), and W the same way, so lo sits in
// the same FP16 range as hi.  Three products are accumulated in FP32:
// W_hi x_hi in one accumulator, W_hi x_lo + W_lo x_hi in another, combined as
// acc + 2^-12 acc' on the vector ALU; W_lo x_lo is below 2^-24 of the
// result and is dropped.  The instruction sums its products in chunks of 8
// and drops products below ~2^-24 of a chunk's largest, so the low-order
// terms get an accumulator of their own: with all four products chained
// through one accumulator the 8192^2 round trip's relative max error was
// 1.2e-5, with this split 9.1e-7 (relative L2 2.9e-7 -> 2.8e-7).  What is
// left is the ~2^-23 relative representation error of the split.
// FP16 range: the kernel scales each line group by a power of two when it
// stages it into LDS (largest |re|, |im| of the group in [2^6, 2^7),
// MFMA_SCALE_LOG2) and undoes it exactly on the final store.  A radix-R pass
// grows magnitudes by at most R, and at most 16 * 8 of growth precedes the
// last MFMA pass of a direction, so every hi operand stays below 2^14 and
// every lo operand below 2^12 * 2^(14 - 11) = 2^15 < 65504; values 2^-30
// below the group's largest still keep their absolute error under 2^-31 of it.
// ---------------------------------------------------------------------------
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

// A operands of one lane for output tile 0 (rows s = 0..7), hi and lo
// halves, kept in LDS (MFMA_W_BYTES).  Tile 1 needs no storage:
// W_16^((s+8) r) = (-1)^r W_16^(s r), and this lane's K slots 2..3 and 6..7
// hold the odd r (r = 4g + q/2), so tile 1's operand is tile 0's with the
// sign bits of its 2nd and 4th dwords flipped.
struct Mfma16W {
    half8_t hi, lo;
};
typedef unsigned int uint4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ half8_t negate_odd_r(half8_t a) {
    uint4_t u = __builtin_bit_cast(uint4_t, a);
    u[1] ^= 0x80008000u;
    u[3] ^= 0x80008000u;
    return __builtin_bit_cast(half8_t, u);
}
constexpr int MFMA_W_BYTES = 64 * (int)sizeof(Mfma16W);
constexpr int MFMA_SCALE_LOG2 = 7;
constexpr float MFMA_LO_SCALE = 4096.0f;  // lo operands hold 2^12 (x - hi)
constexpr int MFMA_MAX_BYTES = 2 * 16 * 4;  // two slots of per-wave maxima (<= 16 waves)

template <bool FWD>
__device__ __forceinline__ Mfma16W mfma16_weights(int lane) {
    Mfma16W w;
    const int i = lane & 15, g = lane >> 4;
    {
        const int srow = i >> 1, cp = i & 1;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int r = 4 * g + (q >> 1), c = q & 1;
            double sn, cs;
            sincospi((double)((srow * r) & 15) / 8.0, &sn, &cs);  // W_16^(s r) = cs - i sn (forward)
            const double wi = FWD ? -sn : sn;
            // real form of the complex product: [[re, -im], [im, re]]
            const double a = (cp == 0) ? (c == 0 ? cs : -wi) : (c == 0 ? wi : cs);
            const _Float16 h = (_Float16)a;
            w.hi[q] = h;
            w.lo[q] = (_Float16)((a - (double)h) * (double)MFMA_LO_SCALE);
        }
    }
    return w;
}

template <typename T, int LN_T, int NS, int LVEC, int LN, int LE, int NT, bool CONTIG, bool FWD, bool PRE,
          bool TOGLOBAL, bool FULLTAB, bool OOP, typename Sink>
__device__ __forceinline__ void mfma_pass16(const Cx<T>* src, Cx<T>* dst, const Cx<T>* ts, const Cx<T>* ta,
                                            const Cx<T>* tb, int tid, const Mfma16W* wtab, Sink&& sink) {
    static_assert(sizeof(T) == 4, "split-precision MFMA pass is FP32 only");
    static_assert(OOP || TOGLOBAL, "the MFMA pass writes each group as soon as it is done");
    using C = Cx<T>;
    constexpr int R = 16;
    constexpr int n = 1 << LN_T;
    constexpr int nbf = n / R;
    constexpr int LNBF = fast::Log2<nbf>::value;
    constexpr int VEC = 1 << LVEC;
    constexpr int N = 1 << LN;
    constexpr int ES = CONTIG ? 1 : VEC;
    static_assert(NT % 64 == 0 && (NT * EPT) == (1 << LE), "work split");
    // stages of at most 128 points: the radix-16 pass is always a stage's
    // first (NS = 1), so it reads no W table (`ts` is the per-pass table)
    static_assert(NS == 1, "MFMA radix-16 pass with an inter-pass twiddle");
    const int lane = tid & 63, wave = tid >> 6, col = lane & 15, g = lane >> 4;
    // Register budget (the fused kernel runs at up to 256 VGPRs): the groups
    // are not unrolled, and each group reads its A operands from LDS.
    const Mfma16W* wl = wtab + lane;
#pragma unroll 1
    for (int q = 0; q < 4; q++) {
        const int u = wave * 64 + q * 16 + col;  // transform handled by this column
        int base, j, kk;
        if (CONTIG) {
            j = u & (nbf - 1);
            const int bt = u >> LNBF;
            base = bt << LN_T;
            kk = bt & ((N >> LN_T) - 1);
        } else {
            const int k1 = u & (VEC - 1);
            const int rest = u >> LVEC;
            j = rest & (nbf - 1);
            const int line = rest >> LNBF;
            base = (line << LN) + k1;
            kk = k1;
        }
        C x[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int r = 4 * g + i;
            x[i] = src[fidx<T, LN>(base + (j + r * nbf) * ES)];
            if (PRE) {
                const int e = ((j + r * nbf) * kk) & (N - 1);
                if constexpr (FULLTAB) x[i] = ctw<FWD>(x[i], ta[e]);
                else x[i] = ctw<FWD>(x[i], cmul(ta[e & 63], tb[e >> 6]));
            }
        }
        half8_t bh, bl;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const _Float16 hr = (_Float16)x[i].x, hi_ = (_Float16)x[i].y;
            bh[2 * i] = hr;
            bh[2 * i + 1] = hi_;
            bl[2 * i] = (_Float16)((x[i].x - (float)hr) * MFMA_LO_SCALE);
            bl[2 * i + 1] = (_Float16)((x[i].y - (float)hi_) * MFMA_LO_SCALE);
        }
        const int jm = j & (NS - 1);
        const int oidx = (j - jm) * R + jm;
        const int ob = TOGLOBAL ? oidx : base + oidx * ES;
        const half8_t whi0 = wl->hi, wlo0 = wl->lo;
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const half8_t whi = t ? negate_odd_r(whi0) : whi0;
            const half8_t wlo = t ? negate_odd_r(wlo0) : wlo0;
            floatx4_t acc = {0.0f, 0.0f, 0.0f, 0.0f}, accx = {0.0f, 0.0f, 0.0f, 0.0f};
            accx = __builtin_amdgcn_mfma_f32_16x16x32_f16(wlo, bh, accx, 0, 0, 0);
            accx = __builtin_amdgcn_mfma_f32_16x16x32_f16(whi, bl, accx, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(whi, bh, acc, 0, 0, 0);
#pragma unroll
            for (int c = 0; c < 4; c++) acc[c] = fmaf(accx[c], 1.0f / MFMA_LO_SCALE, acc[c]);
            // this lane's outputs of tile t: s = 8t + 2g and 8t + 2g + 1
            const int s0 = 8 * t + 2 * g;
            const C y0{acc[0], acc[1]}, y1{acc[2], acc[3]};
            if constexpr (TOGLOBAL) {
                sink(u, ob + s0 * NS, y0);
                sink(u, ob + (s0 + 1) * NS, y1);
            } else {
                dst[fidx<T, LN>(ob + s0 * NS * ES)] = y0;
                dst[fidx<T, LN>(ob + (s0 + 1) * NS * ES)] = y1;
            }
        }
    }
    if constexpr (!TOGLOBAL) __syncthreads();
}

struct NoHook {
    template <typename X>
    __device__ void operator()(X) const {}
};
struct NoSink {
    template <typename C>
    __device__ void operator()(int, int, C) const {}
};

// Run passes PASS..NPASS-1 of a length-2^LN_T transform, all in LDS.  With
// OOP the passes alternate between the buffers `cur` and `alt`; the return
// value is the buffer that holds the result.
template <typename T, int LN_T, int PASS, int NPASS, int LVEC, int LN, int LE, int NT, bool CONTIG, bool FWD, bool PRE,
          bool FULLTAB, bool MF, bool OOP, int HOFF = 0, bool FROMREG = false, bool SF = false, typename Hook = NoHook>
__device__ __forceinline__ Cx<T>* fast_passes_lds(Cx<T>* cur, Cx<T>* alt, const Cx<T>* ts, const Cx<T>* ta,
                                                  const Cx<T>* tb, int tid, const Mfma16W* w, Hook&& hook = Hook{},
                                                  const Cx<T>* rin = nullptr) {
    if constexpr (PASS < NPASS) {
        constexpr int R = fast::RadixS<LN_T, PASS, SF>::value;
        constexpr int NS = fast::StrideS<LN_T, PASS, SF>::value;
        Cx<T>* dst = OOP ? alt : cur;
        static_assert(!(MF && FROMREG), "the MFMA passes read LDS");
        if constexpr (MF && R == 16)
            mfma_pass16<T, LN_T, NS, LVEC, LN, LE, NT, CONTIG, FWD, PRE && PASS == 0, false, FULLTAB, OOP>(
                cur, dst, ts, ta, tb, tid, w, NoSink{});
        else
            fast_pass<T, LN_T, NS, R, LVEC, LN, LE, NT, CONTIG, FWD, PRE && PASS == 0, false, FULLTAB, OOP,
                      kPerThreadPre<MF, FULLTAB, LN, LE>, FROMREG && PASS == 0>(cur, dst, ts, ta, tb, tid, NoSink{},
                                                                               rin);
        hook(std::integral_constant<int, HOFF + PASS + 1>{});
        return fast_passes_lds<T, LN_T, PASS + 1, NPASS, LVEC, LN, LE, NT, CONTIG, FWD, PRE, FULLTAB, MF, OOP, HOFF,
                               false, SF>(dst, OOP ? cur : alt, ts, ta, tb, tid, w, hook);
    } else {
        return cur;
    }
}

// last pass of a direction, straight to HBM through `sink`
template <typename T, int LN_T, int NS, int R, int LVEC, int LN, int LE, int NT, bool CONTIG, bool FWD, bool PRE,
          bool FULLTAB, bool MF, typename Sink>
__device__ __forceinline__ void last_pass(const Cx<T>* cur, const Cx<T>* ts, const Cx<T>* ta,
                                          const Cx<T>* tb, int tid, const Mfma16W* w, Sink&& sink) {
    if constexpr (MF && R == 16)
        mfma_pass16<T, LN_T, NS, LVEC, LN, LE, NT, CONTIG, FWD, PRE, true, FULLTAB, false>(cur, nullptr, ts, ta,
                                                                                           tb, tid, w, sink);
    else
        fast_pass<T, LN_T, NS, R, LVEC, LN, LE, NT, CONTIG, FWD, PRE, true, FULLTAB, false,
                  kPerThreadPre<MF, FULLTAB, LN, LE>>(cur, nullptr, ts, ta, tb, tid, sink);
}

// The periodic 5-point Laplacian at one element from its four neighbours:
// a / b the rows above / below, l / r the x neighbours (line vectors m -+ 1;
// at m = 0 / M - 1 the neighbouring lane's other end, the reference's lane
// shift at the global x edge, laplace.cu:148, 160).  FP64 sums in the
// reference's order (laplace.cu:117-167), FP32 in the first-difference order
// of the stencil kernels, both then times `scale`: the same operations as
// zm_laplace, so the values are bit for bit its own.  The multiply by
// `scale` is unconditional: an unscaled call passes 1, and x * 1 is exact
// (the kernels keep denormals), where a select per value cost 32 VALU
// instructions per thread and group.
template <typename T>
__device__ __forceinline__ Cx<T> lap5(Cx<T> a, Cx<T> b, Cx<T> c, Cx<T> l, Cx<T> r, T scale) {
#pragma clang fp contract(off)
    using C = Cx<T>;
    C s;
    if constexpr (sizeof(T) == 8) {
        s = C{(T)0, (T)0};
        s = C{s.x + a.x, s.y + a.y};
        s = C{s.x + b.x, s.y + b.y};
        s = C{s.x + l.x, s.y + l.y};
        s = C{s.x + r.x, s.y + r.y};
        s = C{s.x - c.x * (T)4, s.y - c.y * (T)4};
    } else {
        const T yx = (a.x - c.x) + (b.x - c.x), yy = (a.y - c.y) + (b.y - c.y);
        const T xx = (l.x - c.x) + (r.x - c.x), xy = (l.y - c.y) + (r.y - c.y);
        s = C{yx + xx, yy + xy};
    }
    s = C{s.x * scale, s.y * scale};
    return s;
}

// Whether the group goes from HBM straight into the first radix pass's
// registers: vector-ALU inverse passes of one-line groups, whose first
// (M-point) pass takes runs of VEC / P contiguous elements per rank across
// the lanes (and FP64 forward ones, below, SFV).  Measured against staging through LDS on the same box: 8192^2
// FP32 0.211 -> 0.207 ms, FP64 0.395 -> 0.384 ms, 4096^2 FP32 unchanged.
// Not the forward passes with their radix-16 pass first, whose (VEC-point,
// stride VEC/16) first pass would take 32-byte pieces per lane group (8192^2
// FP32 0.205 -> 0.291 ms), nor the multi-line groups of 3D (512^3 FP32
// inverse 0.416 -> 0.418 ms).
template <typename T, bool MF, int LVEC, int LM, int LLPI, bool FWD>
constexpr bool kDirectFirstPass = !MF && !FWD && LLPI == 0;

// Fused pass specialised for VEC = 2^LVEC, M = 2^LM, LPI = 2^LLPI lines per
// iteration.  a.lpb is the number of iterations (line groups) per workgroup.
// LAPX (forward x pass): transform scale * the 5-point (2D, one-line groups)
// or 7-point (3D, groups of several rows) Laplacian of the input instead of
// the input (zm_laplace_forward).  The 3D form asks for two waves per SIMD
// (at most 256 VGPRs), which its plane march fits without spills.
template <typename T, int LVEC, int LM, int LLPI, bool FWD, bool FULLTAB, bool MF, bool LAPX = false>
__global__ void __launch_bounds__((1 << (LVEC + LM + LLPI)) / EPT, (LAPX && LLPI > 0) ? 2 : 1)
    line_fast_kernel(const Cx<T>* in, Cx<T>* out, Geo g, LineArgs a, const Cx<T>* __restrict__ tables, T scale,
                     int groups) {
    using C = Cx<T>;
    constexpr int VEC = 1 << LVEC, M = 1 << LM;
    constexpr int LN = LVEC + LM, N = 1 << LN;
    constexpr int LE = LN + LLPI, E = 1 << LE;
    constexpr int NT = E / EPT;
    constexpr int LS = LVEC > LM ? LVEC : LM, S = 1 << LS;
    constexpr int NPV = fast::NumPasses<LVEC>::value, NPM = fast::NumPasses<LM>::value;
    constexpr int NTWO = S + 64 + (N / 64 > 0 ? N / 64 : 1);  // W_S + two-level W_N
    extern __shared__ unsigned char smem_raw[];
    // MF: two LDS data buffers, the radix passes ping-pong between them
    constexpr bool OOP = MF;
    // VEC = 16: the last pass would store 16 contiguous outputs per thread
    // (inverse) or VEC/P-element runs (forward), i.e. partial cache lines per
    // store instruction; instead it stays in LDS and a sweep writes whole rows
    constexpr bool STAGED = LVEC == 4;
    // DIRECT: the group is loaded from HBM straight into the order the first
    // radix pass takes it in, which then runs from registers: no staging
    // writes, reads or barriers
    // SFV (FP64 forward passes of one-line groups): the VEC stage runs its
    // smaller radix first, so that its first pass takes runs of 16
    // contiguous elements across the lanes and can load the group straight
    // into its registers (DIRECT).  8192^2 FP64 forward 0.416 -> 0.399 ms;
    // in FP32 the forward pass was unchanged and the fused Laplacian pass
    // (which stages its rows) slower (0.226 -> 0.230 ms).
    constexpr bool SFV = FWD && !MF && sizeof(T) == 8 && LLPI == 0 && (LVEC % 4) != 0;
    constexpr bool DIRECT = kDirectFirstPass<T, MF, LVEC, LM, LLPI, FWD> || SFV;
    C* lds = reinterpret_cast<C*>(smem_raw);
    C* lds2 = OOP ? lds + fast_padded<T, LN>(E) : lds;
    C* tabs = lds + fast_padded<T, LN>(E) * (OOP ? 2 : 1);
    int tid = threadIdx.x;

    // global table layout: [W_S | A (64) | B (N/64) | W_N (N)]
    // LDS table layout:    [per-pass V (VEC) | per-pass M (M) | PRE]
    // PRE: per-thread table (kPerThreadPre, E = N entries), else W_N (FULLTAB),
    // else A | B.  The per-pass tables hold W_{16 R}^(jm r) at r*16 + jm for
    // the second pass (NS = 16, radix R) of the VEC and M stages.
    constexpr bool PTW = kPerThreadPre<MF, FULLTAB, LN, LE>;
    constexpr int TV = 0, TM = VEC, TP = VEC + M;
    if constexpr (NPV > 1 && SFV) {
        constexpr int R0 = fast::RadixS<LVEC, 0, true>::value, L0 = fast::Log2<R0>::value;
        for (int t = tid; t < VEC; t += NT) tabs[TV + t] = tables[(t & (R0 - 1)) * (t >> L0) * (S / VEC)];
    } else if constexpr (NPV > 1) {
        constexpr int RV = fast::Radix<LVEC, 1>::value;
        for (int t = tid; t < 16 * RV; t += NT) tabs[TV + t] = tables[(t & 15) * (t >> 4) * (S / (16 * RV))];
    }
    if constexpr (NPM > 1) {
        constexpr int RM = fast::Radix<LM, 1>::value;
        for (int t = tid; t < 16 * RM; t += NT) tabs[TM + t] = tables[(t & 15) * (t >> 4) * (S / (16 * RM))];
    }
    if constexpr (PTW) {
        // the PRE pass is the first M pass (forward) or the first VEC pass (inverse)
        constexpr int LN_P = FWD ? LM : LVEC;
        constexpr int RP = fast::Radix<LN_P, 0>::value;
#pragma unroll 1
        for (int k = 0; k < EPT / RP; k++)
#pragma unroll 1
            for (int r = 0; r < RP; r++)
                tabs[TP + (k * RP + r) * NT + tid] =
                    tables[NTWO + pre_exp<LN_P, RP, !FWD, LVEC, LN>(tid + k * NT, r)];
    } else if constexpr (FULLTAB) {
        for (int t = tid; t < N; t += NT) tabs[TP + t] = tables[NTWO + t];
    } else {
        for (int t = tid; t < NTWO - S; t += NT) tabs[TP + t] = tables[S + t];
    }
    const C* ta = tabs + TP;
    const C* tb = tabs + TP + 64;
    // per-pass tables of each stage's second pass
    const C* tsv = tabs + TV;
    const C* tsm = tabs + TM;
    // MF: the per-lane MFMA weights and two slots of per-wave maxima follow
    // the tables in LDS
    Mfma16W* w = reinterpret_cast<Mfma16W*>(tabs + TP + (FULLTAB ? N : NTWO - S));
    float* wmax = reinterpret_cast<float*>(w + 64);
    constexpr int NW = NT / 64;
    static_assert(!MF || NW <= 16, "maxima slots");
    if constexpr (MF) {
        if (tid < 64) w[tid] = mfma16_weights<FWD>(tid);
    }

    const int L = g.l, ll = g.ll, lp = g.lp;
    const int lvp = g.lvec - g.lp;
    const int VP = 1 << lvp;
    int lsg = 0;
    while ((1 << lsg) < a.sg) lsg++;
    const int lpg = a.lines_per_group;
    const C* src0 = in + (int64_t)blockIdx.y * g.field;
    C* dst0 = out + (int64_t)blockIdx.y * g.field;

    // Line group grp = exchange group o, rows q0 .. q0 + LPI - 1 of each of its
    // P member ranks.  Member gm of group o is rank rank0(o) + gm*sg, so
    // element `pos` of member gm's row q0 + l sits at
    //   group_base(grp) + (gm << mshift) + (l << ll) + pos,
    // a uniform 64-bit base plus a 32-bit offset (fields hold < 2^31 elements).
    const int mshift = lsg + g.dims * ll;  // log2(sg * L^dims)
    auto group_base = [&](int grp) -> int64_t {
        const int64_t lam0 = (int64_t)grp << LLPI;
        const int o = (int)(lam0 / lpg);
        const int q0 = (int)(lam0 - (int64_t)o * lpg);
        const int rank0 = ((o >> lsg) << (lsg + lp)) + (o & (a.sg - 1));
        return (int64_t)rank0 * g.local + ((int64_t)q0 << ll);
    };
    auto off = [&](int gm, int l, int pos) -> int { return (gm << mshift) + (l << ll) + pos; };

    // Global accesses are non-temporal: the line group is read and written
    // once.  With plain stores the next kernel (the repack) runs up to 11%
    // slower, as if evicting the dirty lines; measured over whole steps.
    auto ldg = [](const C* p) -> C { return ld_nt(p); };
    auto stg = [](C* p, C v) { st_nt(p, v); };

    // DIRECT: input k of this thread's first radix pass (work item k / R0,
    // input k % R0) is LDS-order element slot_li(k) of the group, whose
    // global offset from the group base is li_off (the inverse of stage()'s
    // map below)
    auto slot_li = [&](int k) -> int {
        if constexpr (FWD) {
            constexpr int R0 = fast::RadixS<LVEC, 0, SFV>::value, nbf = VEC / R0, LNBF = fast::Log2<nbf>::value;
            const int u = tid + (k / R0) * NT, r = k % R0;
            return ((u >> LNBF) << LVEC) + (u & (nbf - 1)) + r * nbf;
        } else {
            constexpr int R0 = fast::Radix<LM, 0>::value, nbf = M / R0, LNBF = fast::Log2<nbf>::value;
            const int u = tid + (k / R0) * NT, r = k % R0;
            const int rest = u >> LVEC;
            return ((rest >> LNBF) << LN) + (u & (VEC - 1)) + ((rest & (nbf - 1)) + r * nbf) * VEC;
        }
    };
    auto li_off = [&](int li) -> int {
        const int l = li >> LN, x = li & (N - 1);
        if constexpr (FWD) {
            return off(x >> ll, l, x & (L - 1));
        } else {
            const int lo = x & (VEC - 1), hi = x >> LVEC;
            return off(lo >> lvp, l, (hi << lvp) + (lo & (VP - 1)));
        }
    };
    // register prefetch of one line group: element e = tid + k*NT (DIRECT:
    // slot_li(k))
    C pre[EPT];
    auto slot_off = [&](int k) -> int {
        if constexpr (DIRECT) {
            return li_off(slot_li(k));
        } else {
            const int e = tid + k * NT;
            const int l = e >> LN, rem = e & (N - 1);
            return off(rem >> ll, l, rem & (L - 1));
        }
    };
    auto fetch_part = [&](auto k0c, auto k1c, int grp) {
        constexpr int K0 = decltype(k0c)::value, K1 = decltype(k1c)::value;
        const C* gb = src0 + group_base(grp);
#pragma unroll
        for (int k = K0; k < K1; k++) pre[k] = ldg(gb + slot_off(k));
    };
    auto fetch = [&](int grp) {
        const C* gb = src0 + group_base(grp);
#pragma unroll
        for (int k = 0; k < EPT; k++) pre[k] = ldg(gb + slot_off(k));
    };
    // MF: largest |re|, |im| of the prefetched group, reduced over the wave
    // and left in maxima slot `slot` for the next stage()
    auto group_max = [&](int slot) {
        if constexpr (MF) {
            float m = 0.0f;
#pragma unroll
            for (int k = 0; k < EPT; k++) m = fmaxf(m, fmaxf(fabsf(pre[k].x), fabsf(pre[k].y)));
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) m = fmaxf(m, __shfl_xor(m, o));
            if ((tid & 63) == 0) wmax[slot * 16 + (tid >> 6)] = m;
        }
    };
    T s0 = 1, inv_s0 = 1;
    auto stage = [&]() {
#pragma unroll
        for (int k = 0; k < EPT; k++) {
            const int e = tid + k * NT;
            int li = e;
            if (!FWD) {
                const int l = e >> LN, rem = e & (N - 1);
                const int gm = rem >> ll, pos = rem & (L - 1);
                li = (l << LN) + ((pos >> lvp) << LVEC) + (gm << lvp) + (pos & (VP - 1));
            }
            if constexpr (MF) lds[fidx<T, LN>(li)] = C{pre[k].x * s0, pre[k].y * s0};
            else lds[fidx<T, LN>(li)] = pre[k];
        }
    };

    // Spread prefetch: HN hook points (after staging and after every LDS pass
    // before the last) take a slice each of the next group's loads.
    constexpr bool SPREAD = (sizeof(T) == 4 && !MF) || LLPI == 0;
    constexpr int NP1 = FWD ? NPV : NPM;
    constexpr int NP2 = STAGED ? (FWD ? NPM : NPV) : ((FWD ? NPM : NPV) - 1);
    constexpr int HN = 1 + NP1 + NP2;
    static_assert(!DIRECT || SPREAD, "the direct first pass takes the next group's loads after it");

    // The line group's transform, from the staged LDS data to HBM, with the
    // next group's loads spread over its pass boundaries by `hook`.
    auto transform_group = [&](int grp, auto& hook, auto direct) __attribute__((always_inline)) {
        C* const gd = dst0 + group_base(grp);
        constexpr bool FR = decltype(direct)::value;

        if constexpr (FWD) {
            // VEC-point DFTs (contiguous), all passes in LDS
            C* cur = fast_passes_lds<T, LVEC, 0, NPV, LVEC, LN, LE, NT, true, true, false, FULLTAB, MF, OOP, 0, FR,
                                     SFV>(lds, lds2, tsv, ta, tb, tid, w, hook, pre);
            // M-point DFTs (stride VEC) with the inter-stage twiddle on the
            // first pass; the last pass stores to HBM
            if constexpr (STAGED) {
                cur = fast_passes_lds<T, LM, 0, NPM, LVEC, LN, LE, NT, false, true, true, FULLTAB, MF, OOP, NP1>(
                    cur, cur == lds ? lds2 : lds, tsm, ta, tb, tid, w, hook);
                // element k2 of transform (line, k1) is at line*N + k1 + k2*VEC;
                // sweep each destination row (rank gq, position fr + k2*VP)
#pragma unroll
                for (int k = 0; k < EPT; k++) {
                    const int e = tid + k * NT;
                    const int l = e >> LN, rem = e & (N - 1);
                    const int gq = rem >> ll, pos = rem & (L - 1);
                    const int k1 = (gq << lvp) + (pos & (VP - 1)), k2 = pos >> lvp;
                    C v = cur[fidx<T, LN>((l << LN) + k1 + (k2 << LVEC))];
                    if constexpr (MF) v = C{v.x * inv_s0, v.y * inv_s0};
                    stg(gd + off(gq, l, pos), v);
                }
            } else {
                cur = fast_passes_lds<T, LM, 0, NPM - 1, LVEC, LN, LE, NT, false, true, true, FULLTAB, MF, OOP, NP1>(
                    cur, cur == lds ? lds2 : lds, tsm, ta, tb, tid, w, hook);
                constexpr int PL = NPM - 1;
                constexpr int RL = fast::Radix<LM, PL>::value;
                constexpr int NSL = 1 << (4 * PL);
                last_pass<T, LM, NSL, RL, LVEC, LN, LE, NT, false, true, NPM == 1, FULLTAB, MF>(
                    cur, tsm, ta, tb, tid, w, [&](int u, int k2, C v) {
                        // transform (line, k1): k1 = u & (VEC-1), line = u >> (LVEC + log2(nbf))
                        constexpr int nbf = (1 << LM) / RL;
                        const int k1 = u & (VEC - 1);
                        const int line = (u >> LVEC) >> fast::Log2<nbf>::value;
                        const int gq = k1 >> lvp, fr = k1 & (VP - 1);
                        if constexpr (MF) v = C{v.x * inv_s0, v.y * inv_s0};
                        stg(gd + off(gq, line, fr + (k2 << lvp)), v);
                    });
            }
        } else {
            // inverse: M-point DFTs first (in LDS), then VEC-point DFTs with
            // the conjugate inter-stage twiddle; the last pass stores to HBM
            C* cur = fast_passes_lds<T, LM, 0, NPM, LVEC, LN, LE, NT, false, false, false, FULLTAB, MF, OOP, 0, FR>(
                lds, lds2, tsm, ta, tb, tid, w, hook, pre);
            const T sc = MF ? scale * inv_s0 : scale;
            if constexpr (STAGED) {
                cur = fast_passes_lds<T, LVEC, 0, NPV, LVEC, LN, LE, NT, true, false, true, FULLTAB, MF, OOP, NP1>(
                    cur, cur == lds ? lds2 : lds, tsv, ta, tb, tid, w, hook);
                // element jj of transform (line, r = gm*B + j) is at
                // line*N + r*VEC + jj = line*N + gm*L + (j*VEC + jj): the
                // group's LDS order is the destination rows' order
#pragma unroll
                for (int k = 0; k < EPT; k++) {
                    const int e = tid + k * NT;
                    const int l = e >> LN, rem = e & (N - 1);
                    C v = cur[fidx<T, LN>(e)];
                    v = C{v.x * sc, v.y * sc};
                    stg(gd + off(rem >> ll, l, rem & (L - 1)), v);
                }
            } else {
                cur = fast_passes_lds<T, LVEC, 0, NPV - 1, LVEC, LN, LE, NT, true, false, true, FULLTAB, MF, OOP, NP1>(
                    cur, cur == lds ? lds2 : lds, tsv, ta, tb, tid, w, hook);
                constexpr int PL = NPV - 1;
                constexpr int RL = fast::Radix<LVEC, PL>::value;
                constexpr int NSL = 1 << (4 * PL);
                last_pass<T, LVEC, NSL, RL, LVEC, LN, LE, NT, true, false, NPV == 1, FULLTAB, MF>(
                    cur, tsv, ta, tb, tid, w, [&](int u, int jj, C v) {
                        // transform (line, r): r = bt & (M-1), line = bt >> LM
                        constexpr int nbf = (1 << LVEC) / RL;
                        const int bt = u >> fast::Log2<nbf>::value;
                        const int r = bt & (M - 1), line = bt >> LM;
                        const int gm = r >> g.lb, j = r & (g.b - 1);
                        v = C{v.x * sc, v.y * sc};
                        stg(gd + off(gm, line, (j << LVEC) + jj), v);
                    });
            }
        }
    };

    if constexpr (LAPX && LLPI == 0) {
        // zm_laplace_forward's x pass: the workgroup walks a contiguous chunk
        // of global rows and transforms each row's Laplacian.  Of every row,
        // thread tid holds the EPT elements (m, v) with v = tid mod VEC and
        // m = (tid div VEC) EPT + k, k < EPT (element m VEC + v of the line):
        // loads stay coalesced along v, and the x neighbours m -+ 1 of its
        // elements are its own but at k = 0 and k = EPT - 1.  Those two come
        // from the neighbouring thread through a small exchange area in LDS,
        // written one row ahead and double-buffered.  The rows above and
        // below are this thread's elements of the previous and the next row,
        // in registers; rows run two ahead (while row gr is transformed, row
        // gr + 2's loads are in flight into the register set row gr came
        // from).  Per row this costs the plain pass's staging writes, two
        // exchange writes and reads, and no barrier beyond the plain pass's.
        static_assert(FWD && LLPI == 0 && !MF && FULLTAB, "the fused Laplacian runs ahead of one-line forward x passes");
        static_assert(M % EPT == 0 && NT == (M / EPT) * VEC, "thread layout of the rows");
        constexpr int NG = M / EPT;  // groups of EPT consecutive m
        // exchange area after the PRE table: [buffer][end][thread]
        C* const xb = tabs + TP + N;
        C rowa[EPT], rowb[EPT], up[EPT];
        const int chunk = (groups + (int)gridDim.x - 1) / (int)gridDim.x;
        // short chunks (two border rows re-read per chunk) run in XCD-aware
        // order, so that a chunk's border rows were just read into the same
        // L2 by its neighbours: 4096^2 (8 rows per chunk) 0.089 -> 0.073 ms;
        // at 8192^2 (32 rows) that order took 0.232 -> 0.261 ms
        const int wg = chunk < 16 ? (int)xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
        const int r0 = wg * chunk;
        const int r1 = r0 + chunk < groups ? r0 + chunk : groups;
        // global row gr -> its line: exchange group o = rid, local row lr
        // (zmodel/utils.h:176-185)
        auto line_of = [&](int gr) -> int {
            gr &= g.n - 1;
            const int rid = (gr >> g.lb) & (g.p - 1);
            const int lr = ((gr >> (g.lb + lp)) << g.lb) + (gr & (g.b - 1));
            return (rid << ll) + lr;
        };
        // The transform's per-thread addresses are recomputed per row (the
        // empty asm on tid in row_step: three rows of registers leave no room
        // for them); the rows' own load, exchange and LDS addresses use this
        // copy and stay hoisted (212 VGPRs, no spills; 8192^2 pass 0.235 ->
        // 0.229 ms against recomputing them too)
        const int tidh = threadIdx.x;
        auto elem = [&](int k) -> int { return (((tidh >> LVEC) * EPT + k) << LVEC) + (tidh & (VEC - 1)); };
        auto fetch_row = [&](C* dst, int gr, auto k0c, auto k1c) __attribute__((always_inline)) {
            constexpr int K0 = decltype(k0c)::value, K1 = decltype(k1c)::value;
            const C* gb = src0 + group_base(line_of(gr));
#pragma unroll
            for (int k = K0; k < K1; k++) {
                const int e = elem(k);
                dst[k] = ldg(gb + off(e >> ll, 0, e & (L - 1)));
            }
        };
        auto put_ends = [&](const C* x, int buf) {
            xb[(buf * 2 + 0) * NT + tidh] = x[0];
            xb[(buf * 2 + 1) * NT + tidh] = x[EPT - 1];
        };
        using KB = std::integral_constant<int, 0>;
        using KE = std::integral_constant<int, EPT>;
        if (r0 < r1) {
            fetch_row(up, r0 - 1, KB{}, KE{});
            fetch_row(rowa, r0, KB{}, KE{});
            fetch_row(rowb, r0 + 1, KB{}, KE{});
            put_ends(rowa, 0);
        }
        __syncthreads();  // tables, the first row's ends
        // x neighbours across the threads' ends: m = 0 / M - 1 wrap to the
        // neighbouring lane's other end, the reference's lane shift at the
        // global x edge (laplace.cu:148, 160)
        const int v = tidh & (VEC - 1), gq = tidh >> LVEC;
        const int lt = gq > 0 ? tidh - VEC : (NG - 1) * VEC + ((v - 1) & (VEC - 1));
        const int rt = gq < NG - 1 ? tidh + VEC : ((v + 1) & (VEC - 1));
        auto row_step = [&](int gr, C* x, C* y, auto bufc) __attribute__((always_inline)) {
            constexpr int BUF = decltype(bufc)::value;
            asm volatile("" : "+v"(tid));  // the transform's addresses, per row
            const C lft = xb[(BUF * 2 + 1) * NT + lt], rgt = xb[(BUF * 2 + 0) * NT + rt];
#pragma unroll
            for (int k = 0; k < EPT; k++) {
                const C l = k > 0 ? x[k - 1] : lft, r = k < EPT - 1 ? x[k + 1] : rgt;
                lds[fidx<T, LN>(elem(k))] = lap5<T>(up[k], y[k], x[k], l, r, scale);
            }
#pragma unroll
            for (int k = 0; k < EPT; k++) up[k] = x[k];
            __syncthreads();
            const bool more = gr + 2 <= r1;  // row r1 is the chunk's last "below"
            auto hook = [&](auto hc) {
                constexpr int h = decltype(hc)::value;
                if constexpr (h == 0) put_ends(y, BUF ^ 1);  // the next row's ends
                if constexpr (h < HN) {
                    constexpr int K0 = h * EPT / HN, K1 = (h + 1) * EPT / HN;
                    if constexpr (K1 > K0) {
                        if (more) fetch_row(x, gr + 2, std::integral_constant<int, K0>{}, std::integral_constant<int, K1>{});
                    }
                }
            };
            hook(std::integral_constant<int, 0>{});
            transform_group(line_of(gr), hook, std::false_type{});
            __syncthreads();  // the next row's writes overwrite LDS
        };
        for (int gr = r0; gr < r1; gr += 2) {
            row_step(gr, rowa, rowb, std::integral_constant<int, 0>{});
            if (gr + 1 < r1) row_step(gr + 1, rowb, rowa, std::integral_constant<int, 1>{});
        }
        return;
    }

    if constexpr (LAPX) {
        // zm_laplace_forward's x pass in 3D: a group is LPI consecutive global
        // rows of one global plane, so the workgroup owns one tile of LPI rows
        // (the x lines of those rows) and marches along global planes z0 ..
        // z1 - 1, transforming each plane's tile of the 7-point Laplacian.
        // Of every tile, thread tid holds rows l < LPI at line elements
        // e = tid + NT k2, k2 < KP: the rows above and below (y -+ 1) are its
        // own registers, the two rows bordering the tile come as halo
        // elements, and the planes before and after (z -+ 1) are registers
        // too (the previous and the prefetched next plane).  The x neighbours
        // come from the raw tile staged in a second LDS buffer, written at the
        // previous plane's first hook point, so they cost no barrier.  The
        // arithmetic is the FP32 stencil's: bit for bit zm_laplace.
        static_assert(FWD && !MF && FULLTAB, "the fused Laplacian runs ahead of forward x passes");
        constexpr int LPI = 1 << LLPI, KP = EPT / LPI;
        static_assert(KP * NT == N && KP >= 1, "tile layout");
        C* const xr = tabs + TP + N;  // the raw tile of the current plane
        const int lpg = a.lines_per_group;
        const int ntile = g.n >> LLPI;
        const int wg = (int)xcd_block(blockIdx.x, gridDim.x);
        const int tile = wg % ntile, zc = wg / ntile;
        const int tz = g.n / ((int)gridDim.x / ntile);
        const int z0 = zc * tz, z1 = z0 + tz;
        const int gy0 = tile << LLPI;
        // global index -> (rank coordinate, local index) (zmodel/utils.h:176-185)
        auto split = [&](int gi, int& r, int& li) {
            gi &= g.n - 1;
            r = (gi >> g.lb) & (g.p - 1);
            li = ((gi >> (g.lb + lp)) << g.lb) + (gi & (g.b - 1));
        };
        int rid0, lr0, ridu, lru, ridd, lrd;
        split(gy0, rid0, lr0);
        split(gy0 - 1, ridu, lru);
        split(gy0 + LPI, ridd, lrd);
        // line group of the tile at plane gz, and the row base of a halo row
        auto group_of = [&](int gz) -> int {
            int did, ldp;
            split(gz, did, ldp);
            return (int)((((int64_t)((did << lp) + rid0)) * lpg + ldp * L + lr0) >> LLPI);
        };
        auto row_base = [&](int gz, int rid, int lr) -> int64_t {
            int did, ldp;
            split(gz, did, ldp);
            return (int64_t)(((did << lp) + rid) << lp) * g.local + ((int64_t)(ldp * L + lr) << ll);
        };
        // the per-thread addresses are recomputed per plane (the empty asm on
        // tid in plane_step), but for the x neighbours' line indices, kept
        // from the start (elemh): 210 VGPRs, two waves per SIMD (with every
        // address hoisted 260 VGPRs, one wave; with none 178 VGPRs and the
        // 512^3 pass 0.570 against 0.497 ms)
        const int tidh = threadIdx.x;
        auto elem = [&](int k2) -> int { return tid + NT * k2; };
        auto elemh = [&](int k2) -> int { return tidh + NT * k2; };
        // The tile's first and last rows are the neighbouring tiles' halo
        // rows: plain loads keep them in L2 for those tiles (which run on the
        // same XCD), where non-temporal ones had them re-read from HBM.
        auto fetch_plane = [&](C* dst, int gz, auto k0c, auto k1c) __attribute__((always_inline)) {
            constexpr int K0 = decltype(k0c)::value, K1 = decltype(k1c)::value;
            const C* gb = src0 + group_base(group_of(gz));
#pragma unroll
            for (int k = K0; k < K1; k++) {
                const int l = k / KP, e = elem(k % KP);
                const C* p = gb + off(e >> ll, l, e & (L - 1));
                dst[k] = (l == 0 || l == LPI - 1) ? *p : ldg(p);
            }
        };
        C up[EPT], rowa[EPT], rowb[EPT], hal[2][2][KP];
        auto fetch_halo = [&](C (*h)[KP], int gz) {
            const C* bu = src0 + row_base(gz, ridu, lru);
            const C* bd = src0 + row_base(gz, ridd, lrd);
#pragma unroll
            for (int k2 = 0; k2 < KP; k2++) {
                const int e = elem(k2);
                h[0][k2] = ldg(bu + off(e >> ll, 0, e & (L - 1)));
                h[1][k2] = ldg(bd + off(e >> ll, 0, e & (L - 1)));
            }
        };
        auto stage_raw = [&](const C* x) {
#pragma unroll
            for (int k = 0; k < EPT; k++) xr[fidx<T, LN>(((k / KP) << LN) + elem(k % KP))] = x[k];
        };
        // x neighbours of line element e (packed column pos = j VEC + jj of
        // member rank gm): j -+ 1, across the rank boundary at j = 0 / B - 1,
        // and at the global x edge the other end's lane jj -+ 1 (the
        // reference's lane shift, laplace3d.cu:90-225)
        auto x_lo = [&](int e) -> int {
            const int gm = e >> ll, pos = e & (L - 1), j = pos >> LVEC, jj = pos & (VEC - 1);
            const int lastj = (g.b - 1) << LVEC;
            return j > 0 ? e - VEC
                         : gm > 0 ? ((gm - 1) << ll) + lastj + jj : ((g.p - 1) << ll) + lastj + ((jj - 1) & (VEC - 1));
        };
        auto x_hi = [&](int e) -> int {
            const int gm = e >> ll, pos = e & (L - 1), j = pos >> LVEC, jj = pos & (VEC - 1);
            return j < g.b - 1 ? e + VEC : gm < g.p - 1 ? ((gm + 1) << ll) + jj : ((jj + 1) & (VEC - 1));
        };
        int xlo[KP], xhi[KP];
#pragma unroll
        for (int k2 = 0; k2 < KP; k2++) {
            xlo[k2] = x_lo(elemh(k2));
            xhi[k2] = x_hi(elemh(k2));
        }
        using KB = std::integral_constant<int, 0>;
        using KE = std::integral_constant<int, EPT>;
        if (z0 < z1) {
            fetch_plane(up, z0 - 1, KB{}, KE{});
            fetch_plane(rowa, z0, KB{}, KE{});
            fetch_plane(rowb, z0 + 1, KB{}, KE{});
            fetch_halo(hal[0], z0);
            stage_raw(rowa);
        }
        __syncthreads();  // tables, the first plane's raw tile
        auto plane_step = [&](int gz, C* x, C* y, auto bufc) __attribute__((always_inline)) {
#pragma clang fp contract(off)
            constexpr int BUF = decltype(bufc)::value;
            asm volatile("" : "+v"(tid));  // the transform's addresses, per plane
#pragma unroll
            for (int k = 0; k < EPT; k++) {
                const int l = k / KP, k2 = k % KP;
                const C c = x[k];
                const C ya = l > 0 ? x[k - KP] : hal[BUF][0][k2];
                const C yb = l < LPI - 1 ? x[k + KP] : hal[BUF][1][k2];
                const C xa = xr[fidx<T, LN>((l << LN) + xlo[k2])];
                const C xb = xr[fidx<T, LN>((l << LN) + xhi[k2])];
                const C za = up[k], zb = y[k];
                const T sy = (ya.x - c.x) + (yb.x - c.x), ty = (ya.y - c.y) + (yb.y - c.y);
                const T sx = (xa.x - c.x) + (xb.x - c.x), tx = (xa.y - c.y) + (xb.y - c.y);
                const T sz = (za.x - c.x) + (zb.x - c.x), tzz = (za.y - c.y) + (zb.y - c.y);
                C s = C{(sy + sx) + sz, (ty + tx) + tzz};
                s = C{s.x * scale, s.y * scale};
                lds[fidx<T, LN>((l << LN) + elem(k2))] = s;
            }
#pragma unroll
            for (int k = 0; k < EPT; k++) up[k] = x[k];
            __syncthreads();  // the Laplacian staged; the raw tile read
            const bool more = gz + 2 <= z1;  // plane z1 is the chunk's last "after"
            auto hook = [&](auto hc) {
                constexpr int h = decltype(hc)::value;
                if constexpr (h == 0) {
                    stage_raw(y);  // the next plane's raw tile
                    if (gz + 1 < z1) fetch_halo(hal[BUF ^ 1], gz + 1);
                }
                if constexpr (h < HN) {
                    constexpr int K0 = h * EPT / HN, K1 = (h + 1) * EPT / HN;
                    if constexpr (K1 > K0) {
                        if (more) fetch_plane(x, gz + 2, std::integral_constant<int, K0>{}, std::integral_constant<int, K1>{});
                    }
                }
            };
            hook(std::integral_constant<int, 0>{});
            transform_group(group_of(gz), hook, std::false_type{});
            __syncthreads();  // the next plane's writes overwrite LDS
        };
        for (int gz = z0; gz < z1; gz += 2) {
            plane_step(gz, rowa, rowb, std::integral_constant<int, 0>{});
            if (gz + 1 < z1) plane_step(gz + 1, rowb, rowa, std::integral_constant<int, 1>{});
        }
        return;
    }

    // Groups run strided over the workgroups (grp = wg + k gridDim.x), so the
    // lines in flight together are consecutive; contiguous chunks per
    // workgroup were slower (8192^2 FP32 forward / inverse 0.200 / 0.205 ->
    // 0.218 / 0.243 ms).  FP32 one-line groups of N = 8192 number the
    // workgroups XCD-aware, so that the 32 workgroups of an XCD take 32
    // consecutive lines: 0.199 / 0.205 -> 0.194 / 0.200 ms; FP64 at N = 8192
    // (forward 0.397 -> 0.414 ms), FP32 at N = 4096 and the 3D groups were
    // slower that way, and the MFMA pass at N = 8192 the same.
    constexpr bool kXcdLines = sizeof(T) == 4 && !MF && LLPI == 0 && LN == 13;
    int grp = kXcdLines ? (int)xcd_block(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    int it = 0;
    if (grp < groups) {
        fetch(grp);
        if constexpr (MF) group_max(0);
    }
    __syncthreads();  // tables, maxima
    for (; grp < groups; grp += gridDim.x, ++it) {
        // The compiler hoists every per-thread global and LDS address out of
        // this loop and keeps them live.  An opaque tid per iteration makes
        // them recomputed per group instead, where that was measured faster:
        //  * FP64, N = 8192: the hoisted addresses spill 6 to 87 VGPRs at the
        //    256-VGPR limit; recomputed, 154-163 VGPRs and no spills (the
        //    pass 0.562 -> 0.467 ms);
        //  * FP32 vector ALU, several lines per group: 228 -> 102-120 VGPRs,
        //    four 256-thread workgroups per CU instead of two (3D N = 512:
        //    0.413 -> 0.401 ms).
        // Elsewhere it only costs (FP64 3D N = 512: 0.725 -> 0.835 ms).
        constexpr bool kRecompute = (sizeof(T) == 8 && LN >= 13) || (sizeof(T) == 4 && !MF && LLPI > 0);
        if constexpr (kRecompute) asm volatile("" : "+v"(tid));
        if constexpr (MF) {
            float m = 0.0f;
#pragma unroll
            for (int v = 0; v < NW; v++) m = fmaxf(m, wmax[(it & 1) * 16 + v]);
            int ex;
            (void)frexpf(m, &ex);  // m = f 2^ex, f in [0.5, 1)
            // a group of FP32 denormals or tiny normals would need a scale
            // beyond 2^127: clamp, so s0 stays finite (its largest value
            // then lands below 2^7, which costs no range)
            ex = ex < MFMA_SCALE_LOG2 - 127 ? MFMA_SCALE_LOG2 - 127 : ex;
            s0 = ldexpf(1.0f, MFMA_SCALE_LOG2 - ex);
            inv_s0 = ldexpf(1.0f, ex - MFMA_SCALE_LOG2);
        }
        if constexpr (!DIRECT) {
            stage();
            __syncthreads();
        }
        const int nxt = grp + gridDim.x;
        // The next group's loads are spread over the pass boundaries (HN
        // points: after staging and after every LDS pass before the last); a
        // burst of EPT loads per thread stalls instruction issue.  Measured:
        //  * FP32 vector ALU: always (N = 8192: 0.234 -> 0.213 ms);
        //  * one-line groups (N >= 4096) of FP64 (N = 8192: 0.466 -> 0.401
        //    ms) and MFMA (0.289 -> 0.260 ms);
        //  * not the multi-line groups of FP64 or MFMA, which were slower
        //    spread (3D N = 512: FP64 0.725 -> 0.871 ms, MFMA 0.456 -> 0.470).
        // MF's maxima of the next group are taken after the last pass, when
        // all of its loads were issued.
        auto hook = [&](auto hc) {
            constexpr int h = decltype(hc)::value;
            if constexpr (SPREAD && h < HN) {
                // DIRECT: the first pass reads pre[], so the first slice
                // waits for it to end (point 1 issues slices 0 and 1)
                constexpr int K0 = (DIRECT && h == 1 ? 0 : h) * EPT / HN, K1 = (h + 1) * EPT / HN;
                if constexpr (K1 > K0 && !(DIRECT && h == 0)) {
                    if (nxt < groups)
                        fetch_part(std::integral_constant<int, K0>{}, std::integral_constant<int, K1>{}, nxt);
                }
            }
        };
        if constexpr (SPREAD) hook(std::integral_constant<int, 0>{});
        else if (nxt < groups) fetch(nxt);

        transform_group(grp, hook, std::bool_constant<DIRECT>{});
        if constexpr (MF) {
            if (nxt < groups) group_max((it + 1) & 1);
        }
        __syncthreads();  // the next stage() overwrites LDS; maxima
    }
}

// host side ------------------------------------------------------------------

namespace {

constexpr size_t kLdsLimit = 160 * 1024;

int lines_per_iter_log2(int n) {
    int l = 0;
    while ((n << l) < 4096) l++;
    return l;
}

// LDS bytes of the specialised kernel for element size `es`
size_t fast_lds(size_t es, int vec, int m, bool fulltab, bool mf = false, bool lapx = false) {
    const int n = vec * m;
    const int e = n << lines_per_iter_log2(n);
    const int shift = n >= 4096 ? (es == 16 ? 6 : 4) : 5;
    // lapx, one-line groups (2D): the exchange area of the rows' thread
    // ends, 2 buffers x 2 ends; several lines (3D): the raw tile
    if (lapx) return fast_lds(es, vec, m, fulltab, mf) + es * (e == n ? 4 * (e / EPT) : e + (e >> shift) + 1);
    // the two per-pass tables, then W_N (or the per-thread PRE table of the
    // same size) or the two-level pair
    const int tab = vec + m + (fulltab ? n : 64 + (n / 64 > 0 ? n / 64 : 1));
    const int padded = e + (e >> shift) + 1;
    return es * (padded * (mf ? 2 : 1) + tab) + (mf ? MFMA_W_BYTES + MFMA_MAX_BYTES : 0);
}

}  // namespace

template <typename T>
bool fast_line_fulltab(int vec, int m, bool mf = false) {
    return fast_lds(sizeof(Cx<T>), vec, m, true, mf) <= kLdsLimit;
}

template <typename T, int LVEC, int LM, int LLPI, bool FWD, bool FULLTAB, bool MF, bool LAPX>
static const void* fast_fn() {
    return (const void*)&line_fast_kernel<T, LVEC, LM, LLPI, FWD, FULLTAB, MF, LAPX>;
}

template <typename T, int LVEC, int LM, int LLPI, bool FWD, bool FULLTAB, bool MF, bool LAPX>
static hipError_t launch_fast_inst(const Cx<T>* in, Cx<T>* out, const Geo& g, const LineArgs& a,
                                   const Cx<T>* tables, int batch, double scale, int grid_x, hipStream_t stream) {
    constexpr int LN = LVEC + LM, N = 1 << LN, E = N << LLPI;
    const size_t lds = fast_lds(sizeof(Cx<T>), 1 << LVEC, 1 << LM, FULLTAB, MF, LAPX);
    const void* fn = fast_fn<T, LVEC, LM, LLPI, FWD, FULLTAB, MF, LAPX>();
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const int groups = (int)(g.field / N) >> LLPI;
    const dim3 grid((unsigned)(grid_x < groups ? grid_x : groups), (unsigned)batch);
    hipLaunchKernelGGL((line_fast_kernel<T, LVEC, LM, LLPI, FWD, FULLTAB, MF, LAPX>), grid, dim3(E / EPT), lds, stream,
                       in, out, g, a, tables, (T)scale, groups);
    return hipGetLastError();
}

// Dispatch over the instantiated (VEC, M) pairs; `op` is called with the
// kernel's template arguments.
// The FP32 vector-ALU pass always has room for the full W_N table (at most
// 136.7 KB of LDS at N = 8192), so its two-level variants are not built.
template <typename T, bool MF>
constexpr bool kBuildTwoLevel = !(sizeof(T) == 4 && !MF);

// The fused Laplacian + x pass (LAPX) is built for the FP32 vector ALU: 2D
// one-line groups (N >= 4096) and 3D groups of several lines of N <= 1024;
// its register sets leave no room in FP64 or beside the MFMA operands.
template <typename T, bool MF, int V, int MM>
constexpr bool kBuildLapx = sizeof(T) == 4 && !MF && (V + MM >= 12 || V + MM <= 10);

// 2D: one-line groups; 3D: groups of LPI = 4096 / N lines that lie in one
// B-block of rows (consecutive global rows)
template <typename T, bool MF>
bool lapx_built(int vec, int m, int dims, int b) {
    if (!(sizeof(T) == 4 && !MF)) return false;
    const int n = vec * m;
    return dims == 2 ? n >= 4096 : (n <= 1024 && (4096 / n) <= b);
}

template <typename T, bool MF, typename Op>
static auto with_fast(int vec, int m, bool fwd, bool lapx, Op&& op) {
#define ZM_LF2(V, MM, F, FT, LX)                                                                          \
    if constexpr ((FT || kBuildTwoLevel<T, MF>) && (!LX || (F && kBuildLapx<T, MF, V, MM>))) {        \
        if (vec == (1 << V) && m == (1 << MM) && fwd == F && lapx == LX &&                                \
            fast_line_fulltab<T>(vec, m, MF) == FT)                                                       \
            return op.template run<V, MM, ((V + MM) >= 12 ? 0 : 12 - (V + MM)), F, FT, MF, LX>();         \
    }
#define ZM_LF1(V, MM, F, FT) ZM_LF2(V, MM, F, FT, false) ZM_LF2(V, MM, F, FT, true)
#define ZM_LF(V, MM) ZM_LF1(V, MM, true, true) ZM_LF1(V, MM, true, false) ZM_LF1(V, MM, false, true) ZM_LF1(V, MM, false, false)
    ZM_LF(4, 4) ZM_LF(4, 5) ZM_LF(4, 6) ZM_LF(4, 7)
    ZM_LF(5, 4) ZM_LF(5, 5) ZM_LF(5, 6) ZM_LF(5, 7)
    ZM_LF(6, 4) ZM_LF(6, 5) ZM_LF(6, 6) ZM_LF(6, 7)
    ZM_LF(7, 4) ZM_LF(7, 5) ZM_LF(7, 6)
#undef ZM_LF
#undef ZM_LF1
#undef ZM_LF2
    return op.none();
}

template <typename T>
struct LaunchOp {
    const Cx<T>* in;
    Cx<T>* out;
    const Geo& g;
    const LineArgs& a;
    const Cx<T>* tables;
    int batch;
    double scale;
    int grid_x;
    hipStream_t stream;
    template <int V, int MM, int LL, bool F, bool FT, bool MF, bool LX>
    hipError_t run() {
        return launch_fast_inst<T, V, MM, LL, F, FT, MF, LX>(in, out, g, a, tables, batch, scale, grid_x, stream);
    }
    hipError_t none() { return hipErrorInvalidValue; }
};

template <typename T>
struct OccOp {
    int vec, m;
    template <int V, int MM, int LL, bool F, bool FT, bool MF, bool LX>
    int run() {
        const void* fn = fast_fn<T, V, MM, LL, F, FT, MF, LX>();
        const size_t lds = fast_lds(sizeof(Cx<T>), vec, m, FT, MF, LX);
        const int nthreads = ((1 << (V + MM)) << LL) / EPT;
        if (lds > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, nthreads, lds) != hipSuccess) return 0;
        return blocks;
    }
    int none() { return 0; }
};

template <typename T>
bool fast_line_supported(int vec, int m) {
    const bool okv = vec == 16 || vec == 32 || vec == 64 || vec == 128;
    const bool okm = m == 16 || m == 32 || m == 64 || m == 128;
    return okv && okm && (int64_t)vec * m <= 8192 &&
           fast_lds(sizeof(Cx<T>), vec, m, false, sizeof(T) == 4) <= kLdsLimit;
}

template <typename T>
int fast_line_occupancy(int vec, int m, bool fwd, bool mf, bool lapx) {
    if (lapx && !(fwd && !mf && sizeof(T) == 4)) return 0;
    if constexpr (sizeof(T) == 4) {
        if (mf) return with_fast<T, true>(vec, m, fwd, lapx, OccOp<T>{vec, m});
    }
    if (mf) return 0;
    return with_fast<T, false>(vec, m, fwd, lapx, OccOp<T>{vec, m});
}

template <typename T>
bool fast_line_lapx(int vec, int m, bool mf, int dims, int b) {
    return fast_line_supported<T>(vec, m) &&
           (mf ? lapx_built<T, true>(vec, m, dims, b) : lapx_built<T, false>(vec, m, dims, b)) &&
           fast_line_fulltab<T>(vec, m, mf) && fast_lds(sizeof(Cx<T>), vec, m, true, mf, true) <= kLdsLimit;
}

int fast_line_lapx_lines_log2(int vec, int m) { return lines_per_iter_log2(vec * m); }

template <typename T>
hipError_t launch_line_fast(const Cx<T>* in, Cx<T>* out, const Geo& g, const LineArgs& a, const Cx<T>* tables,
                            int batch, bool fwd, double scale, int grid_x, bool mf, bool lapx, hipStream_t stream) {
    if (lapx && !(fwd && fast_line_lapx<T>(g.vec, g.m, mf, g.dims, g.b))) return hipErrorInvalidValue;
    if (lapx && g.dims == 3) {
        // the plane march: grid_x = row tiles x plane chunks, chunks dividing N
        const int ntile = g.n >> lines_per_iter_log2(g.vec * g.m);
        if (grid_x <= 0 || grid_x % ntile || g.n % (grid_x / ntile)) return hipErrorInvalidValue;
    }
    const LaunchOp<T> op{in, out, g, a, tables, batch, scale, grid_x, stream};
    if constexpr (sizeof(T) == 4) {
        if (mf) return with_fast<T, true>(g.vec, g.m, fwd, lapx, LaunchOp<T>(op));
    }
    if (mf) return hipErrorInvalidValue;  // the split-precision passes are FP32 only
    return with_fast<T, false>(g.vec, g.m, fwd, lapx, LaunchOp<T>(op));
}

// FP64 twiddle tables of a line (N = vec*m, S = max(vec, m)), rounded for
// FP32:  [W_S^t, t < S | W_N^t, t < 64 | W_N^(64 t), t < N/64 | W_N^t, t < N]
__global__ void line_tables_kernel(double2* t64, float2* t32, int n, int s) {
    const int nb = n / 64 > 0 ? n / 64 : 1;
    const int nt = s + 64 + nb + n;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double num, den;
    if (t < s) {
        num = t;
        den = s;
    } else if (t < s + 64) {
        num = t - s;
        den = n;
    } else if (t < s + 64 + nb) {
        num = 64.0 * (t - s - 64);
        den = n;
    } else {
        num = t - s - 64 - nb;
        den = n;
    }
    double sn, cs;
    sincospi(-2.0 * num / den, &sn, &cs);
    t64[t] = double2{cs, sn};
    t32[t] = float2{(float)cs, (float)sn};
}

int line_table_entries(int vec, int m) {
    const int n = vec * m, s = vec > m ? vec : m;
    return s + 64 + (n / 64 > 0 ? n / 64 : 1) + n;
}

hipError_t launch_line_tables(double2* t64, float2* t32, int vec, int m, hipStream_t stream) {
    const int nt = line_table_entries(vec, m);
    hipLaunchKernelGGL(line_tables_kernel, dim3((nt + 255) / 256), dim3(256), 0, stream, t64, t32, vec * m,
                       vec > m ? vec : m);
    return hipGetLastError();
}

template bool fast_line_supported<float>(int, int);
template bool fast_line_supported<double>(int, int);
template int fast_line_occupancy<float>(int, int, bool, bool, bool);
template int fast_line_occupancy<double>(int, int, bool, bool, bool);
template bool fast_line_lapx<float>(int, int, bool, int, int);
template bool fast_line_lapx<double>(int, int, bool, int, int);
template hipError_t launch_line_fast<float>(const float2*, float2*, const Geo&, const LineArgs&, const float2*, int,
                                            bool, double, int, bool, bool, hipStream_t);
template hipError_t launch_line_fast<double>(const double2*, double2*, const Geo&, const LineArgs&, const double2*,
                                             int, bool, double, int, bool, bool, hipStream_t);

}  // namespace zm
