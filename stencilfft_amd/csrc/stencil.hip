// stencil.hip -- periodic stencils on the real-space packed layout of the
// whole virtual grid: 5-point / 7-point Laplacian (zmodel/stencils/laplace.cu:
// 73-169, laplace3d.cu:90-225) and the 4th-order x / y derivatives
// (zmodel/stencils/dx.cu:10-82, dy.cu:10-103).
//
// The reference packs halos (pack_laplace, laplace.cu:172-191), exchanges them
// with MPI_Sendrecv (laplace.cu:194-256) and reads them from receive buffers.
// On one GPU the neighbour rank's slab is in the same buffer, so a halo read
// is a direct load from that slab: the walks below return, for a step of d
// samples along a global axis, the rank and local index holding the
// neighbour, including the lane shift the reference applies at the global x
// boundary (laplace.cu:148, 160; dx.cu:32, 57).
//
// FP64 arithmetic follows the reference's evaluation order term by term with
// FMA contraction off, so the FP64 results are bit-identical to it.  The
// reference has no FP32 path; FP32 uses a first-difference evaluation order
// that stays accurate on smooth fields (see lap_combine).
#include "common.hpp"
#include "kernels.hpp"
#include <type_traits>

namespace zm {

template <typename E>
struct EOps;
template <>
struct EOps<float> {
    using S = float;
    static __device__ __forceinline__ float zero() { return 0.f; }
    static __device__ __forceinline__ float add(float a, float b) { return a + b; }
    static __device__ __forceinline__ float sub(float a, float b) { return a - b; }
    static __device__ __forceinline__ float mul(float a, float s) { return a * s; }
    static __device__ __forceinline__ float div(float a, float s) { return a / s; }
};
template <>
struct EOps<double> {
    using S = double;
    static __device__ __forceinline__ double zero() { return 0.0; }
    static __device__ __forceinline__ double add(double a, double b) { return a + b; }
    static __device__ __forceinline__ double sub(double a, double b) { return a - b; }
    static __device__ __forceinline__ double mul(double a, double s) { return a * s; }
    static __device__ __forceinline__ double div(double a, double s) { return a / s; }
};
template <typename C, typename T>
struct EOpsC {
    using S = T;
    static __device__ __forceinline__ C zero() { return C{(T)0, (T)0}; }
    static __device__ __forceinline__ C add(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
    static __device__ __forceinline__ C sub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
    static __device__ __forceinline__ C mul(C a, T s) { return C{a.x * s, a.y * s}; }
    static __device__ __forceinline__ C div(C a, T s) { return C{a.x / s, a.y / s}; }
};
template <>
struct EOps<float2> : EOpsC<float2, float> {};
template <>
struct EOps<double2> : EOpsC<double2, double> {};

// Neighbour along a block-cyclic axis (rows in 2D/3D, depth in 3D): local
// index li of the rank at coordinate rc, step d (|d| <= B).  Inside a block it
// stays on the rank; across a block edge it moves to rank rc -+ 1 and, at the
// global edge (rc = 0 or P-1), to the previous / next block of that rank
// (laplace.cu:103-140, dy.cu:34-97).
__device__ __forceinline__ void bc_walk(const Geo& g, int rc, int li, int d, int* rc2, int* li2) {
    const int B = g.b, P = g.p, nblk = g.l >> g.lb;
    const int bk = li >> g.lb, t = (li & (B - 1)) + d;
    if (t >= 0 && t < B) {
        *rc2 = rc;
        *li2 = (bk << g.lb) + t;
    } else if (t < 0) {
        *rc2 = (rc - 1 + P) & (P - 1);
        const int bk2 = (rc == 0) ? ((bk - 1 + nblk) & (nblk - 1)) : bk;
        *li2 = (bk2 << g.lb) + t + B;
    } else {
        *rc2 = (rc + 1) & (P - 1);
        const int bk2 = (rc == P - 1) ? ((bk + 1) & (nblk - 1)) : bk;
        *li2 = (bk2 << g.lb) + t - B;
    }
}

// Neighbour along the packed x axis: vector j, lane jj of the rank at column
// coordinate cid, step d (|d| <= B).  Across the rank edge it moves to the
// neighbour rank's vector on the other side; at the global edge the lane
// shifts by one (laplace.cu:142-165, dx.cu:29-77).
__device__ __forceinline__ void x_walk(const Geo& g, int cid, int j, int jj, int d, int* cid2, int* lc2) {
    const int B = g.b, P = g.p, VEC = g.vec;
    const int t = j + d;
    if (t >= 0 && t < B) {
        *cid2 = cid;
        *lc2 = (t << g.lvec) + jj;
    } else if (t < 0) {
        *cid2 = (cid - 1 + P) & (P - 1);
        const int jj2 = (cid == 0) ? ((jj - 1 + VEC) & (VEC - 1)) : jj;
        *lc2 = ((t + B) << g.lvec) + jj2;
    } else {
        *cid2 = (cid + 1) & (P - 1);
        const int jj2 = (cid == P - 1) ? ((jj + 1) & (VEC - 1)) : jj;
        *lc2 = ((t - B) << g.lvec) + jj2;
    }
}

// coordinates of element i of a field
struct Pt {
    int did, rid, cid;
    int ld, lr, lc;
};

__device__ __forceinline__ Pt decode(const Geo& g, int i) {
    Pt p;
    const int id = (int)(i / g.local);
    const int rem = (int)(i - (int64_t)id * g.local);
    p.lc = rem & (g.l - 1);
    p.lr = (rem >> g.ll) & (g.l - 1);
    p.ld = (g.dims == 3) ? (rem >> (2 * g.ll)) : 0;
    p.cid = id & (g.p - 1);
    p.rid = (id >> g.lp) & (g.p - 1);
    p.did = (g.dims == 3) ? (id >> (2 * g.lp)) : 0;
    return p;
}

__device__ __forceinline__ int64_t encode(const Geo& g, int did, int rid, int cid, int ld, int lr, int lc) {
    const int id = (g.dims == 3) ? ((did << (2 * g.lp)) + (rid << g.lp) + cid) : ((rid << g.lp) + cid);
    return (int64_t)id * g.local + (((int64_t)ld << g.ll) + lr) * g.l + lc;
}

// Sum of the neighbours minus 2d x centre.  REF: the reference's order
// (laplace.cu:117-167: ((((0 + above) + below) + left) + right) - 4 c), used
// for FP64 and bit-identical to it.  Otherwise (FP32) the first-difference
// form ((above - c) + (below - c)) + ((left - c) + (right - c)): for smooth
// data the neighbour differences are exact (Sterbenz), so the only rounding
// falls on the small result instead of cancelling digits of the large terms.
template <bool REF, typename O, typename E, typename S>
__device__ __forceinline__ E lap_combine(E a, E b, E l, E r, E f, E k, E c, bool d3) {
    if (REF) {
        E s = O::zero();
        s = O::add(s, a);
        s = O::add(s, b);
        s = O::add(s, l);
        s = O::add(s, r);
        if (d3) {
            s = O::add(s, f);
            s = O::add(s, k);
            return O::sub(s, O::mul(c, (S)6));
        }
        return O::sub(s, O::mul(c, (S)4));
    }
    E y = O::add(O::sub(a, c), O::sub(b, c));
    E x = O::add(O::sub(l, c), O::sub(r, c));
    E s = O::add(y, x);
    if (d3) s = O::add(s, O::add(O::sub(f, c), O::sub(k, c)));
    return s;
}

template <typename E, typename S>
__global__ void __launch_bounds__(256) laplace_kernel(const E* __restrict__ in, E* __restrict__ out, Geo g, S scale,
                                                      int apply_scale) {
#pragma clang fp contract(off)
    using O = EOps<E>;
    constexpr bool REF = sizeof(S) == 8;
    const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (i >= g.field) return;
    const E* src = in + (int64_t)blockIdx.y * g.field;
    const Pt p = decode(g, (int)i);
    int r2, l2, c2, x2;
    // above / below: block-cyclic rows (laplace.cu:100-140)
    bc_walk(g, p.rid, p.lr, -1, &r2, &l2);
    const E a = src[encode(g, p.did, r2, p.cid, p.ld, l2, p.lc)];
    bc_walk(g, p.rid, p.lr, +1, &r2, &l2);
    const E b = src[encode(g, p.did, r2, p.cid, p.ld, l2, p.lc)];
    // left / right: packed columns (laplace.cu:142-165)
    const int j = p.lc >> g.lvec, jj = p.lc & (g.vec - 1);
    x_walk(g, p.cid, j, jj, -1, &c2, &x2);
    const E l = src[encode(g, p.did, p.rid, c2, p.ld, p.lr, x2)];
    x_walk(g, p.cid, j, jj, +1, &c2, &x2);
    const E r = src[encode(g, p.did, p.rid, c2, p.ld, p.lr, x2)];
    const E c = src[i];
    E f = c, k = c;
    if (g.dims == 3) {
        // front / back: block-cyclic depth (laplace3d.cu:181-219)
        bc_walk(g, p.did, p.ld, -1, &r2, &l2);
        f = src[encode(g, r2, p.rid, p.cid, l2, p.lr, p.lc)];
        bc_walk(g, p.did, p.ld, +1, &r2, &l2);
        k = src[encode(g, r2, p.rid, p.cid, l2, p.lr, p.lc)];
    }
    E s = lap_combine<REF, O, E, S>(a, b, l, r, f, k, c, g.dims == 3);
    if (apply_scale) s = O::mul(s, scale);
    out[(int64_t)blockIdx.y * g.field + i] = s;
}

// [1, -8, 0, 8, -1] / (12 delta) in the reference's order (dx.cu:29-79):
// s = -8 f(-1) + f(-2) + 8 f(+1) - f(+2), then s / (12 delta).  FP32 uses
// 8 (f(+1) - f(-1)) - (f(+2) - f(-2)): differences of nearby values first.
template <typename O, typename E, typename S>
__device__ __forceinline__ E deriv4_combine(E fm1, E fm2, E fp1, E fp2, S denom, bool negate) {
    E s;
    if constexpr (sizeof(S) == 8) {
        s = O::zero();
        s = O::add(s, O::mul(fm1, (S)-8));
        s = O::add(s, fm2);
        s = O::add(s, O::mul(fp1, (S)8));
        s = O::sub(s, fp2);
    } else {
        s = O::sub(O::mul(O::sub(fp1, fm1), (S)8), O::sub(fp2, fm2));
    }
    s = O::div(s, denom);
    if (negate) s = O::mul(s, (S)-1);
    return s;
}

// 4th-order derivative as a march along the derivative's global axis.  AXIS
// 0 = x (packed columns), 1 = y (block-cyclic rows), 2 = z (block-cyclic
// depth, 3D only; the reference has no z derivative).  Every thread owns K
// lines of that axis (256 apart in line index) and walks `T` consecutive
// global positions of them:
//   x: one packed lane jj of one global row; position q in [0, B P) is vector
//      q mod B of column rank q / B, so the march crosses the P ranks of the
//      row, and the halos before q = 0 and after q = B P - 1 are the
//      neighbouring lanes' ends (the reference's lane shift at the global x
//      edge, dx.cu:32,57, is this wrap of the global column jj B P + q);
//   y: one packed column lc of column rank cid (and plane gd in 3D), over
//      global rows, each mapped to its row rank and local row (dy.cu:34-97);
//   z: one (global row, column rank, lc) pencil over global planes.
// Consecutive threads own consecutive packed columns (lanes for x), so every
// load and store of a wave is one contiguous run.  The lines sit in a
// RING-slot register ring that runs AHEAD positions ahead of the output: each
// value is loaded once per strip (plus 4 halo loads at the strip ends),
// (AHEAD - 4) K loads are in flight while one position is computed, and no
// value is exchanged between threads.  Block order (`order`): 0, strips of
// one line set are consecutive blocks in XCD-aware order, so a strip's halo
// loads hit the L2 that its neighbour strip's loads filled; 1, the same in
// plain order; 2, XCD-aware with consecutive line sets (the blocks in flight
// together cover adjacent columns at one strip).  NTL: non-temporal loads.
template <typename E, typename S, int AXIS, int K, int RING, int AHEAD, bool NTL>
__global__ void __launch_bounds__(256) deriv4_march_kernel(const E* __restrict__ in, E* __restrict__ out, Geo g,
                                                           int T, int lstrips, int lines, S denom, int negate,
                                                           int order) {
#pragma clang fp contract(off)
    using O = EOps<E>;
    static_assert(AHEAD < RING && AHEAD > 4, "the slot loaded must hold a dead value");
    const unsigned blk = order == 1 ? blockIdx.x : xcd_block(blockIdx.x, gridDim.x);
    const unsigned nls = gridDim.x >> lstrips;  // line sets
    const int s = order == 2 ? (int)(blk / nls) : (int)(blk & ((1u << lstrips) - 1));
    const int o0 = (int)(order == 2 ? blk % nls : blk >> lstrips) * 256 * K + (int)threadIdx.x;
    // the launch gives K > 1 only when 256 K divides `lines`
    if (o0 >= lines) return;
    const int64_t fb = (int64_t)blockIdx.y * g.field;
    const E* src = in + fb;
    E* dst = out + fb;
    const int N = g.n, L = g.l, P = g.p, B = g.b;
    const int local = (int)g.local;
    // block-cyclic global index -> (rank coordinate, local index)
    auto split = [&](int gi, int& r, int& li) {
        gi &= N - 1;
        r = (gi >> g.lb) & (P - 1);
        li = ((gi >> (g.lb + g.lp)) << g.lb) + (gi & (B - 1));
    };
    const bool d3 = g.dims == 3;
    // base: offset of each line at march position 0 of its axis
    int base[K], jj[K];
    const int p0 = s * T - 2;  // march position of ring index 0
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int o = o0 + 256 * k;
        jj[k] = 0;
        if constexpr (AXIS == 0) {
            jj[k] = o & (g.vec - 1);
            const int gr = (o >> g.lvec) & (N - 1);
            const int gd = d3 ? (o >> (g.lvec + g.ln)) : 0;
            int rid, lr, did = 0, ldp = 0;
            split(gr, rid, lr);
            if (d3) split(gd, did, ldp);
            base[k] = ((did * P + rid) * P) * local + (ldp * L + lr) * L;
        } else if constexpr (AXIS == 1) {
            const int lc = o & (L - 1);
            const int cid = (o >> g.ll) & (P - 1);
            const int gd = d3 ? (o >> (g.ll + g.lp)) : 0;
            int did = 0, ldp = 0;
            if (d3) split(gd, did, ldp);
            base[k] = (did * P * P + cid) * local + ldp * L * L + lc;
        } else {
            const int lc = o & (L - 1);
            const int cid = (o >> g.ll) & (P - 1);
            const int gr = o >> (g.ll + g.lp);
            int rid, lr;
            split(gr, rid, lr);
            base[k] = (rid * P + cid) * local + lr * L + lc;
        }
    }
    // offset of ring index i (march position p0 + i) of line k
    auto at = [&](int i, int k) -> int {
        const int q = p0 + i;
        if constexpr (AXIS == 0) {
            const int bp = B * P;
            if (q >= 0 && q < bp) return base[k] + (q >> g.lb) * local + (q & (B - 1)) * g.vec + jj[k];
            // past an end of the lane: the global column jj B P + q wraps
            // onto the neighbouring lane (periodic in x)
            const int gx = (jj[k] * bp + q) & (N - 1);
            const int q2 = gx & (bp - 1);
            return base[k] + (q2 >> g.lb) * local + (q2 & (B - 1)) * g.vec + (gx >> (g.lb + g.lp));
        } else if constexpr (AXIS == 1) {
            int rid, lr;
            split(q, rid, lr);
            return base[k] + rid * P * local + lr * L;
        } else {
            int did, ldp;
            split(q, did, ldp);
            return base[k] + did * P * P * local + ldp * L * L;
        }
    };
    auto ld = [](const E* p) -> E {
        if constexpr (NTL) return ld_nt(p);
        else return *p;
    };
    E w[RING][K];
#pragma unroll
    for (int i = 0; i < AHEAD; i++)
        if (i < T + 4)
#pragma unroll
            for (int k = 0; k < K; k++) w[i][k] = ld(src + at(i, k));
    for (int t0 = 0; t0 < T; t0 += RING) {
#pragma unroll
        for (int u = 0; u < RING; u++) {
            const int t = t0 + u;
            if (t < T) {
                if (t + AHEAD < T + 4)
#pragma unroll
                    for (int k = 0; k < K; k++) w[(u + AHEAD) & (RING - 1)][k] = ld(src + at(t + AHEAD, k));
                // output position p0 + t + 2 needs ring indices t .. t + 4
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const E r = deriv4_combine<O, E, S>(w[(u + 1) & (RING - 1)][k], w[u & (RING - 1)][k],
                                                        w[(u + 3) & (RING - 1)][k], w[(u + 4) & (RING - 1)][k],
                                                        denom, negate != 0);
                    st_nt(dst + at(t + 2, k), r);
                }
            }
        }
    }
}

// Launch shapes, from sweeps of (K, RING, AHEAD, T, load hint) at 8192^2
// P=4 B=32 and 512^3 P=2 B=16 on the MI355X (one box, all variants side by
// side; complex FP32, FP64 where it differs):
//   x: K=1, ring 16 / 12 ahead; T = 128 in 2D (FP64 32) with non-temporal
//      loads, 16 in 3D (2D 0.209 ms, 3D 0.401 ms; FP64 0.455, 0.781 ms);
//   y: 2D K=2 with non-temporal loads, ring 8 / 6 ahead and T = 256 (FP64
//      ring 16 / 12 ahead, T = 32, plain loads); 3D K=1, ring 16 / 12, T = 16
//      (FP64 32) (2D 0.192 ms, 3D 0.408 ms; FP64 0.435, 0.827 ms);
//   z: K=4, ring 8 / 6 ahead (FP64 5), T = 256, FP32 with non-temporal loads
//      (0.419 -> 0.407 ms; FP64 0.835 ms); 512 or 1024 threads with K = 2
//      or 1, K = 8, and T = 64 or 128 were no faster.
// Longer strips re-read fewer halo positions; shorter ones give more
// workgroups: neither wins everywhere.
// A K that does not divide the line count falls back to K=1.
template <typename E, typename S, int AXIS, int K, int RING, int AHEAD, bool NTL>
static hipError_t launch_march(const E* in, E* out, const Geo& g, int batch, int lines, int len, int tmax, S denom,
                               bool negate, hipStream_t stream, int order = 0) {
    if constexpr (K > 1) {
        if (lines % (256 * K))
            return launch_march<E, S, AXIS, 1, 16, 12, NTL>(in, out, g, batch, lines, len, tmax, denom, negate, stream,
                                                            order);
    }
    const int T = len < tmax ? len : tmax;
    int lstrips = 0;
    while ((T << lstrips) < len) lstrips++;
    const unsigned blocks = (unsigned)(((int64_t)lines + 256 * K - 1) / (256 * K)) << lstrips;
    hipLaunchKernelGGL((deriv4_march_kernel<E, S, AXIS, K, RING, AHEAD, NTL>), dim3(blocks, (unsigned)batch),
                       dim3(256), 0, stream, in, out, g, T, lstrips, lines, denom, negate ? 1 : 0, order);
    return hipGetLastError();
}

template <typename E, typename S>
static hipError_t launch_deriv4(const E* in, E* out, const Geo& g, int batch, int axis, S denom, bool negate,
                                hipStream_t stream) {
    const bool d3 = g.dims == 3;
    const int planes = d3 ? g.n : 1;
    constexpr bool wide = sizeof(E) >= 16;  // complex FP64
    // Block orders from a same-box comparison of the three at 8192^2 (P = 4,
    // B = 32) and 512^3 (P = 2, B = 16), complex FP32 / FP64: plain order for
    // x in 3D (0.390 -> 0.379, 0.787 -> 0.748 ms) and 2D FP64 (0.456 ->
    // 0.432 ms) and for y in 3D (0.405 -> 0.391, 0.819 -> 0.803 ms);
    // consecutive line sets for z in FP64 (0.826 -> 0.799 ms); the rest keep
    // XCD-aware strips (2D FP32 x plain: 0.200 -> 0.218 ms, 2D FP64 y plain:
    // 0.418 -> 0.499 ms)
    if (axis == 0) {
        const int lines = planes * g.n * g.vec, len = g.b * g.p;
        if (d3)
            return launch_march<E, S, 0, 1, 16, 12, false>(in, out, g, batch, lines, len, 16, denom, negate, stream, 1);
        return launch_march<E, S, 0, 1, 16, 12, true>(in, out, g, batch, lines, len, wide ? 32 : 128, denom, negate,
                                                      stream, wide ? 1 : 0);
    }
    if (axis == 1) {
        const int lines = planes * g.n;
        if (d3)
            return launch_march<E, S, 1, 1, 16, 12, false>(in, out, g, batch, lines, g.n, wide ? 32 : 16, denom, negate,
                                                           stream, 1);
        if (wide) return launch_march<E, S, 1, 2, 16, 12, false>(in, out, g, batch, lines, g.n, 32, denom, negate, stream);
        return launch_march<E, S, 1, 2, 8, 6, true>(in, out, g, batch, lines, g.n, 256, denom, negate, stream);
    }
    const int lines = g.n * g.n;
    if (wide)
        return launch_march<E, S, 2, 4, 8, 5, false>(in, out, g, batch, lines, g.n, 256, denom, negate, stream, 2);
    return launch_march<E, S, 2, 4, 8, 6, true>(in, out, g, batch, lines, g.n, 256, denom, negate, stream);
}

// 5/7-point Laplacian as a rolling strip: a workgroup owns one column rank
// cid, one plane gd (3D) and T consecutive global rows gr; every thread keeps
// the rows above, at and below its K = L/256 columns in registers, so each row
// is read once per strip.  The current row is staged in LDS with a VEC-wide
// halo on each side (the neighbour ranks' edge vectors, lane-shifted at the
// global x edge as laplace.cu:148,160 do), so left = lds[pos] and right =
// lds[pos + 2 VEC].  3D: the rows of planes gd - 1 and gd + 1 are loaded
// directly; the strips of those planes are grid.x * grid.y (a multiple of 8)
// blocks away, so they run on the same XCD and the loads hit its L2.
template <typename E, typename S, int K, bool D3, int HK>
__global__ void __launch_bounds__(256) laplace_strip_kernel(const E* __restrict__ in, E* __restrict__ out, Geo g,
                                                            int rows_per_strip, S scale, int apply_scale) {
#pragma clang fp contract(off)
    using O = EOps<E>;
    extern __shared__ unsigned char smem_raw[];
    E* row = reinterpret_cast<E*>(smem_raw);  // [VEC halo | L | VEC halo]
    const int tid = threadIdx.x;
    const int N = g.n, L = g.l, VEC = g.vec, B = g.b, P = g.p;
    const int cid = blockIdx.y;
    const int g0 = blockIdx.x * rows_per_strip;
    const int gd = D3 ? (int)(blockIdx.z & (unsigned)(N - 1)) : 0;
    const int64_t fbase = (int64_t)(D3 ? blockIdx.z >> g.ln : blockIdx.z) * g.field;
    const E* src = in + fbase;
    E* dst = out + fbase;
    const int cl = (cid - 1 + P) & (P - 1), cr = (cid + 1) & (P - 1);
    // 2D streams with non-temporal loads; 3D keeps rows cacheable for the
    // neighbouring planes' strips
    auto ld = [](const E* p) -> E {
        if constexpr (D3) return *p;
        else return ld_nt(p);
    };
    // block-cyclic global index -> (rank coordinate, local index)
    auto split = [&](int gi, int& r, int& li) {
        gi &= N - 1;
        r = (gi >> g.lb) & (P - 1);
        li = ((gi >> (g.lb + g.lp)) << g.lb) + (gi & (B - 1));
    };
    // offset of global row gr of plane z in column rank c
    auto row_off = [&](int z, int gr, int c) -> int64_t {
        int rid, lr, did = 0, ldp = 0;
        split(gr, rid, lr);
        if (D3) split(z, did, ldp);
        const int id = (((did << g.lp) + rid) << g.lp) + c;
        return (int64_t)id * g.local + ((int64_t)ldp * L + lr) * L;
    };
    // One row of lookahead: while row gr is computed, the loads of row gr + 2
    // (the next row's row below), of row gr + 1's halo vectors and, in 3D, of
    // row gr + 1's z-neighbours are in flight.  The rows live in a ring of
    // four register slots and the loop is unrolled by four (strips hold a
    // multiple of 4 rows), so no register copy forces an early wait.
    // Halo vector element jj = tid + 256 k, k < HK (VEC <= 256 HK <= L = 256 K).
    auto load_halo = [&](int gr, E* hl, E* hr) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < HK; k++) {
            const int jj = tid + 256 * k;
            if (jj < VEC) {
                const int jl = (cid == 0) ? ((jj - 1 + VEC) & (VEC - 1)) : jj;
                const int jr = (cid == P - 1) ? ((jj + 1) & (VEC - 1)) : jj;
                hl[k] = src[row_off(gd, gr, cl) + (B - 1) * VEC + jl];
                hr[k] = src[row_off(gd, gr, cr) + jr];
            }
        }
    };
    auto load_row = [&](int z, int gr, E* r) __attribute__((always_inline)) {
        const int64_t o = row_off(z, gr, cid);
#pragma unroll
        for (int k = 0; k < K; k++) r[k] = ld(src + o + tid + 256 * k);
    };
    E rr[4][K], zf[2][K], zb[2][K], hl[2][HK], hr[2][HK];
#pragma unroll
    for (int k = 0; k < K; k++) {
        rr[3][k] = O::zero();
        zf[0][k] = zf[1][k] = zb[0][k] = zb[1][k] = O::zero();
    }
#pragma unroll
    for (int k = 0; k < HK; k++) hl[0][k] = hl[1][k] = hr[0][k] = hr[1][k] = O::zero();
    load_row(gd, g0 - 1, rr[0]);
    load_row(gd, g0, rr[1]);
    load_row(gd, g0 + 1, rr[2]);
    if constexpr (D3) {
        load_row(gd - 1, g0, zf[0]);
        load_row(gd + 1, g0, zb[0]);
    }
    load_halo(g0, hl[0], hr[0]);
    // row t of the strip: up / centre / down in slots t, t+1, t+2 (mod 4), the
    // lookahead row in slot t+3; z-neighbours and halos in slot t (mod 2)
    auto step = [&](int t, auto uc, auto hc) __attribute__((always_inline)) {
        constexpr int U = decltype(uc)::value, C0 = (U + 1) & 3, D = (U + 2) & 3, X = (U + 3) & 3;
        constexpr int H = decltype(hc)::value, HN = H ^ 1;
        const int gr = g0 + t;
        if (t + 1 < rows_per_strip) {
            load_row(gd, gr + 2, rr[X]);
            if constexpr (D3) {
                load_row(gd - 1, gr + 1, zf[HN]);
                load_row(gd + 1, gr + 1, zb[HN]);
            }
            load_halo(gr + 1, hl[HN], hr[HN]);
        }
        // stage the current row and its halos
#pragma unroll
        for (int k = 0; k < K; k++) row[VEC + tid + 256 * k] = rr[C0][k];
#pragma unroll
        for (int k = 0; k < HK; k++) {
            const int jj = tid + 256 * k;
            if (jj < VEC) {
                row[jj] = hl[H][k];
                row[VEC + L + jj] = hr[H][k];
            }
        }
        __syncthreads();
        const int64_t oc = row_off(gd, gr, cid);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int pos = tid + 256 * k;
            const E c = rr[C0][k];
            E acc;
            if constexpr (D3)
                acc = lap_combine<sizeof(S) == 8, O, E, S>(rr[U][k], rr[D][k], row[pos], row[pos + 2 * VEC], zf[H][k],
                                                           zb[H][k], c, true);
            else
                acc = lap_combine<sizeof(S) == 8, O, E, S>(rr[U][k], rr[D][k], row[pos], row[pos + 2 * VEC], c, c, c,
                                                           false);
            if (apply_scale) acc = O::mul(acc, scale);
            if constexpr (D3) dst[oc + pos] = acc;
            else st_nt(dst + oc + pos, acc);
        }
        __syncthreads();
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    for (int t = 0; t < rows_per_strip; t += 4) {
        step(t, I0{}, I0{});
        step(t + 1, I1{}, I1{});
        step(t + 2, I2{}, I0{});
        step(t + 3, I3{}, I1{});
    }
}

// 3D 7-point Laplacian as a march along depth over row tiles (2.5D
// blocking): a workgroup owns one column rank cid, R consecutive global rows
// and a strip of tz consecutive global planes.  The tile of each plane (R
// rows of L packed columns) is read once: the planes before, at and after
// the current one sit in a ring of four register slots (the fourth is the
// lookahead plane in flight), so the z-neighbours cost no extra reads.  The
// current plane goes to LDS with one halo row above and below (rows r0 - 1
// and r0 + R, the only re-read data: 2 / R of a plane) and the VEC-wide halo
// vectors of the neighbouring column ranks, lane-shifted at the global x edge
// (laplace.cu:148,160), so above / below / left / right are LDS reads.  The
// tiles whose rows a tile re-reads as halos (row tiles -+ 1, column ranks
// cid -+ 1) are near it in XCD-aware block order, so those rows were just
// fetched into the same L2.  Against the row strip kernel, which re-reads the rows of
// planes gd +- 1 from L2 (FETCH 1.32x the field at 512^3), this took the
// 512^3 FP32 Laplacian from 0.60 to 0.41 ms (0.40 with 512 threads), and
// the FP64 one from 0.93 to 0.79 ms (1024^3 at L = 512: 8.05 to 6.75 ms).
// NTH threads per workgroup; KE = R L / NTH tile elements per thread; KY, HX:
// y-halo and x-halo elements per thread (guarded).
template <typename E, typename S, int R, int KE, int KY, int HX, int NTH = 256>
__global__ void __launch_bounds__(NTH) laplace_zmarch_kernel(const E* __restrict__ in, E* __restrict__ out, Geo g, int tz,
                                                   int ltiles, S scale, int apply_scale) {
#pragma clang fp contract(off)
    using O = EOps<E>;
    extern __shared__ unsigned char smem_raw[];
    E* tile = reinterpret_cast<E*>(smem_raw);
    const int tid = threadIdx.x;
    const int N = g.n, L = g.l, VEC = g.vec, B = g.b, P = g.p;
    const int W = L + 2 * VEC;
    // logical block = (depth strip, row tile, cid), cid fastest: in XCD-aware
    // order the tiles that trade x halos (cid -+ 1) and y halo rows (row
    // tiles -+ 1) run on one XCD at the same time
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int cid = (int)(blk & (unsigned)(P - 1));
    const int rt = (int)((blk >> g.lp) & ((1u << ltiles) - 1));
    const int zs = (int)(blk >> (g.lp + ltiles));
    const int r0 = rt * R, z0 = zs * tz;
    const int64_t fbase = (int64_t)blockIdx.y * g.field;
    const E* src = in + fbase;
    E* dst = out + fbase;
    const int cl = (cid - 1 + P) & (P - 1), cr = (cid + 1) & (P - 1);
    auto split = [&](int gi, int& r, int& li) {
        gi &= N - 1;
        r = (gi >> g.lb) & (P - 1);
        li = ((gi >> (g.lb + g.lp)) << g.lb) + (gi & (B - 1));
    };
    auto row_off = [&](int z, int gr, int c) -> int {
        int rid, lr, did, ldp;
        split(gr, rid, lr);
        split(z, did, ldp);
        const int id = (((did << g.lp) + rid) << g.lp) + c;
        return id * (int)g.local + (ldp * L + lr) * L;
    };
    const int nyh = 2 * L, nxh = 2 * R * VEC;
    auto load_plane = [&](int z, E* c) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < KE; k++) {
            const int e = tid + NTH * k;
            c[k] = src[row_off(z, r0 + (e >> g.ll), cid) + (e & (L - 1))];
        }
    };
    // The guarded loads go to a temporary and the slot is written on every
    // path: with the slot written under the guard, the double2 instantiations
    // kept the halo slots in scratch memory and waited on each halo load.
    auto load_halo = [&](int z, E* yh, E* xh) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < KY; k++) {
            const int e = tid + NTH * k;
            E v = O::zero();
            if (e < nyh) {
                const int side = e >> g.ll;  // 0: row r0 - 1, 1: row r0 + R
                v = src[row_off(z, side ? r0 + R : r0 - 1, cid) + (e & (L - 1))];
            }
            yh[k] = v;
        }
#pragma unroll
        for (int k = 0; k < HX; k++) {
            const int e = tid + NTH * k;
            E v = O::zero();
            if (e < nxh) {
                const int row = e >> (g.lvec + 1), side = (e >> g.lvec) & 1, jj = e & (VEC - 1);
                if (side == 0) {
                    const int jl = (cid == 0) ? ((jj - 1 + VEC) & (VEC - 1)) : jj;
                    v = src[row_off(z, r0 + row, cl) + (B - 1) * VEC + jl];
                } else {
                    const int jr = (cid == P - 1) ? ((jj + 1) & (VEC - 1)) : jj;
                    v = src[row_off(z, r0 + row, cr) + jr];
                }
            }
            xh[k] = v;
        }
    };
    E pl[4][KE], yh[2][KY > 0 ? KY : 1], xh[2][HX > 0 ? HX : 1];
    load_plane(z0 - 1, pl[0]);
    load_plane(z0, pl[1]);
    load_plane(z0 + 1, pl[2]);
    load_halo(z0, yh[0], xh[0]);
    auto step = [&](int t, auto uc, auto hc) __attribute__((always_inline)) {
        constexpr int U = decltype(uc)::value, C0 = (U + 1) & 3, D = (U + 2) & 3, X = (U + 3) & 3;
        constexpr int H = decltype(hc)::value, HN = H ^ 1;
        const int gd = z0 + t;
        if (t + 1 < tz) {
            load_plane(gd + 2, pl[X]);
            load_halo(gd + 1, yh[HN], xh[HN]);
        }
#pragma unroll
        for (int k = 0; k < KE; k++) {
            const int e = tid + NTH * k;
            tile[((e >> g.ll) + 1) * W + VEC + (e & (L - 1))] = pl[C0][k];
        }
#pragma unroll
        for (int k = 0; k < KY; k++) {
            const int e = tid + NTH * k;
            if (e < nyh) tile[(e >> g.ll) * (R + 1) * W + VEC + (e & (L - 1))] = yh[H][k];
        }
#pragma unroll
        for (int k = 0; k < HX; k++) {
            const int e = tid + NTH * k;
            if (e < nxh) {
                const int row = e >> (g.lvec + 1), side = (e >> g.lvec) & 1, jj = e & (VEC - 1);
                tile[(row + 1) * W + (side ? VEC + L : 0) + jj] = xh[H][k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < KE; k++) {
            const int e = tid + NTH * k;
            const int r = (e >> g.ll) + 1, c = e & (L - 1);
            const int pos = r * W + VEC + c;
            const E a = tile[pos - W], b = tile[pos + W], l = tile[pos - VEC], rr = tile[pos + VEC];
            E acc = lap_combine<sizeof(S) == 8, O, E, S>(a, b, l, rr, pl[U][k], pl[D][k], pl[C0][k], true);
            if (apply_scale) acc = O::mul(acc, scale);
            dst[row_off(gd, r0 + r - 1, cid) + c] = acc;
        }
        __syncthreads();
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    for (int t = 0; t < tz; t += 4) {
        step(t, I0{}, I0{});
        step(t + 1, I1{}, I1{});
        step(t + 2, I2{}, I0{});
        step(t + 3, I3{}, I1{});
    }
}

// Rows per strip: 64 (each row read once, two halo rows per strip), fewer
// for small 2D fields so that the grid still has about 512 workgroups.
static int strip_rows(const Geo& g, int batch) {
    int rows = 64 < g.n ? 64 : g.n;
    if (g.dims == 2)
        while (rows > 4 && (int64_t)(g.n / rows) * g.p * batch < 512) rows >>= 1;
    return rows;
}

template <typename E, typename S>
static hipError_t laplace_strip(const E* in, E* out, const Geo& g, int batch, double coef, hipStream_t stream) {
    const bool d3 = g.dims == 3;
    const int rows = strip_rows(g, batch);
    const dim3 grid((unsigned)(g.n / rows), (unsigned)g.p, (unsigned)(d3 ? batch * g.n : batch));
    const size_t lds = sizeof(E) * (size_t)(g.l + 2 * g.vec);
    const S sc = (S)coef;
    const int apply = coef != 1.0 ? 1 : 0;
    // halo vectors: one slot per thread when VEC <= 256, else K
    const bool one_halo = g.vec <= 256;
#define ZM_STRIP2(KK, HH)                                                                                        \
    if (d3)                                                                                                     \
        hipLaunchKernelGGL((laplace_strip_kernel<E, S, KK, true, HH>), grid, dim3(256), lds, stream, in, out, g,  \
                           rows, sc, apply);                                                                    \
    else                                                                                                        \
        hipLaunchKernelGGL((laplace_strip_kernel<E, S, KK, false, HH>), grid, dim3(256), lds, stream, in, out, g, \
                           rows, sc, apply);
#define ZM_STRIP(KK)                \
    if (one_halo) {                 \
        ZM_STRIP2(KK, 1)            \
    } else {                        \
        ZM_STRIP2(KK, KK)           \
    }
    switch (g.l) {
        case 256: ZM_STRIP(1) break;
        case 512: ZM_STRIP(2) break;
        case 1024: ZM_STRIP(4) break;
        case 2048: ZM_STRIP(8) break;
        default: return hipErrorInvalidValue;
    }
#undef ZM_STRIP2
#undef ZM_STRIP
    return hipGetLastError();
}

static bool strip_ok(const Geo& g, int batch) {
    // the strip loop is unrolled by four rows
    const bool lok = (g.l == 256 || g.l == 512 || g.l == 1024 || g.l == 2048) && strip_rows(g, batch) % 4 == 0;
    if (g.dims == 2) return lok;
    // 3D: grid.z = batch * N planes, and grid.x * grid.y a multiple of the 8 XCDs
    const int strips = g.n / strip_rows(g, batch);
    return lok && (int64_t)batch * g.n <= 65535 && (strips * g.p) % 8 == 0;
}

// Row tile height of the depth march: 8 rows, fewer for long rows so that a
// thread holds at most 8 tile elements per plane.
static int zmarch_rows(const Geo& g) { return g.l <= 256 ? 8 : 2048 / g.l; }

// The depth march covers 3D shapes with 32 <= L <= 1024 whose x halos fit
// two elements per thread (R VEC <= 256), whose planes split into strips of
// a multiple of 4, and whose LDS tile (R + 2 rows of L + 2 VEC elements of
// `es` bytes) fits the default 64 KiB.
static bool zmarch_ok(const Geo& g, size_t es) {
    if (g.dims != 3 || g.l < 32 || g.l > 1024 || g.n < 32) return false;
    const int R = zmarch_rows(g);
    return R * g.vec <= 256 && g.n % R == 0 && (size_t)(R + 2) * (size_t)(g.l + 2 * g.vec) * es <= 65536;
}

template <typename E, typename S, int R, int KE, int KY, int NTH = 256>
static hipError_t laplace_zmarch_inst(const E* in, E* out, const Geo& g, int batch, double coef, hipStream_t stream) {
    if (KE * NTH != R * g.l || KY * NTH < 2 * g.l || 2 * R * g.vec > 2 * NTH) return hipErrorInvalidValue;
    const int ntiles = g.n / R;
    int ltiles = 0;
    while ((1 << ltiles) < ntiles) ltiles++;
    // depth strips: the longest (up to 128 planes) that still gives 512
    // workgroups (two per CU); 512^3 FP32 ran 0.41 ms at 128 planes against
    // 0.45-0.47 ms at 32-64
    int tz = g.n < 128 ? g.n : 128;
    while (tz > 32 && (int64_t)ntiles * g.p * (g.n / tz) * batch < 512) tz >>= 1;
    const unsigned blocks = (unsigned)((int64_t)ntiles * g.p * (g.n / tz));
    const size_t lds = sizeof(E) * (size_t)(R + 2) * (size_t)(g.l + 2 * g.vec);
    const bool two = 2 * R * g.vec > NTH;
    if (two)
        hipLaunchKernelGGL((laplace_zmarch_kernel<E, S, R, KE, KY, 2, NTH>), dim3(blocks, (unsigned)batch), dim3(NTH),
                           lds, stream, in, out, g, tz, ltiles, (S)coef, coef != 1.0 ? 1 : 0);
    else
        hipLaunchKernelGGL((laplace_zmarch_kernel<E, S, R, KE, KY, 1, NTH>), dim3(blocks, (unsigned)batch), dim3(NTH),
                           lds, stream, in, out, g, tz, ltiles, (S)coef, coef != 1.0 ? 1 : 0);
    return hipGetLastError();
}

template <typename E, typename S>
static hipError_t laplace_zmarch(const E* in, E* out, const Geo& g, int batch, double coef, hipStream_t stream) {
    switch (g.l) {
        case 32: return laplace_zmarch_inst<E, S, 8, 1, 1>(in, out, g, batch, coef, stream);
        case 64: return laplace_zmarch_inst<E, S, 8, 2, 1>(in, out, g, batch, coef, stream);
        case 128: return laplace_zmarch_inst<E, S, 8, 4, 1>(in, out, g, batch, coef, stream);
        // 512 threads: 0.405 -> 0.399 ms (FP32) at 512^3
        case 256: return laplace_zmarch_inst<E, S, 8, 4, 1, 512>(in, out, g, batch, coef, stream);
        case 512: return laplace_zmarch_inst<E, S, 4, 8, 4>(in, out, g, batch, coef, stream);
        case 1024: return laplace_zmarch_inst<E, S, 2, 8, 8>(in, out, g, batch, coef, stream);
        default: return hipErrorInvalidValue;
    }
}

template <typename E, typename S>
static hipError_t stencil_dispatch(const void* in, void* out, const Geo& g, int batch, StencilOp op, double coef,
                                   bool negate, hipStream_t stream) {
    const unsigned threads = 256;
    const dim3 grid((unsigned)((g.field + threads - 1) / threads), (unsigned)batch);
    const E* src = static_cast<const E*>(in);
    E* dst = static_cast<E*>(out);
    // 3D: the depth march where its tile fits, else the strip kernel
    if (op == OP_LAPLACE && zmarch_ok(g, sizeof(E))) {
        return laplace_zmarch<E, S>(src, dst, g, batch, coef, stream);
    } else if (op == OP_LAPLACE && strip_ok(g, batch)) {
        return laplace_strip<E, S>(src, dst, g, batch, coef, stream);
    } else if (op == OP_LAPLACE) {
        hipLaunchKernelGGL((laplace_kernel<E, S>), grid, dim3(threads), 0, stream, src, dst, g, (S)coef,
                           coef != 1.0 ? 1 : 0);
    } else {
        const int axis = op == OP_DX4 ? 0 : op == OP_DY4 ? 1 : 2;
        return launch_deriv4<E, S>(src, dst, g, batch, axis, (S)coef, negate, stream);
    }
    return hipGetLastError();
}

template <typename T>
hipError_t launch_stencil(const void* in, void* out, const Geo& g, int batch, StencilOp op, bool complex_elems,
                          double coef, bool negate, hipStream_t stream) {
    if (complex_elems) return stencil_dispatch<Cx<T>, T>(in, out, g, batch, op, coef, negate, stream);
    return stencil_dispatch<T, T>(in, out, g, batch, op, coef, negate, stream);
}

template hipError_t launch_stencil<float>(const void*, void*, const Geo&, int, StencilOp, bool, double, bool,
                                          hipStream_t);
template hipError_t launch_stencil<double>(const void*, void*, const Geo&, int, StencilOp, bool, double, bool,
                                           hipStream_t);

}  // namespace zm
