// fft.hip -- gfx950 kernels of the interleaved block-cyclic distributed FFT on
// one GPU: the P^d virtual ranks are slabs of one buffer, so the reference's
// pack -> MPI_Alltoall -> unpack becomes addressing inside the kernels.
//
// One direction of the reference's transform (zmodel/transforms/fft.cu:207-264)
// is, for every "line" (the P ranks of an exchange group x one local row q),
// one N-point DFT with N = VEC * M decomposed as
//   FFT1a (VEC-point, plan0) -> twiddle W_N^(r k1) -> FFT1b (M-point, plan1).
// Two kernel paths implement it:
//   * line_kernel: the whole line in LDS, one HBM sweep per direction;
//   * stage_a/stage_b: FFT1a + pack + exchange, then unpack + twiddle + FFT1b,
//     two sweeps with the reference's post-exchange buffer in between (used
//     when a line does not fit LDS, or on request).
// repack_kernel is the LDS-tiled repack-transpose of fft.cu:183-187 /
// fft3d.cu:203-220 and its inverses.
#include "common.hpp"
#include "kernels.hpp"

namespace zm {

// ---------------------------------------------------------------------------
// fused line kernel
// ---------------------------------------------------------------------------
//
// Forward: rank g (member of the group) row q holds, at packed column
// j*VEC + jj, the global sample x[jj*M + g*B + j] of the line (zmodel/utils.h:
// 122-128, 166-170).  LDS keeps the line as the VEC x M matrix
// lds[r*VEC + jj], r = g*B + j, which is exactly the concatenation of the P
// rank rows.  After the VEC-point DFTs, the twiddle and the M-point DFTs,
// lds[k2*VEC + k1] = X[k1 + VEC*k2]; rank g' = k1 / VP receives it at
// position (k1 % VP) + k2*VP of row q -- the layout unpack_forward_fft +
// FFT1b leave (fft.cu:162-181, fft_device_plans.h:64-66).
// Inverse: the same maps read backwards, conjugate twiddles, unnormalised
// inverse DFTs, `scale` applied on the final store.
// COMPUTE=false keeps only the data movement (do_compute=false).
template <typename T, bool FWD, bool COMPUTE>
__global__ void __launch_bounds__(512) line_kernel(const Cx<T>* in, Cx<T>* out, Geo g, LineArgs a,
                                                    const Cx<T>* __restrict__ tw, T scale, int apply_scale) {
    using C = Cx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int N = g.n, L = g.l, VEC = g.vec, VP = g.vp, M = g.m, P = g.p;
    const int E = a.lpb * N;
    const int64_t lam0 = (int64_t)blockIdx.x * a.lpb;
    const int o = (int)(lam0 / a.lines_per_group);
    const int q0 = (int)(lam0 % a.lines_per_group);
    const C* src = in + (int64_t)blockIdx.y * g.field;
    C* dst = out + (int64_t)blockIdx.y * g.field;

    for (int e = tid; e < E; e += nt) {
        const int l = e / N, rem = e - l * N, gm = rem / L, pos = rem - gm * L;
        const int64_t addr = (int64_t)member_rank(o, gm, a.sg, P) * g.local + (int64_t)(q0 + l) * L + pos;
        int li = e;
        if (!FWD) {
            const int k2 = pos / VP, fr = pos - k2 * VP;
            li = l * N + k2 * VEC + gm * VP + fr;
        }
        lds[pidx(li)] = src[addr];
    }
    __syncthreads();

    if constexpr (COMPUTE) {
        const Batch bv{VEC, 1, 1, 0, VEC};  // contiguous VEC-point transforms, bt = (l, r)
        const Batch bm{M, VEC, VEC, 1, N};  // stride-VEC M-point transforms, bt = (l, k1)
        if constexpr (FWD) {
            lds_fft<T, true>(lds, tid, nt, bv, tw, N, PreTwiddle{0, 0});
            lds_fft<T, true>(lds, tid, nt, bm, tw, N, PreTwiddle{VEC, 0});
        } else {
            lds_fft<T, false>(lds, tid, nt, bm, tw, N, PreTwiddle{0, 0});
            lds_fft<T, false>(lds, tid, nt, bv, tw, N, PreTwiddle{M, 0});
        }
    }

    for (int e = tid; e < E; e += nt) {
        const int l = e / N, rem = e - l * N, gm = rem / L, pos = rem - gm * L;
        const int64_t addr = (int64_t)member_rank(o, gm, a.sg, P) * g.local + (int64_t)(q0 + l) * L + pos;
        int li = e;
        if (FWD) {
            const int k2 = pos / VP, fr = pos - k2 * VP;
            li = l * N + k2 * VEC + gm * VP + fr;
        }
        C v = lds[pidx(li)];
        if (!FWD && apply_scale) v = C{v.x * scale, v.y * scale};
        dst[addr] = v;
    }
}

// ---------------------------------------------------------------------------
// two-stage path
// ---------------------------------------------------------------------------
//
// The post-exchange buffer of rank g' is the reference's receive buffer
// (fft.cu:221-241): slot s (from member s) of L^d/P elements, inside it the
// depth slice, then row/P, then ((row % P)*B + j)*VP + fr.

// Stage A forward = FFT1a + pack_forward_fft + exchange (fft.cu:207-241):
// a workgroup reads rows q0.. of rank `id`, runs the VEC-point DFTs and writes
// every lane group k1 / VP straight into the receiving rank's slot.
// Stage A inverse = exchange + pack_inverse_fft x scale + FFT1a^-1
// (fft.cu:478-505).
template <typename T, bool FWD, bool COMPUTE>
__global__ void __launch_bounds__(512) stage_a_kernel(const Cx<T>* in, Cx<T>* out, Geo g, StageArgs a,
                                                       const Cx<T>* __restrict__ tw, T scale, int apply_scale) {
    using C = Cx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int L = g.l, VEC = g.vec, VP = g.vp, P = g.p, B = g.b;
    const int E = a.rpb * L;
    const int bpr = a.rows_per_rank / a.rpb;
    const int id = blockIdx.x / bpr;
    const int q0 = (blockIdx.x % bpr) * a.rpb;
    const int gme = (id / a.sg) % P;
    const int base = id - gme * a.sg;
    const int64_t slab = (int64_t)L * L / P;
    const int run = B * VP;  // contiguous run per (row, peer)
    const C* src = in + (int64_t)blockIdx.y * g.field;
    C* dst = out + (int64_t)blockIdx.y * g.field;

    if (FWD) {
        for (int e = tid; e < E; e += nt) lds[pidx(e)] = src[(int64_t)id * g.local + (int64_t)q0 * L + e];
    } else {
        for (int e = tid; e < E; e += nt) {
            const int rho = e / L, w0 = e - rho * L, peer = w0 / run, w = w0 - peer * run;
            const int j = w / VP, fr = w - j * VP;
            const int q = q0 + rho, dep = q / L, row = q - dep * L;
            const int64_t addr = (int64_t)(base + peer * a.sg) * g.local + gme * slab * a.ndep + dep * slab +
                                 (int64_t)(row / P) * L + (row % P) * run + w;
            lds[pidx(rho * L + j * VEC + peer * VP + fr)] = src[addr];
        }
    }
    __syncthreads();

    if constexpr (COMPUTE) {
        const Batch bv{VEC, 1, 1, 0, VEC};
        lds_fft<T, FWD>(lds, tid, nt, bv, tw, g.n, PreTwiddle{0, 0});
    }

    if (FWD) {
        for (int e = tid; e < E; e += nt) {
            const int rho = e / L, w0 = e - rho * L, peer = w0 / run, w = w0 - peer * run;
            const int j = w / VP, fr = w - j * VP;
            const int q = q0 + rho, dep = q / L, row = q - dep * L;
            const int64_t addr = (int64_t)(base + peer * a.sg) * g.local + gme * slab * a.ndep + dep * slab +
                                 (int64_t)(row / P) * L + (row % P) * run + w;
            dst[addr] = lds[pidx(rho * L + j * VEC + peer * VP + fr)];
        }
    } else {
        for (int e = tid; e < E; e += nt) {
            C v = lds[pidx(e)];
            if (apply_scale) v = C{v.x * scale, v.y * scale};
            dst[(int64_t)id * g.local + (int64_t)q0 * L + e] = v;
        }
    }
}

// Stage B forward = unpack_forward_fft x twiddle + FFT1b (fft.cu:243-264,
// fft_device_plans.h:48-90): rows q0.. of rank `id` gather their P slots,
// apply W_N^(r k1) with k1 = g'*VP + fr, run the stride-VP M-point DFTs and
// store the rows contiguously.  Stage B inverse = FFT1b^-1 + unpack_inverse_fft
// x conjugate twiddle (fft.cu:420-431).
template <typename T, bool FWD, bool COMPUTE>
__global__ void __launch_bounds__(512) stage_b_kernel(const Cx<T>* in, Cx<T>* out, Geo g, StageArgs a,
                                                       const Cx<T>* __restrict__ tw) {
    using C = Cx<T>;
    extern __shared__ unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int tid = threadIdx.x, nt = blockDim.x;
    const int N = g.n, L = g.l, VP = g.vp, M = g.m, P = g.p, B = g.b;
    const int E = a.rpb * L;
    const int bpr = a.rows_per_rank / a.rpb;
    const int id = blockIdx.x / bpr;
    const int q0 = (blockIdx.x % bpr) * a.rpb;
    const int gme = (id / a.sg) % P;
    const int64_t slab = (int64_t)L * L / P;
    const int run = B * VP;
    const C* src = in + (int64_t)blockIdx.y * g.field;
    C* dst = out + (int64_t)blockIdx.y * g.field;

    if (FWD) {
        for (int e = tid; e < E; e += nt) {
            const int rho = e / L, w0 = e - rho * L, s = w0 / run, w = w0 - s * run;
            const int q = q0 + rho, dep = q / L, row = q - dep * L;
            const int64_t addr = (int64_t)id * g.local + s * slab * a.ndep + dep * slab + (int64_t)(row / P) * L +
                                 (row % P) * run + w;
            lds[pidx(e)] = src[addr];
        }
    } else {
        for (int e = tid; e < E; e += nt) lds[pidx(e)] = src[(int64_t)id * g.local + (int64_t)q0 * L + e];
    }
    __syncthreads();

    if constexpr (COMPUTE) {
        const Batch bm{M, VP, VP, 1, L};  // bt = (rho, fr), element r at rho*L + r*VP + fr
        if (FWD) lds_fft<T, true>(lds, tid, nt, bm, tw, N, PreTwiddle{VP, gme * VP});
        else lds_fft<T, false>(lds, tid, nt, bm, tw, N, PreTwiddle{0, 0});
    }

    if (FWD) {
        for (int e = tid; e < E; e += nt) dst[(int64_t)id * g.local + (int64_t)q0 * L + e] = lds[pidx(e)];
    } else {
        for (int e = tid; e < E; e += nt) {
            const int rho = e / L, w0 = e - rho * L, s = w0 / run, w = w0 - s * run;
            const int q = q0 + rho, dep = q / L, row = q - dep * L;
            C v = lds[pidx(e)];
            if (COMPUTE) {
                const int r = w0 / VP, fr = w0 - r * VP;
                v = cmulc(v, tw[(r * (gme * VP + fr)) & (N - 1)]);
            }
            const int64_t addr = (int64_t)id * g.local + s * slab * a.ndep + dep * slab + (int64_t)(row / P) * L +
                                 (row % P) * run + w;
            dst[addr] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// repack-transpose
// ---------------------------------------------------------------------------
//
// Forward: out[e*eo + y*syo + pack(x)] = in[e*ei + x*sxi + y] with
// pack(x) = (x % B)*VEC + x / B, per rank.  2D: repack_transpose
// (fft.cu:183-187); 3D: repack_transpose_xy_3d (x = row, y = col, e = depth)
// and repack_transpose_yz_3d (x = depth, y = col, e = row) (fft3d.cu:203-220).
// A tile holds TX values x = rho + B*t (consecutive t, so pack(x) is a
// contiguous run) by TY consecutive y: both the reads and the writes are
// contiguous runs.  FWD=false is the inverse permutation (fft.cu:393-397,
// fft3d.cu:477-494).
template <typename T, bool FWD>
__global__ void __launch_bounds__(256) repack_kernel(const Cx<T>* in, Cx<T>* out, Geo g, RepackArgs a) {
    using C = Cx<T>;
    __shared__ C tile[REPACK_TX][REPACK_TY + 1];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int L = g.l, B = g.b, VEC = g.vec;
    const int TX = a.tx, TY = a.ty;
    int rest = blockIdx.x;
    const int yb = rest % (L / TY);
    rest /= (L / TY);
    const int tb = rest % (VEC / TX);
    rest /= (VEC / TX);
    const int rho = rest % B;
    rest /= B;
    const int e = rest % a.ecount;
    const int id = rest / a.ecount;
    const int t0 = tb * TX, y0 = yb * TY;
    const C* src = in + (int64_t)blockIdx.y * g.field + (int64_t)id * g.local;
    C* dst = out + (int64_t)blockIdx.y * g.field + (int64_t)id * g.local;
    const int64_t xy_base = (int64_t)e * a.ei;  // x/y side
    const int64_t pk_base = (int64_t)e * a.eo;  // packed side

    if (FWD) {
        for (int k = tid; k < TX * TY; k += nt) {
            const int t = k / TY, y = k - t * TY;
            tile[t][y] = src[xy_base + (int64_t)(rho + B * (t0 + t)) * a.sxi + y0 + y];
        }
        __syncthreads();
        for (int k = tid; k < TX * TY; k += nt) {
            const int y = k / TX, t = k - y * TX;
            dst[pk_base + (int64_t)(y0 + y) * a.syo + rho * VEC + t0 + t] = tile[t][y];
        }
    } else {
        for (int k = tid; k < TX * TY; k += nt) {
            const int y = k / TX, t = k - y * TX;
            tile[t][y] = src[pk_base + (int64_t)(y0 + y) * a.syo + rho * VEC + t0 + t];
        }
        __syncthreads();
        for (int k = tid; k < TX * TY; k += nt) {
            const int t = k / TY, y = k - t * TY;
            dst[xy_base + (int64_t)(rho + B * (t0 + t)) * a.sxi + y0 + y] = tile[t][y];
        }
    }
}

// The same permutation with compile-time tiles of TX = 2^LTX values of x by
// TY = 2^LTY values of y, 256 threads, and 16-byte global accesses (two FP32
// or one FP64 complex value): every index is a shift, every thread keeps
// TX*TY/256 values in flight, and each global row segment is TY (x/y side) or
// TX (packed side) contiguous elements.  The LDS pitch TY + 1 puts the
// column reads of the packed side on distinct banks.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    using type = float4;
    static constexpr int n = 2;
    __device__ static void split(const float4& v, float2* c) {
        c[0] = float2{v.x, v.y};
        c[1] = float2{v.z, v.w};
    }
    __device__ static float4 join(const float2* c) { return float4{c[0].x, c[0].y, c[1].x, c[1].y}; }
};
template <>
struct Vec16<double> {
    using type = double2;
    static constexpr int n = 1;
    __device__ static void split(const double2& v, double2* c) { c[0] = v; }
    __device__ static double2 join(const double2* c) { return c[0]; }
};

template <typename T, bool FWD, int LTX, int LTY, int LKR = 0>
__global__ void __launch_bounds__(256) repack_tiled_kernel(const Cx<T>* __restrict__ in, Cx<T>* __restrict__ out,
                                                           Geo g, RepackArgs a, int lecount, int order) {
    using C = Cx<T>;
    using V16 = Vec16<T>;
    using Vt = typename V16::type;
    // KR = 2^LKR consecutive rho per tile (used with TX = VEC only): the
    // tile's TXE = KR TX rows are x = rho0 + j + B (t0 + t), j < KR, t < TX,
    // whose packed positions (rho0 + j) VEC + t form one run of KR VEC values
    constexpr int TX = 1 << LTX, TY = 1 << LTY, V = V16::n, KR = 1 << LKR, TXE = TX * KR;
    constexpr int XA = TY / V, XR = 256 / XA;   // x/y side: accesses per row, rows per sweep
    constexpr int PA = TXE / V, PR = 256 / PA;  // packed side
    static_assert(XR >= 1 && TX % XR == 0 && PR >= 1 && TY % PR == 0, "tile split");
    // non-temporal loads; non-temporal stores for FP64 only (FP32 plain
    // stores measured faster over whole steps, line pass included)
    auto ldv = [](const Vt* p) -> Vt { return ld_nt(p); };
    auto stv = [](Vt* p, Vt v) {
        if constexpr (sizeof(T) == 8) st_nt(p, v);
        else *p = v;
    };
    __shared__ C tile[TXE][TY + 1];
    const int tid = threadIdx.x;
    // block order (`order`, see launch_repack): bit 0 XCD-aware, bit 1 rho
    // fastest, else y tile fastest
    unsigned rest = (order & 1) ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int lyb = g.ll - LTY, ltb = g.lvec - LTX, lrb = g.lb - LKR;
    int yb, tb, rho;
    if (order & 2) {
        rho = (int)(rest & ((1u << lrb) - 1)) << LKR;
        rest >>= lrb;
        yb = (int)(rest & ((1u << lyb) - 1));
        rest >>= lyb;
        tb = (int)(rest & ((1u << ltb) - 1));
        rest >>= ltb;
    } else {
        yb = (int)(rest & ((1u << lyb) - 1));
        rest >>= lyb;
        tb = (int)(rest & ((1u << ltb) - 1));
        rest >>= ltb;
        rho = (int)(rest & ((1u << lrb) - 1)) << LKR;
        rest >>= lrb;
    }
    const int e = (int)(rest & ((1u << lecount) - 1));
    const int id = (int)(rest >> lecount);
    const C* src = in + (int64_t)blockIdx.y * g.field + (int64_t)id * g.local;
    C* dst = out + (int64_t)blockIdx.y * g.field + (int64_t)id * g.local;
    const int t0 = tb << LTX, y0 = yb << LTY;
    // x/y side: tile row p = j TX + t holds x = rho + j + B*(t0 + t),
    // columns y0 .. y0 + TY - 1; sweep i covers rows xr + i XR, so its j and
    // t offsets are compile-time
    const int xc = (tid % XA) * V, xr = tid / XA;
    const C* xy = (FWD ? src : dst) + (int64_t)e * a.ei + y0 + xc;
    // packed side: row y0 + y holds pack(x) = rho*VEC + t0 .. + TXE - 1
    const int pc = (tid % PA) * V, pr = tid / PA;
    const C* pk = (FWD ? dst : src) + (int64_t)e * a.eo + (int64_t)y0 * a.syo + (rho << g.lvec) + t0 + pc;
    const int64_t xstep = (int64_t)g.b * a.sxi;
    const C* xyr = xy + (int64_t)(rho + g.b * (t0 + xr)) * a.sxi;
    constexpr int SW = TX / XR;  // sweeps per rho
    auto xrow = [&](int i) { return xyr + (int64_t)(i / SW) * a.sxi + (int64_t)(i % SW) * XR * xstep; };
    if constexpr (FWD) {
        Vt r[TXE / XR];
#pragma unroll
        for (int i = 0; i < TXE / XR; i++) r[i] = ldv(reinterpret_cast<const Vt*>(xrow(i)));
#pragma unroll
        for (int i = 0; i < TXE / XR; i++) {
            C c[V];
            V16::split(r[i], c);
#pragma unroll
            for (int v = 0; v < V; v++) tile[xr + i * XR][xc + v] = c[v];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TY / PR; i++) {
            const int y = pr + i * PR;
            C c[V];
#pragma unroll
            for (int v = 0; v < V; v++) c[v] = tile[pc + v][y];
            stv(reinterpret_cast<Vt*>(const_cast<C*>(pk) + (int64_t)y * a.syo), V16::join(c));
        }
    } else {
        Vt r[TY / PR];
#pragma unroll
        for (int i = 0; i < TY / PR; i++) r[i] = ldv(reinterpret_cast<const Vt*>(pk + (int64_t)(pr + i * PR) * a.syo));
#pragma unroll
        for (int i = 0; i < TY / PR; i++) {
            C c[V];
            V16::split(r[i], c);
#pragma unroll
            for (int v = 0; v < V; v++) tile[pc + v][pr + i * PR] = c[v];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TXE / XR; i++) {
            const int t = xr + i * XR;
            C c[V];
#pragma unroll
            for (int v = 0; v < V; v++) c[v] = tile[t][xc + v];
            stv(reinterpret_cast<Vt*>(const_cast<C*>(xrow(i))), V16::join(c));
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <typename T, bool FWD, bool COMPUTE>
static hipError_t set_lds_limit(const void* fn, size_t bytes) {
    (void)sizeof(T);
    if (bytes > 65536) return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return hipSuccess;
}

template <typename T>
hipError_t launch_line(const Cx<T>* in, Cx<T>* out, const Geo& g, const LineArgs& a, const Cx<T>* tw, int batch,
                       bool fwd, bool compute, double scale, hipStream_t stream) {
    const int E = a.lpb * g.n;
    const int nthreads = E / EPT;
    const size_t lds = sizeof(Cx<T>) * padded_elems(E);
    const dim3 grid((unsigned)(g.field / g.n / a.lpb), (unsigned)batch);
    const int apply = scale != 1.0;
    const void* fn;
#define ZM_LINE(F, Cm)                                                                                          \
    do {                                                                                                        \
        fn = (const void*)&line_kernel<T, F, Cm>;                                                               \
        hipError_t err = set_lds_limit<T, F, Cm>(fn, lds);                                                      \
        if (err != hipSuccess) return err;                                                                      \
        hipLaunchKernelGGL((line_kernel<T, F, Cm>), grid, dim3(nthreads), lds, stream, in, out, g, a, tw, (T)scale, \
                           apply);                                                                              \
    } while (0)
    if (fwd && compute) ZM_LINE(true, true);
    else if (fwd) ZM_LINE(true, false);
    else if (compute) ZM_LINE(false, true);
    else ZM_LINE(false, false);
#undef ZM_LINE
    return hipGetLastError();
}

template <typename T>
hipError_t launch_stage_a(const Cx<T>* in, Cx<T>* out, const Geo& g, const StageArgs& a, const Cx<T>* tw, int batch,
                          bool fwd, bool compute, double scale, hipStream_t stream) {
    const int E = a.rpb * g.l;
    const int nthreads = E / EPT;
    const size_t lds = sizeof(Cx<T>) * padded_elems(E);
    const dim3 grid((unsigned)(g.nranks * (a.rows_per_rank / a.rpb)), (unsigned)batch);
    const int apply = scale != 1.0;
#define ZM_STA(F, Cm)                                                                                          \
    do {                                                                                                       \
        hipError_t err = set_lds_limit<T, F, Cm>((const void*)&stage_a_kernel<T, F, Cm>, lds);                 \
        if (err != hipSuccess) return err;                                                                     \
        hipLaunchKernelGGL((stage_a_kernel<T, F, Cm>), grid, dim3(nthreads), lds, stream, in, out, g, a, tw,   \
                           (T)scale, apply);                                                                   \
    } while (0)
    if (fwd && compute) ZM_STA(true, true);
    else if (fwd) ZM_STA(true, false);
    else if (compute) ZM_STA(false, true);
    else ZM_STA(false, false);
#undef ZM_STA
    return hipGetLastError();
}

template <typename T>
hipError_t launch_stage_b(const Cx<T>* in, Cx<T>* out, const Geo& g, const StageArgs& a, const Cx<T>* tw, int batch,
                          bool fwd, bool compute, hipStream_t stream) {
    const int E = a.rpb * g.l;
    const int nthreads = E / EPT;
    const size_t lds = sizeof(Cx<T>) * padded_elems(E);
    const dim3 grid((unsigned)(g.nranks * (a.rows_per_rank / a.rpb)), (unsigned)batch);
#define ZM_STB(F, Cm)                                                                                          \
    do {                                                                                                       \
        hipError_t err = set_lds_limit<T, F, Cm>((const void*)&stage_b_kernel<T, F, Cm>, lds);                 \
        if (err != hipSuccess) return err;                                                                     \
        hipLaunchKernelGGL((stage_b_kernel<T, F, Cm>), grid, dim3(nthreads), lds, stream, in, out, g, a, tw);  \
    } while (0)
    if (fwd && compute) ZM_STB(true, true);
    else if (fwd) ZM_STB(true, false);
    else if (compute) ZM_STB(false, true);
    else ZM_STB(false, false);
#undef ZM_STB
    return hipGetLastError();
}

template <typename T>
hipError_t launch_repack(const Cx<T>* in, Cx<T>* out, const Geo& g, const RepackArgs& a, int batch, bool fwd,
                         hipStream_t stream) {
    // tiled path: VEC >= 16 and L >= 32.  A tile is TX = 2^ltx values of x
    // (times KR = 2^lkr consecutive rho) by TY = 2^lty values of y; `order`
    // is the block order (see repack_tiled_kernel).  From same-box sweeps of
    // x tile (16, 32), y tile (32-128), rho per tile and order at 8192^2
    // (P = 4, B = 32) and 512^3 (P = 2, B = 16):
    // * 2D FP32: 16 x 32 tiles, rho fastest, so that the blocks in flight
    //   together write whole packed rows.  32 x 128 tiles in that order had
    //   taken forward / inverse 0.203 / 0.197 -> 0.170 / 0.174 ms against y
    //   tile fastest; 16 x 32: 0.171 -> 0.152 ms (7.1 TB/s), step 1.187 ->
    //   1.139 ms.
    // * 2D FP64: 32 x 64, rho fastest (0.364 -> 0.354 ms against the same
    //   XCD-aware).
    // * 3D FP32: 16 (or 32) x 32 tiles of one rho, y tile fastest,
    //   XCD-aware.  A tile of the xy repack then spans 32 rows of one plane,
    //   one of the yz repack 32 + 16 planes instead of 128 + 32: 0.387 /
    //   0.382 -> 0.356 / 0.352 ms (y tile fastest in plain order had taken
    //   0.402 / 0.465 ms).
    // * 3D FP64: plain order, two rho per tile at VEC = 16 (XCD-aware and
    //   other tiles no better).
    // * VEC = 16 otherwise: two rho per tile, so that the packed side writes
    //   runs of 2 VEC values like the VEC >= 32 shapes (4096^2, P = 4,
    //   B = 64, 32 x 128 tiles: 0.0586 / 0.0574 -> 0.0567 / 0.0542 ms).
    if (g.vec >= 16 && g.l >= 32) {
        constexpr bool fp32 = sizeof(T) == 4;
        int ltx = g.vec >= 32 ? 5 : 4;
        int lty = g.l >= 64 ? 6 : 5;
        int lkr = (g.vec == 16 && g.b >= 2) ? 1 : 0;
        int order;
        if (g.dims == 2) {
            order = 2;
            if (fp32) {
                ltx = 4;
                lty = 5;
            }
        } else if (fp32) {
            order = 1;
            lty = 5;
            lkr = 0;
        } else {
            order = 0;
        }
        // several rho per tile form one packed run only when the x tile is
        // the whole vector (t0 = 0)
        if ((1 << ltx) != g.vec) lkr = 0;
        int lecount = 0;
        while ((1 << lecount) < a.ecount) lecount++;
        const unsigned blocks =
            (unsigned)((int64_t)g.nranks * a.ecount * (g.b >> lkr) * (g.vec >> ltx) * (g.l >> lty));
        const dim3 grid(blocks, (unsigned)batch);
#define ZM_REPACK(F, X, Y, K)                                                                                   \
    if (fwd == F && ltx == X && lty == Y && lkr == K) {                                                         \
        hipLaunchKernelGGL((repack_tiled_kernel<T, F, X, Y, K>), grid, dim3(256), 0, stream, in, out, g, a,      \
                           lecount, order);                                                                     \
        return hipGetLastError();                                                                               \
    }
        ZM_REPACK(true, 5, 6, 0) ZM_REPACK(true, 5, 5, 0) ZM_REPACK(true, 4, 6, 0) ZM_REPACK(true, 4, 5, 0)
        ZM_REPACK(false, 5, 6, 0) ZM_REPACK(false, 5, 5, 0) ZM_REPACK(false, 4, 6, 0) ZM_REPACK(false, 4, 5, 0)
        ZM_REPACK(true, 4, 6, 1) ZM_REPACK(true, 4, 5, 1) ZM_REPACK(false, 4, 6, 1) ZM_REPACK(false, 4, 5, 1)
#undef ZM_REPACK
    }
    const unsigned blocks = (unsigned)(g.nranks * a.ecount * g.b * (g.vec / a.tx) * (g.l / a.ty));
    const dim3 grid(blocks, (unsigned)batch);
    if (fwd) hipLaunchKernelGGL((repack_kernel<T, true>), grid, dim3(256), 0, stream, in, out, g, a);
    else hipLaunchKernelGGL((repack_kernel<T, false>), grid, dim3(256), 0, stream, in, out, g, a);
    return hipGetLastError();
}

template hipError_t launch_line<float>(const float2*, float2*, const Geo&, const LineArgs&, const float2*, int, bool,
                                       bool, double, hipStream_t);
template hipError_t launch_line<double>(const double2*, double2*, const Geo&, const LineArgs&, const double2*, int,
                                        bool, bool, double, hipStream_t);
template hipError_t launch_stage_a<float>(const float2*, float2*, const Geo&, const StageArgs&, const float2*, int,
                                          bool, bool, double, hipStream_t);
template hipError_t launch_stage_a<double>(const double2*, double2*, const Geo&, const StageArgs&, const double2*, int,
                                           bool, bool, double, hipStream_t);
template hipError_t launch_stage_b<float>(const float2*, float2*, const Geo&, const StageArgs&, const float2*, int,
                                          bool, bool, hipStream_t);
template hipError_t launch_stage_b<double>(const double2*, double2*, const Geo&, const StageArgs&, const double2*, int,
                                           bool, bool, hipStream_t);
template hipError_t launch_repack<float>(const float2*, float2*, const Geo&, const RepackArgs&, int, bool, hipStream_t);
template hipError_t launch_repack<double>(const double2*, double2*, const Geo&, const RepackArgs&, int, bool,
                                          hipStream_t);

}  // namespace zm
