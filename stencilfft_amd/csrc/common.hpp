// common.hpp -- device building blocks shared by the gfx950 kernels.
//
// Complex values are HIP's float2 / double2 (interleaved re, im), the layout
// of cufftDoubleComplex that the reference stores (zmodel/utils.h:9).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace zm {

template <typename T>
struct CxOf;
template <>
struct CxOf<float> {
    using type = float2;
    using vec2 = float4;  // two complex values, one 16-byte access
};
template <>
struct CxOf<double> {
    using type = double2;
    using vec2 = double4;
};
template <typename T>
using Cx = typename CxOf<T>::type;

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
    return C{a.x + b.x, a.y + b.y};
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
    return C{a.x - b.x, a.y - b.y};
}
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename C>
__device__ __forceinline__ C cmulc(C a, C b) {
    return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// multiply by exp(-i pi/2) = -i (forward) or +i (inverse)
template <bool FWD, typename C>
__device__ __forceinline__ C mul_quarter(C a) {
    if constexpr (FWD) return C{a.y, -a.x};
    else return C{-a.y, a.x};
}
// a * w for the forward direction, a * conj(w) for the inverse one; tables
// hold w = exp(-2 pi i t / n).
template <bool FWD, typename C>
__device__ __forceinline__ C ctw(C a, C w) {
    if constexpr (FWD) return cmul(a, w);
    else return cmulc(a, w);
}

// ---------------------------------------------------------------------------
// Streaming global accesses: the hot kernels read and write every element once
// per launch, so loads and stores carry the non-temporal hint (a measured
// 5-6 points of HBM efficiency on gfx950 for a plain copy).
// ---------------------------------------------------------------------------
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float ld_nt(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double ld_nt(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float2 ld_nt(const float2* p) {
    const nt_f2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f2*>(p));
    return float2{v[0], v[1]};
}
__device__ __forceinline__ double2 ld_nt(const double2* p) {
    const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2*>(p));
    return double2{v[0], v[1]};
}
__device__ __forceinline__ float4 ld_nt(const float4* p) {
    const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
    return float4{v[0], v[1], v[2], v[3]};
}
__device__ __forceinline__ void st_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
    __builtin_nontemporal_store(nt_f2{v.x, v.y}, reinterpret_cast<nt_f2*>(p));
}
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
    __builtin_nontemporal_store(nt_d2{v.x, v.y}, reinterpret_cast<nt_d2*>(p));
}
__device__ __forceinline__ void st_nt(float4* p, float4 v) {
    __builtin_nontemporal_store(nt_f4{v.x, v.y, v.z, v.w}, reinterpret_cast<nt_f4*>(p));
}

// ---------------------------------------------------------------------------
// XCD-aware block order.  gfx950 dispatches consecutive workgroups round-robin
// over its 8 XCDs, each with its own L2.  For stencils, whose neighbour rows
// sit in the neighbouring blocks, remap so that XCD k runs the k-th contiguous
// eighth of the grid: neighbours then share an L2.  Identity unless the block
// count is a multiple of 8.
// ---------------------------------------------------------------------------
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    if (nb % kXcds) return b;
    return (b % kXcds) * (nb / kXcds) + b / kXcds;
}

// ---------------------------------------------------------------------------
// Register DFTs of radix 2..16 (unnormalised, natural order), written as a
// decimation-in-time recursion with exact constants for the R-th roots of
// unity, R <= 16.  FWD uses exp(-2 pi i / R), the inverse its conjugate.
// ---------------------------------------------------------------------------
struct Roots16 {
    // cos and sin of 2 pi k / 16, k = 0..7
    static constexpr double c[8] = {1.0,
                                    0.923879532511286756128183189396788933,
                                    0.707106781186547524400844362104849039,
                                    0.382683432365089771728459984030398866,
                                    0.0,
                                    -0.382683432365089771728459984030398866,
                                    -0.707106781186547524400844362104849039,
                                    -0.923879532511286756128183189396788933};
    static constexpr double s[8] = {0.0,
                                    0.382683432365089771728459984030398866,
                                    0.707106781186547524400844362104849039,
                                    0.923879532511286756128183189396788933,
                                    1.0,
                                    0.923879532511286756128183189396788933,
                                    0.707106781186547524400844362104849039,
                                    0.382683432365089771728459984030398866};
};

// multiply by W_16^idx = exp(-+ 2 pi i idx / 16), idx in [0, 8)
template <int IDX, bool FWD, typename C>
__device__ __forceinline__ C mul_w16(C a) {
    using T = decltype(a.x);
    if constexpr (IDX == 0) {
        return a;
    } else if constexpr (IDX == 4) {
        return mul_quarter<FWD>(a);
    } else {
        constexpr T c = (T)Roots16::c[IDX];
        constexpr T s = FWD ? -(T)Roots16::s[IDX] : (T)Roots16::s[IDX];
        return C{a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

template <int R, bool FWD, int K = 0, typename C>
__device__ __forceinline__ void combine(C* v, const C* e, const C* o) {
    if constexpr (K < R / 2) {
        C t = mul_w16<K*(16 / R), FWD>(o[K]);
        v[K] = cadd(e[K], t);
        v[K + R / 2] = csub(e[K], t);
        combine<R, FWD, K + 1>(v, e, o);
    }
}

template <int R, bool FWD, typename C>
__device__ __forceinline__ void dft(C* v) {
    static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 16, "radix");
    if constexpr (R == 2) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        C t2 = cadd(v[1], v[3]), t3 = mul_quarter<FWD>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    } else if constexpr (R > 4) {
        C e[R / 2], o[R / 2];
#pragma unroll
        for (int k = 0; k < R / 2; k++) {
            e[k] = v[2 * k];
            o[k] = v[2 * k + 1];
        }
        dft<R / 2, FWD>(e);
        dft<R / 2, FWD>(o);
        combine<R, FWD>(v, e, o);
    }
}

// ---------------------------------------------------------------------------
// LDS layout: one padding element after every 32, so power-of-two strides
// spread over the banks.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int pidx(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int padded_elems(int e) { return e + (e >> 5) + 1; }

// Elements each thread owns during an LDS pass.  A workgroup of E elements
// runs E / EPT threads, so every radix-R pass is exactly EPT / R butterflies
// per thread and a pass can read everything before writing anything back.
constexpr int EPT = 16;

// Description of a batch of equal-length transforms inside LDS: element i of
// transform bt lives at base(bt) + i * es with
// base(bt) = (bt / bdiv) * bso + (bt % bdiv) * bsi.
struct Batch {
    int n;     // transform length (power of two)
    int es;    // element stride
    int bdiv;  // batch split
    int bsi;   // inner batch stride
    int bso;   // outer batch stride
};

// Optional twiddle applied on the first pass: element i of transform bt is
// multiplied by W_N^(k i) with k = koff + bt % bmod (bmod = 0: none).
struct PreTwiddle {
    int bmod;
    int koff;
};

// One in-place Stockham autosort radix-R pass (Govindaraju et al., "High
// performance discrete Fourier transforms on graphics processors", SC08; Van
// Loan, Computational Frameworks for the FFT, 1992, sec. 1.7).  `ns` is the
// product of the radices already applied; the output of the last pass is in
// natural order.  `tw` holds exp(-2 pi i t / twn), twn a multiple of n.
template <typename T, int R, bool FWD>
__device__ __forceinline__ void stockham_pass(Cx<T>* lds, int tid, int nthreads, const Batch& bt_desc, int ns,
                                              const Cx<T>* __restrict__ tw, int twn, PreTwiddle pre) {
    using C = Cx<T>;
    constexpr int BPT = EPT / R;
    const int nbf = bt_desc.n / R;
    const int twstep = twn / (ns * R);
    C v[BPT][R];
    int obase[BPT];
#pragma unroll
    for (int k = 0; k < BPT; k++) {
        const int u = tid + k * nthreads;
        const int bt = u / nbf;
        const int j = u - bt * nbf;
        const int base = (bt / bt_desc.bdiv) * bt_desc.bso + (bt % bt_desc.bdiv) * bt_desc.bsi;
#pragma unroll
        for (int r = 0; r < R; r++) v[k][r] = lds[pidx(base + (j + r * nbf) * bt_desc.es)];
        if (pre.bmod > 0) {
            const int kk = pre.koff + bt % pre.bmod;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int e = ((j + r * nbf) * kk) & (twn - 1);
                v[k][r] = ctw<FWD>(v[k][r], tw[e]);
            }
        }
        const int jm = j & (ns - 1);
        if (ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[k][r] = ctw<FWD>(v[k][r], tw[jm * r * twstep]);
        }
        dft<R, FWD>(v[k]);
        obase[k] = base + ((j - jm) * R + jm) * bt_desc.es;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < BPT; k++) {
#pragma unroll
        for (int r = 0; r < R; r++) lds[pidx(obase[k] + r * ns * bt_desc.es)] = v[k][r];
    }
    __syncthreads();
}

// Full transform of every member of a batch: greedy radix-16 passes, the
// remainder last.  The pre-twiddle (if any) is applied on the first pass.
template <typename T, bool FWD>
__device__ __forceinline__ void lds_fft(Cx<T>* lds, int tid, int nthreads, const Batch& b, const Cx<T>* tw, int twn,
                                        PreTwiddle pre) {
    int ns = 1;
    while (ns < b.n) {
        const int rem = b.n / ns;
        if (rem >= 16) stockham_pass<T, 16, FWD>(lds, tid, nthreads, b, ns, tw, twn, pre), ns *= 16;
        else if (rem == 8) stockham_pass<T, 8, FWD>(lds, tid, nthreads, b, ns, tw, twn, pre), ns *= 8;
        else if (rem == 4) stockham_pass<T, 4, FWD>(lds, tid, nthreads, b, ns, tw, twn, pre), ns *= 4;
        else stockham_pass<T, 2, FWD>(lds, tid, nthreads, b, ns, tw, twn, pre), ns *= 2;
        pre.bmod = 0;
    }
}

// ---------------------------------------------------------------------------
// Geometry of one plan, passed by value to every kernel.
// ---------------------------------------------------------------------------
struct Geo {
    int dims;
    int n, p, b;
    int l, vec, vp, m;
    int nranks;
    int64_t local;  // L^dims
    int64_t field;  // N^dims
    // log2 of the power-of-two extents, for shift-based index decoding
    int ln, lp, lb, ll, lvec;
};

// Rank id of member g of the exchange group `o` along an axis whose member
// stride in world-rank ids is sg (1: row_comm, P: col_comm, P*P: dep_comm).
__host__ __device__ __forceinline__ int member_rank(int o, int g, int sg, int p) {
    return (o / sg) * sg * p + (o % sg) + g * sg;
}

}  // namespace zm
