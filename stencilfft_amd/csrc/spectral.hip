// spectral.hip -- FP64 twiddle generation, reciprocal-space multipliers, the
// device forms of the index maps and the natural <-> packed converters.
#include "common.hpp"
#include "kernels.hpp"
#include "maps.hpp"

namespace zm {

// tw[t] = exp(-2 pi i t / n) in FP64 from sincospi with the exact dyadic
// argument 2t/n; the FP32 table is the correctly rounded FP64 value.  Replaces
// the host std::exp loops of init_forward_twiddles (zmodel/transforms/fft.cu:37-54):
// the reference's L x L table holds W_N^(k l) for k = fr + id*VP, l = fc, and
// the kernels index this length-N table with (k l) mod N instead.
__global__ void twiddle_kernel(double2* tw64, float2* tw32, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s, c;
    sincospi(-2.0 * (double)t / (double)n, &s, &c);
    tw64[t] = double2{c, s};
    if (tw32) tw32[t] = float2{(float)c, (float)s};
}

hipError_t launch_twiddles(double2* tw64, float2* tw32, int n, hipStream_t stream) {
    const unsigned threads = 256;
    hipLaunchKernelGGL(twiddle_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, stream, tw64, tw32, n);
    return hipGetLastError();
}

struct Coord {
    int did, rid, cid, ld, lr, lc;
};

__device__ __forceinline__ Coord decode_coord(const Geo& g, int64_t i) {
    Coord c;
    const int id = (int)(i / g.local);
    const int64_t rem = i - (int64_t)id * g.local;
    c.lc = (int)(rem & (g.l - 1));
    c.lr = (int)((rem >> g.ll) & (g.l - 1));
    c.ld = (g.dims == 3) ? (int)(rem >> (2 * g.ll)) : 0;
    c.cid = id & (g.p - 1);
    c.rid = (id >> g.lp) & (g.p - 1);
    c.did = (g.dims == 3) ? (id >> (2 * g.lp)) : 0;
    return c;
}

__device__ __forceinline__ int signed_wavenumber(int k, int n) { return k <= n / 2 ? k : k - n; }

// Frequencies of element i of the reciprocal layout: 2D kx <- (cid, lr),
// ky <- (rid, lc) (zmodel/utils.h:265-268); 3D kc <- (cid, lr),
// kr <- (rid, ld), kd <- (did, lc) (utils.h:298-302).  Returns the linear
// index ky*N + kx / (kd*N + kr)*N + kc and the three signed wavenumbers.
__device__ __forceinline__ int64_t recip_freq(const Geo& g, const MapShape& ms, int64_t i, int* s0, int* s1,
                                              int* s2) {
    const Coord c = decode_coord(g, i);
    if (g.dims == 2) {
        const int kx = recip_to_freq(ms, c.cid, c.lr);
        const int ky = recip_to_freq(ms, c.rid, c.lc);
        *s0 = signed_wavenumber(kx, g.n);
        *s1 = signed_wavenumber(ky, g.n);
        *s2 = 0;
        return (int64_t)ky * g.n + kx;
    }
    const int kc = recip_to_freq(ms, c.cid, c.lr);
    const int kr = recip_to_freq(ms, c.rid, c.ld);
    const int kd = recip_to_freq(ms, c.did, c.lc);
    *s0 = signed_wavenumber(kc, g.n);
    *s1 = signed_wavenumber(kr, g.n);
    *s2 = signed_wavenumber(kd, g.n);
    return ((int64_t)kd * g.n + kr) * g.n + kc;
}

// Real-space linear index gr*N + gc / (gd*N + gr)*N + gc of element i of the
// packed layout (zmodel/utils.h:166-170, 192-197).
__device__ __forceinline__ int64_t real_linear(const Geo& g, const MapShape& ms, int64_t i) {
    const Coord c = decode_coord(g, i);
    const int gr = local_row_to_global(ms, c.rid, c.lr);
    const int gc = packed_col_to_global(ms, c.cid, c.lc);
    if (g.dims == 2) return (int64_t)gr * g.n + gc;
    const int gd = local_row_to_global(ms, c.did, c.ld);
    return ((int64_t)gd * g.n + gr) * g.n + gc;
}

// Symbol of the periodic 5-/7-point Laplacian (scaled by `lap`) plus, in 2D,
// the 4th-order x-derivative [1, -8, 0, 8, -1] / (12 delta) when delta != 0:
// for the mode exp(2 pi i k x / N), theta = 2 pi k / N,
//   Laplacian: lap * sum_d (2 cos theta_d - 2) = -4 lap sum_d sin^2(theta_d / 2)
//   dx4:       i (8 sin theta_x - sin 2 theta_x) / (6 delta)
// (the sin^2 form has no cancellation at small k).  kx is the packed-column
// (gc) wavenumber, the axis zm_dx4 differentiates.
__device__ __forceinline__ void stencil_symbol(const Geo& g, int s0, int s1, int s2, double lap, double delta,
                                               double* re, double* im) {
    const double inv_n = 1.0 / (double)g.n;
    double acc = 0.0;
    const int s[3] = {s0, s1, s2};
    for (int d = 0; d < g.dims; ++d) {
        const double h = sinpi((double)s[d] * inv_n);  // sin(theta_d / 2)
        acc += h * h;
    }
    *re = -4.0 * lap * acc;
    *im = 0.0;
    if (delta != 0.0) {
        const double t = 2.0 * (double)s0 * inv_n;  // theta_x / pi
        *im = (8.0 * sinpi(t) - sinpi(2.0 * t)) / (6.0 * delta);
    }
}

// kind: 0 Poisson -1/|k|^2 (k = 0 -> 0), 1 heat exp(-p0 |k|^2 p1), 2 constant
// p0, 3 stencil symbol (Laplacian scale p0, dx4 delta p1), 4 Poisson times the
// stencil symbol.  conj = 1 applies the complex conjugate multiplier (the
// adjoint); the multiplier is formed in FP64 for every storage precision.
template <typename T>
__global__ void __launch_bounds__(256) spectral_kernel(const Cx<T>* in, Cx<T>* out, Geo g, int kind, double p0,
                                                       double p1, double kscale, int conj) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.field) return;
    const MapShape ms = make_map_shape(g.n, g.p, g.b);
    int s0, s1, s2;
    recip_freq(g, ms, i, &s0, &s1, &s2);
    const double k2 = kscale * kscale * ((double)s0 * s0 + (double)s1 * s1 + (double)s2 * s2);
    const double poisson = (k2 == 0.0) ? 0.0 : -1.0 / k2;  // 1 / (-|k|^2), mean mode removed
    double mr, mi = 0.0;
    if (kind == 0) mr = poisson;
    else if (kind == 1) mr = exp(-p0 * k2 * p1);  // heat: exp(-nu |k|^2 dt)
    else if (kind == 2) mr = p0;
    else {
        stencil_symbol(g, s0, s1, s2, p0, p1, &mr, &mi);
        if (kind == 4) {
            mr *= poisson;
            mi *= poisson;
        }
    }
    if (conj) mi = -mi;
    const int64_t at = (int64_t)blockIdx.y * g.field + i;
    const Cx<T> v = in[at];
    const double vx = (double)v.x, vy = (double)v.y;
    out[at] = Cx<T>{(T)(vx * mr - vy * mi), (T)(vx * mi + vy * mr)};
}

template <typename T>
hipError_t launch_spectral(const Cx<T>* in, Cx<T>* out, const Geo& g, int batch, int kind, double p0, double p1,
                           double kscale, bool conj, hipStream_t stream) {
    const unsigned threads = 256;
    const dim3 grid((unsigned)((g.field + threads - 1) / threads), (unsigned)batch);
    hipLaunchKernelGGL((spectral_kernel<T>), grid, dim3(threads), 0, stream, in, out, g, kind, p0, p1, kscale,
                       conj ? 1 : 0);
    return hipGetLastError();
}

template hipError_t launch_spectral<float>(const float2*, float2*, const Geo&, int, int, double, double, double,
                                           bool, hipStream_t);
template hipError_t launch_spectral<double>(const double2*, double2*, const Geo&, int, int, double, double, double,
                                            bool, hipStream_t);

__global__ void __launch_bounds__(256) index_table_kernel(int64_t* out, Geo g, int reciprocal) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.field) return;
    const MapShape ms = make_map_shape(g.n, g.p, g.b);
    if (reciprocal) {
        int s0, s1, s2;
        out[i] = recip_freq(g, ms, i, &s0, &s1, &s2);
    } else {
        out[i] = real_linear(g, ms, i);
    }
}

hipError_t launch_index_table(int64_t* out, const Geo& g, bool reciprocal, hipStream_t stream) {
    const unsigned threads = 256;
    hipLaunchKernelGGL(index_table_kernel, dim3((unsigned)((g.field + threads - 1) / threads)), dim3(threads), 0,
                       stream, out, g, reciprocal ? 1 : 0);
    return hipGetLastError();
}

// natural (row-major N^d) <-> real-space packed layout, one thread per packed
// element: the composition of the block-cyclic distribution and vector_pack
// that init_host performs on the host (zmodel/utils.cu:9-36, utils.h:99-108).
template <typename T>
__global__ void __launch_bounds__(256) natural_packed_kernel(const Cx<T>* in, Cx<T>* out, Geo g, int to_packed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.field) return;
    const MapShape ms = make_map_shape(g.n, g.p, g.b);
    const int64_t nat = real_linear(g, ms, i);
    const int64_t f = (int64_t)blockIdx.y * g.field;
    if (to_packed) out[f + i] = in[f + nat];
    else out[f + nat] = in[f + i];
}

template <typename T>
hipError_t launch_natural_packed(const Cx<T>* in, Cx<T>* out, const Geo& g, int batch, bool to_packed,
                                 hipStream_t stream) {
    const unsigned threads = 256;
    const dim3 grid((unsigned)((g.field + threads - 1) / threads), (unsigned)batch);
    hipLaunchKernelGGL((natural_packed_kernel<T>), grid, dim3(threads), 0, stream, in, out, g, to_packed ? 1 : 0);
    return hipGetLastError();
}

template hipError_t launch_natural_packed<float>(const float2*, float2*, const Geo&, int, bool, hipStream_t);
template hipError_t launch_natural_packed<double>(const double2*, double2*, const Geo&, int, bool, hipStream_t);

}  // namespace zm
