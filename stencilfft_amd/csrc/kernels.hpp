// kernels.hpp -- launch interface between the host plan (plan.cpp) and the
// gfx950 kernels (fft.hip, stencil.hip, spectral.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace zm {

// fused line kernel: lines of one exchange axis
struct LineArgs {
    int sg;               // world-rank stride between members of the exchange group
    int lines_per_group;  // L^(dims-1) local rows per rank
    int lpb;              // lines per workgroup
};

// two-stage kernels: rank rows
struct StageArgs {
    int sg;             // world-rank stride between members of the exchange group
    int rows_per_rank;  // L^(dims-1)
    int rpb;            // rows per workgroup
    int ndep;           // depth slices of a slot (1 in 2D, L in 3D)
};

constexpr int REPACK_TX = 32;
constexpr int REPACK_TY = 32;

struct RepackArgs {
    int64_t ei, eo;  // stride of the untouched index on the x/y side and on the packed side
    int64_t sxi;     // stride of x on the x/y side (y has stride 1)
    int64_t syo;     // stride of y on the packed side (pack(x) has stride 1)
    int ecount;      // extent of the untouched index
    int tx, ty;      // tile extents (<= REPACK_TX, REPACK_TY)
};

template <typename T>
hipError_t launch_line(const Cx<T>* in, Cx<T>* out, const Geo& g, const LineArgs& a, const Cx<T>* tw, int batch,
                       bool fwd, bool compute, double scale, hipStream_t stream);
template <typename T>
hipError_t launch_stage_a(const Cx<T>* in, Cx<T>* out, const Geo& g, const StageArgs& a, const Cx<T>* tw, int batch,
                          bool fwd, bool compute, double scale, hipStream_t stream);
template <typename T>
hipError_t launch_stage_b(const Cx<T>* in, Cx<T>* out, const Geo& g, const StageArgs& a, const Cx<T>* tw, int batch,
                          bool fwd, bool compute, hipStream_t stream);
template <typename T>
hipError_t launch_repack(const Cx<T>* in, Cx<T>* out, const Geo& g, const RepackArgs& a, int batch, bool fwd,
                         hipStream_t stream);

// compile-time specialised fused line pass (line_fast.hip)
template <typename T>
bool fast_line_supported(int vec, int m);
template <typename T>
int fast_line_occupancy(int vec, int m, bool fwd, bool mf, bool lapx);
// whether the fused Laplacian + x pass (lapx) exists for this (VEC, M) pair
// in `dims` dimensions with block size `b`
template <typename T>
bool fast_line_lapx(int vec, int m, bool mf, int dims, int b);
// log2 of the lines per group of the specialised pass for N = vec * m
int fast_line_lapx_lines_log2(int vec, int m);
// mf: radix-16 passes as split-precision FP16 MFMA GEMMs (FP32 only).
// lapx (forward x pass, fast_line_lapx): transform `scale` times the 5-point
// (2D) / 7-point (3D) Laplacian of the input (the fused Laplacian + forward
// transform).  3D: grid_x = (N / lines per group) row tiles times the number
// of plane chunks, which must divide N.
template <typename T>
hipError_t launch_line_fast(const Cx<T>* in, Cx<T>* out, const Geo& g, const LineArgs& a, const Cx<T>* tables,
                            int batch, bool fwd, double scale, int grid_x, bool mf, bool lapx, hipStream_t stream);
int line_table_entries(int vec, int m);
hipError_t launch_line_tables(double2* t64, float2* t32, int vec, int m, hipStream_t stream);

// stencils: E = element type (T or Cx<T>)
enum StencilOp { OP_LAPLACE = 0, OP_DX4 = 1, OP_DY4 = 2, OP_DZ4 = 3 };
template <typename T>
hipError_t launch_stencil(const void* in, void* out, const Geo& g, int batch, StencilOp op, bool complex_elems,
                          double coef, bool negate, hipStream_t stream);

// twiddle tables exp(-2 pi i t / n), t in [0, n), computed in FP64
hipError_t launch_twiddles(double2* tw64, float2* tw32, int n, hipStream_t stream);

// reciprocal-space multiplier
template <typename T>
hipError_t launch_spectral(const Cx<T>* in, Cx<T>* out, const Geo& g, int batch, int kind, double p0, double p1,
                           double kscale, bool conj, hipStream_t stream);

// index tables and natural <-> packed conversion
hipError_t launch_index_table(int64_t* out, const Geo& g, bool reciprocal, hipStream_t stream);
template <typename T>
hipError_t launch_natural_packed(const Cx<T>* in, Cx<T>* out, const Geo& g, int batch, bool to_packed,
                                 hipStream_t stream);

}  // namespace zm
