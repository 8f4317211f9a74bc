// zmodel.hpp -- header-only C++ facade with the reference's zmodel names over
// the C ABI of stencilfft.h, for code written against the reference's
// zmodel/transforms/fft.h, zmodel/stencils/laplace.h and zmodel/utils.h.
//
// What changes for a caller of the reference:
//   * the P x P (P x P x P) virtual ranks live in one GPU: a FftBuffers holds
//     the whole virtual grid (rank buffers concatenated in world-rank order),
//     and the MPI communicators / rank ids disappear from the signatures;
//   * shapes are runtime values (zm_shape) instead of -DN_DIM/-DB_DIM/-DP_DIM;
//   * forward_fft reads device_buf0 and writes device_buf1 like the
//     reference (fft.cu:189-198) but leaves device_buf0 intact;
//   * errors throw zmodel::Error (the reference prints and exits,
//     zmodel/device_macros.h:112-152).
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

#include "stencilfft.h"
#include "stencilfft_hdf5.h"

namespace zmodel {

// cufftDoubleComplex layout (zmodel/utils.h:9)
struct Complex {
    double x, y;
};
using Scalar = double;

struct Error : std::runtime_error {
    zm_status status;
    Error(zm_status s, const std::string& where) : std::runtime_error(where + ": " + zm_strerror(s)), status(s) {}
};

inline void check(zm_status s, const char* where) {
    if (s != ZM_OK) throw Error(s, where);
}

inline void check_hip(hipError_t e, const char* where) {
    if (e != hipSuccess) throw Error(ZM_ERR_DEVICE, std::string(where) + " (" + hipGetErrorString(e) + ")");
}

// ---------------------------------------------------------------------------
// plans: replace FftHostPlans + init_plans / destroy_plans
// (zmodel/transforms/fft_host_plans.h:6-132) and FftdxFft1Ctx
// ---------------------------------------------------------------------------
struct FftHostPlans {
    zm_plan* plan = nullptr;
    zm_shape shape{};
    zm_dims dims{};
    zm_precision precision = ZM_FP64;
};

inline void init_plans(FftHostPlans& p, const zm_shape& shape, zm_precision precision = ZM_FP64, int batch = 1,
                       unsigned flags = 0) {
    p.shape = shape;
    p.precision = precision;
    check(zm_shape_check(&shape, &p.dims), "zm_shape_check");
    check(zm_plan_create(&p.plan, &shape, precision, batch, flags), "zm_plan_create");
}

inline void init_plans_3d(FftHostPlans& p, const zm_shape& shape, zm_precision precision = ZM_FP64) {
    init_plans(p, shape, precision);
}

inline void destroy_plans(FftHostPlans& p) {
    if (p.plan) check(zm_plan_destroy(p.plan), "zm_plan_destroy");
    p.plan = nullptr;
}

inline void destroy_plans_3d(FftHostPlans& p) { destroy_plans(p); }

// ---------------------------------------------------------------------------
// buffers: replace FftBuffers / init_buffers / destroy_buffers
// (zmodel/transforms/fft.h:9-34, fft.cu:75-129).  Twiddles live in the plan.
// ---------------------------------------------------------------------------
struct FftBuffers {
    void* host_buf0 = nullptr;  // pinned, one whole virtual grid
    void* host_buf1 = nullptr;
    void* device_buf0 = nullptr;
    void* device_buf1 = nullptr;
    void* device_scratch = nullptr;
    size_t bytes = 0;
};

template <bool include_inverse>
inline void init_buffers(FftBuffers& b, const FftHostPlans& p) {
    b.bytes = zm_field_bytes(p.plan);
    check_hip(hipHostMalloc(&b.host_buf0, b.bytes, hipHostMallocDefault), "hipHostMalloc");
    check_hip(hipHostMalloc(&b.host_buf1, b.bytes, hipHostMallocDefault), "hipHostMalloc");
    check_hip(hipMalloc(&b.device_buf0, b.bytes), "hipMalloc");
    check_hip(hipMalloc(&b.device_buf1, b.bytes), "hipMalloc");
    const size_t s = zm_scratch_bytes(p.plan);
    if (s) check_hip(hipMalloc(&b.device_scratch, s), "hipMalloc");
}

template <bool include_inverse>
inline void destroy_buffers(FftBuffers& b) {
    if (b.host_buf0) check_hip(hipHostFree(b.host_buf0), "hipHostFree");
    if (b.host_buf1) check_hip(hipHostFree(b.host_buf1), "hipHostFree");
    if (b.device_buf0) check_hip(hipFree(b.device_buf0), "hipFree");
    if (b.device_buf1) check_hip(hipFree(b.device_buf1), "hipFree");
    if (b.device_scratch) check_hip(hipFree(b.device_scratch), "hipFree");
    b = FftBuffers{};
}

// ---------------------------------------------------------------------------
// transforms (zmodel/transforms/fft.h:36-60)
// ---------------------------------------------------------------------------

// forward_fft<do_compute>: device_buf0 (real-space packed) -> device_buf1
// (reciprocal layout).  do_compute=false runs the data movement alone.
template <bool do_compute>
inline void forward_fft(FftBuffers& b, FftHostPlans& p, hipStream_t stream = nullptr) {
    if (do_compute) check(zm_forward(p.plan, b.device_buf0, b.device_buf1, b.device_scratch, stream), "forward_fft");
    else check(zm_permute_forward(p.plan, b.device_buf0, b.device_buf1, b.device_scratch, stream), "forward_fft");
}

// inverse_fft<do_compute>: device_buf0 (reciprocal) -> device_buf1 (real
// space), unnormalised inverse DFTs with `scale` applied once (fft.cu:417-418).
template <bool do_compute>
inline void inverse_fft(FftBuffers& b, FftHostPlans& p, Complex scale, hipStream_t stream = nullptr) {
    if (scale.y != 0.0) throw Error(ZM_ERR_INVALID_ARG, "inverse_fft: complex scale");
    if (do_compute)
        check(zm_inverse(p.plan, b.device_buf0, b.device_buf1, scale.x, b.device_scratch, stream), "inverse_fft");
    else check(zm_permute_inverse(p.plan, b.device_buf0, b.device_buf1, b.device_scratch, stream), "inverse_fft");
}

template <bool do_compute>
inline void forward_fft_3d(FftBuffers& b, FftHostPlans& p, hipStream_t stream = nullptr) {
    forward_fft<do_compute>(b, p, stream);
}

template <bool do_compute>
inline void inverse_fft_3d(FftBuffers& b, FftHostPlans& p, Complex scale, hipStream_t stream = nullptr) {
    inverse_fft<do_compute>(b, p, scale, stream);
}

// ---------------------------------------------------------------------------
// stencils (zmodel/stencils/laplace.h:9-38).  On one GPU the halo exchange
// is a direct read of the neighbour slabs, so the pack / communicate steps
// of the reference have nothing left to do.
// ---------------------------------------------------------------------------
struct LaplaceBuffers {};
struct LaplaceBuffers3d {};
inline void init_laplace_buffers(LaplaceBuffers&) {}
inline void destroy_laplace_buffers(LaplaceBuffers&) {}
inline void init_laplace_buffers_3d(LaplaceBuffers3d&) {}
inline void destroy_laplace_buffers_3d(LaplaceBuffers3d&) {}
inline void pack_laplace(const void*, LaplaceBuffers&) {}
inline void communicate_laplace(LaplaceBuffers&) {}

// laplace (laplace.cu:73-169) / laplace_3d: unscaled 5- / 7-point stencil
inline void laplace(FftHostPlans& p, const void* device_in, void* device_out, hipStream_t stream = nullptr) {
    check(zm_laplace(p.plan, device_in, device_out, 1.0, ZM_COMPLEX, stream), "laplace");
}

// dx / dy (dx.cu:10-82, dy.cu:10-103) on real Scalar fields
inline void dx(FftHostPlans& p, const Scalar* device_in, Scalar* device_out, Scalar delta,
               hipStream_t stream = nullptr) {
    check(zm_dx4(p.plan, device_in, device_out, delta, ZM_REAL, stream), "dx");
}

inline void dy(FftHostPlans& p, const Scalar* device_in, Scalar* device_out, Scalar delta,
               hipStream_t stream = nullptr) {
    check(zm_dy4(p.plan, device_in, device_out, delta, ZM_REAL, stream), "dy");
}

// dz: the same stencil along depth on a 3D plan (no reference counterpart)
inline void dz(FftHostPlans& p, const Scalar* device_in, Scalar* device_out, Scalar delta,
               hipStream_t stream = nullptr) {
    check(zm_dz4(p.plan, device_in, device_out, delta, ZM_REAL, stream), "dz");
}

// ---------------------------------------------------------------------------
// utils.h
// ---------------------------------------------------------------------------

// init_host / init_host_3d (zmodel/utils.cu:9-81) for the whole virtual grid
inline void init_host(const zm_shape& s, Complex* host_grid) {
    check(zm_init_host(&s, reinterpret_cast<double*>(host_grid), 2), "init_host");
}

inline void init_host(const zm_shape& s, Scalar* host_grid) { check(zm_init_host(&s, host_grid, 1), "init_host"); }

// init_host_hdf5 / store_host_hdf5 and their _3d forms (zmodel/utils.cu:103-281)
// for the whole virtual grid, in the reference's file format (one N^d dataset
// of {"real", "imag"} doubles, natural order).  Callers link
// libstencilfft_hdf5 as well; nothing else in this header needs it.
inline void init_host_hdf5(const zm_shape& s, Complex* host_grid, const char* filename, const char* dataset_name) {
    check(zm_load_hdf5(&s, reinterpret_cast<double*>(host_grid), filename, dataset_name), "init_host_hdf5");
}
inline void store_host_hdf5(const zm_shape& s, const Complex* host_grid, const char* filename,
                            const char* dataset_name) {
    check(zm_store_hdf5(&s, reinterpret_cast<const double*>(host_grid), filename, dataset_name), "store_host_hdf5");
}

inline int packed_to_block_cyclic(const zm_shape& s, int c) { return zm_packed_to_block_cyclic(&s, c); }
inline int block_cyclic_to_packed(const zm_shape& s, int c) { return zm_block_cyclic_to_packed(&s, c); }
inline void real_to_global(const zm_shape& s, int rid, int cid, int lr, int lc, int& gr, int& gc) {
    zm_real_to_global(&s, rid, cid, lr, lc, &gr, &gc);
}
inline void global_to_real(const zm_shape& s, int gr, int gc, int& rid, int& cid, int& lr, int& lc) {
    zm_global_to_real(&s, gr, gc, &rid, &cid, &lr, &lc);
}
inline void real_to_global_3d(const zm_shape& s, int rid, int cid, int did, int ld, int lr, int lc, int& gd, int& gr,
                              int& gc) {
    zm_real_to_global_3d(&s, rid, cid, did, ld, lr, lc, &gd, &gr, &gc);
}
inline void global_to_real_3d(const zm_shape& s, int gd, int gr, int gc, int& rid, int& cid, int& did, int& ld,
                              int& lr, int& lc) {
    zm_global_to_real_3d(&s, gd, gr, gc, &rid, &cid, &did, &ld, &lr, &lc);
}
inline int lp(const zm_shape& s, int i) { return zm_lp(&s, i); }
inline int ct(const zm_shape& s, int p) { return zm_ct(&s, p); }
inline int lp_inv(const zm_shape& s, int j) { return zm_lp_inv(&s, j); }
inline int ct_inv(const zm_shape& s, int k) { return zm_ct_inv(&s, k); }
inline void reciprocal_to_global(const zm_shape& s, int rid, int cid, int lr, int lc, int& ky, int& kx) {
    zm_reciprocal_to_global(&s, rid, cid, lr, lc, &ky, &kx);
}
inline void global_to_reciprocal(const zm_shape& s, int ky, int kx, int& rid, int& cid, int& lr, int& lc) {
    zm_global_to_reciprocal(&s, ky, kx, &rid, &cid, &lr, &lc);
}
inline void reciprocal_to_global_3d(const zm_shape& s, int rid, int cid, int did, int ld, int lr, int lc, int& kd,
                                    int& kr, int& kc) {
    zm_reciprocal_to_global_3d(&s, rid, cid, did, ld, lr, lc, &kd, &kr, &kc);
}
inline void global_to_reciprocal_3d(const zm_shape& s, int kd, int kr, int kc, int& rid, int& cid, int& did, int& ld,
                                    int& lr, int& lc) {
    zm_global_to_reciprocal_3d(&s, kd, kr, kc, &rid, &cid, &did, &ld, &lr, &lc);
}

// element offset of rank (rid, cid[, did])'s local buffer inside a virtual grid
inline size_t rank_offset(const FftHostPlans& p, int rid, int cid, int did = 0) {
    const int P = p.shape.p;
    const int id = p.shape.dims == 3 ? (did * P * P + rid * P + cid) : (rid * P + cid);
    return (size_t)id * (size_t)p.dims.local;
}

}  // namespace zmodel
