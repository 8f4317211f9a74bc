/*
 * oracle.h -- CPU FP64 restatement of StencilFFT's zmodel hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
 * else.  The product (stencilfft_amd/) never links, loads or calls it.
 *
 * Every function restates one piece of the reference and cites the reference
 * file:line it follows (paths relative to the reference repository root).
 *
 * Data model.  The reference runs one MPI rank per block of the P x P (2D) or
 * P x P x P (3D) virtual grid.  Here the whole virtual grid lives in one array:
 * the P^d rank-local buffers are concatenated in world-rank order, rank
 * id = rid*P + cid (2D, zmodel/transforms/fft.cu:563-567) or
 * id = did*P*P + rid*P + cid (3D, zmodel/transforms/fft3d.cu:677-679), each of
 * L^d elements.  Complex values are interleaved (re, im) doubles, exactly the
 * cufftDoubleComplex layout of zmodel/utils.h:9.
 */
#ifndef STENCILFFT_ORACLE_H
#define STENCILFFT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dims;     /* 2 or 3 */
    int n;        /* N_DIM   (zmodel/utils.h:28-30) */
    int p;        /* P_DIM   (zmodel/utils.h:34-38) */
    int b;        /* B_DIM   (zmodel/utils.h:31-33) */
    int l;        /* LOCAL_DIM = N/P (zmodel/utils.h:47) */
    int vec;      /* VEC = L/B (zmodel/utils.h:60) */
    int vp;       /* VEC/P, the FFT1b stride (zmodel/transforms/fft_host_plans.h:111) */
    int m;        /* L/VP = B*P, the FFT1b size (zmodel/transforms/fft_host_plans.h:108) */
    int nranks;   /* P^dims */
    int64_t local;   /* L^dims elements per rank */
    int64_t total;   /* N^dims elements in the virtual grid */
} zo_shape;

/* Returns 0 for a valid shape, -1 otherwise.  Validity: P | N, B | L, P | VEC,
 * VEC >= 1 (the reference's sweep enforces B <= N/P^2, run.sh:11). */
int zo_shape_init(zo_shape* s, int dims, int n, int p, int b);

/* ---- index maps (zmodel/utils.h:122-320) ---- */
int zo_packed_to_block_cyclic(const zo_shape* s, int packed_col);
int zo_block_cyclic_to_packed(const zo_shape* s, int block_cyclic_col);
void zo_real_to_global(const zo_shape* s, int rid, int cid, int lr, int lc, int* gr, int* gc);
void zo_global_to_real(const zo_shape* s, int gr, int gc, int* rid, int* cid, int* lr, int* lc);
void zo_real_to_global_3d(const zo_shape* s, int rid, int cid, int did, int ld, int lr, int lc,
                          int* gd, int* gr, int* gc);
void zo_global_to_real_3d(const zo_shape* s, int gd, int gr, int gc, int* rid, int* cid, int* did,
                          int* ld, int* lr, int* lc);
int zo_lp(const zo_shape* s, int i);
int zo_ct(const zo_shape* s, int p);
int zo_lp_inv(const zo_shape* s, int j);
int zo_ct_inv(const zo_shape* s, int k);
void zo_reciprocal_to_global(const zo_shape* s, int rid, int cid, int lr, int lc, int* ky, int* kx);
void zo_global_to_reciprocal(const zo_shape* s, int ky, int kx, int* rid, int* cid, int* lr, int* lc);
void zo_reciprocal_to_global_3d(const zo_shape* s, int rid, int cid, int did, int ld, int lr, int lc,
                                int* kd, int* kr, int* kc);
void zo_global_to_reciprocal_3d(const zo_shape* s, int kd, int kr, int kc, int* rid, int* cid, int* did,
                                int* ld, int* lr, int* lc);

/* vector_pack / vector_unpack of one rank (zmodel/utils.h:99-119); elements of
 * `ncomp` doubles (1 = Scalar, 2 = Complex). */
void zo_vector_pack(const zo_shape* s, const double* block_cyclic, double* packed, int64_t num_rows, int ncomp);
void zo_vector_unpack(const zo_shape* s, const double* packed, double* block_cyclic, int64_t num_rows, int ncomp);

/* Whole-grid conversions between the natural N^d row-major field and the
 * virtual-grid buffer, through the maps above. */
void zo_natural_to_real(const zo_shape* s, const double* natural, double* grid, int ncomp);
void zo_real_to_natural(const zo_shape* s, const double* grid, double* natural, int ncomp);
void zo_reciprocal_to_natural(const zo_shape* s, const double* grid, double* natural, int ncomp);
void zo_natural_to_reciprocal(const zo_shape* s, const double* natural, double* grid, int ncomp);

/* init_host / init_host_3d ramp of every rank (zmodel/utils.cu:9-81). */
void zo_init_host(const zo_shape* s, double* grid, int ncomp);

/* Twiddle table of one rank along one axis, the full L^2 layout of
 * init_forward_twiddles / init_inverse_twiddles (zmodel/transforms/fft.cu:37-73).
 * dir = -1 forward, +1 inverse.  index_mode != 0 writes the (k, l) integer
 * pair instead, as __PRINT__TWIDDLES__ does (fft.cu:45-46, 64-65). */
void zo_twiddle_table(const zo_shape* s, int id, int dir, int index_mode, double* out);

/* ---- FFT stages (one call covers every rank of the virtual grid) ---- */
#define ZO_FORWARD (-1)
#define ZO_INVERSE (+1)

/* FFT arithmetic used by zo_fft1a / zo_fft1b: 0 = direct O(n^2) DFT with
 * long-double accumulation (the definition), 1 = iterative radix-2 FFT. */
void zo_set_dft_mode(int mode);
int zo_get_dft_mode(void);
/* Unnormalised length-n DFT of `count` sequences: element i of sequence t is
 * at in[t*dist + i*stride].  Uses the current DFT mode. */
void zo_dft_many(const double* in, double* out, int n, int64_t count, int64_t dist, int64_t stride, int dir);

void zo_fft1a(const zo_shape* s, const double* in, double* out, int dir);
void zo_fft1b(const zo_shape* s, const double* in, double* out, int dir);
void zo_pack_forward(const zo_shape* s, const double* in, double* out);
void zo_pack_inverse(const zo_shape* s, const double* in, double* out, double scale_re, double scale_im,
                     int do_compute);
/* axis 0 = row_comm (cid varies), 1 = col_comm (rid varies), 2 = dep_comm (did varies) */
void zo_alltoall(const zo_shape* s, int axis, const double* send, double* recv);
void zo_unpack_forward(const zo_shape* s, int axis, const double* in, double* out, int do_compute);
void zo_unpack_inverse(const zo_shape* s, int axis, const double* in, double* out, int do_compute);
void zo_repack_transpose(const zo_shape* s, const double* in, double* out);
void zo_inverse_repack_transpose(const zo_shape* s, const double* in, double* out);
void zo_repack_transpose_xy_3d(const zo_shape* s, const double* in, double* out);
void zo_repack_transpose_yz_3d(const zo_shape* s, const double* in, double* out);
void zo_inverse_repack_transpose_xy_3d(const zo_shape* s, const double* in, double* out);
void zo_inverse_repack_transpose_yz_3d(const zo_shape* s, const double* in, double* out);

/* Composite transforms.  `stages`, when not NULL, receives a copy of the
 * grid after every reference stage (11 for 2D, 17 for 3D, in the order the
 * reference runs them); each entry must hold s->total complex values. */
#define ZO_STAGES_2D 11
#define ZO_STAGES_3D 17
void zo_forward_fft(const zo_shape* s, const double* in, double* out, int do_compute, double** stages);
void zo_inverse_fft(const zo_shape* s, const double* in, double* out, double scale_re, double scale_im,
                    int do_compute, double** stages);

/* ---- stencils ---- */
/* 5-point (2D) or 7-point (3D) Laplacian with the reference's halo exchange,
 * unscaled (zmodel/stencils/laplace.cu:73-256, laplace3d.cu:90-361). */
void zo_laplace(const zo_shape* s, const double* in, double* out, int ncomp);
/* 4th-order x / y derivatives (zmodel/stencils/dx.cu:10-161, dy.cu:10-185). 2D only, B >= 2. */
int zo_dx(const zo_shape* s, const double* in, double* out, double delta, int ncomp);
int zo_dy(const zo_shape* s, const double* in, double* out, double delta, int ncomp);

#ifdef __cplusplus
}
#endif
#endif
