/*
 * stencilfft.h -- C ABI of the MI355X-native interleaved block-cyclic FFT and
 * stencil library (libstencilfft.so).
 *
 * One GPU holds the whole P x P (2D) or P x P x P (3D) virtual rank grid of the
 * reference's zmodel path.  A "field" is one device buffer of N^d complex
 * values: the P^d rank-local buffers of L^d = (N/P)^d elements concatenated in
 * world-rank order, rank id = rid*P + cid (2D) or did*P*P + rid*P + cid (3D).
 * Inside a rank buffer the layout is exactly the reference's:
 *   - real space: the vector-packed block-cyclic layout of zmodel/utils.h:99-128
 *     (row lr, packed column lc = j*VEC + lane);
 *   - reciprocal space: the layout forward_fft leaves in device_buf1, indexed
 *     by reciprocal_to_global (zmodel/utils.h:254-320).
 * Complex values are interleaved (re, im): double2 for ZM_FP64, float2 for
 * ZM_FP32 / ZM_FP32_MFMA.  A batch of fields is contiguous, field stride N^d.
 *
 * Conventions (no counterpart in the reference, which exits on error):
 *   - every entry point returns a zm_status; nothing calls exit();
 *   - device pointers are plain pointers; `stream` is a hipStream_t passed as
 *     void* (NULL = the default stream); calls are stream-ordered and
 *     asynchronous, nothing synchronises the device;
 *   - in and out must not alias unless an entry point says otherwise;
 *   - a plan is immutable after creation; concurrent calls on one plan must
 *     use distinct scratch buffers.  Calls that pass scratch = NULL share the
 *     buffer the plan allocated at creation, so they must be ordered on one
 *     stream (or otherwise serialised by the caller).
 */
#ifndef STENCILFFT_H
#define STENCILFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ZM_OK = 0,
    ZM_ERR_INVALID_ARG = 1, /* null pointer, unknown enum, bad batch */
    ZM_ERR_SHAPE = 2,       /* N, P, B violate P | N, B | N/P, P | VEC (zmodel/utils.h:47-60) */
    ZM_ERR_UNSUPPORTED = 3, /* valid for the reference, not built here (non power-of-two sizes, B < 2 for dx/dy) */
    ZM_ERR_DEVICE = 4,      /* no HIP device, or a HIP runtime call failed */
    ZM_ERR_ALIAS = 5,       /* in and out alias where that is not allowed */
    ZM_ERR_LAUNCH = 6,      /* a kernel launch failed */
    ZM_ERR_NOMEM = 7,       /* device allocation failed */
    ZM_ERR_IO = 8,          /* file I/O failed (stencilfft_hdf5.h) */
    ZM_ERR_SCRATCH_LIMIT = 9 /* the plan needs more scratch than ZM_NOALLOC_SCRATCH_BYTES (zm_forward_noalloc) */
} zm_status;

typedef enum {
    ZM_FP64 = 0,       /* the reference's precision (cufftDoubleComplex, zmodel/utils.h:9) */
    ZM_FP32 = 1,       /* FP32 butterflies on the vector ALU, FP64-generated twiddles */
    ZM_FP32_MFMA = 2   /* FP32 data, DFT stages as split-precision (hi/lo FP16) MFMA GEMMs */
} zm_precision;

/* plan flags */
#define ZM_FLAG_TWO_STAGE 1u /* per direction run FFT1a+pack+exchange, then unpack+twiddle+FFT1b,
                                materialising the reference's post-exchange buffer */
#define ZM_FLAG_CALLER_SCRATCH 2u /* the plan allocates no scratch of its own: every transform call
                                     passes one (scratch = NULL returns ZM_ERR_INVALID_ARG) */

/* The shape a plan is built for: zmodel/utils.h:28-71 derives everything else. */
typedef struct {
    int dims; /* 2 or 3 */
    int n;    /* N_DIM: global edge */
    int p;    /* P_DIM: ranks per edge */
    int b;    /* B_DIM: block edge */
} zm_shape;

/* Derived sizes (zmodel/utils.h:47-71, zmodel/transforms/fft_host_plans.h:43-50). */
typedef struct {
    int l;            /* LOCAL_DIM = N/P */
    int vec;          /* VEC = L/B, the FFT1a length */
    int vp;           /* VEC/P, the FFT1b stride */
    int m;            /* B*P, the FFT1b length */
    int nranks;       /* P^dims virtual ranks */
    int64_t local;    /* L^dims elements per rank */
    int64_t field;    /* N^dims elements per field */
} zm_dims;

typedef struct zm_plan zm_plan;

/* ---------------------------------------------------------------------- */
/* shapes and plans                                                        */
/* ---------------------------------------------------------------------- */

/* Host-only validation of a shape; fills `out` when not NULL.  Replaces the
 * compile-time -DN_DIM/-DB_DIM/-DP_DIM build of the reference
 * (Makefile:30-33, zmodel/utils.h:28-38). */
zm_status zm_shape_check(const zm_shape* shape, zm_dims* out);

/* Create a plan on the current HIP device: validates the shape, builds FP64
 * twiddle tables on the GPU, picks the kernel path and, unless `flags` has
 * ZM_FLAG_CALLER_SCRATCH, allocates the plan's own scratch
 * (zm_scratch_bytes).  Unknown flag bits are ZM_ERR_INVALID_ARG.  Replaces
 * init_plans / init_plans_3d (zmodel/transforms/fft_host_plans.h:14-61,
 * 75-121), init_fftdx_fft1 (fft_device_plans.h:37-46) and the twiddle half of
 * init_buffers / init_buffers_3d (zmodel/transforms/fft.cu:75-108,
 * fft3d.cu:70-111).  `batch` fields are transformed per call. */
zm_status zm_plan_create(zm_plan** plan, const zm_shape* shape, zm_precision precision, int batch, unsigned flags);

/* Replaces destroy_plans / destroy_buffers (fft_host_plans.h:63-71, fft.cu:110-129). */
zm_status zm_plan_destroy(zm_plan* plan);

zm_status zm_plan_dims(const zm_plan* plan, zm_shape* shape, zm_dims* dims, zm_precision* precision, int* batch);

/* The kernel path a plan runs (no reference counterpart): a bitwise OR of
 * the flags below. */
#define ZM_PATH_TWO_STAGE 1u   /* two kernels per direction (post-exchange buffer materialised) */
#define ZM_PATH_FAST_LINE 2u   /* compile-time specialised fused line pass */
#define ZM_PATH_MFMA 4u        /* radix-16 stages as split-precision MFMA GEMMs */
zm_status zm_plan_path(const zm_plan* plan, unsigned* path);

/* Bytes of device scratch the transforms need (0 when none). */
size_t zm_scratch_bytes(const zm_plan* plan);

/* Bytes of one batch of fields (batch * N^d * element size). */
size_t zm_field_bytes(const zm_plan* plan);

/* ---------------------------------------------------------------------- */
/* transforms                                                              */
/* ---------------------------------------------------------------------- */

/* Forward 2D/3D FFT, real-space packed layout -> reciprocal layout.
 * Replaces forward_fft<true> (zmodel/transforms/fft.h:36-38, fft.cu:199-339)
 * and forward_fft_3d<true> (fft.h:54-56, fft3d.cu:234-412).  Unlike the
 * reference, `in` is not clobbered.  `scratch` holds zm_scratch_bytes(plan)
 * bytes and overlaps neither `in` nor `out` (only pointer equality is
 * checked here), or is NULL to use the buffer the plan allocated at creation
 * (ZM_ERR_INVALID_ARG for a ZM_FLAG_CALLER_SCRATCH plan). */
zm_status zm_forward(zm_plan* plan, const void* in, void* out, void* scratch, void* stream);

/* Forward transform of `scale` times the periodic 5-point (2D) / 7-point (3D)
 * Laplacian of `in`: the reference's laplace followed by forward_fft
 * (zmodel/main.cu:50-63, laplace.cu:73-169 / laplace3d.cu:90-225, fft.cu:
 * 199-339), bit for bit zm_forward(zm_laplace(in)) with `scale`.  On FP32
 * plans with the specialised path (2D: N >= 4096; 3D: lines of N <= 1024
 * points whose row groups lie in one B-block) the Laplacian is formed inside
 * the x pass, one sweep of the field less; elsewhere it is the two calls.
 * ZM_PATH_TWO_STAGE plans get ZM_ERR_UNSUPPORTED.  Buffers as zm_forward. */
zm_status zm_laplace_forward(zm_plan* plan, const void* in, void* out, double scale, void* scratch, void* stream);

/* Scratch bytes available to zm_forward_noalloc on each GPU. */
#define ZM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* zm_forward without any device allocation: the same kernels, with the
 * library's static device scratch (ZM_NOALLOC_SCRATCH_BYTES on each GPU the
 * library runs on, held from the first zm_plan_create on that GPU until the
 * process ends) in place of a caller's or the plan's buffer.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc); it calls no hipMalloc, hipFree or allocating library routine,
 * and records no per-launch timing events.  A plan whose zm_scratch_bytes
 * exceeds ZM_NOALLOC_SCRATCH_BYTES gets ZM_ERR_SCRATCH_LIMIT before anything
 * is launched (2D complex64 up to N = 2048 fits, complex128 up to N = 2048,
 * 3D complex64 up to N = 128).  Every call on one GPU shares that GPU's
 * buffer, which is fully overwritten before it is read (nothing in it needs
 * clearing) and keeps the last call's contents: calls that overlap on one GPU
 * (two streams, two threads) race, so order them on one stream. */
zm_status zm_forward_noalloc(zm_plan* plan, const void* in, void* out, void* stream);

/* Inverse (unnormalised inverse DFTs, `scale` applied once), reciprocal ->
 * real space.  Replaces inverse_fft<true> (fft.h:40-42, fft.cu:409-514) and
 * inverse_fft_3d<true> (fft.h:58-60, fft3d.cu:508-633), with a real scale
 * (the reference passes Complex{s, 0}, zmodel/main.cu:77-78). */
zm_status zm_inverse(zm_plan* plan, const void* in, void* out, double scale, void* scratch, void* stream);

/* The do_compute=false pipelines: every pack / exchange / unpack / repack of
 * forward_fft<false> / inverse_fft<false> with the FFTs and twiddles left out
 * (fft.cu:131-134, 176-180, 365-369).  Bit-exact data movement. */
zm_status zm_permute_forward(zm_plan* plan, const void* in, void* out, void* scratch, void* stream);
zm_status zm_permute_inverse(zm_plan* plan, const void* in, void* out, void* scratch, void* stream);

/* Adjoints.  forward^H = inverse with scale 1; (inverse with scale s)^H =
 * s * forward (every stage is a permutation, a unit-modulus diagonal or a
 * batch of unnormalised DFTs).  No reference counterpart: they back the
 * gradient checks of the differentiable solves. */
zm_status zm_forward_adjoint(zm_plan* plan, const void* in, void* out, void* scratch, void* stream);
zm_status zm_inverse_adjoint(zm_plan* plan, const void* in, void* out, double scale, void* scratch, void* stream);

/* ---------------------------------------------------------------------- */
/* stencils (real-space packed layout)                                     */
/* ---------------------------------------------------------------------- */

typedef enum { ZM_COMPLEX = 0, ZM_REAL = 1 } zm_element;

/* Periodic 5-point (2D) or 7-point (3D) Laplacian, sum of neighbours minus
 * 2d * centre, multiplied by `scale` (1.0 gives the reference's unscaled
 * result; N*N gives h = 1/N).  Halos are read straight from the neighbour
 * ranks' slabs.  Replaces pack_laplace + communicate_laplace + laplace
 * (zmodel/stencils/laplace.h:34-38, laplace.cu:73-256) and their 3D forms
 * (laplace3d.cu:90-361).  The Laplacian is self-adjoint. */
zm_status zm_laplace(zm_plan* plan, const void* in, void* out, double scale, zm_element element, void* stream);

/* Periodic 4th-order central differences [1, -8, 0, 8, -1] / (12 delta) along
 * x (packed columns) or y (block-cyclic rows); B >= 2.  Replaces dx / dy
 * with their halo packers (zmodel/stencils/dx.cu:10-161, dy.cu:10-185, which
 * are 2D); on a 3D plan they act on every plane gd.  The complex form applies
 * the real stencil to re and im. */
zm_status zm_dx4(zm_plan* plan, const void* in, void* out, double delta, zm_element element, void* stream);
zm_status zm_dy4(zm_plan* plan, const void* in, void* out, double delta, zm_element element, void* stream);

/* The same stencil along z (block-cyclic depth); 3D plans only, B >= 2
 * (ZM_ERR_UNSUPPORTED otherwise).  No reference counterpart: the reference
 * has dx and dy only; its depth axis is distributed like its rows
 * (zmodel/utils.h:193-199). */
zm_status zm_dz4(zm_plan* plan, const void* in, void* out, double delta, zm_element element, void* stream);

/* Transposed stencils: dx4^T = -dx4, dy4^T = -dy4, dz4^T = -dz4 (no
 * reference counterpart). */
zm_status zm_dx4_adjoint(zm_plan* plan, const void* in, void* out, double delta, zm_element element, void* stream);
zm_status zm_dy4_adjoint(zm_plan* plan, const void* in, void* out, double delta, zm_element element, void* stream);
zm_status zm_dz4_adjoint(zm_plan* plan, const void* in, void* out, double delta, zm_element element, void* stream);

/* ---------------------------------------------------------------------- */
/* reciprocal-space multipliers (no reference counterpart; built on        */
/* reciprocal_to_global, zmodel/utils.h:254-320)                           */
/* ---------------------------------------------------------------------- */

typedef enum {
    ZM_SPECTRAL_POISSON = 0,        /* 1 / (-|k|^2), k = 0 mode set to 0 */
    ZM_SPECTRAL_HEAT = 1,           /* exp(-nu |k|^2 dt): p0 = nu, p1 = dt */
    ZM_SPECTRAL_SCALE = 2,          /* constant p0 */
    ZM_SPECTRAL_STENCIL = 3,        /* exact symbol of zm_laplace(scale = p0) + zm_dx4(delta = p1);
                                       p1 = 0 drops the dx4 term (which is 2D only) */
    ZM_SPECTRAL_POISSON_STENCIL = 4 /* POISSON x STENCIL: the stencils applied to the Poisson solution */
} zm_spectral;

/* out = multiplier(k) * in on the reciprocal layout; |k|^2 = (2 pi / length)^2
 * * sum of signed integer wavenumbers squared.  The stencil symbols use
 * theta = 2 pi k / N per axis: Laplacian p0 * sum (2 cos theta - 2), dx4
 * i (8 sin theta_x - sin 2 theta_x) / (6 p1), theta_x the packed-column axis.
 * The multiplier is evaluated in FP64 for every precision; in == out is
 * allowed.  POISSON, HEAT and SCALE are real (self-adjoint); the _adjoint
 * form applies the complex conjugate multiplier. */
zm_status zm_spectral_multiply(zm_plan* plan, const void* in, void* out, zm_spectral kind, double p0, double p1,
                               double length, void* stream);
zm_status zm_spectral_multiply_adjoint(zm_plan* plan, const void* in, void* out, zm_spectral kind, double p0,
                                       double p1, double length, void* stream);

/* ---------------------------------------------------------------------- */
/* index maps (host, pure, thread-safe): zmodel/utils.h:99-320             */
/* ---------------------------------------------------------------------- */

/* zmodel/utils.h:122-128 */
int zm_packed_to_block_cyclic(const zm_shape* s, int packed_col);
int zm_block_cyclic_to_packed(const zm_shape* s, int block_cyclic_col);
/* zmodel/utils.h:166-214 */
void zm_real_to_global(const zm_shape* s, int rid, int cid, int lr, int lc, int* gr, int* gc);
void zm_global_to_real(const zm_shape* s, int gr, int gc, int* rid, int* cid, int* lr, int* lc);
void zm_real_to_global_3d(const zm_shape* s, int rid, int cid, int did, int ld, int lr, int lc, int* gd, int* gr,
                          int* gc);
void zm_global_to_real_3d(const zm_shape* s, int gd, int gr, int gc, int* rid, int* cid, int* did, int* ld, int* lr,
                          int* lc);
/* zmodel/utils.h:232-252 */
int zm_lp(const zm_shape* s, int i);
int zm_ct(const zm_shape* s, int p);
int zm_lp_inv(const zm_shape* s, int j);
int zm_ct_inv(const zm_shape* s, int k);
/* zmodel/utils.h:254-320 */
void zm_reciprocal_to_global(const zm_shape* s, int rid, int cid, int lr, int lc, int* ky, int* kx);
void zm_global_to_reciprocal(const zm_shape* s, int ky, int kx, int* rid, int* cid, int* lr, int* lc);
void zm_reciprocal_to_global_3d(const zm_shape* s, int rid, int cid, int did, int ld, int lr, int lc, int* kd,
                                int* kr, int* kc);
void zm_global_to_reciprocal_3d(const zm_shape* s, int kd, int kr, int kc, int* rid, int* cid, int* did, int* ld,
                                int* lr, int* lc);

/* init_host / init_host_3d (zmodel/utils.cu:9-81) for every rank of the
 * virtual grid: host buffer of N^d elements of `ncomp` doubles (2 = Complex,
 * 1 = Scalar) receives the global linear index of each point in the packed
 * layout (imaginary part 0).  Host only. */
zm_status zm_init_host(const zm_shape* s, double* host_grid, int ncomp);

/* Whole-field index tables on the device: for every element of a field the
 * global linear index of its real-space point (gr*N+gc / (gd*N+gr)*N+gc) or
 * of its frequency (ky*N+kx / (kd*N+kr)*N+kc).  int64 output, one per
 * element.  Device forms of the maps above. */
zm_status zm_real_index_table(zm_plan* plan, int64_t* out, void* stream);
zm_status zm_reciprocal_index_table(zm_plan* plan, int64_t* out, void* stream);

/* Device vector_pack / vector_unpack of whole fields between the natural
 * N^d row-major layout and the real-space packed layout (zmodel/utils.h:99-119
 * composed with the block-cyclic distribution of zmodel/utils.cu:9-36). */
zm_status zm_natural_to_packed(zm_plan* plan, const void* natural, void* packed, void* stream);
zm_status zm_packed_to_natural(zm_plan* plan, const void* packed, void* natural, void* stream);

/* ---------------------------------------------------------------------- */
/* per-stage timing                                                        */
/* ---------------------------------------------------------------------- */

/* Stage kinds timed by zm_timing_*.  Replaces the per-stage wall-clock
 * timers of __PRINT__DETAILED__TIMING__ (TIME_MPI_MAX_NS,
 * zmodel/transforms/fft.cu:8-21, 207-328) with HIP events recorded after
 * every kernel launch on the caller's stream. */
typedef enum {
    ZM_STAGE_LINE = 0,     /* fused FFT1a + exchange + twiddle + FFT1b pass */
    ZM_STAGE_A = 1,        /* FFT1a + pack + exchange */
    ZM_STAGE_B = 2,        /* unpack + twiddle + FFT1b */
    ZM_STAGE_REPACK = 3,   /* repack-transpose */
    ZM_STAGE_LAPLACE = 4,
    ZM_STAGE_DERIV = 5,    /* dx4 / dy4 */
    ZM_STAGE_SPECTRAL = 6,
    ZM_STAGE_CONVERT = 7,  /* layout conversion and index tables */
    ZM_STAGE_LINE_LAPLACE = 8, /* zm_laplace_forward's x pass with the Laplacian formed in it */
    ZM_STAGE_COUNT = 9
} zm_stage;

/* enable != 0 starts recording one event after every subsequent kernel launch
 * of this plan on its stream; a launch's time runs from the previous timed
 * launch's event on the same stream (an event recorded just before it for
 * the first) to its own.  Not thread-safe; meant for benchmarks.  0 stops. */
zm_status zm_timing_enable(zm_plan* plan, int enable);
/* Waits for the recorded events and returns the number of launches of
 * `stage` and their summed duration in milliseconds. */
zm_status zm_timing_read(zm_plan* plan, zm_stage stage, int* count, double* total_ms);
/* Drops every recorded event. */
zm_status zm_timing_reset(zm_plan* plan);

const char* zm_strerror(zm_status status);
const char* zm_version(void);

#ifdef __cplusplus
}
#endif
#endif /* STENCILFFT_H */
