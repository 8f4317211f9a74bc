/*
 * oracle.c -- CPU FP64 restatement of StencilFFT's zmodel hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product.  Each function cites the reference file:line it restates.  The
 * reference's MPI ranks are the slabs of one virtual-grid array (see oracle.h).
 *
 * Pinning (what checks each part independently of this file):
 *   index maps ............ closed forms + bijection tests + images/step1-4.png
 *   pack/alltoall/unpack/
 *   repack (do_compute=0) . images/step1-4.png transcribed layouts (tests/golden)
 *   twiddle table ......... (k, l) index-pair mode + exp closed form
 *   fft1a / fft1b ......... DFT definition, numpy.fft
 *   forward/inverse ....... numpy.fft.fft2/fftn through reciprocal_to_global,
 *                           closed-form DFT of the init_host ramp, round trip
 *   laplace / dx / dy ..... periodic np.roll stencils on the natural layout,
 *                           plane-wave eigenvalues
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

static int ipow(int a, int e) {
    int r = 1;
    while (e-- > 0) r *= a;
    return r;
}

int zo_shape_init(zo_shape* s, int dims, int n, int p, int b) {
    if (!s || (dims != 2 && dims != 3) || n <= 0 || p <= 0 || b <= 0) return -1;
    if (n % p) return -1;
    int l = n / p;
    if (l % b) return -1;
    int vec = l / b;
    if (vec % p) return -1;
    s->dims = dims;
    s->n = n;
    s->p = p;
    s->b = b;
    s->l = l;
    s->vec = vec;
    s->vp = vec / p;
    s->m = b * p;
    s->nranks = ipow(p, dims);
    s->local = (int64_t)l * l * (dims == 3 ? l : 1);
    s->total = s->local * s->nranks;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* index maps                                                               */
/* ------------------------------------------------------------------------ */

/* zmodel/utils.h:122-124 */
int zo_packed_to_block_cyclic(const zo_shape* s, int packed_col) {
    return (packed_col / s->vec) + (packed_col % s->vec) * s->b;
}

/* zmodel/utils.h:126-128 */
int zo_block_cyclic_to_packed(const zo_shape* s, int block_cyclic_col) {
    return (block_cyclic_col % s->b) * s->vec + (block_cyclic_col / s->b);
}

/* zmodel/utils.h:166-170 */
void zo_real_to_global(const zo_shape* s, int rid, int cid, int lr, int lc, int* gr, int* gc) {
    int B = s->b, P = s->p;
    int bcc = zo_packed_to_block_cyclic(s, lc);
    *gr = (lr / B) * B * P + rid * B + (lr % B);
    *gc = (bcc / B) * B * P + cid * B + (bcc % B);
}

/* zmodel/utils.h:176-185 */
void zo_global_to_real(const zo_shape* s, int gr, int gc, int* rid, int* cid, int* lr, int* lc) {
    int B = s->b, P = s->p;
    *rid = (gr / B) % P;
    *lr = (gr / (B * P)) * B + (gr % B);
    *cid = (gc / B) % P;
    int bcc = (gc / (B * P)) * B + (gc % B);
    *lc = zo_block_cyclic_to_packed(s, bcc);
}

/* zmodel/utils.h:192-197 */
void zo_real_to_global_3d(const zo_shape* s, int rid, int cid, int did, int ld, int lr, int lc,
                          int* gd, int* gr, int* gc) {
    int B = s->b, P = s->p;
    int bcc = zo_packed_to_block_cyclic(s, lc);
    *gd = (ld / B) * B * P + did * B + (ld % B);
    *gr = (lr / B) * B * P + rid * B + (lr % B);
    *gc = (bcc / B) * B * P + cid * B + (bcc % B);
}

/* zmodel/utils.h:203-214 */
void zo_global_to_real_3d(const zo_shape* s, int gd, int gr, int gc, int* rid, int* cid, int* did,
                          int* ld, int* lr, int* lc) {
    int B = s->b, P = s->p;
    *did = (gd / B) % P;
    *ld = (gd / (B * P)) * B + (gd % B);
    *rid = (gr / B) % P;
    *lr = (gr / (B * P)) * B + (gr % B);
    *cid = (gc / B) % P;
    int bcc = (gc / (B * P)) * B + (gc % B);
    *lc = zo_block_cyclic_to_packed(s, bcc);
}

/* zmodel/utils.h:232-236 */
int zo_lp(const zo_shape* s, int i) { return (i / s->vp) + (i % s->vp) * s->m; }
/* zmodel/utils.h:238-241 */
int zo_ct(const zo_shape* s, int p) { return (p % s->m) * s->vec + (p / s->m); }
/* zmodel/utils.h:243-247 */
int zo_lp_inv(const zo_shape* s, int j) { return (j % s->m) * s->vp + (j / s->m); }
/* zmodel/utils.h:249-252 */
int zo_ct_inv(const zo_shape* s, int k) { return (k % s->vec) * s->m + (k / s->vec); }

/* zmodel/utils.h:265-268 */
void zo_reciprocal_to_global(const zo_shape* s, int rid, int cid, int lr, int lc, int* ky, int* kx) {
    *kx = zo_ct(s, cid * s->l + zo_lp(s, lr));
    *ky = zo_ct(s, rid * s->l + zo_lp(s, lc));
}

/* zmodel/utils.h:274-282 */
void zo_global_to_reciprocal(const zo_shape* s, int ky, int kx, int* rid, int* cid, int* lr, int* lc) {
    int px = zo_ct_inv(s, kx);
    *cid = px / s->l;
    *lr = zo_lp_inv(s, px % s->l);
    int py = zo_ct_inv(s, ky);
    *rid = py / s->l;
    *lc = zo_lp_inv(s, py % s->l);
}

/* zmodel/utils.h:298-302 */
void zo_reciprocal_to_global_3d(const zo_shape* s, int rid, int cid, int did, int ld, int lr, int lc,
                                int* kd, int* kr, int* kc) {
    *kc = zo_ct(s, cid * s->l + zo_lp(s, lr));
    *kr = zo_ct(s, rid * s->l + zo_lp(s, ld));
    *kd = zo_ct(s, did * s->l + zo_lp(s, lc));
}

/* zmodel/utils.h:308-320 */
void zo_global_to_reciprocal_3d(const zo_shape* s, int kd, int kr, int kc, int* rid, int* cid, int* did,
                                int* ld, int* lr, int* lc) {
    int pc = zo_ct_inv(s, kc);
    *cid = pc / s->l;
    *lr = zo_lp_inv(s, pc % s->l);
    int pr = zo_ct_inv(s, kr);
    *rid = pr / s->l;
    *ld = zo_lp_inv(s, pr % s->l);
    int pd = zo_ct_inv(s, kd);
    *did = pd / s->l;
    *lc = zo_lp_inv(s, pd % s->l);
}

/* zmodel/utils.h:99-108 */
void zo_vector_pack(const zo_shape* s, const double* block_cyclic, double* packed, int64_t num_rows, int ncomp) {
    const int64_t L = s->l;
    for (int64_t i = 0; i < num_rows; i++)
        for (int64_t j = 0; j < s->b; j++)
            for (int64_t jj = 0; jj < s->vec; jj++)
                for (int c = 0; c < ncomp; c++)
                    packed[(i * L + j * s->vec + jj) * ncomp + c] = block_cyclic[(i * L + j + jj * s->b) * ncomp + c];
}

/* zmodel/utils.h:110-119 */
void zo_vector_unpack(const zo_shape* s, const double* packed, double* block_cyclic, int64_t num_rows, int ncomp) {
    const int64_t L = s->l;
    for (int64_t i = 0; i < num_rows; i++)
        for (int64_t j = 0; j < s->b; j++)
            for (int64_t jj = 0; jj < s->vec; jj++)
                for (int c = 0; c < ncomp; c++)
                    block_cyclic[(i * L + j + jj * s->b) * ncomp + c] = packed[(i * L + j * s->vec + jj) * ncomp + c];
}

/* rank id -> (rid, cid, did): fft.cu:563-567 (2D), fft3d.cu:677-679 (3D) */
static void rank_coords(const zo_shape* s, int id, int* rid, int* cid, int* did) {
    int P = s->p;
    if (s->dims == 2) {
        *did = 0;
        *rid = id / P;
        *cid = id % P;
    } else {
        *did = id / (P * P);
        *rid = (id / P) % P;
        *cid = id % P;
    }
}

static int rank_id(const zo_shape* s, int rid, int cid, int did) {
    int P = s->p;
    return s->dims == 2 ? rid * P + cid : did * P * P + rid * P + cid;
}

/* natural[g] <-> grid[id*local + local_offset] through real_to_global(_3d) */
static void real_map(const zo_shape* s, const double* src, double* dst, int ncomp, int to_grid) {
    const int64_t N = s->n, L = s->l;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        int64_t ndep = s->dims == 3 ? L : 1;
        for (int64_t ld = 0; ld < ndep; ld++)
            for (int lr = 0; lr < L; lr++)
                for (int lc = 0; lc < L; lc++) {
                    int64_t nat, loc = (int64_t)id * s->local + (ld * L + lr) * L + lc;
                    if (s->dims == 2) {
                        int gr, gc;
                        zo_real_to_global(s, rid, cid, lr, lc, &gr, &gc);
                        nat = gr * N + gc;
                    } else {
                        int gd, gr, gc;
                        zo_real_to_global_3d(s, rid, cid, did, (int)ld, lr, lc, &gd, &gr, &gc);
                        nat = ((int64_t)gd * N + gr) * N + gc;
                    }
                    for (int c = 0; c < ncomp; c++) {
                        if (to_grid) dst[loc * ncomp + c] = src[nat * ncomp + c];
                        else dst[nat * ncomp + c] = src[loc * ncomp + c];
                    }
                }
    }
}

void zo_natural_to_real(const zo_shape* s, const double* natural, double* grid, int ncomp) {
    real_map(s, natural, grid, ncomp, 1);
}

void zo_real_to_natural(const zo_shape* s, const double* grid, double* natural, int ncomp) {
    real_map(s, grid, natural, ncomp, 0);
}

/* natural[k] <-> grid through reciprocal_to_global(_3d); natural is A[ky][kx]
 * (2D) or A[kd][kr][kc] (3D) as zmodel/utils.h:254-262, 284-297 describe. */
static void recip_map(const zo_shape* s, const double* src, double* dst, int ncomp, int to_grid) {
    const int64_t N = s->n, L = s->l;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        int64_t ndep = s->dims == 3 ? L : 1;
        for (int64_t ld = 0; ld < ndep; ld++)
            for (int lr = 0; lr < L; lr++)
                for (int lc = 0; lc < L; lc++) {
                    int64_t nat, loc = (int64_t)id * s->local + (ld * L + lr) * L + lc;
                    if (s->dims == 2) {
                        int ky, kx;
                        zo_reciprocal_to_global(s, rid, cid, lr, lc, &ky, &kx);
                        nat = ky * N + kx;
                    } else {
                        int kd, kr, kc;
                        zo_reciprocal_to_global_3d(s, rid, cid, did, (int)ld, lr, lc, &kd, &kr, &kc);
                        nat = ((int64_t)kd * N + kr) * N + kc;
                    }
                    for (int c = 0; c < ncomp; c++) {
                        if (to_grid) dst[loc * ncomp + c] = src[nat * ncomp + c];
                        else dst[nat * ncomp + c] = src[loc * ncomp + c];
                    }
                }
    }
}

void zo_reciprocal_to_natural(const zo_shape* s, const double* grid, double* natural, int ncomp) {
    recip_map(s, grid, natural, ncomp, 0);
}

void zo_natural_to_reciprocal(const zo_shape* s, const double* natural, double* grid, int ncomp) {
    recip_map(s, natural, grid, ncomp, 1);
}

/* init_host (zmodel/utils.cu:9-36) and init_host_3d (zmodel/utils.cu:43-78) */
void zo_init_host(const zo_shape* s, double* grid, int ncomp) {
    const int64_t N = s->n, L = s->l, B = s->b, P = s->p;
    double* bc = (double*)malloc(sizeof(double) * s->local * ncomp);
    for (int id = 0; id < s->nranks; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        int64_t row_offset = cid * B;
        int64_t col_offset = rid * (N * B);
        int64_t dep_offset = s->dims == 3 ? did * (N * N * B) : 0;
        int64_t offset = row_offset + col_offset + dep_offset;
        if (s->dims == 2) {
            for (int64_t i = 0; i < L / B; i++)
                for (int64_t ii = 0; ii < B; ii++)
                    for (int64_t j = 0; j < L / B; j++)
                        for (int64_t jj = 0; jj < B; jj++) {
                            int64_t ai = (i * B * (L / B) * B) + (ii * (L / B) * B) + (j * B) + jj;
                            double real = (double)((i * N * B * P) + (ii * N) + (j * B * P) + jj + offset);
                            bc[ai * ncomp] = real;
                            if (ncomp == 2) bc[ai * ncomp + 1] = 0.0;
                        }
            zo_vector_pack(s, bc, grid + id * s->local * ncomp, L, ncomp);
        } else {
            for (int64_t i = 0; i < L / B; i++)
                for (int64_t ii = 0; ii < B; ii++)
                    for (int64_t j = 0; j < L / B; j++)
                        for (int64_t jj = 0; jj < B; jj++)
                            for (int64_t k = 0; k < L / B; k++)
                                for (int64_t kk = 0; kk < B; kk++) {
                                    int64_t ai = (i * B * L * L) + (ii * L * L) + (j * B * L) + (jj * L) + (k * B) + kk;
                                    double real = (double)((i * N * N * B * P) + (ii * N * N) + (j * N * B * P) +
                                                           (jj * N) + (k * B * P) + kk + offset);
                                    bc[ai * ncomp] = real;
                                    if (ncomp == 2) bc[ai * ncomp + 1] = 0.0;
                                }
            zo_vector_pack(s, bc, grid + id * s->local * ncomp, L * L, ncomp);
        }
    }
    free(bc);
}

/* Twiddle value at column `col` of a row of the table built by
 * init_forward_twiddles / init_inverse_twiddles (zmodel/transforms/fft.cu:37-73,
 * fft3d.cu:36-68): index local_row*L + fft_row + fft_col*VP holds
 * exp(dir * 2*pi*i * k*l / N) with k = fft_row + id*VP, l = fft_col.  The
 * expression order follows fft.cu:48 (-2*M_PI*k*l/N_DIM). */
static void twiddle_at(const zo_shape* s, int id, int dir, int col, double* re, double* im) {
    int fft_row = col % s->vp;
    int fft_col = col / s->vp;
    double k = (double)(fft_row + id * s->vp);
    double l = (double)fft_col;
    double ang = (dir < 0) ? (-2 * M_PI * k * l / s->n) : (2 * M_PI * k * l / s->n);
    *re = cos(ang);
    *im = sin(ang);
}

void zo_twiddle_table(const zo_shape* s, int id, int dir, int index_mode, double* out) {
    const int L = s->l, VP = s->vp;
    for (int local_row = 0; local_row < L; local_row++)
        for (int fft_row = 0; fft_row < VP; fft_row++)
            for (int fft_col = 0; fft_col < L / VP; fft_col++) {
                int64_t idx = (int64_t)local_row * L + fft_row + fft_col * VP;
                if (index_mode) {
                    out[2 * idx] = (double)(fft_row + id * VP);
                    out[2 * idx + 1] = (double)fft_col;
                } else {
                    twiddle_at(s, id, dir, fft_row + fft_col * VP, &out[2 * idx], &out[2 * idx + 1]);
                }
            }
}

/* ------------------------------------------------------------------------ */
/* DFT arithmetic (stands for cuFFT plan0 / plan1)                          */
/* ------------------------------------------------------------------------ */

static int g_dft_mode = 0;

void zo_set_dft_mode(int mode) { g_dft_mode = mode; }
int zo_get_dft_mode(void) { return g_dft_mode; }

/* Direct DFT, the definition: X[k] = sum_j x[j] exp(dir*2*pi*i*j*k/n),
 * unnormalised (cuFFT's CUFFT_FORWARD / CUFFT_INVERSE convention). */
static void dft_direct(const double* x, double* X, int n, int64_t stride, const long double* wr,
                       const long double* wi) {
    for (int k = 0; k < n; k++) {
        long double sr = 0.0L, si = 0.0L;
        for (int j = 0; j < n; j++) {
            int t = (int)(((int64_t)j * k) % n);
            long double xr = x[(int64_t)j * stride * 2], xi = x[(int64_t)j * stride * 2 + 1];
            sr += xr * wr[t] - xi * wi[t];
            si += xr * wi[t] + xi * wr[t];
        }
        X[2 * k] = (double)sr;
        X[2 * k + 1] = (double)si;
    }
}

/* Iterative radix-2 decimation-in-time FFT (textbook Cooley-Tukey). */
static void fft_radix2(double* a, int n, const double* wr, const double* wi) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = a[2 * i], ti = a[2 * i + 1];
            a[2 * i] = a[2 * j];
            a[2 * i + 1] = a[2 * j + 1];
            a[2 * j] = tr;
            a[2 * j + 1] = ti;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int step = n / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < len / 2; j++) {
                double w_r = wr[j * step], w_i = wi[j * step];
                double* u = a + 2 * (i + j);
                double* v = a + 2 * (i + j + len / 2);
                double vr = v[0] * w_r - v[1] * w_i;
                double vi = v[0] * w_i + v[1] * w_r;
                v[0] = u[0] - vr;
                v[1] = u[1] - vi;
                u[0] += vr;
                u[1] += vi;
            }
    }
}

void zo_dft_many(const double* in, double* out, int n, int64_t count, int64_t dist, int64_t stride, int dir) {
    int pow2 = n > 0 && (n & (n - 1)) == 0;
    int mode = (g_dft_mode == 1 && pow2) ? 1 : 0;
    long double* lwr = (long double*)malloc(sizeof(long double) * n);
    long double* lwi = (long double*)malloc(sizeof(long double) * n);
    double* wr = (double*)malloc(sizeof(double) * n);
    double* wi = (double*)malloc(sizeof(double) * n);
    for (int t = 0; t < n; t++) {
        long double ang = (long double)dir * 2.0L * 3.141592653589793238462643383279502884L * t / n;
        lwr[t] = cosl(ang);
        lwi[t] = sinl(ang);
        wr[t] = (double)lwr[t];
        wi[t] = (double)lwi[t];
    }
#pragma omp parallel
    {
        double* x = (double*)malloc(sizeof(double) * 2 * n);
        double* X = (double*)malloc(sizeof(double) * 2 * n);
#pragma omp for schedule(static)
        for (int64_t t = 0; t < count; t++) {
            const double* src = in + 2 * t * dist;
            for (int i = 0; i < n; i++) {
                x[2 * i] = src[2 * i * stride];
                x[2 * i + 1] = src[2 * i * stride + 1];
            }
            if (mode == 1) {
                fft_radix2(x, n, wr, wi);
                memcpy(X, x, sizeof(double) * 2 * n);
            } else {
                dft_direct(x, X, n, 1, lwr, lwi);
            }
            double* dst = out + 2 * t * dist;
            for (int i = 0; i < n; i++) {
                dst[2 * i * stride] = X[2 * i];
                dst[2 * i * stride + 1] = X[2 * i + 1];
            }
        }
        free(x);
        free(X);
    }
    free(lwr);
    free(lwi);
    free(wr);
    free(wi);
}

/* plan0 = FFT1a: VEC-point DFT of every contiguous packed vector
 * (zmodel/transforms/fft_host_plans.h:25-39 with batch VEC_TOTAL, and :86-100
 * with batch VEC_TOTAL*LOCAL_DIM in 3D).  Over the whole grid this is
 * total/VEC contiguous vectors. */
void zo_fft1a(const zo_shape* s, const double* in, double* out, int dir) {
    zo_dft_many(in, out, s->vec, s->total / s->vec, s->vec, 1, dir);
}

/* plan1 = FFT1b: for every row of L elements, VP interleaved M-point DFTs;
 * element n of transform t sits at row*L + t + n*VP
 * (zmodel/transforms/fft_host_plans.h:41-60; fft_device_plans.h:56-66;
 * executed per row at zmodel/transforms/fft.cu:254-258). */
void zo_fft1b(const zo_shape* s, const double* in, double* out, int dir) {
    int64_t rows = s->total / s->l;
    for (int64_t row = 0; row < rows; row++)
        zo_dft_many(in + 2 * row * s->l, out + 2 * row * s->l, s->m, s->vp, 1, s->vp, dir);
}

/* ------------------------------------------------------------------------ */
/* pack / exchange / unpack / repack                                        */
/* ------------------------------------------------------------------------ */

static void cmul(double ar, double ai, double br, double bi, double* r, double* i) {
    /* Complex * Complex, zmodel/utils.h:90-92 */
    *r = ar * br - ai * bi;
    *i = ar * bi + ai * br;
}

/* pack_forward_fft (zmodel/transforms/fft.cu:138-160) and pack_forward_fft_3d
 * (zmodel/transforms/fft3d.cu:146-169); the 2D kernel is the 3D one with a
 * single depth slice.  `forward` selects pack (1) or pack_inverse (0,
 * fft.cu:343-370, fft3d.cu:418-445). */
static void pack_common(const zo_shape* s, const double* in, double* out, int forward, double sr, double si,
                        int do_compute) {
    const int64_t L = s->l, B = s->b, VEC = s->vec, VP = s->vp, P = s->p;
    const int64_t ndep = s->dims == 3 ? L : 1;
    const int64_t vec_total = L * L / VEC;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        const double* src = in + 2 * (int64_t)id * s->local;
        double* dst = out + 2 * (int64_t)id * s->local;
        for (int64_t dep = 0; dep < ndep; dep++)
            for (int64_t blk = 0; blk < vec_total; blk++) {
                int64_t row = blk / B, col = blk % B;
                int64_t natural_offset = (row * L) + (col * VEC) + (dep * L * L);
                for (int64_t v = 0; v < VEC; v++) {
                    int64_t split = v / VP, index_in_split = v % VP;
                    int64_t slab_size = (L * L) / P;
                    int64_t slab_vec_row = L / P;
                    int64_t slab_vec_col = L / VP;
                    int64_t slab_offset = (split * slab_size * ndep) + (dep * slab_size);
                    int64_t out_row = (blk / slab_vec_col) % slab_vec_row;
                    int64_t out_col = blk % slab_vec_col;
                    int64_t packed_offset = slab_offset + (out_row * L) + (out_col * VP);
                    if (forward) {
                        dst[2 * (packed_offset + index_in_split)] = src[2 * (natural_offset + v)];
                        dst[2 * (packed_offset + index_in_split) + 1] = src[2 * (natural_offset + v) + 1];
                    } else {
                        const double* x = src + 2 * (packed_offset + index_in_split);
                        double* y = dst + 2 * (natural_offset + v);
                        if (do_compute) cmul(sr, si, x[0], x[1], &y[0], &y[1]);
                        else {
                            y[0] = x[0];
                            y[1] = x[1];
                        }
                    }
                }
            }
    }
}

void zo_pack_forward(const zo_shape* s, const double* in, double* out) { pack_common(s, in, out, 1, 1.0, 0.0, 0); }

void zo_pack_inverse(const zo_shape* s, const double* in, double* out, double scale_re, double scale_im,
                     int do_compute) {
    pack_common(s, in, out, 0, scale_re, scale_im, do_compute);
}

/* Rank id of member g of the communicator along `axis` that contains rank
 * (rid, cid, did).  Comm splits: zmodel/transforms/fft.cu:561-567 (key = world
 * id, so members are ordered by g), zmodel/transforms/fft3d.cu:685-697. */
static int member_id(const zo_shape* s, int axis, int rid, int cid, int did, int g) {
    if (axis == 0) return rank_id(s, rid, g, did);
    if (axis == 1) return rank_id(s, g, cid, did);
    return rank_id(s, rid, cid, g);
}

static int axis_coord(int axis, int rid, int cid, int did) { return axis == 0 ? cid : (axis == 1 ? rid : did); }

/* MPI_Alltoall over row_comm / col_comm / dep_comm (zmodel/transforms/fft.cu:221-241,
 * 285-305; fft3d.cu:260-270, 314-324, 368-378): chunk s of member r's receive
 * buffer is chunk r of member s's send buffer; chunk = L^d / P elements
 * (COMM_SIZE_2D fft.cu:23, COMM_SIZE_3D fft3d.cu:22). */
void zo_alltoall(const zo_shape* s, int axis, const double* send, double* recv) {
    const int64_t chunk = s->local / s->p;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        int r = axis_coord(axis, rid, cid, did);
        for (int src = 0; src < s->p; src++) {
            int sid = member_id(s, axis, rid, cid, did, src);
            memcpy(recv + 2 * ((int64_t)id * s->local + src * chunk),
                   send + 2 * ((int64_t)sid * s->local + r * chunk), sizeof(double) * 2 * chunk);
        }
    }
}

/* unpack_forward_fft (zmodel/transforms/fft.cu:162-181), unpack_forward_fft_3d
 * (fft3d.cu:173-198) and their inverses (fft.cu:372-391, fft3d.cu:449-472).
 * The twiddle table is the one init_buffers builds with id = the rank's
 * coordinate along the exchange axis (fft.cu:95-100, fft3d.cu:94-101). */
static void unpack_common(const zo_shape* s, int axis, const double* in, double* out, int do_compute, int forward) {
    const int64_t L = s->l, P = s->p;
    const int64_t ndep = s->dims == 3 ? L : 1;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        int g = axis_coord(axis, rid, cid, did);
        double* twr = (double*)malloc(sizeof(double) * L);
        double* twi = (double*)malloc(sizeof(double) * L);
        for (int c = 0; c < L; c++) twiddle_at(s, g, forward ? -1 : +1, c, &twr[c], &twi[c]);
        const double* src = in + 2 * (int64_t)id * s->local;
        double* dst = out + 2 * (int64_t)id * s->local;
        for (int64_t dep = 0; dep < ndep; dep++)
            for (int64_t bx = 0; bx < L; bx++)
                for (int64_t t = 0; t < L; t++) {
                    int64_t split_size = L / P;
                    int64_t slab_size = (L * L) / P;
                    int64_t index_in_row_split = bx % split_size;
                    int64_t index_in_vec_split = t % split_size;
                    int64_t row_split = bx / split_size;
                    int64_t vec_split = t / split_size;
                    int64_t packed = (row_split * slab_size * ndep) + (dep * slab_size) + (index_in_row_split * L) + t;
                    int64_t row = vec_split + (index_in_row_split * P);
                    int64_t col = (row_split * split_size) + index_in_vec_split;
                    int64_t natural = (dep * L * L) + (row * L) + col;
                    const double* x = forward ? src + 2 * packed : src + 2 * natural;
                    double* y = forward ? dst + 2 * natural : dst + 2 * packed;
                    if (do_compute) cmul(x[0], x[1], twr[col], twi[col], &y[0], &y[1]);
                    else {
                        y[0] = x[0];
                        y[1] = x[1];
                    }
                }
        free(twr);
        free(twi);
    }
}

void zo_unpack_forward(const zo_shape* s, int axis, const double* in, double* out, int do_compute) {
    unpack_common(s, axis, in, out, do_compute, 1);
}

void zo_unpack_inverse(const zo_shape* s, int axis, const double* in, double* out, int do_compute) {
    unpack_common(s, axis, in, out, do_compute, 0);
}

/* repack_transpose (zmodel/transforms/fft.cu:183-187) and
 * repack_transpose_xy_3d (fft3d.cu:203-209): out[d*L^2 + j*L + (i%B)*VEC + i/B]
 * = in[d*L^2 + i*L + j]; inverses fft.cu:393-397, fft3d.cu:477-483. */
static void repack_xy(const zo_shape* s, const double* in, double* out, int forward) {
    const int64_t L = s->l, B = s->b, VEC = s->vec;
    const int64_t ndep = s->dims == 3 ? L : 1;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        const double* src = in + 2 * (int64_t)id * s->local;
        double* dst = out + 2 * (int64_t)id * s->local;
        for (int64_t d = 0; d < ndep; d++)
            for (int64_t i = 0; i < L; i++)
                for (int64_t j = 0; j < L; j++) {
                    int64_t a = d * L * L + j * L + (i % B) * VEC + (i / B);
                    int64_t b = d * L * L + i * L + j;
                    if (forward) {
                        dst[2 * a] = src[2 * b];
                        dst[2 * a + 1] = src[2 * b + 1];
                    } else {
                        dst[2 * b] = src[2 * a];
                        dst[2 * b + 1] = src[2 * a + 1];
                    }
                }
    }
}

/* repack_transpose_yz_3d (zmodel/transforms/fft3d.cu:214-220):
 * out[c*L^2 + r*L + (d%B)*VEC + d/B] = in[d*L^2 + r*L + c]; inverse fft3d.cu:488-494. */
static void repack_yz(const zo_shape* s, const double* in, double* out, int forward) {
    const int64_t L = s->l, B = s->b, VEC = s->vec;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        const double* src = in + 2 * (int64_t)id * s->local;
        double* dst = out + 2 * (int64_t)id * s->local;
        for (int64_t d = 0; d < L; d++)
            for (int64_t r = 0; r < L; r++)
                for (int64_t c = 0; c < L; c++) {
                    int64_t a = c * L * L + r * L + (d % B) * VEC + (d / B);
                    int64_t b = d * L * L + r * L + c;
                    if (forward) {
                        dst[2 * a] = src[2 * b];
                        dst[2 * a + 1] = src[2 * b + 1];
                    } else {
                        dst[2 * b] = src[2 * a];
                        dst[2 * b + 1] = src[2 * a + 1];
                    }
                }
    }
}

void zo_repack_transpose(const zo_shape* s, const double* in, double* out) { repack_xy(s, in, out, 1); }
void zo_inverse_repack_transpose(const zo_shape* s, const double* in, double* out) { repack_xy(s, in, out, 0); }
void zo_repack_transpose_xy_3d(const zo_shape* s, const double* in, double* out) { repack_xy(s, in, out, 1); }
void zo_repack_transpose_yz_3d(const zo_shape* s, const double* in, double* out) { repack_yz(s, in, out, 1); }
void zo_inverse_repack_transpose_xy_3d(const zo_shape* s, const double* in, double* out) { repack_xy(s, in, out, 0); }
void zo_inverse_repack_transpose_yz_3d(const zo_shape* s, const double* in, double* out) { repack_yz(s, in, out, 0); }

/* ------------------------------------------------------------------------ */
/* composites                                                               */
/* ------------------------------------------------------------------------ */

static void copy_grid(const zo_shape* s, const double* in, double* out) {
    memcpy(out, in, sizeof(double) * 2 * s->total);
}

static void capture(const zo_shape* s, double** stages, int* k, const double* buf) {
    if (stages) copy_grid(s, buf, stages[*k]);
    (*k)++;
}

/* One direction of the forward transform: FFT1a, pack, all-to-all,
 * unpack x twiddle, FFT1b (zmodel/transforms/fft.cu:207-264).  buf0 -> buf1. */
static void forward_direction(const zo_shape* s, int axis, double* buf0, double* buf1, int do_compute,
                              double** stages, int* k) {
    if (do_compute) zo_fft1a(s, buf0, buf1, ZO_FORWARD);
    else copy_grid(s, buf0, buf1); /* placeholder, fft.cu:132-134 */
    capture(s, stages, k, buf1);
    zo_pack_forward(s, buf1, buf0);
    capture(s, stages, k, buf0);
    zo_alltoall(s, axis, buf0, buf1);
    capture(s, stages, k, buf1);
    zo_unpack_forward(s, axis, buf1, buf0, do_compute);
    capture(s, stages, k, buf0);
    if (do_compute) zo_fft1b(s, buf0, buf1, ZO_FORWARD);
    else copy_grid(s, buf0, buf1);
    capture(s, stages, k, buf1);
}

/* forward_fft (zmodel/transforms/fft.cu:199-339) and forward_fft_3d
 * (zmodel/transforms/fft3d.cu:234-412). */
void zo_forward_fft(const zo_shape* s, const double* in, double* out, int do_compute, double** stages) {
    double* buf0 = (double*)malloc(sizeof(double) * 2 * s->total);
    double* buf1 = (double*)malloc(sizeof(double) * 2 * s->total);
    int k = 0;
    copy_grid(s, in, buf0);
    forward_direction(s, 0, buf0, buf1, do_compute, stages, &k);
    if (s->dims == 2) zo_repack_transpose(s, buf1, buf0);
    else zo_repack_transpose_xy_3d(s, buf1, buf0);
    capture(s, stages, &k, buf0);
    forward_direction(s, 1, buf0, buf1, do_compute, stages, &k);
    if (s->dims == 3) {
        zo_repack_transpose_yz_3d(s, buf1, buf0);
        capture(s, stages, &k, buf0);
        forward_direction(s, 2, buf0, buf1, do_compute, stages, &k);
    }
    copy_grid(s, buf1, out);
    free(buf0);
    free(buf1);
}

/* One direction of the inverse transform: FFT1b^-1, unpack_inverse x conj
 * twiddle, all-to-all, pack_inverse x scale, FFT1a^-1
 * (zmodel/transforms/fft.cu:420-460).  buf0 -> buf1. */
static void inverse_direction(const zo_shape* s, int axis, double* buf0, double* buf1, double sr, double si,
                              int do_compute, double** stages, int* k) {
    if (do_compute) zo_fft1b(s, buf0, buf1, ZO_INVERSE);
    else copy_grid(s, buf0, buf1);
    capture(s, stages, k, buf1);
    zo_unpack_inverse(s, axis, buf1, buf0, do_compute);
    capture(s, stages, k, buf0);
    zo_alltoall(s, axis, buf0, buf1);
    capture(s, stages, k, buf1);
    zo_pack_inverse(s, buf1, buf0, sr, si, do_compute);
    capture(s, stages, k, buf0);
    if (do_compute) zo_fft1a(s, buf0, buf1, ZO_INVERSE);
    else copy_grid(s, buf0, buf1);
    capture(s, stages, k, buf1);
}

/* inverse_fft (zmodel/transforms/fft.cu:409-514) and inverse_fft_3d
 * (zmodel/transforms/fft3d.cu:508-633); the scale is applied once, in the
 * last pack_inverse (fft.cu:417-418, 498). */
void zo_inverse_fft(const zo_shape* s, const double* in, double* out, double scale_re, double scale_im,
                    int do_compute, double** stages) {
    double* buf0 = (double*)malloc(sizeof(double) * 2 * s->total);
    double* buf1 = (double*)malloc(sizeof(double) * 2 * s->total);
    int k = 0;
    copy_grid(s, in, buf0);
    if (s->dims == 3) {
        inverse_direction(s, 2, buf0, buf1, 1.0, 0.0, do_compute, stages, &k);
        zo_inverse_repack_transpose_yz_3d(s, buf1, buf0);
        capture(s, stages, &k, buf0);
    }
    inverse_direction(s, 1, buf0, buf1, 1.0, 0.0, do_compute, stages, &k);
    if (s->dims == 2) zo_inverse_repack_transpose(s, buf1, buf0);
    else zo_inverse_repack_transpose_xy_3d(s, buf1, buf0);
    capture(s, stages, &k, buf0);
    inverse_direction(s, 0, buf0, buf1, scale_re, scale_im, do_compute, stages, &k);
    copy_grid(s, buf1, out);
    free(buf0);
    free(buf1);
}

/* ------------------------------------------------------------------------ */
/* stencils                                                                 */
/* ------------------------------------------------------------------------ */

/* element copy / accumulate helpers on `nc` doubles */
#define ELT(ptr, idx) ((ptr) + (int64_t)(idx) * nc)

/* 2D Laplacian: pack_laplace (zmodel/stencils/laplace.cu:172-191),
 * communicate_laplace (laplace.cu:194-256) and laplace (laplace.cu:73-169). */
static void laplace_2d(const zo_shape* s, const double* in, double* out, int nc) {
    const int64_t L = s->l, B = s->b, VEC = s->vec, P = s->p;
    const int64_t prow = L * (L / B); /* PACKED_ROW_COMPLEX, zmodel/utils.h:50 */
    const int64_t pcol = L * VEC;     /* PACKED_COL_COMPLEX, zmodel/utils.h:52 */
    const int nr = s->nranks;
    double* send_above = (double*)malloc(sizeof(double) * nc * prow * nr);
    double* send_below = (double*)malloc(sizeof(double) * nc * prow * nr);
    double* send_left = (double*)malloc(sizeof(double) * nc * pcol * nr);
    double* send_right = (double*)malloc(sizeof(double) * nc * pcol * nr);
    /* pack_laplace<<<L, L>>> */
    for (int id = 0; id < nr; id++) {
        const double* src = in + (int64_t)id * s->local * nc;
        for (int64_t i = 0; i < L; i++)
            for (int64_t j = 0; j < L; j++) {
                if (i < L / B) {
                    memcpy(ELT(send_above + id * prow * nc, i * L + j), ELT(src, i * L * B + j), sizeof(double) * nc);
                    int64_t bottom_offset = L * (B - 1);
                    memcpy(ELT(send_below + id * prow * nc, i * L + j), ELT(src, bottom_offset + i * L * B + j),
                           sizeof(double) * nc);
                }
                if (j < VEC) {
                    memcpy(ELT(send_left + id * pcol * nc, i * VEC + j), ELT(src, i * L + j), sizeof(double) * nc);
                    int64_t right_offset = L - VEC;
                    memcpy(ELT(send_right + id * pcol * nc, i * VEC + j), ELT(src, right_offset + i * L + j),
                           sizeof(double) * nc);
                }
            }
    }
#pragma omp parallel for schedule(static)
    for (int id = 0; id < nr; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        /* communicate_laplace: recv_left <- left's send_right, recv_right <-
         * right's send_left, recv_below <- down's send_above, recv_above <-
         * up's send_below. */
        int left = (int)((cid - 1 + P) % P), right = (int)((cid + 1) % P);
        int up = (int)((rid - 1 + P) % P), down = (int)((rid + 1) % P);
        const double* recv_left = send_right + (int64_t)rank_id(s, rid, left, 0) * pcol * nc;
        const double* recv_right = send_left + (int64_t)rank_id(s, rid, right, 0) * pcol * nc;
        const double* recv_below = send_above + (int64_t)rank_id(s, down, cid, 0) * prow * nc;
        const double* recv_above = send_below + (int64_t)rank_id(s, up, cid, 0) * prow * nc;
        const double* din = in + (int64_t)id * s->local * nc;
        double* dout = out + (int64_t)id * s->local * nc;
        for (int64_t blk = 0; blk < L * B; blk++) {
            int64_t row = blk / B, col = blk % B;
            int64_t my_offset = (row * L) + (col * VEC);
            int top_of_block = (row % B) == 0, top_block = row == 0, top_p = rid == 0;
            int bottom_of_block = (row % B) == B - 1, bottom_block = row == L - 1, bottom_p = rid == P - 1;
            int rightmost_col = col == B - 1, rightmost_p = cid == P - 1;
            int leftmost_col = col == 0, leftmost_p = cid == 0;
            for (int64_t v = 0; v < VEC; v++)
                for (int c = 0; c < nc; c++) {
                    double scratch = 0.0;
                    int64_t off;
                    if (top_of_block) {
                        if (top_p) off = top_block ? ((L / B - 1) * L) + (col * VEC) : ((row / B - 1) * L) + (col * VEC);
                        else off = ((row / B) * L) + (col * VEC);
                        scratch += ELT(recv_above, off + v)[c];
                    } else {
                        off = ((row - 1) * L) + (col * VEC);
                        scratch += ELT(din, off + v)[c];
                    }
                    if (bottom_of_block) {
                        if (bottom_p) off = bottom_block ? (col * VEC) : ((row / B + 1) * L) + (col * VEC);
                        else off = ((row / B) * L) + (col * VEC);
                        scratch += ELT(recv_below, off + v)[c];
                    } else {
                        off = ((row + 1) * L) + (col * VEC);
                        scratch += ELT(din, off + v)[c];
                    }
                    if (leftmost_col) {
                        off = row * VEC;
                        int64_t nv = leftmost_p ? (v - 1 + VEC) % VEC : v;
                        scratch += ELT(recv_left, off + nv)[c];
                    } else {
                        off = (row * L) + ((col - 1) * VEC);
                        scratch += ELT(din, off + v)[c];
                    }
                    if (rightmost_col) {
                        off = row * VEC;
                        int64_t nv = rightmost_p ? (v + 1) % VEC : v;
                        scratch += ELT(recv_right, off + nv)[c];
                    } else {
                        off = (row * L) + ((col + 1) * VEC);
                        scratch += ELT(din, off + v)[c];
                    }
                    scratch -= 4 * ELT(din, my_offset + v)[c];
                    ELT(dout, my_offset + v)[c] = scratch;
                }
        }
    }
    free(send_above);
    free(send_below);
    free(send_left);
    free(send_right);
}

/* 3D Laplacian: pack_laplace_3d (zmodel/stencils/laplace3d.cu:228-262),
 * communicate_laplace_3d (laplace3d.cu:265-361), laplace_3d (laplace3d.cu:90-225). */
static void laplace_3d(const zo_shape* s, const double* in, double* out, int nc) {
    const int64_t L = s->l, B = s->b, VEC = s->vec, P = s->p;
    const int64_t prow = L * (L / B);   /* PACKED_ROW_COMPLEX */
    const int64_t pcol = L * VEC;       /* PACKED_COL_COMPLEX */
    const int64_t face = L * prow;      /* PACKED_FACE_COMPLEX_3D, zmodel/utils.h:70 */
    const int nr = s->nranks;
    double* sa = (double*)calloc((size_t)(nc * face * nr), sizeof(double));
    double* sb = (double*)calloc((size_t)(nc * face * nr), sizeof(double));
    double* sl = (double*)calloc((size_t)(nc * face * nr), sizeof(double));
    double* sr = (double*)calloc((size_t)(nc * face * nr), sizeof(double));
    double* sf = (double*)calloc((size_t)(nc * face * nr), sizeof(double));
    double* sk = (double*)calloc((size_t)(nc * face * nr), sizeof(double));
#pragma omp parallel for schedule(static)
    for (int id = 0; id < nr; id++) {
        const double* src = in + (int64_t)id * s->local * nc;
        double *a = sa + id * face * nc, *b = sb + id * face * nc, *l = sl + id * face * nc;
        double *r = sr + id * face * nc, *f = sf + id * face * nc, *k = sk + id * face * nc;
        for (int64_t dep = 0; dep < L; dep++)
            for (int64_t i = 0; i < L; i++)
                for (int64_t j = 0; j < L; j++) {
                    if (i < L / B) {
                        memcpy(ELT(a, dep * prow + i * L + j), ELT(src, dep * L * L + i * L * B + j), sizeof(double) * nc);
                        int64_t bottom_offset = L * (B - 1);
                        memcpy(ELT(b, dep * prow + i * L + j), ELT(src, dep * L * L + bottom_offset + i * L * B + j),
                               sizeof(double) * nc);
                    }
                    if (j < VEC) {
                        memcpy(ELT(l, dep * pcol + i * VEC + j), ELT(src, dep * L * L + i * L + j), sizeof(double) * nc);
                        int64_t right_offset = L - VEC;
                        memcpy(ELT(r, dep * pcol + i * VEC + j), ELT(src, dep * L * L + right_offset + i * L + j),
                               sizeof(double) * nc);
                    }
                    if (dep < L / B) {
                        memcpy(ELT(f, dep * L * L + i * L + j), ELT(src, dep * L * L * B + i * L + j), sizeof(double) * nc);
                        int64_t back_depth_offset = L * L * (B - 1);
                        memcpy(ELT(k, dep * L * L + i * L + j), ELT(src, back_depth_offset + dep * L * L * B + i * L + j),
                               sizeof(double) * nc);
                    }
                }
    }
#pragma omp parallel for schedule(static)
    for (int id = 0; id < nr; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        int left = (int)((cid - 1 + P) % P), right = (int)((cid + 1) % P);
        int up = (int)((rid - 1 + P) % P), down = (int)((rid + 1) % P);
        int front = (int)((did - 1 + P) % P), back = (int)((did + 1) % P);
        const double* recv_left = sr + (int64_t)rank_id(s, rid, left, did) * face * nc;
        const double* recv_right = sl + (int64_t)rank_id(s, rid, right, did) * face * nc;
        const double* recv_below = sa + (int64_t)rank_id(s, down, cid, did) * face * nc;
        const double* recv_above = sb + (int64_t)rank_id(s, up, cid, did) * face * nc;
        const double* recv_front = sk + (int64_t)rank_id(s, rid, cid, front) * face * nc;
        const double* recv_back = sf + (int64_t)rank_id(s, rid, cid, back) * face * nc;
        const double* din = in + (int64_t)id * s->local * nc;
        double* dout = out + (int64_t)id * s->local * nc;
        for (int64_t dep = 0; dep < L; dep++)
            for (int64_t blk = 0; blk < L * B; blk++) {
                int64_t row = blk / B, col = blk % B;
                int64_t my_offset = (dep * L * L) + (row * L) + (col * VEC);
                int top_of_block = (row % B) == 0, top_block = row == 0, top_p = rid == 0;
                int bottom_of_block = (row % B) == B - 1, bottom_block = row == L - 1, bottom_p = rid == P - 1;
                int rightmost_col = col == B - 1, rightmost_p = cid == P - 1;
                int leftmost_col = col == 0, leftmost_p = cid == 0;
                int front_of_dblock = (dep % B) == 0, front_dblock = dep == 0, front_p = did == 0;
                int back_of_dblock = (dep % B) == B - 1, back_dblock = dep == L - 1, back_p = did == P - 1;
                for (int64_t v = 0; v < VEC; v++)
                    for (int c = 0; c < nc; c++) {
                        double scratch = 0.0;
                        int64_t off;
                        if (top_of_block) {
                            if (top_p)
                                off = top_block ? (dep * prow) + ((L / B - 1) * L) + (col * VEC)
                                                : (dep * prow) + ((row / B - 1) * L) + (col * VEC);
                            else off = (dep * prow) + ((row / B) * L) + (col * VEC);
                            scratch += ELT(recv_above, off + v)[c];
                        } else {
                            off = (dep * L * L) + ((row - 1) * L) + (col * VEC);
                            scratch += ELT(din, off + v)[c];
                        }
                        if (bottom_of_block) {
                            if (bottom_p)
                                off = bottom_block ? (dep * prow) + (col * VEC)
                                                   : (dep * prow) + ((row / B + 1) * L) + (col * VEC);
                            else off = (dep * prow) + ((row / B) * L) + (col * VEC);
                            scratch += ELT(recv_below, off + v)[c];
                        } else {
                            off = (dep * L * L) + ((row + 1) * L) + (col * VEC);
                            scratch += ELT(din, off + v)[c];
                        }
                        if (leftmost_col) {
                            off = (dep * pcol) + (row * VEC);
                            int64_t nv = leftmost_p ? (v - 1 + VEC) % VEC : v;
                            scratch += ELT(recv_left, off + nv)[c];
                        } else {
                            off = (dep * L * L) + (row * L) + ((col - 1) * VEC);
                            scratch += ELT(din, off + v)[c];
                        }
                        if (rightmost_col) {
                            off = (dep * pcol) + (row * VEC);
                            int64_t nv = rightmost_p ? (v + 1) % VEC : v;
                            scratch += ELT(recv_right, off + nv)[c];
                        } else {
                            off = (dep * L * L) + (row * L) + ((col + 1) * VEC);
                            scratch += ELT(din, off + v)[c];
                        }
                        if (front_of_dblock) {
                            if (front_p)
                                off = front_dblock ? ((L / B - 1) * L * L) + (row * L) + (col * VEC)
                                                   : ((dep / B - 1) * L * L) + (row * L) + (col * VEC);
                            else off = ((dep / B) * L * L) + (row * L) + (col * VEC);
                            scratch += ELT(recv_front, off + v)[c];
                        } else {
                            off = ((dep - 1) * L * L) + (row * L) + (col * VEC);
                            scratch += ELT(din, off + v)[c];
                        }
                        if (back_of_dblock) {
                            if (back_p)
                                off = back_dblock ? (row * L) + (col * VEC)
                                                  : ((dep / B + 1) * L * L) + (row * L) + (col * VEC);
                            else off = ((dep / B) * L * L) + (row * L) + (col * VEC);
                            scratch += ELT(recv_back, off + v)[c];
                        } else {
                            off = ((dep + 1) * L * L) + (row * L) + (col * VEC);
                            scratch += ELT(din, off + v)[c];
                        }
                        scratch -= 6 * ELT(din, my_offset + v)[c];
                        ELT(dout, my_offset + v)[c] = scratch;
                    }
            }
    }
    free(sa);
    free(sb);
    free(sl);
    free(sr);
    free(sf);
    free(sk);
}

void zo_laplace(const zo_shape* s, const double* in, double* out, int ncomp) {
    if (s->dims == 2) laplace_2d(s, in, out, ncomp);
    else laplace_3d(s, in, out, ncomp);
}

/* dx (zmodel/stencils/dx.cu:10-82).  Halos as init_packed_dx / pack_dx_*
 * build them (dx.cu:104-161): left0/left1 = vectors B-1 / B-2 of the left
 * rank, right0/right1 = vectors 0 / 1 of the right rank, [row*VEC + lane]. */
int zo_dx(const zo_shape* s, const double* in, double* out, double delta, int nc) {
    if (s->dims != 2 || s->b < 2) return -1;
    const int64_t L = s->l, B = s->b, VEC = s->vec, P = s->p;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        int left_cid = (cid == 0) ? (int)P - 1 : cid - 1;
        int right_cid = (int)((cid + 1) % P);
        const double* lrank = in + (int64_t)rank_id(s, rid, left_cid, 0) * s->local * nc;
        const double* rrank = in + (int64_t)rank_id(s, rid, right_cid, 0) * s->local * nc;
        const double* din = in + (int64_t)id * s->local * nc;
        double* dout = out + (int64_t)id * s->local * nc;
        for (int64_t blk = 0; blk < L * B; blk++) {
            int64_t row = blk / B, col = blk % B;
            int64_t my_offset = (row * L) + (col * VEC);
            int left0_col = col == 0, left1_col = col == 1, right0_col = col == B - 1, right1_col = col == B - 2;
            int left_p = cid == 0, right_p = cid == P - 1;
            for (int64_t v = 0; v < VEC; v++)
                for (int c = 0; c < nc; c++) {
                    /* packed_left0[row*VEC + x] = left rank in[row*L + (L-VEC) + x], etc. */
#define LEFT0(x) ELT(lrank, row * L + (L - VEC) + (x))[c]
#define LEFT1(x) ELT(lrank, row * L + (L - 2 * VEC) + (x))[c]
#define RIGHT0(x) ELT(rrank, row * L + (x))[c]
#define RIGHT1(x) ELT(rrank, row * L + VEC + (x))[c]
                    double scratch = 0.0;
                    int64_t lvi = left_p ? (v - 1 + VEC) % VEC : v;
                    if (left0_col) scratch += (LEFT0(lvi) * -8);
                    else scratch += (ELT(din, (row * L) + ((col - 1) * VEC) + v)[c] * -8);
                    if (left0_col) scratch += LEFT1(lvi);
                    else if (left1_col) scratch += LEFT0(lvi);
                    else scratch += ELT(din, (row * L) + ((col - 2) * VEC) + v)[c];
                    int64_t rvi = right_p ? (v + 1) % VEC : v;
                    if (right0_col) scratch += (RIGHT0(rvi) * 8);
                    else scratch += (ELT(din, (row * L) + ((col + 1) * VEC) + v)[c] * 8);
                    if (right0_col) scratch -= RIGHT1(rvi);
                    else if (right1_col) scratch -= RIGHT0(rvi);
                    else scratch -= ELT(din, (row * L) + ((col + 2) * VEC) + v)[c];
                    scratch /= (12 * delta);
                    ELT(dout, my_offset + v)[c] = scratch;
#undef LEFT0
#undef LEFT1
#undef RIGHT0
#undef RIGHT1
                }
        }
    }
    return 0;
}

/* dy (zmodel/stencils/dy.cu:10-103).  Halos as init_packed_dy / pack_dy_*
 * build them (dy.cu:127-185): above0/above1 = rows B-1 / B-2 of every block of
 * the rank above, below0/below1 = rows 0 / 1 of every block of the rank below,
 * [block*L + col]. */
int zo_dy(const zo_shape* s, const double* in, double* out, double delta, int nc) {
    if (s->dims != 2 || s->b < 2) return -1;
    const int64_t L = s->l, B = s->b, VEC = s->vec, P = s->p;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->nranks; id++) {
        int rid, cid, did;
        rank_coords(s, id, &rid, &cid, &did);
        int above_rid = (rid == 0) ? (int)P - 1 : rid - 1;
        int below_rid = (int)((rid + 1) % P);
        const double* arank = in + (int64_t)rank_id(s, above_rid, cid, 0) * s->local * nc;
        const double* brank = in + (int64_t)rank_id(s, below_rid, cid, 0) * s->local * nc;
        const double* din = in + (int64_t)id * s->local * nc;
        double* dout = out + (int64_t)id * s->local * nc;
        /* packed buffers index [i*L + j] with i = block; map to the source rank's rows */
#define ABOVE0(o) ELT(arank, ((o) / L) * L * B + L * (B - 1) + ((o) % L))[c]
#define ABOVE1(o) ELT(arank, ((o) / L) * L * B + L * (B - 2) + ((o) % L))[c]
#define BELOW0(o) ELT(brank, ((o) / L) * L * B + ((o) % L))[c]
#define BELOW1(o) ELT(brank, ((o) / L) * L * B + L + ((o) % L))[c]
        for (int64_t blk = 0; blk < L * B; blk++) {
            int64_t row = blk / B, col = blk % B;
            int top0_of_block = (row % B) == 0, top1_of_block = (row % B) == 1;
            int bottom0_of_block = (row % B) == B - 1, bottom1_of_block = (row % B) == B - 2;
            int top0_row = row == 0, top1_row = row == 1, bottom0_row = row == L - 1, bottom1_row = row == L - 2;
            int top_p = rid == 0, bottom_p = rid == P - 1;
            for (int64_t t = 0; t < VEC; t++)
                for (int c = 0; c < nc; c++) {
                    int64_t col_offset = (col * VEC) + t;
                    int64_t my_offset = (row * L) + col_offset;
                    double scratch = 0.0;
                    int64_t off;
                    if (top0_of_block) {
                        if (top_p) off = top0_row ? ((L / B - 1) * L) : ((row / B - 1) * L);
                        else off = ((row / B) * L);
                        scratch += (ABOVE0(off + col_offset) * -8);
                    } else {
                        off = ((row - 1) * L);
                        scratch += (ELT(din, off + col_offset)[c] * -8);
                    }
                    if (top0_of_block) {
                        if (top_p) off = top0_row ? ((L / B - 1) * L) : ((row / B - 1) * L);
                        else off = ((row / B) * L);
                        scratch += ABOVE1(off + col_offset);
                    } else if (top1_of_block) {
                        if (top_p) off = top1_row ? ((L / B - 1) * L) : ((row / B - 1) * L);
                        else off = ((row / B) * L);
                        scratch += ABOVE0(off + col_offset);
                    } else {
                        off = ((row - 2) * L);
                        scratch += ELT(din, off + col_offset)[c];
                    }
                    if (bottom0_of_block) {
                        if (bottom_p) off = bottom0_row ? 0 : ((row / B + 1) * L);
                        else off = ((row / B) * L);
                        scratch += (BELOW0(off + col_offset) * 8);
                    } else {
                        off = ((row + 1) * L);
                        scratch += (ELT(din, off + col_offset)[c] * 8);
                    }
                    if (bottom0_of_block) {
                        if (bottom_p) off = bottom0_row ? 0 : ((row / B + 1) * L);
                        else off = ((row / B) * L);
                        scratch -= BELOW1(off + col_offset);
                    } else if (bottom1_of_block) {
                        if (bottom_p) off = bottom1_row ? 0 : ((row / B + 1) * L);
                        else off = ((row / B) * L);
                        scratch -= BELOW0(off + col_offset);
                    } else {
                        off = ((row + 2) * L);
                        scratch -= ELT(din, off + col_offset)[c];
                    }
                    scratch /= (12 * delta);
                    ELT(dout, my_offset)[c] = scratch;
                }
        }
#undef ABOVE0
#undef ABOVE1
#undef BELOW0
#undef BELOW1
    }
    return 0;
}
