// plan.hip -- host side of the C ABI declared in include/stencilfft.h: shape
// validation, plans (twiddle tables, kernel-path choice, scratch), the stage
// sequences of the forward / inverse transforms, stencils and index maps.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/stencilfft.h"
#include "kernels.hpp"
#include "maps.hpp"

using zm::Geo;

struct zm_plan {
    zm_shape shape;
    zm_dims dims;
    zm_precision precision;
    int batch;
    unsigned flags;
    int device;
    Geo geo;
    double2* tw64;  // exp(-2 pi i t / N), t < N
    float2* tw32;
    bool two_stage;
    int lpb;  // lines per workgroup of the fused line kernel
    int rpb;  // rank rows per workgroup of the two-stage kernels
    void* own_scratch;
    size_t scratch_bytes;
    // compile-time specialised line pass (line_fast.hip)
    bool fast;
    double2* ltab64;  // per-line twiddle tables (W_S, W_N two-level)
    float2* ltab32;
    int fast_grid_fwd, fast_grid_inv;
    int fast_grid_lapx;  // the fused Laplacian + x pass's grid (0: not built for this plan)
    // per-stage timing (zm_timing_*)
    bool timing;
    struct Timed {
        int stage;
        hipEvent_t start, stop;
        bool own_start;  // start is this entry's own event, not the previous launch's stop
    };
    std::vector<Timed> timed;
    hipEvent_t last_stop;  // completion event of the previous timed launch
    hipStream_t last_stream;
};

namespace {

// Runs `launch` (returning hipError_t) and, when the plan records per-stage
// timing, records one event after it on `stream`.  A launch's time is the
// interval from the previous timed launch's completion event on the same
// stream (or from an event recorded just before it, for the first launch on
// a stream) to its own: one event per launch, so back-to-back launches are
// timed without doubling the events in the queue.
template <typename F>
hipError_t timed_launch(zm_plan* p, int stage, hipStream_t stream, F&& launch, bool record = true) {
    if (!p->timing || !record) return launch();
    zm_plan::Timed t{stage, nullptr, nullptr, false};
    hipError_t e = hipSuccess;
    if (p->last_stop && p->last_stream == stream) {
        t.start = p->last_stop;
    } else {
        e = hipEventCreate(&t.start);
        if (e == hipSuccess) e = hipEventRecord(t.start, stream);
        t.own_start = true;
    }
    if (e == hipSuccess) e = hipEventCreate(&t.stop);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(t.stop, stream);
    if (e == hipSuccess) {
        p->timed.push_back(t);
        p->last_stop = t.stop;
        p->last_stream = stream;
    } else {
        if (t.own_start && t.start) (void)hipEventDestroy(t.start);
        if (t.stop) (void)hipEventDestroy(t.stop);
    }
    return e;
}

void drop_timing(zm_plan* p) {
    for (auto& t : p->timed) {
        if (t.own_start) (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    p->timed.clear();
    p->last_stop = nullptr;
    p->last_stream = nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of zm_forward_noalloc: a file-scope device array, one copy on every
// GPU the library runs on, resident until the process ends.  In host code the
// name `scratch` is only a stand-in for the symbol, never a device address:
// zm_plan_create looks up the current device's copy once with
// hipGetSymbolAddress, and every use goes through that per-device address.
// ---------------------------------------------------------------------------
namespace noalloc {

static __device__ __align__(256) unsigned char scratch[ZM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex mu;
bool looked_up[kMaxDevices];
void* addr[kMaxDevices];  // this device's copy of `scratch`, or null if the lookup failed

// Called by zm_plan_create with `device` current.
void resolve(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(mu);
    if (looked_up[device]) return;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) ptr = nullptr;
    addr[device] = ptr;
    looked_up[device] = true;
}

void* device_copy(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    return looked_up[device] ? addr[device] : nullptr;
}

}  // namespace noalloc

namespace {

bool is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

int ilog2(int x) {
    int r = 0;
    while ((1 << r) < x) r++;
    return r;
}

zm::MapShape map_shape(const zm_shape* s) { return zm::make_map_shape(s->n, s->p, s->b); }

size_t elem_bytes(zm_precision p) { return p == ZM_FP64 ? sizeof(double2) : sizeof(float2); }

zm_status launch_status(hipError_t e) { return e == hipSuccess ? ZM_OK : ZM_ERR_LAUNCH; }

#define ZM_TRY(expr)                              \
    do {                                          \
        zm_status _st = (expr);                   \
        if (_st != ZM_OK) return _st;             \
    } while (0)

int ipow(int a, int e) {
    int r = 1;
    while (e-- > 0) r *= a;
    return r;
}

// Workgroup size targets (elements held in LDS per workgroup).
constexpr int kLineTarget = 4096;
constexpr int kStageTarget = 4096;
constexpr size_t kMaxLds = 160 * 1024;

}  // namespace

extern "C" {

zm_status zm_shape_check(const zm_shape* s, zm_dims* out) {
    if (!s) return ZM_ERR_INVALID_ARG;
    if ((s->dims != 2 && s->dims != 3) || s->n <= 0 || s->p <= 0 || s->b <= 0) return ZM_ERR_SHAPE;
    // zmodel/utils.h:47 (LOCAL_DIM = N/P), :60 (VEC = L/B), fft_host_plans.h:46-50 (VEC/P)
    if (s->n % s->p) return ZM_ERR_SHAPE;
    const int l = s->n / s->p;
    if (l % s->b) return ZM_ERR_SHAPE;
    const int vec = l / s->b;
    if (vec % s->p) return ZM_ERR_SHAPE;
    if (!is_pow2(s->n) || !is_pow2(s->p) || !is_pow2(s->b)) return ZM_ERR_UNSUPPORTED;
    const double field = std::pow((double)s->n, s->dims);
    if (field >= 2147483648.0) return ZM_ERR_UNSUPPORTED;  // 32-bit element index inside a field
    if (out) {
        out->l = l;
        out->vec = vec;
        out->vp = vec / s->p;
        out->m = s->b * s->p;
        out->nranks = ipow(s->p, s->dims);
        out->local = (int64_t)l * l * (s->dims == 3 ? l : 1);
        out->field = out->local * out->nranks;
    }
    return ZM_OK;
}

zm_status zm_plan_create(zm_plan** plan, const zm_shape* shape, zm_precision precision, int batch, unsigned flags) {
    if (!plan || !shape) return ZM_ERR_INVALID_ARG;
    *plan = nullptr;
    if (precision != ZM_FP64 && precision != ZM_FP32 && precision != ZM_FP32_MFMA) return ZM_ERR_INVALID_ARG;
    if (batch < 1 || batch > 65535) return ZM_ERR_INVALID_ARG;
    if (flags & ~(ZM_FLAG_TWO_STAGE | ZM_FLAG_CALLER_SCRATCH)) return ZM_ERR_INVALID_ARG;
    zm_dims d;
    ZM_TRY(zm_shape_check(shape, &d));
    const bool mfma = precision == ZM_FP32_MFMA;
    // the split-precision radix-16 passes live in the specialised fused line
    // pass only
    if (mfma && ((flags & ZM_FLAG_TWO_STAGE) || !zm::fast_line_supported<float>(d.vec, d.m)))
        return ZM_ERR_UNSUPPORTED;

    const int n = shape->n;
    const int64_t rows = (int64_t)d.local / d.l;  // L^(dims-1) local rows per rank
    const size_t es = elem_bytes(precision);

    // Fused line kernel when a whole line (N elements) fits LDS; otherwise
    // the two-stage path.
    bool two_stage = (flags & ZM_FLAG_TWO_STAGE) != 0;
    int lpb = 1;
    if (!two_stage) {
        if (es * zm::padded_elems(n) > kMaxLds || n / zm::EPT > 512) two_stage = true;
    }
    if (!two_stage) {
        while ((int64_t)lpb * 2 * n <= kLineTarget && lpb * 2 <= rows) lpb *= 2;
        if ((int64_t)lpb * n < zm::EPT) return ZM_ERR_UNSUPPORTED;
    }
    int rpb = 1;
    while ((int64_t)rpb * 2 * d.l <= kStageTarget && rpb * 2 <= rows) rpb *= 2;
    if (two_stage) {
        if ((int64_t)rpb * d.l < zm::EPT || es * zm::padded_elems(rpb * d.l) > kMaxLds || rpb * d.l / zm::EPT > 512)
            return ZM_ERR_UNSUPPORTED;
    }

    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) return ZM_ERR_DEVICE;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) return ZM_ERR_DEVICE;

    noalloc::resolve(device);

    zm_plan* p = new (std::nothrow) zm_plan();
    if (!p) return ZM_ERR_NOMEM;
    p->shape = *shape;
    p->dims = d;
    p->precision = precision;
    p->batch = batch;
    p->flags = flags;
    p->device = device;
    p->two_stage = two_stage;
    p->lpb = lpb;
    p->rpb = rpb;
    Geo& g = p->geo;
    g.dims = shape->dims;
    g.n = n;
    g.p = shape->p;
    g.b = shape->b;
    g.l = d.l;
    g.vec = d.vec;
    g.vp = d.vp;
    g.m = d.m;
    g.nranks = d.nranks;
    g.local = d.local;
    g.field = d.field;
    g.ln = ilog2(n);
    g.lp = ilog2(shape->p);
    g.lb = ilog2(shape->b);
    g.ll = ilog2(d.l);
    g.lvec = ilog2(d.vec);
    p->scratch_bytes = (size_t)batch * (size_t)d.field * es;

    // The plan's own scratch is allocated here, once, so that the plan is
    // immutable after creation; ZM_FLAG_CALLER_SCRATCH plans own none and
    // every transform call must bring one.
    if (!(flags & ZM_FLAG_CALLER_SCRATCH) && hipMalloc(&p->own_scratch, p->scratch_bytes) != hipSuccess) {
        delete p;
        return ZM_ERR_NOMEM;
    }
    if (hipMalloc(&p->tw64, sizeof(double2) * n) != hipSuccess) {
        if (p->own_scratch) (void)hipFree(p->own_scratch);
        delete p;
        return ZM_ERR_NOMEM;
    }
    if (hipMalloc(&p->tw32, sizeof(float2) * n) != hipSuccess) {
        (void)hipFree(p->tw64);
        if (p->own_scratch) (void)hipFree(p->own_scratch);
        delete p;
        return ZM_ERR_NOMEM;
    }
    hipError_t e = zm::launch_twiddles(p->tw64, p->tw32, n, nullptr);
    // the specialised line pass, when the (VEC, M) pair has one and its LDS fits
    const bool fp64 = precision == ZM_FP64;
    if (e == hipSuccess && !two_stage &&
        (fp64 ? zm::fast_line_supported<double>(d.vec, d.m) : zm::fast_line_supported<float>(d.vec, d.m))) {
        const int nt = zm::line_table_entries(d.vec, d.m);
        if (hipMalloc(&p->ltab64, sizeof(double2) * nt) == hipSuccess &&
            hipMalloc(&p->ltab32, sizeof(float2) * nt) == hipSuccess) {
            e = zm::launch_line_tables(p->ltab64, p->ltab32, d.vec, d.m, nullptr);
            int cus = 0;
            if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
            const int of = fp64 ? zm::fast_line_occupancy<double>(d.vec, d.m, true, false, false)
                                : zm::fast_line_occupancy<float>(d.vec, d.m, true, mfma, false);
            const int oi = fp64 ? zm::fast_line_occupancy<double>(d.vec, d.m, false, false, false)
                                : zm::fast_line_occupancy<float>(d.vec, d.m, false, mfma, false);
            // the fused Laplacian + x pass (FP32 vector ALU)
            const bool lx = fp64 ? zm::fast_line_lapx<double>(d.vec, d.m, false, shape->dims, shape->b)
                                 : zm::fast_line_lapx<float>(d.vec, d.m, mfma, shape->dims, shape->b);
            const int ol = !lx ? 0 : fp64 ? zm::fast_line_occupancy<double>(d.vec, d.m, true, false, true)
                                          : zm::fast_line_occupancy<float>(d.vec, d.m, true, mfma, true);
            if (e == hipSuccess && of > 0 && oi > 0) {
                p->fast = true;
                p->fast_grid_fwd = of * cus;
                p->fast_grid_inv = oi * cus;
                p->fast_grid_lapx = ol * cus;
                if (shape->dims == 3 && ol > 0) {
                    // 3D: row tiles times plane chunks, about the resident
                    // workgroups, the chunks at least two planes long
                    const int ntile = n >> zm::fast_line_lapx_lines_log2(d.vec, d.m);
                    int nzc = 1;
                    while (2 * nzc * ntile <= ol * cus && 2 * nzc <= n / 2) nzc *= 2;
                    p->fast_grid_lapx = ntile * nzc;
                }
            }
        }
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess || (mfma && !p->fast)) {
        (void)hipFree(p->tw64);
        (void)hipFree(p->tw32);
        if (p->ltab64) (void)hipFree(p->ltab64);
        if (p->ltab32) (void)hipFree(p->ltab32);
        if (p->own_scratch) (void)hipFree(p->own_scratch);
        delete p;
        return e != hipSuccess ? ZM_ERR_DEVICE : ZM_ERR_UNSUPPORTED;
    }
    *plan = p;
    return ZM_OK;
}

zm_status zm_plan_destroy(zm_plan* p) {
    if (!p) return ZM_ERR_INVALID_ARG;
    zm_status st = ZM_OK;
    if (p->tw64 && hipFree(p->tw64) != hipSuccess) st = ZM_ERR_DEVICE;
    if (p->tw32 && hipFree(p->tw32) != hipSuccess) st = ZM_ERR_DEVICE;
    if (p->ltab64 && hipFree(p->ltab64) != hipSuccess) st = ZM_ERR_DEVICE;
    if (p->ltab32 && hipFree(p->ltab32) != hipSuccess) st = ZM_ERR_DEVICE;
    if (p->own_scratch && hipFree(p->own_scratch) != hipSuccess) st = ZM_ERR_DEVICE;
    drop_timing(p);
    delete p;
    return st;
}

zm_status zm_plan_dims(const zm_plan* p, zm_shape* shape, zm_dims* dims, zm_precision* precision, int* batch) {
    if (!p) return ZM_ERR_INVALID_ARG;
    if (shape) *shape = p->shape;
    if (dims) *dims = p->dims;
    if (precision) *precision = p->precision;
    if (batch) *batch = p->batch;
    return ZM_OK;
}

zm_status zm_plan_path(const zm_plan* p, unsigned* path) {
    if (!p || !path) return ZM_ERR_INVALID_ARG;
    *path = (p->two_stage ? ZM_PATH_TWO_STAGE : 0u) | (p->fast ? ZM_PATH_FAST_LINE : 0u) |
            (p->precision == ZM_FP32_MFMA ? ZM_PATH_MFMA : 0u);
    return ZM_OK;
}

size_t zm_scratch_bytes(const zm_plan* p) { return p ? p->scratch_bytes : 0; }

size_t zm_field_bytes(const zm_plan* p) {
    return p ? (size_t)p->batch * (size_t)p->dims.field * elem_bytes(p->precision) : 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// transform sequences
// ---------------------------------------------------------------------------
namespace {

int axis_stride(const zm_plan* p, int axis) { return axis == 0 ? 1 : (axis == 1 ? p->shape.p : p->shape.p * p->shape.p); }

enum RepackKind { REPACK_2D = 0, REPACK_XY = 1, REPACK_YZ = 2 };

zm::RepackArgs repack_args(const zm_plan* p, RepackKind k) {
    const int64_t L = p->dims.l;
    zm::RepackArgs a{};
    a.tx = p->dims.vec < zm::REPACK_TX ? p->dims.vec : zm::REPACK_TX;
    a.ty = L < zm::REPACK_TY ? (int)L : zm::REPACK_TY;
    if (k == REPACK_2D) {
        a.ecount = 1;
        a.ei = a.eo = 0;
        a.sxi = L;
        a.syo = L;
    } else if (k == REPACK_XY) {
        a.ecount = (int)L;
        a.ei = a.eo = L * L;
        a.sxi = L;
        a.syo = L;
    } else {
        a.ecount = (int)L;
        a.ei = a.eo = L;
        a.sxi = L * L;
        a.syo = L * L;
    }
    return a;
}

template <typename T>
struct Exec {
    zm_plan* p;
    hipStream_t st;
    bool compute;
    const zm::Cx<T>* tw;
    bool record = true;  // per-launch timing events when the plan records them

    const zm::Cx<T>* line_tables() const {
        return p->precision == ZM_FP64 ? (const zm::Cx<T>*)p->ltab64 : (const zm::Cx<T>*)p->ltab32;
    }
    // lapx: the x pass of zm_laplace_forward (fast_grid_lapx != 0 only),
    // transforming `scale` times the input's Laplacian
    zm_status line(const void* in, void* out, int axis, bool fwd, double scale, bool lapx = false) const {
        zm::LineArgs a{axis_stride(p, axis), (int)(p->dims.local / p->dims.l), p->lpb};
        if (p->fast && compute) {
            const int grid = lapx ? p->fast_grid_lapx : fwd ? p->fast_grid_fwd : p->fast_grid_inv;
            return launch_status(timed_launch(p, lapx ? ZM_STAGE_LINE_LAPLACE : ZM_STAGE_LINE, st, [&] {
                return zm::launch_line_fast<T>((const zm::Cx<T>*)in, (zm::Cx<T>*)out, p->geo, a, line_tables(),
                                               p->batch, fwd, scale, grid, p->precision == ZM_FP32_MFMA, lapx, st);
            }, record));
        }
        if (lapx) return ZM_ERR_UNSUPPORTED;
        return launch_status(timed_launch(p, ZM_STAGE_LINE, st, [&] {
            return zm::launch_line<T>((const zm::Cx<T>*)in, (zm::Cx<T>*)out, p->geo, a, tw, p->batch, fwd, compute,
                                      scale, st);
        }, record));
    }
    zm::StageArgs sargs(int axis) const {
        return zm::StageArgs{axis_stride(p, axis), (int)(p->dims.local / p->dims.l), p->rpb,
                             p->shape.dims == 3 ? p->dims.l : 1};
    }
    zm_status stage_a(const void* in, void* out, int axis, bool fwd, double scale) const {
        return launch_status(timed_launch(p, ZM_STAGE_A, st, [&] {
            return zm::launch_stage_a<T>((const zm::Cx<T>*)in, (zm::Cx<T>*)out, p->geo, sargs(axis), tw, p->batch,
                                         fwd, compute, scale, st);
        }, record));
    }
    zm_status stage_b(const void* in, void* out, int axis, bool fwd) const {
        return launch_status(timed_launch(p, ZM_STAGE_B, st, [&] {
            return zm::launch_stage_b<T>((const zm::Cx<T>*)in, (zm::Cx<T>*)out, p->geo, sargs(axis), tw, p->batch,
                                         fwd, compute, st);
        }, record));
    }
    zm_status repack(const void* in, void* out, RepackKind k, bool fwd) const {
        return launch_status(timed_launch(p, ZM_STAGE_REPACK, st, [&] {
            return zm::launch_repack<T>((const zm::Cx<T>*)in, (zm::Cx<T>*)out, p->geo, repack_args(p, k), p->batch,
                                        fwd, st);
        }, record));
    }

    // forward_fft / forward_fft_3d (zmodel/transforms/fft.cu:199-339,
    // fft3d.cu:234-412) as x-pass, repack, y-pass [, repack-yz, z-pass].
    zm_status forward(const void* in, void* out, void* s) const {
        const bool d3 = p->shape.dims == 3;
        if (!p->two_stage) {
            ZM_TRY(line(in, s, 0, true, 1.0));
            ZM_TRY(repack(s, out, d3 ? REPACK_XY : REPACK_2D, true));
            ZM_TRY(line(out, out, 1, true, 1.0));
            if (d3) {
                ZM_TRY(repack(out, s, REPACK_YZ, true));
                ZM_TRY(line(s, out, 2, true, 1.0));
            }
            return ZM_OK;
        }
        if (!d3) {
            ZM_TRY(stage_a(in, out, 0, true, 1.0));
            ZM_TRY(stage_b(out, s, 0, true));
            ZM_TRY(repack(s, out, REPACK_2D, true));
            ZM_TRY(stage_a(out, s, 1, true, 1.0));
            ZM_TRY(stage_b(s, out, 1, true));
            return ZM_OK;
        }
        ZM_TRY(stage_a(in, s, 0, true, 1.0));
        ZM_TRY(stage_b(s, out, 0, true));
        ZM_TRY(repack(out, s, REPACK_XY, true));
        ZM_TRY(stage_a(s, out, 1, true, 1.0));
        ZM_TRY(stage_b(out, s, 1, true));
        ZM_TRY(repack(s, out, REPACK_YZ, true));
        ZM_TRY(stage_a(out, s, 2, true, 1.0));
        ZM_TRY(stage_b(s, out, 2, true));
        return ZM_OK;
    }

    // The reference's laplace followed by forward_fft (zmodel/main.cu:50-63).
    // With the fused pass the Laplacian is formed inside the x pass (one
    // sweep of the field less: 3 instead of 4 in 2D, 5 instead of 6 in 3D);
    // otherwise the stencil writes `out` and the
    // forward chain runs in place on it (its first pass reads `out` whole
    // before the repack writes it).  Either way the result is bit for bit
    // zm_forward(zm_laplace(in)).
    zm_status laplace_forward(const void* in, void* out, void* s, double lap_scale) const {
        if (p->two_stage) return ZM_ERR_UNSUPPORTED;
        if (p->fast_grid_lapx > 0) {
            const bool d3 = p->shape.dims == 3;
            ZM_TRY(line(in, s, 0, true, lap_scale, true));
            ZM_TRY(repack(s, out, d3 ? REPACK_XY : REPACK_2D, true));
            ZM_TRY(line(out, out, 1, true, 1.0));
            if (d3) {
                ZM_TRY(repack(out, s, REPACK_YZ, true));
                ZM_TRY(line(s, out, 2, true, 1.0));
            }
            return ZM_OK;
        }
        ZM_TRY(launch_status(timed_launch(p, ZM_STAGE_LAPLACE, st, [&] {
            return zm::launch_stencil<T>(in, out, p->geo, p->batch, zm::OP_LAPLACE, true, lap_scale, false, st);
        }, record)));
        return forward(out, out, s);
    }

    // inverse_fft / inverse_fft_3d (zmodel/transforms/fft.cu:409-514,
    // fft3d.cu:508-633): the reverse chain, scale applied once on the x-pass.
    zm_status inverse(const void* in, void* out, void* s, double scale) const {
        const bool d3 = p->shape.dims == 3;
        if (!p->two_stage) {
            if (d3) {
                ZM_TRY(line(in, s, 2, false, 1.0));
                ZM_TRY(repack(s, out, REPACK_YZ, false));
                ZM_TRY(line(out, out, 1, false, 1.0));
                ZM_TRY(repack(out, s, REPACK_XY, false));
                ZM_TRY(line(s, out, 0, false, scale));
                return ZM_OK;
            }
            ZM_TRY(line(in, s, 1, false, 1.0));
            ZM_TRY(repack(s, out, REPACK_2D, false));
            ZM_TRY(line(out, out, 0, false, scale));
            return ZM_OK;
        }
        if (!d3) {
            ZM_TRY(stage_b(in, out, 1, false));
            ZM_TRY(stage_a(out, s, 1, false, 1.0));
            ZM_TRY(repack(s, out, REPACK_2D, false));
            ZM_TRY(stage_b(out, s, 0, false));
            ZM_TRY(stage_a(s, out, 0, false, scale));
            return ZM_OK;
        }
        ZM_TRY(stage_b(in, s, 2, false));
        ZM_TRY(stage_a(s, out, 2, false, 1.0));
        ZM_TRY(repack(out, s, REPACK_YZ, false));
        ZM_TRY(stage_b(s, out, 1, false));
        ZM_TRY(stage_a(out, s, 1, false, 1.0));
        ZM_TRY(repack(s, out, REPACK_XY, false));
        ZM_TRY(stage_b(out, s, 0, false));
        ZM_TRY(stage_a(s, out, 0, false, scale));
        return ZM_OK;
    }
};

zm_status check_io(const zm_plan* p, const void* in, const void* out) {
    if (!p || !in || !out) return ZM_ERR_INVALID_ARG;
    if (in == out) return ZM_ERR_ALIAS;
    return ZM_OK;
}

// The caller's scratch, else the one the plan allocated at creation (none
// for ZM_FLAG_CALLER_SCRATCH plans: then NULL is an invalid argument).
zm_status scratch_for(const zm_plan* p, void* scratch, void** s) {
    *s = scratch ? scratch : p->own_scratch;
    return *s ? ZM_OK : ZM_ERR_INVALID_ARG;
}

zm_status run_transform(zm_plan* p, const void* in, void* out, void* scratch, void* stream, bool fwd, bool compute,
                        double scale, bool record = true) {
    ZM_TRY(check_io(p, in, out));
    void* s = nullptr;
    ZM_TRY(scratch_for(p, scratch, &s));
    if (s == in || s == out) return ZM_ERR_ALIAS;
    hipStream_t st = (hipStream_t)stream;
    if (p->precision == ZM_FP64) {
        Exec<double> ex{p, st, compute, (const double2*)p->tw64, record};
        return fwd ? ex.forward(in, out, s) : ex.inverse(in, out, s, scale);
    }
    Exec<float> ex{p, st, compute, (const float2*)p->tw32, record};
    return fwd ? ex.forward(in, out, s) : ex.inverse(in, out, s, scale);
}

}  // namespace

extern "C" {

zm_status zm_forward(zm_plan* p, const void* in, void* out, void* scratch, void* stream) {
    return run_transform(p, in, out, scratch, stream, true, true, 1.0);
}

zm_status zm_forward_noalloc(zm_plan* p, const void* in, void* out, void* stream) {
    if (!p) return ZM_ERR_INVALID_ARG;
    if (p->scratch_bytes > ZM_NOALLOC_SCRATCH_BYTES) return ZM_ERR_SCRATCH_LIMIT;
    void* s = noalloc::device_copy(p->device);
    if (!s) return ZM_ERR_DEVICE;
    // no timing events: creating them is an allocation
    return run_transform(p, in, out, s, stream, true, true, 1.0, false);
}

zm_status zm_laplace_forward(zm_plan* p, const void* in, void* out, double scale, void* scratch, void* stream) {
    ZM_TRY(check_io(p, in, out));
    if (!std::isfinite(scale)) return ZM_ERR_INVALID_ARG;
    void* s = nullptr;
    ZM_TRY(scratch_for(p, scratch, &s));
    if (s == in || s == out) return ZM_ERR_ALIAS;
    const hipStream_t st = (hipStream_t)stream;
    if (p->precision == ZM_FP64) {
        Exec<double> ex{p, st, true, (const double2*)p->tw64};
        return ex.laplace_forward(in, out, s, scale);
    }
    Exec<float> ex{p, st, true, (const float2*)p->tw32};
    return ex.laplace_forward(in, out, s, scale);
}

zm_status zm_inverse(zm_plan* p, const void* in, void* out, double scale, void* scratch, void* stream) {
    return run_transform(p, in, out, scratch, stream, false, true, scale);
}

zm_status zm_permute_forward(zm_plan* p, const void* in, void* out, void* scratch, void* stream) {
    return run_transform(p, in, out, scratch, stream, true, false, 1.0);
}

zm_status zm_permute_inverse(zm_plan* p, const void* in, void* out, void* scratch, void* stream) {
    return run_transform(p, in, out, scratch, stream, false, false, 1.0);
}

zm_status zm_forward_adjoint(zm_plan* p, const void* in, void* out, void* scratch, void* stream) {
    // forward = P_out * DFT * P_in stage by stage, so its adjoint is the
    // inverse chain (conjugate twiddles, inverse DFTs) with scale 1.
    return run_transform(p, in, out, scratch, stream, false, true, 1.0);
}

zm_status zm_inverse_adjoint(zm_plan* p, const void* in, void* out, double scale, void* scratch, void* stream) {
    // (s * inverse)^H = s * forward; the scale goes on a separate pass only
    // when it is not 1.
    zm_status st = run_transform(p, in, out, scratch, stream, true, true, 1.0);
    if (st != ZM_OK || scale == 1.0) return st;
    return zm_spectral_multiply(p, out, out, ZM_SPECTRAL_SCALE, scale, 0.0, 1.0, stream);
}

// ---------------------------------------------------------------------------
// stencils
// ---------------------------------------------------------------------------

static zm_status run_stencil(zm_plan* p, const void* in, void* out, zm::StencilOp op, double coef, bool negate,
                             zm_element element, void* stream) {
    ZM_TRY(check_io(p, in, out));
    if (element != ZM_COMPLEX && element != ZM_REAL) return ZM_ERR_INVALID_ARG;
    const bool cplx = element == ZM_COMPLEX;
    const hipStream_t st = (hipStream_t)stream;
    return launch_status(timed_launch(p, op == zm::OP_LAPLACE ? ZM_STAGE_LAPLACE : ZM_STAGE_DERIV, st, [&] {
        if (p->precision == ZM_FP64) return zm::launch_stencil<double>(in, out, p->geo, p->batch, op, cplx, coef, negate, st);
        return zm::launch_stencil<float>(in, out, p->geo, p->batch, op, cplx, coef, negate, st);
    }));
}

zm_status zm_laplace(zm_plan* p, const void* in, void* out, double scale, zm_element element, void* stream) {
    return run_stencil(p, in, out, zm::OP_LAPLACE, scale, false, element, stream);
}

static zm_status deriv(zm_plan* p, const void* in, void* out, double delta, zm_element element, void* stream,
                       zm::StencilOp op, bool negate) {
    if (!p) return ZM_ERR_INVALID_ARG;
    if (p->shape.b < 2) return ZM_ERR_UNSUPPORTED;  // dx.cu:94-100, dy.cu:117-124 need B >= 2
    if (op == zm::OP_DZ4 && p->shape.dims != 3) return ZM_ERR_UNSUPPORTED;
    if (!(delta != 0.0) || !std::isfinite(delta)) return ZM_ERR_INVALID_ARG;
    return run_stencil(p, in, out, op, 12 * delta, negate, element, stream);
}

zm_status zm_dx4(zm_plan* p, const void* in, void* out, double delta, zm_element element, void* stream) {
    return deriv(p, in, out, delta, element, stream, zm::OP_DX4, false);
}
zm_status zm_dy4(zm_plan* p, const void* in, void* out, double delta, zm_element element, void* stream) {
    return deriv(p, in, out, delta, element, stream, zm::OP_DY4, false);
}
zm_status zm_dx4_adjoint(zm_plan* p, const void* in, void* out, double delta, zm_element element, void* stream) {
    return deriv(p, in, out, delta, element, stream, zm::OP_DX4, true);
}
zm_status zm_dy4_adjoint(zm_plan* p, const void* in, void* out, double delta, zm_element element, void* stream) {
    return deriv(p, in, out, delta, element, stream, zm::OP_DY4, true);
}
zm_status zm_dz4(zm_plan* p, const void* in, void* out, double delta, zm_element element, void* stream) {
    return deriv(p, in, out, delta, element, stream, zm::OP_DZ4, false);
}
zm_status zm_dz4_adjoint(zm_plan* p, const void* in, void* out, double delta, zm_element element, void* stream) {
    return deriv(p, in, out, delta, element, stream, zm::OP_DZ4, true);
}

static zm_status spectral(zm_plan* p, const void* in, void* out, zm_spectral kind, double p0, double p1,
                          double length, void* stream, bool conj) {
    if (!p || !in || !out) return ZM_ERR_INVALID_ARG;
    if ((int)kind < (int)ZM_SPECTRAL_POISSON || (int)kind > (int)ZM_SPECTRAL_POISSON_STENCIL) return ZM_ERR_INVALID_ARG;
    if (!(length > 0.0)) return ZM_ERR_INVALID_ARG;
    // the dx4 symbol is defined on the 2D layout only, like zm_dx4
    if ((kind == ZM_SPECTRAL_STENCIL || kind == ZM_SPECTRAL_POISSON_STENCIL) && p1 != 0.0 && p->shape.dims != 2)
        return ZM_ERR_INVALID_ARG;
    const double kscale = 2.0 * M_PI / length;
    const hipStream_t st = (hipStream_t)stream;
    return launch_status(timed_launch(p, ZM_STAGE_SPECTRAL, st, [&] {
        if (p->precision == ZM_FP64)
            return zm::launch_spectral<double>((const double2*)in, (double2*)out, p->geo, p->batch, (int)kind, p0, p1,
                                               kscale, conj, st);
        return zm::launch_spectral<float>((const float2*)in, (float2*)out, p->geo, p->batch, (int)kind, p0, p1, kscale,
                                          conj, st);
    }));
}

zm_status zm_spectral_multiply(zm_plan* p, const void* in, void* out, zm_spectral kind, double p0, double p1,
                               double length, void* stream) {
    return spectral(p, in, out, kind, p0, p1, length, stream, false);
}

zm_status zm_spectral_multiply_adjoint(zm_plan* p, const void* in, void* out, zm_spectral kind, double p0, double p1,
                                       double length, void* stream) {
    return spectral(p, in, out, kind, p0, p1, length, stream, true);
}

zm_status zm_real_index_table(zm_plan* p, int64_t* out, void* stream) {
    if (!p || !out) return ZM_ERR_INVALID_ARG;
    return launch_status(timed_launch(p, ZM_STAGE_CONVERT, (hipStream_t)stream,
                                      [&] { return zm::launch_index_table(out, p->geo, false, (hipStream_t)stream); }));
}

zm_status zm_reciprocal_index_table(zm_plan* p, int64_t* out, void* stream) {
    if (!p || !out) return ZM_ERR_INVALID_ARG;
    return launch_status(timed_launch(p, ZM_STAGE_CONVERT, (hipStream_t)stream,
                                      [&] { return zm::launch_index_table(out, p->geo, true, (hipStream_t)stream); }));
}

static zm_status natural_packed(zm_plan* p, const void* in, void* out, bool to_packed, void* stream) {
    ZM_TRY(check_io(p, in, out));
    const hipStream_t st = (hipStream_t)stream;
    return launch_status(timed_launch(p, ZM_STAGE_CONVERT, st, [&] {
        if (p->precision == ZM_FP64)
            return zm::launch_natural_packed<double>((const double2*)in, (double2*)out, p->geo, p->batch, to_packed, st);
        return zm::launch_natural_packed<float>((const float2*)in, (float2*)out, p->geo, p->batch, to_packed, st);
    }));
}

zm_status zm_natural_to_packed(zm_plan* p, const void* natural, void* packed, void* stream) {
    return natural_packed(p, natural, packed, true, stream);
}

zm_status zm_packed_to_natural(zm_plan* p, const void* packed, void* natural, void* stream) {
    return natural_packed(p, packed, natural, false, stream);
}

// ---------------------------------------------------------------------------
// host index maps (zmodel/utils.h:122-320)
// ---------------------------------------------------------------------------

int zm_packed_to_block_cyclic(const zm_shape* s, int c) { return zm::packed_to_block_cyclic(map_shape(s), c); }
int zm_block_cyclic_to_packed(const zm_shape* s, int c) { return zm::block_cyclic_to_packed(map_shape(s), c); }

void zm_real_to_global(const zm_shape* s, int rid, int cid, int lr, int lc, int* gr, int* gc) {
    const zm::MapShape m = map_shape(s);
    *gr = zm::local_row_to_global(m, rid, lr);
    *gc = zm::packed_col_to_global(m, cid, lc);
}

void zm_global_to_real(const zm_shape* s, int gr, int gc, int* rid, int* cid, int* lr, int* lc) {
    const zm::MapShape m = map_shape(s);
    zm::global_row_to_local(m, gr, rid, lr);
    zm::global_col_to_local(m, gc, cid, lc);
}

void zm_real_to_global_3d(const zm_shape* s, int rid, int cid, int did, int ld, int lr, int lc, int* gd, int* gr,
                          int* gc) {
    const zm::MapShape m = map_shape(s);
    *gd = zm::local_row_to_global(m, did, ld);
    *gr = zm::local_row_to_global(m, rid, lr);
    *gc = zm::packed_col_to_global(m, cid, lc);
}

void zm_global_to_real_3d(const zm_shape* s, int gd, int gr, int gc, int* rid, int* cid, int* did, int* ld, int* lr,
                          int* lc) {
    const zm::MapShape m = map_shape(s);
    zm::global_row_to_local(m, gd, did, ld);
    zm::global_row_to_local(m, gr, rid, lr);
    zm::global_col_to_local(m, gc, cid, lc);
}

int zm_lp(const zm_shape* s, int i) { return zm::lp(map_shape(s), i); }
int zm_ct(const zm_shape* s, int q) { return zm::ct(map_shape(s), q); }
int zm_lp_inv(const zm_shape* s, int j) { return zm::lp_inv(map_shape(s), j); }
int zm_ct_inv(const zm_shape* s, int k) { return zm::ct_inv(map_shape(s), k); }

void zm_reciprocal_to_global(const zm_shape* s, int rid, int cid, int lr, int lc, int* ky, int* kx) {
    const zm::MapShape m = map_shape(s);
    *kx = zm::recip_to_freq(m, cid, lr);
    *ky = zm::recip_to_freq(m, rid, lc);
}

void zm_global_to_reciprocal(const zm_shape* s, int ky, int kx, int* rid, int* cid, int* lr, int* lc) {
    const zm::MapShape m = map_shape(s);
    zm::freq_to_recip(m, kx, cid, lr);
    zm::freq_to_recip(m, ky, rid, lc);
}

void zm_reciprocal_to_global_3d(const zm_shape* s, int rid, int cid, int did, int ld, int lr, int lc, int* kd, int* kr,
                                int* kc) {
    const zm::MapShape m = map_shape(s);
    *kc = zm::recip_to_freq(m, cid, lr);
    *kr = zm::recip_to_freq(m, rid, ld);
    *kd = zm::recip_to_freq(m, did, lc);
}

void zm_global_to_reciprocal_3d(const zm_shape* s, int kd, int kr, int kc, int* rid, int* cid, int* did, int* ld,
                                int* lr, int* lc) {
    const zm::MapShape m = map_shape(s);
    zm::freq_to_recip(m, kc, cid, lr);
    zm::freq_to_recip(m, kr, rid, ld);
    zm::freq_to_recip(m, kd, did, lc);
}

zm_status zm_timing_enable(zm_plan* p, int enable) {
    if (!p) return ZM_ERR_INVALID_ARG;
    p->timing = enable != 0;
    p->last_stop = nullptr;  // a new timed sequence starts with its own start event
    return ZM_OK;
}

zm_status zm_timing_read(zm_plan* p, zm_stage stage, int* count, double* total_ms) {
    if (!p || !count || !total_ms || (int)stage < 0 || stage >= ZM_STAGE_COUNT) return ZM_ERR_INVALID_ARG;
    int c = 0;
    double tot = 0.0;
    for (auto& t : p->timed) {
        if (t.stage != (int)stage) continue;
        if (hipEventSynchronize(t.stop) != hipSuccess) return ZM_ERR_DEVICE;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) != hipSuccess) return ZM_ERR_DEVICE;
        tot += ms;
        c++;
    }
    *count = c;
    *total_ms = tot;
    return ZM_OK;
}

zm_status zm_timing_reset(zm_plan* p) {
    if (!p) return ZM_ERR_INVALID_ARG;
    drop_timing(p);
    return ZM_OK;
}

zm_status zm_init_host(const zm_shape* s, double* host_grid, int ncomp) {
    zm_dims d;
    if (!host_grid || (ncomp != 1 && ncomp != 2)) return ZM_ERR_INVALID_ARG;
    ZM_TRY(zm_shape_check(s, &d));
    const zm::MapShape m = map_shape(s);
    const int64_t L = d.l, N = s->n;
    // every packed position holds the global linear index of its point: the
    // block-cyclic fill of utils.cu:17-32 followed by vector_pack (utils.h:100-108)
    for (int id = 0; id < d.nranks; id++) {
        const int cid = id % s->p;
        const int rid = (id / s->p) % s->p;
        const int did = s->dims == 3 ? id / (s->p * s->p) : 0;
        const int64_t ndep = s->dims == 3 ? L : 1;
        for (int64_t ld = 0; ld < ndep; ld++)
            for (int64_t lr = 0; lr < L; lr++)
                for (int64_t lc = 0; lc < L; lc++) {
                    const int64_t gr = zm::local_row_to_global(m, rid, (int)lr);
                    const int64_t gc = zm::packed_col_to_global(m, cid, (int)lc);
                    const int64_t gd = s->dims == 3 ? zm::local_row_to_global(m, did, (int)ld) : 0;
                    const double v = (double)((gd * N + gr) * N + gc);
                    const int64_t at = (int64_t)id * d.local + (ld * L + lr) * L + lc;
                    host_grid[at * ncomp] = v;
                    if (ncomp == 2) host_grid[at * ncomp + 1] = 0.0;
                }
    }
    return ZM_OK;
}

const char* zm_strerror(zm_status s) {
    switch (s) {
        case ZM_OK: return "ok";
        case ZM_ERR_INVALID_ARG: return "invalid argument";
        case ZM_ERR_SHAPE: return "invalid shape: need P | N, B | N/P and P | N/(P*B)";
        case ZM_ERR_UNSUPPORTED: return "unsupported configuration";
        case ZM_ERR_DEVICE: return "HIP device or runtime error";
        case ZM_ERR_ALIAS: return "input and output buffers alias";
        case ZM_ERR_LAUNCH: return "kernel launch failed";
        case ZM_ERR_NOMEM: return "device allocation failed";
        case ZM_ERR_IO: return "file I/O failed";
        case ZM_ERR_SCRATCH_LIMIT: return "the plan needs more scratch than the static no-allocation buffer holds";
    }
    return "unknown status";
}

const char* zm_version(void) { return "stencilfft-amd 0.1.0 (gfx950)"; }

}  // extern "C"
