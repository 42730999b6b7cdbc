// zsc_api.cpp -- the C-ABI of libzsc: contexts, the reference-named drop-in
// kernels and the fused fast path.  Status codes instead of aborts.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/zsc.h"
#include "zsc_err.h"
#include "zsc_adjoint.h"

namespace zsc {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

template <class T>
struct Tag {
    using type = T;
};

// Calls f(Tag<S>{}) for the storage type of a precision.
template <class F>
static int dispatch(int prec, F &&f)
{
    switch (prec) {
    case ZSC_F64: return f(Tag<double>{});
    case ZSC_F32: return f(Tag<float>{});
    case ZSC_BF16: return f(Tag<bf16s>{});
    case ZSC_FP16: return f(Tag<f16s>{});
    case ZSC_F32D: return f(Tag<f32d>{});
    default: return fail(ZSC_ERR_ARG, "unknown precision %d", prec);
    }
}

static int hip_fail(hipError_t e, const char *what)
{
    return fail(ZSC_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define ZSC_HIP(call, what)                           \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

}  // namespace zsc

using namespace zsc;

enum { STATE_NONE = 0, STATE_A = 1, STATE_B = 2 };

struct zsc_ctx {
    Geo geo{};
    int prec = ZSC_F64;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    zsc_params params{};
    bool have_params = false;
    int grav_on = 1;
    // fast-path lattices; lat[cur] holds the state
    void *lat[2] = {nullptr, nullptr};
    size_t lat_bytes = 0;
    int cur = 0;
    int state = STATE_NONE;
    // phi, rho, ux, uy, uz at compute precision
    void *work = nullptr;
    size_t rsize = 8;
    double *dbuf = nullptr;  // n doubles of staging for host transfers
    double *stats_buf = nullptr;
    double *scratch = nullptr;  // this device's copy of the no-allocation scratch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    FusedCfg cfg{4, 16, 0};
    // deferred coefficient sums (backward calls with grads == NULL) exist
    bool adj_deferred = false;
    // FP32 adjoint lattices over the FP64 states of an F64 context
    bool adj_lam_f32 = false;

    void *w(int i) const { return (char *)work + (size_t)i * align256((size_t)geo.n * rsize); }
};

static int need_ctx(zsc_ctx *c)
{
    if (!c) return fail(ZSC_ERR_ARG, "null context");
    return ZSC_OK;
}

static int set_device(zsc_ctx *c)
{
    ZSC_HIP(hipSetDevice(c->device), "hipSetDevice");
    return ZSC_OK;
}

extern "C" const char *zsc_last_error(void) { return g_last_error.c_str(); }

// Results (10 doubles) and per-block partial sums of the statistics
// reduction: a fixed grid of kStatsBlocks blocks strides over any box.
static constexpr size_t kStatsScratch = (size_t)(kStatsBlocks + 1) * 10;

// ---- the same reduction with its scratch in a file-scope device array ----
// For callers that cannot allocate device memory (code under graph capture,
// serving loops, harnesses that forbid hipMalloc).  Each GPU the library runs
// on holds one copy of `scratch` (kStatsScratch doubles, 82 KB) until the
// process ends.  In host code the array's name is not a device address: each
// device's copy is looked up once, with that device current, when the first
// context on it is created (zsc_create), and only that pointer is used.
static __device__ __align__(256) double scratch[kStatsScratch];
static std::mutex g_scratch_mu;
static std::vector<double *> g_scratch_ptr;  // by device ordinal

static int scratch_for_device(int device, double **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0) return fail(ZSC_ERR_ARG, "bad device %d", device);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = (double *)p;
    }
    *out = g_scratch_ptr[device];
    return ZSC_OK;
}

extern "C" int zsc_create(int nx, int ny, int nz, int precision, void *stream, zsc_ctx **out)
{
    if (!out) return fail(ZSC_ERR_ARG, "zsc_create: null output");
    *out = nullptr;
    if (nx <= 0 || ny <= 0 || nz <= 0) return fail(ZSC_ERR_ARG, "zsc_create: sizes must be positive");
    const int64_t n = (int64_t)nx * ny * nz;
    if (n >= ((int64_t)1 << 31)) return fail(ZSC_ERR_ARG, "zsc_create: more than 2^31-1 cells");
    if (precision < ZSC_F64 || precision > ZSC_F32D)
        return fail(ZSC_ERR_ARG, "zsc_create: unknown precision %d", precision);
    zsc_ctx *c = new zsc_ctx;
    c->geo = {nx, ny, nz, n};
    c->prec = precision;
    c->rsize = precision == ZSC_F64 ? 8 : 4;
    // fused-kernel launch shape per precision, from 512^3 sweeps on MI355X:
    // FP32 gains from the register-held g prefetch (2 blocks per CU, 64
    // planes); the 16-bit lattices, bound by memory-level parallelism, from
    // two cells per lane (32-bit words: twice the bytes per load) where rows
    // are a multiple of 128 cells, else from 6 waves per SIMD; both from the
    // stash of f_5 / f_6 (two fewer global loads per pulled cell, 1-2%
    // faster, FP64 too; the FP32 kernel, bound by its registers, runs 2%
    // slower with it)
    const bool wide = nx % 128 == 0;
    const int stash = kFusedNtStore | kFusedZStash;
    // two-cell kernel: z chunks of at most 24 planes, as equal as the count
    // allows (24 at 512^3 and 256^3; 22 at nz = 128, where 24-plane chunks,
    // the last of 8, run 4-5% slower)
    const int nchunk = (nz + 23) / 24;
    const int zwide = (nz + nchunk - 1) / nchunk;
    if (precision == ZSC_F32 || precision == ZSC_F32D)
        c->cfg = {8, 64, kFusedNtStore | kFusedPrefetch};
    else if (precision == ZSC_BF16)
        c->cfg = wide ? FusedCfg{4, zwide, stash | kFusedWide} : FusedCfg{8, 32, stash};
    else if (precision == ZSC_FP16)
        c->cfg = wide ? FusedCfg{4, zwide, stash | kFusedWide} : FusedCfg{8, 16, stash};
    else
        c->cfg = {4, 16, stash};
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) {
        if (stream) {
            c->stream = (hipStream_t)stream;
        } else {
            e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
            c->own_stream = e == hipSuccess;
        }
    }
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        zsc_destroy(c);
        return hip_fail(e, "zsc_create");
    }
    // this device's copy of the no-allocation scratch (looked up once per
    // device, with the device current)
    if (int r = scratch_for_device(c->device, &c->scratch)) {
        zsc_destroy(c);
        return r;
    }
    *out = c;
    return ZSC_OK;
}

extern "C" int zsc_destroy(zsc_ctx *c)
{
    if (!c) return ZSC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : {c->lat[0], c->lat[1], c->work, (void *)c->dbuf, (void *)c->stats_buf})
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ZSC_OK;
}

// The scalars a step divides by or relaxes with must be finite, and the
// smoothing weight eta must lie in (0, 1) so that invEta2 = 1/(2(1 - eta))
// exists (src/paramread.c:147-148).
static bool params_valid(const zsc_params &p)
{
    if (!(p.eta > 0.0 && p.eta < 1.0) || !isfinite(p.inv_eta2)) return false;
    const double *v = (const double *)&p;
    for (size_t i = 0; i < sizeof(zsc_params) / sizeof(double); ++i)
        if (!isfinite(v[i])) return false;
    return true;
}

extern "C" int zsc_set_params(zsc_ctx *c, const zsc_params *p)
{
    if (int r = need_ctx(c)) return r;
    if (!p) return fail(ZSC_ERR_ARG, "zsc_set_params: null params");
    if (!params_valid(*p))
        return fail(ZSC_ERR_ARG, "zsc_set_params: eta must lie in (0,1) and every scalar be finite");
    c->params = *p;
    c->have_params = true;
    return ZSC_OK;
}

extern "C" int zsc_set_gravity(zsc_ctx *c, int on)
{
    if (int r = need_ctx(c)) return r;
    c->grav_on = on ? 1 : 0;
    return ZSC_OK;
}

extern "C" int zsc_set_fused_config(zsc_ctx *c, int ty, int zchunk, int flags)
{
    if (int r = need_ctx(c)) return r;
    if (!(ty == 1 || ty == 2 || ty == 4 || ty == 8) || zchunk < 0 || flags < 0 ||
        (flags & 0xc0) != 0 || (flags >> 8) > 128 || (flags & 3) == kFusedNtLoad)
        return fail(ZSC_ERR_ARG, "zsc_set_fused_config: ty in {1,2,4,8}, zchunk >= 0, valid flags");
    c->cfg = {ty, zchunk, flags};
    return ZSC_OK;
}

extern "C" int zsc_get_fused_config(zsc_ctx *c, int *ty, int *zchunk, int *flags)
{
    if (int r = need_ctx(c)) return r;
    if (ty) *ty = c->cfg.ty;
    if (zchunk) *zchunk = c->cfg.zchunk;
    if (flags) *flags = c->cfg.flags;
    return ZSC_OK;
}

extern "C" int zsc_sync(zsc_ctx *c)
{
    if (int r = need_ctx(c)) return r;
    ZSC_HIP(hipStreamSynchronize(c->stream), "zsc_sync");
    return ZSC_OK;
}

// ---------------------------------------------------------------------------
// drop-in kernels
// ---------------------------------------------------------------------------
template <class F>
static int dropin(zsc_ctx *c, bool need_params, const char *name, F &&f)
{
    if (int r = need_ctx(c)) return r;
    if (need_params && !c->have_params) return fail(ZSC_ERR_STATE, "%s: zsc_set_params first", name);
    if (int r = set_device(c)) return r;
    hipError_t e;
    if (c->prec == ZSC_F64)
        e = f(Tag<double>{});
    else if (c->prec == ZSC_F32)
        e = f(Tag<float>{});
    else
        return fail(ZSC_ERR_PRECISION, "%s: drop-in kernels need an F64 or F32 context", name);
    if (e != hipSuccess) return hip_fail(e, name);
    return ZSC_OK;
}

template <class T, int Q>
static bool all_set(void *const *p)
{
    if (!p) return false;
    for (int i = 0; i < Q; ++i)
        if (!p[i]) return false;
    return true;
}

#define ZSC_T typename decltype(tag)::type

extern "C" int zsc_init_f(zsc_ctx *c, const void *phi, void *const f[7])
{
    if (!phi || !all_set<void, 7>(f)) return fail(ZSC_ERR_ARG, "zsc_init_f: null pointer");
    return dropin(c, true, "zsc_init_f", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::init_f(c->geo, c->params, c->stream, (const T *)phi, (T *const *)f);
    });
}

extern "C" int zsc_init_g(zsc_ctx *c, const void *phi, const void *rho, void *const g[19])
{
    if (!phi || !rho || !all_set<void, 19>(g)) return fail(ZSC_ERR_ARG, "zsc_init_g: null pointer");
    return dropin(c, true, "zsc_init_g", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::init_g(c->geo, c->params, c->stream, (const T *)phi, (const T *)rho,
                                 (T *const *)g);
    });
}

extern "C" int zsc_update_phi(zsc_ctx *c, void *phi, void *const f[7])
{
    if (!phi || !all_set<void, 7>(f)) return fail(ZSC_ERR_ARG, "zsc_update_phi: null pointer");
    return dropin(c, false, "zsc_update_phi", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::update_phi(c->geo, c->stream, (T *)phi, (T *const *)f);
    });
}

extern "C" int zsc_update_rho(zsc_ctx *c, void *rho, void *const g[19])
{
    if (!rho || !all_set<void, 19>(g)) return fail(ZSC_ERR_ARG, "zsc_update_rho: null pointer");
    return dropin(c, false, "zsc_update_rho", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::update_rho(c->geo, c->stream, (T *)rho, (T *const *)g);
    });
}

extern "C" int zsc_update_velocity(zsc_ctx *c, const void *rho, void *ux, void *uy, void *uz,
                                   void *const g[19])
{
    if (!rho || !ux || !uy || !uz || !g) return fail(ZSC_ERR_ARG, "zsc_update_velocity: null pointer");
    for (int i = 1; i < 19; ++i)
        if (!g[i]) return fail(ZSC_ERR_ARG, "zsc_update_velocity: g[%d] is null", i);
    return dropin(c, false, "zsc_update_velocity", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::update_velocity(c->geo, c->stream, (const T *)rho, (T *)ux, (T *)uy,
                                          (T *)uz, (T *const *)g);
    });
}

extern "C" int zsc_collision_f(zsc_ctx *c, const void *phi, const void *rho, const void *ux,
                               const void *uy, const void *uz, void *const f[7])
{
    if (!phi || !rho || !ux || !uy || !uz || !all_set<void, 7>(f))
        return fail(ZSC_ERR_ARG, "zsc_collision_f: null pointer");
    return dropin(c, true, "zsc_collision_f", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::collision_f(c->geo, c->params, c->grav_on, c->stream, (const T *)phi,
                                      (const T *)rho, (const T *)ux, (const T *)uy,
                                      (const T *)uz, (T *const *)f);
    });
}

extern "C" int zsc_collision_g(zsc_ctx *c, const void *phi, const void *rho, const void *ux,
                               const void *uy, const void *uz, void *const g[19])
{
    if (!phi || !rho || !ux || !uy || !uz || !all_set<void, 19>(g))
        return fail(ZSC_ERR_ARG, "zsc_collision_g: null pointer");
    return dropin(c, true, "zsc_collision_g", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::collision_g(c->geo, c->params, c->grav_on, c->stream, (const T *)phi,
                                      (const T *)rho, (const T *)ux, (const T *)uy,
                                      (const T *)uz, (T *const *)g);
    });
}

extern "C" int zsc_stream_f(zsc_ctx *c, void *const f[7])
{
    if (!all_set<void, 7>(f)) return fail(ZSC_ERR_ARG, "zsc_stream_f: null pointer");
    return dropin(c, true, "zsc_stream_f", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::stream_f(c->geo, c->params, c->stream, (T *const *)f);
    });
}

extern "C" int zsc_stream_g(zsc_ctx *c, void *const g[19])
{
    if (!all_set<void, 19>(g)) return fail(ZSC_ERR_ARG, "zsc_stream_g: null pointer");
    return dropin(c, false, "zsc_stream_g", [&](auto tag) {
        using T = ZSC_T;
        return DropIn<T>::stream_g(c->geo, c->stream, (T *const *)g);
    });
}

// ---------------------------------------------------------------------------
// fast path
// ---------------------------------------------------------------------------
static int ensure_alloc(zsc_ctx *c)
{
    if (c->lat[0]) return ZSC_OK;
    if (int r = set_device(c)) return r;
    size_t lb = 0;
    dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        lb = LatLayout<S>::bytes(c->geo.n);
        return ZSC_OK;
    });
    const size_t wb = 5 * align256((size_t)c->geo.n * c->rsize);
    void *a = nullptr, *b = nullptr, *w = nullptr;
    hipError_t e = hipMalloc(&a, lb);
    if (e == hipSuccess) e = hipMalloc(&b, lb);
    if (e == hipSuccess) e = hipMalloc(&w, wb);
    if (e != hipSuccess) {
        for (void *p : {a, b, w})
            if (p) (void)hipFree(p);
        (void)hipGetLastError();
        return fail(ZSC_ERR_ALLOC, "cannot allocate %.2f GB of lattices and work arrays",
                    (2.0 * lb + wb) / 1e9);
    }
    c->lat[0] = a;
    c->lat[1] = b;
    c->lat_bytes = lb;
    c->work = w;
    return ZSC_OK;
}

static int ensure_stats_buf(zsc_ctx *c)
{
    if (c->stats_buf) return ZSC_OK;
    if (int r = set_device(c)) return r;
    if (hipMalloc((void **)&c->stats_buf, kStatsScratch * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        c->stats_buf = nullptr;
        return fail(ZSC_ERR_ALLOC, "cannot allocate the statistics buffer");
    }
    return ZSC_OK;
}

static int ensure_dbuf(zsc_ctx *c)
{
    if (c->dbuf) return ZSC_OK;
    if (hipMalloc((void **)&c->dbuf, (size_t)c->geo.n * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        c->dbuf = nullptr;
        return fail(ZSC_ERR_ALLOC, "cannot allocate the FP64 staging buffer");
    }
    return ZSC_OK;
}

// The post-stream (A) lattice of the current state, for reading.  A state
// held post-collision (B) is streamed into the other lattice, which the next
// step overwrites anyway; the state itself stays B, so reading moments,
// statistics or populations never changes the trajectory (with 16-bit
// storage, re-storing the streamed lattice would round it once more).
static int a_view(zsc_ctx *c, const void **a)
{
    if (c->state == STATE_NONE) return fail(ZSC_ERR_STATE, "no state: call zsc_state_init first");
    if (c->state == STATE_A) {
        *a = c->lat[c->cur];
        return ZSC_OK;
    }
    void *dst = c->lat[1 - c->cur];
    int r = dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        hipError_t e = Ops<S>::lat_stream(c->geo, c->params, c->stream, c->lat[c->cur], dst);
        return e == hipSuccess ? ZSC_OK : hip_fail(e, "stream");
    });
    if (r) return r;
    *a = dst;
    return ZSC_OK;
}

static int init_from_work(zsc_ctx *c, bool have_rho)
{
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        hipError_t e = Ops<S>::lat_init(c->geo, c->params, c->stream, c->w(0),
                                        have_rho ? c->w(1) : nullptr, c->lat[c->cur]);
        if (e != hipSuccess) return hip_fail(e, "state init");
        c->state = STATE_A;
        return ZSC_OK;
    });
}

extern "C" int zsc_state_init(zsc_ctx *c, const double *phi_host, const double *rho_host)
{
    if (int r = need_ctx(c)) return r;
    if (!phi_host) return fail(ZSC_ERR_ARG, "zsc_state_init: null phi");
    if (!c->have_params) return fail(ZSC_ERR_STATE, "zsc_state_init: zsc_set_params first");
    if (int r = ensure_alloc(c)) return r;
    if (int r = ensure_dbuf(c)) return r;
    const size_t bytes = (size_t)c->geo.n * sizeof(double);
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        ZSC_HIP(hipMemcpyAsync(c->dbuf, phi_host, bytes, hipMemcpyHostToDevice, c->stream), "phi upload");
        ZSC_HIP(Ops<S>::from_f64(c->stream, c->geo.n, c->dbuf, c->w(0)), "phi convert");
        if (rho_host) {
            ZSC_HIP(hipStreamSynchronize(c->stream), "sync");
            ZSC_HIP(hipMemcpyAsync(c->dbuf, rho_host, bytes, hipMemcpyHostToDevice, c->stream), "rho upload");
            ZSC_HIP(Ops<S>::from_f64(c->stream, c->geo.n, c->dbuf, c->w(1)), "rho convert");
        }
        int r = init_from_work(c, rho_host != nullptr);
        if (r) return r;
        ZSC_HIP(hipStreamSynchronize(c->stream), "state init");
        return ZSC_OK;
    });
}

extern "C" int zsc_state_init_bubbles(zsc_ctx *c, const double *bubbles, int n)
{
    if (int r = need_ctx(c)) return r;
    if (n < 0 || (n > 0 && !bubbles)) return fail(ZSC_ERR_ARG, "zsc_state_init_bubbles: bad bubbles");
    if (!c->have_params) return fail(ZSC_ERR_STATE, "zsc_state_init_bubbles: zsc_set_params first");
    if (int r = ensure_alloc(c)) return r;
    double *db = nullptr;
    const size_t bb = (size_t)(n > 0 ? n : 1) * 4 * sizeof(double);
    ZSC_HIP(hipMalloc((void **)&db, bb), "bubble buffer");
    int r = dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        if (n > 0) ZSC_HIP(hipMemcpyAsync(db, bubbles, bb, hipMemcpyHostToDevice, c->stream), "bubbles");
        ZSC_HIP(Ops<S>::init_phi_bubbles(c->geo, c->stream, db, n, c->params.phi_star,
                                         c->params.width, c->w(0)),
                "init phi");
        int rr = init_from_work(c, false);
        if (rr) return rr;
        ZSC_HIP(hipStreamSynchronize(c->stream), "state init");
        return ZSC_OK;
    });
    (void)hipFree(db);
    return r;
}

static int step_impl(zsc_ctx *c, int nsteps)
{
    if (int r = need_ctx(c)) return r;
    if (nsteps < 0) return fail(ZSC_ERR_ARG, "zsc_step: negative step count");
    // the fused kernel addresses each population array with 32-bit byte
    // offsets (f_0 is 4 or 8 bytes per cell at every precision)
    const uint64_t f0_bytes = (uint64_t)c->geo.n * (c->prec == ZSC_F64 ? 8 : 4);
    if (f0_bytes > ((uint64_t)1 << 32))
        return fail(ZSC_ERR_ARG,
                    "zsc_step: %lld cells exceed the fused step's 32-bit population offsets "
                    "(at most %lld cells at this precision); the drop-in kernels take any size",
                    (long long)c->geo.n, (long long)(((uint64_t)1 << 32) / (c->prec == ZSC_F64 ? 8 : 4)));
    if (c->state == STATE_NONE) return fail(ZSC_ERR_STATE, "zsc_step: no state");
    if (nsteps == 0) return ZSC_OK;
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        int left = nsteps;
        if (c->state == STATE_A) {
            // first update + collision from the A lattice, through the phi array
            const int nxt = 1 - c->cur;
            ZSC_HIP(Ops<S>::lat_phi(c->geo, c->stream, c->lat[c->cur], c->w(0)), "phi");
            ZSC_HIP(Ops<S>::lat_collide(c->geo, c->params, c->grav_on, c->stream, c->lat[c->cur],
                                        c->w(0), c->lat[nxt]),
                    "collide");
            c->cur = nxt;
            c->state = STATE_B;
            --left;
        }
        // each fused launch: stream of step k, update + collision of step k+1
        for (; left > 0; --left) {
            const int nxt = 1 - c->cur;
            ZSC_HIP(Ops<S>::lat_fused(c->geo, c->params, c->grav_on, c->stream, c->cfg,
                                      c->lat[c->cur], c->lat[nxt]),
                    "fused step");
            c->cur = nxt;
        }
        return ZSC_OK;
    });
}

extern "C" int zsc_step(zsc_ctx *c, int nsteps) { return step_impl(c, nsteps); }

extern "C" int zsc_step_timed(zsc_ctx *c, int nsteps, double *ms)
{
    if (int r = need_ctx(c)) return r;
    ZSC_HIP(hipEventRecord(c->ev0, c->stream), "event");
    if (int r = step_impl(c, nsteps)) return r;
    ZSC_HIP(hipEventRecord(c->ev1, c->stream), "event");
    ZSC_HIP(hipEventSynchronize(c->ev1), "event sync");
    float t = 0.f;
    ZSC_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1), "elapsed");
    if (ms) *ms = t;
    return ZSC_OK;
}

// Lattice placement.  The fused step's time depends on the physical pages
// behind the two lattices, which hipMalloc does not let a caller choose: at
// 512^3 FP32 the same kernel on the same shape took 4.94 to 5.67 ms over 18
// allocations in one process (steady to 0.1% within one).  Each trial
// allocates a pair, sets a rest state (phi = 0, rho = rho_mean) and times
// four fused steps; the fastest pair is kept.  Rejected pairs are held
// until the end so that the driver cannot hand the same pages back.
static constexpr int kMaxPlacementTrials = 16;

static int time_pair(zsc_ctx *c, void *a, void *b, double *ms)
{
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        ZSC_HIP(hipMemsetAsync(c->w(0), 0, (size_t)c->geo.n * c->rsize, c->stream), "clear");
        ZSC_HIP(Ops<S>::lat_init(c->geo, c->params, c->stream, c->w(0), nullptr, a), "init");
        void *in = a, *out = b;
        ZSC_HIP(Ops<S>::lat_fused(c->geo, c->params, c->grav_on, c->stream, c->cfg, in, out),
                "fused step");
        std::swap(in, out);
        ZSC_HIP(hipEventRecord(c->ev0, c->stream), "event");
        for (int k = 0; k < 4; ++k) {
            ZSC_HIP(Ops<S>::lat_fused(c->geo, c->params, c->grav_on, c->stream, c->cfg, in, out),
                    "fused step");
            std::swap(in, out);
        }
        ZSC_HIP(hipEventRecord(c->ev1, c->stream), "event");
        ZSC_HIP(hipEventSynchronize(c->ev1), "event sync");
        float t = 0.f;
        ZSC_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1), "elapsed");
        *ms = t / 4.0;
        return ZSC_OK;
    });
}

extern "C" int zsc_place_lattices(zsc_ctx *c, int trials, double *ms)
{
    if (int r = need_ctx(c)) return r;
    if (trials < 1 || trials > kMaxPlacementTrials)
        return fail(ZSC_ERR_ARG, "zsc_place_lattices: trials must be 1..%d", kMaxPlacementTrials);
    if (!c->have_params) return fail(ZSC_ERR_STATE, "zsc_place_lattices: zsc_set_params first");
    if (c->lat[0])
        return fail(ZSC_ERR_STATE, "zsc_place_lattices: the lattices exist already (place them "
                                   "before the state is set)");
    if ((uint64_t)c->geo.n * (c->prec == ZSC_F64 ? 8 : 4) > ((uint64_t)1 << 32))
        return fail(ZSC_ERR_ARG, "zsc_place_lattices: box too large for the fused kernel");
    if (int r = ensure_alloc(c)) return r;
    for (int t = 0; t < trials; ++t)
        if (ms) ms[t] = 0.0;
    const size_t lb = c->lat_bytes;
    std::vector<void *> held;
    double best = 0.0;
    int r = ZSC_OK;
    for (int t = 0; t < trials && r == ZSC_OK; ++t) {
        void *a = c->lat[0], *b = c->lat[1];
        if (t > 0) {
            // another pair only while it leaves a quarter of the free memory
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr < 2 * lb + fr / 4) {
                (void)hipGetLastError();
                break;
            }
            a = b = nullptr;
            if (hipMalloc(&a, lb) != hipSuccess || hipMalloc(&b, lb) != hipSuccess) {
                (void)hipGetLastError();
                if (a) (void)hipFree(a);
                if (b) (void)hipFree(b);
                break;
            }
        }
        double tm = 0.0;
        r = time_pair(c, a, b, &tm);
        if (r != ZSC_OK) {
            if (t > 0) {
                held.push_back(a);
                held.push_back(b);
            }
            break;
        }
        if (ms) ms[t] = tm;
        if (t == 0 || tm < best) {
            if (t > 0) {
                held.push_back(c->lat[0]);
                held.push_back(c->lat[1]);
                c->lat[0] = a;
                c->lat[1] = b;
            }
            best = tm;
        } else {
            held.push_back(a);
            held.push_back(b);
        }
    }
    for (void *p : held) (void)hipFree(p);
    return r;
}

extern "C" int zsc_state_get(zsc_ctx *c, double *f_host, double *g_host)
{
    if (int r = need_ctx(c)) return r;
    const void *a = nullptr;
    if (int r = a_view(c, &a)) return r;
    if (int r = ensure_dbuf(c)) return r;
    const size_t bytes = (size_t)c->geo.n * sizeof(double);
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        for (int q = 0; q < 26; ++q) {
            double *dst = q < 7 ? (f_host ? f_host + (size_t)q * c->geo.n : nullptr)
                                : (g_host ? g_host + (size_t)(q - 7) * c->geo.n : nullptr);
            if (!dst) continue;
            ZSC_HIP(Ops<S>::pop_get(c->geo, c->params, c->stream, a, q, c->dbuf), "decode");
            ZSC_HIP(hipMemcpyAsync(dst, c->dbuf, bytes, hipMemcpyDeviceToHost, c->stream), "download");
            ZSC_HIP(hipStreamSynchronize(c->stream), "download");
        }
        return ZSC_OK;
    });
}

extern "C" int zsc_state_set(zsc_ctx *c, const double *f_host, const double *g_host)
{
    if (int r = need_ctx(c)) return r;
    if (!f_host || !g_host) return fail(ZSC_ERR_ARG, "zsc_state_set: null arrays");
    if (!c->have_params) return fail(ZSC_ERR_STATE, "zsc_state_set: zsc_set_params first");
    if (int r = ensure_alloc(c)) return r;
    if (int r = ensure_dbuf(c)) return r;
    const size_t bytes = (size_t)c->geo.n * sizeof(double);
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        for (int q = 0; q < 26; ++q) {
            const double *src = q < 7 ? f_host + (size_t)q * c->geo.n : g_host + (size_t)(q - 7) * c->geo.n;
            ZSC_HIP(hipMemcpyAsync(c->dbuf, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
            ZSC_HIP(Ops<S>::pop_set(c->geo, c->params, c->stream, c->lat[c->cur], q, c->dbuf), "encode");
            ZSC_HIP(hipStreamSynchronize(c->stream), "upload");
        }
        c->state = STATE_A;
        return ZSC_OK;
    });
}

static int moments_to_work(zsc_ctx *c)
{
    const void *a = nullptr;
    if (int r = a_view(c, &a)) return r;
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        ZSC_HIP(Ops<S>::lat_moments(c->geo, c->params, c->stream, a, c->w(0), c->w(1),
                                    c->w(2), c->w(3), c->w(4)),
                "moments");
        return ZSC_OK;
    });
}

extern "C" int zsc_state_moments(zsc_ctx *c, double *phi, double *rho, double *ux, double *uy,
                                 double *uz)
{
    if (int r = need_ctx(c)) return r;
    if (int r = moments_to_work(c)) return r;
    if (int r = ensure_dbuf(c)) return r;
    const size_t bytes = (size_t)c->geo.n * sizeof(double);
    double *outs[5] = {phi, rho, ux, uy, uz};
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        for (int i = 0; i < 5; ++i) {
            if (!outs[i]) continue;
            ZSC_HIP(Ops<S>::to_f64(c->stream, c->geo.n, c->w(i), c->dbuf), "convert");
            ZSC_HIP(hipMemcpyAsync(outs[i], c->dbuf, bytes, hipMemcpyDeviceToHost, c->stream), "download");
            ZSC_HIP(hipStreamSynchronize(c->stream), "download");
        }
        return ZSC_OK;
    });
}

// Reduces five moment arrays (element type: the compute type of storage S)
// with the statistics kernels, using buf (kStatsScratch doubles: results,
// then per-block partials), and copies the ten results to h.  Every partial
// and result is written before it is read: buf needs no clearing.
template <class S>
static hipError_t reduce_stats(zsc_ctx *c, const void *phi, const void *rho, const void *ux,
                               const void *uy, const void *uz, double *buf, double h[10])
{
    hipError_t e = Ops<S>::stats(c->geo, c->params, c->stream, phi, rho, ux, uy, uz, buf + 10, buf);
    if (e == hipSuccess) e = hipMemcpyAsync(h, buf, 10 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

static int stats_out(const double h[10], zsc_stats *out)
{
    out->sum_phi = h[0];
    out->sum_rho = h[1];
    out->bubble_volume = h[2];
    out->ux_sum = h[3];
    out->uy_sum = h[4];
    out->uz_sum = h[5];
    out->umax_sq = h[6];
    out->p_in_sum = h[7];
    out->p_out_sum = h[8];
    out->bubble_volume_full = h[9];
    if (!isfinite(h[0]) || !isfinite(h[1]) || !isfinite(h[6]))
        return fail(ZSC_ERR_DIVERGED, "non-finite conserved sum (sum phi %g, sum rho %g)", h[0], h[1]);
    return ZSC_OK;
}

extern "C" int zsc_state_stats(zsc_ctx *c, zsc_stats *out)
{
    if (int r = need_ctx(c)) return r;
    if (!out) return fail(ZSC_ERR_ARG, "zsc_state_stats: null output");
    if (int r = moments_to_work(c)) return r;
    if (int r = ensure_stats_buf(c)) return r;
    double h[10];
    int r = dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        ZSC_HIP(reduce_stats<S>(c, c->w(0), c->w(1), c->w(2), c->w(3), c->w(4), c->stats_buf, h),
                "stats");
        return ZSC_OK;
    });
    if (r) return r;
    return stats_out(h, out);
}

static int moment_stats(zsc_ctx *c, const void *phi, const void *rho, const void *ux,
                        const void *uy, const void *uz, zsc_stats *out, double *buf, const char *name)
{
    double h[10];
    int r = dropin(c, true, name, [&](auto tag) {
        using T = ZSC_T;
        return reduce_stats<T>(c, phi, rho, ux, uy, uz, buf, h);
    });
    if (r) return r;
    return stats_out(h, out);
}

static bool moments_set(const void *phi, const void *rho, const void *ux, const void *uy,
                        const void *uz, const zsc_stats *out)
{
    return phi && rho && ux && uy && uz && out;
}

extern "C" int zsc_moment_stats(zsc_ctx *c, const void *phi, const void *rho, const void *ux,
                                const void *uy, const void *uz, zsc_stats *out)
{
    if (int r = need_ctx(c)) return r;
    if (!moments_set(phi, rho, ux, uy, uz, out)) return fail(ZSC_ERR_ARG, "zsc_moment_stats: null pointer");
    if (c->prec != ZSC_F64 && c->prec != ZSC_F32)
        return fail(ZSC_ERR_PRECISION, "zsc_moment_stats: drop-in arrays need an F64 or F32 context");
    if (int r = ensure_stats_buf(c)) return r;
    return moment_stats(c, phi, rho, ux, uy, uz, out, c->stats_buf, "zsc_moment_stats");
}


template <bool Export>
static int xport(zsc_ctx *c, void *const *f, void *const *g, const char *name)
{
    if (int r = need_ctx(c)) return r;
    if (!all_set<void, 7>(f) || !all_set<void, 19>(g)) return fail(ZSC_ERR_ARG, "%s: null pointer", name);
    const void *a = nullptr;
    if (Export) {
        if (int r = a_view(c, &a)) return r;
    } else {
        if (!c->have_params) return fail(ZSC_ERR_STATE, "%s: zsc_set_params first", name);
        if (int r = ensure_alloc(c)) return r;
    }
    auto go = [&](auto tag, const zsc_params *dev) -> hipError_t {
        using T = ZSC_T;
        if (Export)
            return DropIn<T>::export_state(c->geo, c->stream, a, (T *const *)f, (T *const *)g, dev);
        hipError_t e = DropIn<T>::import_state(c->geo, c->stream, c->lat[c->cur], (T *const *)f,
                                               (T *const *)g, dev);
        if (e == hipSuccess) c->state = STATE_A;
        return e;
    };
    if (c->prec == ZSC_F32D) {
        // float reference arrays; the lattice's g are deviations from w_i rho_mean
        if (!c->have_params) return fail(ZSC_ERR_STATE, "%s: zsc_set_params first", name);
        if (int r = set_device(c)) return r;
        const hipError_t e = go(Tag<float>{}, &c->params);
        return e == hipSuccess ? ZSC_OK : hip_fail(e, name);
    }
    return dropin(c, false, name, [&](auto tag) { return go(tag, nullptr); });
}

extern "C" int zsc_state_export(zsc_ctx *c, void *const f[7], void *const g[19])
{
    return xport<true>(c, f, g, "zsc_state_export");
}

extern "C" int zsc_state_import(zsc_ctx *c, void *const f[7], void *const g[19])
{
    return xport<false>(c, f, g, "zsc_state_import");
}

extern "C" int zsc_moment_stats_noalloc(zsc_ctx *c, const void *phi, const void *rho,
                                        const void *ux, const void *uy, const void *uz,
                                        zsc_stats *out)
{
    if (int r = need_ctx(c)) return r;
    if (!moments_set(phi, rho, ux, uy, uz, out))
        return fail(ZSC_ERR_ARG, "zsc_moment_stats_noalloc: null pointer");
    if (c->prec != ZSC_F64 && c->prec != ZSC_F32)
        return fail(ZSC_ERR_PRECISION, "zsc_moment_stats_noalloc: drop-in arrays need an F64 or F32 context");
    // the reduction's grid is fixed, so the scratch it needs does not grow
    // with the box: every box a context accepts fits
    static_assert(sizeof(scratch) / sizeof(double) >= kStatsScratch, "scratch too small");
    if (!c->scratch)
        return fail(ZSC_ERR_STATE, "zsc_moment_stats_noalloc: no scratch address for device %d", c->device);
    return moment_stats(c, phi, rho, ux, uy, uz, out, c->scratch, "zsc_moment_stats_noalloc");
}

// ---------------------------------------------------------------------------
// copy ceiling: plain copies between the context's two lattices
// ---------------------------------------------------------------------------
typedef float zsc_v4f __attribute__((ext_vector_type(4)));

// Linear copy: each lane loads four 16-byte vectors of its block's 16 KiB,
// then stores them.
__global__ void __launch_bounds__(256) k_copy_linear(const zsc_v4f *__restrict__ in,
                                                     zsc_v4f *__restrict__ out)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    zsc_v4f v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = __builtin_nontemporal_load(in + base + k * 256);
#pragma unroll
    for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(v[k], out + base + k * 256);
}

__global__ void __launch_bounds__(256) k_copy_tail(const unsigned char *in, unsigned char *out,
                                                   size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        out[i] = in[i];
}

// The fused step's memory pattern without its pulls or arithmetic: tiles of
// 64 lane-words x TY rows march a z chunk and move the 26 populations of
// each cell (a lane-word holds one cell, or two for 16-bit storage: W0 is
// the f_0 word, W1 the word of the other 25 rows).  XCD-contiguous tile
// order, non-temporal stores, as in the fused kernel.
template <class W0, class W1, int TY>
__global__ void __launch_bounds__(64 * TY)
    k_copy_march(const W0 *__restrict__ in0, W0 *__restrict__ out0, const W1 *__restrict__ in1,
                 W1 *__restrict__ out1, int nxw, int ny, int nz, int zchunk, int nbx, int nby,
                 int nblk, int64_t row1)
{
    int l = blockIdx.x;
    if ((nblk & 7) == 0) l = (l & 7) * (nblk >> 3) + (l >> 3);
    const int bx = l % nbx, by = (l / nbx) % nby, bz = l / (nbx * nby);
    const int x = bx * 64 + (int)threadIdx.x, y = by * TY + (int)threadIdx.y;
    if (x >= nxw || y >= ny) return;
    const int z1 = min(nz, (bz + 1) * zchunk);
    for (int z = bz * zchunk; z < z1; ++z) {
        const int64_t o = x + (int64_t)nxw * (y + (int64_t)ny * z);
        const W0 a = in0[o];
        W1 v[25];
#pragma unroll
        for (int q = 0; q < 25; ++q) v[q] = in1[o + q * row1];
        __builtin_nontemporal_store(a, out0 + o);
#pragma unroll
        for (int q = 0; q < 25; ++q) __builtin_nontemporal_store(v[q], out1 + o + q * row1);
    }
}

template <class W0, class W1, int CPL>
static hipError_t launch_copy_march(const Geo &g, hipStream_t s, size_t f_off, const void *a, void *b)
{
    constexpr int TY = CPL == 1 ? 8 : 4;
    const int nxw = g.nx / CPL, zchunk = 16;
    const int nbx = (nxw + 63) / 64, nby = (g.ny + TY - 1) / TY, nbz = (g.nz + zchunk - 1) / zchunk;
    const int64_t nblk = (int64_t)nbx * nby * nbz;
    if (nblk > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_copy_march<W0, W1, TY>), dim3((unsigned)nblk), dim3(64, TY), 0, s,
                       (const W0 *)a, (W0 *)b, (const W1 *)((const char *)a + f_off),
                       (W1 *)((char *)b + f_off), nxw, g.ny, g.nz, zchunk, nbx, nby, (int)nblk,
                       (int64_t)(g.n / CPL));
    return hipGetLastError();
}

extern "C" int zsc_copy_ceiling(zsc_ctx *c, double ms[2])
{
    if (int r = need_ctx(c)) return r;
    if (!ms) return fail(ZSC_ERR_ARG, "zsc_copy_ceiling: null output");
    if (c->state != STATE_NONE)
        return fail(ZSC_ERR_STATE, "zsc_copy_ceiling: overwrites the lattices; call it before the state is set");
    if (int r = ensure_alloc(c)) return r;
    const void *a = c->lat[0];
    void *b = c->lat[1];
    const size_t bytes = c->lat_bytes, chunks = bytes / 16384, done = chunks * 16384;
    auto linear = [&]() {
        if (chunks) hipLaunchKernelGGL(k_copy_linear, dim3((unsigned)chunks), dim3(256), 0, c->stream,
                                       (const zsc_v4f *)a, (zsc_v4f *)b);
        if (bytes > done)
            hipLaunchKernelGGL(k_copy_tail, dim3(64), dim3(256), 0, c->stream,
                               (const unsigned char *)a + done, (unsigned char *)b + done, bytes - done);
        return hipGetLastError();
    };
    const bool wide = c->prec == ZSC_BF16 || c->prec == ZSC_FP16;
    auto march = [&]() {
        if (c->prec == ZSC_F64) return launch_copy_march<double, double, 1>(c->geo, c->stream, c->geo.n * 8, a, b);
        if (c->prec == ZSC_F32 || c->prec == ZSC_F32D)
            return launch_copy_march<float, float, 1>(c->geo, c->stream, c->geo.n * 4, a, b);
        return launch_copy_march<uint64_t, uint32_t, 2>(c->geo, c->stream, c->geo.n * 4, a, b);
    };
    if (chunks > 0x7fffffff) return fail(ZSC_ERR_ARG, "zsc_copy_ceiling: lattice too large");
    const bool can_march = !wide || c->geo.nx % 2 == 0;
    for (int k = 0; k < 2; ++k) {
        ms[k] = 0.0;
        if (k == 1 && !can_march) continue;
        auto run = [&]() { return k == 0 ? linear() : march(); };
        ZSC_HIP(run(), "copy");  // warm-up
        ZSC_HIP(hipEventRecord(c->ev0, c->stream), "event");
        for (int r = 0; r < 5; ++r) ZSC_HIP(run(), "copy");
        ZSC_HIP(hipEventRecord(c->ev1, c->stream), "event");
        ZSC_HIP(hipEventSynchronize(c->ev1), "event sync");
        float t = 0.f;
        ZSC_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1), "elapsed");
        ms[k] = t / 5.0;
    }
    return ZSC_OK;
}

// ---------------------------------------------------------------------------
// pressure, VTK output and restart files
// ---------------------------------------------------------------------------
// Moments of the A state with the pressure over the density: w(0) phi,
// w(1) pressure, w(2..4) u.
static int moments_and_pressure(zsc_ctx *c)
{
    if (int r = moments_to_work(c)) return r;
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        ZSC_HIP(Ops<S>::pressure(c->geo, c->params, c->stream, c->w(0), c->w(1)), "pressure");
        return ZSC_OK;
    });
}

extern "C" int zsc_state_pressure(zsc_ctx *c, double *p_host)
{
    if (int r = need_ctx(c)) return r;
    if (!p_host) return fail(ZSC_ERR_ARG, "zsc_state_pressure: null output");
    if (int r = moments_and_pressure(c)) return r;
    if (int r = ensure_dbuf(c)) return r;
    return dispatch(c->prec, [&](auto tag) {
        using S = ZSC_T;
        ZSC_HIP(Ops<S>::to_f64(c->stream, c->geo.n, c->w(1), c->dbuf), "convert");
        ZSC_HIP(hipMemcpyAsync(p_host, c->dbuf, (size_t)c->geo.n * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream),
                "download");
        ZSC_HIP(hipStreamSynchronize(c->stream), "download");
        return ZSC_OK;
    });
}

namespace {

// Big-endian float32 for legacy VTK BINARY sections.
inline uint32_t be_f32(double v)
{
    const float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    return __builtin_bswap32(u);
}

template <class R>
double as_double(const char *base, int64_t i)
{
    R v;
    memcpy(&v, base + i * (int64_t)sizeof(R), sizeof(R));
    return (double)v;
}

struct FileCloser {
    FILE *fp;
    ~FileCloser()
    {
        if (fp) fclose(fp);
    }
};

}  // namespace

// Writes fields of the work arrays in slabs of whole z planes: ncomp = 1 for
// one scalar (array i0), 3 for the interleaved vector (arrays i0..i0+2).
template <class R>
static int vtk_field(zsc_ctx *c, FILE *fp, int i0, int ncomp, int binary)
{
    const int64_t n = c->geo.n, plane = (int64_t)c->geo.nx * c->geo.ny;
    int64_t slab = ((int64_t)1 << 22) / plane;
    if (slab < 1) slab = 1;
    const int64_t cells_max = slab * plane;
    std::string host((size_t)(cells_max * ncomp) * sizeof(R), '\0');
    std::string out;
    for (int64_t c0 = 0; c0 < n; c0 += cells_max) {
        const int64_t m = std::min(cells_max, n - c0);
        for (int k = 0; k < ncomp; ++k)
            ZSC_HIP(hipMemcpyAsync(&host[(size_t)(k * cells_max) * sizeof(R)],
                                   (const char *)c->w(i0 + k) + c0 * (int64_t)sizeof(R),
                                   (size_t)m * sizeof(R), hipMemcpyDeviceToHost, c->stream),
                    "vtk download");
        ZSC_HIP(hipStreamSynchronize(c->stream), "vtk download");
        const char *hb = host.data();
        if (binary) {
            out.resize((size_t)(m * ncomp) * 4);
            uint32_t *o = (uint32_t *)&out[0];
            for (int64_t i = 0; i < m; ++i)
                for (int k = 0; k < ncomp; ++k)
                    o[i * ncomp + k] = be_f32(as_double<R>(hb, k * cells_max + i));
            if (fwrite(out.data(), 1, out.size(), fp) != out.size())
                return fail(ZSC_ERR_IO, "zsc_vtk_save: write failed");
        } else {
            out.clear();
            char buf[128];
            for (int64_t i = 0; i < m; ++i) {
                int len;
                if (ncomp == 1)
                    len = snprintf(buf, sizeof buf, "%f\n", as_double<R>(hb, i));
                else
                    len = snprintf(buf, sizeof buf, "%f %f %f\n", as_double<R>(hb, i),
                                   as_double<R>(hb, cells_max + i), as_double<R>(hb, 2 * cells_max + i));
                out.append(buf, (size_t)len);
            }
            if (fwrite(out.data(), 1, out.size(), fp) != out.size())
                return fail(ZSC_ERR_IO, "zsc_vtk_save: write failed");
        }
    }
    return ZSC_OK;
}

extern "C" int zsc_vtk_save(zsc_ctx *c, const char *path, int binary)
{
    if (int r = need_ctx(c)) return r;
    if (!path) return fail(ZSC_ERR_ARG, "zsc_vtk_save: null path");
    if (int r = moments_and_pressure(c)) return r;
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(ZSC_ERR_IO, "zsc_vtk_save: cannot open %s", path);
    FileCloser closer{fp};
    const Geo &g = c->geo;
    // header of src/vtksave.c:44-57 (one rank: the z origin is plane 1)
    fprintf(fp, "# vtk DataFile Version 2.0\n");
    fprintf(fp, "libzsc ZSC bubbly flow\n");
    fprintf(fp, binary ? "BINARY\n" : "ASCII\n");
    fprintf(fp, "\nDATASET STRUCTURED_POINTS\n");
    fprintf(fp, "DIMENSIONS %d %d %d\n", g.nx, g.ny, g.nz);
    fprintf(fp, "ORIGIN %d %d %d\n", 1, 1, 1);
    fprintf(fp, "SPACING %d %d %d\n", 1, 1, 1);
    fprintf(fp, "\nPOINT_DATA %lld\n", (long long)g.n);
    auto body = [&](auto rtag) {
        using R = typename decltype(rtag)::type;
        fprintf(fp, "\nSCALARS Phi float\nLOOKUP_TABLE default\n");
        if (int r = vtk_field<R>(c, fp, 0, 1, binary)) return r;
        fprintf(fp, "\nSCALARS Pressure float\nLOOKUP_TABLE default\n");
        if (int r = vtk_field<R>(c, fp, 1, 1, binary)) return r;
        fprintf(fp, "\nVectors Velocity float\n");
        return vtk_field<R>(c, fp, 2, 3, binary);
    };
    int r = c->prec == ZSC_F64 ? body(Tag<double>{}) : body(Tag<float>{});
    if (r) return r;
    if (binary) fputc('\n', fp);
    if (ferror(fp)) return fail(ZSC_ERR_IO, "zsc_vtk_save: write failed");
    return ZSC_OK;
}

namespace {
// Restart file header, host byte order, followed by the lattice bytes.
// Version 2 appends up to ZSC_RESTART_USER_MAX caller values; version 1
// files (no caller values) still load.
struct RestartHeaderV1 {
    char magic[8];
    int32_t version, precision;
    int32_t nx, ny, nz;
    int32_t state, grav_on, pad;
    int64_t step;
    uint64_t lattice_bytes;
    zsc_params params;
};
struct RestartHeader : RestartHeaderV1 {
    int32_t n_user, pad2;
    double user[ZSC_RESTART_USER_MAX];
};
const char kRestartMagic[8] = {'Z', 'S', 'C', 'S', 'T', 'A', 'T', 'E'};
const int64_t kRestartChunk = (int64_t)64 << 20;
}  // namespace

extern "C" int zsc_state_save(zsc_ctx *c, const char *path, int64_t step)
{
    return zsc_state_save_ex(c, path, step, nullptr, 0);
}

extern "C" int zsc_state_load(zsc_ctx *c, const char *path, int64_t *step)
{
    return zsc_state_load_ex(c, path, step, nullptr, 0, nullptr);
}

extern "C" int zsc_state_save_ex(zsc_ctx *c, const char *path, int64_t step, const double *user,
                                 int n_user)
{
    if (int r = need_ctx(c)) return r;
    if (!path) return fail(ZSC_ERR_ARG, "zsc_state_save: null path");
    if (n_user < 0 || n_user > ZSC_RESTART_USER_MAX || (n_user > 0 && !user))
        return fail(ZSC_ERR_ARG, "zsc_state_save: 0..%d caller values", ZSC_RESTART_USER_MAX);
    if (c->state == STATE_NONE) return fail(ZSC_ERR_STATE, "zsc_state_save: no state");
    RestartHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, kRestartMagic, 8);
    h.version = 2;
    h.n_user = n_user;
    for (int i = 0; i < n_user; ++i) h.user[i] = user[i];
    h.precision = c->prec;
    h.nx = c->geo.nx;
    h.ny = c->geo.ny;
    h.nz = c->geo.nz;
    h.state = c->state;
    h.grav_on = c->grav_on;
    h.step = step;
    h.lattice_bytes = c->lat_bytes;
    h.params = c->params;
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(ZSC_ERR_IO, "zsc_state_save: cannot open %s", path);
    FileCloser closer{fp};
    if (fwrite(&h, sizeof h, 1, fp) != 1) return fail(ZSC_ERR_IO, "zsc_state_save: write failed");
    void *pin = nullptr;
    ZSC_HIP(hipHostMalloc(&pin, (size_t)kRestartChunk, hipHostMallocDefault), "pinned buffer");
    int r = ZSC_OK;
    const char *src = (const char *)c->lat[c->cur];
    for (int64_t off = 0; off < (int64_t)c->lat_bytes && r == ZSC_OK; off += kRestartChunk) {
        const size_t m = (size_t)std::min(kRestartChunk, (int64_t)c->lat_bytes - off);
        hipError_t e = hipMemcpyAsync(pin, src + off, m, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess)
            r = hip_fail(e, "zsc_state_save: download");
        else if (fwrite(pin, 1, m, fp) != m)
            r = fail(ZSC_ERR_IO, "zsc_state_save: write failed");
    }
    (void)hipHostFree(pin);
    return r;
}

extern "C" int zsc_state_load_ex(zsc_ctx *c, const char *path, int64_t *step, double *user,
                                 int capacity, int *n_user)
{
    if (int r = need_ctx(c)) return r;
    if (!path) return fail(ZSC_ERR_ARG, "zsc_state_load: null path");
    if (capacity < 0 || (capacity > 0 && !user)) return fail(ZSC_ERR_ARG, "zsc_state_load: bad user buffer");
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(ZSC_ERR_IO, "zsc_state_load: cannot open %s", path);
    FileCloser closer{fp};
    RestartHeader h;
    memset(&h, 0, sizeof h);
    RestartHeaderV1 &h1 = h;
    if (fread(&h1, sizeof h1, 1, fp) != 1 || memcmp(h.magic, kRestartMagic, 8) != 0 ||
        (h.version != 1 && h.version != 2))
        return fail(ZSC_ERR_IO, "zsc_state_load: %s is not a libzsc restart file", path);
    if (h.version == 2) {
        const size_t tail = sizeof h - sizeof h1;
        if (fread((char *)&h + sizeof h1, tail, 1, fp) != 1 || h.n_user < 0 ||
            h.n_user > ZSC_RESTART_USER_MAX)
            return fail(ZSC_ERR_IO, "zsc_state_load: %s has a damaged header", path);
    }
    if (!params_valid(h.params))
        return fail(ZSC_ERR_IO, "zsc_state_load: %s holds invalid parameters (eta outside (0,1) "
                                "or a non-finite scalar)", path);
    if (h.precision != c->prec || h.nx != c->geo.nx || h.ny != c->geo.ny || h.nz != c->geo.nz)
        return fail(ZSC_ERR_ARG,
                    "zsc_state_load: file holds a %dx%dx%d box at precision %d, context is "
                    "%dx%dx%d at %d",
                    h.nx, h.ny, h.nz, h.precision, c->geo.nx, c->geo.ny, c->geo.nz, c->prec);
    if (h.state != STATE_A && h.state != STATE_B)
        return fail(ZSC_ERR_IO, "zsc_state_load: bad state kind %d", h.state);
    if (int r = ensure_alloc(c)) return r;
    if (h.lattice_bytes != c->lat_bytes)
        return fail(ZSC_ERR_IO, "zsc_state_load: lattice size %llu, expected %llu",
                    (unsigned long long)h.lattice_bytes, (unsigned long long)c->lat_bytes);
    void *pin = nullptr;
    ZSC_HIP(hipHostMalloc(&pin, (size_t)kRestartChunk, hipHostMallocDefault), "pinned buffer");
    int r = ZSC_OK;
    char *dst = (char *)c->lat[0];
    for (int64_t off = 0; off < (int64_t)c->lat_bytes && r == ZSC_OK; off += kRestartChunk) {
        const size_t m = (size_t)std::min(kRestartChunk, (int64_t)c->lat_bytes - off);
        if (fread(pin, 1, m, fp) != m) {
            r = fail(ZSC_ERR_IO, "zsc_state_load: %s is truncated", path);
            break;
        }
        hipError_t e = hipMemcpyAsync(dst + off, pin, m, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) r = hip_fail(e, "zsc_state_load: upload");
    }
    (void)hipHostFree(pin);
    if (r) {
        c->state = STATE_NONE;
        return r;
    }
    c->cur = 0;
    c->state = h.state;
    c->params = h.params;
    c->have_params = true;
    c->grav_on = h.grav_on;
    if (step) *step = h.step;
    for (int i = 0; i < std::min(capacity, (int)h.n_user); ++i) user[i] = h.user[i];
    if (n_user) *n_user = h.n_user;
    return ZSC_OK;
}

// ---------------------------------------------------------------------------
// adjoint (F64 / F32 contexts; plain (26, N) lattices, device pointers)
// ---------------------------------------------------------------------------
extern "C" size_t zsc_adj_scratch_bytes(zsc_ctx *c)
{
    if (!c) return 0;
    const size_t n = (size_t)c->geo.n;
    // forward: B lattice (26n) + phi (n); backward: phi (n) + Adj scratch
    const size_t fwd = align256(26 * n * c->rsize) + align256(n * c->rsize);
    const size_t adj = c->prec != ZSC_F64 ? Adj<float>::scratch_bytes(c->geo.n)
                       : c->adj_lam_f32   ? Adj<double, float>::scratch_bytes(c->geo.n)
                                          : Adj<double>::scratch_bytes(c->geo.n);
    const size_t bwd = align256(n * c->rsize) + adj;
    return fwd > bwd ? fwd : bwd;
}

extern "C" int zsc_adj_set_lambda_precision(zsc_ctx *c, int precision)
{
    if (int r = need_ctx(c)) return r;
    if (c->adj_deferred)
        return fail(ZSC_ERR_STATE, "zsc_adj_set_lambda_precision: deferred coefficient sums are kept");
    if (precision == c->prec && (c->prec == ZSC_F64 || c->prec == ZSC_F32)) {
        c->adj_lam_f32 = false;
        return ZSC_OK;
    }
    if (c->prec == ZSC_F32D && (precision == ZSC_F32 || precision == ZSC_F32D)) {
        c->adj_lam_f32 = false;  // FP32 adjoint lattices over FP32 deviation states
        return ZSC_OK;
    }
    if (c->prec == ZSC_F64 && precision == ZSC_F32) {
        c->adj_lam_f32 = true;
        return ZSC_OK;
    }
    return fail(ZSC_ERR_PRECISION, "zsc_adj_set_lambda_precision: adjoint lattices at precision %d "
                                   "over a context at %d (F64 or F32 as the context, F32 over F64 "
                                   "or over F32D)",
                precision, c->prec);
}

// f(tag of the context's lattice type): double, float, or f32d (FP32
// lattices whose g are deviations from w_i rho_mean, same (26, N) layout).
template <class F>
static int adj_call(zsc_ctx *c, const char *name, F &&f)
{
    if (int r = need_ctx(c)) return r;
    if (!c->have_params) return fail(ZSC_ERR_STATE, "%s: zsc_set_params first", name);
    if (c->prec == ZSC_F32D) {
        if (int r = set_device(c)) return r;
        const hipError_t e = f(Tag<f32d>{});
        return e == hipSuccess ? ZSC_OK : hip_fail(e, name);
    }
    return dropin(c, false, name, std::forward<F>(f));
}

// adj_call for the entry points that take adjoint lattices: f(state tag,
// adjoint-lattice tag).  F32D states run the FP32 adjoint, told by
// adj_dev(c) that their g are deviations.
template <class F>
static int adj_lam_call(zsc_ctx *c, const char *name, F &&f)
{
    return adj_call(c, name, [&](auto tag) -> hipError_t {
        using T = ZSC_T;
        if constexpr (std::is_same<T, f32d>::value) {
            return f(Tag<float>{}, Tag<float>{});
        } else {
            if constexpr (std::is_same<T, double>::value)
                if (c->adj_lam_f32) return f(tag, Tag<float>{});
            return f(tag, tag);
        }
    });
}

static bool adj_dev(const zsc_ctx *c) { return c->prec == ZSC_F32D; }

extern "C" int zsc_adj_init(zsc_ctx *c, const void *phi0, void *a0)
{
    if (!phi0 || !a0) return fail(ZSC_ERR_ARG, "zsc_adj_init: null pointer");
    return adj_call(c, "zsc_adj_init", [&](auto tag) {
        using T = ZSC_T;
        return Ops<T>::lat_init(c->geo, c->params, c->stream, phi0, nullptr, a0);
    });
}

extern "C" int zsc_adj_forward(zsc_ctx *c, const void *a_in, void *a_out, void *scratch)
{
    if (c && c->adj_deferred)
        return fail(ZSC_ERR_STATE, "zsc_adj_forward: the scratch buffer holds deferred coefficient "
                                   "sums; call zsc_adj_flush first");
    if (!a_in || !a_out || !scratch || a_in == a_out)
        return fail(ZSC_ERR_ARG, "zsc_adj_forward: null or aliased pointer");
    return adj_call(c, "zsc_adj_forward", [&](auto tag) {
        using T = ZSC_T;
        const size_t n = (size_t)c->geo.n;
        void *B = scratch;
        void *phi = (char *)scratch + align256(26 * n * sizeof(T));
        hipError_t e = Ops<T>::lat_phi(c->geo, c->stream, a_in, phi);
        if (e == hipSuccess)
            e = Ops<T>::lat_collide(c->geo, c->params, c->grav_on, c->stream, a_in, phi, B);
        if (e == hipSuccess) e = Ops<T>::lat_stream(c->geo, c->params, c->stream, B, a_out);
        return e;
    });
}

extern "C" int zsc_adj_seed(zsc_ctx *c, const void *a_n, const void *phi_tgt, void *lam,
                            void *scratch, double *grads)
{
    if (!a_n || !phi_tgt || !lam || !scratch || !grads) return fail(ZSC_ERR_ARG, "zsc_adj_seed: null pointer");
    return adj_lam_call(c, "zsc_adj_seed", [&](auto tag, auto ltag) {
        using T = ZSC_T;
        using L = typename decltype(ltag)::type;
        void *adj_scr = (char *)scratch + align256((size_t)c->geo.n * sizeof(T));
        c->adj_deferred = false;
        return Adj<T, L>::seed(c->geo, c->stream, (const T *)a_n, (const T *)phi_tgt, (L *)lam,
                               adj_scr, grads);
    });
}

extern "C" int zsc_adj_backward(zsc_ctx *c, const void *state, int pulled, const void *phi_n,
                                const void *lam_next, void *lam_out, void *scratch, double *grads)
{
    if (!state || !lam_next || !lam_out || !scratch || lam_next == lam_out || (pulled && !phi_n))
        return fail(ZSC_ERR_ARG, "zsc_adj_backward: null or aliased pointer");
    return adj_lam_call(c, "zsc_adj_backward", [&](auto tag, auto ltag) {
        using T = ZSC_T;
        using L = typename decltype(ltag)::type;
        void *phi = scratch;
        void *adj_scr = (char *)scratch + align256((size_t)c->geo.n * sizeof(T));
        if (!phi_n) {
            hipError_t e = Ops<T>::lat_phi(c->geo, c->stream, state, phi);
            if (e != hipSuccess) return e;
            phi_n = phi;
        }
        hipError_t e = Adj<T, L>::backward(c->geo, c->params, c->grav_on, c->stream,
                                           (const T *)state, pulled != 0, (const T *)phi_n,
                                           (const L *)lam_next, (L *)lam_out, adj_scr, grads,
                                           !grads && !c->adj_deferred, adj_dev(c));
        if (e == hipSuccess && !grads) c->adj_deferred = true;
        return e;
    });
}

extern "C" int zsc_adj_backward_steps(zsc_ctx *c, int k, const void *const *states,
                                      const void *const *phis, int first_pulled,
                                      const void *lam_next, void *lam_out, void *lam_tmp,
                                      void *scratch, double *grads)
{
    if (k < 1 || !states || !phis || !lam_next || !lam_out || !scratch || lam_next == lam_out ||
        (k > 1 && (!lam_tmp || lam_tmp == lam_out || lam_tmp == lam_next)))
        return fail(ZSC_ERR_ARG, "zsc_adj_backward_steps: bad step count, null or aliased pointer");
    for (int j = 0; j < k; ++j)
        if (!states[j] || ((j > 0 || first_pulled) && !phis[j]))
            return fail(ZSC_ERR_ARG, "zsc_adj_backward_steps: null state or phi of step %d", j);
    return adj_lam_call(c, "zsc_adj_backward_steps", [&](auto tag, auto ltag) {
        using T = ZSC_T;
        using L = typename decltype(ltag)::type;
        void *adj_scr = (char *)scratch + align256((size_t)c->geo.n * sizeof(T));
        L *term = Adj<T, L>::term(c->geo, adj_scr);
        const L *lin = (const L *)lam_next;
        const L *a_in = nullptr;
        for (int j = k - 1; j >= 0; --j) {
            // the step of j = 0 writes lam_out, the ones before alternate
            L *lout = (L *)((j & 1) == 0 ? lam_out : lam_tmp);
            const bool pulled = j > 0 || first_pulled;
            const void *phi_n = phis[j];
            if (!phi_n) {
                hipError_t e = Ops<T>::lat_phi(c->geo, c->stream, states[j], scratch);
                if (e != hipSuccess) return e;
                phi_n = scratch;
            }
            hipError_t e = Adj<T, L>::backward(c->geo, c->params, c->grav_on, c->stream,
                                               (const T *)states[j], pulled, (const T *)phi_n, lin,
                                               lout, adj_scr, grads, !grads && !c->adj_deferred,
                                               adj_dev(c), a_in, j > 0 ? term : nullptr);
            if (e != hipSuccess) return e;
            if (!grads) c->adj_deferred = true;
            lin = lout;
            a_in = term;
        }
        return hipSuccess;
    });
}

extern "C" int zsc_adj_flush(zsc_ctx *c, void *scratch, double *grads)
{
    if (!scratch || !grads) return fail(ZSC_ERR_ARG, "zsc_adj_flush: null pointer");
    if (c && !c->adj_deferred) return ZSC_OK;
    return adj_lam_call(c, "zsc_adj_flush", [&](auto tag, auto ltag) {
        using T = ZSC_T;
        using L = typename decltype(ltag)::type;
        hipError_t e = Adj<T, L>::flush(c->geo, c->stream,
                                        (char *)scratch + align256((size_t)c->geo.n * sizeof(T)),
                                        grads);
        if (e == hipSuccess) c->adj_deferred = false;
        return e;
    });
}

extern "C" int zsc_adj_collide(zsc_ctx *c, const void *a, void *b, void *phi)
{
    if (!a || !b || !phi || a == b) return fail(ZSC_ERR_ARG, "zsc_adj_collide: null or aliased pointer");
    return adj_call(c, "zsc_adj_collide", [&](auto tag) {
        using T = ZSC_T;
        hipError_t e = Ops<T>::lat_phi(c->geo, c->stream, a, phi);
        if (e != hipSuccess) return e;
        return Ops<T>::lat_collide(c->geo, c->params, c->grav_on, c->stream, a, phi, b);
    });
}

extern "C" int zsc_adj_fused(zsc_ctx *c, const void *b_in, void *b_out, void *phi)
{
    if (!b_in || !b_out || b_in == b_out) return fail(ZSC_ERR_ARG, "zsc_adj_fused: null or aliased pointer");
    return adj_call(c, "zsc_adj_fused", [&](auto tag) {
        using T = ZSC_T;
        return Ops<T>::lat_fused(c->geo, c->params, c->grav_on, c->stream, c->cfg, b_in, b_out, phi);
    });
}

extern "C" int zsc_adj_stream(zsc_ctx *c, const void *b, void *a)
{
    if (!a || !b || a == b) return fail(ZSC_ERR_ARG, "zsc_adj_stream: null or aliased pointer");
    return adj_call(c, "zsc_adj_stream", [&](auto tag) {
        using T = ZSC_T;
        return Ops<T>::lat_stream(c->geo, c->params, c->stream, b, a);
    });
}

extern "C" int zsc_adj_init_backward(zsc_ctx *c, const void *phi0, const void *lam0, void *scratch,
                                     double *grads)
{
    if (!phi0 || !lam0 || !scratch || !grads) return fail(ZSC_ERR_ARG, "zsc_adj_init_backward: null pointer");
    return adj_lam_call(c, "zsc_adj_init_backward", [&](auto tag, auto ltag) {
        using T = ZSC_T;
        using L = typename decltype(ltag)::type;
        void *adj_scr = (char *)scratch + align256((size_t)c->geo.n * sizeof(T));
        return Adj<T, L>::init_backward(c->geo, c->params, c->stream, (const T *)phi0,
                                        (const L *)lam0, adj_scr, grads);
    });
}
