// tt_capi.cpp -- C-ABI entry points of the MI355X trace transform.
//
// Host responsibilities: argument checks (the reference asserts or throws;
// this ABI returns TT_ERR_INVALID with a message), per-context device
// workspaces, and the small host-computed tables: rotation trig and T3-T5
// tables whose rounding must match the reference's libm calls bit for bit,
// and Hermite weights at the reference's float sample positions.  All GPU
// work is in tt_kernels.hip.
#include "../../include/tt.h"
#include "../../include/tt_testing.h"
#include "tt_internal.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

// Dynamic LDS of one workgroup: at most the 160 KiB of a gfx950 CU.
constexpr size_t kLdsLimit = 160 * 1024;

tt_status fail(tt_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

#define TT_HIP(expr)                                                           \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess)                                                  \
            return fail(e_ == hipErrorOutOfMemory ? TT_ERR_NOMEM : TT_ERR_DEVICE, \
                        "%s failed: %s", #expr, hipGetErrorString(e_));        \
    } while (0)

#define TT_CHECK(cond, ...)                                                    \
    do {                                                                       \
        if (!(cond))                                                           \
            return fail(TT_ERR_INVALID, __VA_ARGS__);                          \
    } while (0)

// Allocation-failure hook (tt_testing_fail_allocations): the number of
// workspace allocations still allowed to succeed, or -1 (no limit).
std::atomic<long long> g_alloc_budget{-1};

hipError_t workspace_malloc(void **ptr, size_t bytes)
{
    long long left = g_alloc_budget.load();
    while (left >= 0) {
        if (left == 0)
            return hipErrorOutOfMemory;
        if (g_alloc_budget.compare_exchange_weak(left, left - 1))
            break;
    }
    return hipMalloc(ptr, bytes);
}

struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes)
    {
        if (bytes <= cap)
            return hipSuccess;
        if (ptr) {
            hipError_t e = hipFree(ptr);
            if (e != hipSuccess)
                return e;
            ptr = nullptr;
            cap = 0;
        }
        hipError_t e = workspace_malloc(&ptr, bytes);
        if (e == hipSuccess)
            cap = bytes;
        return e;
    }
    void release()
    {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// ---- host tables (reference rounding) -------------------------------------

// c++/src/auxiliary.cpp:180 deg2rad, c++/src/sinogram.cpp:112 angle,
// c++/src/auxiliary.cpp:235-236 std::cos/std::sin of the negated float
// angle (glibc cosf/sinf).
void make_trig(int A, float step, std::vector<float2> &out)
{
    out.resize(A);
    for (int k = 0; k < A; k++) {
        const float deg = (float)((double)k * (double)step);
        const float rad = (float)(deg * M_PI / 180);
        out[k].x = cosf(-rad);
        out[k].y = sinf(-rad);
    }
}

// c++/src/functionals.cpp:136-198 (GCC branch): double log/cos/sin, stored
// as float; entry 0 is never read.
void make_tab345(int n, std::vector<float2> &out)
{
    out.assign((size_t)3 * n, float2{0.0f, 0.0f});
    for (int r = 1; r < n; r++) {
        const double lr = log((double)r);
        out[r] = float2{(float)(r * cos(5.0 * lr)), (float)(r * sin(5.0 * lr))};
        out[(size_t)n + r] = float2{(float)cos(3.0 * lr), (float)sin(3.0 * lr)};
        const double sq = sqrt((double)r);
        out[(size_t)2 * n + r] =
            float2{(float)(sq * cos(4.0 * lr)), (float)(sq * sin(4.0 * lr))};
    }
}

// Hermite function h_n(z) = H_n(z) / (sqrt(2^n n! sqrt(pi)) e^{z^2/2}) in
// double, with the physicists' recursion H_k = 2z H_{k-1} - 2(k-1) H_{k-2}
// (c++/src/functionals.cpp:62-84 evaluates the same recursion, rounding each
// level to float).
double herm_function(unsigned order, double z)
{
    double h0 = 1.0, h1 = 2.0 * z;
    if (order == 0)
        h1 = 1.0;
    for (unsigned k = 2; k <= order; k++) {
        const double h2 = 2.0 * z * h1 - 2.0 * (k - 1) * h0;
        h0 = h1;
        h1 = h2;
    }
    double fact = 1.0;
    for (unsigned k = 2; k <= order; k++)
        fact *= k;
    return h1 / (sqrt(pow(2.0, order) * fact * sqrt(M_PI)) * exp(z * z / 2));
}

// c++/src/functionals.cpp:369-386: the z grid over [-10, 10] split at
// center, with the reference's float step sizes and float z values (the
// sample positions); the weight of sample p is h_order(z_p), kept in double.
void make_hermite(int n, unsigned order, int center, double *w)
{
    const float lo = (float)(10.0 / center);
    const float hi = (float)(10.0 / (double)(n - 1 - center));
    for (int p = 0; p < n; p++) {
        const float z = p < center ? p * lo - 10 : (p - center) * hi;
        w[p] = herm_function(order, (double)z);
    }
}

}  // namespace

// Line data a batched forward keeps for a batched backward of the same batch
// (per image, parts 256-byte aligned): the weighted medians m and sm
// ([A][n] int32 each), the complex T3..T5 integrals ([3][A][n] float2), P3's
// half spectra ([nt][A][2h] float, h = n / 2 + 1) and P2's source rows
// ([nt][A] int32).
struct SavedLayout {
    size_t med_x, med_s, z345, spec, p2src, total;
};

SavedLayout saved_layout(int n, int A, int nt)
{
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t ln = (size_t)n * A;
    const size_t h2 = 2 * (size_t)(n / 2 + 1);
    SavedLayout L;
    L.med_x = 0;
    L.med_s = al(L.med_x + 4 * ln);
    L.z345 = al(L.med_s + 4 * ln);
    L.spec = al(L.z345 + 24 * ln);
    L.p2src = al(L.spec + 4 * (size_t)nt * A * h2);
    L.total = al(L.p2src + 4 * (size_t)nt * A);
    return L;
}

// What a saved block was written for: angles, step and functionals.
std::vector<int64_t> saved_key(int A, float step, const int32_t *tcodes, int nt,
                               const tt_pfunc *pf, int np)
{
    std::vector<int64_t> k = {A, (int64_t)(*(const uint32_t *)&step), nt, np};
    for (int t = 0; t < nt; t++)
        k.push_back(tcodes[t]);
    for (int q = 0; q < np; q++)
        k.push_back(pf[q].code);
    return k;
}

struct SavedEntry {
    int n;
    std::vector<int64_t> key;
    const float *sino;
};

// Event pairs bracketing kernel launches while profiling is on.
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<int> kind;        // kind of pair i
    double ms[TT_KERNEL_KINDS] = {};
    int64_t count[TT_KERNEL_KINDS] = {};

    hipError_t take(hipEvent_t *e)
    {
        if (used == pool.size()) {
            hipEvent_t ev;
            hipError_t err = hipEventCreate(&ev);
            if (err != hipSuccess)
                return err;
            pool.push_back(ev);
        }
        *e = pool[used++];
        return hipSuccess;
    }
    hipError_t begin(int k, hipStream_t s)
    {
        if (!on)
            return hipSuccess;
        hipEvent_t e;
        hipError_t err = take(&e);
        if (err == hipSuccess)
            err = hipEventRecord(e, s);
        kind.push_back(k);
        return err;
    }
    hipError_t end(hipStream_t s)
    {
        if (!on)
            return hipSuccess;
        hipEvent_t e;
        hipError_t err = take(&e);
        return err == hipSuccess ? hipEventRecord(e, s) : err;
    }
    hipError_t collect()
    {
        for (size_t i = 0; i < kind.size(); i++) {
            float t = 0.0f;
            hipError_t err = hipEventSynchronize(pool[2 * i + 1]);
            if (err == hipSuccess)
                err = hipEventElapsedTime(&t, pool[2 * i], pool[2 * i + 1]);
            if (err != hipSuccess)
                return err;
            ms[kind[i]] += t;
            count[kind[i]] += 1;
        }
        kind.clear();
        used = 0;
        return hipSuccess;
    }
    void release()
    {
        for (hipEvent_t e : pool)
            (void)hipEventDestroy(e);
        pool.clear();
        kind.clear();
        used = 0;
    }
};

// A batch ready to replay: descriptors, workspaces and the captured graph.
// It owns copies of the tables it uses so later single-image calls (which
// rewrite the context's tables) cannot change what the graph reads.
struct BatchPlan {
    std::vector<int64_t> key;
    DevBuf desc, imgT, trig, tab, counter, bbox, ckpt, spec, jobs;
    ttk::DftParams dft = {};      // P3 spectra of every image (jobs in `jobs`)
    // launch parameters of the captured stages (reused for timed eager runs)
    ttk::BatchSinoParams sp = {};
    ttk::BatchCircusParams cp = {};
    int nimg = 0, nmax = 0;
    bool need_x = false, need_s = false, do345 = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int waves_per_block = 4, blocks = 0;
    int tile_a = 1, tile_c = 64;  // line tile (angles x columns)
    bool valid = false;

    void drop_graph()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        valid = false;
    }
    void release()
    {
        drop_graph();
        desc.release();
        imgT.release();
        trig.release();
        tab.release();
        counter.release();
        bbox.release();
        ckpt.release();
        spec.release();
        jobs.release();
        dft = {};
        key.clear();
    }
};

// Workspaces of a batched backward (per image: sinograms, gradients,
// medians, T3..T5 integrals, line records), laid out back to back.
struct BackwardPlan {
    BatchPlan g;                  // descriptors, tables, graph
    DevBuf sino, dsino, z, medx, meds, rec, spec, y, p2src, fjobs, sjobs, bjobs;
    int gather_tile = 32, gather_waves = 4;
    ttk::BatchAdjointParams ap = {};
    ttk::DftParams fdp = {}, bdp = {};
    ttk::P3ScaleParams p3s = {};
    bool saved = false, p2src_on = false;
    int cols = 0;
};

struct tt_context {
    int device = 0;
    KernelTimer timer;
    BatchPlan batch;
    BackwardPlan bwd;
    bool autotune = false;
    tt_tiles force = {};  // tiles forced by tt_context_set_tiles (0: not forced)
    std::map<const void *, SavedEntry> saved;  // d_saved blocks a batched forward writes
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    DevBuf imgT, trig, tab345, io_a, io_b, med_x, med_s, sino_ck;
    DevBuf p3_spec;
    DevBuf bw_sino, bw_dsino, bw_z, bw_rec, bw_rec67, bw_part, bw_t6src, bw_t7src, bw_p2src;
    DevBuf nos_med, nos_a, nos_u, nos_v, nos_s, nos_e, nos_c, nos_info, nos_sweeps;
    DevBuf nos_w, nos_m, nos_q, nos_r, nos_dev, nos_sp, nos_ptrs, nos_h;
    double *h_dev = nullptr;  // pinned host copy of the NOS convergence partials
    size_t h_dev_cap = 0;
    DevBuf sort_qa, sort_qb, sort_count, bbox;
    rocblas_handle blas = nullptr;
    int trig_A = -1;
    float trig_step = 0.0f;
    int tab_n = -1;
    std::vector<float2> h_trig, h_tab;
    std::vector<double> h_herm;

    tt_status set_device()
    {
        TT_HIP(hipSetDevice(device));
        return TT_OK;
    }

    tt_status ensure_trig(int A, float step)
    {
        if (trig_A == A && trig_step == step)
            return TT_OK;
        make_trig(A, step, h_trig);
        TT_HIP(trig.ensure(sizeof(float2) * A));
        TT_HIP(hipMemcpyAsync(trig.ptr, h_trig.data(), sizeof(float2) * A,
                              hipMemcpyHostToDevice, stream));
        TT_HIP(hipStreamSynchronize(stream));
        trig_A = A;
        trig_step = step;
        return TT_OK;
    }

    tt_status ensure_tab345(int n)
    {
        if (tab_n == n)
            return TT_OK;
        make_tab345(n, h_tab);
        TT_HIP(tab345.ensure(sizeof(float2) * 3 * (size_t)n));
        TT_HIP(hipMemcpyAsync(tab345.ptr, h_tab.data(), sizeof(float2) * 3 * (size_t)n,
                              hipMemcpyHostToDevice, stream));
        TT_HIP(hipStreamSynchronize(stream));
        tab_n = n;
        return TT_OK;
    }

    // Hermite weight tables of (n, (order, center) list) keys, the most
    // recently used kHermCache kept on the device: a transform's orthonormal
    // sinograms have their own heights and centres, so one table per call
    // would be rebuilt (and the stream synchronised) on every call.
    static constexpr size_t kHermCache = 16;
    struct HermEntry {
        std::vector<int32_t> key;
        DevBuf buf;
        uint64_t used = 0;
    };
    std::vector<HermEntry> herm_cache;
    uint64_t herm_clock = 0;

    tt_status ensure_hermite(int n, const std::vector<int32_t> &pairs, const double **out)
    {
        std::vector<int32_t> key;
        key.reserve(pairs.size() + 1);
        key.push_back(n);
        key.insert(key.end(), pairs.begin(), pairs.end());
        for (HermEntry &e : herm_cache)
            if (e.key == key) {
                e.used = ++herm_clock;
                *out = e.buf.as<double>();
                return TT_OK;
            }
        HermEntry *slot = nullptr;
        if (herm_cache.size() < kHermCache) {
            herm_cache.emplace_back();
            slot = &herm_cache.back();
        } else {
            slot = &herm_cache[0];
            for (HermEntry &e : herm_cache)
                slot = e.used < slot->used ? &e : slot;
            slot->buf.release();  // hipFree waits for the kernels that read it
        }
        const size_t nh = pairs.size() / 2;
        h_herm.assign(nh * (size_t)n, 0.0f);
        for (size_t h = 0; h < nh; h++)
            make_hermite(n, (unsigned)pairs[2 * h], pairs[2 * h + 1],
                         h_herm.data() + h * (size_t)n);
        TT_HIP(slot->buf.ensure(sizeof(double) * (nh ? nh : 1) * (size_t)n));
        if (nh)
            TT_HIP(hipMemcpyAsync(slot->buf.ptr, h_herm.data(),
                                  sizeof(double) * nh * (size_t)n,
                                  hipMemcpyHostToDevice, stream));
        TT_HIP(hipStreamSynchronize(stream));  // h_herm is reused
        slot->key = std::move(key);
        slot->used = ++herm_clock;
        *out = slot->buf.as<double>();
        return TT_OK;
    }
};

namespace {

tt_status check_tcodes(const int32_t *tcodes, int nt)
{
    TT_CHECK(nt >= 0, "nt must be >= 0 (got %d)", nt);
    TT_CHECK(nt == 0 || tcodes, "tcodes is NULL");
    for (int t = 0; t < nt; t++)
        TT_CHECK(tcodes[t] >= 0 && tcodes[t] <= 7,
                 "invalid T-functional code %d at position %d (valid: 0..7)",
                 tcodes[t], t);
    return TT_OK;
}

tt_status check_pfuncs(const tt_pfunc *pf, int np, int n)
{
    TT_CHECK(np >= 0 && np <= ttk::kMaxP, "np must be in 0..%d (got %d)",
             ttk::kMaxP, np);
    TT_CHECK(np == 0 || pf, "pf is NULL");
    for (int q = 0; q < np; q++) {
        TT_CHECK(pf[q].code >= 0 && pf[q].code <= 3,
                 "invalid P-functional code %d at position %d (valid: 0..3)",
                 pf[q].code, q);
        if (pf[q].code == TT_P_HERMITE) {
            TT_CHECK(pf[q].order >= 0, "Hermite order must be >= 0 (got %d)",
                     pf[q].order);
            TT_CHECK(pf[q].center >= 0 && pf[q].center < n,
                     "Hermite center %d unset or outside [0, %d)", pf[q].center,
                     n);
        }
    }
    return TT_OK;
}

}  // namespace

#define TT_TRY(expr)                                                           \
    do {                                                                       \
        tt_status s_ = (expr);                                                 \
        if (s_ != TT_OK)                                                       \
            return s_;                                                         \
    } while (0)

#define TT_BLAS(expr)                                                          \
    do {                                                                       \
        rocblas_status b_ = (expr);                                            \
        if (b_ != rocblas_status_success)                                      \
            return fail(TT_ERR_DEVICE, "%s failed: rocblas status %d", #expr,  \
                        (int)b_);                                              \
    } while (0)

extern "C" {

// ---- scratch of the allocation-free entry point -----------------------------
// tt_circus_noalloc keeps P3's half spectra in this file-scope device array
// (each GPU the library runs on holds its own copy until the process ends).
// Host code never uses the array's name as an address: each device's copy is
// looked up once, with that device current, when the first context is
// created on it, and kernels receive that pointer.
constexpr size_t kNoAllocSpecFloats = (size_t)8 << 20;  // 32 MiB per GPU
static __device__ __align__(256) float scratch[kNoAllocSpecFloats];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
bool g_scratch_done[kMaxDevices];

// Call with `device` current.
void resolve_noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device])
        return;
    void *ptr = nullptr;
    g_scratch_err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? (float *)ptr : nullptr;
    g_scratch_done[device] = true;
}

tt_status noalloc_scratch(int device, float **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || !g_scratch_done[device])
        return fail(TT_ERR_DEVICE, "no scratch address for device %d", device);
    if (!g_scratch_ptr[device])
        return fail(TT_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s",
                    device, hipGetErrorString(g_scratch_err[device]));
    *out = g_scratch_ptr[device];
    return TT_OK;
}

}  // namespace

TT_API const char *tt_last_error(void) { return g_err.c_str(); }

TT_API tt_status tt_context_create(int device, tt_context **out)
{
    TT_CHECK(out, "out is NULL");
    *out = nullptr;
    int count = 0;
    TT_HIP(hipGetDeviceCount(&count));
    TT_CHECK(device >= 0 && device < count, "device %d not in [0, %d)", device,
             count);
    tt_context *ctx = new (std::nothrow) tt_context();
    if (!ctx)
        return fail(TT_ERR_NOMEM, "context allocation failed");
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking);
    if (e == hipSuccess)
        resolve_noalloc_scratch(device);  // tt_circus_noalloc reports a failure
    if (e != hipSuccess) {
        delete ctx;
        return fail(TT_ERR_DEVICE, "context setup failed: %s", hipGetErrorString(e));
    }
    ctx->stream = ctx->own;
    *out = ctx;
    return TT_OK;
}

TT_API tt_status tt_context_destroy(tt_context *ctx)
{
    if (!ctx)
        return TT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    DevBuf *bufs[] = {&ctx->p3_spec,
                      &ctx->imgT, &ctx->trig, &ctx->tab345,
                      &ctx->io_a, &ctx->io_b, &ctx->med_x,
                      &ctx->med_s, &ctx->sino_ck, &ctx->bw_sino, &ctx->bw_dsino, &ctx->bw_z,
                      &ctx->bw_rec, &ctx->bw_rec67, &ctx->bw_part, &ctx->bw_p2src, &ctx->bw_t6src,
                      &ctx->bw_t7src, &ctx->nos_med, &ctx->nos_a, &ctx->nos_u,
                      &ctx->nos_v, &ctx->nos_s, &ctx->nos_e, &ctx->nos_c,
                      &ctx->nos_info, &ctx->nos_sweeps, &ctx->nos_w, &ctx->nos_m,
                      &ctx->nos_q, &ctx->nos_r, &ctx->nos_dev, &ctx->nos_sp,
                      &ctx->nos_ptrs, &ctx->nos_h, &ctx->sort_qa,
                      &ctx->sort_qb, &ctx->sort_count, &ctx->bbox};
    for (DevBuf *b : bufs)
        b->release();
    for (auto &e : ctx->herm_cache)
        e.buf.release();
    ctx->herm_cache.clear();
    if (ctx->h_dev)
        (void)hipHostFree(ctx->h_dev);
    ctx->h_dev = nullptr;
    ctx->h_dev_cap = 0;
    ctx->timer.release();
    ctx->batch.release();
    ctx->bwd.g.release();
    DevBuf *bb[] = {&ctx->bwd.sino, &ctx->bwd.dsino, &ctx->bwd.z, &ctx->bwd.medx,
                    &ctx->bwd.meds, &ctx->bwd.rec, &ctx->bwd.spec, &ctx->bwd.y,
                    &ctx->bwd.p2src, &ctx->bwd.fjobs, &ctx->bwd.sjobs, &ctx->bwd.bjobs};
    for (DevBuf *b : bb)
        b->release();
    if (ctx->blas)
        (void)rocblas_destroy_handle(ctx->blas);
    if (ctx->ev_in)
        (void)hipEventDestroy(ctx->ev_in);
    if (ctx->ev_out)
        (void)hipEventDestroy(ctx->ev_out);
    if (ctx->own)
        (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return TT_OK;
}

TT_API tt_status tt_context_set_stream(tt_context *ctx, void *stream)
{
    TT_CHECK(ctx, "ctx is NULL");
    hipStream_t next = (hipStream_t)stream;  // NULL: HIP's default stream
    if (next != ctx->stream) {
        TT_TRY(ctx->set_device());
        // Tables and workspaces are shared by all work of the context: drain
        // the old stream before work is queued on the new one.
        TT_HIP(hipStreamSynchronize(ctx->stream));
        ctx->stream = next;
    }
    return TT_OK;
}

TT_API tt_status tt_context_synchronize(tt_context *ctx)
{
    TT_CHECK(ctx, "ctx is NULL");
    TT_TRY(ctx->set_device());
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TT_OK;
}

TT_API tt_status tt_pad_geometry(int rows, int cols, int *n, int *dfx, int *dfy)
{
    TT_CHECK(rows > 0 && cols > 0, "image must be non-empty (got %d x %d)", rows,
             cols);
    TT_CHECK(n && dfx && dfy, "output pointer is NULL");
    // c++/src/auxiliary.cpp:259-270
    const float ox = (float)(floor((cols + 1) / 2.0) - 1);
    const float oy = (float)(floor((rows + 1) / 2.0) - 1);
    const float hx = (float)cols - 1 - ox - 1;
    const float hy = (float)rows - 1 - oy - 1;
    const int rlast = (int)ceilf(hypotf(hx, hy)) + 1;
    const int nb = 2 * rlast + 1;
    const float op = (float)(floor((nb + 1) / 2.0) - 1);
    *n = nb;
    *dfx = (int)(op - ox);
    *dfy = (int)(op - oy);
    return TT_OK;
}

TT_API tt_status tt_pad(tt_context *ctx, const float *d_image, int rows,
                        int cols, float *d_padded)
{
    TT_CHECK(ctx && d_image && d_padded, "NULL argument");
    int n, dfx, dfy;
    TT_TRY(tt_pad_geometry(rows, cols, &n, &dfx, &dfy));
    TT_TRY(ctx->set_device());
    TT_HIP(ctx->timer.begin(TT_KERNEL_AUX, ctx->stream));
    TT_HIP(ttk::launch_pad(d_image, rows, cols, d_padded, n, dfx, dfy, ctx->stream));
    TT_HIP(ctx->timer.end(ctx->stream));
    return TT_OK;
}

TT_API size_t tt_saved_bytes(int n, int n_angles, int nt)
{
    if (n < 3 || n_angles < 1 || nt < 1)
        return 0;
    return saved_layout(n, n_angles, nt).total;
}

TT_API tt_status tt_saved_offsets(int n, int n_angles, int nt, size_t *med_x,
                                  size_t *med_s, size_t *z345, size_t *spec,
                                  size_t *p2src)
{
    TT_CHECK(n >= 3 && n_angles >= 1 && nt >= 1, "invalid sizes");
    TT_CHECK(med_x && med_s && z345 && spec && p2src, "NULL argument");
    const SavedLayout L = saved_layout(n, n_angles, nt);
    *med_x = L.med_x;
    *med_s = L.med_s;
    *z345 = L.z345;
    *spec = L.spec;
    *p2src = L.p2src;
    return TT_OK;
}

TT_API int tt_angle_count(float angle_step_deg)
{
    if (!(angle_step_deg > 0.0f))
        return 0;
    return (int)floor(360.0 / (double)angle_step_deg);
}

TT_API tt_status tt_sinograms(tt_context *ctx, const float *d_image, int n,
                              int n_angles, float angle_step_deg,
                              const int32_t *tcodes, int nt, float *d_sino,
                              const tt_sino_aux *aux)
{
    TT_CHECK(ctx, "ctx is NULL");
    TT_CHECK(d_image && (d_sino || nt == 0), "NULL image or sinogram pointer");
    TT_CHECK(n >= 3, "padded side n must be >= 3 (got %d)", n);
    TT_CHECK(n_angles >= 1, "n_angles must be >= 1 (got %d)", n_angles);
    TT_CHECK(angle_step_deg > 0.0f && isfinite(angle_step_deg),
             "angle step must be positive and finite");
    // the T3-T5 tables are staged in LDS (24 bytes per row, 64 KiB maximum)
    TT_CHECK(n <= 2730, "padded side %d above the supported 2730", n);
    TT_TRY(check_tcodes(tcodes, nt));
    tt_sino_aux none = {};
    const tt_sino_aux &ax = aux ? *aux : none;

    int slot[8];
    for (int k = 0; k < 8; k++)
        slot[k] = -1;
    std::vector<std::pair<int, int>> dups;  // (from slot, to slot)
    for (int t = 0; t < nt; t++) {
        if (slot[tcodes[t]] < 0)
            slot[tcodes[t]] = t;
        else
            dups.push_back({slot[tcodes[t]], t});
    }
    TT_TRY(ctx->set_device());
    hipStream_t st = ctx->stream;
    const bool do6 = slot[6] >= 0, do7 = slot[7] >= 0;
    // T6/T7 sort up to 2048 keys per line in a wave's registers
    TT_CHECK(!(do6 || do7) || n <= 2048,
             "T6/T7 support padded sides up to 2048 (got %d)", n);
    const bool need_x = slot[1] >= 0 || slot[2] >= 0 || ax.median_x || do7;
    const bool do345 = slot[3] >= 0 || slot[4] >= 0 || slot[5] >= 0;
    const bool need_s = do345 || ax.median_sqrt || do6;
    const bool any = need_x || need_s || slot[0] >= 0;
    if (!any)
        return TT_OK;
    // T6/T7 take their medians from the T0..T5 line kernel; route them
    // through the workspace when the caller does not ask for them.
    int32_t *med_x = ax.median_x, *med_s = ax.median_sqrt;
    const size_t med_b = sizeof(int32_t) * (size_t)n * n_angles;
    if (do7 && !med_x) {
        TT_HIP(ctx->med_x.ensure(med_b));
        med_x = ctx->med_x.as<int32_t>();
    }
    if (do6 && !med_s) {
        TT_HIP(ctx->med_s.ensure(med_b));
        med_s = ctx->med_s.as<int32_t>();
    }

    TT_TRY(ctx->ensure_trig(n_angles, angle_step_deg));
    if (do345)
        TT_TRY(ctx->ensure_tab345(n));
    TT_HIP(ctx->imgT.ensure(sizeof(float) * (size_t)n * n));
    TT_HIP(ctx->bbox.ensure(ttk::kBboxInts * sizeof(int)));
    TT_HIP(ctx->timer.begin(TT_KERNEL_AUX, st));
    TT_HIP(ttk::launch_transpose(d_image, ctx->imgT.as<float>(), n, st));
    TT_HIP(ttk::launch_image_bbox(d_image, n, ctx->bbox.as<int>(), st));
    TT_HIP(ctx->timer.end(st));

    ttk::SinoParams p = {};
    p.img = d_image;
    p.imgT = ctx->imgT.as<float>();
    p.trig = ctx->trig.as<float2>();
    p.tab345 = do345 ? ctx->tab345.as<float2>() : nullptr;
    p.sino = d_sino;
    p.med_x = med_x;
    p.med_s = med_s;
    p.bbox = ctx->bbox.as<int>();
    TT_HIP(ctx->sino_ck.ensure(sizeof(float2) * ttk::sino_ckpt_blocks(n) * (size_t)n *
                               n_angles));
    p.ckpt = ctx->sino_ck.as<float2>();
    p.n = n;
    p.A = n_angles;
    p.nb = (n + ttk::kWave - 1) / ttk::kWave;
    p.items = p.nb * n_angles;
    for (int k = 0; k < 8; k++)
        p.slot.s[k] = slot[k];
    TT_HIP(ctx->timer.begin(TT_KERNEL_SINO_T05, st));
    TT_HIP(ttk::launch_sino_t05(p, need_x, need_s, do345, st));
    TT_HIP(ctx->timer.end(st));
    if (do6 || do7) {
        ttk::SortTParams q = {};
        q.img = d_image;
        q.imgT = ctx->imgT.as<float>();
        q.trig = ctx->trig.as<float2>();
        q.sino = d_sino;
        q.med_x = med_x;
        q.med_s = med_s;
        q.t6_rank = ax.t6_rank;
        q.t6_src = ax.t6_src;
        q.t7_rank = ax.t7_rank;
        q.t7_src = ax.t7_src;
        q.bbox = ctx->bbox.as<int>();
        q.n = n;
        q.A = n_angles;
        q.slot6 = slot[6];
        q.slot7 = slot[7];
        int32_t *qa = nullptr, *qb = nullptr;
        TT_HIP(ctx->sort_count.ensure(ttk::kSortCounters * sizeof(unsigned int)));
        unsigned int *counts = ctx->sort_count.as<unsigned int>();
        if (n > 512) {
            const size_t qbytes = sizeof(int32_t) * (size_t)n * n_angles;  // a line item per line
            TT_HIP(ctx->sort_qa.ensure(qbytes));
            TT_HIP(ctx->sort_qb.ensure(qbytes));
            qa = ctx->sort_qa.as<int32_t>();
            qb = ctx->sort_qb.as<int32_t>();
        }
        TT_HIP(ctx->timer.begin(TT_KERNEL_SINO_SORT, st));
        TT_HIP(ttk::launch_sino_sort(q, qa, qb, counts, st));
        TT_HIP(ctx->timer.end(st));
    }
    const size_t plane = sizeof(float) * (size_t)n * n_angles;
    for (auto &d : dups)
        TT_HIP(hipMemcpyAsync((char *)d_sino + plane * d.second,
                              (char *)d_sino + plane * d.first, plane,
                              hipMemcpyDeviceToDevice, st));
    return TT_OK;
}

TT_API tt_status tt_sinograms_host(tt_context *ctx, const float *image, int n,
                                   int n_angles, float angle_step_deg,
                                   const int32_t *tcodes, int nt, float *sino)
{
    TT_CHECK(ctx && image && (sino || nt == 0), "NULL argument");
    TT_CHECK(n >= 3 && n_angles >= 1 && nt >= 0, "invalid sizes");
    TT_TRY(ctx->set_device());
    const size_t img_b = sizeof(float) * (size_t)n * n;
    const size_t out_b = sizeof(float) * (size_t)nt * n * n_angles;
    TT_HIP(ctx->io_a.ensure(img_b));
    TT_HIP(ctx->io_b.ensure(out_b ? out_b : 4));
    TT_HIP(hipMemcpyAsync(ctx->io_a.ptr, image, img_b, hipMemcpyHostToDevice,
                          ctx->stream));
    TT_TRY(tt_sinograms(ctx, ctx->io_a.as<float>(), n, n_angles, angle_step_deg,
                        tcodes, nt, ctx->io_b.as<float>(), nullptr));
    if (out_b)
        TT_HIP(hipMemcpyAsync(sino, ctx->io_b.ptr, out_b, hipMemcpyDeviceToHost,
                              ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TT_OK;
}

TT_API tt_status tt_circus(tt_context *ctx, const float *d_sino, int n,
                           int n_cols, int n_sino, const tt_pfunc *pf, int np,
                           float *d_circus, int32_t *d_p2_rank)
{
    TT_CHECK(ctx, "ctx is NULL");
    TT_CHECK(n >= 2, "sinogram height must be >= 2 (got %d)", n);
    TT_CHECK(n_cols >= 1 && n_sino >= 1, "n_cols and n_sino must be >= 1");
    // column, sort buffer, spectrum and twiddles staged in LDS (64 KiB max)
    TT_CHECK(n <= 2730, "sinogram height %d above the supported 2730", n);
    TT_TRY(check_pfuncs(pf, np, n));
    if (np == 0)
        return TT_OK;
    TT_CHECK(d_sino && d_circus, "NULL sinogram or circus pointer");
    TT_TRY(ctx->set_device());

    ttk::CircusParams p = {};
    p.sino = d_sino;
    p.out = d_circus;
    p.p2rank = d_p2_rank;
    p.n = n;
    p.ncols = n_cols;
    p.nsino = n_sino;
    p.np = np;
    std::vector<int32_t> hkey;
    bool p3 = false;
    for (int q = 0; q < np; q++) {
        p.code[q] = pf[q].code;
        p.herm[q] = -1;
        if (pf[q].code == TT_P_HERMITE) {
            p.herm[q] = (int)(hkey.size() / 2);
            hkey.push_back(pf[q].order);
            hkey.push_back(pf[q].center);
        }
        p3 = p3 || pf[q].code == TT_P3;
    }
    if (!hkey.empty())
        TT_TRY(ctx->ensure_hermite(n, hkey, &p.herm_w));
    hipStream_t st = ctx->stream;
    const int h = n / 2 + 1;
    const size_t cols = (size_t)n_sino * n_cols;
    ttk::DftParams dp = {};
    if (p3) {
        // P3's half spectra of every column at once: the DFT as a GEMM on the
        // matrix cores, spec (2h x cols) = W^T x (tt_dft.hip)
        TT_HIP(ctx->p3_spec.ensure(sizeof(float) * cols * 2 * h));
        p.p3_spec = ctx->p3_spec.as<float>();
        p.p3_h = h;
        dp.njobs = 1;
        dp.job.in = d_sino;
        dp.job.out = ctx->p3_spec.as<float>();
        dp.job.n = n;
        dp.job.cols = (int)cols;
        dp.job.ld_in = n;
        dp.job.ld_out = 2 * h;
        ttk::dft_tile_shape(ttk::dft_job_tiles(n, (int)cols, true, 128, 128),
                            ttk::dft_job_tiles(n, (int)cols, true, 64, 128), &dp.bm, &dp.bn);
        dp.tiles = ttk::dft_job_tiles(n, (int)cols, true, dp.bm, dp.bn);
    }
    if (p3) {
        TT_HIP(ctx->timer.begin(TT_KERNEL_P3_DFT, st));
        TT_HIP(ttk::launch_dft_gemm(dp, true, n, st));
        TT_HIP(ctx->timer.end(st));
    }
    TT_HIP(ctx->timer.begin(TT_KERNEL_CIRCUS, st));
    TT_HIP(ttk::launch_circus(p, st));
    TT_HIP(ctx->timer.end(st));
    return TT_OK;
}

TT_API tt_status tt_circus_host(tt_context *ctx, const float *sino, int n,
                                int n_cols, const tt_pfunc *pf, int np,
                                float *circus)
{
    TT_CHECK(ctx && sino && (circus || np == 0), "NULL argument");
    TT_CHECK(n >= 2 && n_cols >= 1, "invalid sizes");
    TT_TRY(ctx->set_device());
    const size_t in_b = sizeof(float) * (size_t)n * n_cols;
    const size_t out_b = sizeof(float) * (size_t)np * n_cols;
    TT_HIP(ctx->io_a.ensure(in_b));
    TT_HIP(ctx->io_b.ensure(out_b ? out_b : 4));
    TT_HIP(hipMemcpyAsync(ctx->io_a.ptr, sino, in_b, hipMemcpyHostToDevice,
                          ctx->stream));
    TT_TRY(tt_circus(ctx, ctx->io_a.as<float>(), n, n_cols, 1, pf, np,
                     ctx->io_b.as<float>(), nullptr));
    if (out_b)
        TT_HIP(hipMemcpyAsync(circus, ctx->io_b.ptr, out_b, hipMemcpyDeviceToHost,
                              ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    return TT_OK;
}

namespace {

int device_cus(int device)
{
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) !=
        hipSuccess)
        cus = 256;
    return cus;
}

// Line tiles (angles x columns) the batched forward's tuner tries.
struct LineTile {
    int a, c;
};
constexpr LineTile kLineTiles[] = {{1, 64}, {2, 64}, {4, 64}, {2, 32},
                                   {4, 32}, {4, 16}, {8, 16}};

LineTile line_tile_of(const tt_context *ctx)
{
    if (ctx->force.line_angles > 0)
        return {ctx->force.line_angles, ctx->force.line_columns};
    return kLineTiles[0];
}

// Wave items of the batched line kernel with tile t: each image's column
// tiles and first item (descriptors in launch order); returns the total.
int layout_line_items(std::vector<ttk::BatchImage> &d, int A, LineTile t)
{
    int items = 0;
    for (ttk::BatchImage &b : d) {
        b.item0 = items;
        items += ttk::line_tile_items(b.n, A, t.a, t.c, &b.nb);
    }
    return items;
}

// First gather tile of each image for tiles of side T; returns the total.
int layout_gather_tiles(std::vector<ttk::BatchImage> &d, int T)
{
    int tiles = 0;
    for (ttk::BatchImage &b : d) {
        b.tile0 = tiles;
        const int t = (b.n + T - 1) / T;
        tiles += t * t;
    }
    return tiles;
}

// The most waves per workgroup (up to 16) whose gather LDS fits, or 0.
int gather_waves_max(int T, int tabn, bool do345, bool do12)
{
    for (int waves : {16, 8, 4, 2, 1})
        if (ttk::gather_lds_bytes(T, waves, tabn, do345, do12, false) <= kLdsLimit)
            return waves;
    return 0;
}

// Workgroup offsets of a DFT job list for its launch's tile shape (larger
// tiles when the list is large enough); returns the workgroups.
int layout_dft_jobs(std::vector<ttk::DftJob> &jobs, bool forward, int *bm, int *bn)
{
    long long t128 = 0, t64 = 0;
    for (const ttk::DftJob &j : jobs) {
        t128 += ttk::dft_job_tiles(j.n, j.cols, forward, 128, 128);
        t64 += ttk::dft_job_tiles(j.n, j.cols, forward, 64, 128);
    }
    ttk::dft_tile_shape(t128, t64, bm, bn);
    int tiles = 0;
    for (ttk::DftJob &j : jobs) {
        j.tile0 = tiles;
        tiles += ttk::dft_job_tiles(j.n, j.cols, forward, *bm, *bn);
    }
    return tiles;
}

// Enqueue the batched pass (used for capture and for autotuning).
// The batched forward's stages on st.  timed: each stage bracketed by the
// context's kernel timer (never while a graph is being captured).
hipError_t enqueue_batch(tt_context *ctx, const ttk::BatchSinoParams &sp,
                         const ttk::BatchCircusParams &cp, int nimg, int nmax,
                         bool need_x, bool need_s, bool do345, int wpb, int blocks,
                         bool with_transpose, bool with_circus, hipStream_t st,
                         bool timed = false)
{
    KernelTimer *T = timed ? &ctx->timer : nullptr;
    hipError_t e = hipSuccess;
    auto stage = [&](int kind, auto launch) {
        if (e != hipSuccess)
            return;
        if (T)
            e = T->begin(kind, st);
        if (e == hipSuccess)
            e = launch();
        if (T && e == hipSuccess)
            e = T->end(st);
    };
    if (with_transpose)
        stage(TT_KERNEL_AUX, [&] {
            hipError_t r = ttk::launch_transpose_batched(sp.images, nimg, nmax, st);
            return r == hipSuccess ? ttk::launch_image_bbox_batched(sp.images, nimg, nmax, st)
                                   : r;
        });
    stage(TT_KERNEL_SINO_T05, [&] {
        hipError_t r = ttk::launch_reset_counter(sp.counter, st);
        return r == hipSuccess
                   ? ttk::launch_sino_t05_batched(sp, need_x, need_s, do345, wpb, blocks, st)
                   : r;
    });
    if (with_circus && ctx->batch.dft.tiles > 0)
        stage(TT_KERNEL_P3_DFT, [&] { return ttk::launch_dft_gemm(ctx->batch.dft, true, nmax, st); });
    if (with_circus && cp.np > 0)
        stage(TT_KERNEL_CIRCUS, [&] { return ttk::launch_circus_batched(cp, st); });
    return e;
}

// The batched backward's stages on st (timed as enqueue_batch).
hipError_t enqueue_backward(tt_context *ctx, const BackwardPlan &P, hipStream_t st,
                            bool timed)
{
    const BatchPlan &B = P.g;
    KernelTimer *T = timed ? &ctx->timer : nullptr;
    hipError_t e = hipSuccess;
    auto stage = [&](int kind, auto launch) {
        if (e != hipSuccess)
            return;
        if (T)
            e = T->begin(kind, st);
        if (e == hipSuccess)
            e = launch();
        if (T && e == hipSuccess)
            e = T->end(st);
    };
    const bool p3 = P.p3s.np3 > 0;
    if (!P.saved) {
        stage(TT_KERNEL_AUX, [&] {
            hipError_t r = ttk::launch_transpose_batched(B.sp.images, B.nimg, B.nmax, st);
            return r == hipSuccess
                       ? ttk::launch_image_bbox_batched(B.sp.images, B.nimg, B.nmax, st)
                       : r;
        });
        stage(TT_KERNEL_SINO_T05, [&] {
            hipError_t r = ttk::launch_reset_counter(B.sp.counter, st);
            return r == hipSuccess ? ttk::launch_sino_t05_batched(B.sp, true, true, B.do345,
                                                                  B.waves_per_block, B.blocks,
                                                                  st)
                                   : r;
        });
        if (p3)
            stage(TT_KERNEL_P3_DFT, [&] { return ttk::launch_dft_gemm(P.fdp, true, B.nmax, st); });
        if (P.p2src_on)
            stage(TT_KERNEL_CIRCUS, [&] { return ttk::launch_p2_source_batched(B.cp, st); });
    }
    if (p3)
        stage(TT_KERNEL_P3_DFT, [&] {
            hipError_t r = ttk::launch_p3_vjp_scale(P.p3s, P.cols, st);
            return r == hipSuccess ? ttk::launch_dft_gemm(P.bdp, false, B.nmax, st) : r;
        });
    stage(TT_KERNEL_CIRCUS_VJP, [&] { return ttk::launch_circus_vjp_batched(P.ap, st); });
    stage(TT_KERNEL_RECORDS, [&] { return ttk::launch_records_batched(P.ap, st); });
    stage(TT_KERNEL_GATHER, [&] {
        return ttk::launch_gather_batched(P.ap, P.gather_tile, P.gather_waves, st);
    });
    return e;
}

}  // namespace

TT_API tt_status tt_context_set_autotune(tt_context *ctx, int enable)
{
    TT_CHECK(ctx, "ctx is NULL");
    if ((enable != 0) != ctx->autotune) {
        ctx->batch.drop_graph();
        ctx->bwd.g.drop_graph();
    }
    ctx->autotune = enable != 0;
    return TT_OK;
}

TT_API tt_status tt_context_set_tiles(tt_context *ctx, const tt_tiles *tiles)
{
    TT_CHECK(ctx, "ctx is NULL");
    tt_tiles t = tiles ? *tiles : tt_tiles{};
    TT_CHECK((t.line_angles == 0 && t.line_columns == 0) ||
                 ttk::line_tile_ok(t.line_angles, t.line_columns),
             "line tile %d x %d: columns must be 64, 32 or 16 and angles a multiple of "
             "64 / columns up to 16",
             t.line_angles, t.line_columns);
    TT_CHECK(t.gather_tile == 0 || t.gather_tile == 16 || t.gather_tile == 32,
             "gather tile %d: must be 16 or 32", t.gather_tile);
    TT_CHECK(t.gather_waves >= 0 && t.gather_waves <= 16,
             "gather waves %d: must be 1..16", t.gather_waves);
    ctx->force = t;
    ctx->batch.drop_graph();
    ctx->bwd.g.drop_graph();
    return TT_OK;
}

TT_API tt_status tt_batched_tiles(tt_context *ctx, tt_tiles *tiles)
{
    TT_CHECK(ctx && tiles, "NULL argument");
    TT_CHECK(ctx->batch.valid || ctx->bwd.g.valid, "no batch has been planned on this context");
    *tiles = tt_tiles{};
    if (ctx->batch.valid) {
        tiles->line_angles = ctx->batch.tile_a;
        tiles->line_columns = ctx->batch.tile_c;
    }
    if (ctx->bwd.g.valid) {
        tiles->gather_tile = ctx->bwd.gather_tile;
        tiles->gather_waves = ctx->bwd.gather_waves;
    }
    return TT_OK;
}

TT_API tt_status tt_batched_launch_shape(tt_context *ctx, int *waves_per_block,
                                         int *blocks)
{
    TT_CHECK(ctx && waves_per_block && blocks, "NULL argument");
    TT_CHECK(ctx->batch.valid, "no batch has been planned on this context");
    *waves_per_block = ctx->batch.waves_per_block;
    *blocks = ctx->batch.blocks;
    return TT_OK;
}

TT_API tt_status tt_forward_batched(tt_context *ctx, const tt_image *images,
                                    int nimg, int n_angles, float angle_step_deg,
                                    const int32_t *tcodes, int nt,
                                    const tt_pfunc *pf, int np)
{
    TT_CHECK(ctx, "ctx is NULL");
    TT_CHECK(images && nimg >= 1, "images is NULL or nimg < 1");
    TT_CHECK(n_angles >= 1, "n_angles must be >= 1 (got %d)", n_angles);
    TT_CHECK(angle_step_deg > 0.0f && isfinite(angle_step_deg),
             "angle step must be positive and finite");
    TT_TRY(check_tcodes(tcodes, nt));
    TT_CHECK(nt >= 1, "at least one T-functional is required");
    int slot[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    for (int t = 0; t < nt; t++) {
        if (tcodes[t] > 5)
            return fail(TT_ERR_UNSUPPORTED,
                        "the batched path computes T0..T5 (got T%d)", tcodes[t]);
        TT_CHECK(slot[tcodes[t]] < 0, "duplicate T-functional %d in a batch",
                 tcodes[t]);
        slot[tcodes[t]] = t;
    }
    TT_CHECK(np >= 0 && np <= ttk::kMaxP, "np must be in 0..%d", ttk::kMaxP);
    for (int q = 0; q < np; q++) {
        TT_CHECK(pf, "pf is NULL");
        if (pf[q].code == TT_P_HERMITE)
            return fail(TT_ERR_UNSUPPORTED,
                        "Hermite P-functionals are not available in the batched path");
        TT_CHECK(pf[q].code >= 1 && pf[q].code <= 3,
                 "invalid P-functional code %d at position %d", pf[q].code, q);
    }
    int nmax = 0;
    for (int i = 0; i < nimg; i++) {
        TT_CHECK(images[i].d_image && images[i].d_sino,
                 "image %d: NULL image or sinogram pointer", i);
        TT_CHECK(images[i].n >= 3 && images[i].n <= 2730,
                 "image %d: padded side %d outside 3..2730", i, images[i].n);
        TT_CHECK(np == 0 || images[i].d_circus, "image %d: NULL circus pointer", i);
        nmax = images[i].n > nmax ? images[i].n : nmax;
    }
    TT_TRY(ctx->set_device());

    std::vector<int64_t> key = {nimg, n_angles, (int64_t)(*(const uint32_t *)&angle_step_deg),
                                nt, np, ctx->autotune};
    for (int t = 0; t < nt; t++)
        key.push_back(tcodes[t]);
    for (int q = 0; q < np; q++)
        key.push_back(pf[q].code);
    for (int i = 0; i < nimg; i++) {
        key.push_back((int64_t)(intptr_t)images[i].d_image);
        key.push_back(images[i].n);
        key.push_back((int64_t)(intptr_t)images[i].d_sino);
        key.push_back((int64_t)(intptr_t)images[i].d_circus);
        key.push_back((int64_t)(intptr_t)images[i].d_saved);
    }
    BatchPlan &B = ctx->batch;
    hipStream_t own = ctx->own;
    // everything below (autotune trials included) reads the caller's images
    if (!ctx->ev_in)
        TT_HIP(hipEventCreateWithFlags(&ctx->ev_in, hipEventDisableTiming));
    if (!ctx->ev_out)
        TT_HIP(hipEventCreateWithFlags(&ctx->ev_out, hipEventDisableTiming));
    if (ctx->stream != own) {
        TT_HIP(hipEventRecord(ctx->ev_in, ctx->stream));
        TT_HIP(hipStreamWaitEvent(own, ctx->ev_in, 0));
    }
    const bool do345 = slot[3] >= 0 || slot[4] >= 0 || slot[5] >= 0;
    const bool need_x = slot[1] >= 0 || slot[2] >= 0;
    const bool need_s = do345;
    if (!B.valid || B.key != key) {
        B.drop_graph();
        // a replay of the previous plan may still read the descriptors
        TT_HIP(hipStreamSynchronize(own));
        // largest images first: the dynamic counter then hands out the long
        // items early and the short ones fill the tail
        std::vector<int> order(nimg);
        for (int i = 0; i < nimg; i++)
            order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return images[a].n > images[b].n;
        });
        std::vector<ttk::BatchImage> d(nimg);
        size_t tot = 0;
        for (int i = 0; i < nimg; i++)
            tot += (size_t)images[i].n * images[i].n;
        TT_HIP(B.imgT.ensure(sizeof(float) * tot));
        TT_HIP(B.bbox.ensure(ttk::kBboxInts * sizeof(int) * (size_t)nimg));
        size_t ck = 0;  // the line kernel's walk checkpoints, image after image
        for (int i = 0; i < nimg; i++)
            ck += (size_t)ttk::sino_ckpt_blocks(images[i].n) * images[i].n * n_angles;
        TT_HIP(B.ckpt.ensure(sizeof(float2) * ck));
        bool p3 = false;
        for (int q = 0; q < np; q++)
            p3 = p3 || pf[q].code == TT_P3;
        size_t spec_tot = 0;
        for (int i = 0; i < nimg && p3; i++)
            spec_tot += (size_t)nt * n_angles * 2 * (images[i].n / 2 + 1);
        if (p3)
            TT_HIP(B.spec.ensure(sizeof(float) * spec_tot));
        std::vector<ttk::DftJob> jobs;
        int dft_tiles = 0;
        size_t off = 0, cko = 0, so = 0;
        int cols = 0;
        bool p2 = false;
        for (int q = 0; q < np; q++)
            p2 = p2 || pf[q].code == TT_P2;
        for (int k = 0; k < nimg; k++) {
            const tt_image &im = images[order[k]];
            ttk::BatchImage &b = d[k];
            b.spec = nullptr;
            b.p2src = nullptr;
            unsigned char *sv = (unsigned char *)im.d_saved;
            const SavedLayout SL = saved_layout(im.n, n_angles, nt);
            if (sv && p2 && nmax <= 2048)
                b.p2src = (int32_t *)(sv + SL.p2src);
            if (p3) {
                const int h2 = 2 * (im.n / 2 + 1);
                b.spec = sv ? (float *)(sv + SL.spec) : B.spec.as<float>() + so;
                so += (size_t)nt * n_angles * h2;
                ttk::DftJob j = {};
                j.in = im.d_sino;
                j.out = b.spec;
                j.n = im.n;
                j.cols = nt * n_angles;
                j.ld_in = im.n;
                j.ld_out = h2;
                jobs.push_back(j);
            }
            b.img = im.d_image;
            b.bbox = B.bbox.as<int>() + ttk::kBboxInts * k;
            b.ckpt = B.ckpt.as<float2>() + cko;
            cko += (size_t)ttk::sino_ckpt_blocks(im.n) * im.n * n_angles;
            b.imgT = B.imgT.as<float>() + off;
            off += (size_t)im.n * im.n;
            b.sino = im.d_sino;
            // with d_saved the medians and T3..T5 integrals a backward needs
            b.med_x = sv && need_x ? (int32_t *)(sv + SL.med_x) : nullptr;
            b.med_s = sv && need_s ? (int32_t *)(sv + SL.med_s) : nullptr;
            b.z345 = sv && do345 ? (float2 *)(sv + SL.z345) : nullptr;
            b.n = im.n;
            b.col0 = cols;
            b.circus = im.d_circus;
            cols += nt * n_angles;
        }
        TT_HIP(B.desc.ensure(sizeof(ttk::BatchImage) * nimg));
        B.dft = {};
        int dft_bm = 64, dft_bn = 128;
        dft_tiles = layout_dft_jobs(jobs, true, &dft_bm, &dft_bn);
        if (p3) {
            TT_HIP(B.jobs.ensure(sizeof(ttk::DftJob) * jobs.size()));
            TT_HIP(hipMemcpy(B.jobs.ptr, jobs.data(), sizeof(ttk::DftJob) * jobs.size(),
                             hipMemcpyHostToDevice));
            B.dft.jobs = B.jobs.as<ttk::DftJob>();
            B.dft.njobs = (int)jobs.size();
            B.dft.tiles = dft_tiles;
            B.dft.bm = dft_bm;
            B.dft.bn = dft_bn;
        }
        std::vector<float2> h;
        make_trig(n_angles, angle_step_deg, h);
        TT_HIP(B.trig.ensure(sizeof(float2) * n_angles));
        TT_HIP(hipMemcpy(B.trig.ptr, h.data(), sizeof(float2) * n_angles,
                         hipMemcpyHostToDevice));
        make_tab345(nmax, h);
        TT_HIP(B.tab.ensure(sizeof(float2) * h.size()));
        TT_HIP(hipMemcpy(B.tab.ptr, h.data(), sizeof(float2) * h.size(),
                         hipMemcpyHostToDevice));
        TT_HIP(B.counter.ensure(sizeof(unsigned int)));

        ttk::BatchSinoParams sp = {};
        sp.images = B.desc.as<ttk::BatchImage>();
        sp.nimg = nimg;
        sp.trig = B.trig.as<float2>();
        sp.tab345 = B.tab.as<float2>();
        sp.tabn = nmax;
        sp.A = n_angles;
        sp.counter = B.counter.as<unsigned int>();
        for (int k = 0; k < 8; k++)
            sp.slot.s[k] = slot[k];
        // descriptors with the line items of tile t
        auto use_tile = [&](LineTile t) {
            sp.items = layout_line_items(d, n_angles, t);
            sp.tile_a = t.a;
            sp.tile_c = t.c;
            return hipMemcpy(B.desc.ptr, d.data(), sizeof(ttk::BatchImage) * nimg,
                             hipMemcpyHostToDevice);
        };
        LineTile best_t = line_tile_of(ctx);
        TT_HIP(use_tile(best_t));
        ttk::BatchCircusParams cp = {};
        cp.images = sp.images;
        cp.nimg = nimg;
        cp.cols = cols;
        cp.nt = nt;
        cp.A = n_angles;
        cp.nmax = nmax;
        cp.np = np;
        cp.with_circus = 1;
        for (int q = 0; q < np; q++)
            cp.code[q] = pf[q].code;

        // line tile and launch shape: forced, the defaults, or the fastest
        // candidates on this batch (tiles at the default shape, then shapes
        // with the chosen tile)
        const int cus = device_cus(ctx->device);
        int best_w = 4, best_b = cus * 4;
        if (ctx->autotune) {
            TT_HIP(ttk::launch_transpose_batched(sp.images, nimg, nmax, own));
            TT_HIP(ttk::launch_image_bbox_batched(sp.images, nimg, nmax, own));
            hipEvent_t e0, e1;
            TT_HIP(hipEventCreate(&e0));
            TT_HIP(hipEventCreate(&e1));
            // line kernel time with w waves per workgroup, blocks workgroups
            // (the mean of two runs after a first)
            auto trial = [&](int w, int blocks, float *ms_out) {
                float tsum = 0.0f;
                for (int rep = 0; rep < 3; rep++) {
                    hipError_t r = hipEventRecord(e0, own);
                    if (r == hipSuccess)
                        r = enqueue_batch(ctx, sp, cp, nimg, nmax, need_x, need_s, do345, w,
                                          blocks, false, false, own);
                    if (r == hipSuccess)
                        r = hipEventRecord(e1, own);
                    if (r == hipSuccess)
                        r = hipEventSynchronize(e1);
                    float ms = 0.0f;
                    if (r == hipSuccess)
                        r = hipEventElapsedTime(&ms, e0, e1);
                    if (r != hipSuccess)
                        return r;
                    if (rep > 0)
                        tsum += ms;
                }
                *ms_out = tsum;
                return hipSuccess;
            };
            float best = 1e30f;
            if (ctx->force.line_angles == 0) {
                for (const LineTile &t : kLineTiles) {
                    float ms = 0.0f;
                    TT_HIP(use_tile(t));
                    TT_HIP(trial(best_w, best_b, &ms));
                    if (ms < best) {
                        best = ms;
                        best_t = t;
                    }
                }
                TT_HIP(use_tile(best_t));
            }
            best = 1e30f;
            const int waves[] = {1, 2, 4};
            const int per_cu[] = {2, 4, 8, 16};
            for (int w : waves)
                for (int pc : per_cu) {
                    const int blocks = cus * pc * 4 / w;
                    float ms = 0.0f;
                    TT_HIP(trial(w, blocks, &ms));
                    if (ms < best) {
                        best = ms;
                        best_w = w;
                        best_b = blocks;
                    }
                }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        B.waves_per_block = best_w;
        B.blocks = best_b;
        B.tile_a = best_t.a;
        B.tile_c = best_t.c;
        B.sp = sp;
        B.cp = cp;
        B.nimg = nimg;
        B.nmax = nmax;
        B.need_x = need_x;
        B.need_s = need_s;
        B.do345 = do345;

        TT_HIP(hipStreamBeginCapture(own, hipStreamCaptureModeThreadLocal));
        hipError_t ce = enqueue_batch(ctx, sp, cp, nimg, nmax, need_x, need_s, do345,
                                      best_w, best_b, true, true, own);
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(own, &g);
        TT_HIP(ce);
        TT_HIP(ee);
        B.graph = g;
        TT_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
        B.key = key;
        B.valid = true;
        const std::vector<int64_t> fkey =
            saved_key(n_angles, angle_step_deg, tcodes, nt, pf, np);
        for (int i = 0; i < nimg; i++)
            if (images[i].d_saved)
                ctx->saved[images[i].d_saved] = SavedEntry{images[i].n, fkey, images[i].d_sino};
    }
    if (ctx->timer.on) {
        // per-stage kernel times: the captured stages launched directly, each
        // bracketed by HIP events
        TT_HIP(enqueue_batch(ctx, B.sp, B.cp, B.nimg, B.nmax, B.need_x, B.need_s, B.do345,
                             B.waves_per_block, B.blocks, true, true, own, true));
    } else {
        TT_HIP(hipGraphLaunch(B.exec, own));
    }
    if (ctx->stream != own) {
        TT_HIP(hipEventRecord(ctx->ev_out, own));
        TT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_out, 0));
    }
    return TT_OK;
}

namespace {

// Hermite weights of the P list into the context (rows in list order).
tt_status prepare_hermite(tt_context *ctx, const tt_pfunc *pf, int np, int n,
                          int *herm_row, const double **herm_w)
{
    std::vector<int32_t> hkey;
    for (int q = 0; q < np; q++) {
        herm_row[q] = -1;
        if (pf[q].code == TT_P_HERMITE) {
            herm_row[q] = (int)(hkey.size() / 2);
            hkey.push_back(pf[q].order);
            hkey.push_back(pf[q].center);
        }
    }
    *herm_w = nullptr;
    if (!hkey.empty())
        TT_TRY(ctx->ensure_hermite(n, hkey, herm_w));
    return TT_OK;
}

// P3's positions in a P list.
int p3_positions(const tt_pfunc *pf, int np, int *q3)
{
    int k = 0;
    for (int q = 0; q < np; q++)
        if (pf[q].code == TT_P3)
            q3[k++] = q;
    return k;
}

tt_status circus_vjp(tt_context *ctx, const float *d_sino, int n, int n_cols,
                     int n_sino, const tt_pfunc *pf, int np, const float *d_dcircus,
                     float *d_dsino)
{
    ttk::CircusVjpParams p = {};
    p.sino = d_sino;
    p.dcircus = d_dcircus;
    p.dsino = d_dsino;
    p.n = n;
    p.ncols = n_cols;
    p.nsino = n_sino;
    p.np = np;
    for (int q = 0; q < np; q++)
        p.code[q] = pf[q].code;
    TT_TRY(prepare_hermite(ctx, pf, np, n, p.herm, &p.herm_w));
    hipStream_t st = ctx->stream;
    ttk::P3ScaleParams sp = {};
    sp.np3 = p3_positions(pf, np, sp.q3);
    const int h = n / 2 + 1, cols = n_cols * n_sino;
    ttk::DftJob j = {};
    if (sp.np3 > 0) {
        // P3's part: spectra (W^T x), scaled per column and frequency, then
        // W y into d_dsino, which the VJP kernel adds the other parts to
        TT_HIP(ctx->p3_spec.ensure(sizeof(float) * (size_t)cols * 2 * h));
        j.in = d_sino;
        j.out = ctx->p3_spec.as<float>();
        j.n = n;
        j.cols = cols;
        j.ld_in = n;
        j.ld_out = 2 * h;
        j.g = d_dcircus;
        j.g_cols = n_cols;
        j.g_stride = np * n_cols;
        sp.job = j;
        sp.job.in = j.out;  // scaled in place
        sp.job.ld_in = 2 * h;
        sp.njobs = 1;
        p.p3_done = 1;
    }
    bool p2 = false;
    for (int q = 0; q < np; q++)
        p2 = p2 || pf[q].code == TT_P2;
    if (p2 && n <= 2048) {
        // P2's source rows by the one-wave register sort
        TT_HIP(ctx->bw_p2src.ensure(sizeof(int32_t) * (size_t)cols));
        p.p2src = ctx->bw_p2src.as<int32_t>();
    }
    if (sp.np3 > 0) {
        TT_HIP(ctx->timer.begin(TT_KERNEL_P3_DFT, st));
        ttk::DftParams f = {};
        f.njobs = 1;
        f.job = j;
        ttk::dft_tile_shape(ttk::dft_job_tiles(n, cols, true, 128, 128),
                            ttk::dft_job_tiles(n, cols, true, 64, 128), &f.bm, &f.bn);
        f.tiles = ttk::dft_job_tiles(n, cols, true, f.bm, f.bn);
        TT_HIP(ttk::launch_dft_gemm(f, true, n, st));
        TT_HIP(ttk::launch_p3_vjp_scale(sp, cols, st));
        ttk::DftParams b = {};
        b.njobs = 1;
        b.job.in = j.out;
        b.job.out = d_dsino;
        b.job.n = n;
        b.job.cols = cols;
        b.job.ld_in = 2 * h;
        b.job.ld_out = n;
        ttk::dft_tile_shape(ttk::dft_job_tiles(n, cols, false, 128, 128),
                            ttk::dft_job_tiles(n, cols, false, 64, 128), &b.bm, &b.bn);
        b.tiles = ttk::dft_job_tiles(n, cols, false, b.bm, b.bn);
        TT_HIP(ttk::launch_dft_gemm(b, false, n, st));
        TT_HIP(ctx->timer.end(st));
    }
    if (p.p2src) {
        TT_HIP(ctx->timer.begin(TT_KERNEL_CIRCUS, st));
        TT_HIP(ttk::launch_p2_source(d_sino, n, cols, ctx->bw_p2src.as<int32_t>(), st));
        TT_HIP(ctx->timer.end(st));
    }
    TT_HIP(ctx->timer.begin(TT_KERNEL_CIRCUS_VJP, st));
    TT_HIP(ttk::launch_circus_vjp(p, st));
    TT_HIP(ctx->timer.end(st));
    return TT_OK;
}

// Forward recompute (sinograms, medians, T3..T5 integrals) into the
// context's adjoint workspace; returns the sinogram workspace.
tt_status forward_for_adjoint(tt_context *ctx, const float *d_image, int n, int A,
                              float step, const int *slot, int nt, float **sino_out)
{
    const size_t lines = (size_t)n * A;
    TT_HIP(ctx->bw_sino.ensure(sizeof(float) * nt * lines));
    TT_HIP(ctx->bw_z.ensure(sizeof(float2) * 3 * lines));
    TT_HIP(ctx->med_x.ensure(sizeof(int32_t) * lines));
    TT_HIP(ctx->med_s.ensure(sizeof(int32_t) * lines));
    TT_TRY(ctx->ensure_trig(A, step));
    TT_TRY(ctx->ensure_tab345(n));
    TT_HIP(ctx->imgT.ensure(sizeof(float) * (size_t)n * n));
    TT_HIP(ctx->bbox.ensure(ttk::kBboxInts * sizeof(int)));
    hipStream_t st = ctx->stream;
    TT_HIP(ctx->timer.begin(TT_KERNEL_AUX, st));
    TT_HIP(ttk::launch_transpose(d_image, ctx->imgT.as<float>(), n, st));
    TT_HIP(ttk::launch_image_bbox(d_image, n, ctx->bbox.as<int>(), st));
    TT_HIP(ctx->timer.end(st));
    ttk::SinoParams p = {};
    p.bbox = ctx->bbox.as<int>();
    p.img = d_image;
    p.imgT = ctx->imgT.as<float>();
    p.trig = ctx->trig.as<float2>();
    p.tab345 = ctx->tab345.as<float2>();
    p.sino = ctx->bw_sino.as<float>();
    p.med_x = ctx->med_x.as<int32_t>();
    p.med_s = ctx->med_s.as<int32_t>();
    p.z345 = ctx->bw_z.as<float2>();
    TT_HIP(ctx->sino_ck.ensure(sizeof(float2) * ttk::sino_ckpt_blocks(n) * (size_t)n * A));
    p.ckpt = ctx->sino_ck.as<float2>();
    p.n = n;
    p.A = A;
    p.nb = (n + ttk::kWave - 1) / ttk::kWave;
    p.items = p.nb * A;
    for (int k = 0; k < 8; k++)
        p.slot.s[k] = slot[k];
    const bool do345 = slot[3] >= 0 || slot[4] >= 0 || slot[5] >= 0;
    TT_HIP(ctx->timer.begin(TT_KERNEL_SINO_T05, st));
    TT_HIP(ttk::launch_sino_t05(p, true, true, do345, st));
    TT_HIP(ctx->timer.end(st));
    if (slot[6] >= 0 || slot[7] >= 0) {
        // T6/T7 values and the sources their adjoints need
        TT_HIP(ctx->bw_t6src.ensure(sizeof(int32_t) * lines));
        TT_HIP(ctx->bw_t7src.ensure(sizeof(int32_t) * lines));
        ttk::SortTParams q = {};
        q.img = d_image;
        q.imgT = ctx->imgT.as<float>();
        q.trig = ctx->trig.as<float2>();
        q.sino = ctx->bw_sino.as<float>();
        q.med_x = ctx->med_x.as<int32_t>();
        q.med_s = ctx->med_s.as<int32_t>();
        q.t6_src = slot[6] >= 0 ? ctx->bw_t6src.as<int32_t>() : nullptr;
        q.t7_src = slot[7] >= 0 ? ctx->bw_t7src.as<int32_t>() : nullptr;
        q.bbox = ctx->bbox.as<int>();
        q.n = n;
        q.A = A;
        q.slot6 = slot[6];
        q.slot7 = slot[7];
        int32_t *qa = nullptr, *qb = nullptr;
        TT_HIP(ctx->sort_count.ensure(ttk::kSortCounters * sizeof(unsigned int)));
        unsigned int *counts = ctx->sort_count.as<unsigned int>();
        if (n > 512) {
            TT_HIP(ctx->sort_qa.ensure(sizeof(int32_t) * lines));  // a line item per line
            TT_HIP(ctx->sort_qb.ensure(sizeof(int32_t) * lines));
            qa = ctx->sort_qa.as<int32_t>();
            qb = ctx->sort_qb.as<int32_t>();
        }
        TT_HIP(ctx->timer.begin(TT_KERNEL_SINO_SORT, st));
        TT_HIP(ttk::launch_sino_sort(q, qa, qb, counts, st));
        TT_HIP(ctx->timer.end(st));
    }
    *sino_out = ctx->bw_sino.as<float>();
    return TT_OK;
}

// Launch shape of the transposed-rotation gather: 32 x 32 pixel tiles, the
// most waves per workgroup (16, 8, 4, 2, 1) whose LDS (the T3..T5 tables once
// plus each wave's accumulators and staged columns) fits the 160 KiB of a
// CU, and for a single image enough angle groups that tiles x groups waves
// fill every SIMD four times over.
struct GatherLaunch {
    int tile, waves, groups;
};

GatherLaunch choose_gather(int tiles32, int A, int cus, int tabn, bool do345, bool do12,
                           bool do67, bool split_angles)
{
    GatherLaunch g = {32, 1, 1};
    for (int waves : {16, 8, 4, 2, 1}) {
        if (ttk::gather_lds_bytes(32, waves, tabn, do345, do12, do67) <= kLdsLimit) {
            g.waves = waves;
            break;
        }
    }
    if (split_angles) {
        // Angle groups: the waves (tiles x groups) run in rounds of what the
        // device holds at once (16 waves per CU: the LDS and 128 VGPRs), each
        // wave over ceil(A / groups) angles; the group count minimises
        // rounds x angles per wave, plus a little for each partial image the
        // reduction reads (a count just above one round's worth would run a
        // second round for a few waves).
        const double cap = 16.0 * cus;
        double best = 1e30;
        const int gmax = A < 32 ? A : 32;
        for (int groups = 1; groups <= gmax; groups++) {
            const double rounds = ceil((double)tiles32 * groups / cap);
            const double cost = rounds * ceil((double)A / groups) + groups / 8.0;
            if (cost < best) {
                best = cost;
                g.groups = groups;
            }
        }
    }
    return g;
}

int tile_count(int n, int tile)
{
    const int t = (n + tile - 1) / tile;
    return t * t;
}


tt_status image_vjp(tt_context *ctx, int n, int A, const int *slot,
                    const float *d_dsino, float *d_dimage)
{
    const size_t lines = (size_t)n * A;
    const bool do67 = slot[6] >= 0 || slot[7] >= 0;
    TT_HIP(ctx->bw_rec.ensure(sizeof(ttk::LineRecord) * lines));
    if (do67)
        TT_HIP(ctx->bw_rec67.ensure(sizeof(ttk::LineRecord67) * lines));
    hipStream_t st = ctx->stream;
    ttk::RecordParams r = {};
    r.dsino = d_dsino;
    r.med_x = ctx->med_x.as<int32_t>();
    r.med_s = ctx->med_s.as<int32_t>();
    r.t6_src = slot[6] >= 0 ? ctx->bw_t6src.as<int32_t>() : nullptr;
    r.t7_src = slot[7] >= 0 ? ctx->bw_t7src.as<int32_t>() : nullptr;
    r.z345 = ctx->bw_z.as<float2>();
    r.rec = ctx->bw_rec.as<ttk::LineRecord>();
    r.rec67 = do67 ? ctx->bw_rec67.as<ttk::LineRecord67>() : nullptr;
    r.n = n;
    r.A = A;
    for (int k = 0; k < 8; k++)
        r.slot.s[k] = slot[k];
    ttk::GatherParams g = {};
    g.rec = r.rec;
    g.rec67 = r.rec67;
    g.trig = ctx->trig.as<float2>();
    g.tab345 = ctx->tab345.as<float2>();
    g.dimg = d_dimage;
    g.n = n;
    g.A = A;
    g.do345 = slot[3] >= 0 || slot[4] >= 0 || slot[5] >= 0;
    g.do12 = slot[1] >= 0 || slot[2] >= 0;
    const GatherLaunch gl = choose_gather(tile_count(n, 32), A, device_cus(ctx->device), n,
                                          g.do345, g.do12, do67, true);
    g.tile = gl.tile;
    g.waves = gl.waves;
    g.groups = gl.groups;
    if (g.groups > 1) {
        TT_HIP(ctx->bw_part.ensure(sizeof(float) * (size_t)g.groups * n * n));
        g.part = ctx->bw_part.as<float>();
    }
    TT_HIP(ctx->timer.begin(TT_KERNEL_RECORDS, st));
    TT_HIP(ttk::launch_records(r, st));
    TT_HIP(ctx->timer.end(st));
    TT_HIP(ctx->timer.begin(TT_KERNEL_GATHER, st));
    TT_HIP(ttk::launch_gather(g, st));
    TT_HIP(ctx->timer.end(st));
    return TT_OK;
}

// Output slot of each T-functional (slot[8]); T6/T7 only where with_sort.
tt_status adjoint_slots(const int32_t *tcodes, int nt, int *slot, bool with_sort)
{
    TT_TRY(check_tcodes(tcodes, nt));
    TT_CHECK(nt >= 1, "at least one T-functional is required");
    for (int k = 0; k < 8; k++)
        slot[k] = -1;
    for (int t = 0; t < nt; t++) {
        if (tcodes[t] > (with_sort ? 7 : 5))
            return fail(TT_ERR_UNSUPPORTED, "the batched adjoint covers T0..T5 (got T%d)",
                        tcodes[t]);
        TT_CHECK(slot[tcodes[t]] < 0, "duplicate T-functional %d", tcodes[t]);
        slot[tcodes[t]] = t;
    }
    return TT_OK;
}

}  // namespace

TT_API tt_status tt_circus_backward(tt_context *ctx, const float *d_sino, int n,
                                    int n_cols, int n_sino, const tt_pfunc *pf,
                                    int np, const float *d_dcircus, float *d_dsino)
{
    TT_CHECK(ctx && d_sino && d_dsino, "NULL argument");
    TT_CHECK(n >= 2 && n <= 1800, "sinogram height %d outside 2..1800", n);
    TT_CHECK(n_cols >= 1 && n_sino >= 1, "n_cols and n_sino must be >= 1");
    TT_TRY(check_pfuncs(pf, np, n));
    TT_CHECK(np == 0 || d_dcircus, "d_dcircus is NULL");
    TT_TRY(ctx->set_device());
    if (np == 0)
        TT_HIP(hipMemsetAsync(d_dsino, 0, sizeof(float) * (size_t)n * n_cols * n_sino,
                              ctx->stream));
    else
        TT_TRY(circus_vjp(ctx, d_sino, n, n_cols, n_sino, pf, np, d_dcircus, d_dsino));
    return TT_OK;
}

TT_API tt_status tt_sinograms_backward(tt_context *ctx, const float *d_image,
                                       int n, int n_angles, float angle_step_deg,
                                       const int32_t *tcodes, int nt,
                                       const float *d_dsino, float *d_dimage)
{
    TT_CHECK(ctx && d_image && d_dsino && d_dimage, "NULL argument");
    TT_CHECK(n >= 3 && n <= 1800, "padded side %d outside 3..1800", n);
    TT_CHECK(n_angles >= 1, "n_angles must be >= 1");
    TT_CHECK(angle_step_deg > 0.0f && isfinite(angle_step_deg),
             "angle step must be positive and finite");
    int slot[8];
    TT_TRY(adjoint_slots(tcodes, nt, slot, true));
    TT_TRY(ctx->set_device());
    float *sino = nullptr;
    TT_TRY(forward_for_adjoint(ctx, d_image, n, n_angles, angle_step_deg, slot, nt,
                               &sino));
    TT_TRY(image_vjp(ctx, n, n_angles, slot, d_dsino, d_dimage));
    return TT_OK;
}

TT_API tt_status tt_backward(tt_context *ctx, const float *d_image, int n,
                             int n_angles, float angle_step_deg,
                             const int32_t *tcodes, int nt, const tt_pfunc *pf,
                             int np, const float *d_dcircus, float *d_dimage)
{
    TT_CHECK(ctx && d_image && d_dimage, "NULL argument");
    TT_CHECK(n >= 3 && n <= 1800, "padded side %d outside 3..1800", n);
    TT_CHECK(n_angles >= 1, "n_angles must be >= 1");
    TT_CHECK(angle_step_deg > 0.0f && isfinite(angle_step_deg),
             "angle step must be positive and finite");
    int slot[8];
    TT_TRY(adjoint_slots(tcodes, nt, slot, true));
    TT_TRY(check_pfuncs(pf, np, n));
    TT_CHECK(np >= 1 && d_dcircus, "tt_backward needs P-functionals and d_dcircus");
    TT_TRY(ctx->set_device());
    float *sino = nullptr;
    TT_TRY(forward_for_adjoint(ctx, d_image, n, n_angles, angle_step_deg, slot, nt,
                               &sino));
    const size_t lines = (size_t)n * n_angles;
    TT_HIP(ctx->bw_dsino.ensure(sizeof(float) * nt * lines));
    TT_TRY(circus_vjp(ctx, sino, n, n_angles, nt, pf, np, d_dcircus,
                      ctx->bw_dsino.as<float>()));
    TT_TRY(image_vjp(ctx, n, n_angles, slot, ctx->bw_dsino.as<float>(), d_dimage));
    return TT_OK;
}

TT_API tt_status tt_resize(tt_context *ctx, const float *d_in, int in_rows,
                           int in_cols, float *d_out, int rows, int cols)
{
    TT_CHECK(ctx && d_in && d_out, "NULL argument");
    TT_CHECK(in_rows >= 2 && in_cols >= 2 && rows >= 1 && cols >= 1,
             "invalid sizes %d x %d -> %d x %d", in_rows, in_cols, rows, cols);
    TT_TRY(ctx->set_device());
    TT_HIP(ttk::launch_resize(d_in, in_rows, in_cols, d_out, rows, cols, ctx->stream));
    return TT_OK;
}

TT_API tt_status tt_zscore(tt_context *ctx, const float *d_in, int n, int count,
                           float *d_out)
{
    TT_CHECK(ctx && d_in && d_out, "NULL argument");
    TT_CHECK(n >= 2 && count >= 1, "need n >= 2 and count >= 1");
    TT_TRY(ctx->set_device());
    TT_HIP(ttk::launch_zscore(d_in, n, count, d_out, ctx->stream));
    return TT_OK;
}

// Polynomial schedule of the polar iteration X <- X q(X^T X), q(y) = c0 +
// c1 y + c2 y^2 + c3 y^3 + c4 y^4, so that each step maps every singular
// value x of X to x q(x^2) and leaves the singular vectors alone (generated
// by csrc/polar_schedule.py): eight growth steps that multiply singular
// values below ~0.01 by 7.28 without letting any fall below 0.1 of the
// upper end, two minimax steps and two steps of fifth-order convergence.
// From X_0 = A / ||A^T A||_F^(1/2) every singular value x in [1e-8, 1.02]
// ends within 2e-15 of 1; smaller ones (condition numbers beyond ~1e8) do
// not converge and the caller falls back to the eigendecomposition route.
static const ttk::PolarStep kPolarSteps[] = {
    {{7.2816756381049732, -66.063394187123038, 209.00402730908877, -256.85470452555404, 107.07762732431864}},
    {{7.2816756381049732, -66.063394187123038, 209.00402730908877, -256.85470452555404, 107.07762732431864}},
    {{7.2816756381049732, -66.063394187123038, 209.00402730908877, -256.85470452555404, 107.07762732431864}},
    {{7.2816756381049732, -66.063394187123038, 209.00402730908877, -256.85470452555404, 107.07762732431864}},
    {{7.2816756381049732, -66.063394187123038, 209.00402730908877, -256.85470452555404, 107.07762732431864}},
    {{7.2816756381049732, -66.063394187123038, 209.00402730908877, -256.85470452555404, 107.07762732431864}},
    {{7.2816756381049732, -66.063394187123038, 209.00402730908877, -256.85470452555404, 107.07762732431864}},
    {{7.2816756381049732, -66.063394187123038, 209.00402730908877, -256.85470452555404, 107.07762732431864}},
    {{8.6756202485986371, -63.534465944226177, 187.09388164456479, -221.83925019608748, 90.517402439679913}},
    {{2.6119822462388274, -3.7135489280995135, 3.3244544471543302, -1.4797355908332768, 0.25480859276722512}},
    {{1.9493886477464955, -1.2740525232146713, 0, 0.52460041621275599, -0.19993651672502183}},
    {{2.4609375, -3.28125, 2.953125, -1.40625, 0.2734375}},
};
constexpr int kPolarStepCount = (int)(sizeof(kPolarSteps) / sizeof(kPolarSteps[0]));
// steps that update X itself; the others run on X^T X.  After four growth
// steps every singular value of X is above ~3e-5 of the largest for a
// condition number of 1e8 (7.28^4 times its start), and the Gram-space
// steps' error, about 1e-16 times the square of X's condition number at the
// switch, stays below the float output's rounding: in FP64 simulations the
// factor's normwise error is 2e-11 / 2e-10 / 6e-9 at condition numbers
// 1e6 / 1e7 / 1e8 (1.4e-9 at 1e8 with six tall steps, 3e-7 with three).
constexpr int kPolarTallSteps = 4;

TT_API tt_status tt_nearest_orthonormal_sinograms(tt_context *ctx,
                                                  const float *d_sino, int n,
                                                  int n_cols, int n_sino, float *d_out,
                                                  int out_capacity_rows,
                                                  int *out_rows, int *centers)
{
    TT_CHECK(ctx && d_sino && d_out && out_rows && centers, "NULL argument");
    TT_CHECK(n >= 2 && n <= 2730 && n_cols >= 1, "invalid sinogram size %d x %d", n,
             n_cols);
    TT_CHECK(n_sino >= 1, "n_sino must be >= 1 (got %d)", n_sino);
    TT_TRY(ctx->set_device());
    hipStream_t st = ctx->stream;
    const int A = n_cols;
    const size_t cols_all = (size_t)n_sino * A;
    // weighted medians of every column of every sinogram, one launch
    TT_HIP(ctx->nos_med.ensure(sizeof(int32_t) * cols_all));
    TT_HIP(ttk::launch_column_medians(d_sino, n, (int)cols_all, ctx->nos_med.as<int32_t>(),
                                      st));
    std::vector<int32_t> med(cols_all);
    TT_HIP(hipMemcpyAsync(med.data(), ctx->nos_med.ptr, sizeof(int32_t) * cols_all,
                          hipMemcpyDeviceToHost, st));
    TT_HIP(hipStreamSynchronize(st));
    // c++/src/circus.cpp:72-84: offsets to floor((n-1)/2), aligned height
    // n + |max| + |min| and the Hermite centre floor((n-1)/2) + max
    const int sc = (int)floor((n - 1) / 2.0);
    int rows_max = 0;
    for (int t = 0; t < n_sino; t++) {
        int mn = 1 << 30, mx = -(1 << 30);
        for (int c = 0; c < A; c++) {
            const int32_t o = med[(size_t)t * A + c] - sc;
            mn = o < mn ? o : mn;
            mx = o > mx ? o : mx;
        }
        const int rows = n + abs(mx) + abs(mn);
        TT_CHECK(rows >= A, "sinogram %d: aligned sinogram %d x %d is wider than tall", t,
                 rows, A);
        TT_CHECK(out_capacity_rows >= rows, "sinogram %d: output holds %d rows, %d needed",
                 t, out_capacity_rows, rows);
        out_rows[t] = rows;
        centers[t] = sc + mx;
        rows_max = rows > rows_max ? rows : rows_max;
    }
    // Every aligned matrix is zero-padded to rows_max rows so one batched
    // solve covers them all; zero rows leave the polar factor's other rows
    // unchanged and stay exactly zero in it.  rows_max is rounded up to
    // kGramChunks row chunks of a multiple of 8 rows, over which the Gram
    // GEMMs are split (more workgroups than the 360 x 360 outputs give).
    constexpr int kGramChunks = 3;
    const int chunk_rows = ((rows_max + kGramChunks - 1) / kGramChunks + 7) / 8 * 8;
    rows_max = chunk_rows * kGramChunks;
    const size_t mA = (size_t)rows_max * A, AA = (size_t)A * A;
    TT_HIP(ctx->nos_a.ensure(sizeof(double) * mA * n_sino));
    TT_HIP(ctx->nos_u.ensure(sizeof(double) * mA * n_sino));
    TT_HIP(ctx->nos_c.ensure(sizeof(double) * mA * n_sino));
    TT_HIP(ctx->nos_v.ensure(sizeof(double) * AA * n_sino));
    TT_HIP(ctx->nos_w.ensure(sizeof(double) * AA * n_sino));
    TT_HIP(ctx->nos_m.ensure(sizeof(double) * AA * n_sino));
    TT_HIP(ctx->nos_s.ensure(sizeof(double) * (size_t)A * n_sino));
    TT_HIP(ctx->nos_e.ensure(sizeof(double) * (size_t)A * n_sino));
    TT_HIP(ctx->nos_info.ensure(sizeof(rocblas_int) * n_sino));
    TT_HIP(ctx->nos_sweeps.ensure(sizeof(rocblas_int) * n_sino));
    TT_HIP(ctx->nos_q.ensure(sizeof(double) * AA * n_sino));
    TT_HIP(ctx->nos_sp.ensure(sizeof(double) * AA * n_sino * kGramChunks));
    TT_HIP(ctx->nos_h.ensure(sizeof(double) * AA * n_sino * 4));
    TT_HIP(ctx->nos_r.ensure(sizeof(double) * 2 * n_sino));
    TT_HIP(ctx->nos_dev.ensure(sizeof(double) * ttk::kNosSlices * n_sino));
    TT_HIP(ctx->timer.begin(TT_KERNEL_NOS, st));
    // row r of aligned column c reads row r - (max - med) of the sinogram
    // column, max the sinogram's largest median (the kernels find it and the
    // aligned heights from the medians themselves)
    TT_HIP(ttk::launch_align_batched(d_sino, n, A, ctx->nos_med.as<int32_t>(), n_sino,
                                     rows_max, mA, ctx->nos_a.as<double>(), st));
    if (!ctx->blas)
        TT_BLAS(rocblas_create_handle(&ctx->blas));
    TT_BLAS(rocblas_set_stream(ctx->blas, st));
    TT_BLAS(rocblas_set_pointer_mode(ctx->blas, rocblas_pointer_mode_host));
    const double one = 1.0, zero = 0.0;
    double *a_d = ctx->nos_a.as<double>(), *g = ctx->nos_v.as<double>();
    const rocblas_stride sA = (rocblas_stride)mA, sG = (rocblas_stride)AA;
    // S = X^T X for the batch: one batched product over the n_sino *
    // kGramChunks row chunks (job table: chunk c of matrix m of each X
    // buffer into its partial product), then the chunk sums in order.  The
    // Gram matrices and every A x A product below are symmetric and run in
    // the matrix-core kernel's symmetric mode.
    double *sp = ctx->nos_sp.as<double>();
    double *xbuf[3] = {a_d, ctx->nos_c.as<double>(), ctx->nos_u.as<double>()};
    const int nb = n_sino * kGramChunks;
    std::vector<ttk::NosGemmJob> jobs((size_t)3 * nb);
    for (int b = 0; b < nb; b++) {
        const size_t off = mA * (size_t)(b / kGramChunks) +
                           (size_t)chunk_rows * (b % kGramChunks);
        for (int w = 0; w < 3; w++)
            jobs[(size_t)w * nb + b] = {xbuf[w] + off, xbuf[w] + off, sp + AA * b};
    }
    // The Gram-space steps after the first form M Q and Q Q in one batched
    // call of 2 n_sino products.  Their buffers: Q in nos_h's first quarter,
    // Q Q in its third, and M alternating between nos_q and the second
    // quarter (see the loop below); one job list for each parity of the step.
    double *hq = ctx->nos_h.as<double>(), *hm = hq + AA * n_sino, *hq2 = hq + 2 * AA * n_sino;
    double *qb = ctx->nos_q.as<double>();
    const size_t mq0 = jobs.size();
    for (int par = 0; par < 2; par++) {
        double *mcur = par == 0 ? qb : hm, *mnext = par == 0 ? hm : qb;
        for (int b = 0; b < 2 * n_sino; b++) {
            const size_t i = (size_t)(b % n_sino);
            const bool mq = b < n_sino;  // M Q, else Q Q
            jobs.push_back({(mq ? mcur : hq) + AA * i, hq + AA * i,
                            (mq ? mnext : hq2) + AA * i});
        }
    }
    TT_HIP(ctx->nos_ptrs.ensure(sizeof(ttk::NosGemmJob) * jobs.size()));
    TT_HIP(hipMemcpyAsync(ctx->nos_ptrs.ptr, jobs.data(), sizeof(ttk::NosGemmJob) * jobs.size(),
                          hipMemcpyHostToDevice, st));
    const ttk::NosGemmJob *jobs_d = ctx->nos_ptrs.as<ttk::NosGemmJob>();
    auto gram = [&](const double *x, double *out) -> tt_status {
        int w = 0;
        while (w < 3 && xbuf[w] != x)
            w++;
        if (w == 3)
            return fail(TT_ERR_INVALID, "internal: unknown NOS buffer");
        const ttk::NosGemmBatch bt{jobs_d + (size_t)w * nb, nullptr, nullptr, nullptr, 0, 0, 0};
        TT_HIP(ttk::launch_nos_gemm(bt, nb, true, true, A, A, chunk_rows, rows_max, rows_max, A,
                                    1.0, 0.0, st));
        TT_HIP(ttk::launch_nos_reduce(sp, kGramChunks, A, n_sino, out, st));
        return TT_OK;
    };
    // Polar factor by the polynomial iteration X <- X q(X^T X) (every step
    // batched GEMMs in FP64 on the matrix cores): from X_0 = A r,
    // r = ||A^T A||_F^(-1/2), the schedule drives every singular value to
    // 1 while the singular vectors stay those of A, so X converges to
    // U V^T.  Rounding in each step perturbs the factor by about
    // eps * kappa(A) relative (the reference's FP32 Jacobi SVD carries
    // about 6e-8 * kappa(A)).  Convergence is checked on X^T X (the last
    // Gram-space S, equal to it up to rounding); a matrix
    // beyond the schedule's reach (kappa above ~1e8, rank deficiency) falls
    // back to the eigendecomposition route below.
    double *result = nullptr;
    {
        double *S = g, *S2 = ctx->nos_w.as<double>(), *T = ctx->nos_m.as<double>();
        double *Q = ctx->nos_q.as<double>(), *scale = ctx->nos_r.as<double>();
        double *r = scale + n_sino, *part = ctx->nos_dev.as<double>();
        double *xa = ctx->nos_c.as<double>(), *xb = ctx->nos_u.as<double>();
        const double *x = a_d;
        TT_TRY(gram(a_d, S));
        TT_HIP(ttk::launch_nos_scales(S, A, n_sino, part, scale, r, st));
        // C = P Q + beta C for batches of A x A matrices (polynomials of
        // one Gram matrix, so the products are symmetric)
        auto mul = [&](const double *P, const double *Qm, double beta, double *C) {
            const ttk::NosGemmBatch bt{nullptr, P, Qm, C, (long long)AA, (long long)AA,
                                       (long long)AA};
            return ttk::launch_nos_gemm(bt, n_sino, false, true, A, A, A, A, A, A, 1.0, beta,
                                        st);
        };
        // Y = X P for the batch's rows_max x A matrices X
        auto tall = [&](const double *X, const double *P, double *Y) {
            return rocblas_dgemm_strided_batched(ctx->blas, rocblas_operation_none,
                                                 rocblas_operation_none, rows_max, A, A, &one,
                                                 X, rows_max, sA, P, A, sG, &zero, Y, rows_max,
                                                 sA, n_sino);
        };
        // Q = q(S) of the step (the first step scales by s and r)
        auto poly = [&](int step) -> tt_status {
            ttk::NosPolyOut po{T, Q, (long long)AA, step == 0 ? scale : nullptr,
                               step == 0 ? r : nullptr, {}};
            for (int j = 0; j < 5; j++)
                po.c[j] = kPolarSteps[step].c[j];
            if (A % 4 == 0) {
                TT_HIP(ttk::launch_nos_square_poly(S, S2, A, n_sino, po, st));
            } else {
                TT_HIP(mul(S, S, 0.0, S2));
                TT_HIP(ttk::launch_nos_poly(S, S2, A, n_sino, kPolarSteps[step], po.scale,
                                            po.r, T, Q, st));
            }
            TT_HIP(mul(S2, T, 1.0, Q));
            return TT_OK;
        };
        // The first kPolarTallSteps steps update X itself; once X is better
        // conditioned the rest run in the Gram space, exact in exact
        // arithmetic as q(S) commutes with S: S <- S q(S)^2 and M <- M q(S),
        // and X <- X M once at the end.  Forming S squares only X's
        // condition number at the switch (at most ~3e4 for kappa(A) = 1e8).
        double *M = ctx->nos_h.as<double>(), *Mn = M + AA * n_sino;
        double *Q2 = M + 2 * AA * n_sino, *Sn = M + 3 * AA * n_sino;
        for (int step = 0; step < kPolarStepCount; step++) {
            if (step > 0 && step <= kPolarTallSteps)
                TT_TRY(gram(x, S));
            TT_TRY(poly(step));
            if (step < kPolarTallSteps) {
                TT_BLAS(tall(x, Q, xa));
                x = xa;
                std::swap(xa, xb);
                continue;
            }
            if (step == kPolarTallSteps) {
                std::swap(M, Q);  // M = Q (Q gets M's buffer)
                TT_HIP(mul(M, M, 0.0, Q2));
            } else {
                // M Q and Q Q together from the parity's job list
                const int par = (step - kPolarTallSteps - 1) & 1;
                if (Q != hq || Q2 != hq2 || M != (par == 0 ? qb : hm) ||
                    Mn != (par == 0 ? hm : qb))
                    return fail(TT_ERR_INVALID, "internal: NOS buffer rotation");
                const ttk::NosGemmBatch bt{jobs_d + mq0 + (size_t)par * 2 * n_sino, nullptr,
                                           nullptr, nullptr, 0, 0, 0};
                TT_HIP(ttk::launch_nos_gemm(bt, 2 * n_sino, false, true, A, A, A, A, A, A, 1.0,
                                            0.0, st));
                std::swap(M, Mn);
            }
            TT_HIP(mul(S, Q2, 0.0, Sn));
            std::swap(S, Sn);
        }
        TT_BLAS(tall(x, M, xa));
        x = xa;
        // convergence: S (= X^T X up to rounding) against I
        TT_HIP(ttk::launch_nos_deviation(S, A, n_sino, part, st));
        // the check's partials into pinned host memory and the float output
        // written ahead of the check: one synchronisation, and none more
        // unless a matrix falls back
        const size_t ndev = (size_t)ttk::kNosSlices * n_sino;
        if (ctx->h_dev_cap < ndev) {
            if (ctx->h_dev)
                TT_HIP(hipHostFree(ctx->h_dev));
            ctx->h_dev = nullptr;
            ctx->h_dev_cap = 0;
            TT_HIP(hipHostMalloc((void **)&ctx->h_dev, sizeof(double) * ndev));
            ctx->h_dev_cap = ndev;
        }
        TT_HIP(hipMemcpyAsync(ctx->h_dev, part, sizeof(double) * ndev,
                              hipMemcpyDeviceToHost, st));
        TT_HIP(ttk::launch_to_float_batched(x, rows_max, mA, A, ctx->nos_med.as<int32_t>(), n,
                                            sc, n_sino, rows_max, d_out,
                                            (size_t)out_capacity_rows * A, st));
        TT_HIP(ctx->timer.end(st));
        TT_HIP(hipStreamSynchronize(st));
        bool converged = true;
        for (size_t i = 0; i < ndev; i++)
            converged = converged && ctx->h_dev[i] <= 1e-10;  // false for NaN
        if (converged)
            return TT_OK;
        TT_HIP(ctx->timer.begin(TT_KERNEL_NOS, st));
    }
    std::vector<rocblas_int> info((size_t)n_sino);
    bool gram_ok = true;
    if (!result) {
        // Eigendecomposition route: P = A (A^T A)^(-1/2) with the Gram
        // matrix's eigenvectors by divide and conquer, M = V diag(lambda^-1/2)
        // V^T and P = A M, then two Newton-Schulz steps
        // P <- P (1.5 I - 0.5 P^T P) that restore orthonormality to rounding
        // (accurate to about eps * kappa(A)^2 relative).
        double *lam = ctx->nos_s.as<double>();
        rocblas_int *info_d = ctx->nos_info.as<rocblas_int>();
        TT_TRY(gram(a_d, g));
        TT_BLAS(rocsolver_dsyevd_strided_batched(ctx->blas, rocblas_evect_original,
                                                 rocblas_fill_upper, A, g, A, sG, lam,
                                                 (rocblas_stride)A, ctx->nos_e.as<double>(),
                                                 (rocblas_stride)A, info_d, n_sino));
        std::vector<double> lam_h((size_t)A * n_sino);
        TT_HIP(hipMemcpyAsync(lam_h.data(), lam, sizeof(double) * lam_h.size(),
                              hipMemcpyDeviceToHost, st));
        TT_HIP(hipMemcpyAsync(info.data(), info_d, sizeof(rocblas_int) * n_sino,
                              hipMemcpyDeviceToHost, st));
        TT_HIP(hipStreamSynchronize(st));
        for (int t = 0; t < n_sino; t++) {
            // ascending eigenvalues; a (numerically) singular Gram matrix has
            // no inverse square root
            const double lo = lam_h[(size_t)t * A], hi = lam_h[(size_t)t * A + A - 1];
            if (info[t] != 0 || !(hi > 0.0) || !(lo > hi * 1e-15))
                gram_ok = false;
        }
        if (gram_ok) {
            double *vs = ctx->nos_w.as<double>(), *m = ctx->nos_m.as<double>();
            double *p0 = ctx->nos_c.as<double>(), *p1 = ctx->nos_u.as<double>();
            TT_HIP(ttk::launch_scale_eigvec(g, lam, A, n_sino, vs, st));
            TT_BLAS(rocblas_dgemm_strided_batched(
                ctx->blas, rocblas_operation_none, rocblas_operation_transpose, A, A, A, &one,
                vs, A, sG, g, A, sG, &zero, m, A, sG, n_sino));
            TT_BLAS(rocblas_dgemm_strided_batched(
                ctx->blas, rocblas_operation_none, rocblas_operation_none, rows_max, A, A, &one,
                a_d, rows_max, sA, m, A, sG, &zero, p0, rows_max, sA, n_sino));
            for (int it = 0; it < 2; it++) {
                TT_TRY(gram(p0, m));
                TT_HIP(ttk::launch_newton_schulz_factor(m, A, n_sino, st));
                TT_BLAS(rocblas_dgemm_strided_batched(
                    ctx->blas, rocblas_operation_none, rocblas_operation_none, rows_max, A, A,
                    &one, p0, rows_max, sA, m, A, sG, &zero, p1, rows_max, sA, n_sino));
                std::swap(p0, p1);
            }
            result = p0;
        } else {
            // Rank-deficient input: thin U from a Jacobi SVD (V from A^T A, U
            // from a QR of A V, orthonormal even where sigma = 0); NOS = U V^T
            TT_BLAS(rocsolver_dgesvdj_strided_batched(
                ctx->blas, rocblas_svect_singular, rocblas_svect_all, rows_max, A, a_d,
                rows_max, sA, 0.0, ctx->nos_e.as<double>(), 100,
                ctx->nos_sweeps.as<rocblas_int>(), lam, (rocblas_stride)A,
                ctx->nos_u.as<double>(), rows_max, sA, g, A, sG, info_d, n_sino));
            TT_BLAS(rocblas_dgemm_strided_batched(
                ctx->blas, rocblas_operation_none, rocblas_operation_none, rows_max, A, A, &one,
                ctx->nos_u.as<double>(), rows_max, sA, g, A, sG, &zero,
                ctx->nos_c.as<double>(), rows_max, sA, n_sino));
            result = ctx->nos_c.as<double>();
            TT_HIP(hipMemcpyAsync(info.data(), info_d, sizeof(rocblas_int) * n_sino,
                                  hipMemcpyDeviceToHost, st));
        }
    }
    TT_HIP(ttk::launch_to_float_batched(result, rows_max, mA, A, ctx->nos_med.as<int32_t>(), n,
                                        sc, n_sino, rows_max, d_out,
                                        (size_t)out_capacity_rows * A, st));
    TT_HIP(ctx->timer.end(st));
    TT_HIP(hipStreamSynchronize(st));
    if (!gram_ok)
        for (int t = 0; t < n_sino; t++)
            if (info[t] != 0)
                return fail(TT_ERR_DEVICE,
                            "sinogram %d: Jacobi SVD did not converge (info %d)", t,
                            (int)info[t]);
    return TT_OK;
}

TT_API tt_status tt_nearest_orthonormal_sinogram(tt_context *ctx,
                                                 const float *d_sino, int n,
                                                 int n_cols, float *d_out,
                                                 int out_capacity_rows,
                                                 int *out_rows, int *center)
{
    return tt_nearest_orthonormal_sinograms(ctx, d_sino, n, n_cols, 1, d_out,
                                            out_capacity_rows, out_rows, center);
}

TT_API tt_status tt_backward_batched(tt_context *ctx, const tt_image_grad *images,
                                     int nimg, int n_angles, float angle_step_deg,
                                     const int32_t *tcodes, int nt,
                                     const tt_pfunc *pf, int np)
{
    TT_CHECK(ctx, "ctx is NULL");
    TT_CHECK(images && nimg >= 1, "images is NULL or nimg < 1");
    TT_CHECK(n_angles >= 1, "n_angles must be >= 1 (got %d)", n_angles);
    TT_CHECK(angle_step_deg > 0.0f && isfinite(angle_step_deg),
             "angle step must be positive and finite");
    int slot[8];
    TT_TRY(adjoint_slots(tcodes, nt, slot, false));
    TT_CHECK(np >= 1 && np <= ttk::kMaxP && pf, "np must be in 1..%d", ttk::kMaxP);
    bool p2 = false;
    for (int q = 0; q < np; q++) {
        if (pf[q].code == TT_P_HERMITE)
            return fail(TT_ERR_UNSUPPORTED,
                        "Hermite P-functionals are not available in the batched path");
        TT_CHECK(pf[q].code >= 1 && pf[q].code <= 3,
                 "invalid P-functional code %d at position %d", pf[q].code, q);
        p2 = p2 || pf[q].code == TT_P2;
    }
    int nmax = 0, nsaved = 0;
    for (int i = 0; i < nimg; i++) {
        TT_CHECK(images[i].d_image && images[i].d_dcircus && images[i].d_dimage,
                 "image %d: NULL pointer", i);
        TT_CHECK(images[i].n >= 3 && images[i].n <= 1800,
                 "image %d: padded side %d outside 3..1800", i, images[i].n);
        nmax = images[i].n > nmax ? images[i].n : nmax;
        nsaved += images[i].d_saved != nullptr;
    }
    TT_CHECK(nsaved == 0 || nsaved == nimg,
             "d_saved is set for %d of %d images (all or none)", nsaved, nimg);
    const bool saved = nsaved > 0;
    if (saved) {
        const std::vector<int64_t> fkey =
            saved_key(n_angles, angle_step_deg, tcodes, nt, pf, np);
        for (int i = 0; i < nimg; i++) {
            TT_CHECK(images[i].d_sino, "image %d: d_saved without d_sino", i);
            auto it = ctx->saved.find(images[i].d_saved);
            TT_CHECK(it != ctx->saved.end() && it->second.n == images[i].n &&
                         it->second.key == fkey && it->second.sino == images[i].d_sino,
                     "image %d: d_saved was not written by tt_forward_batched with this "
                     "image size, its sinograms, angles and functionals", i);
        }
    }
    TT_TRY(ctx->set_device());
    std::vector<int64_t> key = {nimg, n_angles, (int64_t)(*(const uint32_t *)&angle_step_deg),
                                nt, np, saved, ctx->autotune};
    for (int t = 0; t < nt; t++)
        key.push_back(tcodes[t]);
    for (int q = 0; q < np; q++)
        key.push_back(pf[q].code);
    for (int i = 0; i < nimg; i++) {
        key.push_back((int64_t)(intptr_t)images[i].d_image);
        key.push_back(images[i].n);
        key.push_back((int64_t)(intptr_t)images[i].d_dcircus);
        key.push_back((int64_t)(intptr_t)images[i].d_dimage);
        key.push_back((int64_t)(intptr_t)images[i].d_sino);
        key.push_back((int64_t)(intptr_t)images[i].d_saved);
    }
    BackwardPlan &P = ctx->bwd;
    BatchPlan &B = P.g;
    hipStream_t own = ctx->own;
    const bool do345 = slot[3] >= 0 || slot[4] >= 0 || slot[5] >= 0;
    const bool do12 = slot[1] >= 0 || slot[2] >= 0;
    if (!B.valid || B.key != key) {
        B.drop_graph();
        // a replay of the previous plan may still read the descriptors
        TT_HIP(hipStreamSynchronize(own));
        std::vector<int> order(nimg);
        for (int i = 0; i < nimg; i++)
            order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return images[a].n > images[b].n;
        });
        size_t sq = 0, ln = 0, spec_tot = 0;
        for (int i = 0; i < nimg; i++) {
            sq += (size_t)images[i].n * images[i].n;
            ln += (size_t)images[i].n * n_angles;
            spec_tot += (size_t)nt * n_angles * 2 * (images[i].n / 2 + 1);
        }
        ttk::P3ScaleParams p3s = {};
        p3s.np3 = p3_positions(pf, np, p3s.q3);
        const bool p2src = p2 && nmax <= 2048;
        // without saved line data the forward is recomputed into workspaces
        if (!saved) {
            TT_HIP(B.imgT.ensure(sizeof(float) * sq));
            TT_HIP(P.sino.ensure(sizeof(float) * nt * ln));
            TT_HIP(P.z.ensure(sizeof(float2) * 3 * ln));
            TT_HIP(P.medx.ensure(sizeof(int32_t) * ln));
            TT_HIP(P.meds.ensure(sizeof(int32_t) * ln));
            TT_HIP(B.bbox.ensure(ttk::kBboxInts * sizeof(int) * (size_t)nimg));
            size_t ck = 0;  // the line kernel's walk checkpoints, image after image
            for (int i = 0; i < nimg; i++)
                ck += (size_t)ttk::sino_ckpt_blocks(images[i].n) * images[i].n * n_angles;
            TT_HIP(B.ckpt.ensure(sizeof(float2) * ck));
            if (p3s.np3 > 0)
                TT_HIP(P.spec.ensure(sizeof(float) * spec_tot));
            if (p2src)
                TT_HIP(P.p2src.ensure(sizeof(int32_t) * nt * (size_t)n_angles * nimg));
        }
        TT_HIP(P.dsino.ensure(sizeof(float) * nt * ln));
        TT_HIP(P.rec.ensure(sizeof(ttk::LineRecord) * ln));
        if (p3s.np3 > 0)
            TT_HIP(P.y.ensure(sizeof(float) * spec_tot));
        // gather tile: forced, 32 pixels, or tuned below
        int gtile = ctx->force.gather_tile > 0 ? ctx->force.gather_tile : 32;
        int gwaves = ctx->force.gather_waves > 0 ? ctx->force.gather_waves
                                                 : gather_waves_max(gtile, nmax, do345, do12);
        TT_CHECK(gwaves >= 1 && ttk::gather_lds_bytes(gtile, gwaves, nmax, do345, do12,
                                                      false) <= kLdsLimit,
                 "gather tile %d with %d waves per workgroup exceeds the LDS", gtile, gwaves);
        std::vector<ttk::DftJob> fj, sj, bj;
        int ftiles = 0, btiles = 0;
        size_t so = 0;
        std::vector<ttk::BatchImage> d(nimg);
        size_t po = 0, lo = 0, cko = 0;
        int cols = 0;
        for (int k = 0; k < nimg; k++) {
            const tt_image_grad &im = images[order[k]];
            ttk::BatchImage &b = d[k];
            const int h2 = 2 * (im.n / 2 + 1);
            const size_t l = (size_t)im.n * n_angles;
            b.img = im.d_image;
            b.n = im.n;
            b.col0 = cols;
            b.circus = nullptr;
            b.dcircus = im.d_dcircus;
            b.dsino = P.dsino.as<float>() + nt * lo;
            b.rec = P.rec.as<ttk::LineRecord>() + lo;
            b.dimg = im.d_dimage;
            b.line0 = (int)lo;
            b.pix0 = (int)po;
            if (saved) {
                const SavedLayout L = saved_layout(im.n, n_angles, nt);
                unsigned char *base = (unsigned char *)images[order[k]].d_saved;
                b.sino = const_cast<float *>(im.d_sino);
                b.med_x = (int32_t *)(base + L.med_x);
                b.med_s = (int32_t *)(base + L.med_s);
                b.z345 = (float2 *)(base + L.z345);
                b.spec = (float *)(base + L.spec);
                b.p2src = p2src ? (int32_t *)(base + L.p2src) : nullptr;
            } else {
                b.bbox = B.bbox.as<int>() + ttk::kBboxInts * k;
                b.ckpt = B.ckpt.as<float2>() + cko;
                cko += (size_t)ttk::sino_ckpt_blocks(im.n) * im.n * n_angles;
                b.imgT = B.imgT.as<float>() + po;
                b.sino = P.sino.as<float>() + nt * lo;
                b.med_x = P.medx.as<int32_t>() + lo;
                b.med_s = P.meds.as<int32_t>() + lo;
                b.z345 = P.z.as<float2>() + 3 * lo;
                b.spec = p3s.np3 > 0 ? P.spec.as<float>() + so : nullptr;
                b.p2src = p2src ? P.p2src.as<int32_t>() + (size_t)nt * n_angles * k : nullptr;
            }
            if (p3s.np3 > 0) {
                // P3's adjoint: spectra of the sinograms (unless saved),
                // scaled per column into y, then W y into the sinogram gradient
                ttk::DftJob j = {};
                j.in = b.sino;
                j.out = b.spec;
                j.n = im.n;
                j.cols = nt * n_angles;
                j.ld_in = im.n;
                j.ld_out = h2;
                fj.push_back(j);
                ttk::DftJob js = j;
                js.in = b.spec;
                js.ld_in = h2;
                js.out = P.y.as<float>() + so;
                js.g = im.d_dcircus;
                js.g_cols = n_angles;
                js.g_stride = np * n_angles;
                sj.push_back(js);
                ttk::DftJob jb = j;
                jb.in = js.out;
                jb.out = b.dsino;
                jb.ld_in = h2;
                jb.ld_out = im.n;
                bj.push_back(jb);
            }
            so += (size_t)nt * n_angles * h2;
            cols += nt * n_angles;
            po += (size_t)im.n * im.n;
            lo += l;
        }
        TT_CHECK(po < (1ull << 31) && lo < (1ull << 31), "batch too large");
        // the recompute's line tile (forced or the default); gather tiles
        const LineTile lt = line_tile_of(ctx);
        const int items = layout_line_items(d, n_angles, lt);
        int tiles = layout_gather_tiles(d, gtile);
        TT_HIP(B.desc.ensure(sizeof(ttk::BatchImage) * nimg));
        TT_HIP(hipMemcpy(B.desc.ptr, d.data(), sizeof(ttk::BatchImage) * nimg,
                         hipMemcpyHostToDevice));
        ttk::DftParams fdp = {}, bdp = {};
        ftiles = layout_dft_jobs(fj, true, &fdp.bm, &fdp.bn);
        btiles = layout_dft_jobs(bj, false, &bdp.bm, &bdp.bn);
        if (p3s.np3 > 0) {
            TT_HIP(P.fjobs.ensure(sizeof(ttk::DftJob) * fj.size()));
            TT_HIP(P.sjobs.ensure(sizeof(ttk::DftJob) * sj.size()));
            TT_HIP(P.bjobs.ensure(sizeof(ttk::DftJob) * bj.size()));
            TT_HIP(hipMemcpy(P.fjobs.ptr, fj.data(), sizeof(ttk::DftJob) * fj.size(),
                             hipMemcpyHostToDevice));
            TT_HIP(hipMemcpy(P.sjobs.ptr, sj.data(), sizeof(ttk::DftJob) * sj.size(),
                             hipMemcpyHostToDevice));
            TT_HIP(hipMemcpy(P.bjobs.ptr, bj.data(), sizeof(ttk::DftJob) * bj.size(),
                             hipMemcpyHostToDevice));
            fdp.jobs = P.fjobs.as<ttk::DftJob>();
            fdp.njobs = (int)fj.size();
            fdp.tiles = ftiles;
            bdp.jobs = P.bjobs.as<ttk::DftJob>();
            bdp.njobs = (int)bj.size();
            bdp.tiles = btiles;
            p3s.jobs = P.sjobs.as<ttk::DftJob>();
            p3s.njobs = (int)sj.size();
        }
        std::vector<float2> h;
        make_trig(n_angles, angle_step_deg, h);
        TT_HIP(B.trig.ensure(sizeof(float2) * n_angles));
        TT_HIP(hipMemcpy(B.trig.ptr, h.data(), sizeof(float2) * n_angles,
                         hipMemcpyHostToDevice));
        make_tab345(nmax, h);
        TT_HIP(B.tab.ensure(sizeof(float2) * h.size()));
        TT_HIP(hipMemcpy(B.tab.ptr, h.data(), sizeof(float2) * h.size(),
                         hipMemcpyHostToDevice));
        TT_HIP(B.counter.ensure(sizeof(unsigned int)));

        ttk::BatchSinoParams sp = {};
        sp.images = B.desc.as<ttk::BatchImage>();
        sp.nimg = nimg;
        sp.items = items;
        sp.tile_a = lt.a;
        sp.tile_c = lt.c;
        sp.trig = B.trig.as<float2>();
        sp.tab345 = B.tab.as<float2>();
        sp.tabn = nmax;
        sp.A = n_angles;
        sp.counter = B.counter.as<unsigned int>();
        for (int k = 0; k < 8; k++)
            sp.slot.s[k] = slot[k];
        ttk::BatchCircusParams cp = {};
        cp.images = sp.images;
        cp.nimg = nimg;
        cp.cols = cols;
        cp.nt = nt;
        cp.A = n_angles;
        cp.nmax = nmax;
        cp.np = np;
        ttk::BatchAdjointParams ap = {};
        ap.images = sp.images;
        ap.nimg = nimg;
        ap.A = n_angles;
        ap.nt = nt;
        ap.np = np;
        ap.nmax = nmax;
        ap.lines = (int)lo;
        ap.pixels = (int)po;
        ap.tiles = tiles;
        ap.cols = cols;
        ap.trig = sp.trig;
        ap.tab345 = sp.tab345;
        ap.tabn = nmax;
        ap.do345 = do345;
        ap.slot = sp.slot;
        for (int q = 0; q < np; q++)
            ap.code[q] = pf[q].code;
        ap.p3_done = p3s.np3 > 0;
        const int cus = device_cus(ctx->device);
        B.waves_per_block = 4;
        B.blocks = cus * 4;
        B.sp = sp;
        B.cp = cp;
        B.nimg = nimg;
        B.nmax = nmax;
        B.do345 = do345;
        P.ap = ap;
        P.fdp = fdp;
        P.bdp = bdp;
        P.p3s = p3s;
        P.saved = saved;
        P.p2src_on = p2src;
        P.cols = nt * n_angles;
        B.tile_a = lt.a;
        B.tile_c = lt.c;
        P.gather_tile = gtile;
        P.gather_waves = gwaves;
        if (ctx->autotune && ctx->force.gather_tile == 0) {
            // one eager backward writes the line records; then the gather
            // alone with each candidate tile (the most waves that fit, and
            // half as many), the mean of two runs after a first
            TT_HIP(enqueue_backward(ctx, P, own, false));
            hipEvent_t e0, e1;
            TT_HIP(hipEventCreate(&e0));
            TT_HIP(hipEventCreate(&e1));
            float best = 1e30f;
            for (int T : {32, 16}) {
                const int wmax = gather_waves_max(T, nmax, do345, do12);
                for (int w : {wmax, wmax / 2}) {
                    if (w < 1 || (ctx->force.gather_waves > 0 && w != ctx->force.gather_waves))
                        continue;
                    P.ap.tiles = layout_gather_tiles(d, T);
                    TT_HIP(hipMemcpy(B.desc.ptr, d.data(), sizeof(ttk::BatchImage) * nimg,
                                     hipMemcpyHostToDevice));
                    float tsum = 0.0f;
                    for (int rep = 0; rep < 3; rep++) {
                        TT_HIP(hipEventRecord(e0, own));
                        TT_HIP(ttk::launch_gather_batched(P.ap, T, w, own));
                        TT_HIP(hipEventRecord(e1, own));
                        TT_HIP(hipEventSynchronize(e1));
                        float ms = 0.0f;
                        TT_HIP(hipEventElapsedTime(&ms, e0, e1));
                        if (rep > 0)
                            tsum += ms;
                    }
                    if (tsum < best) {
                        best = tsum;
                        P.gather_tile = T;
                        P.gather_waves = w;
                    }
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            tiles = layout_gather_tiles(d, P.gather_tile);
            P.ap.tiles = tiles;
            TT_HIP(hipMemcpy(B.desc.ptr, d.data(), sizeof(ttk::BatchImage) * nimg,
                             hipMemcpyHostToDevice));
        }
        TT_HIP(hipStreamBeginCapture(own, hipStreamCaptureModeThreadLocal));
        hipError_t ce = enqueue_backward(ctx, P, own, false);
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(own, &g);
        TT_HIP(ce);
        TT_HIP(ee);
        B.graph = g;
        TT_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
        B.key = key;
        B.valid = true;
    }
    if (!ctx->ev_in)
        TT_HIP(hipEventCreateWithFlags(&ctx->ev_in, hipEventDisableTiming));
    if (!ctx->ev_out)
        TT_HIP(hipEventCreateWithFlags(&ctx->ev_out, hipEventDisableTiming));
    if (ctx->stream != own) {
        TT_HIP(hipEventRecord(ctx->ev_in, ctx->stream));
        TT_HIP(hipStreamWaitEvent(own, ctx->ev_in, 0));
    }
    if (ctx->timer.on)
        TT_HIP(enqueue_backward(ctx, P, own, true));  // per-stage kernel times
    else
        TT_HIP(hipGraphLaunch(B.exec, own));
    if (ctx->stream != own) {
        TT_HIP(hipEventRecord(ctx->ev_out, own));
        TT_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_out, 0));
    }
    return TT_OK;
}

TT_API tt_status tt_profile_enable(tt_context *ctx, int enable)
{
    TT_CHECK(ctx, "ctx is NULL");
    TT_TRY(ctx->set_device());
    if (!enable && ctx->timer.on)
        TT_HIP(ctx->timer.collect());
    ctx->timer.on = enable != 0;
    return TT_OK;
}

TT_API tt_status tt_profile_read(tt_context *ctx, int kind, double *total_ms,
                                 int64_t *launches)
{
    TT_CHECK(ctx, "ctx is NULL");
    TT_CHECK(kind >= 0 && kind < TT_KERNEL_KINDS, "invalid kernel kind %d", kind);
    TT_CHECK(total_ms && launches, "output pointer is NULL");
    TT_TRY(ctx->set_device());
    TT_HIP(ctx->timer.collect());
    *total_ms = ctx->timer.ms[kind];
    *launches = ctx->timer.count[kind];
    ctx->timer.ms[kind] = 0.0;
    ctx->timer.count[kind] = 0;
    return TT_OK;
}

}  // extern "C"

TT_API tt_status tt_circus_noalloc(tt_context *ctx, const float *d_sino, int n,
                                   int n_cols, int n_sino, const tt_pfunc *pf, int np,
                                   float *d_circus, int32_t *d_p2_rank)
{
    TT_CHECK(ctx, "ctx is NULL");
    TT_CHECK(n >= 2 && n <= 2730, "sinogram height %d outside 2..2730", n);
    TT_CHECK(n_cols >= 1 && n_sino >= 1, "n_cols and n_sino must be >= 1");
    TT_CHECK(np >= 0 && np <= ttk::kMaxP && (np == 0 || pf), "np must be in 0..%d",
             ttk::kMaxP);
    for (int q = 0; q < np; q++) {
        if (pf[q].code == TT_P_HERMITE)
            return fail(TT_ERR_UNSUPPORTED,
                        "tt_circus_noalloc computes P1..P3 (Hermite needs its weight "
                        "tables allocated: use tt_circus)");
        TT_CHECK(pf[q].code >= 1 && pf[q].code <= 3,
                 "invalid P-functional code %d at position %d", pf[q].code, q);
    }
    if (np == 0)
        return TT_OK;
    TT_CHECK(d_sino && d_circus, "NULL sinogram or circus pointer");
    const int h = n / 2 + 1;
    const size_t cols = (size_t)n_sino * n_cols;
    bool p3 = false;
    for (int q = 0; q < np; q++)
        p3 = p3 || pf[q].code == TT_P3;
    TT_CHECK(!p3 || cols * 2 * h <= kNoAllocSpecFloats,
             "P3 spectra of %zu columns of height %d need %zu floats, above the %zu of "
             "the allocation-free scratch",
             cols, n, cols * 2 * h, kNoAllocSpecFloats);
    TT_TRY(ctx->set_device());
    float *spec = nullptr;
    if (p3)
        TT_TRY(noalloc_scratch(ctx->device, &spec));

    ttk::CircusParams p = {};
    p.sino = d_sino;
    p.out = d_circus;
    p.p2rank = d_p2_rank;
    p.n = n;
    p.ncols = n_cols;
    p.nsino = n_sino;
    p.np = np;
    for (int q = 0; q < np; q++) {
        p.code[q] = pf[q].code;
        p.herm[q] = -1;
    }
    hipStream_t st = ctx->stream;
    if (p3) {
        // every spectrum entry is written before it is read: no clearing
        p.p3_spec = spec;
        p.p3_h = h;
        ttk::DftParams dp = {};
        dp.njobs = 1;
        dp.job.in = d_sino;
        dp.job.out = spec;
        dp.job.n = n;
        dp.job.cols = (int)cols;
        dp.job.ld_in = n;
        dp.job.ld_out = 2 * h;
        ttk::dft_tile_shape(ttk::dft_job_tiles(n, (int)cols, true, 128, 128),
                            ttk::dft_job_tiles(n, (int)cols, true, 64, 128), &dp.bm, &dp.bn);
        dp.tiles = ttk::dft_job_tiles(n, (int)cols, true, dp.bm, dp.bn);
        TT_HIP(ttk::launch_dft_gemm(dp, true, n, st));
    }
    TT_HIP(ttk::launch_circus(p, st));
    return TT_OK;
}

// ---- test hooks (include/tt_testing.h) --------------------------------------

TT_API tt_status tt_testing_fail_allocations(long long allowed)
{
    g_alloc_budget.store(allowed < 0 ? -1 : allowed);
    return TT_OK;
}

TT_API tt_status tt_testing_nos_product(tt_context *ctx, int trans_a, int sym, int count,
                                        int m, int n, int k, const double *d_a,
                                        const double *d_b, double *d_c, const double *coef,
                                        double *d_t, double *d_q)
{
    TT_CHECK(ctx && d_a && d_b && d_c, "NULL argument");
    TT_CHECK(count >= 1 && m >= 1 && n >= 1 && k >= 1, "invalid sizes %d x %d by %d (x %d)",
             m, n, k, count);
    TT_CHECK(!sym || m == n, "a symmetric product is square (got %d x %d)", m, n);
    TT_TRY(ctx->set_device());
    hipStream_t st = ctx->stream;
    if (coef) {
        TT_CHECK(sym && !trans_a && m == k && d_a == d_b && d_t && d_q && k % 4 == 0,
                 "the polynomial step squares one symmetric k x k matrix, k a multiple of 4");
        ttk::NosPolyOut po{d_t, d_q, (long long)k * k, nullptr, nullptr, {}};
        for (int j = 0; j < 5; j++)
            po.c[j] = coef[j];
        TT_HIP(ttk::launch_nos_square_poly(d_a, d_c, k, count, po, st));
    } else {
        const ttk::NosGemmBatch bt{nullptr, d_a, d_b, d_c, (long long)m * k, (long long)k * n,
                                   (long long)m * n};
        TT_HIP(ttk::launch_nos_gemm(bt, count, trans_a != 0, sym != 0, m, n, k,
                                    trans_a ? k : m, k, m, 1.0, 0.0, st));
    }
    TT_HIP(hipStreamSynchronize(st));
    return TT_OK;
}

TT_API tt_status tt_testing_short_sqrt(tt_context *ctx, uint32_t lo, uint64_t count,
                                       uint64_t *mismatches, uint32_t *first_bad)
{
    TT_CHECK(ctx && mismatches && first_bad, "NULL argument");
    TT_CHECK(count <= ((uint64_t)1 << 32) - lo, "range past the last bit pattern");
    TT_TRY(ctx->set_device());
    unsigned long long *d_m = nullptr;
    unsigned int *d_f = nullptr;
    TT_HIP(hipMalloc(&d_m, sizeof(unsigned long long) + sizeof(unsigned int)));
    d_f = reinterpret_cast<unsigned int *>(d_m + 1);
    const unsigned long long zero = 0;
    const unsigned int none = 0xFFFFFFFFu;
    hipError_t e = hipMemcpy(d_m, &zero, sizeof zero, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d_f, &none, sizeof none, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = ttk::launch_short_sqrt_check(lo, count, d_m, d_f, ctx->stream);
    unsigned long long m = 0;
    unsigned int f = 0;
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpy(&m, d_m, sizeof m, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(&f, d_f, sizeof f, hipMemcpyDeviceToHost);
    (void)hipFree(d_m);
    TT_HIP(e);
    *mismatches = m;
    *first_bad = f;
    return TT_OK;
}
