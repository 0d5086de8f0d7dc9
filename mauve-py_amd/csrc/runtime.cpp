// runtime.cpp -- context, errors, device buffers, kernel timers and the
// C-ABI entry points that are not tied to one kernel family.
#include "mh_internal.hpp"
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <cstring>
#include <cmath>
#include <chrono>

namespace mh {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int DevBuf::ensure(size_t need)
{
    if (need <= bytes && ptr) return MH_OK;
    release();
    size_t want = need + need / 8;
    if (hipMalloc(&ptr, want) != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        if (hipMalloc(&ptr, need) != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
            bytes = 0;
            set_error("device allocation of " + std::to_string(need) + " bytes failed");
            return MH_ERR_NOMEM;
        }
        want = need;
    }
    bytes = want;
    return MH_OK;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void SmlBufs::release()
{
    if (words) (void)hipFree(words);
    if (keys_pos) (void)hipFree(keys_pos);
    if (keys) (void)hipFree(keys);
    if (pos) (void)hipFree(pos);
    *this = SmlBufs();
}

static hipEvent_t take_event(mh_context *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedLaunch::TimedLaunch(mh_context *c, const char *n) : ctx(c), name(n)
{
    if (!ctx->timing) return;
    a = take_event(ctx);
    b = take_event(ctx);
    if (a) (void)hipEventRecord(a, ctx->stream);
}

TimedLaunch::~TimedLaunch()
{
    if (!ctx->timing || !a || !b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({name, {a, b}});
}

int collect_timers(mh_context *ctx)
{
    for (auto &p : ctx->pending) {
        MH_HIP(hipEventSynchronize(p.second.second));
        float ms = 0;
        MH_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        auto &t = ctx->timers[p.first];
        t.total_ms += ms;
        t.launches += 1;
        ctx->event_pool.push_back(p.second.first);
        ctx->event_pool.push_back(p.second.second);
    }
    ctx->pending.clear();
    return MH_OK;
}

void host_time(mh_context *ctx, const char *name, double ms)
{
    if (!ctx->timing) return;
    auto &t = ctx->timers[name];
    t.total_ms += ms;
    t.launches += 1;
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int pinned(mh_context *ctx, size_t bytes, void **out)
{
    if (bytes > ctx->pinned_bytes) {
        if (ctx->pinned) (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        const size_t want = bytes + bytes / 4 + 4096;
        if (hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            ctx->pinned = nullptr;
            set_error("pinned host allocation failed");
            return MH_ERR_NOMEM;
        }
        ctx->pinned_bytes = want;
    }
    *out = ctx->pinned;
    return MH_OK;
}

int pinned_jobs(mh_context *ctx, size_t bytes, void **out)
{
    if (bytes > ctx->pinned_jobs_bytes) {
        if (ctx->pinned_jobs) (void)hipHostFree(ctx->pinned_jobs);
        ctx->pinned_jobs = nullptr;
        ctx->pinned_jobs_bytes = 0;
        const size_t want = bytes + bytes / 4 + 4096;
        if (hipHostMalloc(&ctx->pinned_jobs, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            ctx->pinned_jobs = nullptr;
            set_error("pinned host allocation failed");
            return MH_ERR_NOMEM;
        }
        ctx->pinned_jobs_bytes = want;
    }
    *out = ctx->pinned_jobs;
    return MH_OK;
}

// ---- seed patterns: LM/SeedMasks.h:44-401 ------------------------------
// Rank-ordered masks per weight (the reference stores each as a pair of 32-bit
// words whose high word is 0).
static const uint32_t kSeedTable[32][6] = {
    {0}, {0}, {0},
    {0xb, 0, 0, 0, 0, 0},
    {0x3b, 0, 0, 0, 0, 0},
    {0x6b, 0x139, 0x193, 0x6b, 0, 0},
    {0x58D, 0x653, 0x1AB, 0xdb, 0, 0},
    {0x1953, 0x588d, 0x688b, 0x17d, 0x164d, 0},
    {0x3927, 0x1CA7, 0x6553, 0xb6d, 0, 0},
    {0x7497, 0x1c927, 0x72a7, 0x6fb, 0x16ed, 0},
    {0x1d297, 0x3A497, 0xE997, 0x6D5B, 0, 0},
    {0x7954f, 0x75257, 0x1c9527, 0x5bed, 0x5b26d, 0},
    {0x7954f, 0x3D32F, 0x768B7, 0x5B56D, 0, 0},
    {0x792a4f, 0x1d64d7, 0x1d3597, 0x1b7db, 0x75ad7, 0},
    {0x1e6acf, 0xF59AF, 0x3D4CAF, 0x35AD6B, 0, 0},
    {0x7ac9af, 0x7b2a6f, 0x79aacf, 0x16df6d, 0x6b5d6b, 0},
    {0xf599af, 0xEE5A77, 0x7CD59F, 0xEB5AD7, 0, 0},
    {0x6dbedb, 0, 0, 0, 0, 0},
    {0x3E6B59F, 0x3EB335F, 0x7B3566F, 0, 0, 0},
    {0x7b974ef, 0x7d6735f, 0x1edd74f, 0, 0, 0},
    {0x1F59B35F, 0x3EDCEDF, 0xFAE675F, 0, 0, 0},
    {0x7ddaddf, 0xaeb3f, 0x7eb76bf, 0, 0, 0},
    {0x003fffff, 0, 0, 0, 0, 0},
    {0x007fffff, 0, 0, 0, 0, 0},
    {0x00ffffff, 0, 0, 0, 0, 0},
    {0x01ffffff, 0, 0, 0, 0, 0},
    {0x03ffffff, 0, 0, 0, 0, 0},
    {0x07ffffff, 0, 0, 0, 0, 0},
    {0x0fffffff, 0, 0, 0, 0, 0},
    {0x1fffffff, 0, 0, 0, 0, 0},
    {0x3fffffff, 0, 0, 0, 0, 0},
    {0x7fffffff, 0, 0, 0, 0, 0},
};

static int64_t solid(int w) { return (int64_t)(((uint64_t)1 << w) - 1); }

int64_t get_seed(int weight, int rank)
{
    if (weight > 31) return solid(32);
    if (weight < 0) return 0;
    if (rank > 5) return solid(weight);
    if (rank < 0 || kSeedTable[weight][rank] == 0) return solid(weight);
    return (int64_t)kSeedTable[weight][rank];
}

int seed_length(int64_t seed)
{
    uint64_t s = (uint64_t)seed;
    if (!s) return 0;
    return 64 - __builtin_clzll(s) - __builtin_ctzll(s);
}

int seed_weight(int64_t seed) { return __builtin_popcountll((uint64_t)seed); }

namespace {

// Worker threads wait for a job generation; each takes task indices from a
// shared counter.  The caller clears the job only once no worker is inside it,
// and a worker that wakes after that sees no job.
class ThreadPool {
public:
    explicit ThreadPool(unsigned nt)
    {
        for (unsigned t = 1; t < nt; ++t) workers_.emplace_back([this]() { loop(); });
    }
    ~ThreadPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &w : workers_) w.join();
    }
    unsigned size() const { return (unsigned)workers_.size() + 1; }
    void run(size_t n, const std::function<void(size_t)> &f)
    {
        // a single task, or a call made from inside a task (a worker, or the
        // caller while it takes part), runs here; so does a call made while
        // another thread's job holds the pool
        if (n < 2 || workers_.empty() || inside_) {
            for (size_t i = 0; i < n; ++i) f(i);
            return;
        }
        std::unique_lock<std::mutex> busy(call_mu_, std::try_to_lock);
        if (!busy.owns_lock()) {
            for (size_t i = 0; i < n; ++i) f(i);
            return;
        }
        struct Inside {
            Inside() { inside_ = true; }
            ~Inside() { inside_ = false; }
        } mark;
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            n_ = n;
            next_.store(0);
            ++gen_;
        }
        cv_.notify_all();
        for (size_t i; (i = next_.fetch_add(1)) < n;) f(i);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&]() { return active_ == 0; });
        job_ = nullptr;
    }

private:
    void loop()
    {
        inside_ = true;
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&]() { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            const std::function<void(size_t)> *f = job_;
            if (!f) continue;
            const size_t n = n_;
            ++active_;
            lk.unlock();
            for (size_t i; (i = next_.fetch_add(1)) < n;) (*f)(i);
            lk.lock();
            if (--active_ == 0) done_.notify_all();
        }
    }
    static thread_local bool inside_;
    std::vector<std::thread> workers_;
    std::mutex mu_, call_mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t)> *job_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0};
    int active_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

thread_local bool ThreadPool::inside_ = false;

ThreadPool &pool()
{
    static ThreadPool p(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    return p;
}

}  // namespace

void host_parallel(size_t n, const std::function<void(size_t)> &f) { pool().run(n, f); }

unsigned host_threads() { return pool().size(); }

}  // namespace mh

using namespace mh;

extern "C" {

const char *mh_last_error(void) { return g_err.c_str(); }

int mh_create(int device, mh_context **out)
{
    MH_CHECK(out != nullptr, MH_ERR_INVALID, "null output pointer");
    int n = 0;
    MH_HIP(hipGetDeviceCount(&n));
    MH_CHECK(device >= 0 && device < n, MH_ERR_INVALID, "no such HIP device");
    MH_HIP(hipSetDevice(device));
    {
        // the no-allocation DP's scratch address on this device (once per device)
        const int rc = mh::dp_scratch_init(device);
        if (rc != MH_OK) return rc;
    }
    mh_context *c = new mh_context();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed");
        return MH_ERR_DEVICE;
    }
    *out = c;
    return MH_OK;
}

void mh_destroy(mh_context *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (mh_context *h : ctx->helpers) mh_destroy(h);
    collect_timers(ctx);
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    for (hipGraphExec_t &g : ctx->dp_graph_fwd)
        if (g) (void)hipGraphExecDestroy(g);
    for (hipGraph_t &g : ctx->dp_graph_fwd_src)
        if (g) (void)hipGraphDestroy(g);
    if (ctx->dp_graph_tb) (void)hipGraphExecDestroy(ctx->dp_graph_tb);
    if (ctx->dp_graph_tb_src) (void)hipGraphDestroy(ctx->dp_graph_tb_src);
    if (ctx->dp_round_host) (void)hipHostFree(ctx->dp_round_host);
    ctx->dp_round.release();
    ctx->dp_tb.release();
    ctx->dp_bnd.release();
    ctx->dp_jobs.release();
    ctx->dp_a.release();
    ctx->dp_b.release();
    ctx->dp_paths.release();
    ctx->dp_res.release();
    for (auto &b : ctx->mf_scratch) b.release();
    for (auto &b : ctx->sort_tmp) b.release();
    for (mh_sml *s : ctx->smls) s->ctx = nullptr;   // lists outliving the context free their own buffers
    for (auto &b : ctx->sml_pool) b.release();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->pinned_jobs) (void)hipHostFree(ctx->pinned_jobs);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int mh_synchronize(mh_context *ctx)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int mh_set_timing(mh_context *ctx, int enabled)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    ctx->timing = enabled != 0;
    return MH_OK;
}

int mh_kernel_time(mh_context *ctx, const char *kernel, double *avg_ms, int64_t *launches)
{
    MH_CHECK(ctx && kernel, MH_ERR_INVALID, "null argument");
    int rc = collect_timers(ctx);
    if (rc != MH_OK) return rc;
    auto it = ctx->timers.find(kernel);
    if (it == ctx->timers.end()) {
        if (avg_ms) *avg_ms = 0;
        if (launches) *launches = 0;
        return MH_OK;
    }
    if (avg_ms) *avg_ms = it->second.launches ? it->second.total_ms / it->second.launches : 0;
    if (launches) *launches = it->second.launches;
    return MH_OK;
}

void mh_kernel_time_reset(mh_context *ctx)
{
    if (!ctx) return;
    collect_timers(ctx);
    ctx->timers.clear();
}

int64_t mh_get_seed(int weight, int seed_rank) { return get_seed(weight, seed_rank); }
int mh_get_seed_length(int64_t seed) { return seed_length(seed); }
int mh_get_seed_weight(int64_t seed) { return seed_weight(seed); }

uint32_t mh_default_seed_weight(uint64_t avg)
{
    uint32_t m = (uint32_t)std::ceil((std::log((double)avg) / std::log(2.0)) / 1.5);
    if (!(m & 1)) ++m;
    m = m < 5 ? 0 : m;
    if (avg == 0) m = 0;
    return m > 31 ? 31 : m;
}

int mh_device_alloc(mh_context *ctx, uint64_t bytes, void **d_ptr)
{
    MH_CHECK(ctx && d_ptr, MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) {
        (void)hipGetLastError();
        set_error("device allocation failed");
        return MH_ERR_NOMEM;
    }
    return MH_OK;
}

int mh_device_free(mh_context *ctx, void *d_ptr)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    if (d_ptr) MH_HIP(hipFree(d_ptr));
    return MH_OK;
}

int mh_memcpy_h2d(mh_context *ctx, void *d_dst, const void *h_src, uint64_t bytes)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    if (!bytes) return MH_OK;
    MH_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int mh_memcpy_d2h(mh_context *ctx, void *h_dst, const void *d_src, uint64_t bytes)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    if (!bytes) return MH_OK;
    MH_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int mh_ascii_to_codes_device(mh_context *ctx, const void *d_ascii, uint64_t len, uint8_t *d_codes)
{
    MH_CHECK(ctx && ((d_ascii && d_codes) || len == 0), MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    return ascii_to_codes(ctx, d_ascii, len, d_codes);
}

// ---- DP ------------------------------------------------------------------

int mh_dp_align_device(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
                       uint64_t njobs, uint8_t *d_paths, mh_dp_result *d_results)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    MH_CHECK(njobs == 0 || (jobs && d_a && d_b && d_paths && d_results), MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    return dp_run(ctx, d_a, d_b, jobs, njobs, d_paths, d_results);
}

int mh_dp_align_device_noalloc(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
                               uint64_t njobs, uint8_t *d_paths, mh_dp_result *d_results)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    MH_CHECK(njobs == 0 || (jobs && d_a && d_b && d_paths && d_results), MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    return dp_run_noalloc(ctx, d_a, d_b, jobs, njobs, d_paths, d_results);
}

uint64_t mh_dp_noalloc_scratch_bytes(void) { return mh::dp_scratch_bytes(); }

int mh_dp_align_banded_device(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
                              uint64_t njobs, const uint32_t *bands, uint32_t flags, uint8_t *d_paths,
                              mh_dp_result *d_results, uint32_t *d_dev, mh_band_stats *stats)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    MH_CHECK(njobs == 0 || (jobs && d_a && d_b && d_paths && d_results), MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    return dp_run_banded(ctx, d_a, d_b, jobs, njobs, bands, flags, d_paths, d_results, d_dev, stats);
}

uint32_t mh_dp_band_region_ex(uint32_t a_len, uint32_t b_len, uint32_t band, uint32_t flags, uint32_t *kind,
                              uint32_t *width)
{
    uint32_t k = MH_REGION_FULL, w = 0, eff = 0;
    if (a_len && b_len) {
        const int force = dp_force_cols();
        uint32_t c = dp_cols_for(a_len, b_len, band, force, (flags & MH_REGION_RERUN) != 0);
        eff = band ? dp_effective_band(a_len, b_len, band, c) : 0u;
        if (eff) {
            const uint32_t wc = dp_window_cols(a_len, b_len, eff, dp_force_kind());
            k = wc ? MH_REGION_WINDOW : MH_REGION_STRIPES;
            w = wc ? 64 * wc : dp_stripe_width(c);
        }
    }
    if (kind) *kind = k;
    if (width) *width = w;
    return eff;
}

uint32_t mh_dp_band_region(uint32_t a_len, uint32_t b_len, uint32_t band, uint32_t *kind, uint32_t *width)
{
    return mh_dp_band_region_ex(a_len, b_len, band, 0u, kind, width);
}

int mh_dp_band_cells(const uint32_t *a_len, const uint32_t *b_len, const uint32_t *band, uint64_t n, uint64_t *cells)
{
    MH_CHECK((a_len && b_len && cells) || n == 0, MH_ERR_INVALID, "null argument");
    auto fill = [&](uint64_t lo, uint64_t hi) {
        for (uint64_t x = lo; x < hi; ++x) cells[x] = dp_item_cells(a_len[x], b_len[x], band ? band[x] : 0u);
    };
    const uint64_t chunk = 65536;
    mh::host_parallel((size_t)((n + chunk - 1) / chunk),
                      [&](size_t k) { fill(k * chunk, std::min<uint64_t>(n, (k + 1) * chunk)); });
    return MH_OK;
}

int mh_dp_align(mh_context *ctx, const uint8_t *a_codes, uint64_t a_bytes, const uint8_t *b_codes,
                uint64_t b_bytes, const mh_dp_job *jobs, uint64_t njobs, uint8_t *paths,
                mh_dp_result *results)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    if (njobs == 0) return MH_OK;
    MH_CHECK(a_codes && b_codes && jobs && paths && results, MH_ERR_INVALID, "null argument");
    uint64_t total = 0;
    for (uint64_t j = 0; j < njobs; ++j) {
        MH_CHECK(jobs[j].a_off + jobs[j].a_len <= a_bytes && jobs[j].b_off + jobs[j].b_len <= b_bytes,
                 MH_ERR_INVALID, "DP job " + std::to_string(j) + " lies outside the code buffers");
        total += (uint64_t)jobs[j].a_len + jobs[j].b_len;
    }
    for (uint64_t x = 0; x < a_bytes; ++x)
        MH_CHECK(a_codes[x] < DP_NCLASS, MH_ERR_INVALID, "A code outside 0..5");
    for (uint64_t x = 0; x < b_bytes; ++x)
        MH_CHECK(b_codes[x] < DP_NCLASS, MH_ERR_INVALID, "B code outside 0..5");
    MH_HIP(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ctx->dp_a.ensure(a_bytes + 64)) != MH_OK) return rc;
    if ((rc = ctx->dp_b.ensure(b_bytes + 64)) != MH_OK) return rc;
    if ((rc = ctx->dp_paths.ensure(total + 64)) != MH_OK) return rc;
    if ((rc = ctx->dp_res.ensure(njobs * sizeof(mh_dp_result) + 64)) != MH_OK) return rc;
    MH_HIP(hipMemcpyAsync(ctx->dp_a.ptr, a_codes, a_bytes, hipMemcpyHostToDevice, ctx->stream));
    MH_HIP(hipMemcpyAsync(ctx->dp_b.ptr, b_codes, b_bytes, hipMemcpyHostToDevice, ctx->stream));
    // jobs with letters other than A,C,G,T (classes 4, 5) take the float DP
    std::vector<uint8_t> wild(njobs, 0);
    for (uint64_t j = 0; j < njobs; ++j) {
        uint8_t m = 0;
        for (uint32_t x = 0; x < jobs[j].a_len; ++x) m |= a_codes[jobs[j].a_off + x];
        for (uint32_t x = 0; x < jobs[j].b_len; ++x) m |= b_codes[jobs[j].b_off + x];
        wild[j] = m > 3;
    }
    rc = dp_run(ctx, ctx->dp_a.as<uint8_t>(), ctx->dp_b.as<uint8_t>(), jobs, njobs, ctx->dp_paths.as<uint8_t>(),
                ctx->dp_res.as<mh_dp_result>(), wild.data());
    if (rc != MH_OK) return rc;
    MH_HIP(hipMemcpyAsync(paths, ctx->dp_paths.ptr, total, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipMemcpyAsync(results, ctx->dp_res.ptr, njobs * sizeof(mh_dp_result), hipMemcpyDeviceToHost,
                          ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t j = 0; j < njobs; ++j)
        MH_CHECK(results[j].path_len > 0, MH_ERR_DEVICE, "trace-back failed for job " + std::to_string(j));
    return MH_OK;
}

}  // extern "C"
