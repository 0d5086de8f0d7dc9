// mh_internal.hpp -- shared host-side state of libmauvehip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdlib>
#include <cstddef>
#include <string>
#include <vector>
#include <set>
#include <map>
#include "mauvehip.h"

#include <functional>

// A match list owned by the caller: matches in GetMatchList order, and the
// sequence names and lengths a match-list file carries.
struct mh_match_list {
    std::vector<mh_match> v;
    int seq_count = 0;
    std::vector<std::string> files;
    std::vector<uint64_t> lengths;
};

namespace mh {

void set_error(const std::string &msg);

#define MH_HIP(call)                                                            \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::mh::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return MH_ERR_DEVICE;                                               \
        }                                                                       \
    } while (0)

#define MH_CHECK(cond, code, msg)                  \
    do {                                           \
        if (!(cond)) {                             \
            ::mh::set_error(msg);                  \
            return (code);                         \
        }                                          \
    } while (0)

// A growable device buffer owned by a context.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    int ensure(size_t need);   // MH_OK or MH_ERR_NOMEM
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Device buffers of one sorted mer list, kept by the context for reuse.
struct SmlBufs {
    uint32_t *words = nullptr;
    uint64_t *keys_pos = nullptr, *keys = nullptr;
    uint32_t *pos = nullptr;
    uint64_t cap_words = 0, cap_n = 0;
    void release();
};

// Host memory kept by the context between calls (uninitialised; grows only),
// so large per-call host arrays do not pay for zero-filling and first-touch
// page faults every time.
struct HostBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf() { std::free(ptr); }
    template <class T> T *get(size_t count)
    {
        const size_t need = count * sizeof(T) + 64;
        if (need > bytes) {
            std::free(ptr);
            ptr = std::malloc(need + need / 4);
            bytes = ptr ? need + need / 4 : 0;
        }
        return static_cast<T *>(ptr);
    }
};

struct KernelTimer {
    double total_ms = 0;
    int64_t launches = 0;
};

}  // namespace mh

#define MH_DP_MAX_VARIANTS 8   // forward kernel variants of the DP (dp.hip)

struct mh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool timing = false;
    std::map<std::string, mh::KernelTimer> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    // DP scratch
    mh::DevBuf dp_tb, dp_bnd, dp_jobs, dp_a, dp_b, dp_paths, dp_res, dp_dev;
    mh::DevBuf dp_wild, dp_fl;   // operand scan; the float DP's job table and rows
    // DP rounds: the descriptor and queue cursors (fixed address), its pinned
    // host copy, the launches captured as graphs per forward variant, and the
    // persistent grids
    mh::DevBuf dp_round;
    void *dp_round_host = nullptr;
    hipGraph_t dp_graph_fwd_src[MH_DP_MAX_VARIANTS] = {}, dp_graph_tb_src = nullptr;
    hipGraphExec_t dp_graph_fwd[MH_DP_MAX_VARIANTS] = {}, dp_graph_tb = nullptr;
    hipGraphNode_t dp_graph_fwd_node[MH_DP_MAX_VARIANTS] = {}, dp_graph_tb_node[2] = {};
    unsigned dp_fwd_grid[MH_DP_MAX_VARIANTS] = {};
    mh::DevBuf ml_jobs, ml_y;   // learned band: job table, network outputs + bands
    // seed-matching scratch
    mh::DevBuf mf_scratch[16];
    mh::DevBuf grid_scratch[11];   // spaced-seed extension by chains
    mh::DevBuf diag_scratch[8];    // long extensions resolved per diagonal
    mh::DevBuf multi_scratch[11];  // the seeds of every genome pair, merged and filtered
    mh::DevBuf heavy_buf, live_buf; // masked mers above the repeat limit; live ranges of a join
    mh::DevBuf run_scratch[4];      // odd solid seeds extended by runs
    mh::DevBuf uq_scratch[8];       // UniqueMatchFinder: merged lists, runs, seeds
    mh::DevBuf uq_rc[16];           // UniqueMatchFinder: packed reverse complements
    uint64_t last_walks = 0, last_long_walks = 0;   // of the last seed extension
    // contexts (own streams and scratch) that run independent searches
    // concurrently for this one, created on first use
    std::vector<mh_context *> helpers;
    // scratch for sorting; the radix sort's look-back status words and the
    // epoch of its last pass
    mh::DevBuf sort_tmp[4];
    mh::DevBuf sort_status;
    uint32_t sort_epoch = 0;
    // pinned host staging for device-to-host transfers; a second one for the
    // DP job tables
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    void *pinned_jobs = nullptr;
    size_t pinned_jobs_bytes = 0;
    mh::DevBuf dp_packed;   // trace-back paths at two bits an edge
    // sorted mer lists: live ones (orphaned when the context goes first) and
    // released buffer sets kept for the next build
    std::set<mh_sml *> smls;
    std::vector<mh::SmlBufs> sml_pool;
    // refinement passes: prefix counts and the rows each pass writes
    mh::HostBuf refine_prefix[2], refine_rows[2][2];
};

struct mh_sml {
    mh_context *ctx = nullptr;
    uint64_t len = 0;          // sequence length
    uint64_t n = 0;            // SMLLength
    uint64_t seed = 0;
    int seed_length = 0, seed_weight = 0;
    uint64_t n_words = 0;
    uint32_t *d_words = nullptr;     // packed sequence
    uint64_t *d_keys_pos = nullptr;  // GetDnaSeedMer per position (position order)
    uint64_t *d_keys = nullptr;      // sorted keys
    uint32_t *d_pos = nullptr;       // sorted positions
    uint64_t cap_words = 0, cap_n = 0;   // capacities of the buffers above
};

namespace mh {

// Kernel timing: record a start event before and a stop event after a launch
// on the context stream; totals are folded in by collect_timers().
struct TimedLaunch {
    mh_context *ctx;
    std::string name;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(mh_context *c, const char *n);
    ~TimedLaunch();
};
int collect_timers(mh_context *ctx);
// Host-side stage timing folded into the same table (wall clock, ms).
void host_time(mh_context *ctx, const char *name, double ms);
double now_ms();
// Runs f(0) .. f(n-1) on the library's host thread pool (the caller takes
// part); tasks are handed out in order as threads free up.  A call made while
// the pool is busy (from a task, or another thread) runs its tasks itself.
void host_parallel(size_t n, const std::function<void(size_t)> &f);
unsigned host_threads();
// Pinned host staging buffer of at least `bytes` (owned by the context).
int pinned(mh_context *ctx, size_t bytes, void **out);
int pinned_jobs(mh_context *ctx, size_t bytes, void **out);   // a second pinned buffer, for DP job tables

// Launchers (return MH_OK or an error code); all run on ctx->stream.
// wild (nullable): per job, nonzero when its operands hold a class above 3
// (the caller knows); otherwise the operands are scanned on the device.
int dp_run(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
           uint64_t njobs, uint8_t *d_paths, mh_dp_result *d_results, const uint8_t *wild = nullptr);
// DP operand classes: 0-3 A,C,G,T, 4 'X', 5 any other letter (libMUSCLE's
// wildcard counts); jobs with a class above 3 run the float DP (dp_float.hip).
constexpr int DP_NCLASS = 6;
// A fixed device scratch (the no-allocation entry point): job tables and
// arenas are laid out in it and nothing is allocated.
struct FixedScratch {
    uint8_t *ptr = nullptr;
    uint64_t bytes = 0;
};
// out[j] = the largest operand code of job j (device scan, synchronises).
int dp_code_max(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs, uint64_t njobs,
                const FixedScratch *fx, std::vector<uint8_t> &out);
// NWSmall in float for the jobs `which` (indices into jobs; slot = path slot
// starts); cells (nullable) accumulates la * lb.
int dp_run_float(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
                 const std::vector<uint64_t> &slot, const std::vector<uint32_t> &which, uint8_t *d_paths,
                 mh_dp_result *d_results, uint32_t *d_dev, uint64_t *cells, const FixedScratch *fx);
// the 6 x 6 float match scores of the classes (row = A's class)
void dp_score6(float *out);
// The no-allocation DP: the same schedule and kernels with every scratch
// buffer in the library's fixed per-device array (dp_scratch_bytes() bytes,
// address resolved by dp_scratch_init for the current device).
int dp_scratch_init(int device);
uint64_t dp_scratch_bytes();
int dp_run_noalloc(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs, uint64_t njobs,
                   uint8_t *d_paths, mh_dp_result *d_results);
// paths[0, n) ('M', 'D', 'I') packed four to a byte into out: edge x at bits
// 2(x mod 4) of byte x / 4, codes M 0, D 1, I 2 (3 for anything else)
int dp_pack_paths(mh_context *ctx, const uint8_t *d_paths, uint64_t n, uint8_t *d_out);
// Banded DP with edge fallback (bands null: all full); d_dev may be null.
int dp_run_banded(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs, uint64_t njobs,
                  const uint32_t *bands, uint32_t flags, uint8_t *d_paths, mh_dp_result *d_results,
                  uint32_t *d_dev, mh_band_stats *stats);
// Stripe shape (columns per lane) and effective band the scheduler uses.
int dp_force_cols();   // MH_DP_COLS = 4 | 8 | 16 | 36 | 40 forces one stripe shape (tuning sweeps)
uint32_t dp_cols_for(uint32_t la, uint32_t lb, uint32_t band, int force, bool wide_only = false);
uint32_t dp_stripe_width(uint32_t cols);   // columns of one stripe of a dp_cols_for shape
uint32_t dp_effective_band(uint32_t la, uint32_t lb, uint32_t band, uint32_t cols);
int dp_force_kind();   // MH_DP_BAND_KIND = stripes | window forces one banded region (tests, sweeps)
uint32_t dp_window_cols(uint32_t la, uint32_t lb, uint32_t band, int kind);
uint64_t dp_item_cells(uint32_t la, uint32_t lb, uint32_t band);   // cells of the scheduled region

int sml_build(mh_context *ctx, const void *d_ascii, uint64_t len, uint64_t seed, mh_sml *s);
// Device buffers for s (s->n_words and s->n set), from the context's pool or fresh.
int sml_alloc(mh_context *ctx, mh_sml *s);
// For a list whose packed words and sorted positions are on the device (a
// list read from a file, positions already checked to be a permutation of
// [0, n)): computes the per-position keys and the sorted keys; *unsorted is
// set when the keys at the given positions are not in ascending order.
int sml_rebuild_keys(mh_context *ctx, mh_sml *s, int *unsorted);
int sml_seed_mers(mh_context *ctx, const mh_sml *s, const uint32_t *d_pos, uint64_t n, uint64_t *d_out);

int find_matches_pairwise(mh_context *ctx, const mh_sml *const *smls, int k,
                          std::vector<mh_match> &out, mh_match_stats *st);

// One seed in HashMatch order (after SetDirection) with its ExtendMatch result.
struct SeedRec {
    int32_t s0, s1;        // seed starts, 1-based, genome 1 negative on the other strand
    int32_t e0, e1;        // extended starts
    uint32_t elen;         // extended length
};
// Seeds stay in device memory owned by the context (valid until the next
// match search); *d_seeds points there.
// live (nullable) gives the index ranges of each list the search reads
// (search_live_ranges); ga, gb are a's and b's list numbers in it.
struct LiveRanges;
int join_and_extend(mh_context *ctx, const mh_sml *a, const mh_sml *b, const LiveRanges *live, int ga, int gb,
                    const SeedRec **d_seeds, uint64_t *ns);
// The merge key of the seeds of one genome pair: the masked mer at each
// seed's genome-a start, shifted up three bits, or'ed with the pair number.
int seed_mers(mh_context *ctx, const mh_sml *a, const SeedRec *d_seeds, uint64_t ns, uint32_t pair,
              uint64_t *d_out);
// Orders n seeds (concatenated over genome pairs) by key = mer << 3 | pair:
// d_seeds_out[i] = d_seeds[sorted index i], d_keys_out[i] = its key.
int merge_pair_seeds(mh_context *ctx, const SeedRec *d_seeds, uint64_t *d_keys, uint64_t n, SeedRec **d_seeds_out,
                     uint64_t **d_keys_out);
// The merged seeds (d_seeds, keys as merge_pair_seeds leaves them) whose
// MemHash bucket is marked in mask[40000], in order, with their pair numbers.
int select_bucket_seeds(mh_context *ctx, const SeedRec *d_seeds, const uint64_t *d_keys, uint64_t n,
                        const uint8_t *mask, int seed_len, std::vector<SeedRec> &seeds, std::vector<uint8_t> &pairs);
// A distinct extension of a clean bucket, and a seed of any other bucket.
struct CleanEnt {
    uint32_t bucket;
    int32_t e0, e1;
    uint32_t elen;
};
struct DirtySeed {
    uint32_t bucket;
    SeedRec r;
};
// Splits the seeds by MemHash bucket: clean buckets resolved on the device
// (distinct extensions, bucket then start order), the rest returned in bucket
// then HashMatch order for the literal replay.
int partition_seeds(mh_context *ctx, const SeedRec *d_seeds, uint64_t ns, int k, bool solid_odd, uint64_t len_a,
                    std::vector<CleanEnt> &clean, std::vector<DirtySeed> &dirty, uint64_t *n_clean_seeds);
int sml_find_mer_dev(mh_context *ctx, const mh_sml *s, uint64_t mer, uint64_t *index);
uint64_t sml_full_key(uint64_t c, int sw);
int ascii_to_codes(mh_context *ctx, const void *d_ascii, uint64_t len, uint8_t *d_codes);
uint64_t sml_compact_key(uint64_t k, int sw);

// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits, in
// the fewest passes of digits up to 10 bits, ping-ponging between (k0,v0) and (k1,v1); *out_k/*out_v get the
// pair of buffers holding the result.  Runs on ctx->stream (onesweep: one
// histogram kernel, then one kernel per pass with decoupled look-back).
int radix_sort_kv(mh_context *ctx, uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, uint64_t n, int bits,
                  uint64_t **out_k, uint32_t **out_v);
// Synchronises ctx->stream and reports MH_ERR_DEVICE if a radix sort's
// look-back wait ran past its bound since the last check.
int sort_check(mh_context *ctx);

// Sorted lists as the restart planner reads them: compact keys
// ((chars << 1) | strand, ordered as the 64-bit keys), host memory.
constexpr int PLAN_MAXK = 16;
struct PlanLists {
    int k = 0, sw = 0, sl = 0;              // list count, seed weight, seed length
    const uint64_t *ck[PLAN_MAXK] = {};     // sorted compact keys
    uint64_t n[PLAN_MAXK] = {};             // SMLLength
    uint64_t len[PLAN_MAXK] = {};           // sequence length (FindMer's bound)
};
// The index ranges of each list that SearchRange reads, restart by restart
// (ascending, disjoint), and the number of restarts.
struct LiveRanges {
    std::vector<uint64_t> lo[PLAN_MAXK], hi[PLAN_MAXK];
    std::vector<uint32_t> ep[PLAN_MAXK];   // the search (restarts before it) that reads each range
    int64_t restarts = 0;
};
// Plans SearchRange's repeat-limit restarts (LM/MatchFinder.cpp:166-340) given
// every masked mer with more than MER_REPEAT_LIMIT entries over all lists
// (ascending, distinct).  restart.cpp.
int plan_restarts(const PlanLists &L, const std::vector<uint64_t> &heavy, LiveRanges *out);
std::vector<uint64_t> heavy_mers_host(const PlanLists &L);
// The live ranges of k device lists: heavy mers found on the device; when there
// are any, the keys come to the host and the restarts are planned.  A list
// with no restart reads [0, n).
int search_live_ranges(mh_context *ctx, const mh_sml *const *smls, int k, LiveRanges *out);

// UniqueMatchFinder over k genomes (2..16) on the device: its seeds (n, in
// HashMatch order), the collisions AddHashEntry counts, and the match list in
// GetMatchList order (starts k a match, 1-based, negative = reverse, 0 =
// absent; lengths).  match.hip.
struct UniqueSeeds {
    int k = 0;
    uint64_t n = 0;
    uint64_t walks = 0;        // ExtendMatch walks run (spaced seeds: one a chain)
    int64_t collisions = 0;
    std::vector<int64_t> starts;
    std::vector<uint64_t> lengths;
};
int unique_seeds(mh_context *ctx, const mh_sml *const *smls, int k, const LiveRanges &live, UniqueSeeds *out);

// host-side seed helpers (LM/SeedMasks.h)
int64_t get_seed(int weight, int rank);
int seed_length(int64_t seed);
int seed_weight(int64_t seed);

}  // namespace mh
