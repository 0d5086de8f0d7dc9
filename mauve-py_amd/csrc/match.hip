// match.hip -- seed join and ungapped extension for two sorted mer lists.
//
// The reference walks the lists with a k-way merge and, for every run of
// equal seed (masked) keys, hashes each pair of genomes that occur exactly
// once in the run (LM/MatchFinder.cpp:172-340, LM/PairwiseMatchFinder.cpp:
// 37-71).  With two genomes that is: every masked key present exactly once
// in each list gives one seed, and the seeds come in ascending key order.
// Here one thread per genome-0 entry tests uniqueness against its
// neighbours, binary-searches the key in genome 1, and flags a seed; a
// tile scan compacts the seeds in key order.
//
// Each seed is then set up as HashMatch does (LM/MemHash.cpp:167-203):
// 1-based starts, genome 1 negated when its strand flag differs from genome
// 0's, and extended with the reference's four-phase ExtendMatch
// (LM/MatchFinder.h:218-374), a pure function of the seed and the two
// sequences.  Extension runs for every seed in parallel; which extended seeds
// enter the match table is decided by the host replay of AddHashEntry.
//
// Runs longer than MER_REPEAT_LIMIT (LM/MatchFinder.cpp:166) make the
// reference restart its merge from recomputed start points (253-276); such
// runs are counted so the host can refuse inputs that contain them.
#include "mh_internal.hpp"

namespace mh {

namespace {

constexpr int JT = 256;              // threads per join/compaction tile
constexpr int JITEMS = 16;
constexpr int JTILE = JT * JITEMS;
constexpr uint32_t MER_REPEAT_LIMIT = 1000;

__device__ __forceinline__ uint64_t lower_bound_m(const uint64_t *__restrict__ ck, uint64_t n, uint64_t m)
{
    uint64_t lo = 0, len = n;
    while (len > 0) {
        const uint64_t half = len >> 1, mid = lo + half;
        if ((ck[mid] >> 1) < m) { lo = mid + 1; len -= half + 1; }
        else len = half;
    }
    return lo;
}

__device__ __forceinline__ uint64_t upper_bound_m(const uint64_t *__restrict__ ck, uint64_t n, uint64_t m)
{
    uint64_t lo = 0, len = n;
    while (len > 0) {
        const uint64_t half = len >> 1, mid = lo + half;
        if ((ck[mid] >> 1) <= m) { lo = mid + 1; len -= half + 1; }
        else len = half;
    }
    return lo;
}

// flag[i] = 1 if genome-0 entry i forms a seed; partner[i] = its genome-1 index
__global__ void join_kernel(const uint64_t *__restrict__ ck0, uint64_t n0, const uint64_t *__restrict__ ck1,
                            uint64_t n1, uint32_t *__restrict__ flag, uint32_t *__restrict__ partner,
                            uint32_t *__restrict__ repeat_runs)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0) return;
    const uint64_t m = ck0[i] >> 1;
    const bool first = i == 0 || (ck0[i - 1] >> 1) != m;
    const bool last = i + 1 == n0 || (ck0[i + 1] >> 1) != m;
    uint32_t f = 0, pj = 0;
    if (first) {
        const uint64_t lo = lower_bound_m(ck1, n1, m);
        const bool hit = lo < n1 && (ck1[lo] >> 1) == m;
        if (last && hit && (lo + 1 == n1 || (ck1[lo + 1] >> 1) != m)) {
            f = 1;
            pj = (uint32_t)lo;
        }
        if (!last || hit) {
            const uint64_t r0 = last ? 1 : upper_bound_m(ck0, n0, m) - i;
            const uint64_t r1 = hit ? upper_bound_m(ck1, n1, m) - lo : 0;
            if (r0 + r1 > MER_REPEAT_LIMIT) atomicAdd(repeat_runs, 1u);
        }
    }
    flag[i] = f;
    partner[i] = pj;
}

// runs of genome 1 alone above the limit
__global__ void repeat1_kernel(const uint64_t *__restrict__ ck1, uint64_t n1, const uint64_t *__restrict__ ck0,
                               uint64_t n0, uint32_t *__restrict__ repeat_runs)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n1) return;
    const uint64_t m = ck1[i] >> 1;
    if (i > 0 && (ck1[i - 1] >> 1) == m) return;
    if (i + MER_REPEAT_LIMIT >= n1 || (ck1[i + MER_REPEAT_LIMIT] >> 1) != m) return;
    const uint64_t lo = lower_bound_m(ck0, n0, m);
    if (lo < n0 && (ck0[lo] >> 1) == m) return;   // counted from genome 0's side
    atomicAdd(repeat_runs, 1u);
}

__global__ __launch_bounds__(JT) void tile_count_kernel(const uint32_t *__restrict__ flag, uint64_t n,
                                                        uint32_t *__restrict__ tile_sum)
{
    __shared__ uint32_t red[JT];
    const int t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * JTILE;
    uint32_t s = 0;
    for (int r = 0; r < JITEMS; ++r) {
        const uint64_t x = base + (uint64_t)r * JT + t;
        if (x < n) s += flag[x];
    }
    red[t] = s;
    __syncthreads();
    for (int off = JT / 2; off > 0; off >>= 1) {
        if (t < off) red[t] += red[t + off];
        __syncthreads();
    }
    if (t == 0) tile_sum[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void excl_scan_kernel(uint32_t *__restrict__ v, uint64_t total,
                                                         uint32_t *__restrict__ grand)
{
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const uint64_t chunk = (total + 1023) / 1024;
    const uint64_t lo = (uint64_t)t * chunk, hi = min(lo + chunk, total);
    uint32_t s = 0;
    for (uint64_t x = lo; x < hi; ++x) s += v[x];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t w = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += w;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (uint64_t x = lo; x < hi; ++x) {
        const uint32_t c = v[x];
        v[x] = run;
        run += c;
    }
    if (t == 1023) grand[0] = part[1023];
}

// compact flagged entries (in index order) into (i, partner) pairs
__global__ __launch_bounds__(JT) void compact_kernel(const uint32_t *__restrict__ flag,
                                                     const uint32_t *__restrict__ partner, uint64_t n,
                                                     const uint32_t *__restrict__ tile_off,
                                                     uint32_t *__restrict__ out_i, uint32_t *__restrict__ out_j)
{
    __shared__ uint32_t sc[JT];
    const int t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * JTILE + (uint64_t)t * JITEMS;   // blocked per thread
    uint32_t f[JITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < JITEMS; ++r) {
        const uint64_t x = base + r;
        f[r] = x < n ? flag[x] : 0u;
        s += f[r];
    }
    sc[t] = s;
    __syncthreads();
    for (int off = 1; off < JT; off <<= 1) {
        const uint32_t w = t >= off ? sc[t - off] : 0u;
        __syncthreads();
        sc[t] += w;
        __syncthreads();
    }
    uint32_t o = tile_off[blockIdx.x] + sc[t] - s;
#pragma unroll
    for (int r = 0; r < JITEMS; ++r) {
        if (f[r]) {
            const uint64_t x = base + r;
            out_i[o] = (uint32_t)x;
            out_j[o] = partner[x];
            ++o;
        }
    }
}

struct ExtEnv {
    const uint64_t *kp0, *kp1;   // GetDnaSeedMer by position
    uint64_t n0, n1;              // SMLLength
    int64_t L0, L1;               // sequence lengths
    int seed_len;
    uint64_t mer_mask;            // GetSeedMask(): top 2*weight bits
};

__device__ __forceinline__ bool lookup(const ExtEnv &e, int g, int64_t pos1, int64_t len, int64_t start,
                                       uint64_t &mer, int &parity)
{
    int64_t p = pos1;
    if (p < 0) p = -p + len - e.seed_len;
    const int64_t idx = p - 1;
    const uint64_t n = g ? e.n1 : e.n0;
    if (idx < 0 || (uint64_t)idx >= n) return false;
    const uint64_t k = (g ? e.kp1 : e.kp0)[idx];
    parity = start < 0 ? (int)(k & 1ull) : (int)!(k & 1ull);
    mer = k & e.mer_mask;
    return true;
}

// ExtendMatch, LM/MatchFinder.h:218-374, for a two-genome seed (both starts
// present, linear sequences).  Returns false if it would read outside the
// lists (the reference reads out of bounds there; never seen in practice).
__device__ bool extend2(const ExtEnv &e, int64_t st[2], int64_t &len)
{
    const int64_t Lg[2] = {e.L0, e.L1};
    int jump = e.seed_len;
    uint32_t extend_limit = 0, extend_attempts = 0;
    bool extend_again = false;
    int64_t guard = 0;
    const int64_t guard_max = 8 * (e.L0 + e.L1) + 64;
    for (int dir = 0; dir < 4; ++dir) {
        int64_t maxlen = dir >= 2 ? (int64_t)e.seed_len : (int64_t)4294967295ll;
        for (int x = 0; x < 2; ++x) {
            if (st[x] < 0) {
                const int64_t rc_len = Lg[x] - len + st[x] + 1;
                if (rc_len < maxlen) maxlen = rc_len;
            } else if (st[x] - 1 < maxlen) {
                maxlen = st[x] - 1;
            }
        }
        int j = 0, i = 2;
        extend_limit = 0;
        extend_attempts = 0;
        while (maxlen - jump >= 0) {
            if (++guard > guard_max) return false;
            len += jump;
            maxlen -= jump;
            for (j = 0; j < 2; ++j)
                if (st[j] > 0) {
                    st[j] -= jump;
                    if (st[j] <= 0) st[j] += Lg[j];
                }
            uint64_t cm, om;
            int cp, op;
            if (!lookup(e, 0, st[0], len, st[0], cm, cp)) return false;
            if (!lookup(e, 1, st[1], len, st[1], om, op)) return false;
            i = 1;
            if (cm != om || cp != op) {
                if (dir < 2) maxlen = 0;
            } else {
                i = 2;
            }
            extend_attempts += jump;
            if (i == 2) extend_limit = extend_attempts;
            if (dir > 1 && extend_attempts == (uint32_t)e.seed_len) break;
        }
        if (i < 2) {
            len -= jump;
            for (; j > 0; j--)
                if (st[j - 1] >= 0) st[j - 1] += jump;
        }
        if (dir > 1 && extend_attempts > 0) {
            if (extend_limit > 0) extend_again = true;
            int unmatched = (int)(extend_attempts - extend_limit);
            if (i < 2) unmatched -= jump;
            len -= unmatched;
            for (j = 0; j < 2; ++j)
                if (st[j] > 0) {
                    st[j] += unmatched;
                    if (st[j] > Lg[j]) st[j] -= Lg[j];
                }
        }
        st[0] = -st[0];
        st[1] = -st[1];
        if (dir >= 1) jump = 1;
        if (dir == 3 && extend_again) {
            dir = -1;
            jump = e.seed_len;
            extend_again = false;
        }
    }
    return true;
}

__global__ void seed_kernel(const uint32_t *__restrict__ pi, const uint32_t *__restrict__ pj, uint64_t ns,
                            const uint64_t *__restrict__ ck0, const uint32_t *__restrict__ pos0,
                            const uint64_t *__restrict__ ck1, const uint32_t *__restrict__ pos1, ExtEnv env,
                            SeedRec *__restrict__ out, uint32_t *__restrict__ err)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ns) return;
    const uint32_t i = pi[x], j = pj[x];
    const int64_t s0 = (int64_t)pos0[i] + 1;
    int64_t s1 = (int64_t)pos1[j] + 1;
    // SetDirection: genome 0 is the reference; flip genome 1 if its strand flag differs
    if ((ck0[i] & 1ull) != (ck1[j] & 1ull)) s1 = -s1;
    int64_t st[2] = {s0, s1};
    int64_t len = env.seed_len;
    const bool ok = extend2(env, st, len);
    if (!ok) atomicAdd(err, 1u);
    SeedRec r;
    r.s0 = (int32_t)s0;
    r.s1 = (int32_t)s1;
    r.e0 = (int32_t)st[0];
    r.e1 = (int32_t)st[1];
    r.elen = (uint32_t)len;
    out[x] = r;
}

}  // namespace

// Seeds in HashMatch order with their extensions; returns the count.
int join_and_extend(mh_context *ctx, const mh_sml *a, const mh_sml *b, std::vector<SeedRec> &seeds,
                    uint32_t *repeat_runs)
{
    const uint64_t n0 = a->n, n1 = b->n;
    seeds.clear();
    *repeat_runs = 0;
    if (n0 == 0 || n1 == 0) return MH_OK;
    MH_CHECK(a->len < (1ull << 31) && b->len < (1ull << 31), MH_ERR_UNSUPPORTED,
             "genomes longer than 2^31 are not supported by the seed matcher");
    const uint64_t ntiles = (n0 + JTILE - 1) / JTILE;
    int rc;
    if ((rc = ctx->mf_scratch[0].ensure(n0 * 4 + 64)) != MH_OK) return rc;    // flag
    if ((rc = ctx->mf_scratch[1].ensure(n0 * 4 + 64)) != MH_OK) return rc;    // partner
    if ((rc = ctx->mf_scratch[2].ensure(ntiles * 4 + 64)) != MH_OK) return rc;
    if ((rc = ctx->mf_scratch[3].ensure(64)) != MH_OK) return rc;             // counters
    uint32_t *flag = ctx->mf_scratch[0].as<uint32_t>();
    uint32_t *partner = ctx->mf_scratch[1].as<uint32_t>();
    uint32_t *tiles = ctx->mf_scratch[2].as<uint32_t>();
    uint32_t *cnt = ctx->mf_scratch[3].as<uint32_t>();   // [0] repeat runs, [1] total seeds, [2] ext errors
    MH_HIP(hipMemsetAsync(cnt, 0, 16, ctx->stream));
    {
        TimedLaunch tl(ctx, "join");
        hipLaunchKernelGGL(join_kernel, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, ctx->stream, a->d_keys, n0,
                           b->d_keys, n1, flag, partner, cnt);
        hipLaunchKernelGGL(repeat1_kernel, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, ctx->stream, b->d_keys,
                           n1, a->d_keys, n0, cnt);
        hipLaunchKernelGGL(tile_count_kernel, dim3((unsigned)ntiles), dim3(JT), 0, ctx->stream, flag, n0, tiles);
        hipLaunchKernelGGL(excl_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, tiles, ntiles, cnt + 1);
        MH_HIP(hipGetLastError());
    }
    uint32_t h[2] = {0, 0};
    MH_HIP(hipMemcpyAsync(h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    *repeat_runs = h[0];
    const uint64_t ns = h[1];
    if (ns == 0) return MH_OK;
    if ((rc = ctx->mf_scratch[4].ensure(ns * 4 + 64)) != MH_OK) return rc;
    if ((rc = ctx->mf_scratch[5].ensure(ns * 4 + 64)) != MH_OK) return rc;
    if ((rc = ctx->mf_scratch[6].ensure(ns * sizeof(SeedRec) + 64)) != MH_OK) return rc;
    uint32_t *oi = ctx->mf_scratch[4].as<uint32_t>();
    uint32_t *oj = ctx->mf_scratch[5].as<uint32_t>();
    SeedRec *d_seeds = ctx->mf_scratch[6].as<SeedRec>();
    hipLaunchKernelGGL(compact_kernel, dim3((unsigned)ntiles), dim3(JT), 0, ctx->stream, flag, partner, n0, tiles, oi,
                       oj);
    ExtEnv env;
    env.kp0 = a->d_keys_pos;
    env.kp1 = b->d_keys_pos;
    env.n0 = n0;
    env.n1 = n1;
    env.L0 = (int64_t)a->len;
    env.L1 = (int64_t)b->len;
    env.seed_len = a->seed_length;
    env.mer_mask = a->seed_weight >= 32 ? ~0ull : (~0ull << (64 - 2 * a->seed_weight));
    {
        TimedLaunch tl(ctx, "extend");
        hipLaunchKernelGGL(seed_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream, oi, oj, ns,
                           a->d_keys, a->d_pos, b->d_keys, b->d_pos, env, d_seeds, cnt + 2);
        MH_HIP(hipGetLastError());
    }
    seeds.resize(ns);
    MH_HIP(hipMemcpyAsync(seeds.data(), d_seeds, ns * sizeof(SeedRec), hipMemcpyDeviceToHost, ctx->stream));
    uint32_t ext_err = 0;
    MH_HIP(hipMemcpyAsync(&ext_err, cnt + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    MH_CHECK(ext_err == 0, MH_ERR_UNSUPPORTED, "seed extension left the sorted mer lists");
    return MH_OK;
}

}  // namespace mh
