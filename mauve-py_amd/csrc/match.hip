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
// Runs longer than MER_REPEAT_LIMIT (LM/MatchFinder.cpp:166) can make the
// reference restart its merge from recomputed start points (253-276).
// heavy_kernel finds every masked mer with more than MER_REPEAT_LIMIT entries
// over all lists; when there is one, the host plans the restarts
// (restart.cpp) and the join reads only the index ranges the reference's
// search reads (live ranges).
#include "extend.hpp"
#include "mhe.hpp"
#include "mh_internal.hpp"

#include <algorithm>
#include <cstring>

namespace mh {

namespace {

constexpr int JT = 256;              // threads per join/compaction tile
constexpr int JITEMS = 16;
constexpr int JTILE = JT * JITEMS;
constexpr uint32_t MER_REPEAT_LIMIT = 1000;
constexpr int64_t LONG_WALK = 64;   // steps of a head's walk before it is resolved per diagonal

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

// the masked mer of each seed's genome-a start (the top 2 * weight bits of its
// position key), shifted up by three bits for the pair number
__global__ void seed_mer_kernel(const SeedRec *__restrict__ seeds, uint64_t ns, const uint64_t *__restrict__ kp,
                                int shift, uint32_t pair, uint64_t *__restrict__ out)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ns) return;
    out[x] = ((kp[seeds[x].s0 - 1] >> shift) << 3) | pair;
}

// Up to PLAN_MAXK sorted lists.
struct KLists {
    const uint64_t *ck[PLAN_MAXK];
    uint64_t n[PLAN_MAXK];
    int k;
};

// Masked mers of more than MER_REPEAT_LIMIT entries over all k lists.  Such a
// run holds more than MER_REPEAT_LIMIT / k entries in some list, so only run
// starts of list i with at least `thresh` = MER_REPEAT_LIMIT / k + 1 entries
// count the others (binary searches); a mer may be reported from several
// lists (the host removes duplicates).
__global__ void heavy_kernel(KLists L, int i, uint32_t thresh, uint64_t *__restrict__ out, uint32_t cap,
                             uint32_t *__restrict__ count)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= L.n[i]) return;
    const uint64_t *ci = L.ck[i];
    const uint64_t m = ci[x] >> 1;
    if (x > 0 && (ci[x - 1] >> 1) == m) return;
    if (x + thresh - 1 >= L.n[i] || (ci[x + thresh - 1] >> 1) != m) return;
    uint64_t total = 0;
    for (int j = 0; j < L.k; ++j) {
        const uint64_t lo = lower_bound_m(L.ck[j], L.n[j], m);
        if (lo < L.n[j] && (L.ck[j][lo] >> 1) == m) total += upper_bound_m(L.ck[j], L.n[j], m) - lo;
    }
    if (total > MER_REPEAT_LIMIT) {
        const uint32_t at = atomicAdd(count, 1u);
        if (at < cap) out[at] = m;
    }
}

// Live ranges of the two lists of a join: n == 0 means the whole list.
struct LivePair {
    const uint64_t *lo0, *hi0, *lo1, *hi1;
    uint32_t c0, c1;
};

// the live range [ilo, ihi) holding index i (false when i is in none)
__device__ __forceinline__ bool live_at(const uint64_t *__restrict__ lo, const uint64_t *__restrict__ hi, uint32_t c,
                                        uint64_t n, uint64_t i, uint64_t &ilo, uint64_t &ihi)
{
    if (c == 0) {
        ilo = 0;
        ihi = n;
        return true;
    }
    uint32_t a = 0, len = c;    // last range with lo <= i
    while (len > 0) {
        const uint32_t half = len >> 1, mid = a + half;
        if (lo[mid] <= i) { a = mid + 1; len -= half + 1; }
        else len = half;
    }
    if (a == 0) return false;
    ilo = lo[a - 1];
    ihi = hi[a - 1];
    return i < ihi;
}

// flag[i] = 1 if genome-0 entry i forms a seed; partner[i] = its genome-1
// index.  A seed is a masked mer with exactly one live entry in each list.
__global__ void join_kernel(const uint64_t *__restrict__ ck0, uint64_t n0, const uint64_t *__restrict__ ck1,
                            uint64_t n1, LivePair live, uint32_t *__restrict__ flag, uint32_t *__restrict__ partner)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0) return;
    uint32_t f = 0, pj = 0;
    uint64_t ilo, ihi;
    if (live_at(live.lo0, live.hi0, live.c0, n0, i, ilo, ihi)) {
        const uint64_t m = ck0[i] >> 1;
        const bool first = i == ilo || (ck0[i - 1] >> 1) != m;
        const bool last = i + 1 == ihi || (ck0[i + 1] >> 1) != m;
        if (first && last) {
            const uint64_t lo = lower_bound_m(ck1, n1, m);
            if (lo < n1 && (ck1[lo] >> 1) == m) {
                uint64_t jlo, jhi;
                if (lo + 1 == n1 || (ck1[lo + 1] >> 1) != m) {
                    if (live_at(live.lo1, live.hi1, live.c1, n1, lo, jlo, jhi)) {
                        f = 1;
                        pj = (uint32_t)lo;
                    }
                } else if (live.c1) {
                    // a run the last restart cut: one live entry left?
                    const uint64_t hi = upper_bound_m(ck1, n1, m);
                    if (live_at(live.lo1, live.hi1, live.c1, n1, hi - 1, jlo, jhi) && (lo > jlo ? lo : jlo) == hi - 1) {
                        f = 1;
                        pj = (uint32_t)(hi - 1);
                    }
                }
            }
        }
    }
    flag[i] = f;
    partner[i] = pj;
}

// ---- tiled join (every list-0 entry live) ---------------------------------
// List 0 is cut into tiles of TJ_I entries.  A tile's masked mers lie in
// [m(first), m(last)]; the list-1 entries with those mers, [jlo, jhi), are
// found once per tile and staged in LDS with the tile, so each entry's search
// in list 1 is a binary search in LDS instead of ~22 dependent global loads.
// The tile is joined twice: once to count its seeds, once (after a scan of
// the counts) to write them at their place in list-0 (HashMatch) order.
constexpr int TJ_T = 256;                 // threads per tile
constexpr int TJ_ITEMS = 8;               // list-0 entries per thread (consecutive)
constexpr int TJ_I = TJ_T * TJ_ITEMS;     // list-0 entries per tile
constexpr int TJ_CAP = 4096;              // list-1 entries a tile can stage
constexpr uint64_t TJ_NONE = ~0ull;       // not a masked mer (those fit 62 bits)

__global__ void tile_bounds_kernel(const uint64_t *__restrict__ ck0, uint64_t n0, const uint64_t *__restrict__ ck1,
                                   uint64_t n1, uint64_t ntiles, uint64_t *__restrict__ jb)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const uint64_t i0 = t * TJ_I, i1 = min(n0, i0 + TJ_I);
    jb[2 * t] = lower_bound_m(ck1, n1, ck0[i0] >> 1);
    jb[2 * t + 1] = upper_bound_m(ck1, n1, ck0[i1 - 1] >> 1);
}

// write == 0: cnt[tile] = the tile's seed count; write == 1: the seeds
// (list-0 index, list-1 index) at off[tile] onwards, in list-0 order
__global__ __launch_bounds__(TJ_T) void tile_join_kernel(const uint64_t *__restrict__ ck0, uint64_t n0,
                                                         const uint64_t *__restrict__ ck1, uint64_t n1,
                                                         const uint64_t *__restrict__ jb, int write,
                                                         uint32_t *__restrict__ cnt, const uint32_t *__restrict__ off,
                                                         uint32_t *__restrict__ out_i, uint32_t *__restrict__ out_j)
{
    __shared__ uint64_t s0[TJ_I + 2];
    __shared__ uint64_t s1[TJ_CAP];
    __shared__ uint32_t sc[TJ_T];
    const int t = threadIdx.x;
    const uint64_t tile = blockIdx.x;
    const uint64_t i0 = tile * TJ_I, i1 = min(n0, i0 + TJ_I), ni = i1 - i0;
    const uint64_t jlo = jb[2 * tile], jhi = jb[2 * tile + 1], nj = jhi - jlo;
    const bool staged = nj <= (uint64_t)TJ_CAP;
    // s0[k] = masked mer of list-0 entry i0 - 1 + k (TJ_NONE outside the list)
    for (uint64_t k = t; k < ni + 2; k += TJ_T) {
        const int64_t idx = (int64_t)i0 - 1 + (int64_t)k;
        s0[k] = (idx >= 0 && (uint64_t)idx < n0) ? ck0[idx] >> 1 : TJ_NONE;
    }
    if (staged)
        for (uint64_t k = t; k < nj; k += TJ_T) s1[k] = ck1[jlo + k] >> 1;
    __syncthreads();
    uint32_t part[TJ_ITEMS];
    uint32_t c = 0;
    // the thread's entries are consecutive, so one binary search finds the
    // first one's lower bound in list 1 and the others advance from there
    auto v1 = [&](uint64_t q) { return staged ? s1[q] : (ck1[jlo + q] >> 1); };
    uint64_t lo = 0;
    {
        const uint64_t k0 = (uint64_t)t * TJ_ITEMS;
        const uint64_t m0 = k0 < ni ? s0[k0 + 1] : TJ_NONE;
        uint64_t len = nj;
        while (len > 0) {
            const uint64_t half = len >> 1, mid = lo + half;
            if (v1(mid) < m0) { lo = mid + 1; len -= half + 1; }
            else len = half;
        }
    }
#pragma unroll
    for (int r = 0; r < TJ_ITEMS; ++r) {
        const uint64_t k = (uint64_t)t * TJ_ITEMS + r;
        part[r] = ~0u;
        if (k >= ni) continue;
        const uint64_t m = s0[k + 1];
        while (lo < nj && v1(lo) < m) ++lo;               // lower bound of m in list-1 entries [jlo, jhi)
        if (s0[k] == m || s0[k + 2] == m) continue;      // not unique in list 0
        if (lo >= nj || v1(lo) != m) continue;
        if (lo + 1 < nj && v1(lo + 1) == m) continue;    // not unique in list 1
        part[r] = (uint32_t)(jlo + lo);
        ++c;
    }
    // exclusive scan of the per-thread counts (threads hold consecutive entries)
    sc[t] = c;
    __syncthreads();
    for (int o = 1; o < TJ_T; o <<= 1) {
        const uint32_t w = t >= o ? sc[t - o] : 0u;
        __syncthreads();
        sc[t] += w;
        __syncthreads();
    }
    if (!write) {
        if (t == TJ_T - 1) cnt[tile] = sc[t];
        return;
    }
    uint32_t o = off[tile] + sc[t] - c;
#pragma unroll
    for (int r = 0; r < TJ_ITEMS; ++r)
        if (part[r] != ~0u) {
            out_i[o] = (uint32_t)(i0 + (uint64_t)t * TJ_ITEMS + r);
            out_j[o] = part[r];
            ++o;
        }
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

__global__ void seed_kernel(const uint32_t *__restrict__ pi, const uint32_t *__restrict__ pj, uint64_t ns,
                            const uint64_t *__restrict__ ck0, const uint32_t *__restrict__ pos0,
                            const uint64_t *__restrict__ ck1, const uint32_t *__restrict__ pos1, ExtEnv env,
                            int64_t budget, SeedRec *__restrict__ out, uint32_t *__restrict__ err,
                            uint32_t *__restrict__ long_flag)
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
    const int rc = extend2(env, st, len, budget);
    if (long_flag) long_flag[x] = rc == EXT_LONG;
    if (rc == EXT_ERR) atomicAdd(err, 1u);
    SeedRec r;
    r.s0 = (int32_t)s0;
    r.s1 = (int32_t)s1;
    r.e0 = (int32_t)st[0];
    r.e1 = (int32_t)st[1];
    r.elen = (uint32_t)len;
    out[x] = r;
}

// ---- spaced seeds: one extension per grid chain -----------------------------
// Two seeds of the same strand relation whose starts differ by the seed length
// along their diagonal (forward: s0 + L, s1 + L; reverse: s0 + L, |s1| - L)
// extend to the same match: the later one's seed-length walk to the left steps
// onto the earlier one's window (a seed, so it matches) and from there repeats
// the earlier one's walk, with one step less of room; to the right the earlier
// one steps onto the later one's window the same way; the single-step phases
// then start from the same match.  So each chain of such seeds is extended once,
// from its first seed, and the result is copied to the others.

__global__ void grid_key_kernel(const uint32_t *__restrict__ oi, const uint32_t *__restrict__ oj, uint64_t ns,
                                const uint64_t *__restrict__ ck0, const uint32_t *__restrict__ pos0,
                                const uint64_t *__restrict__ ck1, const uint32_t *__restrict__ pos1,
                                uint64_t *__restrict__ key, uint32_t *__restrict__ val, int32_t *__restrict__ t1)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ns) return;
    const uint32_t i = oi[x], j = oj[x];
    const int32_t s1 = (int32_t)pos1[j] + 1;
    key[x] = (uint64_t)pos0[i] + 1;
    val[x] = (uint32_t)x;
    t1[x] = (ck0[i] & 1ull) != (ck1[j] & 1ull) ? -s1 : s1;   // the seed's signed start in genome 2
}

// link[r] = sorted index of the seed one seed length before r on its chain, or
// r.  Seeds are sorted by their start in genome 1; the predecessor of (s0, t)
// is (s0 - L, t - L) with t's sign kept (forward: s1 - L; reverse: the start
// in genome 2 moves right, -(s1 + L)).  At most `scan` seeds sharing a start
// are looked at: a seed left without its link only extends by itself.
__global__ void grid_pred_kernel(const uint64_t *__restrict__ sk, const uint32_t *__restrict__ sv,
                                 const int32_t *__restrict__ t1, uint64_t ns, uint32_t seed_len, int scan,
                                 uint32_t *__restrict__ link)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ns) return;
    const uint64_t k = sk[r];
    const int32_t t = t1[sv[r]];
    const int32_t tw = t - (int32_t)seed_len;
    uint32_t out = (uint32_t)r;
    if (k > seed_len && (t < 0 || tw > 0)) {
        const uint64_t want = k - seed_len;
        uint64_t lo = 0, hi = r;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            if (sk[mid] < want) lo = mid + 1;
            else hi = mid;
        }
        for (int c = 0; c < scan && lo < r && sk[lo] == want; ++c, ++lo)
            if (t1[sv[lo]] == tw) {
                out = (uint32_t)lo;
                break;
            }
    }
    link[r] = out;
}

__global__ void grid_jump_kernel(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, uint64_t ns)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ns) return;
    out[r] = in[in[r]];
}

// flag[r]: sorted seed r heads its chain
__global__ void chain_root_flag_kernel(const uint32_t *__restrict__ root, uint64_t ns, uint32_t *__restrict__ flag)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ns) return;
    flag[r] = root[r] == (uint32_t)r;
}

// sorted positions of the heads to their seed indices, in place
__global__ void chain_gather_kernel(const uint32_t *__restrict__ sv, uint64_t nh, uint32_t *__restrict__ heads)
{
    const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    heads[h] = sv[heads[h]];
}

// per seed (original order): is it a chain head, and which seed heads its chain
__global__ void grid_head_kernel(const uint32_t *__restrict__ sv, const uint32_t *__restrict__ root, uint64_t ns,
                                 uint32_t *__restrict__ is_head, uint32_t *__restrict__ head_of)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ns) return;
    const uint32_t x = sv[r];
    is_head[x] = root[r] == (uint32_t)r;
    head_of[x] = sv[root[r]];
}

__global__ void grid_gather_kernel(const uint32_t *__restrict__ heads, uint64_t nh, const uint32_t *__restrict__ oi,
                                   const uint32_t *__restrict__ oj, uint32_t *__restrict__ hoi,
                                   uint32_t *__restrict__ hoj, uint32_t *__restrict__ headpos)
{
    const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const uint32_t x = heads[h];
    hoi[h] = oi[x];
    hoj[h] = oj[x];
    headpos[x] = (uint32_t)h;
}

__global__ void grid_fill_kernel(const uint32_t *__restrict__ oi, const uint32_t *__restrict__ oj, uint64_t ns,
                                 const uint64_t *__restrict__ ck0, const uint32_t *__restrict__ pos0,
                                 const uint64_t *__restrict__ ck1, const uint32_t *__restrict__ pos1,
                                 const uint32_t *__restrict__ head_of, const uint32_t *__restrict__ headpos,
                                 const SeedRec *__restrict__ hres, SeedRec *__restrict__ out)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ns) return;
    const uint32_t i = oi[x], j = oj[x];
    const int64_t s0 = (int64_t)pos0[i] + 1;
    int64_t s1 = (int64_t)pos1[j] + 1;
    if ((ck0[i] & 1ull) != (ck1[j] & 1ull)) s1 = -s1;
    const SeedRec h = hres[headpos[head_of[x]]];
    SeedRec r;
    r.s0 = (int32_t)s0;
    r.s1 = (int32_t)s1;
    r.e0 = h.e0;
    r.e1 = h.e1;
    r.elen = h.elen;
    out[x] = r;
}

// ---- long extensions: both ends resolved per diagonal ----------------------
// ExtendMatch's two ends move independently.  Within one round the left end
// takes seed-length jumps while the next window to the left matches, then a
// single-step phase moves it to the furthest matching window within one seed
// length (or leaves it); the right end does the same to the right; rounds
// repeat while either single-step phase moved.  Each end's moves depend only
// on its own position (the room to the sequence ends and the windows read are
// functions of it, the diagonal fixed), and an end whose single-step phase
// found nothing finds nothing again.  So the final left end of a seed is the
// fixed point of F(p) = J(p) - T(J(p)) from its start, J(p) the last window of
// p's run of matching windows spaced one seed length apart and T(h) the
// largest t <= seed length with window h - t matching (0 if none); the right
// end mirrors it.  For the seeds whose step-by-step walk is long, F is built
// for every window of a range around them on their diagonal and the fixed
// points found by pointer jumping.  A window whose value depends on windows
// beyond a range that was cut short (not by the sequence ends) is flagged, and
// its diagonal is redone with a wider range.

struct DiagGroup {
    int64_t lo;        // first window start in genome 1 (1-based) of the range
    int64_t n;         // windows in the range
    int64_t base;      // first element of the range in the dense arrays
    int64_t dconst;    // forward: s1 - s0; reverse: s0 + s1 (window starts pair as p + q)
    int32_t rev, clip_lo, clip_hi, pad;
};

struct DiagHead {
    uint32_t head;     // index into the head results
    uint32_t group;
    int32_t s0, t1;    // start in genome 1, signed start in genome 2
};

constexpr uint32_t DG_ESC = 0x80000000u, DG_IDX = 0x7fffffffu;

__device__ __forceinline__ uint32_t dg_find(const DiagGroup *__restrict__ g, uint32_t ng, uint64_t e)
{
    uint32_t lo = 0, hi = ng;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if ((uint64_t)g[mid].base <= e) lo = mid;
        else hi = mid;
    }
    return lo;
}

// window e of its range matches (both windows inside the sequences by construction)
__global__ void dg_match_kernel(const DiagGroup *__restrict__ g, uint32_t ng, uint64_t N, ExtEnv env,
                                uint8_t *__restrict__ mg)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    const DiagGroup d = g[dg_find(g, ng, e)];
    const int64_t p = d.lo + (int64_t)(e - (uint64_t)d.base);
    const int64_t q = d.rev ? d.dconst - p : p + d.dconst;
    const uint64_t k0 = env.kp0[p - 1], k1 = env.kp1[q - 1];
    mg[e] = ((k0 ^ k1) & env.mer_mask) == 0 && ((k0 ^ k1) & 1ull) == (uint64_t)d.rev;
}

// one seed-length step of the jump runs, left and right
__global__ void dg_link_kernel(const DiagGroup *__restrict__ g, uint32_t ng, uint64_t N, int L,
                               const uint8_t *__restrict__ mg, uint32_t *__restrict__ cl, uint32_t *__restrict__ cr)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    const DiagGroup d = g[dg_find(g, ng, e)];
    const int64_t off = (int64_t)(e - (uint64_t)d.base);
    uint32_t l = (uint32_t)e, r = (uint32_t)e;
    if (off >= L) {
        if (mg[e - L]) l = (uint32_t)(e - L);
    } else if (d.clip_lo) {
        l |= DG_ESC;
    }
    if (off + L < d.n) {
        if (mg[e + L]) r = (uint32_t)(e + L);
    } else if (d.clip_hi) {
        r |= DG_ESC;
    }
    cl[e] = l;
    cr[e] = r;
}

// pointer jumping; a root keeps its own flag, which the others take with it
__global__ void dg_jump_kernel(uint64_t N, const uint32_t *__restrict__ il, const uint32_t *__restrict__ ir,
                               uint32_t *__restrict__ ol, uint32_t *__restrict__ orr)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    ol[e] = il[il[e] & DG_IDX];
    orr[e] = ir[ir[e] & DG_IDX];
}

// F (left) and its mirror (right) from the jump-run ends hl, hr
__global__ void dg_step_kernel(const DiagGroup *__restrict__ g, uint32_t ng, uint64_t N, int L,
                               const uint8_t *__restrict__ mg, const uint32_t *__restrict__ hl,
                               const uint32_t *__restrict__ hr, uint32_t *__restrict__ fl, uint32_t *__restrict__ fr)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    const DiagGroup d = g[dg_find(g, ng, e)];
    uint32_t h = hl[e], esc = h & DG_ESC;
    h &= DG_IDX;
    int64_t ho = (int64_t)h - d.base;
    int T = 0;
    for (int t = 1; t <= L; ++t) {
        if (ho - t < 0) {
            if (d.clip_lo) esc = DG_ESC;
            break;
        }
        if (mg[h - t]) T = t;
    }
    fl[e] = (h - (uint32_t)T) | esc;
    h = hr[e];
    esc = h & DG_ESC;
    h &= DG_IDX;
    ho = (int64_t)h - d.base;
    T = 0;
    for (int t = 1; t <= L; ++t) {
        if (ho + t >= d.n) {
            if (d.clip_hi) esc = DG_ESC;
            break;
        }
        if (mg[h + t]) T = t;
    }
    fr[e] = (h + (uint32_t)T) | esc;
}

// pointer jumping that collects the flags along the path
__global__ void dg_jump_or_kernel(uint64_t N, const uint32_t *__restrict__ il, const uint32_t *__restrict__ ir,
                                  uint32_t *__restrict__ ol, uint32_t *__restrict__ orr)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N) return;
    uint32_t a = il[e], b = il[a & DG_IDX];
    ol[e] = (b & DG_IDX) | ((a | b) & DG_ESC);
    a = ir[e];
    b = ir[a & DG_IDX];
    orr[e] = (b & DG_IDX) | ((a | b) & DG_ESC);
}

__global__ void dg_heads_kernel(const uint32_t *__restrict__ long_idx, uint64_t nl, const uint32_t *__restrict__ hoi,
                                const uint32_t *__restrict__ hoj, const uint64_t *__restrict__ ck0,
                                const uint32_t *__restrict__ pos0, const uint64_t *__restrict__ ck1,
                                const uint32_t *__restrict__ pos1, DiagHead *__restrict__ out)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nl) return;
    const uint32_t h = long_idx[x], i = hoi[h], j = hoj[h];
    const int32_t s1 = (int32_t)pos1[j] + 1;
    DiagHead r;
    r.head = h;
    r.group = 0;
    r.s0 = (int32_t)pos0[i] + 1;
    r.t1 = (ck0[i] & 1ull) != (ck1[j] & 1ull) ? -s1 : s1;
    out[x] = r;
}

__global__ void dg_result_kernel(const DiagHead *__restrict__ hd, uint64_t nh, const DiagGroup *__restrict__ g,
                                 const uint32_t *__restrict__ fl, const uint32_t *__restrict__ fr, int L,
                                 SeedRec *__restrict__ res, uint8_t *__restrict__ esc)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nh) return;
    const DiagHead h = hd[x];
    const DiagGroup d = g[h.group];
    const uint64_t e = (uint64_t)d.base + (uint64_t)(h.s0 - d.lo);
    const uint32_t a = fl[e], b = fr[e];
    esc[x] = ((a | b) & DG_ESC) != 0;
    const int64_t left = d.lo + ((int64_t)(a & DG_IDX) - d.base);
    const int64_t last = d.lo + ((int64_t)(b & DG_IDX) - d.base);   // start of the rightmost window
    const int64_t len = last - left + L;
    const int64_t s0 = h.s0, s1 = h.t1 < 0 ? -(int64_t)h.t1 : h.t1;
    SeedRec r;
    r.s0 = h.s0;
    r.s1 = h.t1;
    r.e0 = (int32_t)left;
    r.e1 = (int32_t)(d.rev ? -(s0 + s1 + L - left - len) : left + (s1 - s0));
    r.elen = (uint32_t)len;
    res[h.head] = r;
}

// ---- solid odd-length seeds ----------------------------------------------
// With a solid seed of odd length k, a window and its reverse complement are
// never equal, so comparing canonical keys plus strand parity (ExtendMatch's
// test) is exactly comparing the k-character windows themselves (forward, or
// genome 0 against the reverse complement of genome 1).  The four-phase
// extension then stops at the first mismatching character on each side, or at
// a sequence end: the result is the maximal exact match through the seed.
// It is computed here 32 characters per 64-bit XOR of packed words.

__device__ __forceinline__ uint64_t win32(const uint32_t *__restrict__ w, uint64_t pos)
{
    // 32 characters starting at pos (first character in the top two bits)
    const uint64_t q = pos >> 4;
    const uint32_t bit = (uint32_t)(pos & 15) * 2;
    uint64_t m = ((uint64_t)w[q] << 32) | w[q + 1];
    if (bit) m = (m << bit) | (w[q + 2] >> (32 - bit));
    return m;
}

// number of equal characters going right from (pa, pb), at most lim
__device__ __forceinline__ int64_t match_right(const uint32_t *__restrict__ wa, int64_t pa,
                                               const uint32_t *__restrict__ wb, int64_t pb, int64_t lim)
{
    int64_t n = 0;
    while (n < lim) {
        const uint64_t x = win32(wa, (uint64_t)(pa + n)) ^ win32(wb, (uint64_t)(pb + n));
        const int64_t eq = x ? (int64_t)(__clzll(x) >> 1) : 32;
        n += eq;
        if (eq < 32) break;
    }
    return n < lim ? n : lim;
}

// number of equal characters going left from (pa-1, pb-1), at most lim
__device__ __forceinline__ int64_t match_left(const uint32_t *__restrict__ wa, int64_t pa,
                                              const uint32_t *__restrict__ wb, int64_t pb, int64_t lim)
{
    int64_t n = 0;
    while (n < lim) {
        // the 32 characters ending just before pa-n (clamped at 0)
        const int64_t take = lim - n < 32 ? lim - n : 32;
        const uint64_t xa = win32(wa, (uint64_t)(pa - n - take)) >> (64 - 2 * take);
        const uint64_t xb = win32(wb, (uint64_t)(pb - n - take)) >> (64 - 2 * take);
        const uint64_t x = xa ^ xb;
        const int64_t eq = x ? (int64_t)(__ffsll((long long)x) - 1) >> 1 : take;
        n += eq;
        if (eq < take) break;
    }
    return n < lim ? n : lim;
}

__global__ void rc_pack_kernel(const uint32_t *__restrict__ w, uint64_t len, uint64_t n_words,
                               uint32_t *__restrict__ out)
{
    // packed reverse complement: character q of the output is 3 - char(len-1-q)
    const uint64_t ow = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ow >= n_words) return;
    uint32_t v = 0;
    for (int k = 0; k < 16; ++k) {
        const uint64_t q = ow * 16 + k;
        if (q < len) {
            const uint64_t src = len - 1 - q;
            const uint32_t c = (w[src >> 4] >> (30 - 2 * (src & 15))) & 3u;
            v |= (3u - c) << (30 - 2 * k);
        }
    }
    out[ow] = v;
}

// ---- odd solid seeds by runs ------------------------------------------------
// Two seeds one position apart on one diagonal (forward: (p0, p1) and (p0+1,
// p1+1); reverse, in genome 1's forward coordinates: (p0, p1) and (p0+1,
// p1-1)) are exact matches that overlap in k-1 characters, so they extend to
// the same maximal exact match.  A run of such seeds is found through
// genome-0 positions (each position holds at most one seed: its mer is
// unique in list 0): only the run's first seed walks left and its last seed
// walks right, and every seed of the run takes their ends.  The first and
// last seed of the run at each position come from a segmented scan over the
// positions (the nearest head at or before, the nearest tail at or after).

constexpr uint32_t NO_SEED = 0xFFFFFFFFu;
constexpr int RS_T = 256, RS_ITEMS = 16, RS_P = RS_T * RS_ITEMS;   // positions per scan block
constexpr uint8_t RF_SEED = 4, RF_FOLLOW = 1, RF_NEXT = 2;       // position flags

// seed order -> genome-0 position order: pp[p0] = (genome-1 forward position
// with the strand relation in bit 31, seed index)
__global__ void run_scatter_kernel(const uint32_t *__restrict__ pi, const uint32_t *__restrict__ pj, uint64_t ns,
                                   const uint64_t *__restrict__ ck0, const uint32_t *__restrict__ pos0,
                                   const uint64_t *__restrict__ ck1, const uint32_t *__restrict__ pos1,
                                   uint2 *__restrict__ pp)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ns) return;
    const uint32_t i = pi[x], j = pj[x];
    const uint32_t p0 = pos0[i];
    const bool rev = (ck0[i] & 1ull) != (ck1[j] & 1ull);
    pp[p0] = make_uint2(pos1[j] | (rev ? 0x80000000u : 0u), (uint32_t)x);
}

// per position: a seed (RF_SEED), continued from the seed at p - 1
// (RF_FOLLOW: not a head), continued into the seed at p + 1 (RF_NEXT: not a
// tail); forward seeds continue at (p+1, p1+1), reverse ones at (p+1, p1-1)
__global__ void run_flags_kernel(const uint2 *__restrict__ pp, uint64_t L0, uint8_t *__restrict__ flag)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L0) return;
    const uint32_t q = pp[p].x;
    uint8_t f = 0;
    if (q != NO_SEED) {
        f = RF_SEED;
        const uint32_t rev = q & 0x80000000u, p1 = q & 0x7FFFFFFFu;
        if (p > 0) {
            const uint32_t y = pp[p - 1].x;
            if (y != NO_SEED && (y & 0x80000000u) == rev && (y & 0x7FFFFFFFu) == (rev ? p1 + 1 : p1 - 1))
                f |= RF_FOLLOW;
        }
        if (p + 1 < L0) {
            const uint32_t z = pp[p + 1].x;
            if (z != NO_SEED && (z & 0x80000000u) == rev && (z & 0x7FFFFFFFu) == (rev ? p1 - 1 : p1 + 1))
                f |= RF_NEXT;
        }
    }
    flag[p] = f;
}

__device__ __forceinline__ bool rf_head(uint8_t f) { return (f & (RF_SEED | RF_FOLLOW)) == RF_SEED; }
__device__ __forceinline__ bool rf_tail(uint8_t f) { return (f & (RF_SEED | RF_NEXT)) == RF_SEED; }

// per block of RS_P positions: the last head and the first tail in it (-1 / L0)
__global__ __launch_bounds__(RS_T) void run_block_kernel(const uint8_t *__restrict__ flag, uint64_t L0,
                                                         int64_t *__restrict__ bhead, int64_t *__restrict__ btail)
{
    __shared__ int64_t h[RS_T], tl[RS_T];
    const int t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * RS_P + (uint64_t)t * RS_ITEMS;
    int64_t lh = -1, ft = (int64_t)L0;
    for (int r = 0; r < RS_ITEMS; ++r) {
        const uint64_t p = base + r;
        if (p >= L0) break;
        const uint8_t f = flag[p];
        if (rf_head(f)) lh = (int64_t)p;
        if (rf_tail(f) && ft == (int64_t)L0) ft = (int64_t)p;
    }
    h[t] = lh;
    tl[t] = ft;
    __syncthreads();
    for (int o = RS_T / 2; o > 0; o >>= 1) {
        if (t < o) {
            h[t] = max(h[t], h[t + o]);
            tl[t] = min(tl[t], tl[t + o]);
        }
        __syncthreads();
    }
    if (t == 0) {
        bhead[blockIdx.x] = h[0];
        btail[blockIdx.x] = tl[0];
    }
}

// carry-in of every block: the last head before it, the first tail after it
__global__ __launch_bounds__(1024) void run_carry_kernel(const int64_t *__restrict__ bhead,
                                                         const int64_t *__restrict__ btail, uint64_t nb, uint64_t L0,
                                                         int64_t *__restrict__ chead, int64_t *__restrict__ ctail)
{
    __shared__ int64_t ph[1024], pt[1024];
    const int t = threadIdx.x;
    const uint64_t chunk = (nb + 1023) / 1024;
    const uint64_t lo = min((uint64_t)t * chunk, nb), hi = min(lo + chunk, nb);
    int64_t mh = -1, mt = (int64_t)L0;
    for (uint64_t b = lo; b < hi; ++b) mh = max(mh, bhead[b]);
    for (uint64_t b = lo; b < hi; ++b) mt = min(mt, btail[b]);
    ph[t] = mh;
    pt[t] = mt;
    __syncthreads();
    // inclusive max scan forward, inclusive min scan backward
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t a = t >= o ? ph[t - o] : -1;
        const int64_t b = t + o < 1024 ? pt[t + o] : (int64_t)L0;
        __syncthreads();
        ph[t] = max(ph[t], a);
        pt[t] = min(pt[t], b);
        __syncthreads();
    }
    int64_t run = t > 0 ? ph[t - 1] : -1;
    for (uint64_t b = lo; b < hi; ++b) {
        chead[b] = run;
        run = max(run, bhead[b]);
    }
    int64_t rt = t + 1 < 1024 ? pt[t + 1] : (int64_t)L0;
    for (uint64_t b = hi; b > lo; --b) {
        ctail[b - 1] = rt;
        rt = min(rt, btail[b - 1]);
    }
}

// heads walk left, tails walk right (position order)
__global__ void run_walk_kernel(const uint8_t *__restrict__ flag, const uint2 *__restrict__ pp, uint64_t L0,
                                const uint32_t *__restrict__ wa, const uint32_t *__restrict__ wb,
                                const uint32_t *__restrict__ wbrc, int64_t La, int64_t Lb, int k,
                                uint32_t *__restrict__ lft, uint32_t *__restrict__ rgt)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L0) return;
    const uint8_t f = flag[p];
    const bool hd = rf_head(f), tl = rf_tail(f);
    if (!hd && !tl) return;
    const int64_t p0 = (int64_t)p;
    const uint32_t q = pp[p].x;
    const bool rev = (q & 0x80000000u) != 0;
    const int64_t p1f = q & 0x7FFFFFFFu;
    const int64_t p1 = rev ? Lb - p1f - k : p1f;
    const uint32_t *w1 = rev ? wbrc : wb;
    if (hd) lft[p] = (uint32_t)match_left(wa, p0, w1, p1, p0 < p1 ? p0 : p1);
    if (tl) {
        const int64_t ra = La - (p0 + k), rb = Lb - (p1 + k);
        rgt[p] = (uint32_t)match_right(wa, p0 + k, w1, p1 + k, ra < rb ? ra : rb);
    }
}

// every seed's record from its run's head (left end) and tail (right end),
// written at its seed index.  The scan runs over RS_ITEMS consecutive
// positions per thread; each seed's head and tail go through LDS (padded one
// word per 16) so that the reads of pp run over consecutive positions across
// the wave.
__device__ __forceinline__ int rs_lds(int idx) { return idx + (idx >> 4); }

__global__ __launch_bounds__(RS_T) void run_final_kernel(const uint8_t *__restrict__ flag,
                                                         const uint2 *__restrict__ pp, uint64_t L0,
                                                         const int64_t *__restrict__ chead,
                                                         const int64_t *__restrict__ ctail,
                                                         const uint32_t *__restrict__ lft,
                                                         const uint32_t *__restrict__ rgt, int64_t Lb, int k,
                                                         SeedRec *__restrict__ out)
{
    __shared__ int64_t h[RS_T], tt[RS_T];
    __shared__ uint32_t hs[RS_P + RS_P / 16], ts[RS_P + RS_P / 16];
    const int t = threadIdx.x;
    const uint64_t bbase = (uint64_t)blockIdx.x * RS_P;
    const uint64_t base = bbase + (uint64_t)t * RS_ITEMS;
    uint8_t fv[RS_ITEMS];
    int64_t lh = -1, ft = (int64_t)L0;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const uint64_t p = base + r;
        fv[r] = p < L0 ? flag[p] : (uint8_t)0;
        if (rf_head(fv[r])) lh = (int64_t)p;
        if (rf_tail(fv[r]) && ft == (int64_t)L0) ft = (int64_t)p;
    }
    h[t] = lh;
    tt[t] = ft;
    __syncthreads();
    for (int o = 1; o < RS_T; o <<= 1) {
        const int64_t a = t >= o ? h[t - o] : -1;
        const int64_t b = t + o < RS_T ? tt[t + o] : (int64_t)L0;
        __syncthreads();
        h[t] = max(h[t], a);
        tt[t] = min(tt[t], b);
        __syncthreads();
    }
    // tails among this thread's items, and the first tail after them
    uint32_t tmask = 0;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r)
        if (rf_tail(fv[r])) tmask |= 1u << r;
    const int64_t after = min(ctail[blockIdx.x], t + 1 < RS_T ? tt[t + 1] : (int64_t)L0);
    int64_t cur = max(chead[blockIdx.x], t > 0 ? h[t - 1] : (int64_t)-1);
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        if (rf_head(fv[r])) cur = (int64_t)(base + r);
        const uint32_t rest = tmask >> r;
        const int64_t tail = rest ? (int64_t)(base + r) + __ffs(rest) - 1 : after;
        // every run starts with a head and ends with a tail
        const bool ok = (fv[r] & RF_SEED) && cur >= 0 && tail < (int64_t)L0;
        const int li = rs_lds(t * RS_ITEMS + r);
        hs[li] = ok ? (uint32_t)cur : NO_SEED;
        ts[li] = (uint32_t)tail;
    }
    __syncthreads();
    for (int r = 0; r < RS_ITEMS; ++r) {
        const int idx = r * RS_T + t;
        const uint64_t pu = bbase + idx;
        if (pu >= L0) break;
        const uint32_t p0h = hs[rs_lds(idx)];
        if (p0h == NO_SEED) continue;
        const int64_t p0 = (int64_t)pu, p0t = ts[rs_lds(idx)];
        const uint2 q = pp[pu];
        const bool rev = (q.x & 0x80000000u) != 0;
        const int64_t p1f = q.x & 0x7FFFFFFFu;
        const int64_t lh0 = lft[p0h];
        const int64_t left = lh0 + (p0 - (int64_t)p0h);      // this seed's own left extension
        const int64_t L = (p0t + k + (int64_t)rgt[p0t]) - ((int64_t)p0h - lh0);
        SeedRec rec;
        rec.s0 = (int32_t)(p0 + 1);
        rec.e0 = (int32_t)(p0 - left + 1);
        rec.elen = (uint32_t)L;
        if (!rev) {
            rec.s1 = (int32_t)(p1f + 1);
            rec.e1 = (int32_t)(p1f - left + 1);
        } else {
            const int64_t p1 = Lb - p1f - k;
            rec.s1 = (int32_t)(-(p1f + 1));
            rec.e1 = (int32_t)(-(Lb - (p1 - left + L) + 1));
        }
        out[q.y] = rec;
    }
}

// ---- bucket partition of the seeds ----------------------------------------
// AddHashEntry puts a seed in bucket ((offset mod 40000) + 40000) mod 40000
// (LM/MemHash.cpp:213).  Call a bucket clean when all its seeds come from an
// odd solid seed, lie on one strand and share one offset, i.e. one diagonal
// (forward) or one anti-diagonal (reverse: s0 + |s1| + len is constant).
// Then every extension is the maximal exact match through its seed on that
// line, so two extensions are disjoint or equal in genome 0.  On one line,
// Contains (LM/MatchHashEntry.cpp:164-200) reduces to interval containment in
// genome 0 (for the reverse strand its diff_rc == di test holds exactly on
// the anti-diagonal), and strict_start_lessthan_ptr (48-67) decides disjoint
// entries by their genome-0 start, which is positive.  So MheCompare
// (LM/MatchHashEntry.h:121-143) restricted to the bucket orders entries by
// start and calls a seed equivalent exactly to the extension containing it,
// and lower_bound finds that extension whenever it is stored, whatever its
// probe sequence: the first seed of each extension inserts it, every later
// one collides, and the bucket ends up holding the distinct extensions by
// start.  Clean buckets are resolved here by sorting on (bucket, extended
// start) and keeping the first of each run; every other bucket's seeds go to
// the host for the literal replay.

constexpr uint32_t TABLE_SIZE = 40000;   // DEFAULT_MEM_TABLE_SIZE, LM/MemHash.h:30

__device__ __forceinline__ int64_t seed_offset(const SeedRec &r, int k)
{
    return (int64_t)r.s1 - r.s0 - (r.s1 < 0 ? k : 0);   // CalculateOffset, LM/MatchHashEntry.cpp:141-160
}

__device__ __forceinline__ uint32_t bucket_of(int64_t off)
{
    return (uint32_t)(((off % (int64_t)TABLE_SIZE) + TABLE_SIZE) % TABLE_SIZE);
}

__global__ void bucket_key_kernel(const SeedRec *__restrict__ seeds, uint64_t ns, int k, uint64_t *__restrict__ key,
                                  uint32_t *__restrict__ val)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ns) return;
    key[x] = bucket_of(seed_offset(seeds[x], k));
    val[x] = (uint32_t)x;
}

// the seed's line: its offset and strand
__device__ __forceinline__ int64_t seed_line(const SeedRec &r, int k)
{
    return seed_offset(r, k) * 2 + (r.s1 < 0 ? 1 : 0);
}

__global__ void bucket_first_kernel(const uint64_t *__restrict__ skey, const uint32_t *__restrict__ sval,
                                    const SeedRec *__restrict__ seeds, uint64_t ns, int k,
                                    int64_t *__restrict__ first_line)
{
    const uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ns) return;
    const uint64_t b = skey[y];
    if (y == 0 || skey[y - 1] != b) first_line[b] = seed_line(seeds[sval[y]], k);
}

__global__ void bucket_dirty_kernel(const uint64_t *__restrict__ skey, const uint32_t *__restrict__ sval,
                                    const SeedRec *__restrict__ seeds, uint64_t ns, int k, int solid_odd,
                                    const int64_t *__restrict__ first_line, uint32_t *__restrict__ dirty)
{
    const uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ns) return;
    const uint64_t b = skey[y];
    if (!solid_odd || seed_line(seeds[sval[y]], k) != first_line[b]) dirty[b] = 1u;
}

__global__ void key2_kernel(const uint64_t *__restrict__ skey, const uint32_t *__restrict__ sval,
                            const SeedRec *__restrict__ seeds, uint64_t ns, int eb, const uint32_t *__restrict__ dirty,
                            uint64_t *__restrict__ key2, uint32_t *__restrict__ val2, uint32_t *__restrict__ flag_dirty)
{
    const uint64_t y = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= ns) return;
    const uint64_t b = skey[y];
    const uint32_t v = sval[y];
    const bool d = dirty[b] != 0;
    key2[y] = (b << eb) | (d ? ((1ull << eb) - 1) : (uint64_t)(uint32_t)seeds[v].e0);
    val2[y] = v;
    flag_dirty[y] = d ? 1u : 0u;
}

__global__ void distinct_kernel(const uint64_t *__restrict__ key2, uint64_t ns, int eb, const uint32_t *__restrict__ dirty,
                                uint32_t *__restrict__ flag)
{
    const uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= ns) return;
    const uint64_t kz = key2[z];
    flag[z] = (!dirty[kz >> eb] && (z == 0 || key2[z - 1] != kz)) ? 1u : 0u;
}

__global__ __launch_bounds__(JT) void compact_idx_kernel(const uint32_t *__restrict__ flag, uint64_t n,
                                                         const uint32_t *__restrict__ tile_off,
                                                         uint32_t *__restrict__ out_idx)
{
    __shared__ uint32_t sc[JT];
    const int t = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * JTILE + (uint64_t)t * JITEMS;
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
    for (int r = 0; r < JITEMS; ++r)
        if (f[r]) out_idx[o++] = (uint32_t)(base + r);
}

__global__ void gather_clean_kernel(const uint32_t *__restrict__ idx, uint64_t m, int eb, const uint64_t *__restrict__ key2,
                                    const uint32_t *__restrict__ val2, const SeedRec *__restrict__ seeds,
                                    CleanEnt *__restrict__ out)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= m) return;
    const uint32_t z = idx[x];
    const SeedRec r = seeds[val2[z]];
    CleanEnt e;
    e.bucket = (uint32_t)(key2[z] >> eb);
    e.e0 = r.e0;
    e.e1 = r.e1;
    e.elen = r.elen;
    out[x] = e;
}

__global__ void gather_dirty_kernel(const uint32_t *__restrict__ idx, uint64_t m, const uint64_t *__restrict__ skey,
                                    const uint32_t *__restrict__ sval, const SeedRec *__restrict__ seeds,
                                    DirtySeed *__restrict__ out)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= m) return;
    const uint32_t y = idx[x];
    DirtySeed d;
    d.bucket = (uint32_t)skey[y];
    d.r = seeds[sval[y]];
    out[x] = d;
}

}  // namespace

// Seeds in HashMatch order with their extensions; returns the count.
static int compact_flags(mh_context *ctx, const uint32_t *flag, uint64_t n, uint32_t *out_idx, uint32_t *tiles,
                         uint32_t *d_count, uint64_t *count);

// The long walks of the heads in long_idx, per diagonal (the kernels above).
static int resolve_long(mh_context *ctx, const mh_sml *a, const mh_sml *b, const ExtEnv &env,
                        const uint32_t *long_idx, uint64_t nl, const uint32_t *hoi, const uint32_t *hoj,
                        SeedRec *hres)
{
    DevBuf *g = ctx->diag_scratch;
    int rc;
    if ((rc = g[0].ensure(nl * sizeof(DiagHead) + 64)) != MH_OK) return rc;
    DiagHead *d_heads = g[0].as<DiagHead>();
    const dim3 blk(256);
    hipLaunchKernelGGL(dg_heads_kernel, dim3((unsigned)((nl + 255) / 256)), blk, 0, ctx->stream, long_idx, nl, hoi,
                       hoj, a->d_keys, a->d_pos, b->d_keys, b->d_pos, d_heads);
    MH_HIP(hipGetLastError());
    std::vector<DiagHead> hh(nl);
    MH_HIP(hipMemcpyAsync(hh.data(), d_heads, nl * sizeof(DiagHead), hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t L = env.seed_len;
    const int64_t last0 = env.L0 - L + 1, last1 = env.L1 - L + 1;   // last window starts
    auto key = [](const DiagHead &h) {
        const int64_t s1 = h.t1 < 0 ? -(int64_t)h.t1 : h.t1;
        return h.t1 < 0 ? std::make_pair(1, h.s0 + s1) : std::make_pair(0, s1 - h.s0);
    };
    std::sort(hh.begin(), hh.end(), [&](const DiagHead &x, const DiagHead &y) {
        const auto kx = key(x), ky = key(y);
        return kx != ky ? kx < ky : x.s0 < y.s0;
    });
    struct Diag {
        int rev;
        int64_t dconst, glo, ghi, smin, smax, reach;
        size_t h0, h1;
        bool done;
    };
    std::vector<Diag> diags;
    for (size_t i = 0; i < nl;) {
        size_t j = i + 1;
        while (j < nl && key(hh[j]) == key(hh[i])) ++j;
        Diag d;
        d.rev = key(hh[i]).first;
        d.dconst = key(hh[i]).second;
        // window starts p of genome 1 whose partner q is a window start of genome 2
        if (d.rev) {
            d.glo = std::max<int64_t>(1, d.dconst - last1);
            d.ghi = std::min<int64_t>(last0, d.dconst - 1);
        } else {
            d.glo = std::max<int64_t>(1, 1 - d.dconst);
            d.ghi = std::min<int64_t>(last0, last1 - d.dconst);
        }
        d.smin = hh[i].s0;
        d.smax = hh[j - 1].s0;
        MH_CHECK(d.glo <= d.smin && d.smax <= d.ghi, MH_ERR_DEVICE, "seed window outside its sequences");
        d.reach = 16384;
        d.h0 = i;
        d.h1 = j;
        d.done = false;
        diags.push_back(d);
        i = j;
    }
    std::vector<DiagGroup> tab;
    std::vector<DiagHead> hp;
    std::vector<size_t> owner;
    std::vector<uint8_t> esc;
    for (int iter = 0;; ++iter) {
        MH_CHECK(iter < 16, MH_ERR_DEVICE, "long seed extensions did not settle");
        tab.clear();
        hp.clear();
        owner.clear();
        int64_t N = 0, max_n = 1;
        for (size_t k = 0; k < diags.size(); ++k) {
            Diag &d = diags[k];
            if (d.done) continue;
            DiagGroup t;
            t.lo = std::max(d.glo, d.smin - d.reach);
            const int64_t hi = std::min(d.ghi, d.smax + d.reach);
            t.n = hi - t.lo + 1;
            t.base = N;
            t.dconst = d.dconst;
            t.rev = d.rev;
            t.clip_lo = t.lo > d.glo;
            t.clip_hi = hi < d.ghi;
            t.pad = 0;
            N += t.n;
            max_n = std::max(max_n, t.n);
            for (size_t i = d.h0; i < d.h1; ++i) {
                DiagHead h = hh[i];
                h.group = (uint32_t)tab.size();
                hp.push_back(h);
            }
            owner.push_back(k);
            tab.push_back(t);
        }
        if (tab.empty()) break;
        MH_CHECK(N < (int64_t)DG_IDX, MH_ERR_UNSUPPORTED, "long seed extensions span too many windows");
        const uint64_t nh = hp.size(), ng = tab.size();
        if ((rc = g[1].ensure(ng * sizeof(DiagGroup) + 64)) != MH_OK) return rc;
        if ((rc = g[0].ensure(nh * sizeof(DiagHead) + 64)) != MH_OK) return rc;
        if ((rc = g[2].ensure((uint64_t)N + 64)) != MH_OK) return rc;
        for (int q = 3; q < 7; ++q)
            if ((rc = g[q].ensure((uint64_t)N * 4 + 64)) != MH_OK) return rc;
        if ((rc = g[7].ensure(nh + 64)) != MH_OK) return rc;
        DiagGroup *d_tab = g[1].as<DiagGroup>();
        DiagHead *d_hp = g[0].as<DiagHead>();
        uint8_t *mg = g[2].as<uint8_t>(), *d_esc = g[7].as<uint8_t>();
        uint32_t *l0 = g[3].as<uint32_t>(), *r0 = g[4].as<uint32_t>(), *l1 = g[5].as<uint32_t>(),
                 *r1 = g[6].as<uint32_t>();
        MH_HIP(hipMemcpyAsync(d_tab, tab.data(), ng * sizeof(DiagGroup), hipMemcpyHostToDevice, ctx->stream));
        MH_HIP(hipMemcpyAsync(d_hp, hp.data(), nh * sizeof(DiagHead), hipMemcpyHostToDevice, ctx->stream));
        const dim3 grid((unsigned)((N + 255) / 256));
        int rounds = 1;
        while ((1ll << (rounds - 1)) < max_n) ++rounds;
        hipLaunchKernelGGL(dg_match_kernel, grid, blk, 0, ctx->stream, d_tab, (uint32_t)ng, (uint64_t)N, env, mg);
        hipLaunchKernelGGL(dg_link_kernel, grid, blk, 0, ctx->stream, d_tab, (uint32_t)ng, (uint64_t)N, (int)L, mg,
                           l0, r0);
        for (int r = 0; r < rounds; ++r) {
            hipLaunchKernelGGL(dg_jump_kernel, grid, blk, 0, ctx->stream, (uint64_t)N, l0, r0, l1, r1);
            std::swap(l0, l1);
            std::swap(r0, r1);
        }
        hipLaunchKernelGGL(dg_step_kernel, grid, blk, 0, ctx->stream, d_tab, (uint32_t)ng, (uint64_t)N, (int)L, mg,
                           l0, r0, l1, r1);
        std::swap(l0, l1);
        std::swap(r0, r1);
        for (int r = 0; r < rounds; ++r) {
            hipLaunchKernelGGL(dg_jump_or_kernel, grid, blk, 0, ctx->stream, (uint64_t)N, l0, r0, l1, r1);
            std::swap(l0, l1);
            std::swap(r0, r1);
        }
        hipLaunchKernelGGL(dg_result_kernel, dim3((unsigned)((nh + 255) / 256)), blk, 0, ctx->stream, d_hp, nh,
                           d_tab, l0, r0, (int)L, hres, d_esc);
        MH_HIP(hipGetLastError());
        esc.resize(nh);
        MH_HIP(hipMemcpyAsync(esc.data(), d_esc, nh, hipMemcpyDeviceToHost, ctx->stream));
        MH_HIP(hipStreamSynchronize(ctx->stream));
        std::vector<uint8_t> redo(ng, 0);
        for (uint64_t i = 0; i < nh; ++i)
            if (esc[i]) redo[hp[i].group] = 1;
        for (uint64_t k = 0; k < ng; ++k) {
            Diag &d = diags[owner[k]];
            if (redo[k]) d.reach *= 8;
            else d.done = true;
        }
    }
    return MH_OK;
}

// Spaced-seed extension with one ExtendMatch per grid chain (see the kernels
// above): sort the seeds by (diagonal, start), link each to its chain
// predecessor, find chain heads by pointer jumping, extend the heads, copy.
static int extend_by_chains(mh_context *ctx, const mh_sml *a, const mh_sml *b, const uint32_t *oi,
                            const uint32_t *oj, uint64_t ns, const ExtEnv &env, SeedRec *d_seeds, uint32_t *cnt)
{
    DevBuf *g = ctx->grid_scratch;
    int rc;
    for (int q = 0; q < 2; ++q)
        if ((rc = g[q].ensure(ns * 8 + 64)) != MH_OK) return rc;            // keys
    for (int q = 2; q < 8; ++q)
        if ((rc = g[q].ensure(ns * 4 + 64)) != MH_OK) return rc;            // u32 arrays
    if ((rc = g[8].ensure(((ns + JTILE - 1) / JTILE) * 4 + 64)) != MH_OK) return rc;
    if ((rc = g[9].ensure(ns * sizeof(SeedRec) + 64)) != MH_OK) return rc;
    if ((rc = g[10].ensure(ns * 4 + 64)) != MH_OK) return rc;
    int bits = 1;
    while (bits < 32 && (1ull << bits) - 1 < a->len) ++bits;
    const dim3 grid((unsigned)((ns + 255) / 256)), blk(256);
    uint64_t *k0 = g[0].as<uint64_t>(), *k1 = g[1].as<uint64_t>();
    uint32_t *v0 = g[2].as<uint32_t>(), *v1 = g[3].as<uint32_t>();
    int32_t *t1 = g[10].as<int32_t>();
    hipLaunchKernelGGL(grid_key_kernel, grid, blk, 0, ctx->stream, oi, oj, ns, a->d_keys, a->d_pos, b->d_keys,
                       b->d_pos, k0, v0, t1);
    uint64_t *sk = nullptr;
    uint32_t *sv = nullptr;
    if ((rc = radix_sort_kv(ctx, k0, v0, k1, v1, ns, bits, &sk, &sv)) != MH_OK) return rc;
    // the sort's spare value array holds the links, g[4] their jumps
    uint32_t *la = sv == v0 ? v1 : v0, *lb = g[4].as<uint32_t>();
    hipLaunchKernelGGL(grid_pred_kernel, grid, blk, 0, ctx->stream, sk, sv, t1, ns, (uint32_t)env.seed_len, 64, la);
    // links point to earlier sorted entries, so a chain has at most ns seeds and
    // this many doublings reach every head without a check in between
    int rounds = 1;
    while ((1ull << (rounds - 1)) < ns) ++rounds;
    for (int round = 0; round < rounds; ++round) {
        hipLaunchKernelGGL(grid_jump_kernel, grid, blk, 0, ctx->stream, la, lb, ns);
        std::swap(la, lb);
    }
    uint32_t *is_head = g[5].as<uint32_t>(), *head_of = g[6].as<uint32_t>();
    hipLaunchKernelGGL(grid_head_kernel, grid, blk, 0, ctx->stream, sv, la, ns, is_head, head_of);
    // the heads in genome-1 order, so that neighbouring threads walk
    // neighbouring stretches and share cache lines
    uint32_t *heads = g[7].as<uint32_t>();
    hipLaunchKernelGGL(chain_root_flag_kernel, grid, blk, 0, ctx->stream, la, ns, lb);
    uint64_t nh = 0;
    if ((rc = compact_flags(ctx, lb, ns, heads, g[8].as<uint32_t>(), cnt + 3, &nh)) != MH_OK) return rc;
    if (nh)
        hipLaunchKernelGGL(chain_gather_kernel, dim3((unsigned)((nh + 255) / 256)), blk, 0, ctx->stream, sv, nh,
                           heads);
    // head operands reuse the key buffer the sort left spare (8 ns + 64 bytes)
    uint32_t *hoi = reinterpret_cast<uint32_t *>(sk == k0 ? k1 : k0), *hoj = hoi + ns + 16;
    uint32_t *headpos = lb;
    SeedRec *hres = g[9].as<SeedRec>();
    const dim3 hgrid((unsigned)((nh + 255) / 256));
    hipLaunchKernelGGL(grid_gather_kernel, hgrid, blk, 0, ctx->stream, heads, nh, oi, oj, hoi, hoj, headpos);
    // heads: the step-by-step walk up to a budget, the longer ones per diagonal
    uint32_t *long_flag = is_head, *long_idx = heads;   // both free by now
    hipLaunchKernelGGL(seed_kernel, hgrid, blk, 0, ctx->stream, hoi, hoj, nh, a->d_keys, a->d_pos, b->d_keys,
                       b->d_pos, env, (int64_t)LONG_WALK, hres, cnt + 2, long_flag);
    uint64_t nl = 0;
    if ((rc = compact_flags(ctx, long_flag, nh, long_idx, g[8].as<uint32_t>(), cnt + 3, &nl)) != MH_OK) return rc;
    if (nl && (rc = resolve_long(ctx, a, b, env, long_idx, nl, hoi, hoj, hres)) != MH_OK) return rc;
    ctx->last_walks = nh;
    ctx->last_long_walks = nl;
    hipLaunchKernelGGL(grid_fill_kernel, grid, blk, 0, ctx->stream, oi, oj, ns, a->d_keys, a->d_pos, b->d_keys,
                       b->d_pos, head_of, headpos, hres, d_seeds);
    MH_HIP(hipGetLastError());
    return MH_OK;
}

int seed_mers(mh_context *ctx, const mh_sml *a, const SeedRec *d_seeds, uint64_t ns, uint32_t pair,
              uint64_t *d_out)
{
    MH_CHECK(a->seed_weight <= 30 && pair < 8, MH_ERR_UNSUPPORTED,
             "seed weight above 30 with more than two genomes");
    if (!ns) return MH_OK;
    hipLaunchKernelGGL(seed_mer_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream, d_seeds, ns,
                       a->d_keys_pos, 64 - 2 * a->seed_weight, pair, d_out);
    MH_HIP(hipGetLastError());
    return MH_OK;
}

__global__ void gather_seeds_kernel(const SeedRec *__restrict__ in, const uint32_t *__restrict__ idx, uint64_t n,
                                    SeedRec *__restrict__ out)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    out[x] = in[idx[x]];
}

__global__ void iota_kernel(uint32_t *__restrict__ v, uint64_t n)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) v[x] = (uint32_t)x;
}

int merge_pair_seeds(mh_context *ctx, const SeedRec *d_seeds, uint64_t *d_keys, uint64_t n, SeedRec **d_seeds_out,
                     uint64_t **d_keys_out)
{
    DevBuf *m = ctx->multi_scratch;
    int rc;
    if ((rc = m[2].ensure(n * 8 + 64)) != MH_OK) return rc;
    if ((rc = m[3].ensure(n * 4 + 64)) != MH_OK) return rc;
    if ((rc = m[4].ensure(n * 4 + 64)) != MH_OK) return rc;
    if ((rc = m[5].ensure(n * sizeof(SeedRec) + 64)) != MH_OK) return rc;
    const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    hipLaunchKernelGGL(iota_kernel, grid, blk, 0, ctx->stream, m[3].as<uint32_t>(), n);
    uint64_t *sk = nullptr;
    uint32_t *sv = nullptr;
    if ((rc = radix_sort_kv(ctx, d_keys, m[3].as<uint32_t>(), m[2].as<uint64_t>(), m[4].as<uint32_t>(), n, 64, &sk,
                            &sv)) != MH_OK)
        return rc;
    hipLaunchKernelGGL(gather_seeds_kernel, grid, blk, 0, ctx->stream, d_seeds, sv, n, m[5].as<SeedRec>());
    MH_HIP(hipGetLastError());
    *d_seeds_out = m[5].as<SeedRec>();
    *d_keys_out = sk;
    return MH_OK;
}

static int compact_flags(mh_context *ctx, const uint32_t *flag, uint64_t n, uint32_t *out_idx, uint32_t *tiles,
                         uint32_t *d_count, uint64_t *count);

__device__ __forceinline__ uint32_t bucket_of_seed(const SeedRec &r, int sl)
{
    const int64_t off = (int64_t)r.s1 - r.s0 - (r.s1 < 0 ? sl : 0);   // CalculateOffset, LM/MatchHashEntry.cpp:141-160
    return (uint32_t)(((off % 40000) + 40000) % 40000);
}

__global__ void bucket_mask_kernel(const SeedRec *__restrict__ seeds, uint64_t n, const uint8_t *__restrict__ mask,
                                   int sl, uint32_t *__restrict__ flag)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    flag[x] = mask[bucket_of_seed(seeds[x], sl)];
}

__global__ void gather_marked_kernel(const SeedRec *__restrict__ seeds, const uint64_t *__restrict__ keys,
                                     const uint32_t *__restrict__ idx, uint64_t m, SeedRec *__restrict__ out,
                                     uint8_t *__restrict__ pair)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= m) return;
    const uint32_t i = idx[x];
    out[x] = seeds[i];
    pair[x] = (uint8_t)(keys[i] & 7u);
}

int select_bucket_seeds(mh_context *ctx, const SeedRec *d_seeds, const uint64_t *d_keys, uint64_t n,
                        const uint8_t *mask, int seed_len, std::vector<SeedRec> &seeds, std::vector<uint8_t> &pairs)
{
    seeds.clear();
    pairs.clear();
    if (!n) return MH_OK;
    DevBuf *m = ctx->multi_scratch;
    int rc;
    if ((rc = m[6].ensure(40000 + 64)) != MH_OK) return rc;
    if ((rc = m[7].ensure(n * 4 + 64)) != MH_OK) return rc;
    if ((rc = m[8].ensure(n * 4 + 64)) != MH_OK) return rc;
    if ((rc = m[9].ensure(((n + JTILE - 1) / JTILE) * 4 + 128)) != MH_OK) return rc;
    MH_HIP(hipMemcpyAsync(m[6].ptr, mask, 40000, hipMemcpyHostToDevice, ctx->stream));
    const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    hipLaunchKernelGGL(bucket_mask_kernel, grid, blk, 0, ctx->stream, d_seeds, n, m[6].as<uint8_t>(), seed_len,
                       m[7].as<uint32_t>());
    uint32_t *tiles = m[9].as<uint32_t>();
    uint32_t *d_count = tiles + (n + JTILE - 1) / JTILE + 8;
    uint64_t cnt = 0;
    if ((rc = compact_flags(ctx, m[7].as<uint32_t>(), n, m[8].as<uint32_t>(), tiles, d_count, &cnt)) != MH_OK)
        return rc;
    if (!cnt) return MH_OK;
    if ((rc = m[0].ensure(cnt * sizeof(SeedRec) + 64)) != MH_OK) return rc;
    if ((rc = m[10].ensure(cnt + 64)) != MH_OK) return rc;
    hipLaunchKernelGGL(gather_marked_kernel, dim3((unsigned)((cnt + 255) / 256)), blk, 0, ctx->stream, d_seeds,
                       d_keys, m[8].as<uint32_t>(), cnt, m[0].as<SeedRec>(), m[10].as<uint8_t>());
    MH_HIP(hipGetLastError());
    seeds.resize(cnt);
    pairs.resize(cnt);
    MH_HIP(hipMemcpyAsync(seeds.data(), m[0].ptr, cnt * sizeof(SeedRec), hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipMemcpyAsync(pairs.data(), m[10].ptr, cnt, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int search_live_ranges(mh_context *ctx, const mh_sml *const *smls, int k, LiveRanges *out)
{
    MH_CHECK(k >= 1 && k <= PLAN_MAXK, MH_ERR_INVALID, "list count out of range");
    out->restarts = 0;
    for (int g = 0; g < k; ++g) {
        out->lo[g].assign(1, 0);
        out->hi[g].assign(1, smls[g]->n);
        out->ep[g].assign(1, 0);
    }
    KLists L;
    std::memset(&L, 0, sizeof(L));
    L.k = k;
    for (int i = 0; i < k; ++i) {
        L.ck[i] = smls[i]->d_keys;
        L.n[i] = smls[i]->n;
    }
    const uint32_t thresh = (uint32_t)(MER_REPEAT_LIMIT / (uint32_t)k + 1);
    int rc;
    uint32_t cap = 4096, count = 0;
    for (;;) {
        if ((rc = ctx->heavy_buf.ensure((size_t)cap * 8 + 64)) != MH_OK) return rc;
        uint64_t *d_out = ctx->heavy_buf.as<uint64_t>();
        uint32_t *d_cnt = reinterpret_cast<uint32_t *>(d_out + cap);
        MH_HIP(hipMemsetAsync(d_cnt, 0, 4, ctx->stream));
        for (int i = 0; i < k; ++i)
            if (L.n[i])
                hipLaunchKernelGGL(heavy_kernel, dim3((unsigned)((L.n[i] + 255) / 256)), dim3(256), 0, ctx->stream,
                                   L, i, thresh, d_out, cap, d_cnt);
        MH_HIP(hipGetLastError());
        MH_HIP(hipMemcpyAsync(&count, d_cnt, 4, hipMemcpyDeviceToHost, ctx->stream));
        MH_HIP(hipStreamSynchronize(ctx->stream));
        if (count <= cap) break;
        cap = count;
    }
    if (count == 0) return MH_OK;
    std::vector<uint64_t> heavy(count);
    MH_HIP(hipMemcpyAsync(heavy.data(), ctx->heavy_buf.ptr, (size_t)count * 8, hipMemcpyDeviceToHost, ctx->stream));
    // the sorted keys on the host, for the planner's searches
    std::vector<std::vector<uint64_t>> keys(k);
    PlanLists P;
    P.k = k;
    P.sw = smls[0]->seed_weight;
    P.sl = smls[0]->seed_length;
    for (int g = 0; g < k; ++g) {
        keys[g].resize(smls[g]->n);
        if (smls[g]->n)
            MH_HIP(hipMemcpyAsync(keys[g].data(), smls[g]->d_keys, smls[g]->n * 8, hipMemcpyDeviceToHost,
                                  ctx->stream));
        P.ck[g] = keys[g].data();
        P.n[g] = smls[g]->n;
        P.len[g] = smls[g]->len;
    }
    MH_HIP(hipStreamSynchronize(ctx->stream));
    std::sort(heavy.begin(), heavy.end());
    heavy.erase(std::unique(heavy.begin(), heavy.end()), heavy.end());
    return plan_restarts(P, heavy, out);
}

int join_and_extend(mh_context *ctx, const mh_sml *a, const mh_sml *b, const LiveRanges *live, int ga, int gb,
                    const SeedRec **seeds_out, uint64_t *ns_out)
{
    const uint64_t n0 = a->n, n1 = b->n;
    *seeds_out = nullptr;
    *ns_out = 0;
    ctx->last_walks = ctx->last_long_walks = 0;
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
    uint32_t *cnt = ctx->mf_scratch[3].as<uint32_t>();   // [1] total seeds, [2] ext errors
    // live ranges other than the whole list go to the device
    LivePair lp;
    std::memset(&lp, 0, sizeof(lp));
    const bool cut_a = live && !(live->lo[ga].size() == 1 && live->lo[ga][0] == 0 && live->hi[ga][0] == n0);
    const bool cut_b = live && !(live->lo[gb].size() == 1 && live->lo[gb][0] == 0 && live->hi[gb][0] == n1);
    if (cut_a || cut_b) {
        const size_t ca = cut_a ? live->lo[ga].size() : 0, cb = cut_b ? live->lo[gb].size() : 0;
        std::vector<uint64_t> h(2 * ca + 2 * cb + 2, 0);
        for (size_t t = 0; t < ca; ++t) {
            h[t] = live->lo[ga][t];
            h[ca + t] = live->hi[ga][t];
        }
        for (size_t t = 0; t < cb; ++t) {
            h[2 * ca + t] = live->lo[gb][t];
            h[2 * ca + cb + t] = live->hi[gb][t];
        }
        if ((rc = ctx->live_buf.ensure(h.size() * 8 + 64)) != MH_OK) return rc;
        uint64_t *d = ctx->live_buf.as<uint64_t>();
        MH_HIP(hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        MH_HIP(hipStreamSynchronize(ctx->stream));
        lp.lo0 = d;
        lp.hi0 = d + ca;
        lp.lo1 = d + 2 * ca;
        lp.hi1 = d + 2 * ca + cb;
        lp.c0 = (uint32_t)ca;
        lp.c1 = (uint32_t)cb;
        // a list that reads nothing gives no seeds
        if ((cut_a && ca == 0) || (cut_b && cb == 0)) return MH_OK;
    }
    MH_HIP(hipMemsetAsync(cnt, 0, 16, ctx->stream));
    const bool tiled = !cut_a && !cut_b;
    const uint64_t ntj = (n0 + TJ_I - 1) / TJ_I;
    if (tiled) {
        // every entry live: the LDS-tiled join (count pass here, write pass below)
        if ((rc = ctx->mf_scratch[2].ensure(ntj * 4 + 64)) != MH_OK) return rc;
        if ((rc = ctx->mf_scratch[1].ensure(ntj * 16 + 64)) != MH_OK) return rc;
        tiles = ctx->mf_scratch[2].as<uint32_t>();
        TimedLaunch tl(ctx, "join");
        uint64_t *jb = ctx->mf_scratch[1].as<uint64_t>();
        hipLaunchKernelGGL(tile_bounds_kernel, dim3((unsigned)((ntj + 255) / 256)), dim3(256), 0, ctx->stream,
                           a->d_keys, n0, b->d_keys, n1, ntj, jb);
        hipLaunchKernelGGL(tile_join_kernel, dim3((unsigned)ntj), dim3(TJ_T), 0, ctx->stream, a->d_keys, n0,
                           b->d_keys, n1, (const uint64_t *)jb, 0, tiles, (const uint32_t *)nullptr,
                           (uint32_t *)nullptr, (uint32_t *)nullptr);
        hipLaunchKernelGGL(excl_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, tiles, ntj, cnt + 1);
        MH_HIP(hipGetLastError());
    } else {
        TimedLaunch tl(ctx, "join");
        hipLaunchKernelGGL(join_kernel, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, ctx->stream, a->d_keys, n0,
                           b->d_keys, n1, lp, flag, partner);
        hipLaunchKernelGGL(tile_count_kernel, dim3((unsigned)ntiles), dim3(JT), 0, ctx->stream, flag, n0, tiles);
        hipLaunchKernelGGL(excl_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, tiles, ntiles, cnt + 1);
        MH_HIP(hipGetLastError());
    }
    uint32_t h[2] = {0, 0};
    MH_HIP(hipMemcpyAsync(h, cnt, 8, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    const uint64_t ns = h[1];
    if (ns == 0) return MH_OK;
    if ((rc = ctx->mf_scratch[4].ensure(ns * 4 + 64)) != MH_OK) return rc;
    if ((rc = ctx->mf_scratch[5].ensure(ns * 4 + 64)) != MH_OK) return rc;
    if ((rc = ctx->mf_scratch[6].ensure(ns * sizeof(SeedRec) + 64)) != MH_OK) return rc;
    uint32_t *oi = ctx->mf_scratch[4].as<uint32_t>();
    uint32_t *oj = ctx->mf_scratch[5].as<uint32_t>();
    SeedRec *d_seeds = ctx->mf_scratch[6].as<SeedRec>();
    if (tiled) {
        TimedLaunch tl(ctx, "join_write");
        hipLaunchKernelGGL(tile_join_kernel, dim3((unsigned)ntj), dim3(TJ_T), 0, ctx->stream, a->d_keys, n0,
                           b->d_keys, n1, (const uint64_t *)ctx->mf_scratch[1].as<uint64_t>(), 1,
                           (uint32_t *)nullptr, (const uint32_t *)tiles, oi, oj);
        MH_HIP(hipGetLastError());
    } else {
        hipLaunchKernelGGL(compact_kernel, dim3((unsigned)ntiles), dim3(JT), 0, ctx->stream, flag, partner, n0, tiles,
                           oi, oj);
    }
    ExtEnv env;
    env.kp0 = a->d_keys_pos;
    env.kp1 = b->d_keys_pos;
    env.n0 = n0;
    env.n1 = n1;
    env.L0 = (int64_t)a->len;
    env.L1 = (int64_t)b->len;
    env.seed_len = a->seed_length;
    env.mer_mask = a->seed_weight >= 32 ? ~0ull : (~0ull << (64 - 2 * a->seed_weight));
    const bool solid_odd = a->seed_length == a->seed_weight && (a->seed_length & 1);
    if (solid_odd) {
        if ((rc = ctx->mf_scratch[7].ensure(b->n_words * 4 + 64)) != MH_OK) return rc;
        uint32_t *wbrc = ctx->mf_scratch[7].as<uint32_t>();
        // run scratch, indexed by genome-0 position: partner with strand, left
        // and right ends of runs, flags, extensions; block carries
        const uint64_t L0 = a->len, nb = (L0 + RS_P - 1) / RS_P;
        DevBuf *rs = ctx->run_scratch;
        if ((rc = rs[0].ensure(L0 * 4 * 4 + 64)) != MH_OK) return rc;
        if ((rc = rs[1].ensure(L0 + 64)) != MH_OK) return rc;
        if ((rc = rs[2].ensure(nb * 8 * 4 + 64)) != MH_OK) return rc;
        uint2 *pp = rs[0].as<uint2>();
        uint32_t *lft = (uint32_t *)(pp + L0), *rgt = lft + L0;
        uint8_t *fl = rs[1].as<uint8_t>();
        int64_t *bhead = rs[2].as<int64_t>(), *btail = bhead + nb, *chead = btail + nb, *ctail = chead + nb;
        const unsigned gs = (unsigned)((ns + 255) / 256), gp = (unsigned)((L0 + 255) / 256);
        TimedLaunch tl(ctx, "extend");
        hipLaunchKernelGGL(rc_pack_kernel, dim3((unsigned)((b->n_words + 255) / 256)), dim3(256), 0, ctx->stream,
                           b->d_words, b->len, b->n_words, wbrc);
        MH_HIP(hipMemsetAsync(pp, 0xFF, L0 * 8, ctx->stream));
        hipLaunchKernelGGL(run_scatter_kernel, dim3(gs), dim3(256), 0, ctx->stream, oi, oj, ns, a->d_keys, a->d_pos,
                           b->d_keys, b->d_pos, pp);
        hipLaunchKernelGGL(run_flags_kernel, dim3(gp), dim3(256), 0, ctx->stream, pp, L0, fl);
        hipLaunchKernelGGL(run_block_kernel, dim3((unsigned)nb), dim3(RS_T), 0, ctx->stream, fl, L0, bhead, btail);
        hipLaunchKernelGGL(run_carry_kernel, dim3(1), dim3(1024), 0, ctx->stream, bhead, btail, nb, L0, chead, ctail);
        hipLaunchKernelGGL(run_walk_kernel, dim3(gp), dim3(256), 0, ctx->stream, fl, pp, L0, a->d_words, b->d_words,
                           wbrc, (int64_t)a->len, (int64_t)b->len, a->seed_length, lft, rgt);
        hipLaunchKernelGGL(run_final_kernel, dim3((unsigned)nb), dim3(RS_T), 0, ctx->stream, fl, pp, L0, chead,
                           ctail, lft, rgt, (int64_t)b->len, a->seed_length, d_seeds);
        MH_HIP(hipGetLastError());
        ctx->last_walks = ns;
    } else if (a->len < (1ull << 30) && b->len < (1ull << 30)) {
        TimedLaunch tl(ctx, "extend");
        if ((rc = extend_by_chains(ctx, a, b, oi, oj, ns, env, d_seeds, cnt)) != MH_OK) return rc;
    } else {
        TimedLaunch tl(ctx, "extend");
        hipLaunchKernelGGL(seed_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream, oi, oj, ns,
                           a->d_keys, a->d_pos, b->d_keys, b->d_pos, env, (int64_t)0, d_seeds, cnt + 2,
                           (uint32_t *)nullptr);
        MH_HIP(hipGetLastError());
        ctx->last_walks = ns;
    }
    uint32_t ext_err = 0;
    MH_HIP(hipMemcpyAsync(&ext_err, cnt + 2, 4, hipMemcpyDeviceToHost, ctx->stream));
    const double t0 = now_ms();
    MH_HIP(hipStreamSynchronize(ctx->stream));
    host_time(ctx, "host_wait_seeds", now_ms() - t0);
    *seeds_out = d_seeds;
    *ns_out = ns;
    MH_CHECK(ext_err == 0, MH_ERR_UNSUPPORTED, "seed extension left the sorted mer lists");
    return MH_OK;
}


// flags -> compacted indices in order; returns the count (synchronises)
static int compact_flags(mh_context *ctx, const uint32_t *flag, uint64_t n, uint32_t *out_idx, uint32_t *tiles,
                         uint32_t *d_count, uint64_t *count)
{
    const uint64_t ntiles = (n + JTILE - 1) / JTILE;
    hipLaunchKernelGGL(tile_count_kernel, dim3((unsigned)ntiles), dim3(JT), 0, ctx->stream, flag, n, tiles);
    hipLaunchKernelGGL(excl_scan_kernel, dim3(1), dim3(1024), 0, ctx->stream, tiles, ntiles, d_count);
    hipLaunchKernelGGL(compact_idx_kernel, dim3((unsigned)ntiles), dim3(JT), 0, ctx->stream, flag, n, tiles, out_idx);
    MH_HIP(hipGetLastError());
    uint32_t h = 0;
    MH_HIP(hipMemcpyAsync(&h, d_count, 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    *count = h;
    return MH_OK;
}

int partition_seeds(mh_context *ctx, const SeedRec *d_seeds, uint64_t ns, int k, bool solid_odd, uint64_t len_a,
                    std::vector<CleanEnt> &clean, std::vector<DirtySeed> &dirty, uint64_t *n_clean_seeds)
{
    // the second sort key is (bucket << eb) | extended start, eb bits holding
    // every start (1-based, <= len_a) and the all-ones mark of other buckets
    int eb = 1;
    while (eb < 32 && (1ull << eb) - 1 <= len_a) ++eb;
    clean.clear();
    dirty.clear();
    *n_clean_seeds = 0;
    if (ns == 0) return MH_OK;
    const double t0 = now_ms();
    int rc;
    // scratch: 8..15 of mf_scratch
    DevBuf *m = ctx->mf_scratch;
    if ((rc = m[8].ensure(ns * 8 + 64)) != MH_OK) return rc;     // key A
    if ((rc = m[9].ensure(ns * 8 + 64)) != MH_OK) return rc;     // key B
    if ((rc = m[10].ensure(ns * 4 + 64)) != MH_OK) return rc;    // val A
    if ((rc = m[11].ensure(ns * 4 + 64)) != MH_OK) return rc;    // val B
    if ((rc = m[12].ensure(TABLE_SIZE * 12 + 64)) != MH_OK) return rc;   // first_off + dirty
    if ((rc = m[13].ensure(ns * 4 + 64)) != MH_OK) return rc;    // flags
    if ((rc = m[14].ensure(ns * 4 + 64)) != MH_OK) return rc;    // compacted indices
    if ((rc = m[15].ensure(((ns + JTILE - 1) / JTILE) * 4 + 64 + ns * (sizeof(DirtySeed) + 16))) != MH_OK) return rc;
    uint64_t *kA = m[8].as<uint64_t>(), *kB = m[9].as<uint64_t>();
    uint32_t *vA = m[10].as<uint32_t>(), *vB = m[11].as<uint32_t>();
    int64_t *first_off = m[12].as<int64_t>();
    uint32_t *dirty_b = reinterpret_cast<uint32_t *>(first_off + TABLE_SIZE);
    uint32_t *flag = m[13].as<uint32_t>(), *idx = m[14].as<uint32_t>();
    uint32_t *tiles = m[15].as<uint32_t>();
    uint32_t *d_count = tiles + ((ns + JTILE - 1) / JTILE) + 4;
    void *gather = reinterpret_cast<char *>(m[15].ptr) + (((ns + JTILE - 1) / JTILE) * 4 + 64);
    const unsigned grid = (unsigned)((ns + 255) / 256);
    MH_HIP(hipMemsetAsync(dirty_b, 0, TABLE_SIZE * 4, ctx->stream));
    uint64_t *sk = nullptr, *s2 = nullptr;
    uint32_t *sv = nullptr, *v2 = nullptr;
    {
        TimedLaunch tl(ctx, "partition");
        hipLaunchKernelGGL(bucket_key_kernel, dim3(grid), dim3(256), 0, ctx->stream, d_seeds, ns, k, kA, vA);
        if ((rc = radix_sort_kv(ctx, kA, vA, kB, vB, ns, 16, &sk, &sv)) != MH_OK) return rc;
        hipLaunchKernelGGL(bucket_first_kernel, dim3(grid), dim3(256), 0, ctx->stream, sk, sv, d_seeds, ns, k,
                           first_off);
        hipLaunchKernelGGL(bucket_dirty_kernel, dim3(grid), dim3(256), 0, ctx->stream, sk, sv, d_seeds, ns, k,
                           solid_odd ? 1 : 0, first_off, dirty_b);
        // (sk, sv) now hold the bucket order; the other pair is free for the second sort
        uint64_t *k2a = (sk == kA) ? kB : kA;
        uint32_t *v2a = (sv == vA) ? vB : vA;
        hipLaunchKernelGGL(key2_kernel, dim3(grid), dim3(256), 0, ctx->stream, sk, sv, d_seeds, ns, eb, dirty_b, k2a,
                           v2a, flag);
        MH_HIP(hipGetLastError());
        // dirty seeds in (bucket, call) order
        uint64_t nd = 0;
        if ((rc = compact_flags(ctx, flag, ns, idx, tiles, d_count, &nd)) != MH_OK) return rc;
        DirtySeed *d_dirty = static_cast<DirtySeed *>(gather);
        if (nd) {
            hipLaunchKernelGGL(gather_dirty_kernel, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, ctx->stream, idx,
                               nd, sk, sv, d_seeds, d_dirty);
            dirty.resize(nd);
            MH_HIP(hipMemcpyAsync(dirty.data(), d_dirty, nd * sizeof(DirtySeed), hipMemcpyDeviceToHost, ctx->stream));
            MH_HIP(hipStreamSynchronize(ctx->stream));
        }
        *n_clean_seeds = ns - nd;
        if (nd < ns) {
            // second sort: (bucket, extended start); the bucket-order pair is scratch now
            uint64_t *k2b = (k2a == kA) ? kB : kA;
            uint32_t *v2b = (v2a == vA) ? vB : vA;
            if ((rc = radix_sort_kv(ctx, k2a, v2a, k2b, v2b, ns, 16 + eb, &s2, &v2)) != MH_OK) return rc;
            hipLaunchKernelGGL(distinct_kernel, dim3(grid), dim3(256), 0, ctx->stream, s2, ns, eb, dirty_b, flag);
            uint64_t nc = 0;
            if ((rc = compact_flags(ctx, flag, ns, idx, tiles, d_count, &nc)) != MH_OK) return rc;
            CleanEnt *d_clean = static_cast<CleanEnt *>(gather);
            if (nc) {
                hipLaunchKernelGGL(gather_clean_kernel, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, ctx->stream,
                                   idx, nc, eb, s2, v2, d_seeds, d_clean);
                clean.resize(nc);
                MH_HIP(hipMemcpyAsync(clean.data(), d_clean, nc * sizeof(CleanEnt), hipMemcpyDeviceToHost,
                                      ctx->stream));
                MH_HIP(hipStreamSynchronize(ctx->stream));
            }
        }
    }
    host_time(ctx, "host_partition", now_ms() - t0);
    return MH_OK;
}

// ---- UniqueMatchFinder seeds (k genomes) ----------------------------------
// UniqueMatchFinder::EnumerateMatches (MA/UniqueMatchFinder.cpp:36-60) hashes
// each run of one masked mer once, with every genome that holds the mer
// exactly once, when there are at least two.  The live entries of all k
// lists are merged into one order by a radix sort on (search, masked mer,
// genome) -- the search being the number of repeat-limit restarts before the
// range that holds the entry, since SearchRange's runs never span a restart
// -- so that a run's entries are adjacent and grouped by genome.  An entry is
// its genome's only one in the run when its neighbours hold other keys; the
// runs with two or more such entries become seeds, in mer (HashMatch) order.
// Each seed is then set up as HashMatch does (SetDirection, CalculateOffset
// with the seed length) and extended over its genomes (ExtendMatch).

struct UqLists {
    const uint64_t *ck[EXT_MAXK];
    const uint32_t *pos[EXT_MAXK];
    const uint32_t *w[EXT_MAXK];      // packed sequence
    const uint32_t *wrc[EXT_MAXK];    // packed reverse complement
    uint64_t off[EXT_MAXK];           // each list's first entry in the merged numbering
};

__global__ void uq_gather_kernel(const uint64_t *__restrict__ ck, uint64_t lo, uint64_t cnt, uint64_t tag,
                                 uint64_t gbase, uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const uint64_t i = lo + t;
    key[t] = tag | ((ck[i] >> 1) << 4);
    val[t] = (uint32_t)(gbase + i);
}

// flag[e]: entry e is the only one of its genome in its run
__global__ void uq_mark_kernel(const uint64_t *__restrict__ key, uint64_t n, uint32_t *__restrict__ flag)
{
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint64_t k = key[e];
    flag[e] = (e == 0 || key[e - 1] != k) && (e + 1 == n || key[e + 1] != k);
}

// over the genome-only entries U (in order): flag the first entry of each run
// that holds at least two of them
__global__ void uq_leader_kernel(const uint64_t *__restrict__ key, const uint32_t *__restrict__ U, uint64_t m,
                                 uint32_t *__restrict__ flag)
{
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= m) return;
    const uint64_t r = key[U[u]] >> 4;
    const bool first = u == 0 || (key[U[u - 1]] >> 4) != r;
    const bool more = u + 1 < m && (key[U[u + 1]] >> 4) == r;
    flag[u] = first && more;
}

__global__ void uq_seed_kernel(const uint64_t *__restrict__ key, const uint32_t *__restrict__ val,
                               const uint32_t *__restrict__ U, uint64_t m, const uint32_t *__restrict__ lead,
                               uint64_t ns, UqLists ls, ExtEnvK env, int k, int solid_odd, int walk,
                               int32_t *__restrict__ seed_st, int32_t *__restrict__ ext_st,
                               uint32_t *__restrict__ ext_len, int64_t *__restrict__ offset,
                               uint32_t *__restrict__ err)
{
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int sl = env.seed_len;
    int gs[EXT_MAXK];
    int64_t st[EXT_MAXK], p[EXT_MAXK];
    int used = 0, fstrand = 0;
    const uint64_t r = key[U[lead[s]]] >> 4;
    for (uint64_t u = lead[s]; u < m && used < EXT_MAXK; ++u) {
        const uint32_t e = U[u];
        const uint64_t kk = key[e];
        if ((kk >> 4) != r) break;
        const int g = (int)(kk & 15u);
        const uint64_t i = val[e] - ls.off[g];
        const int64_t q = ls.pos[g][i];
        const int strand = (int)(ls.ck[g][i] & 1ull);
        if (used == 0) fstrand = strand;
        gs[used] = g;
        p[used] = q;
        // SetDirection, LM/MemHash.cpp:167-203: a genome whose window is on the
        // other strand than the first genome's is reversed
        st[used] = strand == fstrand ? q + 1 : -(q + 1);
        ++used;
    }
    int64_t off = 0;   // CalculateOffset, LM/MatchHashEntry.cpp:141-160, with the seed length
    for (int x = 1; x < used; ++x) off += st[x] - st[0] - (st[x] < 0 ? (int64_t)sl : 0);
    for (int g = 0; g < k; ++g) seed_st[s * k + g] = 0;
    for (int x = 0; x < used; ++x) seed_st[s * k + gs[x]] = (int32_t)st[x];
    offset[s] = off;
    int64_t len = sl;
    int64_t ex[EXT_MAXK];
    if (solid_odd) {
        // the shortest of the pairwise maximal exact matches with the first genome
        const int f = gs[0];
        const int64_t pf = p[0], Lf = env.L[f];
        int64_t left = pf, right = Lf - (pf + sl);
        int64_t q1[EXT_MAXK];
        for (int x = 1; x < used; ++x) {
            const int g = gs[x];
            const bool rev = st[x] < 0;
            const int64_t Lg = env.L[g];
            q1[x] = rev ? Lg - p[x] - sl : p[x];
            const uint32_t *wg = rev ? ls.wrc[g] : ls.w[g];
            const int64_t lim_l = left < q1[x] ? left : q1[x];
            const int64_t lg = match_left(ls.w[f], pf, wg, q1[x], lim_l);
            if (lg < left) left = lg;
            const int64_t rr = Lg - (q1[x] + sl);
            const int64_t lim_r = right < rr ? right : rr;
            const int64_t rg = match_right(ls.w[f], pf + sl, wg, q1[x] + sl, lim_r);
            if (rg < right) right = rg;
        }
        len = sl + left + right;
        ex[0] = pf - left + 1;
        for (int x = 1; x < used; ++x)
            ex[x] = st[x] > 0 ? q1[x] - left + 1 : -(env.L[gs[x]] - (q1[x] - left + len) + 1);
    } else {
        if (!walk) return;   // extended by chains (uq_chain_*)
        for (int x = 0; x < used; ++x) ex[x] = st[x];
        if (extend_k(env, gs, used, ex, len) != EXT_OK) atomicOr(err, 1u);
    }
    for (int g = 0; g < k; ++g) ext_st[s * k + g] = 0;
    for (int x = 0; x < used; ++x) ext_st[s * k + gs[x]] = (int32_t)ex[x];
    ext_len[s] = (uint32_t)len;
}

// ---- spaced seeds: one k-genome walk per chain ---------------------------
// As for two genomes (grid_* kernels): seed B and seed A with the same
// genomes, where every start of A is B's minus the seed length L (forward:
// the window L to the left; reversed: -(q + 1) - L, the window L to the right
// in that genome's forward coordinates), extend to the same match.  B's first
// seed-length step to the left lands on A's windows, which match (A is a
// seed), and from there its walk is A's; A's first step to the right lands
// on B's windows the same way.  Each chain is walked once, from its first
// seed (its head), and the others copy the head's extension.

// key: the first present genome and its (positive) start; a genome position
// holds one mer, so no two seeds share a key
__global__ void uq_chain_key_kernel(const int32_t *__restrict__ seed_st, uint64_t ns, int k,
                                    uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    int g = 0;
    while (g < k - 1 && seed_st[s * k + g] == 0) ++g;
    key[s] = ((uint64_t)g << 32) | (uint32_t)seed_st[s * k + g];
    val[s] = (uint32_t)s;
}

// link[r] = the sorted index of seed r's chain predecessor, or r
__global__ void uq_chain_pred_kernel(const uint64_t *__restrict__ sk, const uint32_t *__restrict__ sv, uint64_t ns,
                                     const int32_t *__restrict__ seed_st, int k, uint32_t seed_len,
                                     uint32_t *__restrict__ link)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= ns) return;
    const uint64_t key = sk[r];
    uint32_t out = (uint32_t)r;
    if ((key & 0xFFFFFFFFull) > seed_len) {
        const uint64_t want = key - seed_len;
        uint64_t lo = 0, hi = r;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) / 2;
            if (sk[mid] < want) lo = mid + 1;
            else hi = mid;
        }
        if (lo < r && sk[lo] == want) {
            const int32_t *b = seed_st + (uint64_t)sv[r] * k, *a = seed_st + (uint64_t)sv[lo] * k;
            bool same = true;
            for (int g = 0; g < k && same; ++g)
                same = b[g] == 0 ? a[g] == 0 : (a[g] == b[g] - (int32_t)seed_len && (b[g] < 0 || a[g] > 0));
            if (same) out = (uint32_t)lo;
        }
    }
    link[r] = out;
}

// the heads' walks (extend_kt: a seed holds at most KMAX genomes)
template <int KMAX>
__global__ __launch_bounds__(128) void uq_chain_walk_kernel(const uint32_t *__restrict__ heads, uint64_t nh, ExtEnvK env, int k,
                                     const int32_t *__restrict__ seed_st, int32_t *__restrict__ ext_st,
                                     uint32_t *__restrict__ ext_len, uint32_t *__restrict__ err)
{
    const uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const uint64_t s = heads[h];
    const uint64_t *gk[KMAX];
    uint32_t gn[KMAX];
    int64_t gL[KMAX], ex[KMAX];
    int gsl[KMAX];
    int used = 0;
    for (int g = 0; g < k; ++g) {
        const int32_t v = seed_st[s * k + g];
        if (v == 0) continue;
#pragma unroll
        for (int x = 0; x < KMAX; ++x)
            if (x == used) {
                gk[x] = env.kp[g];
                gn[x] = (uint32_t)env.n[g];
                gL[x] = env.L[g];
                ex[x] = v;
                gsl[x] = g;
            }
        ++used;
    }
#pragma unroll
    for (int x = 1; x < KMAX; ++x)
        if (x >= used) {
            gk[x] = gk[0];
            gn[x] = gn[0];
            gL[x] = 0;
            ex[x] = 0;
            gsl[x] = 0;
        }
    int64_t len = env.seed_len;
    if (extend_kt<KMAX>(gk, gn, gL, used, env.seed_len, env.mer_mask, ex, len) != EXT_OK) atomicOr(err, 1u);
    for (int g = 0; g < k; ++g) ext_st[s * k + g] = 0;
#pragma unroll
    for (int x = 0; x < KMAX; ++x)
        if (x < used) ext_st[s * k + gsl[x]] = (int32_t)ex[x];
    ext_len[s] = (uint32_t)len;
}

// every other seed copies its head's extension
__global__ void uq_chain_fill_kernel(const uint32_t *__restrict__ is_head, const uint32_t *__restrict__ head_of,
                                     uint64_t ns, int k, int32_t *__restrict__ ext_st, uint32_t *__restrict__ ext_len)
{
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns || is_head[s]) return;
    const uint64_t h = head_of[s];
    for (int g = 0; g < k; ++g) ext_st[s * k + g] = ext_st[h * k + g];
    ext_len[s] = ext_len[h];
}

// ---- AddHashEntry of the unique seeds, one bucket per thread --------------
// AddHashEntry (LM/MemHash.cpp:209-251) keeps each of the 40000 buckets a
// vector sorted by MheCompare and inserts with std::lower_bound; buckets never
// interact, so each is replayed literally by one thread, its seeds in HashMatch
// order (a stable sort by bucket): the seed is looked up, a collision when the
// entry at the lower bound compares equal, else the seed's extension (offset
// kept, mersize 0) goes in at its own lower bound.  The bucket's vector holds
// seed indices (an entry is its seed's extension).

struct UqSeeds {
    const int32_t *sst, *est;   // seed starts, extension starts (k a seed)
    const uint32_t *elen;
    const int64_t *off;
    int k, sl;
};

// E: EntU, or EntT<8, int32_t> for seeds of at most eight genomes (fewer
// slots for the comparisons to run over)
template <class E> __device__ __forceinline__ void uq_entry(const UqSeeds &q, uint32_t x, bool ext, E &e)
{
    constexpr int MK = (int)(sizeof(e.s) / sizeof(e.s[0]));
    const int32_t *s = (ext ? q.est : q.sst) + (uint64_t)x * q.k;
    e.first = -1;
#pragma unroll
    for (int g = 0; g < MK; ++g) {
        e.s[g] = g < q.k ? s[g] : 0;
        if (e.first < 0 && e.s[g] != 0) e.first = g;
    }
    e.len = ext ? q.elen[x] : (uint64_t)q.sl;
    e.mersize = ext ? 0 : (uint64_t)q.sl;
    e.off = q.off[x];
}

__global__ void uq_bucket_kernel(const int64_t *__restrict__ off, uint64_t ns, uint64_t *__restrict__ key,
                                 uint32_t *__restrict__ val, uint32_t *__restrict__ cnt)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ns) return;
    const int64_t m = off[x] % 40000;
    const uint32_t b = (uint32_t)(m < 0 ? m + 40000 : m);
    key[x] = b;
    val[x] = (uint32_t)x;
    atomicAdd(&cnt[b], 1u);
}

// std::lower_bound over the bucket's vector v[0, n) for q, with libstdc++'s
// probe sequence (MheCompare is not a strict weak order, so the probes must
// be exactly those), run by a whole wavefront: the entries of the next six
// levels of the probe tree from (first, len) -- 63 positions, one a lane --
// are loaded and compared with q at once (a ballot of the comparisons), then
// the first lane takes up to six probes from the ballot; repeated until the
// range is empty.  One round of dependent loads
// serves six probes.  The end of the range is always n or the last probe that
// was not below q, so when the result is not n, *at (first lane) is the entry
// at the result: that probe's, kept from the LDS tree.
template <class E>
__device__ uint32_t uq_lower_bound(const UqSeeds &q, const uint32_t *__restrict__ v, uint32_t n, const E &qe,
                                   E *tree, E *at)
{
    const int lane = threadIdx.x;
    uint32_t first = 0, len = n;
    while (len > 0) {
        const int node = lane + 1;                    // heap index of this lane's probe, 1..63
        bool lt = false;
        if (node < 64) {
            const int lvl = 31 - __builtin_clz((unsigned)node);
            uint32_t f = first, l = len;
            for (int bt = lvl - 1; bt >= 0 && l > 0; --bt) {
                const uint32_t half = l >> 1, mid = f + half;
                if ((node >> bt) & 1) {                   // the probe was below q: right half
                    f = mid + 1;
                    l = l - half - 1;
                } else {
                    l = half;
                }
            }
            if (l > 0) {
                E mine;
                uq_entry(q, v[f + (l >> 1)], true, mine);
                lt = mhe_less_k(mine, qe, q.k);
                tree[node] = mine;
            }
        }
        // every probe's comparison at once, one bit a node (bit node - 1)
        const uint64_t below = __ballot(lt);
        __syncthreads();
        if (lane == 0) {
            int node = 1, keep = 0;
            for (int s = 0; s < 6 && len > 0; ++s) {
                const uint32_t half = len >> 1, mid = first + half;
                if ((below >> (node - 1)) & 1ull) {
                    first = mid + 1;
                    len = len - half - 1;
                    node = 2 * node + 1;
                } else {
                    len = half;
                    keep = node;
                    node = 2 * node;
                }
            }
            if (keep) *at = tree[keep];
        }
        first = (uint32_t)__builtin_amdgcn_readfirstlane((int)first);
        len = (uint32_t)__builtin_amdgcn_readfirstlane((int)len);
        __syncthreads();
    }
    return first;
}

// One wavefront per bucket: std::lower_bound's probes as above (the
// comparisons are the reference's, one at a time, by the first lane); the
// insertion's shift of the vector's tail runs across the wave, 64 entries a
// step from the end (every lane reads its entry before any lane writes).
template <class E>
__global__ __launch_bounds__(64) void uq_replay_kernel(UqSeeds q, const uint32_t *__restrict__ order,
                                                       const uint32_t *__restrict__ bstart, uint32_t *__restrict__ vec,
                                                       uint32_t *__restrict__ nent,
                                                       unsigned long long *__restrict__ coll)
{
    __shared__ E tree[64];
    const uint32_t b = blockIdx.x;
    const int lane = threadIdx.x;
    const uint32_t y0 = bstart[b], y1 = bstart[b + 1];
    uint32_t *v = vec + y0;
    uint32_t ne = 0;
    unsigned long long c = 0;
    E cand, at_it;
    if (y0 < y1) uq_entry(q, order[y0], false, cand);
    for (uint32_t y = y0; y < y1; ++y) {
        const uint32_t x = order[y];
        uint32_t pos = 0xFFFFFFFFu;     // insertion point, or none (a collision)
        // the next seed's entry, loaded alongside this one's probes
        E next = cand;
        if (y + 1 < y1) uq_entry(q, order[y + 1], false, next);
        const uint32_t it = uq_lower_bound(q, v, ne, cand, tree, &at_it);
        bool hit = false;
        if (lane == 0 && it != ne) hit = !mhe_less_k(cand, at_it, q.k);   // !comp(*it, cand) holds at lower_bound
        hit = __builtin_amdgcn_readfirstlane((int)hit) != 0;
        if (hit) {
            if (lane == 0) ++c;
        } else {
            E e;
            uq_entry(q, x, true, e);
            pos = uq_lower_bound(q, v, ne, e, tree, &at_it);
        }
        cand = next;
        pos = (uint32_t)__builtin_amdgcn_readfirstlane((int)pos);
        if (pos == 0xFFFFFFFFu) continue;
        // shift [pos, ne) up by one, 64 entries a step from the end
        for (int64_t j = (int64_t)ne; j > (int64_t)pos; j -= 64) {
            const int64_t i = j - 64 + lane;
            uint32_t t = 0;
            const bool in = i >= (int64_t)pos;
            if (in) t = v[i];
            __syncthreads();
            if (in) v[i + 1] = t;
            __syncthreads();
        }
        if (lane == 0) v[pos] = x;
        __syncthreads();
        ++ne;
    }
    if (lane == 0) {
        nent[b] = ne;
        if (c) atomicAdd(coll, c);
    }
}

// the list in bucket order: bucket b's entries at outpos[b] ...
__global__ void uq_emit_kernel(UqSeeds q, const uint32_t *__restrict__ bstart, const uint32_t *__restrict__ vec,
                               const uint32_t *__restrict__ nent, const uint64_t *__restrict__ outpos,
                               int64_t *__restrict__ starts, uint64_t *__restrict__ lengths)
{
    const uint32_t b = blockIdx.x;
    const uint32_t ne = nent[b];
    for (uint32_t i = threadIdx.x; i < ne; i += blockDim.x) {
        const uint32_t x = vec[bstart[b] + i];
        const uint64_t o = outpos[b] + i;
        for (int g = 0; g < q.k; ++g) starts[o * q.k + g] = q.est[(uint64_t)x * q.k + g];
        lengths[o] = q.elen[x];
    }
}

// Spaced-seed extension of the k-genome seeds with one walk per chain (the
// uq_chain_* kernels): sort the seeds by (first genome, start), link each to
// its chain predecessor, find the heads by pointer jumping, walk the heads,
// copy.  ext_st / ext_len as uq_seed_kernel writes them; *err set by a walk
// that left a list; *walks the heads walked.
static int uq_extend_by_chains(mh_context *ctx, const int32_t *d_sst, uint64_t S, int k, const ExtEnvK &env,
                               int32_t *d_est, uint32_t *d_len, uint32_t *err, uint64_t *walks)
{
    DevBuf *g = ctx->grid_scratch;
    int rc;
    for (int q = 0; q < 2; ++q)
        if ((rc = g[q].ensure(S * 8 + 64)) != MH_OK) return rc;            // keys
    for (int q = 2; q < 8; ++q)
        if ((rc = g[q].ensure(S * 4 + 64)) != MH_OK) return rc;            // u32 arrays
    if ((rc = g[8].ensure(((S + JTILE - 1) / JTILE) * 4 + 64)) != MH_OK) return rc;
    if ((rc = g[10].ensure(64)) != MH_OK) return rc;
    const dim3 grid((unsigned)((S + 255) / 256)), blk(256);
    uint64_t *k0 = g[0].as<uint64_t>(), *k1 = g[1].as<uint64_t>();
    uint32_t *v0 = g[2].as<uint32_t>(), *v1 = g[3].as<uint32_t>();
    hipLaunchKernelGGL(uq_chain_key_kernel, grid, blk, 0, ctx->stream, d_sst, S, k, k0, v0);
    uint64_t *sk = nullptr;
    uint32_t *sv = nullptr;
    if ((rc = radix_sort_kv(ctx, k0, v0, k1, v1, S, 32 + 4, &sk, &sv)) != MH_OK) return rc;
    uint32_t *la = sv == v0 ? v1 : v0, *lb = g[4].as<uint32_t>();
    hipLaunchKernelGGL(uq_chain_pred_kernel, grid, blk, 0, ctx->stream, sk, sv, S, d_sst, k, (uint32_t)env.seed_len,
                       la);
    // links point to earlier sorted entries: this many doublings reach every head
    int rounds = 1;
    while ((1ull << (rounds - 1)) < S) ++rounds;
    for (int round = 0; round < rounds; ++round) {
        hipLaunchKernelGGL(grid_jump_kernel, grid, blk, 0, ctx->stream, la, lb, S);
        std::swap(la, lb);
    }
    uint32_t *is_head = g[5].as<uint32_t>(), *head_of = g[6].as<uint32_t>();
    hipLaunchKernelGGL(grid_head_kernel, grid, blk, 0, ctx->stream, sv, la, S, is_head, head_of);
    // the heads in sorted order (first genome, start), so that neighbouring
    // threads walk neighbouring stretches of the genomes and share cache lines
    uint32_t *sflag = lb, *heads = g[7].as<uint32_t>();
    hipLaunchKernelGGL(chain_root_flag_kernel, grid, blk, 0, ctx->stream, la, S, sflag);
    uint64_t nh = 0;
    if ((rc = compact_flags(ctx, sflag, S, heads, g[8].as<uint32_t>(), g[10].as<uint32_t>(), &nh)) != MH_OK)
        return rc;
    if (nh) {
        hipLaunchKernelGGL(chain_gather_kernel, dim3((unsigned)((nh + 255) / 256)), blk, 0, ctx->stream, sv, nh,
                           heads);
        const dim3 hg((unsigned)((nh + 127) / 128)), hb(128);
        if (k <= 4)
            hipLaunchKernelGGL(uq_chain_walk_kernel<4>, hg, hb, 0, ctx->stream, heads, nh, env, k, d_sst, d_est, d_len,
                               err);
        else if (k <= 8)
            hipLaunchKernelGGL(uq_chain_walk_kernel<8>, hg, hb, 0, ctx->stream, heads, nh, env, k, d_sst, d_est, d_len,
                               err);
        else
            hipLaunchKernelGGL(uq_chain_walk_kernel<16>, hg, hb, 0, ctx->stream, heads, nh, env, k, d_sst, d_est,
                               d_len, err);
    }
    hipLaunchKernelGGL(uq_chain_fill_kernel, grid, blk, 0, ctx->stream, is_head, head_of, S, k, d_est, d_len);
    MH_HIP(hipGetLastError());
    *walks = nh;
    return MH_OK;
}

int unique_seeds(mh_context *ctx, const mh_sml *const *smls, int k, const LiveRanges &live, UniqueSeeds *out)
{
    MH_CHECK(k >= 2 && k <= EXT_MAXK && k <= PLAN_MAXK, MH_ERR_UNSUPPORTED, "UniqueMatchFinder takes 2 to 16 genomes");
    const int w = smls[0]->seed_weight, sl = smls[0]->seed_length;
    int ebits = 0;
    while (ebits < 32 && (1ull << ebits) <= (uint64_t)live.restarts) ++ebits;
    const int bits = 2 * w + 4 + ebits;
    MH_CHECK(bits <= 64, MH_ERR_UNSUPPORTED, "seed weight too large for the merged sort key");
    out->k = k;
    out->n = 0;
    uint64_t N = 0;
    for (int g = 0; g < k; ++g)
        for (size_t r = 0; r < live.lo[g].size(); ++r) N += live.hi[g][r] - live.lo[g][r];
    MH_CHECK(N < (1ull << 32), MH_ERR_UNSUPPORTED, "more than 2^32 sorted-list entries");
    DevBuf *b = ctx->uq_scratch;
    int rc;
    if ((rc = b[0].ensure(N * 8 + 64)) != MH_OK) return rc;
    if ((rc = b[1].ensure(N * 8 + 64)) != MH_OK) return rc;
    if ((rc = b[2].ensure(N * 4 + 64)) != MH_OK) return rc;
    if ((rc = b[3].ensure(N * 4 + 64)) != MH_OK) return rc;
    uint64_t *k0 = b[0].as<uint64_t>(), *k1 = b[1].as<uint64_t>();
    uint32_t *v0 = b[2].as<uint32_t>(), *v1 = b[3].as<uint32_t>();
    UqLists ls;
    std::memset(&ls, 0, sizeof(ls));
    uint64_t at = 0, gbase = 0;
    for (int g = 0; g < k; ++g) {
        ls.ck[g] = smls[g]->d_keys;
        ls.pos[g] = smls[g]->d_pos;
        ls.w[g] = smls[g]->d_words;
        ls.off[g] = gbase;
        for (size_t r = 0; r < live.lo[g].size(); ++r) {
            const uint64_t lo = live.lo[g][r], cnt = live.hi[g][r] - lo;
            const uint64_t tag = ((uint64_t)live.ep[g][r] << (2 * w + 4)) | (uint64_t)g;
            if (cnt)
                hipLaunchKernelGGL(uq_gather_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, ctx->stream,
                                   smls[g]->d_keys, lo, cnt, tag, gbase, k0 + at, v0 + at);
            at += cnt;
        }
        gbase += smls[g]->n;
    }
    MH_HIP(hipGetLastError());
    if (N == 0) return MH_OK;
    uint64_t *sk;
    uint32_t *sv;
    if ((rc = radix_sort_kv(ctx, k0, v0, k1, v1, N, bits, &sk, &sv)) != MH_OK) return rc;
    uint32_t *flag = sk == k0 ? reinterpret_cast<uint32_t *>(k1) : reinterpret_cast<uint32_t *>(k0);
    uint32_t *U = sv == v0 ? v1 : v0;
    if ((rc = b[4].ensure(((N + JTILE - 1) / JTILE) * 4 + 64)) != MH_OK) return rc;
    if ((rc = b[5].ensure(64)) != MH_OK) return rc;
    uint32_t *tiles = b[4].as<uint32_t>(), *d_cnt = b[5].as<uint32_t>();
    hipLaunchKernelGGL(uq_mark_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, ctx->stream, sk, N, flag);
    uint64_t M = 0;
    if ((rc = compact_flags(ctx, flag, N, U, tiles, d_cnt, &M)) != MH_OK) return rc;
    if (M < 2) return MH_OK;
    // leaders: flags and indices after U, in the other key buffer's space
    uint32_t *flag2 = flag;
    if ((rc = b[6].ensure(M * 4 + 64)) != MH_OK) return rc;
    uint32_t *lead = b[6].as<uint32_t>();
    hipLaunchKernelGGL(uq_leader_kernel, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, ctx->stream, sk, U, M,
                       flag2);
    uint64_t S = 0;
    if ((rc = compact_flags(ctx, flag2, M, lead, tiles, d_cnt, &S)) != MH_OK) return rc;
    out->n = S;
    out->walks = S;
    if (S == 0) return MH_OK;
    const bool solid_odd = smls[0]->seed_length == smls[0]->seed_weight && (smls[0]->seed_length & 1);
    ExtEnvK env;
    std::memset(&env, 0, sizeof(env));
    env.seed_len = sl;
    env.mer_mask = w >= 32 ? ~0ull : (~0ull << (64 - 2 * w));
    for (int g = 0; g < k; ++g) {
        env.kp[g] = smls[g]->d_keys_pos;
        env.n[g] = smls[g]->n;
        env.L[g] = (int64_t)smls[g]->len;
        MH_CHECK(smls[g]->len < (1ull << 31), MH_ERR_UNSUPPORTED, "genomes of 2^31 bases or more");
    }
    if (solid_odd) {
        for (int g = 0; g < k; ++g) {
            if ((rc = ctx->uq_rc[g].ensure(smls[g]->n_words * 4 + 64)) != MH_OK) return rc;
            ls.wrc[g] = ctx->uq_rc[g].as<uint32_t>();
            hipLaunchKernelGGL(rc_pack_kernel, dim3((unsigned)((smls[g]->n_words + 255) / 256)), dim3(256), 0,
                               ctx->stream, smls[g]->d_words, smls[g]->len, smls[g]->n_words, ctx->uq_rc[g].as<uint32_t>());
        }
    }
    if ((rc = b[7].ensure(S * (uint64_t)k * 8 + S * 12 + 256)) != MH_OK) return rc;
    int32_t *d_sst = b[7].as<int32_t>();
    int32_t *d_est = d_sst + S * k;
    uint32_t *d_len = reinterpret_cast<uint32_t *>(d_est + S * k);
    int64_t *d_off = reinterpret_cast<int64_t *>((reinterpret_cast<uintptr_t>(d_len + S) + 7) & ~(uintptr_t)7);
    MH_HIP(hipMemsetAsync(d_cnt, 0, 4, ctx->stream));
    {
        TimedLaunch tl(ctx, "extend");
        // spaced seeds: starts and offsets here, one walk per chain below
        // (MH_UQ_CHAINS=0 walks every seed)
        const char *ce = getenv("MH_UQ_CHAINS");
        const bool chains = !solid_odd && !(ce && ce[0] == '0');
        hipLaunchKernelGGL(uq_seed_kernel, dim3((unsigned)((S + 127) / 128)), dim3(128), 0, ctx->stream, sk, sv, U, M,
                           lead, S, ls, env, k, solid_odd ? 1 : 0, chains ? 0 : 1, d_sst, d_est, d_len, d_off, d_cnt);
        MH_HIP(hipGetLastError());
        if (chains && (rc = uq_extend_by_chains(ctx, d_sst, S, k, env, d_est, d_len, d_cnt, &out->walks)) != MH_OK)
            return rc;
    }
    uint32_t err = 0;
    MH_HIP(hipMemcpyAsync(&err, d_cnt, 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    MH_CHECK(err == 0, MH_ERR_UNSUPPORTED, "seed extension left the sorted mer lists");
    // AddHashEntry, bucket by bucket: seeds ordered by bucket (stable), bucket
    // starts, one thread per bucket, the list in bucket order
    const double t0 = now_ms();
    constexpr uint32_t NB = 40000;
    if ((rc = b[0].ensure(S * 8 * 2 + 64)) != MH_OK) return rc;
    if ((rc = b[2].ensure(S * 4 * 2 + (NB + 1) * 4 * 3 + NB * 8 + 256)) != MH_OK) return rc;
    uint64_t *bk0 = b[0].as<uint64_t>(), *bk1 = bk0 + S;
    uint32_t *bv0 = b[2].as<uint32_t>(), *bv1 = bv0 + S;
    uint32_t *bcnt = bv1 + S, *bst = bcnt + (NB + 1), *nent = bst + (NB + 1);
    uint64_t *outpos = reinterpret_cast<uint64_t *>((reinterpret_cast<uintptr_t>(nent + NB + 1) + 7) & ~(uintptr_t)7);
    MH_HIP(hipMemsetAsync(bcnt, 0, (NB + 1) * 4, ctx->stream));
    hipLaunchKernelGGL(uq_bucket_kernel, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, ctx->stream, d_off, S,
                       bk0, bv0, bcnt);
    uint64_t *ok;
    uint32_t *order;
    if ((rc = radix_sort_kv(ctx, bk0, bv0, bk1, bv1, S, 16, &ok, &order)) != MH_OK) return rc;
    std::vector<uint32_t> hc(NB + 1, 0), hs(NB + 1, 0);
    MH_HIP(hipMemcpyAsync(hc.data(), bcnt, NB * 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    for (uint32_t q = 0; q < NB; ++q) hs[q + 1] = hs[q] + hc[q];
    MH_HIP(hipMemcpyAsync(bst, hs.data(), (NB + 1) * 4, hipMemcpyHostToDevice, ctx->stream));
    uint32_t *vec = order == bv0 ? bv1 : bv0;     // the other value buffer: S slots
    unsigned long long *d_coll = reinterpret_cast<unsigned long long *>(ok == bk0 ? bk1 : bk0);
    MH_HIP(hipMemsetAsync(d_coll, 0, 8, ctx->stream));
    UqSeeds qs;
    qs.sst = d_sst;
    qs.est = d_est;
    qs.elen = d_len;
    qs.off = d_off;
    qs.k = k;
    qs.sl = sl;
    {
        TimedLaunch tl(ctx, "partition");
        if (k <= 8)
            hipLaunchKernelGGL((uq_replay_kernel<EntT<8, int32_t>>), dim3(NB), dim3(64), 0, ctx->stream, qs, order, bst, vec,
                               nent, d_coll);
        else
            hipLaunchKernelGGL(uq_replay_kernel<EntU>, dim3(NB), dim3(64), 0, ctx->stream, qs, order, bst, vec, nent,
                               d_coll);
        MH_HIP(hipGetLastError());
    }
    std::vector<uint32_t> hn(NB);
    unsigned long long coll = 0;
    MH_HIP(hipMemcpyAsync(hn.data(), nent, NB * 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipMemcpyAsync(&coll, d_coll, 8, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<uint64_t> hp(NB + 1, 0);
    for (uint32_t q = 0; q < NB; ++q) hp[q + 1] = hp[q] + hn[q];
    const uint64_t nm = hp[NB];
    MH_HIP(hipMemcpyAsync(outpos, hp.data(), NB * 8, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = b[1].ensure(nm * (uint64_t)k * 8 + nm * 8 + 64)) != MH_OK) return rc;
    int64_t *d_ost = b[1].as<int64_t>();
    uint64_t *d_olen = reinterpret_cast<uint64_t *>(d_ost + nm * k);
    hipLaunchKernelGGL(uq_emit_kernel, dim3(NB), dim3(64), 0, ctx->stream, qs, bst, vec, nent, outpos, d_ost, d_olen);
    MH_HIP(hipGetLastError());
    out->starts.resize(nm * k);
    out->lengths.resize(nm);
    if (nm) {
        MH_HIP(hipMemcpyAsync(out->starts.data(), d_ost, nm * k * 8, hipMemcpyDeviceToHost, ctx->stream));
        MH_HIP(hipMemcpyAsync(out->lengths.data(), d_olen, nm * 8, hipMemcpyDeviceToHost, ctx->stream));
    }
    MH_HIP(hipStreamSynchronize(ctx->stream));
    out->collisions = (int64_t)coll;
    host_time(ctx, "host_replay", now_ms() - t0);
    return MH_OK;
}

}  // namespace mh
