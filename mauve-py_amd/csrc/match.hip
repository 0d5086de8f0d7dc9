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
#include <cstring>

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
        if (dir < 2 && maxlen >= jump) {
            // The seed-length jumps of phases 0 and 1: the addresses of the
            // K = maxlen / jump lookups do not depend on their outcomes, so
            // they are issued in groups of PREFETCH and checked in order; the
            // state after the first failing step (or after the K-th) is what
            // the step-by-step loop leaves (len and the forward starts moved by
            // the successful steps only).
            constexpr int PREFETCH = 8;
            const int64_t K = maxlen / jump;
            int64_t ok_steps = K;
            bool done = false;
            for (int64_t k0 = 1; k0 <= K && !done; k0 += PREFETCH) {
                uint64_t m0[PREFETCH], m1[PREFETCH];
                int q0[PREFETCH], q1[PREFETCH];
                bool in[PREFETCH];
                const int cnt = (int)((K - k0 + 1) < PREFETCH ? (K - k0 + 1) : PREFETCH);
#pragma unroll
                for (int u = 0; u < PREFETCH; ++u) {
                    if (u >= cnt) break;
                    const int64_t k = k0 + u;
                    const int64_t ln = len + k * jump;
                    const int64_t p0 = st[0] > 0 ? st[0] - k * jump : st[0];
                    const int64_t p1 = st[1] > 0 ? st[1] - k * jump : st[1];
                    in[u] = lookup(e, 0, p0, ln, st[0], m0[u], q0[u]) && lookup(e, 1, p1, ln, st[1], m1[u], q1[u]);
                }
                for (int u = 0; u < cnt; ++u) {
                    if (++guard > guard_max) return false;
                    if (!in[u]) return false;
                    if (m0[u] != m1[u] || q0[u] != q1[u]) {
                        ok_steps = k0 + u - 1;
                        done = true;
                        break;
                    }
                }
            }
            len += ok_steps * jump;
            for (int x = 0; x < 2; ++x)
                if (st[x] > 0) st[x] -= ok_steps * jump;
            maxlen = 0;   // the loop below does not run
        }
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

__global__ void seed_kernel_solid(const uint32_t *__restrict__ pi, const uint32_t *__restrict__ pj, uint64_t ns,
                                  const uint64_t *__restrict__ ck0, const uint32_t *__restrict__ pos0,
                                  const uint64_t *__restrict__ ck1, const uint32_t *__restrict__ pos1,
                                  const uint32_t *__restrict__ wa, const uint32_t *__restrict__ wb,
                                  const uint32_t *__restrict__ wbrc, int64_t La, int64_t Lb, int k,
                                  SeedRec *__restrict__ out)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ns) return;
    const uint32_t i = pi[x], j = pj[x];
    const int64_t p0 = pos0[i];
    const int64_t p1f = pos1[j];
    const bool rev = (ck0[i] & 1ull) != (ck1[j] & 1ull);
    SeedRec r;
    r.s0 = (int32_t)(p0 + 1);
    if (!rev) {
        const int64_t left = match_left(wa, p0, wb, p1f, p0 < p1f ? p0 : p1f);
        const int64_t ra = La - (p0 + k), rb = Lb - (p1f + k);
        const int64_t right = match_right(wa, p0 + k, wb, p1f + k, ra < rb ? ra : rb);
        r.s1 = (int32_t)(p1f + 1);
        r.e0 = (int32_t)(p0 - left + 1);
        r.e1 = (int32_t)(p1f - left + 1);
        r.elen = (uint32_t)(k + left + right);
    } else {
        // genome 1 window in reverse-complement coordinates
        const int64_t p1 = Lb - p1f - k;
        const int64_t left = match_left(wa, p0, wbrc, p1, p0 < p1 ? p0 : p1);
        const int64_t ra = La - (p0 + k), rb = Lb - (p1 + k);
        const int64_t right = match_right(wa, p0 + k, wbrc, p1 + k, ra < rb ? ra : rb);
        const int64_t L = k + left + right;
        r.s1 = (int32_t)(-(p1f + 1));
        r.e0 = (int32_t)(p0 - left + 1);
        r.e1 = (int32_t)(-(Lb - (p1 - left + L) + 1));
        r.elen = (uint32_t)L;
    }
    out[x] = r;
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
int join_and_extend(mh_context *ctx, const mh_sml *a, const mh_sml *b, const SeedRec **seeds_out, uint64_t *ns_out,
                    uint32_t *repeat_runs)
{
    const uint64_t n0 = a->n, n1 = b->n;
    *seeds_out = nullptr;
    *ns_out = 0;
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
    const bool solid_odd = a->seed_length == a->seed_weight && (a->seed_length & 1);
    if (solid_odd) {
        if ((rc = ctx->mf_scratch[7].ensure(b->n_words * 4 + 64)) != MH_OK) return rc;
        uint32_t *wbrc = ctx->mf_scratch[7].as<uint32_t>();
        TimedLaunch tl(ctx, "extend");
        hipLaunchKernelGGL(rc_pack_kernel, dim3((unsigned)((b->n_words + 255) / 256)), dim3(256), 0, ctx->stream,
                           b->d_words, b->len, b->n_words, wbrc);
        hipLaunchKernelGGL(seed_kernel_solid, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream, oi, oj,
                           ns, a->d_keys, a->d_pos, b->d_keys, b->d_pos, a->d_words, b->d_words, wbrc,
                           (int64_t)a->len, (int64_t)b->len, a->seed_length, d_seeds);
        MH_HIP(hipGetLastError());
    } else {
        TimedLaunch tl(ctx, "extend");
        hipLaunchKernelGGL(seed_kernel, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, ctx->stream, oi, oj, ns,
                           a->d_keys, a->d_pos, b->d_keys, b->d_pos, env, d_seeds, cnt + 2);
        MH_HIP(hipGetLastError());
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

}  // namespace mh
