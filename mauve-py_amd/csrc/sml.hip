// sml.hip -- sorted mer list build on gfx950.
//
// K1a  pack:  ASCII -> 2-bit words, 16 characters per uint32, first character
//             in the two most significant bits (translate32,
//             LM/SortedMerList.cpp:425-460; BasicDNATable 29-47).
// K1b  keys:  canonical seed key per position (GetDnaSeedMer,
//             LM/SortedMerList.cpp:726-769): the pattern-selected characters
//             left-aligned in 64 bits, or the reverse complement with the LSB
//             strand flag set (RevCompMer 597-614), whichever is smaller.
//             Keys are also written in compact form for sorting:
//             (selected chars << 1) | flag, 2w+1 significant bits, which
//             orders exactly like the 64-bit key.
// K2   sort:  onesweep LSD radix sort of (compact key, position) by digits
//             of 8 to 10 bits (the fewest passes): one histogram kernel for every pass, then per pass one
//             kernel whose tiles rank their keys stably in LDS (wave ballots
//             per digit bit), find their digit offsets by decoupled look-back
//             over the preceding tiles and write every digit's run out
//             contiguously.  Positions enter in ascending order and every
//             pass is stable, so equal keys stay in position order.
#include "mh_internal.hpp"
#include <algorithm>
#include <cstring>

namespace mh {
namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;   // 4096 entries per tile

struct SeedPattern {
    int sl, sw, spaced;
    int off[32];          // offsets of the selected characters, first to last
};

__device__ __forceinline__ uint32_t dna_code(uint32_t c)
{
    // BasicDNATable: c/C/b/B/y/Y -> 1, g/G/s/S/k/K -> 2, t/T -> 3, else 0.
    const uint32_t l = c | 0x20u;   // exact fold for letters; other bytes stay distinct
    if (l == 'c' || l == 'b' || l == 'y') return 1;
    if (l == 'g' || l == 's' || l == 'k') return 2;
    if (l == 't') return 3;
    return 0;
}

__global__ void pack_kernel(const uint8_t *__restrict__ ascii, uint64_t len, uint64_t n_words,
                            uint32_t *__restrict__ words, int *__restrict__ err, int aligned16)
{
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n_words) return;
    const uint64_t base = w * 16;
    uint32_t out = 0;
    if (aligned16 && base + 16 <= len) {
        const uint4 v = *reinterpret_cast<const uint4 *>(ascii + base);
        const uint32_t d[4] = {v.x, v.y, v.z, v.w};
        bool gap = false;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const uint32_t c = (d[q] >> (8 * b)) & 0xFFu;
                gap |= (c == '-');
                out |= dna_code(c) << (30 - 2 * (4 * q + b));
            }
        if (gap) atomicOr(err, 1);
    } else {
        for (int k = 0; k < 16; ++k) {
            const uint64_t p = base + k;
            if (p < len) {
                const uint32_t c = ascii[p];
                if (c == '-') atomicOr(err, 1);
                out |= dna_code(c) << (30 - 2 * k);
            }
        }
    }
    words[w] = out;
}

__device__ __forceinline__ uint64_t get_mer(const uint32_t *__restrict__ words, uint64_t pos)
{
    const uint64_t w = pos >> 4;
    const uint32_t bit = (uint32_t)(pos & 15) * 2;
    uint64_t m = ((uint64_t)words[w] << 32) | words[w + 1];
    if (bit) m = (m << bit) | (words[w + 2] >> (32 - bit));
    return m;
}

__global__ void keys_kernel(const uint32_t *__restrict__ words, uint64_t n, SeedPattern sp,
                            uint64_t *__restrict__ keys_pos, uint64_t *__restrict__ ckeys,
                            uint32_t *__restrict__ pos_out)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint64_t mer = get_mer(words, p);
    uint64_t fwd = 0, rc = 0;
    if (!sp.spaced) {
        fwd = mer >> (64 - 2 * sp.sl);
#pragma unroll 4
        for (int k = 0; k < sp.sl; ++k) rc |= (uint64_t)(3u - ((mer >> (62 - 2 * k)) & 3u)) << (2 * k);
    } else {
        for (int k = 0; k < sp.sw; ++k) {
            const uint32_t ch = (uint32_t)(mer >> (62 - 2 * sp.off[k])) & 3u;
            fwd = (fwd << 2) | ch;
            rc |= (uint64_t)(3u - ch) << (2 * k);
        }
    }
    const int sh = 64 - 2 * sp.sw;
    const uint64_t f64 = fwd << sh;
    const uint64_t r64 = (rc << sh) | 1ull;
    const bool use_r = r64 < f64;
    keys_pos[p] = use_r ? r64 : f64;
    ckeys[p] = use_r ? ((rc << 1) | 1ull) : (fwd << 1);
    pos_out[p] = (uint32_t)p;
}

// ---- K2: onesweep LSD radix sort ------------------------------------------
//
// One histogram kernel counts the digits of every pass in a single read of
// the keys (per-workgroup LDS histograms, reduced and scanned by a second
// small kernel into each pass's digit starts).  Then one kernel per pass:
// each workgroup takes the next tile ticket (tiles in ticket order, so a
// tile's predecessors are already running), ranks its 4096 keys stably in LDS,
// publishes its digit counts, and finds each digit's offset among the
// preceding tiles by decoupled look-back over their published counts: a
// predecessor's word carries either its own count (aggregate) or the sum over
// it and every tile before it (inclusive prefix), which ends the walk.  A
// status word packs (epoch, flag, value) into 64 bits, so one atomic load sees
// a consistent pair, and words of earlier passes (other epochs) read as not
// yet published; the array is never cleared between passes.  A wait is
// bounded: past OS_SPIN_MAX polls the tile records an error (the sort's
// caller checks it) and moves on.

constexpr int OS_MAX_BITS = 10;                       // widest digit (radix 1024)
constexpr int OS_MAX_RADIX = 1 << OS_MAX_BITS;
constexpr int OS_MAX_PASSES = 8;                      // 64-bit keys
constexpr int OS_HIST_WGS = 256;                      // workgroups of the histogram
constexpr int OS_HIST_THREADS = 256;
constexpr int OS_HIST_LDS = 7 * OS_MAX_RADIX;          // passes x radix at most (64-bit keys: 7 x 1024)
constexpr uint32_t OS_SPIN_MAX = 1u << 22;
constexpr uint64_t OS_FLAG_A = 1, OS_FLAG_P = 2;
constexpr uint64_t OS_VALUE_MASK = (1ull << 38) - 1;

__device__ __forceinline__ uint64_t os_word(uint32_t epoch, uint64_t flag, uint64_t v)
{
    return ((uint64_t)epoch << 40) | (flag << 38) | v;
}

// Per-workgroup digit counts of every pass (digits of db bits); workgroup 0
// also zeroes the passes' tile tickets (the passes run after this kernel on
// the stream).
__global__ __launch_bounds__(OS_HIST_THREADS) void os_hist_kernel(const uint64_t *__restrict__ keys, uint64_t n,
                                                                 int db, int passes, uint32_t *__restrict__ part,
                                                                 uint32_t *__restrict__ tickets)
{
    __shared__ uint32_t h[OS_HIST_LDS];
    const int t = threadIdx.x, radix = 1 << db;
    const uint32_t mask = (uint32_t)radix - 1u;
    for (int i = t; i < passes * radix; i += OS_HIST_THREADS) h[i] = 0;
    if (blockIdx.x == 0 && t < OS_MAX_PASSES) tickets[t] = 0;
    __syncthreads();
    // a contiguous chunk per workgroup, eight keys a thread loaded at once
    const uint64_t chunk = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = (uint64_t)blockIdx.x * chunk, hi = min<uint64_t>(n, lo + chunk);
    constexpr int U = 8;
    for (uint64_t x0 = lo + t; x0 < hi; x0 += (uint64_t)U * OS_HIST_THREADS) {
        uint64_t k[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t x = x0 + (uint64_t)u * OS_HIST_THREADS;
            k[u] = x < hi ? keys[x] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (x0 + (uint64_t)u * OS_HIST_THREADS >= hi) break;
            for (int q = 0; q < passes; ++q) atomicAdd(&h[q * radix + ((uint32_t)(k[u] >> (db * q)) & mask)], 1u);
        }
    }
    __syncthreads();
    uint32_t *out = part + (uint64_t)blockIdx.x * passes * radix;
    for (int i = t; i < passes * radix; i += OS_HIST_THREADS) out[i] = h[i];
}

// Pass q = blockIdx.x, digit d = threadIdx.x (blockDim = radix): the digit's
// total over the histogram workgroups, then the exclusive scan of the totals
// (the first output position of each digit).
__global__ __launch_bounds__(OS_MAX_RADIX) void os_start_kernel(const uint32_t *__restrict__ part, int nparts,
                                                               int passes, uint32_t *__restrict__ start)
{
    __shared__ uint32_t s[OS_MAX_RADIX];
    const int d = threadIdx.x, q = blockIdx.x, radix = blockDim.x;
    uint32_t sum = 0;
#pragma unroll 16
    for (int w = 0; w < nparts; ++w) sum += part[((uint64_t)w * passes + q) * radix + d];
    s[d] = sum;
    __syncthreads();
    for (int off = 1; off < radix; off <<= 1) {
        const uint32_t v = d >= off ? s[d - off] : 0u;
        __syncthreads();
        s[d] += v;
        __syncthreads();
    }
    start[q * radix + d] = s[d] - sum;
}

// One pass over the DB-bit digit (key >> shift) & (2^DB - 1).  Wave w ranks
// its own contiguous quarter of the tile (16 rounds of 64 keys): per round,
// lanes with the same digit find each other with DB ballots, read the wave's
// running count of that digit from LDS and the round's leader advances it
// (LDS operations of one wave complete in order).  Thread t owns digits t,
// t + 256, ... for the tile's counts, the look-back and the global bases.
template <int DB>
__global__ __launch_bounds__(SORT_THREADS) void os_pass_kernel(const uint64_t *__restrict__ kin,
                                                              const uint32_t *__restrict__ vin,
                                                              uint64_t *__restrict__ kout,
                                                              uint32_t *__restrict__ vout, uint64_t n, int shift,
                                                              const uint32_t *__restrict__ dstart,
                                                              uint64_t *__restrict__ status,
                                                              uint32_t *__restrict__ ticket, uint32_t epoch,
                                                              int *__restrict__ err)
{
    constexpr int NWAVES = SORT_THREADS / 64;
    constexpr int R = 1 << DB, E = R / SORT_THREADS;   // digits a thread owns
    constexpr uint32_t MASK = (uint32_t)R - 1u;
    __shared__ uint64_t sk[SORT_TILE];
    __shared__ uint32_t sv[SORT_TILE];
    __shared__ uint16_t wcnt[NWAVES][R];   // per wave and digit: counts, then offsets (<= 4096)
    __shared__ uint16_t start[R];          // the tile's digit counts, then its digit run starts
    __shared__ uint32_t tile_s;
    // the key tile's LDS is free until the keys are placed: the scan's sums
    uint32_t *part = reinterpret_cast<uint32_t *>(sk);
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (t == 0) tile_s = atomicAdd(ticket, 1u);
    for (int w = 0; w < NWAVES; ++w)
        for (int e = 0; e < E; ++e) wcnt[w][t + e * SORT_THREADS] = 0;
    __syncthreads();
    const uint32_t tile = tile_s;
    const uint64_t base = (uint64_t)tile * SORT_TILE;
    const uint64_t wbase = base + (uint64_t)wv * (SORT_ITEMS * 64);
    const uint64_t lane_lt = (1ull << lane) - 1ull;
    uint64_t key[SORT_ITEMS];
    uint32_t val[SORT_ITEMS], rank[SORT_ITEMS];
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
        const uint64_t x = wbase + (uint64_t)r * 64 + lane;
        key[r] = x < n ? kin[x] : 0ull;
        val[r] = x < n ? vin[x] : 0u;
    }
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
        const bool valid = wbase + (uint64_t)r * 64 + lane < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & MASK;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lane_lt);
        const uint32_t cur = wcnt[wv][d];
        rank[r] = cur + below;
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) wcnt[wv][d] = (uint16_t)(cur + (uint32_t)__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // owned digits: wave offsets within the tile and the tile's counts, which
    // are published at once (tile 0: its count is its inclusive prefix)
    uint32_t cnt[E];
    uint64_t *st = status + (uint64_t)tile * R;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int d = t + e * SORT_THREADS;
        uint32_t c = 0;
        for (int w = 0; w < NWAVES; ++w) {
            const uint32_t x = wcnt[w][d];
            wcnt[w][d] = (uint16_t)c;
            c += x;
        }
        cnt[e] = c;
        __hip_atomic_store(&st[d], os_word(epoch, tile == 0 ? OS_FLAG_P : OS_FLAG_A, c), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // look-back: each owned digit's count over every preceding tile
    uint32_t gb[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int d = t + e * SORT_THREADS;
        uint64_t before = 0;
        if (tile > 0) {
            int64_t j = (int64_t)tile - 1;
            uint32_t spins = 0;
            for (;;) {
                const uint64_t w = __hip_atomic_load(&status[(uint64_t)j * R + d], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t flag = (w >> 38) & 3u;
                if ((uint32_t)(w >> 40) == epoch && flag != 0) {
                    before += w & OS_VALUE_MASK;
                    if (flag == OS_FLAG_P || --j < 0) break;
                    continue;
                }
                if (++spins > OS_SPIN_MAX) {
                    atomicOr(err, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __hip_atomic_store(&st[d], os_word(epoch, OS_FLAG_P, (before + cnt[e]) & OS_VALUE_MASK),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        gb[e] = dstart[d] + (uint32_t)before;
        start[d] = (uint16_t)cnt[e];
    }
    __syncthreads();
    // exclusive scan of the tile's digit counts (the start of each digit run):
    // thread t sums digits [t*E, t*E + E), the 256 sums are scanned, then each
    // thread writes its digits' starts
    uint32_t loc[E], tot = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        loc[e] = start[t * E + e];
        tot += loc[e];
    }
    part[t] = tot;
    __syncthreads();
    for (int off = 1; off < SORT_THREADS; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - tot;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        start[t * E + e] = (uint16_t)run;
        run += loc[e];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_ITEMS; ++r) {
        if (wbase + (uint64_t)r * 64 + lane < n) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & MASK;
            const uint32_t p = start[d] + wcnt[wv][d] + rank[r];
            sk[p] = key[r];
            sv[p] = val[r];
        }
    }
    __syncthreads();
    // the wave counts' LDS (no longer read) takes each digit's offset from a
    // tile slot to its output position
    uint32_t *off = reinterpret_cast<uint32_t *>(&wcnt[0][0]);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int d = t + e * SORT_THREADS;
        off[d] = gb[e] - (uint32_t)start[d];
    }
    __syncthreads();
    const uint32_t tile_n = (uint32_t)min<uint64_t>(SORT_TILE, n - base);
    for (uint32_t p = t; p < tile_n; p += SORT_THREADS) {
        const uint64_t k = sk[p];
        const uint32_t d = (uint32_t)(k >> shift) & MASK;
        const uint64_t g = (uint32_t)(off[d] + p);   // below n < 2^32, so exact mod 2^32
        kout[g] = k;
        vout[g] = sv[p];
    }
}

__global__ void find_mer_kernel(const uint64_t *__restrict__ ckeys, uint64_t last, uint64_t query, int sw,
                                uint64_t *__restrict__ out)
{
    // SortedMerList::bsearch, LM/SortedMerList.cpp:380-394 (one thread); the
    // probed keys are expanded to 64 bits so any query compares as in the
    // reference
    uint64_t start = 0, end = last;
    for (;;) {
        const uint64_t mid = (start + end) / 2;
        const uint64_t c = ckeys[mid];
        const uint64_t mm = ((c >> 1) << (64 - 2 * sw)) | (c & 1ull);
        if (mm == query) { out[0] = mid; return; }
        if (mm < query && mid < end) { start = mid + 1; continue; }
        if (mm > query && start < mid) { end = mid - 1; continue; }
        out[0] = mid;
        return;
    }
}

__global__ void gather_keys_kernel(const uint64_t *__restrict__ keys_pos, uint64_t n,
                                   const uint32_t *__restrict__ pos, uint64_t m, uint64_t *__restrict__ out)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= m) return;
    const uint32_t p = pos[x];
    out[x] = p < n ? keys_pos[p] : ~0ull;
}

// flags a descending neighbour pair of sorted keys
__global__ void check_sorted_kernel(const uint64_t *__restrict__ keys, uint64_t n, int *__restrict__ bad)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x + 1 >= n) return;
    if (keys[x] > keys[x + 1]) atomicOr(bad, 1);
}

// ASCII -> DP operand classes: A,C,G,T (U as T), either case -> 0..3, X -> 4,
// any other letter -> 5 (libMUSCLE's wildcard counts, see dp_float.hip); the
// gap characters '-' and '.' are not residues and flag
__global__ void codes_kernel(const uint8_t *__restrict__ ascii, uint64_t len, uint8_t *__restrict__ codes,
                             int *__restrict__ bad)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= len) return;
    const uint32_t raw = ascii[p], c = raw | 0x20u;
    uint32_t v = 5;
    if (c == 'a') v = 0;
    else if (c == 'c') v = 1;
    else if (c == 'g') v = 2;
    else if (c == 't' || c == 'u') v = 3;
    else if (c == 'x') v = 4;
    if (raw == '-' || raw == '.') atomicOr(bad, 1);
    codes[p] = (uint8_t)v;
}

SeedPattern make_pattern(uint64_t seed)
{
    SeedPattern sp{};
    sp.sl = seed_length((int64_t)seed);
    sp.sw = seed_weight((int64_t)seed);
    sp.spaced = sp.sl != sp.sw;
    // characters are taken from the seed's most significant set bit down
    // (GetSeedMer walks char_mask from bit seed_length-1 to 0)
    const int low = __builtin_ctzll(seed);
    int k = 0;
    for (int q = 0; q < sp.sl; ++q)
        if ((seed >> (low + sp.sl - 1 - q)) & 1ull) sp.off[k++] = q;
    return sp;
}

}  // namespace

// compact key <-> 64-bit key
static inline uint64_t full_key(uint64_t c, int sw) { return ((c >> 1) << (64 - 2 * sw)) | (c & 1ull); }
static inline uint64_t compact_key(uint64_t k, int sw) { return ((k >> (64 - 2 * sw)) << 1) | (k & 1ull); }

// Buffers for s (n_words, n set): a released set from the context's pool
// that is large enough, else fresh allocations.
static int sml_take_buffers(mh_context *ctx, mh_sml *s)
{
    const uint64_t n = s->n ? s->n : 1;
    for (size_t q = 0; q < ctx->sml_pool.size(); ++q) {
        SmlBufs &b = ctx->sml_pool[q];
        if (b.cap_words >= s->n_words && b.cap_n >= n) {
            s->d_words = b.words;
            s->d_keys_pos = b.keys_pos;
            s->d_keys = b.keys;
            s->d_pos = b.pos;
            s->cap_words = b.cap_words;
            s->cap_n = b.cap_n;
            ctx->sml_pool.erase(ctx->sml_pool.begin() + (ptrdiff_t)q);
            return MH_OK;
        }
    }
    // 16 spare elements: a list's buffers can trade places with the sort
    // scratch, which asks for n*8 + 64 and n*4 + 64 bytes
    s->cap_words = s->n_words;
    s->cap_n = n + 16;
    MH_HIP(hipMalloc(&s->d_words, s->cap_words * 4));
    MH_HIP(hipMalloc(&s->d_keys_pos, s->cap_n * 8));
    MH_HIP(hipMalloc(&s->d_keys, s->cap_n * 8));
    MH_HIP(hipMalloc(&s->d_pos, s->cap_n * 4));
    return MH_OK;
}

int sml_build(mh_context *ctx, const void *d_ascii, uint64_t len, uint64_t seed, mh_sml *s)
{
    const int sl = seed_length((int64_t)seed), sw = seed_weight((int64_t)seed);
    MH_CHECK(sl > 0, MH_ERR_INVALID, "Can't have 0 seed length");
    MH_CHECK(sl <= 31, MH_ERR_INVALID, "Mer size is too large");    // LM/DNAFileSML.cpp:51-53
    MH_CHECK(len < (1ull << 32), MH_ERR_UNSUPPORTED, "sequence longer than 2^32 (32-bit SML positions)");
    s->ctx = ctx;
    s->len = len;
    s->seed = seed;
    s->seed_length = sl;
    s->seed_weight = sw;
    s->n = len >= (uint64_t)sl ? len - sl + 1 : 0;
    s->n_words = (len * 2 + 31) / 32 + 2;
    const uint64_t n = s->n;
    int rc;
    if ((rc = sml_take_buffers(ctx, s)) != MH_OK) return rc;
    if ((rc = ctx->sort_tmp[0].ensure((n ? n : 1) * 8 + 64)) != MH_OK) return rc;
    if ((rc = ctx->sort_tmp[1].ensure((n ? n : 1) * 4 + 64)) != MH_OK) return rc;
    if ((rc = ctx->sort_tmp[3].ensure(64)) != MH_OK) return rc;
    int *d_err = ctx->sort_tmp[3].as<int>();
    MH_HIP(hipMemsetAsync(d_err, 0, 4, ctx->stream));
    {
        TimedLaunch tl(ctx, "sml_pack");
        const uint64_t blocks = (s->n_words + 255) / 256;
        hipLaunchKernelGGL(pack_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                           static_cast<const uint8_t *>(d_ascii), len, s->n_words, s->d_words, d_err,
                           (int)(((uintptr_t)d_ascii & 15u) == 0));
        MH_HIP(hipGetLastError());
    }
    int herr = 0;
    MH_HIP(hipMemcpyAsync(&herr, d_err, 4, hipMemcpyDeviceToHost, ctx->stream));
    if (n == 0) {
        MH_HIP(hipStreamSynchronize(ctx->stream));
        MH_CHECK(herr == 0, MH_ERR_GAP_CHAR, "Gap in genome sequence");
        return MH_OK;
    }
    const SeedPattern sp = make_pattern(seed);
    uint64_t *k0 = s->d_keys, *k1 = ctx->sort_tmp[0].as<uint64_t>();
    uint32_t *v0 = s->d_pos, *v1 = ctx->sort_tmp[1].as<uint32_t>();
    {
        TimedLaunch tl(ctx, "sml_keys");
        hipLaunchKernelGGL(keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, s->d_words, n,
                           sp, s->d_keys_pos, k0, v0);
        MH_HIP(hipGetLastError());
    }
    {
        TimedLaunch tl(ctx, "sml_sort");
        uint64_t *rk = nullptr;
        uint32_t *rv = nullptr;
        int rc2 = radix_sort_kv(ctx, k0, v0, k1, v1, n, 2 * sw + 1, &rk, &rv);
        if (rc2 != MH_OK) return rc2;
        k0 = rk;
        v0 = rv;
    }
    if (k0 != s->d_keys) {
        // odd number of passes: the result is in the scratch pair, so the list
        // takes the scratch buffers and the scratch takes the list's (capacity
        // in elements = the smallest of the list's buffers)
        DevBuf &tk = ctx->sort_tmp[0], &tv = ctx->sort_tmp[1];
        const uint64_t cap_k = tk.bytes / 8, cap_v = tv.bytes / 4;
        void *old_k = s->d_keys, *old_v = s->d_pos;
        s->d_keys = static_cast<uint64_t *>(tk.ptr);
        s->d_pos = static_cast<uint32_t *>(tv.ptr);
        tk.ptr = old_k;
        tk.bytes = s->cap_n * 8;
        tv.ptr = old_v;
        tv.bytes = s->cap_n * 4;
        s->cap_n = std::min<uint64_t>(s->cap_n, std::min(cap_k, cap_v));
    }
    MH_HIP(hipStreamSynchronize(ctx->stream));
    MH_CHECK(herr == 0, MH_ERR_GAP_CHAR, "Gap in genome sequence");
    return sort_check(ctx);
}

int sml_alloc(mh_context *ctx, mh_sml *s) { return sml_take_buffers(ctx, s); }

int sml_rebuild_keys(mh_context *ctx, mh_sml *s, int *unsorted)
{
    *unsorted = 0;
    const uint64_t n = s->n;
    if (n == 0) return MH_OK;
    int rc;
    if ((rc = ctx->sort_tmp[0].ensure(n * 8 + 64)) != MH_OK) return rc;
    if ((rc = ctx->sort_tmp[1].ensure(n * 4 + 64)) != MH_OK) return rc;
    if ((rc = ctx->sort_tmp[3].ensure(64)) != MH_OK) return rc;
    uint64_t *ck = ctx->sort_tmp[0].as<uint64_t>();
    int *d_bad = ctx->sort_tmp[3].as<int>();
    const SeedPattern sp = make_pattern(s->seed);
    const dim3 grid((unsigned)((n + 255) / 256));
    // keys per position (and the compact keys in position order), then the
    // list's keys gathered at its sorted positions
    hipLaunchKernelGGL(keys_kernel, grid, dim3(256), 0, ctx->stream, s->d_words, n, sp, s->d_keys_pos, ck,
                       ctx->sort_tmp[1].as<uint32_t>());
    hipLaunchKernelGGL(gather_keys_kernel, grid, dim3(256), 0, ctx->stream, ck, n, s->d_pos, n, s->d_keys);
    MH_HIP(hipMemsetAsync(d_bad, 0, 4, ctx->stream));
    hipLaunchKernelGGL(check_sorted_kernel, grid, dim3(256), 0, ctx->stream, s->d_keys, n, d_bad);
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(unsorted, d_bad, 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int radix_sort_kv(mh_context *ctx, uint64_t *k0, uint32_t *v0, uint64_t *k1, uint32_t *v1, uint64_t n, int bits,
                  uint64_t **out_k, uint32_t **out_v)
{
    *out_k = k0;
    *out_v = v0;
    if (n == 0) return MH_OK;
    MH_CHECK(bits > 0 && bits <= 64, MH_ERR_INVALID, "radix sort: key bits out of range");
    MH_CHECK(n < (1ull << 32), MH_ERR_UNSUPPORTED, "radix sort: more than 2^32 - 1 keys");
    const uint32_t ntiles = (uint32_t)((n + SORT_TILE - 1) / SORT_TILE);
    // the fewest passes of digits up to 10 bits, then the narrowest digit
    // (8 at least) that covers the key in that many passes
    int max_bits = OS_MAX_BITS;
    if (const char *e = getenv("MH_SORT_MAX_DIGIT")) max_bits = std::min(OS_MAX_BITS, std::max(8, atoi(e)));
    const int passes = (bits + max_bits - 1) / max_bits;
    const int db = std::max(8, (bits + passes - 1) / passes);
    const int radix = 1 << db;
    MH_CHECK(passes * radix <= OS_HIST_LDS && db <= OS_MAX_BITS, MH_ERR_INVALID, "radix sort: digit table too large");
    const uint32_t nparts = std::min<uint32_t>(OS_HIST_WGS, ntiles);
    const uint64_t part_n = (uint64_t)nparts * passes * radix;
    int rc;
    if ((rc = ctx->sort_tmp[2].ensure((part_n + (uint64_t)passes * radix + OS_MAX_PASSES) * 4 + 64)) != MH_OK)
        return rc;
    uint32_t *part = ctx->sort_tmp[2].as<uint32_t>();
    uint32_t *dstart = part + part_n;
    uint32_t *tickets = dstart + (uint64_t)passes * radix;
    // the look-back's status words (and the error word in front of them):
    // zeroed when (re)allocated, then never cleared -- each pass has an epoch
    const uint64_t status_bytes = 256 + (uint64_t)ntiles * radix * 8;
    const void *old_ptr = ctx->sort_status.ptr;
    if ((rc = ctx->sort_status.ensure(status_bytes)) != MH_OK) return rc;
    if (ctx->sort_status.ptr != old_ptr) {
        MH_HIP(hipMemsetAsync(ctx->sort_status.ptr, 0, ctx->sort_status.bytes, ctx->stream));
        ctx->sort_epoch = 0;
    }
    int *err = ctx->sort_status.as<int>();
    uint64_t *status = reinterpret_cast<uint64_t *>(ctx->sort_status.as<uint8_t>() + 256);
    hipLaunchKernelGGL(os_hist_kernel, dim3(nparts), dim3(OS_HIST_THREADS), 0, ctx->stream, k0, n, db, passes, part,
                       tickets);
    hipLaunchKernelGGL(os_start_kernel, dim3(passes), dim3(radix), 0, ctx->stream, part, (int)nparts, passes, dstart);
    for (int pass = 0; pass < passes; ++pass) {
        if (++ctx->sort_epoch >= (1u << 24)) {
            MH_HIP(hipMemsetAsync(ctx->sort_status.ptr, 0, ctx->sort_status.bytes, ctx->stream));
            ctx->sort_epoch = 1;
        }
        const uint32_t *ds = dstart + (uint64_t)pass * radix;
        const int shift = db * pass;
        if (db == 8)
            hipLaunchKernelGGL(os_pass_kernel<8>, dim3(ntiles), dim3(SORT_THREADS), 0, ctx->stream, k0, v0, k1, v1, n,
                               shift, ds, status, tickets + pass, ctx->sort_epoch, err);
        else if (db == 9)
            hipLaunchKernelGGL(os_pass_kernel<9>, dim3(ntiles), dim3(SORT_THREADS), 0, ctx->stream, k0, v0, k1, v1, n,
                               shift, ds, status, tickets + pass, ctx->sort_epoch, err);
        else
            hipLaunchKernelGGL(os_pass_kernel<10>, dim3(ntiles), dim3(SORT_THREADS), 0, ctx->stream, k0, v0, k1, v1,
                               n, shift, ds, status, tickets + pass, ctx->sort_epoch, err);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    MH_HIP(hipGetLastError());
    *out_k = k0;
    *out_v = v0;
    return MH_OK;
}

int sort_check(mh_context *ctx)
{
    if (!ctx->sort_status.ptr) return MH_OK;
    int err = 0;
    MH_HIP(hipMemcpyAsync(&err, ctx->sort_status.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    if (err) {
        MH_HIP(hipMemsetAsync(ctx->sort_status.ptr, 0, 4, ctx->stream));
        MH_CHECK(false, MH_ERR_DEVICE, "radix sort: a tile's look-back wait ran past its limit");
    }
    return MH_OK;
}

int sml_seed_mers(mh_context *ctx, const mh_sml *s, const uint32_t *d_pos, uint64_t m, uint64_t *d_out)
{
    if (!m) return MH_OK;
    hipLaunchKernelGGL(gather_keys_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, ctx->stream,
                       s->d_keys_pos, s->n, d_pos, m, d_out);
    MH_HIP(hipGetLastError());
    return MH_OK;
}

int sml_find_mer_dev(mh_context *ctx, const mh_sml *s, uint64_t mer, uint64_t *index)
{
    uint64_t *d = ctx->sort_tmp[3].as<uint64_t>();
    hipLaunchKernelGGL(find_mer_kernel, dim3(1), dim3(1), 0, ctx->stream, s->d_keys, s->len - s->seed_length, mer,
                       s->seed_weight, d);
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(index, d, 8, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int ascii_to_codes(mh_context *ctx, const void *d_ascii, uint64_t len, uint8_t *d_codes)
{
    if (!len) return MH_OK;
    int rc;
    if ((rc = ctx->sort_tmp[3].ensure(64)) != MH_OK) return rc;
    int *d_bad = ctx->sort_tmp[3].as<int>();
    MH_HIP(hipMemsetAsync(d_bad, 0, 4, ctx->stream));
    hipLaunchKernelGGL(codes_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, ctx->stream,
                       static_cast<const uint8_t *>(d_ascii), len, d_codes, d_bad);
    MH_HIP(hipGetLastError());
    int bad = 0;
    MH_HIP(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    MH_CHECK(bad == 0, MH_ERR_GAP_CHAR, "gap character in a DP operand");
    return MH_OK;
}

uint64_t sml_full_key(uint64_t c, int sw) { return full_key(c, sw); }
uint64_t sml_compact_key(uint64_t k, int sw) { return compact_key(k, sw); }

}  // namespace mh
