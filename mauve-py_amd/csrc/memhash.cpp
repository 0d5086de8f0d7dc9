// memhash.cpp -- MemHash insertion replay and the SML / match-finding C ABI.
//
// The reference inserts every seed into one of 40000 buckets, chosen by the
// seed's generalized offset, with std::lower_bound under MheCompare; a seed
// equivalent to the entry found there (containment either way) is dropped
// as a collision, otherwise it is extended and inserted
// (LM/MemHash.cpp:209-251).  MheCompare is not a strict weak order, so the
// result depends on the exact probe sequence of lower_bound and on the
// insertion history of the bucket; both are reproduced here literally.
// Buckets whose outcome is fixed in closed form (one diagonal or
// anti-diagonal, odd solid seed) were resolved on the GPU; the others never interact, so
// they are replayed in parallel host threads, each in the seeds' HashMatch
// order.  Extensions were computed on the GPU.
#include "mh_internal.hpp"
#include "mhe.hpp"
#include <algorithm>
#include <atomic>
#include <thread>
#include <cstring>

namespace mh {

namespace {

constexpr uint32_t TABLE_SIZE = 40000;   // DEFAULT_MEM_TABLE_SIZE, LM/MemHash.h:30; bucket ids are below it

// A two-genome MatchHashEntry: both starts always present in pairwise seeds.
struct Ent {
    int64_t s0, s1;
    uint64_t len;
    int64_t off;
    uint64_t mersize;   // seed length for candidates, 0 for stored copies
};

// MatchHashEntry::Contains (a contains b), LM/MatchHashEntry.cpp:164-200
inline bool contains(const Ent &a, const Ent &b)
{
    if (a.off != b.off) return false;
    const int64_t diff = b.s0 - a.s0;
    if (diff < 0 || a.len < b.len + (uint64_t)diff) return false;
    const int64_t diff_rc = (int64_t)b.len - (int64_t)a.len + diff;
    const int64_t di = b.s1 - a.s1;
    if (b.s1 < 0 && diff_rc == di) return true;
    return diff == di;
}

// strict_start_lessthan_ptr, LM/MatchHashEntry.cpp:48-67
inline bool strict_less(const Ent &a, const Ent &b)
{
    int64_t as = a.s0, bs = b.s0;
    if (as < 0) as = (int64_t)((uint64_t)(-as) + a.len - a.mersize);
    if (bs < 0) bs = (int64_t)((uint64_t)(-bs) + b.len - b.mersize);
    if (as != bs) return as < bs;
    as = a.s1;
    bs = b.s1;
    if (as < 0) as = (int64_t)((uint64_t)(-as) + a.len - a.mersize);
    if (bs < 0) bs = (int64_t)((uint64_t)(-bs) + b.len - b.mersize);
    if (as != bs) return as < bs;
    return false;
}

// MheCompare, LM/MatchHashEntry.h:121-143, with both genomes present in both
// entries (the FirstStart and genome-presence tests never decide)
inline bool mhe_less(const Ent &a, const Ent &b)
{
    if (contains(a, b) || contains(b, a)) return false;
    return strict_less(a, b);
}

// libstdc++ std::lower_bound probe sequence
inline size_t lower_bound_ent(const std::vector<Ent> &v, const Ent &q)
{
    size_t first = 0, len = v.size();
    while (len > 0) {
        const size_t half = len >> 1, mid = first + half;
        if (mhe_less(v[mid], q)) {
            first = mid + 1;
            len = len - half - 1;
        } else {
            len = half;
        }
    }
    return first;
}

template <class E> inline size_t lower_bound_k(const std::vector<E> &v, const E &q, int k)
{
    return lower_bound_at(v.size(), q, k, [&](size_t i) -> const E & { return v[i]; });
}

// b grown to at least need bytes, its first keep bytes kept.
int grow_keep(mh_context *ctx, DevBuf &b, size_t keep, size_t need)
{
    if (need <= b.bytes && b.ptr) return MH_OK;
    DevBuf nb;
    int rc;
    if ((rc = nb.ensure(need + need / 2)) != MH_OK) return rc;
    if (keep) MH_HIP(hipMemcpyAsync(nb.ptr, b.ptr, keep, hipMemcpyDeviceToDevice, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    std::swap(b.ptr, nb.ptr);
    std::swap(b.bytes, nb.bytes);
    return MH_OK;   // nb now holds the old buffer and frees it
}

// PairwiseMatchFinder over three or four genomes (LM/PairwiseMatchFinder.cpp:
// 37-71 with LM/MatchFinder.cpp:172-340 and LM/MemHash.cpp:167-251): every
// run of one masked mer hashes each pair of genomes that hold it exactly once,
// pairs in genome order, runs in mer order.  A pair's seeds and their
// extensions are the two-genome search's (ExtendMatch reads only the genomes
// a seed holds); the insertion into the 40000 buckets is replayed on the
// host for every seed, buckets in parallel, and the list comes out bucket by
// bucket (LM/MemHash.h:182-203).
int find_matches_multi(mh_context *ctx, const mh_sml *const *smls, int k, std::vector<mh_match> &out,
                       mh_match_stats *st)
{
    for (int i = 0; i < k; ++i) {
        MH_CHECK(smls[i], MH_ERR_INVALID, "null sorted mer list");
        MH_CHECK(smls[i]->seed == smls[0]->seed, MH_ERR_INVALID, "Different seed patterns.");
    }
    int rc;
    LiveRanges live;
    if ((rc = search_live_ranges(ctx, smls, k, &live)) != MH_OK) return rc;
    if (st) std::memset(st, 0, sizeof(*st));
    if (st) st->repeat_runs = live.restarts;
    const uint64_t sl = (uint64_t)smls[0]->seed_length;
    // every pair's seeds on the device, with key = mer << 3 | pair (pairs
    // numbered in genome order), then one sort: HashMatch order
    DevBuf *mb = ctx->multi_scratch;
    const bool solid_odd = smls[0]->seed_length == smls[0]->seed_weight && (smls[0]->seed_length & 1);
    std::vector<CleanEnt> clean_p[6];
    std::vector<uint8_t> replay_bucket(TABLE_SIZE, 0);
    uint64_t total = 0;
    int64_t walks = 0, long_walks = 0;
    int pair_a[6], pair_b[6], np = 0;
    for (int a = 0; a < k; ++a)
        for (int b = a + 1; b < k; ++b) {
            const SeedRec *d_seeds = nullptr;
            uint64_t ns = 0;
            if ((rc = join_and_extend(ctx, smls[a], smls[b], &live, a, b, &d_seeds, &ns)) != MH_OK) return rc;
            walks += (int64_t)ctx->last_walks;
            long_walks += (int64_t)ctx->last_long_walks;
            pair_a[np] = a;
            pair_b[np] = b;
            if (ns) {
                // the pair's buckets that the two-genome closed form settles
                std::vector<DirtySeed> dirty_p;
                uint64_t ncs = 0;
                if ((rc = partition_seeds(ctx, d_seeds, ns, (int)sl, solid_odd, smls[a]->len, clean_p[np], dirty_p,
                                          &ncs)) != MH_OK)
                    return rc;
                for (const DirtySeed &d : dirty_p) replay_bucket[d.bucket] = 1;
                if ((rc = grow_keep(ctx, mb[0], total * sizeof(SeedRec), (total + ns) * sizeof(SeedRec) + 64)) != MH_OK)
                    return rc;
                if ((rc = grow_keep(ctx, mb[1], total * 8, (total + ns) * 8 + 64)) != MH_OK) return rc;
                MH_HIP(hipMemcpyAsync(mb[0].as<SeedRec>() + total, d_seeds, ns * sizeof(SeedRec),
                                      hipMemcpyDeviceToDevice, ctx->stream));
                if ((rc = seed_mers(ctx, smls[a], d_seeds, ns, (uint32_t)np, mb[1].as<uint64_t>() + total)) != MH_OK)
                    return rc;
                total += ns;
            }
            ++np;
        }
    const double tg1 = now_ms();
    // only the seeds of buckets to replay come back to the host
    std::vector<SeedRec> hs;
    std::vector<uint8_t> hp;
    if (total) {
        SeedRec *d_sorted = nullptr;
        uint64_t *d_skeys = nullptr;
        if ((rc = merge_pair_seeds(ctx, mb[0].as<SeedRec>(), mb[1].as<uint64_t>(), total, &d_sorted, &d_skeys)) !=
            MH_OK)
            return rc;
        if ((rc = select_bucket_seeds(ctx, d_sorted, d_skeys, total, replay_bucket.data(), (int)sl, hs, hp)) != MH_OK)
            return rc;
    }
    struct Tup {
        int a, b;
        SeedRec r;
    };
    std::vector<Tup> tup(hs.size());
    for (uint64_t x = 0; x < hs.size(); ++x) tup[x] = Tup{pair_a[hp[x]], pair_b[hp[x]], hs[x]};
    const uint64_t settled_seeds = total - hs.size();
    // A bucket is settled in closed form when every genome pair's seeds in it
    // are (a single line, odd solid seed): MheCompare orders entries by
    // genome presence before anything else, so each pair's entries form a
    // block (pairs with the later first genome first, then the later second
    // genome), the comparisons a search makes against other blocks are
    // consistent, and within its block a seed meets only its own line, where
    // the two-genome result holds.  Any other bucket is replayed literally with
    // the seeds of every pair, in HashMatch order.
    std::vector<uint32_t> bucket(tup.size()), cnt(TABLE_SIZE + 1, 0);
    std::vector<EntK> seed(tup.size());
    for (size_t x = 0; x < tup.size(); ++x) {
        EntK c;
        std::memset(&c, 0, sizeof(c));
        c.s[tup[x].a] = tup[x].r.s0;
        c.s[tup[x].b] = tup[x].r.s1;
        c.len = sl;
        c.mersize = sl;
        c.first = tup[x].a;
        c.off = offset_k(c, k);
        seed[x] = c;
        const int64_t m = c.off % (int64_t)TABLE_SIZE;
        bucket[x] = (uint32_t)(m < 0 ? m + TABLE_SIZE : m);
        ++cnt[bucket[x] + 1];
    }
    for (uint32_t q = 0; q < TABLE_SIZE; ++q) cnt[q + 1] += cnt[q];
    std::vector<uint32_t> order(cnt[TABLE_SIZE]);
    {
        std::vector<uint32_t> at(cnt.begin(), cnt.end() - 1);
        for (size_t x = 0; x < tup.size(); ++x) order[at[bucket[x]]++] = (uint32_t)x;
    }
    std::vector<std::vector<EntK>> table(TABLE_SIZE);
    std::atomic<int64_t> collisions{0};
    const uint32_t per = 64;
    host_parallel((TABLE_SIZE + per - 1) / per, [&](size_t t) {
        int64_t coll = 0;
        for (uint32_t q = (uint32_t)t * per; q < std::min<uint32_t>(TABLE_SIZE, (uint32_t)(t + 1) * per); ++q) {
            std::vector<EntK> &v = table[q];
            for (uint32_t y = cnt[q]; y < cnt[q + 1]; ++y) {
                const uint32_t x = order[y];
                const EntK &c = seed[x];
                size_t it = lower_bound_k(v, c, k);
                if (it != v.size() && !mhe_less_k(c, v[it], k)) {   // !comp(*it, c) holds at lower_bound
                    ++coll;
                    continue;
                }
                EntK e = c;
                const SeedRec &r = tup[x].r;
                e.s[tup[x].a] = r.e0;
                e.s[tup[x].b] = r.e1;
                e.len = r.elen;
                e.mersize = 0;       // MatchHashEntry::operator= zeroes m_mersize; the offset is kept
                it = lower_bound_k(v, e, k);
                v.insert(v.begin() + (ptrdiff_t)it, e);
            }
        }
        collisions += coll;
    });
    host_time(ctx, "host_replay", now_ms() - tg1);
    // pair blocks of a settled bucket in MheCompare order
    std::vector<int> block(np);
    for (int p = 0; p < np; ++p) block[p] = p;
    std::sort(block.begin(), block.end(), [&](int x, int y) {
        return pair_a[x] != pair_a[y] ? pair_a[x] > pair_a[y] : pair_b[x] > pair_b[y];
    });
    std::vector<size_t> ci(np, 0);
    uint64_t settled_entries = 0;
    out.clear();
    auto emit = [&](const int64_t *s4, uint64_t len) {
        mh_match m;
        std::memset(&m, 0, sizeof(m));
        for (int i = 0; i < k; ++i) m.start[i] = s4[i];
        m.length = len;
        out.push_back(m);
    };
    for (uint32_t q = 0; q < TABLE_SIZE; ++q) {
        for (int p = 0; p < np; ++p)   // each pair's entries of this bucket (used only when settled)
            while (ci[p] < clean_p[p].size() && clean_p[p][ci[p]].bucket < q) ++ci[p];
        if (replay_bucket[q]) {
            for (const EntK &e : table[q]) emit(e.s, e.len);
            continue;
        }
        for (int p : block) {
            for (size_t y = ci[p]; y < clean_p[p].size() && clean_p[p][y].bucket == q; ++y) {
                int64_t s4[4] = {0, 0, 0, 0};
                s4[pair_a[p]] = clean_p[p][y].e0;
                s4[pair_b[p]] = clean_p[p][y].e1;
                emit(s4, clean_p[p][y].elen);
                ++settled_entries;
            }
        }
    }
    collisions += (int64_t)(settled_seeds - settled_entries);
    if (st) {
        st->seed_tuples = (int64_t)total;
        st->collisions = collisions.load();
        st->matches = (int64_t)out.size();
        st->walks = walks;
        st->long_walks = long_walks;
    }
    return MH_OK;
}

}  // namespace

int find_matches_pairwise(mh_context *ctx, const mh_sml *const *smls, int k, std::vector<mh_match> &out,
                          mh_match_stats *st)
{
    MH_CHECK(k >= 2 && k <= MH_MAX_SEQ, MH_ERR_UNSUPPORTED, "this build matches two to four genomes");
    if (k > 2) return find_matches_multi(ctx, smls, k, out, st);
    const mh_sml *a = smls[0], *b = smls[1];
    MH_CHECK(a && b, MH_ERR_INVALID, "null sorted mer list");
    MH_CHECK(a->seed == b->seed, MH_ERR_INVALID, "Different seed patterns.");   // LM/MatchFinder.cpp:190-194
    const SeedRec *d_seeds = nullptr;
    uint64_t nseeds = 0;
    LiveRanges live;
    int rc = search_live_ranges(ctx, smls, 2, &live);
    if (rc != MH_OK) return rc;
    if ((rc = join_and_extend(ctx, a, b, &live, 0, 1, &d_seeds, &nseeds)) != MH_OK) return rc;
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->repeat_runs = live.restarts;
        st->walks = (int64_t)ctx->last_walks;
        st->long_walks = (int64_t)ctx->last_long_walks;
    }
    const int64_t sl = a->seed_length;
    const bool solid_odd = a->seed_length == a->seed_weight && (a->seed_length & 1);
    std::vector<CleanEnt> clean;
    std::vector<DirtySeed> dirty;
    uint64_t n_clean_seeds = 0;
    if ((rc = partition_seeds(ctx, d_seeds, nseeds, (int)sl, solid_odd, a->len, clean, dirty, &n_clean_seeds)) != MH_OK)
        return rc;
    const double tg1 = now_ms();
    // the remaining buckets: runs of equal bucket in the dirty list, each in HashMatch order
    std::vector<size_t> run_lo;
    for (size_t y = 0; y < dirty.size(); ++y)
        if (y == 0 || dirty[y].bucket != dirty[y - 1].bucket) run_lo.push_back(y);
    const size_t nruns = run_lo.size();
    run_lo.push_back(dirty.size());
    std::vector<std::vector<Ent>> table(nruns);
    std::vector<uint32_t> busy(nruns);
    for (size_t q = 0; q < nruns; ++q) busy[q] = (uint32_t)q;
    std::sort(busy.begin(), busy.end(),
              [&](uint32_t x, uint32_t y) { return run_lo[x + 1] - run_lo[x] > run_lo[y + 1] - run_lo[y]; });
    std::atomic<size_t> next{0};
    std::atomic<int64_t> collisions{0};
    auto work = [&]() {
        int64_t coll = 0;
        for (;;) {
            const size_t w = next.fetch_add(1);
            if (w >= busy.size()) break;
            const uint32_t q = busy[w];
            std::vector<Ent> &v = table[q];
            for (size_t y = run_lo[q]; y < run_lo[q + 1]; ++y) {
                const SeedRec &r = dirty[y].r;
                Ent c;
                c.s0 = r.s0;
                c.s1 = r.s1;
                c.len = (uint64_t)sl;
                c.off = (int64_t)r.s1 - r.s0 - (r.s1 < 0 ? sl : 0);   // CalculateOffset
                c.mersize = (uint64_t)sl;
                size_t it = lower_bound_ent(v, c);
                if (it != v.size() && !mhe_less(c, v[it])) {   // !comp(*it, c) holds at lower_bound
                    ++coll;
                    continue;
                }
                Ent e;
                e.s0 = r.e0;
                e.s1 = r.e1;
                e.len = r.elen;
                e.off = c.off;       // the offset is not recomputed after ExtendMatch
                e.mersize = 0;       // MatchHashEntry::operator= zeroes m_mersize
                it = lower_bound_ent(v, e);
                v.insert(v.begin() + (ptrdiff_t)it, e);
            }
        }
        collisions += coll;
    };
    const unsigned nt = (unsigned)std::min<size_t>(host_threads(), std::max<size_t>(1, busy.size()));
    host_parallel(nt, [&](size_t) { work(); });
    const double tr1 = now_ms();
    host_time(ctx, "host_replay", tr1 - tg1);
    // GetMatchList order: bucket by bucket (LM/MemHash.h:182-203); clean and
    // replayed buckets are disjoint sets of buckets, both already in bucket order
    out.clear();
    out.reserve(clean.size() + dirty.size());
    auto emit = [&](int64_t s0, int64_t s1, uint64_t len) {
        mh_match m;
        std::memset(&m, 0, sizeof(m));
        m.start[0] = s0;
        m.start[1] = s1;
        m.length = len;
        out.push_back(m);
    };
    size_t ci = 0;
    for (size_t q = 0; q <= nruns; ++q) {
        const uint32_t bq = q < nruns ? dirty[run_lo[q]].bucket : TABLE_SIZE;
        for (; ci < clean.size() && clean[ci].bucket < bq; ++ci) emit(clean[ci].e0, clean[ci].e1, clean[ci].elen);
        if (q < nruns)
            for (const Ent &e : table[q]) emit(e.s0, e.s1, e.len);
    }
    host_time(ctx, "host_list", now_ms() - tr1);
    if (st) {
        st->seed_tuples = (int64_t)nseeds;
        st->collisions = collisions.load() + (int64_t)(n_clean_seeds - clean.size());
        st->matches = (int64_t)out.size();
    }
    return MH_OK;
}

// UniqueMatchFinder over 2 to 16 genomes (MA/UniqueMatchFinder.cpp:36-60 with
// LM/MatchFinder.cpp:172-340 and LM/MemHash.cpp:167-251): SearchRange's live
// ranges, then the seeds, their extensions and the bucket insertion on the
// device (unique_seeds); the list comes out bucket by bucket
// (LM/MemHash.h:182-203).
int find_matches_unique(mh_context *ctx, const mh_sml *const *smls, int k, std::vector<int64_t> &starts,
                        std::vector<uint64_t> &lengths, mh_match_stats *st)
{
    MH_CHECK(k >= 2 && k <= 16, MH_ERR_UNSUPPORTED, "UniqueMatchFinder takes 2 to 16 genomes");
    for (int i = 0; i < k; ++i) {
        MH_CHECK(smls[i], MH_ERR_INVALID, "null sorted mer list");
        MH_CHECK(smls[i]->seed == smls[0]->seed, MH_ERR_INVALID, "Different seed patterns.");
    }
    int rc;
    LiveRanges live;
    if ((rc = search_live_ranges(ctx, smls, k, &live)) != MH_OK) return rc;
    UniqueSeeds us;
    if ((rc = unique_seeds(ctx, smls, k, live, &us)) != MH_OK) return rc;
    starts.swap(us.starts);
    lengths.swap(us.lengths);
    if (st) {
        std::memset(st, 0, sizeof(*st));
        st->seed_tuples = (int64_t)us.n;
        st->collisions = us.collisions;
        st->matches = (int64_t)lengths.size();
        st->repeat_runs = live.restarts;
        st->walks = (int64_t)us.walks;
    }
    return MH_OK;
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_sml_create_device(mh_context *ctx, const void *d_seq, uint64_t len, uint64_t seed, mh_sml **out)
{
    MH_CHECK(ctx && out && (d_seq || len == 0), MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    mh_sml *s = new mh_sml();
    s->ctx = ctx;
    ctx->smls.insert(s);
    int rc = sml_build(ctx, d_seq, len, seed, s);
    if (rc != MH_OK) {
        mh_sml_free(s);
        return rc;
    }
    *out = s;
    return MH_OK;
}

int mh_sml_create(mh_context *ctx, const char *seq, uint64_t len, uint64_t seed, mh_sml **out)
{
    MH_CHECK(ctx && out && (seq || len == 0), MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    void *d = nullptr;
    MH_HIP(hipMalloc(&d, len + 16));
    hipError_t e = hipMemcpyAsync(d, seq, len, hipMemcpyHostToDevice, ctx->stream);
    int rc = e == hipSuccess ? mh_sml_create_device(ctx, d, len, seed, out) : MH_ERR_DEVICE;
    if (e != hipSuccess) set_error(std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    return rc;
}

void mh_sml_free(mh_sml *s)
{
    if (!s) return;
    SmlBufs b;
    b.words = s->d_words;
    b.keys_pos = s->d_keys_pos;
    b.keys = s->d_keys;
    b.pos = s->d_pos;
    b.cap_words = s->cap_words;
    b.cap_n = s->cap_n;
    mh_context *ctx = s->ctx;
    if (ctx) {
        // the buffers stay with the context for the next build (a few sets at
        // most); work on them is ordered on the context stream, so no wait
        ctx->smls.erase(s);
        if (b.words && b.keys_pos && b.keys && b.pos && ctx->sml_pool.size() < 4) {
            ctx->sml_pool.push_back(b);
            b = SmlBufs();
        }
    }
    b.release();
    delete s;
}

int mh_sml_get_header(const mh_sml *s, mh_sml_header *h)
{
    MH_CHECK(s && h, MH_ERR_INVALID, "null argument");
    std::memset(h, 0, sizeof(*h));
    h->version = 5;
    h->alphabet_bits = 2;
    h->seed = s->seed;
    h->seed_length = (uint32_t)s->seed_length;
    h->seed_weight = (uint32_t)s->seed_weight;
    h->length = s->len;
    h->sml_length = s->n;
    h->circular = 0;
    return MH_OK;
}

int mh_sml_read(mh_context *ctx, const mh_sml *s, uint64_t offset, uint64_t count, mh_bmer *out, uint64_t *n_read)
{
    MH_CHECK(ctx && s && n_read && (out || count == 0), MH_ERR_INVALID, "null argument");
    *n_read = 0;
    if (offset >= s->n) return MH_OK;
    const uint64_t m = std::min<uint64_t>(count, s->n - offset);
    std::vector<uint64_t> ck(m);
    std::vector<uint32_t> ps(m);
    MH_HIP(hipMemcpyAsync(ck.data(), s->d_keys + offset, m * 8, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipMemcpyAsync(ps.data(), s->d_pos + offset, m * 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t x = 0; x < m; ++x) {
        out[x].position = ps[x];
        out[x].mer = sml_full_key(ck[x], s->seed_weight);
    }
    *n_read = m;
    return MH_OK;
}

int mh_sml_get_seed_mers(mh_context *ctx, const mh_sml *s, const uint32_t *positions, uint64_t n, uint64_t *out)
{
    MH_CHECK(ctx && s && (n == 0 || (positions && out)), MH_ERR_INVALID, "null argument");
    for (uint64_t x = 0; x < n; ++x)
        MH_CHECK(positions[x] < s->n, MH_ERR_INVALID, "position beyond SMLLength");
    if (!n) return MH_OK;
    void *dp = nullptr, *dk = nullptr;
    MH_HIP(hipMalloc(&dp, n * 4));
    if (hipMalloc(&dk, n * 8) != hipSuccess) {
        (void)hipFree(dp);
        set_error("device allocation failed");
        return MH_ERR_NOMEM;
    }
    int rc = MH_OK;
    if (hipMemcpyAsync(dp, positions, n * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) rc = MH_ERR_DEVICE;
    if (rc == MH_OK) rc = sml_seed_mers(ctx, s, (const uint32_t *)dp, n, (uint64_t *)dk);
    if (rc == MH_OK && hipMemcpyAsync(out, dk, n * 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        rc = MH_ERR_DEVICE;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(dp);
    (void)hipFree(dk);
    if (rc == MH_ERR_DEVICE) set_error("copy failed in mh_sml_get_seed_mers");
    return rc;
}

int mh_sml_find_mer(mh_context *ctx, const mh_sml *s, uint64_t mer, uint64_t *index, int *found)
{
    MH_CHECK(ctx && s && index && found, MH_ERR_INVALID, "null argument");
    *found = 0;
    // LM/SortedMerList.cpp:173-175: an empty or too-short list reports "not found"
    if (s->len == 0 || s->len < (uint64_t)s->seed_length) return MH_OK;
    uint64_t idx = 0;
    int rc = sml_find_mer_dev(ctx, s, mer, &idx);
    if (rc != MH_OK) return rc;
    uint64_t ck = 0;
    MH_HIP(hipMemcpyAsync(&ck, s->d_keys + idx, 8, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    *index = idx;
    *found = sml_full_key(ck, s->seed_weight) == mer;
    return MH_OK;
}

int mh_sml_get_sequence(mh_context *ctx, const mh_sml *s, uint32_t *words, uint64_t n_words)
{
    MH_CHECK(ctx && s && words, MH_ERR_INVALID, "null argument");
    const uint64_t m = std::min<uint64_t>(n_words, s->n_words);
    std::memset(words, 0, n_words * 4);
    MH_HIP(hipMemcpyAsync(words, s->d_words, m * 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

// the radix sorts' look-back bound of a search's contexts (its own and the
// helpers that ran its independent searches)
static int sorts_ok(mh_context *ctx)
{
    int rc = sort_check(ctx);
    for (mh_context *h : ctx->helpers)
        if (rc == MH_OK) rc = sort_check(h);
    return rc;
}

int mh_find_matches_pairwise(mh_context *ctx, const mh_sml *const *smls, int k, mh_match *out, uint64_t cap,
                             uint64_t *n_out, mh_match_stats *stats)
{
    MH_CHECK(ctx && smls && n_out && (out || cap == 0), MH_ERR_INVALID, "null argument");
    MH_CHECK(k >= 1 && k <= MH_MAX_SEQ, MH_ERR_INVALID, "genome count out of range");
    MH_HIP(hipSetDevice(ctx->device));
    std::vector<mh_match> v;
    int rc = find_matches_pairwise(ctx, smls, k, v, stats);
    if (rc == MH_OK) rc = sorts_ok(ctx);
    if (rc != MH_OK) return rc;
    *n_out = v.size();
    const uint64_t m = std::min<uint64_t>(cap, v.size());
    if (m) std::memcpy(out, v.data(), m * sizeof(mh_match));
    return MH_OK;
}

}  // extern "C"

extern "C" {

int mh_find_matches_pairwise_list(mh_context *ctx, const mh_sml *const *smls, int k, mh_match_list **out,
                                  mh_match_stats *stats)
{
    MH_CHECK(ctx && smls && out, MH_ERR_INVALID, "null argument");
    MH_CHECK(k >= 1 && k <= MH_MAX_SEQ, MH_ERR_INVALID, "genome count out of range");
    MH_HIP(hipSetDevice(ctx->device));
    mh_match_list *l = new mh_match_list();
    int rc = find_matches_pairwise(ctx, smls, k, l->v, stats);
    if (rc == MH_OK) rc = sorts_ok(ctx);
    if (rc != MH_OK) {
        delete l;
        return rc;
    }
    l->seq_count = k;
    l->files.assign(k, std::string());
    for (int q = 0; q < k; ++q) l->lengths.push_back(smls[q]->len);
    *out = l;
    return MH_OK;
}

struct mh_match_k_list {
    int k = 0;
    std::vector<int64_t> starts;
    std::vector<uint64_t> lengths;
};

int mh_find_matches_unique(mh_context *ctx, const mh_sml *const *smls, int k, mh_match_k_list **out,
                           mh_match_stats *stats)
{
    MH_CHECK(ctx && smls && out, MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    mh_match_k_list *l = new mh_match_k_list();
    int rc = find_matches_unique(ctx, smls, k, l->starts, l->lengths, stats);
    if (rc == MH_OK) rc = sorts_ok(ctx);
    if (rc != MH_OK) {
        delete l;
        return rc;
    }
    l->k = k;
    *out = l;
    return MH_OK;
}

uint64_t mh_match_k_list_size(const mh_match_k_list *l) { return l ? l->lengths.size() : 0; }
int mh_match_k_list_genomes(const mh_match_k_list *l) { return l ? l->k : 0; }
const int64_t *mh_match_k_list_starts(const mh_match_k_list *l)
{
    return l && !l->starts.empty() ? l->starts.data() : nullptr;
}
const uint64_t *mh_match_k_list_lengths(const mh_match_k_list *l)
{
    return l && !l->lengths.empty() ? l->lengths.data() : nullptr;
}
void mh_match_k_list_free(mh_match_k_list *l) { delete l; }

uint64_t mh_match_list_size(const mh_match_list *l) { return l ? l->v.size() : 0; }
const mh_match *mh_match_list_data(const mh_match_list *l) { return l && !l->v.empty() ? l->v.data() : nullptr; }
void mh_match_list_free(mh_match_list *l) { delete l; }

}  // extern "C"
