// restart.cpp -- the repeat-limit restarts of MatchFinder::SearchRange
// (LM/MatchFinder.cpp:166-340, GetBreakpoint 89-126, FindMatchSeeds 139-164),
// planned on the host as the index ranges of each sorted list that the search
// reads.
//
// SearchRange merges the k sorted lists by masked mer and hands every run of
// one masked mer to EnumerateMatches.  When, at the top of an iteration whose
// head still carries the current run's mer, the run already holds more than
// MER_REPEAT_LIMIT entries, the run is dropped and the search starts again
// from new start points: FindMer of the next masked mer in the first list
// where it lands inside the list, GetBreakpoint from there, each start point
// kept no earlier than where its list had got to, and the lists before that
// first one closed.  Between two restarts the search is an ordinary merge, so
// its outcome is fixed by the ranges [start point, first entry of the dropped
// run) of each list, and after the last restart [start point, end): every run
// of those ranges is enumerated, each masked mer inside one range of each
// list.  The join on the device reads only these ranges.
//
// Whether a long run restarts depends on where the checks fall:
//  * the iteration of each block but the first (a block is one list's entries
//    of the mer), with every earlier block counted;
//  * the iteration after a buffer of MER_BUFFER_SIZE entries runs out inside
//    or at the end of a block (the list head keeps the run's mer over the
//    refill), unless the list is exhausted.
// Buffers of list g end at start_point[g] + 10000 t.  The order of the blocks
// is the order of the heads in the sorted idmer list: a head inserted later
// comes before heads of equal mer already there (329-331), and heads from the
// initial, stable sort of the first entries come last, in list order.  A head
// for mer m in list g is inserted when g's block of its previous mer p is
// done, so heads with a larger p come first, and heads with the same p come in
// the reverse of their order in run p.
#include "mh_internal.hpp"

#include <algorithm>
#include <cstring>

namespace mh {

namespace {

constexpr uint64_t MER_REPEAT_LIMIT = 1000;    // LM/MatchFinder.cpp:166
constexpr uint64_t MER_BUFFER_SIZE = 10000;    // LM/MatchFinder.cpp:175

struct Planner {
    const PlanLists &L;
    std::vector<uint64_t> sp;   // start points of the current search

    explicit Planner(const PlanLists &l) : L(l), sp(l.k, 0) {}

    uint64_t mk(int g, uint64_t i) const { return L.ck[g][i] >> 1; }

    // first index of list g (within [0, end)) whose masked mer is >= m
    uint64_t lower_m(int g, uint64_t m, uint64_t end) const
    {
        uint64_t lo = 0, len = end;
        while (len > 0) {
            const uint64_t half = len >> 1, mid = lo + half;
            if (mk(g, mid) < m) { lo = mid + 1; len -= half + 1; }
            else len = half;
        }
        return lo;
    }
    uint64_t upper_m(int g, uint64_t m) const
    {
        uint64_t lo = 0, len = L.n[g];
        while (len > 0) {
            const uint64_t half = len >> 1, mid = lo + half;
            if (mk(g, mid) <= m) { lo = mid + 1; len -= half + 1; }
            else len = half;
        }
        return lo;
    }
    // first entry of the run of masked mer p that ends at index last
    uint64_t run_start(int g, uint64_t p, uint64_t last) const
    {
        uint64_t i = last;
        for (int s = 0; s < 32; ++s) {
            if (i == 0 || mk(g, i - 1) != p) return i;
            --i;
        }
        return lower_m(g, p, i);
    }
    uint64_t block_start(int g, uint64_t p, uint64_t last) const { return std::max(sp[g], run_start(g, p, last)); }

    // SortedMerList::FindMer with its bsearch, LM/SortedMerList.cpp:170-179,
    // 380-394, on compact keys (ordered as the 64-bit keys): 1 found, 0 not
    // found (result = where the search stopped), -1 a list shorter than the
    // seed (result untouched).
    int find_mer(int g, uint64_t q, uint64_t &result) const
    {
        uint64_t last = L.len[g];
        if (last == 0 || last < (uint64_t)L.sl) return -1;
        last -= (uint64_t)L.sl;
        uint64_t start = 0, end = last;
        for (;;) {
            const uint64_t mid = (start + end) / 2;
            const uint64_t mm = L.ck[g][mid];
            if (mm == q) { result = mid; break; }
            if (mm < q && mid < end) { start = mid + 1; continue; }
            if (mm > q && start < mid) { end = mid - 1; continue; }
            result = mid;
            break;
        }
        return L.ck[g][result] == q ? 1 : 0;
    }

    // GetBreakpoint, LM/MatchFinder.cpp:89-126, with its two quirks: in the
    // other lists the backward scan compares the masked entry with the
    // logical value (break_mer.mer && mer_mask) and re-reads the same entry,
    // so it runs (to index -1) only when break_mer's key is 0; otherwise the
    // start is one past the entry FindMer found (cur_matchI + 1).
    int breakpoint(int sar, uint64_t start, std::vector<uint64_t> &bp) const
    {
        const uint64_t brk = L.ck[sar][start];
        uint64_t s = start;
        uint64_t prev = brk;
        while ((prev >> 1) == (brk >> 1)) {
            if (s == 0) {
                s = ~0ull;
                break;
            }
            --s;
            prev = L.ck[sar][s];
        }
        ++s;
        const uint64_t brk64 = full_key(brk);
        for (int i = 0; i < L.k; ++i) {
            if (i == sar) {
                bp[i] = s;
                continue;
            }
            uint64_t cs = 0;
            const int f = find_mer(i, brk, cs);
            MH_CHECK(f >= 0, MH_ERR_UNSUPPORTED,
                     "repeat-limit restart against a list shorter than the seed (GetBreakpoint reads an "
                     "unset index in the reference)");
            if (f == 1) cs = brk64 == 0 ? 0 : cs + 1;
            bp[i] = cs;
        }
        return MH_OK;
    }

    uint64_t full_key(uint64_t c) const { return ((c >> 1) << (64 - 2 * L.sw)) | (c & 1ull); }
    uint64_t compact_key(uint64_t k) const { return ((k >> (64 - 2 * L.sw)) << 1) | (k & 1ull); }

    // Processing order of the blocks of run x for the lists S, whose block
    // starts in x are a (all live).
    std::vector<int> order_of(std::vector<int> S, std::vector<uint64_t> a) const
    {
        if (S.size() <= 1) return S;
        bool flip = false;
        // while every head was inserted after a block of one and the same
        // run p, the order is the reverse of the order in p
        for (;;) {
            bool chain = true;
            uint64_t p = 0;
            for (size_t t = 0; t < S.size() && chain; ++t) {
                if (a[t] == sp[S[t]]) chain = false;
                else {
                    const uint64_t pg = mk(S[t], a[t] - 1);
                    if (t == 0) p = pg;
                    else if (pg != p) chain = false;
                }
            }
            if (!chain) break;
            for (size_t t = 0; t < S.size(); ++t) a[t] = block_start(S[t], p, a[t] - 1);
            flip = !flip;
        }
        std::vector<int> initial;
        std::vector<std::pair<uint64_t, size_t>> ins;   // (previous mer, index into S)
        for (size_t t = 0; t < S.size(); ++t) {
            if (a[t] == sp[S[t]]) initial.push_back(S[t]);
            else ins.push_back({mk(S[t], a[t] - 1), t});
        }
        std::sort(initial.begin(), initial.end());
        std::stable_sort(ins.begin(), ins.end(),
                         [](const std::pair<uint64_t, size_t> &x, const std::pair<uint64_t, size_t> &y) {
                             return x.first > y.first;
                         });
        std::vector<int> out;
        for (size_t q = 0; q < ins.size();) {
            size_t r = q;
            while (r < ins.size() && ins[r].first == ins[q].first) ++r;
            if (r - q == 1) out.push_back(S[ins[q].second]);
            else {
                std::vector<int> G;
                std::vector<uint64_t> aG;
                for (size_t u = q; u < r; ++u) {
                    const size_t t = ins[u].second;
                    G.push_back(S[t]);
                    aG.push_back(block_start(S[t], ins[q].first, a[t] - 1));
                }
                std::vector<int> sub = order_of(G, aG);
                out.insert(out.end(), sub.rbegin(), sub.rend());
            }
            q = r;
        }
        out.insert(out.end(), initial.begin(), initial.end());
        if (flip) std::reverse(out.begin(), out.end());
        return out;
    }
};

}  // namespace

int plan_restarts(const PlanLists &L, const std::vector<uint64_t> &heavy, LiveRanges *out)
{
    const int k = L.k;
    MH_CHECK(k >= 1 && k <= PLAN_MAXK, MH_ERR_INVALID, "restart planning: list count out of range");
    for (int g = 0; g < k; ++g) {
        out->lo[g].clear();
        out->hi[g].clear();
        out->ep[g].clear();
    }
    out->restarts = 0;
    Planner P(L);
    std::vector<int> G;
    std::vector<uint64_t> a, e, pos(k);
    for (uint64_t m : heavy) {
        G.clear();
        a.clear();
        e.clear();
        uint64_t total = 0;
        for (int g = 0; g < k; ++g) {
            const uint64_t lo = P.lower_m(g, m, L.n[g]), hi = P.upper_m(g, m);
            const uint64_t aa = std::max(P.sp[g], lo);
            if (aa < hi) {
                G.push_back(g);
                a.push_back(aa);
                e.push_back(hi);
                total += hi - aa;
            }
        }
        if (total <= MER_REPEAT_LIMIT) continue;
        const std::vector<int> ord = P.order_of(G, a);
        // block of each list in processing order
        std::vector<size_t> at(ord.size());
        for (size_t j = 0; j < ord.size(); ++j)
            at[j] = (size_t)(std::find(G.begin(), G.end(), ord[j]) - G.begin());
        uint64_t cum = 0;
        int fire_j = -1;
        uint64_t fire_pos = 0;
        for (size_t j = 0; j < ord.size() && fire_j < 0; ++j) {
            const int g = ord[j];
            const uint64_t aj = a[at[j]], ej = e[at[j]];
            if (j > 0 && cum > MER_REPEAT_LIMIT) {
                fire_j = (int)j;
                fire_pos = aj;
                break;
            }
            // refills inside the block (or at its end) leave the head on the run's mer
            uint64_t c = P.sp[g] + MER_BUFFER_SIZE * ((aj - P.sp[g]) / MER_BUFFER_SIZE + 1);
            for (; c <= ej && c < L.n[g]; c += MER_BUFFER_SIZE)
                if (cum + (c - aj) > MER_REPEAT_LIMIT) {
                    fire_j = (int)j;
                    fire_pos = c;
                    break;
                }
            cum += ej - aj;
        }
        if (fire_j < 0) continue;
        // where every list had got to (old_starts + mer_baseindex + mer_index)
        for (int g = 0; g < k; ++g) pos[g] = std::max(P.sp[g], P.upper_m(g, m));
        for (size_t j = 0; j < ord.size(); ++j) {
            const int g = ord[j];
            if ((int)j < fire_j) pos[g] = e[at[j]];
            else if ((int)j == fire_j) pos[g] = fire_pos;
            else pos[g] = a[at[j]];
        }
        // the ranges this search read, up to the dropped run
        for (int g = 0; g < k; ++g) {
            const uint64_t E = P.lower_m(g, m, L.n[g]);
            if (P.sp[g] < E) {
                out->lo[g].push_back(P.sp[g]);
                out->hi[g].push_back(E);
                out->ep[g].push_back((uint32_t)out->restarts);
            }
        }
        // LM/MatchFinder.cpp:253-276
        const uint64_t unit = 1ull << (64 - 2 * L.sw);
        const uint64_t next64 = (m << (64 - 2 * L.sw)) + unit;
        const uint64_t nq = P.compact_key(next64);
        uint64_t next_pos = 0;
        int s = 0;
        for (; s < k; ++s) {
            if (P.find_mer(s, nq, next_pos) != 1) ++next_pos;
            if (next_pos < L.n[s]) break;
        }
        std::vector<uint64_t> nsp = P.sp;
        if (s < k) {
            const int rc = P.breakpoint(s, next_pos, nsp);
            if (rc != MH_OK) return rc;
        }
        for (int g = 0; g < k; ++g) {
            if (nsp[g] < pos[g]) nsp[g] = pos[g];
            if (g < s) nsp[g] = L.n[g];
        }
        P.sp = nsp;
        ++out->restarts;
    }
    for (int g = 0; g < k; ++g)
        if (P.sp[g] < L.n[g]) {
            out->lo[g].push_back(P.sp[g]);
            out->hi[g].push_back(L.n[g]);
            out->ep[g].push_back((uint32_t)out->restarts);
        }
    return MH_OK;
}

// masked mers of more than MER_REPEAT_LIMIT entries over all lists (host lists)
std::vector<uint64_t> heavy_mers_host(const PlanLists &L)
{
    std::vector<uint64_t> out;
    std::vector<uint64_t> at(L.k, 0);
    // k-way walk over the runs of every list
    for (;;) {
        uint64_t m = ~0ull;
        bool any = false;
        for (int g = 0; g < L.k; ++g)
            if (at[g] < L.n[g]) {
                const uint64_t x = L.ck[g][at[g]] >> 1;
                if (!any || x < m) m = x;
                any = true;
            }
        if (!any) break;
        uint64_t total = 0;
        for (int g = 0; g < L.k; ++g)
            while (at[g] < L.n[g] && (L.ck[g][at[g]] >> 1) == m) {
                ++at[g];
                ++total;
            }
        if (total > MER_REPEAT_LIMIT) out.push_back(m);
    }
    return out;
}

}  // namespace mh

// Host-only restart planning over sorted 64-bit keys (for checking).
extern "C" int mh_search_live_ranges(int k, const uint64_t *const *keys, const uint64_t *n, const uint64_t *len,
                                     uint64_t seed, uint64_t *ranges, uint64_t cap, uint64_t *n_ranges,
                                     int64_t *restarts)
{
    using namespace mh;
    MH_CHECK(k >= 1 && k <= PLAN_MAXK && keys && n && len && n_ranges, MH_ERR_INVALID, "bad arguments");
    const int sw = seed_weight((int64_t)seed), sl = seed_length((int64_t)seed);
    MH_CHECK(sw >= 1 && sw <= 31, MH_ERR_UNSUPPORTED, "seed weight outside 1..31");
    std::vector<std::vector<uint64_t>> ck(k);
    PlanLists L;
    L.k = k;
    L.sw = sw;
    L.sl = sl;
    for (int g = 0; g < k; ++g) {
        MH_CHECK(n[g] == 0 || keys[g], MH_ERR_INVALID, "null key list");
        ck[g].resize(n[g]);
        for (uint64_t i = 0; i < n[g]; ++i) ck[g][i] = ((keys[g][i] >> (64 - 2 * sw)) << 1) | (keys[g][i] & 1ull);
        L.ck[g] = ck[g].data();
        L.n[g] = n[g];
        L.len[g] = len[g];
    }
    LiveRanges lr;
    const int rc = plan_restarts(L, heavy_mers_host(L), &lr);
    if (rc != MH_OK) return rc;
    uint64_t cnt = 0;
    for (int g = 0; g < k; ++g)
        for (size_t t = 0; t < lr.lo[g].size(); ++t, ++cnt)
            if (ranges && cnt < cap) {
                ranges[3 * cnt] = (uint64_t)g;
                ranges[3 * cnt + 1] = lr.lo[g][t];
                ranges[3 * cnt + 2] = lr.hi[g][t];
            }
    *n_ranges = cnt;
    if (restarts) *restarts = lr.restarts;
    return cnt > cap && ranges ? MH_ERR_INVALID : MH_OK;
}
