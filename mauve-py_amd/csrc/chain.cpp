// chain.cpp -- anchor chain and inter-anchor gap jobs.
//
// In the reference the DP jobs come from the LCB pass and MUSCLE's anchored
// profile-profile ranges (LM/ProgressiveAligner.cpp:1119-1290,
// MU/anchoredpp.cpp:443-552).  That host logic is not part of this library;
// this file gives a simple deterministic producer of the same kind of work so
// the seed-anchoring and DP stages can run end to end: the heaviest chain of
// non-overlapping, collinear forward-strand matches (weighted by length), and
// one DP job per gap between consecutive chain anchors where both sides are
// non-empty and at most max_gap long.
#include "mh_internal.hpp"
#include <algorithm>
#include <cstring>

using namespace mh;

namespace {

struct Anc {
    int64_t s0, s1, len;   // 0-based starts
    uint32_t idx;
};

}  // namespace

extern "C" {

int mh_chain_gap_jobs(const mh_match *matches, uint64_t n, uint64_t len_a, uint64_t len_b, uint32_t max_gap,
                      uint32_t *chain, uint64_t *n_chain, mh_dp_job *jobs, uint64_t cap, uint64_t *n_jobs)
{
    MH_CHECK((matches || n == 0) && n_chain && n_jobs && (jobs || cap == 0), MH_ERR_INVALID, "null argument");
    // anchors in (s0, s1, match index) order; the sort runs on packed keys
    std::vector<std::pair<uint64_t, uint32_t>> ord;
    ord.reserve(n);
    for (uint64_t i = 0; i < n; ++i)
        if (matches[i].start[0] > 0 && matches[i].start[1] > 0)
            ord.push_back({((uint64_t)(matches[i].start[0] - 1) << 32) | (uint64_t)(matches[i].start[1] - 1),
                           (uint32_t)i});
    std::sort(ord.begin(), ord.end());
    std::vector<Anc> a(ord.size());
    for (size_t q = 0; q < ord.size(); ++q) {
        const mh_match &mt = matches[ord[q].second];
        a[q] = {mt.start[0] - 1, mt.start[1] - 1, (int64_t)mt.length, ord[q].second};
    }
    const size_t m = a.size();
    // best[i] = len_i + max best[j] over anchors ending (in both genomes) before i starts.
    // Sweep by start; anchors become eligible once their end0 <= current start0;
    // a Fenwick tree over end1 ranks answers the prefix maximum.
    std::vector<int64_t> ends1(m);
    for (size_t i = 0; i < m; ++i) ends1[i] = a[i].s1 + a[i].len;
    std::vector<int64_t> keys = ends1;
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<std::pair<int64_t, int64_t>> fen(keys.size() + 1, {0, -1});   // (score, anchor)
    auto upd = [&](size_t r, std::pair<int64_t, int64_t> v) {
        for (++r; r <= keys.size(); r += r & (~r + 1))
            if (v.first > fen[r].first) fen[r] = v;
    };
    auto qry = [&](size_t r) {   // max over ranks [0, r)
        std::pair<int64_t, int64_t> b{0, -1};
        for (; r > 0; r -= r & (~r + 1))
            if (fen[r].first > b.first) b = fen[r];
        return b;
    };
    // anchors by end in genome 0 (ties by anchor rank)
    std::vector<uint64_t> by_end(m);
    for (size_t i = 0; i < m; ++i) by_end[i] = ((uint64_t)(a[i].s0 + a[i].len) << 32) | (uint64_t)i;
    std::sort(by_end.begin(), by_end.end());
    std::vector<int64_t> best(m, 0), prev(m, -1);
    size_t e = 0;
    for (size_t i = 0; i < m; ++i) {
        while (e < m && (int64_t)(by_end[e] >> 32) <= a[i].s0) {
            const size_t j = (size_t)(by_end[e++] & 0xFFFFFFFFu);
            if (best[j] > 0) upd(std::lower_bound(keys.begin(), keys.end(), ends1[j]) - keys.begin(), {best[j], (int64_t)j});
        }
        const size_t r = std::upper_bound(keys.begin(), keys.end(), a[i].s1) - keys.begin();
        const auto b = qry(r);
        best[i] = a[i].len + b.first;
        prev[i] = b.second;
    }
    int64_t top = -1, topv = 0;
    for (size_t i = 0; i < m; ++i)
        if (best[i] > topv) { topv = best[i]; top = (int64_t)i; }
    std::vector<size_t> ch;
    for (int64_t i = top; i >= 0; i = prev[i]) ch.push_back((size_t)i);
    std::reverse(ch.begin(), ch.end());
    *n_chain = ch.size();
    if (chain)
        for (size_t q = 0; q < ch.size(); ++q) chain[q] = a[ch[q]].idx;
    // gaps: before the first anchor, between anchors, after the last one
    uint64_t nj = 0;
    int64_t pa = 0, pb = 0;
    auto gap = [&](int64_t ea, int64_t eb) {
        const int64_t ga = ea - pa, gb = eb - pb;
        if (ga > 0 && gb > 0 && ga <= (int64_t)max_gap && gb <= (int64_t)max_gap) {
            if (nj < cap) {
                jobs[nj].a_off = (uint64_t)pa;
                jobs[nj].b_off = (uint64_t)pb;
                jobs[nj].a_len = (uint32_t)ga;
                jobs[nj].b_len = (uint32_t)gb;
            }
            ++nj;
        }
    };
    for (size_t q = 0; q < ch.size(); ++q) {
        const Anc &x = a[ch[q]];
        gap(x.s0, x.s1);
        pa = x.s0 + x.len;
        pb = x.s1 + x.len;
    }
    gap((int64_t)len_a, (int64_t)len_b);
    *n_jobs = nj;
    return MH_OK;
}

}  // extern "C"
