// chain.cpp -- anchor chain, inter-anchor gap jobs and the alignment rows of
// a chain.
//
// In the reference the anchors come from the LCB pass (greedy breakpoint
// elimination over the match list, LM/ProgressiveAligner.cpp), which this
// library does not implement.  This file gives a simple deterministic stand-in
// for it so the path can run end to end: the heaviest chain of non-overlapping,
// collinear forward-strand matches (weighted by length); the DP jobs of the
// gaps between consecutive chain anchors (both sides non-empty and at most
// max_gap long); and the two rows of the chain as one interval, anchors
// aligned and each stretch between them unaligned, which the refinement passes
// (refine.cpp) take as input.  mh_recursion_gaps restates the reference's
// choice of gaps for the recursive anchor search on such rows
// (LM/ProgressiveAligner.cpp:838-905).
#include "mh_internal.hpp"
#include <algorithm>
#include <cstring>
#include <vector>

using namespace mh;

namespace {

struct Anc {
    int64_t s0, s1, len;   // 0-based starts
    uint32_t idx;
};

// Stable LSD radix sort of (key, value) pairs: as few passes of at most 13
// bits as the largest key needs, digits of equal width, every pass's histogram
// from one read of the keys.
void radix_sort_pairs(std::vector<uint64_t> &key, std::vector<uint32_t> &val)
{
    const size_t n = key.size();
    uint64_t mx = 0;
    for (uint64_t k : key) mx |= k;
    int bits = 0;
    while (bits < 64 && (mx >> bits)) ++bits;
    if (!bits || n < 2) return;
    const int passes = (bits + 12) / 13, width = (bits + passes - 1) / passes;
    const uint64_t radix = 1ull << width, mask = radix - 1;
    std::vector<uint32_t> cnt((size_t)passes * radix, 0u);
    for (size_t i = 0; i < n; ++i)
        for (int p = 0; p < passes; ++p) ++cnt[(size_t)p * radix + ((key[i] >> (p * width)) & mask)];
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> v2(n);
    for (int p = 0; p < passes; ++p) {
        uint32_t *c = cnt.data() + (size_t)p * radix;
        uint32_t acc = 0;
        bool one = false;
        for (uint64_t d = 0; d < radix; ++d) {
            if (c[d] == n) one = true;   // every key has this digit: nothing moves
            const uint32_t t = c[d];
            c[d] = acc;
            acc += t;
        }
        if (one) continue;
        const int sh = p * width;
        for (size_t i = 0; i < n; ++i) {
            const uint32_t d = (uint32_t)((key[i] >> sh) & mask);
            k2[c[d]] = key[i];
            v2[c[d]++] = val[i];
        }
        key.swap(k2);
        val.swap(v2);
    }
}

// High bit of each of the 8 bytes at p that holds a residue (not '-').
inline uint64_t residues8(const char *p)
{
    uint64_t v;
    std::memcpy(&v, p, 8);
    v ^= 0x2D2D2D2D2D2D2D2Dull;
    const uint64_t lo7 = 0x7F7F7F7F7F7F7F7Full;
    return (((v & lo7) + lo7) | v) & 0x8080808080808080ull;
}

// Bits of the bytes [0, b) of a group of 8 (b <= 8).
inline uint64_t bytes_below(int b) { return b ? ~0ull >> (64 - 8 * b) : 0ull; }

}  // namespace

extern "C" {

int mh_chain_gap_jobs(const mh_match *matches, uint64_t n, uint64_t len_a, uint64_t len_b, uint32_t max_gap,
                      uint32_t *chain, uint64_t *n_chain, mh_dp_job *jobs, uint64_t cap, uint64_t *n_jobs)
{
    MH_CHECK((matches || n == 0) && n_chain && n_jobs && (jobs || cap == 0), MH_ERR_INVALID, "null argument");
    MH_CHECK(len_a < (1ull << 31) && len_b < (1ull << 31) && n < (1ull << 31), MH_ERR_UNSUPPORTED,
             "chain producer limited to 2^31 positions and matches");
    // anchors in (s0, s1, match index) order (stable sort of packed starts)
    std::vector<uint64_t> key;
    std::vector<uint32_t> val;
    key.reserve(n);
    val.reserve(n);
    uint64_t top1 = len_b;   // the largest start in genome 2 fixes the key's low field
    for (uint64_t i = 0; i < n; ++i)
        if (matches[i].start[1] > 0 && (uint64_t)matches[i].start[1] > top1) top1 = (uint64_t)matches[i].start[1];
    int b1 = 1;
    while (b1 < 32 && (1ull << b1) <= top1) ++b1;
    for (uint64_t i = 0; i < n; ++i)
        if (matches[i].start[0] > 0 && matches[i].start[1] > 0) {
            key.push_back(((uint64_t)(matches[i].start[0] - 1) << b1) | (uint64_t)(matches[i].start[1] - 1));
            val.push_back((uint32_t)i);
        }
    radix_sort_pairs(key, val);
    const size_t m = key.size();
    std::vector<Anc> a(m);
    for (size_t q = 0; q < m; ++q) {
        const mh_match &mt = matches[val[q]];
        a[q] = {mt.start[0] - 1, mt.start[1] - 1, (int64_t)mt.length, val[q]};
    }
    // best[i] = len_i + max best[j] over anchors ending (in both genomes) before i starts.
    // Sweep by start; anchors become eligible once their end0 <= current start0;
    // a Fenwick tree over the distinct end1 values answers the prefix maximum.
    // Ranks come from one sort of every end1 (tag 0) and start1 (tag 1):
    // an end's rank is its distinct value's index, a start's query bound the
    // number of distinct ends <= it.
    std::vector<uint32_t> end_rank(m), qry_rank(m);
    size_t nkeys = 0;
    {
        std::vector<uint64_t> k2(2 * m);
        std::vector<uint32_t> v2(2 * m);
        for (size_t i = 0; i < m; ++i) {
            k2[2 * i] = ((uint64_t)(a[i].s1 + a[i].len) << 1);
            v2[2 * i] = (uint32_t)(2 * i);
            k2[2 * i + 1] = ((uint64_t)a[i].s1 << 1) | 1u;
            v2[2 * i + 1] = (uint32_t)(2 * i + 1);
        }
        radix_sort_pairs(k2, v2);
        uint64_t last = ~0ull;
        for (size_t x = 0; x < 2 * m; ++x) {
            const uint32_t i = v2[x] >> 1;
            if (!(v2[x] & 1u)) {
                if ((k2[x] >> 1) != last) {
                    last = k2[x] >> 1;
                    ++nkeys;
                }
                end_rank[i] = (uint32_t)(nkeys - 1);
            } else {
                qry_rank[i] = (uint32_t)nkeys;
            }
        }
    }
    struct Fen {
        int64_t score, anchor;
    };
    std::vector<Fen> fen(nkeys + 1, Fen{0, -1});
    auto upd = [&](size_t r, Fen v) {
        for (++r; r <= nkeys; r += r & (~r + 1))
            if (v.score > fen[r].score) fen[r] = v;
    };
    auto qry = [&](size_t r) {   // max over ranks [0, r)
        Fen b{0, -1};
        for (; r > 0; r -= r & (~r + 1))
            if (fen[r].score > b.score) b = fen[r];
        return b;
    };
    // anchors by end in genome 0 (ties by anchor rank)
    std::vector<uint64_t> by_end(m);
    std::vector<uint32_t> by_end_i(m);
    for (size_t i = 0; i < m; ++i) {
        by_end[i] = (uint64_t)(a[i].s0 + a[i].len);
        by_end_i[i] = (uint32_t)i;
    }
    radix_sort_pairs(by_end, by_end_i);
    std::vector<int64_t> best(m, 0), prev(m, -1);
    size_t e = 0;
    for (size_t i = 0; i < m; ++i) {
        while (e < m && (int64_t)by_end[e] <= a[i].s0) {
            const size_t j = by_end_i[e++];
            if (best[j] > 0) upd(end_rank[j], Fen{best[j], (int64_t)j});
        }
        const Fen b = qry(qry_rank[i]);
        best[i] = a[i].len + b.score;
        prev[i] = b.anchor;
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

int mh_chain_rows(const mh_match *matches, const uint32_t *chain, uint64_t n_chain, const char *seq_a,
                  uint64_t len_a, const char *seq_b, uint64_t len_b, char *row_a, char *row_b, uint64_t cap,
                  uint64_t *cols)
{
    MH_CHECK(cols && (matches || n_chain == 0) && (chain || n_chain == 0) && (seq_a || len_a == 0) &&
                 (seq_b || len_b == 0),
             MH_ERR_INVALID, "null argument");
    // check the chain: forward anchors, in order, not overlapping in either genome
    uint64_t pa = 0, pb = 0, anchored = 0;
    for (uint64_t q = 0; q < n_chain; ++q) {
        const mh_match &m = matches[chain[q]];
        MH_CHECK(m.start[0] > 0 && m.start[1] > 0, MH_ERR_INVALID, "chain anchor not on the forward strand");
        const uint64_t s0 = (uint64_t)m.start[0] - 1, s1 = (uint64_t)m.start[1] - 1;
        MH_CHECK(s0 >= pa && s1 >= pb && s0 + m.length <= len_a && s1 + m.length <= len_b, MH_ERR_INVALID,
                 "chain anchors out of order, overlapping or outside the genomes");
        pa = s0 + m.length;
        pb = s1 + m.length;
        anchored += m.length;
    }
    const uint64_t n = len_a + len_b - anchored;
    *cols = n;
    MH_CHECK(cap >= n, MH_ERR_INVALID, "row capacity too small (cols holds the size)");
    MH_CHECK(row_a && row_b, MH_ERR_INVALID, "null argument");
    // every anchor's stretch (the unaligned residues before it, genome 1's
    // against gaps, then gaps against genome 2's) and the anchor itself start
    // at a column known from the lengths before it, so blocks of anchors are
    // written in parallel; the last stretch runs to the genome ends
    std::vector<uint64_t> at(n_chain + 1);
    {
        uint64_t c = 0, qa = 0, qb = 0;
        for (uint64_t q = 0; q < n_chain; ++q) {
            const mh_match &m = matches[chain[q]];
            const uint64_t s0 = (uint64_t)m.start[0] - 1, s1 = (uint64_t)m.start[1] - 1;
            at[q] = c;
            c += (s0 - qa) + (s1 - qb) + m.length;
            qa = s0 + m.length;
            qb = s1 + m.length;
        }
        at[n_chain] = c;
    }
    auto piece = [&](uint64_t q) {   // stretch + anchor q, or the final stretch for q == n_chain
        uint64_t pa = 0, pb = 0;
        if (q > 0) {
            const mh_match &m = matches[chain[q - 1]];
            pa = (uint64_t)m.start[0] - 1 + m.length;
            pb = (uint64_t)m.start[1] - 1 + m.length;
        }
        uint64_t ea = len_a, eb = len_b, len = 0;
        if (q < n_chain) {
            const mh_match &m = matches[chain[q]];
            ea = (uint64_t)m.start[0] - 1;
            eb = (uint64_t)m.start[1] - 1;
            len = m.length;
        }
        uint64_t c = at[q];
        const uint64_t ga = ea - pa, gb = eb - pb;
        std::memcpy(row_a + c, seq_a + pa, ga);
        std::memset(row_b + c, '-', ga);
        c += ga;
        std::memset(row_a + c, '-', gb);
        std::memcpy(row_b + c, seq_b + pb, gb);
        c += gb;
        std::memcpy(row_a + c, seq_a + ea, len);
        std::memcpy(row_b + c, seq_b + eb, len);
    };
    const uint64_t per = 512, ntask = (n_chain + 1 + per - 1) / per;
    host_parallel(ntask, [&](size_t k) {
        for (uint64_t q = k * per; q < std::min<uint64_t>(n_chain + 1, (k + 1) * per); ++q) piece(q);
    });
    return MH_OK;
}

int mh_recursion_gaps(const char *row_a, const char *row_b, uint64_t cols, int64_t left_a, int64_t left_b,
                      uint64_t min_gap, mh_anchor_gap *gaps, uint64_t cap, uint64_t *n_gaps)
{
    MH_CHECK(n_gaps && (gaps || cap == 0) && ((row_a && row_b) || cols == 0), MH_ERR_INVALID, "null argument");
    // The scan below keeps, at each column, the residue counts so far (ci, cj),
    // the counts at the last column aligned in both rows (pi, pj) and whether
    // a column since then was not (in_gap).  Chunks of columns first summarise
    // themselves, the state at each chunk start follows from the summaries,
    // and then every chunk scans with its own start state.
    const uint64_t chunk = 1 << 18, nc = (cols + chunk - 1) / chunk;
    struct Sum {
        uint64_t ca = 0, cb = 0;          // residues in the chunk
        bool aligned = false;             // a column aligned in both rows
        uint64_t pa = 0, pb = 0;          // counts (within the chunk) at the last one
        bool gap_after = false;           // a column not aligned after it (or at all)
    };
    std::vector<Sum> sum(nc);
    const uint64_t HIGH = 0x8080808080808080ull;
    host_parallel(nc, [&](size_t k) {
        Sum &x = sum[k];
        const uint64_t hi = std::min(cols, (k + 1) * chunk);
        uint64_t c = k * chunk;
        // groups of 8 columns at once (residue and aligned-column bit masks)
        for (; c + 8 <= hi; c += 8) {
            const uint64_t ra = residues8(row_a + c), rb = residues8(row_b + c), al = ra & rb;
            if (al) {
                const int l = (63 - __builtin_clzll(al)) >> 3;
                x.aligned = true;
                x.pa = x.ca + (uint64_t)__builtin_popcountll(ra & bytes_below(l));
                x.pb = x.cb + (uint64_t)__builtin_popcountll(rb & bytes_below(l));
                x.gap_after = (~al & HIGH & ~bytes_below(l + 1)) != 0;
            } else {
                x.gap_after = true;
            }
            x.ca += (uint64_t)__builtin_popcountll(ra);
            x.cb += (uint64_t)__builtin_popcountll(rb);
        }
        for (; c < hi; ++c) {
            const bool ra = row_a[c] != '-', rb = row_b[c] != '-';
            if (ra && rb) {
                x.aligned = true;
                x.pa = x.ca;
                x.pb = x.cb;
                x.gap_after = false;
            } else {
                x.gap_after = true;
            }
            x.ca += ra;
            x.cb += rb;
        }
    });
    struct State {
        uint64_t ci = 0, cj = 0, pi = 0, pj = 0;
        bool in_gap = false;
    };
    std::vector<State> at(nc + 1);
    for (uint64_t k = 0; k < nc; ++k) {
        State t = at[k];
        if (sum[k].aligned) {
            t.pi = t.ci + sum[k].pa;
            t.pj = t.cj + sum[k].pb;
            t.in_gap = sum[k].gap_after;
        } else {
            t.in_gap = t.in_gap || sum[k].gap_after;
        }
        t.ci += sum[k].ca;
        t.cj += sum[k].cb;
        at[k + 1] = t;
    }
    // the interval's right ends (RightEnd = LeftEnd + residues - 1)
    const int64_t right_a = left_a + (int64_t)at[nc].ci - 1, right_b = left_b + (int64_t)at[nc].cj - 1;
    // a gap closes at an aligned column (or the end) after a run of other
    // columns longer than min_gap in both rows; zero-length anchors at the
    // previous aligned column and (unless a count equals the right end, as the
    // reference compares) at this one, which getInterveningCoordinates turns
    // into the gap
    auto close = [&](const State &t, std::vector<mh_anchor_gap> &out) {
        if (!(t.in_gap && t.ci - t.pi > min_gap && t.cj - t.pj > min_gap)) return;
        mh_anchor_gap g;
        g.lend[0] = left_a + (int64_t)t.pi;
        g.lend[1] = left_b + (int64_t)t.pj;
        const bool right = (int64_t)t.cj != right_b && (int64_t)t.ci != right_a;
        g.rend[0] = right ? left_a + (int64_t)t.ci : right_a + 1;
        g.rend[1] = right ? left_b + (int64_t)t.cj : right_b + 1;
        out.push_back(g);
    };
    std::vector<std::vector<mh_anchor_gap>> found(nc + 1);
    host_parallel(nc, [&](size_t k) {
        State t = at[k];
        const uint64_t hi = std::min(cols, (k + 1) * chunk);
        uint64_t c = k * chunk;
        // groups of 8 columns at once: only a group's first aligned column can
        // close a gap when a gap needs more than 7 residues (the others follow an
        // aligned column less than 8 columns before them)
        if (min_gap >= 8)
            for (; c + 8 <= hi; c += 8) {
                const uint64_t ra = residues8(row_a + c), rb = residues8(row_b + c), al = ra & rb;
                if (al) {
                    const int f = __builtin_ctzll(al) >> 3, l = (63 - __builtin_clzll(al)) >> 3;
                    State u = t;
                    u.ci += (uint64_t)__builtin_popcountll(ra & bytes_below(f));
                    u.cj += (uint64_t)__builtin_popcountll(rb & bytes_below(f));
                    if (f > 0) u.in_gap = true;
                    close(u, found[k]);
                    t.pi = t.ci + (uint64_t)__builtin_popcountll(ra & bytes_below(l));
                    t.pj = t.cj + (uint64_t)__builtin_popcountll(rb & bytes_below(l));
                    t.in_gap = (~al & HIGH & ~bytes_below(l + 1)) != 0;
                } else {
                    t.in_gap = true;
                }
                t.ci += (uint64_t)__builtin_popcountll(ra);
                t.cj += (uint64_t)__builtin_popcountll(rb);
            }
        for (; c < hi; ++c) {
            const bool ra = row_a[c] != '-', rb = row_b[c] != '-';
            if (ra && rb) {
                close(t, found[k]);
                t.pi = t.ci;
                t.pj = t.cj;
                t.in_gap = false;
            } else {
                t.in_gap = true;
            }
            t.ci += ra;
            t.cj += rb;
        }
    });
    close(at[nc], found[nc]);
    uint64_t ng = 0;
    for (const auto &v : found)
        for (const mh_anchor_gap &g : v) {
            if (ng < cap) gaps[ng] = g;
            ++ng;
        }
    *n_gaps = ng;
    return MH_OK;
}

}  // extern "C"
