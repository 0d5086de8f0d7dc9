// Compiled CPU restatement of the host stages of the pairwise whole path, for
// the CPU baseline: the profile and refinement passes over a pair of aligned
// rows, the gaps of the recursive anchor search, the XMFA text and the
// buildIndex LUT read back from it, and a batch of NWSmall jobs.
//
// TEST INFRASTRUCTURE ONLY: the bench's cpu_baseline legs and tests/ call it
// (through oracle/pyoracle.py); nothing in the product links it.  One thread
// per call; -O3 -ffp-contract=off like the rest of the oracle.
//
// The refinement follows oracle/refine_oracle.py function by function (that
// module carries the reference citations: removeLargeGapsPP and
// refineAlignment LM/ProgressiveAligner.cpp:1015-1240, AnchoredProfileProfile
// MU/anchoredpp.cpp:16-552, WindowSmooth/MergeBestCols MU/anchors.cpp:9-186,
// RefineW MU/refinew.cpp:75-200), float32 operation by operation; the DP is
// the oracle's or_nwsmall (MU/nwsmall.cpp).  The LUT follows
// oracle/index_oracle.py (PY/buildindex.py:110-129, PY/indexutils.pyx:27-108),
// the XMFA text the layout of WriteStandardAlignment (LM/IntervalList.h:
// 353-442) for one interval of two genomes.  tests/test_oracle_cpu_path.py
// checks every function here against those modules on the same inputs.
#include "mh_oracle.h"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace {

typedef float F;
const F GAP_OPEN = -400.0f, GAP_EXTEND = 0.0f, TERM_GAP = GAP_OPEN / 2.0f;
const int HOXD[4][4] = {{91, -114, -31, -123}, {-114, 100, -125, -31}, {-31, -125, 100, -114}, {-123, -31, -114, 91}};
const int SMOOTH_WINDOW = 21, ANCHOR_SPACING = 96;
const double MIN_BEST_COL_SCORE = 90.0, MIN_SMOOTH_SCORE = 90.0, SMOOTH_CEIL = 999.0;
const int64_t MAX_GAP_LENGTH = 5000, LCB_HANGOVER = 300;
const double MIN_WINDOW = 200.0, MAX_WINDOW = 20000.0, MIN_DENSITY = 0.5, MAX_DENSITY = 0.9;

struct Err {
    int code;
};

inline bool mu_gap(char c) { return c == '-' || c == '.'; }

inline int letter(char c)
{
    switch (std::toupper((unsigned char)c)) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': case 'U': return 3;
    default: return 99;
    }
}

inline uint8_t dp_class(char c)
{
    switch (std::toupper((unsigned char)c)) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': case 'U': return 3;
    case 'X': return 4;
    default: return 5;
    }
}

// LetterObjScoreXP's per-column score (refine_oracle.column_scores)
void column_scores(const char *r1, const char *r2, int64_t n, std::vector<F> &out)
{
    std::vector<F> mm(n, 0.0f), gg(n, 0.0f);
    int64_t start = 0, end = n - 1;
    for (int64_t c = 0; c < n; ++c)
        if (!(mu_gap(r1[c]) && mu_gap(r2[c]))) { start = c; break; }
    for (int64_t c = n - 1; c >= 0; --c)
        if (!(mu_gap(r1[c]) && mu_gap(r2[c]))) { end = c; break; }
    for (int64_t c = start; c <= end; ++c) {
        const int a = letter(r1[c]);
        if (a >= 4) continue;
        const int b = letter(r2[c]);
        if (b >= 4) continue;
        mm[c] = (F)HOXD[a][b] + 60.0f;
    }
    bool g1 = false, g2 = false;
    int64_t left = 0;
    F cur = 0.0f;
    for (int64_t c = start; c <= end; ++c) {
        const bool x = mu_gap(r1[c]), y = mu_gap(r2[c]);
        if (x && y) continue;
        if (x || y) {
            if ((x && !g1) || (!x && !g2)) {
                left = c;
                cur = cur + (c == start ? TERM_GAP : GAP_OPEN);
                if (x) g1 = true;
                else g2 = true;
            } else {
                cur = cur + GAP_EXTEND;
            }
            continue;
        }
        if (g1 || g2) {
            const F per = cur / (F)(c - left);
            for (int64_t k = left; k < c; ++k) gg[k] = per;
            left = 0xFFFFFFFFll;
            cur = 0.0f;
        }
        g1 = g2 = false;
    }
    if (g1 || g2) {
        cur = cur - GAP_OPEN;
        cur = cur + TERM_GAP;
        const F per = cur / (F)(n - left);
        for (int64_t k = left; k < n; ++k) gg[k] = per;
    }
    out.resize(n);
    for (int64_t c = 0; c < n; ++c) {
        const F s = mm[c] + gg[c];
        const F t = 1.0f * s;
        out[c] = 0.0f + t;
    }
}

// WindowSmooth (refine_oracle.window_smooth)
void window_smooth(const std::vector<F> &score, std::vector<F> &out)
{
    const int64_t n = (int64_t)score.size(), w = SMOOTH_WINDOW, h = w / 2;
    out.assign(n, 0.0f);
    if (n <= w) return;
    auto cl = [](F x) { return (double)x > SMOOTH_CEIL ? (F)SMOOTH_CEIL : x; };
    F tot = 0.0f;
    for (int64_t i = 0; i < w; ++i) tot = tot + cl(score[i]);
    for (int64_t i = h;; ++i) {
        out[i] = tot / (F)w;
        if (i == n - h - 1) break;
        tot = tot - cl(score[i - h]);
        tot = tot + cl(score[i + h + 1]);
    }
}

// FindBestColsComboPP + MergeBestCols (refine_oracle.anchor_cols)
void anchor_cols(const char *r1, const char *r2, int64_t n, std::vector<int64_t> &out)
{
    std::vector<F> s, sm;
    column_scores(r1, r2, n, s);
    window_smooth(s, sm);
    std::vector<int64_t> best;
    for (int64_t i = 0; i < n; ++i)
        if (!((double)s[i] < MIN_BEST_COL_SCORE) && !((double)sm[i] < MIN_SMOOTH_SCORE)) best.push_back(i);
    out.clear();
    size_t k = 0;
    while (k < best.size()) {
        const int64_t b = best[k];
        size_t within = 0;
        for (size_t i = k + 1; i < best.size(); ++i) {
            if (best[i] - b >= ANCHOR_SPACING) break;
            ++within;
        }
        int64_t anchor = b;
        if (within == 1) {
            const int64_t b2 = best[k + 1];
            anchor = s[b] > s[b2] ? b : b2;
        } else if (within > 1) {
            int64_t cd = ANCHOR_SPACING, closest = b;
            for (size_t i = k + 1; i < k + within; ++i) {
                const int64_t d = std::llabs(best[i] - b);
                if (d < cd) { closest = best[i]; cd = d; }
            }
            anchor = closest;
        }
        out.push_back(anchor);
        k += within + 1;
    }
}

std::string residues(const char *row, int64_t lo, int64_t hi)
{
    std::string s;
    s.reserve(hi - lo);
    for (int64_t c = lo; c < hi; ++c)
        if (!mu_gap(row[c])) s.push_back(row[c]);
    return s;
}

struct Dp {
    std::vector<uint8_t> A, B;
    std::vector<char> path;
    uint64_t cells = 0;
    // NWSmall of residue strings a and b, laid out as aligned rows
    void rows(const std::string &a, const std::string &b, std::string &x, std::string &y)
    {
        if (a.empty() || b.empty()) throw Err{-2};
        A.resize(a.size());
        B.resize(b.size());
        for (size_t i = 0; i < a.size(); ++i) A[i] = dp_class(a[i]);
        for (size_t i = 0; i < b.size(); ++i) B[i] = dp_class(b[i]);
        path.resize(a.size() + b.size() + 1);
        float sc;
        const int64_t np = or_nwsmall(A.data(), (uint32_t)A.size(), B.data(), (uint32_t)B.size(), path.data(), &sc);
        if (np < 0) throw Err{-2};
        cells += (uint64_t)a.size() * b.size();
        x.clear();
        y.clear();
        size_t i = 0, j = 0;
        for (int64_t k = 0; k < np; ++k) {
            const char c = path[k];
            if (c == 'M') { x.push_back(a[i++]); y.push_back(b[j++]); }
            else if (c == 'D') { x.push_back(a[i++]); y.push_back('-'); }
            else { x.push_back('-'); y.push_back(b[j++]); }
        }
    }
};

void anchored_profile_profile(Dp &dp, const char *r1, const char *r2, int64_t n, std::string &o1, std::string &o2)
{
    std::vector<int64_t> anchors;
    anchor_cols(r1, r2, n, anchors);
    o1.clear();
    o2.clear();
    std::string x, y;
    for (size_t k = 0; k <= anchors.size(); ++k) {
        const int64_t lo = k > 0 ? anchors[k - 1] : 0, hi = k < anchors.size() ? anchors[k] : n;
        if (hi == lo) continue;
        dp.rows(residues(r1, lo, hi), residues(r2, lo, hi), x, y);
        o1 += x;
        o2 += y;
    }
}

// RefineW over the rows present (one or two)
void refine_w(Dp &dp, const std::vector<std::string> &rows, int64_t window, std::vector<std::string> &out)
{
    const int64_t n = (int64_t)rows[0].size();
    out.assign(rows.size(), std::string());
    std::string x, y;
    for (int64_t f = 0; f < n; f += window) {
        const int64_t t = std::min(f + window - 1, n - 1);
        std::vector<std::string> seqs;
        for (const std::string &r : rows) seqs.push_back(residues(r.data(), f, t + 1));
        int full = 0, only = -1;
        for (size_t k = 0; k < seqs.size(); ++k)
            if (!seqs[k].empty()) { ++full; only = (int)k; }
        if (full > 1) {
            dp.rows(seqs[0], seqs[1], x, y);
            out[0] += x;
            out[1] += y;
        } else if (full == 1) {
            for (size_t j = 0; j < rows.size(); ++j)
                out[j] += (int)j == only ? seqs[only] : std::string(seqs[only].size(), '-');
        } else {
            throw Err{-3};
        }
    }
}

struct Piece {
    int64_t lo, hi;
    bool gap;
};

void remove_large_gaps(const char *r1, const char *r2, int64_t n, std::vector<Piece> &pieces)
{
    pieces.clear();
    int64_t gap_cols = 0, last = -1, base = 0;
    for (int64_t col = 0; col < n; ++col) {
        if (r1[col] != '-' && r2[col] != '-') {
            if (gap_cols > MAX_GAP_LENGTH) {
                if (last >= 0) {
                    const int64_t sp = last + LCB_HANGOVER - base;
                    pieces.push_back({base, base + sp, false});
                    base += sp;
                }
                const int64_t sp = col - LCB_HANGOVER - base;
                pieces.push_back({base, base + sp, true});
                base += sp;
            }
            last = col;
            gap_cols = 0;
        } else {
            ++gap_cols;
        }
    }
    if (gap_cols > MAX_GAP_LENGTH) {
        if (last >= 0) {
            const int64_t sp = last + LCB_HANGOVER - base;
            pieces.push_back({base, base + sp, false});
            base += sp;
        }
        pieces.push_back({base, n, true});
    } else {
        pieces.push_back({base, n, false});
    }
}

// residue counts per row over [lo, hi), from prefix sums
struct Counts {
    std::vector<int64_t> c1, c2;
    Counts(const char *r1, const char *r2, int64_t n) : c1(n + 1, 0), c2(n + 1, 0)
    {
        for (int64_t i = 0; i < n; ++i) {
            c1[i + 1] = c1[i] + (r1[i] != '-');
            c2[i + 1] = c2[i] + (r2[i] != '-');
        }
    }
    int64_t l1(int64_t lo, int64_t hi) const { return c1[hi] - c1[lo]; }
    int64_t l2(int64_t lo, int64_t hi) const { return c2[hi] - c2[lo]; }
};

int dense_enough(const Counts &cn, int64_t lo, int64_t hi)
{
    const int64_t l1 = cn.l1(lo, hi), l2 = cn.l2(lo, hi);
    const int mult = (l1 > 0) + (l2 > 0);
    const int64_t aln = hi - lo;
    const double density = (double)(l1 + l2) / ((double)aln * (double)mult);
    const double threshold =
        ((MAX_DENSITY - MIN_DENSITY) / (MIN_WINDOW - MAX_WINDOW)) * ((double)aln - MAX_WINDOW) + MIN_DENSITY;
    if (density > MAX_DENSITY) return 2;
    if (density > threshold) return 1;
    return 0;
}

void pieces_of(const char *r1, const char *r2, int64_t n, const Counts &cn, std::vector<Piece> &pieces)
{
    remove_large_gaps(r1, r2, n, pieces);
    size_t i = 0;
    while (i < pieces.size()) {
        const Piece p = pieces[i];
        const int d = dense_enough(cn, p.lo, p.hi);
        const int64_t len = p.hi - p.lo;
        if ((d == 0 && (double)len > MAX_WINDOW / 3) || (d == 1 && (double)len > MAX_WINDOW) ||
            (d == 2 && (double)len > MAX_WINDOW * 3)) {
            const int64_t sp = len / 2;
            pieces[i] = {p.lo, p.hi - sp, p.gap};
            pieces.insert(pieces.begin() + i + 1, Piece{p.hi - sp, p.hi, p.gap});
            continue;
        }
        ++i;
    }
}

void refine_alignment(Dp &dp, const std::string &r1, const std::string &r2, bool profile, std::string &o1,
                      std::string &o2)
{
    const int64_t n = (int64_t)r1.size();
    const Counts cn(r1.data(), r2.data(), n);
    std::vector<Piece> pieces;
    pieces_of(r1.data(), r2.data(), n, cn, pieces);
    std::vector<std::vector<std::string>> outs;
    outs.reserve(pieces.size());
    for (const Piece &p : pieces) {
        std::string a = r1.substr(p.lo, p.hi - p.lo), b = r2.substr(p.lo, p.hi - p.lo);
        const int64_t l1 = cn.l1(p.lo, p.hi), l2 = cn.l2(p.lo, p.hi);
        if (p.gap) {
            outs.push_back({a, b});
        } else if (profile) {
            if (l1 > 0 && l2 > 0) {
                std::string x, y;
                anchored_profile_profile(dp, a.data(), b.data(), (int64_t)a.size(), x, y);
                outs.push_back({x, y});
            } else {
                outs.push_back({a, b});
            }
        } else {
            const int d = dense_enough(cn, p.lo, p.hi);
            std::vector<std::string> present;
            if (l1 > 0) present.push_back(a);
            if (l2 > 0) present.push_back(b);
            if (d == 0) {
                outs.push_back(present);
            } else {
                std::vector<std::string> w;
                refine_w(dp, present, d == 1 ? 500 : 200, w);
                outs.push_back(w);
            }
        }
    }
    int64_t new_len = 0;
    for (const auto &o : outs) new_len += (int64_t)o[0].size();
    o1.assign(new_len, '-');
    o2.assign(new_len, '-');
    int64_t pos[2] = {0, 0};
    std::string *rows[2] = {&o1, &o2};
    for (const auto &o : outs)
        for (size_t k = 0; k < o.size(); ++k) {
            std::memcpy(&(*rows[k])[pos[k]], o[k].data(), o[k].size());
            pos[k] += (int64_t)o[k].size();
        }
}

// ---- XMFA text and the LUT read back from it ----

void put_row(std::FILE *fp, const char *row, int64_t n)
{
    for (int64_t k = 0; k < n; k += 80) {
        std::fwrite(row + k, 1, (size_t)std::min<int64_t>(80, n - k), fp);
        std::fputc('\n', fp);
    }
}

struct Entry {
    int64_t num = 0, start = 0;
    std::string seq;
};

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

// the groups of index_oracle.parse_xmfa, entries reduced to number, start
// and sequence
void parse_xmfa(const std::string &t, std::vector<std::vector<Entry>> &groups)
{
    groups.clear();
    const size_t n = t.size();
    size_t i = 0;
    while (i < n) {
        // a group opens with '>' at a line start and runs to the first '='
        if (t[i] == '>' && (i == 0 || t[i - 1] == '\n')) {
            const size_t eq = t.find('=', i);
            if (eq == std::string::npos) break;
            std::vector<Entry> entries;
            size_t p = i;
            while (p < eq) {
                if (!(t[p] == '>' && (p == 0 || t[p - 1] == '\n'))) { ++p; continue; }
                size_t q = p + 1;
                while (q < eq && is_ws(t[q]) && t[q] != '\n') ++q;
                Entry e;
                const size_t d0 = q;
                while (q < eq && std::isdigit((unsigned char)t[q])) e.num = e.num * 10 + (t[q++] - '0');
                if (q == d0 || q >= eq || t[q] != ':') { p = q; continue; }
                ++q;
                const size_t s0 = q;
                while (q < eq && std::isdigit((unsigned char)t[q])) e.start = e.start * 10 + (t[q++] - '0');
                if (q == s0) { p = q; continue; }
                const size_t nl = t.find('\n', q);
                if (nl == std::string::npos || nl > eq) { p = eq; continue; }
                size_t r = nl + 1;
                while (r < n && t[r] != '=' && t[r] != '>') {
                    if (!is_ws(t[r])) e.seq.push_back(t[r]);
                    ++r;
                }
                entries.push_back(std::move(e));
                p = r;
            }
            groups.push_back(std::move(entries));
            const size_t nl = t.find('\n', eq);
            i = nl == std::string::npos ? n : nl + 1;
        } else {
            const size_t nl = t.find('\n', i);
            i = nl == std::string::npos ? n : nl + 1;
        }
    }
}

void fill_from(int32_t *lut, int64_t n, int64_t lower, int64_t idx, int64_t width)
{
    const int64_t lm = lut[lower];
    for (int64_t up = idx + 1; up < n; ++up) {
        if (up - idx > width) return;
        if (lut[up] - lm == up - lower) {
            for (int64_t k = lower; k <= up; ++k) lut[k] = (int32_t)(lm + (k - lower));
            return;
        }
    }
}

}  // namespace

extern "C" {

// The profile pass, then the final refinement pass, of two aligned rows of
// n columns (refine_oracle.gapped_alignment).  o1, o2 hold cap bytes each.
// Returns the new column count; -1 if it exceeds cap, -2 for a DP operand
// without residues ("NWSmall: length=0"), -3 for a window without residues.
// *cells_out (nullable) gets the DP cells computed.
int64_t or_refine_gapped(const char *r1, const char *r2, int64_t n, char *o1, char *o2, int64_t cap,
                         uint64_t *cells_out)
{
    try {
        Dp dp;
        std::string a, b, c, d;
        refine_alignment(dp, std::string(r1, n), std::string(r2, n), true, a, b);
        refine_alignment(dp, a, b, false, c, d);
        if (cells_out) *cells_out = dp.cells;
        if ((int64_t)c.size() > cap) return -1;
        std::memcpy(o1, c.data(), c.size());
        std::memcpy(o2, d.data(), d.size());
        return (int64_t)c.size();
    } catch (const Err &e) {
        return e.code;
    }
}

// Column scores, anchor columns and one AnchoredProfileProfile of two rows,
// for the tests.  or_anchor_cols returns the anchor count (out holds cap).
int64_t or_column_scores(const char *r1, const char *r2, int64_t n, float *out)
{
    std::vector<F> s;
    column_scores(r1, r2, n, s);
    std::memcpy(out, s.data(), n * sizeof(F));
    return n;
}

int64_t or_anchor_cols(const char *r1, const char *r2, int64_t n, int64_t *out, int64_t cap)
{
    std::vector<int64_t> a;
    anchor_cols(r1, r2, n, a);
    for (size_t k = 0; k < a.size() && (int64_t)k < cap; ++k) out[k] = a[k];
    return (int64_t)a.size();
}

int64_t or_refine_pass(const char *r1, const char *r2, int64_t n, int profile, char *o1, char *o2, int64_t cap)
{
    try {
        Dp dp;
        std::string a, b;
        refine_alignment(dp, std::string(r1, n), std::string(r2, n), profile != 0, a, b);
        if ((int64_t)a.size() > cap) return -1;
        std::memcpy(o1, a.data(), a.size());
        std::memcpy(o2, b.data(), b.size());
        return (int64_t)a.size();
    } catch (const Err &e) {
        return e.code;
    }
}

// Writes the XMFA of one interval (two rows of n columns, genome lengths
// len1/len2, both forward from position 1) to path, reads the file back and
// builds the LUT over genome 1 (genome_length entries): column walk,
// fix_zero_idx (seq1/seq2: the genomes), fill_gaps(300), smooth_edges(20).
// Returns the XMFA size in bytes, -1 on an I/O error, -2 when fix_zero_idx
// finds no mapped entry (the reference's IndexError).
int64_t or_xmfa_lut(const char *path, const char *name1, const char *name2, const char *r1, const char *r2,
                    int64_t n, int64_t len1, int64_t len2, const char *seq1, const char *seq2, int32_t *lut,
                    int64_t genome_length)
{
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) return -1;
    std::fprintf(fp, "#FormatVersion Mauve1\n");
    std::fprintf(fp, "#Sequence1File\t%s\n#Sequence1Format\tFastA\n", name1);
    std::fprintf(fp, "#Sequence2File\t%s\n#Sequence2Format\tFastA\n", name2);
    std::fprintf(fp, "> 1:1-%lld + %s\n", (long long)len1, name1);
    put_row(fp, r1, n);
    std::fprintf(fp, "> 2:1-%lld + %s\n", (long long)len2, name2);
    put_row(fp, r2, n);
    std::fprintf(fp, "=\n");
    if (std::fclose(fp) != 0) return -1;

    fp = std::fopen(path, "rb");
    if (!fp) return -1;
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    std::string text((size_t)sz, '\0');
    const size_t got = std::fread(&text[0], 1, (size_t)sz, fp);
    std::fclose(fp);
    if ((long)got != sz) return -1;

    std::vector<std::vector<Entry>> groups;
    parse_xmfa(text, groups);
    const int64_t L = genome_length;
    for (int64_t i = 0; i < L; ++i) lut[i] = -1;
    for (const auto &entries : groups) {
        const Entry *g = nullptr, *r = nullptr;
        for (const Entry &e : entries) {
            if (e.num == 1 && !g) g = &e;
            if (e.num == 2 && !r) r = &e;
        }
        if (!g || !r) continue;
        int64_t gi = g->start - 1, ri = r->start - 1;
        const std::string &gs = g->seq, &rs = r->seq;
        if (rs.size() < gs.size()) return -1;
        for (size_t col = 0; col < gs.size(); ++col) {
            const char base = gs[col], rb = rs[col];
            if (base != '-') ++gi;
            if (rb != '-') ++ri;
            if (base == rb && gi < L) lut[gi >= 0 ? gi : L + gi] = (int32_t)ri;
        }
    }
    // fix_zero_idx
    if (L > 0 && lut[0] == -1) {
        int64_t nxt = 1;
        while (nxt < L && (lut[nxt] == -1 || nxt < 10)) ++nxt;
        if (nxt >= L) return -2;
        const int64_t r = (int64_t)lut[nxt] - nxt;
        if (r > -1 && seq1[0] == seq2[r]) lut[0] = (int32_t)r;
    }
    for (int64_t idx = 1; idx < L; ++idx)
        if (lut[idx] == -1 && lut[idx - 1] != -1) fill_from(lut, L, idx - 1, idx, 300);
    for (int64_t idx = 1; idx < L; ++idx)
        if (lut[idx] != lut[idx - 1] + 1 && lut[idx - 1] != -1) fill_from(lut, L, idx - 1, idx, 20);
    return (int64_t)sz;
}

// The gaps recurseOnPairs searches (pyoracle.recursion_gaps; LM/
// ProgressiveAligner.cpp:838-905) of two aligned rows forward in both genomes
// from left_a/left_b: out receives (lend_a, lend_b, rend_a, rend_b) per gap,
// up to cap gaps.  Returns the gap count.
int64_t or_recursion_gaps(const char *ra, const char *rb, int64_t n, int64_t left_a, int64_t left_b,
                          int64_t min_gap, int64_t *out, int64_t cap)
{
    int64_t na = 0, nb = 0;
    for (int64_t c = 0; c < n; ++c) {
        na += ra[c] != '-';
        nb += rb[c] != '-';
    }
    const int64_t right_a = left_a + na - 1, right_b = left_b + nb - 1;
    int64_t ci = 0, cj = 0, pi = 0, pj = 0, count = 0;
    bool in_gap = false;
    for (int64_t col = 0; col <= n; ++col) {
        if (col == n || (ra[col] != '-' && rb[col] != '-')) {
            if (in_gap && ci - pi > min_gap && cj - pj > min_gap) {
                if (count < cap) {
                    int64_t *o = out + 4 * count;
                    o[0] = left_a + pi;
                    o[1] = left_b + pj;
                    if (cj != right_b && ci != right_a) {
                        o[2] = left_a + ci;
                        o[3] = left_b + cj;
                    } else {
                        o[2] = right_a + 1;
                        o[3] = right_b + 1;
                    }
                }
                ++count;
            }
            pi = ci;
            pj = cj;
            in_gap = false;
        } else {
            in_gap = true;
        }
        if (col < n) {
            ci += ra[col] != '-';
            cj += rb[col] != '-';
        }
    }
    return count;
}

// NWSmall over jobs [lo, hi) of a batch (letter classes; job k is
// A[a_off[k]..+a_len[k]] against B[b_off[k]..+b_len[k]]); scores to
// score_out (nullable).  Returns the cells computed, -1 on an empty operand.
int64_t or_nwsmall_batch(const uint8_t *A, const uint8_t *B, const uint64_t *a_off, const uint32_t *a_len,
                         const uint64_t *b_off, const uint32_t *b_len, int64_t lo, int64_t hi, float *score_out)
{
    uint32_t mx = 0;
    for (int64_t k = lo; k < hi; ++k) mx = std::max(mx, a_len[k] + b_len[k]);
    std::vector<char> path(mx + 1);
    int64_t cells = 0;
    for (int64_t k = lo; k < hi; ++k) {
        float sc;
        if (or_nwsmall(A + a_off[k], a_len[k], B + b_off[k], b_len[k], path.data(), &sc) < 0) return -1;
        if (score_out) score_out[k] = sc;
        cells += (int64_t)a_len[k] * b_len[k];
    }
    return cells;
}

}  // extern "C"
