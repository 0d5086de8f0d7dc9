// Refinement of pairwise gapped alignments: the host job producers around the
// DP and the splicing of its paths, with every DP job of every interval of a
// pass run in one batch on the GPU.
//
// For two genomes progressiveMauve refines each aligned interval twice
// (LM/ProgressiveAligner.cpp:2561-2565 and 3751-3755):
//   profile pass     refineAlignment(profile_aln=true): pieces from
//                    removeLargeGapsPP and the density halving; each piece with
//                    residues of both genomes goes through ProfileAlignFast in
//                    anchored mode (LM/MuscleInterface.cpp:1053-1162) =
//                    AnchoredProfileProfile (MU/anchoredpp.cpp:443-552): anchor
//                    columns from the smoothed column score, the column ranges
//                    between them re-aligned from their residues;
//   refinement pass  refineAlignment(profile_aln=false): each piece through
//                    RefineFast with the window its density picks (0: RefineVert,
//                    which leaves two sequences alone; 500 or 200: RefineW,
//                    MU/refinew.cpp:75-200, each window re-aligned from its
//                    residues).
// Every re-alignment is one NWSmall of genome 1's residues (A) against genome
// 2's (B) with single-sequence profiles, the job mh_dp_align runs.  Pieces are
// put back together as refineAlignment does (LM/ProgressiveAligner.cpp:
// 1222-1237), including its handling of a piece re-aligned from one row: that
// row goes to genome 1 and genome 2's position does not advance.
//
// Column scores are float, operation by operation, as MUSCLE's SCORE.
#include "mh_internal.hpp"

#include <cmath>
#include <cstring>
#include <atomic>
#include <thread>

#pragma clang fp contract(off)

using namespace mh;

namespace {

constexpr float GAP_OPEN = -400.0f;
constexpr float GAP_EXTEND = 0.0f;
constexpr float TERM_GAP = GAP_OPEN / 2.0f;          // TERMGAPS_Half, MU/objscore2.cpp:22-41
constexpr unsigned SMOOTH_WINDOW = 21;               // MU/anchoredpp.cpp:367-368
constexpr unsigned ANCHOR_SPACING = 96;
constexpr double SMOOTH_CEIL = 999.0;                // SetDefaultsSPN_DNA, MU/params.cpp:296-314
constexpr double MIN_BEST_COL_SCORE = 90.0;
constexpr double MIN_SMOOTH_SCORE = 90.0;
constexpr uint64_t MAX_GAP_LENGTH = 5000;            // LM/ProgressiveAligner.cpp:58-63
constexpr uint64_t LCB_HANGOVER = 300;
constexpr double MIN_WINDOW = 200, MAX_WINDOW = 20000, MIN_DENSITY = 0.5, MAX_DENSITY = 0.9;
constexpr unsigned INSANE = 0xFFFFFFFFu;

const int HOXD[4][4] = {{91, -114, -31, -123}, {-114, 100, -125, -31}, {-31, -125, 100, -114}, {-123, -31, -114, 91}};

inline bool mu_gap(char c) { return c == '-' || c == '.'; }   // MUSCLE's IsGapChar
inline unsigned letter(char c)
{
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': case 'U': case 'u': return 3;   // libMUSCLE scores U as T
    default: return 99;
    }
}

// Column classes: 0-3 the letters A C G T (either case), GAPC MUSCLE's gap
// characters '-' and '.', anything else OTHER.
constexpr uint8_t GAPC = 4, OTHER = 5;
struct ColumnCodes {
    uint8_t v[256];
    float mm[6][6];   // the letter score of a column (0 unless both are letters)
    ColumnCodes()
    {
        for (int c = 0; c < 256; ++c) {
            const unsigned l = letter((char)c);
            v[c] = l < 4 ? (uint8_t)l : mu_gap((char)c) ? GAPC : OTHER;
        }
        for (int x = 0; x < 6; ++x)
            for (int y = 0; y < 6; ++y) mm[x][y] = x < 4 && y < 4 ? (float)HOXD[x][y] + 60.0f : 0.0f;
    }
};
const ColumnCodes CODES;

// LetterObjScoreXP's per-column score for two single-sequence profiles
// (weights 1): the NUC_SP letter score, plus each gap's penalty spread
// evenly over the gap's columns (MU/anchoredpp.cpp:16-316).  A column is
// one or the other: letter-score columns carry no gap share, and the columns
// a gap's share covers (from its first column to the next column of two
// residues, all-gap columns included) have no letter score.
void column_scores(const char *r1, const char *r2, unsigned n, std::vector<float> &score)
{
    score.assign(n, 0.0f);
    float *out = score.data();
    const uint8_t *code = CODES.v;
    unsigned start = 0, end = n - 1;
    for (unsigned c = 0; c < n; ++c)
        if (!(code[(uint8_t)r1[c]] == GAPC && code[(uint8_t)r2[c]] == GAPC)) {
            start = c;
            break;
        }
    for (int c = (int)n - 1; c >= 0; --c)
        if (!(code[(uint8_t)r1[c]] == GAPC && code[(uint8_t)r2[c]] == GAPC)) {
            end = (unsigned)c;
            break;
        }
    bool g1 = false, g2 = false;
    unsigned left = 0;
    float cur = 0.0f;
    for (unsigned c = start; c <= end; ++c) {
        const uint8_t a = code[(uint8_t)r1[c]], b = code[(uint8_t)r2[c]];
        const bool x = a == GAPC, y = b == GAPC;
        if (x && y) continue;
        if (x || y) {
            bool &gapping = x ? g1 : g2;
            if (!gapping) {
                left = c;
                cur += c == start ? TERM_GAP : GAP_OPEN;
                gapping = true;
            } else {
                cur += GAP_EXTEND;
            }
            continue;
        }
        if (g1 || g2) {
            const float per = cur / (float)(c - left);
            for (unsigned k = left; k < c; ++k) out[k] = 1.0f * (0.0f + per);
            left = INSANE;
            cur = 0.0f;
        }
        g1 = g2 = false;
        out[c] = 1.0f * (CODES.mm[a][b] + 0.0f);
    }
    if (g1 || g2) {
        cur -= GAP_OPEN;
        cur += TERM_GAP;
        const float per = cur / (float)(n - left);
        for (unsigned k = left; k < n; ++k) out[k] = 1.0f * (0.0f + per);
    }
}

// WindowSmooth, MU/anchors.cpp:9-48.
void window_smooth(const std::vector<float> &s, std::vector<float> &out)
{
    const unsigned n = (unsigned)s.size(), w = SMOOTH_WINDOW, h = w / 2;
    out.assign(n, 0.0f);
    if (n <= w) return;
    auto cl = [](float x) { return (float)(x > SMOOTH_CEIL ? SMOOTH_CEIL : (double)x); };
    float tot = 0.0f;
    for (unsigned i = 0; i < w; ++i) tot += cl(s[i]);
    for (unsigned i = h;; ++i) {
        out[i] = tot / (float)w;
        if (i == n - h - 1) break;
        tot -= cl(s[i - h]);
        tot += cl(s[i + h + 1]);
    }
}

}  // namespace

namespace mh {

// FindAnchorColsPP (MU/anchoredpp.cpp:353-404) for two rows of n columns:
// FindBestColsComboPP, then MergeBestCols (MU/anchors.cpp:137-186).
void refine_anchor_cols(const char *r1, const char *r2, unsigned n, std::vector<unsigned> &anchors)
{
    anchors.clear();
    if (n == 0) return;
    thread_local std::vector<float> score, smooth;
    thread_local std::vector<unsigned> best;
    best.clear();
    column_scores(r1, r2, n, score);
    window_smooth(score, smooth);
    for (unsigned i = 0; i < n; ++i) {
        if ((double)score[i] < MIN_BEST_COL_SCORE) continue;
        if ((double)smooth[i] < MIN_SMOOTH_SCORE) continue;
        best.push_back(i);
    }
    for (size_t k = 0; k < best.size();) {
        const unsigned b = best[k];
        unsigned within = 0;
        for (size_t i = k + 1; i < best.size(); ++i) {
            if (best[i] - b >= ANCHOR_SPACING) break;
            ++within;
        }
        unsigned anchor = b;
        if (within == 1) {
            const unsigned b2 = best[k + 1];
            anchor = score[b] > score[b2] ? b : b2;
        } else if (within > 1) {
            int closest_d = (int)ANCHOR_SPACING;
            unsigned closest = b;
            for (size_t i = k + 1; i < k + within; ++i) {
                int d = (int)(best[i] - b);
                if (d < 0) d = -d;
                if (d < closest_d) {
                    closest = best[i];
                    closest_d = d;
                }
            }
            anchor = closest;
        }
        anchors.push_back(anchor);
        k += within + 1;
    }
}

struct RefinePiece {
    uint64_t lo, hi;
    bool gap;
};

}  // namespace mh

namespace {

template <class F> void parallel_for(size_t n, F f) { host_parallel(n, std::function<void(size_t)>(f)); }

// Gaps ('-', the only gap in Mauve's GappedAlignment) among the 8 bytes at p.
inline unsigned gaps8(const char *p)
{
    uint64_t v;
    std::memcpy(&v, p, 8);
    v ^= 0x2D2D2D2D2D2D2D2Dull;                       // '-' bytes become zero
    const uint64_t lo7 = 0x7F7F7F7F7F7F7F7Full;
    const uint64_t t = ~(((v & lo7) + lo7) | v | lo7);  // high bit of each zero byte
    return (unsigned)__builtin_popcountll(t);
}

// High bit of each of the 8 columns at c that holds residues in both rows.
inline uint64_t aligned8(const char *r1, const char *r2)
{
    uint64_t v1, v2;
    std::memcpy(&v1, r1, 8);
    std::memcpy(&v2, r2, 8);
    v1 ^= 0x2D2D2D2D2D2D2D2Dull;
    v2 ^= 0x2D2D2D2D2D2D2D2Dull;
    const uint64_t lo7 = 0x7F7F7F7F7F7F7F7Full;
    const uint64_t g1 = ~(((v1 & lo7) + lo7) | v1 | lo7), g2 = ~(((v2 & lo7) + lo7) | v2 | lo7);
    return ~(g1 | g2) & 0x8080808080808080ull;
}

// Residue counts of the two rows before any column, from counts kept every 64
// columns and a count inside the block.  Built in parallel chunks.
struct Prefix {
    const char *r[2] = {nullptr, nullptr};
    uint32_t *blk[2] = {nullptr, nullptr};    // n / 64 + 1 entries each, caller's memory
    static uint64_t entries(uint64_t n) { return n / 64 + 2; }
    void build(const char *r1, const char *r2, uint64_t n, uint32_t *m1, uint32_t *m2)
    {
        r[0] = r1;
        r[1] = r2;
        blk[0] = m1;
        blk[1] = m2;
        const uint64_t nb = n / 64, chunk = 1 << 12, nc = (nb + chunk - 1) / chunk;   // chunks of blocks
        std::vector<uint32_t> t1(nc + 1, 0), t2(nc + 1, 0);
        parallel_for(nc, [&](size_t k) {
            uint32_t a = 0, b = 0;
            for (uint64_t q = k * chunk; q < std::min(nb, (k + 1) * chunk); ++q) {
                m1[q] = a;
                m2[q] = b;
                unsigned g1 = 0, g2 = 0;
                for (int x = 0; x < 64; x += 8) {
                    g1 += gaps8(r1 + q * 64 + x);
                    g2 += gaps8(r2 + q * 64 + x);
                }
                a += 64 - g1;
                b += 64 - g2;
            }
            t1[k + 1] = a;
            t2[k + 1] = b;
        });
        for (uint64_t k = 0; k < nc; ++k) {
            t1[k + 1] += t1[k];
            t2[k + 1] += t2[k];
        }
        parallel_for(nc, [&](size_t k) {
            for (uint64_t q = k * chunk; q < std::min(nb, (k + 1) * chunk); ++q) {
                m1[q] += t1[k];
                m2[q] += t2[k];
            }
        });
        m1[nb] = t1[nc];
        m2[nb] = t2[nc];
    }
    uint64_t count(int g, uint64_t x) const
    {
        const uint64_t q = x / 64;
        const char *p = r[g] + q * 64;
        unsigned m = (unsigned)(x - q * 64), res = 0;
        for (; m >= 8; m -= 8, p += 8) res += 8 - gaps8(p);
        for (; m > 0; --m, ++p) res += *p != '-';
        return blk[g][q] + res;
    }
    uint64_t l1(uint64_t lo, uint64_t hi) const { return count(0, hi) - count(0, lo); }
    uint64_t l2(uint64_t lo, uint64_t hi) const { return count(1, hi) - count(1, lo); }
};

// IsDenseEnough (LM/ProgressiveAligner.cpp:1015-1038) from a piece's residue
// counts and alignment length.
int density_of(uint64_t l1, uint64_t l2, uint64_t aln)
{
    double total = 0;
    if (l1) total += (double)l1;
    if (l2) total += (double)l2;
    const unsigned mult = (l1 > 0) + (l2 > 0);
    const double density = total / ((double)aln * (double)mult);
    const double threshold = ((MAX_DENSITY - MIN_DENSITY) / (MIN_WINDOW - MAX_WINDOW)) * ((double)aln - MAX_WINDOW) +
                             MIN_DENSITY;
    if (density > MAX_DENSITY) return 2;
    if (density > threshold) return 1;
    return 0;
}

// The pieces of one interval: removeLargeGapsPP (LM/ProgressiveAligner.cpp:
// 1051-1117), then the halving of pieces too long for their density
// (1142-1166), with the densities from the prefix counts.
void pieces_fast(const char *r1, const char *r2, uint64_t n, const Prefix &pf, std::vector<RefinePiece> &pieces)
{
    pieces.clear();
    // The cuts depend only on the runs of more than MAX_GAP_LENGTH columns not
    // aligned in both rows and the aligned columns around them: find those runs
    // in parallel chunks (runs inside a chunk, plus each chunk's first and last
    // aligned column for the runs across chunks), then replay the scan's
    // decisions over the runs in order.
    struct Run {
        int64_t prev, next;    // aligned columns around the run (-1: none before)
    };
    const uint64_t chunk = 1 << 18, nc = (n + chunk - 1) / chunk;
    std::vector<std::vector<Run>> inner(nc);
    std::vector<int64_t> fst(nc, -1), lst(nc, -1);
    parallel_for(nc, [&](size_t k) {
        const uint64_t lo = k * chunk, hi = std::min(n, lo + chunk);
        int64_t prev = -1;
        // an aligned column from first to last of a group of 8 (runs inside a
        // group are shorter than MAX_GAP_LENGTH)
        auto seen = [&](int64_t first, int64_t last) {
            if (prev < 0) fst[k] = first;
            else if (first - prev - 1 > (int64_t)MAX_GAP_LENGTH) inner[k].push_back(Run{prev, first});
            prev = last;
        };
        uint64_t c = lo;
        for (; c + 8 <= hi; c += 8) {
            const uint64_t al = aligned8(r1 + c, r2 + c);
            if (al) seen((int64_t)c + (__builtin_ctzll(al) >> 3), (int64_t)c + ((63 - __builtin_clzll(al)) >> 3));
        }
        for (; c < hi; ++c)
            if (r1[c] != '-' && r2[c] != '-') seen((int64_t)c, (int64_t)c);
        lst[k] = prev;
    });
    std::vector<Run> runs;
    int64_t last = -1;
    for (uint64_t k = 0; k < nc; ++k) {
        if (fst[k] < 0) continue;
        if (fst[k] - last - 1 > (int64_t)MAX_GAP_LENGTH) runs.push_back(Run{last, fst[k]});
        runs.insert(runs.end(), inner[k].begin(), inner[k].end());
        last = lst[k];
    }
    uint64_t base = 0;
    for (const Run &r : runs) {
        if (r.prev >= 0) {
            const uint64_t sp = (uint64_t)r.prev + LCB_HANGOVER - base;
            pieces.push_back({base, base + sp, false});
            base += sp;
        }
        const uint64_t sp = (uint64_t)r.next - LCB_HANGOVER - base;
        pieces.push_back({base, base + sp, true});
        base += sp;
    }
    const uint64_t tail = last < 0 ? n : n - 1 - (uint64_t)last;
    if (tail > MAX_GAP_LENGTH) {
        if (last >= 0) {
            const uint64_t sp = (uint64_t)last + LCB_HANGOVER - base;
            pieces.push_back({base, base + sp, false});
            base += sp;
        }
        pieces.push_back({base, n, true});
    } else {
        pieces.push_back({base, n, false});
    }
    // the halving loop, depth first as the list insertion makes it
    std::vector<RefinePiece> out;
    std::vector<RefinePiece> stack;
    for (const RefinePiece &p0 : pieces) {
        stack.push_back(p0);
        while (!stack.empty()) {
            const RefinePiece p = stack.back();
            stack.pop_back();
            const uint64_t len = p.hi - p.lo;
            const int d = density_of(pf.l1(p.lo, p.hi), pf.l2(p.lo, p.hi), len);
            if ((d == 0 && (double)len > MAX_WINDOW / 3) || (d == 1 && (double)len > MAX_WINDOW) ||
                (d == 2 && (double)len > MAX_WINDOW * 3)) {
                const uint64_t sp = len / 2;
                stack.push_back(RefinePiece{p.hi - sp, p.hi, p.gap});
                stack.push_back(RefinePiece{p.lo, p.hi - sp, p.gap});
                continue;
            }
            out.push_back(p);
        }
    }
    pieces.swap(out);
}

// The output of one piece in one pass, as a list of segments per output row.
enum SegKind : uint8_t {
    S_COPY,   // the columns [lo, hi) of input row src
    S_RES,    // the residues of input row src in [lo, hi)
    S_GAPS,   // lo gap characters
    S_JOB     // row 0: the A row of DP job lo; row 1: its B row
};
struct Seg {
    SegKind kind;
    uint8_t src;
    uint64_t lo, hi;
};
struct Job {
    uint64_t lo, hi;           // columns: A = residues of row 0, B = residues of row 1
    uint32_t la, lb;
};
struct Task {
    size_t iv;
    RefinePiece p;
    bool whole;                // the interval is copied as it is
    int rows = 2;              // 2, or 1 for a piece re-aligned from one row
    std::vector<Seg> row[2];
    std::vector<Job> jobs;
    uint64_t out_len[2] = {0, 0};
    int err = MH_OK;
    std::string msg;
};

// Plan one piece for one pass.
void plan_piece(const char *r1, const char *r2, const Prefix &pf, bool profile, Task &t)
{
    const RefinePiece &p = t.p;
    const uint64_t len = p.hi - p.lo;
    auto copy_both = [&]() {
        t.row[0].push_back(Seg{S_COPY, 0, p.lo, p.hi});
        t.row[1].push_back(Seg{S_COPY, 1, p.lo, p.hi});
    };
    auto add_job = [&](uint64_t lo, uint64_t hi, uint64_t la, uint64_t lb) {
        t.row[0].push_back(Seg{S_JOB, 0, t.jobs.size(), 0});
        t.row[1].push_back(Seg{S_JOB, 1, t.jobs.size(), 0});
        t.jobs.push_back(Job{lo, hi, (uint32_t)la, (uint32_t)lb});
    };
    if (t.whole) {
        copy_both();
        return;
    }
    const uint64_t l1 = pf.l1(p.lo, p.hi), l2 = pf.l2(p.lo, p.hi);
    if (p.gap || (profile && !(l1 > 0 && l2 > 0))) {
        copy_both();
        return;
    }
    if (len >= (1ull << 31)) {
        t.err = MH_ERR_UNSUPPORTED;
        t.msg = "piece longer than 2^31 columns";
        return;
    }
    if (profile) {
        // AnchoredProfileProfile: the ranges between anchor columns
        std::vector<unsigned> anchors;
        refine_anchor_cols(r1 + p.lo, r2 + p.lo, (unsigned)len, anchors);
        for (size_t k = 0; k <= anchors.size(); ++k) {
            const uint64_t lo = p.lo + (k > 0 ? anchors[k - 1] : 0);
            const uint64_t hi = p.lo + (k < anchors.size() ? anchors[k] : len);
            if (hi == lo) continue;
            const uint64_t la = pf.l1(lo, hi), lb = pf.l2(lo, hi);
            if (la == 0 || lb == 0) {
                t.err = MH_ERR_ZERO_LENGTH;
                t.msg = "NWSmall: length=0 (an anchored range has residues of one genome only)";
                return;
            }
            add_job(lo, hi, la, lb);
        }
        return;
    }
    const int d = density_of(l1, l2, len);
    int present[2], np = 0;
    if (l1) present[np++] = 0;
    if (l2) present[np++] = 1;
    t.rows = np;
    if (d == 0) {
        for (int q = 0; q < np; ++q) t.row[q].push_back(Seg{S_COPY, (uint8_t)present[q], p.lo, p.hi});
        return;
    }
    const uint64_t w = d == 1 ? 500 : 200;
    for (uint64_t f = p.lo; f < p.hi; f += w) {
        const uint64_t e = std::min(f + w, p.hi);
        const uint64_t cnt[2] = {pf.l1(f, e), pf.l2(f, e)};
        if (np == 2 && cnt[0] && cnt[1]) {
            add_job(f, e, cnt[0], cnt[1]);
            continue;
        }
        int full = -1;
        for (int q = 0; q < np; ++q)
            if (cnt[present[q]]) full = q;
        if (full < 0) {
            t.err = MH_ERR_INVALID;
            t.msg = "a refinement window holds no residues";
            return;
        }
        for (int q = 0; q < np; ++q) {
            if (q == full) t.row[q].push_back(Seg{S_RES, (uint8_t)present[q], f, e});
            else t.row[q].push_back(Seg{S_GAPS, 0, cnt[present[full]], 0});
        }
    }
}

// DP operand classes of the residues (A,C,G,T with U as T -> 0..3, X -> 4,
// any other letter -> 5; libMUSCLE's wildcard counts, see dp_float.hip).
struct OperandCodes {
    uint8_t v[256];
    OperandCodes()
    {
        for (int c = 0; c < 256; ++c) {
            const unsigned l = letter((char)c);
            v[c] = l < 4 ? (uint8_t)l : (c == 'X' || c == 'x') ? 4 : 5;
        }
    }
};
const OperandCodes OPCODES;

// Operand classes of the residues of row r in [lo, hi) into out; true when
// any is not A, C, G, T (the job takes the float DP).
bool put_codes(const char *r, uint64_t lo, uint64_t hi, uint8_t *out)
{
    unsigned wild = 0;
    for (uint64_t c = lo; c < hi; ++c) {
        if (r[c] == '-') continue;
        const uint8_t l = OPCODES.v[(uint8_t)r[c]];
        wild |= l > 3;
        *out++ = l;
    }
    return wild != 0;
}

// One pass over all intervals: plan (in parallel over pieces), one batched DP
// through pinned staging, then every piece written straight into the new rows
// at its place (the reassembly of LM/ProgressiveAligner.cpp:1222-1237).
int refine_pass(mh_context *ctx, const std::vector<const char *> &in1, const std::vector<const char *> &in2,
                const std::vector<uint64_t> &lens, bool profile, int slot, std::vector<char *> &o1,
                std::vector<char *> &o2, std::vector<uint64_t> &olen, uint64_t *cells_out)
{
    const size_t n = lens.size();
    double tt = now_ms();
    auto lap = [&](const char *name) {
        const double t2 = now_ms();
        host_time(ctx, name, t2 - tt);
        tt = t2;
    };
    std::vector<Prefix> pf(n);
    std::vector<std::vector<RefinePiece>> pcs(n);
    std::vector<char> whole(n, 0);
    {
        uint64_t tot = 0;
        for (size_t i = 0; i < n; ++i) tot += Prefix::entries(lens[i]);
        uint32_t *m1 = ctx->refine_prefix[0].get<uint32_t>(tot), *m2 = ctx->refine_prefix[1].get<uint32_t>(tot);
        MH_CHECK(m1 && m2, MH_ERR_NOMEM, "host allocation failed");
        for (size_t i = 0; i < n; ++i) {
            pf[i].build(in1[i], in2[i], lens[i], m1, m2);
            m1 += Prefix::entries(lens[i]);
            m2 += Prefix::entries(lens[i]);
        }
    }
    lap("host_refine_prefix");
    parallel_for(n, [&](size_t i) {
        const uint64_t len = lens[i];
        // an interval that holds one genome only is not refined (the
        // Multiplicity() > 1 test of doGappedAlignment, LM/ProgressiveAligner.cpp:1268)
        if (pf[i].l1(0, len) == 0 || pf[i].l2(0, len) == 0) {
            whole[i] = 1;
            pcs[i] = {RefinePiece{0, len, false}};
            return;
        }
        pieces_fast(in1[i], in2[i], len, pf[i], pcs[i]);
    });
    std::vector<Task> tasks;
    std::vector<size_t> first(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        first[i] = tasks.size();
        for (const RefinePiece &p : pcs[i]) {
            Task t;
            t.iv = i;
            t.p = p;
            t.whole = whole[i] != 0;
            tasks.push_back(std::move(t));
        }
    }
    first[n] = tasks.size();
    lap("host_refine_pieces");
    parallel_for(tasks.size(), [&](size_t k) {
        Task &t = tasks[k];
        plan_piece(in1[t.iv], in2[t.iv], pf[t.iv], profile, t);
    });
    for (const Task &t : tasks)
        if (t.err != MH_OK) {
            set_error(t.msg + " (interval " + std::to_string(t.iv) + ")");
            return t.err;
        }
    // every job of the pass in one batch, in task order
    const size_t nt = tasks.size();
    std::vector<uint64_t> jbase(nt + 1, 0), abase(nt + 1, 0), bbase(nt + 1, 0);
    uint64_t cells = 0;
    for (size_t k = 0; k < nt; ++k) {
        uint64_t sa = 0, sb = 0;
        for (const Job &j : tasks[k].jobs) {
            sa += j.la;
            sb += j.lb;
            cells += (uint64_t)j.la * j.lb;
        }
        jbase[k + 1] = jbase[k] + tasks[k].jobs.size();
        abase[k + 1] = abase[k] + sa;
        bbase[k + 1] = bbase[k] + sb;
    }
    if (cells_out) *cells_out += cells;
    const uint64_t njobs = jbase[nt], abytes = abase[nt], bbytes = bbase[nt], pbytes = abytes + bbytes;
    std::vector<mh_dp_job> jobs(njobs);
    // pinned staging: [A codes | B codes | paths | results]
    const uint64_t res_off = (abytes + bbytes + pbytes + 15) & ~15ull;
    uint8_t *stage = nullptr;
    if (njobs) {
        int rc = pinned(ctx, res_off + njobs * sizeof(mh_dp_result) + 64, (void **)&stage);
        if (rc != MH_OK) return rc;
    }
    std::vector<uint8_t> wild(njobs, 0);
    parallel_for(nt, [&](size_t k) {
        const Task &t = tasks[k];
        uint64_t pa = abase[k], pb = bbase[k];
        for (size_t j = 0; j < t.jobs.size(); ++j) {
            const Job &jb = t.jobs[j];
            mh_dp_job &dj = jobs[jbase[k] + j];
            dj.a_off = pa;
            dj.b_off = pb;
            dj.a_len = jb.la;
            dj.b_len = jb.lb;
            const bool w1 = put_codes(in1[t.iv], jb.lo, jb.hi, stage + pa);
            const bool w2 = put_codes(in2[t.iv], jb.lo, jb.hi, stage + abytes + pb);
            wild[jbase[k] + j] = (uint8_t)(w1 || w2);
            pa += jb.la;
            pb += jb.lb;
        }
    });
    lap("host_refine_plan");
    const uint8_t *paths = stage + abytes + bbytes;
    const mh_dp_result *res = reinterpret_cast<const mh_dp_result *>(stage + res_off);
    if (njobs) {
        int rc;
        if ((rc = ctx->dp_a.ensure(abytes + 64)) != MH_OK) return rc;
        if ((rc = ctx->dp_b.ensure(bbytes + 64)) != MH_OK) return rc;
        if ((rc = ctx->dp_paths.ensure(pbytes + 64)) != MH_OK) return rc;
        if ((rc = ctx->dp_res.ensure(njobs * sizeof(mh_dp_result) + 64)) != MH_OK) return rc;
        MH_HIP(hipMemcpyAsync(ctx->dp_a.ptr, stage, abytes, hipMemcpyHostToDevice, ctx->stream));
        MH_HIP(hipMemcpyAsync(ctx->dp_b.ptr, stage + abytes, bbytes, hipMemcpyHostToDevice, ctx->stream));
        rc = dp_run(ctx, ctx->dp_a.as<uint8_t>(), ctx->dp_b.as<uint8_t>(), jobs.data(), njobs,
                    ctx->dp_paths.as<uint8_t>(), ctx->dp_res.as<mh_dp_result>(), wild.data());
        if (rc != MH_OK) return rc;
        // the paths come back at two bits an edge
        if ((rc = ctx->dp_packed.ensure(pbytes / 4 + 64)) != MH_OK) return rc;
        if ((rc = dp_pack_paths(ctx, ctx->dp_paths.as<uint8_t>(), pbytes, ctx->dp_packed.as<uint8_t>())) != MH_OK)
            return rc;
        MH_HIP(hipMemcpyAsync(stage + abytes + bbytes, ctx->dp_packed.ptr, (pbytes + 3) / 4, hipMemcpyDeviceToHost,
                              ctx->stream));
        MH_HIP(hipMemcpyAsync(stage + res_off, ctx->dp_res.ptr, njobs * sizeof(mh_dp_result), hipMemcpyDeviceToHost,
                              ctx->stream));
        MH_HIP(hipStreamSynchronize(ctx->stream));
        for (uint64_t j = 0; j < njobs; ++j)
            MH_CHECK(res[j].path_len > 0, MH_ERR_DEVICE, "trace-back failed for job " + std::to_string(j));
    }
    lap("host_refine_dp");
    // output lengths, then each piece's place in its interval's rows
    parallel_for(nt, [&](size_t k) {
        Task &t = tasks[k];
        for (int q = 0; q < t.rows; ++q) {
            uint64_t m = 0;
            for (const Seg &sg : t.row[q]) {
                switch (sg.kind) {
                case S_COPY: m += sg.hi - sg.lo; break;
                case S_RES: m += sg.src == 0 ? pf[t.iv].l1(sg.lo, sg.hi) : pf[t.iv].l2(sg.lo, sg.hi); break;
                case S_GAPS: m += sg.lo; break;
                case S_JOB: m += res[jbase[k] + sg.lo].path_len; break;
                }
            }
            t.out_len[q] = m;
        }
    });
    std::vector<uint64_t> pos0(nt), pos1(nt), tail(n);
    o1.assign(n, nullptr);
    o2.assign(n, nullptr);
    olen.assign(n, 0);
    uint64_t total_out = 0;
    for (size_t i = 0; i < n; ++i) {
        uint64_t p0 = 0, p1 = 0;
        for (size_t k = first[i]; k < first[i + 1]; ++k) {
            pos0[k] = p0;
            pos1[k] = p1;
            p0 += tasks[k].out_len[0];
            if (tasks[k].rows == 2) p1 += tasks[k].out_len[1];
        }
        olen[i] = p0;
        tail[i] = p1;
        total_out += p0;
    }
    {
        char *b1 = ctx->refine_rows[slot][0].get<char>(total_out), *b2 = ctx->refine_rows[slot][1].get<char>(total_out);
        MH_CHECK(b1 && b2, MH_ERR_NOMEM, "host allocation failed");
        for (size_t i = 0; i < n; ++i) {
            o1[i] = b1;
            o2[i] = b2;
            b1 += olen[i];
            b2 += olen[i];
            // genome 2's row is written up to tail; the rest stays a gap
            // (a piece re-aligned from one row does not advance it)
            std::memset(o2[i] + tail[i], '-', olen[i] - tail[i]);
        }
    }
    parallel_for(nt, [&](size_t k) {
        const Task &t = tasks[k];
        const char *in[2] = {in1[t.iv], in2[t.iv]};
        // a one-row piece writes its row into genome 1's row (see the file comment)
        char *dst[2] = {o1[t.iv] + pos0[k], o2[t.iv] + pos1[k]};
        for (int q = 0; q < t.rows; ++q) {
            char *w = dst[q];
            for (const Seg &sg : t.row[q]) {
                switch (sg.kind) {
                case S_COPY:
                    std::memcpy(w, in[sg.src] + sg.lo, sg.hi - sg.lo);
                    w += sg.hi - sg.lo;
                    break;
                case S_RES:
                    for (uint64_t c = sg.lo; c < sg.hi; ++c)
                        if (in[sg.src][c] != '-') *w++ = in[sg.src][c];
                    break;
                case S_GAPS:
                    std::memset(w, '-', sg.lo);
                    w += sg.lo;
                    break;
                case S_JOB: {
                    const Job &jb = t.jobs[sg.lo];
                    const mh_dp_result &r = res[jbase[k] + sg.lo];
                    const char *src = in[q];
                    uint64_t c = jb.lo;
                    // row 0 takes a residue on M and D, row 1 on M and I
                    // (two-bit codes M 0, D 1, I 2)
                    const uint32_t skip = q == 0 ? 2u : 1u;
                    for (uint64_t at = r.path_start, e = r.path_start + r.path_len; at < e; ++at) {
                        if (((paths[at >> 2] >> (2 * (at & 3))) & 3u) == skip) {
                            *w++ = '-';
                            continue;
                        }
                        while (src[c] == '-') ++c;
                        *w++ = src[c++];
                    }
                    break;
                }
                }
            }
        }
    });
    lap("host_refine_splice");
    return MH_OK;
}

}  // namespace

extern "C" {

int mh_refine_pairwise(mh_context *ctx, uint64_t n, const char *const *rows1, const char *const *rows2,
                       const uint64_t *cols, uint32_t passes, char *const *out1, char *const *out2,
                       const uint64_t *caps, uint64_t *out_cols, mh_refine_stats *stats)
{
    MH_CHECK(ctx && (n == 0 || (rows1 && rows2 && cols && out1 && out2 && caps && out_cols)), MH_ERR_INVALID,
             "null argument");
    MH_CHECK((passes & ~3u) == 0, MH_ERR_INVALID, "unknown refinement pass");
    MH_HIP(hipSetDevice(ctx->device));
    if (stats) std::memset(stats, 0, sizeof(*stats));
    std::vector<const char *> in1(n), in2(n);
    std::vector<uint64_t> lens(n);
    // 1 MiB pieces of every row, for the '.' check and the final copy
    struct Span {
        uint64_t iv, lo, hi;
    };
    std::vector<Span> spans;
    auto pieces_of = [&](const std::vector<uint64_t> &len) {
        spans.clear();
        for (uint64_t i = 0; i < n; ++i)
            for (uint64_t lo = 0; lo < len[i]; lo += 1 << 20) spans.push_back({i, lo, std::min(len[i], lo + (1 << 20))});
    };
    for (uint64_t i = 0; i < n; ++i) {
        MH_CHECK(rows1[i] && rows2[i], MH_ERR_INVALID, "null row");
        in1[i] = rows1[i];
        in2[i] = rows2[i];
        lens[i] = cols[i];
    }
    {
        pieces_of(lens);
        std::atomic<int> dot(0);
        parallel_for(spans.size(), [&](size_t k) {
            const Span &p = spans[k];
            if (std::memchr(rows1[p.iv] + p.lo, '.', p.hi - p.lo) || std::memchr(rows2[p.iv] + p.lo, '.', p.hi - p.lo))
                dot = 1;
        });
        MH_CHECK(!dot, MH_ERR_INVALID, "'.' as a gap character is not supported (use '-')");
    }
    std::vector<char *> o1, o2;
    std::vector<uint64_t> ol;
    int rc;
    if (passes & MH_REFINE_PROFILE) {
        uint64_t cells = 0;
        if ((rc = refine_pass(ctx, in1, in2, lens, true, 0, o1, o2, ol, &cells)) != MH_OK) return rc;
        if (stats) stats->profile_cells = cells;
        for (uint64_t i = 0; i < n; ++i) {
            in1[i] = o1[i];
            in2[i] = o2[i];
            lens[i] = ol[i];
        }
    }
    if (passes & MH_REFINE_WINDOWS) {
        uint64_t cells = 0;
        if ((rc = refine_pass(ctx, in1, in2, lens, false, 1, o1, o2, ol, &cells)) != MH_OK) return rc;
        if (stats) stats->window_cells = cells;
        for (uint64_t i = 0; i < n; ++i) {
            in1[i] = o1[i];
            in2[i] = o2[i];
            lens[i] = ol[i];
        }
    }
    for (uint64_t i = 0; i < n; ++i) out_cols[i] = lens[i];
    for (uint64_t i = 0; i < n; ++i)
        MH_CHECK(lens[i] <= caps[i], MH_ERR_INVALID,
                 "output capacity too small for interval " + std::to_string(i) + " (out_cols holds the sizes)");
    pieces_of(lens);
    parallel_for(spans.size(), [&](size_t k) {
        const Span &p = spans[k];
        std::memcpy(out1[p.iv] + p.lo, in1[p.iv] + p.lo, p.hi - p.lo);
        std::memcpy(out2[p.iv] + p.lo, in2[p.iv] + p.lo, p.hi - p.lo);
    });
    return MH_OK;
}

int mh_refine_anchor_cols(const char *row1, const char *row2, uint64_t cols, uint32_t *anchors, uint64_t cap,
                          uint64_t *n_anchors)
{
    MH_CHECK(row1 && row2 && n_anchors && (anchors || cap == 0), MH_ERR_INVALID, "null argument");
    MH_CHECK(cols < (1ull << 32), MH_ERR_UNSUPPORTED, "more than 2^32 columns");
    std::vector<unsigned> a;
    refine_anchor_cols(row1, row2, (unsigned)cols, a);
    *n_anchors = a.size();
    for (size_t k = 0; k < a.size() && k < cap; ++k) anchors[k] = a[k];
    return MH_OK;
}

int mh_refine_pieces(const char *row1, const char *row2, uint64_t cols, uint64_t *pieces, uint64_t cap,
                     uint64_t *n_pieces)
{
    MH_CHECK(row1 && row2 && n_pieces && (pieces || cap == 0), MH_ERR_INVALID, "null argument");
    std::vector<RefinePiece> p;
    std::vector<uint32_t> m1(Prefix::entries(cols)), m2(Prefix::entries(cols));
    Prefix pf;
    pf.build(row1, row2, cols, m1.data(), m2.data());
    pieces_fast(row1, row2, cols, pf, p);
    *n_pieces = p.size();
    for (size_t k = 0; k < p.size() && k < cap; ++k) {
        pieces[3 * k] = p[k].lo;
        pieces[3 * k + 1] = p[k].hi;
        pieces[3 * k + 2] = p[k].gap;
    }
    return MH_OK;
}

}  // extern "C"
