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
    case 'T': case 't': return 3;
    default: return 99;
    }
}

// LetterObjScoreXP's per-column score for two single-sequence profiles
// (weights 1): the NUC_SP letter score, plus each gap's penalty spread
// evenly over the gap's columns (MU/anchoredpp.cpp:16-316).
void column_scores(const char *r1, const char *r2, unsigned n, std::vector<float> &score)
{
    std::vector<float> mm(n, 0.0f), gg(n, 0.0f);
    unsigned start = 0, end = n - 1;
    for (unsigned c = 0; c < n; ++c)
        if (!(mu_gap(r1[c]) && mu_gap(r2[c]))) {
            start = c;
            break;
        }
    for (int c = (int)n - 1; c >= 0; --c)
        if (!(mu_gap(r1[c]) && mu_gap(r2[c]))) {
            end = (unsigned)c;
            break;
        }
    for (unsigned c = start; c <= end; ++c) {
        const unsigned a = letter(r1[c]);
        if (a >= 4) continue;
        const unsigned b = letter(r2[c]);
        if (b >= 4) continue;
        mm[c] = (float)HOXD[a][b] + 60.0f;
    }
    bool g1 = false, g2 = false;
    unsigned left = 0;
    float cur = 0.0f;
    for (unsigned c = start; c <= end; ++c) {
        const bool x = mu_gap(r1[c]), y = mu_gap(r2[c]);
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
            for (unsigned k = left; k < c; ++k) gg[k] = per;
            left = INSANE;
            cur = 0.0f;
        }
        g1 = g2 = false;
    }
    if (g1 || g2) {
        cur -= GAP_OPEN;
        cur += TERM_GAP;
        const float per = cur / (float)(n - left);
        for (unsigned k = left; k < n; ++k) gg[k] = per;
    }
    score.assign(n, 0.0f);
    for (unsigned c = 0; c < n; ++c) score[c] += 1.0f * (mm[c] + gg[c]);
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
    std::vector<float> score, smooth;
    column_scores(r1, r2, n, score);
    window_smooth(score, smooth);
    std::vector<unsigned> best;
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

// Counts of residues ('-' is the only gap in Mauve's GappedAlignment) in
// each row before every column, so a piece's lengths are two subtractions.
struct Prefix {
    std::vector<uint32_t> c1, c2;
    void build(const char *r1, const char *r2, uint64_t n)
    {
        c1.resize(n + 1);
        c2.resize(n + 1);
        c1[0] = c2[0] = 0;
        for (uint64_t c = 0; c < n; ++c) {
            c1[c + 1] = c1[c] + (r1[c] != '-');
            c2[c + 1] = c2[c] + (r2[c] != '-');
        }
    }
    uint64_t l1(uint64_t lo, uint64_t hi) const { return c1[hi] - c1[lo]; }
    uint64_t l2(uint64_t lo, uint64_t hi) const { return c2[hi] - c2[lo]; }
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
    uint64_t gap_cols = 0, base = 0, last = 0;
    bool have_last = false;
    for (uint64_t col = 0; col < n; ++col) {
        if (r1[col] != '-' && r2[col] != '-') {
            if (gap_cols > MAX_GAP_LENGTH) {
                if (have_last) {
                    const uint64_t sp = last + LCB_HANGOVER - base;
                    pieces.push_back({base, base + sp, false});
                    base += sp;
                }
                const uint64_t sp = col - LCB_HANGOVER - base;
                pieces.push_back({base, base + sp, true});
                base += sp;
            }
            last = col;
            have_last = true;
            gap_cols = 0;
        } else {
            ++gap_cols;
        }
    }
    if (gap_cols > MAX_GAP_LENGTH) {
        if (have_last) {
            const uint64_t sp = last + LCB_HANGOVER - base;
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

// The output of one piece in one pass, as a list of segments per row: a
// segment is literal text or the A/B row of one of the piece's DP jobs.
struct Seg {
    int64_t job;               // -1: literal
    std::string text;
};
struct Task {
    size_t iv;
    RefinePiece p;
    bool whole;                // the interval is copied as it is
    int rows = 2;              // 2, or 1 for a piece re-aligned from one row
    std::vector<Seg> row[2];
    std::vector<std::string> a, b;   // residues of the piece's DP jobs
    std::string out[2];
    int err = MH_OK;
    std::string msg;
};

std::string residues(const char *r, uint64_t lo, uint64_t hi)
{
    std::string s;
    s.reserve(hi - lo);
    for (uint64_t c = lo; c < hi; ++c)
        if (!mu_gap(r[c])) s.push_back(r[c]);
    return s;
}

void literal(Task &t, int k, std::string s) { t.row[k].push_back(Seg{-1, std::move(s)}); }
void add_job(Task &t, std::string a, std::string b)
{
    const int64_t j = (int64_t)t.a.size();
    t.a.push_back(std::move(a));
    t.b.push_back(std::move(b));
    t.row[0].push_back(Seg{j, std::string()});
    t.row[1].push_back(Seg{j, std::string()});
}

// Plan one piece for one pass.
void plan_piece(const char *r1, const char *r2, const Prefix &pf, bool profile, Task &t)
{
    const RefinePiece &p = t.p;
    const uint64_t len = p.hi - p.lo;
    if (t.whole) {
        literal(t, 0, std::string(r1 + p.lo, len));
        literal(t, 1, std::string(r2 + p.lo, len));
        return;
    }
    const uint64_t l1 = pf.l1(p.lo, p.hi), l2 = pf.l2(p.lo, p.hi);
    if (p.gap || (profile && !(l1 > 0 && l2 > 0))) {
        literal(t, 0, std::string(r1 + p.lo, len));
        literal(t, 1, std::string(r2 + p.lo, len));
        return;
    }
    if (profile) {
        // AnchoredProfileProfile: the ranges between anchor columns
        if (len >= (1ull << 32)) {
            t.err = MH_ERR_UNSUPPORTED;
            t.msg = "piece longer than 2^32 columns";
            return;
        }
        std::vector<unsigned> anchors;
        refine_anchor_cols(r1 + p.lo, r2 + p.lo, (unsigned)len, anchors);
        for (size_t k = 0; k <= anchors.size(); ++k) {
            const uint64_t lo = p.lo + (k > 0 ? anchors[k - 1] : 0);
            const uint64_t hi = p.lo + (k < anchors.size() ? anchors[k] : len);
            if (hi == lo) continue;
            std::string a = residues(r1, lo, hi), b = residues(r2, lo, hi);
            if (a.empty() || b.empty()) {
                t.err = MH_ERR_ZERO_LENGTH;
                t.msg = "NWSmall: length=0 (an anchored range has residues of one genome only)";
                return;
            }
            add_job(t, std::move(a), std::move(b));
        }
        return;
    }
    const int d = density_of(l1, l2, len);
    const char *rows[2] = {r1, r2};
    int present[2], np = 0;
    if (l1) present[np++] = 0;
    if (l2) present[np++] = 1;
    t.rows = np;
    if (d == 0) {
        for (int q = 0; q < np; ++q) literal(t, q, std::string(rows[present[q]] + p.lo, len));
        return;
    }
    const uint64_t w = d == 1 ? 500 : 200;
    for (uint64_t f = p.lo; f < p.hi; f += w) {
        const uint64_t e = std::min(f + w, p.hi);
        std::string sq[2];
        for (int q = 0; q < np; ++q) sq[q] = residues(rows[present[q]], f, e);
        if (np == 2 && !sq[0].empty() && !sq[1].empty()) {
            add_job(t, std::move(sq[0]), std::move(sq[1]));
            continue;
        }
        int full = -1;
        for (int q = 0; q < np; ++q)
            if (!sq[q].empty()) full = q;
        if (full < 0) {
            t.err = MH_ERR_INVALID;
            t.msg = "a refinement window holds no residues";
            return;
        }
        const size_t m = sq[full].size();
        for (int q = 0; q < np; ++q) literal(t, q, q == full ? sq[full] : std::string(m, '-'));
    }
}

// Aligned A and B rows of a job from its path.
void job_text(const std::string &a, const std::string &b, const uint8_t *path, uint32_t plen, std::string &ra,
              std::string &rb)
{
    ra.clear();
    rb.clear();
    size_t i = 0, j = 0;
    for (uint32_t e = 0; e < plen; ++e) {
        const uint8_t c = path[e];
        if (c == 'M') {
            ra.push_back(a[i++]);
            rb.push_back(b[j++]);
        } else if (c == 'D') {
            ra.push_back(a[i++]);
            rb.push_back('-');
        } else {
            ra.push_back('-');
            rb.push_back(b[j++]);
        }
    }
}

template <class F> void parallel_for(size_t n, F f)
{
    if (n == 0) return;
    const size_t nt = std::max<size_t>(1, std::min<size_t>(n, std::max(1u, std::thread::hardware_concurrency())));
    if (nt == 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<size_t> next(0);
    for (size_t t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto &x : th) x.join();
}

// One pass over all intervals: plan (in parallel over pieces), one batched
// DP, splice (in parallel over pieces), reassemble.
int refine_pass(mh_context *ctx, std::vector<std::string> &r1, std::vector<std::string> &r2, bool profile)
{
    const size_t n = r1.size();
    std::vector<Prefix> pf(n);
    std::vector<std::vector<RefinePiece>> pcs(n);
    std::vector<char> whole(n, 0);
    parallel_for(n, [&](size_t i) {
        const uint64_t len = r1[i].size();
        pf[i].build(r1[i].data(), r2[i].data(), len);
        // an interval that holds one genome only is not refined (the
        // Multiplicity() > 1 test of doGappedAlignment, LM/ProgressiveAligner.cpp:1268)
        if (pf[i].l1(0, len) == 0 || pf[i].l2(0, len) == 0) {
            whole[i] = 1;
            pcs[i] = {RefinePiece{0, len, false}};
            return;
        }
        pieces_fast(r1[i].data(), r2[i].data(), len, pf[i], pcs[i]);
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
    parallel_for(tasks.size(), [&](size_t k) {
        Task &t = tasks[k];
        plan_piece(r1[t.iv].data(), r2[t.iv].data(), pf[t.iv], profile, t);
    });
    for (const Task &t : tasks)
        if (t.err != MH_OK) {
            set_error(t.msg + " (interval " + std::to_string(t.iv) + ")");
            return t.err;
        }
    // every job of the pass in one batch, in task order
    std::vector<uint64_t> jbase(tasks.size() + 1, 0), abase(tasks.size() + 1, 0), bbase(tasks.size() + 1, 0);
    for (size_t k = 0; k < tasks.size(); ++k) {
        jbase[k + 1] = jbase[k] + tasks[k].a.size();
        uint64_t sa = 0, sb = 0;
        for (size_t j = 0; j < tasks[k].a.size(); ++j) {
            MH_CHECK(tasks[k].a[j].size() < (1ull << 31) && tasks[k].b[j].size() < (1ull << 31), MH_ERR_UNSUPPORTED,
                     "DP operand longer than 2^31");
            sa += tasks[k].a[j].size();
            sb += tasks[k].b[j].size();
        }
        abase[k + 1] = abase[k] + sa;
        bbase[k + 1] = bbase[k] + sb;
    }
    const uint64_t njobs = jbase[tasks.size()], abytes = abase[tasks.size()], bbytes = bbase[tasks.size()];
    std::vector<mh_dp_job> jobs(njobs);
    std::vector<uint8_t> ac(abytes + 1), bc(bbytes + 1);
    std::atomic<int> bad(0);
    parallel_for(tasks.size(), [&](size_t k) {
        const Task &t = tasks[k];
        uint64_t pa = abase[k], pb = bbase[k];
        for (size_t j = 0; j < t.a.size(); ++j) {
            mh_dp_job &jb = jobs[jbase[k] + j];
            jb.a_off = pa;
            jb.b_off = pb;
            jb.a_len = (uint32_t)t.a[j].size();
            jb.b_len = (uint32_t)t.b[j].size();
            for (char c : t.a[j]) {
                const unsigned l = letter(c);
                if (l >= 4) bad = 1;
                ac[pa++] = (uint8_t)(l & 3);
            }
            for (char c : t.b[j]) {
                const unsigned l = letter(c);
                if (l >= 4) bad = 1;
                bc[pb++] = (uint8_t)(l & 3);
            }
        }
    });
    MH_CHECK(!bad, MH_ERR_UNSUPPORTED, "residue other than A,C,G,T in a DP operand");
    uint64_t total = 0;
    for (const mh_dp_job &jb : jobs) total += (uint64_t)jb.a_len + jb.b_len;
    std::vector<uint8_t> paths(total + 1);
    std::vector<mh_dp_result> res(njobs);
    if (njobs) {
        const int rc = mh_dp_align(ctx, ac.data(), abytes, bc.data(), bbytes, jobs.data(), njobs, paths.data(),
                                   res.data());
        if (rc != MH_OK) return rc;
    }
    // splice each piece
    parallel_for(tasks.size(), [&](size_t k) {
        Task &t = tasks[k];
        std::string ra, rb;
        for (int q = 0; q < t.rows; ++q)
            for (const Seg &sg : t.row[q]) {
                if (sg.job < 0) {
                    t.out[q] += sg.text;
                    continue;
                }
                const uint64_t g = jbase[k] + (uint64_t)sg.job;
                job_text(t.a[sg.job], t.b[sg.job], paths.data() + res[g].path_start, res[g].path_len, ra, rb);
                t.out[q] += q == 0 ? ra : rb;
            }
    });
    // reassemble each interval (LM/ProgressiveAligner.cpp:1222-1237)
    parallel_for(n, [&](size_t i) {
        uint64_t new_len = 0;
        for (size_t k = first[i]; k < first[i + 1]; ++k) new_len += tasks[k].out[0].size();
        std::string rows[2] = {std::string(new_len, '-'), std::string(new_len, '-')};
        uint64_t pos[2] = {0, 0};
        for (size_t k = first[i]; k < first[i + 1]; ++k)
            for (int q = 0; q < tasks[k].rows; ++q) {
                rows[q].replace(pos[q], tasks[k].out[q].size(), tasks[k].out[q]);
                pos[q] += tasks[k].out[q].size();
            }
        rows[0].resize(new_len, '-');
        rows[1].resize(new_len, '-');
        r1[i].swap(rows[0]);
        r2[i].swap(rows[1]);
    });
    return MH_OK;
}

}  // namespace

extern "C" {

int mh_refine_pairwise(mh_context *ctx, uint64_t n, const char *const *rows1, const char *const *rows2,
                       const uint64_t *cols, uint32_t passes, char *const *out1, char *const *out2,
                       const uint64_t *caps, uint64_t *out_cols)
{
    MH_CHECK(ctx && (n == 0 || (rows1 && rows2 && cols && out1 && out2 && caps && out_cols)), MH_ERR_INVALID,
             "null argument");
    MH_CHECK((passes & ~3u) == 0, MH_ERR_INVALID, "unknown refinement pass");
    std::vector<std::string> r1(n), r2(n);
    for (uint64_t i = 0; i < n; ++i) {
        MH_CHECK(rows1[i] && rows2[i], MH_ERR_INVALID, "null row");
        r1[i].assign(rows1[i], cols[i]);
        r2[i].assign(rows2[i], cols[i]);
    }
    int rc;
    if ((passes & MH_REFINE_PROFILE) && (rc = refine_pass(ctx, r1, r2, true)) != MH_OK) return rc;
    if ((passes & MH_REFINE_WINDOWS) && (rc = refine_pass(ctx, r1, r2, false)) != MH_OK) return rc;
    for (uint64_t i = 0; i < n; ++i) out_cols[i] = r1[i].size();
    for (uint64_t i = 0; i < n; ++i)
        MH_CHECK(r1[i].size() <= caps[i], MH_ERR_INVALID,
                 "output capacity too small for interval " + std::to_string(i) + " (out_cols holds the sizes)");
    for (uint64_t i = 0; i < n; ++i) {
        std::memcpy(out1[i], r1[i].data(), r1[i].size());
        std::memcpy(out2[i], r2[i].data(), r2[i].size());
    }
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
    Prefix pf;
    pf.build(row1, row2, cols);
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
