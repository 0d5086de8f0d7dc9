// buildIndex's lookup table: the XMFA column walk and the three sequential
// post-passes of PY/buildindex.py:110-136 and PY/indexutils.pyx:27-108.
//
// The LUT maps each position of genome 1 (0-based) to a position of genome 2,
// -1 where unmapped.  The passes modify it in place, left to right, and later
// steps read what earlier steps wrote, so they run on the host (one pass over
// a few MB; the work per entry is a compare except at gap starts).
//
// XMFA grammar (PY/xmfa.py:15-68, the readable stand-in for the parser the
// reference imports): a group runs from a '>' at the start of a line to the
// first '=' after it, and the rest of that line is the group's notes.  Inside
// a group, an entry starts at a '>' at the start of a line:
//     '>' ws* digits ':' digits '-' digits ws* [+|-] ws* notes
// where notes is a run of at least one character that is a space or not
// whitespace, followed by optional whitespace up to the end of the line; the
// entry's sequence is every character up to the next '>' (or the group's
// end), whitespace removed, and it must not be empty.  A '>' line that does
// not fit is skipped.
#include "mh_internal.hpp"

#include <cstring>
#include <thread>

using namespace mh;

namespace {

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }
inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

struct Entry {
    int64_t num = 0, start = 0;
    const char *seq = nullptr, *seq_end = nullptr;   // raw sequence text (with whitespace)
};

// Parses the entry whose '>' is at p (a line start inside [.., end)); on a
// match sets e and returns the position after the entry's sequence, else
// nullptr.
const char *parse_entry(const char *p, const char *end, Entry &e)
{
    const char *q = p + 1;
    while (q < end && is_ws(*q)) ++q;
    int64_t v[3] = {0, 0, 0};
    for (int f = 0; f < 3; ++f) {
        if (q >= end || !is_digit(*q)) return nullptr;
        int64_t x = 0;
        while (q < end && is_digit(*q)) {
            if (x < (int64_t)1 << 58) x = x * 10 + (*q - '0');
            ++q;
        }
        v[f] = x;
        if (f < 2) {
            if (q >= end || *q != (f == 0 ? ':' : '-')) return nullptr;
            ++q;
        }
    }
    while (q < end && is_ws(*q)) ++q;
    if (q >= end || (*q != '+' && *q != '-' && *q != '|')) return nullptr;
    ++q;
    while (q < end && is_ws(*q)) ++q;
    // notes: one or more of ' ' or non-whitespace
    const char *n0 = q;
    while (q < end && (*q == ' ' || !is_ws(*q))) ++q;
    if (q == n0) return nullptr;
    // optional whitespace, then the end of the line
    while (q < end && *q != '\n' && is_ws(*q)) ++q;
    if (q < end && *q != '\n') return nullptr;
    // the sequence: everything up to the next '>' or the group's end
    const char *s = q;
    while (q < end && *q != '>') ++q;
    if (q == s) return nullptr;
    e.num = v[0];
    e.start = v[1];
    e.seq = s;
    e.seq_end = q;
    return q;
}

// The sequence of an entry without whitespace, into out.
void squeeze(const Entry &e, std::string &out)
{
    out.resize((size_t)(e.seq_end - e.seq));
    char *w = &out[0];
    for (const char *c = e.seq; c < e.seq_end; ++c) {
        *w = *c;
        w += !is_ws(*c);
    }
    out.resize((size_t)(w - out.data()));
}

template <class F> void parallel_chunks(uint64_t nc, F f)
{
    const unsigned nt = std::max(1u, std::min<unsigned>((unsigned)nc, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (uint64_t k = t; k < nc; k += nt) f(k);
        });
    for (auto &x : th) x.join();
}

// The column walk of one group (PY/buildindex.py:119-129): each index
// advances on a non-gap character, and equal characters map sequence 1's
// index to sequence 2's.  Long rows are walked in parallel chunks from the
// indices each chunk starts at; a chunk's writes before its first residue of
// sequence 1 (all-gap columns) reuse the previous chunk's last index and are
// applied afterwards in column order, so later columns still win.
void walk(const std::string &gs, const std::string &rs, int64_t g0, int64_t r0, uint64_t genome_length, int32_t *lut)
{
    const uint64_t n = gs.size(), chunk = 1 << 18, nc = (n + chunk - 1) / chunk;
    auto put = [&](int64_t gi, int64_t ri) {
        // index -1 is the last entry (numpy's negative indexing)
        const int64_t at = gi >= 0 ? gi : (int64_t)genome_length + gi;
        if (at >= 0) lut[at] = (int32_t)ri;
    };
    if (nc <= 1 || g0 < 0) {
        int64_t gi = g0, ri = r0;
        for (uint64_t c = 0; c < n; ++c) {
            const char a = gs[c], b = rs[c];
            gi += a != '-';
            ri += b != '-';
            if (a == b && gi < (int64_t)genome_length) put(gi, ri);
        }
        return;
    }
    std::vector<int64_t> cg(nc + 1, 0), cr(nc + 1, 0);
    parallel_chunks(nc, [&](uint64_t k) {
        const uint64_t lo = k * chunk, hi = std::min(n, lo + chunk);
        int64_t a = 0, b = 0;
        for (uint64_t c = lo; c < hi; ++c) {
            a += gs[c] != '-';
            b += rs[c] != '-';
        }
        cg[k + 1] = a;
        cr[k + 1] = b;
    });
    cg[0] = g0;
    cr[0] = r0;
    for (uint64_t k = 0; k < nc; ++k) {
        cg[k + 1] += cg[k];
        cr[k + 1] += cr[k];
    }
    std::vector<std::vector<std::pair<int64_t, int64_t>>> early(nc);
    parallel_chunks(nc, [&](uint64_t k) {
        const uint64_t lo = k * chunk, hi = std::min(n, lo + chunk);
        int64_t gi = cg[k], ri = cr[k];
        for (uint64_t c = lo; c < hi; ++c) {
            const char a = gs[c], b = rs[c];
            gi += a != '-';
            ri += b != '-';
            if (a == b && gi < (int64_t)genome_length) {
                if (gi == cg[k]) early[k].push_back({gi, ri});
                else put(gi, ri);
            }
        }
    });
    for (uint64_t k = 0; k < nc; ++k)
        for (const auto &e : early[k]) put(e.first, e.second);
}

// Starting at a gap start idx whose left neighbour lower = idx-1 is mapped:
// the first upper in (idx, idx + width] with lut[upper] - lut[lower] ==
// upper - lower gets [lower, upper] filled with consecutive values.
inline void fill_from(int32_t *lut, uint64_t n, uint64_t idx, int64_t width)
{
    const uint64_t lower = idx - 1;
    const int64_t lm = lut[lower];
    for (uint64_t up = idx + 1; up < n; ++up) {
        if ((int64_t)(up - idx) > width) return;
        if ((int64_t)lut[up] - lm == (int64_t)(up - lower)) {
            for (uint64_t k = lower; k <= up; ++k) lut[k] = (int32_t)(lm + (int64_t)(k - lower));
            return;
        }
    }
}

}  // namespace

extern "C" {

int mh_index_from_xmfa(const char *xmfa, uint64_t len, uint64_t genome_length, int32_t *lut, uint64_t *n_groups)
{
    MH_CHECK((xmfa || len == 0) && (lut || genome_length == 0), MH_ERR_INVALID, "null argument");
    for (uint64_t k = 0; k < genome_length; ++k) lut[k] = -1;
    const char *end = xmfa + len;
    const char *p = xmfa;
    uint64_t groups = 0;
    std::string gs, rs;
    while (p < end) {
        // next '>' at a line start
        const char *gt = nullptr;
        for (const char *q = p; q < end; ++q) {
            if (*q == '>' && (q == xmfa || q[-1] == '\n')) {
                gt = q;
                break;
            }
            const void *nl = std::memchr(q, '\n', (size_t)(end - q));
            if (!nl) break;
            q = static_cast<const char *>(nl);
        }
        if (!gt) break;
        const char *eq = static_cast<const char *>(std::memchr(gt, '=', (size_t)(end - gt)));
        if (!eq) break;
        if (eq == gt + 1) {         // '>' then '=' directly: no group starts here
            p = gt + 1;
            continue;
        }
        const char *eol = static_cast<const char *>(std::memchr(eq, '\n', (size_t)(end - eq)));
        p = eol ? eol : end;
        ++groups;
        // the first entries of genomes 1 and 2 in [gt, eq)
        Entry g, r;
        bool have_g = false, have_r = false;
        const char *q = gt;
        while (q < eq && !(have_g && have_r)) {
            Entry e;
            const char *nx = (*q == '>' && (q == gt || q[-1] == '\n')) ? parse_entry(q, eq, e) : nullptr;
            if (nx) {
                if (e.num == 1 && !have_g) { g = e; have_g = true; }
                else if (e.num == 2 && !have_r) { r = e; have_r = true; }
                q = nx;
                continue;
            }
            const void *nl = std::memchr(q, '\n', (size_t)(eq - q));
            if (!nl) break;
            q = static_cast<const char *>(nl) + 1;
        }
        if (!have_g || !have_r) continue;
        squeeze(g, gs);
        squeeze(r, rs);
        MH_CHECK(rs.size() >= gs.size(), MH_ERR_INVALID, "string index out of range (genome 2 row shorter than genome 1 row)");
        walk(gs, rs, g.start - 1, r.start - 1, genome_length, lut);
    }
    if (n_groups) *n_groups = groups;
    return MH_OK;
}

int mh_index_fix_zero(int32_t *lut, uint64_t n, const char *genome, uint64_t genome_len, const char *ref,
                      uint64_t ref_len)
{
    MH_CHECK(lut && n && genome && ref, MH_ERR_INVALID, "null argument");
    if (lut[0] != -1) return MH_OK;
    uint64_t nx = 1;
    while (nx < n && (lut[nx] == -1 || nx < 10)) ++nx;
    MH_CHECK(nx < n, MH_ERR_INVALID, "Out of bounds on buffer access (no mapped index at or after 10)");
    const int64_t r = (int64_t)lut[nx] - (int64_t)nx;
    if (r > -1) {
        MH_CHECK(genome_len > 0 && (uint64_t)r < ref_len, MH_ERR_INVALID, "string index out of range");
        if (genome[0] == ref[r]) lut[0] = (int32_t)r;
    }
    return MH_OK;
}

int mh_index_fill_gaps(int32_t *lut, uint64_t n, int max_gap_width)
{
    MH_CHECK(lut || n == 0, MH_ERR_INVALID, "null argument");
    for (uint64_t idx = 1; idx < n; ++idx)
        if (lut[idx] == -1 && lut[idx - 1] != -1) fill_from(lut, n, idx, max_gap_width);
    return MH_OK;
}

int mh_index_smooth_edges(int32_t *lut, uint64_t n, int smoothing_radius)
{
    MH_CHECK(lut || n == 0, MH_ERR_INVALID, "null argument");
    for (uint64_t idx = 1; idx < n; ++idx)
        if ((int64_t)lut[idx] != (int64_t)lut[idx - 1] + 1 && lut[idx - 1] != -1)
            fill_from(lut, n, idx, smoothing_radius);
    return MH_OK;
}

}  // extern "C"
