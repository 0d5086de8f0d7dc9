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

#include <algorithm>
#include <atomic>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

using namespace mh;

namespace {

// ' ', '\t', '\n', '\v', '\f', '\r' (bits 32 and 9-13)
inline bool is_ws(char c)
{
    const uint32_t u = (uint8_t)c;
    return u <= 32 && ((0x100003E00ull >> u) & 1ull);
}
inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

// '-' among the 8 bytes at p (exact zero-byte test of p ^ '-').
inline unsigned gaps8(const char *p)
{
    uint64_t v;
    std::memcpy(&v, p, 8);
    v ^= 0x2D2D2D2D2D2D2D2Dull;
    const uint64_t lo7 = 0x7F7F7F7F7F7F7F7Full;
    return (unsigned)__builtin_popcountll(~(((v & lo7) + lo7) | v | lo7));
}

struct Entry {
    int64_t num = 0, start = 0;
    const char *seq = nullptr, *seq_end = nullptr;   // raw sequence text (with whitespace)
};

// The first c in [p, end), or end: one memchr for short ranges, else the
// lowest hit among 1 MiB pieces searched in parallel.
const char *find_char(const char *p, const char *end, char c)
{
    const uint64_t n = (uint64_t)(end - p), piece = 1 << 20;
    if (n <= 2 * piece) {
        const void *hit = std::memchr(p, c, (size_t)n);
        return hit ? static_cast<const char *>(hit) : end;
    }
    const uint64_t np = (n + piece - 1) / piece;
    std::vector<uint64_t> at(np, n);
    host_parallel(np, [&](size_t k) {
        const uint64_t lo = k * piece, len = std::min(piece, n - lo);
        const void *hit = std::memchr(p + lo, c, (size_t)len);
        if (hit) at[k] = (uint64_t)(static_cast<const char *>(hit) - p);
    });
    for (uint64_t k = 0; k < np; ++k)
        if (at[k] < n) return p + at[k];
    return end;
}

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
    q = find_char(q, end, '>');
    if (q == s) return nullptr;
    e.num = v[0];
    e.start = v[1];
    e.seq = s;
    e.seq_end = q;
    return q;
}

template <class F> void parallel_chunks(uint64_t nc, F f) { host_parallel((size_t)nc, std::function<void(size_t)>(f)); }

// The sequence of an entry without whitespace, into out: per chunk a count of
// the characters kept, then each chunk copies to its offset.
struct Row {
    const char *c;
    uint64_t n;
};

// True when none of the 8 bytes at p is below 33 (so none is whitespace).
inline bool no_ws8(const char *p)
{
    uint64_t v;
    std::memcpy(&v, p, 8);
    return ((v - 0x2121212121212121ull) & ~v & 0x8080808080808080ull) == 0;
}

Row squeeze(const Entry &e, HostBuf &buf)
{
    const uint64_t n = (uint64_t)(e.seq_end - e.seq), chunk = 1 << 18, nc = (n + chunk - 1) / chunk;
    std::vector<uint64_t> at(nc + 1, 0);
    parallel_chunks(nc, [&](uint64_t k) {
        const char *c = e.seq + k * chunk, *ce = e.seq + std::min(n, (k + 1) * chunk);
        uint64_t kept = 0;
        for (; c + 8 <= ce; c += 8) {
            if (no_ws8(c)) {
                kept += 8;
                continue;
            }
            for (int x = 0; x < 8; ++x) kept += !is_ws(c[x]);
        }
        for (; c < ce; ++c) kept += !is_ws(*c);
        at[k + 1] = kept;
    });
    for (uint64_t k = 0; k < nc; ++k) at[k + 1] += at[k];
    char *out = buf.get<char>(at[nc] + 8);
    parallel_chunks(nc, [&](uint64_t k) {
        const char *c = e.seq + k * chunk, *ce = e.seq + std::min(n, (k + 1) * chunk);
        char *w = out + at[k];
        for (; c + 8 <= ce; c += 8) {
            if (no_ws8(c)) {
                std::memcpy(w, c, 8);
                w += 8;
                continue;
            }
            for (int x = 0; x < 8; ++x) {
                *w = c[x];
                w += !is_ws(c[x]);
            }
        }
        for (; c < ce; ++c) {
            *w = *c;
            w += !is_ws(*c);
        }
    });
    return Row{out, at[nc]};
}

// The column walk of one group (PY/buildindex.py:119-129): each index
// advances on a non-gap character, and equal characters map sequence 1's
// index to sequence 2's.  Long rows are walked in parallel chunks from the
// indices each chunk starts at; a chunk's writes before its first residue of
// sequence 1 (all-gap columns) reuse the previous chunk's last index and are
// applied afterwards in column order, so later columns still win.
void walk(Row gr, Row rr, int64_t g0, int64_t r0, uint64_t genome_length, int32_t *lut)
{
    const char *gs = gr.c, *rs = rr.c;
    const uint64_t n = gr.n, chunk = 1 << 18, nc = (n + chunk - 1) / chunk;
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
        uint64_t c = lo;
        for (; c + 8 <= hi; c += 8) {
            a += 8 - (int64_t)gaps8(gs + c);
            b += 8 - (int64_t)gaps8(rs + c);
        }
        for (; c < hi; ++c) {
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
        uint64_t c = lo;
        while (c < hi) {
            // eight equal columns without gaps: both indices step by one each,
            // and every column maps (none is before the chunk's first residue)
            if (c + 8 <= hi && gi >= -1 && gi + 8 < (int64_t)genome_length) {
                uint64_t va, vb;
                std::memcpy(&va, gs + c, 8);
                std::memcpy(&vb, rs + c, 8);
                if (va == vb && gaps8(gs + c) == 0) {
                    for (int x = 1; x <= 8; ++x) lut[gi + x] = (int32_t)(ri + x);
                    gi += 8;
                    ri += 8;
                    c += 8;
                    continue;
                }
            }
            const char a = gs[c], b = rs[c];
            gi += a != '-';
            ri += b != '-';
            if (a == b && gi < (int64_t)genome_length) {
                if (gi == cg[k]) early[k].push_back({gi, ri});
                else put(gi, ri);
            }
            ++c;
        }
    });
    for (uint64_t k = 0; k < nc; ++k)
        for (const auto &e : early[k]) put(e.first, e.second);
}

// Starting at a gap start idx whose left neighbour lower = idx-1 is mapped:
// the first upper in (idx, idx + width] with lut[upper] - lut[lower] ==
// upper - lower, or 0 if none.  Filling sets [lower, upper] to consecutive
// values (its two ends keep theirs).
inline uint64_t fill_end(const int32_t *lut, uint64_t n, uint64_t idx, int64_t width)
{
    const uint64_t lower = idx - 1;
    const int64_t lm = lut[lower];
    const uint64_t stop = std::min<uint64_t>(n, idx + (uint64_t)width + 1);
    for (uint64_t up = idx + 1; up < stop; ++up)
        if ((int64_t)lut[up] - lm == (int64_t)(up - lower)) return up;
    return 0;
}

// A left-to-right pass that, at every idx where trig(lut, idx) holds, fills
// from idx as above (indexutils.pyx's fillGaps and smoothEdges).  In the
// sequential pass the entries a fill writes inside its range no longer
// trigger, and every trigger after a fill's range, and the fill it finds, read
// only values no fill has changed (the range ends keep their values).  So the
// fills are found in parallel on the input, kept left to right unless they
// start inside a kept fill, and written in parallel (the kept ranges overlap
// at most at their ends).
template <class Trig> void fill_pass(int32_t *lut, uint64_t n, int64_t width, Trig trig)
{
    if (n < 2) return;
    const uint64_t chunk = 1 << 18, nc = (n + chunk - 1) / chunk;
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> found(nc);
    parallel_chunks(nc, [&](uint64_t k) {
        const uint64_t hi = std::min(n, (k + 1) * chunk);
        for (uint64_t idx = std::max<uint64_t>(1, k * chunk); idx < hi; ++idx)
            if (trig(lut, idx)) {
                const uint64_t up = fill_end(lut, n, idx, width);
                if (up) found[k].push_back({idx - 1, up});
            }
    });
    std::vector<std::pair<uint64_t, uint64_t>> keep;
    uint64_t last_up = 0;
    for (const auto &v : found)
        for (const auto &f : v)
            if (f.first + 1 > last_up) {
                keep.push_back(f);
                last_up = f.second;
            }
    const uint64_t per = 4096, nk = (keep.size() + per - 1) / per;
    parallel_chunks(nk, [&](uint64_t k) {
        for (uint64_t i = k * per; i < std::min<uint64_t>(keep.size(), (k + 1) * per); ++i) {
            const uint64_t lower = keep[i].first, up = keep[i].second;
            const int64_t lm = lut[lower];
            for (uint64_t x = lower + 1; x < up; ++x) lut[x] = (int32_t)(lm + (int64_t)(x - lower));
        }
    });
}

}  // namespace

extern "C" {

int mh_index_from_xmfa(const char *xmfa, uint64_t len, uint64_t genome_length, int32_t *lut, uint64_t *n_groups)
{
    MH_CHECK((xmfa || len == 0) && (lut || genome_length == 0), MH_ERR_INVALID, "null argument");
    const uint64_t fill_chunk = 1 << 20;
    parallel_chunks((genome_length + fill_chunk - 1) / fill_chunk, [&](uint64_t k) {
        const uint64_t hi = std::min(genome_length, (k + 1) * fill_chunk);
        for (uint64_t x = k * fill_chunk; x < hi; ++x) lut[x] = -1;
    });
    const char *end = xmfa + len;
    const char *p = xmfa;
    uint64_t groups = 0;
    // the squeezed rows: buffers kept between calls (one call at a time)
    static std::mutex mu;
    static HostBuf rows[2];
    std::lock_guard<std::mutex> lock(mu);
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
        const char *eq = find_char(gt, end, '=');
        if (eq == end) break;
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
        const Row gs = squeeze(g, rows[0]), rs = squeeze(r, rows[1]);
        MH_CHECK(rs.n >= gs.n, MH_ERR_INVALID, "string index out of range (genome 2 row shorter than genome 1 row)");
        walk(gs, rs, g.start - 1, r.start - 1, genome_length, lut);
    }
    if (n_groups) *n_groups = groups;
    return MH_OK;
}

int mh_index_from_xmfa_file(const char *path, uint64_t genome_length, int32_t *lut, uint64_t *n_groups)
{
    MH_CHECK(path, MH_ERR_INVALID, "null argument");
    const int fd = ::open(path, O_RDONLY);
    MH_CHECK(fd >= 0, MH_ERR_IO, std::string("cannot open ") + path);
    struct stat st;
    if (::fstat(fd, &st) != 0) {
        ::close(fd);
        MH_CHECK(false, MH_ERR_IO, std::string("cannot stat ") + path);
    }
    const uint64_t n = (uint64_t)st.st_size;
    static std::mutex mu;
    static HostBuf text;
    std::lock_guard<std::mutex> lock(mu);
    char *buf = text.get<char>(n);
    if (!buf) {
        ::close(fd);
        MH_CHECK(false, MH_ERR_NOMEM, "out of host memory");
    }
    // read in parallel 4 MiB pieces
    const uint64_t piece = 1ull << 20, np = (n + piece - 1) / piece;
    std::atomic<int> bad{0};
    parallel_chunks(np, [&](uint64_t k) {
        uint64_t off = k * piece;
        const uint64_t end = std::min(n, off + piece);
        while (off < end) {
            const ssize_t got = ::pread(fd, buf + off, (size_t)(end - off), (off_t)off);
            if (got <= 0) {
                bad = 1;
                return;
            }
            off += (uint64_t)got;
        }
    });
    ::close(fd);
    MH_CHECK(!bad, MH_ERR_IO, std::string("cannot read ") + path);
    return mh_index_from_xmfa(buf, n, genome_length, lut, n_groups);
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
    fill_pass(lut, n, max_gap_width,
              [](const int32_t *t, uint64_t idx) { return t[idx] == -1 && t[idx - 1] != -1; });
    return MH_OK;
}

int mh_index_smooth_edges(int32_t *lut, uint64_t n, int smoothing_radius)
{
    MH_CHECK(lut || n == 0, MH_ERR_INVALID, "null argument");
    fill_pass(lut, n, smoothing_radius, [](const int32_t *t, uint64_t idx) {
        return (int64_t)t[idx] != (int64_t)t[idx - 1] + 1 && t[idx - 1] != -1;
    });
    return MH_OK;
}

}  // extern "C"
