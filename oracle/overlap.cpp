// CPU oracle of the recursive anchoring's match clean-up (TEST INFRASTRUCTURE
// ONLY: linked into oracle/liboracle.so, called from tests/ as the checker).
//
// EliminateOverlaps_v2 (LM/ProgressiveAligner.h:301-394) for matches of two
// genomes, then LengthFilter (LM/MatchList.h:680-693).  C++ because the
// reference sorts with std::sort, whose order of equal elements is libstdc++'s
// (the same header sort is used here on the same input order).
//
// Match semantics restated (LM/UngappedLocalAlignment.h:138-184,
// LM/HybridAbstractMatch.h:217-290): Start is signed (negative = reverse
// strand, 0 = absent), LeftEnd = |Start|; CropStart(n) shortens the match by
// n and moves forward starts right by n; CropEnd(n) shortens it and moves
// reverse starts (more negative) by n; CropLeft/CropRight pick CropStart or
// CropEnd by the strand in the given genome.  For two genomes the copies
// EliminateOverlaps_v2 makes of cropped-off parts keep one genome only after
// processNewMatch (LM/ProgressiveAligner.h:261-271) and are dropped.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <vector>

extern "C" {
#include "mh_oracle.h"
}

namespace {

struct M {
    int64_t start[2];
    int64_t len;
    bool dead;
};

int64_t left_end(const M &m, int s) { return std::llabs(m.start[s]); }

void crop_start(M &m, int64_t n)
{
    m.len -= n;
    for (int s = 0; s < 2; ++s)
        if (m.start[s] > 0) m.start[s] += n;
}

void crop_end(M &m, int64_t n)
{
    m.len -= n;
    for (int s = 0; s < 2; ++s)
        if (m.start[s] < 0) m.start[s] -= n;
}

void crop_left(M &m, int64_t n, int s) { m.start[s] >= 0 ? crop_start(m, n) : crop_end(m, n); }
void crop_right(M &m, int64_t n, int s) { m.start[s] >= 0 ? crop_end(m, n) : crop_start(m, n); }

int multiplicity(const M &m) { return (m.start[0] != 0) + (m.start[1] != 0); }

}  // namespace

extern "C" int64_t or_eliminate_overlaps_pairwise(or_match *ml, int64_t n, uint64_t min_length)
{
    std::vector<M> v((size_t)n);
    for (int64_t i = 0; i < n; ++i) v[i] = M{{ml[i].start[0], ml[i].start[1]}, (int64_t)ml[i].length, false};
    std::vector<int> idx((size_t)n);
    for (int64_t i = 0; i < n; ++i) idx[i] = (int)i;
    if (n >= 2) {
        for (int s = 0; s < 2; ++s) {
            std::sort(idx.begin(), idx.end(), [&](int a, int b) {
                const int64_t as = left_end(v[a], s), bs = left_end(v[b], s);
                if (as == 0 || bs == 0) return bs != 0;
                return as < bs;
            });
            const int64_t cnt = (int64_t)idx.size();
            int64_t i = 0;
            for (; i < cnt; ++i)
                if (v[idx[i]].start[s] != 0) break;
            for (; i < cnt; ++i) {
                if (v[idx[i]].dead) continue;
                for (int64_t j = i + 1; j < cnt; ++j) {
                    if (v[idx[j]].dead) continue;
                    M &a = v[idx[i]], &b = v[idx[j]];
                    bool deleted_a = false;
                    int64_t diff = left_end(b, s) - left_end(a, s) - a.len;
                    if (diff >= 0) break;
                    diff = -diff;
                    const bool a_smaller = multiplicity(b) > multiplicity(a) ||
                                           (multiplicity(b) == multiplicity(a) && b.len > a.len);
                    if (a_smaller) {
                        if (diff >= a.len) {
                            a.dead = true;
                            --i;
                            deleted_a = true;
                        } else {
                            crop_right(a, diff, s);
                        }
                    } else {
                        if (diff >= b.len)
                            b.dead = true;
                        else
                            crop_left(b, diff, s);
                    }
                    if (deleted_a) break;
                }
            }
            // compaction keeps the sorted order
            std::vector<int> keep;
            for (int k : idx)
                if (!v[k].dead) keep.push_back(k);
            idx.swap(keep);
        }
    }
    int64_t out = 0;
    for (int k : idx)
        if ((uint64_t)v[k].len >= min_length) {
            ml[out].start[0] = v[k].start[0];
            ml[out].start[1] = v[k].start[1];
            for (int s = 2; s < OR_MAX_SEQ; ++s) ml[out].start[s] = 0;
            ml[out].length = (uint64_t)v[k].len;
            ++out;
        }
    return out;
}
