// Recursive anchoring in the gaps between anchors: pairwiseAnchorSearch,
// LM/ProgressiveAligner.cpp:590-679, for a list of gaps.
//
// Per gap (genome coordinates from getInterveningCoordinates, 1-based): the
// searched piece of each genome starts at gap_lend and is gap_rend - gap_lend
// - 1 long when that is positive (the reference's length, one short of the
// open interval); the seed is getSeed(getDefaultSeedWeight(average piece
// length), 0) and a gap whose weight is below MIN_DNA_SEED_WEIGHT yields no
// anchors.  Both pieces get a sorted mer list and a pairwise match search on
// the GPU (the MemHash search of two sequences finds what PairwiseMatchFinder
// finds: each run contributes the pair of its two unique seeds), then on the
// host EliminateOverlaps_v2 (LM/ProgressiveAligner.h:301-394), LengthFilter
// (MIN_ANCHOR_LENGTH + 3 = 12) and the shift back to genome coordinates
// (SetLeftEnd(LeftEnd + gap_lend - 1), strand kept).
#include "mh_internal.hpp"

#include <algorithm>
#include <string>
#include <vector>
#include <cstdlib>

using namespace mh;

namespace {

constexpr int MIN_DNA_SEED_WEIGHT = 5;        // LM/SeedMasks.h:375
constexpr uint64_t ANCHOR_MIN_LENGTH = 9 + 3; // MIN_ANCHOR_LENGTH + 3, LM/Aligner.h:266

inline int64_t left_end(const mh_match &m, int g) { return std::llabs(m.start[g]); }

// UngappedLocalAlignment crops (LM/UngappedLocalAlignment.h:138-184 with
// HybridAbstractMatch::MoveStart/MoveEnd, LM/HybridAbstractMatch.h:271-290).
void crop_start(mh_match &m, uint64_t n)
{
    m.length -= n;
    for (int g = 0; g < 2; ++g)
        if (m.start[g] > 0) m.start[g] += (int64_t)n;
}
void crop_end(mh_match &m, uint64_t n)
{
    m.length -= n;
    for (int g = 0; g < 2; ++g)
        if (m.start[g] < 0) m.start[g] -= (int64_t)n;
}

// EliminateOverlaps_v2 for matches of two genomes: in each genome, sorted by
// left end (std::sort, as the reference), an overlapping pair loses the
// overlap from the shorter match (the longer one wins ties of multiplicity;
// equal lengths crop the later one); a match covered entirely is dropped.
// Cropped-off copies keep one genome after processNewMatch and are dropped.
void eliminate_overlaps(std::vector<mh_match> &ml)
{
    if (ml.size() < 2) return;
    std::vector<mh_match *> v;
    for (mh_match &m : ml) v.push_back(&m);
    for (int g = 0; g < 2; ++g) {
        std::sort(v.begin(), v.end(), [g](const mh_match *a, const mh_match *b) {
            const int64_t as = left_end(*a, g), bs = left_end(*b, g);
            if (as == 0 || bs == 0) return bs != 0;
            return as < bs;
        });
        const int64_t n = (int64_t)v.size();
        int64_t i = 0;
        while (i < n && v[i]->start[g] == 0) ++i;
        for (; i < n; ++i) {
            if (!v[i]) continue;
            for (int64_t j = i + 1; j < n; ++j) {
                if (!v[j]) continue;
                mh_match &a = *v[i], &b = *v[j];
                int64_t over = left_end(b, g) - left_end(a, g) - (int64_t)a.length;
                if (over >= 0) break;
                over = -over;
                const int ma = (a.start[0] != 0) + (a.start[1] != 0), mb = (b.start[0] != 0) + (b.start[1] != 0);
                if (mb > ma || (mb == ma && b.length > a.length)) {
                    // the earlier match gives up its right end
                    if ((uint64_t)over >= a.length) {
                        v[i] = nullptr;
                        --i;
                        break;
                    }
                    if (a.start[g] >= 0) crop_end(a, (uint64_t)over);
                    else crop_start(a, (uint64_t)over);
                } else {
                    // the later match gives up its left end
                    if ((uint64_t)over >= b.length) {
                        v[j] = nullptr;
                        continue;
                    }
                    if (b.start[g] >= 0) crop_start(b, (uint64_t)over);
                    else crop_end(b, (uint64_t)over);
                }
            }
        }
        v.erase(std::remove(v.begin(), v.end(), nullptr), v.end());
    }
    std::vector<mh_match> out;
    out.reserve(v.size());
    for (mh_match *m : v) out.push_back(*m);
    ml.swap(out);
}

// One gap: its pieces' sorted mer lists and pairwise search on c, then the
// clean-up and the shift to genome coordinates, into out.
int search_gap(mh_context *c, const uint8_t *const seq[2], const uint64_t len[2], const mh_anchor_gap &gp,
               std::vector<mh_match> &out)
{
    out.clear();
    uint64_t off[2] = {0, 0}, plen[2] = {0, 0};
    for (int g = 0; g < 2; ++g) {
        int64_t diff = gp.rend[g] - gp.lend[g];
        diff = diff > 0 ? diff - 1 : 0;
        if (diff > 0 && gp.lend[g] >= 1 && (uint64_t)(gp.lend[g] + diff - 1) <= len[g]) {
            off[g] = (uint64_t)gp.lend[g] - 1;
            plen[g] = (uint64_t)diff;
        }
    }
    const uint32_t w = mh_default_seed_weight((plen[0] + plen[1]) / 2);
    if ((int)w < MIN_DNA_SEED_WEIGHT) return MH_OK;
    const uint64_t seed = (uint64_t)get_seed((int)w, 0);
    // a piece shorter than the seed has an empty sorted mer list: no seeds pair
    if (std::min(plen[0], plen[1]) < (uint64_t)seed_length((int64_t)seed)) return MH_OK;
    mh_sml *sm[2] = {nullptr, nullptr};
    int rc = MH_OK;
    for (int g = 0; g < 2 && rc == MH_OK; ++g) rc = mh_sml_create_device(c, seq[g] + off[g], plen[g], seed, &sm[g]);
    std::vector<mh_match> found;
    if (rc == MH_OK) rc = find_matches_pairwise(c, sm, 2, found, nullptr);
    mh_sml_free(sm[0]);
    mh_sml_free(sm[1]);
    if (rc != MH_OK) return rc;
    eliminate_overlaps(found);
    for (mh_match m : found) {
        if (m.length < ANCHOR_MIN_LENGTH) continue;
        for (int g = 0; g < 2; ++g) {
            if (m.start[g] == 0) continue;   // LeftEnd NO_MATCH: the reference's shift of 0 is not reached here
            const int64_t le = left_end(m, g) + gp.lend[g] - 1;
            m.start[g] = m.start[g] < 0 ? -le : le;
        }
        out.push_back(m);
    }
    return MH_OK;
}

}  // namespace

extern "C" {

// The gaps are independent searches: up to MAX_HELPERS of them run at once,
// each on a helper context of its own (a stream and scratch of its own), so
// that one search's host steps overlap another's kernels.  The results are
// kept per gap and returned in gap order.
int mh_anchor_gaps(mh_context *ctx, const void *d_seq_a, uint64_t len_a, const void *d_seq_b, uint64_t len_b,
                   const mh_anchor_gap *gaps, uint64_t n_gaps, mh_match_list **out, uint64_t *first)
{
    MH_CHECK(ctx && out && (n_gaps == 0 || (gaps && d_seq_a && d_seq_b && first)), MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    const uint8_t *seq[2] = {static_cast<const uint8_t *>(d_seq_a), static_cast<const uint8_t *>(d_seq_b)};
    const uint64_t len[2] = {len_a, len_b};
    std::vector<std::vector<mh_match>> per(n_gaps);
    constexpr uint64_t MAX_HELPERS = 4;
    const uint64_t nh = std::min<uint64_t>(MAX_HELPERS, n_gaps);
    if (nh <= 1) {
        for (uint64_t gi = 0; gi < n_gaps; ++gi) {
            const int rc = search_gap(ctx, seq, len, gaps[gi], per[gi]);
            if (rc != MH_OK) return rc;
        }
    } else {
        // the sequences may have been written on the caller's stream
        MH_HIP(hipStreamSynchronize(ctx->stream));
        while (ctx->helpers.size() < nh) {
            mh_context *h = nullptr;
            const int rc = mh_create(ctx->device, &h);
            if (rc != MH_OK) return rc;
            ctx->helpers.push_back(h);
        }
        std::vector<int> rcs(nh, MH_OK);
        std::vector<std::string> msgs(nh);
        host_parallel(nh, [&](size_t t) {
            mh_context *h = ctx->helpers[t];
            h->timing = ctx->timing;
            if (hipSetDevice(ctx->device) != hipSuccess) {
                rcs[t] = MH_ERR_DEVICE;
                msgs[t] = "hipSetDevice failed";
                return;
            }
            for (uint64_t gi = t; gi < n_gaps; gi += nh) {
                const int rc = search_gap(h, seq, len, gaps[gi], per[gi]);
                if (rc != MH_OK) {
                    rcs[t] = rc;
                    msgs[t] = mh_last_error();
                    return;
                }
            }
        });
        // the helpers' kernel and host timers count as this context's
        for (mh_context *h : ctx->helpers) {
            collect_timers(h);
            for (auto &kv : h->timers) {
                KernelTimer &k = ctx->timers[kv.first];
                k.total_ms += kv.second.total_ms;
                k.launches += kv.second.launches;
            }
            h->timers.clear();
        }
        for (uint64_t t = 0; t < nh; ++t)
            if (rcs[t] != MH_OK) {
                set_error(msgs[t]);
                return rcs[t];
            }
    }
    mh_match_list *l = new mh_match_list();
    l->seq_count = 2;
    l->files.assign(2, std::string());
    l->lengths = {len_a, len_b};
    for (uint64_t gi = 0; gi < n_gaps; ++gi) {
        first[gi] = l->v.size();
        l->v.insert(l->v.end(), per[gi].begin(), per[gi].end());
    }
    if (first) first[n_gaps] = l->v.size();
    *out = l;
    return MH_OK;
}

}  // extern "C"
