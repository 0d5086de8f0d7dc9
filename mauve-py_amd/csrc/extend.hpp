// extend.hpp -- ungapped seed extension of two genomes, shared by the device
// kernels and host-side checks of them.
//
// ExtEnv describes the two sorted mer lists by position (GetDnaSeedMer
// values, LM/SortedMerList.h) and the sequence lengths.  extend2 is the
// reference's ExtendMatch (LM/MatchFinder.h:218-374) for one seed.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define MH_HD __host__ __device__
#define MH_PRAGMA_UNROLL _Pragma("unroll")
#else
#define MH_HD
#define MH_PRAGMA_UNROLL
#endif

namespace mh {

enum : int { EXT_OK = 0, EXT_ERR = 1, EXT_LONG = 2 };

struct ExtEnv {
    const uint64_t *kp0, *kp1;   // GetDnaSeedMer by position
    uint64_t n0, n1;              // SMLLength
    int64_t L0, L1;               // sequence lengths
    int seed_len;
    uint64_t mer_mask;            // GetSeedMask(): top 2*weight bits
};

MH_HD inline bool lookup(const ExtEnv &e, int g, int64_t pos1, int64_t len, int64_t start, uint64_t &mer, int &parity)
{
    int64_t p = pos1;
    if (p < 0) p = -p + len - e.seed_len;
    const int64_t idx = p - 1;
    const uint64_t n = g ? e.n1 : e.n0;
    if (idx < 0 || (uint64_t)idx >= n) return false;
    const uint64_t k = (g ? e.kp1 : e.kp0)[idx];
    parity = start < 0 ? (int)(k & 1ull) : (int)!(k & 1ull);
    mer = k & e.mer_mask;
    return true;
}

// ExtendMatch, LM/MatchFinder.h:218-374, for a two-genome seed (both starts
// present, linear sequences).  EXT_ERR if it would read outside the lists
// (the reference reads out of bounds there; never seen in practice) or runs
// more than 8 (L0 + L1) steps; EXT_LONG once it has run more than `budget`
// steps (0: no budget), leaving st and len mid-walk.
MH_HD inline int extend2(const ExtEnv &e, int64_t st[2], int64_t &len, int64_t budget)
{
    const int64_t Lg[2] = {e.L0, e.L1};
    int jump = e.seed_len;
    uint32_t extend_limit = 0, extend_attempts = 0;
    bool extend_again = false;
    int64_t guard = 0;
    const int64_t guard_max = 8 * (e.L0 + e.L1) + 64;
    for (int dir = 0; dir < 4; ++dir) {
        int64_t maxlen = dir >= 2 ? (int64_t)e.seed_len : (int64_t)4294967295ll;
        for (int x = 0; x < 2; ++x) {
            if (st[x] < 0) {
                const int64_t rc_len = Lg[x] - len + st[x] + 1;
                if (rc_len < maxlen) maxlen = rc_len;
            } else if (st[x] - 1 < maxlen) {
                maxlen = st[x] - 1;
            }
        }
        int j = 0, i = 2;
        extend_limit = 0;
        extend_attempts = 0;
        if (dir < 2 && maxlen >= jump) {
            // The seed-length jumps of phases 0 and 1: the addresses of the
            // K = maxlen / jump lookups do not depend on their outcomes, so
            // they are issued in groups of PREFETCH and checked in order; the
            // state after the first failing step (or after the K-th) is what
            // the step-by-step loop leaves (len and the forward starts moved by
            // the successful steps only).
            constexpr int PREFETCH = 8;
            const int64_t K = maxlen / jump;
            int64_t ok_steps = K;
            bool done = false;
            for (int64_t k0 = 1; k0 <= K && !done; k0 += PREFETCH) {
                uint64_t m0[PREFETCH], m1[PREFETCH];
                int q0[PREFETCH], q1[PREFETCH];
                bool in[PREFETCH];
                const int cnt = (int)((K - k0 + 1) < PREFETCH ? (K - k0 + 1) : PREFETCH);
MH_PRAGMA_UNROLL
                for (int u = 0; u < PREFETCH; ++u) {
                    if (u >= cnt) break;
                    const int64_t k = k0 + u;
                    const int64_t ln = len + k * jump;
                    const int64_t p0 = st[0] > 0 ? st[0] - k * jump : st[0];
                    const int64_t p1 = st[1] > 0 ? st[1] - k * jump : st[1];
                    in[u] = lookup(e, 0, p0, ln, st[0], m0[u], q0[u]) && lookup(e, 1, p1, ln, st[1], m1[u], q1[u]);
                }
                for (int u = 0; u < cnt; ++u) {
                    if (++guard > guard_max) return EXT_ERR;
                    if (budget && guard > budget) return EXT_LONG;
                    if (!in[u]) return EXT_ERR;
                    if (m0[u] != m1[u] || q0[u] != q1[u]) {
                        ok_steps = k0 + u - 1;
                        done = true;
                        break;
                    }
                }
            }
            len += ok_steps * jump;
            for (int x = 0; x < 2; ++x)
                if (st[x] > 0) st[x] -= ok_steps * jump;
            maxlen = 0;   // the loop below does not run
        }
        while (maxlen - jump >= 0) {
            if (++guard > guard_max) return EXT_ERR;
            if (budget && guard > budget) return EXT_LONG;
            len += jump;
            maxlen -= jump;
            for (j = 0; j < 2; ++j)
                if (st[j] > 0) {
                    st[j] -= jump;
                    if (st[j] <= 0) st[j] += Lg[j];
                }
            uint64_t cm, om;
            int cp, op;
            if (!lookup(e, 0, st[0], len, st[0], cm, cp)) return EXT_ERR;
            if (!lookup(e, 1, st[1], len, st[1], om, op)) return EXT_ERR;
            i = 1;
            if (cm != om || cp != op) {
                if (dir < 2) maxlen = 0;
            } else {
                i = 2;
            }
            extend_attempts += jump;
            if (i == 2) extend_limit = extend_attempts;
            if (dir > 1 && extend_attempts == (uint32_t)e.seed_len) break;
        }
        if (i < 2) {
            len -= jump;
            for (; j > 0; j--)
                if (st[j - 1] >= 0) st[j - 1] += jump;
        }
        if (dir > 1 && extend_attempts > 0) {
            if (extend_limit > 0) extend_again = true;
            int unmatched = (int)(extend_attempts - extend_limit);
            if (i < 2) unmatched -= jump;
            len -= unmatched;
            for (j = 0; j < 2; ++j)
                if (st[j] > 0) {
                    st[j] += unmatched;
                    if (st[j] > Lg[j]) st[j] -= Lg[j];
                }
        }
        st[0] = -st[0];
        st[1] = -st[1];
        if (dir >= 1) jump = 1;
        if (dir == 3 && extend_again) {
            dir = -1;
            jump = e.seed_len;
            extend_again = false;
        }
    }
    return EXT_OK;
}

// The k-genome form of the same walk (the genomes a seed holds; start 0
// absent): ExtendMatch compares every present genome's window with the first
// present genome's, and a step matches when all of them do
// (LM/MatchFinder.h:218-374).  EnvK holds each genome's GetDnaSeedMer by
// position, list length and sequence length.
constexpr int EXT_MAXK = 16;

struct ExtEnvK {
    const uint64_t *kp[EXT_MAXK];
    uint64_t n[EXT_MAXK];
    int64_t L[EXT_MAXK];
    int seed_len;
    uint64_t mer_mask;
};

MH_HD inline bool lookup_k(const ExtEnvK &e, int g, int64_t pos1, int64_t len, uint64_t &mer, int &parity)
{
    int64_t p = pos1;
    if (p < 0) p = -p + len - e.seed_len;
    const int64_t idx = p - 1;
    if (idx < 0 || (uint64_t)idx >= e.n[g]) return false;
    const uint64_t k = e.kp[g][idx];
    parity = pos1 < 0 ? (int)(k & 1ull) : (int)!(k & 1ull);
    mer = k & e.mer_mask;
    return true;
}

// st[x] is the start in genome gs[x], x < used (the present genomes in genome
// order).  EXT_ERR if a lookup would leave a list or the walk runs more than
// 8 (sum of the lengths) steps.
MH_HD inline int extend_k(const ExtEnvK &e, const int *gs, int used, int64_t *st, int64_t &len)
{
    int jump = e.seed_len;
    uint32_t extend_limit = 0, extend_attempts = 0;
    bool extend_again = false;
    int64_t guard = 0, guard_max = 64;
    for (int x = 0; x < used; ++x) guard_max += 8 * e.L[gs[x]];
    for (int dir = 0; dir < 4; ++dir) {
        int64_t maxlen = dir >= 2 ? (int64_t)e.seed_len : (int64_t)4294967295ll;
        for (int x = 0; x < used; ++x) {
            if (st[x] < 0) {
                const int64_t rc_len = e.L[gs[x]] - len + st[x] + 1;
                if (rc_len < maxlen) maxlen = rc_len;
            } else if (st[x] - 1 < maxlen) {
                maxlen = st[x] - 1;
            }
        }
        int j = 0, i = used;
        extend_limit = 0;
        extend_attempts = 0;
        while (maxlen - jump >= 0) {
            if (++guard > guard_max) return EXT_ERR;
            len += jump;
            maxlen -= jump;
            for (j = 0; j < used; ++j)
                if (st[j] > 0) {
                    st[j] -= jump;
                    if (st[j] <= 0) st[j] += e.L[gs[j]];
                }
            uint64_t cm;
            int cp;
            if (!lookup_k(e, gs[0], st[0], len, cm, cp)) return EXT_ERR;
            for (i = 1; i < used; ++i) {
                uint64_t om;
                int op;
                if (!lookup_k(e, gs[i], st[i], len, om, op)) return EXT_ERR;
                if (cm != om || cp != op) {
                    if (dir < 2) maxlen = 0;
                    break;
                }
            }
            extend_attempts += jump;
            if (i == used) extend_limit = extend_attempts;
            if (dir > 1 && extend_attempts == (uint32_t)e.seed_len) break;
        }
        if (i < used) {
            len -= jump;
            for (; j > 0; j--)
                if (st[j - 1] >= 0) st[j - 1] += jump;
        }
        if (dir > 1 && extend_attempts > 0) {
            if (extend_limit > 0) extend_again = true;
            int unmatched = (int)(extend_attempts - extend_limit);
            if (i < used) unmatched -= jump;
            len -= unmatched;
            for (j = 0; j < used; ++j)
                if (st[j] > 0) {
                    st[j] += unmatched;
                    if (st[j] > e.L[gs[j]]) st[j] -= e.L[gs[j]];
                }
        }
        for (int x = 0; x < used; ++x) st[x] = -st[x];
        if (dir >= 1) jump = 1;
        if (dir == 3 && extend_again) {
            dir = -1;
            jump = e.seed_len;
            extend_again = false;
        }
    }
    return EXT_OK;
}

#if defined(__HIPCC__)
// extend_k for a seed of at most KMAX present genomes, with the lookups of a
// step issued together.  The caller passes the present genomes' lists,
// list sizes and sequence lengths in order (gk, gn, gL: slot x is genome
// gs[x]; slots from `used` on repeat slot 0 and are never compared), and st
// with KMAX slots.  A lookup's index is clamped into its list and the miss
// flagged instead of branching around the load, so a step's loads do not
// wait on each other; a miss is an error only where the literal loop, which
// stops at the first genome that differs, would have made that lookup.
// Same walk, result and errors as extend_k.
template <int KMAX>
__device__ inline int extend_kt(const uint64_t *const *gk, const uint32_t *gn, const int64_t *gL, int used,
                                int seed_len, uint64_t mer_mask, int64_t *st, int64_t &len)
{
    int jump = seed_len;
    uint32_t extend_limit = 0, extend_attempts = 0;
    bool extend_again = false;
    int64_t guard = 0, guard_max = 64;
    MH_PRAGMA_UNROLL
    for (int x = 0; x < KMAX; ++x)
        if (x < used) guard_max += 8 * gL[x];
    for (int dir = 0; dir < 4; ++dir) {
        int64_t maxlen = dir >= 2 ? (int64_t)seed_len : (int64_t)4294967295ll;
        MH_PRAGMA_UNROLL
        for (int x = 0; x < KMAX; ++x) {
            if (x >= used) continue;
            if (st[x] < 0) {
                const int64_t rc_len = gL[x] - len + st[x] + 1;
                if (rc_len < maxlen) maxlen = rc_len;
            } else if (st[x] - 1 < maxlen) {
                maxlen = st[x] - 1;
            }
        }
        int i = used;
        extend_limit = 0;
        extend_attempts = 0;
        while (maxlen - jump >= 0) {
            if (++guard > guard_max) return EXT_ERR;
            len += jump;
            maxlen -= jump;
            MH_PRAGMA_UNROLL
            for (int x = 0; x < KMAX; ++x)
                if (x < used && st[x] > 0) {
                    st[x] -= jump;
                    if (st[x] <= 0) st[x] += gL[x];
                }
            // every slot's window (lookup_k), loads issued together
            uint64_t kv[KMAX];
            bool in[KMAX];
            MH_PRAGMA_UNROLL
            for (int x = 0; x < KMAX; ++x) {
                const int64_t p = st[x] < 0 ? -st[x] + len - seed_len : st[x];
                const int64_t idx = p - 1;
                in[x] = x >= used || (idx >= 0 && (uint64_t)idx < (uint64_t)gn[x]);
                kv[x] = gk[x][(x < used && in[x]) ? idx : 0];
            }
            if (!in[0]) return EXT_ERR;
            const uint64_t cm = kv[0] & mer_mask;
            const int cp = st[0] < 0 ? (int)(kv[0] & 1ull) : (int)!(kv[0] & 1ull);
            i = used;
            MH_PRAGMA_UNROLL
            for (int x = 1; x < KMAX; ++x)
                if (x < used && i == used) {
                    if (!in[x]) return EXT_ERR;
                    const int op = st[x] < 0 ? (int)(kv[x] & 1ull) : (int)!(kv[x] & 1ull);
                    if (cm != (kv[x] & mer_mask) || cp != op) i = x;
                }
            if (i < used && dir < 2) maxlen = 0;
            extend_attempts += jump;
            if (i == used) extend_limit = extend_attempts;
            if (dir > 1 && extend_attempts == (uint32_t)seed_len) break;
        }
        if (i < used) {
            len -= jump;
            MH_PRAGMA_UNROLL
            for (int x = 0; x < KMAX; ++x)
                if (x < used && st[x] >= 0) st[x] += jump;
        }
        if (dir > 1 && extend_attempts > 0) {
            if (extend_limit > 0) extend_again = true;
            int unmatched = (int)(extend_attempts - extend_limit);
            if (i < used) unmatched -= jump;
            len -= unmatched;
            MH_PRAGMA_UNROLL
            for (int x = 0; x < KMAX; ++x)
                if (x < used && st[x] > 0) {
                    st[x] += unmatched;
                    if (st[x] > gL[x]) st[x] -= gL[x];
                }
        }
        MH_PRAGMA_UNROLL
        for (int x = 0; x < KMAX; ++x)
            if (x < used) st[x] = -st[x];
        if (dir >= 1) jump = 1;
        if (dir == 3 && extend_again) {
            dir = -1;
            jump = seed_len;
            extend_again = false;
        }
    }
    return EXT_OK;
}
#endif

}  // namespace mh
