/*
 * mh_oracle.c -- CPU oracle (test infrastructure only; see mh_oracle.h).
 *
 * Plain-C restatement of the reference's seed-anchoring and NWSmall paths.
 * Reference paths below are relative to mauve/src/ (LM = libMems/libMems,
 * MU = muscle/libMUSCLE, LG = libGenome/libGenome).
 */
#include "mh_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>

#define GNSEQI_END 4294967295ULL            /* LG/gnDefs.h:169 (UINT32_MAX) */
#define MER_REPEAT_LIMIT 1000               /* LM/MatchFinder.cpp:166        */
#define MER_BUFFER_SIZE 10000               /* LM/MatchFinder.cpp:174        */
#define MEM_TABLE_SIZE 40000                /* LM/MemHash.h:30               */

/* ======================================================================
 * Seed patterns -- LM/SeedMasks.h:44-401
 * ====================================================================== */

/* Rank-ordered spaced-seed masks per weight (low 32-bit word; the high word
 * is 0 for every entry in the reference table), LM/SeedMasks.h:44-257. */
static const uint32_t seed_table[32][6] = {
    {0}, {0}, {0},
    {0xb, 0, 0, 0, 0, 0},
    {0x3b, 0, 0, 0, 0, 0},
    {0x6b, 0x139, 0x193, 0x6b, 0, 0},
    {0x58D, 0x653, 0x1AB, 0xdb, 0, 0},
    {0x1953, 0x588d, 0x688b, 0x17d, 0x164d, 0},
    {0x3927, 0x1CA7, 0x6553, 0xb6d, 0, 0},
    {0x7497, 0x1c927, 0x72a7, 0x6fb, 0x16ed, 0},
    {0x1d297, 0x3A497, 0xE997, 0x6D5B, 0, 0},
    {0x7954f, 0x75257, 0x1c9527, 0x5bed, 0x5b26d, 0},
    {0x7954f, 0x3D32F, 0x768B7, 0x5B56D, 0, 0},
    {0x792a4f, 0x1d64d7, 0x1d3597, 0x1b7db, 0x75ad7, 0},
    {0x1e6acf, 0xF59AF, 0x3D4CAF, 0x35AD6B, 0, 0},
    {0x7ac9af, 0x7b2a6f, 0x79aacf, 0x16df6d, 0x6b5d6b, 0},
    {0xf599af, 0xEE5A77, 0x7CD59F, 0xEB5AD7, 0, 0},
    {0x6dbedb, 0, 0, 0, 0, 0},
    {0x3E6B59F, 0x3EB335F, 0x7B3566F, 0, 0, 0},
    {0x7b974ef, 0x7d6735f, 0x1edd74f, 0, 0, 0},
    {0x1F59B35F, 0x3EDCEDF, 0xFAE675F, 0, 0, 0},
    {0x7ddaddf, 0xaeb3f, 0x7eb76bf, 0, 0, 0},
    {0x003fffff, 0, 0, 0, 0, 0},
    {0x007fffff, 0, 0, 0, 0, 0},
    {0x00ffffff, 0, 0, 0, 0, 0},
    {0x01ffffff, 0, 0, 0, 0, 0},
    {0x03ffffff, 0, 0, 0, 0, 0},
    {0x07ffffff, 0, 0, 0, 0, 0},
    {0x0fffffff, 0, 0, 0, 0, 0},
    {0x1fffffff, 0, 0, 0, 0, 0},
    {0x3fffffff, 0, 0, 0, 0, 0},
    {0x7fffffff, 0, 0, 0, 0, 0},
};

/* LM/SeedMasks.h:276-281 */
static int64_t solid_seed(int weight) { return (int64_t)(((uint64_t)1 << weight) - 1); }

/* LM/SeedMasks.h:298-321 (SOLID_SEED rank = INT_MAX handled by rank > 5) */
int64_t or_get_seed(int weight, int seed_rank)
{
    if (weight > 31) return solid_seed(32);
    if (weight < 0) return 0;
    if (seed_rank > 5) return solid_seed(weight);
    if (seed_table[weight][seed_rank] == 0) return solid_seed(weight);
    return (int64_t)seed_table[weight][seed_rank];
}

/* LM/SeedMasks.h:335-350 */
int or_get_seed_length(int64_t seed)
{
    int right = -1, left = -1;
    uint64_t s = (uint64_t)seed;
    for (int b = 0; b < 64; ++b) {
        if (s & 1) { left = b; if (right == -1) right = b; }
        s >>= 1;
    }
    return left != -1 ? left - right + 1 : 0;
}

/* LM/SeedMasks.h:363-373 */
int or_get_seed_weight(int64_t seed)
{
    int w = 0;
    uint64_t s = (uint64_t)seed;
    for (int b = 0; b < 64; ++b) { w += (int)(s & 1); s >>= 1; }
    return w;
}

/* LM/SeedMasks.h:389-401 */
uint32_t or_default_seed_weight(uint64_t avg)
{
    uint32_t m = (uint32_t)ceil((log((double)avg) / log(2.0)) / 1.5);
    if (!(m & 1)) ++m;
    m = m < 5 ? 0 : m;
    if (avg == 0) m = 0;
    return m > 31 ? 31 : m;
}

/* ======================================================================
 * 2-bit packing and keys -- LM/SortedMerList.cpp
 * ====================================================================== */

/* BasicDNATable, LM/SortedMerList.cpp:29-47: c/C/b/B/y/Y=1, g/G/s/S/k/K=2,
 * t/T=3, every other byte 0.  (The reference indexes with a signed char; bytes
 * >= 128 are outside the table there and map to 0 here.) */
static uint8_t dna_code(unsigned char c)
{
    switch (c) {
    case 'c': case 'C': case 'b': case 'B': case 'y': case 'Y': return 1;
    case 'g': case 'G': case 's': case 'S': case 'k': case 'K': return 2;
    case 't': case 'T': return 3;
    default: return 0;
    }
}

/* SetSequence word count, LM/SortedMerList.cpp:306-311 */
uint64_t or_packed_words(uint64_t len)
{
    uint64_t w = (len * 2) / 32;
    if ((len * 2) % 32) ++w;
    return w + 2;
}

/* translate32, LM/SortedMerList.cpp:425-460: 16 chars per uint32, first char
 * in the two most significant bits.  '-' is an error (433-437).  dest must
 * hold or_packed_words(len) words; words past the data are zeroed here (the
 * reference leaves them uninitialised; no key depends on them). */
int or_translate32(const char *src, uint64_t len, uint32_t *dest)
{
    uint64_t nw = or_packed_words(len);
    memset(dest, 0, nw * sizeof(uint32_t));
    for (uint64_t i = 0; i < len; ++i) {
        if (src[i] == '-') return -1;
        uint32_t t = dna_code((unsigned char)src[i]);
        dest[i >> 4] |= t << (30 - 2 * (i & 15));
    }
    return 0;
}

/* GetMer, LM/SortedMerList.cpp:321-342 */
uint64_t or_get_mer(const uint32_t *seq, uint64_t pos, uint64_t mer_mask)
{
    uint64_t w = (pos * 2) / 32, bit = (pos * 2) % 32;
    uint64_t m = ((uint64_t)seq[w] << 32) | seq[w + 1];
    if (bit > 0) {
        uint32_t nx = seq[w + 2] >> (32 - bit);
        m = (m << bit) | nx;
    }
    return m & mer_mask;
}

/* RevCompMer, LM/SortedMerList.cpp:597-614: complement, reverse the 2-bit
 * groups, left-align to mer_length characters, set the LSB strand flag. */
uint64_t or_revcomp_mer(uint64_t mer_a, int mer_length)
{
    uint64_t b = ~mer_a, c = 0;
    for (int i = 0; i < 64; i += 2) {
        c |= b & 3;
        b >>= 2;
        c <<= 2;
    }
    c <<= 64 - (2 * (mer_length + 1));
    return c | 1;
}

static uint64_t mask_for(int chars)
{
    return chars >= 32 ? ~0ULL : (~0ULL << (64 - 2 * chars));
}

/* GetSeedMer, LM/SortedMerList.cpp:726-762 (seed length <= 31, so the
 * mer_transition branch at 744-748 is never taken). */
uint64_t or_get_seed_mer(const uint32_t *seq, uint64_t pos, uint64_t seed,
                         int seed_length, int seed_weight)
{
    uint64_t mer = or_get_mer(seq, pos, mask_for(seed_length));
    uint64_t out = 0;
    uint64_t amask = 3ULL << 62;
    uint64_t cmask = 1ULL << (seed_length - 1);
    int rsh = 62;
    for (int p = 0; p < seed_length; ++p) {
        if (seed & cmask) {
            out = (out << 2) | ((mer & amask) >> rsh);
        }
        amask >>= 2;
        cmask >>= 1;
        rsh -= 2;
    }
    return out << (64 - 2 * seed_weight);
}

/* GetDnaSeedMer, LM/SortedMerList.cpp:764-769 */
uint64_t or_get_dna_seed_mer(const uint32_t *seq, uint64_t pos, uint64_t seed,
                             int seed_length, int seed_weight)
{
    uint64_t f = or_get_seed_mer(seq, pos, seed, seed_length, seed_weight);
    uint64_t r = or_revcomp_mer(f, seed_weight);
    return f < r ? f : r;
}

/* FillDnaSML, LM/SortedMerList.cpp:617-723 (incremental solid-seed keys;
 * linear sequences only).  keys_out gets len-seed_length+1 keys in position
 * order.  Restated step by step, including its word-streaming bookkeeping. */
void or_fill_dna_sml(const uint32_t *sequence, uint64_t len, int mer_size,
                     uint64_t *keys_out)
{
    if (len < (uint64_t)mer_size) return;
    uint64_t sar_len = len - (mer_size - 1);
    uint32_t dead_bits = 64 - mer_size * 2;
    uint64_t create_mask = ~0ULL << dead_bits;

    uint64_t cur = (((uint64_t)sequence[0] << 32) | sequence[1]) & create_mask;
    uint64_t rcur = 0, b = ~cur;
    for (int i = 0; i < 64; i += 2) { rcur |= b & 3; b >>= 2; rcur <<= 2; }
    rcur <<= dead_bits - 2;
    rcur |= 1;
    uint64_t n = 0;
    keys_out[n++] = cur < rcur ? cur : rcur;

    uint64_t endI = sar_len + mer_size;
    uint32_t rdead_bits = 64 - 2 - dead_bits;
    uint32_t seqI = (mer_size * 2) / 32;
    int32_t cur_bit = 32 - 2 - ((mer_size * 2) % 32);
    uint32_t cur_seq = sequence[seqI];
    uint64_t revalpha_mask = 3ULL << dead_bits;
    for (uint64_t p = mer_size + 1; p < endI; ++p) {
        uint64_t t = cur_seq;
        t >>= cur_bit;
        t &= 3;
        t <<= dead_bits;
        cur = (cur << 2) | t;
        t = ~t;
        t &= revalpha_mask;
        uint64_t tr = t << rdead_bits;
        rcur >>= 2;
        rcur |= tr;
        rcur &= create_mask;
        rcur |= 1;
        keys_out[n++] = cur < rcur ? cur : rcur;
        cur_bit -= 2;
        if (cur_bit < 0) {
            cur_bit += 2;
            cur_seq <<= 16;
            cur_seq <<= 16 - cur_bit;
            seqI++;
            uint64_t ns = sequence[seqI];
            ns >>= cur_bit;
            cur_seq |= (uint32_t)ns;
            cur_bit += 32 - 2;
        }
    }
}

/* FillDnaSeedSML, LM/SortedMerList.cpp:771-783 */
void or_fill_dna_seed_sml(const uint32_t *seq, uint64_t n, uint64_t seed,
                          int seed_length, int seed_weight, uint64_t *keys_out)
{
    for (uint64_t p = 0; p < n; ++p)
        keys_out[p] = or_get_dna_seed_mer(seq, p, seed, seed_length, seed_weight);
}

typedef struct { uint64_t key; uint32_t pos; } kp_t;

static int kp_cmp(const void *a, const void *b)
{
    const kp_t *x = (const kp_t *)a, *y = (const kp_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

/* SortedMerList::Create + Fill + sort, LM/FileSML.cpp:401-429 and
 * LM/MemorySML.cpp:45-60.  The reference's std::sort is not stable; ties
 * (equal 64-bit keys) are put in ascending position order here. */
int64_t or_sml_build(const char *ascii, uint64_t len, uint64_t seed,
                     uint32_t *packed_out, uint64_t *keys_out, uint32_t *pos_out)
{
    int sl = or_get_seed_length((int64_t)seed), sw = or_get_seed_weight((int64_t)seed);
    if (sl == 0 || sl > 31) return -1;          /* LM/SortedMerList.cpp:794-798, DNAFileSML max 31 */
    if (or_translate32(ascii, len, packed_out) != 0) return -1;
    if (len < (uint64_t)sl) return 0;
    uint64_t n = len - sl + 1;
    uint64_t *k = (uint64_t *)malloc(n * sizeof(uint64_t));
    if (sl != sw) or_fill_dna_seed_sml(packed_out, n, seed, sl, sw, k);
    else or_fill_dna_sml(packed_out, len, sl, k);
    kp_t *kp = (kp_t *)malloc(n * sizeof(kp_t));
    for (uint64_t i = 0; i < n; ++i) { kp[i].key = k[i]; kp[i].pos = (uint32_t)i; }
    qsort(kp, n, sizeof(kp_t), kp_cmp);
    for (uint64_t i = 0; i < n; ++i) { keys_out[i] = kp[i].key; pos_out[i] = kp[i].pos; }
    free(kp);
    free(k);
    return (int64_t)n;
}

/* ======================================================================
 * Seed matching
 * ====================================================================== */

typedef struct {
    uint64_t len;          /* sequence length                        */
    uint64_t n;            /* SMLLength                              */
    uint32_t *packed;
    uint64_t *keys;        /* sorted keys                            */
    uint32_t *pos;         /* sorted positions                       */
    uint64_t seed;
    int sl, sw;
    uint64_t seed_mask;    /* GetSeedMask(): top 2*weight bits       */
} sml_t;

static int sml_make(sml_t *s, const char *ascii, uint64_t len, uint64_t seed)
{
    memset(s, 0, sizeof(*s));
    s->len = len;
    s->seed = seed;
    s->sl = or_get_seed_length((int64_t)seed);
    s->sw = or_get_seed_weight((int64_t)seed);
    s->seed_mask = mask_for(s->sw);
    s->packed = (uint32_t *)malloc(or_packed_words(len) * sizeof(uint32_t));
    uint64_t n = len >= (uint64_t)s->sl ? len - s->sl + 1 : 0;
    s->keys = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    s->pos = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    int64_t r = or_sml_build(ascii, len, seed, s->packed, s->keys, s->pos);
    if (r < 0) return -1;
    s->n = (uint64_t)r;
    return 0;
}

static void sml_free(sml_t *s) { free(s->packed); free(s->keys); free(s->pos); }

/* DNAFileSML::GetSeedMer -> GetDnaSeedMer (LM/DNAFileSML.cpp:59-61) */
static uint64_t sml_seed_mer(const sml_t *s, uint64_t pos)
{
    return or_get_dna_seed_mer(s->packed, pos, s->seed, s->sl, s->sw);
}

/* SortedMerList::bsearch, LM/SortedMerList.cpp:380-394 (keys only; the key at
 * each sorted index does not depend on how ties were ordered). */
static uint64_t sml_bsearch(const sml_t *s, uint64_t q, uint64_t start, uint64_t end)
{
    for (;;) {
        uint64_t mid = (start + end) / 2;
        uint64_t mm = s->keys[mid];
        if (mm == q) return mid;
        if (mm < q && mid < end) { start = mid + 1; continue; }
        if (mm > q && start < mid) { end = mid - 1; continue; }
        return mid;
    }
}

/* SortedMerList::FindMer, LM/SortedMerList.cpp:170-179 (linear sequences).
 * Returns -1 for "false without touching result" (too short). */
static int sml_find_mer(const sml_t *s, uint64_t q, uint64_t *result)
{
    uint64_t last = s->len;
    if (last == 0 || last < (uint64_t)s->sl) return -1;
    last -= s->sl;
    *result = sml_bsearch(s, q, 0, last);
    return s->keys[*result] == q;
}

/* FindMer over a sorted key list of a sequence of `len` bases (seed length
 * sl): 1 found, 0 not found, -1 list too short (result untouched). */
int or_find_mer_keys(const uint64_t *keys, uint64_t len, int sl, uint64_t q, uint64_t *result)
{
    sml_t s;
    memset(&s, 0, sizeof(s));
    s.keys = (uint64_t *)keys;
    s.len = len;
    s.sl = sl;
    return sml_find_mer(&s, q, result);
}

/* ---- match entries (LM/MatchHashEntry.{h,cpp}, LM/HybridAbstractMatch.h) ---- */
typedef struct {
    int64_t  start[OR_MAX_SEQ];
    uint64_t length;
    int64_t  offset;
    uint64_t mersize;      /* candidates: seed length; stored copies: 0 (MatchHashEntry::operator=, LM/MatchHashEntry.cpp:118-126) */
} mhe_t;

static int mhe_first_start(const mhe_t *m, int k)
{
    for (int i = 0; i < k; ++i) if (m->start[i] != 0) return i;
    return 0x7fffffff;   /* NO_SEQ */
}

static uint64_t mhe_left_end(const mhe_t *m, int i)
{
    return (uint64_t)(m->start[i] < 0 ? -m->start[i] : m->start[i]);
}

/* CalculateOffset, LM/MatchHashEntry.cpp:141-160 */
static void mhe_calc_offset(mhe_t *m, int k)
{
    int f = mhe_first_start(m, k);
    int64_t ref = m->start[f];
    m->offset = 0;
    for (int i = f + 1; i < k; ++i) {
        if (m->start[i] != 0) {
            int64_t t = m->start[i] - ref;
            if (m->start[i] < 0) t -= (int64_t)m->length;
            m->offset += t;
        }
    }
}

/* Contains, LM/MatchHashEntry.cpp:164-200 (a contains b) */
static int mhe_contains(const mhe_t *a, const mhe_t *b, int k)
{
    if (a->offset != b->offset) return 0;
    int i = mhe_first_start(b, k);
    int64_t diff = b->start[i] - a->start[i];
    if (a->start[i] == 0) return 0;
    if (diff < 0 || a->length < b->length + (uint64_t)diff) return 0;
    int64_t diff_rc = (int64_t)b->length - (int64_t)a->length + diff;
    for (++i; i < k; ++i) {
        int64_t di = b->start[i] - a->start[i];
        if (b->start[i] == 0 && a->start[i] == 0) continue;
        else if (b->start[i] < 0 && diff_rc == di) continue;
        else if (diff != di) return 0;
    }
    return 1;
}

/* strict_start_lessthan_ptr, LM/MatchHashEntry.cpp:48-67 */
static int mhe_strict_less(const mhe_t *a, const mhe_t *b, int k)
{
    int sd = mhe_first_start(a, k) - mhe_first_start(b, k);
    if (sd == 0) {
        for (int i = 0; i < k; ++i) {
            int64_t as = a->start[i], bs = b->start[i];
            if (as < 0) as = (int64_t)((uint64_t)(-as) + a->length - a->mersize);
            if (bs < 0) bs = (int64_t)((uint64_t)(-bs) + b->length - b->mersize);
            int64_t d = as - bs;
            if (d == 0) continue;
            return d < 0;
        }
    }
    return sd < 0;
}

/* MheCompare, LM/MatchHashEntry.h:121-143 */
static int mhe_compare(const mhe_t *a, const mhe_t *b, int k)
{
    int fa = mhe_first_start(a, k), fb = mhe_first_start(b, k);
    if (fa > fb) return 1;
    if (fa == fb) {
        for (int i = fa; i < k; ++i) {
            int ae = mhe_left_end(a, i) != 0, be = mhe_left_end(b, i) != 0;
            if (!ae && be) return 1;
            if (ae && !be) return 0;
        }
        if (mhe_contains(a, b, k) || mhe_contains(b, a, k)) return 0;
        return mhe_strict_less(a, b, k);
    }
    return 0;
}

/* ---- ExtendMatch, LM/MatchFinder.h:218-374 (linear sequences; the
 * subset-match search at 359-364 is disabled in the reference) ---- */
static void extend_match(const sml_t *sm, int k, mhe_t *m)
{
    uint64_t mer_mask = sm[0].seed_mask;
    int seed_len = sm[0].sl;
    int cur_seqs[OR_MAX_SEQ], used = 0;
    for (int s = 0; s < k; ++s) if (m->start[s] != 0) cur_seqs[used++] = s;
    int jump = seed_len;
    uint32_t extend_limit = 0, extend_attempts = 0;
    int extend_again = 0;
    for (int dir = 0; dir < 4; ++dir) {
        int64_t maxlen = (int64_t)GNSEQI_END;
        if (dir >= 2) maxlen = seed_len;
        for (int x = 0; x < used; ++x) {
            int s = cur_seqs[x];
            if (m->start[s] < 0) {
                int64_t rc_len = (int64_t)sm[s].len - (int64_t)m->length + m->start[s] + 1;
                if (rc_len < maxlen) maxlen = rc_len;
            } else if (m->start[s] - 1 < maxlen) {
                maxlen = m->start[s] - 1;
            }
        }
        int j = 0, i = used;
        extend_limit = 0;
        extend_attempts = 0;
        while (maxlen - jump >= 0) {
            m->length += jump;
            maxlen -= jump;
            for (j = 0; j < used; ++j) {
                int s = cur_seqs[j];
                if (m->start[s] > 0) {
                    m->start[s] -= jump;
                    if (m->start[s] <= 0) m->start[s] += (int64_t)sm[s].len;
                }
            }
            int64_t g = m->start[cur_seqs[0]];
            if (g < 0) g = -g + (int64_t)m->length - seed_len;
            if (g - 1 < 0 || (uint64_t)(g - 1) >= sm[cur_seqs[0]].n) { fprintf(stderr, "oracle: extend out of range\n"); abort(); }
            uint64_t cm = sml_seed_mer(&sm[cur_seqs[0]], (uint64_t)(g - 1));
            int parity = m->start[cur_seqs[0]] < 0 ? (int)(cm & 1) : !(cm & 1);
            cm &= mer_mask;
            for (i = 1; i < used; ++i) {
                int s = cur_seqs[i];
                int64_t h = m->start[s];
                if (h < 0) h = -h + (int64_t)m->length - seed_len;
                if (h - 1 < 0 || (uint64_t)(h - 1) >= sm[s].n) { fprintf(stderr, "oracle: extend out of range\n"); abort(); }
                uint64_t om = sml_seed_mer(&sm[s], (uint64_t)(h - 1));
                int op = m->start[s] < 0 ? (int)(om & 1) : !(om & 1);
                om &= mer_mask;
                if (cm != om || parity != op) {
                    if (dir < 2) maxlen = 0;
                    break;
                }
            }
            extend_attempts += jump;
            if (i == used) extend_limit = extend_attempts;
            if (dir > 1 && extend_attempts == (uint32_t)seed_len) break;
        }
        if (i < used) {
            m->length -= jump;
            for (; j > 0; j--) {
                int s = cur_seqs[j - 1];
                if (m->start[s] >= 0) m->start[s] += jump;
            }
        }
        if (dir > 1 && extend_attempts > 0) {
            if (extend_limit > 0) extend_again = 1;
            int unmatched = (int)(extend_attempts - extend_limit);
            if (i < used) unmatched -= jump;
            m->length -= (int64_t)unmatched;
            for (j = 0; j < used; ++j) {
                int s = cur_seqs[j];
                if (m->start[s] > 0) {
                    m->start[s] += unmatched;
                    if (m->start[s] > (int64_t)sm[s].len) m->start[s] -= (int64_t)sm[s].len;
                }
            }
        }
        for (int s = 0; s < k; ++s) m->start[s] = -m->start[s];   /* Invert(), LM/HybridAbstractMatch.h:206-213 */
        if (dir >= 1) jump = 1;
        if (dir == 3 && extend_again) {
            dir = -1;
            jump = seed_len;
            extend_again = 0;
        }
    }
}

/* ---- MemHash table replay (LM/MemHash.cpp:167-251) ---- */
typedef struct {
    mhe_t  **v;
    size_t n, cap;
} bucket_t;

typedef struct {
    int k;
    const sml_t *sm;
    bucket_t *tab;
    mhe_t **pool;          /* every stored entry, for freeing */
    size_t npool, cappool;
    or_match_stats st;
    /* UniqueMatchFinder's EnumerateMatches instead of PairwiseMatchFinder's */
    int unique;
    /* seed-tuple capture mode: record HashMatch calls instead of inserting */
    int tup_mode;
    or_match *tup_out;
    int64_t *tup_off;
    int64_t tup_max, tup_n;
} memhash_t;

/* std::lower_bound as implemented by libstdc++ (bits/stl_algobase.h):
 * the probe sequence matters because MheCompare is not a strict weak order. */
static size_t lower_bound_mhe(bucket_t *b, const mhe_t *val, int k)
{
    size_t first = 0, len = b->n;
    while (len > 0) {
        size_t half = len >> 1, mid = first + half;
        if (mhe_compare(b->v[mid], val, k)) { first = mid + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}

/* AddHashEntry, LM/MemHash.cpp:209-251 */
static void add_hash_entry(memhash_t *h, mhe_t *m)
{
    int k = h->k;
    uint32_t bucket = (uint32_t)(((m->offset % (int64_t)MEM_TABLE_SIZE) + MEM_TABLE_SIZE) % MEM_TABLE_SIZE);
    bucket_t *b = &h->tab[bucket];
    size_t it = lower_bound_mhe(b, m, k);
    if (it != b->n && !mhe_compare(b->v[it], m, k) && !mhe_compare(m, b->v[it], k)) {
        h->st.collisions++;
        return;
    }
    extend_match(h->sm, k, m);
    mhe_t *nm = (mhe_t *)malloc(sizeof(mhe_t));
    *nm = *m;
    nm->mersize = 0;
    if (h->npool == h->cappool) {
        h->cappool = h->cappool ? h->cappool * 2 : 1024;
        h->pool = (mhe_t **)realloc(h->pool, h->cappool * sizeof(mhe_t *));
    }
    h->pool[h->npool++] = nm;
    it = lower_bound_mhe(b, nm, k);
    if (b->n == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 8;
        b->v = (mhe_t **)realloc(b->v, b->cap * sizeof(mhe_t *));
    }
    memmove(&b->v[it + 1], &b->v[it], (b->n - it) * sizeof(mhe_t *));
    b->v[it] = nm;
    b->n++;
}

typedef struct { uint64_t pos; uint64_t mer; int id; } idmer_t;

/* HashMatch + SetDirection, LM/MemHash.cpp:167-203 */
static void hash_match(memhash_t *h, const idmer_t *ent, int cnt)
{
    int k = h->k;
    mhe_t m;
    memset(&m, 0, sizeof(m));
    m.length = (uint64_t)h->sm[0].sl;
    m.mersize = (uint64_t)h->sm[0].sl;
    for (int x = 0; x < cnt; ++x) m.start[ent[x].id] = (int64_t)ent[x].pos + 1;
    int s = 0, ref_fwd = 0;
    for (; s < k; ++s)
        if (m.start[s] != 0) {
            ref_fwd = !(sml_seed_mer(&h->sm[s], (uint64_t)(m.start[s] - 1)) & 1);
            break;
        }
    for (++s; s < k; ++s)
        if (m.start[s] != 0)
            if (ref_fwd == (int)(sml_seed_mer(&h->sm[s], (uint64_t)(m.start[s] - 1)) & 1))
                m.start[s] = -m.start[s];
    mhe_calc_offset(&m, k);
    h->st.hash_calls++;
    if (h->tup_mode) {
        if (h->tup_out && h->tup_n < h->tup_max) {
            for (int i = 0; i < OR_MAX_SEQ; ++i) h->tup_out[h->tup_n].start[i] = i < k ? m.start[i] : 0;
            h->tup_out[h->tup_n].length = m.length;
            if (h->tup_off) h->tup_off[h->tup_n] = m.offset;
        }
        h->tup_n++;
        return;
    }
    add_hash_entry(h, &m);
}

static int idmer_id_cmp(const void *a, const void *b)
{
    const idmer_t *x = (const idmer_t *)a, *y = (const idmer_t *)b;
    return x->id - y->id;
}

/* PairwiseMatchFinder::EnumerateMatches, LM/PairwiseMatchFinder.cpp:37-71.
 * list::sort is stable; a stable merge of ids gives the same result as the
 * insertion sort below. */
static void enumerate_pairwise(memhash_t *h, idmer_t *run, size_t cnt)
{
    /* stable sort by id (insertion sort keeps equal ids in order) */
    for (size_t i = 1; i < cnt; ++i) {
        idmer_t t = run[i];
        size_t j = i;
        while (j > 0 && run[j - 1].id > t.id) { run[j] = run[j - 1]; --j; }
        run[j] = t;
    }
    (void)idmer_id_cmp;
    idmer_t uniq[OR_MAX_SEQ];
    int nu = 0;
    size_t i = 0;
    while (i < cnt) {
        size_t j = i;
        while (j < cnt && run[j].id == run[i].id) ++j;
        if (j - i == 1) uniq[nu++] = run[i];
        i = j;
    }
    for (int a = 0; a < nu; ++a)
        for (int b = a + 1; b < nu; ++b) {
            idmer_t pr[2] = {uniq[a], uniq[b]};
            hash_match(h, pr, 2);
        }
}

/* UniqueMatchFinder::EnumerateMatches, MA/UniqueMatchFinder.cpp:36-60: the
 * run sorted by genome (stable), every genome with exactly one entry kept,
 * one HashMatch of all of them when there are at least two. */
static void enumerate_unique(memhash_t *h, idmer_t *run, size_t cnt)
{
    for (size_t i = 1; i < cnt; ++i) {
        idmer_t t = run[i];
        size_t j = i;
        while (j > 0 && run[j - 1].id > t.id) { run[j] = run[j - 1]; --j; }
        run[j] = t;
    }
    idmer_t uniq[OR_MAX_SEQ];
    int nu = 0;
    size_t i = 0;
    while (i < cnt) {
        size_t j = i;
        while (j < cnt && run[j].id == run[i].id) ++j;
        if (j - i == 1) uniq[nu++] = run[i];
        i = j;
    }
    if (nu >= 2) hash_match(h, uniq, nu);
}

static void enumerate(memhash_t *h, idmer_t *run, size_t cnt)
{
    if (h->unique) enumerate_unique(h, run, cnt);
    else enumerate_pairwise(h, run, cnt);
}

/* Read(), LM/FileSML.cpp:470-500: entries [off, off+size) clipped to SMLLength */
typedef struct { uint64_t pos, mer; } bm_t;

static size_t sml_read(const sml_t *s, uint64_t size, uint64_t off, bm_t *buf)
{
    if (off >= s->n) return 0;
    uint64_t rl = off + size < s->n ? size : s->n - off;
    for (uint64_t j = 0; j < rl; ++j) { buf[j].pos = s->pos[off + j]; buf[j].mer = s->keys[off + j]; }
    return (size_t)rl;
}

/* GetBreakpoint, LM/MatchFinder.cpp:89-126, including its two quirks:
 * the backward scan in the other genomes compares against
 * (break_mer && mer_mask) and re-reads the same entry. */
static void get_breakpoint(const sml_t *sm, int k, int sarI, uint64_t startI, uint64_t *bp)
{
    uint64_t mask = sm[sarI].seed_mask;
    uint64_t brk = sm[sarI].keys[startI];
    uint64_t prev = brk;
    while ((prev & mask) == (brk & mask)) {
        if (startI == 0) { startI--; break; }
        startI--;
        prev = sm[sarI].keys[startI];
    }
    ++startI;
    uint64_t logical = (brk != 0 && mask != 0) ? 1 : 0;
    for (int i = 0; i < k; ++i) {
        if (i == sarI) { bp[i] = startI; continue; }
        uint64_t cs = 0;
        int f = sml_find_mer(&sm[i], brk, &cs);
        if (f < 0) { fprintf(stderr, "oracle: GetBreakpoint on a too-short SML is undefined in the reference\n"); abort(); }
        if (f) {
            int64_t ci = (int64_t)cs;
            uint64_t mm = sm[i].keys[cs];
            while (ci >= 0 && ((mm & mask) == logical)) ci--;
            cs = (uint64_t)(ci + 1);
        }
        bp[i] = cs;
    }
}

/* SearchRange, LM/MatchFinder.cpp:172-340, restated with its buffering and
 * repeat-limit restart.  cur_mers is the sorted idmer list; insertion puts a
 * new entry before the first element whose mer is >= it (329-331). */
static int search_range(memhash_t *h, uint64_t *start_points)
{
    int k = h->k;
    const sml_t *sm = h->sm;
    uint64_t mask = sm[0].seed_mask;
    for (int i = 0; i < k; ++i) if (sm[i].seed != sm[0].seed) return -1;

    bm_t *buf[OR_MAX_SEQ];
    size_t bufn[OR_MAX_SEQ];
    uint32_t mer_index[OR_MAX_SEQ], mer_base[OR_MAX_SEQ];
    idmer_t cur_mers[OR_MAX_SEQ];
    int ncur = 0;
    for (int n = 0; n < k; ++n) {
        buf[n] = (bm_t *)malloc(MER_BUFFER_SIZE * sizeof(bm_t));
        bufn[n] = sml_read(&sm[n], MER_BUFFER_SIZE, start_points[n], buf[n]);
        mer_index[n] = 0;
        mer_base[n] = 0;
        if (bufn[n] > 0) {
            idmer_t e = {buf[n][0].pos, buf[n][0].mer & mask, n};
            cur_mers[ncur++] = e;
        }
    }
    /* list::sort by mer (stable) */
    for (int i = 1; i < ncur; ++i) {
        idmer_t t = cur_mers[i];
        int j = i;
        while (j > 0 && cur_mers[j - 1].mer > t.mer) { cur_mers[j] = cur_mers[j - 1]; --j; }
        cur_mers[j] = t;
    }
    size_t mcap = 1024, mcnt = 0;
    idmer_t *match = (idmer_t *)malloc(mcap * sizeof(idmer_t));
    int ret = 1;
    while (ncur > 0) {
        idmer_t it = cur_mers[0];
        int cid = it.id;
        if (mcnt > 0) {
            if (it.mer > match[0].mer) {
                if (mcnt > 1) enumerate(h, match, mcnt);
                mcnt = 0;
            }
        }
        if (mcnt > MER_REPEAT_LIMIT) {
            uint64_t next_mer = match[0].mer + (~mask + 1);
            uint64_t next_pos = 0;
            int s = 0;
            for (; s < k; ++s) {
                int f = sml_find_mer(&sm[s], next_mer, &next_pos);
                if (f != 1) ++next_pos;
                if (next_pos < sm[s].n) break;
            }
            uint64_t old[OR_MAX_SEQ];
            memcpy(old, start_points, k * sizeof(uint64_t));
            if (s < k) get_breakpoint(sm, k, s, next_pos, start_points);
            for (int sp = 0; sp < k; ++sp) {
                uint64_t cur = (uint64_t)mer_index[sp] + mer_base[sp] + old[sp];
                if (start_points[sp] < cur) start_points[sp] = cur;
                if (sp < s) start_points[sp] = sm[sp].n;
            }
            h->st.restarts++;
            ret = 0;
            goto done;
        }
        uint32_t mi = mer_index[cid];
        int exhausted = mi < bufn[cid] ? 0 : 1;
        while (!exhausted && it.mer == (buf[cid][mi].mer & mask)) {
            if (mcnt == mcap) { mcap *= 2; match = (idmer_t *)realloc(match, mcap * sizeof(idmer_t)); }
            idmer_t e = {buf[cid][mi].pos, buf[cid][mi].mer & mask, cid};
            match[mcnt++] = e;
            ++mi;
            ++mer_index[cid];
            if (mi == bufn[cid]) exhausted = 1;
        }
        if (exhausted) {
            mer_base[cid] += (uint32_t)bufn[cid];
            uint32_t rs = MER_BUFFER_SIZE;
            if ((uint64_t)MER_BUFFER_SIZE + mer_base[cid] > GNSEQI_END) rs = (uint32_t)(GNSEQI_END - mer_base[cid]);
            bufn[cid] = sml_read(&sm[cid], rs, start_points[cid] + mer_base[cid], buf[cid]);
            mer_index[cid] = 0;
            if (bufn[cid] == 0) {
                memmove(&cur_mers[0], &cur_mers[1], (ncur - 1) * sizeof(idmer_t));
                --ncur;
            }
        } else {
            memmove(&cur_mers[0], &cur_mers[1], (ncur - 1) * sizeof(idmer_t));
            --ncur;
            idmer_t e = {buf[cid][mi].pos, buf[cid][mi].mer & mask, cid};
            int p = 0;
            while (p < ncur && cur_mers[p].mer < e.mer) ++p;
            memmove(&cur_mers[p + 1], &cur_mers[p], (ncur - p) * sizeof(idmer_t));
            cur_mers[p] = e;
            ++ncur;
        }
    }
    if (mcnt > 1) enumerate(h, match, mcnt);
done:
    free(match);
    for (int n = 0; n < k; ++n) free(buf[n]);
    return ret;
}

static int build_smls(int k, const char *const *seqs, const uint64_t *lens, uint64_t seed, sml_t *sm)
{
    if (k < 1 || k > OR_MAX_SEQ) return -1;
    int sl = or_get_seed_length((int64_t)seed);
    for (int i = 0; i < k; ++i) {
        if (lens[i] < (uint64_t)sl) return -1;
        if (sml_make(&sm[i], seqs[i], lens[i], seed) != 0) {
            for (int j = 0; j <= i; ++j) sml_free(&sm[j]);
            return -1;
        }
    }
    return 0;
}

static void run_search(memhash_t *h)
{
    uint64_t sp[OR_MAX_SEQ] = {0};
    while (search_range(h, sp) == 0) { }
}

static int64_t find_matches(int k, const char *const *seqs, const uint64_t *lens, uint64_t seed, or_match *out,
                            int64_t max_out, or_match_stats *stats, int unique)
{
    sml_t sm[OR_MAX_SEQ];
    if (build_smls(k, seqs, lens, seed, sm) != 0) return -1;
    memhash_t h;
    memset(&h, 0, sizeof(h));
    h.k = k;
    h.sm = sm;
    h.unique = unique;
    h.tab = (bucket_t *)calloc(MEM_TABLE_SIZE, sizeof(bucket_t));
    run_search(&h);
    /* GetMatchList, LM/MemHash.h:182-203: bucket order, then vector order */
    int64_t n = 0;
    for (uint32_t b = 0; b < MEM_TABLE_SIZE; ++b) {
        for (size_t i = 0; i < h.tab[b].n; ++i) {
            if (n < max_out) {
                const mhe_t *m = h.tab[b].v[i];
                for (int s = 0; s < OR_MAX_SEQ; ++s) out[n].start[s] = s < k ? m->start[s] : 0;
                out[n].length = m->length;
            }
            ++n;
        }
        free(h.tab[b].v);
    }
    h.st.matches = n;
    if (stats) *stats = h.st;
    for (size_t i = 0; i < h.npool; ++i) free(h.pool[i]);
    free(h.pool);
    free(h.tab);
    for (int i = 0; i < k; ++i) sml_free(&sm[i]);
    return n < max_out ? n : max_out;
}

int64_t or_find_matches_pairwise(int k, const char *const *seqs, const uint64_t *lens,
                                 uint64_t seed, or_match *out, int64_t max_out,
                                 or_match_stats *stats)
{
    return find_matches(k, seqs, lens, seed, out, max_out, stats, 0);
}

int64_t or_find_matches_unique(int k, const char *const *seqs, const uint64_t *lens,
                               uint64_t seed, or_match *out, int64_t max_out,
                               or_match_stats *stats)
{
    if (k < 2) return -1;
    return find_matches(k, seqs, lens, seed, out, max_out, stats, 1);
}

static int64_t seed_tuples(int k, const char *const *seqs, const uint64_t *lens, uint64_t seed, or_match *out,
                           int64_t *offsets, int64_t max_out, int unique)
{
    sml_t sm[OR_MAX_SEQ];
    if (build_smls(k, seqs, lens, seed, sm) != 0) return -1;
    memhash_t h;
    memset(&h, 0, sizeof(h));
    h.k = k;
    h.sm = sm;
    h.unique = unique;
    h.tup_mode = 1;
    h.tup_out = out;
    h.tup_off = offsets;
    h.tup_max = max_out;
    run_search(&h);
    for (int i = 0; i < k; ++i) sml_free(&sm[i]);
    return h.tup_n;
}

int64_t or_seed_tuples_pairwise(int k, const char *const *seqs, const uint64_t *lens,
                                uint64_t seed, or_match *out, int64_t *offsets,
                                int64_t max_out)
{
    return seed_tuples(k, seqs, lens, seed, out, offsets, max_out, 0);
}

int64_t or_seed_tuples_unique(int k, const char *const *seqs, const uint64_t *lens,
                              uint64_t seed, or_match *out, int64_t *offsets,
                              int64_t max_out)
{
    if (k < 2) return -1;
    return seed_tuples(k, seqs, lens, seed, out, offsets, max_out, 1);
}

int or_extend_seed(int k, const char *const *seqs, const uint64_t *lens, uint64_t seed,
                   or_match *m)
{
    sml_t sm[OR_MAX_SEQ];
    if (build_smls(k, seqs, lens, seed, sm) != 0) return -1;
    mhe_t e;
    memset(&e, 0, sizeof(e));
    for (int s = 0; s < k; ++s) e.start[s] = m->start[s];
    e.length = m->length;
    e.mersize = (uint64_t)sm[0].sl;
    extend_match(sm, k, &e);
    for (int s = 0; s < k; ++s) m->start[s] = e.start[s];
    m->length = e.length;
    for (int i = 0; i < k; ++i) sml_free(&sm[i]);
    return 0;
}

/* ======================================================================
 * NWSmall -- MU/nwsmall.cpp:500-670, float arithmetic as in the reference
 * ====================================================================== */

#define MINUS_INFINITY (-1e37f)            /* MU/intmath.h:46 */
enum { BIT_MM = 0, BIT_DM = 1, BIT_IM = 2, BIT_xM = 3,   /* MU/types.h:26-46 */
       BIT_DD = 0, BIT_MD = 4, BIT_xD = 4,
       BIT_II = 0, BIT_MI = 8, BIT_xI = 8 };

/* NUC_SP = hoxd + 2*NUC_EXTEND, MU/nucmx.cpp:8-25 */
static const float NUC_SP[4][4] = {
    {91 + 60.f, -114 + 60.f, -31 + 60.f, -123 + 60.f},
    {-114 + 60.f, 100 + 60.f, -125 + 60.f, -31 + 60.f},
    {-31 + 60.f, -125 + 60.f, 100 + 60.f, -114 + 60.f},
    {-123 + 60.f, -31 + 60.f, -114 + 60.f, 91 + 60.f},
};

/* Match scores of single-sequence profile positions for the six operand
 * classes A, C, G, T, X and any other letter, as libMUSCLE computes them:
 * fractional counts (MU/msa2.cpp:19-90: a letter counts 1; a wildcard's
 * DNA branch tests the amino values AX_R and AX_Y, so 'X' (NX_X == AX_R)
 * takes half to G and half to A, and every other wildcard, with any letter
 * FixAlpha turns into 'N' (MU/msa.cpp:795-813), 1/20 to each base; counts are
 * not normalised for DNA, MU/params.cpp:296-310), m_AAScores = sum_j counts[j]
 * * NUC_SP[i][j] (MU/profilefrommsa.cpp:279-285), SortCounts' bubble order
 * (159-183) and ScoreProfPos2SPN (MU/scorepp.cpp:80-92, centre 0), each step
 * rounded to float. */
float OR_SCORE6[OR_NCLASS][OR_NCLASS];
static int score6_ready;

static void score6_init(void)
{
    if (score6_ready) return;
    const float w = 1.0f;
    float cnt[OR_NCLASS][4];
    memset(cnt, 0, sizeof(cnt));
    for (int c = 0; c < 4; ++c) cnt[c][c] = w;
    cnt[4][2] += w / 2;
    cnt[4][0] += w / 2;
    const float f = w / 20;
    for (int l = 0; l < 4; ++l) cnt[5][l] += f;
    float aa[OR_NCLASS][4];
    int order[OR_NCLASS][4];
    for (int c = 0; c < OR_NCLASS; ++c) {
        for (int i = 0; i < 4; ++i) {
            float sum = 0;
            for (int j = 0; j < 4; ++j) {
                const float t = cnt[c][j] * NUC_SP[i][j];
                sum += t;
            }
            aa[c][i] = sum;
        }
        for (int i = 0; i < 4; ++i) order[c][i] = i;
        int any = 1;
        while (any) {
            any = 0;
            for (int n = 0; n < 3; ++n)
                if (cnt[c][order[c][n]] < cnt[c][order[c][n + 1]]) {
                    const int t = order[c][n];
                    order[c][n] = order[c][n + 1];
                    order[c][n + 1] = t;
                    any = 1;
                }
        }
    }
    for (int a = 0; a < OR_NCLASS; ++a)
        for (int b = 0; b < OR_NCLASS; ++b) {
            float score = 0;
            for (int n = 0; n < 4; ++n) {
                const int l = order[a][n];
                const float fc = cnt[a][l];
                if (fc == 0) break;
                const float t = fc * aa[b][l];
                score += t;
            }
            OR_SCORE6[a][b] = score - 0.0f;
        }
    score6_ready = 1;
}

void or_score6(float *out)
{
    score6_init();
    memcpy(out, OR_SCORE6, sizeof(OR_SCORE6));
}

typedef struct { float open, close; } gp_t;

/* Gap scores of a one-sequence profile (MU/profilefrommsa.cpp:285-292:
 * occupancy 1 => open = close = gap_open/2), then SetTermGaps with
 * TERMGAPS_Half (MU/termgaps.cpp:19-33, falls through into the Ext case). */
static void make_gaps(gp_t *g, uint32_t L, float gap_open)
{
    for (uint32_t i = 0; i < L; ++i) { g[i].open = 1.0f * gap_open / 2; g[i].close = 1.0f * gap_open / 2; }
    if (g[0].open != MINUS_INFINITY) g[0].open = 0;
    if (L > 1 && g[L - 1].close != MINUS_INFINITY) g[L - 1].close = 0;
    if (g[0].open != MINUS_INFINITY) g[0].open *= -1;
    if (L > 1 && g[L - 1].close != MINUS_INFINITY) g[L - 1].close *= -1;
}

/* XChar, MU/bittraceback.cpp:58-136 */
static char xchar(uint8_t bits, char t)
{
    if (t == 'M') { int v = bits & BIT_xM; return v == BIT_MM ? 'M' : v == BIT_DM ? 'D' : v == BIT_IM ? 'I' : '?'; }
    if (t == 'D') return (bits & BIT_xD) == BIT_MD ? 'M' : 'D';
    return (bits & BIT_xI) == BIT_MI ? 'M' : 'I';
}

int64_t or_nwsmall_params(const uint8_t *PAc, uint32_t LA, const uint8_t *PBc, uint32_t LB,
                          float gap_open, float e, char *path_out, float *score_out)
{
    if (LA == 0 || LB == 0) return -1;          /* MU/nwsmall.cpp:503-504 Quit */
    for (uint32_t i = 0; i < LA; ++i) if (PAc[i] >= OR_NCLASS) return -3;
    for (uint32_t j = 0; j < LB; ++j) if (PBc[j] >= OR_NCLASS) return -3;
    score6_init();
    gp_t *GA = (gp_t *)malloc(LA * sizeof(gp_t)), *GB = (gp_t *)malloc(LB * sizeof(gp_t));
    make_gaps(GA, LA, gap_open);
    make_gaps(GB, LB, gap_open);
#define S(a, b) (OR_SCORE6[PAc[a]][PBc[b]])
    const uint32_t PCA = LA + 1, PCB = LB + 1;
    float *MCurr = (float *)malloc(PCB * sizeof(float)), *MNext = (float *)malloc(PCB * sizeof(float));
    float *MPrev = (float *)malloc(PCB * sizeof(float)), *DRow = (float *)malloc(PCB * sizeof(float));
    uint8_t *TBm = (uint8_t *)calloc((size_t)PCA * PCB, 1);
#define TB(i, j) TBm[(size_t)(i) * PCB + (j)]
    float Iij;
    Iij = GB[0].open;
    for (uint32_t j = 2; j <= LB; ++j) Iij += e;
    for (uint32_t j = 0; j <= LB; ++j) DRow[j] = MINUS_INFINITY;
    MPrev[0] = 0;
    for (uint32_t j = 1; j <= LB; ++j) MPrev[j] = MINUS_INFINITY;
    MCurr[0] = MINUS_INFINITY;
    MCurr[1] = S(0, 0);
    TB(1, 1) = (TB(1, 1) & ~BIT_xM) | BIT_MM;
    for (uint32_t j = 2; j <= LB; ++j) {
        MCurr[j] = S(0, j - 1) + GB[0].open + (j - 2) * e + GB[j - 2].close;
        TB(1, j) = (TB(1, j) & ~BIT_xM) | BIT_IM;
    }
    for (uint32_t i = 1; i < LA; ++i) {
        Iij = MINUS_INFINITY;
        DRow[0] = GA[0].open + (i - 1) * e;
        MCurr[0] = MINUS_INFINITY;
        if (i == 1) {
            MCurr[1] = S(0, 0);
            TB(i, 1) = (TB(i, 1) & ~BIT_xM) | BIT_MM;
        } else {
            MCurr[1] = S(i - 1, 0) + GA[0].open + (i - 2) * e + GA[i - 2].close;
            TB(i, 1) = (TB(i, 1) & ~BIT_xM) | BIT_DM;
        }
        for (uint32_t j = 1; j < LB; ++j) MNext[j + 1] = S(i, j);
        for (uint32_t j = 1; j <= LB; ++j) {
            /* RECURSE_D, MU/nwsmall.cpp:68-86 */
            {
                float DD = DRow[j] + e, MD = MPrev[j] + GA[i - 1].open;
                if (DD > MD) DRow[j] = DD;
                else { DRow[j] = MD; TB(i, j) &= ~BIT_xD; TB(i, j) |= BIT_MD; }
            }
            /* RECURSE_I, MU/nwsmall.cpp:91-106 */
            {
                Iij += e;
                float MI = MCurr[j - 1] + GB[j - 1].open;
                if (MI >= Iij) { Iij = MI; TB(i, j) &= ~BIT_xI; TB(i, j) |= BIT_MI; }
            }
            if (j == LB) break;           /* j == uLengthB: BTerm has no M step */
            /* RECURSE_M, MU/nwsmall.cpp:111-142 */
            {
                float DM = DRow[j] + GA[i - 1].close, IM = Iij + GB[j - 1].close, MM = MCurr[j];
                TB(i + 1, j + 1) &= ~BIT_xM;
                if (MM >= DM && MM >= IM) { MNext[j + 1] += MM; TB(i + 1, j + 1) |= BIT_MM; }
                else if (DM >= MM && DM >= IM) { MNext[j + 1] += DM; TB(i + 1, j + 1) |= BIT_DM; }
                else { MNext[j + 1] += IM; TB(i + 1, j + 1) |= BIT_IM; }
            }
        }
        float *t = MPrev; MPrev = MCurr; MCurr = MNext; MNext = t;   /* Rotate */
    }
    /* i == uLengthA, MU/nwsmall.cpp:616-637 */
    MCurr[0] = MINUS_INFINITY;
    if (LA > 1)
        MCurr[1] = S(LA - 1, 0) + (LA - 2) * e + GA[0].open + GA[LA - 2].close;
    else
        MCurr[1] = S(LA - 1, 0) + GA[0].open + GA[0].close;
    TB(LA, 1) = (TB(LA, 1) & ~BIT_xM) | BIT_DM;
    DRow[0] = MINUS_INFINITY;
    for (uint32_t j = 1; j <= LB; ++j) {
        float DD = DRow[j] + e, MD = MPrev[j] + GA[LA - 1].open;
        if (DD > MD) DRow[j] = DD;
        else { DRow[j] = MD; TB(LA, j) &= ~BIT_xD; TB(LA, j) |= BIT_MD; }
    }
    Iij = MINUS_INFINITY;
    for (uint32_t j = 1; j <= LB; ++j) {
        Iij += e;
        float MI = MCurr[j - 1] + GB[j - 1].open;
        if (MI >= Iij) { Iij = MI; TB(LA, j) &= ~BIT_xI; TB(LA, j) |= BIT_MI; }
    }
    float MAB = MCurr[LB], DAB = DRow[LB], IAB = Iij;
    float Score = MAB;
    char edge = 'M';
    if (DAB > Score) { Score = DAB; edge = 'D'; }
    if (IAB > Score) { Score = IAB; edge = 'I'; }
    if (score_out) *score_out = Score;

    /* BitTraceBack, MU/bittraceback.cpp:138-209; edges are prepended, so they
     * are collected in reverse and flipped at the end. */
    int64_t n = 0;
    uint32_t pa = LA, pb = LB;
    char t = edge;
    for (;;) {
        path_out[n++] = t;
        char nx = xchar(TB(pa, pb), t);
        if (t == 'M') { --pa; --pb; }
        else if (t == 'D') { --pa; }
        else { --pb; }
        if (pa == 0 && pb == 0) break;
        t = nx;
        if (t == '?' || (t == 'M' && (pa == 0 || pb == 0)) || (t == 'D' && pa == 0) || (t == 'I' && pb == 0)) {
            n = -2;   /* the reference Quit()s here; never observed */
            break;
        }
    }
    if (n > 0)
        for (int64_t a = 0, b = n - 1; a < b; ++a, --b) { char c = path_out[a]; path_out[a] = path_out[b]; path_out[b] = c; }
#undef TB
#undef S
    free(TBm); free(MCurr); free(MNext); free(MPrev); free(DRow); free(GA); free(GB);
    return n;
}

int64_t or_nwsmall(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                   char *path_out, float *score_out)
{
    return or_nwsmall_params(A, LA, B, LB, -400.0f, 0.0f, path_out, score_out);
}

/* Score of a path under the same transition terms NWSmall adds, including
 * its one-row quirk: with LA == 1 the cell M(1,1) carries the close score of
 * position 0 (MU/nwsmall.cpp:623-624; SetTermGaps leaves close[0] alone when
 * the length is 1). */
float or_path_score(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                    const char *path, int64_t plen)
{
    gp_t *GA = (gp_t *)malloc(LA * sizeof(gp_t)), *GB = (gp_t *)malloc(LB * sizeof(gp_t));
    make_gaps(GA, LA, -400.0f);
    make_gaps(GB, LB, -400.0f);
    float s = 0;
    uint32_t i = 0, j = 0;
    char prev = 'S';
    for (int64_t p = 0; p < plen; ++p) {
        char c = path[p];
        if (c == 'M') {
            ++i; ++j;
            s += NUC_SP[A[i - 1]][B[j - 1]];
            if (prev == 'D') s += GA[i - 2].close;
            else if (prev == 'I') s += GB[j - 2].close;
            if (LA == 1 && i == 1 && j == 1) s += GA[0].close;
        } else if (c == 'D') {
            ++i;
            if (prev != 'D') s += GA[i - 1].open;
        } else {
            ++j;
            if (prev != 'I') s += GB[j - 1].open;
        }
        prev = c;
    }
    free(GA); free(GB);
    return s;
}

/* ======================================================================
 * Banded NWSmall over a stripe region (no reference counterpart; the
 * recurrences, boundaries and tie rules are MU/nwsmall.cpp's, written with
 * virtual row/column 0 values)
 * ====================================================================== */

#define RNEG (-(int64_t)1 << 40)

static void region_rows(uint32_t LA, uint32_t LB, uint32_t band, uint32_t sw, uint32_t s, int64_t *r0, int64_t *r1)
{
    if (band == 0) { *r0 = 1; *r1 = LA; return; }
    int64_t c0 = (int64_t)s * sw + 1, c1 = (int64_t)(s + 1) * sw;
    if (c1 > LB) c1 = LB;
    int64_t lo = c0 - band;
    int64_t a0 = lo <= 0 ? 1 : ((int64_t)LA * lo + LB - 1) / LB;
    if (a0 < 1) a0 = 1;
    int64_t a1 = ((int64_t)LA * (c1 + band)) / LB;
    if (a1 > LA) a1 = LA;
    *r0 = a0;
    *r1 = a1;
}

uint32_t or_path_dev(uint32_t LA, uint32_t LB, const char *path, int64_t plen)
{
    int64_t i = 0, j = 0, mx = 0;
    for (int64_t k = 0; k < plen; ++k) {
        if (path[k] == 'M') { ++i; ++j; }
        else if (path[k] == 'D') ++i;
        else ++j;
        int64_t d = i * (int64_t)LB - j * (int64_t)LA;
        if (d < 0) d = -d;
        if (d > mx) mx = d;
    }
    return (uint32_t)((mx + LA - 1) / LA);
}

/* The DP over the cells (i, j) with jlo[i] <= j <= jhi[i] (i = 1..LA); every
 * other cell of rows/columns >= 1 is -infinity. */
static int64_t region_dp(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB, const int64_t *jlo,
                         const int64_t *jhi, char *path_out, int32_t *score_out, uint32_t *dev_out)
{
    const int64_t W = LB + 1;
    const size_t cells = (size_t)(LA + 1) * (size_t)W;
    int64_t *M = (int64_t *)malloc(cells * sizeof(int64_t));
    int64_t *D = (int64_t *)malloc(cells * sizeof(int64_t));
    int64_t *I = (int64_t *)malloc(cells * sizeof(int64_t));
    uint8_t *T = (uint8_t *)calloc(cells, 1);
#define IX(i, j) ((size_t)(i) * W + (j))
    /* gap scores of a one-sequence profile with half terminal gaps
     * (MU/profilefrommsa.cpp:285-292, MU/termgaps.cpp:19-33): open[0] = 0,
     * close[L-1] = 0 when L > 1, otherwise -200; extend 0 */
#define OPEN_A(x) ((x) == 0 ? 0 : -200)
#define OPEN_B(x) ((x) == 0 ? 0 : -200)
#define CLOSE_A(x) (((x) == LA - 1 && LA > 1) ? 0 : -200)
#define CLOSE_B(x) (((x) == LB - 1 && LB > 1) ? 0 : -200)
    /* virtual boundary: M(0,0) = 0; row 0: I(0,j) = 0 (free leading gap),
     * M = D = -inf; column 0: D(i,0) = 0, M = I = -inf */
    M[IX(0, 0)] = 0; D[IX(0, 0)] = RNEG; I[IX(0, 0)] = RNEG;
    for (uint32_t j = 1; j <= LB; ++j) { M[IX(0, j)] = RNEG; D[IX(0, j)] = RNEG; I[IX(0, j)] = 0; }
    for (uint32_t i = 1; i <= LA; ++i) { M[IX(i, 0)] = RNEG; D[IX(i, 0)] = 0; I[IX(i, 0)] = RNEG; }
    for (uint32_t i = 1; i <= LA; ++i) {
        for (uint32_t j = 1; j <= LB; ++j) {
            const size_t x = IX(i, j);
            if ((int64_t)j < jlo[i] || (int64_t)j > jhi[i]) { M[x] = D[x] = I[x] = RNEG; continue; }
            uint8_t tb = 0;
            /* RECURSE_D (MU/nwsmall.cpp:68-86): tie -> M */
            {
                const int64_t dd = D[IX(i - 1, j)], md = M[IX(i - 1, j)] + OPEN_A(i - 1);
                if (dd > md) D[x] = dd; else { D[x] = md; tb |= BIT_MD; }
            }
            /* RECURSE_I (91-106): tie -> M */
            {
                const int64_t ii = I[IX(i, j - 1)], mi = M[IX(i, j - 1)] + OPEN_B(j - 1);
                if (mi >= ii) { I[x] = mi; tb |= BIT_MI; } else I[x] = ii;
            }
            /* RECURSE_M (111-142): M(i,j) = s + max(M, D + close_A[i-2], I + close_B[j-2])
             * at (i-1, j-1), tie order M, D, I; the one-row case (LA == 1) puts
             * close_A[0] on M(1,1) with a D source (616-626) */
            {
                const int64_t s = (int64_t)NUC_SP[A[i - 1]][B[j - 1]];
                if (LA == 1 && i == 1 && j == 1) {
                    M[x] = s + CLOSE_A(0);
                    tb |= BIT_DM;
                } else {
                    const int64_t mm = M[IX(i - 1, j - 1)];
                    const int64_t dm = D[IX(i - 1, j - 1)] + (i >= 2 ? CLOSE_A(i - 2) : -200);
                    const int64_t im = I[IX(i - 1, j - 1)] + (j >= 2 ? CLOSE_B(j - 2) : -200);
                    if (mm >= dm && mm >= im) { M[x] = s + mm; tb |= BIT_MM; }
                    else if (dm >= mm && dm >= im) { M[x] = s + dm; tb |= BIT_DM; }
                    else { M[x] = s + im; tb |= BIT_IM; }
                }
            }
            T[x] = tb;
        }
    }
    /* final pick (641-656): M, then D if greater, then I if greater */
    int64_t sc = M[IX(LA, LB)];
    char t = 'M';
    if (D[IX(LA, LB)] > sc) { sc = D[IX(LA, LB)]; t = 'D'; }
    if (I[IX(LA, LB)] > sc) { sc = I[IX(LA, LB)]; t = 'I'; }
    int64_t n = 0;
    if (sc < RNEG / 2) {
        n = -2;
    } else {
        int64_t pa = LA, pb = LB;
        for (;;) {
            path_out[n++] = t;
            const char nx = xchar(T[IX(pa, pb)], t);
            if (t == 'M') { --pa; --pb; } else if (t == 'D') --pa; else --pb;
            if (pa == 0 && pb == 0) break;
            t = nx;
            if (t == '?' || pa < 0 || pb < 0 || (t == 'M' && (pa == 0 || pb == 0)) || (t == 'D' && pa == 0) ||
                (t == 'I' && pb == 0)) { n = -2; break; }
        }
        if (n > 0)
            for (int64_t a = 0, b = n - 1; a < b; ++a, --b) { char c = path_out[a]; path_out[a] = path_out[b]; path_out[b] = c; }
    }
    if (score_out) *score_out = (int32_t)sc;
    if (dev_out) *dev_out = n > 0 ? or_path_dev(LA, LB, path_out, n) : 0xFFFFFFFFu;
#undef IX
#undef OPEN_A
#undef OPEN_B
#undef CLOSE_A
#undef CLOSE_B
    free(M); free(D); free(I); free(T);
    return n;
}

int64_t or_nwsmall_region(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                          uint32_t band, uint32_t sw, char *path_out, int32_t *score_out, uint32_t *dev_out)
{
    if (LA == 0 || LB == 0) return -1;
    if (sw == 0) sw = LB;
    /* the stripes holding a row are consecutive (their row ranges grow with
     * the stripe index), so each row's cells are one column range */
    int64_t *jlo = (int64_t *)malloc((LA + 1) * sizeof(int64_t)), *jhi = (int64_t *)malloc((LA + 1) * sizeof(int64_t));
    for (uint32_t i = 0; i <= LA; ++i) { jlo[i] = (int64_t)LB + 1; jhi[i] = 0; }
    const uint32_t ns = (LB + sw - 1) / sw;
    for (uint32_t s = 0; s < ns; ++s) {
        int64_t r0, r1;
        region_rows(LA, LB, band, sw, s, &r0, &r1);
        const int64_t c0 = (int64_t)s * sw + 1, c1 = (int64_t)(s + 1) * sw < LB ? (int64_t)(s + 1) * sw : LB;
        for (int64_t i = r0; i <= r1; ++i) {
            if (c0 < jlo[i]) jlo[i] = c0;
            if (c1 > jhi[i]) jhi[i] = c1;
        }
    }
    const int64_t n = region_dp(A, LA, B, LB, jlo, jhi, path_out, score_out, dev_out);
    free(jlo); free(jhi);
    return n;
}

int64_t or_window_base(uint32_t LA, uint32_t LB, uint32_t band, uint32_t width, uint32_t i)
{
    /* ceil((i*LB - h*LA) / LA) - 1 with h = (width-1)/2, clamped to
     * [0, max(0, LB - width)]: the window is centred on the line */
    (void)band;
    const int64_t h = ((int64_t)width - 1) / 2;
    const int64_t num = (int64_t)i * LB - h * LA;
    const int64_t q = num >= 0 ? (num + LA - 1) / LA : -((-num) / LA);
    int64_t b = q - 1;
    const int64_t top = (int64_t)LB > (int64_t)width ? (int64_t)LB - width : 0;
    if (b > top) b = top;
    if (b < 0) b = 0;
    return b;
}

int64_t or_nwsmall_window(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                          uint32_t band, uint32_t width, char *path_out, int32_t *score_out, uint32_t *dev_out)
{
    if (LA == 0 || LB == 0) return -1;
    int64_t *jlo = (int64_t *)malloc((LA + 1) * sizeof(int64_t)), *jhi = (int64_t *)malloc((LA + 1) * sizeof(int64_t));
    jlo[0] = 1;
    jhi[0] = 0;
    for (uint32_t i = 1; i <= LA; ++i) {
        const int64_t b = or_window_base(LA, LB, band, width, i);
        jlo[i] = b + 1;
        jhi[i] = b + width;
    }
    const int64_t n = region_dp(A, LA, B, LB, jlo, jhi, path_out, score_out, dev_out);
    free(jlo); free(jhi);
    return n;
}
