/*
 * mh_oracle.h -- CPU oracle for the seed-anchoring + inter-anchor DP path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms, used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker.  The product library (libmauvehip.so)
 * never links, loads or calls anything in this directory.
 *
 * Every function cites the reference lines it follows (paths relative to
 * mauve/src/: LM = libMems/libMems, MU = muscle/libMUSCLE).
 *
 * Parity pins (what checks each function beyond its own restatement):
 *   seeds (or_get_seed & co)      : known-answer tests on LM/SeedMasks.h tables
 *   or_translate32 / keys / sort  : oracle/_ref harness built from the reference
 *                                   SortedMerList.cpp + DNAMemorySML.cpp, golden
 *                                   vectors under tests/golden, brute force
 *   or_nwsmall                    : oracle/_ref harness built from libMUSCLE
 *                                   (ProfileProfile -> NWSmall), golden vectors,
 *                                   brute-force path-score optimality
 *   or_find_matches (SearchRange, HashMatch, ExtendMatch, AddHashEntry replay):
 *                                   MatchFinder/MemHash need boost (absent), so
 *                                   no reference build; pinned by brute force
 *                                   (all unique shared seeds, naive maximal
 *                                   extension) and invariants -- see tests.
 */
#ifndef MH_ORACLE_H
#define MH_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAX_SEQ 16

/* ---- seed patterns (LM/SeedMasks.h) ---- */
int64_t  or_get_seed(int weight, int seed_rank);
int      or_get_seed_length(int64_t seed);
int      or_get_seed_weight(int64_t seed);
uint32_t or_default_seed_weight(uint64_t avg_sequence_length);

/* ---- 2-bit packing and keys (LM/SortedMerList.cpp) ---- */
uint64_t or_packed_words(uint64_t len);
int      or_translate32(const char *src, uint64_t len, uint32_t *dest);
uint64_t or_get_mer(const uint32_t *seq, uint64_t pos, uint64_t mer_mask);
uint64_t or_revcomp_mer(uint64_t mer, int mer_length);
uint64_t or_get_seed_mer(const uint32_t *seq, uint64_t pos, uint64_t seed,
                         int seed_length, int seed_weight);
uint64_t or_get_dna_seed_mer(const uint32_t *seq, uint64_t pos, uint64_t seed,
                             int seed_length, int seed_weight);
void     or_fill_dna_sml(const uint32_t *seq, uint64_t len, int seed_length,
                         uint64_t *keys_out);
void     or_fill_dna_seed_sml(const uint32_t *seq, uint64_t n, uint64_t seed,
                              int seed_length, int seed_weight, uint64_t *keys_out);

/* Build a sorted mer list: packed words, keys sorted ascending (ties by
 * position) and the matching positions.  Returns number of entries, or -1. */
/* SortedMerList::FindMer (LM/SortedMerList.cpp:170-179) on a sorted key list:
 * 1 found, 0 not found (result = where the bsearch stopped), -1 a list shorter
 * than the seed (result untouched). */
int      or_find_mer_keys(const uint64_t *keys, uint64_t len, int sl, uint64_t q, uint64_t *result);
int64_t  or_sml_build(const char *ascii, uint64_t len, uint64_t seed,
                      uint32_t *packed_out, uint64_t *keys_out, uint32_t *pos_out);

/* ---- seed matching (LM/MatchFinder.cpp, LM/MemHash.cpp, ...) ---- */
typedef struct {
    int64_t  start[OR_MAX_SEQ];   /* 1-based, <0 = reverse strand, 0 = none */
    uint64_t length;
} or_match;

typedef struct {
    int64_t  hash_calls;      /* HashMatch invocations (seed tuples)      */
    int64_t  collisions;      /* AddHashEntry collisions (dropped seeds)  */
    int64_t  restarts;        /* repeat-limit SearchRange restarts        */
    int64_t  matches;         /* entries in the final match list          */
} or_match_stats;

/* EliminateOverlaps_v2 (LM/ProgressiveAligner.h:301-394) for matches of two
 * genomes, then LengthFilter(min_length) (LM/MatchList.h:680-693); ml is
 * rewritten in the resulting order.  Returns the number kept.  (overlap.cpp) */
int64_t or_eliminate_overlaps_pairwise(or_match *ml, int64_t n, uint64_t min_length);

/* Full pairwise match finding over k genomes (k <= OR_MAX_SEQ):
 * SML build + SearchRange + PairwiseMatchFinder::EnumerateMatches +
 * HashMatch + AddHashEntry/ExtendMatch + GetMatchList order.
 * Returns number of matches written (<= max_out) or -1 on error. */
int64_t or_find_matches_pairwise(int k, const char *const *seqs, const uint64_t *lens,
                                 uint64_t seed, or_match *out, int64_t max_out,
                                 or_match_stats *stats);

/* The seed tuples in HashMatch call order (after SetDirection), with their
 * generalized offsets.  Returns count or -1. */
/* UniqueMatchFinder (MA/UniqueMatchFinder.cpp:36-60) over k genomes (2 <= k
 * <= OR_MAX_SEQ): SearchRange's runs as for the pairwise finder, each run
 * hashed once with every genome that holds its mer exactly once (when there
 * are at least two).  Same outputs and conventions as
 * or_find_matches_pairwise. */
int64_t or_find_matches_unique(int k, const char *const *seqs, const uint64_t *lens,
                               uint64_t seed, or_match *out, int64_t max_out,
                               or_match_stats *stats);
int64_t or_seed_tuples_unique(int k, const char *const *seqs, const uint64_t *lens,
                              uint64_t seed, or_match *out, int64_t *offsets,
                              int64_t max_out);
int64_t or_seed_tuples_pairwise(int k, const char *const *seqs, const uint64_t *lens,
                                uint64_t seed, or_match *out, int64_t *offsets,
                                int64_t max_out);

/* ExtendMatch applied to one seed (starts, length = seed length). */
int or_extend_seed(int k, const char *const *seqs, const uint64_t *lens, uint64_t seed,
                   or_match *m);

/* ---- inter-anchor DP (MU/nwsmall.cpp, MU/bittraceback.cpp) ---- */
/* DP operand classes: 0-3 A,C,G,T (either case; U counts as T), 4 'X', 5 any
 * other letter (libMUSCLE's wildcard counts; see or_score6). */
#define OR_NCLASS 6
/* The 6 x 6 float match scores of the classes (row = A's class). */
void or_score6(float *out);

/* A, B: letter codes 0..3 (A,C,G,T).  path_out receives LA+LB bytes max of
 * 'M'/'D'/'I' in alignment order; returns path length, or -1 if LA==0 or
 * LB==0 (the reference Quit()s there).  *score_out gets the best terminal
 * score of the chosen edge (float, integer-valued). */
int64_t or_nwsmall(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                   char *path_out, float *score_out);

/* Operands are classes 0..5 (OR_NCLASS); -3 for a code outside them.
 * Same DP with explicit gap parameters (open is the full open penalty, split
 * in half between open and close as MU/profilefrommsa.cpp:291-292 does). */
int64_t or_nwsmall_params(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                          float gap_open, float gap_extend,
                          char *path_out, float *score_out);

/* Banded NWSmall: the same DP restricted to the region that the library's
 * banded mode evaluates -- columns cut into stripes of `stripe_w`; stripe s
 * (columns c0 = s*w+1 .. c1 = min((s+1)*w, LB)) holds rows
 * ceil(LA*(c0-W)/LB) .. floor(LA*(c1+W)/LB), clipped to [1, LA]; band 0 is
 * the whole matrix.  Cells outside the region are -infinity; the virtual
 * row 0 / column 0 boundary and every tie rule are NWSmall's.  Integer
 * arithmetic (every operand on this path is an integer).  No counterpart in
 * the reference.  Returns the path length, -1 on an empty operand, -2 when
 * the region holds no path.  *dev_out = ceil(max |i*LB - j*LA| / LA) over the
 * path's positions. */
int64_t or_nwsmall_region(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                          uint32_t band, uint32_t stripe_w,
                          char *path_out, int32_t *score_out, uint32_t *dev_out);

/* Banded NWSmall over a sliding window: row i holds the `width` columns
 * base(i)+1 .. base(i)+width, base(i) = ceil((i*LB - h*LA)/LA) - 1 with
 * h = (width-1)/2, clamped to [0, max(0, LB - width)] (or_window_base): the
 * window is centred on the corner-to-corner line and contains every cell
 * within h of it.  `band` does not move the window (the library sizes the
 * width from the band).  Same DP, boundaries, ties and returns as
 * or_nwsmall_region.  No counterpart in the reference. */
int64_t or_window_base(uint32_t LA, uint32_t LB, uint32_t band, uint32_t width, uint32_t i);
int64_t or_nwsmall_window(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                          uint32_t band, uint32_t width,
                          char *path_out, int32_t *score_out, uint32_t *dev_out);

/* ceil(max |i*LB - j*LA| / LA) over the positions of a path. */
uint32_t or_path_dev(uint32_t LA, uint32_t LB, const char *path, int64_t plen);

/* Score of a given path under the same scheme (used by tests). */
float or_path_score(const uint8_t *A, uint32_t LA, const uint8_t *B, uint32_t LB,
                    const char *path, int64_t plen);

#ifdef __cplusplus
}
#endif
#endif
