/*
 * mauvehip.h -- C ABI of libmauvehip.so, the MI355X (gfx950) implementation of
 * progressiveMauve's seed-anchoring and inter-anchor gapped-DP path.
 *
 * Plain C types only (pointers, sizes, fixed-width integers).  Every entry
 * point returns MH_OK (0) or a negative MH_ERR_* code; mh_last_error() gives
 * the message of the calling thread's last failure.  A context owns one HIP
 * stream and its device scratch; calls on one context must not overlap.
 *
 * Reference interfaces each entry point replaces are cited as
 * file:line relative to mauve/src/ (LM = libMems/libMems,
 * MU = muscle/libMUSCLE).
 */
#ifndef MAUVEHIP_H
#define MAUVEHIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_OK                 0
#define MH_ERR_INVALID       -1   /* bad argument                                   */
#define MH_ERR_DEVICE        -2   /* HIP runtime error                              */
#define MH_ERR_NOMEM         -3   /* device or host allocation failed               */
#define MH_ERR_GAP_CHAR      -4   /* '-' in a genome (LM/SortedMerList.cpp:433-437) */
#define MH_ERR_ZERO_LENGTH   -5   /* empty DP operand (MU/nwsmall.cpp:503-504 Quit) */
#define MH_ERR_UNSUPPORTED   -6   /* input outside what this build implements       */
#define MH_ERR_IO            -7   /* file missing, short or of another format       */

#define MH_MAX_SEQ 4              /* genomes per pairwise match search (the CLI uses
                                     PairwiseMatchFinder for <= 4 genomes,
                                     MA/progressiveMauve.cpp:493-504)                */

typedef struct mh_context mh_context;
typedef struct mh_sml mh_sml;

/* ---- context ---------------------------------------------------------- */
int         mh_create(int device, mh_context **out);
void        mh_destroy(mh_context *ctx);
const char *mh_last_error(void);
/* Synchronise the context stream. */
int         mh_synchronize(mh_context *ctx);
/* Average device time (ms) and launch count of a named kernel since the last
 * reset, measured with HIP events on the context stream.  Names:
 * "dp_forward", "dp_traceback", "sml_keys", "sml_sort", "join", "extend". */
int         mh_kernel_time(mh_context *ctx, const char *kernel, double *avg_ms, int64_t *launches);
void        mh_kernel_time_reset(mh_context *ctx);
int         mh_set_timing(mh_context *ctx, int enabled);

/* ---- seed patterns: getSeed / getSeedLength / getSeedWeight /
 *      getDefaultSeedWeight, LM/SeedMasks.h:298-401 ----------------------- */
int64_t  mh_get_seed(int weight, int seed_rank);
int      mh_get_seed_length(int64_t seed);
int      mh_get_seed_weight(int64_t seed);
uint32_t mh_default_seed_weight(uint64_t avg_sequence_length);

/* ---- sorted mer list ---------------------------------------------------- */

/* Same 16-byte layout as struct bmer, LM/SortedMerList.h:43-46. */
typedef struct {
    uint32_t position;
    uint64_t mer;
} mh_bmer;

/* The SMLHeader fields that carry meaning (LM/SortedMerList.h:48-63). */
typedef struct {
    uint32_t version;          /* 5, LM/DNAFileSML.h:58-62   */
    uint32_t alphabet_bits;    /* 2                          */
    uint64_t seed;
    uint32_t seed_length;
    uint32_t seed_weight;
    uint64_t length;           /* sequence length            */
    uint64_t sml_length;       /* SMLLength()                */
    uint8_t  circular;         /* always 0 in this build     */
} mh_sml_header;

/* Replaces FileSML::Create / MemorySML::Create (LM/FileSML.cpp:401-459,
 * LM/MemorySML.cpp:45-60): 2-bit packing, canonical seed keys and the sort by
 * 64-bit key.  The list stays resident on the device.  Keys with equal value
 * are ordered by ascending position (the reference's std::sort leaves them in
 * an unspecified order).  Host ASCII input. */
int mh_sml_create(mh_context *ctx, const char *seq, uint64_t len, uint64_t seed, mh_sml **out);
/* As mh_sml_create, with the ASCII sequence already in device memory. */
int mh_sml_create_device(mh_context *ctx, const void *d_seq, uint64_t len, uint64_t seed, mh_sml **out);
void mh_sml_free(mh_sml *sml);
/* SortedMerList::GetHeader / Length / SMLLength, LM/SortedMerList.h:181-212. */
int mh_sml_get_header(const mh_sml *sml, mh_sml_header *out);
/* FileSML::Read, LM/FileSML.cpp:470-500: entries [offset, offset+count) of the
 * sorted list, clipped to SMLLength; *n_read gets the number written. */
int mh_sml_read(mh_context *ctx, const mh_sml *sml, uint64_t offset, uint64_t count,
                mh_bmer *out, uint64_t *n_read);
/* GetDnaSeedMer at each given position, LM/SortedMerList.cpp:764-769. */
int mh_sml_get_seed_mers(mh_context *ctx, const mh_sml *sml, const uint32_t *positions,
                         uint64_t n, uint64_t *out);
/* SortedMerList::FindMer, LM/SortedMerList.cpp:170-179 (same bsearch probe
 * order, so the same index on a miss). */
int mh_sml_find_mer(mh_context *ctx, const mh_sml *sml, uint64_t mer, uint64_t *index, int *found);
/* The packed sequence words (translate32, LM/SortedMerList.cpp:425-460);
 * n_words = ceil(2L/32)+2, trailing words zero. */
int mh_sml_get_sequence(mh_context *ctx, const mh_sml *sml, uint32_t *words, uint64_t n_words);

/* .sslist files, format version 5 (the SMLHeader of LM/SortedMerList.h:48-63,
 * 2352 bytes, then the packed sequence words, then SMLLength uint32 sorted
 * positions).  mh_sml_save replaces the file writing of FileSML::Create
 * (LM/FileSML.cpp:427-446); a file it writes loads in the reference and the
 * other way round.  MH_ERR_IO if the file cannot be written. */
int mh_sml_save(mh_context *ctx, const mh_sml *sml, const char *path);
/* Replaces FileSML::LoadFile2 (LM/FileSML.cpp:120-192).  *status (may be
 * NULL) gets LoadFile2's code: 0 loaded; 1 cannot open; 2 short header;
 * 3 version is not 5; 4 short sequence data; 5 fewer than SMLLength
 * positions (MH_ERR_IO for 1-5).  Two codes of this build: 6 a header this
 * build does not represent (alphabet_bits != 2, circular, seed fields that do
 * not match the seed, length >= 2^32; MH_ERR_UNSUPPORTED) and 7 positions
 * that are not a permutation of [0, SMLLength) or keys out of order
 * (MH_ERR_INVALID).  The keys are recomputed from the sequence on the
 * device; the loaded list behaves as one built by mh_sml_create. */
int mh_sml_load(mh_context *ctx, const char *path, mh_sml **out, int *status);

/* ---- seed matching ---------------------------------------------------- */

/* A match as the reference's Match (UngappedLocalAlignment<HybridAbstractMatch>):
 * 1-based starts, negative = reverse strand, 0 = absent (LM/AbstractMatch.h:27). */
typedef struct {
    int64_t  start[MH_MAX_SEQ];
    uint64_t length;
} mh_match;

typedef struct {
    int64_t seed_tuples;     /* HashMatch calls                              */
    int64_t collisions;      /* seeds dropped by AddHashEntry                */
    int64_t matches;         /* entries in the match list                    */
    int64_t repeat_runs;     /* SearchRange restarts after runs above
                                MER_REPEAT_LIMIT (LM/MatchFinder.cpp:253-276)  */
    int64_t walks;           /* seeds extended by the step-by-step walk      */
    int64_t long_walks;      /* of those, walks finished per diagonal        */
} mh_match_stats;

/* Replaces PairwiseMatchFinder/MemHash::FindMatches (LM/MemHash.cpp:109-127,
 * LM/PairwiseMatchFinder.cpp:37-71, LM/MatchFinder.cpp:172-340,
 * LM/MemHash.cpp:167-251, LM/MatchFinder.h:218-374): all matches between the
 * k sorted mer lists, in GetMatchList order (LM/MemHash.h:182-203).
 * k is 2 to 4 (MH_MAX_SEQ), as progressiveMauve uses this finder for up to four
 * genomes (MA/progressiveMauve.cpp:493-504).  If *n_out exceeds cap, only cap entries are
 * written and the call still returns MH_OK with the full count. */
int mh_find_matches_pairwise(mh_context *ctx, const mh_sml *const *smls, int k,
                             mh_match *out, uint64_t cap, uint64_t *n_out,
                             mh_match_stats *stats);

/* The index ranges of k sorted lists that MatchFinder::SearchRange reads when
 * runs above MER_REPEAT_LIMIT (1000) make it restart (LM/MatchFinder.cpp:
 * 166-340 with GetBreakpoint 89-126 and FindMatchSeeds 139-164): each restart
 * drops the run and jumps every list to new start points; the matches found
 * are the runs of the ranges read.  The device search plans the same ranges;
 * this host-only entry point plans them from host copies of the sorted 64-bit
 * keys (mh_sml_read order; len = sequence lengths), for checking.  ranges
 * receives (list, lo, hi) triples, list by list, ascending; *restarts the
 * number of restarts.  MH_ERR_INVALID with *n_ranges set when cap is short. */
int mh_search_live_ranges(int k, const uint64_t *const *keys, const uint64_t *n, const uint64_t *len,
                          uint64_t seed, uint64_t *ranges, uint64_t cap, uint64_t *n_ranges, int64_t *restarts);

/* Same search, result kept in a library-owned list (one call, no size guess). */
typedef struct mh_match_list mh_match_list;
int mh_find_matches_pairwise_list(mh_context *ctx, const mh_sml *const *smls, int k,
                                  mh_match_list **out, mh_match_stats *stats);
uint64_t        mh_match_list_size(const mh_match_list *list);
const mh_match *mh_match_list_data(const mh_match_list *list);
void            mh_match_list_free(mh_match_list *list);
/* Replaces UniqueMatchFinder::FindMatches (MA/UniqueMatchFinder.cpp:36-60 over
 * MemHash::FindMatches, LM/MemHash.cpp:109-127, LM/MatchFinder.cpp:172-340,
 * LM/MemHash.cpp:167-251, LM/MatchFinder.h:218-374), the finder
 * progressiveMauve runs for more than four genomes (MA/progressiveMauve.cpp:
 * 493-498): each run of one masked mer is hashed once, with every genome that
 * holds the mer exactly once, when there are at least two.  k is 2 to 16.
 * The list (GetMatchList order) is library-owned: size() matches of k starts
 * each (starts row-major, size() x k; 1-based, negative = reverse strand,
 * 0 = absent) and a length each.  MH_ERR_UNSUPPORTED when a seed's extension
 * would leave a sorted list, or the seed weight leaves no room for the merged
 * sort key (2 weight + 4 + restart bits above 64). */
typedef struct mh_match_k_list mh_match_k_list;
int mh_find_matches_unique(mh_context *ctx, const mh_sml *const *smls, int k,
                           mh_match_k_list **out, mh_match_stats *stats);
uint64_t        mh_match_k_list_size(const mh_match_k_list *list);
int             mh_match_k_list_genomes(const mh_match_k_list *list);
const int64_t  *mh_match_k_list_starts(const mh_match_k_list *list);
const uint64_t *mh_match_k_list_lengths(const mh_match_k_list *list);
void            mh_match_k_list_free(mh_match_k_list *list);

/* Sequence count, file names and lengths a list carries (the genomes it was
 * found in, or the header of a file it was read from; names "" when unknown). */
int             mh_match_list_seq_count(const mh_match_list *list);
const char     *mh_match_list_seq_file(const mh_match_list *list, int i);
uint64_t        mh_match_list_seq_length(const mh_match_list *list, int i);

/* Match-list files, format version 3 (`progressiveMauve --mums` output).
 * Replaces WriteList, LM/MatchList.h:617-662: header lines, then one line per
 * match "<length>\t<starts...>\t<id>\t0\t0".  seq_files[i] NULL (or
 * seq_files NULL) writes "null", seq_lengths NULL writes 0, as the reference
 * does for sequences it has no name or length for.  The id column is
 * 0x<index+1> where the reference prints the match's address.  n == 0 leaves
 * an empty file.  MH_ERR_IO if the file cannot be written. */
int mh_match_list_write(const char *path, const mh_match *matches, uint64_t n, int seq_count,
                        const char *const *seq_files, const uint64_t *seq_lengths);
/* Replaces ReadList, LM/MatchList.h:524-613 (`--match-input`): MH_ERR_IO for
 * a file that is missing or not format 3 with >= 2 sequences, a match line
 * with subset or superset entries, or a match count that differs from the
 * lines read; MH_ERR_UNSUPPORTED above MH_MAX_SEQ sequences. */
int mh_match_list_read(const char *path, mh_match_list **out);

/* ---- inter-anchor gapped DP ----------------------------------------------- */

/* One DP job: A and B as operand codes at the given offsets of the A and B
 * code buffers.  Codes: 0-3 A,C,G,T (U as T), 4 'X', 5 any other letter --
 * the classes libMUSCLE's profile code gives a position (MU/msa2.cpp:19-90:
 * 'X' counts half to A and half to G, every other wildcard and every letter
 * FixAlpha turns into 'N' 1/20 to each base).  A job whose operands are all
 * A,C,G,T runs the int32 DP; any other runs NWSmall in float, as the
 * reference does. */
typedef struct {
    uint64_t a_off;
    uint64_t b_off;
    uint32_t a_len;
    uint32_t b_len;
} mh_dp_job;

typedef struct {
    int32_t  score;        /* terminal score of the chosen edge (NWSmall 641-656),
                              rounded to the nearest integer for float jobs    */
    uint32_t path_len;     /* number of edges                                     */
    uint64_t path_start;   /* offset of the first edge in the path buffer         */
} mh_dp_result;

/* Replaces GlobalAlign -> NWSmall -> BitTraceBack (MU/glbalign.cpp:69-81,
 * MU/nwsmall.cpp:500-670, MU/bittraceback.cpp:138-209) for a batch of
 * one-sequence profiles scored as ProfileAlignFast sets them up (NUC_SP,
 * gap open -400 split as -200/-200, extend 0, half terminal gaps).
 * Job j's edges ('M','D','I', in alignment order) are written inside the slot
 * [sum_{i<j}(a_len+b_len), +a_len+b_len) of `paths`; results[j].path_start
 * gives the first edge.  Host buffers. */
int mh_dp_align(mh_context *ctx, const uint8_t *a_codes, uint64_t a_bytes,
                const uint8_t *b_codes, uint64_t b_bytes,
                const mh_dp_job *jobs, uint64_t njobs,
                uint8_t *paths, mh_dp_result *results);

/* Device-resident variant: d_a, d_b, d_paths, d_results are device pointers;
 * jobs is a host array (the scheduler reads it).  Runs on the context stream;
 * results are complete when it returns. */
int mh_dp_align_device(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b,
                       const mh_dp_job *jobs, uint64_t njobs,
                       uint8_t *d_paths, mh_dp_result *d_results);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same DP as mh_dp_align_device (same schedule, kernels
 * and results), with every scratch buffer -- each round's job table and
 * results, stripe boundary columns and trace arena -- in one fixed device array
 * of mh_dp_noalloc_scratch_bytes() (64 MiB) that the library holds on each GPU
 * until the process ends; no hipMalloc, hipFree or pinned allocation is made.
 * Jobs run in as many rounds as the scratch needs; a single job whose trace
 * and boundaries exceed it is refused with MH_ERR_INVALID (about 100M cells,
 * e.g. 10,000 x 10,000).  The array is shared by every call on one GPU and
 * keeps the last call's contents: calls that overlap on one GPU (several
 * contexts or threads) race and must be serialised by the caller. */
int mh_dp_align_device_noalloc(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b,
                               const mh_dp_job *jobs, uint64_t njobs,
                               uint8_t *d_paths, mh_dp_result *d_results);
uint64_t mh_dp_noalloc_scratch_bytes(void);

/* Banded mode (no counterpart in the reference, whose NWSmall always fills
 * the whole matrix; the full DP stays the default and the judge of it).
 * bands[j] = W > 0 restricts job j to the region of band 2W around the
 * corner-to-corner line (mh_dp_band_region(a_len, b_len, 2W), a superset of
 * the cells within 2W of the line); W = 0 or bands == NULL is the full DP.
 * The result equals the full DP's whenever the full DP's path stays inside
 * that region.  Edge fallback: a job whose path comes within distance W of
 * the line's band edge (d_dev[j] >= W) is rerun in the region of band 4W and
 * accepted within 2W, else rerun with the full DP, unless flags has
 * MH_BAND_NO_FALLBACK (which returns the first region's result).  d_dev
 * (device, nullable) receives each final path's largest distance from the
 * line, ceil(max |i*b_len - j*a_len| / a_len) over its positions (i, j): for
 * full jobs this is the band label. */
#define MH_BAND_NO_FALLBACK 1u
typedef struct {
    uint64_t banded;          /* jobs run with a band in the first pass        */
    uint64_t rerun_doubled;   /* jobs rerun with the band doubled              */
    uint64_t rerun_full;      /* jobs rerun with the full DP                   */
    uint64_t cells;           /* DP cells evaluated over all passes            */
} mh_band_stats;
int mh_dp_align_banded_device(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b,
                              const mh_dp_job *jobs, uint64_t njobs, const uint32_t *bands,
                              uint32_t flags, uint8_t *d_paths, mh_dp_result *d_results,
                              uint32_t *d_dev, mh_band_stats *stats);

/* The region the scheduler gives a job: returns the effective band (0 when
 * the band's region is the whole matrix) and writes the region kind and width.
 *   MH_REGION_FULL     every cell.
 *   MH_REGION_STRIPES  columns cut into stripes of `width`; stripe s (columns
 *                      c0 = s*w+1 .. c1 = min((s+1)*w, b_len)) evaluates rows
 *                      ceil(a_len*(c0-W)/b_len) .. floor(a_len*(c1+W)/b_len),
 *                      clipped to [1, a_len].
 *   MH_REGION_WINDOW   row i evaluates the `width` columns base(i)+1 ..
 *                      base(i)+width, base(i) = ceil((i*b_len - W*a_len)/a_len)
 *                      - 1 clamped to [0, b_len - width].
 * Both banded regions contain every cell within W of the corner-to-corner
 * line; a window reaches at least 2W on each side.  Stripes are the default;
 * MH_DP_BAND_KIND=window in the environment selects the window where it
 * fits (slope <= 4, band <= 127). */
#define MH_REGION_FULL 0u
#define MH_REGION_STRIPES 1u
#define MH_REGION_WINDOW 2u
uint32_t mh_dp_band_region(uint32_t a_len, uint32_t b_len, uint32_t band, uint32_t *kind, uint32_t *width);
/* The same with flags: MH_REGION_RERUN gives the region of the fallback's
 * second pass (band 4W), which takes 64-lane stripes only: a rerun pass holds
 * few jobs and its time is its longest job's. */
#define MH_REGION_RERUN 1u
uint32_t mh_dp_band_region_ex(uint32_t a_len, uint32_t b_len, uint32_t band, uint32_t flags, uint32_t *kind,
                              uint32_t *width);
/* DP cells the scheduler evaluates for each (a_len, b_len, band) -- the
 * region's cells, 64C per row for a window -- for cost models (host only). */
int mh_dp_band_cells(const uint32_t *a_len, const uint32_t *b_len, const uint32_t *band, uint64_t n,
                     uint64_t *cells);

/* ---- learned band (no counterpart in the reference) ----------------------
 *
 * A 272-256-128-1 ReLU MLP with bf16 weights and fp32 accumulation predicts
 * y = log2(1 + label) for a gap, where the label is the full DP path's
 * largest distance from the corner-to-corner line (mh_dp_align_banded_device
 * d_dev of a full run).  Parameters are fp32 masters in one flat array:
 * W1[256][272], b1[256], W2[128][256], b2[128], w3[128], b3. */
#define MH_MLP_IN 272
#define MH_MLP_H1 256
#define MH_MLP_H2 128
#define MH_MLP_NPARAMS (MH_MLP_H1 * MH_MLP_IN + MH_MLP_H1 + MH_MLP_H2 * MH_MLP_H1 + MH_MLP_H2 + MH_MLP_H2 + 1)
typedef struct mh_mlp mh_mlp;
typedef struct {
    uint32_t epochs;
    uint32_t batch;          /* rows per step (<= 65536); the tail of an epoch is dropped */
    float    lr;             /* Adam step size, cosine-decayed to 0 over the run        */
    float    weight_decay;   /* decoupled, per step                                     */
    uint64_t seed;           /* shuffling                                               */
} mh_mlp_train_opts;

/* Per-gap features (float[njobs][272], device) from the DP jobs' operands:
 * log lengths, length difference, 4-mer L1 distance, the 256 normalised 4-mer
 * count differences, and 12 flank seed densities (8-mers of A near each end
 * found in B within 8 of the corner-to-corner line, 6 window sizes). */
int  mh_gap_features_device(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b,
                            const mh_dp_job *jobs, uint64_t njobs, float *d_feat);
int  mh_mlp_create(mh_context *ctx, uint64_t seed, mh_mlp **out);      /* He-uniform init */
void mh_mlp_free(mh_mlp *mlp);
int  mh_mlp_set_params(mh_context *ctx, mh_mlp *mlp, const float *params);   /* resets Adam */
int  mh_mlp_get_params(mh_context *ctx, const mh_mlp *mlp, float *params);
/* y[rows] (device) = network output for feature rows d_x (device). */
int  mh_mlp_forward(mh_context *ctx, mh_mlp *mlp, const float *d_x, uint64_t rows, float *d_y);
/* Loss sum (y-t)^2/(2 rows) and its gradient (host float[MH_MLP_NPARAMS])
 * for one batch of rows <= 65536 (device inputs). */
int  mh_mlp_grad(mh_context *ctx, mh_mlp *mlp, const float *d_x, const float *d_t, uint64_t rows,
                 float *grads, float *loss);
/* Adam training on device features/targets; epoch_loss (host, nullable)
 * receives the mean batch loss of each epoch. */
int  mh_mlp_train(mh_context *ctx, mh_mlp *mlp, const float *d_x, const float *d_t, uint64_t rows,
                  const mh_mlp_train_opts *opts, float *epoch_loss);
/* Bands (host uint32[rows]) = max(min_band, ceil(2^(y + delta))). */
int  mh_band_predict(mh_context *ctx, mh_mlp *mlp, const float *d_feat, uint64_t rows, float delta,
                     uint32_t min_band, uint32_t *bands);

/* ASCII genome (device) -> DP operand codes (mh_dp_job): A,C,G,T,U in either
 * case 0..3, X 4, any other byte 5; the gap characters '-' and '.' are refused
 * with MH_ERR_GAP_CHAR. */
int mh_ascii_to_codes_device(mh_context *ctx, const void *d_ascii, uint64_t len, uint8_t *d_codes);

/* Inter-anchor DP jobs from a match list.  No counterpart in the reference:
 * there the jobs come from the LCB pass and MUSCLE's anchored ranges
 * (LM/ProgressiveAligner.cpp:1119-1290, MU/anchoredpp.cpp:443-552), which this
 * library does not implement.  This producer takes the heaviest chain of
 * non-overlapping collinear forward matches and emits one job per gap whose
 * sides are both in 1..max_gap.  chain (optional) receives the chained match
 * indices; jobs are offsets into the two genomes' code arrays. */
int mh_chain_gap_jobs(const mh_match *matches, uint64_t n, uint64_t len_a, uint64_t len_b, uint32_t max_gap,
                      uint32_t *chain, uint64_t *n_chain, mh_dp_job *jobs, uint64_t cap, uint64_t *n_jobs);

/* Device memory helpers for callers without their own allocator. */
int mh_device_alloc(mh_context *ctx, uint64_t bytes, void **d_ptr);
int mh_device_free(mh_context *ctx, void *d_ptr);
int mh_memcpy_h2d(mh_context *ctx, void *d_dst, const void *h_src, uint64_t bytes);
int mh_memcpy_d2h(mh_context *ctx, void *h_dst, const void *d_src, uint64_t bytes);

/* ---- recursive anchoring in the gaps between anchors --------------------- */

/* A gap between two anchors as getInterveningCoordinates gives it
 * (LM/ProgressiveAligner.cpp:566-587), per genome: lend = first position after
 * the left anchor, rend = first position of the right anchor (1-based). */
typedef struct {
    int64_t lend[2];
    int64_t rend[2];
} mh_anchor_gap;

/* Replaces ProgressiveAligner::pairwiseAnchorSearch
 * (LM/ProgressiveAligner.cpp:590-679) for a list of gaps of the two genomes
 * d_seq_a / d_seq_b (device ASCII).  Per gap: the pieces lend .. lend +
 * (rend - lend - 1) - 1 of each genome, the seed getSeed(getDefaultSeedWeight
 * (average piece length), 0) (no anchors below weight 5), sorted mer lists and
 * the pairwise match search on the GPU, then EliminateOverlaps_v2
 * (LM/ProgressiveAligner.h:301-394) and LengthFilter(12), shifted to genome
 * coordinates.  The anchors of gap i are entries [first[i], first[i+1]) of the
 * returned list (first has n_gaps + 1 entries), in the reference's order. */
int mh_anchor_gaps(mh_context *ctx, const void *d_seq_a, uint64_t len_a, const void *d_seq_b, uint64_t len_b,
                   const mh_anchor_gap *gaps, uint64_t n_gaps, mh_match_list **out, uint64_t *first);

/* The rows of a chain from mh_chain_gap_jobs as one interval over both whole
 * genomes: the chain's anchors aligned column for column, each stretch before,
 * between and after them unaligned (genome 1's residues against gaps, then
 * gaps against genome 2's).  No counterpart in the reference: a stand-in for
 * the alignment the LCB pass hands to the gapped aligner.  *cols = len_a +
 * len_b - anchored bases; MH_ERR_INVALID if cap is smaller (cols set) or the
 * chain is not forward, ordered and non-overlapping. */
int mh_chain_rows(const mh_match *matches, const uint32_t *chain, uint64_t n_chain, const char *seq_a,
                  uint64_t len_a, const char *seq_b, uint64_t len_b, char *row_a, char *row_b, uint64_t cap,
                  uint64_t *cols);
/* The gaps recurseOnPairs hands to pairwiseAnchorSearch
 * (LM/ProgressiveAligner.cpp:838-905) for an interval on the forward strand of
 * both genomes starting at left_a / left_b (1-based): between consecutive
 * columns aligned in both rows (and the interval's ends), every stretch with
 * more than min_gap residues in each genome (the reference's 200,
 * LM/Aligner.cpp:179), as getInterveningCoordinates gives it. */
int mh_recursion_gaps(const char *row_a, const char *row_b, uint64_t cols, int64_t left_a, int64_t left_b,
                      uint64_t min_gap, mh_anchor_gap *gaps, uint64_t cap, uint64_t *n_gaps);

/* ---- refinement of pairwise gapped alignments ---------------------------- */

#define MH_REFINE_PROFILE  1u   /* refineAlignment(profile_aln = true)  */
#define MH_REFINE_WINDOWS  2u   /* refineAlignment(profile_aln = false) */

/* Replaces the gapped-alignment passes progressiveMauve runs over each aligned
 * interval of two genomes (LM/ProgressiveAligner.cpp:1119-1290; for two genomes
 * the profile pass, then the final refinement pass, 2561-2565 and 3751-3755).
 * The profile pass cuts each interval into pieces (removeLargeGapsPP and the
 * density halving, 1015-1117, 1142-1166) and re-aligns each piece that holds
 * both genomes as ProfileAlignFast(anchored) does (LM/MuscleInterface.cpp:
 * 1053-1162, MU/anchoredpp.cpp:353-552): anchor columns from the smoothed
 * column score, then every column range between them re-aligned from its
 * residues.  The refinement pass re-aligns each piece in windows of 500 or 200
 * columns by its density (RefineFast/RefineW, LM/MuscleInterface.cpp:823-894,
 * MU/refinew.cpp:75-200).  Each re-alignment is one NWSmall of genome 1's
 * residues against genome 2's; all of a pass's jobs, over all n intervals,
 * run as one mh_dp_align batch.  Rows are ASCII with '-' gaps, residues
 * in either case, wildcard letters scored as libMUSCLE scores them; an
 * anchored range with residues of one genome only is MH_ERR_ZERO_LENGTH (the
 * reference's NWSmall quits).  out1[i]/out2[i] receive out_cols[i] columns;
 * caps[i] >= 2 * cols[i] always suffices (MH_ERR_INVALID with out_cols set
 * when a capacity is short).  '-' is the gap character.  stats may be NULL. */
typedef struct {
    uint64_t profile_cells;   /* DP cells of the profile pass (anchored ranges) */
    uint64_t window_cells;    /* DP cells of the refinement pass (windows)      */
} mh_refine_stats;

int mh_refine_pairwise(mh_context *ctx, uint64_t n, const char *const *rows1, const char *const *rows2,
                       const uint64_t *cols, uint32_t passes, char *const *out1, char *const *out2,
                       const uint64_t *caps, uint64_t *out_cols, mh_refine_stats *stats);
/* Host-only parts of the above, for checking: FindAnchorColsPP
 * (MU/anchoredpp.cpp:353-404) over two rows, and the pieces of refineAlignment
 * (lo, hi, unaligned flag; three values per piece). */
int mh_refine_anchor_cols(const char *row1, const char *row2, uint64_t cols, uint32_t *anchors, uint64_t cap,
                          uint64_t *n_anchors);
int mh_refine_pieces(const char *row1, const char *row2, uint64_t cols, uint64_t *pieces, uint64_t cap,
                     uint64_t *n_pieces);

/* ---- XMFA output --------------------------------------------------------- */

/* One genome's row of an aligned interval. */
typedef struct {
    int64_t     start;     /* 1-based left end, negative on the reverse strand, 0 absent */
    uint64_t    length;    /* bases of the genome in the interval                       */
    const char *row;       /* aligned row ('-' for gaps), row_len characters            */
    uint64_t    row_len;   /* the interval's alignment length (equal for all its rows)  */
} mh_xmfa_row;

/* Replaces GenericIntervalList::WriteStandardAlignment (LM/IntervalList.h:353-442):
 * rows[iv * seq_count + s] is genome s of interval iv.  Header lines name each
 * genome's file and format (by extension, as libGenome's gnSourceFactory picks
 * it); intervals of alignment length 0 are skipped; a genome absent from an
 * interval (start 0) is written as 0-0 in the first interval only; rows are
 * wrapped at 80 columns; "=" ends each interval.  n_intervals == 0 leaves an
 * empty file.  backbone_file may be NULL. */
int mh_xmfa_write(const char *path, int seq_count, const char *const *seq_files, const char *backbone_file,
                  const mh_xmfa_row *rows, uint64_t n_intervals);

/* ---- genome index LUT (mauve.buildIndex) ------------------------------ */

/* The column walk of buildIndex, PY/buildindex.py:110-129: lut[genome_length]
 * set to -1, then for each XMFA group (PY/xmfa.py:15-68 grammar) holding
 * sequences 1 and 2, every column whose two characters are equal maps the
 * current 0-based position of sequence 1 to that of sequence 2 (each advanced
 * on a non-gap character from start-1; strand ignored).  *n_groups (may be
 * NULL) gets the number of groups.  MH_ERR_INVALID if a sequence-2 row is
 * shorter than its sequence-1 row (the reference's IndexError). */
int mh_index_from_xmfa(const char *xmfa, uint64_t len, uint64_t genome_length, int32_t *lut,
                       uint64_t *n_groups);
/* The same walk over the XMFA file at path, as buildIndex opens it
 * (PY/buildindex.py:110-118); the file is read into a buffer the library keeps
 * between calls.  MH_ERR_IO if it cannot be read. */
int mh_index_from_xmfa_file(const char *path, uint64_t genome_length, int32_t *lut, uint64_t *n_groups);
/* fixZeroIdx, PY/indexutils.pyx:27-43: an unmapped lut[0] takes
 * lut[k] - k from the first mapped k >= 10 when the first bases agree.
 * MH_ERR_INVALID when no such k exists (the reference's IndexError). */
int mh_index_fix_zero(int32_t *lut, uint64_t n, const char *genome, uint64_t genome_len, const char *ref,
                      uint64_t ref_len);
/* fillGaps, PY/indexutils.pyx:46-74: left to right, a run of -1 after a
 * mapped entry is filled with consecutive values when an entry within
 * max_gap_width of the run's start continues the left neighbour's diagonal. */
int mh_index_fill_gaps(int32_t *lut, uint64_t n, int max_gap_width);
/* smoothEdges, PY/indexutils.pyx:77-108: the same repair at every break of
 * consecutive mapping after a mapped entry, within smoothing_radius. */
int mh_index_smooth_edges(int32_t *lut, uint64_t n, int smoothing_radius);

#ifdef __cplusplus
}
#endif
#endif /* MAUVEHIP_H */
