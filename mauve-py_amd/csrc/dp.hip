// dp.hip -- batched NWSmall-exact affine-gap DP for gfx950.
//
// One wavefront per job.  The B axis is cut into stripes of 512 columns;
// lane k owns columns [8k, 8k+8) of the stripe and walks down the rows of A
// one step behind lane k-1 (anti-diagonal skew), so at step t lane k is on
// row i = t - k + 1.  The right-edge values (M, D, I) of each lane's last
// column move to lane k+1 with a DPP wave_shr:1 each step; lane 0 reads the
// stripe's left boundary column from a per-job buffer written by lane 63 of
// the previous stripe, 64 rows at a time.
//
// Scores are int32.  On this path every operand of the reference's float DP
// is an integer (NUC_SP, -200 open/close halves, extend 0, MU/nucmx.cpp:8-25,
// MU/profilefrommsa.cpp:291-292), and MINUS_INFINITY (-1e37, MU/intmath.h:46)
// only ever meets finite values or itself, so a large negative sentinel
// reproduces every comparison.  Cell rules follow MU/nwsmall.cpp:
//   D(i,j) = max(D(i-1,j), M(i-1,j) + open_A[i-1])       tie -> M   (68-86)
//   I(i,j) = max(I(i,j-1), M(i,j-1) + open_B[j-1])       tie -> M   (91-106)
//   M(i,j) = s(i,j) + max(M(i-1,j-1), D(i-1,j-1) + close, I(i-1,j-1) + close)
//            tie order M, D, I                                      (111-142)
// with the boundary rows/columns of 546-569, 579-595, 616-637 expressed as
// virtual row/column 0 values, and the one-row case (LA == 1) where M(1,1)
// carries close_A[0] and a 'D' trace bit (616-626; SetTermGaps leaves
// close[0] alone for a length-1 profile, MU/termgaps.cpp:24).
//
// Trace bits are the reference's 4-bit code (MU/types.h:26-46): bits 0-1 the
// M source (MM=0, DM=1, IM=2), bit 2 MD, bit 3 MI.  Each lane writes one
// 32-bit word (8 cells) per step, so a wavefront stores 256 contiguous bytes
// per step; the word of cell (i, j) is at
//   stripe*(LA+63)*64 + (i-1 + lane)*64 + lane,  lane = ((j-1) mod 512) / 8.
#include "mh_internal.hpp"
#include "dp_tune_table.h"
#include <algorithm>
#include <cmath>
#include <numeric>
#include <cstdlib>
#include <utility>
#include <type_traits>
#include <cstring>
#include <thread>
#include <mutex>

namespace mh {
namespace {

constexpr int NEG = -(1 << 29);
constexpr int SBIAS = 65;           // smallest NUC_SP entry is -125 + 60 = -65

struct DevJob {
    uint64_t a_off, b_off;
    uint32_t la, lb;
    uint64_t tb_off;    // bytes
    uint64_t bnd_off;   // int32 elements
    uint64_t path_off;  // bytes (slot start)
    uint32_t user;      // index in the caller's job array
    uint32_t cols;      // columns per lane (4, 8 or 16): stripe width 64*cols
    uint32_t band;      // 0: the whole matrix; W > 0: the stripes' rows that meet |j - i*lb/la| <= W
    uint32_t pad_;
};

struct DpOut {
    int32_t score;
    int32_t edge;       // 'M', 'D' or 'I'
};

}  // namespace

// Rows [r0, r1] (1-based) that stripe s (columns s*sw+1 .. min((s+1)*sw, lb))
// computes.  Full DP: every row.  Band W: the rows where some column of the
// stripe lies within W of the corner-to-corner line j = i*lb/la, i.e.
// la*(c0-W)/lb <= i <= la*(c1+W)/lb; empty when r1 < r0.
__host__ __device__ inline void stripe_rows(uint32_t la, uint32_t lb, uint32_t band, uint32_t sw, uint32_t s,
                                            int &r0, int &r1)
{
    if (band == 0) {
        r0 = 1;
        r1 = (int)la;
        return;
    }
    const int64_t c0 = (int64_t)s * sw + 1;
    const int64_t c1e = (int64_t)(s + 1) * sw;
    const int64_t c1 = c1e < (int64_t)lb ? c1e : (int64_t)lb;
    const int64_t lo = c0 - (int64_t)band;
    int64_t a0 = lo <= 0 ? 1 : ((int64_t)la * lo + lb - 1) / lb;
    if (a0 < 1) a0 = 1;
    int64_t a1 = ((int64_t)la * (c1 + (int64_t)band)) / lb;
    if (a1 > (int64_t)la) a1 = la;
    r0 = (int)a0;
    r1 = (int)a1;
}

// Trace bytes of one job run by g lanes: per non-empty stripe (rows + g - 1)
// steps of g lanes x cols/2 bytes.
uint64_t dp_trace_bytes(uint32_t la, uint32_t lb, uint32_t band, uint32_t cols, uint32_t g)
{
    const uint32_t sw = g * cols;
    const uint32_t ns = (lb + sw - 1) / sw;
    uint64_t steps = 0;
    for (uint32_t s = 0; s < ns; ++s) {
        int r0, r1;
        stripe_rows(la, lb, band, sw, s, r0, r1);
        if (r1 >= r0) steps += (uint64_t)(r1 - r0 + 1) + (g - 1);
    }
    return steps * g * (cols / 2);
}

namespace {

__device__ __forceinline__ int shr1_old(int v, int old)
{
    // wave_shr:1 -- lane k receives lane k-1's value, lane 0 receives `old`
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}

// The same shift within the job's lanes: the whole wavefront (G = 64) or each
// 16-lane DPP row (G = 16, row_shr:1; the row's first lane receives `old`).
template <int G> __device__ __forceinline__ int shr1_group(int v, int old)
{
    if constexpr (G == 64) return shr1_old(v, old);
    else return __builtin_amdgcn_update_dpp(old, v, 0x111, 0xF, 0xF, false);
}

template <int... Cs, class F>
__device__ __forceinline__ void static_for(std::integer_sequence<int, Cs...>, F &&f)
{
    (f(std::integral_constant<int, Cs>{}), ...);
}

__device__ __forceinline__ int max3i(int a, int b, int c)
{
    int r;
    asm("v_max3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// One job per G lanes: a wavefront (G = 64), or each 16-lane DPP row of a
// wavefront (G = 16: four jobs at once, for jobs at most 16*COLS columns wide,
// whose skew and column padding would leave most of a wavefront idle).  In
// what follows "64" is G.  D and I are kept shifted by the close half (goh):
// D' = D + goh and I' = I + goh, so the M recurrence reads them without an
// add; every comparison is between two values shifted alike, so ties and
// trace bits are unchanged.  Per row a lane runs two passes over its COLS
// columns: M and D right to left (each needs only the previous row's values of
// its own and the left column), then I left to right (it chains along the
// row).  Both update the registers in place.
//
// Stripe s computes rows [r0, r1] (stripe_rows); at step t lane k is on row
// r0 + t - k.  Cells outside those rows are outside the computed region and
// read as NEG, so a banded job is the exact DP restricted to that region;
// with band 0 every stripe runs all rows.  Lanes outside their row range sit
// out the step (exec mask).  Trace bits are the reference's 4-bit code
// (MU/types.h:26-46): bits 0-1 the M source (MM=0, DM=1, IM=2), bit 2 MD,
// bit 3 MI; each lane stores COLS/2 bytes per step, so a wavefront stores
// 32*COLS contiguous bytes per step.  Cell (i, j) of lane k, column c is at
// byte  stripe_base + ((i - r0 + k)*64 + k)*COLS/2 + c/2, nibble c%2.
//
// PIPE: the job's stripes are shared by P wavefronts, wavefront p0 taking
// stripes p0, p0 + P, ...: stripe s reads the boundary column that stripe
// s - 1 (another wavefront) writes.  Each boundary value travels as one
// 64-bit word, the value and a tag (the writing stripe's index + 1), in one
// of two buffers by stripe parity (zeroed by the host before the round):
// written with atomic stores and read with atomic loads at device scope, a
// word is either the one stripe s - 1 wrote or visibly not, so no fence is
// needed.  A reader fetches a chunk of rows one chunk ahead and checks the
// tags when it is about to use them, fetching again the rows not yet
// written.  A wait is bounded (DP_SPIN_MAX polls), past which it sets *err
// and goes on, and the wavefront waits no more (so a stuck job ends within a
// few bounds, its result flagged by the host as an error).  A buffer is
// rewritten by stripe s + 1 only at rows stripe s has used: it computes a row
// after stripe s published it, and stripe s publishes it after reading it.
// Deadlock-free: parts take tickets from their bin in the order they start
// running, so a part waits only on parts that run or have run, except the
// parts of the one job a bin whose parts are not all taken; those are far
// fewer than the resident wavefronts.
// The result is the same as one wavefront's.
constexpr uint32_t DP_SPIN_MAX = 1u << 22;

template <int COLS, int G, bool PIPE = false>
__device__ __forceinline__ void dp_job(const DevJob &jb, const uint8_t *__restrict__ A,
                                       const uint8_t *__restrict__ B, uint8_t *__restrict__ tb,
                                       int *__restrict__ bnd, DpOut *__restrict__ out, uint32_t sr0,
                                       uint32_t sr1, uint32_t sr2, uint32_t sr3, int goh, int (*lds_in)[64],
                                       int (*lds_out)[64], int p0 = 0, int P = 1,
                                       uint32_t *__restrict__ err = nullptr)
{
    constexpr int STRIPE_W = G * COLS;
    constexpr int NW = (COLS + 7) / 8;               // 32-bit trace words per lane per step
    constexpr int LANE_BYTES = COLS / 2;
    const int lane = (int)threadIdx.x & (G - 1);
    const int gb = (int)threadIdx.x & ~(G - 1);    // the job's slots in the LDS rows
    const int la = (int)jb.la, lb = (int)jb.lb;
    const uint8_t *a = A + jb.a_off;
    const uint8_t *b = B + jb.b_off;
    const int nr = (la + 63) & ~63;
    int *bM = bnd + jb.bnd_off;
    int *bD = bM + nr;
    int *bI = bD + nr;
    uint8_t *tbj = tb + jb.tb_off;
    const int nstripes = (lb + STRIPE_W - 1) / STRIPE_W;
    const int fin_lane = ((lb - 1) % STRIPE_W) / COLS;
    const int fin_col = (lb - 1) % COLS;
    const int go2 = 2 * goh;
    int finM = NEG, finD = NEG, finI = NEG;
    int prev_r0 = 1, prev_r1 = 0;     // rows of the previous stripe's last column held in bnd
    uint64_t tbs = 0;                 // byte offset of the stripe's trace block
    bool gave_up = false;             // PIPE: a wait ran past its bound

    for (int s = PIPE ? p0 : 0; s < nstripes; s += PIPE ? P : 1) {
        int r0, r1;
        stripe_rows(jb.la, jb.lb, jb.band, STRIPE_W, s, r0, r1);
        // PIPE: the tagged boundary words read (written by stripe s - 1, tag s) and written (tag s + 1)
        unsigned long long *pin = nullptr, *pout = nullptr;
        if constexpr (PIPE) {
            // stripe s - 1 runs on another wavefront: its rows, and this
            // stripe's trace block after the blocks of the stripes before it
            prev_r0 = 1;
            prev_r1 = 0;
            if (s > 0) stripe_rows(jb.la, jb.lb, jb.band, STRIPE_W, s - 1, prev_r0, prev_r1);
            tbs = 0;
            for (int q = 0; q < s; ++q) {
                int a0, a1;
                stripe_rows(jb.la, jb.lb, jb.band, STRIPE_W, q, a0, a1);
                if (a1 >= a0) tbs += (uint64_t)(a1 - a0 + G) * G * LANE_BYTES;
            }
            unsigned long long *pb = reinterpret_cast<unsigned long long *>(bnd + jb.bnd_off);
            pin = pb + (size_t)((s + 1) & 1) * 3 * nr;
            pout = pb + (size_t)(s & 1) * 3 * nr;
        }
        if (r1 < r0) {
            prev_r0 = 1;
            prev_r1 = 0;
            continue;
        }
        // PIPE: until every lane with want set holds the words stripe s - 1
        // wrote for row i (tag s), fetch them again
        auto settle = [&](bool want, int i, unsigned long long &wM, unsigned long long &wD,
                          unsigned long long &wI) {
            if constexpr (PIPE) {
                uint32_t spins = 0;
                for (;;) {
                    const bool bad = want && ((uint32_t)(wM >> 32) != (uint32_t)s ||
                                              (uint32_t)(wD >> 32) != (uint32_t)s || (uint32_t)(wI >> 32) != (uint32_t)s);
                    if (__ballot(bad) == 0 || gave_up) break;
                    if (++spins > DP_SPIN_MAX) {
                        if (lane == 0) atomicOr(err, 1u);
                        gave_up = true;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                    if (bad) {
                        wM = __hip_atomic_load(pin + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        wD = __hip_atomic_load(pin + nr + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        wI = __hip_atomic_load(pin + 2 * nr + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        };
        const int nrows = r1 - r0 + 1;
        const int j0 = s * STRIPE_W + 1;
        const int jl = j0 + lane * COLS;
        const int rem = lb - (jl - 1);
        const int ncols = rem < 0 ? 0 : (rem > COLS ? COLS : rem);
        const int span = lb - s * STRIPE_W;
        const int nl = span >= STRIPE_W ? G : (span + COLS - 1) / COLS;
        const bool last_stripe = (s == nstripes - 1);
        const bool top0 = (r0 == 1);

        uint32_t bsh[COLS];
#pragma unroll
        for (int c = 0; c < COLS; ++c) bsh[c] = (c < ncols) ? 8u * (uint32_t)b[jl - 1 + c] : 0u;

        // the row above the stripe: virtual row 0 (M = NEG, D' = NEG + goh,
        // I' = 0 + goh) when the stripe starts at row 1, else outside (NEG)
        int Mp[COLS], Dp[COLS], Ip[COLS];
#pragma unroll
        for (int c = 0; c < COLS; ++c) {
            Mp[c] = NEG;
            Dp[c] = top0 ? NEG + goh : NEG;
            Ip[c] = top0 ? goh : NEG;
        }
        // the cell above-left of the lane's first column: M(0,0) = 0 for column 1,
        // virtual row 0 for other columns of row 0, the previous stripe's last
        // column (row r0-1) for lane 0 when it holds that row, else outside
        int lpM = NEG, lpD = NEG, lpI = NEG;
        if (top0) {
            lpM = jl == 1 ? 0 : NEG;
            lpD = NEG + goh;
            lpI = jl == 1 ? NEG + goh : goh;
        } else if constexpr (PIPE) {
            const int i = r0 - 1;
            const bool want = lane == 0 && s > 0 && i >= prev_r0 && i <= prev_r1;
            unsigned long long wM = 0, wD = 0, wI = 0;
            if (want) {
                wM = __hip_atomic_load(pin + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                wD = __hip_atomic_load(pin + nr + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                wI = __hip_atomic_load(pin + 2 * nr + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            settle(want, i, wM, wD, wI);
            if (want) {
                lpM = (int)(uint32_t)wM;
                lpD = (int)(uint32_t)wD;
                lpI = (int)(uint32_t)wI;
            }
        } else if (lane == 0 && s > 0 && r0 - 1 >= prev_r0 && r0 - 1 <= prev_r1) {
            lpM = bM[r0 - 2];
            lpD = bD[r0 - 2];
            lpI = bI[r0 - 2];
        }
        int outM = NEG, outD = NEG, outI = NEG, acode = 0;
        const int ob2_0 = (jl == 1) ? goh : go2;      // open_B[j-1] + goh for the lane's first column

        // lane 0's inputs, G rows per chunk: column jl-1 of the previous stripe
        // where it computed that row (else NEG), or column 0 (M = NEG,
        // D' = 0 + goh, I' = NEG + goh); and the A codes.  The next chunk is
        // loaded one chunk ahead.
        // (PIPE: the words are fetched here, for rows this stripe computes,
        // and settled before use)
        unsigned long long wM = 0, wD = 0, wI = 0;
        bool want = false;
        int wrow = 0;
        auto left_in = [&](int x, int &m, int &d, int &ii, int &ac) {
            const int i = r0 + x;
            if (s == 0) {
                m = NEG; d = goh; ii = NEG + goh;
            } else if (i >= prev_r0 && i <= prev_r1) {
                if constexpr (PIPE) {
                    want = i <= r1;
                    wrow = i;
                    if (want) {
                        wM = __hip_atomic_load(pin + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        wD = __hip_atomic_load(pin + nr + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        wI = __hip_atomic_load(pin + 2 * nr + i - 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    m = NEG; d = NEG; ii = NEG;
                } else {
                    m = bM[i - 1]; d = bD[i - 1]; ii = bI[i - 1];
                }
            } else {
                if constexpr (PIPE) want = false;
                m = NEG; d = NEG; ii = NEG;
            }
            ac = i <= r1 ? a[i - 1] : 0;
        };
        // PIPE: the fetched words, once settled, as the lane's inputs
        auto take = [&](int &m, int &d, int &ii) {
            if constexpr (PIPE) {
                settle(want, wrow, wM, wD, wI);
                if (want) {
                    m = (int)(uint32_t)wM;
                    d = (int)(uint32_t)wD;
                    ii = (int)(uint32_t)wI;
                }
            }
        };
        int nM, nD, nI, nA;
        left_in(lane, nM, nD, nI, nA);

        const int nsteps = nrows + nl - 1;
        for (int t = 0; t < nsteps; ++t) {
            const int q = t & (G - 1);
            if (q == 0) {
                take(nM, nD, nI);
                lds_in[0][gb + lane] = nM;
                lds_in[1][gb + lane] = nD;
                lds_in[2][gb + lane] = nI;
                lds_in[3][gb + lane] = nA;
                left_in(t + G + lane, nM, nD, nI, nA);
            }
            const int rM = shr1_group<G>(outM, lds_in[0][gb + q]);
            const int rD = shr1_group<G>(outD, lds_in[1][gb + q]);
            const int rI = shr1_group<G>(outI, lds_in[2][gb + q]);
            const int ra = shr1_group<G>(acode, lds_in[3][gb + q]);
            acode = ra;

            const int i = r0 + t - lane;
            if (i >= r0 && i <= r1 && lane < nl) {
            const uint32_t srow = (ra & 2) ? ((ra & 1) ? sr3 : sr2) : ((ra & 1) ? sr1 : sr0);
            const int oa2 = (i == 1) ? goh : go2;   // open_A[i-1] + goh
            uint32_t word[NW];
#pragma unroll
            for (int w = 0; w < NW; ++w) word[w] = 0;
            // pass 1, right to left: D'(i,j) and M(i,j) from row i-1
            static_for(std::make_integer_sequence<int, COLS>{}, [&](auto ic) {
                constexpr int c = COLS - 1 - decltype(ic)::value;
                constexpr uint32_t K = 1u << (4 * (c & 7));
                const int md = Mp[c] + oa2;
                const bool tD = md >= Dp[c];
                const int Dn = max(md, Dp[c]);
                const int dM = c ? Mp[c > 0 ? c - 1 : 0] : lpM;
                const int dD = c ? Dp[c > 0 ? c - 1 : 0] : lpD;
                const int dI = c ? Ip[c > 0 ? c - 1 : 0] : lpI;
                const int best = max3i(dM, dD, dI);
                uint32_t srcK = (dM != best) ? ((dD == best) ? K : 2u * K) : 0u;
                const int sc = (int)__builtin_amdgcn_ubfe(srow, bsh[c], 8) - SBIAS;
                int Mn = sc + best;
                if (c == 0 && la == 1 && jl == 1 && i == 1) {
                    // one-row quirk: with LA == 1, M(1,1) carries close_A[0] and a 'D' bit
                    Mn = sc + goh;
                    srcK = K;
                }
                Mp[c] = Mn;
                Dp[c] = Dn;
                word[c >> 3] |= srcK | (tD ? 4u * K : 0u);
            });
            // pass 2, left to right: I'(i,j) from I'(i,j-1) and M(i,j-1)
            int iL = rI, mL = rM;
            static_for(std::make_integer_sequence<int, COLS>{}, [&](auto ic) {
                constexpr int c = decltype(ic)::value;
                constexpr uint32_t K = 1u << (4 * (c & 7));
                const int mi = mL + (c == 0 ? ob2_0 : go2);
                const bool tI = mi >= iL;
                iL = max(mi, iL);
                Ip[c] = iL;
                mL = Mp[c];
                word[c >> 3] |= tI ? 8u * K : 0u;
            });
            outM = Mp[COLS - 1];
            outD = Dp[COLS - 1];
            outI = Ip[COLS - 1];
            lpM = rM;
            lpD = rD;
            lpI = rI;
            uint8_t *dst = tbj + tbs + ((size_t)t * G + lane) * LANE_BYTES;
            if constexpr (COLS == 16) {
                *reinterpret_cast<uint2 *>(dst) = make_uint2(word[0], word[1]);
            } else if constexpr (COLS == 8) {
                *reinterpret_cast<uint32_t *>(dst) = word[0];
            } else if constexpr (COLS == 4) {
                *reinterpret_cast<uint16_t *>(dst) = (uint16_t)word[0];
            } else if constexpr (COLS == 12) {
                // 6 bytes a lane: 2-byte aligned
                uint16_t *h = reinterpret_cast<uint16_t *>(dst);
                h[0] = (uint16_t)word[0];
                h[1] = (uint16_t)(word[0] >> 16);
                h[2] = (uint16_t)word[1];
            } else {
                // an odd number of bytes a lane (6, 10 or 14 columns): byte stores
#pragma unroll
                for (int y = 0; y < LANE_BYTES; ++y) dst[y] = (uint8_t)(word[y >> 2] >> (8 * (y & 3)));
            }
            }
            if (!last_stripe) {
                // the last lane just finished relative row t-(G-1) of the stripe's last column
                const int xl = t - (G - 1);
                if (lane == G - 1 && xl >= 0 && xl < nrows) {
                    const int slot = gb + (xl & (G - 1));
                    lds_out[0][slot] = outM;
                    lds_out[1][slot] = outD;
                    lds_out[2][slot] = outI;
                }
                if (xl >= 0 && xl < nrows && ((xl & (G - 1)) == G - 1 || xl == nrows - 1)) {
                    const int x = (xl & ~(G - 1)) + lane;
                    if (x < nrows) {
                        const int r = r0 - 1 + x;
                        if constexpr (PIPE) {
                            const unsigned long long tag = (unsigned long long)(uint32_t)(s + 1) << 32;
                            __hip_atomic_store(pout + r, tag | (uint32_t)lds_out[0][gb + lane], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                            __hip_atomic_store(pout + nr + r, tag | (uint32_t)lds_out[1][gb + lane], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                            __hip_atomic_store(pout + 2 * nr + r, tag | (uint32_t)lds_out[2][gb + lane],
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        } else {
                            bM[r] = lds_out[0][gb + lane];
                            bD[r] = lds_out[1][gb + lane];
                            bI[r] = lds_out[2][gb + lane];
                        }
                    }
                }
            }
        }
        if (last_stripe && r1 == la && lane == fin_lane) {
            // the lane's registers hold row LA, the last it computed: the
            // final cell's values, picked by way of LDS (the last stripe
            // stages no boundary there), so that no per-column select on
            // fin_col is held through the step loop
#pragma unroll
            for (int c = 0; c < COLS; ++c) {
                lds_out[0][gb + c] = Mp[c];
                lds_out[1][gb + c] = Dp[c];
                lds_out[2][gb + c] = Ip[c];
            }
            finM = lds_out[0][gb + fin_col];
            finD = lds_out[1][gb + fin_col];
            finI = lds_out[2][gb + fin_col];
        }
        prev_r0 = r0;
        prev_r1 = r1;
        tbs += (uint64_t)(nrows + G - 1) * G * LANE_BYTES;
        if constexpr (!PIPE) __threadfence();
    }
    // (PIPE: the wavefront that ran the last stripe)
    if (lane == fin_lane && (!PIPE || (p0 < nstripes && (nstripes - 1 - p0) % P == 0))) {
        // back to unshifted D and I for the final pick, MU/nwsmall.cpp:641-656
        finD -= goh;
        finI -= goh;
        int score = finM, edge = 'M';
        if (finD > score) { score = finD; edge = 'D'; }
        if (finI > score) { score = finI; edge = 'I'; }
        out->score = score;
        out->edge = edge;
    }
}


// Row-packed jobs: four jobs of one shape per wavefront, one per 16-lane row
// (DevJob::cols = GROUP_TAG + COLS, COLS even from 4 to 16: the narrowest
// stripe that holds the job).
constexpr uint32_t GROUP_TAG = 32;
constexpr uint32_t ROWS_MAX_LB = 256;   // widest job run four to a wavefront

// ---- one round of jobs: a length-binned work queue ----------------------
//
// A round's jobs sit in one table in queue order: the 64-lane stripe jobs
// (those with long trace-back walks first, each part largest first), the
// row-packed jobs by shape (16, 14, ..., 4 columns a lane), the sliding-window
// jobs, the one-lane jobs.  The queue cuts that table into bins: stripe jobs
// by trace size (a new bin where it falls below the next power of 4), one bin
// per row shape, one of window jobs, one of one-lane jobs.  A bin hands out
// items (1 stripe or window job, 4 row-packed jobs, 64 one-lane jobs) through
// its own atomic cursor.  One launch drains the bins in order: a wavefront
// takes the next item of the first bin that is not empty (a persistent
// wavefront keeps taking items until the last bin is empty).  Everything a
// launch needs is in the round's descriptor in device memory, so the
// launches of a round take one pointer and are captured in a graph once per
// context; a round only sets the grid sizes of the graph's nodes.
constexpr int ROW_SHAPES = 7;
// QK_PIPE + P - 2 (P = 2 .. 8): 64-lane stripe jobs whose stripes P
// wavefronts share (an item is one wavefront's part: job first + item / P,
// part item % P)
enum : uint32_t { QK_STRIPE = 0, QK_ROWS = 1, QK_WINDOW = QK_ROWS + ROW_SHAPES, QK_TINY = QK_WINDOW + 1,
                  QK_PIPE = QK_TINY + 1 };
constexpr int DPQ_MAXB = 48;
constexpr int DPQ_CURSORS = DPQ_MAXB;

struct DpRound {
    const DevJob *jobs;
    DpOut *outs;
    const uint8_t *A, *B;
    uint8_t *tb;
    int *bnd;
    uint8_t *paths;
    mh_dp_result *res;
    uint32_t *dev;
    uint32_t *cursor;          // DPQ_CURSORS counters, zeroed at the start of every round; the last
                               // one is the round's wait-bound error flag
    uint32_t sr[4];
    int32_t goh;
    uint32_t nbins;
    uint32_t nwave;            // jobs [0, nwave) walk back on a wavefront each
    uint32_t n;                // jobs in the round
    uint32_t kind[DPQ_MAXB], first[DPQ_MAXB], count[DPQ_MAXB];
    uint32_t item0[DPQ_MAXB + 1];   // the bins' first items in queue order (prefix sums)
};

__device__ __forceinline__ uint32_t qk_per_item(uint32_t kind)
{
    return kind == QK_STRIPE || kind == QK_WINDOW || kind >= QK_PIPE ? 1u : (kind == QK_TINY ? 64u : 4u);
}

// ---- sliding-window banded DP ---------------------------------------------
//
// A banded job whose band fits 64*C columns (C = 1, 2, 4, 8) runs as a window
// that slides along the corner-to-corner line: all 64 lanes work on one row,
// lane k holding window columns base(i)+1+kC .. +C, with
//   base(i) = clamp(ceil((i*lb - h*la)/la) - 1, 0, lb - 64C),  h = (64C-1)/2
// so the window is centred on the line and holds every cell within h of it.
// Cells outside the window are -infinity: the result is the exact DP over the
// window region.  Per row: the previous row's M, D', I' shift left by
// s = base(i) - base(i-1) columns (DPP wave_shl:1 per column step), M and D'
// follow from the previous row (right-to-left register pass, as dp_job), and
// I' -- a running maximum along the row since the gap extension is 0 -- is a
// lane-local prefix maximum joined by a cross-lane exclusive prefix maximum
// (DPP row_shr within 16-lane rows, readlane across rows).  Trace: C*4 bits
// per lane per row, 8/C rows per 32-bit word; the word of row i, lane k is
// at ((i-1) / (8/C)) * 64 + k.
constexpr uint32_t WINDOW_TAG = 64;   // DevJob::cols = WINDOW_TAG + C

__host__ __device__ inline int window_base(int la, int lb, int width, int i)
{
    const int64_t num = (int64_t)i * lb - (int64_t)((width - 1) / 2) * la;
    const int64_t q = num >= 0 ? (num + la - 1) / la : -((-num) / la);
    int64_t b = q - 1;
    const int64_t top = lb > width ? (int64_t)lb - width : 0;
    if (b > top) b = top;
    if (b < 0) b = 0;
    return (int)b;
}

__device__ __forceinline__ int shl1_old(int v, int old)
{
    // wave_shl:1 -- lane k receives lane k+1's value, lane 63 receives `old`
    return __builtin_amdgcn_update_dpp(old, v, 0x130, 0xF, 0xF, false);
}

__device__ __forceinline__ int row_shr_old(int v, int old, int n)
{
    switch (n) {
    case 1: return __builtin_amdgcn_update_dpp(old, v, 0x111, 0xF, 0xF, false);
    case 2: return __builtin_amdgcn_update_dpp(old, v, 0x112, 0xF, 0xF, false);
    case 4: return __builtin_amdgcn_update_dpp(old, v, 0x114, 0xF, 0xF, false);
    default: return __builtin_amdgcn_update_dpp(old, v, 0x118, 0xF, 0xF, false);
    }
}

template <int C>
__device__ __forceinline__ void dp_window_job(const DevJob &jb, const uint8_t *__restrict__ A,
                                              const uint8_t *__restrict__ B, uint8_t *__restrict__ tb,
                                              DpOut *__restrict__ out, uint32_t sr0, uint32_t sr1, uint32_t sr2,
                                              uint32_t sr3, int goh)
{
    constexpr int WD = 64 * C;
    constexpr int RPD = 8 / C;   // rows per trace word
    const int lane = threadIdx.x;
    const int la = (int)jb.la, lb = (int)jb.lb;
    const uint8_t *a = A + jb.a_off;
    const uint8_t *b = B + jb.b_off;
    uint32_t *tbj = reinterpret_cast<uint32_t *>(tb + jb.tb_off);
    const int go2 = 2 * goh;
    // base(i) without a division per row: q = ceil((i*lb - h*la)/la) kept with
    // its remainder rem = q*la - (i*lb - h*la) in [0, la)
    const int top = lb > WD ? lb - WD : 0;
    const int dq = lb / la, dr = lb % la;
    int q, rem;
    {
        const int64_t num = (int64_t)lb - (int64_t)((WD - 1) / 2) * la;
        const int64_t qq = num >= 0 ? (num + la - 1) / la : -((-num) / la);
        q = (int)qq;
        rem = (int)(qq * la - num);
    }
    auto clampb = [&](int v) { return v - 1 > top ? top : (v - 1 < 0 ? 0 : v - 1); };
    int bprev = clampb(q);
    int Mp[C], Dp[C], Ip[C];
    uint32_t bsh[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {   // row 0 in the frame of base(1)
        Mp[c] = NEG;
        Dp[c] = NEG + goh;
        Ip[c] = goh;
        const int j0 = bprev + lane * C + c;
        bsh[c] = j0 < lb ? 8u * (uint32_t)b[j0] : 0u;
    }
    // A codes 64 rows at a time, the next chunk loaded one chunk ahead
    int achunk = lane < la ? a[lane] : 0;
    int anext = 64 + lane < la ? a[64 + lane] : 0;
    uint32_t bnx[C];   // B codes of the next row's window, loaded a row ahead
#pragma unroll
    for (int c = 0; c < C; ++c) bnx[c] = bsh[c];
    uint32_t tbw = 0;
    int finM = NEG, finD = NEG, finI = NEG;
    for (int i = 1; i <= la; ++i) {
        if (i > 1) {
            q += dq;
            rem -= dr;
            if (rem < 0) {
                ++q;
                rem += la;
            }
        }
        const int bi = clampb(q);
        const int s = bi - bprev;
        if (s) {
            // this row's window codes, loaded during the previous row
#pragma unroll
            for (int c = 0; c < C; ++c) bsh[c] = bnx[c];
        }
        {   // issue the next row's B codes if its window moves
            int qn = q + dq;
            if (rem - dr < 0) ++qn;
            const int bn = clampb(qn);
            if (bn != bi && i < la) {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int j0 = bn + lane * C + c;
                    bnx[c] = j0 < lb ? 8u * (uint32_t)b[j0] : 0u;
                }
            }
        }
        // (i-1, bi): the cell left of the window on the previous row
        int ldM, ldD, ldI;
        if (i == 1) {
            ldM = bi == 0 ? 0 : NEG;
            ldD = NEG + goh;
            ldI = bi == 0 ? NEG + goh : goh;
        } else if (bi == 0) {
            ldM = NEG;
            ldD = goh;
            ldI = NEG + goh;
        } else {
            ldM = NEG;
            ldD = NEG;
            ldI = NEG;
            for (int st = 0; st < s; ++st) {
                if (st == s - 1) {
                    ldM = __builtin_amdgcn_readlane(Mp[0], 0);
                    ldD = __builtin_amdgcn_readlane(Dp[0], 0);
                    ldI = __builtin_amdgcn_readlane(Ip[0], 0);
                }
                const int eM = shl1_old(Mp[0], NEG), eD = shl1_old(Dp[0], NEG), eI = shl1_old(Ip[0], NEG);
#pragma unroll
                for (int c = 0; c + 1 < C; ++c) {
                    Mp[c] = Mp[c + 1];
                    Dp[c] = Dp[c + 1];
                    Ip[c] = Ip[c + 1];
                }
                Mp[C - 1] = eM;
                Dp[C - 1] = eD;
                Ip[C - 1] = eI;
            }
        }
        if (i > 1 && ((i - 1) & 63) == 0) {
            achunk = anext;
            anext = (i + 63 + lane < la) ? a[i + 63 + lane] : 0;
        }
        const int ra = __builtin_amdgcn_readlane(achunk, (i - 1) & 63);
        const uint32_t srow = (ra & 2) ? ((ra & 1) ? sr3 : sr2) : ((ra & 1) ? sr1 : sr0);
        const int oa2 = (i == 1) ? goh : go2;   // open_A[i-1] + goh
        // diagonal inputs of the lane's first column: lane k-1's last column
        // of the previous row, or the cell left of the window for lane 0
        const int dM0 = shr1_old(Mp[C - 1], ldM), dD0 = shr1_old(Dp[C - 1], ldD), dI0 = shr1_old(Ip[C - 1], ldI);
        uint32_t bits = 0;
        // pass 1, right to left: D'(i,j) and M(i,j) from row i-1
        static_for(std::make_integer_sequence<int, C>{}, [&](auto ic) {
            constexpr int c = C - 1 - decltype(ic)::value;
            constexpr uint32_t K = 1u << (4 * c);
            const int md = Mp[c] + oa2;
            const bool tD = md >= Dp[c];
            const int Dn = max(md, Dp[c]);
            const int dM = c ? Mp[c > 0 ? c - 1 : 0] : dM0;
            const int dD = c ? Dp[c > 0 ? c - 1 : 0] : dD0;
            const int dI = c ? Ip[c > 0 ? c - 1 : 0] : dI0;
            const int best = max3i(dM, dD, dI);
            uint32_t srcK = (dM != best) ? ((dD == best) ? K : 2u * K) : 0u;
            const int sc = (int)__builtin_amdgcn_ubfe(srow, bsh[c], 8) - SBIAS;
            int Mn = sc + best;
            if (c == 0 && la == 1 && i == 1 && bi == 0 && lane == 0) {
                // one-row quirk: with LA == 1, M(1,1) carries close_A[0] and a 'D' bit
                Mn = sc + goh;
                srcK = K;
            }
            Mp[c] = Mn;
            Dp[c] = Dn;
            bits |= srcK | (tD ? 4u * K : 0u);
        });
        // pass 2: I'(i,j) = max(I'(i,j-1), M(i,j-1) + open_B[j-1] + goh) along the row
        const int mleft = shr1_old(Mp[C - 1], NEG);   // M(i, first column - 1); left of the window: NEG
        const int j1 = bi + lane * C + 1;                // the lane's first column
        int cand[C], pre[C];
        int run = NEG;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int mprev = c ? Mp[c > 0 ? c - 1 : 0] : mleft;
            cand[c] = mprev + ((j1 + c == 1) ? goh : go2);
            run = max(run, cand[c]);
            pre[c] = run;
        }
        // exclusive prefix maximum of the lane totals across the wave
        int v = run;
        v = max(v, row_shr_old(v, NEG, 1));
        v = max(v, row_shr_old(v, NEG, 2));
        v = max(v, row_shr_old(v, NEG, 4));
        v = max(v, row_shr_old(v, NEG, 8));
        const int t15 = __builtin_amdgcn_readlane(v, 15), t31 = __builtin_amdgcn_readlane(v, 31);
        const int t47 = __builtin_amdgcn_readlane(v, 47);
        const int r1 = t15, r2 = max(t15, t31), r3 = max(r2, t47);
        v = max(v, lane >= 48 ? r3 : (lane >= 32 ? r2 : (lane >= 16 ? r1 : NEG)));
        const int leftI = bi == 0 ? NEG + goh : NEG;    // I'(i, window start - 1)
        const int excl = shr1_old(v, leftI);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int iprev = c ? max(excl, pre[c > 0 ? c - 1 : 0]) : excl;
            const bool tI = cand[c] >= iprev;
            Ip[c] = max(iprev, cand[c]);
            bits |= tI ? (8u << (4 * c)) : 0u;
        }
        const int sub = (i - 1) % RPD;
        tbw |= bits << (4 * C * sub);
        if (sub == RPD - 1 || i == la) {
            tbj[(size_t)((i - 1) / RPD) * 64 + lane] = tbw;
            tbw = 0;
        }
        if (i == la && lane == 63) {
            finM = Mp[C - 1];
            finD = Dp[C - 1];
            finI = Ip[C - 1];
        }
        bprev = bi;
    }
    if (lane == 63) {
        finD -= goh;
        finI -= goh;
        int score = finM, edge = 'M';
        if (finD > score) { score = finD; edge = 'D'; }
        if (finI > score) { score = finI; edge = 'I'; }
        out->score = score;
        out->edge = edge;
    }
}

// Narrow jobs (lb <= 16): one lane per job.  The lane keeps a whole row of
// M, D' and I' in registers and walks the rows of A, with the same two
// register passes per row as dp_job; columns past lb are computed and
// ignored (nothing flows leftwards).  The left boundary is column 0 and the
// top row 0, as for the first stripe of dp_job.  Trace: 8 bytes per row,
// cell (i, j) at byte (i-1)*8 + (j-1)/2, nibble (j-1)%2.
constexpr int TINY_LB = 16;
constexpr uint32_t TINY_COLS_TAG = 1;   // DevJob::cols of a one-lane job

__device__ __forceinline__ void tiny_job(const DevJob &jb, const uint8_t *__restrict__ A,
                                         const uint8_t *__restrict__ B, uint8_t *__restrict__ tb,
                                         DpOut *__restrict__ out, uint32_t sr0, uint32_t sr1, uint32_t sr2,
                                         uint32_t sr3, int goh)
{
    const int la = (int)jb.la, lb = (int)jb.lb;
    const uint8_t *a = A + jb.a_off;
    const uint8_t *b = B + jb.b_off;
    uint8_t *tbj = tb + jb.tb_off;
    const int go2 = 2 * goh;
    uint32_t bsh[TINY_LB];
#pragma unroll
    for (int c = 0; c < TINY_LB; ++c) bsh[c] = c < lb ? 8u * (uint32_t)b[c] : 0u;
    int Mp[TINY_LB], Dp[TINY_LB], Ip[TINY_LB];
#pragma unroll
    for (int c = 0; c < TINY_LB; ++c) {   // row 0: M = NEG, D' = NEG + goh, I' = 0 + goh
        Mp[c] = NEG;
        Dp[c] = NEG + goh;
        Ip[c] = goh;
    }
    int finM = NEG, finD = NEG, finI = NEG;
    for (int i = 1; i <= la; ++i) {
        const int ra = a[i - 1];
        const uint32_t srow = (ra & 2) ? ((ra & 1) ? sr3 : sr2) : ((ra & 1) ? sr1 : sr0);
        const int oa2 = (i == 1) ? goh : go2;
        // column 0: (i-1, 0) is M(0,0) = 0 on row 1, else M = NEG, D' = 0 + goh
        const int lpM = (i == 1) ? 0 : NEG, lpD = (i == 1) ? NEG + goh : goh, lpI = NEG + goh;
        uint32_t word[2] = {0u, 0u};
        static_for(std::make_integer_sequence<int, TINY_LB>{}, [&](auto ic) {
            constexpr int c = TINY_LB - 1 - decltype(ic)::value;
            constexpr uint32_t K = 1u << (4 * (c & 7));
            const int md = Mp[c] + oa2;
            const bool tD = md >= Dp[c];
            const int Dn = max(md, Dp[c]);
            const int dM = c ? Mp[c > 0 ? c - 1 : 0] : lpM;
            const int dD = c ? Dp[c > 0 ? c - 1 : 0] : lpD;
            const int dI = c ? Ip[c > 0 ? c - 1 : 0] : lpI;
            const int best = max3i(dM, dD, dI);
            uint32_t srcK = (dM != best) ? ((dD == best) ? K : 2u * K) : 0u;
            const int sc = (int)__builtin_amdgcn_ubfe(srow, bsh[c], 8) - SBIAS;
            int Mn = sc + best;
            if (c == 0 && la == 1 && i == 1) {
                // one-row quirk: with LA == 1, M(1,1) carries close_A[0] and a 'D' bit
                Mn = sc + goh;
                srcK = K;
            }
            Mp[c] = Mn;
            Dp[c] = Dn;
            word[c >> 3] |= srcK | (tD ? 4u * K : 0u);
        });
        // column 0 on row i: M = NEG, I' = NEG + goh; open_B[0] + goh = goh
        int iL = NEG + goh, mL = NEG;
        static_for(std::make_integer_sequence<int, TINY_LB>{}, [&](auto ic) {
            constexpr int c = decltype(ic)::value;
            constexpr uint32_t K = 1u << (4 * (c & 7));
            const int mi = mL + (c == 0 ? goh : go2);
            const bool tI = mi >= iL;
            iL = max(mi, iL);
            Ip[c] = iL;
            mL = Mp[c];
            word[c >> 3] |= tI ? 8u * K : 0u;
        });
        *reinterpret_cast<uint2 *>(tbj + (size_t)(i - 1) * 8) = make_uint2(word[0], word[1]);
        if (i == la) {
#pragma unroll
            for (int c = 0; c < TINY_LB; ++c)
                if (c == lb - 1) { finM = Mp[c]; finD = Dp[c]; finI = Ip[c]; }
        }
    }
    finD -= goh;
    finI -= goh;
    int score = finM, edge = 'M';
    if (finD > score) { score = finD; edge = 'D'; }
    if (finI > score) { score = finI; edge = 'I'; }
    out->score = score;
    out->edge = edge;
}

__device__ __forceinline__ int xchar(uint32_t bits, int t)
{
    if (t == 'M') { const uint32_t v = bits & 3u; return v == 0 ? 'M' : (v == 1 ? 'D' : (v == 2 ? 'I' : '?')); }
    if (t == 'D') return (bits & 4u) ? 'M' : 'D';
    return (bits & 8u) ? 'M' : 'I';
}

// BitTraceBack, MU/bittraceback.cpp:138-209: one thread per job walks the
// trace bits from (LA, LB) and writes the edges backwards into the job's
// path slot.  It also records the path's largest distance from the
// corner-to-corner line, ceil(max |i*LB - j*LA| / LA) over the visited
// positions, which is the band label and the edge test of a banded job.
__device__ __forceinline__ void traceback_thread_job(const DevJob *__restrict__ jobs,
                                                     const DpOut *__restrict__ outs,
                                                     const uint8_t *__restrict__ tb, uint8_t *__restrict__ paths,
                                                     mh_dp_result *__restrict__ res, uint32_t *__restrict__ dev,
                                                     uint32_t x)
{
    const DevJob jb = jobs[x];
    const int la = (int)jb.la, lb = (int)jb.lb;
    // stripe layouts: G lanes per job (16 for the row-packed jobs) of cols columns
    const bool rows16 = jb.cols > GROUP_TAG && jb.cols < WINDOW_TAG;
    const int G = rows16 ? 16 : 64;
    const int cols = rows16 ? (int)(jb.cols - GROUP_TAG) : (int)jb.cols, sw = G * cols, lane_bytes = cols / 2;
    const uint8_t *tj = tb + jb.tb_off;
    uint8_t *slot = paths + jb.path_off;
    // trace block of the last stripe: the sum over the stripes before it
    int cur = (lb - 1) / sw, r0 = 1, r1 = 0;
    uint64_t base = 0;
    for (int s = 0; s < cur; ++s) {
        int a0, a1;
        stripe_rows(jb.la, jb.lb, jb.band, sw, s, a0, a1);
        if (a1 >= a0) base += (uint64_t)(a1 - a0 + G) * G * lane_bytes;
    }
    stripe_rows(jb.la, jb.lb, jb.band, sw, cur, r0, r1);
    // stripe layout: column j-1 = cur*sw + k*cols + c, tracked without
    // divisions; dnum = i*lb - j*la, tracked as the walk moves
    const bool stripes = jb.cols != TINY_COLS_TAG && jb.cols <= WINDOW_TAG;
    int k = stripes ? ((lb - 1) % sw) / cols : 0, c = stripes ? (lb - 1) % cols : 0;
    int64_t dnum = 0;
    int pos = la + lb;
    int i = la, j = lb;
    int type = outs[x].edge;
    bool ok = true;
    int64_t maxdev = 0;
    for (;;) {
        if (stripes && i > 0 && j > 0) {
            // Look ahead along the run that keeps the walk's state (M along the
            // diagonal, D up, I along the row): the bytes of up to TB_AHEAD of
            // its cells in this stripe are loaded at once, and the walk moves
            // over all but the last cell that keeps the state (the step below
            // takes that one), instead of one dependent load per cell.
            constexpr int TB_AHEAD = 8;
            uint32_t by[TB_AHEAD], sh[TB_AHEAD];
            int cnt = 0, kl = k, cc = c;
            bool go = true;
#pragma unroll
            for (int d = 0; d < TB_AHEAD; ++d) {
                const int ci = i - (type == 'I' ? 0 : d);
                if (d > 0 && type != 'D' && --cc < 0) {
                    cc = cols - 1;
                    if (--kl < 0) go = false;
                }
                if (ci < r0 || ci > r1 || ci < 1) go = false;
                by[d] = 0;
                sh[d] = 4u * (uint32_t)(cc & 1);
                if (go) {
                    by[d] = tj[base + ((uint64_t)(ci - r0 + kl) * G + kl) * lane_bytes + (cc >> 1)];
                    cnt = d + 1;
                }
            }
            // stay mask over the loaded cells; n = the leading run of ones
            const uint32_t bit = type == 'M' ? 3u : (type == 'D' ? 4u : 8u);
            uint32_t mask = 0;
#pragma unroll
            for (int d = 0; d < TB_AHEAD; ++d)
                if (d < cnt && ((by[d] >> sh[d]) & bit) == 0u) mask |= 1u << d;
            const int n = (int)__builtin_ctz(~mask);
            if (n >= 2) {
                const int adv = n - 1;
                for (int d = 0; d < adv; ++d) slot[pos - 1 - d] = (uint8_t)type;
                pos -= adv;
                const int64_t st = type == 'M' ? (int64_t)la - lb : (type == 'D' ? -(int64_t)lb : (int64_t)la);
                const int64_t d1 = dnum + st, d2 = dnum + (int64_t)adv * st;
                maxdev = max(maxdev, max(d1 < 0 ? -d1 : d1, d2 < 0 ? -d2 : d2));
                dnum = d2;
                if (type != 'I') i -= adv;
                if (type != 'D') {
                    j -= adv;
                    for (int d = 0; d < adv; ++d)
                        if (--c < 0) {
                            c = cols - 1;
                            --k;
                        }
                }
            }
        }
        slot[--pos] = (uint8_t)type;
        uint32_t nib = 0;
        if (i > 0 && j > 0) {
            if (jb.cols == TINY_COLS_TAG) {
                const uint8_t byte = tj[(uint64_t)(i - 1) * 8 + ((j - 1) >> 1)];
                nib = (byte >> (4 * ((j - 1) & 1))) & 15u;
            } else if (jb.cols > WINDOW_TAG) {
                const int C = (int)jb.cols - (int)WINDOW_TAG, rpd = 8 / C;
                const int xw = j - window_base(la, lb, 64 * C, i) - 1;
                if (xw < 0 || xw >= 64 * C) { ok = false; break; }
                const uint32_t w = reinterpret_cast<const uint32_t *>(tj)[(uint64_t)((i - 1) / rpd) * 64 + xw / C];
                nib = (w >> (4 * (C * ((i - 1) % rpd) + xw % C))) & 15u;
            } else {
                if (i < r0 || i > r1) { ok = false; break; }
                const uint8_t byte = tj[base + ((uint64_t)(i - r0 + k) * G + k) * lane_bytes + (c >> 1)];
                nib = (byte >> (4 * (c & 1))) & 15u;
            }
        }
        const int nx = xchar(nib, type);
        bool left = false;
        if (type == 'M') { if (i == 0 || j == 0) { ok = false; break; } --i; --j; dnum += (int64_t)la - lb; left = true; }
        else if (type == 'D') { if (i == 0) { ok = false; break; } --i; dnum -= lb; }
        else { if (j == 0) { ok = false; break; } --j; dnum += la; left = true; }
        if (stripes && left && j > 0 && --c < 0) {
            c = cols - 1;
            if (--k < 0) {   // into the previous stripe: step back over its trace block
                k = G - 1;
                --cur;
                stripe_rows(jb.la, jb.lb, jb.band, sw, cur, r0, r1);
                if (r1 >= r0) base -= (uint64_t)(r1 - r0 + G) * G * lane_bytes;
            }
        }
        maxdev = max(maxdev, dnum < 0 ? -dnum : dnum);
        if (i == 0 && j == 0) break;
        if (pos == 0 || nx == '?') { ok = false; break; }
        type = nx;
    }
    mh_dp_result r;
    r.score = outs[x].score;
    r.path_len = ok ? (uint32_t)(la + lb - pos) : 0u;
    r.path_start = jb.path_off + (uint64_t)pos;
    res[jb.user] = r;
    if (dev) dev[jb.user] = ok ? (uint32_t)((maxdev + la - 1) / la) : 0xFFFFFFFFu;
}

// The same trace-back for stripe-layout jobs by one wavefront per job.  The
// walk itself is sequential (all lanes run it in step), but its trace bytes
// come from LDS: the wave loads a tile of 128 steps x 128 columns of the
// current stripe (128 segments of 64 contiguous bytes, 16-byte aligned) in
// one round of loads and refills it only when the walk leaves it, which
// replaces a dependent global load per step with one per ~100 steps.
constexpr int TBC_STEPS = 128, TBC_SEG = 64;
constexpr uint32_t WAVE_WALK_MIN = 1024;   // a + b above which a job's walk gets a wavefront

__device__ __forceinline__ void traceback_wave_job(const DevJob *__restrict__ jobs, const DpOut *__restrict__ outs,
                                                   const uint8_t *__restrict__ tb, uint8_t *__restrict__ paths,
                                                   mh_dp_result *__restrict__ res, uint32_t *__restrict__ dev,
                                                   uint32_t x, uint8_t *__restrict__ cache)
{
    const int lane = threadIdx.x;
    const DevJob jb = jobs[x];
    const int la = (int)jb.la, lb = (int)jb.lb;
    // G lanes a step: 64, or 16 for the row-packed layout
    const bool rows16 = jb.cols > GROUP_TAG;
    const int G = rows16 ? 16 : 64;
    const int cols = rows16 ? (int)(jb.cols - GROUP_TAG) : (int)jb.cols, sw = G * cols, lane_bytes = cols / 2;
    const int K = TBC_SEG / lane_bytes < G ? TBC_SEG / lane_bytes : G;   // lanes a cached step holds
    const int P = K * lane_bytes / 16;                                   // 16-byte parts a cached step
    const int kalign = 16 / lane_bytes;      // segment starts on a 16-byte boundary
    const uint8_t *tj = tb + jb.tb_off;
    uint8_t *slot = paths + jb.path_off;
    int cur = (lb - 1) / sw, r0 = 1, r1 = 0;
    uint64_t base = 0;
    for (int s = 0; s < cur; ++s) {
        int a0, a1;
        stripe_rows(jb.la, jb.lb, jb.band, sw, s, a0, a1);
        if (a1 >= a0) base += (uint64_t)(a1 - a0 + G) * G * lane_bytes;
    }
    stripe_rows(jb.la, jb.lb, jb.band, sw, cur, r0, r1);
    int c_s = -1, c_tlo = 0, c_thi = -1, c_klo = 0;   // cached stripe, steps and first lane
    const int kshift = kalign == 2 ? 1 : (kalign == 4 ? 2 : 3);
    // column j-1 = cur*sw + k*cols + c, tracked without divisions; dnum =
    // i*lb - j*la, tracked as the walk moves
    int k = ((lb - 1) % sw) / cols, c = (lb - 1) % cols;
    int64_t dnum = 0;
    int pos = la + lb;
    int i = la, j = lb;
    int type = outs[x].edge;
    bool ok = true;
    int64_t maxdev = 0;
    const int cshift = cols == 4 ? 2 : (cols == 8 ? 3 : 4);
    for (;;) {
        // A run of cells that keep the walk's state (M from M along the
        // diagonal, D from D up the column, I from I along the row): lane l
        // reads the l-th cell of the run from the cached tile, and the walk
        // moves over all but the last good cell at once (the step below takes
        // that one, so the run never leaves the stripe or the tile).
        if (i > 0 && j > 0 && cur == c_s) {
            const int ci = i - (type == 'I' ? 0 : lane);
            const int off = k * cols + c - (type == 'D' ? 0 : lane);
            bool good = ci >= r0 && ci <= r1 && ci >= 1 && off >= 0;
            if (good) {
                const int kl = off >> cshift, cl = off & (cols - 1);
                const int t = ci - r0 + kl;
                good = t >= c_tlo && t <= c_thi && kl >= c_klo && kl < c_klo + K;
                if (good) {
                    const uint32_t nb =
                        ((uint32_t)cache[(t - c_tlo) * TBC_SEG + (kl - c_klo) * lane_bytes + (cl >> 1)] >>
                         (4 * (cl & 1))) & 15u;
                    good = type == 'M' ? (nb & 3u) == 0u : (type == 'D' ? (nb & 4u) == 0u : (nb & 8u) == 0u);
                }
            }
            const uint64_t bad = __ballot(!good);
            const int n = bad ? (int)__builtin_ctzll(bad) : 64;
            if (n >= 2) {
                const int adv = n - 1;
                if (lane < adv) slot[pos - 1 - lane] = (uint8_t)type;
                pos -= adv;
                const int64_t st = type == 'M' ? (int64_t)la - lb : (type == 'D' ? -(int64_t)lb : (int64_t)la);
                const int64_t d1 = dnum + st, d2 = dnum + (int64_t)adv * st;
                maxdev = max(maxdev, max(d1 < 0 ? -d1 : d1, d2 < 0 ? -d2 : d2));
                dnum = d2;
                if (type != 'I') i -= adv;
                if (type != 'D') {
                    j -= adv;
                    const int o = k * cols + c - adv;
                    k = o >> cshift;
                    c = o & (cols - 1);
                }
            }
        }
        --pos;
        if (lane == 0) slot[pos] = (uint8_t)type;
        uint32_t nib = 0;
        if (i > 0 && j > 0) {
            if (i < r0 || i > r1) { ok = false; break; }
            const int t = i - r0 + k;
            if (cur != c_s || t < c_tlo || t > c_thi || k < c_klo || k >= c_klo + K) {
                c_s = cur;
                c_thi = t;
                c_tlo = t - TBC_STEPS + 1 < 0 ? 0 : t - TBC_STEPS + 1;
                const int kl = (((k >> kshift) + 1) << kshift) - K;
                c_klo = kl < 0 ? 0 : kl;
                const int nseg = c_thi - c_tlo + 1;
                __syncthreads();
                for (int e = lane; e < nseg * P; e += 64) {
                    const int seg = e / P, part = e - seg * P;
                    const uint8_t *src = tj + base + ((uint64_t)(c_tlo + seg) * G + c_klo) * lane_bytes + part * 16;
                    *reinterpret_cast<uint4 *>(cache + seg * TBC_SEG + part * 16) = *reinterpret_cast<const uint4 *>(src);
                }
                __syncthreads();
            }
            // one LDS byte read by every lane; made wave-uniform so the walk's
            // bookkeeping runs on the scalar unit
            const uint32_t byte = (uint32_t)__builtin_amdgcn_readfirstlane(
                (int)cache[(t - c_tlo) * TBC_SEG + (k - c_klo) * lane_bytes + (c >> 1)]);
            nib = (byte >> (4 * (c & 1))) & 15u;
        }
        const int nx = xchar(nib, type);
        bool left = false;
        if (type == 'M') { if (i == 0 || j == 0) { ok = false; break; } --i; --j; dnum += (int64_t)la - lb; left = true; }
        else if (type == 'D') { if (i == 0) { ok = false; break; } --i; dnum -= lb; }
        else { if (j == 0) { ok = false; break; } --j; dnum += la; left = true; }
        if (left && j > 0 && --c < 0) {
            c = cols - 1;
            if (--k < 0) {   // into the previous stripe: step back over its trace block
                k = G - 1;
                --cur;
                stripe_rows(jb.la, jb.lb, jb.band, sw, cur, r0, r1);
                if (r1 >= r0) base -= (uint64_t)(r1 - r0 + G) * G * lane_bytes;
            }
        }
        maxdev = max(maxdev, dnum < 0 ? -dnum : dnum);
        if (i == 0 && j == 0) break;
        if (pos == 0 || nx == '?') { ok = false; break; }
        type = nx;
    }
    if (lane == 0) {
        mh_dp_result r;
        r.score = outs[x].score;
        r.path_len = ok ? (uint32_t)(la + lb - pos) : 0u;
        r.path_start = jb.path_off + (uint64_t)pos;
        res[jb.user] = r;
        if (dev) dev[jb.user] = ok ? (uint32_t)((maxdev + la - 1) / la) : 0xFFFFFFFFu;
    }
}

// ---- the round's launches ------------------------------------------------

__global__ __launch_bounds__(64) void dp_queue_reset(uint32_t *__restrict__ cursor)
{
    for (int k = threadIdx.x; k < DPQ_CURSORS; k += 64) cursor[k] = 0u;
}

// The forward kernels' arguments: every pointer restrict-qualified (so that
// job records and the bin table are read with scalar loads), the bin table in
// the round's descriptor.
#define DP_FWD_PARAMS                                                                                            \
    const DevJob *__restrict__ jobs, DpOut *__restrict__ outs, const uint8_t *__restrict__ A,                      \
        const uint8_t *__restrict__ B, uint8_t *__restrict__ tb, int *__restrict__ bnd,                          \
        const DpRound *__restrict__ rd, uint32_t *__restrict__ cursor, uint32_t sr0, uint32_t sr1, uint32_t sr2, \
        uint32_t sr3, int goh
#define DP_FWD_ARGS jobs, outs, A, B, tb, bnd, rd, cursor, sr0, sr1, sr2, sr3, goh

template <int MAXC>
__device__ __forceinline__ void run_item(DP_FWD_PARAMS, uint32_t kind, uint32_t first, uint32_t count,
                                         uint32_t item, int (*lds_in)[64], int (*lds_out)[64])
{
    if (kind >= QK_PIPE) {
        const uint32_t P = kind - QK_PIPE + 2;
        const uint32_t x = first + item / P;
        const int p0 = (int)(item % P);
        const DevJob jb = jobs[x];
        uint32_t *err = cursor + (DPQ_CURSORS - 1);
        if constexpr (MAXC >= 16) {
            if (jb.cols == 16) {
                dp_job<16, 64, true>(jb, A, B, tb, bnd, &outs[x], sr0, sr1, sr2, sr3, goh, lds_in, lds_out, p0, (int)P,
                                     err);
                return;
            }
        }
        if (jb.cols == 8)
            dp_job<8, 64, true>(jb, A, B, tb, bnd, &outs[x], sr0, sr1, sr2, sr3, goh, lds_in, lds_out, p0, (int)P,
                                err);
        else
            dp_job<4, 64, true>(jb, A, B, tb, bnd, &outs[x], sr0, sr1, sr2, sr3, goh, lds_in, lds_out, p0, (int)P,
                                err);
    } else if (kind == QK_STRIPE) {
        const uint32_t x = first + item;
        const DevJob jb = jobs[x];
        if constexpr (MAXC >= 16) {
            if (jb.cols == 16) {
                dp_job<16, 64>(jb, A, B, tb, bnd, &outs[x], sr0, sr1, sr2, sr3, goh, lds_in, lds_out);
                return;
            }
        }
        if (jb.cols == 8)
            dp_job<8, 64>(jb, A, B, tb, bnd, &outs[x], sr0, sr1, sr2, sr3, goh, lds_in, lds_out);
        else
            dp_job<4, 64>(jb, A, B, tb, bnd, &outs[x], sr0, sr1, sr2, sr3, goh, lds_in, lds_out);
    } else if (kind == QK_WINDOW) {
        const uint32_t x = first + item;
        const DevJob jb = jobs[x];
        switch (jb.cols - WINDOW_TAG) {
        case 1: dp_window_job<1>(jb, A, B, tb, &outs[x], sr0, sr1, sr2, sr3, goh); break;
        case 2: dp_window_job<2>(jb, A, B, tb, &outs[x], sr0, sr1, sr2, sr3, goh); break;
        case 4: dp_window_job<4>(jb, A, B, tb, &outs[x], sr0, sr1, sr2, sr3, goh); break;
        default: dp_window_job<8>(jb, A, B, tb, &outs[x], sr0, sr1, sr2, sr3, goh); break;
        }
    } else if (kind == QK_TINY) {
        const uint32_t x = first + item * 64 + threadIdx.x;
        if (x < first + count) tiny_job(jobs[x], A, B, tb, &outs[x], sr0, sr1, sr2, sr3, goh);
    } else {
        // four jobs of one shape, one per 16-lane row; a row past the bin's end sits out
        const uint32_t x = first + item * 4 + (threadIdx.x >> 4);
        if (x >= first + count) return;
        const DevJob jb = jobs[x];
        DpOut *o = &outs[x];
        switch (kind - QK_ROWS) {
        case 0: if constexpr (MAXC >= 16) dp_job<16, 16>(jb, A, B, tb, bnd, o, sr0, sr1, sr2, sr3, goh, lds_in, lds_out); break;
        case 1: if constexpr (MAXC >= 14) dp_job<14, 16>(jb, A, B, tb, bnd, o, sr0, sr1, sr2, sr3, goh, lds_in, lds_out); break;
        case 2: if constexpr (MAXC >= 12) dp_job<12, 16>(jb, A, B, tb, bnd, o, sr0, sr1, sr2, sr3, goh, lds_in, lds_out); break;
        case 3: if constexpr (MAXC >= 10) dp_job<10, 16>(jb, A, B, tb, bnd, o, sr0, sr1, sr2, sr3, goh, lds_in, lds_out); break;
        case 4: dp_job<8, 16>(jb, A, B, tb, bnd, o, sr0, sr1, sr2, sr3, goh, lds_in, lds_out); break;
        case 5: dp_job<6, 16>(jb, A, B, tb, bnd, o, sr0, sr1, sr2, sr3, goh, lds_in, lds_out); break;
        default: dp_job<4, 16>(jb, A, B, tb, bnd, o, sr0, sr1, sr2, sr3, goh, lds_in, lds_out); break;
        }
    }
}

// The forward launch.  MAXC is the widest shape it holds (16: every shape;
// 8: stripes and row shapes of at most 8 columns a lane, which need fewer
// registers and so fit more wavefronts a SIMD).  PERSIST: a grid of resident
// wavefronts, each taking the next item from the bins' atomic cursors until
// the last bin is empty.  Otherwise one workgroup per item in queue order
// (workgroup w takes item w; workgroups start in that order), which needs no
// cursor.  The host picks the variant a round's shapes allow, by the tuning
// table.
template <int MAXC, int WPE, bool PERSIST>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void dp_fwd(DP_FWD_PARAMS)
{
    __shared__ int lds_in[4][64];
    __shared__ int lds_out[3][64];
    const uint32_t nb = rd->nbins;
    if constexpr (!PERSIST) {
        const uint32_t w = blockIdx.x;
        uint32_t b = 0;
        while (b + 1 < nb && w >= rd->item0[b + 1]) ++b;
        if (b >= nb || w >= rd->item0[nb]) return;
        uint32_t item = w - rd->item0[b];
        if (rd->kind[b] >= QK_PIPE) {
            // a ticket, so that the parts of a job are taken in order by
            // running wavefronts (a part waits only on parts taken before it)
            uint32_t v = 0;
            if (threadIdx.x == 0) v = atomicAdd(&cursor[b], 1u);
            item = (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
        }
        run_item<MAXC>(DP_FWD_ARGS, rd->kind[b], rd->first[b], rd->count[b], item, lds_in, lds_out);
    } else {
        uint32_t b = 0;
        for (;;) {
            // the next item: the first bin whose cursor is still inside it
            uint32_t item = 0;
            for (; b < nb; ++b) {
                uint32_t v = 0;
                if (threadIdx.x == 0) v = atomicAdd(&cursor[b], 1u);
                item = (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
                if ((uint64_t)item * qk_per_item(rd->kind[b]) < rd->count[b]) break;
            }
            if (b >= nb) return;
            run_item<MAXC>(DP_FWD_ARGS, rd->kind[b], rd->first[b], rd->count[b], item, lds_in, lds_out);
        }
    }
}

// Trace-backs: the jobs with long walks a wavefront each (workgroup x walks
// job x), the others a thread each.  Their pointers come as arguments (not
// from the round's descriptor): with every pointer restrict-qualified the
// walk's job record is read with scalar loads and its bookkeeping stays on the
// scalar unit.
__global__ __launch_bounds__(64) void dp_tb_wave(const DevJob *__restrict__ jobs, const DpOut *__restrict__ outs,
                                                 const uint8_t *__restrict__ tb, uint8_t *__restrict__ paths,
                                                 mh_dp_result *__restrict__ res, uint32_t *__restrict__ dev,
                                                 const uint32_t *__restrict__ idx, uint32_t nwave)
{
    __shared__ __align__(16) uint8_t cache[TBC_STEPS * TBC_SEG];
    if (blockIdx.x >= nwave) return;
    traceback_wave_job(jobs, outs, tb, paths, res, dev, idx[blockIdx.x], cache);
}

__global__ __launch_bounds__(64) void dp_tb_thread(const DevJob *__restrict__ jobs, const DpOut *__restrict__ outs,
                                                   const uint8_t *__restrict__ tb, uint8_t *__restrict__ paths,
                                                   mh_dp_result *__restrict__ res, uint32_t *__restrict__ dev,
                                                   const uint32_t *__restrict__ idx, uint32_t first, uint32_t n)
{
    const uint32_t t = first + blockIdx.x * 64 + threadIdx.x;
    if (t >= n) return;
    traceback_thread_job(jobs, outs, tb, paths, res, dev, idx[t]);
}

// NUC_SP (MU/nucmx.cpp:8-25) rows packed as biased bytes: row a, column b at
// bits [8b, 8b+8) holding NUC_SP[a][b] + 65.
void packed_rows(uint32_t r[4])
{
    static const int hoxd[4][4] = {{91, -114, -31, -123},
                                   {-114, 100, -125, -31},
                                   {-31, -125, 100, -114},
                                   {-123, -31, -114, 91}};
    for (int a = 0; a < 4; ++a) {
        r[a] = 0;
        for (int b = 0; b < 4; ++b) r[a] |= (uint32_t)(hoxd[a][b] + 60 + SBIAS) << (8 * b);
    }
}

}  // namespace

namespace {

struct DpItem {
    uint32_t job;     // index in the caller's job array
    uint32_t band;    // 0 = full matrix
    uint32_t cols;
    uint64_t tb_bytes;
};

// Relative cost per computed cell of the stripe shapes, used to pick a banded
// job's shape (which decides how much of each stripe lies outside the band).
// Per shape alone (one class of B's width at a time, MI355X) 64-lane stripes
// of 16 columns evaluate cells fastest and 16-lane stripes slowest (2-4x),
// but a round's shapes share one forward kernel, and weights taken from those
// runs pick more 16-column stripes, whose wide kernel then slows the round's
// other jobs.  These weights were chosen by timing whole banded runs (learned
// and constant bands, edge fallback included) over 100k training gaps: the
// fastest learned-band run of the sets tried.
constexpr double COST_PER_CELL[3] = {1.34, 1.00, 1.30};   // 64-lane stripes of cols 4, 8, 16
constexpr double COST_PER_CELL_G16[2] = {2.40, 1.50};      // 16-lane stripes of cols 4, 8


}  // namespace

// Scratch of the no-allocation entry point (mh_dp_align_device_noalloc): the
// round's job table, stripe boundary columns and trace arena, in one array
// that every GPU the library runs on holds until the process ends.  Host code
// never uses the array's name: each device's address comes from
// hipGetSymbolAddress, resolved once per device by dp_scratch_init.
constexpr uint64_t DP_SCRATCH_BYTES = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[DP_SCRATCH_BYTES];

namespace {
constexpr int DP_SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[DP_SCRATCH_MAX_DEVICES] = {};
}  // namespace

int dp_scratch_init(int device)
{
    MH_CHECK(device >= 0 && device < DP_SCRATCH_MAX_DEVICES, MH_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return MH_OK;
    int cur = -1;
    MH_HIP(hipGetDevice(&cur));
    MH_CHECK(cur == device, MH_ERR_INVALID, "dp_scratch_init: device is not current");
    void *ptr = nullptr;
    MH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MH_CHECK(ptr != nullptr, MH_ERR_DEVICE, "hipGetSymbolAddress gave a null address for the DP scratch");
    g_scratch_ptr[device] = static_cast<uint8_t *>(ptr);
    return MH_OK;
}

uint64_t dp_scratch_bytes() { return DP_SCRATCH_BYTES; }

// The banded shapes' weights: COST_PER_CELL, COST_PER_CELL_G16, or five
// comma-separated values in MH_DP_BAND_COST (read once per process, for sweeps)
static const double *band_cost_weights()
{
    static double w[5];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int q = 0; q < 5; ++q) w[q] = q < 3 ? COST_PER_CELL[q] : COST_PER_CELL_G16[q - 3];
        const char *e = getenv("MH_DP_BAND_COST");
        if (!e) return;
        double v[5];
        int n = 0;
        for (const char *p = e; *p && n < 5;) {
            char *end = nullptr;
            const double x = strtod(p, &end);
            if (end == p || !(x > 0.0)) return;
            v[n++] = x;
            p = *end == ',' ? end + 1 : end;
        }
        if (n == 5) std::memcpy(w, v, sizeof(v));
    });
    return w;
}

// MH_DP_COLS forces one shape: 4, 8 or 16 (64-lane stripes of that many
// columns a lane) or 36, 40 (16-lane stripes of 4, 8 columns, four jobs a
// wavefront)
int dp_force_cols()
{
    if (const char *e = getenv("MH_DP_COLS")) {
        const int c = atoi(e);
        if (c == 4 || c == 8 || c == 16 || c == (int)GROUP_TAG + 4 || c == (int)GROUP_TAG + 8) return c;
    }
    return 0;
}

// class of B's width in the tuning table
static int dp_tune_class(uint32_t lb)
{
    int c = 0;
    while (c < DP_TUNE_NCLASS - 1 && lb > DP_TUNE_LB_CLASS[c]) ++c;
    return c;
}

// The full DP's shape per class: the tuning table, or MH_DP_SHAPES (a comma
// list of DP_TUNE_NCLASS values, each 0, 4, 8 or 16; read once per process,
// for the tuner's sweeps)
static const uint8_t *shape_table()
{
    static uint8_t t[DP_TUNE_NCLASS];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int c = 0; c < DP_TUNE_NCLASS; ++c) t[c] = DP_TUNE_SHAPE[c];
        const char *e = getenv("MH_DP_SHAPES");
        if (!e) return;
        uint8_t v[DP_TUNE_NCLASS];
        int c = 0;
        for (const char *p = e; *p && c < DP_TUNE_NCLASS;) {
            char *end = nullptr;
            const long x = strtol(p, &end, 10);
            if (end == p || !(x == 0 || x == 4 || x == 8 || x == 16)) return;
            v[c++] = (uint8_t)x;
            p = *end == ',' ? end + 1 : end;
        }
        if (c == DP_TUNE_NCLASS) std::memcpy(t, v, sizeof(v));
    });
    return t;
}

// Stripe shape of a job: the full DP takes the shape the tuning table names
// for B's width (stripes of 8 columns a lane where it names the row-packed
// shape and the job cannot take it); a banded job takes the shape with the
// lowest estimated cost: ceil(lb/w) stripes of min(la, la*(w+2W)/lb + 1) +
// g - 1 steps of w cells (g lanes), times the shape's cost per cell.  force
// (any MH_DP_COLS shape) overrides.  A rerun pass holds few jobs, whose time
// is the longest job's: its banded jobs take 64-lane stripes only
// (wide_only), since a 16-lane stripe runs one job on a quarter of a
// wavefront.
uint32_t dp_cols_for(uint32_t la, uint32_t lb, uint32_t band, int force, bool wide_only)
{
    if (force) return (uint32_t)force;
    if (band == 0) {
        const uint32_t t = shape_table()[dp_tune_class(lb)];
        return t ? t : 8u;
    }
    // 64-lane stripes of 4, 8 or 16 columns a lane, and (unless
    // MH_DP_BAND_G16=0) 16-lane stripes of 4 or 8 columns a lane, four jobs a
    // wavefront: narrower stripes, whose rows follow a narrow band closer
    uint32_t best = 16;
    double best_cost = 0;
    const char *g16e = getenv("MH_DP_BAND_G16");
    const int nshape = wide_only || (g16e && g16e[0] == '0') ? 3 : 5;
    const double *wt = band_cost_weights();
    for (int q = 0; q < nshape; ++q) {
        const uint32_t g = q < 3 ? 64u : 16u;
        const uint32_t c = q < 3 ? 4u << q : 4u << (q - 3), sw = g * c;
        const double ns = (double)((lb + sw - 1) / sw);
        double rows = (double)la * (sw + 2.0 * band) / lb + 1.0;
        if (rows > la) rows = la;
        const double cost = ns * (rows + (g - 1.0)) * sw * wt[q];
        if (q == 0 || cost < best_cost) {
            best = g == 64 ? c : GROUP_TAG + c;
            best_cost = cost;
        }
    }
    return best;
}

// The sliding window spends much of each row on the shift and the cross-lane
// scan: measured on MI355X over 100k learned bands it evaluates its cells
// about 1.45x slower than the stripes evaluate theirs and saves fewer cells
// than that, so the scheduler uses it only when MH_DP_BAND_KIND=window asks.
int dp_force_kind()
{
    if (const char *e = getenv("MH_DP_BAND_KIND")) {
        if (!strcmp(e, "stripes")) return 1;
        if (!strcmp(e, "window")) return 2;
    }
    return 0;
}

// Columns per lane of the sliding window for a banded region (0: use
// stripes): when the window is asked for (kind 2), the narrowest window
// reaching the region band on each side of the line (64C >= 2R+2), if it is
// narrower than B and the slope lb/la is at most 4.
uint32_t dp_window_cols(uint32_t la, uint32_t lb, uint32_t band, int kind)
{
    if (band == 0 || kind != 2 || (uint64_t)lb > 4ull * la) return 0;
    int q = 0;
    while (q < 4 && 64u * (1u << q) < 2u * band + 2u) ++q;
    if (q == 4) return 0;
    const uint32_t c = 1u << q;
    if (64u * c >= lb) return 0;
    return c;
}

uint32_t dp_stripe_width(uint32_t cols) { return cols > GROUP_TAG ? 16 * (cols - GROUP_TAG) : 64 * cols; }

// A band whose region is the whole matrix is the full DP.
uint32_t dp_effective_band(uint32_t la, uint32_t lb, uint32_t band, uint32_t cols)
{
    if (band == 0) return 0;
    const uint32_t sw = cols > GROUP_TAG ? 16 * (cols - GROUP_TAG) : 64 * cols;
    const uint32_t ns = (lb + sw - 1) / sw;
    int f0, f1, l0, l1;
    stripe_rows(la, lb, band, sw, 0, f0, f1);
    stripe_rows(la, lb, band, sw, ns - 1, l0, l1);
    return (l0 == 1 && f1 == (int)la) ? 0u : band;
}

static inline uint64_t align256(uint64_t v) { return (v + 255) & ~255ull; }

// device bytes of a round's descriptor and its queue cursors
constexpr uint64_t DP_ROUND_BYTES = (sizeof(DpRound) + DPQ_CURSORS * 4 + 255) & ~255ull;

// The forward kernel's variants: the widest shape each holds, its
// wavefronts a SIMD, and whether its wavefronts persist.  Rounds with a shape
// wider than 8 columns a lane run the tuning table's wide variant, the others
// its narrow one; MH_DP_VARIANT forces one (for sweeps), where the round's
// shapes allow it.
struct FwdVariant {
    const void *fn;
    int maxc, wpe;
    bool persist;
};
const FwdVariant FWD_VARIANTS[] = {
    {reinterpret_cast<const void *>(&dp_fwd<16, 3, false>), 16, 3, false},
    {reinterpret_cast<const void *>(&dp_fwd<8, 4, false>), 8, 4, false},
    {reinterpret_cast<const void *>(&dp_fwd<16, 2, false>), 16, 2, false},
    {reinterpret_cast<const void *>(&dp_fwd<16, 3, true>), 16, 3, true},
    {reinterpret_cast<const void *>(&dp_fwd<16, 2, true>), 16, 2, true},
    {reinterpret_cast<const void *>(&dp_fwd<8, 4, true>), 8, 4, true},
    {reinterpret_cast<const void *>(&dp_fwd<16, 4, false>), 16, 4, false},
};
constexpr int N_FWD_VARIANTS = (int)(sizeof(FWD_VARIANTS) / sizeof(FWD_VARIANTS[0]));
static_assert(N_FWD_VARIANTS <= MH_DP_MAX_VARIANTS, "context graph slots");
static_assert(DP_TUNE_WIDE_VARIANT < N_FWD_VARIANTS && DP_TUNE_NARROW_VARIANT < N_FWD_VARIANTS, "tuning table");

static int dp_variant_for(bool wide)
{
    for (const char *name : {"MH_DP_VARIANT", wide ? "MH_DP_VARIANT_WIDE" : "MH_DP_VARIANT_NARROW"})
        if (const char *e = getenv(name)) {
            const int v = atoi(e);
            if (v >= 0 && v < N_FWD_VARIANTS && (!wide || FWD_VARIANTS[v].maxc >= 16)) return v;
        }
    return wide ? DP_TUNE_WIDE_VARIANT : DP_TUNE_NARROW_VARIANT;
}

// Graph replay of the round's launches (MH_DP_GRAPH=0 launches them directly)
static bool dp_use_graph()
{
    const char *e = getenv("MH_DP_GRAPH");
    return !(e && e[0] == '0');
}

static int dp_round_staging(mh_context *ctx)
{
    if (ctx->dp_round_host) return MH_OK;
    if (hipHostMalloc(&ctx->dp_round_host, sizeof(DpRound), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        ctx->dp_round_host = nullptr;
        set_error("pinned host allocation failed");
        return MH_ERR_NOMEM;
    }
    return MH_OK;
}

// resident workgroups of a 64-thread kernel over the whole device
static int resident_grid(mh_context *ctx, const void *fn, unsigned *grid)
{
    int cus = 0, per = 0;
    MH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    MH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, 64, 0));
    *grid = (unsigned)std::max(1, cus * std::max(1, per));
    return MH_OK;
}

// grid sizes of a round: forward (items, or the resident grid of a
// persistent variant), wavefront trace-backs, thread trace-backs
struct RoundGrids {
    unsigned fwd, tbw, tbt;
};

static int round_grids(mh_context *ctx, int v, uint64_t items, uint32_t nwave, uint32_t n, RoundGrids *g)
{
    if (FWD_VARIANTS[v].persist) {
        if (!ctx->dp_fwd_grid[v]) {
            unsigned r = 0;
            int rc;
            if ((rc = resident_grid(ctx, FWD_VARIANTS[v].fn, &r)) != MH_OK) return rc;
            ctx->dp_fwd_grid[v] = r;
        }
        g->fwd = ctx->dp_fwd_grid[v];
    } else {
        MH_CHECK(items < (1ull << 31), MH_ERR_INVALID, "too many DP work items in one round");
        g->fwd = (unsigned)std::max<uint64_t>(1, items);
    }
    g->tbw = std::max(1u, nwave);
    g->tbt = std::max(1u, (n - nwave + 63) / 64);
    return MH_OK;
}

// the forward kernels' arguments of a round (DP_FWD_PARAMS)
struct FwdArgs {
    const DevJob *jobs;
    DpOut *outs;
    const uint8_t *A, *B;
    uint8_t *tb;
    int *bnd;
    const DpRound *rd;
    uint32_t *cursor;
    uint32_t sr0, sr1, sr2, sr3;
    int goh;
    void *args[13];
    void set()
    {
        void *a[13] = {&jobs, &outs, &A, &B, &tb, &bnd, &rd, &cursor, &sr0, &sr1, &sr2, &sr3, &goh};
        std::memcpy(args, a, sizeof(a));
    }
};

static int launch_fwd(mh_context *ctx, FwdArgs &a, int v, unsigned grid)
{
    a.set();
    hipLaunchKernelGGL(dp_queue_reset, dim3(1), dim3(64), 0, ctx->stream, a.cursor);
    MH_HIP(hipLaunchKernel(FWD_VARIANTS[v].fn, dim3(grid), dim3(64), a.args, 0, ctx->stream));
    MH_HIP(hipGetLastError());
    return MH_OK;
}

// the trace-back kernels' arguments of a round
struct TbArgs {
    const DevJob *jobs;
    const DpOut *outs;
    const uint8_t *tb;
    uint8_t *paths;
    mh_dp_result *res;
    uint32_t *dev;
    const uint32_t *idx;
    uint32_t nwave, first, n;
    void *wave[8];
    void *thread[9];
    void set()
    {
        void *w[8] = {&jobs, &outs, &tb, &paths, &res, &dev, &idx, &nwave};
        void *t[9] = {&jobs, &outs, &tb, &paths, &res, &dev, &idx, &first, &n};
        std::memcpy(wave, w, sizeof(w));
        std::memcpy(thread, t, sizeof(t));
    }
};

static int launch_tb(mh_context *ctx, TbArgs &a, const RoundGrids &g)
{
    a.set();
    MH_HIP(hipLaunchKernel(reinterpret_cast<const void *>(&dp_tb_wave), dim3(g.tbw), dim3(64), a.wave, 0,
                           ctx->stream));
    MH_HIP(hipLaunchKernel(reinterpret_cast<const void *>(&dp_tb_thread), dim3(g.tbt), dim3(64), a.thread, 0,
                           ctx->stream));
    MH_HIP(hipGetLastError());
    return MH_OK;
}

// Captures f's launches on the context stream: the graph (kept, for its
// nodes), its executable form, and the kernel nodes of the given functions.
template <class F>
static int capture(mh_context *ctx, hipGraph_t *graph, hipGraphExec_t *exec, const void *const *fns,
                   hipGraphNode_t *nodes, int nfns, F f)
{
    MH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = f();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != MH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    MH_HIP(e);
    size_t nn = 0;
    MH_HIP(hipGraphGetNodes(g, nullptr, &nn));
    std::vector<hipGraphNode_t> all(nn);
    MH_HIP(hipGraphGetNodes(g, all.data(), &nn));
    for (int k = 0; k < nfns; ++k) nodes[k] = nullptr;
    for (hipGraphNode_t nd : all) {
        hipGraphNodeType t;
        MH_HIP(hipGraphNodeGetType(nd, &t));
        if (t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams kp;
        MH_HIP(hipGraphKernelNodeGetParams(nd, &kp));
        for (int k = 0; k < nfns; ++k)
            if (kp.func == fns[k]) nodes[k] = nd;
    }
    for (int k = 0; k < nfns; ++k)
        if (!nodes[k]) {
            (void)hipGraphDestroy(g);
            MH_CHECK(false, MH_ERR_DEVICE, "a captured DP launch has no kernel node");
        }
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        MH_HIP(ei);
    }
    *graph = g;
    return MH_OK;
}

// sets a captured kernel node's grid and arguments in the executable graph
static int set_node(hipGraphExec_t exec, hipGraphNode_t node, unsigned grid, void **args)
{
    hipKernelNodeParams kp;
    MH_HIP(hipGraphKernelNodeGetParams(node, &kp));
    kp.gridDim = dim3(grid);
    kp.kernelParams = args;
    kp.extra = nullptr;
    MH_HIP(hipGraphExecKernelNodeSetParams(exec, node, &kp));
    return MH_OK;
}

// The round's forward launch (queue reset + forward kernel) and its
// trace-backs, timed apart; from graphs captured once per context (the
// descriptor's address is fixed; a round sets the nodes' grids) or launched
// directly.
static int dp_launch_round(mh_context *ctx, const DpRound *d_round, int v, bool graph, const RoundGrids &g,
                           FwdArgs &fwa, TbArgs &tba)
{
    int rc;
    if (graph && d_round != ctx->dp_round.as<DpRound>()) graph = false;
    if (graph && !ctx->dp_graph_fwd[v]) {
        const void *fns[1] = {FWD_VARIANTS[v].fn};
        if ((rc = capture(ctx, &ctx->dp_graph_fwd_src[v], &ctx->dp_graph_fwd[v], fns, &ctx->dp_graph_fwd_node[v], 1,
                          [&] { return launch_fwd(ctx, fwa, v, g.fwd); })) != MH_OK)
            return rc;
    }
    if (graph && !ctx->dp_graph_tb) {
        const void *fns[2] = {reinterpret_cast<const void *>(&dp_tb_wave),
                              reinterpret_cast<const void *>(&dp_tb_thread)};
        if ((rc = capture(ctx, &ctx->dp_graph_tb_src, &ctx->dp_graph_tb, fns, ctx->dp_graph_tb_node, 2,
                          [&] { return launch_tb(ctx, tba, g); })) != MH_OK)
            return rc;
    }
    if (graph) {
        fwa.set();
        tba.set();
        if ((rc = set_node(ctx->dp_graph_fwd[v], ctx->dp_graph_fwd_node[v], g.fwd, fwa.args)) != MH_OK) return rc;
        if ((rc = set_node(ctx->dp_graph_tb, ctx->dp_graph_tb_node[0], g.tbw, tba.wave)) != MH_OK) return rc;
        if ((rc = set_node(ctx->dp_graph_tb, ctx->dp_graph_tb_node[1], g.tbt, tba.thread)) != MH_OK) return rc;
    }
    {
        TimedLaunch tl(ctx, "dp_forward");
        if (graph) {
            MH_HIP(hipGraphLaunch(ctx->dp_graph_fwd[v], ctx->stream));
        } else if ((rc = launch_fwd(ctx, fwa, v, g.fwd)) != MH_OK) {
            return rc;
        }
    }
    {
        TimedLaunch tl(ctx, "dp_traceback");
        if (graph) {
            MH_HIP(hipGraphLaunch(ctx->dp_graph_tb, ctx->stream));
        } else if ((rc = launch_tb(ctx, tba, g)) != MH_OK) {
            return rc;
        }
    }
    return MH_OK;
}

static int dp_run_items(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
                        const std::vector<uint64_t> &slot, std::vector<DpItem> &items, uint8_t *d_paths,
                        mh_dp_result *d_results, uint32_t *d_dev, uint64_t *cells,
                        const FixedScratch *fx = nullptr)
{
    const uint64_t nitems = items.size();
    if (nitems == 0) return MH_OK;
    // largest first so the long tails start early: a stable counting sort on
    // the trace size quantised to 1/16 octave (the order only balances load)
    const double ts0 = now_ms();
    {
        auto q = [](uint64_t v) -> uint32_t {
            if (v < 16) return (uint32_t)v;
            const int e = 63 - __builtin_clzll(v);
            return (uint32_t)(e * 16 + ((v >> (e - 4)) & 15u));
        };
        constexpr uint32_t NQ = 64 * 16 + 16;
        std::vector<uint32_t> cnt(NQ + 1, 0);
        for (const DpItem &it : items) ++cnt[NQ - 1 - q(it.tb_bytes)];
        uint32_t acc = 0;
        for (uint32_t &c : cnt) {
            const uint32_t t = c;
            c = acc;
            acc += t;
        }
        std::vector<DpItem> sorted(nitems);
        for (const DpItem &it : items) sorted[cnt[NQ - 1 - q(it.tb_bytes)]++] = it;
        items.swap(sorted);
    }
    host_time(ctx, "host_dp_sort", now_ms() - ts0);
    auto bnd_ints = [&](const DpItem &it) {
        // one-lane and window jobs hand no column over; a job of 16-lane
        // stripes does when it has more than one (banded jobs)
        if (it.cols == TINY_COLS_TAG || it.cols >= WINDOW_TAG) return 0ull;
        const uint64_t sw = it.cols > GROUP_TAG ? 16ull * (it.cols - GROUP_TAG) : 64ull * it.cols;
        const uint64_t ns = (jobs[it.job].b_len + sw - 1) / sw;
        return ns > 1 ? 3ull * ((jobs[it.job].a_len + 63) & ~63ull) : 0ull;
    };
    // footprint in a fixed scratch of a round of `cnt` jobs with trace bytes
    // tbb (each rounded up to 16) and bi boundary ints
    auto fixed_need = [&](uint64_t cnt, uint64_t tbb, uint64_t bi) {
        return DP_ROUND_BYTES + align256(cnt * (sizeof(DevJob) + sizeof(DpOut) + 4)) + align256(bi * 4) + tbb + 256;
    };
    if (fx) {
        for (const DpItem &it : items)
            MH_CHECK(fixed_need(1, (it.tb_bytes + 15) & ~15ull, bnd_ints(it)) <= fx->bytes, MH_ERR_INVALID,
                     "DP job " + std::to_string(it.job) + " (" + std::to_string(jobs[it.job].a_len) + " x " +
                         std::to_string(jobs[it.job].b_len) + ") needs more than the " +
                         std::to_string(fx->bytes >> 20) + " MiB scratch of the no-allocation entry point");
    }
    // trace arena budget: 70% of the memory that is free or already held by
    // this context's arena (at most 192 GB); once grown the arena is kept
    uint64_t budget = 0;
    if (!fx) {
        size_t free_b = 0, total_b = 0;
        MH_HIP(hipMemGetInfo(&free_b, &total_b));
        budget = (uint64_t)((free_b + ctx->dp_tb.bytes + ctx->dp_bnd.bytes) * 0.7);
        const uint64_t cap = 192ull << 30;
        if (budget > cap) budget = cap;
        if (const char *e = getenv("MH_DP_TB_BUDGET_GB")) budget = (uint64_t)(atof(e) * (1ull << 30));
    }
    if (!fx) {
        // size the arena once for the whole call: the largest round fits the budget
        uint64_t need = 0;
        for (const DpItem &it : items) need += it.tb_bytes + bnd_ints(it) * 4;
        if (need > budget) need = budget;
        int rc0;
        if (ctx->dp_tb.bytes < need) {
            ctx->dp_bnd.release();
            if ((rc0 = ctx->dp_tb.ensure(need + 256)) != MH_OK) return rc0;
        }
    }
    uint32_t sr[4];
    packed_rows(sr);
    // pipelined stripe jobs: wavefronts a job by the round's size (up to
    // DP_PIPE_MAX), or MH_DP_PIPE (0 or 1 keeps one wavefront a job); the
    // no-allocation path keeps one
    constexpr uint32_t DP_PIPE_MAX = 8;
    int pipe_env = -1;
    if (const char *e = getenv("MH_DP_PIPE")) pipe_env = std::max(0, std::min((int)DP_PIPE_MAX, atoi(e)));
    int ncu = 0;
    if (!fx && pipe_env != 0 && pipe_env != 1)
        MH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));

    uint64_t first = 0;
    std::vector<DpItem> shaped;
    std::vector<uint8_t> pj;   // wavefronts of each stripe job of a round
    while (first < nitems) {
        // one round: as many jobs (in schedule order) as fit the trace budget
        uint64_t tbb = 0, bi = 0, last = first;
        while (last < nitems) {
            const uint64_t w = items[last].tb_bytes, b = bnd_ints(items[last]);
            if (fx) {
                const uint64_t w16 = (w + 15) & ~15ull;
                if (last > first && fixed_need(last - first + 1, tbb + w16, bi + b) > fx->bytes) break;
                tbb += w16;
            } else {
                if (last > first && tbb + w + (bi + b) * 4 > budget) break;
                tbb += w;
            }
            bi += b;
            ++last;
        }
        const uint64_t n = last - first;
        // stripe jobs, then sliding-window jobs, then one-lane jobs, each in order
        std::stable_partition(items.begin() + (ptrdiff_t)first, items.begin() + (ptrdiff_t)last,
                              [](const DpItem &it) { return it.cols != TINY_COLS_TAG; });
        uint64_t nbig = 0;
        while (nbig < n && items[first + nbig].cols != TINY_COLS_TAG) ++nbig;
        std::stable_partition(items.begin() + (ptrdiff_t)first, items.begin() + (ptrdiff_t)(first + nbig),
                              [](const DpItem &it) { return it.cols < WINDOW_TAG; });
        uint64_t nstripe = 0;
        while (nstripe < nbig && items[first + nstripe].cols < WINDOW_TAG) ++nstripe;
        // wavefront stripes, then the row-packed jobs by shape (16, 14, ..., 4
        // columns a lane), each in order: a stable bucket pass
        auto shape = [](const DpItem &it) {
            return it.cols < GROUP_TAG ? 0 : 1 + (int)(GROUP_TAG + 16 - it.cols) / 2;
        };
        uint64_t rows_at[ROW_SHAPES + 2] = {};
        for (uint64_t x = 0; x < nstripe; ++x) ++rows_at[shape(items[first + x]) + 1];
        for (int q = 0; q <= ROW_SHAPES; ++q) rows_at[q + 1] += rows_at[q];
        {
            uint64_t put[ROW_SHAPES + 1];
            for (int q = 0; q <= ROW_SHAPES; ++q) put[q] = rows_at[q];
            shaped.resize(nstripe);
            for (uint64_t x = 0; x < nstripe; ++x) shaped[put[shape(items[first + x])]++] = items[first + x];
            std::copy(shaped.begin(), shaped.end(), items.begin() + (ptrdiff_t)first);
        }
        // A wavefront of 16-lane stripes runs four jobs whose stripe loops
        // diverge: each stripe costs it the longest of the four, so banded
        // jobs of those shapes go in order of stripe count, then rows a
        // stripe (largest first), and the four of a wavefront step alike.
        // (Full jobs of these shapes hold one stripe of LA rows: their
        // largest-first order is that order already.)
        for (int q : {5, 7}) {   // 16-lane stripes of 8 and of 4 columns a lane
            const uint64_t b0 = first + rows_at[q], b1 = first + rows_at[q + 1];
            bool banded = false;
            for (uint64_t x = b0; x < b1 && !banded; ++x) banded = items[x].band != 0;
            if (!banded) continue;
            std::vector<std::pair<uint64_t, uint32_t>> keyed(b1 - b0);
            for (uint64_t x = b0; x < b1; ++x) {
                const DpItem &it = items[x];
                const mh_dp_job &jb = jobs[it.job];
                const uint64_t sw = 16ull * (it.cols - GROUP_TAG);
                const uint64_t ns = (jb.b_len + sw - 1) / sw;
                uint64_t rows = jb.a_len;
                if (it.band) rows = std::min<uint64_t>(jb.a_len, jb.a_len * (sw + 2ull * it.band) / jb.b_len + 1);
                keyed[x - b0] = {(ns << 32) | rows, (uint32_t)(x - b0)};
            }
            std::stable_sort(keyed.begin(), keyed.end(),
                             [](const std::pair<uint64_t, uint32_t> &a, const std::pair<uint64_t, uint32_t> &b) {
                                 return a.first > b.first;
                             });
            shaped.resize(b1 - b0);
            for (uint64_t x = 0; x < b1 - b0; ++x) shaped[x] = items[b0 + keyed[x].second];
            std::copy(shaped.begin(), shaped.end(), items.begin() + (ptrdiff_t)b0);
        }
        const uint64_t n64 = rows_at[1];   // rows_at[1..]: the row jobs, 16 columns a lane first
        // the forward variant: the wide one when a shape wider than 8 columns a lane is in the round
        bool wide = false;
        for (uint64_t z = 0; z < n64 && !wide; ++z) wide = items[first + z].cols == 16;
        for (int q = 0; q < 4; ++q)
            if (rows_at[q + 2] > rows_at[q + 1]) wide = true;
        const int variant = dp_variant_for(wide);
        // Pipelined stripe jobs.  A round with fewer work items than the
        // forward launch's wavefront slots ends on its longest jobs' latency
        // with SIMDs idle: its 64-lane stripe jobs then run on P_j wavefronts
        // each.  P_j = ceil(W_j / T) (W_j the job's trace bytes, a measure of
        // its cells), at most DP_PIPE_MAX and the job's stripes, with the
        // smallest T whose parts fit the slots; so the parts are of about
        // one size.  MH_DP_PIPE = P instead gives P wavefronts to the first
        // jobs with at least P stripes, as many as the idle slots hold.
        // Pipelined jobs go to the front of the queue, most parts first.
        uint64_t npipe = 0;
        pj.assign(n64, 1);
        if (!fx && ncu > 0) {
            uint64_t items_est = n64 + (nbig - nstripe) + (n - nbig + 63) / 64;
            for (int q = 0; q < ROW_SHAPES; ++q) items_est += (rows_at[q + 2] - rows_at[q + 1] + 3) / 4;
            const uint64_t slots = (uint64_t)ncu * 4 * (uint64_t)FWD_VARIANTS[variant].wpe;
            auto nstr = [&](const DpItem &it) {
                const uint64_t sw = 64ull * it.cols;
                return (jobs[it.job].b_len + sw - 1) / sw;
            };
            if (items_est < slots && pipe_env > 1) {
                const uint64_t kmax = (slots - items_est) / (uint64_t)(pipe_env - 1);
                for (uint64_t x = 0, k = 0; x < n64 && k < kmax; ++x)
                    if (nstr(items[first + x]) >= (uint64_t)pipe_env) {
                        pj[x] = (uint8_t)pipe_env;
                        ++k;
                    }
            } else if (items_est < slots) {
                const uint64_t room = slots - (items_est - n64);   // wavefronts for the stripe jobs
                auto parts = [&](double T, bool set) {
                    uint64_t tot = 0;
                    for (uint64_t x = 0; x < n64; ++x) {
                        const DpItem &it = items[first + x];
                        const uint64_t cap = std::min<uint64_t>(DP_PIPE_MAX, nstr(it));
                        const uint64_t q = std::max<uint64_t>(1, std::min<uint64_t>(cap, (uint64_t)std::ceil((double)it.tb_bytes / T)));
                        tot += q;
                        if (set) pj[x] = (uint8_t)q;
                    }
                    return tot;
                };
                double wmax = 1;
                for (uint64_t x = 0; x < n64; ++x) wmax = std::max(wmax, (double)items[first + x].tb_bytes);
                double lo = wmax / (2.0 * DP_PIPE_MAX), hi = wmax;   // parts(hi) = n64 <= room
                for (int it = 0; it < 48; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    (parts(mid, false) <= room ? hi : lo) = mid;
                }
                parts(hi, true);
            }
            for (uint64_t x = 0; x < n64; ++x) npipe += pj[x] > 1;
            if (npipe) {
                // most parts first, each part count in order; then the others
                shaped.clear();
                std::vector<uint8_t> pj2;
                for (uint32_t P = DP_PIPE_MAX; P >= 2; --P)
                    for (uint64_t x = 0; x < n64; ++x)
                        if (pj[x] == P) {
                            shaped.push_back(items[first + x]);
                            pj2.push_back((uint8_t)P);
                        }
                for (uint64_t x = 0; x < n64; ++x)
                    if (pj[x] <= 1) {
                        shaped.push_back(items[first + x]);
                        pj2.push_back(1);
                    }
                std::copy(shaped.begin(), shaped.end(), items.begin() + (ptrdiff_t)first);
                pj.swap(pj2);
                // (a count under the timers' names: pipelined jobs a round)
                host_time(ctx, "dp_pipe_jobs", (double)npipe);
            }
        }
        DevJob *dj = nullptr;
        std::vector<DevJob> host_jobs;
        if (fx) {
            host_jobs.resize(n + (n * 4 + sizeof(DevJob) - 1) / sizeof(DevJob));
            dj = host_jobs.data();
        } else {
            const int rc0 = pinned_jobs(ctx, n * sizeof(DevJob) + n * 4 + 64, (void **)&dj);
            if (rc0 != MH_OK) return rc0;
        }
        // trace-back order: the jobs whose walks are long and whose layout the
        // wavefront walk reads (64-lane stripes; 16-lane stripes of 4 or 8
        // columns a lane) walk on a wavefront each, the others on a thread each
        uint32_t *tbi = reinterpret_cast<uint32_t *>(dj + n);
        uint64_t nwave = 0;
        {
            auto wave_walk = [&](const DpItem &it) {
                if (jobs[it.job].a_len + jobs[it.job].b_len < WAVE_WALK_MIN) return false;
                if (it.cols == TINY_COLS_TAG || it.cols >= WINDOW_TAG) return false;
                return it.cols < GROUP_TAG || it.cols == GROUP_TAG + 4 || it.cols == GROUP_TAG + 8;
            };
            for (uint64_t x = 0; x < n; ++x)
                if (wave_walk(items[first + x])) tbi[nwave++] = (uint32_t)x;
            uint64_t t = nwave;
            for (uint64_t x = 0; x < n; ++x)
                if (!wave_walk(items[first + x])) tbi[t++] = (uint32_t)x;
        }
        uint64_t tw = 0, bo = 0, bo_pipe = 0;
        for (uint64_t x = 0; x < n; ++x) {
            const DpItem &it = items[first + x];
            const uint32_t j = it.job;
            DevJob &d = dj[x];
            d.a_off = jobs[j].a_off;
            d.b_off = jobs[j].b_off;
            d.la = jobs[j].a_len;
            d.lb = jobs[j].b_len;
            d.tb_off = tw;
            d.bnd_off = bo;
            d.path_off = slot[j];
            d.user = j;
            d.cols = it.cols;
            d.band = it.band;
            d.pad_ = 0;
            tw += (it.tb_bytes + 15) & ~15ull;
            bo += bnd_ints(it);
            if (x < npipe) {
                // tagged 64-bit words, in two buffers (stripes alternate by parity)
                bo += 3 * bnd_ints(it);
                bo_pipe = bo;
            }
            if (cells) *cells += 2 * it.tb_bytes;
        }
        DevJob *d_jobs;
        DpOut *d_outs;
        uint8_t *d_tb;
        int *d_bnd;
        DpRound *d_round;
        if (fx) {
            // [round descriptor + cursors] [jobs | results] [boundary ints] [trace]
            uint8_t *q = fx->ptr;
            d_round = reinterpret_cast<DpRound *>(q);
            q += DP_ROUND_BYTES;
            d_jobs = reinterpret_cast<DevJob *>(q);
            d_outs = reinterpret_cast<DpOut *>(d_jobs + n);
            q += align256(n * (sizeof(DevJob) + sizeof(DpOut) + 4));
            d_bnd = reinterpret_cast<int *>(q);
            q += align256(bo * 4);
            d_tb = q;
            MH_CHECK((uint64_t)(q - fx->ptr) + tw <= fx->bytes, MH_ERR_INVALID, "DP round exceeds the scratch");
        } else {
            int rc;
            if ((rc = ctx->dp_tb.ensure(tw + 256)) != MH_OK) return rc;
            if ((rc = ctx->dp_bnd.ensure(bo * 4 + 256)) != MH_OK) return rc;
            if ((rc = ctx->dp_jobs.ensure(n * (sizeof(DevJob) + sizeof(DpOut) + 4) + 256)) != MH_OK) return rc;
            if ((rc = ctx->dp_round.ensure(DP_ROUND_BYTES)) != MH_OK) return rc;
            d_jobs = ctx->dp_jobs.as<DevJob>();
            d_outs = reinterpret_cast<DpOut *>(d_jobs + n);
            d_tb = ctx->dp_tb.as<uint8_t>();
            d_bnd = ctx->dp_bnd.as<int>();
            d_round = ctx->dp_round.as<DpRound>();
        }
        uint32_t *d_tbi = reinterpret_cast<uint32_t *>(d_outs + n);
        MH_HIP(hipMemcpyAsync(d_jobs, dj, n * sizeof(DevJob), hipMemcpyHostToDevice, ctx->stream));
        MH_HIP(hipMemcpyAsync(d_tbi, tbi, n * 4, hipMemcpyHostToDevice, ctx->stream));
        // the pipelined jobs' boundary words (the first jobs') start untagged
        if (bo_pipe) MH_HIP(hipMemsetAsync(d_bnd, 0, bo_pipe * 4, ctx->stream));
        // the round's descriptor: pointers, the trace-back split and the queue's bins
        DpRound hr;
        std::memset(&hr, 0, sizeof(hr));
        hr.jobs = d_jobs;
        hr.outs = d_outs;
        hr.A = d_a;
        hr.B = d_b;
        hr.tb = d_tb;
        hr.bnd = d_bnd;
        hr.paths = d_paths;
        hr.res = d_results;
        hr.dev = d_dev;
        hr.cursor = reinterpret_cast<uint32_t *>(reinterpret_cast<uint8_t *>(d_round) + sizeof(DpRound));
        for (int k = 0; k < 4; ++k) hr.sr[k] = sr[k];
        hr.goh = -200;
        hr.nwave = (uint32_t)nwave;
        hr.n = (uint32_t)n;
        auto add_bin = [&](uint32_t kind, uint64_t b0, uint64_t cnt) {
            if (cnt == 0) return;
            hr.kind[hr.nbins] = kind;
            hr.first[hr.nbins] = (uint32_t)b0;
            hr.count[hr.nbins] = (uint32_t)cnt;
            ++hr.nbins;
        };
        {
            // stripe jobs: a new bin where the trace size falls below the next power of 4
            // the pipelined jobs first, a bin for each part count, an item a part
            for (uint64_t x = 0; x < npipe;) {
                uint64_t y = x + 1;
                while (y < npipe && pj[y] == pj[x]) ++y;
                add_bin(QK_PIPE + (pj[x] - 2u), x, (y - x) * pj[x]);
                x = y;
            }
            auto cls = [](uint64_t v) { return v ? (63 - __builtin_clzll(v)) / 2 : 0; };
            uint64_t x = npipe;
            while (x < n64) {
                const int c0 = cls(items[first + x].tb_bytes);
                uint64_t y = x + 1;
                while (y < n64 && cls(items[first + y].tb_bytes) == c0) ++y;
                // keep ROW_SHAPES + 2 bins for the row, window and one-lane
                // jobs, and the last cursor for the wait-bound error flag
                if (hr.nbins + 1 >= (uint32_t)(DPQ_MAXB - ROW_SHAPES - 3)) y = n64;
                add_bin(QK_STRIPE, x, y - x);
                x = y;
            }
            for (int q = 0; q < ROW_SHAPES; ++q) add_bin(QK_ROWS + q, rows_at[q + 1], rows_at[q + 2] - rows_at[q + 1]);
            add_bin(QK_WINDOW, nstripe, nbig - nstripe);
            add_bin(QK_TINY, nbig, n - nbig);
        }
        MH_CHECK(hr.nbins < (uint32_t)DPQ_CURSORS, MH_ERR_INVALID, "DP round: too many queue bins");
        uint64_t nitems_q = 0;
        for (uint32_t q = 0; q < hr.nbins; ++q) {
            const uint32_t per =
                hr.kind[q] == QK_STRIPE || hr.kind[q] == QK_WINDOW || hr.kind[q] >= QK_PIPE ? 1u
                                                                                          : (hr.kind[q] == QK_TINY ? 64u : 4u);
            hr.item0[q] = (uint32_t)nitems_q;
            nitems_q += (hr.count[q] + per - 1) / per;
        }
        hr.item0[hr.nbins] = (uint32_t)nitems_q;
        RoundGrids grids;
        {
            int rc;
            if ((rc = round_grids(ctx, variant, nitems_q, hr.nwave, hr.n, &grids)) != MH_OK) return rc;
        }
        FwdArgs fwa;
        fwa.jobs = d_jobs;
        fwa.outs = d_outs;
        fwa.A = d_a;
        fwa.B = d_b;
        fwa.tb = d_tb;
        fwa.bnd = d_bnd;
        fwa.rd = d_round;
        fwa.cursor = hr.cursor;
        fwa.sr0 = sr[0];
        fwa.sr1 = sr[1];
        fwa.sr2 = sr[2];
        fwa.sr3 = sr[3];
        fwa.goh = -200;
        TbArgs tba;
        tba.jobs = d_jobs;
        tba.outs = d_outs;
        tba.tb = d_tb;
        tba.paths = d_paths;
        tba.res = d_results;
        tba.dev = d_dev;
        tba.idx = d_tbi;
        tba.nwave = hr.nwave;
        tba.first = hr.nwave;
        tba.n = hr.n;
        if (fx) {
            MH_HIP(hipMemcpyAsync(d_round, &hr, sizeof(hr), hipMemcpyHostToDevice, ctx->stream));
            int rc;
            if ((rc = dp_launch_round(ctx, d_round, variant, false, grids, fwa, tba)) != MH_OK) return rc;
        } else {
            int rc;
            if ((rc = dp_round_staging(ctx)) != MH_OK) return rc;
            std::memcpy(ctx->dp_round_host, &hr, sizeof(hr));
            MH_HIP(hipMemcpyAsync(d_round, ctx->dp_round_host, sizeof(hr), hipMemcpyHostToDevice, ctx->stream));
            if ((rc = dp_launch_round(ctx, d_round, variant, dp_use_graph(), grids, fwa, tba)) != MH_OK) return rc;
        }
        // the host job table is reused by the next round
        MH_HIP(hipStreamSynchronize(ctx->stream));
        if (npipe) {
            uint32_t err = 0;
            MH_HIP(hipMemcpyAsync(&err, hr.cursor + (DPQ_CURSORS - 1), 4, hipMemcpyDeviceToHost, ctx->stream));
            MH_HIP(hipStreamSynchronize(ctx->stream));
            MH_CHECK(err == 0, MH_ERR_DEVICE, "DP: a pipelined stripe's wait ran past its limit");
        }
        first = last;
    }
    return MH_OK;
}

__global__ void pack_paths_kernel(const uint8_t *__restrict__ p, uint64_t n, uint8_t *__restrict__ out)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k * 4 >= n) return;
    uint32_t v = 0;
    for (int x = 0; x < 4; ++x) {
        const uint64_t at = k * 4 + x;
        const uint8_t e = at < n ? p[at] : 0;
        const uint32_t c = e == 'M' ? 0u : (e == 'D' ? 1u : (e == 'I' ? 2u : 3u));
        v |= c << (2 * x);
    }
    out[k] = (uint8_t)v;
}

int dp_pack_paths(mh_context *ctx, const uint8_t *d_paths, uint64_t n, uint8_t *d_out)
{
    const uint64_t nb = (n + 3) / 4;
    if (!nb) return MH_OK;
    hipLaunchKernelGGL(pack_paths_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, ctx->stream, d_paths, n,
                       d_out);
    MH_HIP(hipGetLastError());
    return MH_OK;
}

static int dp_check_jobs(const mh_dp_job *jobs, uint64_t njobs, std::vector<uint64_t> &slot)
{
    MH_CHECK(njobs < (1ull << 31), MH_ERR_INVALID, "too many DP jobs in one call");
    slot.resize(njobs);
    uint64_t acc = 0;
    for (uint64_t j = 0; j < njobs; ++j) {
        MH_CHECK(jobs[j].a_len > 0 && jobs[j].b_len > 0, MH_ERR_ZERO_LENGTH,
                 "NWSmall: length=0 (job " + std::to_string(j) + ")");
        MH_CHECK(jobs[j].a_len < (1u << 24) && jobs[j].b_len < (1u << 24), MH_ERR_UNSUPPORTED,
                 "DP operand longer than 2^24");
        slot[j] = acc;
        acc += (uint64_t)jobs[j].a_len + jobs[j].b_len;
    }
    return MH_OK;
}

static DpItem make_item(const mh_dp_job &jb, uint32_t job, uint32_t band, int force, bool rerun = false)
{
    DpItem it;
    it.job = job;
    it.cols = dp_cols_for(jb.a_len, jb.b_len, band, force, rerun);
    it.band = dp_effective_band(jb.a_len, jb.b_len, band, it.cols);
    if (it.band == 0 && band != 0) it.cols = dp_cols_for(jb.a_len, jb.b_len, 0, force);
    if (it.band == 0 && !force && jb.b_len <= (uint32_t)TINY_LB) {
        it.cols = TINY_COLS_TAG;
        it.tb_bytes = (uint64_t)jb.a_len * 8;
        return it;
    }
    if (it.band == 0 && !force && jb.b_len <= ROWS_MAX_LB && jb.a_len + jb.b_len < WAVE_WALK_MIN &&
        shape_table()[dp_tune_class(jb.b_len)] == 0) {
        // four jobs per wavefront, one stripe each of the fewest columns a lane
        // (an even number from 4 to 16) that holds the job
        const uint32_t c = std::max(4u, 2u * ((jb.b_len + 31) / 32));
        it.cols = GROUP_TAG + c;
        it.tb_bytes = dp_trace_bytes(jb.a_len, jb.b_len, 0, c, 16);
        return it;
    }
    it.tb_bytes = it.cols > GROUP_TAG ? dp_trace_bytes(jb.a_len, jb.b_len, it.band, it.cols - GROUP_TAG, 16)
                                      : dp_trace_bytes(jb.a_len, jb.b_len, it.band, it.cols, 64);
    if (it.band) {
        const uint32_t wc = dp_window_cols(jb.a_len, jb.b_len, it.band, dp_force_kind());
        if (wc) {
            it.cols = WINDOW_TAG + wc;
            it.tb_bytes = (uint64_t)((jb.a_len + 8 / wc - 1) / (8 / wc)) * 256;
        }
    }
    return it;
}

// items[x] = make_item(jobs[x], x, bands[x]) for every job, in parallel host
// threads for large batches
static void make_items(const mh_dp_job *jobs, uint64_t njobs, const uint32_t *bands, std::vector<DpItem> &items)
{
    const int force = dp_force_cols();
    items.resize(njobs);
    auto fill = [&](uint64_t lo, uint64_t hi) {
        for (uint64_t j = lo; j < hi; ++j) items[j] = make_item(jobs[j], (uint32_t)j, bands ? bands[j] : 0u, force);
    };
    const uint64_t chunk = 8192;
    host_parallel((size_t)((njobs + chunk - 1) / chunk),
                  [&](size_t k) { fill(k * chunk, std::min<uint64_t>(njobs, (k + 1) * chunk)); });
}

uint64_t dp_item_cells(uint32_t la, uint32_t lb, uint32_t band)
{
    if (la == 0 || lb == 0) return 0;
    const DpItem it = make_item(mh_dp_job{0, 0, la, lb}, 0, band, dp_force_cols());
    if (it.cols > WINDOW_TAG) return (uint64_t)la * 64 * (it.cols - WINDOW_TAG);
    if (it.cols == TINY_COLS_TAG) return (uint64_t)la * lb;
    return 2 * it.tb_bytes;
}

// The jobs whose operands hold letters other than A, C, G, T (class > 3):
// from the caller's flags, or a device scan; codes above 5 are refused.
static int split_wild(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
                      uint64_t njobs, const uint8_t *wild, const FixedScratch *fx, std::vector<uint32_t> &out)
{
    out.clear();
    if (wild) {
        for (uint64_t j = 0; j < njobs; ++j)
            if (wild[j]) out.push_back((uint32_t)j);
        return MH_OK;
    }
    std::vector<uint8_t> mx;
    const int rc = dp_code_max(ctx, d_a, d_b, jobs, njobs, fx, mx);
    if (rc != MH_OK) return rc;
    for (uint64_t j = 0; j < njobs; ++j) {
        MH_CHECK(mx[j] < DP_NCLASS, MH_ERR_INVALID,
                 "DP operand code above 5 in job " + std::to_string(j) + " (codes: 0-3 A,C,G,T, 4 X, 5 other)");
        if (mx[j] > 3) out.push_back((uint32_t)j);
    }
    return MH_OK;
}

// drops the items of the float-DP jobs (items[j] is job j's)
static void drop_wild(std::vector<DpItem> &items, const std::vector<uint32_t> &wild)
{
    if (wild.empty()) return;
    size_t w = 0, k = 0;
    for (size_t j = 0; j < items.size(); ++j) {
        if (k < wild.size() && wild[k] == j) {
            ++k;
            continue;
        }
        items[w++] = items[j];
    }
    items.resize(w);
}

int dp_run(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
           uint64_t njobs, uint8_t *d_paths, mh_dp_result *d_results, const uint8_t *wild)
{
    if (njobs == 0) return MH_OK;
    std::vector<uint64_t> slot;
    int rc;
    if ((rc = dp_check_jobs(jobs, njobs, slot)) != MH_OK) return rc;
    std::vector<uint32_t> wj;
    if ((rc = split_wild(ctx, d_a, d_b, jobs, njobs, wild, nullptr, wj)) != MH_OK) return rc;
    std::vector<DpItem> items;
    const double tp0 = now_ms();
    make_items(jobs, njobs, nullptr, items);
    drop_wild(items, wj);
    host_time(ctx, "host_dp_plan", now_ms() - tp0);
    if ((rc = dp_run_items(ctx, d_a, d_b, jobs, slot, items, d_paths, d_results, nullptr, nullptr)) != MH_OK)
        return rc;
    return dp_run_float(ctx, d_a, d_b, jobs, slot, wj, d_paths, d_results, nullptr, nullptr, nullptr);
}

int dp_run_noalloc(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs, uint64_t njobs,
                   uint8_t *d_paths, mh_dp_result *d_results)
{
    uint8_t *base = nullptr;
    if (ctx->device >= 0 && ctx->device < DP_SCRATCH_MAX_DEVICES) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[ctx->device];
    }
    MH_CHECK(base, MH_ERR_INVALID, "the DP scratch of this device was not set up (mh_create does it)");
    if (njobs == 0) return MH_OK;
    std::vector<uint64_t> slot;
    int rc;
    if ((rc = dp_check_jobs(jobs, njobs, slot)) != MH_OK) return rc;
    FixedScratch fx;
    fx.ptr = base;
    fx.bytes = DP_SCRATCH_BYTES;
    std::vector<uint32_t> wj;
    if ((rc = split_wild(ctx, d_a, d_b, jobs, njobs, nullptr, &fx, wj)) != MH_OK) return rc;
    std::vector<DpItem> items;
    make_items(jobs, njobs, nullptr, items);
    drop_wild(items, wj);
    if ((rc = dp_run_items(ctx, d_a, d_b, jobs, slot, items, d_paths, d_results, nullptr, nullptr, &fx)) != MH_OK)
        return rc;
    return dp_run_float(ctx, d_a, d_b, jobs, slot, wj, d_paths, d_results, nullptr, nullptr, &fx);
}

int dp_run_banded(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs, uint64_t njobs,
                  const uint32_t *bands, uint32_t flags, uint8_t *d_paths, mh_dp_result *d_results,
                  uint32_t *d_dev, mh_band_stats *stats)
{
    mh_band_stats st;
    std::memset(&st, 0, sizeof(st));
    if (stats) *stats = st;
    if (njobs == 0) return MH_OK;
    std::vector<uint64_t> slot;
    int rc;
    if ((rc = dp_check_jobs(jobs, njobs, slot)) != MH_OK) return rc;
    if (!d_dev) {
        if ((rc = ctx->dp_dev.ensure(njobs * 4 + 64)) != MH_OK) return rc;
        d_dev = ctx->dp_dev.as<uint32_t>();
    }
    const double tp0 = now_ms();
    // a job with band W evaluates the region of band 2W and accepts its path
    // when it stays within W of the line (the region reaches at least W past
    // the accepted band, where a path that wants to leave shows itself)
    std::vector<uint32_t> region(njobs, 0);
    for (uint64_t j = 0; j < njobs; ++j) {
        const uint64_t w = bands ? bands[j] : 0u;
        region[j] = 2 * w > 0xFFFFFFFFull ? 0u : (uint32_t)(2 * w);
    }
    // jobs with letters other than A,C,G,T run the full DP in float, unbanded
    std::vector<uint32_t> wj;
    if ((rc = split_wild(ctx, d_a, d_b, jobs, njobs, nullptr, nullptr, wj)) != MH_OK) return rc;
    std::vector<DpItem> items;
    make_items(jobs, njobs, region.data(), items);
    std::vector<uint32_t> accept1(njobs, 0);
    {
        size_t k = 0;
        for (uint64_t j = 0; j < njobs; ++j) {
            if (k < wj.size() && wj[k] == j) {
                ++k;
                continue;
            }
            if (items[j].band) {
                accept1[j] = bands[j];
                ++st.banded;
            }
        }
    }
    drop_wild(items, wj);
    host_time(ctx, "host_dp_plan", now_ms() - tp0);
    const int force = dp_force_cols();
    if ((rc = dp_run_items(ctx, d_a, d_b, jobs, slot, items, d_paths, d_results, d_dev, &st.cells)) != MH_OK)
        return rc;
    if ((rc = dp_run_float(ctx, d_a, d_b, jobs, slot, wj, d_paths, d_results, d_dev, &st.cells, nullptr)) != MH_OK)
        return rc;
    if (!(flags & MH_BAND_NO_FALLBACK) && st.banded) {
        // edge fallback: a banded path that reaches distance W is rerun with
        // the band doubled (region 4W, accepted within 2W), and one that
        // reaches 2W too with the full DP
        std::vector<uint32_t> dev(njobs);
        MH_HIP(hipMemcpyAsync(dev.data(), d_dev, njobs * 4, hipMemcpyDeviceToHost, ctx->stream));
        MH_HIP(hipStreamSynchronize(ctx->stream));
        double tq = now_ms();
        host_time(ctx, "band_pass1", tq - tp0);
        std::vector<DpItem> again;
        std::vector<uint32_t> accept2(njobs, 0);
        for (uint64_t j = 0; j < njobs; ++j)
            if (accept1[j] && dev[j] >= accept1[j]) {
                const uint64_t r2 = 4ull * accept1[j];
                again.push_back(make_item(jobs[j], (uint32_t)j, r2 > 0xFFFFFFFFull ? 0u : (uint32_t)r2, force, true));
                if (again.back().band) {
                    accept2[j] = 2 * accept1[j];
                    ++st.rerun_doubled;
                } else {
                    ++st.rerun_full;
                }
            }
        if (!again.empty()) {
            if ((rc = dp_run_items(ctx, d_a, d_b, jobs, slot, again, d_paths, d_results, d_dev, &st.cells)) != MH_OK)
                return rc;
            MH_HIP(hipMemcpyAsync(dev.data(), d_dev, njobs * 4, hipMemcpyDeviceToHost, ctx->stream));
            MH_HIP(hipStreamSynchronize(ctx->stream));
            host_time(ctx, "band_pass2", now_ms() - tq);
            tq = now_ms();
            std::vector<DpItem> full;
            for (uint64_t j = 0; j < njobs; ++j)
                if (accept2[j] && dev[j] >= accept2[j]) {
                    full.push_back(make_item(jobs[j], (uint32_t)j, 0, force));
                    ++st.rerun_full;
                }
            if (!full.empty()) {
                if ((rc = dp_run_items(ctx, d_a, d_b, jobs, slot, full, d_paths, d_results, d_dev, &st.cells)) != MH_OK)
                    return rc;
                MH_HIP(hipStreamSynchronize(ctx->stream));
                host_time(ctx, "band_pass3", now_ms() - tq);
            }
        }
    }
    if (stats) *stats = st;
    return MH_OK;
}

}  // namespace mh
