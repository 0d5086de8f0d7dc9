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
#include <algorithm>
#include <numeric>
#include <cstdlib>
#include <utility>
#include <type_traits>

namespace mh {
namespace {

constexpr int LANES = 64;
constexpr int NEG = -(1 << 29);
constexpr int SBIAS = 65;           // smallest NUC_SP entry is -125 + 60 = -65

struct DevJob {
    uint64_t a_off, b_off;
    uint32_t la, lb;
    uint64_t tb_off;    // 32-bit words
    uint64_t bnd_off;   // int32 elements
    uint64_t path_off;  // bytes (slot start)
    uint32_t user;      // index in the caller's job array
    uint32_t cols;      // columns per lane (8 or 16): stripe width 64*cols
};

struct DpOut {
    int32_t score;
    int32_t edge;       // 'M', 'D' or 'I'
};

__device__ __forceinline__ int shr1_old(int v, int old)
{
    // wave_shr:1 -- lane k receives lane k-1's value, lane 0 receives `old`
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
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

// One wavefront per job.  D and I are kept shifted by the close half (goh):
// D' = D + goh and I' = I + goh, so the M recurrence reads them without an
// add; every comparison is between two values shifted alike, so ties and
// trace bits are unchanged.  Per row a lane runs two passes over its COLS
// columns: M and D right to left (each needs only the previous row's values of
// its own and the left column), then I left to right (it chains along the
// row).  Both update the registers in place.
//
// Lanes outside their row range sit out the step (exec mask).  Trace bits are the
// reference's 4-bit code (MU/types.h:26-46): bits 0-1 the M source (MM=0,
// DM=1, IM=2), bit 2 MD, bit 3 MI; COLS/8 32-bit words per lane per step, so
// a wavefront stores 64*COLS/2 contiguous bytes per step.  Cell (i, j) of lane
// k, column c is in word (stripe*(LA+63) + i-1+k)*64 + k)*COLS/8 + c/8, bits
// 4*(c%8) .. +3.
template <int COLS>
__device__ __forceinline__ void dp_job(const DevJob &jb, const uint8_t *__restrict__ A,
                                       const uint8_t *__restrict__ B, uint32_t *__restrict__ tb,
                                       int *__restrict__ bnd, DpOut *__restrict__ out, uint32_t sr0,
                                       uint32_t sr1, uint32_t sr2, uint32_t sr3, int goh, int (*lds_in)[64],
                                       int (*lds_out)[64])
{
    constexpr int STRIPE_W = LANES * COLS;
    constexpr int NW = COLS / 8;           // trace words per lane per step
    const int lane = threadIdx.x;
    const int la = (int)jb.la, lb = (int)jb.lb;
    const uint8_t *a = A + jb.a_off;
    const uint8_t *b = B + jb.b_off;
    const int nr = (la + 63) & ~63;
    int *bM = bnd + jb.bnd_off;
    int *bD = bM + nr;
    int *bI = bD + nr;
    uint32_t *tbj = tb + jb.tb_off;
    const int nstripes = (lb + STRIPE_W - 1) / STRIPE_W;
    const int fin_lane = ((lb - 1) % STRIPE_W) / COLS;
    const int fin_col = (lb - 1) % COLS;
    const int go2 = 2 * goh;
    int finM = NEG, finD = NEG, finI = NEG;

    for (int s = 0; s < nstripes; ++s) {
        const int j0 = s * STRIPE_W + 1;
        const int jl = j0 + lane * COLS;
        const int rem = lb - (jl - 1);
        const int ncols = rem < 0 ? 0 : (rem > COLS ? COLS : rem);
        const int span = lb - s * STRIPE_W;
        const int nl = span >= STRIPE_W ? LANES : (span + COLS - 1) / COLS;
        const bool last_stripe = (s == nstripes - 1);

        uint32_t bsh[COLS];
#pragma unroll
        for (int c = 0; c < COLS; ++c) bsh[c] = (c < ncols) ? 8u * (uint32_t)b[jl - 1 + c] : 0u;

        // row-0 state: M = NEG, D' = NEG + goh, I' = 0 + goh; left column likewise,
        // or column 0 (M(0,0) = 0, D' = I' = NEG + goh) for the first lane
        const int iM0 = jl == 1 ? 0 : NEG, iI0 = jl == 1 ? NEG + goh : goh;
        int Mp[COLS], Dp[COLS], Ip[COLS];
#pragma unroll
        for (int c = 0; c < COLS; ++c) { Mp[c] = NEG; Dp[c] = NEG + goh; Ip[c] = goh; }
        int lpM = iM0, lpD = NEG + goh, lpI = iI0;
        int outM = NEG, outD = NEG, outI = NEG, acode = 0;
        const int ob2_0 = (jl == 1) ? goh : go2;      // open_B[j-1] + goh for the lane's first column

        // lane 0's inputs, 64 rows per chunk: column jl-1 of the previous stripe
        // (or column 0: M = NEG, D' = 0 + goh, I' = NEG + goh) and the A codes;
        // the next chunk is loaded one chunk ahead
        int nM = NEG, nD = goh, nI = NEG + goh, nA = 0;
        if (s > 0 && lane < la) { nM = bM[lane]; nD = bD[lane]; nI = bI[lane]; }
        if (lane < la) nA = a[lane];

        const int nsteps = la + nl - 1;
        for (int t = 0; t < nsteps; ++t) {
            const int q = t & 63;
            if (q == 0) {
                lds_in[0][lane] = nM;
                lds_in[1][lane] = nD;
                lds_in[2][lane] = nI;
                lds_in[3][lane] = nA;
                const int r = t + 64 + lane;
                nM = NEG; nD = goh; nI = NEG + goh; nA = 0;
                if (s > 0 && r < la) { nM = bM[r]; nD = bD[r]; nI = bI[r]; }
                if (r < la) nA = a[r];
            }
            const int rM = shr1_old(outM, lds_in[0][q]);
            const int rD = shr1_old(outD, lds_in[1][q]);
            const int rI = shr1_old(outI, lds_in[2][q]);
            const int ra = shr1_old(acode, lds_in[3][q]);
            acode = ra;

            const int i = t - lane + 1;
            if (i >= 1 && i <= la && lane < nl) {
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
            uint32_t *dst = tbj + (((size_t)s * (size_t)(la + 63) + (size_t)t) * 64 + lane) * NW;
            if constexpr (NW == 2) {
                *reinterpret_cast<uint2 *>(dst) = make_uint2(word[0], word[1]);
            } else {
                dst[0] = word[0];
            }
            if (last_stripe && i == la && lane == fin_lane) {
#pragma unroll
                for (int c = 0; c < COLS; ++c)
                    if (c == fin_col) { finM = Mp[c]; finD = Dp[c]; finI = Ip[c]; }
            }
            }
            if (!last_stripe) {
                // lane 63 just finished row t-62 of the stripe's last column
                const int i63 = t - 62;
                if (lane == 63 && i63 >= 1 && i63 <= la) {
                    const int slot = (i63 - 1) & 63;
                    lds_out[0][slot] = outM;
                    lds_out[1][slot] = outD;
                    lds_out[2][slot] = outI;
                }
                if (i63 >= 1 && i63 <= la && (((i63 - 1) & 63) == 63 || i63 == la)) {
                    const int r = ((i63 - 1) & ~63) + lane;
                    if (r < la) { bM[r] = lds_out[0][lane]; bD[r] = lds_out[1][lane]; bI[r] = lds_out[2][lane]; }
                }
            }
        }
        __threadfence();
    }
    if (lane == fin_lane) {
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

__global__ __launch_bounds__(64) void dp_forward(const DevJob *__restrict__ jobs, const uint8_t *__restrict__ A,
                                                 const uint8_t *__restrict__ B, uint32_t *__restrict__ tb,
                                                 int *__restrict__ bnd, DpOut *__restrict__ outs, uint32_t sr0,
                                                 uint32_t sr1, uint32_t sr2, uint32_t sr3, int goh)
{
    __shared__ int lds_in[4][64];
    __shared__ int lds_out[3][64];
    const DevJob jb = jobs[blockIdx.x];
    if (jb.cols == 16)
        dp_job<16>(jb, A, B, tb, bnd, &outs[blockIdx.x], sr0, sr1, sr2, sr3, goh, lds_in, lds_out);
    else
        dp_job<8>(jb, A, B, tb, bnd, &outs[blockIdx.x], sr0, sr1, sr2, sr3, goh, lds_in, lds_out);
}

__device__ __forceinline__ int xchar(uint32_t bits, int t)
{
    if (t == 'M') { const uint32_t v = bits & 3u; return v == 0 ? 'M' : (v == 1 ? 'D' : (v == 2 ? 'I' : '?')); }
    if (t == 'D') return (bits & 4u) ? 'M' : 'D';
    return (bits & 8u) ? 'M' : 'I';
}

// BitTraceBack, MU/bittraceback.cpp:138-209: one thread per job walks the
// trace bits from (LA, LB) and writes the edges backwards into the job's
// path slot.
__global__ void dp_traceback(const DevJob *__restrict__ jobs, const DpOut *__restrict__ outs,
                             const uint32_t *__restrict__ tb, uint8_t *__restrict__ paths,
                             mh_dp_result *__restrict__ res, int njobs)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= njobs) return;
    const DevJob jb = jobs[x];
    const int la = (int)jb.la, lb = (int)jb.lb;
    const uint32_t *t = tb + jb.tb_off;
    uint8_t *slot = paths + jb.path_off;
    int pos = la + lb;
    int i = la, j = lb;
    int type = outs[x].edge;
    bool ok = true;
    for (;;) {
        slot[--pos] = (uint8_t)type;
        uint32_t nib = 0;
        if (i > 0 && j > 0) {
            const int cols = (int)jb.cols, wpl = cols / 8, sw = 64 * cols;
            const int s = (j - 1) / sw, cc = (j - 1) % sw, k = cc / cols, c = cc % cols;
            const uint32_t w = t[(((size_t)s * (size_t)(la + 63) + (size_t)(i - 1 + k)) * 64 + k) * wpl + (c >> 3)];
            nib = (w >> (4 * (c & 7))) & 15u;
        }
        const int nx = xchar(nib, type);
        if (type == 'M') { if (i == 0 || j == 0) { ok = false; break; } --i; --j; }
        else if (type == 'D') { if (i == 0) { ok = false; break; } --i; }
        else { if (j == 0) { ok = false; break; } --j; }
        if (i == 0 && j == 0) break;
        if (pos == 0 || nx == '?') { ok = false; break; }
        type = nx;
    }
    mh_dp_result r;
    r.score = outs[x].score;
    r.path_len = ok ? (uint32_t)(la + lb - pos) : 0u;
    r.path_start = jb.path_off + (uint64_t)pos;
    res[jb.user] = r;
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

int dp_run(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
           uint64_t njobs, uint8_t *d_paths, mh_dp_result *d_results)
{
    if (njobs == 0) return MH_OK;
    MH_CHECK(njobs < (1ull << 31), MH_ERR_INVALID, "too many DP jobs in one call");
    std::vector<uint64_t> slot(njobs);
    uint64_t acc = 0;
    for (uint64_t j = 0; j < njobs; ++j) {
        MH_CHECK(jobs[j].a_len > 0 && jobs[j].b_len > 0, MH_ERR_ZERO_LENGTH,
                 "NWSmall: length=0 (job " + std::to_string(j) + ")");
        MH_CHECK(jobs[j].a_len < (1u << 24) && jobs[j].b_len < (1u << 24), MH_ERR_UNSUPPORTED,
                 "DP operand longer than 2^24");
        slot[j] = acc;
        acc += (uint64_t)jobs[j].a_len + jobs[j].b_len;
    }
    // largest jobs first so the long tails start early
    std::vector<uint32_t> order(njobs);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return (uint64_t)jobs[x].a_len * jobs[x].b_len > (uint64_t)jobs[y].a_len * jobs[y].b_len;
    });
    // columns per lane: 8 (512-column stripes) for narrow jobs, 16 (1024) for
    // wide ones; MH_DP_COLS=8|16 forces one shape (used for tuning sweeps)
    int force_cols = 0;
    if (const char *e = getenv("MH_DP_COLS")) force_cols = atoi(e);
    auto cols_of = [&](uint32_t j) -> uint32_t {
        if (force_cols == 8 || force_cols == 16) return (uint32_t)force_cols;
        return jobs[j].b_len > 512 ? 16u : 8u;
    };
    auto tb_words = [&](uint32_t j) {
        const uint32_t c = cols_of(j);
        const uint64_t ns = (jobs[j].b_len + 64 * c - 1) / (64 * c);
        return ns * (uint64_t)(jobs[j].a_len + 63) * 64 * (c / 8);
    };
    auto bnd_ints = [&](uint32_t j) {
        const uint32_t c = cols_of(j);
        const uint64_t ns = (jobs[j].b_len + 64 * c - 1) / (64 * c);
        return ns > 1 ? 3ull * ((jobs[j].a_len + 63) & ~63ull) : 0ull;
    };
    // trace arena budget: 70% of the memory that is free or already held by
    // this context's arena (at most 192 GB); once grown the arena is kept
    size_t free_b = 0, total_b = 0;
    MH_HIP(hipMemGetInfo(&free_b, &total_b));
    uint64_t budget = (uint64_t)((free_b + ctx->dp_tb.bytes + ctx->dp_bnd.bytes) * 0.7);
    const uint64_t cap = 192ull << 30;
    if (budget > cap) budget = cap;
    if (const char *e = getenv("MH_DP_TB_BUDGET_GB")) budget = (uint64_t)(atof(e) * (1ull << 30));
    {
        // size the arena once for the whole call: the largest round fits the budget
        uint64_t need = 0;
        for (uint64_t j = 0; j < njobs; ++j) need += (tb_words(order[j]) + bnd_ints(order[j])) * 4;
        if (need > budget) need = budget;
        int rc0;
        if (ctx->dp_tb.bytes < need) {
            ctx->dp_bnd.release();
            if ((rc0 = ctx->dp_tb.ensure(need + 256)) != MH_OK) return rc0;
        }
    }
    uint32_t sr[4];
    packed_rows(sr);

    uint64_t first = 0;
    std::vector<DevJob> dj;
    while (first < njobs) {
        // one round: as many jobs (in schedule order) as fit the trace budget
        uint64_t tbw = 0, bi = 0, last = first;
        while (last < njobs) {
            const uint64_t w = tb_words(order[last]), b = bnd_ints(order[last]);
            if (last > first && (tbw + w) * 4 + (bi + b) * 4 > budget) break;
            tbw += w;
            bi += b;
            ++last;
        }
        const uint64_t n = last - first;
        dj.resize(n);
        uint64_t tw = 0, bo = 0;
        for (uint64_t x = 0; x < n; ++x) {
            const uint32_t j = order[first + x];
            DevJob &d = dj[x];
            d.a_off = jobs[j].a_off;
            d.b_off = jobs[j].b_off;
            d.la = jobs[j].a_len;
            d.lb = jobs[j].b_len;
            d.tb_off = tw;
            d.bnd_off = bo;
            d.path_off = slot[j];
            d.user = j;
            d.cols = cols_of(j);
            tw += tb_words(j);
            bo += bnd_ints(j);
        }
        int rc;
        if ((rc = ctx->dp_tb.ensure(tw * 4 + 256)) != MH_OK) return rc;
        if ((rc = ctx->dp_bnd.ensure(bo * 4 + 256)) != MH_OK) return rc;
        if ((rc = ctx->dp_jobs.ensure(n * sizeof(DevJob) + n * sizeof(DpOut) + 256)) != MH_OK) return rc;
        DevJob *d_jobs = ctx->dp_jobs.as<DevJob>();
        DpOut *d_outs = reinterpret_cast<DpOut *>(d_jobs + n);
        MH_HIP(hipMemcpyAsync(d_jobs, dj.data(), n * sizeof(DevJob), hipMemcpyHostToDevice, ctx->stream));
        {
            TimedLaunch tl(ctx, "dp_forward");
            hipLaunchKernelGGL(dp_forward, dim3((unsigned)n), dim3(64), 0, ctx->stream, d_jobs, d_a, d_b,
                               ctx->dp_tb.as<uint32_t>(), ctx->dp_bnd.as<int>(), d_outs, sr[0], sr[1], sr[2],
                               sr[3], -200);
            MH_HIP(hipGetLastError());
        }
        {
            TimedLaunch tl(ctx, "dp_traceback");
            hipLaunchKernelGGL(dp_traceback, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, ctx->stream, d_jobs,
                               d_outs, ctx->dp_tb.as<uint32_t>(), d_paths, d_results, (int)n);
            MH_HIP(hipGetLastError());
        }
        // the host job table is reused by the next round
        MH_HIP(hipStreamSynchronize(ctx->stream));
        first = last;
    }
    return MH_OK;
}

}  // namespace mh
