// dp_float.hip -- NWSmall in float for DP jobs whose operands hold letters
// other than A, C, G, T.
//
// libMUSCLE scores a single-sequence profile position from fractional letter
// counts (MU/msa2.cpp:19-90): a residue counts 1 for its letter; the DNA
// branch for wildcards tests the amino values AX_R and AX_Y, so 'X' (NX_X ==
// AX_R) counts half to G and half to A, and every other wildcard -- with any
// letter FixAlpha replaces by 'N' (MU/msa.cpp:795-813) -- counts 1/20 to
// each base; DNA counts are not normalised (MU/params.cpp:296-310).  The
// position's m_AAScores (MU/profilefrommsa.cpp:279-285), SortCounts' order
// (159-183) and ScoreProfPos2SPN (MU/scorepp.cpp:80-92) then give fractional
// match scores, and NWSmall (MU/nwsmall.cpp:500-670) adds them in float.  An
// int32 DP cannot reproduce that rounding, so such jobs run here: one thread
// per job does NWSmall's float recurrences in the reference's order of
// operations (no contraction into fused multiply-adds), keeps the 4-bit
// trace of every cell (MU/types.h:26-46), and walks it back as BitTraceBack
// (MU/bittraceback.cpp:138-209).  Operand classes: 0-3 A,C,G,T (U as T),
// 4 'X', 5 any other letter.  Such jobs are rare (N runs in assemblies); the
// A/C/G/T jobs of a batch keep the int32 kernels.
#include "mh_internal.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace mh {

namespace {

struct Score6 {
    float s[DP_NCLASS][DP_NCLASS];
};

struct FloatJob {
    uint64_t a_off, b_off;
    uint64_t path_off;   // slot start in the path buffer
    uint64_t scr_off;    // bytes into the round's scratch
    uint32_t la, lb;
    uint32_t user;       // index in the caller's job array
    uint32_t pad_;
};

enum : uint32_t { BIT_MM = 0, BIT_DM = 1, BIT_IM = 2, BIT_xM = 3, BIT_MD = 4, BIT_xD = 4, BIT_MI = 8, BIT_xI = 8 };
constexpr float MINUS_INFINITY = -1e37f;   // MU/intmath.h:46

// scratch of one job: four rows of lb + 1 floats, then (la + 1) x (lb + 1) trace bytes
inline uint64_t float_job_bytes(uint32_t la, uint32_t lb)
{
    const uint64_t rows = ((uint64_t)16 * (lb + 1) + 15) & ~15ull;
    return rows + ((((uint64_t)la + 1) * (lb + 1) + 15) & ~15ull);
}

// MU/profilefrommsa.cpp:279-292 with the fractional counts above, one row A
// class and one column B class, each float step rounded as the reference
// rounds it
Score6 make_score6()
{
#pragma clang fp contract(off)
    static const float NUC_SP[4][4] = {
        {91 + 60.f, -114 + 60.f, -31 + 60.f, -123 + 60.f},
        {-114 + 60.f, 100 + 60.f, -125 + 60.f, -31 + 60.f},
        {-31 + 60.f, -125 + 60.f, 100 + 60.f, -114 + 60.f},
        {-123 + 60.f, -31 + 60.f, -114 + 60.f, 91 + 60.f},
    };
    const float w = 1.0f;   // the weight of the one sequence of a profile
    float cnt[DP_NCLASS][4] = {};
    for (int c = 0; c < 4; ++c) cnt[c][c] = w;
    cnt[4][2] += w / 2;     // 'X': G, then A
    cnt[4][0] += w / 2;
    const float f = w / 20; // any other wildcard
    for (int l = 0; l < 4; ++l) cnt[5][l] += f;
    float aa[DP_NCLASS][4];
    int order[DP_NCLASS][4];
    for (int c = 0; c < DP_NCLASS; ++c) {
        for (int i = 0; i < 4; ++i) {
            volatile float sum = 0.0f;
            for (int j = 0; j < 4; ++j) {
                volatile float t = cnt[c][j] * NUC_SP[i][j];
                sum = sum + t;
            }
            aa[c][i] = sum;
        }
        for (int i = 0; i < 4; ++i) order[c][i] = i;
        bool any = true;
        while (any) {
            any = false;
            for (int n = 0; n < 3; ++n)
                if (cnt[c][order[c][n]] < cnt[c][order[c][n + 1]]) {
                    std::swap(order[c][n], order[c][n + 1]);
                    any = true;
                }
        }
    }
    Score6 s;
    for (int a = 0; a < DP_NCLASS; ++a)
        for (int b = 0; b < DP_NCLASS; ++b) {
            volatile float score = 0.0f;
            for (int n = 0; n < 4; ++n) {
                const int l = order[a][n];
                const float fc = cnt[a][l];
                if (fc == 0) break;
                volatile float t = fc * aa[b][l];
                score = score + t;
            }
            s.s[a][b] = score - 0.0f;   // g_scoreCenter is 0 for DNA
        }
    return s;
}

const Score6 &score6()
{
    static const Score6 s = make_score6();
    return s;
}

// MU/termgaps.cpp with TermGaps Half: the first position opens and the last
// closes at -0.0 (0 times -1), every other at half the -400 open
__device__ __forceinline__ float gap_open(uint32_t i) { return i == 0 ? -0.0f : -200.0f; }
__device__ __forceinline__ float gap_close(uint32_t i, uint32_t L) { return (L > 1 && i == L - 1) ? -0.0f : -200.0f; }

__device__ __forceinline__ int xchar(uint32_t bits, int t)
{
    if (t == 'M') {
        const uint32_t v = bits & BIT_xM;
        return v == BIT_MM ? 'M' : v == BIT_DM ? 'D' : v == BIT_IM ? 'I' : '?';
    }
    if (t == 'D') return (bits & BIT_xD) == BIT_MD ? 'M' : 'D';
    return (bits & BIT_xI) == BIT_MI ? 'M' : 'I';
}

// One thread per job: NWSmall's rows in float (MU/nwsmall.cpp:500-670, the
// macros RECURSE_D 68-86, RECURSE_I 91-106, RECURSE_M 111-142), then the
// trace-back; the path is written at the end of the job's slot, as the int32
// trace-back writes it, and dev gets ceil(max |i*lb - j*la| / la).
__global__ __launch_bounds__(64) void dp_float_kernel(const FloatJob *__restrict__ jobs, int njobs,
                                                      const uint8_t *__restrict__ A, const uint8_t *__restrict__ B,
                                                      uint8_t *__restrict__ scr, Score6 S,
                                                      uint8_t *__restrict__ paths, mh_dp_result *__restrict__ res,
                                                      uint32_t *__restrict__ dev)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= njobs) return;
    const FloatJob jb = jobs[x];
    const uint32_t LA = jb.la, LB = jb.lb, PCB = LB + 1;
    const uint8_t *pa = A + jb.a_off, *pb = B + jb.b_off;
    float *MPrev = reinterpret_cast<float *>(scr + jb.scr_off);
    float *MCurr = MPrev + PCB, *MNext = MCurr + PCB, *DRow = MNext + PCB;
    uint8_t *TB = scr + jb.scr_off + (((uint64_t)16 * PCB + 15) & ~15ull);
    const uint64_t tb_n = ((uint64_t)LA + 1) * PCB;
    for (uint64_t q = 0; q < tb_n; ++q) TB[q] = 0;
#define TBC(i, j) TB[(uint64_t)(i) * PCB + (j)]
#define SC(a, b) (S.s[pa[a]][pb[b]])
    const float e = 0.0f;   // gap extend on this path (MU/params.cpp:303-304)
    float Iij = gap_open(0);
    for (uint32_t j = 2; j <= LB; ++j) Iij += e;
    for (uint32_t j = 0; j <= LB; ++j) DRow[j] = MINUS_INFINITY;
    MPrev[0] = 0;
    for (uint32_t j = 1; j <= LB; ++j) MPrev[j] = MINUS_INFINITY;
    MCurr[0] = MINUS_INFINITY;
    MCurr[1] = SC(0, 0);
    TBC(1, 1) = (uint8_t)((TBC(1, 1) & ~BIT_xM) | BIT_MM);
    for (uint32_t j = 2; j <= LB; ++j) {
        MCurr[j] = SC(0, j - 1) + gap_open(0) + (float)(j - 2) * e + gap_close(j - 2, LB);
        TBC(1, j) = (uint8_t)((TBC(1, j) & ~BIT_xM) | BIT_IM);
    }
    for (uint32_t i = 1; i < LA; ++i) {
        Iij = MINUS_INFINITY;
        DRow[0] = gap_open(0) + (float)(i - 1) * e;
        MCurr[0] = MINUS_INFINITY;
        if (i == 1) {
            MCurr[1] = SC(0, 0);
            TBC(i, 1) = (uint8_t)((TBC(i, 1) & ~BIT_xM) | BIT_MM);
        } else {
            MCurr[1] = SC(i - 1, 0) + gap_open(0) + (float)(i - 2) * e + gap_close(i - 2, LA);
            TBC(i, 1) = (uint8_t)((TBC(i, 1) & ~BIT_xM) | BIT_DM);
        }
        for (uint32_t j = 1; j < LB; ++j) MNext[j + 1] = SC(i, j);
        for (uint32_t j = 1; j <= LB; ++j) {
            {   // RECURSE_D
                const float DD = DRow[j] + e, MD = MPrev[j] + gap_open(i - 1);
                if (DD > MD) DRow[j] = DD;
                else {
                    DRow[j] = MD;
                    TBC(i, j) = (uint8_t)((TBC(i, j) & ~BIT_xD) | BIT_MD);
                }
            }
            {   // RECURSE_I
                Iij += e;
                const float MI = MCurr[j - 1] + gap_open(j - 1);
                if (MI >= Iij) {
                    Iij = MI;
                    TBC(i, j) = (uint8_t)((TBC(i, j) & ~BIT_xI) | BIT_MI);
                }
            }
            if (j == LB) break;
            {   // RECURSE_M
                const float DM = DRow[j] + gap_close(i - 1, LA), IM = Iij + gap_close(j - 1, LB), MM = MCurr[j];
                uint8_t t = (uint8_t)(TBC(i + 1, j + 1) & ~BIT_xM);
                if (MM >= DM && MM >= IM) {
                    MNext[j + 1] += MM;
                    t |= BIT_MM;
                } else if (DM >= MM && DM >= IM) {
                    MNext[j + 1] += DM;
                    t |= BIT_DM;
                } else {
                    MNext[j + 1] += IM;
                    t |= BIT_IM;
                }
                TBC(i + 1, j + 1) = t;
            }
        }
        float *r = MPrev;
        MPrev = MCurr;
        MCurr = MNext;
        MNext = r;
    }
    // i == LA (MU/nwsmall.cpp:616-637)
    MCurr[0] = MINUS_INFINITY;
    if (LA > 1) MCurr[1] = SC(LA - 1, 0) + (float)(LA - 2) * e + gap_open(0) + gap_close(LA - 2, LA);
    else MCurr[1] = SC(LA - 1, 0) + gap_open(0) + gap_close(0, LA);
    TBC(LA, 1) = (uint8_t)((TBC(LA, 1) & ~BIT_xM) | BIT_DM);
    DRow[0] = MINUS_INFINITY;
    for (uint32_t j = 1; j <= LB; ++j) {
        const float DD = DRow[j] + e, MD = MPrev[j] + gap_open(LA - 1);
        if (DD > MD) DRow[j] = DD;
        else {
            DRow[j] = MD;
            TBC(LA, j) = (uint8_t)((TBC(LA, j) & ~BIT_xD) | BIT_MD);
        }
    }
    Iij = MINUS_INFINITY;
    for (uint32_t j = 1; j <= LB; ++j) {
        Iij += e;
        const float MI = MCurr[j - 1] + gap_open(j - 1);
        if (MI >= Iij) {
            Iij = MI;
            TBC(LA, j) = (uint8_t)((TBC(LA, j) & ~BIT_xI) | BIT_MI);
        }
    }
    float Score = MCurr[LB];
    int type = 'M';
    if (DRow[LB] > Score) {
        Score = DRow[LB];
        type = 'D';
    }
    if (Iij > Score) {
        Score = Iij;
        type = 'I';
    }
    // BitTraceBack from (LA, LB); the path fills the slot from its end
    uint8_t *slot = paths + jb.path_off;
    int64_t pos = (int64_t)LA + LB, dnum = 0, maxdev = 0;
    uint32_t i = LA, j = LB;
    bool ok = true;
    for (;;) {
        slot[--pos] = (uint8_t)type;
        const int nx = xchar(TBC(i, j), type);
        if (type == 'M') {
            --i;
            --j;
            dnum += (int64_t)LA - LB;
        } else if (type == 'D') {
            --i;
            dnum -= LB;
        } else {
            --j;
            dnum += LA;
        }
        maxdev = max(maxdev, dnum < 0 ? -dnum : dnum);
        if (i == 0 && j == 0) break;
        if (pos == 0 || nx == '?' || (nx == 'M' && (i == 0 || j == 0)) || (nx == 'D' && i == 0) ||
            (nx == 'I' && j == 0)) {
            ok = false;
            break;
        }
        type = nx;
    }
#undef TBC
#undef SC
    mh_dp_result r;
    r.score = (int32_t)rintf(Score);
    r.path_len = ok ? (uint32_t)((int64_t)LA + LB - pos) : 0u;
    r.path_start = jb.path_off + (uint64_t)pos;
    res[jb.user] = r;
    if (dev) dev[jb.user] = ok ? (uint32_t)((maxdev + LA - 1) / LA) : 0xFFFFFFFFu;
}

// flag[j] = the largest operand code of job j (one wavefront per job)
__global__ __launch_bounds__(64) void code_max_kernel(const mh_dp_job *__restrict__ jobs, const uint8_t *__restrict__ A,
                                                      const uint8_t *__restrict__ B, uint8_t *__restrict__ flag)
{
    const mh_dp_job jb = jobs[blockIdx.x];
    uint32_t m = 0;
    for (uint32_t p = threadIdx.x; p < jb.a_len; p += 64) m = max(m, (uint32_t)A[jb.a_off + p]);
    for (uint32_t p = threadIdx.x; p < jb.b_len; p += 64) m = max(m, (uint32_t)B[jb.b_off + p]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
    if (threadIdx.x == 0) flag[blockIdx.x] = (uint8_t)min(m, 255u);
}

}  // namespace

int dp_code_max(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs, uint64_t njobs,
                const FixedScratch *fx, std::vector<uint8_t> &out)
{
    out.assign(njobs, 0);
    if (!njobs) return MH_OK;
    const uint64_t jbytes = (njobs * sizeof(mh_dp_job) + 255) & ~255ull;
    mh_dp_job *d_jobs;
    uint8_t *d_flag;
    if (fx) {
        MH_CHECK(jbytes + njobs + 256 <= fx->bytes, MH_ERR_INVALID,
                 "too many DP jobs in one call for the " + std::to_string(fx->bytes >> 20) +
                     " MiB scratch of the no-allocation entry point");
        d_jobs = reinterpret_cast<mh_dp_job *>(fx->ptr);
        d_flag = fx->ptr + jbytes;
    } else {
        int rc;
        if ((rc = ctx->dp_wild.ensure(jbytes + njobs + 256)) != MH_OK) return rc;
        d_jobs = ctx->dp_wild.as<mh_dp_job>();
        d_flag = ctx->dp_wild.as<uint8_t>() + jbytes;
    }
    MH_HIP(hipMemcpyAsync(d_jobs, jobs, njobs * sizeof(mh_dp_job), hipMemcpyHostToDevice, ctx->stream));
    for (uint64_t first = 0; first < njobs; first += 0x7FFFFFFFull) {
        const uint64_t n = std::min<uint64_t>(njobs - first, 0x7FFFFFFFull);
        hipLaunchKernelGGL(code_max_kernel, dim3((unsigned)n), dim3(64), 0, ctx->stream, d_jobs + first, d_a, d_b,
                           d_flag + first);
    }
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(out.data(), d_flag, njobs, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int dp_run_float(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
                 const std::vector<uint64_t> &slot, const std::vector<uint32_t> &which, uint8_t *d_paths,
                 mh_dp_result *d_results, uint32_t *d_dev, uint64_t *cells, const FixedScratch *fx)
{
    if (which.empty()) return MH_OK;
    const Score6 &S = score6();
    const uint64_t jt = sizeof(FloatJob);
    auto need = [&](uint64_t cnt, uint64_t bytes) { return ((cnt * jt + 255) & ~255ull) + bytes + 256; };
    uint64_t budget;
    if (fx) {
        budget = fx->bytes;
        for (uint32_t j : which)
            MH_CHECK(need(1, float_job_bytes(jobs[j].a_len, jobs[j].b_len)) <= budget, MH_ERR_INVALID,
                     "DP job " + std::to_string(j) + " (" + std::to_string(jobs[j].a_len) + " x " +
                         std::to_string(jobs[j].b_len) + ", letters other than A,C,G,T) needs more than the " +
                         std::to_string(fx->bytes >> 20) + " MiB scratch of the no-allocation entry point");
    } else {
        size_t free_b = 0, total_b = 0;
        MH_HIP(hipMemGetInfo(&free_b, &total_b));
        budget = (uint64_t)((free_b + ctx->dp_fl.bytes) * 0.5);
    }
    std::vector<FloatJob> hj;
    size_t first = 0;
    while (first < which.size()) {
        uint64_t bytes = 0;
        size_t last = first;
        hj.clear();
        while (last < which.size()) {
            const mh_dp_job &q = jobs[which[last]];
            const uint64_t b = float_job_bytes(q.a_len, q.b_len);
            if (last > first && need(last - first + 1, bytes + b) > budget) break;
            FloatJob f;
            f.a_off = q.a_off;
            f.b_off = q.b_off;
            f.path_off = slot[which[last]];
            f.scr_off = bytes;
            f.la = q.a_len;
            f.lb = q.b_len;
            f.user = which[last];
            f.pad_ = 0;
            hj.push_back(f);
            bytes += b;
            if (cells) *cells += (uint64_t)q.a_len * q.b_len;
            ++last;
        }
        const uint64_t n = last - first, tbytes = (n * jt + 255) & ~255ull;
        uint8_t *base;
        if (fx) {
            base = fx->ptr;
        } else {
            int rc;
            if ((rc = ctx->dp_fl.ensure(tbytes + bytes + 256)) != MH_OK) return rc;
            base = ctx->dp_fl.as<uint8_t>();
        }
        FloatJob *d_jobs = reinterpret_cast<FloatJob *>(base);
        MH_HIP(hipMemcpyAsync(d_jobs, hj.data(), n * jt, hipMemcpyHostToDevice, ctx->stream));
        {
            TimedLaunch tl(ctx, "dp_float");
            hipLaunchKernelGGL(dp_float_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, ctx->stream, d_jobs,
                               (int)n, d_a, d_b, base + tbytes, S, d_paths, d_results, d_dev);
            MH_HIP(hipGetLastError());
        }
        // the host job table is reused by the next round
        MH_HIP(hipStreamSynchronize(ctx->stream));
        first = last;
    }
    return MH_OK;
}

void dp_score6(float *out)
{
    std::memcpy(out, score6().s, sizeof(Score6));
}

}  // namespace mh
