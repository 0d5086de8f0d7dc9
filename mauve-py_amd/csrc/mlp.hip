// mlp.hip -- the learned band: per-gap features, a 272-256-128-1 MLP with
// bf16 weights and fp32 accumulation (forward, backward, Adam), and the band
// prediction that feeds the banded DP.  No counterpart in the reference,
// whose NWSmall always fills the whole matrix.
//
// Numerics (the CPU oracle restates exactly these roundings):
//   x_bf = bf16(x); W_bf = bf16(W) for W1, W2, w3 (master weights fp32)
//   h1 = bf16(relu(x_bf W1_bf^T + b1)); h2 = bf16(relu(h1 W2_bf^T + b2))
//   y  = h2 . w3_bf + b3                                   (fp32)
//   loss = sum (y - t)^2 / (2B);  g = bf16((y - t) / B)
//   dz2 = bf16(g w3_bf) where h2 > 0, else 0;  dw3 = g^T h2;  db3 = sum g
//   dW2 = dz2^T h1;  db2 = colsum dz2
//   dz1 = bf16(dz2 W2_bf) where h1 > 0, else 0
//   dW1 = dz1^T x_bf;  db1 = colsum dz1
// bf16 rounding is round-to-nearest-even; products of two bf16 values are
// exact in fp32; every sum accumulates in fp32.
#include "mh_internal.hpp"
#include <cmath>
#include <cstring>
#include <random>

namespace mh {
namespace {

constexpr int NIN = MH_MLP_IN, NH1 = MH_MLP_H1, NH2 = MH_MLP_H2;
// flat parameter layout (fp32): W1 [NH1][NIN], b1, W2 [NH2][NH1], b2, w3 [NH2], b3
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NH1 * NIN;
constexpr int OFF_W2 = OFF_B1 + NH1;
constexpr int OFF_B2 = OFF_W2 + NH2 * NH1;
constexpr int OFF_W3 = OFF_B2 + NH2;
constexpr int OFF_B3 = OFF_W3 + NH2;
constexpr int NPARAM = OFF_B3 + 1;
static_assert(NPARAM == MH_MLP_NPARAMS, "parameter count");
// bf16 copies: W1, W2, w3 back to back
constexpr int BOFF_W1 = 0, BOFF_W2 = NH1 * NIN, BOFF_W3 = BOFF_W2 + NH2 * NH1, NBF = BOFF_W3 + NH2;

__host__ __device__ inline uint16_t to_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float from_bf16(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---- features ----------------------------------------------------------------
//
// Per job (A of length la, B of length lb):
//   f[0] = log2(la)/16, f[1] = log2(lb)/16, f[2] = log2(1 + |la - lb|)/16
//   f[3] = sum_k |cA[k] - cB[k]| / max(1, nA + nB)       (4-mer L1 distance)
//   f[4 + k] = clamp((cA[k] - cB[k]) * 512 / max(512, nA + nB), -8, 8) / 8
//   f[260 + 6*side + s], s = 0..5: flank seed density at scale w = 16 << s:
//       the fraction of the 8-mers of A starting at distance p < w from the
//       gap's left (side 0) or right (side 1) end that occur in B within 8
//       positions of the corner-to-corner line (B start floor(p*lb/la) from
//       the same end, +-8); 0 when no 8-mer fits.
// cA, cB are 4-mer counts (codes 0..3 packed 2 bits per letter, first letter
// high), nA = max(0, la - 3), nB likewise.

struct FeatJob {
    uint64_t a_off, b_off;
    uint32_t la, lb;
};

__device__ inline uint32_t pack8(const uint8_t *s)
{
    uint32_t v = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) v = (v << 2) | (s[q] & 3u);
    return v;
}

__global__ __launch_bounds__(256) void features_kernel(const uint8_t *__restrict__ A, const uint8_t *__restrict__ B,
                                                       const mh_dp_job *__restrict__ jobs, uint64_t njobs,
                                                       float *__restrict__ feat)
{
    __shared__ int ha[256], hb[256];
    __shared__ int hits[2][6], valid[2][6];
    __shared__ int l1part[256];
    const int t = threadIdx.x;
    const uint64_t jx = blockIdx.x;
    if (jx >= njobs) return;
    const mh_dp_job jb = jobs[jx];
    const int la = (int)jb.a_len, lb = (int)jb.b_len;
    const uint8_t *a = A + jb.a_off, *b = B + jb.b_off;
    ha[t] = 0;
    hb[t] = 0;
    if (t < 12) {
        hits[t / 6][t % 6] = 0;
        valid[t / 6][t % 6] = 0;
    }
    __syncthreads();
    for (int p = t; p + 4 <= la; p += 256)
        atomicAdd(&ha[((a[p] & 3) << 6) | ((a[p + 1] & 3) << 4) | ((a[p + 2] & 3) << 2) | (a[p + 3] & 3)], 1);
    for (int p = t; p + 4 <= lb; p += 256)
        atomicAdd(&hb[((b[p] & 3) << 6) | ((b[p + 1] & 3) << 4) | ((b[p + 2] & 3) << 2) | (b[p + 3] & 3)], 1);
    // flank densities: thread t takes p = t and t + 256 on both sides
    if (la >= 8 && lb >= 8) {
        for (int side = 0; side < 2; ++side) {
            for (int p = t; p < 512; p += 256) {
                if (p + 8 > la) continue;
                const uint32_t a8 = pack8(side == 0 ? a + p : a + (la - 8 - p));
                const int64_t qc = ((int64_t)p * lb) / la;
                bool hit = false;
                for (int64_t q = qc - 8; q <= qc + 8; ++q) {
                    if (q < 0 || q + 8 > lb) continue;
                    const uint32_t b8 = pack8(side == 0 ? b + q : b + (lb - 8 - q));
                    hit |= (b8 == a8);
                }
#pragma unroll
                for (int s = 0; s < 6; ++s)
                    if (p < (16 << s)) {
                        atomicAdd(&valid[side][s], 1);
                        if (hit) atomicAdd(&hits[side][s], 1);
                    }
            }
        }
    }
    __syncthreads();
    const int na = la >= 3 ? la - 3 : 0, nb = lb >= 3 ? lb - 3 : 0;
    const int diff = ha[t] - hb[t];
    l1part[t] = diff < 0 ? -diff : diff;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) l1part[t] += l1part[t + off];
        __syncthreads();
    }
    float *f = feat + jx * (uint64_t)NIN;
    const int tot = na + nb;
    float r = (float)diff * 512.0f / (float)(tot > 512 ? tot : 512);
    r = r < -8.0f ? -8.0f : (r > 8.0f ? 8.0f : r);
    f[4 + t] = r / 8.0f;
    if (t == 0) {
        f[0] = log2f((float)la) / 16.0f;
        f[1] = log2f((float)lb) / 16.0f;
        f[2] = log2f(1.0f + (float)(la > lb ? la - lb : lb - la)) / 16.0f;
        f[3] = (float)l1part[0] / (float)(tot > 1 ? tot : 1);
    }
    if (t < 12) {
        const int side = t / 6, s = t % 6;
        const int v = valid[side][s];
        f[260 + t] = v ? (float)hits[side][s] / (float)v : 0.0f;
    }
}

// ---- GEMM on MFMA ---------------------------------------------------------
//
// C[M x N] = A[M x K] . B[K x N], bf16 operands (fp32 sources are rounded on
// load), fp32 accumulation with v_mfma_f32_16x16x16bf16.  A workgroup of 4
// waves computes a 64x64 tile, each wave a 32x32 quarter as 2x2 MFMA blocks;
// K advances 16 at a time through LDS.  Operand element (r, c) is at
// p[(idx ? idx[r] : r) * s0 + c * s1]: strides give the transposes, idx
// gathers rows (a batch of examples).

struct Operand {
    const void *p;
    int64_t s0, s1;
    const uint32_t *idx;
    int f32;
};

enum { EPI_F32 = 0, EPI_BIAS_RELU_BF16 = 1, EPI_MASK_BF16 = 2 };

struct Epilogue {
    int mode;
    const float *bias;      // per column (EPI_BIAS_RELU_BF16)
    float *c32;             // EPI_F32
    uint16_t *cbf;          // bf16 outputs
    int64_t ldc;
    const uint16_t *mask;   // EPI_MASK_BF16: keep where mask > 0
    int64_t ldm;
};

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t load_op(const Operand &o, int64_t r, int64_t c)
{
    const int64_t rr = o.idx ? (int64_t)o.idx[r] : r;
    const int64_t x = rr * o.s0 + c * o.s1;
    return o.f32 ? to_bf16(static_cast<const float *>(o.p)[x]) : static_cast<const uint16_t *>(o.p)[x];
}

constexpr int GT = 64, GK = 16, GKP = GK + 4;   // tile, k step, padded LDS row

__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, Operand A, Operand Bo, Epilogue E)
{
    __shared__ uint16_t As[GT][GKP];   // [m][k]
    __shared__ uint16_t Bs[GT][GKP];   // [n][k]
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += GK) {
        // 64x16 of A and 16x64 of B: 4 elements each per thread
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;          // 0..1023
            const int r = e >> 4, c = e & 15;   // tile row (m or n), k
            const int gm = m0 + r, gn = n0 + r, gk = k0 + c;
            As[r][c] = (gm < M && gk < K) ? load_op(A, gm, gk) : (uint16_t)0;
            Bs[r][c] = (gn < N && gk < K) ? load_op(Bo, gk, gn) : (uint16_t)0;
        }
        __syncthreads();
        const int li = lane & 15, lk = (lane >> 4) * 4;
        bf16x4 av[2], bv[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const uint16_t *pa = &As[wm + 16 * x + li][lk];
            const uint16_t *pb = &Bs[wn + 16 * x + li][lk];
            av[x] = bf16x4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
            bv[x] = bf16x4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
                acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[x], bv[y], acc[x][y], 0, 0, 0);
        __syncthreads();
    }
    // D[i][j] of a 16x16 block: lane holds column j = lane%16, rows 4*(lane/16) + 0..3
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gm = m0 + wm + 16 * x + 4 * (lane >> 4) + q;
                const int gn = n0 + wn + 16 * y + (lane & 15);
                if (gm >= M || gn >= N) continue;
                const float v = acc[x][y][q];
                const int64_t o = (int64_t)gm * E.ldc + gn;
                if (E.mode == EPI_F32) {
                    E.c32[o] = v;
                } else if (E.mode == EPI_BIAS_RELU_BF16) {
                    const float z = v + E.bias[gn];
                    E.cbf[o] = to_bf16(z > 0.f ? z : 0.f);
                } else {
                    const uint16_t mk = E.mask[(int64_t)gm * E.ldm + gn];
                    const bool pos = mk != 0 && !(mk & 0x8000u);
                    E.cbf[o] = pos ? to_bf16(v) : (uint16_t)0;
                }
            }
}

// y[b] = b3 + sum_k h2[b][k] * w3[k]  (fp32, k ascending)
__global__ void head_kernel(const uint16_t *__restrict__ h2, int n, const uint16_t *__restrict__ w3bf,
                            const float *__restrict__ b3, float *__restrict__ y)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float s = 0.f;
    for (int k = 0; k < NH2; ++k) s += from_bf16(h2[(int64_t)r * NH2 + k]) * from_bf16(w3bf[k]);
    y[r] = s + b3[0];
}

// one workgroup: g = bf16((y - t)/n), loss = sum (y - t)^2 / (2n)
__global__ __launch_bounds__(256) void loss_kernel(const float *__restrict__ y, const float *__restrict__ target,
                                                   const uint32_t *__restrict__ idx, int n,
                                                   uint16_t *__restrict__ g, float *__restrict__ loss)
{
    __shared__ float part[256];
    const int t = threadIdx.x;
    float s = 0.f;
    for (int r = t; r < n; r += 256) {
        const float d = y[r] - target[idx ? idx[r] : (uint32_t)r];
        g[r] = to_bf16(d / (float)n);
        s += d * d;
    }
    part[t] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) part[t] += part[t + off];
        __syncthreads();
    }
    if (t == 0) loss[0] = part[0] / (2.0f * (float)n);
}

// dz2[b][k] = bf16(g[b] * w3[k]) where h2 > 0
__global__ void head_back_kernel(const uint16_t *__restrict__ g, const uint16_t *__restrict__ h2, int n,
                                 const uint16_t *__restrict__ w3bf, uint16_t *__restrict__ dz2)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * NH2) return;
    const int r = (int)(e / NH2), k = (int)(e % NH2);
    const uint16_t h = h2[e];
    const bool pos = h != 0 && !(h & 0x8000u);
    dz2[e] = pos ? to_bf16(from_bf16(g[r]) * from_bf16(w3bf[k])) : (uint16_t)0;
}

// out[c] = sum_r v[r][c] (* w[r] when w is given), r ascending, fp32
__global__ void colsum_kernel(const uint16_t *__restrict__ v, int rows, int cols, const uint16_t *__restrict__ w,
                              float *__restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) {
        const float x = from_bf16(v[(int64_t)r * cols + c]);
        s += w ? x * from_bf16(w[r]) : x;
    }
    out[c] = s;
}

__global__ void sum_bf16_kernel(const uint16_t *__restrict__ v, int n, float *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float s = 0.f;
    for (int r = 0; r < n; ++r) s += from_bf16(v[r]);
    out[0] = s;
}

// Adam on the fp32 master weights, then the bf16 copies
__global__ void adam_kernel(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, uint16_t *__restrict__ wbf, float lr, float b1, float b2,
                            float c1, float c2, float eps, float wd)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= NPARAM) return;
    const float gx = g[x];
    const float mx = b1 * m[x] + (1.f - b1) * gx;
    const float vx = b2 * v[x] + (1.f - b2) * gx * gx;
    m[x] = mx;
    v[x] = vx;
    float px = p[x];
    px -= lr * ((mx / c1) / (sqrtf(vx / c2) + eps) + wd * px);
    p[x] = px;
    int bx = -1;
    if (x < OFF_B1) bx = BOFF_W1 + x;
    else if (x >= OFF_W2 && x < OFF_B2) bx = BOFF_W2 + (x - OFF_W2);
    else if (x >= OFF_W3 && x < OFF_B3) bx = BOFF_W3 + (x - OFF_W3);
    if (bx >= 0) wbf[bx] = to_bf16(px);
}

__global__ void refresh_bf16_kernel(const float *__restrict__ p, uint16_t *__restrict__ wbf)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= NBF) return;
    int px;
    if (x < BOFF_W2) px = OFF_W1 + x;
    else if (x < BOFF_W3) px = OFF_W2 + (x - BOFF_W2);
    else px = OFF_W3 + (x - BOFF_W3);
    wbf[x] = to_bf16(p[px]);
}

// band = max(min_band, ceil(2^(y + delta)))
__global__ void band_kernel(const float *__restrict__ y, uint64_t n, float delta, uint32_t min_band,
                            uint32_t *__restrict__ band)
{
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    float e = y[x] + delta;
    e = e > 30.f ? 30.f : e;
    const uint32_t w = (uint32_t)ceilf(exp2f(e));
    band[x] = w > min_band ? w : min_band;
}

}  // namespace
}  // namespace mh

using namespace mh;

struct mh_mlp {
    mh_context *ctx = nullptr;
    DevBuf p, m, v, g, wbf;
    DevBuf h1, h2, y, gy, dz1, dz2, loss, idx;
    uint64_t cap = 0;     // batch rows the activation buffers hold
    uint64_t step = 0;
    uint64_t seed = 0;
};

namespace {

int mlp_reserve(mh_mlp *n, uint64_t rows)
{
    if (rows <= n->cap) return MH_OK;
    int rc;
    if ((rc = n->h1.ensure(rows * NH1 * 2 + 64)) != MH_OK) return rc;
    if ((rc = n->h2.ensure(rows * NH2 * 2 + 64)) != MH_OK) return rc;
    if ((rc = n->y.ensure(rows * 4 + 64)) != MH_OK) return rc;
    if ((rc = n->gy.ensure(rows * 2 + 64)) != MH_OK) return rc;
    if ((rc = n->dz1.ensure(rows * NH1 * 2 + 64)) != MH_OK) return rc;
    if ((rc = n->dz2.ensure(rows * NH2 * 2 + 64)) != MH_OK) return rc;
    n->cap = rows;
    return MH_OK;
}

void launch_gemm(hipStream_t st, int M, int N, int K, const Operand &A, const Operand &B, const Epilogue &E)
{
    hipLaunchKernelGGL(gemm_kernel, dim3((M + GT - 1) / GT, (N + GT - 1) / GT), dim3(256), 0, st, M, N, K, A, B, E);
}

// forward over `rows` examples (features x rows, or x[idx[r]]): h1, h2, y
void mlp_forward_rows(mh_mlp *n, const float *x, const uint32_t *idx, int rows)
{
    hipStream_t st = n->ctx->stream;
    const float *p = n->p.as<float>();
    const uint16_t *wbf = n->wbf.as<uint16_t>();
    Operand X{x, NIN, 1, idx, 1};
    Operand W1t{wbf + BOFF_W1, 1, NIN, nullptr, 0};          // (k = in, n = out) at W1[out][in]
    Epilogue e1{EPI_BIAS_RELU_BF16, p + OFF_B1, nullptr, n->h1.as<uint16_t>(), NH1, nullptr, 0};
    launch_gemm(st, rows, NH1, NIN, X, W1t, e1);
    Operand H1{n->h1.as<uint16_t>(), NH1, 1, nullptr, 0};
    Operand W2t{wbf + BOFF_W2, 1, NH1, nullptr, 0};
    Epilogue e2{EPI_BIAS_RELU_BF16, p + OFF_B2, nullptr, n->h2.as<uint16_t>(), NH2, nullptr, 0};
    launch_gemm(st, rows, NH2, NH1, H1, W2t, e2);
    hipLaunchKernelGGL(head_kernel, dim3((rows + 255) / 256), dim3(256), 0, st, n->h2.as<uint16_t>(), rows,
                       wbf + BOFF_W3, p + OFF_B3, n->y.as<float>());
}

// loss and gradients of one batch into n->g (fp32, flat parameter layout)
void mlp_backward_rows(mh_mlp *n, const float *x, const float *target, const uint32_t *idx, int rows)
{
    hipStream_t st = n->ctx->stream;
    const uint16_t *wbf = n->wbf.as<uint16_t>();
    float *g = n->g.as<float>();
    uint16_t *gy = n->gy.as<uint16_t>(), *h1 = n->h1.as<uint16_t>(), *h2 = n->h2.as<uint16_t>();
    uint16_t *dz1 = n->dz1.as<uint16_t>(), *dz2 = n->dz2.as<uint16_t>();
    hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, st, n->y.as<float>(), target, idx, rows, gy,
                       n->loss.as<float>());
    // head: dw3 = g^T h2, db3 = sum g, dz2
    hipLaunchKernelGGL(colsum_kernel, dim3((NH2 + 127) / 128), dim3(128), 0, st, h2, rows, NH2, gy, g + OFF_W3);
    hipLaunchKernelGGL(sum_bf16_kernel, dim3(1), dim3(64), 0, st, gy, rows, g + OFF_B3);
    hipLaunchKernelGGL(head_back_kernel, dim3((unsigned)(((int64_t)rows * NH2 + 255) / 256)), dim3(256), 0, st, gy,
                       h2, rows, wbf + BOFF_W3, dz2);
    // layer 2: dW2 = dz2^T h1 (M = NH2, N = NH1, K = rows), db2, dz1
    {
        Operand A{dz2, 1, NH2, nullptr, 0};       // (m = out2, k = row) at dz2[row][out2]
        Operand B{h1, NH1, 1, nullptr, 0};        // (k = row, n = in2) at h1[row][in2]
        Epilogue e{EPI_F32, nullptr, g + OFF_W2, nullptr, NH1, nullptr, 0};
        launch_gemm(st, NH2, NH1, rows, A, B, e);
    }
    hipLaunchKernelGGL(colsum_kernel, dim3((NH2 + 127) / 128), dim3(128), 0, st, dz2, rows, NH2, nullptr,
                       g + OFF_B2);
    {
        Operand A{dz2, NH2, 1, nullptr, 0};       // (m = row, k = out2)
        Operand B{wbf + BOFF_W2, NH1, 1, nullptr, 0};   // (k = out2, n = in2) at W2[out2][in2]
        Epilogue e{EPI_MASK_BF16, nullptr, nullptr, dz1, NH1, h1, NH1};
        launch_gemm(st, rows, NH1, NH2, A, B, e);
    }
    // layer 1: dW1 = dz1^T x (M = NH1, N = NIN, K = rows), db1
    {
        Operand A{dz1, 1, NH1, nullptr, 0};
        Operand B{x, NIN, 1, idx, 1};             // (k = row, n = feature) at x[idx[row]][feature]
        Epilogue e{EPI_F32, nullptr, g + OFF_W1, nullptr, NIN, nullptr, 0};
        launch_gemm(st, NH1, NIN, rows, A, B, e);
    }
    hipLaunchKernelGGL(colsum_kernel, dim3((NH1 + 127) / 128), dim3(128), 0, st, dz1, rows, NH1, nullptr,
                       g + OFF_B1);
}

}  // namespace

extern "C" {

int mh_gap_features_device(mh_context *ctx, const uint8_t *d_a, const uint8_t *d_b, const mh_dp_job *jobs,
                           uint64_t njobs, float *d_feat)
{
    MH_CHECK(ctx, MH_ERR_INVALID, "null context");
    if (njobs == 0) return MH_OK;
    MH_CHECK(jobs && d_a && d_b && d_feat, MH_ERR_INVALID, "null argument");
    MH_CHECK(njobs < (1ull << 31), MH_ERR_INVALID, "too many jobs in one call");
    MH_HIP(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ctx->ml_jobs.ensure(njobs * sizeof(mh_dp_job) + 64)) != MH_OK) return rc;
    MH_HIP(hipMemcpyAsync(ctx->ml_jobs.ptr, jobs, njobs * sizeof(mh_dp_job), hipMemcpyHostToDevice, ctx->stream));
    {
        TimedLaunch tl(ctx, "band_features");
        hipLaunchKernelGGL(features_kernel, dim3((unsigned)njobs), dim3(256), 0, ctx->stream, d_a, d_b,
                           ctx->ml_jobs.as<mh_dp_job>(), njobs, d_feat);
        MH_HIP(hipGetLastError());
    }
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int mh_mlp_create(mh_context *ctx, uint64_t seed, mh_mlp **out)
{
    MH_CHECK(ctx && out, MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    mh_mlp *n = new mh_mlp();
    n->ctx = ctx;
    n->seed = seed;
    int rc;
    if ((rc = n->p.ensure(NPARAM * 4 + 64)) != MH_OK || (rc = n->m.ensure(NPARAM * 4 + 64)) != MH_OK ||
        (rc = n->v.ensure(NPARAM * 4 + 64)) != MH_OK || (rc = n->g.ensure(NPARAM * 4 + 64)) != MH_OK ||
        (rc = n->wbf.ensure(NBF * 2 + 64)) != MH_OK || (rc = n->loss.ensure(64)) != MH_OK) {
        delete n;
        return rc;
    }
    // He-uniform weights, zero biases (mt19937_64 from the seed)
    std::vector<float> h(NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_out, int fan_in) {
        const float lim = std::sqrt(6.0f / (float)fan_in);
        std::uniform_real_distribution<float> u(-lim, lim);
        for (int x = 0; x < fan_out * fan_in; ++x) h[off + x] = u(rng);
    };
    fill(OFF_W1, NH1, NIN);
    fill(OFF_W2, NH2, NH1);
    fill(OFF_W3, 1, NH2);
    *out = n;
    return mh_mlp_set_params(ctx, n, h.data());
}

void mh_mlp_free(mh_mlp *n)
{
    delete n;
}

int mh_mlp_set_params(mh_context *ctx, mh_mlp *n, const float *params)
{
    MH_CHECK(ctx && n && params, MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    MH_HIP(hipMemcpyAsync(n->p.ptr, params, NPARAM * 4, hipMemcpyHostToDevice, ctx->stream));
    MH_HIP(hipMemsetAsync(n->m.ptr, 0, NPARAM * 4, ctx->stream));
    MH_HIP(hipMemsetAsync(n->v.ptr, 0, NPARAM * 4, ctx->stream));
    hipLaunchKernelGGL(refresh_bf16_kernel, dim3((NBF + 255) / 256), dim3(256), 0, ctx->stream, n->p.as<float>(),
                       n->wbf.as<uint16_t>());
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(ctx->stream));
    n->step = 0;
    return MH_OK;
}

int mh_mlp_get_params(mh_context *ctx, const mh_mlp *n, float *params)
{
    MH_CHECK(ctx && n && params, MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    MH_HIP(hipMemcpyAsync(params, n->p.ptr, NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int mh_mlp_forward(mh_context *ctx, mh_mlp *n, const float *d_x, uint64_t rows, float *d_y)
{
    MH_CHECK(ctx && n, MH_ERR_INVALID, "null argument");
    if (rows == 0) return MH_OK;
    MH_CHECK(d_x && d_y, MH_ERR_INVALID, "null argument");
    MH_HIP(hipSetDevice(ctx->device));
    const uint64_t chunk = 65536;
    int rc;
    if ((rc = mlp_reserve(n, std::min<uint64_t>(rows, chunk))) != MH_OK) return rc;
    TimedLaunch tl(ctx, "band_mlp_forward");
    for (uint64_t r0 = 0; r0 < rows; r0 += chunk) {
        const int nr = (int)std::min<uint64_t>(chunk, rows - r0);
        mlp_forward_rows(n, d_x + r0 * NIN, nullptr, nr);
        MH_HIP(hipMemcpyAsync(d_y + r0, n->y.ptr, (size_t)nr * 4, hipMemcpyDeviceToDevice, ctx->stream));
    }
    MH_HIP(hipGetLastError());
    return MH_OK;
}

int mh_mlp_grad(mh_context *ctx, mh_mlp *n, const float *d_x, const float *d_t, uint64_t rows, float *grads,
                float *loss)
{
    MH_CHECK(ctx && n && d_x && d_t && grads, MH_ERR_INVALID, "null argument");
    MH_CHECK(rows > 0 && rows <= 65536, MH_ERR_INVALID, "gradient batch must hold 1..65536 rows");
    MH_HIP(hipSetDevice(ctx->device));
    int rc;
    if ((rc = mlp_reserve(n, rows)) != MH_OK) return rc;
    mlp_forward_rows(n, d_x, nullptr, (int)rows);
    mlp_backward_rows(n, d_x, d_t, nullptr, (int)rows);
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(grads, n->g.ptr, NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream));
    float l = 0.f;
    MH_HIP(hipMemcpyAsync(&l, n->loss.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    if (loss) *loss = l;
    return MH_OK;
}

int mh_mlp_train(mh_context *ctx, mh_mlp *n, const float *d_x, const float *d_t, uint64_t rows,
                 const mh_mlp_train_opts *opt, float *epoch_loss)
{
    MH_CHECK(ctx && n && d_x && d_t && opt, MH_ERR_INVALID, "null argument");
    MH_CHECK(rows > 0 && rows < (1ull << 32), MH_ERR_INVALID, "training set must hold 1..2^32-1 rows");
    MH_CHECK(opt->batch > 0 && opt->batch <= 65536 && opt->epochs > 0, MH_ERR_INVALID, "bad batch or epochs");
    MH_HIP(hipSetDevice(ctx->device));
    int rc;
    const uint32_t bs = opt->batch;
    if ((rc = mlp_reserve(n, bs)) != MH_OK) return rc;
    if ((rc = n->idx.ensure(rows * 4 + 64)) != MH_OK) return rc;
    std::vector<uint32_t> perm(rows);
    for (uint64_t r = 0; r < rows; ++r) perm[r] = (uint32_t)r;
    std::mt19937_64 rng(opt->seed);
    DevBuf lossbuf;
    const uint64_t nb = rows / bs;   // full batches only
    MH_CHECK(nb > 0, MH_ERR_INVALID, "training set smaller than one batch");
    if ((rc = lossbuf.ensure(nb * 4 + 64)) != MH_OK) return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    TimedLaunch tl(ctx, "band_mlp_train");
    for (uint32_t ep = 0; ep < opt->epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        MH_HIP(hipMemcpyAsync(n->idx.ptr, perm.data(), rows * 4, hipMemcpyHostToDevice, ctx->stream));
        for (uint64_t bi = 0; bi < nb; ++bi) {
            const uint32_t *idx = n->idx.as<uint32_t>() + bi * bs;
            mlp_forward_rows(n, d_x, idx, (int)bs);
            mlp_backward_rows(n, d_x, d_t, idx, (int)bs);
            ++n->step;
            // cosine decay of the learning rate over the run
            const double prog = ((double)ep * nb + bi) / ((double)opt->epochs * nb);
            const float lr = (float)(opt->lr * 0.5 * (1.0 + std::cos(3.14159265358979 * prog)));
            const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
            hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, ctx->stream,
                               n->p.as<float>(), n->m.as<float>(), n->v.as<float>(), n->g.as<float>(),
                               n->wbf.as<uint16_t>(), lr, b1, b2, c1, c2, eps, opt->weight_decay);
            MH_HIP(hipMemcpyAsync(lossbuf.as<float>() + bi, n->loss.ptr, 4, hipMemcpyDeviceToDevice, ctx->stream));
        }
        MH_HIP(hipGetLastError());
        if (epoch_loss) {
            std::vector<float> l(nb);
            MH_HIP(hipMemcpyAsync(l.data(), lossbuf.ptr, nb * 4, hipMemcpyDeviceToHost, ctx->stream));
            MH_HIP(hipStreamSynchronize(ctx->stream));
            double s = 0;
            for (float v : l) s += v;
            epoch_loss[ep] = (float)(s / nb);
        }
    }
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

int mh_band_predict(mh_context *ctx, mh_mlp *n, const float *d_feat, uint64_t rows, float delta,
                    uint32_t min_band, uint32_t *bands)
{
    MH_CHECK(ctx && n, MH_ERR_INVALID, "null argument");
    if (rows == 0) return MH_OK;
    MH_CHECK(d_feat && bands, MH_ERR_INVALID, "null argument");
    int rc;
    if ((rc = ctx->ml_y.ensure(rows * 8 + 64)) != MH_OK) return rc;
    float *y = ctx->ml_y.as<float>();
    uint32_t *b = reinterpret_cast<uint32_t *>(y + rows);
    if ((rc = mh_mlp_forward(ctx, n, d_feat, rows, y)) != MH_OK) return rc;
    hipLaunchKernelGGL(band_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, ctx->stream, y, rows, delta,
                       min_band, b);
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(bands, b, rows * 4, hipMemcpyDeviceToHost, ctx->stream));
    MH_HIP(hipStreamSynchronize(ctx->stream));
    return MH_OK;
}

}  // extern "C"
