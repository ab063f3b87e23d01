// mlp.hip -- ensemble MLP for the mouth-[Ca2+] surrogate: hand-written BF16
// MFMA GEMM (FP32 accumulation), fused bias/SiLU epilogues, backward pass and
// Adam, behind include/casim_surrogate.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/casim_surrogate.h"
#include "dyad_internal.hpp"

using namespace casim;

namespace casim {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

// float -> bf16 with round-to-nearest-even (finite inputs).
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------------------
// C[M x N] = op(A) op(B) (+ beta C), BF16 operands, FP32 accumulation.
// 64 x 64 tile per workgroup of four waves; each wave owns a 32 x 32 quarter
// as 2 x 2 v_mfma_f32_16x16x32_bf16 tiles.  Both LDS tiles keep k contiguous
// ([m][k] and [n][k]), so a lane's eight k-values of one MFMA operand are one
// 16-byte LDS read; A lanes supply row (lane % 16), B lanes column (lane % 16),
// and the accumulator lane holds rows 4 (lane / 16) .. +3 of column lane % 16.
constexpr int GBM = 64, GBN = 64, GBK = 32, GPAD = 8;

__global__ __launch_bounds__(256) void k_gemm(const uint16_t *__restrict__ A, int lda, int ta,
                                              const uint16_t *__restrict__ B, int ldb, int tb, float *__restrict__ C,
                                              int ldc, int M, int N, int K, float beta) {
    __shared__ __attribute__((aligned(16))) uint16_t As[GBM][GBK + GPAD];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[GBN][GBK + GPAD];
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    f32x4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += GBK) {
        if (!ta) {
            const int m = t >> 2, kk = (t & 3) * 8;
            const int gm = m0 + m;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kk + e;
                As[m][kk + e] = (gm < M && gk < K) ? A[(size_t)gm * lda + gk] : (uint16_t)0;
            }
        } else {
            const int kk = t >> 3, m = (t & 7) * 8;
            const int gk = k0 + kk;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gm = m0 + m + e;
                As[m + e][kk] = (gm < M && gk < K) ? A[(size_t)gk * lda + gm] : (uint16_t)0;
            }
        }
        if (!tb) {
            const int kk = t >> 3, n = (t & 7) * 8;
            const int gk = k0 + kk;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gn = n0 + n + e;
                Bs[n + e][kk] = (gn < N && gk < K) ? B[(size_t)gk * ldb + gn] : (uint16_t)0;
            }
        } else {
            const int n = t >> 2, kk = (t & 3) * 8;
            const int gn = n0 + n;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kk + e;
                Bs[n][kk + e] = (gn < N && gk < K) ? B[(size_t)gn * ldb + gk] : (uint16_t)0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t *>(&As[wm + i * 16 + (l & 15)][(l >> 4) * 8]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const bf16x8_t b = *reinterpret_cast<const bf16x8_t *>(&Bs[wn + j * 16 + (l & 15)][(l >> 4) * 8]);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + i * 16 + (l >> 4) * 4 + r;
                const int col = n0 + wn + j * 16 + (l & 15);
                if (row < M && col < N) {
                    float v = acc[i][j][r];
                    if (beta != 0.f) v += beta * C[(size_t)row * ldc + col];
                    C[(size_t)row * ldc + col] = v;
                }
            }
}

__device__ __forceinline__ float sigmoidf_(float z) { return 1.f / (1.f + __expf(-z)); }

// Z += b (per column); H = bf16(SiLU(Z)) or bf16(Z) for the last layer.
__global__ void k_bias_act(float *__restrict__ Z, const float *__restrict__ b, uint16_t *__restrict__ H, int rows,
                           int cols, int act) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const float z = Z[i] + b[i % cols];
    Z[i] = z;
    if (H) H[i] = f2bf(act ? z * sigmoidf_(z) : z);
}

// Inference of one ensemble member for IMR rows per workgroup, every layer in
// one launch: the normalised BF16 input (k_gather's arithmetic), each hidden
// layer as BF16 MFMA products with FP32 accumulation (k_gemm's operand layout
// and k order), bias, SiLU and BF16 rounding (k_bias_act's), and the linear
// output, with the activations in LDS.  The weights come transposed and
// zero-padded ([n][k], k_wt) so a lane's eight k-values of an MFMA operand are
// one 16-byte load.  Per output element the operations and their order are
// those of k_gather + k_gemm + k_bias_act, so the outputs are the same bits.
// Grid: (row tiles, members); sixteen waves, wave w computes the 16 columns
// 16w .. +15 of a layer, and a lane loads all of the layer's k-blocks of its
// weight row at once (at most IMAXW / 32 = 8) while the layer before runs:
// with few rows a workgroup is a latency chain of layers.
constexpr int IMR = 32, IMAXW = 256, IPADK = 8, IMAXM = 8, IMAXL = 8, IWAVES = IMAXW / 16;
struct InferArgs {
    const uint16_t *wt[IMAXM];  // member's transposed weights
    const float *p[IMAXM];      // member's FP32 parameters (biases at boff)
    int wtoff[IMAXL], boff[IMAXL], kp[IMAXL], fout[IMAXL];
    int nl, n_in;
    const float *mean, *istd;
};

__global__ void k_wt(const uint16_t *__restrict__ W, uint16_t *__restrict__ WT, int fin, int fout, int kp, int np) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)np * kp) return;
    const int n = (int)(i / kp), k = (int)(i % kp);
    WT[i] = (n < fout && k < fin) ? W[(size_t)k * fout + n] : (uint16_t)0;
}

__global__ __launch_bounds__(64 * IWAVES) void k_mlp_infer(const float *__restrict__ X, int rows,
                                                            float *__restrict__ out, int ostride, InferArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t H[2][IMR][IMAXW + IPADK];
    const int m0 = blockIdx.x * IMR, e = blockIdx.y;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const uint16_t *wt = a.wt[e];
    const float *pe = a.p[e];
    // every k-block of the lane's weight row for layer L, and its bias, issued
    // together (blocks past the layer's K repeat its last one, columns past
    // its width the last row; neither is used)
    auto fetch = [&](int L, bf16x8_t (&b)[IMAXW / 32], float &bias) {
        const int K = a.kp[L], N = a.fout[L], NP = (N + 15) & ~15, nk = K >> 5;
        const int col = min(w * 16 + (l & 15), NP - 1);
        const uint16_t *wrow = wt + a.wtoff[L] + (size_t)col * K + (l >> 4) * 8;
#pragma unroll
        for (int kb = 0; kb < IMAXW / 32; ++kb)
            b[kb] = *reinterpret_cast<const bf16x8_t *>(wrow + 32 * (kb < nk ? kb : nk - 1));
        bias = pe[a.boff[L] + min(col, N - 1)];
    };
    bf16x8_t bw[IMAXW / 32], bn[IMAXW / 32];
    float bias, bias_n;
    fetch(0, bw, bias);  // in flight while the input is staged
    const int kp0 = a.kp[0];
    for (int i = t; i < IMR * kp0; i += 64 * IWAVES) {
        const int r = i / kp0, f = i - r * kp0;
        uint16_t v = 0;
        if (m0 + r < rows && f < a.n_in) v = f2bf((X[(size_t)(m0 + r) * a.n_in + f] - a.mean[f]) * a.istd[f]);
        H[0][r][f] = v;
    }
    __syncthreads();
    for (int L = 0; L < a.nl; ++L) {
        const int src = L & 1, K = a.kp[L], N = a.fout[L], NP = (N + 15) & ~15;
        const bool last = L == a.nl - 1;
        const int n0 = w * 16;
        // the next layer's weights are in flight during this layer (the last
        // layer fetches itself again, so that the count of loads in flight is
        // the same every layer)
        fetch(last ? L : L + 1, bn, bias_n);
        if (n0 < NP) {
            const int nk = K >> 5;  // 32-wide k-blocks, at most 8
            f32x4_t acc[2] = {f32x4_t{0.f, 0.f, 0.f, 0.f}, f32x4_t{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
            for (int kb = 0; kb < IMAXW / 32; ++kb) {
                if (kb >= nk) break;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const bf16x8_t av =
                        *reinterpret_cast<const bf16x8_t *>(&H[src][i * 16 + (l & 15)][32 * kb + (l >> 4) * 8]);
                    acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bw[kb], acc[i], 0, 0, 0);
                }
            }
            const int col = n0 + (l & 15);
            if (col < N) {
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = i * 16 + (l >> 4) * 4 + r;
                        const float z = acc[i][r] + bias;
                        if (last) {
                            if (m0 + row < rows) out[(size_t)e * ostride + m0 + row] = z;
                        } else {
                            H[src ^ 1][row][col] = f2bf(z * sigmoidf_(z));
                        }
                    }
            }
        }
        __syncthreads();
#pragma unroll
        for (int kb = 0; kb < IMAXW / 32; ++kb) bw[kb] = bn[kb];
        bias = bias_n;
    }
}

// dZ = dH * SiLU'(Z), kept in FP32 (bias gradient) and BF16 (GEMM operand).
__global__ void k_silu_back(const float *__restrict__ dH, const float *__restrict__ Z, float *__restrict__ dZ,
                            uint16_t *__restrict__ dZb, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = Z[i];
    const float s = sigmoidf_(z);
    const float d = dH[i] * s * (1.f + z * (1.f - s));
    dZ[i] = d;
    dZb[i] = f2bf(d);
}

// out[c] = sum_r X[r][c]
__global__ void k_colsum(const float *__restrict__ X, float *__restrict__ out, int rows, int cols) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; ++r) s += X[(size_t)r * cols + c];
    out[c] = s;
}

// Gather and normalise rows of the training set into a BF16 batch.
__global__ void k_gather(const float *__restrict__ X, const float *__restrict__ Y, const int *__restrict__ idx,
                         int rows, int F, const float *__restrict__ mean, const float *__restrict__ istd,
                         float ym, float yis, uint16_t *__restrict__ Xb, float *__restrict__ Yb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * F) return;
    const int r = (int)(i / F), f = (int)(i % F);
    const int src = idx ? idx[r] : r;
    Xb[i] = f2bf((X[(size_t)src * F + f] - mean[f]) * istd[f]);
    if (f == 0 && Yb) Yb[r] = (Y[src] - ym) * yis;
}

// Mean squared error of pred vs target; dpred = 2 (pred - t) / rows.
__global__ void k_mse(const float *__restrict__ pred, const float *__restrict__ tgt, float *__restrict__ dpred,
                      uint16_t *__restrict__ dpredb, float *__restrict__ sq, int rows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    const float e = pred[i] - tgt[i];
    const float d = 2.f * e / (float)rows;
    dpred[i] = d;
    dpredb[i] = f2bf(d);
    sq[i] = e * e;
}

// Adam on FP32 master weights; refresh the BF16 copy.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ pb, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float w = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = w;
    pb[i] = f2bf(w);
}

__global__ void k_to_bf16(const float *__restrict__ in, uint16_t *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

}  // namespace casim

struct CasimMLP {
    int dev = 0;
    hipStream_t stream = nullptr;
    int n_in = 0, width = 0, depth = 0, members = 0;
    std::vector<int> fin, fout;       // layer fan-in / fan-out
    std::vector<size_t> woff, boff;   // offsets of W_l and b_l in a member's parameters
    size_t P = 0;                     // parameters per member
    std::vector<float *> p, g, am, av;
    std::vector<uint16_t *> pb;
    std::vector<int> steps;
    float *mean = nullptr, *istd = nullptr;
    float ym = 0.f, ys = 1.f;
    float *X = nullptr, *Y = nullptr;
    int *perm_d = nullptr;  // [rows] shuffled row order of the current epoch
    int64_t rows = 0;
    // fused inference (k_mlp_infer): transposed weights, refreshed when the
    // parameters change (pver counts parameter updates, wt_ver the copy's)
    bool fused = true;
    uint16_t *wt = nullptr;  // [members][wt_per]
    size_t wt_per = 0;
    uint64_t pver = 1, wt_ver = 0;
    bool fused_ok() const;
    void refresh_wt();
    void infer(const float *x, int B, float *out, int ostride);
    // batch workspace
    int cap = 0;
    std::vector<uint16_t *> Xl;  // [depth + 1] BF16 activations
    std::vector<float *> Zl;     // [depth + 1] FP32 pre-activations (last = output)
    float *dH = nullptr, *dZ = nullptr, *tgt = nullptr, *sq = nullptr;
    uint16_t *dZb = nullptr;
    int *idx = nullptr;

    void use() { HIPCK(hipSetDevice(dev)); }
    void reserve(int b);
    void gemm(const uint16_t *A, int lda, int ta, const uint16_t *B, int ldb, int tb, float *C, int ldc, int M, int N,
              int K, float beta = 0.f);
    void forward(int e, int B);
    void backward(int e, int B);
    void free_batch();
};

void CasimMLP::gemm(const uint16_t *A, int lda, int ta, const uint16_t *B, int ldb, int tb, float *C, int ldc, int M,
                    int N, int K, float beta) {
    dim3 grid((N + GBN - 1) / GBN, (M + GBM - 1) / GBM);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, stream, A, lda, ta, B, ldb, tb, C, ldc, M, N, K, beta);
    HIPCK(hipGetLastError());
}

bool CasimMLP::fused_ok() const {
    return fused && members <= IMAXM && depth + 1 <= IMAXL && n_in <= IMAXW && width <= IMAXW && width % 64 == 0;
}

void CasimMLP::refresh_wt() {
    if (wt && wt_ver == pver) return;
    std::vector<size_t> off;
    size_t per = 0;
    for (int l = 0; l <= depth; ++l) {
        off.push_back(per);
        per += (size_t)((fout[l] + 15) & ~15) * ((fin[l] + 31) & ~31);
    }
    if (!wt) {
        HIPCK(hipMalloc(&wt, sizeof(uint16_t) * per * members));
        wt_per = per;
    }
    for (int e = 0; e < members; ++e)
        for (int l = 0; l <= depth; ++l) {
            const int kp = (fin[l] + 31) & ~31, np = (fout[l] + 15) & ~15;
            const size_t n = (size_t)np * kp;
            hipLaunchKernelGGL(k_wt, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, pb[e] + woff[l],
                               wt + (size_t)e * wt_per + off[l], fin[l], fout[l], kp, np);
            HIPCK(hipGetLastError());
        }
    wt_ver = pver;
}

// Every member's output (normalised units) for B raw rows x -> out[e * ostride + row].
void CasimMLP::infer(const float *x, int B, float *out, int ostride) {
    refresh_wt();
    InferArgs a{};
    size_t off = 0;
    for (int l = 0; l <= depth; ++l) {
        a.wtoff[l] = (int)off;
        a.kp[l] = (fin[l] + 31) & ~31;
        a.fout[l] = fout[l];
        a.boff[l] = (int)boff[l];
        off += (size_t)((fout[l] + 15) & ~15) * a.kp[l];
    }
    for (int e = 0; e < members; ++e) {
        a.wt[e] = wt + (size_t)e * wt_per;
        a.p[e] = p[e];
    }
    a.nl = depth + 1;
    a.n_in = n_in;
    a.mean = mean;
    a.istd = istd;
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((B + IMR - 1) / IMR), (unsigned)members), dim3(64 * IWAVES), 0, stream, x,
                       B, out, ostride, a);
    HIPCK(hipGetLastError());
}

void CasimMLP::free_batch() {
    for (auto q : Xl) (void)hipFree(q);
    for (auto q : Zl) (void)hipFree(q);
    Xl.clear();
    Zl.clear();
    for (void *q : {(void *)dH, (void *)dZ, (void *)tgt, (void *)sq, (void *)dZb, (void *)idx})
        if (q) (void)hipFree(q);
    dH = dZ = tgt = sq = nullptr;
    dZb = nullptr;
    idx = nullptr;
    cap = 0;
}

void CasimMLP::reserve(int b) {
    if (b <= cap) return;
    free_batch();
    cap = b;
    Xl.resize(depth + 1);
    Zl.resize(depth + 1);
    for (int l = 0; l <= depth; ++l) {
        HIPCK(hipMalloc(&Xl[l], sizeof(uint16_t) * (size_t)b * std::max(fin[l], 1)));
        HIPCK(hipMalloc(&Zl[l], sizeof(float) * (size_t)b * fout[l]));
    }
    const int wmax = std::max(width, n_in);
    HIPCK(hipMalloc(&dH, sizeof(float) * (size_t)b * wmax));
    HIPCK(hipMalloc(&dZ, sizeof(float) * (size_t)b * wmax));
    HIPCK(hipMalloc(&dZb, sizeof(uint16_t) * (size_t)b * wmax));
    HIPCK(hipMalloc(&tgt, sizeof(float) * b));
    HIPCK(hipMalloc(&sq, sizeof(float) * b));
    HIPCK(hipMalloc(&idx, sizeof(int) * b));
}

// Activations of member e for the B rows already in Xl[0].
void CasimMLP::forward(int e, int B) {
    for (int l = 0; l <= depth; ++l) {
        const uint16_t *W = pb[e] + woff[l];
        gemm(Xl[l], fin[l], 0, W, fout[l], 0, Zl[l], fout[l], B, fout[l], fin[l]);
        const size_t n = (size_t)B * fout[l];
        const bool last = l == depth;
        hipLaunchKernelGGL(k_bias_act, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, Zl[l], p[e] + boff[l],
                           last ? (uint16_t *)nullptr : Xl[l + 1], B, fout[l], last ? 0 : 1);
        HIPCK(hipGetLastError());
    }
}

// Gradients of member e into g[e] given dZb/dZ of the output layer.
void CasimMLP::backward(int e, int B) {
    for (int l = depth; l >= 0; --l) {
        // dW_l = X_l^T dZ_l, db_l = column sums of dZ_l
        gemm(Xl[l], fin[l], 1, dZb, fout[l], 0, g[e] + woff[l], fout[l], fin[l], fout[l], B);
        hipLaunchKernelGGL(k_colsum, dim3((fout[l] + 255) / 256), dim3(256), 0, stream, dZ, g[e] + boff[l], B,
                           fout[l]);
        HIPCK(hipGetLastError());
        if (l == 0) break;
        // dH_{l} = dZ_l W_l^T  (the activation feeding layer l)
        gemm(dZb, fout[l], 0, pb[e] + woff[l], fout[l], 1, dH, fin[l], B, fin[l], fout[l]);
        const size_t n = (size_t)B * fin[l];
        hipLaunchKernelGGL(k_silu_back, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, dH, Zl[l - 1], dZ,
                           dZb, n);
        HIPCK(hipGetLastError());
    }
}

namespace casim {

// Every member's output (normalised units) for `rows` raw feature rows on the
// device: x [rows x n_in] -> out [members x rows].  Runs on `stream` after
// the MLP's own stream has drained.
void mlp_forward_device(CasimMLP *m, const float *x, int64_t rows, float *out, hipStream_t stream) {
    m->use();
    HIPCK(hipStreamSynchronize(m->stream));
    hipStream_t keep = m->stream;
    m->stream = stream;
    try {
        const int chunk = (int)std::min<int64_t>(rows, 65536);
        if (!m->fused_ok()) m->reserve(chunk);
        for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
            const int B = (int)std::min<int64_t>(chunk, rows - r0);
            if (m->fused_ok()) {
                m->infer(x + r0 * m->n_in, B, out + r0, (int)rows);
                continue;
            }
            const size_t n = (size_t)B * m->n_in;
            hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                               x + r0 * m->n_in, (const float *)nullptr, (const int *)nullptr, B, m->n_in, m->mean,
                               m->istd, 0.f, 1.f, m->Xl[0], (float *)nullptr);
            HIPCK(hipGetLastError());
            for (int e = 0; e < m->members; ++e) {
                m->forward(e, B);
                HIPCK(hipMemcpyAsync(out + (size_t)e * rows + r0, m->Zl[m->depth], sizeof(float) * B,
                                     hipMemcpyDeviceToDevice, stream));
            }
        }
    } catch (...) {
        m->stream = keep;
        throw;
    }
    m->stream = keep;
}

int mlp_members(const CasimMLP *m) { return m->members; }
int mlp_inputs(const CasimMLP *m) { return m->n_in; }
float mlp_out_mean(const CasimMLP *m) { return m->ym; }
float mlp_out_std(const CasimMLP *m) { return m->ys; }

}  // namespace casim

// ---------------------------------------------------------------------------
// Ensemble prediction that allocates nothing: the activations live in a fixed
// workspace, one per GPU, held from the library's first use of that GPU until
// the process ends.  Layout for up to CASIM_MLP_NOALLOC_MAX_ROWS rows: two BF16
// activation buffers [rows x CASIM_MLP_NOALLOC_MAX_WIDTH] (ping-pong), one FP32
// pre-activation buffer [rows x CASIM_MLP_NOALLOC_MAX_WIDTH], and the members'
// outputs [CASIM_MLP_NOALLOC_MAX_MEMBERS x rows].

namespace casim {
namespace {
constexpr size_t NA_ACT = (size_t)CASIM_MLP_NOALLOC_MAX_ROWS * CASIM_MLP_NOALLOC_MAX_WIDTH;  // elements per activation buffer
constexpr size_t NA_OFF_XB = NA_ACT * sizeof(uint16_t);
constexpr size_t NA_OFF_Z = 2 * NA_ACT * sizeof(uint16_t);
constexpr size_t NA_OFF_OUT = NA_OFF_Z + NA_ACT * sizeof(float);
constexpr size_t NA_BYTES = NA_OFF_OUT + (size_t)CASIM_MLP_NOALLOC_MAX_MEMBERS * CASIM_MLP_NOALLOC_MAX_ROWS * sizeof(float);
static_assert(NA_OFF_XB % 256 == 0 && NA_OFF_Z % 256 == 0 && NA_OFF_OUT % 256 == 0, "workspace alignment");
static_assert(NA_BYTES == CASIM_MLP_NOALLOC_BYTES, "workspace size as declared in casim_surrogate.h");

// The workspace.  In host code this name stands for the symbol only; its
// device address comes from hipGetSymbolAddress (noalloc_workspace below).
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
void *na_ptr[NA_MAX_DEVICES] = {};

// Device address of this GPU's workspace, looked up once per device with that
// device current.
void *noalloc_workspace(int dev) {
    REQUIRE(dev >= 0 && dev < NA_MAX_DEVICES, CASIM_E_DEVICE, "device index out of range for the MLP workspace");
    std::lock_guard<std::mutex> lock(na_mutex);
    if (!na_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw Err(CASIM_E_DEVICE, std::string("hipGetSymbolAddress of the MLP workspace: ") + hipGetErrorString(e));
        na_ptr[dev] = ptr;
    }
    return na_ptr[dev];
}

// out[e * cap + r] = z[r] (the output layer's column) for member e.
__global__ void k_na_collect(const float *__restrict__ z, float *__restrict__ out, int rows) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) out[r] = z[r];
}

// Mean and relative spread of exp(y * ys + ym) over the members, in the
// arithmetic of casim_mlp_predict's host reduction.
__global__ void k_na_reduce(const float *__restrict__ out, int members, int cap, int rows, float ys, float ym,
                            float *__restrict__ mean, float *__restrict__ spread) {
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double s = 0, s2 = 0;
    for (int e = 0; e < members; ++e) {
        const float o = out[(size_t)e * cap + r] * ys + ym;
        const double v = exp((double)o);
        s += v;
        s2 += v * v;
    }
    const double mu = s / members;
    const double var = fmax(0.0, s2 / members - mu * mu);
    if (mean) mean[r] = (float)mu;
    if (spread) spread[r] = (float)(mu > 0 ? sqrt(var) / mu : 0.0);
}
}  // namespace
}  // namespace casim

extern "C" {

int casim_mlp_create(int32_t n_in, int32_t width, int32_t depth, int32_t members, uint64_t seed, CasimMLP **out) {
    CasimMLP *m = nullptr;
    int rc = guard([&] {
        REQUIRE(out && n_in >= 1 && width >= 1 && depth >= 1 && members >= 1, CASIM_E_INVALID, "bad MLP shape");
        *out = nullptr;
        m = new CasimMLP();
        HIPCK(hipGetDevice(&m->dev));
        HIPCK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        (void)casim::noalloc_workspace(m->dev);  // this GPU's workspace of casim_mlp_predict_noalloc
        m->n_in = n_in;
        m->width = width;
        m->depth = depth;
        m->members = members;
        for (int l = 0; l <= depth; ++l) {
            m->fin.push_back(l == 0 ? n_in : width);
            m->fout.push_back(l == depth ? 1 : width);
        }
        size_t off = 0;
        for (int l = 0; l <= depth; ++l) {
            m->woff.push_back(off);
            off += (size_t)m->fin[l] * m->fout[l];
            m->boff.push_back(off);
            off += m->fout[l];
        }
        m->P = off;
        std::mt19937_64 gen(seed);
        std::vector<float> h(m->P);
        for (int e = 0; e < members; ++e) {
            for (int l = 0; l <= depth; ++l) {
                const float lim = std::sqrt(6.f / (float)m->fin[l]);
                std::uniform_real_distribution<float> U(-lim, lim);
                for (size_t i = 0; i < (size_t)m->fin[l] * m->fout[l]; ++i) h[m->woff[l] + i] = U(gen);
                for (int i = 0; i < m->fout[l]; ++i) h[m->boff[l] + i] = 0.f;
            }
            float *pp, *gg, *mm, *vv;
            uint16_t *bb;
            HIPCK(hipMalloc(&pp, sizeof(float) * m->P));
            HIPCK(hipMalloc(&gg, sizeof(float) * m->P));
            HIPCK(hipMalloc(&mm, sizeof(float) * m->P));
            HIPCK(hipMalloc(&vv, sizeof(float) * m->P));
            HIPCK(hipMalloc(&bb, sizeof(uint16_t) * m->P));
            m->p.push_back(pp);
            m->g.push_back(gg);
            m->am.push_back(mm);
            m->av.push_back(vv);
            m->pb.push_back(bb);
            m->steps.push_back(0);
            HIPCK(hipMemcpyAsync(pp, h.data(), sizeof(float) * m->P, hipMemcpyHostToDevice, m->stream));
            HIPCK(hipMemsetAsync(mm, 0, sizeof(float) * m->P, m->stream));
            HIPCK(hipMemsetAsync(vv, 0, sizeof(float) * m->P, m->stream));
            hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m->P + 255) / 256)), dim3(256), 0, m->stream, pp, bb,
                               m->P);
            HIPCK(hipGetLastError());
        }
        std::vector<float> zero(n_in, 0.f), one(n_in, 1.f);
        HIPCK(hipMalloc(&m->mean, sizeof(float) * n_in));
        HIPCK(hipMalloc(&m->istd, sizeof(float) * n_in));
        HIPCK(hipMemcpyAsync(m->mean, zero.data(), sizeof(float) * n_in, hipMemcpyHostToDevice, m->stream));
        HIPCK(hipMemcpyAsync(m->istd, one.data(), sizeof(float) * n_in, hipMemcpyHostToDevice, m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
        *out = m;
    });
    if (rc != CASIM_OK && m) {
        std::string keep = last_error();
        casim_mlp_destroy(m);
        last_error() = keep;
    }
    return rc;
}

int casim_mlp_destroy(CasimMLP *m) {
    if (!m) return CASIM_OK;
    (void)hipSetDevice(m->dev);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    m->free_batch();
    for (auto v : {&m->p, &m->g, &m->am, &m->av})
        for (float *q : *v) (void)hipFree(q);
    for (uint16_t *q : m->pb) (void)hipFree(q);
    if (m->wt) (void)hipFree(m->wt);
    for (void *q : {(void *)m->mean, (void *)m->istd, (void *)m->X, (void *)m->Y, (void *)m->perm_d})
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return CASIM_OK;
}

int64_t casim_mlp_param_count(const CasimMLP *m) { return m ? (int64_t)m->P : 0; }

int casim_mlp_get_params(CasimMLP *m, int32_t member, float *out) {
    return guard([&] {
        REQUIRE(m && out, CASIM_E_INVALID, "null argument");
        REQUIRE(member >= 0 && member < m->members, CASIM_E_RANGE, "member out of range");
        m->use();
        HIPCK(hipMemcpyAsync(out, m->p[member], sizeof(float) * m->P, hipMemcpyDeviceToHost, m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
    });
}

int casim_mlp_set_params(CasimMLP *m, int32_t member, const float *in) {
    return guard([&] {
        REQUIRE(m && in, CASIM_E_INVALID, "null argument");
        REQUIRE(member >= 0 && member < m->members, CASIM_E_RANGE, "member out of range");
        m->use();
        HIPCK(hipMemcpyAsync(m->p[member], in, sizeof(float) * m->P, hipMemcpyHostToDevice, m->stream));
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m->P + 255) / 256)), dim3(256), 0, m->stream, m->p[member],
                           m->pb[member], m->P);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(m->stream));
        ++m->pver;
    });
}

int casim_mlp_set_normalization(CasimMLP *m, const float *in_mean, const float *in_std, float out_mean,
                                float out_std) {
    return guard([&] {
        REQUIRE(m && in_mean && in_std && out_std > 0.f, CASIM_E_INVALID, "bad normalisation");
        m->use();
        std::vector<float> is(m->n_in);
        for (int i = 0; i < m->n_in; ++i) is[i] = in_std[i] > 0.f ? 1.f / in_std[i] : 1.f;
        HIPCK(hipMemcpyAsync(m->mean, in_mean, sizeof(float) * m->n_in, hipMemcpyHostToDevice, m->stream));
        HIPCK(hipMemcpyAsync(m->istd, is.data(), sizeof(float) * m->n_in, hipMemcpyHostToDevice, m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
        m->ym = out_mean;
        m->ys = out_std;
    });
}

int casim_mlp_set_data(CasimMLP *m, const float *x, const float *y, int64_t rows) {
    return guard([&] {
        REQUIRE(m && x && y && rows >= 1, CASIM_E_INVALID, "bad training data");
        m->use();
        if (m->X) (void)hipFree(m->X);
        if (m->Y) (void)hipFree(m->Y);
        if (m->perm_d) (void)hipFree(m->perm_d);
        m->X = m->Y = nullptr;
        m->perm_d = nullptr;
        HIPCK(hipMalloc(&m->X, sizeof(float) * rows * m->n_in));
        HIPCK(hipMalloc(&m->Y, sizeof(float) * rows));
        HIPCK(hipMalloc(&m->perm_d, sizeof(int) * rows));
        HIPCK(hipMemcpyAsync(m->X, x, sizeof(float) * rows * m->n_in, hipMemcpyHostToDevice, m->stream));
        HIPCK(hipMemcpyAsync(m->Y, y, sizeof(float) * rows, hipMemcpyHostToDevice, m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
        m->rows = rows;
    });
}

int casim_mlp_train(CasimMLP *m, int32_t epochs, int32_t batch, float lr, uint64_t seed, float *loss_out) {
    return guard([&] {
        REQUIRE(m && m->X, CASIM_E_STATE, "no training data");
        REQUIRE(epochs >= 1 && batch >= 1, CASIM_E_INVALID, "bad epochs/batch");
        m->use();
        const int B = (int)std::min<int64_t>(batch, m->rows);
        m->reserve(B);
        std::vector<int> perm(m->rows);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float yis = 1.f / m->ys;
        float *sq_h = nullptr;
        std::vector<float> sq(B);
        (void)sq_h;
        ++m->pver;  // the transposed inference weights follow the parameters
        for (int e = 0; e < m->members; ++e) {
            std::mt19937_64 gen(seed * 1315423911ull + (uint64_t)e * 2654435761ull + 17);
            double last = 0.0;
            for (int ep = 0; ep < epochs; ++ep) {
                std::iota(perm.begin(), perm.end(), 0);
                std::shuffle(perm.begin(), perm.end(), gen);
                HIPCK(hipMemcpyAsync(m->perm_d, perm.data(), sizeof(int) * m->rows, hipMemcpyHostToDevice,
                                     m->stream));
                double sum = 0.0;
                int64_t cnt = 0;
                const bool keep = ep == epochs - 1;
                for (int64_t r0 = 0; r0 + B <= m->rows; r0 += B) {
                    const size_t n = (size_t)B * m->n_in;
                    hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, m->X,
                                       m->Y, (const int *)(m->perm_d + r0), B, m->n_in, m->mean, m->istd, m->ym, yis,
                                       m->Xl[0], m->tgt);
                    HIPCK(hipGetLastError());
                    m->forward(e, B);
                    hipLaunchKernelGGL(k_mse, dim3((B + 255) / 256), dim3(256), 0, m->stream, m->Zl[m->depth], m->tgt,
                                       m->dZ, m->dZb, m->sq, B);
                    HIPCK(hipGetLastError());
                    m->backward(e, B);
                    const int t = ++m->steps[e];
                    const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
                    hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->P + 255) / 256)), dim3(256), 0, m->stream, m->p[e],
                                       m->g[e], m->am[e], m->av[e], m->pb[e], m->P, lr, b1, b2, eps, c1, c2);
                    HIPCK(hipGetLastError());
                    if (keep) {
                        HIPCK(hipMemcpyAsync(sq.data(), m->sq, sizeof(float) * B, hipMemcpyDeviceToHost, m->stream));
                        HIPCK(hipStreamSynchronize(m->stream));
                        for (int i = 0; i < B; ++i) sum += sq[i];
                        cnt += B;
                    }
                }
                if (keep) last = cnt ? sum / (double)cnt : 0.0;
            }
            HIPCK(hipStreamSynchronize(m->stream));
            if (loss_out) loss_out[e] = (float)last;
        }
    });
}

int casim_mlp_gradient(CasimMLP *m, int32_t member, const float *x, const float *y, int64_t rows, float *grad,
                       float *loss) {
    return guard([&] {
        REQUIRE(m && x && y && grad && rows >= 1, CASIM_E_INVALID, "bad arguments");
        REQUIRE(member >= 0 && member < m->members, CASIM_E_RANGE, "member out of range");
        m->use();
        const int B = (int)rows;
        m->reserve(B);
        float *dx = nullptr, *dy = nullptr;
        HIPCK(hipMalloc(&dx, sizeof(float) * rows * m->n_in));
        HIPCK(hipMalloc(&dy, sizeof(float) * rows));
        HIPCK(hipMemcpyAsync(dx, x, sizeof(float) * rows * m->n_in, hipMemcpyHostToDevice, m->stream));
        HIPCK(hipMemcpyAsync(dy, y, sizeof(float) * rows, hipMemcpyHostToDevice, m->stream));
        const size_t n = (size_t)B * m->n_in;
        hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, dx, dy,
                           (const int *)nullptr, B, m->n_in, m->mean, m->istd, m->ym, 1.f / m->ys, m->Xl[0], m->tgt);
        m->forward(member, B);
        hipLaunchKernelGGL(k_mse, dim3((B + 255) / 256), dim3(256), 0, m->stream, m->Zl[m->depth], m->tgt, m->dZ,
                           m->dZb, m->sq, B);
        m->backward(member, B);
        std::vector<float> sq(B);
        HIPCK(hipMemcpyAsync(grad, m->g[member], sizeof(float) * m->P, hipMemcpyDeviceToHost, m->stream));
        HIPCK(hipMemcpyAsync(sq.data(), m->sq, sizeof(float) * B, hipMemcpyDeviceToHost, m->stream));
        HIPCK(hipStreamSynchronize(m->stream));
        (void)hipFree(dx);
        (void)hipFree(dy);
        if (loss) {
            double s = 0;
            for (float v : sq) s += v;
            *loss = (float)(s / B);
        }
    });
}

int casim_mlp_predict(CasimMLP *m, const float *x, int64_t rows, float *mean, float *spread, float *member_out) {
    return guard([&] {
        REQUIRE(m && x && rows >= 1, CASIM_E_INVALID, "bad arguments");
        m->use();
        const int chunk = (int)std::min<int64_t>(rows, 65536);
        const bool fused = m->fused_ok();
        if (!fused) m->reserve(chunk);
        float *dx = nullptr, *dout = nullptr;
        HIPCK(hipMalloc(&dx, sizeof(float) * (size_t)chunk * m->n_in));
        if (fused) HIPCK(hipMalloc(&dout, sizeof(float) * (size_t)chunk * m->members));
        std::vector<float> out((size_t)m->members * rows);
        std::vector<float> tmp(chunk);
        for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
            const int B = (int)std::min<int64_t>(chunk, rows - r0);
            HIPCK(hipMemcpyAsync(dx, x + r0 * m->n_in, sizeof(float) * (size_t)B * m->n_in, hipMemcpyHostToDevice,
                                 m->stream));
            if (fused) {
                m->infer(dx, B, dout, B);
                for (int e = 0; e < m->members; ++e) {
                    HIPCK(hipMemcpyAsync(tmp.data(), dout + (size_t)e * B, sizeof(float) * B, hipMemcpyDeviceToHost,
                                         m->stream));
                    HIPCK(hipStreamSynchronize(m->stream));
                    for (int i = 0; i < B; ++i) out[(size_t)e * rows + r0 + i] = tmp[i] * m->ys + m->ym;
                }
                continue;
            }
            const size_t n = (size_t)B * m->n_in;
            hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, dx,
                               (const float *)nullptr, (const int *)nullptr, B, m->n_in, m->mean, m->istd, 0.f, 1.f,
                               m->Xl[0], (float *)nullptr);
            HIPCK(hipGetLastError());
            for (int e = 0; e < m->members; ++e) {
                m->forward(e, B);
                HIPCK(hipMemcpyAsync(tmp.data(), m->Zl[m->depth], sizeof(float) * B, hipMemcpyDeviceToHost,
                                     m->stream));
                HIPCK(hipStreamSynchronize(m->stream));
                for (int i = 0; i < B; ++i) out[(size_t)e * rows + r0 + i] = tmp[i] * m->ys + m->ym;
            }
        }
        (void)hipFree(dx);
        if (dout) (void)hipFree(dout);
        for (int64_t i = 0; i < rows; ++i) {
            double s = 0, s2 = 0;
            for (int e = 0; e < m->members; ++e) {
                const double v = std::exp((double)out[(size_t)e * rows + i]);
                s += v;
                s2 += v * v;
            }
            const double mu = s / m->members;
            const double var = std::max(0.0, s2 / m->members - mu * mu);
            if (mean) mean[i] = (float)mu;
            if (spread) spread[i] = (float)(mu > 0 ? std::sqrt(var) / mu : 0.0);
        }
        if (member_out) std::memcpy(member_out, out.data(), sizeof(float) * out.size());
    });
}

int casim_mlp_set_fused_inference(CasimMLP *m, int32_t on) {
    return guard([&] {
        REQUIRE(m, CASIM_E_INVALID, "null handle");
        REQUIRE(on == 0 || on == 1, CASIM_E_INVALID, "on must be 0 or 1");
        m->fused = on != 0;
    });
}

int casim_mlp_predict_noalloc(CasimMLP *m, const float *x, int64_t rows, float *mean, float *spread, void *stream) {
    return guard([&] {
        REQUIRE(m && x && rows >= 1, CASIM_E_INVALID, "bad arguments");
        REQUIRE(rows <= CASIM_MLP_NOALLOC_MAX_ROWS, CASIM_E_RANGE,
                "rows exceed the workspace of casim_mlp_predict_noalloc (" + std::to_string(CASIM_MLP_NOALLOC_MAX_ROWS) +
                    "); split the call");
        REQUIRE(m->width <= CASIM_MLP_NOALLOC_MAX_WIDTH && m->n_in <= CASIM_MLP_NOALLOC_MAX_WIDTH, CASIM_E_RANGE,
                "layer wider than the workspace of casim_mlp_predict_noalloc (" +
                    std::to_string(CASIM_MLP_NOALLOC_MAX_WIDTH) + ")");
        REQUIRE(m->members <= CASIM_MLP_NOALLOC_MAX_MEMBERS, CASIM_E_RANGE,
                "more members than the workspace of casim_mlp_predict_noalloc holds (" +
                    std::to_string(CASIM_MLP_NOALLOC_MAX_MEMBERS) + ")");
        REQUIRE(m->mean && m->istd, CASIM_E_STATE, "normalisation not set");
        m->use();
        hipStream_t st = stream ? (hipStream_t)stream : m->stream;
        unsigned char *ws = (unsigned char *)casim::noalloc_workspace(m->dev);
        uint16_t *xa = (uint16_t *)ws;
        uint16_t *xb = (uint16_t *)(ws + casim::NA_OFF_XB);
        float *z = (float *)(ws + casim::NA_OFF_Z);
        float *outm = (float *)(ws + casim::NA_OFF_OUT);
        const int B = (int)rows;
        const int cap = CASIM_MLP_NOALLOC_MAX_ROWS;
        const size_t n = (size_t)B * m->n_in;
        for (int e = 0; e < m->members; ++e) {
            hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, (const float *)nullptr,
                               (const int *)nullptr, B, m->n_in, m->mean, m->istd, 0.f, 1.f, xa, (float *)nullptr);
            HIPCK(hipGetLastError());
            uint16_t *in = xa, *nxt = xb;
            for (int l = 0; l <= m->depth; ++l) {
                const int fi = m->fin[l], fo = m->fout[l];
                dim3 grid((fo + GBN - 1) / GBN, (B + GBM - 1) / GBM);
                hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, st, in, fi, 0, m->pb[e] + m->woff[l], fo, 0, z, fo, B,
                                   fo, fi, 0.f);
                HIPCK(hipGetLastError());
                const bool last = l == m->depth;
                const size_t nz = (size_t)B * fo;
                hipLaunchKernelGGL(k_bias_act, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, z,
                                   m->p[e] + m->boff[l], last ? (uint16_t *)nullptr : nxt, B, fo, last ? 0 : 1);
                HIPCK(hipGetLastError());
                std::swap(in, nxt);
            }
            hipLaunchKernelGGL(casim::k_na_collect, dim3((B + 255) / 256), dim3(256), 0, st, (const float *)z,
                               outm + (size_t)e * cap, B);
            HIPCK(hipGetLastError());
        }
        hipLaunchKernelGGL(casim::k_na_reduce, dim3((B + 255) / 256), dim3(256), 0, st, (const float *)outm, m->members,
                           cap, B, m->ys, m->ym, mean, spread);
        HIPCK(hipGetLastError());
    });
}

int casim_gemm_bf16(const void *a, int32_t lda, int32_t trans_a, const void *b, int32_t ldb, int32_t trans_b,
                    float *c, int32_t ldc, int32_t M, int32_t N, int32_t K, float beta, void *stream) {
    return guard([&] {
        REQUIRE(a && b && c && M >= 0 && N >= 0 && K >= 0, CASIM_E_INVALID, "bad GEMM arguments");
        dim3 grid((N + GBN - 1) / GBN, (M + GBM - 1) / GBM);
        if (M == 0 || N == 0) return;
        hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, (hipStream_t)stream, (const uint16_t *)a, lda, trans_a,
                           (const uint16_t *)b, ldb, trans_b, c, ldc, M, N, K, beta);
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize((hipStream_t)stream));
    });
}

}  // extern "C"
