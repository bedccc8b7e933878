// pt_mlp.hip — MLP ensemble surrogate on gfx950 MFMA (include/pandemic/pt_surrogate.h).
//
// All matrix products go through one tiled GEMM kernel on
// v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation.  A 256-thread
// block computes a 64 x 64 tile of C = A * B (four waves, each 32 x 32 = 2 x 2
// MFMA tiles), staging 64 x 64 slices of A and B in LDS per K step (16-byte
// vector reads from global memory); operands may be read transposed, and the
// epilogue fuses bias + GELU (forward),
// bias (output layer), GELU backward (input gradients) or a plain FP32
// store (weight gradients).  Every launch covers all ensemble members
// (blockIdx.z), which train on the same batch.
//
// MFMA lane layout (checked on the device): lane l supplies row l % 16 of A
// and column l % 16 of B with eight K values each (K group l / 16), and
// receives C[4 (l / 16) + r][l % 16], r = 0..3.
//
// GELU is the exact form x * Phi(x), Phi(x) = (1 + erf(x / sqrt 2)) / 2
// (Hendrycks and Gimpel, 2016), through the HIP device erff / expf.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pandemic/pt_surrogate.h"
#include "pt_day.h"

using namespace pt;

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kDims[PT_MLP_LAYERS + 1] = {PT_MLP_IN, PT_MLP_HIDDEN, PT_MLP_HIDDEN, PT_MLP_HIDDEN, PT_MLP_HIDDEN,
                                           PT_MLP_OUT};
constexpr int kTM = 64, kTN = 64, kTK = 64;
constexpr int kGemmThreads = 256;
constexpr int kMaxWidth = PT_MLP_OUT;  // widest layer output

// Offsets of W_l and b_l in a member's parameter block.
struct Layout {
    int64_t w[PT_MLP_LAYERS], b[PT_MLP_LAYERS], count;
    Layout() {
        int64_t o = 0;
        for (int l = 0; l < PT_MLP_LAYERS; ++l) {
            w[l] = o;
            o += (int64_t)kDims[l] * kDims[l + 1];
            b[l] = o;
            o += kDims[l + 1];
        }
        count = o;
    }
};

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

enum Epilogue : int { EPI_BIAS_GELU = 0, EPI_BIAS, EPI_GELU_GRAD, EPI_F32 };

// Operands and outputs of one batched GEMM; *_z are per-member strides in
// elements (0: shared by all members).
struct GemmArgs {
    const bf16* A;
    int64_t lda, a_z;
    const bf16* B;
    int64_t ldb, b_z;
    int M, N, K;
    const float* bias;  // EPI_BIAS_GELU, EPI_BIAS
    int64_t bias_z;
    const float* pre;   // EPI_GELU_GRAD: pre-activations [M][N]
    int64_t pre_z;
    float* out_f32;     // pre-activations / outputs / input gradients / weight gradients
    int64_t f32_z;
    bf16* out_bf16;     // activations / input gradients
    int64_t bf16_z;
    int64_t ldc;
    // EPI_GELU_GRAD (may be null): the FP32 column sums of each block's 64
    // output rows, colpart[z][blockIdx.y][col] (N columns per row tile).
    float* colpart;
};

// C[M][N] = A(m, k) * B(k, n): A(m, k) = TA ? A[k lda + m] : A[m lda + k],
// B(k, n) = TB ? B[n ldb + k] : B[k ldb + n].
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(kGemmThreads) void k_gemm(GemmArgs g) {
    const int z = blockIdx.z;
    const bf16* __restrict__ A = g.A + z * g.a_z;
    const bf16* __restrict__ B = g.B + z * g.b_z;
    const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
    // Tiles stored [row or column][k] with rows of kTK + 8 bf16, so each
    // lane's eight K values are one aligned 16-byte LDS read.
    __shared__ __attribute__((aligned(16))) bf16 sA[kTM][kTK + 8];
    __shared__ __attribute__((aligned(16))) bf16 sB[kTN][kTK + 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = (bf16x8)(bf16)0.0f;
    // Global reads are 16-byte vectors of 8 consecutive elements of the
    // contiguous dimension (every stride, offset and extent here is a
    // multiple of 8, so a vector is wholly inside or wholly outside).  An
    // operand stored K-major (A with TA, B without TB) is read by one wave
    // (wave 0 for A, wave 1 for B) as 8 x 8 blocks, 8 K rows of 8
    // consecutive M or N, transposed in registers and stored as 16-byte LDS
    // rows of 8 K values.  The next K step's vectors are loaded before the
    // current step's products, so their latency overlaps the MFMA work.
    constexpr int kVec = kTM * kTK / 8 / kGemmThreads;  // vectors per thread per operand
    static_assert(kTM == 64 && kTN == 64 && kTK == 64, "the 8 x 8 block loads assume 64 x 64 tiles");
    bf16x8 ra[TA ? 8 : kVec], rb[TB ? kVec : 8];
    const int bk = (lane >> 3) * 8, bmn = (lane & 7) * 8;  // this lane's 8 x 8 block: K rows, M / N columns
    auto load = [&](int k0) {
        if constexpr (TA) {
            if (wave == 0) {
                const int m = m0 + bmn;
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int k = k0 + bk + r;
                    ra[r] = zero8;
                    if (m < g.M && k < g.K) ra[r] = *reinterpret_cast<const bf16x8*>(A + (int64_t)k * g.lda + m);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < kVec; ++q) {
                const int v = tid + q * kGemmThreads;
                const int mm = v / (kTK / 8), kk = (v % (kTK / 8)) * 8;
                const int m = m0 + mm, k = k0 + kk;
                ra[q] = zero8;
                if (m < g.M && k < g.K) ra[q] = *reinterpret_cast<const bf16x8*>(A + (int64_t)m * g.lda + k);
            }
        }
        if constexpr (!TB) {
            if (wave == 1) {
                const int n = n0 + bmn;
#pragma unroll
                for (int r = 0; r < 8; ++r) {
                    const int k = k0 + bk + r;
                    rb[r] = zero8;
                    if (n < g.N && k < g.K) rb[r] = *reinterpret_cast<const bf16x8*>(B + (int64_t)k * g.ldb + n);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < kVec; ++q) {
                const int v = tid + q * kGemmThreads;
                const int nn = v / (kTK / 8), kk = (v % (kTK / 8)) * 8;
                const int n = n0 + nn, k = k0 + kk;
                rb[q] = zero8;
                if (n < g.N && k < g.K) rb[q] = *reinterpret_cast<const bf16x8*>(B + (int64_t)n * g.ldb + k);
            }
        }
    };
    // Row c of the transposed block: element r is x[r][c].
    auto column = [](const bf16x8 (&x)[8], int c) {
        bf16x8 o;
#pragma unroll
        for (int r = 0; r < 8; ++r) o[r] = x[r][c];
        return o;
    };
    auto stage = [&]() {
        if constexpr (TA) {
            if (wave == 0) {
#pragma unroll
                for (int c = 0; c < 8; ++c) *reinterpret_cast<bf16x8*>(&sA[bmn + c][bk]) = column(ra, c);
            }
        } else {
#pragma unroll
            for (int q = 0; q < kVec; ++q) {
                const int v = tid + q * kGemmThreads;
                const int mm = v / (kTK / 8), kk = (v % (kTK / 8)) * 8;
                *reinterpret_cast<bf16x8*>(&sA[mm][kk]) = ra[q];
            }
        }
        if constexpr (!TB) {
            if (wave == 1) {
#pragma unroll
                for (int c = 0; c < 8; ++c) *reinterpret_cast<bf16x8*>(&sB[bmn + c][bk]) = column(rb, c);
            }
        } else {
#pragma unroll
            for (int q = 0; q < kVec; ++q) {
                const int v = tid + q * kGemmThreads;
                const int nn = v / (kTK / 8), kk = (v % (kTK / 8)) * 8;
                *reinterpret_cast<bf16x8*>(&sB[nn][kk]) = rb[q];
            }
        }
    };
    load(0);
    for (int k0 = 0; k0 < g.K; k0 += kTK) {
        stage();
        __syncthreads();
        if (k0 + kTK < g.K) load(k0 + kTK);
#pragma unroll
        for (int kc = 0; kc < kTK; kc += 32) {
            const int kq = kc + 8 * (lane >> 4);
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[wm + ti * 16 + (lane & 15)][kq]);
#pragma unroll
                for (int tj = 0; tj < 2; ++tj) {
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[wn + tj * 16 + (lane & 15)][kq]);
                    acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[ti][tj], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float cs[2] = {0.f, 0.f};  // EPI_GELU_GRAD: this lane's column sums
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm + ti * 16 + 4 * (lane >> 4) + r;
                const int col = n0 + wn + tj * 16 + (lane & 15);
                if (row >= g.M || col >= g.N) continue;
                const float v = acc[ti][tj][r];
                const int64_t o = (int64_t)row * g.ldc + col;
                if (EPI == EPI_BIAS_GELU) {
                    const float pre = v + g.bias[z * g.bias_z + col];
                    g.out_f32[z * g.f32_z + o] = pre;
                    g.out_bf16[z * g.bf16_z + o] = (bf16)gelu(pre);
                } else if (EPI == EPI_BIAS) {
                    g.out_f32[z * g.f32_z + o] = v + g.bias[z * g.bias_z + col];
                } else if (EPI == EPI_GELU_GRAD) {
                    const float d = v * gelu_grad(g.pre[z * g.pre_z + o]);
                    g.out_f32[z * g.f32_z + o] = d;
                    g.out_bf16[z * g.bf16_z + o] = (bf16)d;
                    cs[tj] += d;
                } else {
                    g.out_f32[z * g.f32_z + o] = v;
                }
            }
    if constexpr (EPI == EPI_GELU_GRAD) {
        // Column sums of the tile: this lane's 8 rows, then the 4 lane
        // groups of a 16-column slice, then the two row halves, in that
        // fixed order.
        if (g.colpart) {
            __shared__ float s_cs[2][kTN];
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
                cs[tj] += __shfl_xor(cs[tj], 16, 64);
                cs[tj] += __shfl_xor(cs[tj], 32, 64);
                if (lane < 16) s_cs[wm / 32][wn + tj * 16 + lane] = cs[tj];
            }
            __syncthreads();
            const int col = n0 + tid;
            if (tid < kTN && col < g.N)
                g.colpart[((int64_t)z * gridDim.y + blockIdx.y) * g.N + col] = s_cs[0][tid] + s_cs[1][tid];
        }
    }
}

// Batch gather: xb[b][i] = bf16(X[rows[b]][i]), yb[b][j] = Y[rows[b]][j].
// step (may be null): [0] the step index within a training call, which
// selects the batch's rows order + step[0] * B.
__global__ void k_gather(const float* __restrict__ X, const float* __restrict__ Y, const int32_t* __restrict__ order,
                         const int64_t* __restrict__ step, int B, bf16* __restrict__ xb, float* __restrict__ yb) {
    const int32_t* rows = order + (step ? step[0] * B : 0);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t nx = (int64_t)B * PT_MLP_IN, ny = (int64_t)B * PT_MLP_OUT;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nx + ny; i += stride) {
        if (i < nx) {
            const int64_t b = i / PT_MLP_IN, c = i % PT_MLP_IN;
            xb[i] = (bf16)X[(int64_t)rows[b] * PT_MLP_IN + c];
        } else {
            const int64_t j = i - nx, b = j / PT_MLP_OUT, c = j % PT_MLP_OUT;
            yb[j] = Y[(int64_t)rows[b] * PT_MLP_OUT + c];
        }
    }
}

__global__ void k_to_bf16(const float* __restrict__ x, int64_t n, bf16* __restrict__ y) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = (bf16)x[i];
}

constexpr int kLossThreads = 256;
constexpr int kLossBlocks = 64;

// Mean-squared error of member blockIdx.y: dz = 2 (pred - y) / (B * 720)
// (FP32 and BF16) and per-block partial sums of (pred - y)^2.
__global__ __launch_bounds__(kLossThreads) void k_loss(const float* __restrict__ pred, const float* __restrict__ yb,
                                                       int B, float* __restrict__ dz32, bf16* __restrict__ dz16,
                                                       int64_t dz_z, float* __restrict__ part) {
    const int z = blockIdx.y;
    const int64_t n = (int64_t)B * PT_MLP_OUT;
    const float* p = pred + z * n;
    const float scale = 2.0f / (float)n;
    float s = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * kLossThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLossThreads) {
        const float d = p[i] - yb[i];
        s += d * d;
        const float g = d * scale;
        dz32[z * dz_z + i] = g;
        dz16[z * dz_z + i] = (bf16)g;
    }
    __shared__ float s_w[kLossThreads / 64];
    s = s + __shfl_xor(s, 32, 64);
    s = s + __shfl_xor(s, 16, 64);
    s = s + __shfl_xor(s, 8, 64);
    s = s + __shfl_xor(s, 4, 64);
    s = s + __shfl_xor(s, 2, 64);
    s = s + __shfl_xor(s, 1, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < kLossThreads / 64; ++w) t += s_w[w];
        part[z * gridDim.x + blockIdx.x] = t;
    }
}

// Loss of member threadIdx.x from the block partials, in a fixed order.
// step (may be null): [0] the step index, the loss's column.
__global__ void k_loss_final(const float* __restrict__ part, int nblk, int members, int64_t n,
                             float* __restrict__ loss, int64_t loss_z, const int64_t* __restrict__ step) {
    const int z = threadIdx.x;
    if (z >= members) return;
    float t = 0.f;
    for (int b = 0; b < nblk; ++b) t += part[z * nblk + b];
    loss[z * loss_z + (step ? step[0] : 0)] = t / (float)n;
}

// Bias gradients: db[n] = sum over rows of dz[row][n], per member.  Partial
// sums per row chunk: for a hidden layer the input-gradient GEMM's epilogue
// leaves them (64-row tiles); for the output layer k_colsum_part sums a
// chunk of kColRows rows for 64 columns (4 row groups, added in a fixed
// order) into part[z][chunk][n].  k_colsum_final adds the chunks in order.
// Every order is fixed, so the result does not depend on scheduling.
constexpr int kColCols = 64, kColGroups = 4, kColRows = 128;
__global__ __launch_bounds__(kColCols * kColGroups) void k_colsum_part(const float* __restrict__ dz, int M, int N,
                                                                      int64_t dz_z, float* __restrict__ part) {
    const int z = blockIdx.z, chunk = blockIdx.y, R = gridDim.y;
    const int c = threadIdx.x % kColCols, grp = threadIdx.x / kColCols;
    const int n = blockIdx.x * kColCols + c;
    const float* d = dz + z * dz_z;
    const int m1 = min(M, (chunk + 1) * kColRows);
    float s = 0.f;
    if (n < N)
        for (int m = chunk * kColRows + grp; m < m1; m += kColGroups) s += d[(int64_t)m * N + n];
    __shared__ float s_p[kColGroups][kColCols];
    s_p[grp][c] = s;
    __syncthreads();
    if (grp == 0 && n < N) {
        float t = 0.f;
#pragma unroll
        for (int q = 0; q < kColGroups; ++q) t += s_p[q][c];
        part[((int64_t)z * R + chunk) * N + n] = t;
    }
}

__global__ void k_colsum_final(const float* __restrict__ part, int R, int N, int members, float* __restrict__ db,
                               int64_t db_z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= members * N) return;
    const int z = i / N, n = i % N;
    float t = 0.f;
    for (int r = 0; r < R; ++r) t += part[((int64_t)z * R + r) * N + n];
    db[z * db_z + n] = t;
}

// Adam on the FP32 master parameters, then their BF16 copies.
// Adam with bias corrections 1 - beta^t for t = step[1] + 1 (the Adam
// steps taken before this one, plus this one).
// Bias gradients (colpart set): summed here from layer l's partial sums,
// slice l of colpart ([member][R][width], R = r_out chunks for the output
// layer, r_hidden tiles for the others), in chunk order, the order
// k_colsum_final uses, instead of read from g.  Grid: x over one member's
// count parameters, y = member.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, bf16* __restrict__ pb, int64_t count, float lr,
                       const int64_t* __restrict__ step, const float* __restrict__ colpart, int64_t slice,
                       int r_hidden, int r_out) {
    const double t = (double)(step[1] + 1);
    const float bc1 = (float)(1.0 - pow(0.9, t)), bc2 = (float)(1.0 - pow(0.999, t));
    const int z = blockIdx.y;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < count; j += stride) {
        const int64_t i = (int64_t)z * count + j;
        int bl = -1;
        int64_t bj = 0, o = 0;
#pragma unroll
        for (int l = 0; l < PT_MLP_LAYERS; ++l) {
            o += (int64_t)kDims[l] * kDims[l + 1];
            if (j >= o && j < o + kDims[l + 1]) {
                bl = l;
                bj = j - o;
            }
            o += kDims[l + 1];
        }
        float gi;
        if (colpart && bl >= 0) {
            const int R = bl == PT_MLP_LAYERS - 1 ? r_out : r_hidden;
            const int N = kDims[bl + 1];
            const float* c = colpart + bl * slice + (int64_t)z * R * N + bj;
            float s = 0.f;
            for (int r = 0; r < R; ++r) s += c[(int64_t)r * N];
            gi = s;
        } else {
            gi = g[i];
        }
        const float a = 0.9f * m1[i] + 0.1f * gi;
        const float b = 0.999f * m2[i] + 0.001f * gi * gi;
        m1[i] = a;
        m2[i] = b;
        const float mh = a / bc1, vh = b / bc2;
        const float w = p[i] - lr * mh / (sqrtf(vh) + 1e-8f);
        p[i] = w;
        pb[i] = (bf16)w;
    }
}

// Ends a training step: the next step's index and Adam count.
// The end of a training step: k_loss_final for every member (threads
// 0..members-1, loss column step[0]), then the step index and the Adam
// step count advance.
__global__ void k_loss_step(const float* __restrict__ part, int nblk, int members, int64_t n,
                            float* __restrict__ loss, int64_t loss_z, int64_t* __restrict__ step) {
    const int z = threadIdx.x;
    if (z < members) {
        float t = 0.f;
        for (int b = 0; b < nblk; ++b) t += part[z * nblk + b];
        loss[z * loss_z + step[0]] = t / (float)n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        step[0] += 1;
        step[1] += 1;
    }
}

}  // namespace

struct pt_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int members = 0, batch = 0;
    Layout L;
    int64_t t = 0;  // Adam steps taken
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    bf16* wb = nullptr;
    float *X = nullptr, *Y = nullptr;
    int64_t n_data = 0;
    int32_t* order = nullptr;
    int64_t order_cap = 0;
    bf16* xb = nullptr;
    float* yb = nullptr;
    bf16* act[PT_MLP_LAYERS] = {};   // act[0] = xb; act[1..4] hidden outputs
    float* pre[PT_MLP_LAYERS - 1] = {};
    float* pred = nullptr;
    float* dz32[2] = {};
    bf16* dz16[2] = {};
    float* part = nullptr;
    float* colpart = nullptr;  // bias-gradient partial sums [layers][members][row chunks or tiles][kMaxWidth]
    float* loss = nullptr;
    int64_t loss_cap = 0;
    float* xin = nullptr;  // prediction inputs
    int64_t xin_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    // Training steps: [0] step index within the current call, [1] Adam steps
    // taken.  One step is captured as a HIP graph per training call and
    // replayed (use_graphs), or enqueued kernel by kernel.
    int64_t* d_step = nullptr;
    bool use_graphs = true;
};

namespace {

int64_t act_z(const pt_mlp* m) { return (int64_t)m->batch * PT_MLP_HIDDEN; }
int64_t dz_z(const pt_mlp* m) { return (int64_t)m->batch * kMaxWidth; }

template <bool TA, bool TB, int EPI>
void gemm(pt_mlp* m, const GemmArgs& g) {
    const dim3 grid((g.N + kTN - 1) / kTN, (g.M + kTM - 1) / kTM, m->members);
    k_gemm<TA, TB, EPI><<<grid, kGemmThreads, 0, m->stream>>>(g);
}

GemmArgs args() {
    GemmArgs g;
    memset(&g, 0, sizeof g);
    return g;
}

// Forward pass of every member on rows already in xb (rows <= batch).
void forward(pt_mlp* m, int rows) {
    const int64_t P = m->L.count;
    for (int l = 0; l < PT_MLP_LAYERS; ++l) {
        GemmArgs g = args();
        g.A = m->act[l];
        g.lda = kDims[l];
        g.a_z = l == 0 ? 0 : act_z(m);
        g.B = m->wb + m->L.w[l];
        g.ldb = kDims[l + 1];
        g.b_z = P;
        g.M = rows;
        g.N = kDims[l + 1];
        g.K = kDims[l];
        g.bias = m->params + m->L.b[l];
        g.bias_z = P;
        g.ldc = kDims[l + 1];
        if (l < PT_MLP_LAYERS - 1) {
            g.out_f32 = m->pre[l];
            g.f32_z = act_z(m);
            g.out_bf16 = m->act[l + 1];
            g.bf16_z = act_z(m);
            gemm<false, false, EPI_BIAS_GELU>(m, g);
        } else {
            g.out_f32 = m->pred;
            g.f32_z = (int64_t)m->batch * PT_MLP_OUT;
            gemm<false, false, EPI_BIAS>(m, g);
        }
    }
}

int64_t colpart_slice(const pt_mlp* m);

// Loss and gradients of every member for the batch in xb / yb.
// step (may be null): the training step counters (the loss column).
// bias_sums: also add the bias gradients' partial sums into grads and the
// loss partials into loss (else the partials stay for k_adam, in colpart's
// per-layer slices, and for k_loss_step, in part).
void backward(pt_mlp* m, float* loss, int64_t loss_z, const int64_t* step, bool bias_sums) {
    const int64_t P = m->L.count;
    const int B = m->batch;
    k_loss<<<dim3(kLossBlocks, m->members), kLossThreads, 0, m->stream>>>(m->pred, m->yb, B, m->dz32[0], m->dz16[0],
                                                                          dz_z(m), m->part);
    if (bias_sums)
        k_loss_final<<<1, 64, 0, m->stream>>>(m->part, kLossBlocks, m->members, (int64_t)B * PT_MLP_OUT, loss, loss_z,
                                              step);
    int cur = 0;
    for (int l = PT_MLP_LAYERS - 1; l >= 0; --l) {
        const int nout = kDims[l + 1], nin = kDims[l];
        // dW_l = act_l^T dz_l
        GemmArgs g = args();
        g.A = m->act[l];
        g.lda = nin;
        g.a_z = l == 0 ? 0 : act_z(m);
        g.B = m->dz16[cur];
        g.ldb = nout;
        g.b_z = dz_z(m);
        g.M = nin;
        g.N = nout;
        g.K = B;
        g.out_f32 = m->grads + m->L.w[l];
        g.f32_z = P;
        g.ldc = nout;
        gemm<true, false, EPI_F32>(m, g);
        // db_l: the output layer's dz comes from the loss kernel and is summed
        // here in 128-row chunks; a hidden layer's column sums per 64-row
        // tile came with its dz from the input-gradient GEMM's epilogue.
        int R = (B + kTM - 1) / kTM;
        float* part_l = m->colpart + l * colpart_slice(m);
        if (l == PT_MLP_LAYERS - 1) {
            R = (B + kColRows - 1) / kColRows;
            k_colsum_part<<<dim3((nout + kColCols - 1) / kColCols, R, m->members), kColCols * kColGroups, 0,
                            m->stream>>>(m->dz32[cur], B, nout, dz_z(m), part_l);
        }
        if (bias_sums)
            k_colsum_final<<<(m->members * nout + 255) / 256, 256, 0, m->stream>>>(part_l, R, nout, m->members,
                                                                                  m->grads + m->L.b[l], P);
        if (l == 0) break;
        // dz_{l-1} = (dz_l W_l^T) * gelu'(pre_{l-1})
        GemmArgs h = args();
        h.A = m->dz16[cur];
        h.lda = nout;
        h.a_z = dz_z(m);
        h.B = m->wb + m->L.w[l];
        h.ldb = nout;
        h.b_z = P;
        h.M = B;
        h.N = nin;
        h.K = nout;
        h.pre = m->pre[l - 1];
        h.pre_z = act_z(m);
        h.out_f32 = m->dz32[1 - cur];
        h.f32_z = dz_z(m);
        h.out_bf16 = m->dz16[1 - cur];
        h.bf16_z = dz_z(m);
        h.ldc = nin;
        h.colpart = m->colpart + (l - 1) * colpart_slice(m);
        gemm<false, true, EPI_GELU_GRAD>(m, h);
        cur = 1 - cur;
    }
}

// Floats per layer slice of colpart: members x (row tiles of kTM) x kMaxWidth.
int64_t colpart_slice(const pt_mlp* m) { return (int64_t)m->members * ((m->batch + kTM - 1) / kTM) * kMaxWidth; }

int upload_rows(pt_mlp* m, const int32_t* rows, int64_t n) {
    for (int64_t i = 0; i < n; ++i)
        if (rows[i] < 0 || rows[i] >= m->n_data) return fail(PT_EINVAL, "row %d outside the training set", rows[i]);
    if (n > m->order_cap) {
        dev_free(m->order);
        if (dev_alloc(&m->order, n)) return PT_ENOMEM;
        m->order_cap = n;
    }
    HIP_TRY(hipMemcpyAsync(m->order, rows, sizeof(int32_t) * n, hipMemcpyHostToDevice, m->stream));
    return PT_OK;
}

}  // namespace

extern "C" {

int64_t pt_mlp_param_count(void) { return Layout().count; }

int pt_mlp_create(int32_t members, int32_t batch, int32_t device, pt_mlp** out) {
    if (!out) return fail(PT_EINVAL, "null argument");
    *out = nullptr;
    if (members < 1 || members > 64) return fail(PT_EINVAL, "members %d outside [1, 64]", members);
    if (batch < 64 || batch % 64 || batch > 65536) return fail(PT_EINVAL, "batch %d: a multiple of 64 up to 65536", batch);
    HIP_TRY(hipSetDevice(device));
    pt_mlp* m = new pt_mlp();
    m->device = device;
    m->members = members;
    m->batch = batch;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return fail(PT_EDEVICE, "hipStreamCreate failed");
    }
    const int64_t P = m->L.count * members;
    bool bad = dev_alloc(&m->params, P) || dev_alloc(&m->grads, P) || dev_alloc(&m->m1, P) || dev_alloc(&m->m2, P) ||
               dev_alloc(&m->wb, P) || dev_alloc(&m->xb, (int64_t)batch * PT_MLP_IN) ||
               dev_alloc(&m->yb, (int64_t)batch * PT_MLP_OUT) ||
               dev_alloc(&m->pred, (int64_t)members * batch * PT_MLP_OUT) ||
               dev_alloc(&m->part, (int64_t)members * kLossBlocks) ||
               dev_alloc(&m->colpart, (int64_t)PT_MLP_LAYERS * members * ((batch + kTM - 1) / kTM) * kMaxWidth);
    for (int l = 1; l < PT_MLP_LAYERS && !bad; ++l) bad = dev_alloc(&m->act[l], members * act_z(m)) != 0;
    for (int l = 0; l < PT_MLP_LAYERS - 1 && !bad; ++l) bad = dev_alloc(&m->pre[l], members * act_z(m)) != 0;
    for (int i = 0; i < 2 && !bad; ++i)
        bad = dev_alloc(&m->dz32[i], members * dz_z(m)) || dev_alloc(&m->dz16[i], members * dz_z(m));
    if (bad) {
        const std::string msg = pt_last_error();
        pt_mlp_destroy(m);
        return fail(PT_ENOMEM, "%s", msg.c_str());
    }
    m->act[0] = m->xb;
    (void)hipEventCreate(&m->ev0);
    (void)hipEventCreate(&m->ev1);
    if (dev_alloc(&m->d_step, 2)) {
        pt_mlp_destroy(m);
        return PT_ENOMEM;
    }
    *out = m;
    return PT_OK;
}

int pt_mlp_set_params(pt_mlp* m, const float* params) {
    if (!m || !params) return fail(PT_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(m->device));
    const int64_t P = m->L.count * m->members;
    HIP_TRY(hipMemcpyAsync(m->params, params, sizeof(float) * P, hipMemcpyHostToDevice, m->stream));
    HIP_TRY(hipMemsetAsync(m->m1, 0, sizeof(float) * P, m->stream));
    HIP_TRY(hipMemsetAsync(m->m2, 0, sizeof(float) * P, m->stream));
    k_to_bf16<<<grid_for(P, 256, 8192), 256, 0, m->stream>>>(m->params, P, m->wb);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(m->stream));
    m->t = 0;
    return PT_OK;
}

int pt_mlp_get_params(pt_mlp* m, float* params) {
    if (!m || !params) return fail(PT_EINVAL, "null argument");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(params, m->params, sizeof(float) * m->L.count * m->members, hipMemcpyDeviceToHost));
    return PT_OK;
}

int pt_mlp_set_data(pt_mlp* m, const float* X, const float* Y, int64_t n) {
    if (!m || !X || !Y || n <= 0) return fail(PT_EINVAL, "bad training set");
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipStreamSynchronize(m->stream));
    dev_free(m->X);
    dev_free(m->Y);
    if (dev_alloc(&m->X, n * PT_MLP_IN) || dev_alloc(&m->Y, n * PT_MLP_OUT)) return PT_ENOMEM;
    HIP_TRY(hipMemcpy(m->X, X, sizeof(float) * n * PT_MLP_IN, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->Y, Y, sizeof(float) * n * PT_MLP_OUT, hipMemcpyHostToDevice));
    m->n_data = n;
    return PT_OK;
}

int pt_mlp_train(pt_mlp* m, const int32_t* order, int32_t steps, float lr, float* loss_out) {
    if (!m || !order || steps < 1) return fail(PT_EINVAL, "bad arguments");
    if (m->n_data == 0) return fail(PT_ESTATE, "train before set_data");
    HIP_TRY(hipSetDevice(m->device));
    const int B = m->batch;
    if (int rc = upload_rows(m, order, (int64_t)steps * B)) return rc;
    if ((int64_t)steps * m->members > m->loss_cap) {
        dev_free(m->loss);
        if (dev_alloc(&m->loss, (int64_t)steps * m->members)) return PT_ENOMEM;
        m->loss_cap = (int64_t)steps * m->members;
    }
    const int64_t P = m->L.count * m->members;
    // Every kernel of a step reads the step index and Adam count from
    // d_step, so one captured step serves all of them.
    const int64_t h_step[2] = {0, m->t};
    HIP_TRY(hipMemcpyAsync(m->d_step, h_step, sizeof h_step, hipMemcpyHostToDevice, m->stream));
    auto one_step = [&]() {
        k_gather<<<grid_for((int64_t)B * (PT_MLP_IN + PT_MLP_OUT), 256, 4096), 256, 0, m->stream>>>(
            m->X, m->Y, m->order, m->d_step, B, m->xb, m->yb);
        forward(m, B);
        backward(m, m->loss, steps, m->d_step, false);
        k_adam<<<dim3(grid_for(m->L.count, 256, 2048), m->members), 256, 0, m->stream>>>(
            m->params, m->grads, m->m1, m->m2, m->wb, m->L.count, lr, m->d_step, m->colpart, colpart_slice(m),
            (B + kTM - 1) / kTM, (B + kColRows - 1) / kColRows);
        k_loss_step<<<1, 64, 0, m->stream>>>(m->part, kLossBlocks, m->members, (int64_t)B * PT_MLP_OUT, m->loss,
                                             steps, m->d_step);
    };
    // The instantiated step graph, released on every return.
    struct ExecGuard {
        hipGraphExec_t e = nullptr;
        ~ExecGuard() {
            if (e) (void)hipGraphExecDestroy(e);
        }
    } exec;
    if (m->use_graphs) {
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
        one_step();
        // Ends the capture whatever happened inside it; a failed launch
        // there is reported here.
        const hipError_t ce = hipStreamEndCapture(m->stream, &graph);
        if (ce != hipSuccess || !graph) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(PT_EDEVICE, "training graph capture: %s", hipGetErrorString(ce));
        }
        const hipError_t e = hipGraphInstantiate(&exec.e, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return fail(PT_EDEVICE, "training graph: %s", hipGetErrorString(e));
    }
    HIP_TRY(hipEventRecord(m->ev0, m->stream));
    for (int s = 0; s < steps; ++s) {
        if (exec.e) {
            const hipError_t e = hipGraphLaunch(exec.e, m->stream);
            if (e != hipSuccess) return fail(PT_EDEVICE, "training graph launch: %s", hipGetErrorString(e));
        } else {
            one_step();
        }
    }
    m->t += steps;
    HIP_TRY(hipEventRecord(m->ev1, m->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(m->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    m->last_ms = ms;
    if (loss_out) {
        std::vector<float> h((size_t)steps * m->members);
        HIP_TRY(hipMemcpy(h.data(), m->loss, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
        // device layout [member][step] -> [step][member]
        for (int s = 0; s < steps; ++s)
            for (int k = 0; k < m->members; ++k) loss_out[(size_t)s * m->members + k] = h[(size_t)k * steps + s];
    }
    return PT_OK;
}

int pt_mlp_gradients(pt_mlp* m, const int32_t* rows, float* grads_out, float* loss_out) {
    if (!m || !rows || !grads_out || !loss_out) return fail(PT_EINVAL, "null argument");
    if (m->n_data == 0) return fail(PT_ESTATE, "gradients before set_data");
    HIP_TRY(hipSetDevice(m->device));
    const int B = m->batch;
    if (int rc = upload_rows(m, rows, B)) return rc;
    if (m->members > m->loss_cap) {
        dev_free(m->loss);
        if (dev_alloc(&m->loss, m->members)) return PT_ENOMEM;
        m->loss_cap = m->members;
    }
    k_gather<<<grid_for((int64_t)B * (PT_MLP_IN + PT_MLP_OUT), 256, 4096), 256, 0, m->stream>>>(
        m->X, m->Y, m->order, (const int64_t*)nullptr, B, m->xb, m->yb);
    forward(m, B);
    backward(m, m->loss, 1, nullptr, true);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(m->stream));
    HIP_TRY(hipMemcpy(grads_out, m->grads, sizeof(float) * m->L.count * m->members, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(loss_out, m->loss, sizeof(float) * m->members, hipMemcpyDeviceToHost));
    return PT_OK;
}

int pt_mlp_predict(pt_mlp* m, const float* X, int64_t n, float* out) {
    if (!m || !X || !out || n < 0) return fail(PT_EINVAL, "bad arguments");
    HIP_TRY(hipSetDevice(m->device));
    if (n == 0) return PT_OK;
    if (n * PT_MLP_IN > m->xin_cap) {
        dev_free(m->xin);
        if (dev_alloc(&m->xin, n * PT_MLP_IN)) return PT_ENOMEM;
        m->xin_cap = n * PT_MLP_IN;
    }
    HIP_TRY(hipMemcpyAsync(m->xin, X, sizeof(float) * n * PT_MLP_IN, hipMemcpyHostToDevice, m->stream));
    const int B = m->batch;
    std::vector<float> h((size_t)m->members * B * PT_MLP_OUT);
    for (int64_t i0 = 0; i0 < n; i0 += B) {
        const int rows = (int)std::min<int64_t>(B, n - i0);
        k_to_bf16<<<grid_for((int64_t)rows * PT_MLP_IN, 256, 1024), 256, 0, m->stream>>>(
            m->xin + i0 * PT_MLP_IN, (int64_t)rows * PT_MLP_IN, m->xb);
        forward(m, rows);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(h.data(), m->pred, sizeof(float) * h.size(), hipMemcpyDeviceToHost, m->stream));
        HIP_TRY(hipStreamSynchronize(m->stream));
        for (int k = 0; k < m->members; ++k)
            memcpy(out + ((size_t)k * n + i0) * PT_MLP_OUT, h.data() + (size_t)k * B * PT_MLP_OUT,
                   sizeof(float) * rows * PT_MLP_OUT);
    }
    return PT_OK;
}

int pt_mlp_use_graphs(pt_mlp* m, int32_t on) {
    if (!m) return fail(PT_EINVAL, "null handle");
    m->use_graphs = on != 0;
    return PT_OK;
}

double pt_mlp_last_train_ms(pt_mlp* m) { return m ? m->last_ms : -1.0; }

void pt_mlp_destroy(pt_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    dev_free(m->params);
    dev_free(m->grads);
    dev_free(m->m1);
    dev_free(m->m2);
    dev_free(m->wb);
    dev_free(m->X);
    dev_free(m->Y);
    dev_free(m->order);
    dev_free(m->xb);
    dev_free(m->yb);
    for (int l = 1; l < PT_MLP_LAYERS; ++l) dev_free(m->act[l]);
    for (int l = 0; l < PT_MLP_LAYERS - 1; ++l) dev_free(m->pre[l]);
    for (int i = 0; i < 2; ++i) {
        dev_free(m->dz32[i]);
        dev_free(m->dz16[i]);
    }
    dev_free(m->pred);
    dev_free(m->part);
    dev_free(m->colpart);
    dev_free(m->loss);
    dev_free(m->xin);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    dev_free(m->d_step);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

}  // extern "C"
