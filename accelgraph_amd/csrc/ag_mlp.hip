// ag_mlp.hip -- learned warm start: features, a 3-layer MLP on BF16 MFMA, training, inference.
//
// Network: 5 inputs (zero-padded to 16) -> 64 -> 64 -> 1, ReLU after the hidden layers,
// predicting (standardised) log(r*) - log(r2): a vertex's converged scaled rank relative to its
// scaled rank after two iterations from the uniform start. A wavefront processes 16 samples:
// the two hidden layers and the backward GEMMs run on v_mfma_f32_16x16x16bf16_1k (BF16
// operands, FP32 accumulators); the output layer, the biases and the loss are FP32 on VALU.
// Weights live in FP32 and are rounded to BF16 into LDS at kernel start. Training runs one
// fused forward + backward kernel per minibatch that writes per-wavefront gradient partials;
// a reduction kernel adds them in a fixed order and an Adam kernel updates the FP32 weights.
//
// MFMA 16x16x16 operand layout (64 lanes, lane l, q = l / 16, i = 0..3):
//   A (16 x 16, row-major M x K): lane l holds A[l % 16][4q + i]
//   B (16 x 16, K x N):           lane l holds B[4q + i][l % 16]
//   D (16 x 16):                  lane l, register i holds D[4q + i][l % 16]
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ag_internal.h"
#include "ag_pull_kernel.h"

#define MLP_IN 16
#define MLP_H 64
#define MLP_NF AG_MLP_FEATURES
#define MLP_FSTRIDE 8   // floats per sample in the device feature buffer
#define MLP_WAVES 4
#define MLP_WG (MLP_WAVES * AG_WAVE)

// parameter offsets
#define P_W1 0
#define P_B1 (P_W1 + MLP_IN * MLP_H)
#define P_W2 (P_B1 + MLP_H)
#define P_B2 (P_W2 + MLP_H * MLP_H)
#define P_W3 (P_B2 + MLP_H)
#define P_B3 (P_W3 + MLP_H)
#define P_COUNT (P_B3 + 1)

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

struct MlpNorm {
    float in_mean[8];
    float in_std[8];
    float y_mean, y_std;
};

struct ag_mlp {
    ag_device *dev;
    float *p;        // [P] FP32 parameters
    float *m1, *m2;  // Adam moments
    float *grad;     // [P]
    uint32_t adam_t;
    MlpNorm norm;
    // training buffers (grown on demand)
    float *X;        // [S][8]
    float *y;        // [S]
    uint32_t *perm;  // [S]
    size_t cap;
    float *part;     // [waves][P]
    size_t part_cap;
    double *loss_part;
};

extern "C" uint32_t ag_mlp_param_count(void) { return P_COUNT; }

namespace {

__device__ __forceinline__ uint16_t bf16_rne(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// A fragment from a row-major bf16 matrix M (leading dimension ld): A[r][k] = M[r][k]
__device__ __forceinline__ mlp_s4 frag_a(const uint16_t *M, int ld, int r0, int k0, int l)
{
    const uint16_t *p = M + (r0 + (l & 15)) * ld + k0 + 4 * (l >> 4);
    return mlp_s4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
}
// A fragment of a transposed matrix: A[r][k] = M[k][r]
__device__ __forceinline__ mlp_s4 frag_at(const uint16_t *M, int ld, int r0, int k0, int l)
{
    const int r = r0 + (l & 15), k = k0 + 4 * (l >> 4);
    return mlp_s4{(short)M[k * ld + r], (short)M[(k + 1) * ld + r], (short)M[(k + 2) * ld + r],
                  (short)M[(k + 3) * ld + r]};
}
// B fragment: B[k][n] = M[k][n]
__device__ __forceinline__ mlp_s4 frag_b(const uint16_t *M, int ld, int k0, int n0, int l)
{
    const int n = n0 + (l & 15), k = k0 + 4 * (l >> 4);
    return mlp_s4{(short)M[k * ld + n], (short)M[(k + 1) * ld + n], (short)M[(k + 2) * ld + n],
                  (short)M[(k + 3) * ld + n]};
}
// B fragment of a transposed matrix: B[k][n] = M[n][k]
__device__ __forceinline__ mlp_s4 frag_bt(const uint16_t *M, int ld, int k0, int n0, int l)
{
    const uint16_t *p = M + (n0 + (l & 15)) * ld + k0 + 4 * (l >> 4);
    return mlp_s4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
}

__device__ __forceinline__ mlp_f4 mfma(mlp_s4 a, mlp_s4 b, mlp_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct MlpArgs {
    const float *p;          // parameters
    const float *X;          // [S][8] raw features
    const float *y;          // [S] raw targets (training) or null
    const uint32_t *perm;    // sample order (training) or null
    uint32_t S;              // samples in this call
    uint32_t off;            // first position of this minibatch in perm
    float *out;              // predictions (raw units) or null
    float *part;             // [waves][P] gradient partials (training)
    double *loss_part;       // [waves]
    float inv_b;             // 2 / batch for dL/dy
    MlpNorm norm;
};

// One fused kernel: forward for 16 samples per wavefront, and with TRAIN the backward pass.
template <bool TRAIN>
__global__ __launch_bounds__(MLP_WG) void mlp_kernel(MlpArgs A)
{
    __shared__ uint16_t sW1[MLP_IN * MLP_H];
    __shared__ uint16_t sW2[MLP_H * MLP_H];
    __shared__ float sB1[MLP_H], sB2[MLP_H], sW3[MLP_H];
    __shared__ uint16_t sX[MLP_WAVES][16 * MLP_IN];
    __shared__ uint16_t sH1[MLP_WAVES][16 * MLP_H];
    __shared__ uint16_t sD2[MLP_WAVES][16 * MLP_H];
    __shared__ uint16_t sD1[MLP_WAVES][16 * MLP_H];
    const int tid = threadIdx.x, w = tid / AG_WAVE, l = tid % AG_WAVE, q = l >> 4, c16 = l & 15;
    for (int i = tid; i < MLP_IN * MLP_H; i += MLP_WG)
        sW1[i] = bf16_rne(A.p[P_W1 + i]);
    for (int i = tid; i < MLP_H * MLP_H; i += MLP_WG)
        sW2[i] = bf16_rne(A.p[P_W2 + i]);
    for (int i = tid; i < MLP_H; i += MLP_WG) {
        sB1[i] = A.p[P_B1 + i];
        sB2[i] = A.p[P_B2 + i];
        sW3[i] = A.p[P_W3 + i];
    }
    const float b3 = A.p[P_B3];
    __syncthreads();

    mlp_f4 gW2[4][4], gW1[4];
    float gb1[4], gb2[4], gw3[4], gb3 = 0.0f;
    double loss = 0.0;
    if (TRAIN) {
#pragma unroll
        for (int a = 0; a < 4; a++) {
            gW1[a] = mlp_f4{0, 0, 0, 0};
            gb1[a] = gb2[a] = gw3[a] = 0.0f;
#pragma unroll
            for (int b = 0; b < 4; b++)
                gW2[a][b] = mlp_f4{0, 0, 0, 0};
        }
    }
    const uint32_t ntiles = (A.S + 15) / 16;
    const uint32_t ngroups = (ntiles + MLP_WAVES - 1) / MLP_WAVES;
    uint16_t *X = sX[w], *H1 = sH1[w], *D2 = sD2[w], *D1 = sD1[w];
    for (uint32_t gidx = blockIdx.x; gidx < ngroups; gidx += gridDim.x) {
        const uint32_t tile = gidx * MLP_WAVES + w;
        const uint32_t s0 = tile * 16;
        // standardised features, bf16, [16 samples][16 inputs]
        {
            const int r = l >> 2, k0 = (l & 3) * 4;
            const uint32_t sidx = s0 + r;
            const bool ok = tile < ntiles && sidx < A.S;
            const uint32_t src = ok ? (A.perm ? A.perm[A.off + sidx] : A.off + sidx) : 0u;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int k = k0 + i;
                float v = 0.0f;
                if (ok && k < MLP_NF)
                    v = __fdiv_rn(__fsub_rn(A.X[(size_t)src * MLP_FSTRIDE + k], A.norm.in_mean[k]),
                                  A.norm.in_std[k]);
                X[r * MLP_IN + k] = bf16_rne(v);
            }
        }
        __syncthreads();
        // layer 1: Z1 = X W1 + b1, H1 = relu(Z1)
        const mlp_s4 ax = frag_a(X, MLP_IN, 0, 0, l);
        mlp_f4 h1[4];
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            h1[nb] = mfma(ax, frag_b(sW1, MLP_H, 0, nb * 16, l), mlp_f4{0, 0, 0, 0});
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = h1[nb][i] + sB1[nb * 16 + c16];
                h1[nb][i] = z > 0.0f ? z : 0.0f;
                H1[(4 * q + i) * MLP_H + nb * 16 + c16] = bf16_rne(h1[nb][i]);
            }
        }
        __syncthreads();
        // layer 2: Z2 = H1 W2 + b2, H2 = relu(Z2)
        mlp_f4 h2[4];
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            mlp_f4 acc = mlp_f4{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ks++)
                acc = mfma(frag_a(H1, MLP_H, 0, ks * 16, l), frag_b(sW2, MLP_H, ks * 16, nb * 16, l), acc);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = acc[i] + sB2[nb * 16 + c16];
                h2[nb][i] = z > 0.0f ? z : 0.0f;
            }
        }
        // layer 3: y = H2 w3 + b3 (rows 4q + i; sum over the 64 columns)
        float yv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float t = 0.0f;
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
                t += h2[nb][i] * sW3[nb * 16 + c16];
            yv[i] = group_reduce<float, 16>(t) + b3;
        }
        if (!TRAIN) {
            if (c16 == 0 && tile < ntiles) {
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t sidx = s0 + 4 * q + i;
                    if (sidx < A.S)
                        A.out[A.off + sidx] = yv[i] * A.norm.y_std + A.norm.y_mean;
                }
            }
            __syncthreads();
            continue;
        }
        // loss and dL/dy (standardised target)
        float dy[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t sidx = s0 + 4 * q + i;
            const bool ok = tile < ntiles && sidx < A.S;
            float d = 0.0f;
            if (ok) {
                const uint32_t src = A.perm ? A.perm[A.off + sidx] : A.off + sidx;
                const float t = __fdiv_rn(__fsub_rn(A.y[src], A.norm.y_mean), A.norm.y_std);
                d = yv[i] - t;
                if (c16 == 0)
                    loss += (double)d * (double)d;
            }
            dy[i] = d * A.inv_b;
        }
        // output layer gradients; dZ2 = dy w3 * relu'(Z2)
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                gw3[nb] += dy[i] * h2[nb][i];
                const float dz = h2[nb][i] > 0.0f ? dy[i] * sW3[nb * 16 + c16] : 0.0f;
                gb2[nb] += dz;
                D2[(4 * q + i) * MLP_H + nb * 16 + c16] = bf16_rne(dz);
            }
        }
        if (c16 == 0)
            gb3 += (dy[0] + dy[1]) + (dy[2] + dy[3]);
        __syncthreads();
        // dW2 += H1^T dZ2
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const mlp_s4 a = frag_at(H1, MLP_H, mb * 16, 0, l);
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
                gW2[mb][nb] = mfma(a, frag_b(D2, MLP_H, 0, nb * 16, l), gW2[mb][nb]);
        }
        // dH1 = dZ2 W2^T; dZ1 = dH1 * relu'(Z1)
#pragma unroll
        for (int nb = 0; nb < 4; nb++) {
            mlp_f4 acc = mlp_f4{0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < 4; ks++)
                acc = mfma(frag_a(D2, MLP_H, 0, ks * 16, l), frag_bt(sW2, MLP_H, ks * 16, nb * 16, l), acc);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float dz = h1[nb][i] > 0.0f ? acc[i] : 0.0f;
                gb1[nb] += dz;
                D1[(4 * q + i) * MLP_H + nb * 16 + c16] = bf16_rne(dz);
            }
        }
        __syncthreads();
        // dW1 += X^T dZ1
        {
            const mlp_s4 a = frag_at(X, MLP_IN, 0, 0, l);
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
                gW1[nb] = mfma(a, frag_b(D1, MLP_H, 0, nb * 16, l), gW1[nb]);
        }
        __syncthreads();
    }
    if (!TRAIN)
        return;
    // per-wavefront partials
    float *P = A.part + (size_t)(blockIdx.x * MLP_WAVES + w) * P_COUNT;
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            P[P_W1 + (4 * q + i) * MLP_H + nb * 16 + c16] = gW1[nb][i];
#pragma unroll
            for (int mb = 0; mb < 4; mb++)
                P[P_W2 + (mb * 16 + 4 * q + i) * MLP_H + nb * 16 + c16] = gW2[mb][nb][i];
        }
        // bias / w3 partials: sum over the four row groups q (lanes c16, c16+16, +32, +48)
        float a = gb1[nb], b = gb2[nb], c = gw3[nb];
        a += __shfl_xor(a, 16);
        a += __shfl_xor(a, 32);
        b += __shfl_xor(b, 16);
        b += __shfl_xor(b, 32);
        c += __shfl_xor(c, 16);
        c += __shfl_xor(c, 32);
        if (q == 0) {
            P[P_B1 + nb * 16 + c16] = a;
            P[P_B2 + nb * 16 + c16] = b;
            P[P_W3 + nb * 16 + c16] = c;
        }
    }
    float tb3 = gb3;
    tb3 += __shfl_xor(tb3, 16);
    tb3 += __shfl_xor(tb3, 32);
    double tl = loss;
    tl += __shfl_xor(tl, 16);
    tl += __shfl_xor(tl, 32);
    if (l == 0) {
        P[P_B3] = tb3;
        A.loss_part[blockIdx.x * MLP_WAVES + w] = tl;
    }
}

// grad[i] = sum over partials, in index order
__global__ void mlp_reduce_kernel(const float *part, uint32_t nparts, float *grad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P_COUNT)
        return;
    float t = 0.0f;
    for (uint32_t k = 0; k < nparts; k++)
        t += part[(size_t)k * P_COUNT + i];
    grad[i] = t;
}

__global__ void mlp_adam_kernel(float *p, float *m1, float *m2, const float *grad, float lr,
                                float bc1, float bc2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P_COUNT)
        return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.0f - b1) * g;
    const float v = b2 * m2[i] + (1.0f - b2) * g * g;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
}

// ---------------------------------------------------------------- features

// mean in-degree of the in-neighbours: a pull over the plan with the sources' in-degree
struct NbrDegPol {
    using Acc = float;
    const uint32_t *off;
    float *feat;     // [n][8] device order
    __device__ __forceinline__ float edge(uint32_t, uint32_t s) const
    {
        return __uint2float_rn(off[s + 1] - off[s]);
    }
    __device__ __forceinline__ void finish(uint32_t r, float sum, Red &) const
    {
        const uint32_t d = off[r + 1] - off[r];
        feat[(size_t)r * MLP_FSTRIDE + 2] = d ? log1pf(__fdiv_rn(sum, __uint2float_rn(d))) : 0.0f;
    }
};

__global__ void feat_local_kernel(uint32_t n, const uint32_t *off, const uint32_t *outdeg,
                                  const float *r1, const float *r2, float *feat)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    float *f = feat + (size_t)r * MLP_FSTRIDE;
    f[0] = log1pf(__uint2float_rn(off[r + 1] - off[r]));
    f[1] = log1pf(__uint2float_rn(outdeg[r]));
    f[3] = logf(r1[r]);
    f[4] = logf(r2[r]);
    f[5] = f[6] = f[7] = 0.0f;
}

__global__ void feat_out_kernel(uint32_t n, const float *feat, const uint32_t *d2o, float *out)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    for (int k = 0; k < MLP_NF; k++)
        out[(size_t)d2o[r] * MLP_NF + k] = feat[(size_t)r * MLP_FSTRIDE + k];
}

__global__ void pack_features_kernel(uint32_t S, const float *in5, float *out8)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S)
        return;
    for (int k = 0; k < MLP_FSTRIDE; k++)
        out8[(size_t)i * MLP_FSTRIDE + k] = k < MLP_NF ? in5[(size_t)i * MLP_NF + k] : 0.0f;
}

// warm start, pass 1: rank = r2 * exp(prediction) (the network predicts log(r*) - log(r2));
// per-workgroup FP64 sums of the whole start vector and of its dangling part
__global__ __launch_bounds__(AG_WG) void warm_apply_kernel(uint32_t n, const float *pred,
                                                           const float *feat, const uint32_t *outdeg,
                                                           float *rank, double *part)
{
    __shared__ double s_w[AG_WG / AG_WAVE];
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    double all = 0.0, dang = 0.0;
    if (r < n) {
        const float x = expf(__fadd_rn(pred[r], feat[(size_t)r * MLP_FSTRIDE + 4]));
        rank[r] = x;
        all = (double)x;
        dang = outdeg[r] ? 0.0 : (double)x;
    }
    all = block_reduce<double, AG_WG>(all, s_w);
    dang = block_reduce<double, AG_WG>(dang, s_w);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = all;
        part[2 * blockIdx.x + 1] = dang;
    }
}

// pass 2 (one workgroup): scale alpha that makes the start's total mass satisfy the fixed-point
// balance m = (1 - d) N + d (m - m_dangling), i.e. alpha = (1 - d) N / ((1 - d) m0 + d m0_dangling)
__global__ void warm_mass_kernel(const double *part, uint32_t nblk, double n, double *alpha)
{
    __shared__ double s_w[AG_WG / AG_WAVE];
    double all = 0.0, dang = 0.0;
    for (uint32_t i = threadIdx.x; i < nblk; i += AG_WG) {
        all += part[2 * i];
        dang += part[2 * i + 1];
    }
    all = block_reduce<double, AG_WG>(all, s_w);
    dang = block_reduce<double, AG_WG>(dang, s_w);
    if (threadIdx.x == 0) {
        const double d = 0.85, den = (1.0 - d) * all + d * dang;
        *alpha = den > 0.0 ? (1.0 - d) * n / den : 1.0;
    }
}

// pass 3: rank *= alpha, contribution = rank / out_degree
__global__ void warm_scale_kernel(uint32_t n, const double *alpha, const uint32_t *outdeg,
                                  float *rank, float *contrib)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    const float x = __fmul_rn(rank[r], (float)*alpha);
    rank[r] = x;
    const uint32_t od = outdeg[r];
    contrib[r] = od ? __fdiv_rn(x, __uint2float_rn(od)) : 0.0f;
}

__global__ void restore_kernel(uint32_t n, const float *saved, const uint32_t *outdeg, float *rank,
                               float *contrib)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    const float x = saved[r];
    rank[r] = x;
    const uint32_t od = outdeg[r];
    contrib[r] = od ? __fdiv_rn(x, __uint2float_rn(od)) : 0.0f;
}

} // namespace

// ---------------------------------------------------------------- host side

static int mlp_grow(ag_mlp *m, uint32_t S)
{
    if (m->cap >= S)
        return AG_OK;
    if (m->X) (void)hipFree(m->X);
    if (m->y) (void)hipFree(m->y);
    if (m->perm) (void)hipFree(m->perm);
    m->X = NULL;
    m->y = NULL;
    m->perm = NULL;
    m->cap = 0;
    AG_HIP(hipMalloc((void **)&m->X, (size_t)S * MLP_FSTRIDE * sizeof(float)));
    AG_HIP(hipMalloc((void **)&m->y, (size_t)S * sizeof(float)));
    AG_HIP(hipMalloc((void **)&m->perm, (size_t)S * sizeof(uint32_t)));
    m->cap = S;
    return AG_OK;
}

static uint32_t mlp_grid(const ag_mlp *m, uint32_t S)
{
    const uint32_t groups = ((S + 15) / 16 + MLP_WAVES - 1) / MLP_WAVES;
    uint32_t g = m->dev->num_cus * 4;
    return g < groups ? g : (groups ? groups : 1);
}

static int mlp_part(ag_mlp *m, uint32_t grid)
{
    const size_t need = (size_t)grid * MLP_WAVES;
    if (m->part_cap >= need)
        return AG_OK;
    if (m->part) (void)hipFree(m->part);
    if (m->loss_part) (void)hipFree(m->loss_part);
    m->part = NULL;
    m->loss_part = NULL;
    m->part_cap = 0;
    AG_HIP(hipMalloc((void **)&m->part, need * P_COUNT * sizeof(float)));
    AG_HIP(hipMalloc((void **)&m->loss_part, need * sizeof(double)));
    m->part_cap = need;
    return AG_OK;
}

// upload S raw samples (X [S][5]) into the padded device buffer
static int mlp_upload(ag_mlp *m, const float *X, const float *y, uint32_t S)
{
    int st = mlp_grow(m, S);
    if (st != AG_OK)
        return st;
    hipStream_t s = m->dev->stream;
    // stage X[S][5] in the perm+y area is too small; use a temporary
    float *tmp = NULL;
    AG_HIP(hipMalloc((void **)&tmp, (size_t)S * MLP_NF * sizeof(float)));
    hipError_t e = hipMemcpyAsync(tmp, X, (size_t)S * MLP_NF * sizeof(float), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        pack_features_kernel<<<(S + 255) / 256, 256, 0, s>>>(S, tmp, m->X);
        e = hipGetLastError();
    }
    if (e == hipSuccess && y)
        e = hipMemcpyAsync(m->y, y, (size_t)S * sizeof(float), hipMemcpyHostToDevice, s);
    st = (e == hipSuccess) ? ag_wait_stream(m->dev) : ag_hip_fail(e, "MLP upload", __FILE__, __LINE__);
    (void)hipFree(tmp);
    return st;
}

extern "C" int ag_mlp_create(ag_device *dev, uint32_t seed, ag_mlp **out)
{
    if (!dev || !out)
        return ag_set_error(AG_ERR_INVALID, "null argument");
    *out = NULL;
    AG_HIP(hipSetDevice(dev->ordinal));
    ag_mlp *m = (ag_mlp *)calloc(1, sizeof(ag_mlp));
    if (!m)
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    m->dev = dev;
    for (int k = 0; k < 8; k++) {
        m->norm.in_mean[k] = 0.0f;
        m->norm.in_std[k] = 1.0f;
    }
    m->norm.y_mean = 0.0f;
    m->norm.y_std = 1.0f;
    float *h = (float *)calloc(P_COUNT, sizeof(float));
    if (!h) {
        free(m);
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    }
    // He-uniform init from a 32-bit LCG (deterministic for a seed); unused input rows stay 0
    uint64_t x = 0x9E3779B97F4A7C15ull ^ seed;
    auto rnd = [&]() {
        x = x * 6364136223846793005ull + 1442695040888963407ull;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    for (int k = 0; k < MLP_NF; k++)
        for (int j = 0; j < MLP_H; j++)
            h[P_W1 + k * MLP_H + j] = (float)(rnd() * sqrt(6.0 / MLP_NF));
    for (int k = 0; k < MLP_H * MLP_H; k++)
        h[P_W2 + k] = (float)(rnd() * sqrt(6.0 / MLP_H));
    for (int k = 0; k < MLP_H; k++)
        h[P_W3 + k] = (float)(rnd() * sqrt(3.0 / MLP_H));
    hipError_t e = hipMalloc((void **)&m->p, P_COUNT * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->m1, P_COUNT * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->m2, P_COUNT * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&m->grad, P_COUNT * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(m->p, h, P_COUNT * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(m->m1, 0, P_COUNT * sizeof(float));
    if (e == hipSuccess) e = hipMemset(m->m2, 0, P_COUNT * sizeof(float));
    free(h);
    if (e != hipSuccess) {
        ag_mlp_free(m);
        return ag_hip_fail(e, "MLP allocation", __FILE__, __LINE__);
    }
    *out = m;
    return AG_OK;
}

extern "C" int ag_mlp_free(ag_mlp *m)
{
    if (!m)
        return AG_OK;
    (void)hipSetDevice(m->dev->ordinal);
    (void)hipStreamSynchronize(m->dev->stream);
    void *ptrs[] = {m->p, m->m1, m->m2, m->grad, m->X, m->y, m->perm, m->part, m->loss_part};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    free(m);
    return AG_OK;
}

extern "C" int ag_mlp_set_params(ag_mlp *m, const float *params)
{
    if (!m || !params)
        return ag_set_error(AG_ERR_INVALID, "null argument");
    AG_HIP(hipSetDevice(m->dev->ordinal));
    AG_HIP(hipMemcpy(m->p, params, P_COUNT * sizeof(float), hipMemcpyHostToDevice));
    AG_HIP(hipMemset(m->m1, 0, P_COUNT * sizeof(float)));
    AG_HIP(hipMemset(m->m2, 0, P_COUNT * sizeof(float)));
    m->adam_t = 0;
    return AG_OK;
}

extern "C" int ag_mlp_get_params(ag_mlp *m, float *params)
{
    if (!m || !params)
        return ag_set_error(AG_ERR_INVALID, "null argument");
    AG_HIP(hipSetDevice(m->dev->ordinal));
    AG_HIP(hipStreamSynchronize(m->dev->stream));
    AG_HIP(hipMemcpy(params, m->p, P_COUNT * sizeof(float), hipMemcpyDeviceToHost));
    return AG_OK;
}

extern "C" int ag_mlp_set_normalization(ag_mlp *m, const float *in_mean, const float *in_std,
                                        float y_mean, float y_std)
{
    if (!m || !in_mean || !in_std || !(y_std > 0.0f))
        return ag_set_error(AG_ERR_INVALID, "bad normalisation");
    for (int k = 0; k < MLP_NF; k++) {
        if (!(in_std[k] > 0.0f))
            return ag_set_error(AG_ERR_INVALID, "std of feature %d must be > 0", k);
        m->norm.in_mean[k] = in_mean[k];
        m->norm.in_std[k] = in_std[k];
    }
    m->norm.y_mean = y_mean;
    m->norm.y_std = y_std;
    return AG_OK;
}

static MlpArgs mlp_args(ag_mlp *m, uint32_t S)
{
    MlpArgs A;
    memset(&A, 0, sizeof(A));
    A.p = m->p;
    A.X = m->X;
    A.y = m->y;
    A.S = S;
    A.norm = m->norm;
    A.part = m->part;
    A.loss_part = m->loss_part;
    return A;
}

extern "C" int ag_mlp_loss_grad(ag_mlp *m, const float *X, const float *y, uint32_t S,
                                double *loss, float *grad)
{
    if (!m || !X || !y || S == 0)
        return ag_set_error(AG_ERR_INVALID, "bad argument");
    AG_HIP(hipSetDevice(m->dev->ordinal));
    int st = mlp_upload(m, X, y, S);
    const uint32_t grid = mlp_grid(m, S);
    if (st == AG_OK)
        st = mlp_part(m, grid);
    if (st != AG_OK)
        return st;
    hipStream_t s = m->dev->stream;
    MlpArgs A = mlp_args(m, S);
    A.inv_b = 2.0f / (float)S;
    mlp_kernel<true><<<grid, MLP_WG, 0, s>>>(A);
    mlp_reduce_kernel<<<(P_COUNT + 255) / 256, 256, 0, s>>>(m->part, grid * MLP_WAVES, m->grad);
    AG_HIP(hipGetLastError());
    double *lp = (double *)malloc((size_t)grid * MLP_WAVES * sizeof(double));
    if (!lp)
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    hipError_t e = hipMemcpyAsync(lp, m->loss_part, (size_t)grid * MLP_WAVES * sizeof(double),
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && grad)
        e = hipMemcpyAsync(grad, m->grad, P_COUNT * sizeof(float), hipMemcpyDeviceToHost, s);
    st = (e == hipSuccess) ? ag_wait_stream(m->dev) : ag_hip_fail(e, "MLP grad", __FILE__, __LINE__);
    if (st == AG_OK && loss) {
        double t = 0.0;
        for (uint32_t i = 0; i < grid * MLP_WAVES; i++)
            t += lp[i];
        *loss = t / S;
    }
    free(lp);
    return st;
}

extern "C" int ag_mlp_train(ag_mlp *m, const float *X, const float *y, uint32_t S,
                            uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                            float *loss_per_epoch)
{
    if (!m || !X || !y || S == 0 || batch == 0 || !(lr > 0.0f))
        return ag_set_error(AG_ERR_INVALID, "bad argument");
    AG_HIP(hipSetDevice(m->dev->ordinal));
    int st = mlp_upload(m, X, y, S);
    if (batch > S)
        batch = S;
    const uint32_t grid = mlp_grid(m, batch);
    if (st == AG_OK)
        st = mlp_part(m, grid);
    if (st != AG_OK)
        return st;
    hipStream_t s = m->dev->stream;
    uint32_t *perm = (uint32_t *)malloc((size_t)S * sizeof(uint32_t));
    double *lp = (double *)malloc((size_t)grid * MLP_WAVES * sizeof(double));
    if (!perm || !lp) {
        free(perm);
        free(lp);
        return ag_set_error(AG_ERR_OOM, "host allocation failed");
    }
    for (uint32_t i = 0; i < S; i++)
        perm[i] = i;
    uint64_t x = 0xD1B54A32D192ED03ull ^ ((uint64_t)seed << 1);
    for (uint32_t ep = 0; ep < epochs && st == AG_OK; ep++) {
        for (uint32_t i = S - 1; i > 0; i--) {  // Fisher-Yates with a 64-bit LCG
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            const uint32_t j = (uint32_t)((x >> 33) % (i + 1));
            const uint32_t t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
        hipError_t e = hipMemcpyAsync(m->perm, perm, (size_t)S * sizeof(uint32_t), hipMemcpyHostToDevice, s);
        double epoch_loss = 0.0;
        uint32_t nb = 0;
        for (uint32_t off = 0; off + batch <= S && e == hipSuccess; off += batch, nb++) {
            MlpArgs A = mlp_args(m, batch);
            A.perm = m->perm;
            A.off = off;
            A.inv_b = 2.0f / (float)batch;
            mlp_kernel<true><<<grid, MLP_WG, 0, s>>>(A);
            mlp_reduce_kernel<<<(P_COUNT + 255) / 256, 256, 0, s>>>(m->part, grid * MLP_WAVES, m->grad);
            m->adam_t++;
            const float bc1 = 1.0f - powf(0.9f, (float)m->adam_t);
            const float bc2 = 1.0f - powf(0.999f, (float)m->adam_t);
            mlp_adam_kernel<<<(P_COUNT + 255) / 256, 256, 0, s>>>(m->p, m->m1, m->m2, m->grad, lr, bc1, bc2);
            e = hipGetLastError();
            if (e == hipSuccess && loss_per_epoch)
                e = hipMemcpyAsync(lp, m->loss_part, (size_t)grid * MLP_WAVES * sizeof(double),
                                   hipMemcpyDeviceToHost, s);
            if (e == hipSuccess && loss_per_epoch) {
                st = ag_wait_stream(m->dev);
                if (st != AG_OK)
                    break;
                double t = 0.0;
                for (uint32_t i = 0; i < grid * MLP_WAVES; i++)
                    t += lp[i];
                epoch_loss += t / batch;
            }
        }
        if (e != hipSuccess)
            st = ag_hip_fail(e, "MLP training", __FILE__, __LINE__);
        if (st == AG_OK)
            st = ag_wait_stream(m->dev);
        if (loss_per_epoch)
            loss_per_epoch[ep] = nb ? (float)(epoch_loss / nb) : 0.0f;
    }
    free(perm);
    free(lp);
    return st;
}

static int mlp_predict_device(ag_mlp *m, const float *Xdev, uint32_t S, float *ydev)
{
    const uint32_t grid = mlp_grid(m, S);
    MlpArgs A = mlp_args(m, S);
    A.X = Xdev;
    A.out = ydev;
    A.y = NULL;
    mlp_kernel<false><<<grid, MLP_WG, 0, m->dev->stream>>>(A);
    AG_HIP(hipGetLastError());
    return AG_OK;
}

extern "C" int ag_mlp_predict(ag_mlp *m, const float *X, uint32_t S, float *y_out)
{
    if (!m || !X || !y_out || S == 0)
        return ag_set_error(AG_ERR_INVALID, "bad argument");
    AG_HIP(hipSetDevice(m->dev->ordinal));
    int st = mlp_upload(m, X, NULL, S);
    if (st == AG_OK)
        st = mlp_predict_device(m, m->X, S, m->y);
    if (st != AG_OK)
        return st;
    AG_HIP(hipMemcpyAsync(y_out, m->y, (size_t)S * sizeof(float), hipMemcpyDeviceToHost, m->dev->stream));
    return ag_wait_stream(m->dev);
}

// features of the session's graph into feat[n][8] (device order); r1/r2 = scaled ranks after
// one and two iterations
static int compute_features(ag_dgraph *g, const float *r1, const float *r2, float *feat)
{
    hipStream_t s = g->dev->stream;
    const uint32_t n = g->n;
    feat_local_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, g->off, g->outdeg, r1, r2, feat);
    PullGeom G = ag_make_geom(g);
    NbrDegPol pol{g->off, feat};
    const uint32_t grid = ag_pull_grid(g, AG_TILE_NNZ_MAX * 4u);
    pull_kernel<NbrDegPol><<<grid, AG_WG, 0, s>>>(G, pol);
    if (g->n_hubs)
        pull_finish_kernel<NbrDegPol><<<1, AG_FIN_WG, 0, s>>>(G, pol);
    AG_HIP(hipGetLastError());
    return AG_OK;
}

// scratch for the warm start: feat [n][8], r1, r2, pred (n floats plus room for the mass
// partials and alpha behind them)
static int warm_scratch(ag_dgraph *g, float **feat, float **r1, float **r2, float **pred)
{
    const size_t n = g->n ? g->n : 1;
    const size_t nblk = (n + AG_WG - 1) / AG_WG;
    AG_HIP(hipMalloc((void **)feat, n * MLP_FSTRIDE * sizeof(float)));
    AG_HIP(hipMalloc((void **)r1, n * sizeof(float)));
    AG_HIP(hipMalloc((void **)r2, n * sizeof(float)));
    AG_HIP(hipMalloc((void **)pred, n * sizeof(float) + (2 * nblk + 2) * sizeof(double) + 16));
    return AG_OK;
}

extern "C" int ag_mlp_features(ag_dgraph *g, float *X_out)
{
    if (!g || (g->n && !X_out))
        return ag_set_error(AG_ERR_INVALID, "null argument");
    if (g->n == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    ag_pr_options o;
    ag_pr_options_default(&o);
    o.max_iterations = 2;
    o.stop_rule = AG_STOP_NONE;
    o.init = AG_INIT_UNIFORM;
    ag_pr_session *s = NULL;
    float *feat = NULL, *r1 = NULL, *r2 = NULL, *pred = NULL;
    int st = warm_scratch(g, &feat, &r1, &r2, &pred);
    if (st == AG_OK)
        st = ag_pr_begin(g, &o, &s);
    hipStream_t hs = g->dev->stream;
    if (st == AG_OK)
        st = ag_pr_step(s, 1);
    if (st == AG_OK && hipMemcpyAsync(r1, s->rank, (size_t)g->n * sizeof(float), hipMemcpyDeviceToDevice, hs) != hipSuccess)
        st = ag_set_error(AG_ERR_HIP, "copy r1");
    if (st == AG_OK)
        st = ag_pr_step(s, 1);
    if (st == AG_OK && hipMemcpyAsync(r2, s->rank, (size_t)g->n * sizeof(float), hipMemcpyDeviceToDevice, hs) != hipSuccess)
        st = ag_set_error(AG_ERR_HIP, "copy r2");
    if (st == AG_OK)
        st = compute_features(g, r1, r2, feat);
    if (st == AG_OK) {
        // reuse pred as [n][5] staging would be too small; permute into vec_a
        st = ag_ensure_vec(g, (size_t)g->n * MLP_NF * sizeof(float));
        if (st == AG_OK) {
            feat_out_kernel<<<(g->n + 255) / 256, 256, 0, hs>>>(g->n, feat, g->d2o, (float *)g->vec_a);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess)
                e = hipMemcpyAsync(X_out, g->vec_a, (size_t)g->n * MLP_NF * sizeof(float), hipMemcpyDeviceToHost, hs);
            st = (e == hipSuccess) ? ag_wait_stream(g->dev) : ag_hip_fail(e, "features", __FILE__, __LINE__);
        }
    }
    if (s)
        ag_pr_end(s, NULL, NULL);
    (void)hipStreamSynchronize(hs);
    if (feat) (void)hipFree(feat);
    if (r1) (void)hipFree(r1);
    if (r2) (void)hipFree(r2);
    if (pred) (void)hipFree(pred);
    return st;
}

extern "C" int ag_pr_warm_start(ag_pr_session *s, ag_mlp *m, ag_warm_result *res)
{
    if (!s || !m || !res)
        return ag_set_error(AG_ERR_INVALID, "null argument");
    memset(res, 0, sizeof(*res));
    ag_dgraph *g = s->g;
    if (s->kind_code != 0 || s->o.init != AG_INIT_UNIFORM || s->enqueued != 0 ||
        s->o.direction != AG_DIR_PULL)
        return ag_set_error(AG_ERR_INVALID,
                            "warm start needs a fresh FP32 pull session with the uniform init");
    if (s->o.max_iterations < 4)
        return ag_set_error(AG_ERR_INVALID, "warm start needs max_iterations >= 4");
    if (g->n == 0)
        return AG_OK;
    AG_HIP(hipSetDevice(g->dev->ordinal));
    hipStream_t hs = g->dev->stream;
    const size_t nb = (size_t)g->n * sizeof(float);
    float *feat = NULL, *r1 = NULL, *r2 = NULL, *pred = NULL;
    int st = warm_scratch(g, &feat, &r1, &r2, &pred);
    hipEvent_t e0 = NULL, e1 = NULL;
    if (st == AG_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        st = ag_set_error(AG_ERR_HIP, "hipEventCreate");
    if (st == AG_OK && hipEventRecord(e0, hs) != hipSuccess)
        st = ag_set_error(AG_ERR_HIP, "hipEventRecord");
    if (st == AG_OK)
        st = ag_pr_step(s, 1);
    if (st == AG_OK && hipMemcpyAsync(r1, s->rank, nb, hipMemcpyDeviceToDevice, hs) != hipSuccess)
        st = ag_set_error(AG_ERR_HIP, "copy r1");
    if (st == AG_OK)
        st = ag_pr_step(s, 1);
    if (st == AG_OK && hipMemcpyAsync(r2, s->rank, nb, hipMemcpyDeviceToDevice, hs) != hipSuccess)
        st = ag_set_error(AG_ERR_HIP, "copy r2");
    if (st == AG_OK)
        st = compute_features(g, r1, r2, feat);
    if (st == AG_OK)
        st = mlp_predict_device(m, feat, g->n, pred);
    if (st == AG_OK) {
        // start vector, mass-balanced; iteration 2 reads contribution buffer 0
        const uint32_t nblk = (g->n + AG_WG - 1) / AG_WG;
        double *part = (double *)(((uintptr_t)(pred + g->n) + 15) & ~(uintptr_t)15);
        double *alpha = part + 2 * nblk;
        warm_apply_kernel<<<nblk, AG_WG, 0, hs>>>(g->n, pred, feat, g->outdeg, s->rank, part);
        warm_mass_kernel<<<1, AG_WG, 0, hs>>>(part, nblk, (double)g->n, alpha);
        warm_scale_kernel<<<nblk, AG_WG, 0, hs>>>(g->n, alpha, g->outdeg, s->rank, (float *)s->c[0]);
        if (hipGetLastError() != hipSuccess)
            st = ag_set_error(AG_ERR_HIP, "warm apply launch");
    }
    if (st == AG_OK)
        st = ag_pr_step(s, 1);
    double h[3] = {0, 0, 0};
    if (st == AG_OK && hipMemcpyAsync(h, s->hist_err, 3 * sizeof(double), hipMemcpyDeviceToHost, hs) != hipSuccess)
        st = ag_set_error(AG_ERR_HIP, "history copy");
    if (st == AG_OK)
        st = ag_wait_stream(g->dev);
    if (st == AG_OK) {
        res->uniform_l1 = h[1];
        res->warm_l1 = h[2];
        if (!(h[2] < h[1])) {
            // keep the uniform run: continue from r2; iteration 3 reads buffer 1
            res->fallback = 1;
            restore_kernel<<<(g->n + 255) / 256, 256, 0, hs>>>(g->n, r2, g->outdeg, s->rank,
                                                                (float *)s->c[1]);
            if (hipGetLastError() != hipSuccess)
                st = ag_set_error(AG_ERR_HIP, "restore launch");
        }
    }
    if (st == AG_OK && hipEventRecord(e1, hs) == hipSuccess && ag_wait_stream(g->dev) == AG_OK) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            ms = 0.0f;   // timing is informational
        res->ms = ms;
    }
    (void)hipStreamSynchronize(hs);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (feat) (void)hipFree(feat);
    if (r1) (void)hipFree(r1);
    if (r2) (void)hipFree(r2);
    if (pred) (void)hipFree(pred);
    return st;
}
