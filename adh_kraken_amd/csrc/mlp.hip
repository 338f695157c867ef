/*
 * mlp.hip -- learned warm start of the Newton iterate (gfx950).
 *
 * A per-node MLP, 12 -> 64 -> 64 -> 3 with GELU (erf form), BF16 weights and
 * FP32 arithmetic, predicts the change of (h, u, v) over the next time step
 * from the node's h, hu, hv at t and t-1, the 1-ring neighbour means of h, hu,
 * hv at t, the bed slope and Manning's n (see node_features for the encoding).  It is trained on trajectories that the solver itself
 * produces, on the GPU, by the kernels below:
 *   k_node_static   per-node bed slope and Manning n (area-weighted over the node's elements)
 *   k_collect       features(t, t-1) and target sol(t+1) - sol(t) of every node -> dataset (SoA)
 *   k_stats         per-workgroup sums of each column and its square (normalisation)
 *   k_tile_grad     one wavefront per 64-sample tile: forward, MSE loss, backward; activations
 *                   staged in LDS; the tile's weight gradient is written as one partial row
 *   k_reduce_adam   sums the tile partials in tile order (deterministic) and applies Adam to the
 *                   FP32 master weights, refreshing their BF16 copy
 *   k_predict       inference straight into the Newton initial iterate
 *
 * No counterpart in the reference: there the Newton solve starts from the previous
 * time level (smodel_super_prep_sol.c:2-8, initialize_system newton_tools.c:21).
 * Plain Newton stays the fallback and the arbiter: fe_newton keeps whichever initial
 * iterate, learned or previous level, has the smaller residual (l2 norm by default,
 * max norm with adhk_solver.warm_gate = 0), and runs the unchanged Newton iteration
 * and stopping test from it.
 *
 * Parameter vector (P = 5187 floats): W1[64][12] b1[64] W2[64][64] b2[64] W3[3][64] b3[3].
 * Normalisation (36 floats): x mean [12], x 1/std [12], y mean [3], y std [3], y 1/std [3].
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "adhk_internal.h"

#define NIN ADHK_MLP_NIN
#define HID ADHK_MLP_HID
#define NOUT ADHK_MLP_NOUT
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + HID * NIN)
#define OFF_W2 (OFF_B1 + HID)
#define OFF_B2 (OFF_W2 + HID * HID)
#define OFF_W3 (OFF_B2 + HID)
#define OFF_B3 (OFF_W3 + NOUT * HID)
#define NPARAM (OFF_B3 + NOUT)
#define NNORM 36
#define TILE 64     /* samples per training tile = one wavefront */
#define APAD 68     /* LDS row stride of a 64-wide activation row (16-byte aligned, spreads banks) */
#define MAX_TILES 1024
#define NCOL (NIN + NOUT)
static_assert(TILE == HID, "the bias-gradient sums use one thread per hidden unit");

struct adhk_mlp {
    hipStream_t stream;
    float *master, *m1, *m2, *norm, *partials, *lossp, *loss_hist;
    unsigned short *wbf;
    float *feat, *targ; /* dataset, SoA [NIN][cap] and [NOUT][cap] */
    long long cap, n;
    double *stat_part;
    long long adam_t;
    float h_norm[NNORM];
};

#define HC(x)                                                                                                     \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));                         \
            goto fail;                                                                                            \
        }                                                                                                         \
    } while (0)

/* ---------------- numerics shared by all kernels ---------------- */
__host__ __device__ static inline unsigned short f2bf(float f)
{
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u); /* quiet NaN */
    u += 0x7fffu + ((u >> 16) & 1u);                                                 /* round to nearest even */
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b)
{
    unsigned u = ((unsigned)b) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

static inline unsigned long long splitmix64(unsigned long long *s)
{
    unsigned long long z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* ---------------- features ---------------- */
/* static node features: area-weighted bed slope (x, y) and Manning n over the node's triangles.  The
 * triangles containing node n are the 2-D contributions of its diagonal block, in element order. */
__global__ void k_node_static(int N, int E, const int *__restrict__ diag_blk, const int *__restrict__ cptr,
                              const int *__restrict__ contrib, const int *__restrict__ tri,
                              const double *__restrict__ z, const double *__restrict__ grad,
                              const double *__restrict__ djac, const double *__restrict__ mann, float *nstat)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double sx = 0.0, sy = 0.0, sn = 0.0, sw = 0.0;
    int b = diag_blk[n];
    int j0 = b >= 0 ? cptr[b] : 0, j1 = b >= 0 ? cptr[b + 1] : 0;
    for (int j = j0; j < j1; j++) {
        int code = contrib[j];
        if (code >= 9 * E) continue;
        int e = code / 9;
        double gx = 0.0, gy = 0.0;
        for (int a = 0; a < 3; a++) {
            double za = z[tri[a * E + e]];
            gx += za * grad[a * E + e];
            gy += za * grad[(3 + a) * E + e];
        }
        double w = djac[e];
        sx += w * gx;
        sy += w * gy;
        sn += w * mann[e];
        sw += w;
    }
    nstat[n] = sw > 0.0 ? (float)(sx / sw) : 0.0f;
    nstat[N + n] = sw > 0.0 ? (float)(sy / sw) : 0.0f;
    nstat[2 * N + n] = sw > 0.0 ? (float)(sn / sw) : 0.0f;
}

/* f = [h, hu, hv](t); the change of [h, hu, hv] since t-1; the 1-ring means of [h, hu, hv](t) minus the
 * node's own values; slope x, slope y, n.  The state at t-1 and the neighbour means enter as differences,
 * formed in double: the same information as the raw values, but a BF16-weight layer could not resolve
 * a change of 1e-3 relative between two nearly equal inputs. */
__device__ __forceinline__ void node_features(int n, int N, const double *__restrict__ st,
                                              const double *__restrict__ stm1, const int *__restrict__ brow,
                                              const int *__restrict__ bcol, const float *__restrict__ nstat,
                                              float f[NIN])
{
    const double h = st[3 * n], hu = h * st[3 * n + 1], hv = h * st[3 * n + 2];
    const double h1 = stm1[3 * n], hu1 = h1 * stm1[3 * n + 1], hv1 = h1 * stm1[3 * n + 2];
    f[0] = (float)h;
    f[1] = (float)hu;
    f[2] = (float)hv;
    f[3] = (float)(h - h1);
    f[4] = (float)(hu - hu1);
    f[5] = (float)(hv - hv1);
    double sh = 0.0, su = 0.0, sv = 0.0;
    int cnt = 0;
    for (int j = brow[n]; j < brow[n + 1]; j++) {
        int m = bcol[j];
        if (m == n) continue;
        double hm = st[3 * m];
        sh += hm;
        su += hm * st[3 * m + 1];
        sv += hm * st[3 * m + 2];
        cnt++;
    }
    f[6] = cnt ? (float)(sh / cnt - h) : 0.0f;
    f[7] = cnt ? (float)(su / cnt - hu) : 0.0f;
    f[8] = cnt ? (float)(sv / cnt - hv) : 0.0f;
    f[9] = nstat[n];
    f[10] = nstat[N + n];
    f[11] = nstat[2 * N + n];
}

__global__ void k_features(int N, const double *st, const double *stm1, const int *brow, const int *bcol,
                           const float *nstat, float *out /* [N][NIN] */)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float f[NIN];
    node_features(n, N, st, stm1, brow, bcol, nstat, f);
    for (int k = 0; k < NIN; k++) out[(size_t)n * NIN + k] = f[k];
}

__global__ void k_collect(int N, long long off, long long cap, const double *sol, const double *st,
                          const double *stm1, const int *brow, const int *bcol, const float *nstat, float *feat,
                          float *targ)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float f[NIN];
    node_features(n, N, st, stm1, brow, bcol, nstat, f);
    long long i = off + n;
    for (int k = 0; k < NIN; k++) feat[k * cap + i] = f[k];
    for (int k = 0; k < NOUT; k++) targ[k * cap + i] = (float)(sol[3 * n + k] - st[3 * n + k]);
}

/* per-workgroup sums of every column and its square, in double */
__global__ __launch_bounds__(256) void k_stats(long long n, long long cap, const float *feat, const float *targ,
                                               double *part /* [G][2*NCOL] */)
{
    __shared__ double sh[256];
    for (int c = 0; c < NCOL; c++) {
        const float *col = c < NIN ? feat + c * cap : targ + (c - NIN) * cap;
        double s = 0.0, s2 = 0.0;
        for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
            double x = col[i];
            s += x;
            s2 += x * x;
        }
        for (int pass = 0; pass < 2; pass++) {
            sh[threadIdx.x] = pass ? s2 : s;
            __syncthreads();
            for (int o = 128; o > 0; o >>= 1) {
                if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
                __syncthreads();
            }
            if (threadIdx.x == 0) part[(size_t)blockIdx.x * 2 * NCOL + 2 * c + pass] = sh[0];
            __syncthreads();
        }
    }
}

/* ---------------- training ---------------- */
/* Sample s of a batch: sequential (i = base + s) for gradient checks, else a hashed index into the
 * dataset (with replacement), fixed by (seed, step, s). */
struct Batch {
    const float *feat, *targ;
    long long cap, n, base;
    unsigned long long seed, step;
    int size, sequential;
};

__device__ __forceinline__ long long batch_index(const Batch &B, int s)
{
    if (B.sequential) return B.base + s;
    unsigned long long h = mix64(B.seed ^ mix64(B.step * 0x100000001B3ULL + (unsigned long long)s));
    return (long long)(h % (unsigned long long)B.n);
}

__device__ __forceinline__ void load_weights(const unsigned short *__restrict__ wbf, float *W)
{
    for (int p = threadIdx.x; p < NPARAM; p += blockDim.x) W[p] = bf2f(wbf[p]);
}

/* one wavefront = one tile of TILE samples; partial[tile][NPARAM] = d(loss)/d(params) over the tile,
 * loss = 1/(size*NOUT) * sum_s sum_k (y_sk - t_sk)^2 on normalised outputs.
 * Registers hold one sample's a1 and one accumulator vector; rows of W come from LDS in rolled loops,
 * so every register array is indexed by constants.  z2 is recomputed in the backward pass, which also
 * accumulates W2^T dz2 from the same W2 row. */
__global__ __launch_bounds__(TILE) void k_tile_grad(Batch B, const unsigned short *__restrict__ wbf,
                                                    const float *__restrict__ norm, float *__restrict__ partial,
                                                    float *__restrict__ lossp)
{
    __shared__ float W[NPARAM];
    __shared__ float A[TILE * APAD];  /* a1 rows, later dz1 rows */
    __shared__ float Bs[TILE * APAD]; /* a2 rows, later dz2 rows */
    __shared__ float X[TILE * NIN];
    __shared__ float DY[TILE * 4];
    __shared__ float red[TILE];
    const int t = threadIdx.x, tile = blockIdx.x;
    const int s = tile * TILE + t;
    const bool valid = s < B.size;
    load_weights(wbf, W);
    float x[NIN], tn[NOUT];
    if (valid) {
        long long i = batch_index(B, s);
#pragma unroll
        for (int k = 0; k < NIN; k++) x[k] = (B.feat[k * B.cap + i] - norm[k]) * norm[NIN + k];
#pragma unroll
        for (int k = 0; k < NOUT; k++) tn[k] = (B.targ[k * B.cap + i] - norm[24 + k]) * norm[30 + k];
    } else {
#pragma unroll
        for (int k = 0; k < NIN; k++) x[k] = 0.0f;
#pragma unroll
        for (int k = 0; k < NOUT; k++) tn[k] = 0.0f;
    }
#pragma unroll
    for (int k = 0; k < NIN; k++) X[t * NIN + k] = x[k];
    __syncthreads();
    /* forward: layer 1 (unrolled, a1 in registers) */
    float a1[HID];
#pragma unroll
    for (int i = 0; i < HID; i++) {
        float zz = W[OFF_B1 + i];
#pragma unroll
        for (int j = 0; j < NIN; j++) zz += W[OFF_W1 + i * NIN + j] * x[j];
        a1[i] = gelu(zz);
        A[t * APAD + i] = a1[i];
    }
    /* layer 2 and output, one W2 row at a time */
    float y[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; k++) y[k] = W[OFF_B3 + k];
#pragma unroll 1
    for (int i = 0; i < HID; i++) {
        float zz = W[OFF_B2 + i];
#pragma unroll
        for (int j = 0; j < HID; j++) zz += W[OFF_W2 + i * HID + j] * a1[j];
        float a2 = gelu(zz);
        Bs[t * APAD + i] = a2;
#pragma unroll
        for (int k = 0; k < NOUT; k++) y[k] += W[OFF_W3 + k * HID + i] * a2;
    }
    float l = 0.0f, dy[NOUT];
    const float sc = 2.0f / ((float)B.size * (float)NOUT);
#pragma unroll
    for (int k = 0; k < NOUT; k++) {
        float d = y[k] - tn[k];
        l += d * d;
        dy[k] = valid ? sc * d : 0.0f;
        DY[t * 4 + k] = dy[k];
    }
    red[t] = valid ? l : 0.0f;
    __syncthreads();
    float *P = partial + (size_t)tile * NPARAM;
    /* dW3[k][i] = sum_s dy[s][k] a2[s][i], db3, tile loss */
    for (int e = t; e < NOUT * HID; e += TILE) {
        int k = e / HID, i = e % HID;
        float acc = 0.0f;
        for (int q = 0; q < TILE; q++) acc += DY[q * 4 + k] * Bs[q * APAD + i];
        P[OFF_W3 + e] = acc;
    }
    if (t < NOUT) {
        float acc = 0.0f;
        for (int q = 0; q < TILE; q++) acc += DY[q * 4 + t];
        P[OFF_B3 + t] = acc;
    }
    if (t == 0) {
        float acc = 0.0f;
        for (int q = 0; q < TILE; q++) acc += red[q];
        lossp[tile] = acc;
    }
    __syncthreads(); /* every read of the a2 rows is done */
    /* dz2_i = (W3^T dy)_i gelu'(z2_i), z2 recomputed; da1 += W2[i][:] dz2_i */
    float da1[HID];
#pragma unroll
    for (int j = 0; j < HID; j++) da1[j] = 0.0f;
#pragma unroll 1
    for (int i = 0; i < HID; i++) {
        float zz = W[OFF_B2 + i];
#pragma unroll
        for (int j = 0; j < HID; j++) zz += W[OFF_W2 + i * HID + j] * a1[j];
        float g = 0.0f;
#pragma unroll
        for (int k = 0; k < NOUT; k++) g += W[OFF_W3 + k * HID + i] * dy[k];
        float dz = g * gelu_d(zz);
        Bs[t * APAD + i] = dz;
#pragma unroll
        for (int j = 0; j < HID; j++) da1[j] += W[OFF_W2 + i * HID + j] * dz;
    }
    /* dz1 = da1 * gelu'(z1), z1 recomputed (held in da1) */
#pragma unroll
    for (int j = 0; j < HID; j++) {
        float zz = W[OFF_B1 + j];
#pragma unroll
        for (int k = 0; k < NIN; k++) zz += W[OFF_W1 + j * NIN + k] * x[k];
        da1[j] *= gelu_d(zz);
    }
    __syncthreads();
    /* dW2[i][j] = sum_s dz2[s][i] a1[s][j]: thread t owns rows 8*(t%8).., columns 8*(t/8).. */
    {
        const int i0 = 8 * (t % 8), j0 = 8 * (t / 8);
        float acc[8][8];
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int b = 0; b < 8; b++) acc[a][b] = 0.0f;
#pragma unroll 2
        for (int q = 0; q < TILE; q++) {
            float dz[8], av[8];
#pragma unroll
            for (int a = 0; a < 8; a++) dz[a] = Bs[q * APAD + i0 + a];
#pragma unroll
            for (int b = 0; b < 8; b++) av[b] = A[q * APAD + j0 + b];
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int b = 0; b < 8; b++) acc[a][b] += dz[a] * av[b];
        }
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int b = 0; b < 8; b++) P[OFF_W2 + (i0 + a) * HID + j0 + b] = acc[a][b];
    }
    {
        float acc = 0.0f;
        for (int q = 0; q < TILE; q++) acc += Bs[q * APAD + t];
        P[OFF_B2 + t] = acc;
    }
    __syncthreads(); /* every read of the a1 rows is done */
#pragma unroll
    for (int j = 0; j < HID; j++) A[t * APAD + j] = da1[j];
    __syncthreads();
    /* dW1[i][k] = sum_s dz1[s][i] x[s][k]: 768 entries, 12 per thread; db1 */
    for (int e = t; e < HID * NIN; e += TILE) {
        int i = e / NIN, k = e % NIN;
        float acc = 0.0f;
        for (int q = 0; q < TILE; q++) acc += A[q * APAD + i] * X[q * NIN + k];
        P[OFF_W1 + e] = acc;
    }
    {
        float acc = 0.0f;
        for (int q = 0; q < TILE; q++) acc += A[q * APAD + t];
        P[OFF_B1 + t] = acc;
    }
}

/* grad[p] = sum over tiles in order; adam != 0: Adam step on master, BF16 copy refreshed */
__global__ __launch_bounds__(256) void k_reduce_adam(int ntiles, const float *__restrict__ partial,
                                                     const float *__restrict__ lossp, float *grad_out,
                                                     double *loss_out, int adam, float lr, float b1c, float b2c,
                                                     float *master, float *m1, float *m2, unsigned short *wbf)
{
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == NPARAM && loss_out) {
        double l = 0.0;
        for (int q = 0; q < ntiles; q++) l += lossp[q];
        *loss_out = l;
    }
    if (p >= NPARAM) return;
    float g = 0.0f;
    for (int q = 0; q < ntiles; q++) g += partial[(size_t)q * NPARAM + p];
    if (grad_out) grad_out[p] = g;
    if (!adam) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float m = beta1 * m1[p] + (1.0f - beta1) * g;
    float v = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p] = m;
    m2[p] = v;
    float w = master[p] - lr * (m / b1c) / (sqrtf(v / b2c) + eps);
    master[p] = w;
    wbf[p] = f2bf(w);
}

/* ---------------- inference ---------------- */
__device__ __forceinline__ void mlp_forward(const float *W, const float x[NIN], float y[NOUT])
{
    float a1[HID];
#pragma unroll
    for (int i = 0; i < HID; i++) {
        float zz = W[OFF_B1 + i];
#pragma unroll
        for (int j = 0; j < NIN; j++) zz += W[OFF_W1 + i * NIN + j] * x[j];
        a1[i] = gelu(zz);
    }
#pragma unroll
    for (int k = 0; k < NOUT; k++) y[k] = W[OFF_B3 + k];
#pragma unroll 1
    for (int i = 0; i < HID; i++) {
        float zz = W[OFF_B2 + i];
#pragma unroll
        for (int j = 0; j < HID; j++) zz += W[OFF_W2 + i * HID + j] * a1[j];
        float a2 = gelu(zz);
#pragma unroll
        for (int k = 0; k < NOUT; k++) y[k] += W[OFF_W3 + k * HID + i] * a2;
    }
}

/* raw features in, raw outputs (denormalised) out */
__global__ __launch_bounds__(256) void k_forward(int n, const float *__restrict__ xin,
                                                 const unsigned short *__restrict__ wbf,
                                                 const float *__restrict__ norm, float *__restrict__ yout)
{
    __shared__ float W[NPARAM];
    load_weights(wbf, W);
    __syncthreads();
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[NIN], y[NOUT];
    #pragma unroll

    for (int k = 0; k < NIN; k++) x[k] = (xin[(size_t)s * NIN + k] - norm[k]) * norm[NIN + k];
    mlp_forward(W, x, y);
    #pragma unroll

    for (int k = 0; k < NOUT; k++) yout[(size_t)s * NOUT + k] = y[k] * norm[27 + k] + norm[24 + k];
}

/* sol = sol_old + MLP(features(sol_old, sol_older)) */
__global__ __launch_bounds__(256) void k_predict(int N, double *__restrict__ sol, const double *__restrict__ st,
                                                 const double *__restrict__ stm1, const int *__restrict__ brow,
                                                 const int *__restrict__ bcol, const float *__restrict__ nstat,
                                                 const unsigned short *__restrict__ wbf,
                                                 const float *__restrict__ norm, const int *__restrict__ node_inst,
                                                 const int *__restrict__ work)
{
    __shared__ float W[NPARAM];
    load_weights(wbf, W);
    __syncthreads();
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N || !work[node_inst[n]]) return;
    float f[NIN], y[NOUT];
    node_features(n, N, st, stm1, brow, bcol, nstat, f);
    #pragma unroll

    for (int k = 0; k < NIN; k++) f[k] = (f[k] - norm[k]) * norm[NIN + k];
    mlp_forward(W, f, y);
    #pragma unroll

    for (int k = 0; k < NOUT; k++) sol[3 * n + k] = st[3 * n + k] + (double)(y[k] * norm[27 + k] + norm[24 + k]);
}

/* ---------------- host side ---------------- */
static int ensure_static(adhk_ctx *c)
{
    if (c->nstat) return 0;
    HC(hipMalloc((void **)&c->nstat, 3 * (size_t)(c->N > 0 ? c->N : 1) * sizeof(float)));
    if (c->N > 0) {
        hipLaunchKernelGGL(k_node_static, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->E, c->diag_blk,
                           c->cptr, c->contrib, c->tri, c->z, c->grad, c->djac, c->mann, c->nstat);
        HC(hipGetLastError());
    }
    return 0;
fail:
    return -1;
}

extern "C" int adhk_mlp_param_count(void) { return NPARAM; }

extern "C" void adhk_mlp_destroy(adhk_mlp *m)
{
    if (!m) return;
    void *ptrs[] = {m->master, m->m1, m->m2, m->norm, m->partials, m->lossp, m->loss_hist, m->wbf, m->feat, m->targ,
                    m->stat_part};
    for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++)
        if (ptrs[i]) hipFree(ptrs[i]);
    if (m->stream) hipStreamDestroy(m->stream);
    free(m);
}

static void norm_identity(float *nm)
{
    for (int k = 0; k < NNORM; k++) nm[k] = 0.0f;
    for (int k = 0; k < NIN; k++) nm[NIN + k] = 1.0f;
    for (int k = 0; k < NOUT; k++) nm[27 + k] = nm[30 + k] = 1.0f;
}

static int upload_params(adhk_mlp *m, const float *w)
{
    unsigned short *bf = (unsigned short *)malloc(NPARAM * sizeof(unsigned short));
    if (!bf) return -1;
    for (int p = 0; p < NPARAM; p++) bf[p] = f2bf(w[p]);
    HC(hipMemcpyAsync(m->master, w, NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HC(hipMemcpyAsync(m->wbf, bf, NPARAM * sizeof(unsigned short), hipMemcpyHostToDevice, m->stream));
    HC(hipMemsetAsync(m->m1, 0, NPARAM * sizeof(float), m->stream));
    HC(hipMemsetAsync(m->m2, 0, NPARAM * sizeof(float), m->stream));
    HC(hipStreamSynchronize(m->stream));
    m->adam_t = 0;
    free(bf);
    return 0;
fail:
    free(bf);
    return -1;
}

/* Glorot-uniform weights from SplitMix64 (u = (x >> 11) 2^-53), zero biases; the output layer is
 * scaled by 0.1 so that the untrained prediction is close to the mean change */
extern "C" adhk_mlp *adhk_mlp_create(unsigned long long seed)
{
    adhk_mlp *m = (adhk_mlp *)calloc(1, sizeof(adhk_mlp));
    float *w = NULL;
    if (!m) return NULL;
    HC(hipStreamCreate(&m->stream));
    HC(hipMalloc((void **)&m->master, NPARAM * sizeof(float)));
    HC(hipMalloc((void **)&m->m1, NPARAM * sizeof(float)));
    HC(hipMalloc((void **)&m->m2, NPARAM * sizeof(float)));
    HC(hipMalloc((void **)&m->wbf, NPARAM * sizeof(unsigned short)));
    HC(hipMalloc((void **)&m->norm, NNORM * sizeof(float)));
    HC(hipMalloc((void **)&m->partials, (size_t)MAX_TILES * NPARAM * sizeof(float)));
    HC(hipMalloc((void **)&m->lossp, MAX_TILES * sizeof(float)));
    HC(hipMalloc((void **)&m->loss_hist, sizeof(double)));
    w = (float *)calloc(NPARAM, sizeof(float));
    if (!w) goto fail;
    {
        unsigned long long st = seed;
        const int fan[3][3] = {{OFF_W1, NIN, HID}, {OFF_W2, HID, HID}, {OFF_W3, HID, NOUT}};
        for (int l = 0; l < 3; l++) {
            double lim = sqrt(6.0 / (fan[l][1] + fan[l][2])) * (l == 2 ? 0.1 : 1.0);
            for (int p = 0; p < fan[l][1] * fan[l][2]; p++) {
                double u = (double)(splitmix64(&st) >> 11) * (1.0 / 9007199254740992.0);
                w[fan[l][0] + p] = (float)((2.0 * u - 1.0) * lim);
            }
        }
    }
    norm_identity(m->h_norm);
    HC(hipMemcpy(m->norm, m->h_norm, NNORM * sizeof(float), hipMemcpyHostToDevice));
    if (upload_params(m, w)) goto fail;
    free(w);
    return m;
fail:
    free(w);
    adhk_mlp_destroy(m);
    return NULL;
}

extern "C" int adhk_mlp_get_params(adhk_mlp *m, float *params, float *norm)
{
    if (!m) return -1;
    if (params) HC(hipMemcpy(params, m->master, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    if (norm) memcpy(norm, m->h_norm, sizeof m->h_norm);
    return 0;
fail:
    return -1;
}

extern "C" int adhk_mlp_set_params(adhk_mlp *m, const float *params, const float *norm)
{
    if (!m) return -1;
    if (norm) {
        memcpy(m->h_norm, norm, sizeof m->h_norm);
        HC(hipMemcpy(m->norm, m->h_norm, NNORM * sizeof(float), hipMemcpyHostToDevice));
    }
    if (params && upload_params(m, params)) return -1;
    return 0;
fail:
    return -1;
}

extern "C" int adhk_mlp_data_reserve(adhk_mlp *m, long long cap)
{
    if (!m || cap <= 0) {
        adhk_set_error("adhk_mlp_data_reserve: bad arguments");
        return -1;
    }
    if (m->feat) hipFree(m->feat);
    if (m->targ) hipFree(m->targ);
    m->feat = m->targ = NULL;
    HC(hipMalloc((void **)&m->feat, (size_t)cap * NIN * sizeof(float)));
    HC(hipMalloc((void **)&m->targ, (size_t)cap * NOUT * sizeof(float)));
    m->cap = cap;
    m->n = 0;
    return 0;
fail:
    return -1;
}

extern "C" long long adhk_mlp_data_size(const adhk_mlp *m) { return m ? m->n : -1; }

extern "C" int adhk_mlp_collect(adhk_mlp *m, adhk_sw2 *c)
{
    if (!m || !c) return -1;
    if (m->n + c->N > m->cap) {
        adhk_set_error("adhk_mlp_collect: dataset full (%lld + %d > %lld)", m->n, c->N, m->cap);
        return -1;
    }
    if (ensure_static(c)) return -1;
    hipLaunchKernelGGL(k_collect, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, m->n, m->cap, c->sol,
                       c->sol_old, c->sol_older, c->brow, c->bcol, c->nstat, m->feat, m->targ);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(c->stream));
    m->n += c->N;
    return 0;
fail:
    return -1;
}

extern "C" int adhk_sw2_features(adhk_sw2 *c, float *feat)
{
    float *d = NULL;
    if (!c || !feat) return -1;
    if (ensure_static(c)) return -1;
    HC(hipMalloc((void **)&d, (size_t)(c->N > 0 ? c->N : 1) * NIN * sizeof(float)));
    hipLaunchKernelGGL(k_features, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->sol_old,
                       c->sol_older, c->brow, c->bcol, c->nstat, d);
    HC(hipGetLastError());
    HC(hipMemcpyAsync(feat, d, (size_t)c->N * NIN * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HC(hipStreamSynchronize(c->stream));
    hipFree(d);
    return 0;
fail:
    if (d) hipFree(d);
    return -1;
}

extern "C" int adhk_mlp_fit_normalization(adhk_mlp *m)
{
    const int G = 512;
    double *h = NULL;
    if (!m || m->n <= 0) {
        adhk_set_error("adhk_mlp_fit_normalization: empty dataset");
        return -1;
    }
    if (!m->stat_part) HC(hipMalloc((void **)&m->stat_part, (size_t)G * 2 * NCOL * sizeof(double)));
    hipLaunchKernelGGL(k_stats, dim3(G), dim3(256), 0, m->stream, m->n, m->cap, m->feat, m->targ, m->stat_part);
    HC(hipGetLastError());
    h = (double *)malloc((size_t)G * 2 * NCOL * sizeof(double));
    if (!h) return -1;
    HC(hipMemcpyAsync(h, m->stat_part, (size_t)G * 2 * NCOL * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HC(hipStreamSynchronize(m->stream));
    for (int c = 0; c < NCOL; c++) {
        double s = 0.0, s2 = 0.0;
        for (int g = 0; g < G; g++) {
            s += h[(size_t)g * 2 * NCOL + 2 * c];
            s2 += h[(size_t)g * 2 * NCOL + 2 * c + 1];
        }
        double mean = s / (double)m->n, var = s2 / (double)m->n - mean * mean;
        double sd = var > 0.0 ? sqrt(var) : 0.0;
        int constant = sd <= 1e-7 * (fabs(mean) > 1.0 ? fabs(mean) : 1.0);
        if (c < NIN) {
            m->h_norm[c] = (float)mean;
            m->h_norm[NIN + c] = constant ? 0.0f : (float)(1.0 / sd);
        } else {
            int k = c - NIN;
            m->h_norm[24 + k] = (float)mean;
            m->h_norm[27 + k] = constant ? 1.0f : (float)sd;
            m->h_norm[30 + k] = constant ? 1.0f : (float)(1.0 / sd);
        }
    }
    HC(hipMemcpy(m->norm, m->h_norm, NNORM * sizeof(float), hipMemcpyHostToDevice));
    free(h);
    return 0;
fail:
    free(h);
    return -1;
}

static int tile_grad(adhk_mlp *m, const Batch &B)
{
    int ntiles = (B.size + TILE - 1) / TILE;
    hipLaunchKernelGGL(k_tile_grad, dim3(ntiles), dim3(TILE), 0, m->stream, B, m->wbf, m->norm, m->partials, m->lossp);
    HC(hipGetLastError());
    return ntiles;
fail:
    return -1;
}

/* steps of Adam on minibatches of `batch` samples drawn from the dataset; loss_hist[steps] (or NULL)
 * receives each step's mean squared error on normalised outputs */
extern "C" int adhk_mlp_train(adhk_mlp *m, int steps, int batch, double lr, unsigned long long seed,
                              double *loss_hist)
{
    double *dl = NULL;
    if (!m || steps < 0 || batch <= 0 || batch > MAX_TILES * TILE || m->n <= 0) {
        adhk_set_error("adhk_mlp_train: bad arguments (batch <= %d, dataset not empty)", MAX_TILES * TILE);
        return -1;
    }
    if (loss_hist) HC(hipMalloc((void **)&dl, (size_t)(steps > 0 ? steps : 1) * sizeof(double)));
    for (int it = 0; it < steps; it++) {
        Batch B;
        B.feat = m->feat;
        B.targ = m->targ;
        B.cap = m->cap;
        B.n = m->n;
        B.base = 0;
        B.seed = seed;
        B.step = (unsigned long long)it;
        B.size = batch;
        B.sequential = 0;
        int nt = tile_grad(m, B);
        if (nt < 0) goto fail;
        m->adam_t++;
        float b1c = (float)(1.0 - pow(0.9, (double)m->adam_t)), b2c = (float)(1.0 - pow(0.999, (double)m->adam_t));
        hipLaunchKernelGGL(k_reduce_adam, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, m->stream, nt, m->partials,
                           m->lossp, (float *)NULL, dl ? dl + it : (double *)NULL, 1, (float)lr, b1c, b2c,
                           m->master, m->m1, m->m2, m->wbf);
        HC(hipGetLastError());
    }
    if (dl) {
        HC(hipMemcpyAsync(loss_hist, dl, (size_t)steps * sizeof(double), hipMemcpyDeviceToHost, m->stream));
        HC(hipStreamSynchronize(m->stream));
        for (int it = 0; it < steps; it++) loss_hist[it] /= (double)batch * NOUT;
        hipFree(dl);
        dl = NULL;
    }
    HC(hipStreamSynchronize(m->stream));
    return 0;
fail:
    if (dl) hipFree(dl);
    return -1;
}

/* gradient check surface: n samples (raw features x[n][12], raw targets t[n][3]) in order ->
 * grad[P] of the mean squared error on normalised outputs and the loss itself */
extern "C" int adhk_mlp_loss_grad(adhk_mlp *m, int n, const float *x, const float *t, float *grad, double *loss)
{
    float *df = NULL, *dt_ = NULL, *dg = NULL, *hf = NULL, *ht = NULL;
    double *dl = NULL;
    if (!m || n <= 0 || n > MAX_TILES * TILE || !x || !t) {
        adhk_set_error("adhk_mlp_loss_grad: bad arguments (1 <= n <= %d)", MAX_TILES * TILE);
        return -1;
    }
    hf = (float *)malloc((size_t)n * NIN * sizeof(float));
    ht = (float *)malloc((size_t)n * NOUT * sizeof(float));
    if (!hf || !ht) goto fail;
    for (int s = 0; s < n; s++) {
        for (int k = 0; k < NIN; k++) hf[(size_t)k * n + s] = x[(size_t)s * NIN + k];
        for (int k = 0; k < NOUT; k++) ht[(size_t)k * n + s] = t[(size_t)s * NOUT + k];
    }
    HC(hipMalloc((void **)&df, (size_t)n * NIN * sizeof(float)));
    HC(hipMalloc((void **)&dt_, (size_t)n * NOUT * sizeof(float)));
    HC(hipMalloc((void **)&dg, NPARAM * sizeof(float)));
    HC(hipMalloc((void **)&dl, sizeof(double)));
    HC(hipMemcpyAsync(df, hf, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
    HC(hipMemcpyAsync(dt_, ht, (size_t)n * NOUT * sizeof(float), hipMemcpyHostToDevice, m->stream));
    {
        Batch B;
        B.feat = df;
        B.targ = dt_;
        B.cap = n;
        B.n = n;
        B.base = 0;
        B.seed = 0;
        B.step = 0;
        B.size = n;
        B.sequential = 1;
        int nt = tile_grad(m, B);
        if (nt < 0) goto fail;
        hipLaunchKernelGGL(k_reduce_adam, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, m->stream, nt, m->partials,
                           m->lossp, dg, dl, 0, 0.0f, 1.0f, 1.0f, (float *)NULL, (float *)NULL, (float *)NULL,
                           (unsigned short *)NULL);
        HC(hipGetLastError());
    }
    if (grad) HC(hipMemcpyAsync(grad, dg, NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    {
        double l = 0.0;
        HC(hipMemcpyAsync(&l, dl, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        HC(hipStreamSynchronize(m->stream));
        if (loss) *loss = l / ((double)n * NOUT);
    }
    hipFree(df);
    hipFree(dt_);
    hipFree(dg);
    hipFree(dl);
    free(hf);
    free(ht);
    return 0;
fail:
    if (df) hipFree(df);
    if (dt_) hipFree(dt_);
    if (dg) hipFree(dg);
    if (dl) hipFree(dl);
    free(hf);
    free(ht);
    return -1;
}

/* raw features x[n][12] -> raw outputs y[n][3] (the change of h, u, v) */
extern "C" int adhk_mlp_forward(adhk_mlp *m, int n, const float *x, float *y)
{
    float *dx = NULL, *dy = NULL;
    if (!m || n <= 0 || !x || !y) return -1;
    HC(hipMalloc((void **)&dx, (size_t)n * NIN * sizeof(float)));
    HC(hipMalloc((void **)&dy, (size_t)n * NOUT * sizeof(float)));
    HC(hipMemcpyAsync(dx, x, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_forward, dim3((n + 255) / 256), dim3(256), 0, m->stream, n, dx, m->wbf, m->norm, dy);
    HC(hipGetLastError());
    HC(hipMemcpyAsync(y, dy, (size_t)n * NOUT * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    HC(hipStreamSynchronize(m->stream));
    hipFree(dx);
    hipFree(dy);
    return 0;
fail:
    if (dx) hipFree(dx);
    if (dy) hipFree(dy);
    return -1;
}

extern "C" int adhk_sw2_set_warm_start(adhk_sw2 *c, adhk_mlp *m)
{
    if (!c) return -1;
    c->warm = m;
    if (m && ensure_static(c)) return -1;
    return 0;
}

/* the Newton start: sol = sol_old + MLP(features); the weights are read on the context's stream
 * after the model's own stream has finished its work */
extern "C" int adhk_mlp_predict_into(adhk_mlp *m, adhk_ctx *c)
{
    if (!m || !c) return -1;
    if (ensure_static(c)) return -1;
    HC(hipStreamSynchronize(m->stream));
    hipLaunchKernelGGL(k_predict, dim3((c->N + 255) / 256), dim3(256), 0, c->stream, c->N, c->sol, c->sol_old,
                       c->sol_older, c->brow, c->bcol, c->nstat, m->wbf, m->norm, c->node_inst, c->work);
    HC(hipGetLastError());
    return 0;
fail:
    return -1;
}
