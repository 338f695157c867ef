/*
 * mlp.hip -- learned warm start of the Newton iterate (gfx950).
 *
 * A per-node MLP, 12 -> 64 -> 64 -> 3 with GELU (erf form) and BF16 weights, its 64 x 64 layer on
 * the matrix cores (BF16 operands, FP32 sums) and the rest in FP32, predicts the change of (h, u, v)
 * over the next time step
 * from the node's h, hu, hv at t and t-1, the 1-ring neighbour means of h, hu,
 * hv at t, the bed slope and Manning's n (see node_features for the encoding).  It is trained on trajectories that the solver itself
 * produces, on the GPU, by the kernels below:
 *   k_node_static   per-node bed slope and Manning n (area-weighted over the node's elements)
 *   k_collect       features(t, t-1) and target sol(t+1) - sol(t) of every node -> dataset (SoA)
 *   k_stats         per-workgroup sums of each column and its square (normalisation)
 *   k_tile_grad     one wavefront per 64-sample tile: forward, MSE loss, backward; the hidden
 *                   layer's three 64 x 64 x 64 products on the matrix cores, activations staged in
 *                   LDS; the tile's weight gradient is written as one partial row
 *   k_reduce_adam   sums the tile partials in a fixed order (deterministic) and applies Adam to the
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

/* ---------------- the 64 x 64 layer on the matrix cores ----------------
 * v_mfma_f32_32x32x16_bf16 computes D (32x32, FP32) = A (32x16, BF16) B (16x32, BF16) + C on one wavefront.
 * Lane l holds A[l % 32][8 (l / 32) + v] and B[8 (l / 32) + v][l % 32] (v = 0..7, eight BF16 in 16 bytes) and
 * D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32] in its r-th accumulator (r = 0..15).  A 64 x 64 x 64 product is
 * 2 x 2 output tiles x 4 K-steps = 16 instructions.  The layer runs transposed, Z2^T = W2 A1^T: the M
 * dimension is the hidden unit, N the sample, so lane l ends with 32 hidden units of samples l % 32 and
 * 32 + l % 32, and the output layer needs one exchange with lane l ^ 32 instead of a pass through LDS. */
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
#define A1STR 72 /* BF16 per LDS row of a1: 144 bytes, rows 16-byte aligned */
#define FSTR 68  /* floats per LDS row of a 64-wide FP32 activation tile */

/* the FP32 weights that the VALU layers read, staged in LDS: W1 [64][12], b1, b2, W3 [3][64], b3 */
#define CW1 0
#define CB1 (CW1 + HID * NIN)
#define CB2 (CB1 + HID)
#define CW3 (CB2 + HID)
#define CB3 (CW3 + NOUT * HID)
#define NCOMPACT (CB3 + NOUT)

__device__ __forceinline__ void load_compact(const unsigned short *__restrict__ wbf, float *Wc)
{
    for (int p = threadIdx.x; p < NCOMPACT; p += blockDim.x) {
        int q;
        if (p < CB1)
            q = OFF_W1 + p;
        else if (p < CB2)
            q = OFF_B1 + (p - CB1);
        else if (p < CW3)
            q = OFF_B2 + (p - CB2);
        else if (p < CB3)
            q = OFF_W3 + (p - CW3);
        else
            q = OFF_B3 + (p - CB3);
        Wc[p] = bf2f(wbf[q]);
    }
}

__device__ __forceinline__ bf16x8 ld_bf16x8(const unsigned short *p)
{
    const uint4 u = *(const uint4 *)p;
    return __builtin_bit_cast(bf16x8, u);
}
/* eight FP32 values (any stride) rounded to BF16 as the operand of one instruction */
__device__ __forceinline__ bf16x8 pack_bf16x8(const float v[8])
{
    uint4 u;
    u.x = (unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16);
    u.y = (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16);
    u.z = (unsigned)f2bf(v[4]) | ((unsigned)f2bf(v[5]) << 16);
    u.w = (unsigned)f2bf(v[6]) | ((unsigned)f2bf(v[7]) << 16);
    return __builtin_bit_cast(bf16x8, u);
}
__device__ __forceinline__ int d_row(int r, int l) { return 8 * (r / 4) + 4 * (l / 32) + (r % 4); }

/* layer 1 of the lane's sample (VALU): z1 = W1 x + b1; its BF16 GELU into row l of A1s (and z1 into Z1, when given) */
__device__ __forceinline__ void layer1(const float *Wc, const float x[NIN], unsigned short *A1s, float *Z1)
{
    const int l = threadIdx.x & 63;
    for (int j0 = 0; j0 < HID; j0 += 8) {
        asm volatile("" ::: "memory"); /* one chunk of units' weights live at a time */
        unsigned short hb[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int j = j0 + u;
            float zz = Wc[CB1 + j];
#pragma unroll
            for (int k = 0; k < NIN; k++) zz += Wc[CW1 + j * NIN + k] * x[k];
            if (Z1) Z1[l * FSTR + j] = zz;
            hb[u] = f2bf(gelu(zz));
        }
        uint4 u4;
        u4.x = (unsigned)hb[0] | ((unsigned)hb[1] << 16);
        u4.y = (unsigned)hb[2] | ((unsigned)hb[3] << 16);
        u4.z = (unsigned)hb[4] | ((unsigned)hb[5] << 16);
        u4.w = (unsigned)hb[6] | ((unsigned)hb[7] << 16);
        *(uint4 *)(A1s + l * A1STR + j0) = u4;
    }
}

/* layer 2 on the matrix cores: acc[mt][nt][r] = z2 of hidden unit 32 mt + d_row(r, l), sample 32 nt + l % 32;
 * W2 row-major BF16 (global), a1 from A1s */
__device__ __forceinline__ void layer2(const unsigned short *__restrict__ w2, const unsigned short *A1s, const float *Wc,
                                       f32x16 acc[2][2])
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const float bb = Wc[CB2 + 32 * mt + d_row(r, l)];
            acc[mt][0][r] = bb;
            acc[mt][1][r] = bb;
        }
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
        bf16x8 a[2], b[2];
#pragma unroll
        for (int mt = 0; mt < 2; mt++) a[mt] = ld_bf16x8(w2 + (32 * mt + l % 32) * HID + 16 * ks + 8 * (l / 32));
#pragma unroll
        for (int nt = 0; nt < 2; nt++) b[nt] = ld_bf16x8(A1s + (32 * nt + l % 32) * A1STR + 16 * ks + 8 * (l / 32));
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                acc[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    }
}

/* output layer: a2 = GELU(z2) from the accumulators into A2 ([sample][unit], FP32, this wave's tile), then
 * y of the lane's own sample l from its row; every lane of the workgroup must call it (it synchronises the
 * workgroup).  Writing a2 as it is formed lets the accumulators die one by one instead of holding them, the
 * GELU temporaries and the output weights of 32 units at once. */
__device__ __forceinline__ void layer3(const float *Wc, f32x16 acc[2][2], float y[NOUT], float *A2)
{
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 16; r++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                A2[(32 * nt + l % 32) * FSTR + 32 * mt + d_row(r, l)] = gelu(acc[mt][nt][r]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NOUT; k++) y[k] = Wc[CB3 + k];
    for (int i = 0; i < HID; i++) {
        const float a2 = A2[l * FSTR + i];
#pragma unroll
        for (int k = 0; k < NOUT; k++) y[k] += Wc[CW3 + k * HID + i] * a2;
    }
}

/* forward pass of a wave's 64-sample tile (lane l: sample l) up to the normalised outputs.  Work is this
 * wave's LDS region (TILE x FSTR floats): a1 in BF16 first, then a2 over it (the matrix products have
 * consumed a1 by then).  Every lane of the workgroup must call it (it synchronises the workgroup). */
__device__ __forceinline__ void tile_forward(const float *Wc, const unsigned short *__restrict__ w2, float *Work,
                                             const float x[NIN], float y[NOUT])
{
    unsigned short *A1s = (unsigned short *)Work;
    layer1(Wc, x, A1s, nullptr);
    __syncthreads();
    f32x16 acc[2][2];
    layer2(w2, A1s, Wc, acc);
    layer3(Wc, acc, y, Work);
}

/* one wavefront = one tile of TILE samples; partial[tile][NPARAM] = d(loss)/d(params) over the tile,
 * loss = 1/(size*NOUT) * sum_s sum_k (y_sk - t_sk)^2 on normalised outputs.  The three 64 x 64 x 64
 * products of the hidden layer (z2, dW2 = dz2^T a1, da1 = dz2 W2) run on the matrix cores with BF16 a1
 * and dz2; the 12- and 3-wide layers, the biases and every other sum run in FP32 from LDS. */
__global__ __launch_bounds__(TILE) void k_tile_grad(Batch B, const unsigned short *__restrict__ wbf,
                                                    const float *__restrict__ norm, float *__restrict__ partial,
                                                    float *__restrict__ lossp)
{
    __shared__ float Wc[NCOMPACT];
    __shared__ __align__(16) unsigned short W2T[HID * A1STR]; /* W2 transposed: [j][i] */
    __shared__ __align__(16) unsigned short A1s[TILE * A1STR];
    __shared__ float X[TILE * NIN];
    __shared__ float Z1[TILE * FSTR];  /* z1 [s][j], later dz1 */
    __shared__ float A2[TILE * FSTR];  /* a2 [s][i], later dz2 */
    __shared__ float DY[TILE * 4];
    __shared__ float red[TILE];
    const int t = threadIdx.x, tile = blockIdx.x, l = t;
    const int s = tile * TILE + t;
    const bool valid = s < B.size;
    const unsigned short *w2 = wbf + OFF_W2;
    load_compact(wbf, Wc);
    for (int p = t; p < HID * HID; p += TILE) W2T[(p % HID) * A1STR + p / HID] = w2[p];
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
    __syncthreads(); /* the staged weights */
    /* forward */
    layer1(Wc, x, A1s, Z1);
    __syncthreads();
    f32x16 acc[2][2];
    layer2(w2, A1s, Wc, acc);
    float y[NOUT];
    layer3(Wc, acc, y, A2);
    float lsum = 0.0f, dy[NOUT];
    const float sc = 2.0f / ((float)B.size * (float)NOUT);
#pragma unroll
    for (int k = 0; k < NOUT; k++) {
        float d = y[k] - tn[k];
        lsum += d * d;
        dy[k] = valid ? sc * d : 0.0f;
        DY[t * 4 + k] = dy[k];
    }
    red[t] = valid ? lsum : 0.0f;
    __syncthreads();
    float *P = partial + (size_t)tile * NPARAM;
    /* dW3[k][i] = sum_s dy[s][k] a2[s][i], db3, tile loss */
    for (int e = t; e < NOUT * HID; e += TILE) {
        int k = e / HID, i = e % HID;
        float a = 0.0f;
        for (int q = 0; q < TILE; q++) a += DY[q * 4 + k] * A2[q * FSTR + i];
        P[OFF_W3 + e] = a;
    }
    if (t < NOUT) {
        float a = 0.0f;
        for (int q = 0; q < TILE; q++) a += DY[q * 4 + t];
        P[OFF_B3 + t] = a;
    }
    if (t == 0) {
        float a = 0.0f;
        for (int q = 0; q < TILE; q++) a += red[q];
        lossp[tile] = a;
    }
    /* dz2 = (W3^T dy) gelu'(z2) in the accumulator layout */
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int i = 32 * mt + d_row(r, l);
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const int q = 32 * nt + l % 32;
                float g = 0.0f;
#pragma unroll
                for (int k = 0; k < NOUT; k++) g += Wc[CW3 + k * HID + i] * DY[q * 4 + k];
                acc[mt][nt][r] = g * gelu_d(acc[mt][nt][r]);
            }
        }
    __syncthreads(); /* every read of the a2 rows is done */
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int r = 0; r < 16; r++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++) A2[(32 * nt + l % 32) * FSTR + 32 * mt + d_row(r, l)] = acc[mt][nt][r];
    __syncthreads(); /* dz2 [s][i] complete */
    /* dW2 = dz2^T a1 (M = unit i, N = input unit j, K = sample): A[i][s] = dz2, B[s][j] = a1, both strided */
    {
        f32x16 g2[2][2];
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
                for (int r = 0; r < 16; r++) g2[mt][nt][r] = 0.0f;
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int q0 = 16 * ks + 8 * (l / 32);
            bf16x8 a[2], b[2];
#pragma unroll
            for (int mt = 0; mt < 2; mt++) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = A2[(q0 + u) * FSTR + 32 * mt + l % 32];
                a[mt] = pack_bf16x8(v);
            }
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                uint4 u4;
                const unsigned short *c = A1s + q0 * A1STR + 32 * nt + l % 32;
                u4.x = (unsigned)c[0] | ((unsigned)c[A1STR] << 16);
                u4.y = (unsigned)c[2 * A1STR] | ((unsigned)c[3 * A1STR] << 16);
                u4.z = (unsigned)c[4 * A1STR] | ((unsigned)c[5 * A1STR] << 16);
                u4.w = (unsigned)c[6 * A1STR] | ((unsigned)c[7 * A1STR] << 16);
                b[nt] = __builtin_bit_cast(bf16x8, u4);
            }
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
                    g2[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mt], b[nt], g2[mt][nt], 0, 0, 0);
        }
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                    P[OFF_W2 + (32 * mt + d_row(r, l)) * HID + 32 * nt + l % 32] = g2[mt][nt][r];
    }
    /* db2[i] = sum_s dz2[s][i] (FP32) */
    {
        float a = 0.0f;
        for (int q = 0; q < TILE; q++) a += A2[q * FSTR + t];
        P[OFF_B2 + t] = a;
    }
    /* da1^T = W2^T dz2^T (M = input unit j, N = sample, K = unit i): A[j][i] = W2T rows, B[i][s] = dz2 rows */
    {
        f32x16 d1[2][2];
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
                for (int r = 0; r < 16; r++) d1[mt][nt][r] = 0.0f;
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int i0 = 16 * ks + 8 * (l / 32);
            bf16x8 a[2], b[2];
#pragma unroll
            for (int mt = 0; mt < 2; mt++) a[mt] = ld_bf16x8(W2T + (32 * mt + l % 32) * A1STR + i0);
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = A2[(32 * nt + l % 32) * FSTR + i0 + u];
                b[nt] = pack_bf16x8(v);
            }
#pragma unroll
            for (int mt = 0; mt < 2; mt++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
                    d1[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mt], b[nt], d1[mt][nt], 0, 0, 0);
        }
        /* dz1 = da1 gelu'(z1), over z1 in place (each (sample, unit) pair belongs to one lane) */
#pragma unroll
        for (int mt = 0; mt < 2; mt++)
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    float *z = Z1 + (32 * nt + l % 32) * FSTR + 32 * mt + d_row(r, l);
                    *z = d1[mt][nt][r] * gelu_d(*z);
                }
    }
    __syncthreads(); /* dz1 [s][j] complete */
    /* dW1[j][k] = sum_s dz1[s][j] x[s][k]: 768 entries, 12 per thread; db1 */
    for (int e = t; e < HID * NIN; e += TILE) {
        int j = e / NIN, k = e % NIN;
        float a = 0.0f;
        for (int q = 0; q < TILE; q++) a += Z1[q * FSTR + j] * X[q * NIN + k];
        P[OFF_W1 + e] = a;
    }
    {
        float a = 0.0f;
        for (int q = 0; q < TILE; q++) a += Z1[q * FSTR + t];
        P[OFF_B1 + t] = a;
    }
}

/* grad[p] = sum over the tiles; adam != 0: Adam step on master, BF16 copy refreshed.  A workgroup takes
 * RPW parameters and RG threads per parameter: thread g sums tiles g, g + RG, g + 2 RG, ... in that order,
 * then the RG partial sums are added in g order (deterministic; 32 consecutive parameters per tile row
 * make each load one 128-byte run).  Block 0 also sums the tile losses. */
#define RPW 32
#define RG 8
__global__ __launch_bounds__(RPW * RG) void k_reduce_adam(int ntiles, const float *__restrict__ partial,
                                                          const float *__restrict__ lossp, float *grad_out,
                                                          double *loss_out, int adam, float lr, float b1c, float b2c,
                                                          float *master, float *m1, float *m2, unsigned short *wbf)
{
    __shared__ float part[RG][RPW];
    const int pl = threadIdx.x % RPW, g = threadIdx.x / RPW;
    const int p = blockIdx.x * RPW + pl;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) {
        double l = 0.0;
        for (int q = 0; q < ntiles; q++) l += lossp[q];
        *loss_out = l;
    }
    float a = 0.0f;
    if (p < NPARAM)
        for (int q = g; q < ntiles; q += RG) a += partial[(size_t)q * NPARAM + p];
    part[g][pl] = a;
    __syncthreads();
    if (g != 0 || p >= NPARAM) return;
    float gsum = part[0][pl];
    for (int k = 1; k < RG; k++) gsum += part[k][pl];
    if (grad_out) grad_out[p] = gsum;
    if (!adam) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float m = beta1 * m1[p] + (1.0f - beta1) * gsum;
    float v = beta2 * m2[p] + (1.0f - beta2) * gsum * gsum;
    m1[p] = m;
    m2[p] = v;
    float w = master[p] - lr * (m / b1c) / (sqrtf(v / b2c) + eps);
    master[p] = w;
    wbf[p] = f2bf(w);
}

/* ---------------- inference ---------------- */
/* raw features in, raw outputs (denormalised) out; two 64-sample tiles per workgroup */
#define INF_TPB 128
__global__ __launch_bounds__(INF_TPB) void k_forward(int n, const float *__restrict__ xin,
                                                     const unsigned short *__restrict__ wbf,
                                                     const float *__restrict__ norm, float *__restrict__ yout)
{
    __shared__ float Wc[NCOMPACT];
    __shared__ __align__(16) float Work[(INF_TPB / 64) * TILE * FSTR];
    load_compact(wbf, Wc);
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = s < n; /* every lane takes part in the wave's matrix products */
    float x[NIN], y[NOUT];
#pragma unroll
    for (int k = 0; k < NIN; k++) x[k] = valid ? (xin[(size_t)s * NIN + k] - norm[k]) * norm[NIN + k] : 0.0f;
    tile_forward(Wc, wbf + OFF_W2, Work + (threadIdx.x / 64) * TILE * FSTR, x, y);
    if (!valid) return;
#pragma unroll
    for (int k = 0; k < NOUT; k++) yout[(size_t)s * NOUT + k] = y[k] * norm[27 + k] + norm[24 + k];
}

/* sol = sol_old + MLP(features(sol_old, sol_older)) */
__global__ __launch_bounds__(INF_TPB) void k_predict(int N, double *__restrict__ sol, const double *__restrict__ st,
                                                     const double *__restrict__ stm1, const int *__restrict__ brow,
                                                     const int *__restrict__ bcol, const float *__restrict__ nstat,
                                                     const unsigned short *__restrict__ wbf,
                                                     const float *__restrict__ norm, const int *__restrict__ node_inst,
                                                     const int *__restrict__ work)
{
    __shared__ float Wc[NCOMPACT];
    __shared__ __align__(16) float Work[(INF_TPB / 64) * TILE * FSTR];
    load_compact(wbf, Wc);
    __syncthreads();
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = n < N && work[node_inst[n]]; /* every lane takes part in the wave's matrix products */
    float f[NIN], y[NOUT];
    if (active) {
        node_features(n, N, st, stm1, brow, bcol, nstat, f);
#pragma unroll
        for (int k = 0; k < NIN; k++) f[k] = (f[k] - norm[k]) * norm[NIN + k];
    } else {
#pragma unroll
        for (int k = 0; k < NIN; k++) f[k] = 0.0f;
    }
    tile_forward(Wc, wbf + OFF_W2, Work + (threadIdx.x / 64) * TILE * FSTR, f, y);
    if (!active) return;
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
        hipLaunchKernelGGL(k_reduce_adam, dim3((NPARAM + RPW - 1) / RPW), dim3(RPW * RG), 0, m->stream, nt, m->partials,
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
        hipLaunchKernelGGL(k_reduce_adam, dim3((NPARAM + RPW - 1) / RPW), dim3(RPW * RG), 0, m->stream, nt, m->partials,
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
    hipLaunchKernelGGL(k_forward, dim3((n + INF_TPB - 1) / INF_TPB), dim3(INF_TPB), 0, m->stream, n, dx, m->wbf, m->norm, dy);
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
    hipLaunchKernelGGL(k_predict, dim3((c->N + INF_TPB - 1) / INF_TPB), dim3(INF_TPB), 0, c->stream, c->N, c->sol,
                       c->sol_old,
                       c->sol_older, c->brow, c->bcol, c->nstat, m->wbf, m->norm, c->node_inst, c->work);
    HC(hipGetLastError());
    return 0;
fail:
    return -1;
}
