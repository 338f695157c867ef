/*
 * sgs_setup.c -- the permuted layout of the multicolour symmetric Gauss-Seidel preconditioner, built once
 * per context from the node-block pattern (no counterpart in the reference, whose preconditioner is an
 * exact LU: bcgstab_solver.c:371-424).
 *
 * Colouring: DSatur (colour_dsatur below), then Kempe-chain exchanges that drop colours where they can
 * (colour_reduce).  The rows are then permuted by colour and every colour's run padded to a multiple of 256 rows, so that each 256-row reduction block and each
 * 64-row slice belongs to one colour and a colour's rows form whole workgroups.  Inside a colour the rows
 * are sorted by their U and L widths within windows of 1024 (rows of one colour do not depend on each
 * other, so the order is free), which keeps the SELL padding small on meshes of uneven rows.  The folded matrix's
 * off-diagonal blocks go to U (neighbour of a higher colour) or L (lower colour), each stored as its own
 * SELL-64 over the permuted rows, blocks of a row in ascending natural column order.  A padding slot
 * points at the zero row Np, whose vector entries are zero and never written.
 */
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include "adhk_internal.h"

#define HC(x)                                                                                                     \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));                        \
            goto fail;                                                                                            \
        }                                                                                                         \
    } while (0)

#define SGS_VECS 8 /* r (= s) p Ap As q xt t (= u) bt, see sg_vec in la_kernels.hip */

/* DSatur colouring (Brelaz, "New methods to color the vertices of a graph", CACM 22(4), 1979) over the block
 * pattern: repeatedly the uncoloured node with the most distinct colours among its neighbours (saturation),
 * ties by the most uncoloured neighbours, then the lowest index, takes the smallest colour none of its
 * neighbours holds.  On a triangle mesh a node with two coloured corners of a triangle is saturated by both,
 * so the colouring propagates forced colours: a mesh whose interior nodes all have even degree (the flumes)
 * gets 3 colours, the fewest a triangle allows.  A max-heap with lazy deletion (an entry is current when its
 * node is uncoloured and its keys match).  Returns the colour count, or -1 beyond ADHK_SGS_MAXC. */
typedef struct {
    int sat, udeg, node;
} ds_ent;
static int ds_below(ds_ent a, ds_ent b)
{
    if (a.sat != b.sat) return a.sat < b.sat;
    if (a.udeg != b.udeg) return a.udeg < b.udeg;
    return a.node > b.node;
}
static void heap_push(ds_ent *h, long *n, ds_ent e)
{
    long i = (*n)++;
    h[i] = e;
    while (i > 0) {
        const long up = (i - 1) / 2;
        if (!ds_below(h[up], h[i])) break;
        const ds_ent t = h[up];
        h[up] = h[i];
        h[i] = t;
        i = up;
    }
}
static ds_ent heap_pop(ds_ent *h, long *n)
{
    const ds_ent top = h[0];
    h[0] = h[--(*n)];
    long i = 0;
    for (;;) {
        const long a = 2 * i + 1, b = a + 1;
        long m = i;
        if (a < *n && ds_below(h[m], h[a])) m = a;
        if (b < *n && ds_below(h[m], h[b])) m = b;
        if (m == i) break;
        const ds_ent t = h[m];
        h[m] = h[i];
        h[i] = t;
        i = m;
    }
    return top;
}
static int colour_dsatur(int N, const int *brow, const int *bcol, int *colour)
{
    unsigned long long *seen = calloc((size_t)(N > 0 ? N : 1), sizeof *seen); /* neighbour colours, bit set */
    int *sat = calloc((size_t)(N > 0 ? N : 1), sizeof(int)), *udeg = calloc((size_t)(N > 0 ? N : 1), sizeof(int));
    ds_ent *h = malloc(((size_t)brow[N] + (size_t)N + 1) * sizeof *h); /* every push: a node once, then per neighbour */
    long n = 0;
    int nc = 0;
    if (!seen || !sat || !udeg || !h) {
        nc = -2;
        goto done;
    }
    for (int v = 0; v < N; v++) {
        colour[v] = -1;
        udeg[v] = brow[v + 1] - brow[v] - 1; /* the block row holds the node itself */
        heap_push(h, &n, (ds_ent){0, udeg[v], v});
    }
    while (n > 0) {
        const ds_ent e = heap_pop(h, &n);
        const int v = e.node;
        if (colour[v] >= 0 || e.sat != sat[v] || e.udeg != udeg[v]) continue;
        int c = 0;
        while (c < ADHK_SGS_MAXC && (seen[v] >> c & 1ULL)) c++;
        if (c >= ADHK_SGS_MAXC) {
            nc = -1;
            goto done;
        }
        colour[v] = c;
        if (c + 1 > nc) nc = c + 1;
        for (int b = brow[v]; b < brow[v + 1]; b++) {
            const int j = bcol[b];
            if (j == v || colour[j] >= 0) continue;
            udeg[j]--;
            if (!(seen[j] >> c & 1ULL)) {
                seen[j] |= 1ULL << c;
                sat[j]++;
            }
            heap_push(h, &n, (ds_ent){sat[j], udeg[j], j});
        }
    }
done:
    free(seen);
    free(sat);
    free(udeg);
    free(h);
    return nc;
}

/* Colour reduction by Kempe chains (Kempe 1879, the interchange of the five-colour proof): the colours a
 * and b may be exchanged on any connected component of the subgraph the two colours induce, and the
 * colouring stays proper.  While the top colour t is in use, each node v of colour t in node order moves to
 * the least colour below t that none of its neighbours holds; failing that, for the pairs a < b < t in
 * order, it collects the {a, b} components that hold v's neighbours of colour a (depth first, at most
 * SGS_KEMPE_CAP nodes) and, when none of v's neighbours of colour b lies in them, exchanges a and b there and
 * takes a.  Once no node holds t the colour is gone; a node that no pair frees restores the colouring from
 * before the pass and ends the reduction.  The outcome depends only on the components, not on the order of
 * the search.  Meshes whose DSatur colouring ends with a colour of a few nodes lose it (one such phase
 * costs a launch, two per product, whatever its size); the SGS iteration counts do not change with it.
 * Returns the colour count, or -2 when out of host memory. */
#define SGS_KEMPE_CAP 4096
static int kempe_free(int v, int a, int b, const int *brow, const int *bcol, int *colour, int *seen, int tag,
                      int *stack)
{
    int top = 0, taken = 0;
    for (int e = brow[v]; e < brow[v + 1]; e++) {
        const int j = bcol[e];
        if (j == v || colour[j] != a || seen[j] == tag) continue;
        if (taken == SGS_KEMPE_CAP) return 0;
        seen[j] = tag;
        stack[taken++] = j;
    }
    while (top < taken) { /* the entries below `taken` are the visited set; `top` walks them */
        const int x = stack[top++];
        for (int e = brow[x]; e < brow[x + 1]; e++) {
            const int y = bcol[e];
            if (seen[y] == tag || (colour[y] != a && colour[y] != b)) continue;
            if (colour[y] == b) {
                for (int f = brow[v]; f < brow[v + 1]; f++)
                    if (bcol[f] == y) return 0; /* the chain reaches a b-neighbour of v: no exchange */
            }
            if (taken == SGS_KEMPE_CAP) return 0;
            seen[y] = tag;
            stack[taken++] = y;
        }
    }
    for (int k = 0; k < taken; k++) colour[stack[k]] = a + b - colour[stack[k]];
    colour[v] = a;
    return 1;
}
static int colour_reduce(int N, const int *brow, const int *bcol, int *colour, int nc)
{
    int *seen = calloc((size_t)(N > 0 ? N : 1), sizeof(int)), *stack = malloc(SGS_KEMPE_CAP * sizeof(int));
    int *before = malloc((size_t)(N > 0 ? N : 1) * sizeof(int));
    int tag = 0;
    if (!seen || !stack || !before) nc = -2;
    while (nc > 1) {
        const int t = nc - 1;
        int freed = 1;
        memcpy(before, colour, (size_t)N * sizeof(int));
        for (int v = 0; v < N && freed; v++) {
            if (colour[v] != t) continue;
            unsigned long long held = 0;
            for (int e = brow[v]; e < brow[v + 1]; e++)
                if (bcol[e] != v) held |= 1ULL << colour[bcol[e]];
            const int c = __builtin_ctzll(~held);
            if (c < t) {
                colour[v] = c;
                continue;
            }
            freed = 0;
            for (int a = 0; a < t && !freed; a++)
                for (int b = a + 1; b < t && !freed; b++) {
                    if (tag == INT_MAX) { /* a fresh mark per search; start the marks again before they wrap */
                        memset(seen, 0, (size_t)N * sizeof(int));
                        tag = 0;
                    }
                    freed = kempe_free(v, a, b, brow, bcol, colour, seen, ++tag, stack);
                }
        }
        if (!freed) {
            memcpy(colour, before, (size_t)N * sizeof(int));
            break;
        }
        nc = t;
    }
    free(seen);
    free(stack);
    free(before);
    return nc;
}

/* rows of a colour are sorted by this key within windows of SGS_WINDOW rows: U blocks x 64 + L blocks */
#define SGS_WINDOW 1024
#define ROW_KEY_MAX (64 * 64)
static unsigned row_key(int n, const int *brow, const int *bcol, const int *colour)
{
    unsigned nu = 0, nl = 0;
    for (int b = brow[n]; b < brow[n + 1]; b++) {
        const int j = bcol[b];
        if (j == n) continue;
        if (colour[j] > colour[n])
            nu++;
        else
            nl++;
    }
    nu = nu < 63 ? nu : 63;
    nl = nl < 63 ? nl : 63;
    return nu * 64 + nl;
}

static int cmp_u64(const void *a, const void *b)
{
    const unsigned long long x = *(const unsigned long long *)a, y = *(const unsigned long long *)b;
    return (x > y) - (x < y);
}

/* one part (upper: colour(j) > colour(i), else lower) as SELL-64 over the permuted rows; dst (the natural
 * SELL positions) gets each block's place in the part: q in U, -(q + 2) in L */
static int build_part(adhk_sgs_part *P, int upper, int Np, const int *prow, const int *colour, const int *brow,
                      const int *bcol, const int *blk2sell, const int *posof, int *dst)
{
    const int nsl = Np / 64;
    int *ptr = malloc(((size_t)nsl + 1) * sizeof(int)), *col = NULL;
    if (!ptr) goto oom;
    ptr[0] = 0;
    for (int s = 0; s < nsl; s++) {
        int w = 0;
        for (int l = 0; l < 64; l++) {
            const int n = prow[s * 64 + l];
            if (n < 0) continue;
            int k = 0;
            for (int b = brow[n]; b < brow[n + 1]; b++) {
                const int j = bcol[b];
                if (j != n && (upper ? colour[j] > colour[n] : colour[j] < colour[n])) k++;
            }
            if (k > w) w = k;
        }
        ptr[s + 1] = ptr[s] + w;
    }
    P->nslot = ptr[nsl];
    {
        const size_t npos = (size_t)(P->nslot > 0 ? P->nslot : 1) * 64;
        col = malloc(npos * sizeof(int));
        if (!col) goto oom;
        for (size_t q = 0; q < npos; q++) col[q] = Np; /* the zero row */
        for (int s = 0; s < nsl; s++)
            for (int l = 0; l < 64; l++) {
                const int n = prow[s * 64 + l];
                if (n < 0) continue;
                int k = 0;
                for (int b = brow[n]; b < brow[n + 1]; b++) {
                    const int j = bcol[b];
                    if (j == n || !(upper ? colour[j] > colour[n] : colour[j] < colour[n])) continue;
                    const size_t q = (size_t)(ptr[s] + k) * 64 + l;
                    col[q] = posof[j];
                    dst[blk2sell[b]] = upper ? (int)q : -((int)q + 2);
                    k++;
                }
            }
        HC(hipMalloc((void **)&P->ptr, ((size_t)nsl + 1) * sizeof(int)));
        HC(hipMalloc((void **)&P->col, npos * sizeof(int)));
        HC(hipMemcpy(P->ptr, ptr, ((size_t)nsl + 1) * sizeof(int), hipMemcpyHostToDevice));
        HC(hipMemcpy(P->col, col, npos * sizeof(int), hipMemcpyHostToDevice));
    }
    free(ptr);
    free(col);
    return 0;
oom:
    adhk_set_error("adhk_sgs_setup: out of host memory");
fail:
    free(ptr);
    free(col);
    return -1;
}

int adhk_sgs_setup(adhk_ctx *c)
{
    adhk_sgs *G = &c->sgs;
    const int N = c->N, nnzb = c->nnzb;
    int *brow = NULL, *bcol = NULL, *b2s = NULL, *prow = NULL, *posof = NULL, *cnt = NULL, *fill = NULL, *dst = NULL;
    const size_t nnat = (size_t)c->nslot * 64;
    int ib[2];
    if (G->ready) return 0;
    if (c->ninst != 1) {
        adhk_set_error("adhk_sgs_setup: the multicolour preconditioner serves one mesh alone");
        return -1;
    }
    brow = malloc(((size_t)N + 1) * sizeof(int));
    bcol = malloc((size_t)(nnzb > 0 ? nnzb : 1) * sizeof(int));
    b2s = malloc((size_t)(nnzb > 0 ? nnzb : 1) * sizeof(int));
    G->colour = malloc((size_t)(N > 0 ? N : 1) * sizeof(int));
    posof = malloc((size_t)(N > 0 ? N : 1) * sizeof(int));
    cnt = calloc(ADHK_SGS_MAXC, sizeof(int));
    fill = calloc(ADHK_SGS_MAXC, sizeof(int));
    dst = malloc((nnat > 0 ? nnat : 1) * sizeof(int));
    if (!brow || !bcol || !b2s || !G->colour || !posof || !cnt || !fill || !dst) {
        adhk_set_error("adhk_sgs_setup: out of host memory");
        goto fail;
    }
    HC(hipMemcpy(brow, c->brow, ((size_t)N + 1) * sizeof(int), hipMemcpyDeviceToHost));
    HC(hipMemcpy(bcol, c->bcol, (size_t)nnzb * sizeof(int), hipMemcpyDeviceToHost));
    HC(hipMemcpy(b2s, c->blk2sell, (size_t)nnzb * sizeof(int), hipMemcpyDeviceToHost));
    G->nc = colour_dsatur(N, brow, bcol, G->colour);
    if (G->nc >= 2) G->nc = colour_reduce(N, brow, bcol, G->colour, G->nc);
    if (G->nc < 2) {
        adhk_set_error("adhk_sgs_setup: %s", G->nc == -2 ? "out of host memory"
                                                 : "more than 64 colours, or a pattern without couplings");
        goto fail;
    }
    for (int i = 0; i < N; i++) cnt[G->colour[i]]++;
    G->cs[0] = 0;
    for (int k = 0; k < G->nc; k++) G->cs[k + 1] = G->cs[k] + (cnt[k] + 255) / 256 * 256;
    G->Np = G->cs[G->nc];
    G->nblk = G->Np / 256;
    if ((long)G->Np * 3 + 3 >= 2147483647L / 9) {
        adhk_set_error("adhk_sgs_setup: too many rows");
        goto fail;
    }
    prow = malloc((size_t)G->Np * sizeof(int));
    if (!prow) {
        adhk_set_error("adhk_sgs_setup: out of host memory");
        goto fail;
    }
    for (int i = 0; i < G->Np; i++) prow[i] = -1;
    for (int n = 0; n < N; n++) {
        const int k = G->colour[n], i = G->cs[k] + fill[k]++;
        prow[i] = n;
    }
    /* Rows of one colour are independent of each other, so their order within the colour is free: inside
     * windows of SGS_WINDOW rows, sort them by (U blocks, L blocks), descending, ties by node.  A slice then
     * holds rows of equal widths and pads little on meshes of uneven rows; the windows keep neighbours
     * close in memory. */
    {
        unsigned long long *kv = malloc(SGS_WINDOW * sizeof *kv);
        if (!kv) {
            adhk_set_error("adhk_sgs_setup: out of host memory");
            goto fail;
        }
        for (int k = 0; k < G->nc; k++)
            for (int w0 = G->cs[k]; w0 < G->cs[k] + cnt[k]; w0 += SGS_WINDOW) {
                const int w1 = w0 + SGS_WINDOW < G->cs[k] + cnt[k] ? w0 + SGS_WINDOW : G->cs[k] + cnt[k];
                for (int i = w0; i < w1; i++) /* ascending order of (inverted key, node) */
                    kv[i - w0] = ((unsigned long long)(ROW_KEY_MAX - row_key(prow[i], brow, bcol, G->colour)) << 32) |
                                 (unsigned)prow[i];
                qsort(kv, (size_t)(w1 - w0), sizeof *kv, cmp_u64);
                for (int i = w0; i < w1; i++) prow[i] = (int)(kv[i - w0] & 0xffffffffULL);
            }
        free(kv);
    }
    for (int i = 0; i < G->Np; i++)
        if (prow[i] >= 0) posof[prow[i]] = i;
    for (size_t q = 0; q < nnat; q++) dst[q] = -1;
    if (build_part(&G->U, 1, G->Np, prow, G->colour, brow, bcol, b2s, posof, dst)) goto fail;
    if (build_part(&G->L, 0, G->Np, prow, G->colour, brow, bcol, b2s, posof, dst)) goto fail;
    HC(hipMalloc((void **)&G->dst, (nnat > 0 ? nnat : 1) * sizeof(int)));
    HC(hipMemcpy(G->dst, dst, nnat * sizeof(int), hipMemcpyHostToDevice));
    HC(hipMalloc((void **)&G->prow, (size_t)G->Np * sizeof(int)));
    HC(hipMemcpy(G->prow, prow, (size_t)G->Np * sizeof(int), hipMemcpyHostToDevice));
    ib[0] = 0;
    ib[1] = G->nblk;
    HC(hipMalloc((void **)&G->inst_blk, 2 * sizeof(int)));
    HC(hipMemcpy(G->inst_blk, ib, 2 * sizeof(int), hipMemcpyHostToDevice));
    HC(hipMalloc((void **)&G->partials, 8 * (size_t)G->nblk * sizeof(double)));
    G->ready = 1;
    free(brow);
    free(bcol);
    free(b2s);
    free(prow);
    free(posof);
    free(cnt);
    free(fill);
    free(dst);
    return 0;
fail:
    free(dst);
    free(brow);
    free(bcol);
    free(b2s);
    free(prow);
    free(posof);
    free(cnt);
    free(fill);
    adhk_sgs_free(c);
    return -1;
}

int adhk_sgs_ensure(adhk_ctx *c, int mixed)
{
    adhk_sgs *G = &c->sgs;
    if (adhk_sgs_setup(c)) return -1;
    {
        const size_t vb = mixed ? sizeof(float) : sizeof(double);
        const size_t nvec = (size_t)3 * ((size_t)G->Np + 1);
        void **vp = mixed ? (void **)&G->v32 : (void **)&G->v64;
        void **up = mixed ? (void **)&G->U.val32 : (void **)&G->U.val;
        void **lp = mixed ? (void **)&G->L.val32 : (void **)&G->L.val;
        if (!*vp) {
            /* zeroed once: the zero row and the padding rows stay zero through every product */
            HC(hipMalloc(vp, SGS_VECS * nvec * vb));
            HC(hipMemset(*vp, 0, SGS_VECS * nvec * vb));
        }
        /* zeroed once: the parts' padding slots are never written again (k_sgs_scatter writes real blocks) */
        if (!*up) {
            const size_t nb = (size_t)(G->U.nslot > 0 ? G->U.nslot : 1) * 64 * 9 * vb;
            HC(hipMalloc(up, nb));
            HC(hipMemset(*up, 0, nb));
        }
        if (!*lp) {
            const size_t nb = (size_t)(G->L.nslot > 0 ? G->L.nslot : 1) * 64 * 9 * vb;
            HC(hipMalloc(lp, nb));
            HC(hipMemset(*lp, 0, nb));
        }
    }
    return 0;
fail:
    return -1;
}

static void free_part(adhk_sgs_part *P)
{
    if (P->ptr) (void)hipFree(P->ptr);
    if (P->col) (void)hipFree(P->col);
    if (P->val) (void)hipFree(P->val);
    if (P->val32) (void)hipFree(P->val32);
    memset(P, 0, sizeof *P);
}

void adhk_sgs_free(adhk_ctx *c)
{
    adhk_sgs *G = &c->sgs;
    free_part(&G->U);
    free_part(&G->L);
    if (G->prow) (void)hipFree(G->prow);
    if (G->dst) (void)hipFree(G->dst);
    if (G->inst_blk) (void)hipFree(G->inst_blk);
    if (G->partials) (void)hipFree(G->partials);
    if (G->v64) (void)hipFree(G->v64);
    if (G->v32) (void)hipFree(G->v32);
    free(G->colour);
    memset(G, 0, sizeof *G);
}

int adhk_sw2_colouring(adhk_sw2 *c, int *colour, int *ncolour, int *layout)
{
    if (!c) return -1;
    if (adhk_sgs_setup(c)) return -1;
    if (colour) memcpy(colour, c->sgs.colour, (size_t)c->N * sizeof(int));
    if (ncolour) *ncolour = c->sgs.nc;
    if (layout) {
        layout[0] = c->sgs.Np;
        layout[1] = c->sgs.U.nslot;
        layout[2] = c->sgs.L.nslot;
    }
    return 0;
}
