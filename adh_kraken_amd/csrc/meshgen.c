/*
 * meshgen.c -- unstructured meshes for the SW2 channel workloads: a Poisson-disk point set, a
 * reverse Cuthill-McKee node order, and the bed / roughness / boundary data of a channel with
 * floodplain (straight or meandering).  The triangulation itself (Delaunay) is done by the caller.
 *
 * The reference builds its meshes from files or from create_rectangular_grid
 * (src/tools/create_rectangular_grid.c) and reorders nodes with its own RCM
 * (src/structs/sgrid/sgrid_reorder.c:468-922); the order here is the textbook Cuthill-McKee
 * breadth-first order from a pseudo-peripheral node (George & Liu, "Computer Solution of Large
 * Sparse Positive Definite Systems", 1981), reversed, which keeps a node's neighbours within one
 * breadth-first level of it -- what the SELL-64 slices and their neighbour gathers need.
 *
 * Random numbers: SplitMix64, u = (x >> 11) * 2^-53.  Point sampling draws from the state seeded
 * with spec->seed; the bumps draw cx, cy, amplitude, sigma per bump from the state seeded with
 * spec->seed + 0x5BD1E995.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/adhk.h"

static uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double unif(uint64_t *s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }

#define PD_TRIES 30
#define TWO_PI 6.283185307179586476925

/* boundary points: the four corners and points on each edge at the largest spacing L/m with m
 * intervals that keeps every gap above r (so no two boundary points are closer than r) */
static int edge_count(double len, double r)
{
    int m = (int)floor(len / (r * 1.0001));
    return m < 1 ? 1 : m;
}

int adhk_poisson_disk_bound(double x0, double x1, double y0, double y1, double r)
{
    if (!(x1 > x0) || !(y1 > y0) || !(r > 0)) return -1;
    /* disks of radius r/2 around the points are disjoint: hexagonal packing bounds the count */
    double lx = x1 - x0 + r, ly = y1 - y0 + r;
    double n = lx * ly / (0.5 * sqrt(3.0) * r * r) + 16.0;
    return n > 2e9 ? -1 : (int)n;
}

int adhk_poisson_disk(double x0, double x1, double y0, double y1, double r, unsigned long long seed, int max_n,
                      double *x, double *y)
{
    if (!(x1 > x0) || !(y1 > y0) || !(r > 0) || !x || !y || max_n < 4) return -1;
    const double cell = r / sqrt(2.0);
    const long gx = (long)ceil((x1 - x0) / cell) + 1, gy = (long)ceil((y1 - y0) / cell) + 1;
    int *grid = (int *)malloc((size_t)gx * (size_t)gy * sizeof(int));
    int *active = (int *)malloc((size_t)max_n * sizeof(int));
    if (!grid || !active) {
        free(grid);
        free(active);
        return -1;
    }
    for (long i = 0; i < gx * gy; i++) grid[i] = -1;
    int n = 0, na = 0, overflow = 0;
    const double r2 = r * r;
#define CELLX(px) ((long)(((px) - x0) / cell))
#define CELLY(py) ((long)(((py) - y0) / cell))
#define PUT(px, py)                                                                                                   \
    do {                                                                                                              \
        if (n == max_n) {                                                                                             \
            overflow = 1;                                                                                             \
            break;                                                                                                    \
        }                                                                                                             \
        x[n] = (px);                                                                                                  \
        y[n] = (py);                                                                                                  \
        grid[CELLY(py) * gx + CELLX(px)] = n;                                                                         \
        active[na++] = n;                                                                                             \
        n++;                                                                                                          \
    } while (0)
    const int mx = edge_count(x1 - x0, r), my = edge_count(y1 - y0, r);
    const double sx = (x1 - x0) / mx, sy = (y1 - y0) / my;
    for (int i = 0; i <= mx; i++) PUT(i == mx ? x1 : x0 + i * sx, y0);
    for (int j = 1; j <= my; j++) PUT(x1, j == my ? y1 : y0 + j * sy);
    for (int i = mx - 1; i >= 0; i--) PUT(x0 + i * sx, y1);
    for (int j = my - 1; j >= 1; j--) PUT(x0, y0 + j * sy);
    uint64_t st = (uint64_t)seed;
    while (na > 0 && !overflow) {
        const int ai = (int)(unif(&st) * na);
        const int p = active[ai];
        int placed = 0;
        for (int t = 0; t < PD_TRIES; t++) {
            /* uniform in the annulus r <= d < 2r */
            const double th = TWO_PI * unif(&st);
            const double rho = r * sqrt(1.0 + 3.0 * unif(&st));
            const double qx = x[p] + rho * cos(th), qy = y[p] + rho * sin(th);
            if (!(qx > x0 && qx < x1 && qy > y0 && qy < y1)) continue;
            const long cx = CELLX(qx), cy = CELLY(qy);
            int ok = 1;
            for (long j = cy - 2; j <= cy + 2 && ok; j++) {
                if (j < 0 || j >= gy) continue;
                for (long i = cx - 2; i <= cx + 2; i++) {
                    if (i < 0 || i >= gx) continue;
                    const int o = grid[j * gx + i];
                    if (o >= 0) {
                        const double dx = x[o] - qx, dy = y[o] - qy;
                        if (dx * dx + dy * dy < r2) {
                            ok = 0;
                            break;
                        }
                    }
                }
            }
            if (ok) {
                PUT(qx, qy);
                placed = 1;
                break;
            }
        }
        if (!placed) active[ai] = active[--na];
    }
#undef PUT
#undef CELLX
#undef CELLY
    free(grid);
    free(active);
    return overflow ? -1 : n;
}

/* node adjacency (unique neighbours, ascending) from triangles */
static int build_adjacency(int N, int E, const int *tri, int **ptr_out, int **adj_out)
{
    int *cnt = (int *)calloc((size_t)N + 1, sizeof(int));
    int *adj = (int *)malloc((size_t)E * 6 * sizeof(int));
    if (!cnt || !adj) {
        free(cnt);
        free(adj);
        return -1;
    }
    for (int e = 0; e < E; e++)
        for (int a = 0; a < 3; a++) {
            const int n = tri[3 * e + a];
            if (n < 0 || n >= N) {
                free(cnt);
                free(adj);
                return -1;
            }
            cnt[n + 1] += 2;
        }
    for (int n = 0; n < N; n++) cnt[n + 1] += cnt[n];
    int *fill = (int *)malloc((size_t)N * sizeof(int));
    if (!fill) {
        free(cnt);
        free(adj);
        return -1;
    }
    memcpy(fill, cnt, (size_t)N * sizeof(int));
    for (int e = 0; e < E; e++)
        for (int a = 0; a < 3; a++) {
            const int n = tri[3 * e + a];
            adj[fill[n]++] = tri[3 * e + (a + 1) % 3];
            adj[fill[n]++] = tri[3 * e + (a + 2) % 3];
        }
    /* sort and unique each row in place (rows are short: insertion sort) */
    int w = 0;
    for (int n = 0; n < N; n++) {
        const int b = cnt[n], e = cnt[n + 1];
        for (int i = b + 1; i < e; i++) {
            int v = adj[i], j = i - 1;
            while (j >= b && adj[j] > v) {
                adj[j + 1] = adj[j];
                j--;
            }
            adj[j + 1] = v;
        }
        const int start = w;
        for (int i = b; i < e; i++)
            if (i == b || adj[i] != adj[i - 1]) adj[w++] = adj[i];
        fill[n] = w - start;
    }
    cnt[0] = 0;
    for (int n = 0; n < N; n++) cnt[n + 1] = cnt[n] + fill[n];
    free(fill);
    *ptr_out = cnt;
    *adj_out = adj;
    return 0;
}

/* breadth-first levels from root over unvisited nodes (mark < 0); returns the number of levels and
 * leaves the visit order in order[0..count) */
static int bfs_levels(int root, const int *ptr, const int *adj, const int *mark, int *lev, int *order, int *count)
{
    int head = 0, tail = 0, depth = 0;
    order[tail++] = root;
    lev[root] = 0;
    while (head < tail) {
        const int v = order[head++];
        if (lev[v] + 1 > depth) depth = lev[v] + 1;
        for (int k = ptr[v]; k < ptr[v + 1]; k++) {
            const int u = adj[k];
            if (mark[u] < 0 && lev[u] < 0) {
                lev[u] = lev[v] + 1;
                order[tail++] = u;
            }
        }
    }
    *count = tail;
    return depth;
}

int adhk_rcm_order(int N, int E, const int *tri, int *perm)
{
    if (N <= 0 || E < 0 || !tri || !perm) return -1;
    int *ptr = NULL, *adj = NULL;
    if (build_adjacency(N, E, tri, &ptr, &adj)) return -1;
    int *mark = (int *)malloc((size_t)N * sizeof(int));   /* position in the CM order, -1 = not yet */
    int *lev = (int *)malloc((size_t)N * sizeof(int));
    int *order = (int *)malloc((size_t)N * sizeof(int));
    int *nb = (int *)malloc((size_t)N * sizeof(int));
    if (!mark || !lev || !order || !nb) {
        free(ptr);
        free(adj);
        free(mark);
        free(lev);
        free(order);
        free(nb);
        return -1;
    }
    for (int n = 0; n < N; n++) mark[n] = -1;
    int placed = 0;
    while (placed < N) {
        /* component root: the unplaced node of least degree (lowest index on ties) */
        int root = -1;
        for (int n = 0; n < N; n++)
            if (mark[n] < 0 && (root < 0 || ptr[n + 1] - ptr[n] < ptr[root + 1] - ptr[root])) root = n;
        /* pseudo-peripheral node: repeat BFS from a least-degree node of the last level while the
         * eccentricity grows */
        int cnt, depth;
        for (int n = 0; n < N; n++) lev[n] = -1;
        depth = bfs_levels(root, ptr, adj, mark, lev, order, &cnt);
        for (int it = 0; it < 16; it++) {
            int cand = -1;
            for (int i = 0; i < cnt; i++) {
                const int v = order[i];
                if (lev[v] == depth - 1 && (cand < 0 || ptr[v + 1] - ptr[v] < ptr[cand + 1] - ptr[cand])) cand = v;
            }
            for (int i = 0; i < cnt; i++) lev[order[i]] = -1;
            int cnt2;
            const int d2 = bfs_levels(cand, ptr, adj, mark, lev, order, &cnt2);
            if (d2 <= depth) {
                for (int i = 0; i < cnt2; i++) lev[order[i]] = -1;
                depth = bfs_levels(root, ptr, adj, mark, lev, order, &cnt);
                break;
            }
            root = cand;
            depth = d2;
            cnt = cnt2;
        }
        for (int i = 0; i < cnt; i++) lev[order[i]] = -1;
        /* Cuthill-McKee: breadth-first, neighbours in ascending degree (then index) */
        int head = placed, tail = placed;
        perm[tail++] = root;
        mark[root] = 0;
        while (head < tail) {
            const int v = perm[head++];
            int m = 0;
            for (int k = ptr[v]; k < ptr[v + 1]; k++)
                if (mark[adj[k]] < 0) {
                    nb[m++] = adj[k];
                    mark[adj[k]] = 0;
                }
            for (int i = 1; i < m; i++) {
                const int u = nb[i];
                const int du = ptr[u + 1] - ptr[u];
                int j = i - 1;
                while (j >= 0 && (ptr[nb[j] + 1] - ptr[nb[j]] > du ||
                                  (ptr[nb[j] + 1] - ptr[nb[j]] == du && nb[j] > u))) {
                    nb[j + 1] = nb[j];
                    j--;
                }
                nb[j + 1] = u;
            }
            for (int i = 0; i < m; i++) perm[tail++] = nb[i];
        }
        /* reverse this component's block */
        for (int i = placed, j = tail - 1; i < j; i++, j--) {
            const int t = perm[i];
            perm[i] = perm[j];
            perm[j] = t;
        }
        placed = tail;
    }
    free(ptr);
    free(adj);
    free(mark);
    free(lev);
    free(order);
    free(nb);
    return 0;
}

/* channel section weight at (x, y): 1 inside the channel, 0 on the floodplain, linear across banks */
static double channel_weight(const adhk_channel_spec *s, double x, double y)
{
    double yc = 0.5 * s->width;
    if (s->meander_amp != 0.0 && s->meander_wavelength > 0.0)
        yc += s->meander_amp * sin(TWO_PI * (x - s->meander_phase) / s->meander_wavelength);
    const double d = fabs(y - yc);
    if (s->bank_width <= 0.0) return d <= 0.5 * s->chan_width ? 1.0 : 0.0;
    double w = (0.5 * s->chan_width + 0.5 * s->bank_width - d) / s->bank_width;
    return w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
}

int adhk_channel_bed(const adhk_channel_spec *s, int N, const double *x, const double *y, double *z)
{
    if (!s || N < 0 || !x || !y || !z) return -1;
    const int nb = s->nbumps > 0 ? s->nbumps : 0;
    double *bcx = (double *)malloc((size_t)(4 * nb + 4) * sizeof(double));
    if (!bcx) return -1;
    double *bcy = bcx + nb, *ba = bcy + nb, *bs = ba + nb;
    uint64_t st = (uint64_t)s->seed + 0x5BD1E995ULL;
    for (int k = 0; k < nb; k++) {
        bcx[k] = s->length * unif(&st);
        bcy[k] = s->width * unif(&st);
        ba[k] = s->bump_amp_max * unif(&st);
        bs[k] = s->bump_sigma_min + (s->bump_sigma_max - s->bump_sigma_min) * unif(&st);
    }
    for (int n = 0; n < N; n++) {
        double zz = -s->slope * x[n] - s->chan_depth * channel_weight(s, x[n], y[n]);
        for (int k = 0; k < nb; k++) {
            const double rx = x[n] - bcx[k], ry = y[n] - bcy[k], s2 = 2.0 * bs[k] * bs[k];
            const double q = (rx * rx + ry * ry) / s2;
            if (q < 40.0) zz += ba[k] * exp(-q);
        }
        z[n] = zz;
    }
    free(bcx);
    return 0;
}

int adhk_channel_roughness(const adhk_channel_spec *s, int N, const double *x, const double *y, int E,
                           const int *tri, double *mann)
{
    if (!s || !x || !y || !tri || !mann || E < 0) return -1;
    for (int e = 0; e < E; e++) {
        double cx = 0, cy = 0;
        for (int a = 0; a < 3; a++) {
            const int n = tri[3 * e + a];
            if (n < 0 || n >= N) return -1;
            cx += x[n];
            cy += y[n];
        }
        mann[e] = channel_weight(s, cx / 3.0, cy / 3.0) >= 0.5 ? s->n_channel : s->n_floodplain;
    }
    return 0;
}

double adhk_channel_section_weight(const adhk_channel_spec *s, double x, double y)
{
    return s ? channel_weight(s, x, y) : -1.0;
}
