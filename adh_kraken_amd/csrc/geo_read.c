/*
 * geo_read.c -- the mesh cards of an AdH .geo file.
 *
 * Restates the card layout that sgrid_read reads (src/structs/sgrid/sgrid_read.c:84-130, :540-617):
 *   ND  id x y z [column]            node id (1-based) at (x, y, z)
 *   TRI id n1 n2 n3 [bflag [mat]]    triangle on nodes n1 n2 n3 (1-based)
 *   SEG id n1 n2 [bflag [mat]]       boundary segment on nodes n1 n2 (1-based)
 * Other cards (NODES, ELEMS_*, comments starting with '!' or '#', QUAD/TET/PRISM of the 3-D models)
 * are skipped here; QUAD, TET and PRISM elements are counted so a caller can refuse such a mesh.
 * The bflag field of each element is returned as written (-1 when absent); materials and boundary
 * strings come from the model's .physics / .cover files, as in smodel_super_read.c:143-159.
 * Nodes are placed by their id; elements keep the file order (the reference numbers them by id in
 * the same order).  Reading is single-pass per call; besides a line buffer it allocates a byte per
 * node, so that a repeated node id (which would leave another id unset) is refused.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/adhk.h"

#define LINE_MAX_LEN 4096

static int card_is(const char *line, const char *card)
{
    const size_t n = strlen(card);
    return strncmp(line, card, n) == 0 && (line[n] == ' ' || line[n] == '\t' || line[n] == '\n' || line[n] == '\r' ||
                                           line[n] == '\0');
}

static const char *skip_ws(const char *p)
{
    while (*p == ' ' || *p == '\t') p++;
    return p;
}

int adhk_geo_counts(const char *path, int *nnodes, int *ntri, int *nseg, int *nother)
{
    FILE *f = path ? fopen(path, "r") : NULL;
    if (!f) return -1;
    char line[LINE_MAX_LEN];
    int nd = 0, nt = 0, ns = 0, no = 0, maxid = 0;
    while (fgets(line, sizeof line, f)) {
        const char *p = skip_ws(line);
        if (card_is(p, "ND")) {
            int id;
            if (sscanf(p + 2, "%d", &id) == 1 && id > maxid) maxid = id;
            nd++;
        } else if (card_is(p, "TRI"))
            nt++;
        else if (card_is(p, "SEG"))
            ns++;
        else if (card_is(p, "QUAD") || card_is(p, "TET") || card_is(p, "PRISM"))
            no++;
    }
    fclose(f);
    if (maxid != nd) return -2; /* node ids must be 1..N */
    if (nnodes) *nnodes = nd;
    if (ntri) *ntri = nt;
    if (nseg) *nseg = ns;
    if (nother) *nother = no;
    return 0;
}

int adhk_geo_read(const char *path, int nnodes, double *x, double *y, double *z, int ntri, int *tri, int *tri_flag,
                  int nseg, int *seg, int *seg_flag)
{
    FILE *f = path ? fopen(path, "r") : NULL;
    if (!f || !x || !y || !z || (ntri > 0 && !tri) || (nseg > 0 && !seg)) {
        if (f) fclose(f);
        return -1;
    }
    unsigned char *seen = calloc(nnodes > 0 ? (size_t)nnodes : 1, 1);
    if (!seen) {
        fclose(f);
        return -1;
    }
    char line[LINE_MAX_LEN];
    int it = 0, is = 0, rc = 0, lineno = 0;
    while (fgets(line, sizeof line, f)) {
        lineno++;
        const char *p = skip_ws(line);
        if (card_is(p, "ND")) {
            int id;
            double xx, yy, zz;
            if (sscanf(p + 2, "%d %lf %lf %lf", &id, &xx, &yy, &zz) != 4 || id < 1 || id > nnodes || seen[id - 1]) {
                rc = -lineno - 2;
                break;
            }
            seen[id - 1] = 1;
            x[id - 1] = xx;
            y[id - 1] = yy;
            z[id - 1] = zz;
        } else if (card_is(p, "TRI")) {
            int id, n[3], bflag = -1, mat = -1;
            const int k = sscanf(p + 3, "%d %d %d %d %d %d", &id, &n[0], &n[1], &n[2], &bflag, &mat);
            if (k < 4 || it >= ntri) {
                rc = -lineno - 2;
                break;
            }
            for (int a = 0; a < 3; a++) {
                if (n[a] < 1 || n[a] > nnodes) {
                    rc = -lineno - 2;
                    break;
                }
                tri[3 * it + a] = n[a] - 1;
            }
            if (rc) break;
            if (tri_flag) tri_flag[it] = k >= 5 ? bflag : -1;
            (void)mat;
            it++;
        } else if (card_is(p, "SEG")) {
            int id, n[2], bflag = -1, mat = -1;
            const int k = sscanf(p + 3, "%d %d %d %d %d", &id, &n[0], &n[1], &bflag, &mat);
            if (k < 3 || is >= nseg || n[0] < 1 || n[0] > nnodes || n[1] < 1 || n[1] > nnodes) {
                rc = -lineno - 2;
                break;
            }
            seg[2 * is] = n[0] - 1;
            seg[2 * is + 1] = n[1] - 1;
            if (seg_flag) seg_flag[is] = k >= 4 ? bflag : -1;
            (void)mat;
            is++;
        }
    }
    fclose(f);
    free(seen);
    if (rc) return rc;
    return (it == ntri && is == nseg) ? 0 : -1;
}
