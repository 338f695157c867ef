/*
 * series.c -- time-series boundary values: piecewise-linear (time, value) tables whose average over
 * each time step drives the values of the boundary segments bound to them.
 *
 * Restates the reference's series evaluation: sseries_set_values_average (sseries.c:622-668) sets a
 * series to its average over (t1, t2) -- the area under the piecewise-linear interpolant divided by
 * t2 - t1 -- and advance_time calls it over the step's interval [t_prev - dt, t_prev] after
 * update_dt has moved t_prev to the new level (time_loop.c:74-76, :108-120).  sseries_get_interval
 * (sseries.c:746-790) rejects a time outside the series' range widened by the relative SMALL
 * (= DBL_EPSILON, define.h:31); here such a time makes the step fail with a message.  The reference feeds series 0 to every
 * depth segment (sw2_bc_h.c:155-157); here each segment is bound to its own series and scale, and a
 * segment bound to none keeps the value it was created with.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "adhk_internal.h"

#define SER_SMALL 2.220446049250313080847e-16 /* relative range tolerance of sseries_get_interval */

#define HC(x)                                                                                                     \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) {                                                                                   \
            adhk_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_));                        \
            goto fail;                                                                                            \
        }                                                                                                         \
    } while (0)

struct adhk_series_tab {
    int n;
    double *t, *v;
};

static double ser_eval(const adhk_series_tab *s, int i, double t)
{
    /* linear interpolation on piece i (tc_eval_series); one point: constant */
    if (s->n == 1) return s->v[0];
    const double w = (t - s->t[i]) / (s->t[i + 1] - s->t[i]);
    return s->v[i] + w * (s->v[i + 1] - s->v[i]);
}

static int ser_piece(const adhk_series_tab *s, double t)
{
    int i = 0;
    while (i + 2 < s->n && t > s->t[i + 1]) i++;
    return i;
}

/* average over [t1, t2] (t2 > t1), or the value at t1 when t2 == t1; -1 when outside the range
 * (a one-entry series is a constant at every time) */
static int ser_average(const adhk_series_tab *s, double t1, double t2, double *out)
{
    if (t2 < t1) return -1;
    if (s->n == 1) { /* one entry: a constant at every time */
        *out = s->v[0];
        return 0;
    }
    const double lo = s->t[0] - SER_SMALL * fmax(fabs(s->t[0]), 1.0);
    const double hi = s->t[s->n - 1] + SER_SMALL * fmax(fabs(s->t[s->n - 1]), 1.0);
    if (t1 < lo || t2 > hi) return -1;
    if (t2 == t1) {
        *out = ser_eval(s, ser_piece(s, t1), t1);
        return 0;
    }
    double area = 0.0;
    for (int i = 0; i + 1 < s->n; i++) {
        const double a = fmax(t1, s->t[i]), b = fmin(t2, s->t[i + 1]);
        if (b > a) area += 0.5 * (b - a) * (ser_eval(s, i, a) + ser_eval(s, i, b));
    }
    /* the parts of [t1, t2] that the tolerance lets fall outside the table take its end values */
    if (t1 < s->t[0]) area += (fmin(t2, s->t[0]) - t1) * s->v[0];
    if (t2 > s->t[s->n - 1]) area += (t2 - fmax(t1, s->t[s->n - 1])) * s->v[s->n - 1];
    *out = area / (t2 - t1);
    return 0;
}

int adhk_sw2_set_series(adhk_sw2 *c, int id, int npts, const double *t, const double *v)
{
    if (!c || id < 0 || npts < 1 || !t || !v) {
        adhk_set_error("set_series: bad arguments");
        return -1;
    }
    for (int i = 0; i + 1 < npts; i++)
        if (!(t[i + 1] > t[i])) {
            adhk_set_error("set_series: times must increase strictly (entry %d)", i + 1);
            return -1;
        }
    if (id >= c->nseries) {
        adhk_series_tab *ns = (adhk_series_tab *)realloc(c->series, (size_t)(id + 1) * sizeof *ns);
        if (!ns) return -1;
        memset(ns + c->nseries, 0, (size_t)(id + 1 - c->nseries) * sizeof *ns);
        c->series = ns;
        c->nseries = id + 1;
    }
    adhk_series_tab *s = &c->series[id];
    double *nt = (double *)malloc((size_t)npts * sizeof(double)), *nv = (double *)malloc((size_t)npts * sizeof(double));
    if (!nt || !nv) {
        free(nt);
        free(nv);
        return -1;
    }
    memcpy(nt, t, (size_t)npts * sizeof(double));
    memcpy(nv, v, (size_t)npts * sizeof(double));
    free(s->t);
    free(s->v);
    s->t = nt;
    s->v = nv;
    s->n = npts;
    return 0;
}

int adhk_sw2_bind_series(adhk_sw2 *c, int nsegs, const int *segs, int component, int id, double scale)
{
    if (!c || nsegs < 0 || (nsegs > 0 && !segs) || component < 0 || component > 1 || id < -1 || id >= c->nseries ||
        (id >= 0 && c->series[id].n < 1)) {
        adhk_set_error("bind_series: bad arguments (series %d of %d defined)", id, c ? c->nseries : 0);
        return -1;
    }
    if (!c->h_seg_ser) {
        const int S = c->S;
        c->h_seg_ser = (int *)malloc((size_t)(S > 0 ? 2 * S : 1) * sizeof(int));
        c->h_seg_scale = (double *)malloc((size_t)(S > 0 ? 2 * S : 1) * sizeof(double));
        c->h_seg_val = (double *)malloc((size_t)(S > 0 ? 2 * S : 1) * sizeof(double));
        if (!c->h_seg_ser || !c->h_seg_scale || !c->h_seg_val) return -1;
        for (int i = 0; i < 2 * S; i++) {
            c->h_seg_ser[i] = -1;
            c->h_seg_scale[i] = 0.0;
        }
        if (S > 0) HC(hipMemcpy(c->h_seg_val, c->seg_val, (size_t)S * 2 * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < nsegs; i++) {
        if (segs[i] < 0 || segs[i] >= c->S) {
            adhk_set_error("bind_series: segment %d out of range [0, %d)", segs[i], c->S);
            return -1;
        }
        c->h_seg_ser[2 * segs[i] + component] = id;
        c->h_seg_scale[2 * segs[i] + component] = scale;
    }
    return 0;
fail:
    return -1;
}

int adhk_sw2_set_time(adhk_sw2 *c, double t)
{
    if (!c) return -1;
    c->t_prev = t;
    return 0;
}

double adhk_sw2_get_time(const adhk_sw2 *c) { return c ? c->t_prev : NAN; }

/* set every bound segment to scale * (its series' average over [t1, t2]) on the device */
int adhk_series_apply(adhk_ctx *c, double t1, double t2)
{
    if (!c->h_seg_ser) return 0;
    double *avg = (double *)malloc((size_t)(c->nseries > 0 ? c->nseries : 1) * sizeof(double));
    char *need = (char *)calloc((size_t)(c->nseries > 0 ? c->nseries : 1), 1);
    if (!avg || !need) {
        free(avg);
        free(need);
        return -1;
    }
    for (int s = 0; s < 2 * c->S; s++)
        if (c->h_seg_ser[s] >= 0) need[c->h_seg_ser[s]] = 1;
    for (int k = 0; k < c->nseries; k++)
        if (need[k] && ser_average(&c->series[k], t1, t2, &avg[k])) {
            adhk_set_error("series %d: interval [%g, %g] outside its times [%g, %g]", k, t1, t2, c->series[k].t[0],
                           c->series[k].t[c->series[k].n - 1]);
            free(avg);
            free(need);
            return -1;
        }
    for (int s = 0; s < 2 * c->S; s++)
        if (c->h_seg_ser[s] >= 0) c->h_seg_val[s] = c->h_seg_scale[s] * avg[c->h_seg_ser[s]];
    free(avg);
    free(need);
    HC(hipMemcpyAsync(c->seg_val, c->h_seg_val, (size_t)c->S * 2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HC(hipStreamSynchronize(c->stream));
    return 0;
fail:
    return -1;
}

int adhk_series_average(int npts, const double *t, const double *v, double t1, double t2, double *value)
{
    if (npts < 1 || !t || !v || !value) return -1;
    for (int i = 0; i + 1 < npts; i++)
        if (!(t[i + 1] > t[i])) return -1;
    adhk_series_tab s = {npts, (double *)t, (double *)v};
    if (ser_average(&s, t1, t2, value)) {
        adhk_set_error("series: interval [%g, %g] outside its times [%g, %g]", t1, t2, t[0], t[npts - 1]);
        return -1;
    }
    return 0;
}

void adhk_series_free(adhk_ctx *c)
{
    for (int k = 0; k < c->nseries; k++) {
        free(c->series[k].t);
        free(c->series[k].v);
    }
    free(c->series);
    free(c->h_seg_ser);
    free(c->h_seg_scale);
    free(c->h_seg_val);
    c->series = NULL;
    c->h_seg_ser = NULL;
    c->h_seg_scale = NULL;
    c->h_seg_val = NULL;
    c->nseries = 0;
}
