/*
 * sw2_wetdry.h -- general SW2 triangle residual with the wet/dry sub-element
 * integration, for elements that are partially dry (some nodal depths <= 0).
 *
 * fe_sw2_wet_dry_wrapper (src/fe/fe_wet_dry_wrapper.c:104-304): with one dry
 * node the integrand is taken over the whole triangle and the dry corner
 * sub-triangle is subtracted; with two dry nodes only the wet corner
 * sub-triangle is integrated.  The sub-triangle's vertices sit where the
 * linear depth reaches zero along each edge (FIXED_POS, macro.h:756), its
 * gradients come from GRAD_SHAPE (macro.h:778), and contributions are moved
 * back to the original nodes by RE_DISTRIBUTE (macro.h:788) when requested.
 * fe_sw2_wet_dry_factor (:29-71) returns the wet area fraction.
 *
 * These elements are rare (the wet/dry front), so this path favours a direct
 * statement over speed; the all-wet and all-dry elements take sw2_tri_resid.
 */
#ifndef ADHK_SW2_WETDRY_H
#define ADHK_SW2_WETDRY_H

#include "sw2_device.h"

struct V2d {
    double x, y;
};
struct V3d {
    double x, y, z;
};

enum { WD_CONV = 0, WD_CONT_TEMPORAL, WD_PRESSURE, WD_BODYFORCE, WD_BPRESSURE, WD_INTF, WD_GLS };

/* the wrapped integrands (src/fe/fe_wet_dry_integrations.c) */
__device__ static void wd_integrand(int fn, const V3d *nds, const double *h, const V2d *gp, const V2d *v,
                                    const V2d *vwd, const double *fwd, double djac, const double *vars, double *rhs)
{
    if (fn == WD_CONV) { /* :94-128 */
        double c = vars[1] * vars[0];
        double uu[3] = {vwd[0].x, vwd[1].x, vwd[2].x}, vv[3] = {vwd[0].y, vwd[1].y, vwd[2].y};
        double p1 = (2 * h[0] + h[1] + h[2]), p2 = (h[0] + 2 * h[1] + h[2]), p3 = (h[0] + h[1] + 2 * h[2]);
        double tx = (p1 * uu[0] + p2 * uu[1] + p3 * uu[2]), ty = (p1 * vv[0] + p2 * vv[1] + p3 * vv[2]);
        double con = djac * D_ONE_12 * c;
        double ff1 = (3 * uu[0] + uu[1] + uu[2]), ff2 = (2 * uu[0] + 2 * uu[1] + uu[2]);
        double ff3 = (2 * uu[0] + uu[1] + 2 * uu[2]), ff4 = (uu[0] + 3 * uu[1] + uu[2]);
        double ff5 = (uu[0] + 2 * uu[1] + 2 * uu[2]), ff6 = (uu[0] + uu[1] + 3 * uu[2]);
        double gg1 = (2 * ff1 * vv[0] + ff2 * vv[1] + ff3 * vv[2]);
        double gg2 = (ff2 * vv[0] + 2 * ff4 * vv[1] + ff5 * vv[2]);
        double gg3 = (ff3 * vv[0] + ff5 * vv[1] + 2 * ff6 * vv[2]);
        double cgh = djac * D_ONE_60 * c * (gg1 * h[0] + gg2 * h[1] + gg3 * h[2]);
        double q[3], s[3];
        {
            double a0 = uu[0] * uu[0], a1 = uu[1] * uu[1], a2 = uu[2] * uu[2];
            q[0] = (3 * a0 + 2 * uu[0] * uu[1] + a1 + (2 * uu[0] + uu[1]) * uu[2] + a2);
            q[1] = (a0 + 2 * uu[0] * uu[1] + 3 * a1 + (uu[0] + 2 * uu[1]) * uu[2] + a2);
            q[2] = (a0 + uu[0] * uu[1] + a1 + 2 * (uu[0] + uu[1]) * uu[2] + 3 * a2);
            double b0 = vv[0] * vv[0], b1 = vv[1] * vv[1], b2 = vv[2] * vv[2];
            s[0] = (3 * b0 + 2 * vv[0] * vv[1] + b1 + (2 * vv[0] + vv[1]) * vv[2] + b2);
            s[1] = (b0 + 2 * vv[0] * vv[1] + 3 * b1 + (vv[0] + 2 * vv[1]) * vv[2] + b2);
            s[2] = (b0 + vv[0] * vv[1] + b1 + 2 * (vv[0] + vv[1]) * vv[2] + 3 * b2);
        }
        double cu = djac * D_ONE_30 * c * (q[0] * h[0] + q[1] * h[1] + q[2] * h[2]);
        double cv = djac * D_ONE_30 * c * (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]);
        for (int i = 0; i < 3; i++) {
            double rx = cgh * gp[i].y;
            rx += cu * gp[i].x;
            double ry = cgh * gp[i].x;
            ry += cv * gp[i].y;
            rhs[i * 3] += con * (gp[i].x * tx + gp[i].y * ty);
            rhs[i * 3 + 1] += rx;
            rhs[i * 3 + 2] += ry;
        }
    } else if (fn == WD_CONT_TEMPORAL) { /* :151-170 */
        double fs = h[0] + h[1] + h[2];
        double t1 = djac * vars[1] * D_ONE_12;
        for (int i = 0; i < 3; i++) rhs[i * 3] += t1 * (fs + h[i]);
    } else if (fn == WD_PRESSURE) { /* :189-205 */
        double iff = vars[1] * D_ONE_6 * djac *
                     (h[0] * h[0] + h[0] * h[1] + h[1] * h[1] + h[0] * h[2] + h[1] * h[2] + h[2] * h[2]);
        for (int i = 0; i < 3; i++) {
            rhs[i * 3 + 1] += gp[i].x * iff;
            rhs[i * 3 + 2] += gp[i].y * iff;
        }
    } else if (fn == WD_BODYFORCE) { /* :227-243 */
        double sh = h[0] + h[1] + h[2];
        double cx = djac * D_ONE_12 * vars[1] * v->x, cy = djac * D_ONE_12 * vars[1] * v->y;
        for (int i = 0; i < 3; i++) {
            rhs[i * 3 + 1] += cx * (sh + h[i]);
            rhs[i * 3 + 2] += cy * (sh + h[i]);
        }
    } else if (fn == WD_BPRESSURE) { /* :265-288 */
        double t1 = 1. * D_ONE_12 * vars[1];
        for (int e = 0; e < 3; e++) {
            int a = e, b = (e + 1) % 3;
            double delx = nds[a].x - nds[b].x, dely = nds[b].y - nds[a].y;
            double ha = h[a], hb = h[b];
            double ra = t1 * (3 * ha * ha + 2 * ha * hb + hb * hb);
            double rb = t1 * (ha * ha + 2 * ha * hb + 3 * hb * hb);
            rhs[a * 3 + 1] += dely * ra;
            rhs[b * 3 + 1] += dely * rb;
            rhs[a * 3 + 2] += delx * ra;
            rhs[b * 3 + 2] += delx * rb;
        }
    } else if (fn == WD_INTF) { /* :437-454 */
        double t1 = (h[0] + h[1] + h[2]) / 3.0, t2 = 0.0, t3 = 0.0;
        if (vwd) {
            t2 = (vwd[0].x + vwd[1].x + vwd[2].x) / 3.0;
            t3 = (vwd[0].y + vwd[1].y + vwd[2].y) / 3.0;
        }
        rhs[0] += djac * t1;
        rhs[1] += djac * t2;
        rhs[2] += djac * t3;
    } else { /* WD_GLS :700-722 */
        double dt = vars[0], c = vars[1];
        double h_avg = D_ONE_3 * (fwd[0] + fwd[1] + fwd[2]);
        double u_avg = D_ONE_3 * (vwd[0].x + vwd[1].x + vwd[2].x);
        double v_avg = D_ONE_3 * (vwd[0].y + vwd[1].y + vwd[2].y);
        double dhdx = gp[0].x * fwd[0] + gp[1].x * fwd[1] + gp[2].x * fwd[2];
        double dhdy = gp[0].y * fwd[0] + gp[1].y * fwd[1] + gp[2].y * fwd[2];
        double dudx = gp[0].x * vwd[0].x + gp[1].x * vwd[1].x + gp[2].x * vwd[2].x;
        double dudy = gp[0].y * vwd[0].x + gp[1].y * vwd[1].x + gp[2].y * vwd[2].x;
        double dvdx = gp[0].x * vwd[0].y + gp[1].x * vwd[1].y + gp[2].x * vwd[2].y;
        double dvdy = gp[0].y * vwd[0].y + gp[1].y * vwd[1].y + gp[2].y * vwd[2].y;
        rhs[0] += djac * dt * c * (u_avg * dhdx + h_avg * dudx + v_avg * dhdy + h_avg * dvdy);
        rhs[1] += djac * dt * c * (u_avg * dudx + v_avg * dudy);
        rhs[2] += djac * dt * c * (u_avg * dvdx + v_avg * dvdy);
    }
}

__device__ static double wd_tri_area(V3d a, V3d b, V3d c) /* TRI_AREA, macro.h:771 */
{
    double v0x = b.x - a.x, v0y = b.y - a.y, v1x = c.x - a.x, v1y = c.y - a.y;
    return (v0x * v1y - v0y * v1x) / 2.0;
}

/* fe_sw2_wet_dry_factor, fe_wet_dry_wrapper.c:29-71 */
__device__ static double wd_factor(const V3d *x, const double *h, double djac)
{
    int total = 0, pass = 0;
    for (int i = 0; i < 3; i++) {
        int d = (h[i] <= 0.0) ? 1 : 0;
        total += d;
        pass += (d << i);
    }
    if (total == 3) return 0.0;
    if (total == 0) return 1.0;
    if (total != 1) pass = 7 - pass;
    int only = pass / 2;
    V3d nx[3];
    for (int i = 0; i < 3; i++) {
        if (i != only) {
            double xsi = h[i] / (h[i] - h[only]);
            nx[i].x = x[i].x + xsi * (x[only].x - x[i].x);
            nx[i].y = x[i].y + xsi * (x[only].y - x[i].y);
            nx[i].z = x[i].z;
        } else {
            nx[i] = x[i];
        }
    }
    double a = wd_tri_area(nx[0], nx[1], nx[2]);
    if (a / djac < D_SMALL) return (double)(total % 2);
    double factor = (total >> 1) - (total % 2);
    factor *= a / djac;
    return factor + ((factor < 0) ? 1.0 : 0.0);
}

/* fe_sw2_wet_dry_wrapper, fe_wet_dry_wrapper.c:104-304 */
__device__ static double wd_wrapper(double *rhs, const V3d *x, const double *h, const V2d *gp, const V2d *v,
                                    const V2d *vwd, const double *fwd, double djac, int redistribute,
                                    const double *vars, int fn)
{
    int total = 0, pass = 0;
    for (int i = 0; i < 3; i++) {
        int d = (h[i] <= 0.0) ? 1 : 0;
        total += d;
        pass += (d << i);
    }
    if (total == 3) return 0;
    if (total == 0) {
        wd_integrand(fn, x, h, gp, v, vwd, fwd, djac, vars, rhs);
        return 1.0;
    }
    if (total == 1)
        wd_integrand(fn, x, h, gp, v, vwd, fwd, djac, vars, rhs);
    else
        pass = 7 - pass;
    int only = pass / 2;
    double nh[3] = {0.0, 0.0, 0.0}, nf[3] = {0.0, 0.0, 0.0};
    V3d nx[3];
    V2d nv[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int i = 0; i < 3; i++) {
        if (i != only) {
            double xsi = h[i] / (h[i] - h[only]);
            nx[i].x = x[i].x + xsi * (x[only].x - x[i].x);
            nx[i].y = x[i].y + xsi * (x[only].y - x[i].y);
            nx[i].z = x[i].z;
            if (vwd) {
                nv[i].x = vwd[i].x + xsi * (vwd[only].x - vwd[i].x);
                nv[i].y = vwd[i].y + xsi * (vwd[only].y - vwd[i].y);
            }
            if (fwd) nf[i] = fwd[i] + xsi * (fwd[only] - fwd[i]);
        } else {
            nx[i] = x[i];
            nh[i] = h[i];
            if (fwd) nf[i] = fwd[i];
            if (vwd) nv[i] = vwd[i];
        }
    }
    double elem_djac = djac;
    djac = wd_tri_area(nx[0], nx[1], nx[2]);
    if (djac / elem_djac < D_SMALL) return (double)(total % 2);
    V2d ngp[3];
    if (gp) { /* GRAD_SHAPE, macro.h:778 */
        ngp[0].x = 0.5 * (nx[1].y - nx[2].y) / djac;
        ngp[1].x = 0.5 * (nx[2].y - nx[0].y) / djac;
        ngp[2].x = 0.5 * (nx[0].y - nx[1].y) / djac;
        ngp[0].y = 0.5 * (nx[2].x - nx[1].x) / djac;
        ngp[1].y = 0.5 * (nx[0].x - nx[2].x) / djac;
        ngp[2].y = 0.5 * (nx[1].x - nx[0].x) / djac;
    }
    double nr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    wd_integrand(fn, nx, nh, gp ? ngp : (const V2d *)0, v, vwd ? nv : (const V2d *)0, fwd ? nf : (const double *)0,
                 djac, vars, nr);
    double factor = (total >> 1) - (total % 2);
    if (redistribute) { /* RE_DISTRIBUTE, macro.h:788 */
        for (int i = 0; i < 3; i++) {
            if (i == only) continue;
            double a0x = nx[i].x - x[only].x, a0y = nx[i].y - x[only].y;
            double a1x = x[i].x - x[only].x, a1y = x[i].y - x[only].y;
            double l0 = a0x * a0x + a0y * a0y, l1 = a1x * a1x + a1y * a1y;
            double xsi = sqrt(l0 / l1);
            for (int k = 0; k < 3; k++) {
                nr[only * 3 + k] += (1.0 - xsi) * nr[i * 3 + k];
                nr[i * 3 + k] *= xsi;
            }
        }
    }
    for (int i = 0; i < 3; i++) {
        rhs[i * 3] += factor * nr[i * 3];
        rhs[i * 3 + 1] += factor * nr[i * 3 + 1];
        rhs[i * 3 + 2] += factor * nr[i * 3 + 2];
    }
    factor *= djac / elem_djac;
    return factor + ((factor < 0) ? 1.0 : 0.0);
}

/* fe_sw2_temporal, sw2_body_resid.c:1195-1251 */
__device__ static void wd_temporal(const V3d *nds, double djac, const double *h, const V2d *vel, double wd_factor,
                                   double dt_factor, double *er, int wd_flag)
{
    double con = dt_factor * wd_factor;
    double rhs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double vars[2] = {1.0, dt_factor};
    wd_wrapper(rhs, nds, h, (const V2d *)0, (const V2d *)0, (const V2d *)0, (const double *)0, djac, 1, vars,
               WD_CONT_TEMPORAL);
    double tl = djac * con * D_ONE_3;
    for (int i = 0; i < 3; i++) {
        double hm = h[i] > D_DLL ? h[i] : D_DLL;
        if (wd_flag != 2) er[i * 3] += rhs[i * 3];
        er[i * 3 + 1] += tl * (vel[i].x * hm);
        er[i * 3 + 2] += tl * (vel[i].y * hm);
    }
}

/* fe_sw2_body_resid (sw2_body_resid.c:44-1176), every wet/dry case */
__device__ static void sw2_tri_resid_general(const TriIn &in, const Sw2Par &P, int pnode, int pvar, double pe,
                                             double r[9])
{
    const double djac = in.djac, dt = P.dt, g = P.g, alpha = P.alpha;
    V3d nds[3];
    V2d gp[3], vel[3], velo[3], veloo[3];
    double h[3], ho[3], hoo[3];
    for (int a = 0; a < 3; a++) {
        nds[a].x = in.x[a];
        nds[a].y = in.y[a];
        nds[a].z = in.z[a];
        gp[a].x = in.gx[a];
        gp[a].y = in.gy[a];
        h[a] = in.h[a] + ((pvar == 0 && pnode == a) ? pe : 0.0);
        vel[a].x = in.u[a] + ((pvar == 1 && pnode == a) ? pe : 0.0);
        vel[a].y = in.v[a] + ((pvar == 2 && pnode == a) ? pe : 0.0);
        ho[a] = in.ho[a];
        velo[a].x = in.uo[a];
        velo[a].y = in.vo[a];
        hoo[a] = in.hoo[a];
        veloo[a].x = in.uoo[a];
        veloo[a].y = in.voo[a];
    }
    const bool pflag = !(pvar == 1 || pvar == 2);
    V2d gz = {0., 0.}, gh = {0., 0.};
    for (int i = 0; i < 3; i++) {
        gz.x += in.z[i] * gp[i].x;
        gz.y += in.z[i] * gp[i].y;
        gh.x += h[i] * gp[i].x;
        gh.y += h[i] * gp[i].y;
    }
    /* wd averages (fe_sw2_wd_average, integrations.c:40) */
    double hwd = 0.;
    V2d vwd = {0., 0.};
    {
        double av[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double vars[2] = {1., 1.};
        double f = wd_wrapper(av, nds, h, (const V2d *)0, (const V2d *)0, vel, h, djac, 0, vars, WD_INTF);
        if (f > 0.0) {
            double wa = djac * f;
            hwd = av[0] / wa;
            vwd.x = av[1] / wa;
            vwd.y = av[2] / wa;
        }
    }
    double factor = wd_factor(nds, h, djac);
    double factor_old = wd_factor(nds, ho, djac);
    double factor_older = wd_factor(nds, hoo, djac);
    double vmag = sqrt(vwd.x * vwd.x + vwd.y * vwd.y);
    double rough;
    if (P.manning) {
        double hf = hwd > D_SMALL ? hwd : D_SMALL;
        rough = (in.ks > D_SMALL) ? drag_coef(hf, in.ks) : 0.0;
    } else {
        rough = P.cf;
    }
    for (int k = 0; k < 9; k++) r[k] = 0.0;
    double rhs[9];
    /* advection */
    for (int k = 0; k < 9; k++) rhs[k] = 0.0;
    {
        double vars[2] = {dt, -1};
        wd_wrapper(rhs, nds, h, gp, vel, vel, (const double *)0, djac, 1, vars, WD_CONV);
    }
    for (int k = 0; k < 9; k++) r[k] += rhs[k];
    /* temporal */
    wd_temporal(nds, djac, h, vel, factor, (1. + alpha / 2.), r, -3);
    wd_temporal(nds, djac, ho, velo, factor_old, (-1.) * (1. + alpha), r, in.wd_flag);
    wd_temporal(nds, djac, hoo, veloo, factor_older, (alpha / 2.), r, in.wd_flag);
    /* friction */
    const double resistance = rough / (2.0 * g);
    {
        double fx[3], fy[3];
        for (int i = 0; i < 3; i++) {
            double t1 = sqrt(vel[i].x * vel[i].x + vel[i].y * vel[i].y);
            fx[i] = vel[i].x * t1;
            fy[i] = vel[i].y * t1;
        }
        double t1 = dt * g * resistance * factor;
        double tt = djac * t1 * D_ONE_12;
        double sx = fx[0] + fx[1] + fx[2], sy = fy[0] + fy[1] + fy[2];
        for (int i = 0; i < 3; i++) {
            r[3 * i + 1] += tt * (sx + fx[i]);
            r[3 * i + 2] += tt * (sy + fy[i]);
        }
    }
    const double bsx = resistance * vwd.x * vmag, bsy = resistance * vwd.y * vmag;
    if (pflag) {
        for (int k = 0; k < 9; k++) rhs[k] = 0.0;
        {
            double vars[2] = {1, -D_ONE_2 * g * dt};
            wd_wrapper(rhs, nds, h, gp, (const V2d *)0, (const V2d *)0, (const double *)0, djac, 1, vars,
                       WD_PRESSURE);
        }
        for (int k = 0; k < 9; k++) r[k] += rhs[k];
        for (int k = 0; k < 9; k++) rhs[k] = 0.0;
        {
            double vars[2] = {1, g * dt};
            wd_wrapper(rhs, nds, h, (const V2d *)0, &gz, (const V2d *)0, (const double *)0, djac, 1, vars,
                       WD_BODYFORCE);
        }
        for (int k = 0; k < 9; k++) r[k] += rhs[k];
        for (int k = 0; k < 9; k++) rhs[k] = 0.0;
        {
            double vars[2] = {1, D_ONE_2 * dt * g};
            wd_wrapper(rhs, nds, h, (const V2d *)0, (const V2d *)0, (const V2d *)0, (const double *)0, djac, 1, vars,
                       WD_BPRESSURE);
        }
        for (int k = 0; k < 9; k++) r[k] += rhs[k];
    }
    /* SUPG */
    if (!(fabs(hwd) < D_SMALL)) {
        const double hden = hwd > D_NQS ? hwd : D_NQS;
        const double le = sqrt(djac * factor_old);
        double cel2 = g * hden;
        if (hden <= 0.0) cel2 = 0.;
        const double sr = sqrt(vwd.x * vwd.x + vwd.y * vwd.y + cel2 + D_SMALL);
        const double tau = 0.5 * le / sr;
        double dfdt[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, adv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double vars[2] = {1, 1};
        factor_old = wd_wrapper(dfdt, nds, ho, (const V2d *)0, (const V2d *)0, velo, ho, djac, 0, vars, WD_INTF);
        double dudt = -dfdt[1], dvdt = -dfdt[2], dhdt = -dfdt[0];
        for (int k = 0; k < 9; k++) dfdt[k] = 0.0;
        factor = wd_wrapper(dfdt, nds, h, (const V2d *)0, (const V2d *)0, vel, h, djac, 0, vars, WD_INTF);
        dudt += dfdt[1];
        dvdt += dfdt[2];
        dhdt += dfdt[0];
        vars[0] = dt;
        vars[1] = 1;
        factor = wd_wrapper(adv, nds, h, gp, (const V2d *)0, vel, h, djac, 0, vars, WD_GLS);
        double xp = g * (gh.x + gz.x) * dt * factor;
        double yp = g * (gh.y + gz.y) * dt * factor;
        double xf = g * bsx * dt * factor, yf = g * bsy * dt * factor;
        double c_eq = dhdt + adv[0];
        double x_eq = dudt + adv[1] + (xp + xf / hden) * djac;
        double y_eq = dvdt + adv[2] + (yp + yf / hden) * djac;
        for (int i = 0; i < 3; i++) {
            r[3 * i] += (gp[i].x * (tau * vwd.x * c_eq + tau * hwd * x_eq) + gp[i].y * (tau * vwd.y * c_eq + tau * hwd * y_eq));
            r[3 * i + 1] += (gp[i].x * (tau * g * hwd * c_eq + tau * vwd.x * hwd * x_eq) + gp[i].y * (tau * vwd.y * hwd * x_eq));
            r[3 * i + 2] += (gp[i].y * (tau * g * hwd * c_eq + tau * vwd.y * hwd * y_eq) + gp[i].x * (tau * vwd.x * hwd * y_eq));
        }
    }
}

#endif
