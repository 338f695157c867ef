/*
 * sw2_device.h -- gfx950 device functions for the SW2 elemental residuals.
 *
 * sw2_tri_resid   : fe_sw2_body_resid (src/models/sw2/sw2_body_resid.c:44-1176), triangle branch
 * sw2_seg_resid   : fe_sw2_bc_h (sw2_bc_h.c:57), explicit flow (fe_sw2_boundary_integrals.c:56),
 *                   fe_sw2_bc_vel (sw2_bc_vel.c:58)
 *
 * Each term keeps the reference's operand grouping so that results agree with
 * the CPU oracle to rounding (the compiler may fuse a multiply and an add).
 * Terms the reference evaluates to exact zeros for this configuration space
 * (eddy viscosity, Coriolis, shock capturing, wind/wave/rain, lid pressure,
 * baroclinic density) are not evaluated: adding +0.0 leaves a double unchanged.
 *
 * Wet/dry: each wet/dry-wrapped integral (fe_wet_dry_wrapper.c:104) has three
 * outcomes, all nodes wet (integrand on the full triangle), all dry (nothing),
 * or partially dry (integral over the wet sub-triangle).  The first two are
 * handled here; a partially dry field makes the function return 1.  Callers
 * route such elements to sw2_tri_resid_general (sw2_wetdry.h) instead.
 */
#ifndef ADHK_SW2_DEVICE_H
#define ADHK_SW2_DEVICE_H

#include <hip/hip_runtime.h>

#define D_SMALL 2.220446049250313080847e-16 /* DBL_EPSILON, define.h:31 */
#define D_NQS 1.1920928955078125e-07        /* FLT_EPSILON, define.h:32 */
#define D_ONE_2 0.5e+0
#define D_ONE_3 0.333333333333333333333333
#define D_ONE_6 0.166666666666666666666667
#define D_ONE_12 0.083333333333333333333333
#define D_ONE_30 0.033333333333333333333333
#define D_ONE_60 0.016666666666666666666667
#define D_DLL 1e-6 /* drying_lower_limit, sw2_body_resid.c:85 */

struct Sw2Par {
    double g, dt, alpha, cf;
    int manning;
};

/* gathered element data */
struct TriIn {
    double x[3], y[3], z[3];
    double h[3], u[3], v[3];
    double ho[3], uo[3], vo[3];
    double hoo[3], uoo[3], voo[3];
    double djac, gx[3], gy[3];
    double ks;
    int wd_flag;
};

__device__ __forceinline__ int n_dry3(double a, double b, double c)
{
    return (a <= 0.0 ? 1 : 0) + (b <= 0.0 ? 1 : 0) + (c <= 0.0 ? 1 : 0);
}

/* fr_bedshstr_drag_coef, src/friction/fr_bedshstr_drag_coef.c:18-43 */
__device__ __forceinline__ double drag_coef(double depth, double rh)
{
    depth = depth > 1.E-8 ? depth : 1.E-8;
    rh = rh > 1.E-8 ? rh : 1.E-8;
    double depth_rh = (rh > 29.7 * depth) ? 1. / 29.7 : depth / rh;
    double bta = 29.7 * depth_rh;
    double t = (.4 * bta) / ((bta + 1) * (log(bta + 1) - 1) + 1);
    return 2. * (t * t);
}

/* fe_sw2_temporal, sw2_body_resid.c:1195-1251, for a field that is all wet (ndry == 0) or all dry (3) */
__device__ __forceinline__ void sw2_temporal(const TriIn &in, const double *h, const double *u, const double *v,
                                             int ndry, double wd_factor, double dt_factor, int wd_flag,
                                             double *r)
{
    double con = dt_factor * wd_factor;
    double tl = in.djac * con * D_ONE_3; /* integrate_triangle_phi_f_lump, fe_integrals_2d.c:52 */
    double tc = in.djac * dt_factor * D_ONE_12;
    double hs = h[0] + h[1] + h[2];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double hm = h[i] > D_DLL ? h[i] : D_DLL;
        double uh = u[i] * hm, vh = v[i] * hm;
        if (ndry == 0 && wd_flag != 2) r[3 * i] += tc * (hs + h[i]); /* integrate_triangle_phi_f */
        r[3 * i + 1] += tl * uh;
        r[3 * i + 2] += tl * vh;
    }
}

/* Returns 1 when a partially dry field is met (r is then not the reference value).
 * FD = true is the form the FD Jacobian evaluates 18 times per element: divisions by 3 become
 * products with 1/3, the wet/dry averages skip the djac / djac that cancels, and 1 / (2 g) and
 * 1 / hden are formed once.  Each of these moves a term by at most a few ulps, about 1e-12 of a
 * Jacobian entry after the central difference; the residual itself (FD = false) keeps the
 * reference's operation order. */
/* Values of an all-wet element's FD residual that a perturbation of one variable leaves unchanged, formed
 * once per element node thread and reused by its +-eps evaluations (the same operations on the same
 * operands; the compiler's contraction of multiply-adds may round a shared value one ulp apart from an
 * inline evaluation, far below the central difference's own error):
 *   le        sqrt(djac factor_old)                                   -- no perturbation changes it
 *   rough0    drag coefficient, ihden0 = 1 / hden                     -- unchanged when u or v is perturbed
 *   vmag0     |(u_wd, v_wd)|, t1_0[a] = |(u_a, v_a)|                  -- unchanged when h is perturbed
 *                                                                        (t1_0 of the other nodes also when
 *                                                                        u or v of one node is) */
struct FdShare {
    double le, rough0, ihden0, vmag0, t1_0[3];
};

template <bool FD = false>
__device__ __forceinline__ int sw2_tri_resid(const TriIn &in, const Sw2Par &P, int pnode, int pvar, double pe,
                                             double r[9], const FdShare *sh = nullptr)
{
    /* perturbation, sw2_body_resid.c:107-120 (selects instead of a dynamic index keep the
     * element in registers; an unperturbed entry gets + 0.0, which leaves it unchanged) */
    double h[3], u[3], v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        h[a] = in.h[a] + ((pvar == 0 && pnode == a) ? pe : 0.0);
        u[a] = in.u[a] + ((pvar == 1 && pnode == a) ? pe : 0.0);
        v[a] = in.v[a] + ((pvar == 2 && pnode == a) ? pe : 0.0);
    }
    const bool pflag = !(pvar == 1 || pvar == 2);
    const int dh = n_dry3(h[0], h[1], h[2]);
    const int dho = n_dry3(in.ho[0], in.ho[1], in.ho[2]);
    if ((dh == 1 || dh == 2) || (dho == 1 || dho == 2)) return 1;
    int dhoo = 0;
    if (P.alpha != 0.0) {
        dhoo = n_dry3(in.hoo[0], in.hoo[1], in.hoo[2]);
        if (dhoo == 1 || dhoo == 2) return 1;
    }
    const double djac = in.djac, dt = P.dt, g = P.g;
    const double *gx = in.gx, *gy = in.gy;

    /* gradients, grad2d_phi_f fe_integrals_2d.c:139 */
    double gzx = 0., gzy = 0., ghx = 0., ghy = 0.;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        gzx += in.z[i] * gx[i];
        gzy += in.z[i] * gy[i];
        ghx += h[i] * gx[i];
        ghy += h[i] * gy[i];
    }
    /* wet/dry averages, fe_sw2_wd_average fe_wet_dry_integrations.c:40 */
    double hwd = 0., uwd = 0., vwd = 0., factor = 0.;
    if (dh == 0 && FD) {
        hwd = (h[0] + h[1] + h[2]) * D_ONE_3;
        uwd = (u[0] + u[1] + u[2]) * D_ONE_3;
        vwd = (v[0] + v[1] + v[2]) * D_ONE_3;
        factor = 1.0;
    } else if (dh == 0) {
        double a0 = djac * ((h[0] + h[1] + h[2]) / 3.0);
        double a1 = djac * ((u[0] + u[1] + u[2]) / 3.0);
        double a2 = djac * ((v[0] + v[1] + v[2]) / 3.0);
        double wa = djac * 1.0;
        hwd = a0 / wa;
        uwd = a1 / wa;
        vwd = a2 / wa;
        factor = 1.0;
    }
    const double factor_old = (dho == 0) ? 1.0 : 0.0;
    const double factor_older = (dhoo == 0) ? 1.0 : 0.0;
    const bool hsame = FD && sh && pvar != 0, vsame = FD && sh && pvar == 0; /* h / (u, v) unperturbed */
    const double vmag = vsame ? sh->vmag0 : sqrt(uwd * uwd + vwd * vwd);
    /* friction coefficient, sw2_body_resid.c:238-241 / sw2_get_roughness.c:40-45 */
    double rough;
    if (hsame) {
        rough = sh->rough0;
    } else if (P.manning) {
        double hf = hwd > D_SMALL ? hwd : D_SMALL;
        rough = (in.ks > D_SMALL) ? drag_coef(hf, in.ks) : 0.0;
    } else {
        rough = P.cf;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) r[k] = 0.0;

    /* advection, :413-434 with fe_sw2_wd_convection_triangle (integrations.c:94) */
    if (dh == 0) {
        const double c = (-1.0) * dt;
        double p1 = (2 * h[0] + h[1] + h[2]), p2 = (h[0] + 2 * h[1] + h[2]), p3 = (h[0] + h[1] + 2 * h[2]);
        double tx = (p1 * u[0] + p2 * u[1] + p3 * u[2]);
        double ty = (p1 * v[0] + p2 * v[1] + p3 * v[2]);
        double con = djac * D_ONE_12 * c;
        /* integrate_triangle_dphi_f_g_h (f = u, g = v) */
        double ff1 = (3 * u[0] + u[1] + u[2]), ff2 = (2 * u[0] + 2 * u[1] + u[2]), ff3 = (2 * u[0] + u[1] + 2 * u[2]);
        double ff4 = (u[0] + 3 * u[1] + u[2]), ff5 = (u[0] + 2 * u[1] + 2 * u[2]), ff6 = (u[0] + u[1] + 3 * u[2]);
        double gg1 = (2 * ff1 * v[0] + ff2 * v[1] + ff3 * v[2]);
        double gg2 = (ff2 * v[0] + 2 * ff4 * v[1] + ff5 * v[2]);
        double gg3 = (ff3 * v[0] + ff5 * v[1] + 2 * ff6 * v[2]);
        double cgh = djac * D_ONE_60 * c * (gg1 * h[0] + gg2 * h[1] + gg3 * h[2]);
        /* integrate_triangle_dphi_f_f_h (f = u) and (f = v) */
        double uu0 = u[0] * u[0], uu1 = u[1] * u[1], uu2 = u[2] * u[2];
        double q1 = (3 * uu0 + 2 * u[0] * u[1] + uu1 + (2 * u[0] + u[1]) * u[2] + uu2);
        double q2 = (uu0 + 2 * u[0] * u[1] + 3 * uu1 + (u[0] + 2 * u[1]) * u[2] + uu2);
        double q3 = (uu0 + u[0] * u[1] + uu1 + 2 * (u[0] + u[1]) * u[2] + 3 * uu2);
        double cu = djac * D_ONE_30 * c * (q1 * h[0] + q2 * h[1] + q3 * h[2]);
        double vv0 = v[0] * v[0], vv1 = v[1] * v[1], vv2 = v[2] * v[2];
        double s1 = (3 * vv0 + 2 * v[0] * v[1] + vv1 + (2 * v[0] + v[1]) * v[2] + vv2);
        double s2 = (vv0 + 2 * v[0] * v[1] + 3 * vv1 + (v[0] + 2 * v[1]) * v[2] + vv2);
        double s3 = (vv0 + v[0] * v[1] + vv1 + 2 * (v[0] + v[1]) * v[2] + 3 * vv2);
        double cv = djac * D_ONE_30 * c * (s1 * h[0] + s2 * h[1] + s3 * h[2]);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            r[3 * i] = con * (gx[i] * tx + gy[i] * ty);
            double rx = cgh * gy[i];
            rx += cu * gx[i];
            double ry = cgh * gx[i];
            ry += cv * gy[i];
            r[3 * i + 1] = rx;
            r[3 * i + 2] = ry;
        }
    }
    /* temporal, :456-463 */
    sw2_temporal(in, h, u, v, dh, factor, (1. + P.alpha / 2.), -3, r);
    sw2_temporal(in, in.ho, in.uo, in.vo, dho, factor_old, (-1.) * (1. + P.alpha), in.wd_flag, r);
    if (P.alpha != 0.0) sw2_temporal(in, in.hoo, in.uoo, in.voo, dhoo, factor_older, (P.alpha / 2.), in.wd_flag, r);

    /* friction, :520-539 */
    const double resistance = FD ? rough * (0.5 / g) : rough / (2.0 * g);
    {
        double fx[3], fy[3];
        if (FD && sh) {
            /* the perturbed node's speed is the one magnitude a u or v perturbation changes */
            double up = u[0], vp = v[0];
#pragma unroll
            for (int i = 1; i < 3; i++) {
                up = (i == pnode) ? u[i] : up;
                vp = (i == pnode) ? v[i] : vp;
            }
            const double tp = vsame ? 0.0 : sqrt(up * up + vp * vp);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const double t1 = (!vsame && i == pnode) ? tp : sh->t1_0[i];
                fx[i] = u[i] * t1;
                fy[i] = v[i] * t1;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double t1 = sqrt(u[i] * u[i] + v[i] * v[i]);
                fx[i] = u[i] * t1;
                fy[i] = v[i] * t1;
            }
        }
        double t1 = dt * g * resistance * factor;
        double tt = djac * t1 * D_ONE_12;
        double sx = fx[0] + fx[1] + fx[2], sy = fy[0] + fy[1] + fy[2];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            r[3 * i + 1] += tt * (sx + fx[i]);
            r[3 * i + 2] += tt * (sy + fy[i]);
        }
    }
    const double bsx = resistance * uwd * vmag;
    const double bsy = resistance * vwd * vmag;

    if (pflag && dh == 0) {
        /* body pressure, :710-739 (integrate_triangle_dphi_f_f) */
        {
            double c = -D_ONE_2 * g * dt;
            double iff = c * D_ONE_6 * djac *
                         (h[0] * h[0] + h[0] * h[1] + h[1] * h[1] + h[0] * h[2] + h[1] * h[2] + h[2] * h[2]);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                r[3 * i + 1] += gx[i] * iff;
                r[3 * i + 2] += gy[i] * iff;
            }
        }
        /* bed slope, :747-768 (integrate_triangle_phi_h_df) */
        {
            double c = g * dt;
            double sh = h[0] + h[1] + h[2];
            double cx = djac * D_ONE_12 * c * gzx;
            double cy = djac * D_ONE_12 * c * gzy;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                r[3 * i + 1] += cx * (sh + h[i]);
                r[3 * i + 2] += cy * (sh + h[i]);
            }
        }
        /* boundary pressure, :776-804 (fe_sw2_wd_boundaryPressure_triangle, integrations.c:265) */
        {
            double c = D_ONE_2 * dt * g;
            double t1 = 1. * D_ONE_12 * c;
            double bx[3] = {0., 0., 0.}, by[3] = {0., 0., 0.};
#pragma unroll
            for (int e = 0; e < 3; e++) {
                int a = e, b = (e + 1) % 3;
                double delx = in.x[a] - in.x[b];
                double dely = in.y[b] - in.y[a];
                double ha = h[a], hb = h[b];
                double ra = t1 * (3 * ha * ha + 2 * ha * hb + hb * hb);
                double rb = t1 * (ha * ha + 2 * ha * hb + 3 * hb * hb);
                bx[a] += dely * ra;
                bx[b] += dely * rb;
                by[a] += delx * ra;
                by[b] += delx * rb;
            }
#pragma unroll
            for (int i = 0; i < 3; i++) {
                r[3 * i + 1] += bx[i];
                r[3 * i + 2] += by[i];
            }
        }
    }

    /* SUPG, :921-1142 (triangle branch; tau from get_supg_tau.c:43-80, le = sqrt(area), :173) */
    if (!(fabs(hwd) < D_SMALL)) {
        const double hden = hwd > D_NQS ? hwd : D_NQS;
        const double le = (FD && sh) ? sh->le : sqrt(djac * factor_old);
        const double cel2 = g * hden;
        const double sr = sqrt(uwd * uwd + vwd * vwd + cel2 + D_SMALL);
        const double tau = 0.5 * le / sr;
        double dhdt = 0., dudt = 0., dvdt = 0.;
        if (dho == 0) {
            dhdt = -(djac * (FD ? (in.ho[0] + in.ho[1] + in.ho[2]) * D_ONE_3 : (in.ho[0] + in.ho[1] + in.ho[2]) / 3.0));
            dudt = -(djac * (FD ? (in.uo[0] + in.uo[1] + in.uo[2]) * D_ONE_3 : (in.uo[0] + in.uo[1] + in.uo[2]) / 3.0));
            dvdt = -(djac * (FD ? (in.vo[0] + in.vo[1] + in.vo[2]) * D_ONE_3 : (in.vo[0] + in.vo[1] + in.vo[2]) / 3.0));
        }
        dudt += djac * (FD ? (u[0] + u[1] + u[2]) * D_ONE_3 : (u[0] + u[1] + u[2]) / 3.0);
        dvdt += djac * (FD ? (v[0] + v[1] + v[2]) * D_ONE_3 : (v[0] + v[1] + v[2]) / 3.0);
        dhdt += djac * (FD ? (h[0] + h[1] + h[2]) * D_ONE_3 : (h[0] + h[1] + h[2]) / 3.0);
        /* fe_sw2_wd_gls_convection_triangle, integrations.c:700-722 */
        double havg = D_ONE_3 * (h[0] + h[1] + h[2]);
        double uavg = D_ONE_3 * (u[0] + u[1] + u[2]);
        double vavg = D_ONE_3 * (v[0] + v[1] + v[2]);
        double dhdx = gx[0] * h[0] + gx[1] * h[1] + gx[2] * h[2];
        double dhdy = gy[0] * h[0] + gy[1] * h[1] + gy[2] * h[2];
        double dudx = gx[0] * u[0] + gx[1] * u[1] + gx[2] * u[2];
        double dudy = gy[0] * u[0] + gy[1] * u[1] + gy[2] * u[2];
        double dvdx = gx[0] * v[0] + gx[1] * v[1] + gx[2] * v[2];
        double dvdy = gy[0] * v[0] + gy[1] * v[1] + gy[2] * v[2];
        double k = djac * dt * 1;
        double a0 = k * (uavg * dhdx + havg * dudx + vavg * dhdy + havg * dvdy);
        double a1 = k * (uavg * dudx + vavg * dudy);
        double a2 = k * (uavg * dvdx + vavg * dvdy);
        const double fct = 1.0; /* wrapper factor of an all-wet field */
        double xp = g * (ghx + gzx) * dt * fct;
        double yp = g * (ghy + gzy) * dt * fct;
        double xf = g * bsx * dt * fct;
        double yf = g * bsy * dt * fct;
        double c_eq = dhdt + a0;
        const double ihden = FD ? (hsame ? sh->ihden0 : 1.0 / hden) : 0.0;
        double x_eq = dudt + a1 + (xp + (FD ? xf * ihden : xf / hden)) * djac;
        double y_eq = dvdt + a2 + (yp + (FD ? yf * ihden : yf / hden)) * djac;
        const double gf = g;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            r[3 * i] += (gx[i] * (tau * uwd * c_eq + tau * hwd * x_eq) + gy[i] * (tau * vwd * c_eq + tau * hwd * y_eq));
            r[3 * i + 1] += (gx[i] * (tau * gf * hwd * c_eq + tau * uwd * hwd * x_eq) + gy[i] * (tau * vwd * hwd * x_eq));
            r[3 * i + 2] += (gy[i] * (tau * gf * hwd * c_eq + tau * vwd * hwd * y_eq) + gx[i] * (tau * uwd * hwd * y_eq));
        }
    }
    return 0;
}

/* the shared values of an all-wet element (the FD residual's unperturbed operands, formed the same way) */
__device__ __forceinline__ void fd_share(const TriIn &in, const Sw2Par &P, FdShare &sh)
{
    const double *h = in.h, *u = in.u, *v = in.v;
    const double hwd = (h[0] + h[1] + h[2]) * D_ONE_3;
    const double uwd = (u[0] + u[1] + u[2]) * D_ONE_3;
    const double vwd = (v[0] + v[1] + v[2]) * D_ONE_3;
    const int dho = n_dry3(in.ho[0], in.ho[1], in.ho[2]);
    const double factor_old = (dho == 0) ? 1.0 : 0.0;
    sh.le = sqrt(in.djac * factor_old);
    sh.vmag0 = sqrt(uwd * uwd + vwd * vwd);
    if (P.manning) {
        double hf = hwd > D_SMALL ? hwd : D_SMALL;
        sh.rough0 = (in.ks > D_SMALL) ? drag_coef(hf, in.ks) : 0.0;
    } else {
        sh.rough0 = P.cf;
    }
    const double hden = hwd > D_NQS ? hwd : D_NQS;
    sh.ihden0 = 1.0 / hden;
#pragma unroll
    for (int i = 0; i < 3; i++) sh.t1_0[i] = sqrt(u[i] * u[i] + v[i] * v[i]);
}

/* ---- 1-D boundary residuals ---- */
__device__ __forceinline__ void line_phi_f_g_h(double djac, double c, double f1, double f2, double g1, double g2,
                                               double h1, double h2, double *in)
{ /* fe_integrals_1d.c:58 */
    double t1 = D_ONE_60 * djac * c;
    in[0] += t1 * ((3 * (4 * f1 + f2) * g1 + (3 * f1 + 2 * f2) * g2) * h1 +
                   ((3 * f1 + 2 * f2) * g1 + (2 * f1 + 3 * f2) * g2) * h2);
    in[1] += t1 * (((3 * f1 + 2 * f2) * g1 + (2 * f1 + 3 * f2) * g2) * h1 +
                   ((2 * f1 + 3 * f2) * g1 + 3 * (f1 + 4 * f2) * g2) * h2);
}

/* fe_sw2_1d_implicit_flow, fe_sw2_boundary_integrals.c:125-174 */
__device__ __forceinline__ void implicit_flow(double *er, double nx, double ny, double djac, double dt,
                                              const double *vx, const double *vy, const double *h)
{
    double rc[2] = {0, 0}, rx[2] = {0, 0}, ry[2] = {0, 0};
    double ur1 = vx[0], ur2 = vx[1], vr1 = vy[0], vr2 = vy[1];
    double d0 = vx[0] * nx + vy[0] * ny, d1 = vx[1] * nx + vy[1] * ny;
    if (d0 < D_SMALL) { ur1 = 0.; vr1 = 0.; }
    if (d1 < D_SMALL) { ur2 = 0.; vr2 = 0.; }
    double t1 = djac * D_ONE_12 * 1.; /* integrate_line_phi_h_v_dot_n, fe_integrals_1d.c:51 */
    rc[0] += t1 * (nx * ((3 * h[0] + h[1]) * vx[0] + (h[0] + h[1]) * vx[1]) +
                   ny * ((3 * h[0] + h[1]) * vy[0] + (h[0] + h[1]) * vy[1]));
    rc[1] += t1 * (nx * ((h[0] + h[1]) * vx[0] + (h[0] + 3 * h[1]) * vx[1]) +
                   ny * ((h[0] + h[1]) * vy[0] + (h[0] + 3 * h[1]) * vy[1]));
    line_phi_f_g_h(djac, nx, ur1, ur2, vx[0], vx[1], h[0], h[1], rx);
    line_phi_f_g_h(djac, ny, vr1, vr2, vy[0], vy[1], h[0], h[1], ry);
    line_phi_f_g_h(djac, ny, ur1, ur2, vy[0], vy[1], h[0], h[1], rx);
    line_phi_f_g_h(djac, nx, vr1, vr2, vx[0], vx[1], h[0], h[1], ry);
    for (int i = 0; i < 2; i++) {
        er[i * 3] += dt * rc[i];
        er[i * 3 + 1] += dt * rx[i];
        er[i * 3 + 2] += dt * ry[i];
    }
}

/* fe_sw2_1d_pressure, fe_sw2_boundary_integrals.c:260-284 */
__device__ __forceinline__ void pressure_1d(double *er, double nx, double ny, double djac, double dt,
                                            const double *h, double g)
{
    double t = djac * D_ONE_12 * 1.;
    double r0 = t * (3 * h[0] * h[0] + 2 * h[0] * h[1] + h[1] * h[1]);
    double r1 = t * (h[0] * h[0] + 2 * h[0] * h[1] + 3 * h[1] * h[1]);
    double t1 = dt * D_ONE_2 * g;
    er[1] += t1 * nx * r0;
    er[2] += t1 * ny * r0;
    er[4] += t1 * nx * r1;
    er[5] += t1 * ny * r1;
}

__device__ __forceinline__ void sw2_seg_resid(int kind, double val0, double val1, double nx, double ny, double djac,
                                              const double *hin, const double *uin, const double *vin,
                                              const Sw2Par &P, int pnode, int pvar, double pe, double er[6])
{
    double h[2] = {hin[0], hin[1]}, u[2] = {uin[0], uin[1]}, v[2] = {vin[0], vin[1]};
    bool pflag = true;
    if (pvar == 0) {
        h[pnode] += pe;
    } else if (pvar == 1) {
        u[pnode] += pe;
        pflag = false;
    } else if (pvar == 2) {
        v[pnode] += pe;
        pflag = false;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) er[k] = 0.0;
    const double dt = P.dt, g = P.g;
    if (kind == ADHK_BC_DEPTH || kind == ADHK_BC_DEPTH_NODAL) { /* fe_sw2_bc_h, sw2_bc_h.c:57-178 */
        double hd[2] = {val0, kind == ADHK_BC_DEPTH ? val0 : val1};
        if (hd[0] < D_SMALL || hd[1] < D_SMALL || h[0] < D_SMALL || h[1] < D_SMALL) return;
        implicit_flow(er, nx, ny, djac, dt, u, v, hd);
        if (pflag) {
            pressure_1d(er, nx, ny, djac, dt, hd, g);
            pressure_1d(er, nx, ny, djac, dt, h, -g);
        }
    } else if (kind == ADHK_BC_FLUX) { /* fe_sw2_1d_explicit_flow, fe_sw2_boundary_integrals.c:56-94 */
        double flux = val0;
        double h_avg = D_ONE_2 * (h[0] + h[1]);
        if (h_avg < D_SMALL) h_avg = D_SMALL;
        double rc = D_ONE_2 * djac * 1.;
        double rx[2], ry[2];
        if (flux < 0) {
            double us = 0., vs = 0.;
            if (h_avg > 0) {
                us = flux * nx / h_avg;
                vs = flux * ny / h_avg;
            }
            rx[0] = rx[1] = D_ONE_2 * djac * us;
            ry[0] = ry[1] = D_ONE_2 * djac * vs;
        } else {
            double t1 = D_ONE_6 * djac * 1.;
            rx[0] = t1 * (2 * u[0] + u[1]);
            rx[1] = t1 * (u[0] + 2 * u[1]);
            ry[0] = t1 * (2 * v[0] + v[1]);
            ry[1] = t1 * (v[0] + 2 * v[1]);
        }
        double t1 = dt * flux;
        for (int i = 0; i < 2; i++) {
            er[i * 3] += t1 * rc;
            er[i * 3 + 1] += t1 * rx[i];
            er[i * 3 + 2] += t1 * ry[i];
        }
    } else { /* fe_sw2_bc_vel, sw2_bc_vel.c:58-162 */
        double vx[2] = {val0, val0}, vy[2] = {val1, val1};
        implicit_flow(er, nx, ny, djac, dt, vx, vy, h);
    }
}

/* NUM_DIFF_EPSILON, macro.h:669-673 */
__device__ __forceinline__ double num_diff_eps(double var, double p)
{
    double e = fabs(var) * (1. + p);
    e -= fabs(var);
    if (e < p) e = p;
    return e;
}

#endif
