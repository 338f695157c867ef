"""Independent numpy statement of the SW2 triangle residual by numerical quadrature.

Pins the closed-form P1 integrals of the oracle (src/fe/fe_integrals_2d.c and
fe_integrals_1d.c as used by sw2_body_resid.c) against the weak form they
integrate, evaluated with a 7-point degree-5 triangle rule (exact for the
cubic integrands h*u*u) and 3-point Gauss-Legendre on edges.  Only the
all-wet branch is stated.  The SUPG term (sw2_body_resid.c:921-1142) is stated
from the shallow-water system itself: its strong residuals (time differences,
convection, pressure, friction) are integrated by quadrature, and they are
weighted by the primitive-variable flux Jacobians A_x, A_y of the quasi-linear
operator, obtained by central differences of that operator (flux_jacobians),
not typed from the oracle.
"""
import numpy as np

S15 = np.sqrt(15.0)
A1, B1 = (6 - S15) / 21, (9 + 2 * S15) / 21
A2, B2 = (6 + S15) / 21, (9 - 2 * S15) / 21
W1, W2 = (155 - S15) / 1200, (155 + S15) / 1200
BARY = np.array([[1 / 3, 1 / 3, 1 / 3],
                 [A1, A1, B1], [A1, B1, A1], [B1, A1, A1],
                 [A2, A2, B2], [A2, B2, A2], [B2, A2, A2]])
WTS = np.array([9 / 40, W1, W1, W1, W2, W2, W2])
GL_X = np.array([0.5 - np.sqrt(0.15), 0.5, 0.5 + np.sqrt(0.15)])
GL_W = np.array([5 / 18, 8 / 18, 5 / 18])


def p1_grad(x, y):
    area2 = (x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0])
    gx = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]]) / area2
    gy = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]]) / area2
    return 0.5 * area2, gx, gy


def quad(area, f_at_points):
    return area * np.dot(WTS, f_at_points)


def quasi_linear_operator(q, qx, qy, g):
    """The shallow-water system in primitive variables q = (h, u, v) without sources and time derivative:
    continuity h_t + (hu)_x + (hv)_y and momentum u_t + u u_x + v u_y + g h_x, v_t + u v_x + v v_y + g h_y."""
    h, u, v = q
    hx, ux, vx = qx
    hy, uy, vy = qy
    return np.array([u * hx + h * ux + v * hy + h * vy,
                     u * ux + v * uy + g * hx,
                     u * vx + v * vy + g * hy])


def flux_jacobians(q, g, delta=1.0):
    """A_x = dF/dq_x and A_y = dF/dq_y of the quasi-linear operator at q, by central differences
    (the operator is linear in the gradients, so the differences are exact up to rounding)."""
    q = np.asarray(q, float)
    Ax, Ay = np.zeros((3, 3)), np.zeros((3, 3))
    z = np.zeros(3)
    for j in range(3):
        e = np.zeros(3)
        e[j] = delta
        Ax[:, j] = (quasi_linear_operator(q, e, z, g) - quasi_linear_operator(q, -e, z, g)) / (2 * delta)
        Ay[:, j] = (quasi_linear_operator(q, z, e, g) - quasi_linear_operator(q, z, -e, g)) / (2 * delta)
    return Ax, Ay


def supg_term(tau, gx, gy, hb, ub, vb, g, R):
    """Streamline-upwind test-function term of node i: tau * sum_k dphi_i/dk * diag(1, h, h) A_k(h, u, v) R,
    where R = (c, x, y) are the element-integrated strong residuals of continuity and of the two momentum
    equations in primitive variables; the diag(1, h, h) weights the momentum rows by the depth, as the
    conservative momentum unknowns h u, h v are tested."""
    Ax, Ay = flux_jacobians((hb, ub, vb), g)
    D = np.diag([1.0, hb, hb])
    wx, wy = D @ Ax @ R, D @ Ay @ R
    out = np.zeros(9)
    for i in range(3):
        out[3 * i:3 * i + 3] = tau * (gx[i] * wx + gy[i] * wy)
    return out


def body_residual(x, y, z, h, u, v, ho, uo, vo, dt, g, cf, dll=1e-6, supg=True):
    """9-vector [node*3 + (c, x, y)] of the SW2 body residual for an all-wet triangle (supg=False leaves
    out the streamline-upwind term)."""
    A, gx, gy = p1_grad(x, y)
    phi = BARY  # (7 points, 3 basis)
    hq, uq, vq = phi @ h, phi @ u, phi @ v
    zx, zy = gx @ z, gy @ z
    r = np.zeros(9)
    for i in range(3):
        # advection: -dt int grad(phi_i) . (h v), (h u u, h u v), (h u v, h v v)
        r[3 * i] += -dt * quad(A, gx[i] * hq * uq + gy[i] * hq * vq)
        r[3 * i + 1] += -dt * quad(A, gx[i] * hq * uq * uq + gy[i] * hq * uq * vq)
        r[3 * i + 2] += -dt * quad(A, gx[i] * hq * uq * vq + gy[i] * hq * vq * vq)
        # temporal: consistent continuity, lumped momentum (backward Euler weights 1, -1)
        r[3 * i] += quad(A, phi[:, i] * (hq - phi @ ho))
        r[3 * i + 1] += A / 3 * (u[i] * max(h[i], dll) - uo[i] * max(ho[i], dll))
        r[3 * i + 2] += A / 3 * (v[i] * max(h[i], dll) - vo[i] * max(ho[i], dll))
        # friction on the nodal interpolant of u|v|: dt (cf/2) int phi_i I(u|v|)
        sp = np.sqrt(u ** 2 + v ** 2)
        r[3 * i + 1] += dt * cf / 2 * quad(A, phi[:, i] * (phi @ (u * sp)))
        r[3 * i + 2] += dt * cf / 2 * quad(A, phi[:, i] * (phi @ (v * sp)))
        # pressure -1/2 g dt int dphi_i h^2, bed slope g dt int phi_i h grad z
        r[3 * i + 1] += -0.5 * g * dt * gx[i] * quad(A, hq ** 2) + g * dt * zx * quad(A, phi[:, i] * hq)
        r[3 * i + 2] += -0.5 * g * dt * gy[i] * quad(A, hq ** 2) + g * dt * zy * quad(A, phi[:, i] * hq)
    # boundary pressure 1/2 g dt int_{edges} phi_i h^2 n
    for a, b in ((0, 1), (1, 2), (2, 0)):
        L = np.hypot(x[b] - x[a], y[b] - y[a])
        nx, ny = (y[b] - y[a]) / L, -(x[b] - x[a]) / L
        for s, w in zip(GL_X, GL_W):
            hs = (1 - s) * h[a] + s * h[b]
            for node, ph in ((a, 1 - s), (b, s)):
                r[3 * node + 1] += 0.5 * g * dt * w * L * ph * hs * hs * nx
                r[3 * node + 2] += 0.5 * g * dt * w * L * ph * hs * hs * ny
    # SUPG on element averages
    hb, ub, vb = h.mean(), u.mean(), v.mean()
    if abs(hb) >= np.finfo(float).eps:
        hden = max(hb, np.finfo(np.float32).eps)
        tau = 0.5 * np.sqrt(A) / np.sqrt(ub * ub + vb * vb + g * hden + np.finfo(float).eps)
        dhdt = quad(A, hq - phi @ ho)
        dudt = quad(A, uq - phi @ uo)
        dvdt = quad(A, vq - phi @ vo)
        hx, hy, ux, uy, vx, vy = gx @ h, gy @ h, gx @ u, gy @ u, gx @ v, gy @ v
        a0 = dt * quad(A, np.full(7, ub * hx + hb * ux + vb * hy + hb * vy))
        a1 = dt * quad(A, np.full(7, ub * ux + vb * uy))
        a2 = dt * quad(A, np.full(7, ub * vx + vb * vy))
        res = cf / (2 * g)
        vmag = np.hypot(ub, vb)
        xf, yf = g * res * ub * vmag * dt, g * res * vb * vmag * dt
        xp, yp = g * (hx + zx) * dt, g * (hy + zy) * dt
        c_eq = dhdt + a0
        x_eq = dudt + a1 + (xp + xf / hden) * A
        y_eq = dvdt + a2 + (yp + yf / hden) * A
        if supg:
            r += supg_term(tau, gx, gy, hb, ub, vb, g, np.array([c_eq, x_eq, y_eq]))
    return r


def wet_pieces(h):
    """Triangles covering the wet part {h > 0} of the P1 depth field, as 3x3 barycentric vertex rows,
    and the wet polygon (counter-clockwise when the element is)."""
    corners = np.eye(3)
    poly = []
    for k in range(3):
        a, b = corners[k], corners[(k + 1) % 3]
        ha, hb = h[k], h[(k + 1) % 3]
        if ha > 0:
            poly.append(a)
        if (ha > 0) != (hb > 0):
            poly.append(a + ha / (ha - hb) * (b - a))
    tris = [np.array([poly[0], poly[i], poly[i + 1]]) for i in range(1, len(poly) - 1)]
    return tris, poly


def wet_quad(A, pieces, fn):
    """sum over the wet pieces of the 7-point rule; fn(P) gets the points' barycentric coordinates (7x3)."""
    tot = 0.0
    for V in pieces:
        P = BARY @ V
        tot = tot + A * abs(np.linalg.det(V)) * np.tensordot(WTS, fn(P), axes=(0, 0))
    return tot


def wet_body_residual(x, y, z, h, u, v, ho, uo, vo, dt, g, cf, dll=1e-6, supg=True):
    """SW2 body residual of a triangle that may be partly dry (h <= 0 at some nodes).

    What the wet/dry wrapper (fe_wet_dry_wrapper.c:104-304, with RE_DISTRIBUTE) integrates:
    every wrapped term is the weak form over the wet region {h > 0} of the linear depth,
    tested with the ORIGINAL P1 basis functions; the lumped momentum storage and the
    friction are full-element terms scaled by the wet area fraction; the SUPG term uses
    wet-region averages and integrals.  Stated here by clipping the element and applying the
    7-point rule on each wet piece (the integrands are at most cubic on a piece).
    """
    A, gx, gy = p1_grad(x, y)
    zx, zy = gx @ z, gy @ z
    wet, poly = wet_pieces(h)
    wet_o, _ = wet_pieces(ho)
    area = wet_quad(A, wet, lambda P: np.ones(len(P)))
    area_o = wet_quad(A, wet_o, lambda P: np.ones(len(P)))
    fac, fac_o = area / A, area_o / A
    r = np.zeros(9)

    def Q(fn, pieces=wet):
        return wet_quad(A, pieces, fn)

    for i in range(3):
        # advection
        r[3 * i] += -dt * Q(lambda P: (gx[i] * (P @ h) * (P @ u) + gy[i] * (P @ h) * (P @ v)))
        r[3 * i + 1] += -dt * Q(lambda P: (P @ h) * (P @ u) * (gx[i] * (P @ u) + gy[i] * (P @ v)))
        r[3 * i + 2] += -dt * Q(lambda P: (P @ h) * (P @ v) * (gx[i] * (P @ u) + gy[i] * (P @ v)))
        # storage: continuity over the wet regions of each level, lumped momentum scaled by the wet fraction
        r[3 * i] += Q(lambda P: P[:, i] * (P @ h)) - Q(lambda P: P[:, i] * (P @ ho), wet_o)
        r[3 * i + 1] += A / 3 * (fac * u[i] * max(h[i], dll) - fac_o * uo[i] * max(ho[i], dll))
        r[3 * i + 2] += A / 3 * (fac * v[i] * max(h[i], dll) - fac_o * vo[i] * max(ho[i], dll))
        # friction on the whole element, scaled by the wet fraction
        sp = np.sqrt(u ** 2 + v ** 2)
        full = [np.eye(3)]
        r[3 * i + 1] += fac * dt * cf / 2 * wet_quad(A, full, lambda P: P[:, i] * (P @ (u * sp)))
        r[3 * i + 2] += fac * dt * cf / 2 * wet_quad(A, full, lambda P: P[:, i] * (P @ (v * sp)))
        # pressure and bed slope over the wet region
        r[3 * i + 1] += -0.5 * g * dt * gx[i] * Q(lambda P: (P @ h) ** 2) + g * dt * zx * Q(lambda P: P[:, i] * (P @ h))
        r[3 * i + 2] += -0.5 * g * dt * gy[i] * Q(lambda P: (P @ h) ** 2) + g * dt * zy * Q(lambda P: P[:, i] * (P @ h))
    # boundary pressure on the wet polygon's edges (h = 0 on the wet/dry line)
    X = np.array([x, y])
    for k in range(len(poly)):
        a, b = poly[k], poly[(k + 1) % len(poly)]
        pa, pb = X @ a, X @ b
        L = np.hypot(*(pb - pa))
        if L == 0.0:
            continue
        nx, ny = (pb[1] - pa[1]) / L, -(pb[0] - pa[0]) / L
        for s, w in zip(GL_X, GL_W):
            bc = (1 - s) * a + s * b
            hs = bc @ h
            for node in range(3):
                r[3 * node + 1] += 0.5 * g * dt * w * L * bc[node] * hs * hs * nx
                r[3 * node + 2] += 0.5 * g * dt * w * L * bc[node] * hs * hs * ny
    # SUPG with wet-region averages
    hb = Q(lambda P: P @ h) / area if area > 0 else 0.0
    ub = Q(lambda P: P @ u) / area if area > 0 else 0.0
    vb = Q(lambda P: P @ v) / area if area > 0 else 0.0
    if abs(hb) >= np.finfo(float).eps:
        hden = max(hb, np.finfo(np.float32).eps)
        tau = 0.5 * np.sqrt(A * fac_o) / np.sqrt(ub * ub + vb * vb + g * hden + np.finfo(float).eps)
        dhdt = Q(lambda P: P @ h) - Q(lambda P: P @ ho, wet_o)
        dudt = Q(lambda P: P @ u) - Q(lambda P: P @ uo, wet_o)
        dvdt = Q(lambda P: P @ v) - Q(lambda P: P @ vo, wet_o)
        hx, hy, ux, uy, vx, vy = gx @ h, gy @ h, gx @ u, gy @ u, gx @ v, gy @ v
        a0 = dt * Q(lambda P: (P @ u) * hx + (P @ h) * ux + (P @ v) * hy + (P @ h) * vy)
        a1 = dt * Q(lambda P: (P @ u) * ux + (P @ v) * uy)
        a2 = dt * Q(lambda P: (P @ u) * vx + (P @ v) * vy)
        res = cf / (2 * g)
        vmag = np.hypot(ub, vb)
        xf, yf = g * res * ub * vmag * dt * fac, g * res * vb * vmag * dt * fac
        xp, yp = g * (hx + zx) * dt * fac, g * (hy + zy) * dt * fac
        c_eq = dhdt + a0
        x_eq = dudt + a1 + (xp + xf / hden) * A
        y_eq = dvdt + a2 + (yp + yf / hden) * A
        if supg:
            r += supg_term(tau, gx, gy, hb, ub, vb, g, np.array([c_eq, x_eq, y_eq]))
    return r
