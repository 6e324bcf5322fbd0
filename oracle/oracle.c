/*
 * oracle.c -- CPU restatement of the visionary-sim per-substep hot path.
 *
 * TEST INFRASTRUCTURE (see oracle.h). Every function follows one WGSL kernel
 * of the reference, operation by operation, in the evaluation order this
 * project fixes for f32 (left-to-right sums, `dot(a,a) = (ax*ax + ay*ay) +
 * az*az`, WGSL matrix products column by column, the mat3.js determinant).
 * The reference's own bit-level order belongs to the browser's WGSL compiler
 * and is not recoverable, so this file is the definition the HIP kernels are
 * held to bit-for-bit; against the reference itself the bar is 1e-5 relative.
 *
 * Build: REAL=float (liboracle_f32.so) and REAL=double (liboracle_f64.so),
 * both with -ffp-contract=off and no fast-math (oracle/Makefile).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef REAL real;

/* Correctly rounded square root and exact floor in the working precision. */
#define SQRT(x) ((real)sqrt((double)(x)))
#define FLOOR(x) ((real)floor((double)(x)))
#define FABS(x) ((real)fabs((double)(x)))

typedef struct { real x, y, z; } v3;
typedef struct { v3 c[3]; } m3; /* column-major, like WGSL mat3x3f */

static inline v3 mk(real x, real y, real z) { v3 r = { x, y, z }; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scale(v3 a, real s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline real dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 cross(v3 a, v3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real length3(v3 a) { return SQRT(dot(a, a)); }

/* Fused multiply-add in the working precision (fmaf for the f32 build: one
 * rounding, as the GPU's v_fma_f32). */
static inline real FMA(real a, real b, real c)
{
    return sizeof(real) == 4 ? (real)fmaf((float)a, (float)b, (float)c) : (real)fma((double)a, (double)b, (double)c);
}
static inline real dotf(v3 a, v3 b) { return FMA(a.z, b.z, FMA(a.y, b.y, a.x * b.x)); }
static inline real dotacc(v3 a, v3 b, real acc) { return FMA(a.z, b.z, FMA(a.y, b.y, FMA(a.x, b.x, acc))); }
static inline v3 crossf(v3 a, v3 b)
{
    return mk(FMA(a.y, b.z, -(a.z * b.y)), FMA(a.z, b.x, -(a.x * b.z)), FMA(a.x, b.y, -(a.y * b.x)));
}

/* (m * n)[:, j] = m[:,0] * n[j].x + m[:,1] * n[j].y + m[:,2] * n[j].z */
static m3 mmul(m3 m, m3 n)
{
    m3 r;
    for (int j = 0; j < 3; j++) {
        r.c[j] = add(add(scale(m.c[0], n.c[j].x), scale(m.c[1], n.c[j].y)), scale(m.c[2], n.c[j].z));
    }
    return r;
}

static m3 mscale(m3 m, real s)
{
    m3 r = { { scale(m.c[0], s), scale(m.c[1], s), scale(m.c[2], s) } };
    return r;
}

static m3 transpose(m3 m)
{
    m3 r = { { mk(m.c[0].x, m.c[1].x, m.c[2].x),
               mk(m.c[0].y, m.c[1].y, m.c[2].y),
               mk(m.c[0].z, m.c[1].z, m.c[2].z) } };
    return r;
}

/* mat3Determinant, discretize/mat3.js:30-36, m[col*3+row]. */
static real det3(m3 m)
{
    real m0 = m.c[0].x, m1 = m.c[0].y, m2 = m.c[0].z;
    real m3_ = m.c[1].x, m4 = m.c[1].y, m5 = m.c[1].z;
    real m6 = m.c[2].x, m7 = m.c[2].y, m8 = m.c[2].z;
    return m0 * (m4 * m8 - m7 * m5) - m3_ * (m1 * m8 - m7 * m2) + m6 * (m1 * m5 - m4 * m2);
}

/* KUHN_TETS, discretize/lattice/kuhn.js:28-35; corner c = (c&1, (c>>1)&1, (c>>2)&1)
 * per corners.js:10-19. */
static const int KUHN_TETS[6][4] = {
    { 0, 1, 3, 7 }, { 0, 5, 1, 7 }, { 0, 3, 2, 7 }, { 0, 2, 6, 7 }, { 0, 4, 5, 7 }, { 0, 6, 4, 7 }
};

/* Unit-cube rest shape inverses (kuhn.js:78-80 = mat3Inverse of the corner
 * edge matrices). Every Dm has det 1 and integer entries, so the inverse is
 * the adjugate and exact in any precision. Built once, lazily. */
static m3 KUHN_INV[6];
/* Every Kuhn tetrahedron is a path of three axis steps from corner 0 to
 * corner 7: PATH[t] lists its corners in path order, AXIS[t][j] the axis of
 * step j. */
static int PATH[6][4], AXIS[6][3];
static real STRETCH_LEN[28];
static int STRETCH_A[28], STRETCH_B[28];
static int tables_ready;

static void build_tables(void)
{
    if (tables_ready) return;
    for (int t = 0; t < 6; t++) {
        double e[3][3]; /* e[k] = corner(t,k+1) - corner(t,0), as columns */
        int c0 = KUHN_TETS[t][0];
        for (int k = 0; k < 3; k++) {
            int c = KUHN_TETS[t][k + 1];
            e[k][0] = (double)((c & 1) - (c0 & 1));
            e[k][1] = (double)(((c >> 1) & 1) - ((c0 >> 1) & 1));
            e[k][2] = (double)(((c >> 2) & 1) - ((c0 >> 2) & 1));
        }
        /* column-major m[col*3+row] */
        double m[9] = { e[0][0], e[0][1], e[0][2], e[1][0], e[1][1], e[1][2], e[2][0], e[2][1], e[2][2] };
        double det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
                     m[6] * (m[1] * m[5] - m[4] * m[2]);
        double s = 1.0 / det;
        double inv[9] = {
            (m[4] * m[8] - m[7] * m[5]) * s, (m[7] * m[2] - m[1] * m[8]) * s, (m[1] * m[5] - m[4] * m[2]) * s,
            (m[6] * m[5] - m[3] * m[8]) * s, (m[0] * m[8] - m[6] * m[2]) * s, (m[3] * m[2] - m[0] * m[5]) * s,
            (m[3] * m[7] - m[6] * m[4]) * s, (m[6] * m[1] - m[0] * m[7]) * s, (m[0] * m[4] - m[3] * m[1]) * s
        };
        for (int col = 0; col < 3; col++) {
            KUHN_INV[t].c[col] = mk((real)inv[col * 3], (real)inv[col * 3 + 1], (real)inv[col * 3 + 2]);
        }
    }
    for (int t = 0; t < 6; t++) {
        int at = 0;
        PATH[t][0] = 0;
        for (int j = 0; j < 3; j++) {
            for (int a = 0; a < 3; a++) {
                int next = at | (1 << a);
                if (next == at) continue;
                int member = 0;
                for (int v = 0; v < 4; v++) member |= KUHN_TETS[t][v] == next;
                if (member) {
                    AXIS[t][j] = a;
                    at = next;
                    break;
                }
            }
            PATH[t][j + 1] = at;
        }
    }
    /* STRETCH_EDGES, stretch.js:5-7: every corner pair i<j with its unit-cell
     * length, stored as an f32 literal by the kernel generator. */
    int k = 0;
    for (int i = 0; i < 8; i++) {
        for (int j = i + 1; j < 8; j++) {
            double dx = (double)((i & 1) - (j & 1));
            double dy = (double)(((i >> 1) & 1) - ((j >> 1) & 1));
            double dz = (double)(((i >> 2) & 1) - ((j >> 2) & 1));
            STRETCH_A[k] = i;
            STRETCH_B[k] = j;
            STRETCH_LEN[k] = (real)(float)sqrt(dx * dx + dy * dy + dz * dz);
            k++;
        }
    }
    tables_ready = 1;
}

/* xpbdCouple, sim/constraints/wgsl/xpbdCouple.wgsl:58-71. */
static void xpbd_couple(real cD, real cH, real aD, real aH, real relax, real rD, real rH,
                        real gx, real gy, real gz, real *lD, real *lH)
{
    real stiffened = (real)1 + relax;
    real a00 = gx * stiffened + aD;
    real a01 = gy * stiffened;
    real a11 = gz * stiffened + aH;
    real det = a00 * a11 - a01 * a01;
    if (det < (real)(float)1e-20) {
        *lD = 0;
        *lH = 0;
        return;
    }
    real bD = -(cD + rD * relax);
    real bH = -(cH + rH * relax);
    *lD = (a11 * bD - a01 * bH) / det;
    *lH = (a00 * bH - a01 * bD) / det;
}

/* One Kuhn tetrahedron of project.wgsl:75-153 on four gathered corners.
 * aD/aH are the compliances already divided by dt^2. */
static void project_tet(v3 *p, const v3 *prev, const real *w, const int *tet, m3 B,
                        real gamma, real aD, real aH, real relax)
{
    int i0 = tet[0], i1 = tet[1], i2 = tet[2], i3 = tet[3];
    real w0 = w[i0], w1 = w[i1], w2 = w[i2], w3 = w[i3];
    if (((w0 + w1) + w2) + w3 == (real)0) return;

    m3 Bt = transpose(B);
    v3 base = p[i0];
    m3 Ds = { { sub(p[i1], base), sub(p[i2], base), sub(p[i3], base) } };
    m3 F = mmul(Ds, B);

    real cDev = SQRT((dot(F.c[0], F.c[0]) + dot(F.c[1], F.c[1])) + dot(F.c[2], F.c[2]));
    real cHyd = det3(F) - gamma;

    real k = cDev > (real)(float)1e-9 ? (real)1 / cDev : (real)0;
    m3 mDev = mmul(mscale(F, k), Bt);
    m3 H = { { cross(F.c[1], F.c[2]), cross(F.c[2], F.c[0]), cross(F.c[0], F.c[1]) } };
    m3 mHyd = mmul(H, Bt);

    v3 d1 = mDev.c[0], d2 = mDev.c[1], d3 = mDev.c[2];
    v3 d0 = neg(add(add(d1, d2), d3));
    v3 h1 = mHyd.c[0], h2 = mHyd.c[1], h3 = mHyd.c[2];
    v3 h0 = neg(add(add(h1, h2), h3));

    real gx = ((w0 * dot(d0, d0) + w1 * dot(d1, d1)) + w2 * dot(d2, d2)) + w3 * dot(d3, d3);
    real gy = ((w0 * dot(d0, h0) + w1 * dot(d1, h1)) + w2 * dot(d2, h2)) + w3 * dot(d3, h3);
    real gz = ((w0 * dot(h0, h0) + w1 * dot(h1, h1)) + w2 * dot(h2, h2)) + w3 * dot(h3, h3);

    v3 x0 = sub(p[i0], prev[i0]);
    v3 x1 = sub(p[i1], prev[i1]);
    v3 x2 = sub(p[i2], prev[i2]);
    v3 x3 = sub(p[i3], prev[i3]);
    real rD = ((dot(d0, x0) + dot(d1, x1)) + dot(d2, x2)) + dot(d3, x3);
    real rH = ((dot(h0, x0) + dot(h1, x1)) + dot(h2, x2)) + dot(h3, x3);

    real lD, lH;
    xpbd_couple(cDev, cHyd, aD, aH, relax, rD, rH, gx, gy, gz, &lD, &lH);
    p[i0] = add(p[i0], scale(add(scale(d0, lD), scale(h0, lH)), w0));
    p[i1] = add(p[i1], scale(add(scale(d1, lD), scale(h1, lH)), w1));
    p[i2] = add(p[i2], scale(add(scale(d2, lD), scale(h2, lH)), w2));
    p[i3] = add(p[i3], scale(add(scale(d3, lD), scale(h3, lH)), w3));
}

/* The same projection in the evaluation order the GPU kernel uses. The
 * mathematics is project.wgsl's, rewritten so that it needs no square root,
 * one division, and no 3x3 scaling:
 *
 * - Every Kuhn tetrahedron is an axis path P0 = corner 0, P1, P2, P3 = corner
 *   7, so F = Ds * Dm^-1 = s * E, with E the path's edge differences A_j =
 *   p(P_j+1) - p(P_j) placed by axis and s = 1/h; the cofactor columns are
 *   H = s^2 * K, K = (E1 x E2, E2 x E0, E0 x E1) (K_j below: the column of
 *   step j's axis), and det F = s^3 * det E.
 * - (dC/dF) * B^T gives each path corner the difference of the per-axis
 *   gradients into and out of it: d = c * e with c = s / |E| and e = -A_0,
 *   A_0 - A_1, A_1 - A_2, A_2 by path position, h = s^3 * f with f the same
 *   differences of the K_j.
 * - Sums over the corners are taken over the path's edges instead:
 *   sum_i w_i |e_i|^2 = (W0 + W1)|A0|^2 + (W1 + W2)|A1|^2 + (W2 + W3)|A2|^2
 *   - 2 (W1 A0.A1 + W2 A1.A2) (W_j the weight at path position j), the same
 *   for f over K (its dot products from the edges' by Binet-Cauchy), and
 *   sum_i e_i . (p_i - q_i) = sum_j A_j . (A_j - Q_j) = |E|^2 - sum_j A_j . Q_j
 *   with Q_j = q(P_j+1) - q(P_j) the edge at the substep's start; for f,
 *   sum_j K_j . (A_j - Q_j) = 3 det E - sum_j K_j . Q_j (A_j . K_j = det E).
 *   The subtraction cancels most of |E|^2 (3 det E); the rate enters the
 *   block as r Re beside |E|^2, so the multipliers' error stays far below a
 *   position's rounding.
 * - With u = c * lambda_D and v = s^3 * lambda_H, the first row of the 2x2
 *   block of xpbdCouple.wgsl:58-71 multiplied by 1/c^2 = |E|^2 h^2 and the
 *   second by h^6, the block is symmetric:
 *       [Ge st + aD h^2 |E|^2,  Gef st         ] [u]   [-(|E|^2 + r Re)      ]
 *       [Gef st,                Gf st + aH h^6 ] [v] = [-(det E - gamma h^3 + r Rf)]
 *   (G.. = sum w e.e, e.f, f.f; R. = sum e.x, f.x; st = 1 + r), using
 *   c * C_D = s^2. Its determinant times s^6 is the original one times
 *   |E|^2 h^2, so the singular-block test compares that with 1e-20 h^2
 *   |E|^2. Step j's edge moves by G_j = A_j u + K_j v, and corner P_j by
 *   W_j (G_j-1 - G_j) (G_-1 = G_3 = 0), the correction formed before it
 *   meets the position (one rounding at the position's scale).
 * - A tetrahedron with |F| <= 1e-9 (|E|^2 <= 1e-18 h^2) has no deviatoric
 *   gradient: u = 0, and the block is the original diagonal one scaled,
 *   [aD, 0; 0, Gf st + aH h^6], whose determinant times s^6 is tested
 *   against 1e-20.
 * Products feeding a sum are fused (one rounding), and the block divides
 * once. */
static void project_tet_fused(v3 *p, const v3 *q, const real *w, int t, real h, real gamma, real aD, real aH,
                              real relax)
{
    const int *path = PATH[t];
    real W0 = w[path[0]], W1 = w[path[1]], W2 = w[path[2]], W3 = w[path[3]];
    /* Weights are inverse masses, never negative: the sum is 0 exactly when
     * the WGSL's is. */
    real wend = W0 + W3, wmid = W1 + W2;
    if (wend + wmid == (real)0) return;
    /* Per-body constants (the kernel forms them once per tile). */
    real s = (real)1 / h, s2 = s * s, s3 = s2 * s, s6 = s3 * s3, h2 = h * h, h3 = h2 * h, h6 = h3 * h3;
    real aDh2 = aD * h2, aHh6 = aH * h6, gh3 = gamma * h3;
    real devMin = (real)(float)1e-18 * h2, detMin = (real)(float)1e-20 * h2;
    v3 A[3], E[3], K[3], Kp[3];
    for (int j = 0; j < 3; j++) {
        A[j] = sub(p[path[j + 1]], p[path[j]]);
        E[AXIS[t][j]] = A[j];
    }
    K[0] = crossf(E[1], E[2]);
    K[1] = crossf(E[2], E[0]);
    K[2] = crossf(E[0], E[1]);
    real n[3];
    for (int j = 0; j < 3; j++) {
        Kp[j] = K[AXIS[t][j]];
        n[j] = dotf(A[j], A[j]);
    }
    real d01 = dotf(A[0], A[1]), d12 = dotf(A[1], A[2]), d02 = dotf(A[0], A[2]);
    real e2 = (n[0] + n[1]) + n[2];
    real detE = dotf(E[0], K[0]);
    int dev = e2 > devMin;
    real W01 = W0 + W1, W23 = W2 + W3;
    real ge = FMA(W23, n[2], FMA(wmid, n[1], W01 * n[0]));
    ge = FMA((real)-2, FMA(W2, d12, W1 * d01), ge);
    /* The K_j are the path's edges crossed pairwise (K_0 = +-A1 x A2, K_1 =
     * +-A2 x A0, K_2 = +-A0 x A1, one sign for all three: + when the path's
     * axes run cyclically), so by Binet-Cauchy |K_0|^2 = |A1|^2 |A2|^2 -
     * (A1.A2)^2, K_0.K_1 = (A1.A2)(A0.A2) - (A0.A1)|A2|^2, and so on. */
    real m0 = FMA(n[1], n[2], -(d12 * d12)), m1 = FMA(n[2], n[0], -(d02 * d02)), m2 = FMA(n[0], n[1], -(d01 * d01));
    real k01 = FMA(d12, d02, -(d01 * n[2])), k12 = FMA(d02, d01, -(d12 * n[0]));
    real gf = FMA(W23, m2, FMA(wmid, m1, W01 * m0));
    gf = FMA((real)-2, FMA(W2, k12, W1 * k01), gf);
    v3 Q0[3];
    for (int j = 0; j < 3; j++) Q0[j] = sub(q[path[j + 1]], q[path[j]]);
    real re = e2 - dotacc(A[2], Q0[2], dotacc(A[1], Q0[1], dotf(A[0], Q0[0])));
    real rf = FMA((real)3, detE, -dotacc(Kp[2], Q0[2], dotacc(Kp[1], Q0[1], dotf(Kp[0], Q0[0]))));
    /* e_i . f_i is det E at the path's ends and 2 det E inside it (E_a . K_a =
     * det E, E_a . K_b = 0 for a != b), so Gef = det E * sum of the weights
     * by path position (1, 2, 2, 1). */
    real gef = detE * FMA((real)2, wmid, wend);
    real st = (real)1 + relax;
    real Ab = dev ? FMA(ge, st, aDh2 * e2) : aD;
    real Bb = dev ? gef * st : (real)0;
    real Db = FMA(gf, st, aHh6);
    real det = FMA(Ab, Db, -(Bb * Bb));
    if (det * s6 < (dev ? detMin * e2 : (real)(float)1e-20)) return;
    real P = dev ? -FMA(relax, re, e2) : (real)0;
    real Q = -FMA(relax, rf, detE - gh3);
    real inv = (real)1 / det;
    real u = FMA(Db, P, -(Bb * Q)) * inv;
    real v = FMA(Ab, Q, -(Bb * P)) * inv;
    v3 G[3];
    for (int j = 0; j < 3; j++) G[j] = mk(FMA(Kp[j].x, v, A[j].x * u), FMA(Kp[j].y, v, A[j].y * u), FMA(Kp[j].z, v, A[j].z * u));
    v3 c[4] = { neg(G[0]), sub(G[0], G[1]), sub(G[1], G[2]), G[2] };
    real W[4] = { W0, W1, W2, W3 };
    for (int j = 0; j < 4; j++) {
        v3 *pj = &p[path[j]];
        *pj = mk(FMA(c[j].x, W[j], pj->x), FMA(c[j].y, W[j], pj->y), FMA(c[j].z, W[j], pj->z));
    }
}

/* Kuhn tetrahedron t of a cell of edge h, corners given in KUHN_TETS order:
 * either evaluation (literal = WGSL formulas of project.wgsl, else the fused
 * GPU order). Alphas already over dt^2. */
void o_project_kuhn_tet(const REAL *p12, const REAL *prev12, const REAL *w4, int t, double h, double alpha_dev,
                        double alpha_hyd, double gamma, double relax, int literal, REAL *out12)
{
    build_tables();
    v3 p[8], q[8];
    real w[8] = { 0 };
    for (int c = 0; c < 8; c++) p[c] = q[c] = mk(0, 0, 0);
    for (int i = 0; i < 4; i++) {
        int c = KUHN_TETS[t][i];
        p[c] = mk(p12[i * 3], p12[i * 3 + 1], p12[i * 3 + 2]);
        q[c] = mk(prev12[i * 3], prev12[i * 3 + 1], prev12[i * 3 + 2]);
        w[c] = w4[i];
    }
    real s = (real)1 / (real)h;
    if (literal)
        project_tet(p, q, w, KUHN_TETS[t], mscale(KUHN_INV[t], s), (real)gamma, (real)alpha_dev, (real)alpha_hyd,
                    (real)relax);
    else
        project_tet_fused(p, q, w, t, (real)h, (real)gamma, (real)alpha_dev, (real)alpha_hyd, (real)relax);
    for (int i = 0; i < 4; i++) {
        int c = KUHN_TETS[t][i];
        out12[i * 3] = p[c].x;
        out12[i * 3 + 1] = p[c].y;
        out12[i * 3 + 2] = p[c].z;
    }
}

/* Known-answer entry for one tetrahedron, the shape of `projectCoupled` in the
 * reference's elastic.test.js:111-127 (B column-major; alphas already over dt^2). */
void o_project_tet(const REAL *p12, const REAL *prev12, const REAL *w4, const REAL *B9,
                   double alpha_dev, double alpha_hyd, double gamma, double relax, REAL *out12)
{
    build_tables();
    v3 p[4], prev[4];
    real w[4];
    for (int i = 0; i < 4; i++) {
        p[i] = mk(p12[i * 3], p12[i * 3 + 1], p12[i * 3 + 2]);
        prev[i] = mk(prev12[i * 3], prev12[i * 3 + 1], prev12[i * 3 + 2]);
        w[i] = w4[i];
    }
    m3 B;
    for (int c = 0; c < 3; c++) B.c[c] = mk(B9[c * 3], B9[c * 3 + 1], B9[c * 3 + 2]);
    static const int order[4] = { 0, 1, 2, 3 };
    project_tet(p, prev, w, order, B, (real)gamma, (real)alpha_dev, (real)alpha_hyd, (real)relax);
    for (int i = 0; i < 4; i++) {
        out12[i * 3] = p[i].x;
        out12[i * 3 + 1] = p[i].y;
        out12[i * 3 + 2] = p[i].z;
    }
}

static inline v3 load3(const real *x, const real *y, const real *z, uint32_t n) { return mk(x[n], y[n], z[n]); }

/* project.wgsl:35-184 for one cell; cellCoords/cornerNode from cells.wgsl:39-53. */
static void project_cell(o_scene *sc, const o_body *cp, uint32_t lin, real dt, int limit_only, int literal)
{
    uint32_t cx = lin % cp->cell_dims[0];
    uint32_t cy = (lin / cp->cell_dims[0]) % cp->cell_dims[1];
    uint32_t cz = lin / (cp->cell_dims[0] * cp->cell_dims[1]);
    uint32_t ids[8];
    v3 p[8], prev[8];
    real w[8];
    for (int c = 0; c < 8; c++) {
        uint32_t nx = cx + (uint32_t)(c & 1), ny = cy + (uint32_t)((c >> 1) & 1), nz = cz + (uint32_t)((c >> 2) & 1);
        uint32_t n = cp->node_offset + nx + cp->node_dims[0] * (ny + cp->node_dims[1] * nz);
        ids[c] = n;
        p[c] = load3(sc->px, sc->py, sc->pz, n);
        prev[c] = load3(sc->qx, sc->qy, sc->qz, n);
        w[c] = sc->w[n];
    }
    real maxStretch = (real)cp->max_stretch;
    real h = (real)cp->h;
    if (!limit_only) {
        real invDt2 = (real)1 / (dt * dt);
        real aD = (real)cp->alpha_dev * invDt2;
        real aH = (real)cp->alpha_hyd * invDt2;
        real relax = (real)cp->relaxation / dt;
        real restScale = (real)1 / h;
        real gamma = (real)cp->gamma;
        for (int t = 0; t < 6; t++) {
            if (literal) {
                m3 B = mscale(KUHN_INV[t], restScale);
                project_tet(p, prev, w, KUHN_TETS[t], B, gamma, aD, aH, relax);
            } else {
                project_tet_fused(p, prev, w, t, h, gamma, aD, aH, relax);
            }
        }
    }
    if (maxStretch > (real)1) {
        for (int sweep = 0; sweep < 2; sweep++) {
            for (int e = 0; e < 28; e++) {
                int a = STRETCH_A[e], b = STRETCH_B[e];
                v3 delta = sub(p[b], p[a]);
                real len = length3(delta);
                real maximum = (STRETCH_LEN[e] * h) * maxStretch;
                real weight = w[a] + w[b];
                if (len > maximum && weight > (real)0) {
                    v3 corr = scale(delta, (len - maximum) / (len * weight));
                    p[a] = add(p[a], scale(corr, w[a]));
                    p[b] = sub(p[b], scale(corr, w[b]));
                }
            }
        }
    }
    for (int c = 0; c < 8; c++) {
        sc->px[ids[c]] = p[c].x;
        sc->py[ids[c]] = p[c].y;
        sc->pz[ids[c]] = p[c].z;
    }
}

/* One colour pass (one `sim-elastic-<c>` dispatch, elasticCells.js:181-196). */
void o_elastic_color(o_scene *sc, int color, int limit_only)
{
    build_tables();
    real dt = (real)sc->dt;
    int start = sc->color_start[color], count = sc->color_count[color];
    for (int s = 0; s < count; s++) {
        uint32_t body = sc->cells[(size_t)(start + s) * 2];
        uint32_t lin = sc->cells[(size_t)(start + s) * 2 + 1];
        const o_body *cp = &sc->bodies[body];
        if (limit_only && !((real)cp->max_stretch > (real)1)) continue;
        project_cell(sc, cp, lin, dt, limit_only, sc->literal);
    }
}

/* predict.wgsl:19-36 */
void o_predict(o_scene *sc)
{
    real dt = (real)sc->dt;
    real gx = (real)sc->gravity[0], gy = (real)sc->gravity[1], gz = (real)sc->gravity[2];
    for (int n = 0; n < sc->n_nodes; n++) {
        if (sc->w[n] == (real)0) continue;
        real vx = sc->vx[n] + gx * dt, vy = sc->vy[n] + gy * dt, vz = sc->vz[n] + gz * dt;
        real px = sc->px[n], py = sc->py[n], pz = sc->pz[n];
        sc->qx[n] = px;
        sc->qy[n] = py;
        sc->qz[n] = pz;
        sc->px[n] = px + vx * dt;
        sc->py[n] = py + vy * dt;
        sc->pz[n] = pz + vz * dt;
        sc->vx[n] = vx;
        sc->vy[n] = vy;
        sc->vz[n] = vz;
    }
}

/* driveHard.wgsl:15-35 and driveCompliant.wgsl:24-41 */
void o_drive(o_scene *sc)
{
    real dt = (real)sc->dt;
    v3 target = mk((real)sc->target[0], (real)sc->target[1], (real)sc->target[2]);
    v3 tvel = mk((real)sc->target_vel[0], (real)sc->target_vel[1], (real)sc->target_vel[2]);
    for (int i = 0; i < sc->n_driven; i++) {
        uint32_t n = sc->d_node[i];
        v3 off = mk(sc->d_off[i * 3], sc->d_off[i * 3 + 1], sc->d_off[i * 3 + 2]);
        v3 goal = add(target, off);
        v3 here = load3(sc->px, sc->py, sc->pz, n);
        if (sc->drive_law == 0) {
            v3 step = scale(tvel, dt);
            v3 toGoal = sub(goal, here);
            v3 next = add(here, dot(step, step) >= dot(toGoal, toGoal) ? toGoal : step);
            sc->qx[n] = here.x;
            sc->qy[n] = here.y;
            sc->qz[n] = here.z;
            sc->px[n] = next.x;
            sc->py[n] = next.y;
            sc->pz[n] = next.z;
            v3 v = sub(next, here);
            sc->vx[n] = v.x / dt;
            sc->vy[n] = v.y / dt;
            sc->vz[n] = v.z / dt;
        } else {
            real dt2 = dt * dt;
            real rt = (real)sc->response_time;
            real tau2 = rt * rt;
            real frac = dt2 / (dt2 + tau2);
            v3 np = add(here, scale(sub(goal, here), frac));
            sc->px[n] = np.x;
            sc->py[n] = np.y;
            sc->pz[n] = np.z;
        }
    }
}

/* gridCoord / bucketAt, contacts.wgsl:84-91 */
static void grid_coord(const o_scene *sc, v3 p, uint32_t out[3])
{
    real o[3] = { (real)sc->grid_origin[0], (real)sc->grid_origin[1], (real)sc->grid_origin[2] };
    real size = (real)sc->bucket_size;
    real q[3] = { p.x, p.y, p.z };
    for (int a = 0; a < 3; a++) {
        real last = (real)sc->grid_dims[a] - (real)1;
        real g = FLOOR((q[a] - o[a]) / size);
        real c = g < (real)0 ? (real)0 : g; /* clamp = min(max(g, 0), last) */
        c = c > last ? last : c;
        out[a] = (uint32_t)c;
    }
}

static uint32_t bucket_at(const o_scene *sc, const uint32_t c[3])
{
    return c[0] + sc->grid_dims[0] * (c[1] + sc->grid_dims[1] * c[2]);
}

/* The four contact passes (bodyContact.js:137-148): bin.wgsl:9-23 into
 * the canonical bucket order sort.wgsl:14-53 establishes (ascending record
 * index, the lowest `bucket_slots` kept), collide.wgsl:16-125, settle.wgsl:20-36.
 * A record's corrections are summed as collide.wgsl sums them, one after
 * another in mate order (literal), or (the GPU's order) row by row: the
 * corrections of each row of three neighbour buckets (dz, dy), in mate
 * order, into a row sum, and the nine row sums in row order into the
 * record's. The same mates and corrections; only the sum's association
 * differs. Hit counts are integers either way. */
void o_contact(o_scene *sc)
{
    int S = sc->n_surf;
    if (S == 0) return;
    uint32_t slots = sc->bucket_slots;
    uint32_t B = sc->grid_dims[0] * sc->grid_dims[1] * sc->grid_dims[2];
    uint32_t *count = (uint32_t *)calloc(B, sizeof(uint32_t));
    uint32_t *slot = (uint32_t *)malloc((size_t)B * slots * sizeof(uint32_t));
    real *delta = (real *)malloc((size_t)S * 4 * sizeof(real));

    for (int s = 0; s < S; s++) {
        uint32_t node = sc->s_node[s];
        uint32_t c[3];
        grid_coord(sc, load3(sc->px, sc->py, sc->pz, node), c);
        uint32_t b = bucket_at(sc, c);
        if (sc->key_out) sc->key_out[s] = b;
        if (count[b] < slots) slot[(size_t)b * slots + count[b]] = (uint32_t)s;
        count[b]++;
    }

    for (int s = 0; s < S; s++) {
        uint32_t node = sc->s_node[s];
        uint32_t body = sc->s_body[s];
        real radius = sc->s_radius[s];
        real wi = sc->w[node];
        v3 pi = load3(sc->px, sc->py, sc->pz, node);
        v3 movedI = sub(pi, load3(sc->qx, sc->qy, sc->qz, node));
        real frictionI = (real)sc->bodies[body].friction;
        v3 sum = mk(0, 0, 0);
        real hits = 0;
        int32_t npairs = 0;
        uint32_t cc[3];
        grid_coord(sc, pi, cc);
        int cxi = (int)cc[0], cyi = (int)cc[1], czi = (int)cc[2];
        int dims[3] = { (int)sc->grid_dims[0], (int)sc->grid_dims[1], (int)sc->grid_dims[2] };
        for (int dz = -1; dz <= 1; dz++) {
            int z = czi + dz;
            if (z < 0 || z >= dims[2]) continue;
            for (int dy = -1; dy <= 1; dy++) {
                int y = cyi + dy;
                if (y < 0 || y >= dims[1]) continue;
                v3 row = mk(0, 0, 0);
                for (int dx = -1; dx <= 1; dx++) {
                    int x = cxi + dx;
                    if (x < 0 || x >= dims[0]) continue;
                    uint32_t bc[3] = { (uint32_t)x, (uint32_t)y, (uint32_t)z };
                    uint32_t b = bucket_at(sc, bc);
                    uint32_t held = count[b] < slots ? count[b] : slots;
                    for (uint32_t k = 0; k < held; k++) {
                        uint32_t other = slot[(size_t)b * slots + k];
                        uint32_t mateBody = sc->s_body[other];
                        if (mateBody == body) continue;
                        uint32_t mate = sc->s_node[other];
                        v3 pj = load3(sc->px, sc->py, sc->pz, mate);
                        v3 d = sub(pi, pj);
                        real dist2 = dot(d, d);
                        real reach = radius + sc->s_radius[other];
                        if (dist2 >= reach * reach) continue;
                        if (sc->pairs && npairs < sc->pair_cap) sc->pairs[(size_t)s * sc->pair_cap + npairs] = other;
                        npairs++;
                        real wj = sc->w[mate];
                        real total = wi + wj;
                        if (total <= (real)0) continue;
                        real share = wi / total;
                        real dist = SQRT(dist2);
                        v3 normal = mk(0, node < mate ? (real)1 : (real)-1, 0);
                        if (dist > (real)(float)1e-9) normal = mk(d.x / dist, d.y / dist, d.z / dist);
                        real depth = reach - dist;
                        v3 corr = scale(normal, depth * share);
                        v3 sliding = sub(movedI, sub(pj, load3(sc->qx, sc->qy, sc->qz, mate)));
                        v3 tangent = sub(sliding, scale(normal, dot(sliding, normal)));
                        real slid = length3(tangent);
                        if (slid > (real)(float)1e-9) {
                            real fj = (real)sc->bodies[mateBody].friction;
                            real mu = frictionI < fj ? frictionI : fj;
                            real lim = (mu * depth) / slid;
                            corr = sub(corr, scale(tangent, share * (lim < (real)1 ? lim : (real)1)));
                        }
                        if (sc->literal) sum = add(sum, corr);
                        else row = add(row, corr);
                        hits += (real)1;
                    }
                }
                if (!sc->literal) sum = add(sum, row);
            }
        }
        if (sc->pair_count) sc->pair_count[s] = npairs;
        delta[s * 4] = sum.x;
        delta[s * 4 + 1] = sum.y;
        delta[s * 4 + 2] = sum.z;
        delta[s * 4 + 3] = hits;
    }

    real gain = (real)sc->gain;
    for (int s = 0; s < S; s++) {
        real h = delta[s * 4 + 3];
        if (h > (real)0) {
            uint32_t node = sc->s_node[s];
            real f = gain / h;
            sc->px[node] = sc->px[node] + delta[s * 4] * f;
            sc->py[node] = sc->py[node] + delta[s * 4 + 1] * f;
            sc->pz[node] = sc->pz[node] + delta[s * 4 + 2] * f;
        }
    }
    free(count);
    free(slot);
    free(delta);
}

/* resolvePlane, finalize.wgsl:29-42 */
static v3 resolve_plane(v3 p, v3 prev, v3 normal, real depth, real friction)
{
    if (depth <= (real)0) return p;
    v3 q = add(p, scale(normal, depth));
    v3 delta = sub(q, prev);
    v3 tangent = sub(delta, scale(normal, dot(delta, normal)));
    real len = length3(tangent);
    if (len > (real)(float)1e-9) {
        real lim = (friction * depth) / len;
        q = sub(q, scale(tangent, lim < (real)1 ? lim : (real)1));
    }
    return q;
}

/* The sphere collider, an extension with no reference counterpart: a node
 * inside the ball is moved out along the radius to the surface (one on the
 * centre leaves along +y), with resolvePlane's friction (finalize.wgsl:29-42)
 * in the tangent plane there. */
static v3 resolve_sphere(v3 p, v3 prev, v3 c, real r, real friction)
{
    v3 d = sub(p, c);
    real len = length3(d);
    if (!(len < r)) return p;
    v3 n = len > (real)(float)1e-9 ? mk(d.x / len, d.y / len, d.z / len) : mk(0, 1, 0);
    return resolve_plane(p, prev, n, r - len, friction);
}

/* bodyOfNode, state.wgsl:39-47 */
static uint32_t body_of_node(const o_scene *sc, uint32_t node)
{
    for (int b = 0; b + 1 < sc->n_bodies; b++) {
        if (node < sc->bodies[b].node_offset + sc->bodies[b].node_count) return (uint32_t)b;
    }
    return (uint32_t)(sc->n_bodies - 1);
}

/* The damping factor exp(-damping * dt) of finalize.wgsl:72: the f32 product
 * of the two uniforms, exponentiated once in double and rounded. */
static real damp_factor(double damping, double dtd)
{
    real a = -((real)damping * (real)dtd);
    return (real)exp((double)a);
}

/* finalize.wgsl:44-74 */
void o_finalize(o_scene *sc)
{
    real dt = (real)sc->dt;
    real tmin[3] = { (real)sc->tank_min[0], (real)sc->tank_min[1], (real)sc->tank_min[2] };
    real tmax[3] = { (real)sc->tank_max[0], (real)sc->tank_max[1], (real)sc->tank_max[2] };
    for (int n = 0; n < sc->n_nodes; n++) {
        if (sc->w[n] == (real)0) continue;
        const o_body *bd = &sc->bodies[body_of_node(sc, (uint32_t)n)];
        real fr = (real)bd->friction;
        v3 prev = load3(sc->qx, sc->qy, sc->qz, (uint32_t)n);
        v3 p = load3(sc->px, sc->py, sc->pz, (uint32_t)n);
        p = resolve_plane(p, prev, mk(0, 1, 0), tmin[1] - p.y, fr);
        p = resolve_plane(p, prev, mk(1, 0, 0), tmin[0] - p.x, fr);
        p = resolve_plane(p, prev, mk(-1, 0, 0), p.x - tmax[0], fr);
        p = resolve_plane(p, prev, mk(0, 0, 1), tmin[2] - p.z, fr);
        p = resolve_plane(p, prev, mk(0, 0, -1), p.z - tmax[2], fr);
        if ((real)sc->sphere[3] > (real)0) {
            p = resolve_sphere(p, prev, mk((real)sc->sphere[0], (real)sc->sphere[1], (real)sc->sphere[2]),
                               (real)sc->sphere[3], fr);
        }
        sc->px[n] = p.x;
        sc->py[n] = p.y;
        sc->pz[n] = p.z;
        /* (p - prev) / dt * exp(-damping dt) as (p - prev) * (exp(-damping dt) / dt),
         * the quotient rounded once: one product a component on the GPU. */
        real rate = damp_factor(bd->damping, sc->dt) / dt;
        v3 v = sub(p, prev);
        sc->vx[n] = v.x * rate;
        sc->vy[n] = v.y * rate;
        sc->vz[n] = v.z * rate;
    }
}

/* Stepper.sequence (stepper.js:101-108) in the solve order of simulator.js:83:
 * predict, [drive], elastic colours (+3 stretch sweeps when any body limits
 * stretch, elasticCells.js:197-214), [contact], finalize. */
void o_substep(o_scene *sc)
{
    build_tables();
    o_predict(sc);
    if (sc->n_driven > 0) o_drive(sc);
    int limited = 0;
    for (int b = 0; b < sc->n_bodies; b++) limited |= sc->bodies[b].max_stretch > 1.0;
    for (int c = 0; c < 8; c++) {
        if (sc->color_count[c] > 0) o_elastic_color(sc, c, 0);
    }
    if (limited) {
        for (int sweep = 0; sweep < 3; sweep++) {
            for (int c = 0; c < 8; c++) {
                if (sc->color_count[c] > 0) o_elastic_color(sc, c, 1);
            }
        }
    }
    if (sc->contact && sc->n_surf > 0) o_contact(sc);
    o_finalize(sc);
}

void o_substeps(o_scene *sc, int count)
{
    for (int i = 0; i < count; i++) o_substep(sc);
}

/* deformationAt (sample.wgsl:40-80) followed by deformedCovariance
 * (covariance.wgsl:30-42) for every splat. */
void o_ellipsoid_axes(const REAL cov6[6], REAL rot[4], REAL scale3[3]);

/* a + l * d, per component, one rounding each. */
static inline v3 lerpf(v3 a, v3 d, real l) { return mk(FMA(l, d.x, a.x), FMA(l, d.y, a.y), FMA(l, d.z, a.z)); }

/* deformationAt (sample.wgsl:40-80) and deformedCovariance
 * (covariance.wgsl:18-42) per splat, in the evaluation order the GPU uses.
 * The trilinear sum over the eight corners and its gradient are taken apart
 * axis by axis (the same mathematics: the weight of corner (fx, fy, fz) is
 * f_x f_y f_z with f = local or 1 - local, and its gradient the same product
 * with one factor replaced by +-1/h): along x each corner pair gives its
 * difference D and its interpolant P = p0 + lx D; along y the P give E and
 * Q, the D give Dy; along z, Q gives the position and, with E and Dy, the
 * three gradient columns. The rotation R(q) (xyzw), A = F R diag(s) and
 * A A^T are sums of products, each fused. */
void o_deform_splats(const o_scene *sc, const o_splats *sp)
{
    for (int i = 0; i < sp->n; i++) {
        const o_body *bd = &sc->bodies[sp->body[i]];
        real h = (real)bd->h;
        real inv = (real)1 / h;
        real org[3] = { (real)bd->origin[0], (real)bd->origin[1], (real)bd->origin[2] };
        real rest[3] = { sp->mx[i], sp->my[i], sp->mz[i] };
        real local[3];
        uint32_t base[3];
        for (int a = 0; a < 3; a++) {
            real grid = (rest[a] - org[a]) * inv;
            real lastc = (real)bd->cell_dims[a] - (real)1;
            real cell = FLOOR(grid);
            cell = cell < (real)0 ? (real)0 : cell;
            cell = cell > lastc ? lastc : cell;
            local[a] = grid - cell;
            base[a] = (uint32_t)cell;
        }
        uint32_t nx = bd->node_dims[0], nxy = bd->node_dims[0] * bd->node_dims[1];
        uint32_t n0 = bd->node_offset + base[0] + nx * base[1] + nxy * base[2];
        v3 P[4], D[4]; /* j = fy + 2 fz */
        for (int j = 0; j < 4; j++) {
            uint32_t n = n0 + ((j & 1) ? nx : 0u) + ((j & 2) ? nxy : 0u);
            v3 p0 = load3(sc->px, sc->py, sc->pz, n), p1 = load3(sc->px, sc->py, sc->pz, n + 1);
            D[j] = sub(p1, p0);
            P[j] = lerpf(p0, D[j], local[0]);
        }
        v3 E[2], Q[2], Dy[2];
        for (int fz = 0; fz < 2; fz++) {
            E[fz] = sub(P[1 + 2 * fz], P[2 * fz]);
            Q[fz] = lerpf(P[2 * fz], E[fz], local[1]);
            Dy[fz] = lerpf(D[2 * fz], sub(D[1 + 2 * fz], D[2 * fz]), local[1]);
        }
        v3 Z = sub(Q[1], Q[0]);
        v3 pos = lerpf(Q[0], Z, local[2]);
        v3 col0 = scale(lerpf(Dy[0], sub(Dy[1], Dy[0]), local[2]), inv);
        v3 col1 = scale(lerpf(E[0], sub(E[1], E[0]), local[2]), inv);
        v3 col2 = scale(Z, inv);
        sp->ox[i] = pos.x;
        sp->oy[i] = pos.y;
        sp->oz[i] = pos.z;

        real x = sp->quat[i * 4], y = sp->quat[i * 4 + 1], z = sp->quat[i * 4 + 2], w = sp->quat[i * 4 + 3];
        v3 r0 = mk(FMA((real)-2, FMA(y, y, z * z), (real)1), (real)2 * FMA(x, y, z * w), (real)2 * FMA(x, z, -(y * w)));
        v3 r1 = mk((real)2 * FMA(x, y, -(z * w)), FMA((real)-2, FMA(x, x, z * z), (real)1), (real)2 * FMA(y, z, x * w));
        v3 r2 = mk((real)2 * FMA(x, z, y * w), (real)2 * FMA(y, z, -(x * w)), FMA((real)-2, FMA(x, x, y * y), (real)1));
        /* gradient * basis[k] = col0*b.x + col1*b.y + col2*b.z, then * scale_k */
        v3 a0 = scale(lerpf(lerpf(scale(col0, r0.x), col1, r0.y), col2, r0.z), sp->sx[i]);
        v3 a1 = scale(lerpf(lerpf(scale(col0, r1.x), col1, r1.y), col2, r1.z), sp->sy[i]);
        v3 a2 = scale(lerpf(lerpf(scale(col0, r2.x), col1, r2.y), col2, r2.z), sp->sz[i]);
        /* a * transpose(a): (r, c) = a0[r]a0[c] + a1[r]a1[c] + a2[r]a2[c] */
        size_t n = (size_t)sp->n;
        sp->cov[0 * n + i] = FMA(a2.x, a2.x, FMA(a1.x, a1.x, a0.x * a0.x));
        sp->cov[1 * n + i] = FMA(a2.x, a2.y, FMA(a1.x, a1.y, a0.x * a0.y));
        sp->cov[2 * n + i] = FMA(a2.x, a2.z, FMA(a1.x, a1.z, a0.x * a0.z));
        sp->cov[3 * n + i] = FMA(a2.y, a2.y, FMA(a1.y, a1.y, a0.y * a0.y));
        sp->cov[4 * n + i] = FMA(a2.y, a2.z, FMA(a1.y, a1.z, a0.y * a0.z));
        sp->cov[5 * n + i] = FMA(a2.z, a2.z, FMA(a1.z, a1.z, a0.z * a0.z));
        if (sp->rot) {
            REAL c6[6], q[4], s3[3];
            for (int k = 0; k < 6; k++) c6[k] = sp->cov[(size_t)k * n + i];
            o_ellipsoid_axes(c6, q, s3);
            for (int k = 0; k < 4; k++) sp->rot[(size_t)k * n + i] = q[k];
            for (int k = 0; k < 3; k++) sp->axis[(size_t)k * n + i] = s3[k];
        }
    }
}

/* A covariance taken apart into the engine's splat format: a unit rotation
 * (xyzw) and three axis lengths. Cyclic Jacobi, four sweeps of the three
 * cyclic planes (p, p+1, other = p+2) in order, skipping an exactly-zero
 * off-diagonal; the smaller angle, tangent t = sign(theta)/(|theta| +
 * sqrt(theta^2+1)); the plane rotation joined onto the accumulated
 * quaternion through the half-angle tangent -sin/(1+cos) about +other, and
 * renormalised; axis lengths sqrt(max(eigenvalue, 0))
 * (apps/simulator/src/render/ellipsoid.wgsl:32-91, ellipsoid.js:40-149).
 *
 * The symmetric matrix is held as diagonal d[k] and off-diagonal o[k] of the
 * pair (k, k+1 mod 3), so plane p reads o[p] and writes o[other] and o[r].
 * The quaternion product with (0.., half, 1) keeps only its non-zero terms
 * (zero products change no non-zero sum), and normalisation multiplies by the
 * reciprocal length. Arithmetic otherwise as written, in REAL. */
void o_ellipsoid_axes(const REAL cov6[6], REAL rot[4], REAL scale3[3])
{
    real d[3] = { cov6[0], cov6[3], cov6[5] };
    real o[3] = { cov6[1], cov6[4], cov6[2] };  /* (0,1), (1,2), (2,0) */
    real q[4] = { 0, 0, 0, 1 };
    for (int sweep = 0; sweep < 4; sweep++) {
        for (int p = 0; p < 3; p++) {
            const int r = (p + 1) % 3, other = (p + 2) % 3;
            const real apr = o[p];
            if (apr == (real)0) continue;
            const real app = d[p], arr = d[r];
            const real theta = (arr - app) / ((real)2 * apr);
            const real sign = theta >= (real)0 ? (real)1 : (real)-1;
            const real tangent = sign / (FABS(theta) + SQRT(theta * theta + (real)1));
            const real cosine = (real)1 / SQRT(tangent * tangent + (real)1);
            const real sine = tangent * cosine;
            const real apo = o[other], aro = o[r];
            o[other] = cosine * apo - sine * aro;
            o[r] = sine * apo + cosine * aro;
            d[p] = app - tangent * apr;
            d[r] = arr + tangent * apr;
            o[p] = 0;
            const real half = -(sine / ((real)1 + cosine));
            real x = q[0], y = q[1], z = q[2], w = q[3];
            real nx, ny, nz, nw;
            if (other == 0) {        /* q * (half, 0, 0, 1) */
                nx = w * half + x; ny = y + z * half; nz = z - y * half; nw = w - x * half;
            } else if (other == 1) { /* q * (0, half, 0, 1) */
                nx = x - z * half; ny = w * half + y; nz = x * half + z; nw = w - y * half;
            } else {                 /* q * (0, 0, half, 1) */
                nx = x + y * half; ny = y - x * half; nz = w * half + z; nw = w - z * half;
            }
            const real inv = (real)1 / SQRT(((nx * nx + ny * ny) + nz * nz) + nw * nw);
            q[0] = nx * inv; q[1] = ny * inv; q[2] = nz * inv; q[3] = nw * inv;
        }
    }
    for (int k = 0; k < 4; k++) rot[k] = q[k];
    for (int k = 0; k < 3; k++) scale3[k] = SQRT(d[k] > (real)0 ? d[k] : (real)0);
}

/* nearestNodeToRay (packages/visionary-sim/src/sim/pick.js:41-68): the
 * grabbable node (mass > 0; `mass` is the material inverse mass with static
 * nodes zeroed) with the smallest depth along the ray among those in front of
 * it and within `radius` of the line, the lower index on equal depth; -1 for
 * a miss. f64 arithmetic over the f32 positions, as the JS does it. */
int64_t o_pick_ray(const float *px, const float *py, const float *pz, const float *mass, uint32_t n,
                   const double origin[3], const double dir[3], double radius, double *depth_out)
{
    const double r2 = radius * radius;
    int64_t best = -1;
    double best_depth = INFINITY;
    for (uint32_t i = 0; i < n; i++) {
        if (!(mass[i] > 0.0f)) continue;
        const double dx = (double)px[i] - origin[0], dy = (double)py[i] - origin[1], dz = (double)pz[i] - origin[2];
        const double depth = dx * dir[0] + dy * dir[1] + dz * dir[2];
        if (depth <= 0 || depth >= best_depth) continue;
        const double off = dx * dx + dy * dy + dz * dz - depth * depth;
        if (off > r2) continue;
        best = i;
        best_depth = depth;
    }
    if (depth_out) *depth_out = best < 0 ? 0.0 : best_depth;
    return best;
}

typedef struct { double d2; uint32_t node; } grip_cand;

static int grip_cmp(const void *a, const void *b)
{
    const grip_cand *x = (const grip_cand *)a, *y = (const grip_cand *)b;
    if (x->d2 != y->d2) return x->d2 < y->d2 ? -1 : 1;
    return x->node < y->node ? -1 : (x->node > y->node);
}

/* withinRadius (grip.js:99-138): every grabbable node of [begin, end) within
 * `radius` of the anchor, nearest first with ties by index, truncated to
 * `limit` entries (anchor included, first, zero offset). Returns the entries
 * written; *found counts all of them (0 when the anchor is not grabbable). */
uint32_t o_grip_radius(const float *px, const float *py, const float *pz, const float *mass, uint32_t begin,
                       uint32_t end, uint32_t anchor, double radius, uint32_t limit, uint32_t *nodes,
                       double *offsets, uint32_t *found)
{
    *found = 0;
    if (!(mass[anchor] > 0.0f)) return 0;
    const double ax = px[anchor], ay = py[anchor], az = pz[anchor];
    const double r2 = radius * radius;
    grip_cand *c = (grip_cand *)malloc(sizeof(grip_cand) * (end - begin + 1));
    uint32_t m = 0;
    for (uint32_t i = begin; i < end; i++) {
        if (i == anchor || !(mass[i] > 0.0f)) continue;
        const double dx = (double)px[i] - ax, dy = (double)py[i] - ay, dz = (double)pz[i] - az;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 <= r2) {
            c[m].d2 = d2;
            c[m].node = i;
            m++;
        }
    }
    qsort(c, m, sizeof(grip_cand), grip_cmp);
    *found = m + 1;
    const uint32_t keep = m < (limit > 0 ? limit - 1 : 0) ? m : (limit > 0 ? limit - 1 : 0);
    nodes[0] = anchor;
    offsets[0] = offsets[1] = offsets[2] = 0.0;
    for (uint32_t k = 0; k < keep; k++) {
        const uint32_t i = c[k].node;
        nodes[1 + k] = i;
        offsets[3 * (1 + k)] = (double)px[i] - ax;
        offsets[3 * (1 + k) + 1] = (double)py[i] - ay;
        offsets[3 * (1 + k) + 2] = (double)pz[i] - az;
    }
    free(c);
    return keep + 1;
}

/* Residuals of one cell's six tetrahedra at the current positions, for the
 * error report: out[12] = (C_D, C_H) per tet. */
void o_cell_residuals(const o_scene *sc, uint32_t body, uint32_t lin, REAL *out12)
{
    build_tables();
    const o_body *cp = &sc->bodies[body];
    uint32_t cx = lin % cp->cell_dims[0];
    uint32_t cy = (lin / cp->cell_dims[0]) % cp->cell_dims[1];
    uint32_t cz = lin / (cp->cell_dims[0] * cp->cell_dims[1]);
    v3 p[8];
    for (int c = 0; c < 8; c++) {
        uint32_t n = cp->node_offset + (cx + (uint32_t)(c & 1)) +
                     cp->node_dims[0] * ((cy + (uint32_t)((c >> 1) & 1)) + cp->node_dims[1] * (cz + (uint32_t)((c >> 2) & 1)));
        p[c] = load3(sc->px, sc->py, sc->pz, n);
    }
    real rs = (real)1 / (real)cp->h;
    for (int t = 0; t < 6; t++) {
        const int *tet = KUHN_TETS[t];
        m3 B = mscale(KUHN_INV[t], rs);
        m3 Ds = { { sub(p[tet[1]], p[tet[0]]), sub(p[tet[2]], p[tet[0]]), sub(p[tet[3]], p[tet[0]]) } };
        m3 F = mmul(Ds, B);
        out12[t * 2] = SQRT((dot(F.c[0], F.c[0]) + dot(F.c[1], F.c[1])) + dot(F.c[2], F.c[2]));
        out12[t * 2 + 1] = det3(F) - (real)cp->gamma;
    }
}

/* The tensile limit's edge table (STRETCH_EDGES, stretch.js:5-7) as the
 * cell sweep reads it: corner pairs and unit-cell lengths, 28 entries. */
void o_stretch_edges(int32_t *a, int32_t *b, double *len)
{
    build_tables();
    for (int e = 0; e < 28; e++) {
        a[e] = STRETCH_A[e];
        b[e] = STRETCH_B[e];
        len[e] = (double)STRETCH_LEN[e];
    }
}

int o_real_size(void) { return (int)sizeof(REAL); }
