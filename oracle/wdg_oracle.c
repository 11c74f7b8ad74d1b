/*
 * wdg_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Plain-C restatement of the reference's 2D wave DG path with approximate
 * (collocated Gauss-Lobatto) quadrature.  Each routine names the reference
 * file:line it follows.  It keeps the reference's data layouts and the order
 * of its floating-point sums, so that its results can stand in for the
 * reference (which cannot be built in this image: its headers include a
 * cmake-generated wdg_config.hpp and it links LAPACK).
 *
 * Differences from the reference, all deliberate:
 *   - absorbing BC computes p_ext after u_ext (the reference reads u_ext
 *     before writing it, WaveEquation.hpp:77-84);
 *   - BC code 2 (Dirichlet p = 0) is added;
 *   - edge keys are 64-bit (the reference's `min + nv*max` in int overflows
 *     above ~46k vertices, Mesh2D.cpp:34-38);
 *   - the eigen-decomposition of n.A (LAPACK dgeevx_/dgels_ in
 *     EdgeFlux.cpp:7-89) is done with a cyclic Jacobi rotation method, which is
 *     valid because n.A is symmetric for the wave equation (checked at run
 *     time);
 *   - Legendre polynomials use the Bonnet recurrence (the reference calls
 *     std::legendre); Gauss-Lobatto/Legendre node tables are those of
 *     QuadratureRule.cpp:37-46,103-112 (n <= 9 / n <= 10 only).
 */
#include "wdg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ===================================================================== */
/* basis                                                                 */
/* ===================================================================== */

/* Bonnet recurrence for P_l(x). */
static double legendre_p(int l, double x)
{
    if (l == 0) return 1.0;
    double pm2 = 1.0, pm1 = x, p = x;
    for (int k = 2; k <= l; ++k) {
        p = ((2.0 * k - 1.0) * x * pm1 - (k - 1.0) * pm2) / k;
        pm2 = pm1;
        pm1 = p;
    }
    return p;
}

/* P_l'(x) for |x| < 1 via (x^2-1) P_l' = l (x P_l - P_{l-1}). */
static double legendre_dp(int l, double x)
{
    return l * (x * legendre_p(l, x) - legendre_p(l - 1, x)) / (x * x - 1.0);
}

/* Node table: QuadratureRule.cpp:37-46 (values as printed there). */
static const double GLL_NODES[10][9] = {
    {0}, {0},
    {-1, 1},
    {-1, 0, 1},
    {-1, -0.447213595499958, 0.447213595499958, 1},
    {-1, -0.654653670707977, 0, 0.654653670707977, 1},
    {-1, -0.765055323929465, -0.285231516480645, 0.285231516480645, 0.765055323929465, 1},
    {-1, -0.830223896278567, -0.468848793470714, 0.0, 0.468848793470714, 0.830223896278567, 1},
    {-1, -0.871740148509607, -0.591700181433142, -0.209299217902479, 0.2092992179024789, 0.591700181433142, 0.871740148509607, 1},
    {-1, -0.899757995411460, -0.677186279510738, -0.363117463826178, 0, 0.363117463826178, 0.677186279510738, 0.899757995411460, 1}};

/* QuadratureRule.cpp:32-97: nodes from the table, w_i = 2/(n(n-1) P_{n-1}(x_i)^2) (:94). */
int orc_gauss_lobatto(int n, double *x, double *w)
{
    if (n < 2 || n > 9) return -1;
    for (int i = 0; i < n; ++i) {
        x[i] = GLL_NODES[n][i];
        const double P = legendre_p(n - 1, x[i]);
        w[i] = 2.0 / (n * (n - 1) * (P * P));
    }
    return 0;
}

/* Node table: QuadratureRule.cpp:103-112. */
static const double GL_NODES[11][10] = {
    {0},
    {0.0},
    {-0.577350269189625764509149, 0.577350269189625764509149},
    {-0.774596669241483377035853, 0.0, 0.774596669241483377035853},
    {-0.861136311594052575223946, -0.339981043584856264802666, 0.339981043584856264802666, 0.861136311594052575223946},
    {-0.906179845938663992797627, -0.538469310105683091036314, 0.0, 0.538469310105683091036314, 0.906179845938663992797627},
    {-0.932469514203152027812302, -0.661209386466264513661400, -0.238619186083196908630502, 0.238619186083196908630502, 0.661209386466264513661400, 0.932469514203152027812302},
    {-0.949107912342758524526190, -0.741531185599394439863865, -0.405845151377397166906606, 0.0, 0.405845151377397166906606, 0.741531185599394439863865, 0.949107912342758524526190},
    {-0.960289856497536231683561, -0.796666477413626739591554, -0.525532409916328985817739, -0.183434642495649804939476, 0.183434642495649804939476, 0.525532409916328985817739, 0.796666477413626739591554, 0.960289856497536231683561},
    {-0.968160239507626089835576, -0.836031107326635794299430, -0.613371432700590397308702, -0.324253423403808929038538, 0.0, 0.324253423403808929038538, 0.613371432700590397308702, 0.836031107326635794299430, 0.968160239507626089835576},
    {-0.973906528517171720077964, -0.865063366688984510732097, -0.679409568299024406234327, -0.433395394129247190799266, -0.148874338981631210884826, 0.148874338981631210884826, 0.433395394129247190799266, 0.679409568299024406234327, 0.865063366688984510732097, 0.973906528517171720077964}};

/* QuadratureRule.cpp:99-162: w_i = 2 / ((1 - x_i^2) P_n'(x_i)^2). */
/* Gauss-Legendre rules: the reference's table for n <= 10 (QuadratureRule.cpp:
 * 103-112); above, the roots of P_n by Newton's method from the Chebyshev-like
 * guesses cos(pi (k + 3/4) / (n + 1/2)), ascending (the reference computes
 * them by Golub-Welsch, QuadratureRule.cpp:114-162; both converge to the
 * roots to rounding). */
int orc_gauss_legendre(int n, double *x, double *w)
{
    if (n < 1 || n > 32) return -1;
    for (int i = 0; i < n; ++i) {
        if (n <= 10) {
            x[i] = GL_NODES[n][i];
        } else {
            double t = -cos(M_PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; ++it) {
                const double dt = legendre_p(n, t) / legendre_dp(n, t);
                t -= dt;
                if (fabs(dt) < 1e-16) break;
            }
            x[i] = t;
        }
        const double dP = legendre_dp(n, x[i]);
        w[i] = 2.0 / (1.0 - x[i] * x[i]) / (dP * dP);
    }
    return 0;
}

/* lagrange_interpolation.cpp:52-72 */
static void barycentric_weights(double *w, const double *x, int n)
{
    for (int i = 0; i < n; ++i) {
        w[i] = 1.0;
        for (int j = 0; j < n; ++j)
            if (i != j) w[i] *= x[i] - x[j];
        w[i] = 1.0 / w[i];
    }
    double wmin = w[0], wmax = w[0];
    for (int i = 1; i < n; ++i) {
        if (w[i] < wmin) wmin = w[i];
        if (w[i] > wmax) wmax = w[i];
    }
    const double diff = wmax - wmin;
    for (int i = 0; i < n; ++i) w[i] /= diff;
}

/* lagrange_interpolation.cpp:74-94 */
static double lagrange_interp(double t, const double *x, const double *w, const double *y, int n)
{
    double A = 0.0, B = 0.0;
    const double eps = 2.220446049250313e-16;
    for (int i = 0; i < n; ++i) {
        const double xdiff = t - x[i];
        if (t == x[i] || fabs(xdiff) <= eps) return y[i];
        const double C = w[i] / xdiff;
        A += C * y[i];
        B += C;
    }
    return A / B;
}

/* lagrange_interpolation.cpp:96-139 */
static double lagrange_deriv(double x0, const double *x, const double *w, const double *y, int n)
{
    int atnode = 0, i = 0;
    double A = 0.0, B = 0.0;
    const double p = lagrange_interp(x0, x, w, y, n);
    const double eps = 2.220446049250313e-16;
    for (int j = 0; j < n; ++j) {
        if (x0 == x[j] || fabs(x0 - x[j]) <= eps) {
            atnode = 1;
            B = -w[j];
            i = j;
        }
    }
    if (atnode) {
        for (int j = 0; j < n; ++j)
            if (j != i) A += w[j] * (p - y[j]) / (x0 - x[j]);
    } else {
        for (int j = 0; j < n; ++j) {
            const double t = w[j] / (x0 - x[j]);
            A += t * (p - y[j]) / (x0 - x[j]);
            B += t;
        }
    }
    return A / B;
}

/* lagrange_interpolation.cpp:14-31: B(i, j) = phi_j(xeval_i), shape (neval, ngrid). */
void orc_lagrange_basis(double *B, int ngrid, const double *xgrid, int neval, const double *xeval)
{
    double w[64], y[64];
    barycentric_weights(w, xgrid, ngrid);
    for (int j = 0; j < ngrid; ++j) y[j] = 0.0;
    for (int j = 0; j < ngrid; ++j) {
        y[j] = 1.0;
        for (int i = 0; i < neval; ++i) B[i + neval * j] = lagrange_interp(xeval[i], xgrid, w, y, ngrid);
        y[j] = 0.0;
    }
}

/* lagrange_interpolation.cpp:33-50: D(i, j) = phi_j'(xeval_i), shape (neval, ngrid). */
void orc_lagrange_basis_deriv(double *D, int ngrid, const double *xgrid, int neval, const double *xeval)
{
    double w[64], y[64];
    barycentric_weights(w, xgrid, ngrid);
    for (int j = 0; j < ngrid; ++j) y[j] = 0.0;
    for (int j = 0; j < ngrid; ++j) {
        y[j] = 1.0;
        for (int i = 0; i < neval; ++i) D[i + neval * j] = lagrange_deriv(xeval[i], xgrid, w, y, ngrid);
        y[j] = 0.0;
    }
}

/* ===================================================================== */
/* mesh                                                                  */
/* ===================================================================== */

typedef struct {
    int el[2], side[2], delta, type; /* type: 0 interior, 1 boundary */
    double n[2], meas, x0[2], dx[2];
} orc_edge;

struct orc_mesh {
    int nv, nel, ne, nI, nB;
    double *corners; /* (2, 4, nel) */
    orc_edge *edges;
    int *interior, *boundary;
};

/* Mesh2D.cpp:10-11 */
static const int EMAP1[4] = {0, 1, 3, 0};
static const int EMAP2[4] = {1, 2, 2, 3};

/* Edge.hpp:107-122 (StraightEdge constructor). */
static void straight_edge(orc_edge *E, const double *x0, const double *x1, int side)
{
    E->x0[0] = x0[0];
    E->x0[1] = x0[1];
    E->dx[0] = x1[0] - x0[0];
    E->dx[1] = x1[1] - x0[1];
    const double s = hypot(E->dx[0], E->dx[1]);
    const double sgn = (side == 2 || side == 3) ? -1.0 : 1.0;
    E->n[0] = sgn * E->dx[1] / s;
    E->n[1] = -sgn * E->dx[0] / s;
    E->meas = s / 2;
}

/* Mesh2D.cpp:5-92 (from_vertices), with a 64-bit open-addressing edge map. */
orc_mesh *orc_mesh_from_vertices(int nv, const double *x, int nel, const int *elems)
{
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    m->nv = nv;
    m->nel = nel;
    m->corners = (double *)malloc(sizeof(double) * 8 * (size_t)nel);
    for (int el = 0; el < nel; ++el)
        for (int i = 0; i < 4; ++i) {
            const int c = elems[i + 4 * el];
            m->corners[0 + 2 * (i + 4 * el)] = x[2 * c];
            m->corners[1 + 2 * (i + 4 * el)] = x[2 * c + 1];
        }

    size_t cap = 16;
    while (cap < (size_t)8 * nel + 16) cap <<= 1;
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * cap);
    int *vals = (int *)malloc(sizeof(int) * cap);
    for (size_t i = 0; i < cap; ++i) keys[i] = -1;

    m->edges = (orc_edge *)calloc((size_t)4 * nel + 1, sizeof(orc_edge));
    int ne = 0;
    for (int el = 0; el < nel; ++el) {
        for (int s = 0; s < 4; ++s) {
            const int C0 = elems[EMAP1[s] + 4 * el];
            const int C1 = elems[EMAP2[s] + 4 * el];
            const int64_t lo = C0 < C1 ? C0 : C1, hi = C0 < C1 ? C1 : C0;
            const int64_t k = lo + (int64_t)nv * hi;
            size_t h = (size_t)((uint64_t)k * 0x9E3779B97F4A7C15ull) & (cap - 1);
            while (keys[h] != -1 && keys[h] != k) h = (h + 1) & (cap - 1);
            if (keys[h] == -1) {
                keys[h] = k;
                vals[h] = ne;
                orc_edge *E = &m->edges[ne++];
                straight_edge(E, x + 2 * C0, x + 2 * C1, s);
                E->el[0] = el;
                E->side[0] = s;
                E->el[1] = -1;
                E->side[1] = -1;
                E->type = 1;
                E->delta = 1;
            } else {
                orc_edge *E = &m->edges[vals[h]];
                const int e0 = E->el[0], s0 = E->side[0];
                const int n1 = elems[EMAP1[s0] + 4 * e0];
                E->el[1] = el;
                E->side[1] = s;
                E->type = 0;
                E->delta = (C0 == n1) ? 1 : -1;
            }
        }
    }
    free(keys);
    free(vals);
    m->ne = ne;
    m->interior = (int *)malloc(sizeof(int) * (ne + 1));
    m->boundary = (int *)malloc(sizeof(int) * (ne + 1));
    for (int e = 0; e < ne; ++e) {
        if (m->edges[e].type == 1) m->boundary[m->nB++] = e;
        else m->interior[m->nI++] = e;
    }
    return m;
}

void orc_mesh_free(orc_mesh *m)
{
    if (!m) return;
    free(m->corners);
    free(m->edges);
    free(m->interior);
    free(m->boundary);
    free(m);
}

int orc_mesh_n_elem(const orc_mesh *m) { return m->nel; }
int orc_mesh_n_edges(const orc_mesh *m, int type) { return type == 0 ? m->nI : m->nB; }

static const orc_edge *edge_of(const orc_mesh *m, int i, int type)
{
    return &m->edges[type == 0 ? m->interior[i] : m->boundary[i]];
}

void orc_mesh_edges(const orc_mesh *m, int type, int *out)
{
    const int ne = orc_mesh_n_edges(m, type);
    for (int i = 0; i < ne; ++i) {
        const orc_edge *E = edge_of(m, i, type);
        out[5 * i + 0] = E->el[0];
        out[5 * i + 1] = E->side[0];
        out[5 * i + 2] = E->el[1];
        out[5 * i + 3] = E->side[1];
        out[5 * i + 4] = E->delta;
    }
}

void orc_mesh_edge_geometry(const orc_mesh *m, int type, double *out)
{
    const int ne = orc_mesh_n_edges(m, type);
    for (int i = 0; i < ne; ++i) {
        const orc_edge *E = edge_of(m, i, type);
        out[5 * i + 0] = E->n[0];
        out[5 * i + 1] = E->n[1];
        out[5 * i + 2] = E->meas;
        out[5 * i + 3] = E->x0[0] + E->dx[0] * 0.5; /* Edge.hpp:142-148 at xi = 0 */
        out[5 * i + 4] = E->x0[1] + E->dx[1] * 0.5;
    }
}

/* Mesh2D.cpp:751-767 */
double orc_mesh_min_h(const orc_mesh *m)
{
    double h = INFINITY;
    for (int e = 0; e < m->ne; ++e) {
        const double len = 2.0 * m->edges[e].meas;
        if (len < h) h = len;
    }
    return h;
}

/* Element.cpp:21-27 */
static void quad_jacobian(const double *c, double xi, double eta, double *J)
{
    /* c = (2, 4) corners: c[0 + 2*k], c[1 + 2*k] */
    J[0] = 0.25 * ((1.0 - eta) * (c[2] - c[0]) + (1.0 + eta) * (c[4] - c[6]));
    J[1] = 0.25 * ((1.0 - eta) * (c[3] - c[1]) + (1.0 + eta) * (c[5] - c[7]));
    J[2] = 0.25 * ((1.0 - xi) * (c[6] - c[0]) + (1.0 + xi) * (c[4] - c[2]));
    J[3] = 0.25 * ((1.0 - xi) * (c[7] - c[1]) + (1.0 + xi) * (c[5] - c[3]));
}

/* Element.cpp:5-19 */
static void quad_coords(const double *c, double xi, double eta, double *X)
{
    const double b[4] = {0.25 * (1.0 - xi) * (1.0 - eta), 0.25 * (1.0 + xi) * (1.0 - eta),
                         0.25 * (1.0 + xi) * (1.0 + eta), 0.25 * (1.0 - xi) * (1.0 + eta)};
    X[0] = 0.0;
    X[1] = 0.0;
    for (int i = 0; i < 4; ++i) {
        X[0] += c[2 * i] * b[i];
        X[1] += c[2 * i + 1] * b[i];
    }
}

/* Mesh2D.cpp:643-700 (set_element_metric + physical_coordinates) */
void orc_mesh_nodes(const orc_mesh *m, int nq, const double *xq, double *out)
{
    for (int el = 0; el < m->nel; ++el)
        for (int j = 0; j < nq; ++j)
            for (int i = 0; i < nq; ++i)
                quad_coords(m->corners + 8 * (size_t)el, xq[i], xq[j], out + 2 * (i + nq * (j + (size_t)nq * el)));
}

/* Element.hpp:56-61 (measure = det J) */
void orc_mesh_detj(const orc_mesh *m, int nq, const double *xq, double *out)
{
    for (int el = 0; el < m->nel; ++el)
        for (int j = 0; j < nq; ++j)
            for (int i = 0; i < nq; ++i) {
                double J[4];
                quad_jacobian(m->corners + 8 * (size_t)el, xq[i], xq[j], J);
                out[i + nq * (j + (size_t)nq * el)] = J[0] * J[3] - J[1] * J[2];
            }
}

/* ===================================================================== */
/* wave operator                                                         */
/* ===================================================================== */

/* Wave equation flux tensor A(c, d, k), WaveEquation.cpp:10-18. */
static const double WAVE_A[18] = {0.0, 1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0,
                                  0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0};

/* approx_quad = false tables (exact-quadrature operators) */
struct orc_exact {
    int nqv, nqe, nqm;
    double xv[16], wv[16], Pv[16 * 16], Dv[16 * 16]; /* Div<false> rule; P(i, k) = phi_k(xq_i), (nq, n) */
    double xe[16], we[16], Pe[16 * 16];              /* EdgeFlux<false> rule */
    double *L;                                       /* MassMatrix<false> Cholesky factors (n^2, n^2, nel) */
};

struct orc_wave {
    const orc_mesh *m;
    int n;
    double x[16], w[16], D[256];
    int leg;       /* 1: Gauss-Legendre basis, 0: Gauss-Lobatto */
    double P[16];  /* phi_k(-1), the Legendre face interpolation */
    int *bc;
    double a, b;
    int *v2eI, *v2eB; /* (n, 2, ne) */
    double *FsI, *FsB; /* per edge and side: (3, 3, 2, ne) = Fs(c, d) */
    double *uI, *uB;   /* face vectors (n, 3, 2, ne), FEMVector.cpp:81-87 */
    struct orc_exact *ex; /* approx_quad = false tables, or NULL */
};

/* Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (column-major):
 * A = R diag(e) R^T.  Stands in for real_eig (eig.cpp:29-74). */
static void sym_eig3(const double *A0, double *R, double *e)
{
    double A[9];
    memcpy(A, A0, sizeof(A));
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        const double off = A[3] * A[3] + A[6] * A[6] + A[7] * A[7];
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = A[p + 3 * q];
                if (fabs(apq) < 1e-300) continue;
                const double app = A[p + 3 * p], aqq = A[q + 3 * q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) { /* A <- A G */
                    const double akp = A[k + 3 * p], akq = A[k + 3 * q];
                    A[k + 3 * p] = c * akp - s * akq;
                    A[k + 3 * q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) { /* A <- G^T A */
                    const double apk = A[p + 3 * k], aqk = A[q + 3 * k];
                    A[p + 3 * k] = c * apk - s * aqk;
                    A[q + 3 * k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) { /* R <- R G */
                    const double rkp = R[k + 3 * p], rkq = R[k + 3 * q];
                    R[k + 3 * p] = c * rkp - s * rkq;
                    R[k + 3 * q] = s * rkp + c * rkq;
                }
            }
    }
    for (int i = 0; i < 3; ++i) e[i] = A[i + 3 * i];
}

/* EdgeFlux.cpp:31-89 (_ComputeFlux::flux): Fs = (R diag(a/2 l + sgn b |l|) R^{-1})^T.
 * nA is symmetric here, so R^{-1} = R^T and Fs(c, d) = sum_k R(c,k) coef_k R(d,k). */
static void compute_flux(const double *nA, double a, double b, int side, double *Fs)
{
    const double sgn = (side == 0) ? 1.0 : -1.0;
    if (b == 0.0) {
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) Fs[i + 3 * j] = 0.5 * a * nA[j + 3 * i];
        return;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (nA[i + 3 * j] != nA[j + 3 * i]) abort(); /* symmetric n.A required */
    double R[9], e[3];
    sym_eig3(nA, R, e);
    double coef[3];
    for (int k = 0; k < 3; ++k) coef[k] = 0.5 * a * e[k] + sgn * b * fabs(e[k]);
    for (int d = 0; d < 3; ++d)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += R[c + 3 * k] * coef[k] * R[d + 3 * k];
            Fs[c + 3 * d] = s;
        }
}

void orc_edge_flux(double n0, double n1, double a, double b, int side, double *Fs)
{
    double nA[9];
    for (int d = 0; d < 9; ++d) nA[d] = n0 * WAVE_A[d] + n1 * WAVE_A[9 + d];
    compute_flux(nA, a, b, side, Fs);
}

/* FaceProlongator.cpp:508-514 (mapV2E) */
static int map_v2e(int nc, int i, int f, int el)
{
    const int mm = (f == 0 || f == 2) ? i : (f == 1) ? (nc - 1) : 0;
    const int nn = (f == 1 || f == 3) ? i : (f == 2) ? (nc - 1) : 0;
    return mm + nc * (nn + nc * el);
}

/* FaceProlongator.cpp:497-539 (serial Lobatto constructor) */
static int *make_v2e(const orc_mesh *m, int n, int type)
{
    const int ne = orc_mesh_n_edges(m, type);
    int *v2e = (int *)malloc(sizeof(int) * ((size_t)2 * n * ne + 1));
    for (size_t k = 0; k < (size_t)2 * n * ne; ++k) v2e[k] = -1;
    for (int e = 0; e < ne; ++e) {
        const orc_edge *E = edge_of(m, e, type);
        for (int i = 0; i < n; ++i) v2e[i + n * (0 + 2 * (size_t)e)] = map_v2e(n, i, E->side[0], E->el[0]);
        if (type == 0)
            for (int i = 0; i < n; ++i) {
                const int j = (E->delta > 0) ? i : (n - 1 - i);
                v2e[i + n * (1 + 2 * (size_t)e)] = map_v2e(n, j, E->side[1], E->el[1]);
            }
    }
    return v2e;
}

/* EdgeFlux.cpp:91-139 (setup_flux_2d_cc_c), one Fs per edge and side: straight
 * edges have a constant normal, so every node of an edge gets the same Fs. The
 * node weight w_i ds_i is applied in flux_apply exactly as the stored table
 * F(i,c,d,s,e) = (W(i) ds(i,e)) Fs(c,d) would be. */
static double *make_fs(const orc_mesh *m, int type, double a, double b)
{
    const int ne = orc_mesh_n_edges(m, type);
    double *Fs = (double *)malloc(sizeof(double) * ((size_t)18 * ne + 1));
    for (int e = 0; e < ne; ++e) {
        const orc_edge *E = edge_of(m, e, type);
        double nA[9];
        for (int d = 0; d < 9; ++d) nA[d] = E->n[0] * WAVE_A[d] + E->n[1] * WAVE_A[9 + d];
        for (int s = 0; s < 2; ++s) compute_flux(nA, a, b, s, Fs + 9 * (s + 2 * (size_t)e));
    }
    return Fs;
}

/* WaveEquation.cpp:5-30 with a basis of n Gauss-Lobatto (leg = 0) or
 * Gauss-Legendre (leg = 1) points; the Legendre face interpolation weights
 * P(k) = phi_k(-1) are LegendreFaceProlongator's (FaceProlongator.cpp:717-725). */
orc_wave *orc_wave_create_basis(const orc_mesh *m, int n, int leg, const int *bc, double a, double b)
{
    if (n < 2 || n > 16) return NULL;
    orc_wave *W = (orc_wave *)calloc(1, sizeof(orc_wave));
    W->m = m;
    W->n = n;
    W->a = a;
    W->b = b;
    W->leg = leg ? 1 : 0;
    if ((leg ? orc_gauss_legendre(n, W->x, W->w) : orc_gauss_lobatto(n, W->x, W->w)) != 0) {
        free(W);
        return NULL;
    }
    {
        const double xm1 = -1.0;
        orc_lagrange_basis(W->P, n, W->x, 1, &xm1);
    }
    orc_lagrange_basis_deriv(W->D, n, W->x, n, W->x);
    W->bc = (int *)malloc(sizeof(int) * (m->nB + 1));
    for (int i = 0; i < m->nB; ++i) W->bc[i] = bc ? bc[i] : 1;
    W->v2eI = make_v2e(m, n, 0);
    W->v2eB = make_v2e(m, n, 1);
    W->FsI = make_fs(m, 0, a, b);
    W->FsB = make_fs(m, 1, a, b);
    W->uI = (double *)malloc(sizeof(double) * ((size_t)6 * n * m->nI + 1));
    W->uB = (double *)malloc(sizeof(double) * ((size_t)6 * n * m->nB + 1));
    return W;
}

orc_wave *orc_wave_create(const orc_mesh *m, int n, const int *bc, double a, double b)
{
    return orc_wave_create_basis(m, n, 0, bc, a, b);
}

void orc_wave_free(orc_wave *W)
{
    if (!W) return;
    free(W->bc);
    free(W->v2eI);
    free(W->v2eB);
    free(W->FsI);
    free(W->FsB);
    free(W->uI);
    free(W->uB);
    if (W->ex) free(W->ex->L);
    free(W->ex);
    free(W);
}

long long orc_wave_ndof(const orc_wave *W) { return 3LL * W->n * W->n * W->m->nel; }

/* Div.cpp:5-53 (op table, built here per element on the fly) and
 * Div.cpp:131-179 (div_2d_c): du(d,i,j) += sum_l D(l,i) F(l,d,j) + D(l,j) G(l,d,i). */
static void div_2d(const orc_wave *W, const double *u, double *du)
{
    const int n = W->n, c2d = n * n, nel = W->m->nel;
    const double *D = W->D, *w = W->w, *x = W->x;
#pragma omp parallel
    {
        double Fq[2 * 16 * 3 * 16];
        double op[2 * 3 * 3];
#pragma omp for schedule(static)
        for (int el = 0; el < nel; ++el) {
            const double *c = W->m->corners + 8 * (size_t)el;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const int node = i + n * j;
                    double J[4];
                    quad_jacobian(c, x[i], x[j], J);
                    const double Wt = w[node / n] * w[node % n];
                    const double Y_eta = J[3] * Wt, X_eta = J[2] * Wt;
                    const double Y_xi = J[1] * Wt, X_xi = J[0] * Wt;
                    for (int d = 0; d < 3; ++d)
                        for (int cc = 0; cc < 3; ++cc) {
                            const double a0 = WAVE_A[cc + 3 * d], a1 = WAVE_A[9 + cc + 3 * d];
                            op[0 + 2 * (d + 3 * cc)] = Y_eta * a0 - X_eta * a1;
                            op[1 + 2 * (d + 3 * cc)] = -Y_xi * a0 + X_xi * a1;
                        }
                    const double *uij = u + 3 * (node + (size_t)c2d * el);
                    for (int d = 0; d < 3; ++d) {
                        double Fd = 0.0, Gd = 0.0;
                        for (int cc = 0; cc < 3; ++cc) {
                            Fd += op[0 + 2 * (cc + 3 * d)] * uij[cc];
                            Gd += op[1 + 2 * (cc + 3 * d)] * uij[cc];
                        }
                        Fq[0 + 2 * (i + n * (d + 3 * j))] = Fd;
                        Fq[1 + 2 * (j + n * (d + 3 * i))] = Gd;
                    }
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int d = 0; d < 3; ++d) {
                        double divF = 0.0;
                        for (int l = 0; l < n; ++l)
                            divF += D[l + n * i] * Fq[0 + 2 * (l + n * (d + 3 * j))] +
                                    D[l + n * j] * Fq[1 + 2 * (l + n * (d + 3 * i))];
                        du[d + 3 * (i + n * (j + (size_t)n * el))] += divF;
                    }
        }
    }
}

static void div_2d_q(const orc_wave *W, const double *u, double *du);
void orc_wave_div(const orc_wave *W, const double *u, double *du)
{
    if (W->ex)
        div_2d_q(W, u, du);
    else
        div_2d(W, u, du);
}

/* DivF2D<true>::action (DivF.hpp:207-274) for the linear flux functor
 * f(x, u) = (A u, B u): per node the physical flux Fq(d, 0) = (A u)_d,
 * Fq(d, 1) = (B u)_d, the contravariant fluxes F(0, i, d, j) = Y_eta Fq(d, 0)
 * - X_eta Fq(d, 1) and F(1, j, d, i) = -Y_xi Fq(d, 0) + X_xi Fq(d, 1) (times
 * w_i w_j), then divF(d, i, j) += sum_l D(l, i) F(0, l, d, j) + D(l, j) F(1, l, d, i).
 * A and B are (nvar, nvar) column-major; u, divF are (nvar, n, n, nel). */
void orc_divf_linear(const orc_wave *W, int nvar, const double *A, const double *B, const double *u, double *divF)
{
    const int n = W->n, nn = n * n, nel = W->m->nel;
    const double *D = W->D, *w = W->w, *x = W->x;
#pragma omp parallel
    {
        double *F = (double *)malloc(sizeof(double) * 2 * (size_t)nn * nvar);
        double Fq[2 * 16];
#pragma omp for schedule(static)
        for (int el = 0; el < nel; ++el) {
            const double *c = W->m->corners + 8 * (size_t)el;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double J[4];
                    quad_jacobian(c, x[i], x[j], J);
                    const double w_ij = w[i] * w[j];
                    const double Y_eta = J[3] * w_ij, X_eta = J[2] * w_ij;
                    const double Y_xi = J[1] * w_ij, X_xi = J[0] * w_ij;
                    const double *Uq = u + (size_t)nvar * (i + n * (j + (size_t)n * el));
                    for (int d = 0; d < nvar; ++d) {
                        double fx = 0.0, fy = 0.0;
                        for (int e = 0; e < nvar; ++e) {
                            fx += A[d + nvar * e] * Uq[e];
                            fy += B[d + nvar * e] * Uq[e];
                        }
                        Fq[d] = fx;
                        Fq[nvar + d] = fy;
                    }
                    for (int d = 0; d < nvar; ++d) {
                        F[0 + 2 * (i + n * (d + nvar * j))] = Y_eta * Fq[d] - X_eta * Fq[nvar + d];
                        F[1 + 2 * (j + n * (d + nvar * i))] = -Y_xi * Fq[d] + X_xi * Fq[nvar + d];
                    }
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int d = 0; d < nvar; ++d) {
                        double dF = 0.0;
                        for (int l = 0; l < n; ++l)
                            dF += D[l + n * i] * F[0 + 2 * (l + n * (d + nvar * j))] +
                                  D[l + n * j] * F[1 + 2 * (l + n * (d + nvar * i))];
                        divF[d + nvar * (i + n * (j + (size_t)n * el))] += dF;
                    }
        }
        free(F);
    }
}

/* Gauss-Legendre bases (FaceProlongator.cpp:732-738, mapV2E of the Legendre
 * prolongator): the node k rows into the element from the side node j of
 * element side f is j + k * inward_step(f). */
static int inward_step(int n, int f) { return f == 0 ? n : f == 1 ? -1 : f == 2 ? -n : 1; }

/* Value of variable d (of nvar) on element side f at the side node j:
 * u(d, j) for Gauss-Lobatto (lobatto_action_2d, FaceProlongator.cpp:565-590),
 * sum_k u(d, j_k) P(k) for Gauss-Legendre (legendre_action_2d, :809-835). */
static double face_value(const orc_wave *W, const double *u, int nvar, int j, int f, int d)
{
    if (!W->leg) return u[d + nvar * (size_t)j];
    const int st = inward_step(W->n, f);
    double v = 0.0;
    for (int k = 0; k < W->n; ++k) v += u[d + nvar * (size_t)(j + k * st)] * W->P[k];
    return v;
}

/* The transpose: u(d, j) += Y (lobatto_transpose_2d, :650-672) or
 * u(d, j_k) += Y P(k) (legendre_transpose_2d, :908-934). */
static void face_add(const orc_wave *W, double *u, int nvar, int j, int f, int d, double Y)
{
    if (!W->leg) {
        u[d + nvar * (size_t)j] += Y;
        return;
    }
    const int st = inward_step(W->n, f);
    for (int k = 0; k < W->n; ++k) u[d + nvar * (size_t)(j + k * st)] += Y * W->P[k];
}

/* FaceProlongator.cpp:565-590 (lobatto_action_2d) / :809-835 (legendre_action_2d) */
static void prolong(const orc_wave *W, int type, const double *u, double *uf)
{
    const int n = W->n, ne = orc_mesh_n_edges(W->m, type);
    const int *v2e = type == 0 ? W->v2eI : W->v2eB;
    const int n_sides = type == 0 ? 2 : 1;
    memset(uf, 0, sizeof(double) * (size_t)6 * n * ne);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i) {
                const int j = v2e[i + n * (side + 2 * (size_t)e)];
                for (int d = 0; d < 3; ++d) uf[i + n * (d + 3 * (side + 2 * (size_t)e))] = face_value(W, u, 3, j, f, d);
            }
        }
}

/* EdgeFlux.cpp:289-325 (flux_action_2d_c), in place. */
static void flux_apply(const orc_wave *W, int type, double *uf)
{
    const int n = W->n, ne = orc_mesh_n_edges(W->m, type);
    const double *Fs = type == 0 ? W->FsI : W->FsB;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
        const double ds = edge_of(W->m, e, type)->meas;
        double acc[16 * 3];
        for (int k = 0; k < 3 * n; ++k) acc[k] = 0.0;
        for (int s = 0; s < 2; ++s)
            for (int d = 0; d < 3; ++d)
                for (int c = 0; c < 3; ++c)
                    for (int i = 0; i < n; ++i) {
                        const double F = (W->w[i] * ds) * Fs[c + 3 * d + 9 * (s + 2 * (size_t)e)];
                        acc[i + n * d] += F * uf[i + n * (c + 3 * (s + 2 * (size_t)e))];
                    }
        for (int d = 0; d < 3; ++d)
            for (int i = 0; i < n; ++i) {
                uf[i + n * (d + 3 * (0 + 2 * (size_t)e))] = acc[i + n * d];
                uf[i + n * (d + 3 * (1 + 2 * (size_t)e))] = -acc[i + n * d];
            }
    }
}

/* FaceProlongator.cpp:650-672 (lobatto_transpose_2d) / :908-934
 * (legendre_transpose_2d); serial: nodes receive contributions from several
 * edges, in edge order. */
static void lift(const orc_wave *W, int type, const double *uf, double *du)
{
    const int n = W->n, ne = orc_mesh_n_edges(W->m, type);
    const int *v2e = type == 0 ? W->v2eI : W->v2eB;
    const int n_sides = type == 0 ? 2 : 1;
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < 3; ++d) {
                    const int j = v2e[i + n * (side + 2 * (size_t)e)];
                    face_add(W, du, 3, j, f, d, uf[i + n * (d + 3 * (side + 2 * (size_t)e))]);
                }
        }
}

/* ===================================================================== */
/* exact quadrature (approx_quad = false)                                 */
/* ===================================================================== */

/* linalg.cpp:5-30 (chol): lower-triangular factor in place, column-major. */
static int chol(int m, double *a)
{
    for (int k = 0; k < m; ++k) {
        for (int j = k + 1; j < m; ++j) {
            const double s = a[j + m * k] / a[k + m * k];
            for (int i = j; i < m; ++i) a[i + m * j] -= s * a[i + m * k];
        }
        if (a[k + m * k] == 0) return 0;
        const double s = 1.0 / sqrt(a[k + m * k]);
        for (int i = k; i < m; ++i) a[i + m * k] *= s;
    }
    return 1;
}

/* linalg.cpp:32-66 (solve_chol): x (n, m) <- (L L')^{-1} x. */
static void solve_chol(int m, const double *a, int n, double *x)
{
    for (int j = 0; j < m; ++j) {
        for (int d = 0; d < n; ++d) x[d + n * j] /= a[j + m * j];
        for (int i = j + 1; i < m; ++i)
            for (int d = 0; d < n; ++d) x[d + n * i] -= a[i + m * j] * x[d + n * j];
    }
    for (int j = m - 1; j >= 0; --j) {
        for (int d = 0; d < n; ++d) x[d + n * j] /= a[j + m * j];
        for (int i = 0; i < j; ++i)
            for (int d = 0; d < n; ++d) x[d + n * i] -= a[j + m * i] * x[d + n * j];
    }
}

/* linalg.cpp:68-98 (mult_chol): x (n, m) <- L L' x. */
static void mult_chol(int m, const double *a, int n, double *x)
{
    for (int i = 0; i < m; ++i)
        for (int d = 0; d < n; ++d) {
            double y = 0.0;
            for (int j = i; j < m; ++j) y += a[j + m * i] * x[d + n * j];
            x[d + n * i] = y;
        }
    for (int i = m - 1; i >= 0; --i)
        for (int d = 0; d < n; ++d) {
            double y = 0.0;
            for (int j = 0; j <= i; ++j) y += a[i + m * j] * x[d + n * j];
            x[d + n * i] = y;
        }
}

/* Div.cpp:469-588 (div_2d_q) with the constant-coefficient op of
 * Div.cpp:198-240 (op_2d_cc) evaluated per quadrature point. */
static void div_2d_q(const orc_wave *W, const double *u, double *du)
{
    const struct orc_exact *X = W->ex;
    const int n = W->n, nq = X->nqv, nel = W->m->nel;
    const double *P = X->Pv, *D = X->Dv, *w = X->wv;
#pragma omp parallel
    {
        double Pu[16 * 3 * 16], Uq[3 * 16 * 16], Fq[2 * 3 * 16 * 16], Df[16 * 3 * 16], Pg[16 * 3 * 16];
#pragma omp for schedule(static)
        for (int el = 0; el < nel; ++el) {
            const double *c = W->m->corners + 8 * (size_t)el;
            const double *ue = u + 3 * (size_t)n * n * el;
            /* Pu(l, d, i) = sum_k Pt(k, i) u(d, k, l) */
            for (int i = 0; i < nq; ++i)
                for (int l = 0; l < n; ++l)
                    for (int d = 0; d < 3; ++d) {
                        double pu = 0.0;
                        for (int k = 0; k < n; ++k) pu += P[i + nq * k] * ue[d + 3 * (k + n * l)];
                        Pu[l + n * (d + 3 * i)] = pu;
                    }
            /* Uq(d, i, j) = sum_l Pt(l, j) Pu(l, d, i) */
            for (int j = 0; j < nq; ++j)
                for (int i = 0; i < nq; ++i)
                    for (int d = 0; d < 3; ++d) {
                        double uq = 0.0;
                        for (int l = 0; l < n; ++l) uq += P[j + nq * l] * Pu[l + n * (d + 3 * i)];
                        Uq[d + 3 * (i + nq * j)] = uq;
                    }
            /* contravariant flux with op(k, c, d, i, j) = W (Y_eta a(d,c,0) - X_eta a(d,c,1)), ... */
            for (int j = 0; j < nq; ++j)
                for (int i = 0; i < nq; ++i) {
                    double J[4];
                    quad_jacobian(c, X->xv[i], X->xv[j], J);
                    const double Wt = w[i] * w[j];
                    const double X_xi = J[0] * Wt, Y_xi = J[1] * Wt, X_eta = J[2] * Wt, Y_eta = J[3] * Wt;
                    for (int d = 0; d < 3; ++d) {
                        double Fd = 0.0, Gd = 0.0;
                        for (int cc = 0; cc < 3; ++cc) {
                            const double aq0 = WAVE_A[d + 3 * cc], aq1 = WAVE_A[9 + d + 3 * cc];
                            const double op0 = Y_eta * aq0 - X_eta * aq1, op1 = -Y_xi * aq0 + X_xi * aq1;
                            Fd += op0 * Uq[cc + 3 * (i + nq * j)];
                            Gd += op1 * Uq[cc + 3 * (i + nq * j)];
                        }
                        Fq[0 + 2 * (d + 3 * (i + nq * j))] = Fd;
                        Fq[1 + 2 * (d + 3 * (i + nq * j))] = Gd;
                    }
                }
            /* Df(j, d, k) = sum_i D(i, k) F(i, j); Pg(j, d, k) = sum_i P(i, k) G(i, j) */
            for (int j = 0; j < nq; ++j)
                for (int k = 0; k < n; ++k)
                    for (int d = 0; d < 3; ++d) {
                        double df = 0.0, pg = 0.0;
                        for (int i = 0; i < nq; ++i) {
                            df += D[i + nq * k] * Fq[0 + 2 * (d + 3 * (i + nq * j))];
                            pg += P[i + nq * k] * Fq[1 + 2 * (d + 3 * (i + nq * j))];
                        }
                        Df[j + nq * (d + 3 * k)] = df;
                        Pg[j + nq * (d + 3 * k)] = pg;
                    }
            for (int l = 0; l < n; ++l)
                for (int k = 0; k < n; ++k)
                    for (int d = 0; d < 3; ++d) {
                        double divF = 0.0;
                        for (int j = 0; j < nq; ++j)
                            divF += P[j + nq * l] * Df[j + nq * (d + 3 * k)] + D[j + nq * l] * Pg[j + nq * (d + 3 * k)];
                        du[d + 3 * (k + n * (l + (size_t)n * el))] += divF;
                    }
        }
    }
}

/* EdgeFlux.cpp:548-620 (flux_action_2d_q) with the constant-coefficient
 * table of EdgeFlux.cpp:375-420 (F = W(i) ds Fs), on Lobatto face traces. */
static void flux_apply_q(const orc_wave *W, int type, double *uf)
{
    const struct orc_exact *X = W->ex;
    const int n = W->n, nq = X->nqe, ne = orc_mesh_n_edges(W->m, type);
    const double *Fs = type == 0 ? W->FsI : W->FsB, *P = X->Pe;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < ne; ++e) {
        const double ds = edge_of(W->m, e, type)->meas;
        double acc[16 * 3], Uq[3 * 16];
        for (int k = 0; k < 3 * nq; ++k) acc[k] = 0.0;
        for (int s = 0; s < 2; ++s) {
            for (int i = 0; i < nq; ++i)
                for (int d = 0; d < 3; ++d) {
                    double v = 0.0;
                    for (int j = 0; j < n; ++j) v += uf[j + n * (d + 3 * (s + 2 * (size_t)e))] * P[i + nq * j];
                    Uq[d + 3 * i] = v;
                }
            for (int i = 0; i < nq; ++i) {
                const double wq = X->we[i] * ds;
                for (int d = 0; d < 3; ++d) {
                    double flx = 0.0;
                    for (int cc = 0; cc < 3; ++cc)
                        flx += (wq * Fs[cc + 3 * d + 9 * (s + 2 * (size_t)e)]) * Uq[cc + 3 * i];
                    acc[i + nq * d] += flx;
                }
            }
        }
        for (int d = 0; d < 3; ++d)
            for (int i = 0; i < n; ++i) {
                double flx = 0.0;
                for (int j = 0; j < nq; ++j) flx += acc[j + nq * d] * P[j + nq * i];
                uf[i + n * (d + 3 * (0 + 2 * (size_t)e))] = flx;
                uf[i + n * (d + 3 * (1 + 2 * (size_t)e))] = -flx;
            }
    }
}

orc_wave *orc_wave_create_exact_basis(const orc_mesh *m, int n, int leg, const int *bc, double a, double b, int nqv,
                                      int nqe, int nqm)
{
    /* default rules of Div<false>, EdgeFlux<false>, MassMatrix<false> for
     * bilinear elements (Div.cpp:320-327, EdgeFlux.cpp:503-510, MassMatrix.cpp:92-97) */
    if (nqv <= 0) nqv = 1 + ((n - 2) + (n - 1) + 1) / 2;
    if (nqe <= 0) nqe = 1 + (2 * (n - 1) + 2) / 2;
    if (nqm <= 0) nqm = 1 + (2 * (n - 1) + 2) / 2;
    if (nqv > 10 || nqe > 10 || nqm > 10) return NULL;
    orc_wave *W = orc_wave_create_basis(m, n, leg, bc, a, b);
    if (!W) return NULL;
    struct orc_exact *X = (struct orc_exact *)calloc(1, sizeof(struct orc_exact));
    X->nqv = nqv;
    X->nqe = nqe;
    X->nqm = nqm;
    orc_gauss_legendre(nqv, X->xv, X->wv);
    orc_lagrange_basis(X->Pv, n, W->x, nqv, X->xv);
    orc_lagrange_basis_deriv(X->Dv, n, W->x, nqv, X->xv);
    orc_gauss_legendre(nqe, X->xe, X->we);
    orc_lagrange_basis(X->Pe, n, W->x, nqe, X->xe);
    /* MassMatrix.cpp:85-133: lower triangle of M = (B_i B_j w detJ) summed over
     * the tensor rule, then Cholesky per element */
    double xm[16], wm[16], B[16 * 16];
    orc_gauss_legendre(nqm, xm, wm);
    orc_lagrange_basis(B, n, W->x, nqm, xm);
    const int c2d = n * n, q2d = nqm * nqm, nel = m->nel;
    X->L = (double *)calloc((size_t)c2d * c2d * nel, sizeof(double));
    int ok = 1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
    for (int el = 0; el < nel; ++el) {
        double *M = X->L + (size_t)c2d * c2d * el;
        double detJ[16 * 16];
        for (int p = 0; p < q2d; ++p) {
            double J[4];
            quad_jacobian(m->corners + 8 * (size_t)el, xm[p % nqm], xm[p / nqm], J);
            detJ[p] = J[0] * J[3] - J[1] * J[2];
        }
        for (int j = 0; j < c2d; ++j)
            for (int i = j; i < c2d; ++i) {
                double mij = 0.0;
                for (int p = 0; p < q2d; ++p) {
                    const double bi = B[p % nqm + nqm * (i % n)] * B[p / nqm + nqm * (i / n)];
                    const double bj = B[p % nqm + nqm * (j % n)] * B[p / nqm + nqm * (j / n)];
                    const double w = detJ[p] * wm[p % nqm] * wm[p / nqm];
                    mij += bi * bj * w;
                }
                M[i + c2d * j] = mij;
            }
        ok = ok && chol(c2d, M);
    }
    W->ex = X;
    if (!ok) {
        orc_wave_free(W);
        return NULL;
    }
    return W;
}

orc_wave *orc_wave_create_exact(const orc_mesh *m, int n, const int *bc, double a, double b, int nqv, int nqe,
                                int nqm)
{
    return orc_wave_create_exact_basis(m, n, 0, bc, a, b, nqv, nqe, nqm);
}

/* WaveEquation.cpp:54-68 */
void orc_wave_action(const orc_wave *W, const double *u, double *du)
{
    if (W->ex) {
        div_2d_q(W, u, du);
        prolong(W, 0, u, W->uI);
        flux_apply_q(W, 0, W->uI);
        lift(W, 0, W->uI, du);
        return;
    }
    div_2d(W, u, du);
    prolong(W, 0, u, W->uI);
    flux_apply(W, 0, W->uI);
    lift(W, 0, W->uI, du);
}

/* WaveEquation.hpp:64-69 */
static void reflect_2d(double *v_ext, const double *v_int, const double *n)
{
    v_ext[0] = v_int[0];
    v_ext[1] = (n[1] * n[1] - n[0] * n[0]) * v_int[1] - 2.0 * n[0] * n[1] * v_int[2];
    v_ext[2] = -2.0 * n[0] * n[1] * v_int[1] + (n[0] * n[0] - n[1] * n[1]) * v_int[2];
}

/* WaveEquation.hpp:77-84, with p_ext evaluated after u_ext. */
static void absorb_2d(double *v_ext, const double *v_int, const double *n)
{
    const double outgoing = v_int[0] + n[0] * v_int[1] + n[1] * v_int[2];
    const double tangential_velocity = n[1] * v_int[1] - n[0] * v_int[2];
    v_ext[1] = n[1] * tangential_velocity + 0.5 * n[0] * outgoing;
    v_ext[2] = -n[0] * tangential_velocity + 0.5 * n[1] * outgoing;
    v_ext[0] = n[0] * v_ext[1] + n[1] * v_ext[2];
}

/* Dirichlet p = 0 by the mirror state (no reference counterpart). */
static void dirichlet_2d(double *v_ext, const double *v_int, const double *n)
{
    (void)n;
    v_ext[0] = -v_int[0];
    v_ext[1] = v_int[1];
    v_ext[2] = v_int[2];
}

/* WaveEquation.cpp:144-167 (apply_bc_2d) and :190-215 (WaveBC::action) */
void orc_wavebc_action(const orc_wave *W, const double *u, double *du)
{
    const int n = W->n, nB = W->m->nB;
    double *uB = W->uB;
    prolong(W, 1, u, uB);
    for (int e = 0; e < nB; ++e) {
        const double *nrm = edge_of(W->m, e, 1)->n;
        void (*ext)(double *, const double *, const double *) =
            W->bc[e] == 0 ? absorb_2d : (W->bc[e] == 2 ? dirichlet_2d : reflect_2d);
        for (int i = 0; i < n; ++i) {
            double v_int[3], v_ext[3];
            for (int d = 0; d < 3; ++d) v_int[d] = uB[i + n * (d + 3 * (0 + 2 * (size_t)e))];
            ext(v_ext, v_int, nrm);
            for (int d = 0; d < 3; ++d) uB[i + n * (d + 3 * (1 + 2 * (size_t)e))] = v_ext[d];
        }
    }
    if (W->ex)
        flux_apply_q(W, 1, uB);
    else
        flux_apply(W, 1, uB);
    lift(W, 1, uB, du);
}

/* MassMatrix.cpp:6-28 (m = w_i w_j detJ) and :52-67 (action) */
void orc_mass_action(const orc_wave *W, int nvar, const double *xv, double *y)
{
    const int n = W->n, nel = W->m->nel;
    if (W->ex) { /* MassMatrix.cpp:193-210 */
        const int c2d = n * n;
#pragma omp parallel for schedule(static)
        for (int el = 0; el < nel; ++el) {
            const size_t o = (size_t)nvar * c2d * el;
            for (int i = 0; i < nvar * c2d; ++i) y[o + i] = xv[o + i];
            mult_chol(c2d, W->ex->L + (size_t)c2d * c2d * el, nvar, y + o);
        }
        return;
    }
#pragma omp parallel for schedule(static)
    for (int el = 0; el < nel; ++el)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double J[4];
                quad_jacobian(W->m->corners + 8 * (size_t)el, W->x[i], W->x[j], J);
                const double detJ = J[0] * J[3] - J[1] * J[2];
                const double mm = W->w[i] * W->w[j] * detJ;
                const size_t node = i + n * (j + (size_t)n * el);
                for (int d = 0; d < nvar; ++d) y[d + nvar * node] = mm * xv[d + nvar * node];
            }
}

/* MassMatrix.cpp:69-82 (inv: x /= m) */
void orc_mass_inv(const orc_wave *W, int nvar, double *xv)
{
    const int n = W->n, nel = W->m->nel;
    if (W->ex) { /* MassMatrix.cpp:212-227 */
        const int c2d = n * n;
#pragma omp parallel for schedule(static)
        for (int el = 0; el < nel; ++el)
            solve_chol(c2d, W->ex->L + (size_t)c2d * c2d * el, nvar, xv + (size_t)nvar * c2d * el);
        return;
    }
#pragma omp parallel for schedule(static)
    for (int el = 0; el < nel; ++el)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double J[4];
                quad_jacobian(W->m->corners + 8 * (size_t)el, W->x[i], W->x[j], J);
                const double detJ = J[0] * J[3] - J[1] * J[2];
                const double mm = W->w[i] * W->w[j] * detJ;
                const size_t node = i + n * (j + (size_t)n * el);
                for (int d = 0; d < nvar; ++d) xv[d + nvar * node] /= mm;
            }
}

/* WaveHoltz.cpp:57-70 (time_derivative lambda) */
void orc_rhs(const orc_wave *W, const double *u, double *vt, const double *F, double s)
{
    const long long nd = orc_wave_ndof(W);
    memset(vt, 0, sizeof(double) * (size_t)nd);
    orc_wave_action(W, u, vt);
    orc_wavebc_action(W, u, vt);
    if (F)
        for (long long i = 0; i < nd; ++i) vt[i] += F[i] * s;
    orc_mass_inv(W, 3, vt);
}

/* ode.hpp:66-93 (RungeKutta4::step, Carpenter-Kennedy (5,4) 2N-storage) */
static const double RK4A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double RK4B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};
static const double RK4C[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                               2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

typedef struct {
    double *p, *du;
} rk_work;

static void rk_step(const orc_wave *W, rk_work *wk, double dt, double *t, double *u, const double *F,
                    double omega)
{
    const long long nd = orc_wave_ndof(W);
    double *p = wk->p, *du = wk->du;
    for (int stage = 0; stage < 5; ++stage) {
        const double ts = (stage == 0) ? *t : *t + RK4C[stage] * dt;
        const double s = (F && omega > 0) ? -1.0 / omega * sin(omega * ts) : 0.0;
        orc_rhs(W, u, p, (F && omega > 0) ? F : NULL, s);
        if (stage == 0) {
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < nd; ++i) {
                du[i] = dt * p[i];
                u[i] += RK4B[0] * du[i];
            }
        } else {
            const double a = RK4A[stage], b = RK4B[stage];
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < nd; ++i) {
                du[i] = a * du[i] + dt * p[i];
                u[i] += b * du[i];
            }
        }
    }
    *t += dt;
}

void orc_rk4_steps(const orc_wave *W, double dt, double *t, double *u, int n_steps, const double *F,
                   double omega)
{
    const long long nd = orc_wave_ndof(W);
    rk_work wk;
    wk.p = (double *)malloc(sizeof(double) * (size_t)nd);
    wk.du = (double *)malloc(sizeof(double) * (size_t)nd);
    for (int k = 0; k < n_steps; ++k) rk_step(W, &wk, dt, t, u, F, omega);
    free(wk.p);
    free(wk.du);
}

/* ===================================================================== */
/* WaveHoltz                                                             */
/* ===================================================================== */

struct orc_wh {
    const orc_wave *W;
    double omega, T, dt;
    int nt;
};

/* WaveHoltz.cpp:5-30 */
orc_wh *orc_waveholtz_create(const orc_wave *W, double omega)
{
    orc_wh *H = (orc_wh *)calloc(1, sizeof(orc_wh));
    H->W = W;
    H->omega = omega;
    H->T = 2.0 * M_PI / omega;
    const double h = orc_mesh_min_h(W->m);
    const double p = W->n;
    H->dt = 0.5 * h / (p * p);
    H->nt = (int)ceil(H->T / H->dt);
    H->dt = H->T / H->nt;
    return H;
}

void orc_waveholtz_free(orc_wh *H) { free(H); }

void orc_waveholtz_info(const orc_wh *H, double *T, double *dt, int *nt)
{
    *T = H->T;
    *dt = H->dt;
    *nt = H->nt;
}

/* WaveHoltz.hpp:48-51 */
double orc_waveholtz_K(const orc_wh *H, double t) { return H->omega / M_PI * (cos(H->omega * t) - 0.25); }

/* WaveHoltz.cpp:55-89 */
void orc_waveholtz_pi0(const orc_wh *H, double *u, const double *F)
{
    const long long nd = orc_wave_ndof(H->W);
    double *p = (double *)calloc((size_t)nd, sizeof(double));
    rk_work wk;
    wk.p = (double *)malloc(sizeof(double) * (size_t)nd);
    wk.du = (double *)malloc(sizeof(double) * (size_t)nd);
    for (long long i = 0; i < nd; ++i) u[i] = 0.0;
    double t = 0.0;
    for (int it = 1; it <= H->nt; ++it) {
        rk_step(H->W, &wk, H->dt, &t, p, F, H->omega);
        const double c = ((it == H->nt) ? 0.5 : 1.0) * H->dt * orc_waveholtz_K(H, t);
        for (long long i = 0; i < nd; ++i) u[i] += c * p[i];
    }
    free(p);
    free(wk.p);
    free(wk.du);
}

/* WaveHoltz.cpp:91-121 */
void orc_waveholtz_S(const orc_wh *H, double *u)
{
    const long long nd = orc_wave_ndof(H->W);
    double *p = (double *)malloc(sizeof(double) * (size_t)nd);
    rk_work wk;
    wk.p = (double *)malloc(sizeof(double) * (size_t)nd);
    wk.du = (double *)malloc(sizeof(double) * (size_t)nd);
    const double c0 = 0.5 * H->dt * orc_waveholtz_K(H, 0.0);
    for (long long i = 0; i < nd; ++i) {
        p[i] = u[i];
        u[i] *= c0;
    }
    double t = 0.0;
    for (int it = 1; it <= H->nt; ++it) {
        rk_step(H->W, &wk, H->dt, &t, p, NULL, 0.0);
        const double c = ((it == H->nt) ? 0.5 : 1.0) * H->dt * orc_waveholtz_K(H, t);
        for (long long i = 0; i < nd; ++i) u[i] += c * p[i];
    }
    free(p);
    free(wk.p);
    free(wk.du);
}

/* WaveHoltz.cpp:123-145: H(0,i) = p, H(1,i) = -(1/omega) (M^{-1}(A+B)U)_p */
void orc_waveholtz_postprocess(const orc_wh *H, double *Hout, const double *u)
{
    const long long nd = orc_wave_ndof(H->W), npt = nd / 3;
    double *P = (double *)malloc(sizeof(double) * (size_t)nd);
    orc_rhs(H->W, u, P, NULL, 0.0);
    for (long long i = 0; i < npt; ++i) {
        Hout[2 * i] = u[3 * i];
        Hout[2 * i + 1] = -1.0 / H->omega * P[3 * i];
    }
    free(P);
}

/* ===================================================================== */
/* load vector                                                           */
/* ===================================================================== */


/* ===================================================================== */
/* DivF2D / EdgeFluxF2D with caller functors (C callbacks)                */
/* ===================================================================== */

/* DivF2D<true>::action (DivF.hpp:207-274) with the functor f(x, u, F),
 * F (nvar, 2) column-major, at every node: divF += (F(x, u), grad v). */
void orc_divf_functor(const orc_wave *W, int nvar, orc_flux_fn f, void *ctx, const double *u, double *divF)
{
    const int n = W->n, nel = W->m->nel;
    const double *D = W->D, *w = W->w, *x = W->x;
    double *F = (double *)malloc(sizeof(double) * 2 * (size_t)n * n * nvar);
    double Fq[16], X[2];
    for (int el = 0; el < nel; ++el) {
        const double *c = W->m->corners + 8 * (size_t)el;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                double J[4];
                quad_jacobian(c, x[i], x[j], J);
                quad_coords(c, x[i], x[j], X);
                const double w_ij = w[i] * w[j];
                const double Y_eta = J[3] * w_ij, X_eta = J[2] * w_ij;
                const double Y_xi = J[1] * w_ij, X_xi = J[0] * w_ij;
                for (int d = 0; d < 2 * nvar; ++d) Fq[d] = 0.0;
                f(X, u + (size_t)nvar * (i + n * (j + (size_t)n * el)), Fq, ctx);
                for (int d = 0; d < nvar; ++d) {
                    F[0 + 2 * (i + n * (d + nvar * j))] = Y_eta * Fq[d] - X_eta * Fq[nvar + d];
                    F[1 + 2 * (j + n * (d + nvar * i))] = -Y_xi * Fq[d] + X_xi * Fq[nvar + d];
                }
            }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < nvar; ++d) {
                    double dF = 0.0;
                    for (int l = 0; l < n; ++l)
                        dF += D[l + n * i] * F[0 + 2 * (l + n * (d + nvar * j))] +
                              D[l + n * j] * F[1 + 2 * (l + n * (d + nvar * i))];
                    divF[d + nvar * (i + n * (j + (size_t)n * el))] += dF;
                }
    }
    free(F);
}

/* prol->action, EdgeFluxF2D<true>::action with the functor fh(x, n, uL, uR)
 * (EdgeFluxF.hpp:136-171: fB(side 0) = +fh w ds, fB(side 1) = -fh w ds),
 * a scale (cl2d.cpp:190-194 uses -1) and prol->t (FaceProlongator.cpp:565-672,
 * 809-934).  x is the edge node on el0's side, n el0's outward normal.  On
 * boundary edges (type 1) side 1 is ext(x, n, u) if given, else a copy of
 * side 0 (cl2d.cpp:198-209). */
void orc_edgefluxf_functor(const orc_wave *W, int nvar, orc_nflux_fn fh_fn, orc_ext_fn ext, void *ctx, int type,
                           double scale, const double *u, double *divF)
{
    const int n = W->n, ne = orc_mesh_n_edges(W->m, type);
    const int *v2e = type == 0 ? W->v2eI : W->v2eB;
    const int n_sides = type == 0 ? 2 : 1;
    double *uf = (double *)calloc((size_t)2 * n * nvar * ne + 1, sizeof(double)); /* (n, nvar, 2, ne) */
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i) {
                const int j = v2e[i + n * (side + 2 * (size_t)e)];
                for (int d = 0; d < nvar; ++d)
                    uf[i + n * (d + nvar * (side + 2 * (size_t)e))] = face_value(W, u, nvar, j, f, d);
            }
        }
    double uL[8], uR[8], fh[8], X[2];
    for (int e = 0; e < ne; ++e) {
        const orc_edge *E = edge_of(W->m, e, type);
        const int el0 = E->el[0], s0 = E->side[0];
        const double *c = W->m->corners + 8 * (size_t)el0;
        for (int i = 0; i < n; ++i) {
            /* edge node i in el0's order, at the basis abscissa along side s0 */
            const double xi = (s0 == 0 || s0 == 2) ? W->x[i] : (s0 == 1 ? 1.0 : -1.0);
            const double eta = (s0 == 1 || s0 == 3) ? W->x[i] : (s0 == 2 ? 1.0 : -1.0);
            quad_coords(c, xi, eta, X);
            for (int d = 0; d < nvar; ++d) uL[d] = uf[i + n * (d + nvar * (0 + 2 * (size_t)e))];
            if (type == 0) {
                for (int d = 0; d < nvar; ++d) uR[d] = uf[i + n * (d + nvar * (1 + 2 * (size_t)e))];
            } else if (ext) {
                ext(X, E->n, uL, uR, ctx);
            } else {
                for (int d = 0; d < nvar; ++d) uR[d] = uL[d];
            }
            for (int d = 0; d < nvar; ++d) fh[d] = 0.0;
            fh_fn(X, E->n, uL, uR, fh, ctx);
            for (int d = 0; d < nvar; ++d) {
                const double fl = fh[d] * W->w[i] * E->meas;
                uf[i + n * (d + nvar * (0 + 2 * (size_t)e))] = scale * fl;
                uf[i + n * (d + nvar * (1 + 2 * (size_t)e))] = -scale * fl;
            }
        }
    }
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < nvar; ++d) {
                    const int j = v2e[i + n * (side + 2 * (size_t)e)];
                    face_add(W, divF, nvar, j, f, d, uf[i + n * (d + nvar * (side + 2 * (size_t)e))]);
                }
        }
    free(uf);
}


/* DivF2D<false>::action (DivF.hpp:275-392) with the functor f(x, u, F): u
 * interpolated to the nq x nq Gauss-Legendre points, F~ there with the
 * metric of the bilinear map, projected back with D and P. */
void orc_divf_functor_q(const orc_wave *W, int nvar, int nq, orc_flux_fn f, void *ctx, const double *u, double *divF)
{
    const int n = W->n, nel = W->m->nel;
    double xq[32], wq[32], P[32 * 16], D[32 * 16]; /* (nq, n) column-major */
    orc_gauss_legendre(nq, xq, wq);
    orc_lagrange_basis(P, n, W->x, nq, xq);
    orc_lagrange_basis_deriv(D, n, W->x, nq, xq);
    double *Pu = (double *)malloc(sizeof(double) * (size_t)n * nvar * nq);
    double *F = (double *)malloc(sizeof(double) * 2 * (size_t)nq * nvar * nq);
    double *Df = (double *)malloc(sizeof(double) * (size_t)nq * nvar * n);
    double *Pg = (double *)malloc(sizeof(double) * (size_t)nq * nvar * n);
    double Uq[8], Fq[16], X[2], J[4];
    for (int el = 0; el < nel; ++el) {
        const double *c = W->m->corners + 8 * (size_t)el;
        const double *ue = u + (size_t)nvar * n * n * el;
        for (int i = 0; i < nq; ++i)
            for (int l = 0; l < n; ++l)
                for (int d = 0; d < nvar; ++d) {
                    double pu = 0.0;
                    for (int k = 0; k < n; ++k) pu += P[i + nq * k] * ue[d + nvar * (k + n * l)];
                    Pu[l + n * (d + nvar * i)] = pu;
                }
        for (int j = 0; j < nq; ++j)
            for (int i = 0; i < nq; ++i) {
                quad_jacobian(c, xq[i], xq[j], J);
                quad_coords(c, xq[i], xq[j], X);
                const double w_ij = wq[i] * wq[j];
                const double Y_eta = J[3] * w_ij, X_eta = J[2] * w_ij;
                const double Y_xi = J[1] * w_ij, X_xi = J[0] * w_ij;
                for (int d = 0; d < nvar; ++d) {
                    double uq = 0.0;
                    for (int l = 0; l < n; ++l) uq += P[j + nq * l] * Pu[l + n * (d + nvar * i)];
                    Uq[d] = uq;
                }
                for (int d = 0; d < 2 * nvar; ++d) Fq[d] = 0.0;
                f(X, Uq, Fq, ctx);
                for (int d = 0; d < nvar; ++d) {
                    F[0 + 2 * (i + nq * (d + nvar * j))] = Y_eta * Fq[d] - X_eta * Fq[nvar + d];
                    F[1 + 2 * (i + nq * (d + nvar * j))] = -Y_xi * Fq[d] + X_xi * Fq[nvar + d];
                }
            }
        for (int j = 0; j < nq; ++j)
            for (int k = 0; k < n; ++k)
                for (int d = 0; d < nvar; ++d) {
                    double df = 0.0, pg = 0.0;
                    for (int i = 0; i < nq; ++i) {
                        df += D[i + nq * k] * F[0 + 2 * (i + nq * (d + nvar * j))];
                        pg += P[i + nq * k] * F[1 + 2 * (i + nq * (d + nvar * j))];
                    }
                    Df[j + nq * (d + nvar * k)] = df;
                    Pg[j + nq * (d + nvar * k)] = pg;
                }
        for (int l = 0; l < n; ++l)
            for (int k = 0; k < n; ++k)
                for (int d = 0; d < nvar; ++d) {
                    double dF = 0.0;
                    for (int j = 0; j < nq; ++j)
                        dF += P[j + nq * l] * Df[j + nq * (d + nvar * k)] + D[j + nq * l] * Pg[j + nq * (d + nvar * k)];
                    divF[d + nvar * (k + n * (l + (size_t)n * el))] += dF;
                }
    }
    free(Pu);
    free(F);
    free(Df);
    free(Pg);
}

/* prol->action, EdgeFluxF2D<false>::action (EdgeFluxF.hpp:182-250) with the
 * functor fh(x, n, uL, uR) at nq Gauss-Legendre edge points, a scale and
 * prol->t.  Edge points run in el0's side order. */
void orc_edgefluxf_functor_q(const orc_wave *W, int nvar, int nq, orc_nflux_fn fh_fn, orc_ext_fn ext, void *ctx,
                             int type, double scale, const double *u, double *divF)
{
    const int n = W->n, ne = orc_mesh_n_edges(W->m, type);
    const int *v2e = type == 0 ? W->v2eI : W->v2eB;
    const int n_sides = type == 0 ? 2 : 1;
    double xq[32], wq[32], P[32 * 16];
    orc_gauss_legendre(nq, xq, wq);
    orc_lagrange_basis(P, n, W->x, nq, xq);
    double *uf = (double *)calloc((size_t)2 * n * nvar * ne + 1, sizeof(double)); /* (n, nvar, 2, ne) */
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i) {
                const int j = v2e[i + n * (side + 2 * (size_t)e)];
                for (int d = 0; d < nvar; ++d)
                    uf[i + n * (d + nvar * (side + 2 * (size_t)e))] = face_value(W, u, nvar, j, f, d);
            }
        }
    double uL[8], uR[8], fh[8], X[2], Fq[32 * 8];
    for (int e = 0; e < ne; ++e) {
        const orc_edge *E = edge_of(W->m, e, type);
        const int el0 = E->el[0], s0 = E->side[0];
        const double *c = W->m->corners + 8 * (size_t)el0;
        for (int i = 0; i < nq; ++i) {
            const double xi = (s0 == 0 || s0 == 2) ? xq[i] : (s0 == 1 ? 1.0 : -1.0);
            const double eta = (s0 == 1 || s0 == 3) ? xq[i] : (s0 == 2 ? 1.0 : -1.0);
            quad_coords(c, xi, eta, X);
            for (int d = 0; d < nvar; ++d) {
                double a = 0.0, b = 0.0;
                for (int k = 0; k < n; ++k) {
                    a += P[i + nq * k] * uf[k + n * (d + nvar * (0 + 2 * (size_t)e))];
                    b += P[i + nq * k] * uf[k + n * (d + nvar * (1 + 2 * (size_t)e))];
                }
                uL[d] = a;
                uR[d] = b;
            }
            if (type == 1) {
                if (ext)
                    ext(X, E->n, uL, uR, ctx);
                else
                    for (int d = 0; d < nvar; ++d) uR[d] = uL[d];
            }
            for (int d = 0; d < nvar; ++d) fh[d] = 0.0;
            fh_fn(X, E->n, uL, uR, fh, ctx);
            for (int d = 0; d < nvar; ++d) Fq[i + nq * d] = fh[d];
        }
        for (int d = 0; d < nvar; ++d)
            for (int k = 0; k < n; ++k) {
                double fl = 0.0;
                for (int i = 0; i < nq; ++i) fl += Fq[i + nq * d] * P[i + nq * k] * (wq[i] * E->meas);
                uf[k + n * (d + nvar * (0 + 2 * (size_t)e))] = scale * fl;
                uf[k + n * (d + nvar * (1 + 2 * (size_t)e))] = -scale * fl;
            }
    }
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < nvar; ++d) {
                    const int j = v2e[i + n * (side + 2 * (size_t)e)];
                    face_add(W, divF, nvar, j, f, d, uf[i + n * (d + nvar * (side + 2 * (size_t)e))]);
                }
        }
    free(uf);
}

/* LinearFunctional.hpp:62-74 (constructor) and :136-202 (2D action). */
/* LinearFunctional<2>::action(n_var, f, F) (LinearFunctional.hpp:136-202)
 * with f already evaluated at the mapped Gauss-Legendre(nq) points:
 * fq is (nvar, nq, nq, nel) column-major, fq(d, k, l, el) = f_d(x(xq_k, xq_l)).
 * Each value is weighted by detJ w_k w_l, then projected onto the basis along
 * y (index l -> j) and then along x (k -> i), B(l, j) = phi_j(xq_l).  F is
 * (nvar, n, n, nel). */
void orc_linear_functional(const orc_mesh *m, int n, int nq, int nvar, const double *fq, double *F)
{
    double xb[16], wb[16], xq[24], wq[24], B[16 * 24];
    if (n < 2 || n > 16 || nq < 1 || nq > 24 || nvar < 1 || nvar > 8) {
        if (nvar >= 1 && n >= 1)
            for (size_t q = 0; q < (size_t)nvar * n * n * m->nel; ++q) F[q] = NAN;
        return;
    }
    orc_gauss_lobatto(n, xb, wb);
    orc_gauss_legendre(nq, xq, wq);
    orc_lagrange_basis(B, n, xb, nq, xq); /* B(l, j) = phi_j(xq_l), (nq, n) */
    const int nel = m->nel;
#pragma omp parallel for schedule(static)
    for (int el = 0; el < nel; ++el) {
        const double *c = m->corners + 8 * (size_t)el;
        double Fq[8 * 24 * 24], ProjF[8 * 24 * 16];
        for (int l = 0; l < nq; ++l)
            for (int k = 0; k < nq; ++k) {
                double J[4];
                quad_jacobian(c, xq[k], xq[l], J);
                const double detJ = J[0] * J[3] - J[1] * J[2];
                const double wt = detJ * wq[k] * wq[l];
                for (int d = 0; d < nvar; ++d)
                    Fq[d + nvar * (k + nq * l)] = fq[d + nvar * (k + nq * (l + (size_t)nq * el))] * wt;
            }
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < nq; ++k)
                for (int d = 0; d < nvar; ++d) {
                    double s = 0.0;
                    for (int l = 0; l < nq; ++l) s += Fq[d + nvar * (k + nq * l)] * B[l + nq * j];
                    ProjF[d + nvar * (k + nq * j)] = s;
                }
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < nvar; ++d) {
                    double s = 0.0;
                    for (int k = 0; k < nq; ++k) s += ProjF[d + nvar * (k + nq * j)] * B[k + nq * i];
                    F[d + nvar * (i + n * (j + (size_t)n * el))] = s;
                }
    }
}

/* The same functional for a Gaussian sum in variable `var` of 3: the sources
 * are evaluated at the mapped Gauss-Legendre points (the functor the examples
 * pass, waveholtz2d.cpp:137) and handed to orc_linear_functional. */
void orc_load_vector_gauss(const orc_mesh *m, int n, int nq, int nsrc, const double *src, int var, double *F)
{
    double xq[24], wq[24];
    orc_gauss_legendre(nq, xq, wq);
    const int nel = m->nel;
    const size_t npt = (size_t)nq * nq;
    double *fq = calloc(3 * npt * nel, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int el = 0; el < nel; ++el) {
        const double *c = m->corners + 8 * (size_t)el;
        for (int l = 0; l < nq; ++l)
            for (int k = 0; k < nq; ++k) {
                double X[2];
                quad_coords(c, xq[k], xq[l], X);
                double fv = 0.0;
                for (int s = 0; s < nsrc; ++s) {
                    const double dx = X[0] - src[4 * s], dy = X[1] - src[4 * s + 1], sg = src[4 * s + 2];
                    fv += src[4 * s + 3] * exp(-(dx * dx + dy * dy) / (2.0 * sg * sg));
                }
                fq[var + 3 * (k + nq * (l + (size_t)nq * el))] = fv;
            }
    }
    orc_linear_functional(m, n, nq, 3, fq, F);
    free(fq);
}

/* ===================================================================== */
/* reductions (linalg.cpp:102-130)                                       */
/* ===================================================================== */

double orc_norm(long long n, const double *x)
{
    double r = 0.0;
    for (long long i = 0; i < n; ++i) r += x[i] * x[i];
    return sqrt(r);
}

double orc_error(long long n, const double *x, const double *y)
{
    double r = 0.0;
    for (long long i = 0; i < n; ++i) {
        const double d = x[i] - y[i];
        r += d * d;
    }
    return sqrt(r);
}

/* ===================================================================== */
/* Advection<true>, AdvectionHomogeneousBC<true> (Advection.hpp:15-273)   */
/* ===================================================================== */

/* Cyclic Jacobi eigen-decomposition of a symmetric m x m matrix, m <= 8
 * (column-major): A = R diag(e) R^T.  Stands in for real_eig (eig.cpp:29-74)
 * on the symmetric matrices this restatement accepts. */
static void sym_eig(int m, const double *A0, double *R, double *e)
{
    double A[64];
    memcpy(A, A0, sizeof(double) * m * m);
    for (int i = 0; i < m * m; ++i) R[i] = (i % (m + 1) == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int q = 0; q < m; ++q)
            for (int p = 0; p < q; ++p) off += A[p + m * q] * A[p + m * q];
        if (off < 1e-300) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[p + m * q];
                if (fabs(apq) < 1e-300) continue;
                const double app = A[p + m * p], aqq = A[q + m * q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < m; ++k) {
                    const double akp = A[k + m * p], akq = A[k + m * q];
                    A[k + m * p] = c * akp - s * akq;
                    A[k + m * q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    const double apk = A[p + m * k], aqk = A[q + m * k];
                    A[p + m * k] = c * apk - s * aqk;
                    A[q + m * k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double rkp = R[k + m * p], rkq = R[k + m * q];
                    R[k + m * p] = c * rkp - s * rkq;
                    R[k + m * q] = s * rkp + c * rkq;
                }
            }
    }
    for (int i = 0; i < m; ++i) e[i] = A[i + m * i];
}

/* Eigenvalues of a real m x m matrix (m <= 8, column-major) by the shifted
 * QR iteration on the full matrix, A <- R Q + mu I with A - mu I = Q R
 * (Householder), mu the Wilkinson shift of the trailing 2 x 2 block, and
 * deflation of the last row or the last two rows once their entries left of
 * the diagonal block are negligible.  A 2 x 2 block whose eigenvalues are
 * complex by more than tol ends with 1; no convergence ends with 3. */
static int qr_eigenvalues(int m, const double *A0, double tol, double *e)
{
    double A[64];
    memcpy(A, A0, sizeof(double) * m * m);
    double nrm = 0.0;
    for (int i = 0; i < m * m; ++i) nrm += A[i] * A[i];
    nrm = sqrt(nrm);
    const double small = 4.0 * DBL_EPSILON * nrm;
    int k = m, its = 0;
    while (k > 0) {
        if (k == 1) {
            e[0] = A[0];
            break;
        }
        double off1 = 0.0, off2 = 0.0;
        for (int j = 0; j < k - 1; ++j) off1 += A[(k - 1) + m * j] * A[(k - 1) + m * j];
        for (int j = 0; j < k - 2; ++j)
            off2 += A[(k - 1) + m * j] * A[(k - 1) + m * j] + A[(k - 2) + m * j] * A[(k - 2) + m * j];
        const double p = A[(k - 2) + m * (k - 2)], q = A[(k - 2) + m * (k - 1)];
        const double r = A[(k - 1) + m * (k - 2)], t = A[(k - 1) + m * (k - 1)];
        if (sqrt(off1) <= small) {
            e[k - 1] = t;
            --k;
            its = 0;
            continue;
        }
        const double half = 0.5 * (p + t), disc = 0.25 * (p - t) * (p - t) + q * r;
        if (sqrt(off2) <= small) {
            if (disc < 0.0 && sqrt(-disc) > tol) return 1;
            const double w = disc > 0.0 ? sqrt(disc) : 0.0;
            e[k - 2] = half + w;
            e[k - 1] = half - w;
            k -= 2;
            its = 0;
            continue;
        }
        if (++its > 400) return 3;
        double mu;
        if (its % 11 == 0) {
            mu = t + sqrt(off1); /* exceptional shift */
        } else if (disc >= 0.0) {
            const double w = sqrt(disc);
            mu = fabs(half + w - t) < fabs(half - w - t) ? half + w : half - w;
        } else {
            mu = half;
        }
        /* Householder QR of the k x k block of A - mu I: the reflectors v_c are
         * kept in V; R overwrites B */
        double B[64], V[64], beta[8];
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < k; ++i) B[i + m * j] = A[i + m * j] - (i == j ? mu : 0.0);
        for (int c = 0; c < k; ++c) {
            double tail = 0.0;
            for (int i = c + 1; i < k; ++i) tail += B[i + m * c] * B[i + m * c];
            beta[c] = 0.0;
            if (tail == 0.0) continue;
            const double x0 = B[c + m * c], nx = sqrt(x0 * x0 + tail), alpha = x0 > 0.0 ? -nx : nx;
            for (int i = 0; i < k; ++i) V[i + m * c] = i < c ? 0.0 : (i == c ? x0 - alpha : B[i + m * c]);
            beta[c] = 2.0 / ((x0 - alpha) * (x0 - alpha) + tail);
            for (int j = c; j < k; ++j) {
                double d = 0.0;
                for (int i = c; i < k; ++i) d += V[i + m * c] * B[i + m * j];
                d *= beta[c];
                for (int i = c; i < k; ++i) B[i + m * j] -= d * V[i + m * c];
            }
        }
        /* A <- R Q + mu I, Q = H_0 H_1 ... H_{k-1}: apply the reflectors to R's columns */
        for (int c = 0; c < k; ++c) {
            if (beta[c] == 0.0) continue;
            for (int i = 0; i < k; ++i) {
                double d = 0.0;
                for (int j = c; j < k; ++j) d += B[i + m * j] * V[j + m * c];
                d *= beta[c];
                for (int j = c; j < k; ++j) B[i + m * j] -= d * V[j + m * c];
            }
        }
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < k; ++i) A[i + m * j] = B[i + m * j] + (i == j ? mu : 0.0);
    }
    return 0;
}

/* Null space of the m x m matrix M by Gaussian elimination with complete
 * pivoting, pivots of magnitude <= tol counting as zero: writes the basis
 * (unit columns) into N and returns its dimension. */
static int null_space(int m, const double *M0, double tol, double *N)
{
    double M[64];
    int col[8];
    memcpy(M, M0, sizeof(double) * m * m);
    for (int j = 0; j < m; ++j) col[j] = j;
    int rank = 0;
    for (; rank < m; ++rank) {
        int pi = rank, pj = rank;
        for (int j = rank; j < m; ++j)
            for (int i = rank; i < m; ++i)
                if (fabs(M[i + m * j]) > fabs(M[pi + m * pj])) {
                    pi = i;
                    pj = j;
                }
        if (fabs(M[pi + m * pj]) <= tol) break;
        for (int j = 0; j < m; ++j) {
            const double x = M[rank + m * j];
            M[rank + m * j] = M[pi + m * j];
            M[pi + m * j] = x;
        }
        for (int i = 0; i < m; ++i) {
            const double x = M[i + m * rank];
            M[i + m * rank] = M[i + m * pj];
            M[i + m * pj] = x;
        }
        const int c = col[rank];
        col[rank] = col[pj];
        col[pj] = c;
        for (int i = rank + 1; i < m; ++i) {
            const double f = M[i + m * rank] / M[rank + m * rank];
            for (int j = rank; j < m; ++j) M[i + m * j] -= f * M[rank + m * j];
        }
    }
    /* free variables rank..m-1: set one to 1, back-substitute the pivots */
    for (int f = rank; f < m; ++f) {
        double y[8] = {0.0};
        y[f] = 1.0;
        for (int i = rank - 1; i >= 0; --i) {
            double v = 0.0;
            for (int j = i + 1; j < m; ++j) v -= M[i + m * j] * y[j];
            y[i] = v / M[i + m * i];
        }
        double nn = 0.0;
        for (int j = 0; j < m; ++j) nn += y[j] * y[j];
        nn = sqrt(nn);
        for (int j = 0; j < m; ++j) N[col[j] + m * (f - rank)] = y[j] / nn;
    }
    return m - rank;
}

/* real_eig (eig.cpp:29-74) for a general real n.A: eigenvalues e and right
 * eigenvectors R (columns).  Eigenvalues closer than tol form one repeated
 * eigenvalue whose eigenvectors span the null space of nA - lambda I.
 * Returns 0, 1 for complex eigenvalues, 2 if the eigenvectors of a repeated
 * eigenvalue are fewer than its multiplicity, 3 if the iteration fails. */
static int real_eig_general(int m, const double *nA, double *R, double *e)
{
    double nrm = 0.0;
    for (int i = 0; i < m * m; ++i) nrm += nA[i] * nA[i];
    const double tol = 1e-7 * sqrt(nrm);
    double w[8];
    const int rc = qr_eigenvalues(m, nA, tol, w);
    if (rc) return rc;
    for (int i = 1; i < m; ++i) /* insertion sort */
        for (int j = i; j > 0 && w[j] < w[j - 1]; --j) {
            const double x = w[j];
            w[j] = w[j - 1];
            w[j - 1] = x;
        }
    int nv = 0;
    for (int i = 0; i < m;) {
        int j = i + 1;
        while (j < m && w[j] - w[j - 1] <= tol) ++j;
        double lam = 0.0, M[64], N[64];
        for (int q = i; q < j; ++q) lam += w[q];
        lam /= (j - i);
        for (int q = 0; q < m * m; ++q) M[q] = nA[q] - (q % (m + 1) == 0 ? lam : 0.0);
        const int dim = null_space(m, M, tol, N);
        if (dim < j - i) return 2;
        for (int q = 0; q < j - i; ++q) {
            memcpy(R + m * nv, N + m * q, sizeof(double) * m);
            e[nv++] = lam;
        }
        i = j;
    }
    return 0;
}

/* _ComputeFlux::flux (EdgeFlux.cpp:31-89): Fs(c, d) = (a/2 nA + sgn b |nA|)(d, c).
 * Symmetric n.A: Jacobi eigenvectors, R^{-1} = R^T.  General n.A: the
 * reference's steps, F <- R', F <- diag(a/2 e + sgn b |e|) F, then
 * F <- R^{-T} F (the dgels solve, here elimination with partial pivoting).
 * Returns 0, 1 for complex eigenvalues of n.A, 2 for linearly dependent
 * eigenvectors, 3 if the eigenvalue iteration fails. */
static int compute_flux_m(int m, const double *nA, double a, double b, int side, double *Fs)
{
    const double sgn = (side == 0) ? 1.0 : -1.0;
    if (b == 0.0) {
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) Fs[i + m * j] = 0.5 * a * nA[j + m * i];
        return 0;
    }
    int sym = 1;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            if (nA[i + m * j] != nA[j + m * i]) sym = 0;
    if (!sym) {
        double R[64], e[8], T[64];
        const int rc = real_eig_general(m, nA, R, e);
        if (rc) return rc;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                Fs[i + m * j] = (0.5 * a * e[i] + sgn * b * fabs(e[i])) * R[j + m * i];
                T[i + m * j] = R[j + m * i];
            }
        double tmax = 0.0;
        for (int i = 0; i < m * m; ++i) tmax = fmax(tmax, fabs(T[i]));
        for (int c = 0; c < m; ++c) {
            int piv = c;
            for (int r = c + 1; r < m; ++r)
                if (fabs(T[r + m * c]) > fabs(T[piv + m * c])) piv = r;
            if (fabs(T[piv + m * c]) <= 1e-12 * tmax) return 2;
            for (int j = 0; j < m; ++j) {
                double x = T[c + m * j];
                T[c + m * j] = T[piv + m * j];
                T[piv + m * j] = x;
                x = Fs[c + m * j];
                Fs[c + m * j] = Fs[piv + m * j];
                Fs[piv + m * j] = x;
            }
            for (int r = c + 1; r < m; ++r) {
                const double f = T[r + m * c] / T[c + m * c];
                for (int j = c; j < m; ++j) T[r + m * j] -= f * T[c + m * j];
                for (int j = 0; j < m; ++j) Fs[r + m * j] -= f * Fs[c + m * j];
            }
        }
        for (int c = m - 1; c >= 0; --c)
            for (int j = 0; j < m; ++j) {
                double v = Fs[c + m * j];
                for (int q = c + 1; q < m; ++q) v -= T[c + m * q] * Fs[q + m * j];
                Fs[c + m * j] = v / T[c + m * c];
            }
        return 0;
    }
    double R[64], e[8], coef[8];
    sym_eig(m, nA, R, e);
    for (int k = 0; k < m; ++k) coef[k] = 0.5 * a * e[k] + sgn * b * fabs(e[k]);
    for (int d = 0; d < m; ++d)
        for (int c = 0; c < m; ++c) {
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += R[c + m * k] * coef[k] * R[d + m * k];
            Fs[c + m * d] = s;
        }
    return 0;
}

int orc_flux_matrix(int m, const double *nA, double a, double b, int side, double *Fs)
{
    if (m < 1 || m > 8) return 3;
    return compute_flux_m(m, nA, a, b, side, Fs);
}

struct orc_adv {
    const orc_wave *W; /* mesh, basis and face maps */
    int nvar, cc;
    double *A;         /* (nvar, nvar, 2) or (nvar, nvar, 2, n, n, nel) */
    double *FI, *FB;   /* F(i, c, d, s, e) = W(i) ds Fs(c, d), (n, nvar, nvar, 2, ne) */
    double *uf;        /* face vector (n, nvar, 2, ne) */
};

/* A(k, l, dim) at global node j (Div.cpp:21-37 index _i, _el) */
static const double *coef_at(const struct orc_adv *V, int j)
{
    const int v2 = V->nvar * V->nvar;
    return V->cc ? V->A : V->A + (size_t)2 * v2 * j;
}

/* setup_flux_2d_cc_c / setup_flux_2d_c (EdgeFlux.cpp:91-186): the
 * coefficient on side s is the face-prolonged A of that side; the exterior
 * side of a boundary edge has A = 0. */
static double *adv_flux_table(const struct orc_adv *V, int type, double a, double b, int *err)
{
    const orc_wave *W = V->W;
    const int n = W->n, m = V->nvar, v2 = m * m, ne = orc_mesh_n_edges(W->m, type);
    const int *v2e = type == 0 ? W->v2eI : W->v2eB;
    double *F = (double *)malloc(sizeof(double) * ((size_t)2 * n * v2 * ne + 1));
    double nA[64], Fs[64];
    for (int e = 0; e < ne; ++e) {
        const orc_edge *E = edge_of(W->m, e, type);
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < n; ++i) {
                const int j = v2e[i + n * (s + 2 * (size_t)e)];
                if (j < 0 && !V->cc) {
                    for (int d = 0; d < v2; ++d) nA[d] = 0.0;
                } else if (V->cc || !W->leg) {
                    const double *A = coef_at(V, j < 0 ? 0 : j);
                    for (int d = 0; d < v2; ++d) nA[d] = E->n[0] * A[d] + E->n[1] * A[v2 + d];
                } else {
                    /* the coefficient's face values come from the same
                     * prolongator (Advection.hpp:77-78) */
                    for (int d = 0; d < v2; ++d)
                        nA[d] = E->n[0] * face_value(W, V->A, 2 * v2, j, E->side[s], d) +
                                E->n[1] * face_value(W, V->A, 2 * v2, j, E->side[s], v2 + d);
                }
                if (compute_flux_m(m, nA, a, b, s, Fs) != 0) *err = 1;
                const double w = W->w[i] * E->meas;
                for (int d = 0; d < m; ++d)
                    for (int c = 0; c < m; ++c) F[i + n * (c + m * (d + m * (s + 2 * (size_t)e)))] = w * Fs[c + m * d];
            }
    }
    return F;
}

orc_adv *orc_advection_create(const orc_wave *W, int nvar, const double *A, int constant_coefficient)
{
    if (nvar < 1 || nvar > 8) return NULL;
    orc_adv *V = (orc_adv *)calloc(1, sizeof(orc_adv));
    V->W = W;
    V->nvar = nvar;
    V->cc = constant_coefficient;
    const size_t nA = (size_t)2 * nvar * nvar * (constant_coefficient ? 1 : (size_t)W->n * W->n * W->m->nel);
    V->A = (double *)malloc(sizeof(double) * nA);
    memcpy(V->A, A, sizeof(double) * nA);
    int err = 0;
    /* take negative a, b since -div is computed (Advection.hpp:86-87, 212) */
    V->FI = adv_flux_table(V, 0, -1.0, -0.5, &err);
    V->FB = adv_flux_table(V, 1, -1.0, -0.5, &err);
    const int ne = W->m->nI > W->m->nB ? W->m->nI : W->m->nB;
    V->uf = (double *)malloc(sizeof(double) * ((size_t)2 * W->n * nvar * ne + 1));
    if (err) {
        orc_advection_free(V);
        return NULL;
    }
    return V;
}

void orc_advection_free(orc_adv *V)
{
    if (!V) return;
    free(V->A);
    free(V->FI);
    free(V->FB);
    free(V->uf);
    free(V);
}

/* Div<true>::action with the (nvar, nvar, 2[, n, n, nel]) coefficient
 * (Div.cpp:5-53, 131-179): f_d = sum_c A(d, c, 0) u_c, g_d = sum_c A(d, c, 1) u_c. */
static void adv_div(const orc_adv *V, const double *u, double *divF)
{
    const orc_wave *W = V->W;
    const int n = W->n, nn = n * n, nel = W->m->nel, m = V->nvar, v2 = m * m;
    const double *D = W->D, *w = W->w, *x = W->x;
#pragma omp parallel
    {
        double *F = (double *)malloc(sizeof(double) * 2 * (size_t)nn * m);
#pragma omp for schedule(static)
        for (int el = 0; el < nel; ++el) {
            const double *c = W->m->corners + 8 * (size_t)el;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double J[4];
                    quad_jacobian(c, x[i], x[j], J);
                    const double Wt = w[j] * w[i];
                    const double Y_eta = J[3] * Wt, X_eta = J[2] * Wt;
                    const double Y_xi = J[1] * Wt, X_xi = J[0] * Wt;
                    const int node = i + n * (j + n * el);
                    const double *A = coef_at(V, node);
                    const double *uq = u + (size_t)m * node;
                    for (int d = 0; d < m; ++d) {
                        double Fd = 0.0, Gd = 0.0;
                        for (int cc = 0; cc < m; ++cc) {
                            const double a0 = A[d + m * cc], a1 = A[v2 + d + m * cc];
                            Fd += (Y_eta * a0 - X_eta * a1) * uq[cc];
                            Gd += (-Y_xi * a0 + X_xi * a1) * uq[cc];
                        }
                        F[0 + 2 * (i + n * (d + m * j))] = Fd;
                        F[1 + 2 * (j + n * (d + m * i))] = Gd;
                    }
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int d = 0; d < m; ++d) {
                        double dF = 0.0;
                        for (int l = 0; l < n; ++l)
                            dF += D[l + n * i] * F[0 + 2 * (l + n * (d + m * j))] +
                                  D[l + n * j] * F[1 + 2 * (l + n * (d + m * i))];
                        divF[d + m * (i + n * (j + (size_t)n * el))] += dF;
                    }
        }
        free(F);
    }
}

/* face_prol->action, Flx->action, face_prol->t (Advection.hpp:127-140, 252-270):
 * prolongation (FaceProlongator.cpp:565-590), flux_action_2d_c
 * (EdgeFlux.cpp:289-325) and the transpose (FaceProlongator.cpp:650-672). */
static void adv_faces(const orc_adv *V, int type, const double *u, double *divF)
{
    const orc_wave *W = V->W;
    const int n = W->n, m = V->nvar, ne = orc_mesh_n_edges(W->m, type);
    const int *v2e = type == 0 ? W->v2eI : W->v2eB;
    const int n_sides = type == 0 ? 2 : 1;
    const double *F = type == 0 ? V->FI : V->FB;
    double *uf = V->uf;
    memset(uf, 0, sizeof(double) * (size_t)2 * n * m * ne);
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i) {
                const int j = v2e[i + n * (side + 2 * (size_t)e)];
                for (int d = 0; d < m; ++d)
                    uf[i + n * (d + m * (side + 2 * (size_t)e))] = face_value(W, u, m, j, f, d);
            }
        }
    double acc[16 * 8];
    for (int e = 0; e < ne; ++e) {
        for (int k = 0; k < n * m; ++k) acc[k] = 0.0;
        for (int s = 0; s < 2; ++s)
            for (int d = 0; d < m; ++d)
                for (int c = 0; c < m; ++c)
                    for (int i = 0; i < n; ++i)
                        acc[i + n * d] += F[i + n * (c + m * (d + m * (s + 2 * (size_t)e)))] *
                                          uf[i + n * (c + m * (s + 2 * (size_t)e))];
        for (int d = 0; d < m; ++d)
            for (int i = 0; i < n; ++i) {
                uf[i + n * (d + m * (0 + 2 * (size_t)e))] = acc[i + n * d];
                uf[i + n * (d + m * (1 + 2 * (size_t)e))] = -acc[i + n * d];
            }
    }
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < m; ++d) {
                    const int j = v2e[i + n * (side + 2 * (size_t)e)];
                    face_add(W, divF, m, j, f, d, uf[i + n * (d + m * (side + 2 * (size_t)e))]);
                }
        }
}

void orc_advection_action(const orc_adv *V, const double *u, double *divF)
{
    adv_div(V, u, divF);
    adv_faces(V, 0, u, divF);
}

void orc_advection_bc_action(const orc_adv *V, const double *u, double *divF) { adv_faces(V, 1, u, divF); }

/* ===================================================================== */
/* DivF2D<true> / EdgeFluxF2D<true> with the compressible Euler flux and    */
/* the local Lax-Friedrichs numerical flux of examples/cl2d.cpp:57-111     */
/* ===================================================================== */

/* cl2d.cpp:57-81 (F): q = (rho, rho u, rho v, E), f and g the x and y fluxes */
static void euler_flux(double gamma, const double *q, double *f, double *g)
{
    const double rho = q[0], u = q[1] / q[0], v = q[2] / q[0], E = q[3];
    const double p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    f[0] = rho * u;
    f[1] = rho * u * u + p;
    f[2] = rho * u * v;
    f[3] = u * (E + p);
    g[0] = rho * v;
    g[1] = rho * u * v;
    g[2] = rho * v * v + p;
    g[3] = v * (E + p);
}

/* cl2d.cpp:84-96 (speed) */
static double euler_speed(double gamma, const double *q)
{
    const double rho = q[0], u = q[1] / q[0], v = q[2] / q[0], E = q[3];
    const double p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    return hypot(u, v) + sqrt(gamma * p / rho);
}

/* cl2d.cpp:98-111 (numerical_flux) */
static void euler_llf(double gamma, const double *n, const double *qL, const double *qR, double *fh)
{
    double fL[4], gL[4], fR[4], gR[4];
    euler_flux(gamma, qL, fL, gL);
    euler_flux(gamma, qR, fR, gR);
    const double sL = euler_speed(gamma, qL), sR = euler_speed(gamma, qR);
    const double c = sL > sR ? sL : sR;
    for (int d = 0; d < 4; ++d)
        fh[d] = 0.5 * n[0] * (fL[d] + fR[d]) + 0.5 * n[1] * (gL[d] + gR[d]) + 0.5 * c * (qL[d] - qR[d]);
}

/* DivF2D<true>::action (DivF.hpp:207-274) with the Euler flux: divF += (F(q), grad v) */
void orc_divf_euler(const orc_wave *W, double gamma, const double *u, double *divF)
{
    const int n = W->n, nn = n * n, nel = W->m->nel;
    const double *D = W->D, *w = W->w, *x = W->x;
#pragma omp parallel
    {
        double *F = (double *)malloc(sizeof(double) * 2 * (size_t)nn * 4);
#pragma omp for schedule(static)
        for (int el = 0; el < nel; ++el) {
            const double *c = W->m->corners + 8 * (size_t)el;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    double J[4];
                    quad_jacobian(c, x[i], x[j], J);
                    const double w_ij = w[i] * w[j];
                    const double Y_eta = J[3] * w_ij, X_eta = J[2] * w_ij;
                    const double Y_xi = J[1] * w_ij, X_xi = J[0] * w_ij;
                    double f[4], g[4];
                    euler_flux(gamma, u + 4 * (i + n * (j + (size_t)n * el)), f, g);
                    for (int d = 0; d < 4; ++d) {
                        F[0 + 2 * (i + n * (d + 4 * j))] = Y_eta * f[d] - X_eta * g[d];
                        F[1 + 2 * (j + n * (d + 4 * i))] = -Y_xi * f[d] + X_xi * g[d];
                    }
                }
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i)
                    for (int d = 0; d < 4; ++d) {
                        double dF = 0.0;
                        for (int l = 0; l < n; ++l)
                            dF += D[l + n * i] * F[0 + 2 * (l + n * (d + 4 * j))] +
                                  D[l + n * j] * F[1 + 2 * (l + n * (d + 4 * i))];
                        divF[d + 4 * (i + n * (j + (size_t)n * el))] += dF;
                    }
        }
        free(F);
    }
}

/* The face part of cl2d.cpp's time derivative (cl2d.cpp:180-215):
 * prolongation (FaceProlongator.cpp:565-590), EdgeFluxF2D<true>::action with
 * the LLF flux (EdgeFluxF.hpp:136-171: fB = +-fh w ds), the sign flip, and
 * the transpose (FaceProlongator.cpp:650-672).  type 0: interior edges;
 * type 1: boundary edges with the exterior state a copy of the interior one. */
void orc_edgefluxf_euler(const orc_wave *W, double gamma, int type, const double *u, double *divF)
{
    const int n = W->n, ne = orc_mesh_n_edges(W->m, type);
    const int *v2e = type == 0 ? W->v2eI : W->v2eB;
    const int n_sides = type == 0 ? 2 : 1;
    double *uf = (double *)calloc((size_t)8 * n * ne + 1, sizeof(double)); /* (n, 4, 2, ne) */
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i) {
                const int j = v2e[i + n * (side + 2 * (size_t)e)];
                for (int d = 0; d < 4; ++d)
                    uf[i + n * (d + 4 * (side + 2 * (size_t)e))] = face_value(W, u, 4, j, f, d);
            }
        }
    if (type == 1)
        for (int e = 0; e < ne; ++e)
            for (int d = 0; d < 4; ++d)
                for (int i = 0; i < n; ++i)
                    uf[i + n * (d + 4 * (1 + 2 * (size_t)e))] = uf[i + n * (d + 4 * (0 + 2 * (size_t)e))];
    for (int e = 0; e < ne; ++e) {
        const orc_edge *E = edge_of(W->m, e, type);
        for (int i = 0; i < n; ++i) {
            double qL[4], qR[4], fh[4];
            for (int d = 0; d < 4; ++d) {
                qL[d] = uf[i + n * (d + 4 * (0 + 2 * (size_t)e))];
                qR[d] = uf[i + n * (d + 4 * (1 + 2 * (size_t)e))];
            }
            euler_llf(gamma, E->n, qL, qR, fh);
            for (int d = 0; d < 4; ++d) {
                const double fl = fh[d] * W->w[i] * E->meas;
                uf[i + n * (d + 4 * (0 + 2 * (size_t)e))] = -fl;
                uf[i + n * (d + 4 * (1 + 2 * (size_t)e))] = fl;
            }
        }
    }
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < 4; ++d) {
                    const int j = v2e[i + n * (side + 2 * (size_t)e)];
                    face_add(W, divF, 4, j, f, d, uf[i + n * (d + 4 * (side + 2 * (size_t)e))]);
                }
        }
    free(uf);
}


/* ===================================================================== */
/* Advection<false> (Advection.hpp:59-93, 190-215): Div<false> with the   */
/* coefficient (Div.cpp:312-360, op_2d_cc / op_2d :200-310) and           */
/* EdgeFlux<false> (EdgeFlux.cpp:375-640) with a = -1, b = -0.5           */
/* ===================================================================== */

struct orc_advq {
    const orc_wave *W;
    int nvar, cc, nqv, nqe;
    double *A;
    double xv[32], wv[32], Pv[32 * 16], Dv[32 * 16]; /* (nq, n) column-major */
    double xe[32], we[32], Pe[32 * 16];
};

orc_advq *orc_advection_q_create(const orc_wave *W, int nvar, const double *A, int constant_coefficient, int nqv,
                                 int nqe)
{
    const int n = W->n;
    if (nvar < 1 || nvar > 8) return NULL;
    /* default rules (Div.cpp:319-327, EdgeFlux.cpp:503-511; bilinear elements) */
    if (nqv <= 0) nqv = 1 + ((n - 2) + (n - 1) + 1 + (constant_coefficient ? 0 : n - 1)) / 2;
    if (nqe <= 0) nqe = 1 + (2 * (n - 1) + 2 + (constant_coefficient ? 0 : n - 1)) / 2;
    if (nqv > 32 || nqe > 32) return NULL;
    orc_advq *V = (orc_advq *)calloc(1, sizeof(orc_advq));
    V->W = W;
    V->nvar = nvar;
    V->cc = constant_coefficient ? 1 : 0;
    V->nqv = nqv;
    V->nqe = nqe;
    const size_t na = (size_t)2 * nvar * nvar * (V->cc ? 1 : (size_t)n * n * W->m->nel);
    V->A = (double *)malloc(sizeof(double) * na);
    memcpy(V->A, A, sizeof(double) * na);
    orc_gauss_legendre(nqv, V->xv, V->wv);
    orc_lagrange_basis(V->Pv, n, W->x, nqv, V->xv);
    orc_lagrange_basis_deriv(V->Dv, n, W->x, nqv, V->xv);
    orc_gauss_legendre(nqe, V->xe, V->we);
    orc_lagrange_basis(V->Pe, n, W->x, nqe, V->xe);
    return V;
}

void orc_advection_q_free(orc_advq *V)
{
    if (!V) return;
    free(V->A);
    free(V);
}

void orc_advection_q_rules(const orc_advq *V, int *nqv, int *nqe)
{
    *nqv = V->nqv;
    *nqe = V->nqe;
}

static void advq_div(const orc_advq *V, const double *u, double *divF)
{
    const orc_wave *W = V->W;
    const int n = W->n, nn = n * n, nel = W->m->nel, m = V->nvar, v2 = m * m, nq = V->nqv;
    const double *P = V->Pv, *D = V->Dv;
    double *Pu = (double *)malloc(sizeof(double) * (size_t)n * m * nq);
    double *F = (double *)malloc(sizeof(double) * 2 * (size_t)nq * m * nq);
    double *Df = (double *)malloc(sizeof(double) * (size_t)nq * m * n);
    double *Pg = (double *)malloc(sizeof(double) * (size_t)nq * m * n);
    double Aq[128], J[4];
    for (int el = 0; el < nel; ++el) {
        const double *c = W->m->corners + 8 * (size_t)el;
        const double *ue = u + (size_t)m * nn * el;
        for (int i = 0; i < nq; ++i)
            for (int l = 0; l < n; ++l)
                for (int d = 0; d < m; ++d) {
                    double pu = 0.0;
                    for (int k = 0; k < n; ++k) pu += P[i + nq * k] * ue[d + m * (k + n * l)];
                    Pu[l + n * (d + m * i)] = pu;
                }
        for (int j = 0; j < nq; ++j)
            for (int i = 0; i < nq; ++i) {
                /* the coefficient at the point: A or its interpolation (op_2d) */
                for (int t = 0; t < 2 * v2; ++t) {
                    if (V->cc) {
                        Aq[t] = V->A[t];
                        continue;
                    }
                    double a = 0.0;
                    for (int l = 0; l < n; ++l)
                        for (int k = 0; k < n; ++k)
                            a += P[i + nq * k] * P[j + nq * l] * V->A[t + 2 * v2 * (k + n * (l + (size_t)n * el))];
                    Aq[t] = a;
                }
                quad_jacobian(c, V->xv[i], V->xv[j], J);
                const double w_ij = V->wv[i] * V->wv[j];
                const double Y_eta = J[3] * w_ij, X_eta = J[2] * w_ij, Y_xi = J[1] * w_ij, X_xi = J[0] * w_ij;
                double uq[8];
                for (int d = 0; d < m; ++d) {
                    double s = 0.0;
                    for (int l = 0; l < n; ++l) s += P[j + nq * l] * Pu[l + n * (d + m * i)];
                    uq[d] = s;
                }
                for (int d = 0; d < m; ++d) {
                    double Fd = 0.0, Gd = 0.0;
                    for (int cc = 0; cc < m; ++cc) {
                        const double a0 = Aq[d + m * cc], a1 = Aq[v2 + d + m * cc];
                        Fd += (Y_eta * a0 - X_eta * a1) * uq[cc];
                        Gd += (-Y_xi * a0 + X_xi * a1) * uq[cc];
                    }
                    F[0 + 2 * (i + nq * (d + m * j))] = Fd;
                    F[1 + 2 * (i + nq * (d + m * j))] = Gd;
                }
            }
        for (int j = 0; j < nq; ++j)
            for (int k = 0; k < n; ++k)
                for (int d = 0; d < m; ++d) {
                    double df = 0.0, pg = 0.0;
                    for (int i = 0; i < nq; ++i) {
                        df += D[i + nq * k] * F[0 + 2 * (i + nq * (d + m * j))];
                        pg += P[i + nq * k] * F[1 + 2 * (i + nq * (d + m * j))];
                    }
                    Df[j + nq * (d + m * k)] = df;
                    Pg[j + nq * (d + m * k)] = pg;
                }
        for (int l = 0; l < n; ++l)
            for (int k = 0; k < n; ++k)
                for (int d = 0; d < m; ++d) {
                    double dF = 0.0;
                    for (int j = 0; j < nq; ++j)
                        dF += P[j + nq * l] * Df[j + nq * (d + m * k)] + D[j + nq * l] * Pg[j + nq * (d + m * k)];
                    divF[d + m * (k + n * (l + (size_t)n * el))] += dF;
                }
    }
    free(Pu);
    free(F);
    free(Df);
    free(Pg);
}

/* prolongation of u and of the coefficient, setup_flux_2d_(cc_)q and
 * flux_action_2d_q at the edge rule, the transpose.  Boundary edges: the
 * exterior side has u = 0 and (variable coefficient) A = 0. */
static void advq_faces(const orc_advq *V, int type, const double *u, double *divF, int *err)
{
    const orc_wave *W = V->W;
    const int n = W->n, m = V->nvar, v2 = m * m, nq = V->nqe, ne = orc_mesh_n_edges(W->m, type);
    const int *v2e = type == 0 ? W->v2eI : W->v2eB;
    const int n_sides = type == 0 ? 2 : 1;
    const double *P = V->Pe;
    double *uf = (double *)calloc((size_t)2 * n * m * ne + 1, sizeof(double));       /* (n, m, 2, ne) */
    double *af = (double *)calloc((size_t)2 * n * 2 * v2 * ne + 1, sizeof(double));  /* (n, 2 v2, 2, ne) */
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i) {
                const int j = v2e[i + n * (side + 2 * (size_t)e)];
                for (int d = 0; d < m; ++d) uf[i + n * (d + m * (side + 2 * (size_t)e))] = face_value(W, u, m, j, f, d);
                if (!V->cc)
                    for (int t = 0; t < 2 * v2; ++t)
                        af[i + n * (t + 2 * v2 * (side + 2 * (size_t)e))] = face_value(W, V->A, 2 * v2, j, f, t);
            }
        }
    double nA[64], Fs[64], acc[32 * 8], uq[8];
    for (int e = 0; e < ne; ++e) {
        const orc_edge *E = edge_of(W->m, e, type);
        for (int t = 0; t < nq * m; ++t) acc[t] = 0.0;
        for (int s = 0; s < 2; ++s)
            for (int i = 0; i < nq; ++i) {
                const double w = V->we[i] * E->meas;
                for (int t = 0; t < v2; ++t) {
                    double a0, a1;
                    if (V->cc) {
                        a0 = V->A[t];
                        a1 = V->A[v2 + t];
                    } else {
                        a0 = a1 = 0.0;
                        for (int k = 0; k < n; ++k) {
                            a0 += af[k + n * (t + 2 * v2 * (s + 2 * (size_t)e))] * P[i + nq * k];
                            a1 += af[k + n * (v2 + t + 2 * v2 * (s + 2 * (size_t)e))] * P[i + nq * k];
                        }
                    }
                    nA[t] = E->n[0] * a0 + E->n[1] * a1;
                }
                if (compute_flux_m(m, nA, -1.0, -0.5, s, Fs) != 0) *err = 1;
                for (int d = 0; d < m; ++d) {
                    double t = 0.0;
                    for (int k = 0; k < n; ++k) t += uf[k + n * (d + m * (s + 2 * (size_t)e))] * P[i + nq * k];
                    uq[d] = t;
                }
                for (int d = 0; d < m; ++d) {
                    double flx = 0.0;
                    for (int c2 = 0; c2 < m; ++c2) flx += (w * Fs[c2 + m * d]) * uq[c2];
                    acc[i + nq * d] += flx;
                }
            }
        for (int d = 0; d < m; ++d)
            for (int k = 0; k < n; ++k) {
                double flx = 0.0;
                for (int i = 0; i < nq; ++i) flx += acc[i + nq * d] * P[i + nq * k];
                uf[k + n * (d + m * (0 + 2 * (size_t)e))] = flx;
                uf[k + n * (d + m * (1 + 2 * (size_t)e))] = -flx;
            }
    }
    for (int e = 0; e < ne; ++e)
        for (int side = 0; side < n_sides; ++side) {
            const int f = edge_of(W->m, e, type)->side[side];
            for (int i = 0; i < n; ++i)
                for (int d = 0; d < m; ++d) {
                    const int j = v2e[i + n * (side + 2 * (size_t)e)];
                    face_add(W, divF, m, j, f, d, uf[i + n * (d + m * (side + 2 * (size_t)e))]);
                }
        }
    free(uf);
    free(af);
}

int orc_advection_q_action(const orc_advq *V, const double *u, double *divF)
{
    int err = 0;
    advq_div(V, u, divF);
    advq_faces(V, 0, u, divF, &err);
    return err;
}

int orc_advection_q_bc_action(const orc_advq *V, const double *u, double *divF)
{
    int err = 0;
    advq_faces(V, 1, u, divF, &err);
    return err;
}

