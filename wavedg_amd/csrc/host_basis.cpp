#include "host_basis.hpp"

#include <cmath>

namespace wdg {

namespace {

// P_l(x) by the three-term recurrence.
double legendre(int l, double x)
{
    double a = 1.0, b = x;
    if (l == 0) return a;
    for (int k = 1; k < l; ++k) {
        const double c = ((2 * k + 1) * x * b - k * a) / (k + 1);
        a = b;
        b = c;
    }
    return b;
}

// Table values: QuadratureRule.cpp:37-46.
const double kGll4[] = {-1, -0.447213595499958, 0.447213595499958, 1};
const double kGll5[] = {-1, -0.654653670707977, 0, 0.654653670707977, 1};
const double kGll6[] = {-1, -0.765055323929465, -0.285231516480645, 0.285231516480645, 0.765055323929465, 1};
const double kGll7[] = {-1, -0.830223896278567, -0.468848793470714, 0.0, 0.468848793470714, 0.830223896278567, 1};
const double kGll8[] = {-1, -0.871740148509607, -0.591700181433142, -0.209299217902479,
                        0.2092992179024789, 0.591700181433142, 0.871740148509607, 1};
const double kGll9[] = {-1, -0.899757995411460, -0.677186279510738, -0.363117463826178, 0,
                        0.363117463826178, 0.677186279510738, 0.899757995411460, 1};

// Table values: QuadratureRule.cpp:103-112.
const double kGl[11][10] = {
    {},
    {0.0},
    {-0.577350269189625764509149, 0.577350269189625764509149},
    {-0.774596669241483377035853, 0.0, 0.774596669241483377035853},
    {-0.861136311594052575223946, -0.339981043584856264802666, 0.339981043584856264802666, 0.861136311594052575223946},
    {-0.906179845938663992797627, -0.538469310105683091036314, 0.0, 0.538469310105683091036314, 0.906179845938663992797627},
    {-0.932469514203152027812302, -0.661209386466264513661400, -0.238619186083196908630502, 0.238619186083196908630502,
     0.661209386466264513661400, 0.932469514203152027812302},
    {-0.949107912342758524526190, -0.741531185599394439863865, -0.405845151377397166906606, 0.0,
     0.405845151377397166906606, 0.741531185599394439863865, 0.949107912342758524526190},
    {-0.960289856497536231683561, -0.796666477413626739591554, -0.525532409916328985817739, -0.183434642495649804939476,
     0.183434642495649804939476, 0.525532409916328985817739, 0.796666477413626739591554, 0.960289856497536231683561},
    {-0.968160239507626089835576, -0.836031107326635794299430, -0.613371432700590397308702, -0.324253423403808929038538,
     0.0, 0.324253423403808929038538, 0.613371432700590397308702, 0.836031107326635794299430, 0.968160239507626089835576},
    {-0.973906528517171720077964, -0.865063366688984510732097, -0.679409568299024406234327, -0.433395394129247190799266,
     -0.148874338981631210884826, 0.148874338981631210884826, 0.433395394129247190799266, 0.679409568299024406234327,
     0.865063366688984510732097, 0.973906528517171720077964}};

// Barycentric weights lambda_j = 1 / prod_{k != j} (x_j - x_k).
std::vector<double> bary(int n, const double* x)
{
    std::vector<double> lam(n, 1.0);
    for (int j = 0; j < n; ++j) {
        for (int k = 0; k < n; ++k)
            if (k != j) lam[j] *= (x[j] - x[k]);
        lam[j] = 1.0 / lam[j];
    }
    return lam;
}

}  // namespace

bool gll_rule(int n, std::vector<double>& x, std::vector<double>& w)
{
    const double* tab = nullptr;
    static const double k2[] = {-1, 1}, k3[] = {-1, 0, 1};
    switch (n) {
    case 2: tab = k2; break;
    case 3: tab = k3; break;
    case 4: tab = kGll4; break;
    case 5: tab = kGll5; break;
    case 6: tab = kGll6; break;
    case 7: tab = kGll7; break;
    case 8: tab = kGll8; break;
    case 9: tab = kGll9; break;
    default: return false;
    }
    x.assign(tab, tab + n);
    w.resize(n);
    for (int i = 0; i < n; ++i) {
        const double P = legendre(n - 1, x[i]);
        w[i] = 2.0 / (n * (n - 1) * (P * P));
    }
    return true;
}

bool gauss_legendre_rule(int n, std::vector<double>& x, std::vector<double>& w)
{
    if (n < 1 || n > 20) return false;
    if (n <= 10) {
        x.assign(kGl[n], kGl[n] + n);
    } else {
        // roots of P_n by Newton's method, ascending
        x.resize(n);
        for (int i = 0; i < n; ++i) {
            double t = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; ++it) {
                const double dP = n * (t * legendre(n, t) - legendre(n - 1, t)) / (t * t - 1.0);
                const double dt = legendre(n, t) / dP;
                t -= dt;
                if (std::fabs(dt) < 1e-16) break;
            }
            x[i] = t;
        }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) {
        // P_n'(x) = n (x P_n - P_{n-1}) / (x^2 - 1)
        const double xi = x[i];
        const double dP = n * (xi * legendre(n, xi) - legendre(n - 1, xi)) / (xi * xi - 1.0);
        w[i] = 2.0 / ((1.0 - xi * xi) * dP * dP);
    }
    return true;
}

void lagrange_derivative_matrix(int ngrid, const double* xg, int neval, const double* xe, double* D)
{
    const std::vector<double> lam = bary(ngrid, xg);
    for (int i = 0; i < neval; ++i) {
        const double t = xe[i];
        int at = -1;
        for (int k = 0; k < ngrid; ++k)
            if (t == xg[k]) at = k;
        if (at >= 0) {
            // phi_j'(x_a) = (lam_j / lam_a) / (x_a - x_j), j != a; diagonal by the row-sum identity.
            double diag = 0.0;
            for (int j = 0; j < ngrid; ++j) {
                if (j == at) continue;
                const double v = (lam[j] / lam[at]) / (xg[at] - xg[j]);
                D[i + neval * j] = v;
                diag -= v;
            }
            D[i + neval * at] = diag;
        } else {
            // phi_j(t) = l(t) lam_j / (t - x_j); phi_j' = phi_j (sum_k 1/(t - x_k) - 1/(t - x_j))
            double l = 1.0, s = 0.0;
            for (int k = 0; k < ngrid; ++k) {
                l *= (t - xg[k]);
                s += 1.0 / (t - xg[k]);
            }
            for (int j = 0; j < ngrid; ++j) {
                const double phi = l * lam[j] / (t - xg[j]);
                D[i + neval * j] = phi * (s - 1.0 / (t - xg[j]));
            }
        }
    }
}

void lagrange_interpolation_matrix(int ngrid, const double* xg, int neval, const double* xe, double* B)
{
    const std::vector<double> lam = bary(ngrid, xg);
    for (int i = 0; i < neval; ++i) {
        const double t = xe[i];
        int at = -1;
        for (int k = 0; k < ngrid; ++k)
            if (t == xg[k]) at = k;
        if (at >= 0) {
            for (int j = 0; j < ngrid; ++j) B[i + neval * j] = (j == at) ? 1.0 : 0.0;
            continue;
        }
        // second barycentric form: phi_j(t) = (lam_j/(t-x_j)) / sum_k lam_k/(t-x_k)
        double den = 0.0;
        for (int k = 0; k < ngrid; ++k) den += lam[k] / (t - xg[k]);
        for (int j = 0; j < ngrid; ++j) B[i + neval * j] = (lam[j] / (t - xg[j])) / den;
    }
}

bool mass_fast_diag(int n, const double* xb, int nq, const double* xq, const double* wq, double* A, double* B,
                    double* V, double* lam)
{
    std::vector<double> P((size_t)nq * n);
    lagrange_interpolation_matrix(n, xb, nq, xq, P.data());
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double a = 0.0, b = 0.0;
            for (int q = 0; q < nq; ++q) {
                const double pp = wq[q] * P[q + nq * i] * P[q + nq * j];
                a += pp;
                b += pp * xq[q];
            }
            A[i + n * j] = a;
            B[i + n * j] = b;
        }
    // A = R R' (R lower triangular)
    std::vector<double> R((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        double d = A[j + n * j];
        for (int k = 0; k < j; ++k) d -= R[j + n * k] * R[j + n * k];
        if (!(d > 1e-14 * A[j + n * j])) return false;
        const double rjj = std::sqrt(d);
        R[j + n * j] = rjj;
        for (int i = j + 1; i < n; ++i) {
            double v = A[i + n * j];
            for (int k = 0; k < j; ++k) v -= R[i + n * k] * R[j + n * k];
            R[i + n * j] = v / rjj;
        }
    }
    // C = R^{-1} B R^{-T}: Y = R^{-1} B column by column, then C = R^{-1} Y'
    auto lower_solve = [&](double* x) {  // x <- R^{-1} x
        for (int i = 0; i < n; ++i) {
            double v = x[i];
            for (int k = 0; k < i; ++k) v -= R[i + n * k] * x[k];
            x[i] = v / R[i + n * i];
        }
    };
    std::vector<double> Y(B, B + (size_t)n * n), C((size_t)n * n), col(n);
    for (int j = 0; j < n; ++j) lower_solve(&Y[(size_t)n * j]);
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) col[i] = Y[j + n * i];  // row j of Y
        lower_solve(col.data());
        for (int i = 0; i < n; ++i) C[i + n * j] = col[i];
    }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < j; ++i) C[i + n * j] = C[j + n * i] = 0.5 * (C[i + n * j] + C[j + n * i]);
    // cyclic Jacobi: C = U diag(lam) U'
    std::vector<double> U((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) U[i + n * i] = 1.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                tot += C[i + n * j] * C[i + n * j];
                if (i != j) off += C[i + n * j] * C[i + n * j];
            }
        if (off <= 1e-34 * tot) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = C[p + n * q];
                if (apq == 0.0) continue;
                const double theta = (C[q + n * q] - C[p + n * p]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {  // columns p, q
                    const double ckp = C[k + n * p], ckq = C[k + n * q];
                    C[k + n * p] = c * ckp - s * ckq;
                    C[k + n * q] = s * ckp + c * ckq;
                }
                for (int k = 0; k < n; ++k) {  // rows p, q
                    const double cpk = C[p + n * k], cqk = C[q + n * k];
                    C[p + n * k] = c * cpk - s * cqk;
                    C[q + n * k] = s * cpk + c * cqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double ukp = U[k + n * p], ukq = U[k + n * q];
                    U[k + n * p] = c * ukp - s * ukq;
                    U[k + n * q] = s * ukp + c * ukq;
                }
            }
    }
    // V = R^{-T} U
    for (int j = 0; j < n; ++j) {
        lam[j] = C[j + n * j];
        for (int i = n - 1; i >= 0; --i) {
            double v = U[i + n * j];
            for (int k = i + 1; k < n; ++k) v -= R[k + n * i] * V[k + n * j];
            V[i + n * j] = v / R[i + n * i];
        }
    }
    return true;
}

bool symmetric_abs(int m, const double* C, double* absC)
{
    if (m < 1 || m > 8) return false;
    double A[64], R[64];
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            if (C[i + m * j] != C[j + m * i]) return false;
            A[i + m * j] = C[i + m * j];
            R[i + m * j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int q = 0; q < m; ++q)
            for (int p = 0; p < q; ++p) off += A[p + m * q] * A[p + m * q];
        if (off < 1e-300) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[p + m * q];
                if (std::fabs(apq) < 1e-300) continue;
                const double theta = (A[q + m * q] - A[p + m * p]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
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
    for (int d = 0; d < m; ++d)
        for (int c = 0; c < m; ++c) {
            double v = 0.0;
            for (int k = 0; k < m; ++k) v += R[c + m * k] * std::fabs(A[k + m * k]) * R[d + m * k];
            absC[c + m * d] = v;
        }
    return true;
}

}  // namespace wdg
