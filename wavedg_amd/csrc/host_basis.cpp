#include "host_basis.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

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

namespace {

// General real m x m matrices, m <= 8, column-major: M(i, j) = M[i + m j].

// Diagonal similarity by powers of two that evens out row and column norms
// (Parlett and Reinsch, "Balancing a matrix for calculation of eigenvalues
// and eigenvectors", Numer. Math. 13, 1969; LAPACK's scaling-only balance):
// B <- D^{-1} B D with D = diag(d).  Exact in floating point.
void balance(int m, double* B, double* d)
{
    for (int i = 0; i < m; ++i) d[i] = 1.0;
    for (bool again = true; again;) {
        again = false;
        for (int i = 0; i < m; ++i) {
            double c = 0.0, r = 0.0;
            for (int j = 0; j < m; ++j)
                if (j != i) {
                    c += std::fabs(B[j + m * i]);
                    r += std::fabs(B[i + m * j]);
                }
            if (c == 0.0 || r == 0.0) continue;
            const double s = c + r;
            double f = 1.0;
            while (c < 0.5 * r) {
                f *= 2.0;
                c *= 4.0;
            }
            while (c >= 2.0 * r) {
                f *= 0.5;
                c *= 0.25;
            }
            if ((c + r) / f < 0.95 * s) {
                again = true;
                d[i] *= f;
                for (int j = 0; j < m; ++j) {
                    B[i + m * j] /= f;
                    B[j + m * i] *= f;
                }
            }
        }
    }
}

// Householder reduction to upper Hessenberg form (similarity; only the
// eigenvalues are used afterwards).
void hessenberg(int m, double* H)
{
    for (int k = 0; k + 2 < m; ++k) {
        double v[8] = {0.0}, nx = 0.0;
        for (int i = k + 1; i < m; ++i) nx += H[i + m * k] * H[i + m * k];
        nx = std::sqrt(nx);
        if (nx == 0.0) continue;
        const double alpha = H[k + 1 + m * k] > 0.0 ? -nx : nx;
        double vv = 0.0;
        for (int i = k + 1; i < m; ++i) {
            v[i] = H[i + m * k] - (i == k + 1 ? alpha : 0.0);
            vv += v[i] * v[i];
        }
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        for (int j = 0; j < m; ++j) {
            double t = 0.0;
            for (int i = k + 1; i < m; ++i) t += v[i] * H[i + m * j];
            t *= beta;
            for (int i = k + 1; i < m; ++i) H[i + m * j] -= t * v[i];
        }
        for (int i = 0; i < m; ++i) {
            double t = 0.0;
            for (int j = k + 1; j < m; ++j) t += H[i + m * j] * v[j];
            t *= beta;
            for (int j = k + 1; j < m; ++j) H[i + m * j] -= t * v[j];
        }
        for (int i = k + 2; i < m; ++i) H[i + m * k] = 0.0;
    }
}

// Eigenvalues of the real 2 x 2 block [[a, b], [c, d]]: the pair mid +- r
// written as d + z and d - bc / z (no cancellation), or a complex pair.
void eig2(double a, double b, double c, double d, double* wr, double* wi)
{
    const double p = 0.5 * (a - d), bc = b * c, disc = p * p + bc;
    if (disc >= 0.0) {
        const double r = std::sqrt(disc), z = p >= 0.0 ? p + r : p - r;
        if (z == 0.0) {
            wr[0] = wr[1] = d;
        } else {
            wr[0] = d + z;
            wr[1] = d - bc / z;
        }
        wi[0] = wi[1] = 0.0;
    } else {
        wr[0] = wr[1] = 0.5 * (a + d);
        wi[0] = std::sqrt(-disc);
        wi[1] = -wi[0];
    }
}

// Reflector I - beta v v' with v = x - alpha e1 mapping x (n = 2 or 3) onto
// alpha e1; beta = 0 when x is already a multiple of e1.
double reflector(int n, const double* x, double* v)
{
    double tail = 0.0;
    for (int i = 1; i < n; ++i) tail += x[i] * x[i];
    if (tail == 0.0) return 0.0;
    const double nx = std::sqrt(x[0] * x[0] + tail), alpha = x[0] > 0.0 ? -nx : nx;
    v[0] = x[0] - alpha;
    for (int i = 1; i < n; ++i) v[i] = x[i];
    return 2.0 / (v[0] * v[0] + tail);
}

// Eigenvalues of an upper Hessenberg matrix by the Francis double-shift QR
// iteration with deflation (Golub and Van Loan, Matrix Computations, 4th ed.,
// Algorithms 7.5.1-7.5.2), transforms restricted to the active window.
// False if the iteration does not converge.
bool francis_eigenvalues(int m, double* H, double* wr, double* wi)
{
    const double eps = std::numeric_limits<double>::epsilon();
    double norm = 0.0;
    for (int i = 0; i < m * m; ++i) norm += H[i] * H[i];
    norm = std::sqrt(norm);
    auto h = [&](int i, int j) -> double& { return H[i + m * j]; };
    int hi = m - 1, its = 0, total = 0;
    while (hi >= 0) {
        // a subdiagonal negligible next to its diagonal neighbours, or next
        // to the whole matrix (a repeated eigenvalue leaves subdiagonals at
        // rounding level that no shift reduces relative to the diagonal)
        int l = hi;
        for (; l > 0; --l) {
            const double s = std::fabs(h(l - 1, l - 1)) + std::fabs(h(l, l));
            if (std::fabs(h(l, l - 1)) <= eps * s || std::fabs(h(l, l - 1)) <= 4.0 * eps * norm) {
                h(l, l - 1) = 0.0;
                break;
            }
        }
        if (l == hi) {
            wr[hi] = h(hi, hi);
            wi[hi] = 0.0;
            --hi;
            its = 0;
            continue;
        }
        if (l == hi - 1) {
            eig2(h(hi - 1, hi - 1), h(hi - 1, hi), h(hi, hi - 1), h(hi, hi), wr + hi - 1, wi + hi - 1);
            hi -= 2;
            its = 0;
            continue;
        }
        if (++total > 100 * m) return false;
        ++its;
        double s, t;
        if (its % 10 == 0) {  // exceptional shifts h(hi, hi) + 0.75 sig +- i 0.66 sig against cycling
            const double sig = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2)), x = h(hi, hi) + 0.75 * sig;
            s = 2.0 * x;
            t = x * x + 0.4375 * sig * sig;
        } else {
            s = h(hi - 1, hi - 1) + h(hi, hi);
            t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
        }
        double x[3] = {h(l, l) * h(l, l) + h(l, l + 1) * h(l + 1, l) - s * h(l, l) + t,
                       h(l + 1, l) * (h(l, l) + h(l + 1, l + 1) - s), h(l + 1, l) * h(l + 2, l + 1)};
        for (int k = l; k + 2 <= hi; ++k) {
            double v[3];
            const double beta = reflector(3, x, v);
            if (beta != 0.0) {
                for (int j = std::max(l, k - 1); j <= hi; ++j) {
                    const double d = beta * (v[0] * h(k, j) + v[1] * h(k + 1, j) + v[2] * h(k + 2, j));
                    for (int q = 0; q < 3; ++q) h(k + q, j) -= d * v[q];
                }
                for (int i = l; i <= std::min(k + 3, hi); ++i) {
                    const double d = beta * (h(i, k) * v[0] + h(i, k + 1) * v[1] + h(i, k + 2) * v[2]);
                    for (int q = 0; q < 3; ++q) h(i, k + q) -= d * v[q];
                }
            }
            x[0] = h(k + 1, k);
            x[1] = h(k + 2, k);
            if (k + 3 <= hi) x[2] = h(k + 3, k);
        }
        double v[2];
        const double beta = reflector(2, x, v);
        if (beta != 0.0) {
            for (int j = std::max(l, hi - 2); j <= hi; ++j) {
                const double d = beta * (v[0] * h(hi - 1, j) + v[1] * h(hi, j));
                h(hi - 1, j) -= d * v[0];
                h(hi, j) -= d * v[1];
            }
            for (int i = l; i <= hi; ++i) {
                const double d = beta * (h(i, hi - 1) * v[0] + h(i, hi) * v[1]);
                h(i, hi - 1) -= d * v[0];
                h(i, hi) -= d * v[1];
            }
        }
    }
    return true;
}

// One-sided Jacobi SVD (Hestenes): M V = U diag(sig), V orthogonal.  M is
// overwritten by U diag(sig).
void jacobi_svd(int m, double* M, double* V, double* sig)
{
    for (int i = 0; i < m * m; ++i) V[i] = (i % (m + 1) == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int p = 0; p + 1 < m; ++p)
            for (int q = p + 1; q < m; ++q) {
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int k = 0; k < m; ++k) {
                    al += M[k + m * p] * M[k + m * p];
                    be += M[k + m * q] * M[k + m * q];
                    ga += M[k + m * p] * M[k + m * q];
                }
                if (std::fabs(ga) <= 1e-15 * std::sqrt(al * be) || ga == 0.0) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int k = 0; k < m; ++k) {
                    const double mp = M[k + m * p], mq = M[k + m * q];
                    M[k + m * p] = c * mp - s * mq;
                    M[k + m * q] = s * mp + c * mq;
                    const double vp = V[k + m * p], vq = V[k + m * q];
                    V[k + m * p] = c * vp - s * vq;
                    V[k + m * q] = s * vp + c * vq;
                }
            }
        if (!rotated) break;
    }
    for (int j = 0; j < m; ++j) {
        double s = 0.0;
        for (int k = 0; k < m; ++k) s += M[k + m * j] * M[k + m * j];
        sig[j] = std::sqrt(s);
    }
}

}  // namespace

int real_abs(int m, const double* C, double* absC)
{
    if (m < 1 || m > 8) return 3;
    bool sym = true;
    for (int j = 0; j < m && sym; ++j)
        for (int i = 0; i < j; ++i)
            if (C[i + m * j] != C[j + m * i]) {
                sym = false;
                break;
            }
    if (sym) return symmetric_abs(m, C, absC) ? 0 : 3;

    double B[64], d[8], H[64], wr[8], wi[8];
    for (int i = 0; i < m * m; ++i) B[i] = C[i];
    balance(m, B, d);
    double nb = 0.0;
    for (int i = 0; i < m * m; ++i) nb += B[i] * B[i];
    nb = std::sqrt(nb);
    for (int i = 0; i < m * m; ++i) H[i] = B[i];
    hessenberg(m, H);
    if (!francis_eigenvalues(m, H, wr, wi)) return 3;
    // Real eigenvalues only (eig.cpp:63-72).  Eigenvalues within tol of each
    // other are one repeated eigenvalue: rounding moves a repeated eigenvalue
    // by O(eps |C|) if it has a full set of eigenvectors, and splits it into a
    // real or complex pair O(sqrt(eps |C|)) apart if it has not, so the
    // tolerance sits between the two; the eigenvector count below then
    // refuses the second kind instead of building |C| from nearly parallel
    // eigenvectors.
    const double tol = 1e-7 * nb;
    for (int i = 0; i < m; ++i)
        if (std::fabs(wi[i]) > tol) return 1;
    std::sort(wr, wr + m);

    // eigenvectors: for each cluster of equal eigenvalues (multiplicity k), the
    // k right singular vectors of B - lambda I with the smallest singular values
    double R[64], lam[8];
    int nv = 0;
    for (int i = 0; i < m;) {
        int j = i + 1;
        while (j < m && wr[j] - wr[j - 1] <= tol) ++j;
        const int k = j - i;
        double mean = 0.0;
        for (int q = i; q < j; ++q) mean += wr[q];
        mean /= k;
        double M[64], V[64], sig[8];
        for (int q = 0; q < m * m; ++q) M[q] = B[q] - (q % (m + 1) == 0 ? mean : 0.0);
        jacobi_svd(m, M, V, sig);
        int idx[8];
        for (int q = 0; q < m; ++q) idx[q] = q;
        std::sort(idx, idx + m, [&](int a, int b) { return sig[a] < sig[b]; });
        if (sig[idx[k - 1]] > tol) return 2;  // fewer than k independent eigenvectors
        for (int q = 0; q < k; ++q) {
            for (int r = 0; r < m; ++r) R[r + m * nv] = V[r + m * idx[q]];
            lam[nv++] = std::fabs(mean);
        }
        i = j;
    }

    // |B| R = R diag|lam|: solve R' X = (R diag|lam|)' by elimination with
    // partial pivoting, |B| = X'
    double A[64], X[64];
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            A[i + m * j] = R[j + m * i];
            X[i + m * j] = R[j + m * i] * lam[i];
        }
    double amax = 0.0;
    for (int i = 0; i < m * m; ++i) amax = std::max(amax, std::fabs(A[i]));
    for (int c = 0; c < m; ++c) {
        int piv = c;
        for (int r = c + 1; r < m; ++r)
            if (std::fabs(A[r + m * c]) > std::fabs(A[piv + m * c])) piv = r;
        if (std::fabs(A[piv + m * c]) <= 1e-12 * amax) return 2;
        if (piv != c)
            for (int j = 0; j < m; ++j) {
                std::swap(A[c + m * j], A[piv + m * j]);
                std::swap(X[c + m * j], X[piv + m * j]);
            }
        for (int r = c + 1; r < m; ++r) {
            const double f = A[r + m * c] / A[c + m * c];
            if (f == 0.0) continue;
            for (int j = c; j < m; ++j) A[r + m * j] -= f * A[c + m * j];
            for (int j = 0; j < m; ++j) X[r + m * j] -= f * X[c + m * j];
        }
    }
    for (int c = m - 1; c >= 0; --c)
        for (int j = 0; j < m; ++j) {
            double v = X[c + m * j];
            for (int q = c + 1; q < m; ++q) v -= A[c + m * q] * X[q + m * j];
            X[c + m * j] = v / A[c + m * c];
        }
    // |C| = D |B| D^{-1}
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) absC[i + m * j] = d[i] * X[j + m * i] / d[j];
    return 0;
}

}  // namespace wdg
