// Host-side 1D basis data for the tensor-product Gauss-Lobatto (GLL) nodal
// basis: nodes, weights, the Lagrange differentiation matrix, and the
// Gauss-Legendre rule used by load vectors.
#pragma once

#include <vector>

namespace wdg {

// GLL nodes/weights for n points (2 <= n <= 9).  Node values follow the table
// of QuadratureRule.cpp:37-46 (15 significant digits, including its slightly
// asymmetric n = 8 row) so that operators built on them agree with the
// reference to rounding; weights are 2 / (n (n-1) P_{n-1}(x_i)^2)
// (QuadratureRule.cpp:94).
bool gll_rule(int n, std::vector<double>& x, std::vector<double>& w);

// Gauss-Legendre nodes/weights for 1 <= n <= 10 (QuadratureRule.cpp:103-112).
bool gauss_legendre_rule(int n, std::vector<double>& x, std::vector<double>& w);

// D(i, j) = phi_j'(xe_i) (column-major, neval x ngrid), phi_j the Lagrange
// polynomials on xg (lagrange_interpolation.cpp:33-50).
void lagrange_derivative_matrix(int ngrid, const double* xg, int neval, const double* xe, double* D);

// B(i, j) = phi_j(xe_i) (column-major, neval x ngrid) (lagrange_interpolation.cpp:14-31).
void lagrange_interpolation_matrix(int ngrid, const double* xg, int neval, const double* xe, double* B);

}  // namespace wdg
