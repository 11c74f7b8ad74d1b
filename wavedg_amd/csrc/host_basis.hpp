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

// The exact mass matrix of a bilinear element factorises because det J is
// affine in (xi, eta): det J = d0 + d1 xi + d2 eta, so with the 1D matrices
//   A(i, j) = sum_q w_q phi_i(x_q) phi_j(x_q),  B(i, j) = sum_q w_q x_q phi_i(x_q) phi_j(x_q)
// (rule x_q, w_q; basis phi on xb) the element matrix on node (i + n j) is
//   M = d0 A(x)A + d1 B(x)A + d2 A(x)B   (first factor on the xi index).
// V and lam solve B V = A V diag(lam) with V' A V = I, hence
//   M^{-1} = (V(x)V) diag(1 / (d0 + d1 lam_a + d2 lam_b)) (V(x)V)'.
// All n x n column-major.  Returns false when A is not positive definite
// (a rule too small for the basis).
bool mass_fast_diag(int n, const double* xb, int nq, const double* xq, const double* wq, double* A, double* B,
                    double* V, double* lam);

// |C| = R |diag(e)| R' for a symmetric m x m matrix C = R diag(e) R'
// (column-major, m <= 8) by cyclic Jacobi; false if C is not symmetric.
bool symmetric_abs(int m, const double* C, double* absC);

// |C| = R diag(|e|) R^{-1} for a real m x m matrix C = R diag(e) R^{-1}
// (column-major, m <= 8) with real eigenvalues and a full set of
// eigenvectors: the factorisation _ComputeFlux::flux builds from real_eig
// (EdgeFlux.cpp:31-89, eig.cpp:29-74).  Symmetric C takes the cyclic Jacobi
// route of symmetric_abs; otherwise C is balanced, its eigenvalues come from
// the Francis QR iteration, the eigenvectors of each group of equal
// eigenvalues from the null space of C - lambda I (Jacobi SVD), and |C| from
// one solve with the eigenvector matrix.  Returns 0; 1 if C has complex
// eigenvalues; 2 if its eigenvectors are linearly dependent; 3 if m is out
// of range or the iteration does not converge.
int real_abs(int m, const double* C, double* absC);

}  // namespace wdg
