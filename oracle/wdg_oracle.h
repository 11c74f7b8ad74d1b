/*
 * wdg_oracle.h -- CPU oracle for the 2D wave DG operator and WaveHoltz.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (arotem3/waveDG, approx_quad = true, Gauss-Lobatto basis) used to
 * check the HIP implementation.  Only tests, the smoke check and the CPU
 * baseline leg of bench.py may load it.  The product library never links it.
 *
 * Layouts follow the reference (column-major, first index fastest):
 *   state u        (3, n, n, n_elem)   p, u, v per GLL node      (docs/docs.md:55-61)
 *   vertices       (2, n_vertices)
 *   elements       (4, n_elem)         counter-clockwise corners (Mesh2D.cpp:5-29)
 *   edges          (5, n_edges)        el0, side0, el1, side1, delta
 *
 * Boundary condition codes: 0 absorbing (characteristic outflow, with the
 * exterior pressure computed after the exterior velocity), 1 reflecting,
 * 2 Dirichlet p = 0 (mirror state p_ext = -p, u_ext = u).
 *
 * Pinning status of each function is listed in oracle/README.md.
 */
#ifndef WDG_ORACLE_H
#define WDG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- basis ---------------------------------------------------------- */
int orc_gauss_lobatto(int n, double *x, double *w);
int orc_gauss_legendre(int n, double *x, double *w);
void orc_lagrange_basis(double *B, int ngrid, const double *xgrid, int neval, const double *xeval);
void orc_lagrange_basis_deriv(double *D, int ngrid, const double *xgrid, int neval, const double *xeval);

/* ---- mesh ----------------------------------------------------------- */
typedef struct orc_mesh orc_mesh;
orc_mesh *orc_mesh_from_vertices(int nv, const double *x, int nel, const int *elems);
void orc_mesh_free(orc_mesh *m);
int orc_mesh_n_elem(const orc_mesh *m);
int orc_mesh_n_edges(const orc_mesh *m, int type); /* type 0 interior, 1 boundary */
void orc_mesh_edges(const orc_mesh *m, int type, int *out /* (5, ne) */);
void orc_mesh_edge_geometry(const orc_mesh *m, int type, double *out /* (5, ne): nx, ny, meas, xmid, ymid */);
double orc_mesh_min_h(const orc_mesh *m);
void orc_mesh_nodes(const orc_mesh *m, int nq, const double *xq, double *out /* (2, nq, nq, nel) */);
void orc_mesh_detj(const orc_mesh *m, int nq, const double *xq, double *out /* (nq, nq, nel) */);

/* ---- wave operator -------------------------------------------------- */
/* Fs(c, d) (3x3 column-major) of EdgeFlux for the wave equation at unit
 * normal (n0, n1) and side 0/1 (EdgeFlux.cpp:31-89), reference (a, b) sign. */
void orc_edge_flux(double n0, double n1, double a, double b, int side, double *Fs);
typedef struct orc_wave orc_wave;
/* a, b follow the reference EdgeFlux sign convention (upwind: a=-1, b=-0.5). */
orc_wave *orc_wave_create(const orc_mesh *m, int n, const int *bc, double a, double b);
/* the same on a basis of n Gauss-Lobatto (leg = 0) or Gauss-Legendre (leg = 1) points */
orc_wave *orc_wave_create_basis(const orc_mesh *m, int n, int leg, const int *bc, double a, double b);
/* approx_quad = false (Div<false>, EdgeFlux<false>, MassMatrix<false>) with
 * Gauss-Legendre rules of nqv (volume), nqe (edges) and nqm (mass matrix)
 * points; 0 selects the reference's default for bilinear elements (n, n+1,
 * n+1).  Every other function then uses the exact-quadrature operators. */
orc_wave *orc_wave_create_exact(const orc_mesh *m, int n, const int *bc, double a, double b, int nqv, int nqe,
                                int nqm);
orc_wave *orc_wave_create_exact_basis(const orc_mesh *m, int n, int leg, const int *bc, double a, double b, int nqv,
                                      int nqe, int nqm);
void orc_wave_free(orc_wave *W);
long long orc_wave_ndof(const orc_wave *W);
void orc_wave_action(const orc_wave *W, const double *u, double *du);   /* du += A u   */
void orc_wavebc_action(const orc_wave *W, const double *u, double *du); /* du += B u   */
/* Div<true>::action alone (volume term of WaveEquation::action). */
void orc_wave_div(const orc_wave *W, const double *u, double *du);
/* DivF2D<true>::action with the linear flux f(x, u) = (A u, B u), nvar <= 16. */
void orc_divf_linear(const orc_wave *W, int nvar, const double *A, const double *B, const double *u, double *divF);
void orc_mass_action(const orc_wave *W, int nvar, const double *x, double *y);
void orc_mass_inv(const orc_wave *W, int nvar, double *x);
/* vt = M^{-1} (A u + B u + F s); F may be NULL. */
void orc_rhs(const orc_wave *W, const double *u, double *vt, const double *F, double s);
/* n_steps of the low-storage RK4 on u' = M^{-1}(A+B)u (+ forcing F*sin-term if
 * omega > 0 and F != NULL, as WaveHoltz::pi0). t is advanced. */
void orc_rk4_steps(const orc_wave *W, double dt, double *t, double *u, int n_steps,
                   const double *F, double omega);
void orc_set_threads(int n);

/* ---- WaveHoltz ------------------------------------------------------ */
typedef struct orc_wh orc_wh;
orc_wh *orc_waveholtz_create(const orc_wave *W, double omega);
void orc_waveholtz_free(orc_wh *H);
void orc_waveholtz_info(const orc_wh *H, double *T, double *dt, int *nt);
double orc_waveholtz_K(const orc_wh *H, double t);
void orc_waveholtz_pi0(const orc_wh *H, double *u, const double *F);
void orc_waveholtz_S(const orc_wh *H, double *u);
void orc_waveholtz_postprocess(const orc_wh *H, double *Hout, const double *u);

/* ---- linear functional (load vector) ----------------------------------
 * (f, phi_ij) with Gauss-Legendre(nq) quadrature (nq <= 24, nvar <= 8) for
 * values of f given at the mapped points: fq (nvar, nq, nq, nel), F
 * (nvar, n, n, nel). */
void orc_linear_functional(const orc_mesh *m, int n, int nq, int nvar, const double *fq, double *F);
/* The same for f = sum of Gaussians in
 * variable `var`: f(x) = sum_k amp_k exp(-|x - c_k|^2 / (2 sigma_k^2)).
 * src is (4, nsrc): cx, cy, sigma, amp. */
void orc_load_vector_gauss(const orc_mesh *m, int n, int nq, int nsrc, const double *src,
                           int var, double *F);

/* ---- Advection<true>, AdvectionHomogeneousBC<true> (Advection.hpp:15-273) -
 * On the mesh and basis of W, nvar <= 8 variables, coefficient A of shape
 * (nvar, nvar, 2) (constant_coefficient) or (nvar, nvar, 2, n, n, nel), flux
 * constants a = -1, b = -0.5.  n.A must have real eigenvalues and a full
 * set of eigenvectors at every face node (real_eig, eig.cpp:29-74); create
 * returns NULL otherwise.  action: divF += Div(u) + interior fluxes; bc_action:
 * divF += boundary fluxes with u = 0 outside. */
typedef struct orc_adv orc_adv;
/* _ComputeFlux::flux (EdgeFlux.cpp:31-89) for one m x m normal matrix n.A
 * (column-major, m <= 8): Fs(c, d) = (a/2 nA + sgn b |nA|)(d, c), sgn = +1
 * on side 0 and -1 on side 1.  Returns 0; 1 if n.A has complex eigenvalues;
 * 2 if its eigenvectors are linearly dependent; 3 if m is out of range or
 * the eigenvalue iteration fails. */
int orc_flux_matrix(int m, const double *nA, double a, double b, int side, double *Fs);
orc_adv *orc_advection_create(const orc_wave *W, int nvar, const double *A, int constant_coefficient);
void orc_advection_free(orc_adv *V);
void orc_advection_action(const orc_adv *V, const double *u, double *divF);
void orc_advection_bc_action(const orc_adv *V, const double *u, double *divF);

/* ---- compressible Euler flux of examples/cl2d.cpp ----------------------
 * q = (rho, rho u, rho v, E) per node, arrays (4, n, n, nel).
 * orc_divf_euler: DivF2D<true>::action with F, divF += (F(q), grad v).
 * orc_edgefluxf_euler: divF += the negated, lifted EdgeFluxF2D<true> flux with
 * the local Lax-Friedrichs numerical flux on interior (type 0) or boundary
 * (type 1, exterior = interior copy) edges, as cl2d.cpp's time derivative. */
void orc_divf_euler(const orc_wave *W, double gamma, const double *u, double *divF);
void orc_edgefluxf_euler(const orc_wave *W, double gamma, int type, const double *u, double *divF);

/* DivF2D<true> / EdgeFluxF2D<true> with caller functors in the reference's
 * signatures (plus a context pointer), nvar <= 8.  orc_divf_functor:
 * divF += (F(x, u), grad v), F (nvar, 2) column-major.
 * orc_edgefluxf_functor: divF += scale * lifted (+-fh w ds) on interior
 * (type 0) or boundary (type 1) edges; ext (may be NULL: copy) gives the
 * exterior state of boundary edges. */
typedef void (*orc_flux_fn)(const double *x, const double *u, double *F, void *ctx);
typedef void (*orc_nflux_fn)(const double *x, const double *n, const double *uL, const double *uR, double *fh,
                             void *ctx);
typedef void (*orc_ext_fn)(const double *x, const double *n, const double *u, double *u_ext, void *ctx);
void orc_divf_functor(const orc_wave *W, int nvar, orc_flux_fn f, void *ctx, const double *u, double *divF);
void orc_edgefluxf_functor(const orc_wave *W, int nvar, orc_nflux_fn fh, orc_ext_fn ext, void *ctx, int type,
                           double scale, const double *u, double *divF);
/* Advection<false> / AdvectionHomogeneousBC<false>: coefficient as for
 * orc_advection_create, Gauss-Legendre rules of nqv x nqv volume and nqe edge
 * points (0 = the reference's defaults for bilinear elements).  The actions
 * return nonzero if a normal matrix has no real eigen-decomposition. */
typedef struct orc_advq orc_advq;
orc_advq *orc_advection_q_create(const orc_wave *W, int nvar, const double *A, int constant_coefficient, int nqv,
                                 int nqe);
void orc_advection_q_free(orc_advq *V);
void orc_advection_q_rules(const orc_advq *V, int *nqv, int *nqe);
int orc_advection_q_action(const orc_advq *V, const double *u, double *divF);
int orc_advection_q_bc_action(const orc_advq *V, const double *u, double *divF);
/* the approx_quad = false forms with Gauss-Legendre rules of nq points */
void orc_divf_functor_q(const orc_wave *W, int nvar, int nq, orc_flux_fn f, void *ctx, const double *u, double *divF);
void orc_edgefluxf_functor_q(const orc_wave *W, int nvar, int nq, orc_nflux_fn fh, orc_ext_fn ext, void *ctx,
                             int type, double scale, const double *u, double *divF);

/* ---- reductions ------------------------------------------------------ */
double orc_norm(long long n, const double *x);
double orc_error(long long n, const double *x, const double *y);

#ifdef __cplusplus
}
#endif
#endif
