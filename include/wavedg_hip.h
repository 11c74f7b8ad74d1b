/*
 * wavedg_hip.h -- C ABI of the MI355X (gfx950) implementation of the 2D wave
 * DG operator and the WaveHoltz iteration of arotem3/waveDG.
 *
 * Plain pointers and sizes only.  Every function returns a status code
 * (WDG_OK = 0); on failure wdg_hip_last_error() holds a message for the
 * calling thread.  The C++ drop-in classes (wavedg_amd/include/wavedg_hip/dg.hpp)
 * turn nonzero codes into std::runtime_error, as wdg_error does in the
 * reference (source/wdg_error.cpp:5-20).
 *
 * Memory: arguments named d_* are device pointers (hipMalloc'd, on the
 * handle's device); everything else is host memory.  `stream` is a
 * hipStream_t (NULL = default stream); device work is enqueued on it and the
 * call returns without synchronising, except functions that return a host
 * value (norms, residuals), which synchronise the stream.
 *
 * Layouts (column-major, first index fastest, as in the reference,
 * docs/docs.md:55-61):
 *   state        (3, n, n, n_elem)  (p, u, v) per basis node (Gauss-Lobatto unless
 *                                     the handle was made with a Gauss-Legendre basis)
 *   vertices     (2, n_vertices);   elements (4, n_elem), corners counter-clockwise
 *   H            (2, n, n, n_elem)  real/imag part of the Helmholtz solution
 *
 * Precision: WDG_F64 runs everything in double.  WDG_F32 keeps the state, the
 * right-hand side and the Runge-Kutta stages in float; WaveHoltz keeps its
 * filter accumulation, fixed-point update and reductions in double (the
 * WaveHoltz functions always take double arrays).
 *
 * Thread safety: a handle is used by one thread at a time (the reference's
 * operators hold mutable scratch and are not re-entrant either,
 * WaveEquation.hpp:61).  Calls on one handle must be serialised on one stream.
 */
#ifndef WAVEDG_HIP_H
#define WAVEDG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    WDG_OK = 0,
    WDG_ERR_INVALID = 1,   /* bad argument or handle */
    WDG_ERR_HIP = 2,       /* a HIP runtime call failed */
    WDG_ERR_ALLOC = 3,     /* device allocation failed */
    WDG_ERR_MESH = 4,      /* malformed or inverted mesh, missing file */
    WDG_ERR_NONFINITE = 5  /* a residual or norm is NaN/Inf (e.g. dt above the CFL limit) */
};

enum { WDG_F64 = 0, WDG_F32 = 1 };
/* basis point families, in the order of QuadratureRule::QuadratureType
 * (QuadratureRule.hpp:26-30) */
enum { WDG_GAUSS_LEGENDRE = 0, WDG_GAUSS_LOBATTO = 1 };
enum { WDG_INTERIOR = 0, WDG_BOUNDARY = 1 };  /* FaceType, Mesh2D.hpp */
enum { WDG_BC_ABSORB = 0, WDG_BC_REFLECT = 1, WDG_BC_DIRICHLET = 2 };

typedef struct wdg_mesh_s *wdg_mesh;
typedef struct wdg_wave2d_s *wdg_wave2d;
typedef struct wdg_waveholtz_s *wdg_waveholtz;
typedef struct wdg_advection_s *wdg_advection;
typedef struct wdg_flux_s *wdg_flux;

const char *wdg_hip_last_error(void);
int wdg_hip_version(int *major, int *minor);

/* ---- device memory helpers (plumbing for FFI callers without a HIP runtime) */
int wdg_hip_malloc(void **d_ptr, size_t bytes);
int wdg_hip_free(void *d_ptr);
int wdg_hip_memcpy_htod(void *d_dst, const void *src, size_t bytes);
int wdg_hip_memcpy_dtoh(void *dst, const void *d_src, size_t bytes);
int wdg_hip_memset(void *d_ptr, int value, size_t bytes);
int wdg_hip_synchronize(void *stream);

/* ---- mesh (host topology) -------------------------------------------
 * Replaces Mesh2D::from_vertices / uniform_rect / from_file
 * (wavedg/Mesh2D.hpp:339-355; source/Mesh2D.cpp:5-163).  Edge numbering,
 * interior/boundary lists and delta follow Mesh2D.cpp:5-92; the edge key is
 * 64-bit. */
int wdg_hip_mesh_from_vertices(int n_vertices, const double *x, int n_elem, const int *elements, wdg_mesh *out);
int wdg_hip_mesh_uniform_rect(int nx, double ax, double bx, int ny, double ay, double by, wdg_mesh *out);
int wdg_hip_mesh_from_file(const char *dir, wdg_mesh *out);
int wdg_hip_mesh_destroy(wdg_mesh mesh);
/* Mesh2D::n_elem, n_edges(type) (Mesh2D.hpp:230-252) */
int wdg_hip_mesh_counts(wdg_mesh mesh, int *n_elem, int *n_interior_edges, int *n_boundary_edges);
/* Mesh2D::edge(i, type) (Mesh2D.hpp:292-311): out is (5, n_edges) = el0, side0, el1, side1, delta */
int wdg_hip_mesh_edges(wdg_mesh mesh, int face_type, int *out);
/* edge_metrics(..).normals()/measures() at the edge midpoint and the midpoint
 * itself (Mesh2D.cpp:702-749): out is (5, n_edges) = nx, ny, ds, xmid, ymid */
int wdg_hip_mesh_edge_geometry(wdg_mesh mesh, int face_type, double *out);
/* Mesh2D::min_h (Mesh2D.cpp:751-767) */
int wdg_hip_mesh_min_h(wdg_mesh mesh, double *h);
/* element_metrics(basis).physical_coordinates() (Mesh2D.cpp:687-697) on the
 * n_colloc Gauss-Lobatto nodes: out is (2, n, n, n_elem) */
int wdg_hip_mesh_nodes(wdg_mesh mesh, int n_colloc, double *out);

/* ---- basis -----------------------------------------------------------
 * QuadratureRule::quadrature_rule(n, GaussLobatto) (QuadratureRule.cpp:32-97,164-184)
 * and lagrange_basis_deriv (lagrange_interpolation.cpp:33-50): x, w (n), D (n, n). */
int wdg_hip_gll(int n, double *x, double *w, double *D);
/* QuadratureRule::quadrature_rule(n, GaussLegendre) (QuadratureRule.cpp:99-148,164-184):
 * the rule the approx_quad = false operators integrate with; x, w (n), n in 1..20
 * (tabulated up to 10 as the reference; above, Newton-refined roots of P_n). */
int wdg_hip_gauss_legendre(int n, double *x, double *w);
/* _ComputeFlux::flux (EdgeFlux.cpp:31-89), host only: the m x m flux matrix
 * F = (a/2 nA + sgn b |nA|)^T of one normal matrix nA (column-major, m <= 8),
 * sgn = +1 for side 0 and -1 for side 1, |nA| = R diag(|e|) R^{-1} from the
 * real eigen-decomposition nA = R diag(e) R^{-1} (eig.cpp:29-74).  Fails
 * when nA has complex eigenvalues or linearly dependent eigenvectors (a
 * repeated eigenvalue without a full set of eigenvectors).  The advection
 * operators build their face tables with the same routine. */
int wdg_hip_flux_matrix(int m, const double *nA, double a, double b, int side, double *F);

/* ---- wave operator ---------------------------------------------------
 * One handle holds WaveEquation (WaveEquation.cpp:5-30), WaveBC
 * (WaveEquation.cpp:70-105) and MassMatrix<true> (MassMatrix.cpp:6-28) on the
 * current HIP device, for approx_quad = true and a Gauss-Lobatto basis of
 * n_colloc points (2..9).  bc has one code per boundary edge (NULL = all
 * reflecting).  (a, b) use the README/BASELINE convention: upwind a = 1,
 * b = 1/2; central a = 1, b = 0 (the reference's EdgeFlux uses the negated
 * pair, EdgeFlux.hpp:51).  The mesh may be destroyed after this call. */
int wdg_hip_wave2d_create(wdg_mesh mesh, int n_colloc, const int *bc, double a, double b, int precision,
                          wdg_wave2d *out);
/* The same handle with approx_quad = false: Div<false> (Div.cpp:312-360,
 * 469-588), EdgeFlux<false> (EdgeFlux.cpp:375-620) and MassMatrix<false>
 * (MassMatrix.cpp:85-227, the exact block mass matrix), FP64, with
 * Gauss-Legendre rules of nq_volume, nq_edge and nq_mass points (<= 10;
 * 0 = the reference's default for bilinear elements: n, n + 1, n + 1).  Every
 * operator, RK and WaveHoltz entry point then uses these operators (DivF and
 * the schedule autotuner are for approx_quad = true handles).  The mass matrix
 * is applied and inverted from the element's affine det J (no per-element
 * factors are stored); a mass rule too small for the basis is refused. */
int wdg_hip_wave2d_create_exact(wdg_mesh mesh, int n_colloc, const int *bc, double a, double b, int nq_volume,
                                int nq_edge, int nq_mass, wdg_wave2d *out);
/* The general constructor, WaveEquation(mesh, basis, approx_quad, quad)
 * (WaveEquation.hpp:28; WaveEquation.cpp:5-30): a basis of n_colloc
 * Gauss-Lobatto (WDG_GAUSS_LOBATTO) or Gauss-Legendre (WDG_GAUSS_LEGENDRE)
 * points and either quadrature mode.  A Gauss-Legendre basis takes side
 * traces by interpolation along the element and lifts face terms back with
 * the same weights, LegendreFaceProlongator (FaceProlongator.cpp:717-769,
 * 809-835, 908-934).  approx_quad = 0 needs precision WDG_F64; nq_* as in
 * wdg_hip_wave2d_create_exact (ignored for approx_quad = 1).  Every operator
 * built on a handle (DivF2D, EdgeFluxF2D, flux functors, Advection, the load
 * vector) follows its basis, and so do the RK-stage schedules and their
 * autotuner; batched WaveHoltz takes Gauss-Lobatto bases. */
int wdg_hip_wave2d_create_basis(wdg_mesh mesh, int n_colloc, int basis_type, const int *bc, double a, double b,
                                int precision, int approx_quad, int nq_volume, int nq_edge, int nq_mass,
                                wdg_wave2d *out);
/* the handle's basis: type, nodes x[n_colloc] and weights w[n_colloc] (any may be NULL) */
int wdg_hip_wave2d_basis(wdg_wave2d op, int *basis_type, double *x, double *w);
int wdg_hip_wave2d_quadrature(wdg_wave2d op, int *exact, int *nq_volume, int *nq_edge, int *nq_mass);
int wdg_hip_wave2d_destroy(wdg_wave2d op);
/* ndof = 3 n^2 n_elem */
int wdg_hip_wave2d_info(wdg_wave2d op, long long *ndof, int *n_colloc, int *n_elem, int *precision);
/* WaveEquation::action (WaveEquation.cpp:54-68, Operator.hpp:19): d_y += A d_x.
 * Volume divergence (Div<true>, Div.cpp:131-179), interior traces
 * (lobatto_action_2d, FaceProlongator.cpp:565-590), upwind/central flux
 * (flux_action_2d_c, EdgeFlux.cpp:289-325) and lift (lobatto_transpose_2d,
 * FaceProlongator.cpp:650-672) in one kernel.  d_x and d_y must not overlap. */
int wdg_hip_wave2d_action(wdg_wave2d op, const void *d_x, void *d_y, void *stream);
/* DivF2D<true>::action (wavedg/DivF.hpp:207-274) for the linear flux functor
 * f(x, u) = (A u, B u): d_divF += (F, grad v) over the handle's mesh and
 * basis, for n_var <= 8 variables in the layout (n_var, n, n, n_elem) and
 * the handle's precision.  A and B are host (n_var, n_var) column-major
 * matrices (x- and y-flux).  With the wave equation's matrices it is the
 * volume term of WaveEquation::action. */
int wdg_hip_divf2d_linear(wdg_wave2d op, int n_var, const double *A, const double *B, const void *d_u, void *d_divF,
                          void *stream);
/* LinearFunctional2D::action(n_var, f, F) (wavedg/LinearFunctional.hpp:
 * 136-202) for a function the caller has evaluated at the quadrature points:
 * d_fq (n_var, nq, nq, n_elem) holds f at the nq-point Gauss-Legendre rule
 * mapped into every element (wdg_hip_mesh_gl_points; nq = 0 means n_colloc,
 * the reference's default); d_F (n_var, n, n, n_elem) <- (f, phi_ij).
 * The first call with a given nq builds the rule's table on the handle with
 * a blocking upload (not capturable in a graph); later calls with that nq
 * only enqueue the kernel on `stream`. */
int wdg_hip_linear_functional(wdg_wave2d op, int n_var, int nq, const double *d_fq, double *d_F, void *stream);
/* element_metrics(quad).physical_coordinates() (Mesh2D.cpp:687-697) for the
 * nq-point Gauss-Legendre rule: out is (2, nq, nq, n_elem), host memory. */
int wdg_hip_mesh_gl_points(wdg_mesh mesh, int nq, double *out);
/* DivF2D<true>::action (wavedg/DivF.hpp:207-274) with the compressible
 * Euler flux of examples/cl2d.cpp:57-81 for q = (rho, rho u, rho v, E) per
 * node, (4, n, n, n_elem), ratio of specific heats gamma:
 * d_divF += (F(q), grad v).  FP64 approx_quad = true handles. */
int wdg_hip_divf2d_euler(wdg_wave2d op, double gamma, const void *d_q, void *d_divF, void *stream);
/* The face part of examples/cl2d.cpp's time derivative (cl2d.cpp:180-215):
 * EdgeFluxF2D<true>::action (wavedg/EdgeFluxF.hpp:136-171) with the local
 * Lax-Friedrichs flux (cl2d.cpp:98-111) on the edges of face_type
 * (WDG_INTERIOR, or WDG_BOUNDARY with the exterior state a copy of the
 * interior one), negated and lifted: d_divF -= sum over those edges of
 * (fh, v).  FP64 approx_quad = true handles. */
int wdg_hip_edgefluxf2d_euler(wdg_wave2d op, double gamma, int face_type, const void *d_q, void *d_divF,
                              void *stream);
/* Advection<true> and AdvectionHomogeneousBC<true> (wavedg/Advection.hpp:
 * 15-273) on the mesh and basis of an FP64 approx_quad = true wave handle
 * (which must outlive it): n_var <= 8 variables in the layout
 * (n_var, n, n, n_elem), coefficient A a host array of shape (n_var, n_var, 2)
 * (constant_coefficient != 0) or (n_var, n_var, 2, n, n, n_elem), column-major,
 * A(k, l, d) the (k, l) entry of the d-direction flux matrix; upwind flux
 * (a = -1, b = -0.5 of EdgeFlux).  n.A must have real eigenvalues and a
 * full set of eigenvectors at every face node, as wdg_hip_flux_matrix
 * requires; otherwise create fails. */
int wdg_hip_advection_create(wdg_wave2d op, int n_var, const double *A, int constant_coefficient, wdg_advection *out);
/* Advection::action (Advection.hpp:127-140): d_divF += Div(u) + interior edge fluxes. */
int wdg_hip_advection_action(wdg_advection adv, const void *d_u, void *d_divF, void *stream);
/* AdvectionHomogeneousBC::action (Advection.hpp:252-270): d_divF += boundary
 * edge fluxes with u = 0 outside the domain. */
int wdg_hip_advection_bc_action(wdg_advection adv, const void *d_u, void *d_divF, void *stream);
/* Advection<false> / AdvectionHomogeneousBC<false> (Advection.hpp:59-93,
 * 190-215): Div<false> with the coefficient (Div.cpp:312-360, 200-310) and
 * EdgeFlux<false> (EdgeFlux.cpp:375-640) on Gauss-Legendre rules of
 * nq_volume^2 and nq_edge points (0 = the reference's defaults for bilinear
 * elements; at most 20); the coefficient is interpolated to the points as
 * the reference does.  Any FP64 wave handle of the mesh and basis; the
 * action entry points above then run the quadrature kernels. */
int wdg_hip_advection_create_exact(wdg_wave2d op, int n_var, const double *A, int constant_coefficient,
                                   int nq_volume, int nq_edge, wdg_advection *out);
int wdg_hip_advection_destroy(wdg_advection adv);

/* ---- DivF2D / EdgeFluxF2D with flux functors --------------------------
 * The reference's DivF2D<true>::action(n_var, f, u, divF) and
 * EdgeFluxF2D<true>::action(n_var, flux, uB, fB) (DivF.hpp:207-274,
 * EdgeFluxF.hpp:136-171) take C++ functors.  Here the functors are HIP
 * source text in the reference's signatures, compiled for the GPU at run
 * time (hipRTC) with the handle's order, n_var and basis, on first use:
 *   __device__ void flux(const double *x, const double *u, double *F);
 *       F (n_var, 2) column-major: F[d] = f_d, F[n_var + d] = g_d
 *   __device__ void numerical_flux(const double *x, const double *n,
 *                                  const double *uL, const double *uR, double *fh);
 *   and optionally, after #define WDG_EXTERIOR_STATE,
 *   __device__ void exterior_state(const double *x, const double *n,
 *                                  const double *u, double *u_ext);
 * A source may define only the functors it uses.  n_var <= 8, FP64
 * approx_quad = true handles of either basis; the wave handle must outlive
 * the flux handle. */
int wdg_hip_flux_create(wdg_wave2d op, int n_var, const char *source, wdg_flux *out);
/* The same functors with approx_quad = false: DivF2D<false> (DivF.hpp:
 * 275-392) and EdgeFluxF2D<false> (EdgeFluxF.hpp:182-250) interpolate u to
 * Gauss-Legendre rules of nq_volume x nq_volume and nq_edge points (0 = the
 * reference's default of 2 n_colloc; at most 20) and evaluate the functors
 * there.  Any FP64 wave handle of the mesh and basis serves. */
int wdg_hip_flux_create_exact(wdg_wave2d op, int n_var, const char *source, int nq_volume, int nq_edge,
                              wdg_flux *out);
/* compile now rather than on first use: which 0 = flux, 1 = numerical_flux;
 * a source that does not compile returns WDG_ERR_INVALID with the compiler log.
 * Call it before capturing wdg_hip_divf2d / wdg_hip_edgefluxf2d into a HIP
 * graph: the first use loads a code module, which a capture cannot record. */
int wdg_hip_flux_compile(wdg_flux fx, int which);
/* compile a source for (n_colloc, n_var, basis_type) without a device or a
 * handle, to check it: WDG_ERR_INVALID with the compiler log if it fails */
int wdg_hip_flux_check(const char *source, int n_colloc, int n_var, int basis_type, int which);
/* d_divF += (F(x, u), grad v) at every node, u and divF (n_var, n, n, n_elem) */
int wdg_hip_divf2d(wdg_flux fx, const void *d_u, void *d_divF, void *stream);
/* Face prolongation, EdgeFluxF2D<true>::action and the transpose, for
 * interior or boundary (face_type) edges: with el0's outward normal n and
 * uL = el0's trace, fh = numerical_flux(x, n, uL, uR); el0's nodes receive
 * scale * fh w ds and el1's -scale * fh w ds (cl2d.cpp:186-216 uses
 * scale = -1).  Boundary edges take uR = exterior_state(x, n, uL), or uL
 * itself when the source does not define WDG_EXTERIOR_STATE. */
int wdg_hip_edgefluxf2d(wdg_flux fx, int face_type, double scale, const void *d_u, void *d_divF, void *stream);
int wdg_hip_flux_destroy(wdg_flux fx);
/* WaveBC::action (WaveEquation.cpp:190-215): d_y += B d_x (boundary flux and lift). */
int wdg_hip_wavebc_action(wdg_wave2d op, const void *d_x, void *d_y, void *stream);
/* MassMatrix<true>::inv (MassMatrix.cpp:69-82): d_x <- M^{-1} d_x, in place. */
int wdg_hip_mass_inv(wdg_wave2d op, void *d_x, void *stream);
/* MassMatrix<true>::action (MassMatrix.cpp:52-67): d_y <- M d_x. */
int wdg_hip_mass_action(wdg_wave2d op, const void *d_x, void *d_y, void *stream);
/* MassMatrix<true>::action / inv with n_var (MassMatrix.cpp:52-82): d_y <- M d_x
 * (inverse = 0) or M^{-1} d_x (inverse != 0) for arrays of n_var variables
 * per node, (n_var, n, n, n_elem); d_x == d_y is allowed.  Exact-quadrature
 * handles take n_var = 3 only. */
int wdg_hip_mass_apply(wdg_wave2d op, int n_var, const void *d_x, void *d_y, int inverse, void *stream);
/* The time derivative of the examples and WaveHoltz (WaveHoltz.cpp:57-70):
 * d_y <- M^{-1} (A d_x + B d_x + s d_F), fused.  d_F may be NULL; it has the
 * handle's precision. */
int wdg_hip_wave2d_rhs(wdg_wave2d op, const void *d_x, void *d_y, const void *d_F, double s, void *stream);
/* n_steps of ode::RungeKutta4::step (ode.hpp:66-93) on u' = M^{-1}(A+B)u
 * [- F sin(omega t)/omega if d_F != NULL and omega > 0, as in WaveHoltz::pi0],
 * each stage one fused kernel.  *t is advanced by n_steps dt.  d_u is updated
 * in place; d_F has the handle's precision. */
int wdg_hip_rk4_steps(wdg_wave2d op, double dt, double *t, void *d_u, int n_steps, const void *d_F, double omega,
                      void *stream);
/* Same as wdg_hip_rk4_steps with d_F = NULL, captured once into a HIP graph
 * of `steps_per_graph` steps and replayed (n_steps must be a multiple of it). */
int wdg_hip_rk4_steps_graph(wdg_wave2d op, double dt, double *t, void *d_u, int n_steps, int steps_per_graph,
                            void *stream);
/* ---- RK-stage schedules and their autotuner (no counterpart in the
 * reference, whose operators run on the CPU).  A schedule sets the
 * workgroup size of the RK stage kernels: about threads[k] threads, i.e.
 * threads[k] / n^2 whole elements per workgroup (one wavefront per element
 * for n = 8 and 64 threads).  Every schedule computes the same values in the
 * same order; the results are bitwise identical.  Setting a schedule and
 * the autotuner refuse approx_quad = 0 handles, whose fused stage kernel
 * has one workgroup shape. */
int wdg_hip_schedule_count(int *count, int *threads);
int wdg_hip_wave2d_set_schedule(wdg_wave2d op, int schedule);
int wdg_hip_wave2d_get_schedule(wdg_wave2d op, int *schedule, int *threads_per_workgroup);
/* Times `reps` launches of the RK stage kernel under every schedule on
 * private buffers (the handle's state is untouched), admits only schedules
 * whose output equals schedule 0 bit for bit, selects the fastest for this
 * handle and, for a Gauss-Lobatto basis, records it for (architecture,
 * n_colloc, precision): Gauss-Lobatto handles created afterwards in this
 * process start from it (Gauss-Legendre handles start from schedule 0 and
 * keep their tuned schedule on the handle).  ms (may be NULL) gets
 * the average launch time per schedule, -1 for a schedule not admitted.
 * Synchronises the device. */
int wdg_hip_wave2d_autotune(wdg_wave2d op, int reps, int *schedule, double *ms);
/* The tuned-schedule table as a text file: "arch n_colloc precision schedule
 * threads" per line. */
int wdg_hip_tune_cache_save(const char *path);
int wdg_hip_tune_cache_load(const char *path);
/* Average duration in milliseconds of the stage kernels launched by the last
 * wdg_hip_rk4_steps call on this handle (HIP events around them on `stream`).
 * Enabled by wdg_hip_wave2d_set_timing(op, 1). */
int wdg_hip_wave2d_set_timing(wdg_wave2d op, int enable);
int wdg_hip_wave2d_last_timing(wdg_wave2d op, double *ms_total, int *n_kernels);

/* ---- WaveHoltz -------------------------------------------------------
 * WaveHoltz(omega, ...) (WaveHoltz.cpp:5-30): T = 2 pi/omega,
 * dt = 0.5 min_h / n^2, nt = ceil(T/dt), dt = T/nt. */
int wdg_hip_waveholtz_create(wdg_wave2d op, double omega, wdg_waveholtz *out);
int wdg_hip_waveholtz_destroy(wdg_waveholtz wh);
int wdg_hip_waveholtz_info(wdg_waveholtz wh, double *T, double *dt, int *nt);
/* WaveHoltz::pi0 (WaveHoltz.cpp:55-89): d_w <- Pi(0) for the load vector d_F. */
int wdg_hip_waveholtz_pi0(wdg_waveholtz wh, double *d_w, const double *d_F, void *stream);
/* WaveHoltz::S (WaveHoltz.cpp:91-121): d_w <- S d_w, in place. */
int wdg_hip_waveholtz_S(wdg_waveholtz wh, double *d_w, void *stream);
/* One fixed-point step of the caller loop (examples/waveholtz2d.cpp:165-185):
 * d_w <- S d_w + d_pi0, *rel_change = ||w_new - w_old||_2 / pi0_norm
 * (linalg.cpp:116-130), all in double on the device.  Synchronises `stream`. */
int wdg_hip_waveholtz_iterate(wdg_waveholtz wh, double *d_w, const double *d_pi0, double pi0_norm,
                              double *rel_change, void *stream);
/* No counterpart in the reference.  enable != 0 (the default): iterate and
 * solve replay the step's launches (the nt RK4 steps of S, the update and the
 * residual sum) from a HIP graph captured on the first call for the given
 * d_w, d_pi0 and stage schedule; enable = 0: launch them one by one.  Both
 * give identical results. */
int wdg_hip_waveholtz_set_graph(wdg_waveholtz wh, int enable);
/* The caller loop of examples/waveholtz2d.cpp:165-185 run to a tolerance:
 * iterate with `wh` until rel_change <= tol, or, when `wh_refine` is not NULL,
 * until the residual stops decreasing (it stalls at a precision floor when
 * `wh` runs an FP32 state), then continue with `wh_refine` (an FP64 handle on
 * the same mesh, order and omega) until rel_change <= tol.  iters[0] counts
 * the `wh` iterations, iters[1] the refinement iterations; *final_residual is
 * the last rel_change.  Returns WDG_OK even if max_iter is reached (check
 * *final_residual). */
int wdg_hip_waveholtz_solve(wdg_waveholtz wh, wdg_waveholtz wh_refine, double *d_w, const double *d_pi0,
                            double pi0_norm, double tol, int max_iter, int *iters, double *final_residual,
                            void *stream);
/* WaveHoltz::postprocess (WaveHoltz.cpp:123-145): d_H (2, n, n, n_elem). */
int wdg_hip_waveholtz_postprocess(wdg_waveholtz wh, double *d_H, const double *d_w, void *stream);

/* ---- batched WaveHoltz -------------------------------------------------
 * Many independent WaveHoltz problems (own mesh, order, bc, omega, forcing)
 * solved together: one kernel launch per RK stage covers every problem,
 * with steps replayed from a captured HIP graph.  Per problem the sequence
 * is the reference's: pi0 (WaveHoltz.cpp:55-89), then W <- S(W) + pi0 with
 * the residual ||W_k - W_{k-1}|| / ||pi0|| (examples/waveholtz2d.cpp:165-185)
 * until it is <= tol, in double for the filter, update and residual (the RK
 * state in the batch precision).  Upwind flux.  The first update gives
 * W_1 = S(0) + pi0 = pi0 without applying S to zero, or, after
 * wdg_hip_batch_set_initial, W_1 = the given initial iterate (warm start).
 * Iteration counts are the number of W updates, as in the example loop. */
typedef struct wdg_batch_s *wdg_batch;
typedef struct {
    wdg_mesh mesh;          /* borrowed for the create call only */
    int n_colloc;           /* Gauss-Lobatto points per direction, 2..9 */
    const int *bc;          /* one code per boundary edge, NULL = reflecting */
    double omega;
    int n_sources;          /* Gaussian sources of the p-equation load vector */
    const double *sources;  /* (4, n_sources): cx, cy, sigma, amp */
} wdg_batch_problem;
int wdg_hip_batch_create(int n_problems, const wdg_batch_problem *problems, int precision, wdg_batch *out);
int wdg_hip_batch_destroy(wdg_batch b);
int wdg_hip_batch_info(wdg_batch b, int k, long long *ndof, int *nt, int *n_colloc);
/* Initial iterate of problem k (device, double, ndof); NULL clears it. */
int wdg_hip_batch_set_initial(wdg_batch b, int k, const double *d_W0);
/* Runs every problem to tol (or max_iter updates), steps_per_graph RK steps
 * per graph replay.  Synchronises `stream`. */
int wdg_hip_batch_solve(wdg_batch b, double tol, int max_iter, int steps_per_graph, void *stream);
/* Retire a problem also when its residual has not improved by 3 % for
 * `window` consecutive updates (0, the default, disables this).  For an
 * FP32 batch whose rounding floor lies above the tolerance: its iterates
 * then start an FP64 batch (wdg_hip_batch_set_initial). */
int wdg_hip_batch_set_stall_window(wdg_batch b, int window);
/* Solution (device, double, ndof; may be NULL), number of W updates and the
 * last residual of problem k. */
int wdg_hip_batch_result(wdg_batch b, int k, double *d_W, int *iterations, double *residual);
/* Wall time (HIP events) and RK steps of the last wdg_hip_batch_solve. */
int wdg_hip_batch_timing(wdg_batch b, double *seconds, long long *steps);

/* ---- load vector -----------------------------------------------------
 * LinearFunctional2D::action (LinearFunctional.hpp:136-202) with
 * Gauss-Legendre(n_colloc) quadrature, for f = sum of Gaussians in variable
 * `var` (0 = p): f(x) = sum_k amp_k exp(-|x - c_k|^2 / (2 sigma_k^2)).
 * sources is (4, n_sources) host memory: cx, cy, sigma, amp.  d_F (double) is
 * overwritten. */
int wdg_hip_load_vector_gauss(wdg_wave2d op, int n_sources, const double *sources, int var, double *d_F,
                              void *stream);

/* ---- files ------------------------------------------------------------
 * The examples' to_file (examples/examples.hpp:14-19): n values of a device
 * array written as raw native-endian float64 (FP32 data converted), the
 * format of solution/x.00000, solution/u*.00000 (examples/wave2d.cpp:191-204,
 * examples/waveholtz2d.cpp:194-197).  from_file reads such a file (its size
 * must be n values) into a device array of the given precision. */
int wdg_hip_to_file(const char *path, const void *d_x, long long n, int precision);
int wdg_hip_from_file(const char *path, void *d_x, long long n, int precision);
/* WaveHoltz checkpoint (no counterpart in the reference): the iterate W as
 * a to_file record at `path` and a text record `path`.meta holding the
 * order, element count, ndof, omega, iteration count and residual.  resume
 * checks order, mesh size and omega against the handle and loads W; passing
 * it to wdg_hip_waveholtz_solve continues the iteration where it stopped. */
int wdg_hip_waveholtz_checkpoint(wdg_waveholtz wh, const char *path, const double *d_W, int iterations,
                                 double residual);
int wdg_hip_waveholtz_resume(wdg_waveholtz wh, const char *path, double *d_W, int *iterations, double *residual);

/* ---- time integrators (ode.hpp:12-186) --------------------------------
 * ForwardEuler (method 0), RungeKutta4 (1), RungeKutta2 (2), SSPRK3 (3) for
 * any right-hand side the caller evaluates on the device: a step is, for each
 * stage k, f(p, t + c_k dt, stage input) then wdg_hip_ode_update(method, k,
 * ...).  The stage input is u for stage 0 and for every RK4 stage, w (the
 * second register) for the later RK2 and SSPRK3 stages.  ode_info gives the
 * stage count and the c_k. */
int wdg_hip_ode_info(int method, int *n_stages, double *c);
int wdg_hip_ode_update(int method, int stage, long long n, double dt, double *d_u, double *d_w, const double *d_p,
                       void *stream);

/* ---- reductions (linalg.cpp:102-130), double, deterministic ---------- */
int wdg_hip_norm(const double *d_x, long long n, double *out, void *stream);
int wdg_hip_error(const double *d_x, const double *d_y, long long n, double *out, void *stream);
/* wdg_hip_error without any device allocation.  It has no counterpart in
 * the reference.  It is for callers that cannot allocate device memory:
 * code under stream capture, tight solver loops, and harnesses that forbid
 * hipMalloc.  Its scratch is a fixed array of 8,200 bytes per GPU, which
 * holds 1024 partial sums and the total.  The reduction strides over any
 * length with a fixed grid, so every n >= 0 fits; n < 0 is refused.  The
 * result equals wdg_hip_error's bit for bit.  All calls on one GPU share
 * that GPU's scratch, so calls that overlap on one GPU (two host threads,
 * or two streams without ordering between them) race: order them.  The
 * call blocks until the result is on the host, as wdg_hip_error does. */
int wdg_hip_error_noalloc(const double *d_x, const double *d_y, long long n, double *out, void *stream);
/* Converts between the handle's precision and double (n values). */
int wdg_hip_convert(const void *d_src, int src_precision, void *d_dst, int dst_precision, long long n, void *stream);

/* ---- learned warm start -------------------------------------------------
 * No counterpart in the reference.  A four-level convolutional
 * encoder-decoder (widths base, 2 base, 4 base, 8 base; 3x3 convolutions
 * with bias and ReLU, 2x2 average pooling, nearest upsampling, skip
 * concatenation, 1x1 output layer) on M x M element images: pixel e of an
 * image is element e = ex + M ey of a uniform M x M quad grid, its channels
 * the element's nodal values.  Images are NHWC, [batch][M][M][ld]; inputs
 * bf16 (uint16 bit patterns) with in_ld channels, outputs and targets fp32
 * with out_ld = out_channels rounded up to 16 channels (the padding is
 * ignored by the loss).  Weights are bf16 with an fp32 master copy, all
 * products on bf16 MFMA with fp32 accumulation; training minimises the mean
 * over the batch of ||y - t||^2 / ||t||^2 with Adam (0.9, 0.999, 1e-8).
 * The parameter vector holds, per layer in network order, weights
 * [tap][cout][cin] then bias [cout]. */
typedef struct wdg_cnn_s *wdg_cnn;
int wdg_hip_cnn_create(int M, int in_ld, int out_channels, int base_width, int max_batch, wdg_cnn *out);
int wdg_hip_cnn_destroy(wdg_cnn n);
int wdg_hip_cnn_info(wdg_cnn n, long long *n_params, int *in_ld, int *out_ld);
/* Host fp32 parameter vectors; set_params also resets the Adam state. */
int wdg_hip_cnn_set_params(wdg_cnn n, const float *params);
int wdg_hip_cnn_get_params(wdg_cnn n, float *params);
/* d_y (device, fp32) = network(d_x) for `batch` images. */
int wdg_hip_cnn_forward(wdg_cnn n, int batch, const uint16_t *d_x, float *d_y, void *stream);
/* Loss (host, may be NULL) and its gradient w.r.t. the parameters (host
 * fp32, n_params, may be NULL) for targets d_t.  Synchronises `stream`. */
int wdg_hip_cnn_loss_grad(wdg_cnn n, int batch, const uint16_t *d_x, const float *d_t, float *loss, float *grad,
                          void *stream);
/* One Adam step on the loss of this batch; loss (host) may be NULL, which
 * leaves the call asynchronous. */
int wdg_hip_cnn_train_step(wdg_cnn n, int batch, const uint16_t *d_x, const float *d_t, float lr, float *loss,
                           void *stream);
int wdg_hip_f32_to_bf16(const float *d_src, uint16_t *d_dst, long long n, void *stream);
/* Input image of one problem (nn = n^2 channels per block):
 *   [0, nn)        gain * f(node i + n j), f the sum of the Gaussian sources
 *                  (device (4, n_sources): cx, cy, sigma, amp);
 *   [nn, 3nn)      only if green_gain != 0: green_gain * sum_k m_k
 *                  exp(-omega^2 sigma_k^2 / 2) J0 (then Y0) of omega r_k,
 *                  m_k = amp_k 2 pi sigma_k^2, r_k = sqrt(|x - c_k|^2 + sigma_k^2):
 *                  the free-space Helmholtz response of the sources;
 *   then omega_feature, and the element centre mapped from box = {x0, y0,
 *   x1, y1} to [-1, 1] (two channels); the rest 0.
 * d_nodes are the GLL nodes as from wdg_hip_mesh_nodes, on the device. */
int wdg_hip_ws_features(const double *d_nodes, int n_colloc, int nel, int n_sources, const double *d_sources,
                        double gain, double omega, double omega_feature, double green_gain, const double *box, int ld,
                        uint16_t *d_x, void *stream);
/* d_t[e][c] = d_W[e][c] / scale for the 3 n^2 values of element e, padding 0. */
int wdg_hip_ws_dg_to_image(const double *d_W, int n_colloc, int nel, double scale, int ld, float *d_t, void *stream);
/* d_W[e][c] = scale * d_y[e][c]: a network output as a WaveHoltz iterate. */
int wdg_hip_ws_image_to_dg(const float *d_y, int ld, int n_colloc, int nel, double scale, double *d_W, void *stream);

#ifdef __cplusplus
}
#endif
#endif
