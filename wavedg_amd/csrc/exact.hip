// Exact-quadrature operators (approx_quad = false) in FP64: Div<false>
// (Div.cpp:312-360, 469-588), EdgeFlux<false> on Lobatto or Legendre face
// traces (EdgeFlux.cpp:375-420, 491-620) and MassMatrix<false> (MassMatrix.cpp:
// 85-133, 193-227).  Tables are column-major (n_quad, n) with
// P(q, k) = phi_k(x_q) and D(q, k) = phi_k'(x_q).  The separate volume, face
// and mass kernels take one element per workgroup and serve any rule sizes;
// exact_stage_kernel fuses a whole RHS / RK stage for the default rules.
//
//  * volume: u -> quadrature points by the two tensor-product passes of
//    div_2d_q, contravariant fluxes there, and back with D and P;
//  * faces: for each side the element's own trace and the other trace
//    (neighbour nodes, or the boundary state built at the side's nodes) are
//    interpolated to the edge rule, the closed-form flux
//      c_p = w ds [ha n.(uA+uB) + hb (pA-pB)],  c_u = n w ds [ha (pA+pB) + hb n.(uA-uB)]
//    is evaluated there (own outward normal: +f on el0, -f on el1 of the
//    reference's edge flux) and projected onto the side's nodes; corner
//    nodes add their two sides in side order, so every output has one writer.
//    A Gauss-Legendre basis (KArgs::leg) takes each side node's trace as
//    sum_m P(m) u(m rows in) and lifts back along the same line with P;
//  * mass matrix: det J of a bilinear element is affine in (xi, eta), so the
//    exact element matrix is d0 A(x)A + d1 B(x)A + d2 A(x)B with 1D matrices
//    A, B of the mass rule; it is applied and inverted (by fast
//    diagonalisation) with tensor passes from three numbers per element
//    instead of the reference's stored Cholesky factors (linalg.cpp:5-98), and
//    the kernel finishes the RK stage or right-hand side in place.
#include <hip/hip_runtime.h>

#include "wave2d_kernels.hpp"

namespace wdg {
namespace {

constexpr int kQMax = 10;  // points of any rule
constexpr int kBlk = 128;  // threads per element workgroup

// Own outward unit normal (n0, n1) and ds = |side| / 2 of side s of the
// element with bilinear map coefficients g (its sides are straight).
__device__ __forceinline__ void side_normal(const double* g, int s, double& n0, double& n1, double& ds)
{
    const double sh = (s == 0 || s == 3) ? -1.0 : 1.0;
    const double dx = 2.0 * ((s & 1) ? g[2] + sh * g[4] : g[0] + sh * g[4]);
    const double dy = 2.0 * ((s & 1) ? g[3] + sh * g[5] : g[1] + sh * g[5]);
    const double l2 = dx * dx + dy * dy, il = rsqrt_nr(l2);
    const double sil = (s >= 2) ? -il : il;
    n0 = dy * sil;
    n1 = -dx * sil;
    ds = 0.5 * (l2 * il);
}

// Exterior state at a boundary node for boundary code `code` and outward
// normal (n0, n1).
__device__ __forceinline__ void boundary_state(int code, double n0, double n1, double pA, double uA, double vA,
                                               double& pB, double& uB, double& vB)
{
    if (code == 1) {  // reflecting (WaveEquation.hpp:64-69)
        pB = pA;
        uB = (n1 * n1 - n0 * n0) * uA - 2.0 * n0 * n1 * vA;
        vB = -2.0 * n0 * n1 * uA + (n0 * n0 - n1 * n1) * vA;
    } else if (code == 0) {  // characteristic outflow, p_ext after u_ext
        const double o = pA + n0 * uA + n1 * vA, tang = n1 * uA - n0 * vA;
        uB = n1 * tang + 0.5 * n0 * o;
        vB = -n0 * tang + 0.5 * n1 * o;
        pB = n0 * uB + n1 * vB;
    } else {  // Dirichlet p = 0 (mirror state)
        pB = -pA;
        uB = uA;
        vB = vA;
    }
}

template <int N>
__global__ __launch_bounds__(kBlk) void exact_volume_kernel(const KArgs a, const ExactTables t, int accumulate)
{
    constexpr int NN = N * N;
    __shared__ double s_u[3 * NN], s_Pu[N * 3 * kQMax], s_Uq[3 * kQMax * kQMax], s_F[2 * 3 * kQMax * kQMax];
    __shared__ double s_Df[kQMax * 3 * N], s_Pg[kQMax * 3 * N], s_P[kQMax * N], s_D[kQMax * N], s_x[kQMax],
        s_w[kQMax];
    const int e = blockIdx.x, tid = threadIdx.x, nq = t.nqv;
    const double* u = static_cast<const double*>(a.u_in) + (size_t)e * 3 * NN;
    for (int k = tid; k < 3 * NN; k += kBlk) s_u[k] = u[k];
    for (int k = tid; k < nq * N; k += kBlk) {
        s_P[k] = t.Pv[k];
        s_D[k] = t.Dv[k];
    }
    if (tid < nq) {
        s_x[tid] = t.xv[tid];
        s_w[tid] = t.wv[tid];
    }
    __syncthreads();
    // Pu(l, d, i) = sum_k P(i, k) u(d, k, l)
    for (int k = tid; k < N * 3 * nq; k += kBlk) {
        const int l = k % N, d = (k / N) % 3, i = k / (3 * N);
        double pu = 0.0;
        for (int kk = 0; kk < N; ++kk) pu += s_P[i + nq * kk] * s_u[d + 3 * (kk + N * l)];
        s_Pu[l + N * (d + 3 * i)] = pu;
    }
    __syncthreads();
    // Uq(d, i, j) = sum_l P(j, l) Pu(l, d, i)
    for (int k = tid; k < 3 * nq * nq; k += kBlk) {
        const int d = k % 3, i = (k / 3) % nq, j = k / (3 * nq);
        double uq = 0.0;
        for (int l = 0; l < N; ++l) uq += s_P[j + nq * l] * s_Pu[l + N * (d + 3 * i)];
        s_Uq[d + 3 * (i + nq * j)] = uq;
    }
    __syncthreads();
    // contravariant fluxes of the wave equation: f = (u, p, 0), g = (v, 0, p)
    const double2* gp = reinterpret_cast<const double2*>(a.geom + 6 * (size_t)e);
    const double2 g_a = gp[0], g_c = gp[1], g_h = gp[2];
    for (int k = tid; k < nq * nq; k += kBlk) {
        const int i = k % nq, j = k / nq;
        const double xi = s_x[i], eta = s_x[j], W = s_w[i] * s_w[j];
        const double X_xi = (g_a.x + g_h.x * eta) * W, Y_xi = (g_a.y + g_h.y * eta) * W;
        const double X_eta = (g_c.x + g_h.x * xi) * W, Y_eta = (g_c.y + g_h.y * xi) * W;
        const double p = s_Uq[0 + 3 * k], uu = s_Uq[1 + 3 * k], vv = s_Uq[2 + 3 * k];
        double* F = s_F + 2 * 3 * k;  // (2, d)
        F[0 + 2 * 0] = Y_eta * uu - X_eta * vv;
        F[1 + 2 * 0] = -Y_xi * uu + X_xi * vv;
        F[0 + 2 * 1] = Y_eta * p;
        F[1 + 2 * 1] = -Y_xi * p;
        F[0 + 2 * 2] = -X_eta * p;
        F[1 + 2 * 2] = X_xi * p;
    }
    __syncthreads();
    // Df(j, d, k) = sum_i D(i, k) F(d, i, j); Pg(j, d, k) = sum_i P(i, k) G(d, i, j)
    for (int k = tid; k < nq * 3 * N; k += kBlk) {
        const int j = k % nq, d = (k / nq) % 3, kk = k / (3 * nq);
        double df = 0.0, pg = 0.0;
        for (int i = 0; i < nq; ++i) {
            df += s_D[i + nq * kk] * s_F[0 + 2 * (d + 3 * (i + nq * j))];
            pg += s_P[i + nq * kk] * s_F[1 + 2 * (d + 3 * (i + nq * j))];
        }
        s_Df[j + nq * (d + 3 * kk)] = df;
        s_Pg[j + nq * (d + 3 * kk)] = pg;
    }
    __syncthreads();
    double* y = static_cast<double*>(a.u_out) + (size_t)e * 3 * NN;
    for (int k = tid; k < 3 * NN; k += kBlk) {
        const int d = k % 3, kk = (k / 3) % N, l = k / (3 * N);
        double divF = 0.0;
        for (int j = 0; j < nq; ++j)
            divF += s_P[j + nq * l] * s_Df[j + nq * (d + 3 * kk)] + s_D[j + nq * l] * s_Pg[j + nq * (d + 3 * kk)];
        y[k] = accumulate ? y[k] + divF : divF;
    }
}

// Trace (p, u, v) on element side s at the side node whose state starts at q:
// the node's values for a Gauss-Lobatto basis, sum_m P(m) u(m rows in) for a
// Gauss-Legendre basis (legendre_action_2d, FaceProlongator.cpp:809-835).
template <int N>
__device__ __forceinline__ void side_trace(const double* q, int s, int leg, const double* P, double& p, double& uu,
                                           double& vv)
{
    if (!leg) {
        p = q[0];
        uu = q[1];
        vv = q[2];
        return;
    }
    const int st = 3 * inward_stride<N>(s);
    p = uu = vv = 0.0;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        const double w = P[m];
        p = fma(w, q[m * st + 0], p);
        uu = fma(w, q[m * st + 1], uu);
        vv = fma(w, q[m * st + 2], vv);
    }
}

template <int N>
__global__ __launch_bounds__(64) void exact_face_kernel(const KArgs a, const ExactTables t, int interior, int boundary)
{
    constexpr int NN = N * N;
    __shared__ double s_out[3 * NN], s_A[3 * N], s_B[3 * N], s_c[3 * kQMax], s_P[kQMax * N], s_w[kQMax];
    const int e = blockIdx.x, tid = threadIdx.x, nq = t.nqe;
    const double* u = static_cast<const double*>(a.u_in);
    for (int k = tid; k < 3 * NN; k += 64) s_out[k] = 0.0;
    for (int k = tid; k < nq * N; k += 64) s_P[k] = t.Pe[k];
    if (tid < nq) s_w[tid] = t.we[tid];
    const double* g = a.geom + 6 * (size_t)e;
    bool any = false;
    for (int s = 0; s < 4; ++s) {
        const int nb = a.nbr[4 * (size_t)e + s], inf = a.info[4 * (size_t)e + s];
        const bool is_int = nb >= 0;
        if (is_int ? !interior : !boundary) continue;  // uniform over the workgroup
        any = true;
        double n0, n1, ds;
        side_normal(g, s, n0, n1, ds);
        __syncthreads();  // the previous side's traces and fluxes are consumed
        if (tid < N) {
            const int k = tid;
            double pA, uA, vA;
            side_trace<N>(u + ((size_t)e * NN + side_node<N>(s, k)) * 3, s, a.leg, t.Pf, pA, uA, vA);
            double pB, uB, vB;
            if (is_int) {
                const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
                side_trace<N>(u + ((size_t)nb * NN + side_node<N>(inf & 3, kk)) * 3, inf & 3, a.leg, t.Pf, pB, uB,
                              vB);
            } else {
                boundary_state((inf >> 8) & 0xff, n0, n1, pA, uA, vA, pB, uB, vB);
            }
            s_A[0 * N + k] = pA;
            s_A[1 * N + k] = uA;
            s_A[2 * N + k] = vA;
            s_B[0 * N + k] = pB;
            s_B[1 * N + k] = uB;
            s_B[2 * N + k] = vB;
        }
        __syncthreads();
        if (tid < nq) {
            const int q = tid;
            double A[3], B[3];
            for (int d = 0; d < 3; ++d) {
                double va = 0.0, vb = 0.0;
                for (int k = 0; k < N; ++k) {
                    va += s_A[d * N + k] * s_P[q + nq * k];
                    vb += s_B[d * N + k] * s_P[q + nq * k];
                }
                A[d] = va;
                B[d] = vb;
            }
            const double wds = s_w[q] * ds;
            s_c[0 * kQMax + q] = wds * (a.ha * (n0 * (A[1] + B[1]) + n1 * (A[2] + B[2])) + a.hb * (A[0] - B[0]));
            const double cf = wds * (a.ha * (A[0] + B[0]) + a.hb * (n0 * (A[1] - B[1]) + n1 * (A[2] - B[2])));
            s_c[1 * kQMax + q] = n0 * cf;
            s_c[2 * kQMax + q] = n1 * cf;
        }
        __syncthreads();
        if (tid < N) {
            // lane k owns the line of nodes through side node k
            const int k = tid, node = side_node<N>(s, k), st = inward_stride<N>(s);
            for (int d = 0; d < 3; ++d) {
                double v = 0.0;
                for (int q = 0; q < nq; ++q) v += s_c[d * kQMax + q] * s_P[q + nq * k];
                if (a.leg) {
                    for (int m = 0; m < N; ++m) s_out[d + 3 * (node + m * st)] += v * t.Pf[m];
                } else {
                    s_out[d + 3 * node] += v;
                }
            }
        }
    }
    if (!any) return;  // uniform
    __syncthreads();
    double* y = static_cast<double*>(a.u_out) + (size_t)e * 3 * NN;
    for (int k = tid; k < 3 * NN; k += 64) y[k] += s_out[k];
}

// det J = d0 + d1 xi + d2 eta of the element's bilinear map (the xi eta terms
// of x_xi y_eta - y_xi x_eta cancel).
__device__ __forceinline__ void detj_affine(const double* g, double& d0, double& d1, double& d2)
{
    d0 = g[0] * g[3] - g[1] * g[2];
    d1 = g[0] * g[5] - g[1] * g[4];
    d2 = g[3] * g[4] - g[2] * g[5];
}

// Mass-matrix application per element, M = d0 A(x)A + d1 B(x)A + d2 A(x)B:
//   kid < 0: -1 y = M x, -2 y = M^{-1} x (x = a.u_in, y = a.u_out; may alias);
//   kid >= 0: the RHS / RK-stage epilogue of wave_kernel for kModeOf[kid],
//             with the operator's output r (before M^{-1}) in `r`.
// M^{-1} by fast diagonalisation: (V(x)V) diag(1 / (d0 + d1 lam_a + d2 lam_b)) (V(x)V)'.
template <int N>
__global__ __launch_bounds__(kBlk) void exact_mass_kernel(const KArgs a, const double* r, const ExactMassTables m,
                                                          int kid, int nv)
{
    constexpr int C = N * N;
    constexpr int MV = DivFArgs::kMaxVar;
    __shared__ double s_x[MV * C], s_t[MV * C], s_t2[MV * C], s_V[C], s_W[C], s_lam[N];
    const int e = blockIdx.x, tid = threadIdx.x;
    const size_t base = (size_t)e * nv * C;
    const int mode = kid >= 0 ? kModeOf[kid] : 0;
    const bool force = kid >= 0 && (mode & M_FORCE);
    const bool mult = kid == -1;
    for (int k = tid; k < nv * C; k += kBlk) {
        double v = kid < 0 ? static_cast<const double*>(a.u_in)[base + k] : r[base + k];
        if (force) v += static_cast<const double*>(a.force)[base + k] * a.fs;
        s_x[k] = v;
    }
    for (int k = tid; k < C; k += kBlk) {
        s_V[k] = mult ? m.A[k] : m.V[k];
        s_W[k] = m.B[k];
    }
    if (tid < N) s_lam[tid] = m.lam[tid];
    double d0, d1, d2;
    detj_affine(a.geom + 6 * (size_t)e, d0, d1, d2);
    __syncthreads();
    if (mult) {
        // (d0 A x + d1 B x, d2 A x) along xi, then A and B along eta
        for (int k = tid; k < nv * C; k += kBlk) {
            const int d = k % nv, i = (k / nv) % N, j = k / (nv * N);
            double xa = 0.0, xb = 0.0;
            for (int l = 0; l < N; ++l) {
                const double x = s_x[d + nv * (l + N * j)];
                xa += s_V[i + N * l] * x;
                xb += s_W[i + N * l] * x;
            }
            s_t[k] = d0 * xa + d1 * xb;
            s_t2[k] = d2 * xa;
        }
        __syncthreads();
        for (int k = tid; k < nv * C; k += kBlk) {
            const int d = k % nv, i = (k / nv) % N, j = k / (nv * N);
            double y = 0.0;
            for (int l = 0; l < N; ++l)
                y += s_V[j + N * l] * s_t[d + nv * (i + N * l)] + s_W[j + N * l] * s_t2[d + nv * (i + N * l)];
            s_x[k] = y;  // s_x is not read in this pass
        }
    } else {
        // t(d, a, j) = sum_i V(i, a) x(d, i, j)
        for (int k = tid; k < nv * C; k += kBlk) {
            const int d = k % nv, aa = (k / nv) % N, j = k / (nv * N);
            double t = 0.0;
            for (int i = 0; i < N; ++i) t += s_V[i + N * aa] * s_x[d + nv * (i + N * j)];
            s_t[k] = t;
        }
        __syncthreads();
        // x(d, a, b) = sum_j V(j, b) t(d, a, j) / (d0 + d1 lam_a + d2 lam_b)
        for (int k = tid; k < nv * C; k += kBlk) {
            const int d = k % nv, aa = (k / nv) % N, b = k / (nv * N);
            double t = 0.0;
            for (int j = 0; j < N; ++j) t += s_V[j + N * b] * s_t[d + nv * (aa + N * j)];
            s_x[k] = t / (d0 + d1 * s_lam[aa] + d2 * s_lam[b]);
        }
        __syncthreads();
        // t(d, i, b) = sum_a V(i, a) x(d, a, b)
        for (int k = tid; k < nv * C; k += kBlk) {
            const int d = k % nv, i = (k / nv) % N, b = k / (nv * N);
            double t = 0.0;
            for (int aa = 0; aa < N; ++aa) t += s_V[i + N * aa] * s_x[d + nv * (aa + N * b)];
            s_t[k] = t;
        }
        __syncthreads();
        // x(d, i, j) = sum_b V(j, b) t(d, i, b)
        for (int k = tid; k < nv * C; k += kBlk) {
            const int d = k % nv, i = (k / nv) % N, j = k / (nv * N);
            double t = 0.0;
            for (int b = 0; b < N; ++b) t += s_V[j + N * b] * s_t[d + nv * (i + N * b)];
            s_x[k] = t;
        }
    }
    __syncthreads();
    const int out = mode & M_OUTMASK;
    for (int k = tid; k < nv * C; k += kBlk) {
        const size_t g = base + k;
        const double rhs = s_x[k];
        if (kid < 0 || out == M_OUT_RHS) {
            static_cast<double*>(a.u_out)[g] = rhs;
            continue;
        }
        double* du = static_cast<double*>(a.du);
        const double d = (out == M_OUT_STAGE0 ? 0.0 : a.rk_a * du[g]) + a.dt * rhs;
        du[g] = d;
        const double w = static_cast<const double*>(a.u_in)[g] + a.rk_b * d;
        static_cast<double*>(a.u_out)[g] = w;
        if (mode & M_ACC) a.acc[g] += a.ac * w;
    }
}

// ---------------------------------------------------------------- fused stage
// The whole right-hand side / RK stage for the default rules (QV = N volume,
// QE = N + 1 edge points) in one kernel.  E elements share an NT-lane
// workgroup.  In every tensor pass a lane owns one (element, variable, line),
// keeps the line in registers and reads the 1D tables as uniform operands
// through the scalar cache (row-contiguous copies).  Two LDS regions per
// element carry the transposes between passes:
//   RA: u -> u at the volume points -> edge-point traces -> t1 -> M^{-1} r
//   RB: xi-interpolated u -> Df, then Pg -> staged traces -> edge fluxes and
//       their projections -> t3
// The operator output r stays in the registers of the lane that owns its
// line from the volume pass to the mass matrix; the face projections are
// added there.  Arrays read along a line of lanes have odd strides in
// doubles (the L* pitches), so a wave's reads spread over the LDS banks.
template <int N>
struct StageShape {
    static constexpr int cmax(int x, int y) { return x > y ? x : y; }
    static constexpr int odd(int x) { return x | 1; }
    static constexpr int QV = N, QE = N + 1, NN = N * N;
    static constexpr int NT = (N == 6 || N == 8) ? 128 : 64;  // fills the lanes best
    static constexpr int E = NT / (3 * N);                     // 3N line lanes per element
    static constexpr int LS = odd(3 * N);    // (3, N, N) arrays: d + 3 i + LS j
    static constexpr int LQ = odd(3 * QV);   // u at the volume points: d + 3 i + LQ j
    static constexpr int LP = odd(N);        // xi-interpolated u: (d + 3 i) LP + l
    static constexpr int LD = odd(QV);       // Df or Pg: (d + 3 k) LD + j
    static constexpr int LT = odd(N);        // staged side traces: row LT
    static constexpr int LE = odd(QE);       // edge-point traces and fluxes: row LE
    static constexpr int SA = cmax(cmax(LS * N, LQ * QV), 24 * LE);
    static constexpr int SB = cmax(cmax(3 * QV * LP, 3 * N * LD), cmax(cmax(24 * LT, 12 * LE + 12 * N), LS * N));
    static constexpr int S = SA + SB;
};

// out_o = sum_k T[o NK + k] x[k] for o = 0 .. NO - 1, with T a uniform
// table read through the scalar cache.  The row of output o + 1 is requested
// while output o is computed; the compiler barrier keeps later rows from all
// being hoisted to the top, which would exhaust the scalar registers.
template <int NO, int NK, typename Store>
__device__ __forceinline__ void contract_rows(const double* __restrict__ T, const double (&x)[NK], Store store)
{
    double r[2][NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) r[0][k] = T[k];
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        if (o + 1 < NO) {
#pragma unroll
            for (int k = 0; k < NK; ++k) r[(o + 1) % 2][k] = T[(o + 1) * NK + k];
        }
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < NK; ++k) acc += r[o % 2][k] * x[k];
        store(o, acc);
        asm volatile("" ::: "memory");
    }
}

// At n = 8 on the Gauss-Lobatto basis the kernel fits 96 VGPRs without
// spilling, so it asks for 5 waves per SIMD instead of the 4 that its
// natural 97 allow (the LDS tiles allow 4.5 on average).
template <int N, bool LEG>
__global__ __launch_bounds__(StageShape<N>::NT, (N == 8 && !LEG) ? 5 : 1) void exact_stage_kernel(
    const KArgs a, const double* __restrict__ tab, int kid)
{
    using Sh = StageShape<N>;
    constexpr int QV = Sh::QV, QE = Sh::QE, NN = Sh::NN, E = Sh::E, NT = Sh::NT;
    constexpr int LS = Sh::LS, LQ = Sh::LQ, LP = Sh::LP, LD = Sh::LD, LT = Sh::LT, LE = Sh::LE;
    constexpr int NS = (E * 3 * NN + NT - 1) / NT;  // state values per lane
    static_assert(QV == N, "the line lanes of the volume passes assume QV == N");
    __shared__ double s_mem[E * Sh::S];
    __shared__ double s_geo[E * 6];
    // the epilogue's scalars, parked in LDS so that they hold no scalar
    // registers through the passes (which would spill them)
    __shared__ double s_par[4];
    __shared__ double* s_ptr[3];
    // each XCD walks one contiguous run of elements, so neighbour traces and
    // the block's own side traces are read through the L2 that loaded them
    const int lane = threadIdx.x, e0 = xcd_block(blockIdx.x, gridDim.x) * E, nel = a.nel;
    const int mode = kModeOf[kid];
    if (lane == 0) {
        s_par[0] = a.rk_a;
        s_par[1] = a.rk_b;
        s_par[2] = a.dt;
        s_par[3] = a.ac;
        s_ptr[0] = static_cast<double*>(a.u_out);
        s_ptr[1] = static_cast<double*>(a.du);
        s_ptr[2] = a.acc;
    }
    const double* __restrict__ u = static_cast<const double*>(a.u_in);
    auto RA = [&](int e) { return s_mem + e * Sh::S; };
    auto RB = [&](int e) { return s_mem + e * Sh::S + Sh::SA; };
    auto sidx = [](int d, int i, int j) { return d + 3 * i + LS * j; };
    const int nvalid = nel - e0 < E ? nel - e0 : E;
    // line lanes: element le, variable ld, line lk
    const int le = lane / (3 * N), lj = lane % (3 * N), ld = lj % 3, lk = lj / 3;
    const bool lon = le < nvalid;

    // the block's elements are contiguous in HBM: every load is issued before
    // the first wait
    const double* ublk = u + (size_t)e0 * 3 * NN;
    double v[NS];
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int it = lane + r * NT;
        v[r] = it < nvalid * 3 * NN ? ublk[it] : 0.0;
    }
    const double gv = lane < nvalid * 6 ? a.geom[6 * (size_t)e0 + lane] : 0.0;
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int it = lane + r * NT, e = it / (3 * NN), k = it % (3 * NN);
        if (it < nvalid * 3 * NN) RA(e)[k + (LS - 3 * N) * (k / (3 * N))] = v[r];
    }
    if (lane < nvalid * 6) s_geo[lane] = gv;
    __syncthreads();
    // ---- volume: the passes of div_2d_q (Div.cpp:469-588)
    // PU(l, d, i) = sum_k P(i, k) u(d, k, l)          lane (d, l)
    if (lon) {
        double x[N];
#pragma unroll
        for (int k = 0; k < N; ++k) x[k] = RA(le)[sidx(ld, k, lk)];
        double* PU = RB(le);
        contract_rows<QV, N>(tab + kTabPVT, x, [&](int i, double acc) { PU[(ld + 3 * i) * LP + lk] = acc; });
    }
    __syncthreads();
    // UQ(d, i, j) = sum_l P(j, l) PU(l, d, i)          lane (d, i)
    if (lon) {
        const double* PU = RB(le) + (ld + 3 * lk) * LP;
        double x[N];
#pragma unroll
        for (int l = 0; l < N; ++l) x[l] = PU[l];
        double* UQ = RA(le);
        contract_rows<QV, N>(tab + kTabPVT, x, [&](int jq, double acc) { UQ[ld + 3 * lk + LQ * jq] = acc; });
    }
    __syncthreads();
    // contravariant fluxes of f = (u, p, 0), g = (v, 0, p) at the points, then
    //   Df(j, d, k) = sum_i D(i, k) F(d, i, j)   (lane (d, j)),  r  = sum_j P(j, l) Df(j, d, k)
    //   Pg(j, d, k) = sum_i P(i, k) G(d, i, j),                  r += sum_j D(j, l) Pg(j, d, k)
    // with r(d, k, :) in the registers of lane (d, k)
    double R[N];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        if (lon) {
            const int jq = lk;
            const double* g = s_geo + 6 * le;
            const double ax = g[0], ay = g[1], cx = g[2], cy = g[3], hx = g[4], hy = g[5];
            const double eta = tab[kTabXV + jq], wj = tab[kTabWV + jq];
            const double* UQ = RA(le) + LQ * jq;
            double acc[N];
#pragma unroll
            for (int k = 0; k < N; ++k) acc[k] = 0.0;
#pragma unroll 2
            for (int i = 0; i < QV; ++i) {
                const double xi = tab[kTabXV + i], W = tab[kTabWV + i] * wj;
                const double p = UQ[3 * i], uu = UQ[3 * i + 1], vv = UQ[3 * i + 2];
                const double s1 = ld == 0 ? uu : (ld == 1 ? p : 0.0);
                const double s2 = ld == 0 ? vv : (ld == 2 ? p : 0.0);
                double f;
                if (half == 0) {
                    const double X_eta = (cx + hx * xi) * W, Y_eta = (cy + hy * xi) * W;
                    f = Y_eta * s1 - X_eta * s2;
                } else {
                    const double X_xi = (ax + hx * eta) * W, Y_xi = (ay + hy * eta) * W;
                    f = -Y_xi * s1 + X_xi * s2;
                }
#pragma unroll
                for (int k = 0; k < N; ++k) acc[k] += tab[(half == 0 ? kTabDVT : kTabPVT) + k + N * i] * f;
            }
            double* DP = RB(le);
#pragma unroll
            for (int k = 0; k < N; ++k) DP[(ld + 3 * k) * LD + jq] = acc[k];
        }
        __syncthreads();
        if (lon) {
            const double* DP = RB(le) + (ld + 3 * lk) * LD;
            double x[QV];
#pragma unroll
            for (int jq = 0; jq < QV; ++jq) x[jq] = DP[jq];
            contract_rows<N, QV>(tab + (half == 0 ? kTabPV : kTabDV), x, [&](int l, double acc) {
                if (half == 0)
                    R[l] = acc;
                else
                    R[l] += acc;
            });
        }
        __syncthreads();
    }
    // ---- faces (EdgeFlux<false>): own and exterior traces at the side's nodes
    {
        constexpr int NF = (E * 4 * N + NT - 1) / NT;
        int nb[NF], inf[NF];
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int it = lane + r * NT, e = it / (4 * N), s = (it % (4 * N)) / N;
            const bool ok = it < nvalid * 4 * N;
            nb[r] = ok ? a.nbr[4 * (size_t)(e0 + e) + s] : -1;
            inf[r] = ok ? a.info[4 * (size_t)(e0 + e) + s] : 0;
        }
        double tr[NF][6];
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int it = lane + r * NT, e = it / (4 * N), j = it % (4 * N), s = j / N, k = j % N;
            if (it >= nvalid * 4 * N) continue;
            side_trace<N>(u + ((size_t)(e0 + e) * NN + side_node<N>(s, k)) * 3, s, LEG, tab + kTabPF, tr[r][0],
                          tr[r][1], tr[r][2]);
            if (nb[r] >= 0) {
                const int kk = ((inf[r] >> 2) & 1) ? (N - 1 - k) : k;
                side_trace<N>(u + ((size_t)nb[r] * NN + side_node<N>(inf[r] & 3, kk)) * 3, inf[r] & 3, LEG,
                              tab + kTabPF, tr[r][3], tr[r][4], tr[r][5]);
            }
        }
#pragma unroll
        for (int r = 0; r < NF; ++r) {
            const int it = lane + r * NT, e = it / (4 * N), j = it % (4 * N), s = j / N, k = j % N;
            if (it >= nvalid * 4 * N) continue;
            if (nb[r] < 0) {
                double n0, n1, ds;
                side_normal(s_geo + 6 * e, s, n0, n1, ds);
                boundary_state((inf[r] >> 8) & 0xff, n0, n1, tr[r][0], tr[r][1], tr[r][2], tr[r][3], tr[r][4],
                               tr[r][5]);
            }
            double* T = RB(e) + 6 * LT * s + k;  // rows: own p, u, v, exterior p, u, v
#pragma unroll
            for (int c = 0; c < 6; ++c) T[c * LT] = tr[r][c];
        }
    }
    __syncthreads();
    // traces at the edge points
    for (int it = lane; it < E * 24; it += NT) {
        const int e = it / 24, j = it % 24;  // row j = 6 s + 3 (own | exterior) + d
        if (e >= nvalid) continue;
        const double* T = RB(e) + LT * j;
        double x[N];
#pragma unroll
        for (int k = 0; k < N; ++k) x[k] = T[k];
        double* Q = RA(e) + LE * j;
        contract_rows<QE, N>(tab + kTabPET, x, [&](int q, double acc) { Q[q] = acc; });
    }
    __syncthreads();
    // closed-form flux at every edge point, weighted by w ds
    for (int it = lane; it < E * 4 * QE; it += NT) {
        const int e = it / (4 * QE), j = it % (4 * QE), s = j / QE, q = j % QE;
        if (e >= nvalid) continue;
        double n0, n1, ds;
        side_normal(s_geo + 6 * e, s, n0, n1, ds);
        const double* Q = RA(e) + 6 * LE * s + q;
        const double A0 = Q[0 * LE], A1 = Q[1 * LE], A2 = Q[2 * LE];
        const double B0 = Q[3 * LE], B1 = Q[4 * LE], B2 = Q[5 * LE];
        const double wds = tab[kTabWE + q] * ds;
        double* Cf = RB(e) + 3 * LE * s + q;
        Cf[0 * LE] = wds * (a.ha * (n0 * (A1 + B1) + n1 * (A2 + B2)) + a.hb * (A0 - B0));
        const double cf = wds * (a.ha * (A0 + B0) + a.hb * (n0 * (A1 - B1) + n1 * (A2 - B2)));
        Cf[1 * LE] = n0 * cf;
        Cf[2 * LE] = n1 * cf;
    }
    __syncthreads();
    // projection onto each side's nodes: FP(s, d, k)
    for (int it = lane; it < E * 12; it += NT) {
        const int e = it / 12, j = it % 12;  // j = 3 s + d
        if (e >= nvalid) continue;
        const double* Cf = RB(e) + LE * j;
        double c[QE];
#pragma unroll
        for (int q = 0; q < QE; ++q) c[q] = Cf[q];
        double* FP = RB(e) + 12 * LE + N * j;
        contract_rows<N, QE>(tab + kTabPE, c, [&](int k, double val) { FP[k] = val; });
    }
    __syncthreads();
    // ---- M^{-1} (r + s F) by fast diagonalisation
    // lane (d, i): r(d, i, :) += faces and forcing; t1(d, i, b) = sum_j V(j, b) r(d, i, j)
    if (lon) {
        if (mode & M_FORCE) {
            const double* F = static_cast<const double*>(a.force) + (size_t)(e0 + le) * 3 * NN + ld + 3 * lk;
#pragma unroll
            for (int l = 0; l < N; ++l) R[l] += F[3 * N * l] * a.fs;
        }
        const double* FP = RB(le) + 12 * LE + N * ld;  // side s at + 3 N s
        if constexpr (LEG) {
            // node (i, l) is row l of side 0, row N-1-i of side 1, row N-1-l
            // of side 2 and row i of side 3 (legendre_transpose_2d,
            // FaceProlongator.cpp:908-934)
            const double* Pf = tab + kTabPF;
            const double f0 = FP[0 * 3 * N + lk], f2 = FP[2 * 3 * N + lk];
            const double p1 = Pf[N - 1 - lk], p3 = Pf[lk];
#pragma unroll
            for (int l = 0; l < N; ++l)
                R[l] += Pf[l] * f0 + Pf[N - 1 - l] * f2 + p1 * FP[1 * 3 * N + l] + p3 * FP[3 * 3 * N + l];
        } else {
            R[0] += FP[0 * 3 * N + lk];
            R[N - 1] += FP[2 * 3 * N + lk];
            if (lk == N - 1) {
#pragma unroll
                for (int l = 0; l < N; ++l) R[l] += FP[1 * 3 * N + l];
            }
            if (lk == 0) {
#pragma unroll
                for (int l = 0; l < N; ++l) R[l] += FP[3 * 3 * N + l];
            }
        }
        double* T1 = RA(le);
        contract_rows<N, N>(tab + kTabV, R, [&](int b, double t) { T1[sidx(ld, lk, b)] = t; });
    }
    __syncthreads();
    // lane (d, b): t2(d, a, b) = sum_i V(i, a) t1(d, i, b) / (d0 + d1 lam_a + d2 lam_b),
    //              t3(d, i, b) = sum_a V(i, a) t2(d, a, b)
    if (lon) {
        double x[N];
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = RA(le)[sidx(ld, i, lk)];
        double d0, d1, d2;
        detj_affine(s_geo + 6 * le, d0, d1, d2);
        const double db = d0 + d2 * tab[kTabLam + lk];
        double t2[N];
        contract_rows<N, N>(tab + kTabV, x, [&](int aa, double t) { t2[aa] = t * recip(db + d1 * tab[kTabLam + aa]); });
        double* T3 = RB(le);
        contract_rows<N, N>(tab + kTabVT, t2, [&](int i, double t) { T3[sidx(ld, i, lk)] = t; });
    }
    __syncthreads();
    // lane (d, i): y(d, i, j) = sum_b V(j, b) t3(d, i, b)
    if (lon) {
        double x[N];
#pragma unroll
        for (int b = 0; b < N; ++b) x[b] = RB(le)[sidx(ld, lk, b)];
        double* Y = RA(le);
        contract_rows<N, N>(tab + kTabVT, x, [&](int j, double t) { Y[sidx(ld, lk, j)] = t; });
    }
    __syncthreads();
    // ---- epilogue: right-hand side, or the RK stage (and WaveHoltz accumulation)
    const int out = mode & M_OUTMASK;
    const size_t gb = (size_t)e0 * 3 * NN;
    const volatile double* vpar = s_par;
    double* const volatile* vptr = s_ptr;
    double* const uo = vptr[0] + gb;
    if (out == M_OUT_RHS) {
#pragma unroll
        for (int r = 0; r < NS; ++r) {
            const int it = lane + r * NT, e = it / (3 * NN), k = it % (3 * NN);
            if (it < nvalid * 3 * NN) uo[it] = RA(e)[k + (LS - 3 * N) * (k / (3 * N))];
        }
        return;
    }
    const double rk_a = vpar[0], rk_b = vpar[1], dt = vpar[2], ac = vpar[3];
    double* const du = vptr[1] + gb;
    double* const acc = vptr[2] + gb;
    double dv[NS], av[NS];
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int it = lane + r * NT;
        const bool ok = it < nvalid * 3 * NN;
        dv[r] = ok && out != M_OUT_STAGE0 ? du[it] : 0.0;
        av[r] = ok && (mode & M_ACC) ? acc[it] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NS; ++r) {
        const int it = lane + r * NT, e = it / (3 * NN), k = it % (3 * NN);
        if (it >= nvalid * 3 * NN) continue;
        const double rhs = RA(e)[k + (LS - 3 * N) * (k / (3 * N))];
        const double dd = (out == M_OUT_STAGE0 ? 0.0 : rk_a * dv[r]) + dt * rhs;
        du[it] = dd;
        const double w = v[r] + rk_b * dd;
        uo[it] = w;
        if (mode & M_ACC) acc[it] = av[r] + ac * w;
    }
}

template <int N>
void launch_stage(int kid, const KArgs& a, const double* tab, hipStream_t s)
{
    constexpr int E = StageShape<N>::E;
    if (a.leg)
        hipLaunchKernelGGL((exact_stage_kernel<N, true>), dim3((a.nel + E - 1) / E), dim3(StageShape<N>::NT), 0, s, a,
                           tab, kid);
    else
        hipLaunchKernelGGL((exact_stage_kernel<N, false>), dim3((a.nel + E - 1) / E), dim3(StageShape<N>::NT), 0, s,
                           a, tab, kid);
}

template <int N>
void launch_all(int what, const KArgs& a, const ExactTables& t, const ExactMassTables& m, const double* r, int p0,
                int p1, hipStream_t s)
{
    const dim3 grid(a.nel);
    switch (what) {
    case 0: hipLaunchKernelGGL(exact_volume_kernel<N>, grid, dim3(kBlk), 0, s, a, t, p0); break;
    case 1: hipLaunchKernelGGL(exact_face_kernel<N>, grid, dim3(64), 0, s, a, t, p0, p1); break;
    default: hipLaunchKernelGGL(exact_mass_kernel<N>, grid, dim3(kBlk), 0, s, a, r, m, p0, p1); break;
    }
}

void dispatch(int n, int what, const KArgs& a, const ExactTables& t, const ExactMassTables& m, const double* r, int p0,
              int p1, hipStream_t s)
{
    if (a.nel <= 0) return;
    switch (n) {
    case 2: launch_all<2>(what, a, t, m, r, p0, p1, s); break;
    case 3: launch_all<3>(what, a, t, m, r, p0, p1, s); break;
    case 4: launch_all<4>(what, a, t, m, r, p0, p1, s); break;
    case 5: launch_all<5>(what, a, t, m, r, p0, p1, s); break;
    case 6: launch_all<6>(what, a, t, m, r, p0, p1, s); break;
    case 7: launch_all<7>(what, a, t, m, r, p0, p1, s); break;
    case 8: launch_all<8>(what, a, t, m, r, p0, p1, s); break;
    case 9: launch_all<9>(what, a, t, m, r, p0, p1, s); break;
    default: break;
    }
}

}  // namespace

void launch_exact_volume(int n, const KArgs& a, const ExactTables& t, int accumulate, hipStream_t s)
{
    dispatch(n, 0, a, t, ExactMassTables{}, nullptr, accumulate, 0, s);
}

void launch_exact_face(int n, const KArgs& a, const ExactTables& t, int interior, int boundary, hipStream_t s)
{
    dispatch(n, 1, a, t, ExactMassTables{}, nullptr, interior, boundary, s);
}

void launch_exact_mass(int n, const KArgs& a, const double* r, const ExactMassTables& m, int kid, hipStream_t s,
                       int nv)
{
    dispatch(n, 2, a, ExactTables{}, m, r, kid, kid < 0 ? nv : 3, s);
}

void launch_exact_stage(int n, int kid, const KArgs& a, const double* tab, hipStream_t s)
{
    if (a.nel <= 0) return;
    switch (n) {
    case 2: launch_stage<2>(kid, a, tab, s); break;
    case 3: launch_stage<3>(kid, a, tab, s); break;
    case 4: launch_stage<4>(kid, a, tab, s); break;
    case 5: launch_stage<5>(kid, a, tab, s); break;
    case 6: launch_stage<6>(kid, a, tab, s); break;
    case 7: launch_stage<7>(kid, a, tab, s); break;
    case 8: launch_stage<8>(kid, a, tab, s); break;
    case 9: launch_stage<9>(kid, a, tab, s); break;
    default: break;
    }
}

}  // namespace wdg
