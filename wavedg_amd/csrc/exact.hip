// Exact-quadrature operators (approx_quad = false) in FP64: Div<false>
// (Div.cpp:312-360, 469-588), EdgeFlux<false> on Lobatto face traces
// (EdgeFlux.cpp:375-420, 491-620) and MassMatrix<false> (MassMatrix.cpp:
// 85-133, 193-227).  One element per workgroup throughout; tables are
// column-major (n_quad, n) with P(q, k) = phi_k(x_q) and D(q, k) = phi_k'(x_q).
//
//  * volume: u -> quadrature points by the two tensor-product passes of
//    div_2d_q, contravariant fluxes there, and back with D and P;
//  * faces: for each side the element's own trace and the other trace
//    (neighbour nodes, or the boundary state built at the Lobatto nodes) are
//    interpolated to the edge rule, the closed-form flux
//      c_p = w ds [ha n.(uA+uB) + hb (pA-pB)],  c_u = n w ds [ha (pA+pB) + hb n.(uA-uB)]
//    is evaluated there (own outward normal: +f on el0, -f on el1 of the
//    reference's edge flux) and projected onto the side's nodes; corner
//    nodes add their two sides in side order, so every output has one writer;
//  * mass matrix: a setup kernel integrates M = (phi_i phi_j det J) with the
//    mass rule and factors it (Cholesky, linalg.cpp:5-30) in LDS; the apply
//    kernel solves or multiplies with the factor (linalg.cpp:32-98) and
//    finishes the RK stage or right-hand side in place.
#include <hip/hip_runtime.h>

#include "wave2d_kernels.hpp"

namespace wdg {
namespace {

constexpr int kQMax = 10;  // points of any rule
constexpr int kBlk = 128;  // threads per element workgroup

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
    const double2* gp = reinterpret_cast<const double2*>(a.geom + 6 * (size_t)e);
    const double2 g_a = gp[0], g_c = gp[1], g_h = gp[2];
    bool any = false;
    for (int s = 0; s < 4; ++s) {
        const int nb = a.nbr[4 * (size_t)e + s], inf = a.info[4 * (size_t)e + s];
        const bool is_int = nb >= 0;
        if (is_int ? !interior : !boundary) continue;  // uniform over the workgroup
        any = true;
        // own outward unit normal and ds = |side| / 2 (straight side of the bilinear map)
        const double sh = (s == 0 || s == 3) ? -1.0 : 1.0;
        const double dx = 2.0 * ((s & 1) ? g_c.x + sh * g_h.x : g_a.x + sh * g_h.x);
        const double dy = 2.0 * ((s & 1) ? g_c.y + sh * g_h.y : g_a.y + sh * g_h.y);
        const double l2 = dx * dx + dy * dy, il = rsqrt_nr(l2);
        const double sil = (s >= 2) ? -il : il;
        const double n0 = dy * sil, n1 = -dx * sil, ds = 0.5 * (l2 * il);
        __syncthreads();  // the previous side's traces and fluxes are consumed
        if (tid < N) {
            const int k = tid;
            const double* qa = u + ((size_t)e * NN + side_node<N>(s, k)) * 3;
            const double pA = qa[0], uA = qa[1], vA = qa[2];
            double pB, uB, vB;
            if (is_int) {
                const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
                const double* qb = u + ((size_t)nb * NN + side_node<N>(inf & 3, kk)) * 3;
                pB = qb[0];
                uB = qb[1];
                vB = qb[2];
            } else {
                const int code = (inf >> 8) & 0xff;
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
            const int k = tid, node = side_node<N>(s, k);
            for (int d = 0; d < 3; ++d) {
                double v = 0.0;
                for (int q = 0; q < nq; ++q) v += s_c[d * kQMax + q] * s_P[q + nq * k];
                s_out[d + 3 * node] += v;
            }
        }
    }
    if (!any) return;  // uniform
    __syncthreads();
    double* y = static_cast<double*>(a.u_out) + (size_t)e * 3 * NN;
    for (int k = tid; k < 3 * NN; k += 64) y[k] += s_out[k];
}

// M(i, j) = sum over the mass rule of phi_i phi_j w det J (lower triangle) and
// its Cholesky factor, written to L (c2d, c2d) column-major, upper part zero.
template <int N>
__global__ __launch_bounds__(256) void exact_mass_setup_kernel(const double* geom, const ExactMassTables m, double* L)
{
    constexpr int C = N * N;
    __shared__ double s_M[C * C];
    __shared__ double s_B[kQMax * N], s_x[kQMax], s_w[kQMax], s_detJ[kQMax * kQMax];
    const int e = blockIdx.x, tid = threadIdx.x, nq = m.nq;
    for (int k = tid; k < nq * N; k += 256) s_B[k] = m.B[k];
    if (tid < nq) {
        s_x[tid] = m.x[tid];
        s_w[tid] = m.w[tid];
    }
    __syncthreads();
    const double* g = geom + 6 * (size_t)e;
    for (int p = tid; p < nq * nq; p += 256) {
        const double xi = s_x[p % nq], eta = s_x[p / nq];
        const double J0 = g[0] + g[4] * eta, J1 = g[1] + g[5] * eta, J2 = g[2] + g[4] * xi, J3 = g[3] + g[5] * xi;
        s_detJ[p] = J0 * J3 - J1 * J2;
    }
    __syncthreads();
    for (int k = tid; k < C * C; k += 256) {
        const int i = k % C, j = k / C;
        double mij = 0.0;
        if (i >= j)
            for (int p = 0; p < nq * nq; ++p) {
                const double bi = s_B[p % nq + nq * (i % N)] * s_B[p / nq + nq * (i / N)];
                const double bj = s_B[p % nq + nq * (j % N)] * s_B[p / nq + nq * (j / N)];
                mij += bi * bj * (s_detJ[p] * s_w[p % nq] * s_w[p / nq]);
            }
        s_M[k] = mij;
    }
    __syncthreads();
    for (int k = 0; k < C; ++k) {
        const double akk = s_M[k + C * k];
        const int rem = C - k - 1;
        for (int q = tid; q < rem * rem; q += 256) {
            const int j = k + 1 + q / rem, i = k + 1 + q % rem;
            if (i >= j) s_M[i + C * j] -= (s_M[j + C * k] / akk) * s_M[i + C * k];
        }
        __syncthreads();
        const double sc = 1.0 / sqrt(akk);
        for (int i = k + tid; i < C; i += 256) s_M[i + C * k] *= sc;
        __syncthreads();
    }
    double* out = L + (size_t)e * C * C;
    for (int k = tid; k < C * C; k += 256) out[k] = s_M[k];
}

// Mass-matrix application per element with the factor L:
//   kid < 0: -1 y = M x, -2 y = M^{-1} x (x = a.u_in, y = a.u_out; may alias);
//   kid >= 0: the RHS / RK-stage epilogue of wave_kernel for kModeOf[kid],
//             with the operator's output r (before M^{-1}) in `r`.
template <int N>
__global__ __launch_bounds__(kBlk) void exact_mass_kernel(const KArgs a, const double* r, const double* L, int kid)
{
    constexpr int C = N * N;
    __shared__ double s_x[3 * C], s_t[3 * C];
    const int e = blockIdx.x, tid = threadIdx.x;
    const double* Le = L + (size_t)e * C * C;
    const size_t base = (size_t)e * 3 * C;
    const int mode = kid >= 0 ? kModeOf[kid] : 0;
    const bool force = kid >= 0 && (mode & M_FORCE);
    for (int k = tid; k < 3 * C; k += kBlk) {
        double v = kid < 0 ? static_cast<const double*>(a.u_in)[base + k] : r[base + k];
        if (force) v += static_cast<const double*>(a.force)[base + k] * a.fs;
        s_x[k] = v;
    }
    __syncthreads();
    if (kid == -1) {
        // y = L (L' x)   (mult_chol, linalg.cpp:68-98)
        for (int k = tid; k < 3 * C; k += kBlk) {
            const int d = k % 3, i = k / 3;
            double y = 0.0;
            for (int j = i; j < C; ++j) y += Le[j + C * i] * s_x[d + 3 * j];
            s_t[k] = y;
        }
        __syncthreads();
        for (int k = tid; k < 3 * C; k += kBlk) {
            const int d = k % 3, i = k / 3;
            double y = 0.0;
            for (int j = 0; j <= i; ++j) y += Le[i + C * j] * s_t[d + 3 * j];
            s_x[k] = y;
        }
        __syncthreads();
    } else {
        // x = (L L')^{-1} x   (solve_chol, linalg.cpp:32-66)
        for (int j = 0; j < C; ++j) {
            if (tid < 3) s_x[tid + 3 * j] /= Le[j + C * j];
            __syncthreads();
            for (int k = tid; k < 3 * (C - j - 1); k += kBlk) {
                const int d = k % 3, i = j + 1 + k / 3;
                s_x[d + 3 * i] -= Le[i + C * j] * s_x[d + 3 * j];
            }
            __syncthreads();
        }
        for (int j = C - 1; j >= 0; --j) {
            if (tid < 3) s_x[tid + 3 * j] /= Le[j + C * j];
            __syncthreads();
            for (int k = tid; k < 3 * j; k += kBlk) {
                const int d = k % 3, i = k / 3;
                s_x[d + 3 * i] -= Le[j + C * i] * s_x[d + 3 * j];
            }
            __syncthreads();
        }
    }
    const int out = mode & M_OUTMASK;
    for (int k = tid; k < 3 * C; k += kBlk) {
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

template <int N>
void launch_all(int what, const KArgs& a, const ExactTables& t, const ExactMassTables& m, const double* r, double* L,
                int p0, int p1, hipStream_t s)
{
    const dim3 grid(a.nel);
    switch (what) {
    case 0: hipLaunchKernelGGL(exact_volume_kernel<N>, grid, dim3(kBlk), 0, s, a, t, p0); break;
    case 1: hipLaunchKernelGGL(exact_face_kernel<N>, grid, dim3(64), 0, s, a, t, p0, p1); break;
    case 2: hipLaunchKernelGGL(exact_mass_setup_kernel<N>, grid, dim3(256), 0, s, a.geom, m, L); break;
    default: hipLaunchKernelGGL(exact_mass_kernel<N>, grid, dim3(kBlk), 0, s, a, r, L, p0); break;
    }
}

void dispatch(int n, int what, const KArgs& a, const ExactTables& t, const ExactMassTables& m, const double* r,
              double* L, int p0, int p1, hipStream_t s)
{
    if (a.nel <= 0) return;
    switch (n) {
    case 2: launch_all<2>(what, a, t, m, r, L, p0, p1, s); break;
    case 3: launch_all<3>(what, a, t, m, r, L, p0, p1, s); break;
    case 4: launch_all<4>(what, a, t, m, r, L, p0, p1, s); break;
    case 5: launch_all<5>(what, a, t, m, r, L, p0, p1, s); break;
    case 6: launch_all<6>(what, a, t, m, r, L, p0, p1, s); break;
    case 7: launch_all<7>(what, a, t, m, r, L, p0, p1, s); break;
    case 8: launch_all<8>(what, a, t, m, r, L, p0, p1, s); break;
    case 9: launch_all<9>(what, a, t, m, r, L, p0, p1, s); break;
    default: break;
    }
}

}  // namespace

void launch_exact_volume(int n, const KArgs& a, const ExactTables& t, int accumulate, hipStream_t s)
{
    dispatch(n, 0, a, t, ExactMassTables{}, nullptr, nullptr, accumulate, 0, s);
}

void launch_exact_face(int n, const KArgs& a, const ExactTables& t, int interior, int boundary, hipStream_t s)
{
    dispatch(n, 1, a, t, ExactMassTables{}, nullptr, nullptr, interior, boundary, s);
}

void launch_exact_mass_setup(int n, const KArgs& a, const ExactMassTables& m, double* L, hipStream_t s)
{
    dispatch(n, 2, a, ExactTables{}, m, nullptr, L, 0, 0, s);
}

void launch_exact_mass(int n, const KArgs& a, const double* r, const double* L, int kid, hipStream_t s)
{
    dispatch(n, 3, a, ExactTables{}, ExactMassTables{}, r, const_cast<double*>(L), kid, 0, s);
}

}  // namespace wdg
