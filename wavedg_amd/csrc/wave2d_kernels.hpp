// Fused gfx950 kernels for the 2D wave DG operator (nodal Gauss-Lobatto,
// collocated quadrature).
//
// One thread owns one GLL node of one element; a workgroup owns EPB whole
// elements (EPB n^2 <= 256 threads), so every element's 3 n^2 state values
// arrive as one contiguous, coalesced block of the (3, n, n, n_elem) layout
// and sit in LDS for the tensor-product contractions.  Face coupling reads
// the n neighbour trace values of each side straight from the input state
// (the neighbour's block is in L2: element blocks of one XCD are contiguous,
// see xcd_block), computes the flux in closed form and lifts it into the
// owning node, so no trace buffer and no atomics are needed and every output
// value has exactly one writer.
//
// The same kernel body serves every entry point through the MODE bits:
//   action   (y += A x)                         WaveEquation::action
//   bc       (y += B x)                         WaveBC::action
//   rhs      (y = M^{-1}(A x + B x + s F))      the RK time derivative
//   stage    (du = a du + dt rhs; u' = u + b du [; acc += c u'])
//            one low-storage RK4 stage, with the WaveHoltz filter
//            accumulation folded into the last stage of a step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wdg {

struct KArgs {
    const void* u_in;       // state read by this element and its neighbours
    void* u_out;            // y (action/bc/rhs) or the next stage state
    void* du;               // RK register (stage modes)
    const double* corners;  // (2, 4, nel)
    const int32_t* nbr;     // (4, nel)
    const int32_t* info;    // (4, nel), see host_mesh.hpp
    const double* basis;    // x[N], w[N], D[N*N] (column-major)
    const void* force;      // load vector (rhs/stage with M_FORCE)
    double fs;              // forcing scale
    double* acc;            // WaveHoltz accumulator (M_ACC)
    double ac;              // accumulation weight
    double rk_a, rk_b, dt;  // stage coefficients
    double ha, hb;          // flux: f = w ds [ha nA (u0+u1) + hb |nA| (u0-u1)]
    int nel;
};

enum : int {
    M_VOL = 1,
    M_IF = 2,
    M_BF = 4,
    M_OUT_ACCUM = 0 << 4,
    M_OUT_RHS = 1 << 4,
    M_OUT_STAGE0 = 2 << 4,
    M_OUT_STAGE = 3 << 4,
    M_OUTMASK = 3 << 4,
    M_FORCE = 1 << 8,
    M_ACC = 1 << 9,
};

// Entry points (dispatch ids) and their mode bits.
enum KernelId : int {
    K_ACTION = 0,
    K_BC,
    K_RHS,
    K_RHS_F,
    K_STAGE0,
    K_STAGE0_F,
    K_STAGE,
    K_STAGE_F,
    K_STAGE_ACC,
    K_STAGE_F_ACC,
    K_COUNT
};

constexpr int kModeOf[K_COUNT] = {
    M_VOL | M_IF | M_OUT_ACCUM,
    M_BF | M_OUT_ACCUM,
    M_VOL | M_IF | M_BF | M_OUT_RHS,
    M_VOL | M_IF | M_BF | M_OUT_RHS | M_FORCE,
    M_VOL | M_IF | M_BF | M_OUT_STAGE0,
    M_VOL | M_IF | M_BF | M_OUT_STAGE0 | M_FORCE,
    M_VOL | M_IF | M_BF | M_OUT_STAGE,
    M_VOL | M_IF | M_BF | M_OUT_STAGE | M_FORCE,
    M_VOL | M_IF | M_BF | M_OUT_STAGE | M_ACC,
    M_VOL | M_IF | M_BF | M_OUT_STAGE | M_FORCE | M_ACC,
};

template <int N>
struct Shape {
    static constexpr int NN = N * N;
    static constexpr int EPB = (256 / NN) > 0 ? (256 / NN) : 1;  // elements per workgroup
    static constexpr int NT = EPB * NN;                          // threads per workgroup
};

// Workgroups are dealt to the 8 XCDs round-robin (b -> XCD b mod 8).  Give
// each XCD one contiguous run of elements so that an element's neighbours
// (+-1, +-row in structured meshes) were loaded through the same L2.
__device__ __forceinline__ int xcd_block(int b, int nblk)
{
    const int x = b & 7, q = nblk >> 3, r = nblk & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// Node (k along side s, counted from the side's first corner) -> flat node.
template <int N>
__device__ __forceinline__ int side_node(int s, int k)
{
    const int i = (s == 0 || s == 2) ? k : (s == 1 ? N - 1 : 0);
    const int j = (s == 1 || s == 3) ? k : (s == 2 ? N - 1 : 0);
    return i + N * j;
}

template <int N, typename R, int MODE>
__global__ __launch_bounds__(Shape<N>::NT) void wave_kernel(const KArgs a)
{
    constexpr int NN = N * N;
    constexpr int EPB = Shape<N>::EPB;
    constexpr int NT = Shape<N>::NT;
    constexpr int SZ = EPB * 3 * NN;
    constexpr bool VOL = (MODE & M_VOL) != 0;
    constexpr bool IFACE = (MODE & M_IF) != 0;
    constexpr bool BFACE = (MODE & M_BF) != 0;
    constexpr int OUT = MODE & M_OUTMASK;
    constexpr bool FORCE = (MODE & M_FORCE) != 0;
    constexpr bool ACC = (MODE & M_ACC) != 0;

    __shared__ R s_u[SZ];
    __shared__ R s_F[VOL ? SZ : 1];
    __shared__ R s_G[VOL ? SZ : 1];
    __shared__ R s_D[NN];
    __shared__ double s_x[N], s_w[N];

    const R* __restrict__ u_in = static_cast<const R*>(a.u_in);
    const int tid = threadIdx.x;
    const int e0 = xcd_block(blockIdx.x, gridDim.x) * EPB;
    const int nloc = min(EPB, a.nel - e0);

    {
        const size_t gbase = (size_t)e0 * 3 * NN;
        const int cnt = nloc * 3 * NN;
        for (int k = tid; k < cnt; k += NT) s_u[k] = u_in[gbase + k];
        for (int k = tid; k < NN; k += NT) s_D[k] = (R)a.basis[2 * N + k];
        if (tid < N) {
            s_x[tid] = a.basis[tid];
            s_w[tid] = a.basis[N + tid];
        }
    }
    __syncthreads();

    const int el = tid / NN;
    const int node = tid - el * NN;
    const int i = node % N, j = node / N;
    const bool active = el < nloc;
    const int e = e0 + el;

    double c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) c[k] = active ? a.corners[8 * (size_t)e + k] : 0.0;

    // Bilinear map Jacobian at (x_i, x_j): [dx/dxi, dy/dxi, dx/deta, dy/deta] (Element.cpp:21-27).
    const double xi = s_x[i], eta = s_x[j];
    const double J0 = 0.25 * ((1.0 - eta) * (c[2] - c[0]) + (1.0 + eta) * (c[4] - c[6]));
    const double J1 = 0.25 * ((1.0 - eta) * (c[3] - c[1]) + (1.0 + eta) * (c[5] - c[7]));
    const double J2 = 0.25 * ((1.0 - xi) * (c[6] - c[0]) + (1.0 + xi) * (c[4] - c[2]));
    const double J3 = 0.25 * ((1.0 - xi) * (c[7] - c[1]) + (1.0 + xi) * (c[5] - c[3]));

    R p = 0, uu = 0, vv = 0;
    if (active) {
        p = s_u[el * 3 * NN + 3 * node + 0];
        uu = s_u[el * 3 * NN + 3 * node + 1];
        vv = s_u[el * 3 * NN + 3 * node + 2];
    }
    R r0 = 0, r1 = 0, r2 = 0;

    if constexpr (VOL) {
        // Weak-form volume term (Div.cpp:5-53, 131-179): contravariant fluxes
        // F = W (y_eta f - x_eta g), G = W (-y_xi f + x_xi g) with f = (u, p, 0),
        // g = (v, 0, p), then du(i,j) += sum_l D(l,i) F(l,j) + D(l,j) G(i,l).
        const double Wt = s_w[j] * s_w[i];
        const R Ye = (R)(J3 * Wt), Xe = (R)(J2 * Wt), Yx = (R)(J1 * Wt), Xx = (R)(J0 * Wt);
        R* F = s_F + el * 3 * NN;
        R* G = s_G + el * 3 * NN;
        F[0 * NN + j * N + i] = Ye * uu - Xe * vv;
        F[1 * NN + j * N + i] = Ye * p;
        F[2 * NN + j * N + i] = -(Xe * p);
        G[0 * NN + i * N + j] = Xx * vv - Yx * uu;
        G[1 * NN + i * N + j] = -(Yx * p);
        G[2 * NN + i * N + j] = Xx * p;
        __syncthreads();
        const R* Fr = F + j * N;
        const R* Gc = G + i * N;
#pragma unroll
        for (int l = 0; l < N; ++l) {
            const R di = s_D[l + N * i], dj = s_D[l + N * j];
            r0 += di * Fr[l] + dj * Gc[l];
            r1 += di * Fr[NN + l] + dj * Gc[NN + l];
            r2 += di * Fr[2 * NN + l] + dj * Gc[2 * NN + l];
        }
    }

    if constexpr (IFACE || BFACE) {
        if (active && (i == 0 || i == N - 1 || j == 0 || j == N - 1)) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const bool on = (s == 0) ? (j == 0) : (s == 1) ? (i == N - 1) : (s == 2) ? (j == N - 1) : (i == 0);
                if (!on) continue;
                const int nb = a.nbr[4 * (size_t)e + s];
                if (nb >= 0 ? !IFACE : !BFACE) continue;
                const int inf = a.info[4 * (size_t)e + s];
                const int k = (s == 0 || s == 2) ? i : j;
                const int first = (inf >> 3) & 1, flip = (inf >> 2) & 1, s0 = (inf >> 4) & 3;
                // Side corners in traversal order (Mesh2D.cpp:10-11); the edge
                // normal and length are taken from el0's traversal exactly as
                // StraightEdge does (Edge.hpp:107-122), so both elements of an
                // edge see bitwise identical n and ds.
                const int va = (s == 0 || s == 3) ? 0 : (s == 1 ? 1 : 3);
                const int vb = (s == 0) ? 1 : (s == 3 ? 3 : 2);
                double dx = c[2 * vb] - c[2 * va], dy = c[2 * vb + 1] - c[2 * va + 1];
                if (!first && flip) {
                    dx = -dx;
                    dy = -dy;
                }
                const double len = hypot(dx, dy);
                const double sg = (s0 >= 2) ? -1.0 : 1.0;
                const double nx = sg * dy / len, ny = -sg * dx / len;
                const int kk = flip ? (N - 1 - k) : k;  // matching node on the other side
                const int ie = first ? k : kk;          // node index in el0's order
                const R wds = (R)(s_w[ie] * (len / 2));
                const R n0 = (R)nx, n1 = (R)ny;

                R pB, uB, vB;
                if (nb >= 0) {
                    const R* q = u_in + ((size_t)nb * NN + side_node<N>(inf & 3, kk)) * 3;
                    pB = q[0];
                    uB = q[1];
                    vB = q[2];
                } else {
                    const int code = (inf >> 8) & 0xff;
                    if (code == 1) {  // reflecting (WaveEquation.hpp:64-69)
                        pB = p;
                        uB = (n1 * n1 - n0 * n0) * uu - (R)2 * n0 * n1 * vv;
                        vB = -(R)2 * n0 * n1 * uu + (n0 * n0 - n1 * n1) * vv;
                    } else if (code == 0) {  // characteristic outflow (WaveEquation.hpp:77-84, p_ext after u_ext)
                        const R out = p + n0 * uu + n1 * vv;
                        const R tang = n1 * uu - n0 * vv;
                        uB = n1 * tang + (R)0.5 * n0 * out;
                        vB = -n0 * tang + (R)0.5 * n1 * out;
                        pB = n0 * uB + n1 * vB;
                    } else {  // Dirichlet p = 0 (mirror state)
                        pB = -p;
                        uB = uu;
                        vB = vv;
                    }
                }
                // (u0, u1) in the edge's orientation: u0 belongs to el0.
                const R p0 = first ? p : pB, p1 = first ? pB : p;
                const R u0 = first ? uu : uB, u1 = first ? uB : uu;
                const R v0 = first ? vv : vB, v1 = first ? vB : vv;
                // Closed form of EdgeFlux<true> for the wave equation (EdgeFlux.cpp:31-139,
                // 289-325): with C = n.A, |C| = C^2 for unit n, so
                //   f_p = w ds [ha n.(u0+u1) + hb (p0-p1)]
                //   f_u = w ds n [ha (p0+p1) + hb n.(u0-u1)]
                const R ha = (R)a.ha, hb = (R)a.hb;
                const R fp = wds * (ha * (n0 * (u0 + u1) + n1 * (v0 + v1)) + hb * (p0 - p1));
                const R fc = wds * (ha * (p0 + p1) + hb * (n0 * (u0 - u1) + n1 * (v0 - v1)));
                if (first) {
                    r0 += fp;
                    r1 += n0 * fc;
                    r2 += n1 * fc;
                } else {
                    r0 -= fp;
                    r1 -= n0 * fc;
                    r2 -= n1 * fc;
                }
            }
        }
    }

    if (!active) return;
    const size_t g = ((size_t)e * NN + node) * 3;
    if constexpr (OUT == M_OUT_ACCUM) {
        R* y = static_cast<R*>(a.u_out);
        y[g + 0] += r0;
        y[g + 1] += r1;
        y[g + 2] += r2;
    } else {
        // MassMatrix<true>: m = w_i w_j det J (MassMatrix.cpp:6-28).
        const double m = s_w[i] * s_w[j] * (J0 * J3 - J1 * J2);
        if constexpr (FORCE) {
            const R* Fv = static_cast<const R*>(a.force);
            const R fs = (R)a.fs;
            r0 += Fv[g + 0] * fs;
            r1 += Fv[g + 1] * fs;
            r2 += Fv[g + 2] * fs;
        }
        const R im = (R)(1.0 / m);
        r0 *= im;
        r1 *= im;
        r2 *= im;
        if constexpr (OUT == M_OUT_RHS) {
            R* y = static_cast<R*>(a.u_out);
            y[g + 0] = r0;
            y[g + 1] = r1;
            y[g + 2] = r2;
        } else {
            // Low-storage RK stage (ode.hpp:74-90).
            R* du = static_cast<R*>(a.du);
            R* uo = static_cast<R*>(a.u_out);
            const R dt = (R)a.dt, rb = (R)a.rk_b;
            R d0, d1, d2;
            if constexpr (OUT == M_OUT_STAGE0) {
                d0 = dt * r0;
                d1 = dt * r1;
                d2 = dt * r2;
            } else {
                const R ra = (R)a.rk_a;
                d0 = ra * du[g + 0] + dt * r0;
                d1 = ra * du[g + 1] + dt * r1;
                d2 = ra * du[g + 2] + dt * r2;
            }
            du[g + 0] = d0;
            du[g + 1] = d1;
            du[g + 2] = d2;
            const R w0 = p + rb * d0, w1 = uu + rb * d1, w2 = vv + rb * d2;
            uo[g + 0] = w0;
            uo[g + 1] = w1;
            uo[g + 2] = w2;
            if constexpr (ACC) {
                // WaveHoltz filter accumulation in double (WaveHoltz.cpp:83-87).
                const double ac = a.ac;
                a.acc[g + 0] += ac * (double)w0;
                a.acc[g + 1] += ac * (double)w1;
                a.acc[g + 2] += ac * (double)w2;
            }
        }
    }
}

// Load vector of a Gaussian-sum forcing with Gauss-Legendre(N) quadrature
// (LinearFunctional.hpp:136-202).  tab = xq[N], wq[N], B[N*N] with
// B(l, j) = phi_j(xq_l) column-major.
template <int N>
__global__ __launch_bounds__(Shape<N>::NT) void load_vector_kernel(const double* corners, const double* tab,
                                                                   const double* src, int nsrc, int var,
                                                                   double* F, int nel)
{
    constexpr int NN = N * N;
    constexpr int EPB = Shape<N>::EPB;
    __shared__ double s_q[EPB * NN], s_p[EPB * NN];
    const int tid = threadIdx.x;
    const int el = tid / NN, node = tid - el * NN;
    const int e = blockIdx.x * EPB + el;
    const bool active = e < nel;
    const int k = node % N, l = node / N;  // quad point (k, l)
    double fq = 0.0;
    if (active) {
        const double* c = corners + 8 * (size_t)e;
        const double xi = tab[k], eta = tab[l];
        const double b0 = 0.25 * (1.0 - xi) * (1.0 - eta), b1 = 0.25 * (1.0 + xi) * (1.0 - eta);
        const double b2 = 0.25 * (1.0 + xi) * (1.0 + eta), b3 = 0.25 * (1.0 - xi) * (1.0 + eta);
        const double X = c[0] * b0 + c[2] * b1 + c[4] * b2 + c[6] * b3;
        const double Y = c[1] * b0 + c[3] * b1 + c[5] * b2 + c[7] * b3;
        double f = 0.0;
        for (int s = 0; s < nsrc; ++s) {
            const double dx = X - src[4 * s], dy = Y - src[4 * s + 1], sg = src[4 * s + 2];
            f += src[4 * s + 3] * exp(-(dx * dx + dy * dy) / (2.0 * sg * sg));
        }
        const double J0 = 0.25 * ((1.0 - eta) * (c[2] - c[0]) + (1.0 + eta) * (c[4] - c[6]));
        const double J1 = 0.25 * ((1.0 - eta) * (c[3] - c[1]) + (1.0 + eta) * (c[5] - c[7]));
        const double J2 = 0.25 * ((1.0 - xi) * (c[6] - c[0]) + (1.0 + xi) * (c[4] - c[2]));
        const double J3 = 0.25 * ((1.0 - xi) * (c[7] - c[1]) + (1.0 + xi) * (c[5] - c[3]));
        fq = f * ((J0 * J3 - J1 * J2) * tab[N + k] * tab[N + l]);
    }
    s_q[el * NN + k + N * l] = fq;
    __syncthreads();
    {
        // ProjF(k, j) = sum_l Fq(k, l) B(l, j); this thread: (k, j = l)
        const int jj = l;
        double s = 0.0;
        for (int m = 0; m < N; ++m) s += s_q[el * NN + k + N * m] * tab[2 * N + m + N * jj];
        s_p[el * NN + k + N * jj] = s;
    }
    __syncthreads();
    if (!active) return;
    {
        // F(i, j) = sum_k ProjF(k, j) B(k, i); this thread: (i = k, j = l)
        const int ii = k, jj = l;
        double s = 0.0;
        for (int m = 0; m < N; ++m) s += s_p[el * NN + m + N * jj] * tab[2 * N + m + N * ii];
        double* o = F + 3 * ((size_t)e * NN + ii + N * jj);
        o[0] = var == 0 ? s : 0.0;
        o[1] = var == 1 ? s : 0.0;
        o[2] = var == 2 ? s : 0.0;
    }
}

// Host-side launchers, one translation unit per N (wave2d_kernels_n.hip).
#define WDG_DECLARE_LAUNCH(NV)                                                                              \
    void launch_wave_n##NV(int kid, int f32, const KArgs& a, hipStream_t s);                              \
    void launch_load_vector_n##NV(const double* corners, const double* tab, const double* src, int nsrc, \
                                  int var, double* F, int nel, hipStream_t s);
WDG_DECLARE_LAUNCH(2)
WDG_DECLARE_LAUNCH(3)
WDG_DECLARE_LAUNCH(4)
WDG_DECLARE_LAUNCH(5)
WDG_DECLARE_LAUNCH(6)
WDG_DECLARE_LAUNCH(7)
WDG_DECLARE_LAUNCH(8)
WDG_DECLARE_LAUNCH(9)
#undef WDG_DECLARE_LAUNCH

}  // namespace wdg
