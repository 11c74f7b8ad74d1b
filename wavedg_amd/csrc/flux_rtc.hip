// DivF2D<true> and EdgeFluxF2D<true> with flux functors given as source
// text and compiled for gfx950 at run time with hipRTC (DivF.hpp:207-274,
// EdgeFluxF.hpp:136-171).  The reference takes C++ functors; a C ABI cannot
// carry a functor to the GPU, so the caller passes the functor's body in the
// reference's signature as HIP source:
//
//   __device__ void flux(const double* x, const double* u, double* F);
//       F (n_var, 2) column-major: F[d] = f_d(x, u), F[n_var + d] = g_d(x, u)
//   __device__ void numerical_flux(const double* x, const double* n,
//                                  const double* uL, const double* uR, double* fh);
//   optionally, with #define WDG_EXTERIOR_STATE,
//   __device__ void exterior_state(const double* x, const double* n,
//                                  const double* u, double* u_ext);
//
// The library appends its two kernels (below) to that text and compiles each
// one on first use with the handle's order, variable count and basis as
// macros, so the functor inlines into the kernel.  The kernels follow the
// element-local layout of the other operators: one lane per node (volume) or
// per side node (faces, sides in order, every output with one writer).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "wave2d_kernels.hpp"

namespace wdg {
namespace {

const char* const kRtcKernels = R"WDGRTC(
namespace wdg_rtc {
constexpr int N = WDG_N, NN = N * N, NV = WDG_NVAR;

// Bilinear map of corners c (2, 4) at (xi, eta): X and
// J = (x_xi, y_xi, x_eta, y_eta) (Element.cpp:5-27).
__device__ inline void corner_map(const double* c, double xi, double eta, double* X, double* J)
{
    const double b0 = 0.25 * (1.0 - xi) * (1.0 - eta), b1 = 0.25 * (1.0 + xi) * (1.0 - eta);
    const double b2 = 0.25 * (1.0 + xi) * (1.0 + eta), b3 = 0.25 * (1.0 - xi) * (1.0 + eta);
    X[0] = c[0] * b0 + c[2] * b1 + c[4] * b2 + c[6] * b3;
    X[1] = c[1] * b0 + c[3] * b1 + c[5] * b2 + c[7] * b3;
    J[0] = 0.25 * ((1.0 - eta) * (c[2] - c[0]) + (1.0 + eta) * (c[4] - c[6]));
    J[1] = 0.25 * ((1.0 - eta) * (c[3] - c[1]) + (1.0 + eta) * (c[5] - c[7]));
    J[2] = 0.25 * ((1.0 - xi) * (c[6] - c[0]) + (1.0 + xi) * (c[4] - c[2]));
    J[3] = 0.25 * ((1.0 - xi) * (c[7] - c[1]) + (1.0 + xi) * (c[5] - c[3]));
}

__device__ inline int side_node(int s, int k)
{
    const int i = (s == 0 || s == 2) ? k : (s == 1 ? N - 1 : 0);
    const int j = (s == 1 || s == 3) ? k : (s == 2 ? N - 1 : 0);
    return i + N * j;
}

__device__ inline int inward_stride(int s) { return s == 0 ? N : s == 1 ? -1 : s == 2 ? -N : 1; }

// Values of the NV variables on element side s at side node k of element e:
// the node's values (Gauss-Lobatto), or sum_m P(m) u(m rows in) (Gauss-Legendre).
__device__ inline void side_trace(const double* u, int e, int s, int k, const double* P, double* out)
{
    const int node = side_node(s, k);
    const double* q = u + (size_t)NV * ((size_t)e * NN + node);
    if (WDG_LEG) {
        const int st = NV * inward_stride(s);
        for (int d = 0; d < NV; ++d) {
            double t = 0.0;
            for (int m = 0; m < N; ++m) t += q[d + m * st] * P[m];
            out[d] = t;
        }
    } else {
        for (int d = 0; d < NV; ++d) out[d] = q[d];
    }
}
}  // namespace wdg_rtc

#ifdef WDG_WANT_DIVF
// DivF2D<true>::action (DivF.hpp:207-274): divF += sum_l D(l,i) F~(l,j) + D(l,j) G~(i,l)
// with the contravariant fluxes F~ = w_i w_j (y_eta f - x_eta g),
// G~ = w_i w_j (-y_xi f + x_xi g) of flux(x, u) at every node.  E elements
// of N^2 lanes per workgroup.
extern "C" __global__ void __launch_bounds__(256) wdg_divf_rtc(const double* __restrict__ u, double* __restrict__ out,
                                                                const double* __restrict__ corners,
                                                                const double* __restrict__ basis, int nel)
{
    using namespace wdg_rtc;
    constexpr int E = 256 / NN;
    __shared__ double sF[E * NV * NN], sG[E * NV * NN];
    __shared__ double sB[2 * N + NN];  // x, w, D of the basis
    const int tid = threadIdx.x, el = tid / NN, node = tid - el * NN, i = node % N, j = node / N;
    const int e = blockIdx.x * E + el;
    const bool active = el < E && e < nel;
    for (int t = tid; t < 2 * N + NN; t += E * NN) sB[t] = basis[t];
    __syncthreads();
    const double* xb = sB;
    const double* wb = sB + N;
    const double* D = sB + 2 * N;
    if (active) {
        double X[2], J[4];
        corner_map(corners + 8 * (size_t)e, xb[i], xb[j], X, J);
        const double wij = wb[i] * wb[j];
        const double Ye = J[3] * wij, Xe = J[2] * wij, Yx = J[1] * wij, Xx = J[0] * wij;
        double uq[NV], Fq[2 * NV];
        for (int d = 0; d < NV; ++d) uq[d] = u[(size_t)NV * ((size_t)e * NN + node) + d];
        for (int d = 0; d < 2 * NV; ++d) Fq[d] = 0.0;
        flux(X, uq, Fq);
        for (int d = 0; d < NV; ++d) {
            sF[(el * NV + d) * NN + j * N + i] = Ye * Fq[d] - Xe * Fq[NV + d];
            sG[(el * NV + d) * NN + i * N + j] = -Yx * Fq[d] + Xx * Fq[NV + d];
        }
    }
    __syncthreads();
    if (!active) return;
    for (int d = 0; d < NV; ++d) {
        double r = 0.0;
        for (int l = 0; l < N; ++l)
            r += D[l + N * i] * sF[(el * NV + d) * NN + j * N + l] + D[l + N * j] * sG[(el * NV + d) * NN + i * N + l];
        out[(size_t)NV * ((size_t)e * NN + node) + d] += r;
    }
}
#endif

#ifdef WDG_WANT_EDGE
// EdgeFluxF2D<true>::action (EdgeFluxF.hpp:136-171) between the face
// prolongation and its transpose, element by element: at side node k of
// side s, with el0's normal n and uL = el0's trace, fh = numerical_flux(x, n,
// uL, uR), and this element receives scale * (+fh w ds) if it is el0 and
// scale * (-fh w ds) if el1, lifted into the side's node (or, for a
// Gauss-Legendre basis, along its line with weights P).  Boundary sides take
// the exterior state from exterior_state(), or a copy of the interior trace.
// One lane per (element, side, side node), all sides at once, E = 64 / 4N
// elements per workgroup; the nodes then add their sides in side order.
extern "C" __global__ void __launch_bounds__(64) wdg_edgeflux_rtc(const double* __restrict__ u,
                                                                   double* __restrict__ out,
                                                                   const double* __restrict__ corners,
                                                                   const int* __restrict__ nbr,
                                                                   const int* __restrict__ info,
                                                                   const double* __restrict__ basis, int nel,
                                                                   int interior, int boundary, double scale)
{
    using namespace wdg_rtc;
    constexpr int NF = 4 * N, E = 64 / NF > 0 ? 64 / NF : 1;
    __shared__ double s_f[E * NF * NV];  // each side node's contribution, this element's order
    __shared__ int s_any[E];             // the element has a side in this launch
    const int tid = threadIdx.x, e0 = blockIdx.x * E;
    const int nloc = nel - e0 < E ? nel - e0 : E;
    if (tid < E) s_any[tid] = 0;
    {
        bool act = false;
        for (int t = tid; t < 4 * nloc; t += 64) {
            const int nb = nbr[4 * ((size_t)e0 + t / 4) + (t % 4)];
            act = act || (nb >= 0 ? interior != 0 : boundary != 0);
        }
        if (!__syncthreads_or(act)) return;
    }
    const double* xb = basis;
    const double* wb = basis + N;
    const double* P = basis + 2 * N + NN;
    const int el = tid / NF, f = tid - el * NF, s = f / N, k = f - (f / N) * N, e = e0 + el;
    double c[NV];
    for (int d = 0; d < NV; ++d) c[d] = 0.0;
    if (tid < E * NF && el < nloc) {
        const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
        const bool is_int = nb >= 0;
        if (is_int ? interior : boundary) {
            s_any[el] = 1;
            const double* cr = corners + 8 * (size_t)e;
            // side vector in traversal order (sides 2, 3 run 3->2, 0->3)
            const int a0 = s == 0 ? 0 : s == 1 ? 1 : s == 2 ? 3 : 0;
            const int a1 = s == 0 ? 1 : s == 1 ? 2 : s == 2 ? 2 : 3;
            const double dx = cr[2 * a1] - cr[2 * a0], dy = cr[2 * a1 + 1] - cr[2 * a0 + 1];
            const double len = sqrt(dx * dx + dy * dy), sg = s >= 2 ? -1.0 : 1.0;
            double nrm[2] = {sg * dy / len, -sg * dx / len};  // this element's outward normal
            const double ds = 0.5 * len;
            const double xi = s == 0 || s == 2 ? xb[k] : (s == 1 ? 1.0 : -1.0);
            const double eta = s == 1 || s == 3 ? xb[k] : (s == 2 ? 1.0 : -1.0);
            double X[2], J[4];
            corner_map(cr, xi, eta, X, J);
            double uA[NV], uB[NV], fh[NV];
            side_trace(u, e, s, k, P, uA);
            if (is_int) {
                const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
                side_trace(u, nb, inf & 3, kk, P, uB);
            } else {
#ifdef WDG_EXTERIOR_STATE
                exterior_state(X, nrm, uA, uB);
#else
                for (int d = 0; d < NV; ++d) uB[d] = uA[d];
#endif
            }
            const bool first = (inf >> 3) & 1;
            const int ie = (!first && ((inf >> 2) & 1)) ? (N - 1 - k) : k;  // el0's node order
            for (int d = 0; d < NV; ++d) fh[d] = 0.0;
            if (first) {
                numerical_flux(X, nrm, uA, uB, fh);
            } else {
                nrm[0] = -nrm[0];
                nrm[1] = -nrm[1];
                numerical_flux(X, nrm, uB, uA, fh);
            }
            const double sgn = first ? scale : -scale;
            for (int d = 0; d < NV; ++d) c[d] = sgn * (fh[d] * wb[ie] * ds);
        }
    }
    if (tid < E * NF) {
        for (int d = 0; d < NV; ++d) s_f[(el * NF + f) * NV + d] = c[d];
    }
    __syncthreads();
    double* yb = out + (size_t)e0 * NV * NN;
    for (int t = tid; t < nloc * NV * NN; t += 64) {
        const int el2 = t / (NV * NN), r = t - el2 * NV * NN, node = r / NV, d = r - node * NV;
        if (!s_any[el2]) continue;
        const int i = node % N, j = node / N;
        const double* F = s_f + el2 * NF * NV + d;
        double add;
        if (WDG_LEG) {
            add = P[j] * F[(0 * N + i) * NV] + P[N - 1 - i] * F[(1 * N + j) * NV] +
                  P[N - 1 - j] * F[(2 * N + i) * NV] + P[i] * F[(3 * N + j) * NV];
        } else {
            if (!(i == 0 || j == 0 || i == N - 1 || j == N - 1)) continue;  // no side reaches an inner node
            add = 0.0;
            if (j == 0) add += F[(0 * N + i) * NV];
            if (i == N - 1) add += F[(1 * N + j) * NV];
            if (j == N - 1) add += F[(2 * N + i) * NV];
            if (i == 0) add += F[(3 * N + j) * NV];
        }
        yb[t] += add;
    }
}
#endif

#ifdef WDG_WANT_DIVF_Q
// DivF2D<false>::action (DivF.hpp:275-392): u interpolated to the NQ x NQ
// Gauss-Legendre points, the contravariant fluxes of flux(x, u) there with
// the metric of the bilinear map, then divF(d, k, l) += sum_j P(j, l) Df(j, d, k)
// + D(j, l) Pg(j, d, k) with Df(j, d, k) = sum_i D(i, k) F~(i, d, j) and
// Pg(j, d, k) = sum_i P(i, k) G~(i, d, j).  EQ elements per 256-lane
// workgroup (as many as the largest pass fills); tab = xq[NQ], wq[NQ],
// P[NQ * N], D[NQ * N] (P(i, k) at i + NQ k), kept in LDS.
namespace wdg_rtc {
constexpr int NQ = WDG_NQ;
constexpr int SPU = N * NV * NQ, SDP = 2 * NQ * NV * N, SW = SPU > SDP ? SPU : SDP, SF = 2 * NQ * NV * NQ;
constexpr int WMAX = (NQ * NQ > N * NQ ? NQ * NQ : N * NQ) * NV > NN * NV ? (NQ * NQ > N * NQ ? NQ * NQ : N * NQ) * NV
                                                                          : NN * NV;
constexpr int EQ = 256 / WMAX > 0 ? 256 / WMAX : 1;
}  // namespace wdg_rtc
extern "C" __global__ void __launch_bounds__(256) wdg_divf_q_rtc(const double* __restrict__ u,
                                                                  double* __restrict__ out,
                                                                  const double* __restrict__ corners,
                                                                  const double* __restrict__ tab, int nel)
{
    using namespace wdg_rtc;
    __shared__ double sM[EQ * (SW + SF)];  // per element: Pu, then Df and Pg; F
    __shared__ double sT[2 * NQ + 2 * NQ * N];
    const int tid = threadIdx.x, e0 = blockIdx.x * EQ;
    const int nloc = nel - e0 < EQ ? nel - e0 : EQ;
    for (int t = tid; t < 2 * NQ + 2 * NQ * N; t += 256) sT[t] = tab[t];
    __syncthreads();
    const double* xq = sT;
    const double* wq = sT + NQ;
    const double* P = sT + 2 * NQ;
    const double* D = sT + 2 * NQ + NQ * N;
    // Pu(l, d, i) = sum_k P(i, k) u(d, k, l)
    for (int t = tid; t < nloc * SPU; t += 256) {
        const int el = t / SPU, r = t - el * SPU;
        const int l = r % N, d = (r / N) % NV, i = r / (N * NV);
        const double* ue = u + (size_t)NV * NN * (e0 + el);
        double pu = 0.0;
        for (int k = 0; k < N; ++k) pu += P[i + NQ * k] * ue[d + NV * (k + N * l)];
        sM[el * (SW + SF) + r] = pu;
    }
    __syncthreads();
    for (int t = tid; t < nloc * NQ * NQ; t += 256) {
        const int el = t / (NQ * NQ), r = t - el * NQ * NQ, i = r % NQ, j = r / NQ;
        const double* sW = sM + el * (SW + SF);
        double* sF = sM + el * (SW + SF) + SW;
        double X[2], J[4];
        corner_map(corners + 8 * (size_t)(e0 + el), xq[i], xq[j], X, J);
        const double wij = wq[i] * wq[j];
        const double Ye = J[3] * wij, Xe = J[2] * wij, Yx = J[1] * wij, Xx = J[0] * wij;
        double uq[NV], Fq[2 * NV];
        for (int d = 0; d < NV; ++d) {
            double s = 0.0;
            for (int l = 0; l < N; ++l) s += P[j + NQ * l] * sW[l + N * (d + NV * i)];
            uq[d] = s;
        }
        for (int d = 0; d < 2 * NV; ++d) Fq[d] = 0.0;
        flux(X, uq, Fq);
        for (int d = 0; d < NV; ++d) {
            sF[0 + 2 * (i + NQ * (d + NV * j))] = Ye * Fq[d] - Xe * Fq[NV + d];
            sF[1 + 2 * (i + NQ * (d + NV * j))] = -Yx * Fq[d] + Xx * Fq[NV + d];
        }
    }
    __syncthreads();
    for (int t = tid; t < nloc * NQ * NV * N; t += 256) {
        const int el = t / (NQ * NV * N), r = t - el * NQ * NV * N;
        const int j = r % NQ, d = (r / NQ) % NV, k = r / (NQ * NV);
        double* sW = sM + el * (SW + SF);
        const double* sF = sW + SW;
        double df = 0.0, pg = 0.0;
        for (int i = 0; i < NQ; ++i) {
            df += D[i + NQ * k] * sF[0 + 2 * (i + NQ * (d + NV * j))];
            pg += P[i + NQ * k] * sF[1 + 2 * (i + NQ * (d + NV * j))];
        }
        sW[2 * r] = df;
        sW[2 * r + 1] = pg;
    }
    __syncthreads();
    for (int t = tid; t < nloc * NN * NV; t += 256) {
        const int el = t / (NN * NV), r = t - el * NN * NV;
        const int d = r % NV, k = (r / NV) % N, l = r / (NV * N);
        const double* sW = sM + el * (SW + SF);
        double dF = 0.0;
        for (int j = 0; j < NQ; ++j) {
            const int b = 2 * (j + NQ * (d + NV * k));
            dF += P[j + NQ * l] * sW[b] + D[j + NQ * l] * sW[b + 1];
        }
        out[(size_t)NV * NN * e0 + t] += dF;
    }
}
#endif

#ifdef WDG_WANT_EDGE_Q
// EdgeFluxF2D<false>::action (EdgeFluxF.hpp:182-250) between the face
// prolongation and its transpose, element by element: each side's traces at
// the side nodes (own order) are interpolated to the NQ Gauss-Legendre edge
// points, fh = numerical_flux(x, n, uL, uR) with el0's normal and uL = el0's
// trace, and sum_q (+-scale) fh_q w_q ds P(q, k) is lifted into side node k.
// tab = xq[NQ], wq[NQ], P[NQ * N].  E elements per 64-lane workgroup.
extern "C" __global__ void __launch_bounds__(64) wdg_edgeflux_q_rtc(const double* __restrict__ u,
                                                                     double* __restrict__ out,
                                                                     const double* __restrict__ corners,
                                                                     const int* __restrict__ nbr,
                                                                     const int* __restrict__ info,
                                                                     const double* __restrict__ basis,
                                                                     const double* __restrict__ tab, int nel,
                                                                     int interior, int boundary, double scale)
{
    using namespace wdg_rtc;
    constexpr int NQ = WDG_NQ, W = NQ > N ? NQ : N, E = 64 / W;
    {
        // a workgroup whose elements have no side in this launch (e.g. away
        // from the boundary) has nothing to do
        bool act = false;
        for (int t = threadIdx.x; t < 4 * (nel - (int)blockIdx.x * E < E ? nel - (int)blockIdx.x * E : E); t += 64) {
            const int nb = nbr[4 * ((size_t)blockIdx.x * E + t / 4) + (t % 4)];
            act = act || (nb >= 0 ? interior != 0 : boundary != 0);
        }
        if (!__syncthreads_or(act)) return;
    }
    __shared__ double s_out[E * NV * NN];
    __shared__ double sA[E * N * NV], sB[E * N * NV], sC[E * NQ * NV];
    __shared__ int s_ok[E], s_any[E];
    const int tid = threadIdx.x, e0 = blockIdx.x * E;
    const int nloc = nel - e0 < E ? nel - e0 : E;
    const double* xq = tab;
    const double* wq = tab + NQ;
    const double* Pq = tab + 2 * NQ;
    const double* Pf = basis + 2 * N + NN;  // Gauss-Legendre face interpolation
    for (int t = tid; t < E * NV * NN; t += 64) s_out[t] = 0.0;
    if (tid < E) s_any[tid] = 0;
    for (int s = 0; s < 4; ++s) {
        __syncthreads();  // the previous side's sA, sB, sC are consumed
        // traces at the side nodes in this element's order
        for (int t = tid; t < E * N; t += 64) {
            const int el = t / N, k = t - el * N, e = e0 + el;
            if (el >= nloc) continue;
            const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
            const bool on = nb >= 0 ? interior : boundary;
            if (k == 0) s_ok[el] = on;
            if (!on) continue;
            s_any[el] = 1;
            side_trace(u, e, s, k, Pf, sA + (el * N + k) * NV);
            if (nb >= 0) side_trace(u, nb, inf & 3, ((inf >> 2) & 1) ? (N - 1 - k) : k, Pf, sB + (el * N + k) * NV);
        }
        __syncthreads();
        // numerical flux at the edge points
        for (int t = tid; t < E * NQ; t += 64) {
            const int el = t / NQ, q = t - el * NQ, e = e0 + el;
            if (el >= nloc || !s_ok[el]) continue;
            const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
            const double* c = corners + 8 * (size_t)e;
            const int a0 = s == 0 ? 0 : s == 1 ? 1 : s == 2 ? 3 : 0;
            const int a1 = s == 0 ? 1 : s == 1 ? 2 : s == 2 ? 2 : 3;
            const double dx = c[2 * a1] - c[2 * a0], dy = c[2 * a1 + 1] - c[2 * a0 + 1];
            const double len = sqrt(dx * dx + dy * dy), sg = s >= 2 ? -1.0 : 1.0;
            double nrm[2] = {sg * dy / len, -sg * dx / len};
            const double ds = 0.5 * len;
            const double xi = s == 0 || s == 2 ? xq[q] : (s == 1 ? 1.0 : -1.0);
            const double eta = s == 1 || s == 3 ? xq[q] : (s == 2 ? 1.0 : -1.0);
            double X[2], J[4];
            corner_map(c, xi, eta, X, J);
            double uA[NV], uB[NV], fh[NV];
            for (int d = 0; d < NV; ++d) {
                double a = 0.0, b = 0.0;
                for (int k = 0; k < N; ++k) {
                    a += Pq[q + NQ * k] * sA[(el * N + k) * NV + d];
                    b += Pq[q + NQ * k] * sB[(el * N + k) * NV + d];
                }
                uA[d] = a;
                uB[d] = b;
            }
            if (nb < 0) {
#ifdef WDG_EXTERIOR_STATE
                exterior_state(X, nrm, uA, uB);
#else
                for (int d = 0; d < NV; ++d) uB[d] = uA[d];
#endif
            }
            const bool first = (inf >> 3) & 1;
            for (int d = 0; d < NV; ++d) fh[d] = 0.0;
            if (first) {
                numerical_flux(X, nrm, uA, uB, fh);
            } else {
                nrm[0] = -nrm[0];
                nrm[1] = -nrm[1];
                numerical_flux(X, nrm, uB, uA, fh);
            }
            const double sgn = first ? scale : -scale;
            for (int d = 0; d < NV; ++d) sC[(el * NQ + q) * NV + d] = sgn * (fh[d] * (wq[q] * ds));
        }
        __syncthreads();
        // projection onto the side nodes and lift
        for (int t = tid; t < E * N; t += 64) {
            const int el = t / N, k = t - el * N;
            if (el >= nloc || !s_ok[el]) continue;
            const int node = side_node(s, k), st = inward_stride(s);
            for (int d = 0; d < NV; ++d) {
                double fl = 0.0;
                for (int q = 0; q < NQ; ++q) fl += sC[(el * NQ + q) * NV + d] * Pq[q + NQ * k];
                if (WDG_LEG) {
                    for (int m = 0; m < N; ++m) s_out[(el * NN + node + m * st) * NV + d] += fl * Pf[m];
                } else {
                    s_out[(el * NN + node) * NV + d] += fl;
                }
            }
        }
    }
    __syncthreads();
    double* yb = out + (size_t)e0 * NV * NN;
    for (int t = tid; t < nloc * NV * NN; t += 64)
        if (s_any[t / (NV * NN)]) yb[t] += s_out[t];
}
#endif
)WDGRTC";

struct Compiled {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
};

}  // namespace

// Compiles the caller's source with the library's kernel for one order,
// variable count and basis into a gfx950 code object (no device needed).
bool rtc_build(const std::string& user, int n, int nv, int leg, int kind, int nq, std::vector<char>& code,
               std::string& err)
{
    static const char* const kWant[4] = {"-DWDG_WANT_DIVF", "-DWDG_WANT_EDGE", "-DWDG_WANT_DIVF_Q",
                                         "-DWDG_WANT_EDGE_Q"};
    if (kind < 0 || kind > 3) {
        err = "bad kernel kind";
        return false;
    }
    const std::string src = user + "\n" + kRtcKernels;
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "wdg_flux.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        err = "hiprtcCreateProgram failed";
        return false;
    }
    const std::string dn = "-DWDG_N=" + std::to_string(n), dv = "-DWDG_NVAR=" + std::to_string(nv),
                      dl = "-DWDG_LEG=" + std::to_string(leg ? 1 : 0), dq = "-DWDG_NQ=" + std::to_string(nq);
    // the same code-generation options as the library's own kernels; only
    // the kernel asked for is compiled, so a source may define one functor
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops", dn.c_str(), dv.c_str(), dl.c_str(), dq.c_str(), kWant[kind]};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        err = "flux source does not compile: " + log;
        return false;
    }
    size_t sz = 0;
    hiprtcGetCodeSize(prog, &sz);
    code.assign(sz, 0);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    return true;
}

namespace {

// rtc_build, then load the kernel on the current device.
bool rtc_compile(const std::string& user, int n, int nv, int leg, int kind, int nq, Compiled& out, std::string& err)
{
    static const char* const kName[4] = {"wdg_divf_rtc", "wdg_edgeflux_rtc", "wdg_divf_q_rtc", "wdg_edgeflux_q_rtc"};
    const char* kernel = kName[kind];
    std::vector<char> code;
    if (!rtc_build(user, n, nv, leg, kind, nq, code, err)) return false;
    if (hipModuleLoadData(&out.mod, code.data()) != hipSuccess) {
        err = "hipModuleLoadData failed for the compiled flux kernels";
        return false;
    }
    if (hipModuleGetFunction(&out.fn, out.mod, kernel) != hipSuccess) {
        (void)hipModuleUnload(out.mod);
        out.mod = nullptr;
        err = std::string("compiled flux module lacks ") + kernel;
        return false;
    }
    return true;
}

}  // namespace

// elements per workgroup of wdg_divf_q_rtc (EQ there)
int rtc_divf_q_elems(int n, int nv, int nq)
{
    const int a = nq * nq > n * nq ? nq * nq : n * nq;
    const int w = a * nv > n * n * nv ? a * nv : n * n * nv;
    return 256 / w > 0 ? 256 / w : 1;
}

struct RtcFlux {
    std::string source;
    int n = 0, nv = 0, leg = 0;
    int nqv = 0, nqe = 0;                 // > 0: the quadrature (approx_quad = false) forms
    double* d_tabv = nullptr;             // xq, wq, P, D of the volume rule
    double* d_tabe = nullptr;             // xq, wq, P of the edge rule
    Compiled divf, edge;
    std::mutex mu;
    ~RtcFlux()
    {
        if (divf.mod) (void)hipModuleUnload(divf.mod);
        if (edge.mod) (void)hipModuleUnload(edge.mod);
        if (d_tabv) (void)hipFree(d_tabv);
        if (d_tabe) (void)hipFree(d_tabe);
    }
};

RtcFlux* rtc_flux_create(const char* source, int n, int nv, int leg)
{
    auto* f = new RtcFlux();
    f->source = source;
    f->n = n;
    f->nv = nv;
    f->leg = leg;
    return f;
}

bool rtc_flux_set_rules(RtcFlux* f, int nqv, const std::vector<double>& tabv, int nqe, const std::vector<double>& tabe,
                        std::string& err)
{
    auto upload = [&](double*& d, const std::vector<double>& h) {
        if (hipMalloc(&d, sizeof(double) * h.size()) != hipSuccess) return false;
        return hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!upload(f->d_tabv, tabv) || !upload(f->d_tabe, tabe)) {
        err = "cannot upload the quadrature tables";
        return false;
    }
    f->nqv = nqv;
    f->nqe = nqe;
    return true;
}

void rtc_flux_destroy(RtcFlux* f) { delete f; }

bool rtc_flux_prepare(RtcFlux* f, int which, std::string& err)
{
    std::lock_guard<std::mutex> lk(f->mu);
    Compiled& c = which == 0 ? f->divf : f->edge;
    if (c.fn) return true;
    const bool q = f->nqv > 0;
    const int kind = (which == 0 ? 0 : 1) + (q ? 2 : 0);
    return rtc_compile(f->source, f->n, f->nv, f->leg, kind, which == 0 ? f->nqv : f->nqe, c, err);
}

bool rtc_divf(RtcFlux* f, const double* u, double* out, const double* corners, const double* basis, int nel,
              hipStream_t s, std::string& err)
{
    if (!rtc_flux_prepare(f, 0, err)) return false;
    if (nel <= 0) return true;
    const int NN = f->n * f->n, E = f->nqv > 0 ? rtc_divf_q_elems(f->n, f->nv, f->nqv) : 256 / NN;
    const double* tab = f->nqv > 0 ? f->d_tabv : basis;
    void* params[] = {&u, &out, &corners, &tab, &nel};
    const unsigned threads = f->nqv > 0 ? 256 : E * NN;
    if (hipModuleLaunchKernel(f->divf.fn, (nel + E - 1) / E, 1, 1, threads, 1, 1, 0, s, params, nullptr) !=
        hipSuccess) {
        err = "flux divergence kernel launch failed";
        return false;
    }
    return true;
}

bool rtc_edgeflux(RtcFlux* f, const double* u, double* out, const double* corners, const int32_t* nbr,
                  const int32_t* info, const double* basis, int nel, int interior, int boundary, double scale,
                  hipStream_t s, std::string& err)
{
    if (!rtc_flux_prepare(f, 1, err)) return false;
    if (nel <= 0) return true;
    if (f->nqe > 0) {
        const int E = 64 / (f->nqe > f->n ? f->nqe : f->n);
        const double* tab = f->d_tabe;
        void* params[] = {&u, &out, &corners, &nbr, &info, &basis, &tab, &nel, &interior, &boundary, &scale};
        if (hipModuleLaunchKernel(f->edge.fn, (nel + E - 1) / E, 1, 1, 64, 1, 1, 0, s, params, nullptr) !=
            hipSuccess) {
            err = "edge flux kernel launch failed";
            return false;
        }
        return true;
    }
    const int E = 64 / (4 * f->n) > 0 ? 64 / (4 * f->n) : 1;  // wdg_edgeflux_rtc's elements per workgroup
    void* params[] = {&u, &out, &corners, &nbr, &info, &basis, &nel, &interior, &boundary, &scale};
    if (hipModuleLaunchKernel(f->edge.fn, (nel + E - 1) / E, 1, 1, 64, 1, 1, 0, s, params, nullptr) != hipSuccess) {
        err = "edge flux kernel launch failed";
        return false;
    }
    return true;
}

}  // namespace wdg
