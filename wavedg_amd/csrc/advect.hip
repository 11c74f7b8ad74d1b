// Face terms of Advection<true> and AdvectionHomogeneousBC<true>
// (Advection.hpp:15-273): EdgeFlux<true> with a = -1, b = -0.5
// (EdgeFlux.cpp:91-186, 289-325) and the transpose of the face prolongation
// (FaceProlongator.cpp:650-672), element by element.  For element e, side s
// and node k the host precomputes, with this element's outward normal n,
//   H_own   = w_k ds (a/2 n.A(own node) + b |n.A(own node)|)
//   H_other = w_k ds (a/2 n.A(other node) - b |n.A(other node)|)
// which is +-(the reference's flux table) of whichever side this element is on,
// so y(node) += H_own u(node) + H_other u(other node), and on boundary sides
// (u = 0 outside) y(node) += H_own u(node).  Sides are taken one round at a
// time and corner nodes get their two sides in side order: every output has
// one writer.  With a Gauss-Legendre basis (P = phi_m(-1) non-null) "the
// node" of a side is the line of nodes through it: u(node) stands for
// sum_m P(m) u(m rows in) and the result is lifted back along that line with
// the same weights (legendre_action_2d / legendre_transpose_2d,
// FaceProlongator.cpp:809-835, 908-934).
#include <hip/hip_runtime.h>

#include "wave2d_kernels.hpp"

namespace wdg {
namespace {

constexpr int kMaxVar = DivFArgs::kMaxVar;

// One lane per (element, side node, variable) for all four sides at once,
// E = max(1, 64 / (4 N nv)) elements per 64-lane workgroup (lanes loop when
// an element needs more); each lane leaves its contribution in LDS and the
// nodes then add their sides in side order (only the nodes a side reaches).
template <int N>
__global__ __launch_bounds__(64) void advect_face_kernel(const double* __restrict__ u, double* __restrict__ y,
                                                         const int32_t* __restrict__ nbr,
                                                         const int32_t* __restrict__ info,
                                                         const double* __restrict__ H, const double* __restrict__ P,
                                                         int nv, int nel, int interior, int boundary)
{
    constexpr int NN = N * N, NF = 4 * N;
    constexpr int kF = NF * kMaxVar > 64 ? NF * kMaxVar : 64;
    __shared__ double s_f[kF];  // E NF nv contributions (E NF nv <= max(64, NF nv))
    __shared__ int s_any[64];   // the element has a side in this launch
    const int per = NF * nv, E = per < 64 ? 64 / per : 1;
    const int tid = threadIdx.x, e0 = blockIdx.x * E, vn = nv * NN, v2 = nv * nv;
    const int nloc = nel - e0 < E ? nel - e0 : E;
    if (tid < E) s_any[tid] = 0;
    {
        // a workgroup whose elements have no side in this launch (e.g. away
        // from the boundary) has nothing to do
        bool act = false;
        for (int t = tid; t < 4 * nloc; t += 64) {
            const int nb = nbr[4 * ((size_t)e0 + t / 4) + (t % 4)];
            act = act || (nb >= 0 ? interior != 0 : boundary != 0);
        }
        if (!__syncthreads_or(act)) return;
    }
    for (int it = tid; it < E * per; it += 64) {
        const int el = it / per, r = it - el * per, f = r / nv, d = r - f * nv, s = f / N, k = f - s * N;
        const int e = e0 + el;
        double acc = 0.0;
        if (el < nloc) {
            const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
            const bool is_int = nb >= 0;
            if (is_int ? interior : boundary) {
                s_any[el] = 1;
                const int node = side_node<N>(s, k);
                const double* uA = u + (size_t)nv * ((size_t)e * NN + node);
                const double* Hk = H + (((size_t)e * 4 + s) * N + k) * 2 * v2;
                if (P) {
                    const int st = nv * inward_stride<N>(s);
                    for (int c = 0; c < nv; ++c) {
                        double t = 0.0;
                        for (int m = 0; m < N; ++m) t += uA[c + m * st] * P[m];
                        acc += Hk[d + nv * c] * t;
                    }
                } else {
                    for (int c = 0; c < nv; ++c) acc += Hk[d + nv * c] * uA[c];
                }
                if (is_int) {
                    const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
                    const double* uB = u + (size_t)nv * ((size_t)nb * NN + side_node<N>(inf & 3, kk));
                    if (P) {
                        const int st = nv * inward_stride<N>(inf & 3);
                        for (int c = 0; c < nv; ++c) {
                            double t = 0.0;
                            for (int m = 0; m < N; ++m) t += uB[c + m * st] * P[m];
                            acc += Hk[v2 + d + nv * c] * t;
                        }
                    } else {
                        for (int c = 0; c < nv; ++c) acc += Hk[v2 + d + nv * c] * uB[c];
                    }
                }
            }
        }
        s_f[(el * NF + f) * nv + d] = acc;
    }
    __syncthreads();
    double* yb = y + (size_t)e0 * vn;
    for (int t = tid; t < nloc * vn; t += 64) {
        const int el = t / vn, r = t - el * vn, node = r / nv, d = r - node * nv;
        if (!s_any[el]) continue;
        const int i = node % N, j = node / N;
        const double* F = s_f + el * NF * nv + d;
        double add;
        if (P) {
            add = P[j] * F[(0 * N + i) * nv] + P[N - 1 - i] * F[(1 * N + j) * nv] +
                  P[N - 1 - j] * F[(2 * N + i) * nv] + P[i] * F[(3 * N + j) * nv];
        } else {
            if (!(i == 0 || j == 0 || i == N - 1 || j == N - 1)) continue;  // no side reaches an inner node
            add = 0.0;
            if (j == 0) add += F[(0 * N + i) * nv];
            if (i == N - 1) add += F[(1 * N + j) * nv];
            if (j == N - 1) add += F[(2 * N + i) * nv];
            if (i == 0) add += F[(3 * N + j) * nv];
        }
        yb[t] += add;
    }
}

// Compressible Euler flux and local Lax-Friedrichs flux of examples/cl2d.cpp
// (F :57-81, speed :84-96, numerical_flux :98-111).
__device__ __forceinline__ void euler_fg(double gamma, const double* q, double* f, double* g, double& speed)
{
    const double rho = q[0], u = q[1] / q[0], v = q[2] / q[0], E = q[3];
    const double p = (gamma - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    f[0] = rho * u;
    f[1] = rho * u * u + p;
    f[2] = rho * u * v;
    f[3] = u * (E + p);
    g[0] = rho * v;
    g[1] = rho * u * v;
    g[2] = rho * v * v + p;
    g[3] = v * (E + p);
    speed = hypot(u, v) + sqrt(gamma * p / rho);
}

// The four conserved variables on element side s at the side node whose
// state starts at p: its values (Gauss-Lobatto), or sum_m P(m) q(m rows in).
template <int N>
__device__ __forceinline__ void trace4(const double* p, int s, const double* P, double* out)
{
    if (!P) {
#pragma unroll
        for (int d = 0; d < 4; ++d) out[d] = p[d];
        return;
    }
    const int st = 4 * inward_stride<N>(s);
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        double t = 0.0;
        for (int m = 0; m < N; ++m) t += p[d + m * st] * P[m];
        out[d] = t;
    }
}

// EdgeFluxF2D<true> with the LLF flux, negated and lifted as in cl2d.cpp's
// time derivative (cl2d.cpp:180-215), element by element: at node k of side
// s, with this element's outward normal n, y -= w_k ds LLF(n, q_own, q_other)
// (the flux is odd in (n, qL, qR), so el0 and el1 of an edge agree with the
// reference's +-fh); on boundary sides q_other = q_own.  One lane per
// (element, side, side node), all four sides at once, E = 64 / 4N elements
// per workgroup; each lane leaves its contribution in LDS and the nodes then
// add their sides in side order 0..3 (corner nodes two), so every output has
// one writer.  A Gauss-Legendre basis (P non-null) interpolates the traces
// along the element and lifts every side back along its lines with P.
template <int N>
__global__ __launch_bounds__(64) void euler_face_kernel(double gamma, const double* __restrict__ q,
                                                        double* __restrict__ y, const double* __restrict__ geom,
                                                        const int32_t* __restrict__ nbr,
                                                        const int32_t* __restrict__ info,
                                                        const double* __restrict__ basis,
                                                        const double* __restrict__ P, int nel, int interior,
                                                        int boundary)
{
    constexpr int NN = N * N, NF = 4 * N, E = 64 / NF > 0 ? 64 / NF : 1;
    __shared__ double s_f[E * NF * 4];  // each side node's contribution, this element's order
    __shared__ int s_any[E];            // the element has a side in this launch
    const int tid = threadIdx.x, e0 = blockIdx.x * E;
    const int nloc = nel - e0 < E ? nel - e0 : E;
    if (tid < E) s_any[tid] = 0;
    {
        // a workgroup whose elements have no side in this launch (e.g. away
        // from the boundary) has nothing to do
        bool act = false;
        for (int t = tid; t < 4 * nloc; t += 64) {
            const int nb = nbr[4 * ((size_t)e0 + t / 4) + (t % 4)];
            act = act || (nb >= 0 ? interior != 0 : boundary != 0);
        }
        if (!__syncthreads_or(act)) return;
    }
    const int el = tid / NF, f = tid - el * NF, s = f / N, k = f - (f / N) * N, e = e0 + el;
    double c[4] = {0.0, 0.0, 0.0, 0.0};
    if (tid < E * NF && el < nloc) {
        const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
        const bool is_int = nb >= 0;
        if (is_int ? interior : boundary) {
            s_any[el] = 1;
            const double* g = geom + 6 * (size_t)e;
            const double sh = (s == 0 || s == 3) ? -1.0 : 1.0;
            const double dx = 2.0 * ((s & 1) ? g[2] + sh * g[4] : g[0] + sh * g[4]);
            const double dy = 2.0 * ((s & 1) ? g[3] + sh * g[5] : g[1] + sh * g[5]);
            const double len = sqrt(dx * dx + dy * dy), sg = (s >= 2) ? -1.0 : 1.0;
            const double n0 = sg * dy / len, n1 = -sg * dx / len, wds = basis[N + k] * (0.5 * len);
            const int node = side_node<N>(s, k);
            double qa[4], qb[4];
            trace4<N>(q + 4 * ((size_t)e * NN + node), s, P, qa);
            if (is_int) {
                const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
                trace4<N>(q + 4 * ((size_t)nb * NN + side_node<N>(inf & 3, kk)), inf & 3, P, qb);
            } else {
#pragma unroll
                for (int d = 0; d < 4; ++d) qb[d] = qa[d];
            }
            double fa[4], ga[4], fb[4], gb[4], sa, sb;
            euler_fg(gamma, qa, fa, ga, sa);
            euler_fg(gamma, qb, fb, gb, sb);
            const double cm = sa > sb ? sa : sb;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const double fh = 0.5 * n0 * (fa[d] + fb[d]) + 0.5 * n1 * (ga[d] + gb[d]) + 0.5 * cm * (qa[d] - qb[d]);
                c[d] = -(fh * wds);
            }
        }
    }
    if (tid < E * NF) {
#pragma unroll
        for (int d = 0; d < 4; ++d) s_f[(el * NF + f) * 4 + d] = c[d];
    }
    __syncthreads();
    for (int t = tid; t < nloc * 4 * NN; t += 64) {
        const int el2 = t / (4 * NN), r = t - el2 * 4 * NN, node = r / 4, d = r - node * 4;
        if (!s_any[el2]) continue;
        const int i = node % N, j = node / N;
        const double* F = s_f + el2 * NF * 4 + d;
        double add;
        if (P) {
            add = P[j] * F[(0 * N + i) * 4] + P[N - 1 - i] * F[(1 * N + j) * 4] + P[N - 1 - j] * F[(2 * N + i) * 4] +
                  P[i] * F[(3 * N + j) * 4];
        } else {
            if (!(i == 0 || j == 0 || i == N - 1 || j == N - 1)) continue;  // no side reaches an inner node
            add = 0.0;
            if (j == 0) add += F[(0 * N + i) * 4];
            if (i == N - 1) add += F[(1 * N + j) * 4];
            if (j == N - 1) add += F[(2 * N + i) * 4];
            if (i == 0) add += F[(3 * N + j) * 4];
        }
        y[(size_t)(e0 + el2) * 4 * NN + r] += add;
    }
}

// Advection<false> volume term (Div<false>, Div.cpp:312-360, 469-588): u
// interpolated to the nq x nq Gauss-Legendre points, the fluxes op(., 0) u and
// op(., 1) u there, projected back with D and P.  op (2, nv, nv, nq, nq, nel)
// holds the metric-weighted coefficient w_i w_j (y_eta A_x - x_eta A_y,
// -y_xi A_x + x_xi A_y) of every point, as the reference's op_2d table.
// tab = P (nq, N) then D (nq, N).  E elements per 256-lane workgroup (as many
// as the largest pass fills); the dynamic LDS holds, per element, Pu or
// Df/Pg, and F.
template <int N>
__global__ __launch_bounds__(256) void advect_volume_q_kernel(const double* __restrict__ u, double* __restrict__ out,
                                                              const double* __restrict__ op,
                                                              const double* __restrict__ tab, int nq, int nv, int E,
                                                              int nel)
{
    constexpr int NN = N * N;
    extern __shared__ double s_dyn[];
    const int tid = threadIdx.x, v2 = nv * nv, e0 = blockIdx.x * E;
    const int nloc = nel - e0 < E ? nel - e0 : E;
    const int spu = N * nv * nq, sdp = 2 * nq * nv * N, sw = spu > sdp ? spu : sdp, sf = 2 * nq * nv * nq;
    // the rule's P and D in LDS after the per-element regions
    double* sT = s_dyn + E * (sw + sf);
    for (int t = tid; t < 2 * nq * N; t += 256) sT[t] = tab[t];
    __syncthreads();
    const double* P = sT;
    const double* D = sT + nq * N;
    // per element: sW (Pu, then Df and Pg) and sF
    for (int t = tid; t < nloc * spu; t += 256) {
        const int el = t / spu, r = t - el * spu;
        const int l = r % N, d = (r / N) % nv, i = r / (N * nv);
        const double* ue = u + (size_t)nv * NN * (e0 + el);
        double pu = 0.0;
        for (int k = 0; k < N; ++k) pu += P[i + nq * k] * ue[d + nv * (k + N * l)];
        s_dyn[el * (sw + sf) + r] = pu;
    }
    __syncthreads();
    for (int t = tid; t < nloc * nq * nq * nv; t += 256) {
        const int el = t / (nq * nq * nv), r = t - el * nq * nq * nv;
        const int d = r % nv, i = (r / nv) % nq, j = r / (nv * nq);
        const double* sW = s_dyn + el * (sw + sf);
        double* sF = s_dyn + el * (sw + sf) + sw;
        const double* o = op + (size_t)2 * v2 * (i + nq * (j + (size_t)nq * (e0 + el)));
        double f = 0.0, g = 0.0;
        for (int c = 0; c < nv; ++c) {
            double uq = 0.0;
            for (int l = 0; l < N; ++l) uq += P[j + nq * l] * sW[l + N * (c + nv * i)];
            f += o[2 * (d + nv * c)] * uq;
            g += o[2 * (d + nv * c) + 1] * uq;
        }
        sF[0 + 2 * (i + nq * (d + nv * j))] = f;
        sF[1 + 2 * (i + nq * (d + nv * j))] = g;
    }
    __syncthreads();
    for (int t = tid; t < nloc * nq * nv * N; t += 256) {
        const int el = t / (nq * nv * N), r = t - el * nq * nv * N;
        const int j = r % nq, d = (r / nq) % nv, k = r / (nq * nv);
        double* sW = s_dyn + el * (sw + sf);
        const double* sF = sW + sw;
        double df = 0.0, pg = 0.0;
        for (int i = 0; i < nq; ++i) {
            df += D[i + nq * k] * sF[0 + 2 * (i + nq * (d + nv * j))];
            pg += P[i + nq * k] * sF[1 + 2 * (i + nq * (d + nv * j))];
        }
        sW[2 * r] = df;
        sW[2 * r + 1] = pg;
    }
    __syncthreads();
    for (int t = tid; t < nloc * NN * nv; t += 256) {
        const int el = t / (NN * nv), r = t - el * NN * nv;
        const int d = r % nv, k = (r / nv) % N, l = r / (nv * N);
        const double* sW = s_dyn + el * (sw + sf);
        double dF = 0.0;
        for (int j = 0; j < nq; ++j) {
            const int b = 2 * (j + nq * (d + nv * k));
            dF += P[j + nq * l] * sW[b] + D[j + nq * l] * sW[b + 1];
        }
        out[(size_t)nv * NN * e0 + t] += dF;
    }
}

// Advection<false> face terms (EdgeFlux<false>, EdgeFlux.cpp:375-640): per
// element side the traces at the side nodes (this element's order; lines of
// nodes for a Gauss-Legendre basis, P non-null) are interpolated to the nq
// edge points, c_q = H_own(q) u_own(q) + H_other(q) u_other(q) with H
// (2, nv, nv) per (element, side, point) already weighted by w_q ds, and
// sum_q Pe(q, k) c_q is lifted into side node k.  tab = Pe (nq, N).
// E = 64 / max(N, nq) elements per 64-lane workgroup; dynamic LDS.
template <int N>
__global__ __launch_bounds__(64) void advect_face_q_kernel(const double* __restrict__ u, double* __restrict__ y,
                                                            const int32_t* __restrict__ nbr,
                                                            const int32_t* __restrict__ info,
                                                            const double* __restrict__ Hq,
                                                            const double* __restrict__ tab,
                                                            const double* __restrict__ P, int nq, int nv, int nel,
                                                            int interior, int boundary)
{
    constexpr int NN = N * N;
    extern __shared__ double s_dyn[];
    const int W = nq > N ? nq : N, E = 64 / W, v2 = nv * nv;
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
    double* s_out = s_dyn;                 // E nv NN
    double* sA = s_out + E * nv * NN;      // E N nv
    double* sB = sA + E * N * nv;          // E N nv
    double* sC = sB + E * N * nv;          // E nq nv
    double* sP = sC + E * nq * nv;         // the edge rule's P (nq, N)
    __shared__ int s_any[64];  // the element has a side in this launch
    const int tid = threadIdx.x, e0 = blockIdx.x * E;
    const int nloc = nel - e0 < E ? nel - e0 : E;
    for (int t = tid; t < E * nv * NN; t += 64) s_out[t] = 0.0;
    for (int t = tid; t < nq * N; t += 64) sP[t] = tab[t];
    if (tid < E) s_any[tid] = 0;
    for (int s = 0; s < 4; ++s) {
        __syncthreads();
        for (int t = tid; t < E * N; t += 64) {
            const int el = t / N, k = t - el * N, e = e0 + el;
            if (el >= nloc) continue;
            const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
            if (nb >= 0 ? !interior : !boundary) continue;
            s_any[el] = 1;
            const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
            for (int d = 0; d < nv; ++d) {
                const double* qa = u + (size_t)nv * ((size_t)e * NN + side_node<N>(s, k)) + d;
                double a = 0.0, b = 0.0;
                if (P) {
                    const int st = nv * inward_stride<N>(s);
                    for (int m = 0; m < N; ++m) a += qa[m * st] * P[m];
                } else {
                    a = qa[0];
                }
                if (nb >= 0) {
                    const double* qb = u + (size_t)nv * ((size_t)nb * NN + side_node<N>(inf & 3, kk)) + d;
                    if (P) {
                        const int st = nv * inward_stride<N>(inf & 3);
                        for (int m = 0; m < N; ++m) b += qb[m * st] * P[m];
                    } else {
                        b = qb[0];
                    }
                }
                sA[(el * N + k) * nv + d] = a;
                sB[(el * N + k) * nv + d] = b;  // u = 0 outside the domain
            }
        }
        __syncthreads();
        for (int t = tid; t < E * nq * nv; t += 64) {
            const int el = t / (nq * nv), r = t - el * nq * nv, q = r / nv, d = r - q * nv, e = e0 + el;
            if (el >= nloc) continue;
            const int nb = nbr[4 * (size_t)e + s];
            if (nb >= 0 ? !interior : !boundary) continue;
            const double* H = Hq + (((size_t)e * 4 + s) * nq + q) * 2 * v2;
            double c = 0.0;
            for (int cc = 0; cc < nv; ++cc) {
                double a = 0.0, b = 0.0;
                for (int k = 0; k < N; ++k) {
                    a += sP[q + nq * k] * sA[(el * N + k) * nv + cc];
                    b += sP[q + nq * k] * sB[(el * N + k) * nv + cc];
                }
                c += H[d + nv * cc] * a + H[v2 + d + nv * cc] * b;
            }
            sC[(el * nq + q) * nv + d] = c;
        }
        __syncthreads();
        for (int t = tid; t < E * N * nv; t += 64) {
            const int el = t / (N * nv), r = t - el * N * nv, k = r / nv, d = r - k * nv, e = e0 + el;
            if (el >= nloc) continue;
            const int nb = nbr[4 * (size_t)e + s];
            if (nb >= 0 ? !interior : !boundary) continue;
            double fl = 0.0;
            for (int q = 0; q < nq; ++q) fl += sC[(el * nq + q) * nv + d] * sP[q + nq * k];
            const int node = side_node<N>(s, k);
            if (P) {
                // lane (k, d) owns variable d of the line through side node k
                const int st = inward_stride<N>(s);
                for (int m = 0; m < N; ++m) s_out[el * nv * NN + d + nv * (node + m * st)] += fl * P[m];
            } else {
                s_out[el * nv * NN + d + nv * node] += fl;
            }
        }
    }
    __syncthreads();
    double* yb = y + (size_t)e0 * nv * NN;
    for (int t = tid; t < nloc * nv * NN; t += 64)
        if (s_any[t / (nv * NN)]) yb[t] += s_out[t];
}

template <int N>
__global__ __launch_bounds__(64) void linear_functional_kernel(const double* __restrict__ corners,
                                                               const double* __restrict__ tab, int nq, int nv,
                                                               const double* __restrict__ fq, double* __restrict__ F)
{
    constexpr int kQ = 10;
    __shared__ double s_g[kMaxVar * kQ * kQ], s_p[kMaxVar * kQ * N];
    const int e = blockIdx.x, tid = threadIdx.x;
    const double* xq = tab;
    const double* wq = tab + nq;
    const double* B = tab + 2 * nq;  // B(k, i) at k + nq i
    const double* c = corners + 8 * (size_t)e;
    // f w_k w_l det J at the points (the reference scales Fq by detJ W(k) W(l))
    for (int t = tid; t < nv * nq * nq; t += 64) {
        const int d = t % nv, k = (t / nv) % nq, l = t / (nv * nq);
        const double xi = xq[k], eta = xq[l];
        const double J0 = 0.25 * ((1.0 - eta) * (c[2] - c[0]) + (1.0 + eta) * (c[4] - c[6]));
        const double J1 = 0.25 * ((1.0 - eta) * (c[3] - c[1]) + (1.0 + eta) * (c[5] - c[7]));
        const double J2 = 0.25 * ((1.0 - xi) * (c[6] - c[0]) + (1.0 + xi) * (c[4] - c[2]));
        const double J3 = 0.25 * ((1.0 - xi) * (c[7] - c[1]) + (1.0 + xi) * (c[5] - c[3]));
        s_g[t] = fq[(size_t)nv * nq * nq * e + t] * ((J0 * J3 - J1 * J2) * wq[k] * wq[l]);
    }
    __syncthreads();
    // project along y: P(d, k, j) = sum_l g(d, k, l) B(l, j)
    for (int t = tid; t < nv * nq * N; t += 64) {
        const int d = t % nv, k = (t / nv) % nq, j = t / (nv * nq);
        double s = 0.0;
        for (int l = 0; l < nq; ++l) s += s_g[d + nv * (k + nq * l)] * B[l + nq * j];
        s_p[t] = s;
    }
    __syncthreads();
    // project along x: F(d, i, j) = sum_k P(d, k, j) B(k, i)
    for (int t = tid; t < nv * N * N; t += 64) {
        const int d = t % nv, i = (t / nv) % N, j = t / (nv * N);
        double s = 0.0;
        for (int k = 0; k < nq; ++k) s += s_p[d + nv * (k + nq * j)] * B[k + nq * i];
        F[(size_t)nv * N * N * e + t] = s;
    }
}

}  // namespace

void launch_linear_functional(int n, int nq, int nv, const double* corners, const double* tab, const double* fq,
                              double* F, int nel, hipStream_t s)
{
    if (nel <= 0) return;
#define WDG_LF_CASE(NV_)                                                                                      \
    case NV_:                                                                                                 \
        hipLaunchKernelGGL(linear_functional_kernel<NV_>, dim3(nel), dim3(64), 0, s, corners, tab, nq, nv, fq, F); \
        break;
    switch (n) {
        WDG_LF_CASE(2)
        WDG_LF_CASE(3)
        WDG_LF_CASE(4)
        WDG_LF_CASE(5)
        WDG_LF_CASE(6)
        WDG_LF_CASE(7)
        WDG_LF_CASE(8)
        WDG_LF_CASE(9)
    default: break;
    }
#undef WDG_LF_CASE
}

// elements per workgroup of euler_face_kernel: one lane per side node
static int euler_elems(int n) { return 64 / (4 * n) > 0 ? 64 / (4 * n) : 1; }

void launch_euler_face(int n, double gamma, const double* q, double* y, const double* geom, const int32_t* nbr,
                       const int32_t* info, const double* basis, const double* P, int nel, int interior, int boundary,
                       hipStream_t s)
{
    if (nel <= 0) return;
    const dim3 block(64);
#define WDG_EULER_CASE(NV_)                                                                                   \
    case NV_:                                                                                                 \
        hipLaunchKernelGGL(euler_face_kernel<NV_>, dim3((nel + euler_elems(NV_) - 1) / euler_elems(NV_)), block, 0, \
                           s, gamma, q, y, geom, nbr, info, basis, P, nel, interior, boundary);               \
        break;
    switch (n) {
        WDG_EULER_CASE(2)
        WDG_EULER_CASE(3)
        WDG_EULER_CASE(4)
        WDG_EULER_CASE(5)
        WDG_EULER_CASE(6)
        WDG_EULER_CASE(7)
        WDG_EULER_CASE(8)
        WDG_EULER_CASE(9)
    default: break;
    }
#undef WDG_EULER_CASE
}

void launch_advect_face(int n, int nv, const double* u, double* y, const int32_t* nbr, const int32_t* info,
                        const double* H, const double* P, int nel, int interior, int boundary, hipStream_t s)
{
    if (nel <= 0 || nv < 1 || nv > kMaxVar) return;
    const int per = 4 * n * nv, E = per < 64 ? 64 / per : 1;  // advect_face_kernel's elements per workgroup
    const dim3 grid((nel + E - 1) / E), block(64);
#define WDG_ADVECT_CASE(NV_)                                                                                  \
    case NV_:                                                                                                 \
        hipLaunchKernelGGL(advect_face_kernel<NV_>, grid, block, 0, s, u, y, nbr, info, H, P, nv, nel, interior, \
                           boundary);                                                                         \
        break;
    switch (n) {
        WDG_ADVECT_CASE(2)
        WDG_ADVECT_CASE(3)
        WDG_ADVECT_CASE(4)
        WDG_ADVECT_CASE(5)
        WDG_ADVECT_CASE(6)
        WDG_ADVECT_CASE(7)
        WDG_ADVECT_CASE(8)
        WDG_ADVECT_CASE(9)
    default: break;
    }
#undef WDG_ADVECT_CASE
}

// elements per workgroup of the volume kernel: the largest pass fills the
// 256 lanes at least once
static int advect_volume_q_elems(int n, int nq, int nv)
{
    int w = nq * nq * nv;
    w = w > n * nv * nq ? w : n * nv * nq;
    w = w > n * n * nv ? w : n * n * nv;
    const int e = 256 / w;
    return e < 1 ? 1 : e;
}

size_t advect_volume_q_lds(int n, int nq, int nv)
{
    const int spu = n * nv * nq, sdp = 2 * nq * nv * n;
    return sizeof(double) * ((size_t)advect_volume_q_elems(n, nq, nv) *
                                 ((size_t)(spu > sdp ? spu : sdp) + 2 * (size_t)nq * nv * nq) +
                             2 * (size_t)nq * n);
}

size_t advect_face_q_lds(int n, int nq, int nv)
{
    const int W = nq > n ? nq : n, E = 64 / W;
    return sizeof(double) * ((size_t)E * nv * n * n + 2 * (size_t)E * n * nv + (size_t)E * nq * nv + (size_t)nq * n);
}

void launch_advect_q(int n, int nv, int nq_volume, int nq_edge, const double* u, double* y, const double* op,
                     const double* tabv, const int32_t* nbr, const int32_t* info, const double* Hq,
                     const double* tabe, const double* P, int nel, int volume, int interior, int boundary,
                     hipStream_t s)
{
    if (nel <= 0 || nv < 1 || nv > kMaxVar) return;
    const size_t lv = advect_volume_q_lds(n, nq_volume, nv), lf = advect_face_q_lds(n, nq_edge, nv);
    const int Ev = advect_volume_q_elems(n, nq_volume, nv);
    const int W = nq_edge > n ? nq_edge : n, E = 64 / W;
#define WDG_ADVQ_CASE(NV_)                                                                                       \
    case NV_:                                                                                                    \
        if (volume)                                                                                              \
            hipLaunchKernelGGL(advect_volume_q_kernel<NV_>, dim3((nel + Ev - 1) / Ev), dim3(256), lv, s, u, y, op, \
                               tabv, nq_volume, nv, Ev, nel);                                                    \
        if (interior || boundary)                                                                                \
            hipLaunchKernelGGL(advect_face_q_kernel<NV_>, dim3((nel + E - 1) / E), dim3(64), lf, s, u, y, nbr,     \
                               info, Hq, tabe, P, nq_edge, nv, nel, interior, boundary);                          \
        break;
    switch (n) {
        WDG_ADVQ_CASE(2)
        WDG_ADVQ_CASE(3)
        WDG_ADVQ_CASE(4)
        WDG_ADVQ_CASE(5)
        WDG_ADVQ_CASE(6)
        WDG_ADVQ_CASE(7)
        WDG_ADVQ_CASE(8)
        WDG_ADVQ_CASE(9)
    default: break;
    }
#undef WDG_ADVQ_CASE
}

}  // namespace wdg
