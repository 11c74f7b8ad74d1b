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
// one writer.
#include <hip/hip_runtime.h>

#include "wave2d_kernels.hpp"

namespace wdg {
namespace {

constexpr int kMaxVar = DivFArgs::kMaxVar;

// E = max(1, 64 / (N nv)) elements per 64-lane workgroup, one lane per
// (element, node of the side, variable) in each side round.
template <int N>
__global__ __launch_bounds__(64) void advect_face_kernel(const double* __restrict__ u, double* __restrict__ y,
                                                         const int32_t* __restrict__ nbr,
                                                         const int32_t* __restrict__ info,
                                                         const double* __restrict__ H, int nv, int nel, int interior,
                                                         int boundary)
{
    constexpr int NN = N * N;
    constexpr int kOut = kMaxVar * NN > 64 * N ? kMaxVar * NN : 64 * N;
    __shared__ double s_out[kOut];
    const int per = N * nv, E = per < 64 ? 64 / per : 1;
    const int tid = threadIdx.x, e0 = blockIdx.x * E, vn = nv * NN, v2 = nv * nv;
    const int nloc = nel - e0 < E ? nel - e0 : E;
    for (int k = tid; k < E * vn; k += 64) s_out[k] = 0.0;
    __syncthreads();
    for (int s = 0; s < 4; ++s) {
        for (int it = tid; it < E * per; it += 64) {
            const int el = it / per, r = it - el * per, k = r / nv, d = r - k * nv, e = e0 + el;
            if (el >= nloc) continue;
            const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
            const bool is_int = nb >= 0;
            if (is_int ? !interior : !boundary) continue;
            const int node = side_node<N>(s, k);
            const double* uA = u + (size_t)nv * ((size_t)e * NN + node);
            const double* Hk = H + (((size_t)e * 4 + s) * N + k) * 2 * v2;
            double acc = 0.0;
            for (int c = 0; c < nv; ++c) acc += Hk[d + nv * c] * uA[c];
            if (is_int) {
                const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
                const double* uB = u + (size_t)nv * ((size_t)nb * NN + side_node<N>(inf & 3, kk));
                for (int c = 0; c < nv; ++c) acc += Hk[v2 + d + nv * c] * uB[c];
            }
            s_out[el * vn + d + nv * node] += acc;
        }
        __syncthreads();
    }
    double* yb = y + (size_t)e0 * vn;
    for (int k = tid; k < nloc * vn; k += 64) yb[k] += s_out[k];
}

}  // namespace

void launch_advect_face(int n, int nv, const double* u, double* y, const int32_t* nbr, const int32_t* info,
                        const double* H, int nel, int interior, int boundary, hipStream_t s)
{
    if (nel <= 0 || nv < 1 || nv > kMaxVar) return;
    const int per = n * nv, E = per < 64 ? 64 / per : 1;
    const dim3 grid((nel + E - 1) / E), block(64);
#define WDG_ADVECT_CASE(NV_)                                                                                  \
    case NV_:                                                                                                 \
        hipLaunchKernelGGL(advect_face_kernel<NV_>, grid, block, 0, s, u, y, nbr, info, H, nv, nel, interior, \
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

}  // namespace wdg
