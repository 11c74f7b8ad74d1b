// Face terms of Advection<true> and AdvectionHomogeneousBC<true>
// (Advection.hpp:15-273): EdgeFlux<true> with a = -1, b = -0.5
// (EdgeFlux.cpp:91-186, 289-325) and the transpose of the face prolongation
// (FaceProlongator.cpp:650-672), element by element.  For element e, side s
// and node k the host precomputes, with this element's outward normal n,
//   H_own   = w_k ds (a/2 n.A(own node) + b |n.A(own node)|)
//   H_other = w_k ds (a/2 n.A(other node) - b |n.A(other node)|)
// which is +-(the reference's flux table) of whichever side this element is on,
// so y(node) += H_own u(node) + H_other u(other node), and on boundary sides
// (u = 0 outside) y(node) += H_own u(node).  Corner nodes take their two
// sides in side order: every output has one writer.
#include <hip/hip_runtime.h>

#include "wave2d_kernels.hpp"

namespace wdg {
namespace {

constexpr int kMaxVar = DivFArgs::kMaxVar;

template <int N>
__global__ __launch_bounds__(64) void advect_face_kernel(const double* __restrict__ u, double* __restrict__ y,
                                                         const int32_t* __restrict__ nbr,
                                                         const int32_t* __restrict__ info,
                                                         const double* __restrict__ H, int nv, int interior,
                                                         int boundary)
{
    constexpr int NN = N * N;
    __shared__ double s_out[kMaxVar * NN];
    const int e = blockIdx.x, tid = threadIdx.x, vn = nv * NN, v2 = nv * nv;
    for (int k = tid; k < vn; k += 64) s_out[k] = 0.0;
    __syncthreads();
    for (int s = 0; s < 4; ++s) {
        const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
        const bool is_int = nb >= 0;
        if (is_int ? !interior : !boundary) continue;  // uniform over the workgroup
        for (int it = tid; it < N * nv; it += 64) {
            const int k = it / nv, d = it - k * nv, node = side_node<N>(s, k);
            const double* uA = u + (size_t)nv * ((size_t)e * NN + node);
            const double* Hk = H + (((size_t)e * 4 + s) * N + k) * 2 * v2;
            double acc = 0.0;
            for (int c = 0; c < nv; ++c) acc += Hk[d + nv * c] * uA[c];
            if (is_int) {
                const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
                const double* uB = u + (size_t)nv * ((size_t)nb * NN + side_node<N>(inf & 3, kk));
                for (int c = 0; c < nv; ++c) acc += Hk[v2 + d + nv * c] * uB[c];
            }
            s_out[d + nv * node] += acc;
        }
        __syncthreads();
    }
    double* ye = y + (size_t)e * vn;
    for (int k = tid; k < vn; k += 64) ye[k] += s_out[k];
}

}  // namespace

void launch_advect_face(int n, int nv, const double* u, double* y, const int32_t* nbr, const int32_t* info,
                        const double* H, int nel, int interior, int boundary, hipStream_t s)
{
    if (nel <= 0) return;
    const dim3 grid(nel), block(64);
    switch (n) {
    case 2: hipLaunchKernelGGL(advect_face_kernel<2>, grid, block, 0, s, u, y, nbr, info, H, nv, interior, boundary); break;
    case 3: hipLaunchKernelGGL(advect_face_kernel<3>, grid, block, 0, s, u, y, nbr, info, H, nv, interior, boundary); break;
    case 4: hipLaunchKernelGGL(advect_face_kernel<4>, grid, block, 0, s, u, y, nbr, info, H, nv, interior, boundary); break;
    case 5: hipLaunchKernelGGL(advect_face_kernel<5>, grid, block, 0, s, u, y, nbr, info, H, nv, interior, boundary); break;
    case 6: hipLaunchKernelGGL(advect_face_kernel<6>, grid, block, 0, s, u, y, nbr, info, H, nv, interior, boundary); break;
    case 7: hipLaunchKernelGGL(advect_face_kernel<7>, grid, block, 0, s, u, y, nbr, info, H, nv, interior, boundary); break;
    case 8: hipLaunchKernelGGL(advect_face_kernel<8>, grid, block, 0, s, u, y, nbr, info, H, nv, interior, boundary); break;
    case 9: hipLaunchKernelGGL(advect_face_kernel<9>, grid, block, 0, s, u, y, nbr, info, H, nv, interior, boundary); break;
    default: break;
    }
}

}  // namespace wdg
