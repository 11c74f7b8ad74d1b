// DivF2D<true>::action (DivF.hpp:207-274) on the GPU for the linear flux
// functor f(x, u) = (A u, B u) with constant (n_var, n_var) matrices, or
// matrices given per node (the volume term of Advection<true>,
// Div.cpp:5-53, 131-179, with a variable coefficient): the
// wave equation's volume term is the case A, B = its flux matrices, linear
// advection the case A = a_x I, B = a_y I.  One thread per node; a
// workgroup holds Shape<N>::EPB whole elements.  Each thread evaluates the
// physical flux at its node and the contravariant fluxes
//   F = w_i w_j (y_eta f - x_eta g),  G = w_i w_j (-y_xi f + x_xi g)
// into LDS, then contracts divF(d, i, j) += sum_l D(l, i) F(l, j) + D(l, j) G(i, l)
// (the weak form with D^T, as Div).
#include <hip/hip_runtime.h>

#include "wave2d_kernels.hpp"

namespace wdg {
namespace {

constexpr int kMaxVar = DivFArgs::kMaxVar;

// The compressible Euler flux of examples/cl2d.cpp:57-81 (F): q = (rho, rho u,
// rho v, E), f and g the x- and y-fluxes.
template <typename R>
__device__ __forceinline__ void euler_flux(R gamma, const R* q, R* f, R* g)
{
    const R rho = q[0], u = q[1] / q[0], v = q[2] / q[0], E = q[3];
    const R p = (gamma - (R)1) * (E - (R)0.5 * rho * (u * u + v * v));
    f[0] = rho * u;
    f[1] = rho * u * u + p;
    f[2] = rho * u * v;
    f[3] = u * (E + p);
    g[0] = rho * v;
    g[1] = rho * u * v;
    g[2] = rho * v * v + p;
    g[3] = v * (E + p);
}

// VS >= n_var: the LDS stride between variables, so that few variables leave
// room for more resident workgroups
template <int N, typename R, int VS>
__global__ __launch_bounds__(Shape<N>::NT) void divf_linear_kernel(const DivFArgs a)
{
    constexpr int NN = N * N, EPB = Shape<N>::EPB, NT = Shape<N>::NT;
    __shared__ R s_F[EPB * VS * NN];
    __shared__ R s_G[EPB * VS * NN];
    __shared__ R s_D[NN];
    __shared__ double s_xw[2 * N];
    const int tid = threadIdx.x;
    const int el = tid / NN, node = tid - el * NN, i = node % N, j = node / N;
    const int e = blockIdx.x * EPB + el;
    const bool active = tid < NT && e < a.nel;
    if (tid < NN) s_D[tid] = (R)a.basis[2 * N + tid];
    if (tid < 2 * N) s_xw[tid] = a.basis[tid];
    __syncthreads();
    const int nv = a.nv;
    if (active) {
        const double* gq = a.geom + 6 * (size_t)e;  // ax, ay, cx, cy, hx, hy of the bilinear map
        const double xi = s_xw[i], eta = s_xw[j], wij = s_xw[N + i] * s_xw[N + j];
        const double x_xi = gq[0] + gq[4] * eta, y_xi = gq[1] + gq[5] * eta;
        const double x_eta = gq[2] + gq[4] * xi, y_eta = gq[3] + gq[5] * xi;
        const R Ye = (R)(y_eta * wij), Xe = (R)(x_eta * wij), Yx = (R)(y_xi * wij), Xx = (R)(x_xi * wij);
        const R* uq = static_cast<const R*>(a.u) + (size_t)nv * (node + (size_t)NN * e);
        R uv[kMaxVar];
#pragma unroll
        for (int d = 0; d < kMaxVar; ++d) uv[d] = d < nv ? uq[d] : (R)0;
        const double* An = a.Anode ? a.Anode + (size_t)2 * nv * nv * (node + (size_t)NN * e) : nullptr;
        R ef[4], eg[4];
        if (a.kind == 1) euler_flux((R)a.gamma, uv, ef, eg);
        for (int d = 0; d < nv; ++d) {
            R fx = 0, fy = 0;
            if (a.kind == 1) {
                fx = ef[d & 3];
                fy = eg[d & 3];
            } else if (An) {
                for (int k = 0; k < nv; ++k) {
                    fx = fma((R)An[d + nv * k], uv[k], fx);
                    fy = fma((R)An[nv * nv + d + nv * k], uv[k], fy);
                }
            } else {
#pragma unroll
                for (int k = 0; k < kMaxVar; ++k) {
                    fx = fma((R)a.A[d + kMaxVar * k], uv[k], fx);
                    fy = fma((R)a.B[d + kMaxVar * k], uv[k], fy);
                }
            }
            s_F[(el * VS + d) * NN + j * N + i] = Ye * fx - Xe * fy;
            s_G[(el * VS + d) * NN + i * N + j] = Xx * fy - Yx * fx;
        }
    }
    __syncthreads();
    if (!active) return;
    R* out = static_cast<R*>(a.out) + (size_t)nv * (node + (size_t)NN * e);
    for (int d = 0; d < nv; ++d) {
        const R* Fr = s_F + (el * VS + d) * NN + j * N;
        const R* Gc = s_G + (el * VS + d) * NN + i * N;
        R r = 0;
#pragma unroll
        for (int l = 0; l < N; ++l) {
            r = fma(s_D[l + N * i], Fr[l], r);
            r = fma(s_D[l + N * j], Gc[l], r);
        }
        out[d] += r;
    }
}

template <int N>
void launch_n(int f32, const DivFArgs& a, hipStream_t s)
{
    const dim3 grid((a.nel + Shape<N>::EPB - 1) / Shape<N>::EPB), block(Shape<N>::NT);
#define WDG_DIVF_VS(VS_)                                                                   \
    if (f32)                                                                               \
        hipLaunchKernelGGL((divf_linear_kernel<N, float, VS_>), grid, block, 0, s, a);     \
    else                                                                                   \
        hipLaunchKernelGGL((divf_linear_kernel<N, double, VS_>), grid, block, 0, s, a);
    if (a.nv <= 1) {
        WDG_DIVF_VS(1)
    } else if (a.nv <= 2) {
        WDG_DIVF_VS(2)
    } else if (a.nv <= 4) {
        WDG_DIVF_VS(4)
    } else {
        WDG_DIVF_VS(8)
    }
#undef WDG_DIVF_VS
}

}  // namespace

void launch_divf_linear(int n, int f32, const DivFArgs& a, hipStream_t s)
{
    if (a.nel <= 0) return;
    switch (n) {
    case 2: launch_n<2>(f32, a, s); break;
    case 3: launch_n<3>(f32, a, s); break;
    case 4: launch_n<4>(f32, a, s); break;
    case 5: launch_n<5>(f32, a, s); break;
    case 6: launch_n<6>(f32, a, s); break;
    case 7: launch_n<7>(f32, a, s); break;
    case 8: launch_n<8>(f32, a, s); break;
    case 9: launch_n<9>(f32, a, s); break;
    default: break;
    }
}

}  // namespace wdg
