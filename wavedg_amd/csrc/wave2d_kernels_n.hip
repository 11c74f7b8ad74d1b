// Instantiates the wave kernels for one polynomial order: compiled once per
// N = WDG_N in 2..9 so the eight orders build in parallel.
#include "wave2d_kernels.hpp"

#ifndef WDG_N
#error "compile with -DWDG_N=<n_colloc>"
#endif

namespace wdg {

namespace {

// XM: mode bits added to every kernel (M_LEG for Gauss-Legendre bases,
// M_NT for non-temporal stage output stores).
template <typename R, int TGT, int XM>
void launch_modes(int kid, const KArgs& a, hipStream_t s)
{
    constexpr int N = WDG_N;
    using S = Shape<N, TGT>;
    const dim3 grid((a.nel + S::EPB - 1) / S::EPB), block(S::NT);
    switch (kid) {
#define WDG_CASE(K)                                                                   \
    case K:                                                                           \
        hipLaunchKernelGGL((wave_kernel<N, R, kModeOf[K] | XM, TGT>), grid, block, 0, s, a); \
        break;
        WDG_CASE(K_STAGE0)
        WDG_CASE(K_STAGE0_F)
        WDG_CASE(K_STAGE)
        WDG_CASE(K_STAGE_F)
        WDG_CASE(K_STAGE_ACC)
        WDG_CASE(K_STAGE_F_ACC)
#undef WDG_CASE
    default: break;
    }
}

// The RK stage kernels of one schedule; handles whose stage arrays exceed the
// Infinity Cache (KArgs::nt_store) take the non-temporal-store variants.
template <typename R, int TGT, int XM = 0>
void launch_sched(int kid, const KArgs& a, hipStream_t s)
{
    if (a.nt_store)
        launch_modes<R, TGT, XM | M_NT>(kid, a, s);
    else
        launch_modes<R, TGT, XM>(kid, a, s);
}

// Gauss-Legendre bases: the RK stage kernels run with the schedule picked
// for the handle, the other entry points with the default schedule.
template <typename R>
void launch_legendre(int kid, int sched, const KArgs& a, hipStream_t s)
{
    constexpr int N = WDG_N;
    switch (kid) {
    case K_STAGE0:
    case K_STAGE0_F:
    case K_STAGE:
    case K_STAGE_F:
    case K_STAGE_ACC:
    case K_STAGE_F_ACC:
        switch (sched) {
        case 1: launch_sched<R, kSchedTgt[1], M_LEG>(kid, a, s); return;
        case 2: launch_sched<R, kSchedTgt[2], M_LEG>(kid, a, s); return;
        case 3: launch_sched<R, kSchedTgt[3], M_LEG>(kid, a, s); return;
        default: launch_sched<R, kSchedTgt[0], M_LEG>(kid, a, s); return;
        }
    default: break;
    }
    const int nblk = (a.nel + Shape<N>::EPB - 1) / Shape<N>::EPB;
    const dim3 grid(nblk), block(Shape<N>::NT);
    switch (kid) {
#define WDG_CASE(K)                                                                      \
    case K:                                                                              \
        hipLaunchKernelGGL((wave_kernel<N, R, kModeOf[K] | M_LEG>), grid, block, 0, s, a); \
        break;
        WDG_CASE(K_ACTION)
        WDG_CASE(K_BC)
        WDG_CASE(K_RHS)
        WDG_CASE(K_RHS_F)
#undef WDG_CASE
    default: break;
    }
}

template <typename R>
void launch_typed(int kid, int sched, const KArgs& a, hipStream_t s)
{
    constexpr int N = WDG_N;
    if (a.leg) {
        launch_legendre<R>(kid, sched, a, s);
        return;
    }
    switch (kid) {
    case K_STAGE0:
    case K_STAGE0_F:
    case K_STAGE:
    case K_STAGE_F:
    case K_STAGE_ACC:
    case K_STAGE_F_ACC:
        // the RK stage kernels run with the schedule picked for this order
        switch (sched) {
        case 1: launch_sched<R, kSchedTgt[1]>(kid, a, s); return;
        case 2: launch_sched<R, kSchedTgt[2]>(kid, a, s); return;
        case 3: launch_sched<R, kSchedTgt[3]>(kid, a, s); return;
        default: launch_sched<R, kSchedTgt[0]>(kid, a, s); return;
        }
    default: break;
    }
    const int nblk = (a.nel + Shape<N>::EPB - 1) / Shape<N>::EPB;
    const dim3 grid(nblk), block(Shape<N>::NT);
    switch (kid) {
#define WDG_CASE(K)                                                              \
    case K:                                                                      \
        hipLaunchKernelGGL((wave_kernel<N, R, kModeOf[K]>), grid, block, 0, s, a); \
        break;
        WDG_CASE(K_ACTION)
        WDG_CASE(K_BC)
        WDG_CASE(K_RHS)
        WDG_CASE(K_RHS_F)
#undef WDG_CASE
    default: break;
    }
}

}  // namespace

#define WDG_CAT2(a, b) a##b
#define WDG_CAT(a, b) WDG_CAT2(a, b)

void WDG_CAT(launch_wave_n, WDG_N)(int kid, int f32, int sched, const KArgs& a, hipStream_t s)
{
    if (a.nel <= 0) return;
    if (f32) {
        // the FP32 kernels take their scalars as float kernel arguments
        KArgs b = a;
        b.ha32 = (float)a.ha;
        b.hb32 = (float)a.hb;
        b.dt32 = (float)a.dt;
        b.rk_a32 = (float)a.rk_a;
        b.rk_b32 = (float)a.rk_b;
        b.fs32 = (float)a.fs;
        launch_typed<float>(kid, sched, b, s);
    } else
        launch_typed<double>(kid, sched, a, s);
}

void WDG_CAT(launch_load_vector_n, WDG_N)(const double* corners, const double* tab, const double* src, int nsrc,
                                           int var, double* F, int nel, hipStream_t s)
{
    constexpr int N = WDG_N;
    if (nel <= 0) return;
    const int nblk = (nel + Shape<N>::EPB - 1) / Shape<N>::EPB;
    hipLaunchKernelGGL((load_vector_kernel<N>), dim3(nblk), dim3(Shape<N>::NT), 0, s, corners, tab, src, nsrc, var,
                       F, nel);
}

}  // namespace wdg
