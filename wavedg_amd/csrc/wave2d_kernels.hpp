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

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>
#include <type_traits>

#include "host_faces.hpp"

namespace wdg {

struct KArgs {
    const void* u_in;       // state read by this element and its neighbours
    void* u_out;            // y (action/bc/rhs) or the next stage state
    void* du;               // RK register (stage modes)
    const double* geom;     // (6, nel): ax, ay, cx, cy, hx, hy of the bilinear map
    const int32_t* nbr;     // (4, nel)
    const int32_t* info;    // (4, nel), see host_mesh.hpp
    const double* basis;    // x[N], w[N], D[N*N] (column-major), P[N] = phi_k(-1)
    const void* force;      // load vector (rhs/stage with M_FORCE)
    double fs;              // forcing scale
    double* acc;            // WaveHoltz accumulator (M_ACC)
    double ac;              // accumulation weight
    double rk_a, rk_b, dt;  // stage coefficients
    double ha, hb;          // flux: f = w ds [ha nA (u0+u1) + hb |nA| (u0-u1)]
    int nel;
    const float* geom32;    // FP32 handles: geom and basis rounded to float once
    const float* basis32;
    int leg;                // 1: Gauss-Legendre basis (faces interpolate with P)
    const void* fgeo;       // (4, nel) FaceGeom<R>: side normal and half length (build_face_tables)
    const FaceNb* fnb;      // (4, nel) neighbour trace base and flags (build_face_tables)
    float ha32, hb32, dt32, rk_a32, rk_b32, fs32;  // the scalars above rounded to float (FP32 handles)
    uint32_t sr_seed;       // FP32 stages: seed of the state update's stochastic rounding (sr_seed())
    int nt_store;           // 1: stage outputs as non-temporal stores (arrays beyond the Infinity Cache)
};

// Stochastic rounding of the FP32 state update u + b du.  Rounding to nearest
// loses the same fraction of an ulp stage after stage when b du is small and
// slowly varying next to u, so the FP32 error grows with the number of
// stages; rounding up with probability equal to the discarded fraction makes
// each rounding error zero on average.  The sum of two floats is exact in
// double (to within 2^-29 of the smaller one); adding a uniform 29-bit integer
// to the 29 mantissa bits a float drops, then clearing them, rounds the
// magnitude up with exactly that probability (for either sign).
__host__ __device__ inline uint32_t sr_seed(int step, int stage)
{
    return (uint32_t)(step * 5 + stage + 1) * 0x9E3779B9u;
}
__device__ __forceinline__ uint32_t sr_hash(uint32_t idx, uint32_t seed)
{
    uint32_t h = (idx ^ seed) * 0x85EBCA6Bu;  // MurmurHash3 finaliser
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    return h ^ (h >> 16);
}
__device__ __forceinline__ float add_sr(float u, float bd, uint32_t r29)
{
    unsigned long long b = (unsigned long long)__double_as_longlong((double)u + (double)bd);
    b = (b + r29) & ~0x1FFFFFFFull;
    return (float)__longlong_as_double((long long)b);
}
// u + bd for the three variables of a node: stochastic rounding in float
// (random bits from the node index and seed), plain arithmetic in double.
template <typename R>
__device__ __forceinline__ void stage_update(R p, R uu, R vv, R bd0, R bd1, R bd2, uint32_t idx, uint32_t seed,
                                             R& w0, R& w1, R& w2)
{
    if constexpr (std::is_same<R, float>::value) {
        const uint32_t h = sr_hash(idx, seed);
        w0 = add_sr(p, bd0, h >> 3);
        w1 = add_sr(uu, bd1, (h * 0x9E3779B1u) >> 3);
        w2 = add_sr(vv, bd2, (h * 0x27D4EB2Fu) >> 3);
    } else {
        w0 = p + bd0;
        w1 = uu + bd1;
        w2 = vv + bd2;
    }
}


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
    M_LEG = 1 << 10,  // Gauss-Legendre basis
    M_NT = 1 << 11,   // stage outputs as non-temporal stores (KArgs::nt_store)
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

#ifndef WDG_MIN_WAVES
#define WDG_MIN_WAVES 1  // occupancy hint: minimum waves per SIMD
#endif
#ifndef WDG_MIN_WAVES_F32
#define WDG_MIN_WAVES_F32 WDG_MIN_WAVES  // the same for the FP32 stage kernels
#endif
// Stage outputs u' and du, which no workgroup of the same launch reads.
// When the stage's arrays are larger than the 256 MB Infinity Cache the
// M_NT instantiations send them as non-temporal stores, streaming to HBM
// without displacing the neighbour traces (2-5 % faster at 512^2, P = 7);
// arrays that fit stay cached for the next stage (non-temporal stores there
// cost up to 25 %).
template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T* p, T v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

// Workgroup schedule: about TGT threads, i.e. EPB = TGT / n^2 whole elements
// (at least one).  TGT = 64 at n = 8 is one wavefront per element; smaller
// orders pack several elements into each wavefront.
template <int N, int TGT = 256>
struct Shape {
    static constexpr int NN = N * N;
    static constexpr int EPB = (TGT / NN) > 0 ? (TGT / NN) : 1;  // elements per workgroup
    static constexpr int NT = EPB * NN;                          // threads per workgroup
};

// Schedules the stage kernels can be launched with (index -> TGT).
constexpr int kSchedules = 4;
constexpr int kSchedTgt[kSchedules] = {256, 512, 128, 64};

// Workgroups are dealt to the 8 XCDs round-robin (b -> XCD b mod 8).  Give
// each XCD one contiguous run of elements so that an element's neighbours
// (+-1, +-row in structured meshes) were loaded through the same L2.
__device__ __forceinline__ int xcd_block(int b, int nblk)
{
    const int x = b & 7, q = nblk >> 3, r = nblk & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// 1/x to full double precision: hardware reciprocal estimate and two Newton
// steps (x > 0 here: it is a mass-matrix entry).
__device__ __forceinline__ double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// 1/x in single precision: reciprocal estimate and one Newton step.
__device__ __forceinline__ float recip(float x)
{
    float r = __builtin_amdgcn_rcpf(x);
    return fmaf(r, fmaf(-x, r, 1.0f), r);
}

// 1/sqrt(x) from the hardware estimate and Newton steps (two in double,
// one in single precision).
__device__ __forceinline__ double rsqrt_nr(double x)
{
    double r = __builtin_amdgcn_rsq(x);
    r = r * (1.5 - 0.5 * x * r * r);
    return r * (1.5 - 0.5 * x * r * r);
}
__device__ __forceinline__ float rsqrt_nr(float x)
{
    const float r = __builtin_amdgcn_rsqf(x);
    return r * (1.5f - 0.5f * x * r * r);
}

// Node (k along side s, counted from the side's first corner) -> flat node.
template <int N>
__device__ __forceinline__ int side_node(int s, int k)
{
    // sides 0 and 2 run along i (step 1), sides 1 and 3 along j (step N),
    // starting at node 0 (sides 0, 3), N-1 (side 1) or N(N-1) (side 2)
    const int start = s == 1 ? N - 1 : (s == 2 ? N * (N - 1) : 0);
    return start + k * ((s & 1) ? N : 1);
}

// Step from a node of side s one row into the element: node m of the line
// through side position k is side_node(s, k) + m * inward_stride(s).  With a
// Gauss-Legendre basis the trace at side position k is
// sum_m P(m) u(side_node(s, k) + m * inward_stride(s)), P(m) = phi_m(-1)
// (mapV2E and legendre_action_2d, FaceProlongator.cpp:717-769, 809-835).
template <int N>
__device__ __forceinline__ int inward_stride(int s)
{
    return s == 0 ? N : s == 1 ? -1 : s == 2 ? -N : 1;
}

// LDS carve-up of apply_operator for one workgroup of Shape<N, TGT> elements.
template <int N, typename R, bool VOL, bool FACE, int TGT = 256, bool LEG = false>
struct OpLds {
    static constexpr int NN = N * N, EPB = Shape<N, TGT>::EPB, NF = 4 * N;
    static constexpr int NXW = LEG ? 3 * N : 2 * N;  // x[N], w[N] (, P[N])
    static constexpr size_t n_u = (size_t)EPB * 3 * NN;
    static constexpr size_t n_F = VOL ? n_u : 0, n_G = VOL ? n_u : 0;
    static constexpr size_t n_S = FACE ? (size_t)EPB * (NF + 1) * 3 : 0;  // slot NF of an element: zeros
    static constexpr size_t off_xw = 0;                                  // double x[N], w[N] (, P[N])
    static constexpr size_t off_u = ((NXW * 8 + 15) / 16) * 16;
    static constexpr size_t off_F = off_u + n_u * sizeof(R);
    static constexpr size_t off_G = off_F + n_F * sizeof(R);
    static constexpr size_t off_S = off_G + n_G * sizeof(R);
    static constexpr size_t off_D = ((off_S + n_S * sizeof(R) + 15) / 16) * 16;  // 16-byte rows for wide reads
    static constexpr size_t bytes = ((off_D + NN * sizeof(R) + 15) / 16) * 16;
};

// Result of the operator at one node: own state, the weak-form right-hand
// side r = (A + B) u before M^{-1}, and the mass-matrix entry w_i w_j det J.
template <typename R>
struct NodeResult {
    R p, uu, vv;
    R r0, r1, r2;
    R m;
};

// The operator for one workgroup of elements [e0, e0 + nloc) of one mesh,
// run by threads 0 .. Shape<N, TGT>::NT-1 of the workgroup (threads beyond only
// join the barriers).  Element indices in nbr are local to u_in / geom.
// The node's values are at [gb + g, gb + g + 3) of every state-shaped array:
// gb is the workgroup's first element (uniform, 64-bit), g the node's 32-bit
// offset from it.  `prefetch(active, gb, g)` is called after the state loads
// are issued and
// before the first barrier, so the caller's extra loads overlap them.
// GT and BT, the types of the geometry and basis tables, are double, or
// float for FP32 handles that keep float copies (the rounding is the same as
// converting in the kernel, so results do not change).  LEG: Gauss-Legendre
// basis, whose side traces interpolate along the element (inward_stride) and
// whose face contributions lift back with the same weights P.
template <int N, typename R, bool VOL, bool IFACE, bool BFACE, int TGT = 256, bool LEG = false, typename GT = double,
          typename BT = double, typename Prefetch>
__device__ __forceinline__ NodeResult<R> apply_operator(const R* __restrict__ u_in, const GT* __restrict__ geom,
                                                        const FaceGeom<R>* __restrict__ fgeo,
                                                        const FaceNb* __restrict__ fnb,
                                                        const BT* __restrict__ basis, R ha, R hb,
                                                        int e0, int nloc, unsigned char* lds, int& el_out,
                                                        size_t& gb_out, uint32_t& g_out, bool& active_out,
                                                        Prefetch&& prefetch)
{
    constexpr int NN = N * N;
    constexpr int NT = Shape<N, TGT>::NT;
    constexpr int NF = 4 * N;               // face nodes per element
    constexpr int FP = (NF + NN - 1) / NN;  // face nodes per thread (1 for N >= 4)
    constexpr bool FACE = IFACE || BFACE;
    using L = OpLds<N, R, VOL, FACE, TGT, LEG>;
    BT* s_xw = reinterpret_cast<BT*>(lds + L::off_xw);
    R* s_u = reinterpret_cast<R*>(lds + L::off_u);
    R* s_F = reinterpret_cast<R*>(lds + L::off_F);
    R* s_G = reinterpret_cast<R*>(lds + L::off_G);
    R* s_S = reinterpret_cast<R*>(lds + L::off_S);
    R* s_D = reinterpret_cast<R*>(lds + L::off_D);
    const BT* s_x = s_xw;
    const BT* s_w = s_xw + N;
    const BT* s_P = s_xw + 2 * N;  // LEG only

    const int tid = threadIdx.x;
    const bool worker = tid < NT;

    // ---- prologue: issue every global load before the first LDS write so
    // that they are all in flight together (clamped indices keep the
    // unconditional loads of a partial last block in bounds).
    const int cnt = nloc * 3 * NN;
    R ld[3];
    {
        const R* ub = u_in + (size_t)e0 * 3 * NN;
#pragma unroll
        for (int r = 0; r < 3; ++r) ld[r] = ub[min(tid + r * NT, cnt - 1)];
    }
    const R dval = (R)basis[2 * N + min(tid, NN - 1)];
    const BT xwv = (LEG && tid >= 2 * N) ? basis[NN + min(tid, 3 * N - 1)] : basis[min(tid, 2 * N - 1)];
    const int el = worker ? tid / NN : 0;
    const int node = worker ? tid - el * NN : 0;
    const int i = node % N, j = node / N;
    const bool active = worker && el < nloc;
    int e = e0 + (active ? el : 0);
    if constexpr (NN % 64 == 0) e = __builtin_amdgcn_readfirstlane(e);  // one element per wave
    const size_t gb = (size_t)e0 * NN * 3;
    const uint32_t g = (uint32_t)(((active ? el : 0) * NN + (active ? node : 0)) * 3);
    el_out = el;
    gb_out = gb;
    g_out = g;
    active_out = active;

    // Bilinear element map x(xi, eta) = x0 + a xi + c eta + h xi eta (per
    // coordinate), so dx/dxi = a + h eta, dx/deta = c + h xi (Element.cpp:21-27).
    // Volume metrics are evaluated in the state precision R (double for FP64
    // handles); face normals and half lengths come from the fgeo table.
    using GT2 = typename std::conditional<sizeof(GT) == 8, double2, float2>::type;
    const GT2* gp = reinterpret_cast<const GT2*>(geom + 6 * (size_t)e);
    const GT2 g_a = gp[0], g_c = gp[1], g_h = gp[2];
    const R ax = (R)g_a.x, ay = (R)g_a.y, cx = (R)g_c.x, cy = (R)g_c.y, hx = (R)g_h.x, hy = (R)g_h.y;

    // Face lanes: face node f = node + p NN (side f / N, position f % N),
    // with the side's table entries and the neighbour's trace.
    R nbv[FP][3];
    int finfo[FP], fnb3[FP];
    // FP32 kernels of orders N >= 6 read each side's normal and half length
    // from the fgeo table.  Lower orders, with more elements per DoF, and the
    // FP64 kernels, which are bound by memory rather than by VALU issue,
    // compute them from the map coefficients instead of spending HBM bytes
    // on the table (3 % faster at P = 7 in FP64).
    constexpr bool FTAB = N >= 6 && sizeof(R) == 4;
    FaceGeom<R> fgv[FP];
    if constexpr (FACE) {
#pragma unroll
        for (int p = 0; p < FP; ++p) {
            const int f = node + p * NN;
            const bool fon = active && f < NF;
            const int s = fon ? f / N : 0, k = fon ? f - (f / N) * N : 0;
            const FaceNb fx = fnb[4 * (size_t)e + s];
            if constexpr (FTAB) fgv[p] = fgeo[4 * (size_t)e + s];
            fnb3[p] = fx.x;
            finfo[p] = fx.y;
            nbv[p][0] = nbv[p][1] = nbv[p][2] = (R)0;
            if (IFACE && fon && fx.x >= 0) {
                const R* q = u_in + (fx.x + k * (fx.y >> 16));
                if constexpr (LEG) {
                    const int st = 3 * inward_stride<N>(fx.y & 3);
#pragma unroll
                    for (int m = 0; m < N; ++m) {
                        const R Pm = (R)basis[2 * N + NN + m];
                        nbv[p][0] = fma(Pm, q[m * st + 0], nbv[p][0]);
                        nbv[p][1] = fma(Pm, q[m * st + 1], nbv[p][1]);
                        nbv[p][2] = fma(Pm, q[m * st + 2], nbv[p][2]);
                    }
                } else {
                    nbv[p][0] = q[0];
                    nbv[p][1] = q[1];
                    nbv[p][2] = q[2];
                }
            }
        }
    }
    prefetch(active, gb, g);

    if (worker) {
#pragma unroll
        for (int r = 0; r < 3; ++r) s_u[tid + r * NT] = ld[r];
        if (tid < NN) s_D[tid] = dval;
        if (tid < L::NXW) s_xw[tid] = xwv;
    }
    __syncthreads();

    const R xi = (R)s_x[i], eta = (R)s_x[j];
    const R J0 = ax + hx * eta, J1 = ay + hy * eta, J2 = cx + hx * xi, J3 = cy + hy * xi;
    const R Wt = (R)s_w[j] * (R)s_w[i];

    NodeResult<R> out;
    out.p = s_u[el * 3 * NN + 3 * node + 0];
    out.uu = s_u[el * 3 * NN + 3 * node + 1];
    out.vv = s_u[el * 3 * NN + 3 * node + 2];
    out.m = Wt * (J0 * J3 - J1 * J2);  // MassMatrix<true>: w_i w_j det J (MassMatrix.cpp:6-28)
    const R p = out.p, uu = out.uu, vv = out.vv;

    if constexpr (VOL) {
        // Contravariant fluxes F = W (y_eta f - x_eta g), G = W (-y_xi f + x_xi g),
        // f = (u, p, 0), g = (v, 0, p) (Div.cpp:5-53).
        if (worker) {
            const R Ye = J3 * Wt, Xe = J2 * Wt, Yx = J1 * Wt, Xx = J0 * Wt;
            R* F = s_F + el * 3 * NN;
            R* G = s_G + el * 3 * NN;
            F[0 * NN + j * N + i] = Ye * uu - Xe * vv;
            F[1 * NN + j * N + i] = Ye * p;
            F[2 * NN + j * N + i] = -(Xe * p);
            G[0 * NN + i * N + j] = Xx * vv - Yx * uu;
            G[1 * NN + i * N + j] = -(Yx * p);
            G[2 * NN + i * N + j] = Xx * p;
        }
    }

    if constexpr (FACE) {
        // One lane per face node: closed form of EdgeFlux<true> for the wave
        // equation (EdgeFlux.cpp:31-139, 289-325; |n.A| = (n.A)^2 for unit n)
        // lifted into this element (lobatto_transpose_2d,
        // FaceProlongator.cpp:650-672).  Written with this element's outward
        // normal n (= -n of el0 on the el1 side, Edge.hpp:107-122), own trace
        // A and the other trace B, the contribution is
        //   c_p = w ds [ha n.(uA+uB) + hb (pA-pB)],  c_u = n w ds [ha (pA+pB) + hb n.(uA-uB)]
        // which equals +f on el0 and -f on el1 for the reference's edge flux f.
#pragma unroll
        for (int pp = 0; pp < FP; ++pp) {
            const int f = node + pp * NN;
            if (!(active && f < NF)) continue;
            const int s = f / N, k = f - s * N;
            const int inf = finfo[pp];
            const bool interior = fnb3[pp] >= 0;
            R c0 = 0, c1 = 0, c2 = 0;
            if (interior ? IFACE : BFACE) {
                R n0, n1, hl;
                if constexpr (FTAB) {
                    n0 = fgv[pp].n0;
                    n1 = fgv[pp].n1;
                    hl = fgv[pp].hl;
                } else {
                    // side s runs along 2(a -+ h) (s = 0, 2) or 2(c +- h) (s = 1, 3),
                    // sides 2 and 3 clockwise
                    const R sh = (s == 0 || s == 3) ? (R)-1 : (R)1;
                    const R dx = (R)2 * ((s & 1) ? cx + sh * hx : ax + sh * hx);
                    const R dy = (R)2 * ((s & 1) ? cy + sh * hy : ay + sh * hy);
                    const R l2 = dx * dx + dy * dy;
                    const R il = rsqrt_nr(l2);  // 1/len
                    const R sil = (s >= 2) ? -il : il;
                    n0 = dy * sil;
                    n1 = -dx * sil;
                    hl = (R)0.5 * (l2 * il);
                }
                const int ie = ((inf >> 6) & 1) ? (N - 1 - k) : k;  // node index in el0's order
                const R wds = (R)s_w[ie] * hl;
                const int ln = side_node<N>(s, k);
                R pA, uA, vA;
                if constexpr (LEG) {
                    pA = uA = vA = (R)0;
                    const int st = 3 * inward_stride<N>(s);
                    const R* q = s_u + el * 3 * NN + 3 * ln;
#pragma unroll
                    for (int m = 0; m < N; ++m) {
                        const R Pm = (R)s_P[m];
                        pA = fma(Pm, q[m * st + 0], pA);
                        uA = fma(Pm, q[m * st + 1], uA);
                        vA = fma(Pm, q[m * st + 2], vA);
                    }
                } else {
                    pA = s_u[el * 3 * NN + 3 * ln + 0];
                    uA = s_u[el * 3 * NN + 3 * ln + 1];
                    vA = s_u[el * 3 * NN + 3 * ln + 2];
                }
                R pB, uB, vB;
                if (interior) {
                    pB = nbv[pp][0];
                    uB = nbv[pp][1];
                    vB = nbv[pp][2];
                } else {
                    const int code = (inf >> 8) & 0xff;
                    if (code == 1) {  // reflecting (WaveEquation.hpp:64-69)
                        pB = pA;
                        uB = (n1 * n1 - n0 * n0) * uA - (R)2 * n0 * n1 * vA;
                        vB = -(R)2 * n0 * n1 * uA + (n0 * n0 - n1 * n1) * vA;
                    } else if (code == 0) {  // characteristic outflow, p_ext after u_ext (WaveEquation.hpp:77-84)
                        const R o = pA + n0 * uA + n1 * vA;
                        const R tang = n1 * uA - n0 * vA;
                        uB = n1 * tang + (R)0.5 * n0 * o;
                        vB = -n0 * tang + (R)0.5 * n1 * o;
                        pB = n0 * uB + n1 * vB;
                    } else {  // Dirichlet p = 0 (mirror state)
                        pB = -pA;
                        uB = uA;
                        vB = vA;
                    }
                }
                c0 = wds * (ha * (n0 * (uA + uB) + n1 * (vA + vB)) + hb * (pA - pB));
                const R cf = wds * (ha * (pA + pB) + hb * (n0 * (uA - uB) + n1 * (vA - vB)));
                c1 = n0 * cf;
                c2 = n1 * cf;
            }
            R* S = s_S + (el * (NF + 1) + f) * 3;
            S[0] = c0;
            S[1] = c1;
            S[2] = c2;
        }
        if (active && node < 3) s_S[(el * (NF + 1) + NF) * 3 + node] = (R)0;
    }
    __syncthreads();

    R r0 = 0, r1 = 0, r2 = 0;
    if constexpr (VOL) {
        // du(i,j) += sum_l D(l,i) F(l,j) + D(l,j) G(i,l) (Div.cpp:131-179)
        const R* Fr = s_F + el * 3 * NN + j * N;
        const R* Gc = s_G + el * 3 * NN + i * N;
#pragma unroll
        for (int l = 0; l < N; ++l) {
            const R di = s_D[l + N * i], dj = s_D[l + N * j];
            r0 = fma(di, Fr[l], r0);
            r1 = fma(di, Fr[NN + l], r1);
            r2 = fma(di, Fr[2 * NN + l], r2);
            r0 = fma(dj, Gc[l], r0);
            r1 = fma(dj, Gc[NN + l], r1);
            r2 = fma(dj, Gc[2 * NN + l], r2);
        }
    }
    if constexpr (FACE) {
        const R* S = s_S + el * (NF + 1) * 3;
        if constexpr (LEG) {
            // node (i, j) is row j of side 0, row N-1-i of side 1, row N-1-j
            // of side 2 and row i of side 3 (legendre_transpose_2d,
            // FaceProlongator.cpp:908-934)
            const R q0 = (R)s_P[j], q1 = (R)s_P[N - 1 - i], q2 = (R)s_P[N - 1 - j], q3 = (R)s_P[i];
            const R* S0 = S + (0 * N + i) * 3;
            const R* S1 = S + (1 * N + j) * 3;
            const R* S2 = S + (2 * N + i) * 3;
            const R* S3 = S + (3 * N + j) * 3;
            r0 += q0 * S0[0] + q1 * S1[0] + q2 * S2[0] + q3 * S3[0];
            r1 += q0 * S0[1] + q1 * S1[1] + q2 * S2[1] + q3 * S3[1];
            r2 += q0 * S0[2] + q1 * S1[2] + q2 * S2[2] + q3 * S3[2];
        } else {
            // node (i, j) lies on at most one side of each direction: side 0
            // (j = 0) or 2 (j = N-1), and side 3 (i = 0) or 1 (i = N-1); other
            // nodes read the zero slot NF
            const int fa = j == 0 ? i : (j == N - 1 ? 2 * N + i : NF);
            const int fb = i == N - 1 ? N + j : (i == 0 ? 3 * N + j : NF);
            const R* Sa = S + fa * 3;
            const R* Sb = S + fb * 3;
            r0 += Sa[0] + Sb[0];
            r1 += Sa[1] + Sb[1];
            r2 += Sa[2] + Sb[2];
        }
    }
    out.r0 = r0;
    out.r1 = r1;
    out.r2 = r2;
    return out;
}

template <int N, typename R, int MODE, int TGT = 256>
__global__ __launch_bounds__((Shape<N, TGT>::NT), (sizeof(R) == 4 ? WDG_MIN_WAVES_F32 : WDG_MIN_WAVES)) void wave_kernel(
    const KArgs a)
{
    constexpr int EPB = Shape<N, TGT>::EPB;
    constexpr bool VOL = (MODE & M_VOL) != 0;
    constexpr bool IFACE = (MODE & M_IF) != 0;
    constexpr bool BFACE = (MODE & M_BF) != 0;
    constexpr int OUT = MODE & M_OUTMASK;
    constexpr bool FORCE = (MODE & M_FORCE) != 0;
    constexpr bool ACC = (MODE & M_ACC) != 0;
    constexpr bool READ_DU = OUT == M_OUT_STAGE;
    constexpr bool LEG = (MODE & M_LEG) != 0;
    constexpr bool NT = (MODE & M_NT) != 0;
    using L = OpLds<N, R, VOL, IFACE || BFACE, TGT, LEG>;
    __shared__ __align__(16) unsigned char lds[L::bytes];

    const int e0 = xcd_block(blockIdx.x, gridDim.x) * EPB;
    const int nloc = min(EPB, a.nel - e0);
    if constexpr (BFACE && !VOL && !IFACE) {
        // boundary terms alone (WaveBC::action): a workgroup whose elements
        // have no boundary side adds nothing, so it leaves y untouched
        const int t = threadIdx.x;
        const bool bnd = t < 4 * nloc && a.nbr[4 * (size_t)e0 + t] < 0;
        if (!__syncthreads_or(bnd)) return;
    }

    // RK register / output / forcing / accumulator values, loaded early
    R pre0 = 0, pre1 = 0, pre2 = 0, fv0 = 0, fv1 = 0, fv2 = 0;
    double ac0 = 0, ac1 = 0, ac2 = 0;
    // per-lane pointers are formed once (uniform base + zero-extended 32-bit
    // offset), so the three variables sit at immediate offsets
    auto prefetch = [&](bool act, size_t gb, uint32_t g) {
        if (!act) return;
        const size_t o = gb + (size_t)g;
        if constexpr (READ_DU) {
            const R* du = static_cast<const R*>(a.du) + o;
            pre0 = du[0];
            pre1 = du[1];
            pre2 = du[2];
        } else if constexpr (OUT == M_OUT_ACCUM) {
            const R* y = static_cast<const R*>(a.u_out) + o;
            pre0 = y[0];
            pre1 = y[1];
            pre2 = y[2];
        }
        if constexpr (FORCE) {
            const R* Fv = static_cast<const R*>(a.force) + o;
            fv0 = Fv[0];
            fv1 = Fv[1];
            fv2 = Fv[2];
        }
        if constexpr (ACC) {
            const double* ac = a.acc + o;
            ac0 = ac[0];
            ac1 = ac[1];
            ac2 = ac[2];
        }
    };
    int el;
    size_t gb;
    uint32_t g;
    bool active;
    NodeResult<R> q;
    constexpr bool F32 = std::is_same<R, float>::value;
    const FaceGeom<R>* fgeo = static_cast<const FaceGeom<R>*>(a.fgeo);
    if constexpr (F32)
        q = apply_operator<N, R, VOL, IFACE, BFACE, TGT, LEG>(static_cast<const R*>(a.u_in), a.geom32, fgeo, a.fnb,
                                                         a.basis32, a.ha32, a.hb32, e0, nloc, lds, el, gb, g, active,
                                                         prefetch);
    else
        q = apply_operator<N, R, VOL, IFACE, BFACE, TGT, LEG>(static_cast<const R*>(a.u_in), a.geom, fgeo, a.fnb,
                                                         a.basis, (R)a.ha, (R)a.hb, e0, nloc, lds, el, gb, g, active,
                                                         prefetch);
    // FP64 stage outputs with streaming stores at orders whose elements do not
    // fill whole wavefronts: the workgroup's elements are one contiguous block
    // per array, staged in the workgroup's state tile and stored as 16-byte
    // pieces (EPB 3 n^2 a multiple of the values per piece keeps every
    // workgroup block 16-byte aligned).  Every thread of the workgroup takes
    // part in its barriers.
    constexpr int VPR = 16 / (int)sizeof(R);  // values per 16-byte piece
    constexpr bool WG_BLOCKS = NT && !F32 && N != 8 && (OUT == M_OUT_STAGE0 || OUT == M_OUT_STAGE) &&
                               (Shape<N, TGT>::EPB * 3 * N * N) % VPR == 0;
    if constexpr (!WG_BLOCKS) {
        if (!active) return;
    }
    const size_t o = gb + (size_t)g;
    R r0 = q.r0, r1 = q.r1, r2 = q.r2;
    if constexpr (OUT == M_OUT_ACCUM) {
        R* y = static_cast<R*>(a.u_out) + o;
        y[0] = pre0 + r0;
        y[1] = pre1 + r1;
        y[2] = pre2 + r2;
    } else {
        const R im = recip(q.m);
        if constexpr (FORCE) {
            const R fs = F32 ? (R)a.fs32 : (R)a.fs;
            r0 += fv0 * fs;
            r1 += fv1 * fs;
            r2 += fv2 * fs;
        }
        r0 *= im;
        r1 *= im;
        r2 *= im;
        if constexpr (OUT == M_OUT_RHS) {
            R* y = static_cast<R*>(a.u_out) + o;
            y[0] = r0;
            y[1] = r1;
            y[2] = r2;
        } else {
            // low-storage RK stage (ode.hpp:74-90)
            R* du = static_cast<R*>(a.du) + o;
            R* uo = static_cast<R*>(a.u_out) + o;
            const R dt = F32 ? (R)a.dt32 : (R)a.dt, rb = F32 ? (R)a.rk_b32 : (R)a.rk_b;
            R d0, d1, d2;
            if constexpr (OUT == M_OUT_STAGE0) {
                d0 = dt * r0;
                d1 = dt * r1;
                d2 = dt * r2;
            } else {
                const R ra = F32 ? (R)a.rk_a32 : (R)a.rk_a;
                d0 = ra * pre0 + dt * r0;
                d1 = ra * pre1 + dt * r1;
                d2 = ra * pre2 + dt * r2;
            }
            R w0, w1, w2;
            typedef R piece_t __attribute__((ext_vector_type(VPR)));
            if constexpr (NT && N == 8 && !F32) {
                // One element per wavefront: its 64 nodes x 3 doubles are one
                // contiguous 1536-byte block per array.  Strided per-lane
                // non-temporal stores leave partial lines that reach HBM as
                // extra write traffic, so the wave stages each block in its
                // element's state tile (free once the operator has returned:
                // every read of it precedes the operator's last barrier) and
                // stores it as contiguous 16-byte pieces.  LDS accesses of one
                // wavefront complete in order; the wave barrier keeps the
                // compiler from reordering them.  (FP32 lanes already store
                // their 12 contiguous bytes as one access; staging them
                // measured slower.)
                constexpr int PIECES = 3 * 64 / VPR;
                const int lane = threadIdx.x & 63;
                R* tile = reinterpret_cast<R*>(lds + L::off_u) + (size_t)el * 3 * L::NN;
                auto store_block = [&](R* dst, R v0, R v1, R v2) {
                    tile[3 * lane + 0] = v0;
                    tile[3 * lane + 1] = v1;
                    tile[3 * lane + 2] = v2;
                    __builtin_amdgcn_wave_barrier();
                    const piece_t* t2 = reinterpret_cast<const piece_t*>(tile);
                    piece_t* d2 = reinterpret_cast<piece_t*>(dst);
#pragma unroll
                    for (int k = lane; k < PIECES; k += 64) __builtin_nontemporal_store(t2[k], d2 + k);
                    __builtin_amdgcn_wave_barrier();
                };
                const size_t eb = gb + (size_t)el * 3 * L::NN;  // the element's first value
                store_block(static_cast<R*>(a.du) + eb, d0, d1, d2);
                stage_update(q.p, q.uu, q.vv, rb * d0, rb * d1, rb * d2, (uint32_t)gb + g, a.sr_seed, w0, w1, w2);
                store_block(static_cast<R*>(a.u_out) + eb, w0, w1, w2);
            } else if constexpr (WG_BLOCKS) {
                R* tile = reinterpret_cast<R*>(lds + L::off_u);  // the workgroup's block, node-major
                const int nval = nloc * 3 * L::NN;               // values of the workgroup's elements
                auto store_block = [&](R* dst, R v0, R v1, R v2) {
                    if (active) {
                        tile[g + 0] = v0;
                        tile[g + 1] = v1;
                        tile[g + 2] = v2;
                    }
                    __syncthreads();
                    const piece_t* t2 = reinterpret_cast<const piece_t*>(tile);
                    piece_t* d2 = reinterpret_cast<piece_t*>(dst);
                    const int npc = nval / VPR;
                    for (int k = threadIdx.x; k < npc; k += blockDim.x) __builtin_nontemporal_store(t2[k], d2 + k);
                    for (int k = npc * VPR + threadIdx.x; k < nval; k += blockDim.x)
                        __builtin_nontemporal_store(tile[k], dst + k);
                    __syncthreads();
                };
                store_block(static_cast<R*>(a.du) + gb, d0, d1, d2);
                stage_update(q.p, q.uu, q.vv, rb * d0, rb * d1, rb * d2, (uint32_t)gb + g, a.sr_seed, w0, w1, w2);
                store_block(static_cast<R*>(a.u_out) + gb, w0, w1, w2);
                if (!active) return;
            } else {
                st_stream<NT>(du + 0, d0);
                st_stream<NT>(du + 1, d1);
                st_stream<NT>(du + 2, d2);
                stage_update(q.p, q.uu, q.vv, rb * d0, rb * d1, rb * d2, (uint32_t)gb + g, a.sr_seed, w0, w1, w2);
                st_stream<NT>(uo + 0, w0);
                st_stream<NT>(uo + 1, w1);
                st_stream<NT>(uo + 2, w2);
            }
            if constexpr (ACC) {
                // WaveHoltz filter accumulation in double (WaveHoltz.cpp:83-87)
                const double acw = a.ac;
                double* ac = a.acc + o;
                ac[0] = ac0 + acw * (double)w0;
                ac[1] = ac1 + acw * (double)w1;
                ac[2] = ac2 + acw * (double)w2;
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

// LinearFunctional2D::action (LinearFunctional.hpp:136-202) for values f at
// the nq x nq Gauss-Legendre points of every element, fq (nv, nq, nq, nel):
// F(d, i, j) = sum_{k,l} B(k, i) B(l, j) w_k w_l det J(k, l) fq(d, k, l), with
// tab = xq[nq], wq[nq], B[nq * n] (B(k, i) = phi_i(xq_k) at k + nq i).
// One element per 64-lane workgroup (advect.hip).
void launch_linear_functional(int n, int nq, int nv, const double* corners, const double* tab, const double* fq,
                              double* F, int nel, hipStream_t s);

// DivF2D<true> with the linear flux f(x, u) = (A u, B u) (divf.hip): A, B
// are (kMaxVar, kMaxVar) column-major with the leading (nv, nv) block used.
struct DivFArgs {
    static constexpr int kMaxVar = 8;
    const double* geom;   // (6, nel) bilinear map coefficients
    const double* basis;  // x[N], w[N], D[N*N]
    const void* u;        // (nv, n, n, nel), handle precision
    void* out;            // divF += ...
    int nel, nv;
    double A[kMaxVar * kMaxVar], B[kMaxVar * kMaxVar];
    // if set, the coefficient varies by node: (nv, nv, 2, n, n, nel) with
    // f_d = sum_c Anode(d, c, 0) u_c, g_d = sum_c Anode(d, c, 1) u_c (FP64)
    const double* Anode;
    // kind 1: the compressible Euler flux of q = (rho, rho u, rho v, E) with
    // ratio of specific heats gamma (nv = 4), instead of the linear flux
    int kind;
    double gamma;
};
void launch_divf_linear(int n, int f32, const DivFArgs& a, hipStream_t s);

// Advection<true> / AdvectionHomogeneousBC<true> face terms (advect.hip):
// y(node) += H_own u(node) + H_other u(neighbour node) for every node of the
// selected sides, H per (element, side, node): (2, nv, nv) column-major.
// P: face interpolation phi_m(-1) of a Gauss-Legendre basis, or null.
void launch_advect_face(int n, int nv, const double* u, double* y, const int32_t* nbr, const int32_t* info,
                        const double* H, const double* P, int nel, int interior, int boundary, hipStream_t s);
// Advection<false> (advect.hip): volume term on nq_volume^2 points with the
// metric-weighted coefficient table op (2, nv, nv, nq, nq, nel) and tabv = P,
// D (nq, n); face terms on nq_edge points with Hq (2, nv, nv) per (element,
// side, point) and tabe = P (nq, n); P = the Gauss-Legendre face
// interpolation or null.  The launches need advect_*_q_lds bytes of dynamic
// LDS (at most 64 KiB, checked by the caller).
size_t advect_volume_q_lds(int n, int nq, int nv);
size_t advect_face_q_lds(int n, int nq, int nv);
void launch_advect_q(int n, int nv, int nq_volume, int nq_edge, const double* u, double* y, const double* op,
                     const double* tabv, const int32_t* nbr, const int32_t* info, const double* Hq,
                     const double* tabe, const double* P, int nel, int volume, int interior, int boundary,
                     hipStream_t s);
// y -= w ds LLF(n, q_own, q_other) at every node of the selected sides for the
// compressible Euler flux (advect.hip); boundary sides take q_other = q_own.
void launch_euler_face(int n, double gamma, const double* q, double* y, const double* geom, const int32_t* nbr,
                       const int32_t* info, const double* basis, const double* P, int nel, int interior, int boundary,
                       hipStream_t s);

// DivF2D<true> / EdgeFluxF2D<true> with flux functors compiled from source
// at run time (flux_rtc.hip).  Kernels compile on first use; the functions
// return false with a message on a compile or launch failure.
struct RtcFlux;
// compile only (no device): the code object of one kernel for (n, nv, leg);
// kind 0 collocated divergence, 1 collocated edge flux, 2 and 3 the same with
// quadrature rules of nq points (DivF2D<false>, EdgeFluxF2D<false>)
bool rtc_build(const std::string& user, int n, int nv, int leg, int kind, int nq, std::vector<char>& code,
               std::string& err);
RtcFlux* rtc_flux_create(const char* source, int n, int nv, int leg);
// elements per workgroup of the quadrature volume kernel
int rtc_divf_q_elems(int n, int nv, int nq);
// switch to the quadrature forms: tabv = xq, wq, P, D (nqv points), tabe =
// xq, wq, P (nqe points), P and D column-major (nq, n)
bool rtc_flux_set_rules(RtcFlux* f, int nqv, const std::vector<double>& tabv, int nqe, const std::vector<double>& tabe,
                        std::string& err);
void rtc_flux_destroy(RtcFlux* f);
// which: 0 the divergence kernel, 1 the edge-flux kernel
bool rtc_flux_prepare(RtcFlux* f, int which, std::string& err);
// out += DivF2D<true>(flux, u)
bool rtc_divf(RtcFlux* f, const double* u, double* out, const double* corners, const double* basis, int nel,
              hipStream_t s, std::string& err);
// out += scale * lifted EdgeFluxF2D<true>(numerical_flux, traces of u) on
// interior and / or boundary sides
bool rtc_edgeflux(RtcFlux* f, const double* u, double* out, const double* corners, const int32_t* nbr,
                  const int32_t* info, const double* basis, int nel, int interior, int boundary, double scale,
                  hipStream_t s, std::string& err);

// Exact-quadrature operators, FP64 (exact.hip).  Tables are device arrays,
// column-major (n_quad, n): P(q, k) = phi_k(x_q), D(q, k) = phi_k'(x_q).
// One table block per handle, offsets in doubles: volume rule x, w, P, D;
// edge rule x, w, P; mass matrices A, B, V, lam (ExactMassTables); and the
// transposes of P, D (volume), P (edge) and V, row-major copies whose rows
// are contiguous for the fused stage kernel's scalar loads; and the face
// interpolation weights of a Gauss-Legendre basis.
enum : int {
    kTabXV = 0,
    kTabWV = 10,
    kTabPV = 20,
    kTabDV = 120,
    kTabXE = 220,
    kTabWE = 230,
    kTabPE = 240,
    kTabA = 340,
    kTabB = 440,
    kTabV = 540,
    kTabLam = 640,
    kTabPVT = 650,
    kTabDVT = 750,
    kTabPET = 850,
    kTabVT = 950,
    kTabPF = 1050,  // phi_k(-1): face interpolation of a Gauss-Legendre basis
    kTabSize = 1060,
};
struct ExactTables {
    const double* xv;  // volume rule (Div<false>)
    const double* wv;
    const double* Pv;
    const double* Dv;
    const double* xe;  // edge rule (EdgeFlux<false>)
    const double* we;
    const double* Pe;
    int nqv, nqe;
    const double* Pf;  // (n) phi_k(-1), read for Gauss-Legendre bases (KArgs::leg)
};
struct ExactMassTables {  // MassMatrix<false> by fast diagonalisation (host mass_fast_diag)
    const double* A;    // (n, n) sum_q w_q phi_i phi_j
    const double* B;    // (n, n) sum_q w_q x_q phi_i phi_j
    const double* V;    // (n, n) V' A V = I, V' B V = diag(lam)
    const double* lam;  // (n)
};
// a.u_out (+)= Div<false>(a.u_in)
void launch_exact_volume(int n, const KArgs& a, const ExactTables& t, int accumulate, hipStream_t s);
// a.u_out += lifted EdgeFlux<false> of interior and/or boundary sides
void launch_exact_face(int n, const KArgs& a, const ExactTables& t, int interior, int boundary, hipStream_t s);
// kid -1: y = M x, -2: y = M^{-1} x for nv <= 8 variables; kid >= 0: the
// RHS / RK-stage epilogue of that kernel id applied to the operator output r
void launch_exact_mass(int n, const KArgs& a, const double* r, const ExactMassTables& m, int kid, hipStream_t s,
                       int nv = 3);
// The whole RHS / RK stage (kid >= K_RHS) for the default rules (n volume
// points, n + 1 edge points) in one kernel; tab is the handle's table block.
void launch_exact_stage(int n, int kid, const KArgs& a, const double* tab, hipStream_t s);

// Host-side launchers, one translation unit per N (wave2d_kernels_n.hip).
#define WDG_DECLARE_LAUNCH(NV)                                                                              \
    void launch_wave_n##NV(int kid, int f32, int sched, const KArgs& a, hipStream_t s);                              \
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
