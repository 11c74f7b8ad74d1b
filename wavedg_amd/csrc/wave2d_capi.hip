// C ABI (include/wavedg_hip.h): handles, drivers and the small kernels that
// are not part of the fused operator (mass matrix, reductions, WaveHoltz
// bookkeeping, precision conversion).
#include "../../include/wavedg_hip.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "host_basis.hpp"
#include "host_mesh.hpp"
#include "host_sync.hpp"
#include "wave2d_kernels.hpp"

using wdg::KArgs;

// ------------------------------------------------------------------ errors
namespace wdg {
thread_local std::string g_last_error;

std::mutex& capture_mutex()
{
    static std::mutex m;
    return m;
}
}  // namespace wdg
using wdg::g_last_error;

namespace {

struct Status : std::exception {
    int code;
    std::string msg;
    Status(int c, std::string m) : code(c), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Status(code, msg); }

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(WDG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded(F&& f)
{
    try {
        f();
        return WDG_OK;
    } catch (const Status& s) {
        g_last_error = s.msg;
        return s.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return WDG_ERR_ALLOC;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return WDG_ERR_INVALID;
    } catch (...) {
        g_last_error = "unknown error";
        return WDG_ERR_INVALID;
    }
}

void* dmalloc(size_t bytes)
{
    void* p = nullptr;
    if (bytes == 0) bytes = 8;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(WDG_ERR_ALLOC, "hipMalloc(" + std::to_string(bytes) + " bytes) failed: " + hipGetErrorString(e));
    }
    return p;
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

// Device buffer owned by one scope.
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) : p(dmalloc(bytes)) {}
    ~DevBuf() { dfree(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// Carpenter-Kennedy (5,4) 2N-storage coefficients (ode.hpp:68-70).
constexpr double RK4A[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                            -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
constexpr double RK4B[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                            1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                            2277821191437.0 / 14882151754819.0};
constexpr double RK4C[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                            2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

}  // namespace

// ------------------------------------------------------------------ handles

struct wdg_wave2d_s {
    int dev = 0;
    int n = 0, nel = 0, prec = WDG_F64;
    long long ndof = 0;
    size_t rsz = 8;
    double ha = -0.5, hb = -0.5;
    double min_h = 0.0;
    std::vector<double> x, w;  // basis nodes and weights
    int leg = 0;               // 1: Gauss-Legendre basis, 0: Gauss-Lobatto
    double* d_corners = nullptr;
    double* d_geom = nullptr;
    float* d_geom32 = nullptr;   // FP32 handles: geometry and basis rounded to float
    float* d_basis32 = nullptr;
    int32_t* d_nbr = nullptr;
    int32_t* d_info = nullptr;
    std::map<int, double*> d_lftab;  // LinearFunctional tables per rule size (wdg_hip_linear_functional)
    void* d_fgeo = nullptr;      // FaceGeom<R> per element side (handle precision)
    wdg::FaceNb* d_fnb = nullptr;       // neighbour trace base and flags per element side
    double* d_basis = nullptr;
    double* d_lvtab = nullptr;
    void* d_tmp = nullptr;  // RK ping-pong partner of the caller's state
    void* d_du = nullptr;   // RK register
    void* d_r1 = nullptr;   // scratch state-sized buffers (postprocess, conversions)
    void* d_r2 = nullptr;
    double* d_partial = nullptr;
    double* d_result = nullptr;
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int last_n = 0;
    hipStream_t cap_stream = nullptr;
    std::map<std::tuple<void*, int, double>, hipGraphExec_t> graphs;
    int sched = 0;     // workgroup schedule of the RK stage kernels (wdg::kSchedTgt)
    std::string arch;  // device architecture name, key of the tuned-schedule table
    // approx_quad = false: exact-quadrature operators (FP64)
    bool exact = false;
    int nqv = 0, nqe = 0, nqm = 0;
    double* d_extab = nullptr;  // quadrature and mass tables, offsets wdg::kTab*
    void* d_exr = nullptr;      // operator output before M^{-1} (custom rules only)
    bool exact_fused() const { return nqv == n && nqe == n + 1; }
    wdg::ExactTables tables() const
    {
        const double* t = d_extab;
        return wdg::ExactTables{t + wdg::kTabXV, t + wdg::kTabWV, t + wdg::kTabPV, t + wdg::kTabDV,
                                t + wdg::kTabXE, t + wdg::kTabWE, t + wdg::kTabPE, nqv, nqe, t + wdg::kTabPF};
    }
    wdg::ExactMassTables mass_tables() const
    {
        const double* t = d_extab;
        return wdg::ExactMassTables{t + wdg::kTabA, t + wdg::kTabB, t + wdg::kTabV, t + wdg::kTabLam};
    }

    ~wdg_wave2d_s()
    {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        dfree(d_corners);
        dfree(d_geom);
        dfree(d_geom32);
        dfree(d_basis32);
        dfree(d_nbr);
        dfree(d_info);
        dfree(d_fgeo);
        dfree(d_fnb);
        for (auto& kv : d_lftab) dfree(kv.second);
        dfree(d_basis);
        dfree(d_lvtab);
        dfree(d_extab);
        dfree(d_exr);
        dfree(d_tmp);
        dfree(d_du);
        dfree(d_r1);
        dfree(d_r2);
        dfree(d_partial);
        dfree(d_result);
    }

    void use_device() const
    {
        int cur = -1;
        hip_check(hipGetDevice(&cur), "hipGetDevice");
        if (cur != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
    }
    size_t state_bytes() const { return (size_t)ndof * rsz; }
    // device phi_m(-1) of a Gauss-Legendre basis, null for Gauss-Lobatto
    const double* face_interp() const { return leg ? d_basis + 2 * n + n * n : nullptr; }
    void* ensure(void*& p, size_t bytes)
    {
        if (!p) p = dmalloc(bytes);
        return p;
    }
    KArgs base_args() const
    {
        KArgs a{};
        a.geom = d_geom;
        a.nbr = d_nbr;
        a.info = d_info;
        a.fgeo = d_fgeo;
        a.fnb = d_fnb;
        a.basis = d_basis;
        a.geom32 = d_geom32;
        a.basis32 = d_basis32;
        a.leg = leg;
        // u, u' and du of a stage (3 state arrays) beyond the 256 MB Infinity Cache
        a.nt_store = 3.0 * (double)ndof * (double)rsz > 256.0 * 1024 * 1024 ? 1 : 0;
        a.ha = ha;
        a.hb = hb;
        a.nel = nel;
        return a;
    }
};

struct wdg_waveholtz_s {
    wdg_wave2d op = nullptr;
    double omega = 0, T = 0, dt = 0;
    int nt = 0;
    void* d_p = nullptr;       // RK state (handle precision)
    double* d_wprev = nullptr;
    void* d_F = nullptr;       // forcing in handle precision (F32 only)
    // One fixed-point update (S, W += pi0, residual sum) captured as a HIP
    // graph for the arrays and schedule in it_key, replayed by iterate.
    bool use_graph = true;
    hipGraphExec_t it_graph = nullptr;
    std::tuple<double*, const double*, int> it_key{nullptr, nullptr, -1};
    ~wdg_waveholtz_s()
    {
        if (it_graph) (void)hipGraphExecDestroy(it_graph);
        dfree(d_p);
        dfree(d_wprev);
        dfree(d_F);
    }
    double K(double t) const { return omega / M_PI * (std::cos(omega * t) - 0.25); }  // WaveHoltz.hpp:48-51
};

namespace {

void check_op(wdg_wave2d op)
{
    if (!op) fail(WDG_ERR_INVALID, "null wave2d handle");
    op->use_device();
}

void launch_wave(int n, int kid, int f32, int sched, const KArgs& a, hipStream_t s)
{
    switch (n) {
    case 2: wdg::launch_wave_n2(kid, f32, sched, a, s); break;
    case 3: wdg::launch_wave_n3(kid, f32, sched, a, s); break;
    case 4: wdg::launch_wave_n4(kid, f32, sched, a, s); break;
    case 5: wdg::launch_wave_n5(kid, f32, sched, a, s); break;
    case 6: wdg::launch_wave_n6(kid, f32, sched, a, s); break;
    case 7: wdg::launch_wave_n7(kid, f32, sched, a, s); break;
    case 8: wdg::launch_wave_n8(kid, f32, sched, a, s); break;
    case 9: wdg::launch_wave_n9(kid, f32, sched, a, s); break;
    default: fail(WDG_ERR_INVALID, "n_colloc out of range 2..9");
    }
    hip_check(hipGetLastError(), "wave kernel launch");
}

// One operator / RK-stage evaluation: the fused kernel; for exact-quadrature
// handles the fused exact stage kernel (default rules), or volume + face
// kernels into a scratch output and the mass-matrix kernel that finishes the
// right-hand side or stage.
void apply_op(wdg_wave2d op, int kid, const KArgs& a, hipStream_t s)
{
    if (!op->exact) {
        launch_wave(op->n, kid, op->prec == WDG_F32, op->sched, a, s);
        return;
    }
    const wdg::ExactTables t = op->tables();
    if (kid == wdg::K_ACTION) {
        wdg::launch_exact_volume(op->n, a, t, 1, s);
        wdg::launch_exact_face(op->n, a, t, 1, 0, s);
    } else if (kid == wdg::K_BC) {
        wdg::launch_exact_face(op->n, a, t, 0, 1, s);
    } else if (op->exact_fused()) {
        wdg::launch_exact_stage(op->n, kid, a, op->d_extab, s);
    } else {
        op->ensure(op->d_exr, op->state_bytes());
        KArgs r = a;
        r.u_out = op->d_exr;
        wdg::launch_exact_volume(op->n, r, t, 0, s);
        wdg::launch_exact_face(op->n, r, t, 1, 1, s);
        wdg::launch_exact_mass(op->n, a, static_cast<const double*>(op->d_exr), op->mass_tables(), kid, s);
    }
    hip_check(hipGetLastError(), "exact-quadrature kernel launch");
}

// ------------------------------------------------------------ small kernels
template <typename R>
__global__ void k_mass(const R* x, R* y, const double* __restrict__ corners, const double* __restrict__ basis, int N,
                       long long nnodes, int nv, int inv)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnodes) return;
    const int NN = N * N;
    const long long e = q / NN;
    const int node = (int)(q - e * NN);
    const int i = node % N, j = node / N;
    const double* c = corners + 8 * e;
    const double xi = basis[i], eta = basis[j];
    const double J0 = 0.25 * ((1.0 - eta) * (c[2] - c[0]) + (1.0 + eta) * (c[4] - c[6]));
    const double J1 = 0.25 * ((1.0 - eta) * (c[3] - c[1]) + (1.0 + eta) * (c[5] - c[7]));
    const double J2 = 0.25 * ((1.0 - xi) * (c[6] - c[0]) + (1.0 + xi) * (c[4] - c[2]));
    const double J3 = 0.25 * ((1.0 - xi) * (c[7] - c[1]) + (1.0 + xi) * (c[5] - c[3]));
    const double m = basis[N + i] * basis[N + j] * (J0 * J3 - J1 * J2);
    for (int d = 0; d < nv; ++d) {  // x and y may alias (in-place inverse)
        if (inv)
            y[nv * q + d] = (R)((double)x[nv * q + d] / m);
        else
            y[nv * q + d] = (R)(m * (double)x[nv * q + d]);
    }
}

__device__ double block_sum(double v)
{
    __shared__ double sh[kRedThreads];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(kRedThreads) void k_sumsq(const double* __restrict__ x, const double* __restrict__ y,
                                                       long long n, double* __restrict__ partial)
{
    double s = 0.0;
    for (long long k = (long long)blockIdx.x * kRedThreads + threadIdx.x; k < n;
         k += (long long)gridDim.x * kRedThreads) {
        const double d = y ? x[k] - y[k] : x[k];
        s += d * d;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedThreads) void k_final(const double* __restrict__ partial, int np,
                                                       double* __restrict__ out)
{
    double s = 0.0;
    for (int k = threadIdx.x; k < np; k += kRedThreads) s += partial[k];
    s = block_sum(s);
    if (threadIdx.x == 0) *out = s;
}

// S initialisation (WaveHoltz.cpp:103-107): p = w; [wprev = w]; w *= c0.
template <typename R>
__global__ void k_wh_init(double* __restrict__ w, R* __restrict__ p, double* __restrict__ wprev, double c0,
                          long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double v = w[k];
    p[k] = (R)v;
    if (wprev) wprev[k] = v;
    w[k] = v * c0;
}

// Fixed-point update and residual partial sums (waveholtz2d.cpp:172-175).
__global__ __launch_bounds__(kRedThreads) void k_wh_finalize(double* __restrict__ w, const double* __restrict__ pi0,
                                                             const double* __restrict__ wprev, long long n,
                                                             double* __restrict__ partial)
{
    double s = 0.0;
    for (long long k = (long long)blockIdx.x * kRedThreads + threadIdx.x; k < n;
         k += (long long)gridDim.x * kRedThreads) {
        const double v = w[k] + pi0[k];
        w[k] = v;
        const double d = v - wprev[k];
        s += d * d;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename S, typename D>
__global__ void k_convert(const S* __restrict__ src, D* __restrict__ dst, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = (D)src[k];
}

template <typename R>
__global__ void k_zero(R* __restrict__ x, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = (R)0;
}

// postprocess (WaveHoltz.cpp:138-144): H(0,i) = U(0,i), H(1,i) = -1/omega P(0,i).
template <typename R>
__global__ void k_post(const double* __restrict__ U, const R* __restrict__ P, double* __restrict__ H, double omega,
                       long long npts)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npts) return;
    H[2 * k] = U[3 * k];
    H[2 * k + 1] = -1.0 / omega * (double)P[3 * k];
}

unsigned grid_for(long long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

double reduce_sumsq(const double* x, const double* y, long long n, double* partial, double* result, hipStream_t s)
{
    hipLaunchKernelGGL(k_sumsq, dim3(kRedBlocks), dim3(kRedThreads), 0, s, x, y, n, partial);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(kRedThreads), 0, s, partial, kRedBlocks, result);
    hip_check(hipGetLastError(), "reduction launch");
    double h = 0.0;
    hip_check(hipMemcpyAsync(&h, result, sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    return h;
}

template <typename R>
void zero_state(R* x, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero<R>, dim3(grid_for(n)), dim3(256), 0, s, x, n);
}

void zero_any(void* x, int prec, long long n, hipStream_t s)
{
    if (prec == WDG_F32)
        zero_state(static_cast<float*>(x), n, s);
    else
        zero_state(static_cast<double*>(x), n, s);
}

// Runs n_steps RK4 steps starting from state `u` (ping-pong with op->d_tmp).
// force: load vector in handle precision or nullptr (then omega unused).
// acc: WaveHoltz accumulator or nullptr; coef(it, t_after) gives its weight.
// Returns the buffer holding the final state.
template <typename Coef>
void* run_steps(wdg_wave2d op, double dt, double& t, void* u, int n_steps, const void* force, double omega,
                double* acc, Coef coef, hipStream_t s)
{
    op->ensure(op->d_tmp, op->state_bytes());
    op->ensure(op->d_du, op->state_bytes());
    KArgs a = op->base_args();
    a.du = op->d_du;
    a.dt = dt;
    a.force = force;
    void* cur = u;
    void* nxt = op->d_tmp;
    for (int it = 0; it < n_steps; ++it) {
        const double t_after = t + dt;
        for (int st = 0; st < 5; ++st) {
            a.u_in = cur;
            a.u_out = nxt;
            a.rk_a = RK4A[st];
            a.rk_b = RK4B[st];
            a.sr_seed = wdg::sr_seed(it, st);
            int kid;
            if (st == 0) {
                kid = force ? wdg::K_STAGE0_F : wdg::K_STAGE0;
            } else if (st == 4 && acc) {
                kid = force ? wdg::K_STAGE_F_ACC : wdg::K_STAGE_ACC;
                a.acc = acc;
                a.ac = coef(it, t_after);
            } else {
                kid = force ? wdg::K_STAGE_F : wdg::K_STAGE;
            }
            if (force) {
                const double ts = (st == 0) ? t : t + RK4C[st] * dt;
                a.fs = -1.0 / omega * std::sin(omega * ts);  // WaveHoltz.cpp:65
            }
            apply_op(op, kid, a, s);
            std::swap(cur, nxt);
        }
        t += dt;
    }
    return cur;
}

struct NoCoef {
    double operator()(int, double) const { return 0.0; }
};

// ------------------------------------------------------- schedule autotuner
// Tuned RK-stage schedules, process-wide: (architecture, n_colloc,
// precision) -> index into wdg::kSchedTgt.  New handles start from it.
std::mutex g_tune_mu;
std::map<std::tuple<std::string, int, int>, int> g_tune;

std::string device_arch(int dev)
{
    hipDeviceProp_t p;
    hip_check(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
    std::string a = p.gcnArchName;
    const size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

int tuned_schedule(const std::string& arch, int n, int prec)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find({arch, n, prec});
    return it == g_tune.end() ? 0 : it->second;
}

// Deterministic pseudo-random fill in [-1, 1] (SplitMix64 of the index).
template <typename R>
__global__ void k_fill_hash(R* x, long long n, unsigned long long seed)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = seed + (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = (R)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}

// Counts positions where a and b differ bitwise.
template <typename U>
__global__ void k_count_diff(const U* a, const U* b, long long n, unsigned int* count)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a[i] != b[i]) atomicAdd(count, 1u);
}

void fill_hash(int prec, void* x, long long n, unsigned long long seed, hipStream_t s)
{
    const unsigned g = (unsigned)((n + 255) / 256);
    if (prec == WDG_F32)
        hipLaunchKernelGGL(k_fill_hash<float>, dim3(g), dim3(256), 0, s, static_cast<float*>(x), n, seed);
    else
        hipLaunchKernelGGL(k_fill_hash<double>, dim3(g), dim3(256), 0, s, static_cast<double*>(x), n, seed);
}

unsigned int count_diff(int prec, const void* a, const void* b, long long n, unsigned int* d_count, hipStream_t s)
{
    hip_check(hipMemsetAsync(d_count, 0, sizeof(unsigned int), s), "hipMemsetAsync");
    const unsigned g = (unsigned)((n + 255) / 256);
    if (prec == WDG_F32)
        hipLaunchKernelGGL(k_count_diff<uint32_t>, dim3(g), dim3(256), 0, s, static_cast<const uint32_t*>(a),
                           static_cast<const uint32_t*>(b), n, d_count);
    else
        hipLaunchKernelGGL(k_count_diff<uint64_t>, dim3(g), dim3(256), 0, s, static_cast<const uint64_t*>(a),
                           static_cast<const uint64_t*>(b), n, d_count);
    unsigned int h = 0;
    hip_check(hipMemcpyAsync(&h, d_count, sizeof(h), hipMemcpyDeviceToHost, s), "copy count");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    return h;
}

void start_timing(wdg_wave2d op, hipStream_t s)
{
    if (!op->timing) return;
    if (!op->ev0) hip_check(hipEventCreate(&op->ev0), "hipEventCreate");
    if (!op->ev1) hip_check(hipEventCreate(&op->ev1), "hipEventCreate");
    hip_check(hipEventRecord(op->ev0, s), "hipEventRecord");
}

void stop_timing(wdg_wave2d op, hipStream_t s, int n_kernels)
{
    if (!op->timing) return;
    hip_check(hipEventRecord(op->ev1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(op->ev1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, op->ev0, op->ev1), "hipEventElapsedTime");
    op->last_ms = ms;
    op->last_n = n_kernels;
}

void check_wh(wdg_waveholtz wh)
{
    if (!wh) fail(WDG_ERR_INVALID, "null waveholtz handle");
    check_op(wh->op);
}

// S(w) with optional wprev copy; accumulation in w.
void wh_apply_S(wdg_waveholtz wh, double* d_w, double* d_wprev, hipStream_t s)
{
    wdg_wave2d op = wh->op;
    wh->op->ensure(wh->d_p, op->state_bytes());
    const double c0 = 0.5 * wh->dt * wh->K(0.0);
    const long long n = op->ndof;
    if (op->prec == WDG_F32)
        hipLaunchKernelGGL(k_wh_init<float>, dim3(grid_for(n)), dim3(256), 0, s, d_w, static_cast<float*>(wh->d_p),
                           d_wprev, c0, n);
    else
        hipLaunchKernelGGL(k_wh_init<double>, dim3(grid_for(n)), dim3(256), 0, s, d_w,
                           static_cast<double*>(wh->d_p), d_wprev, c0, n);
    hip_check(hipGetLastError(), "k_wh_init launch");
    double t = 0.0;
    const int nt = wh->nt;
    const double dt = wh->dt;
    auto coef = [wh, nt, dt](int it, double t_after) {
        return ((it + 1 == nt) ? 0.5 : 1.0) * dt * wh->K(t_after);  // WaveHoltz.cpp:115-119
    };
    run_steps(op, dt, t, wh->d_p, nt, nullptr, 0.0, d_w, coef, s);
}

}  // namespace

// ====================================================================== API
extern "C" {

const char* wdg_hip_last_error(void) { return g_last_error.c_str(); }

int wdg_hip_version(int* major, int* minor)
{
    if (major) *major = 0;
    if (minor) *minor = 1;
    return WDG_OK;
}

int wdg_hip_malloc(void** d_ptr, size_t bytes)
{
    return guarded([&] {
        if (!d_ptr) fail(WDG_ERR_INVALID, "null output pointer");
        *d_ptr = dmalloc(bytes);
    });
}

int wdg_hip_free(void* d_ptr)
{
    return guarded([&] { hip_check(hipFree(d_ptr), "hipFree"); });
}

int wdg_hip_memcpy_htod(void* d_dst, const void* src, size_t bytes)
{
    return guarded([&] { hip_check(hipMemcpy(d_dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D"); });
}

int wdg_hip_memcpy_dtoh(void* dst, const void* d_src, size_t bytes)
{
    return guarded([&] { hip_check(hipMemcpy(dst, d_src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H"); });
}

int wdg_hip_memset(void* d_ptr, int value, size_t bytes)
{
    return guarded([&] { hip_check(hipMemset(d_ptr, value, bytes), "hipMemset"); });
}

int wdg_hip_synchronize(void* stream)
{
    return guarded([&] { hip_check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize"); });
}

// ---------------------------------------------------------------- mesh
int wdg_hip_mesh_from_vertices(int n_vertices, const double* x, int n_elem, const int* elements, wdg_mesh* out)
{
    return guarded([&] {
        if (!out) fail(WDG_ERR_INVALID, "null output handle");
        try {
            *out = new wdg_mesh_s{wdg::HostMesh::from_vertices(n_vertices, x, n_elem, elements)};
        } catch (const std::invalid_argument& e) {
            fail(WDG_ERR_MESH, e.what());
        }
    });
}

int wdg_hip_mesh_uniform_rect(int nx, double ax, double bx, int ny, double ay, double by, wdg_mesh* out)
{
    return guarded([&] {
        if (!out) fail(WDG_ERR_INVALID, "null output handle");
        try {
            *out = new wdg_mesh_s{wdg::HostMesh::uniform_rect(nx, ax, bx, ny, ay, by)};
        } catch (const std::invalid_argument& e) {
            fail(WDG_ERR_MESH, e.what());
        }
    });
}

int wdg_hip_mesh_from_file(const char* dir, wdg_mesh* out)
{
    return guarded([&] {
        if (!out || !dir) fail(WDG_ERR_INVALID, "null argument");
        try {
            *out = new wdg_mesh_s{wdg::HostMesh::from_file(dir)};
        } catch (const std::exception& e) {
            fail(WDG_ERR_MESH, e.what());
        }
    });
}

int wdg_hip_mesh_destroy(wdg_mesh mesh)
{
    delete mesh;
    return WDG_OK;
}

int wdg_hip_mesh_counts(wdg_mesh mesh, int* n_elem, int* n_interior, int* n_boundary)
{
    return guarded([&] {
        if (!mesh) fail(WDG_ERR_INVALID, "null mesh handle");
        if (n_elem) *n_elem = mesh->mesh.n_elem();
        if (n_interior) *n_interior = mesh->mesh.n_interior();
        if (n_boundary) *n_boundary = mesh->mesh.n_boundary();
    });
}

int wdg_hip_mesh_edges(wdg_mesh mesh, int face_type, int* out)
{
    return guarded([&] {
        if (!mesh || !out) fail(WDG_ERR_INVALID, "null argument");
        const bool interior = face_type == WDG_INTERIOR;
        const int ne = interior ? mesh->mesh.n_interior() : mesh->mesh.n_boundary();
        for (int i = 0; i < ne; ++i) {
            const wdg::HostEdge& E = mesh->mesh.edge(i, interior);
            out[5 * i + 0] = E.el[0];
            out[5 * i + 1] = E.side[0];
            out[5 * i + 2] = E.el[1];
            out[5 * i + 3] = E.side[1];
            out[5 * i + 4] = E.delta;
        }
    });
}

int wdg_hip_mesh_edge_geometry(wdg_mesh mesh, int face_type, double* out)
{
    return guarded([&] {
        if (!mesh || !out) fail(WDG_ERR_INVALID, "null argument");
        const bool interior = face_type == WDG_INTERIOR;
        const int ne = interior ? mesh->mesh.n_interior() : mesh->mesh.n_boundary();
        for (int i = 0; i < ne; ++i) {
            const wdg::HostEdge& E = mesh->mesh.edge(i, interior);
            out[5 * i + 0] = E.normal[0];
            out[5 * i + 1] = E.normal[1];
            out[5 * i + 2] = E.half_length;
            out[5 * i + 3] = E.x0[0] + E.dx[0] * 0.5;
            out[5 * i + 4] = E.x0[1] + E.dx[1] * 0.5;
        }
    });
}

int wdg_hip_mesh_min_h(wdg_mesh mesh, double* h)
{
    return guarded([&] {
        if (!mesh || !h) fail(WDG_ERR_INVALID, "null argument");
        *h = mesh->mesh.min_h();
    });
}

int wdg_hip_mesh_nodes(wdg_mesh mesh, int n_colloc, double* out)
{
    return guarded([&] {
        if (!mesh || !out) fail(WDG_ERR_INVALID, "null argument");
        std::vector<double> x, w;
        if (!wdg::gll_rule(n_colloc, x, w)) fail(WDG_ERR_INVALID, "n_colloc out of range 2..9");
        mesh->mesh.nodes(n_colloc, x.data(), out);
    });
}

int wdg_hip_gll(int n, double* x, double* w, double* D)
{
    return guarded([&] {
        std::vector<double> xv, wv;
        if (!wdg::gll_rule(n, xv, wv)) fail(WDG_ERR_INVALID, "n out of range 2..9");
        if (x) std::memcpy(x, xv.data(), sizeof(double) * n);
        if (w) std::memcpy(w, wv.data(), sizeof(double) * n);
        if (D) wdg::lagrange_derivative_matrix(n, xv.data(), n, xv.data(), D);
    });
}

// |C| of a normal matrix C = n.A with real eigenvalues and a full set of
// eigenvectors (real_eig, eig.cpp:29-74); refusals name the failure as the
// reference's EdgeFlux errors do (EdgeFlux.cpp:59-60, 86-88).
static void abs_normal_matrix(int m, const double* C, double* absC, const char* where)
{
    const int rc = wdg::real_abs(m, C, absC);
    if (rc == 1) fail(WDG_ERR_INVALID, std::string("n.A has complex eigenvalues ") + where);
    if (rc == 2) fail(WDG_ERR_INVALID, std::string("eigenvectors of n.A are linearly dependent ") + where);
    if (rc != 0) fail(WDG_ERR_INVALID, std::string("eigen-decomposition of n.A failed ") + where);
}

int wdg_hip_flux_matrix(int m, const double* nA, double a, double b, int side, double* F)
{
    return guarded([&] {
        if (m < 1 || m > 8) fail(WDG_ERR_INVALID, "m out of range 1..8");
        if (!nA || !F) fail(WDG_ERR_INVALID, "null pointer");
        if (side != 0 && side != 1) fail(WDG_ERR_INVALID, "side must be 0 or 1");
        double absC[64] = {0.0};
        if (b != 0.0) abs_normal_matrix(m, nA, absC, "");
        const double sgn = side == 0 ? 1.0 : -1.0;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) F[i + m * j] = 0.5 * a * nA[j + m * i] + sgn * b * absC[j + m * i];
    });
}

int wdg_hip_gauss_legendre(int n, double* x, double* w)
{
    return guarded([&] {
        std::vector<double> xv, wv;
        if (!wdg::gauss_legendre_rule(n, xv, wv)) fail(WDG_ERR_INVALID, "n out of range 1..20");
        if (x) std::memcpy(x, xv.data(), sizeof(double) * n);
        if (w) std::memcpy(w, wv.data(), sizeof(double) * n);
    });
}

// ---------------------------------------------------------------- operator
// approx_quad = true handle on a Gauss-Lobatto (basis_type WDG_GAUSS_LOBATTO)
// or Gauss-Legendre (WDG_GAUSS_LEGENDRE) basis of n_colloc points.
static int create_collocated(wdg_mesh mesh, int n_colloc, int basis_type, const int* bc, double a, double b,
                             int precision, wdg_wave2d* out)
{
    return guarded([&] {
        if (!mesh || !out) fail(WDG_ERR_INVALID, "null argument");
        if (precision != WDG_F64 && precision != WDG_F32) fail(WDG_ERR_INVALID, "precision must be WDG_F64 or WDG_F32");
        if (basis_type != WDG_GAUSS_LOBATTO && basis_type != WDG_GAUSS_LEGENDRE)
            fail(WDG_ERR_INVALID, "basis_type must be WDG_GAUSS_LOBATTO or WDG_GAUSS_LEGENDRE");
        if (!std::isfinite(a) || !std::isfinite(b)) fail(WDG_ERR_INVALID, "flux coefficients must be finite");
        const wdg::HostMesh& m = mesh->mesh;
        auto op = std::make_unique<wdg_wave2d_s>();
        // validate everything on the host before touching the device
        op->leg = basis_type == WDG_GAUSS_LEGENDRE ? 1 : 0;
        if (n_colloc < 2 || n_colloc > 9) fail(WDG_ERR_INVALID, "n_colloc out of range 2..9");
        const bool rule_ok = op->leg ? wdg::gauss_legendre_rule(n_colloc, op->x, op->w)
                                     : wdg::gll_rule(n_colloc, op->x, op->w);
        if (!rule_ok) fail(WDG_ERR_INVALID, "n_colloc out of range 2..9");
        if (m.min_corner_detj() <= 0.0) fail(WDG_ERR_MESH, "mesh has an inverted or degenerate element (det J <= 0)");
        wdg::FaceTables ft;
        std::string err;
        if (!m.face_tables(bc, ft, err)) fail(WDG_ERR_INVALID, err);
        hip_check(hipGetDevice(&op->dev), "hipGetDevice");
        op->n = n_colloc;
        op->nel = m.n_elem();
        op->ndof = 3LL * n_colloc * n_colloc * op->nel;
        op->prec = precision;
        op->rsz = precision == WDG_F32 ? 4 : 8;
        op->ha = -0.5 * a;  // reference sign convention (EdgeFlux.hpp:51)
        op->hb = -b;
        op->min_h = m.min_h();
        op->arch = device_arch(op->dev);
        // the tuned table holds Gauss-Lobatto schedules; Gauss-Legendre handles start from 0
        op->sched = op->leg ? 0 : tuned_schedule(op->arch, n_colloc, precision);

        const int n = n_colloc;
        // x[n], w[n], D[n*n], P[n] = phi_k(-1): the face interpolation of a
        // Gauss-Legendre basis (LegendreFaceProlongator, FaceProlongator.cpp:717-725)
        std::vector<double> basis(3 * n + n * n);
        std::memcpy(basis.data(), op->x.data(), sizeof(double) * n);
        std::memcpy(basis.data() + n, op->w.data(), sizeof(double) * n);
        wdg::lagrange_derivative_matrix(n, op->x.data(), n, op->x.data(), basis.data() + 2 * n);
        const double xm1 = -1.0;
        wdg::lagrange_interpolation_matrix(n, op->x.data(), 1, &xm1, basis.data() + 2 * n + n * n);
        std::vector<double> xq, wq;
        wdg::gauss_legendre_rule(n, xq, wq);
        std::vector<double> lv(2 * n + n * n);
        std::memcpy(lv.data(), xq.data(), sizeof(double) * n);
        std::memcpy(lv.data() + n, wq.data(), sizeof(double) * n);
        wdg::lagrange_interpolation_matrix(n, op->x.data(), n, xq.data(), lv.data() + 2 * n);

        // bilinear map coefficients a, c, h per element (see wave2d_kernels.hpp)
        std::vector<double> geom(6 * (size_t)op->nel);
        for (int e = 0; e < op->nel; ++e) {
            const double* c = &m.corners()[8 * (size_t)e];
            double* gq = &geom[6 * (size_t)e];
            gq[0] = 0.25 * ((c[2] - c[0]) + (c[4] - c[6]));
            gq[1] = 0.25 * ((c[3] - c[1]) + (c[5] - c[7]));
            gq[2] = 0.25 * ((c[6] - c[0]) + (c[4] - c[2]));
            gq[3] = 0.25 * ((c[7] - c[1]) + (c[5] - c[3]));
            gq[4] = 0.25 * ((c[0] - c[2]) + (c[4] - c[6]));
            gq[5] = 0.25 * ((c[1] - c[3]) + (c[5] - c[7]));
        }
        op->d_geom = static_cast<double*>(dmalloc(sizeof(double) * geom.size()));
        hip_check(hipMemcpy(op->d_geom, geom.data(), sizeof(double) * geom.size(), hipMemcpyHostToDevice),
                  "upload geometry");
        op->d_corners = static_cast<double*>(dmalloc(sizeof(double) * m.corners().size()));
        op->d_nbr = static_cast<int32_t*>(dmalloc(sizeof(int32_t) * ft.nbr.size()));
        op->d_info = static_cast<int32_t*>(dmalloc(sizeof(int32_t) * ft.info.size()));
        op->d_basis = static_cast<double*>(dmalloc(sizeof(double) * basis.size()));
        op->d_lvtab = static_cast<double*>(dmalloc(sizeof(double) * lv.size()));
        op->d_partial = static_cast<double*>(dmalloc(sizeof(double) * kRedBlocks));
        op->d_result = static_cast<double*>(dmalloc(sizeof(double)));
        hip_check(hipMemcpy(op->d_corners, m.corners().data(), sizeof(double) * m.corners().size(),
                            hipMemcpyHostToDevice), "upload corners");
        hip_check(hipMemcpy(op->d_nbr, ft.nbr.data(), sizeof(int32_t) * ft.nbr.size(), hipMemcpyHostToDevice),
                  "upload nbr");
        hip_check(hipMemcpy(op->d_info, ft.info.data(), sizeof(int32_t) * ft.info.size(), hipMemcpyHostToDevice),
                  "upload info");
        hip_check(hipMemcpy(op->d_basis, basis.data(), sizeof(double) * basis.size(), hipMemcpyHostToDevice),
                  "upload basis");
        hip_check(hipMemcpy(op->d_lvtab, lv.data(), sizeof(double) * lv.size(), hipMemcpyHostToDevice),
                  "upload load-vector table");
        {
            std::vector<wdg::FaceNb> fnb;
            if (precision == WDG_F32) {
                std::vector<wdg::FaceGeom<float>> fg;
                wdg::build_face_tables(n, op->nel, geom.data(), ft.nbr.data(), ft.info.data(), fg, fnb);
                op->d_fgeo = dmalloc(sizeof(fg[0]) * fg.size());
                hip_check(hipMemcpy(op->d_fgeo, fg.data(), sizeof(fg[0]) * fg.size(), hipMemcpyHostToDevice),
                          "upload face geometry");
            } else {
                // FP64 kernels compute side normals from the map coefficients: only the neighbour table
                std::vector<wdg::FaceGeom<double>> fg;
                wdg::build_face_tables(n, op->nel, geom.data(), ft.nbr.data(), ft.info.data(), fg, fnb);
            }
            op->d_fnb = static_cast<wdg::FaceNb*>(dmalloc(sizeof(wdg::FaceNb) * fnb.size()));
            hip_check(hipMemcpy(op->d_fnb, fnb.data(), sizeof(wdg::FaceNb) * fnb.size(), hipMemcpyHostToDevice),
                      "upload face neighbours");
        }
        if (precision == WDG_F32) {
            const std::vector<float> g32(geom.begin(), geom.end()), b32(basis.begin(), basis.end());
            op->d_geom32 = static_cast<float*>(dmalloc(sizeof(float) * g32.size()));
            op->d_basis32 = static_cast<float*>(dmalloc(sizeof(float) * b32.size()));
            hip_check(hipMemcpy(op->d_geom32, g32.data(), sizeof(float) * g32.size(), hipMemcpyHostToDevice),
                      "upload float geometry");
            hip_check(hipMemcpy(op->d_basis32, b32.data(), sizeof(float) * b32.size(), hipMemcpyHostToDevice),
                      "upload float basis");
        }
        *out = op.release();
    });
}

int wdg_hip_wave2d_create(wdg_mesh mesh, int n_colloc, const int* bc, double a, double b, int precision,
                          wdg_wave2d* out)
{
    return create_collocated(mesh, n_colloc, WDG_GAUSS_LOBATTO, bc, a, b, precision, out);
}

static int create_exact(wdg_mesh mesh, int n_colloc, int basis_type, const int* bc, double a, double b,
                        int nq_volume, int nq_edge, int nq_mass, wdg_wave2d* out)
{
    if (!out) {
        g_last_error = "null argument";
        return WDG_ERR_INVALID;
    }
    *out = nullptr;
    // the reference's default rules for bilinear elements (Div.cpp:320-327,
    // EdgeFlux.cpp:503-510, MassMatrix.cpp:92-97)
    if (nq_volume <= 0) nq_volume = 1 + ((n_colloc - 2) + (n_colloc - 1) + 1) / 2;
    if (nq_edge <= 0) nq_edge = 1 + (2 * (n_colloc - 1) + 2) / 2;
    if (nq_mass <= 0) nq_mass = 1 + (2 * (n_colloc - 1) + 2) / 2;
    if (nq_volume > 10 || nq_edge > 10 || nq_mass > 10) {
        g_last_error = "quadrature rules have at most 10 points";
        return WDG_ERR_INVALID;
    }
    wdg_wave2d op = nullptr;
    const int rc = create_collocated(mesh, n_colloc, basis_type, bc, a, b, WDG_F64, &op);
    if (rc) return rc;
    const int rc2 = guarded([&] {
        const int n = n_colloc;
        std::vector<double> tab(wdg::kTabSize, 0.0);
        auto rule = [&](int nq, std::vector<double>& xq, std::vector<double>& wq) {
            if (!wdg::gauss_legendre_rule(nq, xq, wq)) fail(WDG_ERR_INVALID, "no Gauss-Legendre rule of that size");
        };
        std::vector<double> xq, wq;
        rule(nq_volume, xq, wq);
        std::copy(xq.begin(), xq.end(), tab.begin() + wdg::kTabXV);
        std::copy(wq.begin(), wq.end(), tab.begin() + wdg::kTabWV);
        wdg::lagrange_interpolation_matrix(n, op->x.data(), nq_volume, xq.data(), tab.data() + wdg::kTabPV);
        wdg::lagrange_derivative_matrix(n, op->x.data(), nq_volume, xq.data(), tab.data() + wdg::kTabDV);
        rule(nq_edge, xq, wq);
        std::copy(xq.begin(), xq.end(), tab.begin() + wdg::kTabXE);
        std::copy(wq.begin(), wq.end(), tab.begin() + wdg::kTabWE);
        wdg::lagrange_interpolation_matrix(n, op->x.data(), nq_edge, xq.data(), tab.data() + wdg::kTabPE);
        rule(nq_mass, xq, wq);
        if (!wdg::mass_fast_diag(n, op->x.data(), nq_mass, xq.data(), wq.data(), tab.data() + wdg::kTabA,
                                 tab.data() + wdg::kTabB, tab.data() + wdg::kTabV, tab.data() + wdg::kTabLam))
            fail(WDG_ERR_INVALID, "the mass rule is too small for the basis (singular mass matrix)");
        auto transpose = [&](int src, int rows, int cols, int dst) {  // (rows, cols) column-major
            for (int c = 0; c < cols; ++c)
                for (int r = 0; r < rows; ++r) tab[dst + c + cols * r] = tab[src + r + rows * c];
        };
        transpose(wdg::kTabPV, nq_volume, n, wdg::kTabPVT);
        transpose(wdg::kTabDV, nq_volume, n, wdg::kTabDVT);
        transpose(wdg::kTabPE, nq_edge, n, wdg::kTabPET);
        transpose(wdg::kTabV, n, n, wdg::kTabVT);
        const double xm1 = -1.0;
        wdg::lagrange_interpolation_matrix(n, op->x.data(), 1, &xm1, tab.data() + wdg::kTabPF);
        op->exact = true;
        op->nqv = nq_volume;
        op->nqe = nq_edge;
        op->nqm = nq_mass;
        op->d_extab = static_cast<double*>(dmalloc(sizeof(double) * tab.size()));
        hip_check(hipMemcpy(op->d_extab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice),
                  "upload quadrature tables");
    });
    if (rc2) {
        wdg_hip_wave2d_destroy(op);
        return rc2;
    }
    *out = op;
    return WDG_OK;
}

int wdg_hip_wave2d_create_exact(wdg_mesh mesh, int n_colloc, const int* bc, double a, double b, int nq_volume,
                                int nq_edge, int nq_mass, wdg_wave2d* out)
{
    return create_exact(mesh, n_colloc, WDG_GAUSS_LOBATTO, bc, a, b, nq_volume, nq_edge, nq_mass, out);
}

int wdg_hip_wave2d_create_basis(wdg_mesh mesh, int n_colloc, int basis_type, const int* bc, double a, double b,
                                int precision, int approx_quad, int nq_volume, int nq_edge, int nq_mass,
                                wdg_wave2d* out)
{
    if (approx_quad) return create_collocated(mesh, n_colloc, basis_type, bc, a, b, precision, out);
    if (precision != WDG_F64) {
        if (out) *out = nullptr;
        g_last_error = "approx_quad = false handles are FP64";
        return WDG_ERR_INVALID;
    }
    if (basis_type != WDG_GAUSS_LOBATTO && basis_type != WDG_GAUSS_LEGENDRE) {
        if (out) *out = nullptr;
        g_last_error = "basis_type must be WDG_GAUSS_LOBATTO or WDG_GAUSS_LEGENDRE";
        return WDG_ERR_INVALID;
    }
    return create_exact(mesh, n_colloc, basis_type, bc, a, b, nq_volume, nq_edge, nq_mass, out);
}

int wdg_hip_wave2d_basis(wdg_wave2d op, int* basis_type, double* x, double* w)
{
    return guarded([&] {
        check_op(op);
        if (basis_type) *basis_type = op->leg ? WDG_GAUSS_LEGENDRE : WDG_GAUSS_LOBATTO;
        if (x) std::memcpy(x, op->x.data(), sizeof(double) * op->n);
        if (w) std::memcpy(w, op->w.data(), sizeof(double) * op->n);
    });
}

int wdg_hip_wave2d_quadrature(wdg_wave2d op, int* exact, int* nq_volume, int* nq_edge, int* nq_mass)
{
    return guarded([&] {
        check_op(op);
        if (exact) *exact = op->exact ? 1 : 0;
        if (nq_volume) *nq_volume = op->nqv;
        if (nq_edge) *nq_edge = op->nqe;
        if (nq_mass) *nq_mass = op->nqm;
    });
}

int wdg_hip_wave2d_destroy(wdg_wave2d op)
{
    if (op) {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != op->dev) (void)hipSetDevice(op->dev);
        delete op;
        if (cur >= 0) (void)hipSetDevice(cur);
    }
    return WDG_OK;
}

int wdg_hip_wave2d_info(wdg_wave2d op, long long* ndof, int* n_colloc, int* n_elem, int* precision)
{
    return guarded([&] {
        if (!op) fail(WDG_ERR_INVALID, "null wave2d handle");
        if (ndof) *ndof = op->ndof;
        if (n_colloc) *n_colloc = op->n;
        if (n_elem) *n_elem = op->nel;
        if (precision) *precision = op->prec;
    });
}

static int op_launch(wdg_wave2d op, int kid, const void* x, void* y, const void* F, double s, void* stream)
{
    return guarded([&] {
        check_op(op);
        if (!x || !y) fail(WDG_ERR_INVALID, "null device pointer");
        if (x == y) fail(WDG_ERR_INVALID, "input and output must not alias");
        KArgs a = op->base_args();
        a.u_in = x;
        a.u_out = y;
        a.force = F;
        a.fs = s;
        apply_op(op, kid, a, as_stream(stream));
    });
}

int wdg_hip_wave2d_action(wdg_wave2d op, const void* d_x, void* d_y, void* stream)
{
    return op_launch(op, wdg::K_ACTION, d_x, d_y, nullptr, 0.0, stream);
}

int wdg_hip_wavebc_action(wdg_wave2d op, const void* d_x, void* d_y, void* stream)
{
    return op_launch(op, wdg::K_BC, d_x, d_y, nullptr, 0.0, stream);
}

int wdg_hip_wave2d_rhs(wdg_wave2d op, const void* d_x, void* d_y, const void* d_F, double s, void* stream)
{
    return op_launch(op, d_F ? wdg::K_RHS_F : wdg::K_RHS, d_x, d_y, d_F, s, stream);
}

static int mass_launch(wdg_wave2d op, const void* x, void* y, int inv, void* stream, int nv = 3)
{
    return guarded([&] {
        check_op(op);
        if (!x || !y) fail(WDG_ERR_INVALID, "null device pointer");
        if (nv < 1 || nv > 64) fail(WDG_ERR_INVALID, "n_var out of range 1..64");
        const long long nn = op->ndof / 3;
        const hipStream_t s = as_stream(stream);
        if (op->exact) {
            if (nv > wdg::DivFArgs::kMaxVar) fail(WDG_ERR_INVALID, "the exact mass matrix takes n_var <= 8");
            KArgs a = op->base_args();
            a.u_in = x;
            a.u_out = y;
            wdg::launch_exact_mass(op->n, a, nullptr, op->mass_tables(), inv ? -2 : -1, s, nv);
            hip_check(hipGetLastError(), "mass kernel launch");
            return;
        }
        if (op->prec == WDG_F32)
            hipLaunchKernelGGL(k_mass<float>, dim3(grid_for(nn)), dim3(256), 0, s, static_cast<const float*>(x),
                               static_cast<float*>(y), op->d_corners, op->d_basis, op->n, nn, nv, inv);
        else
            hipLaunchKernelGGL(k_mass<double>, dim3(grid_for(nn)), dim3(256), 0, s, static_cast<const double*>(x),
                               static_cast<double*>(y), op->d_corners, op->d_basis, op->n, nn, nv, inv);
        hip_check(hipGetLastError(), "mass kernel launch");
    });
}

int wdg_hip_mass_inv(wdg_wave2d op, void* d_x, void* stream) { return mass_launch(op, d_x, d_x, 1, stream); }

int wdg_hip_mass_action(wdg_wave2d op, const void* d_x, void* d_y, void* stream)
{
    return mass_launch(op, d_x, d_y, 0, stream);
}

int wdg_hip_mass_apply(wdg_wave2d op, int n_var, const void* d_x, void* d_y, int inverse, void* stream)
{
    return mass_launch(op, d_x, d_y, inverse ? 1 : 0, stream, n_var);
}

int wdg_hip_rk4_steps(wdg_wave2d op, double dt, double* t, void* d_u, int n_steps, const void* d_F, double omega,
                      void* stream)
{
    return guarded([&] {
        check_op(op);
        if (!t || !d_u) fail(WDG_ERR_INVALID, "null argument");
        if (n_steps < 0) fail(WDG_ERR_INVALID, "n_steps must be >= 0");
        if (!(dt > 0.0) || !std::isfinite(dt)) fail(WDG_ERR_INVALID, "dt must be positive and finite");
        if (d_F && !(omega > 0.0)) fail(WDG_ERR_INVALID, "forcing requires omega > 0");
        const hipStream_t s = as_stream(stream);
        start_timing(op, s);
        double tt = *t;
        void* fin = run_steps(op, dt, tt, d_u, n_steps, d_F, omega, nullptr, NoCoef{}, s);
        stop_timing(op, s, 5 * n_steps);
        if (fin != d_u)
            hip_check(hipMemcpyAsync(d_u, fin, op->state_bytes(), hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
        *t = tt;
    });
}

int wdg_hip_rk4_steps_graph(wdg_wave2d op, double dt, double* t, void* d_u, int n_steps, int steps_per_graph,
                            void* stream)
{
    return guarded([&] {
        check_op(op);
        if (!t || !d_u) fail(WDG_ERR_INVALID, "null argument");
        if (steps_per_graph <= 0 || (steps_per_graph & 1))
            fail(WDG_ERR_INVALID, "steps_per_graph must be a positive even number");
        if (n_steps < 0 || n_steps % steps_per_graph)
            fail(WDG_ERR_INVALID, "n_steps must be a non-negative multiple of steps_per_graph");
        if (!(dt > 0.0) || !std::isfinite(dt)) fail(WDG_ERR_INVALID, "dt must be positive and finite");
        const hipStream_t s = as_stream(stream);
        const auto key = std::make_tuple(d_u, steps_per_graph, dt);
        auto it = op->graphs.find(key);
        if (it == op->graphs.end()) {
            std::lock_guard<std::mutex> lk(wdg::capture_mutex());
            if (!op->cap_stream) hip_check(hipStreamCreateWithFlags(&op->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
            op->ensure(op->d_tmp, op->state_bytes());
            op->ensure(op->d_du, op->state_bytes());
            hipGraph_t g = nullptr;
            hip_check(hipStreamBeginCapture(op->cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            double tc = 0.0;
            try {
                run_steps(op, dt, tc, d_u, steps_per_graph, nullptr, 0.0, nullptr, NoCoef{}, op->cap_stream);
            } catch (...) {
                (void)hipStreamEndCapture(op->cap_stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(op->cap_stream, &g), "hipStreamEndCapture");
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(ie, "hipGraphInstantiate");
            it = op->graphs.emplace(key, ex).first;
        }
        start_timing(op, s);
        for (int k = 0; k < n_steps / steps_per_graph; ++k) hip_check(hipGraphLaunch(it->second, s), "hipGraphLaunch");
        stop_timing(op, s, 5 * n_steps);
        double tt = *t;
        for (int k = 0; k < n_steps; ++k) tt += dt;
        *t = tt;
    });
}

int wdg_hip_wave2d_set_timing(wdg_wave2d op, int enable)
{
    return guarded([&] {
        if (!op) fail(WDG_ERR_INVALID, "null wave2d handle");
        op->timing = enable != 0;
    });
}

int wdg_hip_wave2d_last_timing(wdg_wave2d op, double* ms_total, int* n_kernels)
{
    return guarded([&] {
        if (!op) fail(WDG_ERR_INVALID, "null wave2d handle");
        if (ms_total) *ms_total = op->last_ms;
        if (n_kernels) *n_kernels = op->last_n;
    });
}

// ---------------------------------------------------------------- WaveHoltz
int wdg_hip_waveholtz_create(wdg_wave2d op, double omega, wdg_waveholtz* out)
{
    return guarded([&] {
        check_op(op);
        if (!out) fail(WDG_ERR_INVALID, "null output handle");
        if (!(omega > 0.0) || !std::isfinite(omega)) fail(WDG_ERR_INVALID, "omega must be positive and finite");
        auto wh = std::make_unique<wdg_waveholtz_s>();
        wh->op = op;
        wh->omega = omega;
        // WaveHoltz.cpp:24-29
        wh->T = 2.0 * M_PI / omega;
        const double h = op->min_h;
        const double p = op->n;
        wh->dt = 0.5 * h / (p * p);
        wh->nt = (int)std::ceil(wh->T / wh->dt);
        wh->dt = wh->T / wh->nt;
        *out = wh.release();
    });
}

int wdg_hip_waveholtz_destroy(wdg_waveholtz wh)
{
    delete wh;
    return WDG_OK;
}

int wdg_hip_waveholtz_info(wdg_waveholtz wh, double* T, double* dt, int* nt)
{
    return guarded([&] {
        if (!wh) fail(WDG_ERR_INVALID, "null waveholtz handle");
        if (T) *T = wh->T;
        if (dt) *dt = wh->dt;
        if (nt) *nt = wh->nt;
    });
}

int wdg_hip_waveholtz_pi0(wdg_waveholtz wh, double* d_w, const double* d_F, void* stream)
{
    return guarded([&] {
        check_wh(wh);
        if (!d_w || !d_F) fail(WDG_ERR_INVALID, "null device pointer");
        wdg_wave2d op = wh->op;
        const hipStream_t s = as_stream(stream);
        const long long n = op->ndof;
        op->ensure(wh->d_p, op->state_bytes());
        const void* force = d_F;
        if (op->prec == WDG_F32) {
            op->ensure(wh->d_F, op->state_bytes());
            hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(256), 0, s, d_F,
                               static_cast<float*>(wh->d_F), n);
            force = wh->d_F;
        }
        zero_any(wh->d_p, op->prec, n, s);
        zero_state(d_w, n, s);
        hip_check(hipGetLastError(), "pi0 init launch");
        double t = 0.0;
        const int nt = wh->nt;
        const double dt = wh->dt;
        auto coef = [wh, nt, dt](int it, double t_after) { return ((it + 1 == nt) ? 0.5 : 1.0) * dt * wh->K(t_after); };
        run_steps(op, dt, t, wh->d_p, nt, force, wh->omega, d_w, coef, s);
    });
}

int wdg_hip_waveholtz_S(wdg_waveholtz wh, double* d_w, void* stream)
{
    return guarded([&] {
        check_wh(wh);
        if (!d_w) fail(WDG_ERR_INVALID, "null device pointer");
        wh_apply_S(wh, d_w, nullptr, as_stream(stream));
    });
}

int wdg_hip_waveholtz_iterate(wdg_waveholtz wh, double* d_w, const double* d_pi0, double pi0_norm,
                              double* rel_change, void* stream)
{
    return guarded([&] {
        check_wh(wh);
        if (!d_w || !d_pi0 || !rel_change) fail(WDG_ERR_INVALID, "null argument");
        if (!(pi0_norm > 0.0)) fail(WDG_ERR_INVALID, "pi0_norm must be positive");
        wdg_wave2d op = wh->op;
        const hipStream_t s = as_stream(stream);
        {
            std::lock_guard<std::mutex> lk(wdg::capture_mutex());
            if (!wh->d_wprev) wh->d_wprev = static_cast<double*>(dmalloc(sizeof(double) * op->ndof));
            op->ensure(op->d_tmp, op->state_bytes());
            op->ensure(op->d_du, op->state_bytes());
            op->ensure(wh->d_p, op->state_bytes());
        }
        // S(W) with W_prev = W, then W += pi0 and the residual's sum of
        // squares (waveholtz2d.cpp:170-175); the same launches eager or as a
        // replayed graph, so both give identical results
        auto enqueue = [&](hipStream_t q) {
            wh_apply_S(wh, d_w, wh->d_wprev, q);
            hipLaunchKernelGGL(k_wh_finalize, dim3(kRedBlocks), dim3(kRedThreads), 0, q, d_w, d_pi0, wh->d_wprev,
                               op->ndof, op->d_partial);
            hipLaunchKernelGGL(k_final, dim3(1), dim3(kRedThreads), 0, q, op->d_partial, kRedBlocks, op->d_result);
            hip_check(hipGetLastError(), "finalize launch");
        };
        if (wh->use_graph) {
            const auto key = std::make_tuple(d_w, d_pi0, op->sched);
            if (!wh->it_graph || wh->it_key != key) {
                std::lock_guard<std::mutex> lk(wdg::capture_mutex());
                if (wh->it_graph) (void)hipGraphExecDestroy(wh->it_graph);
                wh->it_graph = nullptr;
                if (!op->cap_stream)
                    hip_check(hipStreamCreateWithFlags(&op->cap_stream, hipStreamNonBlocking), "hipStreamCreate");
                hipGraph_t g = nullptr;
                hip_check(hipStreamBeginCapture(op->cap_stream, hipStreamCaptureModeThreadLocal),
                          "hipStreamBeginCapture");
                try {
                    enqueue(op->cap_stream);
                } catch (...) {
                    (void)hipStreamEndCapture(op->cap_stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                hip_check(hipStreamEndCapture(op->cap_stream, &g), "hipStreamEndCapture");
                hipGraphExec_t ex = nullptr;
                const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                hip_check(ie, "hipGraphInstantiate");
                wh->it_graph = ex;
                wh->it_key = key;
            }
            hip_check(hipGraphLaunch(wh->it_graph, s), "hipGraphLaunch");
        } else {
            enqueue(s);
        }
        double ss = 0.0;
        hip_check(hipMemcpyAsync(&ss, op->d_result, sizeof(double), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        *rel_change = std::sqrt(ss) / pi0_norm;
        if (!std::isfinite(*rel_change)) fail(WDG_ERR_NONFINITE, "WaveHoltz residual is not finite (check dt / CFL)");
    });
}

int wdg_hip_waveholtz_set_graph(wdg_waveholtz wh, int enable)
{
    return guarded([&] {
        check_wh(wh);
        wh->use_graph = enable != 0;
    });
}

int wdg_hip_waveholtz_solve(wdg_waveholtz wh, wdg_waveholtz wh_refine, double* d_w, const double* d_pi0,
                            double pi0_norm, double tol, int max_iter, int* iters, double* final_residual,
                            void* stream)
{
    return guarded([&] {
        check_wh(wh);
        if (!d_w || !d_pi0 || !iters || !final_residual) fail(WDG_ERR_INVALID, "null argument");
        if (!(tol > 0.0) || max_iter < 1) fail(WDG_ERR_INVALID, "tol must be > 0 and max_iter >= 1");
        if (wh_refine) {
            check_wh(wh_refine);
            if (wh_refine->op->ndof != wh->op->ndof || wh_refine->nt != wh->nt || wh_refine->omega != wh->omega)
                fail(WDG_ERR_INVALID, "refinement handle must discretise the same problem");
        }
        iters[0] = iters[1] = 0;
        double r = 0.0, best = INFINITY;
        int since_best = 0;
        int rc = WDG_OK;
        // first phase: stop at tol, or (with a refinement handle) when the
        // residual has not improved on its best value for 3 iterations
        while (iters[0] + iters[1] < max_iter) {
            rc = wdg_hip_waveholtz_iterate(wh, d_w, d_pi0, pi0_norm, &r, stream);
            if (rc != WDG_OK) fail(rc, g_last_error);
            ++iters[0];
            if (r <= tol) break;
            if (r < 0.97 * best) {
                best = r;
                since_best = 0;
            } else if (++since_best >= 3 && wh_refine) {
                break;
            }
        }
        if (wh_refine && r > tol) {
            while (iters[0] + iters[1] < max_iter) {
                rc = wdg_hip_waveholtz_iterate(wh_refine, d_w, d_pi0, pi0_norm, &r, stream);
                if (rc != WDG_OK) fail(rc, g_last_error);
                ++iters[1];
                if (r <= tol) break;
            }
        }
        *final_residual = r;
    });
}

int wdg_hip_waveholtz_postprocess(wdg_waveholtz wh, double* d_H, const double* d_w, void* stream)
{
    return guarded([&] {
        check_wh(wh);
        if (!d_H || !d_w) fail(WDG_ERR_INVALID, "null device pointer");
        wdg_wave2d op = wh->op;
        const hipStream_t s = as_stream(stream);
        const long long n = op->ndof;
        op->ensure(op->d_r2, op->state_bytes());
        const void* x = d_w;
        if (op->prec == WDG_F32) {
            op->ensure(op->d_r1, op->state_bytes());
            hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(256), 0, s, d_w,
                               static_cast<float*>(op->d_r1), n);
            x = op->d_r1;
        }
        KArgs a = op->base_args();
        a.u_in = x;
        a.u_out = op->d_r2;
        apply_op(op, wdg::K_RHS, a, s);
        if (op->prec == WDG_F32)
            hipLaunchKernelGGL(k_post<float>, dim3(grid_for(n / 3)), dim3(256), 0, s, d_w,
                               static_cast<const float*>(op->d_r2), d_H, wh->omega, n / 3);
        else
            hipLaunchKernelGGL(k_post<double>, dim3(grid_for(n / 3)), dim3(256), 0, s, d_w,
                               static_cast<const double*>(op->d_r2), d_H, wh->omega, n / 3);
        hip_check(hipGetLastError(), "postprocess launch");
    });
}

// ---------------------------------------------------------------- load vector
int wdg_hip_load_vector_gauss(wdg_wave2d op, int n_sources, const double* sources, int var, double* d_F,
                              void* stream)
{
    return guarded([&] {
        check_op(op);
        if (!d_F || (n_sources > 0 && !sources)) fail(WDG_ERR_INVALID, "null argument");
        if (n_sources < 0 || var < 0 || var > 2) fail(WDG_ERR_INVALID, "bad n_sources or var");
        for (int k = 0; k < n_sources; ++k)
            if (!(sources[4 * k + 2] > 0.0)) fail(WDG_ERR_INVALID, "source sigma must be positive");
        const hipStream_t s = as_stream(stream);
        double* d_src = static_cast<double*>(dmalloc(sizeof(double) * 4 * (n_sources > 0 ? n_sources : 1)));
        if (n_sources > 0)
            hip_check(hipMemcpy(d_src, sources, sizeof(double) * 4 * n_sources, hipMemcpyHostToDevice), "upload sources");
        switch (op->n) {
        case 2: wdg::launch_load_vector_n2(op->d_corners, op->d_lvtab, d_src, n_sources, var, d_F, op->nel, s); break;
        case 3: wdg::launch_load_vector_n3(op->d_corners, op->d_lvtab, d_src, n_sources, var, d_F, op->nel, s); break;
        case 4: wdg::launch_load_vector_n4(op->d_corners, op->d_lvtab, d_src, n_sources, var, d_F, op->nel, s); break;
        case 5: wdg::launch_load_vector_n5(op->d_corners, op->d_lvtab, d_src, n_sources, var, d_F, op->nel, s); break;
        case 6: wdg::launch_load_vector_n6(op->d_corners, op->d_lvtab, d_src, n_sources, var, d_F, op->nel, s); break;
        case 7: wdg::launch_load_vector_n7(op->d_corners, op->d_lvtab, d_src, n_sources, var, d_F, op->nel, s); break;
        case 8: wdg::launch_load_vector_n8(op->d_corners, op->d_lvtab, d_src, n_sources, var, d_F, op->nel, s); break;
        case 9: wdg::launch_load_vector_n9(op->d_corners, op->d_lvtab, d_src, n_sources, var, d_F, op->nel, s); break;
        default: break;
        }
        const hipError_t le = hipGetLastError();
        (void)hipStreamSynchronize(s);
        (void)hipFree(d_src);
        hip_check(le, "load vector launch");
    });
}

// ------------------------------------------------------------ ode updates
// The vector updates of ode::ForwardEuler, RungeKutta2, RungeKutta4 and
// SSPRK3 (ode.hpp:12-186) between right-hand-side evaluations; w is the
// second register (RK4 du, RK2 / SSPRK3 y).
__global__ void k_ode_update(int method, int stage, long long n, double dt, double a, double b, double* u, double* w,
                             const double* p)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    switch (method) {
    case 0:  // ForwardEuler
        u[i] += dt * p[i];
        break;
    case 1: {  // RungeKutta4, stage `stage` with a = rk4a[stage], b = rk4b[stage]
        const double d = (stage == 0 ? 0.0 : a * w[i]) + dt * p[i];
        w[i] = d;
        u[i] += b * d;
        break;
    }
    case 2:  // RungeKutta2 (midpoint)
        if (stage == 0)
            w[i] = u[i] + (0.5 * dt) * p[i];
        else
            u[i] += dt * p[i];
        break;
    default:  // SSPRK3
        if (stage == 0)
            w[i] = u[i] + dt * p[i];
        else if (stage == 1)
            w[i] = 0.25 * (w[i] + 3.0 * u[i] + dt * p[i]);
        else
            u[i] = (u[i] + 2.0 * w[i] + 2.0 * dt * p[i]) / 3.0;
        break;
    }
}

int wdg_hip_ode_info(int method, int* n_stages, double* c)
{
    return guarded([&] {
        if (method < 0 || method > 3) fail(WDG_ERR_INVALID, "method: 0 ForwardEuler, 1 RK4, 2 RK2, 3 SSPRK3");
        static const int ns[4] = {1, 5, 2, 3};
        if (n_stages) *n_stages = ns[method];
        if (!c) return;
        for (int k = 0; k < ns[method]; ++k) {
            if (method == 1)
                c[k] = RK4C[k];
            else if (method == 2)
                c[k] = 0.5 * k;
            else if (method == 3)
                c[k] = k == 0 ? 0.0 : (k == 1 ? 1.0 : 0.5);
            else
                c[k] = 0.0;
        }
    });
}

int wdg_hip_ode_update(int method, int stage, long long n, double dt, double* d_u, double* d_w, const double* d_p,
                       void* stream)
{
    return guarded([&] {
        static const int n_stages[4] = {1, 5, 2, 3};
        if (method < 0 || method > 3) fail(WDG_ERR_INVALID, "method: 0 ForwardEuler, 1 RK4, 2 RK2, 3 SSPRK3");
        if (stage < 0 || stage >= n_stages[method]) fail(WDG_ERR_INVALID, "stage out of range for the method");
        if (n < 0 || !d_u || !d_p || (method != 0 && !d_w)) fail(WDG_ERR_INVALID, "bad argument");
        if (n == 0) return;
        const double a = method == 1 ? RK4A[stage] : 0.0, b = method == 1 ? RK4B[stage] : 0.0;
        hipLaunchKernelGGL(k_ode_update, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), method, stage, n, dt, a,
                           b, d_u, d_w, d_p);
        hip_check(hipGetLastError(), "ode update launch");
    });
}

// ---------------------------------------------------------------- reductions
int wdg_hip_norm(const double* d_x, long long n, double* out, void* stream)
{
    return guarded([&] {
        if (!d_x || !out || n < 0) fail(WDG_ERR_INVALID, "bad argument");
        double* buf = static_cast<double*>(dmalloc(sizeof(double) * (kRedBlocks + 1)));
        double v = 0.0;
        try {
            v = reduce_sumsq(d_x, nullptr, n, buf, buf + kRedBlocks, as_stream(stream));
        } catch (...) {
            (void)hipFree(buf);
            throw;
        }
        (void)hipFree(buf);
        *out = std::sqrt(v);
        if (!std::isfinite(*out)) fail(WDG_ERR_NONFINITE, "norm is not finite");
    });
}

int wdg_hip_error(const double* d_x, const double* d_y, long long n, double* out, void* stream)
{
    return guarded([&] {
        if (!d_x || !d_y || !out || n < 0) fail(WDG_ERR_INVALID, "bad argument");
        double* buf = static_cast<double*>(dmalloc(sizeof(double) * (kRedBlocks + 1)));
        double v = 0.0;
        try {
            v = reduce_sumsq(d_x, d_y, n, buf, buf + kRedBlocks, as_stream(stream));
        } catch (...) {
            (void)hipFree(buf);
            throw;
        }
        (void)hipFree(buf);
        *out = std::sqrt(v);
        if (!std::isfinite(*out)) fail(WDG_ERR_NONFINITE, "error norm is not finite");
    });
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free residual
// Device scratch of wdg_hip_error_noalloc: the kRedBlocks partial sums of
// k_sumsq, then the total.  The reduction strides over any n with a fixed
// grid, so this fixed size (8,200 bytes) serves every length.  Each GPU holds
// its own copy until the process ends; in host code `scratch` names a host
// stand-in, so kernels get the address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[kRedBlocks + 1];

namespace {
constexpr int kScratchDevices = 64;
std::once_flag g_scratch_once[kScratchDevices];
double* g_scratch_ptr[kScratchDevices] = {};
hipError_t g_scratch_err[kScratchDevices] = {};

// This device's `scratch`, looked up once per device with that device current.
double* device_scratch()
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kScratchDevices) fail(WDG_ERR_INVALID, "device index beyond the scratch table");
    std::call_once(g_scratch_once[dev], [dev] {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g_scratch_err[dev] == hipSuccess) g_scratch_ptr[dev] = static_cast<double*>(p);
    });
    hip_check(g_scratch_err[dev], "hipGetSymbolAddress(scratch)");
    if (!g_scratch_ptr[dev]) fail(WDG_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int wdg_hip_error_noalloc(const double* d_x, const double* d_y, long long n, double* out, void* stream)
{
    return guarded([&] {
        if (!d_x || !d_y || !out || n < 0) fail(WDG_ERR_INVALID, "bad argument");
        double* buf = device_scratch();
        // k_sumsq writes every one of the kRedBlocks partials and k_final the
        // total, so nothing left by an earlier call is read.
        const double v = reduce_sumsq(d_x, d_y, n, buf, buf + kRedBlocks, as_stream(stream));
        *out = std::sqrt(v);
        if (!std::isfinite(*out)) fail(WDG_ERR_NONFINITE, "error norm is not finite");
    });
}

int wdg_hip_convert(const void* d_src, int src_precision, void* d_dst, int dst_precision, long long n, void* stream)
{
    return guarded([&] {
        if (!d_src || !d_dst || n < 0) fail(WDG_ERR_INVALID, "bad argument");
        const hipStream_t s = as_stream(stream);
        const unsigned g = grid_for(n);
        if (src_precision == WDG_F64 && dst_precision == WDG_F32)
            hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(256), 0, s, static_cast<const double*>(d_src),
                               static_cast<float*>(d_dst), n);
        else if (src_precision == WDG_F32 && dst_precision == WDG_F64)
            hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(256), 0, s, static_cast<const float*>(d_src),
                               static_cast<double*>(d_dst), n);
        else if (src_precision == dst_precision && (src_precision == WDG_F64 || src_precision == WDG_F32))
            hip_check(hipMemcpyAsync(d_dst, d_src, (size_t)n * (src_precision == WDG_F32 ? 4 : 8),
                                     hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
        else
            fail(WDG_ERR_INVALID, "precision must be WDG_F64 or WDG_F32");
        hip_check(hipGetLastError(), "convert launch");
    });
}


int wdg_hip_divf2d_linear(wdg_wave2d op, int n_var, const double* A, const double* B, const void* d_u, void* d_divF,
                          void* stream)
{
    return guarded([&] {
        check_op(op);
        if (n_var < 1 || n_var > wdg::DivFArgs::kMaxVar) fail(WDG_ERR_INVALID, "n_var out of range 1..8");
        if (op->exact) fail(WDG_ERR_INVALID, "DivF2D is implemented for approx_quad = true handles");
        if (!A || !B || !d_u || !d_divF) fail(WDG_ERR_INVALID, "null argument");
        wdg::DivFArgs a{};
        a.geom = op->d_geom;
        a.basis = op->d_basis;
        a.u = d_u;
        a.out = d_divF;
        a.nel = op->nel;
        a.nv = n_var;
        for (int c = 0; c < n_var; ++c)
            for (int r = 0; r < n_var; ++r) {
                a.A[r + wdg::DivFArgs::kMaxVar * c] = A[r + n_var * c];
                a.B[r + wdg::DivFArgs::kMaxVar * c] = B[r + n_var * c];
            }
        wdg::launch_divf_linear(op->n, op->prec == WDG_F32, a, as_stream(stream));
        hip_check(hipGetLastError(), "divf kernel launch");
    });
}

// ------------------------------------------------------------ linear functional
int wdg_hip_mesh_gl_points(wdg_mesh mesh, int nq, double* out)
{
    return guarded([&] {
        if (!mesh || !out) fail(WDG_ERR_INVALID, "null argument");
        std::vector<double> x, w;
        if (!wdg::gauss_legendre_rule(nq, x, w)) fail(WDG_ERR_INVALID, "nq out of range 1..20");
        mesh->mesh.nodes(nq, x.data(), out);
    });
}

int wdg_hip_linear_functional(wdg_wave2d op, int n_var, int nq, const double* d_fq, double* d_F, void* stream)
{
    return guarded([&] {
        check_op(op);
        if (!d_fq || !d_F) fail(WDG_ERR_INVALID, "null device pointer");
        if (n_var < 1 || n_var > wdg::DivFArgs::kMaxVar) fail(WDG_ERR_INVALID, "n_var out of range 1..8");
        if (nq <= 0) nq = op->n;  // the reference's default rule: Gauss-Legendre of n_colloc points
        if (nq > 10) fail(WDG_ERR_INVALID, "nq out of range 1..10");  // the kernel's tables hold 10 points
        auto it = op->d_lftab.find(nq);
        if (it == op->d_lftab.end()) {
            // the rule's table, built and uploaded (synchronously) on the first call with this nq and
            // kept on the handle, so later calls only enqueue the kernel
            std::vector<double> xq, wq;
            if (!wdg::gauss_legendre_rule(nq, xq, wq)) fail(WDG_ERR_INVALID, "nq out of range 1..10");
            std::vector<double> tab(2 * (size_t)nq + (size_t)nq * op->n);
            std::copy(xq.begin(), xq.end(), tab.begin());
            std::copy(wq.begin(), wq.end(), tab.begin() + nq);
            wdg::lagrange_interpolation_matrix(op->n, op->x.data(), nq, xq.data(), tab.data() + 2 * nq);
            double* d = static_cast<double*>(dmalloc(sizeof(double) * tab.size()));
            const hipError_t e = hipMemcpy(d, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                dfree(d);
                hip_check(e, "upload quadrature table");
            }
            it = op->d_lftab.emplace(nq, d).first;
        }
        const hipStream_t s = as_stream(stream);
        wdg::launch_linear_functional(op->n, nq, n_var, op->d_corners, it->second, d_fq, d_F, op->nel, s);
        hip_check(hipGetLastError(), "linear functional launch");
    });
}

// ------------------------------------------------------------ Euler flux
int wdg_hip_divf2d_euler(wdg_wave2d op, double gamma, const void* d_q, void* d_divF, void* stream)
{
    return guarded([&] {
        check_op(op);
        if (op->exact || op->prec != WDG_F64) fail(WDG_ERR_INVALID, "needs an FP64 approx_quad = true wave handle");
        if (!d_q || !d_divF) fail(WDG_ERR_INVALID, "null device pointer");
        if (!(gamma > 1.0)) fail(WDG_ERR_INVALID, "gamma must exceed 1");
        wdg::DivFArgs a{};
        a.geom = op->d_geom;
        a.basis = op->d_basis;
        a.u = d_q;
        a.out = d_divF;
        a.nel = op->nel;
        a.nv = 4;
        a.kind = 1;
        a.gamma = gamma;
        wdg::launch_divf_linear(op->n, 0, a, as_stream(stream));
        hip_check(hipGetLastError(), "divf kernel launch");
    });
}

int wdg_hip_edgefluxf2d_euler(wdg_wave2d op, double gamma, int face_type, const void* d_q, void* d_divF,
                              void* stream)
{
    return guarded([&] {
        check_op(op);
        if (op->exact || op->prec != WDG_F64) fail(WDG_ERR_INVALID, "needs an FP64 approx_quad = true wave handle");
        if (!d_q || !d_divF) fail(WDG_ERR_INVALID, "null device pointer");
        if (!(gamma > 1.0)) fail(WDG_ERR_INVALID, "gamma must exceed 1");
        if (face_type != WDG_INTERIOR && face_type != WDG_BOUNDARY) fail(WDG_ERR_INVALID, "bad face type");
        wdg::launch_euler_face(op->n, gamma, static_cast<const double*>(d_q), static_cast<double*>(d_divF),
                               op->d_geom, op->d_nbr, op->d_info, op->d_basis, op->face_interp(), op->nel,
                               face_type == WDG_INTERIOR ? 1 : 0, face_type == WDG_BOUNDARY ? 1 : 0,
                               as_stream(stream));
        hip_check(hipGetLastError(), "edge flux kernel launch");
    });
}

// ------------------------------------------------------------ advection
struct wdg_advection_s {
    wdg_wave2d op = nullptr;
    int nv = 0, cc = 1;
    double A0[64] = {}, A1[64] = {};  // constant coefficient, padded to DivFArgs::kMaxVar
    double* d_Anode = nullptr;        // variable coefficient (nv, nv, 2, n, n, nel)
    double* d_H = nullptr;            // face matrices (2, nv, nv) per element, side, node
    // approx_quad = false: Gauss-Legendre rules and their tables
    bool exact = false;
    int nqv = 0, nqe = 0;
    double* d_op = nullptr;    // (2, nv, nv, nqv, nqv, nel) metric-weighted coefficient
    double* d_tabv = nullptr;  // P, D (nqv, n)
    double* d_tabe = nullptr;  // P (nqe, n)
    ~wdg_advection_s()
    {
        dfree(d_Anode);
        dfree(d_H);
        dfree(d_op);
        dfree(d_tabv);
        dfree(d_tabe);
    }
};

int wdg_hip_advection_create(wdg_wave2d op, int n_var, const double* A, int constant_coefficient,
                             wdg_advection* out)
{
    return guarded([&] {
        check_op(op);
        if (!A || !out) fail(WDG_ERR_INVALID, "null argument");
        if (n_var < 1 || n_var > wdg::DivFArgs::kMaxVar) fail(WDG_ERR_INVALID, "n_var out of range 1..8");
        if (op->exact || op->prec != WDG_F64)
            fail(WDG_ERR_INVALID, "advection needs an FP64 approx_quad = true wave handle");
        const int N = op->n, NN = N * N, nel = op->nel, m = n_var, v2 = m * m;
        auto adv = std::make_unique<wdg_advection_s>();
        adv->op = op;
        adv->nv = m;
        adv->cc = constant_coefficient ? 1 : 0;
        std::vector<double> geom(6 * (size_t)nel);
        std::vector<int32_t> nbr(4 * (size_t)nel), info(4 * (size_t)nel);
        hip_check(hipMemcpy(geom.data(), op->d_geom, sizeof(double) * geom.size(), hipMemcpyDeviceToHost), "geometry");
        hip_check(hipMemcpy(nbr.data(), op->d_nbr, sizeof(int32_t) * nbr.size(), hipMemcpyDeviceToHost), "nbr");
        hip_check(hipMemcpy(info.data(), op->d_info, sizeof(int32_t) * info.size(), hipMemcpyDeviceToHost), "info");
        auto side_node = [&](int s, int k) {
            const int i = (s == 0 || s == 2) ? k : (s == 1 ? N - 1 : 0);
            const int j = (s == 1 || s == 3) ? k : (s == 2 ? N - 1 : 0);
            return i + N * j;
        };
        // the coefficient on side s of element e at side node k: the node's
        // value, or for a Gauss-Legendre basis the interpolation along the
        // element, as the reference prolongs it (Advection.hpp:77-78)
        std::vector<double> Pf(N, 0.0), Aface(2 * v2);
        if (op->leg) {
            const double xm1 = -1.0;
            wdg::lagrange_interpolation_matrix(N, op->x.data(), 1, &xm1, Pf.data());
        }
        auto coef = [&](int e, int s, int k) -> const double* {
            if (adv->cc) return A;
            const int node = side_node(s, k);
            if (!op->leg) return A + (size_t)2 * v2 * (node + (size_t)NN * e);
            const int st = s == 0 ? N : s == 1 ? -1 : s == 2 ? -N : 1;
            for (int q = 0; q < 2 * v2; ++q) {
                double t = 0.0;
                for (int mm = 0; mm < N; ++mm) t += A[(size_t)2 * v2 * (node + mm * st + (size_t)NN * e) + q] * Pf[mm];
                Aface[q] = t;
            }
            return Aface.data();
        };
        // H_own = w ds (a/2 C + b |C|), H_other = w ds (a/2 C' - b |C'|), a = -1, b = -0.5,
        // C = n.A at this node, C' = n.A at the other side's node, n this element's outward normal
        std::vector<double> H((size_t)nel * 4 * N * 2 * v2, 0.0);
        double C[64], absC[64];
        auto normal_matrix = [&](const double* Ak, double n0, double n1) {
            for (int q = 0; q < v2; ++q) C[q] = n0 * Ak[q] + n1 * Ak[v2 + q];
            abs_normal_matrix(m, C, absC, "at a face node");
        };
        for (int e = 0; e < nel; ++e) {
            const double* g = &geom[6 * (size_t)e];
            for (int s = 0; s < 4; ++s) {
                const double sh = (s == 0 || s == 3) ? -1.0 : 1.0;
                const double dx = 2.0 * ((s & 1) ? g[2] + sh * g[4] : g[0] + sh * g[4]);
                const double dy = 2.0 * ((s & 1) ? g[3] + sh * g[5] : g[1] + sh * g[5]);
                const double len = std::sqrt(dx * dx + dy * dy), sg = (s >= 2) ? -1.0 : 1.0;
                const double n0 = sg * dy / len, n1 = -sg * dx / len, ds = 0.5 * len;
                const int nb = nbr[4 * (size_t)e + s], inf = info[4 * (size_t)e + s];
                for (int k = 0; k < N; ++k) {
                    double* Hk = &H[(((size_t)e * 4 + s) * N + k) * 2 * v2];
                    const double wds = op->w[k] * ds;
                    normal_matrix(coef(e, s, k), n0, n1);
                    for (int q = 0; q < v2; ++q) Hk[q] = wds * (-0.5 * C[q] - 0.5 * absC[q]);
                    if (nb >= 0) {
                        const int kk = ((inf >> 2) & 1) ? (N - 1 - k) : k;
                        normal_matrix(coef(nb, inf & 3, kk), n0, n1);
                        for (int q = 0; q < v2; ++q) Hk[v2 + q] = wds * (-0.5 * C[q] + 0.5 * absC[q]);
                    }
                }
            }
        }
        adv->d_H = static_cast<double*>(dmalloc(sizeof(double) * H.size()));
        hip_check(hipMemcpy(adv->d_H, H.data(), sizeof(double) * H.size(), hipMemcpyHostToDevice), "upload H");
        if (adv->cc) {
            for (int c = 0; c < m; ++c)
                for (int r = 0; r < m; ++r) {
                    adv->A0[r + wdg::DivFArgs::kMaxVar * c] = A[r + m * c];
                    adv->A1[r + wdg::DivFArgs::kMaxVar * c] = A[v2 + r + m * c];
                }
        } else {
            const size_t na = (size_t)2 * v2 * NN * nel;
            adv->d_Anode = static_cast<double*>(dmalloc(sizeof(double) * na));
            hip_check(hipMemcpy(adv->d_Anode, A, sizeof(double) * na, hipMemcpyHostToDevice), "upload A");
        }
        *out = adv.release();
    });
}

int wdg_hip_advection_create_exact(wdg_wave2d op, int n_var, const double* A, int constant_coefficient,
                                   int nq_volume, int nq_edge, wdg_advection* out)
{
    return guarded([&] {
        check_op(op);
        if (!A || !out) fail(WDG_ERR_INVALID, "null argument");
        if (n_var < 1 || n_var > wdg::DivFArgs::kMaxVar) fail(WDG_ERR_INVALID, "n_var out of range 1..8");
        if (op->prec != WDG_F64) fail(WDG_ERR_INVALID, "advection needs an FP64 wave handle");
        const int N = op->n, NN = N * N, nel = op->nel, m = n_var, v2 = m * m;
        const bool cc = constant_coefficient != 0;
        // the reference's default rules for bilinear elements (Div.cpp:319-327, EdgeFlux.cpp:503-511)
        if (nq_volume <= 0) nq_volume = 1 + ((N - 2) + (N - 1) + 1 + (cc ? 0 : N - 1)) / 2;
        if (nq_edge <= 0) nq_edge = 1 + (2 * (N - 1) + 2 + (cc ? 0 : N - 1)) / 2;
        if (nq_volume > 20 || nq_edge > 20) fail(WDG_ERR_INVALID, "quadrature rules have at most 20 points");
        if (wdg::advect_volume_q_lds(N, nq_volume, m) > 65536 || wdg::advect_face_q_lds(N, nq_edge, m) > 65536)
            fail(WDG_ERR_INVALID, "rules too large for n_var (local memory of one workgroup)");
        std::vector<double> xv, wv, xe, we;
        wdg::gauss_legendre_rule(nq_volume, xv, wv);
        wdg::gauss_legendre_rule(nq_edge, xe, we);
        std::vector<double> tabv(2 * (size_t)nq_volume * N), tabe((size_t)nq_edge * N), Pf(N, 0.0);
        wdg::lagrange_interpolation_matrix(N, op->x.data(), nq_volume, xv.data(), tabv.data());
        wdg::lagrange_derivative_matrix(N, op->x.data(), nq_volume, xv.data(), tabv.data() + nq_volume * N);
        wdg::lagrange_interpolation_matrix(N, op->x.data(), nq_edge, xe.data(), tabe.data());
        if (op->leg) {
            const double xm1 = -1.0;
            wdg::lagrange_interpolation_matrix(N, op->x.data(), 1, &xm1, Pf.data());
        }
        std::vector<double> geom(6 * (size_t)nel);
        std::vector<int32_t> nbr(4 * (size_t)nel), info(4 * (size_t)nel);
        hip_check(hipMemcpy(geom.data(), op->d_geom, sizeof(double) * geom.size(), hipMemcpyDeviceToHost), "geometry");
        hip_check(hipMemcpy(nbr.data(), op->d_nbr, sizeof(int32_t) * nbr.size(), hipMemcpyDeviceToHost), "nbr");
        hip_check(hipMemcpy(info.data(), op->d_info, sizeof(int32_t) * info.size(), hipMemcpyDeviceToHost), "info");
        auto node_coef = [&](int e, int node) { return cc ? A : A + (size_t)2 * v2 * (node + (size_t)NN * e); };
        // volume: op(k, d, c, i, j, el) = w_i w_j (metric terms) A at the point (Div.cpp:200-310)
        const int nqv = nq_volume, nqe = nq_edge;
        std::vector<double> opt((size_t)2 * v2 * nqv * nqv * nel);
        std::vector<double> Aq(2 * v2);
        for (int e = 0; e < nel; ++e) {
            const double* g = &geom[6 * (size_t)e];
            for (int j = 0; j < nqv; ++j)
                for (int i = 0; i < nqv; ++i) {
                    for (int t = 0; t < 2 * v2; ++t) {
                        if (cc) {
                            Aq[t] = A[t];
                            continue;
                        }
                        double a = 0.0;
                        for (int l = 0; l < N; ++l)
                            for (int k = 0; k < N; ++k)
                                a += tabv[i + nqv * k] * tabv[j + nqv * l] * node_coef(e, k + N * l)[t];
                        Aq[t] = a;
                    }
                    const double xi = xv[i], eta = xv[j], W = wv[i] * wv[j];
                    const double Xx = (g[0] + g[4] * eta) * W, Yx = (g[1] + g[5] * eta) * W;
                    const double Xe = (g[2] + g[4] * xi) * W, Ye = (g[3] + g[5] * xi) * W;
                    double* o = &opt[(size_t)2 * v2 * (i + nqv * (j + (size_t)nqv * e))];
                    for (int c = 0; c < m; ++c)
                        for (int d = 0; d < m; ++d) {
                            const double ax = Aq[d + m * c], ay = Aq[v2 + d + m * c];
                            o[2 * (d + m * c)] = Ye * ax - Xe * ay;
                            o[2 * (d + m * c) + 1] = -Yx * ax + Xx * ay;
                        }
                }
        }
        // faces: H_own = w ds (a/2 C + b|C|), H_other = w ds (a/2 C' - b|C'|),
        // a = -1, b = -1/2, C = n.A on this side and C' on the other at the
        // edge point, n this element's outward normal (EdgeFlux.cpp:375-490)
        auto side_node = [&](int sd, int k) {
            const int i = (sd == 0 || sd == 2) ? k : (sd == 1 ? N - 1 : 0);
            const int j = (sd == 1 || sd == 3) ? k : (sd == 2 ? N - 1 : 0);
            return i + N * j;
        };
        auto trace_coef = [&](int e, int sd, int k, double* out2) {  // 2 v2 values on side sd at side node k
            if (cc) {
                for (int t = 0; t < 2 * v2; ++t) out2[t] = A[t];
                return;
            }
            const int node = side_node(sd, k);
            if (!op->leg) {
                for (int t = 0; t < 2 * v2; ++t) out2[t] = node_coef(e, node)[t];
                return;
            }
            const int st = sd == 0 ? N : sd == 1 ? -1 : sd == 2 ? -N : 1;
            for (int t = 0; t < 2 * v2; ++t) {
                double a = 0.0;
                for (int mm = 0; mm < N; ++mm) a += node_coef(e, node + mm * st)[t] * Pf[mm];
                out2[t] = a;
            }
        };
        std::vector<double> Hq((size_t)nel * 4 * nqe * 2 * v2, 0.0), tro((size_t)N * 2 * v2), trb((size_t)N * 2 * v2);
        double C[64], absC[64], Af[128];
        auto normal_matrix = [&](const double* Ak, double n0, double n1) {
            for (int q = 0; q < v2; ++q) C[q] = n0 * Ak[q] + n1 * Ak[v2 + q];
            abs_normal_matrix(m, C, absC, "at an edge point");
        };
        for (int e = 0; e < nel; ++e) {
            const double* g = &geom[6 * (size_t)e];
            for (int sd = 0; sd < 4; ++sd) {
                const double sh = (sd == 0 || sd == 3) ? -1.0 : 1.0;
                const double dx = 2.0 * ((sd & 1) ? g[2] + sh * g[4] : g[0] + sh * g[4]);
                const double dy = 2.0 * ((sd & 1) ? g[3] + sh * g[5] : g[1] + sh * g[5]);
                const double len = std::sqrt(dx * dx + dy * dy), sg = (sd >= 2) ? -1.0 : 1.0;
                const double n0 = sg * dy / len, n1 = -sg * dx / len, ds = 0.5 * len;
                const int nb = nbr[4 * (size_t)e + sd], inf = info[4 * (size_t)e + sd];
                for (int k = 0; k < N; ++k) {
                    trace_coef(e, sd, k, &tro[(size_t)2 * v2 * k]);
                    if (nb >= 0) trace_coef(nb, inf & 3, ((inf >> 2) & 1) ? (N - 1 - k) : k, &trb[(size_t)2 * v2 * k]);
                }
                for (int q = 0; q < nqe; ++q) {
                    double* H = &Hq[(((size_t)e * 4 + sd) * nqe + q) * 2 * v2];
                    const double wds = we[q] * ds;
                    for (int t = 0; t < 2 * v2; ++t) {
                        double a = 0.0;
                        for (int k = 0; k < N; ++k) a += tabe[q + nqe * k] * tro[(size_t)2 * v2 * k + t];
                        Af[t] = a;
                    }
                    normal_matrix(Af, n0, n1);
                    for (int t = 0; t < v2; ++t) H[t] = wds * (-0.5 * C[t] - 0.5 * absC[t]);
                    if (nb >= 0) {
                        for (int t = 0; t < 2 * v2; ++t) {
                            double a = 0.0;
                            for (int k = 0; k < N; ++k) a += tabe[q + nqe * k] * trb[(size_t)2 * v2 * k + t];
                            Af[t] = a;
                        }
                        normal_matrix(Af, n0, n1);
                        for (int t = 0; t < v2; ++t) H[v2 + t] = wds * (-0.5 * C[t] + 0.5 * absC[t]);
                    }
                }
            }
        }
        auto adv = std::make_unique<wdg_advection_s>();
        adv->op = op;
        adv->nv = m;
        adv->cc = cc ? 1 : 0;
        adv->exact = true;
        adv->nqv = nqv;
        adv->nqe = nqe;
        auto upload = [&](double*& d, const std::vector<double>& h, const char* what) {
            d = static_cast<double*>(dmalloc(sizeof(double) * h.size()));
            hip_check(hipMemcpy(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice), what);
        };
        upload(adv->d_op, opt, "upload volume table");
        upload(adv->d_H, Hq, "upload face table");
        upload(adv->d_tabv, tabv, "upload volume rule");
        upload(adv->d_tabe, tabe, "upload edge rule");
        *out = adv.release();
    });
}

static int advection_launch(wdg_advection adv, const void* d_u, void* d_divF, bool volume, bool interior,
                            bool boundary, void* stream)
{
    return guarded([&] {
        if (!adv) fail(WDG_ERR_INVALID, "null advection handle");
        if (!d_u || !d_divF) fail(WDG_ERR_INVALID, "null device pointer");
        wdg_wave2d op = adv->op;
        check_op(op);
        const hipStream_t s = as_stream(stream);
        if (adv->exact) {
            wdg::launch_advect_q(op->n, adv->nv, adv->nqv, adv->nqe, static_cast<const double*>(d_u),
                                 static_cast<double*>(d_divF), adv->d_op, adv->d_tabv, op->d_nbr, op->d_info,
                                 adv->d_H, adv->d_tabe, op->face_interp(), op->nel, volume ? 1 : 0, interior ? 1 : 0,
                                 boundary ? 1 : 0, s);
            hip_check(hipGetLastError(), "advection kernel launch");
            return;
        }
        if (volume) {
            wdg::DivFArgs a{};
            a.geom = op->d_geom;
            a.basis = op->d_basis;
            a.u = d_u;
            a.out = d_divF;
            a.nel = op->nel;
            a.nv = adv->nv;
            std::memcpy(a.A, adv->A0, sizeof(a.A));
            std::memcpy(a.B, adv->A1, sizeof(a.B));
            a.Anode = adv->d_Anode;
            wdg::launch_divf_linear(op->n, 0, a, s);
        }
        wdg::launch_advect_face(op->n, adv->nv, static_cast<const double*>(d_u), static_cast<double*>(d_divF),
                                op->d_nbr, op->d_info, adv->d_H, op->face_interp(), op->nel, interior ? 1 : 0,
                                boundary ? 1 : 0, s);
        hip_check(hipGetLastError(), "advection kernel launch");
    });
}

int wdg_hip_advection_action(wdg_advection adv, const void* d_u, void* d_divF, void* stream)
{
    return advection_launch(adv, d_u, d_divF, true, true, false, stream);
}

int wdg_hip_advection_bc_action(wdg_advection adv, const void* d_u, void* d_divF, void* stream)
{
    return advection_launch(adv, d_u, d_divF, false, false, true, stream);
}

int wdg_hip_advection_destroy(wdg_advection adv)
{
    delete adv;
    return WDG_OK;
}

// ------------------------------------------------ flux functors (hipRTC)
struct wdg_flux_s {
    wdg_wave2d op = nullptr;  // borrowed: the wave handle outlives the flux
    int nv = 0;
    wdg::RtcFlux* f = nullptr;
    ~wdg_flux_s() { wdg::rtc_flux_destroy(f); }
};

int wdg_hip_flux_create(wdg_wave2d op, int n_var, const char* source, wdg_flux* out)
{
    return guarded([&] {
        check_op(op);
        if (!source || !out) fail(WDG_ERR_INVALID, "null argument");
        if (n_var < 1 || n_var > wdg::DivFArgs::kMaxVar) fail(WDG_ERR_INVALID, "n_var out of range 1..8");
        if (op->exact || op->prec != WDG_F64)
            fail(WDG_ERR_INVALID, "flux functors need an FP64 approx_quad = true wave handle");
        auto fx = std::make_unique<wdg_flux_s>();
        fx->op = op;
        fx->nv = n_var;
        fx->f = wdg::rtc_flux_create(source, op->n, n_var, op->leg);
        *out = fx.release();
    });
}

// Local memory of the run-time compiled quadrature volume kernel (per
// workgroup) fits the 64 KiB a workgroup may hold.
static bool flux_q_fits(int n, int n_var, int nq)
{
    const int sp = n * n_var * nq, sdp = 2 * nq * n_var * n;
    const size_t lds = (size_t)wdg::rtc_divf_q_elems(n, n_var, nq) * (2 * nq * n_var * nq + std::max(sp, sdp)) +
                       2 * nq + 2 * (size_t)nq * n;
    return lds * sizeof(double) <= 65536;
}

int wdg_hip_flux_check(const char* source, int n_colloc, int n_var, int basis_type, int which)
{
    return guarded([&] {
        if (!source) fail(WDG_ERR_INVALID, "null source");
        if (n_colloc < 2 || n_colloc > 9) fail(WDG_ERR_INVALID, "n_colloc out of range 2..9");
        if (n_var < 1 || n_var > wdg::DivFArgs::kMaxVar) fail(WDG_ERR_INVALID, "n_var out of range 1..8");
        if (basis_type != WDG_GAUSS_LOBATTO && basis_type != WDG_GAUSS_LEGENDRE) fail(WDG_ERR_INVALID, "bad basis");
        if (which != 0 && which != 1) fail(WDG_ERR_INVALID, "which must be 0 (flux) or 1 (numerical_flux)");
        std::vector<char> code;
        std::string err;
        for (int nq : {0, 2 * n_colloc}) {  // collocated and quadrature forms
            if (nq && which == 0 && !flux_q_fits(n_colloc, n_var, nq)) continue;  // refused at creation anyway
            const int kind = (which == 0 ? 0 : 1) + (nq ? 2 : 0);
            if (!wdg::rtc_build(source, n_colloc, n_var, basis_type == WDG_GAUSS_LEGENDRE, kind, nq ? nq : n_colloc,
                                code, err))
                fail(WDG_ERR_INVALID, err);
        }
    });
}

int wdg_hip_flux_create_exact(wdg_wave2d op, int n_var, const char* source, int nq_volume, int nq_edge,
                              wdg_flux* out)
{
    return guarded([&] {
        check_op(op);
        if (!source || !out) fail(WDG_ERR_INVALID, "null argument");
        if (n_var < 1 || n_var > wdg::DivFArgs::kMaxVar) fail(WDG_ERR_INVALID, "n_var out of range 1..8");
        if (op->prec != WDG_F64) fail(WDG_ERR_INVALID, "flux functors need an FP64 wave handle");
        const int n = op->n;
        // the reference's default rule, 2 n points (DivF.cpp:56-60, EdgeFluxF.cpp:51-56)
        if (nq_volume <= 0) nq_volume = 2 * n;
        if (nq_edge <= 0) nq_edge = 2 * n;
        if (nq_volume > 20 || nq_edge > 20) fail(WDG_ERR_INVALID, "quadrature rules have at most 20 points");
        if (!flux_q_fits(n, n_var, nq_volume))
            fail(WDG_ERR_INVALID, "volume rule too large for n_var (local memory of one workgroup)");
        auto table = [&](int nq, bool with_d) {
            std::vector<double> xq, wq;
            if (!wdg::gauss_legendre_rule(nq, xq, wq)) fail(WDG_ERR_INVALID, "no Gauss-Legendre rule of that size");
            std::vector<double> t(2 * nq + (with_d ? 2 : 1) * nq * n);
            std::copy(xq.begin(), xq.end(), t.begin());
            std::copy(wq.begin(), wq.end(), t.begin() + nq);
            wdg::lagrange_interpolation_matrix(n, op->x.data(), nq, xq.data(), t.data() + 2 * nq);
            if (with_d) wdg::lagrange_derivative_matrix(n, op->x.data(), nq, xq.data(), t.data() + 2 * nq + nq * n);
            return t;
        };
        const std::vector<double> tv = table(nq_volume, true), te = table(nq_edge, false);
        auto fx = std::make_unique<wdg_flux_s>();
        fx->op = op;
        fx->nv = n_var;
        fx->f = wdg::rtc_flux_create(source, n, n_var, op->leg);
        std::string err;
        if (!wdg::rtc_flux_set_rules(fx->f, nq_volume, tv, nq_edge, te, err)) fail(WDG_ERR_HIP, err);
        *out = fx.release();
    });
}

int wdg_hip_flux_compile(wdg_flux fx, int which)
{
    return guarded([&] {
        if (!fx) fail(WDG_ERR_INVALID, "null flux handle");
        if (which != 0 && which != 1) fail(WDG_ERR_INVALID, "which must be 0 (flux) or 1 (numerical_flux)");
        check_op(fx->op);
        std::string err;
        if (!wdg::rtc_flux_prepare(fx->f, which, err)) fail(WDG_ERR_INVALID, err);
    });
}

int wdg_hip_divf2d(wdg_flux fx, const void* d_u, void* d_divF, void* stream)
{
    return guarded([&] {
        if (!fx) fail(WDG_ERR_INVALID, "null flux handle");
        if (!d_u || !d_divF) fail(WDG_ERR_INVALID, "null device pointer");
        if (d_u == d_divF) fail(WDG_ERR_INVALID, "input and output must not alias");
        wdg_wave2d op = fx->op;
        check_op(op);
        std::string err;
        if (!wdg::rtc_divf(fx->f, static_cast<const double*>(d_u), static_cast<double*>(d_divF), op->d_corners,
                           op->d_basis, op->nel, as_stream(stream), err))
            fail(WDG_ERR_INVALID, err);
        hip_check(hipGetLastError(), "flux divergence kernel");
    });
}

int wdg_hip_edgefluxf2d(wdg_flux fx, int face_type, double scale, const void* d_u, void* d_divF, void* stream)
{
    return guarded([&] {
        if (!fx) fail(WDG_ERR_INVALID, "null flux handle");
        if (!d_u || !d_divF) fail(WDG_ERR_INVALID, "null device pointer");
        if (d_u == d_divF) fail(WDG_ERR_INVALID, "input and output must not alias");
        if (face_type != WDG_INTERIOR && face_type != WDG_BOUNDARY) fail(WDG_ERR_INVALID, "bad face type");
        wdg_wave2d op = fx->op;
        check_op(op);
        std::string err;
        if (!wdg::rtc_edgeflux(fx->f, static_cast<const double*>(d_u), static_cast<double*>(d_divF), op->d_corners,
                               op->d_nbr, op->d_info, op->d_basis, op->nel, face_type == WDG_INTERIOR ? 1 : 0,
                               face_type == WDG_BOUNDARY ? 1 : 0, scale, as_stream(stream), err))
            fail(WDG_ERR_INVALID, err);
        hip_check(hipGetLastError(), "edge flux kernel");
    });
}

int wdg_hip_flux_destroy(wdg_flux fx)
{
    delete fx;
    return WDG_OK;
}

// ------------------------------------------------------------ file formats
int wdg_hip_to_file(const char* path, const void* d_x, long long n, int precision)
{
    return guarded([&] {
        if (!path || (n > 0 && !d_x) || n < 0) fail(WDG_ERR_INVALID, "bad arguments");
        if (precision != WDG_F64 && precision != WDG_F32) fail(WDG_ERR_INVALID, "bad precision");
        FILE* f = std::fopen(path, "wb");
        if (!f) fail(WDG_ERR_INVALID, std::string("cannot open ") + path + " for writing");
        const long long chunk = 1LL << 22;  // values per host staging buffer
        std::vector<double> buf((size_t)std::min(std::max(n, 1LL), chunk));
        std::vector<float> fbuf(precision == WDG_F32 ? buf.size() : 0);
        bool ok = true;
        for (long long o = 0; o < n && ok; o += chunk) {
            const long long m = std::min(chunk, n - o);
            hipError_t e;
            if (precision == WDG_F32) {
                e = hipMemcpy(fbuf.data(), static_cast<const float*>(d_x) + o, sizeof(float) * m, hipMemcpyDeviceToHost);
                for (long long k = 0; k < m; ++k) buf[k] = fbuf[k];
            } else {
                e = hipMemcpy(buf.data(), static_cast<const double*>(d_x) + o, sizeof(double) * m,
                              hipMemcpyDeviceToHost);
            }
            if (e != hipSuccess) {
                std::fclose(f);
                hip_check(e, "copy to host");
            }
            ok = std::fwrite(buf.data(), sizeof(double), (size_t)m, f) == (size_t)m;
        }
        if (std::fclose(f) != 0 || !ok) fail(WDG_ERR_INVALID, std::string("write failed: ") + path);
    });
}

int wdg_hip_from_file(const char* path, void* d_x, long long n, int precision)
{
    return guarded([&] {
        if (!path || (n > 0 && !d_x) || n < 0) fail(WDG_ERR_INVALID, "bad arguments");
        if (precision != WDG_F64 && precision != WDG_F32) fail(WDG_ERR_INVALID, "bad precision");
        FILE* f = std::fopen(path, "rb");
        if (!f) fail(WDG_ERR_INVALID, std::string("cannot open ") + path);
        std::fseek(f, 0, SEEK_END);
        const long long bytes = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (bytes != n * (long long)sizeof(double)) {
            std::fclose(f);
            fail(WDG_ERR_INVALID, std::string(path) + " holds " + std::to_string(bytes / 8) + " values, expected " +
                                      std::to_string(n));
        }
        const long long chunk = 1LL << 22;
        std::vector<double> buf((size_t)std::min(std::max(n, 1LL), chunk));
        std::vector<float> fbuf(precision == WDG_F32 ? buf.size() : 0);
        for (long long o = 0; o < n; o += chunk) {
            const long long m = std::min(chunk, n - o);
            if (std::fread(buf.data(), sizeof(double), (size_t)m, f) != (size_t)m) {
                std::fclose(f);
                fail(WDG_ERR_INVALID, std::string("read failed: ") + path);
            }
            hipError_t e;
            if (precision == WDG_F32) {
                for (long long k = 0; k < m; ++k) fbuf[k] = (float)buf[k];
                e = hipMemcpy(static_cast<float*>(d_x) + o, fbuf.data(), sizeof(float) * m, hipMemcpyHostToDevice);
            } else {
                e = hipMemcpy(static_cast<double*>(d_x) + o, buf.data(), sizeof(double) * m, hipMemcpyHostToDevice);
            }
            if (e != hipSuccess) {
                std::fclose(f);
                hip_check(e, "copy to device");
            }
        }
        std::fclose(f);
    });
}

int wdg_hip_waveholtz_checkpoint(wdg_waveholtz wh, const char* path, const double* d_W, int iterations,
                                 double residual)
{
    return guarded([&] {
        if (!wh || !path || !d_W) fail(WDG_ERR_INVALID, "null argument");
        const wdg_wave2d op = wh->op;
        const int rc = wdg_hip_to_file(path, d_W, op->ndof, WDG_F64);
        if (rc) fail(rc, g_last_error);
        const std::string meta = std::string(path) + ".meta";
        FILE* f = std::fopen(meta.c_str(), "w");
        if (!f) fail(WDG_ERR_INVALID, "cannot open " + meta + " for writing");
        std::fprintf(f, "waveholtz-checkpoint 1\nn_colloc %d\nn_elem %d\nndof %lld\nomega %.17g\n"
                        "iterations %d\nresidual %.17g\n",
                     op->n, op->nel, op->ndof, wh->omega, iterations, residual);
        if (std::fclose(f) != 0) fail(WDG_ERR_INVALID, "write failed: " + meta);
    });
}

int wdg_hip_waveholtz_resume(wdg_waveholtz wh, const char* path, double* d_W, int* iterations, double* residual)
{
    return guarded([&] {
        if (!wh || !path || !d_W) fail(WDG_ERR_INVALID, "null argument");
        const wdg_wave2d op = wh->op;
        const std::string meta = std::string(path) + ".meta";
        std::ifstream f(meta);
        if (!f) fail(WDG_ERR_INVALID, "cannot open " + meta);
        std::map<std::string, std::string> kv;
        std::string k, v;
        while (f >> k >> v) kv[k] = v;
        auto need = [&](const char* key) {
            auto it = kv.find(key);
            if (it == kv.end()) fail(WDG_ERR_INVALID, meta + ": missing " + key);
            return it->second;
        };
        if (need("waveholtz-checkpoint") != "1") fail(WDG_ERR_INVALID, meta + ": unknown version");
        if (std::stoi(need("n_colloc")) != op->n || std::stoi(need("n_elem")) != op->nel ||
            std::stoll(need("ndof")) != op->ndof)
            fail(WDG_ERR_INVALID, meta + ": mesh or order differs from the handle");
        const double om = std::stod(need("omega"));
        if (std::fabs(om - wh->omega) > 1e-12 * std::fabs(wh->omega))
            fail(WDG_ERR_INVALID, meta + ": omega differs from the handle");
        const int rc = wdg_hip_from_file(path, d_W, op->ndof, WDG_F64);
        if (rc) fail(rc, g_last_error);
        if (iterations) *iterations = std::stoi(need("iterations"));
        if (residual) *residual = std::stod(need("residual"));
    });
}

int wdg_hip_schedule_count(int* count, int* threads)
{
    return guarded([&] {
        if (count) *count = wdg::kSchedules;
        if (threads)
            for (int k = 0; k < wdg::kSchedules; ++k) threads[k] = wdg::kSchedTgt[k];
    });
}

int wdg_hip_wave2d_set_schedule(wdg_wave2d op, int schedule)
{
    return guarded([&] {
        check_op(op);
        if (op->exact) fail(WDG_ERR_INVALID, "schedules apply to the approx_quad = true kernels only");
        if (schedule < 0 || schedule >= wdg::kSchedules) fail(WDG_ERR_INVALID, "schedule out of range");
        op->sched = schedule;
        for (auto& kv : op->graphs) (void)hipGraphExecDestroy(kv.second);
        op->graphs.clear();
    });
}

int wdg_hip_wave2d_get_schedule(wdg_wave2d op, int* schedule, int* threads_per_workgroup)
{
    return guarded([&] {
        check_op(op);
        if (schedule) *schedule = op->sched;
        if (threads_per_workgroup) {
            const int nn = op->n * op->n, epb = std::max(1, wdg::kSchedTgt[op->sched] / nn);
            *threads_per_workgroup = epb * nn;
        }
    });
}

int wdg_hip_wave2d_autotune(wdg_wave2d op, int reps, int* schedule, double* ms)
{
    return guarded([&] {
        check_op(op);
        if (op->exact) fail(WDG_ERR_INVALID, "schedules apply to the approx_quad = true kernels only");
        if (reps < 1) fail(WDG_ERR_INVALID, "reps must be positive");
        const hipStream_t s = nullptr;
        const long long n = op->ndof;
        const size_t bytes = op->state_bytes();
        const int f32 = op->prec == WDG_F32;
        // private buffers: input state, RK register, outputs of the default schedule
        DevBuf u(bytes), du(bytes), du_ref(bytes), out(bytes), out_ref(bytes), cnt(sizeof(unsigned int));
        KArgs a = op->base_args();
        a.u_in = u.p;
        a.u_out = out.p;
        a.du = du.p;
        a.rk_a = RK4A[2];  // any stage 1-4 coefficients
        a.rk_b = RK4B[2];
        a.dt = 0.25 * op->min_h / (op->n * op->n);
        std::vector<double> t(wdg::kSchedules, -1.0);
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        int best = 0;
        for (int k = 0; k < wdg::kSchedules; ++k) {
            // parity: one stage from the same input must equal the default schedule bit for bit
            // (the timing below overwrites both states, so refill them)
            fill_hash(op->prec, u.p, n, 11, s);
            fill_hash(op->prec, du.p, n, 23, s);
            launch_wave(op->n, wdg::K_STAGE, f32, k, a, s);
            if (k == 0) {
                hip_check(hipMemcpyAsync(out_ref.p, out.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
                hip_check(hipMemcpyAsync(du_ref.p, du.p, bytes, hipMemcpyDeviceToDevice, s), "copy");
            } else {
                const unsigned d1 = count_diff(op->prec, out.p, out_ref.p, n, static_cast<unsigned*>(cnt.p), s);
                const unsigned d2 = count_diff(op->prec, du.p, du_ref.p, n, static_cast<unsigned*>(cnt.p), s);
                if (d1 || d2) continue;  // not eligible
            }
            // timing: stages ping-pong between two states as in the RK loop;
            // best of three trials of `reps` launches
            KArgs b = a;
            auto flip = [&b] {
                void* in = const_cast<void*>(b.u_in);
                b.u_in = b.u_out;
                b.u_out = in;
            };
            for (int w = 0; w < 2; ++w) {
                launch_wave(op->n, wdg::K_STAGE, f32, k, b, s);
                flip();
            }
            double tk = 1e300;
            for (int trial = 0; trial < 3; ++trial) {
                hip_check(hipEventRecord(e0, s), "hipEventRecord");
                for (int r = 0; r < reps; ++r) {
                    launch_wave(op->n, wdg::K_STAGE, f32, k, b, s);
                    flip();
                }
                hip_check(hipEventRecord(e1, s), "hipEventRecord");
                hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                float el = 0.f;
                hip_check(hipEventElapsedTime(&el, e0, e1), "hipEventElapsedTime");
                tk = std::min(tk, (double)el / reps);
            }
            t[k] = tk;
            if (t[k] < t[best]) best = k;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        op->sched = best;
        for (auto& kv : op->graphs) (void)hipGraphExecDestroy(kv.second);
        op->graphs.clear();
        if (!op->leg) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[{op->arch, op->n, op->prec}] = best;
        }
        if (schedule) *schedule = best;
        if (ms)
            for (int k = 0; k < wdg::kSchedules; ++k) ms[k] = t[k];
    });
}

int wdg_hip_tune_cache_save(const char* path)
{
    return guarded([&] {
        if (!path) fail(WDG_ERR_INVALID, "null path");
        std::ofstream f(path);
        if (!f) fail(WDG_ERR_INVALID, std::string("cannot write ") + path);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        f << "# arch n_colloc precision(0=f64,1=f32) schedule threads\n";
        for (const auto& kv : g_tune)
            f << std::get<0>(kv.first) << ' ' << std::get<1>(kv.first) << ' ' << std::get<2>(kv.first) << ' '
              << kv.second << ' ' << wdg::kSchedTgt[kv.second] << '\n';
    });
}

int wdg_hip_tune_cache_load(const char* path)
{
    return guarded([&] {
        if (!path) fail(WDG_ERR_INVALID, "null path");
        std::ifstream f(path);
        if (!f) fail(WDG_ERR_INVALID, std::string("cannot read ") + path);
        std::map<std::tuple<std::string, int, int>, int> add;
        std::string line;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream is(line);
            std::string arch;
            int n = 0, prec = 0, sched = 0;
            if (!(is >> arch >> n >> prec >> sched) || n < 2 || n > 9 || (prec != WDG_F64 && prec != WDG_F32) ||
                sched < 0 || sched >= wdg::kSchedules)
                fail(WDG_ERR_INVALID, "malformed tune cache line: " + line);
            add[{arch, n, prec}] = sched;
        }
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto& kv : add) g_tune[kv.first] = kv.second;
    });
}

}  // extern "C"
