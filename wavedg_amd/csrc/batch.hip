// Batched WaveHoltz: many independent solves (different meshes, orders and
// frequencies) advanced together, one kernel launch per RK stage for the
// whole batch, with the steps captured in a HIP graph.
//
// Every workgroup belongs to one problem (block table -> problem table) and
// runs the order-N operator of apply_operator on its elements.  A device step
// counter G drives everything else, so the captured graph is replayed
// unchanged:
//   * problem k is in period it = G / nt_k at local step s = G % nt_k; the
//     ping-pong buffer of stage st is buf[(G + st) & 1];
//   * period 0 is WaveHoltz::pi0 (forcing on, zero initial state,
//     WaveHoltz.cpp:55-89); periods it >= 1 apply S (WaveHoltz.cpp:91-121)
//     to the current iterate;
//   * the last stage of a period's last step finishes the period in place:
//     pi0 (period 0) or W <- S(W) + pi0 (examples/waveholtz2d.cpp:170-175),
//     the residual's per-workgroup partial sum, and the next period's start
//     (state = W, accumulator = 1/2 dt K(0) W);
//   * a small kernel per step reduces the partials of problems whose period
//     ended (fixed order, deterministic), records the residual, retires
//     converged problems, and its last workgroup advances G.
// The filter coefficients and forcing scales are tabulated on the host with
// the same arithmetic as the single-problem driver.
#include "../../include/wavedg_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_basis.hpp"
#include "host_mesh.hpp"
#include "host_sync.hpp"
#include "wave2d_kernels.hpp"

namespace {

struct BatchBlock {
    int prob, N, e0, nloc, gid;  // gid: index in the full table (residual partials)
};

struct BatchProb {
    long long soff;  // offset of the problem's state in the concatenated arrays
    long long coff;  // filter-coefficient table offset (nt + 1 entries)
    long long foff;  // forcing-scale table offset (5 nt entries)
    const double* basis;
    double ha, hb;
    int eoff, nel, N, nt, blk_off, nblk;
};

struct BatchArgs {
    const BatchBlock* blocks;
    const BatchProb* probs;
    void* buf0;
    void* buf1;
    void* du;
    const void* F;   // load vectors, state precision
    double* acc;     // filter accumulators
    double* wprev;   // current iterate W_k
    double* pi0;
    const double* winit;
    const int* warm;
    const double* geom;
    const int32_t* nbr;
    const int32_t* info;
    const void* fgeo;  // wdg::FaceGeom<R> per element side
    const wdg::FaceNb* fnb;
    const double* ctab;
    const double* ftab;
    const int* active;
    double* partial;
    const long long* G;
    double rk_a, rk_b;
    int stage;
};

constexpr int kBT = 256;  // threads per workgroup

template <typename R>
constexpr size_t max_op_lds()
{
    size_t m = 0;
    m = std::max(m, wdg::OpLds<2, R, true, true>::bytes);
    m = std::max(m, wdg::OpLds<3, R, true, true>::bytes);
    m = std::max(m, wdg::OpLds<4, R, true, true>::bytes);
    m = std::max(m, wdg::OpLds<5, R, true, true>::bytes);
    m = std::max(m, wdg::OpLds<6, R, true, true>::bytes);
    m = std::max(m, wdg::OpLds<7, R, true, true>::bytes);
    m = std::max(m, wdg::OpLds<8, R, true, true>::bytes);
    m = std::max(m, wdg::OpLds<9, R, true, true>::bytes);
    return m;
}

// KIND 0: RK stage 0, 1: stages 1-3, 2: stage 4 (accumulate, finish periods)
template <int N, typename R, int KIND>
__device__ __forceinline__ void batch_body(const BatchArgs& a, const BatchBlock& blk, const BatchProb& pr, int s,
                                           long long it, long long G, unsigned char* lds)
{
    const int st = a.stage;
    const R* u_in = static_cast<const R*>(((G + st) & 1) ? a.buf1 : a.buf0) + pr.soff;
    R* u_out = static_cast<R*>(((G + st + 1) & 1) ? a.buf1 : a.buf0) + pr.soff;
    R* du = static_cast<R*>(a.du) + pr.soff;
    const bool forced = it == 0;
    const bool period_end = KIND == 2 && s + 1 == pr.nt;

    R pre0 = 0, pre1 = 0, pre2 = 0, f0 = 0, f1 = 0, f2 = 0;
    double ac0 = 0, ac1 = 0, ac2 = 0, x0 = 0, x1 = 0, x2 = 0;
    auto prefetch = [&](bool act, size_t gb, uint32_t g) {
        if (!act) return;
        if constexpr (KIND != 0) {
            const R* dub = du + gb;
            pre0 = dub[g + 0];
            pre1 = dub[g + 1];
            pre2 = dub[g + 2];
        }
        if (forced) {
            const R* F = static_cast<const R*>(a.F) + pr.soff + gb;
            f0 = F[g + 0];
            f1 = F[g + 1];
            f2 = F[g + 2];
        }
        if constexpr (KIND == 2) {
            const double* ac = a.acc + pr.soff + gb;
            ac0 = ac[g + 0];
            ac1 = ac[g + 1];
            ac2 = ac[g + 2];
            if (period_end) {
                // period 0 hands over to the initial iterate; later periods need pi0 and W_k
                const double* x = (it == 0 ? a.winit : a.pi0) + pr.soff + gb;
                x0 = x[g + 0];
                x1 = x[g + 1];
                x2 = x[g + 2];
            }
        }
    };
    int el;
    size_t gb;
    uint32_t g;
    bool active;
    const wdg::NodeResult<R> q = wdg::apply_operator<N, R, true, true, true>(
        u_in, a.geom + 6 * (size_t)pr.eoff, static_cast<const wdg::FaceGeom<R>*>(a.fgeo) + 4 * (size_t)pr.eoff,
        a.fnb + 4 * (size_t)pr.eoff, pr.basis, (R)pr.ha, (R)pr.hb, blk.e0, blk.nloc, lds, el, gb, g, active, prefetch);
    (void)el;

    double part = 0.0;
    if (active) {
        R r0 = q.r0, r1 = q.r1, r2 = q.r2;
        if (forced) {
            const R fs = (R)a.ftab[pr.foff + 5 * (long long)s + st];  // -sin(omega t)/omega (WaveHoltz.cpp:65)
            r0 += f0 * fs;
            r1 += f1 * fs;
            r2 += f2 * fs;
        }
        const R im = (R)wdg::recip(q.m);
        r0 *= im;
        r1 *= im;
        r2 *= im;
        const R dt = (R)a.ctab[pr.coff + pr.nt + 1];  // dt is stored after the coefficients
        const R rb = (R)a.rk_b;
        R d0, d1, d2;
        if constexpr (KIND == 0) {
            d0 = dt * r0;
            d1 = dt * r1;
            d2 = dt * r2;
        } else {
            const R ra = (R)a.rk_a;
            d0 = ra * pre0 + dt * r0;
            d1 = ra * pre1 + dt * r1;
            d2 = ra * pre2 + dt * r2;
        }
        du += gb;
        u_out += gb;
        du[g + 0] = d0;
        du[g + 1] = d1;
        du[g + 2] = d2;
        R w0, w1, w2;
        wdg::stage_update(q.p, q.uu, q.vv, rb * d0, rb * d1, rb * d2, (uint32_t)gb + g, wdg::sr_seed((int)s, st), w0,
                          w1, w2);
        if constexpr (KIND != 2) {
            u_out[g + 0] = w0;
            u_out[g + 1] = w1;
            u_out[g + 2] = w2;
        } else {
            const double c = a.ctab[pr.coff + s + 1];
            const double n0 = ac0 + c * (double)w0, n1 = ac1 + c * (double)w1, n2 = ac2 + c * (double)w2;
            double* ac = a.acc + pr.soff + gb;
            if (!period_end) {
                u_out[g + 0] = w0;
                u_out[g + 1] = w1;
                u_out[g + 2] = w2;
                ac[g + 0] = n0;
                ac[g + 1] = n1;
                ac[g + 2] = n2;
            } else {
                double W0, W1, W2;
                double* wp = a.wprev + pr.soff + gb;
                if (it == 0) {
                    double* pz = a.pi0 + pr.soff + gb;
                    pz[g + 0] = n0;
                    pz[g + 1] = n1;
                    pz[g + 2] = n2;
                    part = n0 * n0 + n1 * n1 + n2 * n2;  // ||pi0||^2
                    const bool warm = a.warm[blk.prob] != 0;
                    W0 = warm ? x0 : n0;  // W_1 = S(0) + pi0 = pi0 unless an initial iterate is given
                    W1 = warm ? x1 : n1;
                    W2 = warm ? x2 : n2;
                } else {
                    W0 = n0 + x0;
                    W1 = n1 + x1;
                    W2 = n2 + x2;
                    const double e0 = W0 - wp[g + 0], e1 = W1 - wp[g + 1], e2 = W2 - wp[g + 2];
                    part = e0 * e0 + e1 * e1 + e2 * e2;
                }
                wp[g + 0] = W0;
                wp[g + 1] = W1;
                wp[g + 2] = W2;
                const double c0 = a.ctab[pr.coff];  // 1/2 dt K(0)
                ac[g + 0] = W0 * c0;
                ac[g + 1] = W1 * c0;
                ac[g + 2] = W2 * c0;
                u_out[g + 0] = (R)W0;
                u_out[g + 1] = (R)W1;
                u_out[g + 2] = (R)W2;
            }
        }
    }
    if (period_end) {  // uniform per workgroup
        double* red = reinterpret_cast<double*>(lds + max_op_lds<R>());
        __syncthreads();
        red[threadIdx.x] = part;
        __syncthreads();
        for (int w = kBT / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.partial[blk.gid] = red[0];
    }
}

template <typename R, int KIND>
__global__ __launch_bounds__(kBT) void batch_stage_kernel(const BatchArgs a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const BatchBlock blk = a.blocks[blockIdx.x];
    if (!a.active[blk.prob]) return;  // uniform per workgroup
    const BatchProb pr = a.probs[blk.prob];
    const long long G = *a.G;
    const int s = (int)(G % pr.nt);
    const long long it = G / pr.nt;
    switch (blk.N) {
    case 2: batch_body<2, R, KIND>(a, blk, pr, s, it, G, lds); break;
    case 3: batch_body<3, R, KIND>(a, blk, pr, s, it, G, lds); break;
    case 4: batch_body<4, R, KIND>(a, blk, pr, s, it, G, lds); break;
    case 5: batch_body<5, R, KIND>(a, blk, pr, s, it, G, lds); break;
    case 6: batch_body<6, R, KIND>(a, blk, pr, s, it, G, lds); break;
    case 7: batch_body<7, R, KIND>(a, blk, pr, s, it, G, lds); break;
    case 8: batch_body<8, R, KIND>(a, blk, pr, s, it, G, lds); break;
    case 9: batch_body<9, R, KIND>(a, blk, pr, s, it, G, lds); break;
    default: break;
    }
}

// One workgroup per problem: finish the residual of problems whose period
// ended at step G, retire converged ones; the last workgroup advances G.
__global__ __launch_bounds__(kBT) void batch_reduce_kernel(const BatchProb* probs, const int* prob_ids,
                                                           const double* partial, int* active, long long* G,
                                                           unsigned int* done, double* pn, double* hist, int* iters,
                                                           int hist_len, double tol, int max_iter, double* best,
                                                           int* stall, int stall_window)
{
    __shared__ double red[kBT];
    const int k = prob_ids[blockIdx.x];
    const long long g = *G;
    const BatchProb pr = probs[k];
    if (active[k] && (g + 1) % pr.nt == 0) {
        const long long it = g / pr.nt;
        double sum = 0.0;
        for (int b = threadIdx.x; b < pr.nblk; b += kBT) sum += partial[pr.blk_off + b];
        red[threadIdx.x] = sum;
        __syncthreads();
        for (int w = kBT / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double res;
            if (it == 0) {
                pn[k] = sqrt(red[0]);
                res = 1.0;  // ||W_1 - W_0|| / ||pi0|| with W_0 = 0, W_1 = pi0
            } else {
                res = sqrt(red[0]) / pn[k];
            }
            if (it < hist_len) hist[(long long)k * hist_len + it] = res;
            iters[k] = (int)it + 1;
            bool stalled = false;
            if (stall_window > 0 && it >= 1) {
                // retire a problem whose residual stopped improving (3 % over the window)
                if (res < 0.97 * best[k]) {
                    best[k] = res;
                    stall[k] = 0;
                } else {
                    stalled = ++stall[k] >= stall_window;
                }
            }
            if ((it >= 1 && res <= tol) || it + 1 >= max_iter || !(res == res) || stalled) active[k] = 0;
        }
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned int ticket = atomicAdd(done, 1u);
        if (ticket == gridDim.x - 1) {
            *done = 0u;
            *G = g + 1;
            __threadfence();
        }
    }
}

__global__ void batch_zero_kernel(double* x, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = 0.0;
}

template <typename R>
__global__ void batch_zero_r(R* x, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = (R)0;
}

template <typename S, typename D>
__global__ void batch_convert(const S* x, D* y, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = (D)x[k];
}

__global__ void batch_init_state(int* active, int n, long long* G, unsigned int* done, int* iters, double* best,
                                 int* stall)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        active[k] = 1;
        iters[k] = 0;
        best[k] = 1e300;
        stall[k] = 0;
    }
    if (k == 0) {
        *G = 0;
        *done = 0u;
    }
}


struct Fail {
    int code;
    std::string msg;
};

void bcheck(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Fail{WDG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

template <typename F>
int bguard(F&& f)
{
    try {
        f();
        return WDG_OK;
    } catch (const Fail& x) {
        wdg::g_last_error = x.msg;
        return x.code;
    } catch (const std::exception& x) {
        wdg::g_last_error = x.what();
        return WDG_ERR_INVALID;
    }
}

template <typename T>
T* dalloc(size_t n)
{
    void* p = nullptr;
    if (hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) {
        (void)hipGetLastError();
        throw Fail{WDG_ERR_ALLOC, "hipMalloc failed (" + std::to_string(sizeof(T) * n) + " bytes)"};
    }
    return static_cast<T*>(p);
}

// Carpenter-Kennedy (5,4) 2N-storage coefficients (ode.hpp:68-70).
constexpr double kA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                          -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
constexpr double kB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                          1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                          2277821191437.0 / 14882151754819.0};
constexpr double kC[5] = {0.0, 1432997174477.0 / 9575080441755.0, 2526269341429.0 / 6820363962896.0,
                          2006345519317.0 / 3224310063776.0, 2802321613138.0 / 2924317926251.0};

}  // namespace

struct wdg_batch_s {
    int dev = 0, prec = WDG_F64, nprob = 0, nblk = 0, hist_len = 0;
    size_t rsz = 8, lds = 0;
    long long ntot = 0;  // total state length
    std::vector<BatchProb> probs;
    std::vector<BatchBlock> all_blocks;  // full workgroup table
    int n_run_blocks = 0, n_run_probs = 0;  // compacted tables of the problems still running
    int* d_prob_ids = nullptr;
    std::vector<long long> ndof;
    std::vector<double> omega;
    BatchBlock* d_blocks = nullptr;
    BatchProb* d_probs = nullptr;
    double* d_basis = nullptr;
    double* d_geom = nullptr;
    int32_t* d_nbr = nullptr;
    int32_t* d_info = nullptr;
    void* d_fgeo = nullptr;
    wdg::FaceNb* d_fnb = nullptr;
    double* d_ctab = nullptr;
    double* d_ftab = nullptr;
    void* d_buf0 = nullptr;
    void* d_buf1 = nullptr;
    void* d_du = nullptr;
    void* d_F = nullptr;
    double* d_acc = nullptr;
    double* d_wprev = nullptr;
    double* d_pi0 = nullptr;
    double* d_winit = nullptr;
    int* d_warm = nullptr;
    int* d_active = nullptr;
    int* d_iters = nullptr;
    double* d_best = nullptr;  // best residual so far (stall detection)
    int* d_stall = nullptr;    // iterations since the last 3 % improvement
    int stall_window = 0;      // 0: no stall exit
    double* d_partial = nullptr;
    double* d_pn = nullptr;
    double* d_hist = nullptr;
    long long* d_G = nullptr;
    unsigned int* d_done = nullptr;
    hipStream_t cap = nullptr;
    hipGraphExec_t graph = nullptr;
    double last_seconds = 0.0;
    long long last_steps = 0;
    ~wdg_batch_s()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (cap) (void)hipStreamDestroy(cap);
        void* ptrs[] = {d_best, d_stall, d_prob_ids, d_blocks, d_probs, d_basis, d_geom, d_nbr, d_info, d_fgeo, d_fnb, d_ctab, d_ftab, d_buf0, d_buf1, d_du, d_F,
                        d_acc, d_wprev, d_pi0, d_winit, d_warm, d_active, d_iters, d_partial, d_pn, d_hist, d_G,
                        d_done};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
    BatchArgs args(int stage) const
    {
        BatchArgs a{};
        a.blocks = d_blocks;
        a.probs = d_probs;
        a.buf0 = d_buf0;
        a.buf1 = d_buf1;
        a.du = d_du;
        a.F = d_F;
        a.acc = d_acc;
        a.wprev = d_wprev;
        a.pi0 = d_pi0;
        a.winit = d_winit;
        a.warm = d_warm;
        a.geom = d_geom;
        a.nbr = d_nbr;
        a.info = d_info;
        a.fgeo = d_fgeo;
        a.fnb = d_fnb;
        a.ctab = d_ctab;
        a.ftab = d_ftab;
        a.active = d_active;
        a.partial = d_partial;
        a.G = d_G;
        a.rk_a = kA[stage];
        a.rk_b = kB[stage];
        a.stage = stage;
        return a;
    }
};

namespace {

template <typename R>
void launch_step(const wdg_batch_s& b, hipStream_t s, double tol, int max_iter)
{
    for (int st = 0; st < 5; ++st) {
        const BatchArgs a = b.args(st);
        if (st == 0)
            hipLaunchKernelGGL((batch_stage_kernel<R, 0>), dim3(b.n_run_blocks), dim3(kBT), b.lds, s, a);
        else if (st < 4)
            hipLaunchKernelGGL((batch_stage_kernel<R, 1>), dim3(b.n_run_blocks), dim3(kBT), b.lds, s, a);
        else
            hipLaunchKernelGGL((batch_stage_kernel<R, 2>), dim3(b.n_run_blocks), dim3(kBT), b.lds, s, a);
    }
    hipLaunchKernelGGL(batch_reduce_kernel, dim3(b.n_run_probs), dim3(kBT), 0, s, b.d_probs, b.d_prob_ids, b.d_partial,
                       b.d_active, b.d_G, b.d_done, b.d_pn, b.d_hist, b.d_iters, b.hist_len, tol, max_iter, b.d_best,
                       b.d_stall, b.stall_window);
}

// Tables of the problems still running (host flags), uploaded for the next capture.
void compact(wdg_batch_s& b, const std::vector<int>& run, hipStream_t s)
{
    std::vector<BatchBlock> blocks;
    std::vector<int> ids;
    for (int k = 0; k < b.nprob; ++k) {
        if (!run[k]) continue;
        ids.push_back(k);
        const BatchProb& pr = b.probs[k];
        blocks.insert(blocks.end(), b.all_blocks.begin() + pr.blk_off, b.all_blocks.begin() + pr.blk_off + pr.nblk);
    }
    b.n_run_blocks = (int)blocks.size();
    b.n_run_probs = (int)ids.size();
    if (blocks.empty()) return;
    bcheck(hipMemcpyAsync(b.d_blocks, blocks.data(), sizeof(BatchBlock) * blocks.size(), hipMemcpyHostToDevice, s),
           "upload blocks");
    bcheck(hipMemcpyAsync(b.d_prob_ids, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, s), "upload ids");
    bcheck(hipStreamSynchronize(s), "hipStreamSynchronize");
}

template <typename Fn>
void capture(wdg_batch_s& b, int steps, Fn&& step)
{
    std::lock_guard<std::mutex> lk(wdg::capture_mutex());
    if (b.graph) (void)hipGraphExecDestroy(b.graph);
    b.graph = nullptr;
    if (!b.cap) bcheck(hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking), "hipStreamCreate");
    hipGraph_t gph = nullptr;
    bcheck(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    for (int k = 0; k < steps; ++k) step(b.cap);
    bcheck(hipStreamEndCapture(b.cap, &gph), "hipStreamEndCapture");
    const hipError_t ie = hipGraphInstantiate(&b.graph, gph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gph);
    bcheck(ie, "hipGraphInstantiate");
}

}  // namespace

extern "C" {

int wdg_hip_batch_create(int n_problems, const wdg_batch_problem* problems, int precision, wdg_batch* out)
{
    return bguard([&] {
        if (!out || !problems || n_problems <= 0) throw Fail{WDG_ERR_INVALID, "bad batch arguments"};
        if (precision != WDG_F64 && precision != WDG_F32) throw Fail{WDG_ERR_INVALID, "precision must be WDG_F64 or WDG_F32"};
        auto b = std::make_unique<wdg_batch_s>();
        bcheck(hipGetDevice(&b->dev), "hipGetDevice");
        b->prec = precision;
        b->rsz = precision == WDG_F32 ? 4 : 8;
        b->nprob = n_problems;
        std::vector<BatchBlock> blocks;
        std::vector<double> basis_all, geom, ctab, ftab;
        std::vector<int32_t> nbr, info;
        std::vector<wdg::FaceGeom<float>> fg32;
        std::vector<wdg::FaceNb> fnb;
        std::vector<long long> basis_off(10, -1);
        std::vector<std::vector<double>> src_list(n_problems);
        long long soff = 0, coff = 0, foff = 0;
        int eoff = 0;
        for (int k = 0; k < n_problems; ++k) {
            const wdg_batch_problem& P = problems[k];
            if (!P.mesh) throw Fail{WDG_ERR_INVALID, "problem " + std::to_string(k) + ": null mesh"};
            const wdg::HostMesh& m = P.mesh->mesh;
            std::vector<double> x, w;
            if (!wdg::gll_rule(P.n_colloc, x, w))
                throw Fail{WDG_ERR_INVALID, "problem " + std::to_string(k) + ": n_colloc out of range 2..9"};
            if (!(P.omega > 0.0) || !std::isfinite(P.omega))
                throw Fail{WDG_ERR_INVALID, "problem " + std::to_string(k) + ": omega must be positive"};
            if (m.min_corner_detj() <= 0.0)
                throw Fail{WDG_ERR_MESH, "problem " + std::to_string(k) + ": inverted element"};
            wdg::FaceTables ft;
            std::string err;
            if (!m.face_tables(P.bc, ft, err)) throw Fail{WDG_ERR_INVALID, "problem " + std::to_string(k) + ": " + err};
            const int n = P.n_colloc, NN = n * n, nel = m.n_elem();
            if (basis_off[n] < 0) {
                basis_off[n] = (long long)basis_all.size();
                basis_all.insert(basis_all.end(), x.begin(), x.end());
                basis_all.insert(basis_all.end(), w.begin(), w.end());
                std::vector<double> D(NN);
                wdg::lagrange_derivative_matrix(n, x.data(), n, x.data(), D.data());
                basis_all.insert(basis_all.end(), D.begin(), D.end());
            }
            // WaveHoltz schedule (WaveHoltz.cpp:24-29)
            const double T = 2.0 * M_PI / P.omega;
            const double h = m.min_h();
            const double pp = n;
            double dt = 0.5 * h / (pp * pp);
            const int nt = (int)std::ceil(T / dt);
            dt = T / nt;
            auto K = [&](double t) { return P.omega / M_PI * (std::cos(P.omega * t) - 0.25); };
            BatchProb pr{};
            pr.soff = soff;
            pr.coff = coff;
            pr.foff = foff;
            pr.ha = -0.5 * 1.0;  // upwind, reference sign convention (EdgeFlux.hpp:51)
            pr.hb = -0.5;
            pr.eoff = eoff;
            pr.nel = nel;
            pr.N = n;
            pr.nt = nt;
            pr.blk_off = (int)blocks.size();
            // coefficient table: c0, c_1..c_nt, dt (same arithmetic as the single-problem driver)
            ctab.push_back(0.5 * dt * K(0.0));
            double t = 0.0;
            for (int it = 0; it < nt; ++it) {
                for (int st = 0; st < 5; ++st) {
                    const double ts = (st == 0) ? t : t + kC[st] * dt;
                    ftab.push_back(-1.0 / P.omega * std::sin(P.omega * ts));
                }
                const double t_after = t + dt;
                ctab.push_back(((it + 1 == nt) ? 0.5 : 1.0) * dt * K(t_after));
                t += dt;
            }
            ctab.push_back(dt);
            const int epb = 256 / NN > 0 ? 256 / NN : 1;
            for (int e0 = 0; e0 < nel; e0 += epb)
                blocks.push_back(BatchBlock{k, n, e0, std::min(epb, nel - e0), (int)blocks.size()});
            pr.nblk = (int)blocks.size() - pr.blk_off;
            for (int e = 0; e < nel; ++e) {
                const double* c = &m.corners()[8 * (size_t)e];
                geom.push_back(0.25 * ((c[2] - c[0]) + (c[4] - c[6])));
                geom.push_back(0.25 * ((c[3] - c[1]) + (c[5] - c[7])));
                geom.push_back(0.25 * ((c[6] - c[0]) + (c[4] - c[2])));
                geom.push_back(0.25 * ((c[7] - c[1]) + (c[5] - c[3])));
                geom.push_back(0.25 * ((c[0] - c[2]) + (c[4] - c[6])));
                geom.push_back(0.25 * ((c[1] - c[3]) + (c[5] - c[7])));
            }
            nbr.insert(nbr.end(), ft.nbr.begin(), ft.nbr.end());
            info.insert(info.end(), ft.info.begin(), ft.info.end());
            {
                // side normals, half lengths and neighbour trace bases, local to the problem's state
                const double* gk = geom.data() + geom.size() - 6 * (size_t)nel;
                std::vector<wdg::FaceNb> fk;
                if (precision == WDG_F32) {
                    std::vector<wdg::FaceGeom<float>> gq;
                    wdg::build_face_tables(n, nel, gk, ft.nbr.data(), ft.info.data(), gq, fk);
                    fg32.insert(fg32.end(), gq.begin(), gq.end());
                } else {
                    std::vector<wdg::FaceGeom<double>> gq;  // unused: FP64 kernels compute normals
                    wdg::build_face_tables(n, nel, gk, ft.nbr.data(), ft.info.data(), gq, fk);
                }
                fnb.insert(fnb.end(), fk.begin(), fk.end());
            }
            b->probs.push_back(pr);
            b->ndof.push_back(3LL * NN * nel);
            b->omega.push_back(P.omega);
            if (P.n_sources < 0 || (P.n_sources > 0 && !P.sources))
                throw Fail{WDG_ERR_INVALID, "problem " + std::to_string(k) + ": bad sources"};
            src_list[k].assign(P.sources, P.sources + 4 * P.n_sources);
            soff += 3LL * NN * nel;
            coff += nt + 2;
            foff += 5LL * nt;
            eoff += nel;
        }
        b->ntot = soff;
        b->nblk = (int)blocks.size();
        b->all_blocks = blocks;
        b->d_prob_ids = dalloc<int>(n_problems);
        // device arrays
        b->d_blocks = dalloc<BatchBlock>(blocks.size());
        b->d_basis = dalloc<double>(basis_all.size());
        b->d_geom = dalloc<double>(geom.size());
        b->d_nbr = dalloc<int32_t>(nbr.size());
        b->d_info = dalloc<int32_t>(info.size());
        b->d_ctab = dalloc<double>(ctab.size());
        b->d_ftab = dalloc<double>(ftab.size());
        bcheck(hipMemcpy(b->d_blocks, blocks.data(), sizeof(BatchBlock) * blocks.size(), hipMemcpyHostToDevice), "upload");
        bcheck(hipMemcpy(b->d_basis, basis_all.data(), sizeof(double) * basis_all.size(), hipMemcpyHostToDevice), "upload");
        bcheck(hipMemcpy(b->d_geom, geom.data(), sizeof(double) * geom.size(), hipMemcpyHostToDevice), "upload");
        bcheck(hipMemcpy(b->d_nbr, nbr.data(), sizeof(int32_t) * nbr.size(), hipMemcpyHostToDevice), "upload");
        bcheck(hipMemcpy(b->d_info, info.data(), sizeof(int32_t) * info.size(), hipMemcpyHostToDevice), "upload");
        if (precision == WDG_F32) {
            b->d_fgeo = dalloc<wdg::FaceGeom<float>>(fg32.size());
            bcheck(hipMemcpy(b->d_fgeo, fg32.data(), sizeof(fg32[0]) * fg32.size(), hipMemcpyHostToDevice), "upload");
        }  // FP64 kernels compute side normals from the map coefficients (no face-geometry table)
        b->d_fnb = dalloc<wdg::FaceNb>(fnb.size());
        bcheck(hipMemcpy(b->d_fnb, fnb.data(), sizeof(wdg::FaceNb) * fnb.size(), hipMemcpyHostToDevice), "upload");
        bcheck(hipMemcpy(b->d_ctab, ctab.data(), sizeof(double) * ctab.size(), hipMemcpyHostToDevice), "upload");
        bcheck(hipMemcpy(b->d_ftab, ftab.data(), sizeof(double) * ftab.size(), hipMemcpyHostToDevice), "upload");
        for (int k = 0; k < n_problems; ++k) b->probs[k].basis = b->d_basis + basis_off[b->probs[k].N];
        b->d_probs = dalloc<BatchProb>(n_problems);
        bcheck(hipMemcpy(b->d_probs, b->probs.data(), sizeof(BatchProb) * n_problems, hipMemcpyHostToDevice), "upload");
        const size_t sb = b->rsz * (size_t)b->ntot;
        b->d_buf0 = dalloc<char>(sb);
        b->d_buf1 = dalloc<char>(sb);
        b->d_du = dalloc<char>(sb);
        b->d_F = dalloc<char>(sb);
        b->d_acc = dalloc<double>(b->ntot);
        b->d_wprev = dalloc<double>(b->ntot);
        b->d_pi0 = dalloc<double>(b->ntot);
        b->d_winit = dalloc<double>(b->ntot);
        b->d_warm = dalloc<int>(n_problems);
        b->d_active = dalloc<int>(n_problems);
        b->d_iters = dalloc<int>(n_problems);
        b->d_best = dalloc<double>(n_problems);
        b->d_stall = dalloc<int>(n_problems);
        b->d_partial = dalloc<double>(b->nblk);
        b->d_pn = dalloc<double>(n_problems);
        b->d_G = dalloc<long long>(1);
        b->d_done = dalloc<unsigned int>(1);
        bcheck(hipMemset(b->d_warm, 0, sizeof(int) * n_problems), "hipMemset");
        bcheck(hipMemset(b->d_winit, 0, sizeof(double) * b->ntot), "hipMemset");
        b->lds = (precision == WDG_F32 ? max_op_lds<float>() : max_op_lds<double>()) + kBT * sizeof(double);
        // load vectors, one problem at a time with the single-problem kernels
        double* Fd = dalloc<double>(b->ntot);
        double* d_src = nullptr;
        for (int k = 0; k < n_problems; ++k) {
            const BatchProb& pr = b->probs[k];
            const int ns = (int)(src_list[k].size() / 4);
            if (d_src) (void)hipFree(d_src);
            d_src = dalloc<double>(std::max<size_t>(src_list[k].size(), 4));
            if (ns) bcheck(hipMemcpy(d_src, src_list[k].data(), sizeof(double) * src_list[k].size(), hipMemcpyHostToDevice), "upload");
            // Gauss-Legendre table for this order
            const int n = pr.N;
            std::vector<double> xq, wq, xg, wg;
            wdg::gauss_legendre_rule(n, xq, wq);
            wdg::gll_rule(n, xg, wg);
            std::vector<double> tab(2 * n + n * n);
            std::memcpy(tab.data(), xq.data(), sizeof(double) * n);
            std::memcpy(tab.data() + n, wq.data(), sizeof(double) * n);
            wdg::lagrange_interpolation_matrix(n, xg.data(), n, xq.data(), tab.data() + 2 * n);
            double* d_tab = dalloc<double>(tab.size());
            bcheck(hipMemcpy(d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice), "upload");
            // corners of this problem
            const wdg::HostMesh& m = problems[k].mesh->mesh;
            double* d_c = dalloc<double>(m.corners().size());
            bcheck(hipMemcpy(d_c, m.corners().data(), sizeof(double) * m.corners().size(), hipMemcpyHostToDevice), "upload");
            switch (n) {
            case 2: wdg::launch_load_vector_n2(d_c, d_tab, d_src, ns, 0, Fd + pr.soff, pr.nel, nullptr); break;
            case 3: wdg::launch_load_vector_n3(d_c, d_tab, d_src, ns, 0, Fd + pr.soff, pr.nel, nullptr); break;
            case 4: wdg::launch_load_vector_n4(d_c, d_tab, d_src, ns, 0, Fd + pr.soff, pr.nel, nullptr); break;
            case 5: wdg::launch_load_vector_n5(d_c, d_tab, d_src, ns, 0, Fd + pr.soff, pr.nel, nullptr); break;
            case 6: wdg::launch_load_vector_n6(d_c, d_tab, d_src, ns, 0, Fd + pr.soff, pr.nel, nullptr); break;
            case 7: wdg::launch_load_vector_n7(d_c, d_tab, d_src, ns, 0, Fd + pr.soff, pr.nel, nullptr); break;
            case 8: wdg::launch_load_vector_n8(d_c, d_tab, d_src, ns, 0, Fd + pr.soff, pr.nel, nullptr); break;
            case 9: wdg::launch_load_vector_n9(d_c, d_tab, d_src, ns, 0, Fd + pr.soff, pr.nel, nullptr); break;
            default: break;
            }
            bcheck(hipGetLastError(), "load vector launch");
            bcheck(hipDeviceSynchronize(), "load vector");
            (void)hipFree(d_tab);
            (void)hipFree(d_c);
        }
        if (d_src) (void)hipFree(d_src);
        const unsigned gr = (unsigned)((b->ntot + 255) / 256);
        if (precision == WDG_F32)
            hipLaunchKernelGGL((batch_convert<double, float>), dim3(gr), dim3(256), 0, 0, Fd, static_cast<float*>(b->d_F), b->ntot);
        else
            bcheck(hipMemcpy(b->d_F, Fd, sizeof(double) * b->ntot, hipMemcpyDeviceToDevice), "copy F");
        bcheck(hipDeviceSynchronize(), "convert F");
        (void)hipFree(Fd);
        *out = b.release();
    });
}

int wdg_hip_batch_destroy(wdg_batch b)
{
    delete b;
    return WDG_OK;
}

int wdg_hip_batch_info(wdg_batch b, int k, long long* ndof, int* nt, int* n_colloc)
{
    return bguard([&] {
        if (!b || k < 0 || k >= b->nprob) throw Fail{WDG_ERR_INVALID, "bad batch handle or index"};
        if (ndof) *ndof = b->ndof[k];
        if (nt) *nt = b->probs[k].nt;
        if (n_colloc) *n_colloc = b->probs[k].N;
    });
}

int wdg_hip_batch_set_initial(wdg_batch b, int k, const double* d_W0)
{
    return bguard([&] {
        if (!b || k < 0 || k >= b->nprob) throw Fail{WDG_ERR_INVALID, "bad batch handle or index"};
        const int warm = d_W0 ? 1 : 0;
        if (d_W0)
            bcheck(hipMemcpy(b->d_winit + b->probs[k].soff, d_W0, sizeof(double) * b->ndof[k], hipMemcpyDeviceToDevice),
                   "copy initial iterate");
        bcheck(hipMemcpy(b->d_warm + k, &warm, sizeof(int), hipMemcpyHostToDevice), "set warm flag");
    });
}

int wdg_hip_batch_solve(wdg_batch b, double tol, int max_iter, int steps_per_graph, void* stream)
{
    return bguard([&] {
        if (!b) throw Fail{WDG_ERR_INVALID, "null batch handle"};
        if (!(tol > 0.0) || max_iter < 1 || steps_per_graph < 1) throw Fail{WDG_ERR_INVALID, "bad solve arguments"};
        int cur = -1;
        bcheck(hipGetDevice(&cur), "hipGetDevice");
        if (cur != b->dev) bcheck(hipSetDevice(b->dev), "hipSetDevice");
        const hipStream_t s = static_cast<hipStream_t>(stream);
        // reset: zero state and accumulators, all problems active
        if (b->hist_len < max_iter + 1) {
            std::lock_guard<std::mutex> lk(wdg::capture_mutex());
            if (b->d_hist) (void)hipFree(b->d_hist);
            b->hist_len = max_iter + 1;
            b->d_hist = dalloc<double>((size_t)b->hist_len * b->nprob);
        }
        const unsigned gr = (unsigned)((b->ntot + 255) / 256);
        if (b->prec == WDG_F32) {
            hipLaunchKernelGGL(batch_zero_r<float>, dim3(gr), dim3(256), 0, s, static_cast<float*>(b->d_buf0), b->ntot);
        } else {
            hipLaunchKernelGGL(batch_zero_r<double>, dim3(gr), dim3(256), 0, s, static_cast<double*>(b->d_buf0), b->ntot);
        }
        hipLaunchKernelGGL(batch_zero_kernel, dim3(gr), dim3(256), 0, s, b->d_acc, b->ntot);
        hipLaunchKernelGGL(batch_zero_kernel, dim3(gr), dim3(256), 0, s, b->d_wprev, b->ntot);
        hipLaunchKernelGGL(batch_init_state, dim3((b->nprob + 255) / 256), dim3(256), 0, s, b->d_active, b->nprob,
                           b->d_G, b->d_done, b->d_iters, b->d_best, b->d_stall);
        bcheck(hipGetLastError(), "batch reset");
        std::vector<int> act(b->nprob, 1), run(b->nprob, 1);
        auto step = [&](hipStream_t cs) {
            if (b->prec == WDG_F32)
                launch_step<float>(*b, cs, tol, max_iter);
            else
                launch_step<double>(*b, cs, tol, max_iter);
        };
        compact(*b, run, s);
        capture(*b, steps_per_graph, step);
        long long max_steps = 0;
        for (const BatchProb& pr : b->probs) max_steps = std::max(max_steps, (long long)pr.nt * (max_iter + 1));
        hipEvent_t e0, e1;
        bcheck(hipEventCreate(&e0), "hipEventCreate");
        bcheck(hipEventCreate(&e1), "hipEventCreate");
        bcheck(hipEventRecord(e0, s), "hipEventRecord");
        long long steps = 0;
        while (steps < max_steps) {
            bcheck(hipGraphLaunch(b->graph, s), "hipGraphLaunch");
            steps += steps_per_graph;
            bcheck(hipMemcpyAsync(act.data(), b->d_active, sizeof(int) * b->nprob, hipMemcpyDeviceToHost, s), "copy");
            bcheck(hipStreamSynchronize(s), "hipStreamSynchronize");
            if (act == run) continue;
            run = act;
            bool any = false;
            for (int v : run) any = any || v;
            if (!any) break;
            // problems retired: drop their workgroups from the launches
            compact(*b, run, s);
            capture(*b, steps_per_graph, step);
        }
        bcheck(hipEventRecord(e1, s), "hipEventRecord");
        bcheck(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.f;
        bcheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->last_seconds = ms * 1e-3;
        b->last_steps = steps;
    });
}

int wdg_hip_batch_set_stall_window(wdg_batch b, int window)
{
    return bguard([&] {
        if (!b || window < 0) throw Fail{WDG_ERR_INVALID, "bad batch handle or window"};
        b->stall_window = window;
    });
}

int wdg_hip_batch_result(wdg_batch b, int k, double* d_W, int* iterations, double* residual)
{
    return bguard([&] {
        if (!b || k < 0 || k >= b->nprob) throw Fail{WDG_ERR_INVALID, "bad batch handle or index"};
        std::lock_guard<std::mutex> lk(wdg::capture_mutex());  // null-stream copies
        if (d_W)
            bcheck(hipMemcpy(d_W, b->d_wprev + b->probs[k].soff, sizeof(double) * b->ndof[k], hipMemcpyDeviceToDevice),
                   "copy result");
        int it = 0;
        bcheck(hipMemcpy(&it, b->d_iters + k, sizeof(int), hipMemcpyDeviceToHost), "copy iterations");
        if (iterations) *iterations = it;
        if (residual) {
            double r = NAN;
            if (it >= 1 && b->d_hist)
                bcheck(hipMemcpy(&r, b->d_hist + (size_t)k * b->hist_len + (it - 1), sizeof(double),
                                 hipMemcpyDeviceToHost), "copy residual");
            *residual = r;
        }
    });
}

int wdg_hip_batch_timing(wdg_batch b, double* seconds, long long* steps)
{
    return bguard([&] {
        if (!b) throw Fail{WDG_ERR_INVALID, "null batch handle"};
        if (seconds) *seconds = b->last_seconds;
        if (steps) *steps = b->last_steps;
    });
}

}  // extern "C"
