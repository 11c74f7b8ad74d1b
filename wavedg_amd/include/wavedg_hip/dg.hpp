// C++ drop-in classes for the 2D wave path of waveDG on MI355X.
//
// Same names, constructor arguments and semantics as the reference
// (wavedg/Operator.hpp, WaveEquation.hpp, MassMatrix.hpp, WaveHoltz.hpp,
// Mesh2D.hpp, QuadratureRule.hpp), in namespace dg::hip, implemented on the
// C ABI of libwavedg_hip.so only (include/wavedg_hip.h).  Errors are thrown as
// dg::hip::Error (a std::runtime_error), as wdg_error throws in the
// reference's serial build (source/wdg_error.cpp:5-20).
//
// Arrays passed to the reference-style methods (action, inv, pi0, S,
// postprocess) are host arrays in the reference layout (3, n, n, n_elem); each
// call copies them to the device and back.  The *_device methods take device
// pointers (hipMalloc'd) and an optional hipStream_t and do no copies; they
// are the ones to use in time loops.
//
// Bases are Gauss-Lobatto or Gauss-Legendre rules, as in the reference.
// approx_quad = true collocates the integrals on the basis nodes (FP64 or
// FP32); approx_quad = false integrates them exactly with Gauss-Legendre
// rules, chosen as the reference chooses them when quad is null (FP64 only).
#pragma once

#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/wavedg_hip.h"

namespace dg {
namespace hip {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
    int code;
};

inline void check(int rc)
{
    if (rc != WDG_OK) throw Error(rc, wdg_hip_last_error());
}

enum class FaceType { INTERIOR = WDG_INTERIOR, BOUNDARY = WDG_BOUNDARY };
enum class Precision { F64 = WDG_F64, F32 = WDG_F32 };

// Device buffer owned by the drop-in objects for host-array calls.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { resize(bytes); }
    ~DeviceBuffer()
    {
        if (p_) wdg_hip_free(p_);
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void resize(size_t bytes)
    {
        if (bytes <= n_) return;
        if (p_) wdg_hip_free(p_);
        p_ = nullptr;
        check(wdg_hip_malloc(&p_, bytes));
        n_ = bytes;
    }
    void* get() const { return p_; }
    void upload(const void* src, size_t bytes) { check(wdg_hip_memcpy_htod(p_, src, bytes)); }
    void download(void* dst, size_t bytes) const { check(wdg_hip_memcpy_dtoh(dst, p_, bytes)); }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
};

// QuadratureRule::quadrature_rule(n, type) (QuadratureRule.hpp:77,
// QuadratureRule.cpp:164-184): cached rules, Gauss-Legendre by default as in
// the reference.  Either family serves as a basis; Gauss-Legendre rules also
// as the quad argument of the approx_quad = false operators.
class QuadratureRule {
public:
    enum QuadratureType { GaussLegendre, GaussLobatto };
    static const QuadratureRule* quadrature_rule(int n, QuadratureType type = GaussLegendre)
    {
        static std::unordered_map<int, std::unique_ptr<QuadratureRule>> cache;
        const int key = 2 * n + (type == GaussLobatto ? 1 : 0);
        auto it = cache.find(key);
        if (it == cache.end())
            it = cache.emplace(key, std::unique_ptr<QuadratureRule>(new QuadratureRule(n, type))).first;
        return it->second.get();
    }
    const int n;
    const double* x;
    const double* w;
    const QuadratureType type;

private:
    QuadratureRule(int n_, QuadratureType t) : n(n_), x(nullptr), w(nullptr), type(t), x_(n_ > 0 ? n_ : 0), w_(x_.size())
    {
        if (t == GaussLobatto)
            check(wdg_hip_gll(n_, x_.data(), w_.data(), nullptr));
        else
            check(wdg_hip_gauss_legendre(n_, x_.data(), w_.data()));
        x = x_.data();
        w = w_.data();
    }
    std::vector<double> x_, w_;
};

// Mesh2D (Mesh2D.hpp:339-355, Mesh2D.cpp:5-163).
class Mesh2D {
public:
    static Mesh2D from_vertices(int nx, const double* x, int nel, const int* elems)
    {
        wdg_mesh m = nullptr;
        check(wdg_hip_mesh_from_vertices(nx, x, nel, elems, &m));
        return Mesh2D(m);
    }
    static Mesh2D uniform_rect(int nx, double ax, double bx, int ny, double ay, double by)
    {
        wdg_mesh m = nullptr;
        check(wdg_hip_mesh_uniform_rect(nx, ax, bx, ny, ay, by, &m));
        return Mesh2D(m);
    }
    static Mesh2D from_file(const std::string& dir)
    {
        wdg_mesh m = nullptr;
        check(wdg_hip_mesh_from_file(dir.c_str(), &m));
        return Mesh2D(m);
    }
    int n_elem() const { return counts()[0]; }
    int n_edges(FaceType t) const { return counts()[t == FaceType::INTERIOR ? 1 : 2]; }
    double min_h() const
    {
        double h = 0;
        check(wdg_hip_mesh_min_h(h_.get(), &h));
        return h;
    }
    // element_metrics(basis).physical_coordinates(): (2, n, n, n_elem)
    std::vector<double> physical_coordinates(const QuadratureRule* basis) const
    {
        std::vector<double> x(2 * (size_t)basis->n * basis->n * n_elem());
        if (basis->type == QuadratureRule::GaussLobatto)
            check(wdg_hip_mesh_nodes(h_.get(), basis->n, x.data()));
        else
            check(wdg_hip_mesh_gl_points(h_.get(), basis->n, x.data()));
        return x;
    }
    // (5, n_edges): el0, side0, el1, side1, delta in Mesh2D::edge(i, type) order
    std::vector<int> edges(FaceType t) const
    {
        std::vector<int> out(5 * (size_t)n_edges(t));
        check(wdg_hip_mesh_edges(h_.get(), (int)t, out.data()));
        return out;
    }
    // (5, n_edges): nx, ny, ds, x_mid, y_mid
    std::vector<double> edge_geometry(FaceType t) const
    {
        std::vector<double> out(5 * (size_t)n_edges(t));
        check(wdg_hip_mesh_edge_geometry(h_.get(), (int)t, out.data()));
        return out;
    }
    wdg_mesh handle() const { return h_.get(); }

private:
    explicit Mesh2D(wdg_mesh m) : h_(m, [](wdg_mesh p) { wdg_hip_mesh_destroy(p); }) {}
    std::vector<int> counts() const
    {
        std::vector<int> c(3);
        check(wdg_hip_mesh_counts(h_.get(), &c[0], &c[1], &c[2]));
        return c;
    }
    std::shared_ptr<wdg_mesh_s> h_;
};

// Operator / InvertibleOperator (Operator.hpp:10-43).
class Operator {
public:
    virtual ~Operator() = default;
    virtual void action(const double* x, double* y) const = 0;
    virtual void action(int n_var, const double* x, double* y) const = 0;
};

class InvertibleOperator : public Operator {
public:
    virtual void inv(double* x) const = 0;
    virtual void inv(int n_var, double* x) const = 0;
};

namespace detail {
// Size of a user-given rule for the exact-quadrature operators; 0 (null rule)
// lets the library choose as the reference does (Div.cpp:319-327,
// EdgeFlux.cpp:503-513, MassMatrix.cpp:92-98).
inline int exact_rule_size(const QuadratureRule* q)
{
    if (!q) return 0;
    if (q->type != QuadratureRule::GaussLegendre)
        throw std::invalid_argument("dg::hip: approx_quad = false integrates with Gauss-Legendre rules only");
    return q->n;
}

// One device discretisation (geometry, connectivity, basis) per object.  With
// approx_quad = false, quad sets the volume and edge rules (WaveEquation.cpp:
// 25-29 hands the same rule to Div<false> and EdgeFlux<false>) and mass_quad
// the MassMatrix<false> rule.
class Wave2D {
public:
    Wave2D(const Mesh2D& mesh, const QuadratureRule* basis, const int* bc, double a, double b, Precision p,
           bool approx_quad = true, const QuadratureRule* quad = nullptr, const QuadratureRule* mass_quad = nullptr)
    {
        if (!basis) throw std::invalid_argument("dg::hip: null basis");
        const int type = basis->type == QuadratureRule::GaussLobatto ? WDG_GAUSS_LOBATTO : WDG_GAUSS_LEGENDRE;
        wdg_wave2d h = nullptr;
        if (approx_quad) {
            check(wdg_hip_wave2d_create_basis(mesh.handle(), basis->n, type, bc, a, b, (int)p, 1, 0, 0, 0, &h));
        } else {
            if (p != Precision::F64) throw std::invalid_argument("dg::hip: approx_quad = false runs in FP64 only");
            const int nq = exact_rule_size(quad);
            check(wdg_hip_wave2d_create_basis(mesh.handle(), basis->n, type, bc, a, b, (int)p, 0, nq, nq,
                                              exact_rule_size(mass_quad), &h));
        }
        h_.reset(h);
        check(wdg_hip_wave2d_info(h, &ndof_, nullptr, nullptr, nullptr));
        rsz_ = p == Precision::F32 ? 4 : 8;
    }
    wdg_wave2d get() const { return h_.get(); }
    long long ndof() const { return ndof_; }
    size_t state_bytes() const { return (size_t)ndof_ * rsz_; }
    mutable DeviceBuffer x, y;

private:
    struct Del {
        void operator()(wdg_wave2d p) const { wdg_hip_wave2d_destroy(p); }
    };
    std::unique_ptr<wdg_wave2d_s, Del> h_;
    long long ndof_ = 0;
    size_t rsz_ = 8;
};
}  // namespace detail

// WaveEquation (WaveEquation.hpp:28-52): Dw <- Dw + A w, interior faces only.
class WaveEquation : public Operator {
public:
    WaveEquation(const Mesh2D& mesh, const QuadratureRule* basis, bool approx_quad = false,
                 const QuadratureRule* quad = nullptr, double a = 1.0, double b = 0.5)
        : op_(mesh, basis, nullptr, a, b, Precision::F64, approx_quad, quad)
    {
    }
    void action(const double* w, double* Dw) const override
    {
        const size_t bytes = op_.state_bytes();
        op_.x.resize(bytes);
        op_.y.resize(bytes);
        op_.x.upload(w, bytes);
        op_.y.upload(Dw, bytes);
        action_device(op_.x.get(), op_.y.get());
        op_.y.download(Dw, bytes);
    }
    void action(int, const double* w, double* Dw) const override { action(w, Dw); }
    void action_device(const void* d_w, void* d_Dw, void* stream = nullptr) const
    {
        check(wdg_hip_wave2d_action(op_.get(), d_w, d_Dw, stream));
    }

private:
    detail::Wave2D op_;
};

// WaveBC (WaveEquation.hpp:102-122): Bw <- Bw + B w.  bc codes per boundary
// edge: 0 absorbing, 1 reflecting, 2 Dirichlet p = 0.
class WaveBC : public Operator {
public:
    WaveBC(const Mesh2D& mesh, const int* bc, const QuadratureRule* basis, bool approx_quad,
           const QuadratureRule* quad = nullptr, double a = 1.0, double b = 0.5)
        : op_(mesh, basis, bc, a, b, Precision::F64, approx_quad, quad)
    {
    }
    void action(const double* w, double* Bw) const override
    {
        const size_t bytes = op_.state_bytes();
        op_.x.resize(bytes);
        op_.y.resize(bytes);
        op_.x.upload(w, bytes);
        op_.y.upload(Bw, bytes);
        action_device(op_.x.get(), op_.y.get());
        op_.y.download(Bw, bytes);
    }
    void action(int, const double* w, double* Bw) const override { action(w, Bw); }
    void action_device(const void* d_w, void* d_Bw, void* stream = nullptr) const
    {
        check(wdg_hip_wavebc_action(op_.get(), d_w, d_Bw, stream));
    }

private:
    detail::Wave2D op_;
};

// MassMatrix<true> (MassMatrix.hpp:77, MassMatrix.cpp:6-82), the diagonal
// w_i w_j det J, and MassMatrix<false> (MassMatrix.cpp:85-227), the exact
// block matrix (applied and inverted on the GPU without stored factors).
// The n_var overloads take any number of variables per node (3 only for
// <false>); quad is ignored for <true>, as in the reference.
template <bool ApproxQuadrature>
class MassMatrix : public InvertibleOperator {
public:
    MassMatrix(const Mesh2D& mesh, const QuadratureRule* basis, const QuadratureRule* quad = nullptr)
        : op_(mesh, basis, nullptr, 1.0, 0.5, Precision::F64, ApproxQuadrature, nullptr, ApproxQuadrature ? nullptr : quad)
    {
    }
    void action(const double* x, double* y) const override
    {
        const size_t bytes = op_.state_bytes();
        op_.x.resize(bytes);
        op_.y.resize(bytes);
        op_.x.upload(x, bytes);
        check(wdg_hip_mass_action(op_.get(), op_.x.get(), op_.y.get(), nullptr));
        op_.y.download(y, bytes);
    }
    void action(int n_var, const double* x, double* y) const override
    {
        const size_t bytes = op_.state_bytes() / 3 * n_var;
        op_.x.resize(bytes);
        op_.y.resize(bytes);
        op_.x.upload(x, bytes);
        check(wdg_hip_mass_apply(op_.get(), n_var, op_.x.get(), op_.y.get(), 0, nullptr));
        op_.y.download(y, bytes);
    }
    void inv(double* x) const override
    {
        const size_t bytes = op_.state_bytes();
        op_.x.resize(bytes);
        op_.x.upload(x, bytes);
        inv_device(op_.x.get());
        op_.x.download(x, bytes);
    }
    void inv(int n_var, double* x) const override
    {
        const size_t bytes = op_.state_bytes() / 3 * n_var;
        op_.x.resize(bytes);
        op_.x.upload(x, bytes);
        check(wdg_hip_mass_apply(op_.get(), n_var, op_.x.get(), op_.x.get(), 1, nullptr));
        op_.x.download(x, bytes);
    }
    void inv_device(void* d_x, void* stream = nullptr) const { check(wdg_hip_mass_inv(op_.get(), d_x, stream)); }
    void inv_device(int n_var, void* d_x, void* stream = nullptr) const
    {
        check(wdg_hip_mass_apply(op_.get(), n_var, d_x, d_x, 1, stream));
    }

private:
    detail::Wave2D op_;
};

// Advection<true> (Advection.hpp:15-140) and AdvectionHomogeneousBC<true>
// (Advection.hpp:142-270): n_var <= 8 variables, coefficient `a` of shape
// (n_var, n_var, 2) if constant_coefficient else (n_var, n_var, 2, n, n,
// n_elem), column-major as in the reference; upwind flux.  n.A must have
// real eigenvalues and a full set of eigenvectors at every face node, as
// real_eig requires (eig.cpp:29-74).  <false> integrates on Gauss-Legendre rules of
// quad->n points (the reference's automatic sizes when quad is null).
namespace detail {
class AdvectionHandle {
public:
    AdvectionHandle(const Mesh2D& mesh, const QuadratureRule* basis, int n_var, const double* a, bool cc,
                    bool approx_quad, const QuadratureRule* quad)
        : op_(mesh, basis, nullptr, 1.0, 0.5, Precision::F64), n_var_(n_var)
    {
        wdg_advection h = nullptr;
        if (approx_quad) {
            check(wdg_hip_advection_create(op_.get(), n_var, a, cc ? 1 : 0, &h));
        } else {
            const int nq = exact_rule_size(quad);
            check(wdg_hip_advection_create_exact(op_.get(), n_var, a, cc ? 1 : 0, nq, nq, &h));
        }
        h_.reset(h);
    }
    wdg_advection get() const { return h_.get(); }
    size_t bytes() const { return (size_t)op_.ndof() / 3 * n_var_ * sizeof(double); }
    mutable DeviceBuffer x, y;

private:
    struct Del {
        void operator()(wdg_advection p) const { wdg_hip_advection_destroy(p); }
    };
    Wave2D op_;  // declared before h_: destroyed after it
    int n_var_;
    std::unique_ptr<wdg_advection_s, Del> h_;
};
}  // namespace detail

template <bool ApproxQuadrature>
class Advection : public Operator {
public:
    Advection(int n_var, const Mesh2D& mesh, const QuadratureRule* basis, const double* a, bool constant_coefficient,
              const QuadratureRule* quad = nullptr)
        : h_(mesh, basis, n_var, a, constant_coefficient, ApproxQuadrature, quad)
    {
    }
    void action(const double* u, double* divF) const override
    {
        const size_t bytes = h_.bytes();
        h_.x.resize(bytes);
        h_.y.resize(bytes);
        h_.x.upload(u, bytes);
        h_.y.upload(divF, bytes);
        action_device(h_.x.get(), h_.y.get());
        h_.y.download(divF, bytes);
    }
    void action(int, const double* u, double* divF) const override { action(u, divF); }
    void action_device(const void* d_u, void* d_divF, void* stream = nullptr) const
    {
        check(wdg_hip_advection_action(h_.get(), d_u, d_divF, stream));
    }

private:
    detail::AdvectionHandle h_;
};

template <bool ApproxQuadrature>
class AdvectionHomogeneousBC : public Operator {
public:
    AdvectionHomogeneousBC(int n_var, const Mesh2D& mesh, const QuadratureRule* basis, const double* a,
                           bool constant_coefficient, const QuadratureRule* quad = nullptr)
        : h_(mesh, basis, n_var, a, constant_coefficient, ApproxQuadrature, quad)
    {
    }
    void action(const double* u, double* divF) const override
    {
        const size_t bytes = h_.bytes();
        h_.x.resize(bytes);
        h_.y.resize(bytes);
        h_.x.upload(u, bytes);
        h_.y.upload(divF, bytes);
        action_device(h_.x.get(), h_.y.get());
        h_.y.download(divF, bytes);
    }
    void action(int, const double* u, double* divF) const override { action(u, divF); }
    void action_device(const void* d_u, void* d_divF, void* stream = nullptr) const
    {
        check(wdg_hip_advection_bc_action(h_.get(), d_u, d_divF, stream));
    }

private:
    detail::AdvectionHandle h_;
};

// DivF2D and EdgeFluxF2D (DivF.hpp:207-392, EdgeFluxF.hpp:136-250; <true>
// collocated, <false> on Gauss-Legendre rules of quad->n points, 2 n when
// quad is null) with the functors as HIP source text: the reference's lambdas
// become FluxSource strings holding
//   __device__ void flux(const double* x, const double* u, double* F);
//   __device__ void numerical_flux(const double* x, const double* n,
//                                  const double* uL, const double* uR, double* fh);
// (and optionally exterior_state, see wdg_hip_flux_create), compiled for
// the GPU on first use and cached per (source, n_var).  EdgeFluxF2D::action
// covers what the reference does with a face prolongator around the flux:
// prolongation, the numerical flux times w ds (+ on el0, - on el1), a scale
// (cl2d.cpp uses -1) and the transpose into divF.
struct FluxSource {
    std::string source;
};

namespace detail {
// approx_quad = false: nq points of the Gauss-Legendre rule (0 = 2 n, the
// reference's default, DivF.cpp:56-60, EdgeFluxF.cpp:51-56)
class FluxCache {
public:
    FluxCache(const Mesh2D& mesh, const QuadratureRule* basis, bool approx_quad, const QuadratureRule* quad)
        : op_(mesh, basis, nullptr, 1.0, 0.5, Precision::F64), approx_(approx_quad),
          nq_(approx_quad ? 0 : exact_rule_size(quad))
    {
    }
    wdg_flux get(int n_var, const FluxSource& f) const
    {
        const std::string key = std::to_string(n_var) + ":" + f.source;
        auto it = cache_.find(key);
        if (it != cache_.end()) return it->second.get();
        wdg_flux h = nullptr;
        if (approx_)
            check(wdg_hip_flux_create(op_.get(), n_var, f.source.c_str(), &h));
        else
            check(wdg_hip_flux_create_exact(op_.get(), n_var, f.source.c_str(), nq_, nq_, &h));
        std::unique_ptr<wdg_flux_s, Del> p(h);
        return cache_.emplace(key, std::move(p)).first->second.get();
    }
    size_t bytes(int n_var) const { return (size_t)op_.ndof() / 3 * n_var * sizeof(double); }
    mutable DeviceBuffer x, y;

private:
    struct Del {
        void operator()(wdg_flux p) const { wdg_hip_flux_destroy(p); }
    };
    Wave2D op_;  // declared before cache_: destroyed after it
    bool approx_;
    int nq_;
    mutable std::unordered_map<std::string, std::unique_ptr<wdg_flux_s, Del>> cache_;
};
}  // namespace detail

template <bool ApproxQuadrature>
class DivF2D {
public:
    DivF2D(const Mesh2D& mesh, const QuadratureRule* basis, const QuadratureRule* quad = nullptr)
        : c_(mesh, basis, ApproxQuadrature, quad)
    {
    }
    // divF += (F(x, u), grad v); host arrays (n_var, n, n, n_elem)
    void action(int n_var, const FluxSource& f, const double* u, double* divF) const
    {
        const size_t bytes = c_.bytes(n_var);
        c_.x.resize(bytes);
        c_.y.resize(bytes);
        c_.x.upload(u, bytes);
        c_.y.upload(divF, bytes);
        action_device(n_var, f, c_.x.get(), c_.y.get());
        c_.y.download(divF, bytes);
    }
    void action_device(int n_var, const FluxSource& f, const void* d_u, void* d_divF, void* stream = nullptr) const
    {
        check(wdg_hip_divf2d(c_.get(n_var, f), d_u, d_divF, stream));
    }

private:
    detail::FluxCache c_;
};

template <bool ApproxQuadrature>
class EdgeFluxF2D {
public:
    EdgeFluxF2D(const Mesh2D& mesh, FaceType face_type, const QuadratureRule* basis,
                const QuadratureRule* quad = nullptr)
        : c_(mesh, basis, ApproxQuadrature, quad), type_(face_type)
    {
    }
    // divF += scale * lifted (+-fh w ds) on this object's faces; host arrays
    void action(int n_var, const FluxSource& f, const double* u, double* divF, double scale = 1.0) const
    {
        const size_t bytes = c_.bytes(n_var);
        c_.x.resize(bytes);
        c_.y.resize(bytes);
        c_.x.upload(u, bytes);
        c_.y.upload(divF, bytes);
        action_device(n_var, f, c_.x.get(), c_.y.get(), scale);
        c_.y.download(divF, bytes);
    }
    void action_device(int n_var, const FluxSource& f, const void* d_u, void* d_divF, double scale = 1.0,
                       void* stream = nullptr) const
    {
        check(wdg_hip_edgefluxf2d(c_.get(n_var, f), type_ == FaceType::INTERIOR ? WDG_INTERIOR : WDG_BOUNDARY, scale,
                                  d_u, d_divF, stream));
    }

private:
    detail::FluxCache c_;
    FaceType type_;
};

// ode::ForwardEuler, RungeKutta2, RungeKutta4, SSPRK3 (ode.hpp:12-186) on
// device arrays: the functor is called as f(d_F, t, d_u) with device
// pointers and evaluates the right-hand side on the GPU (e.g. with the
// *_device methods); the vector updates between stages run on the GPU.
namespace ode {
class Integrator {
public:
    template <typename Func>
    void step(double dt, Func f, double& t, double* d_u, void* stream = nullptr) const
    {
        double* p = static_cast<double*>(p_.get());
        double* w = static_cast<double*>(w_.get());
        for (int k = 0; k < ns_; ++k) {
            const double* x = (k == 0 || method_ <= 1) ? d_u : w;
            f(p, t + c_[k] * dt, x);
            check(wdg_hip_ode_update(method_, k, n_, dt, d_u, w, p, stream));
        }
        t += dt;
    }

protected:
    Integrator(int method, long long n) : method_(method), n_(n), p_(8 * (size_t)n), w_(method ? 8 * (size_t)n : 0)
    {
        check(wdg_hip_ode_info(method, &ns_, c_));
    }

private:
    int method_;
    long long n_;
    int ns_ = 0;
    double c_[5] = {};
    DeviceBuffer p_, w_;
};
class ForwardEuler : public Integrator {
public:
    explicit ForwardEuler(long long n) : Integrator(0, n) {}
};
class RungeKutta4 : public Integrator {
public:
    explicit RungeKutta4(long long n) : Integrator(1, n) {}
};
class RungeKutta2 : public Integrator {
public:
    explicit RungeKutta2(long long n) : Integrator(2, n) {}
};
class SSPRK3 : public Integrator {
public:
    explicit SSPRK3(long long n) : Integrator(3, n) {}
};
}  // namespace ode

// The semi-discrete wave system u' = M^{-1}(A + B) u advanced by the
// low-storage RK4 of ode::RungeKutta4 (ode.hpp:46-93), one fused kernel per
// stage.  Arrays are device arrays of the chosen precision.
class WaveSystem {
public:
    WaveSystem(const Mesh2D& mesh, const QuadratureRule* basis, const int* bc, Precision p = Precision::F64,
               double a = 1.0, double b = 0.5, bool approx_quad = true)
        : op_(mesh, basis, bc, a, b, p, approx_quad)
    {
    }
    long long ndof() const { return op_.ndof(); }
    // d_y <- M^{-1}(A d_x + B d_x + s d_F)
    void rhs_device(const void* d_x, void* d_y, const void* d_F = nullptr, double s = 0.0,
                    void* stream = nullptr) const
    {
        check(wdg_hip_wave2d_rhs(op_.get(), d_x, d_y, d_F, s, stream));
    }
    // n RungeKutta4::step calls; t advances
    void rk4_steps_device(double dt, double& t, void* d_u, int n, void* stream = nullptr) const
    {
        check(wdg_hip_rk4_steps(op_.get(), dt, &t, d_u, n, nullptr, 0.0, stream));
    }
    wdg_wave2d handle() const { return op_.get(); }

private:
    detail::Wave2D op_;
};

// WaveHoltz (WaveHoltz.hpp:35-90, WaveHoltz.cpp:5-145).  With Precision::F32
// the RK state runs in float while the filter accumulation, the fixed-point
// update and the norms stay in double; the arrays passed in are always double.
// approx_quad defaults to false (exact quadrature) as in the reference.
class WaveHoltz {
public:
    WaveHoltz(double omega, const Mesh2D& mesh, const QuadratureRule* basis, const int* boundary_conditions,
              bool approx_quad = false, Precision p = Precision::F64)
        : omega_(omega), op_(mesh, basis, boundary_conditions, 1.0, 0.5, p, approx_quad)
    {
        wdg_waveholtz h = nullptr;
        check(wdg_hip_waveholtz_create(op_.get(), omega, &h));
        wh_.reset(h);
        check(wdg_hip_waveholtz_info(h, &T_, &dt_, &nt_));
    }
    double K(double t) const { return omega_ / M_PI * (std::cos(omega_ * t) - 0.25); }
    double T() const { return T_; }
    double dt() const { return dt_; }
    int nt() const { return nt_; }
    long long ndof() const { return op_.ndof(); }

    void pi0(double* w, const double* F) const
    {
        const size_t bytes = 8 * (size_t)op_.ndof();
        op_.x.resize(bytes);
        op_.y.resize(bytes);
        op_.y.upload(F, bytes);
        pi0_device(static_cast<double*>(op_.x.get()), static_cast<const double*>(op_.y.get()));
        op_.x.download(w, bytes);
    }
    void S(double* w) const
    {
        const size_t bytes = 8 * (size_t)op_.ndof();
        op_.x.resize(bytes);
        op_.x.upload(w, bytes);
        S_device(static_cast<double*>(op_.x.get()));
        op_.x.download(w, bytes);
    }
    void postprocess(double* H, const double* w) const
    {
        const size_t bytes = 8 * (size_t)op_.ndof();
        op_.x.resize(bytes);
        op_.y.resize(bytes);
        op_.x.upload(w, bytes);
        check(wdg_hip_waveholtz_postprocess(wh_.get(), static_cast<double*>(op_.y.get()),
                                            static_cast<const double*>(op_.x.get()), nullptr));
        op_.y.download(H, 2 * bytes / 3);
    }

    void pi0_device(double* d_w, const double* d_F, void* stream = nullptr) const
    {
        check(wdg_hip_waveholtz_pi0(wh_.get(), d_w, d_F, stream));
    }
    void S_device(double* d_w, void* stream = nullptr) const { check(wdg_hip_waveholtz_S(wh_.get(), d_w, stream)); }
    // d_w <- S d_w + d_pi0; returns ||w_new - w_old|| / pi0_norm (waveholtz2d.cpp:165-185)
    double iterate_device(double* d_w, const double* d_pi0, double pi0_norm, void* stream = nullptr) const
    {
        double r = 0.0;
        check(wdg_hip_waveholtz_iterate(wh_.get(), d_w, d_pi0, pi0_norm, &r, stream));
        return r;
    }
    // Checkpoint of a device iterate (raw float64 + a .meta record) and its
    // resumption; returns the recorded iteration count.
    void checkpoint_device(const std::string& path, const double* d_W, int iterations, double residual) const
    {
        check(wdg_hip_waveholtz_checkpoint(wh_.get(), path.c_str(), d_W, iterations, residual));
    }
    int resume_device(const std::string& path, double* d_W, double* residual = nullptr) const
    {
        int it = 0;
        check(wdg_hip_waveholtz_resume(wh_.get(), path.c_str(), d_W, &it, residual));
        return it;
    }
    wdg_wave2d wave_handle() const { return op_.get(); }

private:
    struct Del {
        void operator()(wdg_waveholtz p) const { wdg_hip_waveholtz_destroy(p); }
    };
    double omega_;
    detail::Wave2D op_;
    std::unique_ptr<wdg_waveholtz_s, Del> wh_;
    double T_ = 0, dt_ = 0;
    int nt_ = 0;
};

// LinearFunctional2D (LinearFunctional.hpp:136-202) with Gauss-Legendre(n)
// quadrature, for forcing functions that are sums of Gaussians in one
// variable: f(x) = sum_k amp_k exp(-|x - c_k|^2 / (2 sigma_k^2)).  sources is
// (4, n_sources): cx, cy, sigma, amp.
class LinearFunctional2D {
public:
    // quad: a Gauss-Legendre rule for the integrals, or null for the
    // reference's default (n_colloc points, LinearFunctional.hpp:42-53)
    LinearFunctional2D(const Mesh2D& mesh, const QuadratureRule* basis, const QuadratureRule* quad = nullptr)
        : op_(mesh, basis, nullptr, 1.0, 0.5, Precision::F64), mesh_(mesh), n_(basis->n),
          nq_(detail::exact_rule_size(quad))
    {
    }
    // LinearFunctional2D::action(n_var, f, F) (LinearFunctional.hpp:136-202):
    // f(const double x[2], double fx[n_var]) is evaluated on the host at every
    // quadrature point; the projection F = (f, phi) runs on the GPU.
    template <typename Func>
    void action(int n_var, Func f, double* F) const
    {
        const int nq = nq_ > 0 ? nq_ : n_;
        const int nel = mesh_.n_elem();
        std::vector<double> x(2 * (size_t)nq * nq * nel), fq((size_t)n_var * nq * nq * nel);
        check(wdg_hip_mesh_gl_points(mesh_.handle(), nq, x.data()));
        for (size_t p = 0; p < (size_t)nq * nq * nel; ++p) f(&x[2 * p], &fq[(size_t)n_var * p]);
        DeviceBuffer d_fq(8 * fq.size());
        const size_t bytes = 8 * (size_t)n_var * n_ * n_ * nel;
        op_.x.resize(bytes);
        d_fq.upload(fq.data(), 8 * fq.size());
        check(wdg_hip_linear_functional(op_.get(), n_var, nq, static_cast<const double*>(d_fq.get()),
                                        static_cast<double*>(op_.x.get()), nullptr));
        op_.x.download(F, bytes);
    }
    void gaussians(int n_sources, const double* sources, int var, double* F) const
    {
        const size_t bytes = 8 * (size_t)op_.ndof();
        op_.x.resize(bytes);
        gaussians_device(n_sources, sources, var, static_cast<double*>(op_.x.get()));
        op_.x.download(F, bytes);
    }
    void gaussians_device(int n_sources, const double* sources, int var, double* d_F, void* stream = nullptr) const
    {
        check(wdg_hip_load_vector_gauss(op_.get(), n_sources, sources, var, d_F, stream));
    }

private:
    detail::Wave2D op_;
    Mesh2D mesh_;
    int n_, nq_;
};

// norm / error (linalg.cpp:102-130) of device arrays, in double.
// The examples' to_file (examples/examples.hpp:14-19) for a device array:
// raw native-endian float64.
inline void to_file_device(const std::string& fname, long long n, const double* d_u)
{
    check(wdg_hip_to_file(fname.c_str(), d_u, n, WDG_F64));
}

inline double norm_device(long long n, const double* d_x, void* stream = nullptr)
{
    double r = 0;
    check(wdg_hip_norm(d_x, n, &r, stream));
    return r;
}
inline double error_device(long long n, const double* d_x, const double* d_y, void* stream = nullptr)
{
    double r = 0;
    check(wdg_hip_error(d_x, d_y, n, &r, stream));
    return r;
}

}  // namespace hip
}  // namespace dg
