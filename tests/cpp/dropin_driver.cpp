// Drives the dg::hip drop-in classes the way examples/waveholtz2d.cpp and
// examples/wave2d.cpp drive the reference, and writes raw float64 results for
// tests/test_gpu_dropin.py to compare with the oracle.
//
// usage: dropin_driver <out_dir> <mesh_dir>
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "wavedg_amd/include/wavedg_hip/dg.hpp"

using namespace dg::hip;

static void to_file(const std::string& fname, size_t n, const double* u)
{
    std::ofstream out(fname, std::ios::out | std::ios::binary);
    out.write(reinterpret_cast<const char*>(u), n * sizeof(double));
}

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <out_dir> <mesh_dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1], mesh_dir = argv[2];
    try {
        // --- WaveHoltz, as examples/waveholtz2d.cpp:92-190 (smaller mesh)
        const int n_colloc = 5;
        auto basis = QuadratureRule::quadrature_rule(n_colloc, QuadratureRule::GaussLobatto);
        Mesh2D mesh = Mesh2D::uniform_rect(8, -1.0, 1.0, 8, -1.0, 1.0);
        const int nB = mesh.n_edges(FaceType::BOUNDARY);
        std::vector<double> geo = mesh.edge_geometry(FaceType::BOUNDARY);
        std::vector<int> bc(nB);
        for (int e = 0; e < nB; ++e) {
            const bool left = std::abs(geo[5 * e + 3] + 1.0) < 1e-12, bottom = std::abs(geo[5 * e + 4] + 1.0) < 1e-12;
            bc[e] = (left || bottom) ? 1 : 0;  // reflecting on the left and bottom walls, absorbing elsewhere
        }
        const double omega = 10.0;
        WaveHoltz WH(omega, mesh, basis, bc.data(), true);
        const long long ndof = WH.ndof();
        // forcing (w/pi) exp(-w r^2), w = (10 pi)^2, centred at (-0.7, -0.1) (waveholtz2d.cpp:41-47)
        const double w2 = std::pow(10.0 * M_PI, 2);
        const double src[4] = {-0.7, -0.1, 1.0 / std::sqrt(2.0 * w2), w2 / M_PI};
        LinearFunctional2D L(mesh, basis);
        std::vector<double> F(ndof), pi0(ndof), W(ndof, 0.0), Wprev(ndof), H(2 * ndof / 3);
        L.gaussians(1, src, 0, F.data());
        WH.pi0(pi0.data(), F.data());
        double pi_zero = 0;
        for (double v : pi0) pi_zero += v * v;
        pi_zero = std::sqrt(pi_zero);
        std::FILE* log = std::fopen((dir + "/errs.txt").c_str(), "w");
        for (int it = 1; it <= 3; ++it) {
            Wprev = W;
            WH.S(W.data());
            double err = 0;
            for (long long i = 0; i < ndof; ++i) {
                W[i] += pi0[i];
                err += (W[i] - Wprev[i]) * (W[i] - Wprev[i]);
            }
            std::fprintf(log, "%.17g\n", std::sqrt(err) / pi_zero);
        }
        std::fclose(log);
        WH.postprocess(H.data(), W.data());
        to_file(dir + "/F.bin", ndof, F.data());
        to_file(dir + "/pi0.bin", ndof, pi0.data());
        to_file(dir + "/W.bin", ndof, W.data());
        to_file(dir + "/H.bin", H.size(), H.data());
        std::vector<double> bcd(bc.begin(), bc.end());
        to_file(dir + "/bc.bin", bcd.size(), bcd.data());

        // --- wave2d time derivative as examples/wave2d.cpp:164-177, host arrays
        Mesh2D mesh2 = Mesh2D::from_file(mesh_dir);
        auto basis2 = QuadratureRule::quadrature_rule(4, QuadratureRule::GaussLobatto);
        const int nB2 = mesh2.n_edges(FaceType::BOUNDARY);
        std::vector<int> bc2(nB2, 1);
        WaveEquation a(mesh2, basis2, true);
        WaveBC b(mesh2, bc2.data(), basis2, true);
        MassMatrix<true> m(mesh2, basis2);
        const size_t nd2 = 3 * (size_t)16 * mesh2.n_elem();
        std::vector<double> u(nd2), dudt(nd2);
        for (size_t i = 0; i < nd2; ++i) u[i] = std::sin(0.37 * i) + 0.25 * std::cos(1.3 * i);
        for (size_t i = 0; i < nd2; ++i) dudt[i] = 0.0;
        a.action(u.data(), dudt.data());
        b.action(u.data(), dudt.data());
        m.inv(3, dudt.data());
        to_file(dir + "/u2.bin", nd2, u.data());
        to_file(dir + "/dudt2.bin", nd2, dudt.data());

        // --- the same with approx_quad = false, the reference's default
        // (Div<false>, EdgeFlux<false>, MassMatrix<false>)
        WaveEquation ax(mesh2, basis2);
        WaveBC bx(mesh2, bc2.data(), basis2, false);
        MassMatrix<false> mx(mesh2, basis2);
        std::fill(dudt.begin(), dudt.end(), 0.0);
        ax.action(u.data(), dudt.data());
        bx.action(u.data(), dudt.data());
        mx.inv(3, dudt.data());
        to_file(dir + "/dudt2_exact.bin", nd2, dudt.data());
        // a user-given 6-point Gauss-Legendre rule for the volume and edge
        // integrals, and a 7-point rule for the mass matrix
        auto q6 = QuadratureRule::quadrature_rule(6);
        WaveEquation aq(mesh2, basis2, false, q6);
        WaveBC bq(mesh2, bc2.data(), basis2, false, q6);
        std::fill(dudt.begin(), dudt.end(), 0.0);
        aq.action(u.data(), dudt.data());
        bq.action(u.data(), dudt.data());
        to_file(dir + "/AB2_q6.bin", nd2, dudt.data());
        MassMatrix<false> mq(mesh2, basis2, QuadratureRule::quadrature_rule(7));
        mq.action(u.data(), dudt.data());
        to_file(dir + "/M2_q7.bin", nd2, dudt.data());
        bool refused = false;
        try {
            WaveEquation bad(mesh2, basis2, false, QuadratureRule::quadrature_rule(6, QuadratureRule::GaussLobatto));
        } catch (const std::invalid_argument&) {
            refused = true;
        }
        if (!refused) {
            std::fprintf(stderr, "a Gauss-Lobatto quad for approx_quad = false was accepted\n");
            return 1;
        }
        // WaveHoltz with its default approx_quad = false: pi0 and two iterations
        WaveHoltz WHx(omega, mesh, basis, bc.data());
        std::vector<double> pi0x(ndof), Wx(ndof, 0.0);
        WHx.pi0(pi0x.data(), F.data());
        for (int it = 0; it < 2; ++it) {
            WHx.S(Wx.data());
            for (long long i = 0; i < ndof; ++i) Wx[i] += pi0x[i];
        }
        to_file(dir + "/pi0_exact.bin", ndof, pi0x.data());
        to_file(dir + "/W_exact.bin", ndof, Wx.data());
        // --- scalar advection as examples/advec2d.cpp builds it (GLL basis,
        // approx_quad = true, variable coefficient c = r (y, -x)), one time
        // derivative m \ (a u + bc u)
        {
            const int n3 = 5;
            auto basis3 = QuadratureRule::quadrature_rule(n3, QuadratureRule::GaussLobatto);
            Mesh2D mesh3 = Mesh2D::uniform_rect(6, -1.0, 1.0, 6, -1.0, 1.0);
            const std::vector<double> x3 = mesh3.physical_coordinates(basis3);  // (2, n, n, nel)
            const size_t np = (size_t)n3 * n3 * mesh3.n_elem();
            std::vector<double> c(2 * np), u3(np), du3(np, 0.0);
            for (size_t q = 0; q < np; ++q) {
                const double x = x3[2 * q], y = x3[2 * q + 1], r = std::hypot(x, y);
                c[2 * q] = r * y;
                c[2 * q + 1] = -r * x;
                u3[q] = r * std::exp(-10.0 * r * r) * std::cos(3.0 * std::atan2(y, x));
            }
            // the initial condition projected as advec2d does: LF.action then m.inv
            LinearFunctional2D LF3(mesh3, basis3);
            std::vector<double> u_lf(np);
            LF3.action(1, [](const double x[2], double* F) {
                const double r = std::hypot(x[0], x[1]);
                *F = r * std::exp(-10.0 * r * r) * std::cos(3.0 * std::atan2(x[1], x[0]));
            }, u_lf.data());
            to_file(dir + "/adv_lf.bin", np, u_lf.data());
            Advection<true> adv(1, mesh3, basis3, c.data(), false);
            AdvectionHomogeneousBC<true> adv_bc(1, mesh3, basis3, c.data(), false);
            MassMatrix<true> m3(mesh3, basis3);
            adv.action(u3.data(), du3.data());
            adv_bc.action(u3.data(), du3.data());
            m3.inv(1, du3.data());
            to_file(dir + "/adv_c.bin", c.size(), c.data());
            to_file(dir + "/adv_u.bin", np, u3.data());
            to_file(dir + "/adv_dudt.bin", np, du3.data());
            // three SSPRK3 steps on the device, as advec2d's time loop
            DeviceBuffer du(8 * np);
            du.upload(u3.data(), 8 * np);
            ode::SSPRK3 rk(np);
            auto time_derivative = [&](double* dudt, double, const double* uu) {
                check(wdg_hip_memset(dudt, 0, 8 * np));
                adv.action_device(uu, dudt);
                adv_bc.action_device(uu, dudt);
                m3.inv_device(1, dudt);
            };
            double t = 0.0;
            const double dt = 0.01;
            for (int it = 0; it < 3; ++it) rk.step(dt, time_derivative, t, static_cast<double*>(du.get()));
            du.download(du3.data(), 8 * np);
            to_file(dir + "/adv_ssprk3.bin", np, du3.data());
            if (std::abs(t - 0.03) > 1e-15) {
                std::fprintf(stderr, "SSPRK3 time %.17g\n", t);
                return 1;
            }
        }
        // --- examples/advec2d.cpp as written: Gauss-Legendre basis, the
        // coefficient and the initial condition projected by
        // LinearFunctional2D then MassMatrix<true>::inv, one time derivative
        {
            const int n4 = 5;
            auto basis4 = QuadratureRule::quadrature_rule(n4, QuadratureRule::GaussLegendre);
            Mesh2D mesh4 = Mesh2D::uniform_rect(6, -1.0, 1.0, 6, -1.0, 1.0);
            const size_t np = (size_t)n4 * n4 * mesh4.n_elem();
            MassMatrix<true> m4(mesh4, basis4);
            LinearFunctional2D LF4(mesh4, basis4);
            std::vector<double> c(2 * np), u4(np), du4(np, 0.0);
            LF4.action(2, [](const double x[2], double* cc) {
                const double r = std::hypot(x[0], x[1]);
                cc[0] = r * x[1];
                cc[1] = -r * x[0];
            }, c.data());
            m4.inv(2, c.data());
            LF4.action(1, [](const double x[2], double* F) {
                const double r = std::hypot(x[0], x[1]);
                *F = r * std::exp(-10.0 * r * r) * std::cos(3.0 * std::atan2(x[1], x[0]));
            }, u4.data());
            m4.inv(1, u4.data());
            Advection<true> adv(1, mesh4, basis4, c.data(), false);
            AdvectionHomogeneousBC<true> adv_bc(1, mesh4, basis4, c.data(), false);
            adv.action(u4.data(), du4.data());
            adv_bc.action(u4.data(), du4.data());
            m4.inv(1, du4.data());
            to_file(dir + "/advgl_x.bin", 2 * np, mesh4.physical_coordinates(basis4).data());
            to_file(dir + "/advgl_c.bin", c.size(), c.data());
            to_file(dir + "/advgl_u.bin", np, u4.data());
            to_file(dir + "/advgl_dudt.bin", np, du4.data());
            // ... and with approx_quad = false (advec2d's other setting)
            Advection<false> adv_x(1, mesh4, basis4, c.data(), false);
            AdvectionHomogeneousBC<false> adv_bc_x(1, mesh4, basis4, c.data(), false);
            MassMatrix<false> m4x(mesh4, basis4);
            std::fill(du4.begin(), du4.end(), 0.0);
            adv_x.action(u4.data(), du4.data());
            adv_bc_x.action(u4.data(), du4.data());
            m4x.inv(1, du4.data());
            to_file(dir + "/advgl_dudt_exact.bin", np, du4.data());
        }
        // --- examples/cl2d.cpp's time derivative with the flux functors as
        // source: DivF2D, EdgeFluxF2D on interior and boundary faces (scale
        // -1, exterior = interior), MassMatrix<true>::inv, on the isentropic
        // vortex (cl2d.cpp:37-55)
        {
            const FluxSource F{R"(
#define GAMMA 1.4
__device__ inline void euler_fg(const double* q, double* f, double* g)
{
    const double rho = q[0], u = q[1] / rho, v = q[2] / rho, E = q[3];
    const double p = (GAMMA - 1.0) * (E - 0.5 * rho * (u * u + v * v));
    f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = u * (E + p);
    g[0] = rho * v; g[1] = rho * u * v; g[2] = rho * v * v + p; g[3] = v * (E + p);
}
__device__ inline double euler_speed(const double* q)
{
    const double rho = q[0], u = q[1] / rho, v = q[2] / rho;
    const double p = (GAMMA - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v));
    return hypot(u, v) + sqrt(GAMMA * p / rho);
}
__device__ void flux(const double* x, const double* q, double* F) { euler_fg(q, F, F + 4); }
__device__ void numerical_flux(const double* x, const double* n, const double* qL, const double* qR, double* fh)
{
    double fL[4], gL[4], fR[4], gR[4];
    euler_fg(qL, fL, gL);
    euler_fg(qR, fR, gR);
    const double cL = euler_speed(qL), cR = euler_speed(qR), c = cL > cR ? cL : cR;
    for (int d = 0; d < 4; ++d)
        fh[d] = 0.5 * n[0] * (fL[d] + fR[d]) + 0.5 * n[1] * (gL[d] + gR[d]) + 0.5 * c * (qL[d] - qR[d]);
}
)"};
            const int n5 = 4;
            auto basis5 = QuadratureRule::quadrature_rule(n5, QuadratureRule::GaussLobatto);
            Mesh2D mesh5 = Mesh2D::uniform_rect(10, 0.0, 10.0, 10, -6.0, 6.0);
            const size_t np = (size_t)n5 * n5 * mesh5.n_elem();
            const std::vector<double> x5 = mesh5.physical_coordinates(basis5);
            std::vector<double> q(4 * np), dqdt(4 * np, 0.0);
            const double g = 1.4;
            for (size_t k = 0; k < np; ++k) {
                const double x = x5[2 * k], y = x5[2 * k + 1];
                const double r = (x - 5.0) * (x - 5.0) + y * y, b = 2.5 * std::exp(1.0 - r) / M_PI;
                const double u = 1.0 - b * y, v = b * (x - 5.0);
                const double rho = std::pow(1.0 - (g - 1.0) / (4.0 * g) * b * b, 1.0 / (g - 1.0));
                const double p = std::pow(rho, g);
                q[4 * k] = rho;
                q[4 * k + 1] = rho * u;
                q[4 * k + 2] = rho * v;
                q[4 * k + 3] = p / (g - 1.0) + 0.5 * rho * (u * u + v * v);
            }
            DivF2D<true> div(mesh5, basis5);
            EdgeFluxF2D<true> interior_flux(mesh5, FaceType::INTERIOR, basis5);
            EdgeFluxF2D<true> boundary_flux(mesh5, FaceType::BOUNDARY, basis5);
            MassMatrix<true> m5(mesh5, basis5);
            div.action(4, F, q.data(), dqdt.data());
            interior_flux.action(4, F, q.data(), dqdt.data(), -1.0);
            boundary_flux.action(4, F, q.data(), dqdt.data(), -1.0);
            m5.inv(4, dqdt.data());
            to_file(dir + "/cl_q.bin", q.size(), q.data());
            to_file(dir + "/cl_dqdt.bin", dqdt.size(), dqdt.data());
            // the same with approx_quad = false, as cl2d builds it when asked
            DivF2D<false> div_x(mesh5, basis5);
            EdgeFluxF2D<false> interior_x(mesh5, FaceType::INTERIOR, basis5);
            EdgeFluxF2D<false> boundary_x(mesh5, FaceType::BOUNDARY, basis5);
            MassMatrix<false> m5x(mesh5, basis5);
            std::fill(dqdt.begin(), dqdt.end(), 0.0);
            div_x.action(4, F, q.data(), dqdt.data());
            interior_x.action(4, F, q.data(), dqdt.data(), -1.0);
            boundary_x.action(4, F, q.data(), dqdt.data(), -1.0);
            m5x.inv(4, dqdt.data());
            to_file(dir + "/cl_dqdt_exact.bin", dqdt.size(), dqdt.data());
        }
        std::printf("dropin_driver ok: ndof=%lld nt=%d\n", ndof, WH.nt());
    } catch (const Error& e) {
        std::fprintf(stderr, "dg::hip::Error %d: %s\n", e.code, e.what());
        return 1;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
