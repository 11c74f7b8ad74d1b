// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the mesh
// topology and face tables (host_mesh.cpp, host_faces.hpp), the basis
// (host_basis.cpp) and the CPU oracle (oracle/wdg_oracle.c) on small meshes.
// Built and run by tests/test_sanitize.py; any sanitizer report aborts.
//
// usage: sanitize_driver <golden_mesh_dir>   (tests/golden/meshes)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "wavedg_amd/csrc/host_basis.hpp"
#include "wavedg_amd/csrc/host_faces.hpp"
#include "wavedg_amd/csrc/host_mesh.hpp"

extern "C" {
#include "oracle/wdg_oracle.h"
}

namespace wdg {
thread_local std::string g_last_error;
}

static int g_fail = 0;
#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++g_fail;                                                             \
        }                                                                         \
    } while (0)

static int side_node(int N, int s, int k)
{
    const int i = (s == 0 || s == 2) ? k : (s == 1 ? N - 1 : 0);
    const int j = (s == 1 || s == 3) ? k : (s == 2 ? N - 1 : 0);
    return i + N * j;
}

// Face tables against the mesh: every interior side node's neighbour entry
// names the neighbour node at the same physical point.
static void check_faces(const wdg::HostMesh& m, int N)
{
    const int nel = m.n_elem();
    std::vector<int> bc(m.n_boundary());
    for (int b = 0; b < m.n_boundary(); ++b) bc[b] = b % 3;
    wdg::FaceTables ft;
    std::string err;
    CHECK(m.face_tables(bc.data(), ft, err));
    std::vector<double> geom(6 * (size_t)nel);
    for (int e = 0; e < nel; ++e) {
        const double* c = &m.corners()[8 * (size_t)e];
        double* g = &geom[6 * (size_t)e];
        g[0] = 0.25 * ((c[2] - c[0]) + (c[4] - c[6]));
        g[1] = 0.25 * ((c[3] - c[1]) + (c[5] - c[7]));
        g[2] = 0.25 * ((c[6] - c[0]) + (c[4] - c[2]));
        g[3] = 0.25 * ((c[7] - c[1]) + (c[5] - c[3]));
        g[4] = 0.25 * ((c[0] - c[2]) + (c[4] - c[6]));
        g[5] = 0.25 * ((c[1] - c[3]) + (c[5] - c[7]));
    }
    std::vector<wdg::FaceGeom<double>> fg;
    std::vector<wdg::FaceGeom<float>> fg32;
    std::vector<wdg::FaceNb> fnb, fnb32;
    wdg::build_face_tables(N, nel, geom.data(), ft.nbr.data(), ft.info.data(), fg, fnb);
    wdg::build_face_tables(N, nel, geom.data(), ft.nbr.data(), ft.info.data(), fg32, fnb32);
    std::vector<double> x, w;
    CHECK(wdg::gll_rule(N, x, w));
    std::vector<double> X(2 * (size_t)N * N * nel);
    m.nodes(N, x.data(), X.data());
    const int NN = N * N;
    int n_bnd = 0;
    for (int e = 0; e < nel; ++e)
        for (int s = 0; s < 4; ++s) {
            const size_t es = 4 * (size_t)e + s;
            CHECK(fnb[es].x == fnb32[es].x && fnb[es].y == fnb32[es].y);
            CHECK(std::abs(fg[es].n0 * fg[es].n0 + fg[es].n1 * fg[es].n1 - 1.0) < 1e-14);
            CHECK(fg[es].hl > 0.0);
            if (ft.nbr[es] < 0) {
                CHECK(fnb[es].x == -1);
                ++n_bnd;
                continue;
            }
            const int nb = ft.nbr[es];
            const int step = fnb[es].y >> 16;
            for (int k = 0; k < N; ++k) {
                const int own = e * NN + side_node(N, s, k);
                const int q = fnb[es].x + k * step;
                CHECK(q % 3 == 0);
                const int other = q / 3;
                CHECK(other / NN == nb);
                CHECK(std::abs(X[2 * (size_t)own] - X[2 * (size_t)other]) < 1e-12);
                CHECK(std::abs(X[2 * (size_t)own + 1] - X[2 * (size_t)other + 1]) < 1e-12);
            }
            // the two sides of an edge see opposite normals and the same length
            int ns = -1;
            for (int t = 0; t < 4; ++t)
                if (ft.nbr[4 * (size_t)nb + t] == e && (ft.info[4 * (size_t)nb + t] & 3) == s) ns = t;
            CHECK(ns >= 0);
            if (ns >= 0) {
                const auto& o = fg[4 * (size_t)nb + ns];
                CHECK(std::abs(o.n0 + fg[es].n0) < 1e-14 && std::abs(o.n1 + fg[es].n1) < 1e-14);
                CHECK(std::abs(o.hl - fg[es].hl) < 1e-14 * fg[es].hl);
            }
        }
    CHECK(n_bnd == m.n_boundary());
}

static void run_oracle(const wdg::HostMesh& m, int n)
{
    const int nel = m.n_elem();
    std::vector<double> v(m.corners().begin(), m.corners().end());  // (2, 4, nel): distinct vertices per element
    std::vector<int> q(4 * (size_t)nel);
    for (int i = 0; i < 4 * nel; ++i) q[i] = i;
    orc_mesh* om = orc_mesh_from_vertices(4 * nel, v.data(), nel, q.data());
    // every element stands alone here (no shared vertices), so all sides are boundary
    CHECK(om && orc_mesh_n_edges(om, 1) == 4 * nel);
    orc_mesh_free(om);
    (void)n;
}

static void run_oracle_mesh(int nv, const double* xv, int nel, const int* el, int n)
{
    orc_mesh* om = orc_mesh_from_vertices(nv, xv, nel, el);
    CHECK(om != nullptr);
    if (!om) return;
    const int nB = orc_mesh_n_edges(om, 1);
    std::vector<int> bc(nB);
    for (int b = 0; b < nB; ++b) bc[b] = b % 3;
    for (int leg = 0; leg < 2; ++leg) {
        orc_wave* W = orc_wave_create_basis(om, n, leg, bc.data(), -1.0, -0.5);
        CHECK(W != nullptr);
        if (!W) continue;
        const long long nd = orc_wave_ndof(W);
        std::vector<double> u(nd), du(nd, 0.0), F(nd, 0.0);
        for (long long i = 0; i < nd; ++i) u[i] = std::sin(0.37 * (double)i);
        orc_wave_action(W, u.data(), du.data());
        orc_wavebc_action(W, u.data(), du.data());
        orc_mass_inv(W, 3, du.data());
        const double src[4] = {0.3, 0.4, 0.2, 1.5};
        orc_load_vector_gauss(om, n, n, 1, src, 0, F.data());
        double t = 0.0;
        orc_rk4_steps(W, 1e-3, &t, u.data(), 2, F.data(), 3.0);
        CHECK(std::isfinite(orc_norm(nd, u.data())));
        orc_wh* H = orc_waveholtz_create(W, 40.0);
        std::vector<double> p(nd, 0.0), Hout(2 * nd / 3);
        orc_waveholtz_pi0(H, p.data(), F.data());
        orc_waveholtz_S(H, p.data());
        orc_waveholtz_postprocess(H, Hout.data(), p.data());
        CHECK(std::isfinite(orc_norm(nd, p.data())));
        orc_waveholtz_free(H);
        const double A[9 * 2] = {0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0};
        orc_adv* V = orc_advection_create(W, 3, A, 1);
        if (V) {
            std::vector<double> y(nd, 0.0);
            orc_advection_action(V, u.data(), y.data());
            orc_advection_bc_action(V, u.data(), y.data());
            orc_advection_free(V);
        }
        std::vector<double> q4(4 * nd / 3), y4(4 * nd / 3, 0.0);
        for (size_t i = 0; i < q4.size(); ++i) q4[i] = (i % 4 == 0 || i % 4 == 3) ? 1.5 + 0.1 * std::sin((double)i) : 0.1;
        orc_divf_euler(W, 1.4, q4.data(), y4.data());
        orc_edgefluxf_euler(W, 1.4, 0, q4.data(), y4.data());
        orc_wave_free(W);
        orc_wave* X = orc_wave_create_exact_basis(om, n, leg, bc.data(), -1.0, -0.5, 0, 0, 0);
        CHECK(X != nullptr);
        if (X) {
            std::vector<double> dx(nd, 0.0);
            orc_wave_action(X, u.data(), dx.data());
            orc_wavebc_action(X, u.data(), dx.data());
            orc_mass_inv(X, 3, dx.data());
            CHECK(std::isfinite(orc_norm(nd, dx.data())));
            orc_wave_free(X);
        }
    }
    orc_mesh_free(om);
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <golden_mesh_dir>\n", argv[0]);
        return 2;
    }
    // basis
    for (int n = 2; n <= 9; ++n) {
        std::vector<double> x, w, D(n * n), B(n * n);
        CHECK(wdg::gll_rule(n, x, w));
        wdg::lagrange_derivative_matrix(n, x.data(), n, x.data(), D.data());
        wdg::lagrange_interpolation_matrix(n, x.data(), n, x.data(), B.data());
        for (int i = 0; i < n; ++i) CHECK(std::abs(B[i + n * i] - 1.0) < 1e-14);
        std::vector<double> xq, wq, A(n * n), Bm(n * n), V(n * n), lam(n);
        CHECK(wdg::gauss_legendre_rule(n + 1, xq, wq));
        CHECK(wdg::mass_fast_diag(n, x.data(), n + 1, xq.data(), wq.data(), A.data(), Bm.data(), V.data(), lam.data()));
        double ox[16], ow[16];
        CHECK(orc_gauss_lobatto(n, ox, ow) == 0);
        std::vector<double> od(n * n);
        orc_lagrange_basis_deriv(od.data(), n, ox, n, ox);
    }
    for (int n = 1; n <= 20; ++n) {
        std::vector<double> x, w;
        CHECK(wdg::gauss_legendre_rule(n, x, w));
    }
    {
        const double C[9] = {2, 1, 0, 1, 2, 1, 0, 1, 2};
        double absC[9];
        CHECK(wdg::symmetric_abs(3, C, absC));
    }
    // |C| of non-symmetric matrices C = L diag(e) L^{-1} (L unit lower
    // triangular, repeated and zero eigenvalues), host routine against the
    // oracle's flux matrix; complex and defective matrices are refused
    {
        unsigned seed = 12345u;
        auto rnd = [&seed] {
            seed = seed * 1664525u + 1013904223u;
            return (seed >> 8) / double(1u << 24) - 0.5;
        };
        for (int m = 1; m <= 8; ++m)
            for (int trial = 0; trial < 20; ++trial) {
                double L[64] = {0}, Li[64] = {0}, e[8], C[64], absC[64], F[64];
                for (int j = 0; j < m; ++j)
                    for (int i = 0; i < m; ++i) L[i + m * j] = i == j ? 1.0 : (i > j ? rnd() : 0.0);
                for (int j = 0; j < m; ++j) {  // columns of L^{-1} by forward substitution
                    for (int i = 0; i < m; ++i) {
                        double v = i == j ? 1.0 : 0.0;
                        for (int k = 0; k < i; ++k) v -= L[i + m * k] * Li[k + m * j];
                        Li[i + m * j] = v;
                    }
                }
                for (int i = 0; i < m; ++i) e[i] = trial % 2 ? std::floor(4.0 * rnd()) : rnd();
                for (int j = 0; j < m; ++j)
                    for (int i = 0; i < m; ++i) {
                        double v = 0.0;
                        for (int k = 0; k < m; ++k) v += L[i + m * k] * e[k] * Li[k + m * j];
                        C[i + m * j] = v;
                    }
                CHECK(wdg::real_abs(m, C, absC) == 0);
                CHECK(orc_flux_matrix(m, C, -1.0, -0.5, 0, F) == 0);
                double worst = 0.0, scale = 1e-300;
                for (int j = 0; j < m; ++j)
                    for (int i = 0; i < m; ++i) {
                        scale = std::fmax(scale, std::fabs(C[i + m * j]));
                        worst = std::fmax(worst, std::fabs(F[j + m * i] - (-0.5 * C[i + m * j] - 0.5 * absC[i + m * j])));
                    }
                CHECK(worst <= 1e-9 * scale);
            }
        const double rot[4] = {0.0, 1.0, -1.0, 0.0}, jordan[4] = {1.0, 0.0, 1.0, 1.0};
        double out[4];
        CHECK(wdg::real_abs(2, rot, out) == 1);
        CHECK(wdg::real_abs(2, jordan, out) == 2);
        CHECK(orc_flux_matrix(2, rot, -1.0, -0.5, 0, out) == 1);
        CHECK(orc_flux_matrix(2, jordan, -1.0, -0.5, 0, out) == 2);
        CHECK(wdg::real_abs(9, rot, out) == 3);
    }
    // meshes: structured, perturbed, L-shaped, and the reference's own
    std::vector<wdg::HostMesh> meshes;
    meshes.push_back(wdg::HostMesh::uniform_rect(5, -1.0, 1.0, 4, -1.0, 1.0));
    {
        const int nx = 6, ny = 5;
        std::vector<double> xv;
        std::vector<int> el;
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const bool in = i > 0 && i < nx && j > 0 && j < ny;
                xv.push_back(i / (double)nx + (in ? 0.03 * std::sin(7.0 * i + 3.0 * j) : 0.0));
                xv.push_back(j / (double)ny + (in ? 0.03 * std::cos(5.0 * i - 2.0 * j) : 0.0));
            }
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (i >= nx / 2 && j >= ny / 2 + 1) continue;  // an L-shaped cut
                const int v0 = i + (nx + 1) * j;
                el.insert(el.end(), {v0, v0 + 1, v0 + 1 + (nx + 1), v0 + (nx + 1)});
            }
        meshes.push_back(wdg::HostMesh::from_vertices((int)xv.size() / 2, xv.data(), (int)el.size() / 4, el.data()));
        for (int n : {3, 5}) run_oracle_mesh((int)xv.size() / 2, xv.data(), (int)el.size() / 4, el.data(), n);
    }
    const std::string gold = argv[1];
    for (const char* name : {"skew", "L", "Square"}) meshes.push_back(wdg::HostMesh::from_file(gold + "/" + name));
    for (const auto& m : meshes) {
        CHECK(m.min_corner_detj() > 0.0);
        CHECK(m.min_h() > 0.0);
        for (int N = 2; N <= 9; ++N) check_faces(m, N);
        run_oracle(m, 4);
    }
    // malformed input is refused without touching memory out of range
    {
        bool threw = false;
        try {
            const double xv[6] = {0, 0, 1, 0, 1, 1};
            const int el[4] = {0, 1, 2, 7};
            (void)wdg::HostMesh::from_vertices(3, xv, 1, el);
        } catch (const std::exception&) {
            threw = true;
        }
        CHECK(threw);
    }
    if (g_fail) {
        std::fprintf(stderr, "%d checks failed\n", g_fail);
        return 1;
    }
    std::printf("sanitize driver ok\n");
    return 0;
}
