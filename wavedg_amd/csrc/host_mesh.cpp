#include "host_mesh.hpp"

#include <cmath>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <unordered_map>

namespace wdg {

namespace {
// First and second corner of each side, in the traversal direction.
constexpr int kSideA[4] = {0, 1, 3, 0};
constexpr int kSideB[4] = {1, 2, 2, 3};
}  // namespace

HostMesh HostMesh::from_vertices(int nv, const double* x, int nel, const int* elems)
{
    if (nv <= 0 || nel <= 0 || !x || !elems) throw std::invalid_argument("mesh: empty vertex or element array");
    HostMesh m;
    m.nv_ = nv;
    m.nel_ = nel;
    m.corners_.resize(8 * (size_t)nel);
    for (int e = 0; e < nel; ++e)
        for (int k = 0; k < 4; ++k) {
            const int v = elems[4 * (size_t)e + k];
            if (v < 0 || v >= nv) throw std::invalid_argument("mesh: element references a vertex out of range");
            m.corners_[8 * (size_t)e + 2 * k] = x[2 * (size_t)v];
            m.corners_[8 * (size_t)e + 2 * k + 1] = x[2 * (size_t)v + 1];
        }

    std::unordered_map<uint64_t, int> seen;
    seen.reserve(4 * (size_t)nel);
    m.edges_.reserve(2 * (size_t)nel + 64);
    for (int e = 0; e < nel; ++e) {
        for (int s = 0; s < 4; ++s) {
            const int va = elems[4 * (size_t)e + kSideA[s]];
            const int vb = elems[4 * (size_t)e + kSideB[s]];
            const uint64_t key = (uint64_t)std::min(va, vb) * (uint64_t)nv + (uint64_t)std::max(va, vb);
            auto it = seen.find(key);
            if (it == seen.end()) {
                HostEdge E{};
                E.el[0] = e;
                E.el[1] = -1;
                E.side[0] = s;
                E.side[1] = -1;
                E.delta = 1;
                E.interior = false;
                E.x0[0] = x[2 * (size_t)va];
                E.x0[1] = x[2 * (size_t)va + 1];
                E.dx[0] = x[2 * (size_t)vb] - E.x0[0];
                E.dx[1] = x[2 * (size_t)vb + 1] - E.x0[1];
                const double len = std::hypot(E.dx[0], E.dx[1]);
                const double sg = (s >= 2) ? -1.0 : 1.0;  // sides 2, 3 run clockwise
                E.normal[0] = sg * E.dx[1] / len;
                E.normal[1] = -sg * E.dx[0] / len;
                E.half_length = len / 2;
                seen.emplace(key, (int)m.edges_.size());
                m.edges_.push_back(E);
            } else {
                HostEdge& E = m.edges_[it->second];
                if (E.interior) throw std::invalid_argument("mesh: an edge is shared by more than two elements");
                const int first_a = elems[4 * (size_t)E.el[0] + kSideA[E.side[0]]];
                E.el[1] = e;
                E.side[1] = s;
                E.interior = true;
                E.delta = (va == first_a) ? 1 : -1;
            }
        }
    }
    for (int i = 0; i < (int)m.edges_.size(); ++i) (m.edges_[i].interior ? m.interior_ : m.boundary_).push_back(i);
    return m;
}

HostMesh HostMesh::uniform_rect(int nx, double ax, double bx, int ny, double ay, double by)
{
    if (nx <= 0 || ny <= 0) throw std::invalid_argument("uniform_rect: nx, ny must be positive");
    // vertex (i, j) -> i + (nx+1) j; element (i, j) corners counter-clockwise (Mesh2D.cpp:130-163)
    std::vector<double> v(2 * (size_t)(nx + 1) * (ny + 1));
    std::vector<int> q(4 * (size_t)nx * ny);
    const double dx = (bx - ax) / nx, dy = (by - ay) / ny;
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            const size_t k = i + (size_t)(nx + 1) * j;
            v[2 * k] = ax + dx * i;
            v[2 * k + 1] = ay + dy * j;
        }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t e = i + (size_t)nx * j;
            const int v00 = i + (nx + 1) * j;
            q[4 * e + 0] = v00;
            q[4 * e + 1] = v00 + 1;
            q[4 * e + 2] = v00 + 1 + (nx + 1);
            q[4 * e + 3] = v00 + (nx + 1);
        }
    return from_vertices((nx + 1) * (ny + 1), v.data(), nx * ny, q.data());
}

HostMesh HostMesh::from_file(const std::string& dir)
{
    // info.txt: "n_pts n_elem"; coordinates.txt: x y per line; elements.txt: 4 vertex ids per line
    // (Mesh2D.cpp:94-128).
    std::ifstream info(dir + "/info.txt");
    if (!info) throw std::runtime_error("Mesh2D::from_file error: cannot open file: " + dir + "/info.txt");
    int np = 0, ne = 0;
    info >> np >> ne;
    if (np <= 0 || ne <= 0) throw std::runtime_error("Mesh2D::from_file error: bad info.txt in " + dir);
    std::ifstream cf(dir + "/coordinates.txt");
    if (!cf) throw std::runtime_error("Mesh2D::from_file error: cannot open file: " + dir + "/coordinates.txt");
    std::vector<double> v(2 * (size_t)np);
    for (size_t i = 0; i < v.size(); ++i)
        if (!(cf >> v[i])) throw std::runtime_error("Mesh2D::from_file error: short coordinates.txt in " + dir);
    std::ifstream ef(dir + "/elements.txt");
    if (!ef) throw std::runtime_error("Mesh2D::from_file error: cannot open file: " + dir + "/elements.txt");
    std::vector<int> q(4 * (size_t)ne);
    for (size_t i = 0; i < q.size(); ++i)
        if (!(ef >> q[i])) throw std::runtime_error("Mesh2D::from_file error: short elements.txt in " + dir);
    return from_vertices(np, v.data(), ne, q.data());
}

double HostMesh::min_h() const
{
    double h = std::numeric_limits<double>::infinity();
    for (const HostEdge& E : edges_) h = std::min(h, 2.0 * E.half_length);
    return h;
}

void HostMesh::nodes(int nq, const double* xq, double* out) const
{
    for (int e = 0; e < nel_; ++e) {
        const double* c = &corners_[8 * (size_t)e];
        for (int j = 0; j < nq; ++j)
            for (int i = 0; i < nq; ++i) {
                const double xi = xq[i], eta = xq[j];
                const double b0 = 0.25 * (1.0 - xi) * (1.0 - eta), b1 = 0.25 * (1.0 + xi) * (1.0 - eta);
                const double b2 = 0.25 * (1.0 + xi) * (1.0 + eta), b3 = 0.25 * (1.0 - xi) * (1.0 + eta);
                double* o = out + 2 * (i + nq * (j + (size_t)nq * e));
                o[0] = c[0] * b0 + c[2] * b1 + c[4] * b2 + c[6] * b3;
                o[1] = c[1] * b0 + c[3] * b1 + c[5] * b2 + c[7] * b3;
            }
    }
}

double HostMesh::min_corner_detj() const
{
    double mn = std::numeric_limits<double>::infinity();
    for (int e = 0; e < nel_; ++e) {
        const double* c = &corners_[8 * (size_t)e];
        for (int k = 0; k < 4; ++k) {
            const int kp = (k + 1) & 3, km = (k + 3) & 3;
            const double ax = c[2 * kp] - c[2 * k], ay = c[2 * kp + 1] - c[2 * k + 1];
            const double bx = c[2 * km] - c[2 * k], by = c[2 * km + 1] - c[2 * k + 1];
            mn = std::min(mn, ax * by - ay * bx);
        }
    }
    return mn;
}

bool HostMesh::face_tables(const int* bc, FaceTables& t, std::string& err) const
{
    t.nbr.assign(4 * (size_t)nel_, -1);
    t.info.assign(4 * (size_t)nel_, 0);
    for (int k : interior_) {
        const HostEdge& E = edges_[k];
        const int flip = E.delta < 0 ? 1 : 0;
        const int s0 = E.side[0], s1 = E.side[1];
        t.nbr[4 * (size_t)E.el[0] + s0] = E.el[1];
        t.info[4 * (size_t)E.el[0] + s0] = s1 | (flip << 2) | (1 << 3) | (s0 << 4);
        t.nbr[4 * (size_t)E.el[1] + s1] = E.el[0];
        t.info[4 * (size_t)E.el[1] + s1] = s0 | (flip << 2) | (0 << 3) | (s0 << 4);
    }
    for (size_t b = 0; b < boundary_.size(); ++b) {
        const HostEdge& E = edges_[boundary_[b]];
        const int code = bc ? bc[b] : 1;
        if (code < 0 || code > 2) {
            err = "WaveBC error: specified boundary condition not implemented (boundary edge " + std::to_string(b) +
                  ", code " + std::to_string(code) + ").";
            return false;
        }
        t.info[4 * (size_t)E.el[0] + E.side[0]] = (1 << 3) | (E.side[0] << 4) | (code << 8);
    }
    return true;
}

}  // namespace wdg
