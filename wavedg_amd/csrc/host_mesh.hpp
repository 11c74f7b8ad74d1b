// Host-side quadrilateral mesh topology.
//
// Builds the same edge list as the reference Mesh2D::from_vertices
// (Mesh2D.cpp:5-92): edges are numbered in order of first discovery while
// walking elements and their sides 0..3; an edge seen a second time becomes
// interior, with delta = +1 when both elements traverse it in the same
// direction.  Interior and boundary edge lists keep that numbering, so a
// per-boundary-edge array (the bc codes) is indexed exactly as in the
// reference.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace wdg {

struct HostEdge {
    int el[2];    // el[1] = -1 on the boundary
    int side[2];  // 0 bottom (v0->v1), 1 right (v1->v2), 2 top (v3->v2), 3 left (v0->v3)
    int delta;    // +1 same orientation, -1 reversed
    bool interior;
    double x0[2], dx[2];  // first element's traversal of the edge
    double normal[2];     // unit, outward from el[0]
    double half_length;   // ds = length / 2
};

// Per element side, for the GPU kernels.
//   nbr  : neighbour element (or -1 on the boundary)
//   info : bit 0-1 neighbour side, bit 2 reversed node order, bit 3 this
//          element is el[0] of the edge, bit 4-5 side of el[0], bit 8-15 bc code
struct FaceTables {
    std::vector<int32_t> nbr, info;
};

class HostMesh {
public:
    static HostMesh from_vertices(int nv, const double* x, int nel, const int* elems);
    static HostMesh uniform_rect(int nx, double ax, double bx, int ny, double ay, double by);
    static HostMesh from_file(const std::string& dir);

    int n_elem() const { return nel_; }
    int n_vertices() const { return nv_; }
    int n_interior() const { return (int)interior_.size(); }
    int n_boundary() const { return (int)boundary_.size(); }
    const HostEdge& edge(int i, bool interior) const { return edges_[interior ? interior_[i] : boundary_[i]]; }
    const std::vector<double>& corners() const { return corners_; }  // (2, 4, nel)
    double min_h() const;
    void nodes(int nq, const double* xq, double* out) const;  // (2, nq, nq, nel)
    // bc: one code per boundary edge (nullptr = all reflecting).  Returns false on a bad code.
    bool face_tables(const int* bc, FaceTables& t, std::string& err) const;
    // Smallest signed Jacobian determinant over the element corners (<= 0 means an inverted element).
    double min_corner_detj() const;

private:
    int nv_ = 0, nel_ = 0;
    std::vector<double> corners_;
    std::vector<HostEdge> edges_;
    std::vector<int> interior_, boundary_;
};

// Thread-local message for the C ABI's wdg_hip_last_error().
extern thread_local std::string g_last_error;

}  // namespace wdg

// Opaque mesh handle of the C ABI.
struct wdg_mesh_s {
    wdg::HostMesh mesh;
};
