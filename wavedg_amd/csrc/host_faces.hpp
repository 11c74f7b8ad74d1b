// Per-element-side tables of the fused wave kernels (wave2d_kernels.hpp),
// built on the host from the bilinear map coefficients and the face tables
// of host_mesh.hpp.  Host-only code: no HIP headers.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace wdg {

// Per element side, in the handle's precision: outward unit normal of the
// side and half its length (ds of the straight side, Edge.hpp:107-122).
// The kernels read it for FP32 handles of n_colloc >= 6.
template <typename R>
struct FaceGeom {
    R n0, n1, hl;
};

// Per element side: x = 3 * (flat node index in the neighbour of the node
// facing this side's node 0), or -1 on the boundary; y = the side's info word
// (host_mesh.hpp) with bit 6 set when the side's nodes run opposite to the
// edge's first element (weights are read in that element's order) and bits
// 16-31 = the signed step of x from one side node to the next.  A trace value
// of side node k is u_in[x + k * (y >> 16) + c].
struct alignas(8) FaceNb {
    int32_t x, y;
};

// Both tables for a mesh of nel elements with n_colloc = N: geom6 (6, nel)
// bilinear map coefficients ax, ay, cx, cy, hx, hy; nbr, info (4, nel).
template <typename R>
inline void build_face_tables(int N, int nel, const double* geom6, const int32_t* nbr, const int32_t* info,
                              std::vector<FaceGeom<R>>& fg, std::vector<FaceNb>& fnb)
{
    fg.resize(4 * (size_t)nel);
    fnb.resize(4 * (size_t)nel);
    const int NN = N * N;
    auto side_node_h = [N](int s, int k) {
        const int i = (s == 0 || s == 2) ? k : (s == 1 ? N - 1 : 0);
        const int j = (s == 1 || s == 3) ? k : (s == 2 ? N - 1 : 0);
        return i + N * j;
    };
    for (int e = 0; e < nel; ++e) {
        const double* g = geom6 + 6 * (size_t)e;  // ax, ay, cx, cy, hx, hy
        for (int s = 0; s < 4; ++s) {
            const size_t es = 4 * (size_t)e + s;
            // side vector in traversal order: 2 (a -+ h) for sides 0, 2 and
            // 2 (c +- h) for sides 1, 3; sides 2 and 3 run clockwise
            const double sh = (s == 0 || s == 3) ? -1.0 : 1.0;
            const double dx = 2.0 * ((s & 1) ? g[2] + sh * g[4] : g[0] + sh * g[4]);
            const double dy = 2.0 * ((s & 1) ? g[3] + sh * g[5] : g[1] + sh * g[5]);
            const double len = std::sqrt(dx * dx + dy * dy);
            const double sg = (s >= 2) ? -1.0 : 1.0;
            FaceGeom<R> r;
            r.n0 = (R)(sg * dy / len);
            r.n1 = (R)(-sg * dx / len);
            r.hl = (R)(0.5 * len);
            fg[es] = r;
            const int inf = info[es];
            const int flip = (inf >> 2) & 1, first = (inf >> 3) & 1;
            int x = -1, step = 0;
            if (nbr[es] >= 0) {
                const int ns = inf & 3;
                const int dir = (ns == 0 || ns == 2) ? 1 : N;
                x = 3 * (int)((long long)nbr[es] * NN + side_node_h(ns, flip ? N - 1 : 0));
                step = 3 * (flip ? -dir : dir);
            }
            fnb[es] = FaceNb{x, (inf & 0xffbf) | ((!first && flip) ? 64 : 0) | (int)((unsigned)step << 16)};
        }
    }
}

}  // namespace wdg
