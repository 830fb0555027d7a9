// The convex-chunk library (host): fracture patterns of the cube [-1,1]^3, each a Voronoi partition of 2-6 seeds,
// each cell a convex polytope. Generated once from the reference's generator and seed (engine/src/hull.rs build :86-134,
// gpu/gpu_world.js hullLibrary :2113-2208), so a hull id names the same cell as in the reference.
#ifndef SF_HULLS_H
#define SF_HULLS_H
#include <array>
#include <cstdint>
#include <vector>

namespace sf {

// limits of one hull (patterns whose cells exceed them are skipped, as the reference skips them)
constexpr int HULL_MAXV = 16, HULL_MAXF = 12, HULL_MAXFV = 8, HULL_MAXS = 24, HULL_MAXD = 12;

struct HullCell {
    using V3d = std::array<double, 3>;
    // in the cell's own box, mapped to [-1,1]^3: vertices, faces (outward unit normal n and offset d of n.x <= d, the
    // vertex loop), edges, distinct edge directions (unit)
    std::vector<V3d> v;
    std::vector<V3d> fn;
    std::vector<double> fd;
    std::vector<std::vector<int>> fv;
    std::vector<int> fother;  // what lies across each face: a cube face 0..5 (+x -x +y -y +z -z) or 6 + a cell's slot
    std::vector<std::array<int, 2>> seg;
    std::vector<V3d> dir;
    // the box in the cube's coordinates (centre, half size); the cell's volume as a share of its box and of the
    // cube; centroid and covariance per unit volume about the centroid, in box coordinates
    V3d bc, bh;
    double box_vol, cube_vol;
    V3d cen;
    double cov[3][3];
    int pattern, slot;  // (-1, 0 for hull 0, the whole cube)
};

struct HullLib {
    std::vector<HullCell> cells;           // cell 0: the whole cube
    std::vector<std::vector<int>> patterns;
};

// the library of 16 patterns from seed 7 (built once, on first use)
const HullLib& hull_lib();

// The device tables, field by field (hull h's entries at h * the field's stride): vertices (HULL_MAXV float4: x y z
// 0), planes (HULL_MAXF float4: n, d), edge directions (HULL_MAXD float4), each face's vertex indices (HULL_MAXF x
// HULL_MAXFV bytes) and vertex count (HULL_MAXF bytes), edges (HULL_MAXS pairs of bytes), and the counts (int4: nv,
// nf, nseg, ndir). f32 copies of the library's f64 values.
struct HullTables {
    int count = 0;
    std::vector<float> v4, pl4, dir4;
    std::vector<uint8_t> fidx, fcnt, seg;
    std::vector<int32_t> head;
};
const HullTables& hull_tables();

}  // namespace sf
#endif
