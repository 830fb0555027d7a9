// The convex-chunk library (engine/src/hull.rs :86-134, gpu/gpu_world.js hullLibrary :2113-2208): the same generator
// (a 32-bit multiply-xorshift stream from seed 7), the same Voronoi construction from half-spaces and the same limits,
// so the patterns and their cells are the reference's. FP64 throughout; the device receives f32 copies of each cell
// mapped into its own bounding box.
#include "hulls.h"
#include "../../include/sf.h"
#include "sf_host.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace sf {
namespace {

using V3d = std::array<double, 3>;
V3d sub(const V3d& a, const V3d& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
double dot(const V3d& a, const V3d& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
V3d cross(const V3d& a, const V3d& b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
double len(const V3d& a) { return std::sqrt(dot(a, a)); }
V3d unit(const V3d& a) { const double l = len(a); return {a[0] / l, a[1] / l, a[2] / l}; }
// a monotonic stand-in for the angle atan2(y, x), 0..4 once round (only the order of angles is used)
double turn(double y, double x) { const double p = x / (std::fabs(x) + std::fabs(y)); return y < 0.0 ? 3.0 + p : 1.0 - p; }

struct Plane { V3d n; double d; int other; };
struct Poly {
    std::vector<V3d> v;
    std::vector<std::pair<int, std::vector<int>>> faces;  // (plane, vertex loop)
    std::vector<std::array<int, 2>> seg;
    std::vector<V3d> dir;
};

// the polytope of the half-spaces n.x <= d: corners from triples of planes, each plane's corners in angular order,
// the edges, the distinct edge directions
Poly polytope(const std::vector<Plane>& pl) {
    Poly P;
    const size_t np = pl.size();
    for (size_t i = 0; i < np; i++)
        for (size_t j = i + 1; j < np; j++)
            for (size_t k = j + 1; k < np; k++) {
                const V3d bc = cross(pl[j].n, pl[k].n);
                const double det = dot(pl[i].n, bc);
                if (std::fabs(det) < 1e-9) continue;
                const V3d ca = cross(pl[k].n, pl[i].n), ab = cross(pl[i].n, pl[j].n);
                V3d x;
                for (int q = 0; q < 3; q++) x[q] = (pl[i].d * bc[q] + pl[j].d * ca[q] + pl[k].d * ab[q]) / det;
                bool inside = true;
                for (const Plane& p : pl) if (!(dot(p.n, x) <= p.d + 1e-7)) { inside = false; break; }
                if (!inside) continue;
                bool dup = false;
                for (const V3d& w : P.v) if (len(sub(w, x)) < 1e-6) { dup = true; break; }
                if (!dup) P.v.push_back(x);
            }
    for (size_t pi = 0; pi < np; pi++) {
        std::vector<int> idx;
        for (size_t i = 0; i < P.v.size(); i++) if (std::fabs(dot(pl[pi].n, P.v[i]) - pl[pi].d) < 1e-6) idx.push_back((int)i);
        if (idx.size() < 3) continue;
        V3d c = {0, 0, 0};
        for (int i : idx) for (int q = 0; q < 3; q++) c[q] += P.v[i][q];
        for (int q = 0; q < 3; q++) c[q] /= (double)idx.size();
        const V3d u = unit(sub(P.v[idx[0]], c)), w = cross(pl[pi].n, u);
        std::vector<double> ang(P.v.size(), 0.0);
        for (int i : idx) { const V3d r = sub(P.v[i], c); ang[i] = turn(dot(r, w), dot(r, u)); }
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ang[a] < ang[b]; });
        P.faces.push_back({(int)pi, idx});
    }
    for (const auto& f : P.faces) {
        const std::vector<int>& L = f.second;
        for (size_t k = 0; k < L.size(); k++) {
            const int i = L[k], j = L[(k + 1) % L.size()];
            bool have = false;
            for (const auto& e : P.seg) if ((e[0] == i && e[1] == j) || (e[0] == j && e[1] == i)) { have = true; break; }
            if (!have) P.seg.push_back({i, j});
        }
    }
    for (const auto& e : P.seg) {
        const V3d d = unit(sub(P.v[e[1]], P.v[e[0]]));
        bool have = false;
        for (const V3d& x : P.dir) if (std::fabs(dot(x, d)) > 0.9995) { have = true; break; }
        if (!have) P.dir.push_back(d);
    }
    return P;
}

// volume, centroid and covariance per unit volume about the centroid of a convex polytope: tetrahedra from the
// vertices' mean to each face's fan of triangles (the integral of x x^T over a tetrahedron is V/20 times the sum of
// its corners' outer products plus the outer product of their sum)
void mass_props(const std::vector<V3d>& v, const std::vector<std::vector<int>>& faces, double& vol, V3d& cen, double cov[3][3]) {
    V3d o = {0, 0, 0};
    for (const V3d& p : v) for (int q = 0; q < 3; q++) o[q] += p[q];
    for (int q = 0; q < 3; q++) o[q] /= (double)v.size();
    vol = 0.0;
    V3d m1 = {0, 0, 0};
    double m2[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (const auto& f : faces)
        for (size_t t = 1; t + 1 < f.size(); t++) {
            const V3d P[4] = {o, v[f[0]], v[f[t]], v[f[t + 1]]};
            const double tv = std::fabs(dot(sub(P[1], o), cross(sub(P[2], o), sub(P[3], o)))) / 6.0;
            vol += tv;
            V3d S = {0, 0, 0};
            for (const V3d& p : P) for (int q = 0; q < 3; q++) S[q] += p[q];
            for (int q = 0; q < 3; q++) m1[q] += tv * S[q] / 4.0;
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double pp = 0.0;
                    for (const V3d& p : P) pp += p[r] * p[c];
                    m2[r][c] += tv / 20.0 * (pp + S[r] * S[c]);
                }
        }
    for (int q = 0; q < 3; q++) cen[q] = m1[q] / vol;
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) cov[r][c] = m2[r][c] / vol - cen[r] * cen[c];
}

// a cell from its cube-coordinate polytope: the geometry mapped into its bounding box (x_box = (x - c) / h; a plane
// n.x <= d becomes (n h / |n h|).x_box <= (d - n.c) / |n h|), volumes, centroid and covariance
HullCell make_cell(const Poly& P, const std::vector<Plane>& pl, int pattern, int slot) {
    HullCell C;
    V3d lo = {1e300, 1e300, 1e300}, hi = {-1e300, -1e300, -1e300};
    for (const V3d& p : P.v) for (int q = 0; q < 3; q++) { lo[q] = std::min(lo[q], p[q]); hi[q] = std::max(hi[q], p[q]); }
    for (int q = 0; q < 3; q++) { C.bc[q] = (hi[q] + lo[q]) / 2.0; C.bh[q] = (hi[q] - lo[q]) / 2.0; }
    C.pattern = pattern; C.slot = slot;
    for (const V3d& p : P.v) C.v.push_back({(p[0] - C.bc[0]) / C.bh[0], (p[1] - C.bc[1]) / C.bh[1], (p[2] - C.bc[2]) / C.bh[2]});
    std::vector<std::vector<int>> loops;
    for (const auto& f : P.faces) {
        const Plane& p = pl[f.first];
        const V3d n0 = {p.n[0] * C.bh[0], p.n[1] * C.bh[1], p.n[2] * C.bh[2]};
        const double l = len(n0);
        C.fn.push_back({n0[0] / l, n0[1] / l, n0[2] / l});
        C.fd.push_back((p.d - dot(p.n, C.bc)) / l);
        C.fv.push_back(f.second);
        C.fother.push_back(p.other);
        loops.push_back(f.second);
    }
    C.seg = P.seg;
    for (const V3d& d : P.dir) C.dir.push_back(unit({d[0] / C.bh[0], d[1] / C.bh[1], d[2] / C.bh[2]}));
    double vb;
    mass_props(C.v, loops, vb, C.cen, C.cov);
    C.box_vol = vb / 8.0;
    double vc, cc[3][3];
    V3d cenc;
    mass_props(P.v, loops, vc, cenc, cc);
    C.cube_vol = vc / 8.0;
    return C;
}

HullLib build(int n_patterns, uint32_t seed) {
    uint32_t a = seed;
    auto rnd = [&]() -> double {
        a += 0x6D2B79F5u;
        uint32_t t = a;
        t = (t ^ (t >> 15)) * (t | 1u);
        t ^= t + (t ^ (t >> 7)) * (t | 61u);
        return (double)(t ^ (t >> 14)) / 4294967296.0;
    };
    std::vector<Plane> cube;
    const double axes[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int k = 0; k < 6; k++) cube.push_back({{axes[k][0], axes[k][1], axes[k][2]}, 1.0, k});
    HullLib L;
    {
        HullCell whole = make_cell(polytope(cube), cube, -1, 0);
        // (the cube's own moments exactly: unit volume share, centroid 0, covariance 1/3 per axis)
        whole.box_vol = 1.0; whole.cube_vol = 1.0; whole.cen = {0, 0, 0};
        for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) whole.cov[r][c] = r == c ? 1.0 / 3.0 : 0.0;
        L.cells.push_back(whole);
    }
    int p = 0;
    for (int tries = 0; p < n_patterns && tries < 40000; tries++) {
        const int k = 2 + (p % 5);
        std::vector<V3d> seeds(k);
        for (int i = 0; i < k; i++) for (int q = 0; q < 3; q++) seeds[i][q] = rnd() * 1.6 - 0.8;
        std::vector<HullCell> out;
        bool ok = true;
        for (int i = 0; i < k && ok; i++) {
            std::vector<Plane> pl = cube;
            for (int j = 0; j < k; j++) {
                if (j == i) continue;
                const V3d n = unit(sub(seeds[j], seeds[i]));
                const V3d m = {(seeds[i][0] + seeds[j][0]) / 2.0, (seeds[i][1] + seeds[j][1]) / 2.0, (seeds[i][2] + seeds[j][2]) / 2.0};
                pl.push_back({n, dot(n, m), 6 + j});
            }
            const Poly P = polytope(pl);
            V3d lo = {1e300, 1e300, 1e300}, hi = {-1e300, -1e300, -1e300};
            for (const V3d& x : P.v) for (int q = 0; q < 3; q++) { lo[q] = std::min(lo[q], x[q]); hi[q] = std::max(hi[q], x[q]); }
            bool small = false;
            for (int q = 0; q < 3; q++) if ((hi[q] - lo[q]) / 2.0 < 0.15) small = true;
            bool big_face = false;
            for (const auto& f : P.faces) if ((int)f.second.size() > HULL_MAXFV) big_face = true;
            if ((int)P.v.size() > HULL_MAXV || (int)P.faces.size() > HULL_MAXF || (int)P.seg.size() > HULL_MAXS ||
                (int)P.dir.size() > HULL_MAXD || big_face || small) { ok = false; break; }
            out.push_back(make_cell(P, pl, (int)L.patterns.size(), i));
        }
        if (!ok) continue;
        std::vector<int> ids;
        for (HullCell& c : out) { L.cells.push_back(std::move(c)); ids.push_back((int)L.cells.size() - 1); }
        L.patterns.push_back(ids);
        p++;
    }
    return L;
}

HullTables pack(const HullLib& L) {
    HullTables T;
    const int H = (int)L.cells.size();
    T.count = H;
    T.v4.assign((size_t)H * HULL_MAXV * 4, 0.0f);
    T.pl4.assign((size_t)H * HULL_MAXF * 4, 0.0f);
    T.dir4.assign((size_t)H * HULL_MAXD * 4, 0.0f);
    T.fidx.assign((size_t)H * HULL_MAXF * HULL_MAXFV, 0);
    T.fcnt.assign((size_t)H * HULL_MAXF, 0);
    T.seg.assign((size_t)H * HULL_MAXS * 2, 0);
    T.head.assign((size_t)H * 4, 0);
    for (int h = 0; h < H; h++) {
        const HullCell& C = L.cells[h];
        for (size_t k = 0; k < C.v.size(); k++) for (int q = 0; q < 3; q++) T.v4[((size_t)h * HULL_MAXV + k) * 4 + q] = (float)C.v[k][q];
        for (size_t f = 0; f < C.fv.size(); f++) {
            for (int q = 0; q < 3; q++) T.pl4[((size_t)h * HULL_MAXF + f) * 4 + q] = (float)C.fn[f][q];
            T.pl4[((size_t)h * HULL_MAXF + f) * 4 + 3] = (float)C.fd[f];
            T.fcnt[(size_t)h * HULL_MAXF + f] = (uint8_t)C.fv[f].size();
            for (size_t i = 0; i < C.fv[f].size(); i++) T.fidx[((size_t)h * HULL_MAXF + f) * HULL_MAXFV + i] = (uint8_t)C.fv[f][i];
        }
        for (size_t e = 0; e < C.seg.size(); e++) {
            T.seg[((size_t)h * HULL_MAXS + e) * 2] = (uint8_t)C.seg[e][0];
            T.seg[((size_t)h * HULL_MAXS + e) * 2 + 1] = (uint8_t)C.seg[e][1];
        }
        for (size_t d = 0; d < C.dir.size(); d++) for (int q = 0; q < 3; q++) T.dir4[((size_t)h * HULL_MAXD + d) * 4 + q] = (float)C.dir[d][q];
        T.head[4 * (size_t)h] = (int32_t)C.v.size(); T.head[4 * (size_t)h + 1] = (int32_t)C.fv.size();
        T.head[4 * (size_t)h + 2] = (int32_t)C.seg.size(); T.head[4 * (size_t)h + 3] = (int32_t)C.dir.size();
    }
    return T;
}

std::once_flag g_once;
HullLib* g_lib = nullptr;
HullTables* g_tab = nullptr;
void init() {
    std::call_once(g_once, [] {
        g_lib = new HullLib(build(16, 7));
        g_tab = new HullTables(pack(*g_lib));
    });
}

}  // namespace

const HullLib& hull_lib() { init(); return *g_lib; }
const HullTables& hull_tables() { init(); return *g_tab; }

}  // namespace sf

// ---------------------------------------------------------------- C ABI
extern "C" {
int sf_hull_count(void) { return (int)sf::hull_lib().cells.size(); }
int sf_hull_pattern_count(void) { return (int)sf::hull_lib().patterns.size(); }
int sf_hull_pattern(int p, int32_t* ids) {
    const sf::HullLib& L = sf::hull_lib();
    if (p < 0 || p >= (int)L.patterns.size()) { sf::set_error(SF_E_ARG, "no such pattern"); return SF_E_ARG; }
    const std::vector<int>& P = L.patterns[p];
    if (ids) for (size_t k = 0; k < P.size(); k++) ids[k] = P[k];
    return (int)P.size();
}
int sf_hull_cell(int h, double* out) {
    const sf::HullLib& L = sf::hull_lib();
    if (h < 0 || h >= (int)L.cells.size() || !out) { sf::set_error(SF_E_ARG, "no such hull"); return SF_E_ARG; }
    const sf::HullCell& C = L.cells[h];
    out[0] = (double)C.v.size(); out[1] = (double)C.fv.size(); out[2] = (double)C.seg.size(); out[3] = (double)C.dir.size();
    for (int q = 0; q < 3; q++) { out[4 + q] = C.bc[q]; out[7 + q] = C.bh[q]; out[12 + q] = C.cen[q]; }
    out[10] = C.box_vol; out[11] = C.cube_vol;
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) out[15 + 3 * r + c] = C.cov[r][c];
    out[24] = (double)C.pattern; out[25] = (double)C.slot;
    return 0;
}
int sf_hull_vertices(int h, int cube, double* out) {
    const sf::HullLib& L = sf::hull_lib();
    if (h < 0 || h >= (int)L.cells.size()) { sf::set_error(SF_E_ARG, "no such hull"); return SF_E_ARG; }
    const sf::HullCell& C = L.cells[h];
    if (out)
        for (size_t k = 0; k < C.v.size(); k++)
            for (int q = 0; q < 3; q++) out[3 * k + q] = cube ? C.bc[q] + C.bh[q] * C.v[k][q] : C.v[k][q];
    return (int)C.v.size();
}
int sf_hull_face(int h, int f, int cube, double* xyz, double* plane4, int32_t* other) {
    const sf::HullLib& L = sf::hull_lib();
    if (h < 0 || h >= (int)L.cells.size() || f < 0 || f >= (int)L.cells[h].fv.size()) { sf::set_error(SF_E_ARG, "no such face"); return SF_E_ARG; }
    const sf::HullCell& C = L.cells[h];
    const std::vector<int>& F = C.fv[f];
    if (xyz)
        for (size_t k = 0; k < F.size(); k++)
            for (int q = 0; q < 3; q++) xyz[3 * k + q] = cube ? C.bc[q] + C.bh[q] * C.v[F[k]][q] : C.v[F[k]][q];
    if (plane4) { for (int q = 0; q < 3; q++) plane4[q] = C.fn[f][q]; plane4[3] = C.fd[f]; }
    if (other) *other = C.fother[f];
    return (int)F.size();
}
int sf_hull_edges(int h, int32_t* pairs, float* dirs) {
    const sf::HullLib& L = sf::hull_lib();
    if (h < 0 || h >= (int)L.cells.size()) { sf::set_error(SF_E_ARG, "no such hull"); return SF_E_ARG; }
    const sf::HullCell& C = L.cells[h];
    if (pairs) for (size_t e = 0; e < C.seg.size(); e++) { pairs[2 * e] = C.seg[e][0]; pairs[2 * e + 1] = C.seg[e][1]; }
    if (dirs) for (size_t d = 0; d < C.dir.size(); d++) for (int q = 0; q < 3; q++) dirs[3 * d + q] = (float)C.dir[d][q];
    return (int)C.seg.size();
}
// the device tables in one call (for a checker that runs the same narrow phase): sizes in hulls first
int sf_hull_tables(float* v4, float* pl4, float* dir4, uint8_t* fidx, uint8_t* fcnt, uint8_t* seg, int32_t* head) {
    const sf::HullTables& T = sf::hull_tables();
    if (v4) std::memcpy(v4, T.v4.data(), sizeof(float) * T.v4.size());
    if (pl4) std::memcpy(pl4, T.pl4.data(), sizeof(float) * T.pl4.size());
    if (dir4) std::memcpy(dir4, T.dir4.data(), sizeof(float) * T.dir4.size());
    if (fidx) std::memcpy(fidx, T.fidx.data(), T.fidx.size());
    if (fcnt) std::memcpy(fcnt, T.fcnt.data(), T.fcnt.size());
    if (seg) std::memcpy(seg, T.seg.data(), T.seg.size());
    if (head) std::memcpy(head, T.head.data(), sizeof(int32_t) * T.head.size());
    return T.count;
}
}
