// Spring-network structural analysis on the device (engine/src/structure.rs:191-783 over the brick store of
// engine/src/brick.rs).
//
// The reference analyses one building on the CPU: each brick is cut into cells (rigid bodies with 6 unknowns), joints
// and in-brick sections become springs, and a conjugate-gradient solve preconditioned by a multigrid V-cycle over
// rigid blocks gives every face's force and moment, which are then checked against the materials' strengths.
//
// Here the host builds the cell/link graph and the block hierarchy of each queued case (C++ below, several cases on
// worker threads), packs them into one arena, and the whole solve of a case — every conjugate-gradient iteration with
// its V-cycle over all levels — runs in ONE workgroup of 1,024 threads (k_net_solve): a batch of cases is one launch
// with a workgroup per case. The level vectors stay in L2/HBM, every sum a cell receives is gathered in the order the
// reference adds it (springs in link order, blocks in coupling order, children in cell order), and dot products are
// 1,024 strided partial sums then a halving tree. The face checks run afterwards as one thread per joint or inner face
// (k_net_checks). f64 throughout, as the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>
#include <map>

#include "../../include/sf.h"
#include "sf_host.h"

namespace {

constexpr uint32_t NONE = 0xffffffffu;
constexpr double TK = 0.125;
constexpr float TKF = 0.125f;
constexpr int CELL_TICKS = 16;
constexpr int MAX_LEVELS = 8;
constexpr int MAX_IT = 2000;
constexpr double TOL = 1e-5;
constexpr double OMEGA = 0.8;
#ifndef SF_NET_THREADS
#define SF_NET_THREADS 1024
#endif
constexpr int NT_SOLVE = SF_NET_THREADS;  // threads of the one-workgroup solve (1024 or 512)
constexpr int NPART = 1024;               // a dot product's strided partial sums (the order the results depend on)
static_assert(NT_SOLVE == 1024 || NT_SOLVE == 512, "the dot's partial sums assume 1,024 or 512 threads");
constexpr int COARSE_SWEEPS = 30;
enum : int { CONCRETE = 0, MASONRY = 1, STEEL = 2, GLASS = 3, TIMBER = 4 };
enum : int { F_ALIVE = 1, F_INFILL = 2, F_DYNAMIC = 4, F_MONO = 8 };

// ---------------------------------------------------------------- materials (structure.rs:20-113)
float mat_density(int m) { return m == CONCRETE ? 2400.0f : m == MASONRY ? 1900.0f : m == STEEL ? 7850.0f : m == GLASS ? 2500.0f : 500.0f; }
void mat_strength(int m, float* s) {
    static const float STEEL_AREA = 0.16f, STEEL_SHEAR = 0.10f;
    switch (m) {
    case CONCRETE: s[0] = 30e6f; s[1] = 4e6f; s[2] = 4e6f; break;
    case MASONRY: s[0] = 10e6f; s[1] = 0.3e6f; s[2] = 0.5e6f; break;
    case STEEL: s[0] = 250e6f * STEEL_AREA; s[1] = 250e6f * STEEL_AREA; s[2] = 150e6f * STEEL_SHEAR; break;
    case GLASS: s[0] = 50e6f; s[1] = 7e6f; s[2] = 7e6f; break;
    default: s[0] = 21e6f; s[1] = 14e6f; s[2] = 4e6f; break;
    }
}
float mat_flexural(int m) { return m == CONCRETE ? 12e6f : m == MASONRY ? 0.5e6f : m == STEEL ? 250e6f * 0.12f : m == GLASS ? 40e6f : 20e6f; }

// ---------------------------------------------------------------- the brick store
struct Store {
    std::vector<int32_t> lo;                 // min corner, 3 per brick (ticks)
    std::vector<uint8_t> ext;                // size, 3 per brick (1..255 ticks)
    std::vector<uint8_t> mat, paint, flags;
    std::vector<uint16_t> sid;
    std::vector<uint32_t> ja, jb;            // joint ends (NONE: removed), ja on the min side
    std::vector<uint16_t> jarea;
    std::vector<uint8_t> jaxis, jdmg;
    std::vector<std::vector<uint32_t>> inc;  // per brick: its live joints, the newest first
    int live = 0;
    // carried for the store's save format (brick.rs:116-145); the solve does not read them
    std::vector<uint32_t> body;              // owning body (NONE: none)
    std::vector<uint8_t> damage, hull;       // damage byte, convex cell (0: the whole box)
    std::vector<uint32_t> jfree;             // removed joint slots, in the order they were freed
    size_t size() const { return mat.size(); }
    int32_t mn(size_t i, int k) const { return lo[3 * i + k]; }
    int32_t mx(size_t i, int k) const { return lo[3 * i + k] + ext[3 * i + k]; }
    bool timber_floor(size_t i) const { return mat[i] == TIMBER && paint[i] == 8; }
    float flexural_of(size_t i) const { return timber_floor(i) ? 2.4e6f : mat_flexural(mat[i]); }
    double modulus(size_t i) const {
        switch (mat[i]) {
        case CONCRETE: return 30e9;
        case MASONRY: return 5e9;
        case STEEL: return 200e9 * 0.17;
        case GLASS: return 70e9;
        default: return paint[i] == 8 ? 1e9 : 11e9;
        }
    }
    bool hinge(size_t x, size_t y) const {
        int a = mat[x], b = mat[y];
        return ((timber_floor(x) || timber_floor(y)) && a != b) || (a == STEEL && b == MASONRY) || (a == MASONRY && b == STEEL);
    }
    // joint strength factors (structure.rs:54-64)
    void connection(uint32_t j, float* c) const {
        int ma = mat[ja[j]], mb = mat[jb[j]], pa = paint[ja[j]], pb = paint[jb[j]];
        c[0] = 1.0f; c[1] = 0.7f; c[2] = 0.7f;
        if (pa == pb && ma == mb) { c[1] = 1.0f; c[2] = 1.0f; }
        else if (ma == TIMBER && mb == TIMBER) { c[1] = 0.1f; c[2] = 0.25f; }
        else if (ma == CONCRETE && mb == CONCRETE) {
            bool slab_a = pa == 0, slab_b = pb == 0, sup_a = pa == 1 || pa == 3, sup_b = pb == 1 || pb == 3;
            if ((slab_a && sup_b) || (slab_b && sup_a)) { c[1] = 0.5f; c[2] = 0.45f; }
        }
    }
};

// the spatial-index rank a brick has in the reference's neighbour query (brick.rs:171-180, :259-281): the level whose
// cell is the first at least as large as the brick's largest side, then its cell coordinates, newest first
struct QRank { int level; int32_t cx, cy, cz; uint32_t idx; };
QRank query_rank(const Store& S, uint32_t i) {
    int32_t big = 0;
    for (int k = 0; k < 3; k++) big = std::max<int32_t>(big, S.ext[3 * i + k]);
    int l = 3;
    for (int q = 0; q < 4; q++) if ((8 << (2 * q)) >= big) { l = q; break; }
    int32_t c = 8 << (2 * l);
    auto fl = [c](int32_t a) { int32_t q = a / c; return (a % c < 0) ? q - 1 : q; };
    return {l, fl(S.mn(i, 0)), fl(S.mn(i, 1)), fl(S.mn(i, 2)), i};
}
bool rank_less(const QRank& a, const QRank& b) {
    if (a.level != b.level) return a.level < b.level;
    if (a.cx != b.cx) return a.cx < b.cx;
    if (a.cy != b.cy) return a.cy < b.cy;
    if (a.cz != b.cz) return a.cz < b.cz;
    return a.idx > b.idx;
}
// shared face of two boxes (brick.rs:53-66): the axis, the area in tick^2 and whether p is on the min side
bool shared_face(const Store& S, uint32_t p, uint32_t q, int* ax, uint32_t* area, bool* p_low) {
    for (int a = 0; a < 3; a++) {
        bool pl = S.mx(p, a) == S.mn(q, a), ql = S.mx(q, a) == S.mn(p, a);
        if (!pl && !ql) continue;
        int o1 = (a + 1) % 3, o2 = (a + 2) % 3;
        int32_t e1 = std::min(S.mx(p, o1), S.mx(q, o1)) - std::max(S.mn(p, o1), S.mn(q, o1));
        int32_t e2 = std::min(S.mx(p, o2), S.mx(q, o2)) - std::max(S.mn(p, o2), S.mn(q, o2));
        if (e1 > 0 && e2 > 0) { *ax = a; *area = (uint32_t)(e1 * e2); *p_low = pl; return true; }
        return false;
    }
    return false;
}
// Every joint from geometry (brick.rs:447-458). Touching pairs come from a sweep over the bricks sorted by their min x;
// the joints are then numbered in the order the reference creates them: by the lower brick, then by the other
// brick's rank in the neighbour query.
int build_joints(Store& S) {
    size_t n = S.size();
    S.ja.clear(); S.jb.clear(); S.jarea.clear(); S.jaxis.clear(); S.jdmg.clear(); S.jfree.clear();
    S.inc.assign(n, {});
    S.live = 0;
    std::vector<uint32_t> ord;
    for (uint32_t i = 0; i < n; i++) if (S.flags[i] & F_ALIVE) ord.push_back(i);
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return S.mn(a, 0) != S.mn(b, 0) ? S.mn(a, 0) < S.mn(b, 0) : a < b; });
    struct J { uint32_t lo, hi; QRank r; };
    std::vector<J> js;
    for (size_t s = 0; s < ord.size(); s++) {
        uint32_t a = ord[s];
        for (size_t t = s + 1; t < ord.size() && S.mn(ord[t], 0) <= S.mx(a, 0); t++) {
            uint32_t b = ord[t];
            if (S.mn(b, 1) > S.mx(a, 1) || S.mx(b, 1) < S.mn(a, 1) || S.mn(b, 2) > S.mx(a, 2) || S.mx(b, 2) < S.mn(a, 2)) continue;
            uint32_t i = std::min(a, b), o = std::max(a, b);
            if (S.flags[o] & F_DYNAMIC) continue;  // the query returns static bricks only
            int ax; uint32_t area; bool pl;
            if (!shared_face(S, i, o, &ax, &area, &pl)) continue;
            js.push_back({i, o, query_rank(S, o)});
        }
    }
    std::sort(js.begin(), js.end(), [](const J& x, const J& y) { return x.lo != y.lo ? x.lo < y.lo : rank_less(x.r, y.r); });
    for (const J& q : js) {
        int ax; uint32_t area; bool pl;
        shared_face(S, q.lo, q.hi, &ax, &area, &pl);
        uint32_t a = pl ? q.lo : q.hi, b = pl ? q.hi : q.lo;
        uint32_t j = (uint32_t)S.ja.size();
        S.ja.push_back(a); S.jb.push_back(b); S.jarea.push_back((uint16_t)area); S.jaxis.push_back((uint8_t)ax); S.jdmg.push_back(0);
        S.inc[a].push_back(j); S.inc[b].push_back(j);
        S.live++;
    }
    for (auto& v : S.inc) std::reverse(v.begin(), v.end());
    return S.live;
}

// ---------------------------------------------------------------- 6 x 6 blocks (row-major)
// a product that skips zero entries of the left factor (structure.rs:622), so -0 and the order of sums match
void mm6(const double* a, const double* b, double* c) {
    for (int i = 0; i < 36; i++) c[i] = 0.0;
    for (int i = 0; i < 6; i++)
        for (int k = 0; k < 6; k++) {
            double v = a[6 * i + k];
            if (v == 0.0) continue;
            for (int j = 0; j < 6; j++) c[6 * i + j] += v * b[6 * k + j];
        }
}
void tr6(const double* a, double* c) { for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) c[6 * j + i] = a[6 * i + j]; }
// rigid map from a cell's (u, theta) to a point at offset d (structure.rs:624-630)
void rigid6(const double* d, double* t) {
    for (int i = 0; i < 36; i++) t[i] = (i % 7 == 0) ? 1.0 : 0.0;
    t[4] = d[2]; t[5] = -d[1]; t[9] = -d[2]; t[11] = d[0]; t[15] = d[1]; t[16] = -d[0];
}
// Products with a rigid map t = rigid6(d), whose only non-zero entries are its unit diagonal and the six of [d]x. Each
// sum keeps mm6's order of k and its start from +0.0; a term with a zero factor adds +-0 and never changes such a sum,
// so leaving those terms out gives mm6's result bit for bit (and mm6's own skips of zero left factors are such terms).
// RK[j]: the rows k of t with a non-zero in column j, in increasing order.
const int RN[6] = {1, 1, 1, 3, 3, 3};
const int RK[6][3] = {{0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {1, 2, 3}, {0, 2, 4}, {0, 1, 5}};
// c = m t
void mul_rigid(const double* m, const double* t, double* c) {
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int r = 0; r < RN[j]; r++) { int k = RK[j][r]; s += m[6 * i + k] * t[6 * k + j]; }
            c[6 * i + j] = s;
        }
}
// c = t' m
void rigid_t_mul(const double* t, const double* m, double* c) {
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double s = 0.0;
            for (int r = 0; r < RN[i]; r++) { int k = RK[i][r]; s += t[6 * k + i] * m[6 * k + j]; }
            c[6 * i + j] = s;
        }
}
// c = diag(d6) t
void diag_rigid(const double* d6, const double* t, double* c) {
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) c[6 * i + j] = 0.0 + d6[i] * t[6 * i + j];
}
// Gauss-Jordan with partial pivoting (structure.rs:765-779)
void inv6(const double* a, double* out) {
    double m[36];
    std::memcpy(m, a, sizeof(m));
    for (int i = 0; i < 36; i++) out[i] = (i % 7 == 0) ? 1.0 : 0.0;
    for (int c = 0; c < 6; c++) {
        int p = c;
        for (int r = c + 1; r < 6; r++) if (std::fabs(m[6 * r + c]) > std::fabs(m[6 * p + c])) p = r;
        if (std::fabs(m[6 * p + c]) < 1e-300) continue;
        if (p != c) for (int j = 0; j < 6; j++) { std::swap(m[6 * c + j], m[6 * p + j]); std::swap(out[6 * c + j], out[6 * p + j]); }
        double d = m[6 * c + c];
        for (int j = 0; j < 6; j++) { m[6 * c + j] /= d; out[6 * c + j] /= d; }
        for (int r = 0; r < 6; r++) {
            if (r == c) continue;
            double q = m[6 * r + c];
            if (q == 0.0) continue;
            for (int j = 0; j < 6; j++) { m[6 * r + j] -= q * m[6 * c + j]; out[6 * r + j] -= q * out[6 * c + j]; }
        }
    }
}

struct Sect { double area, iu, iw, zu, zw, j, rmax; };
Sect rect_section(double eu, double ew, double area) {
    eu = std::fmax(eu, 1e-3); ew = std::fmax(ew, 1e-3);
    return {std::fmax(area, 1e-6), eu * ew * ew * ew / 12.0, ew * eu * eu * eu / 12.0, ew / 2.0, eu / 2.0,
            eu * ew * (eu * eu + ew * ew) / 12.0, 0.5 * std::sqrt(eu * eu + ew * ew)};
}

// ---------------------------------------------------------------- a case on the host
struct Level {
    int n = 0;
    std::vector<double> diag, inv, reg, pos;   // 36n, 36n, n, 3n
    std::vector<uint8_t> fixed;
    std::vector<uint32_t> oa, oc;              // couplings between cells (a < c on coarse levels)
    std::vector<double> om;                    // 36 per coupling
    std::vector<uint32_t> up;                  // (coarse levels) the finer level's cell -> this level's
};
struct Check {                                 // one face to check on the device
    int32_t kind;                              // 0 joint (sub-links summed about its centre), 1 face inside a brick
    int32_t axis, hinge, l0, l1;               // l0..l1: the joint's links (in the case's link-group list) or the link
    float cap[3], flex;
    double c[3];
    Sect sec;
};
struct Case {
    // inputs
    std::vector<uint32_t> set;
    std::vector<float> load;
    std::vector<uint8_t> sink;
    float dir[3] = {0.0f, -1.0f, 0.0f};
    std::vector<float> points;                 // 4 per point
    sf_warm* warm = nullptr;
    // cells
    std::vector<uint32_t> cbrick;
    std::vector<int32_t> clo, chi;             // 3 per cell
    std::vector<float> cload;
    std::vector<uint8_t> csink, live;
    // links
    std::vector<uint32_t> la, lb, lj;
    std::vector<int32_t> lax, flo, fhi;
    std::vector<double> spring;                // 15 per link: k[3] kr[3] ra[3] rb[3] c[3]
    std::vector<double> f, x0;                 // 6 per cell
    bool has_x0 = false;
    std::vector<Level> lv;
    std::vector<Check> checks;
    std::vector<int32_t> glinks;               // the joints' links, joint after joint
    int njoint = 0;                            // checks[0..njoint) are joints, the rest inner faces
    // results
    std::vector<double> x;
    int iterations = 0;
    std::vector<float> ratio, flow, moment;    // per check
    float max_ratio = 0.0f;
    std::string error;
};

uint64_t warm_key(uint32_t brick, int32_t x, int32_t z) {
    return ((uint64_t)brick << 32) | ((uint64_t)(uint32_t)(x + 32768) & 0xffffull) << 16 | ((uint64_t)(uint32_t)(z + 32768) & 0xffffull);
}

}  // namespace

struct sf_warm { std::unordered_map<uint64_t, std::array<double, 6>> m; };
struct sf_bricks { Store s; };

namespace {

// The cells of the case's bricks, the links between them, their springs (structure.rs:191-318, :465-524).
void build_graph(const Store& S, Case& C) {
    size_t ns = C.set.size();
    std::vector<uint32_t> first(S.size(), NONE), slot(S.size(), NONE);
    std::vector<std::vector<int32_t>> xs(ns), zs(ns);
    for (size_t k = 0; k < ns; k++) slot[C.set[k]] = (uint32_t)k;
    auto cuts = [](int32_t a, int32_t b, std::vector<int32_t>& v) {
        int32_t n = std::max<int32_t>((b - a + CELL_TICKS - 1) / CELL_TICKS, 1);
        v.clear();
        v.push_back(a);
        for (int32_t q = 1; q < n; q++) v.push_back(a + (b - a) * q / n);
        v.push_back(b);
    };
    for (size_t k = 0; k < ns; k++) {
        uint32_t i = C.set[k];
        std::vector<int32_t>&X = xs[k], &Z = zs[k];
        if (S.flags[i] & F_INFILL) { X = {S.mn(i, 0), S.mx(i, 0)}; Z = {S.mn(i, 2), S.mx(i, 2)}; }
        else {
            cuts(S.mn(i, 0), S.mx(i, 0), X);
            cuts(S.mn(i, 2), S.mx(i, 2), Z);
            // the edges of what it rests on: the moment peaks there
            for (uint32_t j : S.inc[i]) {
                uint32_t o = S.ja[j] == i ? S.jb[j] : S.ja[j];
                if (S.jaxis[j] != 1 || S.mx(o, 1) != S.mn(i, 1) || (S.flags[o] & F_INFILL)) continue;
                for (int32_t v : {S.mn(o, 0), S.mx(o, 0)}) if (v > S.mn(i, 0) && v < S.mx(i, 0)) X.push_back(v);
                for (int32_t v : {S.mn(o, 2), S.mx(o, 2)}) if (v > S.mn(i, 2) && v < S.mx(i, 2)) Z.push_back(v);
            }
            std::sort(X.begin(), X.end()); X.erase(std::unique(X.begin(), X.end()), X.end());
            std::sort(Z.begin(), Z.end()); Z.erase(std::unique(Z.begin(), Z.end()), Z.end());
        }
        first[i] = (uint32_t)C.cbrick.size();
        int nx = (int)X.size() - 1, nz = (int)Z.size() - 1;
        float area = (float)((S.mx(i, 0) - S.mn(i, 0)) * (S.mx(i, 2) - S.mn(i, 2)));
        for (int cz = 0; cz < nz; cz++)
            for (int cx = 0; cx < nx; cx++) {
                int32_t lo[3] = {X[cx], S.mn(i, 1), Z[cz]}, hi[3] = {X[cx + 1], S.mx(i, 1), Z[cz + 1]};
                C.cbrick.push_back(i);
                for (int d = 0; d < 3; d++) { C.clo.push_back(lo[d]); C.chi.push_back(hi[d]); }
                C.cload.push_back(C.load[k] * ((float)((hi[0] - lo[0]) * (hi[2] - lo[2])) / area));
                C.csink.push_back(C.sink[k] ? 1 : 0);
            }
        auto link = [&](uint32_t a, uint32_t b, uint32_t j, int ax, const int32_t* lo, const int32_t* hi) {
            C.la.push_back(a); C.lb.push_back(b); C.lj.push_back(j); C.lax.push_back(ax);
            for (int d = 0; d < 3; d++) { C.flo.push_back(lo[d]); C.fhi.push_back(hi[d]); }
        };
        for (int cz = 0; cz < nz; cz++)
            for (int cx = 0; cx < nx; cx++) {
                uint32_t id = first[i] + (uint32_t)(cz * nx + cx);
                const int32_t *lo = &C.clo[3 * id], *hi = &C.chi[3 * id];
                if (cx + 1 < nx) { int32_t a[3] = {hi[0], lo[1], lo[2]}, b[3] = {hi[0], hi[1], hi[2]}; link(id, id + 1, NONE, 0, a, b); }
                if (cz + 1 < nz) { int32_t a[3] = {lo[0], lo[1], hi[2]}, b[3] = {hi[0], hi[1], hi[2]}; link(id, id + (uint32_t)nx, NONE, 2, a, b); }
            }
    }
    // across joints, cell to cell where the face overlaps (the cells of the far brick found by its cut lists)
    for (size_t k = 0; k < ns; k++) {
        uint32_t i = C.set[k];
        for (uint32_t j : S.inc[i]) {
            uint32_t o = S.ja[j] == i ? S.jb[j] : S.ja[j];
            if (first[o] == NONE || S.ja[j] != i) continue;
            int ax = S.jaxis[j];
            size_t ko = slot[o];
            const std::vector<int32_t>&OX = xs[ko], &OZ = zs[ko];
            int nax = (int)xs[k].size() - 1, naz = (int)zs[k].size() - 1, nbx = (int)OX.size() - 1, nbz = (int)OZ.size() - 1;
            int32_t plane = S.mx(i, ax);
            auto span = [](const std::vector<int32_t>& c, int32_t lo, int32_t hi, int* a, int* z) {
                int n = (int)c.size() - 1;
                int p = (int)(std::upper_bound(c.begin(), c.end(), lo) - c.begin());
                int q = (int)(std::lower_bound(c.begin(), c.end(), hi) - c.begin());
                *a = std::min(std::max(p - 1, 0), n);
                *z = std::max(std::min(q, n), *a);
            };
            for (int ca = 0; ca < nax * naz; ca++) {
                uint32_t ia = first[i] + (uint32_t)ca;
                const int32_t *plo = &C.clo[3 * ia], *phi = &C.chi[3 * ia];
                if (ax != 1 && phi[ax] != plane) continue;
                int x0 = 0, x1 = nbx, z0 = 0, z1 = nbz;
                if (ax != 0) span(OX, plo[0], phi[0], &x0, &x1);
                if (ax != 2) span(OZ, plo[2], phi[2], &z0, &z1);
                for (int cz = z0; cz < z1; cz++)
                    for (int cx = x0; cx < x1; cx++) {
                        uint32_t ib = first[o] + (uint32_t)(cz * nbx + cx);
                        const int32_t *qlo = &C.clo[3 * ib], *qhi = &C.chi[3 * ib];
                        int32_t lo[3], hi[3];
                        bool ok = true;
                        for (int d = 0; d < 3; d++) {
                            if (d == ax) { lo[d] = plane; hi[d] = plane; continue; }
                            lo[d] = std::max(plo[d], qlo[d]); hi[d] = std::min(phi[d], qhi[d]);
                            if (hi[d] <= lo[d]) ok = false;
                        }
                        if (!ok) continue;
                        C.la.push_back(ia); C.lb.push_back(ib); C.lj.push_back(j); C.lax.push_back(ax);
                        for (int d = 0; d < 3; d++) { C.flo.push_back(lo[d]); C.fhi.push_back(hi[d]); }
                    }
            }
        }
    }
    // point loads: (index into the set, x, z in ticks, N) onto the cell under the point
    for (size_t q = 0; q + 4 <= C.points.size(); q += 4) {
        int k = (int)C.points[q];
        if (k < 0 || (size_t)k >= ns) continue;
        uint32_t i = C.set[k];
        if (first[i] == NONE) continue;
        const std::vector<int32_t>&X = xs[k], &Z = zs[k];
        int nx = (int)X.size() - 1, nz = (int)Z.size() - 1;
        float px = C.points[q + 1], pz = C.points[q + 2];
        int cx = 0, cz = 0;
        while (cx < (int)X.size() && (float)X[cx] <= px) cx++;
        while (cz < (int)Z.size() && (float)Z[cz] <= pz) cz++;
        cx = std::min(std::max(cx - 1, 0), nx - 1);
        cz = std::min(std::max(cz - 1, 0), nz - 1);
        C.cload[first[i] + (uint32_t)(cz * nx + cx)] += C.points[q + 3];
    }
    size_t n = C.cbrick.size(), nl = C.la.size();
    // cells reached from a support through links
    C.live.assign(n, 0);
    {
        std::vector<uint32_t> off(n + 1, 0), adj(2 * nl);
        for (size_t l = 0; l < nl; l++) { off[C.la[l] + 1]++; off[C.lb[l] + 1]++; }
        for (size_t c = 0; c < n; c++) off[c + 1] += off[c];
        std::vector<uint32_t> fill(off.begin(), off.end() - 1);
        for (size_t l = 0; l < nl; l++) { adj[fill[C.la[l]]++] = C.lb[l]; adj[fill[C.lb[l]]++] = C.la[l]; }
        std::vector<uint32_t> q;
        for (size_t c = 0; c < n; c++) if (C.csink[c]) { C.live[c] = 1; q.push_back((uint32_t)c); }
        for (size_t h = 0; h < q.size(); h++)
            for (uint32_t e = off[q[h]]; e < off[q[h] + 1]; e++) if (!C.live[adj[e]]) { C.live[adj[e]] = 1; q.push_back(adj[e]); }
    }
    // springs
    C.spring.assign(15 * nl, 0.0);
    for (size_t l = 0; l < nl; l++) {
        int ax = C.lax[l], u = (ax + 1) % 3, w = (ax + 2) % 3;
        const int32_t *lo = &C.flo[3 * l], *hi = &C.fhi[3 * l];
        double eu = (double)(hi[u] - lo[u]) * TK, ew = (double)(hi[w] - lo[w]) * TK;
        double c[3], pa[3], pb[3];
        for (int d = 0; d < 3; d++) {
            c[d] = 0.5 * (double)(lo[d] + hi[d]) * TK;
            pa[d] = 0.5 * (double)(C.clo[3 * C.la[l] + d] + C.chi[3 * C.la[l] + d]) * TK;
            pb[d] = 0.5 * (double)(C.clo[3 * C.lb[l] + d] + C.chi[3 * C.lb[l] + d]) * TK;
        }
        Sect sec = rect_section(eu, ew, eu * ew);
        double nrm[3] = {0.0, 0.0, 0.0};
        nrm[ax] = 1.0;
        double da = std::fmax(std::fabs((c[0] - pa[0]) * nrm[0] + (c[1] - pa[1]) * nrm[1] + (c[2] - pa[2]) * nrm[2]), 0.5 * TK);
        double db = std::fmax(std::fabs((c[0] - pb[0]) * nrm[0] + (c[1] - pb[1]) * nrm[1] + (c[2] - pb[2]) * nrm[2]), 0.5 * TK);
        uint32_t ia = C.cbrick[C.la[l]], ib = C.cbrick[C.lb[l]];
        double comp = da / S.modulus(ia) + db / S.modulus(ib);
        double gcomp = comp / 0.4;
        double* s = &C.spring[15 * l];
        s[ax] = sec.area / comp; s[u] = sec.area / gcomp; s[w] = sec.area / gcomp;
        s[3 + ax] = sec.j / gcomp; s[3 + u] = sec.iu / comp; s[3 + w] = sec.iw / comp;
        uint32_t j = C.lj[l];
        int fa = S.flags[ia] | S.flags[ib];
        int ma = S.mat[ia], mb = S.mat[ib];
        bool pocket = j != NONE && ax != 1 && ((ma == STEEL && mb == MASONRY) || (ma == MASONRY && mb == STEEL));
        double fac = ((fa & F_INFILL) || pocket) ? 0.02 : ((fa & F_MONO) ? 0.05 : 1.0);
        if (j != NONE) { int dm = S.jdmg[j]; fac *= ((dm & 1) ? 0.5 : 1.0) * (1.0 - 0.6 * (double)(dm >> 1) / 127.0); }
        for (int q = 0; q < 6; q++) s[q] *= fac;
        if (j != NONE && S.hinge(ia, ib)) { s[3 + u] = 0.0; s[3 + w] = 0.0; }
        if (S.timber_floor(ia) && S.timber_floor(ib) && ax == 2) { s[3] *= 0.05; s[5] *= 0.05; }
        if (S.timber_floor(ia) && S.timber_floor(ib) && ax == 0) { s[3] *= 0.05; }
        for (int d = 0; d < 3; d++) { s[6 + d] = c[d] - pa[d]; s[9 + d] = c[d] - pb[d]; s[12 + d] = c[d]; }
    }
}

void finish_level(Level& L) {
    L.inv.assign(36 * (size_t)L.n, 0.0);
    L.reg.assign(L.n, 0.0);
    for (int k = 0; k < L.n; k++) {
        if (L.fixed[k]) continue;
        double* d = &L.diag[36 * (size_t)k];
        double big = 0.0;
        for (int i = 0; i < 6; i++) big = std::fmax(big, std::fabs(d[7 * i]));
        big = std::fmax(big, 1.0);
        L.reg[k] = 1e-7 * big;
        for (int i = 0; i < 6; i++) d[7 * i] += L.reg[k];
        inv6(d, &L.inv[36 * (size_t)k]);
    }
}
// the finest level from the springs: K_aa = B_a' D B_a, K_bb = B_b' D B_b, K_ab = -B_a' D B_b (structure.rs:592-616)
void fine_level(Case& C, Level& L) {
    size_t n = C.cbrick.size(), nl = C.la.size();
    L.n = (int)n;
    L.diag.assign(36 * n, 0.0);
    L.fixed.resize(n);
    L.pos.resize(3 * n);
    for (size_t k = 0; k < n; k++) {
        L.fixed[k] = C.csink[k] || !C.live[k];
        for (int d = 0; d < 3; d++) L.pos[3 * k + d] = 0.5 * (double)(C.clo[3 * k + d] + C.chi[3 * k + d]) * TK;
    }
    for (size_t l = 0; l < nl; l++) {
        const double* s = &C.spring[15 * l];
        double ba[36], bb[36], dba[36], dbb[36], m[36];
        rigid6(s + 6, ba); rigid6(s + 9, bb);
        diag_rigid(s, ba, dba); diag_rigid(s, bb, dbb);   // s[0..5]: the spring's k and kr, its diagonal D
        uint32_t a = C.la[l], b = C.lb[l];
        if (!L.fixed[a]) { rigid_t_mul(ba, dba, m); for (int q = 0; q < 36; q++) L.diag[36 * (size_t)a + q] += m[q]; }
        if (!L.fixed[b]) { rigid_t_mul(bb, dbb, m); for (int q = 0; q < 36; q++) L.diag[36 * (size_t)b + q] += m[q]; }
        if (!L.fixed[a] && !L.fixed[b]) {
            rigid_t_mul(ba, dbb, m);
            L.oa.push_back(a); L.oc.push_back(b);
            for (int q = 0; q < 36; q++) L.om.push_back(-m[q]);
        }
    }
    finish_level(L);
}
// Cells grouped into blocks of `bin` metres, each block rigid: K_c = T' K T (structure.rs:686-722). The couplings
// between blocks are listed in the order they first appear.
void coarse_level(const Level& F, double bin, Level& L) {
    std::unordered_map<uint64_t, uint32_t> blocks;
    L.up.assign(F.n, NONE);
    std::vector<double> cnt;
    for (int k = 0; k < F.n; k++) {
        if (F.fixed[k]) continue;
        const double* p = &F.pos[3 * (size_t)k];
        int64_t b[3] = {(int64_t)std::floor(p[0] / bin), (int64_t)std::floor(p[1] / bin), (int64_t)std::floor(p[2] / bin)};
        uint64_t key = ((uint64_t)(b[0] + (1 << 20)) & 0x1fffffull) << 42 | ((uint64_t)(b[1] + (1 << 20)) & 0x1fffffull) << 21 | ((uint64_t)(b[2] + (1 << 20)) & 0x1fffffull);
        auto it = blocks.emplace(key, (uint32_t)cnt.size());
        if (it.second) { cnt.push_back(0.0); for (int d = 0; d < 3; d++) L.pos.push_back(0.0); }
        uint32_t c = it.first->second;
        L.up[k] = c;
        for (int d = 0; d < 3; d++) L.pos[3 * c + d] += p[d];
        cnt[c] += 1.0;
    }
    L.n = (int)cnt.size();
    for (int c = 0; c < L.n; c++) for (int d = 0; d < 3; d++) L.pos[3 * c + d] /= cnt[c];
    std::vector<double> T(36 * (size_t)F.n, 0.0);
    for (int k = 0; k < F.n; k++) {
        if (L.up[k] == NONE) continue;
        uint32_t c = L.up[k];
        double d[3] = {F.pos[3 * k] - L.pos[3 * c], F.pos[3 * k + 1] - L.pos[3 * c + 1], F.pos[3 * k + 2] - L.pos[3 * c + 2]};
        rigid6(d, &T[36 * (size_t)k]);
    }
    L.diag.assign(36 * (size_t)std::max(L.n, 1), 0.0);
    for (int k = 0; k < F.n; k++) {
        if (L.up[k] == NONE) continue;
        double a[36], m[36];
        mul_rigid(&F.diag[36 * (size_t)k], &T[36 * (size_t)k], a);
        rigid_t_mul(&T[36 * (size_t)k], a, m);
        for (int q = 0; q < 36; q++) L.diag[36 * (size_t)L.up[k] + q] += m[q];
    }
    std::unordered_map<uint64_t, uint32_t> pairs;
    for (size_t o = 0; o < F.oa.size(); o++) {
        uint32_t ua = L.up[F.oa[o]], uc = L.up[F.oc[o]];
        if (ua == NONE || uc == NONE) continue;
        double a[36], q[36], qt[36];
        mul_rigid(&F.om[36 * o], &T[36 * (size_t)F.oc[o]], a);
        rigid_t_mul(&T[36 * (size_t)F.oa[o]], a, q);
        tr6(q, qt);
        if (ua == uc) {
            double* d = &L.diag[36 * (size_t)ua];
            for (int z = 0; z < 36; z++) d[z] += q[z];
            for (int z = 0; z < 36; z++) d[z] += qt[z];
            continue;
        }
        uint32_t lo = std::min(ua, uc), hi = std::max(ua, uc);
        auto it = pairs.emplace((uint64_t)lo << 32 | hi, (uint32_t)L.oa.size());
        if (it.second) { L.oa.push_back(lo); L.oc.push_back(hi); L.om.resize(L.om.size() + 36, 0.0); }
        double* acc = &L.om[36 * (size_t)it.first->second];
        const double* add = ua < uc ? q : qt;
        for (int z = 0; z < 36; z++) acc[z] += add[z];
    }
    L.fixed.assign(std::max(L.n, 1), 0);
    finish_level(L);
}
void build_levels(Case& C) {
    C.lv.clear();
    C.lv.emplace_back();
    fine_level(C, C.lv[0]);
    double bin = 3.0;
    while (C.lv.back().n > 64 && (int)C.lv.size() < MAX_LEVELS) {
        Level c;
        coarse_level(C.lv.back(), bin, c);
        if ((double)c.n > 0.7 * (double)C.lv.back().n) {
            bin *= 2.0;
            if (bin > 256.0) break;
            continue;
        }
        C.lv.push_back(std::move(c));
        bin *= 2.0;
    }
}
// loads, warm start and the checks of a case
void build_rest(const Store& S, Case& C) {
    size_t n = C.cbrick.size();
    C.f.assign(6 * n, 0.0);
    for (size_t k = 0; k < n; k++) {
        double w = (double)C.cload[k];
        for (int d = 0; d < 3; d++) C.f[6 * k + d] = w * (double)C.dir[d];
    }
    C.has_x0 = C.warm != nullptr;
    if (C.warm) {
        C.x0.assign(6 * n, 0.0);
        const Level& L = C.lv[0];
        for (size_t k = 0; k < n; k++) {
            if (L.fixed[k]) continue;
            auto it = C.warm->m.find(warm_key(C.cbrick[k], C.clo[3 * k], C.clo[3 * k + 2]));
            if (it != C.warm->m.end()) for (int q = 0; q < 6; q++) C.x0[6 * k + q] = it->second[q];
        }
    }
    // joints: their live links grouped in the order each joint first appears
    size_t nl = C.la.size();
    std::unordered_map<uint32_t, uint32_t> gslot;
    std::vector<std::vector<int32_t>> groups;
    std::vector<uint32_t> gjoint;
    for (size_t l = 0; l < nl; l++) {
        if (C.lj[l] == NONE || !C.live[C.la[l]] || !C.live[C.lb[l]]) continue;
        auto it = gslot.emplace(C.lj[l], (uint32_t)groups.size());
        if (it.second) { groups.emplace_back(); gjoint.push_back(C.lj[l]); }
        groups[it.first->second].push_back((int32_t)l);
    }
    C.checks.clear();
    C.glinks.clear();
    for (size_t g = 0; g < groups.size(); g++) {
        uint32_t j = gjoint[g];
        int ax = S.jaxis[j], u = (ax + 1) % 3, w = (ax + 2) % 3;
        int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
        for (int32_t l : groups[g])
            for (int d = 0; d < 3; d++) { lo[d] = std::min(lo[d], C.flo[3 * l + d]); hi[d] = std::max(hi[d], C.fhi[3 * l + d]); }
        Check K{};
        K.kind = 0; K.axis = ax;
        K.l0 = (int32_t)C.glinks.size();
        C.glinks.insert(C.glinks.end(), groups[g].begin(), groups[g].end());
        K.l1 = (int32_t)C.glinks.size();
        for (int d = 0; d < 3; d++) K.c[d] = 0.5 * (double)(lo[d] + hi[d]) * TK;
        K.sec = rect_section((double)(hi[u] - lo[u]) * TK, (double)(hi[w] - lo[w]) * TK, (double)S.jarea[j] * TK * TK);
        uint32_t ia = S.ja[j], ib = S.jb[j];
        float conn[3], sa[3], sb[3], res[3] = {1.0f, 1.0f, 1.0f};
        S.connection(j, conn);
        if (S.jdmg[j] & 1) {
            if (S.mat[ia] == MASONRY || S.mat[ib] == MASONRY) { res[1] = 0.05f; res[2] = 0.4f; }
            else { res[0] = 0.5f; res[1] = 0.5f; res[2] = 0.5f; }
        }
        mat_strength(S.mat[ia], sa); mat_strength(S.mat[ib], sb);
        for (int q = 0; q < 3; q++) K.cap[q] = std::fmin(sa[q], sb[q]) * conn[q] * res[q];
        K.flex = std::fmin(S.flexural_of(ia), S.flexural_of(ib)) * conn[1] * res[1];
        K.hinge = S.hinge(ia, ib) ? 1 : 0;
        C.checks.push_back(K);
    }
    C.njoint = (int)C.checks.size();
    for (size_t l = 0; l < nl; l++) {
        if (C.lj[l] != NONE || !C.live[C.la[l]]) continue;
        uint32_t i = C.cbrick[C.la[l]];
        int ax = C.lax[l], u = (ax + 1) % 3, w = (ax + 2) % 3;
        double eu = (double)(C.fhi[3 * l + u] - C.flo[3 * l + u]) * TK, ew = (double)(C.fhi[3 * l + w] - C.flo[3 * l + w]) * TK;
        Check K{};
        K.kind = 1; K.axis = ax; K.l0 = (int32_t)l; K.l1 = (int32_t)l + 1;
        K.sec = rect_section(eu, ew, eu * ew);
        mat_strength(S.mat[i], K.cap);
        K.flex = S.flexural_of(i);
        C.checks.push_back(K);
    }
}

// ---------------------------------------------------------------- the device solve
// A level on the device. An apply (y = K x) runs in two phases: every spring (finest level) or coupling (coarse
// levels) writes what it adds to each of its two cells into that cell's run of the contribution buffer (`cpos`: the
// two slots; the runs are in link or coupling order), then every cell sums its own run in order — the additions the
// reference makes, in its order.
struct DLevel {
    int n, nc;              // cells; springs (finest) or couplings (coarse)
    const double* diag;     // (coarse) 36 per cell
    const double* inv;      // 36 per cell
    const double* reg;      // (finest) the diagonal's small addition per cell
    const uint8_t* fixed;   // (finest)
    const int32_t* ioff;    // per cell: its run in the contribution buffer
    const int32_t* cpos;    // per spring or coupling: its slots for end a and end b
    const double* om;       // (coarse) 36 per coupling
    const int32_t* oends;   // (coarse) the couplings' cells
    const int32_t* choff;   // (coarse) per cell: the finer level's cells in it, in order
    const int32_t* ch;
    const int32_t* upf;     // (coarse) per finer cell: its block here (-1: fixed)
    const double* dupf;     // (coarse) per finer cell: its centre minus its block's
    double *r, *z, *kz, *buf;
};
struct DCase {
    int nl, n0;
    int lds;                // bytes of the coarsest level kept in LDS (0: it stays in global memory)
    int route;              // 0: solved by one workgroup (k_net_solve); 1: a kernel per step over the chip
    DLevel lv[MAX_LEVELS];
    const double* sp;       // 15 per link: k[3] kr[3] ra[3] rb[3] face centre[3]
    const int32_t* ends;
    const double* f;
    const double* x0;       // warm start or null
    double *x, *p, *ap;
    int32_t* out_it;
};
struct DCheck {
    int32_t cs, kind, axis, hinge, l0, l1;
    float cap[3], flex;
    double c[3];
    double sec[7];
    const int32_t* glinks;
};

// A spring's 15 numbers (k[3] kr[3] ra[3] rb[3] face centre[3]) are stored field after field over all the case's
// springs (field q of spring s at q * ld + s), and a 6 x 6 block's 36 entries entry after entry over the level's
// blocks, so the lanes of a wavefront read consecutive words.
__device__ inline void spring_fields(const double* sp, size_t ld, int s, double* S) {
    for (int q = 0; q < 12; q++) S[q] = sp[(size_t)q * ld + s];
}
__device__ inline void spring_load(const double* s, const double* xa, const double* xb, double* f, double* m) {
    const double *A = s + 6, *B = s + 9;
    double pa0 = xa[0] + xa[4] * A[2] - xa[5] * A[1], pa1 = xa[1] + xa[5] * A[0] - xa[3] * A[2], pa2 = xa[2] + xa[3] * A[1] - xa[4] * A[0];
    double pb0 = xb[0] + xb[4] * B[2] - xb[5] * B[1], pb1 = xb[1] + xb[5] * B[0] - xb[3] * B[2], pb2 = xb[2] + xb[3] * B[1] - xb[4] * B[0];
    f[0] = s[0] * (pb0 - pa0); f[1] = s[1] * (pb1 - pa1); f[2] = s[2] * (pb2 - pa2);
    m[0] = s[3] * (xb[3] - xa[3]); m[1] = s[4] * (xb[4] - xa[4]); m[2] = s[5] * (xb[5] - xa[5]);
}
// The threads sharing a loop: one workgroup's (the one-workgroup solve) or a whole grid's (the solve of one large case
// spread over the chip, a kernel per step).
struct Span { int first, stride; };
__device__ inline Span wg_span() { return {(int)threadIdx.x, NT_SOLVE}; }
__device__ inline Span grid_span() { return {(int)(blockIdx.x * blockDim.x + threadIdx.x), (int)(gridDim.x * blockDim.x)}; }

// y = K x on the finest level (structure.rs:636-650), first half: each spring's force on end b and its opposite on
// end a, with the moments of the arms, into the two cells' runs
__device__ void fine_pull(const DCase& C, const DLevel& L, const double* x, Span t) {
    for (int s = t.first; s < L.nc; s += t.stride) {
        double S[12], f[3], m[3];
        spring_fields(C.sp, (size_t)L.nc, s, S);
        spring_load(S, x + 6 * (size_t)C.ends[2 * s], x + 6 * (size_t)C.ends[2 * s + 1], f, m);
        const double *ra = S + 6, *rb = S + 9;
        double* eb = L.buf + 6 * (size_t)L.cpos[2 * s + 1];
        double* ea = L.buf + 6 * (size_t)L.cpos[2 * s];
        eb[0] = f[0]; eb[1] = f[1]; eb[2] = f[2];
        eb[3] = m[0] + rb[1] * f[2] - rb[2] * f[1];
        eb[4] = m[1] + rb[2] * f[0] - rb[0] * f[2];
        eb[5] = m[2] + rb[0] * f[1] - rb[1] * f[0];
        ea[0] = -f[0]; ea[1] = -f[1]; ea[2] = -f[2];
        ea[3] = -(m[0] + ra[1] * f[2] - ra[2] * f[1]);
        ea[4] = -(m[1] + ra[2] * f[0] - ra[0] * f[2]);
        ea[5] = -(m[2] + ra[0] * f[1] - ra[1] * f[0]);
    }
}
// second half: each cell's reg x plus its run in order (fixed cells: 0)
__device__ void fine_sum(const DLevel& L, const double* x, double* y, Span t) {
    for (int k = t.first; k < L.n; k += t.stride) {
        double* yk = y + 6 * (size_t)k;
        if (L.fixed[k]) {
            for (int i = 0; i < 6; i++) yk[i] = 0.0;
            continue;
        }
        double rg = L.reg[k], acc[6];
        for (int i = 0; i < 6; i++) acc[i] = rg * x[6 * (size_t)k + i];
        for (int e = L.ioff[k]; e < L.ioff[k + 1]; e++) {
            const double* b = L.buf + 6 * (size_t)e;
            for (int i = 0; i < 6; i++) acc[i] += b[i];
        }
        for (int i = 0; i < 6; i++) yk[i] = acc[i];
    }
}
// y = K x on a coarse level (structure.rs:651-660): each coupling's products for its two blocks, then each block's own
// 6 x 6 and its run in order
__device__ void coarse_pull(const DLevel& L, const double* x, Span t) {
    for (int o = t.first; o < L.nc; o += t.stride) {
        const double* M = L.om + o;
        const size_t ld = (size_t)L.nc;
        const double* xa = x + 6 * (size_t)L.oends[2 * o];
        const double* xc = x + 6 * (size_t)L.oends[2 * o + 1];
        double* ea = L.buf + 6 * (size_t)L.cpos[2 * o];
        double* ec = L.buf + 6 * (size_t)L.cpos[2 * o + 1];
        for (int i = 0; i < 6; i++) {
            double s = 0.0, u = 0.0;
            for (int j = 0; j < 6; j++) { s += M[(size_t)(6 * i + j) * ld] * xc[j]; u += M[(size_t)(6 * j + i) * ld] * xa[j]; }
            ea[i] = s;
            ec[i] = u;
        }
    }
}
__device__ void coarse_sum(const DLevel& L, const double* x, double* y, Span t) {
    for (int k = t.first; k < L.n; k += t.stride) {
        const double* m = L.diag + k;
        const size_t ld = (size_t)L.n;
        const double* v = x + 6 * (size_t)k;
        double acc[6];
        for (int i = 0; i < 6; i++) { double s = 0.0; for (int j = 0; j < 6; j++) s += m[(size_t)(6 * i + j) * ld] * v[j]; acc[i] = s; }
        for (int e = L.ioff[k]; e < L.ioff[k + 1]; e++) {
            const double* b = L.buf + 6 * (size_t)e;
            for (int i = 0; i < 6; i++) acc[i] += b[i];
        }
        for (int i = 0; i < 6; i++) y[6 * (size_t)k + i] = acc[i];
    }
}
// z = 0 + w inv r (the first sweep from zero) or z += w inv (r - kz) (structure.rs:661-663)
__device__ void jacobi(const DLevel& L, bool fine, const double* kz, bool first, Span t) {
    for (int k = t.first; k < L.n; k += t.stride) {
        double* z = L.z + 6 * (size_t)k;
        if (fine && L.fixed[k]) {
            if (first) for (int i = 0; i < 6; i++) z[i] = 0.0;
            continue;
        }
        double rr[6];
        const double* r = L.r + 6 * (size_t)k;
        if (first) for (int j = 0; j < 6; j++) rr[j] = r[j];
        else for (int j = 0; j < 6; j++) rr[j] = r[j] - kz[6 * (size_t)k + j];
        const double* iv = L.inv + k;
        const size_t ld = (size_t)L.n;
        for (int i = 0; i < 6; i++) {
            double s = 0.0;
            for (int j = 0; j < 6; j++) s += iv[(size_t)(6 * i + j) * ld] * rr[j];
            double w = OMEGA * s;
            z[i] = (first ? 0.0 : z[i]) + w;
        }
    }
}
// the residual r - K z of level L restricted to its blocks on level N: forces summed, their moments about the block's
// centre added
__device__ void restrict_to(const DLevel& L, const DLevel& N, Span t) {
    for (int c = t.first; c < N.n; c += t.stride) {
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int e = N.choff[c]; e < N.choff[c + 1]; e++) {
            int k = N.ch[e];
            double v[6];
            for (int i = 0; i < 6; i++) v[i] = L.r[6 * (size_t)k + i] - L.kz[6 * (size_t)k + i];
            const double* d = N.dupf + 3 * (size_t)k;
            acc[0] += v[0]; acc[1] += v[1]; acc[2] += v[2];
            acc[3] += v[3] + d[1] * v[2] - d[2] * v[1];
            acc[4] += v[4] + d[2] * v[0] - d[0] * v[2];
            acc[5] += v[5] + d[0] * v[1] - d[1] * v[0];
        }
        for (int i = 0; i < 6; i++) N.r[6 * (size_t)c + i] = acc[i];
    }
}
// each cell of level L moves with its block on level N
__device__ void prolong_from(const DLevel& L, const DLevel& N, Span t) {
    for (int k = t.first; k < L.n; k += t.stride) {
        int c = N.upf[k];
        if (c < 0) continue;
        const double* d = N.dupf + 3 * (size_t)k;
        const double* zc = N.z + 6 * (size_t)c;
        double* z = L.z + 6 * (size_t)k;
        z[0] += zc[0] + zc[4] * d[2] - zc[5] * d[1];
        z[1] += zc[1] + zc[5] * d[0] - zc[3] * d[2];
        z[2] += zc[2] + zc[3] * d[1] - zc[4] * d[0];
        for (int i = 3; i < 6; i++) z[i] += zc[i];
    }
}
__device__ inline void wg_apply(const DCase& C, const DLevel* lv, int l, const double* x, double* y) {
    if (l == 0) { fine_pull(C, lv[0], x, wg_span()); __syncthreads(); fine_sum(lv[0], x, y, wg_span()); }
    else { coarse_pull(lv[l], x, wg_span()); __syncthreads(); coarse_sum(lv[l], x, y, wg_span()); }
}
// the coarsest level's smoothing: a first sweep from zero, then 29 more (structure.rs:733-741)
__device__ void coarsest_sweeps(const DCase& C, const DLevel* lv) {
    int l = C.nl - 1;
    const DLevel& L = lv[l];
    jacobi(L, l == 0, nullptr, true, wg_span());
    __syncthreads();
    for (int s = 1; s < COARSE_SWEEPS; s++) {
        wg_apply(C, lv, l, L.z, L.kz);
        __syncthreads();
        jacobi(L, l == 0, L.kz, false, wg_span());
        __syncthreads();
    }
}
// a halving tree over the workgroup's values in red (the last six steps within the first wavefront)
__device__ double tree_sum(double* red) {
    __syncthreads();
    for (int st = NT_SOLVE / 2; st >= 64; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 64) {
        double v = red[threadIdx.x];
        for (int st = 32; st > 0; st >>= 1) {
            double o = __shfl_down(v, st, 64);
            if ((int)threadIdx.x < st) v += o;
        }
        if (threadIdx.x == 0) red[0] = v;
    }
    __syncthreads();
    double v = red[0];
    __syncthreads();
    return v;
}
__device__ inline double strided_sum(const double* a, const double* b, int n6, int t) {
    double s = 0.0;
    for (int e = t; e < n6; e += NPART) s += a[e] * b[e];
    return s;
}
// (with 512 threads each holds two of the 1,024 partial sums and adds them: the tree's first step)
__device__ double block_dot(const double* a, const double* b, int n6, double* red) {
    double v = strided_sum(a, b, n6, threadIdx.x);
    if (NT_SOLVE < NPART) v += strided_sum(a, b, n6, threadIdx.x + NT_SOLVE);
    red[threadIdx.x] = v;
    return tree_sum(red);
}
// one V-cycle: smooth, correct on the next level, smooth (structure.rs:725-764), as a loop down and up the levels
__device__ void vcycle(const DCase& C, const DLevel* lv) {
    int nl = C.nl;
    for (int l = 0; l < nl - 1; l++) {
        const DLevel& L = lv[l];
        jacobi(L, l == 0, nullptr, true, wg_span());
        __syncthreads();
        wg_apply(C, lv, l, L.z, L.kz);
        __syncthreads();
        restrict_to(L, lv[l + 1], wg_span());
        __syncthreads();
    }
    coarsest_sweeps(C, lv);
    for (int l = nl - 2; l >= 0; l--) {
        const DLevel& L = lv[l];
        prolong_from(L, lv[l + 1], wg_span());
        __syncthreads();
        wg_apply(C, lv, l, L.z, L.kz);
        __syncthreads();
        jacobi(L, l == 0, L.kz, false, wg_span());
        __syncthreads();
    }
}
// copy 8-byte words into LDS with the whole workgroup
__device__ inline void lds_copy(void* dst, const void* src, size_t bytes) {
    const uint64_t* s = (const uint64_t*)src;
    uint64_t* d = (uint64_t*)dst;
    for (size_t i = threadIdx.x; i < bytes / 8; i += NT_SOLVE) d[i] = s[i];
}
// the level descriptors into LDS; the coarsest level's blocks, couplings and vectors too when the case has room for
// them (its 30 sweeps a cycle then run from LDS). with_r: bring the coarsest level's right-hand side along.
__device__ void load_levels(const DCase& C, DLevel* lv, uint8_t* dyn, bool with_r) {
    if (threadIdx.x < MAX_LEVELS) lv[threadIdx.x] = C.lv[threadIdx.x];
    __syncthreads();
    if (!C.lds) return;
    const DLevel& G = C.lv[C.nl - 1];
    size_t n = (size_t)G.n, nc = (size_t)G.nc;
    double* dp = (double*)dyn;
    double *diag = dp, *inv = diag + 36 * n, *om = inv + 36 * n, *buf = om + 36 * nc, *r = buf + 12 * nc, *z = r + 6 * n,
           *kz = z + 6 * n;
    int32_t* ip = (int32_t*)(kz + 6 * n);
    int32_t *ioff = ip, *cpos = ioff + (n + 1), *oends = cpos + 2 * nc;
    lds_copy(diag, G.diag, 8 * 36 * n);
    lds_copy(inv, G.inv, 8 * 36 * n);
    lds_copy(om, G.om, 8 * 36 * nc);
    if (with_r) lds_copy(r, G.r, 8 * 6 * n);
    for (size_t i = threadIdx.x; i < n + 1; i += NT_SOLVE) ioff[i] = G.ioff[i];
    for (size_t i = threadIdx.x; i < 2 * nc; i += NT_SOLVE) { cpos[i] = G.cpos[i]; oends[i] = G.oends[i]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        DLevel& L = lv[C.nl - 1];
        L.diag = diag; L.inv = inv; L.om = om; L.buf = buf; L.r = r; L.z = z; L.kz = kz;
        L.ioff = ioff; L.cpos = cpos; L.oends = oends;
    }
    __syncthreads();
}
// the whole conjugate-gradient solve of one case in one workgroup (structure.rs:566-617)
__global__ __launch_bounds__(NT_SOLVE) void k_net_solve(const DCase* cases) {
    __shared__ double red[NT_SOLVE];
    __shared__ DLevel lv[MAX_LEVELS];
    extern __shared__ __align__(16) uint8_t dyn[];
    const DCase& C = cases[blockIdx.x];
    if (C.route) return;   // solved a kernel per step
    load_levels(C, lv, dyn, false);
    const DLevel& F = lv[0];
    int n6 = 6 * C.n0;
    double *r = F.r, *z = F.z;
    for (int k = threadIdx.x; k < C.n0; k += NT_SOLVE)
        for (int i = 0; i < 6; i++) r[6 * k + i] = F.fixed[k] ? 0.0 : C.f[6 * k + i];
    __syncthreads();
    vcycle(C, lv);
    double r0 = fmax(block_dot(r, z, n6, red), 1e-30);
    if (C.x0) {
        for (int e = threadIdx.x; e < n6; e += NT_SOLVE) C.x[e] = C.x0[e];
        __syncthreads();
        wg_apply(C, lv, 0, C.x, C.ap);
        __syncthreads();
        for (int e = threadIdx.x; e < n6; e += NT_SOLVE) r[e] -= C.ap[e];
        __syncthreads();
        vcycle(C, lv);
    } else {
        for (int e = threadIdx.x; e < n6; e += NT_SOLVE) C.x[e] = 0.0;
    }
    for (int e = threadIdx.x; e < n6; e += NT_SOLVE) C.p[e] = z[e];
    __syncthreads();
    double rz = block_dot(r, z, n6, red);
    int it = 0;
    if (!(rz < TOL * r0)) {
        while (it < MAX_IT) {
            wg_apply(C, lv, 0, C.p, C.ap);
            __syncthreads();
            double pap = block_dot(C.p, C.ap, n6, red);
            if (pap <= 0.0) break;
            double alpha = rz / pap;
            for (int e = threadIdx.x; e < n6; e += NT_SOLVE) { C.x[e] += alpha * C.p[e]; r[e] -= alpha * C.ap[e]; }
            it++;
            __syncthreads();
            vcycle(C, lv);
            double rz2 = block_dot(r, z, n6, red);
            if (rz2 < TOL * r0) break;
            double beta = rz2 / rz;
            rz = rz2;
            for (int e = threadIdx.x; e < n6; e += NT_SOLVE) C.p[e] = z[e] + beta * C.p[e];
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) *C.out_it = it;
}

// ---- the solve of one large case spread over the chip: the same steps as k_net_solve, each a kernel over a grid,
// the scalars of the iteration in a small device record. Every kernel returns at once after the solve has ended
// (converged, pap <= 0 or MAX_IT), so the host may queue iterations in blocks and stop on the first block that ends.
struct NetState { int32_t done, it, pad[2]; double r0, rz, alpha, beta; double part[NPART]; };
enum : int { DOT_R0 = 0, DOT_RZ = 1, DOT_PAP = 2, DOT_RZ2 = 3 };
#define MK_GUARD(S) do { if ((S)->done) return; } while (0)
__global__ void k_mk_init_r(const DCase* C, const NetState* S) {
    const DLevel& F = C->lv[0];
    Span t = grid_span();
    for (int k = t.first; k < C->n0; k += t.stride)
        for (int i = 0; i < 6; i++) F.r[6 * (size_t)k + i] = F.fixed[k] ? 0.0 : C->f[6 * (size_t)k + i];
}
__global__ void k_mk_pull(const DCase* C, const NetState* S, int l, int which) {
    MK_GUARD(S);
    const DLevel& L = C->lv[l];
    const double* x = which == 0 ? L.z : (which == 1 ? C->p : C->x);
    if (l == 0) fine_pull(*C, L, x, grid_span()); else coarse_pull(L, x, grid_span());
}
__global__ void k_mk_sum(const DCase* C, const NetState* S, int l, int which) {
    MK_GUARD(S);
    const DLevel& L = C->lv[l];
    const double* x = which == 0 ? L.z : (which == 1 ? C->p : C->x);
    double* y = which == 0 ? L.kz : C->ap;
    if (l == 0) fine_sum(L, x, y, grid_span()); else coarse_sum(L, x, y, grid_span());
}
__global__ void k_mk_jacobi(const DCase* C, const NetState* S, int l, int first) {
    MK_GUARD(S);
    const DLevel& L = C->lv[l];
    jacobi(L, l == 0, L.kz, first != 0, grid_span());
}
__global__ void k_mk_restrict(const DCase* C, const NetState* S, int l) {
    MK_GUARD(S);
    restrict_to(C->lv[l], C->lv[l + 1], grid_span());
}
__global__ void k_mk_prolong(const DCase* C, const NetState* S, int l) {
    MK_GUARD(S);
    prolong_from(C->lv[l], C->lv[l + 1], grid_span());
}
__global__ __launch_bounds__(NT_SOLVE) void k_mk_coarsest(const DCase* C, const NetState* S) {
    MK_GUARD(S);
    __shared__ DLevel lv[MAX_LEVELS];
    extern __shared__ __align__(16) uint8_t dyn[];
    load_levels(*C, lv, dyn, true);
    coarsest_sweeps(*C, lv);
    if (C->lds) {
        const DLevel& G = C->lv[C->nl - 1];
        for (int e = threadIdx.x; e < 6 * G.n; e += NT_SOLVE) G.z[e] = lv[C->nl - 1].z[e];
    }
}
// the 1,024 strided partial sums of a dot product (4 x 256 threads)
__global__ void k_mk_dot_part(const DCase* C, NetState* S, int which) {
    MK_GUARD(S);
    const DLevel& F = C->lv[0];
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const double *a = which == 1 ? C->p : F.r, *b = which == 1 ? C->ap : F.z;
    if (t < NPART) S->part[t] = strided_sum(a, b, 6 * C->n0, t);
}
// the tree over the partial sums, then the iteration's scalar step
__global__ __launch_bounds__(NT_SOLVE) void k_mk_dot_end(NetState* S, int mode) {
    if (S->done) return;
    __shared__ double red[NT_SOLVE];
    double v0 = S->part[threadIdx.x];
    if (NT_SOLVE < NPART) v0 += S->part[threadIdx.x + NT_SOLVE];
    red[threadIdx.x] = v0;
    double v = tree_sum(red);
    if (threadIdx.x != 0) return;
    if (mode == DOT_R0) S->r0 = fmax(v, 1e-30);
    else if (mode == DOT_RZ) { S->rz = v; if (v < TOL * S->r0) S->done = 1; }
    else if (mode == DOT_PAP) { if (v <= 0.0) S->done = 1; else { S->alpha = S->rz / v; S->it += 1; } }
    else {
        if (v < TOL * S->r0) S->done = 1;
        else { S->beta = v / S->rz; S->rz = v; if (S->it >= MAX_IT) S->done = 1; }
    }
}
// x = x0 (or 0) with r -= K x0 to follow; p = z; x += alpha p, r -= alpha ap; p = z + beta p
__global__ void k_mk_vec(const DCase* C, const NetState* S, int op) {
    MK_GUARD(S);
    const DLevel& F = C->lv[0];
    Span t = grid_span();
    int n6 = 6 * C->n0;
    for (int e = t.first; e < n6; e += t.stride) {
        if (op == 0) C->x[e] = C->x0 ? C->x0[e] : 0.0;
        else if (op == 1) F.r[e] -= C->ap[e];
        else if (op == 2) C->p[e] = F.z[e];
        else if (op == 3) { C->x[e] += S->alpha * C->p[e]; F.r[e] -= S->alpha * C->ap[e]; }
        else C->p[e] = F.z[e] + S->beta * C->p[e];
    }
}
__global__ void k_mk_state_init(NetState* S) {
    if (threadIdx.x == 0 && blockIdx.x == 0) { S->done = 0; S->it = 0; }
}
__global__ void k_mk_finish(const DCase* C, const NetState* S) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *C->out_it = S->it;
}
// bytes of LDS the coarsest level takes (the layout load_levels uses)
size_t coarse_lds_bytes(int n, int nc) { return 8 * (90 * (size_t)n + 48 * (size_t)nc) + 4 * ((size_t)n + 1 + 4 * (size_t)nc); }
// stress over strength of a face under force f and moment m in its frame (structure.rs:532-550)
__device__ float section_ratio(const double* s, const double* f, const double* m, const float* cap, float flex) {
    double area = s[0], iu = s[1], iw = s[2], zu = s[3], zw = s[4], j = s[5], rmax = s[6];
    double sn = f[0] / area;
    double bend = fabs(m[1]) * zu / iu + fabs(m[2]) * zw / iw;
    double top = sn + bend, bottom = sn - bend;
    double r = 0.0;
    if (top > 0.0) {
        double tens = fmax(sn, 0.0);
        double fe = (tens * (double)cap[1] + bend * (double)flex) / fmax(tens + bend, 1e-9);
        r = fmax(r, top / fmax(fe, 1.0));
    }
    if (bottom < 0.0) r = fmax(r, -bottom / (double)cap[0]);
    double v = sqrt(f[1] * f[1] + f[2] * f[2]) / area;
    double tor = fabs(m[0]) * rmax / j;
    double cohesion = (double)cap[2] + 0.4 * fmax(-sn, 0.0);
    r = fmax(r, (v + tor) / cohesion);
    return (float)r;
}
__global__ void k_net_checks(const DCheck* checks, int n, const DCase* cases, float* ratio, float* flow, float* moment) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const DCheck& K = checks[t];
        const DCase& C = cases[K.cs];
        int ax = K.axis, u = (ax + 1) % 3, w = (ax + 2) % 3;
        double ft[3] = {0.0, 0.0, 0.0}, mt[3] = {0.0, 0.0, 0.0};
        if (K.kind == 0) {
            for (int e = K.l0; e < K.l1; e++) {
                int l = K.glinks[e];
                const size_t ld = (size_t)C.lv[0].nc;
                double S[12], fo[3], mo[3];
                spring_fields(C.sp, ld, l, S);
                spring_load(S, C.x + 6 * (size_t)C.ends[2 * l], C.x + 6 * (size_t)C.ends[2 * l + 1], fo, mo);
                double rr[3] = {C.sp[12 * ld + l] - K.c[0], C.sp[13 * ld + l] - K.c[1], C.sp[14 * ld + l] - K.c[2]};
                for (int q = 0; q < 3; q++) { ft[q] += fo[q]; mt[q] += mo[q]; }
                double rx[3] = {rr[1] * fo[2] - rr[2] * fo[1], rr[2] * fo[0] - rr[0] * fo[2], rr[0] * fo[1] - rr[1] * fo[0]};
                for (int q = 0; q < 3; q++) mt[q] += rx[q];
            }
        } else {
            int l = K.l0;
            double S[12];
            spring_fields(C.sp, (size_t)C.lv[0].nc, l, S);
            spring_load(S, C.x + 6 * (size_t)C.ends[2 * l], C.x + 6 * (size_t)C.ends[2 * l + 1], ft, mt);
        }
        double fl[3] = {ft[ax], ft[u], ft[w]}, ml[3] = {mt[ax], mt[u], mt[w]};
        if (K.hinge) { ml[0] = 0.0; ml[1] = 0.0; ml[2] = 0.0; }
        ratio[t] = section_ratio(K.sec, fl, ml, K.cap, K.flex);
        flow[t] = (float)ft[ax];
        moment[t] = (float)(mt[u] * mt[u] + mt[w] * mt[w]);
    }
}

// ---------------------------------------------------------------- packing a batch into one device arena
// The arena holds the cases' data (uploaded) then their work vectors (never uploaded: every one is written before it
// is read). A first pass only measures, so the host buffer is allocated once.
struct Arena {
    std::vector<uint8_t> host;
    size_t size = 0;
    bool measure = true;
    size_t put(const void* p, size_t bytes) {
        size_t o = (size + 255) & ~(size_t)255;
        size = o + bytes;
        if (!measure && bytes && p) std::memcpy(host.data() + o, p, bytes);
        return o;
    }
    size_t reserve(size_t bytes) { size_t o = (size + 255) & ~(size_t)255; size = o + bytes; return o; }
};
template <class T> size_t put_vec(Arena& A, const std::vector<T>& v) { return A.put(v.data(), sizeof(T) * v.size()); }
// records of `rec` numbers each, stored field after field (field q of record i at q * n + i)
size_t put_fields(Arena& A, const std::vector<double>& aos, size_t rec) {
    size_t n = aos.size() / rec;
    size_t o = A.put(nullptr, sizeof(double) * aos.size());
    if (!A.measure) {
        double* d = reinterpret_cast<double*>(A.host.data() + o);
        for (size_t i = 0; i < n; i++)
            for (size_t q = 0; q < rec; q++) d[q * n + i] = aos[rec * i + q];
    }
    return o;
}

struct CaseOffsets {                           // arena offsets of one case (device pointers = base + offset)
    size_t sp, ends, f, x0, reg0, fixed0, glinks;
    size_t diag[MAX_LEVELS], inv[MAX_LEVELS], ioff[MAX_LEVELS], cpos[MAX_LEVELS], om[MAX_LEVELS], oends[MAX_LEVELS];
    size_t choff[MAX_LEVELS], ch[MAX_LEVELS], upf[MAX_LEVELS], dupf[MAX_LEVELS];
    // work vectors
    size_t x, p, ap, it, buf, state;
    size_t r[MAX_LEVELS], z[MAX_LEVELS], kz[MAX_LEVELS];
};
// per cell, the run of its ends' contributions (end a: 2 e, end b: 2 e + 1 in `cpos`), in the order of the edges
void contribution_runs(int n, const std::vector<uint32_t>& a, const std::vector<uint32_t>& b, std::vector<int32_t>& off,
                       std::vector<int32_t>& cpos) {
    size_t ne = a.size();
    off.assign((size_t)n + 1, 0);
    cpos.assign(2 * std::max<size_t>(ne, 1), 0);
    for (size_t e = 0; e < ne; e++) { off[a[e] + 1]++; off[b[e] + 1]++; }
    for (int c = 0; c < n; c++) off[c + 1] += off[c];
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (size_t e = 0; e < ne; e++) { cpos[2 * e] = fill[a[e]]++; cpos[2 * e + 1] = fill[b[e]]++; }
}
void pack_data(const Case& C, Arena& A, CaseOffsets& O) {
    size_t nl = C.la.size();
    O.sp = put_fields(A, C.spring, 15);
    std::vector<int32_t> ends(2 * std::max<size_t>(nl, 1), 0);
    for (size_t l = 0; l < nl; l++) { ends[2 * l] = (int32_t)C.la[l]; ends[2 * l + 1] = (int32_t)C.lb[l]; }
    O.ends = put_vec(A, ends);
    O.f = put_vec(A, C.f);
    O.x0 = C.has_x0 ? put_vec(A, C.x0) : 0;
    const Level& F = C.lv[0];
    O.reg0 = put_vec(A, F.reg);
    O.fixed0 = put_vec(A, F.fixed);
    std::vector<int32_t> off, cpos;
    contribution_runs(F.n, C.la, C.lb, off, cpos);
    O.ioff[0] = put_vec(A, off);
    O.cpos[0] = put_vec(A, cpos);
    O.inv[0] = put_fields(A, F.inv, 36);
    for (size_t l = 1; l < C.lv.size(); l++) {
        const Level& L = C.lv[l];
        const Level& P = C.lv[l - 1];
        size_t no = L.oa.size();
        O.diag[l] = put_fields(A, L.diag, 36);
        O.inv[l] = put_fields(A, L.inv, 36);
        contribution_runs(L.n, L.oa, L.oc, off, cpos);
        O.ioff[l] = put_vec(A, off);
        O.cpos[l] = put_vec(A, cpos);
        O.om[l] = put_fields(A, L.om, 36);
        std::vector<int32_t> oends(2 * std::max<size_t>(no, 1), 0);
        for (size_t o = 0; o < no; o++) { oends[2 * o] = (int32_t)L.oa[o]; oends[2 * o + 1] = (int32_t)L.oc[o]; }
        O.oends[l] = put_vec(A, oends);
        std::vector<int32_t> choff((size_t)L.n + 1, 0), upf(std::max(P.n, 1), -1);
        std::vector<double> dupf(3 * (size_t)std::max(P.n, 1), 0.0);
        for (int k = 0; k < P.n; k++) if (L.up[k] != NONE) choff[L.up[k] + 1]++;
        for (int c = 0; c < L.n; c++) choff[c + 1] += choff[c];
        std::vector<int32_t> ch(std::max(choff[L.n], 1), 0);
        std::vector<int32_t> cf(choff.begin(), choff.end() - 1);
        for (int k = 0; k < P.n; k++) {
            uint32_t c = L.up[k];
            if (c == NONE) continue;
            upf[k] = (int32_t)c;
            ch[cf[c]++] = k;
            for (int d = 0; d < 3; d++) dupf[3 * k + d] = P.pos[3 * k + d] - L.pos[3 * c + d];
        }
        O.choff[l] = put_vec(A, choff);
        O.ch[l] = put_vec(A, ch);
        O.upf[l] = put_vec(A, upf);
        O.dupf[l] = put_vec(A, dupf);
    }
    O.glinks = put_vec(A, C.glinks);
}
void pack_work(const Case& C, Arena& A, CaseOffsets& O) {
    size_t n0 = std::max<size_t>(C.cbrick.size(), 1);
    O.x = A.reserve(sizeof(double) * 6 * n0);
    O.p = A.reserve(sizeof(double) * 6 * n0);
    O.ap = A.reserve(sizeof(double) * 6 * n0);
    O.it = A.reserve(sizeof(int32_t));
    O.state = A.reserve(sizeof(NetState));
    size_t nb = 2 * std::max<size_t>(C.la.size(), 1);
    for (size_t l = 1; l < C.lv.size(); l++) nb = std::max(nb, 2 * C.lv[l].oa.size());
    O.buf = A.reserve(sizeof(double) * 6 * nb);
    for (size_t l = 0; l < C.lv.size(); l++) {
        size_t n = std::max(C.lv[l].n, 1);
        O.r[l] = A.reserve(sizeof(double) * 6 * n);
        O.z[l] = A.reserve(sizeof(double) * 6 * n);
        O.kz[l] = A.reserve(sizeof(double) * 6 * n);
    }
}

template <class T> T* at(uint8_t* base, size_t off) { return reinterpret_cast<T*>(base + off); }
double since(std::chrono::steady_clock::time_point t) { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count(); }

void parallel_for(int n, int threads, const std::function<void(int)>& fn) {
    if (n <= 1 || threads <= 1) { for (int i = 0; i < n; i++) fn(i); return; }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(threads, n); t++)
        pool.emplace_back([&]() { for (int i = next++; i < n; i = next++) fn(i); });
    for (auto& th : pool) th.join();
}

}  // namespace

struct sf_net {
    std::vector<Case> cases;
    std::vector<const Store*> stores;
    uint8_t* dev = nullptr;
    size_t dev_bytes = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int host_threads = 0;
    double ms[4] = {0, 0, 0, 0};   // host build, upload, device solve, device checks
    double host_ms[4] = {0, 0, 0, 0};  // summed over cases: cells and links, levels, loads and checks; packing
    size_t lds_budget = 0;             // bytes of dynamic LDS a workgroup may take for the coarsest level
    int route_min_cells = -1;          // a case of at least this many cells, in a batch of at most route_max_cases,
    int route_max_cases = 16;          // is solved a kernel per step over the whole chip (-1: never; the default, that
                                       // route being switched off: it made a GPU memory fault whose cause is not found)
    int block_its = 6;                 // iterations queued between checks of a per-step solve
    int32_t* pinned = nullptr;         // the per-step solve's state read-back
    ~sf_net() {
        if (dev) (void)hipFree(dev);
        if (pinned) (void)hipHostFree(pinned);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace {
int ensure_device(sf_net* N, size_t bytes) {
    if (!N->st) {
        SF_HIP(hipStreamCreateWithFlags(&N->st, hipStreamNonBlocking));
        for (auto& e : N->ev) SF_HIP(hipEventCreate(&e));
        // dynamic LDS for the coarsest level: what a workgroup may hold beyond the kernel's own 10 KB, at most 128 KB
        int dev = 0, lds = 0;
        SF_HIP(hipGetDevice(&dev));
        SF_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        // (both kernels that take the coarsest level into LDS must be allowed the same dynamic size: a launch beyond
        // a kernel's allowance leaves the kernel's LDS accesses outside its allocation)
        size_t budget = lds > 16384 ? std::min<size_t>((size_t)lds - 16384, 131072) : 0;
        if (budget > 49152) {
            const void* ks[2] = {(const void*)k_net_solve, (const void*)k_mk_coarsest};
            for (const void* k : ks)
                if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)budget) != hipSuccess) {
                    (void)hipGetLastError();
                    budget = 49152;
                }
        }
        N->lds_budget = budget;
        SF_HIP(hipHostMalloc((void**)&N->pinned, 64, hipHostMallocDefault));
    }
    if (bytes > N->dev_bytes) {
        if (N->dev) SF_HIP(hipFree(N->dev));
        N->dev = nullptr;
        size_t want = std::max(bytes, N->dev_bytes * 3 / 2);
        SF_HIP(hipMalloc(&N->dev, want));
        N->dev_bytes = want;
    }
    return 0;
}
}  // namespace

namespace {
dim3 mk_grid(size_t n) { return dim3((unsigned)std::min<size_t>(std::max<size_t>((n + 255) / 256, 1), 1024)); }
// one V-cycle of a case solved a kernel per step (the order of vcycle's loops)
int mk_vcycle(sf_net* N, const DCase& h, const DCase* C, const NetState* S) {
    hipStream_t st = N->st;
    int nl = h.nl;
    auto apply = [&](int l, int which) {
        size_t np = h.lv[l].nc, ns = h.lv[l].n;
        hipLaunchKernelGGL(k_mk_pull, mk_grid(np), dim3(256), 0, st, C, S, l, which);
        hipLaunchKernelGGL(k_mk_sum, mk_grid(ns), dim3(256), 0, st, C, S, l, which);
    };
    for (int l = 0; l < nl - 1; l++) {
        hipLaunchKernelGGL(k_mk_jacobi, mk_grid(h.lv[l].n), dim3(256), 0, st, C, S, l, 1);
        apply(l, 0);
        hipLaunchKernelGGL(k_mk_restrict, mk_grid(h.lv[l + 1].n), dim3(256), 0, st, C, S, l);
    }
    hipLaunchKernelGGL(k_mk_coarsest, dim3(1), dim3(NT_SOLVE), (size_t)h.lds, st, C, S);
    for (int l = nl - 2; l >= 0; l--) {
        hipLaunchKernelGGL(k_mk_prolong, mk_grid(h.lv[l].n), dim3(256), 0, st, C, S, l);
        apply(l, 0);
        hipLaunchKernelGGL(k_mk_jacobi, mk_grid(h.lv[l].n), dim3(256), 0, st, C, S, l, 0);
    }
    SF_HIP(hipGetLastError());
    return 0;
}
// the conjugate-gradient solve of one case, a kernel per step (the order of k_net_solve)
int mk_solve(sf_net* N, const DCase& h, const DCase* C, NetState* S) {
    hipStream_t st = N->st;
    size_t n6 = 6 * (size_t)h.n0;
    auto dot = [&](int which, int mode) {
        hipLaunchKernelGGL(k_mk_dot_part, dim3(NPART / 256), dim3(256), 0, st, C, S, which);
        hipLaunchKernelGGL(k_mk_dot_end, dim3(1), dim3(NT_SOLVE), 0, st, S, mode);
    };
    auto vec = [&](int op) { hipLaunchKernelGGL(k_mk_vec, mk_grid(n6), dim3(256), 0, st, C, S, op); };
    auto apply0 = [&](int which) {
        hipLaunchKernelGGL(k_mk_pull, mk_grid(h.lv[0].nc), dim3(256), 0, st, C, S, 0, which);
        hipLaunchKernelGGL(k_mk_sum, mk_grid(h.lv[0].n), dim3(256), 0, st, C, S, 0, which);
    };
    hipLaunchKernelGGL(k_mk_state_init, dim3(1), dim3(64), 0, st, S);
    hipLaunchKernelGGL(k_mk_init_r, mk_grid(h.n0), dim3(256), 0, st, C, S);
    if (mk_vcycle(N, h, C, S)) return -1;
    dot(0, DOT_R0);
    vec(0);
    if (h.x0) {
        apply0(2);
        vec(1);
        if (mk_vcycle(N, h, C, S)) return -1;
    }
    vec(2);
    dot(0, DOT_RZ);
    for (;;) {
        for (int k = 0; k < N->block_its; k++) {
            apply0(1);
            dot(1, DOT_PAP);
            vec(3);
            if (mk_vcycle(N, h, C, S)) return -1;
            dot(0, DOT_RZ2);
            vec(4);
        }
        SF_HIP(hipMemcpyAsync(N->pinned, S, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        SF_HIP(hipStreamSynchronize(st));
        if (N->pinned[0]) break;
    }
    hipLaunchKernelGGL(k_mk_finish, dim3(1), dim3(64), 0, st, C, S);
    SF_HIP(hipGetLastError());
    return 0;
}
}  // namespace


// ---------------------------------------------------------------- the store's byte format (brick.rs:100-145)
// Each array as its length (u32, little endian) and its elements, padded with zeros to 8 bytes: x y z (i16), sx sy sz
// mat paint flags (u8), structure (u16), body jhead (u32), damage hull (u8), ja jb jna jnb (u32), jarea (u16), jaxis
// jdmg (u8), the free joint slots (u32), the live joint count (one u32), the spatial index's cell keys (u64) and first
// bricks (u32), and each brick's next in its cell (u32). The joint lists (jhead, jna, jnb) are each brick's joints
// newest first; the index puts each brick in the cell of its min corner at the finest of four levels (1, 4, 16, 64 m)
// as large as the brick, later bricks first in a cell. A map's entry order carries no meaning: keys are written in
// ascending order, and any order is read.
namespace {
struct Out {
    std::vector<uint8_t> b;
    template <class T> void put(const T* v, size_t n) {
        const uint32_t len = (uint32_t)n;
        const uint8_t* p = reinterpret_cast<const uint8_t*>(&len);
        b.insert(b.end(), p, p + 4);
        const uint8_t* q = reinterpret_cast<const uint8_t*>(v);
        b.insert(b.end(), q, q + n * sizeof(T));
        while (b.size() % 8) b.push_back(0);
    }
    template <class T> void put(const std::vector<T>& v) { put(v.data(), v.size()); }
};
struct In {
    const uint8_t* src;
    size_t n, at = 0;
    bool ok = true;
    template <class T> std::vector<T> get() {
        std::vector<T> v;
        if (!ok || at + 4 > n) { ok = false; return v; }
        uint32_t len;
        std::memcpy(&len, src + at, 4);
        at += 4;
        if ((uint64_t)len * sizeof(T) > n - at) { ok = false; return v; }
        v.resize(len);
        if (len) std::memcpy(v.data(), src + at, (size_t)len * sizeof(T));
        at += (size_t)len * sizeof(T);
        at = (at + 7) & ~(size_t)7;
        return v;
    }
};
constexpr int IDX_LEVELS = 4;
inline int32_t lcell(int l) { return 8 << (2 * l); }
inline int32_t fdiv(int32_t a, int32_t c) { int32_t q = a / c; return (a % c != 0 && ((a < 0) != (c < 0))) ? q - 1 : q; }
inline uint64_t ckey(int l, int32_t x, int32_t y, int32_t z) {
    return ((uint64_t)l << 60) | (((uint64_t)(x + (1 << 19)) & 0xfffffull) << 40) | (((uint64_t)(y + (1 << 19)) & 0xfffffull) << 20) |
           ((uint64_t)(z + (1 << 19)) & 0xfffffull);
}
std::vector<uint8_t> save_store(const Store& S) {
    const size_t n = S.size(), nj = S.ja.size();
    std::vector<int16_t> x(n), y(n), z(n);
    std::vector<uint8_t> sx(n), sy(n), sz(n);
    std::vector<uint32_t> jhead(n, NONE), jna(nj, NONE), jnb(nj, NONE);
    for (size_t i = 0; i < n; i++) {
        x[i] = (int16_t)S.lo[3 * i]; y[i] = (int16_t)S.lo[3 * i + 1]; z[i] = (int16_t)S.lo[3 * i + 2];
        sx[i] = S.ext[3 * i]; sy[i] = S.ext[3 * i + 1]; sz[i] = S.ext[3 * i + 2];
        const std::vector<uint32_t>& L = S.inc[i];
        if (!L.empty()) jhead[i] = L[0];
        for (size_t k = 0; k < L.size(); k++) {
            const uint32_t j = L[k], nx = k + 1 < L.size() ? L[k + 1] : NONE;
            if (S.ja[j] == (uint32_t)i) jna[j] = nx; else jnb[j] = nx;
        }
    }
    // (free slots keep the values they had: ends NONE; their next pointers are never read)
    std::map<uint64_t, uint32_t> cells;
    std::vector<uint32_t> cnext(n, NONE);
    for (size_t i = 0; i < n; i++) {
        int32_t m = std::max<int32_t>(std::max<int32_t>(S.ext[3 * i], S.ext[3 * i + 1]), S.ext[3 * i + 2]);
        int l = 0;
        while (l < IDX_LEVELS - 1 && lcell(l) < m) l++;
        const int32_t c = lcell(l);
        const uint64_t key = ckey(l, fdiv(S.lo[3 * i], c), fdiv(S.lo[3 * i + 1], c), fdiv(S.lo[3 * i + 2], c));
        auto it = cells.find(key);
        cnext[i] = it == cells.end() ? NONE : it->second;
        cells[key] = (uint32_t)i;
    }
    std::vector<uint64_t> keys;
    std::vector<uint32_t> vals;
    for (const auto& kv : cells) { keys.push_back(kv.first); vals.push_back(kv.second); }
    const uint32_t live = (uint32_t)S.live;
    Out o;
    o.put(x); o.put(y); o.put(z); o.put(sx); o.put(sy); o.put(sz);
    o.put(S.mat); o.put(S.paint); o.put(S.flags);
    o.put(S.sid); o.put(S.body); o.put(jhead); o.put(S.damage); o.put(S.hull);
    o.put(S.ja); o.put(S.jb); o.put(jna); o.put(jnb);
    o.put(S.jarea); o.put(S.jaxis); o.put(S.jdmg);
    o.put(S.jfree); o.put(&live, 1);
    o.put(keys); o.put(vals); o.put(cnext);
    return o.b;
}
// the store from its bytes; false (with the reason) on a malformed or inconsistent input
bool load_store(const uint8_t* src, size_t nbytes, Store& S, const char** why) {
    In r{src, nbytes};
    auto x = r.get<int16_t>(), y = r.get<int16_t>(), z = r.get<int16_t>();
    auto sx = r.get<uint8_t>(), sy = r.get<uint8_t>(), sz = r.get<uint8_t>();
    S.mat = r.get<uint8_t>(); S.paint = r.get<uint8_t>(); S.flags = r.get<uint8_t>();
    S.sid = r.get<uint16_t>(); S.body = r.get<uint32_t>();
    auto jhead = r.get<uint32_t>();
    S.damage = r.get<uint8_t>(); S.hull = r.get<uint8_t>();
    S.ja = r.get<uint32_t>(); S.jb = r.get<uint32_t>();
    auto jna = r.get<uint32_t>(), jnb = r.get<uint32_t>();
    S.jarea = r.get<uint16_t>(); S.jaxis = r.get<uint8_t>(); S.jdmg = r.get<uint8_t>();
    S.jfree = r.get<uint32_t>();
    auto live = r.get<uint32_t>();
    auto keys = r.get<uint64_t>();
    auto vals = r.get<uint32_t>();
    auto cnext = r.get<uint32_t>();
    if (!r.ok) { *why = "truncated"; return false; }
    const size_t n = x.size(), nj = S.ja.size();
    for (size_t m : {y.size(), z.size(), sx.size(), sy.size(), sz.size(), S.mat.size(), S.paint.size(), S.flags.size(),
                     S.sid.size(), S.body.size(), jhead.size(), S.damage.size(), S.hull.size(), cnext.size()})
        if (m != n) { *why = "brick arrays of unequal lengths"; return false; }
    for (size_t m : {S.jb.size(), jna.size(), jnb.size(), S.jarea.size(), S.jaxis.size(), S.jdmg.size()})
        if (m != nj) { *why = "joint arrays of unequal lengths"; return false; }
    if (live.size() != 1 || keys.size() != vals.size()) { *why = "malformed counts"; return false; }
    S.lo.resize(3 * n); S.ext.resize(3 * n);
    for (size_t i = 0; i < n; i++) {
        S.lo[3 * i] = x[i]; S.lo[3 * i + 1] = y[i]; S.lo[3 * i + 2] = z[i];
        S.ext[3 * i] = sx[i]; S.ext[3 * i + 1] = sy[i]; S.ext[3 * i + 2] = sz[i];
        if (!sx[i] || !sy[i] || !sz[i] || S.mat[i] > 4) { *why = "a brick's size or material out of range"; return false; }
    }
    // each brick's joint list from its linked list (bounded by the joint count: a cycle is malformed)
    S.inc.assign(n, {});
    int nlive = 0;
    for (size_t j = 0; j < nj; j++) {
        if (S.ja[j] == NONE) continue;
        if (S.ja[j] >= n || S.jb[j] >= n) { *why = "a joint's end out of range"; return false; }
        nlive++;
    }
    for (size_t i = 0; i < n; i++) {
        uint32_t j = jhead[i];
        size_t steps = 0;
        while (j != NONE) {
            if (j >= nj || steps++ > nj || (S.ja[j] != (uint32_t)i && S.jb[j] != (uint32_t)i)) { *why = "a joint list that does not close"; return false; }
            S.inc[i].push_back(j);
            j = S.ja[j] == (uint32_t)i ? jna[j] : jnb[j];
        }
    }
    for (uint32_t j : S.jfree) if (j >= nj || S.ja[j] != NONE) { *why = "a free joint slot in use"; return false; }
    if (live[0] != (uint32_t)nlive) { *why = "the live joint count disagrees"; return false; }
    S.live = nlive;
    return true;
}
}  // namespace

extern "C" {

sf_bricks* sf_bricks_create(void) { return new sf_bricks(); }
int64_t sf_bricks_save(const sf_bricks* b, uint8_t* out, int64_t cap) {
    if (!b) { sf::set_error(SF_E_ARG, "sf_bricks_save: null store"); return SF_E_ARG; }
    const std::vector<uint8_t> v = save_store(b->s);
    if (out) {
        if (cap < (int64_t)v.size()) { sf::set_error(SF_E_CAP, "sf_bricks_save: buffer too small"); return SF_E_CAP; }
        std::memcpy(out, v.data(), v.size());
    }
    return (int64_t)v.size();
}
sf_bricks* sf_bricks_load(const uint8_t* src, int64_t n) {
    if (!src || n < 0) { sf::set_error(SF_E_ARG, "sf_bricks_load: null input"); return nullptr; }
    sf_bricks* b = new sf_bricks();
    const char* why = "";
    if (!load_store(src, (size_t)n, b->s, &why)) {
        delete b;
        sf::set_error(SF_E_ARG, std::string("sf_bricks_load: ") + why);
        return nullptr;
    }
    return b;
}
int sf_bricks_set_cell(sf_bricks* b, int i, int body, int damage, int hull) {
    if (!b || i < 0 || (size_t)i >= b->s.size() || damage < 0 || damage > 255 || hull < 0 || hull > 255) return SF_E_ARG;
    b->s.body[i] = body < 0 ? NONE : (uint32_t)body;
    b->s.damage[i] = (uint8_t)damage;
    b->s.hull[i] = (uint8_t)hull;
    return 0;
}
int sf_bricks_box(const sf_bricks* b, int i, int32_t* out6) {
    if (!b || !out6 || i < 0 || (size_t)i >= b->s.size()) return SF_E_ARG;
    for (int k = 0; k < 3; k++) { out6[k] = b->s.lo[3 * (size_t)i + k]; out6[3 + k] = b->s.ext[3 * (size_t)i + k]; }
    return 0;
}
int sf_bricks_cell(const sf_bricks* b, int i, int32_t* out3) {
    if (!b || !out3 || i < 0 || (size_t)i >= b->s.size()) return SF_E_ARG;
    out3[0] = b->s.body[i] == NONE ? -1 : (int32_t)b->s.body[i];
    out3[1] = b->s.damage[i];
    out3[2] = b->s.hull[i];
    return 0;
}
void sf_bricks_destroy(sf_bricks* b) { delete b; }
int sf_bricks_len(const sf_bricks* b) { return b ? (int)b->s.size() : SF_E_ARG; }
int sf_bricks_add(sf_bricks* b, const int32_t* min3, const int32_t* size3, int mat, int paint, int structure) {
    if (!b || !min3 || !size3) { sf::set_error(SF_E_ARG, "sf_bricks_add: null argument"); return SF_E_ARG; }
    for (int k = 0; k < 3; k++)
        if (size3[k] < 1 || size3[k] > 255 || min3[k] < -32768 || min3[k] > 32767) {
            sf::set_error(SF_E_ARG, "sf_bricks_add: sizes 1..255 ticks, corners within 16 bits");
            return SF_E_ARG;
        }
    if (mat < 0 || mat > 4) { sf::set_error(SF_E_ARG, "sf_bricks_add: material 0..4"); return SF_E_ARG; }
    Store& S = b->s;
    for (int k = 0; k < 3; k++) { S.lo.push_back(min3[k]); S.ext.push_back((uint8_t)size3[k]); }
    S.mat.push_back((uint8_t)mat); S.paint.push_back((uint8_t)paint); S.flags.push_back(F_ALIVE); S.sid.push_back((uint16_t)structure);
    S.inc.emplace_back();
    S.body.push_back(NONE); S.damage.push_back(0); S.hull.push_back(0);
    return (int)S.size() - 1;
}
int sf_bricks_set_flags(sf_bricks* b, int i, int flags) {
    if (!b || i < 0 || (size_t)i >= b->s.size()) return SF_E_ARG;
    b->s.flags[i] = (uint8_t)flags;
    return 0;
}
int sf_bricks_flags(const sf_bricks* b, int i) { return (!b || i < 0 || (size_t)i >= b->s.size()) ? SF_E_ARG : b->s.flags[i]; }
float sf_bricks_mass(const sf_bricks* b, int i) {
    if (!b || i < 0 || (size_t)i >= b->s.size()) return -1.0f;
    const Store& S = b->s;
    float share = S.mat[i] == STEEL ? 0.16f : 1.0f;
    int vol = S.ext[3 * i] * S.ext[3 * i + 1] * S.ext[3 * i + 2];
    return (float)vol * TKF * TKF * TKF * mat_density(S.mat[i]) * 1.0f * share;
}
int sf_bricks_remove(sf_bricks* b, int i) {
    if (!b || i < 0 || (size_t)i >= b->s.size()) return SF_E_ARG;
    Store& S = b->s;
    for (uint32_t j : S.inc[i]) {
        uint32_t o = S.ja[j] == (uint32_t)i ? S.jb[j] : S.ja[j];
        auto& v = S.inc[o];
        v.erase(std::find(v.begin(), v.end(), j));
        S.ja[j] = NONE; S.jb[j] = NONE;
        S.jfree.push_back(j);
        S.live--;
    }
    S.inc[i].clear();
    S.flags[i] &= (uint8_t)~F_ALIVE;
    return 0;
}
int sf_bricks_build_joints(sf_bricks* b) { return b ? build_joints(b->s) : SF_E_ARG; }
int sf_bricks_joints(const sf_bricks* b, int32_t* out5, int cap) {
    if (!b) return SF_E_ARG;
    const Store& S = b->s;
    int n = (int)S.ja.size();
    for (int j = 0; j < n && j < cap && out5; j++) {
        out5[5 * j] = S.ja[j] == NONE ? -1 : (int32_t)S.ja[j];
        out5[5 * j + 1] = S.jb[j] == NONE ? -1 : (int32_t)S.jb[j];
        out5[5 * j + 2] = S.jaxis[j]; out5[5 * j + 3] = S.jarea[j]; out5[5 * j + 4] = S.jdmg[j];
    }
    return n;
}
int sf_bricks_set_joint_damage(sf_bricks* b, int j, int dmg) {
    if (!b || j < 0 || (size_t)j >= b->s.ja.size() || dmg < 0 || dmg > 255) return SF_E_ARG;
    b->s.jdmg[j] = (uint8_t)dmg;
    return 0;
}

sf_warm* sf_warm_create(void) { return new sf_warm(); }
void sf_warm_destroy(sf_warm* w) { delete w; }
int sf_warm_size(const sf_warm* w) { return w ? (int)w->m.size() : SF_E_ARG; }

sf_net* sf_net_create(void) { return new sf_net(); }
void sf_net_destroy(sf_net* N) { delete N; }
int sf_net_clear(sf_net* N) {
    if (!N) return SF_E_ARG;
    N->cases.clear();
    N->stores.clear();
    return 0;
}
int sf_net_set_route(sf_net* N, int min_cells, int max_cases, int block_its) {
    if (!N || max_cases < 0 || block_its < 1 || block_its > 64) return SF_E_ARG;
    N->route_min_cells = min_cells;
    N->route_max_cases = max_cases;
    N->block_its = block_its;
    return 0;
}
int sf_net_set_host_threads(sf_net* N, int n) {
    if (!N || n < 0) return SF_E_ARG;
    N->host_threads = n;
    return 0;
}
int sf_net_add(sf_net* N, const sf_bricks* b, int n, const uint32_t* set, const float* load, const uint8_t* sink,
               const float* dir3, int npoint, const float* point4, sf_warm* warm) {
    if (!N || !b || n < 0 || (n && (!set || !load || !sink)) || !dir3 || npoint < 0 || (npoint && !point4)) {
        sf::set_error(SF_E_ARG, "sf_net_add: bad argument");
        return SF_E_ARG;
    }
    const Store& S = b->s;
    std::vector<uint8_t> seen(S.size(), 0);
    for (int k = 0; k < n; k++) {
        if (set[k] >= S.size() || !(S.flags[set[k]] & F_ALIVE) || seen[set[k]] || !std::isfinite(load[k])) {
            sf::set_error(SF_E_ARG, "sf_net_add: the set must list live bricks once each, with finite loads");
            return SF_E_ARG;
        }
        seen[set[k]] = 1;
    }
    for (int d = 0; d < 3; d++) if (!std::isfinite(dir3[d])) { sf::set_error(SF_E_ARG, "sf_net_add: direction not finite"); return SF_E_ARG; }
    Case C;
    C.set.assign(set, set + n);
    C.load.assign(load, load + n);
    C.sink.assign(sink, sink + n);
    for (int d = 0; d < 3; d++) C.dir[d] = dir3[d];
    C.points.assign(point4, point4 + 4 * (size_t)npoint);
    C.warm = warm;
    N->cases.push_back(std::move(C));
    N->stores.push_back(&S);
    return (int)N->cases.size() - 1;
}
int sf_net_solve(sf_net* N) {
    if (!N) return SF_E_ARG;
    if (N->cases.empty()) return 0;
    auto t0 = std::chrono::steady_clock::now();
    int nc = (int)N->cases.size();
    int threads = N->host_threads > 0 ? N->host_threads : (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::array<double, 3>> part(nc);
    parallel_for(nc, threads, [&](int c) {
        Case& C = N->cases[c];
        auto t = std::chrono::steady_clock::now();
        build_graph(*N->stores[c], C);
        part[c][0] = since(t);
        t = std::chrono::steady_clock::now();
        build_levels(C);
        part[c][1] = since(t);
        t = std::chrono::steady_clock::now();
        build_rest(*N->stores[c], C);
        part[c][2] = since(t);
    });
    for (int q = 0; q < 4; q++) N->host_ms[q] = 0.0;
    for (int c = 0; c < nc; c++) for (int q = 0; q < 3; q++) N->host_ms[q] += part[c][q];
    auto tp = std::chrono::steady_clock::now();
    size_t ncheck = 0;
    for (auto& C : N->cases) ncheck += C.checks.size();
    // the data (uploaded), then the case and check records (uploaded), then the work vectors
    std::vector<CaseOffsets> O(nc);
    Arena A;
    size_t o_cases = 0, o_checks = 0, up = 0;
    for (int pass = 0; pass < 2; pass++) {
        A.size = 0;
        A.measure = pass == 0;
        for (int c = 0; c < nc; c++) pack_data(N->cases[c], A, O[c]);
        o_cases = A.reserve(sizeof(DCase) * nc);
        o_checks = A.reserve(sizeof(DCheck) * std::max<size_t>(ncheck, 1));
        up = A.size;
        if (pass == 0) A.host.resize(up);
    }
    for (int c = 0; c < nc; c++) pack_work(N->cases[c], A, O[c]);
    size_t o_out = A.reserve(3 * sizeof(float) * std::max<size_t>(ncheck, 1));
    N->host_ms[3] = since(tp);
    if (ensure_device(N, A.size)) return SF_E_HIP;
    uint8_t* D = N->dev;
    DCase* hc = at<DCase>(A.host.data(), o_cases);
    DCheck* hk = at<DCheck>(A.host.data(), o_checks);
    size_t kk = 0, dyn = 0;
    for (int c = 0; c < nc; c++) {
        const Case& C = N->cases[c];
        const CaseOffsets& P = O[c];
        DCase& d = hc[c];
        std::memset(&d, 0, sizeof(d));
        d.nl = (int)C.lv.size(); d.n0 = (int)C.cbrick.size();
        d.sp = at<double>(D, P.sp); d.ends = at<int32_t>(D, P.ends); d.f = at<double>(D, P.f);
        d.x0 = C.has_x0 ? at<double>(D, P.x0) : nullptr;
        d.x = at<double>(D, P.x); d.p = at<double>(D, P.p); d.ap = at<double>(D, P.ap); d.out_it = at<int32_t>(D, P.it);
        for (int l = 0; l < d.nl; l++) {
            DLevel& L = d.lv[l];
            L.n = C.lv[l].n;
            L.nc = l == 0 ? (int)C.la.size() : (int)C.lv[l].oa.size();
            L.inv = at<double>(D, P.inv[l]);
            L.ioff = at<int32_t>(D, P.ioff[l]); L.cpos = at<int32_t>(D, P.cpos[l]);
            L.r = at<double>(D, P.r[l]); L.z = at<double>(D, P.z[l]); L.kz = at<double>(D, P.kz[l]); L.buf = at<double>(D, P.buf);
            if (l == 0) { L.reg = at<double>(D, P.reg0); L.fixed = at<uint8_t>(D, P.fixed0); continue; }
            L.diag = at<double>(D, P.diag[l]); L.om = at<double>(D, P.om[l]); L.oends = at<int32_t>(D, P.oends[l]);
            L.choff = at<int32_t>(D, P.choff[l]); L.ch = at<int32_t>(D, P.ch[l]);
            L.upf = at<int32_t>(D, P.upf[l]); L.dupf = at<double>(D, P.dupf[l]);
        }
        if (d.nl >= 2) {
            size_t b = coarse_lds_bytes(d.lv[d.nl - 1].n, d.lv[d.nl - 1].nc);
            if (b <= N->lds_budget) { d.lds = (int)b; dyn = std::max(dyn, b); }
        }
        d.route = (N->route_min_cells >= 0 && nc <= N->route_max_cases && d.n0 >= N->route_min_cells && d.nl >= 2) ? 1 : 0;
        for (const Check& K : C.checks) {
            DCheck& q = hk[kk++];
            q.cs = c; q.kind = K.kind; q.axis = K.axis; q.hinge = K.hinge; q.l0 = K.l0; q.l1 = K.l1;
            for (int z = 0; z < 3; z++) { q.cap[z] = K.cap[z]; q.c[z] = K.c[z]; }
            q.flex = K.flex;
            const double sv[7] = {K.sec.area, K.sec.iu, K.sec.iw, K.sec.zu, K.sec.zw, K.sec.j, K.sec.rmax};
            for (int z = 0; z < 7; z++) q.sec[z] = sv[z];
            q.glinks = at<int32_t>(D, P.glinks);
        }
    }
    N->ms[0] = since(t0);
    t0 = std::chrono::steady_clock::now();
    SF_HIP(hipMemcpyAsync(D, A.host.data(), up, hipMemcpyHostToDevice, N->st));
    SF_HIP(hipEventRecord(N->ev[0], N->st));
    hipLaunchKernelGGL(k_net_solve, dim3(nc), dim3(NT_SOLVE), dyn, N->st, at<DCase>(D, o_cases));
    SF_HIP(hipGetLastError());
    for (int c = 0; c < nc; c++)
        if (hc[c].route && mk_solve(N, hc[c], at<DCase>(D, o_cases) + c, at<NetState>(D, O[c].state))) return SF_E_HIP;
    SF_HIP(hipEventRecord(N->ev[1], N->st));
    float* out = at<float>(D, o_out);
    if (ncheck) {
        hipLaunchKernelGGL(k_net_checks, dim3(sf_grid(ncheck)), dim3(256), 0, N->st, at<DCheck>(D, o_checks), (int)ncheck,
                           at<DCase>(D, o_cases), out, out + ncheck, out + 2 * ncheck);
        SF_HIP(hipGetLastError());
    }
    SF_HIP(hipEventRecord(N->ev[2], N->st));
    std::vector<float> res(3 * std::max<size_t>(ncheck, 1));
    SF_HIP(hipMemcpyAsync(res.data(), out, sizeof(float) * 3 * ncheck, hipMemcpyDeviceToHost, N->st));
    for (int c = 0; c < nc; c++) {
        Case& C = N->cases[c];
        C.x.resize(6 * C.cbrick.size());
        SF_HIP(hipMemcpyAsync(C.x.data(), at<double>(D, O[c].x), sizeof(double) * C.x.size(), hipMemcpyDeviceToHost, N->st));
        SF_HIP(hipMemcpyAsync(&C.iterations, at<int32_t>(D, O[c].it), sizeof(int32_t), hipMemcpyDeviceToHost, N->st));
    }
    SF_HIP(hipStreamSynchronize(N->st));
    float e1 = 0.0f, e2 = 0.0f;
    SF_HIP(hipEventElapsedTime(&e1, N->ev[0], N->ev[1]));
    SF_HIP(hipEventElapsedTime(&e2, N->ev[1], N->ev[2]));
    N->ms[2] = e1; N->ms[3] = e2;
    N->ms[1] = since(t0) - e1 - e2;
    kk = 0;
    for (int c = 0; c < nc; c++) {
        Case& C = N->cases[c];
        size_t m = C.checks.size();
        C.ratio.assign(res.begin() + kk, res.begin() + kk + m);
        C.flow.assign(res.begin() + ncheck + kk, res.begin() + ncheck + kk + m);
        C.moment.assign(res.begin() + 2 * ncheck + kk, res.begin() + 2 * ncheck + kk + m);
        kk += m;
        float mr = 0.0f;
        for (float r : C.ratio) if (r > mr) mr = r;
        C.max_ratio = mr;
        if (C.warm) {
            C.warm->m.clear();
            const Level& F = C.lv[0];
            for (size_t k = 0; k < C.cbrick.size(); k++) {
                if (F.fixed[k]) continue;
                std::array<double, 6> v;
                for (int q = 0; q < 6; q++) v[q] = C.x[6 * k + q];
                C.warm->m[warm_key(C.cbrick[k], C.clo[3 * k], C.clo[3 * k + 2])] = v;
            }
        }
    }
    return 0;
}
int sf_net_count(const sf_net* N) { return N ? (int)N->cases.size() : SF_E_ARG; }
// [max_ratio, iterations, cells, links, levels, worst-in ratio, brick, axis, flow, moment^2, joints checked, inner faces]
int sf_net_result(const sf_net* N, int c, double* out12) {
    if (!N || c < 0 || (size_t)c >= N->cases.size() || !out12) return SF_E_ARG;
    const Case& C = N->cases[c];
    for (int q = 0; q < 12; q++) out12[q] = 0.0;
    out12[0] = C.max_ratio; out12[1] = C.iterations; out12[2] = (double)C.cbrick.size(); out12[3] = (double)C.la.size();
    out12[4] = (double)C.lv.size();
    float worst = 0.0f;
    for (size_t k = C.njoint; k < C.checks.size(); k++) {
        if (!(C.ratio[k] > worst)) continue;
        worst = C.ratio[k];
        int l = C.checks[k].l0;
        out12[5] = worst; out12[6] = C.cbrick[C.la[l]]; out12[7] = C.lax[l]; out12[8] = C.flow[k]; out12[9] = C.moment[k];
    }
    out12[10] = C.njoint; out12[11] = (double)(C.checks.size() - C.njoint);
    return 0;
}
int sf_net_levels(const sf_net* N, int c, int32_t* out8) {
    if (!N || c < 0 || (size_t)c >= N->cases.size() || !out8) return SF_E_ARG;
    const Case& C = N->cases[c];
    for (int q = 0; q < 8; q++) out8[q] = q < (int)C.lv.size() ? C.lv[q].n : 0;
    return (int)C.lv.size();
}
int sf_net_joints(const sf_net* N, int c, int32_t* jid, float* ratio, int cap) {
    if (!N || c < 0 || (size_t)c >= N->cases.size()) return SF_E_ARG;
    const Case& C = N->cases[c];
    for (int k = 0; k < C.njoint && k < cap; k++) {
        if (jid) jid[k] = (int32_t)C.lj[C.glinks[C.checks[k].l0]];
        if (ratio) ratio[k] = C.ratio[k];
    }
    return C.njoint;
}
int sf_net_inner(const sf_net* N, int c, int32_t* bat3, float* ratio, int cap) {
    if (!N || c < 0 || (size_t)c >= N->cases.size()) return SF_E_ARG;
    const Case& C = N->cases[c];
    int n = (int)C.checks.size() - C.njoint;
    for (int k = 0; k < n && k < cap; k++) {
        int l = C.checks[C.njoint + k].l0;
        if (bat3) { bat3[3 * k] = (int32_t)C.cbrick[C.la[l]]; bat3[3 * k + 1] = C.lax[l]; bat3[3 * k + 2] = C.flo[3 * l + C.lax[l]]; }
        if (ratio) ratio[k] = C.ratio[C.njoint + k];
    }
    return n;
}
int sf_net_cells(const sf_net* N, int c, int32_t* cell7, double* x6, int cap) {
    if (!N || c < 0 || (size_t)c >= N->cases.size()) return SF_E_ARG;
    const Case& C = N->cases[c];
    int n = (int)C.cbrick.size();
    for (int k = 0; k < n && k < cap; k++) {
        if (cell7) {
            cell7[7 * k] = (int32_t)C.cbrick[k];
            for (int d = 0; d < 3; d++) { cell7[7 * k + 1 + d] = C.clo[3 * k + d]; cell7[7 * k + 4 + d] = C.chi[3 * k + d]; }
        }
        if (x6) for (int q = 0; q < 6; q++) x6[6 * k + q] = C.x[6 * k + q];
    }
    return n;
}
// [host build ms, upload and read-back ms, device solve ms, device checks ms] of the last sf_net_solve
int sf_net_ms(const sf_net* N, double* out4) {
    if (!N || !out4) return SF_E_ARG;
    for (int q = 0; q < 4; q++) out4[q] = N->ms[q];
    return 0;
}
// the host build of the last sf_net_solve, summed over its cases (thread ms): cells, links and springs; levels; loads,
// warm starts and checks; packing into the arena (wall ms)
int sf_net_host_ms(const sf_net* N, double* out4) {
    if (!N || !out4) return SF_E_ARG;
    for (int q = 0; q < 4; q++) out4[q] = N->host_ms[q];
    return 0;
}

}  // extern "C"
