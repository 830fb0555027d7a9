// Seeded scene generators (host only): the reference's own test scenes restated exactly, the compound-body
// builder, and the synthetic workloads (rubble pile, frame/wall buildings, the brick tower).
#include "../../include/sf.h"
#include "sf_host.h"
#include "hulls.h"
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include <array>

namespace {

// ---- the engine's own f32 helpers, for scenes that must equal the reference's bit for bit (engine/src/math.rs)
inline float e_floor(float x) { float t = (float)(int32_t)x; return t > x ? t - 1.0f : t; }
inline float e_sqrt(float x) { return x <= 0.0f ? 0.0f : std::sqrt(x); }
inline float e_sin(float x) {
    const float tau = 6.28318530717958647692f, pi = 3.14159265358979323846f;
    float a = x - tau * e_floor(x / tau + 0.5f);
    if (a > pi * 0.5f) a = pi - a; else if (a < -pi * 0.5f) a = -pi - a;
    float a2 = a * a;
    return a * (1.0f - a2 / 6.0f * (1.0f - a2 / 20.0f * (1.0f - a2 / 42.0f * (1.0f - a2 / 72.0f))));
}
inline float e_cos(float x) { return e_sin(x + 1.57079632679489661923f); }
// Quat::from_axis_angle (math.rs:35)
void axis_angle(float ax, float ay, float az, float ang, float* q) {
    float l = e_sqrt(ax * ax + ay * ay + az * az);
    if (l > 1e-12f) { float k = 1.0f / l; ax *= k; ay *= k; az *= k; } else { ax = ay = az = 0.0f; }
    float s = e_sin(ang * 0.5f), c = e_cos(ang * 0.5f);
    q[0] = ax * s; q[1] = ay * s; q[2] = az * s; q[3] = c;
}
// World::add_box (phys.rs:139-148): mass and local inverse inertia of a solid box of density rho
void box_record(const float* p, const float* q, const float* h, float rho, float mu, float* body, float* shape) {
    std::memset(body, 0, sizeof(float) * SF_BODY_F);
    std::memset(shape, 0, sizeof(float) * SF_SHAPE_F);
    float m = rho * 8.0f * h[0] * h[1] * h[2];
    float im = 0.0f, ii[3] = {0.0f, 0.0f, 0.0f};
    if (m > 0.0f) {
        float x2 = 4.0f * h[0] * h[0], y2 = 4.0f * h[1] * h[1], z2 = 4.0f * h[2] * h[2];
        im = 1.0f / m;
        ii[0] = 12.0f / (m * (y2 + z2)); ii[1] = 12.0f / (m * (x2 + z2)); ii[2] = 12.0f / (m * (x2 + y2));
    }
    for (int k = 0; k < 3; k++) body[k] = p[k];
    for (int k = 0; k < 4; k++) body[3 + k] = q[k];
    body[13] = im; body[14] = ii[0]; body[15] = ii[1]; body[16] = ii[2];
    body[17] = m > 0.0f ? 1.0f : 0.0f;
    shape[6] = 1.0f;  // identity local rotation
    shape[7] = h[0]; shape[8] = h[1]; shape[9] = h[2]; shape[10] = mu;
}

// ---- seeded randomness for the synthetic workloads
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0xD1B54A32D192ED03ull) {}
    uint64_t next() { uint64_t z = (s += 0x9E3779B97F4A7C15ull); z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull; z = (z ^ (z >> 27)) * 0x94D049BB133111EBull; return z ^ (z >> 31); }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    uint32_t below(uint32_t n) { return (uint32_t)((next() >> 32) * (uint64_t)n >> 32); }
    double normal() { double u = uni(), v = uni(); if (u < 1e-300) u = 1e-300; return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v); }
};

// material table of the GPU engine (gpu/gpu_world.js:15-25): density, friction, restitution; codes 1..5 as in
// src/lib.rs, 0 the ground
const double DENSITY[6] = {0.0, 2400.0, 1900.0, 7850.0, 2500.0, 500.0};
const float FRICTION[6] = {0.6f, 0.6f, 0.6f, 0.45f, 0.5f, 0.5f};
const float RESTITUTION[6] = {0.0f, 0.1f, 0.05f, 0.25f, 0.2f, 0.3f};

// cyclic Jacobi eigen-decomposition of a symmetric 3x3 (columns of V are the eigenvectors)
void eig_sym(double A[3][3], double vals[3], double V[3][3]) {
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        if (off < 1e-30) break;
        for (int p = 0; p < 2; p++) for (int q = p + 1; q < 3; q++) {
            if (std::fabs(A[p][q]) < 1e-300) continue;
            double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
            double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
            double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
            for (int k = 0; k < 3; k++) { double akp = A[k][p], akq = A[k][q]; A[k][p] = c * akp - s * akq; A[k][q] = s * akp + c * akq; }
            for (int k = 0; k < 3; k++) { double apk = A[p][k], aqk = A[q][k]; A[p][k] = c * apk - s * aqk; A[q][k] = s * apk + c * aqk; }
            for (int k = 0; k < 3; k++) { double vkp = V[k][p], vkq = V[k][q]; V[k][p] = c * vkp - s * vkq; V[k][q] = s * vkp + c * vkq; }
        }
    }
    for (int i = 0; i < 3; i++) vals[i] = A[i][i];
    // right-handed frame
    double det = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) - V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) + V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
    if (det < 0) for (int k = 0; k < 3; k++) V[k][2] = -V[k][2];
}
// rotation matrix (columns) to quaternion (x, y, z, w)
void m2q(const double R[3][3], double q[4]) {
    double tr = R[0][0] + R[1][1] + R[2][2];
    if (tr > 0) { double s = std::sqrt(tr + 1.0) * 2; q[3] = 0.25 * s; q[0] = (R[2][1] - R[1][2]) / s; q[1] = (R[0][2] - R[2][0]) / s; q[2] = (R[1][0] - R[0][1]) / s; }
    else if (R[0][0] > R[1][1] && R[0][0] > R[2][2]) { double s = std::sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]) * 2; q[3] = (R[2][1] - R[1][2]) / s; q[0] = 0.25 * s; q[1] = (R[0][1] + R[1][0]) / s; q[2] = (R[0][2] + R[2][0]) / s; }
    else if (R[1][1] > R[2][2]) { double s = std::sqrt(1.0 + R[1][1] - R[0][0] - R[2][2]) * 2; q[3] = (R[0][2] - R[2][0]) / s; q[0] = (R[0][1] + R[1][0]) / s; q[1] = 0.25 * s; q[2] = (R[1][2] + R[2][1]) / s; }
    else { double s = std::sqrt(1.0 + R[2][2] - R[0][0] - R[1][1]) * 2; q[3] = (R[1][0] - R[0][1]) / s; q[0] = (R[0][2] + R[2][0]) / s; q[1] = (R[1][2] + R[2][1]) / s; q[2] = 0.25 * s; }
    double l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int k = 0; k < 4; k++) q[k] /= l;
}

// Compound body of 0.5 m bricks at lattice cells (gpu/gpu_world.js buildBodies :2212-2249, FP64): total mass,
// centre of mass, inertia of each brick about its centre (m (2h)^2 / 6 per axis) moved to the centre of mass,
// principal frame by Jacobi rotations; shapes at their offsets in that frame.
void build_compound(int nb, const int32_t* xyz, const uint8_t* mat, const double origin[3], float* body, float* shapes) {
    const double S = 0.5, h = 0.25;
    double M = 0, com[3] = {0, 0, 0};
    std::vector<double> ms(nb);
    for (int k = 0; k < nb; k++) {
        ms[k] = DENSITY[mat[k]] * 8.0 * h * h * h;
        M += ms[k];
        for (int a = 0; a < 3; a++) com[a] += ms[k] * (origin[a] + (xyz[3 * k + a] + 0.5) * S);
    }
    for (int a = 0; a < 3; a++) com[a] /= M;
    double I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < nb; k++) {
        double d[3];
        for (int a = 0; a < 3; a++) d[a] = origin[a] + (xyz[3 * k + a] + 0.5) * S - com[a];
        double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double own = ms[k] * (2.0 * h) * (2.0 * h) / 6.0;
        for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) I[r][c] += (r == c ? own : 0.0) + ms[k] * ((r == c ? dd : 0.0) - d[r] * d[c]);
    }
    double vals[3], V[3][3];
    eig_sym(I, vals, V);
    double q[4];
    m2q(V, q);
    std::memset(body, 0, sizeof(float) * SF_BODY_F);
    for (int a = 0; a < 3; a++) body[a] = (float)com[a];
    for (int a = 0; a < 4; a++) body[3 + a] = (float)q[a];
    body[13] = (float)(1.0 / M);
    for (int a = 0; a < 3; a++) body[14 + a] = (float)(vals[a] > 1e-12 ? 1.0 / vals[a] : 0.0);
    body[17] = 1.0f;
    // shapes: local position V^T (p - com), local rotation conj(q) (the lattice is axis aligned in the world)
    for (int k = 0; k < nb; k++) {
        float* s = shapes + (size_t)k * SF_SHAPE_F;
        double d[3];
        for (int a = 0; a < 3; a++) d[a] = origin[a] + (xyz[3 * k + a] + 0.5) * S - com[a];
        for (int a = 0; a < 3; a++) s[a] = (float)(V[0][a] * d[0] + V[1][a] * d[1] + V[2][a] * d[2]);
        s[3] = (float)-q[0]; s[4] = (float)-q[1]; s[5] = (float)-q[2]; s[6] = (float)q[3];
        s[7] = s[8] = s[9] = (float)h;
        s[10] = FRICTION[mat[k]];
        s[11] = RESTITUTION[mat[k]];
        s[12] = 0.0f;  // (a whole box)
    }
}


// ---- the site-scale scene: a 320 x 320 m site (lattice indices 0..639 around 320) of 37 structures, one per 40 m
// lot: 16 frame towers of 8-16 storeys (3.5 m each), 12 three-storey masonry row blocks with concrete floors, a 56 m
// single-brick-ring masonry chimney, an arched masonry viaduct and 7 steel-frame halls with glass walls.
struct SiteStruct { std::vector<uint8_t> mat, infill; std::vector<int32_t> xyz; };
void site_add(SiteStruct& S, int x, int y, int z, uint8_t m, uint8_t inf = 0) {
    S.xyz.push_back(x); S.xyz.push_back(y); S.xyz.push_back(z); S.mat.push_back(m); S.infill.push_back(inf);
}
std::vector<SiteStruct> make_site(uint32_t seed) {
    Rng rng(seed);
    std::vector<std::array<int, 2>> lots;
    for (int cx = 40; cx < 600; cx += 80) for (int cz = 40; cz < 600; cz += 80) lots.push_back({cx, cz});
    for (size_t k = lots.size(); k > 1; k--) std::swap(lots[k - 1], lots[rng.below((uint32_t)k)]);  // Fisher-Yates
    std::vector<SiteStruct> out;
    size_t lot = 0;
    for (int t = 0; t < 16; t++) {  // frame towers
        const int cx = lots[lot][0], cz = lots[lot][1]; lot++;
        const int w = 22 + (int)rng.below(10), storeys = 8 + (int)rng.below(9);
        const uint32_t bseed = 1 + (uint32_t)(rng.next() % 0x3FFFFFFF);
        const int n = sf_scene_building(w, w, storeys, 6, 0, bseed, nullptr, nullptr, nullptr);
        std::vector<uint8_t> m(n), inf(n);
        std::vector<int32_t> xyz(3 * (size_t)n);
        sf_scene_building(w, w, storeys, 6, 0, bseed, m.data(), xyz.data(), inf.data());
        SiteStruct S;
        for (int i = 0; i < n; i++) site_add(S, xyz[3 * i] + cx - w / 2, xyz[3 * i + 1], xyz[3 * i + 2] + cz - w / 2, m[i], inf[i]);
        out.push_back(std::move(S));
    }
    for (int t = 0; t < 12; t++) {  // row blocks: per storey a 6-layer masonry wall ring and a concrete slab
        const int cx = lots[lot][0], cz = lots[lot][1]; lot++;
        const int lx = 30 + (int)rng.below(18), lz = 8 + (int)rng.below(4);
        const int x0 = cx - lx / 2, z0 = cz - lz / 2;
        SiteStruct S;
        for (int st = 0; st < 3; st++) {
            for (int y = 7 * st; y < 7 * st + 6; y++)
                for (int x = 0; x < lx; x++) for (int z = 0; z < lz; z++)
                    if (x == 0 || x == lx - 1 || z == 0 || z == lz - 1) site_add(S, x0 + x, y, z0 + z, 2);
            for (int x = 0; x < lx; x++) for (int z = 0; z < lz; z++) site_add(S, x0 + x, 7 * st + 6, z0 + z, 1);
        }
        out.push_back(std::move(S));
    }
    {  // the chimney: a ring of radius 6 bricks, 112 bricks (56 m) tall
        const int cx = lots[lot][0], cz = lots[lot][1]; lot++;
        std::vector<std::array<int, 2>> ring;
        for (int k = 0; k < 400; k++) {
            const double a = 2.0 * 3.14159265358979 * k / 399.0;
            std::array<int, 2> c{(int)std::lround(cx + 6.0 * std::cos(a)), (int)std::lround(cz + 6.0 * std::sin(a))};
            if (std::find(ring.begin(), ring.end(), c) == ring.end()) ring.push_back(c);
        }
        std::sort(ring.begin(), ring.end());
        SiteStruct S;
        for (int y = 0; y < 112; y++) for (auto& c : ring) site_add(S, c[0], y, c[1], 2);
        out.push_back(std::move(S));
    }
    {  // the viaduct: 9 masonry piers of 2 x 4 bricks, 20 tall, 8 arches between them and a deck, along x
        const int cx = lots[lot][0], cz = lots[lot][1]; lot++;
        const int span = 8, x0 = cx - 4 * span;
        std::vector<std::array<int, 3>> cells;
        for (int p = 0; p < 9; p++) for (int y = 0; y < 20; y++) for (int dx = 0; dx < 2; dx++) for (int dz = 0; dz < 4; dz++)
            cells.push_back({x0 + p * span + dx, y, cz + dz});
        for (int a = 0; a < 8; a++) {
            const int ax0 = x0 + a * span + 2;
            for (int x = ax0; x < ax0 + span - 2; x++) {
                const double u = (x - ax0 + 0.5) / (span - 2);
                const int ytop = 14 + (int)std::lround(5.0 * std::sin(3.14159265358979 * u));
                for (int y = ytop; y < 20; y++) for (int dz = 0; dz < 4; dz++) cells.push_back({x, y, cz + dz});
            }
        }
        for (int x = x0; x < x0 + 8 * span + 2; x++) for (int dz = 0; dz < 4; dz++) cells.push_back({x, 20, cz + dz});
        std::sort(cells.begin(), cells.end());
        cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
        SiteStruct S;
        for (auto& c : cells) site_add(S, c[0], c[1], c[2], 2);
        out.push_back(std::move(S));
    }
    for (int t = 0; t < 7; t++) {  // steel halls: columns every 6 bricks and at the far edges, a roof grid, glass walls
        const int cx = lots[lot][0], cz = lots[lot][1]; lot++;
        const int w = 20 + (int)rng.below(10), h = 14;
        SiteStruct S;
        for (int x = 0; x < w; x++) for (int z = 0; z < w; z++) {
            const bool col = (x % 6 == 0 || x == w - 1) && (z % 6 == 0 || z == w - 1);
            const bool per = x == 0 || x == w - 1 || z == 0 || z == w - 1;
            for (int y = 0; y < h; y++) {
                if (col) site_add(S, cx - w / 2 + x, y, cz - w / 2 + z, 3);
                else if (per && y < h - 1) site_add(S, cx - w / 2 + x, y, cz - w / 2 + z, 4, 1);
            }
            if (x % 3 == 0 || z % 3 == 0) site_add(S, cx - w / 2 + x, h, cz - w / 2 + z, 3);
        }
        out.push_back(std::move(S));
    }
    return out;
}
}  // namespace

extern "C" {

int sf_scene_pile(int n, float* bodies, int32_t* shape_body, float* shapes, int* nshape) {
    if (n < 0) { sf::set_error(SF_E_ARG, "n < 0"); return SF_E_ARG; }
    if (nshape) *nshape = n;
    if (!bodies) return n;
    // phys.rs:632-646
    int side = (int)e_sqrt((float)n / 20.0f); if (side < 1) side = 1;
    const float h[3] = {0.25f, 0.125f, 0.125f};
    uint32_t seed = 1u;
    auto rnd = [&]() { seed = seed * 1664525u + 1013904223u; return (float)(seed >> 8) / 16777216.0f; };
    int k = 0;
    for (int layer = 0; k < n; layer++) {
        for (int i = 0; i < side && k < n; i++) for (int j = 0; j < side && k < n; j++) {
            float px = (float)i * 0.62f + rnd() * 0.05f;
            float py = 0.2f + (float)layer * 0.3f;
            float pz = (float)j * 0.36f + rnd() * 0.05f;
            float ax = rnd() - 0.5f, ay = rnd() - 0.5f, az = rnd() - 0.5f;
            float ang = rnd() * 0.6f;
            float q[4]; axis_angle(ax, ay, az, ang, q);
            float p[3] = {px, py, pz};
            box_record(p, q, h, 1900.0f, 0.6f, bodies + (size_t)k * SF_BODY_F, shapes + (size_t)k * SF_SHAPE_F);
            shape_body[k] = k;
            k++;
        }
    }
    return n;
}

int sf_scene_stack(int n, float* bodies, int32_t* shape_body, float* shapes, int* nshape) {
    if (n < 0) { sf::set_error(SF_E_ARG, "n < 0"); return SF_E_ARG; }
    if (nshape) *nshape = n;
    if (!bodies) return n;
    const float h[3] = {0.25f, 0.25f, 0.25f}, q[4] = {0, 0, 0, 1};
    for (int k = 0; k < n; k++) {
        float p[3] = {0.0f, 0.25f + 0.5f * (float)k, 0.0f};
        box_record(p, q, h, 1900.0f, 0.6f, bodies + (size_t)k * SF_BODY_F, shapes + (size_t)k * SF_SHAPE_F);
        shape_body[k] = k;
    }
    return n;
}

// A body from shapes given in the world (gpu/gpu_world.js buildBodies :2212-2249), FP64: each shape's mass is the
// density times its box volume times its hull's share of the box (none on a fixed body); the centre of mass is taken
// over the shapes' centroids (a chunk's centroid, not its box centre); each shape's inertia about its centroid comes
// from its covariance (a box: h^2 / 3 per axis) stretched by h, turned into the world and moved to the centre of mass;
// the principal frame by Jacobi rotations; the shapes at their offsets in that frame, rotations relative to it.
// in: n records of SF_SHAPE_IN floats (p[3] q[4] h[3] material 1..5, hull id); out: body[SF_BODY_F], n shapes.
int sf_build_body(int n, const float* in, int fixed, const float* v, const float* w, float* body, float* shapes) {
    if (n <= 0 || !in || !body || !shapes) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    const sf::HullLib& HL = sf::hull_lib();
    for (int k = 0; k < n; k++) {
        const float* f = in + (size_t)k * SF_SHAPE_IN;
        const int mat = (int)f[10], hull = (int)f[11];
        if (mat < 1 || mat > 5 || hull < 0 || hull >= (int)HL.cells.size()) { sf::set_error(SF_E_ARG, "material 1..5 and a hull of the library"); return SF_E_ARG; }
        for (int q = 0; q < SF_SHAPE_IN; q++) if (!std::isfinite(f[q])) { sf::set_error(SF_E_ARG, "NaN or inf in the input"); return SF_E_ARG; }
    }
    auto q2m = [](const float* q, double R[3][3]) {
        const double x = q[0], y = q[1], z = q[2], ww = q[3];
        R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - z * ww); R[0][2] = 2 * (x * z + y * ww);
        R[1][0] = 2 * (x * y + z * ww); R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - x * ww);
        R[2][0] = 2 * (x * z - y * ww); R[2][1] = 2 * (y * z + x * ww); R[2][2] = 1 - 2 * (x * x + y * y);
    };
    std::vector<double> ms(n);
    std::vector<std::array<double, 3>> mc(n);
    double M = 0, com[3] = {0, 0, 0};
    for (int k = 0; k < n; k++) {
        const float* f = in + (size_t)k * SF_SHAPE_IN;
        const sf::HullCell& C = HL.cells[(int)f[11]];
        ms[k] = (fixed ? 0.0 : DENSITY[(int)f[10]]) * 8.0 * (double)f[7] * (double)f[8] * (double)f[9] * C.box_vol;
        double R[3][3];
        q2m(f + 3, R);
        const double o[3] = {C.cen[0] * f[7], C.cen[1] * f[8], C.cen[2] * f[9]};
        for (int i = 0; i < 3; i++) mc[k][i] = (double)f[i] + R[i][0] * o[0] + R[i][1] * o[1] + R[i][2] * o[2];
        M += ms[k];
        for (int i = 0; i < 3; i++) com[i] += ms[k] * mc[k][i];
    }
    if (M > 0) { for (int i = 0; i < 3; i++) com[i] /= M; }
    else {
        for (int i = 0; i < 3; i++) com[i] = 0.0;
        for (int k = 0; k < n; k++) for (int i = 0; i < 3; i++) com[i] += (double)in[(size_t)k * SF_SHAPE_IN + i] / n;
    }
    double qb[4] = {0, 0, 0, 1}, il[3] = {0, 0, 0};
    if (M > 0) {
        double I[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int k = 0; k < n; k++) {
            const float* f = in + (size_t)k * SF_SHAPE_IN;
            const sf::HullCell& C = HL.cells[(int)f[11]];
            const double h[3] = {f[7], f[8], f[9]};
            double Cm[3][3], Il[3][3], R[3][3];
            for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) Cm[r][c] = ms[k] * C.cov[r][c] * h[r] * h[c];
            const double tr = Cm[0][0] + Cm[1][1] + Cm[2][2];
            for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) Il[r][c] = (r == c ? tr : 0.0) - Cm[r][c];
            q2m(f + 3, R);
            const double d[3] = {mc[k][0] - com[0], mc[k][1] - com[1], mc[k][2] - com[2]};
            const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) {
                    double t = 0.0;
                    for (int a2 = 0; a2 < 3; a2++) for (int b2 = 0; b2 < 3; b2++) t += R[r][a2] * Il[a2][b2] * R[c][b2];
                    I[r][c] += t + ms[k] * ((r == c ? dd : 0.0) - d[r] * d[c]);
                }
        }
        double vals[3], V[3][3];
        eig_sym(I, vals, V);
        m2q(V, qb);
        for (int i = 0; i < 3; i++) il[i] = vals[i] > 1e-12 ? 1.0 / vals[i] : 0.0;
    }
    float qf[4] = {(float)qb[0], (float)qb[1], (float)qb[2], (float)qb[3]};
    double Rb[3][3];
    q2m(qf, Rb);
    std::memset(body, 0, sizeof(float) * SF_BODY_F);
    for (int i = 0; i < 3; i++) body[i] = (float)com[i];
    for (int i = 0; i < 4; i++) body[3 + i] = qf[i];
    if (v) for (int i = 0; i < 3; i++) body[7 + i] = v[i];
    if (w) for (int i = 0; i < 3; i++) body[10 + i] = w[i];
    body[13] = M > 0 ? (float)(1.0 / M) : 0.0f;
    for (int i = 0; i < 3; i++) body[14 + i] = (float)il[i];
    body[17] = M > 0 ? 1.0f : 0.0f;
    // shapes: local position Rb^T (p - com), local rotation conj(qb) * q
    for (int k = 0; k < n; k++) {
        const float* f = in + (size_t)k * SF_SHAPE_IN;
        float* o = shapes + (size_t)k * SF_SHAPE_F;
        std::memset(o, 0, sizeof(float) * SF_SHAPE_F);
        const double d[3] = {f[0] - com[0], f[1] - com[1], f[2] - com[2]};
        for (int a2 = 0; a2 < 3; a2++) o[a2] = (float)(Rb[0][a2] * d[0] + Rb[1][a2] * d[1] + Rb[2][a2] * d[2]);
        const double ax = -qb[0], ay = -qb[1], az = -qb[2], aw = qb[3], bx = f[3], by = f[4], bz = f[5], bw = f[6];
        o[3] = (float)(aw * bx + ax * bw + ay * bz - az * by);
        o[4] = (float)(aw * by - ax * bz + ay * bw + az * bx);
        o[5] = (float)(aw * bz + ax * by - ay * bx + az * bw);
        o[6] = (float)(aw * bw - ax * bx - ay * by - az * bz);
        o[7] = f[7]; o[8] = f[8]; o[9] = f[9];
        o[10] = FRICTION[(int)f[10]];
        o[11] = RESTITUTION[(int)f[10]];
        o[12] = f[11];
    }
    return 0;
}

// The pieces of a brick (centre p, rotation q, half size h) broken by a pattern of the hull library, one shape record
// (SF_SHAPE_IN) per cell, tiling it (gpu/gpu_world.js GpuWorld.hullShapes :2257-2263): each cell's box centre turned
// with the brick, its half size the cell box's times the brick's. Returns the piece count.
int sf_hull_shapes(const float* p, const float* q, const float* h, int pattern, int mat, float* out) {
    const sf::HullLib& HL = sf::hull_lib();
    if (!p || !q || !h || pattern < 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    const std::vector<int>& P = HL.patterns[pattern % (int)HL.patterns.size()];
    if (!out) return (int)P.size();
    const double x = q[0], y = q[1], z = q[2], ww = q[3];
    const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * ww), 2 * (x * z + y * ww)},
                            {2 * (x * y + z * ww), 1 - 2 * (x * x + z * z), 2 * (y * z - x * ww)},
                            {2 * (x * z - y * ww), 2 * (y * z + x * ww), 1 - 2 * (x * x + y * y)}};
    for (size_t k = 0; k < P.size(); k++) {
        const sf::HullCell& C = HL.cells[P[k]];
        const double o[3] = {C.bc[0] * h[0], C.bc[1] * h[1], C.bc[2] * h[2]};
        float* r = out + k * SF_SHAPE_IN;
        for (int i = 0; i < 3; i++) r[i] = (float)(p[i] + R[i][0] * o[0] + R[i][1] * o[1] + R[i][2] * o[2]);
        for (int i = 0; i < 4; i++) r[3 + i] = q[i];
        for (int i = 0; i < 3; i++) r[7 + i] = (float)(C.bh[i] * h[i]);
        r[10] = (float)mat;
        r[11] = (float)P[k];
    }
    return (int)P.size();
}

int sf_build_compound(int nb, const int32_t* xyz, const uint8_t* mat, float* body, float* shapes) {
    if (nb <= 0 || !xyz || !mat || !body || !shapes) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    for (int k = 0; k < nb; k++) if (mat[k] < 1 || mat[k] > 5) { sf::set_error(SF_E_ARG, "material must be 1..5"); return SF_E_ARG; }
    const double origin[3] = {0, 0, 0};
    build_compound(nb, xyz, mat, origin, body, shapes);
    return 0;
}

// Rubble pile: a partition of a lattice box into polycubes grown together (round robin, one brick per body per
// pass, to a size drawn from a truncated geometric distribution) so the packing reaches the target without the
// jamming of random sequential placement.
int sf_scene_rubble(int nbody, float p, int maxsize, float bx, float by, float bz, float vsig, float wsig, uint32_t seed,
                    float* bodies, int32_t* shape_body, float* shapes, int* nshape) {
    if (nbody <= 0 || p <= 0.0f || p > 1.0f || maxsize < 1 || bx <= 0 || by <= 0 || bz <= 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    const int NX = (int)(bx / 0.5f), NY = (int)(by / 0.5f), NZ = (int)(bz / 0.5f);
    const size_t NC = (size_t)NX * NY * NZ;
    Rng rng(seed);
    std::vector<int> target(nbody);
    for (int b = 0; b < nbody; b++) {
        int k = 1;
        while (k < maxsize && rng.uni() >= (double)p) k++;
        target[b] = k;
    }
    std::vector<int32_t> owner(NC, -1);
    std::vector<std::vector<int32_t>> cells(nbody);
    // seeds at distinct random cells
    for (int b = 0; b < nbody; b++) {
        for (int tries = 0; tries < 1000; tries++) {
            size_t c = (size_t)(rng.next() % NC);
            if (owner[c] < 0) { owner[c] = b; cells[b].push_back((int32_t)c); break; }
        }
    }
    auto coord = [&](size_t c, int& x, int& y, int& z) { x = (int)(c % NX); y = (int)((c / NX) % NY); z = (int)(c / ((size_t)NX * NY)); };
    const int D6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    bool grew = true;
    while (grew) {
        grew = false;
        for (int b = 0; b < nbody; b++) {
            if (cells[b].empty() || (int)cells[b].size() >= target[b]) continue;
            // a random free face neighbour of a random brick of the body (a few tries)
            for (int t = 0; t < 12; t++) {
                int32_t c = cells[b][rng.below((uint32_t)cells[b].size())];
                int x, y, z; coord((size_t)c, x, y, z);
                const int* d = D6[rng.below(6)];
                int nx = x + d[0], ny = y + d[1], nz = z + d[2];
                if (nx < 0 || ny < 0 || nz < 0 || nx >= NX || ny >= NY || nz >= NZ) continue;
                size_t nc = (size_t)nx + (size_t)NX * ((size_t)ny + (size_t)NY * nz);
                if (owner[nc] >= 0) continue;
                owner[nc] = b; cells[b].push_back((int32_t)nc); grew = true;
                break;
            }
        }
    }
    int total = 0, nb_out = 0;
    for (int b = 0; b < nbody; b++) if (!cells[b].empty()) { total += (int)cells[b].size(); nb_out++; }
    if (nshape) *nshape = total;
    if (!bodies) return nb_out;
    const double origin[3] = {-0.5 * NX * 0.5, 0.0, -0.5 * NZ * 0.5};
    int si = 0, bi = 0;
    std::vector<int32_t> xyz;
    std::vector<uint8_t> mat;
    for (int b = 0; b < nbody; b++) {
        if (cells[b].empty()) continue;
        // keep the shapes in lattice order within a body
        std::sort(cells[b].begin(), cells[b].end());
        uint8_t m = (uint8_t)(1 + rng.below(5));
        int n = (int)cells[b].size();
        xyz.resize(3 * (size_t)n); mat.assign((size_t)n, m);
        for (int k = 0; k < n; k++) { int x, y, z; coord((size_t)cells[b][k], x, y, z); xyz[3 * k] = x; xyz[3 * k + 1] = y; xyz[3 * k + 2] = z; }
        float* body = bodies + (size_t)bi * SF_BODY_F;
        build_compound(n, xyz.data(), mat.data(), origin, body, shapes + (size_t)si * SF_SHAPE_F);
        for (int a = 0; a < 3; a++) body[7 + a] = (float)(vsig * rng.normal());
        for (int a = 0; a < 3; a++) body[10 + a] = (float)(wsig * rng.normal());
        for (int k = 0; k < n; k++) shape_body[si + k] = bi;
        si += n; bi++;
    }
    return nb_out;
}

// A frame/wall building on the lattice. Storeys are 7 brick layers: 6 of columns and walls, then a concrete slab.
// Columns are 2x2 concrete at every `bay` bricks and along the far edges; the perimeter between columns is
// masonry with a glass band (infill) in layers 2-4. `remove` ground-storey columns are taken out at random.
int sf_scene_building(int wx, int wz, int storeys, int bay, int remove, uint32_t seed, uint8_t* mat, int32_t* xyz, uint8_t* infill) {
    if (wx < 4 || wz < 4 || storeys < 1 || bay < 3) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    auto col_line = [&](int v, int w) { int m = v % bay; return m == 0 || m == 1 || v >= w - 2; };
    // the ground-storey columns: (cx, cz) of each column's corner
    std::vector<std::array<int, 2>> cols;
    for (int x = 0; x < wx; x++) for (int z = 0; z < wz; z++)
        if (col_line(x, wx) && col_line(z, wz) && (x == 0 || !col_line(x - 1, wx)) && (z == 0 || !col_line(z - 1, wz))) cols.push_back({x, z});
    Rng rng(seed);
    std::vector<uint8_t> gone(cols.size(), 0);
    int nrem = std::min<int>(remove, (int)cols.size() - 1);
    for (int r = 0; r < nrem; r++) {
        for (int t = 0; t < 100; t++) { uint32_t k = rng.below((uint32_t)cols.size()); if (!gone[k]) { gone[k] = 1; break; } }
    }
    auto removed = [&](int x, int y, int z) {
        if (y >= 6) return false;
        for (size_t k = 0; k < cols.size(); k++) if (gone[k] && x >= cols[k][0] && x < cols[k][0] + 2 && z >= cols[k][1] && z < cols[k][1] + 2) return true;
        return false;
    };
    int n = 0;
    for (int y = 0; y < 7 * storeys; y++) {
        int ly = y % 7;
        for (int z = 0; z < wz; z++) for (int x = 0; x < wx; x++) {
            uint8_t m = 0, inf = 0;
            bool column = col_line(x, wx) && col_line(z, wz);
            bool perim = x == 0 || z == 0 || x == wx - 1 || z == wz - 1;
            if (ly == 6) m = 1;
            else if (column) m = 1;
            else if (perim) { if (ly >= 2 && ly <= 4) { m = 4; inf = 1; } else m = 2; }
            if (!m || (column && removed(x, y, z))) continue;
            if (mat) { mat[n] = m; xyz[3 * n] = x; xyz[3 * n + 1] = y; xyz[3 * n + 2] = z; infill[n] = inf; }
            n++;
        }
    }
    return n;
}

int sf_scene_tower(uint8_t* mat, int32_t* xyz, uint8_t* infill) {
    int n = 0;
    for (int y = 0; y < 40; y++) for (int z = 0; z < 12; z++) for (int x = 0; x < 12; x++) {
        bool col = (x < 2 || x >= 10) && (z < 2 || z >= 10);
        bool slab = y % 6 == 5;
        if (mat) { mat[n] = (col || slab) ? 1 : 2; xyz[3 * n] = x; xyz[3 * n + 1] = y; xyz[3 * n + 2] = z; infill[n] = 0; }
        n++;
    }
    return n;
}

int sf_scene_site(uint32_t seed, int k, uint8_t* mat, int32_t* xyz, uint8_t* infill) {
    static uint32_t cached_seed = 0;
    static std::vector<SiteStruct> cached;
    if (cached.empty() || cached_seed != seed) { cached = make_site(seed); cached_seed = seed; }
    if (k < 0) return (int)cached.size();
    if (k >= (int)cached.size()) { sf::set_error(SF_E_ARG, "no such structure"); return SF_E_ARG; }
    const SiteStruct& S = cached[k];
    const int n = (int)S.mat.size();
    if (mat) std::memcpy(mat, S.mat.data(), n);
    if (xyz) std::memcpy(xyz, S.xyz.data(), sizeof(int32_t) * 3 * (size_t)n);
    if (infill) std::memcpy(infill, S.infill.data(), n);
    return n;
}

}  // extern "C"
