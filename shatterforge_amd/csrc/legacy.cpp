// The legacy engine's own exports (src/lib.rs, tools/debug_exports.rs) on one default brick world, for a caller
// written against the reference's singleton: same names (prefixed sf_legacy_ because `step` also exists in the box
// core's export set), same argument meaning, pointer results that stay valid until the next call. Everything is
// built on the handle API (sf_lat_*, sf_load_*), so these are views and conveniences, not a second engine.
//
// The map: the reference's init builds its hand-made city (build_map, src/lib.rs:3528-3590, game content); this one
// builds the seeded site scene (sf_scene_site) and gives every structure its converged load field up front, where
// the reference warms the structure nearest the player in the background (src/lib.rs:3688-3698).
#include "../../include/sf.h"
#include "sf_host.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int MB = 500000;          // brick pool of the default world (src/lib.rs:121)
constexpr int BODIES = 32768;       // body slots: enough for the site's collapse (the reference's 16,000, :122, drops)
constexpr int MAXCG = 8;            // thrown charges held at once (:147)
constexpr float S = 0.5f;           // brick size (m)

struct Legacy {
    sf_lat* lat = nullptr;
    int scene = 0, nstruct = 0, nbricks = 0;
    std::vector<int32_t> xyz;       // every brick's lattice cell as built (scene brick order)
    std::vector<int> lo, hi;        // structure k: bricks [lo, hi)
    std::vector<float> out_st, stats, dbg, bodies, bricks;
    float charges[MAXCG][3] = {};
    bool charged[MAXCG] = {};
    std::vector<float> events;
    ~Legacy() { if (lat) sf_lat_destroy(lat); }
};
Legacy& G() { static Legacy g; return g; }

void refresh(Legacy& L) {
    L.stats.assign(16, 0.0f);
    L.out_st.assign(12 * (size_t)std::max(L.nstruct, 1), 0.0f);
    if (!L.lat) return;
    sf_lat_stats(L.lat, L.scene, L.stats.data());
    sf_lat_structs(L.lat, L.scene, L.out_st.data());
}
// the scene's bodies and bricks as the handle reports them (SF_LAT_BODY_F / SF_LAT_BRICK_F floats each)
bool pull(Legacy& L) {
    if (!L.lat) return false;
    int32_t c[3] = {0, 0, 0};
    if (sf_lat_counts(L.lat, L.scene, c) < 0) return false;
    L.bricks.assign((size_t)std::max(c[0], 1) * SF_LAT_BRICK_F, 0.0f);
    L.bodies.assign((size_t)std::max(c[1], 1) * SF_LAT_BODY_F, 0.0f);
    if (sf_lat_bricks(L.lat, L.scene, L.bricks.data()) < 0 || sf_lat_bodies(L.lat, L.scene, L.bodies.data()) < 0) return false;
    L.bodies.resize((size_t)c[1] * SF_LAT_BODY_F);
    L.bricks.resize((size_t)c[0] * SF_LAT_BRICK_F);
    return true;
}
// the debug exports' shared 512-float buffer (tools/debug_exports.rs:1), cleared
float* dbg_clear(Legacy& L) {
    L.dbg.assign(512, 0.0f);
    return L.dbg.data();
}

}  // namespace

extern "C" {

uint32_t sf_legacy_init(void) {
    Legacy& L = G();
    if (L.lat) { sf_lat_destroy(L.lat); L.lat = nullptr; }
    L.xyz.clear(); L.lo.clear(); L.hi.clear();
    for (int k = 0; k < MAXCG; k++) L.charged[k] = false;
    L.events.clear();
    const uint32_t seed = 1;
    const int ns = sf_scene_site(seed, -1, nullptr, nullptr, nullptr);
    if (ns <= 0) return 0;
    L.lat = sf_lat_create(MB, BODIES, 1, 16, 8);
    if (!L.lat) return 0;
    L.scene = sf_lat_scene_add(L.lat, MB, BODIES, ns, 0x9E3779B9u);
    if (L.scene < 0) { sf_lat_destroy(L.lat); L.lat = nullptr; return 0; }
    sf_load* load = sf_load_create();
    std::vector<uint8_t> mat, inf;
    std::vector<int32_t> xyz;
    int nb = 0;
    for (int k = 0; k < ns; k++) {
        const int n = sf_scene_site(seed, k, nullptr, nullptr, nullptr);
        mat.resize(n); inf.resize(n); xyz.resize(3 * (size_t)n);
        sf_scene_site(seed, k, mat.data(), xyz.data(), inf.data());
        if (sf_lat_add_structure(L.lat, L.scene, n, xyz.data(), mat.data(), inf.data()) < 0 ||
            sf_load_add(load, n, mat.data(), xyz.data(), inf.data(), nullptr) < 0) {
            sf_load_destroy(load); sf_lat_destroy(L.lat); L.lat = nullptr; return 0;
        }
        L.lo.push_back(nb); L.hi.push_back(nb + n);
        L.xyz.insert(L.xyz.end(), xyz.begin(), xyz.end());
        nb += n;
    }
    L.nstruct = ns; L.nbricks = nb;
    bool ok = sf_lat_finalize(L.lat, L.scene) >= 0 && sf_load_solve(load, 0.0f, 1e-6, 200000) >= 0;
    std::vector<double> phi64;
    std::vector<float> phi;
    for (int k = 0; ok && k < ns; k++) {
        const int n = L.hi[k] - L.lo[k];
        phi64.resize(n); phi.resize(n);
        ok = sf_load_field(load, k, phi64.data()) >= 0;
        for (int i = 0; ok && i < n; i++) phi[i] = (float)phi64[i];
        ok = ok && sf_lat_set_field(L.lat, L.scene, k, phi.data()) >= 0;
    }
    sf_load_destroy(load);
    if (!ok) { sf_lat_destroy(L.lat); L.lat = nullptr; return 0; }
    refresh(L);
    return (uint32_t)nb;
}

void sf_legacy_step(float dt) {
    Legacy& L = G();
    if (!L.lat) return;
    sf_lat_step(L.lat, std::min(std::max(dt, 0.0f), 1.0f / 30.0f));
    refresh(L);
}

void sf_legacy_explode(float x, float y, float z, float r, float p) {
    Legacy& L = G();
    if (!L.lat) return;
    sf_lat_explode(L.lat, L.scene, x, y, z, r, p);
    refresh(L);
}

// the charge is placed where it is thrown from: its flight and sticking (charges_step) are game behaviour; the
// velocity is accepted for the signature's sake
uint32_t sf_legacy_throw_charge(float ox, float oy, float oz, float vx, float vy, float vz) {
    Legacy& L = G();
    (void)vx; (void)vy; (void)vz;
    if (!std::isfinite(ox) || !std::isfinite(oy) || !std::isfinite(oz)) return 0;
    for (int k = 0; k < MAXCG; k++)
        if (!L.charged[k]) { L.charged[k] = true; L.charges[k][0] = ox; L.charges[k][1] = oy; L.charges[k][2] = oz; return 1; }
    return 0;
}

// every held charge goes off with radius 3 and power 1.4 (src/lib.rs:3942-3954); returns how many
uint32_t sf_legacy_detonate(void) {
    Legacy& L = G();
    uint32_t n = 0;
    for (int k = 0; k < MAXCG; k++) {
        if (!L.charged[k]) continue;
        L.charged[k] = false;
        if (L.lat) sf_lat_explode(L.lat, L.scene, L.charges[k][0], L.charges[k][1], L.charges[k][2], 3.0f, 1.4f);
        n++;
    }
    refresh(L);
    return n;
}

uint32_t sf_legacy_n_structs(void) { return (uint32_t)G().nstruct; }
const float* sf_legacy_ptr_structs(void) { Legacy& L = G(); if (L.out_st.empty()) refresh(L); return L.out_st.data(); }
const float* sf_legacy_ptr_stats(void) { Legacy& L = G(); if (L.stats.empty()) refresh(L); return L.stats.data(); }
// events (sound, particles, the HUD's break counter) are produced by the render and game side: none here
void sf_legacy_clear_events(void) { G().events.clear(); }
uint32_t sf_legacy_n_events(void) { return 0; }
const float* sf_legacy_ptr_events(void) { Legacy& L = G(); L.events.assign(5, 0.0f); return L.events.data(); }

// per height of structure s: [standing, detached (in a piece), removed] bricks, and the x/z cell extents of the
// detached ones at 300..303 (tools/debug_exports.rs:3-19)
const float* sf_legacy_dbg_hist(uint32_t s) {
    Legacy& L = G();
    float* d = dbg_clear(L);
    d[300] = 1e9f; d[301] = -1e9f; d[302] = 1e9f; d[303] = -1e9f;
    if (!pull(L) || s >= (uint32_t)L.nstruct) return d;
    for (int i = L.lo[s]; i < L.hi[s]; i++) {
        const int y = L.xyz[3 * (size_t)i + 1];
        if (y < 0 || y >= 100) continue;
        const float* b = &L.bricks[(size_t)i * SF_LAT_BRICK_F];
        if (b[0] == 0.0f) d[y * 3 + 2] += 1.0f;
        else if (b[1] < 0.0f) d[y * 3] += 1.0f;
        else {
            d[y * 3 + 1] += 1.0f;
            const float x = (float)L.xyz[3 * (size_t)i], z = (float)L.xyz[3 * (size_t)i + 2];
            d[300] = std::min(d[300], x); d[301] = std::max(d[301], x); d[302] = std::min(d[302], z); d[303] = std::max(d[303], z);
        }
    }
    return d;
}

// the 8 largest pieces: (bricks, vertical extent + S, horizontal extent + S) (tools/debug_exports.rs:20-40)
const float* sf_legacy_dbg_bodies(void) {
    Legacy& L = G();
    float* d = dbg_clear(L);
    if (!pull(L)) return d;
    const size_t nb = L.bodies.size() / SF_LAT_BODY_F;
    std::vector<float> lo(3 * nb, 1e9f), hi(3 * nb, -1e9f);
    for (size_t i = 0; i < L.bricks.size() / SF_LAT_BRICK_F; i++) {
        const float* b = &L.bricks[i * SF_LAT_BRICK_F];
        if (b[0] == 0.0f || b[1] < 0.0f) continue;
        const size_t k = (size_t)b[1];
        if (k >= nb) continue;
        for (int a = 0; a < 3; a++) { lo[3 * k + a] = std::min(lo[3 * k + a], b[2 + a]); hi[3 * k + a] = std::max(hi[3 * k + a], b[2 + a]); }
    }
    for (size_t k = 0; k < nb; k++) {
        const float* B = &L.bodies[k * SF_LAT_BODY_F];
        if (B[0] == 0.0f) continue;
        const float c = B[2];
        int pos = 8;
        for (int t = 0; t < 8; t++) if (c > d[t * 3]) { pos = t; break; }
        if (pos == 8) continue;
        for (int t = 7; t > pos; t--) for (int q = 0; q < 3; q++) d[t * 3 + q] = d[(t - 1) * 3 + q];
        d[pos * 3] = c;
        d[pos * 3 + 1] = hi[3 * k + 1] - lo[3 * k + 1] + S;
        d[pos * 3 + 2] = std::max(hi[3 * k] - lo[3 * k], hi[3 * k + 2] - lo[3 * k + 2]) + S;
    }
    return d;
}

// per dominant material: pieces, bricks, largest, pieces of 50+ bricks, pieces of 1-2 (tools/debug_exports.rs:41-56)
const float* sf_legacy_dbg_matstats(void) {
    Legacy& L = G();
    float* d = dbg_clear(L);
    if (L.lat) sf_lat_matstats(L.lat, L.scene, d);
    return d;
}

// fastest awake piece (horizontal speed, its bricks, vertical speed), pieces over 5 m/s and their bricks, pieces
// rising over 3 m/s, awake pieces, fastest rim spin (speed, bricks, rad/s) (tools/debug_exports.rs:57-71)
const float* sf_legacy_dbg_speed(void) {
    Legacy& L = G();
    float* d = dbg_clear(L);
    if (!pull(L)) return d;
    const size_t nb = L.bodies.size() / SF_LAT_BODY_F;
    // the rim radius: the largest brick distance from the centre plus 1.8 half sizes (setup_body, src/lib.rs:880)
    std::vector<float> r2(nb, 0.0f);
    for (size_t i = 0; i < L.bricks.size() / SF_LAT_BRICK_F; i++) {
        const float* b = &L.bricks[i * SF_LAT_BRICK_F];
        if (b[0] == 0.0f || b[1] < 0.0f || (size_t)b[1] >= nb) continue;
        const float* B = &L.bodies[(size_t)b[1] * SF_LAT_BODY_F];
        const float dx = b[2] - B[4], dy = b[3] - B[5], dz = b[4] - B[6];
        r2[(size_t)b[1]] = std::max(r2[(size_t)b[1]], dx * dx + dy * dy + dz * dz);
    }
    for (size_t k = 0; k < nb; k++) {
        const float* B = &L.bodies[k * SF_LAT_BODY_F];
        if (B[0] == 0.0f || B[1] != 0.0f) continue;
        const float hs = std::sqrt(B[11] * B[11] + B[13] * B[13]);
        if (hs > d[0]) { d[0] = hs; d[1] = B[2]; d[2] = B[12]; }
        if (hs > 5.0f) { d[3] += 1.0f; d[4] += B[2]; }
        if (B[12] > 3.0f) d[5] += 1.0f;
        d[6] += 1.0f;
        const float wl = std::sqrt(B[14] * B[14] + B[15] * B[15] + B[16] * B[16]);
        const float tip = wl * (std::sqrt(r2[k]) + 0.25f * 1.8f);
        if (tip > d[7]) { d[7] = tip; d[8] = B[2]; d[9] = wl; }
    }
    return d;
}

// the largest piece: bricks, position, tilt of its up axis (degrees), speed, spin, asleep, worst joint ratio,
// load-field residual, contacts, 0 (tools/debug_exports.rs:72-87); the pointer is to its first of 12 values
const float* sf_legacy_dbg_bigbody(void) {
    Legacy& L = G();
    float* d = dbg_clear(L);
    if (!pull(L)) return d + 200;
    for (size_t k = 0; k < L.bodies.size() / SF_LAT_BODY_F; k++) {
        const float* B = &L.bodies[k * SF_LAT_BODY_F];
        if (B[0] == 0.0f || !(B[2] > d[200])) continue;
        d[200] = B[2]; d[201] = B[4]; d[202] = B[5]; d[203] = B[6];
        const float qx = B[7], qz = B[9];
        const float up_y = 1.0f - 2.0f * (qx * qx + qz * qz);  // the rotated up axis' y component
        d[204] = std::acos(std::min(1.0f, std::max(-1.0f, up_y))) * 57.3f;
        d[205] = std::sqrt(B[11] * B[11] + B[12] * B[12] + B[13] * B[13]);
        d[206] = std::sqrt(B[14] * B[14] + B[15] * B[15] + B[16] * B[16]);
        d[207] = B[1]; d[208] = B[19]; d[209] = B[20]; d[210] = B[3]; d[211] = 0.0f;
    }
    return d + 200;
}

// structure s's 12 most stressed standing bricks: (ratio, x, y, z, material), and its worst ratio at 60
// (tools/debug_exports.rs:89-105)
const float* sf_legacy_dbg_top(uint32_t s) {
    Legacy& L = G();
    float* d = dbg_clear(L);
    if (!pull(L) || s >= (uint32_t)L.nstruct) return d;
    float worst = 0.0f;
    for (int i = L.lo[s]; i < L.hi[s]; i++) {
        const float* b = &L.bricks[(size_t)i * SF_LAT_BRICK_F];
        if (b[0] == 0.0f || b[1] >= 0.0f) continue;
        const float r = b[11];
        worst = std::max(worst, r);
        int pos = 12;
        for (int t = 0; t < 12; t++) if (r > d[t * 5]) { pos = t; break; }
        if (pos == 12) continue;
        for (int t = 11; t > pos; t--) for (int q = 0; q < 5; q++) d[t * 5 + q] = d[(t - 1) * 5 + q];
        d[pos * 5] = r;
        d[pos * 5 + 1] = (float)L.xyz[3 * (size_t)i]; d[pos * 5 + 2] = (float)L.xyz[3 * (size_t)i + 1];
        d[pos * 5 + 3] = (float)L.xyz[3 * (size_t)i + 2]; d[pos * 5 + 4] = b[0];
    }
    d[60] = worst;
    return d;
}

void sf_legacy_dbg_wake_all(void) { Legacy& L = G(); if (L.lat) sf_lat_wake_all(L.lat, L.scene); }

// the default world's handle and scene, for a caller mixing both APIs (no counterpart in the reference)
sf_lat* sf_legacy_world(int* scene) { Legacy& L = G(); if (scene) *scene = L.scene; return L.lat; }

}  // extern "C"
