// Maths of the lattice (legacy brick) engine: f32 vectors, row-major 3x3 matrices and quaternions, each helper in
// the operation order of the reference's src/lib.rs:25-108 (sums left to right). Built with -ffp-contract=off and
// IEEE-rounded sqrt/division, so host and device evaluate every expression to the same bits.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define LAT_HD __host__ __device__ __forceinline__

namespace lat {

constexpr float S = 0.5f, H = 0.25f, RB = 0.27f, GR = 9.81f, MU = 0.65f, GROUNDK = 4.0f;
constexpr int NX2 = 320;  // lattice index of x = 0 (and z = 0)
constexpr int SLUG = 6, STEEL = 3, GLASS = 4;
constexpr uint32_t NONE = 0xFFFFFFFFu;

struct LV { float x, y, z; };
struct LM { float m[9]; };
struct LQ { float x, y, z, w; };

LAT_HD LV lv(float x, float y, float z) { LV r; r.x = x; r.y = y; r.z = z; return r; }
LAT_HD LV lv4(float4 a) { return lv(a.x, a.y, a.z); }
LAT_HD float4 f4(LV a, float w = 0.0f) { return make_float4(a.x, a.y, a.z, w); }
LAT_HD LQ lq4(float4 a) { LQ q; q.x = a.x; q.y = a.y; q.z = a.z; q.w = a.w; return q; }
LAT_HD float4 f4q(LQ q) { return make_float4(q.x, q.y, q.z, q.w); }
LAT_HD float minf(float a, float b) { return a < b ? a : b; }
LAT_HD float maxf(float a, float b) { return a > b ? a : b; }
LAT_HD float clampf(float x, float a, float b) { return minf(maxf(x, a), b); }
LAT_HD LV add(LV a, LV b) { return lv(a.x + b.x, a.y + b.y, a.z + b.z); }
LAT_HD LV sub(LV a, LV b) { return lv(a.x - b.x, a.y - b.y, a.z - b.z); }
LAT_HD LV mul(LV a, float s) { return lv(a.x * s, a.y * s, a.z * s); }
LAT_HD LV madd(LV a, LV b, float s) { return lv(a.x + b.x * s, a.y + b.y * s, a.z + b.z * s); }
LAT_HD LV neg(LV a) { return lv(-a.x, -a.y, -a.z); }
LAT_HD float dot(LV a, LV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LAT_HD LV cross(LV a, LV b) { return lv(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
LAT_HD float len(LV a) { return sqrtf(dot(a, a)); }
LAT_HD LV mv(const LM& m, LV v) {
    const float* a = m.m;
    return lv(a[0] * v.x + a[1] * v.y + a[2] * v.z, a[3] * v.x + a[4] * v.y + a[5] * v.z, a[6] * v.x + a[7] * v.y + a[8] * v.z);
}
LAT_HD LM mmul(const LM& a, const LM& b) {
    LM r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i * 3 + j] = a.m[i * 3] * b.m[j] + a.m[i * 3 + 1] * b.m[3 + j] + a.m[i * 3 + 2] * b.m[6 + j];
    return r;
}
LAT_HD LM mt(const LM& a) {
    LM r;
    r.m[0] = a.m[0]; r.m[1] = a.m[3]; r.m[2] = a.m[6];
    r.m[3] = a.m[1]; r.m[4] = a.m[4]; r.m[5] = a.m[7];
    r.m[6] = a.m[2]; r.m[7] = a.m[5]; r.m[8] = a.m[8];
    return r;
}
LAT_HD LM minv(const LM& m) {
    const float* a = m.m;
    float c0 = a[4] * a[8] - a[5] * a[7];
    float c1 = a[5] * a[6] - a[3] * a[8];
    float c2 = a[3] * a[7] - a[4] * a[6];
    float det = a[0] * c0 + a[1] * c1 + a[2] * c2;
    LM r;
    if (fabsf(det) < 1e-12f) { for (int k = 0; k < 9; k++) r.m[k] = 0.0f; return r; }
    float id = 1.0f / det;
    r.m[0] = c0 * id; r.m[1] = (a[2] * a[7] - a[1] * a[8]) * id; r.m[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    r.m[3] = c1 * id; r.m[4] = (a[0] * a[8] - a[2] * a[6]) * id; r.m[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    r.m[6] = c2 * id; r.m[7] = (a[1] * a[6] - a[0] * a[7]) * id; r.m[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return r;
}
LAT_HD LM q_to_m(LQ q) {
    float x = q.x, y = q.y, z = q.z, w = q.w;
    LM r;
    r.m[0] = 1.0f - 2.0f * (y * y + z * z); r.m[1] = 2.0f * (x * y - z * w); r.m[2] = 2.0f * (x * z + y * w);
    r.m[3] = 2.0f * (x * y + z * w); r.m[4] = 1.0f - 2.0f * (x * x + z * z); r.m[5] = 2.0f * (y * z - x * w);
    r.m[6] = 2.0f * (x * z - y * w); r.m[7] = 2.0f * (y * z + x * w); r.m[8] = 1.0f - 2.0f * (x * x + y * y);
    return r;
}
LAT_HD LQ q_int(LQ q, LV w, float dt) {
    LV qv = lv(q.x, q.y, q.z);
    LV d = add(mul(w, q.w), cross(w, qv));
    float dw = -dot(w, qv);
    float h = 0.5f * dt;
    LQ r;
    r.x = q.x + h * d.x; r.y = q.y + h * d.y; r.z = q.z + h * d.z; r.w = q.w + h * dw;
    float l = sqrtf(r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w);
    if (l > 1e-9f) { r.x /= l; r.y /= l; r.z /= l; r.w /= l; }
    else { r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 1.0f; }
    return r;
}
LAT_HD LV lat_pos(int x, int y, int z) { return lv((float)(x - NX2) * S + H, (float)y * S + H, (float)(z - NX2) * S + H); }
LAT_HD int fcell(float c) { return (int)floorf(c / S); }
LAT_HD uint32_t mix32(uint32_t h) { h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16; return h; }
// claim priority of a contact (its index among its scene's contacts) in colouring round r
LAT_HD uint64_t claim_prio(uint32_t k, int r) {
    return ((uint64_t)(r + 1) << 56) | ((uint64_t)(mix32(k * 0x9E3779B1u + (uint32_t)r * 0x85EBCA77u) >> 8) << 32) | (uint64_t)(0xFFFFFFFFu - k);
}

// material tables (src/lib.rs:143-186)
struct Tables {
    float mass[9], capv[9], caph[9], tens[9], stiff[9], blast[9], eb[9], pulv[9], duct[9];
};
__host__ __device__ inline const Tables& tables() {
    static constexpr Tables T = {
        {0.0f, 300.0f, 230.0f, 250.0f, 60.0f, 70.0f, 2000.0f, 8.0f, 44.0f},
        {0.0f, 140000.0f, 50000.0f, 150000.0f, 1500.0f, 9000.0f, 1e12f, 1e12f, 1e12f},
        {0.0f, 12000.0f, 12000.0f, 90000.0f, 400.0f, 5000.0f, 1e12f, 1e12f, 1e12f},
        {0.0f, 0.35f, 0.1f, 0.9f, 0.5f, 0.7f, 1.0f, 1.0f, 1.0f},
        {0.0f, 1.0f, 0.8f, 2.0f, 0.02f, 0.4f, 1.0f, 1.0f, 1.0f},
        {0.0f, 0.75f, 0.85f, 0.35f, 1.4f, 0.95f, 0.0f, 0.0f, 0.0f},
        {0.0f, 15000.0f, 12000.0f, 1.5e6f, 300.0f, 6000.0f, 1e12f, 1e12f, 1e12f},
        {0.0f, 0.7f, 0.55f, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f, 0.0f},
        {0.0f, 6.0f, 12.0f, 3.0f, 40.0f, 8.0f, 1.0f, 1.0f, 1.0f}};
    return T;
}
LAT_HD float bond_cap(int mi, int mj, int axis) {
    const Tables& T = tables();
    return axis == 1 ? minf(T.capv[mi], T.capv[mj]) : minf(T.caph[mi], T.caph[mj]);
}
constexpr int DIRS[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

}  // namespace lat
