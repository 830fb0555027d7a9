/* Legacy lattice engine: maths helpers and material tables, restated from src/lib.rs:25-108 and :143-186.
 *
 * TEST INFRASTRUCTURE ONLY: part of the CPU oracle. Every expression keeps the reference's operation order
 * (left-associative sums, the same products), so that a build with -ffp-contract=off reproduces it bit for bit.
 */
#ifndef SF_ORACLE_LEGACY_H
#define SF_ORACLE_LEGACY_H
#include <math.h>
#include <stdint.h>

#define LG_S 0.5f
#define LG_H 0.25f
#define LG_NX2 320 /* NX / 2, NZ / 2 (src/lib.rs:113-115) */
#define LG_RB 0.27f
#define LG_GR 9.81f
#define LG_MU 0.65f
#define LG_GROUNDK 4.0f
#define LG_SLUG 6
#define LG_STEEL 3
#define LG_GLASS 4

static const float LG_MASS[9] = {0.0f, 300.0f, 230.0f, 250.0f, 60.0f, 70.0f, 2000.0f, 8.0f, 44.0f};
static const float LG_CAPV[9] = {0.0f, 140000.0f, 50000.0f, 150000.0f, 1500.0f, 9000.0f, 1e12f, 1e12f, 1e12f};
static const float LG_CAPH[9] = {0.0f, 12000.0f, 12000.0f, 90000.0f, 400.0f, 5000.0f, 1e12f, 1e12f, 1e12f};
static const float LG_TENS[9] = {0.0f, 0.35f, 0.1f, 0.9f, 0.5f, 0.7f, 1.0f, 1.0f, 1.0f};
static const float LG_STIFF[9] = {0.0f, 1.0f, 0.8f, 2.0f, 0.02f, 0.4f, 1.0f, 1.0f, 1.0f};
static const float LG_BLAST[9] = {0.0f, 0.75f, 0.85f, 0.35f, 1.4f, 0.95f, 0.0f, 0.0f, 0.0f};
static const float LG_EB[9] = {0.0f, 15000.0f, 12000.0f, 1.5e6f, 300.0f, 6000.0f, 1e12f, 1e12f, 1e12f};
static const float LG_PULV[9] = {0.0f, 0.7f, 0.55f, 0.0f, 1.0f, 0.5f, 0.0f, 0.0f, 0.0f};
static const float LG_DUCT[9] = {0.0f, 6.0f, 12.0f, 3.0f, 40.0f, 8.0f, 1.0f, 1.0f, 1.0f};
static const int LG_DIRS[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

typedef struct { float x, y, z; } LV;
typedef struct { float m[9]; } LM;
typedef struct { float x, y, z, w; } LQ;

static inline float lg_minf(float a, float b) { return a < b ? a : b; }
static inline float lg_maxf(float a, float b) { return a > b ? a : b; }
static inline float lg_clampf(float x, float a, float b) { return lg_minf(lg_maxf(x, a), b); }
static inline float lg_abs(float x) { return fabsf(x); }
static inline LV lv(float x, float y, float z) { LV r = {x, y, z}; return r; }
static inline LV lg_add(LV a, LV b) { return lv(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline LV lg_sub(LV a, LV b) { return lv(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline LV lg_mul(LV a, float s) { return lv(a.x * s, a.y * s, a.z * s); }
static inline LV lg_madd(LV a, LV b, float s) { return lv(a.x + b.x * s, a.y + b.y * s, a.z + b.z * s); }
static inline float lg_dot(LV a, LV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline LV lg_cross(LV a, LV b) { return lv(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float lg_len(LV a) { return sqrtf(lg_dot(a, a)); }
static inline LV lg_norm(LV a) { float l = lg_len(a); return l > 1e-9f ? lg_mul(a, 1.0f / l) : lv(0.0f, 1.0f, 0.0f); }
static inline LV lg_mv(const LM* m, LV v) {
    const float* a = m->m;
    return lv(a[0] * v.x + a[1] * v.y + a[2] * v.z, a[3] * v.x + a[4] * v.y + a[5] * v.z, a[6] * v.x + a[7] * v.y + a[8] * v.z);
}
static inline LM lg_mmul(const LM* a, const LM* b) {
    LM r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i * 3 + j] = a->m[i * 3] * b->m[j] + a->m[i * 3 + 1] * b->m[3 + j] + a->m[i * 3 + 2] * b->m[6 + j];
    return r;
}
static inline LM lg_mt(const LM* a) {
    LM r = {{a->m[0], a->m[3], a->m[6], a->m[1], a->m[4], a->m[7], a->m[2], a->m[5], a->m[8]}};
    return r;
}
static inline LM lg_minv(const LM* m) {
    const float* a = m->m;
    float c0 = a[4] * a[8] - a[5] * a[7];
    float c1 = a[5] * a[6] - a[3] * a[8];
    float c2 = a[3] * a[7] - a[4] * a[6];
    float det = a[0] * c0 + a[1] * c1 + a[2] * c2;
    LM r;
    if (lg_abs(det) < 1e-12f) { for (int k = 0; k < 9; k++) r.m[k] = 0.0f; return r; }
    float id = 1.0f / det;
    r.m[0] = c0 * id; r.m[1] = (a[2] * a[7] - a[1] * a[8]) * id; r.m[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    r.m[3] = c1 * id; r.m[4] = (a[0] * a[8] - a[2] * a[6]) * id; r.m[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    r.m[6] = c2 * id; r.m[7] = (a[1] * a[6] - a[0] * a[7]) * id; r.m[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return r;
}
static inline LM lg_q_to_m(LQ q) {
    float x = q.x, y = q.y, z = q.z, w = q.w;
    LM r = {{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - z * w), 2.0f * (x * z + y * w),
             2.0f * (x * y + z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - x * w),
             2.0f * (x * z - y * w), 2.0f * (y * z + x * w), 1.0f - 2.0f * (x * x + y * y)}};
    return r;
}
static inline LQ lg_q_int(LQ q, LV w, float dt) {
    LV qv = lv(q.x, q.y, q.z);
    LV d = lg_add(lg_mul(w, q.w), lg_cross(w, qv));
    float dw = -lg_dot(w, qv);
    float h = 0.5f * dt;
    LQ r = {q.x + h * d.x, q.y + h * d.y, q.z + h * d.z, q.w + h * dw};
    float l = sqrtf(r.x * r.x + r.y * r.y + r.z * r.z + r.w * r.w);
    if (l > 1e-9f) { r.x /= l; r.y /= l; r.z /= l; r.w /= l; }
    else { r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 1.0f; }
    return r;
}
static inline LV lg_lat_pos(int x, int y, int z) {
    return lv((float)(x - LG_NX2) * LG_S + LG_H, (float)y * LG_S + LG_H, (float)(z - LG_NX2) * LG_S + LG_H);
}
static inline int lg_fcell(float c) { return (int)floorf(c / LG_S); }
static inline float lg_bond_cap(int mi, int mj, int axis) {
    return axis == 1 ? lg_minf(LG_CAPV[mi], LG_CAPV[mj]) : lg_minf(LG_CAPH[mi], LG_CAPH[mj]);
}
/* xorshift32 as rnd() (src/lib.rs:323-328) */
static inline float lg_rnd(uint32_t* s) {
    uint32_t x = *s;
    x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    *s = x;
    return (float)(x >> 8) / 16777216.0f;
}
static inline LV lg_rndv(uint32_t* s) {
    for (;;) {
        float a = lg_rnd(s) * 2.0f - 1.0f, b = lg_rnd(s) * 2.0f - 1.0f, c = lg_rnd(s) * 2.0f - 1.0f;
        LV v = lv(a, b, c);
        float l = lg_dot(v, v);
        if (l > 0.01f && l <= 1.0f) return lg_mul(v, 1.0f / sqrtf(l));
    }
}
#endif
