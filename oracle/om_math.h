/* Oracle maths: f32 vectors, quaternions and 3x3 matrices, restated from the reference's engine/src/math.rs
 * (V3 ops :9-27, Quat :33-60, M3 :64-77, sqrt/fabs/sin/cos/floor :80-99).
 *
 * TEST INFRASTRUCTURE ONLY. This file belongs to the CPU oracle: only tests, the smoke check and the benchmark's
 * CPU-baseline leg may use it, and only as the checker. The product path never links it.
 *
 * Every expression keeps the reference's operation order (sums left to right, products as written) and the build
 * uses -ffp-contract=off, so each f32 result is the one the Rust code computes. */
#ifndef OM_MATH_H
#define OM_MATH_H
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef struct { float x, y, z; } V3;
typedef struct { float x, y, z, w; } Quat;
typedef struct { V3 c[3]; } M3; /* columns; for a rotation the body's axes in world space */

static inline V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 vmul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float vlen2(V3 a) { return vdot(a, a); }
/* math.rs:80-83 native branch: sqrt of a non-positive number is 0 */
static inline float om_sqrt(float x) { return x <= 0.0f ? 0.0f : sqrtf(x); }
static inline float vlen(V3 a) { return om_sqrt(vlen2(a)); }
static inline V3 vnorm(V3 a) { float l = vlen(a); return l > 1e-12f ? vmul(a, 1.0f / l) : v3(0, 0, 0); }
static inline float om_fabs(float x) { uint32_t u; memcpy(&u, &x, 4); u &= 0x7fffffffu; memcpy(&x, &u, 4); return x; }
/* Rust f32::max / f32::min for non-NaN operands (the NaN operand loses) */
static inline float om_max(float a, float b) { return (b > a || a != a) ? b : a; }
static inline float om_min(float a, float b) { return (b < a || a != a) ? b : a; }
static inline V3 vmaxv(V3 a, V3 b) { return v3(om_max(a.x, b.x), om_max(a.y, b.y), om_max(a.z, b.z)); }
static inline float vget(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* math.rs:94 floor through an i32 truncation */
static inline float om_floor(float x) { float t = (float)(int32_t)x; return t > x ? t - 1.0f : t; }
/* math.rs:85-93 range-reduced 7th-order polynomial sine */
static inline float om_sin(float x) {
    const float tau = 6.28318530717958647692f, pi = 3.14159265358979323846f;
    float a = x - tau * om_floor(x / tau + 0.5f);
    if (a > pi * 0.5f) a = pi - a; else if (a < -pi * 0.5f) a = -pi - a;
    float a2 = a * a;
    return a * (1.0f - a2 / 6.0f * (1.0f - a2 / 20.0f * (1.0f - a2 / 42.0f * (1.0f - a2 / 72.0f))));
}
static inline float om_cos(float x) { return om_sin(x + 1.57079632679489661923f); }

static const Quat Q_ID = {0, 0, 0, 1};
static inline Quat quat(float x, float y, float z, float w) { Quat q = {x, y, z, w}; return q; }
/* math.rs:36-41 */
static inline Quat qmul(Quat s, Quat o) {
    return quat(s.w * o.x + s.x * o.w + s.y * o.z - s.z * o.y,
                s.w * o.y - s.x * o.z + s.y * o.w + s.z * o.x,
                s.w * o.z + s.x * o.y - s.y * o.x + s.z * o.w,
                s.w * o.w - s.x * o.x - s.y * o.y - s.z * o.z);
}
static inline Quat qnormalized(Quat q) {
    float l = om_sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    float k = l > 0.0f ? 1.0f / l : 1.0f;
    return quat(q.x * k, q.y * k, q.z * k, q.w * k);
}
/* math.rs:44-47 first-order integration, renormalised */
static inline Quat qintegrate(Quat q, V3 w, float h) {
    Quat dq = qmul(quat(w.x, w.y, w.z, 0.0f), q);
    return qnormalized(quat(q.x + 0.5f * h * dq.x, q.y + 0.5f * h * dq.y, q.z + 0.5f * h * dq.z, q.w + 0.5f * h * dq.w));
}
static inline Quat qconj(Quat q) { return quat(-q.x, -q.y, -q.z, q.w); }
static inline Quat q_axis_angle(V3 a, float ang) {
    a = vnorm(a);
    float s = om_sin(ang * 0.5f), c = om_cos(ang * 0.5f);
    return quat(a.x * s, a.y * s, a.z * s, c);
}
/* math.rs:54-59 */
static inline M3 q_to_m3(Quat q) {
    float x = q.x, y = q.y, z = q.z, w = q.w;
    M3 m;
    m.c[0] = v3(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + z * w), 2.0f * (x * z - y * w));
    m.c[1] = v3(2.0f * (x * y - z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + x * w));
    m.c[2] = v3(2.0f * (x * z + y * w), 2.0f * (y * z - x * w), 1.0f - 2.0f * (x * x + y * y));
    return m;
}
static inline V3 m3_mulv(const M3* m, V3 v) { return vadd(vadd(vmul(m->c[0], v.x), vmul(m->c[1], v.y)), vmul(m->c[2], v.z)); }
static inline V3 m3_tmulv(const M3* m, V3 v) { return v3(vdot(m->c[0], v), vdot(m->c[1], v), vdot(m->c[2], v)); }
/* math.rs:71-75: R diag(d) R^T, column by column */
static inline M3 m3_sandwich(const M3* r, V3 d) {
    V3 a = vmul(r->c[0], d.x), b = vmul(r->c[1], d.y), c = vmul(r->c[2], d.z);
    M3 o;
    for (int j = 0; j < 3; j++)
        o.c[j] = vadd(vadd(vmul(a, vget(r->c[0], j)), vmul(b, vget(r->c[1], j))), vmul(c, vget(r->c[2], j)));
    return o;
}
static inline M3 m3_diag(V3 d) { M3 m; m.c[0] = v3(d.x, 0, 0); m.c[1] = v3(0, d.y, 0); m.c[2] = v3(0, 0, d.z); return m; }
static inline M3 m3_identity(void) { return m3_diag(v3(1, 1, 1)); }
/* a * b for column matrices: column j of the product is a applied to column j of b */
static inline M3 m3_mul(const M3* a, const M3* b) { M3 o; for (int j = 0; j < 3; j++) o.c[j] = m3_mulv(a, b->c[j]); return o; }
#endif
