// Device maths for the box step: f32 vectors, quaternions and 3x3 rotation matrices.
// Every helper evaluates in the order the reference's engine/src/math.rs does (sums left to right), and the
// library is built with -ffp-contract=off and IEEE-rounded sqrt/division, so a kernel's f32 results are the
// reference's bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SF_HD __host__ __device__ __forceinline__

struct V3 { float x, y, z; };
struct Qt { float x, y, z, w; };
struct M3 { V3 c[3]; };  // columns (a rotation's columns are the body axes in world space)

SF_HD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
SF_HD V3 v3f4(float4 a) { return v3(a.x, a.y, a.z); }
SF_HD float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
SF_HD V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
SF_HD V3 vsel(bool c, V3 a, V3 b) { return v3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }
SF_HD V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
SF_HD V3 vmul(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
SF_HD V3 vneg(V3 a) { return v3(-a.x, -a.y, -a.z); }
SF_HD float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SF_HD V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
SF_HD float vlen2(V3 a) { return vdot(a, a); }
SF_HD float sf_sqrt(float x) { return x <= 0.0f ? 0.0f : sqrtf(x); }
SF_HD float vlen(V3 a) { return sf_sqrt(vlen2(a)); }
SF_HD V3 vnorm(V3 a) { float l = vlen(a); return l > 1e-12f ? vmul(a, 1.0f / l) : v3(0.0f, 0.0f, 0.0f); }
SF_HD float sf_abs(float x) { return fabsf(x); }
// max/min of two non-NaN values keeping the first on ties (the NaN operand loses)
SF_HD float sf_max(float a, float b) { return (b > a || a != a) ? b : a; }
SF_HD float sf_min(float a, float b) { return (b < a || a != a) ? b : a; }
SF_HD V3 vmaxv(V3 a, V3 b) { return v3(sf_max(a.x, b.x), sf_max(a.y, b.y), sf_max(a.z, b.z)); }
SF_HD float vget(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// floor through an i32 truncation (the engine's own floor)
SF_HD float sf_floor(float x) { float t = (float)(int32_t)x; return t > x ? t - 1.0f : t; }

SF_HD Qt qt(float x, float y, float z, float w) { Qt q; q.x = x; q.y = y; q.z = z; q.w = w; return q; }
SF_HD Qt qf4(float4 a) { return qt(a.x, a.y, a.z, a.w); }
SF_HD float4 f4q(Qt q) { return make_float4(q.x, q.y, q.z, q.w); }
SF_HD Qt qmul(Qt s, Qt o) {
    return qt(s.w * o.x + s.x * o.w + s.y * o.z - s.z * o.y,
              s.w * o.y - s.x * o.z + s.y * o.w + s.z * o.x,
              s.w * o.z + s.x * o.y - s.y * o.x + s.z * o.w,
              s.w * o.w - s.x * o.x - s.y * o.y - s.z * o.z);
}
SF_HD Qt qnormalized(Qt q) {
    float l = sf_sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    float k = l > 0.0f ? 1.0f / l : 1.0f;
    return qt(q.x * k, q.y * k, q.z * k, q.w * k);
}
// q advanced by angular velocity w over h: first order, renormalised
SF_HD Qt qintegrate(Qt q, V3 w, float h) {
    Qt dq = qmul(qt(w.x, w.y, w.z, 0.0f), q);
    return qnormalized(qt(q.x + 0.5f * h * dq.x, q.y + 0.5f * h * dq.y, q.z + 0.5f * h * dq.z, q.w + 0.5f * h * dq.w));
}
SF_HD Qt qconj(Qt q) { return qt(-q.x, -q.y, -q.z, q.w); }
SF_HD M3 q_to_m3(Qt q) {
    float x = q.x, y = q.y, z = q.z, w = q.w;
    M3 m;
    m.c[0] = v3(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + z * w), 2.0f * (x * z - y * w));
    m.c[1] = v3(2.0f * (x * y - z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + x * w));
    m.c[2] = v3(2.0f * (x * z + y * w), 2.0f * (y * z - x * w), 1.0f - 2.0f * (x * x + y * y));
    return m;
}
SF_HD V3 mulv(const M3& m, V3 v) { return vadd(vadd(vmul(m.c[0], v.x), vmul(m.c[1], v.y)), vmul(m.c[2], v.z)); }
SF_HD V3 tmulv(const M3& m, V3 v) { return v3(vdot(m.c[0], v), vdot(m.c[1], v), vdot(m.c[2], v)); }
// R diag(d) R^T built column by column
SF_HD M3 sandwich(const M3& r, V3 d) {
    V3 a = vmul(r.c[0], d.x), b = vmul(r.c[1], d.y), c = vmul(r.c[2], d.z);
    M3 o;
    for (int j = 0; j < 3; j++) o.c[j] = vadd(vadd(vmul(a, vget(r.c[0], j)), vmul(b, vget(r.c[1], j))), vmul(c, vget(r.c[2], j)));
    return o;
}
SF_HD M3 m3mul(const M3& a, const M3& b) { M3 o; for (int j = 0; j < 3; j++) o.c[j] = mulv(a, b.c[j]); return o; }

SF_HD uint32_t mix32(uint32_t h) { h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16; return h; }
// hash of a manifold's key (shape a, shape b or -1 for the ground): colour choices and priorities go by it
SF_HD uint32_t key_hash(int32_t a, int32_t b) { return mix32((uint32_t)a * 0x9E3779B9u + mix32((uint32_t)b ^ 0x5BD1E995u)); }
// a manifold's sort key: shape a, then the ground (b = -1), a chunk's second ground patch (b = -2), partners by index
SF_HD uint64_t man_key(int32_t a, int32_t b) { return ((uint64_t)(uint32_t)a << 32) | (uint64_t)(b >= 0 ? (uint32_t)b + 3u : (uint32_t)(-b)); }
