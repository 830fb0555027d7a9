/* CPU oracle: the legacy load-path solve of the reference's src/lib.rs, static structures:
 *   material tables MASS/CAPV/CAPH/TENS/STIFF (lib.rs:166-172), GROUNDK (:168),
 *   linked (:378-383), sor_brick (:1005-1026), sor_struct (:1029-1040), cond/cond_w (:1142-1151),
 *   stress_brick (:1153-1227, the ratio and the mode of each bond of a standing brick),
 *   conn_struct's grounded search (:1085-1098).
 * The FP64 solve (conjugate gradients) solves the same linear system the SOR sweeps iterate on: for every brick,
 * sum_j k_ij (phi_i - phi_j) + g_i phi_i = MASS_i + ext_i with g_i = GROUNDK * STIFF_i on grounded footings.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): tests, the smoke check and the benchmark's CPU-baseline leg use it. */
#include "sf_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const float MASS[9] = {0.0f, 300.0f, 230.0f, 250.0f, 60.0f, 70.0f, 2000.0f, 8.0f, 44.0f};
static const float CAPV[9] = {0.0f, 140000.0f, 50000.0f, 150000.0f, 1500.0f, 9000.0f, 1e12f, 1e12f, 1e12f};
static const float CAPH[9] = {0.0f, 12000.0f, 12000.0f, 90000.0f, 400.0f, 5000.0f, 1e12f, 1e12f, 1e12f};
static const float TENS[9] = {0.0f, 0.35f, 0.1f, 0.9f, 0.5f, 0.7f, 1.0f, 1.0f, 1.0f};
static const float STIFF[9] = {0.0f, 1.0f, 0.8f, 2.0f, 0.02f, 0.4f, 1.0f, 1.0f, 1.0f};
static const float GROUNDK = 4.0f;

/* lib.rs:646-649 */
static float bond_cap(int mi, int mj, int axis) {
    float a = axis == 1 ? CAPV[mi] : CAPH[mi], b = axis == 1 ? CAPV[mj] : CAPH[mj];
    return a < b ? a : b;
}

/* lib.rs:668-684 (the neighbour table and bond capacities of standing bricks; all bricks of one structure).
 * infill: 1 for glazing / partition bricks (GLASS is always infill, lib.rs:635). */
void ol_finalize(int n, const uint8_t* mat, const int32_t* xyz, const uint8_t* infill, int32_t* nb6, float* bond) {
    /* lattice lookup by sorting cells: a hash from packed coordinates to the brick index */
    size_t cap = 1; while (cap < (size_t)n * 2 + 2) cap <<= 1;
    int64_t* keys = (int64_t*)malloc(sizeof(int64_t) * cap);
    int32_t* vals = (int32_t*)malloc(sizeof(int32_t) * cap);
    for (size_t k = 0; k < cap; k++) keys[k] = -1;
    #define PACK(x, y, z) ((((int64_t)(x) + 4096) << 32) | (((int64_t)(y) + 4096) << 16) | ((int64_t)(z) + 4096))
    for (int i = 0; i < n; i++) {
        int64_t key = PACK(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        size_t h = (size_t)((uint64_t)key * 0x9E3779B97F4A7C15ull >> 20) & (cap - 1);
        while (keys[h] != -1) h = (h + 1) & (cap - 1);
        keys[h] = key; vals[h] = i;
    }
    static const int DIRS[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int i = 0; i < n; i++) {
        for (int d = 0; d < 6; d++) {
            int64_t key = PACK(xyz[3 * i] + DIRS[d][0], xyz[3 * i + 1] + DIRS[d][1], xyz[3 * i + 2] + DIRS[d][2]);
            size_t h = (size_t)((uint64_t)key * 0x9E3779B97F4A7C15ull >> 20) & (cap - 1);
            int j = -1;
            while (keys[h] != -1) { if (keys[h] == key) { j = vals[h]; break; } h = (h + 1) & (cap - 1); }
            nb6[6 * i + d] = j;
        }
    }
    #undef PACK
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            int j = nb6[6 * i + 2 * a];
            int gap = j >= 0 && infill[i] != infill[j] && (a != 1 || infill[i] == 1);
            bond[4 * i + a] = (j >= 0 && !gap && mat[j] != 0) ? bond_cap(mat[i], mat[j], a) : 0.0f;
        }
        bond[4 * i + 3] = xyz[3 * i + 1] == 0 ? CAPV[mat[i]] * 2.0f : 0.0f;
    }
    free(keys); free(vals);
}

static inline float minf_(float a, float b) { return a < b ? a : b; }
static inline float maxf_(float a, float b) { return a > b ? a : b; }

/* lib.rs:378-383: neighbour in direction d joined by an intact bond (the bond of the lower brick of the pair) */
static inline int linked(const ol_struct* s, int i, int d, float* hp_out) {
    int j = s->nb6[6 * i + d];
    if (j < 0) return -1;
    float hp = (d & 1) == 0 ? s->bond[4 * i + (d >> 1)] : s->bond[4 * j + (d >> 1)];
    if (hp > 0.0f) { if (hp_out) *hp_out = hp; return j; }
    return -1;
}
/* lib.rs:1142 / :1146-1151 for standing bricks */
static inline float cond(int mi, int mj, int axis) { return minf_(STIFF[mi], STIFF[mj]) * (axis == 1 ? 1.0f : 0.3f); }
static inline int grounded_footing(const ol_struct* s, int i) { return s->xyz[3 * i + 1] == 0 && s->bond[4 * i + 3] > 0.0f; }

/* lib.rs:1005-1026 (standing brick) */
static float sor_brick(const ol_struct* s, float* phi, int i, float omega) {
    int mi = s->mat[i];
    float sk = 0.0f, skp = 0.0f;
    for (int d = 0; d < 6; d++) {
        int j = linked(s, i, d, NULL);
        if (j < 0) continue;
        float k = cond(mi, s->mat[j], d >> 1);
        sk += k; skp += k * phi[j];
    }
    if (grounded_footing(s, i)) sk += GROUNDK * STIFF[mi];
    if (sk > 0.0f) {
        float target = (MASS[mi] + (s->ext ? s->ext[i] : 0.0f) + skp) / sk;
        float d = omega * (target - phi[i]);
        phi[i] += d;
        return fabsf(d);
    }
    return 0.0f;
}

/* lib.rs:1029-1040 */
float ol_sor_sweep(const ol_struct* s, float* phi, float omega) {
    float md = 0.0f, mp = 1.0f;
    for (int i = 0; i < s->n; i++) {
        float d = sor_brick(s, phi, i, omega);
        if (d > md) md = d;
        if (phi[i] > mp) mp = phi[i];
    }
    return md / mp;
}

/* lib.rs:1085-1098: breadth-first search from grounded footings through linked bonds */
int ol_grounded(const ol_struct* s, uint8_t* out) {
    int* q = (int*)malloc(sizeof(int) * (size_t)(s->n ? s->n : 1));
    int qh = 0, qt = 0, cnt = 0;
    memset(out, 0, (size_t)s->n);
    for (int i = 0; i < s->n; i++) if (grounded_footing(s, i)) { out[i] = 1; q[qt++] = i; }
    while (qh < qt) {
        int i = q[qh++];
        for (int d = 0; d < 6; d++) { int j = linked(s, i, d, NULL); if (j >= 0 && !out[j]) { out[j] = 1; q[qt++] = j; } }
    }
    cnt = qt;
    free(q);
    return cnt;
}

/* y = A x over the grounded bricks (others are held at zero) */
static void apply_a(const ol_struct* s, const uint8_t* g, const double* x, double* y) {
    for (int i = 0; i < s->n; i++) {
        if (!g[i]) { y[i] = 0.0; continue; }
        int mi = s->mat[i];
        double acc = 0.0, sk = 0.0;
        for (int d = 0; d < 6; d++) {
            int j = linked(s, i, d, NULL);
            if (j < 0) continue;
            double k = (double)cond(mi, s->mat[j], d >> 1);
            sk += k; acc -= k * x[j];
        }
        if (grounded_footing(s, i)) sk += (double)(GROUNDK * STIFF[mi]);
        y[i] = sk * x[i] + acc;
    }
}
static double diag_a(const ol_struct* s, int i) {
    int mi = s->mat[i]; double sk = 0.0;
    for (int d = 0; d < 6; d++) { int j = linked(s, i, d, NULL); if (j >= 0) sk += (double)cond(mi, s->mat[j], d >> 1); }
    if (grounded_footing(s, i)) sk += (double)(GROUNDK * STIFF[mi]);
    return sk;
}
static double rhs(const ol_struct* s, int i) { return (double)MASS[s->mat[i]] + (s->ext ? (double)s->ext[i] : 0.0); }

double ol_residual64(const ol_struct* s, const double* phi) {
    uint8_t* g = (uint8_t*)malloc((size_t)(s->n ? s->n : 1));
    ol_grounded(s, g);
    double* y = (double*)malloc(sizeof(double) * (size_t)(s->n ? s->n : 1));
    apply_a(s, g, phi, y);
    double rm = 0.0, bm = 0.0;
    for (int i = 0; i < s->n; i++) {
        if (!g[i]) continue;
        double b = rhs(s, i), r = fabs(b - y[i]);
        if (r > rm) rm = r;
        if (fabs(b) > bm) bm = fabs(b);
    }
    free(g); free(y);
    return bm > 0.0 ? rm / bm : 0.0;
}

/* Jacobi-preconditioned conjugate gradients in FP64 over the grounded bricks */
int ol_solve64(const ol_struct* s, double* phi, double rel_tol, int max_iter) {
    int n = s->n;
    size_t nn = (size_t)(n ? n : 1);
    uint8_t* g = (uint8_t*)malloc(nn);
    ol_grounded(s, g);
    double *r = (double*)malloc(sizeof(double) * nn), *z = (double*)malloc(sizeof(double) * nn);
    double *p = (double*)malloc(sizeof(double) * nn), *ap = (double*)malloc(sizeof(double) * nn);
    double *dinv = (double*)malloc(sizeof(double) * nn);
    double bnorm = 0.0;
    for (int i = 0; i < n; i++) {
        if (!g[i]) { phi[i] = 0.0; dinv[i] = 0.0; continue; }
        double d = diag_a(s, i); dinv[i] = d > 0.0 ? 1.0 / d : 0.0;
        double b = rhs(s, i); if (fabs(b) > bnorm) bnorm = fabs(b);
    }
    apply_a(s, g, phi, ap);
    double rz = 0.0;
    for (int i = 0; i < n; i++) { r[i] = g[i] ? rhs(s, i) - ap[i] : 0.0; z[i] = dinv[i] * r[i]; p[i] = z[i]; rz += r[i] * z[i]; }
    int it = 0, ok = -1;
    for (; it < max_iter; it++) {
        double rm = 0.0; for (int i = 0; i < n; i++) if (fabs(r[i]) > rm) rm = fabs(r[i]);
        if (bnorm == 0.0 || rm <= rel_tol * bnorm) { ok = it; break; }
        apply_a(s, g, p, ap);
        double pap = 0.0; for (int i = 0; i < n; i++) pap += p[i] * ap[i];
        if (pap <= 0.0) break;
        double alpha = rz / pap;
        for (int i = 0; i < n; i++) { phi[i] += alpha * p[i]; r[i] -= alpha * ap[i]; }
        /* every 50 iterations replace the recurrence residual by the true one */
        if ((it + 1) % 50 == 0) { apply_a(s, g, phi, ap); for (int i = 0; i < n; i++) r[i] = g[i] ? rhs(s, i) - ap[i] : 0.0; }
        double rz1 = 0.0; for (int i = 0; i < n; i++) { z[i] = dinv[i] * r[i]; rz1 += r[i] * z[i]; }
        double beta = rz1 / rz; rz = rz1;
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    free(g); free(r); free(z); free(p); free(ap); free(dinv);
    return ok;
}

/* lib.rs:1153-1227 for a standing brick (axis_up = 1 for the vertical bond, 0 otherwise; hpw = the bond) */
void ol_verdicts64(const ol_struct* s, const double* phi, double* ratio4, uint8_t* mode4, uint8_t* fail4) {
    for (int i = 0; i < s->n; i++) {
        int mi = s->mat[i];
        for (int a = 0; a < 4; a++) { ratio4[4 * i + a] = 0.0; mode4[4 * i + a] = 0; fail4[4 * i + a] = 0; }
        for (int a = 0; a < 3; a++) {
            float hp = 0.0f;
            int j = linked(s, i, 2 * a, &hp);
            if (j < 0) continue;
            double f = (double)cond(mi, s->mat[j], a) * (phi[i] - phi[j]);
            double u = a == 1 ? 1.0 : 0.0;
            double up = f > 0.0 ? (u > 0.0 ? u : 0.0) : (-u > 0.0 ? -u : 0.0);
            double tens = (double)minf_(TENS[mi], TENS[s->mat[j]]);
            double capm = 1.0 - (1.0 - tens) * up;
            double r = fabs(f) / ((double)hp * capm);
            ratio4[4 * i + a] = r;
            mode4[4 * i + a] = a == 1 ? (f > 0.0 ? 2 : 1) : 3;
            fail4[4 * i + a] = r > 1.0;
        }
        if (grounded_footing(s, i)) {
            double f = (double)(GROUNDK * STIFF[mi]) * phi[i];
            double r = f / (double)s->bond[4 * i + 3];
            ratio4[4 * i + 3] = r; mode4[4 * i + 3] = 4; fail4[4 * i + 3] = r > 1.0;
        }
    }
}

void ol_verdicts32(const ol_struct* s, const float* phi, float* ratio4, uint8_t* mode4) {
    for (int i = 0; i < s->n; i++) {
        int mi = s->mat[i];
        for (int a = 0; a < 4; a++) { ratio4[4 * i + a] = 0.0f; mode4[4 * i + a] = 0; }
        for (int a = 0; a < 3; a++) {
            float hp = 0.0f;
            int j = linked(s, i, 2 * a, &hp);
            if (j < 0) continue;
            float f = cond(mi, s->mat[j], a) * (phi[i] - phi[j]);
            float u = a == 1 ? 1.0f : 0.0f;
            float up = f > 0.0f ? maxf_(u, 0.0f) : maxf_(-u, 0.0f);
            float tens = minf_(TENS[mi], TENS[s->mat[j]]);
            float capm = 1.0f - (1.0f - tens) * up;
            float r = fabsf(f) / (hp * capm);
            ratio4[4 * i + a] = r;
            mode4[4 * i + a] = a == 1 ? (f > 0.0f ? 2 : 1) : 3;
        }
        if (grounded_footing(s, i)) {
            float f = GROUNDK * STIFF[mi] * phi[i];
            ratio4[4 * i + 3] = f / s->bond[4 * i + 3]; mode4[4 * i + 3] = 4;
        }
    }
}
