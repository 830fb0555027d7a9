/* Spring-network structural analysis: C restatement of engine/src/structure.rs:191-783 over the brick store of
 * engine/src/brick.rs (boxes on the 0.125 m tick grid, joints as shared faces with exact areas).
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): tests, the smoke check and the benchmark's CPU-baseline leg load it as the
 * checker; the product library never links or calls it.
 *
 * Restated, in the reference's operation order (f64, no FMA contraction; f32 where the reference computes in f32):
 *   - Bricks::add / index / query / build_joints / face_contact / add_joint / remove_joint / remove
 *     (brick.rs:53-66, :157-198, :259-281, :284-344, :447-458), without the fracture-cell (hull) paths: every brick
 *     here is a whole box (hull 0, vol_share 1);
 *   - analyse_warm (structure.rs:191-424): cells cut at 2 m and at the edges of the supports a brick rests on,
 *     links inside bricks and across joints, point loads, liveness from the supports, springs, the solve, and the
 *     checks of joints (sub-faces summed about the joint's centre) and of faces inside bricks;
 *   - spring / section_ratio / solve_network / Level / vcycle / invert6 (structure.rs:432-764).
 * Two orders, differing only in how a dot product sums: OS_ORDER_REFERENCE sums element after element
 * (structure.rs:611); OS_ORDER_DEVICE sums as the device's one-workgroup solve does (1,024 strided partial sums,
 * then a halving tree).
 * The one place the reference's order cannot be restated: the coarse levels' coupling list comes out of a hash map's
 * iteration (structure.rs:726, hashbrown 0.15.5); here it is in the order the pairs first appear, which changes
 * only the order of additions in the coarse levels' products. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sf_oracle.h"

#define NONE 0xffffffffu
#define TICKF 0.125f
#define TK 0.125
#define CELL 16
#define LEVELS 4
enum { M_CONCRETE = 0, M_MASONRY = 1, M_STEEL = 2, M_GLASS = 3, M_TIMBER = 4 };
enum { F_ALIVE = 1, F_INFILL = 2, F_DYNAMIC = 4, F_MONO = 8 };
#define STEEL_AREA 0.16f
#define STEEL_SHEAR 0.10f
#define STEEL_FLEX 0.12f
#define STEEL_STIFF 0.17
#define MAX_IT 2000
#define TOL 1e-5
#define OMEGA 0.8
#define DEV_THREADS 1024

/* ---------------------------------------------------------------- growable arrays */
#define GROW(p, n, cap)                                                        \
    do {                                                                       \
        if ((n) >= (cap)) {                                                    \
            (cap) = (cap) ? 2 * (cap) : 64;                                    \
            void* q_ = realloc((p), sizeof(*(p)) * (size_t)(cap));             \
            if (!q_) abort();                                                  \
            (p) = q_;                                                          \
        }                                                                      \
    } while (0)

static void* zalloc(size_t n) { void* p = calloc(n ? n : 1, 1); if (!p) abort(); return p; }

/* ---------------------------------------------------------------- u64 -> u32 map (open addressing) */
typedef struct { uint64_t* k; uint32_t* v; uint8_t* used; size_t cap, n; } Map64;
static uint64_t mix64(uint64_t x) { x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33; return x; }
static void map_init(Map64* m, size_t cap) {
    size_t c = 64; while (c < 2 * cap) c <<= 1;
    m->cap = c; m->n = 0;
    m->k = zalloc(sizeof(uint64_t) * c); m->v = zalloc(sizeof(uint32_t) * c); m->used = zalloc(c);
}
static void map_free(Map64* m) { free(m->k); free(m->v); free(m->used); memset(m, 0, sizeof(*m)); }
static uint32_t* map_slot(Map64* m, uint64_t key, int create) {
    if (create && 2 * (m->n + 1) > m->cap) {
        Map64 o = *m;
        map_init(m, o.cap);
        for (size_t i = 0; i < o.cap; i++) if (o.used[i]) *map_slot(m, o.k[i], 1) = o.v[i];
        map_free(&o);
    }
    size_t h = mix64(key) & (m->cap - 1);
    for (;;) {
        if (!m->used[h]) {
            if (!create) return NULL;
            m->used[h] = 1; m->k[h] = key; m->v[h] = NONE; m->n++;
            return &m->v[h];
        }
        if (m->k[h] == key) return &m->v[h];
        h = (h + 1) & (m->cap - 1);
    }
}

/* ---------------------------------------------------------------- the brick store (brick.rs) */
struct os_store {
    int n, cap;
    int32_t *x, *y, *z;
    uint8_t *sx, *sy, *sz, *mat, *paint, *flags;
    uint16_t* structure;
    uint32_t *jhead, *cnext;
    int nj, jcap;
    uint32_t *ja, *jb, *jna, *jnb;
    uint16_t* jarea;
    uint8_t *jaxis, *jdmg;
    uint32_t* jfree;
    int nfree, fcap;
    int live_joints;
    Map64 cells;
};
typedef struct { int32_t mn[3], mx[3]; } TBox;

static TBox tbox(const os_store* b, uint32_t i) {
    TBox t;
    t.mn[0] = b->x[i]; t.mn[1] = b->y[i]; t.mn[2] = b->z[i];
    t.mx[0] = t.mn[0] + b->sx[i]; t.mx[1] = t.mn[1] + b->sy[i]; t.mx[2] = t.mn[2] + b->sz[i];
    return t;
}
static int32_t lcell(int l) { return 8 << (2 * l); }
static int32_t div_euclid(int32_t a, int32_t c) { int32_t q = a / c; if (a % c < 0) q -= 1; return q; }
static uint64_t ckey(int l, int32_t c0, int32_t c1, int32_t c2) {
    return ((uint64_t)l << 60) | (((uint64_t)(uint32_t)(c0 + (1 << 19)) & 0xfffffull) << 40) |
           (((uint64_t)(uint32_t)(c1 + (1 << 19)) & 0xfffffull) << 20) | ((uint64_t)(uint32_t)(c2 + (1 << 19)) & 0xfffffull);
}

os_store* os_create(void) {
    os_store* b = zalloc(sizeof(os_store));
    map_init(&b->cells, 64);
    return b;
}
void os_destroy(os_store* b) {
    if (!b) return;
    free(b->x); free(b->y); free(b->z); free(b->sx); free(b->sy); free(b->sz); free(b->mat); free(b->paint);
    free(b->flags); free(b->structure); free(b->jhead); free(b->cnext);
    free(b->ja); free(b->jb); free(b->jna); free(b->jnb); free(b->jarea); free(b->jaxis); free(b->jdmg); free(b->jfree);
    map_free(&b->cells);
    free(b);
}
int os_len(const os_store* b) { return b->n; }

/* the cell of the brick's min corner at the finest level whose cell is at least the brick's largest side
 * (brick.rs:171-180) */
static void os_index(os_store* b, uint32_t i) {
    TBox t = tbox(b, i);
    int32_t m = 0;
    for (int k = 0; k < 3; k++) if (t.mx[k] - t.mn[k] > m) m = t.mx[k] - t.mn[k];
    int l = LEVELS - 1;
    for (int q = 0; q < LEVELS; q++) if (lcell(q) >= m) { l = q; break; }
    int32_t c = lcell(l);
    uint32_t* head = map_slot(&b->cells, ckey(l, div_euclid(t.mn[0], c), div_euclid(t.mn[1], c), div_euclid(t.mn[2], c)), 1);
    b->cnext[i] = *head;
    *head = i;
}

int os_add(os_store* b, const int32_t* mn, const int32_t* size, int mat, int paint, int structure) {
    for (int k = 0; k < 3; k++) if (size[k] < 1 || size[k] > 255) return -1;
    if (b->n >= b->cap) {
        int c = b->cap ? 2 * b->cap : 256;
#define RE(p) do { void* q_ = realloc(p, sizeof(*(p)) * (size_t)c); if (!q_) abort(); p = q_; } while (0)
        RE(b->x); RE(b->y); RE(b->z); RE(b->sx); RE(b->sy); RE(b->sz); RE(b->mat); RE(b->paint); RE(b->flags);
        RE(b->structure); RE(b->jhead); RE(b->cnext);
#undef RE
        b->cap = c;
    }
    uint32_t i = (uint32_t)b->n++;
    b->x[i] = mn[0]; b->y[i] = mn[1]; b->z[i] = mn[2];
    b->sx[i] = (uint8_t)size[0]; b->sy[i] = (uint8_t)size[1]; b->sz[i] = (uint8_t)size[2];
    b->mat[i] = (uint8_t)mat; b->paint[i] = (uint8_t)paint; b->flags[i] = F_ALIVE; b->structure[i] = (uint16_t)structure;
    b->jhead[i] = NONE; b->cnext[i] = NONE;
    os_index(b, i);
    return (int)i;
}
void os_set_flags(os_store* b, int i, int flags) { if (i >= 0 && i < b->n) b->flags[i] = (uint8_t)flags; }
int os_flags(const os_store* b, int i) { return (i >= 0 && i < b->n) ? b->flags[i] : -1; }

static uint32_t joint_next(const os_store* b, uint32_t j, uint32_t k) { return b->ja[j] == k ? b->jna[j] : b->jnb[j]; }
static void set_next(os_store* b, uint32_t j, uint32_t k, uint32_t n) { if (b->ja[j] == k) b->jna[j] = n; else b->jnb[j] = n; }
static uint32_t joint_other(const os_store* b, uint32_t j, uint32_t k) { return b->ja[j] == k ? b->jb[j] : b->ja[j]; }

static uint32_t add_joint(os_store* b, uint32_t a, uint32_t c, int axis, uint32_t area, int dmg) {
    uint32_t j;
    if (b->nfree > 0) j = b->jfree[--b->nfree];
    else {
        if (b->nj >= b->jcap) {
            int cc = b->jcap ? 2 * b->jcap : 256;
#define RE(p) do { void* q_ = realloc(p, sizeof(*(p)) * (size_t)cc); if (!q_) abort(); p = q_; } while (0)
            RE(b->ja); RE(b->jb); RE(b->jna); RE(b->jnb); RE(b->jarea); RE(b->jaxis); RE(b->jdmg);
#undef RE
            b->jcap = cc;
        }
        j = (uint32_t)b->nj++;
    }
    b->ja[j] = a; b->jb[j] = c; b->jarea[j] = (uint16_t)area; b->jaxis[j] = (uint8_t)axis; b->jdmg[j] = (uint8_t)dmg;
    b->jna[j] = b->jhead[a]; b->jhead[a] = j;
    b->jnb[j] = b->jhead[c]; b->jhead[c] = j;
    b->live_joints++;
    return j;
}
static void unlink_joint(os_store* b, uint32_t j, uint32_t k) {
    uint32_t prev = NONE, cur = b->jhead[k];
    while (cur != j) { prev = cur; cur = joint_next(b, cur, k); }
    uint32_t nxt = joint_next(b, j, k);
    if (prev == NONE) b->jhead[k] = nxt; else set_next(b, prev, k, nxt);
}
static void remove_joint(os_store* b, uint32_t j) {
    uint32_t a = b->ja[j], c = b->jb[j];
    unlink_joint(b, j, a); unlink_joint(b, j, c);
    b->ja[j] = NONE; b->jb[j] = NONE;
    GROW(b->jfree, b->nfree, b->fcap);
    b->jfree[b->nfree++] = j;
    b->live_joints--;
}
/* brick.rs:182-190 (the fracture-surface marks of the neighbours are left out: no analysis reads them) */
void os_remove(os_store* b, int i) {
    if (i < 0 || i >= b->n) return;
    while (b->jhead[i] != NONE) remove_joint(b, b->jhead[i]);
    b->flags[i] &= (uint8_t)~F_ALIVE;
}

/* brick.rs:53-66 */
static int face_contact(const TBox* p, const TBox* q, int* ax_out, uint32_t* area, int* p_low_out) {
    for (int ax = 0; ax < 3; ax++) {
        int p_low = p->mx[ax] == q->mn[ax];
        int q_low = q->mx[ax] == p->mn[ax];
        if (!p_low && !q_low) continue;
        int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
        int32_t a = (p->mx[o1] < q->mx[o1] ? p->mx[o1] : q->mx[o1]) - (p->mn[o1] > q->mn[o1] ? p->mn[o1] : q->mn[o1]);
        int32_t c = (p->mx[o2] < q->mx[o2] ? p->mx[o2] : q->mx[o2]) - (p->mn[o2] > q->mn[o2] ? p->mn[o2] : q->mn[o2]);
        if (a > 0 && c > 0) { *ax_out = ax; *area = (uint32_t)(a * c); *p_low_out = p_low; return 1; }
        return 0;
    }
    return 0;
}
/* brick.rs:259-281: per level, the cells that can hold the min corner of a touching brick, each cell's list newest
 * first */
static int query(const os_store* b, const int32_t* mn, const int32_t* mx, uint32_t** out, int* cap) {
    int n = 0;
    const int t = 1;
    for (int l = 0; l < LEVELS; l++) {
        int32_t c = lcell(l);
        int32_t c0[3], c1[3];
        for (int k = 0; k < 3; k++) { c0[k] = div_euclid(mn[k] - c - t, c); c1[k] = div_euclid(mx[k] - 1 + t, c); }
        for (int32_t cx = c0[0]; cx <= c1[0]; cx++)
            for (int32_t cy = c0[1]; cy <= c1[1]; cy++)
                for (int32_t cz = c0[2]; cz <= c1[2]; cz++) {
                    uint32_t* h = map_slot((Map64*)&b->cells, ckey(l, cx, cy, cz), 0);
                    if (!h) continue;
                    for (uint32_t i = *h; i != NONE; i = b->cnext[i]) {
                        if ((b->flags[i] & (F_ALIVE | F_DYNAMIC)) != F_ALIVE) continue;
                        TBox o = tbox(b, i);
                        int ok = 1;
                        for (int k = 0; k < 3; k++) if (!(o.mn[k] <= mx[k] && o.mx[k] >= mn[k])) ok = 0;
                        if (ok) { GROW(*out, n, *cap); (*out)[n++] = i; }
                    }
                }
    }
    return n;
}
/* brick.rs:447-458 */
int os_build_joints(os_store* b) {
    b->nj = 0; b->nfree = 0; b->live_joints = 0;
    for (int i = 0; i < b->n; i++) b->jhead[i] = NONE;
    uint32_t* cand = NULL; int ccap = 0;
    for (int i = 0; i < b->n; i++) {
        if (!(b->flags[i] & F_ALIVE)) continue;
        TBox t = tbox(b, (uint32_t)i);
        int nc = query(b, t.mn, t.mx, &cand, &ccap);
        for (int k = 0; k < nc; k++) {
            uint32_t o = cand[k];
            if (o <= (uint32_t)i) continue;
            TBox q = tbox(b, o);
            int ax, pl; uint32_t area;
            if (face_contact(&t, &q, &ax, &area, &pl)) {
                if (pl) add_joint(b, (uint32_t)i, o, ax, area, 0); else add_joint(b, o, (uint32_t)i, ax, area, 0);
            }
        }
    }
    free(cand);
    return b->live_joints;
}
int os_njoint_slots(const os_store* b) { return b->nj; }
int os_live_joints(const os_store* b) { return b->live_joints; }
/* per joint slot: ja, jb (-1: a free slot), axis, area (tick^2), damage */
int os_joints(const os_store* b, int32_t* out5, int cap) {
    for (int j = 0; j < b->nj && j < cap; j++) {
        out5[5 * j + 0] = b->ja[j] == NONE ? -1 : (int32_t)b->ja[j];
        out5[5 * j + 1] = b->jb[j] == NONE ? -1 : (int32_t)b->jb[j];
        out5[5 * j + 2] = b->jaxis[j]; out5[5 * j + 3] = b->jarea[j]; out5[5 * j + 4] = b->jdmg[j];
    }
    return b->nj;
}
void os_set_joint_damage(os_store* b, int j, int dmg) { if (j >= 0 && j < b->nj) b->jdmg[j] = (uint8_t)dmg; }

/* ---------------------------------------------------------------- materials (structure.rs:20-113) */
static float density(int m) { switch (m) { case M_CONCRETE: return 2400.0f; case M_MASONRY: return 1900.0f; case M_STEEL: return 7850.0f; case M_GLASS: return 2500.0f; default: return 500.0f; } }
static void strength(int m, float* s) {
    switch (m) {
    case M_CONCRETE: s[0] = 30e6f; s[1] = 4e6f; s[2] = 4e6f; return;
    case M_MASONRY: s[0] = 10e6f; s[1] = 0.3e6f; s[2] = 0.5e6f; return;
    case M_STEEL: s[0] = 250e6f * STEEL_AREA; s[1] = 250e6f * STEEL_AREA; s[2] = 150e6f * STEEL_SHEAR; return;
    case M_GLASS: s[0] = 50e6f; s[1] = 7e6f; s[2] = 7e6f; return;
    default: s[0] = 21e6f; s[1] = 14e6f; s[2] = 4e6f; return;
    }
}
static float flexural(int m) { switch (m) { case M_CONCRETE: return 12e6f; case M_MASONRY: return 0.5e6f; case M_STEEL: return 250e6f * STEEL_FLEX; case M_GLASS: return 40e6f; default: return 20e6f; } }
static float flexural_of(const os_store* b, uint32_t i) { return (b->mat[i] == M_TIMBER && b->paint[i] == 8) ? 2.4e6f : flexural(b->mat[i]); }
static int tfloor(const os_store* b, uint32_t i) { return b->mat[i] == M_TIMBER && b->paint[i] == 8; }
static int hinge(const os_store* b, uint32_t x, uint32_t y) {
    int mx = b->mat[x], my = b->mat[y];
    return ((tfloor(b, x) || tfloor(b, y)) && mx != my) || (mx == M_STEEL && my == M_MASONRY) || (mx == M_MASONRY && my == M_STEEL);
}
static void connection(const os_store* b, uint32_t j, float* c) {
    uint32_t a = b->ja[j], o = b->jb[j];
    int ma = b->mat[a], mc = b->mat[o], pa = b->paint[a], pc = b->paint[o];
    c[0] = 1.0f; c[1] = 0.7f; c[2] = 0.7f;
    if (pa == pc && ma == mc) { c[1] = 1.0f; c[2] = 1.0f; return; }
    if (ma == M_TIMBER && mc == M_TIMBER) { c[1] = 0.1f; c[2] = 0.25f; return; }
    if (ma == M_CONCRETE && mc == M_CONCRETE) {
        int sa = pa == 0, sc = pc == 0, ua = pa == 1 || pa == 3, uc = pc == 1 || pc == 3;
        if ((sa && uc) || (sc && ua)) { c[1] = 0.5f; c[2] = 0.45f; return; }
    }
}
static double modulus_of(const os_store* b, uint32_t i) {
    switch (b->mat[i]) {
    case M_CONCRETE: return 30e9; case M_MASONRY: return 5e9; case M_STEEL: return 200e9 * STEEL_STIFF; case M_GLASS: return 70e9;
    default: return b->paint[i] == 8 ? 1e9 : 11e9;
    }
}
/* structure.rs:66-71 */
float os_mass(const os_store* b, int i) {
    float share = b->mat[i] == M_STEEL ? STEEL_AREA : 1.0f;
    return (float)(b->sx[i] * b->sy[i] * b->sz[i]) * TICKF * TICKF * TICKF * density(b->mat[i]) * (float)1.0 * share;
}

/* ---------------------------------------------------------------- sections and springs */
typedef struct { double area, iu, iw, zu, zw, j, rmax; } Sect;
static Sect sect_rect(double eu, double ew, double area) {
    eu = fmax(eu, 1e-3); ew = fmax(ew, 1e-3);
    Sect s;
    s.area = fmax(area, 1e-6);
    s.iu = eu * ew * ew * ew / 12.0; s.iw = ew * eu * eu * eu / 12.0;
    s.zu = ew / 2.0; s.zw = eu / 2.0;
    s.j = eu * ew * (eu * eu + ew * ew) / 12.0;
    s.rmax = 0.5 * sqrt(eu * eu + ew * ew);
    return s;
}
typedef struct { uint32_t brick; int32_t lo[3], hi[3]; float load; int sink; double c[3]; } Node;
typedef struct { uint32_t a, b, joint; int axis; int32_t flo[3], fhi[3]; } Link;
typedef struct { double c[3], ra[3], rb[3], k[3], kr[3]; } Spring;

/* structure.rs:465-524 (axis-plane faces only) */
static Spring make_spring(const os_store* b, const Node* nodes, const Link* l) {
    const Node *na = &nodes[l->a], *nb = &nodes[l->b];
    const double t = TK;
    int ax = l->axis, u = (ax + 1) % 3, w = (ax + 2) % 3;
    double eu = (double)(l->fhi[u] - l->flo[u]) * t, ew = (double)(l->fhi[w] - l->flo[w]) * t;
    double c[3];
    for (int d = 0; d < 3; d++) c[d] = 0.5 * (double)(l->flo[d] + l->fhi[d]) * t;
    Sect sec = sect_rect(eu, ew, eu * ew);
    double n[3] = {0.0, 0.0, 0.0};
    n[ax] = 1.0;
    double la = fmax(fabs((c[0] - na->c[0]) * n[0] + (c[1] - na->c[1]) * n[1] + (c[2] - na->c[2]) * n[2]), 0.5 * t);
    double lb = fmax(fabs((c[0] - nb->c[0]) * n[0] + (c[1] - nb->c[1]) * n[1] + (c[2] - nb->c[2]) * n[2]), 0.5 * t);
    double ea = modulus_of(b, na->brick), eb = modulus_of(b, nb->brick);
    double comp = la / ea + lb / eb;
    double gcomp = comp / 0.4;
    double kl[3] = {sec.area / comp, sec.area / gcomp, sec.area / gcomp};
    double krl[3] = {sec.j / gcomp, sec.iu / comp, sec.iw / comp};
    Spring s;
    s.k[ax] = kl[0]; s.k[u] = kl[1]; s.k[w] = kl[2];
    s.kr[ax] = krl[0]; s.kr[u] = krl[1]; s.kr[w] = krl[2];
    uint32_t ia = na->brick, ib = nb->brick;
    int infill = ((b->flags[ia] | b->flags[ib]) & F_INFILL) != 0;
    int mono = ((b->flags[ia] | b->flags[ib]) & F_MONO) != 0;
    int mi = b->mat[ia], mo = b->mat[ib];
    int pocket = l->joint != NONE && ax != 1 && ((mi == M_STEEL && mo == M_MASONRY) || (mi == M_MASONRY && mo == M_STEEL));
    double f = (infill || pocket) ? 0.02 : (mono ? 0.05 : 1.0);
    if (l->joint != NONE) {
        int dm = b->jdmg[l->joint];
        f *= ((dm & 1) ? 0.5 : 1.0) * (1.0 - 0.6 * (double)(dm >> 1) / 127.0);
    }
    for (int i = 0; i < 3; i++) { s.k[i] *= f; s.kr[i] *= f; }
    if (l->joint != NONE && hinge(b, ia, ib)) { s.kr[u] = 0.0; s.kr[w] = 0.0; }
    if (tfloor(b, ia) && tfloor(b, ib) && ax == 2) { s.kr[0] *= 0.05; s.kr[2] *= 0.05; }
    if (tfloor(b, ia) && tfloor(b, ib) && ax == 0) { s.kr[0] *= 0.05; }
    for (int d = 0; d < 3; d++) { s.c[d] = c[d]; s.ra[d] = c[d] - na->c[d]; s.rb[d] = c[d] - nb->c[d]; }
    return s;
}
/* relative displacement at the face centre and relative rotation (structure.rs:453-457) */
static void strain(const Spring* s, const double* xa, const double* xb, double* d, double* r) {
    const double *A = s->ra, *B = s->rb;
    double pa[3] = {xa[0] + xa[4] * A[2] - xa[5] * A[1], xa[1] + xa[5] * A[0] - xa[3] * A[2], xa[2] + xa[3] * A[1] - xa[4] * A[0]};
    double pb[3] = {xb[0] + xb[4] * B[2] - xb[5] * B[1], xb[1] + xb[5] * B[0] - xb[3] * B[2], xb[2] + xb[3] * B[1] - xb[4] * B[0]};
    for (int i = 0; i < 3; i++) { d[i] = pb[i] - pa[i]; r[i] = xb[3 + i] - xa[3 + i]; }
}
static void force(const Spring* s, const double* d, const double* r, double* f, double* m) {
    for (int i = 0; i < 3; i++) { f[i] = s->k[i] * d[i]; m[i] = s->kr[i] * r[i]; }
}
/* structure.rs:532-550 */
static float section_ratio(const Sect* s, const double* f, const double* m, const float* cap, float flex) {
    double sn = f[0] / s->area;
    double bend = fabs(m[1]) * s->zu / s->iu + fabs(m[2]) * s->zw / s->iw;
    double top = sn + bend, bottom = sn - bend;
    double r = 0.0;
    if (top > 0.0) {
        double tens = fmax(sn, 0.0);
        double fe = (tens * (double)cap[1] + bend * (double)flex) / fmax(tens + bend, 1e-9);
        r = fmax(r, top / fmax(fe, 1.0));
    }
    if (bottom < 0.0) r = fmax(r, -bottom / (double)cap[0]);
    double v = sqrt(f[1] * f[1] + f[2] * f[2]) / s->area;
    double tor = fabs(m[0]) * s->rmax / s->j;
    double cohesion = (double)cap[2] + 0.4 * fmax(-sn, 0.0);
    r = fmax(r, (v + tor) / cohesion);
    return (float)r;
}

/* ---------------------------------------------------------------- the solve (structure.rs:560-764) */
typedef double M6[36];
static void mmul(const double* a, const double* b, double* c) {
    for (int i = 0; i < 36; i++) c[i] = 0.0;
    for (int i = 0; i < 6; i++)
        for (int k = 0; k < 6; k++) {
            double v = a[6 * i + k];
            if (v != 0.0) for (int j = 0; j < 6; j++) c[6 * i + j] += v * b[6 * k + j];
        }
}
static void trm(const double* a, double* c) { for (int i = 0; i < 6; i++) for (int j = 0; j < 6; j++) c[6 * j + i] = a[6 * i + j]; }
static void rigid(const double* d, double* t) {
    for (int i = 0; i < 36; i++) t[i] = 0.0;
    for (int i = 0; i < 6; i++) t[7 * i] = 1.0;
    t[0 * 6 + 4] = d[2]; t[0 * 6 + 5] = -d[1]; t[1 * 6 + 3] = -d[2]; t[1 * 6 + 5] = d[0]; t[2 * 6 + 3] = d[1]; t[2 * 6 + 4] = -d[0];
}
static void invert6(const double* m0, double* inv) {
    double m[36];
    memcpy(m, m0, sizeof(m));
    for (int i = 0; i < 36; i++) inv[i] = 0.0;
    for (int i = 0; i < 6; i++) inv[7 * i] = 1.0;
    for (int c = 0; c < 6; c++) {
        int p = c;
        for (int r = c + 1; r < 6; r++) if (fabs(m[6 * r + c]) > fabs(m[6 * p + c])) p = r;
        if (fabs(m[6 * p + c]) < 1e-300) continue;
        if (p != c) for (int j = 0; j < 6; j++) { double t = m[6 * c + j]; m[6 * c + j] = m[6 * p + j]; m[6 * p + j] = t; t = inv[6 * c + j]; inv[6 * c + j] = inv[6 * p + j]; inv[6 * p + j] = t; }
        double d = m[6 * c + c];
        for (int j = 0; j < 6; j++) { m[6 * c + j] /= d; inv[6 * c + j] /= d; }
        for (int r = 0; r < 6; r++) {
            if (r == c) continue;
            double q = m[6 * r + c];
            if (q != 0.0) for (int j = 0; j < 6; j++) { m[6 * r + j] -= q * m[6 * c + j]; inv[6 * r + j] -= q * inv[6 * c + j]; }
        }
    }
}

typedef struct {
    int n;
    double* diag;     /* n x 36 */
    int noff, offcap;
    uint32_t *oa, *oc;
    double* om;       /* noff x 36 */
    uint8_t* fixed;
    double* pos;      /* n x 3 */
    double* inv;      /* n x 36 */
    uint32_t* up;     /* the finer level's cell -> this level's cell (on the coarse level) */
    double* reg;
    /* the finest level keeps its springs */
    int nsp;
    const Spring* sp;
    const uint32_t* ends;   /* 2 per spring */
} Level;

static void level_free(Level* l) { free(l->diag); free(l->oa); free(l->oc); free(l->om); free(l->fixed); free(l->pos); free(l->inv); free(l->up); free(l->reg); memset(l, 0, sizeof(*l)); }
static void push_off(Level* l, uint32_t a, uint32_t c, const double* m) {
    if (l->noff >= l->offcap) {
        int cc = l->offcap ? 2 * l->offcap : 256;
        l->oa = realloc(l->oa, sizeof(uint32_t) * (size_t)cc); l->oc = realloc(l->oc, sizeof(uint32_t) * (size_t)cc);
        l->om = realloc(l->om, sizeof(double) * 36 * (size_t)cc);
        if (!l->oa || !l->oc || !l->om) abort();
        l->offcap = cc;
    }
    l->oa[l->noff] = a; l->oc[l->noff] = c; memcpy(l->om + 36 * (size_t)l->noff, m, sizeof(double) * 36);
    l->noff++;
}
static void level_finish(Level* l) {
    l->inv = zalloc(sizeof(double) * 36 * (size_t)l->n);
    l->reg = zalloc(sizeof(double) * (size_t)l->n);
    for (int k = 0; k < l->n; k++) {
        if (l->fixed[k]) continue;
        double* d = l->diag + 36 * (size_t)k;
        double big = 0.0;
        for (int i = 0; i < 6; i++) big = fmax(big, fabs(d[7 * i]));
        big = fmax(big, 1.0);
        l->reg[k] = 1e-7 * big;
        for (int i = 0; i < 6; i++) d[7 * i] += l->reg[k];
        invert6(d, l->inv + 36 * (size_t)k);
    }
}
static void level_from_springs(Level* l, int n, const Spring* sp, const uint32_t* ends, int nsp, const uint8_t* fixed, const double* pos) {
    memset(l, 0, sizeof(*l));
    l->n = n;
    l->diag = zalloc(sizeof(double) * 36 * (size_t)n);
    l->fixed = zalloc((size_t)n); memcpy(l->fixed, fixed, (size_t)n);
    l->pos = zalloc(sizeof(double) * 3 * (size_t)n); memcpy(l->pos, pos, sizeof(double) * 3 * (size_t)n);
    for (int s = 0; s < nsp; s++) {
        uint32_t a = ends[2 * s], c = ends[2 * s + 1];
        double d[36], ba[36], bb[36], dba[36], dbb[36], t[36], m[36];
        for (int i = 0; i < 36; i++) d[i] = 0.0;
        for (int i = 0; i < 3; i++) { d[7 * i] = sp[s].k[i]; d[7 * (3 + i)] = sp[s].kr[i]; }
        rigid(sp[s].ra, ba); rigid(sp[s].rb, bb);
        mmul(d, ba, dba); mmul(d, bb, dbb);
        if (!fixed[a]) { trm(ba, t); mmul(t, dba, m); for (int i = 0; i < 36; i++) l->diag[36 * (size_t)a + i] += m[i]; }
        if (!fixed[c]) { trm(bb, t); mmul(t, dbb, m); for (int i = 0; i < 36; i++) l->diag[36 * (size_t)c + i] += m[i]; }
        if (!fixed[a] && !fixed[c]) { trm(ba, t); mmul(t, dbb, m); for (int i = 0; i < 36; i++) m[i] = -m[i]; push_off(l, a, c, m); }
    }
    level_finish(l);
    l->sp = sp; l->ends = ends; l->nsp = nsp;
}
/* y = K x (fixed cells: 0), structure.rs:636-660 */
static void level_apply(const Level* l, const double* x, double* y) {
    if (l->nsp > 0) {
        for (int k = 0; k < l->n; k++) for (int i = 0; i < 6; i++) y[6 * k + i] = l->reg[k] * x[6 * k + i];
        for (int s = 0; s < l->nsp; s++) {
            const Spring* S = &l->sp[s];
            uint32_t a = l->ends[2 * s], c = l->ends[2 * s + 1];
            double d[3], r[3], f[3], m[3];
            strain(S, x + 6 * (size_t)a, x + 6 * (size_t)c, d, r);
            force(S, d, r, f, m);
            const double *ra = S->ra, *rb = S->rb;
            double mb[3] = {m[0] + rb[1] * f[2] - rb[2] * f[1], m[1] + rb[2] * f[0] - rb[0] * f[2], m[2] + rb[0] * f[1] - rb[1] * f[0]};
            double ma[3] = {m[0] + ra[1] * f[2] - ra[2] * f[1], m[1] + ra[2] * f[0] - ra[0] * f[2], m[2] + ra[0] * f[1] - ra[1] * f[0]};
            for (int i = 0; i < 3; i++) {
                y[6 * c + i] += f[i]; y[6 * c + 3 + i] += mb[i];
                y[6 * a + i] -= f[i]; y[6 * a + 3 + i] -= ma[i];
            }
        }
        for (int k = 0; k < l->n; k++) if (l->fixed[k]) for (int i = 0; i < 6; i++) y[6 * k + i] = 0.0;
        return;
    }
    for (int k = 0; k < l->n; k++) {
        if (l->fixed[k]) { for (int i = 0; i < 6; i++) y[6 * k + i] = 0.0; continue; }
        const double* m = l->diag + 36 * (size_t)k;
        const double* v = x + 6 * (size_t)k;
        for (int i = 0; i < 6; i++) { double s = 0.0; for (int j = 0; j < 6; j++) s += m[6 * i + j] * v[j]; y[6 * k + i] = s; }
    }
    for (int o = 0; o < l->noff; o++) {
        uint32_t a = l->oa[o], c = l->oc[o];
        const double* m = l->om + 36 * (size_t)o;
        double xa[6], xc[6];
        memcpy(xa, x + 6 * (size_t)a, sizeof(xa)); memcpy(xc, x + 6 * (size_t)c, sizeof(xc));
        for (int i = 0; i < 6; i++) {
            double s = 0.0, t = 0.0;
            for (int j = 0; j < 6; j++) { s += m[6 * i + j] * xc[j]; t += m[6 * j + i] * xa[j]; }
            y[6 * a + i] += s; y[6 * c + i] += t;
        }
    }
}
static void level_jacobi(const Level* l, const double* r, double* z, double w) {
    for (int k = 0; k < l->n; k++) {
        if (l->fixed[k]) continue;
        const double* iv = l->inv + 36 * (size_t)k;
        for (int i = 0; i < 6; i++) { double s = 0.0; for (int j = 0; j < 6; j++) s += iv[6 * i + j] * r[6 * k + j]; z[6 * k + i] += w * s; }
    }
}
/* structure.rs:686-722 */
static void level_coarsen(const Level* f, double bin, Level* l) {
    memset(l, 0, sizeof(*l));
    int n = f->n;
    Map64 map; map_init(&map, 64);
    uint32_t* up = zalloc(sizeof(uint32_t) * (size_t)n);
    int nc = 0, pcap = 0;
    double* pos = NULL; double* cnt = NULL; int ccap = 0;
    /* the block key: three cell indices of 21 bits each in one 64-bit key (blocks of at least 3 m; a structure
     * spans far fewer than 2^20 blocks a side) */
    for (int k = 0; k < n; k++) {
        if (f->fixed[k]) { up[k] = NONE; continue; }
        const double* p = f->pos + 3 * (size_t)k;
        int64_t kx = (int64_t)floor(p[0] / bin), ky = (int64_t)floor(p[1] / bin), kz = (int64_t)floor(p[2] / bin);
        uint64_t key = (((uint64_t)(kx + (1 << 20)) & 0x1fffffull) << 42) | (((uint64_t)(ky + (1 << 20)) & 0x1fffffull) << 21) | ((uint64_t)(kz + (1 << 20)) & 0x1fffffull);
        uint32_t* s = map_slot(&map, key, 1);
        if (*s == NONE) {
            *s = (uint32_t)nc;
            if (nc >= pcap) { pcap = pcap ? 2 * pcap : 256; pos = realloc(pos, sizeof(double) * 3 * (size_t)pcap); if (!pos) abort(); }
            if (nc >= ccap) { ccap = ccap ? 2 * ccap : 256; cnt = realloc(cnt, sizeof(double) * (size_t)ccap); if (!cnt) abort(); }
            pos[3 * nc] = 0.0; pos[3 * nc + 1] = 0.0; pos[3 * nc + 2] = 0.0; cnt[nc] = 0.0;
            nc++;
        }
        uint32_t c = *s;
        up[k] = c;
        for (int d = 0; d < 3; d++) pos[3 * c + d] += p[d];
        cnt[c] += 1.0;
    }
    map_free(&map);
    for (int c = 0; c < nc; c++) for (int d = 0; d < 3; d++) pos[3 * c + d] /= cnt[c];
    free(cnt);
    double* T = zalloc(sizeof(double) * 36 * (size_t)n);
    for (int k = 0; k < n; k++) {
        if (up[k] == NONE) continue;
        uint32_t c = up[k];
        double d[3] = {f->pos[3 * k] - pos[3 * c], f->pos[3 * k + 1] - pos[3 * c + 1], f->pos[3 * k + 2] - pos[3 * c + 2]};
        rigid(d, T + 36 * (size_t)k);
    }
    l->n = nc;
    l->diag = zalloc(sizeof(double) * 36 * (size_t)(nc ? nc : 1));
    for (int k = 0; k < n; k++) {
        if (up[k] == NONE) continue;
        double a[36], b[36], t[36];
        mmul(f->diag + 36 * (size_t)k, T + 36 * (size_t)k, a);
        trm(T + 36 * (size_t)k, t);
        mmul(t, a, b);
        for (int i = 0; i < 36; i++) l->diag[36 * (size_t)up[k] + i] += b[i];
    }
    Map64 pairs; map_init(&pairs, 64);
    for (int o = 0; o < f->noff; o++) {
        uint32_t a = f->oa[o], c = f->oc[o];
        uint32_t ua = up[a], uc = up[c];
        if (ua == NONE || uc == NONE) continue;
        double q[36], t[36], mt[36];
        mmul(f->om + 36 * (size_t)o, T + 36 * (size_t)c, mt);
        trm(T + 36 * (size_t)a, t);
        mmul(t, mt, q);
        if (ua == uc) {
            double qt[36];
            trm(q, qt);
            for (int i = 0; i < 36; i++) l->diag[36 * (size_t)ua + i] += q[i];
            for (int i = 0; i < 36; i++) l->diag[36 * (size_t)ua + i] += qt[i];
            continue;
        }
        uint32_t lo = ua < uc ? ua : uc, hi = ua < uc ? uc : ua;
        uint32_t* s = map_slot(&pairs, ((uint64_t)lo << 32) | hi, 1);
        if (*s == NONE) { double z[36] = {0}; *s = (uint32_t)l->noff; push_off(l, lo, hi, z); }
        double* acc = l->om + 36 * (size_t)*s;
        if (ua < uc) { for (int i = 0; i < 36; i++) acc[i] += q[i]; }
        else { double qt[36]; trm(q, qt); for (int i = 0; i < 36; i++) acc[i] += qt[i]; }
    }
    map_free(&pairs);
    free(T);
    l->fixed = zalloc((size_t)(nc ? nc : 1));
    l->pos = pos ? pos : zalloc(sizeof(double) * 3);
    level_finish(l);
    l->up = up;
}

typedef struct { Level lv[8]; int nl; int order; } Net;

static double dot(const Net* N, const double* a, const double* b, int n6) {
    if (N->order == OS_ORDER_REFERENCE) {
        double s = 0.0;
        for (int e = 0; e < n6; e++) s += a[e] * b[e];
        return s;
    }
    static double part[DEV_THREADS];
    for (int t = 0; t < DEV_THREADS; t++) {
        double s = 0.0;
        for (int e = t; e < n6; e += DEV_THREADS) s += a[e] * b[e];
        part[t] = s;
    }
    for (int st = DEV_THREADS / 2; st > 0; st >>= 1)
        for (int t = 0; t < st; t++) part[t] += part[t + st];
    return part[0];
}
/* structure.rs:725-764 */
static void vcycle(const Net* N, int li, const double* r, double* z) {
    const Level* lv = &N->lv[li];
    int n = lv->n;
    memset(z, 0, sizeof(double) * 6 * (size_t)n);
    double* kz = zalloc(sizeof(double) * 6 * (size_t)n);
    double* rr = zalloc(sizeof(double) * 6 * (size_t)n);
    int coarsest = li + 1 >= N->nl;
    int sweeps = coarsest ? 30 : 1;
    for (int s = 0; s < sweeps; s++) {
        if (s == 0) { level_jacobi(lv, r, z, OMEGA); continue; }
        level_apply(lv, z, kz);
        for (int e = 0; e < 6 * n; e++) rr[e] = r[e] - kz[e];
        level_jacobi(lv, rr, z, OMEGA);
    }
    if (!coarsest) {
        const Level* nx = &N->lv[li + 1];
        level_apply(lv, z, kz);
        double* rc = zalloc(sizeof(double) * 6 * (size_t)nx->n);
        for (int k = 0; k < n; k++) {
            uint32_t c = nx->up[k];
            if (c == NONE) continue;
            double v[6];
            for (int i = 0; i < 6; i++) v[i] = r[6 * k + i] - kz[6 * k + i];
            double d[3] = {lv->pos[3 * k] - nx->pos[3 * c], lv->pos[3 * k + 1] - nx->pos[3 * c + 1], lv->pos[3 * k + 2] - nx->pos[3 * c + 2]};
            for (int i = 0; i < 3; i++) rc[6 * c + i] += v[i];
            rc[6 * c + 3] += v[3] + d[1] * v[2] - d[2] * v[1];
            rc[6 * c + 4] += v[4] + d[2] * v[0] - d[0] * v[2];
            rc[6 * c + 5] += v[5] + d[0] * v[1] - d[1] * v[0];
        }
        double* zc = zalloc(sizeof(double) * 6 * (size_t)nx->n);
        vcycle(N, li + 1, rc, zc);
        for (int k = 0; k < n; k++) {
            uint32_t c = nx->up[k];
            if (c == NONE) continue;
            double d[3] = {lv->pos[3 * k] - nx->pos[3 * c], lv->pos[3 * k + 1] - nx->pos[3 * c + 1], lv->pos[3 * k + 2] - nx->pos[3 * c + 2]};
            const double* t = zc + 6 * (size_t)c + 3;
            z[6 * k + 0] += zc[6 * c + 0] + t[1] * d[2] - t[2] * d[1];
            z[6 * k + 1] += zc[6 * c + 1] + t[2] * d[0] - t[0] * d[2];
            z[6 * k + 2] += zc[6 * c + 2] + t[0] * d[1] - t[1] * d[0];
            for (int i = 3; i < 6; i++) z[6 * k + i] += zc[6 * c + i];
        }
        free(rc); free(zc);
        level_apply(lv, z, kz);
        for (int e = 0; e < 6 * n; e++) rr[e] = r[e] - kz[e];
        level_jacobi(lv, rr, z, OMEGA);
    }
    free(kz); free(rr);
}

/* structure.rs:566-617: returns the iterations; x holds the solution */
static int solve_network(Net* N, const Spring* sp, const uint32_t* ends, int nsp, const uint8_t* fixed, const double* pos,
                         const double* f, int n, double* x, const double* x0) {
    level_from_springs(&N->lv[0], n, sp, ends, nsp, fixed, pos);
    N->nl = 1;
    double bin = 3.0;
    while (N->lv[N->nl - 1].n > 64 && N->nl < 8) {
        Level c;
        level_coarsen(&N->lv[N->nl - 1], bin, &c);
        if ((double)c.n > 0.7 * (double)N->lv[N->nl - 1].n) {
            level_free(&c);
            bin *= 2.0;
            if (bin > 256.0) break;
            continue;
        }
        N->lv[N->nl++] = c;
        bin *= 2.0;
    }
    size_t n6 = 6 * (size_t)n;
    double* r = zalloc(sizeof(double) * n6);
    double* ap = zalloc(sizeof(double) * n6);
    double* z = zalloc(sizeof(double) * n6);
    double* p = zalloc(sizeof(double) * n6);
    for (int k = 0; k < n; k++) for (int i = 0; i < 6; i++) r[6 * k + i] = fixed[k] ? 0.0 : f[6 * k + i];
    vcycle(N, 0, r, z);
    double r0 = fmax(dot(N, r, z, (int)n6), 1e-30);
    if (x0) {
        memcpy(x, x0, sizeof(double) * n6);
        level_apply(&N->lv[0], x, ap);
        for (size_t e = 0; e < n6; e++) r[e] -= ap[e];
        vcycle(N, 0, r, z);
    } else memset(x, 0, sizeof(double) * n6);
    memcpy(p, z, sizeof(double) * n6);
    double rz = dot(N, r, z, (int)n6);
    int it = 0;
    if (!(rz < TOL * r0)) {
        while (it < MAX_IT) {
            level_apply(&N->lv[0], p, ap);
            double pap = dot(N, p, ap, (int)n6);
            if (pap <= 0.0) break;
            double alpha = rz / pap;
            for (size_t e = 0; e < n6; e++) { x[e] += alpha * p[e]; r[e] -= alpha * ap[e]; }
            it++;
            vcycle(N, 0, r, z);
            double rz2 = dot(N, r, z, (int)n6);
            if (rz2 < TOL * r0) break;
            double beta = rz2 / rz;
            rz = rz2;
            for (size_t e = 0; e < n6; e++) p[e] = z[e] + beta * p[e];
        }
    }
    free(r); free(ap); free(z); free(p);
    return it;
}

/* ---------------------------------------------------------------- warm starts (structure.rs:178-189) */
struct os_warm { Map64 map; double* x; int n, cap; };
os_warm* os_warm_create(void) { os_warm* w = zalloc(sizeof(os_warm)); map_init(&w->map, 64); return w; }
void os_warm_destroy(os_warm* w) { if (!w) return; map_free(&w->map); free(w->x); free(w); }
static uint64_t warm_key(uint32_t brick, int32_t x, int32_t z) {
    return ((uint64_t)brick << 32) | (((uint64_t)(uint32_t)(x + 32768) & 0xffffull) << 16) | ((uint64_t)(uint32_t)(z + 32768) & 0xffffull);
}
int os_warm_size(const os_warm* w) { return w ? w->n : 0; }

/* ---------------------------------------------------------------- analyse (structure.rs:191-424) */
struct os_result {
    float max_ratio;
    int iterations, nodes, nlinks, nlevels;
    int level_n[8];
    float worst_ratio, worst_flow, worst_moment;
    int worst_brick, worst_axis;
    int nj; int32_t* jid; float* jratio;                 /* joints checked, in order */
    int ni; int32_t* ibrick; int32_t* iaxis; int32_t* iat; float* iratio;   /* faces inside bricks */
    double* x;                                          /* node displacements (6 per node) */
    int32_t* node;                                      /* brick, lo[3], hi[3] per node */
    float* nload;
    uint8_t* nfixed;
    uint32_t* ends;                                     /* 2 per link */
    double* sp;                                         /* k[3] kr[3] ra[3] rb[3] per link */
};
void os_result_free(os_result* r) {
    if (!r) return;
    free(r->jid); free(r->jratio); free(r->ibrick); free(r->iaxis); free(r->iat); free(r->iratio); free(r->x); free(r->node);
    free(r->nload); free(r->nfixed); free(r->ends); free(r->sp); free(r);
}

static void cut(int32_t lo, int32_t hi, int32_t** v, int* nv) {
    int32_t n = (hi - lo + CELL - 1) / CELL;
    if (n < 1) n = 1;
    *v = zalloc(sizeof(int32_t) * (size_t)(n + 1 + 64));
    int k = 0;
    (*v)[k++] = lo;
    for (int32_t q = 1; q < n; q++) (*v)[k++] = lo + (hi - lo) * q / n;
    (*v)[k++] = hi;
    *nv = k;
}
static int cmp_i32(const void* a, const void* b) { int32_t x = *(const int32_t*)a, y = *(const int32_t*)b; return (x > y) - (x < y); }
static void merge_sorted(int32_t** base, int* nb, const int32_t* extra, int ne) {
    int32_t* v = zalloc(sizeof(int32_t) * (size_t)(*nb + ne));
    memcpy(v, *base, sizeof(int32_t) * (size_t)*nb);
    memcpy(v + *nb, extra, sizeof(int32_t) * (size_t)ne);
    int n = *nb + ne;
    qsort(v, (size_t)n, sizeof(int32_t), cmp_i32);
    int m = 0;
    for (int i = 0; i < n; i++) if (m == 0 || v[m - 1] != v[i]) v[m++] = v[i];
    free(*base);
    *base = v; *nb = m;
}
/* partition_point over a sorted cut list */
static int pp_le(const int32_t* c, int n, int32_t x) { int k = 0; while (k < n && c[k] <= x) k++; return k; }
static int pp_lt(const int32_t* c, int n, int32_t x) { int k = 0; while (k < n && c[k] < x) k++; return k; }
static void range_of(const int32_t* c, int nc, int32_t lo, int32_t hi, int* a, int* z) {
    int n = nc - 1;
    int aa = pp_le(c, nc, lo); aa = aa > 0 ? aa - 1 : 0; if (aa > n) aa = n;
    int zz = pp_lt(c, nc, hi); if (zz > n) zz = n;
    *a = aa; *z = zz > aa ? zz : aa;
}

os_result* os_analyse(const os_store* b, int nset, const uint32_t* set, const float* load, const uint8_t* sink,
                      const float* dir, int npoint, const float* point4, int order, os_warm* warm) {
    os_result* R = zalloc(sizeof(os_result));
    Node* nodes = NULL; int nn = 0, ncap = 0;
    Link* links = NULL; int nl = 0, lcap = 0;
    uint32_t* first = zalloc(sizeof(uint32_t) * (size_t)(b->n ? b->n : 1));
    for (int i = 0; i < b->n; i++) first[i] = NONE;
    uint32_t* slot = zalloc(sizeof(uint32_t) * (size_t)(b->n ? b->n : 1));
    int* gx = zalloc(sizeof(int) * (size_t)(nset ? nset : 1));
    int* gz = zalloc(sizeof(int) * (size_t)(nset ? nset : 1));
    int32_t** cx_ = zalloc(sizeof(int32_t*) * (size_t)(nset ? nset : 1));
    int32_t** cz_ = zalloc(sizeof(int32_t*) * (size_t)(nset ? nset : 1));
    int* ncx = zalloc(sizeof(int) * (size_t)(nset ? nset : 1));
    int* ncz = zalloc(sizeof(int) * (size_t)(nset ? nset : 1));
    for (int k = 0; k < nset; k++) slot[set[k]] = (uint32_t)k;
    for (int k = 0; k < nset; k++) {
        uint32_t i = set[k];
        TBox t = tbox(b, i);
        int infill = (b->flags[i] & F_INFILL) != 0;
        int32_t *xs, *zs; int nxs, nzs;
        if (infill) {
            xs = zalloc(sizeof(int32_t) * 2); zs = zalloc(sizeof(int32_t) * 2);
            xs[0] = t.mn[0]; xs[1] = t.mx[0]; zs[0] = t.mn[2]; zs[1] = t.mx[2]; nxs = 2; nzs = 2;
        } else {
            cut(t.mn[0], t.mx[0], &xs, &nxs);
            cut(t.mn[2], t.mx[2], &zs, &nzs);
            int32_t* ex = NULL; int nex = 0, excap = 0;
            int32_t* ez = NULL; int nez = 0, ezcap = 0;
            for (uint32_t j = b->jhead[i]; j != NONE; j = joint_next(b, j, i)) {
                uint32_t oi = joint_other(b, j, i);
                TBox o = tbox(b, oi);
                if (b->jaxis[j] == 1 && o.mx[1] == t.mn[1] && !(b->flags[oi] & F_INFILL)) {
                    int32_t vx[2] = {o.mn[0], o.mx[0]}, vz[2] = {o.mn[2], o.mx[2]};
                    for (int q = 0; q < 2; q++) if (vx[q] > t.mn[0] && vx[q] < t.mx[0]) { GROW(ex, nex, excap); ex[nex++] = vx[q]; }
                    for (int q = 0; q < 2; q++) if (vz[q] > t.mn[2] && vz[q] < t.mx[2]) { GROW(ez, nez, ezcap); ez[nez++] = vz[q]; }
                }
            }
            merge_sorted(&xs, &nxs, ex, nex);
            merge_sorted(&zs, &nzs, ez, nez);
            free(ex); free(ez);
        }
        first[i] = (uint32_t)nn;
        gx[k] = nxs - 1; gz[k] = nzs - 1;
        cx_[k] = xs; cz_[k] = zs; ncx[k] = nxs; ncz[k] = nzs;
        float vol = (float)((t.mx[0] - t.mn[0]) * (t.mx[2] - t.mn[2]));
        for (int cz = 0; cz < nzs - 1; cz++)
            for (int cx = 0; cx < nxs - 1; cx++) {
                Node nd;
                nd.brick = i;
                nd.lo[0] = xs[cx]; nd.lo[1] = t.mn[1]; nd.lo[2] = zs[cz];
                nd.hi[0] = xs[cx + 1]; nd.hi[1] = t.mx[1]; nd.hi[2] = zs[cz + 1];
                float share = (float)((nd.hi[0] - nd.lo[0]) * (nd.hi[2] - nd.lo[2])) / vol;
                for (int d = 0; d < 3; d++) nd.c[d] = 0.5 * (double)(nd.lo[d] + nd.hi[d]) * TK;
                nd.load = load[k] * share;
                nd.sink = sink[k] != 0;
                GROW(nodes, nn, ncap);
                nodes[nn++] = nd;
            }
        int nx = gx[k], nz = gz[k];
        uint32_t base = first[i];
        for (int cz = 0; cz < nz; cz++)
            for (int cx = 0; cx < nx; cx++) {
                uint32_t id = base + (uint32_t)(cz * nx + cx);
                Node nd = nodes[id];
                if (cx + 1 < nx) {
                    Link L = {id, id + 1, NONE, 0, {nd.hi[0], nd.lo[1], nd.lo[2]}, {nd.hi[0], nd.hi[1], nd.hi[2]}};
                    GROW(links, nl, lcap); links[nl++] = L;
                }
                if (cz + 1 < nz) {
                    Link L = {id, id + (uint32_t)nx, NONE, 2, {nd.lo[0], nd.lo[1], nd.hi[2]}, {nd.hi[0], nd.hi[1], nd.hi[2]}};
                    GROW(links, nl, lcap); links[nl++] = L;
                }
            }
    }
    /* links across joints */
    for (int k = 0; k < nset; k++) {
        uint32_t i = set[k];
        for (uint32_t j = b->jhead[i]; j != NONE; j = joint_next(b, j, i)) {
            uint32_t o = joint_other(b, j, i);
            if (first[o] == NONE || b->ja[j] != i) continue;
            int ax = b->jaxis[j];
            int ko = (int)slot[o];
            int na0 = gx[k], na1 = gz[k], nb0 = gx[ko], nb1 = gz[ko];
            int32_t plane = tbox(b, i).mx[ax];
            for (int ca = 0; ca < na0 * na1; ca++) {
                const Node* p = &nodes[first[i] + (uint32_t)ca];
                if (ax != 1 && p->hi[ax] != plane) continue;
                int x0, x1, z0, z1;
                if (ax == 0) { x0 = 0; x1 = nb0; } else range_of(cx_[ko], ncx[ko], p->lo[0], p->hi[0], &x0, &x1);
                if (ax == 2) { z0 = 0; z1 = nb1; } else range_of(cz_[ko], ncz[ko], p->lo[2], p->hi[2], &z0, &z1);
                for (int cz = z0; cz < z1; cz++)
                    for (int cx = x0; cx < x1; cx++) {
                        uint32_t ia = first[i] + (uint32_t)ca, ib = first[o] + (uint32_t)(cz * nb0 + cx);
                        const Node *P = &nodes[ia], *Q = &nodes[ib];
                        Link L;
                        L.a = ia; L.b = ib; L.joint = j; L.axis = ax;
                        int ok = 1;
                        for (int d = 0; d < 3; d++) {
                            if (d == ax) { L.flo[d] = plane; L.fhi[d] = plane; continue; }
                            L.flo[d] = P->lo[d] > Q->lo[d] ? P->lo[d] : Q->lo[d];
                            L.fhi[d] = P->hi[d] < Q->hi[d] ? P->hi[d] : Q->hi[d];
                            if (L.fhi[d] <= L.flo[d]) ok = 0;
                        }
                        if (!ok) continue;
                        GROW(links, nl, lcap); links[nl++] = L;
                    }
            }
        }
    }
    /* point loads: (index into the set, x, z in ticks, N) */
    for (int q = 0; q < npoint; q++) {
        int k = (int)point4[4 * q];
        if (k < 0 || k >= nset) continue;
        uint32_t i = set[k];
        if (first[i] == NONE) continue;
        int nx = gx[k], nz = gz[k];
        float x = point4[4 * q + 1], z = point4[4 * q + 2], f = point4[4 * q + 3];
        int cx = 0, cz = 0;
        { int c = 0; while (c < ncx[k] && (float)cx_[k][c] <= x) c++; cx = c > 0 ? c - 1 : 0; if (cx > nx - 1) cx = nx - 1; }
        { int c = 0; while (c < ncz[k] && (float)cz_[k][c] <= z) c++; cz = c > 0 ? c - 1 : 0; if (cz > nz - 1) cz = nz - 1; }
        nodes[first[i] + (uint32_t)(cz * nx + cx)].load += f;
    }
    int n = nn;
    /* cells connected to a support */
    uint8_t* live = zalloc((size_t)(n ? n : 1));
    {
        uint32_t* head = zalloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
        uint32_t* nxt = zalloc(sizeof(uint32_t) * (size_t)(2 * nl + 1));
        for (int k = 0; k < n; k++) head[k] = NONE;
        for (int l = 0; l < nl; l++) {
            nxt[2 * l] = head[links[l].a]; head[links[l].a] = (uint32_t)(2 * l);
            nxt[2 * l + 1] = head[links[l].b]; head[links[l].b] = (uint32_t)(2 * l + 1);
        }
        uint32_t* st = zalloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
        int sn = 0;
        for (int k = 0; k < n; k++) if (nodes[k].sink) { live[k] = 1; st[sn++] = (uint32_t)k; }
        while (sn > 0) {
            uint32_t k = st[--sn];
            for (uint32_t e = head[k]; e != NONE; e = nxt[e]) {
                uint32_t o = (e & 1) ? links[e / 2].a : links[e / 2].b;
                if (!live[o]) { live[o] = 1; st[sn++] = o; }
            }
        }
        free(head); free(nxt); free(st);
    }
    Spring* springs = zalloc(sizeof(Spring) * (size_t)(nl ? nl : 1));
    uint32_t* ends = zalloc(sizeof(uint32_t) * 2 * (size_t)(nl ? nl : 1));
    for (int l = 0; l < nl; l++) { springs[l] = make_spring(b, nodes, &links[l]); ends[2 * l] = links[l].a; ends[2 * l + 1] = links[l].b; }
    uint8_t* fixed = zalloc((size_t)(n ? n : 1));
    for (int k = 0; k < n; k++) fixed[k] = nodes[k].sink || !live[k];
    double* f = zalloc(sizeof(double) * 6 * (size_t)(n ? n : 1));
    for (int k = 0; k < n; k++) {
        double w = (double)nodes[k].load;
        f[6 * k + 0] = w * (double)dir[0]; f[6 * k + 1] = w * (double)dir[1]; f[6 * k + 2] = w * (double)dir[2];
    }
    double* pos = zalloc(sizeof(double) * 3 * (size_t)(n ? n : 1));
    for (int k = 0; k < n; k++) for (int d = 0; d < 3; d++) pos[3 * k + d] = nodes[k].c[d];
    double* x0 = NULL;
    if (warm) {
        x0 = zalloc(sizeof(double) * 6 * (size_t)(n ? n : 1));
        for (int k = 0; k < n; k++) {
            if (fixed[k]) continue;
            uint32_t* s = map_slot(&warm->map, warm_key(nodes[k].brick, nodes[k].lo[0], nodes[k].lo[2]), 0);
            if (s) memcpy(x0 + 6 * (size_t)k, warm->x + 6 * (size_t)*s, sizeof(double) * 6);
        }
    }
    double* x = zalloc(sizeof(double) * 6 * (size_t)(n ? n : 1));
    Net* N = zalloc(sizeof(Net));
    N->order = order;
    int it = n > 0 ? solve_network(N, springs, ends, nl, fixed, pos, f, n, x, x0) : 0;
    R->nlevels = N->nl;
    for (int q = 0; q < N->nl && q < 8; q++) R->level_n[q] = N->lv[q].n;
    for (int q = 0; q < N->nl; q++) level_free(&N->lv[q]);
    free(N);
    if (warm) {
        map_free(&warm->map); map_init(&warm->map, 64);
        warm->n = 0;
        for (int k = 0; k < n; k++) {
            if (fixed[k]) continue;
            uint32_t* s = map_slot(&warm->map, warm_key(nodes[k].brick, nodes[k].lo[0], nodes[k].lo[2]), 1);
            if (*s == NONE) {
                if (warm->n >= warm->cap) { warm->cap = warm->cap ? 2 * warm->cap : 256; warm->x = realloc(warm->x, sizeof(double) * 6 * (size_t)warm->cap); if (!warm->x) abort(); }
                *s = (uint32_t)warm->n++;
            }
            memcpy(warm->x + 6 * (size_t)*s, x + 6 * (size_t)k, sizeof(double) * 6);
        }
    }
    R->iterations = it; R->nodes = n; R->nlinks = nl;
    /* checks: joints (their sub-links summed about the joint's face centre), then faces inside bricks */
    {
        uint32_t* jslot = zalloc(sizeof(uint32_t) * (size_t)(b->nj ? b->nj : 1));
        for (int j = 0; j < b->nj; j++) jslot[j] = NONE;
        uint32_t* order_j = zalloc(sizeof(uint32_t) * (size_t)(nl ? nl : 1));
        int ng = 0;
        /* parts: links of each joint in link order, as a linked list through nextl */
        uint32_t* headl = zalloc(sizeof(uint32_t) * (size_t)(nl ? nl : 1));
        uint32_t* taill = zalloc(sizeof(uint32_t) * (size_t)(nl ? nl : 1));
        uint32_t* nextl = zalloc(sizeof(uint32_t) * (size_t)(nl ? nl : 1));
        for (int l = 0; l < nl; l++) {
            nextl[l] = NONE;
            const Link* L = &links[l];
            if (L->joint == NONE || !live[L->a] || !live[L->b]) continue;
            uint32_t s = jslot[L->joint];
            if (s == NONE) { s = (uint32_t)ng++; jslot[L->joint] = s; order_j[s] = L->joint; headl[s] = (uint32_t)l; taill[s] = (uint32_t)l; }
            else { nextl[taill[s]] = (uint32_t)l; taill[s] = (uint32_t)l; }
        }
        R->nj = ng;
        R->jid = zalloc(sizeof(int32_t) * (size_t)(ng ? ng : 1));
        R->jratio = zalloc(sizeof(float) * (size_t)(ng ? ng : 1));
        float maxr = 0.0f;
        for (int s = 0; s < ng; s++) {
            uint32_t j = order_j[s];
            int ax = b->jaxis[j], u = (ax + 1) % 3, w = (ax + 2) % 3;
            int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
            for (uint32_t l = headl[s]; l != NONE; l = nextl[l])
                for (int d = 0; d < 3; d++) {
                    if (links[l].flo[d] < lo[d]) lo[d] = links[l].flo[d];
                    if (links[l].fhi[d] > hi[d]) hi[d] = links[l].fhi[d];
                }
            double c[3];
            for (int d = 0; d < 3; d++) c[d] = 0.5 * (double)(lo[d] + hi[d]) * TK;
            double ft[3] = {0, 0, 0}, mt[3] = {0, 0, 0};
            for (uint32_t l = headl[s]; l != NONE; l = nextl[l]) {
                double d[3], r[3], fo[3], mo[3];
                strain(&springs[l], x + 6 * (size_t)links[l].a, x + 6 * (size_t)links[l].b, d, r);
                force(&springs[l], d, r, fo, mo);
                const double* p = springs[l].c;
                double rr[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
                for (int q = 0; q < 3; q++) { ft[q] += fo[q]; mt[q] += mo[q]; }
                double rx[3] = {rr[1] * fo[2] - rr[2] * fo[1], rr[2] * fo[0] - rr[0] * fo[2], rr[0] * fo[1] - rr[1] * fo[0]};
                for (int q = 0; q < 3; q++) mt[q] += rx[q];
            }
            uint32_t ia = b->ja[j], ib = b->jb[j];
            int ma = b->mat[ia], mb = b->mat[ib];
            Sect sect = sect_rect((double)(hi[u] - lo[u]) * TK, (double)(hi[w] - lo[w]) * TK, (double)b->jarea[j] * TK * TK);
            float conn[3]; connection(b, j, conn);
            int cracked_masonry = ma == M_MASONRY || mb == M_MASONRY;
            float residual[3] = {1.0f, 1.0f, 1.0f};
            if (b->jdmg[j] & 1) {
                if (cracked_masonry) { residual[0] = 1.0f; residual[1] = 0.05f; residual[2] = 0.4f; }
                else { residual[0] = 0.5f; residual[1] = 0.5f; residual[2] = 0.5f; }
            }
            float sa[3], sb[3]; strength(ma, sa); strength(mb, sb);
            float cap[3];
            for (int q = 0; q < 3; q++) cap[q] = fminf(sa[q], sb[q]) * conn[q] * residual[q];
            float flex = fminf(flexural_of(b, ia), flexural_of(b, ib)) * conn[1] * residual[1];
            double fl[3] = {ft[ax], ft[u], ft[w]}, ml[3] = {mt[ax], mt[u], mt[w]};
            if (hinge(b, ia, ib)) { ml[0] = 0.0; ml[1] = 0.0; ml[2] = 0.0; }
            float ratio = section_ratio(&sect, fl, ml, cap, flex);
            R->jid[s] = (int32_t)j; R->jratio[s] = ratio;
            if (ratio > maxr) maxr = ratio;
        }
        free(jslot); free(order_j); free(headl); free(taill); free(nextl);
        int ni = 0;
        R->ibrick = zalloc(sizeof(int32_t) * (size_t)(nl ? nl : 1));
        R->iaxis = zalloc(sizeof(int32_t) * (size_t)(nl ? nl : 1));
        R->iat = zalloc(sizeof(int32_t) * (size_t)(nl ? nl : 1));
        R->iratio = zalloc(sizeof(float) * (size_t)(nl ? nl : 1));
        R->worst_ratio = 0.0f; R->worst_brick = 0; R->worst_axis = 0; R->worst_flow = 0.0f; R->worst_moment = 0.0f;
        for (int l = 0; l < nl; l++) {
            const Link* L = &links[l];
            if (L->joint != NONE || !live[L->a]) continue;
            uint32_t i = nodes[L->a].brick;
            int u = (L->axis + 1) % 3, w = (L->axis + 2) % 3;
            double eu = (double)(L->fhi[u] - L->flo[u]) * TK, ew = (double)(L->fhi[w] - L->flo[w]) * TK;
            double d[3], r[3], fo[3], mo[3];
            strain(&springs[l], x + 6 * (size_t)L->a, x + 6 * (size_t)L->b, d, r);
            force(&springs[l], d, r, fo, mo);
            Sect sect = sect_rect(eu, ew, eu * ew);
            double fl[3] = {fo[L->axis], fo[u], fo[w]}, ml[3] = {mo[L->axis], mo[u], mo[w]};
            float st[3]; strength(b->mat[i], st);
            float ratio = section_ratio(&sect, fl, ml, st, flexural_of(b, i));
            if (ratio > R->worst_ratio) {
                R->worst_ratio = ratio; R->worst_brick = (int)i; R->worst_axis = L->axis;
                R->worst_flow = (float)fo[L->axis]; R->worst_moment = (float)(mo[u] * mo[u] + mo[w] * mo[w]);
            }
            if (ratio > maxr) maxr = ratio;
            R->ibrick[ni] = (int32_t)i; R->iaxis[ni] = L->axis; R->iat[ni] = L->flo[L->axis]; R->iratio[ni] = ratio;
            ni++;
        }
        R->ni = ni;
        R->max_ratio = maxr;
    }
    R->x = x;
    R->node = zalloc(sizeof(int32_t) * 7 * (size_t)(n ? n : 1));
    R->nload = zalloc(sizeof(float) * (size_t)(n ? n : 1));
    R->nfixed = zalloc((size_t)(n ? n : 1));
    for (int k = 0; k < n; k++) {
        R->node[7 * k] = (int32_t)nodes[k].brick;
        for (int d = 0; d < 3; d++) { R->node[7 * k + 1 + d] = nodes[k].lo[d]; R->node[7 * k + 4 + d] = nodes[k].hi[d]; }
        R->nload[k] = nodes[k].load;
        R->nfixed[k] = fixed[k];
    }
    R->ends = ends;
    R->sp = zalloc(sizeof(double) * 12 * (size_t)(nl ? nl : 1));
    for (int l = 0; l < nl; l++)
        for (int d = 0; d < 3; d++) {
            R->sp[12 * l + d] = springs[l].k[d]; R->sp[12 * l + 3 + d] = springs[l].kr[d];
            R->sp[12 * l + 6 + d] = springs[l].ra[d]; R->sp[12 * l + 9 + d] = springs[l].rb[d];
        }
    for (int k = 0; k < nset; k++) { free(cx_[k]); free(cz_[k]); }
    free(cx_); free(cz_); free(ncx); free(ncz); free(gx); free(gz); free(first); free(slot);
    free(nodes); free(links); free(live); free(springs); free(fixed); free(f); free(pos); free(x0);
    return R;
}

/* ---------------------------------------------------------------- result readers */
/* [max_ratio, iterations, nodes, links, levels, worst-in ratio, brick, axis, flow, moment^2, joints checked, inner
 *  faces checked] */
void os_result_scalars(const os_result* R, double* out12) {
    out12[0] = R->max_ratio; out12[1] = R->iterations; out12[2] = R->nodes; out12[3] = R->nlinks; out12[4] = R->nlevels;
    out12[5] = R->worst_ratio; out12[6] = R->worst_brick; out12[7] = R->worst_axis; out12[8] = R->worst_flow;
    out12[9] = R->worst_moment; out12[10] = R->nj; out12[11] = R->ni;
}
void os_result_levels(const os_result* R, int32_t* out8) { for (int q = 0; q < 8; q++) out8[q] = q < R->nlevels ? R->level_n[q] : 0; }
int os_result_joints(const os_result* R, int32_t* jid, float* ratio, int cap) {
    for (int s = 0; s < R->nj && s < cap; s++) { jid[s] = R->jid[s]; ratio[s] = R->jratio[s]; }
    return R->nj;
}
int os_result_inner(const os_result* R, int32_t* bat3, float* ratio, int cap) {
    for (int s = 0; s < R->ni && s < cap; s++) { bat3[3 * s] = R->ibrick[s]; bat3[3 * s + 1] = R->iaxis[s]; bat3[3 * s + 2] = R->iat[s]; ratio[s] = R->iratio[s]; }
    return R->ni;
}
int os_result_nodes(const os_result* R, int32_t* node7, float* load, uint8_t* fixed, double* x6, int cap) {
    for (int k = 0; k < R->nodes && k < cap; k++) {
        if (node7) memcpy(node7 + 7 * (size_t)k, R->node + 7 * (size_t)k, sizeof(int32_t) * 7);
        if (load) load[k] = R->nload[k];
        if (fixed) fixed[k] = R->nfixed[k];
        if (x6) memcpy(x6 + 6 * (size_t)k, R->x + 6 * (size_t)k, sizeof(double) * 6);
    }
    return R->nodes;
}
/* the springs: ends (2 per link) and k[3] kr[3] ra[3] rb[3] per link */
int os_result_springs(const os_result* R, uint32_t* ends2, double* sp12, int cap) {
    for (int l = 0; l < R->nlinks && l < cap; l++) {
        ends2[2 * l] = R->ends[2 * l]; ends2[2 * l + 1] = R->ends[2 * l + 1];
        memcpy(sp12 + 12 * (size_t)l, R->sp + 12 * (size_t)l, sizeof(double) * 12);
    }
    return R->nlinks;
}
