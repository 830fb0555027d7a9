/* Lattice world: the legacy brick engine's physics frame, restated from src/lib.rs.
 *
 * TEST INFRASTRUCTURE ONLY. This is the CPU oracle for the lattice path; only tests, the smoke check and the
 * benchmark's CPU baseline use it. The library never calls it.
 *
 * What is restated (src/lib.rs line ranges):
 *   bricks, bonds, finalize_brick ........ :620-696    setup_body / rebuild_body / kill_body .. :844-1002
 *   sor_brick / sor_struct / sor_body .... :1004-1051  warm_struct / warm_step ............... :1061-1079
 *   conn_struct / connectivity ........... :1085-1141  cond_w / axis_up / stress_brick ........ :1142-1227
 *   structural (loads, g - a, sweeps) .... :1228-1365  update_cache / gen_contacts ........... :1368-1471
 *   plough / prep / solve ................ :1493-1601  impact_damage / crush_at ............... :1610-1702
 *   integrate / pinch_check / substep .... :1703-1783  explode_at / blast_brick .............. :2679-2716
 *   step (frame orchestration) ........... :3639-3721
 * Out of scope (render and game side): particles, shards, dust, events, mesh sets, zones, agents, vehicles,
 * crumb clearing; the random stream is drawn only by physical decisions (plough velocities, the pulverised share of
 * a crush, dusting of single-brick pieces).
 *
 * Two orders:
 *   LL_ORDER_SEQ  the reference's orders: contacts solved one after another in generation order (10 Gauss-Seidel
 *                 passes), SOR sweeps in place in brick order, stress judged brick after brick.
 *   LL_ORDER_DEV  the GPU's orders: in each of the 10 passes, first every body's static chain (its contacts against
 *                 static bricks, the ground and sleeping bodies, in generation order, as the reference solves them),
 *                 then the body-body contacts coloured by claim rounds and solved colour after colour, the rest by
 *                 a Jacobi pass with mass splitting (per-body sums as a 64-lane halving tree in contact order); SOR
 *                 sweeps red-black by lattice parity; stress judged on the field at the start of the pass, crushes
 *                 applied afterwards in brick order.
 * Common to both (deviations from the reference, all explicit): no contact cap (MAXC); the load-field sweeps of
 * loose pieces have no per-frame budget (standing structures keep the reference's 160k and 120k brick-update budgets);
 * sphere neighbours are found in the 27 cells around a brick by exact cell (no hash-bucket collisions or duplicates).
 */
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include "legacy.h"
#include "sf_oracle.h"

#define NONE 0xFFFFFFFFu
#define LL_CHAIN (-2)      /* colour of a contact solved in its body's static chain */

typedef struct {
    int a, b, bo;          /* body a; dynamic partner body or -1; body of the other brick (-1: static lattice or ground) */
    uint32_t ba, bb;       /* brick of a; other brick or NONE */
    int on, col;
    LV n, p, ra, rb, t1, t2;
    float depth, mn, mt1, mt2, jn, jt1, jt2, bias, vn0;
    float smn, smt1, smt2; /* overflow: effective masses with split inverse masses */
} LC;

typedef struct LW {
    int order, K, R;
    uint32_t rng, frame;
    float time, blast_t, blast_r, blast_p, hsub;
    LV blast_c;
    int crush_n;
    uint32_t fail[4];      /* joints failed under load: compression (crushed), tension, shear, footing */
    /* bricks */
    int nb, bcap;
    uint8_t *mat, *sid, *infill;
    int32_t *bx, *by, *bz, *body;
    uint32_t *next, *nb6, *sup, *stamp, *q;
    LV *local, *wpos;
    float *bond, *bmax, *phi, *ext, *extn, *stress;
    uint32_t stamp_id;
    int gx0, gy0, gz0, gnx, gny, gnz;
    int32_t *grid;
    /* structures */
    int st_n, st_cap;
    int *st_lo, *st_hi, *st_min, *st_max, *st_top0, *st_top, *st_warm;
    uint32_t *st_init, *st_now;
    float *st_active, *st_settle, *st_maxr, *st_resid;
    uint8_t *st_solved, *st_conn, *st_broke;
    int conn_any;
    /* bodies */
    int hw, cursor, dcap;
    uint8_t *alive, *asleep, *dirty, *wake;
    uint32_t *first, *count, *nc;
    LV *pos, *v, *w, *vpre, *wpre, *vlast, *wlast;
    LQ *qq;
    LM *rot, *iil, *iiw;
    float *m, *im, *sleep, *rad, *cool, *lt, *resid, *maxr, *nmax;
    /* contacts */
    int ncs, ccap;
    LC* cs;
    /* scratch */
    uint32_t *lst, *lst2, *comp, *cnt;
    uint64_t* hkeys; int nh;
    float stats[16];
} LW;

static void* xcalloc(size_t n, size_t s) { void* p = calloc(n ? n : 1, s); if (!p) abort(); return p; }
/* whether one part of the frame runs in the device's order: all of them for LL_ORDER_DEV, none for LL_ORDER_SEQ,
 * the flagged ones for LL_ORDER_MIX | flags (mixed orders attribute a difference between the two to its part) */
static inline int dev_part(const LW* w, int part) {
    return w->order == LL_ORDER_DEV || ((w->order & LL_ORDER_MIX) && (w->order & part));
}

/* ------------------------------------------------------------------ creation */
LW* ll_create(int order, int brick_cap, int body_cap, int ncolour, int rounds, uint32_t seed) {
    LW* w = (LW*)xcalloc(1, sizeof(LW));
    w->order = order; w->K = ncolour < 1 ? 1 : (ncolour > 32 ? 32 : ncolour); w->R = rounds;
    w->rng = seed ? seed : 0x9E3779B9u;
    int B = brick_cap, D = body_cap;
    w->bcap = B; w->dcap = D;
    w->mat = xcalloc(B, 1); w->sid = xcalloc(B, 1); w->infill = xcalloc(B, 1);
    w->bx = xcalloc(B, 4); w->by = xcalloc(B, 4); w->bz = xcalloc(B, 4); w->body = xcalloc(B, 4);
    w->next = xcalloc(B, 4); w->nb6 = xcalloc((size_t)B * 6, 4); w->sup = xcalloc(B, 4); w->stamp = xcalloc(B, 4); w->q = xcalloc(B, 4);
    w->local = xcalloc(B, sizeof(LV)); w->wpos = xcalloc(B, sizeof(LV));
    w->bond = xcalloc((size_t)B * 4, 4); w->bmax = xcalloc((size_t)B * 4, 4);
    w->phi = xcalloc(B, 4); w->ext = xcalloc(B, 4); w->extn = xcalloc(B, 4); w->stress = xcalloc(B, 4);
    w->lst = xcalloc(B, 4); w->lst2 = xcalloc(B, 4); w->comp = xcalloc(B, 4); w->cnt = xcalloc((size_t)B + 2, 4);
    w->hkeys = xcalloc(B, 8);
    w->st_cap = 256;
    int S = w->st_cap;
    w->st_lo = xcalloc(S, 4); w->st_hi = xcalloc(S, 4); w->st_min = xcalloc(S * 3, 4); w->st_max = xcalloc(S * 3, 4);
    w->st_top0 = xcalloc(S, 4); w->st_top = xcalloc(S, 4); w->st_warm = xcalloc(S, 4);
    w->st_init = xcalloc(S, 4); w->st_now = xcalloc(S, 4);
    w->st_active = xcalloc(S, 4); w->st_settle = xcalloc(S, 4); w->st_maxr = xcalloc(S, 4); w->st_resid = xcalloc(S, 4);
    w->st_solved = xcalloc(S, 1); w->st_conn = xcalloc(S, 1); w->st_broke = xcalloc(S, 1);
    w->alive = xcalloc(D, 1); w->asleep = xcalloc(D, 1); w->dirty = xcalloc(D, 1); w->wake = xcalloc(D, 1);
    w->first = xcalloc(D, 4); w->count = xcalloc(D, 4); w->nc = xcalloc(D, 4);
    w->pos = xcalloc(D, sizeof(LV)); w->v = xcalloc(D, sizeof(LV)); w->w = xcalloc(D, sizeof(LV));
    w->vpre = xcalloc(D, sizeof(LV)); w->wpre = xcalloc(D, sizeof(LV)); w->vlast = xcalloc(D, sizeof(LV)); w->wlast = xcalloc(D, sizeof(LV));
    w->qq = xcalloc(D, sizeof(LQ)); w->rot = xcalloc(D, sizeof(LM)); w->iil = xcalloc(D, sizeof(LM)); w->iiw = xcalloc(D, sizeof(LM));
    w->m = xcalloc(D, 4); w->im = xcalloc(D, 4); w->sleep = xcalloc(D, 4); w->rad = xcalloc(D, 4); w->cool = xcalloc(D, 4);
    w->lt = xcalloc(D, 4); w->resid = xcalloc(D, 4); w->maxr = xcalloc(D, 4); w->nmax = xcalloc(D, 4);
    for (int b = 0; b < D; b++) w->first[b] = NONE;
    w->ccap = 1024;
    w->cs = xcalloc(w->ccap, sizeof(LC));
    return w;
}

void ll_destroy(LW* w) {
    if (!w) return;
    void* ps[] = {w->mat, w->sid, w->infill, w->bx, w->by, w->bz, w->body, w->next, w->nb6, w->sup, w->stamp, w->q, w->local, w->wpos,
                  w->bond, w->bmax, w->phi, w->ext, w->extn, w->stress, w->lst, w->lst2, w->comp, w->cnt, w->hkeys, w->grid,
                  w->st_lo, w->st_hi, w->st_min, w->st_max, w->st_top0, w->st_top, w->st_warm, w->st_init, w->st_now,
                  w->st_active, w->st_settle, w->st_maxr, w->st_resid, w->st_solved, w->st_conn, w->st_broke,
                  w->alive, w->asleep, w->dirty, w->wake, w->first, w->count, w->nc, w->pos, w->v, w->w, w->vpre, w->wpre,
                  w->vlast, w->wlast, w->qq, w->rot, w->iil, w->iiw, w->m, w->im, w->sleep, w->rad, w->cool, w->lt, w->resid,
                  w->maxr, w->nmax, w->cs};
    for (size_t k = 0; k < sizeof(ps) / sizeof(ps[0]); k++) free(ps[k]);
    free(w);
}

/* ------------------------------------------------------------------ lattice grid (cell -> brick + 1) */
static int32_t gget(const LW* w, int x, int y, int z) {
    x -= w->gx0; y -= w->gy0; z -= w->gz0;
    if (x < 0 || y < 0 || z < 0 || x >= w->gnx || y >= w->gny || z >= w->gnz) return 0;
    return w->grid[((size_t)y * w->gnz + z) * w->gnx + x];
}
static void gset(LW* w, int x, int y, int z, int32_t v) {
    x -= w->gx0; y -= w->gy0; z -= w->gz0;
    if (x < 0 || y < 0 || z < 0 || x >= w->gnx || y >= w->gny || z >= w->gnz) return;
    w->grid[((size_t)y * w->gnz + z) * w->gnx + x] = v;
}
static int static_brick(const LW* w, int x, int y, int z) {
    int32_t g = gget(w, x, y, z);
    if (!g) return -1;
    int j = g - 1;
    return (w->mat[j] != 0 && w->body[j] < 0) ? j : -1;
}
static inline int nbr(const LW* w, int i, int d) { uint32_t j = w->nb6[6 * i + d]; return j == NONE ? -1 : (int)j; }
/* (src/lib.rs:380-385) */
static inline int linked(const LW* w, int i, int d, float* hp_out) {
    int j = nbr(w, i, d);
    if (j < 0 || w->mat[j] == 0 || w->body[j] != w->body[i]) return -1;
    float hp = (d & 1) == 0 ? w->bond[4 * i + (d >> 1)] : w->bond[4 * j + (d >> 1)];
    if (hp > 0.0f) { if (hp_out) *hp_out = hp; return j; }
    return -1;
}
static inline LV bpos(const LW* w, int i) { return w->body[i] < 0 ? lg_lat_pos(w->bx[i], w->by[i], w->bz[i]) : w->wpos[i]; }
static inline int in_struct(const LW* w, int i, int s) { return w->sid[i] == s && w->mat[i] != 0 && w->mat[i] < LG_SLUG && w->body[i] < 0; }

static int add_brick(LW* w, int x, int y, int z, int m, int sid, int infill) {
    if (w->nb >= w->bcap) return -1;
    int i = w->nb++;
    w->mat[i] = (uint8_t)m; w->bx[i] = x; w->by[i] = y; w->bz[i] = z;
    w->sid[i] = (uint8_t)sid; w->infill[i] = (uint8_t)((m == LG_GLASS || infill) ? 1 : 0);
    w->body[i] = -1; w->next[i] = NONE;
    w->local[i] = lg_lat_pos(x, y, z); w->wpos[i] = w->local[i];
    return i;
}

int ll_add_structure(LW* w, int n, const int32_t* xyz, const uint8_t* mat, const uint8_t* infill) {
    if (w->st_n >= w->st_cap || w->st_n >= 255) return -1;
    int s = w->st_n++;
    w->st_lo[s] = w->nb;
    for (int k = 0; k < n; k++) add_brick(w, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], mat[k], s, infill ? infill[k] : 0);
    w->st_hi[s] = w->nb;
    return s;
}

static int alloc_body(LW* w) {
    for (int k = 0; k < w->dcap; k++) {
        int b = (w->cursor + k) % w->dcap;
        if (!w->alive[b]) {
            w->cursor = (b + 1) % w->dcap;
            if (b + 1 > w->hw) w->hw = b + 1;
            return b;
        }
    }
    return -1;
}

/* (src/lib.rs:844-891): bricks lst2[s..e) become body b */
static void setup_body(LW* w, int b, int s, int e, LV ppos, LM prot, LQ pq, LV pv, LV pw) {
    w->lt[b] = 0.0f; w->resid[b] = 1.0f; w->maxr[b] = 0.0f;
    float mass = 0.0f;
    LV cl = lv(0, 0, 0);
    for (int t = s; t < e; t++) { int i = w->lst2[t]; float m = LG_MASS[w->mat[i]]; mass += m; cl = lg_madd(cl, w->local[i], m); }
    cl = lg_mul(cl, 1.0f / mass);
    float it[9] = {0};
    float rad = 0.0f;
    const float cube = LG_S * LG_S / 6.0f;
    for (int t = s; t < e; t++) {
        int i = w->lst2[t];
        float m = LG_MASS[w->mat[i]];
        LV r = lg_sub(w->local[i], cl);
        w->local[i] = r;
        float rr = lg_dot(r, r);
        rad = lg_maxf(rad, rr);
        it[0] += m * (rr - r.x * r.x + cube); it[4] += m * (rr - r.y * r.y + cube); it[8] += m * (rr - r.z * r.z + cube);
        it[1] -= m * r.x * r.y; it[2] -= m * r.x * r.z; it[5] -= m * r.y * r.z;
        w->body[i] = b;
        w->next[i] = t + 1 < e ? w->lst2[t + 1] : NONE;
    }
    it[3] = it[1]; it[6] = it[2]; it[7] = it[5];
    LM itm; memcpy(itm.m, it, sizeof it);
    LM iil = lg_minv(&itm);
    LV off = lg_mv(&prot, cl);
    w->alive[b] = 1;
    w->first[b] = w->lst2[s];
    w->count[b] = (uint32_t)(e - s);
    w->pos[b] = lg_add(ppos, off);
    w->qq[b] = pq; w->rot[b] = prot;
    w->v[b] = lg_add(pv, lg_cross(pw, off));
    w->w[b] = pw;
    w->vlast[b] = w->v[b]; w->wlast[b] = pw; w->nmax[b] = 1.0f;
    w->m[b] = mass; w->im[b] = 1.0f / mass;
    w->iil[b] = iil;
    LM t1 = lg_mmul(&prot, &iil), pt = lg_mt(&prot);
    w->iiw[b] = lg_mmul(&t1, &pt);
    w->sleep[b] = 0.0f; w->asleep[b] = 0; w->dirty[b] = 0; w->wake[b] = 0; w->cool[b] = 0.15f;
    w->rad[b] = sqrtf(rad) + LG_H * 1.8f;
    for (int t = s; t < e; t++) { int i = w->lst2[t]; w->wpos[i] = lg_add(w->pos[b], lg_mv(&prot, w->local[i])); }
}

static const LM IDM = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
static const LQ IDQ = {0, 0, 0, 1};

int ll_add_body(LW* w, int n, const int32_t* xyz, const uint8_t* mat, const float* v3, const float* w3) {
    if (n <= 0) return -1;
    int b = alloc_body(w);
    if (b < 0) return -1;
    for (int k = 0; k < n; k++) {
        int i = add_brick(w, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], mat[k], 255, 0);
        if (i < 0) return -1;
        w->lst2[k] = (uint32_t)i;
    }
    setup_body(w, b, 0, n, lv(0, 0, 0), IDM, IDQ, lv(v3[0], v3[1], v3[2]), lv(w3[0], w3[1], w3[2]));
    return b;
}

/* (src/lib.rs:668-696): neighbour table and bond capacities; the ground bond of a footing */
static void finalize_brick(LW* w, int i) {
    if (w->mat[i] == 0) return;
    for (int d = 0; d < 6; d++) {
        int32_t g = gget(w, w->bx[i] + LG_DIRS[d][0], w->by[i] + LG_DIRS[d][1], w->bz[i] + LG_DIRS[d][2]);
        w->nb6[6 * i + d] = g ? (uint32_t)(g - 1) : NONE;
    }
    for (int a = 0; a < 3; a++) {
        int j = nbr(w, i, a * 2);
        int gap = j >= 0 && w->infill[i] != w->infill[j] && (a != 1 || w->infill[i] == 1);
        if (j >= 0 && !gap && w->mat[j] != 0 && w->body[j] == w->body[i] && w->sid[j] == w->sid[i]) {
            float c = lg_bond_cap(w->mat[i], w->mat[j], a);
            w->bond[4 * i + a] = c; w->bmax[4 * i + a] = c;
        } else { w->bond[4 * i + a] = 0.0f; w->bmax[4 * i + a] = 0.0f; }
    }
    if (w->by[i] == 0 && w->body[i] < 0) { float c = LG_CAPV[w->mat[i]] * 2.0f; w->bond[4 * i + 3] = c; w->bmax[4 * i + 3] = c; }
    else { w->bond[4 * i + 3] = 0.0f; w->bmax[4 * i + 3] = 0.0f; }
}

void ll_finalize(LW* w) {
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int i = 0; i < w->nb; i++) {
        int c[3] = {w->bx[i], w->by[i], w->bz[i]};
        for (int k = 0; k < 3; k++) { if (c[k] < lo[k]) lo[k] = c[k]; if (c[k] > hi[k]) hi[k] = c[k]; }
    }
    if (w->nb == 0) { lo[0] = lo[1] = lo[2] = 0; hi[0] = hi[1] = hi[2] = 0; }
    w->gx0 = lo[0] - 1; w->gy0 = lo[1] - 1; w->gz0 = lo[2] - 1;
    w->gnx = hi[0] - lo[0] + 3; w->gny = hi[1] - lo[1] + 3; w->gnz = hi[2] - lo[2] + 3;
    free(w->grid);
    w->grid = xcalloc((size_t)w->gnx * w->gny * w->gnz, 4);
    for (int i = 0; i < w->nb; i++) if (w->mat[i] && !gget(w, w->bx[i], w->by[i], w->bz[i])) gset(w, w->bx[i], w->by[i], w->bz[i], i + 1);
    for (int i = 0; i < w->nb; i++) finalize_brick(w, i);
    for (int s = 0; s < w->st_n; s++) {
        uint32_t cnt = 0; int top = 0;
        int mn[3] = {1 << 30, 1 << 30, 1 << 30}, mx[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
        for (int i = w->st_lo[s]; i < w->st_hi[s]; i++) {
            if (!in_struct(w, i, s)) continue;
            cnt++; if (w->by[i] + 1 > top) top = w->by[i] + 1;
            int c[3] = {w->bx[i], w->by[i], w->bz[i]};
            for (int k = 0; k < 3; k++) { if (c[k] < mn[k]) mn[k] = c[k]; if (c[k] > mx[k]) mx[k] = c[k]; }
        }
        w->st_init[s] = w->st_now[s] = cnt; w->st_top0[s] = w->st_top[s] = top;
        for (int k = 0; k < 3; k++) { w->st_min[3 * s + k] = mn[k]; w->st_max[3 * s + k] = mx[k]; }
    }
}

/* ------------------------------------------------------------------ structures (src/lib.rs:1061-1079, :1143) */
static void warm_struct(LW* w, int s) {
    if (w->st_solved[s]) return;
    w->st_solved[s] = 1; w->st_warm[s] = 560; w->st_settle[s] = 1e9f;
}
static void st_activate(LW* w, int s) {
    if (s < 0 || s >= w->st_n) return;
    if (!w->st_solved[s]) warm_struct(w, s);
    if (w->st_active[s] < 4.0f) w->st_active[s] = 4.0f;
}
void ll_warm_all(LW* w) { for (int s = 0; s < w->st_n; s++) warm_struct(w, s); }
/* dbg_wake_all (tools/debug_exports.rs:107): every structure's load field and joints live */
void ll_wake_all(LW* w) { for (int s = 0; s < w->st_n; s++) st_activate(w, s); }
static void st_static_change(LW* w, int i) { int s = w->sid[i]; if (s < w->st_n) st_activate(w, s); }

/* ------------------------------------------------------------------ topology changes (src/lib.rs:697-769, :816-843) */
static void remove_brick(LW* w, int i) {
    if (w->mat[i] == 0) return;
    if (gget(w, w->bx[i], w->by[i], w->bz[i]) == i + 1) gset(w, w->bx[i], w->by[i], w->bz[i], 0);
    w->mat[i] = 0;
    if (w->body[i] < 0) {
        int s = w->sid[i];
        if (s < w->st_n) { w->st_conn[s] = 1; w->conn_any = 1; }
        st_static_change(w, i);
    } else w->dirty[w->body[i]] = 1;
}
static void remove_brick_raw(LW* w, int i) {
    if (gget(w, w->bx[i], w->by[i], w->bz[i]) == i + 1) gset(w, w->bx[i], w->by[i], w->bz[i], 0);
    w->mat[i] = 0; w->body[i] = -1;
}
static uint32_t damage_brick(LW* w, int i, float frac, float absolute) {
    uint32_t broke = 0;
    for (int d = 0; d < 6; d++) {
        int j = nbr(w, i, d);
        if (j < 0 || w->mat[j] == 0) continue;
        int o = (d & 1) == 0 ? i : j, a = d >> 1;
        float* bd = &w->bond[4 * o + a];
        if (*bd > 0.0f) {
            *bd -= w->bmax[4 * o + a] * frac + absolute;
            if (*bd <= w->bmax[4 * o + a] * 0.03f) { *bd = 0.0f; broke++; }
        }
    }
    float* g = &w->bond[4 * i + 3];
    if (*g > 0.0f) {
        *g -= w->bmax[4 * i + 3] * frac * 0.5f + absolute * 0.5f;
        if (*g <= 0.0f) { *g = 0.0f; broke++; }
    }
    if (w->body[i] < 0) {
        st_static_change(w, i);
        if (broke > 0) { int s = w->sid[i]; if (s < w->st_n) { w->st_conn[s] = 1; w->conn_any = 1; w->st_broke[s] = 1; } }
    } else if (broke > 0) w->dirty[w->body[i]] = 1;
    return broke;
}
static void sever_brick(LW* w, int i) {
    for (int k = 0; k < 4; k++) w->bond[4 * i + k] = 0.0f;
    static const int ds[3] = {1, 3, 5};
    for (int t = 0; t < 3; t++) { int j = nbr(w, i, ds[t]); if (j >= 0) w->bond[4 * j + (ds[t] >> 1)] = 0.0f; }
}
static void wake_body(LW* w, int b) { w->asleep[b] = 0; w->sleep[b] = 0.0f; }
static void kill_body(LW* w, int b) {
    uint32_t k = w->first[b];
    while (k != NONE) {
        uint32_t nx = w->next[k];
        if (w->body[k] == b) {
            if (gget(w, w->bx[k], w->by[k], w->bz[k]) == (int32_t)k + 1) gset(w, w->bx[k], w->by[k], w->bz[k], 0);
            w->mat[k] = 0; w->body[k] = -1;
        }
        k = nx;
    }
    w->alive[b] = 0; w->first[b] = NONE;
}
static void clamp_vel(LW* w, int b) {
    float v = lg_len(w->v[b]);
    if (v > 60.0f) w->v[b] = lg_mul(w->v[b], 60.0f / v);
    float a = lg_len(w->w[b]);
    float amax = lg_minf(25.0f, 16.0f / lg_maxf(w->rad[b], 0.3f));
    if (a > amax) w->w[b] = lg_mul(w->w[b], amax / a);
}
static void impulse_at(LW* w, int b, LV p, LV j) {
    w->v[b] = lg_madd(w->v[b], j, w->im[b]);
    LV r = lg_sub(p, w->pos[b]);
    w->w[b] = lg_add(w->w[b], lg_mv(&w->iiw[b], lg_cross(r, j)));
}
static void apply_blast_to_body(LW* w, int b) {
    if (w->blast_t <= 0.0f) return;
    uint32_t k = w->first[b];
    while (k != NONE) {
        LV p = w->wpos[k];
        LV d = lg_sub(p, w->blast_c);
        float dist = lg_len(d);
        if (dist < w->blast_r) {
            float f = w->blast_p * 1600.0f * (1.0f - dist / w->blast_r);
            LV dir = dist > 1e-3f ? lg_mul(d, 1.0f / dist) : lv(0, 1, 0);
            dir = lg_add(dir, lv(0.0f, 0.35f, 0.0f));
            float mm = lg_maxf(LG_MASS[w->mat[k]], 25.0f);
            impulse_at(w, b, p, lg_mul(dir, f * mm / 300.0f));
        }
        k = w->next[k];
    }
    clamp_vel(w, b);
}
static void wake_near(LW* w, LV p, float r) {
    for (int b = 0; b < w->hw; b++)
        if (w->alive[b] && w->asleep[b] && lg_len(lg_sub(w->pos[b], p)) < r + w->rad[b]) wake_body(w, b);
}

/* (src/lib.rs:930-1002) */
static void rebuild_body(LW* w, int id) {
    w->dirty[id] = 0;
    int n = 0;
    uint32_t k = w->first[id];
    while (k != NONE) {
        uint32_t nx = w->next[k];
        if (w->body[k] == id) { if (w->mat[k] != 0) w->lst[n++] = k; else w->body[k] = -1; }
        k = nx;
    }
    if (n == 0) { w->alive[id] = 0; w->first[id] = NONE; return; }
    uint32_t sid = ++w->stamp_id;
    uint32_t ncomp = 0;
    for (int t = 0; t < n; t++) {
        uint32_t s = w->lst[t];
        if (w->stamp[s] == sid) continue;
        w->stamp[s] = sid; w->comp[s] = ncomp;
        int qh = 0, qt = 0;
        w->q[qt++] = s;
        while (qh < qt) {
            int i = (int)w->q[qh++];
            for (int d = 0; d < 6; d++) {
                int j = linked(w, i, d, NULL);
                if (j >= 0 && w->stamp[j] != sid) { w->stamp[j] = sid; w->comp[j] = ncomp; w->q[qt++] = (uint32_t)j; }
            }
        }
        ncomp++;
    }
    int nc = (int)ncomp;
    for (int c = 0; c <= nc; c++) w->cnt[c] = 0;
    for (int t = 0; t < n; t++) w->cnt[w->comp[w->lst[t]] + 1]++;
    for (int c = 0; c < nc; c++) w->cnt[c + 1] += w->cnt[c];
    for (int t = 0; t < n; t++) { int i = w->lst[t]; uint32_t c = w->comp[i]; w->lst2[w->cnt[c]++] = (uint32_t)i; }
    LV ppos = w->pos[id], pv = w->v[id], pw = w->w[id];
    LM prot = w->rot[id];
    LQ pq = w->qq[id];
    for (int c = 0; c < nc; c++) {
        int s = c == 0 ? 0 : (int)w->cnt[c - 1];
        int e = (int)w->cnt[c];
        if (c > 0 && e - s == 1 && w->stats[3] > 1500.0f) {
            int i = w->lst2[s];
            int m = w->mat[i];
            if (m != LG_STEEL && lg_rnd(&w->rng) < lg_minf((w->stats[3] - 1500.0f) / 2000.0f, 0.85f)) { remove_brick_raw(w, i); continue; }
        }
        int b = c == 0 ? id : alloc_body(w);
        if (b < 0) { for (int t = s; t < e; t++) remove_brick_raw(w, w->lst2[t]); continue; }
        float keep_cool = w->cool[id], keep_lt = w->lt[id];
        setup_body(w, b, s, e, ppos, prot, pq, pv, pw);
        if (c == 0) { w->cool[b] = keep_cool; w->lt[b] = keep_lt; }
    }
}

/* (src/lib.rs:1085-1141) */
static void conn_struct(LW* w, int s) {
    uint32_t sid = ++w->stamp_id;
    int qh = 0, qt = 0;
    for (int i = w->st_lo[s]; i < w->st_hi[s]; i++)
        if (in_struct(w, i, s) && w->by[i] == 0 && w->bond[4 * i + 3] > 0.0f) { w->stamp[i] = sid; w->q[qt++] = (uint32_t)i; }
    while (qh < qt) {
        int i = (int)w->q[qh++];
        for (int d = 0; d < 6; d++) {
            int j = linked(w, i, d, NULL);
            if (j >= 0 && w->stamp[j] != sid) { w->stamp[j] = sid; w->q[qt++] = (uint32_t)j; }
        }
    }
    int retry = 0;
    for (int st = w->st_lo[s]; st < w->st_hi[s]; st++) {
        if (!in_struct(w, st, s) || w->stamp[st] == sid) continue;
        int n = 0;
        w->stamp[st] = sid; w->lst2[n++] = (uint32_t)st;
        int h = 0;
        while (h < n) {
            int i = (int)w->lst2[h++];
            for (int d = 0; d < 6; d++) {
                int j = linked(w, i, d, NULL);
                if (j >= 0 && w->stamp[j] != sid) { w->stamp[j] = sid; w->lst2[n++] = (uint32_t)j; }
            }
        }
        int b = alloc_body(w);
        if (b < 0) { retry = 1; continue; }
        for (int t = 0; t < n; t++) {
            int i = w->lst2[t];
            w->local[i] = lg_lat_pos(w->bx[i], w->by[i], w->bz[i]);
            w->phi[i] = 0.0f; w->ext[i] = 0.0f; w->extn[i] = 0.0f;
        }
        setup_body(w, b, 0, n, lv(0, 0, 0), IDM, IDQ, lv(0, 0, 0), lv(0, 0, 0));
        apply_blast_to_body(w, b);
        wake_near(w, w->pos[b], w->rad[b] + 1.5f);
    }
    if (retry) { w->st_conn[s] = 1; w->conn_any = 1; }
    uint32_t cnt = 0; int top = 0;
    for (int i = w->st_lo[s]; i < w->st_hi[s]; i++) if (in_struct(w, i, s)) { cnt++; if (w->by[i] + 1 > top) top = w->by[i] + 1; }
    w->st_now[s] = cnt; w->st_top[s] = top;
}
static void connectivity(LW* w) {
    if (!w->conn_any) return;
    w->conn_any = 0;
    for (int s = 0; s < w->st_n; s++) if (w->st_conn[s]) { w->st_conn[s] = 0; conn_struct(w, s); }
}

/* ------------------------------------------------------------------ rigid-body substep */
static void update_cache(LW* w) {
    for (int b = 0; b < w->hw; b++) {
        if (!w->alive[b] || w->asleep[b]) continue;
        LM r = lg_q_to_m(w->qq[b]);
        w->rot[b] = r;
        LM t = lg_mmul(&r, &w->iil[b]), rt = lg_mt(&r);
        w->iiw[b] = lg_mmul(&t, &rt);
        LV pos = w->pos[b];
        for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) w->wpos[k] = lg_add(pos, lg_mv(&r, w->local[k]));
    }
}

/* exact-cell neighbour lists: every brick of a live body (mat != 0), keyed by cell then by its place in body order */
#define HK_CELL(x, y, z) ((((uint64_t)((x) + 1024) & 2047) << 46) | (((uint64_t)((y) + 1024) & 2047) << 35) | (((uint64_t)((z) + 1024) & 2047) << 24))
static int cmp_u64(const void* a, const void* b) { uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b; return x < y ? -1 : x > y; }
static void build_hash(LW* w) {
    int n = 0;
    for (int b = 0; b < w->hw; b++) {
        if (!w->alive[b]) continue;
        for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) {
            if (w->mat[k] == 0) continue;
            LV p = w->wpos[k];
            uint64_t cell = HK_CELL(lg_fcell(p.x), lg_fcell(p.y), lg_fcell(p.z));
            w->hkeys[n] = cell | (uint64_t)n;   /* low 24 bits: place in body order */
            w->lst[n] = k;                      /* place -> brick */
            n++;
        }
    }
    w->nh = n;
    qsort(w->hkeys, (size_t)n, 8, cmp_u64);
}
static int hash_lower(const LW* w, uint64_t cell) {
    int lo = 0, hi = w->nh;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (w->hkeys[mid] < cell) lo = mid + 1; else hi = mid; }
    return lo;
}

static void push_contact(LW* w, int a, int b, int bo, uint32_t ba, uint32_t bb, LV p, LV n, float depth) {
    if (w->ncs >= w->ccap) { w->ccap *= 2; w->cs = realloc(w->cs, sizeof(LC) * (size_t)w->ccap); if (!w->cs) abort(); }
    LC* c = &w->cs[w->ncs++];
    memset(c, 0, sizeof *c);
    c->a = a; c->b = b; c->bo = bo; c->ba = ba; c->bb = bb; c->p = p; c->n = n; c->depth = depth; c->on = 1;
    c->col = b < 0 ? LL_CHAIN : -1;
    w->nc[a]++;
    if (b >= 0) w->nc[b]++;
}

/* (src/lib.rs:1396-1471) */
static void gen_contacts(LW* w) {
    w->ncs = 0;
    for (int b = 0; b < w->hw; b++) w->nc[b] = 0;
    for (int a = 0; a < w->hw; a++) {
        if (!w->alive[a] || w->asleep[a]) continue;
        LM rot = w->rot[a];
        float speed = lg_len(w->v[a]) + lg_len(w->w[a]) * w->rad[a] * 0.5f;
        for (uint32_t i = w->first[a]; i != NONE; i = w->next[i]) {
            LV p = w->wpos[i];
            if (p.y < 0.5f) {
                for (int c = 0; c < 8; c++) {
                    LV l = lv((c & 1) == 0 ? -LG_H : LG_H, (c & 2) == 0 ? -LG_H : LG_H, (c & 4) == 0 ? -LG_H : LG_H);
                    LV cp = lg_add(p, lg_mv(&rot, l));
                    if (cp.y < 0.0f) { push_contact(w, a, -1, -1, i, NONE, cp, lv(0, 1, 0), -cp.y); w->sup[i] = w->frame; }
                }
            }
            int cx = lg_fcell(p.x) + LG_NX2, cy = lg_fcell(p.y), cz = lg_fcell(p.z) + LG_NX2;
            for (int dz = -1; dz <= 1; dz++) for (int dy = -1; dy <= 1; dy++) for (int dx = -1; dx <= 1; dx++) {
                int x = cx + dx, y = cy + dy, z = cz + dz;
                int j = static_brick(w, x, y, z);
                if (j < 0) continue;
                LV c = lg_lat_pos(x, y, z);
                LV q = lv(lg_clampf(p.x, c.x - LG_H, c.x + LG_H), lg_clampf(p.y, c.y - LG_H, c.y + LG_H), lg_clampf(p.z, c.z - LG_H, c.z + LG_H));
                LV d = lg_sub(p, q);
                float dist = lg_len(d);
                if (dist >= LG_RB) continue;
                if (dist > 1e-4f) {
                    if (d.y > 0.3f * dist) w->sup[i] = w->frame;
                    push_contact(w, a, -1, -1, i, (uint32_t)j, q, lg_mul(d, 1.0f / dist), LG_RB - dist);
                } else {
                    LV r = lg_sub(p, c);
                    float rv[3] = {r.x, r.y, r.z};
                    int ax = 0;
                    for (int t = 1; t < 3; t++) if (lg_abs(rv[t]) > lg_abs(rv[ax])) ax = t;
                    float nv[3] = {0, 0, 0};
                    nv[ax] = rv[ax] >= 0.0f ? 1.0f : -1.0f;
                    push_contact(w, a, -1, -1, i, (uint32_t)j, p, lv(nv[0], nv[1], nv[2]), LG_H + LG_RB - lg_abs(rv[ax]));
                }
            }
            int hx = lg_fcell(p.x), hy = lg_fcell(p.y), hz = lg_fcell(p.z);
            for (int dz = -1; dz <= 1; dz++) for (int dy = -1; dy <= 1; dy++) for (int dx = -1; dx <= 1; dx++) {
                uint64_t cell = HK_CELL(hx + dx, hy + dy, hz + dz);
                for (int e = hash_lower(w, cell); e < w->nh && (w->hkeys[e] & ~0xFFFFFFull) == cell; e++) {
                    int j = (int)w->lst[w->hkeys[e] & 0xFFFFFFull];
                    int bj = w->body[j];
                    if (bj == a || w->mat[j] == 0) continue;
                    int other_awake = !w->asleep[bj];
                    if (other_awake && (uint32_t)j < i) continue;
                    LV d = lg_sub(p, w->wpos[j]);
                    float dd = lg_dot(d, d);
                    if (dd >= 4.0f * LG_RB * LG_RB || dd < 1e-10f) continue;
                    float dist = sqrtf(dd);
                    LV n = lg_mul(d, 1.0f / dist);
                    LV cp = lg_madd(w->wpos[j], n, dist * 0.5f);
                    if (n.y > 0.3f) w->sup[i] = w->frame; else if (n.y < -0.3f) w->sup[j] = w->frame;
                    if (other_awake) push_contact(w, a, bj, bj, i, (uint32_t)j, cp, n, 2.0f * LG_RB - dist);
                    else {
                        if (speed > 0.7f) w->wake[bj] = 1;
                        push_contact(w, a, -1, bj, i, (uint32_t)j, cp, n, 2.0f * LG_RB - dist);
                    }
                }
            }
        }
    }
}

static inline LV vel_at(const LW* w, int b, LV r) { return b < 0 ? lv(0, 0, 0) : lg_add(w->v[b], lg_cross(w->w[b], r)); }
static inline float eff_k(const LW* w, int a, LV ra, LV n) { return w->im[a] + lg_dot(n, lg_cross(lg_mv(&w->iiw[a], lg_cross(ra, n)), ra)); }
static float eff_mass(const LW* w, int a, int b, LV ra, LV rb, LV n) {
    float k = eff_k(w, a, ra, n);
    if (b >= 0) k += eff_k(w, b, rb, n);
    return k > 1e-12f ? 1.0f / k : 0.0f;
}
/* the same with every inverse mass and inertia split over the body's overflow contacts */
static float eff_mass_split(const LW* w, int a, float na, int b, float nb, LV ra, LV rb, LV n) {
    float k = eff_k(w, a, ra, n) * na;
    if (b >= 0) k += eff_k(w, b, rb, n) * nb;
    return k > 1e-12f ? 1.0f / k : 0.0f;
}
static void apply_imp(LW* w, int a, int b, LV ra, LV rb, LV p) {
    w->v[a] = lg_madd(w->v[a], p, w->im[a]);
    w->w[a] = lg_add(w->w[a], lg_mv(&w->iiw[a], lg_cross(ra, p)));
    if (b >= 0) {
        w->v[b] = lg_madd(w->v[b], p, -w->im[b]);
        w->w[b] = lg_sub(w->w[b], lg_mv(&w->iiw[b], lg_cross(rb, p)));
    }
}

/* (src/lib.rs:1493-1550) */
static void plough(LW* w) {
    for (int k = 0; k < w->ncs; k++) {
        LC c = w->cs[k];
        if (c.b >= 0 || c.bb == NONE || c.bo >= 0) continue;
        int j = (int)c.bb;
        if (w->mat[j] == 0 || w->body[j] >= 0) { w->cs[k].on = 0; continue; }
        int a = c.a;
        LV ra = lg_sub(c.p, w->pos[a]);
        float vn = -lg_dot(vel_at(w, a, ra), c.n);
        if (vn < 2.0f) continue;
        float meff = w->m[a] * 3.0f / lg_maxf((float)w->nc[a], 3.0f);
        float mom = lg_minf(meff, w->m[a]) * vn;
        int mj = w->mat[j];
        if (mj >= LG_SLUG || mom < LG_CAPV[mj] * 0.35f) continue;
        if (w->count[a] > 40) {
            float f = lg_clampf(mom / (LG_CAPV[mj] * 1.5f), 0.15f, 0.9f);
            damage_brick(w, j, f, 0.0f);
            continue;
        }
        sever_brick(w, j);
        int s = w->sid[j];
        if (s < w->st_n) { w->st_conn[s] = 1; w->conn_any = 1; }
        st_static_change(w, j);
        int b = alloc_body(w);
        if (b < 0) { w->cs[k].on = 0; continue; }
        w->local[j] = lg_lat_pos(w->bx[j], w->by[j], w->bz[j]);
        w->lst2[0] = (uint32_t)j;
        LV rv = lg_rndv(&w->rng);
        LV v = lg_madd(lg_mul(w->v[a], 0.75f), rv, 2.0f);
        LV ww = lg_mul(lg_rndv(&w->rng), 6.0f);
        setup_body(w, b, 0, 1, lv(0, 0, 0), IDM, IDQ, v, lv(0, 0, 0));
        w->w[b] = ww;
        float m = w->m[a];
        float f = m / (m + LG_MASS[mj] * 1.5f);
        w->v[a] = lg_mul(w->v[a], f);
        w->cs[k].on = 0;
        for (int t = k + 1; t < w->ncs; t++) if (w->cs[t].bb == (uint32_t)j) w->cs[t].on = 0;
    }
}

/* (src/lib.rs:1551-1577) */
static void prep_contacts(LW* w, float h) {
    for (int k = 0; k < w->ncs; k++) {
        LC* c = &w->cs[k];
        if (!c->on) continue;
        int a = c->a;
        c->ra = lg_sub(c->p, w->pos[a]);
        c->rb = c->b >= 0 ? lg_sub(c->p, w->pos[c->b]) : lv(0, 0, 0);
        c->mn = eff_mass(w, a, c->b, c->ra, c->rb, c->n);
        LV dv = lg_sub(vel_at(w, a, c->ra), vel_at(w, c->b, c->rb));
        float vn = lg_dot(dv, c->n);
        c->vn0 = lg_maxf(-vn, 0.0f);
        LV vt = lg_sub(dv, lg_mul(c->n, vn));
        float lt = lg_len(vt);
        if (lt > 1e-4f) c->t1 = lg_mul(vt, 1.0f / lt);
        else {
            LV t = lg_abs(c->n.x) < 0.9f ? lv(1, 0, 0) : lv(0, 1, 0);
            LV x = lg_cross(c->n, t);
            c->t1 = lg_mul(x, 1.0f / lg_len(x));
        }
        c->t2 = lg_cross(c->n, c->t1);
        c->mt1 = eff_mass(w, a, c->b, c->ra, c->rb, c->t1);
        c->mt2 = eff_mass(w, a, c->b, c->ra, c->rb, c->t2);
        float bias = lg_minf(0.2f / h * lg_maxf(c->depth - 0.01f, 0.0f), 1.0f);
        if (vn < -2.0f) bias = lg_maxf(bias, -0.12f * vn);
        c->bias = bias;
    }
}

/* one contact, Gauss-Seidel (src/lib.rs:1579-1601) */
static void solve_one(LW* w, LC* c) {
    int a = c->a;
    LV dv = lg_sub(vel_at(w, a, c->ra), vel_at(w, c->b, c->rb));
    float vn = lg_dot(dv, c->n);
    float lam = c->mn * (-vn + c->bias);
    float old = c->jn;
    c->jn = lg_maxf(old + lam, 0.0f);
    float d = c->jn - old;
    apply_imp(w, a, c->b, c->ra, c->rb, lg_mul(c->n, d));
    float lim = LG_MU * c->jn;
    dv = lg_sub(vel_at(w, a, c->ra), vel_at(w, c->b, c->rb));
    float l1 = -c->mt1 * lg_dot(dv, c->t1);
    float o1 = c->jt1; c->jt1 = lg_clampf(o1 + l1, -lim, lim);
    float l2 = -c->mt2 * lg_dot(dv, c->t2);
    float o2 = c->jt2; c->jt2 = lg_clampf(o2 + l2, -lim, lim);
    LV p = lg_add(lg_mul(c->t1, c->jt1 - o1), lg_mul(c->t2, c->jt2 - o2));
    apply_imp(w, a, c->b, c->ra, c->rb, p);
}

/* ------------------------------------------------------------------ device order: colouring and overflow */
static inline uint32_t mix32(uint32_t h) { h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16; return h; }
/* claim priority of contact k (its index among the scene's contacts) in round r: later rounds outrank earlier ones */
static inline uint64_t claim_prio(uint32_t k, int r) {
    return ((uint64_t)(r + 1) << 56) | ((uint64_t)(mix32(k * 0x9E3779B1u + (uint32_t)r * 0x85EBCA77u) >> 8) << 32) | (uint64_t)(0xFFFFFFFFu - k);
}
static void colour_contacts(LW* w, uint32_t* mask, uint64_t* claim, int* prop) {
    const int K = w->K;
    const uint32_t full = K >= 32 ? 0xFFFFFFFFu : ((1u << K) - 1u);
    for (int b = 0; b < w->hw; b++) mask[b] = 0;
    memset(claim, 0, sizeof(uint64_t) * (size_t)w->hw * K);
    for (int r = 0; r < w->R; r++) {
        for (int k = 0; k < w->ncs; k++) {
            LC* c = &w->cs[k];
            prop[k] = -1;
            if (!c->on || c->col != -1) continue;
            uint32_t used = mask[c->a] | (c->b >= 0 ? mask[c->b] : 0u);
            uint32_t fr = ~used & full;
            if (!fr) continue;
            int col = __builtin_ctz(fr);
            prop[k] = col;
            uint64_t pr = claim_prio((uint32_t)k, r);
            uint64_t* ca = &claim[(size_t)c->a * K + col];
            if (pr > *ca) *ca = pr;
            if (c->b >= 0) { uint64_t* cb = &claim[(size_t)c->b * K + col]; if (pr > *cb) *cb = pr; }
        }
        for (int k = 0; k < w->ncs; k++) {
            LC* c = &w->cs[k];
            int col = prop[k];
            if (col < 0) continue;
            uint64_t pr = claim_prio((uint32_t)k, r);
            if (claim[(size_t)c->a * K + col] != pr) continue;
            if (c->b >= 0 && claim[(size_t)c->b * K + col] != pr) continue;
            c->col = col;
        }
        for (int k = 0; k < w->ncs; k++) {
            LC* c = &w->cs[k];
            if (prop[k] < 0 || c->col != prop[k]) continue;
            mask[c->a] |= 1u << c->col;
            if (c->b >= 0) mask[c->b] |= 1u << c->col;
        }
    }
}
/* 64-lane halving-tree sum of one chunk of vectors (zero padded) */
static LV tree64(const LV* x, int n) {
    LV t[64];
    for (int l = 0; l < 64; l++) t[l] = l < n ? x[l] : lv(0, 0, 0);
    for (int h = 32; h >= 1; h >>= 1) for (int l = 0; l < h; l++) t[l] = lg_add(t[l], t[l + h]);
    return t[0];
}

typedef struct { int* off; int* ent; LV* dv; LV* dw; int* novf; } OvfLists;

/* per-body lists of overflow contact sides in contact order: entry = contact * 2 + side (0: body a, 1: body b) */
static void build_ovf_lists(LW* w, OvfLists* L) {
    int nbd = w->hw;
    for (int b = 0; b <= nbd; b++) L->off[b] = 0;
    for (int k = 0; k < w->ncs; k++) {
        LC* c = &w->cs[k];
        if (!c->on || c->col != -1) continue;
        L->off[c->a + 1]++;
        if (c->b >= 0) L->off[c->b + 1]++;
    }
    for (int b = 0; b < nbd; b++) { L->novf[b] = L->off[b + 1]; L->off[b + 1] += L->off[b]; }
    int* fill = (int*)xcalloc((size_t)nbd + 1, sizeof(int));
    for (int k = 0; k < w->ncs; k++) {
        LC* c = &w->cs[k];
        if (!c->on || c->col != -1) continue;
        L->ent[L->off[c->a] + fill[c->a]++] = 2 * k;
        if (c->b >= 0) L->ent[L->off[c->b] + fill[c->b]++] = 2 * k + 1;
    }
    free(fill);
    for (int k = 0; k < w->ncs; k++) {
        LC* c = &w->cs[k];
        if (!c->on || c->col != -1) continue;
        float na = (float)L->novf[c->a], nb = c->b >= 0 ? (float)L->novf[c->b] : 0.0f;
        c->smn = eff_mass_split(w, c->a, na, c->b, nb, c->ra, c->rb, c->n);
        c->smt1 = eff_mass_split(w, c->a, na, c->b, nb, c->ra, c->rb, c->t1);
        c->smt2 = eff_mass_split(w, c->a, na, c->b, nb, c->ra, c->rb, c->t2);
    }
}

/* one Jacobi pass over the overflow contacts: each sees private copies of its bodies with split inverse masses; the
 * body then takes the plain sum of its contacts' velocity changes (the mean of the copies) */
static void overflow_pass(LW* w, OvfLists* L) {
    for (int k = 0; k < w->ncs; k++) {
        LC* c = &w->cs[k];
        if (!c->on || c->col != -1) continue;
        int a = c->a, b = c->b;
        float na = (float)L->novf[a], nb = b >= 0 ? (float)L->novf[b] : 0.0f;
        LV va = w->v[a], wa = w->w[a], vb = lv(0, 0, 0), wb = lv(0, 0, 0);
        if (b >= 0) { vb = w->v[b]; wb = w->w[b]; }
        float ima = w->im[a] * na, imb = b >= 0 ? w->im[b] * nb : 0.0f;
        LV dv = lg_sub(lg_add(va, lg_cross(wa, c->ra)), b >= 0 ? lg_add(vb, lg_cross(wb, c->rb)) : lv(0, 0, 0));
        float vn = lg_dot(dv, c->n);
        float lam = c->smn * (-vn + c->bias);
        float old = c->jn;
        c->jn = lg_maxf(old + lam, 0.0f);
        LV pn = lg_mul(c->n, c->jn - old);
        va = lg_madd(va, pn, ima);
        wa = lg_add(wa, lg_mul(lg_mv(&w->iiw[a], lg_cross(c->ra, pn)), na));
        if (b >= 0) { vb = lg_madd(vb, pn, -imb); wb = lg_sub(wb, lg_mul(lg_mv(&w->iiw[b], lg_cross(c->rb, pn)), nb)); }
        float lim = LG_MU * c->jn;
        dv = lg_sub(lg_add(va, lg_cross(wa, c->ra)), b >= 0 ? lg_add(vb, lg_cross(wb, c->rb)) : lv(0, 0, 0));
        float l1 = -c->smt1 * lg_dot(dv, c->t1);
        float o1 = c->jt1; c->jt1 = lg_clampf(o1 + l1, -lim, lim);
        float l2 = -c->smt2 * lg_dot(dv, c->t2);
        float o2 = c->jt2; c->jt2 = lg_clampf(o2 + l2, -lim, lim);
        LV pt = lg_add(lg_mul(c->t1, c->jt1 - o1), lg_mul(c->t2, c->jt2 - o2));
        LV p = lg_add(pn, pt);
        L->dv[2 * k] = lg_mul(p, w->im[a]);
        L->dw[2 * k] = lg_mv(&w->iiw[a], lg_cross(c->ra, p));
        if (b >= 0) {
            L->dv[2 * k + 1] = lg_mul(p, -w->im[b]);
            LV t = lg_mv(&w->iiw[b], lg_cross(c->rb, p));
            L->dw[2 * k + 1] = lv(-t.x, -t.y, -t.z);
        }
    }
    LV bufv[64], bufw[64];
    for (int b = 0; b < w->hw; b++) {
        int o = L->off[b], n = L->novf[b];
        if (!n) continue;
        LV av = lv(0, 0, 0), aw = lv(0, 0, 0);
        for (int e0 = 0; e0 < n; e0 += 64) {
            int m = n - e0 < 64 ? n - e0 : 64;
            for (int l = 0; l < m; l++) { int ent = L->ent[o + e0 + l]; bufv[l] = L->dv[ent]; bufw[l] = L->dw[ent]; }
            av = lg_add(av, tree64(bufv, m));
            aw = lg_add(aw, tree64(bufw, m));
        }
        w->v[b] = lg_add(w->v[b], av);
        w->w[b] = lg_add(w->w[b], aw);
    }
}

static void solve_contacts(LW* w) {
    if (!dev_part(w, LL_MIX_SOLVE)) {
        for (int it = 0; it < 10; it++) for (int k = 0; k < w->ncs; k++) if (w->cs[k].on) solve_one(w, &w->cs[k]);
        return;
    }
    int nbd = w->hw > 0 ? w->hw : 1, nc = w->ncs > 0 ? w->ncs : 1;
    uint32_t* mask = xcalloc(nbd, 4);
    uint64_t* claim = xcalloc((size_t)nbd * w->K, 8);
    int* prop = xcalloc(nc, sizeof(int));
    colour_contacts(w, mask, claim, prop);
    OvfLists L;
    L.off = xcalloc((size_t)nbd + 1, sizeof(int)); L.novf = xcalloc(nbd, sizeof(int));
    L.ent = xcalloc((size_t)nc * 2, sizeof(int));
    L.dv = xcalloc((size_t)nc * 2, sizeof(LV)); L.dw = xcalloc((size_t)nc * 2, sizeof(LV));
    build_ovf_lists(w, &L);
    int any_ovf = 0;
    for (int b = 0; b < w->hw; b++) if (L.novf[b]) { any_ovf = 1; break; }
    for (int it = 0; it < 10; it++) {
        /* static chains: each body's contacts against static bricks, the ground and sleeping bodies, in generation
         * order (bodies do not meet here, so this loop equals one chain per body run side by side) */
        for (int k = 0; k < w->ncs; k++) if (w->cs[k].on && w->cs[k].col == LL_CHAIN) solve_one(w, &w->cs[k]);
        for (int col = 0; col < w->K; col++)
            for (int k = 0; k < w->ncs; k++) if (w->cs[k].on && w->cs[k].col == col) solve_one(w, &w->cs[k]);
        if (any_ovf) overflow_pass(w, &L);
    }
    free(mask); free(claim); free(prop); free(L.off); free(L.novf); free(L.ent); free(L.dv); free(L.dw);
}

/* (src/lib.rs:1746-1768) */
static void pinch_check(LW* w) {
    for (int b = 0; b < w->hw; b++) {
        if (!w->alive[b] || w->asleep[b]) continue;
        LV dw = lg_sub(w->w[b], w->wpre[b]);
        float tip = lg_len(dw) * w->rad[b];
        if (tip > 6.0f) w->w[b] = lg_madd(w->wpre[b], dw, 6.0f / tip);
        LV dv = lg_sub(w->v[b], w->vpre[b]);
        float m = lg_len(dv);
        if (m < 8.0f) continue;
        uint32_t f = w->first[b];
        if (f == NONE || w->mat[f] >= LG_SLUG) continue;
        if (w->count[b] <= 6 && w->blast_t <= 0.0f) kill_body(w, b);
        else w->v[b] = lg_madd(w->vpre[b], dv, 8.0f / m);
    }
}

/* (src/lib.rs:1643-1702) */
static uint32_t crush_at(LW* w, int start, float e, LV p, uint32_t cap) {
    int m0 = w->mat[start];
    if (m0 == 0 || m0 >= LG_SLUG) return 0;
    if (e < LG_EB[m0] * 0.15f) return 0;
    int body = w->body[start];
    int sid = w->sid[start];
    uint32_t st = ++w->stamp_id;
    w->q[0] = (uint32_t)start;
    int qn = 1; w->stamp[start] = st;
    int head = 0;
    uint32_t crushed = 0;
    uint32_t maxn = cap < 600 ? cap : 600;
    int ncr = 0;
    while (head < qn && crushed < maxn) {
        int best = head; float bd = 1e18f;
        int lim = head + 24 < qn ? head + 24 : qn;
        for (int t = head; t < lim; t++) { int i = (int)w->q[t]; LV d = lg_sub(bpos(w, i), p); float dd = lg_dot(d, d); if (dd < bd) { bd = dd; best = t; } }
        uint32_t tmp = w->q[head]; w->q[head] = w->q[best]; w->q[best] = tmp;
        int i = (int)w->q[head]; head++;
        int m = w->mat[i];
        if (m == 0) continue;
        float cost = LG_EB[m];
        if (e < cost) { damage_brick(w, i, 0.6f * e / cost, 0.0f); break; }
        e -= cost;
        w->lst2[ncr++] = (uint32_t)i; crushed++;
        for (int d = 0; d < 6; d++) {
            int j = linked(w, i, d, NULL);
            if (j < 0 || w->stamp[j] == st || w->body[j] != body || w->sid[j] != sid) continue;
            if (qn >= w->bcap) break;
            w->stamp[j] = st; w->q[qn++] = (uint32_t)j;
        }
    }
    if (ncr == 0) return 0;
    for (int t = 0; t < ncr; t++) {
        int i = (int)w->lst2[t];
        int m = w->mat[i];
        if (m == 0) continue;
        if (lg_rnd(&w->rng) < LG_PULV[m]) remove_brick(w, i);
        else {
            sever_brick(w, i);
            if (body < 0) { int s = w->sid[i]; if (s < w->st_n) { w->st_conn[s] = 1; w->conn_any = 1; } st_static_change(w, i); }
            else w->dirty[body] = 1;
        }
    }
    return crushed;
}

/* (src/lib.rs:1610-1642) */
static void impact_damage(LW* w) {
    uint32_t budget = 2500;
    for (int k = 0; k < w->ncs; k++) {
        LC c = w->cs[k];
        if (!c.on) continue;
        int ia = (int)c.ba;
        int ma = w->mat[ia];
        if (ma == 0 || ma >= LG_SLUG) continue;
        if (c.bb != NONE && c.bo < 0 && c.jn * c.vn0 > 4000.0f) {
            int j = (int)c.bb;
            if (w->mat[j] != 0 && w->body[j] < 0) st_activate(w, w->sid[j]);
        }
        if (w->mat[ia] == 0) continue;
        if (c.vn0 < 1.0f) continue;
        float e = 0.5f * c.jn * c.vn0;
        if (e < 150.0f) continue;
        int jb = c.bb != NONE ? (int)c.bb : -1;
        int mb = jb >= 0 ? w->mat[jb] : 0;
        float soft_a = 1.0f / LG_EB[ma];
        float soft_b = (jb < 0 || mb == 0 || mb >= LG_SLUG) ? 0.0f : 1.0f / LG_EB[mb];
        if (soft_a + soft_b <= 0.0f) continue;
        float ea = e * soft_a / (soft_a + soft_b);
        if (ea > 0.0f && budget > 0) { uint32_t n = crush_at(w, ia, ea, c.p, budget); budget -= n < budget ? n : budget; }
        if (soft_b > 0.0f && budget > 0) { uint32_t n = crush_at(w, jb, e - ea, c.p, budget); budget -= n < budget ? n : budget; }
    }
}

/* (src/lib.rs:1703-1745) */
static void integrate(LW* w, float h) {
    for (int b = 0; b < w->hw; b++) {
        if (!w->alive[b] || w->asleep[b]) continue;
        clamp_vel(w, b);
        if (w->cool[b] > 0.0f) w->cool[b] -= h;
        w->pos[b] = lg_madd(w->pos[b], w->v[b], h);
        w->qq[b] = lg_q_int(w->qq[b], w->w[b], h);
        w->v[b] = lg_mul(w->v[b], 1.0f - 0.01f * h);
        w->w[b] = lg_mul(w->w[b], 1.0f - 0.05f * h);
        int small = w->count[b] <= 3;
        if (w->nc[b] > 0 && small) {
            w->w[b] = lg_mul(w->w[b], 1.0f - lg_minf(7.0f * h, 0.5f));
            w->v[b] = lg_mul(w->v[b], 1.0f - lg_minf(2.0f * h, 0.3f));
        }
        float v2 = lg_dot(w->v[b], w->v[b]);
        float rr = w->rad[b];
        float w2 = lg_dot(w->w[b], w->w[b]) * rr * rr;
        float vt = small ? 0.16f : 0.06f, wt = small ? 0.3f : 0.04f;
        if (v2 < vt && w2 < wt && w->nc[b] > 0 && w->maxr[b] < 0.95f) {
            w->sleep[b] += h;
            if (w->sleep[b] > 0.45f) { w->asleep[b] = 1; w->v[b] = lv(0, 0, 0); w->w[b] = lv(0, 0, 0); }
        } else w->sleep[b] = 0.0f;
        LV p = w->pos[b];
        if (p.y < -15.0f || lg_abs(p.x) > 190.0f || lg_abs(p.z) > 190.0f || p.y > 120.0f) kill_body(w, b);
    }
    for (int b = 0; b < w->hw; b++) if (w->wake[b]) { w->wake[b] = 0; if (w->alive[b]) wake_body(w, b); }
}

static void substep(LW* w, float h) {
    w->hsub = h;
    update_cache(w);
    for (int b = 0; b < w->hw; b++) if (w->alive[b] && !w->asleep[b]) w->v[b].y -= LG_GR * h;
    build_hash(w);
    gen_contacts(w);
    plough(w);
    prep_contacts(w, h);
    for (int b = 0; b < w->hw; b++) if (w->alive[b]) { w->vpre[b] = w->v[b]; w->wpre[b] = w->w[b]; }
    solve_contacts(w);
    pinch_check(w);
    impact_damage(w);
    integrate(w, h);
}

/* ------------------------------------------------------------------ load field and joints */
static inline float cond_w(const LW* w, int i, int j, int a) {
    float k = lg_minf(LG_STIFF[w->mat[i]], LG_STIFF[w->mat[j]]);
    if (w->body[i] < 0) return k * (a == 1 ? 1.0f : 0.3f);
    float u = lg_abs(w->rot[w->body[i]].m[3 + a]);
    return k * (0.3f + 0.7f * u);
}
static inline float axis_up(const LW* w, int i, int a) {
    if (w->body[i] < 0) return a == 1 ? 1.0f : 0.0f;
    return w->rot[w->body[i]].m[3 + a];
}
/* (src/lib.rs:1005-1026) */
static float sor_brick(LW* w, int i, float omega) {
    int mi = w->mat[i];
    float sk = 0.0f, skp = 0.0f;
    for (int d = 0; d < 6; d++) {
        int j = linked(w, i, d, NULL);
        if (j < 0) continue;
        float k = cond_w(w, i, j, d >> 1);
        sk += k; skp += k * w->phi[j];
    }
    if (w->body[i] < 0) { if (w->by[i] == 0 && w->bond[4 * i + 3] > 0.0f) sk += LG_GROUNDK * LG_STIFF[mi]; }
    else if (w->frame - w->sup[i] < 10u) sk += LG_GROUNDK * LG_STIFF[mi];
    if (sk > 0.0f) {
        float target = (LG_MASS[mi] + w->ext[i] + skp) / sk;
        float d = omega * (target - w->phi[i]);
        w->phi[i] += d;
        return lg_abs(d);
    }
    return 0.0f;
}
/* one sweep over a list of bricks; returns max |update| / max(1, max phi) */
static float sweep_list(LW* w, const uint32_t* list, int n, float omega) {
    float md = 0.0f, mp = 1.0f;
    if (!dev_part(w, LL_MIX_SWEEP)) {
        for (int t = 0; t < n; t++) {
            int i = (int)list[t];
            float d = sor_brick(w, i, omega);
            if (d > md) md = d;
            if (w->phi[i] > mp) mp = w->phi[i];
        }
    } else {
        for (int par = 0; par < 2; par++)
            for (int t = 0; t < n; t++) {
                int i = (int)list[t];
                if (((w->bx[i] + w->by[i] + w->bz[i]) & 1) != par) continue;
                float d = sor_brick(w, i, omega);
                if (d > md) md = d;
                if (w->phi[i] > mp) mp = w->phi[i];
            }
    }
    return md / mp;
}
static int struct_list(LW* w, int s, uint32_t* out) {
    int n = 0;
    for (int i = w->st_lo[s]; i < w->st_hi[s]; i++) if (in_struct(w, i, s)) out[n++] = (uint32_t)i;
    return n;
}
static int body_list(LW* w, int b, uint32_t* out) {
    int n = 0;
    for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) if (w->mat[k] != 0) out[n++] = k;
    return n;
}
static int body_loaded(const LW* w, int b) {
    return w->alive[b] && w->count[b] >= 8 && !w->asleep[b] && w->nc[b] > 0 && w->mat[w->first[b]] < LG_SLUG;
}

/* stress and fatigue of brick i's three positive-axis bonds and its footing (src/lib.rs:1153-1227).
 * ev_crush: when non-null (device order), a compression crush is recorded there instead of being applied. */
static float stress_brick(LW* w, int i, float dt, int live, int* ev_crush) {
    int mi = w->mat[i];
    int dynb = w->body[i] >= 0;
    int s = w->sid[i];
    float maxr = 0.0f;
    for (int a = 0; a < 3; a++) {
        float hp = 0.0f;
        int j = linked(w, i, a * 2, &hp);
        if (j < 0) continue;
        float f = cond_w(w, i, j, a) * (w->phi[i] - w->phi[j]);
        float u = axis_up(w, i, a);
        float up = f > 0.0f ? lg_maxf(u, 0.0f) : lg_maxf(-u, 0.0f);
        float tens = lg_minf(LG_TENS[mi], LG_TENS[w->mat[j]]);
        float capm = 1.0f - (1.0f - tens) * up;
        float hpw = hp;
        if (dynb) {
            float cv = lg_minf(LG_CAPV[mi], LG_CAPV[w->mat[j]]), ch = lg_minf(LG_CAPH[mi], LG_CAPH[w->mat[j]]);
            float lat = a == 1 ? cv : ch;
            float au = lg_abs(u);
            hpw = hp * (cv * au + ch * (1.0f - au)) / lg_maxf(lat, 1.0f);
        }
        float r = lg_abs(f) / (hpw * capm);
        if (r > w->stress[i]) w->stress[i] = r;
        if (r > w->stress[j]) w->stress[j] = r;
        if (r > maxr) maxr = r;
        if (live && r > 1.0f) {
            float nhp = dynb ? hp - hp * dt * LG_DUCT[mi] * lg_minf(r - 0.85f, 40.0f) : hp - hp * dt * 2.2f * lg_minf(r - 0.85f, 2.5f);
            if (nhp < w->bmax[4 * i + a] * 0.08f) {
                if (up == 0.0f && lg_abs(u) > 0.5f) {
                    int k = LG_CAPV[w->mat[j]] < LG_CAPV[mi] ? j : i;
                    if (ev_crush) { ev_crush[0] = a; ev_crush[1] = k; return maxr > r ? maxr : r; }
                    if (w->crush_n < 300) {
                        w->crush_n++;
                        w->fail[0]++;
                        remove_brick(w, k);
                        return maxr > r ? maxr : r;
                    }
                    w->fail[0]++;
                } else w->fail[up > 0.0f ? 1 : 2]++;
                w->bond[4 * i + a] = 0.0f;
                if (dynb) w->dirty[w->body[i]] = 1;
                else if (s < w->st_n) { w->st_conn[s] = 1; w->conn_any = 1; w->st_broke[s] = 1; }
            } else w->bond[4 * i + a] = nhp;
        }
    }
    if (!dynb && w->by[i] == 0 && w->bond[4 * i + 3] > 0.0f) {
        float f = LG_GROUNDK * LG_STIFF[mi] * w->phi[i];
        float r = f / w->bond[4 * i + 3];
        if (r > w->stress[i]) w->stress[i] = r;
        if (r > maxr) maxr = r;
        if (live && r > 1.0f) {
            w->bond[4 * i + 3] -= w->bond[4 * i + 3] * dt * 2.2f * (r - 0.85f);
            if (w->bond[4 * i + 3] < w->bmax[4 * i + 3] * 0.08f) {
                w->fail[3]++;
                w->bond[4 * i + 3] = 0.0f;
                if (s < w->st_n) { w->st_conn[s] = 1; w->conn_any = 1; w->st_broke[s] = 1; }
            }
        }
    }
    return maxr;
}
/* judge a list of bricks; in device order every brick is judged on the state at the start of the pass and the
 * crushes are applied afterwards in list order */
static float stress_list(LW* w, const uint32_t* list, int n, float dt, int live) {
    float mr = 0.0f;
    if (!dev_part(w, LL_MIX_STRESS)) {
        for (int t = 0; t < n; t++) { float r = stress_brick(w, (int)list[t], dt, live, NULL); if (r > mr) mr = r; }
        return mr;
    }
    int* ev = (int*)xcalloc((size_t)(n ? n : 1) * 2, sizeof(int));
    for (int t = 0; t < n; t++) { ev[2 * t] = -1; float r = stress_brick(w, (int)list[t], dt, live, &ev[2 * t]); if (r > mr) mr = r; }
    for (int t = 0; t < n; t++) {
        if (ev[2 * t] < 0) continue;
        int i = (int)list[t], a = ev[2 * t], k = ev[2 * t + 1];
        w->fail[0]++;
        if (w->crush_n < 300) { w->crush_n++; remove_brick(w, k); continue; }
        w->bond[4 * i + a] = 0.0f;
        if (w->body[i] >= 0) w->dirty[w->body[i]] = 1;
        else { int s = w->sid[i]; if (s < w->st_n) { w->st_conn[s] = 1; w->conn_any = 1; w->st_broke[s] = 1; } }
    }
    free(ev);
    return mr;
}

/* (src/lib.rs:1228-1365) */
static void structural(LW* w, float dt) {
    int act[256];
    int na = 0;
    for (int s = 0; s < w->st_n; s++) if (w->st_active[s] > 0.0f && w->st_solved[s]) act[na++] = s;
    for (int t = 0; t < na; t++) { int s = act[t]; for (int i = w->st_lo[s]; i < w->st_hi[s]; i++) w->extn[i] = 0.0f; }
    for (int b = 0; b < w->hw; b++) {
        if (!w->alive[b]) continue;
        uint32_t ns = 0, nground = 0;
        for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) {
            LV p = w->wpos[k];
            if (p.y < LG_RB + 0.05f) nground++;
            float yb = p.y - LG_RB - 0.02f;
            int j = static_brick(w, lg_fcell(p.x) + LG_NX2, lg_fcell(yb), lg_fcell(p.z) + LG_NX2);
            if (j >= 0 && w->sid[j] < w->st_n && w->st_active[w->sid[j]] > 0.0f) ns++;
        }
        if (ns == 0) continue;
        float share = w->m[b] / (float)(ns + nground);
        for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) {
            LV p = w->wpos[k];
            float yb = p.y - LG_RB - 0.02f;
            int j = static_brick(w, lg_fcell(p.x) + LG_NX2, lg_fcell(yb), lg_fcell(p.z) + LG_NX2);
            if (j >= 0 && w->sid[j] < w->st_n && w->st_active[w->sid[j]] > 0.0f) w->extn[j] += share;
        }
    }
    for (int t = 0; t < na; t++) { int s = act[t]; for (int i = w->st_lo[s]; i < w->st_hi[s]; i++) w->ext[i] = w->ext[i] * 0.8f + w->extn[i] * 0.2f; }
    for (int b = 0; b < w->hw; b++) {
        if (!w->alive[b] || w->count[b] < 8) continue;
        int any = 0;
        for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) if (w->frame - w->sup[k] < 10u) { any = 1; break; }
        if (!any) for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) w->phi[k] = 0.0f;
    }
    uint32_t* list = w->lst;
    int64_t budget = 160000;
    for (int t = 0; t < na; t++) {
        int s = act[t];
        int size = w->st_hi[s] - w->st_lo[s];
        int nmin = w->st_resid[s] < 1e-4f ? (((w->frame + (uint32_t)s) % 3u) == 0 ? 1 : 0) : (size > 15000 ? 1 : 6);
        if (nmin == 0) continue;
        int n = struct_list(w, s, list);
        int k = 0;
        for (;;) {
            w->st_resid[s] = sweep_list(w, list, n, 1.85f);
            k++;
            budget -= size;
            if (k >= nmin && (w->st_resid[s] < 4e-4f || k >= 90 || budget <= 0)) break;
        }
    }
    float idt = 1.0f / lg_maxf(dt, 1e-3f);
    for (int b = 0; b < w->hw; b++) {
        if (!w->alive[b]) continue;
        LV al = lg_mul(lg_sub(w->v[b], w->vlast[b]), idt);
        LV aa = lg_mul(lg_sub(w->w[b], w->wlast[b]), idt);
        w->vlast[b] = w->v[b]; w->wlast[b] = w->w[b];
        if (!body_loaded(w, b)) continue;
        LV pb = w->pos[b];
        float nm = 0.0f;
        for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) {
            if (w->mat[k] == 0) continue;
            LV a = lg_add(al, lg_cross(aa, lg_sub(w->wpos[k], pb)));
            LV ge = lv(-a.x, -LG_GR - a.y, -a.z);
            float n = lg_minf(lg_len(ge) / LG_GR, 150.0f);
            if (n > nm) nm = n;
            w->ext[k] = LG_MASS[w->mat[k]] * (n - 1.0f);
        }
        w->nmax[b] = nm;
    }
    for (int b = 0; b < w->hw; b++) {
        if (!body_loaded(w, b) || w->count[b] > 60000) continue;
        if (w->resid[b] < 2e-4f && w->nmax[b] < 1.5f && ((w->frame + (uint32_t)b) % 4u) != 0) continue;
        int n = body_list(w, b, list);
        int k = 0;
        for (;;) {
            float r = sweep_list(w, list, n, 1.8f);
            w->resid[b] = r;
            k++;
            if (k >= 2 && (r < 4e-4f || k >= 60)) break;
        }
    }
    float gmax = 0.0f;
    for (int t = 0; t < na; t++) {
        int s = act[t];
        int n = struct_list(w, s, list);
        for (int e = 0; e < n; e++) w->stress[list[e]] = 0.0f;
        int live = w->st_settle[s] <= 0.0f && w->st_resid[s] < 1.5e-3f;
        float mr = stress_list(w, list, n, dt, live);
        w->st_maxr[s] = mr;
        if (mr > gmax) gmax = mr;
        if (w->st_settle[s] > 0.0f && w->st_warm[s] == 0) w->st_settle[s] -= dt;
        if (w->st_warm[s] == 0) w->st_active[s] -= dt;
        if (w->st_broke[s]) { w->st_broke[s] = 0; w->st_active[s] = 4.0f; }
        if (mr > 0.98f && w->st_active[s] < 1.0f) w->st_active[s] = 1.0f;
    }
    for (int b = 0; b < w->hw; b++) {
        if (!body_loaded(w, b) || w->count[b] > 60000) continue;
        int n = body_list(w, b, list);
        for (int e = 0; e < n; e++) w->stress[list[e]] = 0.0f;
        w->lt[b] += dt;
        int settled = w->resid[b] < 2e-4f && w->nmax[b] < 1.5f;
        if (settled && ((w->frame + (uint32_t)b) % 4u) != 0) continue;
        float dtb = settled ? dt * 4.0f : dt;
        int live = (w->lt[b] > 0.3f && (w->resid[b] < 1e-2f || w->lt[b] > 0.6f)) || w->nmax[b] > 2.5f;
        float bm = stress_list(w, list, n, dtb, live);
        w->maxr[b] = bm;
        if (bm > gmax) gmax = bm;
    }
    for (int b = 0; b < w->hw; b++) if (w->alive[b] && !body_loaded(w, b)) w->lt[b] = lg_maxf(w->lt[b] - 2.0f * dt, 0.0f);
    w->stats[6] = gmax;
    w->stats[13] = (float)na;
}

/* warm-up sweeps spread over frames, 120k brick updates per frame (warm_step, src/lib.rs:1068-1079) */
static void warm_step(LW* w) {
    int64_t budget = 120000;
    for (int s = 0; s < w->st_n; s++) {
        if (w->st_warm[s] <= 0 || budget <= 0) continue;
        int n = struct_list(w, s, w->lst);
        int size = w->st_hi[s] - w->st_lo[s];
        while (w->st_warm[s] > 0 && budget > 0) {
            float om = w->st_warm[s] > 60 ? 1.9f : 1.0f;
            w->st_resid[s] = sweep_list(w, w->lst, n, om);
            w->st_warm[s]--;
            budget -= size;
            if (w->st_warm[s] == 0) w->st_settle[s] = 0.3f;
        }
    }
}

static void write_stats(LW* w) {
    uint32_t dynb = 0, awake = 0, bodies = 0, stat = 0, unsolved = 0;
    for (int b = 0; b < w->hw; b++) if (w->alive[b]) { bodies++; dynb += w->count[b]; if (!w->asleep[b]) awake++; }
    for (int s = 0; s < w->st_n; s++) { stat += w->st_now[s]; if (!w->st_solved[s]) unsolved++; }
    w->stats[0] = (float)stat; w->stats[1] = (float)dynb; w->stats[2] = (float)bodies; w->stats[3] = (float)awake;
    w->stats[4] = (float)w->ncs; w->stats[14] = (float)unsolved; w->stats[15] = (float)w->st_n;
}

/* (src/lib.rs:3639-3721) */
void ll_step(LW* w, float dt) {
    dt = lg_clampf(dt, 0.0f, 1.0f / 30.0f);
    w->time += dt;
    w->crush_n = 0;
    w->frame += 1;
    if (w->blast_t > 0.0f) w->blast_t -= dt;
    for (int k = 0; k < 2; k++) substep(w, dt * 0.5f);
    for (int b = 0; b < w->hw; b++) if (w->alive[b] && w->dirty[b]) rebuild_body(w, b);
    connectivity(w);
    update_cache(w);
    structural(w, dt);
    warm_step(w);
    connectivity(w);
    update_cache(w);
    write_stats(w);
}

/* (src/lib.rs:2679-2716) */
static void blast_brick(LW* w, int i, LV c, float r, float power) {
    int m = w->mat[i];
    if (m == 0 || m >= LG_SLUG) return;
    LV p = bpos(w, i);
    float d = lg_len(lg_sub(p, c));
    float rr = r * LG_BLAST[m];
    if (d < rr * 0.5f) remove_brick(w, i);
    else if (d < rr * 1.35f) { float f = 1.0f - d / (rr * 1.35f); damage_brick(w, i, power * f * f * 1.3f, 0.0f); }
}
static int struct_near(const LW* w, int s, LV c, float r) {
    LV a = lg_sub(c, lv(r, r, r)), b = lg_add(c, lv(r, r, r));
    int x0 = lg_fcell(a.x) + LG_NX2, y0 = lg_fcell(a.y), z0 = lg_fcell(a.z) + LG_NX2;
    int x1 = lg_fcell(b.x) + LG_NX2, y1 = lg_fcell(b.y), z1 = lg_fcell(b.z) + LG_NX2;
    const int *mn = &w->st_min[3 * s], *mx = &w->st_max[3 * s];
    return !(x1 < mn[0] || x0 > mx[0] || y1 < mn[1] || y0 > mx[1] || z1 < mn[2] || z0 > mx[2]);
}
static int bodies_near(const LW* w, int b, LV c, float r) { return w->alive[b] && lg_len(lg_sub(w->pos[b], c)) < r + w->rad[b]; }
void ll_explode(LW* w, float x, float y, float z, float r, float power) {
    LV c = lv(x, y, z);
    w->blast_c = c; w->blast_r = r * 2.2f; w->blast_p = power; w->blast_t = 0.25f;
    float reach = r * 1.4f * 1.35f + 0.5f;
    for (int s = 0; s < w->st_n; s++) {
        if (!struct_near(w, s, c, reach)) continue;
        for (int i = w->st_lo[s]; i < w->st_hi[s]; i++) if (in_struct(w, i, s)) blast_brick(w, i, c, r, power);
    }
    for (int b = 0; b < w->hw; b++) {
        if (!bodies_near(w, b, c, reach)) continue;
        uint32_t k = w->first[b];
        while (k != NONE) { uint32_t nx = w->next[k]; blast_brick(w, (int)k, c, r, power); k = nx; }
    }
    for (int b = 0; b < w->hw; b++) {
        if (!bodies_near(w, b, c, w->blast_r)) continue;
        wake_body(w, b);
        apply_blast_to_body(w, b);
    }
}

/* a structure's load field taken from a converged solve: the warm-up is skipped (warm_struct then warm_step) */
void ll_set_field(LW* w, int s, const float* phi) {
    if (s < 0 || s >= w->st_n) return;
    for (int i = w->st_lo[s]; i < w->st_hi[s]; i++) w->phi[i] = phi[i - w->st_lo[s]];
    w->st_solved[s] = 1; w->st_warm[s] = 0; w->st_settle[s] = 0.3f; w->st_resid[s] = 0.0f;
}
void ll_failures(const LW* w, uint32_t* out4) { for (int k = 0; k < 4; k++) out4[k] = w->fail[k]; }

/* ------------------------------------------------------------------ readers */
void ll_stats(const LW* w, float* out16) { memcpy(out16, w->stats, sizeof w->stats); }
int ll_nbricks(const LW* w) { return w->nb; }
int ll_nbodies(const LW* w) { return w->hw; }
int ll_nstructs(const LW* w) { return w->st_n; }
/* per body slot: alive, asleep, count, nc, pos[3], q[4], v[3], w[3], m, sleep, maxr, resid, nmax (LL_BODY_F floats) */
void ll_bodies(const LW* w, float* out) {
    for (int b = 0; b < w->hw; b++) {
        float* o = out + (size_t)b * LL_BODY_F;
        o[0] = w->alive[b]; o[1] = w->asleep[b]; o[2] = (float)w->count[b]; o[3] = (float)w->nc[b];
        o[4] = w->pos[b].x; o[5] = w->pos[b].y; o[6] = w->pos[b].z;
        o[7] = w->qq[b].x; o[8] = w->qq[b].y; o[9] = w->qq[b].z; o[10] = w->qq[b].w;
        o[11] = w->v[b].x; o[12] = w->v[b].y; o[13] = w->v[b].z; o[14] = w->w[b].x; o[15] = w->w[b].y; o[16] = w->w[b].z;
        o[17] = w->m[b]; o[18] = w->sleep[b]; o[19] = w->maxr[b]; o[20] = w->resid[b]; o[21] = w->nmax[b];
    }
}
/* per brick: mat, body, wpos[3], phi, ext, bond[4], stress, sup age (LL_BRICK_F floats) */
void ll_bricks(const LW* w, float* out) {
    for (int i = 0; i < w->nb; i++) {
        float* o = out + (size_t)i * LL_BRICK_F;
        o[0] = w->mat[i]; o[1] = (float)w->body[i];
        o[2] = w->wpos[i].x; o[3] = w->wpos[i].y; o[4] = w->wpos[i].z;
        o[5] = w->phi[i]; o[6] = w->ext[i];
        for (int k = 0; k < 4; k++) o[7 + k] = w->bond[4 * i + k];
        o[11] = w->stress[i]; o[12] = (float)(w->frame - w->sup[i]);
    }
}
/* 12 floats per structure (src/lib.rs:3733-3748) */
void ll_structs(const LW* w, float* out) {
    for (int s = 0; s < w->st_n; s++) {
        float* o = out + 12 * s;
        LV pa = lg_lat_pos(w->st_min[3 * s], w->st_min[3 * s + 1], w->st_min[3 * s + 2]);
        LV pb = lg_lat_pos(w->st_max[3 * s], w->st_max[3 * s + 1], w->st_max[3 * s + 2]);
        o[0] = pa.x - LG_H; o[1] = pa.z - LG_H; o[2] = pb.x + LG_H; o[3] = pb.z + LG_H;
        o[4] = (float)w->st_init[s]; o[5] = (float)w->st_now[s]; o[6] = (float)w->st_top0[s]; o[7] = (float)w->st_top[s];
        o[8] = 0.0f; o[9] = 0.0f; o[10] = w->st_active[s] > 0.0f ? 1.0f : 0.0f; o[11] = (float)w->st_top0[s] * LG_S;
    }
}
/* per dominant material: pieces, bricks, largest, pieces >= 50, pieces <= 2 (tools/debug_exports.rs:41-56) */
void ll_matstats(const LW* w, float* out50) {
    for (int k = 0; k < 50; k++) out50[k] = 0.0f;
    for (int b = 0; b < w->hw; b++) {
        if (!w->alive[b]) continue;
        uint32_t cnt[9] = {0}, tot = 0;
        for (uint32_t k = w->first[b]; k != NONE; k = w->next[k]) if (w->mat[k] != 0 && w->mat[k] < 9) { cnt[w->mat[k]]++; tot++; }
        if (!tot) continue;
        int m = 1;
        for (int q = 1; q < 9; q++) if (cnt[q] > cnt[m]) m = q;
        if (m * 5 + 4 >= 50) continue;
        out50[m * 5] += 1.0f; out50[m * 5 + 1] += (float)tot;
        if ((float)tot > out50[m * 5 + 2]) out50[m * 5 + 2] = (float)tot;
        if (tot >= 50) out50[m * 5 + 3] += 1.0f;
        if (tot <= 2) out50[m * 5 + 4] += 1.0f;
    }
}
/* contacts of the last substep: a, b, bo, ba, bb, on, col, n[3], p[3], depth, jn, jt1, jt2, vn0 (LL_CONTACT_F) */
int ll_contacts(const LW* w, float* out, int cap) {
    if (out) for (int k = 0; k < w->ncs && k < cap; k++) {
        const LC* c = &w->cs[k];
        float* o = out + (size_t)k * LL_CONTACT_F;
        o[0] = (float)c->a; o[1] = (float)c->b; o[2] = (float)c->bo; o[3] = (float)(int32_t)c->ba; o[4] = c->bb == NONE ? -1.0f : (float)c->bb;
        o[5] = (float)c->on; o[6] = (float)c->col;
        o[7] = c->n.x; o[8] = c->n.y; o[9] = c->n.z; o[10] = c->p.x; o[11] = c->p.y; o[12] = c->p.z;
        o[13] = c->depth; o[14] = c->jn; o[15] = c->jt1; o[16] = c->jt2; o[17] = c->vn0;
    }
    return w->ncs;
}
uint32_t ll_rng(const LW* w) { return w->rng; }

/* ------------------------------------------------------------------ test hooks
 * Pieces of a substep run on their own, so the tests can pin each against what geometry and mechanics fix
 * (a sphere on a cube face, edge or corner; a single resting contact; the friction box; the pinch limits; a crush's
 * energy budget; the g - a load of a piece in free fall). None of these is part of a frame. */
void ll_set_body(LW* w, int b, const float* p3, const float* q4, const float* v3, const float* w3) {
    if (b < 0 || b >= w->hw || !w->alive[b]) return;
    if (p3) w->pos[b] = lv(p3[0], p3[1], p3[2]);
    if (q4) { w->qq[b].x = q4[0]; w->qq[b].y = q4[1]; w->qq[b].z = q4[2]; w->qq[b].w = q4[3]; }
    if (v3) w->v[b] = lv(v3[0], v3[1], v3[2]);
    if (w3) w->w[b] = lv(w3[0], w3[1], w3[2]);
    w->asleep[b] = 0;
    update_cache(w);
}
/* update_cache, build_hash and gen_contacts; returns the contact count (read them with ll_contacts) */
int ll_probe_contacts(LW* w) { update_cache(w); build_hash(w); gen_contacts(w); return w->ncs; }
/* prep_contacts for substep length h */
void ll_probe_prep(LW* w, float h) { w->hsub = h; prep_contacts(w, h); }
/* `passes` Gauss-Seidel passes over the contacts in generation order (solve_one, src/lib.rs:1579-1601) */
void ll_probe_solve(LW* w, int passes) {
    for (int it = 0; it < passes; it++) for (int k = 0; k < w->ncs; k++) if (w->cs[k].on) solve_one(w, &w->cs[k]);
}
/* body b's velocity and spin before the solve are set to vpre/wpre, then pinch_check runs (src/lib.rs:1746-1768) */
void ll_probe_pinch(LW* w, int b, const float* vpre3, const float* wpre3) {
    if (b < 0 || b >= w->hw) return;
    w->vpre[b] = lv(vpre3[0], vpre3[1], vpre3[2]);
    w->wpre[b] = lv(wpre3[0], wpre3[1], wpre3[2]);
    pinch_check(w);
}
/* crush_at from brick `start` with energy e at point p3 (src/lib.rs:1643-1702); returns the bricks crushed */
uint32_t ll_probe_crush(LW* w, int start, float e, const float* p3, uint32_t cap) {
    if (start < 0 || start >= w->nb) return 0;
    return crush_at(w, start, e, lv(p3[0], p3[1], p3[2]), cap);
}
/* the body velocity and spin a frame ago (the g - a loading differentiates against them) */
void ll_set_last(LW* w, int b, const float* v3, const float* w3) {
    if (b < 0 || b >= w->hw) return;
    w->vlast[b] = lv(v3[0], v3[1], v3[2]);
    w->wlast[b] = lv(w3[0], w3[1], w3[2]);
}
/* structural(dt) alone: resting loads, g - a, sweeps, joints (src/lib.rs:1228-1365) */
void ll_probe_structural(LW* w, float dt) { update_cache(w); structural(w, dt); }
/* per body slot: contacts counted for it in the last contact pass */
void ll_set_nc(LW* w, int b, uint32_t nc) { if (b >= 0 && b < w->hw) w->nc[b] = nc; }
/* put a body to sleep where it is (its bricks keep their world positions) */
void ll_set_asleep(LW* w, int b) { if (b >= 0 && b < w->hw && w->alive[b]) { w->asleep[b] = 1; w->v[b] = lv(0, 0, 0); w->w[b] = lv(0, 0, 0); } }
