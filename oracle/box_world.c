/* CPU oracle: box rigid-body step, restated from the reference's engine/src/phys.rs (World::step :152-243,
 * warm_start :245-262, solve :264-324, sleep_islands :326-354, broadphase_and_collide :356-409, wake :411-416,
 * collide_ground :418-432, collide_pair :434-448, push_manifold :450-498, tangents :501-504, reduce4 :507-527,
 * collide_boxes :531-628), generalised to compound bodies (one body, several box shapes at local poses, as in
 * gpu/gpu_world.js buildBodies :2212-2249) and per-shape friction mu = sqrt(mu_a mu_b) (gpu_world.js:601-603).
 * With one shape per body at the identity local pose and equal frictions every f32 result equals phys.rs.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): tests, the smoke check and the benchmark's CPU-baseline leg use it. */
#include "sf_oracle.h"
#include "om_math.h"
#include <stdlib.h>
#include <stdio.h>

#define NONE_U 0xFFFFFFFFu

typedef struct { V3 ra, rb, la, raxn, rbxn, iaxn, ibxn; float adj, nm, jn, rv; } Cp;  /* rv: approach speed at build */
typedef struct {
    int32_t a, b;   /* shapes (b = -1: ground) */
    int32_t ba, bb; /* bodies (bb = -1: ground) */
    int32_t sa, sb; /* solver slots: body + 1 when the body is in the solver, 0 = static dummy */
    V3 n, t1, t2; float mu; int cnt; Cp cp[4];
    V3 fa, fb, la_f, lb_f, e0;
    float tm1, tm2, twm, rad, jt1, jt2, jtw; int stick;
    V3 ca1, ca2, cb1, cb2, ia1, ia2, ib1, ib2, ian, ibn; float ima, imb;
    float rest; int pushed;  /* the pair's restitution; set once a relax pass pushed (gpu_world.js M_OUT .x, .y) */
    V3 cen; float smin;      /* centre of its points, least separation (contact reduction) */
    int red;                 /* reduced: kept for the next step's warm start, not coloured, solved or warm-started */
} Man;

typedef struct {
    V3 p; Quat q; V3 v, w; float inv_m; V3 inv_i; int awake; float sleep_t; V3 sleep_p; Quat sleep_q;
    M3 rot;
} Body;
typedef struct { int body; V3 lp; Quat lq; M3 lrot; V3 h; float mu, e; int hull; V3 wp; M3 wrot; V3 lo, hi; } Shape;

/* the chunk library's f32 tables, given by the caller (field by field, hull h at h times each stride) */
#define OH_V 16
#define OH_F 12
#define OH_FV 8
#define OH_S 24
#define OH_D 12
typedef struct { int count; float *v4, *pl4, *dir4; uint8_t *fidx, *fcnt, *seg; int32_t* head; } Hulls;

/* a debris particle (gpu_world.js:1676-1684 record): position, radius; velocity, state (0 free, 1 moving, 2 asleep);
 * drawing rotation; sub-step start position, rest timer; spin, friction; corrected position, touch (1 touching, 2
 * bearing a moving body); where the rest timer last restarted; ground flags; candidate lists */
#define OP_NN 24
#define OP_NS 12
typedef struct {
    V3 x; float r; V3 v; float st; Quat q; V3 xs; float t; V3 w; float mu; V3 xc; float touch; V3 anchor;
    uint32_t g_last, g_now; int nn, ns; int32_t nb[OP_NN], sh[OP_NS]; uint32_t cell;
} Particle;

struct ob_world {
    int order;
    float prm[OB_NPARAM];
    int nb, ns;
    Body* bd; Shape* sh;
    /* solver arrays indexed by slot (0 = static dummy, body i -> i + 1) */
    V3 *sv, *sw, *sdp, *sth, *sp0; float* sim; M3* sii; Quat* sq; int32_t* sidx;
    Man *man, *old; int nman, nold, capman, capold;
    int32_t* old_sorted; /* indices of old manifolds sorted by key */
    int32_t* cand; int ncand, capcand;
    int32_t* colour; int capcolour;
    int32_t* island;
    int32_t stats[8];
    uint8_t* wake_flag;
    int32_t *bl_off, *bl_ent, *ovf_index;  /* device order: per-slot contact lists, overflow position per manifold */
    uint32_t *lvl, *lvl2;  /* shock-propagation heights per slot (the two copies the passes alternate between) */
    int any_rest;          /* some shape has a restitution above zero */
    int nred;              /* manifolds reduced in the last step */
    Hulls hl;              /* convex-chunk tables (count 0: none given) */
    /* particle debris */
    Particle* pt; int np, capp; int32_t* pfree; int nfree;
    float blast[16][5]; int nblast; int pwake;
    int32_t pstats[4];
};

/* parameter indices (same as SF_P_* in include/sf.h) */
enum { P_GX, P_GY, P_GZ, P_SUBSTEPS, P_HERTZ, P_DAMPING, P_MAXPUSH, P_FRICTION, P_SPEC, P_SLEEP_LIN, P_SLEEP_ANG,
       P_SLEEP_TIME, P_SLEEP_DIST, P_SLEEP_ROT, P_ALLOW_SLEEP, P_CELL, P_LEAK, P_ANCHOR, P_ANCHOR_PUSH, P_ANCHOR_SLOP,
       P_COL_LIMIT, P_COL_ROUNDS, P_COL_SPREAD, P_TAIL_MIN, P_SOLVE_NT, P_RATIO, P_SHOCK, P_REST_THR, P_REST_ITERS,
       P_COL_SUPPORT, P_PSUB, P_PCELL, P_RED_K };

static void default_params(float* p) {
    /* phys.rs:101-105 */
    p[P_GX] = 0.0f; p[P_GY] = -9.81f; p[P_GZ] = 0.0f; p[P_SUBSTEPS] = 12; p[P_HERTZ] = 0.0f; p[P_DAMPING] = 10.0f;
    p[P_MAXPUSH] = 3.0f; p[P_FRICTION] = 0.6f; p[P_SPEC] = 0.02f; p[P_SLEEP_LIN] = 0.05f; p[P_SLEEP_ANG] = 0.08f;
    p[P_SLEEP_TIME] = 0.5f; p[P_SLEEP_DIST] = 0.01f; p[P_SLEEP_ROT] = 0.02f; p[P_ALLOW_SLEEP] = 1.0f; p[P_CELL] = 1.0f;
    p[P_LEAK] = 1.0f; p[P_ANCHOR] = 1.0f; p[P_ANCHOR_PUSH] = 3.0f; p[P_ANCHOR_SLOP] = 0.0f;
    p[P_COL_LIMIT] = 16; p[P_COL_ROUNDS] = 32; p[P_COL_SPREAD] = 4; p[P_TAIL_MIN] = 1024; p[P_SOLVE_NT] = 64;
    /* the GPU world's extensions (gpu_world.js), off: ratio 20, 6 level passes and 2 support rounds turn them on;
     * restitution acts where shapes carry one (threshold 1 m/s, 8 iterations, docs/GPU.md:41) */
    p[P_RATIO] = 0.0f; p[P_SHOCK] = 0.0f; p[P_REST_THR] = 1.0f; p[P_REST_ITERS] = 8.0f; p[P_COL_SUPPORT] = 0.0f;
    /* particle debris (gpu_world.js:2411-2412): 4 sub-steps, grid cell 2 x 0.3 m + 0.1 m */
    p[P_PSUB] = 4.0f; p[P_PCELL] = 0.7f;
    /* contact reduction (gpu_world.js:841): off; the GPU world reduces body pairs with more than 32 manifolds */
    p[P_RED_K] = 0.0f;
}

ob_world* ob_create(int order) {
    ob_world* w = (ob_world*)calloc(1, sizeof(ob_world));
    w->order = order;
    default_params(w->prm);
    return w;
}
static void free_hulls(Hulls* h) {
    free(h->v4); free(h->pl4); free(h->dir4); free(h->fidx); free(h->fcnt); free(h->seg); free(h->head);
    memset(h, 0, sizeof(Hulls));
}
int ob_set_hulls(ob_world* w, int count, const float* v4, const float* pl4, const float* dir4, const uint8_t* fidx,
                 const uint8_t* fcnt, const uint8_t* seg, const int32_t* head) {
    free_hulls(&w->hl);
    if (count <= 0) return 0;
    Hulls* h = &w->hl;
    h->count = count;
    h->v4 = (float*)malloc(sizeof(float) * 4 * OH_V * (size_t)count); memcpy(h->v4, v4, sizeof(float) * 4 * OH_V * (size_t)count);
    h->pl4 = (float*)malloc(sizeof(float) * 4 * OH_F * (size_t)count); memcpy(h->pl4, pl4, sizeof(float) * 4 * OH_F * (size_t)count);
    h->dir4 = (float*)malloc(sizeof(float) * 4 * OH_D * (size_t)count); memcpy(h->dir4, dir4, sizeof(float) * 4 * OH_D * (size_t)count);
    h->fidx = (uint8_t*)malloc((size_t)OH_F * OH_FV * count); memcpy(h->fidx, fidx, (size_t)OH_F * OH_FV * count);
    h->fcnt = (uint8_t*)malloc((size_t)OH_F * count); memcpy(h->fcnt, fcnt, (size_t)OH_F * count);
    h->seg = (uint8_t*)malloc((size_t)OH_S * 2 * count); memcpy(h->seg, seg, (size_t)OH_S * 2 * count);
    h->head = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)count); memcpy(h->head, head, sizeof(int32_t) * 4 * (size_t)count);
    return count;
}

void ob_destroy(ob_world* w) {
    free_hulls(&w->hl);
    if (!w) return;
    free(w->bd); free(w->sh); free(w->sv); free(w->sw); free(w->sdp); free(w->sth); free(w->sp0); free(w->sim);
    free(w->sii); free(w->sq); free(w->sidx); free(w->man); free(w->old); free(w->old_sorted); free(w->cand);
    free(w->colour); free(w->island); free(w->wake_flag); free(w->bl_off); free(w->bl_ent); free(w->ovf_index);
    free(w->lvl); free(w->lvl2); free(w->pt); free(w->pfree); free(w);
}
void ob_set_params(ob_world* w, const float* prm, int n) { for (int i = 0; i < n && i < OB_NPARAM; i++) w->prm[i] = prm[i]; }
int ob_nbody(const ob_world* w) { return w->nb; }

int ob_load(ob_world* w, int nbody, const float* bodies, int nshape, const int32_t* shape_body, const float* shapes) {
    w->nb = nbody; w->ns = nshape;
    w->bd = (Body*)realloc(w->bd, sizeof(Body) * (size_t)(nbody > 0 ? nbody : 1));
    w->sh = (Shape*)realloc(w->sh, sizeof(Shape) * (size_t)(nshape > 0 ? nshape : 1));
    for (int i = 0; i < nbody; i++) {
        const float* f = bodies + (size_t)i * OB_BODY_F;
        Body* b = &w->bd[i];
        b->p = v3(f[0], f[1], f[2]); b->q = quat(f[3], f[4], f[5], f[6]); b->v = v3(f[7], f[8], f[9]);
        b->w = v3(f[10], f[11], f[12]); b->inv_m = f[13]; b->inv_i = v3(f[14], f[15], f[16]);
        b->awake = f[17] > 0.5f; b->sleep_t = f[18]; b->sleep_p = b->p; b->sleep_q = b->q;
        b->rot = q_to_m3(b->q);
    }
    for (int s = 0; s < nshape; s++) {
        const float* f = shapes + (size_t)s * OB_SHAPE_F;
        Shape* sh = &w->sh[s];
        sh->body = shape_body[s];
        if (sh->body < 0 || sh->body >= nbody) return -1;
        if (s > 0 && shape_body[s] < shape_body[s - 1]) return -2;
        sh->lp = v3(f[0], f[1], f[2]); sh->lq = quat(f[3], f[4], f[5], f[6]); sh->h = v3(f[7], f[8], f[9]); sh->mu = f[10];
        sh->e = f[11];
        sh->hull = (int)f[12];
        if (sh->hull < 0 || (sh->hull > 0 && sh->hull >= w->hl.count)) return -3;
        sh->lrot = q_to_m3(sh->lq);
    }
    w->any_rest = 0;
    for (int s = 0; s < nshape; s++) if (w->sh[s].e > 0.0f) w->any_rest = 1;
    size_t n1 = (size_t)nbody + 1;
    w->sv = (V3*)realloc(w->sv, sizeof(V3) * n1); w->sw = (V3*)realloc(w->sw, sizeof(V3) * n1);
    w->sdp = (V3*)realloc(w->sdp, sizeof(V3) * n1); w->sth = (V3*)realloc(w->sth, sizeof(V3) * n1);
    w->sp0 = (V3*)realloc(w->sp0, sizeof(V3) * n1); w->sim = (float*)realloc(w->sim, sizeof(float) * n1);
    w->sii = (M3*)realloc(w->sii, sizeof(M3) * n1); w->sq = (Quat*)realloc(w->sq, sizeof(Quat) * n1);
    w->sidx = (int32_t*)realloc(w->sidx, sizeof(int32_t) * (size_t)(nbody > 0 ? nbody : 1));
    w->island = (int32_t*)realloc(w->island, sizeof(int32_t) * (size_t)(nbody > 0 ? nbody : 1));
    w->wake_flag = (uint8_t*)realloc(w->wake_flag, (size_t)(nbody > 0 ? nbody : 1));
    w->lvl = (uint32_t*)realloc(w->lvl, sizeof(uint32_t) * n1); w->lvl2 = (uint32_t*)realloc(w->lvl2, sizeof(uint32_t) * n1);
    for (size_t i = 0; i < n1; i++) { w->lvl[i] = 0; w->lvl2[i] = 0; }
    for (int i = 0; i < nbody; i++) w->island[i] = i;
    w->nman = 0; w->nold = 0;
    return 0;
}

void ob_read_bodies(const ob_world* w, float* out) {
    for (int i = 0; i < w->nb; i++) {
        const Body* b = &w->bd[i]; float* f = out + (size_t)i * OB_BODY_F;
        f[0] = b->p.x; f[1] = b->p.y; f[2] = b->p.z; f[3] = b->q.x; f[4] = b->q.y; f[5] = b->q.z; f[6] = b->q.w;
        f[7] = b->v.x; f[8] = b->v.y; f[9] = b->v.z; f[10] = b->w.x; f[11] = b->w.y; f[12] = b->w.z;
        f[13] = b->inv_m; f[14] = b->inv_i.x; f[15] = b->inv_i.y; f[16] = b->inv_i.z;
        f[17] = b->awake ? 1.0f : 0.0f; f[18] = b->sleep_t; f[19] = 0.0f;
    }
}
void ob_stats(const ob_world* w, int32_t* out8) { for (int i = 0; i < 8; i++) out8[i] = w->stats[i]; }
int ob_reduced(const ob_world* w) { return w->nred; }

/* ---------------------------------------------------------------- geometry helpers */
static void tangents(V3 n, V3* t1, V3* t2) {
    /* phys.rs:501-504 */
    V3 t = om_fabs(n.x) > 0.57f ? v3(n.y, -n.x, 0.0f) : v3(0.0f, n.z, -n.y);
    *t1 = vnorm(t);
    *t2 = vcross(n, *t1);
}

typedef struct { V3 p; float s; } Pt;

/* phys.rs:507-527 */
static int reduce4(Pt* pts, int cnt, V3 n) {
    if (cnt <= 4) return cnt;
    int i0 = 0; for (int k = 1; k < cnt; k++) if (pts[k].s < pts[i0].s) i0 = k;
    V3 p0 = pts[i0].p;
    int i1 = 0; float best = -1.0f;
    for (int k = 0; k < cnt; k++) { float d = vlen2(vsub(pts[k].p, p0)); if (d > best) { best = d; i1 = k; } }
    V3 p1 = pts[i1].p; V3 e = vsub(p1, p0);
    int i2 = 0; float b2 = -3.40282347e38f;
    for (int k = 0; k < cnt; k++) { float a = vdot(vcross(vsub(pts[k].p, p0), e), n); if (a > b2) { b2 = a; i2 = k; } }
    int i3 = 0; float b3 = -3.40282347e38f;
    for (int k = 0; k < cnt; k++) { float a = -vdot(vcross(vsub(pts[k].p, p0), e), n); if (a > b3) { b3 = a; i3 = k; } }
    int idx[4] = {i0, i1, i2, i3};
    Pt sel[4] = {pts[i0], pts[i1], pts[i2], pts[i3]};
    int out = 0;
    for (int k = 0; k < 4; k++) {
        int dup = 0; for (int q = 0; q < k; q++) if (idx[q] == idx[k]) dup = 1;
        if (dup) continue;
        pts[out++] = sel[k];
    }
    return out;
}

/* phys.rs:531-628. Returns the point count (0 = no contact); *n_out the normal from A to B. */
static int collide_boxes(V3 pa, const M3* ra, V3 ha, V3 pb, const M3* rb, V3 hb, float margin, Pt* out, V3* n_out) {
    V3 t = vsub(pb, pa);
    float ac[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) ac[i][j] = om_fabs(vdot(ra->c[i], rb->c[j])) + 1e-6f;
    float hav[3] = {ha.x, ha.y, ha.z}, hbv[3] = {hb.x, hb.y, hb.z};
    float sa = -3.40282347e38f; int ia = 0;
    for (int i = 0; i < 3; i++) {
        float s = om_fabs(vdot(t, ra->c[i])) - (hav[i] + hbv[0] * ac[i][0] + hbv[1] * ac[i][1] + hbv[2] * ac[i][2]);
        if (s > margin) return 0;
        if (s > sa) { sa = s; ia = i; }
    }
    float sbm = -3.40282347e38f; int jb = 0;
    for (int j = 0; j < 3; j++) {
        float s = om_fabs(vdot(t, rb->c[j])) - (hbv[j] + hav[0] * ac[0][j] + hav[1] * ac[1][j] + hav[2] * ac[2][j]);
        if (s > margin) return 0;
        if (s > sbm) { sbm = s; jb = j; }
    }
    float se = -3.40282347e38f; int ei = 0, ej = 0; V3 el = v3(0, 0, 0);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        V3 l = vcross(ra->c[i], rb->c[j]);
        float len2 = vlen2(l);
        if (len2 < 1e-8f) continue;
        l = vmul(l, 1.0f / om_sqrt(len2));
        float r = hav[0] * om_fabs(vdot(ra->c[0], l)) + hav[1] * om_fabs(vdot(ra->c[1], l)) + hav[2] * om_fabs(vdot(ra->c[2], l))
                + hbv[0] * om_fabs(vdot(rb->c[0], l)) + hbv[1] * om_fabs(vdot(rb->c[1], l)) + hbv[2] * om_fabs(vdot(rb->c[2], l));
        float s = om_fabs(vdot(t, l)) - r;
        if (s > margin) return 0;
        if (s > se) { se = s; ei = i; ej = j; el = l; }
    }
    const float face_tol = 0.005f;
    int use_b = sbm > sa + face_tol;
    float face_best = use_b ? sbm : sa;
    if (se > face_best + 0.01f) {
        V3 n = vdot(t, el) < 0.0f ? vneg(el) : el;
        V3 ca = pa, cb = pb;
        for (int k = 0; k < 3; k++) {
            if (k != ei) { float s = vdot(ra->c[k], n) > 0.0f ? 1.0f : -1.0f; ca = vadd(ca, vmul(ra->c[k], s * hav[k])); }
            if (k != ej) { float s = vdot(rb->c[k], n) > 0.0f ? -1.0f : 1.0f; cb = vadd(cb, vmul(rb->c[k], s * hbv[k])); }
        }
        V3 d1 = ra->c[ei], d2 = rb->c[ej];
        V3 r = vsub(ca, cb);
        float bb = vdot(d1, d2), c = vdot(d1, r), f = vdot(d2, r);
        float den = 1.0f - bb * bb;
        float s = den > 1e-6f ? (bb * f - c) / den : 0.0f;
        s = om_min(om_max(s, -hav[ei]), hav[ei]);
        float tt = bb * s + f;
        tt = om_min(om_max(tt, -hbv[ej]), hbv[ej]);
        s = om_min(om_max(bb * tt - c, -hav[ei]), hav[ei]);
        V3 p1 = vadd(ca, vmul(d1, s)), p2 = vadd(cb, vmul(d2, tt));
        float sep = vdot(vsub(p2, p1), n);
        out[0].p = vmul(vadd(p1, p2), 0.5f); out[0].s = sep;
        *n_out = n;
        return 1;
    }
    V3 pr, pi_; const M3 *rr, *ri; float hr[3], hi[3]; int k, flip;
    if (use_b) { pr = pb; rr = rb; pi_ = pa; ri = ra; k = jb; flip = 1; for (int q = 0; q < 3; q++) { hr[q] = hbv[q]; hi[q] = hav[q]; } }
    else { pr = pa; rr = ra; pi_ = pb; ri = rb; k = ia; flip = 0; for (int q = 0; q < 3; q++) { hr[q] = hav[q]; hi[q] = hbv[q]; } }
    V3 n = rr->c[k];
    if (vdot(vsub(pi_, pr), n) < 0.0f) n = vneg(n);
    int j = 0; float bd = -1.0f;
    for (int q = 0; q < 3; q++) { float d = om_fabs(vdot(ri->c[q], n)); if (d > bd) { bd = d; j = q; } }
    float sgn = vdot(ri->c[j], n) > 0.0f ? -1.0f : 1.0f;
    V3 fc = vadd(pi_, vmul(ri->c[j], sgn * hi[j]));
    int u = (j + 1) % 3, v = (j + 2) % 3;
    V3 eu = vmul(ri->c[u], hi[u]), ev = vmul(ri->c[v], hi[v]);
    V3 poly[8], tmp[8];
    poly[0] = vadd(vadd(fc, eu), ev); poly[1] = vadd(vsub(fc, eu), ev); poly[2] = vsub(vsub(fc, eu), ev); poly[3] = vsub(vadd(fc, eu), ev);
    int pn = 4;
    int u2 = (k + 1) % 3, v2 = (k + 2) % 3;
    int axs[4] = {u2, u2, v2, v2}; float sgs[4] = {1.0f, -1.0f, 1.0f, -1.0f};
    for (int pl = 0; pl < 4; pl++) {
        V3 dir = vmul(rr->c[axs[pl]], sgs[pl]);
        float off = vdot(dir, pr) + hr[axs[pl]];
        int tn = 0;
        for (int q = 0; q < pn; q++) {
            V3 p0 = poly[q], p1 = poly[(q + 1) % pn];
            float d0 = vdot(p0, dir) - off, d1 = vdot(p1, dir) - off;
            if (d0 <= 0.0f) { if (tn < 8) tmp[tn++] = p0; }
            if ((d0 <= 0.0f) != (d1 <= 0.0f)) {
                float tt = d0 / (d0 - d1);
                if (tn < 8) tmp[tn++] = vadd(p0, vmul(vsub(p1, p0), tt));
            }
        }
        pn = tn; for (int q = 0; q < pn; q++) poly[q] = tmp[q];
        if (pn == 0) return 0;
    }
    V3 rc = vadd(pr, vmul(n, hr[k]));
    int cnt = 0;
    for (int q = 0; q < pn; q++) {
        V3 p = poly[q];
        float sep = vdot(vsub(p, rc), n);
        if (sep <= margin) { out[cnt].p = vsub(p, vmul(n, 0.5f * sep)); out[cnt].s = sep; cnt++; }
    }
    if (cnt == 0) return 0;
    *n_out = flip ? vneg(n) : n;
    return cnt;
}

/* ---------------------------------------------------------------- per-step preparation */
static float shape_reach(const Shape* s) { return vlen(s->lp) + vlen(s->h); }

/* phys.rs:158-172 per shape: world pose, AABB with the speculative margin */
static void prep_shapes(ob_world* w, float dt) {
    const float spec = w->prm[P_SPEC];
    for (int i = 0; i < w->nb; i++) w->bd[i].rot = q_to_m3(w->bd[i].q);
    for (int s = 0; s < w->ns; s++) {
        Shape* sh = &w->sh[s]; const Body* b = &w->bd[sh->body];
        sh->wrot = m3_mul(&b->rot, &sh->lrot);
        sh->wp = vadd(b->p, m3_mulv(&b->rot, sh->lp));
        const M3* r = &sh->wrot; V3 h = sh->h;
        V3 e = v3(om_fabs(r->c[0].x) * h.x + om_fabs(r->c[1].x) * h.y + om_fabs(r->c[2].x) * h.z,
                  om_fabs(r->c[0].y) * h.x + om_fabs(r->c[1].y) * h.y + om_fabs(r->c[2].y) * h.z,
                  om_fabs(r->c[0].z) * h.x + om_fabs(r->c[1].z) * h.y + om_fabs(r->c[2].z) * h.z);
        float m = b->awake ? (vlen(b->v) + vlen(b->w) * shape_reach(sh)) * dt + 0.5f * spec : 0.5f * spec;
        e = vadd(e, v3(m, m, m));
        sh->lo = vsub(sh->wp, e); sh->hi = vadd(sh->wp, e);
    }
}

static int pair_ok(const ob_world* w, int ia, int ib) {
    const Shape *a = &w->sh[ia], *b = &w->sh[ib];
    if (ia == ib || a->body == b->body) return 0;
    const Body *ba = &w->bd[a->body], *bb = &w->bd[b->body];
    if (!ba->awake && !bb->awake) return 0;
    if (ba->inv_m == 0.0f && bb->inv_m == 0.0f) return 0;
    if (a->lo.x > b->hi.x || b->lo.x > a->hi.x || a->lo.y > b->hi.y || b->lo.y > a->hi.y || a->lo.z > b->hi.z || b->lo.z > a->hi.z) return 0;
    return 1;
}

/* phys.rs:366-404: hashed cells (every cell a shape's box overlaps), counting sort by slot in insertion order,
 * pairs reported in the cell of the max of the two min corners. Returns pairs (a < b) in slot order. */
typedef struct { uint32_t h; int32_t i; int32_t c[3]; } Ent;
static int32_t* hashed_pairs(ob_world* w, int* npairs) {
    const float inv_cell = 1.0f / w->prm[P_CELL];
    size_t ne = 0, cap = 1024;
    Ent* tmp = (Ent*)malloc(sizeof(Ent) * cap);
    for (int i = 0; i < w->ns; i++) {
        const Shape* s = &w->sh[i];
        int c0[3] = {(int32_t)om_floor(s->lo.x * inv_cell), (int32_t)om_floor(s->lo.y * inv_cell), (int32_t)om_floor(s->lo.z * inv_cell)};
        int c1[3] = {(int32_t)om_floor(s->hi.x * inv_cell), (int32_t)om_floor(s->hi.y * inv_cell), (int32_t)om_floor(s->hi.z * inv_cell)};
        for (int x = c0[0]; x <= c1[0]; x++) for (int y = c0[1]; y <= c1[1]; y++) for (int z = c0[2]; z <= c1[2]; z++) {
            if (ne == cap) { cap *= 2; tmp = (Ent*)realloc(tmp, sizeof(Ent) * cap); }
            uint32_t h = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)z * 83492791u;
            tmp[ne].h = h; tmp[ne].i = i; tmp[ne].c[0] = x; tmp[ne].c[1] = y; tmp[ne].c[2] = z; ne++;
        }
    }
    size_t tbl = 1; while (tbl < ne * 2) tbl <<= 1;
    uint32_t mask = (uint32_t)(tbl - 1);
    uint32_t* cnt = (uint32_t*)calloc(tbl + 1, sizeof(uint32_t));
    for (size_t e = 0; e < ne; e++) cnt[(tmp[e].h & mask) + 1]++;
    for (size_t k = 0; k < tbl; k++) cnt[k + 1] += cnt[k];
    Ent* ent = (Ent*)malloc(sizeof(Ent) * (ne ? ne : 1));
    uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * tbl);
    for (size_t k = 0; k < tbl; k++) fill[k] = cnt[k];
    for (size_t e = 0; e < ne; e++) { uint32_t s = tmp[e].h & mask; ent[fill[s]++] = tmp[e]; }
    size_t pcap = 1024, np = 0;
    int32_t* pairs = (int32_t*)malloc(sizeof(int32_t) * 2 * pcap);
    for (size_t s = 0; s < tbl; s++) {
        uint32_t lo = cnt[s], hi = cnt[s + 1];
        if (hi - lo < 2) continue;
        for (uint32_t x = lo; x < hi; x++) for (uint32_t y = x + 1; y < hi; y++) {
            const Ent *ea = &ent[x], *eb = &ent[y];
            if (ea->c[0] != eb->c[0] || ea->c[1] != eb->c[1] || ea->c[2] != eb->c[2] || ea->i == eb->i) continue;
            if (!pair_ok(w, ea->i, eb->i)) continue;
            V3 m = vmaxv(w->sh[ea->i].lo, w->sh[eb->i].lo);
            if ((int32_t)om_floor(m.x * inv_cell) != ea->c[0] || (int32_t)om_floor(m.y * inv_cell) != ea->c[1] ||
                (int32_t)om_floor(m.z * inv_cell) != ea->c[2]) continue;
            if (np == pcap) { pcap *= 2; pairs = (int32_t*)realloc(pairs, sizeof(int32_t) * 2 * pcap); }
            pairs[2 * np] = ea->i < eb->i ? ea->i : eb->i; pairs[2 * np + 1] = ea->i < eb->i ? eb->i : ea->i; np++;
        }
    }
    free(tmp); free(cnt); free(ent); free(fill);
    *npairs = (int)np;
    return pairs;
}

static int cmp_pair(const void* x, const void* y) {
    const int32_t *a = (const int32_t*)x, *b = (const int32_t*)y;
    if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
    uint32_t ua = (uint32_t)a[1] + 1u, ub = (uint32_t)b[1] + 1u; /* ground (-1) first */
    return ua < ub ? -1 : (ua > ub ? 1 : 0);
}

int ob_bruteforce_pairs(ob_world* w, float dt, int32_t* out, int cap) {
    prep_shapes(w, dt);
    int np = 0;
    for (int a = 0; a < w->ns; a++) {
        const Shape* s = &w->sh[a];
        if (w->bd[s->body].awake && s->lo.y < 0.0f) { if (np < cap) { out[2 * np] = a; out[2 * np + 1] = -1; } np++; }
        for (int b = a + 1; b < w->ns; b++) if (pair_ok(w, a, b)) { if (np < cap) { out[2 * np] = a; out[2 * np + 1] = b; } np++; }
    }
    return np;
}

/* ---------------------------------------------------------------- manifolds */
/* sort key: shape a, then the ground (b = -1), a chunk's second ground patch (b = -2), partners by index */
static uint64_t mkey64(int32_t a, int32_t b) { return ((uint64_t)(uint32_t)a << 32) | (uint64_t)(b >= 0 ? (uint32_t)b + 3u : (uint32_t)(-b)); }

static int find_old(const ob_world* w, int32_t a, int32_t b) {
    uint64_t k = mkey64(a, b);
    int lo = 0, hi = w->nold - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        const Man* m = &w->old[w->old_sorted[mid]];
        uint64_t km = mkey64(m->a, m->b);
        if (km == k) return w->old_sorted[mid];
        if (km < k) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

static Man* new_man(ob_world* w) {
    if (w->nman == w->capman) { w->capman = w->capman ? w->capman * 2 : 1024; w->man = (Man*)realloc(w->man, sizeof(Man) * (size_t)w->capman); }
    Man* m = &w->man[w->nman++];
    memset(m, 0, sizeof(Man));
    return m;
}

static float mix_mu(float a, float b) { return a == b ? a : om_sqrt(a * b); }

/* phys.rs:450-498 */
static void push_manifold(ob_world* w, int a, int b, V3 n, const Pt* pts, int cnt) {
    const Shape* sa_ = &w->sh[a];
    const Body* ba = &w->bd[sa_->body];
    V3 pa = ba->p; const M3* ra = &ba->rot;
    const Body* bb = b >= 0 ? &w->bd[w->sh[b].body] : NULL;
    V3 pb = bb ? bb->p : v3(0, 0, 0);
    int32_t sa = w->sidx[sa_->body], sb = bb ? w->sidx[w->sh[b].body] : 0;
    if (sa == 0 && sb == 0) return;
    Man* m = new_man(w);
    m->a = a; m->b = b; m->ba = sa_->body; m->bb = bb ? w->sh[b].body : -1; m->sa = sa; m->sb = sb;
    m->n = n; tangents(n, &m->t1, &m->t2);
    m->mu = mix_mu(sa_->mu, b >= 0 ? w->sh[b].mu : w->prm[P_FRICTION]);
    /* the larger restitution of the two (the ground's is 0), gpu_world.js:601-603 */
    m->rest = om_max(sa_->e, b >= 0 ? w->sh[b].e : 0.0f); m->pushed = 0;
    m->cnt = cnt; m->stick = w->prm[P_ANCHOR] > 0.5f;
    int oi = find_old(w, a, b);
    const Man* o = oi >= 0 ? &w->old[oi] : NULL;
    M3 rbm = bb ? bb->rot : m3_identity();
    V3 cen = v3(0, 0, 0); for (int k = 0; k < cnt; k++) cen = vadd(cen, pts[k].p); cen = vmul(cen, 1.0f / (float)cnt);
    m->cen = cen; m->smin = pts[0].s;
    for (int k = 1; k < cnt; k++) m->smin = pts[k].s < m->smin ? pts[k].s : m->smin;
    float rad = 0.0f; for (int k = 0; k < cnt; k++) rad += vlen(vsub(pts[k].p, cen)); m->rad = rad / (float)cnt;
    int anchored = 0;
    if (o) {
        if (o->stick) {
            V3 wa = vadd(pa, m3_mulv(ra, o->la_f));
            V3 wb = bb ? vadd(pb, m3_mulv(&rbm, o->lb_f)) : o->lb_f;
            float lim = om_max(0.05f, m->rad);
            if (vlen2(vsub(wa, cen)) < lim * lim && vlen2(vsub(wb, wa)) < 0.03f * 0.03f) {
                m->la_f = o->la_f; m->lb_f = o->lb_f; m->fa = vsub(wa, pa); m->fb = vsub(wb, pb); m->e0 = vsub(wb, wa); anchored = 1;
            }
        }
        V3 f = vadd(vmul(o->t1, o->jt1), vmul(o->t2, o->jt2));
        m->jt1 = vdot(f, m->t1); m->jt2 = vdot(f, m->t2); m->jtw = o->jtw * vdot(o->n, n);
    }
    if (!anchored) {
        m->fa = vsub(cen, pa); m->fb = vsub(cen, pb); m->la_f = m3_tmulv(ra, m->fa);
        m->lb_f = bb ? m3_tmulv(&rbm, m->fb) : cen;
        m->e0 = v3(0, 0, 0);
    }
    for (int k = 0; k < cnt; k++) {
        Cp* c = &m->cp[k];
        c->ra = vsub(pts[k].p, pa);
        c->rb = bb ? vsub(pts[k].p, pb) : pts[k].p;
        c->adj = pts[k].s;
        c->la = m3_tmulv(ra, c->ra);
        /* the point's normal velocity at the step start (negative: approaching), gpu_world.js:662 */
        {
            V3 pva = vadd(w->sv[sa], vcross(w->sw[sa], c->ra)), pvb = vadd(w->sv[sb], vcross(w->sw[sb], c->rb));
            c->rv = vdot(vsub(pvb, pva), n);
        }
        if (o) {
            float best = 0.03f * 0.03f; int bi = -1;
            for (int q = 0; q < o->cnt; q++) { float d = vlen2(vsub(o->cp[q].la, c->la)); if (d < best) { best = d; bi = q; } }
            if (bi >= 0) c->jn = o->cp[bi].jn;
        }
    }
}


/* ---------------------------------------------------------------- convex chunks (gpu/gpu_world.js:402-530) */
/* A shape with hull t > 0 is cell t of the chunk library in its box stretched by its half size h: vertex k in the
 * world is p + R (h * v_k); a face's world plane is R (n / h) normalised. A box paired with a chunk is the whole-cube
 * cell (t = 0). */
typedef struct { V3 p; M3 R; V3 h; int t; } HS;
static V3 hull_w(const Hulls* H, const HS* s, int k) {
    const float* v = H->v4 + 4 * ((size_t)s->t * OH_V + k);
    return vadd(s->p, m3_mulv(&s->R, v3(s->h.x * v[0], s->h.y * v[1], s->h.z * v[2])));
}
static void hull_face(const Hulls* H, const HS* s, int f, V3* n, float* d) {
    const float* pl = H->pl4 + 4 * ((size_t)s->t * OH_F + f);
    V3 nw = m3_mulv(&s->R, v3(pl[0] / s->h.x, pl[1] / s->h.y, pl[2] / s->h.z));
    float l = vlen(nw);
    *n = vmul(nw, 1.0f / l);
    *d = pl[3] / l + vdot(*n, s->p);
}
static int hull_fv(const Hulls* H, const HS* s, int f, int i) { return H->fidx[((size_t)s->t * OH_F + f) * OH_FV + i]; }
static int hull_fn(const Hulls* H, const HS* s, int f) { return H->fcnt[(size_t)s->t * OH_F + f]; }
static V3 hull_dir(const Hulls* H, const HS* s, int k) {
    const float* d = H->dir4 + 4 * ((size_t)s->t * OH_D + k);
    return vnorm(m3_mulv(&s->R, v3(s->h.x * d[0], s->h.y * d[1], s->h.z * d[2])));
}
/* up to 8 points; past 8 a deeper point replaces the shallowest */
static void keep8(Pt* pts, int* cnt, Pt t) {
    if (*cnt < 8) { pts[(*cnt)++] = t; return; }
    int hi = 0;
    for (int q = 1; q < 8; q++) if (pts[q].s > pts[hi].s) hi = q;
    if (t.s < pts[hi].s) pts[hi] = t;
}
/* gpu_world.js hull_ground :421-435: the corners below the margin */
static int hull_ground(const Hulls* H, const HS* s, float margin, Pt* pts) {
    int nv = H->head[4 * s->t], cnt = 0;
    for (int k = 0; k < nv; k++) {
        V3 x = hull_w(H, s, k);
        if (!(x.y < margin)) continue;
        Pt t; t.p = v3(x.x, 0.5f * x.y, x.z); t.s = x.y;
        keep8(pts, &cnt, t);
    }
    return cnt;
}
/* gpu_world.js hull_hull :437-530: face normals of both, then edge-direction cross products; a clipped reference
 * face (B's only if better by 0.005) or, for an edge axis better than both faces by 0.01, the closest points of the
 * support edges. Normal from A to B. */
static int hull_hull(const Hulls* H, const HS* A, const HS* B, float margin, Pt* pts, V3* n_out) {
    const int32_t *ha = H->head + 4 * A->t, *hb = H->head + 4 * B->t;
    V3 VA[OH_V], VB[OH_V], DA[OH_D], DB[OH_D];
    for (int k = 0; k < ha[0]; k++) VA[k] = hull_w(H, A, k);
    for (int k = 0; k < hb[0]; k++) VB[k] = hull_w(H, B, k);
    for (int k = 0; k < ha[3]; k++) DA[k] = hull_dir(H, A, k);
    for (int k = 0; k < hb[3]; k++) DB[k] = hull_dir(H, B, k);
    float sa = -3.40282347e38f, sb = -3.40282347e38f, se = -3.40282347e38f;
    int fa = 0, fb = 0, ei = 0, ej = 0;
    V3 el = v3(0, 0, 0);
    for (int f = 0; f < ha[1]; f++) {
        V3 pn; float pd; hull_face(H, A, f, &pn, &pd);
        float lo = 3.40282347e38f;
        for (int k = 0; k < hb[0]; k++) lo = om_min(lo, vdot(VB[k], pn));
        float sp = lo - pd;
        if (sp > margin) return 0;
        if (sp > sa) { sa = sp; fa = f; }
    }
    for (int f = 0; f < hb[1]; f++) {
        V3 pn; float pd; hull_face(H, B, f, &pn, &pd);
        float lo = 3.40282347e38f;
        for (int k = 0; k < ha[0]; k++) lo = om_min(lo, vdot(VA[k], pn));
        float sp = lo - pd;
        if (sp > margin) return 0;
        if (sp > sb) { sb = sp; fb = f; }
    }
    for (int i = 0; i < ha[3]; i++)
        for (int j = 0; j < hb[3]; j++) {
            V3 l0 = vcross(DA[i], DB[j]);
            float l2 = vdot(l0, l0);
            if (l2 < 1e-6f) continue;
            V3 l = vmul(l0, 1.0f / om_sqrt(l2));
            float alo = 3.40282347e38f, ahi = -3.40282347e38f, blo = 3.40282347e38f, bhi = -3.40282347e38f;
            for (int k = 0; k < ha[0]; k++) { float d = vdot(VA[k], l); alo = om_min(alo, d); ahi = om_max(ahi, d); }
            for (int k = 0; k < hb[0]; k++) { float d = vdot(VB[k], l); blo = om_min(blo, d); bhi = om_max(bhi, d); }
            float s1 = blo - ahi, s2 = alo - bhi, sp = s1;
            if (s2 > s1) { sp = s2; l = vneg(l); }
            if (sp > margin) return 0;
            if (sp > se) { se = sp; el = l; ei = i; ej = j; }
        }
    int use_b = sb > sa + 0.005f;
    float fbest = use_b ? sb : sa;
    if (se > fbest + 0.01f) {
        V3 da = DA[ei], db = DB[ej];
        V3 a0 = A->p, a1 = A->p, b0 = B->p, b1 = B->p;
        float best = -3.40282347e38f;
        for (int k = 0; k < ha[2]; k++) {
            int x = H->seg[((size_t)A->t * OH_S + k) * 2], y = H->seg[((size_t)A->t * OH_S + k) * 2 + 1];
            V3 p0 = VA[x], p1 = VA[y];
            if (fabsf(vdot(vnorm(vsub(p1, p0)), da)) < 0.999f) continue;
            float m = vdot(vadd(p0, p1), el);
            if (m > best) { best = m; a0 = p0; a1 = p1; }
        }
        best = 3.40282347e38f;
        for (int k = 0; k < hb[2]; k++) {
            int x = H->seg[((size_t)B->t * OH_S + k) * 2], y = H->seg[((size_t)B->t * OH_S + k) * 2 + 1];
            V3 p0 = VB[x], p1 = VB[y];
            if (fabsf(vdot(vnorm(vsub(p1, p0)), db)) < 0.999f) continue;
            float m = vdot(vadd(p0, p1), el);
            if (m < best) { best = m; b0 = p0; b1 = p1; }
        }
        V3 d1 = vsub(a1, a0), d2 = vsub(b1, b0), r = vsub(a0, b0);
        float aa = vdot(d1, d1), ee = vdot(d2, d2), ff = vdot(d2, r), cc = vdot(d1, r), bb = vdot(d1, d2);
        float den = aa * ee - bb * bb;
        float s1 = 0.0f;
        if (den > 1e-9f) s1 = om_min(om_max((bb * ff - cc * ee) / den, 0.0f), 1.0f);
        float t1 = (bb * s1 + ff) / om_max(ee, 1e-9f);
        if (t1 < 0.0f) { t1 = 0.0f; s1 = om_min(om_max(-cc / om_max(aa, 1e-9f), 0.0f), 1.0f); }
        else if (t1 > 1.0f) { t1 = 1.0f; s1 = om_min(om_max((bb - cc) / om_max(aa, 1e-9f), 0.0f), 1.0f); }
        V3 p1 = vadd(a0, vmul(d1, s1)), p2 = vadd(b0, vmul(d2, t1));
        pts[0].p = vmul(vadd(p1, p2), 0.5f);
        pts[0].s = vdot(vsub(p2, p1), el);
        *n_out = el;
        return 1;
    }
    const HS* Rs = use_b ? B : A;
    const HS* Is = use_b ? A : B;
    int fr = use_b ? fb : fa, nfi = use_b ? ha[1] : hb[1];
    V3 n; float rw;
    hull_face(H, Rs, fr, &n, &rw);
    int fi = 0;
    float bd = 3.40282347e38f;
    for (int f = 0; f < nfi; f++) { V3 fnv; float fdv; hull_face(H, Is, f, &fnv, &fdv); float d = vdot(fnv, n); if (d < bd) { bd = d; fi = f; } }
    V3 poly[16], tmp[16];
    int pn = hull_fn(H, Is, fi);
    for (int q = 0; q < pn; q++) poly[q] = hull_w(H, Is, hull_fv(H, Is, fi, q));
    int rn = hull_fn(H, Rs, fr);
    V3 rc = v3(0, 0, 0);
    for (int q = 0; q < rn; q++) rc = vadd(rc, hull_w(H, Rs, hull_fv(H, Rs, fr, q)));
    rc = vmul(rc, 1.0f / (float)rn);
    for (int e = 0; e < rn; e++) {
        V3 v0 = hull_w(H, Rs, hull_fv(H, Rs, fr, e)), v1 = hull_w(H, Rs, hull_fv(H, Rs, fr, (e + 1) % rn));
        V3 sn = vnorm(vcross(vsub(v1, v0), n));
        if (vdot(sn, vsub(rc, v0)) > 0.0f) sn = vneg(sn);
        int tn = 0;
        for (int q = 0; q < pn; q++) {
            V3 p0 = poly[q], p1 = poly[(q + 1) % pn];
            float d0 = vdot(vsub(p0, v0), sn), d1 = vdot(vsub(p1, v0), sn);
            if (d0 <= 0.0f && tn < 16) tmp[tn++] = p0;
            if ((d0 <= 0.0f) != (d1 <= 0.0f) && tn < 16) tmp[tn++] = vadd(p0, vmul(vsub(p1, p0), d0 / (d0 - d1)));
        }
        pn = tn;
        for (int q = 0; q < pn; q++) poly[q] = tmp[q];
        if (pn == 0) return 0;
    }
    int cnt = 0;
    for (int q = 0; q < pn; q++) {
        float sep = vdot(poly[q], n) - rw;
        if (sep > margin) continue;
        Pt t; t.p = vsub(poly[q], vmul(n, 0.5f * sep)); t.s = sep;
        keep8(pts, &cnt, t);
    }
    *n_out = use_b ? vneg(n) : n;
    return cnt;
}
static HS hshape(const Shape* s) { HS h; h.p = s->wp; h.R = s->wrot; h.h = s->h; h.t = s->hull; return h; }

/* One pair's contact by the chunk path (hull_hull, or hull_ground when b is NULL) or by the box SAT (collide_boxes;
 * the ground's eight corners), before reduce4: shapes as p[3] q[4] h[3] hull; out: n[3], count, then up to 8 x (p[3],
 * separation). For the checks that pin the chunk path against the box path and closed forms. */
static HS hs_of(const float* f) {
    HS h; h.p = v3(f[0], f[1], f[2]); h.R = q_to_m3(quat(f[3], f[4], f[5], f[6])); h.h = v3(f[7], f[8], f[9]); h.t = (int)f[10];
    return h;
}
int ob_pair_contact(const ob_world* w, int chunk_path, const float* a, const float* b, float margin, float* out) {
    Pt pts[8]; V3 n = v3(0, 0, 0); int cnt = 0;
    HS A = hs_of(a);
    if (chunk_path) {
        if (b) { HS B = hs_of(b); cnt = hull_hull(&w->hl, &A, &B, margin, pts, &n); }
        else { cnt = hull_ground(&w->hl, &A, margin, pts); n = v3(0.0f, -1.0f, 0.0f); }
    } else {
        if (b) { HS B = hs_of(b); cnt = collide_boxes(A.p, &A.R, A.h, B.p, &B.R, B.h, margin, pts, &n); }
        else {
            for (int k = 0; k < 8; k++) {
                V3 sg = v3((k & 1) ? 1.0f : -1.0f, (k & 2) ? 1.0f : -1.0f, (k & 4) ? 1.0f : -1.0f);
                V3 c = vadd(A.p, m3_mulv(&A.R, v3(sg.x * A.h.x, sg.y * A.h.y, sg.z * A.h.z)));
                if (c.y < margin) { pts[cnt].p = v3(c.x, 0.5f * c.y, c.z); pts[cnt].s = c.y; cnt++; }
            }
            n = v3(0.0f, -1.0f, 0.0f);
        }
    }
    out[0] = n.x; out[1] = n.y; out[2] = n.z; out[3] = (float)cnt;
    for (int k = 0; k < cnt; k++) { out[4 + 4 * k] = pts[k].p.x; out[5 + 4 * k] = pts[k].p.y; out[6 + 4 * k] = pts[k].p.z; out[7 + 4 * k] = pts[k].s; }
    return cnt;
}

/* phys.rs:411-416 */
static void wake_body(ob_world* w, int bi) {
    Body* b = &w->bd[bi];
    if (b->awake || b->inv_m == 0.0f) return;
    b->awake = 1; b->sleep_t = 0.0f; b->sleep_p = b->p; b->sleep_q = b->q;
}

/* phys.rs:418-432 */
static void collide_ground(ob_world* w, int i) {
    const Shape* s = &w->sh[i]; const Body* b = &w->bd[s->body];
    Pt pts[8]; int cnt = 0;
    float margin = w->prm[P_SPEC] + om_max(-b->v.y, 0.0f) * (1.0f / 60.0f);
    if (s->hull) {
        /* a chunk's margin also covers its corners' sweep from spin (gpu_world.js ground_margin :776-778) */
        HS hs = hshape(s);
        margin = w->prm[P_SPEC] + (om_max(-b->v.y, 0.0f) + vlen(b->w) * shape_reach(s)) * (1.0f / 60.0f);
        cnt = hull_ground(&w->hl, &hs, margin, pts);
        if (cnt == 0) return;
        /* up to 8 corners as two manifolds of 4 in the order found (gpu_world.js narrow_hull :813-823); the second
         * keyed as b = -2 */
        V3 n = v3(0.0f, -1.0f, 0.0f);
        push_manifold(w, i, -1, n, pts, cnt < 4 ? cnt : 4);
        if (cnt > 4) push_manifold(w, i, -2, n, pts + 4, cnt - 4);
        return;
    }
    for (int k = 0; k < 8; k++) {
        V3 sg = v3((k & 1) ? 1.0f : -1.0f, (k & 2) ? 1.0f : -1.0f, (k & 4) ? 1.0f : -1.0f);
        V3 c = vadd(s->wp, m3_mulv(&s->wrot, v3(sg.x * s->h.x, sg.y * s->h.y, sg.z * s->h.z)));
        if (c.y < margin) { pts[cnt].p = v3(c.x, 0.5f * c.y, c.z); pts[cnt].s = c.y; cnt++; }
    }
    if (cnt == 0) return;
    V3 n = v3(0.0f, -1.0f, 0.0f);
    cnt = reduce4(pts, cnt, n);
    push_manifold(w, i, -1, n, pts, cnt);
}

/* phys.rs:434-448; `snap` holds the step-start awake/still flags in the coloured order (NULL: live flags) */
static void collide_pair(ob_world* w, int ia, int ib, const uint8_t* snap) {
    const Shape *sa = &w->sh[ia], *sb = &w->sh[ib];
    const Body *a = &w->bd[sa->body], *b = &w->bd[sb->body];
    int a_awake, b_awake, a_still, b_still;
    if (snap) { a_awake = snap[sa->body] & 1; b_awake = snap[sb->body] & 1; a_still = (snap[sa->body] >> 1) & 1; b_still = (snap[sb->body] >> 1) & 1; }
    else { a_awake = a->awake; b_awake = b->awake; a_still = a->sleep_t != 0.0f; b_still = b->sleep_t != 0.0f; }
    float rel = vlen(vsub(b->v, a->v)) + vlen(a->w) * shape_reach(sa) + vlen(b->w) * shape_reach(sb);
    float margin = w->prm[P_SPEC] + rel * (1.0f / 60.0f);
    Pt pts[8]; V3 n;
    int cnt;
    if (sa->hull || sb->hull) { HS A = hshape(sa), B = hshape(sb); cnt = hull_hull(&w->hl, &A, &B, margin, pts, &n); }
    else cnt = collide_boxes(sa->wp, &sa->wrot, sa->h, sb->wp, &sb->wrot, sb->h, margin, pts, &n);
    if (!cnt) return;
    if (snap) {
        if (a_awake && !b_awake && !a_still) w->wake_flag[sb->body] = 1;
        if (b_awake && !a_awake && !b_still) w->wake_flag[sa->body] = 1;
    } else {
        if (a_awake && !b_awake && !a_still) wake_body(w, sb->body);
        if (b_awake && !a_awake && !b_still) wake_body(w, sa->body);
    }
    cnt = reduce4(pts, cnt, n);
    push_manifold(w, ia, ib, n, pts, cnt);
}

/* ---------------------------------------------------------------- contact reduction */
/* gpu_world.js:834-920. A body pair (the ground counting as one body) touching through more than K manifolds keeps
 * only representative ones: for each of the 26 directions d to the neighbours of a cube cell (corners, edges, faces)
 * the manifold whose points' centre c lies furthest out (c . d largest), and the deepest (least separation). Ties go
 * to the later manifold (manifold order is the canonical candidate order, so later = larger shape key). The others
 * are marked reduced: colouring, the solve, the warm start and the level and island passes skip them, while their
 * points and impulses stay in the step's manifolds for the next step's warm-start match. */
static inline uint32_t f2o(float x) { uint32_t b; memcpy(&b, &x, 4); return (b & 0x80000000u) ? ~b : (b | 0x80000000u); }
/* direction k of 26 (the cell offsets j = 0..26 in x-fastest order, the centre j = 13 skipped) */
static inline V3 red_dir(int k) {
    int j = k >= 13 ? k + 1 : k;
    return v3((float)(j % 3) - 1.0f, (float)((j / 3) % 3) - 1.0f, (float)(j / 9) - 1.0f);
}
static inline uint64_t pair_key(const Man* m) { return ((uint64_t)(uint32_t)m->ba << 32) | (uint64_t)(m->bb >= 0 ? (uint32_t)m->bb : 0xFFFFFFFFu); }
static const Man* g_red_base;
static int cmp_red(const void* x, const void* y) {
    int a = *(const int32_t*)x, b = *(const int32_t*)y;
    uint64_t ka = pair_key(&g_red_base[a]), kb = pair_key(&g_red_base[b]);
    if (ka != kb) return ka < kb ? -1 : 1;
    return a < b ? -1 : (a > b ? 1 : 0);
}
static int reduce_pairs(ob_world* w) {
    int M = w->nman, K = w->prm[P_RED_K] >= 1.0f ? (int)w->prm[P_RED_K] : 0, nred = 0;
    if (K > 0 && K < 8) K = 8;  /* (smaller thresholds act as 8) */
    for (int m = 0; m < M; m++) w->man[m].red = 0;
    if (K <= 0 || M <= K) return 0;
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)M);
    for (int m = 0; m < M; m++) idx[m] = m;
    g_red_base = w->man;
    qsort(idx, (size_t)M, sizeof(int32_t), cmp_red);
    for (int r0 = 0; r0 < M;) {
        int r1 = r0 + 1;
        uint64_t key = pair_key(&w->man[idx[r0]]);
        while (r1 < M && pair_key(&w->man[idx[r1]]) == key) r1++;
        if (r1 - r0 > K) {
            uint64_t best[27] = {0};
            for (int t = r0; t < r1; t++) {
                const Man* m = &w->man[idx[t]];
                for (int k = 0; k < 26; k++) {
                    V3 d = red_dir(k);
                    float sc = m->cen.x * d.x + m->cen.y * d.y + m->cen.z * d.z;
                    uint64_t v = ((uint64_t)f2o(sc) << 32) | (uint32_t)idx[t];
                    if (v > best[k]) best[k] = v;
                }
                uint64_t v = ((uint64_t)f2o(-m->smin) << 32) | (uint32_t)idx[t];
                if (v > best[26]) best[26] = v;
            }
            for (int t = r0; t < r1; t++) w->man[idx[t]].red = 1;
            for (int k = 0; k < 27; k++) w->man[(uint32_t)best[k]].red = 0;
            for (int t = r0; t < r1; t++) nred += w->man[idx[t]].red;
        }
        r0 = r1;
    }
    free(idx);
    return nred;
}

/* ---------------------------------------------------------------- colouring (device order) */
/* Deterministic claim rounds. Each round every uncoloured manifold picks one of the lowest `spread` colours free at
 * both its moving bodies (by a hash of its key and the round) and claims it at each with a unique priority; it takes
 * the colour where its claim is the highest at both. Manifolds with no free colour, or none after the last round,
 * go to the overflow (solved one by one in manifold order).
 * Supports first (gpu_world.js:954-985): with RS = P_COL_SUPPORT > 0 the first RS rounds are for supports alone
 * (manifolds with a side that is not moving: the ground, a static or sleeping body), and supports outrank the
 * others in every round. A round in which nothing claims changes nothing more in its phase: the support phase then
 * goes on to the first round for all, the last phase ends. */
static inline uint32_t mix32(uint32_t h) { h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16; return h; }
static inline uint32_t key_hash(int32_t a, int32_t b) { return mix32((uint32_t)a * 0x9E3779B9u + mix32((uint32_t)b ^ 0x5BD1E995u)); }

static void colour_manifolds(ob_world* w) {
    int K = (int)w->prm[P_COL_LIMIT]; if (K < 1) K = 1; if (K > 64) K = 64;
    int R = (int)w->prm[P_COL_ROUNDS]; int spread = (int)w->prm[P_COL_SPREAD]; if (spread < 1) spread = 1;
    int RS = (int)w->prm[P_COL_SUPPORT]; if (RS < 0) RS = 0;
    int M = w->nman;
    if (w->capcolour < M) { w->capcolour = M; w->colour = (int32_t*)realloc(w->colour, sizeof(int32_t) * (size_t)M); }
    uint64_t* mask = (uint64_t*)calloc((size_t)w->nb + 1, sizeof(uint64_t));
    uint64_t* claim = (uint64_t*)calloc(((size_t)w->nb + 1) * (size_t)K, sizeof(uint64_t));
    int32_t* cand = (int32_t*)malloc(sizeof(int32_t) * (size_t)(M ? M : 1));
    uint64_t* pri = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(M ? M : 1));
    const uint64_t all = K == 64 ? ~0ull : ((1ull << K) - 1ull);
    for (int m = 0; m < M; m++) w->colour[m] = w->man[m].red ? -2 : -1;
    for (int r = 0; r < R; r++) {
        int any = 0;
        memset(claim, 0, sizeof(uint64_t) * ((size_t)w->nb + 1) * (size_t)K);
        for (int m = 0; m < M; m++) {
            cand[m] = -1;
            if (w->colour[m] != -1) continue;
            const Man* mm = &w->man[m];
            const int support = !mm->sa || !mm->sb;
            if (r < RS && !support) continue;
            uint64_t used = (mm->sa ? mask[mm->sa] : 0) | (mm->sb ? mask[mm->sb] : 0);
            uint64_t fr = all & ~used;
            if (!fr) { w->colour[m] = K; continue; }
            int nf = __builtin_popcountll(fr);
            uint32_t kh = key_hash(mm->a, mm->b);
            int lim = nf < spread ? nf : spread;
            int pick = (int)(mix32(kh + (uint32_t)r * 0x165667B1u) % (uint32_t)lim);
            for (int q = 0; q < pick; q++) fr &= fr - 1;
            int c = __builtin_ctzll(fr);
            cand[m] = c;
            pri[m] = ((uint64_t)(RS > 0 && support) << 56) | ((uint64_t)(mix32(kh ^ ((uint32_t)r * 0x27D4EB2Fu)) >> 8) << 32) |
                     (uint64_t)(0xFFFFFFFFu - (uint32_t)m);
            if (mm->sa) { uint64_t* cl = &claim[(size_t)mm->sa * K + c]; if (pri[m] > *cl) *cl = pri[m]; }
            if (mm->sb) { uint64_t* cl = &claim[(size_t)mm->sb * K + c]; if (pri[m] > *cl) *cl = pri[m]; }
            any = 1;
        }
        if (!any) { if (r < RS) { r = RS - 1; continue; } break; }
        for (int m = 0; m < M; m++) {
            if (cand[m] < 0) continue;
            const Man* mm = &w->man[m]; int c = cand[m];
            if (mm->sa && claim[(size_t)mm->sa * K + c] != pri[m]) continue;
            if (mm->sb && claim[(size_t)mm->sb * K + c] != pri[m]) continue;
            w->colour[m] = c;
        }
        for (int m = 0; m < M; m++) {
            if (cand[m] < 0 || w->colour[m] != cand[m]) continue;
            const Man* mm = &w->man[m];
            if (mm->sa) mask[mm->sa] |= 1ull << cand[m];
            if (mm->sb) mask[mm->sb] |= 1ull << cand[m];
        }
    }
    int used = 0, ovf = 0;
    for (int m = 0; m < M; m++) {
        if (w->colour[m] == -1) w->colour[m] = K;
        if (w->colour[m] == K) ovf++; else if (w->colour[m] + 1 > used) used = w->colour[m] + 1;
    }
    w->stats[4] = used; w->stats[5] = ovf;
    free(mask); free(claim); free(cand); free(pri);
}

/* ---------------------------------------------------------------- solver */
/* phys.rs:245-262 */
static void warm_start(ob_world* w) {
    for (int mi = 0; mi < w->nman; mi++) {
        const Man* m = &w->man[mi];
        if (m->red) continue;
        int sa = m->sa, sb = m->sb;
        V3 va = w->sv[sa], wa = w->sw[sa], vb = w->sv[sb], wb = w->sw[sb];
        float pn = 0.0f;
        for (int k = 0; k < m->cnt; k++) {
            const Cp* c = &m->cp[k];
            pn += c->jn; wa = vsub(wa, vmul(c->iaxn, c->jn)); wb = vadd(wb, vmul(c->ibxn, c->jn));
        }
        V3 p = vadd(vadd(vmul(m->n, pn), vmul(m->t1, m->jt1)), vmul(m->t2, m->jt2));
        va = vsub(va, vmul(p, m->ima)); vb = vadd(vb, vmul(p, m->imb));
        wa = vsub(wa, vadd(vadd(vmul(m->ia1, m->jt1), vmul(m->ia2, m->jt2)), vmul(m->ian, m->jtw)));
        wb = vadd(wb, vadd(vadd(vmul(m->ib1, m->jt1), vmul(m->ib2, m->jt2)), vmul(m->ibn, m->jtw)));
        w->sv[sa] = va; w->sw[sa] = wa; w->sv[sb] = vb; w->sw[sb] = wb;
        w->sv[0] = v3(0, 0, 0); w->sw[0] = v3(0, 0, 0);
    }
}

/* Device-order sums: a body's contributions (one per contact, its contacts in manifold order) are added in chunks of
 * 64 by a fixed halving tree (x[i] += x[i + h] for h = 32, 16, ..., 1, empty slots zero), the chunk totals in order.
 * The device forms the same tree across the 64 lanes of a wavefront. */
static V3 tree64(V3* x) {
    for (int h = 32; h >= 1; h >>= 1) for (int i = 0; i < h; i++) x[i] = vadd(x[i], x[i + h]);
    return x[0];
}

/* per-slot contact lists in manifold order (entries m * 2 + side) */
static void build_lists(const ob_world* w, int32_t* off, int32_t* ent) {
    size_t n1 = (size_t)w->nb + 1;
    for (size_t s = 0; s <= n1; s++) off[s] = 0;
    for (int m = 0; m < w->nman; m++) {
        if (w->man[m].red) continue;
        if (w->man[m].sa) off[w->man[m].sa + 1]++;
        if (w->man[m].sb) off[w->man[m].sb + 1]++;
    }
    for (size_t s = 0; s < n1; s++) off[s + 1] += off[s];
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * n1);
    for (size_t s = 0; s < n1; s++) fill[s] = off[s];
    for (int m = 0; m < w->nman; m++) {
        if (w->man[m].red) continue;
        if (w->man[m].sa) ent[fill[w->man[m].sa]++] = 2 * m;
        if (w->man[m].sb) ent[fill[w->man[m].sb]++] = 2 * m + 1;
    }
    free(fill);
}

/* the warm-start contribution of one contact to one side's body: the same vectors phys.rs:245-262 applies, the
 * side's w terms summed in point order then the friction/twist term */
static void warm_contrib(const Man* m, int side, V3* dv, V3* dw) {
    float pn = 0.0f;
    V3 acc = v3(0, 0, 0);
    for (int k = 0; k < m->cnt; k++) {
        const Cp* c = &m->cp[k];
        pn += c->jn;
        V3 t = side ? vmul(c->ibxn, c->jn) : vneg(vmul(c->iaxn, c->jn));
        acc = k == 0 ? t : vadd(acc, t);
    }
    V3 p = vadd(vadd(vmul(m->n, pn), vmul(m->t1, m->jt1)), vmul(m->t2, m->jt2));
    if (side) {
        *dv = vmul(p, m->imb);
        *dw = vadd(acc, vadd(vadd(vmul(m->ib1, m->jt1), vmul(m->ib2, m->jt2)), vmul(m->ibn, m->jtw)));
    } else {
        *dv = vneg(vmul(p, m->ima));
        *dw = vadd(acc, vneg(vadd(vadd(vmul(m->ia1, m->jt1), vmul(m->ia2, m->jt2)), vmul(m->ian, m->jtw))));
    }
}

/* device-order warm start: each moving body adds the tree sum of its contacts' contributions */
static void warm_start_tree(ob_world* w, const int32_t* off, const int32_t* ent) {
    V3 xv[64], xw[64];
    for (int s = 1; s <= w->nb; s++) {
        int o = off[s], n = off[s + 1] - off[s];
        if (!n) continue;
        V3 av = v3(0, 0, 0), aw = v3(0, 0, 0);
        for (int c0 = 0; c0 < n; c0 += 64) {
            for (int i = 0; i < 64; i++) {
                if (c0 + i < n) { int e = ent[o + c0 + i]; warm_contrib(&w->man[e >> 1], e & 1, &xv[i], &xw[i]); }
                else { xv[i] = v3(0, 0, 0); xw[i] = v3(0, 0, 0); }
            }
            av = vadd(av, tree64(xv));
            aw = vadd(aw, tree64(xw));
        }
        w->sv[s] = vadd(w->sv[s], av);
        w->sw[s] = vadd(w->sw[s], aw);
    }
}

typedef struct {
    int use_bias; float inv_h, bias_rate, mass_scale, imp_scale, max_push, anchor_push, anchor_slop;
    float ratio; int shock; float rest_thr; int rest_iters;
} SolveK;

static inline float slop_fn(float e, float slop) { return e > slop ? e - slop : (e < -slop ? e + slop : 0.0f); }

/* Mass-ratio limit (gpu_world.js:1155-1160): across one contact the heavier side responds as if at most R times
 * heavier than the lighter; the factors multiply the two inverse masses (1 when either side is static or R is 0). */
static inline void ratio_scale(float ia, float ib, float R, float* sa, float* sb) {
    *sa = 1.0f; *sb = 1.0f;
    if (R <= 0.0f || ia <= 0.0f || ib <= 0.0f) return;
    if (ia * R < ib) *sa = ib / (R * ia);
    else if (ib * R < ia) *sb = ia / (R * ib);
}
/* The limited side keeps the part of a change d to its velocity x that brings |x| down to its least, and takes the
 * rest divided by sc: the limit only takes energy from the heavy body (gpu_world.js:1166-1177). */
static inline V3 slowing(V3 x, V3 d, float sc) {
    float dd = vdot(d, d);
    if (dd == 0.0f) return d;
    float t = om_min(om_max(-vdot(x, d) / dd, 0.0f), 1.0f);
    return vmul(d, t + (1.0f - t) / sc);
}
/* a slot's shock-propagation height: 255 for the static world (slot 0, or no inverse mass) */
static inline uint32_t slot_height(const ob_world* w, int s) { return s && w->sim[s] > 0.0f ? w->lvl[s] : 255u; }

/* phys.rs:268-323 on velocity copies of the two bodies. fa, fb = 0: the manifold as the reference solves it.
 * fa, fb > 0: mass splitting for the overflow pass: each side's inverse mass and inertia are scaled by the number of
 * overflow manifolds at that body (each manifold acts on its own copy of the body), and the effective masses are
 * recomputed from the scaled ones.
 * The GPU world's extensions (gpu_world.js:1207-1329): with a mass-ratio limit the inverse masses (as split) are
 * scaled by ratio_scale as well, and each limited side keeps its change only as `slowing` allows. In the relax pass
 * with levels (ha, hb: the sides' heights), a contact within 45 degrees of vertical whose lower body sits on fewer
 * levels holds that body where a point is separating: the upper side alone takes the point's impulse. */
static void solve_core(Man* m, const SolveK* K, V3* pva, V3* pwa, V3* pvb, V3* pwb, V3 tha, V3 thb, V3 dpab, float fa, float fb,
                       uint32_t ha, uint32_t hb) {
    V3 va = *pva, wa = *pwa, vb = *pvb, wb = *pwb;
    const V3 va0 = va, wa0 = wa, vb0 = vb, wb0 = wb;
    const int split = fa > 0.0f;
    float sca, scb;
    ratio_scale(split ? m->ima * fa : m->ima, split ? m->imb * fb : m->imb, K->ratio, &sca, &scb);
    const int rec = split || sca != 1.0f || scb != 1.0f;
    const float xa = (split ? fa : 1.0f) * sca, xb = (split ? fb : 1.0f) * scb;
    V3 n = m->n, t1 = m->t1, t2 = m->t2;
    float dn = vdot(dpab, n);
    int hold_a = 0, hold_b = 0;
    if (!K->use_bias && K->shock) { hold_a = n.y > 0.7f && ha > hb; hold_b = n.y < -0.7f && hb > ha; }
    float ima = rec ? m->ima * xa : m->ima, imb = rec ? m->imb * xb : m->imb;
    float tot = 0.0f;
    for (int k = 0; k < m->cnt; k++) {
        Cp* c = &m->cp[k];
        V3 iaxn = rec ? vmul(c->iaxn, xa) : c->iaxn, ibxn = rec ? vmul(c->ibxn, xb) : c->ibxn;
        float nm = c->nm;
        if (rec) { float kk = ima + imb + vdot(c->raxn, iaxn) + vdot(c->rbxn, ibxn); nm = kk > 0.0f ? 1.0f / kk : 0.0f; }
        float s = c->adj + dn + vdot(thb, c->rbxn) - vdot(tha, c->raxn);
        float bias, ms, is;
        if (s > 0.0f) { bias = s * K->inv_h; ms = 1.0f; is = 0.0f; }
        else if (K->use_bias) { bias = om_max(K->bias_rate * s, -K->max_push); ms = K->mass_scale; is = K->imp_scale; }
        else { bias = 0.0f; ms = 1.0f; is = 0.0f; }
        float vn = vdot(vsub(vb, va), n) + vdot(wb, c->rbxn) - vdot(wa, c->raxn);
        const int ga = hold_a && vn > 0.0f, gb = hold_b && vn > 0.0f;  /* this point holds side a (b) */
        if (ga) { float kk = imb + vdot(c->rbxn, ibxn); nm = kk > 0.0f ? 1.0f / kk : 0.0f; }
        if (gb) { float kk = ima + vdot(c->raxn, iaxn); nm = kk > 0.0f ? 1.0f / kk : 0.0f; }
        float imp = -nm * ms * (vn + bias) - is * c->jn;
        float nj = om_max(c->jn + imp, 0.0f);
        imp = nj - c->jn; c->jn = nj; tot += nj;
        if (!ga) { va = vsub(va, vmul(n, imp * ima)); wa = vsub(wa, vmul(iaxn, imp)); }
        if (!gb) { vb = vadd(vb, vmul(n, imp * imb)); wb = vadd(wb, vmul(ibxn, imp)); }
    }
    if (!K->use_bias && tot > 0.0f) m->pushed = 1;
    V3 ia1 = m->ia1, ia2 = m->ia2, ib1 = m->ib1, ib2 = m->ib2, ian = m->ian, ibn = m->ibn;
    float tm1 = m->tm1, tm2 = m->tm2, twm = m->twm;
    if (rec) {
        ia1 = vmul(ia1, xa); ia2 = vmul(ia2, xa); ian = vmul(ian, xa); ib1 = vmul(ib1, xb); ib2 = vmul(ib2, xb); ibn = vmul(ibn, xb);
        float k1 = ima + imb + vdot(m->ca1, ia1) + vdot(m->cb1, ib1);
        float k2 = ima + imb + vdot(m->ca2, ia2) + vdot(m->cb2, ib2);
        tm1 = k1 > 0.0f ? 1.0f / k1 : 0.0f; tm2 = k2 > 0.0f ? 1.0f / k2 : 0.0f;
        float kt = vdot(n, ian) + vdot(n, ibn);
        twm = kt > 0.0f ? 1.0f / kt : 0.0f;
    }
    float lim = m->mu * tot;
    V3 dv = vsub(vb, va);
    float v1 = vdot(dv, t1) + vdot(wb, m->cb1) - vdot(wa, m->ca1);
    float v2 = vdot(dv, t2) + vdot(wb, m->cb2) - vdot(wa, m->ca2);
    int anchored = K->use_bias && m->stick;
    float b1, b2, ms, is;
    if (anchored) {
        float e1 = vdot(m->e0, t1) + vdot(dpab, t1) + vdot(thb, m->cb1) - vdot(tha, m->ca1);
        float e2 = vdot(m->e0, t2) + vdot(dpab, t2) + vdot(thb, m->cb2) - vdot(tha, m->ca2);
        float ap = K->anchor_push, sl = K->anchor_slop;
        b1 = om_min(om_max(K->bias_rate * slop_fn(e1, sl), -ap), ap);
        b2 = om_min(om_max(K->bias_rate * slop_fn(e2, sl), -ap), ap);
        ms = K->mass_scale; is = K->imp_scale;
    } else { b1 = 0.0f; b2 = 0.0f; ms = 1.0f; is = 0.0f; }
    float o1 = m->jt1, o2 = m->jt2;
    float j1 = o1 - tm1 * ms * (v1 + b1) - is * o1;
    float j2 = o2 - tm2 * ms * (v2 + b2) - is * o2;
    float l2 = j1 * j1 + j2 * j2;
    if (l2 > lim * lim) { float kk = lim / om_sqrt(l2); j1 *= kk; j2 *= kk; m->stick = 0; }
    m->jt1 = j1; m->jt2 = j2;
    float d1 = j1 - o1, d2 = j2 - o2;
    V3 p = vadd(vmul(t1, d1), vmul(t2, d2));
    va = vsub(va, vmul(p, ima)); wa = vsub(wa, vadd(vmul(ia1, d1), vmul(ia2, d2)));
    vb = vadd(vb, vmul(p, imb)); wb = vadd(wb, vadd(vmul(ib1, d1), vmul(ib2, d2)));
    float wl = lim * m->rad;
    float ot = m->jtw;
    float jt = om_min(om_max(ot - twm * vdot(vsub(wb, wa), n), -wl), wl);
    m->jtw = jt;
    float dtw = jt - ot;
    wa = vsub(wa, vmul(ian, dtw)); wb = vadd(wb, vmul(ibn, dtw));
    if (sca != 1.0f) { va = vadd(va0, slowing(va0, vsub(va, va0), sca)); wa = vadd(wa0, slowing(wa0, vsub(wa, wa0), sca)); }
    if (scb != 1.0f) { vb = vadd(vb0, slowing(vb0, vsub(vb, vb0), scb)); wb = vadd(wb0, slowing(wb0, vsub(wb, wb0), scb)); }
    *pva = va; *pwa = wa; *pvb = vb; *pwb = wb;
}

/* Restitution after the sub-steps (gpu_world.js:1333-1363): in a manifold with a restitution that pushed during the
 * step, the points that approached faster than the threshold at the step start have their normal velocity driven to
 * -rest times that approach, the points iterated together (a point's bounce moves the others). fa, fb as in
 * solve_core (mass splitting for the overflow), with the mass-ratio limit. */
static void rest_core(Man* m, const SolveK* K, V3* pva, V3* pwa, V3* pvb, V3* pwb, float fa, float fb) {
    if (m->rest <= 0.0f || !m->pushed) return;
    int any = 0;
    for (int k = 0; k < m->cnt; k++) if (!(m->cp[k].rv > -K->rest_thr)) any = 1;
    if (!any) return;
    V3 va = *pva, wa = *pwa, vb = *pvb, wb = *pwb;
    const V3 va0 = va, wa0 = wa, vb0 = vb, wb0 = wb;
    const int split = fa > 0.0f;
    float sca, scb;
    ratio_scale(split ? m->ima * fa : m->ima, split ? m->imb * fb : m->imb, K->ratio, &sca, &scb);
    const int rec = split || sca != 1.0f || scb != 1.0f;
    const float xa = (split ? fa : 1.0f) * sca, xb = (split ? fb : 1.0f) * scb;
    float ima = rec ? m->ima * xa : m->ima, imb = rec ? m->imb * xb : m->imb;
    V3 n = m->n;
    for (int it = 0; it < K->rest_iters; it++) {
        for (int k = 0; k < m->cnt; k++) {
            Cp* c = &m->cp[k];
            if (c->rv > -K->rest_thr) continue;
            V3 iaxn = rec ? vmul(c->iaxn, xa) : c->iaxn, ibxn = rec ? vmul(c->ibxn, xb) : c->ibxn;
            float nm = c->nm;
            if (rec) { float kk = ima + imb + vdot(c->raxn, iaxn) + vdot(c->rbxn, ibxn); nm = kk > 0.0f ? 1.0f / kk : 0.0f; }
            float vn = vdot(vsub(vb, va), n) + vdot(wb, c->rbxn) - vdot(wa, c->raxn);
            float nj = om_max(c->jn - nm * (vn + m->rest * c->rv), 0.0f);
            float imp = nj - c->jn; c->jn = nj;
            va = vsub(va, vmul(n, imp * ima)); wa = vsub(wa, vmul(iaxn, imp));
            vb = vadd(vb, vmul(n, imp * imb)); wb = vadd(wb, vmul(ibxn, imp));
        }
    }
    if (sca != 1.0f) { va = vadd(va0, slowing(va0, vsub(va, va0), sca)); wa = vadd(wa0, slowing(wa0, vsub(wa, wa0), sca)); }
    if (scb != 1.0f) { vb = vadd(vb0, slowing(vb0, vsub(vb, vb0), scb)); wb = vadd(wb0, slowing(wb0, vsub(wb, wb0), scb)); }
    *pva = va; *pwa = wa; *pvb = vb; *pwb = wb;
}

/* one manifold of the sequential (or colour-ordered) solve: the reference's manifold loop body */
static void solve_one(ob_world* w, Man* m, const SolveK* K) {
    if (m->red) return;
    int sa = m->sa, sb = m->sb;
    V3 va = w->sv[sa], wa = w->sw[sa], vb = w->sv[sb], wb = w->sw[sb];
    V3 dpab = vsub(w->sdp[sb], w->sdp[sa]);
    solve_core(m, K, &va, &wa, &vb, &wb, w->sth[sa], w->sth[sb], dpab, 0.0f, 0.0f, slot_height(w, sa), slot_height(w, sb));
    w->sv[sa] = va; w->sw[sa] = wa; w->sv[sb] = vb; w->sw[sb] = wb;
    w->sv[0] = v3(0, 0, 0); w->sw[0] = v3(0, 0, 0);
}
static void rest_one(ob_world* w, Man* m, const SolveK* K) {
    if (m->red) return;
    int sa = m->sa, sb = m->sb;
    V3 va = w->sv[sa], wa = w->sw[sa], vb = w->sv[sb], wb = w->sw[sb];
    rest_core(m, K, &va, &wa, &vb, &wb, 0.0f, 0.0f);
    w->sv[sa] = va; w->sw[sa] = wa; w->sv[sb] = vb; w->sw[sb] = wb;
    w->sv[0] = v3(0, 0, 0); w->sw[0] = v3(0, 0, 0);
}

/* The overflow pass (manifolds the colouring left over): Jacobi with mass splitting. Every overflow manifold solves
 * against its own copies of its bodies, all starting from the velocities before the pass; each body then takes the
 * average of its copies, its copies' changes summed in manifold order. */
static void overflow_pass(ob_world* w, const SolveK* K, const int32_t* ovf, int novf, const int32_t* ocnt, V3* dv, int rest) {
    size_t n1 = (size_t)w->nb + 1;
    V3* v0 = (V3*)malloc(sizeof(V3) * n1);
    V3* w0 = (V3*)malloc(sizeof(V3) * n1);
    memcpy(v0, w->sv, sizeof(V3) * n1); memcpy(w0, w->sw, sizeof(V3) * n1);
    for (int i = 0; i < novf; i++) {
        Man* m = &w->man[ovf[i]];
        int sa = m->sa, sb = m->sb;
        V3 va = v0[sa], wa = w0[sa], vb = v0[sb], wb = w0[sb];
        V3 dpab = vsub(w->sdp[sb], w->sdp[sa]);
        float fa = sa ? (float)ocnt[sa] : 1.0f, fb = sb ? (float)ocnt[sb] : 1.0f;
        if (rest) rest_core(m, K, &va, &wa, &vb, &wb, fa, fb);
        else solve_core(m, K, &va, &wa, &vb, &wb, w->sth[sa], w->sth[sb], dpab, fa, fb, slot_height(w, sa), slot_height(w, sb));
        dv[4 * i + 0] = vsub(va, v0[sa]); dv[4 * i + 1] = vsub(wa, w0[sa]);
        dv[4 * i + 2] = vsub(vb, v0[sb]); dv[4 * i + 3] = vsub(wb, w0[sb]);
    }
    /* each body: the tree sum (over its full contact list, zero for contacts not in the overflow) of its copies'
     * changes, averaged */
    V3 xv[64], xw[64];
    for (size_t s = 1; s < n1; s++) {
        if (!ocnt[s]) continue;
        int o = w->bl_off[s], n = w->bl_off[s + 1] - w->bl_off[s];
        V3 av = v3(0, 0, 0), aw = v3(0, 0, 0);
        for (int c0 = 0; c0 < n; c0 += 64) {
            for (int i = 0; i < 64; i++) {
                xv[i] = v3(0, 0, 0); xw[i] = v3(0, 0, 0);
                if (c0 + i >= n) continue;
                int e = w->bl_ent[o + c0 + i], mi = e >> 1;
                int oi = w->ovf_index[mi];
                if (oi < 0) continue;
                xv[i] = dv[4 * oi + ((e & 1) ? 2 : 0)];
                xw[i] = dv[4 * oi + ((e & 1) ? 3 : 1)];
            }
            av = vadd(av, tree64(xv));
            aw = vadd(aw, tree64(xw));
        }
        float inv = 1.0f / (float)ocnt[s];
        w->sv[s] = vadd(v0[s], vmul(av, inv));
        w->sw[s] = vadd(w0[s], vmul(aw, inv));
    }
    free(v0); free(w0);
}

static void solve_pass(ob_world* w, const SolveK* K, const int32_t* order, int ncol_order, const int32_t* ovf, int novf,
                       const int32_t* ocnt, V3* dv) {
    for (int i = 0; i < ncol_order; i++) solve_one(w, &w->man[order ? order[i] : i], K);
    if (novf) overflow_pass(w, K, ovf, novf, ocnt, dv, 0);
}
/* the restitution pass: the colours in order, then the overflow (Jacobi with mass splitting, as the solve's) */
static void rest_pass(ob_world* w, const SolveK* K, const int32_t* order, int ncol_order, const int32_t* ovf, int novf,
                      const int32_t* ocnt, V3* dv) {
    for (int i = 0; i < ncol_order; i++) rest_one(w, &w->man[order ? order[i] : i], K);
    if (novf) overflow_pass(w, K, ovf, novf, ocnt, dv, 1);
}

/* Shock-propagation levels (gpu_world.js:1178-1203): a body's height is 255 minus its number of near-vertical
 * contacts above the static world (static and sleeping bodies 255, a body on them 254, ...; 0: no chain reached it).
 * Each pass reads one copy and raises the other with maxima over the manifolds (carrying both bodies' read heights,
 * and lifting the upper body of a contact within 45 degrees of vertical to one below its support's), so the result
 * does not depend on the manifold order. The count of passes is even, so the heights end in lvl. */
static void level_passes(ob_world* w, int passes) {
    size_t n1 = (size_t)w->nb + 1;
    for (size_t i = 0; i < n1; i++) { w->lvl[i] = 0; w->lvl2[i] = 0; }
    for (int k = 0; k < passes; k++) {
        uint32_t* rd = (k & 1) ? w->lvl2 : w->lvl;
        uint32_t* wr = (k & 1) ? w->lvl : w->lvl2;
        for (int mi = 0; mi < w->nman; mi++) {
            const Man* m = &w->man[mi];
            if (m->red) continue;
            int sa = m->sa, sb = m->sb;
            int da = sa && w->sim[sa] > 0.0f, db = sb && w->sim[sb] > 0.0f;
            uint32_t ha = da ? rd[sa] : 255u, hb = db ? rd[sb] : 255u;
            if (da && rd[sa] > wr[sa]) wr[sa] = rd[sa];
            if (db && rd[sb] > wr[sb]) wr[sb] = rd[sb];
            if (m->n.y > 0.7f && ha > 1u && db && ha - 1u > wr[sb]) wr[sb] = ha - 1u;
            if (m->n.y < -0.7f && hb > 1u && da && hb - 1u > wr[sa]) wr[sa] = hb - 1u;
        }
    }
}

/* phys.rs:326-354, plus island labels = minimum body index of each island */
static int uf_find(int32_t* uf, int x) { while (uf[x] != x) { int p = uf[uf[x]]; uf[x] = p; x = p; } return x; }
static void sleep_islands(ob_world* w, float dt) {
    int n = w->nb;
    int32_t* uf = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; i++) uf[i] = i;
    for (int mi = 0; mi < w->nman; mi++) {
        const Man* m = &w->man[mi];
        if (m->sa == 0 || m->sb == 0 || m->red) continue;
        int ra = uf_find(uf, m->ba), rb = uf_find(uf, m->bb);
        if (ra != rb) uf[ra] = rb;
    }
    for (int i = 0; i < n; i++) w->island[i] = n;
    for (int i = 0; i < n; i++) { int r = uf_find(uf, i); if (i < w->island[r]) w->island[r] = i; }
    for (int i = 0; i < n; i++) { int r = uf_find(uf, i); w->island[i] = w->island[r]; }
    /* (the root's slot is rewritten by its own turn in the loop above; islands[r] then already holds the minimum) */
    float l2 = w->prm[P_SLEEP_LIN] * w->prm[P_SLEEP_LIN], a2 = w->prm[P_SLEEP_ANG] * w->prm[P_SLEEP_ANG];
    for (int i = 0; i < n; i++) {
        Body* b = &w->bd[i];
        if (!b->awake) continue;
        Quat dq = qmul(b->q, qconj(b->sleep_q));
        float ang = 2.0f * om_sqrt(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z);
        int near = vlen2(vsub(b->p, b->sleep_p)) < w->prm[P_SLEEP_DIST] * w->prm[P_SLEEP_DIST] && ang < w->prm[P_SLEEP_ROT];
        if (near && vlen2(b->v) < 100.0f * l2 && vlen2(b->w) < 100.0f * a2) b->sleep_t += dt;
        else { b->sleep_t = 0.0f; b->sleep_p = b->p; b->sleep_q = b->q; }
    }
    if (w->prm[P_ALLOW_SLEEP] > 0.5f) {
        float* min_t = (float*)malloc(sizeof(float) * (size_t)(n ? n : 1));
        for (int i = 0; i < n; i++) min_t[i] = 3.40282347e38f;
        for (int i = 0; i < n; i++) if (w->bd[i].awake) { int r = uf_find(uf, i); min_t[r] = om_min(min_t[r], w->bd[i].sleep_t); }
        for (int i = 0; i < n; i++) {
            if (!w->bd[i].awake) continue;
            int r = uf_find(uf, i);
            if (min_t[r] >= w->prm[P_SLEEP_TIME]) { Body* b = &w->bd[i]; b->awake = 0; b->v = v3(0, 0, 0); b->w = v3(0, 0, 0); }
        }
        free(min_t);
    }
    free(uf);
}

static const Man* g_sort_base;
static int cmp_old(const void* x, const void* y) {
    const Man *a = &g_sort_base[*(const int32_t*)x], *b = &g_sort_base[*(const int32_t*)y];
    uint64_t ka = mkey64(a->a, a->b), kb = mkey64(b->a, b->b);
    return ka < kb ? -1 : (ka > kb ? 1 : 0);
}

static void particle_stage(ob_world* w, float dt);

/* phys.rs:152-243 */
void ob_step(ob_world* w, float dt) {
    const float* P = w->prm;
    const int coloured = w->order == OB_ORDER_COLOURED;
    int n = w->nb;
    /* ---- prepare bodies, AABBs, solver slots for awake ones */
    prep_shapes(w, dt);
    w->sv[0] = v3(0, 0, 0); w->sw[0] = v3(0, 0, 0); w->sdp[0] = v3(0, 0, 0); w->sth[0] = v3(0, 0, 0);
    w->sim[0] = 0.0f; w->sii[0] = m3_diag(v3(0, 0, 0)); w->sq[0] = Q_ID; w->sp0[0] = v3(0, 0, 0);
    int nsolver = 0;
    for (int i = 0; i < n; i++) {
        Body* b = &w->bd[i];
        if (b->awake) {
            int s = i + 1;
            w->sidx[i] = s; nsolver++;
            w->sv[s] = b->v; w->sw[s] = b->w; w->sdp[s] = v3(0, 0, 0); w->sth[s] = v3(0, 0, 0);
            w->sim[s] = b->inv_m; w->sii[s] = m3_sandwich(&b->rot, b->inv_i); w->sq[s] = b->q; w->sp0[s] = b->p;
        } else w->sidx[i] = 0;
    }
    /* ---- broadphase + narrowphase */
    Man* t = w->old; w->old = w->man; w->man = t;
    int tc = w->capold; w->capold = w->capman; w->capman = tc;
    w->nold = w->nman; w->nman = 0;
    w->old_sorted = (int32_t*)realloc(w->old_sorted, sizeof(int32_t) * (size_t)(w->nold ? w->nold : 1));
    for (int i = 0; i < w->nold; i++) w->old_sorted[i] = i;
    g_sort_base = w->old;
    qsort(w->old_sorted, (size_t)w->nold, sizeof(int32_t), cmp_old);

    int np = 0;
    int32_t* pairs = hashed_pairs(w, &np);
    w->stats[1] = np;
    int ngc = 0;
    uint8_t* snap = NULL;
    if (coloured) {
        snap = (uint8_t*)malloc((size_t)(n ? n : 1));
        for (int i = 0; i < n; i++) { snap[i] = (uint8_t)((w->bd[i].awake ? 1 : 0) | (w->bd[i].sleep_t != 0.0f ? 2 : 0)); w->wake_flag[i] = 0; }
        /* canonical candidates: per shape, the ground first, then partners by index */
        int nc = np;
        for (int s = 0; s < w->ns; s++) if (w->bd[w->sh[s].body].awake && w->sh[s].lo.y < 0.0f) nc++;
        if (w->capcand < nc) { w->capcand = nc; w->cand = (int32_t*)realloc(w->cand, sizeof(int32_t) * 2 * (size_t)nc); }
        int k = 0;
        for (int s = 0; s < w->ns; s++) if (w->bd[w->sh[s].body].awake && w->sh[s].lo.y < 0.0f) { w->cand[2 * k] = s; w->cand[2 * k + 1] = -1; k++; ngc++; }
        memcpy(w->cand + 2 * k, pairs, sizeof(int32_t) * 2 * (size_t)np);
        qsort(w->cand, (size_t)nc, 2 * sizeof(int32_t), cmp_pair);
        w->ncand = nc;
        for (int c = 0; c < nc; c++) {
            if (w->cand[2 * c + 1] < 0) collide_ground(w, w->cand[2 * c]);
            else collide_pair(w, w->cand[2 * c], w->cand[2 * c + 1], snap);
        }
        int nw = 0;
        for (int i = 0; i < n; i++) if (w->wake_flag[i]) { if (!w->bd[i].awake && w->bd[i].inv_m != 0.0f) nw++; wake_body(w, i); }
        w->stats[7] = nw;
        free(snap);
    } else {
        int nc = np;
        for (int s = 0; s < w->ns; s++) if (w->bd[w->sh[s].body].awake && w->sh[s].lo.y < 0.0f) nc++;
        if (w->capcand < nc) { w->capcand = nc; w->cand = (int32_t*)realloc(w->cand, sizeof(int32_t) * 2 * (size_t)nc); }
        int k = 0;
        for (int s = 0; s < w->ns; s++) if (w->bd[w->sh[s].body].awake && w->sh[s].lo.y < 0.0f) { w->cand[2 * k] = s; w->cand[2 * k + 1] = -1; k++; ngc++; collide_ground(w, s); }
        for (int p = 0; p < np; p++) { w->cand[2 * k] = pairs[2 * p]; w->cand[2 * k + 1] = pairs[2 * p + 1]; k++; collide_pair(w, pairs[2 * p], pairs[2 * p + 1], NULL); }
        w->ncand = nc;
        w->stats[7] = 0;
    }
    w->stats[6] = ngc;
    free(pairs);
    w->nred = reduce_pairs(w);
    /* ---- constraint masses (phys.rs:180-211) */
    int sub = (int)P[P_SUBSTEPS]; if (sub < 1) sub = 1;
    float h = dt / (float)sub;
    float inv_h = 1.0f / h;
    float hz = P[P_HERTZ] > 0.0f ? P[P_HERTZ] : inv_h / 6.0f;
    float omega = 2.0f * 3.14159265358979323846f * hz;
    float a1 = 2.0f * P[P_DAMPING] + h * omega;
    float a2 = h * omega * a1;
    float a3 = 1.0f / (1.0f + a2);
    float bias_rate = omega / a1, mass_scale = a2 * a3, imp_scale = a3 * P[P_LEAK];
    int pts = 0;
    for (int mi = 0; mi < w->nman; mi++) {
        Man* m = &w->man[mi];
        float ima = w->sim[m->sa]; M3 iia = w->sii[m->sa];
        float imb = w->sim[m->sb]; M3 iib = w->sii[m->sb];
        m->ima = ima; m->imb = imb;
        V3 nn = m->n;
        for (int k = 0; k < m->cnt; k++) {
            Cp* c = &m->cp[k];
            c->raxn = vcross(c->ra, nn); c->rbxn = vcross(c->rb, nn);
            c->iaxn = m3_mulv(&iia, c->raxn); c->ibxn = m3_mulv(&iib, c->rbxn);
            float kk = ima + imb + vdot(c->raxn, c->iaxn) + vdot(c->rbxn, c->ibxn);
            c->nm = kk > 0.0f ? 1.0f / kk : 0.0f;
        }
        m->ca1 = vcross(m->fa, m->t1); m->ca2 = vcross(m->fa, m->t2); m->cb1 = vcross(m->fb, m->t1); m->cb2 = vcross(m->fb, m->t2);
        m->ia1 = m3_mulv(&iia, m->ca1); m->ia2 = m3_mulv(&iia, m->ca2); m->ib1 = m3_mulv(&iib, m->cb1); m->ib2 = m3_mulv(&iib, m->cb2);
        float k1 = ima + imb + vdot(m->ca1, m->ia1) + vdot(m->cb1, m->ib1);
        float k2 = ima + imb + vdot(m->ca2, m->ia2) + vdot(m->cb2, m->ib2);
        m->tm1 = k1 > 0.0f ? 1.0f / k1 : 0.0f; m->tm2 = k2 > 0.0f ? 1.0f / k2 : 0.0f;
        m->ian = m3_mulv(&iia, nn); m->ibn = m3_mulv(&iib, nn);
        float kt = vdot(nn, m->ian) + vdot(nn, m->ibn);
        m->twm = kt > 0.0f ? 1.0f / kt : 0.0f;
        pts += m->cnt;
    }
    w->stats[2] = w->nman; w->stats[3] = pts;
    /* ---- solve order */
    int32_t* order = NULL;
    int ncol_order = w->nman, novf = 0;
    int32_t *ovf = NULL, *ocnt = NULL;
    V3* dv = NULL;
    if (coloured) {
        colour_manifolds(w);
        int K = (int)P[P_COL_LIMIT]; if (K < 1) K = 1; if (K > 64) K = 64;
        order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w->nman ? w->nman : 1));
        ovf = (int32_t*)malloc(sizeof(int32_t) * (size_t)(w->nman ? w->nman : 1));
        ocnt = (int32_t*)calloc((size_t)w->nb + 1, sizeof(int32_t));
        int k = 0;
        for (int c = 0; c < K; c++) for (int mi = 0; mi < w->nman; mi++) if (w->colour[mi] == c) order[k++] = mi;
        ncol_order = k;
        for (int mi = 0; mi < w->nman; mi++) if (w->colour[mi] == K) {
            ovf[novf++] = mi;
            if (w->man[mi].sa) ocnt[w->man[mi].sa]++;
            if (w->man[mi].sb) ocnt[w->man[mi].sb]++;
        }
        dv = (V3*)malloc(sizeof(V3) * 4 * (size_t)(novf ? novf : 1));
        w->bl_off = (int32_t*)realloc(w->bl_off, sizeof(int32_t) * ((size_t)w->nb + 2));
        w->bl_ent = (int32_t*)realloc(w->bl_ent, sizeof(int32_t) * 2 * (size_t)(w->nman ? w->nman : 1));
        w->ovf_index = (int32_t*)realloc(w->ovf_index, sizeof(int32_t) * (size_t)(w->nman ? w->nman : 1));
        build_lists(w, w->bl_off, w->bl_ent);
        for (int mi = 0; mi < w->nman; mi++) w->ovf_index[mi] = -1;
        for (int i = 0; i < novf; i++) w->ovf_index[ovf[i]] = i;
    } else { w->stats[4] = 0; w->stats[5] = 0; }
    /* ---- sub-steps (phys.rs:214-230) */
    V3 g = v3(P[P_GX], P[P_GY], P[P_GZ]);
    float wmax = 0.25f * 3.14159265358979323846f * inv_h;
    /* (level passes: an even count, at most 64; restitution iterations at most 64) */
    int shock = (int)P[P_SHOCK]; if (shock < 0) shock = 0; if (shock > 64) shock = 64; shock = (shock + 1) & ~1;
    int riters = (int)P[P_REST_ITERS]; if (riters < 0) riters = 0; if (riters > 64) riters = 64;
    SolveK kb = {1, inv_h, bias_rate, mass_scale, imp_scale, P[P_MAXPUSH], P[P_ANCHOR_PUSH], P[P_ANCHOR_SLOP],
                 P[P_RATIO] > 0.0f ? P[P_RATIO] : 0.0f, shock > 0, P[P_REST_THR], riters};
    SolveK kr = kb; kr.use_bias = 0;
    if (shock) level_passes(w, shock);
    for (int s = 0; s < sub; s++) {
        for (int i = 0; i < n; i++) {
            int si = w->sidx[i]; if (!si) continue;
            w->sv[si] = vadd(w->sv[si], vmul(g, h));
            V3 ww = w->sw[si]; float wl = vlen2(ww);
            if (wl > wmax * wmax) w->sw[si] = vmul(ww, wmax / om_sqrt(wl));
        }
        if (coloured) warm_start_tree(w, w->bl_off, w->bl_ent); else warm_start(w);
        solve_pass(w, &kb, order, ncol_order, ovf, novf, ocnt, dv);
        for (int i = 0; i < n; i++) {
            int si = w->sidx[i]; if (!si) continue;
            V3 v = w->sv[si], ww = w->sw[si];
            w->sdp[si] = vadd(w->sdp[si], vmul(v, h)); w->sth[si] = vadd(w->sth[si], vmul(ww, h));
            w->sq[si] = qintegrate(w->sq[si], ww, h);
        }
        solve_pass(w, &kr, order, ncol_order, ovf, novf, ocnt, dv);
    }
    /* (the pass is skipped when no manifold may bounce: a restitution and a point approaching past the threshold) */
    int bounce = 0;
    if (w->any_rest)
        for (int mi = 0; mi < w->nman && !bounce; mi++) {
            const Man* m = &w->man[mi];
            if (m->rest > 0.0f && !m->red) for (int k = 0; k < m->cnt; k++) if (!(m->cp[k].rv > -kr.rest_thr)) bounce = 1;
        }
    if (bounce) rest_pass(w, &kr, order, ncol_order, ovf, novf, ocnt, dv);
    free(order); free(ovf); free(ocnt); free(dv);
    /* ---- write back */
    for (int i = 0; i < n; i++) {
        int si = w->sidx[i]; if (!si) continue;
        Body* b = &w->bd[i];
        b->v = w->sv[si]; b->w = w->sw[si]; b->p = vadd(w->sp0[si], w->sdp[si]); b->q = w->sq[si];
    }
    sleep_islands(w, dt);
    w->stats[0] = nsolver;
    particle_stage(w, dt);
}

int ob_candidates(const ob_world* w, int32_t* out, int cap) {
    for (int i = 0; i < w->ncand && i < cap; i++) { out[2 * i] = w->cand[2 * i]; out[2 * i + 1] = w->cand[2 * i + 1]; }
    return w->ncand;
}

int ob_manifolds(const ob_world* w, int32_t* keys, float* data, int cap) {
    for (int i = 0; i < w->nman && i < cap; i++) {
        const Man* m = &w->man[i];
        keys[3 * i] = m->a; keys[3 * i + 1] = m->b; keys[3 * i + 2] = m->cnt;
        float* f = data + (size_t)i * OB_MAN_F;
        memset(f, 0, sizeof(float) * OB_MAN_F);
        f[0] = m->n.x; f[1] = m->n.y; f[2] = m->n.z; f[3] = (float)m->cnt;
        for (int k = 0; k < m->cnt; k++) {
            const Cp* c = &m->cp[k]; float* q = f + 4 + 8 * k;
            q[0] = c->raxn.x; q[1] = c->raxn.y; q[2] = c->raxn.z; q[3] = c->rbxn.x; q[4] = c->rbxn.y; q[5] = c->rbxn.z;
            q[6] = c->adj; q[7] = c->jn;
        }
        f[36] = m->jt1; f[37] = m->jt2; f[38] = m->jtw; f[39] = (float)m->stick;
    }
    return w->nman;
}

int ob_colours(const ob_world* w, int32_t* out, int cap) {
    if (w->order != OB_ORDER_COLOURED) return 0;
    for (int i = 0; i < w->nman && i < cap; i++) out[i] = w->colour[i];
    return w->nman;
}

void ob_islands(const ob_world* w, int32_t* out) { for (int i = 0; i < w->nb; i++) out[i] = w->island[i]; }
void ob_levels(const ob_world* w, int32_t* out) {
    for (int i = 0; i < w->nb; i++) { int s = w->sidx[i]; out[i] = (int32_t)slot_height(w, s); }
}

/* ---------------------------------------------------------------- particle debris (gpu_world.js:1671-2038) */
/* Position-based debris: each sub-step predicts, pushes each moving particle out of the other particles (Jacobi:
 * the pushes averaged and over-relaxed, against the predicted positions of all), then out of the ground and every
 * candidate rigid shape in turn, each in full, with Coulomb friction from the push; the velocity is the move over the
 * sub-step. Particles never push bodies. A particle at rest for 0.5 s sleeps; a blast, a wake request or a moving shape
 * near it wakes it. Candidates are gathered once per step: other particles from a hashed grid (each slot's particles
 * in index order, the 27 cells round the particle's in z-y-x order, the first 24 in reach), shapes from the box
 * step's cells (the 12 lowest indices whose step-start box meets the particle's reach). Shapes are seen at their
 * step-start pose, their bodies' velocities at the end of the step. */
static float op_pcell(const ob_world* w) { return w->prm[P_PCELL] > 0.1f ? w->prm[P_PCELL] : 0.7f; }
static uint32_t op_grid_size(int np) { uint32_t g = 256; while (g < 2u * (uint32_t)np) g <<= 1; return g; }
static uint32_t op_hash(int x, int y, int z, uint32_t mask) {
    return (((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u)) & mask;
}
static int op_cell(float v, float pc) { return (int32_t)om_floor(v / pc); }

int ob_particles_capacity(ob_world* w, int cap) {
    if (cap < w->np) return -1;
    w->pt = (Particle*)realloc(w->pt, sizeof(Particle) * (size_t)(cap > 0 ? cap : 1));
    w->pfree = (int32_t*)realloc(w->pfree, sizeof(int32_t) * (size_t)(cap > 0 ? cap : 1));
    w->capp = cap;
    return 0;
}
/* records of 8 floats: p[3] v[3] radius friction; the slot of each (-1: full, or not finite); returns the count added */
int ob_add_particles(ob_world* w, int n, const float* rec, int32_t* slots) {
    int added = 0;
    float rmax = op_pcell(w) * 0.5f - 0.05f;
    for (int k = 0; k < n; k++) {
        const float* f = rec + (size_t)8 * k;
        int fin = 1;
        for (int a = 0; a < 8; a++) if (!isfinite(f[a])) fin = 0;
        int i = -1;
        if (fin) { if (w->nfree) i = w->pfree[--w->nfree]; else if (w->np < w->capp) i = w->np++; }
        slots[k] = i;
        if (i < 0) continue;
        Particle* p = &w->pt[i];
        memset(p, 0, sizeof(Particle));
        p->x = v3(f[0], f[1], f[2]); p->v = v3(f[3], f[4], f[5]); p->r = om_min(f[6], rmax); p->mu = f[7];
        p->st = 1.0f; p->q = quat(0, 0, 0, 1); p->xs = p->x; p->anchor = p->x; p->xc = p->x;
        added++;
    }
    return added;
}
void ob_kill_particle(ob_world* w, int i) {
    if (i < 0 || i >= w->np || w->pt[i].st == 0.0f) return;
    w->pt[i].st = 0.0f;
    w->pfree[w->nfree++] = i;
}
void ob_blast_particles(ob_world* w, float cx, float cy, float cz, float reach, float speed) {
    if (w->nblast >= 16) return;
    float* b = w->blast[w->nblast++];
    b[0] = cx; b[1] = cy; b[2] = cz; b[3] = reach; b[4] = speed;
}
void ob_wake_particles(ob_world* w) { w->pwake = 1; }
/* 12 floats per slot: x y z vx, vy vz qx qy, qz qw state radius (gpu_world.js:2891-2896) */
int ob_read_particles(const ob_world* w, float* out) {
    for (int i = 0; i < w->np; i++) {
        const Particle* p = &w->pt[i]; float* o = out + (size_t)12 * i;
        o[0] = p->x.x; o[1] = p->x.y; o[2] = p->x.z; o[3] = p->v.x; o[4] = p->v.y; o[5] = p->v.z;
        o[6] = p->q.x; o[7] = p->q.y; o[8] = p->q.z; o[9] = p->q.w; o[10] = p->st; o[11] = p->r;
    }
    return w->np;
}
/* [slots in use (high-water), moving at the stage's start, asleep at its end, free slots] */
void ob_particle_stats(const ob_world* w, int32_t* out4) { for (int k = 0; k < 4; k++) out4[k] = w->pstats[k]; }

/* whether shape s is in the box step's cell grid: at most 8 cells along each axis of its step-start box, finite */
static int op_in_grid(const ob_world* w, int s) {
    const float ic = 1.0f / w->prm[P_CELL];
    const Shape* sh = &w->sh[s];
    float fl[3] = {sh->lo.x * ic, sh->lo.y * ic, sh->lo.z * ic}, fh[3] = {sh->hi.x * ic, sh->hi.y * ic, sh->hi.z * ic};
    for (int a = 0; a < 3; a++) {
        if (!(om_fabs(fl[a]) < 1.0e8f) || !(om_fabs(fh[a]) < 1.0e8f) || !(fh[a] - fl[a] < 8.0f)) return 0;
        int n = (int32_t)om_floor(fh[a]) - (int32_t)om_floor(fl[a]) + 1;
        if (n < 1 || n > 8) return 0;
    }
    return 1;
}
/* a moving shape (awake body with mass) whose step-start box meets [lo, hi] */
static int op_near_moving(const ob_world* w, V3 lo, V3 hi) {
    for (int s = 0; s < w->ns; s++) {
        const Shape* sh = &w->sh[s]; const Body* b = &w->bd[sh->body];
        if (!b->awake || b->inv_m == 0.0f) continue;
        if (sh->lo.x > hi.x || lo.x > sh->hi.x || sh->lo.y > hi.y || lo.y > sh->hi.y || sh->lo.z > hi.z || lo.z > sh->hi.z) continue;
        if (!op_in_grid(w, s)) continue;
        return 1;
    }
    return 0;
}

/* friction: take back the sliding of this sub-step, up to mu times the push (gpu_world.js:1765-1769) */
static V3 op_fric(V3 disp, V3 n, float pen, float mu) {
    V3 t = vsub(disp, vmul(n, vdot(disp, n)));
    float tl = vlen(t);
    if (tl < 1e-9f) return v3(0, 0, 0);
    return vmul(t, -om_min(mu * pen / tl, 1.0f));
}
static V3 op_cap(V3 v) { float l = vlen(v); return l > 40.0f ? vmul(v, 40.0f / l) : v; }

static void particle_stage(ob_world* w, float dt) {
    const float* P = w->prm;
    int np = w->np;
    w->pstats[0] = np; w->pstats[1] = 0; w->pstats[2] = 0; w->pstats[3] = w->nfree;
    int nbl = w->nblast < 3 ? w->nblast : 3;
    if (np == 0) { w->nblast -= nbl; memmove(w->blast, w->blast + nbl, sizeof(w->blast[0]) * (size_t)w->nblast); w->pwake = 0; return; }
    int psub = (int)P[P_PSUB]; if (psub < 1) psub = 1; if (psub > 16) psub = 16;
    const float h = dt / (float)psub, pc = op_pcell(w);
    const V3 g = v3(P[P_GX], P[P_GY], P[P_GZ]);
    const float mug = P[P_FRICTION];
    const uint32_t G = op_grid_size(np), mask = G - 1u;
    /* ---- blasts, wakes, grid (gpu_world.js:1696-1713) */
    uint32_t* cnt = (uint32_t*)calloc(G + 1, sizeof(uint32_t));
    int moving = 0;
    for (int i = 0; i < np; i++) {
        Particle* p = &w->pt[i];
        if (p->st == 0.0f || p->st >= 3.0f) continue;
        V3 v = p->v; int hit = 0;
        for (int b = 0; b < nbl; b++) {
            const float* bl = w->blast[b];
            V3 d = vsub(p->x, v3(bl[0], bl[1], bl[2])); float l = vlen(d);
            if (l < bl[3]) { hit = 1; v = vadd(v, vmul(vadd(vmul(d, 1.0f / om_max(l, 1e-3f)), v3(0.0f, 0.3f, 0.0f)), bl[4] * (1.0f - l / bl[3]))); }
        }
        if (p->st == 2.0f && (hit || w->pwake || op_near_moving(w, vsub(p->x, v3(0.5f, 0.5f, 0.5f)), vadd(p->x, v3(0.5f, 0.5f, 0.5f))))) { p->st = 1.0f; p->t = 0.0f; }
        if (p->st == 1.0f) { moving++; p->v = v; }
        p->cell = op_hash(op_cell(p->x.x, pc), op_cell(p->x.y, pc), op_cell(p->x.z, pc), mask);
        cnt[p->cell + 1]++;
    }
    w->nblast -= nbl; memmove(w->blast, w->blast + nbl, sizeof(w->blast[0]) * (size_t)w->nblast); w->pwake = 0;
    w->pstats[1] = moving;
    for (uint32_t k = 0; k < G; k++) cnt[k + 1] += cnt[k];
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
    uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * G);
    memcpy(fill, cnt, sizeof(uint32_t) * G);
    for (int i = 0; i < np; i++) {  /* index order: each slot's list ascending */
        const Particle* p = &w->pt[i];
        if (p->st == 0.0f || p->st >= 3.0f) continue;
        list[fill[p->cell]++] = i;
    }
    if (moving) {
        /* ---- candidates (gpu_world.js:1774-1816) */
        for (int i = 0; i < np; i++) {
            Particle* p = &w->pt[i];
            if (p->st != 1.0f) continue;
            const V3 xi = p->x; const float ri = p->r;
            const float m = vlen(p->v) * dt + 0.02f;
            int cx = op_cell(xi.x, pc), cy = op_cell(xi.y, pc), cz = op_cell(xi.z, pc);
            uint32_t seen[27]; int ns = 0; int nn = 0;
            for (int z = cz - 1; z <= cz + 1; z++) for (int y = cy - 1; y <= cy + 1; y++) for (int x = cx - 1; x <= cx + 1; x++) {
                uint32_t hc = op_hash(x, y, z, mask);
                int dup = 0; for (int q = 0; q < ns; q++) if (seen[q] == hc) dup = 1;
                if (dup) continue;
                seen[ns++] = hc;
                for (uint32_t e = cnt[hc]; e < cnt[hc + 1]; e++) {
                    int j = list[e]; if (j == i) continue;
                    const Particle* q = &w->pt[j];
                    float rr = ri + q->r + m + vlen(q->v) * dt + 0.01f;
                    V3 d = vsub(xi, q->x);
                    if (vdot(d, d) < rr * rr && nn < OP_NN) p->nb[nn++] = j;
                }
            }
            p->nn = nn;
            /* shapes: the 12 lowest indices whose step-start box meets the particle's reach */
            float e = ri + m + 0.01f;
            int nsh = 0;
            for (int sidx = 0; sidx < w->ns && nsh < OP_NS; sidx++) {
                const Shape* sh = &w->sh[sidx];
                if (xi.x + e < sh->lo.x || xi.y + e < sh->lo.y || xi.z + e < sh->lo.z ||
                    xi.x - e > sh->hi.x || xi.y - e > sh->hi.y || xi.z - e > sh->hi.z) continue;
                if (!op_in_grid(w, sidx)) continue;
                p->sh[nsh++] = sidx;
            }
            p->ns = nsh;
        }
        /* ---- sub-steps */
        const float cap = 2.0f * h;
        for (int sub = 0; sub < psub; sub++) {
            for (int i = 0; i < np; i++) {  /* predict (:1755-1763) */
                Particle* p = &w->pt[i];
                if (p->st != 1.0f) continue;
                p->v = op_cap(vadd(p->v, vmul(g, h)));
                p->xs = p->x;
                p->x = vadd(p->x, vmul(p->v, h));
            }
            for (int i = 0; i < np; i++) {  /* push out (:1818-1924) */
                Particle* p = &w->pt[i];
                if (p->st != 1.0f) continue;
                const V3 xi = p->x, xs = p->xs; const float ri = p->r, mui = p->mu;
                V3 corr = v3(0, 0, 0); float n = 0.0f, touch = 0.0f; int bearing = 0, grounded = 0;
                for (int q = 0; q < p->nn; q++) {
                    const Particle* o = &w->pt[p->nb[q]];
                    V3 d = vsub(xi, o->x); float d2 = vdot(d, d), rr = ri + o->r;
                    if (d2 >= (rr + 0.01f) * (rr + 0.01f) || d2 < 1e-12f) continue;
                    float dist = om_sqrt(d2); V3 nn = vmul(d, 1.0f / dist);
                    if (o->g_last && nn.y > 0.5f) grounded = 1;
                    touch = 1.0f;
                    if (dist >= rr) continue;
                    const int om = o->st == 1.0f;
                    const float wgt = om ? 0.5f : 1.0f;
                    const V3 dj = om ? vsub(o->x, o->xs) : v3(0, 0, 0);
                    const V3 rel = vsub(vsub(xi, xs), dj);
                    float pen = om_min((rr - dist) * wgt, (om_max(-vdot(rel, nn), 0.0f) + cap) * wgt);
                    corr = vadd(corr, vadd(vmul(nn, pen), vmul(op_fric(rel, nn, pen, mui), wgt)));
                    n += 1.0f;
                }
                V3 x = vadd(xi, vmul(corr, n > 1.5f ? 1.5f / n : 1.0f));
                float pg0 = ri - x.y;
                if (pg0 > -0.01f) { touch = 1.0f; grounded = 1; }
                if (pg0 > 0.0f) {
                    V3 nn = v3(0.0f, 1.0f, 0.0f);
                    x = vadd(x, vadd(vmul(nn, pg0), op_fric(vsub(x, xs), nn, pg0, 0.5f * (mui + mug))));
                }
                for (int q = 0; q < p->ns; q++) {
                    const Shape* sh = &w->sh[p->sh[q]];
                    const Body* bd = &w->bd[sh->body];
                    V3 dl = vsub(x, sh->wp);
                    V3 lc = m3_tmulv(&sh->wrot, dl);
                    V3 hh = sh->h;
                    V3 cl = v3(om_min(om_max(lc.x, -hh.x), hh.x), om_min(om_max(lc.y, -hh.y), hh.y), om_min(om_max(lc.z, -hh.z), hh.z));
                    V3 df = vsub(lc, cl); float d2 = vdot(df, df);
                    V3 nl; float pen;
                    if (d2 > 1e-12f) {
                        float dist = om_sqrt(d2);
                        if (dist >= ri + 0.01f) continue;
                        touch = 1.0f;
                        if (dist >= ri) continue;
                        nl = vmul(df, 1.0f / dist); pen = ri - dist;
                    } else {
                        V3 dd = v3(hh.x - om_fabs(lc.x), hh.y - om_fabs(lc.y), hh.z - om_fabs(lc.z));
                        if (dd.x <= dd.y && dd.x <= dd.z) { nl = v3(lc.x >= 0.0f ? 1.0f : -1.0f, 0.0f, 0.0f); pen = dd.x + ri; }
                        else if (dd.y <= dd.z) { nl = v3(0.0f, lc.y >= 0.0f ? 1.0f : -1.0f, 0.0f); pen = dd.y + ri; }
                        else { nl = v3(0.0f, 0.0f, lc.z >= 0.0f ? 1.0f : -1.0f); pen = dd.z + ri; }
                    }
                    V3 nn = m3_mulv(&sh->wrot, nl);
                    V3 vb = vadd(bd->v, vcross(bd->w, vsub(x, bd->p)));
                    V3 vbh = vmul(vb, h);
                    float push = om_min(pen, om_max(vdot(vsub(vbh, vsub(x, xs)), nn), 0.0f) + cap);
                    float mu = 0.5f * (mui + sh->mu);
                    x = vadd(x, vadd(vmul(nn, push), op_fric(vsub(vsub(x, xs), vbh), nn, push, mu)));
                    touch = 1.0f;
                    const int fixed = !(bd->inv_m > 0.0f) || !bd->awake;
                    if (fixed && nn.y > 0.5f) grounded = 1;
                    if (!fixed && nn.y < -0.3f) bearing = 1;
                }
                p->xc = x; p->touch = bearing ? 2.0f : touch; p->g_now = (uint32_t)grounded;
            }
            for (int i = 0; i < np; i++) {  /* velocity from the move (:1994-2001) */
                Particle* p = &w->pt[i];
                if (p->st != 1.0f) continue;
                p->g_last = p->g_now;
                p->v = op_cap(vmul(vsub(p->xc, p->xs), 1.0f / h));
                p->x = p->xc;
            }
        }
        /* ---- rest timers, spin, drawing rotation (:2002-2033) */
        for (int i = 0; i < np; i++) {
            Particle* p = &w->pt[i];
            if (p->st != 1.0f) continue;
            V3 v = p->v, ww = p->w; float t = 0.0f, st = 1.0f;
            if (p->touch > 0.0f) ww = vmul(vcross(v3(0.0f, 1.0f, 0.0f), v), 1.0f / om_max(p->r, 0.01f));
            float dr = vlen(vsub(p->x, p->anchor));
            if (p->touch < 1.5f && dr < 0.01f) t = p->t + dt;
            else if (p->touch < 1.5f && dr < 2.0f * p->r &&
                     !op_near_moving(w, vsub(p->x, v3(p->r, p->r, p->r)), vadd(p->x, v3(p->r, p->r, p->r)))) t = p->t + 0.25f * dt;
            else p->anchor = p->x;
            if (t > 0.5f || p->x.y < -50.0f) { st = 2.0f; v = v3(0, 0, 0); ww = v3(0, 0, 0); }
            p->q = qintegrate(p->q, ww, dt);
            p->v = v; p->st = st; p->t = t; p->w = ww;
        }
    }
    int asleep = 0;
    for (int i = 0; i < np; i++) if (w->pt[i].st == 2.0f) asleep++;
    w->pstats[2] = asleep;
    free(cnt); free(list); free(fill);
}
