/* CPU oracle for the physics hot path: C restatement of the reference's box core (engine/src/phys.rs) extended to
 * compound bodies and per-shape friction as in gpu/gpu_world.js (buildBodies :2212-2249, emit :601-603), the
 * colouring and island labelling the device path uses, and the legacy load-path solve of src/lib.rs
 * (sor_brick :1005-1026, cond_w :1142-1151, stress_brick :1153-1227).
 *
 * TEST INFRASTRUCTURE ONLY: tests, the smoke check and the benchmark's CPU-baseline leg load it as the checker.
 * The product library never links or calls it. */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- box world (engine/src/phys.rs) ---- */
typedef struct ob_world ob_world;

/* Solve/pair order. OB_ORDER_REFERENCE reproduces phys.rs exactly: pairs in hash-table slot order
 * (phys.rs:376-404), ground manifolds first, sequential Gauss-Seidel in manifold order, wakes applied at once
 * (phys.rs:441-444). OB_ORDER_COLOURED is the device order: candidates per shape (ground, then partners by index),
 * colour-bucketed solve in colour order then the overflow in manifold order, wakes from the step-start snapshot. */
enum { OB_ORDER_REFERENCE = 0, OB_ORDER_COLOURED = 1 };

/* Parameter block, phys.rs:74-106 defaults plus the colouring limits and the GPU world's solver extensions (mass-ratio
 * limit, shock-propagation level passes, restitution threshold and iterations, supports-first colouring rounds; all
 * off by default). Index names in sf.h (SF_P_*). */
#define OB_NPARAM 33

/* body record: p[3] q[4] v[3] w[3] inv_m inv_i[3] awake sleep_t reserved  (20 floats) */
#define OB_BODY_F 20
/* shape record: lp[3] lq[4] h[3] mu e hull  (13 floats: friction, restitution, convex cell 0 = whole box); shape_body[] gives the owning body */
#define OB_SHAPE_F 13

ob_world* ob_create(int order);
void ob_destroy(ob_world* w);
void ob_set_params(ob_world* w, const float* prm, int n);
/* Bodies and shapes; shapes of one body must be contiguous and in body order. */
/* The convex-chunk tables (f32, field by field as the library's narrow phase reads them: per hull 16 vertices x 4,
 * 12 planes x 4, 12 directions x 4 floats, 12 x 8 face vertex indices, 12 face counts, 24 x 2 edge ends, 4 counts
 * nv nf nseg ndir), copied; before ob_load of shapes whose hull (record float 12) is above 0. */
int ob_set_hulls(ob_world* w, int count, const float* v4, const float* pl4, const float* dir4, const uint8_t* fidx,
                 const uint8_t* fcnt, const uint8_t* seg, const int32_t* head);
int ob_load(ob_world* w, int nbody, const float* bodies, int nshape, const int32_t* shape_body, const float* shapes);
/* one pair's contact points before reduce4, by the chunk path (chunk_path != 0) or the box SAT; shapes p[3] q[4] h[3]
 * hull (b NULL: the ground); out: n[3] count, then up to 8 x (p[3], separation); returns the count */
int ob_pair_contact(const ob_world* w, int chunk_path, const float* a, const float* b, float margin, float* out);
void ob_step(ob_world* w, float dt);
int ob_nbody(const ob_world* w);
/* read back body state (OB_BODY_F floats per body) */
void ob_read_bodies(const ob_world* w, float* out);
/* stats: [awake solver bodies, pairs, manifolds, points, colours used, overflow, ground candidates, wakes] */
void ob_stats(const ob_world* w, int32_t* out8);
/* manifolds the last step's contact reduction marked (SF_P_RED_K > 0; gpu/gpu_world.js:834-920) */
int ob_reduced(const ob_world* w);
/* last step's candidate list in canonical order: (a, b) with b = -1 for the ground; returns the count */
int ob_candidates(const ob_world* w, int32_t* out, int cap);
/* last step's manifolds: keys (a, b) and point counts; plus per manifold: normal[3], npts, and up to 4 points of
 * (ra x n [3], rb x n [3], adj, jn) after the solve, jt1, jt2, jtw, stick */
int ob_manifolds(const ob_world* w, int32_t* keys, float* data, int cap);
#define OB_MAN_F 40
/* colour per manifold (K = overflow) for the last step (coloured order only) */
int ob_colours(const ob_world* w, int32_t* out, int cap);
/* island label per body (min body index of its island, -1 for bodies not awake at the step start) */
void ob_islands(const ob_world* w, int32_t* out);
/* shock-propagation heights of the last step per body (255 - level; 255 for bodies not moving, 0 for bodies no
 * support chain reached), all 0 when the passes are off */
void ob_levels(const ob_world* w, int32_t* out);
/* particle debris (gpu/gpu_world.js:1671-2038, addParticles/killParticle/blastParticles/readParticles :2870-2896):
 * capacity, records of 8 floats (p[3] v[3] radius friction) returning their slots, kill, blast (applied at the next
 * step, three a step), wake all, read 12 floats per slot (x y z vx, vy vz qx qy, qz qw state radius), stats
 * [slots, moving at the stage's start, asleep at its end, free slots] */
int ob_particles_capacity(ob_world* w, int cap);
int ob_add_particles(ob_world* w, int n, const float* rec, int32_t* slots);
void ob_kill_particle(ob_world* w, int i);
void ob_blast_particles(ob_world* w, float cx, float cy, float cz, float reach, float speed);
void ob_wake_particles(ob_world* w);
int ob_read_particles(const ob_world* w, float* out);
void ob_particle_stats(const ob_world* w, int32_t* out4);
/* brute-force O(n^2) candidate pairs for the current pose (test pin for the broadphase) */
int ob_bruteforce_pairs(ob_world* w, float dt, int32_t* out, int cap);

/* ---- load path (src/lib.rs structures / SOR) ---- */
/* A structure: n bricks on the lattice with material codes 1..5 (concrete, masonry, steel, glass, timber), lattice
 * coordinates, per-brick bond capacities bond[4] (+x,+y,+z, ground) and external load ext. nb6[6] neighbour brick
 * index or -1 (DIRS order +x,-x,+y,-y,+z,-z). */
typedef struct {
    int n;
    const uint8_t* mat;
    const int32_t* xyz;   /* 3 per brick */
    const int32_t* nb6;   /* 6 per brick */
    const float* bond;    /* 4 per brick */
    const float* ext;     /* 1 per brick (may be NULL) */
} ol_struct;

/* One lexicographic SOR sweep in brick order (sor_struct, lib.rs:1029-1040); returns max|d| / max(phi, 1). */
float ol_sor_sweep(const ol_struct* s, float* phi, float omega);
/* FP64 solve of the same linear system by conjugate gradients to rel_tol; returns iterations (or -1). */
int ol_solve64(const ol_struct* s, double* phi, double rel_tol, int max_iter);
/* FP64 residual max|b - A phi| / max|b| */
double ol_residual64(const ol_struct* s, const double* phi);
/* Joint verdicts from a field (stress_brick, lib.rs:1153-1227, static bricks): per brick 4 bonds (+x,+y,+z, ground):
 * ratio (FP64), mode (0 none, 1 compression, 2 tension, 3 shear, 4 ground) and fail = ratio > 1. */
void ol_verdicts64(const ol_struct* s, const double* phi, double* ratio4, uint8_t* mode4, uint8_t* fail4);
/* The reference's f32 verdict from an f32 field (same function, f32 arithmetic as the reference). */
void ol_verdicts32(const ol_struct* s, const float* phi, float* ratio4, uint8_t* mode4);
/* grounded connectivity: 1 for bricks reachable from a grounded brick through linked bonds (conn_struct :1085-1098) */
int ol_grounded(const ol_struct* s, uint8_t* out);
/* neighbour table and bond capacities of one standing structure (finalize_brick :668-684, bond_cap :646-649) */
void ol_finalize(int n, const uint8_t* mat, const int32_t* xyz, const uint8_t* infill, int32_t* nb6, float* bond);

/* ---------------------------------------------------------------- lattice world (lattice.c): the legacy brick
 * engine's physics frame (src/lib.rs), in the reference's sequential orders or the GPU's orders. */
typedef struct LW LW;
enum { LL_ORDER_SEQ = 0, LL_ORDER_DEV = 1 };
/* mixed orders (test use): LL_ORDER_MIX | any of the part flags runs those parts in the device's order and the rest
 * in the reference's: the contact solve, the SOR sweeps, the joint judgement */
enum { LL_ORDER_MIX = 0x10, LL_MIX_SOLVE = 0x1, LL_MIX_SWEEP = 0x2, LL_MIX_STRESS = 0x4 };
#define LL_BODY_F 22
#define LL_BRICK_F 13
#define LL_CONTACT_F 18
LW* ll_create(int order, int brick_cap, int body_cap, int ncolour, int rounds, uint32_t seed);
void ll_destroy(LW* w);
/* a standing structure of lattice bricks (add_brick, lib.rs:629-645); returns its index */
int ll_add_structure(LW* w, int n, const int32_t* xyz, const uint8_t* mat, const uint8_t* infill);
/* a loose body of lattice bricks with the given velocity and spin (setup_body, lib.rs:844-891); returns its slot */
int ll_add_body(LW* w, int n, const int32_t* xyz, const uint8_t* mat, const float* v3, const float* w3);
/* neighbour tables, bonds and structure counts once every brick is in (finalize_brick, lib.rs:668-696) */
void ll_finalize(LW* w);
/* ask every structure for its load field (warm_struct, lib.rs:1061-1066); the sweeps run in the next frame */
void ll_warm_all(LW* w);
/* every structure's load field and joints live (dbg_wake_all, tools/debug_exports.rs:107) */
void ll_wake_all(LW* w);
void ll_step(LW* w, float dt);
void ll_explode(LW* w, float x, float y, float z, float r, float power);
void ll_stats(const LW* w, float* out16);
int ll_nbricks(const LW* w);
int ll_nbodies(const LW* w);
int ll_nstructs(const LW* w);
void ll_bodies(const LW* w, float* out);
void ll_bricks(const LW* w, float* out);
void ll_structs(const LW* w, float* out);
void ll_matstats(const LW* w, float* out50);
int ll_contacts(const LW* w, float* out, int cap);
uint32_t ll_rng(const LW* w);
/* a structure's load field from a converged solve (brick order of the structure); marks it solved */
void ll_set_field(LW* w, int s, const float* phi);
/* joints failed under load since creation: compression (crushed), tension, shear, footing */
void ll_failures(const LW* w, uint32_t* out4);
/* test hooks: a piece of a substep run on its own (see lattice.c) */
void ll_set_body(LW* w, int b, const float* p3, const float* q4, const float* v3, const float* w3);
int ll_probe_contacts(LW* w);
void ll_probe_prep(LW* w, float h);
void ll_probe_solve(LW* w, int passes);
void ll_probe_pinch(LW* w, int b, const float* vpre3, const float* wpre3);
uint32_t ll_probe_crush(LW* w, int start, float e, const float* p3, uint32_t cap);
void ll_set_last(LW* w, int b, const float* v3, const float* w3);
void ll_probe_structural(LW* w, float dt);
void ll_set_nc(LW* w, int b, uint32_t nc);
void ll_set_asleep(LW* w, int b);

/* ---- spring-network structural analysis (engine/src/structure.rs:191-783) over a brick store (engine/src/brick.rs) */
typedef struct os_store os_store;
typedef struct os_result os_result;
typedef struct os_warm os_warm;
/* dot-product order: element after element (the reference), or as the device's one-workgroup solve sums */
enum { OS_ORDER_REFERENCE = 0, OS_ORDER_DEVICE = 1 };
os_store* os_create(void);
void os_destroy(os_store* b);
int os_len(const os_store* b);
/* a brick: min corner and size in ticks (0.125 m, sizes 1..255), material (0 concrete, 1 masonry, 2 steel, 3 glass,
 * 4 timber), paint, structure; returns its index (-1: bad size) */
int os_add(os_store* b, const int32_t* min3, const int32_t* size3, int mat, int paint, int structure);
/* flags: 1 alive, 2 infill, 4 dynamic, 8 cast monolithically */
void os_set_flags(os_store* b, int i, int flags);
int os_flags(const os_store* b, int i);
void os_remove(os_store* b, int i);
int os_build_joints(os_store* b);
int os_njoint_slots(const os_store* b);
int os_live_joints(const os_store* b);
/* per joint slot: ja, jb (-1 for a free slot), axis, area in tick^2, damage */
int os_joints(const os_store* b, int32_t* out5, int cap);
void os_set_joint_damage(os_store* b, int j, int dmg);
float os_mass(const os_store* b, int i);
os_warm* os_warm_create(void);
void os_warm_destroy(os_warm* w);
int os_warm_size(const os_warm* w);
/* analyse the bricks `set` under per-brick loads (N) along dir, supported at the `sink` bricks; point loads as
 * (index into set, x, z in ticks, N); warm: the last solution to start from and to keep (may be NULL) */
os_result* os_analyse(const os_store* b, int n, const uint32_t* set, const float* load, const uint8_t* sink,
                      const float* dir3, int npoint, const float* point4, int order, os_warm* warm);
void os_result_free(os_result* r);
/* [max_ratio, iterations, nodes, links, levels, worst-in ratio, brick, axis, flow, moment^2, joints checked, inner
 *  faces checked] */
void os_result_scalars(const os_result* r, double* out12);
void os_result_levels(const os_result* r, int32_t* out8);
int os_result_joints(const os_result* r, int32_t* jid, float* ratio, int cap);
int os_result_inner(const os_result* r, int32_t* bat3, float* ratio, int cap);
int os_result_nodes(const os_result* r, int32_t* node7, float* load, uint8_t* fixed, double* x6, int cap);
int os_result_springs(const os_result* r, uint32_t* ends2, double* sp12, int cap);

#ifdef __cplusplus
}
#endif
#endif
