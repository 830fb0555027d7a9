/* libsf_hip: MI355X-native physics hot path of the Shatterforge engine behind a plain C ABI.
 *
 * Three groups of entry points:
 *   1. the box world (rigid step: broadphase, SAT contacts, colouring, coloured soft-step solve, union-find sleep),
 *      handle based (sf_world_*) plus the reference's singleton exports (sf_pile_setup ... sf_pose);
 *   2. the load path (per-structure scalar-potential SOR, joint verdicts) (sf_load_*);
 *   3. seeded scene generators used by tests and the benchmark (sf_scene_*).
 * Plain pointers and sizes only. Functions never abort the process: failures return a negative code (or NULL) and
 * leave a message for sf_last_error_message(). Calls on one handle are not thread safe. Host pointers returned by
 * the library stay valid until the next mutating call on the same handle. */
#ifndef SF_H
#define SF_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors and device */
#define SF_OK 0
#define SF_E_ARG (-1)      /* bad argument (size, pointer, NaN input) */
#define SF_E_HIP (-2)      /* a HIP runtime call failed */
#define SF_E_CAP (-3)      /* input larger than the handle's capacity */
#define SF_E_NODEV (-4)    /* no GPU */
int sf_last_error(void);
const char* sf_last_error_message(void);
int sf_device_count(void);

/* ---------------------------------------------------------------- box world */
/* Parameter block (floats). Defaults: engine/src/phys.rs:101-105 (Params::default) plus the colouring limits; the GPU
 * world's solver extensions (gpu/gpu_world.js:1152-1203, :954-985) are off by default, so the defaults are the
 * reference's CPU model. The GPU world's values: SF_P_RATIO 20, SF_P_SHOCK 6, SF_P_COL_SUPPORT 2, SF_P_RED_K 32. */
enum {
    SF_P_GX, SF_P_GY, SF_P_GZ,      /* gravity (0, -9.81, 0) */
    SF_P_SUBSTEPS,                  /* 12 */
    SF_P_HERTZ,                     /* contact stiffness, 0 = a sixth of the sub-step rate */
    SF_P_DAMPING,                   /* 10 */
    SF_P_MAXPUSH,                   /* 3 m/s */
    SF_P_FRICTION,                  /* 0.6: the ground's friction (shapes carry their own) */
    SF_P_SPEC,                      /* speculative distance 0.02 m */
    SF_P_SLEEP_LIN, SF_P_SLEEP_ANG, SF_P_SLEEP_TIME, SF_P_SLEEP_DIST, SF_P_SLEEP_ROT, SF_P_ALLOW_SLEEP,
    SF_P_CELL,                      /* broadphase cell 1 m */
    SF_P_LEAK, SF_P_ANCHOR, SF_P_ANCHOR_PUSH, SF_P_ANCHOR_SLOP,
    SF_P_COL_LIMIT,                 /* colours (<= 64, default 16); contacts beyond go to the overflow: a Jacobi pass with mass splitting */
    SF_P_COL_ROUNDS,                /* colouring rounds (32) */
    SF_P_COL_SPREAD,                /* a contact picks among the lowest this-many free colours (4) */
    SF_P_TAIL_MIN,                  /* colour buckets smaller than this are solved in one workgroup (1024) */
    SF_P_SOLVE_NT,                  /* threads per workgroup of the colour-bucket solve: 64, 128 or 256 (64); results do not depend on it */
    SF_P_RATIO,                     /* mass-ratio limit: across a contact the heavier side acts at most this many times heavier (0 = off) */
    SF_P_SHOCK,                     /* shock-propagation level passes: the relax pass holds the lower body of a separating vertical contact (0 = off) */
    SF_P_REST_THRESHOLD,            /* restitution acts on points that approached faster than this (1 m/s) */
    SF_P_REST_ITERS,                /* restitution iterations per manifold (8) */
    SF_P_COL_SUPPORT,               /* colouring rounds for supports alone before the rounds for all; supports outrank the rest (0 = off) */
    SF_P_PSUB,                      /* particle debris: sub-steps per step (4) */
    SF_P_PCELL,                     /* particle debris: grid cell, twice the largest radius + 0.1 m (0.7) */
    SF_P_RED_K,                     /* contact reduction: a body pair with more manifolds than this keeps the 26 furthest out and the deepest (0 = off; 1-7 act as 8) */
    SF_NPARAM
};
/* body record, SF_BODY_F floats: p[3] q[4] v[3] w[3] inv_m inv_i[3] (principal, local) awake sleep_t reserved */
#define SF_BODY_F 20
/* shape record, SF_SHAPE_F floats: local position[3] local rotation q[4] half extents[3] friction restitution hull
 * (a pair bounces with the larger of its two restitutions, the ground's being 0: gpu/gpu_world.js:601-603; hull 0 is
 * the whole box, hull k > 0 a convex cell of the chunk library in that box, gpu/gpu_world.js:402-410) */
#define SF_SHAPE_F 13
/* a shape given in the world, for the body builder: position[3] rotation q[4] half extents[3] material (1..5) hull */
#define SF_SHAPE_IN 12
/* statistics of the last step */
enum { SF_S_AWAKE, SF_S_PAIRS, SF_S_MANIFOLDS, SF_S_POINTS, SF_S_COLOURS, SF_S_OVERFLOW, SF_S_GROUND, SF_S_WAKES,
       SF_S_ENTRIES, SF_S_CANDIDATES, SF_S_DROPPED, SF_S_ROUNDS, SF_S_REDUCED, SF_NSTAT };
/* per-manifold record from sf_world_manifolds, SF_MAN_F floats: n[3] count, then 4 points of
 * (raxn[3], rbxn[3], separation, normal impulse), then jt1 jt2 jtw stick */
#define SF_MAN_F 40

typedef struct sf_world sf_world;
/* Capacity in bodies and shapes; params may be NULL (defaults). Replaces engine/src/phys.rs World::new (:132). */
sf_world* sf_world_create(int body_cap, int shape_cap, const float* params);
void sf_world_destroy(sf_world* w);
int sf_world_set_params(sf_world* w, const float* params, int n);
/* Load bodies and shapes (shapes of a body contiguous, bodies in order). Replaces World::add_box (phys.rs:139) and
 * GpuWorld.addBodies (gpu/gpu_world.js:2656). Rejects NaN/inf input like addBodies (:2657-2659). */
int sf_world_load(sf_world* w, int nbody, const float* bodies, int nshape, const int32_t* shape_body, const float* shapes);
/* One 60 Hz step; returns the number of bodies in the solver (phys.rs World::step :152, wasm step :39) */
int sf_world_step(sf_world* w, float dt);
/* `nsteps` steps, one after another, exactly as that many sf_world_step calls (each step reads its counters and colour
 * plan back to the host once, so the steps are not enqueued ahead of each other); returns 0 or an error */
int sf_world_step_many(sf_world* w, float dt, int nsteps);
int sf_world_sync(sf_world* w);
int sf_world_read_bodies(sf_world* w, float* out);
int sf_world_stats(sf_world* w, int32_t* out);
int sf_world_nbody(sf_world* w);
/* last step's candidates (a, b; b = -1 ground) in canonical order, count returned */
int sf_world_candidates(sf_world* w, int32_t* out, int cap);
/* last step's manifolds: keys (a, b, count) and SF_MAN_F floats each */
int sf_world_manifolds(sf_world* w, int32_t* keys, float* data, int cap);
/* colour of each manifold of the last step (SF_P_COL_LIMIT = overflow, -2 = reduced) */
int sf_world_colours(sf_world* w, int32_t* out, int cap);
/* island label (minimum body index of its island) of each body after the last step */
int sf_world_islands(sf_world* w, int32_t* out);
/* 10 floats per body (position, rotation, half extents of the first shape), like engine/src/wasm.rs pose (:47-53) */
const float* sf_world_pose(sf_world* w);
/* Particle debris (GpuWorld addParticles / killParticle / blastParticles / wakeAll / readParticles,
 * gpu/gpu_world.js:2863-2896, stages :1671-2038): position-based crumbs after each step, pushed out of each other, the
 * ground and the world's shapes (one way: they never push a body), asleep after 0.5 s at rest. Capacity first (slots
 * in use are kept); records of 8 floats: p[3] v[3] radius friction, the slot of each into `slots` (-1: full or not
 * finite), the count added returned; a killed slot is reused; blasts apply at the next step, three a step; reading
 * gives 12 floats per slot in use: x y z vx, vy vz qx qy, qz qw state radius (state 0 free, 1 moving, 2 asleep);
 * stats: slots in use, moving at the last step's start, asleep at its end, free slots. */
int sf_world_particles_capacity(sf_world* w, int cap);
int sf_world_add_particles(sf_world* w, int n, const float* rec8, int32_t* slots);
int sf_world_kill_particle(sf_world* w, int i);
int sf_world_blast_particles(sf_world* w, float cx, float cy, float cz, float reach, float speed);
int sf_world_wake_particles(sf_world* w);
int sf_world_read_particles(sf_world* w, float* out12);
int sf_world_particle_stats(sf_world* w, int32_t* out4);
/* device ms of the last step's particle stage (with sf_world_set_timing on; 0 otherwise) */
int sf_world_particle_ms(sf_world* w, float* ms);
/* device time of the last step's stages in ms: broad, narrow, colour, solve, sleep (when timing is on) */
int sf_world_set_timing(sf_world* w, int on);
int sf_world_stage_ms(sf_world* w, float* out5);
/* live timing of the dominant kernel (the colour-bucket solve): an event pair around each launch; stats are the
 * summed device ms, the launches and the manifold solves they performed since timing was switched on */
int sf_world_set_kernel_timing(sf_world* w, int on);
int sf_world_kernel_stats(sf_world* w, double* ms, int64_t* launches, int64_t* items);
/* Launch mode of the sub-steps (no counterpart in the reference, whose sub-step dispatches are encoded one by one,
 * gpu/gpu_world.js:2599-2604): on (the default), the chain of sub-step launches is captured once as a HIP graph and
 * replayed each step while its launch shapes hold; off, the kernels are launched one by one. Results are the same.
 * While kernel timing is on the launches go one by one. graph_stats: captures and replays so far. */
int sf_world_set_graph(sf_world* w, int on);
int sf_world_graph_stats(sf_world* w, int64_t* captures, int64_t* replays);

/* ---- the reference's box-core exports (engine/src/wasm.rs:27-53) on one default world */
void sf_pile_setup(uint32_t n, uint32_t substeps);              /* wasm.rs:27 pile_setup */
void sf_stack_setup(uint32_t n, uint32_t substeps, uint32_t sleep); /* wasm.rs:33 stack_setup */
uint32_t sf_step(float dt);                                     /* wasm.rs:39 step */
uint32_t sf_body_count(void);                                   /* wasm.rs:40 */
uint32_t sf_manifolds(void);                                    /* wasm.rs:41 */
uint32_t sf_points(void);                                       /* wasm.rs:42 */
float sf_max_speed(void);                                       /* wasm.rs:43 */
float sf_body_pos(uint32_t i, uint32_t axis);                   /* wasm.rs:44 */
const float* sf_pose(void);                                     /* wasm.rs:47 */

/* ---------------------------------------------------------------- scenes (host only, no GPU needed) */
/* Each generator is called twice: with NULL outputs it returns the body count and writes the shape count to
 * *nshape; then with arrays of those sizes it fills them. */
/* engine/src/phys.rs:632-646 pile(n): brick-sized boxes in a loose grid (LCG seed 1) */
int sf_scene_pile(int n, float* bodies, int32_t* shape_body, float* shapes, int* nshape);
/* phys.rs:654-661 / wasm.rs:33-38: n 0.5 m cubes stacked exactly */
int sf_scene_stack(int n, float* bodies, int32_t* shape_body, float* shapes, int* nshape);
/* Rubble pile: nbody polycubes of 0.5 m bricks, sizes ~ Geometric(p) truncated at maxsize, grown on the lattice,
 * materials uniform over concrete/masonry/steel/glass/timber, placed without overlap in a box of the given size
 * (m) on the ground, velocities N(0, vsig), angular N(0, wsig). Compound mass/COM/inertia in the principal frame. */
int sf_scene_rubble(int nbody, float p, int maxsize, float bx, float by, float bz, float vsig, float wsig, uint32_t seed,
                    float* bodies, int32_t* shape_body, float* shapes, int* nshape);
/* Compound body from bricks: nb bricks at lattice cells xyz (0.5 m), materials 1..5; writes one body record and nb
 * shape records (principal frame, gpu/gpu_world.js buildBodies :2212-2249). Returns 0. */
int sf_build_compound(int nb, const int32_t* xyz, const uint8_t* mat, float* body, float* shapes);
/* A body from n shapes given in the world (SF_SHAPE_IN floats each), as gpu/gpu_world.js buildBodies (:2212-2249)
 * builds it: masses by material density and hull volume (none when fixed), centre of mass over the shapes' centroids,
 * inertia from each shape's covariance, principal frame; v, w may be NULL. Writes the body and n shape records. */
int sf_build_body(int n, const float* in, int fixed, const float* v, const float* w, float* body, float* shapes);
/* Convex-chunk library (engine/src/hull.rs, gpu/gpu_world.js hullLibrary :2113-2208): 16 fracture patterns of the
 * cube [-1,1]^3 from seed 7, each a Voronoi partition of 2-6 seeds; hull 0 is the whole cube. */
int sf_hull_count(void);
int sf_hull_pattern_count(void);
/* the hull ids of pattern p's cells (ids may be NULL); returns their count */
int sf_hull_pattern(int p, int32_t* ids);
/* 26 doubles: nv nf nseg ndir, box centre[3] and half size[3] in the cube, volume share of its box and of the cube,
 * centroid[3] and covariance per unit volume[9] in box coordinates, pattern, slot */
int sf_hull_cell(int h, double* out);
/* vertices (3 doubles each) in box coordinates, or (cube != 0) in the cube's; returns the count */
int sf_hull_vertices(int h, int cube, double* out);
/* face f: its polygon in order (3 doubles per vertex; box or cube coordinates), its plane (n, d of n.x <= d, box
 * coordinates) and what lies across it (a cube face 0..5: +x -x +y -y +z -z, or 6 + the slot of a cell of the
 * pattern); returns the vertex count */
int sf_hull_face(int h, int f, int cube, double* xyz, double* plane4, int32_t* other);
/* edges (pairs of vertex indices) and distinct edge directions (3 floats each, box coordinates); returns the edges */
int sf_hull_edges(int h, int32_t* pairs, float* dirs);
/* the narrow phase's f32 tables of every hull (sizes per hull: 16 x 4, 12 x 4, 12 x 4 floats, 96, 12, 48 bytes, 4
 * ints); returns the hull count */
int sf_hull_tables(float* v4, float* pl4, float* dir4, uint8_t* fidx, uint8_t* fcnt, uint8_t* seg, int32_t* head);
/* the pieces of a brick (centre p, rotation q, half size h) broken by pattern `pattern`: one SF_SHAPE_IN record per
 * cell, tiling the brick (gpu/gpu_world.js GpuWorld.hullShapes :2257-2263); out may be NULL; returns the count */
int sf_hull_shapes(const float* p, const float* q, const float* h, int pattern, int mat, float* out);

/* ---------------------------------------------------------------- load path (src/lib.rs structures, SOR) */
/* A standing structure of bricks on the 0.5 m lattice: material 1..5 per brick, lattice cell, infill flag.
 * Bonds and neighbours are built as finalize_brick does (src/lib.rs:668-684). */
#define SF_MODE_NONE 0
#define SF_MODE_COMPRESSION 1
#define SF_MODE_TENSION 2
#define SF_MODE_SHEAR 3
#define SF_MODE_GROUND 4
typedef struct sf_load sf_load;
/* one batch of independent structures solved together */
sf_load* sf_load_create(void);
void sf_load_destroy(sf_load* L);
/* add a structure; ext (external load per brick, kg) may be NULL; returns the structure id */
int sf_load_add(sf_load* L, int n, const uint8_t* mat, const int32_t* xyz, const uint8_t* infill, const float* ext);
/* Solve every structure to max|b - A phi| / max|b| <= rel_tol in FP64 (FP32 sweeps, FP64 residual and
 * refinement). omega <= 0 picks the over-relaxation per structure. Replaces sor_struct / warm_struct
 * (src/lib.rs:1029-1079). Returns the number of structures that converged. */
int sf_load_solve(sf_load* L, float omega, double rel_tol, int max_sweeps);
/* the same solve split so that it overlaps other work: begin launches it on the handle's own stream and returns;
 * ready is 1 once it has finished (0 before); end waits for it and returns what sf_load_solve returns. Field and
 * verdicts are readable after end. (No counterpart in the reference: a serving loop admits queries while the
 * scenes already in flight keep stepping.) */
int sf_load_solve_begin(sf_load* L, float omega, double rel_tol, int max_sweeps);
int sf_load_solve_ready(sf_load* L);
int sf_load_solve_end(sf_load* L);
/* per structure: sweeps done, final relative residual, grounded bricks */
int sf_load_info(sf_load* L, int sid, int32_t* sweeps, double* resid, int32_t* grounded);
/* the FP64 field of a structure (n doubles) */
int sf_load_field(sf_load* L, int sid, double* phi);
/* bond verdicts of a structure: 4 bonds per brick (+x,+y,+z,ground): ratio, mode, fail (stress_brick,
 * src/lib.rs:1153-1227) */
int sf_load_verdicts(sf_load* L, int sid, double* ratio4, uint8_t* mode4, uint8_t* fail4);
/* neighbour table (6 per brick, -1 none) and bonds (4 per brick) of a structure as built */
int sf_load_topology(sf_load* L, int sid, int32_t* nb6, float* bond4);
int sf_load_count(sf_load* L);
/* device time of the last sf_load_solve in ms and the FP32 / FP64 sweep counts summed over structures */
int sf_load_last_timing(sf_load* L, float* ms, int64_t* sweeps32, int64_t* sweeps64, int64_t* bricks);
/* FP32 sweeps from packed 16-byte rows (on = 1, the default) or from the full neighbour and conductance arrays
 * (on = 0); both give the same field bit for bit. No counterpart in the reference: a switch for tests and
 * measurement. */
int sf_load_set_packed(sf_load* L, int on);
/* One structure's solve and verdicts without allocating device memory. No counterpart in the reference: it is for
 * callers that cannot allocate (serving loops and harnesses that forbid hipMalloc while they run). Inputs as
 * sf_load_add, outputs as sf_load_field / sf_load_verdicts / sf_load_info (ratio4, mode4, fail4, sweeps, resid may be
 * NULL); omega, rel_tol and max_sweeps as sf_load_solve. The work goes on `stream` (a hipStream_t; NULL: the null
 * stream) and the call returns once it is done: 1 converged, 0 not, or a negative code. Every device array lives in
 * one fixed scratch array per GPU (about 10 MB, held from the first call until the process ends), so a structure
 * may have at most sf_load_static_capacity() bricks (65,000); a larger one is refused with SF_E_CAP. Calls on one
 * GPU share its scratch: the library runs them one at a time. Results equal the handle path's bit for bit. */
int sf_load_solve_static(int n, const uint8_t* mat, const int32_t* xyz, const uint8_t* infill, const float* ext,
                         float omega, double rel_tol, int max_sweeps, void* stream, double* phi, double* ratio4,
                         uint8_t* mode4, uint8_t* fail4, int32_t* sweeps, double* resid);
int64_t sf_load_static_capacity(void);

/* Frame/wall building for load tests: footprint wx x wz bricks, storeys of 7 brick layers, columns every
 * `bay` bricks, slabs, perimeter walls of masonry with glass infill; `remove` columns removed at random.
 * Returns the brick count; call with NULL outputs for the size. */
int sf_scene_building(int wx, int wz, int storeys, int bay, int remove, uint32_t seed, uint8_t* mat, int32_t* xyz, uint8_t* infill);
/* the tower of the plain-method workload: 12x12 footprint x 40 high, 2x2 concrete corner columns, a concrete
 * slab every 6th layer, masonry elsewhere */
int sf_scene_tower(uint8_t* mat, int32_t* xyz, uint8_t* infill);
/* The site-scale scene: 37 structures on a 320 x 320 m site (16 frame towers of 8-16 storeys, 12 three-storey row
 * blocks, a 56 m ring chimney, an arched viaduct, 7 steel halls), about 380k bricks, seeded. k = -1 returns the
 * structure count; k >= 0 returns structure k's brick count and, with non-NULL arrays, its bricks (lattice cells,
 * materials 1..5, infill flags). Stands in for the reference's hand-built map (build_map, src/lib.rs:3528-3590). */
int sf_scene_site(uint32_t seed, int k, uint8_t* mat, int32_t* xyz, uint8_t* infill);


/* ---------------------------------------------------------------- lattice world (the legacy brick engine)
 * One handle holds one or many independent scenes (a standing structure of lattice bricks plus loose bodies made of
 * bricks) that step together: sf_lat_step advances every scene by one frame of the reference's step
 * (src/lib.rs:3639-3721: two substeps of sphere-per-brick contacts and a 10-iteration impulse solve, then the
 * structural pass: resting loads, g - a loading, SOR load field, joint fatigue and failure, fracture bookkeeping).
 * Indices given and returned are scene local. Calls on one handle are not thread safe. */
typedef struct sf_lat sf_lat;
#define SF_LAT_BODY_F 22
#define SF_LAT_BRICK_F 13
#define SF_LAT_CONTACT_F 18
/* capacities are totals over the scenes; colours: colour limit of the contact solve (1..32), rounds: claim rounds */
sf_lat* sf_lat_create(int brick_cap, int body_cap, int max_scenes, int colours, int rounds);
void sf_lat_destroy(sf_lat* L);
/* a scene with its own brick, body and structure capacity and random seed (rnd(), src/lib.rs:323); returns its index */
int sf_lat_scene_add(sf_lat* L, int brick_cap, int body_cap, int max_structs, uint32_t seed);
/* empty a scene for reuse (same capacities) */
int sf_lat_scene_clear(sf_lat* L, int scene, uint32_t seed);
/* A scene's brick and body ranges go back to the handle's pool (the scene stays, empty, without capacity), and a
 * released scene takes new ranges of the given capacities from the pool (first fit; SF_E_CAP when none is free that
 * large). Lets a server keep scenes of uneven size in one handle. (No counterpart in the reference.) */
int sf_lat_scene_release(sf_lat* L, int scene);
int sf_lat_scene_place(sf_lat* L, int scene, int brick_cap, int body_cap, uint32_t seed);
/* device bytes a scene holds in the handle's arrays: its share of every per-brick, per-body, per-contact (four per
 * brick) and per-structure array, and its static grid. What a server admits against its HBM budget. */
int64_t sf_lat_bytes_resident(sf_lat* L, int scene);
/* the same rule for a scene not placed yet: its brick, body and structure capacities (without the grid, which
 * finalize sizes) */
int64_t sf_lat_bytes_for(sf_lat* L, int brick_cap, int body_cap, int max_structs);
/* a standing structure (add_brick, src/lib.rs:629-645): lattice cells xyz[3n], materials 1..8, infill flags */
int sf_lat_add_structure(sf_lat* L, int scene, int n, const int32_t* xyz, const uint8_t* mat, const uint8_t* infill);
/* a loose body of lattice bricks with velocity v3 and spin w3 (setup_body, src/lib.rs:844-891); returns its slot */
int sf_lat_add_body(sf_lat* L, int scene, int n, const int32_t* xyz, const uint8_t* mat, const float* v3, const float* w3);
/* neighbour tables, bonds (finalize_brick, src/lib.rs:668-696) and upload, once every brick of the scene is in */
int sf_lat_finalize(sf_lat* L, int scene);
/* ask every structure of the scene for its load field (warm_struct, src/lib.rs:1061); the sweeps run next frame */
int sf_lat_warm_all(sf_lat* L, int scene);
/* every structure's load field and joints live (dbg_wake_all, tools/debug_exports.rs:107) */
int sf_lat_wake_all(sf_lat* L, int scene);
/* a charge (explode_at, src/lib.rs:2697-2716) */
int sf_lat_explode(sf_lat* L, int scene, float x, float y, float z, float r, float power);
/* one frame of every scene (step, src/lib.rs:3639) */
int sf_lat_step(sf_lat* L, float dt);
/* ptr_stats layout (src/lib.rs:3718-3720): 0 static bricks, 1 bricks in bodies, 2 bodies, 3 awake, 4 contacts,
 * 6 worst joint ratio, 13 structures with a live load field, 14 structures never solved, 15 structures */
int sf_lat_stats(sf_lat* L, int scene, float* out16);
/* bricks, body slots, structures of a scene */
int sf_lat_counts(sf_lat* L, int scene, int32_t* out3);
int sf_lat_bodies(sf_lat* L, int scene, float* out);
int sf_lat_bricks(sf_lat* L, int scene, float* out);
/* 12 floats per structure (ptr_structs, src/lib.rs:3733-3748) */
int sf_lat_structs(sf_lat* L, int scene, float* out);
/* per dominant material: pieces, bricks, largest piece, pieces >= 50 bricks, pieces <= 2 (dbg_matstats,
 * tools/debug_exports.rs:41-56) */
int sf_lat_matstats(sf_lat* L, int scene, float* out50);
int sf_lat_contacts(sf_lat* L, int scene, float* out, int cap);
int sf_lat_rng(sf_lat* L, int scene, uint32_t* out);
/* a structure's load field from a converged solve (sf_load_field, brick order of the structure): the structure is
 * solved and its warm-up sweeps skipped (warm_struct/warm_step, src/lib.rs:1061-1079) */
int sf_lat_set_field(sf_lat* L, int scene, int structure, const float* phi);
/* joints failed under load since the scene was created: compression (crushed), tension, shear, footing
 * (stress_brick, src/lib.rs:1153-1227) */
int sf_lat_failures(sf_lat* L, int scene, uint32_t* out4);
/* wall time of the last sf_lat_step in ms and the host-device synchronisations it made */
int sf_lat_timing(sf_lat* L, double* ms_frame, int32_t* syncs);
/* wall ms of the last frame by phase: substep launches and syncs, plough, impact, topology (rebuild and
 * connectivity), structural, warm-up, statistics, all */
int sf_lat_phase_ms(sf_lat* L, double* out8);
/* finer split of the last frame's host work, ms: topology (body pull, rebuilds, connectivity, delta flush, slot
 * lists, scene upload, scenes visited), structural (upload, launches, read-back wait, event application, flush),
 * then connectivity's device search (launch and wait) */
int sf_lat_subphase_ms(sf_lat* L, double* out16);
/* the contact solve of a substep runs as one workgroup per scene when no scene has more than max_contacts contacts
 * (default 32768), else as colour-by-colour launches over all scenes; 0: always the latter. Both give the same
 * bits (no counterpart in the reference: the solve route is a device choice) */
int sf_lat_set_fused(sf_lat* L, int max_contacts);
/* worker threads of the frame's per-scene host bookkeeping (body rebuilds, connectivity), 1-16 (0: SF_HOST_THREADS,
 * else the cores, at most 16); results do not depend on it (no counterpart in the reference: a host choice) */
int sf_lat_set_host_threads(sf_lat* L, int n);
/* connectivity (src/lib.rs conn_struct :1085-1141): the search for the bricks of a structure that still reach the
 * ground runs on the device (on, the default: a breadth-first search per structure, its bits copied back once per
 * pass) or on the host (off); the detached parts are then made bodies on the host in the reference's order. Results
 * do not depend on it (no counterpart in the reference: a placement choice, a switch for tests) */
int sf_lat_set_conn_device(sf_lat* L, int on);
/* structures of at most max_bricks bricks are swept with their potential in the workgroup's LDS from packed 16-byte
 * stencil rows; larger ones from the full neighbour and conductance rows with the potential in global memory.
 * -1 (the default) takes the device's LDS limit (about 40k bricks on gfx950); 0 sends every structure down the
 * global path. Both give the same bits (no counterpart in the reference: a storage choice, a switch for tests). */
int sf_lat_set_sor_lds(sf_lat* L, int max_bricks);

/* ---- the legacy engine's own exports (src/lib.rs, tools/debug_exports.rs) on one default brick world. Prefixed
 * because `step` also names the box core's export. Pointer results are library buffers valid until the next call.
 * init builds the seeded site scene (sf_scene_site; the reference's build_map is game content) with every
 * structure's load field converged; charges are placed where thrown (their flight is game behaviour); events
 * (sound, particles, HUD) are not produced, so n_events is 0. */
uint32_t sf_legacy_init(void);                          /* init() -> bricks, src/lib.rs:3614-3638 */
void sf_legacy_step(float dt);                          /* step(dt), :3639-3721 (dt clamped to [0, 1/30]) */
void sf_legacy_explode(float x, float y, float z, float r, float p);  /* explode, :3855-3856 */
uint32_t sf_legacy_throw_charge(float ox, float oy, float oz, float vx, float vy, float vz);  /* :3933-3941; 1 held, 0 full (8) */
uint32_t sf_legacy_detonate(void);                      /* detonate() -> charges fired (r 3, power 1.4), :3942-3954 */
uint32_t sf_legacy_n_structs(void);                     /* n_structs, :4051 */
const float* sf_legacy_ptr_structs(void);               /* ptr_structs: 12 f32 per structure, :3733-3748, :4053 */
const float* sf_legacy_ptr_stats(void);                 /* ptr_stats: 16 f32, :3718-3720 */
void sf_legacy_clear_events(void);                      /* clear_events, :4026 */
uint32_t sf_legacy_n_events(void);                      /* n_events, :4027 */
const float* sf_legacy_ptr_events(void);                /* ptr_events, :4028 */
const float* sf_legacy_dbg_hist(uint32_t s);            /* tools/debug_exports.rs:3-19 */
const float* sf_legacy_dbg_bodies(void);                /* :20-40 */
const float* sf_legacy_dbg_matstats(void);              /* :41-56 */
const float* sf_legacy_dbg_speed(void);                 /* :57-71 */
const float* sf_legacy_dbg_bigbody(void);               /* :72-87 (12 values) */
const float* sf_legacy_dbg_top(uint32_t s);             /* :89-105 */
void sf_legacy_dbg_wake_all(void);                      /* :107 */
/* the default world's handle and its scene, to mix both APIs (no counterpart in the reference) */
sf_lat* sf_legacy_world(int* scene);



/* ---------------------------------------------------------------- served what-if stream (no counterpart in the
 * reference, which answers one scene at a time on the player's machine). A query is a standing structure plus charges:
 * load-path solve of the intact structure (its converged field becomes the scene's load field), the charges, then
 * `frames` frames of the brick step; the answer gives joints failed by mode, the fragment table, bricks standing and
 * pieces. Continuous batching: live queries are scenes of one lattice handle stepping in one batched frame; a query is
 * admitted when a slot is free and its scene's resident bytes (sf_lat_bytes_for at its size) fit the HBM budget beside
 * the live ones, its scene then takes ranges of its own size from the handle's pool and gives them back when it
 * retires; admitted queries' load solves run as one batch on their own stream while the frames go on. An answer is
 * the same alone and in any mix (every order in a frame is scene local). Calls on one server are not thread safe. */
typedef struct sf_serve sf_serve;
typedef struct {
    uint64_t id;
    int32_t bricks, standing, pieces, load_sweeps;
    uint32_t failed[4];         /* joints failed: compression (crushed), tension, shear, footing */
    float fragments[25];        /* per dominant material 1..5: pieces, bricks, largest, pieces of 50+, pieces of 1-2 */
    double load_resid;          /* the load solve's final relative FP64 residual */
    double latency_ms, queue_ms; /* submit to answer, submit to admission */
} sf_answer;
/* slots: scenes in flight at most; pool_bricks: the handle's brick pool shared by the live scenes; slot_bodies: body
 * slots per scene; hbm_budget: bytes the live scenes may hold (<= 0: the whole pool's); max_solving: load-solve
 * batches in flight */
sf_serve* sf_serve_create(int slots, int pool_bricks, int slot_bodies, int frames, int64_t hbm_budget, int max_solving);
void sf_serve_destroy(sf_serve* S);
/* resident bytes a query of this many bricks would hold (what admission charges) */
int64_t sf_serve_query_bytes(sf_serve* S, int bricks);
/* queue a query (charges: 5 floats each, x y z radius power); SF_E_CAP at once for one that could never fit */
int sf_serve_submit(sf_serve* S, uint64_t id, int n, const int32_t* xyz, const uint8_t* mat, const uint8_t* infill,
                    int ncharge, const float* charges5, uint32_t seed);
/* admit, start the admitted queries' load solves, bring in those solved, advance every live scene one frame, retire
 * the finished; returns the live scenes */
int sf_serve_tick(sf_serve* S, float dt);
/* move up to cap finished answers out (oldest first); returns how many */
int sf_serve_poll(sf_serve* S, sf_answer* out, int cap);
/* queued, solving, live, answers waiting, bytes in use, budget, brick-steps, batched frames */
int sf_serve_status(sf_serve* S, int64_t* out8);
/* wall-clock ms the server has spent in admission (placing scenes, starting load solves), activation (scenes built,
 * converged fields set, charges), frames and retirement (answers read, ranges released), since it was created */
int sf_serve_timing(sf_serve* S, double* out4);
/* the server's lattice handle (read-only use: statistics, timing) */
sf_lat* sf_serve_world(sf_serve* S);

/* ---- structural analysis: the spring-network solve of the new engine (engine/src/structure.rs:191-783) over its
 * brick store (engine/src/brick.rs). A brick is a box on the 0.125 m tick grid (min corner and size in ticks, sizes
 * 1..255); joints are the faces two bricks share, with exact areas. Materials: 0 concrete, 1 masonry, 2 steel,
 * 3 glass, 4 timber; paint as the reference's elements (0 slab, 1 column, 3 core, 8 timber floor, ...); flags:
 * 1 alive, 2 infill, 4 dynamic, 8 cast monolithically. Whole boxes only: the fracture cells (convex hulls) of the
 * reference's store are not represented. */
typedef struct sf_bricks sf_bricks;
typedef struct sf_warm sf_warm;
typedef struct sf_net sf_net;
/* Bricks::new / add (brick.rs:147-170); add returns the brick's index (SF_E_ARG for a size outside 1..255, a corner
 * outside 16 bits or a material outside 0..4) */
sf_bricks* sf_bricks_create(void);
void sf_bricks_destroy(sf_bricks* b);
int sf_bricks_len(const sf_bricks* b);
int sf_bricks_add(sf_bricks* b, const int32_t* min3, const int32_t* size3, int mat, int paint, int structure);
int sf_bricks_set_flags(sf_bricks* b, int i, int flags);
int sf_bricks_flags(const sf_bricks* b, int i);
/* structure.rs mass (:66-71): the brick's mass in kg (a steel member weighs its section) */
float sf_bricks_mass(const sf_bricks* b, int i);
/* Bricks::remove (brick.rs:182-190): its joints go, the record stays dead */
int sf_bricks_remove(sf_bricks* b, int i);
/* Bricks::build_joints (brick.rs:447-458): every joint from geometry, numbered as the reference numbers them; returns
 * the live joints */
int sf_bricks_build_joints(sf_bricks* b);
/* per joint slot: ja (min side), jb (-1 for a removed joint), axis, area in tick^2, damage; returns the slots */
int sf_bricks_joints(const sf_bricks* b, int32_t* out5, int cap);
/* a joint's damage byte (bit 0 cracked, bits 1..7 wear), as the game layer sets it */
int sf_bricks_set_joint_damage(sf_bricks* b, int j, int dmg);
/* Bricks::save / Bricks::load (brick.rs:116-145): the store as bytes, each array as its u32 length and elements padded
 * to 8 bytes, in the reference's order and types (the spatial index's keys ascending); save returns the byte count
 * (out may be NULL; SF_E_CAP when cap is smaller), load returns a new store or NULL for malformed bytes */
int64_t sf_bricks_save(const sf_bricks* b, uint8_t* out, int64_t cap);
sf_bricks* sf_bricks_load(const uint8_t* src, int64_t n);
/* the per-brick fields the store carries for its format: owning body (-1 none), damage byte, convex cell */
int sf_bricks_set_cell(sf_bricks* b, int i, int body, int damage, int hull);
int sf_bricks_cell(const sf_bricks* b, int i, int32_t* out3);
/* a brick's box: min corner[3] and size[3] in ticks */
int sf_bricks_box(const sf_bricks* b, int i, int32_t* out6);
/* structure.rs Warm (:178-189): the last solution of a load case, per cell; a case given one starts from it and keeps
 * its own solution in it */
sf_warm* sf_warm_create(void);
void sf_warm_destroy(sf_warm* w);
int sf_warm_size(const sf_warm* w);
/* Analyses queued, solved together: each case is analyse_warm (structure.rs:191) of the bricks `set` (live, once
 * each) under per-brick loads (N) along dir3, supported at the bricks whose sink is set, plus point loads of
 * (index into set, x, z in ticks, N). sf_net_add returns the case's index; sf_net_solve builds every case's cells,
 * links, springs and multigrid levels on host threads and runs all their conjugate-gradient solves in one launch
 * (a 1,024-thread workgroup per case), then the face checks. The cases and their results stay until sf_net_clear; the
 * bricks a case names must not change until then. */
sf_net* sf_net_create(void);
void sf_net_destroy(sf_net* N);
int sf_net_add(sf_net* N, const sf_bricks* b, int n, const uint32_t* set, const float* load, const uint8_t* sink,
               const float* dir3, int npoint, const float* point4, sf_warm* warm);
int sf_net_solve(sf_net* N);
int sf_net_clear(sf_net* N);
int sf_net_count(const sf_net* N);
/* host threads that build the cases (0: up to 16) */
int sf_net_set_host_threads(sf_net* N, int n);
/* The route of a case: in a batch of at most max_cases cases, one of at least min_cells cells is solved a kernel per
 * step over the whole chip, block_its iterations queued between checks (min_cells -1: every case in its own
 * workgroup). Default -1, 16, 6: the kernel-per-step route is switched off, having made a GPU memory fault (a
 * 1,760-cell case with a warm start) whose cause is not yet found; its tests are skipped. */
int sf_net_set_route(sf_net* N, int min_cells, int max_cases, int block_its);
/* Analysis (structure.rs:119-133): [max_ratio, iterations, cells, links, levels, worst inner face: ratio, brick, axis,
 * flow N, moment^2, joints checked, inner faces checked] */
int sf_net_result(const sf_net* N, int c, double* out12);
/* cells per multigrid level; returns the levels */
int sf_net_levels(const sf_net* N, int c, int32_t* out8);
/* every joint checked, in the reference's order, with its stress over strength (> 1: it fails); returns the count */
int sf_net_joints(const sf_net* N, int c, int32_t* jid, float* ratio, int cap);
/* every face inside a brick checked: (brick, axis, plane in ticks) and its ratio (> 1: it cracks there, > 0.7: it
 * wears); returns the count */
int sf_net_inner(const sf_net* N, int c, int32_t* bat3, float* ratio, int cap);
/* the cells (brick, lo[3], hi[3] in ticks) and their displacements and rotations (6 per cell) */
int sf_net_cells(const sf_net* N, int c, int32_t* cell7, double* x6, int cap);
/* [host build ms, upload and read-back ms, device solve ms, device checks ms] of the last sf_net_solve */
int sf_net_ms(const sf_net* N, double* out4);
/* the host build of the last sf_net_solve summed over its cases (thread ms): cells, links and springs; multigrid
 * levels; loads, warm starts and checks; then packing (wall ms) */
int sf_net_host_ms(const sf_net* N, double* out4);

#ifdef __cplusplus
}
#endif
#endif
