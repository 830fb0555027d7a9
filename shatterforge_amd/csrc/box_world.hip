// Box world on the GPU: the rigid step of the reference's box core (engine/src/phys.rs World::step :152-243) with
// compound bodies (gpu/gpu_world.js :2212-2249), as hand-written gfx950 kernels.
//
// One step, all on one stream, counts kept on the device:
//   prep_bodies / prep_shapes    world poses, AABBs with the speculative margin, solver slots (phys.rs:158-172)
//   broadphase                   every cell a shape's AABB covers goes into a hashed counting-sort table; each
//                                shape then finds its partners of higher index in its cells, reporting a pair only
//                                in the cell of the max of the two min corners (phys.rs:366-404), ground first,
//                                partners sorted: the candidate list is canonical and deterministic
//   narrow                       one lane per candidate: 15-axis SAT with face clipping or ground corners
//                                (phys.rs:418-432, :531-628), reduce4, wake requests (:441-444)
//   build                        compacted manifolds, warm start from last step's sorted keys, anchored patch
//                                friction, effective masses (:180-211, :450-498)
//   colour                       deterministic claim rounds (one of the lowest free colours at both bodies, the
//                                highest priority claim wins); contacts left over go to the overflow
//   sub-steps                    gravity + spin clamp + warm-start gather per body (its contacts in manifold order),
//                                coloured soft-step solve, integrate, relax (:214-230, :245-324)
//   write back, islands, sleep   union-find by atomic hooking of the larger root under the smaller (labels = the
//                                island's minimum body index), rest timers, island minimum (:231-237, :326-354)
// Layouts are struct-of-arrays of float4 so every lane's loads are coalesced.
#include "sf_dev.h"
#include "sf_host.h"
#include "sf_scan.h"
#include "hulls.h"
#include "../../include/sf.h"
#include <vector>
#include <string>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <array>

namespace {

constexpr int NT = 256;  // threads per block of the grid-stride kernels
constexpr int sf_hull_maxv = sf::HULL_MAXV, sf_hull_maxf = sf::HULL_MAXF, sf_hull_maxfv = sf::HULL_MAXFV,
              sf_hull_maxs = sf::HULL_MAXS, sf_hull_maxd = sf::HULL_MAXD;

// device counters
// (C_ACTIVE0 + g: set when a manifold claimed in the rounds the host checks as group g)
// (C_NREST: manifolds of the step that may bounce, a restitution and a point approaching past the threshold)
enum { C_NE, C_NCAND, C_NMAN, C_NPTS, C_NPAIR, C_NGROUND, C_NWAKE, C_NOVF, C_NCOL, C_DROP, C_AWAKE, C_NSLOT, C_NBIGC, C_NBIGL,
       C_NREST, C_NRED, C_NREDP, C_NHULL, C_ACTIVE0, C_NACTIVE = 16, C_NCTR = C_ACTIVE0 + C_NACTIVE };

// manifold rows (each row is a float4 array of mcap entries)
// (the inertia terms, a body's world inverse inertia times an arm or the normal, are not stored: the solve forms them
// from the bodies' inverse inertia, fixed for the step, with the build's operations; MR_JN holds the four points'
// normal impulses. Spare lanes: MR_LAF.w the pair's restitution, MR_LBF.w set once a relax pass pushed, each point's
// third row .w its normal velocity at the build, for the restitution pass.)
enum { MR_N, MR_T1, MR_T2, MR_E0, MR_J, MR_CA1, MR_CA2, MR_CB1, MR_CB2, MR_JN, MR_LAF, MR_LBF, MR_P0 };
constexpr int MR_PER_PT = 3;  // raxn|separation, rbxn|nm, la|-
constexpr int MR_ROWS = MR_P0 + 4 * MR_PER_PT;

// A step's manifolds. Canonical index m (candidate order): header, sort key, and pos[m], the manifold's position
// in colour order. The rows are stored by position, so a colour bucket is one contiguous, coalesced range.
struct ManBuf {
    int4* hdr;        // by m: shape a, shape b (-1 ground), slot a, slot b
    uint64_t* key;    // by m: sort key (a, b + 1)
    uint32_t* pos;    // by m: position in colour order
    int4* hdrp;       // by position: the header again
    float4* row;      // by position: MR_ROWS x mcap
};

struct StepK {
    float dt, h, inv_h, bias_rate, mass_scale, imp_scale, max_push, anchor_push, anchor_slop;
    float spec, inv_cell, friction, sleep_lin2, sleep_ang2, sleep_time, sleep_dist2, sleep_rot, wmax;
    float gx, gy, gz;
    int anchor, allow_sleep, K, spread, use_bias;
    // the GPU world's extensions: mass-ratio limit (0 off), shock-propagation passes (0 off, else even), restitution
    // threshold and iterations, whether any shape bounces, support-only colouring rounds
    float ratio, rest_thr;
    int shock, rest_iters, any_rest, col_support;
    int red_k;  // contact reduction: body pairs with more manifolds than this are reduced (0 off, else at least 8)
};

// the chunk library's f32 tables on the device (hull h's entries at h times each field's stride, as sf::HullTables)
struct HullDev { const float4 *v, *pl, *dir; const uint8_t *fidx, *fcnt, *seg; const int4* head; };

struct Dev {
    int nb, ns;
    uint32_t ecap, tcap, tmask, ccap, mcap;
    // bodies
    float4 *bp, *bq, *bv, *bw, *bii, *bsp, *bsq, *brot;  // bp: p|inv_m, bii: inv_i|sleep_t, brot: 3 per body
    int32_t *bflag, *bslot;
    uint32_t* wake;
    // solver slots (nb + 1, slot 0 = the static dummy, all zero)
    // solver slots: one 128-byte record each (SLOT_F4 float4: v|inv_m, w, position change, rotation change, q,
    // world inverse inertia rows), so the two bodies a manifold gathers are two cache lines, not a dozen
    float4* sbod;
    float4* sp0;
    // shapes
    int32_t* shb;
    float4 *slp, *slr, *shh, *swp, *swr, *slo, *shi;  // slp: lp|reach, slr/swr: 3 per shape, shh: h|mu
    float* srest;                                     // restitution per shape
    int32_t* shull;                                   // hull id per shape (0: the whole box)
    HullDev hl;
    uint32_t* hlist;                                  // this step's candidates with a convex chunk
    // broadphase
    uint32_t *ecnt, *eoff, *ekey;
    int32_t* eshape;
    uint32_t *tcnt, *tstart, *tfill, *tkey;
    int32_t* tent;    // shape of each table entry
    uint32_t* tidx;   // cell entry of each table entry
    uint32_t *soff, *slist;  // occupied table slots: offset in the list (scan of the flags), the list
    uint32_t *epc, *erel;  // per cell entry: partners homed there, offset within the shape's candidate list
    int32_t* ekeep;        // per cell entry: its first ENTRY_KEEP partners, as the count pass found them
    uint32_t* bigc;        // shapes with more than 16 partners this step
    uint32_t* bigl;        // solver slots whose contact lists are longer than a wavefront this step
    uint32_t *ccnt, *coff;
    int2* cand;
    float4* nar;  // 5 per candidate: n|count, points p|sep
    float4* nar2; // 5 per shape: a chunk's second patch of ground corners (its ground candidate has two manifolds)
    uint32_t *hit, *hoff;
    ManBuf cur, old;
    float4* wrec;  // by body contact-list entry, WREC_F4 float4: the warm start's velocity and spin change, then the
                   // overflow velocity change v, w (zero outside the overflow)
    // colouring
    int32_t *mcol, *mcand;
    unsigned long long *mpri, *bmask, *claim;
    uint32_t *colcnt, *coloff, *colfill, *mbin, *subcnt, *suboff, *ovcnt;
    uint32_t* chunkc;  // per colour, per chunk of NT manifolds: count, then its exclusive offset within the colour
    uint32_t* ptpart;  // contact points per workgroup of the build (summed by k_sum_points)
    // per-slot manifold lists (m * 2 + side), in manifold order
    uint32_t *blcnt, *bloff, *blfill, *blist;
    uint32_t *blinv, *minv;  // the list entry of (colour position, side), of (manifold, side); ~0u: the static world
    // shock-propagation heights per solver slot (255 - level), the two copies the level passes alternate between
    uint32_t *lvl, *lvl2;
    // contact reduction: body-pair table (key, manifold count, the pair's place among the reduced pairs + 1), each
    // manifold's table slot and reduced flag, and per reduced pair the RED_D extreme picks (score << 32 | manifold)
    unsigned long long* rkey;
    uint32_t *rcnt, *rbig, *rslot, *rflag;
    unsigned long long* rext;
    uint32_t rmask;
    // islands
    int32_t* parent;
    uint32_t* sst;     // sleep state bits per body (SS_*)
    uint32_t* ctr;
};

constexpr int SLOT_F4 = 8;
__device__ __forceinline__ float4& SV(const Dev& D, int s) { return D.sbod[(size_t)s * SLOT_F4 + 0]; }
__device__ __forceinline__ float4& SW(const Dev& D, int s) { return D.sbod[(size_t)s * SLOT_F4 + 1]; }
__device__ __forceinline__ float4& SDP(const Dev& D, int s) { return D.sbod[(size_t)s * SLOT_F4 + 2]; }
__device__ __forceinline__ float4& STH(const Dev& D, int s) { return D.sbod[(size_t)s * SLOT_F4 + 3]; }
__device__ __forceinline__ float4& SQ(const Dev& D, int s) { return D.sbod[(size_t)s * SLOT_F4 + 4]; }
__device__ __forceinline__ float4* SII(const Dev& D, int s) { return D.sbod + (size_t)s * SLOT_F4 + 5; }
constexpr int WREC_F4 = 4;
__device__ __forceinline__ float4* WREC(const Dev& D, uint32_t e) { return D.wrec + (size_t)e * WREC_F4; }
__device__ __forceinline__ float4* OVL(const Dev& D, uint32_t e) { return D.wrec + (size_t)e * WREC_F4 + 2; }
__device__ __forceinline__ float rdl(float x, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l)); }
#define GRID_STRIDE(i, n) for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

__device__ __forceinline__ M3 load_m3(const float4* p) { M3 r; r.c[0] = v3f4(p[0]); r.c[1] = v3f4(p[1]); r.c[2] = v3f4(p[2]); return r; }
__device__ __forceinline__ void store_m3(float4* p, const M3& r) { p[0] = f4(r.c[0], 0.0f); p[1] = f4(r.c[1], 0.0f); p[2] = f4(r.c[2], 0.0f); }
__device__ __forceinline__ float4& MROW(const Dev& D, const ManBuf& M, int r, uint32_t p) { return M.row[(size_t)r * D.mcap + p]; }
__device__ __forceinline__ uint32_t cell_key(int x, int y, int z) { return ((uint32_t)x & 1023u) | (((uint32_t)y & 1023u) << 10) | (((uint32_t)z & 1023u) << 20); }
__device__ __forceinline__ uint32_t slot_of(uint32_t key, uint32_t mask) { return mix32(key * 0x9E3779B1u) & mask; }

// Cells a box covers. A box over more than 8 cells along an axis (or 512 in all, or not finite) is left out of the
// grid and counted as dropped: no sane body is that large or that fast, and it bounds every lane's loop.
constexpr int MAX_CELLS_AXIS = 8;
struct CellRange { int x0, y0, z0, x1, y1, z1; uint32_t n; };
__device__ __forceinline__ CellRange cell_range(V3 lo, V3 hi, float inv_cell) {
    CellRange c;
    c.n = 0;
    c.x0 = c.y0 = c.z0 = 0; c.x1 = c.y1 = c.z1 = -1;
    float fl[3] = {lo.x * inv_cell, lo.y * inv_cell, lo.z * inv_cell}, fh[3] = {hi.x * inv_cell, hi.y * inv_cell, hi.z * inv_cell};
    for (int a = 0; a < 3; a++) if (!(fabsf(fl[a]) < 1.0e8f) || !(fabsf(fh[a]) < 1.0e8f) || !(fh[a] - fl[a] < (float)MAX_CELLS_AXIS)) return c;
    c.x0 = (int32_t)sf_floor(fl[0]); c.y0 = (int32_t)sf_floor(fl[1]); c.z0 = (int32_t)sf_floor(fl[2]);
    c.x1 = (int32_t)sf_floor(fh[0]); c.y1 = (int32_t)sf_floor(fh[1]); c.z1 = (int32_t)sf_floor(fh[2]);
    int nx = c.x1 - c.x0 + 1, ny = c.y1 - c.y0 + 1, nz = c.z1 - c.z0 + 1;
    if (nx < 1 || ny < 1 || nz < 1 || nx > MAX_CELLS_AXIS || ny > MAX_CELLS_AXIS || nz > MAX_CELLS_AXIS) { c.x1 = c.x0 - 1; return c; }
    c.n = (uint32_t)(nx * ny * nz);
    return c;
}

// ---------------------------------------------------------------- prepare
// A statistics count added from a whole workgroup at once (every thread calls it with its own share): one atomic per
// workgroup, not one per wavefront on a single address.
__device__ __forceinline__ void block_count(uint32_t v, uint32_t* dst) {
    __shared__ uint32_t part[NT / 64];
    for (int h = 32; h >= 1; h >>= 1) v += __shfl_xor(v, h, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < NT / 64; k++) t += part[k];
        if (t) atomicAdd(dst, t);
    }
    __syncthreads();
}

__global__ void __launch_bounds__(NT) k_prep_bodies(Dev D) {
    uint32_t nawake = 0;
    GRID_STRIDE(i, (uint32_t)D.nb) {
        float4 q4 = D.bq[i];
        M3 R = q_to_m3(qf4(q4));
        store_m3(D.brot + 3 * (size_t)i, R);
        int awake = D.bflag[i] & 1;
        int slot = awake ? (int)i + 1 : 0;
        D.bslot[i] = slot;
        D.wake[i] = 0u;
        D.parent[i] = (int32_t)i;
        D.sst[i] = 0u;
        if (awake) {
            float4 p = D.bp[i];
            SV(D, slot) = make_float4(D.bv[i].x, D.bv[i].y, D.bv[i].z, p.w);
            SW(D, slot) = make_float4(D.bw[i].x, D.bw[i].y, D.bw[i].z, 0.0f);
            SDP(D, slot) = make_float4(0, 0, 0, 0);
            STH(D, slot) = make_float4(0, 0, 0, 0);
            SQ(D, slot) = q4;
            D.sp0[slot] = make_float4(p.x, p.y, p.z, 0.0f);
            store_m3(SII(D, slot), sandwich(R, v3f4(D.bii[i])));
            nawake++;
        }
    }
    block_count(nawake, &D.ctr[C_AWAKE]);
}

__global__ void __launch_bounds__(NT) k_prep_shapes(Dev D, StepK K) {
    GRID_STRIDE(s, (uint32_t)D.ns) {
        int b = D.shb[s];
        M3 R = load_m3(D.brot + 3 * (size_t)b);
        M3 Rs = m3mul(R, load_m3(D.slr + 3 * (size_t)s));
        float4 lp4 = D.slp[s];
        V3 p = v3f4(D.bp[b]);
        V3 wp = vadd(p, mulv(R, v3f4(lp4)));
        V3 h = v3f4(D.shh[s]);
        V3 e = v3(sf_abs(Rs.c[0].x) * h.x + sf_abs(Rs.c[1].x) * h.y + sf_abs(Rs.c[2].x) * h.z,
                  sf_abs(Rs.c[0].y) * h.x + sf_abs(Rs.c[1].y) * h.y + sf_abs(Rs.c[2].y) * h.z,
                  sf_abs(Rs.c[0].z) * h.x + sf_abs(Rs.c[1].z) * h.y + sf_abs(Rs.c[2].z) * h.z);
        float m = (D.bflag[b] & 1) ? (vlen(v3f4(D.bv[b])) + vlen(v3f4(D.bw[b])) * lp4.w) * K.dt + 0.5f * K.spec : 0.5f * K.spec;
        e = vadd(e, v3(m, m, m));
        V3 lo = vsub(wp, e), hi = vadd(wp, e);
        D.swp[s] = f4(wp, 0.0f);
        store_m3(D.swr + 3 * (size_t)s, Rs);
        // (the AABB's w carries what the pair test needs: the body, and its awake / static flags)
        int flags = (D.bflag[b] & 1) | (D.bp[b].w == 0.0f ? 2 : 0);
        D.slo[s] = f4(lo, __int_as_float(b));
        D.shi[s] = f4(hi, __int_as_float(flags));
        CellRange cr = cell_range(lo, hi, K.inv_cell);
        D.ecnt[s] = cr.n;
        if (!cr.n) atomicAdd(&D.ctr[C_DROP], 1u);
    }
}

// ---------------------------------------------------------------- broadphase
__global__ void __launch_bounds__(NT) k_emit_entries(Dev D, StepK K) {
    GRID_STRIDE(s, (uint32_t)D.ns) {
        CellRange cr = cell_range(v3f4(D.slo[s]), v3f4(D.shi[s]), K.inv_cell);
        if (!cr.n) continue;
        uint32_t e = D.eoff[s];
        for (int x = cr.x0; x <= cr.x1; x++) for (int y = cr.y0; y <= cr.y1; y++) for (int z = cr.z0; z <= cr.z1; z++, e++) {
            if (e >= D.ecap) { atomicAdd(&D.ctr[C_DROP], 1u); continue; }
            uint32_t key = cell_key(x, y, z);
            D.ekey[e] = key;
            D.eshape[e] = (int32_t)s;
            atomicAdd(&D.tcnt[slot_of(key, D.tmask)], 1u);
        }
    }
}

__global__ void __launch_bounds__(NT) k_fill_table(Dev D) {
    uint32_t ne = min(D.ctr[C_NE], D.ecap);
    GRID_STRIDE(e, ne) {
        uint32_t key = D.ekey[e];
        uint32_t sl = slot_of(key, D.tmask);
        uint32_t pos = D.tstart[sl] + atomicAdd(&D.tfill[sl], 1u);
        D.tkey[pos] = key;
        D.tent[pos] = D.eshape[e];
        D.tidx[pos] = e;
    }
}
// the table's occupied slots in slot order (flags into tfill, free once the table is filled, then scanned)
__global__ void __launch_bounds__(NT) k_slot_flag(Dev D) { GRID_STRIDE(sl, D.tcap) D.tfill[sl] = D.tcnt[sl] ? 1u : 0u; }
__global__ void __launch_bounds__(NT) k_slot_list(Dev D) { GRID_STRIDE(sl, D.tcap) if (D.tcnt[sl]) D.slist[D.soff[sl]] = sl; }

// Candidates by cell entry: one thread per (shape, cell) entry finds the shape's partners whose home is that cell
// (count, then write), one thread per shape gathers its entries' counts and puts its ground candidate first, then
// one wavefront per shape sorts its partners ascending: the same canonical list as a walk of every cell per shape.
__device__ __forceinline__ int ground_cand(const Dev& D, uint32_t a) { return (__float_as_int(D.shi[a].w) & 1) && D.slo[a].y < 0.0f; }
// the cell of entry e of shape a (entries are emitted x-, y-, z-major over the shape's cell range)
__device__ __forceinline__ void entry_cell(const Dev& D, const StepK& K, uint32_t a, uint32_t e, int& x, int& y, int& z) {
    CellRange cr = cell_range(v3f4(D.slo[a]), v3f4(D.shi[a]), K.inv_cell);
    uint32_t k = e - D.eoff[a];
    uint32_t ny = (uint32_t)(cr.y1 - cr.y0 + 1), nz = (uint32_t)(cr.z1 - cr.z0 + 1);
    z = cr.z0 + (int)(k % nz); y = cr.y0 + (int)((k / nz) % ny); x = cr.x0 + (int)(k / (nz * ny));
}
// Partners b > a of shape a whose pair's home (the cell holding the larger of the two lower corners) is cell (x, y, z),
// under the pair rules of phys.rs:388-404: different bodies, not both resting, not both static, overlapping boxes.
// emit(j, b) for the j-th found (cell-table order); returns the count.
template <class Emit>
__device__ __forceinline__ uint32_t cell_partners(const Dev& D, const StepK& K, int a, int x, int y, int z, Emit emit) {
    float4 lo4 = D.slo[a], hi4 = D.shi[a];
    V3 lo = v3f4(lo4), hi = v3f4(hi4);
    int ba = __float_as_int(lo4.w), fa = __float_as_int(hi4.w);
    uint32_t key = cell_key(x, y, z), sl = slot_of(key, D.tmask);
    uint32_t s0 = D.tstart[sl], s1 = s0 + D.tcnt[sl], n = 0;
    // four table entries at a time: their keys and shapes, then the boxes of all four (from a valid index whatever
    // they hold, so the loads are not held behind the tests), then the tests and the emits in table order
    constexpr int U = 4;
    for (uint32_t q0 = s0; q0 < s1; q0 += U) {
        uint32_t kq[U];
        int bq[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool in = q0 + u < s1;
            kq[u] = in ? D.tkey[q0 + u] : ~key;
            bq[u] = in ? D.tent[q0 + u] : a;
        }
        float4 lq[U], hq[U];
#pragma unroll
        for (int u = 0; u < U; u++) { lq[u] = D.slo[bq[u]]; hq[u] = D.shi[bq[u]]; }
#pragma unroll
        for (int u = 0; u < U; u++) {
            int b = bq[u];
            if (kq[u] != key || b <= a) continue;
            int bb = __float_as_int(lq[u].w), fb = __float_as_int(hq[u].w);
            if (bb == ba) continue;
            if (!((fa | fb) & 1)) continue;          // both resting
            if ((fa & fb & 2)) continue;             // both static
            V3 lob = v3f4(lq[u]), hib = v3f4(hq[u]);
            if (lo.x > hib.x || lob.x > hi.x || lo.y > hib.y || lob.y > hi.y || lo.z > hib.z || lob.z > hi.z) continue;
            V3 m = vmaxv(lo, lob);
            if ((int32_t)sf_floor(m.x * K.inv_cell) != x || (int32_t)sf_floor(m.y * K.inv_cell) != y || (int32_t)sf_floor(m.z * K.inv_cell) != z) continue;
            emit(n, b);
            n++;
        }
    }
    return n;
}
// the count pass keeps up to ENTRY_KEEP partners of each entry, so the write pass walks the cell again only for
// entries with more
constexpr uint32_t ENTRY_KEEP = 8;
__global__ void __launch_bounds__(NT) k_cand_count(Dev D, StepK K) {
    uint32_t ne = min(D.ctr[C_NE], D.ecap);
    GRID_STRIDE(e, ne) {
        uint32_t a = (uint32_t)D.eshape[e];
        int x, y, z;
        entry_cell(D, K, a, e, x, y, z);
        int32_t* keep = D.ekeep + (size_t)e * ENTRY_KEEP;
        D.epc[e] = cell_partners(D, K, (int)a, x, y, z, [&](uint32_t j, int b) { if (j < ENTRY_KEEP) keep[j] = b; });
    }
}
// The count pass by table slot, one wavefront each: the slot's entries (a cell's, and any cell hashed to the same
// slot) go to the lanes once, and each entry in turn is broadcast to all lanes, every lane testing its own pair; a lane
// keeps its entry's partners in slot order, as cell_partners finds them. Slots of more than 64 entries take the
// per-entry walk.
__global__ void __launch_bounds__(NT) k_cand_slot(Dev D, StepK K) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    const uint32_t nslot = D.ctr[C_NSLOT];
    for (uint32_t w = wave; w < nslot; w += nwaves) {
        const uint32_t sl = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.slist[w]);  // (the wave's own slot)
        const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.tstart[sl]);
        const uint32_t n = (uint32_t)__builtin_amdgcn_readfirstlane((int)D.tcnt[sl]);
        if (n > 64) {
            for (uint32_t q = s0 + lane; q < s0 + n; q += 64) {
                uint32_t e = D.tidx[q];
                uint32_t a = (uint32_t)D.eshape[e];
                int x, y, z;
                entry_cell(D, K, a, e, x, y, z);
                int32_t* keep = D.ekeep + (size_t)e * ENTRY_KEEP;
                D.epc[e] = cell_partners(D, K, (int)a, x, y, z, [&](uint32_t j, int b) { if (j < ENTRY_KEEP) keep[j] = b; });
            }
            continue;
        }
        const bool mine = (uint32_t)lane < n;
        const uint32_t q = s0 + (mine ? (uint32_t)lane : 0u);
        const uint32_t key = D.tkey[q], e = D.tidx[q];
        const int a = D.tent[q];
        const float4 lo4 = D.slo[a], hi4 = D.shi[a];
        const V3 lo = v3f4(lo4), hi = v3f4(hi4);
        const int ba = __float_as_int(lo4.w), fa = __float_as_int(hi4.w);
        int x, y, z;
        entry_cell(D, K, (uint32_t)a, e, x, y, z);
        int32_t* keep = D.ekeep + (size_t)e * ENTRY_KEEP;
        uint32_t np = 0;
        for (uint32_t j = 0; j < n; j++) {
            const uint32_t kj = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)j);
            const int b = __builtin_amdgcn_readlane(a, (int)j);
            const float lbx = rdl(lo.x, (int)j), lby = rdl(lo.y, (int)j), lbz = rdl(lo.z, (int)j);
            const float hbx = rdl(hi.x, (int)j), hby = rdl(hi.y, (int)j), hbz = rdl(hi.z, (int)j);
            const int bb = __builtin_amdgcn_readlane(ba, (int)j), fb = __builtin_amdgcn_readlane(fa, (int)j);
            if (!mine || kj != key || b <= a) continue;
            if (bb == ba) continue;
            if (!((fa | fb) & 1)) continue;          // both resting
            if ((fa & fb & 2)) continue;             // both static
            if (lo.x > hbx || lbx > hi.x || lo.y > hby || lby > hi.y || lo.z > hbz || lbz > hi.z) continue;
            V3 m = vmaxv(lo, v3(lbx, lby, lbz));
            if ((int32_t)sf_floor(m.x * K.inv_cell) != x || (int32_t)sf_floor(m.y * K.inv_cell) != y || (int32_t)sf_floor(m.z * K.inv_cell) != z) continue;
            if (np < ENTRY_KEEP) keep[np] = b;
            np++;
        }
        if (mine) D.epc[e] = np;
    }
}

// per shape: its candidate count (ground first) and each entry's offset within the shape's list
__global__ void __launch_bounds__(NT) k_cand_shape(Dev D) {
    uint32_t ng = 0, np = 0;
    GRID_STRIDE(a, (uint32_t)D.ns) {
        uint32_t g = (uint32_t)ground_cand(D, a), o = g, e0 = D.eoff[a], n = D.ecnt[a];
        uint32_t ne = min(D.ctr[C_NE], D.ecap);
        for (uint32_t k = 0; k < n; k++) {
            uint32_t e = e0 + k;
            if (e >= ne) break;
            uint32_t c = D.epc[e];
            D.erel[e] = o;
            o += c;
        }
        D.ccnt[a] = o;
        ng += g;
        np += o - g;
    }
    block_count(ng, &D.ctr[C_NGROUND]);
    block_count(np, &D.ctr[C_NPAIR]);
}
__global__ void __launch_bounds__(NT) k_cand_write(Dev D, StepK K) {
    uint32_t ne = min(D.ctr[C_NE], D.ecap);
    GRID_STRIDE(e, ne) {
        uint32_t a = (uint32_t)D.eshape[e];
        if (!D.epc[e]) continue;
        if (D.coff[a] + D.ccnt[a] > D.ccap) { atomicAdd(&D.ctr[C_DROP], D.epc[e]); continue; }  // the shape's list is dropped whole
        uint32_t o = D.coff[a] + D.erel[e];
        int2* out = D.cand + o;
        const uint32_t n = D.epc[e];
        if (n <= ENTRY_KEEP) {
            const int32_t* keep = D.ekeep + (size_t)e * ENTRY_KEEP;
            for (uint32_t j = 0; j < n; j++) out[j] = make_int2((int)a, keep[j]);
        } else {
            int x, y, z;
            entry_cell(D, K, a, e, x, y, z);
            cell_partners(D, K, (int)a, x, y, z, [&](uint32_t j, int b) { out[j] = make_int2((int)a, b); });
        }
    }
}
// per shape: the ground candidate first, the partners ascending after it. Up to 16 partners: one thread, a bitonic
// network in registers; more: the shape goes to a list sorted by k_cand_finish_big, a wavefront per shape.
__device__ __forceinline__ void cas_i(int& x, int& y) { int lo = min(x, y), hi = max(x, y); x = lo; y = hi; }
__global__ void __launch_bounds__(NT) k_cand_finish(Dev D) {
    GRID_STRIDE(a, (uint32_t)D.ns) {
        uint32_t o = D.coff[a], n = D.ccnt[a];
        uint32_t g = (uint32_t)ground_cand(D, a);
        if (!n) continue;
        if (o + n > D.ccap) { if (g) atomicAdd(&D.ctr[C_DROP], 1u); continue; }
        if (g) D.cand[o] = make_int2((int)a, -1);
        o += g; n -= g;
        if (n < 2) continue;
        if (n > 16) { D.bigc[atomicAdd(&D.ctr[C_NBIGC], 1u)] = a; continue; }
        int v[16];
#pragma unroll
        for (int k = 0; k < 16; k++) v[k] = (uint32_t)k < n ? D.cand[o + k].y : 0x7FFFFFFF;
#pragma unroll
        for (int k = 2; k <= 16; k <<= 1)
#pragma unroll
            for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    int l = i ^ j;
                    if (l > i) { if ((i & k) == 0) cas_i(v[i], v[l]); else cas_i(v[l], v[i]); }
                }
#pragma unroll
        for (int k = 0; k < 16; k++) if ((uint32_t)k < n) D.cand[o + k].y = v[k];
    }
}
// the shapes with more than 16 partners: one workgroup each, a bitonic sort in LDS (lists past 4,096 entries, which
// no sane scene has, by insertion)
__global__ void __launch_bounds__(NT) k_cand_finish_big(Dev D) {
    __shared__ int sh[4096];
    const uint32_t nbig = min(D.ctr[C_NBIGC], (uint32_t)D.ns);
    for (uint32_t t = blockIdx.x; t < nbig; t += gridDim.x) {
        const uint32_t a = D.bigc[t];
        const uint32_t g = (uint32_t)ground_cand(D, a);
        const uint32_t o = D.coff[a] + g, n = D.ccnt[a] - g;
        if (n > 4096) {
            if (threadIdx.x == 0) {
                for (uint32_t i = 1; i < n; i++) {
                    int2 v = D.cand[o + i];
                    uint32_t j = i;
                    while (j > 0 && D.cand[o + j - 1].y > v.y) { D.cand[o + j] = D.cand[o + j - 1]; j--; }
                    D.cand[o + j] = v;
                }
            }
            continue;
        }
        uint32_t n2 = 32;
        while (n2 < n) n2 <<= 1;
        for (uint32_t i = threadIdx.x; i < n2; i += NT) sh[i] = i < n ? D.cand[o + i].y : 0x7FFFFFFF;
        __syncthreads();
        for (uint32_t k = 2; k <= n2; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < n2; i += NT) {
                    uint32_t l = i ^ j;
                    if (l > i) {
                        int x = sh[i], y = sh[l];
                        bool up = (i & k) == 0;
                        if ((x > y) == up) { sh[i] = y; sh[l] = x; }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t i = threadIdx.x; i < n; i += NT) D.cand[o + i].y = sh[i];
        __syncthreads();
    }
}

// ---------------------------------------------------------------- narrow phase
struct Pt { V3 p; float s; };

__device__ __forceinline__ void tangents(V3 n, V3& t1, V3& t2) {
    V3 t = sf_abs(n.x) > 0.57f ? v3(n.y, -n.x, 0.0f) : v3(0.0f, n.z, -n.y);
    t1 = vnorm(t);
    t2 = vcross(n, t1);
}

// Eight-slot arrays kept in registers: every loop over them runs to eight with a guard and every access at a
// computed position is an unrolled select (a dynamically indexed private array lives in scratch memory).
// (bitwise picks: a select chain over an array is folded back into an indexed access, which needs the array in memory)
__device__ __forceinline__ float pick(bool c, float a, float b) {
    const int m = -(int)c;
    return __int_as_float((__float_as_int(a) & m) | (__float_as_int(b) & ~m));
}
__device__ __forceinline__ V3 pick(bool c, V3 a, V3 b) { return v3(pick(c, a.x, b.x), pick(c, a.y, b.y), pick(c, a.z, b.z)); }
__device__ __forceinline__ Pt pick(bool c, Pt a, Pt b) { Pt r; r.p = pick(c, a.p, b.p); r.s = pick(c, a.s, b.s); return r; }
template <class T> __device__ __forceinline__ void put8(T (&a)[8], int n, const T& v) {
#pragma unroll
    for (int j = 0; j < 8; j++) a[j] = pick(j == n, v, a[j]);
}
template <class T> __device__ __forceinline__ T get8(const T (&a)[8], int n) {
    T r = a[0];
#pragma unroll
    for (int j = 1; j < 8; j++) r = pick(j == n, a[j], r);
    return r;
}
__device__ __forceinline__ float sel3f(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
__device__ __forceinline__ V3 row3(M3 m, int i) {
    V3 a = m.c[0], b = m.c[1], c = m.c[2];
    return v3(i == 0 ? a.x : (i == 1 ? b.x : c.x), i == 0 ? a.y : (i == 1 ? b.y : c.y), i == 0 ? a.z : (i == 1 ? b.z : c.z));
}

// keep the deepest point, the farthest from it and the two spanning the most area (phys.rs:507-527)
__device__ __forceinline__ int reduce4(Pt (&pts)[8], int cnt, V3 n) {
    if (cnt <= 4) return cnt;
    int i0 = 0;
#pragma unroll
    for (int k = 1; k < 8; k++) if (k < cnt && pts[k].s < get8(pts, i0).s) i0 = k;
    V3 p0 = get8(pts, i0).p;
    int i1 = 0; float best = -1.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) { if (k >= cnt) continue; float d = vlen2(vsub(pts[k].p, p0)); if (d > best) { best = d; i1 = k; } }
    V3 e = vsub(get8(pts, i1).p, p0);
    int i2 = 0; float b2 = -3.40282347e38f;
#pragma unroll
    for (int k = 0; k < 8; k++) { if (k >= cnt) continue; float a = vdot(vcross(vsub(pts[k].p, p0), e), n); if (a > b2) { b2 = a; i2 = k; } }
    int i3 = 0; float b3 = -3.40282347e38f;
#pragma unroll
    for (int k = 0; k < 8; k++) { if (k >= cnt) continue; float a = -vdot(vcross(vsub(pts[k].p, p0), e), n); if (a > b3) { b3 = a; i3 = k; } }
    const int idx[4] = {i0, i1, i2, i3};
    const Pt sel[4] = {get8(pts, i0), get8(pts, i1), get8(pts, i2), get8(pts, i3)};
    int out = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        bool dup = false;
#pragma unroll
        for (int q = 0; q < k; q++) dup |= idx[q] == idx[k];
        if (!dup) { put8(pts, out, sel[k]); out++; }
    }
    return out;
}

// 15-axis SAT; face contact clips the incident face against the reference face's side planes; edge contact gives
// the closest points of the two edges (phys.rs:531-628). Returns the point count, normal from A to B.
__device__ __forceinline__ int collide_boxes(V3 pa, M3 ra, V3 ha, V3 pb, M3 rb, V3 hb, float margin, Pt (&out)[8], V3& n_out) {
    V3 t = vsub(pb, pa);
    float ac[3][3];
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) ac[i][j] = sf_abs(vdot(ra.c[i], rb.c[j])) + 1e-6f;
    float hav[3] = {ha.x, ha.y, ha.z}, hbv[3] = {hb.x, hb.y, hb.z};
    float sa = -3.40282347e38f; int ia = 0;
    for (int i = 0; i < 3; i++) {
        float s = sf_abs(vdot(t, ra.c[i])) - (hav[i] + hbv[0] * ac[i][0] + hbv[1] * ac[i][1] + hbv[2] * ac[i][2]);
        if (s > margin) return 0;
        if (s > sa) { sa = s; ia = i; }
    }
    float sbm = -3.40282347e38f; int jb = 0;
    for (int j = 0; j < 3; j++) {
        float s = sf_abs(vdot(t, rb.c[j])) - (hbv[j] + hav[0] * ac[0][j] + hav[1] * ac[1][j] + hav[2] * ac[2][j]);
        if (s > margin) return 0;
        if (s > sbm) { sbm = s; jb = j; }
    }
    float se = -3.40282347e38f; int ei = 0, ej = 0; V3 el = v3(0, 0, 0);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) {
        V3 l = vcross(ra.c[i], rb.c[j]);
        float len2 = vlen2(l);
        if (len2 < 1e-8f) continue;
        l = vmul(l, 1.0f / sf_sqrt(len2));
        float r = hav[0] * sf_abs(vdot(ra.c[0], l)) + hav[1] * sf_abs(vdot(ra.c[1], l)) + hav[2] * sf_abs(vdot(ra.c[2], l))
                + hbv[0] * sf_abs(vdot(rb.c[0], l)) + hbv[1] * sf_abs(vdot(rb.c[1], l)) + hbv[2] * sf_abs(vdot(rb.c[2], l));
        float s = sf_abs(vdot(t, l)) - r;
        if (s > margin) return 0;
        if (s > se) { se = s; ei = i; ej = j; el = l; }
    }
    const float face_tol = 0.005f;
    bool use_b = sbm > sa + face_tol;
    float face_best = use_b ? sbm : sa;
    if (se > face_best + 0.01f) {
        V3 n = vdot(t, el) < 0.0f ? vneg(el) : el;
        V3 ca = pa, cb = pb;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (k != ei) { float s = vdot(ra.c[k], n) > 0.0f ? 1.0f : -1.0f; ca = vadd(ca, vmul(ra.c[k], s * hav[k])); }
            if (k != ej) { float s = vdot(rb.c[k], n) > 0.0f ? -1.0f : 1.0f; cb = vadd(cb, vmul(rb.c[k], s * hbv[k])); }
        }
        V3 d1 = row3(ra, ei), d2 = row3(rb, ej);
        const float hae = sel3f(ha, ei), hbe = sel3f(hb, ej);
        V3 r = vsub(ca, cb);
        float bb = vdot(d1, d2), c = vdot(d1, r), f = vdot(d2, r);
        float den = 1.0f - bb * bb;
        float s = den > 1e-6f ? (bb * f - c) / den : 0.0f;
        s = sf_min(sf_max(s, -hae), hae);
        float tt = bb * s + f;
        tt = sf_min(sf_max(tt, -hbe), hbe);
        s = sf_min(sf_max(bb * tt - c, -hae), hae);
        V3 p1 = vadd(ca, vmul(d1, s)), p2 = vadd(cb, vmul(d2, tt));
        out[0].p = vmul(vadd(p1, p2), 0.5f);
        out[0].s = vdot(vsub(p2, p1), n);
        n_out = n;
        return 1;
    }
    V3 pr = use_b ? pb : pa, pi = use_b ? pa : pb;
    M3 rr, ri;
#pragma unroll
    for (int q = 0; q < 3; q++) { rr.c[q] = use_b ? rb.c[q] : ra.c[q]; ri.c[q] = use_b ? ra.c[q] : rb.c[q]; }
    const V3 hr = use_b ? hb : ha, hi = use_b ? ha : hb;
    int k = use_b ? jb : ia;
    V3 n = row3(rr, k);
    if (vdot(vsub(pi, pr), n) < 0.0f) n = vneg(n);
    int j = 0; float bd = -1.0f;
#pragma unroll
    for (int q = 0; q < 3; q++) { float d = sf_abs(vdot(ri.c[q], n)); if (d > bd) { bd = d; j = q; } }
    const V3 rj = row3(ri, j);
    float sgn = vdot(rj, n) > 0.0f ? -1.0f : 1.0f;
    V3 fc = vadd(pi, vmul(rj, sgn * sel3f(hi, j)));
    int u = (j + 1) % 3, v = (j + 2) % 3;
    V3 eu = vmul(row3(ri, u), sel3f(hi, u)), ev = vmul(row3(ri, v), sel3f(hi, v));
    V3 poly[8], tmp[8];
#pragma unroll
    for (int q = 4; q < 8; q++) { poly[q] = v3(0, 0, 0); tmp[q] = v3(0, 0, 0); }
    poly[0] = vadd(vadd(fc, eu), ev); poly[1] = vadd(vsub(fc, eu), ev); poly[2] = vsub(vsub(fc, eu), ev); poly[3] = vsub(vadd(fc, eu), ev);
#pragma unroll
    for (int q = 0; q < 4; q++) tmp[q] = v3(0, 0, 0);
    int pn = 4;
    int u2 = (k + 1) % 3, v2 = (k + 2) % 3;
#pragma unroll
    for (int pl = 0; pl < 4; pl++) {
        int ax = pl < 2 ? u2 : v2;
        float sg = (pl & 1) ? -1.0f : 1.0f;
        V3 dir = vmul(row3(rr, ax), sg);
        float off = vdot(dir, pr) + sel3f(hr, ax);
        int tn = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (q >= pn) continue;
            V3 p0 = poly[q], p1 = (q + 1 < pn && q < 7) ? poly[q < 7 ? q + 1 : 0] : poly[0];
            float d0 = vdot(p0, dir) - off, d1 = vdot(p1, dir) - off;
            if (d0 <= 0.0f) { if (tn < 8) { put8(tmp, tn, p0); tn++; } }
            if ((d0 <= 0.0f) != (d1 <= 0.0f)) {
                float tt = d0 / (d0 - d1);
                if (tn < 8) { put8(tmp, tn, vadd(p0, vmul(vsub(p1, p0), tt))); tn++; }
            }
        }
        pn = tn;
#pragma unroll
        for (int q = 0; q < 8; q++) poly[q] = tmp[q];
        if (pn == 0) return 0;
    }
    V3 rc = vadd(pr, vmul(n, sel3f(hr, k)));
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (q >= pn) continue;
        float sep = vdot(vsub(poly[q], rc), n);
        if (sep <= margin) { Pt t; t.p = vsub(poly[q], vmul(n, 0.5f * sep)); t.s = sep; put8(out, cnt, t); cnt++; }
    }
    if (cnt == 0) return 0;
    n_out = use_b ? vneg(n) : n;
    return cnt;
}

// (with chunks in the world, candidates with one go to the list k_narrow_hull takes; their outputs come from there)
__global__ void __launch_bounds__(NT) k_narrow(Dev D, StepK K, int any_hull) {
    uint32_t nc = min(D.ctr[C_NCAND], D.ccap);
    GRID_STRIDE(i, nc) {
        int2 c = D.cand[i];
        if (any_hull && (D.shull[c.x] || (c.y >= 0 && D.shull[c.y]))) { D.hlist[atomicAdd(&D.ctr[C_NHULL], 1u)] = i; continue; }
        Pt pts[8];
        V3 n = v3(0, 0, 0);
        int cnt = 0;
        int ba = D.shb[c.x];
        int slot_b = 0;
        if (c.y < 0) {
            // ground: the shape's corners under the speculative margin (phys.rs:418-432)
            V3 wp = v3f4(D.swp[c.x]);
            M3 R = load_m3(D.swr + 3 * (size_t)c.x);
            V3 h = v3f4(D.shh[c.x]);
            float margin = K.spec + sf_max(-D.bv[ba].y, 0.0f) * (1.0f / 60.0f);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                V3 sg = v3((k & 1) ? 1.0f : -1.0f, (k & 2) ? 1.0f : -1.0f, (k & 4) ? 1.0f : -1.0f);
                V3 cp = vadd(wp, mulv(R, v3(sg.x * h.x, sg.y * h.y, sg.z * h.z)));
                if (cp.y < margin) { Pt t; t.p = v3(cp.x, 0.5f * cp.y, cp.z); t.s = cp.y; put8(pts, cnt, t); cnt++; }
            }
            n = v3(0.0f, -1.0f, 0.0f);
        } else {
            int bb = D.shb[c.y];
            slot_b = D.bslot[bb];
            V3 va = v3f4(D.bv[ba]), vb = v3f4(D.bv[bb]);
            float rel = vlen(vsub(vb, va)) + vlen(v3f4(D.bw[ba])) * D.slp[c.x].w + vlen(v3f4(D.bw[bb])) * D.slp[c.y].w;
            float margin = K.spec + rel * (1.0f / 60.0f);
            M3 Ra = load_m3(D.swr + 3 * (size_t)c.x), Rb = load_m3(D.swr + 3 * (size_t)c.y);
            cnt = collide_boxes(v3f4(D.swp[c.x]), Ra, v3f4(D.shh[c.x]), v3f4(D.swp[c.y]), Rb, v3f4(D.shh[c.y]), margin, pts, n);
            if (cnt) {
                // wakes from the step-start flags: a moving awake body wakes a resting one it touches
                int fa = D.bflag[ba], fb = D.bflag[bb];
                bool a_awake = fa & 1, b_awake = fb & 1;
                bool a_still = D.bii[ba].w != 0.0f, b_still = D.bii[bb].w != 0.0f;
                if (a_awake && !b_awake && !a_still) D.wake[bb] = 1u;
                if (b_awake && !a_awake && !b_still) D.wake[ba] = 1u;
            }
        }
        if (cnt) cnt = reduce4(pts, cnt, n);
        bool hit = cnt > 0 && !(D.bslot[ba] == 0 && slot_b == 0);
        D.hit[i] = hit ? 1u : 0u;
        float4* o = D.nar + 5 * (size_t)i;
        o[0] = f4(n, (float)cnt);
#pragma unroll
        for (int k = 0; k < 4; k++) if (k < cnt) o[1 + k] = f4(pts[k].p, pts[k].s);  // (at most 4 after reduce4)
    }
}


// ---------------------------------------------------------------- convex chunks (gpu/gpu_world.js:402-530)
// A shape with hull t > 0 is cell t of the chunk library in its box, stretched by the shape's half size: vertex k in
// the world is p + R (h * v_k); a face's world plane is R (n / h) normalised, offset by its d and the position. A box
// (t = 0) paired with a chunk takes the same path as the whole-cube cell. Per-lane arrays (corners, directions, clip
// polygons) are indexed at run time: this path serves the chunk pairs only, listed apart from the box pairs.
struct HS { V3 p; M3 R; V3 h; int t; };
__device__ __forceinline__ V3 hull_w(const HullDev& H, const HS& s, int k) {
    const V3 v = v3f4(H.v[s.t * sf_hull_maxv + k]);
    return vadd(s.p, mulv(s.R, v3(s.h.x * v.x, s.h.y * v.y, s.h.z * v.z)));
}
__device__ __forceinline__ float4 hull_face(const HullDev& H, const HS& s, int f) {
    const float4 pl = H.pl[s.t * sf_hull_maxf + f];
    const V3 nw = mulv(s.R, v3(pl.x / s.h.x, pl.y / s.h.y, pl.z / s.h.z));
    const float l = vlen(nw);
    const V3 n = vmul(nw, 1.0f / l);
    return f4(n, pl.w / l + vdot(n, s.p));
}
__device__ __forceinline__ int hull_fv(const HullDev& H, const HS& s, int f, int i) { return H.fidx[(s.t * sf_hull_maxf + f) * sf_hull_maxfv + i]; }
__device__ __forceinline__ int hull_fn(const HullDev& H, const HS& s, int f) { return H.fcnt[s.t * sf_hull_maxf + f]; }
__device__ __forceinline__ V3 hull_dir(const HullDev& H, const HS& s, int k) {
    const V3 d = v3f4(H.dir[s.t * sf_hull_maxd + k]);
    return vnorm(mulv(s.R, v3(s.h.x * d.x, s.h.y * d.y, s.h.z * d.z)));
}
// keep up to 8 points; past 8, a deeper point replaces the shallowest
__device__ __forceinline__ void keep8(Pt* pts, int& cnt, Pt t) {
    if (cnt < 8) { pts[cnt++] = t; return; }
    int hi = 0;
    for (int q = 1; q < 8; q++) if (pts[q].s > pts[hi].s) hi = q;
    if (t.s < pts[hi].s) pts[hi] = t;
}
// the chunk's corners below the margin (normal from the chunk into the ground)
__device__ int hull_ground(const HullDev& H, const HS& s, float margin, Pt* pts) {
    const int nv = H.head[s.t].x;
    int cnt = 0;
    for (int k = 0; k < nv; k++) {
        const V3 x = hull_w(H, s, k);
        if (!(x.y < margin)) continue;
        Pt t; t.p = v3(x.x, 0.5f * x.y, x.z); t.s = x.y;
        keep8(pts, cnt, t);
    }
    return cnt;
}
// Separating axes: both hulls' face normals, then the cross products of their edge directions (the sign that gives
// the larger separation). The best face (B's only if clearly better) gives a reference face whose side planes clip
// the other hull's face most against it; an edge axis clearly better than both gives the closest points of the two
// support edges. Normal from A to B; a point's separation is negative in overlap.
__device__ int hull_hull(const HullDev& H, const HS& A, const HS& B, float margin, Pt* pts, V3& n_out) {
    const int4 ha = H.head[A.t], hb = H.head[B.t];
    V3 VA[sf_hull_maxv], VB[sf_hull_maxv], DA[sf_hull_maxd], DB[sf_hull_maxd];
    for (int k = 0; k < ha.x; k++) VA[k] = hull_w(H, A, k);
    for (int k = 0; k < hb.x; k++) VB[k] = hull_w(H, B, k);
    for (int k = 0; k < ha.w; k++) DA[k] = hull_dir(H, A, k);
    for (int k = 0; k < hb.w; k++) DB[k] = hull_dir(H, B, k);
    float sa = -3.40282347e38f, sb = -3.40282347e38f, se = -3.40282347e38f;
    int fa = 0, fb = 0, ei = 0, ej = 0;
    V3 el = v3(0, 0, 0);
    for (int f = 0; f < ha.y; f++) {
        const float4 pl = hull_face(H, A, f);
        float lo = 3.40282347e38f;
        for (int k = 0; k < hb.x; k++) lo = sf_min(lo, vdot(VB[k], v3f4(pl)));
        const float sp = lo - pl.w;
        if (sp > margin) return 0;
        if (sp > sa) { sa = sp; fa = f; }
    }
    for (int f = 0; f < hb.y; f++) {
        const float4 pl = hull_face(H, B, f);
        float lo = 3.40282347e38f;
        for (int k = 0; k < ha.x; k++) lo = sf_min(lo, vdot(VA[k], v3f4(pl)));
        const float sp = lo - pl.w;
        if (sp > margin) return 0;
        if (sp > sb) { sb = sp; fb = f; }
    }
    for (int i = 0; i < ha.w; i++)
        for (int j = 0; j < hb.w; j++) {
            const V3 l0 = vcross(DA[i], DB[j]);
            const float l2 = vdot(l0, l0);
            if (l2 < 1e-6f) continue;
            V3 l = vmul(l0, 1.0f / sf_sqrt(l2));
            float alo = 3.40282347e38f, ahi = -3.40282347e38f, blo = 3.40282347e38f, bhi = -3.40282347e38f;
            for (int k = 0; k < ha.x; k++) { const float d = vdot(VA[k], l); alo = sf_min(alo, d); ahi = sf_max(ahi, d); }
            for (int k = 0; k < hb.x; k++) { const float d = vdot(VB[k], l); blo = sf_min(blo, d); bhi = sf_max(bhi, d); }
            const float s1 = blo - ahi, s2 = alo - bhi;
            float sp = s1;
            if (s2 > s1) { sp = s2; l = vneg(l); }
            if (sp > margin) return 0;
            if (sp > se) { se = sp; el = l; ei = i; ej = j; }
        }
    const bool use_b = sb > sa + 0.005f;
    const float fbest = use_b ? sb : sa;
    if (se > fbest + 0.01f) {
        // edge against edge: each hull's edge along the axis's direction furthest out (A) / in (B) along the axis
        const V3 da = DA[ei], db = DB[ej];
        V3 a0 = A.p, a1 = A.p, b0 = B.p, b1 = B.p;
        float best = -3.40282347e38f;
        for (int k = 0; k < ha.z; k++) {
            const int x = H.seg[(A.t * sf_hull_maxs + k) * 2], y = H.seg[(A.t * sf_hull_maxs + k) * 2 + 1];
            const V3 p0 = VA[x], p1 = VA[y];
            if (sf_abs(vdot(vnorm(vsub(p1, p0)), da)) < 0.999f) continue;
            const float m = vdot(vadd(p0, p1), el);
            if (m > best) { best = m; a0 = p0; a1 = p1; }
        }
        best = 3.40282347e38f;
        for (int k = 0; k < hb.z; k++) {
            const int x = H.seg[(B.t * sf_hull_maxs + k) * 2], y = H.seg[(B.t * sf_hull_maxs + k) * 2 + 1];
            const V3 p0 = VB[x], p1 = VB[y];
            if (sf_abs(vdot(vnorm(vsub(p1, p0)), db)) < 0.999f) continue;
            const float m = vdot(vadd(p0, p1), el);
            if (m < best) { best = m; b0 = p0; b1 = p1; }
        }
        const V3 d1 = vsub(a1, a0), d2 = vsub(b1, b0), r = vsub(a0, b0);
        const float aa = vdot(d1, d1), ee = vdot(d2, d2), ff = vdot(d2, r), cc = vdot(d1, r), bb = vdot(d1, d2);
        const float den = aa * ee - bb * bb;
        float s1 = 0.0f;
        if (den > 1e-9f) s1 = sf_min(sf_max((bb * ff - cc * ee) / den, 0.0f), 1.0f);
        float t1 = (bb * s1 + ff) / sf_max(ee, 1e-9f);
        if (t1 < 0.0f) { t1 = 0.0f; s1 = sf_min(sf_max(-cc / sf_max(aa, 1e-9f), 0.0f), 1.0f); }
        else if (t1 > 1.0f) { t1 = 1.0f; s1 = sf_min(sf_max((bb - cc) / sf_max(aa, 1e-9f), 0.0f), 1.0f); }
        const V3 p1 = vadd(a0, vmul(d1, s1)), p2 = vadd(b0, vmul(d2, t1));
        pts[0].p = vmul(vadd(p1, p2), 0.5f);
        pts[0].s = vdot(vsub(p2, p1), el);
        n_out = el;
        return 1;
    }
    // a face: reference Rs (the hull whose face it is), incident Is (the other's face most against its normal)
    const HS& Rs = use_b ? B : A;
    const HS& Is = use_b ? A : B;
    const int fr = use_b ? fb : fa, nfi = use_b ? ha.y : hb.y;
    const float4 rpl = hull_face(H, Rs, fr);
    const V3 n = v3f4(rpl);
    int fi = 0;
    float bd = 3.40282347e38f;
    for (int f = 0; f < nfi; f++) { const float d = vdot(v3f4(hull_face(H, Is, f)), n); if (d < bd) { bd = d; fi = f; } }
    V3 poly[16], tmp[16];
    int pn = hull_fn(H, Is, fi);
    for (int q = 0; q < pn; q++) poly[q] = hull_w(H, Is, hull_fv(H, Is, fi, q));
    const int rn = hull_fn(H, Rs, fr);
    V3 rc = v3(0, 0, 0);
    for (int q = 0; q < rn; q++) rc = vadd(rc, hull_w(H, Rs, hull_fv(H, Rs, fr, q)));
    rc = vmul(rc, 1.0f / (float)rn);
    for (int e = 0; e < rn; e++) {
        const V3 v0 = hull_w(H, Rs, hull_fv(H, Rs, fr, e)), v1 = hull_w(H, Rs, hull_fv(H, Rs, fr, (e + 1) % rn));
        V3 sn = vnorm(vcross(vsub(v1, v0), n));
        if (vdot(sn, vsub(rc, v0)) > 0.0f) sn = vneg(sn);
        int tn = 0;
        for (int q = 0; q < pn; q++) {
            const V3 p0 = poly[q], p1 = poly[(q + 1) % pn];
            const float d0 = vdot(vsub(p0, v0), sn), d1 = vdot(vsub(p1, v0), sn);
            if (d0 <= 0.0f && tn < 16) tmp[tn++] = p0;
            if ((d0 <= 0.0f) != (d1 <= 0.0f) && tn < 16) tmp[tn++] = vadd(p0, vmul(vsub(p1, p0), d0 / (d0 - d1)));
        }
        pn = tn;
        for (int q = 0; q < pn; q++) poly[q] = tmp[q];
        if (pn == 0) return 0;
    }
    int cnt = 0;
    for (int q = 0; q < pn; q++) {
        const float sep = vdot(poly[q], n) - rpl.w;
        if (sep > margin) continue;
        Pt t; t.p = vsub(poly[q], vmul(n, 0.5f * sep)); t.s = sep;
        keep8(pts, cnt, t);
    }
    n_out = use_b ? vneg(n) : n;
    return cnt;
}

// the listed chunk candidates, one lane each: the box narrow phase's margins, wakes and reduce4, chunk geometry
__global__ void __launch_bounds__(NT) k_narrow_hull(Dev D, StepK K) {
    const uint32_t nl = min(D.ctr[C_NHULL], D.ccap);
    GRID_STRIDE(li, nl) {
        const uint32_t i = D.hlist[li];
        const int2 c = D.cand[i];
        Pt pts[8];
        V3 n = v3(0, 0, 0);
        int cnt = 0;
        const int ba = D.shb[c.x];
        int slot_b = 0;
        HS A;
        A.p = v3f4(D.swp[c.x]); A.R = load_m3(D.swr + 3 * (size_t)c.x); A.h = v3f4(D.shh[c.x]); A.t = D.shull[c.x];
        if (c.y < 0) {
            // (a chunk's margin also covers its corners' sweep from spin, gpu/gpu_world.js ground_margin :776-778)
            const float margin = K.spec + (sf_max(-D.bv[ba].y, 0.0f) + vlen(v3f4(D.bw[ba])) * D.slp[c.x].w) * (1.0f / 60.0f);
            cnt = hull_ground(D.hl, A, margin, pts);
            n = v3(0.0f, -1.0f, 0.0f);
            // up to 8 corners as two manifolds of 4 in the order found (gpu/gpu_world.js narrow_hull :813-823: one
            // patch of 4 left the other corners free to swing into the ground); no reduce4
            const bool hit = cnt > 0 && D.bslot[ba] != 0;
            D.hit[i] = hit ? (cnt > 4 ? 2u : 1u) : 0u;
            float4* o = D.nar + 5 * (size_t)i;
            o[0] = f4(n, (float)min(cnt, 4));
            for (int k = 0; k < 4; k++) if (k < cnt) o[1 + k] = f4(pts[k].p, pts[k].s);
            if (cnt > 4) {
                float4* o2 = D.nar2 + 5 * (size_t)c.x;
                o2[0] = f4(n, (float)(cnt - 4));
                for (int k = 4; k < 8; k++) if (k < cnt) o2[k - 3] = f4(pts[k].p, pts[k].s);
            }
            continue;
        } else {
            const int bb = D.shb[c.y];
            slot_b = D.bslot[bb];
            const V3 va = v3f4(D.bv[ba]), vb = v3f4(D.bv[bb]);
            const float rel = vlen(vsub(vb, va)) + vlen(v3f4(D.bw[ba])) * D.slp[c.x].w + vlen(v3f4(D.bw[bb])) * D.slp[c.y].w;
            const float margin = K.spec + rel * (1.0f / 60.0f);
            HS B;
            B.p = v3f4(D.swp[c.y]); B.R = load_m3(D.swr + 3 * (size_t)c.y); B.h = v3f4(D.shh[c.y]); B.t = D.shull[c.y];
            cnt = hull_hull(D.hl, A, B, margin, pts, n);
            if (cnt) {
                const int fa = D.bflag[ba], fb = D.bflag[bb];
                const bool a_awake = fa & 1, b_awake = fb & 1;
                const bool a_still = D.bii[ba].w != 0.0f, b_still = D.bii[bb].w != 0.0f;
                if (a_awake && !b_awake && !a_still) D.wake[bb] = 1u;
                if (b_awake && !a_awake && !b_still) D.wake[ba] = 1u;
            }
        }
        if (cnt) cnt = reduce4(pts, cnt, n);
        const bool hit = cnt > 0 && !(D.bslot[ba] == 0 && slot_b == 0);
        D.hit[i] = hit ? 1u : 0u;
        float4* o = D.nar + 5 * (size_t)i;
        o[0] = f4(n, (float)cnt);
        for (int k = 0; k < 4; k++) if (k < cnt) o[1 + k] = f4(pts[k].p, pts[k].s);
    }
}

__global__ void __launch_bounds__(NT) k_apply_wake(Dev D) {
    GRID_STRIDE(i, (uint32_t)D.nb) {
        if (!D.wake[i]) continue;
        if ((D.bflag[i] & 1) || D.bp[i].w == 0.0f) continue;
        D.bflag[i] |= 1;
        float4 ii = D.bii[i]; ii.w = 0.0f; D.bii[i] = ii;
        D.bsp[i] = D.bp[i];
        D.bsq[i] = D.bq[i];
        atomicAdd(&D.ctr[C_NWAKE], 1u);
    }
}

// the j-th manifold of candidate i (hit[i] of them): its normal|count and 4 points (j = 1: a chunk's second patch of
// ground corners, kept by shape)
__device__ __forceinline__ const float4* cand_pts(const Dev& D, uint32_t i, uint32_t j) {
    return j ? D.nar2 + 5 * (size_t)D.cand[i].x : D.nar + 5 * (size_t)i;
}

__device__ int find_old(const Dev& D, uint64_t key, uint32_t nold) {
    int lo = 0, hi = (int)nold - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        uint64_t k = D.old.key[mid];
        if (k == key) return mid;
        if (k < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

__device__ __forceinline__ float mix_mu(float a, float b) { return a == b ? a : sf_sqrt(a * b); }

// The warm-start record of one side of a manifold: the vectors the warm start adds to that side's body, w += t_k
// (k < count) then v += tv, w += tf (signs by side), formed from the accumulated impulses with the true masses,
// exactly as phys.rs:245-262 forms them. It goes to the body's contact-list entry (8 float4 each), so a body's warm
// start reads its records in one contiguous run.
__device__ __forceinline__ void write_warm_side(float4* w, int side, int cnt, float4 n4, float4 t14, float4 t24,
                                                float j1, float j2, float jt, const float* jn, const V3* ixn, V3 i1,
                                                V3 i2, V3 in, float im) {
    float pn = 0.0f;
    V3 t[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k < cnt) { pn += jn[k]; t[k] = vmul(ixn[k], jn[k]); }
        else t[k] = v3(0, 0, 0);
    }
    V3 pp = vadd(vadd(vmul(v3f4(n4), pn), vmul(v3f4(t14), j1)), vmul(v3f4(t24), j2));
    V3 tf = vadd(vadd(vmul(i1, j1), vmul(i2, j2)), vmul(in, jt));
    // the body's spin change is the point terms then tf, added in point order; side a subtracts (its terms negated,
    // which is exact)
    auto sg = [&](V3 x) { return side ? x : vneg(x); };
    V3 dw = sg(t[0]);
    if (cnt > 1) dw = vadd(dw, sg(t[1]));
    if (cnt > 2) dw = vadd(dw, sg(t[2]));
    if (cnt > 3) dw = vadd(dw, sg(t[3]));
    dw = vadd(dw, sg(tf));
    w[0] = f4(sg(vmul(pp, im)), 0.0f);
    w[1] = f4(dw, 0.0f);
}
// both sides of the manifold (ea, eb: the list entries of its two bodies, ~0u for the static world); iaxn, ibxn, ia*,
// ib*: the inertia terms of side a and side b (unsplit)
__device__ __forceinline__ void write_warm(const Dev& D, uint32_t ea, uint32_t eb, int cnt, float4 n4, float4 t14,
                                           float4 t24, float j1, float j2, float jt, const float* jn, const V3* iaxn,
                                           const V3* ibxn, V3 ia1, V3 ia2, V3 ian, V3 ib1, V3 ib2, V3 ibn, float ima,
                                           float imb) {
    if (ea != ~0u) write_warm_side(WREC(D, ea), 0, cnt, n4, t14, t24, j1, j2, jt, jn, iaxn, ia1, ia2, ian, ima);
    if (eb != ~0u) write_warm_side(WREC(D, eb), 1, cnt, n4, t14, t24, j1, j2, jt, jn, ibxn, ib1, ib2, ibn, imb);
}

// push_manifold (phys.rs:450-498) and the constraint masses (:190-211), one manifold per lane, written straight into
// its place in colour order
__global__ void __launch_bounds__(NT) k_build(Dev D, StepK K, uint32_t nold) {
    __shared__ uint32_t wsum[NT / 64];
    uint32_t nc = min(D.ctr[C_NCAND], D.ccap);
    uint32_t npts = 0;  // this thread's contact points (the block's sum goes to ptpart, not to one global counter)
    uint32_t nbounce = 0;  // this thread's manifolds that may bounce
    GRID_STRIDE(i, nc) for (uint32_t j = 0; j < D.hit[i]; j++) {
        uint32_t m = D.hoff[i] + j;
        if (m >= D.mcap) { atomicAdd(&D.ctr[C_DROP], 1u); continue; }
        const uint32_t p = D.cur.pos[m];  // its place in colour order
        const bool red = D.mcol[m] == -2;  // reduced: built for the next step's warm start only
        int2 c = D.cand[i];
        int a = c.x, b = j ? -2 : c.y;  // (-2: a chunk's second patch of ground corners)
        int ba = D.shb[a];
        int bb = b >= 0 ? D.shb[b] : -1;
        V3 pa = v3f4(D.bp[ba]);
        M3 ra = load_m3(D.brot + 3 * (size_t)ba);
        V3 pb = bb >= 0 ? v3f4(D.bp[bb]) : v3(0, 0, 0);
        M3 rbm;
        if (bb >= 0) rbm = load_m3(D.brot + 3 * (size_t)bb);
        int sa = D.bslot[ba], sb = bb >= 0 ? D.bslot[bb] : 0;
        const float4* nr = cand_pts(D, i, j);
        float4 n4 = nr[0];
        V3 n = v3f4(n4);
        int cnt = (int)n4.w;
        Pt pts[4];
        // (point loops run to the fixed four with k < cnt tests, so the per-point arrays stay in registers)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k < cnt) { float4 q4 = nr[1 + k]; pts[k].p = v3f4(q4); pts[k].s = q4.w; }
            else { pts[k].p = v3(0, 0, 0); pts[k].s = 0.0f; }
        }
        V3 t1, t2;
        tangents(n, t1, t2);
        float mu = mix_mu(D.shh[a].w, b >= 0 ? D.shh[b].w : K.friction);
        // the larger restitution of the two, the ground's being 0 (gpu_world.js:601-603)
        const float rest = sf_max(D.srest[a], b >= 0 ? D.srest[b] : 0.0f);
        bool stick = K.anchor != 0;
        uint64_t key = man_key(a, b);
        int oi = find_old(D, key, nold);
        uint32_t opos = oi >= 0 ? D.old.pos[oi] : 0u;
        V3 cen = v3(0, 0, 0);
#pragma unroll
        for (int k = 0; k < 4; k++) if (k < cnt) cen = vadd(cen, pts[k].p);
        cen = vmul(cen, 1.0f / (float)cnt);
        float rad = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) if (k < cnt) rad += vlen(vsub(pts[k].p, cen));
        rad = rad / (float)cnt;
        V3 fa, fb, la_f, lb_f, e0 = v3(0, 0, 0);
        float jt1 = 0.0f, jt2 = 0.0f, jtw = 0.0f;
        bool anchored = false;
        if (oi >= 0) {
            const ManBuf& O = D.old;
            float4 oj = MROW(D, O, MR_J, opos);
            bool ostick = MROW(D, O, MR_CB2, opos).w != 0.0f;
            if (ostick) {
                V3 olaf = v3f4(MROW(D, O, MR_LAF, opos)), olbf = v3f4(MROW(D, O, MR_LBF, opos));
                V3 wa = vadd(pa, mulv(ra, olaf));
                V3 wb = bb >= 0 ? vadd(pb, mulv(rbm, olbf)) : olbf;
                float lim = sf_max(0.05f, rad);
                if (vlen2(vsub(wa, cen)) < lim * lim && vlen2(vsub(wb, wa)) < 0.03f * 0.03f) {
                    la_f = olaf; lb_f = olbf; fa = vsub(wa, pa); fb = vsub(wb, pb); e0 = vsub(wb, wa); anchored = true;
                }
            }
            V3 ot1 = v3f4(MROW(D, O, MR_T1, opos)), ot2 = v3f4(MROW(D, O, MR_T2, opos)), on = v3f4(MROW(D, O, MR_N, opos));
            V3 f = vadd(vmul(ot1, oj.x), vmul(ot2, oj.y));
            jt1 = vdot(f, t1); jt2 = vdot(f, t2); jtw = oj.z * vdot(on, n);
        }
        if (!anchored) {
            fa = vsub(cen, pa); fb = vsub(cen, pb); la_f = tmulv(ra, fa);
            lb_f = bb >= 0 ? tmulv(rbm, fb) : cen;
            e0 = v3(0, 0, 0);
        }
        // masses (the static side: slot 0, zero inverse mass and inertia)
        const float4 sva = SV(D, sa), svb = SV(D, sb);
        const V3 swa = v3f4(SW(D, sa)), swb = v3f4(SW(D, sb));
        float ima = sva.w, imb = svb.w;
        M3 iia = load_m3(SII(D, sa)), iib = load_m3(SII(D, sb));
        const ManBuf& C = D.cur;
        float jnk[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        V3 iaxk[4], ibxk[4];
        int ocnt = 0;
        bool bounce = false;
        V3 ola[4];  // last step's point anchors, loaded once for all the new points
        if (oi >= 0) {
            ocnt = min(max((int)MROW(D, D.old, MR_CB1, opos).w, 0), 4);
#pragma unroll
            for (int q = 0; q < 4; q++) ola[q] = v3f4(MROW(D, D.old, MR_P0 + MR_PER_PT * q + 2, opos));
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k >= cnt) { iaxk[k] = v3(0, 0, 0); ibxk[k] = v3(0, 0, 0); continue; }
            V3 rap = vsub(pts[k].p, pa);
            V3 rbp = bb >= 0 ? vsub(pts[k].p, pb) : pts[k].p;
            V3 la = tmulv(ra, rap);
            float jn = 0.0f;
            if (oi >= 0) {
                float best = 0.03f * 0.03f; int bi = -1;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (q >= ocnt) continue;
                    float d = vlen2(vsub(ola[q], la));
                    if (d < best) { best = d; bi = q; }
                }
                if (bi >= 0) jn = reinterpret_cast<const float*>(&MROW(D, D.old, MR_JN, opos))[bi];
            }
            V3 raxn = vcross(rap, n), rbxn = vcross(rbp, n);
            V3 iaxn = mulv(iia, raxn), ibxn = mulv(iib, rbxn);
            float kk = ima + imb + vdot(raxn, iaxn) + vdot(rbxn, ibxn);
            float nm = kk > 0.0f ? 1.0f / kk : 0.0f;
            // the point's normal velocity at the step start (negative: approaching), gpu_world.js:662
            const V3 pva = vadd(v3f4(sva), vcross(swa, rap)), pvb = vadd(v3f4(svb), vcross(swb, rbp));
            const float rv = vdot(vsub(pvb, pva), n);
            if (rest > 0.0f && !(rv > -K.rest_thr) && !red) bounce = true;
            MROW(D, C, MR_P0 + MR_PER_PT * k + 0, p) = f4(raxn, pts[k].s);
            MROW(D, C, MR_P0 + MR_PER_PT * k + 1, p) = f4(rbxn, nm);
            iaxk[k] = iaxn; ibxk[k] = ibxn; jnk[k] = jn;
            MROW(D, C, MR_P0 + MR_PER_PT * k + 2, p) = f4(la, rv);
        }
        V3 ca1 = vcross(fa, t1), ca2 = vcross(fa, t2), cb1 = vcross(fb, t1), cb2 = vcross(fb, t2);
        V3 ia1 = mulv(iia, ca1), ia2 = mulv(iia, ca2), ib1 = mulv(iib, cb1), ib2 = mulv(iib, cb2);
        float k1 = ima + imb + vdot(ca1, ia1) + vdot(cb1, ib1);
        float k2 = ima + imb + vdot(ca2, ia2) + vdot(cb2, ib2);
        float tm1 = k1 > 0.0f ? 1.0f / k1 : 0.0f, tm2 = k2 > 0.0f ? 1.0f / k2 : 0.0f;
        V3 ian = mulv(iia, n), ibn = mulv(iib, n);
        float kt = vdot(n, ian) + vdot(n, ibn);
        float twm = kt > 0.0f ? 1.0f / kt : 0.0f;
        C.hdrp[p] = make_int4(a, b, sa, sb);
        MROW(D, C, MR_N, p) = f4(n, mu);
        MROW(D, C, MR_T1, p) = f4(t1, tm1);
        MROW(D, C, MR_T2, p) = f4(t2, tm2);
        MROW(D, C, MR_E0, p) = f4(e0, twm);
        MROW(D, C, MR_J, p) = make_float4(jt1, jt2, jtw, rad);
        MROW(D, C, MR_CA1, p) = f4(ca1, ima);
        MROW(D, C, MR_CA2, p) = f4(ca2, imb);
        MROW(D, C, MR_CB1, p) = f4(cb1, (float)cnt);
        MROW(D, C, MR_CB2, p) = f4(cb2, stick ? 1.0f : 0.0f);
        MROW(D, C, MR_JN, p) = make_float4(jnk[0], jnk[1], jnk[2], jnk[3]);
        MROW(D, C, MR_LAF, p) = f4(la_f, rest);
        MROW(D, C, MR_LBF, p) = f4(lb_f, 0.0f);  // (.w: not pushed yet)
        npts += (uint32_t)cnt;
        nbounce += bounce ? 1u : 0u;
        // the bodies' list entries of this position, their first warm-start records (from the impulses carried over
        // from the last step) and their overflow changes cleared
        const uint32_t ea = sa && !red ? D.minv[2 * m] : ~0u, eb = sb && !red ? D.minv[2 * m + 1] : ~0u;
        D.blinv[2 * p] = ea;
        D.blinv[2 * p + 1] = eb;
        write_warm(D, ea, eb, cnt, f4(n, mu), f4(t1, tm1), f4(t2, tm2), jt1, jt2, jtw, jnk, iaxk, ibxk, ia1, ia2, ian, ib1,
                   ib2, ibn, ima, imb);
        const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (ea != ~0u) { OVL(D, ea)[0] = z; OVL(D, ea)[1] = z; }
        if (eb != ~0u) { OVL(D, eb)[0] = z; OVL(D, eb)[1] = z; }
    }
    for (int h = 32; h >= 1; h >>= 1) npts += __shfl_xor(npts, h, 64);
    if (__ballot(nbounce != 0) && (threadIdx.x & 63) == 0) atomicAdd(&D.ctr[C_NREST], 1u);  // (a flag: any is enough)
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = npts;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < NT / 64; k++) t += wsum[k];
        D.ptpart[blockIdx.x] = t;
    }
}
// the contact points of the step: the sum of k_build's per-block partials
__global__ void __launch_bounds__(NT) k_sum_points(Dev D, uint32_t nparts) {
    __shared__ uint32_t wsum[NT / 64];
    uint32_t t = 0;
    for (uint32_t i = threadIdx.x; i < nparts; i += NT) t += D.ptpart[i];
    for (int h = 32; h >= 1; h >>= 1) t += __shfl_xor(t, h, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0;
        for (int k = 0; k < NT / 64; k++) a += wsum[k];
        D.ctr[C_NPTS] = a;
    }
}

// the manifolds' headers (shapes, body slots) and keys, and the bodies' contact counts (reduced manifolds take no
// place in the bodies' lists), before colouring
__global__ void __launch_bounds__(NT) k_man_hdr(Dev D, StepK K) {
    uint32_t nc = min(D.ctr[C_NCAND], D.ccap);
    GRID_STRIDE(i, nc) for (uint32_t j = 0; j < D.hit[i]; j++) {
        uint32_t m = D.hoff[i] + j;
        if (m >= D.mcap) continue;
        int2 c = D.cand[i];
        const int cb = j ? -2 : c.y;
        int ba = D.shb[c.x];
        int sa = D.bslot[ba], sb = cb >= 0 ? D.bslot[D.shb[cb]] : 0;
        D.cur.hdr[m] = make_int4(c.x, cb, sa, sb);
        D.cur.key[m] = man_key(c.x, cb);
        if (K.red_k && D.rflag[m]) continue;
        if (sa) atomicAdd(&D.blcnt[sa], 1u);
        if (sb) atomicAdd(&D.blcnt[sb], 1u);
    }
}

// ---------------------------------------------------------------- contact reduction (gpu/gpu_world.js:834-920)
// A body pair (the ground counting as one body) touching through more than red_k manifolds keeps only representative
// ones: for each of the 26 directions to the neighbours of a cube cell the manifold whose points' centre lies furthest
// out, and the deepest (least separation); ties go to the later manifold. The rest are reduced: not coloured, not
// solved, not in the warm start, but built into the step's records (after the overflow) so the next step's warm
// start still finds their impulses. Three passes over the manifolds: count per pair in an open-addressing table
// (the pair's (red_k + 1)-th manifold hands the pair its place among the reduced pairs), score with 64-bit atomic
// maxima of (ordered score << 32 | manifold), pick. Which thread wins a race changes no result.
constexpr int RED_D = 27;  // 26 directions and the deepest
constexpr int RED_KMIN = 8;  // smaller thresholds act as this (it bounds the reduced pairs at a ninth of the manifolds)
__device__ __forceinline__ uint64_t red_pair(const Dev& D, int2 c) {
    const uint32_t ba = (uint32_t)D.shb[c.x], bb = c.y >= 0 ? (uint32_t)D.shb[c.y] : 0xFFFFFFFFu;
    return ((uint64_t)ba << 32) | bb;
}
// a float's bits as an unsigned integer in the float's order
__device__ __forceinline__ uint32_t f2o(float x) { const uint32_t b = __float_as_uint(x); return (b & 0x80000000u) ? ~b : (b | 0x80000000u); }
// direction k of 26: the cell offsets j = 0..26 in x-fastest order, the centre (j = 13) skipped
__device__ __forceinline__ V3 red_dir(int k) {
    const int j = k >= 13 ? k + 1 : k;
    return v3((float)(j % 3) - 1.0f, (float)((j / 3) % 3) - 1.0f, (float)(j / 9) - 1.0f);
}
__global__ void __launch_bounds__(NT) k_red_count(Dev D, StepK K) {
    const uint32_t nc = min(D.ctr[C_NCAND], D.ccap);
    GRID_STRIDE(i, nc) for (uint32_t j = 0; j < D.hit[i]; j++) {
        const uint32_t m = D.hoff[i] + j;
        if (m >= D.mcap) continue;
        D.rflag[m] = 0u;
        const uint64_t key = red_pair(D, D.cand[i]);
        // (the table has at least twice as many slots as there are manifolds, so a free slot is always found)
        uint32_t h = mix32((uint32_t)key * 0x9E3779B1u ^ mix32((uint32_t)(key >> 32))) & D.rmask;
        while (true) {
            const unsigned long long old = atomicCAS(&D.rkey[h], ~0ull, (unsigned long long)key);
            if (old == ~0ull || old == key) break;
            h = (h + 1u) & D.rmask;
        }
        D.rslot[m] = h;
        if (atomicAdd(&D.rcnt[h], 1u) == (uint32_t)K.red_k) {
            const uint32_t b = atomicAdd(&D.ctr[C_NREDP], 1u);
            D.rbig[h] = b + 1u;
            for (int k = 0; k < RED_D; k++) D.rext[(size_t)b * RED_D + k] = 0ull;
        }
    }
}
__global__ void __launch_bounds__(NT) k_red_score(Dev D) {
    if (!D.ctr[C_NREDP]) return;
    const uint32_t nc = min(D.ctr[C_NCAND], D.ccap);
    GRID_STRIDE(i, nc) for (uint32_t j = 0; j < D.hit[i]; j++) {
        const uint32_t m = D.hoff[i] + j;
        if (m >= D.mcap) continue;
        const uint32_t b = D.rbig[D.rslot[m]];
        if (!b) continue;
        const float4* nr = cand_pts(D, i, j);
        const int cnt = (int)nr[0].w;
        V3 c = v3(0, 0, 0);
        float smin = nr[1].w;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k >= cnt) continue;
            const float4 q = nr[1 + k];
            c = vadd(c, v3f4(q));
            if (k > 0) smin = q.w < smin ? q.w : smin;
        }
        c = vmul(c, 1.0f / (float)cnt);
        unsigned long long* e = D.rext + (size_t)(b - 1u) * RED_D;
        for (int k = 0; k < RED_D - 1; k++) atomicMax(&e[k], ((unsigned long long)f2o(vdot(c, red_dir(k))) << 32) | m);
        atomicMax(&e[RED_D - 1], ((unsigned long long)f2o(-smin) << 32) | m);
    }
}
__global__ void __launch_bounds__(NT) k_red_pick(Dev D) {
    if (!D.ctr[C_NREDP]) return;
    const uint32_t nc = min(D.ctr[C_NCAND], D.ccap);
    uint32_t nred = 0;
    GRID_STRIDE(i, nc) for (uint32_t j = 0; j < D.hit[i]; j++) {
        const uint32_t m = D.hoff[i] + j;
        if (m >= D.mcap) continue;
        const uint32_t b = D.rbig[D.rslot[m]];
        if (!b) continue;
        const unsigned long long* e = D.rext + (size_t)(b - 1u) * RED_D;
        bool kept = false;
        for (int k = 0; k < RED_D; k++) kept |= (uint32_t)e[k] == m;
        if (!kept) { D.rflag[m] = 1u; nred++; }
    }
    block_count(nred, &D.ctr[C_NRED]);
}

// ---------------------------------------------------------------- colouring
// (with reduction on, the pair table's slots the step used are cleared here, after the last pass that reads them)
__global__ void __launch_bounds__(NT) k_col_init(Dev D, StepK K) {
    uint32_t nm = min(D.ctr[C_NMAN], D.mcap);
    GRID_STRIDE(m, nm) {
        const bool red = K.red_k && D.rflag[m];
        D.mcol[m] = red ? -2 : -1;
        if (K.red_k) { const uint32_t h = D.rslot[m]; D.rkey[h] = ~0ull; D.rcnt[h] = 0u; D.rbig[h] = 0u; }
    }
}

// (the activity flag of the round's group is set by one plain store per wavefront that claimed: tens of thousands of
// atomic adds on one address serialise)
// Supports first (gpu/gpu_world.js:954-985): in rounds r < rs only supports claim (manifolds with a side that is not
// moving: the ground, a static or sleeping body), and with rs > 0 supports outrank the others in every round. A
// claim's priority: the round + 1 (bits 57-62: a claim left from an earlier round never wins), the support bit (56),
// a hash of the key and round, the manifold index.
__global__ void __launch_bounds__(NT) k_col_claim(Dev D, int K, int spread, int r, int rs, int flag) {
    uint32_t nm = min(D.ctr[C_NMAN], D.mcap);
    const unsigned long long all = K == 64 ? ~0ull : ((1ull << K) - 1ull);
    bool claimed = false;
    GRID_STRIDE(m, nm) {
        if (D.mcol[m] != -1) continue;
        int4 hd = D.cur.hdr[m];
        int sa = hd.z, sb = hd.w;
        const bool support = !sa || !sb;
        if (r < rs && !support) { D.mcand[m] = -1; continue; }  // (no claim this round: resolve passes it over)
        unsigned long long used = (sa ? D.bmask[sa] : 0ull) | (sb ? D.bmask[sb] : 0ull);
        unsigned long long fr = all & ~used;
        if (!fr) { D.mcol[m] = K; continue; }
        int nf = __popcll(fr);
        uint32_t kh = key_hash(hd.x, hd.y);
        int lim = nf < spread ? nf : spread;
        int pick = (int)(mix32(kh + (uint32_t)r * 0x165667B1u) % (uint32_t)lim);
        for (int q = 0; q < pick; q++) fr &= fr - 1;
        int c = __ffsll((long long)fr) - 1;
        unsigned long long pri = ((unsigned long long)(r + 1) << 57) | ((unsigned long long)(rs > 0 && support) << 56) |
                                 ((unsigned long long)(mix32(kh ^ ((uint32_t)r * 0x27D4EB2Fu)) >> 8) << 32) |
                                 (unsigned long long)(0xFFFFFFFFu - m);
        D.mcand[m] = c;
        D.mpri[m] = pri;
        if (sa) atomicMax(&D.claim[(size_t)sa * K + c], pri);
        if (sb) atomicMax(&D.claim[(size_t)sb * K + c], pri);
        claimed = true;
    }
    if (__ballot(claimed) && (threadIdx.x & 63) == 0) D.ctr[C_ACTIVE0 + flag] = 1u;
}

__global__ void __launch_bounds__(NT) k_col_resolve(Dev D, int K) {
    uint32_t nm = min(D.ctr[C_NMAN], D.mcap);
    GRID_STRIDE(m, nm) {
        if (D.mcol[m] != -1) continue;
        int4 hd = D.cur.hdr[m];
        int sa = hd.z, sb = hd.w, c = D.mcand[m];
        if (c < 0) continue;
        unsigned long long pri = D.mpri[m];
        if (sa && D.claim[(size_t)sa * K + c] != pri) continue;
        if (sb && D.claim[(size_t)sb * K + c] != pri) continue;
        D.mcol[m] = c;
        if (sa) atomicOr(&D.bmask[sa], 1ull << c);
        if (sb) atomicOr(&D.bmask[sb], 1ull << c);
    }
}

// The overflow bucket is placed in OVS spatial sub-buckets: a Morton index of its first moving body's 4 m cell, each
// axis taken modulo 4, so the contiguous eighth of the bucket that one XCD solves covers a compact region and its
// bodies' records are reused from that XCD's L2. Within the overflow (a Jacobi pass) the order changes no result.
constexpr int OVS = 64;
constexpr int NBIN = 64 + OVS + 1;  // colour bins (up to 64 colours), the overflow's sub-buckets, the reduced manifolds
__device__ __forceinline__ uint32_t ovf_key(const Dev& D, int sa, int sb) {
    const float4 p = D.sp0[sa ? sa : sb];
    const uint32_t x = (uint32_t)(int)floorf(p.x * 0.25f) & 3u, y = (uint32_t)(int)floorf(p.y * 0.25f) & 3u,
                   z = (uint32_t)(int)floorf(p.z * 0.25f) & 3u;
    return (x & 1u) | ((y & 1u) << 1) | ((z & 1u) << 2) | ((x & 2u) << 2) | ((y & 2u) << 3) | ((z & 2u) << 4);
}
// colour histogram: per-block counts in LDS, stored per bin per chunk (a few colours take every manifold, so
// global atomics, per manifold or per block, would serialise on a handful of addresses); the colour totals come
// from the chunk scan. Bins 0..K-1 are the colours, K + k the overflow's sub-bucket k; mbin keeps each manifold's bin.
__global__ void __launch_bounds__(NT) k_col_count(Dev D, int K) {
    __shared__ uint32_t hist[NBIN];
    uint32_t nm = min(D.ctr[C_NMAN], D.mcap);
    for (uint32_t b0 = blockIdx.x * NT; b0 < nm; b0 += gridDim.x * NT) {
        for (int c = threadIdx.x; c < K + OVS + 1; c += NT) hist[c] = 0u;
        __syncthreads();
        uint32_t m = b0 + threadIdx.x;
        if (m < nm) {
            int c = D.mcol[m];
            if (c == -2) c = K + OVS;  // (reduced: after the overflow)
            else {
                if (c < 0) { c = K; D.mcol[m] = K; }
                if (c == K) {
                    const int4 hd = D.cur.hdr[m];
                    c = K + (int)ovf_key(D, hd.z, hd.w);
                }
            }
            atomicAdd(&hist[c], 1u);
            D.mbin[m] = (uint32_t)c;
        }
        __syncthreads();
        const uint32_t nchunk = (D.mcap + NT - 1) / NT, chunk = b0 / NT;
        for (int c = threadIdx.x; c < K + OVS + 1; c += NT) D.chunkc[(size_t)c * nchunk + chunk] = hist[c];
        __syncthreads();
    }
}

// per bin, the exclusive scan of its chunk counts in chunk order (one block per bin): a colour's total into its
// bucket count, an overflow sub-bucket's into subcnt, the reduced manifolds' into subcnt[OVS]
__global__ void __launch_bounds__(NT) k_col_chunk_scan(Dev D, int K) {
    __shared__ uint32_t ws[NT / 64];
    const int c = blockIdx.x;
    const uint32_t nm = min(D.ctr[C_NMAN], D.mcap), nchunk = (D.mcap + NT - 1) / NT, used = (nm + NT - 1) / NT;
    uint32_t* a = D.chunkc + (size_t)c * nchunk;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < used; b0 += NT) {
        uint32_t i = b0 + threadIdx.x, v = i < used ? a[i] : 0u, x = v;
        for (int d = 1; d < 64; d <<= 1) { uint32_t y = __shfl_up(x, d, 64); if (lane >= d) x += y; }
        if (lane == 63) ws[wv] = x;
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (int k = 0; k < NT / 64; k++) { if (k < wv) before += ws[k]; all += ws[k]; }
        if (i < used) a[i] = carry + before + x - v;
        carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (c < K) {
            D.colcnt[c] = carry;
            if (carry) atomicMax(&D.ctr[C_NCOL], (uint32_t)c + 1u);
        } else D.subcnt[c - K] = carry;
    }
}
// colour-ordered list, stable: within a colour (and within an overflow sub-bucket) manifolds keep their canonical
// order (chunk offsets from the scan, ranks inside a chunk from wavefront ballots)
__global__ void __launch_bounds__(NT) k_col_fill(Dev D, int K) {
    __shared__ uint32_t wc[NT / 64][NBIN];
    const uint32_t nm = min(D.ctr[C_NMAN], D.mcap), nchunk = (D.mcap + NT - 1) / NT;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (blockIdx.x == 0 && threadIdx.x == 0) D.ctr[C_NOVF] = D.colcnt[K];
    for (uint32_t b0 = blockIdx.x * NT; b0 < nm; b0 += gridDim.x * NT) {
        for (int q = threadIdx.x; q < (NT / 64) * NBIN; q += NT) wc[q / NBIN][q % NBIN] = 0u;
        __syncthreads();
        uint32_t m = b0 + threadIdx.x;
        int c = m < nm ? (int)D.mbin[m] : -1;
        uint32_t rank = 0;
        unsigned long long todo = __ballot(c >= 0);
        while (todo) {
            int l = __ffsll((long long)todo) - 1;
            int cc = __shfl(c, l, 64);
            unsigned long long same = __ballot(c == cc);
            if (c == cc) rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
            if (lane == 0) wc[wv][cc] = (uint32_t)__popcll(same);
            todo &= ~same;
        }
        __syncthreads();
        if (c >= 0) {
            uint32_t before = 0;
            for (int k = 0; k < wv; k++) before += wc[k][c];
            const uint32_t base = c < K ? D.coloff[c] : (c < K + OVS ? D.coloff[K] + D.suboff[c - K] : D.coloff[K] + D.colcnt[K]);
            D.cur.pos[m] = base + D.chunkc[(size_t)c * nchunk + b0 / NT] + before + rank;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- per-body contact lists
__global__ void __launch_bounds__(NT) k_bl_fill(Dev D) {
    uint32_t nm = min(D.ctr[C_NMAN], D.mcap);
    GRID_STRIDE(m, nm) {
        if (D.mcol[m] == -2) continue;
        int4 hd = D.cur.hdr[m];
        if (hd.z) D.blist[D.bloff[hd.z] + atomicAdd(&D.blfill[hd.z], 1u)] = m * 2u;
        if (hd.w) D.blist[D.bloff[hd.w] + atomicAdd(&D.blfill[hd.w], 1u)] = m * 2u + 1u;
    }
}
// Each body's contact list into manifold order: one wavefront per body, a 64-lane bitonic sort in registers (lists
// longer than a wavefront are sorted by k_bl_sort_big).
__global__ void __launch_bounds__(NT) k_bl_sort(Dev D) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t s = wave; s < (uint32_t)D.nb + 1u; s += nwaves) {
        uint32_t o = D.bloff[s], n = D.blcnt[s];
        if (n < 2) continue;
        if (n <= 64) {
            uint32_t v = (uint32_t)lane < n ? D.blist[o + lane] : 0xFFFFFFFFu;
            for (int k = 2; k <= 64; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    uint32_t u = __shfl_xor(v, j, 64);
                    bool up = (lane & k) == 0, lower = (lane & j) == 0;
                    v = (lower == up) ? min(v, u) : max(v, u);
                }
            }
            if ((uint32_t)lane < n) D.blist[o + lane] = v;
        }
    }
}

// the slots whose lists are longer than a wavefront, into bigl (one thread per slot, one atomic per wavefront)
__global__ void __launch_bounds__(NT) k_bl_big(Dev D) {
    const int lane = threadIdx.x & 63;
    for (uint32_t s0 = (blockIdx.x * NT + threadIdx.x) & ~63u; s0 < (uint32_t)D.nb + 1u; s0 += gridDim.x * NT) {
        const uint32_t s = s0 + lane;
        const bool big = s < (uint32_t)D.nb + 1u && D.blcnt[s] > 64u;
        const unsigned long long m = __ballot(big);
        if (!m) continue;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&D.ctr[C_NBIGL], (uint32_t)__popcll(m));
        base = (uint32_t)__shfl((int)base, 0, 64);
        if (big) D.bigl[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = s;
    }
}
// lists longer than a wavefront (the largest bodies, listed by k_bl_big): one workgroup each, bitonic sort in LDS
__global__ void __launch_bounds__(NT) k_bl_sort_big(Dev D) {
    __shared__ uint32_t sh[2048];
    const uint32_t nbig = min(D.ctr[C_NBIGL], (uint32_t)D.nb + 1u);
    for (uint32_t t = blockIdx.x; t < nbig; t += gridDim.x) {
        const uint32_t s = D.bigl[t];
        uint32_t o = D.bloff[s], n = D.blcnt[s];
        if (n > 2048) {
            if (threadIdx.x == 0) {
                for (uint32_t i = 1; i < n; i++) {
                    uint32_t v = D.blist[o + i];
                    uint32_t j = i;
                    while (j > 0 && D.blist[o + j - 1] > v) { D.blist[o + j] = D.blist[o + j - 1]; j--; }
                    D.blist[o + j] = v;
                }
            }
            continue;
        }
        uint32_t n2 = 128;
        while (n2 < n) n2 <<= 1;
        for (uint32_t i = threadIdx.x; i < n2; i += NT) sh[i] = i < n ? D.blist[o + i] : 0xFFFFFFFFu;
        __syncthreads();
        for (uint32_t k = 2; k <= n2; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < n2; i += NT) {
                    uint32_t l = i ^ j;
                    if (l > i) {
                        uint32_t x = sh[i], y = sh[l];
                        bool up = (i & k) == 0;
                        if ((x > y) == up) { sh[i] = y; sh[l] = x; }
                    }
                }
                __syncthreads();
            }
        }
        for (uint32_t i = threadIdx.x; i < n; i += NT) D.blist[o + i] = sh[i];
        __syncthreads();
    }
}

// the list entry of each (manifold, side): the inverse of the sorted lists
__global__ void __launch_bounds__(NT) k_bl_inv(Dev D) {
    uint32_t nm = min(D.ctr[C_NMAN], D.mcap);
    GRID_STRIDE(e, 2 * nm) {
        if (e >= D.bloff[D.nb] + D.blcnt[D.nb]) continue;
        D.minv[D.blist[e]] = e;
    }
}

// ---------------------------------------------------------------- solver
// gravity, spin clamp, then this body's warm start gathered over its contacts in manifold order: the same
// sequence of f32 updates the reference's manifold loop applies to it (phys.rs:217-222, :245-262)
__device__ __forceinline__ V3 rdl3(V3 a, int l) { return v3(rdl(a.x, l), rdl(a.y, l), rdl(a.z, l)); }

// Cross-lane steps of the trees, on the vector ALU where the hardware allows: lane i takes lane i + H within its row
// of 16 lanes (a DPP row shift, H = 1, 2, 4, 8), or lane i ^ 16 within its 32 lanes (a swizzle). Lanes whose source
// lies outside the row keep their own value; in a halving tree they never feed lane 0 of a row.
template <int H>
__device__ __forceinline__ float row_down(float x) {
    static_assert(H == 1 || H == 2 || H == 4 || H == 8, "row shift of 1, 2, 4 or 8 lanes");
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), 0x100 + H, 0xF, 0xF, false));
}
__device__ __forceinline__ float xor16(float x) {
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), 0x1F | (0x10 << 10)));
}
template <int H>
__device__ __forceinline__ V3 v3_row_down(V3 x) { return v3(row_down<H>(x.x), row_down<H>(x.y), row_down<H>(x.z)); }
__device__ __forceinline__ V3 v3_xor16(V3 x) { return v3(xor16(x.x), xor16(x.y), xor16(x.z)); }
// the tree's steps h = 16 ... 1 within each 32 lanes (lane 0 and lane 32 end with their half's total)
__device__ __forceinline__ V3 tree32(V3 x) {
    x = vadd(x, v3_xor16(x));
    x = vadd(x, v3_row_down<8>(x));
    x = vadd(x, v3_row_down<4>(x));
    x = vadd(x, v3_row_down<2>(x));
    return vadd(x, v3_row_down<1>(x));
}
// Fixed halving tree across a wavefront (x_i += x_{i+h}, h = 32 ... 1); lane 0's total is returned to every lane.
// The oracle sums in exactly this tree.
__device__ __forceinline__ V3 wave_tree(V3 x) {
    x = vadd(x, v3(__shfl_down(x.x, 32, 64), __shfl_down(x.y, 32, 64), __shfl_down(x.z, 32, 64)));
    return rdl3(tree32(x), 0);
}
// The same tree for a list of at most 32 on a half wavefront: its first step adds the zeros of lanes 32-63 (x + 0,
// kept: it turns -0 into +0 as that step does), then halving within the 32 lanes; every lane of the half gets the total.
__device__ __forceinline__ V3 half_tree(V3 x) {
    x = tree32(vadd(x, v3(0.0f, 0.0f, 0.0f)));
    const V3 lo = rdl3(x, 0), hi = rdl3(x, 32);
    return vsel((threadIdx.x & 32) != 0, hi, lo);
}
// What the per-body kernels read of a body's solver slot. A body's slot is its index + 1 while it is awake and 0
// otherwise (k_prep_bodies), so these loads are issued on index + 1 in the same round trip as bslot, which then only
// says whether the body takes part (s = 0: it does not; its list then counts as empty).
struct BodyIn {
    int s;                // the slot, 0 for a body that is not awake
    uint32_t o, n, cnt;   // contact-list offset and length, overflow manifolds
    float4 v4, w4;        // SV, SW
};
__device__ __forceinline__ BodyIn body_in(const Dev& D, uint32_t i) {
    const int q = (int)i + 1;
    BodyIn b;
    const int s = D.bslot[i];
    const uint32_t o = D.bloff[q], n = D.blcnt[q], cnt = D.ovcnt[q];
    b.v4 = SV(D, q); b.w4 = SW(D, q);
    b.s = s; b.o = o; b.n = s ? n : 0u; b.cnt = s ? cnt : 0u;
    return b;
}
__device__ __forceinline__ BodyIn body_pick(bool hi, const BodyIn& b0, const BodyIn& b1) {
    BodyIn b;
    b.s = hi ? b1.s : b0.s; b.o = hi ? b1.o : b0.o; b.n = hi ? b1.n : b0.n; b.cnt = hi ? b1.cnt : b0.cnt;
    b.v4 = hi ? b1.v4 : b0.v4; b.w4 = hi ? b1.w4 : b0.w4;
    return b;
}
// Per-body wavefront work over every body, two bodies at once when both have at most 32 contacts (a half wavefront
// each, one chunk, half_tree), else one after the other on the whole wavefront (wave_tree): body(b, lane, half)
// (b.s = 0, a resting body, in a half: it takes part in the shuffles and writes nothing).
template <class Body>
__device__ __forceinline__ void body_pairs(const Dev& D, Body body) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    const uint32_t nb = (uint32_t)D.nb, npair = (nb + 1) / 2;
    for (uint32_t pi = wave; pi < npair; pi += nwaves) {
        const uint32_t i0 = 2 * pi, i1 = i0 + 1 < nb ? i0 + 1 : i0;
        const BodyIn b0 = body_in(D, i0);
        BodyIn b1 = body_in(D, i1);
        if (i1 == i0) { b1.s = 0; b1.n = 0u; b1.cnt = 0u; }  // an odd count: the last wavefront's second half is empty
        if (b0.n <= 32 && b1.n <= 32) {
            body(body_pick(lane >> 5, b0, b1), lane & 31, true);
        } else {
            if (b0.s) body(b0, lane, false);
            if (b1.s) body(b1, lane, false);
        }
    }
}
// the tree sums of a body's list entries (offset o, length n): f(e) gives the two vectors of entry e; half: one
// chunk of 32 lanes
template <class F>
__device__ __forceinline__ void list_sums(uint32_t o, uint32_t n, int lane, bool half, F f, V3& a, V3& b) {
    a = v3(0, 0, 0); b = v3(0, 0, 0);
    if (half) {
        V3 x = v3(0, 0, 0), y = v3(0, 0, 0);
        if ((uint32_t)lane < n) f(o + lane, x, y);
        a = vadd(a, half_tree(x));
        b = vadd(b, half_tree(y));
        return;
    }
    for (uint32_t e0 = 0; e0 < n; e0 += 64) {
        uint32_t e = e0 + lane;
        V3 x = v3(0, 0, 0), y = v3(0, 0, 0);
        if (e < n) f(o + e, x, y);
        a = vadd(a, wave_tree(x));
        b = vadd(b, wave_tree(y));
    }
}

// Gravity, spin clamp, then the warm start (phys.rs:217-222, :245-262): one wavefront per body, each lane one contact
// (the body's contacts in manifold order, 64 at a time); a contact's contribution is its record's terms summed in
// point order, and the body adds the tree sum of its contacts' contributions.
// one body's warm start (its wavefront; every lane holds the same b on entry; nothing is written for b.s = 0)
__device__ __forceinline__ void warm_body(const Dev& D, const StepK& K, const BodyIn& b, int lane) {
    const V3 gh = vmul(v3(K.gx, K.gy, K.gz), K.h);
    const float4 v4 = b.v4;
    V3 v = vadd(v3f4(v4), gh), w = v3f4(b.w4);
    float wl = vlen2(w);
    if (wl > K.wmax * K.wmax) w = vmul(w, K.wmax / sf_sqrt(wl));
    const int s = b.s;
    const uint32_t o = b.o, n = b.n;
    V3 av = v3(0, 0, 0), aw = v3(0, 0, 0);
    for (uint32_t e0 = 0; e0 < n; e0 += 64) {
        uint32_t e = e0 + lane;
        V3 dv = v3(0, 0, 0), dw = v3(0, 0, 0);
        if (e < n) {
            const float4* wr = WREC(D, o + e);
            dv = v3f4(wr[0]);
            dw = v3f4(wr[1]);
        }
        av = vadd(av, wave_tree(dv));
        aw = vadd(aw, wave_tree(dw));
    }
    if (s && lane == 0) {
        SV(D, s) = f4(vadd(v, av), v4.w);
        SW(D, s) = f4(vadd(w, aw), 0.0f);
    }
}
__global__ void __launch_bounds__(NT) k_warm(Dev D, StepK K) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t i = wave; i < (uint32_t)D.nb; i += nwaves) warm_body(D, K, body_in(D, i), lane);
}

__device__ __forceinline__ float slop_fn(float e, float slop) { return e > slop ? e - slop : (e < -slop ? e + slop : 0.0f); }

// Mass-ratio limit (gpu_world.js:1155-1160): across one contact the heavier side responds as if at most R times heavier
// than the lighter; the factors multiply the two inverse masses (1 when either side is static or R is 0).
__device__ __forceinline__ void ratio_scale(float ia, float ib, float R, float& sa, float& sb) {
    sa = 1.0f; sb = 1.0f;
    if (R <= 0.0f || ia <= 0.0f || ib <= 0.0f) return;
    if (ia * R < ib) sa = ib / (R * ia);
    else if (ib * R < ia) sb = ia / (R * ib);
}
// the limited side keeps the part of a change d to its velocity x that brings |x| down to its least and takes the rest
// divided by sc: the limit only takes energy from the heavy body (gpu_world.js:1166-1177)
__device__ __forceinline__ V3 slowing(V3 x, V3 d, float sc) {
    const float dd = vdot(d, d);
    if (dd == 0.0f) return d;
    const float t = sf_min(sf_max(-vdot(x, d) / dd, 0.0f), 1.0f);
    return vmul(d, t + (1.0f - t) / sc);
}
// a slot's shock-propagation height, 255 for the static world (v4: the slot's v|inv_m)
__device__ __forceinline__ uint32_t slot_height(const Dev& D, int s, const float4& v4) { return s && v4.w > 0.0f ? D.lvl[s] : 255u; }

// The soft-step solve of one manifold (phys.rs:268-323) on velocity copies of its two bodies. fa, fb = 0: as the
// reference solves it. fa, fb > 0 (the overflow pass): mass splitting, each side's inverse mass and inertia scaled
// by the number of overflow manifolds at that body and the effective masses recomputed from the scaled ones.
// The GPU world's extensions (gpu_world.js:1207-1329): with a mass-ratio limit the inverse masses (as split) are
// scaled by ratio_scale as well and each limited side keeps its change only as `slowing` allows; in the relax pass
// with levels (ha, hb: the sides' heights) a contact within 45 degrees of vertical whose lower body sits on fewer
// levels holds that body where a point is separating, the upper side alone taking the point's impulse.
__device__ __forceinline__ void solve_core(const Dev& D, uint32_t m, const StepK& K, V3& va, V3& wa, V3& vb, V3& wb,
                                           V3 tha, V3 thb, V3 dpab, float fa, float fb, uint32_t ea, uint32_t eb,
                                           const M3& iia, const M3& iib, uint32_t ha, uint32_t hb) {
    const ManBuf& C = D.cur;
    const bool split = fa > 0.0f;
    const V3 va0 = va, wa0 = wa, vb0 = vb, wb0 = wb;
    // every load of the record up front (they are independent; the stores come at the end)
    float4 n4 = MROW(D, C, MR_N, m), t14 = MROW(D, C, MR_T1, m), t24 = MROW(D, C, MR_T2, m);
    float4 ca14 = MROW(D, C, MR_CA1, m), ca24 = MROW(D, C, MR_CA2, m), cb14 = MROW(D, C, MR_CB1, m), cb24 = MROW(D, C, MR_CB2, m);
    float4 e04 = MROW(D, C, MR_E0, m), j4 = MROW(D, C, MR_J, m), jn4 = MROW(D, C, MR_JN, m);
    float4 r0[4], r1[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        r0[k] = MROW(D, C, MR_P0 + MR_PER_PT * k + 0, m);
        r1[k] = MROW(D, C, MR_P0 + MR_PER_PT * k + 1, m);
    }
    const float jnv[4] = {jn4.x, jn4.y, jn4.z, jn4.w};
    V3 iax0[4], ibx0[4];  // the points' inertia terms, unsplit
    V3 n = v3f4(n4), t1 = v3f4(t14), t2 = v3f4(t24);
    float dn = vdot(dpab, n);
    float sca, scb;
    ratio_scale(split ? ca14.w * fa : ca14.w, split ? ca24.w * fb : ca24.w, K.ratio, sca, scb);
    const bool rec = split || sca != 1.0f || scb != 1.0f;
    const float xa = (split ? fa : 1.0f) * sca, xb = (split ? fb : 1.0f) * scb;
    bool hold_a = false, hold_b = false;
    if (!K.use_bias && K.shock) { hold_a = n.y > 0.7f && ha > hb; hold_b = n.y < -0.7f && hb > ha; }
    float ima = rec ? ca14.w * xa : ca14.w, imb = rec ? ca24.w * xb : ca24.w;
    int cnt = (int)cb14.w;
    float tot = 0.0f;
    float njs[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    // Every point is evaluated and only the first cnt are applied (selects, not a branch): a branch lets the compiler
    // move each point's row loads after the previous point's updates, one memory round trip per point; this way the
    // rows load with the rest of the record. For the points applied the operations are the same.
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool on = k < cnt;
        V3 raxn = v3f4(r0[k]), rbxn = v3f4(r1[k]);
        iax0[k] = mulv(iia, raxn); ibx0[k] = mulv(iib, rbxn);
        V3 iaxn = rec ? vmul(iax0[k], xa) : iax0[k], ibxn = rec ? vmul(ibx0[k], xb) : ibx0[k];
        float nm = r1[k].w;
        if (rec) { float kk = ima + imb + vdot(raxn, iaxn) + vdot(rbxn, ibxn); nm = kk > 0.0f ? 1.0f / kk : 0.0f; }
        float jn = jnv[k];
        float s = r0[k].w + dn + vdot(thb, rbxn) - vdot(tha, raxn);
        float bias, ms, is;
        if (s > 0.0f) { bias = s * K.inv_h; ms = 1.0f; is = 0.0f; }
        else if (K.use_bias) { bias = sf_max(K.bias_rate * s, -K.max_push); ms = K.mass_scale; is = K.imp_scale; }
        else { bias = 0.0f; ms = 1.0f; is = 0.0f; }
        float vn = vdot(vsub(vb, va), n) + vdot(wb, rbxn) - vdot(wa, raxn);
        const bool ga = hold_a && vn > 0.0f, gb = hold_b && vn > 0.0f;  // this point holds side a (b)
        if (ga) { float kk = imb + vdot(rbxn, ibxn); nm = kk > 0.0f ? 1.0f / kk : 0.0f; }
        if (gb) { float kk = ima + vdot(raxn, iaxn); nm = kk > 0.0f ? 1.0f / kk : 0.0f; }
        float imp = -nm * ms * (vn + bias) - is * jn;
        float nj = sf_max(jn + imp, 0.0f);
        imp = nj - jn;
        const float tot1 = tot + nj;
        const V3 va1 = ga ? va : vsub(va, vmul(n, imp * ima)), wa1 = ga ? wa : vsub(wa, vmul(iaxn, imp));
        const V3 vb1 = gb ? vb : vadd(vb, vmul(n, imp * imb)), wb1 = gb ? wb : vadd(wb, vmul(ibxn, imp));
        tot = on ? tot1 : tot;
        njs[k] = on ? nj : jnv[k];
        va = vsel(on, va1, va); wa = vsel(on, wa1, wa); vb = vsel(on, vb1, vb); wb = vsel(on, wb1, wb);
    }
    V3 ca1 = v3f4(ca14), ca2 = v3f4(ca24), cb1 = v3f4(cb14), cb2 = v3f4(cb24);
    const V3 ia10 = mulv(iia, ca1), ia20 = mulv(iia, ca2), ib10 = mulv(iib, cb1), ib20 = mulv(iib, cb2);
    const V3 ian0 = mulv(iia, n), ibn0 = mulv(iib, n);
    V3 ia1 = ia10, ia2 = ia20, ib1 = ib10, ib2 = ib20, ian = ian0, ibn = ibn0;
    float tm1 = t14.w, tm2 = t24.w, twm = e04.w;
    if (rec) {
        ia1 = vmul(ia1, xa); ia2 = vmul(ia2, xa); ian = vmul(ian, xa); ib1 = vmul(ib1, xb); ib2 = vmul(ib2, xb); ibn = vmul(ibn, xb);
        float k1 = ima + imb + vdot(ca1, ia1) + vdot(cb1, ib1);
        float k2 = ima + imb + vdot(ca2, ia2) + vdot(cb2, ib2);
        tm1 = k1 > 0.0f ? 1.0f / k1 : 0.0f; tm2 = k2 > 0.0f ? 1.0f / k2 : 0.0f;
        float kt = vdot(n, ian) + vdot(n, ibn);
        twm = kt > 0.0f ? 1.0f / kt : 0.0f;
    }
    float lim = n4.w * tot;
    V3 dv = vsub(vb, va);
    float v1 = vdot(dv, t1) + vdot(wb, cb1) - vdot(wa, ca1);
    float v2 = vdot(dv, t2) + vdot(wb, cb2) - vdot(wa, ca2);
    bool stick = cb24.w != 0.0f;
    bool anchored = K.use_bias && stick;
    float b1, b2, ms, is;
    if (anchored) {
        V3 e0 = v3f4(e04);
        float e1 = vdot(e0, t1) + vdot(dpab, t1) + vdot(thb, cb1) - vdot(tha, ca1);
        float e2 = vdot(e0, t2) + vdot(dpab, t2) + vdot(thb, cb2) - vdot(tha, ca2);
        float ap = K.anchor_push, sl = K.anchor_slop;
        b1 = sf_min(sf_max(K.bias_rate * slop_fn(e1, sl), -ap), ap);
        b2 = sf_min(sf_max(K.bias_rate * slop_fn(e2, sl), -ap), ap);
        ms = K.mass_scale; is = K.imp_scale;
    } else { b1 = 0.0f; b2 = 0.0f; ms = 1.0f; is = 0.0f; }
    float o1 = j4.x, o2 = j4.y;
    float j1 = o1 - tm1 * ms * (v1 + b1) - is * o1;
    float j2 = o2 - tm2 * ms * (v2 + b2) - is * o2;
    float l2 = j1 * j1 + j2 * j2;
    bool unstick = false;
    if (l2 > lim * lim) { float kk = lim / sf_sqrt(l2); j1 *= kk; j2 *= kk; unstick = true; }
    float d1 = j1 - o1, d2 = j2 - o2;
    V3 p = vadd(vmul(t1, d1), vmul(t2, d2));
    va = vsub(va, vmul(p, ima)); wa = vsub(wa, vadd(vmul(ia1, d1), vmul(ia2, d2)));
    vb = vadd(vb, vmul(p, imb)); wb = vadd(wb, vadd(vmul(ib1, d1), vmul(ib2, d2)));
    float wl = lim * j4.w;
    float ot = j4.z;
    float jt = sf_min(sf_max(ot - twm * vdot(vsub(wb, wa), n), -wl), wl);
    float dtw = jt - ot;
    wa = vsub(wa, vmul(ian, dtw));
    wb = vadd(wb, vmul(ibn, dtw));
    if (sca != 1.0f) { va = vadd(va0, slowing(va0, vsub(va, va0), sca)); wa = vadd(wa0, slowing(wa0, vsub(wa, wa0), sca)); }
    if (scb != 1.0f) { vb = vadd(vb0, slowing(vb0, vsub(vb, vb0), scb)); wb = vadd(wb0, slowing(wb0, vsub(wb, wb0), scb)); }
    // stores: impulses and the stick flag (and, for the restitution pass, that a relax pass pushed)
    MROW(D, C, MR_JN, m) = make_float4(njs[0], njs[1], njs[2], njs[3]);
    MROW(D, C, MR_J, m) = make_float4(j1, j2, jt, j4.w);
    if (unstick) reinterpret_cast<float*>(&MROW(D, C, MR_CB2, m))[3] = 0.0f;
    if (K.any_rest && !K.use_bias && tot > 0.0f) reinterpret_cast<float*>(&MROW(D, C, MR_LBF, m))[3] = 1.0f;
    // the relax pass ends the sub-step: its impulses are the next sub-step's warm start
    if (!K.use_bias)
        write_warm(D, ea, eb, cnt, n4, t14, t24, j1, j2, jt, njs, iax0, ibx0, ia10, ia20, ian0, ib10, ib20, ibn0, ca14.w, ca24.w);
}

// one manifold of the colour-ordered solve (p: its position in colour order)
__device__ void solve_one(const Dev& D, uint32_t m, const StepK& K) {
    // (the list entries are loaded beside the header, not after it)
    uint2 e2 = K.use_bias ? make_uint2(~0u, ~0u) : reinterpret_cast<const uint2*>(D.blinv)[m];
    int4 hd = D.cur.hdrp[m];
    int sa = hd.z, sb = hd.w;
    float4 va4 = SV(D, sa), vb4 = SV(D, sb);
    V3 va = v3f4(va4), wa = v3f4(SW(D, sa)), vb = v3f4(vb4), wb = v3f4(SW(D, sb));
    V3 dpab = vsub(v3f4(SDP(D, sb)), v3f4(SDP(D, sa)));
    // the relax pass writes the next warm start into the bodies' list entries
    uint32_t ea = sa ? e2.x : ~0u, eb = sb ? e2.y : ~0u;
    const M3 iia = load_m3(SII(D, sa)), iib = load_m3(SII(D, sb));
    uint32_t ha = 255u, hb = 255u;
    if (K.shock && !K.use_bias) { ha = slot_height(D, sa, va4); hb = slot_height(D, sb, vb4); }
    solve_core(D, m, K, va, wa, vb, wb, v3f4(STH(D, sa)), v3f4(STH(D, sb)), dpab, 0.0f, 0.0f, ea, eb, iia, iib, ha, hb);
    if (sa) { SV(D, sa) = f4(va, va4.w); SW(D, sa) = f4(wa, 0.0f); }
    if (sb) { SV(D, sb) = f4(vb, vb4.w); SW(D, sb) = f4(wb, 0.0f); }
}

// one colour bucket: its manifolds share no moving body, so every lane solves one
// One colour bucket. A bucket holds tens of thousands of manifolds, so workgroups are one wavefront each to spread
// them over every CU. Workgroups go to the 8 XCDs round robin (block b to XCD b % 8); block b takes the b-th slice of
// its XCD's contiguous eighth of the bucket, so an XCD's L2 holds the bodies of one region of the bucket instead of
// all of them. Within a colour no two manifolds share a moving body, so any assignment gives the same result.
constexpr int SOLVE_NT = 64;
// The grid may be larger than the bucket needs (a captured launch sized with headroom): the first G blocks, G from
// the bucket's count, do the work, spread over the XCDs like a grid of exactly G. NTB threads per workgroup (64 by
// default; 128 and 256 for the launch-shape tuner): any value gives the same results.
template <int NTB>
__global__ void __launch_bounds__(NTB) k_solve_bucket(Dev D, StepK K, int c) {
    const uint32_t n = D.colcnt[c], o = D.coloff[c];
    const uint32_t G = min(gridDim.x, max((n + NTB - 1) / NTB, 1u)), b = blockIdx.x;
    if (b >= G) return;
    const uint32_t x = b & 7u, q = G >> 3, r = G & 7u;
    const uint32_t lb = x * q + min(x, r) + (b >> 3);  // a bijection of [0, G)
    for (uint32_t i = lb * NTB + threadIdx.x; i < n; i += G * NTB) solve_one(D, o + i, K);
}
int solve_nt_of(const float* prm) { int v = (int)prm[SF_P_SOLVE_NT]; return v == 128 || v == 256 ? v : SOLVE_NT; }
void launch_solve_bucket(int nt, unsigned grid, hipStream_t s, const Dev& D, const StepK& K, int c) {
    if (nt == 256) hipLaunchKernelGGL(k_solve_bucket<256>, dim3(grid), dim3(256), 0, s, D, K, c);
    else if (nt == 128) hipLaunchKernelGGL(k_solve_bucket<128>, dim3(grid), dim3(128), 0, s, D, K, c);
    else hipLaunchKernelGGL(k_solve_bucket<SOLVE_NT>, dim3(grid), dim3(SOLVE_NT), 0, s, D, K, c);
}

// the small colours from c0 on, one after another in one workgroup
__global__ void __launch_bounds__(NT) k_solve_tail(Dev D, StepK K, int c0) {
    int ncol = (int)D.ctr[C_NCOL];
    for (int c = c0; c < ncol; c++) {
        uint32_t n = D.colcnt[c], o = D.coloff[c];
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) solve_one(D, o + i, K);
        __syncthreads();
    }
}

// overflow (Jacobi with mass splitting): every overflow manifold against copies of its bodies taken before the pass.
// Block b works on a slice of its XCD's contiguous eighth of the bucket, as in k_solve_bucket, so an XCD's manifolds
// lie in a few of the bucket's spatial sub-buckets and share their bodies' records in that XCD's L2.
__global__ void __launch_bounds__(NT) k_ovf_solve(Dev D, StepK K) {
    const uint32_t n = D.colcnt[K.K], o = D.coloff[K.K];
    const uint32_t G = min(gridDim.x, max((n + NT - 1) / NT, 1u)), b = blockIdx.x;
    if (b >= G) return;
    const uint32_t x = b & 7u, q = G >> 3, r = G & 7u;
    const uint32_t lb = x * q + min(x, r) + (b >> 3);  // a bijection of [0, G)
    for (uint32_t i = lb * NT + threadIdx.x; i < n; i += G * NT) {
        uint32_t m = o + i;
        uint2 e2 = reinterpret_cast<const uint2*>(D.blinv)[m];
        int4 hd = D.cur.hdrp[m];
        int sa = hd.z, sb = hd.w;
        const float4 va4 = SV(D, sa), vb4 = SV(D, sb);
        V3 va0 = v3f4(va4), wa0 = v3f4(SW(D, sa)), vb0 = v3f4(vb4), wb0 = v3f4(SW(D, sb));
        V3 va = va0, wa = wa0, vb = vb0, wb = wb0;
        V3 dpab = vsub(v3f4(SDP(D, sb)), v3f4(SDP(D, sa)));
        float fa = sa ? (float)D.ovcnt[sa] : 1.0f, fb = sb ? (float)D.ovcnt[sb] : 1.0f;
        uint32_t ea = sa ? e2.x : ~0u, eb = sb ? e2.y : ~0u;
        const M3 iia = load_m3(SII(D, sa)), iib = load_m3(SII(D, sb));
        uint32_t ha = 255u, hb = 255u;
        if (K.shock && !K.use_bias) { ha = slot_height(D, sa, va4); hb = slot_height(D, sb, vb4); }
        solve_core(D, m, K, va, wa, vb, wb, v3f4(STH(D, sa)), v3f4(STH(D, sb)), dpab, fa, fb,
                   K.use_bias ? ~0u : ea, K.use_bias ? ~0u : eb, iia, iib, ha, hb);
        // each side's change into its body's list entry
        if (ea != ~0u) { OVL(D, ea)[0] = f4(vsub(va, va0), 0.0f); OVL(D, ea)[1] = f4(vsub(wa, wa0), 0.0f); }
        if (eb != ~0u) { OVL(D, eb)[0] = f4(vsub(vb, vb0), 0.0f); OVL(D, eb)[1] = f4(vsub(wb, wb0), 0.0f); }
    }
}
// each body takes the average of its copies: the tree sum of the changes over its contacts (zero for contacts not
// in the overflow), one wavefront per body reading its list entries' changes in one contiguous run
// one body's average of its overflow copies (its wavefront): v, w (every lane) take the tree sums of the changes
// over its contacts (zero for contacts not in the overflow) over the count
__device__ __forceinline__ void ovf_average(const Dev& D, uint32_t o, uint32_t n, int lane, uint32_t cnt, V3& v, V3& w) {
    V3 av = v3(0, 0, 0), aw = v3(0, 0, 0);
    for (uint32_t e0 = 0; e0 < n; e0 += 64) {
        uint32_t e = e0 + lane;
        V3 dvv = v3(0, 0, 0), dww = v3(0, 0, 0);
        if (e < n) {
            const float4* dv = OVL(D, o + e);
            dvv = v3f4(dv[0]);
            dww = v3f4(dv[1]);
        }
        av = vadd(av, wave_tree(dvv));
        aw = vadd(aw, wave_tree(dww));
    }
    float inv = 1.0f / (float)cnt;
    v = vadd(v, vmul(av, inv));
    w = vadd(w, vmul(aw, inv));
}
// (rest_gate: after the restitution pass, which left the changes of the last relax pass in place when no manifold
// of the step may bounce: then nothing to apply)
__global__ void __launch_bounds__(NT) k_ovf_apply(Dev D, StepK K, int rest_gate) {
    if (rest_gate && !D.ctr[C_NREST]) return;
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t i = wave; i < (uint32_t)D.nb; i += nwaves) {
        const BodyIn b = body_in(D, i);
        if (!b.cnt) continue;  // (zero for a body that is not awake)
        V3 v = v3f4(b.v4), w = v3f4(b.w4);
        ovf_average(D, b.o, b.n, lane, b.cnt, v, w);
        if (lane == 0) { SV(D, b.s) = f4(v, b.v4.w); SW(D, b.s) = f4(w, 0.0f); }
    }
}
// the bias pass's end: the overflow average, then the position integration (k_integrate's update), per body
__device__ __forceinline__ void integrate_slot(const Dev& D, const StepK& K, int s, V3 v, V3 w) {
    SDP(D, s) = f4(vadd(v3f4(SDP(D, s)), vmul(v, K.h)), 0.0f);
    STH(D, s) = f4(vadd(v3f4(STH(D, s)), vmul(w, K.h)), 0.0f);
    SQ(D, s) = f4q(qintegrate(qf4(SQ(D, s)), w, K.h));
}
__global__ void __launch_bounds__(NT) k_ovf_apply_integrate(Dev D, StepK K) {
    body_pairs(D, [&](const BodyIn& b, int l, bool half) {
        const int s = b.s;
        V3 v = v3f4(b.v4), w = v3f4(b.w4);
        V3 ov, ow;
        list_sums(b.o, b.n, l, half, [&](uint32_t e, V3& x, V3& y) { const float4* d = OVL(D, e); x = v3f4(d[0]); y = v3f4(d[1]); }, ov, ow);
        if (!s) return;
        if (b.cnt) {
            float inv = 1.0f / (float)b.cnt;
            v = vadd(v, vmul(ov, inv));
            w = vadd(w, vmul(ow, inv));
            if (l == 0) { SV(D, s) = f4(v, b.v4.w); SW(D, s) = f4(w, 0.0f); }
        }
        if (l == 0) integrate_slot(D, K, s, v, w);
    });
}
// The relax pass's end and the next sub-step's start: the overflow average, then the warm start, per body. Both
// sets of per-contact records are loaded in one pass over the body's list (they do not depend on each other), and
// each tree sums its records as ovf_average and warm_body do.
__global__ void __launch_bounds__(NT) k_ovf_apply_warm(Dev D, StepK K) {
    const V3 gh = vmul(v3(K.gx, K.gy, K.gz), K.h);
    body_pairs(D, [&](const BodyIn& b, int l, bool half) {
        const int s = b.s;
        const float4 v4 = b.v4;
        V3 v = v3f4(v4), w = v3f4(b.w4);
        V3 ov, ow, av, aw;
        list_sums(b.o, b.n, l, half, [&](uint32_t e, V3& x, V3& y) { const float4* d = OVL(D, e); x = v3f4(d[0]); y = v3f4(d[1]); }, ov, ow);
        list_sums(b.o, b.n, l, half, [&](uint32_t e, V3& x, V3& y) { const float4* d = WREC(D, e); x = v3f4(d[0]); y = v3f4(d[1]); }, av, aw);
        if (!s) return;
        if (b.cnt) {
            float inv = 1.0f / (float)b.cnt;
            v = vadd(v, vmul(ov, inv));
            w = vadd(w, vmul(ow, inv));
        }
        // the warm start (warm_body's operations)
        V3 vw = vadd(v, gh);
        float wl = vlen2(w);
        if (wl > K.wmax * K.wmax) w = vmul(w, K.wmax / sf_sqrt(wl));
        if (l == 0) {
            SV(D, s) = f4(vadd(vw, av), v4.w);
            SW(D, s) = f4(vadd(w, aw), 0.0f);
        }
    });
}
// ---------------------------------------------------------------- shock propagation and restitution
// One level pass (gpu_world.js:1178-1203): heights are read from rd and raised in wr, in two launches. Each pass's
// output is at least its input at every body a manifold touches, so carrying the read heights over (the maximum of
// the two copies) is a copy, k_level_copy; then, over the manifolds, the upper body of a contact within 45 degrees of
// vertical is lifted to one below its support's with atomic maxima (k_level_lift; a lift a plain load shows to be no
// higher is skipped: heights only rise), so the result does not depend on the order threads run in.
__global__ void __launch_bounds__(NT) k_level_copy(Dev D, const uint32_t* rd, uint32_t* wr) {
    GRID_STRIDE(s, (uint32_t)D.nb + 1u) wr[s] = rd[s];
}
__global__ void __launch_bounds__(NT) k_level_lift(Dev D, const uint32_t* rd, uint32_t* wr) {
    const uint32_t nm = min(D.ctr[C_NMAN], D.mcap) - D.ctr[C_NRED];  // (the reduced manifolds: the last positions)
    GRID_STRIDE(p, nm) {
        const float ny = MROW(D, D.cur, MR_N, p).y;
        if (!(ny > 0.7f) && !(ny < -0.7f)) continue;
        const int4 hd = D.cur.hdrp[p];
        const int sa = hd.z, sb = hd.w;
        const bool da = sa && SV(D, sa).w > 0.0f, db = sb && SV(D, sb).w > 0.0f;
        if (ny > 0.7f && db) {
            const uint32_t ha = da ? rd[sa] : 255u;
            if (ha > 1u && ha - 1u > wr[sb]) atomicMax(&wr[sb], ha - 1u);
        }
        if (ny < -0.7f && da) {
            const uint32_t hb = db ? rd[sb] : 255u;
            if (hb > 1u && hb - 1u > wr[sa]) atomicMax(&wr[sa], hb - 1u);
        }
    }
}

// Restitution after the sub-steps (gpu_world.js:1333-1363): in a manifold with a restitution that pushed during the
// step, the points that approached faster than the threshold at the step start have their normal velocity driven to
// -rest times that approach, the points iterated together. rest_live: whether manifold m (a position) takes part
// (loads only the flags and the points' approach speeds).
__device__ __forceinline__ bool rest_live(const Dev& D, uint32_t m, const StepK& K) {
    const ManBuf& C = D.cur;
    const float rest = MROW(D, C, MR_LAF, m).w, pushed = MROW(D, C, MR_LBF, m).w;
    if (!(rest > 0.0f) || pushed == 0.0f) return false;
    const int cnt = (int)MROW(D, C, MR_CB1, m).w;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; k++) if (k < cnt && !(MROW(D, C, MR_P0 + MR_PER_PT * k + 2, m).w > -K.rest_thr)) any = true;
    return any;
}
// fa, fb as in solve_core (mass splitting in the overflow), with the mass-ratio limit
__device__ __forceinline__ void rest_core(const Dev& D, uint32_t m, const StepK& K, V3& va, V3& wa, V3& vb, V3& wb,
                                          float fa, float fb, const M3& iia, const M3& iib) {
    const ManBuf& C = D.cur;
    const bool split = fa > 0.0f;
    const V3 va0 = va, wa0 = wa, vb0 = vb, wb0 = wb;
    const float4 n4 = MROW(D, C, MR_N, m), jn4 = MROW(D, C, MR_JN, m);
    const float ima0 = MROW(D, C, MR_CA1, m).w, imb0 = MROW(D, C, MR_CA2, m).w, rest = MROW(D, C, MR_LAF, m).w;
    const int cnt = (int)MROW(D, C, MR_CB1, m).w;
    float4 r0[4], r1[4];
    float rv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        r0[k] = MROW(D, C, MR_P0 + MR_PER_PT * k + 0, m);
        r1[k] = MROW(D, C, MR_P0 + MR_PER_PT * k + 1, m);
        rv[k] = MROW(D, C, MR_P0 + MR_PER_PT * k + 2, m).w;
    }
    float sca, scb;
    ratio_scale(split ? ima0 * fa : ima0, split ? imb0 * fb : imb0, K.ratio, sca, scb);
    const bool rec = split || sca != 1.0f || scb != 1.0f;
    const float xa = (split ? fa : 1.0f) * sca, xb = (split ? fb : 1.0f) * scb;
    const float ima = rec ? ima0 * xa : ima0, imb = rec ? imb0 * xb : imb0;
    const V3 n = v3f4(n4);
    float jn[4] = {jn4.x, jn4.y, jn4.z, jn4.w};
    V3 iaxn[4], ibxn[4];
    float nm[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const V3 raxn = v3f4(r0[k]), rbxn = v3f4(r1[k]);
        iaxn[k] = mulv(iia, raxn); ibxn[k] = mulv(iib, rbxn);
        if (rec) { iaxn[k] = vmul(iaxn[k], xa); ibxn[k] = vmul(ibxn[k], xb); }
        nm[k] = r1[k].w;
        if (rec) { float kk = ima + imb + vdot(raxn, iaxn[k]) + vdot(rbxn, ibxn[k]); nm[k] = kk > 0.0f ? 1.0f / kk : 0.0f; }
    }
    for (int it = 0; it < K.rest_iters; it++) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool on = k < cnt && !(rv[k] > -K.rest_thr);
            const V3 raxn = v3f4(r0[k]), rbxn = v3f4(r1[k]);
            const float vn = vdot(vsub(vb, va), n) + vdot(wb, rbxn) - vdot(wa, raxn);
            const float nj = sf_max(jn[k] - nm[k] * (vn + rest * rv[k]), 0.0f);
            const float imp = nj - jn[k];
            const V3 va1 = vsub(va, vmul(n, imp * ima)), wa1 = vsub(wa, vmul(iaxn[k], imp));
            const V3 vb1 = vadd(vb, vmul(n, imp * imb)), wb1 = vadd(wb, vmul(ibxn[k], imp));
            jn[k] = on ? nj : jn[k];
            va = vsel(on, va1, va); wa = vsel(on, wa1, wa); vb = vsel(on, vb1, vb); wb = vsel(on, wb1, wb);
        }
    }
    if (sca != 1.0f) { va = vadd(va0, slowing(va0, vsub(va, va0), sca)); wa = vadd(wa0, slowing(wa0, vsub(wa, wa0), sca)); }
    if (scb != 1.0f) { vb = vadd(vb0, slowing(vb0, vsub(vb, vb0), scb)); wb = vadd(wb0, slowing(wb0, vsub(wb, wb0), scb)); }
    MROW(D, C, MR_JN, m) = make_float4(jn[0], jn[1], jn[2], jn[3]);
}
__device__ void rest_one(const Dev& D, uint32_t m, const StepK& K) {
    if (!rest_live(D, m, K)) return;
    const int4 hd = D.cur.hdrp[m];
    const int sa = hd.z, sb = hd.w;
    const float4 va4 = SV(D, sa), vb4 = SV(D, sb);
    V3 va = v3f4(va4), wa = v3f4(SW(D, sa)), vb = v3f4(vb4), wb = v3f4(SW(D, sb));
    const M3 iia = load_m3(SII(D, sa)), iib = load_m3(SII(D, sb));
    rest_core(D, m, K, va, wa, vb, wb, 0.0f, 0.0f, iia, iib);
    if (sa) { SV(D, sa) = f4(va, va4.w); SW(D, sa) = f4(wa, 0.0f); }
    if (sb) { SV(D, sb) = f4(vb, vb4.w); SW(D, sb) = f4(wb, 0.0f); }
}
// one colour bucket of the restitution pass, spread over the XCDs as k_solve_bucket's
__global__ void __launch_bounds__(NT) k_rest_bucket(Dev D, StepK K, int c) {
    if (!D.ctr[C_NREST]) return;  // (no manifold of the step may bounce)
    const uint32_t n = D.colcnt[c], o = D.coloff[c];
    const uint32_t G = min(gridDim.x, max((n + NT - 1) / NT, 1u)), b = blockIdx.x;
    if (b >= G) return;
    const uint32_t x = b & 7u, q = G >> 3, r = G & 7u;
    const uint32_t lb = x * q + min(x, r) + (b >> 3);
    for (uint32_t i = lb * NT + threadIdx.x; i < n; i += G * NT) rest_one(D, o + i, K);
}
// the small colours from c0 on, one after another in one workgroup
__global__ void __launch_bounds__(NT) k_rest_tail(Dev D, StepK K, int c0) {
    if (!D.ctr[C_NREST]) return;
    int ncol = (int)D.ctr[C_NCOL];
    for (int c = c0; c < ncol; c++) {
        uint32_t n = D.colcnt[c], o = D.coloff[c];
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) rest_one(D, o + i, K);
        __syncthreads();
    }
}
// the overflow's restitution: Jacobi with mass splitting as its solve (every overflow manifold writes its changes,
// zero when it takes no part, and k_ovf_apply averages them)
__global__ void __launch_bounds__(NT) k_rest_ovf(Dev D, StepK K) {
    if (!D.ctr[C_NREST]) return;  // (then k_ovf_apply, gated the same way, applies nothing)
    const uint32_t n = D.colcnt[K.K], o = D.coloff[K.K];
    const uint32_t G = min(gridDim.x, max((n + NT - 1) / NT, 1u)), b = blockIdx.x;
    if (b >= G) return;
    const uint32_t x = b & 7u, q = G >> 3, r = G & 7u;
    const uint32_t lb = x * q + min(x, r) + (b >> 3);
    for (uint32_t i = lb * NT + threadIdx.x; i < n; i += G * NT) {
        const uint32_t m = o + i;
        const uint2 e2 = reinterpret_cast<const uint2*>(D.blinv)[m];
        const int4 hd = D.cur.hdrp[m];
        const int sa = hd.z, sb = hd.w;
        const uint32_t ea = sa ? e2.x : ~0u, eb = sb ? e2.y : ~0u;
        V3 dva = v3(0, 0, 0), dwa = v3(0, 0, 0), dvb = v3(0, 0, 0), dwb = v3(0, 0, 0);
        if (rest_live(D, m, K)) {
            const V3 va0 = v3f4(SV(D, sa)), wa0 = v3f4(SW(D, sa)), vb0 = v3f4(SV(D, sb)), wb0 = v3f4(SW(D, sb));
            V3 va = va0, wa = wa0, vb = vb0, wb = wb0;
            const float fa = sa ? (float)D.ovcnt[sa] : 1.0f, fb = sb ? (float)D.ovcnt[sb] : 1.0f;
            const M3 iia = load_m3(SII(D, sa)), iib = load_m3(SII(D, sb));
            rest_core(D, m, K, va, wa, vb, wb, fa, fb, iia, iib);
            dva = vsub(va, va0); dwa = vsub(wa, wa0); dvb = vsub(vb, vb0); dwb = vsub(wb, wb0);
        }
        if (ea != ~0u) { OVL(D, ea)[0] = f4(dva, 0.0f); OVL(D, ea)[1] = f4(dwa, 0.0f); }
        if (eb != ~0u) { OVL(D, eb)[0] = f4(dvb, 0.0f); OVL(D, eb)[1] = f4(dwb, 0.0f); }
    }
}

__global__ void __launch_bounds__(NT) k_ovf_count(Dev D, int K) {
    uint32_t nm = min(D.ctr[C_NMAN], D.mcap);
    GRID_STRIDE(m, nm) {
        if (D.mcol[m] != K) continue;
        int4 hd = D.cur.hdr[m];
        if (hd.z) atomicAdd(&D.ovcnt[hd.z], 1u);
        if (hd.w) atomicAdd(&D.ovcnt[hd.w], 1u);
    }
}

__global__ void __launch_bounds__(NT) k_integrate(Dev D, StepK K) {
    GRID_STRIDE(i, (uint32_t)D.nb) {
        int s = D.bslot[i];
        if (!s) continue;
        integrate_slot(D, K, s, v3f4(SV(D, s)), v3f4(SW(D, s)));
    }
}

// ---------------------------------------------------------------- write back, islands, sleep
__global__ void __launch_bounds__(NT) k_writeback(Dev D) {
    GRID_STRIDE(i, (uint32_t)D.nb) {
        int s = D.bslot[i];
        if (!s) continue;
        float4 p = D.bp[i];
        V3 np = vadd(v3f4(D.sp0[s]), v3f4(SDP(D, s)));
        D.bp[i] = f4(np, p.w);
        D.bv[i] = f4(v3f4(SV(D, s)), 0.0f);
        D.bw[i] = f4(v3f4(SW(D, s)), 0.0f);
        D.bq[i] = SQ(D, s);
    }
}

// ---------------------------------------------------------------- islands and sleep
// Union-find over the contact graph of moving bodies; an island's root is its smallest body index. Parents only move
// to smaller indices of the same island, so whatever a racing store leaves behind is still an ancestor. Loads and
// stores are device-scope relaxed atomics: coherent across the XCDs' L2s without the cost of system scope.
__device__ __forceinline__ int par_load(const int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void par_store(int32_t* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// root of x with path halving
__device__ __forceinline__ int uf_find(int32_t* par, int x) {
    while (true) {
        int p = par_load(&par[x]);
        if (p == x) return x;
        int g = par_load(&par[p]);
        if (g != p) par_store(&par[x], g);
        x = g;
    }
}

// Sleep state of a body (sst): an island sleeps when every member's rest timer has reached the sleep time. Bodies
// whose timer has not are "restless"; the others are "sleepy". An island with a restless member is connected, so each
// of its sleepy components (under the contacts between two sleepy bodies) has a contact with a restless body, and an
// island without one is a single sleepy component with no such contact. The step therefore unites only sleepy
// bodies and taints the components that touch a restless body: the same decisions as a minimum of the timers over
// full islands, without the union-find work (and its hot-root atomics) over islands that cannot sleep.
enum : uint32_t { SS_SLEEPY = 1u, SS_RESTLESS_NBR = 2u, SS_TAINT = 4u };

// rest timer: still = within a small distance and angle of the pose where the timer last restarted
__global__ void __launch_bounds__(NT) k_sleep_timer(Dev D, StepK K) {
    GRID_STRIDE(i, (uint32_t)D.nb) {
        if (!(D.bflag[i] & 1)) continue;
        Qt q = qf4(D.bq[i]);
        Qt dq = qmul(q, qconj(qf4(D.bsq[i])));
        float ang = 2.0f * sf_sqrt(dq.x * dq.x + dq.y * dq.y + dq.z * dq.z);
        V3 p = v3f4(D.bp[i]);
        float4 ii = D.bii[i];
        bool near = vlen2(vsub(p, v3f4(D.bsp[i]))) < K.sleep_dist2 && ang < K.sleep_rot;
        if (near && vlen2(v3f4(D.bv[i])) < 100.0f * K.sleep_lin2 && vlen2(v3f4(D.bw[i])) < 100.0f * K.sleep_ang2) ii.w += K.dt;
        else { ii.w = 0.0f; D.bsp[i] = D.bp[i]; D.bsq[i] = D.bq[i]; }
        D.bii[i] = ii;
        if (K.allow_sleep && ii.w >= K.sleep_time) D.sst[i] = SS_SLEEPY;
    }
}

// A first forest without atomics: each body points at its smallest-index neighbour across a contact between two
// moving bodies (a wavefront per body, min over its contact list). With all == 0 only sleepy bodies and sleepy
// neighbours take part, and a sleepy body with a restless neighbour is marked.
__global__ void __launch_bounds__(NT) k_uf_init(Dev D, int all) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t i = wave; i < (uint32_t)D.nb; i += nwaves) {
        int s = D.bslot[i];
        if (!s) continue;
        if (!all && !(D.sst[i] & SS_SLEEPY)) continue;
        uint32_t o = D.bloff[s], n = D.blcnt[s];
        int best = (int)i;
        bool restless = false;
        for (uint32_t e0 = 0; e0 < n; e0 += 64) {
            uint32_t e = e0 + lane;
            int nbr = (int)i;
            if (e < n) {
                uint32_t ent = D.blist[o + e];
                int4 hd = D.cur.hdr[ent >> 1];
                int other = (ent & 1u) ? hd.z : hd.w;
                if (hd.z && hd.w) {
                    if (all || (D.sst[other - 1] & SS_SLEEPY)) nbr = other - 1;
                    else restless = true;
                }
            }
#pragma unroll
            for (int h = 32; h >= 1; h >>= 1) nbr = min(nbr, __shfl_xor(nbr, h, 64));
            best = min(best, nbr);
        }
        bool any_restless = __ballot(restless) != 0ull;
        if (lane == 0) {
            D.parent[i] = best;
            if (any_restless) D.sst[i] |= SS_RESTLESS_NBR;
        }
    }
}

// pointer jumping: every body's parent replaced by its grandparent until it reaches a root (each store is an
// ancestor, so the lanes may run it concurrently)
__global__ void __launch_bounds__(NT) k_uf_jump(Dev D, int all) {
    GRID_STRIDE(i, (uint32_t)D.nb) {
        if (!all && !(D.sst[i] & SS_SLEEPY)) continue;
        int p = par_load(&D.parent[i]);
        for (int it = 0; it < 64; it++) {
            int g = par_load(&D.parent[p]);
            if (g == p) break;
            p = g;
            par_store(&D.parent[i], p);
        }
    }
}

// hook the larger root under the smaller, for every contact between two moving (all == 0: two sleepy) bodies
__global__ void __launch_bounds__(NT) k_uf_hook(Dev D, int all) {
    uint32_t nm = min(D.ctr[C_NMAN], D.mcap);
    GRID_STRIDE(m, nm) {
        if (D.mcol[m] == -2) continue;
        int4 hd = D.cur.hdr[m];
        if (hd.z == 0 || hd.w == 0) continue;
        int a = hd.z - 1, b = hd.w - 1;
        if (!all && !(D.sst[a] & D.sst[b] & SS_SLEEPY)) continue;
        while (true) {
            a = uf_find(D.parent, a);
            b = uf_find(D.parent, b);
            if (a == b) break;
            if (a < b) { int t = a; a = b; b = t; }
            // lower the larger root's parent; if it had been hooked elsewhere meanwhile, join that parent next
            int old = atomicMin(&D.parent[a], b);
            if (old == a) break;
            a = old;
        }
    }
}

// every body's parent made its root (all == 0: sleepy bodies only, each carrying its restless mark to the root)
__global__ void __launch_bounds__(NT) k_uf_flatten(Dev D, int all) {
    GRID_STRIDE(i, (uint32_t)D.nb) {
        uint32_t ss = all ? 0u : D.sst[i];
        if (!all && !(ss & SS_SLEEPY)) continue;
        int r = uf_find(D.parent, (int)i);
        D.parent[i] = r;
        if ((ss & SS_RESTLESS_NBR) && !(__hip_atomic_load(&D.sst[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & SS_TAINT))
            atomicOr(&D.sst[r], SS_TAINT);
    }
}

__global__ void __launch_bounds__(NT) k_uf_reset(Dev D) { GRID_STRIDE(i, (uint32_t)D.nb) D.parent[i] = (int32_t)i; }

__global__ void __launch_bounds__(NT) k_sleep_apply(Dev D) {
    GRID_STRIDE(i, (uint32_t)D.nb) {
        if (!(D.sst[i] & SS_SLEEPY)) continue;
        if (!(D.sst[D.parent[i]] & SS_TAINT)) {
            D.bflag[i] &= ~1;
            D.bv[i] = make_float4(0, 0, 0, 0);
            D.bw[i] = make_float4(0, 0, 0, 0);
        }
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) { for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v; }
// the step's counters and per-step tables cleared in one launch
__global__ void __launch_bounds__(NT) k_step_clear(Dev D, int K) {
    const size_t nb1 = (size_t)D.nb + 1, ncl = nb1 * (size_t)K, n = max((size_t)D.tcap, ncl);
    for (size_t i = blockIdx.x * (size_t)NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT) {
        if (i < (size_t)C_NCTR) D.ctr[i] = 0u;
        if (i < (size_t)D.tcap) { D.tcnt[i] = 0u; D.tfill[i] = 0u; }
        if (i < nb1) { D.blcnt[i] = 0u; D.blfill[i] = 0u; D.bmask[i] = 0ull; D.ovcnt[i] = 0u; D.lvl[i] = 0u; D.lvl2[i] = 0u; }
        if (i < ncl) D.claim[i] = 0ull;
        if (i < 72) { D.colcnt[i] = 0u; D.colfill[i] = 0u; }
    }
}
__global__ void k_fill_u64(unsigned long long* p, size_t n) { for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull; }

// ---------------------------------------------------------------- particle debris (gpu/gpu_world.js:1671-2038)
// Position-based debris, after the body step: every sub-step predicts, pushes each moving particle out of the other
// particles (Jacobi: the pushes averaged and over-relaxed, against the predicted positions of all), then out of the
// ground and every candidate shape in turn, each in full, with Coulomb friction from the push; the velocity is the
// move over the sub-step. Particles never push bodies. A particle at rest for 0.5 s sleeps; a blast, a wake request
// or a moving shape near it wakes it. Candidates are gathered once per step: particles from a hashed grid (each
// slot's particles in index order, the 27 cells round the particle's in z-y-x order, the first 24 in reach), shapes
// from this step's cell table (the 12 lowest indices whose step-start box meets the particle's reach). Shapes are
// seen at their step-start pose, their bodies' velocities at the end of the step. One lane per particle; records
// are struct-of-arrays of float4, candidate lists stored by rank ([k][capacity]) so a wavefront's loads coalesce.
constexpr int P_NN = 24, P_NS = 12;
struct PDev {
    int np, cap;
    uint32_t G, gmask;
    float4 *x, *v, *q, *s, *w, *c, *a;  // x|r, v|state, q, xs|t, w|mu, xc|touch, anchor|-
    uint32_t *gl, *gn;                  // on the ground (directly or through grounded particles): last sub-step, this one
    int2* cnt;                          // neighbours, shapes
    int32_t *nb, *sh;                   // [P_NN][cap], [P_NS][cap]
    uint32_t* cell;                     // grid slot of each particle
    uint32_t *gcnt, *gstart, *gfill;    // [G]
    int32_t* list;                      // [cap]
    uint32_t* ctr;                      // moving at the stage's start, asleep at its end
};
struct PStepK {
    float dt, h, pc, mug, gx, gy, gz, cell;
    int psub, nblast, wake;
    float4 bc[3];  // blast centre | reach
    float bs[3];   // blast speed
};

__device__ __forceinline__ uint32_t p_hash(int x, int y, int z, uint32_t mask) {
    return (((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u)) & mask;
}
__device__ __forceinline__ int p_cellc(float v, float pc) { return (int32_t)sf_floor(v / pc); }
__device__ __forceinline__ V3 p_fric(V3 disp, V3 n, float pen, float mu) {
    const V3 t = vsub(disp, vmul(n, vdot(disp, n)));
    const float tl = vlen(t);
    if (tl < 1e-9f) return v3(0, 0, 0);
    return vmul(t, -sf_min(mu * pen / tl, 1.0f));
}
__device__ __forceinline__ V3 p_capv(V3 v) { const float l = vlen(v); return l > 40.0f ? vmul(v, 40.0f / l) : v; }
// Each shape of this step's cell table whose step-start box meets [lo, hi], once (a shape sits in every cell its box
// covers; it is visited in the first of the box's cells it shares): f(shape).
template <class F>
__device__ __forceinline__ void p_table_shapes(const Dev& D, const PStepK& K, V3 lo, V3 hi, F f) {
    const float ic = 1.0f / K.cell;
    const int x0 = (int32_t)sf_floor(lo.x * ic), y0 = (int32_t)sf_floor(lo.y * ic), z0 = (int32_t)sf_floor(lo.z * ic);
    const int x1 = (int32_t)sf_floor(hi.x * ic), y1 = (int32_t)sf_floor(hi.y * ic), z1 = (int32_t)sf_floor(hi.z * ic);
    for (int x = x0; x <= x1; x++) for (int y = y0; y <= y1; y++) for (int z = z0; z <= z1; z++) {
        const uint32_t key = cell_key(x, y, z), sl = slot_of(key, D.tmask);
        const uint32_t e0 = D.tstart[sl], e1 = e0 + D.tcnt[sl];
        for (uint32_t e = e0; e < e1; e++) {
            if (D.tkey[e] != key) continue;
            const int s = D.tent[e];
            const float4 slo = D.slo[s], shi = D.shi[s];
            if (slo.x > hi.x || lo.x > shi.x || slo.y > hi.y || lo.y > shi.y || slo.z > hi.z || lo.z > shi.z) continue;
            // the first cell of [lo, hi] the shape's box covers: visit it there only
            const int fx = max(x0, (int32_t)sf_floor(slo.x * ic)), fy = max(y0, (int32_t)sf_floor(slo.y * ic)), fz = max(z0, (int32_t)sf_floor(slo.z * ic));
            if (fx != x || fy != y || fz != z) continue;
            f(s);
        }
    }
}
// a moving shape (awake body with mass, as of now) whose step-start box meets [lo, hi]
__device__ __forceinline__ bool p_near_moving(const Dev& D, const PStepK& K, V3 lo, V3 hi) {
    bool found = false;
    p_table_shapes(D, K, lo, hi, [&](int s) {
        const int b = D.shb[s];
        if ((D.bflag[b] & 1) && D.bp[b].w > 0.0f) found = true;
    });
    return found;
}

__global__ void __launch_bounds__(NT) k_p_prep(Dev D, PDev P, PStepK K) {
    uint32_t nmov = 0;
    GRID_STRIDE(i, (uint32_t)P.np) {
        float4 v4 = P.v[i];
        if (v4.w == 0.0f || v4.w >= 3.0f) continue;
        const float4 x4 = P.x[i];
        const V3 x = v3f4(x4);
        V3 v = v3f4(v4);
        bool hit = false;
        for (int b = 0; b < K.nblast; b++) {
            const V3 d = vsub(x, v3f4(K.bc[b]));
            const float l = vlen(d);
            if (l < K.bc[b].w) { hit = true; v = vadd(v, vmul(vadd(vmul(d, 1.0f / sf_max(l, 1e-3f)), v3(0.0f, 0.3f, 0.0f)), K.bs[b] * (1.0f - l / K.bc[b].w))); }
        }
        if (v4.w == 2.0f && (hit || K.wake || p_near_moving(D, K, vsub(x, v3(0.5f, 0.5f, 0.5f)), vadd(x, v3(0.5f, 0.5f, 0.5f))))) {
            v4.w = 1.0f;
            const float4 s4 = P.s[i];
            P.s[i] = make_float4(s4.x, s4.y, s4.z, 0.0f);
        }
        if (v4.w == 1.0f) { nmov++; P.v[i] = f4(v, 1.0f); }
        const uint32_t hc = p_hash(p_cellc(x.x, K.pc), p_cellc(x.y, K.pc), p_cellc(x.z, K.pc), P.gmask);
        P.cell[i] = hc;
        atomicAdd(&P.gcnt[hc], 1u);
    }
    block_count(nmov, &P.ctr[0]);
}
__global__ void __launch_bounds__(NT) k_p_scatter(PDev P) {
    GRID_STRIDE(i, (uint32_t)P.np) {
        const float st = P.v[i].w;
        if (st == 0.0f || st >= 3.0f) continue;
        const uint32_t hc = P.cell[i];
        P.list[P.gstart[hc] + atomicAdd(&P.gfill[hc], 1u)] = (int32_t)i;
    }
}
// each grid slot's particles in index order (the lists are a few long: insertion sort, a thread per slot)
__global__ void __launch_bounds__(NT) k_p_sort(PDev P) {
    GRID_STRIDE(hc, P.G) {
        const uint32_t s0 = P.gstart[hc], n = P.gcnt[hc];
        for (uint32_t a = 1; a < n; a++) {
            const int32_t v = P.list[s0 + a];
            uint32_t b = a;
            while (b > 0 && P.list[s0 + b - 1] > v) { P.list[s0 + b] = P.list[s0 + b - 1]; b--; }
            P.list[s0 + b] = v;
        }
    }
}
__global__ void __launch_bounds__(NT) k_p_gather(Dev D, PDev P, PStepK K) {
    GRID_STRIDE(i, (uint32_t)P.np) {
        const float4 v4 = P.v[i];
        if (v4.w != 1.0f) continue;
        const float4 x4 = P.x[i];
        const V3 xi = v3f4(x4);
        const float ri = x4.w;
        const float m = vlen(v3f4(v4)) * K.dt + 0.02f;
        const int cx = p_cellc(xi.x, K.pc), cy = p_cellc(xi.y, K.pc), cz = p_cellc(xi.z, K.pc);
        int nn = 0;
        for (int k = 0; k < 27; k++) {
            const int z = cz - 1 + k / 9, y = cy - 1 + (k / 3) % 3, x = cx - 1 + k % 3;
            const uint32_t hc = p_hash(x, y, z, P.gmask);
            bool dup = false;  // (a slot already visited through another cell)
            for (int k2 = 0; k2 < k; k2++) dup |= p_hash(cx - 1 + k2 % 3, cy - 1 + (k2 / 3) % 3, cz - 1 + k2 / 9, P.gmask) == hc;
            if (dup) continue;
            const uint32_t e0 = P.gstart[hc], e1 = e0 + P.gcnt[hc];
            for (uint32_t e = e0; e < e1; e++) {
                const int j = P.list[e];
                if (j == (int)i) continue;
                const float4 xj = P.x[j];
                const float rr = ri + xj.w + m + vlen(v3f4(P.v[j])) * K.dt + 0.01f;
                const V3 d = vsub(xi, v3f4(xj));
                if (vdot(d, d) < rr * rr && nn < P_NN) { P.nb[(size_t)nn * P.cap + i] = j; nn++; }
            }
        }
        // the 12 lowest shape indices whose step-start box meets the particle's reach (kept sorted in registers)
        const float e = ri + m + 0.01f;
        int best[P_NS];
#pragma unroll
        for (int k = 0; k < P_NS; k++) best[k] = 0x7fffffff;
        p_table_shapes(D, K, vsub(xi, v3(e, e, e)), vadd(xi, v3(e, e, e)), [&](int s) {
            if (s >= best[P_NS - 1]) return;
            int c = s;
#pragma unroll
            for (int k = 0; k < P_NS; k++) { const int lo = min(c, best[k]), hi = max(c, best[k]); best[k] = lo; c = hi; }
        });
        int ns = 0;
#pragma unroll
        for (int k = 0; k < P_NS; k++) if (best[k] != 0x7fffffff) { P.sh[(size_t)k * P.cap + i] = best[k]; ns = k + 1; }
        P.cnt[i] = make_int2(nn, ns);
    }
}
__global__ void __launch_bounds__(NT) k_p_integrate(PDev P, PStepK K) {
    GRID_STRIDE(i, (uint32_t)P.np) {
        const float4 v4 = P.v[i];
        if (v4.w != 1.0f) continue;
        const V3 v = p_capv(vadd(v3f4(v4), vmul(v3(K.gx, K.gy, K.gz), K.h)));
        const float4 x4 = P.x[i];
        P.s[i] = make_float4(x4.x, x4.y, x4.z, P.s[i].w);
        P.x[i] = f4(vadd(v3f4(x4), vmul(v, K.h)), x4.w);
        P.v[i] = f4(v, 1.0f);
    }
}
__global__ void __launch_bounds__(NT) k_p_contact(Dev D, PDev P, PStepK K) {
    GRID_STRIDE(i, (uint32_t)P.np) {
        if (P.v[i].w != 1.0f) continue;
        const float4 x4 = P.x[i];
        const V3 xi = v3f4(x4), xs = v3f4(P.s[i]);
        const float ri = x4.w, mui = P.w[i].w;
        const int2 cn = P.cnt[i];
        const float cap = 2.0f * K.h;
        V3 corr = v3(0, 0, 0);
        float n = 0.0f, touch = 0.0f;
        bool bearing = false, grounded = false;
        for (int q = 0; q < cn.x; q++) {
            const int j = P.nb[(size_t)q * P.cap + i];
            const float4 xj = P.x[j];
            const V3 d = vsub(xi, v3f4(xj));
            const float d2 = vdot(d, d), rr = ri + xj.w;
            if (d2 >= (rr + 0.01f) * (rr + 0.01f) || d2 < 1e-12f) continue;
            const float dist = sf_sqrt(d2);
            const V3 nn = vmul(d, 1.0f / dist);
            if (P.gl[j] && nn.y > 0.5f) grounded = true;
            touch = 1.0f;
            if (dist >= rr) continue;
            const bool om = P.v[j].w == 1.0f;
            const float wgt = om ? 0.5f : 1.0f;
            const V3 dj = om ? vsub(v3f4(xj), v3f4(P.s[j])) : v3(0, 0, 0);
            const V3 rel = vsub(vsub(xi, xs), dj);
            const float pen = sf_min((rr - dist) * wgt, (sf_max(-vdot(rel, nn), 0.0f) + cap) * wgt);
            corr = vadd(corr, vadd(vmul(nn, pen), vmul(p_fric(rel, nn, pen, mui), wgt)));
            n += 1.0f;
        }
        V3 x = vadd(xi, vmul(corr, n > 1.5f ? 1.5f / n : 1.0f));
        const float pg0 = ri - x.y;
        if (pg0 > -0.01f) { touch = 1.0f; grounded = true; }
        if (pg0 > 0.0f) {
            const V3 nn = v3(0.0f, 1.0f, 0.0f);
            x = vadd(x, vadd(vmul(nn, pg0), p_fric(vsub(x, xs), nn, pg0, 0.5f * (mui + K.mug))));
        }
        for (int q = 0; q < cn.y; q++) {
            const int sj = P.sh[(size_t)q * P.cap + i];
            const int b = D.shb[sj];
            const V3 wp = v3f4(D.swp[sj]);
            const M3 Rm = load_m3(D.swr + 3 * (size_t)sj);
            const float4 hh4 = D.shh[sj];
            const V3 hh = v3f4(hh4);
            const V3 lc = tmulv(Rm, vsub(x, wp));
            const V3 cl = v3(sf_min(sf_max(lc.x, -hh.x), hh.x), sf_min(sf_max(lc.y, -hh.y), hh.y), sf_min(sf_max(lc.z, -hh.z), hh.z));
            const V3 df = vsub(lc, cl);
            const float d2 = vdot(df, df);
            V3 nl;
            float pen;
            if (d2 > 1e-12f) {
                const float dist = sf_sqrt(d2);
                if (dist >= ri + 0.01f) continue;
                touch = 1.0f;
                if (dist >= ri) continue;
                nl = vmul(df, 1.0f / dist); pen = ri - dist;
            } else {
                const V3 dd = v3(hh.x - sf_abs(lc.x), hh.y - sf_abs(lc.y), hh.z - sf_abs(lc.z));
                if (dd.x <= dd.y && dd.x <= dd.z) { nl = v3(lc.x >= 0.0f ? 1.0f : -1.0f, 0.0f, 0.0f); pen = dd.x + ri; }
                else if (dd.y <= dd.z) { nl = v3(0.0f, lc.y >= 0.0f ? 1.0f : -1.0f, 0.0f); pen = dd.y + ri; }
                else { nl = v3(0.0f, 0.0f, lc.z >= 0.0f ? 1.0f : -1.0f); pen = dd.z + ri; }
            }
            const V3 nn = mulv(Rm, nl);
            const float4 bp = D.bp[b];
            const V3 vb = vadd(v3f4(D.bv[b]), vcross(v3f4(D.bw[b]), vsub(x, v3f4(bp))));
            const V3 vbh = vmul(vb, K.h);
            const float push = sf_min(pen, sf_max(vdot(vsub(vbh, vsub(x, xs)), nn), 0.0f) + cap);
            const float mu = 0.5f * (mui + hh4.w);
            x = vadd(x, vadd(vmul(nn, push), p_fric(vsub(vsub(x, xs), vbh), nn, push, mu)));
            touch = 1.0f;
            const bool fixed = !(bp.w > 0.0f) || !(D.bflag[b] & 1);
            if (fixed && nn.y > 0.5f) grounded = true;
            if (!fixed && nn.y < -0.3f) bearing = true;
        }
        P.c[i] = f4(x, bearing ? 2.0f : touch);
        P.gn[i] = grounded ? 1u : 0u;
    }
}
__global__ void __launch_bounds__(NT) k_p_apply(PDev P, PStepK K) {
    GRID_STRIDE(i, (uint32_t)P.np) {
        if (P.v[i].w != 1.0f) continue;
        P.gl[i] = P.gn[i];
        const float4 c4 = P.c[i], x4 = P.x[i];
        P.v[i] = f4(p_capv(vmul(vsub(v3f4(c4), v3f4(P.s[i])), 1.0f / K.h)), 1.0f);
        P.x[i] = f4(v3f4(c4), x4.w);
    }
}
__global__ void __launch_bounds__(NT) k_p_finish(Dev D, PDev P, PStepK K) {
    uint32_t nasleep = 0;
    GRID_STRIDE(i, (uint32_t)P.np) {
        float4 v4 = P.v[i];
        if (v4.w == 1.0f) {
            const float4 x4 = P.x[i], s4 = P.s[i], w4 = P.w[i], a4 = P.a[i];
            const float touch = P.c[i].w, r = x4.w;
            const V3 x = v3f4(x4);
            V3 v = v3f4(v4), ww = v3f4(w4);
            float t = 0.0f, st = 1.0f;
            if (touch > 0.0f) ww = vmul(vcross(v3(0.0f, 1.0f, 0.0f), v), 1.0f / sf_max(r, 0.01f));
            const float dr = vlen(vsub(x, v3f4(a4)));
            if (touch < 1.5f && dr < 0.01f) t = s4.w + K.dt;
            else if (touch < 1.5f && dr < 2.0f * r && !p_near_moving(D, K, vsub(x, v3(r, r, r)), vadd(x, v3(r, r, r)))) t = s4.w + 0.25f * K.dt;
            else P.a[i] = f4(x, 0.0f);
            if (t > 0.5f || x.y < -50.0f) { st = 2.0f; v = v3(0, 0, 0); ww = v3(0, 0, 0); }
            P.q[i] = f4q(qintegrate(qf4(P.q[i]), ww, K.dt));
            v4 = f4(v, st);
            P.v[i] = v4;
            P.s[i] = make_float4(s4.x, s4.y, s4.z, t);
            P.w[i] = f4(ww, w4.w);
        }
        if (v4.w == 2.0f) nasleep++;
    }
    block_count(nasleep, &P.ctr[1]);
}

// ---------------------------------------------------------------- host
struct HostWorld {
    int dev = 0;
    long steps = 0;   // steps since the last load
    hipStream_t st = nullptr;
    int body_cap = 0, shape_cap = 0;
    float prm[SF_NPARAM];
    bool any_rest = false;  // some loaded shape has a restitution above zero
    bool any_hull = false;  // some loaded shape is a convex chunk
    Dev D{};
    std::vector<void*> allocs;
    uint32_t nold = 0;
    int32_t stats[SF_NSTAT] = {0};
    std::vector<float> pose;
    uint32_t* scan_tmp = nullptr;
    uint32_t* h_ctr = nullptr;     // pinned
    uint32_t* h_col = nullptr;     // pinned colour counts
    bool timing = false;
    // the sub-step launch chain as a captured graph, replayed while its launches keep their shape (the key: the
    // device tables, the step constants, the colour plan and each launch's grid, sized with headroom)
    // (two kept: the manifold buffers swap every step, so consecutive steps alternate between two device tables)
    struct SubGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<uint8_t> key;
        std::vector<unsigned> grid;  // per big colour, then the overflow pass
        int64_t used = 0;
        void clear() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr; graph = nullptr; key.clear(); grid.clear();
        }
    };
    bool use_graph = true;
    SubGraph sg[2];
    int64_t graph_captures = 0, graph_replays = 0;
    // live timing of the solve-bucket kernel: an event pair around each launch
    bool ktime = false;
    std::vector<hipEvent_t> kev;
    size_t kev_used = 0;
    double ksolve_ms = 0.0;
    int64_t ksolve_launches = 0, ksolve_items = 0;
    hipEvent_t ev[6] = {};
    float stage_ms[5] = {0};
    std::vector<int32_t> h_shb;
    std::vector<float> h_h;  // half extents of the first shape of each body (pose())
    // particle debris: device records, free slots (reused last in, first out), queued blasts, wake request, stats
    PDev P{};
    uint32_t Gmax = 0;
    uint32_t* pscan = nullptr;
    std::vector<void*> pallocs;
    std::vector<int32_t> pfree;
    std::vector<std::array<float, 5>> pblast;
    bool pwake = false;
    uint32_t* h_pctr = nullptr;  // pinned: moving, asleep
    int32_t pstats[4] = {0, 0, 0, 0};
    hipEvent_t pev[2] = {};      // around the particle stage, when timing is on
    bool ptimed = false;
    float pstage_ms = 0.0f;

    ~HostWorld() {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : allocs) (void)hipFree(p);
        for (void* p : pallocs) (void)hipFree(p);
        if (h_pctr) (void)hipHostFree(h_pctr);
        if (h_ctr) (void)hipHostFree(h_ctr);
        if (h_col) (void)hipHostFree(h_col);
        for (auto& e : ev) if (e) (void)hipEventDestroy(e);
        for (auto& e : pev) if (e) (void)hipEventDestroy(e);
        for (auto& e : kev) (void)hipEventDestroy(e);
        for (auto& g : sg) g.clear();
        if (st) (void)hipStreamDestroy(st);
    }
    template <class T> int alloc(T*& p, size_t n) {
        void* q = nullptr;
        SF_HIP(hipMalloc(&q, sizeof(T) * (n ? n : 1)));
        SF_HIP(hipMemset(q, 0, sizeof(T) * (n ? n : 1)));
        allocs.push_back(q);
        p = (T*)q;
        return 0;
    }
};

void default_params(float* p) {
    p[SF_P_GX] = 0.0f; p[SF_P_GY] = -9.81f; p[SF_P_GZ] = 0.0f; p[SF_P_SUBSTEPS] = 12; p[SF_P_HERTZ] = 0.0f;
    p[SF_P_DAMPING] = 10.0f; p[SF_P_MAXPUSH] = 3.0f; p[SF_P_FRICTION] = 0.6f; p[SF_P_SPEC] = 0.02f;
    p[SF_P_SLEEP_LIN] = 0.05f; p[SF_P_SLEEP_ANG] = 0.08f; p[SF_P_SLEEP_TIME] = 0.5f; p[SF_P_SLEEP_DIST] = 0.01f;
    p[SF_P_SLEEP_ROT] = 0.02f; p[SF_P_ALLOW_SLEEP] = 1.0f; p[SF_P_CELL] = 1.0f; p[SF_P_LEAK] = 1.0f;
    p[SF_P_ANCHOR] = 1.0f; p[SF_P_ANCHOR_PUSH] = 3.0f; p[SF_P_ANCHOR_SLOP] = 0.0f; p[SF_P_COL_LIMIT] = 16;
    p[SF_P_COL_ROUNDS] = 32; p[SF_P_COL_SPREAD] = 4; p[SF_P_TAIL_MIN] = 1024; p[SF_P_SOLVE_NT] = 64;
    p[SF_P_RATIO] = 0.0f; p[SF_P_SHOCK] = 0.0f; p[SF_P_REST_THRESHOLD] = 1.0f; p[SF_P_REST_ITERS] = 8.0f;
    p[SF_P_COL_SUPPORT] = 0.0f; p[SF_P_PSUB] = 4.0f; p[SF_P_PCELL] = 0.7f; p[SF_P_RED_K] = 0.0f;
}

int world_alloc(HostWorld* w) {
    Dev& D = w->D;
    int nb = w->body_cap, ns = w->shape_cap;
    D.ecap = (uint32_t)std::min<uint64_t>(8ull * ns + 4096, 0x7FFFFFFFull);
    uint32_t t = 1; while (t < 2 * D.ecap) t <<= 1;
    D.tcap = t; D.tmask = t - 1;
    D.ccap = (uint32_t)(12ull * ns + 4096);
    D.mcap = D.ccap;
    size_t n1 = (size_t)nb + 1;
    int r = 0;
    r |= w->alloc(D.bp, nb); r |= w->alloc(D.bq, nb); r |= w->alloc(D.bv, nb); r |= w->alloc(D.bw, nb);
    r |= w->alloc(D.bii, nb); r |= w->alloc(D.bsp, nb); r |= w->alloc(D.bsq, nb); r |= w->alloc(D.brot, 3 * (size_t)nb);
    r |= w->alloc(D.bflag, nb); r |= w->alloc(D.bslot, nb); r |= w->alloc(D.wake, nb);
    r |= w->alloc(D.sbod, SLOT_F4 * n1); r |= w->alloc(D.sp0, n1);
    r |= w->alloc(D.shb, ns); r |= w->alloc(D.slp, ns); r |= w->alloc(D.slr, 3 * (size_t)ns); r |= w->alloc(D.shh, ns);
    r |= w->alloc(D.swp, ns); r |= w->alloc(D.swr, 3 * (size_t)ns); r |= w->alloc(D.slo, ns); r |= w->alloc(D.shi, ns);
    r |= w->alloc(D.srest, ns); r |= w->alloc(D.lvl, n1); r |= w->alloc(D.lvl2, n1);
    r |= w->alloc(D.shull, ns);
    r |= w->alloc(D.ecnt, ns); r |= w->alloc(D.eoff, ns); r |= w->alloc(D.ekey, D.ecap); r |= w->alloc(D.eshape, D.ecap);
    r |= w->alloc(D.tcnt, D.tcap); r |= w->alloc(D.tstart, D.tcap); r |= w->alloc(D.tfill, D.tcap); r |= w->alloc(D.tent, D.ecap); r |= w->alloc(D.tkey, D.ecap);
    r |= w->alloc(D.tidx, D.ecap); r |= w->alloc(D.soff, D.tcap); r |= w->alloc(D.slist, D.ecap);
    r |= w->alloc(D.epc, D.ecap); r |= w->alloc(D.erel, D.ecap); r |= w->alloc(D.ekeep, (size_t)ENTRY_KEEP * D.ecap);
    r |= w->alloc(D.bigc, (size_t)ns); r |= w->alloc(D.bigl, (size_t)nb + 1);
    r |= w->alloc(D.ccnt, ns); r |= w->alloc(D.coff, ns); r |= w->alloc(D.cand, D.ccap); r |= w->alloc(D.nar, 5 * (size_t)D.ccap);
    r |= w->alloc(D.hit, D.ccap); r |= w->alloc(D.hoff, D.ccap); r |= w->alloc(D.hlist, D.ccap);
    r |= w->alloc(D.nar2, 5 * (size_t)ns);
    {
        // the chunk library's tables, once per world
        const sf::HullTables& T = sf::hull_tables();
        float4 *v = nullptr, *pl = nullptr, *dir = nullptr;
        uint8_t *fidx = nullptr, *fcnt = nullptr, *seg = nullptr;
        int4* head = nullptr;
        r |= w->alloc(v, T.v4.size() / 4); r |= w->alloc(pl, T.pl4.size() / 4); r |= w->alloc(dir, T.dir4.size() / 4);
        r |= w->alloc(fidx, T.fidx.size()); r |= w->alloc(fcnt, T.fcnt.size()); r |= w->alloc(seg, T.seg.size());
        r |= w->alloc(head, T.head.size() / 4);
        if (r) return r;
        SF_HIP(hipMemcpy(v, T.v4.data(), sizeof(float) * T.v4.size(), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(pl, T.pl4.data(), sizeof(float) * T.pl4.size(), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(dir, T.dir4.data(), sizeof(float) * T.dir4.size(), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(fidx, T.fidx.data(), T.fidx.size(), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(fcnt, T.fcnt.data(), T.fcnt.size(), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(seg, T.seg.data(), T.seg.size(), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(head, T.head.data(), sizeof(int32_t) * T.head.size(), hipMemcpyHostToDevice));
        D.hl = HullDev{v, pl, dir, fidx, fcnt, seg, head};
    }
    for (ManBuf* M : {&D.cur, &D.old}) {
        r |= w->alloc(M->hdr, D.mcap); r |= w->alloc(M->row, (size_t)MR_ROWS * D.mcap); r |= w->alloc(M->key, D.mcap);
        r |= w->alloc(M->pos, D.mcap); r |= w->alloc(M->hdrp, D.mcap);
    }
    r |= w->alloc(D.wrec, (size_t)2 * WREC_F4 * D.mcap);  // two list entries per manifold
    r |= w->alloc(D.mcol, D.mcap); r |= w->alloc(D.mcand, D.mcap); r |= w->alloc(D.mpri, D.mcap);
    r |= w->alloc(D.bmask, n1); r |= w->alloc(D.claim, n1 * 64);
    r |= w->alloc(D.colcnt, 72); r |= w->alloc(D.coloff, 72); r |= w->alloc(D.colfill, 72);
    r |= w->alloc(D.chunkc, NBIN * (size_t)((D.mcap + NT - 1) / NT + 1));
    r |= w->alloc(D.ptpart, 4096);  // k_build's grid: sf_grid(ccap) is at most 4096 workgroups
    r |= w->alloc(D.mbin, D.mcap); r |= w->alloc(D.subcnt, OVS + 1); r |= w->alloc(D.suboff, OVS + 1);
    {
        uint32_t rt = 1; while (rt < 2 * D.mcap) rt <<= 1;
        D.rmask = rt - 1;
        r |= w->alloc(D.rkey, rt); r |= w->alloc(D.rcnt, rt); r |= w->alloc(D.rbig, rt);
        r |= w->alloc(D.rslot, D.mcap); r |= w->alloc(D.rflag, D.mcap);
        r |= w->alloc(D.rext, (size_t)RED_D * (D.mcap / (RED_KMIN + 1) + 1));  // (a reduced pair has more than RED_KMIN)
        if (!r) SF_HIP(hipMemset(D.rkey, 0xFF, sizeof(unsigned long long) * rt));  // (free slots hold ~0)
    }
    r |= w->alloc(D.ovcnt, n1);
    r |= w->alloc(D.blcnt, n1); r |= w->alloc(D.bloff, n1); r |= w->alloc(D.blfill, n1); r |= w->alloc(D.blist, 2 * (size_t)D.mcap); r |= w->alloc(D.blinv, 2 * (size_t)D.mcap); r |= w->alloc(D.minv, 2 * (size_t)D.mcap);
    r |= w->alloc(D.parent, nb); r |= w->alloc(D.sst, nb); r |= w->alloc(D.ctr, C_NCTR);
    uint32_t big = std::max<uint32_t>(std::max<uint32_t>(D.tcap, D.ccap), (uint32_t)n1);
    r |= w->alloc(w->scan_tmp, sf::scan_scratch_words(big));
    if (r) return r;
    SF_HIP(hipHostMalloc((void**)&w->h_ctr, sizeof(uint32_t) * C_NCTR, hipHostMallocDefault));
    SF_HIP(hipHostMalloc((void**)&w->h_col, sizeof(uint32_t) * 72, hipHostMallocDefault));
    for (auto& e : w->ev) SF_HIP(hipEventCreate(&e));
    // the zero fills ran on the null stream, which the world's non-blocking stream does not wait for
    SF_HIP(hipDeviceSynchronize());
    return 0;
}

StepK make_step(const float* P, float dt) {
    StepK K{};
    int sub = (int)P[SF_P_SUBSTEPS]; if (sub < 1) sub = 1;
    float h = dt / (float)sub;
    float inv_h = 1.0f / h;
    float hz = P[SF_P_HERTZ] > 0.0f ? P[SF_P_HERTZ] : inv_h / 6.0f;
    float omega = 2.0f * 3.14159265358979323846f * hz;
    float a1 = 2.0f * P[SF_P_DAMPING] + h * omega;
    float a2 = h * omega * a1;
    float a3 = 1.0f / (1.0f + a2);
    K.dt = dt; K.h = h; K.inv_h = inv_h;
    K.bias_rate = omega / a1; K.mass_scale = a2 * a3; K.imp_scale = a3 * P[SF_P_LEAK];
    K.max_push = P[SF_P_MAXPUSH]; K.anchor_push = P[SF_P_ANCHOR_PUSH]; K.anchor_slop = P[SF_P_ANCHOR_SLOP];
    K.spec = P[SF_P_SPEC]; K.inv_cell = 1.0f / P[SF_P_CELL]; K.friction = P[SF_P_FRICTION];
    K.sleep_lin2 = P[SF_P_SLEEP_LIN] * P[SF_P_SLEEP_LIN]; K.sleep_ang2 = P[SF_P_SLEEP_ANG] * P[SF_P_SLEEP_ANG];
    K.sleep_time = P[SF_P_SLEEP_TIME]; K.sleep_dist2 = P[SF_P_SLEEP_DIST] * P[SF_P_SLEEP_DIST]; K.sleep_rot = P[SF_P_SLEEP_ROT];
    K.wmax = 0.25f * 3.14159265358979323846f * inv_h;
    K.gx = P[SF_P_GX]; K.gy = P[SF_P_GY]; K.gz = P[SF_P_GZ];
    K.anchor = P[SF_P_ANCHOR] > 0.5f; K.allow_sleep = P[SF_P_ALLOW_SLEEP] > 0.5f;
    K.K = (int)P[SF_P_COL_LIMIT]; if (K.K < 1) K.K = 1; if (K.K > 64) K.K = 64;
    K.spread = (int)P[SF_P_COL_SPREAD]; if (K.spread < 1) K.spread = 1;
    K.use_bias = 1;
    K.ratio = P[SF_P_RATIO] > 0.0f ? P[SF_P_RATIO] : 0.0f;
    K.shock = std::min(std::max((int)P[SF_P_SHOCK], 0), 64);
    K.shock = (K.shock + 1) & ~1;  // an even count: the heights end in lvl
    K.rest_thr = P[SF_P_REST_THRESHOLD];
    K.rest_iters = std::min(std::max((int)P[SF_P_REST_ITERS], 0), 64);
    K.col_support = std::max((int)P[SF_P_COL_SUPPORT], 0);
    K.red_k = P[SF_P_RED_K] >= 1.0f ? std::max((int)P[SF_P_RED_K], RED_KMIN) : 0;
    K.any_rest = 0;  // (set from the loaded shapes)
    return K;
}


// Everything up to the colour counts; the sub-steps are planned from the counts read back by the caller.
int step_front(HostWorld* w, const StepK& K) {
    Dev& D = w->D;
    hipStream_t s = w->st;
    const unsigned gb = sf_grid(D.nb), gs = sf_grid(D.ns), gc = sf_grid(D.ccap), gm = sf_grid(D.mcap);
    // old manifolds = last step's
    std::swap(D.cur, D.old);
    // counters and per-step tables
    hipLaunchKernelGGL(k_step_clear, dim3(sf_grid(std::max<size_t>(D.tcap, ((size_t)D.nb + 1) * K.K))), dim3(NT), 0, s, D, K.K);
    if (w->timing) SF_HIP(hipEventRecord(w->ev[0], s));
    hipLaunchKernelGGL(k_prep_bodies, dim3(gb), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_prep_shapes, dim3(gs), dim3(NT), 0, s, D, K);
    sf::scan_u32(D.ecnt, D.eoff, D.ctr + C_NE, nullptr, (uint32_t)D.ns, (uint32_t)D.ns, w->scan_tmp, s);
    hipLaunchKernelGGL(k_emit_entries, dim3(gs), dim3(NT), 0, s, D, K);
    sf::scan_u32(D.tcnt, D.tstart, nullptr, nullptr, D.tcap, D.tcap, w->scan_tmp, s);
    hipLaunchKernelGGL(k_fill_table, dim3(sf_grid(D.ecap)), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_slot_flag, dim3(sf_grid(D.tcap)), dim3(NT), 0, s, D);
    sf::scan_u32(D.tfill, D.soff, D.ctr + C_NSLOT, nullptr, D.tcap, D.tcap, w->scan_tmp, s);
    hipLaunchKernelGGL(k_slot_list, dim3(sf_grid(D.tcap)), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_cand_slot, dim3(sf_grid((uint64_t)D.ecap * 64, NT, 65536)), dim3(NT), 0, s, D, K);
    hipLaunchKernelGGL(k_cand_shape, dim3(gs), dim3(NT), 0, s, D);
    sf::scan_u32(D.ccnt, D.coff, D.ctr + C_NCAND, nullptr, (uint32_t)D.ns, (uint32_t)D.ns, w->scan_tmp, s);
    hipLaunchKernelGGL(k_cand_write, dim3(sf_grid(D.ecap)), dim3(NT), 0, s, D, K);
    hipLaunchKernelGGL(k_cand_finish, dim3(gs), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_cand_finish_big, dim3(1024), dim3(NT), 0, s, D);
    if (w->timing) SF_HIP(hipEventRecord(w->ev[1], s));
    hipLaunchKernelGGL(k_narrow, dim3(gc), dim3(NT), 0, s, D, K, w->any_hull ? 1 : 0);
    if (w->any_hull) hipLaunchKernelGGL(k_narrow_hull, dim3(std::min(gc, 1024u)), dim3(NT), 0, s, D, K);
    sf::scan_u32(D.hit, D.hoff, D.ctr + C_NMAN, D.ctr + C_NCAND, 0, D.ccap, w->scan_tmp, s);
    hipLaunchKernelGGL(k_apply_wake, dim3(gb), dim3(NT), 0, s, D);
    if (K.red_k) {
        hipLaunchKernelGGL(k_red_count, dim3(gc), dim3(NT), 0, s, D, K);
        hipLaunchKernelGGL(k_red_score, dim3(gc), dim3(NT), 0, s, D);
        hipLaunchKernelGGL(k_red_pick, dim3(gc), dim3(NT), 0, s, D);
    }
    hipLaunchKernelGGL(k_man_hdr, dim3(gc), dim3(NT), 0, s, D, K);
    if (w->timing) SF_HIP(hipEventRecord(w->ev[2], s));
    // colouring rounds (checked for progress every few rounds and at the end of the support rounds: rounds in which
    // nothing claimed change nothing more in their phase, so a quiet group ends the support phase or the colouring)
    int R = (int)w->prm[SF_P_COL_ROUNDS];
    const int RS = K.col_support;
    hipLaunchKernelGGL(k_col_init, dim3(gm), dim3(NT), 0, s, D, K);
    int rounds = 0, flag = 0;
    for (int r = 0; r < R; r++) {
        hipLaunchKernelGGL(k_col_claim, dim3(gm), dim3(NT), 0, s, D, K.K, K.spread, r, RS, flag);
        hipLaunchKernelGGL(k_col_resolve, dim3(gm), dim3(NT), 0, s, D, K.K);
        rounds++;
        if (((r & 3) == 3 || r == RS - 1) && r + 1 < R) {
            SF_HIP(hipMemcpyAsync(w->h_ctr, D.ctr, sizeof(uint32_t) * C_NCTR, hipMemcpyDeviceToHost, s));
            SF_HIP(hipStreamSynchronize(s));
            const bool active = w->h_ctr[C_ACTIVE0 + flag] != 0;
            flag = std::min(flag + 1, C_NACTIVE - 1);
            if (flag == C_NACTIVE - 1) {  // (the last flag is reused: cleared before its next group)
                w->h_ctr[C_ACTIVE0 + flag] = 0;
                SF_HIP(hipMemcpyAsync(D.ctr + C_ACTIVE0 + flag, w->h_ctr + C_ACTIVE0 + flag, sizeof(uint32_t), hipMemcpyHostToDevice, s));
            }
            if (!active) {
                if (r < RS) { r = RS - 1; continue; }
                break;
            }
        }
    }
    w->stats[SF_S_ROUNDS] = rounds;
    hipLaunchKernelGGL(k_col_count, dim3(gm), dim3(NT), 0, s, D, K.K);
    hipLaunchKernelGGL(k_col_chunk_scan, dim3(K.K + OVS + 1), dim3(NT), 0, s, D, K.K);
    sf::scan_u32(D.subcnt, D.suboff, D.colcnt + K.K, nullptr, OVS, OVS, w->scan_tmp, s);
    sf::scan_u32(D.colcnt, D.coloff, nullptr, nullptr, (uint32_t)K.K + 1, 72, w->scan_tmp, s);
    hipLaunchKernelGGL(k_col_fill, dim3(gm), dim3(NT), 0, s, D, K.K);
    hipLaunchKernelGGL(k_ovf_count, dim3(gm), dim3(NT), 0, s, D, K.K);
    // per-slot contact lists in manifold order, then the manifolds built straight into colour order
    sf::scan_u32(D.blcnt, D.bloff, nullptr, nullptr, (uint32_t)D.nb + 1, (uint32_t)D.nb + 1, w->scan_tmp, s);
    hipLaunchKernelGGL(k_bl_fill, dim3(gm), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_bl_sort, dim3(sf_grid(((uint64_t)D.nb + 1) * 64, NT, 65536)), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_bl_big, dim3(sf_grid((uint64_t)D.nb + 1)), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_bl_sort_big, dim3(1024), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_bl_inv, dim3(sf_grid(2ull * D.mcap)), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_build, dim3(gc), dim3(NT), 0, s, D, K, w->nold);
    hipLaunchKernelGGL(k_sum_points, dim3(1), dim3(NT), 0, s, D, gc);
    // shock-propagation heights: an even number of passes alternating between the two copies, ending in lvl
    for (int k = 0; k < K.shock; k++) {
        const uint32_t* rd = (k & 1) ? D.lvl2 : D.lvl;
        uint32_t* wr = (k & 1) ? D.lvl : D.lvl2;
        hipLaunchKernelGGL(k_level_copy, dim3(sf_grid((uint64_t)D.nb + 1)), dim3(NT), 0, s, D, rd, wr);
        hipLaunchKernelGGL(k_level_lift, dim3(gm), dim3(NT), 0, s, D, rd, wr);
    }
    if (w->timing) SF_HIP(hipEventRecord(w->ev[3], s));
    SF_HIP(hipMemcpyAsync(w->h_col, D.colcnt, sizeof(uint32_t) * 72, hipMemcpyDeviceToHost, s));
    SF_HIP(hipMemcpyAsync(w->h_ctr, D.ctr, sizeof(uint32_t) * C_NCTR, hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    return 0;
}

// one colour bucket's launch; with kernel timing on, an event pair around it
int launch_bucket(HostWorld* w, const StepK& K, int c, unsigned grid) {
    hipStream_t s = w->st;
    if (w->ktime) {
        while (w->kev.size() < w->kev_used + 2) { hipEvent_t e; SF_HIP(hipEventCreate(&e)); w->kev.push_back(e); }
        SF_HIP(hipEventRecord(w->kev[w->kev_used], s));
    }
    launch_solve_bucket(solve_nt_of(w->prm), grid, s, w->D, K, c);
    if (w->ktime) {
        SF_HIP(hipEventRecord(w->kev[w->kev_used + 1], s));
        w->kev_used += 2;
        w->ksolve_launches++;
        w->ksolve_items += w->h_col[c];
    }
    return 0;
}

// The sub-steps' launches with the given grids (one per big colour, then the overflow pass); the kernels read the
// counts on the device, so grids larger than the counts need are correct.
int substep_chain(HostWorld* w, const StepK& K, int big, bool tail, bool ovf, int sub, const unsigned* grid) {
    Dev& D = w->D;
    hipStream_t s = w->st;
    const unsigned gb = sf_grid(D.nb), gw = sf_grid((uint64_t)D.nb * 64, NT, 65536);
    StepK Kb = K; Kb.use_bias = 1;
    StepK Kr = K; Kr.use_bias = 0;
    for (int it = 0; it < sub; it++) {
        // (with an overflow pass, the warm start of sub-steps after the first runs with the last one's average)
        if (it == 0 || !ovf) hipLaunchKernelGGL(k_warm, dim3(gw), dim3(NT), 0, s, D, K);
        for (int c = 0; c < big; c++) { int r = launch_bucket(w, Kb, c, grid[c]); if (r) return r; }
        if (tail) hipLaunchKernelGGL(k_solve_tail, dim3(1), dim3(NT), 0, s, D, Kb, big);
        if (ovf) {
            hipLaunchKernelGGL(k_ovf_solve, dim3(grid[big]), dim3(NT), 0, s, D, Kb);
            hipLaunchKernelGGL(k_ovf_apply_integrate, dim3(gw), dim3(NT), 0, s, D, K);
        } else {
            hipLaunchKernelGGL(k_integrate, dim3(gb), dim3(NT), 0, s, D, K);
        }
        for (int c = 0; c < big; c++) { int r = launch_bucket(w, Kr, c, grid[c]); if (r) return r; }
        if (tail) hipLaunchKernelGGL(k_solve_tail, dim3(1), dim3(NT), 0, s, D, Kr, big);
        if (ovf) {
            hipLaunchKernelGGL(k_ovf_solve, dim3(grid[big]), dim3(NT), 0, s, D, Kr);
            if (it + 1 < sub) hipLaunchKernelGGL(k_ovf_apply_warm, dim3(gw), dim3(NT), 0, s, D, K);
            else hipLaunchKernelGGL(k_ovf_apply, dim3(gw), dim3(NT), 0, s, D, Kr, 0);
        }
    }
    // restitution after the sub-steps: the colours in order, then the overflow (Jacobi with mass splitting)
    if (K.any_rest && K.rest_iters > 0) {
        for (int c = 0; c < big; c++) hipLaunchKernelGGL(k_rest_bucket, dim3(sf_grid(grid[c] * (unsigned)solve_nt_of(w->prm), NT, 16384)), dim3(NT), 0, s, D, Kr, c);
        if (tail) hipLaunchKernelGGL(k_rest_tail, dim3(1), dim3(NT), 0, s, D, Kr, big);
        if (ovf) {
            hipLaunchKernelGGL(k_rest_ovf, dim3(grid[big]), dim3(NT), 0, s, D, Kr);
            hipLaunchKernelGGL(k_ovf_apply, dim3(gw), dim3(NT), 0, s, D, Kr, 1);
        }
    }
    return 0;
}

// The sub-steps as one graph launch. A captured graph is replayed while the device tables, the step constants and
// the colour plan are those it was captured with and every launch's grid still covers its count without being more
// than 4x too large; otherwise the chain is captured again, with grids of 1.5x the counts, in place of the graph
// used least recently.
int solve_graph(HostWorld* w, const StepK& K, int big, bool tail, uint32_t novf, int sub) {
    hipStream_t s = w->st;
    const int nt = solve_nt_of(w->prm);
    std::vector<uint8_t> key(sizeof(Dev) + sizeof(StepK) + 5 * sizeof(int));
    int plan[5] = {big, tail ? 1 : 0, novf ? 1 : 0, sub, nt};
    memcpy(key.data(), &w->D, sizeof(Dev));
    memcpy(key.data() + sizeof(Dev), &K, sizeof(StepK));
    memcpy(key.data() + sizeof(Dev) + sizeof(StepK), plan, sizeof(plan));
    std::vector<unsigned> need(big + 1);
    for (int c = 0; c < big; c++) need[c] = sf_grid(w->h_col[c], nt, 16384);
    need[big] = sf_grid(novf ? novf : 1);
    auto fits = [&](const HostWorld::SubGraph& g) {
        if (!g.exec || g.key != key || g.grid.size() != need.size()) return false;
        for (size_t k = 0; k < need.size(); k++) if (!(need[k] <= g.grid[k] && 4ull * need[k] >= g.grid[k])) return false;
        return true;
    };
    HostWorld::SubGraph* use = fits(w->sg[0]) ? &w->sg[0] : (fits(w->sg[1]) ? &w->sg[1] : nullptr);
    const int64_t tick = w->graph_captures + w->graph_replays + 1;
    if (!use) {
        use = w->sg[0].used <= w->sg[1].used ? &w->sg[0] : &w->sg[1];  // the one used least recently
        use->clear();
        use->grid.resize(need.size());
        for (int c = 0; c < big; c++) use->grid[c] = std::min(16384u, need[c] + need[c] / 2);
        use->grid[big] = std::min(4096u, need[big] + need[big] / 2);
        SF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = substep_chain(w, K, big, tail, novf != 0, sub, use->grid.data());
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc || e != hipSuccess || !g) { if (g) (void)hipGraphDestroy(g); sf::set_error(SF_E_HIP, "graph capture failed"); return SF_E_HIP; }
        use->graph = g;
        SF_HIP(hipGraphInstantiate(&use->exec, use->graph, nullptr, nullptr, 0));
        use->key = key;
        w->graph_captures++;
    } else {
        w->graph_replays++;
    }
    use->used = tick;
    SF_HIP(hipGraphLaunch(use->exec, s));
    return 0;
}

int step_back(HostWorld* w, StepK K) {
    Dev& D = w->D;
    hipStream_t s = w->st;
    const unsigned gb = sf_grid(D.nb), gm = sf_grid(D.mcap), gw = sf_grid((uint64_t)D.nb * 64, NT, 65536);
    int ncol = (int)w->h_ctr[C_NCOL];
    int big = 0;
    const int tail_min = (int)w->prm[SF_P_TAIL_MIN];
    while (big < ncol && (int)w->h_col[big] >= tail_min) big++;
    bool tail = big < ncol;
    uint32_t novf = w->h_col[K.K];
    int sub = (int)w->prm[SF_P_SUBSTEPS]; if (sub < 1) sub = 1;
    if (w->use_graph && !w->ktime) {
        int r = solve_graph(w, K, big, tail, novf, sub);
        if (r) return r;
    } else {
        // launched one by one, each grid sized to its count
        const int nt = solve_nt_of(w->prm);
        std::vector<unsigned> grid(big + 1);
        for (int c = 0; c < big; c++) grid[c] = sf_grid(w->h_col[c], nt, 16384);
        grid[big] = sf_grid(novf ? novf : 1);
        int r = substep_chain(w, K, big, tail, novf != 0, sub, grid.data());
        if (r) return r;
    }
    if (w->timing) SF_HIP(hipEventRecord(w->ev[4], s));
    hipLaunchKernelGGL(k_writeback, dim3(gb), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_sleep_timer, dim3(gb), dim3(NT), 0, s, D, K);
    if (K.allow_sleep) {
        hipLaunchKernelGGL(k_uf_init, dim3(gw), dim3(NT), 0, s, D, 0);
        hipLaunchKernelGGL(k_uf_jump, dim3(gb), dim3(NT), 0, s, D, 0);
        hipLaunchKernelGGL(k_uf_hook, dim3(gm), dim3(NT), 0, s, D, 0);
        hipLaunchKernelGGL(k_uf_flatten, dim3(gb), dim3(NT), 0, s, D, 0);
        hipLaunchKernelGGL(k_sleep_apply, dim3(gb), dim3(NT), 0, s, D);
    }
    if (w->timing) SF_HIP(hipEventRecord(w->ev[5], s));
    SF_HIP(hipGetLastError());
    return 0;
}

int finish_stats(HostWorld* w) {
    uint32_t* c = w->h_ctr;
    w->nold = std::min(c[C_NMAN], w->D.mcap);
    w->stats[SF_S_AWAKE] = (int32_t)c[C_AWAKE];
    w->stats[SF_S_PAIRS] = (int32_t)c[C_NPAIR];
    w->stats[SF_S_MANIFOLDS] = (int32_t)std::min(c[C_NMAN], w->D.mcap);
    w->stats[SF_S_POINTS] = (int32_t)c[C_NPTS];
    w->stats[SF_S_COLOURS] = (int32_t)c[C_NCOL];
    w->stats[SF_S_OVERFLOW] = (int32_t)c[C_NOVF];
    w->stats[SF_S_GROUND] = (int32_t)c[C_NGROUND];
    w->stats[SF_S_WAKES] = (int32_t)c[C_NWAKE];
    w->stats[SF_S_ENTRIES] = (int32_t)c[C_NE];
    w->stats[SF_S_CANDIDATES] = (int32_t)c[C_NCAND];
    w->stats[SF_S_DROPPED] = (int32_t)c[C_DROP];
    w->stats[SF_S_REDUCED] = (int32_t)c[C_NRED];
    return 0;
}

// ---- particle debris (host side)
uint32_t p_grid_size(int np) { uint32_t g = 256; while (g < 2u * (uint32_t)np) g <<= 1; return g; }
int particles_alloc(HostWorld* w, int cap) {
    if (cap < w->P.np) { sf::set_error(SF_E_ARG, "capacity below the particles in use"); return SF_E_ARG; }
    SF_HIP(hipStreamSynchronize(w->st));
    PDev old = w->P;
    std::vector<void*> old_allocs;
    old_allocs.swap(w->pallocs);
    uint32_t* old_scan = w->pscan;
    PDev& P = w->P;
    const uint32_t G = p_grid_size(cap);
    auto get = [&](auto*& ptr, size_t n) -> int {
        void* q = nullptr;
        SF_HIP(hipMalloc(&q, sizeof(*ptr) * (n ? n : 1)));
        SF_HIP(hipMemset(q, 0, sizeof(*ptr) * (n ? n : 1)));
        w->pallocs.push_back(q);
        ptr = (std::remove_reference_t<decltype(ptr)>)q;
        return 0;
    };
    const size_t c = (size_t)cap;
    int r = 0;
    r |= get(P.x, c); r |= get(P.v, c); r |= get(P.q, c); r |= get(P.s, c); r |= get(P.w, c); r |= get(P.c, c); r |= get(P.a, c);
    r |= get(P.gl, c); r |= get(P.gn, c); r |= get(P.cnt, c); r |= get(P.nb, (size_t)P_NN * c); r |= get(P.sh, (size_t)P_NS * c);
    r |= get(P.cell, c); r |= get(P.gcnt, G); r |= get(P.gstart, G); r |= get(P.gfill, G); r |= get(P.list, c); r |= get(P.ctr, 4);
    r |= get(w->pscan, sf::scan_scratch_words(G));
    if (r) return r;
    if (!w->h_pctr) SF_HIP(hipHostMalloc((void**)&w->h_pctr, sizeof(uint32_t) * 4, hipHostMallocDefault));
    SF_HIP(hipDeviceSynchronize());  // (the zero fills ran on the null stream)
    P.cap = cap;
    w->Gmax = G;
    if (old.np) {  // keep the slots in use
        const size_t n = (size_t)old.np;
        float4* src[7] = {old.x, old.v, old.q, old.s, old.w, old.c, old.a};
        float4* dst[7] = {P.x, P.v, P.q, P.s, P.w, P.c, P.a};
        for (int k = 0; k < 7; k++) SF_HIP(hipMemcpy(dst[k], src[k], sizeof(float4) * n, hipMemcpyDeviceToDevice));
        SF_HIP(hipMemcpy(P.gl, old.gl, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice));
        SF_HIP(hipMemcpy(P.gn, old.gn, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice));
    }
    P.np = old.np;
    for (void* q : old_allocs) (void)hipFree(q);
    (void)old_scan;
    return 0;
}
// Records of 8 floats: p[3] v[3] radius friction (addParticles, gpu/gpu_world.js:2870-2883: moving, at rest
// rotation, the radius capped at half a grid cell less 5 cm). Slots of the new particles into `slots` (-1: full or
// not finite).
int particles_add(HostWorld* w, int n, const float* rec, int32_t* slots) {
    PDev& P = w->P;
    float pc = w->prm[SF_P_PCELL] > 0.1f ? w->prm[SF_P_PCELL] : 0.7f;
    const float rmax = pc * 0.5f - 0.05f;
    std::vector<int32_t> at;
    std::vector<float4> rx, rv, rq, rs, rw, rc, ra;
    int added = 0;
    for (int k = 0; k < n; k++) {
        const float* f = rec + (size_t)8 * k;
        bool fin = true;
        for (int a = 0; a < 8; a++) if (!std::isfinite(f[a])) fin = false;
        int i = -1;
        if (fin) {
            if (!w->pfree.empty()) { i = w->pfree.back(); w->pfree.pop_back(); }
            else if (P.np < P.cap) i = P.np++;
        }
        slots[k] = i;
        if (i < 0) continue;
        const float r = rmax < f[6] ? rmax : f[6];
        at.push_back(i);
        rx.push_back(make_float4(f[0], f[1], f[2], r));
        rv.push_back(make_float4(f[3], f[4], f[5], 1.0f));
        rq.push_back(make_float4(0.0f, 0.0f, 0.0f, 1.0f));
        rs.push_back(make_float4(f[0], f[1], f[2], 0.0f));
        rw.push_back(make_float4(0.0f, 0.0f, 0.0f, f[7]));
        rc.push_back(make_float4(f[0], f[1], f[2], 0.0f));
        ra.push_back(make_float4(f[0], f[1], f[2], 0.0f));
        added++;
    }
    SF_HIP(hipStreamSynchronize(w->st));
    const uint32_t zero = 0u;
    for (size_t t = 0; t < at.size(); t++) {
        const int i = at[t];
        SF_HIP(hipMemcpy(P.x + i, &rx[t], sizeof(float4), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P.v + i, &rv[t], sizeof(float4), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P.q + i, &rq[t], sizeof(float4), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P.s + i, &rs[t], sizeof(float4), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P.w + i, &rw[t], sizeof(float4), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P.c + i, &rc[t], sizeof(float4), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P.a + i, &ra[t], sizeof(float4), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P.gl + i, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(P.gn + i, &zero, sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    return added;
}
// the stage's launches, after the body step (on the world's stream)
int particle_stage(HostWorld* w, float dt) {
    PDev& P = w->P;
    const int nbl = (int)std::min<size_t>(w->pblast.size(), 3);
    PStepK K{};
    int psub = (int)w->prm[SF_P_PSUB]; if (psub < 1) psub = 1; if (psub > 16) psub = 16;
    K.dt = dt; K.h = dt / (float)psub; K.psub = psub;
    K.pc = w->prm[SF_P_PCELL] > 0.1f ? w->prm[SF_P_PCELL] : 0.7f;
    K.mug = w->prm[SF_P_FRICTION]; K.gx = w->prm[SF_P_GX]; K.gy = w->prm[SF_P_GY]; K.gz = w->prm[SF_P_GZ];
    K.cell = w->prm[SF_P_CELL];
    K.nblast = nbl; K.wake = w->pwake ? 1 : 0;
    for (int b = 0; b < nbl; b++) {
        const auto& e = w->pblast[b];
        K.bc[b] = make_float4(e[0], e[1], e[2], e[3]);
        K.bs[b] = e[4];
    }
    w->pblast.erase(w->pblast.begin(), w->pblast.begin() + nbl);
    w->pwake = false;
    w->pstats[0] = P.np; w->pstats[3] = (int32_t)w->pfree.size();
    if (!P.np) { w->pstats[1] = 0; w->pstats[2] = 0; return 0; }
    P.G = p_grid_size(P.np); P.gmask = P.G - 1u;
    hipStream_t s = w->st;
    const Dev& D = w->D;
    w->ptimed = false;
    if (w->timing) {
        for (auto& e : w->pev) if (!e) SF_HIP(hipEventCreate(&e));
        SF_HIP(hipEventRecord(w->pev[0], s));
    }
    const unsigned gp = sf_grid((uint64_t)P.np);
    hipLaunchKernelGGL(k_fill_u32, dim3(sf_grid(P.G)), dim3(NT), 0, s, P.gcnt, 0u, (size_t)P.G);
    hipLaunchKernelGGL(k_fill_u32, dim3(sf_grid(P.G)), dim3(NT), 0, s, P.gfill, 0u, (size_t)P.G);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, P.ctr, 0u, (size_t)4);
    hipLaunchKernelGGL(k_p_prep, dim3(gp), dim3(NT), 0, s, D, P, K);
    sf::scan_u32(P.gcnt, P.gstart, nullptr, nullptr, P.G, P.G, w->pscan, s);
    hipLaunchKernelGGL(k_p_scatter, dim3(gp), dim3(NT), 0, s, P);
    hipLaunchKernelGGL(k_p_sort, dim3(sf_grid(P.G)), dim3(NT), 0, s, P);
    hipLaunchKernelGGL(k_p_gather, dim3(gp), dim3(NT), 0, s, D, P, K);
    for (int k = 0; k < psub; k++) {
        hipLaunchKernelGGL(k_p_integrate, dim3(gp), dim3(NT), 0, s, P, K);
        hipLaunchKernelGGL(k_p_contact, dim3(gp), dim3(NT), 0, s, D, P, K);
        hipLaunchKernelGGL(k_p_apply, dim3(gp), dim3(NT), 0, s, P, K);
    }
    hipLaunchKernelGGL(k_p_finish, dim3(gp), dim3(NT), 0, s, D, P, K);
    if (w->timing) { SF_HIP(hipEventRecord(w->pev[1], s)); w->ptimed = true; }
    SF_HIP(hipMemcpyAsync(w->h_pctr, P.ctr, sizeof(uint32_t) * 4, hipMemcpyDeviceToHost, s));
    SF_HIP(hipGetLastError());
    return 1;  // (stats read after the step's synchronisation)
}

int world_step(HostWorld* w, float dt) {
    if (!(dt >= 0.0f) || dt > 1.0f) { sf::set_error(SF_E_ARG, "dt out of range"); return SF_E_ARG; }
    SF_HIP(hipSetDevice(w->dev));
    StepK K = make_step(w->prm, dt);
    K.any_rest = w->any_rest ? 1 : 0;
    int r = step_front(w, K);
    if (r) return r;
    r = step_back(w, K);
    if (r) return r;
    const int pst = particle_stage(w, dt);
    if (pst < 0) return pst;
    SF_HIP(hipStreamSynchronize(w->st));
    if (pst > 0) { w->pstats[1] = (int32_t)w->h_pctr[0]; w->pstats[2] = (int32_t)w->h_pctr[1]; }
    w->pstage_ms = 0.0f;
    if (w->ptimed) SF_HIP(hipEventElapsedTime(&w->pstage_ms, w->pev[0], w->pev[1]));
    w->steps++;
    finish_stats(w);
    if (w->timing) {
        for (int k = 0; k < 5; k++) { float ms = 0; (void)hipEventElapsedTime(&ms, w->ev[k], w->ev[k + 1]); w->stage_ms[k] = ms; }
    }
    if (w->ktime) {
        for (size_t k = 0; k < w->kev_used; k += 2) { float ms = 0; SF_HIP(hipEventElapsedTime(&ms, w->kev[k], w->kev[k + 1])); w->ksolve_ms += ms; }
        w->kev_used = 0;
    }
    if (w->h_ctr[C_DROP]) { sf::set_error(SF_E_CAP, "capacity exceeded: candidates or manifolds dropped"); }
    return w->stats[SF_S_AWAKE];
}

bool finite_all(const float* p, size_t n) { for (size_t i = 0; i < n; i++) if (!std::isfinite(p[i])) return false; return true; }

int world_load(HostWorld* w, int nb, const float* bodies, int ns, const int32_t* shb, const float* shapes) {
    if (nb < 0 || ns < 0 || nb > w->body_cap || ns > w->shape_cap) { sf::set_error(SF_E_CAP, "more bodies or shapes than the world's capacity"); return SF_E_CAP; }
    if ((nb && !bodies) || (ns && (!shb || !shapes))) { sf::set_error(SF_E_ARG, "null input"); return SF_E_ARG; }
    if (!finite_all(bodies, (size_t)nb * SF_BODY_F) || !finite_all(shapes, (size_t)ns * SF_SHAPE_F)) { sf::set_error(SF_E_ARG, "NaN or inf in the input"); return SF_E_ARG; }
    w->steps = 0;
    for (int s = 0; s < ns; s++) {
        if (shb[s] < 0 || shb[s] >= nb || (s > 0 && shb[s] < shb[s - 1])) { sf::set_error(SF_E_ARG, "shapes must name bodies in order"); return SF_E_ARG; }
    }
    SF_HIP(hipSetDevice(w->dev));
    Dev& D = w->D;
    D.nb = nb; D.ns = ns;
    std::vector<float4> bp(nb), bq(nb), bv(nb), bw(nb), bii(nb);
    std::vector<int32_t> fl(nb);
    for (int i = 0; i < nb; i++) {
        const float* f = bodies + (size_t)i * SF_BODY_F;
        bp[i] = make_float4(f[0], f[1], f[2], f[13]);
        bq[i] = make_float4(f[3], f[4], f[5], f[6]);
        bv[i] = make_float4(f[7], f[8], f[9], 0.0f);
        bw[i] = make_float4(f[10], f[11], f[12], 0.0f);
        bii[i] = make_float4(f[14], f[15], f[16], f[18]);
        fl[i] = f[17] > 0.5f ? 1 : 0;
    }
    std::vector<float4> slp(ns), slr(3 * (size_t)ns), shh(ns);
    std::vector<float> srest(ns);
    std::vector<int32_t> shull(ns);
    w->any_rest = false;
    w->any_hull = false;
    const int nhull = (int)sf::hull_lib().cells.size();
    for (int s = 0; s < ns; s++) {
        const float hf = shapes[(size_t)s * SF_SHAPE_F + 12];
        if (!(hf >= 0.0f && hf < (float)nhull && hf == (float)(int)hf)) { sf::set_error(SF_E_ARG, "a shape's hull is not in the chunk library"); return SF_E_ARG; }
        shull[s] = (int32_t)hf;
        if (shull[s]) w->any_hull = true;
    }
    w->h_h.assign(3 * (size_t)nb, 0.0f);
    std::vector<int> seen(nb, 0);
    for (int s = 0; s < ns; s++) {
        const float* f = shapes + (size_t)s * SF_SHAPE_F;
        // reach = |lp| + |h| (the shape's farthest extent from the body centre, for speculative margins)
        float lpx = f[0], lpy = f[1], lpz = f[2], hx = f[7], hy = f[8], hz = f[9];
        float l1 = lpx * lpx + lpy * lpy + lpz * lpz, l2 = hx * hx + hy * hy + hz * hz;
        float reach = (l1 <= 0.0f ? 0.0f : sqrtf(l1)) + (l2 <= 0.0f ? 0.0f : sqrtf(l2));
        slp[s] = make_float4(lpx, lpy, lpz, reach);
        float x = f[3], y = f[4], z = f[5], ww = f[6];
        slr[3 * s + 0] = make_float4(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + z * ww), 2.0f * (x * z - y * ww), 0.0f);
        slr[3 * s + 1] = make_float4(2.0f * (x * y - z * ww), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + x * ww), 0.0f);
        slr[3 * s + 2] = make_float4(2.0f * (x * z + y * ww), 2.0f * (y * z - x * ww), 1.0f - 2.0f * (x * x + y * y), 0.0f);
        shh[s] = make_float4(hx, hy, hz, f[10]);
        srest[s] = f[11];
        if (f[11] > 0.0f) w->any_rest = true;
        if (!seen[shb[s]]) { seen[shb[s]] = 1; w->h_h[3 * shb[s]] = hx; w->h_h[3 * shb[s] + 1] = hy; w->h_h[3 * shb[s] + 2] = hz; }
    }
    hipStream_t s = w->st;
    SF_HIP(hipMemcpyAsync(D.bp, bp.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.bq, bq.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.bv, bv.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.bw, bw.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.bii, bii.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.bsp, bp.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.bsq, bq.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.bflag, fl.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.shb, shb, sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.slp, slp.data(), sizeof(float4) * ns, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.slr, slr.data(), sizeof(float4) * 3 * ns, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.shh, shh.data(), sizeof(float4) * ns, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.srest, srest.data(), sizeof(float) * ns, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(D.shull, shull.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice, s));
    SF_HIP(hipStreamSynchronize(s));
    w->nold = 0;
    w->h_shb.assign(shb, shb + ns);
    return 0;
}

int world_read_bodies(HostWorld* w, float* out) {
    Dev& D = w->D;
    int nb = D.nb;
    std::vector<float4> bp(nb), bq(nb), bv(nb), bw(nb), bii(nb);
    std::vector<int32_t> fl(nb);
    SF_HIP(hipSetDevice(w->dev));
    SF_HIP(hipStreamSynchronize(w->st));
    SF_HIP(hipMemcpy(bp.data(), D.bp, sizeof(float4) * nb, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(bq.data(), D.bq, sizeof(float4) * nb, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(bv.data(), D.bv, sizeof(float4) * nb, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(bw.data(), D.bw, sizeof(float4) * nb, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(bii.data(), D.bii, sizeof(float4) * nb, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(fl.data(), D.bflag, sizeof(int32_t) * nb, hipMemcpyDeviceToHost));
    for (int i = 0; i < nb; i++) {
        float* f = out + (size_t)i * SF_BODY_F;
        f[0] = bp[i].x; f[1] = bp[i].y; f[2] = bp[i].z; f[3] = bq[i].x; f[4] = bq[i].y; f[5] = bq[i].z; f[6] = bq[i].w;
        f[7] = bv[i].x; f[8] = bv[i].y; f[9] = bv[i].z; f[10] = bw[i].x; f[11] = bw[i].y; f[12] = bw[i].z;
        f[13] = bp[i].w; f[14] = bii[i].x; f[15] = bii[i].y; f[16] = bii[i].z; f[17] = (fl[i] & 1) ? 1.0f : 0.0f;
        f[18] = bii[i].w; f[19] = 0.0f;
    }
    return 0;
}

}  // namespace

// ---------------------------------------------------------------- C ABI
struct sf_world { HostWorld h; };

extern "C" {

sf_world* sf_world_create(int body_cap, int shape_cap, const float* params) {
    if (body_cap <= 0 || shape_cap <= 0) { sf::set_error(SF_E_ARG, "capacities must be positive"); return nullptr; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { sf::set_error(SF_E_NODEV, "no GPU visible"); return nullptr; }
    sf_world* w = new sf_world();
    HostWorld* h = &w->h;
    SF_HIP_NULL(hipGetDevice(&h->dev));
    default_params(h->prm);
    if (params) for (int i = 0; i < SF_NPARAM; i++) h->prm[i] = params[i];
    h->body_cap = body_cap; h->shape_cap = shape_cap;
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) { delete w; sf::set_error(SF_E_HIP, "stream"); return nullptr; }
    if (world_alloc(h) != 0) { delete w; return nullptr; }
    return w;
}
void sf_world_destroy(sf_world* w) { delete w; }
int sf_world_set_params(sf_world* w, const float* params, int n) {
    if (!w || !params || n < 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    for (int i = 0; i < n && i < SF_NPARAM; i++) w->h.prm[i] = params[i];
    return 0;
}
int sf_world_load(sf_world* w, int nbody, const float* bodies, int nshape, const int32_t* shape_body, const float* shapes) {
    if (!w) { sf::set_error(SF_E_ARG, "null world"); return SF_E_ARG; }
    return world_load(&w->h, nbody, bodies, nshape, shape_body, shapes);
}
int sf_world_step(sf_world* w, float dt) {
    if (!w) { sf::set_error(SF_E_ARG, "null world"); return SF_E_ARG; }
    return world_step(&w->h, dt);
}
int sf_world_step_many(sf_world* w, float dt, int nsteps) {
    if (!w || nsteps < 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    for (int i = 0; i < nsteps; i++) { int r = world_step(&w->h, dt); if (r < 0) return r; }
    return 0;
}
int sf_world_sync(sf_world* w) { if (!w) return SF_E_ARG; SF_HIP(hipStreamSynchronize(w->h.st)); return 0; }
int sf_world_read_bodies(sf_world* w, float* out) {
    if (!w || !out) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    return world_read_bodies(&w->h, out);
}
int sf_world_stats(sf_world* w, int32_t* out) {
    if (!w || !out) return SF_E_ARG;
    for (int i = 0; i < SF_NSTAT; i++) out[i] = w->h.stats[i];
    return 0;
}
int sf_world_nbody(sf_world* w) { return w ? w->h.D.nb : SF_E_ARG; }
int sf_world_candidates(sf_world* w, int32_t* out, int cap) {
    if (!w) return SF_E_ARG;
    HostWorld* h = &w->h;
    int n = (int)std::min(h->h_ctr[C_NCAND], h->D.ccap);
    if (out && cap > 0) SF_HIP(hipMemcpy(out, h->D.cand, sizeof(int2) * std::min(n, cap), hipMemcpyDeviceToHost));
    return n;
}
int sf_world_manifolds(sf_world* w, int32_t* keys, float* data, int cap) {
    if (!w) return SF_E_ARG;
    HostWorld* h = &w->h;
    Dev& D = h->D;
    int n = (int)std::min(h->h_ctr[C_NMAN], D.mcap);
    if (!keys || !data || cap <= 0) return n;
    int k = std::min(n, cap);
    std::vector<int4> hdr(k);
    std::vector<uint32_t> pos(n);
    std::vector<float4> rows((size_t)MR_ROWS * n);
    SF_HIP(hipMemcpy(hdr.data(), D.cur.hdr, sizeof(int4) * k, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(pos.data(), D.cur.pos, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    for (int r = 0; r < MR_ROWS; r++) SF_HIP(hipMemcpy(rows.data() + (size_t)r * n, D.cur.row + (size_t)r * D.mcap, sizeof(float4) * n, hipMemcpyDeviceToHost));
    auto R = [&](int r, int m) { return rows[(size_t)r * n + pos[m]]; };
    for (int m = 0; m < k; m++) {
        int cnt = (int)R(MR_CB1, m).w;
        keys[3 * m] = hdr[m].x; keys[3 * m + 1] = hdr[m].y; keys[3 * m + 2] = cnt;
        float* f = data + (size_t)m * SF_MAN_F;
        std::memset(f, 0, sizeof(float) * SF_MAN_F);
        float4 n4 = R(MR_N, m);
        f[0] = n4.x; f[1] = n4.y; f[2] = n4.z; f[3] = (float)cnt;
        for (int q = 0; q < cnt; q++) {
            float4 a = R(MR_P0 + MR_PER_PT * q, m), b = R(MR_P0 + MR_PER_PT * q + 1, m), jn = R(MR_JN, m);
            float4 c = make_float4(0.0f, 0.0f, 0.0f, q == 0 ? jn.x : (q == 1 ? jn.y : (q == 2 ? jn.z : jn.w)));
            float* o = f + 4 + 8 * q;
            o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = b.x; o[4] = b.y; o[5] = b.z; o[6] = a.w; o[7] = c.w;
        }
        float4 j = R(MR_J, m);
        f[36] = j.x; f[37] = j.y; f[38] = j.z; f[39] = R(MR_CB2, m).w;
    }
    return n;
}
int sf_world_colours(sf_world* w, int32_t* out, int cap) {
    if (!w) return SF_E_ARG;
    HostWorld* h = &w->h;
    int n = (int)std::min(h->h_ctr[C_NMAN], h->D.mcap);
    if (out && cap > 0) SF_HIP(hipMemcpy(out, h->D.mcol, sizeof(int32_t) * std::min(n, cap), hipMemcpyDeviceToHost));
    return n;
}
// Island labels of the last step (smallest body index of each island of moving bodies in contact, a body's own
// index otherwise). The step itself unites only bodies that could fall asleep; the full labels are computed here,
// from the last step's contacts, on demand.
int sf_world_islands(sf_world* w, int32_t* out) {
    if (!w || !out) return SF_E_ARG;
    HostWorld* h = &w->h;
    Dev& D = h->D;
    if (h->steps == 0) {
        for (int i = 0; i < D.nb; i++) out[i] = i;
        return 0;
    }
    SF_HIP(hipSetDevice(h->dev));
    hipStream_t s = h->st;
    const unsigned gb = sf_grid(D.nb), gm = sf_grid(D.mcap), gw = sf_grid((uint64_t)D.nb * 64, NT, 65536);
    hipLaunchKernelGGL(k_uf_reset, dim3(gb), dim3(NT), 0, s, D);
    hipLaunchKernelGGL(k_uf_init, dim3(gw), dim3(NT), 0, s, D, 1);
    hipLaunchKernelGGL(k_uf_jump, dim3(gb), dim3(NT), 0, s, D, 1);
    hipLaunchKernelGGL(k_uf_hook, dim3(gm), dim3(NT), 0, s, D, 1);
    hipLaunchKernelGGL(k_uf_flatten, dim3(gb), dim3(NT), 0, s, D, 1);
    SF_HIP(hipGetLastError());
    SF_HIP(hipMemcpyAsync(out, D.parent, sizeof(int32_t) * D.nb, hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    return 0;
}
const float* sf_world_pose(sf_world* w) {
    if (!w) return nullptr;
    HostWorld* h = &w->h;
    int nb = h->D.nb;
    std::vector<float> b((size_t)nb * SF_BODY_F);
    if (world_read_bodies(h, b.data()) != 0) return nullptr;
    h->pose.resize((size_t)nb * 10);
    for (int i = 0; i < nb; i++) {
        const float* f = b.data() + (size_t)i * SF_BODY_F;
        float* o = h->pose.data() + (size_t)i * 10;
        for (int k = 0; k < 7; k++) o[k] = f[k];
        o[7] = h->h_h[3 * i]; o[8] = h->h_h[3 * i + 1]; o[9] = h->h_h[3 * i + 2];
    }
    return h->pose.data();
}
// ---- particle debris (gpu/gpu_world.js addParticles/killParticle/blastParticles/readParticles :2870-2896)
int sf_world_particles_capacity(sf_world* w, int cap) {
    if (!w || cap < 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    SF_HIP(hipSetDevice(w->h.dev));
    return particles_alloc(&w->h, cap);
}
int sf_world_add_particles(sf_world* w, int n, const float* rec8, int32_t* slots) {
    if (!w || n < 0 || (n && (!rec8 || !slots))) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    SF_HIP(hipSetDevice(w->h.dev));
    return particles_add(&w->h, n, rec8, slots);
}
int sf_world_kill_particle(sf_world* w, int i) {
    if (!w || i < 0 || i >= w->h.P.np) { sf::set_error(SF_E_ARG, "no such particle"); return SF_E_ARG; }
    HostWorld* h = &w->h;
    SF_HIP(hipSetDevice(h->dev));
    SF_HIP(hipStreamSynchronize(h->st));
    float4 v4;
    SF_HIP(hipMemcpy(&v4, h->P.v + i, sizeof(float4), hipMemcpyDeviceToHost));
    if (v4.w == 0.0f) return 0;
    v4.w = 0.0f;
    SF_HIP(hipMemcpy(h->P.v + i, &v4, sizeof(float4), hipMemcpyHostToDevice));
    h->pfree.push_back(i);
    return 0;
}
int sf_world_blast_particles(sf_world* w, float cx, float cy, float cz, float reach, float speed) {
    if (!w) return SF_E_ARG;
    if (w->h.pblast.size() < 16) w->h.pblast.push_back({cx, cy, cz, reach, speed});
    return 0;
}
int sf_world_wake_particles(sf_world* w) { if (!w) return SF_E_ARG; w->h.pwake = true; return 0; }
int sf_world_read_particles(sf_world* w, float* out12) {
    if (!w) return SF_E_ARG;
    HostWorld* h = &w->h;
    const int n = h->P.np;
    if (!out12 || !n) return n;
    SF_HIP(hipSetDevice(h->dev));
    SF_HIP(hipStreamSynchronize(h->st));
    std::vector<float4> x(n), v(n), q(n);
    SF_HIP(hipMemcpy(x.data(), h->P.x, sizeof(float4) * n, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(v.data(), h->P.v, sizeof(float4) * n, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(q.data(), h->P.q, sizeof(float4) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        float* o = out12 + (size_t)12 * i;
        o[0] = x[i].x; o[1] = x[i].y; o[2] = x[i].z; o[3] = v[i].x; o[4] = v[i].y; o[5] = v[i].z;
        o[6] = q[i].x; o[7] = q[i].y; o[8] = q[i].z; o[9] = q[i].w; o[10] = v[i].w; o[11] = x[i].w;
    }
    return n;
}
int sf_world_particle_ms(sf_world* w, float* ms) {
    if (!w || !ms) return SF_E_ARG;
    *ms = w->h.pstage_ms;
    return 0;
}
int sf_world_particle_stats(sf_world* w, int32_t* out4) {
    if (!w || !out4) return SF_E_ARG;
    for (int k = 0; k < 4; k++) out4[k] = w->h.pstats[k];
    return 0;
}
int sf_world_set_timing(sf_world* w, int on) { if (!w) return SF_E_ARG; w->h.timing = on != 0; return 0; }
int sf_world_set_graph(sf_world* w, int on) { if (!w) return SF_E_ARG; w->h.use_graph = on != 0; return 0; }
int sf_world_graph_stats(sf_world* w, int64_t* captures, int64_t* replays) {
    if (!w) return SF_E_ARG;
    if (captures) *captures = w->h.graph_captures;
    if (replays) *replays = w->h.graph_replays;
    return 0;
}
int sf_world_set_kernel_timing(sf_world* w, int on) {
    if (!w) return SF_E_ARG;
    w->h.ktime = on != 0;
    w->h.ksolve_ms = 0.0; w->h.ksolve_launches = 0; w->h.ksolve_items = 0;
    return 0;
}
int sf_world_kernel_stats(sf_world* w, double* ms, int64_t* launches, int64_t* items) {
    if (!w) return SF_E_ARG;
    if (ms) *ms = w->h.ksolve_ms;
    if (launches) *launches = w->h.ksolve_launches;
    if (items) *items = w->h.ksolve_items;
    return 0;
}
int sf_world_stage_ms(sf_world* w, float* out5) { if (!w || !out5) return SF_E_ARG; for (int k = 0; k < 5; k++) out5[k] = w->h.stage_ms[k]; return 0; }

// ---- the reference's singleton box-core exports (engine/src/wasm.rs:27-53)
static sf_world* g_world = nullptr;
static uint32_t g_substeps = 12;
static void replace_default(const std::vector<float>& bodies, const std::vector<int32_t>& shb, const std::vector<float>& shapes, const float* prm) {
    if (g_world) sf_world_destroy(g_world);
    int nb = (int)(bodies.size() / SF_BODY_F), ns = (int)shb.size();
    g_world = sf_world_create(nb > 0 ? nb : 1, ns > 0 ? ns : 1, prm);
    if (g_world) sf_world_load(g_world, nb, bodies.data(), ns, shb.data(), shapes.data());
}
void sf_pile_setup(uint32_t n, uint32_t substeps) {
    int ns = 0;
    int nb = sf_scene_pile((int)n, nullptr, nullptr, nullptr, &ns);
    std::vector<float> b((size_t)nb * SF_BODY_F), s((size_t)ns * SF_SHAPE_F);
    std::vector<int32_t> sb(ns);
    sf_scene_pile((int)n, b.data(), sb.data(), s.data(), &ns);
    float prm[SF_NPARAM]; default_params(prm);
    prm[SF_P_ALLOW_SLEEP] = 0.0f; prm[SF_P_SUBSTEPS] = (float)substeps;
    g_substeps = substeps;
    replace_default(b, sb, s, prm);
}
void sf_stack_setup(uint32_t n, uint32_t substeps, uint32_t sleep) {
    int ns = 0;
    int nb = sf_scene_stack((int)n, nullptr, nullptr, nullptr, &ns);
    std::vector<float> b((size_t)nb * SF_BODY_F), s((size_t)ns * SF_SHAPE_F);
    std::vector<int32_t> sb(ns);
    sf_scene_stack((int)n, b.data(), sb.data(), s.data(), &ns);
    float prm[SF_NPARAM]; default_params(prm);
    prm[SF_P_ALLOW_SLEEP] = sleep ? 1.0f : 0.0f; prm[SF_P_SUBSTEPS] = (float)substeps;
    replace_default(b, sb, s, prm);
}
uint32_t sf_step(float dt) { if (!g_world) return 0; int r = sf_world_step(g_world, dt); return r < 0 ? 0u : (uint32_t)r; }
uint32_t sf_body_count(void) { return g_world ? (uint32_t)g_world->h.D.nb : 0u; }
uint32_t sf_manifolds(void) { return g_world ? (uint32_t)g_world->h.stats[SF_S_MANIFOLDS] : 0u; }
uint32_t sf_points(void) { return g_world ? (uint32_t)g_world->h.stats[SF_S_POINTS] : 0u; }
float sf_max_speed(void) {
    if (!g_world) return 0.0f;
    int nb = g_world->h.D.nb;
    std::vector<float> b((size_t)nb * SF_BODY_F);
    if (world_read_bodies(&g_world->h, b.data()) != 0) return 0.0f;
    float m = 0.0f;
    for (int i = 0; i < nb; i++) { const float* f = b.data() + (size_t)i * SF_BODY_F; float v = f[7] * f[7] + f[8] * f[8] + f[9] * f[9]; v = v <= 0.0f ? 0.0f : sqrtf(v); if (v > m) m = v; }
    return m;
}
float sf_body_pos(uint32_t i, uint32_t axis) {
    const float* p = g_world ? sf_world_pose(g_world) : nullptr;
    if (!p || i >= (uint32_t)g_world->h.D.nb || axis > 2) return 0.0f;
    return p[(size_t)i * 10 + axis];
}
const float* sf_pose(void) { return g_world ? sf_world_pose(g_world) : nullptr; }

}  // extern "C"
