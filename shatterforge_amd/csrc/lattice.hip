// Lattice world on the GPU: the legacy brick engine's physics frame (src/lib.rs step :3639-3721) for one or many
// independent scenes at once, as hand-written gfx950 kernels plus host bookkeeping for the bursty topology changes.
//
// A frame (all scenes of a world together):
//   2 substeps of h = dt/2 (substep, src/lib.rs:1769-1783):
//     cache           rotation, world inertia, gravity per awake body; world brick positions (update_cache :1368)
//     hash            every live body brick into a counting-sort table keyed by (scene, cell); a cell's bricks in
//                     body order (build_hash :1383)
//     contacts        one lane per awake brick, counted then written: ground corners, spheres against the 27 lattice
//                     cells of the standing structure, spheres against bricks of other bodies in the 27 cells
//                     (gen_contacts :1396-1471); wake requests, support stamps, contact counts
//     plough          candidates flagged on the device, applied by the host in contact order (plough :1493-1550)
//     prep            effective masses, tangents, bias (prep_contacts :1551-1577)
//     colour          claim rounds: each contact proposes the lowest colour free at its moving bodies; the highest
//                     priority claim wins; the rest go to a Jacobi pass with mass splitting
//     solve x 10      colour after colour, Gauss-Seidel within a colour is race free (solve :1578-1601); then the
//                     overflow pass: private body copies with split masses, per-body sums as a 64-lane tree
//     pinch           velocity-change limits, pulverised small pieces (pinch_check :1746-1768)
//     impact          candidates flagged on the device, crushed by the host in contact order (impact_damage :1610)
//     integrate       drag, rolling resistance, per-body sleep, out-of-bounds (integrate :1703-1745), wakes
//   host: rebuild split bodies, detach unsupported parts of structures (rebuild_body :930, connectivity :1138)
//   structural (structural :1228-1365): resting loads onto structures, g - a loading of loaded pieces, red-black
//     SOR sweeps (one workgroup per structure or piece, adaptive count), joint judgement with fatigue; crushes and
//     cracks go back to the host as events
//   warm-up sweeps of newly solved structures (warm_step :1068), connectivity again
//
// Scenes never interact: every table is keyed by scene, every order and priority is scene local, so a scene's
// result does not depend on which other scenes share the world.
#include "lattice_dev.h"
#include "sf_host.h"
#include "sf_scan.h"
#include "../../include/sf.h"
#include <vector>
#include <string>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <chrono>
#include <atomic>
#include <thread>
#include <mutex>
#include <condition_variable>
#include <functional>
#include <memory>

using namespace lat;

namespace {

#define GS(i, n) for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (uint32_t)(n); i += gridDim.x * blockDim.x)
// launch on the world's stream (host code with a Lat* L in scope)
#define LAUNCH(kern, grid, block, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, L->st, __VA_ARGS__)

__device__ __forceinline__ LM load_m(const float4* p) {
    float4 a = p[0], b = p[1], c = p[2];
    LM r;
    r.m[0] = a.x; r.m[1] = a.y; r.m[2] = a.z; r.m[3] = b.x; r.m[4] = b.y; r.m[5] = b.z; r.m[6] = c.x; r.m[7] = c.y; r.m[8] = c.z;
    return r;
}
__device__ __forceinline__ void store_m(float4* p, const LM& m) {
    p[0] = make_float4(m.m[0], m.m[1], m.m[2], 0.0f);
    p[1] = make_float4(m.m[3], m.m[4], m.m[5], 0.0f);
    p[2] = make_float4(m.m[6], m.m[7], m.m[8], 0.0f);
}
__device__ __forceinline__ float fmax_bits(float a) { return a; }
__device__ __forceinline__ void atomic_max_pos(uint32_t* p, float v) { if (v > 0.0f) atomicMax(p, __float_as_uint(v)); }

// ---------------------------------------------------------------- cache
__global__ void __launch_bounds__(NT) k_cache(LD D, int gravity, float h) {
    GS(ub, D.nub) {
        const uint32_t b = D.ubody[ub];
        if (!D.alive[b] || D.asleep[b]) continue;
        LM r = q_to_m(lq4(D.q[b]));
        store_m(D.rot + 3 * (size_t)b, r);
        LM iil = load_m(D.iil + 3 * (size_t)b);
        LM t = mmul(r, iil), rt = mt(r);
        store_m(D.iiw + 3 * (size_t)b, mmul(t, rt));
        if (gravity) { float4 v = D.v[b]; v.y -= GR * h; D.v[b] = v; }
    }
}
__global__ void __launch_bounds__(NT) k_wpos(LD D) {
    GS(k, D.ndyn) {
        int b = D.dyn_body[k];
        if (!D.alive[b] || D.asleep[b]) continue;
        uint32_t i = D.dyn[k];
        LM r = load_m(D.rot + 3 * (size_t)b);
        D.wpos[i] = f4(add(lv4(D.pos[b]), mv(r, lv4(D.local[i]))));
    }
}

// ---------------------------------------------------------------- hash of body bricks by (scene, cell)
__device__ __forceinline__ unsigned long long cell_key(int scene, int x, int y, int z) {
    return ((unsigned long long)(uint32_t)scene << 33) | ((unsigned long long)((x + 1024) & 2047) << 22) |
           ((unsigned long long)((y + 1024) & 2047) << 11) | (unsigned long long)((z + 1024) & 2047);
}
__device__ __forceinline__ uint32_t key_slot(unsigned long long k, uint32_t mask) {
    return mix32((uint32_t)k ^ mix32((uint32_t)(k >> 32) * 0x9E3779B9u + 0x7F4A7C15u)) & mask;
}
__global__ void __launch_bounds__(NT) k_hash_count(LD D) {
    GS(k, D.ndyn) {
        uint32_t i = D.dyn[k];
        int b = D.dyn_body[k];
        if (!D.alive[b] || D.mat[i] == 0) { D.dslot[k] = NONE; continue; }
        LV p = lv4(D.wpos[i]);
        unsigned long long key = cell_key(D.bd_scene[b], fcell(p.x), fcell(p.y), fcell(p.z));
        uint32_t s = key_slot(key, D.tmask);
        D.dkey[k] = key;
        D.dslot[k] = s;
        atomicAdd(&D.hcnt[s], 1u);
    }
}
__global__ void __launch_bounds__(NT) k_hash_fill(LD D) {
    GS(k, D.ndyn) {
        uint32_t s = D.dslot[k];
        if (s == NONE) continue;
        D.hent[D.hoff[s] + atomicAdd(&D.hfill[s], 1u)] = k;
    }
}
// every segment of val[off[s] .. off[s] + cnt[s]) sorted ascending (in-cell and per-brick lists in a fixed order):
// each block takes NT consecutive segments; its waves sort the segments of 2-64 values by a bitonic network across
// the lanes, then the block sorts the larger ones one by one (bitonic in LDS up to 4096 values, insertion beyond)
__global__ void __launch_bounds__(NT) k_segsort_all(const uint32_t* off, const uint32_t* cnt, uint32_t* val, uint32_t nseg) {
    __shared__ uint32_t sh[4096];
    __shared__ uint32_t big[NT];
    __shared__ uint32_t nbig;
    const int lane = threadIdx.x & 63;
    for (uint32_t base = blockIdx.x * NT; base < nseg; base += gridDim.x * NT) {
        if (threadIdx.x == 0) nbig = 0u;
        __syncthreads();
        uint32_t s0 = base + threadIdx.x;
        uint32_t my_n = s0 < nseg ? cnt[s0] : 0u, my_o = s0 < nseg ? off[s0] : 0u;
        if (my_n > 64u) big[atomicAdd(&nbig, 1u)] = s0;
        unsigned long long todo = __ballot(my_n >= 2u && my_n <= 64u);
        while (todo) {
            int l = __ffsll((long long)todo) - 1;
            todo &= todo - 1;
            uint32_t o = __shfl(my_o, l, 64), n = __shfl(my_n, l, 64);
            uint32_t v = (uint32_t)lane < n ? val[o + lane] : 0xFFFFFFFFu;
            for (int k = 2; k <= 64; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1) {
                    uint32_t u = __shfl_xor(v, j, 64);
                    bool up = (lane & k) == 0, lower = (lane & j) == 0;
                    v = (lower == up) ? min(v, u) : max(v, u);
                }
            if ((uint32_t)lane < n) val[o + lane] = v;
        }
        __syncthreads();
        const uint32_t nb = nbig;
        for (uint32_t q = 0; q < nb; q++) {
            uint32_t s = big[q];
            uint32_t o = off[s], n = cnt[s];
            if (n > 4096) {
                if (threadIdx.x == 0)
                    for (uint32_t i = 1; i < n; i++) {
                        uint32_t v = val[o + i], j = i;
                        while (j > 0 && val[o + j - 1] > v) { val[o + j] = val[o + j - 1]; j--; }
                        val[o + j] = v;
                    }
                __syncthreads();
                continue;
            }
            uint32_t n2 = 128;
            while (n2 < n) n2 <<= 1;
            for (uint32_t i = threadIdx.x; i < n2; i += NT) sh[i] = i < n ? val[o + i] : 0xFFFFFFFFu;
            __syncthreads();
            for (uint32_t k = 2; k <= n2; k <<= 1)
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
            for (uint32_t i = threadIdx.x; i < n; i += NT) val[o + i] = sh[i];
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------- finalize (src/lib.rs:668-696)
// Each brick of a scene (one thread each): its six neighbours from the scene's lattice grid, its three positive-axis
// bond capacities (finalize_brick's rule: a neighbour of the same body and structure, not across an infill gap,
// min CAPV vertical / min CAPH lateral) and the footing bond of a standing brick on the ground (2 CAPV).
struct GridK { int32_t gx0, gy0, gz0, gnx, gny, gnz; uint32_t goff; };
__device__ __forceinline__ int32_t grid_at(const LD& D, const GridK& G, int x, int y, int z) {
    x -= G.gx0; y -= G.gy0; z -= G.gz0;
    if (x < 0 || y < 0 || z < 0 || x >= G.gnx || y >= G.gny || z >= G.gnz) return 0;
    return D.grid[G.goff + ((size_t)y * G.gnz + z) * G.gnx + x];
}
__global__ void __launch_bounds__(NT) k_finalize(LD D, GridK G, uint32_t b0, uint32_t nb) {
    const Tables& T = tables();
    GS(t, nb) {
        const uint32_t i = b0 + (uint32_t)t;
        const int mi = D.mat[i];
        if (mi == 0) continue;
        const int4 c = D.bxyz[i];
        uint32_t nbr[6];
        for (int d = 0; d < 6; d++) {
            const int32_t g = grid_at(D, G, c.x + DIRS[d][0], c.y + DIRS[d][1], c.z + DIRS[d][2]);
            nbr[d] = g ? (uint32_t)(g - 1) : NONE;
            D.nb6[6 * (size_t)i + d] = nbr[d];
        }
        float bd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        const int bi = D.body[i], si = D.sid[i], fi = D.infill[i];
        for (int a = 0; a < 3; a++) {
            const uint32_t j = nbr[2 * a];
            if (j == NONE) continue;
            const bool gap = fi != D.infill[j] && (a != 1 || fi == 1);
            if (!gap && D.mat[j] != 0 && D.body[j] == bi && D.sid[j] == si) bd[a] = bond_cap(mi, D.mat[j], a);
        }
        if (c.y == 0 && bi < 0) bd[3] = T.capv[mi] * 2.0f;
        const float4 b4 = make_float4(bd[0], bd[1], bd[2], bd[3]);
        D.bond[i] = b4;
        D.bmax[i] = b4;
    }
}

// ---------------------------------------------------------------- contacts
__device__ __forceinline__ int static_brick(const LD& D, const SceneK& S, int x, int y, int z) {
    x -= S.gx0; y -= S.gy0; z -= S.gz0;
    if (x < 0 || y < 0 || z < 0 || x >= S.gnx || y >= S.gny || z >= S.gnz) return -1;
    int32_t g = D.grid[S.goff + ((size_t)y * S.gnz + z) * S.gnx + x];
    if (!g) return -1;
    int j = g - 1;
    return (D.mat[j] != 0 && D.body[j] < 0) ? j : -1;
}

// Contacts of one awake dynamic brick (gen_contacts, src/lib.rs:1396-1471) by one wavefront: lanes 0-7 test the
// ground corners, lanes 8-34 the 27 lattice cells and lanes 35-61 the 27 hash cells, each in the sequential walk's
// order; an exclusive prefix sum over the lanes' counts puts every contact where the sequential walk would.
// WRITE = false counts (ccnt); WRITE = true counts again, then writes from coff with the side effects (support
// stamps, wake requests, per-body contact counts).
template <bool WRITE>
__global__ void __launch_bounds__(NT) k_contacts(LD D) {
    // the write pass counts contacts per body from zero and records each body's contact range
    if (!WRITE) GS(ub, D.nub) { const uint32_t b = D.ubody[ub]; D.nc[b] = 0u; D.cfirst[b] = 0xFFFFFFFFu; D.cend[b] = 0u; }
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t k = wave; k < D.ndyn; k += nwaves) {
        const int a = D.dyn_body[k];
        if (!D.alive[a] || D.asleep[a]) { if (!WRITE && lane == 0) D.ccnt[k] = 0; continue; }
        const uint32_t i = D.dyn[k];
        const int sc = D.bd_scene[a];
        const SceneK S = D.sc[sc];
        const LV p = lv4(D.wpos[i]);
        // one lane's probe: emit(b, bo, bb, point, normal, depth) for each contact in order; `side` also performs
        // the stamps and wake requests
        auto probe = [&](auto&& emit, bool side) {
            if (lane < 8) {
                if (p.y < 0.5f) {
                    LM rot = load_m(D.rot + 3 * (size_t)a);
                    int c = lane;
                    LV l = lv((c & 1) == 0 ? -H : H, (c & 2) == 0 ? -H : H, (c & 4) == 0 ? -H : H);
                    LV cp = add(p, mv(rot, l));
                    if (cp.y < 0.0f) { emit(-1, -1, NONE, cp, lv(0.0f, 1.0f, 0.0f), -cp.y); if (side) D.sup[i] = S.frame; }
                }
            } else if (lane < 35) {
                int c = lane - 8, dz = c / 9 - 1, dy = (c / 3) % 3 - 1, dx = c % 3 - 1;
                int x = fcell(p.x) + NX2 + dx, y = fcell(p.y) + dy, z = fcell(p.z) + NX2 + dz;
                int j = static_brick(D, S, x, y, z);
                if (j >= 0) {
                    LV cc = lat_pos(x, y, z);
                    LV q = lv(clampf(p.x, cc.x - H, cc.x + H), clampf(p.y, cc.y - H, cc.y + H), clampf(p.z, cc.z - H, cc.z + H));
                    LV d = sub(p, q);
                    float dist = len(d);
                    if (dist < RB) {
                        if (dist > 1e-4f) {
                            if (side && d.y > 0.3f * dist) D.sup[i] = S.frame;
                            emit(-1, -1, (uint32_t)j, q, mul(d, 1.0f / dist), RB - dist);
                        } else {
                            LV r = sub(p, cc);
                            float rv[3] = {r.x, r.y, r.z};
                            int ax = 0;
                            for (int t = 1; t < 3; t++) if (fabsf(rv[t]) > fabsf(rv[ax])) ax = t;
                            float nv[3] = {0.0f, 0.0f, 0.0f};
                            nv[ax] = rv[ax] >= 0.0f ? 1.0f : -1.0f;
                            emit(-1, -1, (uint32_t)j, p, lv(nv[0], nv[1], nv[2]), H + RB - fabsf(rv[ax]));
                        }
                    }
                }
            } else if (lane < 62) {
                int c = lane - 35, dz = c / 9 - 1, dy = (c / 3) % 3 - 1, dx = c % 3 - 1;
                unsigned long long key = cell_key(sc, fcell(p.x) + dx, fcell(p.y) + dy, fcell(p.z) + dz);
                uint32_t sl = key_slot(key, D.tmask);
                uint32_t e0 = D.hoff[sl], e1 = e0 + D.hcnt[sl];
                float speed = 0.0f;
                if (side) speed = len(lv4(D.v[a])) + len(lv4(D.w[a])) * D.rad[a] * 0.5f;
                for (uint32_t e = e0; e < e1; e++) {
                    uint32_t k2 = D.hent[e];
                    if (D.dkey[k2] != key) continue;
                    uint32_t j = D.dyn[k2];
                    int bj = D.body[j];
                    if (bj == a || D.mat[j] == 0) continue;
                    bool other_awake = !D.asleep[bj];
                    if (other_awake && j < i) continue;
                    LV d = sub(p, lv4(D.wpos[j]));
                    float dd = dot(d, d);
                    if (dd >= 4.0f * RB * RB || dd < 1e-10f) continue;
                    float dist = sqrtf(dd);
                    LV nn = mul(d, 1.0f / dist);
                    LV cp = madd(lv4(D.wpos[j]), nn, dist * 0.5f);
                    if (side) { if (nn.y > 0.3f) D.sup[i] = S.frame; else if (nn.y < -0.3f) D.sup[j] = S.frame; }
                    if (other_awake) emit(bj, bj, j, cp, nn, 2.0f * RB - dist);
                    else {
                        if (side && speed > 0.7f) D.wake[bj] = 1;
                        emit(-1, bj, j, cp, nn, 2.0f * RB - dist);
                    }
                }
            }
        };
        uint32_t n = 0;
        probe([&](int, int, uint32_t, LV, LV, float) { n++; }, false);
        // exclusive prefix over the lanes, and the wave's total
        uint32_t x = n;
        for (int d = 1; d < 64; d <<= 1) { uint32_t y = __shfl_up(x, d, 64); if (lane >= d) x += y; }
        if (!WRITE) {
            if (lane == 63) D.ccnt[k] = x;
            continue;
        }
        if (lane == 63 && x > 0) { atomicMin(&D.cfirst[a], D.coff[k]); atomicMax(&D.cend[a], D.coff[k] + x); }
        uint32_t out = D.coff[k] + x - n, m = 0;
        probe([&](int b, int bo, uint32_t bb, LV cp, LV nn, float depth) {
            uint32_t c = out + m;
            if (c < D.ccap) {
                D.ca[c] = a; D.cb[c] = b; D.cbo[c] = bo; D.cba[c] = i; D.cbb[c] = bb; D.cscene[c] = sc;
                D.cp[c] = f4(cp); D.cn[c] = f4(nn); D.cdepth[c] = depth;
                D.con[c] = 1; D.ccol[c] = b < 0 ? CHAIN : -1; D.cjn[c] = 0.0f; D.cjt1[c] = 0.0f; D.cjt2[c] = 0.0f; D.cvn0[c] = 0.0f;
            }
            atomicAdd(&D.nc[a], 1u);
            if (b >= 0) atomicAdd(&D.nc[b], 1u);
            m++;
        }, true);
    }
}

__device__ __forceinline__ LV vel_at(const LD& D, int b, LV r) { return b < 0 ? lv(0.0f, 0.0f, 0.0f) : add(lv4(D.v[b]), cross(lv4(D.w[b]), r)); }

__global__ void __launch_bounds__(NT) k_plough_flag(LD D) {
    // first contact of every scene (its first dynamic slot's offset)
    if (blockIdx.x == 0)
        for (uint32_t s = threadIdx.x; s < D.nscene; s += NT) {
            SceneK S = D.sc[s];
            D.sc_c0[s] = S.dyn0 < D.ndyn ? D.coff[S.dyn0] : D.ctr[C_NCS];
        }
    uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    const Tables& T = tables();
    GS(c, ncs) {
        if (!D.con[c] || D.cb[c] >= 0 || D.cbb[c] == NONE || D.cbo[c] >= 0) continue;
        uint32_t j = D.cbb[c];
        bool cand = D.mat[j] == 0 || D.body[j] >= 0;
        if (!cand) {
            int a = D.ca[c];
            LV ra = sub(lv4(D.cp[c]), lv4(D.pos[a]));
            float vn = -dot(vel_at(D, a, ra), lv4(D.cn[c]));
            if (vn < 2.0f) continue;
            float meff = D.m[a] * 3.0f / maxf((float)D.nc[a], 3.0f);
            float mom = minf(meff, D.m[a]) * vn;
            int mj = D.mat[j];
            if (mj >= SLUG || mom < T.capv[mj] * 0.35f) continue;
        }
        atomicAdd(&D.ctr[C_NPLOUGH], 1u);
        atomicAdd(&D.xpcnt[D.cscene[c]], 1u);
        D.pflag[c] = 1;
    }
}

__device__ __forceinline__ float eff_k(const LD& D, int a, LV r, LV n) {
    LM iiw = load_m(D.iiw + 3 * (size_t)a);
    return D.im[a] + dot(n, cross(mv(iiw, cross(r, n)), r));
}
__device__ __forceinline__ float eff_mass(const LD& D, int a, int b, LV ra, LV rb, LV n) {
    float k = eff_k(D, a, ra, n);
    if (b >= 0) k += eff_k(D, b, rb, n);
    return k > 1e-12f ? 1.0f / k : 0.0f;
}
__device__ __forceinline__ float eff_mass_split(const LD& D, int a, float na, int b, float nb, LV ra, LV rb, LV n) {
    float k = eff_k(D, a, ra, n) * na;
    if (b >= 0) k += eff_k(D, b, rb, n) * nb;
    return k > 1e-12f ? 1.0f / k : 0.0f;
}

__global__ void __launch_bounds__(NT) k_prep(LD D, float h) {
    // velocities before the solve (prep only reads them)
    GS(ub, D.nub) { const uint32_t b = D.ubody[ub]; if (D.alive[b]) { D.vpre[b] = D.v[b]; D.wpre[b] = D.w[b]; } }
    uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    GS(c, ncs) {
        if (!D.con[c]) continue;
        int a = D.ca[c], b = D.cb[c];
        LV p = lv4(D.cp[c]), n = lv4(D.cn[c]);
        LV ra = sub(p, lv4(D.pos[a]));
        LV rb = b >= 0 ? sub(p, lv4(D.pos[b])) : lv(0.0f, 0.0f, 0.0f);
        D.cra[c] = f4(ra); D.crb[c] = f4(rb);
        D.cmn[c] = eff_mass(D, a, b, ra, rb, n);
        LV dv = sub(vel_at(D, a, ra), vel_at(D, b, rb));
        float vn = dot(dv, n);
        D.cvn0[c] = maxf(-vn, 0.0f);
        LV vt = sub(dv, mul(n, vn));
        float lt = len(vt);
        LV t1;
        if (lt > 1e-4f) t1 = mul(vt, 1.0f / lt);
        else {
            LV t = fabsf(n.x) < 0.9f ? lv(1.0f, 0.0f, 0.0f) : lv(0.0f, 1.0f, 0.0f);
            LV x = cross(n, t);
            t1 = mul(x, 1.0f / len(x));
        }
        LV t2 = cross(n, t1);
        D.ct1[c] = f4(t1); D.ct2[c] = f4(t2);
        D.cmt1[c] = eff_mass(D, a, b, ra, rb, t1);
        D.cmt2[c] = eff_mass(D, a, b, ra, rb, t2);
        float bias = minf(0.2f / h * maxf(D.cdepth[c] - 0.01f, 0.0f), 1.0f);
        if (vn < -2.0f) bias = maxf(bias, -0.12f * vn);
        D.cbias[c] = bias;
    }
}

// ---------------------------------------------------------------- colouring
__global__ void __launch_bounds__(NT) k_col_clear(LD D) {
    GS(x, (size_t)D.nbd * D.K) {
        D.claim[x] = 0ull;
        if (x < D.nbd) D.mask[x] = 0u;
    }
}
__global__ void __launch_bounds__(NT) k_col_claim(LD D, int r) {
    uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    const uint32_t full = D.K >= 32 ? 0xFFFFFFFFu : ((1u << D.K) - 1u);
    GS(c, ncs) {
        D.cprop[c] = -1;
        if (!D.con[c] || D.ccol[c] != -1) continue;
        int a = D.ca[c], b = D.cb[c];
        uint32_t used = D.mask[a] | (b >= 0 ? D.mask[b] : 0u);
        uint32_t fr = ~used & full;
        if (!fr) continue;
        int col = __ffs(fr) - 1;
        D.cprop[c] = col;
        unsigned long long pr = claim_prio(c - D.sc_c0[D.cscene[c]], r);
        atomicMax(&D.claim[(size_t)a * D.K + col], pr);
        if (b >= 0) atomicMax(&D.claim[(size_t)b * D.K + col], pr);
    }
}
__global__ void __launch_bounds__(NT) k_col_resolve(LD D, int r) {
    uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    GS(c, ncs) {
        int col = D.cprop[c];
        if (col < 0) continue;
        int a = D.ca[c], b = D.cb[c];
        unsigned long long pr = claim_prio(c - D.sc_c0[D.cscene[c]], r);
        if (D.claim[(size_t)a * D.K + col] != pr) continue;
        if (b >= 0 && D.claim[(size_t)b * D.K + col] != pr) continue;
        D.ccol[c] = col;
        atomicOr(&D.mask[a], 1u << col);
        if (b >= 0) atomicOr(&D.mask[b], 1u << col);
    }
}
// contacts bucketed by colour (overflow last), one atomic per bucket per block
__global__ void __launch_bounds__(NT) k_col_count(LD D) {
    __shared__ uint32_t h[KMAX + 1];
    uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    for (uint32_t t = threadIdx.x; t <= D.K; t += NT) h[t] = 0;
    __syncthreads();
    GS(c, ncs) if (D.con[c] && D.ccol[c] != CHAIN) atomicAdd(&h[D.ccol[c] < 0 ? D.K : (uint32_t)D.ccol[c]], 1u);
    __syncthreads();
    for (uint32_t t = threadIdx.x; t <= D.K; t += NT) if (h[t]) atomicAdd(&D.ctr[C_COLCNT + t], h[t]);
}
__global__ void k_col_scan(LD D) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t o = 0;
        for (uint32_t t = 0; t <= D.K; t++) { D.coloff[t] = o; o += D.ctr[C_COLCNT + t]; D.colfill[t] = 0; }
    }
}
__global__ void __launch_bounds__(NT) k_col_fill(LD D) {
    __shared__ uint32_t cnt[KMAX + 1], base[KMAX + 1];
    uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    for (uint32_t c0 = blockIdx.x * NT; c0 < ncs; c0 += gridDim.x * NT) {
        for (uint32_t t = threadIdx.x; t <= D.K; t += NT) cnt[t] = 0;
        __syncthreads();
        uint32_t c = c0 + threadIdx.x;
        bool ok = c < ncs && D.con[c] && D.ccol[c] != CHAIN;
        uint32_t bin = 0, r = 0;
        if (ok) { bin = D.ccol[c] < 0 ? D.K : (uint32_t)D.ccol[c]; r = atomicAdd(&cnt[bin], 1u); }
        __syncthreads();
        for (uint32_t t = threadIdx.x; t <= D.K; t += NT) base[t] = cnt[t] ? atomicAdd(&D.colfill[t], cnt[t]) : 0u;
        __syncthreads();
        if (ok) D.collist[D.coloff[bin] + base[bin] + r] = c;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- overflow lists (per body, in contact order)
__global__ void __launch_bounds__(NT) k_ovf_count(LD D) {
    uint32_t n = D.ctr[C_COLCNT + D.K];
    GS(t, n) {
        uint32_t c = D.collist[D.coloff[D.K] + t];
        atomicAdd(&D.ovcnt[D.ca[c]], 1u);
        if (D.cb[c] >= 0) atomicAdd(&D.ovcnt[D.cb[c]], 1u);
    }
}
__global__ void __launch_bounds__(NT) k_ovf_fill(LD D) {
    uint32_t n = D.ctr[C_COLCNT + D.K];
    GS(t, n) {
        uint32_t c = D.collist[D.coloff[D.K] + t];
        int a = D.ca[c], b = D.cb[c];
        D.ovent[D.ovoff[a] + atomicAdd(&D.ovfill[a], 1u)] = 2u * c;
        if (b >= 0) D.ovent[D.ovoff[b] + atomicAdd(&D.ovfill[b], 1u)] = 2u * c + 1u;
    }
}
__global__ void __launch_bounds__(NT) k_ovf_prep(LD D) {
    uint32_t n = D.ctr[C_COLCNT + D.K];
    GS(t, n) {
        uint32_t c = D.collist[D.coloff[D.K] + t];
        int a = D.ca[c], b = D.cb[c];
        float na = (float)D.ovcnt[a], nb = b >= 0 ? (float)D.ovcnt[b] : 0.0f;
        LV ra = lv4(D.cra[c]), rb = lv4(D.crb[c]);
        D.csmn[c] = eff_mass_split(D, a, na, b, nb, ra, rb, lv4(D.cn[c]));
        D.csmt1[c] = eff_mass_split(D, a, na, b, nb, ra, rb, lv4(D.ct1[c]));
        D.csmt2[c] = eff_mass_split(D, a, na, b, nb, ra, rb, lv4(D.ct2[c]));
    }
}

// ---------------------------------------------------------------- solve
__device__ __forceinline__ void apply_imp(LD& D, int a, int b, LV ra, LV rb, LV p) {
    D.v[a] = f4(madd(lv4(D.v[a]), p, D.im[a]));
    D.w[a] = f4(add(lv4(D.w[a]), mv(load_m(D.iiw + 3 * (size_t)a), cross(ra, p))));
    if (b >= 0) {
        D.v[b] = f4(madd(lv4(D.v[b]), p, -D.im[b]));
        D.w[b] = f4(sub(lv4(D.w[b]), mv(load_m(D.iiw + 3 * (size_t)b), cross(rb, p))));
    }
}
// one contact of a colour bucket, Gauss-Seidel (solve, src/lib.rs:1579-1601)
__device__ __forceinline__ void solve_contact(LD& D, uint32_t c) {
    int a = D.ca[c], b = D.cb[c];
    LV ra = lv4(D.cra[c]), rb = lv4(D.crb[c]), nn = lv4(D.cn[c]), t1 = lv4(D.ct1[c]), t2 = lv4(D.ct2[c]);
    LV dv = sub(vel_at(D, a, ra), vel_at(D, b, rb));
    float vn = dot(dv, nn);
    float lam = D.cmn[c] * (-vn + D.cbias[c]);
    float old = D.cjn[c];
    float jn = maxf(old + lam, 0.0f);
    float d = jn - old;
    apply_imp(D, a, b, ra, rb, mul(nn, d));
    float lim = MU * jn;
    dv = sub(vel_at(D, a, ra), vel_at(D, b, rb));
    float l1 = -D.cmt1[c] * dot(dv, t1);
    float o1 = D.cjt1[c], j1 = clampf(o1 + l1, -lim, lim);
    float l2 = -D.cmt2[c] * dot(dv, t2);
    float o2 = D.cjt2[c], j2 = clampf(o2 + l2, -lim, lim);
    LV p = add(mul(t1, j1 - o1), mul(t2, j2 - o2));
    apply_imp(D, a, b, ra, rb, p);
    D.cjn[c] = jn; D.cjt1[c] = j1; D.cjt2[c] = j2;
}
__global__ void __launch_bounds__(NT) k_solve_colour(LD D, int col) {
    uint32_t n = D.ctr[C_COLCNT + col], o = D.coloff[col];
    GS(t, n) solve_contact(D, D.collist[o + t]);
}
// A body's static chain: its contacts against static bricks, the ground and sleeping bodies, in generation order
// (the reference's order for them, src/lib.rs:1580), solved one after another by one lane. They touch no other
// moving body, so every body's chain runs beside the others. A body's contacts (as side a) are the contiguous run
// its bricks' slots wrote, [cfirst, cend) (recorded by the write pass; a body made after it, by a plough, has none:
// the body check keeps a stale range from reaching another body's contacts).
__device__ __forceinline__ void chain_range(const LD& D, uint32_t b, uint32_t& c0, uint32_t& c1) {
    c0 = D.cfirst[b];
    c1 = D.cend[b];
}
// solve_contact with b < 0 on a body held in registers (the same operations in the same order)
__device__ __forceinline__ void chain_contact(LD& D, uint32_t c, LV& va, LV& wa, const LM& ia, float ima) {
    const LV ra = lv4(D.cra[c]), nn = lv4(D.cn[c]), t1 = lv4(D.ct1[c]), t2 = lv4(D.ct2[c]);
    const LV zero = lv(0.0f, 0.0f, 0.0f);
    LV dv = sub(add(va, cross(wa, ra)), zero);
    float vn = dot(dv, nn);
    float lam = D.cmn[c] * (-vn + D.cbias[c]);
    float old = D.cjn[c];
    float jn = maxf(old + lam, 0.0f);
    float d = jn - old;
    LV p = mul(nn, d);
    va = madd(va, p, ima); wa = add(wa, mv(ia, cross(ra, p)));
    float lim = MU * jn;
    dv = sub(add(va, cross(wa, ra)), zero);
    float l1 = -D.cmt1[c] * dot(dv, t1);
    float o1 = D.cjt1[c], j1 = clampf(o1 + l1, -lim, lim);
    float l2 = -D.cmt2[c] * dot(dv, t2);
    float o2 = D.cjt2[c], j2 = clampf(o2 + l2, -lim, lim);
    p = add(mul(t1, j1 - o1), mul(t2, j2 - o2));
    va = madd(va, p, ima); wa = add(wa, mv(ia, cross(ra, p)));
    D.cjn[c] = jn; D.cjt1[c] = j1; D.cjt2[c] = j2;
}
// one body's chain on its state in registers; false when the body has no chain contact (its state untouched)
__device__ __forceinline__ bool chain_body(LD& D, uint32_t b, LV& v, LV& w, const LM& ia, float ima) {
    uint32_t c0, c1;
    chain_range(D, b, c0, c1);
    bool any = false;
    for (uint32_t c = c0; c < c1; c++)
        if (D.con[c] && D.ccol[c] == CHAIN && D.ca[c] == (int)b) { chain_contact(D, c, v, w, ia, ima); any = true; }
    return any;
}
__global__ void __launch_bounds__(NT) k_chain(LD D) {
    GS(ub, D.nub) {
        const uint32_t b = D.ubody[ub];
        if (D.cfirst[b] >= D.cend[b]) continue;
        LV v = lv4(D.v[b]), w = lv4(D.w[b]);
        const LM ia = load_m(D.iiw + 3 * (size_t)b);
        if (chain_body(D, b, v, w, ia, D.im[b])) { D.v[b] = f4(v); D.w[b] = f4(w); }
    }
}
// one overflow contact, Jacobi: private copies of its bodies with inverse masses scaled by their overflow counts
// na, nb; the velocity changes go to odv/odw for the per-body sums
__device__ __forceinline__ void ovf_contact(LD& D, uint32_t c, float na, float nb) {
    int a = D.ca[c], b = D.cb[c];
    LV ra = lv4(D.cra[c]), rb = lv4(D.crb[c]), nn = lv4(D.cn[c]), t1 = lv4(D.ct1[c]), t2 = lv4(D.ct2[c]);
    LV va = lv4(D.v[a]), wa = lv4(D.w[a]), vb = lv(0.0f, 0.0f, 0.0f), wb = lv(0.0f, 0.0f, 0.0f);
    if (b >= 0) { vb = lv4(D.v[b]); wb = lv4(D.w[b]); }
    LM iia = load_m(D.iiw + 3 * (size_t)a);
    LM iib;
    if (b >= 0) iib = load_m(D.iiw + 3 * (size_t)b);
    float ima = D.im[a] * na, imb = b >= 0 ? D.im[b] * nb : 0.0f;
    LV dv = sub(add(va, cross(wa, ra)), b >= 0 ? add(vb, cross(wb, rb)) : lv(0.0f, 0.0f, 0.0f));
    float vn = dot(dv, nn);
    float lam = D.csmn[c] * (-vn + D.cbias[c]);
    float old = D.cjn[c];
    float jn = maxf(old + lam, 0.0f);
    LV pn = mul(nn, jn - old);
    va = madd(va, pn, ima);
    wa = add(wa, mul(mv(iia, cross(ra, pn)), na));
    if (b >= 0) { vb = madd(vb, pn, -imb); wb = sub(wb, mul(mv(iib, cross(rb, pn)), nb)); }
    float lim = MU * jn;
    dv = sub(add(va, cross(wa, ra)), b >= 0 ? add(vb, cross(wb, rb)) : lv(0.0f, 0.0f, 0.0f));
    float l1 = -D.csmt1[c] * dot(dv, t1);
    float o1 = D.cjt1[c], j1 = clampf(o1 + l1, -lim, lim);
    float l2 = -D.csmt2[c] * dot(dv, t2);
    float o2 = D.cjt2[c], j2 = clampf(o2 + l2, -lim, lim);
    LV pt = add(mul(t1, j1 - o1), mul(t2, j2 - o2));
    LV p = add(pn, pt);
    D.cjn[c] = jn; D.cjt1[c] = j1; D.cjt2[c] = j2;
    D.odv[2 * c] = f4(mul(p, D.im[a]));
    D.odw[2 * c] = f4(mv(iia, cross(ra, p)));
    if (b >= 0) {
        D.odv[2 * c + 1] = f4(mul(p, -D.im[b]));
        D.odw[2 * c + 1] = f4(neg(mv(iib, cross(rb, p))));
    }
}
__global__ void __launch_bounds__(NT) k_ovf_solve(LD D) {
    uint32_t n = D.ctr[C_COLCNT + D.K], o = D.coloff[D.K];
    GS(t, n) {
        uint32_t c = D.collist[o + t];
        int a = D.ca[c], b = D.cb[c];
        ovf_contact(D, c, (float)D.ovcnt[a], b >= 0 ? (float)D.ovcnt[b] : 0.0f);
    }
}
__device__ __forceinline__ LV wave_tree(LV x) {
    for (int h = 32; h >= 1; h >>= 1) {
        float ux = __shfl_down(x.x, h, 64), uy = __shfl_down(x.y, h, 64), uz = __shfl_down(x.z, h, 64);
        x = add(x, lv(ux, uy, uz));
    }
    return lv(__shfl(x.x, 0, 64), __shfl(x.y, 0, 64), __shfl(x.z, 0, 64));
}
// one body's overflow sum by one wave: its entries in contact order, 64 at a time as halving trees
__device__ __forceinline__ void ovf_body(LD& D, uint32_t b, uint32_t n, int lane) {
    uint32_t o = D.ovoff[b];
    LV av = lv(0.0f, 0.0f, 0.0f), aw = lv(0.0f, 0.0f, 0.0f);
    for (uint32_t e0 = 0; e0 < n; e0 += 64) {
        LV xv = lv(0.0f, 0.0f, 0.0f), xw = lv(0.0f, 0.0f, 0.0f);
        if (e0 + lane < n) { uint32_t ent = D.ovent[o + e0 + lane]; xv = lv4(D.odv[ent]); xw = lv4(D.odw[ent]); }
        av = add(av, wave_tree(xv));
        aw = add(aw, wave_tree(xw));
    }
    if (lane == 0) { D.v[b] = f4(add(lv4(D.v[b]), av)); D.w[b] = f4(add(lv4(D.w[b]), aw)); }
}
__global__ void __launch_bounds__(NT) k_ovf_apply(LD D) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t b = wave; b < D.nbd; b += nwaves) {
        uint32_t n = D.ovcnt[b];
        if (n) ovf_body(D, b, n, lane);
    }
}

// ---------------------------------------------------------------- fused per-scene solve
// One workgroup per scene runs the whole contact solve of a substep: claim-round colouring, colour buckets, the
// overflow lists and split masses, then the 10 iterations colour after colour with the overflow pass, separated by
// workgroup barriers instead of kernel boundaries. The arithmetic and every order are those of the kernels above
// (k_col_claim ... k_ovf_apply), so both routes give the same bits; this one needs one launch instead of about 150.
// Locations updated by atomics (claims, masks, overflow counts) are read with device-scope loads.
constexpr int FS_NT = 512;
__device__ __forceinline__ uint32_t ld_dev(const uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ unsigned long long ld_dev(const unsigned long long* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// exclusive scan of one value per thread over the workgroup (wsum: 17 words of LDS); every thread must call it
__device__ __forceinline__ uint32_t wg_excl_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) { uint32_t y = __shfl_up(x, d, 64); if (lane >= d) x += y; }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) { uint32_t o = 0; for (int k = 0; k < FS_NT / 64; k++) { uint32_t t = wsum[k]; wsum[k] = o; o += t; } wsum[FS_NT / 64] = o; }
    __syncthreads();
    *total = wsum[FS_NT / 64];
    return wsum[wv] + x - v;
}
// A contact's solve record, in colour-bucket order: its bodies as scene-local slots, arms, normal, tangents,
// effective masses (split ones for the overflow bucket), bias and accumulated impulses; one load per pass.
struct SRec {
    int a, b;               // scene-local body slots (b < 0: static)
    float4 r0, r1, r2, r3, r4, r5;  // ra|mn, rb|bias, n|mt1, t1|mt2, t2|na, jn jt1 jt2|nb
};
__device__ __forceinline__ SRec srec_load(const LD& D, uint32_t q) {
    SRec r;
    int2 ab = D.rq_ab[q];
    r.a = ab.x; r.b = ab.y;
    r.r0 = D.rq_0[q]; r.r1 = D.rq_1[q]; r.r2 = D.rq_2[q]; r.r3 = D.rq_3[q]; r.r4 = D.rq_4[q]; r.r5 = D.rq_5[q];
    return r;
}
// the scene's bodies in LDS: velocities, angular velocities, world inverse inertia rows, inverse mass
struct LBody {
    float4 *v, *w, *ii;
    float* im;
    __device__ __forceinline__ LM iiw(int b) const {
        float4 x = ii[3 * b], y = ii[3 * b + 1], z = ii[3 * b + 2];
        LM m;
        m.m[0] = x.x; m.m[1] = x.y; m.m[2] = x.z; m.m[3] = y.x; m.m[4] = y.y; m.m[5] = y.z; m.m[6] = z.x; m.m[7] = z.y; m.m[8] = z.z;
        return m;
    }
};
// solve_contact on LDS bodies (same operations in the same order)
__device__ __forceinline__ void srec_solve(const LD& D, const LBody& B, SRec& r, uint32_t q) {
    const int a = r.a, b = r.b;
    const LV ra = lv4(r.r0), rb = lv4(r.r1), nn = lv4(r.r2), t1 = lv4(r.r3), t2 = lv4(r.r4);
    LV va = lv4(B.v[a]), wa = lv4(B.w[a]), vb = lv(0.0f, 0.0f, 0.0f), wb = lv(0.0f, 0.0f, 0.0f);
    if (b >= 0) { vb = lv4(B.v[b]); wb = lv4(B.w[b]); }
    const float ima = B.im[a], imb = b >= 0 ? B.im[b] : 0.0f;
    const LM ia = B.iiw(a);
    LM ib;
    if (b >= 0) ib = B.iiw(b);
    LV dv = sub(add(va, cross(wa, ra)), b >= 0 ? add(vb, cross(wb, rb)) : lv(0.0f, 0.0f, 0.0f));
    float vn = dot(dv, nn);
    float lam = r.r0.w * (-vn + r.r1.w);
    float old = r.r5.x;
    float jn = maxf(old + lam, 0.0f);
    float d = jn - old;
    LV p = mul(nn, d);
    va = madd(va, p, ima); wa = add(wa, mv(ia, cross(ra, p)));
    if (b >= 0) { vb = madd(vb, p, -imb); wb = sub(wb, mv(ib, cross(rb, p))); }
    float lim = MU * jn;
    dv = sub(add(va, cross(wa, ra)), b >= 0 ? add(vb, cross(wb, rb)) : lv(0.0f, 0.0f, 0.0f));
    float l1 = -r.r2.w * dot(dv, t1);
    float o1 = r.r5.y, j1 = clampf(o1 + l1, -lim, lim);
    float l2 = -r.r3.w * dot(dv, t2);
    float o2 = r.r5.z, j2 = clampf(o2 + l2, -lim, lim);
    p = add(mul(t1, j1 - o1), mul(t2, j2 - o2));
    va = madd(va, p, ima); wa = add(wa, mv(ia, cross(ra, p)));
    if (b >= 0) { vb = madd(vb, p, -imb); wb = sub(wb, mv(ib, cross(rb, p))); }
    B.v[a] = f4(va); B.w[a] = f4(wa);
    if (b >= 0) { B.v[b] = f4(vb); B.w[b] = f4(wb); }
    r.r5.x = jn; r.r5.y = j1; r.r5.z = j2;
    D.rq_5[q] = r.r5;
}
// ovf_contact on LDS bodies; the velocity changes go to odv/odw under the contact's own index
__device__ __forceinline__ void srec_ovf(const LD& D, const LBody& B, SRec& r, uint32_t q) {
    const int a = r.a, b = r.b;
    const float na = r.r4.w, nb = r.r5.w;
    const LV ra = lv4(r.r0), rb = lv4(r.r1), nn = lv4(r.r2), t1 = lv4(r.r3), t2 = lv4(r.r4);
    LV va = lv4(B.v[a]), wa = lv4(B.w[a]), vb = lv(0.0f, 0.0f, 0.0f), wb = lv(0.0f, 0.0f, 0.0f);
    if (b >= 0) { vb = lv4(B.v[b]); wb = lv4(B.w[b]); }
    LM iia = B.iiw(a);
    LM iib;
    if (b >= 0) iib = B.iiw(b);
    float ima = B.im[a] * na, imb = b >= 0 ? B.im[b] * nb : 0.0f;
    LV dv = sub(add(va, cross(wa, ra)), b >= 0 ? add(vb, cross(wb, rb)) : lv(0.0f, 0.0f, 0.0f));
    float vn = dot(dv, nn);
    float lam = r.r0.w * (-vn + r.r1.w);
    float old = r.r5.x;
    float jn = maxf(old + lam, 0.0f);
    LV pn = mul(nn, jn - old);
    va = madd(va, pn, ima);
    wa = add(wa, mul(mv(iia, cross(ra, pn)), na));
    if (b >= 0) { vb = madd(vb, pn, -imb); wb = sub(wb, mul(mv(iib, cross(rb, pn)), nb)); }
    float lim = MU * jn;
    dv = sub(add(va, cross(wa, ra)), b >= 0 ? add(vb, cross(wb, rb)) : lv(0.0f, 0.0f, 0.0f));
    float l1 = -r.r2.w * dot(dv, t1);
    float o1 = r.r5.y, j1 = clampf(o1 + l1, -lim, lim);
    float l2 = -r.r3.w * dot(dv, t2);
    float o2 = r.r5.z, j2 = clampf(o2 + l2, -lim, lim);
    LV pt = add(mul(t1, j1 - o1), mul(t2, j2 - o2));
    LV p = add(pn, pt);
    r.r5.x = jn; r.r5.y = j1; r.r5.z = j2;
    D.rq_5[q] = r.r5;
    const uint32_t c = D.rq_c[q];
    D.odv[2 * c] = f4(mul(p, B.im[a]));
    D.odw[2 * c] = f4(mv(iia, cross(ra, p)));
    if (b >= 0) {
        D.odv[2 * c + 1] = f4(mul(p, -B.im[b]));
        D.odw[2 * c + 1] = f4(neg(mv(iib, cross(rb, p))));
    }
}
// ovf_body into LDS bodies
__device__ __forceinline__ void lbody_ovf(const LD& D, const LBody& B, uint32_t lb, uint32_t b, uint32_t n, int lane) {
    uint32_t o = D.ovoff[b];
    LV av = lv(0.0f, 0.0f, 0.0f), aw = lv(0.0f, 0.0f, 0.0f);
    for (uint32_t e0 = 0; e0 < n; e0 += 64) {
        LV xv = lv(0.0f, 0.0f, 0.0f), xw = lv(0.0f, 0.0f, 0.0f);
        if (e0 + lane < n) { uint32_t ent = D.ovent[o + e0 + lane]; xv = lv4(D.odv[ent]); xw = lv4(D.odw[ent]); }
        av = add(av, wave_tree(xv));
        aw = add(aw, wave_tree(xw));
    }
    if (lane == 0) { B.v[lb] = f4(add(lv4(B.v[lb]), av)); B.w[lb] = f4(add(lv4(B.w[lb]), aw)); }
}

__global__ void __launch_bounds__(FS_NT) k_scene_solve(LD D, const int32_t* list, uint32_t lds_cap) {
    __shared__ uint32_t cnt[KMAX + 1], off[KMAX + 2], fil[KMAX + 1], wsum[FS_NT / 64 + 1], nov;
    __shared__ uint32_t ps_off[KMAX + 3], ps_cnt[KMAX + 3], ps_type[KMAX + 3], npass, nchain;
    extern __shared__ float4 lbody[];
    const int s = list[blockIdx.x];
    const SceneK S = D.sc[s];
    const uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    const uint32_t c0 = min(D.sc_c0[s], ncs);
    const uint32_t c1 = s + 1 < (int)D.nscene ? min(D.sc_c0[s + 1], ncs) : ncs;
    if (c1 <= c0) return;
    const uint32_t tid = threadIdx.x, K = D.K;
    const uint32_t full = K >= 32 ? 0xFFFFFFFFu : ((1u << K) - 1u);
    const uint32_t bd0 = (uint32_t)S.bd0, hw = (uint32_t)S.hw;
    // clear the scene's bodies
    for (uint32_t lb = tid; lb < hw; lb += FS_NT) { uint32_t b = bd0 + lb; D.mask[b] = 0u; D.ovcnt[b] = 0u; D.ovfill[b] = 0u; }
    for (uint32_t x = tid; x < hw * K; x += FS_NT) D.claim[(size_t)bd0 * K + x] = 0ull;
    if (tid <= K) { cnt[tid] = 0u; fil[tid] = 0u; }
    if (tid == 0) { nov = 0u; nchain = 0u; }
    __syncthreads();
    // colouring: claim rounds (k_col_claim, k_col_resolve)
    for (int r = 0; r < (int)D.R; r++) {
        for (uint32_t c = c0 + tid; c < c1; c += FS_NT) {
            int prop = -1;
            if (D.con[c] && D.ccol[c] == -1) {
                int a = D.ca[c], b = D.cb[c];
                uint32_t used = ld_dev(&D.mask[a]) | (b >= 0 ? ld_dev(&D.mask[b]) : 0u);
                uint32_t fr = ~used & full;
                if (fr) {
                    int col = __ffs(fr) - 1;
                    prop = col;
                    unsigned long long pr = claim_prio(c - c0, r);
                    atomicMax(&D.claim[(size_t)a * K + col], pr);
                    if (b >= 0) atomicMax(&D.claim[(size_t)b * K + col], pr);
                }
            }
            D.cprop[c] = prop;
        }
        __syncthreads();
        for (uint32_t c = c0 + tid; c < c1; c += FS_NT) {
            int col = D.cprop[c];
            if (col < 0) continue;
            int a = D.ca[c], b = D.cb[c];
            unsigned long long pr = claim_prio(c - c0, r);
            if (ld_dev(&D.claim[(size_t)a * K + col]) != pr) continue;
            if (b >= 0 && ld_dev(&D.claim[(size_t)b * K + col]) != pr) continue;
            D.ccol[c] = col;
            atomicOr(&D.mask[a], 1u << col);
            if (b >= 0) atomicOr(&D.mask[b], 1u << col);
        }
        __syncthreads();
    }
    // colour buckets in the scene's own stretch of collist, overflow last
    for (uint32_t c = c0 + tid; c < c1; c += FS_NT) {
        if (!D.con[c]) continue;
        if (D.ccol[c] == CHAIN) nchain = 1u;
        else atomicAdd(&cnt[D.ccol[c] < 0 ? K : (uint32_t)D.ccol[c]], 1u);
    }
    __syncthreads();
    if (tid == 0) { uint32_t o = 0; for (uint32_t t = 0; t <= K; t++) { off[t] = o; o += cnt[t]; } off[K + 1] = o; }
    __syncthreads();
    for (uint32_t c = c0 + tid; c < c1; c += FS_NT) {
        if (!D.con[c] || D.ccol[c] == CHAIN) continue;
        uint32_t bin = D.ccol[c] < 0 ? K : (uint32_t)D.ccol[c];
        D.collist[c0 + off[bin] + atomicAdd(&fil[bin], 1u)] = c;
    }
    __syncthreads();
    // overflow lists: per body, its overflow contact sides in contact order (k_ovf_count ... k_ovf_prep)
    const uint32_t no = cnt[K], oo = c0 + off[K];
    if (no) {
        for (uint32_t t = tid; t < no; t += FS_NT) {
            uint32_t c = D.collist[oo + t];
            atomicAdd(&D.ovcnt[D.ca[c]], 1u);
            if (D.cb[c] >= 0) atomicAdd(&D.ovcnt[D.cb[c]], 1u);
        }
        __syncthreads();
        const uint32_t chunk = (hw + FS_NT - 1) / FS_NT, lo = min(tid * chunk, hw), hi = min(lo + chunk, hw);
        uint32_t sum = 0, nb = 0;
        for (uint32_t lb = lo; lb < hi; lb++) { uint32_t n = ld_dev(&D.ovcnt[bd0 + lb]); sum += n; nb += n ? 1u : 0u; }
        uint32_t tot;
        uint32_t base = wg_excl_scan(sum, wsum, &tot);
        uint32_t bslot = nb ? atomicAdd(&nov, nb) : 0u;
        for (uint32_t lb = lo, o = 2 * c0 + base; lb < hi; lb++) {
            uint32_t n = ld_dev(&D.ovcnt[bd0 + lb]);
            D.ovoff[bd0 + lb] = o;
            o += n;
            if (n) D.ovbody[bd0 + bslot++] = bd0 + lb;
        }
        __syncthreads();
        for (uint32_t t = tid; t < no; t += FS_NT) {
            uint32_t c = D.collist[oo + t];
            int a = D.ca[c], b = D.cb[c];
            D.ovent[D.ovoff[a] + atomicAdd(&D.ovfill[a], 1u)] = 2u * c;
            if (b >= 0) D.ovent[D.ovoff[b] + atomicAdd(&D.ovfill[b], 1u)] = 2u * c + 1u;
        }
        __syncthreads();
        for (uint32_t k = tid; k < nov; k += FS_NT) {
            uint32_t b = D.ovbody[bd0 + k], o = D.ovoff[b], n = ld_dev(&D.ovcnt[b]);
            for (uint32_t x = 1; x < n; x++) {
                uint32_t v = D.ovent[o + x], y = x;
                while (y > 0 && D.ovent[o + y - 1] > v) { D.ovent[o + y] = D.ovent[o + y - 1]; y--; }
                D.ovent[o + y] = v;
            }
        }
        for (uint32_t t = tid; t < no; t += FS_NT) {
            uint32_t c = D.collist[oo + t];
            int a = D.ca[c], b = D.cb[c];
            float na = (float)ld_dev(&D.ovcnt[a]), nb2 = b >= 0 ? (float)ld_dev(&D.ovcnt[b]) : 0.0f;
            LV ra = lv4(D.cra[c]), rb = lv4(D.crb[c]);
            D.csmn[c] = eff_mass_split(D, a, na, b, nb2, ra, rb, lv4(D.cn[c]));
            D.csmt1[c] = eff_mass_split(D, a, na, b, nb2, ra, rb, lv4(D.ct1[c]));
            D.csmt2[c] = eff_mass_split(D, a, na, b, nb2, ra, rb, lv4(D.ct2[c]));
        }
        __syncthreads();
    }
    const int lane = tid & 63;
    if (hw <= lds_cap) {
        // bodies into LDS, contacts into solve records in bucket order
        const LBody B{lbody, lbody + hw, lbody + 2 * hw, (float*)(lbody + 5 * hw)};
        for (uint32_t lb = tid; lb < hw; lb += FS_NT) {
            uint32_t b = bd0 + lb;
            B.v[lb] = D.v[b]; B.w[lb] = D.w[b];
            B.ii[3 * lb] = D.iiw[3 * (size_t)b]; B.ii[3 * lb + 1] = D.iiw[3 * (size_t)b + 1]; B.ii[3 * lb + 2] = D.iiw[3 * (size_t)b + 2];
            B.im[lb] = D.im[b];
        }
        const uint32_t nrec = off[K + 1];
        for (uint32_t q = tid; q < nrec; q += FS_NT) {
            uint32_t c = D.collist[c0 + q], pos = c0 + q;
            int a = D.ca[c], b = D.cb[c];
            const bool ovf = D.ccol[c] < 0;
            float na = 0.0f, nb2 = 0.0f;
            if (ovf) { na = (float)ld_dev(&D.ovcnt[a]); if (b >= 0) nb2 = (float)ld_dev(&D.ovcnt[b]); }
            D.rq_c[pos] = c;
            D.rq_ab[pos] = make_int2(a - (int)bd0, b >= 0 ? b - (int)bd0 : -1);
            const LV ra = lv4(D.cra[c]), rb = lv4(D.crb[c]), nn = lv4(D.cn[c]), t1 = lv4(D.ct1[c]), t2 = lv4(D.ct2[c]);
            D.rq_0[pos] = f4(ra, ovf ? D.csmn[c] : D.cmn[c]);
            D.rq_1[pos] = f4(rb, D.cbias[c]);
            D.rq_2[pos] = f4(nn, ovf ? D.csmt1[c] : D.cmt1[c]);
            D.rq_3[pos] = f4(t1, ovf ? D.csmt2[c] : D.cmt2[c]);
            D.rq_4[pos] = f4(t2, na);
            D.rq_5[pos] = make_float4(D.cjn[c], D.cjt1[c], D.cjt2[c], nb2);
        }
        if (tid == 0) {
            uint32_t np = 0;
            if (nchain) { ps_type[np] = 3; ps_off[np] = 0; ps_cnt[np] = hw; np++; }
            for (uint32_t col = 0; col < K; col++) if (cnt[col]) { ps_type[np] = 0; ps_off[np] = c0 + off[col]; ps_cnt[np] = cnt[col]; np++; }
            if (no) { ps_type[np] = 1; ps_off[np] = oo; ps_cnt[np] = no; np++; ps_type[np] = 2; ps_off[np] = 0; ps_cnt[np] = nov; np++; }
            npass = np;
        }
        __syncthreads();
        // 10 iterations of the passes; each thread's record of the next pass is loaded before this pass's barrier
        const uint32_t np = npass, total = 10 * np;
        SRec cur, nxt;
        bool has = false, hasn = false;
        if (np && ps_type[0] < 2 && tid < ps_cnt[0]) { cur = srec_load(D, ps_off[0] + tid); has = true; }
        for (uint32_t x = 0; x < total; x++) {
            const uint32_t ps = x % np, ty = ps_type[ps], po = ps_off[ps], pn = ps_cnt[ps];
            hasn = false;
            if (x + 1 < total && np > 1) {  // (with one pass, its records are the ones being updated: kept in cur)
                uint32_t q = (x + 1) % np;
                if (ps_type[q] < 2 && tid < ps_cnt[q]) { nxt = srec_load(D, ps_off[q] + tid); hasn = true; }
            }
            if (ty == 0) {
                if (has) srec_solve(D, B, cur, po + tid);
                for (uint32_t t = tid + FS_NT; t < pn; t += FS_NT) { SRec r = srec_load(D, po + t); srec_solve(D, B, r, po + t); }
            } else if (ty == 1) {
                if (has) srec_ovf(D, B, cur, po + tid);
                for (uint32_t t = tid + FS_NT; t < pn; t += FS_NT) { SRec r = srec_load(D, po + t); srec_ovf(D, B, r, po + t); }
            } else if (ty == 2) {
                for (uint32_t k = tid >> 6; k < pn; k += FS_NT / 64) {
                    uint32_t b = D.ovbody[bd0 + k];
                    lbody_ovf(D, B, b - bd0, b, ld_dev(&D.ovcnt[b]), lane);
                }
            } else {
                // the static chains: a lane per body, its state in registers for the whole chain
                for (uint32_t lb = tid; lb < pn; lb += FS_NT) {
                    LV v = lv4(B.v[lb]), w = lv4(B.w[lb]);
                    if (chain_body(D, bd0 + lb, v, w, B.iiw(lb), B.im[lb])) { B.v[lb] = f4(v); B.w[lb] = f4(w); }
                }
            }
            __syncthreads();
            if (np > 1) { cur = nxt; has = hasn; }
        }
        // impulses back to the contacts, bodies back to global memory
        for (uint32_t q = tid; q < nrec; q += FS_NT) {
            uint32_t pos = c0 + q, c = D.rq_c[pos];
            float4 j = D.rq_5[pos];
            D.cjn[c] = j.x; D.cjt1[c] = j.y; D.cjt2[c] = j.z;
        }
        for (uint32_t lb = tid; lb < hw; lb += FS_NT) { D.v[bd0 + lb] = B.v[lb]; D.w[bd0 + lb] = B.w[lb]; }
        return;
    }
    // 10 iterations: the static chains, colour after colour, then the overflow pass
    for (int it = 0; it < 10; it++) {
        if (nchain) {
            for (uint32_t lb = tid; lb < hw; lb += FS_NT) {
                const uint32_t b = bd0 + lb;
                LV v = lv4(D.v[b]), w = lv4(D.w[b]);
                if (chain_body(D, b, v, w, load_m(D.iiw + 3 * (size_t)b), D.im[b])) { D.v[b] = f4(v); D.w[b] = f4(w); }
            }
            __syncthreads();
        }
        for (uint32_t col = 0; col < K; col++) {
            const uint32_t n = cnt[col], o = c0 + off[col];
            if (!n) continue;
            for (uint32_t t = tid; t < n; t += FS_NT) solve_contact(D, D.collist[o + t]);
            __syncthreads();
        }
        if (no) {
            for (uint32_t t = tid; t < no; t += FS_NT) {
                uint32_t c = D.collist[oo + t];
                int a = D.ca[c], b = D.cb[c];
                ovf_contact(D, c, (float)ld_dev(&D.ovcnt[a]), b >= 0 ? (float)ld_dev(&D.ovcnt[b]) : 0.0f);
            }
            __syncthreads();
            for (uint32_t k = tid >> 6; k < nov; k += FS_NT / 64) {
                uint32_t b = D.ovbody[bd0 + k];
                ovf_body(D, b, ld_dev(&D.ovcnt[b]), lane);
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------- pinch, impact, integrate
__device__ void kill_body_dev(LD& D, int b) {
    uint32_t o = D.boff[b], n = D.bcnt[b];
    for (uint32_t t = 0; t < n; t++) {
        uint32_t i = D.dyn[o + t];
        if (D.body[i] == b) { D.mat[i] = 0; D.body[i] = -1; }
    }
    D.alive[b] = 0;
    uint32_t s = atomicAdd(&D.ctr[C_NKILL], 1u);
    if (s < D.nbd) D.klist[s] = b;
}
__global__ void __launch_bounds__(NT) k_pinch(LD D) {
    GS(ub, D.nub) {
        const uint32_t b = D.ubody[ub];
        if (!D.alive[b] || D.asleep[b]) continue;
        LV w = lv4(D.w[b]), wpre = lv4(D.wpre[b]);
        LV dw = sub(w, wpre);
        float tip = len(dw) * D.rad[b];
        if (tip > 6.0f) D.w[b] = f4(madd(wpre, dw, 6.0f / tip));
        LV v = lv4(D.v[b]), vpre = lv4(D.vpre[b]);
        LV dv = sub(v, vpre);
        float m = len(dv);
        if (m < 8.0f) continue;
        if (D.bcnt[b] == 0) continue;
        uint32_t f = D.dyn[D.boff[b]];
        if (D.mat[f] >= SLUG) continue;
        if (D.count[b] <= 6u && D.sc[D.bd_scene[b]].blast_t <= 0.0f) kill_body_dev(D, b);
        else D.v[b] = f4(madd(vpre, dv, 8.0f / m));
    }
}
__global__ void __launch_bounds__(NT) k_impact_flag(LD D) {
    uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    GS(c, ncs) {
        if (!D.con[c]) continue;
        int ma = D.mat[D.cba[c]];
        if (ma == 0 || ma >= SLUG) continue;
        float jn = D.cjn[c], vn0 = D.cvn0[c];
        bool hit = D.cbb[c] != NONE && D.cbo[c] < 0 && jn * vn0 > 4000.0f;
        bool crush = !(vn0 < 1.0f) && !(0.5f * jn * vn0 < 150.0f);
        if (!hit && !crush) continue;
        atomicAdd(&D.ctr[C_NIMPACT], 1u);
        atomicAdd(&D.xicnt[D.cscene[c]], 1u);
        D.iflag[c] = 1;
    }
}
__device__ __forceinline__ void integrate_body(LD& D, uint32_t b, float h);
// integrate (src/lib.rs:1703-1745), then the body's wake request from this substep's contacts, both per body
__global__ void __launch_bounds__(NT) k_integrate(LD D, float h) {
    GS(ub, D.nub) {
        const uint32_t b = D.ubody[ub];
        integrate_body(D, b, h);
        if (D.wake[b]) { D.wake[b] = 0; if (D.alive[b]) { D.asleep[b] = 0; D.sleep[b] = 0.0f; } }
    }
}
__device__ __forceinline__ void integrate_body(LD& D, uint32_t b, float h) {
    {
        if (!D.alive[b] || D.asleep[b]) return;
        LV v = lv4(D.v[b]), w = lv4(D.w[b]);
        float vl = len(v);
        if (vl > 60.0f) v = mul(v, 60.0f / vl);
        float a = len(w);
        float amax = minf(25.0f, 16.0f / maxf(D.rad[b], 0.3f));
        if (a > amax) w = mul(w, amax / a);
        if (D.cool[b] > 0.0f) D.cool[b] -= h;
        LV pos = madd(lv4(D.pos[b]), v, h);
        D.q[b] = f4q(q_int(lq4(D.q[b]), w, h));
        v = mul(v, 1.0f - 0.01f * h);
        w = mul(w, 1.0f - 0.05f * h);
        bool small = D.count[b] <= 3u;
        if (D.nc[b] > 0u && small) {
            w = mul(w, 1.0f - minf(7.0f * h, 0.5f));
            v = mul(v, 1.0f - minf(2.0f * h, 0.3f));
        }
        float v2 = dot(v, v);
        float rr = D.rad[b];
        float w2 = dot(w, w) * rr * rr;
        float vt = small ? 0.16f : 0.06f, wt = small ? 0.3f : 0.04f;
        if (v2 < vt && w2 < wt && D.nc[b] > 0u && D.maxr[b] < 0.95f) {
            D.sleep[b] += h;
            if (D.sleep[b] > 0.45f) { D.asleep[b] = 1; v = lv(0.0f, 0.0f, 0.0f); w = lv(0.0f, 0.0f, 0.0f); }
        } else D.sleep[b] = 0.0f;
        D.pos[b] = f4(pos); D.v[b] = f4(v); D.w[b] = f4(w);
        if (pos.y < -15.0f || fabsf(pos.x) > 190.0f || fabsf(pos.z) > 190.0f || pos.y > 120.0f) kill_body_dev(D, b);
    }
}

// ---------------------------------------------------------------- load field
__device__ __forceinline__ int linked(const LD& D, uint32_t i, int d, float* hp_out) {
    uint32_t j = D.nb6[6 * (size_t)i + d];
    if (j == NONE || D.mat[j] == 0 || D.body[j] != D.body[i]) return -1;
    float4 bi = (d & 1) == 0 ? D.bond[i] : D.bond[j];
    int ax = d >> 1;
    float hp = ax == 0 ? bi.x : (ax == 1 ? bi.y : bi.z);
    if (hp > 0.0f) { if (hp_out) *hp_out = hp; return (int)j; }
    return -1;
}
__device__ __forceinline__ float rot_el(const LD& D, int b, int row, int col) {
    float4 r = D.rot[3 * (size_t)b + row];
    return col == 0 ? r.x : (col == 1 ? r.y : r.z);
}
__device__ __forceinline__ float cond_w(const LD& D, uint32_t i, uint32_t j, int a) {
    const Tables& T = tables();
    float k = minf(T.stiff[D.mat[i]], T.stiff[D.mat[j]]);
    int b = D.body[i];
    if (b < 0) return k * (a == 1 ? 1.0f : 0.3f);
    float u = fabsf(rot_el(D, b, 1, a));
    return k * (0.3f + 0.7f * u);
}
__device__ __forceinline__ float axis_up(const LD& D, uint32_t i, int a) {
    int b = D.body[i];
    if (b < 0) return a == 1 ? 1.0f : 0.0f;
    return rot_el(D, b, 1, a);
}
__device__ float sor_brick(LD& D, uint32_t i, float omega, uint32_t frame) {
    const Tables& T = tables();
    int mi = D.mat[i];
    float sk = 0.0f, skp = 0.0f;
    for (int d = 0; d < 6; d++) {
        int j = linked(D, i, d, nullptr);
        if (j < 0) continue;
        float k = cond_w(D, i, (uint32_t)j, d >> 1);
        sk += k; skp += k * D.phi[j];
    }
    if (D.body[i] < 0) { if (D.bxyz[i].y == 0 && D.bond[i].w > 0.0f) sk += GROUNDK * T.stiff[mi]; }
    else if (frame - D.sup[i] < 10u) sk += GROUNDK * T.stiff[mi];
    if (sk > 0.0f) {
        float target = (T.mass[mi] + D.ext[i] + skp) / sk;
        float d = omega * (target - D.phi[i]);
        D.phi[i] += d;
        return fabsf(d);
    }
    return 0.0f;
}
__device__ __forceinline__ bool in_struct(const LD& D, uint32_t i, int s_local) {
    int m = D.mat[i];
    return D.sid[i] == s_local && m != 0 && m < SLUG && D.body[i] < 0;
}

// ---------------------------------------------------------------- structure connectivity (src/lib.rs:1085-1100)
// The bricks of a structure that reach the ground through linked bonds are the connected components (under "linked":
// same body, bond above zero; symmetric between a brick and its neighbour) that hold a grounded footing. Found for every
// listed structure at once (list[k] = global index, scene-local index; blockIdx.y = k) by union-find over the
// structures' bricks: each brick hooks, for its three positive directions, the larger root under the smaller with
// atomic minima (parents only move to smaller indices of the same component, so whatever a racing store leaves is still
// an ancestor); once the forest is final the footings mark their roots with this pass's token, and a brick is reached
// when its root is marked (roots found by walking up without stores: a path-halving store racing a root store could
// leave a brick pointing at an ancestor below its root). The set equals the host's breadth-first search, whatever order lanes run
// in. The result: one bit per brick of each structure's range at out + nlist + woff[k]; out[k] = 1 when a linked brick
// lay outside the range (bonds never leave a structure; the host then searches that structure itself).
__device__ __forceinline__ uint32_t cc_load(const uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void cc_store(uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ uint32_t cc_find(uint32_t* par, uint32_t x) {
    while (true) {
        const uint32_t p = cc_load(&par[x]);
        if (p == x) return x;
        const uint32_t g = cc_load(&par[p]);
        if (g != p) cc_store(&par[x], g);
        x = g;
    }
}
// joins the trees of a and b: the larger root is hooked under the smaller by an atomic minimum, retried from the
// value found when another hook got there first (the forest ends the same whatever order the hooks land in)
__device__ __forceinline__ void cc_union(uint32_t* par, uint32_t a, uint32_t b) {
    while (true) {
        a = cc_find(par, a);
        b = cc_find(par, b);
        if (a == b) return;
        if (a < b) { const uint32_t t = a; a = b; b = t; }
        const uint32_t old = atomicMin(&par[a], b);
        if (old == a) return;
        a = old;
    }
}
// (each pass: blockIdx.y the listed structure, blockIdx.x and the lanes stride its brick range)
#define CC_RANGE(k, lo, hi)                                                                   \
    const int2 e_ = list[blockIdx.y];                                                         \
    const uint32_t lo = (uint32_t)D.st_lo[e_.x], hi = (uint32_t)D.st_hi[e_.x];              \
    const uint32_t k = blockIdx.y;                                                           \
    (void)k;
__global__ void __launch_bounds__(NT) k_cc_init(LD D, const int2* list) {
    CC_RANGE(k, lo, hi)
    for (uint32_t i = lo + blockIdx.x * NT + threadIdx.x; i < hi; i += gridDim.x * NT) D.cq[i] = i;
}
__global__ void __launch_bounds__(NT) k_cc_hook(LD D, const int2* list, uint32_t* out) {
    CC_RANGE(k, lo, hi)
    bool outside = false;
    for (uint32_t i = lo + blockIdx.x * NT + threadIdx.x; i < hi; i += gridDim.x * NT) {
        if (D.mat[i] == 0) continue;
        for (int d = 0; d < 6; d += 2) {
            const int jj = linked(D, i, d, nullptr);
            if (jj < 0) continue;
            const uint32_t j = (uint32_t)jj;
            if (j < lo || j >= hi) { outside = true; continue; }
            cc_union(D.cq, i, j);
        }
        // (the other direction of a lateral or vertical link is the neighbour's positive one: d odd adds nothing)
    }
    if (outside) out[k] = 1u;
}
// the forest is final after k_cc_hook: a root found by walking up without stores is the component's
__device__ __forceinline__ uint32_t cc_root(const uint32_t* par, uint32_t x) {
    while (true) { const uint32_t p = cc_load(&par[x]); if (p == x) return x; x = p; }
}
__global__ void __launch_bounds__(NT) k_cc_root(LD D, const int2* list, uint32_t token) {
    CC_RANGE(k, lo, hi)
    for (uint32_t i = lo + blockIdx.x * NT + threadIdx.x; i < hi; i += gridDim.x * NT)
        if (in_struct(D, i, e_.y) && D.bxyz[i].y == 0 && D.bond[i].w > 0.0f) cc_store(&D.creach[cc_root(D.cq, i)], token);
}
__global__ void __launch_bounds__(NT) k_cc_bits(LD D, const int2* list, const uint32_t* woff, uint32_t nlist, uint32_t* out, uint32_t token) {
    CC_RANGE(k, lo, hi)
    const uint32_t n = hi - lo, words = (n + 31u) / 32u;
    uint32_t* bits = out + nlist + woff[k];
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t c = wave * 64u; c < n; c += nwaves * 64u) {
        const uint32_t i = lo + c + (uint32_t)lane;
        const bool r = c + (uint32_t)lane < n && D.mat[i] != 0 && cc_load(&D.creach[cc_root(D.cq, i)]) == token;
        const unsigned long long m = __ballot(r);
        if (lane == 0) {
            bits[c / 32u] = (uint32_t)m;
            if (c / 32u + 1u < words) bits[c / 32u + 1u] = (uint32_t)(m >> 32);
        }
    }
}
// The parts of split bodies (rebuild_body, src/lib.rs:930-1002): the live bricks of every body to rebuild, listed by
// the host, joined over their linked bonds by the same union-find; each listed brick's root (the smallest brick index
// of its part) comes back. A linked neighbour of a listed brick is a live brick of the same body, so it is listed too.
__global__ void __launch_bounds__(NT) k_split_init(LD D, const uint32_t* list, uint32_t n) { GS(t, n) D.cq[list[t]] = list[t]; }
__global__ void __launch_bounds__(NT) k_split_hook(LD D, const uint32_t* list, uint32_t n) {
    GS(t, n) {
        const uint32_t i = list[t];
        for (int d = 0; d < 6; d += 2) {
            const int j = linked(D, i, d, nullptr);
            if (j >= 0) cc_union(D.cq, i, (uint32_t)j);
        }
    }
}
__global__ void __launch_bounds__(NT) k_split_root(LD D, const uint32_t* list, uint32_t n, uint32_t* out) {
    GS(t, n) out[t] = cc_root(D.cq, list[t]);
}
// ---------------------------------------------------------------- a charge (explode_at, src/lib.rs:2679-2716)
// The bricks a charge may reach, listed by the host in the reference's order (each structure near the charge brick by
// brick, then each body near it along its brick list) with their structure slot (-1: a body's brick). k_blast_event
// gives each listed brick its event (blast_brick, :2679-2696): removed within half its material's blast radius,
// damaged by power (1 - d / 1.35 rr)^2 1.3 within 1.35 rr, and marks the brick as listed this pass (creach = token,
// cq = its rank in the order). k_blast_bonds then applies the damage to every bond a listed brick owns (the positive
// x, y, z links and the footing): a link's two ends each damage it in rank order, each only while the other end is
// not yet removed and the bond is still above zero (damage_brick, :725-771), so the f32 result is the sequential one.
// k_blast_remove clears removed bricks; k_blast_bodies gives each body near the charge the blast impulse brick by brick
// along its list, wakes it and limits its speeds (apply_blast_to_body :898-916, clamp_vel :917-929).
enum : uint32_t { BL_NONE = 0, BL_REMOVE = 1, BL_DAMAGE = 2, BL_BROKE = 4 };
struct BlastParams { float cx, cy, cz, r, power, br, bp, pad; uint32_t token, n, pad1, pad2; };
__global__ void __launch_bounds__(NT) k_blast_event(LD D, BlastParams P, const int2* ent, uint32_t* ev, float* frac) {
    const Tables& T = tables();
    GS(t, P.n) {
        const uint32_t i = (uint32_t)ent[t].x;
        const int sl = ent[t].y;
        const int m = D.mat[i];
        uint32_t e = BL_NONE;
        float fr = 0.0f;
        const bool active = sl >= 0 ? in_struct(D, i, sl) : (m != 0 && m < SLUG);
        if (active) {
            const LV p = D.body[i] < 0 ? lat_pos(D.bxyz[i].x, D.bxyz[i].y, D.bxyz[i].z) : lv4(D.wpos[i]);
            const float d = len(sub(p, lv(P.cx, P.cy, P.cz)));
            const float rr = P.r * T.blast[m];
            if (d < rr * 0.5f) e = BL_REMOVE;
            else if (d < rr * 1.35f) { const float f = 1.0f - d / (rr * 1.35f); e = BL_DAMAGE; fr = P.power * f * f * 1.3f; }
        }
        // the entry acts for the brick (its events, the bonds it owns): a live brick of the structure, or of the body
        if (m != 0 && (sl < 0 || (D.sid[i] == sl && D.body[i] < 0))) { D.creach[i] = P.token; D.cq[i] = t; }
        ev[t] = e;
        frac[t] = fr;
    }
}
__device__ __forceinline__ float& f4_ref(float4& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : (a == 2 ? v.z : v.w)); }
__device__ __forceinline__ float f4_get(const float4& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : (a == 2 ? v.z : v.w)); }
__global__ void __launch_bounds__(NT) k_blast_bonds(LD D, BlastParams P, const int2* ent, uint32_t* ev, const float* frac, float4* bond_out) {
    GS(t, P.n) {
        const uint32_t o = (uint32_t)ent[t].x;
        float4 bd = D.bond[o];
        if (D.creach[o] == P.token && D.cq[o] == t) {   // (the entry that acts for the brick)
            const float4 bm = D.bmax[o];
            const uint32_t eo = ev[t] & 3u;
            // alive at rank r: not empty, and not removed by its own event before r
            auto alive = [&](uint32_t y, uint32_t r) {
                if (D.mat[y] == 0) return false;
                if (D.creach[y] != P.token) return true;
                const uint32_t ry = D.cq[y];
                return !((ev[ry] & 3u) == BL_REMOVE && ry < r);
            };
            for (int a = 0; a < 3; a++) {
                const uint32_t j = D.nb6[6 * (size_t)o + 2 * a];
                if (j == NONE) continue;
                const bool jl = D.creach[j] == P.token;
                const uint32_t tj = jl ? D.cq[j] : 0u;
                const uint32_t ej = jl ? (ev[tj] & 3u) : BL_NONE;
                // the link's damage events in rank order: (rank, acting entry, other end)
                uint32_t r0 = t, r1 = tj;
                bool h0 = eo == BL_DAMAGE, h1 = ej == BL_DAMAGE;
                uint32_t x0 = o, x1 = j, y0 = j, y1 = o;
                if (h0 && h1 && r1 < r0) { uint32_t q = r0; r0 = r1; r1 = q; q = x0; x0 = x1; x1 = q; q = y0; y0 = y1; y1 = q; bool hh = h0; h0 = h1; h1 = hh; }
                else if (!h0 && h1) { r0 = r1; x0 = x1; y0 = y1; h0 = true; h1 = false; }
                float v = f4_get(bd, a);
                const float vm = f4_get(bm, a);
                for (int k = 0; k < 2; k++) {
                    const bool h = k == 0 ? h0 : h1;
                    if (!h) continue;
                    const uint32_t rk = k == 0 ? r0 : r1, xk = k == 0 ? x0 : x1, yk = k == 0 ? y0 : y1;
                    if (!alive(yk, rk)) continue;
                    if (v > 0.0f) {
                        v -= vm * frac[rk] + 0.0f;
                        if (v <= vm * 0.03f) { v = 0.0f; atomicOr(&ev[rk], BL_BROKE); }
                    }
                    (void)xk;
                }
                f4_ref(bd, a) = v;
            }
            if (eo == BL_DAMAGE && bd.w > 0.0f) {   // the footing: only the brick's own event
                bd.w -= bm.w * frac[t] * 0.5f + 0.0f * 0.5f;
                if (bd.w <= 0.0f) { bd.w = 0.0f; atomicOr(&ev[t], BL_BROKE); }
            }
            D.bond[o] = bd;
        }
        bond_out[t] = bd;
    }
}
__global__ void __launch_bounds__(NT) k_blast_remove(LD D, BlastParams P, const int2* ent, const uint32_t* ev) {
    GS(t, P.n) if ((ev[t] & 3u) == BL_REMOVE) D.mat[ent[t].x] = 0;
}
// one wavefront per body (lane 0): its list is body slot, first and last entry of `list`
__global__ void __launch_bounds__(64) k_blast_bodies(LD D, BlastParams P, const int4* bodies, const uint32_t* list, float4* vw_out) {
    if (threadIdx.x != 0) return;
    const Tables& T = tables();
    const int4 B = bodies[blockIdx.x];
    const int b = B.x;
    D.asleep[b] = 0; D.sleep[b] = 0.0f;
    LV v = lv4(D.v[b]), w = lv4(D.w[b]);
    const LV bp = lv4(D.pos[b]), c = lv(P.cx, P.cy, P.cz);
    const LM iiw = load_m(D.iiw + 3 * (size_t)b);
    const float im = D.im[b];
    for (int t = B.y; t < B.z; t++) {
        const uint32_t k = list[t];
        const LV p = lv4(D.wpos[k]);
        const LV d = sub(p, c);
        const float dist = len(d);
        if (dist < P.br) {
            const float f = P.bp * 1600.0f * (1.0f - dist / P.br);
            LV dir = dist > 1e-3f ? mul(d, 1.0f / dist) : lv(0.0f, 1.0f, 0.0f);
            dir = add(dir, lv(0.0f, 0.35f, 0.0f));
            const float mm = maxf(T.mass[D.mat[k]], 25.0f);
            const LV j = mul(dir, f * mm / 300.0f);
            v = madd(v, j, im);
            const LV r = sub(p, bp);
            w = add(w, mv(iiw, cross(r, j)));
        }
    }
    const float vl = len(v);
    if (vl > 60.0f) v = mul(v, 60.0f / vl);
    const float a = len(w);
    const float amax = minf(25.0f, 16.0f / maxf(D.rad[b], 0.3f));
    if (a > amax) w = mul(w, amax / a);
    D.v[b] = f4(v); D.w[b] = f4(w);
    vw_out[2 * (size_t)blockIdx.x] = f4(v); vw_out[2 * (size_t)blockIdx.x + 1] = f4(w);
}

// ---------------------------------------------------------------- impacts: the crush walk on the device
// (impact_damage, src/lib.rs:1610-1642; crush_at, :1643-1702) One wavefront per scene takes the scene's impact
// candidates in contact order (the flags k_impact_flag set, cleared as they are read) with the scene's budget of 2,500
// crushed bricks. A crush walks from the hit brick: of the next 24 queued bricks the one nearest the contact point is
// crushed next (its energy cost paid), its linked neighbours of the same body or structure queued; when the energy left
// is below a brick's cost that brick is damaged instead and the walk ends. Each crushed brick is then pulverised (with
// the material's share, from the scene's random stream) or severed. The queue (at most 1 + 6 x 600 bricks) and the
// queued bricks' distances live in LDS; lane 0 makes the walk's decisions and writes, the lanes pick the nearest of
// the window and test the six neighbours. What changed is logged for the host mirror: bricks, bodies to rebuild,
// structures to activate, check and mark broken.
constexpr int XQ = 4096, XCR = 600;
enum : uint32_t { XL_BRICK = 0, XL_DIRTY = 1, XL_ACT = 2, XL_CONN = 3, XL_BROKE = 4, XL_VEL = 5 };
struct CrushP { uint32_t* log; uint32_t* nlog; uint32_t* rng; uint32_t logcap, pad; };
__device__ __forceinline__ void x_log(const CrushP& P, uint32_t kind, uint32_t idx) {
    const uint32_t k = atomicAdd(P.nlog, 1u);
    if (k < P.logcap) { P.log[2 * (size_t)k] = kind; P.log[2 * (size_t)k + 1] = idx; }
}
__device__ __forceinline__ void x_static(const LD& D, const CrushP& P, const SceneK& K, uint32_t i) {
    const int sl = D.sid[i];
    if (sl < K.nst) x_log(P, XL_ACT, (uint32_t)(K.st0 + sl));
}
// damage_brick (src/lib.rs:725-771) with no absolute part
__device__ void x_damage(LD& D, const CrushP& P, const SceneK& K, uint32_t i, float frac) {
    uint32_t broke = 0;
    for (int d = 0; d < 6; d++) {
        const uint32_t j = D.nb6[6 * (size_t)i + d];
        if (j == NONE || D.mat[j] == 0) continue;
        const uint32_t o = (d & 1) == 0 ? i : j;
        const int a = d >> 1;
        float4 b = D.bond[o];
        float& bd = f4_ref(b, a);
        if (bd > 0.0f) {
            const float bm = f4_get(D.bmax[o], a);
            bd -= bm * frac + 0.0f;
            if (bd <= bm * 0.03f) { bd = 0.0f; broke++; }
            D.bond[o] = b;
            x_log(P, XL_BRICK, o);
        }
    }
    float4 b = D.bond[i];
    if (b.w > 0.0f) {
        b.w -= D.bmax[i].w * frac * 0.5f + 0.0f * 0.5f;
        if (b.w <= 0.0f) { b.w = 0.0f; broke++; }
        D.bond[i] = b;
        x_log(P, XL_BRICK, i);
    }
    if (D.body[i] < 0) {
        x_static(D, P, K, i);
        const int sl = D.sid[i];
        if (broke > 0 && sl < K.nst) { x_log(P, XL_CONN, (uint32_t)(K.st0 + sl)); x_log(P, XL_BROKE, (uint32_t)(K.st0 + sl)); }
    } else if (broke > 0) x_log(P, XL_DIRTY, (uint32_t)D.body[i]);
}
__device__ __forceinline__ float x_rnd(uint32_t& x) { x ^= x << 13; x ^= x >> 17; x ^= x << 5; return (float)(x >> 8) / 16777216.0f; }
__device__ __forceinline__ float x_dist2(const LD& D, uint32_t i, LV p) {
    const LV q = D.body[i] < 0 ? lat_pos(D.bxyz[i].x, D.bxyz[i].y, D.bxyz[i].z) : lv4(D.wpos[i]);
    const LV d = sub(q, p);
    return dot(d, d);
}
struct XWalk { uint32_t* Q; float* DD; uint32_t* CR; uint32_t* shu; };
// crush_at from brick `start` with energy e at point p, at most cap bricks; every lane calls it (uniform arguments)
__device__ uint32_t x_crush(LD& D, const CrushP& P, const SceneK& K, const XWalk& W, uint32_t& rng, uint32_t start, float e, LV p,
                            uint32_t cap) {
    const Tables& T = tables();
    const int lane = threadIdx.x;
    const int m0 = D.mat[start];
    if (m0 == 0 || m0 >= SLUG) return 0;
    if (e < T.eb[m0] * 0.15f) return 0;
    const int bodyv = D.body[start];
    const int sidv = D.sid[start];
    if (lane == 0) {
        const uint32_t st0 = D.xtok[blockIdx.x] + 1u;
        D.xtok[blockIdx.x] = st0;
        W.shu[0] = st0;
        W.Q[0] = start; W.DD[0] = x_dist2(D, start, p);
        D.xstamp[start] = st0;
    }
    __syncthreads();
    const uint32_t st = W.shu[0];
    int qn = 1, head = 0, ncr = 0;
    uint32_t crushed = 0;
    const uint32_t maxn = cap < 600u ? cap : 600u;
    while (head < qn && crushed < maxn) {
        const int lim = min(head + 24, qn);
        // the nearest of the window: the first of the smallest distances below 1e18 (none: the window's first)
        unsigned long long key = ~0ull;
        if (lane < lim - head) {
            const float v = W.DD[head + lane];
            if (v < 1e18f) key = ((unsigned long long)__float_as_uint(v) << 32) | (unsigned)lane;
        }
        for (int h = 32; h >= 1; h >>= 1) { const unsigned long long o = __shfl_xor(key, h, 64); key = o < key ? o : key; }
        const int best = key == ~0ull ? head : head + (int)(key & 0xffffffffull);
        __syncthreads();
        if (lane == 0 && best != head) {
            const uint32_t q = W.Q[head]; W.Q[head] = W.Q[best]; W.Q[best] = q;
            const float f = W.DD[head]; W.DD[head] = W.DD[best]; W.DD[best] = f;
        }
        __syncthreads();
        const uint32_t i = W.Q[head];
        head++;
        const int mm = D.mat[i];
        if (mm == 0) continue;
        const float cost = T.eb[mm];
        if (e < cost) {
            if (lane == 0) x_damage(D, P, K, i, 0.6f * e / cost);
            __syncthreads();
            break;
        }
        e -= cost;
        if (lane == 0) W.CR[ncr] = i;
        ncr++;
        crushed++;
        // the six neighbours in order: linked, not queued by this walk, of the same body and structure
        int j = -1;
        if (lane < 6) {
            const int jj = linked(D, i, lane, nullptr);
            if (jj >= 0 && D.xstamp[jj] != st && D.body[jj] == bodyv && D.sid[jj] == sidv) j = jj;
        }
        const unsigned long long am = __ballot(j >= 0);
        if (j >= 0) {
            const int pos = qn + __popcll(am & ((1ull << lane) - 1ull));
            if (pos < XQ) { W.Q[pos] = (uint32_t)j; W.DD[pos] = x_dist2(D, (uint32_t)j, p); D.xstamp[j] = st; }
        }
        qn = min(qn + __popcll(am), XQ);
        __syncthreads();
    }
    if (ncr == 0) return 0;
    if (lane == 0) {
        for (int t = 0; t < ncr; t++) {
            const uint32_t i = W.CR[t];
            const int mm = D.mat[i];
            if (mm == 0) continue;
            if (x_rnd(rng) < T.pulv[mm]) {
                // remove_brick
                D.mat[i] = 0;
                x_log(P, XL_BRICK, i);
                if (D.body[i] < 0) {
                    const int sl = D.sid[i];
                    if (sl < K.nst) x_log(P, XL_CONN, (uint32_t)(K.st0 + sl));
                    x_static(D, P, K, i);
                } else x_log(P, XL_DIRTY, (uint32_t)D.body[i]);
            } else {
                // sever_brick
                D.bond[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                x_log(P, XL_BRICK, i);
                for (int d = 1; d < 6; d += 2) {
                    const uint32_t jn = D.nb6[6 * (size_t)i + d];
                    if (jn == NONE) continue;
                    float4 b = D.bond[jn];
                    f4_ref(b, d >> 1) = 0.0f;
                    D.bond[jn] = b;
                    x_log(P, XL_BRICK, jn);
                }
                if (bodyv < 0) {
                    const int sl = D.sid[i];
                    if (sl < K.nst) x_log(P, XL_CONN, (uint32_t)(K.st0 + sl));
                    x_static(D, P, K, i);
                } else x_log(P, XL_DIRTY, (uint32_t)bodyv);
            }
        }
    }
    __syncthreads();
    return crushed;
}
__global__ void __launch_bounds__(64) k_crush(LD D, CrushP P) {
    __shared__ uint32_t Q[XQ];
    __shared__ float DD[XQ];
    __shared__ uint32_t CR[XCR];
    __shared__ uint32_t shu[4];
    const XWalk W{Q, DD, CR, shu};
    const Tables& T = tables();
    const int s = blockIdx.x, lane = threadIdx.x;
    if (s >= (int)D.nscene) return;
    const SceneK K = D.sc[s];
    const uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    const uint32_t c0 = min(D.sc_c0[s], ncs), c1 = s + 1 < (int)D.nscene ? min(D.sc_c0[s + 1], ncs) : ncs;
    uint32_t left = D.xicnt[s];   // (the scene's candidates; at least as many as it flagged)
    if (left == 0) return;
    uint32_t rng = P.rng[s];
    uint32_t bud = 2500u;
    for (uint32_t cb = c0; cb < c1 && left > 0; cb += 64) {
        const uint32_t cl = cb + (uint32_t)lane;
        const bool f = cl < c1 && D.iflag[cl] != 0;
        unsigned long long m = __ballot(f);
        if (f) D.iflag[cl] = 0;
        left -= min(left, (uint32_t)__popcll(m));
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const uint32_t c = cb + (uint32_t)l;
            if (!D.con[c]) continue;
            const uint32_t ia = D.cba[c], bb = D.cbb[c];
            const int ma = D.mat[ia];
            if (ma == 0 || ma >= SLUG) continue;
            const float jn = D.cjn[c], vn0 = D.cvn0[c];
            if (bb != NONE && D.cbo[c] < 0 && jn * vn0 > 4000.0f) {
                if (lane == 0 && D.mat[bb] != 0 && D.body[bb] < 0) {
                    const int sl = D.sid[bb];
                    if (sl >= 0 && sl < K.nst) x_log(P, XL_ACT, (uint32_t)(K.st0 + sl));
                }
            }
            if (D.mat[ia] == 0) continue;
            if (vn0 < 1.0f) continue;
            const float e = 0.5f * jn * vn0;
            if (e < 150.0f) continue;
            const int jb = bb != NONE ? (int)bb : -1;
            const int mb = jb >= 0 ? D.mat[jb] : 0;
            const float soft_a = 1.0f / T.eb[ma];
            const float soft_b = (jb < 0 || mb == 0 || mb >= SLUG) ? 0.0f : 1.0f / T.eb[mb];
            if (soft_a + soft_b <= 0.0f) continue;
            const float ea = e * soft_a / (soft_a + soft_b);
            const float4 p4 = D.cp[c];
            const LV p = lv(p4.x, p4.y, p4.z);
            if (ea > 0.0f && bud > 0) { const uint32_t k = x_crush(D, P, K, W, rng, ia, ea, p, bud); bud -= k < bud ? k : bud; }
            if (soft_b > 0.0f && bud > 0) { const uint32_t k = x_crush(D, P, K, W, rng, (uint32_t)jb, e - ea, p, bud); bud -= k < bud ? k : bud; }
        }
    }
    if (lane == 0) { P.rng[s] = rng; D.xicnt[s] = 0u; }
}
__global__ void __launch_bounds__(NT) k_crush_gather(LD D, const uint32_t* log, const uint32_t* nlog, uint32_t cap, float4* bond_out, uint32_t* mat_out) {
    const uint32_t n = min(*nlog, cap);
    GS(k, n) {
        if (log[2 * (size_t)k] != XL_BRICK) continue;
        const uint32_t i = log[2 * (size_t)k + 1];
        bond_out[k] = D.bond[i];
        mat_out[k] = D.mat[i];
    }
}

// block max of two values; every thread gets the result
template <int BT>
__device__ __forceinline__ void block_max2(float& a, float& b, float* sh) {
    for (int h = 32; h >= 1; h >>= 1) { a = fmaxf(a, __shfl_xor(a, h, 64)); b = fmaxf(b, __shfl_xor(b, h, 64)); }
    int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { sh[wv] = a; sh[BT / 64 + wv] = b; }
    __syncthreads();
    a = sh[0]; b = sh[BT / 64];
    for (int k = 1; k < BT / 64; k++) { a = fmaxf(a, sh[k]); b = fmaxf(b, sh[BT / 64 + k]); }
    __syncthreads();
}
// the same maximum with one barrier: it also serves as the barrier that ends the caller's last pass, and the
// caller's next barrier comes before sh is written again
template <int BT>
__device__ __forceinline__ void block_max2_tail(float& a, float& b, float* sh) {
    for (int h = 32; h >= 1; h >>= 1) { a = fmaxf(a, __shfl_xor(a, h, 64)); b = fmaxf(b, __shfl_xor(b, h, 64)); }
    int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { sh[wv] = a; sh[BT / 64 + wv] = b; }
    __syncthreads();
    a = sh[0]; b = sh[BT / 64];
    for (int k = 1; k < BT / 64; k++) { a = fmaxf(a, sh[k]); b = fmaxf(b, sh[BT / 64 + k]); }
}

// One workgroup per scene: its structures in order sharing the frame's brick-update budget (src/lib.rs:1268-1281,
// :1068-1079). mode 0: adaptive sweeps of live structures (st_nmin = sweeps asked for); mode 1: warm-up sweeps
// (st_nmin = warm-up sweeps left, written back).
constexpr int ST_NT = 1024;
// the stencil of every brick of the structures that sweep this frame (mode 0: live ones, mode 1: warming ones);
// links and conductances do not change during the sweeps, so each brick update becomes a gather of six potentials
// Structures of at most pk_fit bricks (their potential then sits in the sweep's LDS) get packed rows; their linked
// neighbours lie in the structure's own index range, as bonds join bricks of one structure only.
__global__ void __launch_bounds__(NT) k_stencil(LD D, int mode, uint32_t pk_fit) {
    const Tables& T = tables();
    for (uint32_t g = blockIdx.y; g < D.nst; g += gridDim.y) {
        if (D.st_nmin[g] <= 0) continue;
        if (mode == 0 && !D.st_act[g]) continue;
        int sc = D.st_scene[g];
        int sl = (int)g - D.sc[sc].st0;
        const int lo = D.st_lo[g];
        const bool pk = (uint32_t)(D.st_hi[g] - lo) <= pk_fit;
        int4 rb = D.st_rb[g];
        uint32_t o = rb.x, n = (uint32_t)(rb.y + rb.w);
        for (uint32_t t = blockIdx.x * NT + threadIdx.x; t < n; t += gridDim.x * NT) {
            uint32_t p = o + t, i = D.st_red[p];
            int nbr[6];
            float kk[6];
            float sk = 0.0f;
            bool ok = in_struct(D, i, sl);
            for (int d = 0; d < 6; d++) {
                int j = ok ? linked(D, i, d, nullptr) : -1;
                nbr[d] = j;
                kk[d] = 0.0f;
                if (j < 0) continue;
                float k = cond_w(D, i, (uint32_t)j, d >> 1);
                kk[d] = k;
                sk += k;
            }
            int mi = D.mat[i];
            if (ok && D.bxyz[i].y == 0 && D.bond[i].w > 0.0f) sk += GROUNDK * T.stiff[mi];
            if (pk) {
                uint32_t rel[6], w = (uint32_t)mi & 15u;
                for (int d = 0; d < 6; d++) {
                    rel[d] = nbr[d] < 0 ? 0xFFFFu : (uint32_t)(nbr[d] - lo);
                    if (nbr[d] >= 0) w |= ((uint32_t)D.mat[nbr[d]] & 15u) << (4 + 4 * d);
                }
                D.sten_pk[p] = make_uint4(rel[0] | rel[1] << 16, rel[2] | rel[3] << 16, rel[4] | rel[5] << 16, w);
            } else {
                D.sten_na[p] = make_int4(nbr[0], nbr[1], nbr[2], nbr[3]);
                D.sten_nb[p] = make_int2(nbr[4], nbr[5]);
                D.sten_ka[p] = make_float4(kk[0], kk[1], kk[2], kk[3]);
                D.sten_kb[p] = make_float2(kk[4], kk[5]);
            }
            D.sten_s[p] = make_float2(ok ? sk : -1.0f, ok ? T.mass[mi] + D.ext[i] : 0.0f);
        }
    }
}
// one brick's update from its stencil, in sor_brick's operation order (src/lib.rs:1005-1026)
// The potential of the structure being swept: in LDS when the structure fits (the workgroup copies it in before its
// sweeps and back after), else in global memory.
template <bool LDSP> struct Phi;
template <> struct Phi<true> {
    float* p; int lo; uint32_t n; float* g;
    __device__ __forceinline__ float ld(int i) const { uint32_t k = (uint32_t)(i - lo); return k < n ? p[k] : g[i]; }
    __device__ __forceinline__ void st(int i, float v) const { uint32_t k = (uint32_t)(i - lo); if (k < n) p[k] = v; else g[i] = v; }
};
template <> struct Phi<false> {
    float* p; int lo; uint32_t n; float* g;
    __device__ __forceinline__ float ld(int i) const { return g[i]; }
    __device__ __forceinline__ void st(int i, float v) const { g[i] = v; }
};
// one red-black sweep over a structure's stencil (sor_brick in order within a colour, src/lib.rs:1005-1026); returns
// max|change| / max(phi, 1) over its bricks. U stencil rows per thread are loaded together before any is applied:
// rows of one colour only read bricks of the other colour, so the order of application within a colour is free.
// The packed rows' half-sweep (LDS potential): the same operations on the same values, each conductance read from
// ktab[(mat_i * 9 + mat_j) * 2 + vertical].
__device__ __forceinline__ void half_sweep_pk(const LD& D, uint32_t o, uint32_t n, float om, const Phi<true>& phi,
                                              const float* ktab, float& md, float& mp) {
    constexpr int U = 8;
    for (uint32_t t = threadIdx.x; t < n; t += ST_NT * U) {
        float2 sr[U];
        uint4 q[U];
        uint32_t ii[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint32_t qq = t + u * ST_NT;
            const bool in = qq < n;
            uint32_t p = o + (in ? qq : 0u);
            sr[u] = in ? D.sten_s[p] : make_float2(-1.0f, 0.0f);
            q[u] = D.sten_pk[p]; ii[u] = D.st_red[p];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (sr[u].x < 0.0f) continue;
            const uint32_t rel[6] = {q[u].x & 0xFFFFu, q[u].x >> 16, q[u].y & 0xFFFFu, q[u].y >> 16, q[u].z & 0xFFFFu, q[u].z >> 16};
            const uint32_t w = q[u].w, row = (w & 15u) * 18u;
            float skp = 0.0f;
#pragma unroll
            for (int d = 0; d < 6; d++)
                if (rel[d] != 0xFFFFu) skp += ktab[row + ((w >> (4 + 4 * d)) & 15u) * 2u + ((d >> 1) == 1 ? 1u : 0u)] * phi.p[rel[d]];
            const uint32_t k = ii[u] - (uint32_t)phi.lo;
            float pi = phi.p[k], po = pi;
            if (sr[u].x > 0.0f) {
                float target = (sr[u].y + skp) / sr[u].x;
                float d = om * (target - pi);
                po = pi + d;
                phi.p[k] = po;
                md = fmaxf(md, fabsf(d));
            } else {
                md = fmaxf(md, 0.0f);
            }
            mp = fmaxf(mp, po);
        }
    }
}
template <bool LDSP>
__device__ float sweep_struct(LD& D, uint32_t g, float om, const Phi<LDSP>& phi, float* sh, const float* ktab) {
    int4 rb = D.st_rb[g];
    float md = 0.0f, mp = 1.0f;
    constexpr int U = 4;
    for (int par = 0; par < 2; par++) {
        uint32_t o = par == 0 ? rb.x : rb.z, n = par == 0 ? rb.y : rb.w;
        if constexpr (LDSP) {
            half_sweep_pk(D, o, n, om, phi, ktab, md, mp);
            if (par == 0) __syncthreads();
            continue;
        }
        for (uint32_t t = threadIdx.x; t < n; t += ST_NT * U) {
            float2 sr[U];
            int4 na[U];
            int2 nb[U];
            float4 ka[U];
            float2 kb[U];
            uint32_t ii[U];
            // (every row of the batch loads at once: loading the rest only after the diagonal says the brick is in the
            // structure would cost a second round trip)
#pragma unroll
            for (int u = 0; u < U; u++) {
                uint32_t q = t + u * ST_NT;
                const bool in = q < n;
                uint32_t p = o + (in ? q : 0u);
                sr[u] = in ? D.sten_s[p] : make_float2(-1.0f, 0.0f);
                na[u] = D.sten_na[p]; nb[u] = D.sten_nb[p]; ka[u] = D.sten_ka[p]; kb[u] = D.sten_kb[p]; ii[u] = D.st_red[p];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (sr[u].x < 0.0f) continue;
                int nbr[6] = {na[u].x, na[u].y, na[u].z, na[u].w, nb[u].x, nb[u].y};
                float kk[6] = {ka[u].x, ka[u].y, ka[u].z, ka[u].w, kb[u].x, kb[u].y};
                float skp = 0.0f;
#pragma unroll
                for (int d = 0; d < 6; d++) if (nbr[d] >= 0) skp += kk[d] * phi.ld(nbr[d]);
                int i = (int)ii[u];
                float pi = phi.ld(i), po = pi;
                if (sr[u].x > 0.0f) {
                    float target = (sr[u].y + skp) / sr[u].x;
                    float d = om * (target - pi);
                    po = pi + d;
                    phi.st(i, po);
                    md = fmaxf(md, fabsf(d));
                } else {
                    md = fmaxf(md, 0.0f);
                }
                mp = fmaxf(mp, po);
            }
        }
        if (par == 0) __syncthreads();
    }
    block_max2_tail<ST_NT>(md, mp, sh);
    return md / mp;
}
// one structure's sweeps of a frame: adaptive (mode 0, structural :1268-1281) or warm-up (mode 1, warm_step :1068)
template <bool LDSP>
__device__ void sor_struct_frame(LD& D, uint32_t g, int mode, const Phi<LDSP>& phi, long long& budget, float* sh,
                                 const float* ktab) {
    int nmin = D.st_nmin[g];
    long long size = D.st_hi[g] - D.st_lo[g];
    float resid = D.st_resid[g];
    if (LDSP) {
        for (uint32_t k = threadIdx.x; k < phi.n; k += ST_NT) phi.p[k] = phi.g[phi.lo + (int)k];
        __syncthreads();
    }
    if (mode == 0) {
        int k = 0;
        for (;;) {
            resid = sweep_struct<LDSP>(D, g, 1.85f, phi, sh, ktab);
            k++;
            budget -= size;
            if (k >= nmin && (resid < 4e-4f || k >= 90 || budget <= 0)) break;
        }
    } else {
        int warm = nmin;
        while (warm > 0 && budget > 0) {
            resid = sweep_struct<LDSP>(D, g, warm > 60 ? 1.9f : 1.0f, phi, sh, ktab);
            warm--;
            budget -= size;
        }
        if (threadIdx.x == 0) D.st_nmin[g] = warm;
    }
    if (LDSP) for (uint32_t k = threadIdx.x; k < phi.n; k += ST_NT) phi.g[phi.lo + (int)k] = phi.p[k];
    if (threadIdx.x == 0) D.st_resid[g] = resid;
    __syncthreads();
}
// one workgroup per scene: its structures in order, within the frame's budget of brick updates; lds_cap floats of
// dynamic LDS hold the potential of a structure of up to that many bricks
__global__ void __launch_bounds__(ST_NT) k_sor_scene(LD D, int mode, uint32_t lds_cap) {
    __shared__ float sh[2 * ST_NT / 64];
    __shared__ float ktab[9 * 9 * 2];  // cond_w of two static bricks by (mat_i, mat_j, vertical)
    extern __shared__ float lph[];
    {
        const Tables& T = tables();
        for (int t = threadIdx.x; t < 9 * 9 * 2; t += ST_NT) {
            const int mi = t / 18, mj = (t / 2) % 9;
            float k = minf(T.stiff[mi], T.stiff[mj]);
            ktab[t] = k * ((t & 1) ? 1.0f : 0.3f);
        }
        __syncthreads();
    }
    for (uint32_t sc = blockIdx.x; sc < D.nscene; sc += gridDim.x) {
        const SceneK S = D.sc[sc];
        long long budget = mode == 0 ? 160000 : 120000;
        for (int sl = 0; sl < S.nst; sl++) {
            uint32_t g = (uint32_t)(S.st0 + sl);
            if (D.st_nmin[g] <= 0) continue;
            if (mode == 0 && !D.st_act[g]) continue;
            int lo = D.st_lo[g];
            uint32_t size = (uint32_t)(D.st_hi[g] - lo);
            if (size <= lds_cap) sor_struct_frame<true>(D, g, mode, Phi<true>{lph, lo, size, D.phi}, budget, sh, ktab);
            else sor_struct_frame<false>(D, g, mode, Phi<false>{lph, lo, 0u, D.phi}, budget, sh, ktab);
        }
    }
}

// structures: clear the new resting loads, then mix them in
// the structural pass's clears in one launch: per structure maximum ratio and flags, per scene maximum, resting-list
// counts, new resting loads of active structures, and the event counters
__global__ void __launch_bounds__(NT) k_struct_clear(LD D, uint32_t ns, uint32_t nsc) {
    const uint32_t n = max(max(D.nb, ns), nsc);
    GS(i, n) {
        if (i < ns) { D.st_maxr[i] = 0u; D.st_flags[i] = 0u; }
        if (i < nsc) D.sc_gmax[i] = 0u;
        if (i < D.nb) {
            D.lcnt[i] = 0u; D.lfill[i] = 0u;
            int sc = D.bxyz[i].w, s = D.sid[i];
            const SceneK& S = D.sc[sc];
            if (s < S.nst && D.st_act[S.st0 + s]) D.extn[i] = 0.0f;
        }
        if (i == 0) { D.ctr[C_NEV] = 0u; D.ctr[C_NBW] = 0u; }
    }
}
// resting loads: a body's weight shared between the bricks that rest on a structure brick or on the ground
__global__ void __launch_bounds__(NT) k_rest_count(LD D) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t ub = wave; ub < D.nub; ub += nwaves) {
        const uint32_t b = D.ubody[ub];
        if (!D.alive[b]) continue;
        const SceneK S = D.sc[D.bd_scene[b]];
        uint32_t o = D.boff[b], n = D.bcnt[b];
        uint32_t ns = 0, ng = 0;
        for (uint32_t t = lane; t < n; t += 64) {
            uint32_t i = D.dyn[o + t];
            LV p = lv4(D.wpos[i]);
            if (p.y < RB + 0.05f) ng++;
            float yb = p.y - RB - 0.02f;
            int j = static_brick(D, S, fcell(p.x) + NX2, fcell(yb), fcell(p.z) + NX2);
            bool hit = j >= 0 && D.sid[j] < S.nst && D.st_on[S.st0 + D.sid[j]];
            D.ltarget[o + t] = hit ? j : -1;
            if (hit) ns++;
        }
        for (int h = 32; h >= 1; h >>= 1) { ns += __shfl_xor(ns, h, 64); ng += __shfl_xor(ng, h, 64); }
        if (lane == 0) D.bshare[b] = ns == 0 ? -1.0f : D.m[b] / (float)(ns + ng);
    }
}
__global__ void __launch_bounds__(NT) k_rest_emit(LD D, int fill) {
    GS(k, D.ndyn) {
        int b = D.dyn_body[k];
        if (!D.alive[b] || D.bshare[b] < 0.0f) continue;
        int j = D.ltarget[k];
        if (j < 0) continue;
        if (!fill) atomicAdd(&D.lcnt[j], 1u);
        else D.lent[D.loff[j] + atomicAdd(&D.lfill[j], 1u)] = k;
    }
}
// resting loads summed per brick from its list, then mixed into the brick's external load (structure bricks only)
__global__ void __launch_bounds__(NT) k_rest_sum_mix(LD D) {
    GS(j, D.nb) {
        uint32_t n = D.lcnt[j];
        float en = D.extn[j];
        if (n) {
            float s = en;
            for (uint32_t e = 0; e < n; e++) s += D.bshare[D.dyn_body[D.lent[D.loff[j] + e]]];
            D.extn[j] = s;
            en = s;
        }
        int sc = D.bxyz[j].w, sl = D.sid[j];
        const SceneK& S = D.sc[sc];
        if (sl < S.nst && D.st_act[S.st0 + sl]) D.ext[j] = D.ext[j] * 0.8f + en * 0.2f;
    }
}
// pieces of 8+ bricks with no support stamped for 10 frames start their load field over
__global__ void __launch_bounds__(NT) k_unsupported(LD D) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t ub = wave; ub < D.nub; ub += nwaves) {
        const uint32_t b = D.ubody[ub];
        if (!D.alive[b] || D.count[b] < 8u) continue;
        uint32_t frame = D.sc[D.bd_scene[b]].frame;
        uint32_t o = D.boff[b], n = D.bcnt[b];
        bool any = false;
        for (uint32_t t = lane; t < n; t += 64) if (frame - D.sup[D.dyn[o + t]] < 10u) any = true;
        if (__ballot(any)) continue;
        for (uint32_t t = lane; t < n; t += 64) D.phi[D.dyn[o + t]] = 0.0f;
    }
}
__device__ __forceinline__ bool body_loaded(const LD& D, uint32_t b) {
    return D.alive[b] && D.count[b] >= 8u && !D.asleep[b] && D.nc[b] > 0u && D.bcnt[b] > 0u && D.mat[D.dyn[D.boff[b]]] < SLUG;
}
// g - a loading of loaded pieces (src/lib.rs:1284-1308); velocities remembered for every live body
__global__ void __launch_bounds__(NT) k_ga(LD D, float dt) {
    const Tables& T = tables();
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    float idt = 1.0f / maxf(dt, 1e-3f);
    for (uint32_t ub = wave; ub < D.nub; ub += nwaves) {
        const uint32_t b = D.ubody[ub];
        if (!D.alive[b]) continue;
        LV v = lv4(D.v[b]), w = lv4(D.w[b]);
        LV al = mul(sub(v, lv4(D.vlast[b])), idt);
        LV aa = mul(sub(w, lv4(D.wlast[b])), idt);
        bool loaded = body_loaded(D, b);
        if (lane == 0) { D.vlast[b] = f4(v); D.wlast[b] = f4(w); }
        if (!loaded) continue;
        LV pb = lv4(D.pos[b]);
        uint32_t o = D.boff[b], n = D.bcnt[b];
        float nm = 0.0f;
        for (uint32_t t = lane; t < n; t += 64) {
            uint32_t i = D.dyn[o + t];
            if (D.mat[i] == 0) continue;
            LV a = add(al, cross(aa, sub(lv4(D.wpos[i]), pb)));
            LV ge = lv(-a.x, -GR - a.y, -a.z);
            float nn = minf(len(ge) / GR, 150.0f);
            if (nn > nm) nm = nn;
            D.ext[i] = T.mass[D.mat[i]] * (nn - 1.0f);
        }
        for (int h = 32; h >= 1; h >>= 1) nm = fmaxf(nm, __shfl_xor(nm, h, 64));
        if (lane == 0) D.nmax[b] = nm;
    }
}
// which pieces sweep their load field this frame
__global__ void __launch_bounds__(NT) k_body_sor_prep(LD D) {
    GS(ub, D.nub) {
        const uint32_t b = D.ubody[ub];
        uint8_t go = 0;
        if (body_loaded(D, b) && D.count[b] <= 60000u) {
            const SceneK& S = D.sc[D.bd_scene[b]];
            uint32_t lb = b - (uint32_t)S.bd0;
            go = !(D.resid[b] < 2e-4f && D.nmax[b] < 1.5f && ((S.frame + lb) % 4u) != 0u);
        }
        D.bd_sor[b] = go;
    }
}
constexpr int BD_NT = 256;
__global__ void __launch_bounds__(BD_NT) k_sor_body(LD D) {
    __shared__ float sh[2 * BD_NT / 64];
    for (uint32_t ub = blockIdx.x; ub < D.nub; ub += gridDim.x) {
        const uint32_t b = D.ubody[ub];
        if (!D.bd_sor[b]) continue;
        uint32_t frame = D.sc[D.bd_scene[b]].frame;
        uint32_t o = D.boff[b], n = D.bcnt[b];
        int k = 0;
        for (;;) {
            float md = 0.0f, mp = 1.0f;
            for (int par = 0; par < 2; par++) {
                for (uint32_t t = threadIdx.x; t < n; t += BD_NT) {
                    uint32_t i = D.dyn[o + t];
                    if (D.mat[i] == 0) continue;
                    int4 c = D.bxyz[i];
                    if (((c.x + c.y + c.z) & 1) != par) continue;
                    float d = sor_brick(D, i, 1.8f, frame);
                    md = fmaxf(md, d);
                    mp = fmaxf(mp, D.phi[i]);
                }
                __syncthreads();
            }
            block_max2<BD_NT>(md, mp, sh);
            float r = md / mp;
            k++;
            if (threadIdx.x == 0) D.resid[b] = r;
            if (k >= 2 && (r < 4e-4f || k >= 60)) break;
        }
        __syncthreads();
    }
}

// joints: clear stress where it is judged, decide which pieces are judged
__global__ void __launch_bounds__(NT) k_stress_clear(LD D) {
    GS(i, D.nb) {
        int sc = D.bxyz[i].w;
        const SceneK& S = D.sc[sc];
        int s = D.sid[i];
        if (D.body[i] < 0) {
            if (s < S.nst && D.st_act[S.st0 + s] && in_struct(D, i, s)) D.stress[i] = 0.0f;
        } else {
            int b = D.body[i];
            if (D.mat[i] != 0 && body_loaded(D, (uint32_t)b) && D.count[b] <= 60000u) D.stress[i] = 0.0f;
        }
    }
}
__global__ void __launch_bounds__(NT) k_body_judge_prep(LD D, float dt) {
    GS(ub, D.nub) {
        const uint32_t b = D.ubody[ub];
        uint8_t judge = 0;
        if (body_loaded(D, b) && D.count[b] <= 60000u) {
            const SceneK& S = D.sc[D.bd_scene[b]];
            uint32_t lb = b - (uint32_t)S.bd0;
            D.lt[b] += dt;
            bool settled = D.resid[b] < 2e-4f && D.nmax[b] < 1.5f;
            if (!(settled && ((S.frame + lb) % 4u) != 0u)) {
                judge = 1;
                D.bd_dt[b] = settled ? dt * 4.0f : dt;
                D.bd_live[b] = (D.lt[b] > 0.3f && (D.resid[b] < 1e-2f || D.lt[b] > 0.6f)) || D.nmax[b] > 2.5f;
                D.maxr[b] = 0.0f;
            }
        }
        D.bd_judge[b] = judge;
    }
}
// stress of brick i's three positive-axis bonds and its footing (src/lib.rs:1153-1227), judged on the field as it
// stands; a bearing joint failing in compression is reported for the host to crush
__device__ float stress_eval(LD& D, uint32_t i, float dt, bool live, unsigned long long key, int gst) {
    const Tables& T = tables();
    int mi = D.mat[i];
    int bi = D.body[i];
    bool dynb = bi >= 0;
    float maxr = 0.0f;
    float4 bd = D.bond[i];
    float4 bm = D.bmax[i];
    for (int a = 0; a < 3; a++) {
        float hp = 0.0f;
        int j = linked(D, i, a * 2, &hp);
        if (j < 0) continue;
        float f = cond_w(D, i, (uint32_t)j, a) * (D.phi[i] - D.phi[j]);
        float u = axis_up(D, i, a);
        float up = f > 0.0f ? maxf(u, 0.0f) : maxf(-u, 0.0f);
        float tens = minf(T.tens[mi], T.tens[D.mat[j]]);
        float capm = 1.0f - (1.0f - tens) * up;
        float hpw = hp;
        if (dynb) {
            float cv = minf(T.capv[mi], T.capv[D.mat[j]]), ch = minf(T.caph[mi], T.caph[D.mat[j]]);
            float lat = a == 1 ? cv : ch;
            float au = fabsf(u);
            hpw = hp * (cv * au + ch * (1.0f - au)) / maxf(lat, 1.0f);
        }
        float r = fabsf(f) / (hpw * capm);
        atomic_max_pos((uint32_t*)&D.stress[i], r);
        atomic_max_pos((uint32_t*)&D.stress[j], r);
        if (r > maxr) maxr = r;
        if (live && r > 1.0f) {
            float bmaxa = a == 0 ? bm.x : (a == 1 ? bm.y : bm.z);
            float nhp = dynb ? hp - hp * dt * T.duct[mi] * minf(r - 0.85f, 40.0f) : hp - hp * dt * 2.2f * minf(r - 0.85f, 2.5f);
            if (nhp < bmaxa * 0.08f) {
                if (up == 0.0f && fabsf(u) > 0.5f) {
                    int k = T.capv[D.mat[j]] < T.capv[mi] ? j : (int)i;
                    uint32_t e = atomicAdd(&D.ctr[C_NEV], 1u);
                    if (e < D.evcap) { D.evkey[e] = key; D.ev[e] = make_int4((int)i, a, k, gst); }
                    return maxr > r ? maxr : r;
                }
                nhp = 0.0f;
                atomicAdd(&D.sc_fail[4 * D.bxyz[i].w + (up > 0.0f ? 1 : 2)], 1u);
                if (dynb) D.bd_dirty[bi] = 1;
                else atomicOr(&D.st_flags[gst], 3u);
            }
            if (a == 0) bd.x = nhp; else if (a == 1) bd.y = nhp; else bd.z = nhp;
            D.bond[i] = bd;
            uint32_t e = atomicAdd(&D.ctr[C_NBW], 1u);
            if (e < D.bwcap) { D.bwi[e] = 4u * i + (uint32_t)a; D.bwv[e] = nhp; }
        }
    }
    if (!dynb && D.bxyz[i].y == 0 && bd.w > 0.0f) {
        float f = GROUNDK * T.stiff[mi] * D.phi[i];
        float r = f / bd.w;
        atomic_max_pos((uint32_t*)&D.stress[i], r);
        if (r > maxr) maxr = r;
        if (live && r > 1.0f) {
            float g = bd.w - bd.w * dt * 2.2f * (r - 0.85f);
            if (g < bm.w * 0.08f) { g = 0.0f; atomicOr(&D.st_flags[gst], 3u); atomicAdd(&D.sc_fail[4 * D.bxyz[i].w + 3], 1u); }
            bd.w = g;
            D.bond[i] = bd;
            uint32_t e = atomicAdd(&D.ctr[C_NBW], 1u);
            if (e < D.bwcap) { D.bwi[e] = 4u * i + 3u; D.bwv[e] = g; }
        }
    }
    return maxr;
}
// max of v over the lanes sharing each distinct key (key < 0: none), one atomic per key per wave
__device__ __forceinline__ void wave_keyed_max(int key, float v, uint32_t* dst, uint32_t* dst2, const int32_t* key2of) {
    const int lane = threadIdx.x & 63;
    unsigned long long active = __ballot(key >= 0);
    while (active) {
        int leader = __ffsll((long long)active) - 1;
        int k = __shfl(key, leader, 64);
        bool mine = key == k;
        float x = mine ? v : 0.0f;
        for (int h = 32; h >= 1; h >>= 1) x = fmaxf(x, __shfl_xor(x, h, 64));
        if (lane == leader) {
            atomic_max_pos(&dst[k], x);
            if (dst2) atomic_max_pos(&dst2[key2of ? key2of[k] : 0], x);
        }
        active &= ~__ballot(mine);
    }
}
__global__ void __launch_bounds__(NT) k_stress_static(LD D, float dt) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t base = wave * 64; base < D.nb; base += nwaves * 64) {
        uint32_t i = base + lane;
        int gst = -1;
        float r = 0.0f;
        if (i < D.nb && D.body[i] < 0) {
            int sc = D.bxyz[i].w;
            const SceneK& S = D.sc[sc];
            int s = D.sid[i];
            if (s < S.nst && D.st_act[S.st0 + s] && in_struct(D, i, s)) {
                gst = S.st0 + s;
                bool live = D.st_live[gst] != 0 && D.st_resid[gst] < 1.5e-3f;
                unsigned long long key = ((unsigned long long)(uint32_t)gst << 40) | (unsigned long long)i;
                r = stress_eval(D, i, dt, live, key, gst);
            }
        }
        wave_keyed_max(gst, r, D.st_maxr, D.sc_gmax, D.st_scene);
    }
}
__global__ void __launch_bounds__(NT) k_stress_body(LD D) {
    const int lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * NT + threadIdx.x) >> 6, nwaves = (gridDim.x * NT) >> 6;
    for (uint32_t base = wave * 64; base < D.ndyn; base += nwaves * 64) {
        uint32_t k = base + lane;
        int b = -1;
        float r = 0.0f;
        if (k < D.ndyn) {
            int bb = D.dyn_body[k];
            uint32_t i = D.dyn[k];
            if (D.bd_judge[bb] && D.mat[i] != 0) {
                b = bb;
                unsigned long long key = (1ull << 63) | ((unsigned long long)(uint32_t)b << 40) | (unsigned long long)k;
                r = stress_eval(D, i, D.bd_dt[b], D.bd_live[b] != 0, key, -1);
            }
        }
        wave_keyed_max(b, r, (uint32_t*)D.maxr, D.sc_gmax, D.bd_scene);
    }
}
__global__ void __launch_bounds__(NT) k_lt_decay(LD D, float dt) {
    GS(ub, D.nub) {
        const uint32_t b = D.ubody[ub];
        if (D.alive[b] && !body_loaded(D, b)) D.lt[b] = maxf(D.lt[b] - 2.0f * dt, 0.0f);
    }
}

// per-scene counts: bodies, bricks in bodies, awake bodies
__global__ void __launch_bounds__(NT) k_scene_stats(LD D, uint32_t* out) {
    GS(ub, D.nub) {
        const uint32_t b = D.ubody[ub];
        if (!D.alive[b]) continue;
        int s = D.bd_scene[b];
        atomicAdd(&out[3 * s], 1u);
        atomicAdd(&out[3 * s + 1], D.count[b]);
        if (!D.asleep[b]) atomicAdd(&out[3 * s + 2], 1u);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) { GS(i, n) p[i] = v; }
// a substep's start: the hash counts and fill cursors, and every counter but the killed-body count
__global__ void __launch_bounds__(NT) k_hash_clear(LD D) {
    GS(i, D.tmask + 1u) { D.hcnt[i] = 0u; D.hfill[i] = 0u; }
    if (blockIdx.x == 0 && threadIdx.x < (uint32_t)C_NCTR && threadIdx.x != (uint32_t)C_NKILL) D.ctr[threadIdx.x] = 0u;
}
// the dynamic brick lists from one uploaded block: k brick indices, k owning bodies, then (at o_csr) nc records of
// (body, first, count) for the live bodies' brick ranges
__global__ void __launch_bounds__(NT) k_csr_set(LD D, const uint8_t* blk, uint32_t k, uint32_t nc, uint32_t o_csr) {
    GS(t, max(k, nc)) {
        if (t < k) { D.dyn[t] = reinterpret_cast<const uint32_t*>(blk)[t]; D.dyn_body[t] = reinterpret_cast<const int32_t*>(blk)[k + t]; }
        if (t < nc) { int4 x = reinterpret_cast<const int4*>(blk + o_csr)[t]; D.boff[x.x] = (uint32_t)x.y; D.bcnt[x.x] = (uint32_t)x.z; }
    }
}
__global__ void __launch_bounds__(NT) k_nc_clear(LD D) { GS(ub, D.nub) D.nc[D.ubody[ub]] = 0u; }
// The structural pass's results in one block for one copy to the host: the counters, then per structure residual,
// maximum ratio and flags, per scene maximum ratio, per used body slot its dirty flag (taken: cleared here), and the
// first kev crush events and kbw bond writes (the host copies the rest when there are more).
struct PackRB { uint32_t ns, nsc, o_rs, o_mr, o_fl, o_gm, o_bd, o_ek, o_ev, o_bi, o_bv, kev, kbw; };
__global__ void __launch_bounds__(NT) k_pack_rb(LD D, uint8_t* out, PackRB P) {
    const uint32_t ns = P.ns, nsc = P.nsc, o_rs = P.o_rs, o_mr = P.o_mr, o_fl = P.o_fl, o_gm = P.o_gm, o_bd = P.o_bd;
    const uint32_t nev = min(min(D.ctr[C_NEV], D.evcap), P.kev), nbw = min(min(D.ctr[C_NBW], D.bwcap), P.kbw);
    const uint32_t n = max(max(max(D.nub, ns), max(nsc, (uint32_t)C_NCTR)), max(P.kev, P.kbw));
    GS(i, n) {
        if (i < nev) {
            reinterpret_cast<unsigned long long*>(out + P.o_ek)[i] = D.evkey[i];
            reinterpret_cast<int4*>(out + P.o_ev)[i] = D.ev[i];
        }
        if (i < nbw) {
            reinterpret_cast<uint32_t*>(out + P.o_bi)[i] = D.bwi[i];
            reinterpret_cast<float*>(out + P.o_bv)[i] = D.bwv[i];
        }
        if (i < (uint32_t)C_NCTR) reinterpret_cast<uint32_t*>(out)[i] = D.ctr[i];
        if (i < ns) {
            reinterpret_cast<float*>(out + o_rs)[i] = D.st_resid[i];
            reinterpret_cast<uint32_t*>(out + o_mr)[i] = D.st_maxr[i];
            reinterpret_cast<uint32_t*>(out + o_fl)[i] = D.st_flags[i];
        }
        if (i < nsc) reinterpret_cast<uint32_t*>(out + o_gm)[i] = D.sc_gmax[i];
        if (i < D.nub) { uint32_t b = D.ubody[i]; out[o_bd + i] = D.bd_dirty[b]; D.bd_dirty[b] = 0; }
    }
}
// the structures' sweep controls from one uploaded block: sweeps asked for (int32), then active, on, live (bytes)
__global__ void __launch_bounds__(NT) k_st_ctl(LD D, const uint8_t* blk, uint32_t ns) {
    GS(g, ns) {
        D.st_nmin[g] = reinterpret_cast<const int32_t*>(blk)[g];
        D.st_act[g] = blk[4 * (size_t)ns + g];
        D.st_on[g] = blk[5 * (size_t)ns + g];
        D.st_live[g] = blk[6 * (size_t)ns + g];
    }
}
__global__ void k_fill_u8(uint8_t* p, uint8_t v, size_t n) { GS(i, n) p[i] = v; }

}  // namespace

// ====================================================================== host
namespace {

// Host worker pool for per-scene bookkeeping: scenes share no host state but the per-brick and per-body arrays, which
// are partitioned by scene, so their topology work runs on separate threads with identical results.
class Pool {
  public:
    explicit Pool(int workers) {
        for (int k = 0; k < workers; k++) th_.emplace_back([this, k] { loop(k); });
    }
    ~Pool() {
        { std::lock_guard<std::mutex> g(m_); stop_ = true; }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    int workers() const { return (int)th_.size(); }
    // f(i, slot) for i in [0, n): slot 0 is the calling thread, 1.. the workers; returns when all have finished
    void run(int n, const std::function<void(int, int)>& f) {
        if (n <= 0) return;
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = &f; n_ = n; next_.store(0); busy_ = (int)th_.size(); gen_++;
        }
        cv_.notify_all();
        work(0);
        std::unique_lock<std::mutex> g(m_);
        done_.wait(g, [this] { return busy_ == 0; });
        job_ = nullptr;
    }
  private:
    void work(int slot) {
        for (;;) {
            int i = next_.fetch_add(1);
            if (i >= n_) break;
            (*job_)(i, slot);
        }
    }
    void loop(int k) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            work(k + 1);
            std::lock_guard<std::mutex> g(m_);
            if (--busy_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int, int)>* job_ = nullptr;
    std::atomic<int> next_{0};
    int n_ = 0, busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// per-thread scratch of the topology searches and the thread's lists of touched bricks and bodies
struct Scratch {
    std::vector<uint32_t> q, lst, lst2, cnt, tlist, btlist;
    void reserve(size_t bricks) {
        if (q.size() >= bricks + 2) return;
        q.assign(bricks + 2, 0); lst.assign(bricks + 2, 0); lst2.assign(bricks + 2, 0); cnt.assign(bricks + 3, 0);
    }
};
thread_local Scratch* t_scr = nullptr;  // set while a pool worker (or the caller, as slot 0) runs a scene's work

// A body to set up on the device (setup_body, src/lib.rs:844-929): its slot, its bricks [s, e) of the job brick list
// in the reference's order, and the parent's pose and motion it inherits. flags: 1 a part detached from a structure
// (its bricks' load field cleared), 2 the slot keeps its cooldown and life timer (the first part of a split body).
struct SetupJob {
    int32_t b, s, e, flags;
    float4 ppos, pq, pv, pw;
    float4 prot[3];
    float keep_cool, keep_lt, pad0, pad1;
};
// what comes back of a set-up body: position, velocity, inverse inertia (body and world frames), mass, radius
struct SetupOut { float4 pos, v; float4 iil[3], iiw[3]; float m, im, rad, pad; };

struct HScene {
    int b0 = 0, bcap = 0, nb = 0;          // bricks [b0, b0 + nb), capacity bcap
    int bd0 = 0, dcap = 0, hw = 0, cursor = 0;
    int st0 = 0, nst = 0;
    uint32_t rng = 0x9E3779B9u, frame = 0, stamp_id = 0;
    float time = 0.0f, blast_t = 0.0f, blast_r = 0.0f, blast_p = 0.0f;
    LV blast_c{0.0f, 0.0f, 0.0f};
    int crush_n = 0;
    bool conn_any = false, finalized = false, active = true;
    int gx0 = 0, gy0 = 0, gz0 = 0, gnx = 1, gny = 1, gnz = 1;
    size_t goff = 0, gcap = 0;             // grid region of the scene and its size
    float stats[16] = {};
    uint32_t crushed_joints = 0;           // compression failures (applied on the host)
    uint32_t dyn0 = 0, ndyn = 0;
    uint32_t c0 = 0, ncs = 0;              // contact range of the last substep
    bool host_dirty = false;               // host mirror changed since the last push
    // bodies to set up on the device (decided by the host in the reference's order), their bricks, and the bodies a
    // connectivity pass detached (blasted and their neighbours woken once they are set up)
    std::vector<SetupJob> sjobs;
    std::vector<uint32_t> sbricks;
    std::vector<float4> sin;
    std::vector<int> conn_new;
    // bodies to rebuild: their live bricks in list order (rlist) and per body its slot, first entry and count
    std::vector<uint32_t> rlist;
    std::vector<int3> rbody;
};
struct HStruct {
    int scene = 0, lo = 0, hi = 0, mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0}, top0 = 0, top = 0, warm = 0;
    uint32_t init = 0, now = 0;
    float active = 0.0f, settle = 0.0f, maxr = 0.0f, resid = 0.0f;
    bool solved = false, conn = false, broke = false;
    int4 rb{0, 0, 0, 0};
};


// a device buffer zero-filled on the world's stream. Growing it keeps the old block until release: freeing memory
// waits for the whole device, including other streams' work (a load-path solve running beside the frame).
template <class T> struct DBuf {
    T* p = nullptr; size_t n = 0;
    std::vector<void*> old;
    int alloc(size_t cnt, hipStream_t st) {
        if (p) old.push_back((void*)p);
        p = nullptr; n = cnt;
        if (!cnt) cnt = 1;
        if (hipMalloc(&p, sizeof(T) * cnt) != hipSuccess) { p = nullptr; n = 0; sf::set_error(SF_E_HIP, "hipMalloc failed"); return -1; }
        return hipMemsetAsync(p, 0, sizeof(T) * cnt, st) == hipSuccess ? 0 : -1;
    }
    void release() {
        if (p) (void)hipFree(p);
        for (void* q : old) (void)hipFree(q);
        old.clear(); p = nullptr; n = 0;
    }
};

struct BrickDelta { uint32_t i; int32_t body; uint32_t mat; uint32_t full; float4 bond, local, wpos; float phi, ext, extn, pad2; };
// a body's full state, packed for one copy each way
struct BodyRec {
    uint32_t b; uint8_t alive, asleep, wake, pad0; uint32_t count, nc;
    float4 pos, q, v, w, vlast, wlast;
    float4 rot[3], iil[3], iiw[3];
    float m, im, sleep, rad, cool, lt, resid, maxr, nmax, pad1, pad2, pad3;
};
// pinned host staging: copies to and from it are asynchronous DMA
template <class T> struct HBuf {
    T* p = nullptr; size_t n = 0;
    int reserve(size_t cnt) {
        if (cnt <= n) return 0;
        release();
        size_t c = std::max<size_t>(cnt, 4096);
        if (hipHostMalloc((void**)&p, sizeof(T) * c, hipHostMallocDefault) != hipSuccess) { p = nullptr; return -1; }
        n = c;
        return 0;
    }
    void release() { if (p) (void)hipHostFree(p); p = nullptr; n = 0; }
};

// A host-to-device block staged through two pinned buffers: begin() hands out the buffer not in flight (waiting for
// the copy that last used it), end() queues the copy into the device block and marks the buffer in flight.
struct Stage {
    HBuf<uint8_t> h[2];
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool rec[2] = {false, false};
    int cur = 0;
    DBuf<uint8_t> d;
    uint8_t* begin(size_t bytes) {
        cur ^= 1;
        if (!ev[cur] && hipEventCreateWithFlags(&ev[cur], hipEventDisableTiming) != hipSuccess) { sf::set_error(SF_E_HIP, "event"); return nullptr; }
        if (rec[cur] && hipEventSynchronize(ev[cur]) != hipSuccess) { sf::set_error(SF_E_HIP, "event wait"); return nullptr; }
        if (h[cur].reserve(bytes)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return nullptr; }
        return h[cur].p;
    }
    int end(size_t bytes, hipStream_t st) {
        if (d.n < bytes && d.alloc(std::max<size_t>(bytes + bytes / 2, 4096), st)) return -1;
        if (hipMemcpyAsync(d.p, h[cur].p, bytes, hipMemcpyHostToDevice, st) != hipSuccess) { sf::set_error(SF_E_HIP, "copy"); return -1; }
        if (hipEventRecord(ev[cur], st) != hipSuccess) { sf::set_error(SF_E_HIP, "event"); return -1; }
        rec[cur] = true;
        return 0;
    }
    void release() {
        for (int k = 0; k < 2; k++) { h[k].release(); if (ev[k]) (void)hipEventDestroy(ev[k]); ev[k] = nullptr; rec[k] = false; }
        d.release();
    }
};

struct Lat {
    int dev = 0;
    hipStream_t st = nullptr;
    int K = 16, R = 8;
    int fuse_max = 32768;
    int lds_max = 0;           // LDS bytes a workgroup may have (device attribute)
    int64_t sor_lds = -1;      // caller's limit on the structures swept with their potential in LDS (-1: none)
    size_t per_brick = 0, per_body = 0, per_struct = 0, per_contact = 0;  // device bytes per unit (dalloc tallies)
    std::vector<std::pair<int, int>> free_b, free_d;  // unplaced brick and body ranges (start, length), by start
    bool lds_attr = false, lds_attr_fs = false;      // scenes of at most this many contacts are solved by one workgroup each (0: never)
    std::vector<int32_t> slist, slist_up;  // scenes with contacts this substep; the list on the device
    Stage stg_ctl, stg_csr, stg_delta, stg_brec;  // uploads: sweep controls, dynamic brick lists, brick and body deltas
    // structure connectivity on the device: on (default) the grounded search runs as union-find kernels (k_cc_*),
    // off the host searches; per structure its place in this pass's list (-1: none), the list's bit offsets, results
    bool conn_device = true;
    uint32_t ctoken = 0;
    Stage stg_conn, stg_setup, stg_crush;
    DBuf<uint8_t> d_setup_out, d_crush;
    HBuf<uint8_t> h_setup_out, h_crush;
    DBuf<uint32_t> d_creach_out;
    HBuf<uint32_t> h_creach_out;
    std::vector<int> reach_idx;
    std::vector<uint32_t> reach_woff;
    size_t reach_n = 0, split_off = 0;   // the read-back's structure statuses, then bits, then split roots from split_off
    std::vector<size_t> sc_split;        // per scene: its first split root
    DBuf<uint8_t> d_rbk;     // the structural results packed for one copy
    uint32_t kev_pack = 64, kbw_pack = 256;  // crush events and bond writes the packed block carries
    int bcap = 0, dcap = 0, scap = 0;
    int nb = 0, nbd = 0;                   // bricks / body slots handed out to scenes
    std::vector<HScene> sc;
    std::vector<HStruct> sts;
    // host mirror: bricks
    std::vector<uint8_t> mat, sid, infill;
    std::vector<int32_t> body;
    std::vector<int4> bxyz;
    std::vector<uint32_t> next, nb6, stamp, q;
    std::vector<float4> bond, bmax, local, wpos;
    std::vector<float> phi, ext, extn;
    std::vector<int32_t> grid;
    // host mirror: bodies
    std::vector<uint8_t> alive, asleep, dirty, wake;
    std::vector<uint32_t> first, count, nc;
    std::vector<float4> pos, qq, v, w, vlast, wlast;
    std::vector<LM> rot, iil, iiw;
    std::vector<float> m, im, sleep, rad, cool, lt, resid, maxr, nmax;
    // scratch
    std::vector<uint32_t> lst, lst2, comp, cnt;
    // dynamic slots
    std::vector<uint32_t> dyn;
    std::vector<int4> csr;
    std::vector<int32_t> dyn_body;
    uint32_t ndyn = 0;
    std::atomic<bool> lists_dirty{true};
    std::atomic<bool> ub_dirty{true};  // a scene's body high-water mark moved: the used-slot list is rebuilt
    // host workers: per-slot scratch (slot 0 is the calling thread's)
    std::unique_ptr<Pool> pool;
    int host_threads = 0;  // worker pool size (0: SF_HOST_THREADS, else the cores, at most 16)
    std::vector<std::unique_ptr<Scratch>> scr;
    std::vector<uint32_t> ubody;
    // device
    LD D{};
    std::vector<void*> owned;
    DBuf<SceneK> d_sc;
    DBuf<int32_t> d_grid;
    DBuf<uint32_t> d_scstat;
    DBuf<int32_t> d_slist;
    // host phases' transfers: brick deltas and body records up, body records and brick positions down
    DBuf<BodyRec> d_brec;
    DBuf<uint32_t> d_idx;
    HBuf<BodyRec> h_brec;
    HBuf<uint32_t> h_idx;
    // pinned read-back: counters, scene first contacts, killed bodies, the structural pass's results, scene
    // statistics (read at the next wait), and the scene table (two buffers, each reused once its copy is done)
    HBuf<uint32_t> h_ctrbuf, h_scstat;
    HBuf<int32_t> h_kl;
    HBuf<uint8_t> h_rb;
    uint32_t* h_ctr = nullptr;
    HBuf<SceneK> h_sk[2];
    hipEvent_t ev_sk[2] = {nullptr, nullptr}, ev_stat = nullptr;
    int sk_cur = 0;
    bool sk_recorded[2] = {false, false}, stat_pending = false;
    size_t stat_n = 0;
    std::vector<SceneK> sk_last;
    uint32_t ccap = 0, tsize = 0;
    // bricks whose material, body or bonds the host changed since the last flush
    std::vector<uint8_t> touched;
    std::vector<uint32_t> tlist;
    // what the host wrote of a brick: 1 material, body, bonds; 2 local and world position (body set-up); 4 load
    // field (detachment zeroes it). Fields not written stay as the device has them.
    void touch(int i, uint8_t what = 1) { if (!touched[i]) (t_scr ? t_scr->tlist : tlist).push_back((uint32_t)i); touched[i] |= what; }
    // bodies whose state the host changed since the last flush
    std::vector<uint8_t> btouched;
    std::vector<uint32_t> btlist;
    void btouch(int b) { if (!btouched[b]) { btouched[b] = 1; (t_scr ? t_scr->btlist : btlist).push_back((uint32_t)b); } }
    // timing
    double ms_frame = 0.0, ms_host = 0.0;
    double ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double sp[16] = {0};       // finer split of the host phases (sf_lat_subphase_ms)
    int syncs = 0;

    ~Lat() {
        for (void* p : owned) if (p) (void)hipFree(p);
        d_sc.release(); d_grid.release(); d_scstat.release(); d_slist.release();
        d_brec.release(); d_idx.release(); d_rbk.release(); stg_ctl.release(); stg_csr.release(); stg_delta.release(); stg_brec.release(); stg_conn.release(); stg_setup.release(); d_setup_out.release(); h_setup_out.release(); stg_crush.release(); d_crush.release(); h_crush.release();
        h_brec.release(); h_idx.release();
        h_ctrbuf.release(); h_scstat.release(); h_kl.release(); h_rb.release(); h_sk[0].release(); h_sk[1].release();
        for (hipEvent_t e : {ev_sk[0], ev_sk[1], ev_stat}) if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }
    // `tally` += the array's bytes per unit of its size (per brick, body, structure or contact): the resident bytes
    // of a scene follow from its capacities (sf_lat_bytes_resident)
    template <class T> int dalloc(T*& p, size_t n, size_t* tally = nullptr, size_t per = 0) {
        if (tally) *tally += sizeof(T) * per;
        p = nullptr;
        if (hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)) != hipSuccess) { sf::set_error(SF_E_HIP, "hipMalloc failed"); return -1; }
        owned.push_back((void*)p);
        return hipMemsetAsync(p, 0, sizeof(T) * (n ? n : 1), st) == hipSuccess ? 0 : -1;
    }
    // a replaced array stays allocated (in `owned`) until the world is destroyed: freeing waits for the whole device
    template <class T> void dfree(T*& p) { p = nullptr; }
};

int lat_alloc(Lat* L) {
    const size_t B = L->bcap, N = L->dcap, S = L->scap, NS = 256 * (size_t)S;
    LD& D = L->D;
    int r = 0;
    r |= L->dalloc(D.mat, B, &L->per_brick, 1); r |= L->dalloc(D.sid, B, &L->per_brick, 1); r |= L->dalloc(D.infill, B, &L->per_brick, 1); r |= L->dalloc(D.body, B, &L->per_brick, 1); r |= L->dalloc(D.bxyz, B, &L->per_brick, 1);
    r |= L->dalloc(D.nb6, 6 * B, &L->per_brick, 6); r |= L->dalloc(D.bond, B, &L->per_brick, 1); r |= L->dalloc(D.bmax, B, &L->per_brick, 1); r |= L->dalloc(D.local, B, &L->per_brick, 1);
    r |= L->dalloc(D.wpos, B, &L->per_brick, 1); r |= L->dalloc(D.phi, B, &L->per_brick, 1); r |= L->dalloc(D.ext, B, &L->per_brick, 1); r |= L->dalloc(D.extn, B, &L->per_brick, 1);
    r |= L->dalloc(D.stress, B, &L->per_brick, 1); r |= L->dalloc(D.sup, B, &L->per_brick, 1);
    r |= L->dalloc(D.creach, B, &L->per_brick, 1); r |= L->dalloc(D.cq, B, &L->per_brick, 1);
    r |= L->dalloc(D.xstamp, B, &L->per_brick, 1); r |= L->dalloc(D.xtok, S); r |= L->dalloc(D.xicnt, S); r |= L->dalloc(D.xpcnt, S);
    r |= L->dalloc(D.bd_scene, N, &L->per_body, 1); r |= L->dalloc(D.alive, N, &L->per_body, 1); r |= L->dalloc(D.asleep, N, &L->per_body, 1); r |= L->dalloc(D.wake, N, &L->per_body, 1);
    r |= L->dalloc(D.pos, N, &L->per_body, 1); r |= L->dalloc(D.q, N, &L->per_body, 1); r |= L->dalloc(D.v, N, &L->per_body, 1); r |= L->dalloc(D.w, N, &L->per_body, 1);
    r |= L->dalloc(D.vpre, N, &L->per_body, 1); r |= L->dalloc(D.wpre, N, &L->per_body, 1); r |= L->dalloc(D.vlast, N, &L->per_body, 1); r |= L->dalloc(D.wlast, N, &L->per_body, 1);
    r |= L->dalloc(D.rot, 3 * N, &L->per_body, 3); r |= L->dalloc(D.iil, 3 * N, &L->per_body, 3); r |= L->dalloc(D.iiw, 3 * N, &L->per_body, 3);
    r |= L->dalloc(D.m, N, &L->per_body, 1); r |= L->dalloc(D.im, N, &L->per_body, 1); r |= L->dalloc(D.sleep, N, &L->per_body, 1); r |= L->dalloc(D.rad, N, &L->per_body, 1);
    r |= L->dalloc(D.cool, N, &L->per_body, 1); r |= L->dalloc(D.lt, N, &L->per_body, 1); r |= L->dalloc(D.resid, N, &L->per_body, 1); r |= L->dalloc(D.maxr, N, &L->per_body, 1);
    r |= L->dalloc(D.nmax, N, &L->per_body, 1); r |= L->dalloc(D.nc, N, &L->per_body, 1); r |= L->dalloc(D.cfirst, N, &L->per_body, 1); r |= L->dalloc(D.cend, N, &L->per_body, 1); r |= L->dalloc(D.count, N, &L->per_body, 1); r |= L->dalloc(D.boff, N, &L->per_body, 1);
    r |= L->dalloc(D.bcnt, N, &L->per_body, 1); r |= L->dalloc(D.ubody, N, &L->per_body, 1);
    r |= L->dalloc(D.dyn, B, &L->per_brick, 1); r |= L->dalloc(D.dyn_body, B, &L->per_brick, 1);
    r |= L->dalloc(D.st_scene, NS, &L->per_struct, 1); r |= L->dalloc(D.st_lo, NS, &L->per_struct, 1); r |= L->dalloc(D.st_hi, NS, &L->per_struct, 1); r |= L->dalloc(D.st_act, NS, &L->per_struct, 1);
    r |= L->dalloc(D.st_on, NS, &L->per_struct, 1); r |= L->dalloc(D.st_live, NS, &L->per_struct, 1); r |= L->dalloc(D.st_nmin, NS, &L->per_struct, 1); r |= L->dalloc(D.st_resid, NS, &L->per_struct, 1);
    r |= L->dalloc(D.st_maxr, NS, &L->per_struct, 1); r |= L->dalloc(D.st_flags, NS, &L->per_struct, 1); r |= L->dalloc(D.st_red, B, &L->per_brick, 1); r |= L->dalloc(D.st_rb, NS, &L->per_struct, 1);
    r |= L->dalloc(D.sten_na, B, &L->per_brick, 1); r |= L->dalloc(D.sten_nb, B, &L->per_brick, 1); r |= L->dalloc(D.sten_ka, B, &L->per_brick, 1); r |= L->dalloc(D.sten_kb, B, &L->per_brick, 1); r |= L->dalloc(D.sten_s, B, &L->per_brick, 1);
    r |= L->dalloc(D.sten_pk, B, &L->per_brick, 1);
    r |= L->dalloc(D.sc_gmax, S); r |= L->dalloc(D.sc_fail, 4 * S);
    r |= L->dalloc(D.dslot, B, &L->per_brick, 1); r |= L->dalloc(D.dkey, B, &L->per_brick, 1); r |= L->dalloc(D.hent, B, &L->per_brick, 1);
    r |= L->dalloc(D.ccnt, B, &L->per_brick, 1); r |= L->dalloc(D.coff, B, &L->per_brick, 1);
    r |= L->dalloc(D.mask, N, &L->per_body, 1); r |= L->dalloc(D.claim, N * (size_t)L->K, &L->per_body, (size_t)L->K);
    r |= L->dalloc(D.colfill, KMAX + 1); r |= L->dalloc(D.coloff, KMAX + 1);
    r |= L->dalloc(D.ovcnt, N, &L->per_body, 1); r |= L->dalloc(D.ovoff, N, &L->per_body, 1); r |= L->dalloc(D.ovfill, N, &L->per_body, 1); r |= L->dalloc(D.ovbody, N, &L->per_body, 1);
    r |= L->dalloc(D.lcnt, B, &L->per_brick, 1); r |= L->dalloc(D.loff, B, &L->per_brick, 1); r |= L->dalloc(D.lfill, B, &L->per_brick, 1); r |= L->dalloc(D.lent, B, &L->per_brick, 1);
    r |= L->dalloc(D.ltarget, B, &L->per_brick, 1); r |= L->dalloc(D.lshare, B, &L->per_brick, 1); r |= L->dalloc(D.bshare, N, &L->per_body, 1);
    r |= L->dalloc(D.bd_sor, N, &L->per_body, 1); r |= L->dalloc(D.bd_judge, N, &L->per_body, 1); r |= L->dalloc(D.bd_live, N, &L->per_body, 1); r |= L->dalloc(D.bd_dt, N, &L->per_body, 1);
    r |= L->dalloc(D.bd_dirty, N, &L->per_body, 1);
    // the counters, then each scene's first contact: one copy brings back both
    r |= L->dalloc(D.ctr, CTR_WORDS + S + 1);
    D.sc_c0 = D.ctr + CTR_WORDS;
    D.pcap = (uint32_t)std::max<size_t>(4096, B);
    r |= L->dalloc(D.klist, N, &L->per_body, 1);
    D.evcap = (uint32_t)std::max<size_t>(4096, B); D.bwcap = (uint32_t)std::max<size_t>(4096, 4 * B);
    r |= L->dalloc(D.evkey, D.evcap, &L->per_brick, 1); r |= L->dalloc(D.ev, D.evcap, &L->per_brick, 1); r |= L->dalloc(D.bwi, D.bwcap, &L->per_brick, 4); r |= L->dalloc(D.bwv, D.bwcap, &L->per_brick, 4);
    r |= L->dalloc(D.scan_tmp, sf::scan_scratch_words((uint32_t)std::max<size_t>(std::max(2 * B + 4096, N), 1u << 16)));
    r |= L->d_sc.alloc(S, L->st); r |= L->d_scstat.alloc(3 * S, L->st); r |= L->d_slist.alloc(S, L->st);
    D.sc = L->d_sc.p;
    D.K = (uint32_t)L->K; D.R = (uint32_t)L->R;
    return r ? -1 : 0;
}

// contact arrays sized to `cap`
int lat_contacts_alloc(Lat* L, uint32_t cap) {
    LD& D = L->D;
    auto fr = [&](auto*& p) { L->dfree(p); };
    fr(D.ca); fr(D.cb); fr(D.cbo); fr(D.cscene); fr(D.ccol); fr(D.cprop); fr(D.cba); fr(D.cbb); fr(D.con);
    fr(D.cn); fr(D.cp); fr(D.cra); fr(D.crb); fr(D.ct1); fr(D.ct2);
    fr(D.cdepth); fr(D.cmn); fr(D.cmt1); fr(D.cmt2); fr(D.cjn); fr(D.cjt1); fr(D.cjt2); fr(D.cbias); fr(D.cvn0);
    fr(D.csmn); fr(D.csmt1); fr(D.csmt2); fr(D.collist); fr(D.ovent); fr(D.odv); fr(D.odw);
    fr(D.rq_c); fr(D.rq_ab); fr(D.rq_0); fr(D.rq_1); fr(D.rq_2); fr(D.rq_3); fr(D.rq_4); fr(D.rq_5); fr(D.iflag); fr(D.pflag);
    int r = 0;
    L->per_contact = 0;
    r |= L->dalloc(D.ca, cap, &L->per_contact, 1); r |= L->dalloc(D.cb, cap, &L->per_contact, 1); r |= L->dalloc(D.cbo, cap, &L->per_contact, 1); r |= L->dalloc(D.cscene, cap, &L->per_contact, 1);
    r |= L->dalloc(D.ccol, cap, &L->per_contact, 1); r |= L->dalloc(D.cprop, cap, &L->per_contact, 1); r |= L->dalloc(D.cba, cap, &L->per_contact, 1); r |= L->dalloc(D.cbb, cap, &L->per_contact, 1);
    r |= L->dalloc(D.con, cap, &L->per_contact, 1);
    r |= L->dalloc(D.cn, cap, &L->per_contact, 1); r |= L->dalloc(D.cp, cap, &L->per_contact, 1); r |= L->dalloc(D.cra, cap, &L->per_contact, 1); r |= L->dalloc(D.crb, cap, &L->per_contact, 1);
    r |= L->dalloc(D.ct1, cap, &L->per_contact, 1); r |= L->dalloc(D.ct2, cap, &L->per_contact, 1);
    r |= L->dalloc(D.cdepth, cap, &L->per_contact, 1); r |= L->dalloc(D.cmn, cap, &L->per_contact, 1); r |= L->dalloc(D.cmt1, cap, &L->per_contact, 1); r |= L->dalloc(D.cmt2, cap, &L->per_contact, 1);
    r |= L->dalloc(D.cjn, cap, &L->per_contact, 1); r |= L->dalloc(D.cjt1, cap, &L->per_contact, 1); r |= L->dalloc(D.cjt2, cap, &L->per_contact, 1); r |= L->dalloc(D.cbias, cap, &L->per_contact, 1);
    r |= L->dalloc(D.cvn0, cap, &L->per_contact, 1); r |= L->dalloc(D.csmn, cap, &L->per_contact, 1); r |= L->dalloc(D.csmt1, cap, &L->per_contact, 1); r |= L->dalloc(D.csmt2, cap, &L->per_contact, 1);
    r |= L->dalloc(D.collist, cap, &L->per_contact, 1); r |= L->dalloc(D.ovent, 2 * (size_t)cap, &L->per_contact, 2);
    r |= L->dalloc(D.odv, 2 * (size_t)cap, &L->per_contact, 2); r |= L->dalloc(D.odw, 2 * (size_t)cap, &L->per_contact, 2);
    r |= L->dalloc(D.rq_c, cap, &L->per_contact, 1); r |= L->dalloc(D.rq_ab, cap, &L->per_contact, 1); r |= L->dalloc(D.rq_0, cap, &L->per_contact, 1); r |= L->dalloc(D.rq_1, cap, &L->per_contact, 1);
    r |= L->dalloc(D.rq_2, cap, &L->per_contact, 1); r |= L->dalloc(D.rq_3, cap, &L->per_contact, 1); r |= L->dalloc(D.rq_4, cap, &L->per_contact, 1); r |= L->dalloc(D.rq_5, cap, &L->per_contact, 1);
    r |= L->dalloc(D.iflag, cap, &L->per_contact, 1); r |= L->dalloc(D.pflag, cap, &L->per_contact, 1);
    D.ccap = cap;
    L->ccap = cap;
    return r ? -1 : 0;
}
int lat_hash_alloc(Lat* L, uint32_t ndyn) {
    uint32_t t = 1024;
    while (t < 2 * ndyn) t <<= 1;
    if (t <= L->tsize) return 0;
    LD& D = L->D;
    L->dfree(D.hcnt); L->dfree(D.hoff); L->dfree(D.hfill);
    int r = 0;
    r |= L->dalloc(D.hcnt, t); r |= L->dalloc(D.hoff, t); r |= L->dalloc(D.hfill, t);
    L->tsize = t;
    D.tmask = t - 1;
    return r ? -1 : 0;
}

// ---------------------------------------------------------------- host mirror helpers (the reference's bookkeeping)
// the scratch of the running thread: a pool worker's own, else the world's
struct ScrRef { std::vector<uint32_t> &q, &lst, &lst2, &cnt; };
inline ScrRef scr(Lat* L) { if (t_scr) return {t_scr->q, t_scr->lst, t_scr->lst2, t_scr->cnt}; return {L->q, L->lst, L->lst2, L->cnt}; }
inline LV hv(float4 a) { return lv(a.x, a.y, a.z); }
inline float4 hf(LV a) { return make_float4(a.x, a.y, a.z, 0.0f); }
inline float& bond_ref(Lat* L, int i, int a) { float4& b = L->bond[i]; return a == 0 ? b.x : (a == 1 ? b.y : (a == 2 ? b.z : b.w)); }
inline float bmax_of(Lat* L, int i, int a) { float4 b = L->bmax[i]; return a == 0 ? b.x : (a == 1 ? b.y : (a == 2 ? b.z : b.w)); }
inline float rnd(HScene& S) { uint32_t x = S.rng; x ^= x << 13; x ^= x >> 17; x ^= x << 5; S.rng = x; return (float)(x >> 8) / 16777216.0f; }
LV rndv(HScene& S) {
    for (;;) {
        float a = rnd(S) * 2.0f - 1.0f, b = rnd(S) * 2.0f - 1.0f, c = rnd(S) * 2.0f - 1.0f;
        LV x = lv(a, b, c);
        float l = dot(x, x);
        if (l > 0.01f && l <= 1.0f) return mul(x, 1.0f / sqrtf(l));
    }
}
inline int hnbr(Lat* L, int i, int d) { uint32_t j = L->nb6[6 * (size_t)i + d]; return j == NONE ? -1 : (int)j; }
int hlinked(Lat* L, int i, int d) {
    int j = hnbr(L, i, d);
    if (j < 0 || L->mat[j] == 0 || L->body[j] != L->body[i]) return -1;
    float hp = (d & 1) == 0 ? bond_ref(L, i, d >> 1) : bond_ref(L, j, d >> 1);
    return hp > 0.0f ? j : -1;
}
inline LV hbpos(Lat* L, int i) { return L->body[i] < 0 ? lat_pos(L->bxyz[i].x, L->bxyz[i].y, L->bxyz[i].z) : hv(L->wpos[i]); }
inline bool hin_struct(Lat* L, int i, int sl) { int m = L->mat[i]; return L->sid[i] == sl && m != 0 && m < SLUG && L->body[i] < 0; }
int32_t hgget(Lat* L, const HScene& S, int x, int y, int z) {
    x -= S.gx0; y -= S.gy0; z -= S.gz0;
    if (x < 0 || y < 0 || z < 0 || x >= S.gnx || y >= S.gny || z >= S.gnz) return 0;
    return L->grid[S.goff + ((size_t)y * S.gnz + z) * S.gnx + x];
}

void st_warm(Lat* L, int g) { HStruct& T = L->sts[g]; if (T.solved) return; T.solved = true; T.warm = 560; T.settle = 1e9f; }
void st_activate(Lat* L, HScene& S, int sl) {
    if (sl < 0 || sl >= S.nst) return;
    HStruct& T = L->sts[S.st0 + sl];
    if (!T.solved) st_warm(L, S.st0 + sl);
    if (T.active < 4.0f) T.active = 4.0f;
}
void st_static_change(Lat* L, HScene& S, int i) { int s = L->sid[i]; if (s < S.nst) st_activate(L, S, s); }
void remove_brick(Lat* L, HScene& S, int i) {
    if (L->mat[i] == 0) return;
    L->mat[i] = 0;
    L->touch(i);
    if (L->body[i] < 0) {
        int s = L->sid[i];
        if (s < S.nst) { L->sts[S.st0 + s].conn = true; S.conn_any = true; }
        st_static_change(L, S, i);
    } else L->dirty[L->body[i]] = 1;
}
void remove_brick_raw(Lat* L, int i) { L->mat[i] = 0; L->body[i] = -1; L->touch(i); }
uint32_t damage_brick(Lat* L, HScene& S, int i, float frac, float absolute) {
    uint32_t broke = 0;
    for (int d = 0; d < 6; d++) {
        int j = hnbr(L, i, d);
        if (j < 0 || L->mat[j] == 0) continue;
        int o = (d & 1) == 0 ? i : j, a = d >> 1;
        float& bd = bond_ref(L, o, a);
        if (bd > 0.0f) {
            L->touch(o);
            bd -= bmax_of(L, o, a) * frac + absolute;
            if (bd <= bmax_of(L, o, a) * 0.03f) { bd = 0.0f; broke++; }
        }
    }
    float& g = bond_ref(L, i, 3);
    if (g > 0.0f) {
        L->touch(i);
        g -= bmax_of(L, i, 3) * frac * 0.5f + absolute * 0.5f;
        if (g <= 0.0f) { g = 0.0f; broke++; }
    }
    if (L->body[i] < 0) {
        st_static_change(L, S, i);
        if (broke > 0) { int s = L->sid[i]; if (s < S.nst) { HStruct& T = L->sts[S.st0 + s]; T.conn = true; S.conn_any = true; T.broke = true; } }
    } else if (broke > 0) L->dirty[L->body[i]] = 1;
    return broke;
}
void sever_brick(Lat* L, int i) {
    L->bond[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    L->touch(i);
    const int ds[3] = {1, 3, 5};
    for (int t = 0; t < 3; t++) { int j = hnbr(L, i, ds[t]); if (j >= 0) { bond_ref(L, j, ds[t] >> 1) = 0.0f; L->touch(j); } }
}
void wake_body(Lat* L, int b) { L->asleep[b] = 0; L->sleep[b] = 0.0f; L->btouch(b); }
void kill_body_host(Lat* L, int b) {
    uint32_t k = L->first[b];
    while (k != NONE) {
        uint32_t nx = L->next[k];
        if (L->body[k] == b) { L->mat[k] = 0; L->body[k] = -1; }
        k = nx;
    }
    L->alive[b] = 0; L->first[b] = NONE;
    L->lists_dirty = true;
}
void clamp_vel(Lat* L, int b) {
    L->btouch(b);
    LV vv = hv(L->v[b]);
    float vl = len(vv);
    if (vl > 60.0f) L->v[b] = hf(mul(vv, 60.0f / vl));
    LV ww = hv(L->w[b]);
    float a = len(ww);
    float amax = minf(25.0f, 16.0f / maxf(L->rad[b], 0.3f));
    if (a > amax) L->w[b] = hf(mul(ww, amax / a));
}
void impulse_at(Lat* L, int b, LV p, LV j) {
    L->btouch(b);
    L->v[b] = hf(madd(hv(L->v[b]), j, L->im[b]));
    LV r = sub(p, hv(L->pos[b]));
    L->w[b] = hf(add(hv(L->w[b]), mv(L->iiw[b], cross(r, j))));
}
void apply_blast_to_body(Lat* L, HScene& S, int b) {
    if (S.blast_t <= 0.0f) return;
    for (uint32_t k = L->first[b]; k != NONE; k = L->next[k]) {
        LV p = hv(L->wpos[k]);
        LV d = sub(p, S.blast_c);
        float dist = len(d);
        if (dist < S.blast_r) {
            float f = S.blast_p * 1600.0f * (1.0f - dist / S.blast_r);
            LV dir = dist > 1e-3f ? mul(d, 1.0f / dist) : lv(0.0f, 1.0f, 0.0f);
            dir = add(dir, lv(0.0f, 0.35f, 0.0f));
            float mm = maxf(tables().mass[L->mat[k]], 25.0f);
            impulse_at(L, b, p, mul(dir, f * mm / 300.0f));
        }
    }
    clamp_vel(L, b);
}
void wake_near(Lat* L, HScene& S, LV p, float r) {
    for (int lb = 0; lb < S.hw; lb++) {
        int b = S.bd0 + lb;
        if (L->alive[b] && L->asleep[b] && len(sub(hv(L->pos[b]), p)) < r + L->rad[b]) wake_body(L, b);
    }
}
int alloc_body(Lat* L, HScene& S) {
    for (int k = 0; k < S.dcap; k++) {
        int lb = (S.cursor + k) % S.dcap;
        int b = S.bd0 + lb;
        if (!L->alive[b]) {
            S.cursor = (lb + 1) % S.dcap;
            if (lb + 1 > S.hw) { S.hw = lb + 1; L->ub_dirty = true; }
            return b;
        }
    }
    return -1;
}
const LM IDM = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
// Bodies set up on the device, one wavefront each (setup_body, src/lib.rs:844-929). Lane 0 sums the masses and the
// mass-weighted positions brick after brick in the job's order (the reference's f32 order), then the inertia about
// the centre with each brick's own S^2/6 term, and forms the body: inverse inertia, world inertia, radius, the
// velocity of the new centre. Then every lane writes its share of the bricks: position in the body frame, world
// position, body. `in` holds each brick's position in the parent's frame (xyz) and its mass (w).
// the set-up's sums, one brick at a time in the job's order: mass and mass-weighted position; then, about the
// centre, the inertia with each brick's own S^2/6 term and the largest squared distance
__device__ __forceinline__ void setup_acc_mass(float4 a, float& mass, LV& cl) { mass += a.w; cl = madd(cl, lv(a.x, a.y, a.z), a.w); }
__device__ __forceinline__ void setup_acc_inertia(float4 a, LV cl, float* it, float& radv) {
    const float cube = S * S / 6.0f;
    float mm = a.w;
    LV r = sub(lv(a.x, a.y, a.z), cl);
    float rr = dot(r, r);
    radv = maxf(radv, rr);
    it[0] += mm * (rr - r.x * r.x + cube); it[4] += mm * (rr - r.y * r.y + cube); it[8] += mm * (rr - r.z * r.z + cube);
    it[1] -= mm * r.x * r.y; it[2] -= mm * r.x * r.z; it[5] -= mm * r.y * r.z;
}
// the body from its sums: inverse inertia, world inertia, radius, the new centre's position and velocity; written to
// the body's slot and to `o`; returns the body's position
__device__ LV setup_write(LD& D, const SetupJob& J, const LM& prot, float mass, LV cl, float* it, float radv, SetupOut& o) {
    it[3] = it[1]; it[6] = it[2]; it[7] = it[5];
    LM itm;
    for (int k = 0; k < 9; k++) itm.m[k] = it[k];
    const LM iilm = minv(itm);
    const LV off = mv(prot, cl);
    const float4 pos = f4(add(lv4(J.ppos), off));
    const float4 v = f4(add(lv4(J.pv), cross(lv4(J.pw), off)));
    const LM iiw = mmul(mmul(prot, iilm), mt(prot));
    const float rad = sqrtf(radv) + H * 1.8f;
    const int b = J.b;
    const bool keep = (J.flags & 2) != 0;
    D.alive[b] = 1; D.count[b] = (uint32_t)(J.e - J.s);
    D.pos[b] = pos; D.q[b] = J.pq; D.v[b] = v; D.w[b] = f4(lv4(J.pw)); D.vlast[b] = v; D.wlast[b] = f4(lv4(J.pw));
    store_m(D.rot + 3 * (size_t)b, prot); store_m(D.iil + 3 * (size_t)b, iilm); store_m(D.iiw + 3 * (size_t)b, iiw);
    D.m[b] = mass; D.im[b] = 1.0f / mass; D.rad[b] = rad; D.nmax[b] = 1.0f;
    D.lt[b] = keep ? J.keep_lt : 0.0f; D.resid[b] = 1.0f; D.maxr[b] = 0.0f;
    D.sleep[b] = 0.0f; D.asleep[b] = 0; D.wake[b] = 0; D.cool[b] = keep ? J.keep_cool : 0.15f;
    o.pos = pos; o.v = v;
    store_m(o.iil, iilm); store_m(o.iiw, iiw);
    o.m = mass; o.im = 1.0f / mass; o.rad = rad; o.pad = 0.0f;
    return lv4(pos);
}
// one brick of a set-up body: its position in the body frame and in the world, its body (and, leaving a structure,
// its load field cleared)
__device__ __forceinline__ void setup_brick(LD& D, const SetupJob& J, const LM& prot, LV cl, LV bpos, uint32_t i, float4 a, float4* bout) {
    const LV r = sub(lv(a.x, a.y, a.z), cl);
    const float4 loc = f4(r), wp = f4(add(bpos, mv(prot, r)));
    D.local[i] = loc; D.wpos[i] = wp; D.body[i] = J.b;
    if (J.flags & 1) { D.phi[i] = 0.0f; D.ext[i] = 0.0f; D.extn[i] = 0.0f; }
    if (bout) { bout[0] = loc; bout[1] = wp; }
}
__global__ void __launch_bounds__(64) k_setup_bodies(LD D, const SetupJob* jobs, const uint32_t* idx, const float4* in,
                                                     SetupOut* out, float4* bout) {
    constexpr int CH = 256;            // bricks staged in LDS at a time: the lanes load a chunk, lane 0 sums it
    const SetupJob J = jobs[blockIdx.x];
    __shared__ float sh[16];
    __shared__ float4 buf[CH];
    const LM prot = load_m(J.prot);
    float mass = 0.0f;
    LV cl = lv(0.0f, 0.0f, 0.0f);
    for (int c0 = J.s; c0 < J.e; c0 += CH) {
        const int m = min(CH, J.e - c0);
        for (int k = (int)threadIdx.x; k < m; k += 64) buf[k] = in[c0 + k];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < m; k++) setup_acc_mass(buf[k], mass, cl);
        __syncthreads();
    }
    cl = mul(cl, 1.0f / mass);
    float it[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    float radv = 0.0f;
    for (int c0 = J.s; c0 < J.e; c0 += CH) {
        const int m = min(CH, J.e - c0);
        for (int k = (int)threadIdx.x; k < m; k += 64) buf[k] = in[c0 + k];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < m; k++) setup_acc_inertia(buf[k], cl, it, radv);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        SetupOut o;
        const LV bp = setup_write(D, J, prot, mass, cl, it, radv, o);
        out[blockIdx.x] = o;
        sh[0] = cl.x; sh[1] = cl.y; sh[2] = cl.z; sh[3] = bp.x; sh[4] = bp.y; sh[5] = bp.z;
    }
    __syncthreads();
    const LV cen = lv(sh[0], sh[1], sh[2]), bpos = lv(sh[3], sh[4], sh[5]);
    for (int t = J.s + (int)threadIdx.x; t < J.e; t += 64) setup_brick(D, J, prot, cen, bpos, idx[t], in[t], bout + 2 * (size_t)t);
}
float4 m_row(const LM& m, int k) { return make_float4(m.m[3 * k], m.m[3 * k + 1], m.m[3 * k + 2], 0.0f); }
void row_m(LM& m, int k, float4 r) { m.m[3 * k] = r.x; m.m[3 * k + 1] = r.y; m.m[3 * k + 2] = r.z; }
// ---------------------------------------------------------------- the plough on the device (plough, src/lib.rs:1493-1550)
// One wavefront per scene takes its plough candidates in contact order (the flags k_plough_flag set, cleared as read),
// with the scene's random stream and body-slot cursor. A fast, heavy body striking a static brick: a large body
// damages the brick; otherwise the brick is severed and leaves as a body of its own (a new slot from the scene's
// cursor, set up as setup_body does, thrown forward with a random push and spin) and the body slows by the momentum it
// hands over. Contacts made void (the brick no longer static, no slot left, or the brick ploughed away: every later
// contact of the scene against it) are switched off. Lane 0 decides and writes; the lanes switch contacts off. What
// changed is logged for the host mirror, and each new body's record comes back whole.
struct NewRec { int32_t b, j, pad0, pad1; SetupOut o; float4 loc, wp, w; };
__device__ LV x_rndv(uint32_t& x) {
    for (int k = 0; k < (1 << 20); k++) {
        const float a = x_rnd(x) * 2.0f - 1.0f, b = x_rnd(x) * 2.0f - 1.0f, c = x_rnd(x) * 2.0f - 1.0f;
        const LV v = lv(a, b, c);
        const float l = dot(v, v);
        if (l > 0.01f && l <= 1.0f) return mul(v, 1.0f / sqrtf(l));
    }
    return lv(0.0f, 1.0f, 0.0f);   // (a stream stuck at zero: never drawn from a seeded scene)
}
__global__ void __launch_bounds__(64) k_plough(LD D, CrushP P, uint32_t* cur, uint32_t* hwv, NewRec* nrec, uint32_t* nnew, uint32_t ncap) {
    const int s = blockIdx.x, lane = threadIdx.x;
    if (s >= (int)D.nscene) return;
    const Tables& T = tables();
    const SceneK K = D.sc[s];
    const uint32_t ncs = min(D.ctr[C_NCS], D.ccap);
    const uint32_t c0 = min(D.sc_c0[s], ncs), c1 = s + 1 < (int)D.nscene ? min(D.sc_c0[s + 1], ncs) : ncs;
    uint32_t left = D.xpcnt[s];   // (the scene's candidates; at least as many as it flagged)
    if (left == 0) return;
    uint32_t rng = P.rng[s], cursor = cur[s], hw = hwv[s];
    const LM idm = {{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    for (uint32_t cb = c0; cb < c1 && left > 0; cb += 64) {
        const uint32_t cl = cb + (uint32_t)lane;
        const bool f = cl < c1 && D.pflag[cl] != 0;
        unsigned long long m = __ballot(f);
        if (f) D.pflag[cl] = 0;
        left -= min(left, (uint32_t)__popcll(m));
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const uint32_t c = cb + (uint32_t)l;
            int off = -2;   // -2: the contact stays; -1: it alone is switched off; j >= 0: it and later ones against j
            if (lane == 0) {
                const uint32_t j = D.cbb[c];
                if (D.mat[j] == 0 || D.body[j] >= 0) off = -1;
                else {
                    const int a = D.ca[c];
                    const float4 p4 = D.cp[c], n4 = D.cn[c];
                    const LV p = lv(p4.x, p4.y, p4.z), nn = lv(n4.x, n4.y, n4.z);
                    const LV ra = sub(p, lv4(D.pos[a]));
                    const float vn = -dot(add(lv4(D.v[a]), cross(lv4(D.w[a]), ra)), nn);
                    const int mj = D.mat[j];
                    float mom = 0.0f;
                    bool go = !(vn < 2.0f);
                    if (go) {
                        const float meff = D.m[a] * 3.0f / maxf((float)D.nc[a], 3.0f);
                        mom = minf(meff, D.m[a]) * vn;
                        go = !(mj >= SLUG || mom < T.capv[mj] * 0.35f);
                    }
                    if (go && D.count[a] > 40u) {
                        const float fr = clampf(mom / (T.capv[mj] * 1.5f), 0.15f, 0.9f);
                        x_damage(D, P, K, j, fr);
                        go = false;
                    }
                    if (go) {
                        // sever_brick; the brick's structure to check and to solve again
                        D.bond[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        x_log(P, XL_BRICK, j);
                        for (int d = 1; d < 6; d += 2) {
                            const uint32_t jn = D.nb6[6 * (size_t)j + d];
                            if (jn == NONE) continue;
                            float4 b4 = D.bond[jn];
                            f4_ref(b4, d >> 1) = 0.0f;
                            D.bond[jn] = b4;
                            x_log(P, XL_BRICK, jn);
                        }
                        const int sl = D.sid[j];
                        if (sl < K.nst) x_log(P, XL_CONN, (uint32_t)(K.st0 + sl));
                        x_static(D, P, K, j);
                        // alloc_body: the first free slot from the scene's cursor
                        int b = -1;
                        const int dcap = K.nbd;
                        for (int k = 0; k < dcap; k++) {
                            const int lb = (int)((cursor + (uint32_t)k) % (uint32_t)dcap);
                            if (!D.alive[K.bd0 + lb]) {
                                b = K.bd0 + lb;
                                cursor = (uint32_t)((lb + 1) % dcap);
                                if ((uint32_t)(lb + 1) > hw) hw = (uint32_t)(lb + 1);
                                break;
                            }
                        }
                        if (b < 0) off = -1;
                        else {
                            const LV rv = x_rndv(rng);
                            const LV vv = madd(mul(lv4(D.v[a]), 0.75f), rv, 2.0f);
                            const LV ww = mul(x_rndv(rng), 6.0f);
                            SetupJob J;
                            J.b = b; J.s = 0; J.e = 1; J.flags = 0;   // (the brick keeps its load field, as the reference's plough leaves it)
                            J.ppos = make_float4(0.0f, 0.0f, 0.0f, 0.0f); J.pq = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
                            J.pv = f4(vv); J.pw = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                            J.prot[0] = make_float4(1.0f, 0.0f, 0.0f, 0.0f); J.prot[1] = make_float4(0.0f, 1.0f, 0.0f, 0.0f);
                            J.prot[2] = make_float4(0.0f, 0.0f, 1.0f, 0.0f);
                            J.keep_cool = 0.0f; J.keep_lt = 0.0f; J.pad0 = 0.0f; J.pad1 = 0.0f;
                            const LV loc = lat_pos(D.bxyz[j].x, D.bxyz[j].y, D.bxyz[j].z);
                            const float4 a4 = make_float4(loc.x, loc.y, loc.z, T.mass[D.mat[j]]);
                            float mass = 0.0f;
                            LV cen = lv(0.0f, 0.0f, 0.0f);
                            setup_acc_mass(a4, mass, cen);
                            cen = mul(cen, 1.0f / mass);
                            float it[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                            float radv = 0.0f;
                            setup_acc_inertia(a4, cen, it, radv);
                            NewRec r;
                            const LV bp = setup_write(D, J, idm, mass, cen, it, radv, r.o);
                            float4 bo[2];
                            setup_brick(D, J, idm, cen, bp, j, a4, bo);
                            D.w[b] = f4(ww);
                            r.b = b; r.j = (int)j; r.pad0 = r.pad1 = 0; r.loc = bo[0]; r.wp = bo[1]; r.w = f4(ww);
                            const uint32_t k = atomicAdd(nnew, 1u);
                            if (k < ncap) nrec[k] = r;
                            // the striking body hands over the brick's momentum
                            const float mm = D.m[a];
                            const float fa = mm / (mm + T.mass[mj] * 1.5f);
                            D.v[a] = f4(mul(lv4(D.v[a]), fa));
                            x_log(P, XL_VEL, (uint32_t)a);
                            off = (int)j;
                        }
                    }
                }
            }
            off = __shfl(off, 0, 64);
            if (off != -2) {
                if (lane == 0) D.con[c] = 0;
                if (off >= 0)
                    for (uint32_t cc = c + 1u + (uint32_t)lane; cc < c1; cc += 64)
                        if (D.cbb[cc] == (uint32_t)off) D.con[cc] = 0;
            }
            __syncthreads();
        }
    }
    if (lane == 0) { P.rng[s] = rng; cur[s] = cursor; hwv[s] = hw; D.xpcnt[s] = 0u; }
}
__global__ void __launch_bounds__(NT) k_plough_gather(LD D, const uint32_t* log, const uint32_t* nlog, uint32_t cap, float4* v_out, uint32_t* mat_out) {
    const uint32_t n = min(*nlog, cap);
    GS(k, n) {
        const uint32_t kind = log[2 * (size_t)k], i = log[2 * (size_t)k + 1];
        if (kind == XL_BRICK) { v_out[k] = D.bond[i]; mat_out[k] = D.mat[i]; }
        else if (kind == XL_VEL) v_out[k] = D.v[i];
    }
}
// A new body in slot b from the bricks list[0, n) (in the reference's order), queued for the device's set-up
// (k_setup_bodies; setup_body, src/lib.rs:844-929). The host keeps the slot's list structure and the constants the
// set-up writes; the arithmetic comes back with run_setups. lattice: the bricks leave a structure, their positions in
// the parent's frame are their lattice cells and their load field is cleared. keep: the slot keeps its cooldown and
// life timer.
void setup_queue(Lat* L, HScene& S, int b, const uint32_t* list, int n, bool lattice, LV ppos, const LM& prot, LQ pq,
                 LV pv, LV pw, bool keep) {
    const Tables& T = tables();
    SetupJob J;
    std::memset(&J, 0, sizeof(J));
    J.b = b; J.s = (int)S.sbricks.size(); J.e = J.s + n; J.flags = (lattice ? 1 : 0) | (keep ? 2 : 0);
    J.ppos = hf(ppos); J.pq = f4q(pq); J.pv = hf(pv); J.pw = hf(pw);
    for (int k = 0; k < 3; k++) J.prot[k] = m_row(prot, k);
    J.keep_cool = L->cool[b]; J.keep_lt = L->lt[b];
    for (int t = 0; t < n; t++) {
        const int i = (int)list[t];
        const LV p = lattice ? lat_pos(L->bxyz[i].x, L->bxyz[i].y, L->bxyz[i].z) : hv(L->local[i]);
        S.sbricks.push_back((uint32_t)i);
        S.sin.push_back(make_float4(p.x, p.y, p.z, T.mass[L->mat[i]]));
        L->body[i] = b;
        L->next[i] = t + 1 < n ? list[t + 1] : NONE;
        if (lattice) { L->phi[i] = 0.0f; L->ext[i] = 0.0f; L->extn[i] = 0.0f; }
        L->touch(i);
    }
    L->alive[b] = 1; L->first[b] = list[0]; L->count[b] = (uint32_t)n;
    L->qq[b] = f4q(pq); L->rot[b] = prot; L->w[b] = hf(pw); L->wlast[b] = hf(pw);
    if (!keep) { L->lt[b] = 0.0f; L->cool[b] = 0.15f; }
    L->resid[b] = 1.0f; L->maxr[b] = 0.0f; L->nmax[b] = 1.0f;
    L->sleep[b] = 0.0f; L->asleep[b] = 0; L->dirty[b] = 0; L->wake[b] = 0;
    L->lists_dirty = true;
    S.sjobs.push_back(J);
}
// The queued set-ups of the given scenes: one upload, one launch (a wavefront per body), one copy back into the host
// mirror, one wait. The device holds the bodies and bricks once the launch has run; the mirrored bodies are marked
// for the next flush like every body the host has written.
int run_setups(Lat* L, const std::vector<int>& scenes) {
    size_t nj = 0, nbk = 0;
    for (int si : scenes) { nj += L->sc[si].sjobs.size(); nbk += L->sc[si].sbricks.size(); }
    if (!nj) return 0;
    const size_t bj = sizeof(SetupJob) * nj, bin = sizeof(float4) * nbk, bytes = bj + bin + sizeof(uint32_t) * nbk;
    uint8_t* hs = L->stg_setup.begin(bytes);
    if (!hs) return -1;
    SetupJob* hj = reinterpret_cast<SetupJob*>(hs);
    float4* hin = reinterpret_cast<float4*>(hs + bj);
    uint32_t* hidx = reinterpret_cast<uint32_t*>(hs + bj + bin);
    size_t kj = 0, kb = 0;
    for (int si : scenes) {
        HScene& S = L->sc[si];
        for (const SetupJob& J : S.sjobs) { hj[kj] = J; hj[kj].s += (int)kb; hj[kj].e += (int)kb; kj++; }
        std::memcpy(hin + kb, S.sin.data(), sizeof(float4) * S.sin.size());
        std::memcpy(hidx + kb, S.sbricks.data(), sizeof(uint32_t) * S.sbricks.size());
        kb += S.sbricks.size();
    }
    if (L->stg_setup.end(bytes, L->st)) return -1;
    const size_t outb = sizeof(SetupOut) * nj + 2 * sizeof(float4) * nbk;
    if (L->d_setup_out.n < outb && L->d_setup_out.alloc(outb + outb / 2, L->st)) return -1;
    if (L->h_setup_out.reserve(outb)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    const uint8_t* d = L->stg_setup.d.p;
    SetupOut* dout = reinterpret_cast<SetupOut*>(L->d_setup_out.p);
    float4* dbout = reinterpret_cast<float4*>(L->d_setup_out.p + sizeof(SetupOut) * nj);
    LAUNCH(k_setup_bodies, (unsigned)nj, 64, L->D, reinterpret_cast<const SetupJob*>(d), reinterpret_cast<const uint32_t*>(d + bj + bin),
           reinterpret_cast<const float4*>(d + bj), dout, dbout);
    SF_HIP(hipGetLastError());
    SF_HIP(hipMemcpyAsync(L->h_setup_out.p, L->d_setup_out.p, outb, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipStreamSynchronize(L->st));
    L->syncs++;
    const SetupOut* ho = reinterpret_cast<const SetupOut*>(L->h_setup_out.p);
    const float4* hb = reinterpret_cast<const float4*>(L->h_setup_out.p + sizeof(SetupOut) * nj);
    kj = 0; kb = 0;
    for (int si : scenes) {
        HScene& S = L->sc[si];
        for (const SetupJob& J : S.sjobs) {
            const SetupOut& o = ho[kj++];
            const int b = J.b;
            L->pos[b] = o.pos; L->v[b] = o.v; L->vlast[b] = o.v;
            for (int k = 0; k < 3; k++) { row_m(L->iil[b], k, o.iil[k]); row_m(L->iiw[b], k, o.iiw[k]); }
            L->m[b] = o.m; L->im[b] = o.im; L->rad[b] = o.rad;
            L->btouch(b);
        }
        for (size_t t = 0; t < S.sbricks.size(); t++) {
            const uint32_t i = S.sbricks[t];
            L->local[i] = hb[2 * (kb + t)]; L->wpos[i] = hb[2 * (kb + t) + 1];
        }
        kb += S.sbricks.size();
        S.sjobs.clear(); S.sbricks.clear(); S.sin.clear();
    }
    return 0;
}
// rebuild_body (src/lib.rs:930-1002), first half: the body's live bricks in list order, its removed ones unlinked.
// The parts are found on the device (k_split_*) for every listed body at once.
void rebuild_list(Lat* L, HScene& S, int id) {
    L->dirty[id] = 0;
    const int off = (int)S.rlist.size();
    uint32_t k = L->first[id];
    while (k != NONE) {
        uint32_t nx = L->next[k];
        if (L->body[k] == id) { if (L->mat[k] != 0) S.rlist.push_back(k); else { L->body[k] = -1; L->touch(k); } }
        k = nx;
    }
    L->lists_dirty = true;
    L->btouch(id);
    const int n = (int)S.rlist.size() - off;
    if (n == 0) { L->alive[id] = 0; L->first[id] = NONE; return; }
    S.rbody.push_back(make_int3(id, off, n));
}
// second half, from the parts' roots: parts numbered in the order their first brick comes in the list (the order the
// reference's search finds them), each part's bricks in list order; single bricks of a heavily loaded scene dusted,
// each part a body (the first keeps the slot), set up on the device
void rebuild_split(Lat* L, HScene& S, int id, const uint32_t* list, const uint32_t* root, int n) {
    ScrRef W = scr(L);
    uint32_t sid = ++S.stamp_id;
    uint32_t ncomp = 0;
    for (int t = 0; t < n; t++) {
        const uint32_t r = root[t];
        if (L->stamp[r] != sid) { L->stamp[r] = sid; L->comp[r] = ncomp++; }
        W.lst[t] = list[t];
        W.q[t] = L->comp[r];
    }
    int ncmp = (int)ncomp;
    for (int c = 0; c <= ncmp; c++) W.cnt[c] = 0;
    for (int t = 0; t < n; t++) W.cnt[W.q[t] + 1]++;
    for (int c = 0; c < ncmp; c++) W.cnt[c + 1] += W.cnt[c];
    for (int t = 0; t < n; t++) { int i = W.lst[t]; uint32_t c = W.q[t]; W.lst2[W.cnt[c]++] = (uint32_t)i; }
    LV ppos = hv(L->pos[id]), pv = hv(L->v[id]), pw = hv(L->w[id]);
    LM prot = L->rot[id];
    LQ pq = lq4(L->qq[id]);
    for (int c = 0; c < ncmp; c++) {
        int s0 = c == 0 ? 0 : (int)W.cnt[c - 1];
        int e = (int)W.cnt[c];
        if (c > 0 && e - s0 == 1 && S.stats[3] > 1500.0f) {
            int i = W.lst2[s0];
            int mt_ = L->mat[i];
            if (mt_ != STEEL && rnd(S) < minf((S.stats[3] - 1500.0f) / 2000.0f, 0.85f)) { remove_brick_raw(L, i); continue; }
        }
        int b = c == 0 ? id : alloc_body(L, S);
        if (b < 0) { for (int t = s0; t < e; t++) remove_brick_raw(L, W.lst2[t]); continue; }
        setup_queue(L, S, b, W.lst2.data() + s0, e - s0, false, ppos, prot, pq, pv, pw, c == 0);
    }
}
void conn_struct(Lat* L, HScene& S, int sl) {
    ScrRef W = scr(L);
    HStruct& T = L->sts[S.st0 + sl];
    uint32_t sid = ++S.stamp_id;
    const int g = S.st0 + sl;
    const int ri = g < (int)L->reach_idx.size() ? L->reach_idx[g] : -1;
    if (ri >= 0 && L->h_creach_out.p[ri] == 0u) {
        // the grounded bricks as the device's search found them (the same set the search below finds)
        const uint32_t* bits = L->h_creach_out.p + L->reach_n + L->reach_woff[ri];
        for (int i = T.lo; i < T.hi; i++) { const uint32_t k = (uint32_t)(i - T.lo); if ((bits[k >> 5] >> (k & 31u)) & 1u) L->stamp[i] = sid; }
    } else {
        int qh = 0, qt = 0;
        for (int i = T.lo; i < T.hi; i++)
            if (hin_struct(L, i, sl) && L->bxyz[i].y == 0 && L->bond[i].w > 0.0f) { L->stamp[i] = sid; W.q[qt++] = (uint32_t)i; }
        while (qh < qt) {
            int i = (int)W.q[qh++];
            for (int d = 0; d < 6; d++) {
                int j = hlinked(L, i, d);
                if (j >= 0 && L->stamp[j] != sid) { L->stamp[j] = sid; W.q[qt++] = (uint32_t)j; }
            }
        }
    }
    bool retry = false;
    for (int s0 = T.lo; s0 < T.hi; s0++) {
        if (!hin_struct(L, s0, sl) || L->stamp[s0] == sid) continue;
        int n = 0;
        L->stamp[s0] = sid; W.lst2[n++] = (uint32_t)s0;
        int h = 0;
        while (h < n) {
            int i = (int)W.lst2[h++];
            for (int d = 0; d < 6; d++) {
                int j = hlinked(L, i, d);
                if (j >= 0 && L->stamp[j] != sid) { L->stamp[j] = sid; W.lst2[n++] = (uint32_t)j; }
            }
        }
        int b = alloc_body(L, S);
        if (b < 0) { retry = true; continue; }
        LQ idq; idq.x = 0.0f; idq.y = 0.0f; idq.z = 0.0f; idq.w = 1.0f;
        setup_queue(L, S, b, W.lst2.data(), n, true, lv(0.0f, 0.0f, 0.0f), IDM, idq, lv(0.0f, 0.0f, 0.0f), lv(0.0f, 0.0f, 0.0f), false);
        S.conn_new.push_back(b);   // blasted, and its neighbours woken, once it is set up (conn_finish)
    }
    if (retry) { T.conn = true; S.conn_any = true; }
    uint32_t cntv = 0; int top = 0;
    for (int i = T.lo; i < T.hi; i++) if (hin_struct(L, i, sl)) { cntv++; if (L->bxyz[i].y + 1 > top) top = L->bxyz[i].y + 1; }
    T.now = cntv; T.top = top;
}
void connectivity(Lat* L, HScene& S) {
    if (!S.conn_any) return;
    S.conn_any = false;
    for (int sl = 0; sl < S.nst; sl++) { HStruct& T = L->sts[S.st0 + sl]; if (T.conn) { T.conn = false; conn_struct(L, S, sl); } }
}
// the detached parts, once set up, in the order they were found: the blast's impulse, then their neighbours woken
// (src/lib.rs:1118-1121; a later part, already set up, is awake and is blasted on its own turn, so the order holds)
void conn_finish(Lat* L, HScene& S) {
    for (int b : S.conn_new) {
        apply_blast_to_body(L, S, b);
        wake_near(L, S, hv(L->pos[b]), L->rad[b] + 1.5f);
    }
    S.conn_new.clear();
}
}  // namespace

namespace {

enum : int { G_MAT = 1, G_BOND = 2, G_BRICK = 4, G_BODY = 8, G_ALL = 15, G_WPOS = 16 };

template <class T> int cp_down(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) SF_HIP(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    return 0;
}
template <class T> int cp_up(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) SF_HIP(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return 0;
}
std::vector<float4> g_m3;  // staging for matrices

int pull(Lat* L, int si, int groups) {
    HScene& S = L->sc[si];
    LD& D = L->D;
    hipStream_t s = L->st;
    size_t b0 = S.b0, nb = S.nb, d0 = S.bd0, nd = S.hw;
    int r = 0;
    if (groups & G_MAT) { r |= cp_down(&L->mat[b0], D.mat + b0, nb, s); r |= cp_down(&L->body[b0], D.body + b0, nb, s); }
    if (groups & G_BOND) r |= cp_down(&L->bond[b0], D.bond + b0, nb, s);
    if ((groups & G_WPOS) && !(groups & G_BRICK)) r |= cp_down(&L->wpos[b0], D.wpos + b0, nb, s);
    if (groups & G_BRICK) {
        r |= cp_down(&L->wpos[b0], D.wpos + b0, nb, s); r |= cp_down(&L->phi[b0], D.phi + b0, nb, s);
        r |= cp_down(&L->ext[b0], D.ext + b0, nb, s); r |= cp_down(&L->extn[b0], D.extn + b0, nb, s);
    }
    std::vector<float4> rot(3 * nd), iiw(3 * nd);
    if ((groups & G_BODY) && nd) {
        r |= cp_down(&L->alive[d0], D.alive + d0, nd, s); r |= cp_down(&L->asleep[d0], D.asleep + d0, nd, s);
        r |= cp_down(&L->wake[d0], D.wake + d0, nd, s);
        r |= cp_down(&L->pos[d0], D.pos + d0, nd, s); r |= cp_down(&L->qq[d0], D.q + d0, nd, s);
        r |= cp_down(&L->v[d0], D.v + d0, nd, s); r |= cp_down(&L->w[d0], D.w + d0, nd, s);
        r |= cp_down(&L->vlast[d0], D.vlast + d0, nd, s); r |= cp_down(&L->wlast[d0], D.wlast + d0, nd, s);
        r |= cp_down(rot.data(), D.rot + 3 * d0, 3 * nd, s); r |= cp_down(iiw.data(), D.iiw + 3 * d0, 3 * nd, s);
        r |= cp_down(&L->sleep[d0], D.sleep + d0, nd, s); r |= cp_down(&L->cool[d0], D.cool + d0, nd, s);
        r |= cp_down(&L->lt[d0], D.lt + d0, nd, s); r |= cp_down(&L->resid[d0], D.resid + d0, nd, s);
        r |= cp_down(&L->maxr[d0], D.maxr + d0, nd, s); r |= cp_down(&L->nmax[d0], D.nmax + d0, nd, s);
        r |= cp_down(&L->nc[d0], D.nc + d0, nd, s);
    }
    if (r) return r;
    SF_HIP(hipStreamSynchronize(s));
    L->syncs++;
    if ((groups & G_BODY) && nd)
        for (size_t b = 0; b < nd; b++) {
            LM a, c;
            for (int k = 0; k < 3; k++) {
                float4 x = rot[3 * b + k], y = iiw[3 * b + k];
                a.m[3 * k] = x.x; a.m[3 * k + 1] = x.y; a.m[3 * k + 2] = x.z;
                c.m[3 * k] = y.x; c.m[3 * k + 1] = y.y; c.m[3 * k + 2] = y.z;
            }
            L->rot[d0 + b] = a; L->iiw[d0 + b] = c;
        }
    return 0;
}
int push(Lat* L, int si, int groups) {
    HScene& S = L->sc[si];
    LD& D = L->D;
    hipStream_t s = L->st;
    size_t b0 = S.b0, nb = S.nb, d0 = S.bd0, nd = S.hw;
    int r = 0;
    if (groups & G_MAT) { r |= cp_up(D.mat + b0, &L->mat[b0], nb, s); r |= cp_up(D.body + b0, &L->body[b0], nb, s); }
    if (groups & G_BOND) r |= cp_up(D.bond + b0, &L->bond[b0], nb, s);
    if (groups & G_BRICK) {
        r |= cp_up(D.local + b0, &L->local[b0], nb, s); r |= cp_up(D.wpos + b0, &L->wpos[b0], nb, s);
        r |= cp_up(D.phi + b0, &L->phi[b0], nb, s); r |= cp_up(D.ext + b0, &L->ext[b0], nb, s);
        r |= cp_up(D.extn + b0, &L->extn[b0], nb, s);
    }
    std::vector<float4> rot(3 * nd), iil(3 * nd), iiw(3 * nd);
    if ((groups & G_BODY) && nd) {
        for (size_t b = 0; b < nd; b++)
            for (int k = 0; k < 3; k++) {
                const LM &a = L->rot[d0 + b], &c = L->iil[d0 + b], &e = L->iiw[d0 + b];
                rot[3 * b + k] = make_float4(a.m[3 * k], a.m[3 * k + 1], a.m[3 * k + 2], 0.0f);
                iil[3 * b + k] = make_float4(c.m[3 * k], c.m[3 * k + 1], c.m[3 * k + 2], 0.0f);
                iiw[3 * b + k] = make_float4(e.m[3 * k], e.m[3 * k + 1], e.m[3 * k + 2], 0.0f);
            }
        r |= cp_up(D.alive + d0, &L->alive[d0], nd, s); r |= cp_up(D.asleep + d0, &L->asleep[d0], nd, s);
        r |= cp_up(D.wake + d0, &L->wake[d0], nd, s);
        r |= cp_up(D.pos + d0, &L->pos[d0], nd, s); r |= cp_up(D.q + d0, &L->qq[d0], nd, s);
        r |= cp_up(D.v + d0, &L->v[d0], nd, s); r |= cp_up(D.w + d0, &L->w[d0], nd, s);
        r |= cp_up(D.vlast + d0, &L->vlast[d0], nd, s); r |= cp_up(D.wlast + d0, &L->wlast[d0], nd, s);
        r |= cp_up(D.rot + 3 * d0, rot.data(), 3 * nd, s); r |= cp_up(D.iil + 3 * d0, iil.data(), 3 * nd, s);
        r |= cp_up(D.iiw + 3 * d0, iiw.data(), 3 * nd, s);
        r |= cp_up(D.m + d0, &L->m[d0], nd, s); r |= cp_up(D.im + d0, &L->im[d0], nd, s);
        r |= cp_up(D.sleep + d0, &L->sleep[d0], nd, s); r |= cp_up(D.rad + d0, &L->rad[d0], nd, s);
        r |= cp_up(D.cool + d0, &L->cool[d0], nd, s); r |= cp_up(D.lt + d0, &L->lt[d0], nd, s);
        r |= cp_up(D.resid + d0, &L->resid[d0], nd, s); r |= cp_up(D.maxr + d0, &L->maxr[d0], nd, s);
        r |= cp_up(D.nmax + d0, &L->nmax[d0], nd, s); r |= cp_up(D.count + d0, &L->count[d0], nd, s);
        r |= cp_up(D.nc + d0, &L->nc[d0], nd, s);
    }
    if (r) return r;
    SF_HIP(hipStreamSynchronize(s));  // staging vectors go out of scope
    L->syncs++;
    return 0;
}

int upload_scenes(Lat* L);
// dynamic slots: every live body's bricks in list order, bodies in index order, scenes in order
int rebuild_csr(Lat* L) {
    if (!L->lists_dirty) return 0;
    L->lists_dirty = false;
    uint32_t k = 0;
    std::vector<int4>& csr = L->csr;
    csr.clear();
    for (size_t si = 0; si < L->sc.size(); si++) {
        HScene& S = L->sc[si];
        S.dyn0 = k;
        for (int lb = 0; lb < S.hw; lb++) {
            int b = S.bd0 + lb;
            if (!L->alive[b]) continue;
            uint32_t o = k;
            for (uint32_t i = L->first[b]; i != NONE; i = L->next[i]) { L->dyn[k] = i; L->dyn_body[k] = b; k++; }
            csr.push_back(make_int4(b, (int)o, (int)(k - o), 0));
        }
        S.ndyn = k - S.dyn0;
    }
    L->ndyn = k;
    L->D.ndyn = k;
    if (lat_hash_alloc(L, k)) return -1;
    // the lists and the live bodies' brick ranges (the ranges of dead bodies are never read) in one upload
    const size_t nc = csr.size(), o_csr = (8 * (size_t)k + 15) & ~(size_t)15, bytes = o_csr + sizeof(int4) * nc;
    if (k || nc) {
        uint8_t* blk = L->stg_csr.begin(bytes);
        if (!blk) return -1;
        memcpy(blk, L->dyn.data(), sizeof(uint32_t) * k);
        memcpy(blk + 4 * (size_t)k, L->dyn_body.data(), sizeof(int32_t) * k);
        memcpy(blk + o_csr, csr.data(), sizeof(int4) * nc);
        if (L->stg_csr.end(bytes, L->st)) return -1;
        LAUNCH(k_csr_set, sf_grid(std::max<size_t>(k, nc)), NT, L->D, (const uint8_t*)L->stg_csr.d.p, k, (uint32_t)nc, (uint32_t)o_csr);
    }
    return upload_scenes(L);  // the scenes' first dynamic slots moved
}

int upload_scenes(Lat* L) {
    if (L->ub_dirty) {
        L->ub_dirty = false;
        L->ubody.clear();
        for (const HScene& S : L->sc) for (int lb = 0; lb < S.hw; lb++) L->ubody.push_back((uint32_t)(S.bd0 + lb));
        L->D.nub = (uint32_t)L->ubody.size();
        if (cp_up(L->D.ubody, L->ubody.data(), L->ubody.size(), L->st)) return -1;
    }
    std::vector<SceneK> k(L->sc.size());
    for (size_t si = 0; si < L->sc.size(); si++) {
        const HScene& S = L->sc[si];
        SceneK& x = k[si];
        x.frame = S.frame; x.blast_t = S.blast_t;
        x.gx0 = S.gx0; x.gy0 = S.gy0; x.gz0 = S.gz0; x.gnx = S.gnx; x.gny = S.gny; x.gnz = S.gnz; x.goff = (uint32_t)S.goff;
        x.bd0 = S.bd0; x.nbd = S.dcap; x.dyn0 = S.dyn0; x.ndyn = S.ndyn; x.st0 = S.st0; x.nst = S.nst; x.b0 = S.b0; x.nb = S.nb; x.hw = S.hw;
    }
    // unchanged since the last upload: nothing to do; else an asynchronous copy from one of two pinned buffers
    if (k.size() == L->sk_last.size() && !k.empty() && memcmp(k.data(), L->sk_last.data(), sizeof(SceneK) * k.size()) == 0) return 0;
    L->sk_last = k;
    if (k.empty()) return 0;
    int b = L->sk_cur ^= 1;
    if (!L->ev_sk[b]) SF_HIP(hipEventCreateWithFlags(&L->ev_sk[b], hipEventDisableTiming));
    if (L->sk_recorded[b]) SF_HIP(hipEventSynchronize(L->ev_sk[b]));
    if (L->h_sk[b].reserve(k.size())) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    memcpy(L->h_sk[b].p, k.data(), sizeof(SceneK) * k.size());
    SF_HIP(hipMemcpyAsync(L->d_sc.p, L->h_sk[b].p, sizeof(SceneK) * k.size(), hipMemcpyHostToDevice, L->st));
    SF_HIP(hipEventRecord(L->ev_sk[b], L->st));
    L->sk_recorded[b] = true;
    return 0;
}


int fill_u32(Lat* L, uint32_t* p, uint32_t v, size_t n) { if (n) LAUNCH(k_fill_u32, sf_grid(n), NT, p, v, n); return 0; }
// every counter but the killed-body count, which accumulates until the host has drained the list
__global__ void k_reset_ctr(uint32_t* c) { uint32_t i = threadIdx.x; if (i < (uint32_t)C_NCTR && i != (uint32_t)C_NKILL) c[i] = 0u; }
void reset_ctr(Lat* L) { LAUNCH(k_reset_ctr, 1, 64, L->D.ctr); }
int finish_stats(Lat* L);
// the device counters into pinned memory (and with nsc, the first contact of each scene after them, at
// h_ctr + CTR_WORDS); the wait also completes every read-back queued before it
int read_ctr(Lat* L, size_t nsc = 0) {
    SF_HIP(hipMemcpyAsync(L->h_ctr, L->D.ctr, sizeof(uint32_t) * (nsc ? CTR_WORDS + nsc : (size_t)C_NCTR), hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipStreamSynchronize(L->st));
    L->syncs++;
    return finish_stats(L);
}

// bodies killed on the device (pinch, out of bounds): the host mirror follows
int drain_kills(Lat* L) {
    uint32_t n = L->h_ctr[C_NKILL];
    if (!n) return 0;
    n = std::min<uint32_t>(n, (uint32_t)L->nbd);
    if (L->h_kl.reserve(n)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    SF_HIP(hipMemcpyAsync(L->h_kl.p, L->D.klist, sizeof(int32_t) * n, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipStreamSynchronize(L->st));
    L->syncs++;
    for (uint32_t t = 0; t < n; t++) kill_body_host(L, L->h_kl.p[t]);
    L->h_ctr[C_NKILL] = 0;
    return fill_u32(L, L->D.ctr + C_NKILL, 0, 1);
}


__global__ void k_brick_delta(LD D, const BrickDelta* d, uint32_t n) {
    GS(t, n) {
        BrickDelta x = d[t];
        D.mat[x.i] = (uint8_t)x.mat; D.body[x.i] = x.body; D.bond[x.i] = x.bond;
        if (x.full & 2u) { D.local[x.i] = x.local; D.wpos[x.i] = x.wpos; }
        if (x.full & 4u) { D.phi[x.i] = x.phi; D.ext[x.i] = x.ext; D.extn[x.i] = x.extn; }
    }
}
__global__ void k_body_pack(LD D, const uint32_t* idx, uint32_t n, BodyRec* out) {
    GS(t, n) {
        uint32_t b = idx[t];
        BodyRec r;
        r.b = b; r.alive = D.alive[b]; r.asleep = D.asleep[b]; r.wake = D.wake[b]; r.pad0 = 0; r.count = D.count[b]; r.nc = D.nc[b];
        r.pos = D.pos[b]; r.q = D.q[b]; r.v = D.v[b]; r.w = D.w[b]; r.vlast = D.vlast[b]; r.wlast = D.wlast[b];
        for (int k = 0; k < 3; k++) { r.rot[k] = D.rot[3 * (size_t)b + k]; r.iil[k] = D.iil[3 * (size_t)b + k]; r.iiw[k] = D.iiw[3 * (size_t)b + k]; }
        r.m = D.m[b]; r.im = D.im[b]; r.sleep = D.sleep[b]; r.rad = D.rad[b]; r.cool = D.cool[b]; r.lt = D.lt[b];
        r.resid = D.resid[b]; r.maxr = D.maxr[b]; r.nmax = D.nmax[b]; r.pad1 = r.pad2 = r.pad3 = 0.0f;
        out[t] = r;
    }
}
__global__ void k_body_unpack(LD D, const BodyRec* in, uint32_t n) {
    GS(t, n) {
        BodyRec r = in[t];
        uint32_t b = r.b;
        D.alive[b] = r.alive; D.asleep[b] = r.asleep; D.wake[b] = r.wake; D.count[b] = r.count; D.nc[b] = r.nc;
        D.pos[b] = r.pos; D.q[b] = r.q; D.v[b] = r.v; D.w[b] = r.w; D.vlast[b] = r.vlast; D.wlast[b] = r.wlast;
        for (int k = 0; k < 3; k++) { D.rot[3 * (size_t)b + k] = r.rot[k]; D.iil[3 * (size_t)b + k] = r.iil[k]; D.iiw[3 * (size_t)b + k] = r.iiw[k]; }
        D.m[b] = r.m; D.im[b] = r.im; D.sleep[b] = r.sleep; D.rad[b] = r.rad; D.cool[b] = r.cool; D.lt[b] = r.lt;
        D.resid[b] = r.resid; D.maxr[b] = r.maxr; D.nmax[b] = r.nmax;
    }
}

// the scene's body slots in use, from the device into the host mirror (one packing kernel, one copy)
// the body records of the given scenes (every slot up to each scene's high-water mark) into the host mirror: one
// pack, one copy, one wait for all of them
int pull_bodies(Lat* L, const std::vector<int>& scenes) {
    size_t n = 0;
    for (int si : scenes) n += (size_t)L->sc[si].hw;
    if (!n) return 0;
    if (L->h_idx.reserve(n) || L->h_brec.reserve(n)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    if ((L->d_idx.n < n && L->d_idx.alloc(std::max<size_t>(n, 4096), L->st)) || (L->d_brec.n < n && L->d_brec.alloc(std::max<size_t>(n, 4096), L->st))) {
        sf::set_error(SF_E_HIP, "hipMalloc failed");
        return -1;
    }
    size_t k = 0;
    for (int si : scenes) { const HScene& S = L->sc[si]; for (int lb = 0; lb < S.hw; lb++) L->h_idx.p[k++] = (uint32_t)(S.bd0 + lb); }
    SF_HIP(hipMemcpyAsync(L->d_idx.p, L->h_idx.p, sizeof(uint32_t) * n, hipMemcpyHostToDevice, L->st));
    LAUNCH(k_body_pack, sf_grid(n), NT, L->D, (const uint32_t*)L->d_idx.p, (uint32_t)n, L->d_brec.p);
    SF_HIP(hipMemcpyAsync(L->h_brec.p, L->d_brec.p, sizeof(BodyRec) * n, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipStreamSynchronize(L->st));
    L->syncs++;
    for (size_t t = 0; t < n; t++) {
        const BodyRec& x = L->h_brec.p[t];
        int b = (int)x.b;
        L->alive[b] = x.alive; L->asleep[b] = x.asleep; L->wake[b] = x.wake; L->nc[b] = x.nc;
        L->pos[b] = x.pos; L->qq[b] = x.q; L->v[b] = x.v; L->w[b] = x.w; L->vlast[b] = x.vlast; L->wlast[b] = x.wlast;
        for (int k2 = 0; k2 < 3; k2++) { row_m(L->rot[b], k2, x.rot[k2]); row_m(L->iiw[b], k2, x.iiw[k2]); }
        L->sleep[b] = x.sleep; L->cool[b] = x.cool; L->lt[b] = x.lt; L->resid[b] = x.resid; L->maxr[b] = x.maxr; L->nmax[b] = x.nmax;
    }
    return 0;
}
int pull_bodies(Lat* L, int si) { return pull_bodies(L, std::vector<int>{si}); }
// upload what the host changed: bricks (material, body, bonds, local and world position, load field) and bodies
int flush_delta(Lat* L) {
    size_t n = L->tlist.size(), nb = L->btlist.size();
    if (n) {
        BrickDelta* hd = reinterpret_cast<BrickDelta*>(L->stg_delta.begin(sizeof(BrickDelta) * n));
        if (!hd) return -1;
        for (size_t t = 0; t < n; t++) {
            uint32_t i = L->tlist[t];
            BrickDelta& x = hd[t];
            x.i = i; x.body = L->body[i]; x.mat = L->mat[i]; x.full = L->touched[i]; x.bond = L->bond[i]; x.local = L->local[i]; x.wpos = L->wpos[i];
            x.phi = L->phi[i]; x.ext = L->ext[i]; x.extn = L->extn[i]; x.pad2 = 0.0f;
            L->touched[i] = 0;
        }
        L->tlist.clear();
        if (L->stg_delta.end(sizeof(BrickDelta) * n, L->st)) return -1;
        LAUNCH(k_brick_delta, sf_grid(n), NT, L->D, (const BrickDelta*)L->stg_delta.d.p, (uint32_t)n);
    }
    if (nb) {
        BodyRec* hb = reinterpret_cast<BodyRec*>(L->stg_brec.begin(sizeof(BodyRec) * nb));
        if (!hb) return -1;
        for (size_t t = 0; t < nb; t++) {
            uint32_t b = L->btlist[t];
            BodyRec& x = hb[t];
            x.b = b; x.alive = L->alive[b]; x.asleep = L->asleep[b]; x.wake = L->wake[b]; x.pad0 = 0; x.count = L->count[b]; x.nc = L->nc[b];
            x.pos = L->pos[b]; x.q = L->qq[b]; x.v = L->v[b]; x.w = L->w[b]; x.vlast = L->vlast[b]; x.wlast = L->wlast[b];
            for (int k = 0; k < 3; k++) { x.rot[k] = m_row(L->rot[b], k); x.iil[k] = m_row(L->iil[b], k); x.iiw[k] = m_row(L->iiw[b], k); }
            x.m = L->m[b]; x.im = L->im[b]; x.sleep = L->sleep[b]; x.rad = L->rad[b]; x.cool = L->cool[b]; x.lt = L->lt[b];
            x.resid = L->resid[b]; x.maxr = L->maxr[b]; x.nmax = L->nmax[b]; x.pad1 = x.pad2 = x.pad3 = 0.0f;
            L->btouched[b] = 0;
        }
        L->btlist.clear();
        if (L->stg_brec.end(sizeof(BodyRec) * nb, L->st)) return -1;
        LAUNCH(k_body_unpack, sf_grid(nb), NT, L->D, (const BodyRec*)L->stg_brec.d.p, (uint32_t)nb);
    }
    return 0;
}

double ms_since(std::chrono::steady_clock::time_point t) { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count(); }
// (plough, src/lib.rs:1493-1550) on the device (k_plough, a wavefront per scene with the scenes' random streams and
// slot cursors); its log and the new bodies' records come back and the host mirror follows them
int plough_phase(Lat* L, uint32_t n) {
    auto t0 = std::chrono::steady_clock::now();
    struct Acc { Lat* L; std::chrono::steady_clock::time_point t; ~Acc() { L->ph[1] += ms_since(t); } } acc{L, t0};
    if (!n) return 0;
    const size_t ns = L->sc.size();
    // log entries per candidate: a damaged brick (7 bricks and 4 marks) or a severed one (4 bricks, 2 marks) and the
    // striking body's velocity; new bodies: at most one per candidate
    const size_t cap = 12 * (size_t)n + 1024, ncap = (size_t)n + 16;
    const size_t o_cur = (ns * 4 + 15) / 16 * 16, o_hw = o_cur + o_cur, o_cnt = o_hw + o_cur, o_log = o_cnt + 16,
                 o_v = o_log + 8 * cap, o_mat = o_v + 16 * cap, o_new = (o_mat + 4 * cap + 15) / 16 * 16,
                 bytes = o_new + sizeof(NewRec) * ncap;
    if (L->d_crush.n < bytes && L->d_crush.alloc(2 * bytes, L->st)) return -1;
    uint8_t* d = L->d_crush.p;
    uint32_t* hs = reinterpret_cast<uint32_t*>(L->stg_crush.begin(o_log));
    if (!hs) return -1;
    std::memset(hs, 0, o_log);
    for (size_t si = 0; si < ns; si++) {
        const HScene& S = L->sc[si];
        hs[si] = S.rng;
        hs[o_cur / 4 + si] = (uint32_t)S.cursor;
        hs[o_hw / 4 + si] = (uint32_t)S.hw;
    }
    if (L->stg_crush.end(o_log, L->st)) return -1;
    SF_HIP(hipMemcpyAsync(d, L->stg_crush.d.p, o_log, hipMemcpyDeviceToDevice, L->st));
    CrushP P;
    P.log = reinterpret_cast<uint32_t*>(d + o_log); P.nlog = reinterpret_cast<uint32_t*>(d + o_cnt);
    P.rng = reinterpret_cast<uint32_t*>(d); P.logcap = (uint32_t)cap; P.pad = 0;
    uint32_t* dnnew = reinterpret_cast<uint32_t*>(d + o_cnt) + 1;
    NewRec* dnew = reinterpret_cast<NewRec*>(d + o_new);
    LAUNCH(k_plough, (unsigned)ns, 64, L->D, P, reinterpret_cast<uint32_t*>(d + o_cur), reinterpret_cast<uint32_t*>(d + o_hw), dnew, dnnew, (uint32_t)ncap);
    LAUNCH(k_plough_gather, sf_grid(std::min<size_t>(cap, 1u << 20)), NT, L->D, (const uint32_t*)P.log, (const uint32_t*)P.nlog, (uint32_t)cap,
           reinterpret_cast<float4*>(d + o_v), reinterpret_cast<uint32_t*>(d + o_mat));
    SF_HIP(hipGetLastError());
    if (L->h_crush.reserve(o_log)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    uint8_t* h = L->h_crush.p;
    SF_HIP(hipMemcpyAsync(h, d, o_log, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipStreamSynchronize(L->st));
    L->syncs++;
    const uint32_t* hh = reinterpret_cast<const uint32_t*>(h);
    const uint32_t nlog = hh[o_cnt / 4], nnew = hh[o_cnt / 4 + 1];
    if (nlog > cap || nnew > ncap) { sf::set_error(SF_E_CAP, "plough log overflow"); return SF_E_CAP; }
    for (size_t si = 0; si < ns; si++) {
        HScene& S = L->sc[si];
        S.rng = hh[si];
        S.cursor = (int)hh[o_cur / 4 + si];
        const int hw = (int)hh[o_hw / 4 + si];
        if (hw > S.hw) { S.hw = hw; L->ub_dirty = true; }
    }
    if (nlog || nnew) {
        std::vector<uint32_t> lg(2 * (size_t)nlog), mt_(nlog);
        std::vector<float4> vv(nlog);
        std::vector<NewRec> nr(nnew);
        if (nlog) {
            SF_HIP(hipMemcpyAsync(lg.data(), d + o_log, 8 * (size_t)nlog, hipMemcpyDeviceToHost, L->st));
            SF_HIP(hipMemcpyAsync(vv.data(), d + o_v, 16 * (size_t)nlog, hipMemcpyDeviceToHost, L->st));
            SF_HIP(hipMemcpyAsync(mt_.data(), d + o_mat, 4 * (size_t)nlog, hipMemcpyDeviceToHost, L->st));
        }
        if (nnew) SF_HIP(hipMemcpyAsync(nr.data(), dnew, sizeof(NewRec) * nnew, hipMemcpyDeviceToHost, L->st));
        SF_HIP(hipStreamSynchronize(L->st));
        L->syncs++;
        for (uint32_t k = 0; k < nlog; k++) {
            const uint32_t kind = lg[2 * (size_t)k], x = lg[2 * (size_t)k + 1];
            if (kind == XL_BRICK) { L->mat[x] = (uint8_t)mt_[k]; L->bond[x] = vv[k]; continue; }
            if (kind == XL_VEL) { L->v[x] = vv[k]; continue; }
            if (kind == XL_DIRTY) { L->dirty[x] = 1; continue; }
            HStruct& T = L->sts[x];
            HScene& S = L->sc[T.scene];
            if (kind == XL_ACT) st_activate(L, S, (int)x - S.st0);
            else if (kind == XL_CONN) { T.conn = true; S.conn_any = true; }
            else if (kind == XL_BROKE) T.broke = true;
        }
        // the new bodies in the mirror: the slot's list and constants as setup_body leaves them, its numbers from the device
        for (const NewRec& r : nr) {
            const int b = r.b, j = r.j;
            L->body[j] = b; L->next[j] = NONE; L->local[j] = r.loc; L->wpos[j] = r.wp;
            L->alive[b] = 1; L->first[b] = (uint32_t)j; L->count[b] = 1;
            L->qq[b] = make_float4(0.0f, 0.0f, 0.0f, 1.0f); L->rot[b] = IDM;
            L->pos[b] = r.o.pos; L->v[b] = r.o.v; L->vlast[b] = r.o.v; L->w[b] = r.w; L->wlast[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            for (int k = 0; k < 3; k++) { row_m(L->iil[b], k, r.o.iil[k]); row_m(L->iiw[b], k, r.o.iiw[k]); }
            L->m[b] = r.o.m; L->im[b] = r.o.im; L->rad[b] = r.o.rad;
            L->lt[b] = 0.0f; L->cool[b] = 0.15f; L->resid[b] = 1.0f; L->maxr[b] = 0.0f; L->nmax[b] = 1.0f;
            L->sleep[b] = 0.0f; L->asleep[b] = 0; L->dirty[b] = 0; L->wake[b] = 0;
            L->lists_dirty = true;
        }
    }
    if (flush_delta(L)) return -1;
    return rebuild_csr(L);
}

// (src/lib.rs:1610-1642) in contact order, a budget of 2500 crushed bricks per scene and substep: the walks run on the
// device (k_crush, a wavefront per scene, with the scenes' random streams), their log comes back in one copy (two when
// it is long) and the host mirror follows it: bricks' materials and bonds, bodies to rebuild, structures to activate,
// check and mark broken.
int impact_phase(Lat* L, uint32_t n) {
    struct Acc { Lat* L; std::chrono::steady_clock::time_point t; ~Acc() { L->ph[2] += ms_since(t); } } acc{L, std::chrono::steady_clock::now()};
    if (!n) return 0;
    const size_t ns = L->sc.size();
    // log entries: per scene at most 2,500 crushed bricks a phase, each 6 entries (brick and three neighbours severed,
    // a structure or body mark), and per candidate two damaged walk ends (7 bricks and 4 marks each) and a mark
    const size_t cap = 15016 * ns + 23 * (size_t)std::max<uint32_t>(n, 4096) + 1024;
    const size_t o_rng = 0, o_n = (ns * 4 + 15) / 16 * 16, o_log = o_n + 16, o_bond = o_log + 8 * cap, o_mat = o_bond + 16 * cap,
                 bytes = o_mat + 4 * cap;
    if (L->d_crush.n < bytes && L->d_crush.alloc(2 * bytes, L->st)) return -1;   // (doubled: growing allocates and waits)
    const size_t pref = 2048;   // log entries brought back with the first copy
    const size_t hb = o_log + pref * (8 + 16 + 4);
    if (L->h_crush.reserve(std::max(hb, o_log + 8))) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    uint8_t* d = L->d_crush.p;
    uint32_t* hr = reinterpret_cast<uint32_t*>(L->stg_crush.begin(o_log));
    if (!hr) return -1;
    std::memset(hr, 0, o_log);
    for (size_t si = 0; si < ns; si++) hr[si] = L->sc[si].rng;
    if (L->stg_crush.end(o_log, L->st)) return -1;
    SF_HIP(hipMemcpyAsync(d, L->stg_crush.d.p, o_log, hipMemcpyDeviceToDevice, L->st));
    CrushP P;
    P.log = reinterpret_cast<uint32_t*>(d + o_log); P.nlog = reinterpret_cast<uint32_t*>(d + o_n);
    P.rng = reinterpret_cast<uint32_t*>(d + o_rng); P.logcap = (uint32_t)cap; P.pad = 0;
    float4* dbond = reinterpret_cast<float4*>(d + o_bond);
    uint32_t* dmat = reinterpret_cast<uint32_t*>(d + o_mat);
    LAUNCH(k_crush, (unsigned)ns, 64, L->D, P);
    LAUNCH(k_crush_gather, sf_grid(std::min<size_t>(cap, 1u << 20)), NT, L->D, (const uint32_t*)P.log, (const uint32_t*)P.nlog, (uint32_t)cap, dbond, dmat);
    SF_HIP(hipGetLastError());
    uint8_t* h = L->h_crush.p;
    const size_t h_log = o_log, h_bond = h_log + 8 * pref, h_mat = h_bond + 16 * pref;
    SF_HIP(hipMemcpyAsync(h, d, o_log, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipMemcpyAsync(h + h_log, d + o_log, 8 * pref, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipMemcpyAsync(h + h_bond, d + o_bond, 16 * pref, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipMemcpyAsync(h + h_mat, d + o_mat, 4 * pref, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipStreamSynchronize(L->st));
    L->syncs++;
    const uint32_t* hrng = reinterpret_cast<const uint32_t*>(h);
    const uint32_t nlog = reinterpret_cast<const uint32_t*>(h + o_n)[0];
    if (nlog > cap) { sf::set_error(SF_E_CAP, "crush log overflow"); return SF_E_CAP; }
    for (size_t si = 0; si < ns; si++) L->sc[si].rng = hrng[si];
    std::vector<uint32_t> xlog;
    std::vector<float4> xbond;
    std::vector<uint32_t> xmat;
    const uint32_t* lg = reinterpret_cast<const uint32_t*>(h + h_log);
    const float4* bd = reinterpret_cast<const float4*>(h + h_bond);
    const uint32_t* mt_ = reinterpret_cast<const uint32_t*>(h + h_mat);
    if (nlog > pref) {
        xlog.resize(2 * (size_t)nlog); xbond.resize(nlog); xmat.resize(nlog);
        SF_HIP(hipMemcpyAsync(xlog.data(), d + o_log, 8 * (size_t)nlog, hipMemcpyDeviceToHost, L->st));
        SF_HIP(hipMemcpyAsync(xbond.data(), d + o_bond, 16 * (size_t)nlog, hipMemcpyDeviceToHost, L->st));
        SF_HIP(hipMemcpyAsync(xmat.data(), d + o_mat, 4 * (size_t)nlog, hipMemcpyDeviceToHost, L->st));
        SF_HIP(hipStreamSynchronize(L->st));
        L->syncs++;
        lg = xlog.data(); bd = xbond.data(); mt_ = xmat.data();
    }
    for (uint32_t k = 0; k < nlog; k++) {
        const uint32_t kind = lg[2 * (size_t)k], x = lg[2 * (size_t)k + 1];
        if (kind == XL_BRICK) { L->mat[x] = (uint8_t)mt_[k]; L->bond[x] = bd[k]; continue; }
        if (kind == XL_DIRTY) { L->dirty[x] = 1; continue; }
        HStruct& T = L->sts[x];
        HScene& S = L->sc[T.scene];
        if (kind == XL_ACT) st_activate(L, S, (int)x - S.st0);
        else if (kind == XL_CONN) { T.conn = true; S.conn_any = true; }
        else if (kind == XL_BROKE) T.broke = true;
    }
    if (flush_delta(L)) return -1;
    return rebuild_csr(L);
}

int scan(Lat* L, const uint32_t* in, uint32_t* out, uint32_t* total, uint32_t n) {
    sf::scan_u32(in, out, total, nullptr, n, n, L->D.scan_tmp, L->st);
    return 0;
}
void segsort(Lat* L, const uint32_t* off, const uint32_t* cnt, uint32_t* val, uint32_t nseg) {
    if (!nseg) return;
    LAUNCH(k_segsort_all, sf_grid(nseg, NT, 4096), NT, off, cnt, val, nseg);
}

int substep(Lat* L, float h) {
    LD& D = L->D;
    const unsigned gb = sf_grid(D.nub), gd = sf_grid(std::max<uint32_t>(L->ndyn, 1));
    const unsigned gw = sf_grid((uint64_t)std::max<uint32_t>(L->ndyn, 1) * 64, NT, 65536);  // a wave per dynamic brick
    LAUNCH(k_cache, gb, NT, D, 1, h);
    LAUNCH(k_wpos, gd, NT, D);
    // hash
    LAUNCH(k_hash_clear, sf_grid(L->tsize), NT, D);
    LAUNCH(k_hash_count, gd, NT, D);
    scan(L, D.hcnt, D.hoff, D.ctr + C_NCTR0, L->tsize);
    LAUNCH(k_hash_fill, gd, NT, D);
    segsort(L, D.hoff, D.hcnt, D.hent, L->tsize);
    // contacts: count, size, write. The write pass runs without waiting for the count: the contact arrays hold four
    // per brick, and a substep that needs more grows them and writes again (the pass is repeatable: contact counts
    // restart from zero, stamps and wake flags are idempotent)
    LAUNCH(k_contacts<false>, gw, NT, D);
    scan(L, D.ccnt, D.coff, D.ctr + C_NCS, L->ndyn);
    const uint32_t* c0 = L->h_ctr + CTR_WORDS;
    for (int pass = 0;; pass++) {
        if (pass > 0) LAUNCH(k_nc_clear, gb, NT, D);  // (the count pass cleared them the first time)
        LAUNCH(k_contacts<true>, gw, NT, D);
        LAUNCH(k_plough_flag, sf_grid(L->ccap, NT, 1024), NT, D);
        if (read_ctr(L, L->sc.size())) return -1;
        uint32_t need = L->h_ctr[C_NCS];
        if (need <= L->ccap) break;
        if (pass > 0 || lat_contacts_alloc(L, std::max<uint32_t>(need + need / 2, 1u << 16))) { sf::set_error(SF_E_CAP, "contact arrays"); return SF_E_CAP; }
        LAUNCH(k_fill_u32, 1, 64, D.ctr + C_NPLOUGH, 0u, (size_t)1);
    }
    if (drain_kills(L)) return -1;
    const uint32_t ncs = L->h_ctr[C_NCS];
    for (size_t si = 0; si < L->sc.size(); si++) {
        L->sc[si].c0 = c0[si];
        uint32_t end = si + 1 < L->sc.size() ? c0[si + 1] : ncs;
        L->sc[si].ncs = end >= c0[si] ? end - c0[si] : 0;
    }
    if (L->h_ctr[C_NPLOUGH] && plough_phase(L, L->h_ctr[C_NPLOUGH])) return -1;
    const unsigned gc = sf_grid(std::max<uint32_t>(ncs, 1));
    LAUNCH(k_prep, std::max(gc, gb), NT, D, h);
    // every scene small enough: one workgroup per scene solves it whole; else the colour-by-colour launches
    bool fused = L->fuse_max > 0;
    L->slist.clear();
    for (size_t si = 0; si < L->sc.size(); si++) {
        uint32_t n = L->sc[si].ncs;
        if (n > (uint32_t)L->fuse_max) fused = false;
        if (n) L->slist.push_back((int32_t)si);
    }
    if (fused) {
        if (!L->slist.empty()) {
            if (L->slist != L->slist_up) {
                SF_HIP(hipMemcpyAsync(L->d_slist.p, L->slist.data(), sizeof(int32_t) * L->slist.size(), hipMemcpyHostToDevice, L->st));
                L->slist_up = L->slist;
            }
            // bodies of scenes up to the LDS capacity (84 bytes each) are solved in LDS
            if (!L->lds_max) {
                int v = 0;
                if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, L->dev) != hipSuccess || v <= 0) v = 65536;
                L->lds_max = v;
            }
            uint32_t fit = (uint32_t)std::max(0, (L->lds_max - 2048) / 84), cap = 0;
            for (int32_t si : L->slist) { uint32_t hw = (uint32_t)L->sc[si].hw; if (hw <= fit) cap = std::max(cap, hw); }
            size_t bytes = 84 * (size_t)cap;
            if (bytes > 65536 && !L->lds_attr_fs) { (void)hipFuncSetAttribute((const void*)k_scene_solve, hipFuncAttributeMaxDynamicSharedMemorySize, L->lds_max - 2048); L->lds_attr_fs = true; }
            hipLaunchKernelGGL(k_scene_solve, dim3((unsigned)L->slist.size()), dim3(FS_NT), bytes, L->st, D, (const int32_t*)L->d_slist.p, cap);
        }
    } else {
        // colour
        LAUNCH(k_col_clear, sf_grid((uint64_t)L->nbd * L->K), NT, D);
        for (int r = 0; r < L->R; r++) {
            LAUNCH(k_col_claim, gc, NT, D, r);
            LAUNCH(k_col_resolve, gc, NT, D, r);
        }
        LAUNCH(k_col_count, gc, NT, D);
        LAUNCH(k_col_scan, 1, 64, D);
        LAUNCH(k_col_fill, gc, NT, D);
        // overflow lists
        fill_u32(L, D.ovcnt, 0, L->nbd);
        fill_u32(L, D.ovfill, 0, L->nbd);
        LAUNCH(k_ovf_count, gc, NT, D);
        scan(L, D.ovcnt, D.ovoff, D.ctr + C_NCTR0, L->nbd);
        LAUNCH(k_ovf_fill, gc, NT, D);
        segsort(L, D.ovoff, D.ovcnt, D.ovent, L->nbd);
        LAUNCH(k_ovf_prep, gc, NT, D);
        if (read_ctr(L)) return -1;
        const uint32_t* cc = L->h_ctr + C_COLCNT;
        for (int it = 0; it < 10; it++) {
            LAUNCH(k_chain, gb, NT, D);
            for (int c = 0; c < L->K; c++) if (cc[c]) LAUNCH(k_solve_colour, sf_grid(cc[c]), NT, D, c);
            if (cc[L->K]) {
                LAUNCH(k_ovf_solve, sf_grid(cc[L->K]), NT, D);
                LAUNCH(k_ovf_apply, sf_grid((uint64_t)L->nbd * 64, NT, 65536), NT, D);
            }
        }
    }
    LAUNCH(k_pinch, gb, NT, D);
    LAUNCH(k_impact_flag, gc, NT, D);
    if (read_ctr(L)) return -1;
    if (drain_kills(L)) return -1;
    if (L->h_ctr[C_NIMPACT] && impact_phase(L, L->h_ctr[C_NIMPACT])) return -1;
    LAUNCH(k_integrate, gb, NT, D, h);
    SF_HIP(hipGetLastError());
    return 0;
}

// the scenes' structure sweeps; the largest structure that fits the workgroup's LDS sizes the potential copy
// the largest structure whose potential fits the sweep workgroup's LDS (beside its own arrays); below 0xFFFF, so
// the packed rows' 16-bit indices cover it
uint32_t sor_fit(Lat* L) {
    if (!L->lds_max) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, L->dev) != hipSuccess || v <= 0) v = 65536;
        L->lds_max = v;
    }
    uint32_t fit = (uint32_t)std::min(0xFFFE, std::max(0, (L->lds_max - 1024) / (int)sizeof(float)));
    return L->sor_lds >= 0 ? (uint32_t)std::min<int64_t>(fit, L->sor_lds) : fit;
}
void sor_scene_launch(Lat* L, int mode) {
    uint32_t fit = sor_fit(L), cap = 0;
    for (const HStruct& T : L->sts) { uint32_t n = (uint32_t)(T.hi - T.lo); if (n <= fit) cap = std::max(cap, n); }
    size_t bytes = sizeof(float) * (size_t)cap;
    if (bytes > 65536 && !L->lds_attr) { (void)hipFuncSetAttribute((const void*)k_sor_scene, hipFuncAttributeMaxDynamicSharedMemorySize, L->lds_max - 1024); L->lds_attr = true; }
    hipLaunchKernelGGL(k_sor_scene, dim3((unsigned)std::min<size_t>(L->sc.size(), 2048)), dim3(ST_NT), bytes, L->st, L->D, mode, cap);
}

int structures_upload(Lat* L, bool warm) {
    size_t ns = L->sts.size();
    if (!ns) return 0;
    // one block (sweeps asked for, then the active, on and live flags) from one of two pinned buffers
    const size_t bytes = 7 * ns;
    uint8_t* blk = L->stg_ctl.begin(bytes);
    if (!blk) return -1;
    int32_t* nmin = reinterpret_cast<int32_t*>(blk);
    uint8_t *act = blk + 4 * ns, *on = act + ns, *live = on + ns;
    for (size_t g = 0; g < ns; g++) {
        const HStruct& T = L->sts[g];
        const HScene& S = L->sc[T.scene];
        int sl = (int)g - S.st0;
        act[g] = T.active > 0.0f && T.solved;
        on[g] = T.active > 0.0f;
        live[g] = T.settle <= 0.0f;
        if (warm) nmin[g] = T.warm;
        else if (!act[g]) nmin[g] = 0;
        else {
            int size = T.hi - T.lo;
            nmin[g] = T.resid < 1e-4f ? (((S.frame + (uint32_t)sl) % 3u) == 0 ? 1 : 0) : (size > 15000 ? 1 : 6);
        }
    }
    if (!warm) {
        for (HScene& S : L->sc) {
            int na = 0;
            for (int sl = 0; sl < S.nst; sl++) na += act[S.st0 + sl];
            S.stats[13] = (float)na;
        }
    }
    if (L->stg_ctl.end(bytes, L->st)) return -1;
    LAUNCH(k_st_ctl, sf_grid(ns), NT, L->D, (const uint8_t*)L->stg_ctl.d.p, (uint32_t)ns);
    return 0;
}

int structural(Lat* L, float dt) {
    LD& D = L->D;
    const unsigned gb = sf_grid(D.nub), gd = sf_grid(std::max<uint32_t>(L->ndyn, 1)), gbr = sf_grid(L->nb);
    const unsigned gw = sf_grid((uint64_t)D.nub * 64, NT, 65536);
    auto t = std::chrono::steady_clock::now();
    if (structures_upload(L, false)) return -1;
    L->sp[7] += ms_since(t);
    t = std::chrono::steady_clock::now();
    size_t ns = L->sts.size();
    LAUNCH(k_struct_clear, sf_grid(std::max<size_t>(std::max<size_t>(L->nb, ns), L->sc.size())), NT, D, (uint32_t)ns, (uint32_t)L->sc.size());
    LAUNCH(k_rest_count, gw, NT, D);
    LAUNCH(k_rest_emit, gd, NT, D, 0);
    scan(L, D.lcnt, D.loff, D.ctr + C_NCTR0, L->nb);
    LAUNCH(k_rest_emit, gd, NT, D, 1);
    segsort(L, D.loff, D.lcnt, D.lent, L->nb);
    LAUNCH(k_rest_sum_mix, gbr, NT, D);
    LAUNCH(k_unsupported, gw, NT, D);
    if (ns) {
        hipLaunchKernelGGL(k_stencil, dim3(64, (unsigned)std::min<size_t>(ns, 4096)), dim3(NT), 0, L->st, D, 0, sor_fit(L));
        sor_scene_launch(L, 0);
    }
    LAUNCH(k_ga, gw, NT, D, dt);
    LAUNCH(k_body_sor_prep, gb, NT, D);
    LAUNCH(k_sor_body, std::min<uint32_t>(std::max<uint32_t>(D.nub, 1), 8192), BD_NT, D);
    LAUNCH(k_stress_clear, gbr, NT, D);
    LAUNCH(k_body_judge_prep, gb, NT, D, dt);
    LAUNCH(k_stress_static, gbr, NT, D, dt);
    LAUNCH(k_stress_body, gd, NT, D);
    LAUNCH(k_lt_decay, gb, NT, D, dt);
    SF_HIP(hipGetLastError());
    // results back: per structure resid / max ratio / flags, per scene max ratio, events
    const size_t nsc = L->sc.size();
    L->sp[8] += ms_since(t);
    t = std::chrono::steady_clock::now();
    // packed on the device into one block, one copy into pinned memory (its wait completes every read-back
    // queued before it)
    auto al = [](size_t x) { return (x + 63) & ~(size_t)63; };
    const size_t o_rs = al(4 * (size_t)C_NCTR), o_mr = o_rs + al(4 * ns), o_fl = o_mr + al(4 * ns), o_gm = o_fl + al(4 * ns);
    const size_t o_bd = o_gm + al(4 * nsc), o_ek = al(o_bd + D.nub), o_ev = o_ek + al(8 * (size_t)L->kev_pack);
    const size_t o_bi = o_ev + al(16 * (size_t)L->kev_pack), o_bv = o_bi + al(4 * (size_t)L->kbw_pack);
    const size_t total = o_bv + 4 * (size_t)L->kbw_pack;
    if (L->h_rb.reserve(total + 64)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    if (L->d_rbk.n < total && L->d_rbk.alloc(std::max<size_t>(total + total / 2, 4096), L->st)) return -1;
    float* resid = (float*)(L->h_rb.p + o_rs);
    uint32_t* maxr = (uint32_t*)(L->h_rb.p + o_mr);
    uint32_t* flags = (uint32_t*)(L->h_rb.p + o_fl);
    uint32_t* gmax = (uint32_t*)(L->h_rb.p + o_gm);
    uint8_t* bdirty = L->h_rb.p + o_bd;
    int r = 0;
    PackRB P{(uint32_t)ns, (uint32_t)nsc, (uint32_t)o_rs, (uint32_t)o_mr, (uint32_t)o_fl, (uint32_t)o_gm, (uint32_t)o_bd,
             (uint32_t)o_ek, (uint32_t)o_ev, (uint32_t)o_bi, (uint32_t)o_bv, L->kev_pack, L->kbw_pack};
    const size_t npk = std::max(std::max<size_t>(std::max<size_t>(D.nub, ns), std::max<size_t>(nsc, C_NCTR)),
                                (size_t)std::max(L->kev_pack, L->kbw_pack));
    LAUNCH(k_pack_rb, sf_grid(npk), NT, D, L->d_rbk.p, P);
    SF_HIP(hipMemcpyAsync(L->h_rb.p, L->d_rbk.p, total, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipStreamSynchronize(L->st));
    L->syncs++;
    memcpy(L->h_ctr, L->h_rb.p, sizeof(uint32_t) * C_NCTR);
    if (finish_stats(L)) return -1;
    if (drain_kills(L)) return -1;
    uint32_t nev = std::min(L->h_ctr[C_NEV], D.evcap), nbw = L->h_ctr[C_NBW];
    bool bw_overflow = nbw > D.bwcap;
    // events and bond writes: from the packed block when they fit its share, else copied now
    const unsigned long long* ekey = reinterpret_cast<const unsigned long long*>(L->h_rb.p + o_ek);
    const int4* ev = reinterpret_cast<const int4*>(L->h_rb.p + o_ev);
    const uint32_t* bwi = reinterpret_cast<const uint32_t*>(L->h_rb.p + o_bi);
    const float* bwv = reinterpret_cast<const float*>(L->h_rb.p + o_bv);
    std::vector<unsigned long long> ekey_x;
    std::vector<int4> ev_x;
    std::vector<uint32_t> bwi_x;
    std::vector<float> bwv_x;
    bool more = false;
    if (nev > L->kev_pack) {
        ekey_x.resize(nev); ev_x.resize(nev);
        r |= cp_down(ekey_x.data(), D.evkey, nev, L->st); r |= cp_down(ev_x.data(), D.ev, nev, L->st);
        ekey = ekey_x.data(); ev = ev_x.data(); more = true;
    }
    if (nbw && !bw_overflow && nbw > L->kbw_pack) {
        bwi_x.resize(nbw); bwv_x.resize(nbw);
        r |= cp_down(bwi_x.data(), D.bwi, nbw, L->st); r |= cp_down(bwv_x.data(), D.bwv, nbw, L->st);
        bwi = bwi_x.data(); bwv = bwv_x.data(); more = true;
    }
    if (r) return r;
    if (more) { SF_HIP(hipStreamSynchronize(L->st)); L->syncs++; }
    // the next frame's shares follow this frame's counts
    L->kev_pack = std::min<uint32_t>(std::max<uint32_t>(2 * nev + 64, 64u), 16384u);
    L->kbw_pack = std::min<uint32_t>(std::max<uint32_t>(2 * std::min(nbw, D.bwcap) + 256, 256u), 65536u);
    L->sp[9] += ms_since(t);
    t = std::chrono::steady_clock::now();
    if (nev > 0 && L->h_ctr[C_NEV] > D.evcap) { sf::set_error(SF_E_CAP, "structural event buffer overflow"); return SF_E_CAP; }
    if (bw_overflow) { for (size_t si = 0; si < L->sc.size(); si++) if (pull(L, (int)si, G_BOND)) return -1; }
    else for (uint32_t e = 0; e < nbw; e++) bond_ref(L, (int)(bwi[e] >> 2), (int)(bwi[e] & 3u)) = bwv[e];
    for (uint32_t ub = 0; ub < D.nub; ub++) if (bdirty[ub]) L->dirty[L->ubody[ub]] = 1;
    std::vector<uint32_t> order(nev);
    for (uint32_t e = 0; e < nev; e++) order[e] = e;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return ekey[a] < ekey[b]; });
    std::vector<uint8_t> changed(L->sc.size(), 0);
    auto apply_event = [&](const int4& e) {
        int i = e.x, a = e.y, k = e.z;
        int si = L->bxyz[i].w;
        HScene& S = L->sc[si];
        changed[si] = 1;
        S.crushed_joints++;
        if (S.crush_n < 300) { S.crush_n++; remove_brick(L, S, k); return; }
        bond_ref(L, i, a) = 0.0f;
        L->touch(i);
        if (L->body[i] >= 0) L->dirty[L->body[i]] = 1;
        else { int s = L->sid[i]; if (s < S.nst) { HStruct& T = L->sts[S.st0 + s]; T.conn = true; S.conn_any = true; T.broke = true; } }
    };
    // structures in order: their crushes, then their timers; then the pieces' crushes
    uint32_t e = 0;
    for (size_t g = 0; g < ns; g++) {
        HStruct& T = L->sts[g];
        HScene& S = L->sc[T.scene];
        T.resid = resid[g];
        if (!(T.active > 0.0f && T.solved)) continue;
        if (flags[g] & 1u) { T.conn = true; S.conn_any = true; }
        if (flags[g] & 2u) T.broke = true;
        while (e < nev && !(ekey[order[e]] >> 63) && (int)(ekey[order[e]] >> 40) == (int)g) apply_event(ev[order[e++]]);
        float mr; memcpy(&mr, &maxr[g], 4);
        T.maxr = mr;
        if (T.settle > 0.0f && T.warm == 0) T.settle -= dt;
        if (T.warm == 0) T.active -= dt;
        if (T.broke) { T.broke = false; T.active = 4.0f; }
        if (mr > 0.98f && T.active < 1.0f) T.active = 1.0f;
    }
    for (; e < nev; e++) apply_event(ev[order[e]]);
    for (size_t si = 0; si < L->sc.size(); si++) {
        HScene& S = L->sc[si];
        memcpy(&S.stats[6], &gmax[si], 4);
    }
    (void)changed;
    L->sp[10] += ms_since(t);
    t = std::chrono::steady_clock::now();
    r = flush_delta(L);
    L->sp[11] += ms_since(t);
    return r;
}
}  // namespace

namespace {

// the frame statistics (ptr_stats, src/lib.rs:3718-3720): the host's fields now; the device counts (dynamic bricks,
// bodies, awake bodies) are copied back asynchronously and taken at the next wait (finish_stats)
int scene_stats(Lat* L) {
    size_t S = L->sc.size();
    if (finish_stats(L)) return -1;
    fill_u32(L, L->d_scstat.p, 0, 3 * S);
    LAUNCH(k_scene_stats, sf_grid(L->D.nub), NT, L->D, L->d_scstat.p);
    if (L->h_scstat.reserve(3 * S)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    if (!L->ev_stat) SF_HIP(hipEventCreateWithFlags(&L->ev_stat, hipEventDisableTiming));
    SF_HIP(hipMemcpyAsync(L->h_scstat.p, L->d_scstat.p, sizeof(uint32_t) * 3 * S, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipEventRecord(L->ev_stat, L->st));
    L->stat_pending = true;
    L->stat_n = S;
    for (size_t si = 0; si < S; si++) {
        HScene& sc = L->sc[si];
        uint32_t stat = 0, unsolved = 0;
        for (int sl = 0; sl < sc.nst; sl++) { const HStruct& T = L->sts[sc.st0 + sl]; stat += T.now; if (!T.solved) unsolved++; }
        sc.stats[0] = (float)stat; sc.stats[4] = (float)sc.ncs; sc.stats[14] = (float)unsolved; sc.stats[15] = (float)sc.nst;
    }
    return 0;
}
int finish_stats(Lat* L) {
    if (!L->stat_pending) return 0;
    L->stat_pending = false;
    SF_HIP(hipEventSynchronize(L->ev_stat));
    const uint32_t* o = L->h_scstat.p;
    for (size_t si = 0; si < L->stat_n && si < L->sc.size(); si++) {
        HScene& sc = L->sc[si];
        sc.stats[1] = (float)o[3 * si + 1]; sc.stats[2] = (float)o[3 * si]; sc.stats[3] = (float)o[3 * si + 2];
    }
    return 0;
}

// the world's worker pool, made on first use: SF_HOST_THREADS threads in all (default: the cores, at most 16)
Pool* host_pool(Lat* L) {
    if (!L->pool) {
        int n = (int)std::thread::hardware_concurrency();
        if (const char* e = getenv("SF_HOST_THREADS")) n = atoi(e);
        if (L->host_threads > 0) n = L->host_threads;
        n = std::max(1, std::min(n, 16));
        if (n < 2) return nullptr;
        L->pool.reset(new Pool(n - 1));
        L->scr.clear();
        for (int k = 0; k < n; k++) L->scr.emplace_back(new Scratch());
    }
    return L->pool.get();
}
// the workers' touched bricks and bodies join the world's lists
void merge_touched(Lat* L) {
    for (auto& w : L->scr) {
        L->tlist.insert(L->tlist.end(), w->tlist.begin(), w->tlist.end()); w->tlist.clear();
        L->btlist.insert(L->btlist.end(), w->btlist.begin(), w->btlist.end()); w->btlist.clear();
    }
}

// The device's searches of a topology pass, in one upload, one copy back and one wait: the parts of every body to
// rebuild (k_split_*, over the bricks rebuild_list listed) and the grounded bricks of every structure the host will
// check for detached parts (k_cc_*). The host's changes so far go down first. With the grounded search off
// (conn_device false) the host searches the structures itself (conn_struct).
int device_search(Lat* L, const std::vector<int>& need) {
    L->reach_idx.clear();
    L->reach_n = 0;
    std::vector<int2> list;
    L->reach_woff.clear();
    uint32_t words = 0;
    int maxn = 0;
    size_t nsplit = 0;
    for (int si : need) {
        const HScene& S = L->sc[si];
        nsplit += S.rlist.size();
        if (!L->conn_device || !S.conn_any) continue;
        for (int sl = 0; sl < S.nst; sl++) {
            const HStruct& T = L->sts[S.st0 + sl];
            if (!T.conn || T.hi <= T.lo || list.size() >= 65535) continue;  // (past a grid's height: the host searches)
            if (L->reach_idx.empty()) L->reach_idx.assign(L->sts.size(), -1);
            L->reach_idx[S.st0 + sl] = (int)list.size();
            list.push_back(make_int2(S.st0 + sl, sl));
            L->reach_woff.push_back(words);
            words += (uint32_t)((T.hi - T.lo + 31) / 32);
            maxn = std::max(maxn, T.hi - T.lo);
        }
    }
    if (list.empty() && !nsplit) return 0;
    if (flush_delta(L)) return -1;
    const size_t nl = list.size(), bytes = sizeof(int2) * nl + sizeof(uint32_t) * (nl + nsplit);
    uint8_t* hs = L->stg_conn.begin(bytes);
    if (!hs) return -1;
    std::memcpy(hs, list.data(), sizeof(int2) * nl);
    std::memcpy(hs + sizeof(int2) * nl, L->reach_woff.data(), sizeof(uint32_t) * nl);
    uint32_t* hsplit = reinterpret_cast<uint32_t*>(hs + sizeof(int2) * nl + sizeof(uint32_t) * nl);
    for (int si : need) {
        const HScene& S = L->sc[si];
        std::memcpy(hsplit, S.rlist.data(), sizeof(uint32_t) * S.rlist.size());
        hsplit += S.rlist.size();
    }
    if (L->stg_conn.end(bytes, L->st)) return -1;
    // out: per listed structure its status, then the structures' bits, then the split roots
    const size_t outn = nl + words + nsplit;
    if (L->d_creach_out.n < outn && L->d_creach_out.alloc(outn + outn / 2, L->st)) return -1;
    if (L->h_creach_out.reserve(outn)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
    const int2* dl = reinterpret_cast<const int2*>(L->stg_conn.d.p);
    const uint32_t* dw = reinterpret_cast<const uint32_t*>(L->stg_conn.d.p + sizeof(int2) * nl);
    const uint32_t* ds = dw + nl;
    if (nsplit) {
        const unsigned g = sf_grid(nsplit);
        LAUNCH(k_split_init, g, NT, L->D, ds, (uint32_t)nsplit);
        LAUNCH(k_split_hook, g, NT, L->D, ds, (uint32_t)nsplit);
        LAUNCH(k_split_root, g, NT, L->D, ds, (uint32_t)nsplit, L->d_creach_out.p + nl + words);
    }
    if (nl) {
        const uint32_t token = ++L->ctoken;
        fill_u32(L, L->d_creach_out.p, 0u, nl);
        const dim3 grid((unsigned)std::min<int>((maxn + NT - 1) / NT, 64), (unsigned)nl);
        hipLaunchKernelGGL(k_cc_init, grid, dim3(NT), 0, L->st, L->D, dl);
        hipLaunchKernelGGL(k_cc_hook, grid, dim3(NT), 0, L->st, L->D, dl, L->d_creach_out.p);
        hipLaunchKernelGGL(k_cc_root, grid, dim3(NT), 0, L->st, L->D, dl, token);
        hipLaunchKernelGGL(k_cc_bits, grid, dim3(NT), 0, L->st, L->D, dl, dw, (uint32_t)nl, L->d_creach_out.p, token);
    }
    SF_HIP(hipGetLastError());
    SF_HIP(hipMemcpyAsync(L->h_creach_out.p, L->d_creach_out.p, sizeof(uint32_t) * outn, hipMemcpyDeviceToHost, L->st));
    SF_HIP(hipStreamSynchronize(L->st));
    L->syncs++;
    L->reach_n = nl;
    L->split_off = nl + words;
    return 0;
}

int host_topology(Lat* L, bool rebuild) {
    struct Acc { Lat* L; std::chrono::steady_clock::time_point t; ~Acc() { L->ph[3] += ms_since(t); } } acc{L, std::chrono::steady_clock::now()};
    std::vector<int> need;
    for (size_t si = 0; si < L->sc.size(); si++) {
        HScene& S = L->sc[si];
        bool nd = S.conn_any;
        if (rebuild) for (int lb = 0; lb < S.hw && !nd; lb++) { int b = S.bd0 + lb; nd = L->alive[b] && L->dirty[b]; }
        if (nd) need.push_back((int)si);
    }
    auto t = std::chrono::steady_clock::now();
    if (pull_bodies(L, need)) return -1;
    L->sp[0] += ms_since(t);
    t = std::chrono::steady_clock::now();
    // per scene: list the bodies to rebuild; the device then finds their parts and the structures' grounded bricks; per
    // scene the split bodies and the detached parts are decided in the reference's order and set up on the device;
    // last, per scene, the detached parts are blasted and their neighbours woken
    auto list_work = [&](int si) {
        HScene& S = L->sc[si];
        if (rebuild) for (int lb = 0; lb < S.hw; lb++) { int b = S.bd0 + lb; if (L->alive[b] && L->dirty[b]) rebuild_list(L, S, b); }
    };
    auto split_work = [&](int si) {
        HScene& S = L->sc[si];
        const uint32_t* roots = L->h_creach_out.p + L->split_off + L->sc_split[si];
        for (const int3& r : S.rbody) rebuild_split(L, S, r.x, S.rlist.data() + r.y, roots + r.y, r.z);
        S.rlist.clear(); S.rbody.clear();
    };
    auto conn_work = [&](int si) { connectivity(L, L->sc[si]); };
    Pool* pool = need.size() >= 2 ? host_pool(L) : nullptr;
    auto run = [&](const std::function<void(int)>& fn) {
        if (pool) {
            size_t maxb = 0;
            for (int si : need) maxb = std::max<size_t>(maxb, (size_t)L->sc[si].bcap);
            for (auto& w : L->scr) w->reserve(maxb);
            pool->run((int)need.size(), [&](int k, int slot) {
                t_scr = L->scr[slot].get();
                fn(need[k]);
                t_scr = nullptr;
            });
            merge_touched(L);
        } else {
            for (int si : need) fn(si);
        }
    };
    if (rebuild) run(list_work);
    L->sc_split.assign(L->sc.size(), 0);
    {
        size_t o = 0;
        for (int si : need) { L->sc_split[si] = o; o += L->sc[si].rlist.size(); }
    }
    L->sp[1] += ms_since(t);
    t = std::chrono::steady_clock::now();
    if (device_search(L, need)) return -1;
    L->sp[12] += ms_since(t);
    t = std::chrono::steady_clock::now();
    run([&](int si) { split_work(si); conn_work(si); });
    L->reach_idx.clear();
    if (run_setups(L, need)) return -1;
    run([&](int si) { conn_finish(L, L->sc[si]); });
    L->sp[2] += ms_since(t);
    t = std::chrono::steady_clock::now();
    if (flush_delta(L)) return -1;
    L->sp[3] += ms_since(t);
    t = std::chrono::steady_clock::now();
    if (rebuild_csr(L)) return -1;
    L->sp[4] += ms_since(t);
    t = std::chrono::steady_clock::now();
    int r = upload_scenes(L);
    L->sp[5] += ms_since(t);
    L->sp[6] += (double)need.size();
    return r;
}

int frame(Lat* L, float dt) {
    auto t0 = std::chrono::steady_clock::now();
    L->syncs = 0;
    for (double& x : L->ph) x = 0.0;
    for (double& x : L->sp) x = 0.0;
    dt = clampf(dt, 0.0f, 1.0f / 30.0f);
    for (HScene& S : L->sc) {
        S.time += dt; S.crush_n = 0; S.frame += 1;
        if (S.blast_t > 0.0f) S.blast_t -= dt;
    }
    if (rebuild_csr(L) || upload_scenes(L)) return -1;
    auto ts = std::chrono::steady_clock::now();
    for (int k = 0; k < 2; k++) if (substep(L, dt * 0.5f)) return -1;
    L->ph[0] = ms_since(ts) - L->ph[1] - L->ph[2];
    if (read_ctr(L) || drain_kills(L)) return -1;
    if (host_topology(L, true)) return -1;
    LD& D = L->D;
    const unsigned gb = sf_grid(D.nub), gd = sf_grid(std::max<uint32_t>(L->ndyn, 1));
    LAUNCH(k_cache, gb, NT, D, 0, 0.0f);
    LAUNCH(k_wpos, gd, NT, D);
    ts = std::chrono::steady_clock::now();
    if (structural(L, dt)) return -1;
    L->ph[4] = ms_since(ts);
    ts = std::chrono::steady_clock::now();
    // warm-up sweeps of structures asked for a load field
    bool any_warm = false;
    for (const HStruct& T : L->sts) if (T.warm > 0) { any_warm = true; break; }
    if (any_warm) {
        if (structures_upload(L, true)) return -1;
        hipLaunchKernelGGL(k_stencil, dim3(64, (unsigned)std::min<size_t>(L->sts.size(), 4096)), dim3(NT), 0, L->st, D, 1, sor_fit(L));
        sor_scene_launch(L, 1);
        std::vector<float> resid(L->sts.size());
        std::vector<int32_t> left(L->sts.size());
        if (cp_down(resid.data(), D.st_resid, resid.size(), L->st) || cp_down(left.data(), D.st_nmin, left.size(), L->st)) return -1;
        SF_HIP(hipStreamSynchronize(L->st));
        for (size_t g = 0; g < L->sts.size(); g++) {
            HStruct& T = L->sts[g];
            if (T.warm <= 0) continue;
            T.resid = resid[g];
            T.warm = left[g];
            if (T.warm == 0) T.settle = 0.3f;
        }
    }
    L->ph[5] = ms_since(ts);
    if (host_topology(L, false)) return -1;
    LAUNCH(k_cache, gb, NT, D, 0, 0.0f);
    LAUNCH(k_wpos, sf_grid(std::max<uint32_t>(L->ndyn, 1)), NT, D);
    ts = std::chrono::steady_clock::now();
    if (scene_stats(L)) return -1;
    L->ph[6] = ms_since(ts);
    SF_HIP(hipGetLastError());
    L->ms_frame = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    L->ph[7] = L->ms_frame;
    return 0;
}

// ---------------------------------------------------------------- scene set-up
int add_brick(Lat* L, HScene& S, int x, int y, int z, int m, int sid, int infill, int si) {
    if (S.nb >= S.bcap) { sf::set_error(SF_E_CAP, "scene brick capacity exceeded"); return -1; }
    int i = S.b0 + S.nb++;
    L->mat[i] = (uint8_t)m; L->bxyz[i] = make_int4(x, y, z, si);
    L->sid[i] = (uint8_t)sid; L->infill[i] = (uint8_t)((m == GLASS || infill) ? 1 : 0);
    L->body[i] = -1; L->next[i] = NONE;
    L->local[i] = hf(lat_pos(x, y, z)); L->wpos[i] = L->local[i];
    L->bond[i] = make_float4(0, 0, 0, 0); L->bmax[i] = make_float4(0, 0, 0, 0);
    L->phi[i] = 0.0f; L->ext[i] = 0.0f; L->extn[i] = 0.0f;
    return i;
}
int finalize(Lat* L, int si) {
    HScene& S = L->sc[si];
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int i = S.b0; i < S.b0 + S.nb; i++) {
        int c[3] = {L->bxyz[i].x, L->bxyz[i].y, L->bxyz[i].z};
        for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], c[k]); hi[k] = std::max(hi[k], c[k]); }
    }
    if (S.nb == 0) { lo[0] = lo[1] = lo[2] = 0; hi[0] = hi[1] = hi[2] = 0; }
    S.gx0 = lo[0] - 1; S.gy0 = lo[1] - 1; S.gz0 = lo[2] - 1;
    S.gnx = hi[0] - lo[0] + 3; S.gny = hi[1] - lo[1] + 3; S.gnz = hi[2] - lo[2] + 3;
    size_t gsz = (size_t)S.gnx * S.gny * S.gnz;
    if (gsz > S.gcap) { S.goff = L->grid.size(); S.gcap = gsz; L->grid.resize(S.goff + gsz, 0); }
    std::fill(L->grid.begin() + S.goff, L->grid.begin() + S.goff + S.gcap, 0);
    // one brick per lattice cell: the neighbour table, the bonds and the sweeps' packed rows (16-bit indices within
    // a structure) all rest on it
    for (int i = S.b0; i < S.b0 + S.nb; i++) {
        if (!L->mat[i]) continue;
        int4 c = L->bxyz[i];
        size_t idx = S.goff + ((size_t)(c.y - S.gy0) * S.gnz + (c.z - S.gz0)) * S.gnx + (c.x - S.gx0);
        if (L->grid[idx]) {
            sf::set_error(SF_E_ARG, "sf_lat_finalize: two bricks in lattice cell (" + std::to_string(c.x) + ", " +
                                        std::to_string(c.y) + ", " + std::to_string(c.z) + ")");
            return SF_E_ARG;
        }
        L->grid[idx] = i + 1;
    }
    int red_at = S.b0;
    std::vector<int32_t> stsc(S.nst), stlo(S.nst), sthi(S.nst);
    std::vector<int4> strb(S.nst);
    for (int sl = 0; sl < S.nst; sl++) {
        HStruct& T = L->sts[S.st0 + sl];
        uint32_t cnt = 0; int top = 0;
        int mn[3] = {1 << 30, 1 << 30, 1 << 30}, mx[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
        int r0 = red_at;
        for (int par = 0; par < 2; par++) {
            int p0 = red_at;
            for (int i = T.lo; i < T.hi; i++) {
                if (!hin_struct(L, i, sl)) continue;
                int4 c = L->bxyz[i];
                if (((c.x + c.y + c.z) & 1) != par) continue;
                L->lst[red_at - S.b0] = (uint32_t)i;
                red_at++;
            }
            if (par == 0) { T.rb.x = p0; T.rb.y = red_at - p0; } else { T.rb.z = p0; T.rb.w = red_at - p0; }
        }
        (void)r0;
        for (int i = T.lo; i < T.hi; i++) {
            if (!hin_struct(L, i, sl)) continue;
            cnt++; top = std::max(top, L->bxyz[i].y + 1);
            int c[3] = {L->bxyz[i].x, L->bxyz[i].y, L->bxyz[i].z};
            for (int k = 0; k < 3; k++) { mn[k] = std::min(mn[k], c[k]); mx[k] = std::max(mx[k], c[k]); }
        }
        T.init = T.now = cnt; T.top0 = T.top = top;
        for (int k = 0; k < 3; k++) { T.mn[k] = mn[k]; T.mx[k] = mx[k]; }
        stsc[sl] = si; stlo[sl] = T.lo; sthi[sl] = T.hi; strb[sl] = T.rb;
    }
    LD& D = L->D;
    hipStream_t s = L->st;
    size_t b0 = S.b0, nb = S.nb;
    int r = 0;
    r |= cp_up(D.mat + b0, &L->mat[b0], nb, s); r |= cp_up(D.sid + b0, &L->sid[b0], nb, s); r |= cp_up(D.body + b0, &L->body[b0], nb, s);
    r |= cp_up(D.bxyz + b0, &L->bxyz[b0], nb, s); r |= cp_up(D.infill + b0, &L->infill[b0], nb, s);
    r |= cp_up(D.local + b0, &L->local[b0], nb, s); r |= cp_up(D.wpos + b0, &L->wpos[b0], nb, s);
    r |= cp_up(D.phi + b0, &L->phi[b0], nb, s); r |= cp_up(D.ext + b0, &L->ext[b0], nb, s); r |= cp_up(D.extn + b0, &L->extn[b0], nb, s);
    std::vector<uint32_t> zero(nb, 0u);
    r |= cp_up(D.sup + b0, zero.data(), nb, s);
    r |= cp_up((uint32_t*)D.stress + b0, zero.data(), nb, s);
    if (red_at > S.b0) r |= cp_up(D.st_red + S.b0, L->lst.data(), (size_t)(red_at - S.b0), s);
    if (S.nst) {
        r |= cp_up(D.st_scene + S.st0, stsc.data(), S.nst, s); r |= cp_up(D.st_lo + S.st0, stlo.data(), S.nst, s);
        r |= cp_up(D.st_hi + S.st0, sthi.data(), S.nst, s); r |= cp_up(D.st_rb + S.st0, strb.data(), S.nst, s);
        std::vector<float> res(S.nst, 0.0f);
        r |= cp_up(D.st_resid + S.st0, res.data(), S.nst, s);
    }
    if (r) return -1;
    SF_HIP(hipStreamSynchronize(s));
    // the grid: all scenes' regions
    if (L->d_grid.n < L->grid.size()) {
        if (L->d_grid.alloc(L->grid.size() + L->grid.size() / 2, L->st)) return -1;
        r |= cp_up(L->d_grid.p, L->grid.data(), L->grid.size(), s);
    } else {
        r |= cp_up(L->d_grid.p + S.goff, L->grid.data() + S.goff, S.gcap, s);
    }
    SF_HIP(hipStreamSynchronize(s));
    D.grid = L->d_grid.p;
    // neighbours and bonds on the device, then mirrored to the host for the fracture bookkeeping
    if (nb) {
        const GridK G{S.gx0, S.gy0, S.gz0, S.gnx, S.gny, S.gnz, (uint32_t)S.goff};
        LAUNCH(k_finalize, sf_grid(nb), NT, D, G, (uint32_t)b0, (uint32_t)nb);
        r |= cp_down(&L->nb6[6 * b0], D.nb6 + 6 * b0, 6 * nb, s);
        r |= cp_down(&L->bond[b0], D.bond + b0, nb, s); r |= cp_down(&L->bmax[b0], D.bmax + b0, nb, s);
        if (r) return -1;
        SF_HIP(hipStreamSynchronize(s));
    }
    std::vector<int32_t> bsc(S.dcap, si);
    r |= cp_up(D.bd_scene + S.bd0, bsc.data(), S.dcap, s);
    if (r) return -1;
    if (push(L, si, G_BODY)) return -1;
    // everything the host set up is on the device now: nothing left to flush
    for (uint32_t i : L->tlist) L->touched[i] = 0;
    L->tlist.clear();
    for (uint32_t b : L->btlist) L->btouched[b] = 0;
    L->btlist.clear();
    S.finalized = true;
    L->lists_dirty = true;
    if (rebuild_csr(L) || upload_scenes(L)) return -1;
    return 0;
}

// ---------------------------------------------------------------- blast (src/lib.rs:2679-2716)
void blast_brick(Lat* L, HScene& S, int i, LV c, float r, float power) {
    int m = L->mat[i];
    if (m == 0 || m >= SLUG) return;
    LV p = hbpos(L, i);
    float d = len(sub(p, c));
    float rr = r * tables().blast[m];
    if (d < rr * 0.5f) remove_brick(L, S, i);
    else if (d < rr * 1.35f) { float f = 1.0f - d / (rr * 1.35f); damage_brick(L, S, i, power * f * f * 1.3f, 0.0f); }
}
// A charge (explode_at, src/lib.rs:2697-2716), its bricks and bodies on the device (k_blast_*): the host lists what
// the charge may reach in the reference's order, one launch per stage, one copy back into the host mirror, one wait;
// the structures and bodies it disturbed are then marked in that order (for connectivity, the structural solve and
// the next rebuild).
int explode(Lat* L, int si, float x, float y, float z, float r, float power) {
    HScene& S = L->sc[si];
    if (pull_bodies(L, si) || flush_delta(L)) return -1;
    LV c = lv(x, y, z);
    S.blast_c = c; S.blast_r = r * 2.2f; S.blast_p = power; S.blast_t = 0.25f;
    float reach = r * 1.4f * 1.35f + 0.5f;
    std::vector<int2> ent;      // (brick, structure slot; -1: a body's brick) in the order the reference visits them
    for (int sl = 0; sl < S.nst; sl++) {
        HStruct& T = L->sts[S.st0 + sl];
        LV a = sub(c, lv(reach, reach, reach)), b = add(c, lv(reach, reach, reach));
        int x0 = fcell(a.x) + NX2, y0 = fcell(a.y), z0 = fcell(a.z) + NX2;
        int x1 = fcell(b.x) + NX2, y1 = fcell(b.y), z1 = fcell(b.z) + NX2;
        if (x1 < T.mn[0] || x0 > T.mx[0] || y1 < T.mn[1] || y0 > T.mx[1] || z1 < T.mn[2] || z0 > T.mx[2]) continue;
        for (int i = T.lo; i < T.hi; i++) ent.push_back(make_int2(i, sl));
    }
    for (int lb = 0; lb < S.hw; lb++) {
        int b = S.bd0 + lb;
        if (!(L->alive[b] && len(sub(hv(L->pos[b]), c)) < reach + L->rad[b])) continue;
        for (uint32_t k = L->first[b]; k != NONE; k = L->next[k]) ent.push_back(make_int2((int)k, -1));
    }
    std::vector<int4> bl;       // bodies the blast pushes: slot, their bricks [y, z) of `chain`
    std::vector<uint32_t> chain;
    for (int lb = 0; lb < S.hw; lb++) {
        int b = S.bd0 + lb;
        if (!(L->alive[b] && len(sub(hv(L->pos[b]), c)) < S.blast_r + L->rad[b])) continue;
        const int s0 = (int)chain.size();
        for (uint32_t k = L->first[b]; k != NONE; k = L->next[k]) chain.push_back(k);
        bl.push_back(make_int4(b, s0, (int)chain.size(), 0));
    }
    const size_t n = ent.size(), nbl = bl.size(), nch = chain.size();
    if (n || nbl) {
        const size_t bytes = sizeof(int4) * nbl + sizeof(int2) * n + sizeof(uint32_t) * nch;
        uint8_t* hs = L->stg_conn.begin(bytes);
        if (!hs) return -1;
        std::memcpy(hs, bl.data(), sizeof(int4) * nbl);
        std::memcpy(hs + sizeof(int4) * nbl, ent.data(), sizeof(int2) * n);
        std::memcpy(hs + sizeof(int4) * nbl + sizeof(int2) * n, chain.data(), sizeof(uint32_t) * nch);
        if (L->stg_conn.end(bytes, L->st)) return -1;
        // device out: bonds (float4 per entry), the bodies' velocities and spins, events (u32 per entry), fractions
        const size_t outb = sizeof(float4) * (n + 2 * nbl) + sizeof(uint32_t) * n, devb = outb + sizeof(float) * n;
        if (L->d_setup_out.n < devb && L->d_setup_out.alloc(devb + devb / 2, L->st)) return -1;
        if (L->h_setup_out.reserve(outb)) { sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return -1; }
        float4* dbond = reinterpret_cast<float4*>(L->d_setup_out.p);
        float4* dvw = dbond + n;
        uint32_t* dev = reinterpret_cast<uint32_t*>(dvw + 2 * nbl);
        float* dfr = reinterpret_cast<float*>(dev + n);
        const uint8_t* d = L->stg_conn.d.p;
        const int4* dbl = reinterpret_cast<const int4*>(d);
        const int2* dent = reinterpret_cast<const int2*>(d + sizeof(int4) * nbl);
        const uint32_t* dch = reinterpret_cast<const uint32_t*>(d + sizeof(int4) * nbl + sizeof(int2) * n);
        BlastParams P;
        std::memset(&P, 0, sizeof(P));
        P.cx = x; P.cy = y; P.cz = z; P.r = r; P.power = power; P.br = S.blast_r; P.bp = S.blast_p;
        P.token = ++L->ctoken; P.n = (uint32_t)n;
        if (n) {
            const unsigned g = sf_grid(n);
            LAUNCH(k_blast_event, g, NT, L->D, P, dent, dev, dfr);
            LAUNCH(k_blast_bonds, g, NT, L->D, P, dent, dev, (const float*)dfr, dbond);
            LAUNCH(k_blast_remove, g, NT, L->D, P, dent, (const uint32_t*)dev);
        }
        if (nbl) LAUNCH(k_blast_bodies, (unsigned)nbl, 64, L->D, P, dbl, dch, dvw);
        SF_HIP(hipGetLastError());
        SF_HIP(hipMemcpyAsync(L->h_setup_out.p, L->d_setup_out.p, outb, hipMemcpyDeviceToHost, L->st));
        SF_HIP(hipStreamSynchronize(L->st));
        L->syncs++;
        const float4* hbond = reinterpret_cast<const float4*>(L->h_setup_out.p);
        const float4* hvw = hbond + n;
        const uint32_t* hev = reinterpret_cast<const uint32_t*>(hvw + 2 * nbl);
        // the mirror: bonds and removed bricks; then what each event disturbed, in order (blast_brick, remove_brick,
        // damage_brick: the structure to check and to solve again, the body to rebuild)
        for (size_t t = 0; t < n; t++) L->bond[ent[t].x] = hbond[t];
        for (size_t t = 0; t < n; t++) {
            const int i = ent[t].x;
            const uint32_t e = hev[t] & 3u;
            const bool broke = (hev[t] & BL_BROKE) != 0;
            if (e == BL_NONE) continue;
            if (e == BL_REMOVE) L->mat[i] = 0;
            if (L->body[i] < 0) {
                const int sl = L->sid[i];
                if (sl < S.nst && (e == BL_REMOVE || broke)) {
                    HStruct& T = L->sts[S.st0 + sl];
                    T.conn = true; S.conn_any = true;
                    if (e == BL_DAMAGE) T.broke = true;
                }
                st_static_change(L, S, i);
            } else if (e == BL_REMOVE || broke) L->dirty[L->body[i]] = 1;
        }
        for (size_t k = 0; k < nbl; k++) {
            const int b = bl[k].x;
            L->v[b] = hvw[2 * k]; L->w[b] = hvw[2 * k + 1];
            L->asleep[b] = 0; L->sleep[b] = 0.0f;
        }
    }
    if (flush_delta(L)) return -1;
    return upload_scenes(L);
}

int add_body(Lat* L, int si, int n, const int32_t* xyz, const uint8_t* mat, const float* v3, const float* w3) {
    HScene& S = L->sc[si];
    if (n <= 0) return SF_E_ARG;
    int b = alloc_body(L, S);
    if (b < 0) { sf::set_error(SF_E_CAP, "scene body capacity exceeded"); return SF_E_CAP; }
    if (S.nb + n > S.bcap) { sf::set_error(SF_E_CAP, "scene brick capacity exceeded"); return SF_E_CAP; }
    for (int k = 0; k < n; k++) {
        int i = add_brick(L, S, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], mat[k], 255, 0, si);
        L->lst2[k] = (uint32_t)i;
    }
    LQ idq; idq.x = 0.0f; idq.y = 0.0f; idq.z = 0.0f; idq.w = 1.0f;
    setup_queue(L, S, b, L->lst2.data(), n, true, lv(0.0f, 0.0f, 0.0f), IDM, idq, lv(v3[0], v3[1], v3[2]), lv(w3[0], w3[1], w3[2]), false);
    if (run_setups(L, std::vector<int>{si})) return -1;
    return b - S.bd0;
}

// First-fit placement of a scene's brick or body range in the handle's pool of free ranges (sorted by start) and
// the return of a range, merged with its free neighbours. The handle's arrays are sized once; scenes come and go.
int pool_take(std::vector<std::pair<int, int>>& fl, int n) {
    for (size_t k = 0; k < fl.size(); k++) {
        if (fl[k].second < n) continue;
        const int s = fl[k].first;
        fl[k].first += n; fl[k].second -= n;
        if (!fl[k].second) fl.erase(fl.begin() + (long)k);
        return s;
    }
    return -1;
}
void pool_give(std::vector<std::pair<int, int>>& fl, int s, int n) {
    if (n <= 0) return;
    auto it = std::lower_bound(fl.begin(), fl.end(), std::make_pair(s, 0));
    it = fl.insert(it, std::make_pair(s, n));
    if (it + 1 != fl.end() && it->first + it->second == (it + 1)->first) { it->second += (it + 1)->second; fl.erase(it + 1); }
    if (it != fl.begin() && (it - 1)->first + (it - 1)->second == it->first) { (it - 1)->second += it->second; fl.erase(it); }
}
// a scene's ranges marked as its own on the device and the host: scene of each body slot, scene of each brick,
// no structure yet
int claim_ranges(Lat* L, int si) {
    HScene& S = L->sc[si];
    LD& D = L->D;
    L->nb = std::max(L->nb, S.b0 + S.bcap); L->nbd = std::max(L->nbd, S.bd0 + S.dcap);
    D.nb = (uint32_t)L->nb; D.nbd = (uint32_t)L->nbd;
    std::vector<int32_t> bsc(S.dcap, si);
    if (S.dcap && cp_up(D.bd_scene + S.bd0, bsc.data(), S.dcap, L->st)) return -1;
    std::vector<int4> bx(S.bcap, make_int4(0, 0, 0, si));
    if (S.bcap && cp_up(D.bxyz + S.b0, bx.data(), S.bcap, L->st)) return -1;
    std::vector<uint8_t> sids(S.bcap, 255);
    if (S.bcap && cp_up(D.sid + S.b0, sids.data(), S.bcap, L->st)) return -1;
    for (int k = 0; k < S.bcap; k++) { L->bxyz[S.b0 + k] = make_int4(0, 0, 0, si); L->sid[S.b0 + k] = 255; }
    return 0;
}

void clear_scene(Lat* L, int si, uint32_t seed) {
    HScene& S = L->sc[si];
    // the search stamps restart with the scene's stamp counter, so the old ones must go
    for (int i = S.b0; i < S.b0 + S.bcap; i++) { L->mat[i] = 0; L->body[i] = -1; L->next[i] = NONE; L->sid[i] = 255; L->stamp[i] = 0; }
    for (int b = S.bd0; b < S.bd0 + S.dcap; b++) { L->alive[b] = 0; L->first[b] = NONE; L->asleep[b] = 0; L->dirty[b] = 0; L->wake[b] = 0; L->nc[b] = 0; }
    int st0 = S.st0, smax = (int)(L->sts.size()) ;
    (void)smax;
    for (int sl = 0; sl < S.nst; sl++) { HStruct fresh; fresh.scene = si; L->sts[st0 + sl] = fresh; }
    int b0 = S.b0, bcap = S.bcap, bd0 = S.bd0, dcap = S.dcap, stc = S.st0;
    size_t goff = S.goff, gcap = S.gcap;
    int gnx = S.gnx, gny = S.gny, gnz = S.gnz;
    bool had_grid = S.finalized;
    S = HScene();
    S.b0 = b0; S.bcap = bcap; S.bd0 = bd0; S.dcap = dcap; S.st0 = stc;
    S.goff = goff; S.gcap = gcap;  // the grid region is kept for the next finalize
    S.rng = seed ? seed : 0x9E3779B9u;
    if (had_grid) { for (size_t k = 0; k < (size_t)gnx * gny * gnz; k++) L->grid[goff + k] = 0; }
    L->lists_dirty = true;
    L->ub_dirty = true;
}

struct Handle {
    Lat L;
    std::vector<int> smax;  // structure capacity per scene
};

}  // namespace

struct sf_lat { Handle h; };

extern "C" {

sf_lat* sf_lat_create(int brick_cap, int body_cap, int max_scenes, int colours, int rounds) {
    if (brick_cap <= 0 || body_cap <= 0 || max_scenes <= 0 || colours < 1 || colours > KMAX || rounds < 0) {
        sf::set_error(SF_E_ARG, "bad capacities or colouring parameters"); return nullptr;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { sf::set_error(SF_E_NODEV, "no GPU visible"); return nullptr; }
    sf_lat* h = new sf_lat();
    Lat* L = &h->h.L;
    SF_HIP_NULL(hipGetDevice(&L->dev));
    L->bcap = brick_cap; L->dcap = body_cap; L->scap = max_scenes; L->K = colours; L->R = rounds;
    L->free_b.assign(1, std::make_pair(0, brick_cap));
    L->free_d.assign(1, std::make_pair(0, body_cap));
    if (hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking) != hipSuccess) { delete h; sf::set_error(SF_E_HIP, "stream"); return nullptr; }
    if (lat_alloc(L) || lat_contacts_alloc(L, (uint32_t)std::max<int>(1 << 16, 4 * brick_cap)) || lat_hash_alloc(L, 1024) || L->d_grid.alloc(1, L->st)) { delete h; return nullptr; }
    L->D.grid = L->d_grid.p;
    if (L->h_ctrbuf.reserve(CTR_WORDS + (size_t)max_scenes + 1)) { delete h; sf::set_error(SF_E_HIP, "hipHostMalloc failed"); return nullptr; }
    L->h_ctr = L->h_ctrbuf.p;
    memset(L->h_ctr, 0, sizeof(uint32_t) * C_NCTR);
    size_t B = brick_cap, N = body_cap;
    L->mat.assign(B, 0); L->sid.assign(B, 255); L->infill.assign(B, 0); L->body.assign(B, -1); L->bxyz.assign(B, make_int4(0, 0, 0, 0));
    L->next.assign(B, NONE); L->nb6.assign(6 * B, NONE); L->stamp.assign(B, 0); L->q.assign(B, 0);
    L->bond.assign(B, make_float4(0, 0, 0, 0)); L->bmax = L->bond; L->local = L->bond; L->wpos = L->bond;
    L->phi.assign(B, 0.0f); L->ext = L->phi; L->extn = L->phi;
    L->alive.assign(N, 0); L->asleep.assign(N, 0); L->dirty.assign(N, 0); L->wake.assign(N, 0);
    L->first.assign(N, NONE); L->count.assign(N, 0); L->nc.assign(N, 0);
    L->pos.assign(N, make_float4(0, 0, 0, 0)); L->qq.assign(N, make_float4(0, 0, 0, 1)); L->v = L->pos; L->w = L->pos; L->vlast = L->pos; L->wlast = L->pos;
    LM z; for (int k = 0; k < 9; k++) z.m[k] = 0.0f;
    L->rot.assign(N, z); L->iil = L->rot; L->iiw = L->rot;
    L->m.assign(N, 0.0f); L->im = L->m; L->sleep = L->m; L->rad = L->m; L->cool = L->m; L->lt = L->m; L->resid = L->m; L->maxr = L->m; L->nmax = L->m;
    L->touched.assign(B, 0);
    L->btouched.assign(N, 0);
    L->lst.assign(B + 1, 0); L->lst2.assign(B + 1, 0); L->comp.assign(B + 1, 0); L->cnt.assign(B + 2, 0);
    L->dyn.assign(B, 0); L->dyn_body.assign(B, 0);
    return h;
}
void sf_lat_destroy(sf_lat* h) { delete h; }

int sf_lat_scene_add(sf_lat* h, int brick_cap, int body_cap, int max_structs, uint32_t seed) {
    if (!h || brick_cap <= 0 || body_cap <= 0 || max_structs < 0 || max_structs > 255) { sf::set_error(SF_E_ARG, "bad scene capacities"); return SF_E_ARG; }
    Lat* L = &h->h.L;
    if ((int)L->sc.size() >= L->scap || (L->sts.size() + max_structs) > 256 * (size_t)L->scap) {
        sf::set_error(SF_E_CAP, "world capacity exceeded"); return SF_E_CAP;
    }
    const int b0 = pool_take(L->free_b, brick_cap);
    if (b0 < 0) { sf::set_error(SF_E_CAP, "world brick capacity exceeded"); return SF_E_CAP; }
    const int bd0 = pool_take(L->free_d, body_cap);
    if (bd0 < 0) { pool_give(L->free_b, b0, brick_cap); sf::set_error(SF_E_CAP, "world body capacity exceeded"); return SF_E_CAP; }
    HScene S;
    S.b0 = b0; S.bcap = brick_cap; S.bd0 = bd0; S.dcap = body_cap; S.st0 = (int)L->sts.size();
    S.rng = seed ? seed : 0x9E3779B9u;
    int si = (int)L->sc.size();
    L->sc.push_back(S);
    h->h.smax.push_back(max_structs);
    for (int k = 0; k < max_structs; k++) { HStruct T; T.scene = si; L->sts.push_back(T); }
    LD& D = L->D;
    D.nscene = (uint32_t)L->sc.size(); D.nst = (uint32_t)L->sts.size();
    if (claim_ranges(L, si)) return SF_E_HIP;
    SF_HIP(hipStreamSynchronize(L->st));
    if (upload_scenes(L)) return SF_E_HIP;
    return si;
}
int sf_lat_scene_release(sf_lat* h, int scene) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size()) { sf::set_error(SF_E_ARG, "bad scene"); return SF_E_ARG; }
    Lat* L = &h->h.L;
    int r = sf_lat_scene_clear(h, scene, 1);
    if (r < 0) return r;
    HScene& S = L->sc[scene];
    pool_give(L->free_b, S.b0, S.bcap);
    pool_give(L->free_d, S.bd0, S.dcap);
    S.bcap = 0; S.dcap = 0;
    L->ub_dirty = true;
    L->lists_dirty = true;
    if (rebuild_csr(L) || upload_scenes(L)) return SF_E_HIP;
    return 0;
}
int sf_lat_scene_place(sf_lat* h, int scene, int brick_cap, int body_cap, uint32_t seed) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size() || brick_cap <= 0 || body_cap <= 0) { sf::set_error(SF_E_ARG, "bad scene or capacities"); return SF_E_ARG; }
    Lat* L = &h->h.L;
    HScene& S = L->sc[scene];
    if (S.bcap || S.dcap) { sf::set_error(SF_E_ARG, "scene still placed: release it first"); return SF_E_ARG; }
    const int b0 = pool_take(L->free_b, brick_cap);
    if (b0 < 0) { sf::set_error(SF_E_CAP, "no free brick range that large"); return SF_E_CAP; }
    const int bd0 = pool_take(L->free_d, body_cap);
    if (bd0 < 0) { pool_give(L->free_b, b0, brick_cap); sf::set_error(SF_E_CAP, "no free body range that large"); return SF_E_CAP; }
    S.b0 = b0; S.bcap = brick_cap; S.bd0 = bd0; S.dcap = body_cap;
    if (claim_ranges(L, scene)) return SF_E_HIP;
    SF_HIP(hipStreamSynchronize(L->st));
    return sf_lat_scene_clear(h, scene, seed);
}
int64_t sf_lat_bytes_for(sf_lat* h, int brick_cap, int body_cap, int max_structs) {
    if (!h || brick_cap < 0 || body_cap < 0 || max_structs < 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    const Lat* L = &h->h.L;
    return (int64_t)brick_cap * (int64_t)(L->per_brick + 4 * L->per_contact) + (int64_t)body_cap * (int64_t)L->per_body +
           (int64_t)max_structs * (int64_t)L->per_struct;
}
int64_t sf_lat_bytes_resident(sf_lat* h, int scene) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size()) { sf::set_error(SF_E_ARG, "bad scene"); return SF_E_ARG; }
    const Lat* L = &h->h.L;
    const HScene& S = L->sc[scene];
    // its bricks' and bodies' shares of every per-brick and per-body array, four contacts per brick (the contact
    // arrays' sizing), its structures' rows and its static grid
    return (int64_t)S.bcap * (int64_t)(L->per_brick + 4 * L->per_contact) + (int64_t)S.dcap * (int64_t)L->per_body +
           (int64_t)h->h.smax[scene] * (int64_t)L->per_struct + (int64_t)S.gcap * (int64_t)sizeof(int32_t);
}
int sf_lat_scene_clear(sf_lat* h, int scene, uint32_t seed) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size()) { sf::set_error(SF_E_ARG, "bad scene"); return SF_E_ARG; }
    Lat* L = &h->h.L;
    HScene& S = L->sc[scene];
    size_t b0 = S.b0, bc = S.bcap, d0 = S.bd0, dc = S.dcap;
    clear_scene(L, scene, seed);
    LD& D = L->D;
    int r = 0;
    uint32_t zf[4] = {0, 0, 0, 0};
    r |= cp_up(D.sc_fail + 4 * (size_t)scene, zf, 4, L->st);
    r |= cp_up(D.mat + b0, &L->mat[b0], bc, L->st); r |= cp_up(D.body + b0, &L->body[b0], bc, L->st);
    r |= cp_up(D.sid + b0, &L->sid[b0], bc, L->st);
    r |= cp_up(D.alive + d0, &L->alive[d0], dc, L->st); r |= cp_up(D.asleep + d0, &L->asleep[d0], dc, L->st);
    r |= cp_up(D.wake + d0, &L->wake[d0], dc, L->st);
    if (r) return SF_E_HIP;
    SF_HIP(hipStreamSynchronize(L->st));
    if (rebuild_csr(L) || upload_scenes(L)) return SF_E_HIP;
    return 0;
}
int sf_lat_add_structure(sf_lat* h, int scene, int n, const int32_t* xyz, const uint8_t* mat, const uint8_t* infill) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size() || n < 0 || (n && (!xyz || !mat))) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    Lat* L = &h->h.L;
    HScene& S = L->sc[scene];
    if (S.finalized) { sf::set_error(SF_E_ARG, "scene already finalized"); return SF_E_ARG; }
    if (S.nst >= h->h.smax[scene]) { sf::set_error(SF_E_CAP, "scene structure capacity exceeded"); return SF_E_CAP; }
    if (S.nb + n > S.bcap) { sf::set_error(SF_E_CAP, "scene brick capacity exceeded"); return SF_E_CAP; }
    for (int k = 0; k < n; k++) if (mat[k] < 1 || mat[k] > 8) { sf::set_error(SF_E_ARG, "material out of range"); return SF_E_ARG; }
    int sl = S.nst++;
    HStruct& T = L->sts[S.st0 + sl];
    T = HStruct(); T.scene = scene;
    T.lo = S.b0 + S.nb;
    for (int k = 0; k < n; k++) add_brick(L, S, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], mat[k], sl, infill ? infill[k] : 0, scene);
    T.hi = S.b0 + S.nb;
    return sl;
}
int sf_lat_add_body(sf_lat* h, int scene, int n, const int32_t* xyz, const uint8_t* mat, const float* v3, const float* w3) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size() || n <= 0 || !xyz || !mat || !v3 || !w3) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    if (h->h.L.sc[scene].finalized) { sf::set_error(SF_E_ARG, "scene already finalized"); return SF_E_ARG; }
    for (int k = 0; k < n; k++) if (mat[k] < 1 || mat[k] > 8) { sf::set_error(SF_E_ARG, "material out of range"); return SF_E_ARG; }
    return add_body(&h->h.L, scene, n, xyz, mat, v3, w3);
}
int sf_lat_finalize(sf_lat* h, int scene) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size()) { sf::set_error(SF_E_ARG, "bad scene"); return SF_E_ARG; }
    int r = finalize(&h->h.L, scene);
    return r == SF_E_ARG ? SF_E_ARG : (r ? SF_E_HIP : 0);
}
int sf_lat_warm_all(sf_lat* h, int scene) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size()) { sf::set_error(SF_E_ARG, "bad scene"); return SF_E_ARG; }
    Lat* L = &h->h.L;
    HScene& S = L->sc[scene];
    for (int sl = 0; sl < S.nst; sl++) st_warm(L, S.st0 + sl);
    return 0;
}
int sf_lat_wake_all(sf_lat* h, int scene) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size()) { sf::set_error(SF_E_ARG, "bad scene"); return SF_E_ARG; }
    Lat* L = &h->h.L;
    HScene& S = L->sc[scene];
    for (int sl = 0; sl < S.nst; sl++) st_activate(L, S, sl);
    return 0;
}
int sf_lat_explode(sf_lat* h, int scene, float x, float y, float z, float r, float power) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size() || !(r > 0.0f)) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    return explode(&h->h.L, scene, x, y, z, r, power) ? SF_E_HIP : 0;
}
int sf_lat_step(sf_lat* h, float dt) {
    if (!h) return SF_E_ARG;
    if (!(dt >= 0.0f)) { sf::set_error(SF_E_ARG, "dt out of range"); return SF_E_ARG; }
    SF_HIP(hipSetDevice(h->h.L.dev));
    int r = frame(&h->h.L, dt);
    return r ? (sf::last_error() ? sf::last_error() : SF_E_HIP) : 0;
}
int sf_lat_stats(sf_lat* h, int scene, float* out16) {
    if (!h || !out16 || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    if (finish_stats(&h->h.L)) return SF_E_HIP;
    memcpy(out16, h->h.L.sc[scene].stats, sizeof(float) * 16);
    return 0;
}
int sf_lat_counts(sf_lat* h, int scene, int32_t* out3) {
    if (!h || !out3 || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    const HScene& S = h->h.L.sc[scene];
    out3[0] = S.nb; out3[1] = S.hw; out3[2] = S.nst;
    return 0;
}
// per body slot of the scene: alive, asleep, count, nc, pos[3], q[4], v[3], w[3], m, sleep, maxr, resid, nmax
int sf_lat_bodies(sf_lat* h, int scene, float* out) {
    if (!h || !out || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    Lat* L = &h->h.L;
    if (pull(L, scene, G_BODY)) return SF_E_HIP;
    const HScene& S = L->sc[scene];
    for (int lb = 0; lb < S.hw; lb++) {
        int b = S.bd0 + lb;
        float* o = out + (size_t)lb * SF_LAT_BODY_F;
        o[0] = L->alive[b]; o[1] = L->asleep[b]; o[2] = (float)L->count[b]; o[3] = (float)L->nc[b];
        o[4] = L->pos[b].x; o[5] = L->pos[b].y; o[6] = L->pos[b].z;
        o[7] = L->qq[b].x; o[8] = L->qq[b].y; o[9] = L->qq[b].z; o[10] = L->qq[b].w;
        o[11] = L->v[b].x; o[12] = L->v[b].y; o[13] = L->v[b].z; o[14] = L->w[b].x; o[15] = L->w[b].y; o[16] = L->w[b].z;
        o[17] = L->m[b]; o[18] = L->sleep[b]; o[19] = L->maxr[b]; o[20] = L->resid[b]; o[21] = L->nmax[b];
    }
    return S.hw;
}
// per brick: mat, body (scene-local, -1 static), wpos[3], phi, ext, bond[4], stress, frames since support
int sf_lat_bricks(sf_lat* h, int scene, float* out) {
    if (!h || !out || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    Lat* L = &h->h.L;
    if (pull(L, scene, G_MAT | G_BOND | G_BRICK)) return SF_E_HIP;
    const HScene& S = L->sc[scene];
    std::vector<float> stress(S.nb);
    std::vector<uint32_t> sup(S.nb);
    if (cp_down(stress.data(), L->D.stress + S.b0, S.nb, L->st) || cp_down(sup.data(), L->D.sup + S.b0, S.nb, L->st)) return SF_E_HIP;
    SF_HIP(hipStreamSynchronize(L->st));
    for (int k = 0; k < S.nb; k++) {
        int i = S.b0 + k;
        float* o = out + (size_t)k * SF_LAT_BRICK_F;
        o[0] = L->mat[i]; o[1] = L->body[i] < 0 ? -1.0f : (float)(L->body[i] - S.bd0);
        o[2] = L->wpos[i].x; o[3] = L->wpos[i].y; o[4] = L->wpos[i].z;
        o[5] = L->phi[i]; o[6] = L->ext[i];
        o[7] = L->bond[i].x; o[8] = L->bond[i].y; o[9] = L->bond[i].z; o[10] = L->bond[i].w;
        o[11] = stress[k]; o[12] = (float)(S.frame - sup[k]);
    }
    return S.nb;
}
int sf_lat_structs(sf_lat* h, int scene, float* out) {
    if (!h || !out || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    Lat* L = &h->h.L;
    const HScene& S = L->sc[scene];
    for (int sl = 0; sl < S.nst; sl++) {
        const HStruct& T = L->sts[S.st0 + sl];
        float* o = out + 12 * sl;
        LV pa = lat_pos(T.mn[0], T.mn[1], T.mn[2]), pb = lat_pos(T.mx[0], T.mx[1], T.mx[2]);
        o[0] = pa.x - H; o[1] = pa.z - H; o[2] = pb.x + H; o[3] = pb.z + H;
        o[4] = (float)T.init; o[5] = (float)T.now; o[6] = (float)T.top0; o[7] = (float)T.top;
        o[8] = 0.0f; o[9] = 0.0f; o[10] = T.active > 0.0f ? 1.0f : 0.0f; o[11] = (float)T.top0 * lat::S;
    }
    return S.nst;
}
int sf_lat_matstats(sf_lat* h, int scene, float* out50) {
    if (!h || !out50 || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    Lat* L = &h->h.L;
    if (pull(L, scene, G_MAT | G_BODY)) return SF_E_HIP;
    const HScene& S = L->sc[scene];
    for (int k = 0; k < 50; k++) out50[k] = 0.0f;
    for (int lb = 0; lb < S.hw; lb++) {
        int b = S.bd0 + lb;
        if (!L->alive[b]) continue;
        uint32_t cnt[9] = {0}, tot = 0;
        for (uint32_t k = L->first[b]; k != NONE; k = L->next[k]) if (L->mat[k] != 0 && L->mat[k] < 9) { cnt[L->mat[k]]++; tot++; }
        if (!tot) continue;
        int m = 1;
        for (int q = 1; q < 9; q++) if (cnt[q] > cnt[m]) m = q;
        if (m * 5 + 4 >= 50) continue;
        out50[m * 5] += 1.0f; out50[m * 5 + 1] += (float)tot;
        if ((float)tot > out50[m * 5 + 2]) out50[m * 5 + 2] = (float)tot;
        if (tot >= 50) out50[m * 5 + 3] += 1.0f;
        if (tot <= 2) out50[m * 5 + 4] += 1.0f;
    }
    return 0;
}
// contacts of the last substep: a, b, bo, ba, bb (scene-local; -1 none), on, colour, n[3], p[3], depth, jn, jt1, jt2, vn0
int sf_lat_contacts(sf_lat* h, int scene, float* out, int cap) {
    if (!h || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    Lat* L = &h->h.L;
    const HScene& S = L->sc[scene];
    int n = (int)S.ncs;
    if (!out || cap <= 0) return n;
    int m = std::min(n, cap);
    LD& D = L->D;
    size_t c0 = S.c0;
    std::vector<int32_t> a(m), b(m), bo(m), col(m);
    std::vector<uint32_t> ba(m), bb(m);
    std::vector<uint8_t> on(m);
    std::vector<float4> nn(m), pp(m);
    std::vector<float> depth(m), jn(m), jt1(m), jt2(m), vn0(m);
    int r = 0;
    r |= cp_down(a.data(), D.ca + c0, m, L->st); r |= cp_down(b.data(), D.cb + c0, m, L->st); r |= cp_down(bo.data(), D.cbo + c0, m, L->st);
    r |= cp_down(col.data(), D.ccol + c0, m, L->st); r |= cp_down(ba.data(), D.cba + c0, m, L->st); r |= cp_down(bb.data(), D.cbb + c0, m, L->st);
    r |= cp_down(on.data(), D.con + c0, m, L->st); r |= cp_down(nn.data(), D.cn + c0, m, L->st); r |= cp_down(pp.data(), D.cp + c0, m, L->st);
    r |= cp_down(depth.data(), D.cdepth + c0, m, L->st); r |= cp_down(jn.data(), D.cjn + c0, m, L->st);
    r |= cp_down(jt1.data(), D.cjt1 + c0, m, L->st); r |= cp_down(jt2.data(), D.cjt2 + c0, m, L->st); r |= cp_down(vn0.data(), D.cvn0 + c0, m, L->st);
    if (r) return SF_E_HIP;
    SF_HIP(hipStreamSynchronize(L->st));
    for (int k = 0; k < m; k++) {
        float* o = out + (size_t)k * SF_LAT_CONTACT_F;
        o[0] = (float)(a[k] - S.bd0); o[1] = b[k] < 0 ? -1.0f : (float)(b[k] - S.bd0); o[2] = bo[k] < 0 ? -1.0f : (float)(bo[k] - S.bd0);
        o[3] = (float)((int)ba[k] - S.b0); o[4] = bb[k] == NONE ? -1.0f : (float)((int)bb[k] - S.b0);
        o[5] = on[k]; o[6] = (float)col[k];
        o[7] = nn[k].x; o[8] = nn[k].y; o[9] = nn[k].z; o[10] = pp[k].x; o[11] = pp[k].y; o[12] = pp[k].z;
        o[13] = depth[k]; o[14] = jn[k]; o[15] = jt1[k]; o[16] = jt2[k]; o[17] = vn0[k];
    }
    return n;
}
int sf_lat_set_field(sf_lat* h, int scene, int sl, const float* phi) {
    if (!h || !phi || scene < 0 || scene >= (int)h->h.L.sc.size()) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    Lat* L = &h->h.L;
    HScene& S = L->sc[scene];
    if (sl < 0 || sl >= S.nst || !S.finalized) { sf::set_error(SF_E_ARG, "bad structure or scene not finalized"); return SF_E_ARG; }
    HStruct& T = L->sts[S.st0 + sl];
    for (int i = T.lo; i < T.hi; i++) L->phi[i] = phi[i - T.lo];
    T.solved = true; T.warm = 0; T.settle = 0.3f; T.resid = 0.0f;
    float z = 0.0f;
    if (cp_up(L->D.phi + T.lo, &L->phi[T.lo], (size_t)(T.hi - T.lo), L->st) || cp_up(L->D.st_resid + S.st0 + sl, &z, 1, L->st)) return SF_E_HIP;
    SF_HIP(hipStreamSynchronize(L->st));
    return 0;
}
int sf_lat_failures(sf_lat* h, int scene, uint32_t* out4) {
    if (!h || !out4 || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    Lat* L = &h->h.L;
    uint32_t d[4];
    if (cp_down(d, L->D.sc_fail + 4 * (size_t)scene, 4, L->st)) return SF_E_HIP;
    SF_HIP(hipStreamSynchronize(L->st));
    out4[0] = L->sc[scene].crushed_joints; out4[1] = d[1]; out4[2] = d[2]; out4[3] = d[3];
    return 0;
}
int sf_lat_rng(sf_lat* h, int scene, uint32_t* out) {
    if (!h || !out || scene < 0 || scene >= (int)h->h.L.sc.size()) return SF_E_ARG;
    *out = h->h.L.sc[scene].rng;
    return 0;
}
int sf_lat_phase_ms(sf_lat* h, double* out8) {
    if (!h || !out8) return SF_E_ARG;
    for (int k = 0; k < 8; k++) out8[k] = h->h.L.ph[k];
    return 0;
}
int sf_lat_timing(sf_lat* h, double* ms_frame, int32_t* syncs) {
    if (!h) return SF_E_ARG;
    if (ms_frame) *ms_frame = h->h.L.ms_frame;
    if (syncs) *syncs = h->h.L.syncs;
    return 0;
}
int sf_lat_subphase_ms(sf_lat* h, double* out16) {
    if (!h || !out16) return SF_E_ARG;
    for (int k = 0; k < 16; k++) out16[k] = h->h.L.sp[k];
    return 0;
}
int sf_lat_set_host_threads(sf_lat* h, int n) {
    if (!h || n < 0 || n > 16) { sf::set_error(SF_E_ARG, "host threads: 0 (default) to 16"); return SF_E_ARG; }
    Lat& L = h->h.L;
    if (n == L.host_threads) return 0;
    L.pool.reset();
    L.scr.clear();
    L.host_threads = n;
    return 0;
}
int sf_lat_set_fused(sf_lat* h, int max_contacts) {
    if (!h || max_contacts < 0) return SF_E_ARG;
    h->h.L.fuse_max = max_contacts;
    return 0;
}
int sf_lat_set_conn_device(sf_lat* h, int on) {
    if (!h) return SF_E_ARG;
    h->h.L.conn_device = on != 0;
    return 0;
}
int sf_lat_set_sor_lds(sf_lat* h, int max_bricks) {
    if (!h || max_bricks < -1) { sf::set_error(SF_E_ARG, "sf_lat_set_sor_lds: max_bricks must be >= -1"); return SF_E_ARG; }
    h->h.L.sor_lds = max_bricks;
    return 0;
}

}  // extern "C"
