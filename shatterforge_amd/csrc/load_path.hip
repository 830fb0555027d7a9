// Load path on the GPU: the per-structure scalar-potential solve of the reference's legacy engine
// (src/lib.rs sor_brick :1005-1026, sor_struct :1029-1040, warm_struct/warm_step :1061-1079) and the joint
// verdicts of stress_brick (:1153-1227), for a batch of independent structures in one launch.
//
// Each structure is the linear system sum_j k_ij (phi_i - phi_j) + g_i phi_i = MASS_i + ext_i over its grounded
// bricks (k = min stiffness x 1 vertical / 0.3 lateral, g = 4 x stiffness on footings). Bricks are reordered red
// (x+y+z even) then black: lattice neighbours always differ in colour, so a half-sweep updates one colour from the
// other and is order-independent. Precision: the FP64 field phi is corrected by FP32 red-black SOR sweeps on
// A delta = r, where r = b - A phi is evaluated in FP64 (iterative refinement); convergence is judged on the FP64
// residual max|r| / max|b|. One workgroup owns one structure for the whole solve, so the only barriers are
// workgroup barriers, and every structure of the batch runs in the same launch.
#include "sf_host.h"
#include "../../include/sf.h"
#include <hip/hip_runtime.h>
#include <vector>
#include <cmath>
#include <cstring>
#include <algorithm>
#include <unordered_map>
#include <chrono>
#include <mutex>

namespace {

const float MASS[9] = {0.0f, 300.0f, 230.0f, 250.0f, 60.0f, 70.0f, 2000.0f, 8.0f, 44.0f};
const float CAPV[9] = {0.0f, 140000.0f, 50000.0f, 150000.0f, 1500.0f, 9000.0f, 1e12f, 1e12f, 1e12f};
const float CAPH[9] = {0.0f, 12000.0f, 12000.0f, 90000.0f, 400.0f, 5000.0f, 1e12f, 1e12f, 1e12f};
const float STIFF[9] = {0.0f, 1.0f, 0.8f, 2.0f, 0.02f, 0.4f, 1.0f, 1.0f, 1.0f};
const float GROUNDK = 4.0f;
const int DIRS[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

constexpr int LT = 1024;  // threads per structure workgroup

// one structure as stored on the host (input order) plus its device placement
struct Struct {
    int n = 0;
    std::vector<uint8_t> mat, infill;
    std::vector<int32_t> xyz, nb6;
    std::vector<float> bond, ext;
    std::vector<uint8_t> grounded;
    int ngrounded = 0;
    // device placement: bricks base..base+n in device order (red, then black); perm[device k] = input index
    int base = 0, nred = 0;
    std::vector<int32_t> perm, inv;
    float omega = 1.85f;
    // results
    int sweeps = 0;
    double resid = 1.0;
};

struct DevInfo { int base, n, nred, max_sweeps; float omega; double bmax; };

struct LoadDev {
    int nstruct;
    const DevInfo* info;
    const int32_t* nbr;    // 6 x N: device index of the linked neighbour or -1
    const float* kc;       // 6 x N: conductance of that bond (f32, as the reference computes it)
    const float* sk32;     // diagonal in f32 (sum of the bonds' conductances in order, plus the footing)
    const double* sk64;    // the same diagonal summed in f64
    const float* b;        // MASS + ext (0 for bricks not grounded)
    const uint8_t* act;    // 1 = grounded (solved)
    double* phi;           // FP64 field
    float* delta;          // FP32 correction
    float* r32;            // FP32 copy of the FP64 residual
    int* out_sweeps;
    double* out_resid;
    size_t N;
    double tol;
    int inner;             // FP32 sweeps between FP64 residual evaluations
    int lds_floats;        // structures of up to this many bricks keep their correction in LDS
    // Packed sweep rows, used by the structures whose correction is in LDS when pk_codes is set: per brick, x/y/z hold
    // the six neighbours' indices within the structure as 16-bit halves (0xFFFF: no bond), w the six bonds'
    // conductances as 5-bit codes into ktab (bits 5d..5d+4) and the brick's grounded flag in bit 30. A sweep then
    // reads 16 B of stencil per brick instead of 49 (six indices, six conductances, the flag).
    const uint4* row;
    const float* ktab;     // the batch's distinct conductances (code 0: 0.0f); PK_CODES entries
    int pk_codes;          // 1: rows and table are valid
    // Block b solves structure order[b]: largest first. A CU holds one structure workgroup when the corrections of
    // the largest structures fill its LDS, so the biggest solves start in the first round of workgroups and the
    // small ones fill in behind them instead of a big one starting last.
    const int32_t* order;
};
constexpr int PK_CODES = 32;
constexpr uint32_t PK_NONE = 0xFFFFu;

__device__ __forceinline__ double block_max(double v, double* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) { double t = __shfl_xor(v, o, 64); v = t > v ? t : v; }
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double m = 0.0;
    for (int k = 0; k < LT / 64; k++) m = sh[k] > m ? sh[k] : m;
    __syncthreads();
    return m;
}

// the FP32 correction of a structure: in LDS when it fits (neighbour reads then cost no memory round trip and the
// stencil loads, which cannot alias it, overlap), else in global memory
template <bool LDSD> struct Corr;
template <> struct Corr<true> {
    float* p; size_t base;
    __device__ __forceinline__ float& operator()(size_t g) const { return p[(uint32_t)(g - base)]; }
};
template <> struct Corr<false> {
    float* p; size_t base;
    __device__ __forceinline__ float& operator()(size_t g) const { return p[g]; }
};

// One half-sweep of the packed rows (the correction in LDS, neighbour k at lds[k]): the operations of the unpacked
// half-sweep below, on the same f32 values, with U bricks per thread whose rows load together.
__device__ __forceinline__ void half_sweep_packed(const LoadDev& L, int base, int i0, int i1, float om, float* del,
                                                  const float* ktab) {
    constexpr int U = 8;
    for (int i = i0 + threadIdx.x; i < i1; i += LT * U) {
        uint4 q[U];
        float rr[U], ss[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ii = i + u * LT;
            const size_t g = (size_t)base + (ii < i1 ? ii : i);
            q[u] = L.row[g]; rr[u] = L.r32[g]; ss[u] = L.sk32[g];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ii = i + u * LT;
            if (ii >= i1 || !((q[u].w >> 30) & 1u)) continue;
            const uint32_t rel[6] = {q[u].x & 0xFFFFu, q[u].x >> 16, q[u].y & 0xFFFFu, q[u].y >> 16, q[u].z & 0xFFFFu, q[u].z >> 16};
            float skp = 0.0f;
#pragma unroll
            for (int d = 0; d < 6; d++) if (rel[d] != PK_NONE) skp += ktab[(q[u].w >> (5 * d)) & 31u] * del[rel[d]];
            float target = (rr[u] + skp) / ss[u];
            del[ii] += om * (target - del[ii]);
        }
    }
}

// the whole solve of one structure by one workgroup
template <bool LDSD>
__device__ void solve_struct(const LoadDev& L, int s, float* lds, double* sh, const float* ktab) {
    const DevInfo I = L.info[s];
    const int base = I.base, n = I.n, nred = I.nred;
    const size_t N = L.N;
    const float om = I.omega;
    const Corr<LDSD> del{LDSD ? lds : L.delta, (size_t)base};
    int sweeps = 0;
    double rel = 1.0;
    while (true) {
        // FP64 residual r = b - A phi, kept in f32 for the correction sweeps
        double rmax = 0.0;
        for (int i = threadIdx.x; i < n; i += LT) {
            size_t g = (size_t)base + i;
            double r = 0.0;
            if (L.act[g]) {
                double acc = L.sk64[g] * L.phi[g];
                for (int d = 0; d < 6; d++) {
                    int j = L.nbr[d * N + g];
                    if (j >= 0) acc -= (double)L.kc[d * N + g] * L.phi[j];
                }
                r = (double)L.b[g] - acc;
            }
            L.r32[g] = (float)r;
            del(g) = 0.0f;
            double a = fabs(r);
            rmax = a > rmax ? a : rmax;
        }
        rmax = block_max(rmax, sh);
        rel = I.bmax > 0.0 ? rmax / I.bmax : 0.0;
        if (rel <= L.tol || sweeps >= I.max_sweeps) break;
        // FP32 red-black SOR on A delta = r
        for (int it = 0; it < L.inner && sweeps < I.max_sweeps; it++, sweeps++) {
            for (int half = 0; half < 2; half++) {
                int i0 = half ? nred : 0, i1 = half ? n : nred;
                if (LDSD && L.pk_codes) {
                    half_sweep_packed(L, base, i0, i1, om, lds, ktab);
                    __syncthreads();
                    continue;
                }
                // U bricks of one colour per thread at a time: their stencils are loaded together, then the
                // neighbours (all of the other colour, so no brick of this batch is read after it is written)
                constexpr int U = 4;
                for (int i = i0 + threadIdx.x; i < i1; i += LT * U) {
                    int jj[U][6];
                    float kk[U][6], rr[U], ss[U];
                    bool on[U];
                    // (the rows load whether or not the brick is active, from a valid index, so they are in flight
                    // with the activity flag instead of after it)
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        int ii = i + u * LT;
                        const bool in = ii < i1;
                        size_t g = (size_t)base + (in ? ii : i);
                        on[u] = in && L.act[g];
#pragma unroll
                        for (int d = 0; d < 6; d++) { int jv = L.nbr[d * N + g]; float kv = L.kc[d * N + g]; jj[u][d] = on[u] ? jv : -1; kk[u][d] = on[u] ? kv : 0.0f; }
                        float rv = L.r32[g], sv = L.sk32[g];
                        rr[u] = on[u] ? rv : 0.0f;
                        ss[u] = on[u] ? sv : 1.0f;
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (!on[u]) continue;
                        size_t g = (size_t)base + i + u * LT;
                        float skp = 0.0f;
#pragma unroll
                        for (int d = 0; d < 6; d++) if (jj[u][d] >= 0) skp += kk[u][d] * del((size_t)jj[u][d]);
                        float target = (rr[u] + skp) / ss[u];
                        del(g) += om * (target - del(g));
                    }
                }
                __syncthreads();
            }
        }
        for (int i = threadIdx.x; i < n; i += LT) {
            size_t g = (size_t)base + i;
            L.phi[g] += (double)del(g);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { L.out_sweeps[s] = sweeps; L.out_resid[s] = rel; }
}
__global__ void __launch_bounds__(LT) k_load_solve(LoadDev L) {
    __shared__ double sh[LT / 64];
    __shared__ float ktab[PK_CODES];
    extern __shared__ float lds[];
    if ((int)blockIdx.x >= L.nstruct) return;
    const int s = L.order[blockIdx.x];
    if (L.pk_codes && threadIdx.x < PK_CODES) ktab[threadIdx.x] = L.ktab[threadIdx.x];
    __syncthreads();
    if (L.info[s].n <= L.lds_floats) solve_struct<true>(L, s, lds, sh, ktab);
    else solve_struct<false>(L, s, lds, sh, ktab);
}

struct Verd { double* ratio; uint8_t* mode; uint8_t* fail; };

// stress_brick (src/lib.rs:1153-1227) for standing bricks, from the FP64 field: +x, +y, +z bonds and the footing
__global__ void k_verdicts(const int32_t* nbr, const float* kc, const float* hp, const uint8_t* mat, const uint8_t* footing,
                           const double* phi, const float* gk, size_t N, Verd V) {
    const float TENSd[9] = {0.0f, 0.35f, 0.1f, 0.9f, 0.5f, 0.7f, 1.0f, 1.0f, 1.0f};
    for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < N; g += (size_t)gridDim.x * blockDim.x) {
        for (int a = 0; a < 3; a++) {
            int j = nbr[(2 * a) * N + g];
            double r = 0.0; uint8_t mode = 0, fail = 0;
            if (j >= 0) {
                double f = (double)kc[(2 * a) * N + g] * (phi[g] - phi[j]);
                double u = a == 1 ? 1.0 : 0.0;
                double up = f > 0.0 ? (u > 0.0 ? u : 0.0) : (-u > 0.0 ? -u : 0.0);
                float ti = TENSd[mat[g]], tj = TENSd[mat[j]];
                double tens = (double)(ti < tj ? ti : tj);
                double capm = 1.0 - (1.0 - tens) * up;
                r = fabs(f) / ((double)hp[4 * g + a] * capm);
                mode = a == 1 ? (f > 0.0 ? SF_MODE_TENSION : SF_MODE_COMPRESSION) : SF_MODE_SHEAR;
                fail = r > 1.0;
            }
            V.ratio[4 * g + a] = r; V.mode[4 * g + a] = mode; V.fail[4 * g + a] = fail;
        }
        double r = 0.0; uint8_t mode = 0, fail = 0;
        if (footing[g]) {
            r = (double)gk[g] * phi[g] / (double)hp[4 * g + 3];
            mode = SF_MODE_GROUND; fail = r > 1.0;
        }
        V.ratio[4 * g + 3] = r; V.mode[4 * g + 3] = mode; V.fail[4 * g + 3] = fail;
    }
}

float bond_cap(int mi, int mj, int axis) { float a = axis == 1 ? CAPV[mi] : CAPH[mi], b = axis == 1 ? CAPV[mj] : CAPH[mj]; return a < b ? a : b; }
float cond(int mi, int mj, int axis) { float a = STIFF[mi], b = STIFF[mj]; return (a < b ? a : b) * (axis == 1 ? 1.0f : 0.3f); }

// finalize_brick (src/lib.rs:668-684) for a standing structure: neighbour table and bond capacities
void build_topology(Struct& S) {
    const int n = S.n;
    std::unordered_map<int64_t, int32_t> at;
    at.reserve((size_t)n * 2);
    auto pack = [](int x, int y, int z) { return (((int64_t)x + 4096) << 32) | (((int64_t)y + 4096) << 16) | ((int64_t)z + 4096); };
    for (int i = 0; i < n; i++) at[pack(S.xyz[3 * i], S.xyz[3 * i + 1], S.xyz[3 * i + 2])] = i;
    S.nb6.assign(6 * (size_t)n, -1);
    S.bond.assign(4 * (size_t)n, 0.0f);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 6; d++) {
            auto it = at.find(pack(S.xyz[3 * i] + DIRS[d][0], S.xyz[3 * i + 1] + DIRS[d][1], S.xyz[3 * i + 2] + DIRS[d][2]));
            if (it != at.end()) S.nb6[6 * i + d] = it->second;
        }
    for (int i = 0; i < n; i++) {
        for (int a = 0; a < 3; a++) {
            int j = S.nb6[6 * i + 2 * a];
            bool gap = j >= 0 && S.infill[i] != S.infill[j] && (a != 1 || S.infill[i] == 1);
            S.bond[4 * i + a] = (j >= 0 && !gap) ? bond_cap(S.mat[i], S.mat[j], a) : 0.0f;
        }
        S.bond[4 * i + 3] = S.xyz[3 * i + 1] == 0 ? CAPV[S.mat[i]] * 2.0f : 0.0f;
    }
}
// linked (src/lib.rs:378-383): the neighbour in direction d joined by an intact bond
inline int linked(const Struct& S, int i, int d) {
    int j = S.nb6[6 * i + d];
    if (j < 0) return -1;
    float hp = (d & 1) == 0 ? S.bond[4 * i + (d >> 1)] : S.bond[4 * j + (d >> 1)];
    return hp > 0.0f ? j : -1;
}
inline bool footing(const Struct& S, int i) { return S.xyz[3 * i + 1] == 0 && S.bond[4 * i + 3] > 0.0f; }
// conn_struct's grounded search (src/lib.rs:1085-1098)
void grounded(Struct& S) {
    S.grounded.assign(S.n, 0);
    std::vector<int> q;
    q.reserve(S.n);
    for (int i = 0; i < S.n; i++) if (footing(S, i)) { S.grounded[i] = 1; q.push_back(i); }
    for (size_t h = 0; h < q.size(); h++) {
        int i = q[h];
        for (int d = 0; d < 6; d++) { int j = linked(S, i, d); if (j >= 0 && !S.grounded[j]) { S.grounded[j] = 1; q.push_back(j); } }
    }
    S.ngrounded = (int)q.size();
}

struct HostLoad {
    int dev = 0;
    std::vector<Struct> st;
    bool dirty = true;
    size_t N = 0;
    // device arrays
    int32_t* nbr = nullptr; float* kc = nullptr; float* sk32 = nullptr; double* sk64 = nullptr; float* b = nullptr;
    uint8_t *act = nullptr, *mat = nullptr, *foot = nullptr; float* hp = nullptr; float* gk = nullptr;
    double* phi = nullptr; float* delta = nullptr; float* r32 = nullptr;
    DevInfo* info = nullptr; int* osw = nullptr; double* ores = nullptr;
    uint4* row = nullptr; float* ktab = nullptr; bool pk = false;  // packed sweep rows (LoadDev::row)
    int32_t* order = nullptr;                                       // LoadDev::order
    double* vratio = nullptr; uint8_t *vmode = nullptr, *vfail = nullptr;
    bool verdicts_fresh = false;
    float last_ms = 0.0f;
    int64_t last_sweeps32 = 0, last_sweeps64 = 0;
    std::vector<void*> allocs;
    // a solve in flight on the handle's own stream (sf_load_solve_begin ... sf_load_solve_end)
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool pending = false;
    double tol = 0.0;
    int inner = 32;
    bool use_packed = true;  // sf_load_set_packed
    std::vector<DevInfo> hinfo;
    ~HostLoad() {
        if (pending && e1) (void)hipEventSynchronize(e1);
        free_dev();
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void free_dev() { for (void* p : allocs) (void)hipFree(p); allocs.clear(); }
    template <class T> int alloc(T*& p, size_t n) {
        void* q = nullptr;
        SF_HIP(hipMalloc(&q, sizeof(T) * (n ? n : 1)));
        allocs.push_back(q);
        p = (T*)q;
        return 0;
    }
};

// Every structure laid out for the device: red bricks then black, neighbour indices remapped, packed rows coded
// against the batch's table of distinct conductances. Host arrays only; the handle path and the allocation-free path
// copy them to their own device memory.
struct Layout {
    size_t N = 0;
    std::vector<int32_t> nbr;
    std::vector<float> kc, sk32, b, hp, gk;
    std::vector<double> sk64;
    std::vector<uint8_t> act, mat, foot;
    std::vector<DevInfo> info;
    std::vector<uint4> row;
    std::vector<float> ktab;
    std::vector<int32_t> order;
    bool pk = true;
};
void make_layout(std::vector<Struct>& st, Layout& O) {
    size_t N = 0;
    for (auto& S : st) { S.base = (int)N; N += (size_t)S.n; }
    O.N = N;
    O.nbr.assign(6 * N, -1);
    O.kc.assign(6 * N, 0.0f); O.sk32.assign(N, 0.0f); O.b.assign(N, 0.0f); O.hp.assign(4 * N, 0.0f); O.gk.assign(N, 0.0f);
    O.sk64.assign(N, 0.0);
    O.act.assign(N, 0); O.mat.assign(N, 0); O.foot.assign(N, 0);
    O.info.assign(st.size(), DevInfo{});
    O.row.assign(N, make_uint4(0u, 0u, 0u, 0u));
    O.ktab.assign(1, 0.0f);  // code 0: no bond
    std::unordered_map<uint32_t, uint32_t> kcode{{0u, 0u}};
    O.pk = true;
    auto& nbr = O.nbr; auto& kc = O.kc; auto& row = O.row; auto& ktab = O.ktab;
    for (size_t s = 0; s < st.size(); s++) {
        Struct& S = st[s];
        S.perm.clear(); S.inv.assign(S.n, 0);
        for (int i = 0; i < S.n; i++) if (((S.xyz[3 * i] + S.xyz[3 * i + 1] + S.xyz[3 * i + 2]) & 1) == 0) S.perm.push_back(i);
        S.nred = (int)S.perm.size();
        for (int i = 0; i < S.n; i++) if (((S.xyz[3 * i] + S.xyz[3 * i + 1] + S.xyz[3 * i + 2]) & 1) != 0) S.perm.push_back(i);
        for (int k = 0; k < S.n; k++) S.inv[S.perm[k]] = k;
        double bmax = 0.0;
        for (int k = 0; k < S.n; k++) {
            int i = S.perm[k];
            size_t g = (size_t)S.base + k;
            int mi = S.mat[i];
            O.mat[g] = (uint8_t)mi;
            O.act[g] = S.grounded[i];
            for (int a = 0; a < 4; a++) O.hp[4 * g + a] = S.bond[4 * i + a];
            O.foot[g] = footing(S, i);
            O.gk[g] = GROUNDK * STIFF[mi];
            float s32 = 0.0f; double s64 = 0.0;
            for (int d = 0; d < 6; d++) {
                int j = linked(S, i, d);
                if (j < 0) continue;
                float k = cond(mi, S.mat[j], d >> 1);
                nbr[d * N + g] = S.base + S.inv[j];
                kc[d * N + g] = k;
                s32 += k; s64 += (double)k;
            }
            if (O.foot[g]) { s32 += GROUNDK * STIFF[mi]; s64 += (double)(GROUNDK * STIFF[mi]); }
            O.sk32[g] = s32; O.sk64[g] = s64;
            if (O.act[g]) {
                O.b[g] = MASS[mi] + (S.ext.empty() ? 0.0f : S.ext[i]);
                bmax = std::max(bmax, (double)std::fabs(O.b[g]));
            }
        }
        // packed rows of a structure whose indices fit 16 bits, conductances coded in the batch's table
        if (S.n < (int)PK_NONE) {
            for (int k = 0; k < S.n; k++) {
                size_t g = (size_t)S.base + k;
                uint32_t rel[6], codes = O.act[g] ? (1u << 30) : 0u;
                for (int d = 0; d < 6; d++) {
                    int j = nbr[d * N + g];
                    rel[d] = j < 0 ? PK_NONE : (uint32_t)(j - S.base);
                    if (j < 0) continue;
                    uint32_t bits;
                    memcpy(&bits, &kc[d * N + g], 4);
                    auto it = kcode.find(bits);
                    uint32_t c;
                    if (it != kcode.end()) c = it->second;
                    else if (ktab.size() < (size_t)PK_CODES) { c = (uint32_t)ktab.size(); kcode[bits] = c; ktab.push_back(kc[d * N + g]); }
                    else { O.pk = false; c = 0; }
                    codes |= c << (5 * d);
                }
                row[g] = make_uint4(rel[0] | rel[1] << 16, rel[2] | rel[3] << 16, rel[4] | rel[5] << 16, codes);
            }
        }
        O.info[s].base = S.base; O.info[s].n = S.n; O.info[s].nred = S.nred; O.info[s].bmax = bmax;
        O.info[s].omega = S.omega;
        O.info[s].max_sweeps = 0;
    }
    O.order.resize(st.size());
    for (size_t s = 0; s < O.order.size(); s++) O.order[s] = (int32_t)s;
    std::stable_sort(O.order.begin(), O.order.end(), [&](int32_t a, int32_t b) { return st[a].n > st[b].n; });
    ktab.resize(PK_CODES, 0.0f);
}

// a standing structure from its bricks: topology (finalize_brick), grounded bricks (conn_struct's search) and the
// over-relaxation for a lattice of its height and span, 2 / (1 + sin(pi / L)) held to [1.6, 1.95]
int make_struct(int n, const uint8_t* mat, const int32_t* xyz, const uint8_t* infill, const float* ext, Struct& S) {
    if (n <= 0 || !mat || !xyz || !infill) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    S.n = n;
    S.mat.assign(mat, mat + n); S.infill.assign(infill, infill + n); S.xyz.assign(xyz, xyz + 3 * (size_t)n);
    for (int i = 0; i < n; i++) if (S.mat[i] < 1 || S.mat[i] > 5) { sf::set_error(SF_E_ARG, "material must be 1..5"); return SF_E_ARG; }
    if (ext) {
        S.ext.assign(ext, ext + n);
        for (float e : S.ext) if (!std::isfinite(e)) { sf::set_error(SF_E_ARG, "non-finite load"); return SF_E_ARG; }
    }
    build_topology(S);
    grounded(S);
    int ext_max = 1;
    int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-(1 << 30), -(1 << 30), -(1 << 30)};
    for (int i = 0; i < n; i++) for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], S.xyz[3 * i + a]); hi[a] = std::max(hi[a], S.xyz[3 * i + a]); }
    for (int a = 0; a < 3; a++) ext_max = std::max(ext_max, hi[a] - lo[a] + 1);
    double om = 2.0 / (1.0 + std::sin(3.14159265358979 / (double)ext_max));
    S.omega = (float)std::min(1.95, std::max(1.6, om));
    return 0;
}
// the largest right-hand side over a structure's grounded bricks (the residual's scale)
double struct_bmax(const Struct& S) {
    double bmax = 0.0;
    for (int i = 0; i < S.n; i++) if (S.grounded[i]) bmax = std::max(bmax, (double)std::fabs(MASS[S.mat[i]] + (S.ext.empty() ? 0.0f : S.ext[i])));
    return bmax;
}
// dynamic LDS for the corrections: the largest structure of the batch that fits the workgroup's limit (at most
// PK_NONE - 1 bricks, so the packed rows' 16-bit indices cover every structure with its correction in LDS)
int lds_floats_for(const std::vector<Struct>& st, int dev, size_t* bytes) {
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lds_max <= 0) lds_max = 65536;
    int fit = std::min((int)PK_NONE - 1, std::max(0, (lds_max - 1024 - PK_CODES * (int)sizeof(float)) / (int)sizeof(float))), nmax = 0;
    for (const Struct& S : st) if (S.n <= fit) nmax = std::max(nmax, S.n);
    *bytes = sizeof(float) * (size_t)nmax;
    if (*bytes > 65536) (void)hipFuncSetAttribute((const void*)k_load_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)*bytes);
    return nmax;
}

// the handle's device arrays, allocated for the batch and filled from its layout
int upload(HostLoad* L) {
    L->free_dev();
    Layout O;
    make_layout(L->st, O);
    const size_t N = O.N;
    L->N = N;
    int r = 0;
    r |= L->alloc(L->nbr, 6 * N); r |= L->alloc(L->kc, 6 * N); r |= L->alloc(L->sk32, N); r |= L->alloc(L->sk64, N);
    r |= L->alloc(L->b, N); r |= L->alloc(L->act, N); r |= L->alloc(L->mat, N); r |= L->alloc(L->foot, N);
    r |= L->alloc(L->hp, 4 * N); r |= L->alloc(L->gk, N); r |= L->alloc(L->phi, N); r |= L->alloc(L->delta, N);
    r |= L->alloc(L->r32, N); r |= L->alloc(L->info, O.info.size()); r |= L->alloc(L->osw, O.info.size());
    r |= L->alloc(L->ores, O.info.size()); r |= L->alloc(L->vratio, 4 * N); r |= L->alloc(L->vmode, 4 * N); r |= L->alloc(L->vfail, 4 * N);
    r |= L->alloc(L->row, N); r |= L->alloc(L->ktab, PK_CODES); r |= L->alloc(L->order, O.info.size());
    if (r) return r;
    L->pk = O.pk;
    SF_HIP(hipMemcpy(L->order, O.order.data(), sizeof(int32_t) * O.order.size(), hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->row, O.row.data(), sizeof(uint4) * N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->ktab, O.ktab.data(), sizeof(float) * PK_CODES, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->nbr, O.nbr.data(), sizeof(int32_t) * 6 * N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->kc, O.kc.data(), sizeof(float) * 6 * N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->sk32, O.sk32.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->sk64, O.sk64.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->b, O.b.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->act, O.act.data(), N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->mat, O.mat.data(), N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->foot, O.foot.data(), N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->hp, O.hp.data(), sizeof(float) * 4 * N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->gk, O.gk.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    SF_HIP(hipMemcpy(L->info, O.info.data(), sizeof(DevInfo) * O.info.size(), hipMemcpyHostToDevice));
    L->dirty = false;
    return 0;
}

// ---------------------------------------------------------------- the allocation-free solve
// One structure's solve with every device array in a fixed, statically allocated scratch array instead of memory
// allocated for the call: the layout (stencil rows, diagonals, right-hand side, flags), the FP64 field, the FP32
// residual and correction, the verdicts and the result words. Each GPU holds its own copy of `scratch` from the
// first use until the process ends. Host code never passes `scratch` itself to a kernel: its name is a host-side
// stand-in; the device address comes from hipGetSymbolAddress, once per device, and is kept in g_scratch.
constexpr int STATIC_MAX_BRICKS = 65000;
constexpr size_t STATIC_BYTES_PER_BRICK = 160;  // 159 bytes of arrays per brick (see carve_static), rounded up
constexpr size_t STATIC_SCRATCH_BYTES = (size_t)STATIC_MAX_BRICKS * STATIC_BYTES_PER_BRICK + (64u << 10);
static __device__ __align__(256) uint8_t scratch[STATIC_SCRATCH_BYTES];

constexpr int STATIC_MAX_DEVICES = 64;
struct ScratchDev { std::mutex mu; void* ptr = nullptr; };
ScratchDev g_scratch[STATIC_MAX_DEVICES];

__global__ void k_zero64(double* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}
// the arrays of one structure of n bricks, 256-byte aligned, out of the scratch array at `base`; returns the bytes
struct StaticArrays {
    int32_t* nbr; float *kc, *sk32, *b, *hp, *gk, *delta, *r32, *ktab; double *sk64, *phi, *vratio, *ores;
    uint8_t *act, *mat, *foot, *vmode, *vfail; uint4* row; DevInfo* info; int* osw; int32_t* order;
};
size_t carve_static(uint8_t* base, size_t n, StaticArrays& A) {
    size_t off = 0;
    auto take = [&](auto*& p, size_t count) {
        off = (off + 255) & ~(size_t)255;
        p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(base + off);
        off += sizeof(*p) * (count ? count : 1);
    };
    take(A.row, n); take(A.nbr, 6 * n); take(A.kc, 6 * n); take(A.hp, 4 * n); take(A.vratio, 4 * n);
    take(A.sk64, n); take(A.phi, n); take(A.sk32, n); take(A.b, n); take(A.gk, n); take(A.delta, n); take(A.r32, n);
    take(A.vmode, 4 * n); take(A.vfail, 4 * n); take(A.act, n); take(A.mat, n); take(A.foot, n);
    take(A.ktab, PK_CODES); take(A.info, 1); take(A.osw, 1); take(A.ores, 1); take(A.order, 1);
    return off;
}
static_assert((size_t)STATIC_MAX_BRICKS * 159 + 23 * 256 + PK_CODES * 4 + 64 <= STATIC_SCRATCH_BYTES, "scratch too small");

}  // namespace

struct sf_load { HostLoad h; };

extern "C" {

// (structures can be added and their topology read without a GPU; solving needs one)
sf_load* sf_load_create(void) {
    sf_load* L = new sf_load();
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) (void)hipGetDevice(&L->h.dev);
    return L;
}
void sf_load_destroy(sf_load* L) { delete L; }
int sf_load_count(sf_load* L) { return L ? (int)L->h.st.size() : SF_E_ARG; }
int sf_load_set_packed(sf_load* L, int on) {
    if (!L) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    if (L->h.pending) { int r = sf_load_solve_end(L); if (r < 0) return r; }
    L->h.use_packed = on != 0;
    return 0;
}

int sf_load_add(sf_load* L, int n, const uint8_t* mat, const int32_t* xyz, const uint8_t* infill, const float* ext) {
    if (!L) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    Struct S;
    int r = make_struct(n, mat, xyz, infill, ext, S);
    if (r) return r;
    L->h.st.push_back(std::move(S));
    L->h.dirty = true;
    return (int)L->h.st.size() - 1;
}

int sf_load_solve_begin(sf_load* L, float omega, double rel_tol, int max_sweeps) {
    if (!L || !(rel_tol > 0.0) || max_sweeps <= 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    HostLoad* H = &L->h;
    if (H->pending) { int r = sf_load_solve_end(L); if (r < 0) return r; }
    H->verdicts_fresh = false;
    if (H->st.empty()) return 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { sf::set_error(SF_E_NODEV, "no GPU visible"); return SF_E_NODEV; }
    SF_HIP(hipSetDevice(H->dev));
    if (!H->stream) {
        SF_HIP(hipStreamCreateWithFlags(&H->stream, hipStreamNonBlocking));
        SF_HIP(hipEventCreate(&H->e0)); SF_HIP(hipEventCreate(&H->e1));
    }
    if (omega > 0.0f) for (auto& S : H->st) S.omega = omega;
    if (H->dirty) { int r = upload(H); if (r) return r; }
    H->hinfo.resize(H->st.size());
    for (size_t s = 0; s < H->st.size(); s++) {
        const Struct& S = H->st[s];
        H->hinfo[s] = DevInfo{S.base, S.n, S.nred, max_sweeps, S.omega, struct_bmax(S)};
    }
    SF_HIP(hipMemcpy(H->info, H->hinfo.data(), sizeof(DevInfo) * H->hinfo.size(), hipMemcpyHostToDevice));
    SF_HIP(hipMemsetAsync(H->phi, 0, sizeof(double) * H->N, H->stream));
    LoadDev D;
    D.nstruct = (int)H->st.size(); D.info = H->info; D.nbr = H->nbr; D.kc = H->kc; D.sk32 = H->sk32; D.sk64 = H->sk64;
    D.b = H->b; D.act = H->act; D.phi = H->phi; D.delta = H->delta; D.r32 = H->r32; D.out_sweeps = H->osw;
    D.out_resid = H->ores; D.N = H->N; D.tol = rel_tol; D.inner = H->inner;
    D.row = H->row; D.ktab = H->ktab; D.pk_codes = H->pk && H->use_packed ? 1 : 0; D.order = H->order;
    H->tol = rel_tol;
    size_t lds_bytes = 0;
    D.lds_floats = lds_floats_for(H->st, H->dev, &lds_bytes);
    SF_HIP(hipEventRecord(H->e0, H->stream));
    hipLaunchKernelGGL(k_load_solve, dim3(D.nstruct), dim3(LT), lds_bytes, H->stream, D);
    SF_HIP(hipGetLastError());
    // verdicts from the converged field
    Verd V{H->vratio, H->vmode, H->vfail};
    hipLaunchKernelGGL(k_verdicts, dim3(sf_grid(H->N)), dim3(256), 0, H->stream, H->nbr, H->kc, H->hp, H->mat, H->foot, H->phi, H->gk, H->N, V);
    SF_HIP(hipGetLastError());
    SF_HIP(hipEventRecord(H->e1, H->stream));
    H->pending = true;
    return 0;
}

int sf_load_solve_ready(sf_load* L) {
    if (!L) return SF_E_ARG;
    HostLoad* H = &L->h;
    if (!H->pending) return 1;
    hipError_t e = hipEventQuery(H->e1);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    sf::set_error(SF_E_HIP, hipGetErrorString(e));
    return SF_E_HIP;
}

int sf_load_solve_end(sf_load* L) {
    if (!L) return SF_E_ARG;
    HostLoad* H = &L->h;
    if (!H->pending) return H->verdicts_fresh ? (int)std::count_if(H->st.begin(), H->st.end(), [&](const Struct& S) { return S.resid <= H->tol; }) : 0;
    H->pending = false;
    SF_HIP(hipEventSynchronize(H->e1));
    float ms = 0.0f;
    SF_HIP(hipEventElapsedTime(&ms, H->e0, H->e1));
    H->last_ms = ms;
    std::vector<int> sw(H->st.size());
    std::vector<double> rs(H->st.size());
    SF_HIP(hipMemcpy(sw.data(), H->osw, sizeof(int) * sw.size(), hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(rs.data(), H->ores, sizeof(double) * rs.size(), hipMemcpyDeviceToHost));
    int ok = 0;
    H->last_sweeps32 = 0; H->last_sweeps64 = 0;
    for (size_t s = 0; s < H->st.size(); s++) {
        H->st[s].sweeps = sw[s]; H->st[s].resid = rs[s];
        if (rs[s] <= H->tol) ok++;
        H->last_sweeps32 += (int64_t)sw[s] * H->st[s].n;
        H->last_sweeps64 += (int64_t)((sw[s] + H->inner - 1) / H->inner + 1) * H->st[s].n;
    }
    H->verdicts_fresh = true;
    return ok;
}

int sf_load_solve(sf_load* L, float omega, double rel_tol, int max_sweeps) {
    int r = sf_load_solve_begin(L, omega, rel_tol, max_sweeps);
    if (r < 0) return r;
    return sf_load_solve_end(L);
}

int sf_load_info(sf_load* L, int sid, int32_t* sweeps, double* resid, int32_t* grounded_out) {
    if (!L || sid < 0 || sid >= (int)L->h.st.size()) { sf::set_error(SF_E_ARG, "bad structure id"); return SF_E_ARG; }
    const Struct& S = L->h.st[sid];
    if (sweeps) *sweeps = S.sweeps;
    if (resid) *resid = S.resid;
    if (grounded_out) *grounded_out = S.ngrounded;
    return 0;
}

int sf_load_field(sf_load* L, int sid, double* phi) {
    if (!L || !phi || sid < 0 || sid >= (int)L->h.st.size() || L->h.dirty || L->h.pending) { sf::set_error(SF_E_ARG, "bad structure id or not solved"); return SF_E_ARG; }
    const Struct& S = L->h.st[sid];
    std::vector<double> d(S.n);
    SF_HIP(hipMemcpy(d.data(), L->h.phi + S.base, sizeof(double) * S.n, hipMemcpyDeviceToHost));
    for (int k = 0; k < S.n; k++) phi[S.perm[k]] = d[k];
    return 0;
}

int sf_load_verdicts(sf_load* L, int sid, double* ratio4, uint8_t* mode4, uint8_t* fail4) {
    if (!L || sid < 0 || sid >= (int)L->h.st.size() || !L->h.verdicts_fresh) { sf::set_error(SF_E_ARG, "bad structure id or not solved"); return SF_E_ARG; }
    const Struct& S = L->h.st[sid];
    std::vector<double> r(4 * (size_t)S.n);
    std::vector<uint8_t> m(4 * (size_t)S.n), f(4 * (size_t)S.n);
    SF_HIP(hipMemcpy(r.data(), L->h.vratio + 4 * (size_t)S.base, sizeof(double) * 4 * S.n, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(m.data(), L->h.vmode + 4 * (size_t)S.base, 4 * (size_t)S.n, hipMemcpyDeviceToHost));
    SF_HIP(hipMemcpy(f.data(), L->h.vfail + 4 * (size_t)S.base, 4 * (size_t)S.n, hipMemcpyDeviceToHost));
    for (int k = 0; k < S.n; k++) {
        int i = S.perm[k];
        for (int a = 0; a < 4; a++) {
            if (ratio4) ratio4[4 * i + a] = r[4 * k + a];
            if (mode4) mode4[4 * i + a] = m[4 * k + a];
            if (fail4) fail4[4 * i + a] = f[4 * k + a];
        }
    }
    return 0;
}

int sf_load_topology(sf_load* L, int sid, int32_t* nb6, float* bond4) {
    if (!L || sid < 0 || sid >= (int)L->h.st.size()) { sf::set_error(SF_E_ARG, "bad structure id"); return SF_E_ARG; }
    const Struct& S = L->h.st[sid];
    if (nb6) std::memcpy(nb6, S.nb6.data(), sizeof(int32_t) * S.nb6.size());
    if (bond4) std::memcpy(bond4, S.bond.data(), sizeof(float) * S.bond.size());
    return 0;
}

int sf_load_last_timing(sf_load* L, float* ms, int64_t* sweeps32, int64_t* sweeps64, int64_t* bricks) {
    if (!L) return SF_E_ARG;
    if (ms) *ms = L->h.last_ms;
    if (sweeps32) *sweeps32 = L->h.last_sweeps32;
    if (sweeps64) *sweeps64 = L->h.last_sweeps64;
    if (bricks) *bricks = (int64_t)L->h.N;
    return 0;
}

int sf_load_solve_static(int n, const uint8_t* mat, const int32_t* xyz, const uint8_t* infill, const float* ext,
                         float omega, double rel_tol, int max_sweeps, void* stream, double* phi, double* ratio4,
                         uint8_t* mode4, uint8_t* fail4, int32_t* sweeps, double* resid) {
    if (!phi || !(rel_tol > 0.0) || max_sweeps <= 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    if (n > STATIC_MAX_BRICKS) {
        sf::set_error(SF_E_CAP, "sf_load_solve_static: " + std::to_string(n) + " bricks, at most " + std::to_string(STATIC_MAX_BRICKS));
        return SF_E_CAP;
    }
    std::vector<Struct> st(1);
    int r = make_struct(n, mat, xyz, infill, ext, st[0]);
    if (r) return r;
    if (omega > 0.0f) st[0].omega = omega;
    Layout O;
    make_layout(st, O);
    int ndev = 0, dev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { sf::set_error(SF_E_NODEV, "no GPU visible"); return SF_E_NODEV; }
    SF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= STATIC_MAX_DEVICES) { sf::set_error(SF_E_ARG, "device index beyond the scratch table"); return SF_E_ARG; }
    ScratchDev& G = g_scratch[dev];
    std::lock_guard<std::mutex> lock(G.mu);  // one call at a time per GPU: they share its scratch
    if (!G.ptr) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) { sf::set_error(SF_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)); return SF_E_HIP; }
        G.ptr = p;
    }
    const size_t N = O.N;
    StaticArrays A;
    if (carve_static(static_cast<uint8_t*>(G.ptr), N, A) > STATIC_SCRATCH_BYTES) { sf::set_error(SF_E_CAP, "scratch layout"); return SF_E_CAP; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevInfo info = O.info[0];
    info.max_sweeps = max_sweeps;
    info.bmax = struct_bmax(st[0]);
    const int32_t order0 = 0;
    SF_HIP(hipMemcpyAsync(A.row, O.row.data(), sizeof(uint4) * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.ktab, O.ktab.data(), sizeof(float) * PK_CODES, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.nbr, O.nbr.data(), sizeof(int32_t) * 6 * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.kc, O.kc.data(), sizeof(float) * 6 * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.sk32, O.sk32.data(), sizeof(float) * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.sk64, O.sk64.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.b, O.b.data(), sizeof(float) * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.act, O.act.data(), N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.mat, O.mat.data(), N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.foot, O.foot.data(), N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.hp, O.hp.data(), sizeof(float) * 4 * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.gk, O.gk.data(), sizeof(float) * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.info, &info, sizeof(DevInfo), hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(A.order, &order0, sizeof(int32_t), hipMemcpyHostToDevice, s));
    // the field starts at zero (a fill kernel, so a caller may capture the call in a graph)
    hipLaunchKernelGGL(k_zero64, dim3(sf_grid(N)), dim3(256), 0, s, A.phi, N);
    LoadDev D;
    D.nstruct = 1; D.info = A.info; D.nbr = A.nbr; D.kc = A.kc; D.sk32 = A.sk32; D.sk64 = A.sk64;
    D.b = A.b; D.act = A.act; D.phi = A.phi; D.delta = A.delta; D.r32 = A.r32; D.out_sweeps = A.osw;
    D.out_resid = A.ores; D.N = N; D.tol = rel_tol; D.inner = 32;
    D.row = A.row; D.ktab = A.ktab; D.pk_codes = O.pk ? 1 : 0; D.order = A.order;
    size_t lds_bytes = 0;
    D.lds_floats = lds_floats_for(st, dev, &lds_bytes);
    hipLaunchKernelGGL(k_load_solve, dim3(1), dim3(LT), lds_bytes, s, D);
    SF_HIP(hipGetLastError());
    Verd V{A.vratio, A.vmode, A.vfail};
    hipLaunchKernelGGL(k_verdicts, dim3(sf_grid(N)), dim3(256), 0, s, A.nbr, A.kc, A.hp, A.mat, A.foot, A.phi, A.gk, N, V);
    SF_HIP(hipGetLastError());
    std::vector<double> d(N), vr(ratio4 ? 4 * N : 0);
    std::vector<uint8_t> vm(mode4 ? 4 * N : 0), vf(fail4 ? 4 * N : 0);
    int osw = 0;
    double ores = 0.0;
    SF_HIP(hipMemcpyAsync(d.data(), A.phi, sizeof(double) * N, hipMemcpyDeviceToHost, s));
    if (ratio4) SF_HIP(hipMemcpyAsync(vr.data(), A.vratio, sizeof(double) * 4 * N, hipMemcpyDeviceToHost, s));
    if (mode4) SF_HIP(hipMemcpyAsync(vm.data(), A.vmode, 4 * N, hipMemcpyDeviceToHost, s));
    if (fail4) SF_HIP(hipMemcpyAsync(vf.data(), A.vfail, 4 * N, hipMemcpyDeviceToHost, s));
    SF_HIP(hipMemcpyAsync(&osw, A.osw, sizeof(int), hipMemcpyDeviceToHost, s));
    SF_HIP(hipMemcpyAsync(&ores, A.ores, sizeof(double), hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    const Struct& S = st[0];
    for (int k = 0; k < S.n; k++) {
        const int i = S.perm[k];
        phi[i] = d[k];
        for (int a = 0; a < 4; a++) {
            if (ratio4) ratio4[4 * i + a] = vr[4 * k + a];
            if (mode4) mode4[4 * i + a] = vm[4 * k + a];
            if (fail4) fail4[4 * i + a] = vf[4 * k + a];
        }
    }
    if (sweeps) *sweeps = osw;
    if (resid) *resid = ores;
    return ores <= rel_tol ? 1 : 0;
}
int64_t sf_load_static_capacity(void) { return STATIC_MAX_BRICKS; }

}  // extern "C"
