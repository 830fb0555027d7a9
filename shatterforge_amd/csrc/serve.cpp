// Served what-if stream: demolition queries answered by continuous batching on one GPU, behind the C ABI
// (sf_serve_*). A query is a standing structure (lattice bricks) plus charges; it is answered in three parts, all on
// the GPU: the load-path solve of the intact structure (FP32 sweeps, FP64 residual and refinement; sf_load), whose
// converged field becomes the scene's load field; the charges (explode_at, src/lib.rs:2697-2716); then `frames`
// frames of the brick step (src/lib.rs:3639-3721). The answer: joints failed by mode, the fragment table per dominant
// material (tools/debug_exports.rs:41-56), bricks standing and pieces made.
//
// Continuous batching: every live query is a scene of one lattice handle and all live scenes advance together in one
// batched frame. Queries that arrive while others are on the GPU are admitted when a scene slot is free and the
// bytes their scene will hold on the device (sf_lat_bytes_resident's rule at the query's own size) fit the HBM
// budget beside the live ones; an admitted query's scene takes brick and body ranges of its own size from the
// handle's pool (sf_lat_scene_place) and gives them back when it retires. The admitted queries' load-path solves run
// as one batch on the load handle's own stream while the live scenes keep stepping (at most `max_solving` batches at
// once). Every order in a frame is scene local, so a query's answer is the same alone and in any mix.
#include "../../include/sf.h"
#include "sf_host.h"
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <memory>
#include <vector>

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Query {
    uint64_t id = 0;
    int n = 0;
    std::vector<uint8_t> mat, infill;
    std::vector<int32_t> xyz;
    std::vector<float> charges;   // 5 per charge: x, y, z, radius, power
    uint32_t seed = 1;
    int64_t bytes = 0;
    double t_submit = 0.0, t_admit = 0.0;
    int slot = -1, frame = 0, sweeps = 0;
    double resid = 1.0;
};

struct Batch {
    sf_load* load = nullptr;
    std::vector<std::unique_ptr<Query>> qs;
};

}  // namespace

struct sf_serve {
    sf_lat* lat = nullptr;
    int slots = 0, slot_bodies = 0, frames = 0, max_solving = 3, margin = 64;
    int64_t budget = 0, in_use = 0, brick_steps = 0, frames_run = 0;
    double ms[4] = {0, 0, 0, 0};   // wall ms in admission, activation, frames, retirement
    std::deque<int> free_slots;
    std::deque<std::unique_ptr<Query>> queue;
    std::vector<Batch> pending;
    std::vector<std::unique_ptr<Query>> live;   // by slot
    std::vector<sf_answer> done;
    ~sf_serve() {
        for (Batch& b : pending) if (b.load) sf_load_destroy(b.load);
        if (lat) sf_lat_destroy(lat);
    }
};

namespace {

// the resident bytes a query's scene will hold: its bricks (plus a margin for the pieces' bookkeeping) and the
// slot's bodies, by the handle's own per-unit tallies
int64_t query_bytes(sf_serve* S, int n) { return sf_lat_bytes_for(S->lat, n + S->margin, S->slot_bodies, 2); }

int admit(sf_serve* S) {
    if ((int)S->pending.size() >= S->max_solving) return 0;
    Batch b;
    while (!S->queue.empty() && !S->free_slots.empty()) {
        Query& q = *S->queue.front();
        if (S->in_use + q.bytes > S->budget) break;   // first come, first served: later arrivals wait behind it
        const int slot = S->free_slots.front();
        int r = sf_lat_scene_place(S->lat, slot, q.n + S->margin, S->slot_bodies, q.seed);
        if (r == SF_E_CAP) break;                     // no free range that large yet: wait for a retirement
        if (r < 0) return r;
        S->free_slots.pop_front();
        S->in_use += q.bytes;
        q.slot = slot;
        q.t_admit = now_ms();
        b.qs.push_back(std::move(S->queue.front()));
        S->queue.pop_front();
    }
    if (b.qs.empty()) return 0;
    b.load = sf_load_create();
    if (!b.load) return SF_E_HIP;
    for (auto& q : b.qs) {
        int r = sf_load_add(b.load, q->n, q->mat.data(), q->xyz.data(), q->infill.data(), nullptr);
        if (r < 0) { sf_load_destroy(b.load); return r; }
    }
    int r = sf_load_solve_begin(b.load, 0.0f, 1e-6, 200000);
    if (r < 0) { sf_load_destroy(b.load); return r; }
    S->pending.push_back(std::move(b));
    return 0;
}

// batches whose solve has finished (all of them when `wait`) join the frame: scene, converged field, charges
int activate(sf_serve* S, bool wait) {
    std::vector<Batch> keep;
    for (Batch& b : S->pending) {
        int ready = wait ? 1 : sf_load_solve_ready(b.load);
        if (ready < 0) return ready;
        if (!ready) { keep.push_back(std::move(b)); continue; }
        int r = sf_load_solve_end(b.load);
        if (r < 0) return r;
        std::vector<double> f64;
        std::vector<float> f32;
        for (size_t k = 0; k < b.qs.size(); k++) {
            Query& q = *b.qs[k];
            const int st = sf_lat_add_structure(S->lat, q.slot, q.n, q.xyz.data(), q.mat.data(), q.infill.data());
            if (st < 0) return st;
            if ((r = sf_lat_finalize(S->lat, q.slot)) < 0) return r;
            int32_t sw = 0;
            sf_load_info(b.load, (int)k, &sw, &q.resid, nullptr);
            q.sweeps = sw;
            f64.resize(q.n); f32.resize(q.n);
            if ((r = sf_load_field(b.load, (int)k, f64.data())) < 0) return r;
            for (int i = 0; i < q.n; i++) f32[i] = (float)f64[i];
            if ((r = sf_lat_set_field(S->lat, q.slot, st, f32.data())) < 0) return r;
            for (size_t c = 0; c + 4 < q.charges.size(); c += 5)
                if ((r = sf_lat_explode(S->lat, q.slot, q.charges[c], q.charges[c + 1], q.charges[c + 2], q.charges[c + 3], q.charges[c + 4])) < 0) return r;
            q.frame = 0;
            S->live[q.slot] = std::move(b.qs[k]);
        }
        sf_load_destroy(b.load);
        b.load = nullptr;
    }
    S->pending = std::move(keep);
    return 0;
}

int retire(sf_serve* S) {
    for (int slot = 0; slot < S->slots; slot++) {
        if (!S->live[slot] || S->live[slot]->frame < S->frames) continue;
        Query& q = *S->live[slot];
        sf_answer a;
        std::memset(&a, 0, sizeof a);
        a.id = q.id;
        a.bricks = q.n;
        float st[16], ms[50];
        int r = 0;
        r |= sf_lat_stats(S->lat, slot, st);
        r |= sf_lat_matstats(S->lat, slot, ms);
        r |= sf_lat_failures(S->lat, slot, a.failed);
        if (r < 0) return SF_E_HIP;
        a.standing = (int32_t)st[0];
        a.pieces = (int32_t)st[2];
        std::memcpy(a.fragments, ms + 5, sizeof(float) * 25);   // materials 1..5
        a.load_sweeps = q.sweeps;
        a.load_resid = q.resid;
        const double t = now_ms();
        a.latency_ms = t - q.t_submit;
        a.queue_ms = q.t_admit - q.t_submit;
        S->done.push_back(a);
        S->in_use -= q.bytes;
        S->live[slot].reset();
        if ((r = sf_lat_scene_release(S->lat, slot)) < 0) return r;   // its ranges back to the pool
        S->free_slots.push_back(slot);
    }
    return 0;
}

int nlive(const sf_serve* S) {
    int n = 0;
    for (const auto& q : S->live) n += q ? 1 : 0;
    return n;
}

}  // namespace

extern "C" {

sf_serve* sf_serve_create(int slots, int pool_bricks, int slot_bodies, int frames, int64_t hbm_budget, int max_solving) {
    if (slots <= 0 || pool_bricks <= 0 || slot_bodies <= 0 || frames <= 0 || max_solving <= 0) {
        sf::set_error(SF_E_ARG, "sf_serve_create: bad argument"); return nullptr;
    }
    std::unique_ptr<sf_serve> S(new sf_serve());
    S->slots = slots; S->slot_bodies = slot_bodies; S->frames = frames; S->max_solving = max_solving;
    S->lat = sf_lat_create(pool_bricks + slots, slots * slot_bodies + slots, slots, 16, 8);
    if (!S->lat) return nullptr;
    for (int k = 0; k < slots; k++) {
        // every slot starts without ranges: a query's scene takes its own from the pool when it is admitted
        const int sc = sf_lat_scene_add(S->lat, 1, 1, 2, 1);
        if (sc < 0 || sf_lat_scene_release(S->lat, sc) < 0) return nullptr;
        S->free_slots.push_back(sc);
    }
    S->live.resize(slots);
    S->budget = hbm_budget > 0 ? hbm_budget : sf_lat_bytes_for(S->lat, pool_bricks, slots * slot_bodies, 2 * slots);
    return S.release();
}

void sf_serve_destroy(sf_serve* S) { delete S; }

int64_t sf_serve_query_bytes(sf_serve* S, int bricks) {
    if (!S || bricks <= 0) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    return query_bytes(S, bricks);
}

int sf_serve_submit(sf_serve* S, uint64_t id, int n, const int32_t* xyz, const uint8_t* mat, const uint8_t* infill,
                    int ncharge, const float* charges5, uint32_t seed) {
    if (!S || n <= 0 || !xyz || !mat || ncharge < 0 || (ncharge && !charges5)) { sf::set_error(SF_E_ARG, "sf_serve_submit: bad argument"); return SF_E_ARG; }
    std::unique_ptr<Query> q(new Query());
    q->id = id; q->n = n; q->seed = seed ? seed : 1;
    q->bytes = query_bytes(S, n);
    // refused now what could never be admitted: a scene larger than the budget or the handle's whole pool
    if (q->bytes > S->budget) {
        sf::set_error(SF_E_CAP, "query " + std::to_string(id) + ": " + std::to_string(q->bytes) + " resident bytes, the budget is " + std::to_string(S->budget));
        return SF_E_CAP;
    }
    q->mat.assign(mat, mat + n);
    q->infill.assign(infill ? infill : mat, (infill ? infill : mat) + n);
    if (!infill) std::fill(q->infill.begin(), q->infill.end(), 0);
    q->xyz.assign(xyz, xyz + 3 * (size_t)n);
    q->charges.assign(charges5, charges5 + 5 * (size_t)ncharge);
    for (uint8_t m : q->mat) if (m < 1 || m > 5) { sf::set_error(SF_E_ARG, "material must be 1..5"); return SF_E_ARG; }
    for (float c : q->charges) if (!(c == c) || c > 1e30f || c < -1e30f) { sf::set_error(SF_E_ARG, "non-finite charge"); return SF_E_ARG; }
    q->t_submit = now_ms();
    S->queue.push_back(std::move(q));
    return 0;
}

int sf_serve_tick(sf_serve* S, float dt) {
    if (!S) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    double t0 = now_ms();
    int r = admit(S);
    if (r < 0) return r;
    double t1 = now_ms();
    if ((r = activate(S, nlive(S) == 0)) < 0) return r;
    double t2 = now_ms();
    const int n = nlive(S);
    if (n) {
        if ((r = sf_lat_step(S->lat, dt)) < 0) return r;
        S->frames_run++;
        for (auto& q : S->live) if (q) { q->frame++; S->brick_steps += q->n; }
    }
    double t3 = now_ms();
    if ((r = retire(S)) < 0) return r;
    double t4 = now_ms();
    S->ms[0] += t1 - t0; S->ms[1] += t2 - t1; S->ms[2] += t3 - t2; S->ms[3] += t4 - t3;
    return nlive(S);
}

int sf_serve_poll(sf_serve* S, sf_answer* out, int cap) {
    if (!S || cap < 0 || (cap && !out)) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    const int n = std::min<int>(cap, (int)S->done.size());
    if (n) {
        std::memcpy(out, S->done.data(), sizeof(sf_answer) * (size_t)n);
        S->done.erase(S->done.begin(), S->done.begin() + n);
    }
    return n;
}

int sf_serve_timing(sf_serve* S, double* out4) {
    if (!S || !out4) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    for (int k = 0; k < 4; k++) out4[k] = S->ms[k];
    return 0;
}
int sf_serve_status(sf_serve* S, int64_t* out8) {
    if (!S || !out8) { sf::set_error(SF_E_ARG, "bad argument"); return SF_E_ARG; }
    int64_t solving = 0;
    for (const Batch& b : S->pending) solving += (int64_t)b.qs.size();
    out8[0] = (int64_t)S->queue.size(); out8[1] = solving; out8[2] = nlive(S); out8[3] = (int64_t)S->done.size();
    out8[4] = S->in_use; out8[5] = S->budget; out8[6] = S->brick_steps; out8[7] = S->frames_run;
    return 0;
}

sf_lat* sf_serve_world(sf_serve* S) { return S ? S->lat : nullptr; }

}  // extern "C"
