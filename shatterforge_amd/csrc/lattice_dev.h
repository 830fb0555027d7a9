// Device state of the lattice (legacy brick) engine: struct-of-arrays for bricks, bodies, contacts and the
// per-scene tables, shared by the kernels and the host engine in lattice.hip.
#pragma once
#include "lat_math.h"

namespace lat {

constexpr int NT = 256;
constexpr int KMAX = 32;  // colour limit
constexpr int CHAIN = -2;  // colour of a contact solved in its body's static chain (against static bricks, the ground
                           // or a sleeping body; generation order, one lane per body)

// device counters
enum : int {
    C_NCS = 0,      // contacts this substep
    C_NPLOUGH,      // plough candidates
    C_NIMPACT,      // impact candidates
    C_NKILL,        // bodies killed on the device (pinch, out of bounds)
    C_NEV,          // structural crush events
    C_NBW,          // bond writes
    C_NCOL,         // colour bucket counts follow at C_COLCNT
    C_NCTR0,
    C_COLCNT = 16,  // KMAX + 1 counts
    C_NCTR = C_COLCNT + KMAX + 1
};
constexpr int CTR_WORDS = 64;  // the counters' block; the scenes' first contacts follow it
static_assert(C_NCTR <= CTR_WORDS, "counter block");

// Per-scene values a substep or frame reads (uploaded by the host before each frame)
struct SceneK {
    uint32_t frame;
    float blast_t;
    int32_t gx0, gy0, gz0, gnx, gny, gnz;
    uint32_t goff;         // grid offset
    int32_t bd0, nbd;      // body slot range
    uint32_t dyn0, ndyn;   // dynamic slot range
    int32_t st0, nst;      // structure range (global indices)
    int32_t b0, nb;        // brick range
    int32_t hw;            // body slots in use (high-water mark)
};

struct LD {
    uint32_t nb, nbd, ndyn, nscene, nst, K, R, ccap, tmask;
    uint32_t nub;              // body slots in use: every scene's slots below its high-water mark
    uint32_t* ubody;
    // bricks
    uint8_t* mat;
    uint8_t* sid;
    uint8_t* infill;       // glass or marked infill (finalize's gap rule)
    int32_t* body;
    int4* bxyz;            // lattice x, y, z, scene
    uint32_t* nb6;
    float4* bond;
    float4* bmax;
    float4* local;
    float4* wpos;
    float* phi;
    float* ext;
    float* extn;
    float* stress;
    uint32_t* sup;
    // bodies
    int32_t* bd_scene;
    uint8_t* alive;
    uint8_t* asleep;
    uint8_t* wake;
    float4 *pos, *q, *v, *w, *vpre, *wpre, *vlast, *wlast;
    float4 *rot, *iil, *iiw;   // three rows each
    float *m, *im, *sleep, *rad, *cool, *lt, *resid, *maxr, *nmax;
    uint32_t *nc, *count, *boff, *bcnt;
    uint32_t *cfirst, *cend;   // a body's contacts as side a in the current substep: [cfirst, cend), contiguous
    uint32_t* dyn;             // dynamic slot -> brick (body-major, brick-list order)
    int32_t* dyn_body;
    // scenes and structures
    SceneK* sc;
    int32_t* grid;
    int32_t* st_scene;
    int32_t* st_lo;
    int32_t* st_hi;
    uint8_t* st_act;           // 1 for structures whose load field is solved this frame
    uint8_t* st_on;            // st_active > 0
    uint8_t* st_live;
    int32_t* st_nmin;          // sweeps asked for (0: none)
    float* st_resid;
    uint32_t* st_maxr;         // f32 bits, atomic max
    uint32_t* st_flags;        // bit0 conn, bit1 broke
    uint32_t* st_red;          // per structure: red bricks then black bricks (offsets in st_rb)
    int4* st_rb;               // red offset, red count, black offset, black count
    // per position of st_red: the brick's linked neighbours and conductances in link order, their sum with the
    // footing term and the source M + ext (sk < 0: not part of the structure any more)
    int4* sten_na;             // neighbours of directions 0-3 (-1: not linked)
    int2* sten_nb;             // directions 4-5
    float4* sten_ka;
    float2* sten_kb;
    float2* sten_s;            // sk, rhs
    // Packed rows of the structures whose potential fits the sweep workgroup's LDS (the full rows above serve the
    // others): x/y/z the six neighbours' indices relative to the structure's first brick as 16-bit halves (0xFFFF:
    // not linked), w the brick's material in bits 0-3 and neighbour d's in bits 4+4d..7+4d (the conductance is
    // min(stiff_i, stiff_j) x 1 vertical / 0.3 lateral, rebuilt from a table)
    uint4* sten_pk;
    uint32_t* sc_c0;           // first contact of each scene this substep
    uint32_t* sc_gmax;         // f32 bits per scene
    uint32_t* sc_fail;         // per scene: joints failed in tension, shear, at the footing (4 slots, [0] unused)
    // hash
    uint32_t *hcnt, *hoff, *hfill, *hent, *dslot;
    unsigned long long* dkey;
    // contacts
    uint32_t *ccnt, *coff;
    int32_t *ca, *cb, *cbo, *cscene, *ccol, *cprop;
    uint32_t *cba, *cbb;
    uint8_t* con;
    float4 *cn, *cp, *cra, *crb, *ct1, *ct2;
    float *cdepth, *cmn, *cmt1, *cmt2, *cjn, *cjt1, *cjt2, *cbias, *cvn0, *csmn, *csmt1, *csmt2;
    // colouring and buckets
    uint32_t* mask;
    unsigned long long* claim;
    uint32_t *colfill, *coloff, *collist;
    // overflow lists
    uint32_t *ovcnt, *ovoff, *ovfill, *ovent;
    uint32_t* ovbody;          // per scene: its bodies with overflow entries (at the scene's first body slot)
    float4 *odv, *odw;
    // solve records of the workgroup-per-scene route, in colour-bucket order (see SRec)
    uint32_t* rq_c;
    int2* rq_ab;
    float4 *rq_0, *rq_1, *rq_2, *rq_3, *rq_4, *rq_5;
    // structural scratch
    uint32_t *lcnt, *loff, *lfill, *lent;
    int32_t* ltarget;
    float* lshare;
    float* bshare;
    uint8_t* bd_sor;           // body load-field sweep this frame
    uint8_t* bd_judge;         // body judged this frame
    uint8_t* bd_live;
    float* bd_dt;
    uint8_t* bd_dirty;
    // events
    uint32_t* ctr;
    int32_t* klist;            // killed bodies
    unsigned long long* evkey; // crush events: order key
    int4* ev;                  // brick, axis, target, kind
    uint32_t* bwi;             // bond writes: brick * 4 + axis
    float* bwv;
    uint32_t evcap, bwcap, pcap;
    // structure connectivity: per brick the token of the last search that reached it, and the search queue (each
    // structure's part at its own brick range)
    uint32_t* creach;
    uint32_t* cq;
    // impacts: per contact the candidate flag (cleared as the crush walk visits it); per brick the stamp of the last
    // crush walk that queued it; per scene the walks' counter (the stamps' source)
    uint8_t* iflag;
    uint8_t* pflag;            // plough candidates, likewise
    uint32_t* xstamp;
    uint32_t* xtok;
    uint32_t* xicnt;           // per scene: impact and plough candidates flagged (zeroed by the walks)
    uint32_t* xpcnt;
    // scan scratch
    uint32_t* scan_tmp;
};

}  // namespace lat
