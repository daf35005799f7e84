// Shared device/host definitions of the orgpu library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/orgpu.h"

// ---------------------------------------------------------------------------
// Packed search key (one u64 per (slot, node)), minimised with atomicMin:
//
//   bits 63..56  keepout edges on the path (counting up to KEY_KO_MAX)
//   bits 55..25  FP32 bit pattern of the path cost without keepout edges
//                (non-negative floats order like their bit patterns)
//   bits 24..0   predecessor field: predecessor node index + 1 (lowest
//                index wins ties); 0 for a source, so that a source keeps
//                "no predecessor" against an equal-cost path into it
//
// The reference replaces the cost of a keepout edge by 1e9
// (UPF:3182, 3220) and sums in float64; with path costs far below 1e9 that
// orders paths by (number of keepout edges, remaining cost), which is what the
// two top fields encode without the FP32 absorption a 1e9 summand would cause.
// A path whose count reaches KEY_KO_MAX is reported instead of routed (the
// field could no longer order it).
// ---------------------------------------------------------------------------
#define KEY_INF 0xFFFFFFFFFFFFFFFFull
#define KEY_PRED_BITS 25
#define KEY_PRED_MASK 0x01FFFFFFu
#define KEY_NOPRED 0u
#define KEY_MAX_NODES ((1ll << 25) - 2)
#define KEY_KO_MAX 255u

__host__ __device__ inline uint64_t key_pack(uint32_t ko, float d, uint32_t pred)
{
    uint32_t db;
    __builtin_memcpy(&db, &d, 4);
    return ((uint64_t)ko << 56) | ((uint64_t)(db & 0x7FFFFFFFu) << 25) | (uint64_t)(pred & KEY_PRED_MASK);
}
__host__ __device__ inline uint32_t key_ko(uint64_t k) { return (uint32_t)(k >> 56); }
__host__ __device__ inline float key_dist(uint64_t k)
{
    uint32_t db = (uint32_t)((k >> 25) & 0x7FFFFFFFu);
    float d;
    __builtin_memcpy(&d, &db, 4);
    return d;
}
// predecessor field (KEY_NOPRED or node + 1)
__host__ __device__ inline uint32_t key_pred(uint64_t k) { return (uint32_t)(k & KEY_PRED_MASK); }
// (keepout count, distance) part, comparable as an integer
__host__ __device__ inline uint64_t key_cost(uint64_t k) { return k >> 25; }
// one more keepout edge, counting up to KEY_KO_MAX
__host__ __device__ inline uint32_t ko_add(uint32_t ko) { return ko < KEY_KO_MAX ? ko + 1u : KEY_KO_MAX; }

// bucket of a key relative to the slot's base bucket; h: lower bound of the
// remaining cost to the destination (0 for plain Dijkstra order)
__host__ __device__ inline uint64_t key_bucket_rel(uint64_t k, float inv_delta, float h = 0.0f)
{
    float q = (key_dist(k) + h) * inv_delta;
    uint32_t b = q >= 4.0e9f ? 4000000000u : (uint32_t)q;
    return ((uint64_t)key_ko(k) << 32) | b;
}

// ---------------------------------------------------------------------------
// Search engine layout
// ---------------------------------------------------------------------------
#define MAX_RING 32
// physical bucket bitmaps: ring [0, nb), OUT, OVF (and one spare)
#define MAX_BUFS (MAX_RING + 3)

enum { MODE_IDLE = 0, MODE_RELAX = 1, MODE_REDIST = 2 };

struct Ctrl {
    uint64_t cur;              // current global bucket
    uint64_t ovf_min;          // lower bound of buckets held in OVF
    uint32_t mode;
    uint32_t finished;         // last-workgroup counter
    uint32_t ring[MAX_RING];   // ring position -> physical buffer
    uint32_t out_buf, ovf_buf, ovf_next_buf;
    uint32_t spare_list;       // chunk-list storage not bound to a bitmap
    uint32_t count[MAX_BUFS];  // items per physical buffer
    uint32_t active;           // slots seeded and not finished
    uint32_t epoch;            // bumps when any slot finishes
    uint64_t rounds, items, edges;
    uint32_t list_id[MAX_BUFS];     // bitmap -> chunk-list storage
};

// A bucket's chunk list is kept in LIST_BINS bins by item region (bin of
// chunk ch = ch * LIST_BINS / chunks); the bins in order form the list a
// round walks, so neighbouring chunks are visited by the same XCD without
// sorting anything.
#define LIST_BINS 64

struct SlotInfo {
    uint64_t base;      // bucket of distance 0 for this slot's search
    int32_t src, dst;
    int32_t net_id;
    int32_t roi_index;  // -1: whole graph
    int32_t use_keepout;
    uint32_t state;     // 0 free/draining, 1 searching, 2 dst reached, 3 dst unreachable
    int32_t req;        // request index
    int32_t pending;    // items of this slot still queued (pushes - pops)
    int32_t dx, dy, dz; // destination coordinates (goal-directed bucketing)
    int32_t ms;         // multi-source / multi-sink search (targets below)
    int32_t t_off, t_n; // its targets in RoundArgs::ms_tgt
    int32_t pad0, pad1;
};

struct RoundArgs {
    Ctrl *ctrl;
    SlotInfo *slots;
    int32_t K;
    int64_t N;
    int64_t NP;               // per-slot item stride: N rounded up to whole chunks
    uint32_t cps;             // chunks per slot (NP / 2048)
    uint32_t nb;
    float inv_delta;
    float h_xy, h_z;          // goal-directed lower bounds (0, 0: Dijkstra order)
    int32_t local_batches;    // same-bucket batches a wave expands per round
    int32_t X, plane;         // node = z*plane + y*X + x
    const int32_t *indptr;
    const int32_t *indices;
    const float *cost;
    const uint8_t *kind;
    const uint32_t *roi;      // [n_roi][words]
    int64_t roi_words;
    const int32_t *ko_owner;  // [N] or null
    uint64_t *keys;           // [K][NP]
    uint32_t *bits;           // [MAX_BUFS][bitmap_words]
    uint64_t bitmap_words;
    uint32_t *summ;           // [MAX_BUFS][summ_words]: one bit per 64-word chunk
    uint64_t summ_words;
    const orgpu_ms_point *ms_tgt;  // targets of multi-source requests
    uint32_t *tbits;          // [K][tbits_words]: target nodes of each multi-source slot
    int64_t tbits_words;
    uint32_t *touch_bits;     // [K][touch_words]: key chunks (2048 keys) a search set
    uint32_t *touch_list;     // [K][touch_cap]: the same chunks as a list
    uint32_t *touch_n;        // [K]
    int64_t touch_words, touch_cap;
    uint32_t *lists;          // [MAX_BUFS + 1][LIST_BINS][bin_cap] chunk lists
    uint32_t *list_cnt;       // [MAX_BUFS + 1][LIST_BINS] entries per bin
    uint64_t chunks;          // chunks per bitmap (64 words each)
    uint64_t bin_cap;         // chunks per bin
    struct RoundPart *parts;  // [grid] per-workgroup partials of a round
};

// What one workgroup of a round leaves for the control kernel.
struct RoundPart {
    int32_t buf_delta[MAX_BUFS];   // set-bit change per bucket bitmap
    uint32_t pad;
    unsigned long long items, edges;
    unsigned long long ovf_min;    // lowest overflow bucket it saw
};

#define HIP_CHECK_RET(expr)                                                   \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) return orgpu_fail_hip(_e, #expr, __FILE__, __LINE__); \
    } while (0)

int orgpu_fail(int code, const char *fmt, ...);
int orgpu_fail_hip(hipError_t e, const char *what, const char *file, int line);
