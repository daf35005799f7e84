// Host-side state of one orgpu handle: stream and the search engine scratch.
#pragma once

#include "common.h"

struct SeedReq;

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    // capacities reserved
    int64_t cap_nodes = 0;
    int32_t cap_slots = 0, cap_nb = 0, cap_path = 0;
    int32_t cap_bslots = 0;   // slots the bucket-round engine's structures hold (<= cap_slots)
    uint64_t words = 0;
    // device scratch
    uint64_t *keys = nullptr;
    uint32_t *bits = nullptr;
    uint32_t *summ = nullptr;
    uint32_t *lists = nullptr, *list_cnt = nullptr;
    uint32_t *touch_bits = nullptr, *touch_list = nullptr, *touch_n = nullptr;
    int64_t keys_n = -1;   // node count the keys' slot layout was last reset for
    int64_t keys_cap = 0;  // keys allocated
    // multi-source / multi-sink calls
    orgpu_ms_point *ms_seed = nullptr, *ms_tgt = nullptr;
    int64_t ms_seed_cap = 0, ms_tgt_cap = 0;
    uint32_t *tbits = nullptr;
    int64_t tbits_cap = 0;
    double *ms_total = nullptr, *h_ms_total = nullptr;
    int32_t *ms_choice = nullptr, *h_ms_choice = nullptr;
    RoundPart *parts = nullptr;
    uint64_t summ_words = 0, chunks = 0;
    Ctrl *ctrl = nullptr;
    SlotInfo *slots = nullptr;
    SeedReq *seed = nullptr;
    int32_t *slot_list = nullptr;
    int32_t *stage_nodes = nullptr, *stage_edges = nullptr;
    // finished paths packed back to back in src -> dst order: nodes in
    // [0, cap_slots*cap_path), edges in the second half; one copy per poll
    int32_t *pack = nullptr;
    uint32_t *pack_n = nullptr, *pack_off = nullptr;
    uint32_t *h_pack_off = nullptr, *h_pack_n = nullptr;
    int32_t *lens = nullptr;
    float *dists = nullptr;
    unsigned long long *counters = nullptr;
    // pinned host mirrors
    Ctrl *h_ctrl = nullptr;
    SlotInfo *h_slots = nullptr;
    SeedReq *h_seed = nullptr;
    int32_t *h_slot_list = nullptr, *h_lens = nullptr, *h_stage = nullptr;
    float *h_dists = nullptr;
    unsigned long long *h_counters = nullptr;
    hipEvent_t ev_poll = nullptr, ev_t0 = nullptr, ev_t1 = nullptr, ev_r0 = nullptr, ev_r1 = nullptr;
    // a poll's rounds captured as one HIP graph (launch_mode 1): the
    // capture stream only records; the executable is launched on `stream`
    hipStream_t capture_stream = nullptr;
    hipGraphExec_t round_exec = nullptr;
    unsigned char round_exec_args[sizeof(RoundArgs)] = {};
    int32_t round_exec_rpp = 0, round_exec_grid = 0;
    // congestion-state helpers: nonzero staging and host->device uploads
    int32_t *sp_idx = nullptr;
    float *sp_val = nullptr;
    unsigned long long *sp_count = nullptr, *h_sp_count = nullptr;
    int64_t sp_cap = 0;
    int32_t *up_buf = nullptr;
    int64_t up_cap = 0;
    // ROI BFS scratch: per workgroup slot a dst-wave bitmap and two node lists
    uint32_t *roi_b = nullptr;
    int32_t *roi_la = nullptr, *roi_lb = nullptr;
    int64_t roi_n = 0;
    int32_t roi_slots = 0;
    int64_t roi_limit = 32ll << 30;   // bytes the ROI BFS scratch may take (orgpu_set_roi_scratch)
    // ROI call staging, grown on demand and kept (no allocation per call)
    orgpu_roi_req *roi_req = nullptr;
    orgpu_roi_info *roi_info = nullptr;
    uint32_t *roi_next = nullptr;
    int64_t roi_req_cap = 0;
    uint8_t *roi_T = nullptr;   // closed-form ROI: via-hop table of the lattice
    // reference-order search (sssp_ref.hip): per slot 16 B per node and a heap
    void *ref_state = nullptr, *ref_heap = nullptr, *ref_ctl = nullptr;
    int64_t ref_nodes = 0, ref_hcap = 0;
    int32_t ref_slots = 0;
    uint32_t ref_sid = 0;       // search ids below this one are used
    // bucket-round searches leave touch records for the next seeding of a
    // slot; the solo engine resets them first (engine_prepare_keys)
    bool touch_dirty = false;
    // one-wave-per-search engine (sssp_solo.hip): per-slot bucket-list arena,
    // request and result staging, control words
    uint64_t *solo_arena = nullptr;
    uint32_t *solo_links = nullptr;
    uint32_t solo_nch = 0;
    int32_t solo_slots = 0;
    orgpu_net_req *solo_req = nullptr;
    int32_t *solo_out = nullptr, *h_solo_out = nullptr;
    int32_t solo_req_cap = 0;
    uint32_t *solo_ctl = nullptr, *h_solo_ctl = nullptr;
    uint4 *solo_recs = nullptr;       // adjacency records, two uint4 per key index
    int64_t solo_recs_cap = 0;        // key indices they cover
};

struct orgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Engine en;
};

int engine_reserve(Engine &en, int64_t n_nodes, int32_t slots, int32_t n_buckets, int64_t qcap, int32_t max_path);
void engine_release(Engine &en);
void ref_release(Engine &en);
void solo_release(Engine &en);
// Makes every key of the reserved slots KEY_INF for a graph of n_nodes nodes
// (whole fill when the node count changed, else the chunks pending in the
// bucket engine's touch records) and clears those records.
int engine_keys_alloc(Engine &en, uint64_t need);
int engine_prepare_keys(Engine &en, int64_t n_nodes);
// the multi-source part of a route call (null for two-pin requests)
struct MsCall {
    const orgpu_ms_req *reqs;
    const orgpu_ms_point *seeds, *targets;
    int64_t n_seeds, n_targets;
    double *path_total;
    int32_t *path_target;
};

int engine_route(Engine &en, const orgpu_graph *g, const orgpu_net_req *nets, int32_t n_nets,
                 const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner, float keepout_cost,
                 const orgpu_sssp_params *p, int32_t *path_nodes, int32_t *path_edges,
                 int64_t path_capacity, int64_t *path_offset, float *path_dist, orgpu_sssp_stats *stats,
                 const MsCall *ms = nullptr);
int engine_reserve_solo(Engine &en, int64_t n_nodes, int32_t X, int32_t Y, int32_t L, int32_t slots,
                        int32_t list_cap);
int engine_route_async(Engine &en, const orgpu_graph *g, const orgpu_net_req *d_reqs, int32_t n,
                       const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner,
                       const orgpu_sssp_params *p, int32_t *path_nodes, int32_t *path_edges, int64_t capacity,
                       uint32_t *path_off, int32_t *path_len, float *path_dist, uint32_t *used);
int engine_route_solo(Engine &en, const orgpu_graph *g, const orgpu_net_req *nets, int32_t n_nets,
                      const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner, float keepout_cost,
                      const orgpu_sssp_params *p, int32_t *path_nodes, int32_t *path_edges, int64_t path_capacity,
                      int64_t *path_offset, float *path_dist, orgpu_sssp_stats *stats);
