// batch_plan.h -- host-side batch engine shared by rs3d_batch_execute and the request
// server: device descriptors of per-query ray work and the plan that runs them in one
// persistent launch.
#pragma once

#include <vector>

#include "axis_grid.h"
#include "rs3d_internal.h"
#include "traverse.h"

namespace rs3d {

enum { BQ_CONTAIN = 0, BQ_EDGES = 1, BQ_OBJECTS = 2 };

struct BatchDesc {
    const BNode* nodes;
    const TriRec* tris;
    const float* pts;
    int32_t* out;
    EdgeArgs E;
    // object rays (intersection fallback): source mesh, first triangle per object,
    // per-object parity scratch of max_targets entries
    const float* src_v;
    const uint32_t* src_i;
    const int* first_tri;
    int* cand_ids;
    uint8_t* cand_par;
    int max_targets;
    AGArgs grid;        // containment on the BVH's axis grid when grid.cells is set
    OccArgs occ;        // edge rays: the target's segment occupancy bits (bits null: none)
    uint64_t n;         // rays
    uint64_t ray_base;  // offset of this query's rays in the reorder (walk queries only)
    uint32_t task_rays; // rays per task (TASK_RAYS, or GRID_TASK_RAYS on an axis grid)
    int kind;
    int axis;
    int swap;
    int qid;
};

// Work distribution: every query is cut into tasks of TASK_RAYS consecutive rays; task
// t belongs to the query q with task_off[q] <= t < task_off[q + 1]. Tasks are dealt from
// NSTRIPE queue heads (task t sits in stripe t % NSTRIPE), each on its own 256-byte line,
// so the atomics of the persistent waves spread over many addresses; a wave starts on the
// stripe of its workgroup index and moves on when a stripe runs dry.
constexpr int TASK_RAYS = 64;
constexpr int GRID_TASK_RAYS = 1024;  // containment on an axis grid: four 256-point sessions per task
constexpr int NSTRIPE = 32;
constexpr int STRIPE_PITCH = 64;  // 32-bit words between heads

// Host-side descriptor plus the containment Morton frame (lo_u, lo_w, scale_u, scale_w).
struct BatchItem : BatchDesc {
    float lo_sc[4] = {0, 0, 0, 0};
    BatchItem() : BatchDesc{} {}
    BatchItem(const BatchDesc& d) : BatchDesc(d) {}
};

class BatchPlan {
public:
    explicit BatchPlan(rs3d_ctx* ctx) : ctx_(ctx) {}
    void reset(rs3d_ctx* ctx) {
        ctx_ = ctx;
        items.clear();
        depth = 24;
        contain_rays = 0;
        any_pairs = false;
        num_qids = 0;
        slot_hint = 1u << 20;
    }
    int add_contain(const rs3d_bvh* b, const float* pts, uint64_t n, int32_t* out, int axis, int qid);
    int add_edges(const rs3d_bvh* target, const rs3d_edges* e, int swap, int qid);
    int add_objects(const rs3d_bvh* target, const rs3d_mesh_view* src, const int* first_tri,
                    int* cand_ids, uint8_t* cand_par, int max_targets, int swap, int qid);
    // keys_out: sorted unique (qid << 48 | a << 24 | b); offsets[q]..offsets[q+1] per query
    int run(DevBuf& keys_out, uint64_t* n_keys, std::vector<uint64_t>& offsets, uint64_t* overflow);
    // Graph form (after a successful run): capture the device pass with the pair-table
    // size that run settled on, then replay it; *grow on replay means the table was too
    // small for this execution's pairs (run again, then capture again).
    int capture(hipGraphExec_t* exec);
    int replay(hipGraphExec_t exec, DevBuf& keys_out, uint64_t* n_keys,
               std::vector<uint64_t>& offsets, uint64_t* overflow, bool* grow);
    void release() {
        meta.release();
        table.release();
        task_q.release();
    }
    int num_qids = 0;
    uint64_t slot_hint = 1u << 20;
    bool split_kinds = true;  // one persistent launch per ray-program kind
    // false: queries use axis grids and occupancy bits that already exist but build none
    // (the request server builds them itself when its memory budget allows)
    bool build_structures = true;

private:
    int base(BatchDesc& d, const rs3d_bvh* b, int kind, int qid);
    int prepare();
    int device_pass(uint64_t slots);
    template <int KINDS, int AX>
    int launch_batch(const BatchDesc* dp, const uint64_t* tp, uint64_t t0, uint64_t t1,
                     unsigned int* heads, const uint32_t* perm, const HashArgs& H,
                     unsigned long long* bad);
    int finish(uint64_t slots, DevBuf& keys_out, uint64_t* n_keys, std::vector<uint64_t>& offsets,
               uint64_t* overflow, bool* grow);
    rs3d_ctx* ctx_;
    std::vector<BatchItem> items;
    int depth = 24;
    uint64_t contain_rays = 0;
    bool any_pairs = false;
    DevBuf meta, table, task_q;
    std::vector<uint8_t> host_meta;  // staging of the uploaded tables (outlives the copy)
    std::vector<uint64_t> task_off;
    uint32_t nq = 0, n_contain = 0;
    uint32_t axis_first[4] = {0, 0, 0, 0};  // containment items of axis a: [axis_first[a], axis_first[a+1])
    uint64_t ntasks = 0, slots_used = 0;
    size_t o_losc = 0, o_seg = 0, o_task = 0, o_head = 0, o_bad = 0, meta_bytes = 0;
    bool segmented = false;   // per-query segments in the reorder sort
    uint64_t sorted_rays = 0;  // keys in the reorder sort (with segment padding)
};

}  // namespace rs3d
