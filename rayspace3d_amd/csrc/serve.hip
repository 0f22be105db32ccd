// serve.hip -- request server with continuous batching over resident BVHs.
//
// Datasets (mesh, BVH, unique edges, first triangle per object, grid histogram for the
// estimator) stay resident in HBM. Requests (containment: points against a dataset;
// overlap and intersection: a query dataset against a resident dataset) queue FIFO. Each
// call of rs3d_server_step is one scheduling iteration: it admits queued requests while
// a slot is free and their estimated device memory fits the budget (pair tables are
// sized from the grid-histogram estimate of the dataset pair), runs every admitted
// request in one persistent launch of the batch engine, and completes them. Requests that
// arrive while a step runs are admitted at the next iteration. Each request's result is
// identical to running it alone (the batch engine runs the same per-ray programs and
// keeps one pair set per request).
//
// The memory budget covers what the server keeps resident -- per dataset the first
// triangle of each object and the grid histogram, and the axis grids and segment
// occupancy bits of the datasets' BVHs -- plus the per-step memory of the admitted
// requests. Grids and bits only speed requests up (results are identical on the BVH
// walk), so the server builds one only when its upper-bound size fits what the budget
// leaves, before the step's launch; otherwise the request runs on the walk. The BVHs and
// mesh arrays themselves belong to the caller and are not counted.
#include <algorithm>
#include <cstring>
#include <deque>
#include <set>
#include <utility>
#include <vector>

#include "batch_plan.h"

namespace rs3d {

struct ServerDataset {
    const rs3d_bvh* bvh = nullptr;
    rs3d_mesh_view mesh{};
    rs3d_edges edges{};
    bool has_edges = false;
    DevBuf first_tri;   // int per object (0x7fffffff: no triangle)
    DevBuf grid;        // GridCell[res^3]
    int res = 0;
    float gmin[3] = {0, 0, 0}, gmax[3] = {0, 0, 0};
};

struct ServerRequest {
    rs3d_request req{};
    int64_t estimate = 0;
    uint64_t mem = 0;
    int state = 0;  // 0 queued, 1 done, 2 failed
    std::vector<uint64_t> keys;
    uint64_t overflow = 0;
};

}  // namespace rs3d

struct rs3d_server {
    rs3d_ctx* ctx = nullptr;
    int max_slots = 64;
    uint64_t budget = 0;
    std::vector<rs3d::ServerDataset*> ds;
    std::vector<rs3d::ServerRequest> reqs;
    std::deque<int64_t> queue;
    rs3d::BatchPlan plan{nullptr};
    rs3d::DevBuf keys, scratch;
    uint64_t peak_mem = 0;
    uint64_t steps = 0;
};

namespace rs3d {
int first_triangles_dev(rs3d_ctx* ctx, const int32_t* tri_obj, uint64_t nT, uint64_t n_obj, int* first);
int batch_keys_to_pairs(rs3d_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out);
}  // namespace rs3d

using namespace rs3d;

// Device bytes resident for the server's datasets: its own per-dataset arrays and the
// accelerating structures that exist for their BVHs (each BVH counted once).
static uint64_t resident_bytes(const rs3d_server* s) {
    uint64_t b = 0;
    std::set<const rs3d_bvh*> seen;
    for (const ServerDataset* d : s->ds) {
        b += d->first_tri.bytes + d->grid.bytes;
        if (!seen.insert(d->bvh).second) continue;
        for (int ax = 0; ax < 3; ++ax) {
            rs3d_axis_grid_info gi;
            if (rs3d_axis_grid_query(d->bvh, ax, &gi) == RS3D_OK && gi.state == 1) b += gi.device_bytes;
        }
        if (d->bvh->socc_state == 1) b += seg_occ_bytes(d->bvh);
    }
    return b;
}

static uint64_t pow2_ge(uint64_t v) {
    uint64_t p = 1024;
    while (p < v) p <<= 1;
    return p;
}

extern "C" {

int rs3d_server_create(rs3d_ctx* ctx, int max_slots, uint64_t memory_budget, rs3d_server** out) {
    if (!ctx || !out || max_slots <= 0 || max_slots >= (1 << 15)) {
        set_error("rs3d_server_create: bad arguments");
        return RS3D_E_ARG;
    }
    rs3d_server* s = new rs3d_server();
    s->ctx = ctx;
    s->max_slots = max_slots;
    s->budget = memory_budget ? memory_budget : (8ull << 30);
    *out = s;
    return RS3D_OK;
}

int rs3d_server_destroy(rs3d_server* s) {
    if (!s) return RS3D_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    for (auto* d : s->ds) {
        d->first_tri.release();
        d->grid.release();
        delete d;
    }
    s->keys.release();
    s->scratch.release();
    s->plan.release();
    delete s;
    return RS3D_OK;
}

int rs3d_server_add_dataset(rs3d_server* s, const rs3d_mesh_view* mesh, const rs3d_bvh* bvh,
                            const rs3d_edges* edges, int grid_resolution, int* id) {
    if (!s || !mesh || !bvh || !id || grid_resolution < 0 || grid_resolution > 512) {
        set_error("rs3d_server_add_dataset: bad arguments");
        return RS3D_E_ARG;
    }
    ServerDataset* d = new ServerDataset();
    d->bvh = bvh;
    d->mesh = *mesh;
    if (edges && edges->num_edges) {
        d->edges = *edges;
        d->has_edges = true;
    }
    const uint64_t nob = mesh->num_objects ? mesh->num_objects : 1;
    int rc = d->first_tri.ensure(nob * sizeof(int));
    if (rc == RS3D_OK && mesh->num_objects)
        rc = first_triangles_dev(s->ctx, mesh->tri_to_obj, mesh->num_triangles, mesh->num_objects,
                                 d->first_tri.as<int>());
    if (rc == RS3D_OK && grid_resolution > 0) {
        d->res = grid_resolution;
        const uint64_t nc = (uint64_t)grid_resolution * grid_resolution * grid_resolution;
        rc = d->grid.ensure(nc * sizeof(GridCellRec));
        float b[6];
        if (rc == RS3D_OK)
            rc = rs3d_grid_stats(s->ctx, mesh->vertices, mesh->num_vertices, mesh->indices,
                                 mesh->num_triangles, mesh->tri_to_obj, mesh->num_objects,
                                 grid_resolution, -1.0f, d->grid.ptr, b);
        if (rc == RS3D_OK)
            for (int k = 0; k < 3; ++k) {
                d->gmin[k] = b[k];
                d->gmax[k] = b[3 + k];
            }
    }
    if (rc != RS3D_OK) {
        d->first_tri.release();
        d->grid.release();
        delete d;
        return rc;
    }
    s->ds.push_back(d);
    *id = (int)s->ds.size() - 1;
    return RS3D_OK;
}

int rs3d_server_submit(rs3d_server* s, const rs3d_request* r, int64_t* ticket) {
    if (!s || !r || !ticket || r->dataset < 0 || r->dataset >= (int)s->ds.size()) {
        set_error("rs3d_server_submit: bad arguments");
        return RS3D_E_ARG;
    }
    ServerRequest q;
    q.req = *r;
    const ServerDataset* a = s->ds[r->dataset];
    if (r->kind == RS3D_REQ_CONTAIN) {
        if ((r->num_points && (!r->points || !r->out_obj)) || r->axis < 0 || r->axis > 2) {
            set_error("rs3d_server_submit: bad containment request");
            return RS3D_E_ARG;
        }
        q.mem = r->num_points * 8;  // reorder keys and permutation
    } else if (r->kind == RS3D_REQ_OVERLAP || r->kind == RS3D_REQ_INTERSECT) {
        if (r->query_dataset < 0 || r->query_dataset >= (int)s->ds.size()) {
            set_error("rs3d_server_submit: bad query dataset");
            return RS3D_E_ARG;
        }
        const ServerDataset* b = s->ds[r->query_dataset];
        if (!a->has_edges || !b->has_edges) {
            set_error("rs3d_server_submit: overlap and intersection requests need edge payloads");
            return RS3D_E_ARG;
        }
        // pair table share from the grid-histogram estimate (both grids must exist and match)
        int64_t est = 0;
        if (a->res && a->res == b->res) {
            const uint32_t res[3] = {(uint32_t)a->res, (uint32_t)a->res, (uint32_t)a->res};
            const uint64_t nc = (uint64_t)a->res * a->res * a->res;
            RS3D_TRY(rs3d_estimate_pairs(s->ctx, b->grid.ptr, a->grid.ptr, nc, a->gmin, a->gmax, res,
                                         0.001f, 0.8f, 1, &est));
        }
        q.estimate = std::max<int64_t>(est, 256);
        q.mem = pow2_ge(4 * (uint64_t)q.estimate) * 8 * 3;  // table at <= 1/4 load + compaction + sort
        if (r->kind == RS3D_REQ_INTERSECT)
            q.mem += (a->mesh.num_objects + b->mesh.num_objects) * 256ull * 5;
    } else {
        set_error("rs3d_server_submit: unknown request kind %d", r->kind);
        return RS3D_E_ARG;
    }
    s->reqs.push_back(std::move(q));
    *ticket = (int64_t)s->reqs.size() - 1;
    s->queue.push_back(*ticket);
    return RS3D_OK;
}

int rs3d_server_step(rs3d_server* s, int* admitted, int* completed) {
    if (!s) {
        set_error("rs3d_server_step: null server");
        return RS3D_E_ARG;
    }
    if (admitted) *admitted = 0;
    if (completed) *completed = 0;
    if (s->queue.empty()) return RS3D_OK;
    // admission: FIFO, slots and the memory budget (the head request is always admitted).
    // A request whose dataset lacks a grid (containment) or occupancy bits (overlap,
    // intersection) plans their build when the bound of their size fits the budget next to
    // what is resident and admitted; otherwise it runs without them.
    std::vector<int64_t> batch;
    uint64_t mem = 0, est_total = 0, scratch_bytes = 0;
    const uint64_t resident = resident_bytes(s);
    std::vector<std::pair<const rs3d_bvh*, int>> builds;  // (bvh, axis); axis -1: occupancy bits
    uint64_t build_bytes = 0;
    auto planned = [&](const rs3d_bvh* b, int ax) {
        return std::find(builds.begin(), builds.end(), std::make_pair(b, ax)) != builds.end();
    };
    while (!s->queue.empty() && (int)batch.size() < s->max_slots) {
        const ServerRequest& r = s->reqs[s->queue.front()];
        if (!batch.empty() && resident + build_bytes + mem + r.mem > s->budget) break;
        std::vector<std::pair<const rs3d_bvh*, int>> want;
        uint64_t want_bytes = 0;
        if (r.req.kind == RS3D_REQ_CONTAIN) {
            const rs3d_bvh* b = s->ds[r.req.dataset]->bvh;
            if (s->ctx->opt_axis_grid && b->num_tris && b->agrid_state[r.req.axis] == 0 && !planned(b, r.req.axis)) {
                want.emplace_back(b, r.req.axis);
                want_bytes += axis_grid_bytes_bound(b);
            }
        } else if (s->ctx->opt_seg_occupancy) {
            for (const rs3d_bvh* b : {s->ds[r.req.dataset]->bvh, s->ds[r.req.query_dataset]->bvh})
                if (b->num_tris && b->socc_state == 0 && !planned(b, -1) &&
                    std::find(want.begin(), want.end(), std::make_pair(b, -1)) == want.end()) {
                    want.emplace_back(b, -1);
                    want_bytes += seg_occ_bytes_bound(b);
                }
        }
        if (resident + build_bytes + want_bytes + mem + r.mem <= s->budget) {
            builds.insert(builds.end(), want.begin(), want.end());
            build_bytes += want_bytes;
        }
        mem += r.mem;
        est_total += (uint64_t)r.estimate;
        if (r.req.kind == RS3D_REQ_INTERSECT)
            scratch_bytes += (s->ds[r.req.dataset]->mesh.num_objects + s->ds[r.req.query_dataset]->mesh.num_objects) * 256ull * 5 + 512;
        batch.push_back(s->queue.front());
        s->queue.pop_front();
    }
    for (const auto& bw : builds) {
        rs3d_bvh* b = const_cast<rs3d_bvh*>(bw.first);
        if (bw.second >= 0)
            RS3D_TRY(rs3d_axis_grid_prepare(s->ctx, b, bw.second, nullptr));
        else
            RS3D_TRY(rs3d_seg_occupancy_prepare(s->ctx, b, nullptr));
    }
    s->peak_mem = std::max(s->peak_mem, (builds.empty() ? resident : resident_bytes(s)) + mem);
    RS3D_TRY(s->scratch.ensure(scratch_bytes + 256));
    uint8_t* scr = s->scratch.as<uint8_t>();
    size_t so = 0;
    BatchPlan& P = s->plan;
    P.reset(s->ctx);
    P.build_structures = false;  // only what exists now (built above when the budget allowed)
    for (size_t k = 0; k < batch.size(); ++k) {
        const rs3d_request& r = s->reqs[batch[k]].req;
        const int qid = (int)k;
        ServerDataset* a = s->ds[r.dataset];
        if (r.kind == RS3D_REQ_CONTAIN) {
            RS3D_TRY(P.add_contain(a->bvh, r.points, r.num_points, r.out_obj, r.axis, qid));
            continue;
        }
        ServerDataset* b = s->ds[r.query_dataset];
        // pairs are (query object, resident object)
        RS3D_TRY(P.add_edges(a->bvh, &b->edges, 0, qid));
        RS3D_TRY(P.add_edges(b->bvh, &a->edges, 1, qid));
        if (r.kind == RS3D_REQ_INTERSECT) {
            const uint64_t nb = b->mesh.num_objects, na = a->mesh.num_objects;
            int* ids_b = reinterpret_cast<int*>(scr + so);
            uint8_t* par_b = scr + so + nb * 256 * 4;
            so += ((nb * 256 * 5) + 255) & ~255ull;
            int* ids_a = reinterpret_cast<int*>(scr + so);
            uint8_t* par_a = scr + so + na * 256 * 4;
            so += ((na * 256 * 5) + 255) & ~255ull;
            RS3D_TRY(P.add_objects(a->bvh, &b->mesh, b->first_tri.as<int>(), ids_b, par_b, 256, 0, qid));
            RS3D_TRY(P.add_objects(b->bvh, &a->mesh, a->first_tri.as<int>(), ids_a, par_a, 256, 1, qid));
        }
    }
    P.num_qids = (int)batch.size();
    P.slot_hint = pow2_ge(4 * std::max<uint64_t>(est_total, 1024));
    uint64_t n = 0, over = 0;
    std::vector<uint64_t> offs;
    int rc = P.run(s->keys, &n, offs, &over);
    if (rc != RS3D_OK) {
        for (int64_t t : batch) s->reqs[t].state = 2;
        return rc;
    }
    std::vector<uint64_t> hk(n);
    if (n) {
        RS3D_TRY(batch_keys_to_pairs(s->ctx, s->keys.as<uint64_t>(), n, s->keys.as<uint64_t>()));
        RS3D_HIP(hipMemcpyAsync(hk.data(), s->keys.ptr, n * sizeof(uint64_t), hipMemcpyDeviceToHost, s->ctx->stream));
    }
    RS3D_HIP(hipStreamSynchronize(s->ctx->stream));
    for (size_t k = 0; k < batch.size(); ++k) {
        ServerRequest& r = s->reqs[batch[k]];
        const uint64_t lo = k < offs.size() ? offs[k] : n, hi = k + 1 < offs.size() ? offs[k + 1] : n;
        r.keys.assign(hk.begin() + lo, hk.begin() + hi);
        r.overflow = over;
        r.state = 1;
    }
    ++s->steps;
    if (admitted) *admitted = (int)batch.size();
    if (completed) *completed = (int)batch.size();
    return RS3D_OK;
}

int rs3d_server_result(rs3d_server* s, int64_t ticket, int* done, uint64_t* keys, uint64_t max_keys,
                       uint64_t* n_keys) {
    if (!s || ticket < 0 || ticket >= (int64_t)s->reqs.size() || !done) {
        set_error("rs3d_server_result: bad arguments");
        return RS3D_E_ARG;
    }
    const ServerRequest& r = s->reqs[ticket];
    *done = r.state;
    if (n_keys) *n_keys = r.keys.size();
    if (r.state == 1 && keys) {
        if (r.keys.size() > max_keys) {
            set_error("rs3d_server_result: %zu pairs exceed the capacity %llu", r.keys.size(),
                      (unsigned long long)max_keys);
            return RS3D_E_CAPACITY;
        }
        if (!r.keys.empty()) memcpy(keys, r.keys.data(), r.keys.size() * sizeof(uint64_t));
    }
    return RS3D_OK;
}

int rs3d_server_stats(rs3d_server* s, uint64_t* queued, uint64_t* peak_memory, uint64_t* steps) {
    if (!s) {
        set_error("rs3d_server_stats: null server");
        return RS3D_E_ARG;
    }
    if (queued) *queued = s->queue.size();
    if (peak_memory) *peak_memory = s->peak_mem;
    if (steps) *steps = s->steps;
    return RS3D_OK;
}

}  // extern "C"
