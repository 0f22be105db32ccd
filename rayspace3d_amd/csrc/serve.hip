// serve.hip -- request server with continuous batching over resident BVHs.
//
// Datasets (mesh, BVH, unique edges, first triangle per object, grid histogram for the
// estimator) stay resident in HBM. Requests (containment: points against a dataset;
// overlap and intersection: a query dataset against a resident dataset) queue FIFO. Each
// call of rs3d_server_step is one scheduling iteration: it admits queued requests while
// a slot is free and their estimated device memory fits the budget (pair tables are
// sized from the grid-histogram estimate of the dataset pair; the step computes the
// estimates of the requests submitted since the last one in one launch), runs every admitted
// request in one persistent launch of the batch engine, and completes them. Requests that
// arrive while a step runs are admitted at the next iteration. Each request's result is
// identical to running it alone (the batch engine runs the same per-ray programs and
// keeps one pair set per request).
//
// The memory budget covers what the server keeps resident -- per dataset the first
// triangle of each object, the grid histogram and its non-empty cells, and the axis grids and segment
// occupancy bits of the datasets' BVHs -- plus the per-step memory of the admitted
// requests. Grids and bits only speed requests up (results are identical on the BVH
// walk), so the server builds one only when its upper-bound size fits what the budget
// leaves, before the step's launch; otherwise the request runs on the walk. The BVHs and
// mesh arrays themselves belong to the caller and are not counted.
#include <algorithm>
#include <cmath>
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
    // the histogram's non-empty cells (device, increasing index) and its touch-weighted
    // means (for the estimate's alpha correction): a request's estimate visits only the
    // non-empty cells of the sparser of its two histograms
    DevBuf nz;
    uint64_t nnz = 0;
    float g_size = 0.0f, g_ratio = 1.0f;
};

// One pending estimate of rs3d_server_step's batched estimate kernel.
struct SvEstPair {
    const GridCellRec* a;  // query dataset's histogram
    const GridCellRec* b;  // target dataset's histogram
    const uint32_t* idx;   // non-empty cells of the sparser one
    uint64_t n;
    float cell_volume;
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
    rs3d::DevBuf keys, scratch, est;
    uint64_t peak_mem = 0;
    uint64_t steps = 0;
};

namespace rs3d {
int first_triangles_dev(rs3d_ctx* ctx, const int32_t* tri_obj, uint64_t nT, uint64_t n_obj, int* first);
int batch_keys_to_pairs(rs3d_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out);
}  // namespace rs3d

using namespace rs3d;

__global__ __launch_bounds__(256) void k_sv_nz_flags(const GridCellRec* __restrict__ cells, uint64_t n,
                                                     uint8_t* __restrict__ flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = cells[i].touch > 0 ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_sv_gather(const GridCellRec* __restrict__ cells,
                                                   const uint32_t* __restrict__ idx, uint64_t n,
                                                   GridCellRec* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cells[idx[i]];
}

// The raw estimate (rs3d_estimate_raw's sum) of each pending pair over the listed cells:
// block (x, y) sums its share of pair y's cells into partial[y * gridDim.x + x].
__global__ __launch_bounds__(256) void k_sv_estimate(const SvEstPair* __restrict__ pairs, float eps, float gamma,
                                                     double* __restrict__ partial) {
    __shared__ double red[256];
    const SvEstPair P = pairs[blockIdx.y];
    double s = 0.0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < P.n;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = P.idx[k];
        const GridCellRec a = P.a[i], b = P.b[i];
        if (a.touch == 0 || b.touch == 0) continue;
        const float cs = a.avg + b.avg + eps;
        const float mink = cs * cs * cs;
        float prob = mink / P.cell_volume;
        const float ratio = sqrtf(a.ratio * b.ratio);
        prob *= powf(ratio, gamma);
        prob = fminf(prob, 1.0f);
        s += (double)((float)a.touch * (float)b.touch * prob);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(uint64_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// The dataset's non-empty histogram cells (kept on the device) and touch-weighted means.
static int sparse_histogram(rs3d_ctx* ctx, ServerDataset* d) {
    const uint64_t nc = (uint64_t)d->res * d->res * d->res;
    DevBuf tmp;
    const size_t fl = (nc + 255) & ~(size_t)255;
    RS3D_TRY(tmp.ensure(fl + nc * sizeof(uint32_t) + nc * sizeof(GridCellRec)));
    uint8_t* flags = tmp.as<uint8_t>();
    uint32_t* idx = reinterpret_cast<uint32_t*>(flags + fl);
    GridCellRec* cells = reinterpret_cast<GridCellRec*>(idx + nc);
    const unsigned nb = (unsigned)((nc + 255) / 256);
    k_sv_nz_flags<<<nb, 256, 0, ctx->stream>>>(d->grid.as<GridCellRec>(), nc, flags);
    int rc = rs3d_compact_flags(ctx, flags, nc, idx, &d->nnz);
    if (rc == RS3D_OK) rc = d->nz.ensure(std::max<uint64_t>(d->nnz, 1) * sizeof(uint32_t));
    std::vector<GridCellRec> h(d->nnz);
    if (rc == RS3D_OK && d->nnz) {
        k_sv_gather<<<(unsigned)((d->nnz + 255) / 256), 256, 0, ctx->stream>>>(d->grid.as<GridCellRec>(), idx,
                                                                              d->nnz, cells);
        hipError_t e = hipMemcpyAsync(d->nz.ptr, idx, d->nnz * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(h.data(), cells, d->nnz * sizeof(GridCellRec), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_error("server: histogram copy failed: %s", hipGetErrorString(e));
            rc = RS3D_E_HIP;
        }
    }
    tmp.release();
    if (rc != RS3D_OK) return rc;
    double sz = 0.0, ra = 0.0, cnt = 0.0;
    for (const GridCellRec& c : h) {
        sz += (double)c.avg * (double)c.touch;
        ra += (double)c.ratio * (double)c.touch;
        cnt += (double)c.touch;
    }
    d->g_size = cnt > 0 ? (float)(sz / cnt) : 0.0f;
    d->g_ratio = cnt > 0 ? (float)(ra / cnt) : 1.0f;
    return RS3D_OK;
}

// Device bytes resident for the server's datasets: its own per-dataset arrays and the
// accelerating structures that exist for their BVHs (each BVH counted once).
static uint64_t resident_bytes(const rs3d_server* s) {
    uint64_t b = 0;
    std::set<const rs3d_bvh*> seen;
    for (const ServerDataset* d : s->ds) {
        b += d->first_tri.bytes + d->grid.bytes + d->nz.bytes;
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

// Device memory of an overlap or intersection request, from its estimate: its pair table
// share at <= 1/4 load, compaction and sort; intersection adds its per-object parity state.
static uint64_t request_mem(const ServerRequest& q, const ServerDataset* a, const ServerDataset* b) {
    uint64_t m = pow2_ge(4 * (uint64_t)q.estimate) * 8 * 3;
    if (q.req.kind == RS3D_REQ_INTERSECT) m += (a->mesh.num_objects + b->mesh.num_objects) * 256ull * 5;
    return m;
}

// The pending estimates of the first queued requests (up to 4 x max_slots), in one launch:
// rs3d_estimate_pairs (alpha-corrected; the target's bounds set the cell volume) over the
// non-empty cells of the sparser histogram of each pair.
static int estimate_pending(rs3d_server* s) {
    std::vector<int64_t> pend;
    std::vector<SvEstPair> pairs;
    const size_t look = std::min<size_t>({s->queue.size(), (size_t)s->max_slots * 4, (size_t)65535});
    uint64_t most = 0;
    for (size_t k = 0; k < look; ++k) {
        const ServerRequest& q = s->reqs[s->queue[k]];
        if (q.estimate >= 0) continue;
        const ServerDataset* a = s->ds[q.req.dataset];
        const ServerDataset* b = s->ds[q.req.query_dataset];
        const ServerDataset* sm = a->nnz <= b->nnz ? a : b;
        const float w = a->gmax[0] - a->gmin[0], h = a->gmax[1] - a->gmin[1], d = a->gmax[2] - a->gmin[2];
        const float r = (float)a->res;
        pairs.push_back(SvEstPair{b->grid.as<GridCellRec>(), a->grid.as<GridCellRec>(), sm->nz.as<uint32_t>(),
                                  sm->nnz, (w / r) * (h / r) * (d / r)});
        pend.push_back(s->queue[k]);
        most = std::max(most, sm->nnz);
    }
    if (pend.empty()) return RS3D_OK;
    const unsigned bpp = (unsigned)std::min<uint64_t>(std::max<uint64_t>((most + 8191) / 8192, 1), 64);
    const size_t pb = (pairs.size() * sizeof(SvEstPair) + 255) & ~(size_t)255;
    RS3D_TRY(s->est.ensure(pb + pairs.size() * bpp * sizeof(double)));
    SvEstPair* dpairs = s->est.as<SvEstPair>();
    double* dpart = reinterpret_cast<double*>(s->est.as<uint8_t>() + pb);
    RS3D_HIP(hipMemcpyAsync(dpairs, pairs.data(), pairs.size() * sizeof(SvEstPair), hipMemcpyHostToDevice,
                            s->ctx->stream));
    k_sv_estimate<<<dim3(bpp, (unsigned)pairs.size()), 256, 0, s->ctx->stream>>>(dpairs, 0.001f, 0.8f, dpart);
    RS3D_HIP(hipGetLastError());
    std::vector<double> part(pairs.size() * bpp);
    RS3D_HIP(hipMemcpyAsync(part.data(), dpart, part.size() * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
    RS3D_HIP(hipStreamSynchronize(s->ctx->stream));
    for (size_t k = 0; k < pend.size(); ++k) {
        ServerRequest& q = s->reqs[pend[k]];
        const ServerDataset* a = s->ds[q.req.dataset];
        const ServerDataset* b = s->ds[q.req.query_dataset];
        double raw = 0.0;
        for (unsigned j = 0; j < bpp; ++j) raw += part[k * bpp + j];
        const float rawf = (float)raw;
        const float e1 = b->g_size * std::cbrt(b->g_ratio), e2 = a->g_size * std::cbrt(a->g_ratio);
        const float comb = e1 + e2;
        const float mink = comb * comb * comb;
        float cv = pairs[k].cell_volume;
        if (cv < 1e-9f) cv = 1e-9f;
        float alpha = mink / cv;
        if (alpha < 1.0f) alpha = 1.0f;
        q.estimate = std::max<int64_t>((int64_t)(rawf / alpha), 256);
        q.mem = request_mem(q, a, b);
    }
    return RS3D_OK;
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
        d->nz.release();
        delete d;
    }
    s->keys.release();
    s->est.release();
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
        if (rc == RS3D_OK) rc = sparse_histogram(s->ctx, d);
    }
    if (rc != RS3D_OK) {
        d->first_tri.release();
        d->grid.release();
        d->nz.release();
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
        // pair table share from the grid-histogram estimate (both grids must exist and match);
        // the next step computes the pending estimates of its queue in one launch
        if (a->res && a->res == b->res) {
            q.estimate = -1;
        } else {
            q.estimate = 256;
            q.mem = request_mem(q, a, b);
        }
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
    RS3D_TRY(estimate_pending(s));
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

int rs3d_server_request_estimate(rs3d_server* s, int64_t ticket, int64_t* estimate) {
    if (!s || ticket < 0 || ticket >= (int64_t)s->reqs.size() || !estimate) {
        set_error("rs3d_server_request_estimate: bad arguments");
        return RS3D_E_ARG;
    }
    *estimate = s->reqs[ticket].estimate;
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
