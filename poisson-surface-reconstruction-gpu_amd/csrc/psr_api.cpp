// C ABI and orchestration (SolvePoissionReconstructionMesh,
// mesh/PoissonReconstruction.cpp:195-245): octree -> field -> divergence ->
// per-depth CG -> implicit function -> geometry -> marching cubes.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "psr_internal.h"


namespace psr {

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return -1.f;
    return ms;
}

Recorder*& current_recorder() {
    static thread_local Recorder* r = nullptr;
    return r;
}

void Arena::free_block(void* p) {
    if (defer_) deferred_.push_back(p);
    else PSR_HIP_CHECK(hipFreeAsync(p, stream_));  // after the work queued on it
}

void Arena::defer_frees(bool on) {
    defer_ = on;
    if (on) return;
    for (void* p : deferred_) PSR_HIP_CHECK(hipFreeAsync(p, stream_));
    deferred_.clear();
}

void* Arena::get(const std::string& name, size_t bytes) {
    Block& b = blocks_[name];
    b.need = std::max(b.need, bytes);
    if (b.bytes < bytes) {
        if (b.ptr) {
            free_block(b.ptr);
            total_ -= b.bytes;
        }
        size_t want = bytes + bytes / 8 + 256;
        PSR_HIP_CHECK(hipMallocAsync(&b.ptr, want, stream_));
        b.bytes = want;
        total_ += want;
        ++allocs_;
    }
    return b.ptr;
}

void* Arena::get_keep(const std::string& name, size_t bytes, size_t keepBytes) {
    Block& b = blocks_[name];
    b.need = std::max(b.need, bytes);
    if (b.bytes < bytes) {
        size_t want = bytes + bytes / 2 + 256;
        void* p = nullptr;
        PSR_HIP_CHECK(hipMallocAsync(&p, want, stream_));
        if (b.ptr) {
            if (keepBytes)
                PSR_HIP_CHECK(hipMemcpyAsync(p, b.ptr, std::min(keepBytes, b.bytes), hipMemcpyDeviceToDevice, stream_));
            free_block(b.ptr);
            total_ -= b.bytes;
        }
        b.ptr = p;
        b.bytes = want;
        total_ += want;
        ++allocs_;
    }
    return b.ptr;
}

void Arena::release() {
    for (void* p : deferred_) (void)hipFreeAsync(p, stream_);
    deferred_.clear();
    for (auto& kv : blocks_)
        if (kv.second.ptr) (void)hipFreeAsync(kv.second.ptr, stream_);
    if (!blocks_.empty()) (void)hipStreamSynchronize(stream_);
    blocks_.clear();
    total_ = 0;
}

}  // namespace psr

namespace {

int fail(psr_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->lastError = msg;
    return code;
}


}  // namespace

namespace psr {

int call_begin(Ctx& c, const psr_points& pts, std::string& err) {
    const int N = int(pts.count);
    if (N < 8) return err = "need at least 8 points", PSR_ERR_INVALID;
    if (c.cfg.max_points && pts.count > c.cfg.max_points)
        return err = "point count exceeds max_points", PSR_ERR_CAPACITY;
    if (!pts.xyz || !pts.normal) return err = "null point buffer", PSR_ERR_INVALID;
    if (pts.stride_bytes && (pts.stride_bytes < 12 || pts.stride_bytes % 4)) return err = "bad stride", PSR_ERR_INVALID;
    std::memset(&c.stats, 0, sizeof(c.stats));
    c.arena.begin_run();
    c.clear_scan_error();
    c.mesh.valid = false;
    c.mesh.weldedNv = c.mesh.weldedNt = 0;
    return PSR_OK;
}

int call_end(Ctx& c, psr_mesh* out, bool timed, std::string& err) {
    finish_mesh(c);
    psr_stats& st = c.stats;
    const Octree& o = c.oct;
    st.depth = o.L.D;
    st.n_points = uint32_t(o.N);
    st.n_nodes = o.M;
    for (int d = 0; d <= o.L.D; ++d) st.nodes_per_depth[d] = o.L.count[d];
    const DevResults& r = *c.hostResults;
    if (r.cgErr) return err = "CG group barrier timed out", PSR_ERR_INTERNAL;
    if (c.scan_failed()) return err = "scan look-back timed out", PSR_ERR_INTERNAL;
    for (int d = 0; d <= o.L.D; ++d) {
        st.cg_iterations[d] = r.cgIters[d];
        st.cg_outer[d] = r.cgIters[PSR_MAX_DEPTH + 1 + d];
        st.cg_nnz[d] = r.cgNnz[d];
        st.cg_updated_residual[d] = r.cgResid[d];
        st.cg_true_relres[d] = c.cfg.precision == PSR_PRECISION_MIXED ? r.cgTrueRel[d] : -1.0;
    }
    st.iso_value = r.iso;
    st.q1_fired = r.q1Fired;
    st.n_vertex_nodes = c.geo.V;
    st.n_edge_nodes = c.geo.E;
    st.n_face_nodes = c.geo.F;
    if (timed) {
        st.ms_octree = elapsed(c.ev[0], c.ev[1]);
        st.ms_field = elapsed(c.ev[1], c.ev[2]);
        st.ms_solve = elapsed(c.ev[2], c.ev[3]);
        st.ms_mesh = elapsed(c.ev[3], c.ev[4]);
        st.ms_total = elapsed(c.ev[0], c.ev[4]);
        st.ms_cg_kernel = elapsed(c.ev[8], c.ev[9]);
        c.cgMsSum += st.ms_cg_kernel;
    }
    st.ms_cg_kernel_sum = c.cgMsSum;
    st.n_runs = ++c.runs;
    st.device_bytes = c.arena.bytes_needed();
    out->vertices = c.mesh.vertices;
    out->triangles = c.mesh.triangles;
    out->n_vertices = uint32_t(c.mesh.nv);
    out->n_triangles = uint32_t(c.mesh.nt);
    c.mesh.valid = true;
    c.cgClean = true;  // the neighbour-column kernel cleared the CG's controls after the solve
    return PSR_OK;
}

}  // namespace psr

namespace {

int run(psr_ctx* ctx, const psr_points* pts, psr_mesh* out) {
    using namespace psr;
    Ctx& c = ctx->c;
    hipStream_t s = c.stream;
    int rc = call_begin(c, *pts, ctx->lastError);
    if (rc != PSR_OK) return rc;
    PSR_HIP_CHECK(hipEventRecord(c.ev[0], s));
    build_octree(c, *pts);
    PSR_HIP_CHECK(hipEventRecord(c.ev[1], s));
    compute_field(c);
    PSR_HIP_CHECK(hipEventRecord(c.ev[2], s));
    solve_depths(c);
    points_implicit(c);
    PSR_HIP_CHECK(hipEventRecord(c.ev[3], s));
    build_geometry(c);
    extract_mesh(c);
    PSR_HIP_CHECK(hipEventRecord(c.ev[4], s));
    c.wait();  // the results reached hostResults with the vertex values (build_geometry)
    return call_end(c, out, true, ctx->lastError);
}

// Host copy of a device int array, transformed column-wise into row-major [M][w].
int export_columns(psr::Ctx& c, const int* dev, int M, int w, void* dst) {
    std::vector<int> tmp(size_t(M) * w);
    PSR_HIP_CHECK(hipMemcpy(tmp.data(), dev, sizeof(int) * tmp.size(), hipMemcpyDeviceToHost));
    int* o = static_cast<int*>(dst);
    for (int i = 0; i < M; ++i)
        for (int k = 0; k < w; ++k) o[size_t(i) * w + k] = tmp[size_t(k) * M + i];
    (void)c;
    return PSR_OK;
}

}  // namespace

extern "C" {

void psr_default_config(psr_config* cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->depth = 7;
    cfg->max_points = 0;
    cfg->precision = PSR_PRECISION_FP32;
    cfg->cg_tol = 1e-5f;
    cfg->cg_max_iter = 10000;
    cfg->refine_rel_tol = 1e-5f;
    cfg->refine_max_outer = 4;
    cfg->subdiv_chunk_nodes = 0;
    cfg->batch_cg_plan = -1;
}

int psr_create(const psr_config* cfg, void* hip_stream, psr_ctx** out) {
    if (!out) return PSR_ERR_INVALID;
    *out = nullptr;
    psr_config def;
    psr_default_config(&def);
    const psr_config& k = cfg ? *cfg : def;
    if (k.depth < 2 || k.depth > 9) return PSR_ERR_INVALID;
    if (k.precision != PSR_PRECISION_FP32 && k.precision != PSR_PRECISION_MIXED) return PSR_ERR_INVALID;
    if (k.precision == PSR_PRECISION_MIXED && !(k.refine_rel_tol > 0.f)) return PSR_ERR_INVALID;
    psr_ctx* ctx = new (std::nothrow) psr_ctx();
    if (!ctx) return PSR_ERR_INTERNAL;
    try {
        psr::Ctx& c = ctx->c;
        c.cfg = k;
        // tuning override of the CG row-block size (rows per workgroup)
        if (const char* e = std::getenv("PSR_CG_ROWS_PER_WG")) c.cgRowsPerWg = std::max(0, std::atoi(e));
        // test override of the CG barriers' spin bound (forces their time-out path)
        if (const char* e = std::getenv("PSR_CG_SPIN_LIMIT")) c.cgSpinLimit = std::max(0, std::atoi(e));
        PSR_HIP_CHECK(hipGetDevice(&c.device));
        if (hip_stream) {
            c.stream = static_cast<hipStream_t>(hip_stream);
        } else {
            PSR_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
            c.ownStream = true;
        }
        c.arena.set_stream(c.stream);
        for (auto& e : c.ev) PSR_HIP_CHECK(hipEventCreate(&e));
        PSR_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c.hostPinned), psr::kMailWords * sizeof(int),
                                    hipHostMallocMapped | hipHostMallocCoherent));
        PSR_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c.mail), c.hostPinned, 0));
        PSR_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c.hostResults), sizeof(psr::DevResults),
                                    hipHostMallocMapped | hipHostMallocCoherent));
        PSR_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c.mailResults), c.hostResults, 0));
        std::memset(c.hostResults, 0, sizeof(psr::DevResults));
        psr::build_tables(c);  // one-time, like ComputeVectorField's constructor (ComputeVectorField.cpp:11-15)
        PSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    } catch (const psr::HipError& e) {
        psr_destroy(ctx);
        return PSR_ERR_HIP;
    } catch (...) {
        psr_destroy(ctx);
        return PSR_ERR_INTERNAL;
    }
    *out = ctx;
    return PSR_OK;
}

int psr_reconstruct(psr_ctx* ctx, const psr_points* points, psr_mesh* out) {
    if (!ctx || !points || !out) return PSR_ERR_INVALID;
    try {
        return run(ctx, points, out);
    } catch (const psr::HipError& e) {
        return fail(ctx, PSR_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(ctx, PSR_ERR_INVALID, e.what());
    } catch (const std::exception& e) {
        return fail(ctx, PSR_ERR_INTERNAL, e.what());
    }
}

int psr_weld_mesh(psr_ctx* ctx, psr_mesh* out) {
    if (!ctx || !out) return PSR_ERR_INVALID;
    psr::Ctx& c = ctx->c;
    if (!c.mesh.valid) return fail(ctx, PSR_ERR_INVALID, "no completed reconstruction to weld");
    try {
        c.clear_scan_error();
        psr::weld_mesh(c);
        if (c.scan_failed()) return fail(ctx, PSR_ERR_INTERNAL, "scan look-back timed out");
    } catch (const psr::HipError& e) {
        return fail(ctx, PSR_ERR_HIP, e.what());
    } catch (const std::exception& e) {
        return fail(ctx, PSR_ERR_INTERNAL, e.what());
    }
    out->vertices = c.mesh.weldedVertices;
    out->triangles = c.mesh.weldedTriangles;
    out->n_vertices = uint32_t(c.mesh.weldedNv);
    out->n_triangles = uint32_t(c.mesh.weldedNt);
    return PSR_OK;
}

int psr_mesh_normals(psr_ctx* ctx, const psr_mesh* mesh, float* normals) {
    if (!ctx || !mesh || (mesh->n_vertices && (!normals || !mesh->vertices)) ||
        (mesh->n_triangles && !mesh->triangles) || mesh->n_vertices > uint32_t(INT32_MAX) ||
        mesh->n_triangles > uint32_t(INT32_MAX / 3))
        return PSR_ERR_INVALID;
    try {
        ctx->c.clear_scan_error();
        psr::mesh_normals(ctx->c, mesh->vertices, int(mesh->n_vertices), mesh->triangles, int(mesh->n_triangles),
                          normals);
        if (ctx->c.scan_failed()) return fail(ctx, PSR_ERR_INTERNAL, "scan look-back timed out");
    } catch (const psr::HipError& e) {
        return fail(ctx, PSR_ERR_HIP, e.what());
    } catch (const std::exception& e) {
        return fail(ctx, PSR_ERR_INTERNAL, e.what());
    }
    return PSR_OK;
}

int psr_mesh_normals_static(const psr_mesh* mesh, float* normals, void* hip_stream) {
    if (!mesh || (mesh->n_vertices && (!normals || !mesh->vertices)) || (mesh->n_triangles && !mesh->triangles) ||
        mesh->n_vertices > uint32_t(INT32_MAX) || mesh->n_triangles > uint32_t(INT32_MAX / 3))
        return PSR_ERR_INVALID;
    std::string err;
    try {
        return psr::mesh_normals_static(mesh->vertices, int(mesh->n_vertices), mesh->triangles,
                                        int(mesh->n_triangles), normals, static_cast<hipStream_t>(hip_stream), err);
    } catch (const psr::HipError&) {
        return PSR_ERR_HIP;
    } catch (const std::exception&) {
        return PSR_ERR_INTERNAL;
    }
}

uint32_t psr_mesh_normals_static_max_triangles(void) { return uint32_t(psr::mesh_normals_static_max_triangles()); }

int psr_estimate_normals(psr_ctx* ctx, const float* xyz, uint32_t count, uint32_t stride_bytes, int k,
                         const float viewpoint[3], float* normals, float* curvature, int32_t* knn) {
    if (!ctx || (count && (!xyz || !normals)) || count > uint32_t(INT32_MAX) / 16 ||
        (stride_bytes && (stride_bytes < 12 || stride_bytes % 4)))
        return PSR_ERR_INVALID;
    const float3 vp = viewpoint ? make_float3(viewpoint[0], viewpoint[1], viewpoint[2]) : make_float3(0.f, 0.f, 0.f);
    try {
        ctx->c.clear_scan_error();
        psr::estimate_normals(ctx->c, xyz, int(count), stride_bytes ? stride_bytes : 12, k, vp, normals, curvature,
                              knn);
        if (ctx->c.scan_failed()) return fail(ctx, PSR_ERR_INTERNAL, "scan look-back timed out");
    } catch (const psr::HipError& e) {
        return fail(ctx, PSR_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(ctx, PSR_ERR_INVALID, e.what());
    } catch (const std::exception& e) {
        return fail(ctx, PSR_ERR_INTERNAL, e.what());
    }
    return PSR_OK;
}

int psr_vertex_colors(psr_ctx* ctx, const psr_mesh* mesh, const float* xyz, const float* point_colors,
                      uint32_t count, uint32_t stride_bytes, float* colors) {
    if (!ctx || !mesh || (mesh->n_vertices && (!mesh->vertices || !colors || !xyz || !point_colors)) ||
        count > uint32_t(INT32_MAX) / 16 || mesh->n_vertices > uint32_t(INT32_MAX) / 4 ||
        (stride_bytes && (stride_bytes < 12 || stride_bytes % 4)))
        return PSR_ERR_INVALID;
    try {
        ctx->c.clear_scan_error();
        psr::vertex_colors(ctx->c, mesh->vertices, int(mesh->n_vertices), xyz, stride_bytes ? stride_bytes : 12,
                           point_colors, int(count), colors);
        if (ctx->c.scan_failed()) return fail(ctx, PSR_ERR_INTERNAL, "scan look-back timed out");
    } catch (const psr::HipError& e) {
        return fail(ctx, PSR_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(ctx, PSR_ERR_INVALID, e.what());
    } catch (const std::exception& e) {
        return fail(ctx, PSR_ERR_INTERNAL, e.what());
    }
    return PSR_OK;
}

int psr_reconstruct_batch(psr_ctx* ctx, const psr_points* clouds, uint32_t count, psr_mesh* outs) {
    if (!ctx || (count && (!clouds || !outs))) return PSR_ERR_INVALID;
    try {
        if (!ctx->batch) ctx->batch = psr::batch_create(ctx);
        std::string err;
        const int rc = psr::batch_run(ctx->batch, clouds, count, outs, ctx->batchStats, err);
        if (rc != PSR_OK) return fail(ctx, rc, err);
    } catch (const psr::HipError& e) {
        return fail(ctx, PSR_ERR_HIP, e.what());
    } catch (const std::exception& e) {
        return fail(ctx, PSR_ERR_INTERNAL, e.what());
    }
    return PSR_OK;
}

psr_ctx* psr_batch_cloud(psr_ctx* ctx, uint32_t index) {
    if (!ctx || index >= ctx->batchStats.size()) return nullptr;
    return psr::batch_cloud(ctx->batch, index);
}

int psr_batch_cg_plan(const psr_ctx* ctx) { return ctx ? psr::batch_last_plan(ctx->batch) : 0; }

int psr_get_batch_stats(psr_ctx* ctx, uint32_t index, psr_stats* out) {
    if (!ctx || !out || index >= ctx->batchStats.size()) return PSR_ERR_INVALID;
    *out = ctx->batchStats[index];
    return PSR_OK;
}

int psr_sync(psr_ctx* ctx) {
    if (!ctx) return PSR_ERR_INVALID;
    return hipStreamSynchronize(ctx->c.stream) == hipSuccess ? PSR_OK : PSR_ERR_HIP;
}

int psr_get_stats(psr_ctx* ctx, psr_stats* out) {
    if (!ctx || !out) return PSR_ERR_INVALID;
    *out = ctx->c.stats;
    return PSR_OK;
}

int psr_export(psr_ctx* ctx, int id, void* dst, uint64_t bytes, uint64_t* needed) {
    if (!ctx) return PSR_ERR_INVALID;
    psr::Ctx& c = ctx->c;
    const psr::Octree& o = c.oct;
    const int M = o.M, N = o.N, D = o.L.D;
    const int nD = M ? o.L.count[D] : 0;
    struct Src {
        const void* p;
        size_t bytes;
        int cols;  // >0: int column-major [cols][M] exported row-major
    } src{nullptr, 0, 0};
    std::vector<int> hostTmp;
    switch (id) {
        case PSR_ARRAY_SORTED_POS: src = {o.spos, size_t(16) * N, 0}; break;
        case PSR_ARRAY_SORTED_NRM: src = {o.snrm, size_t(16) * N, 0}; break;
        case PSR_ARRAY_CODES: src = {o.codes, size_t(4) * N, 0}; break;
        case PSR_ARRAY_ORDER: src = {o.order, size_t(4) * N, 0}; break;
        case PSR_ARRAY_POINT2NODE: src = {o.point2node, size_t(4) * N, 0}; break;
        case PSR_ARRAY_NODE_KEY: src = {o.key, size_t(4) * M, 0}; break;
        case PSR_ARRAY_NODE_PIDX: src = {o.pidx, size_t(4) * M, 0}; break;
        case PSR_ARRAY_NODE_PNUM: src = {o.pnum, size_t(4) * M, 0}; break;
        case PSR_ARRAY_NODE_PARENT: src = {o.parent, size_t(4) * M, 0}; break;
        case PSR_ARRAY_NODE_DIDX: src = {o.didx, size_t(4) * M, 0}; break;
        case PSR_ARRAY_NODE_DNUM: src = {o.dnum, size_t(4) * M, 0}; break;
        case PSR_ARRAY_NODE_ENCIDX: src = {o.encIdx, size_t(4) * M, 0}; break;
        case PSR_ARRAY_NODE_NEIGHS: src = {o.neigh, size_t(4) * 27 * M, 27}; break;
        case PSR_ARRAY_NODE_CHILDREN: {
            if (needed) *needed = size_t(32) * M;
            if (!dst) return PSR_OK;
            if (bytes < size_t(32) * M) return PSR_ERR_CAPACITY;
            std::vector<int> cb(M);
            if (hipMemcpy(cb.data(), o.childBase, 4 * size_t(M), hipMemcpyDeviceToHost) != hipSuccess)
                return PSR_ERR_HIP;
            int* out = static_cast<int*>(dst);
            for (int i = 0; i < M; ++i)
                for (int k = 0; k < 8; ++k) out[8 * size_t(i) + k] = cb[i] < 0 ? -1 : cb[i] + k;
            return PSR_OK;
        }
        case PSR_ARRAY_LEVEL_BASE:
        case PSR_ARRAY_LEVEL_COUNT: {
            size_t n = (id == PSR_ARRAY_LEVEL_BASE) ? D + 2 : D + 1;
            if (needed) *needed = 4 * n;
            if (!dst) return PSR_OK;
            if (bytes < 4 * n) return PSR_ERR_CAPACITY;
            std::memcpy(dst, id == PSR_ARRAY_LEVEL_BASE ? o.L.base : o.L.count, 4 * n);
            return PSR_OK;
        }
        case PSR_ARRAY_NORM_PARAMS: src = {o.normParams, 16, 0}; break;
        case PSR_ARRAY_VFIELD: {
            if (needed) *needed = size_t(12) * nD;
            if (!dst) return PSR_OK;
            if (bytes < size_t(12) * nD) return PSR_ERR_CAPACITY;
            std::vector<float> t(size_t(3) * nD);
            if (c.fld.vhalf) {  // FP16 storage of V * 2^-(D+1)
                std::vector<__half> h(size_t(3) * nD);
                if (hipMemcpy(h.data(), c.fld.vhalf, 6 * size_t(nD), hipMemcpyDeviceToHost) != hipSuccess)
                    return PSR_ERR_HIP;
                for (size_t i = 0; i < h.size(); ++i) t[i] = __half2float(h[i]) * float(2 << D);
            } else if (hipMemcpy(t.data(), c.fld.vfield, 12 * size_t(nD), hipMemcpyDeviceToHost) != hipSuccess) {
                return PSR_ERR_HIP;
            }
            float* out = static_cast<float*>(dst);
            for (int i = 0; i < nD; ++i)
                for (int a = 0; a < 3; ++a) out[3 * size_t(i) + a] = t[size_t(a) * nD + i];
            return PSR_OK;
        }
        case PSR_ARRAY_DIVERGENCE: src = {c.fld.div, size_t(4) * M, 0}; break;
        case PSR_ARRAY_DX: src = {c.fld.dx, size_t(4) * M, 0}; break;
        case PSR_ARRAY_CHI_POINTS: src = {c.fld.chiPts, size_t(4) * N, 0}; break;
        case PSR_ARRAY_VERTS: {
            int V = c.geo.V;
            if (needed) *needed = size_t(24) * V;
            if (!dst) return PSR_OK;
            if (bytes < size_t(24) * V) return PSR_ERR_CAPACITY;
            std::vector<float4> pos(V);
            std::vector<int> own(V), kind(V);
            if (hipMemcpy(pos.data(), c.geo.vtxPos, 16 * size_t(V), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(own.data(), c.geo.vtxOwner, 4 * size_t(V), hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(kind.data(), c.geo.vtxKind, 4 * size_t(V), hipMemcpyDeviceToHost) != hipSuccess)
                return PSR_ERR_HIP;
            int* out = static_cast<int*>(dst);
            for (int v = 0; v < V; ++v) {
                std::memcpy(out + 6 * size_t(v), &pos[v], 12);
                out[6 * size_t(v) + 3] = own[v];
                out[6 * size_t(v) + 4] = kind[v];
                std::memcpy(out + 6 * size_t(v) + 5, &pos[v].w, 4);
            }
            return PSR_OK;
        }
        case PSR_ARRAY_EDGES:
        case PSR_ARRAY_FACES: {
            bool e = id == PSR_ARRAY_EDGES;
            int n = e ? c.geo.E : c.geo.F, w = e ? 2 : 3;
            if (needed) *needed = size_t(4) * w * n;
            if (!dst) return PSR_OK;
            if (bytes < size_t(4) * w * n) return PSR_ERR_CAPACITY;
            std::vector<int> a(n), b(n), h(n);
            if (hipMemcpy(a.data(), e ? c.geo.edgeKind : c.geo.faceKind, 4 * size_t(n), hipMemcpyDeviceToHost) !=
                    hipSuccess ||
                hipMemcpy(b.data(), e ? c.geo.edgeOwner : c.geo.faceOwner, 4 * size_t(n), hipMemcpyDeviceToHost) !=
                    hipSuccess)
                return PSR_ERR_HIP;
            if (!e && hipMemcpy(h.data(), c.geo.faceHasParent, 4 * size_t(n), hipMemcpyDeviceToHost) != hipSuccess)
                return PSR_ERR_HIP;
            int* out = static_cast<int*>(dst);
            for (int i = 0; i < n; ++i) {
                out[w * size_t(i)] = a[i];
                out[w * size_t(i) + 1] = b[i];
                if (!e) out[w * size_t(i) + 2] = h[i];
            }
            return PSR_OK;
        }
        case PSR_ARRAY_NODE_VERTICES: src = {c.geo.nodeVert, size_t(4) * 8 * M, 8}; break;
        case PSR_ARRAY_NODE_FACES: src = {c.geo.nodeFace, size_t(4) * 6 * M, 6}; break;
        case PSR_ARRAY_NODE_EDGES: {
            size_t need = size_t(48) * M;
            if (needed) *needed = need;
            if (!dst) return PSR_OK;
            if (bytes < need) return PSR_ERR_CAPACITY;
            std::vector<int> t(size_t(12) * nD);
            if (hipMemcpy(t.data(), c.geo.nodeEdge, 4 * t.size(), hipMemcpyDeviceToHost) != hipSuccess)
                return PSR_ERR_HIP;
            int* out = static_cast<int*>(dst);
            std::memset(out, 0, need);
            int off = o.L.base[D];
            for (int i = 0; i < nD; ++i)
                for (int k = 0; k < 12; ++k) out[12 * size_t(off + i) + k] = t[size_t(k) * nD + i];
            return PSR_OK;
        }
        case PSR_ARRAY_VVALUE: src = {c.geo.vvalue, size_t(4) * c.geo.V, 0}; break;
        case PSR_ARRAY_MESH_VERTICES: src = {c.mesh.vertices, size_t(12) * c.mesh.nv, 0}; break;
        case PSR_ARRAY_MESH_TRIANGLES: src = {c.mesh.triangles, size_t(12) * c.mesh.nt, 0}; break;
        case PSR_ARRAY_MESH_VERTEX_KEYS: src = {c.mesh.keys, size_t(4) * c.mesh.nv, 0}; break;
        case PSR_ARRAY_WELDED_VERTICES: src = {c.mesh.weldedVertices, size_t(12) * c.mesh.weldedNv, 0}; break;
        case PSR_ARRAY_WELDED_TRIANGLES: src = {c.mesh.weldedTriangles, size_t(12) * c.mesh.weldedNt, 0}; break;
        case PSR_ARRAY_TABLE_DOT: src = {c.tables.dot, size_t(8) * c.tables.res * c.tables.res, 0}; break;
        case PSR_ARRAY_TABLE_DDOT: src = {c.tables.ddot, size_t(8) * c.tables.res * c.tables.res, 0}; break;
        case PSR_ARRAY_TABLE_D2DOT: src = {c.tables.d2dot, size_t(8) * c.tables.res * c.tables.res, 0}; break;
        case PSR_ARRAY_CUBE: src = {c.arena.get("mc_cube", 4), size_t(4) * nD, 0}; break;
        default: return fail(ctx, PSR_ERR_INVALID, "unknown array id");
    }
    if (needed) *needed = src.bytes;
    if (!dst) return PSR_OK;
    if (bytes < src.bytes) return PSR_ERR_CAPACITY;
    if (!src.p || src.bytes == 0) return PSR_OK;
    try {
        PSR_HIP_CHECK(hipStreamSynchronize(c.stream));
        if (src.cols > 0) return export_columns(c, static_cast<const int*>(src.p), M, src.cols, dst);
        PSR_HIP_CHECK(hipMemcpy(dst, src.p, src.bytes, hipMemcpyDeviceToHost));
    } catch (const psr::HipError& e) {
        return fail(ctx, PSR_ERR_HIP, e.what());
    }
    return PSR_OK;
}

void psr_destroy(psr_ctx* ctx) {
    if (!ctx) return;
    if (ctx->batch) psr::batch_destroy(ctx->batch);
    ctx->batch = nullptr;
    psr::Ctx& c = ctx->c;
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    if (c.ev[9]) psr::cg_forget_event(c.ev[9]);
    c.arena.release();
    for (auto& e : c.ev)
        if (e) (void)hipEventDestroy(e);
    if (c.hostPinned) (void)hipHostFree(c.hostPinned);
    if (c.hostResults) (void)hipHostFree(c.hostResults);
    if (c.hostStage) (void)hipHostFree(c.hostStage);
    if (c.ownStream && c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
}

const char* psr_status_string(int status) {
    switch (status) {
        case PSR_OK: return "ok";
        case PSR_ERR_INVALID: return "invalid argument";
        case PSR_ERR_CAPACITY: return "capacity exceeded";
        case PSR_ERR_HIP: return "HIP error";
        case PSR_ERR_INTERNAL: return "internal error";
        default: return "unknown status";
    }
}

const char* psr_last_error(const psr_ctx* ctx) { return ctx ? ctx->lastError.c_str() : ""; }

}  // extern "C"
