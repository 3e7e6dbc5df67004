// Reconstruction context shared by the stage drivers.
#pragma once
#include <cstddef>
#include <map>
#include <string>
#include <vector>

#include "../../include/psr.h"
#include "psr_common.h"
#include "psr_prims.h"

#include <algorithm>

#include <hip/hip_fp16.h>

namespace psr {

// Grow-only device arena: one allocation per named buffer, reused across
// calls; a buffer is reallocated only when a call needs more than it holds.
// Named device buffers of one context, grown on demand with stream-ordered
// allocation on the context's stream (no device-wide synchronisation, so
// contexts on other streams -- batch workers -- keep running).
class Arena {
  public:
    ~Arena() { release(); }
    void set_stream(hipStream_t s) { stream_ = s; }
    void* get(const std::string& name, size_t bytes);
    // As get, but a reallocation carries over the first keepBytes of the old contents.
    void* get_keep(const std::string& name, size_t bytes, size_t keepBytes);
    void release();
    size_t bytes_reserved() const { return total_; }
    // Starts a run's count of the bytes its buffers need (the largest request
    // of each buffer name); bytes_needed() is that sum.
    void begin_run() {
        for (auto& kv : blocks_) kv.second.need = 0;
    }
    size_t bytes_needed() const {
        size_t s = 0;
        for (const auto& kv : blocks_) s += kv.second.need;
        return s;
    }
    int allocations() const { return allocs_; }
    // While on, a buffer replaced by a larger one is freed only when deferral
    // ends (a batch records launches that still name it); turning it off
    // queues those frees on the stream.
    void defer_frees(bool on);

  private:
    void free_block(void* p);
    struct Block {
        void* ptr = nullptr;
        size_t bytes = 0;
        size_t need = 0;
    };
    std::map<std::string, Block> blocks_;
    bool defer_ = false;
    std::vector<void*> deferred_;
    hipStream_t stream_ = nullptr;
    size_t total_ = 0;
    int allocs_ = 0;
};

// 1-D inner-product tables for the configured depth, [a + res * b].
struct Tables {
    int res = 0;
    double* dot = nullptr;
    double* ddot = nullptr;
    double* d2dot = nullptr;
};

// Point and node arrays of the octree (structure of arrays).
struct Octree {
    int N = 0;
    Levels L{};
    int M = 0;
    int U[PSR_MAX_DEPTH + 1] = {0};  // occupied nodes per level
    float4* spos = nullptr;          // sorted normalised positions
    float4* snrm = nullptr;          // sorted scaled normals
    int* codes = nullptr;            // sorted Morton codes
    int* order = nullptr;            // sorted -> input index
    int* point2node = nullptr;       // sorted point -> level-local depth-D node
    float4* basis = nullptr;         // per node: 1-D basis centres (x, y, z) and inverse width
    int* q1Fired = nullptr;          // device flag: the level-local child 0 parent rule changed a parent
    int* key = nullptr;
    int* pidx = nullptr;
    int* pnum = nullptr;
    int* parent = nullptr;
    int* childBase = nullptr;        // children[k] = childBase + k, or -1
    int* didx = nullptr;
    int* dnum = nullptr;
    int* neigh = nullptr;            // [27][M]
    int* encIdx = nullptr;
    int* depth = nullptr;            // depth of each node
    float* normParams = nullptr;     // corner xyz, edge
};

struct Field {
    float* vfield = nullptr;  // [3][nD] (FP32 form)
    __half* vhalf = nullptr;  // [3][nD], V * 2^-(D+1) (mixed-precision form)
    float* div = nullptr;     // [M]
    float* dx = nullptr;      // [M]
    float* chiPts = nullptr;  // [N]
    float4* nbrCol = nullptr;  // [27][M]: centre (x, y, z) and d_x of neighbour i of each node (0 where absent);
                               // written only when the level-local child 0 rule changed a parent
    float* dCol = nullptr;     // [27][M]: d_x of neighbour i of each node (0 where absent), 4 B a term
    double* isoSum = nullptr; // device scalar
    float iso = 0.f;
    PointsChiArgs chiArgs{};  // the points' implicit values, launched with the mesh stage's first kernel
    int chiBlocks = 0;
};

struct Geometry {
    int* vtxOwner = nullptr;  // compacted vertices: owner node
    int* vtxKind = nullptr;
    float4* vtxPos = nullptr;
    int* edgeOwner = nullptr;
    int* edgeKind = nullptr;
    int* faceOwner = nullptr;
    int* faceKind = nullptr;
    int* faceHasParent = nullptr;
    int* nodeVert = nullptr;  // [8][M], 1-based
    int* nodeEdge = nullptr;  // [12][nD] (depth-D nodes, level-local), 1-based
    int* nodeFace = nullptr;  // [6][M], 1-based
    int* faceMarks = nullptr; // per face: an emitted triangle's edge lies on it (6M, the face bound)
    int2* edgeOf = nullptr;   // per depth-D node: first owned edge, owned-edge mask
    int* counts = nullptr;    // device: V, E, F
    int V = 0, E = 0, F = 0;  // host copies, known once the mesh stage has waited
    float* vvalue = nullptr;
    float* dxGrid = nullptr;  // dense per-depth d_x grids, filled when the octree is regular (or null)
};

// Dense per-depth d_x grids are built before the mesh stage's wait up to this
// depth (9.6 MB at depth 7, 77 MB at depth 8), after it beyond.
constexpr int kDenseEarlyDepth = 8;
inline size_t dense_grid_cells(int D) { return ((size_t(1) << (3 * (D + 1))) - 1) / 7; }

struct Mesh {
    float* vertices = nullptr;  // xyz
    int* triangles = nullptr;
    unsigned* keys = nullptr;   // per vertex: the global key of its depth-D edge (segment_key)
    int nv = 0, nt = 0;
    bool valid = false;         // the last reconstruction on this context completed
    // the welded form (weld_mesh): one vertex per depth-D edge
    float* weldedVertices = nullptr;
    int* weldedTriangles = nullptr;
    int weldedNv = 0, weldedNt = 0;
    // Counts still in the mailbox when extraction queued its last kernels without
    // waiting: kept vertices = slot 3 + slot 1, triangles = triBefore + slot 2,
    // each within its allocated bound.
    bool deferred = false;
    int triBefore = 0;
    size_t vCap = 0, tCap = 0;
};

// The depth-D marching cubes' outputs and the counts read after its wait.
struct MainMc {
    int* vexAddr = nullptr;   // crossing index per depth-D edge
    int* cube = nullptr;      // cube category per depth-D node (exported)
    int* hasInter = nullptr;  // surface-edge intersection flag per face
    int* roots = nullptr;     // subdivision roots, ascending node index
    int* rootPos = nullptr;   // per node above depth D: position in roots, or -1
    int nCross = 0, nTri = 0, nRoots = 0;
    int rootsPerDepth[PSR_MAX_DEPTH + 1] = {0};
    bool q1Fired = false;  // the level-local child 0 rule changed a parent somewhere
};

// Small per-call results written by the kernels and read back in one copy.
struct DevResults {
    int cgIters[2 * (PSR_MAX_DEPTH + 1)];  // inner iterations | outer steps, per depth
    int cgNnz[PSR_MAX_DEPTH + 1];          // stored operator entries per depth
    int cgErr;                             // a CG group barrier timed out
    unsigned cgSlotMask[PSR_MAX_DEPTH + 1];  // per depth: the neighbour slots some row stores
    int q1Fired;                           // the level-local child 0 rule changed a parent
    float iso;                             // iso-value
    double cgResid[PSR_MAX_DEPTH + 1];     // final updated residual norm per depth
    double cgTrueRel[PSR_MAX_DEPTH + 1];   // mixed form: final true relative residual per depth
};
static_assert(offsetof(DevResults, cgErr) == offsetof(DevResults, cgNnz) + sizeof(int) * (PSR_MAX_DEPTH + 1) &&
                  offsetof(DevResults, cgErr) == sizeof(int) * (3 * (PSR_MAX_DEPTH + 1)) &&
                  offsetof(DevResults, cgSlotMask) == offsetof(DevResults, cgErr) + sizeof(int),
              "the operator counts, the CG error flag and the slot masks are cleared as one range");
constexpr int kResultsClearWords = 2 * (PSR_MAX_DEPTH + 1) + 1;  // from cgNnz

// Read-back mailbox slots (Ctx::hostPinned / Ctx::mail).  Each stage's
// kernels write the values the host needs next; the host reads them after
// Ctx::wait(), before the next stage's kernels are queued.
constexpr int kMailWords = 64;
constexpr int kMailNorm = 32;                // octree: 4 normalisation parameters (float bits)
constexpr int kMailLevelStart = 4;           // depth-D MC: roots before each level
constexpr int kMailQ1 = PSR_MAX_DEPTH + 8;   // depth-D MC: the level-local child 0 rule fired
constexpr int kMailGeo = 40;                 // geometry: vertex, edge and face counts
constexpr int kMailScanErr = 63;             // any call: a scan's look-back reached its spin bound
static_assert(kMailQ1 < kMailNorm && kMailNorm + 4 <= kMailGeo && kMailGeo + 3 <= kMailWords,
              "mailbox slots overlap");

struct Ctx {
    psr_config cfg{};
    hipStream_t stream = nullptr;
    bool ownStream = false;
    int device = 0;
    Arena arena;
    Tables tables;
    Octree oct;
    Field fld;
    Geometry geo;
    Mesh mesh;
    MainMc mc;  // the depth-D marching cubes of the call in progress
    psr_stats stats{};
    int* scratch = nullptr;
    size_t scratchInts = 0;
    // CG group controls and slots: zero unless a call left them in use
    bool cgClean = false;
    void* cgCtl = nullptr;
    void* cgSlots = nullptr;
    size_t cgCtlWords = 0, cgSlotWords = 0;
    hipEvent_t ev[16] = {};
    int* hostPinned = nullptr;   // read-back mailbox: pinned, coherent, kMailWords ints
    int* mail = nullptr;         // device view of hostPinned: kernels write read-back values there
    DevResults* hostResults = nullptr;  // pinned, coherent mirror of results()
    DevResults* mailResults = nullptr;  // device view of hostResults: written once the results are final
    double cgMsSum = 0.0;               // CG kernel time summed over completed runs
    uint32_t runs = 0;                  // completed runs
    void* hostStage = nullptr;          // pinned, mapped staging for small uploads
    void* stageDev = nullptr;           // device view of hostStage
    size_t hostStageBytes = 0;
    int cgMaxWgs = 0;            // resident workgroups of the CG kernels with several rows per thread
    int cgMaxWgsRow = 0;         // resident workgroups of the one-row-per-thread CG kernels
    int cgMaxWgsMinres = 0;      // resident workgroups of the MINRES kernel (indefinite depths, mixed form)
    int cgRowsPerWg = 0;         // rows per CG workgroup (0: default)
    int cgRowTarget = 0;         // batch: rows a CG workgroup takes (0: the single-cloud plan)
    int cgSpinLimit = -1;        // spin bound of the CG barriers (-1: the default)
    bool indefinite[PSR_MAX_DEPTH + 1] = {};  // the depth's operator after the 1e-6 drop is indefinite
    template <class T>
    T* buf(const char* name, size_t count) {
        return static_cast<T*>(arena.get(name, sizeof(T) * (count ? count : 1)));
    }
    template <class T>
    T* buf_keep(const char* name, size_t count, size_t keepCount) {
        return static_cast<T*>(arena.get_keep(name, sizeof(T) * (count ? count : 1), sizeof(T) * keepCount));
    }
    DevResults* results() { return buf<DevResults>("results", 1); }
    // The scan-error word of the mailbox: cleared by the host when a call
    // starts (no kernel of the context is in flight then), read after its
    // last wait.
    void clear_scan_error() { hostPinned[kMailScanErr] = 0; }
    bool scan_failed() const { return __atomic_load_n(&hostPinned[kMailScanErr], __ATOMIC_ACQUIRE) != 0; }
    // Waits for the stream by polling it: a blocking wait adds the host thread's
    // wake-up latency to every read-back.  Errors surface through the final
    // synchronisation.
    void wait() {
        while (hipStreamQuery(stream) == hipErrorNotReady) {
        }
        PSR_HIP_CHECK(hipStreamSynchronize(stream));
    }
    // Pinned, mapped, coherent staging of at least `bytes` (device view in stageDev); only
    // reused after a stream synchronisation.
    void* stage(size_t bytes) {
        if (bytes > hostStageBytes) {
            if (hostStage) PSR_HIP_CHECK(hipHostFree(hostStage));
            hostStage = stageDev = nullptr;
            hostStageBytes = 0;
            PSR_HIP_CHECK(hipHostMalloc(&hostStage, bytes + bytes / 2 + 4096, hipHostMallocMapped | hipHostMallocCoherent));
            PSR_HIP_CHECK(hipHostGetDevicePointer(&stageDev, hostStage, 0));
            hostStageBytes = bytes + bytes / 2 + 4096;
        }
        return hostStage;
    }
    ScanState scan;
    // The stream's scan state, able to scan n elements (zeroed when it grows).
    ScanState& scan_for(size_t n) { return scan_for_tiles(scan_tiles(n)); }
    // the scan state for a launch of `tiles` workgroups that scan their own counts
    ScanState& scan_for_tiles(size_t tiles) {
        if (tiles > scan.capacity || scan.epoch + kScanRing >= kScanEpochMax) {
            const size_t cap = std::max<size_t>(tiles + tiles / 2 + 64, 1024);
            scan.status = static_cast<unsigned long long*>(arena.get("scan_status", cap * sizeof(unsigned long long)));
            scan.counters = static_cast<unsigned*>(arena.get("scan_counters", kScanRing * sizeof(unsigned)));
            PSR_HIP_CHECK(hipMemsetAsync(scan.status, 0, cap * sizeof(unsigned long long), stream));
            PSR_HIP_CHECK(hipMemsetAsync(scan.counters, 0, kScanRing * sizeof(unsigned), stream));
            scan.capacity = cap;
            scan.epoch = 0;
        }
        scan.err = mail + kMailScanErr;
        return scan;
    }
    int* scratch_for(size_t ints) {
        if (ints > scratchInts) {
            scratchInts = ints + ints / 4 + 1024;
            scratch = static_cast<int*>(arena.get("scratch", scratchInts * sizeof(int)));
        }
        return scratch;
    }
};

// Thrown by a stage driver that cannot be recorded for a batch (it waits
// for the host in the middle); the batch runs that cloud's stage on its own.
struct NotRecordable {};

// Stage drivers.
void build_tables(Ctx& c);
void build_octree(Ctx& c, const psr_points& pts);
// build_octree in two parts: up to the level counts, and after their wait
void build_octree_a(Ctx& c, const psr_points& pts);
void build_octree_b(Ctx& c);
void compute_field(Ctx& c);
void solve_depths(Ctx& c);
// Drops a context's CG completion event from the batch-worker set (before the
// event is destroyed).
void cg_forget_event(hipEvent_t e);
// A batch's merged CG launches take the whole device: cg_admit_device waits
// until no CG launch of another context is in flight; cg_hold_device then
// holds the device until `done` (recorded after the last of them) fires.
void cg_admit_device(hipEvent_t self);
// Resident workgroups of the batched several-rows CG kernel of c's precision.
int cg_batch_budget(Ctx& c);
void cg_hold_device(hipEvent_t done);
void points_implicit(Ctx& c);
void build_geometry(Ctx& c);
void extract_mesh(Ctx& c);
// extract_mesh in two parts: the depth-D marching cubes, and after their wait
// the subdivision and the output
void extract_mesh_a(Ctx& c);
void extract_mesh_b(Ctx& c);
// Mesh counts of a deferred extraction, read from the mailbox after the final wait.
void finish_mesh(Ctx& c);
// Checks a cloud and resets a context's per-call state (PSR_OK, or a status
// with the message in err).
int call_begin(Ctx& c, const psr_points& pts, std::string& err);
// After a call's last wait: statistics and the mesh (timed: the stage events
// were recorded on the context's stream).
int call_end(Ctx& c, psr_mesh* out, bool timed, std::string& err);
float elapsed(hipEvent_t a, hipEvent_t b);
// Welds the last mesh into mesh.welded*; returns once the counts are known.
void weld_mesh(Ctx& c);
// Per-vertex normals of a device mesh into out (3 floats per vertex); returns once written.
void mesh_normals(Ctx& c, const float* V, int nv, const int* T, int nt, float* out);
// The same with every scratch buffer in a file-scope device array and no
// context: allocates nothing; returns a status (err: its message).
int mesh_normals_static(const float* V, int nv, const int* T, int nt, float* out, hipStream_t s, std::string& err);
size_t mesh_normals_static_max_triangles();
// Normals (and curvature, k-nearest-neighbour indices when non-null) of an
// unoriented point cloud; returns once written.
void estimate_normals(Ctx& c, const float* xyz, int n, unsigned stride, int k, float3 vp, float* normals,
                      float* curvature, int* knn);
// Colours of nv mesh vertices from their 4 nearest of n coloured points; returns once written.
void vertex_colors(Ctx& c, const float* V, int nv, const float* xyz, unsigned stride, const float* pc, int n,
                   float* out);

}  // namespace psr

struct psr_ctx;
namespace psr {
class Batch;
Batch* batch_create(psr_ctx* owner);
void batch_destroy(Batch* b);
// Reconstructs count clouds with one launch per stage (psr_batch.cpp); outs[i]
// and stats[i] are written for every cloud that succeeds, the first failure is
// the status (its message in err).
int batch_run(Batch* b, const psr_points* pts, uint32_t count, psr_mesh* outs, std::vector<psr_stats>& stats,
              std::string& err);
// The CG plan of the batch's last call (psr_config.batch_cg_plan).
int batch_last_plan(const Batch* b);
// The context of cloud i of the last batch call (null if none).
psr_ctx* batch_cloud(const Batch* b, uint32_t i);
}

// The C ABI's context: one reconstruction context, and the batched form's
// state (its clouds' contexts share this context's stream).
struct psr_ctx {
    psr::Ctx c;
    std::string lastError;
    psr::Batch* batch = nullptr;
    std::vector<psr_stats> batchStats;
};
