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
    int allocations() const { return allocs_; }

  private:
    struct Block {
        void* ptr = nullptr;
        size_t bytes = 0;
    };
    std::map<std::string, Block> blocks_;
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
    float4* nbrCol = nullptr;  // [27][M]: centre (x, y, z) and d_x of neighbour i of each node (0 where absent)
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

// Small per-call results written by the kernels and read back in one copy.
struct DevResults {
    int cgIters[2 * (PSR_MAX_DEPTH + 1)];  // inner iterations | outer steps, per depth
    int cgNnz[PSR_MAX_DEPTH + 1];          // stored operator entries per depth
    int cgErr;                             // a CG group barrier timed out
    int q1Fired;                           // the level-local child 0 rule changed a parent
    float iso;                             // iso-value
    double cgResid[PSR_MAX_DEPTH + 1];     // final updated residual norm per depth
    double cgTrueRel[PSR_MAX_DEPTH + 1];   // mixed form: final true relative residual per depth
};
static_assert(offsetof(DevResults, cgErr) == offsetof(DevResults, cgNnz) + sizeof(int) * (PSR_MAX_DEPTH + 1) &&
                  offsetof(DevResults, cgErr) == sizeof(int) * (3 * (PSR_MAX_DEPTH + 1)),
              "the operator counts and the CG error flag are cleared as one range");

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
    int cgRowsPerWg = 0;         // rows per CG workgroup (0: default)
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

// Stage drivers.
void build_tables(Ctx& c);
void build_octree(Ctx& c, const psr_points& pts);
void compute_field(Ctx& c);
void solve_depths(Ctx& c);
// Drops a context's CG completion event from the batch-worker set (before the
// event is destroyed).
void cg_forget_event(hipEvent_t e);
void points_implicit(Ctx& c);
void build_geometry(Ctx& c);
void extract_mesh(Ctx& c);
// Mesh counts of a deferred extraction, read from the mailbox after the final wait.
void finish_mesh(Ctx& c);
// Welds the last mesh into mesh.welded*; returns once the counts are known.
void weld_mesh(Ctx& c);
// Per-vertex normals of a device mesh into out (3 floats per vertex); returns once written.
void mesh_normals(Ctx& c, const float* V, int nv, const int* T, int nt, float* out);
// Normals (and curvature, k-nearest-neighbour indices when non-null) of an
// unoriented point cloud; returns once written.
void estimate_normals(Ctx& c, const float* xyz, int n, unsigned stride, int k, float3 vp, float* normals,
                      float* curvature, int* knn);
// Colours of nv mesh vertices from their 4 nearest of n coloured points; returns once written.
void vertex_colors(Ctx& c, const float* V, int nv, const float* xyz, unsigned stride, const float* pc, int n,
                   float* out);

}  // namespace psr
