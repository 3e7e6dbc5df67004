/* C ABI of the MI355X Poisson surface reconstruction.
 *
 * Replaces the reference's class entry point
 *   void SparseSurfelFusion::PoissonReconstruction::SolvePoissionReconstructionMesh(
 *        DeviceArrayView<DepthSurfel> denseSurfel);
 *   (mesh/PoissonReconstruction.h:124, mesh/PoissonReconstruction.cpp:195-245)
 * and its output getters
 *   ComputeTriangleIndices::GetRebuildMeshVertices / GetRebuildMeshTriangleIndices
 *   (mesh/ComputeTriangleIndices.h:331-342).
 * Plain pointers and sizes only.  Input point buffers are device pointers
 * borrowed for the call; the mesh buffers are device memory owned by the
 * context and valid until the next psr_reconstruct or psr_destroy.
 */
#ifndef PSR_H
#define PSR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psr_ctx psr_ctx;

/* Status codes.  A CG that stops at its iteration limit is reported in
 * psr_stats, not as an error (the reference is silent there,
 * mesh/solver/CGAlgorithm.cu:36). */
enum {
    PSR_OK = 0,
    PSR_ERR_INVALID = 1,  /* bad argument or degenerate input (reference divides by 0, BuildOctree.cu:149) */
    PSR_ERR_CAPACITY = 2, /* input exceeds a configured capacity (reference throws, DeviceBufferArray.h:125-132) */
    PSR_ERR_HIP = 3,      /* HIP runtime error (reference calls exit(1), safe_call.hpp:58-71) */
    PSR_ERR_INTERNAL = 4
};

enum {
    PSR_PRECISION_FP32 = 0, /* the reference's FP32 CG (CGAlgorithm.cu:10-80) */
    PSR_PRECISION_MIXED = 1 /* FP16 field storage, compensated FP32 dots, FP64 outer refinement */
};

typedef struct {
    int depth;             /* maximum octree depth D, 2..9 (MAX_DEPTH_OCTREE, GlobalConfigs.h:34) */
    uint32_t max_points;   /* capacity; 0 = no limit (MAX_SURFEL_COUNT, GlobalConfigs.h:7) */
    int precision;         /* PSR_PRECISION_* */
    float cg_tol;          /* absolute tolerance on the updated residual norm, 1e-5 (CGAlgorithm.cu:298) */
    int cg_max_iter;       /* 10000 (CGAlgorithm.cu:15) */
    float refine_rel_tol;  /* PSR_PRECISION_MIXED: target true relative residual per depth, 1e-5 */
    int refine_max_outer;  /* PSR_PRECISION_MIXED: outer refinement steps, 4 */
    uint32_t subdiv_chunk_nodes; /* virtual depth-D cells per subdivision working set; 0 = 16M
                                    (about 3 GB); results do not depend on it */
    uint32_t batch_workers;      /* psr_reconstruct_batch: clouds merged into each launch (a pass);
                                    0 = 64 */
    int32_t batch_cg_plan;       /* psr_reconstruct_batch, how the clouds' CG solves are laid out:
                                    0 = the single-cloud plan (256 rows a workgroup, in as many waves
                                    of the resident budget as needed: meshes equal psr_reconstruct's
                                    bit for bit); w >= 1 = every depth's rows split evenly over
                                    workgroups so that all solves fit w waves; -1 (default) = picked
                                    per batch shape by the built-in tuner, which times the plans on
                                    the first calls of each shape */
} psr_config;

typedef struct {
    const float* xyz;      /* device pointer, x y z floats at the start of each record */
    const float* normal;   /* device pointer, nx ny nz floats at the start of each record */
    uint32_t count;
    uint32_t stride_bytes; /* 12 for packed float3; 64 for the reference's DepthSurfel (xyz at +16, normal at +32) */
} psr_points;

typedef struct {
    const float* vertices;     /* device, 3 floats per vertex, normalised frame (reference Q11) */
    const int32_t* triangles;  /* device, 3 vertex indices per triangle */
    uint32_t n_vertices;
    uint32_t n_triangles;
} psr_mesh;

#define PSR_STATS_MAX_DEPTH 10
typedef struct {
    int depth;
    uint32_t n_points;
    uint32_t n_nodes;
    uint32_t nodes_per_depth[PSR_STATS_MAX_DEPTH + 1];
    int cg_iterations[PSR_STATS_MAX_DEPTH + 1];
    double cg_updated_residual[PSR_STATS_MAX_DEPTH + 1];
    double cg_true_relres[PSR_STATS_MAX_DEPTH + 1]; /* PSR_PRECISION_MIXED: final FP64 ||b - A x|| / ||b||; -1 otherwise */
    float iso_value;
    uint32_t n_vertex_nodes, n_edge_nodes, n_face_nodes;
    uint32_t n_subdivision_roots, n_coarse_roots;
    uint32_t n_mesh_vertices, n_mesh_triangles;
    int q1_fired;         /* the reference's level-local child 0 parent rule applied somewhere */
    float ms_octree, ms_field, ms_solve, ms_mesh, ms_total;
    float ms_cg_kernel;   /* the concurrent CG kernel alone, hipEvents on the context's stream */
    uint32_t cg_workgroups;
    int cg_nnz[PSR_STATS_MAX_DEPTH + 1]; /* stored operator entries per depth */
    int cg_outer[PSR_STATS_MAX_DEPTH + 1]; /* PSR_PRECISION_MIXED: outer refinement steps per depth; 1 otherwise */
    int cg_row_form;      /* 1: one row per CG thread, stencil held in registers; 0: several rows per thread */
    double ms_cg_kernel_sum; /* ms_cg_kernel summed over the context's completed runs */
    uint32_t n_runs;         /* completed runs of the context */
    uint64_t device_bytes;   /* device memory the run's buffers needed (the context's fixed tables and
                                workspace excluded); the context's grow-only arena holds at least this */
} psr_stats;

/* Arrays that psr_export can copy to host memory (for parity checks). */
enum {
    PSR_ARRAY_SORTED_POS = 0,   /* float4 per sorted point (w = 0) */
    PSR_ARRAY_SORTED_NRM,       /* float4 per sorted point */
    PSR_ARRAY_CODES,            /* int32 Morton code per sorted point */
    PSR_ARRAY_ORDER,            /* int32 input index per sorted point */
    PSR_ARRAY_POINT2NODE,       /* int32 level-local depth-D node per sorted point */
    PSR_ARRAY_NODE_KEY,         /* int32 per node */
    PSR_ARRAY_NODE_PIDX,
    PSR_ARRAY_NODE_PNUM,
    PSR_ARRAY_NODE_PARENT,
    PSR_ARRAY_NODE_CHILDREN,    /* int32 [M][8] */
    PSR_ARRAY_NODE_NEIGHS,      /* int32 [M][27] */
    PSR_ARRAY_NODE_DIDX,
    PSR_ARRAY_NODE_DNUM,
    PSR_ARRAY_NODE_ENCIDX,
    PSR_ARRAY_LEVEL_BASE,       /* int32 [D+2] */
    PSR_ARRAY_LEVEL_COUNT,      /* int32 [D+1] */
    PSR_ARRAY_NORM_PARAMS,      /* float corner xyz, edge */
    PSR_ARRAY_VFIELD,           /* float [nD][3] */
    PSR_ARRAY_DIVERGENCE,       /* float [M] */
    PSR_ARRAY_DX,               /* float [M] */
    PSR_ARRAY_CHI_POINTS,       /* float [N] */
    PSR_ARRAY_VERTS,            /* int32 [V][6]: x y z (float bits), owner, kind, depth */
    PSR_ARRAY_EDGES,            /* int32 [E][2]: kind, owner */
    PSR_ARRAY_FACES,            /* int32 [F][3]: kind, owner, hasParentFace */
    PSR_ARRAY_NODE_VERTICES,    /* int32 [M][8], 1-based */
    PSR_ARRAY_NODE_EDGES,       /* int32 [M][12], 1-based, 0 above depth D */
    PSR_ARRAY_NODE_FACES,       /* int32 [M][6], 1-based */
    PSR_ARRAY_VVALUE,           /* float [V] */
    PSR_ARRAY_MESH_VERTICES,    /* float [n_vertices][3] */
    PSR_ARRAY_MESH_TRIANGLES,   /* int32 [n_triangles][3] */
    PSR_ARRAY_TABLE_DOT,        /* double [res][res] */
    PSR_ARRAY_TABLE_DDOT,
    PSR_ARRAY_TABLE_D2DOT,
    PSR_ARRAY_CUBE,             /* int32 [nD] depth-D cube categories */
    PSR_ARRAY_MESH_VERTEX_KEYS, /* uint32 [n_vertices]: each vertex's depth-D edge, ((x S + y) S + z) 3 + axis,
                                   S = 2^D + 1, (x, y, z) the edge's lower grid corner */
    PSR_ARRAY_WELDED_VERTICES,  /* float [welded n_vertices][3], after psr_weld_mesh */
    PSR_ARRAY_WELDED_TRIANGLES, /* int32 [n_triangles][3], after psr_weld_mesh */
    PSR_ARRAY_COUNT_
};

void psr_default_config(psr_config* cfg);
/* hip_stream: a hipStream_t to work on, or NULL for a context-owned stream. */
int psr_create(const psr_config* cfg, void* hip_stream, psr_ctx** out);
/* Runs the whole pipeline; on return the mesh counts are final and the buffers
 * are ready on the context's stream. */
int psr_reconstruct(psr_ctx* ctx, const psr_points* points, psr_mesh* out);
/* Reconstructs count independent clouds of uneven size with one launch per
 * stage: the clouds' sorts run as one segmented sort, every stage's kernels
 * take all clouds' work in one grid, and the per-depth CG solves of all clouds
 * run as independent groups of one persistent launch (waves within the
 * resident budget), each group stopping on its own convergence.  A phase's
 * launches are replayed from a captured HIP graph when their structure
 * recurs.  cfg.batch_workers clouds are merged per pass (0: 64).  outs[i]
 * receives cloud i's mesh (device buffers owned by the context, valid until
 * the next batch call or psr_destroy); the call returns when every mesh is
 * complete.  A failing cloud stops no other: the call returns the first
 * failure (its cloud named in psr_last_error) with the others' meshes
 * written.  No counterpart in the reference, which reconstructs one cloud per
 * call (mesh/PoissonReconstruction.cpp:195-245). */
int psr_reconstruct_batch(psr_ctx* ctx, const psr_points* clouds, uint32_t count, psr_mesh* outs);
/* Statistics of cloud `index` of the last batch. */
int psr_get_batch_stats(psr_ctx* ctx, uint32_t index, psr_stats* out);
/* The context that reconstructed cloud `index` of the last batch (for
 * psr_export and psr_get_stats; owned by ctx, valid until its next batch call
 * or psr_destroy), or NULL. */
psr_ctx* psr_batch_cloud(psr_ctx* ctx, uint32_t index);
/* The CG plan the last batch call ran with (see psr_config.batch_cg_plan). */
int psr_batch_cg_plan(const psr_ctx* ctx);
int psr_sync(psr_ctx* ctx);
int psr_get_stats(psr_ctx* ctx, psr_stats* out);
/* Copies array `array_id` of the last reconstruction to host memory.  With
 * host_dst == NULL only *needed (bytes) is written. */
int psr_export(psr_ctx* ctx, int array_id, void* host_dst, uint64_t bytes, uint64_t* needed);
/* Welds the mesh of the last psr_reconstruct on ctx: one vertex per depth-D
 * edge (the first of the vertices carrying that edge's key, in ascending key
 * order), triangle corners remapped (a corner index outside the mesh becomes
 * -1).  out receives device buffers owned by the context, valid until the
 * next reconstruction.  No counterpart in the reference, which leaves the
 * duplicates of its marching-cubes batches in place
 * (mesh/ComputeTriangleIndices.cu:1247-1282). */
int psr_weld_mesh(psr_ctx* ctx, psr_mesh* out);
/* Per-vertex normals of a mesh in device memory (any psr_mesh: the last
 * reconstruction's or its welded form), written to normals, a DEVICE buffer
 * of 3 * n_vertices floats; returns once they are written.  Follows
 * mesh/DrawMesh.cu:80-177: each triangle's unit normal cross(A - O, B - O),
 * summed over the triangles touching a vertex, divided by their count and
 * normalised.  Sums run in triangle order (the reference adds with float
 * atomics, in no fixed order); a vertex no triangle touches gets (0, 0, 0)
 * (the reference divides 0 by 0 there); a triangle with a corner outside the
 * mesh contributes nothing. */
int psr_mesh_normals(psr_ctx* ctx, const psr_mesh* mesh, float* normals);
/* psr_mesh_normals without a context and without allocating: no hipMalloc,
 * hipFree or allocating library call.  Its scratch (face normals, sort
 * buffers, histograms, scan state) lives in one file-scope device array of
 * about 17 MB that each GPU holds until the process ends.  No counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  Meshes of up to
 * psr_mesh_normals_static_max_triangles() triangles and 2^24 - 1 vertices
 * (PSR_ERR_CAPACITY beyond).  Stream-ordered on hip_stream (NULL: the null
 * stream); returns once the normals are written.  Every call on a device
 * shares that device's array: calls that overlap on one device race. */
int psr_mesh_normals_static(const psr_mesh* mesh, float* normals, void* hip_stream);
uint32_t psr_mesh_normals_static_max_triangles(void);
/* Normals of an unoriented point cloud in device memory (xyz at the start of
 * each stride_bytes record), the step before psr_reconstruct: replaces the
 * reference's PCL NormalEstimationOMP (mesh/PoissonReconstruction.cpp:137-186,
 * k = 10 from a kd-tree, viewpoint (0, 0, 0)).  For each point: its k nearest
 * points (itself included; exact, ties to the lower index), their covariance,
 * the eigenvector of the smallest eigenvalue, flipped so that
 * (viewpoint - p) . n >= 0.  normals: DEVICE, 3 floats per point; curvature
 * (lambda_0 / sum lambda, may be NULL): DEVICE, 1 float per point; knn (may be
 * NULL): DEVICE, k int32 per point, nearest first.  1 <= k <= 16.  Returns
 * once written.  No counterpart on the GPU in the reference. */
int psr_estimate_normals(psr_ctx* ctx, const float* xyz, uint32_t count, uint32_t stride_bytes, int k,
                         const float viewpoint[3], float* normals, float* curvature, int32_t* knn);
/* Per-vertex colours of a device mesh from coloured points (mesh/DrawMesh.cu:
 * 35-79, 179-211): each vertex takes its 4 nearest points (exact; the
 * reference searches all points by brute force) and weights each point's
 * colour by its squared distance over the four's sum, as the reference does.
 * xyz: DEVICE, in the mesh's frame, xyz at the start of each stride_bytes
 * record; point_colors: DEVICE, 3 floats per point; colors: DEVICE, 3 floats
 * per vertex.  Returns once written. */
int psr_vertex_colors(psr_ctx* ctx, const psr_mesh* mesh, const float* xyz, const float* point_colors,
                      uint32_t count, uint32_t stride_bytes, float* colors);
/* Writes a binary little-endian PLY (float x y z; uchar-counted int32 face
 * lists) from HOST arrays.  No counterpart in the reference, which renders
 * the mesh (mesh/DrawMesh.cpp) and writes no file. */
int psr_write_ply(const char* path, const float* vertices, uint32_t n_vertices, const int32_t* triangles,
                  uint32_t n_triangles);
void psr_destroy(psr_ctx* ctx);
const char* psr_status_string(int status);
const char* psr_last_error(const psr_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* PSR_H */
