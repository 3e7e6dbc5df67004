// Header-only C++ facade over the C ABI (psr.h) with the reference's class and
// method names, for code written against
//   SparseSurfelFusion::PoissonReconstruction::SolvePoissionReconstructionMesh(
//       DeviceArrayView<DepthSurfel>)          (mesh/PoissonReconstruction.h:124)
// and the mesh getters of ComputeTriangleIndices
//   GetRebuildMeshVertices / GetRebuildMeshTriangleIndices
//   (mesh/ComputeTriangleIndices.h:331-342),
// which here sit on PoissonReconstruction itself (they replace the OpenGL
// DrawRebuildMesh, mesh/PoissonReconstruction.h:129).
//
// The value types below have the reference's layouts (DepthSurfel 64 B with
// xyz at +16 and the normal at +32, base/CommonTypes.h:113-118; Point3D<float>
// 12 B, mesh/Geometry.h:19-20; TriangleIndex 12 B, mesh/Geometry.h:235-238),
// so a caller's device buffers pass through unchanged.  Errors throw
// std::runtime_error with the library's message instead of exit(1).
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "psr.h"

namespace SparseSurfelFusion {

struct PixelCoordinate {
    unsigned int row, col;
};

struct alignas(16) Float4 {
    float x, y, z, w;
};

struct DepthSurfel {
    PixelCoordinate pixelCoordinate;
    Float4 VertexAndConfidence;
    Float4 NormalAndRadius;
    Float4 ColorAndTime;
};
static_assert(sizeof(DepthSurfel) == 64, "DepthSurfel layout");
static_assert(offsetof(DepthSurfel, VertexAndConfidence) == 16, "DepthSurfel layout");
static_assert(offsetof(DepthSurfel, NormalAndRadius) == 32, "DepthSurfel layout");

template <class T>
struct Point3D {
    T coords[3];
};

class TriangleIndex {
  public:
    int idx[3];
};

// Non-owning view of a device array (base/DeviceReadWrite/DeviceArrayView.h).
template <class T>
class DeviceArrayView {
  public:
    DeviceArrayView() : deviceArray(nullptr), deviceArraySize(0) {}
    DeviceArrayView(const T* arr, size_t size) : deviceArray(arr), deviceArraySize(size) {}
    DeviceArrayView(const T* arr, size_t start, size_t end) : deviceArray(arr + start), deviceArraySize(end - start) {}
    size_t Size() const { return deviceArraySize; }
    size_t ByteSize() const { return deviceArraySize * sizeof(T); }
    const T* RawPtr() const { return deviceArray; }

  private:
    const T* deviceArray;
    size_t deviceArraySize;
};

class PoissonReconstruction {
  public:
    // hipStream: a hipStream_t to run on, or nullptr for a stream of its own.
    explicit PoissonReconstruction(const psr_config* cfg = nullptr, void* hipStream = nullptr) {
        psr_config c;
        psr_default_config(&c);
        if (cfg) c = *cfg;
        int rc = psr_create(&c, hipStream, &ctx_);
        if (rc != PSR_OK) throw std::runtime_error(std::string("psr_create: ") + psr_status_string(rc));
    }
    ~PoissonReconstruction() { psr_destroy(ctx_); }
    PoissonReconstruction(const PoissonReconstruction&) = delete;
    PoissonReconstruction& operator=(const PoissonReconstruction&) = delete;

    // Oriented points as the reference's DepthSurfel records (device memory).
    void SolvePoissionReconstructionMesh(DeviceArrayView<DepthSurfel> denseSurfel) {
        const char* base = reinterpret_cast<const char*>(denseSurfel.RawPtr());
        psr_points p;
        p.xyz = reinterpret_cast<const float*>(base + offsetof(DepthSurfel, VertexAndConfidence));
        p.normal = reinterpret_cast<const float*>(base + offsetof(DepthSurfel, NormalAndRadius));
        p.count = static_cast<uint32_t>(denseSurfel.Size());
        p.stride_bytes = sizeof(DepthSurfel);
        run(p);
    }

    // Packed float3 positions and normals (device memory), count points each.
    void SolvePoissionReconstructionMesh(const float* xyz, const float* normals, uint32_t count) {
        psr_points p{xyz, normals, count, 12};
        run(p);
    }

    // Valid until the next reconstruction or destruction (normalised frame, as the reference).
    DeviceArrayView<Point3D<float>> GetRebuildMeshVertices() const {
        return DeviceArrayView<Point3D<float>>(reinterpret_cast<const Point3D<float>*>(mesh_.vertices),
                                               mesh_.n_vertices);
    }
    DeviceArrayView<TriangleIndex> GetRebuildMeshTriangleIndices() const {
        return DeviceArrayView<TriangleIndex>(reinterpret_cast<const TriangleIndex*>(mesh_.triangles),
                                              mesh_.n_triangles);
    }

    // One vertex per depth-D edge (psr_weld_mesh; the reference keeps its
    // batches' duplicates).  The welded views are valid until the next
    // reconstruction.
    void WeldMesh() { check(psr_weld_mesh(ctx_, &welded_)); }
    DeviceArrayView<Point3D<float>> GetWeldedMeshVertices() const {
        return DeviceArrayView<Point3D<float>>(reinterpret_cast<const Point3D<float>*>(welded_.vertices),
                                               welded_.n_vertices);
    }
    DeviceArrayView<TriangleIndex> GetWeldedMeshTriangleIndices() const {
        return DeviceArrayView<TriangleIndex>(reinterpret_cast<const TriangleIndex*>(welded_.triangles),
                                              welded_.n_triangles);
    }

    // Per-vertex normals of the mesh (or of its welded form) into a device
    // buffer of 3 floats per vertex (DrawMesh::CalculateMeshNormals,
    // mesh/DrawMesh.cpp).
    void CalculateMeshNormals(float* normalsDevice, bool welded = false) {
        check(psr_mesh_normals(ctx_, welded ? &welded_ : &mesh_, normalsDevice));
    }

    // Normals of an unoriented cloud in device memory, before a reconstruction
    // (the reference's host-side PCL step, PoissonReconstruction.cpp:137-186).
    void CalculatePointCloudNormal(const float* xyzDevice, uint32_t count, float* normalsDevice, int k = 10) {
        const float vp[3] = {0.f, 0.f, 0.f};
        check(psr_estimate_normals(ctx_, xyzDevice, count, 12, k, vp, normalsDevice, nullptr, nullptr));
    }

    // The mesh (or its welded form) as a binary PLY file, in the normalised frame.
    void WriteMeshPly(const char* path, bool welded = false) {
        const int vid = welded ? PSR_ARRAY_WELDED_VERTICES : PSR_ARRAY_MESH_VERTICES;
        const int tid = welded ? PSR_ARRAY_WELDED_TRIANGLES : PSR_ARRAY_MESH_TRIANGLES;
        uint64_t nvb = 0, ntb = 0;
        check(psr_export(ctx_, vid, nullptr, 0, &nvb));
        check(psr_export(ctx_, tid, nullptr, 0, &ntb));
        std::vector<float> v(nvb / sizeof(float));
        std::vector<int32_t> t(ntb / sizeof(int32_t));
        check(psr_export(ctx_, vid, v.data(), nvb, &nvb));
        check(psr_export(ctx_, tid, t.data(), ntb, &ntb));
        check(psr_write_ply(path, v.data(), uint32_t(v.size() / 3), t.data(), uint32_t(t.size() / 3)));
    }

    psr_stats Stats() const {
        psr_stats s;
        psr_get_stats(ctx_, &s);
        return s;
    }
    psr_ctx* Context() const { return ctx_; }

  private:
    void check(int rc) const {
        if (rc != PSR_OK) throw std::runtime_error(std::string(psr_status_string(rc)) + ": " + psr_last_error(ctx_));
    }
    void run(const psr_points& p) { check(psr_reconstruct(ctx_, &p, &mesh_)); }
    psr_ctx* ctx_ = nullptr;
    psr_mesh mesh_{};
    psr_mesh welded_{};
};

}  // namespace SparseSurfelFusion
