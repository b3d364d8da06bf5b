// voxmi_voxelizer.hpp -- C++ host class over the C ABI in voxmi.h with the
// member names, argument meaning and return conventions of the reference's
// cuda::Voxelizer (apecoraro/VoxVizThesis VoxViz/Voxelizer/Voxelizer/
// Voxelizer.h 22-244), so that main.cu's Voxelize() (3225-3341) and
// PagedGigaVoxelOctTreeGenerator::generate (4135-4405) compile against it with
// the type names changed.  Header only; links against libvoxmi.so.
#ifndef VOXMI_VOXELIZER_HPP
#define VOXMI_VOXELIZER_HPP

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "voxmi.h"

namespace voxmi {

using vec2 = std::array<float, 2>;
using vec3 = std::array<float, 3>;
using uvec3 = std::array<uint32_t, 3>;

// cudaTextureAddressMode numbering (main.cu OpenGLModeToCudaMode 3207-3223).
enum AddressMode { AddressModeWrap = 0, AddressModeClamp = 1, AddressModeMirror = 2,
                   AddressModeBorder = 3 };

class Voxelizer {
public:
    // Voxelizer(voxDim, brickDim, voxChunkDim=uvec3(256)) (Voxelizer.h 122-124),
    // with the reference's default chunk.  A chunk of at least the voxel
    // space (N + 4*((N/B)>>1)) processes it in one pass.
    Voxelizer(const uvec3& voxDim, const uvec3& brickDim,
              const uvec3& voxChunkDim = uvec3{256u, 256u, 256u})
        : _h(vox_create(voxDim.data(), brickDim.data(), voxChunkDim.data())) {
        // (voxDim / brickDim) >> 1 per axis (Voxelizer.cu 143-144).
        for (int a = 0; a < 3; ++a)
            _extra[a] = brickDim[a] ? (voxDim[a] / brickDim[a]) >> 1 : 0u;
    }
    ~Voxelizer() { vox_destroy(_h); }
    Voxelizer(const Voxelizer&) = delete;
    Voxelizer& operator=(const Voxelizer&) = delete;

    // addTriangleGroup (Voxelizer.h 128-162): the vectors are read by
    // allocateTriangleMemory and must stay alive until then.  V3 / V2 are any
    // tightly packed 3- / 2-float types (std::array, glm::vec3 / glm::vec2).
    template <class V3, class V2, class B3>
    void addTriangleGroup(const std::vector<V3>* pVerts, const std::vector<V3>* pVertNormals,
                          const std::vector<V2>* pUVs, const B3& minXYZ, const B3& maxXYZ,
                          bool isTerrain, unsigned char* pImageData = nullptr, int imageWidth = 0,
                          int imageHeight = 0, int texAddressMode0 = AddressModeClamp,
                          int texAddressMode1 = AddressModeClamp) {
        static_assert(sizeof(V3) == 3 * sizeof(float), "vertex type must be 3 packed floats");
        static_assert(sizeof(V2) == 2 * sizeof(float), "uv type must be 2 packed floats");
        vox_group g{};
        g.verts = pVerts->empty() ? nullptr : reinterpret_cast<const float*>(pVerts->data());
        g.vnormals = pVertNormals && !pVertNormals->empty()
                         ? reinterpret_cast<const float*>(pVertNormals->data()) : nullptr;
        g.uvs = pUVs && !pUVs->empty() ? reinterpret_cast<const float*>(pUVs->data()) : nullptr;
        g.ntri = pVerts->size() / 3;
        for (int a = 0; a < 3; ++a) {
            g.bbox_min[a] = minXYZ[a];
            g.bbox_max[a] = maxXYZ[a];
        }
        g.is_terrain = isTerrain ? 1 : 0;
        g.rgba8 = pImageData;
        g.width = imageWidth;
        g.height = imageHeight;
        g.addr0 = texAddressMode0;
        g.addr1 = texAddressMode1;
        vox_add_group(_h, &g);
    }

    template <class B3>
    void setVoxelizationParams(const B3& p, const B3& deltaP) {
        const float pf[3] = {p[0], p[1], p[2]};
        const float df[3] = {deltaP[0], deltaP[1], deltaP[2]};
        vox_set_params(_h, pf, df);
    }
    bool initCuda() { return vox_init(_h) == 1; }
    bool resetCuda() { return vox_reset(_h) == 1; }
    bool allocateTriangleMemory() { return vox_allocate_triangle_memory(_h) == 1; }
    void deallocateTriangleMemory() { vox_deallocate_triangle_memory(_h); }
    bool allocateVoxelMemory() { return vox_allocate_voxel_memory(_h) == 1; }
    void deallocateVoxelMemory() { vox_deallocate_voxel_memory(_h); }
    bool computeEdgesFaceNormalsAndBounds() {
        return vox_compute_edges_face_normals_and_bounds(_h) == 1;
    }
    // 1 ok, 0 error, -1 no triangle group overlapped the grid.
    int generateVoxelsAndOctTree(const std::string& outputDir, bool outputBinary,
                                 bool outputCompressed) {
        return vox_generate(_h, outputDir.empty() ? nullptr : outputDir.c_str(),
                            outputBinary ? 1 : 0, outputCompressed ? 1 : 0);
    }
    const std::string& getErrorMessage() {
        _err = vox_get_error(_h);
        return _err;
    }
    // Not in the reference: generate (no files) on a worker thread; wait()
    // returns its 1 / 0 / -1, or 2 while it runs when block is false.
    bool generateAsync(bool outputCompressed = false) {
        return vox_generate_async(_h, outputCompressed ? 1 : 0) != 0;
    }
    int wait(bool block = true) { return vox_wait(_h, block ? 1 : 0); }
    bool releaseWorkMemory() { return vox_release_work_memory(_h) != 0; }
    // Border voxels added on each side of the grid, per axis (Voxelizer.h 179);
    // the tiler sizes its triangle overlap from it (main.cu 4171-4175).
    const uvec3& getExtraVoxChunk() const { return _extra; }

    // Not in the reference: the device-resident result of the last generate.
    bool getResult(vox_result& out) { return vox_get_result(_h, &out) == 1; }
    bool getStats(vox_stats& out) { return vox_get_stats(_h, &out) == 1; }
    vox_ctx* handle() { return _h; }

private:
    vox_ctx* _h;
    uvec3 _extra{};
    std::string _err;
};

}  // namespace voxmi

#endif
