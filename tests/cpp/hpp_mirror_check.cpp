// Compiles main.cu-style host code against include/voxmi_voxelizer.hpp and
// exercises the calls that need no GPU: construction, getExtraVoxChunk,
// parameter and group registration, and the error reporting of an invalid
// chunk.  Prints one line per check; exit status 0 when all hold.
#include <cstdio>
#include <cstring>
#include <vector>

#include "voxmi_voxelizer.hpp"

// A stand-in with glm::vec3's layout and operator[], as main.cu passes.
struct v3 { float x, y, z; float operator[](int i) const { return (&x)[i]; } };
struct v2 { float x, y; };

int main() {
    int bad = 0;
    {
        voxmi::Voxelizer vz(voxmi::uvec3{128, 128, 128}, voxmi::uvec3{8, 8, 8});
        const voxmi::uvec3& e = vz.getExtraVoxChunk();
        std::printf("extra %u %u %u\n", e[0], e[1], e[2]);
        bad |= !(e[0] == 8 && e[1] == 8 && e[2] == 8);
        std::vector<v3> verts = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
        std::vector<v2> uvs = {{0, 0}, {1, 0}, {0, 1}};
        vz.setVoxelizationParams(v3{0, 0, 0}, v3{1.f / 128, 1.f / 128, 1.f / 128});
        vz.addTriangleGroup(&verts, (std::vector<v3>*)nullptr, &uvs, v3{0, 0, 0}, v3{1, 1, 0},
                            false);
        std::printf("error '%s'\n", vz.getErrorMessage().c_str());
        bad |= !vz.getErrorMessage().empty();
    }
    {
        // the reference's default chunk (256) with main.cu's 1024^3 / 16^3 grid
        voxmi::Voxelizer vz(voxmi::uvec3{1024, 1024, 1024}, voxmi::uvec3{16, 16, 16});
        std::printf("default chunk error '%s'\n", vz.getErrorMessage().c_str());
        bad |= !vz.getErrorMessage().empty();
    }
    {
        // a chunk that is not a multiple of 2^(levels-1) = 64
        voxmi::Voxelizer vz(voxmi::uvec3{1024, 1024, 1024}, voxmi::uvec3{16, 16, 16},
                            voxmi::uvec3{100, 100, 100});
        std::printf("chunk error '%s'\n", vz.getErrorMessage().c_str());
        bad |= vz.getErrorMessage().find("chunk") == std::string::npos;
        bad |= vz.initCuda();      // must refuse after a construction error
        bad |= vz.generateAsync();  // and start no worker
        std::printf("wait without a worker %d\n", vz.wait(false));
        bad |= vz.wait(false) != 0;
    }
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad;
}
