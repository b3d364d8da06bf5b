// gvformat.cpp -- GigaVoxels file writers: tree.gvx and voxels_<depth>.gvb.
//
// tree.gvx follows Voxelizer::writeOctTree / WriteOctTreeNode
// (Voxelizer.cu 2816-3061): depth-first XML from the root, children in z, y,
// x loops, LEAF-CONST nodes without children, default iostream formatting.
// voxels_<depth>.gvb follows VoxelBrickWriter::startBricksFile and
// WriteBinaryBrick / WriteUByteBinaryBrick (VoxelBrickWriter.cu 71-124,
// 725-818, 509-588): int32 compressed flag, then per brick in index order
// u64 dims, u32 borders, int32 0, u32 size, RGBA8, int32 0, u32 size, RGB8;
// compressed (WriteCompressedBinaryBrick, 590-723): the same with flags 1
// and BC3 colour / BC1 normal blocks.  Text bricks, voxels_<depth>.gvt
// (StartTextBricksFile / WriteTextBrick / EndTextBricksFile, 45-68, 207-325,
// 96-103): XML-like rows of the brick's (B+2)^3 window of its mip level,
// colours as their four bytes written as characters (the reference streams
// unsigned chars), float normals in default iostream formatting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/voxmi.h"

namespace {

struct HostTree {
    int L = 0;
    uint32_t B[3];
    std::vector<std::vector<uint8_t>> types;
    std::vector<std::vector<uint32_t>> cc, bidx;
};

void write_node(std::ofstream& f, const HostTree& t, int d, uint32_t x, uint32_t y, uint32_t z,
                uint64_t xo, uint64_t yo, uint64_t zo) {
    const uint32_t Rd = 1u << d;
    const uint64_t i = ((uint64_t)z * Rd + y) * Rd + x;
    const uint8_t type = t.types[d][i];
    for (int k = 0; k <= d; ++k) f << "    ";
    if (type == 1) {
        // a NON-CONST node whose brick was never completed (several chunks,
        // its border in a chunk that is not processed) has no index; the
        // reference's getBrickIndex answers 0 for it (VoxelBrickWriter.h 143-149)
        const uint32_t bi = t.bidx[d][i] == 0xffffffffu ? 0u : t.bidx[d][i];
        f << "<Node " << "MipMapX=\"" << xo << "\" " << "MipMapY=\"" << yo << "\" "
          << "MipMapZ=\"" << zo << "\" " << "Type=\"NON-CONST\" " << "Brick=\"" << bi
          << "\" " << "Depth=\"" << d << "\"";
    } else {
        const uint32_t c = t.cc[d][i];
        f << "<Node " << "MipMapX=\"" << xo << "\" " << "MipMapY=\"" << yo << "\" "
          << "MipMapZ=\"" << zo << "\" " << "Type=\"" << (type == 0 ? "CONST" : "LEAF-CONST")
          << "\" " << "ColorR=\"" << static_cast<float>(c & 255u) / 255.0 << "\" "
          << "ColorG=\"" << static_cast<float>((c >> 8) & 255u) / 255.0 << "\" "
          << "ColorB=\"" << static_cast<float>((c >> 16) & 255u) / 255.0 << "\" "
          << "ColorA=\"" << static_cast<float>(c >> 24) / 255.0 << "\" " << "Depth=\"" << d
          << "\"";
    }
    if (d + 1 < t.L && type != 2) {
        f << ">" << std::endl;
        for (uint32_t cz = 0; cz < 2; ++cz)
            for (uint32_t cy = 0; cy < 2; ++cy)
                for (uint32_t cx = 0; cx < 2; ++cx)
                    write_node(f, t, d + 1, 2 * x + cx, 2 * y + cy, 2 * z + cz,
                               (xo << 1) + cx * t.B[0], (yo << 1) + cy * t.B[1],
                               (zo << 1) + cz * t.B[2]);
        for (int k = 0; k <= d; ++k) f << "    ";
        f << "</Node>" << std::endl;
    } else {
        f << " />" << std::endl;
    }
}

// One brick of a text bricks file: the brick's window of (B+2)^3 voxels,
// x0..: its first voxel in the level (dense colours / normals of extent S),
// written at (wx, wy, wz) with the row coordinates counted from there.
void write_text_brick(std::ofstream& f, const uint32_t* col, const float* nrm, const uint32_t S[3],
                      const uint32_t x0[3], const uint64_t w0[3], const uint32_t e[3]) {
    f << " <Brick " << "X=\"" << w0[0] << "\" " << "Y=\"" << w0[1] << "\" "
      << "Z=\"" << w0[2] << "\" " << "DimX=\"" << e[0] << "\" " << "DimY=\"" << e[1] << "\" "
      << "DimZ=\"" << e[2] << "\" " << "BorderX=\"" << 1 << "\" " << "BorderY=\"" << 1 << "\" "
      << "BorderZ=\"" << 1 << "\" " << "/>" << std::endl;
    f << "  <BrickColors>" << std::endl;
    for (uint32_t z = 0; z < e[2]; ++z)
        for (uint32_t y = 0; y < e[1]; ++y) {
            f << "    <Row Y=\"" << w0[1] + y << "\" Z=\"" << w0[2] + z << "\">" << std::endl;
            const uint64_t row = ((uint64_t)(x0[2] + z) * S[1] + (x0[1] + y)) * S[0] + x0[0];
            for (uint32_t x = 0; x < e[0]; ++x) {
                const uint32_t c = col[row + x];
                f << "(" << (char)(c & 255u) << "," << (char)((c >> 8) & 255u) << ","
                  << (char)((c >> 16) & 255u) << "," << (char)(c >> 24) << ") ";
            }
            f << std::endl << "    </Row>" << std::endl;
        }
    f << "  </BrickColors>" << std::endl;
    f << "  <BrickGradients>" << std::endl;
    for (uint32_t z = 0; z < e[2]; ++z)
        for (uint32_t y = 0; y < e[1]; ++y) {
            f << "    <Row YZ=\"" << w0[1] + y << " " << w0[2] + z << "\">" << std::endl;
            const uint64_t row = ((uint64_t)(x0[2] + z) * S[1] + (x0[1] + y)) * S[0] + x0[0];
            for (uint32_t x = 0; x < e[0]; ++x) {
                const float* n = nrm + 3 * (row + x);
                f << "(" << n[0] << "," << n[1] << "," << n[2] << ") ";
            }
            f << std::endl << "    </Row>" << std::endl;
        }
    f << "  </BrickGradients>" << std::endl;
    f << " </Brick>" << std::endl;
}

}  // namespace

// Declared in voxmi.h; defined here with access to the result through the
// public C ABI only.
extern "C" int vox_write_files_impl(vox_ctx* h, const char* out_dir, const float p[3],
                                    const float delta[3], const uint32_t N[3], int binary,
                                    const uint32_t extra[3], const uint32_t chunk[3], int multi);

extern "C" int vox_write_files_impl(vox_ctx* h, const char* out_dir, const float p[3],
                                    const float delta[3], const uint32_t N[3], int binary,
                                    const uint32_t extra[3], const uint32_t chunk[3], int multi) {
    vox_result r;
    if (!vox_get_result(h, &r)) return 0;
    vox_compressed cz;
    const bool compressed = vox_get_compressed(h, &cz) == 1;
    mkdir(out_dir, 0755);
    HostTree t;
    t.L = r.levels;
    for (int a = 0; a < 3; ++a) t.B[a] = r.brick_dim[a];
    t.types.resize(t.L);
    t.cc.resize(t.L);
    t.bidx.resize(t.L);
    const uint64_t bvox = (uint64_t)(t.B[0] + 2) * (t.B[1] + 2) * (t.B[2] + 2);
    // brick border of each depth's level: extra at level 0, then halved, at
    // least 1 (Voxelizer.cu 1270-1272)
    std::vector<std::array<uint32_t, 3>> bord(t.L);
    for (int k = 0; k < t.L; ++k)
        for (int a = 0; a < 3; ++a)
            bord[k][a] = k == 0 ? extra[a] : std::max(bord[k - 1][a] >> 1, 1u);
    for (int d = 0; d < t.L; ++d) {
        const uint64_t n = r.nodes[d];
        t.types[d].resize(n);
        t.cc[d].resize(n);
        t.bidx[d].resize(n);
        if (!vox_copy_to_host(h, t.types[d].data(), r.types[d], n) ||
            !vox_copy_to_host(h, t.cc[d].data(), r.const_colors[d], n * 4) ||
            !vox_copy_to_host(h, t.bidx[d].data(), r.brick_index[d], n * 4))
            return 0;
        if (!binary) {
            // voxels_<depth>.gvt from the dense level (L-1-d) on the host
            std::ofstream f(std::string(out_dir) + "/voxels_" + std::to_string(d) + ".gvt",
                            std::fstream::out);
            if (!f.is_open()) return 0;
            f << "<Bricks>" << std::endl;
            const uint64_t nb = r.bricks[d];
            if (nb > 0) {
                const int k = t.L - 1 - d;
                uint32_t S[3];
                if (!vox_copy_level(h, k, nullptr, nullptr, S)) return 0;
                const uint64_t nv = (uint64_t)S[0] * S[1] * S[2];
                std::vector<uint32_t> col(nv);
                std::vector<float> nrm(3 * nv);
                std::vector<uint32_t> bn(nb);
                if (!vox_copy_level(h, k, col.data(), nrm.data(), S) ||
                    !vox_copy_to_host(h, bn.data(), r.brick_node[d], nb * 4))
                    return 0;
                const uint32_t Rd = 1u << d;
                const uint32_t e[3] = {t.B[0] + 2, t.B[1] + 2, t.B[2] + 2};
                for (uint64_t b = 0; b < nb; ++b) {
                    const uint32_t nx = bn[b] % Rd, ny = (bn[b] / Rd) % Rd, nz = bn[b] / (Rd * Rd);
                    const uint32_t m[3] = {nx * t.B[0] + bord[k][0], ny * t.B[1] + bord[k][1],
                                           nz * t.B[2] + bord[k][2]};
                    const uint32_t x0[3] = {m[0] - 1, m[1] - 1, m[2] - 1};
                    // a brick reaching the chunk's far side was stored as a
                    // partial brick and is written from its own buffer, whose
                    // window starts at 0 (VoxelBrickWriter.cu 1179-1190);
                    // the others from their chunk's level at their offset in
                    // it (Voxelizer.cu 1128-1130, 1182-1253: the node's place
                    // less the chunk's origin at this level).  With several
                    // chunks a brick is partial exactly when its window
                    // [x0, x0 + B + 1] leaves one chunk of C >> k voxels; the
                    // window's voxels are the level's at those places either
                    // way (the chunks' levels laid side by side).
                    bool partial = false;
                    uint32_t origin[3] = {0u, 0u, 0u};
                    for (int a = 0; a < 3; ++a) {
                        if (multi) {
                            const uint32_t Ck = chunk[a] >> k;
                            const uint32_t c0 = x0[a] / Ck, c1 = (x0[a] + t.B[a] + 1u) / Ck;
                            partial = partial || c0 != c1;
                            origin[a] = c0 * Ck;
                        } else {
                            partial = partial || m[a] + t.B[a] >= S[a];
                        }
                    }
                    const uint64_t w0[3] = {partial ? 0u : x0[0] - origin[0],
                                            partial ? 0u : x0[1] - origin[1],
                                            partial ? 0u : x0[2] - origin[2]};
                    write_text_brick(f, col.data(), nrm.data(), S, x0, w0, e);
                }
            }
            f << "</Bricks>" << std::endl;
            if (f.fail()) return 0;
            continue;
        }
        // voxels_<depth>.gvb
        std::string fn = std::string(out_dir) + "/voxels_" + std::to_string(d) + ".gvb";
        std::ofstream f(fn, std::ios_base::out | std::ios_base::binary);
        if (!f.is_open()) return 0;
        const int32_t cflag = compressed ? 1 : 0;
        f.write((const char*)&cflag, 4);
        const uint64_t nb = r.bricks[d];
        const uint64_t chunk = 4096;
        if (compressed) {
            // WriteCompressedBinaryBrick: BC3 colours, BC1 normals per brick
            const uint64_t cb = cz.blocks_per_brick * 16, nbb = cz.blocks_per_brick * 8;
            std::vector<uint8_t> cc, cn;
            for (uint64_t b0 = 0; b0 < nb; b0 += chunk) {
                const uint64_t m = std::min<uint64_t>(chunk, nb - b0);
                cc.resize(m * cb);
                cn.resize(m * nbb);
                if (!vox_copy_to_host(h, cc.data(), cz.colors[d] + b0 * cb, m * cb) ||
                    !vox_copy_to_host(h, cn.data(), cz.normals[d] + b0 * nbb, m * nbb))
                    return 0;
                for (uint64_t j = 0; j < m; ++j) {
                    const uint64_t dims[3] = {t.B[0] + 2, t.B[1] + 2, t.B[2] + 2};
                    const uint32_t border[3] = {1, 1, 1};
                    f.write((const char*)dims, sizeof(dims));
                    f.write((const char*)border, sizeof(border));
                    uint32_t sz = (uint32_t)cb;
                    f.write((const char*)&cflag, 4);
                    f.write((const char*)&sz, 4);
                    f.write((const char*)&cc[j * cb], cb);
                    sz = (uint32_t)nbb;
                    f.write((const char*)&cflag, 4);
                    f.write((const char*)&sz, 4);
                    f.write((const char*)&cn[j * nbb], nbb);
                }
            }
            if (f.fail()) return 0;
            continue;
        }
        std::vector<uint32_t> col, nrm;
        std::vector<uint8_t> rgb(bvox * 3);
        for (uint64_t b0 = 0; b0 < nb; b0 += chunk) {
            const uint64_t m = std::min<uint64_t>(chunk, nb - b0);
            col.resize(m * bvox);
            nrm.resize(m * bvox);
            if (!vox_copy_to_host(h, col.data(), r.brick_colors[d] + b0 * bvox, m * bvox * 4) ||
                !vox_copy_to_host(h, nrm.data(), r.brick_normals[d] + b0 * bvox, m * bvox * 4))
                return 0;
            for (uint64_t j = 0; j < m; ++j) {
                const uint64_t dims[3] = {t.B[0] + 2, t.B[1] + 2, t.B[2] + 2};
                const uint32_t border[3] = {1, 1, 1};
                f.write((const char*)dims, sizeof(dims));
                f.write((const char*)border, sizeof(border));
                int32_t zero = 0;
                uint32_t sz = (uint32_t)(bvox * 4);
                f.write((const char*)&zero, 4);
                f.write((const char*)&sz, 4);
                f.write((const char*)&col[j * bvox], bvox * 4);
                for (uint64_t v = 0; v < bvox; ++v) {
                    const uint32_t q = nrm[j * bvox + v];
                    rgb[3 * v] = (uint8_t)(q & 255u);
                    rgb[3 * v + 1] = (uint8_t)((q >> 8) & 255u);
                    rgb[3 * v + 2] = (uint8_t)((q >> 16) & 255u);
                }
                sz = (uint32_t)(bvox * 3);
                f.write((const char*)&zero, 4);
                f.write((const char*)&sz, 4);
                f.write((const char*)rgb.data(), bvox * 3);
            }
        }
        if (f.fail()) return 0;
    }
    std::ofstream tf(std::string(out_dir) + "/tree.gvx", std::fstream::out);
    if (!tf.is_open()) return 0;
    tf << "<GigaVoxelsOctTree " << "MaxDepth=\"" << t.L << "\" " << "X=\"" << p[0] << "\" "
       << "Y=\"" << p[1] << "\" " << "Z=\"" << p[2] << "\" " << "DeltaX=\"" << delta[0] << "\" "
       << "DeltaY=\"" << delta[1] << "\" " << "DeltaZ=\"" << delta[2] << "\" "
       << "VolumeXSize=\"" << N[0] << "\" " << "VolumeYSize=\"" << N[1] << "\" "
       << "VolumeZSize=\"" << N[2] << "\" " << "BrickXSize=\"" << t.B[0] << "\" "
       << "BrickYSize=\"" << t.B[1] << "\" " << "BrickZSize=\"" << t.B[2] << "\" "
       << "Binary=\"" << (binary ? "YES" : "NO") << "\" " << "Compressed=\""
       << (compressed ? "YES" : "NO") << "\" >"
       << std::endl;
    write_node(tf, t, 0, 0, 0, 0, 0, 0, 0);
    tf << "</GigaVoxelsOctTree>" << std::endl;
    return tf.fail() ? 0 : 1;
}
