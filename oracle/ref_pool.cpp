// ref_pool.cpp -- CPU restatement of the GigaVoxels runtime's file loading
// and node-pool / brick-pool construction.  TEST INFRASTRUCTURE ONLY: the
// oracle for the HIP node-pool builder.  The reference sections followed are
// listed in ref_pool.h.
#include "ref_pool.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

std::string g_err;

// ---- a small XML event reader (elements and attributes only) -------------
struct Ev {
    bool start;  // start element (a self-closing tag yields start + end)
    std::string name;
    std::map<std::string, std::string> attrs;
};

bool parse_xml(const std::string& s, std::vector<Ev>& evs) {
    size_t i = 0;
    while ((i = s.find('<', i)) != std::string::npos) {
        size_t j = s.find('>', i);
        if (j == std::string::npos) return false;
        std::string tag = s.substr(i + 1, j - i - 1);
        i = j + 1;
        if (tag.empty() || tag[0] == '?' || tag[0] == '!') continue;
        if (tag[0] == '/') {
            size_t e = tag.find_first_of(" \t\r\n", 1);
            evs.push_back({false, tag.substr(1, e == std::string::npos ? e : e - 1), {}});
            continue;
        }
        bool selfClose = tag.back() == '/';
        if (selfClose) tag.pop_back();
        Ev ev;
        ev.start = true;
        size_t k = tag.find_first_of(" \t\r\n");
        ev.name = tag.substr(0, k);
        while (k != std::string::npos && k < tag.size()) {
            size_t eq = tag.find('=', k);
            if (eq == std::string::npos) break;
            size_t ks = tag.find_first_not_of(" \t\r\n", k);
            std::string key = tag.substr(ks, eq - ks);
            size_t q0 = tag.find('"', eq);
            size_t q1 = q0 == std::string::npos ? q0 : tag.find('"', q0 + 1);
            if (q1 == std::string::npos) return false;
            ev.attrs[key] = tag.substr(q0 + 1, q1 - q0 - 1);
            k = q1 + 1;
        }
        evs.push_back(ev);
        if (selfClose) evs.push_back({false, ev.name, {}});
    }
    return true;
}

int to_int(const Ev& e, const char* k) {
    auto it = e.attrs.find(k);
    return it == e.attrs.end() ? 0 : std::atoi(it->second.c_str());
}
// QString::toFloat: parse as double, then narrow to float.
float to_float(const Ev& e, const char* k) {
    auto it = e.attrs.find(k);
    return it == e.attrs.end() ? 0.f : (float)std::strtod(it->second.c_str(), nullptr);
}
std::string to_str(const Ev& e, const char* k) {
    auto it = e.attrs.find(k);
    return it == e.attrs.end() ? std::string() : it->second;
}

// ---- the runtime's node (GigaVoxelsOctTree::Node) ---------------------------
struct Node {
    uint32_t w[2] = {0, 0};       // GpuDataStruct: two 32-bit words
    size_t childNodeBlockIndex = 0;  // node index of the first child (0 = none)
    int brickTex[3] = {0, 0, 0};
};

}  // namespace

struct ref_pool {
    std::vector<Node> nodes;       // OctTreeNodePool::m_nodePool
    std::vector<std::vector<size_t>> nonConst;  // per depth, brick index -> node
    uint32_t ndim[3] = {0, 0, 0};
    std::vector<uint32_t> tex;     // node texture
    uint32_t B[3] = {0, 0, 0}, slots[3] = {0, 0, 0}, bdim[3] = {0, 0, 0};
    std::vector<uint32_t> bcol, bnrm;  // brick pool textures
    int depths = 0;
    std::vector<std::vector<uint32_t>> fcol, fnrm;  // bricks read from .gvb
    std::vector<std::vector<uint8_t>> fccomp, fncomp;  // compressed payloads
    std::vector<uint64_t> fcount;
    int compressed = 0;
    int32_t flags = 0;
};

namespace {

// Node::setNodeTypeFlag / setChildNodesPtr / setBrickPtr / setBrickIsOnGpuFlag
// (GigaVoxelsOctTree.cpp 1317-1378): bits are OR-ed in.
void set_const(Node& n, bool c) {
    if (c) n.w[0] |= 0x40000000u;
    else n.w[0] &= ~0x40000000u;
}
void set_child_ptr(Node& n, uint32_t x, uint32_t y, uint32_t z) {
    n.w[0] |= x << 20;
    n.w[0] |= y << 10;
    n.w[0] |= z;
}
void set_brick_ptr(Node& n, uint32_t x, uint32_t y, uint32_t z) {
    n.w[1] |= x << 20;
    n.w[1] |= y << 10;
    n.w[1] |= z;
}

// OctTreeNodePool::allocateChildNodeBlock (GigaVoxelsOctTreeNodePool.cpp 49-74).
size_t allocate_child_block(ref_pool* p, long parent) {
    size_t idx = p->nodes.size();
    if (parent >= 0) p->nodes[parent].childNodeBlockIndex = idx;
    p->nodes.resize(idx + 8);
    return idx;
}

// LoadOctTreeNode (GigaVoxelsReader.cpp 277-372).  pos indexes the node's
// start event; on return it indexes the last event consumed.
bool load_node(ref_pool* p, const std::vector<Ev>& evs, size_t& pos, size_t cur, int maxDepth,
               int depth) {
    const Ev& e = evs[pos];
    if (to_int(e, "Depth") != depth) {
        g_err = "Depth attribute does not match the tree depth";
        return false;
    }
    const std::string type = to_str(e, "Type");
    if (type == "CONST" || type == "LEAF-CONST") {
        set_const(p->nodes[cur], true);
        // setConstantValue computes the colour and never stores it.
        const uint8_t r = (uint8_t)(to_float(e, "ColorR") * 255.0f);
        const uint8_t g = (uint8_t)(to_float(e, "ColorG") * 255.0f);
        const uint8_t b = (uint8_t)(to_float(e, "ColorB") * 255.0f);
        const uint8_t a = (uint8_t)(to_float(e, "ColorA") * 255.0f);
        if (p->flags & REF_POOL_STORE_CONST)
            p->nodes[cur].w[1] = r | (g << 8) | (b << 16) | ((uint32_t)a << 24);
    } else {
        set_const(p->nodes[cur], false);
        const int bi = to_int(e, "Brick");
        auto& nc = p->nonConst.at(depth);
        if (bi >= (int)nc.size()) nc.resize(bi + 1, (size_t)-1);
        nc[bi] = cur;
    }
    if (depth != maxDepth && type != "LEAF-CONST") {
        const size_t childStart = allocate_child_block(p, (long)cur);
        for (int i = 0; i < 8; ++i) {
            // next start element
            ++pos;
            while (pos < evs.size() && !evs[pos].start) ++pos;
            if (pos >= evs.size()) {
                g_err = "parsing next element failed";
                return false;
            }
            if (evs[pos].name != "Node") {
                g_err = "invalid node encountered";
                return false;
            }
            if (!load_node(p, evs, pos, childStart + i, maxDepth, depth + 1)) return false;
        }
    } else {
        // skipCurrentElement: to the matching end element
        int level = 0;
        for (++pos; pos < evs.size(); ++pos) {
            if (evs[pos].start) ++level;
            else if (level-- == 0) break;
        }
    }
    return true;
}

// LoadBinaryBricks (GigaVoxelsReader.cpp 106-221).
bool load_bricks(ref_pool* p, const std::string& path, int depth) {
    std::ifstream f(path, std::ios_base::in | std::ios_base::binary);
    if (!f.is_open()) {
        g_err = "failed to open " + path;
        return false;
    }
    int32_t compressed = 0;
    f.read((char*)&compressed, 4);
    if (compressed) p->compressed = 1;
    auto& nc = p->nonConst[depth];
    uint64_t count = 0;
    for (size_t i = 0; i < nc.size(); ++i) {
        if (nc[i] == (size_t)-1) break;
        uint64_t dims[3];
        uint32_t border[3];
        f.read((char*)dims, sizeof(dims));
        f.read((char*)border, sizeof(border));
        int32_t cc = 0;
        uint32_t size = 0;
        f.read((char*)&cc, 4);
        if (cc != compressed) {
            g_err = "compression mismatch (colors)";
            return false;
        }
        f.read((char*)&size, 4);
        if (size == 0) {
            g_err = "size of color data is zero";
            return false;
        }
        std::vector<uint8_t> col(size);
        f.read((char*)col.data(), size);
        if ((uint32_t)f.gcount() != size) {
            g_err = "short color read";
            return false;
        }
        f.read((char*)&cc, 4);
        if (cc != compressed) {
            g_err = "compression mismatch (gradients)";
            return false;
        }
        f.read((char*)&size, 4);
        if (size == 0) {
            g_err = "size of gradient data is zero";
            return false;
        }
        std::vector<uint8_t> nrm(size);
        f.read((char*)nrm.data(), size);
        if ((uint32_t)f.gcount() != size) {
            g_err = "short gradient read";
            return false;
        }
        const uint64_t nv = dims[0] * dims[1] * dims[2];
        if (compressed) {
            // compressed payloads are kept as bytes (the reader only sizes them)
            p->fccomp[depth].insert(p->fccomp[depth].end(), col.begin(), col.end());
            p->fncomp[depth].insert(p->fncomp[depth].end(), nrm.begin(), nrm.end());
            Node& n = p->nodes[nc[i]];
            n.brickTex[0] = (int)border[0];
            n.brickTex[1] = (int)border[1];
            n.brickTex[2] = (int)border[2];
            ++count;
            continue;
        }
        if (col.size() != nv * 4 || nrm.size() != nv * 3) {
            g_err = "brick size does not match its dimensions";
            return false;
        }
        for (uint64_t v = 0; v < nv; ++v) {
            uint32_t c;
            std::memcpy(&c, &col[4 * v], 4);
            p->fcol[depth].push_back(c);
            p->fnrm[depth].push_back(nrm[3 * v] | (nrm[3 * v + 1] << 8) |
                                     ((uint32_t)nrm[3 * v + 2] << 16));
        }
        Node& n = p->nodes[nc[i]];
        n.brickTex[0] = (int)border[0];
        n.brickTex[1] = (int)border[1];
        n.brickTex[2] = (int)border[2];
        ++count;
    }
    p->fcount[depth] = count;
    if (f.peek() != std::char_traits<char>::eof()) {
        g_err = "trailing bytes after the last brick in " + path;
        return false;
    }
    return true;
}

// BrickPool::initBricks(root, 1) (GigaVoxelsBrickPool.cpp 270-345): only the
// root's brick is uploaded, into slot 0, when the root is NON-CONST.
void init_root_brick(ref_pool* p) {
    Node& root = p->nodes[0];
    if (root.w[0] & 0x40000000u) return;
    if (p->slots[0] * p->slots[1] * p->slots[2] == 0 || p->fcount.empty() || p->fcount[0] == 0 ||
        p->compressed)
        return;
    const uint32_t e[3] = {p->B[0] + 2, p->B[1] + 2, p->B[2] + 2};
    for (uint32_t z = 0; z < e[2]; ++z)
        for (uint32_t y = 0; y < e[1]; ++y)
            for (uint32_t x = 0; x < e[0]; ++x) {
                const uint64_t src = ((uint64_t)z * e[1] + y) * e[0] + x;
                const uint64_t dst = ((uint64_t)z * p->bdim[1] + y) * p->bdim[0] + x;
                p->bcol[dst] = p->fcol[0][src];
                p->bnrm[dst] = p->fnrm[0][src];
            }
    // uploadBrick: pointer = slot origin + brick border (from the .gvb header)
    set_brick_ptr(root, (uint32_t)root.brickTex[0], (uint32_t)root.brickTex[1],
                  (uint32_t)root.brickTex[2]);
    root.w[0] |= 0x80000000u;
}

// OctTreeNodePool::create3DTexture (GigaVoxelsOctTreeNodePool.cpp 90-223).
bool create_texture(ref_pool* p) {
    const size_t numNodes = p->nodes.size();
    size_t cubeRoot = (size_t)std::pow((double)numNodes, 1.0 / 3.0);
    if (cubeRoot % 2 != 0) {
        --cubeRoot;
        if (cubeRoot == 0) cubeRoot = 2;
    }
    size_t dx = cubeRoot, dy = cubeRoot, dz = cubeRoot;
    while (dx * dy * dz < numNodes) dx += 2;
    p->ndim[0] = (uint32_t)dx;
    p->ndim[1] = (uint32_t)dy;
    p->ndim[2] = (uint32_t)dz;
    p->tex.assign(dx * dy * dz * 2, 0u);
    const size_t bx = dx >> 1, by = dy >> 1;
    const size_t nblocks = numNodes / 8;
    for (size_t b = 0; b < nblocks; ++b) {
        const size_t z = b / by / bx;
        const size_t y = (b - z * by * bx) / bx;
        const size_t x = b - z * by * bx - y * bx;
        for (size_t i = 0; i < 8; ++i) {
            Node& n = p->nodes[b * 8 + i];
            size_t cb = n.childNodeBlockIndex;
            if (cb != 0) {
                cb >>= 3;
                const size_t cz = cb / by / bx;
                const size_t cy = (cb - cz * by * bx) / bx;
                const size_t cx = cb - cz * by * bx - cy * bx;
                if (2 * cx > 1023 || 2 * cy > 1023 || 2 * cz > 1023) {
                    g_err = "child pointer exceeds 10 bits";
                    return false;
                }
                set_child_ptr(n, (uint32_t)(2 * cx), (uint32_t)(2 * cy), (uint32_t)(2 * cz));
            }
            // texel of node i of the block: x fastest
            const size_t tx = 2 * x + (i & 1), ty = 2 * y + ((i >> 1) & 1), tz = 2 * z + (i >> 2);
            const size_t t = (tz * dy + ty) * dx + tx;
            p->tex[2 * t] = n.w[0];
            p->tex[2 * t + 1] = n.w[1];
        }
    }
    return true;
}

}  // namespace

extern "C" const char* ref_pool_error(void) { return g_err.c_str(); }

extern "C" ref_pool* ref_pool_load(const char* dir, const uint32_t slots[3], int32_t flags) {
    g_err.clear();
    const std::string d(dir);
    std::ifstream f(d + "/tree.gvx");
    if (!f.is_open()) {
        g_err = "cannot open tree.gvx";
        return nullptr;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::vector<Ev> evs;
    if (!parse_xml(ss.str(), evs)) {
        g_err = "malformed xml";
        return nullptr;
    }
    ref_pool* p = new ref_pool;
    p->flags = flags;
    bool ok = false;
    for (size_t pos = 0; pos < evs.size(); ++pos) {
        if (!evs[pos].start) continue;
        if (evs[pos].name == "GigaVoxelsOctTree") {
            p->depths = to_int(evs[pos], "MaxDepth");
            p->B[0] = (uint32_t)to_int(evs[pos], "BrickXSize");
            p->B[1] = (uint32_t)to_int(evs[pos], "BrickYSize");
            p->B[2] = (uint32_t)to_int(evs[pos], "BrickZSize");
            if (to_str(evs[pos], "Binary") != "YES") {
                g_err = "binary bricks are the only supported format";
                delete p;
                return nullptr;
            }
        } else if (evs[pos].name == "Node" && p->depths > 0) {
            allocate_child_block(p, -1);  // root block, root = node 0
            p->nonConst.assign(p->depths, {});
            if (!load_node(p, evs, pos, 0, p->depths - 1, 0)) {
                delete p;
                return nullptr;
            }
            ok = true;
            break;
        }
    }
    if (!ok) {
        g_err = g_err.empty() ? "no root node" : g_err;
        delete p;
        return nullptr;
    }
    p->fcol.assign(p->depths, {});
    p->fnrm.assign(p->depths, {});
    p->fccomp.assign(p->depths, {});
    p->fncomp.assign(p->depths, {});
    p->fcount.assign(p->depths, 0);
    for (int i = 0; i < p->depths; ++i) {
        if (p->nonConst[i].empty()) continue;
        if (!load_bricks(p, d + "/voxels_" + std::to_string(i) + ".gvb", i)) {
            delete p;
            return nullptr;
        }
    }
    // BrickPool::initSpecial: (B + border 2) * slots texels per axis
    for (int a = 0; a < 3; ++a) {
        p->slots[a] = slots[a];
        p->bdim[a] = (p->B[a] + 2) * slots[a];
    }
    p->bcol.assign((size_t)p->bdim[0] * p->bdim[1] * p->bdim[2], 0u);
    p->bnrm.assign(p->bcol.size(), 0u);
    // GigaVoxelsRenderer.cpp 285-286: uploadInitialBricks, then the node texture
    init_root_brick(p);
    if (!create_texture(p)) {
        delete p;
        return nullptr;
    }
    return p;
}

extern "C" void ref_pool_node_dims(const ref_pool* p, uint32_t dims[3], uint64_t* blocks) {
    for (int a = 0; a < 3; ++a) dims[a] = p->ndim[a];
    *blocks = p->nodes.size() / 8;
}
extern "C" const uint32_t* ref_pool_nodes(const ref_pool* p) { return p->tex.data(); }
extern "C" void ref_pool_brick_dims(const ref_pool* p, uint32_t dims[3]) {
    for (int a = 0; a < 3; ++a) dims[a] = p->bdim[a];
}
extern "C" const uint32_t* ref_pool_brick_colors(const ref_pool* p) { return p->bcol.data(); }
extern "C" const uint32_t* ref_pool_brick_normals(const ref_pool* p) { return p->bnrm.data(); }
extern "C" int32_t ref_pool_depths(const ref_pool* p) { return p->depths; }
extern "C" uint64_t ref_pool_file_bricks(const ref_pool* p, int32_t d) { return p->fcount[d]; }
extern "C" const uint32_t* ref_pool_file_brick_colors(const ref_pool* p, int32_t d) {
    return p->fcol[d].data();
}
extern "C" const uint32_t* ref_pool_file_brick_normals(const ref_pool* p, int32_t d) {
    return p->fnrm[d].data();
}
extern "C" int32_t ref_pool_compressed(const ref_pool* p) { return p->compressed; }
extern "C" const uint8_t* ref_pool_file_compressed(const ref_pool* p, int32_t d, int32_t normals,
                                                   uint64_t* nbytes) {
    const auto& v = normals ? p->fncomp[d] : p->fccomp[d];
    *nbytes = v.size();
    return v.data();
}
extern "C" void ref_pool_free(ref_pool* p) { delete p; }
