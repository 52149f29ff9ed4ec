// Host check build of libfhp's per-path device code (test infrastructure, not shipped).
//
// Compiled for the CPU with FHP_HOST_CHECK under AddressSanitizer (build.py): the scene is built by
// the library's own host builder (sceneFromDesc, buildBVH4, buildLights, buildLightBVH), every
// array the kernels read is bound as an exact-size heap allocation, each traversal stack is a
// single lane whose LDS part and overflow part are exact-size allocations of the launch's depth,
// and the kernels' grid-loop items (warItem, geomItem) run for every path id. Any access outside
// an array, a stack or a gradient row is reported by the sanitizer with its stack trace.
//
// Usage: hostcheck <scene.bin> <steps.txt>
//   scene.bin: named arrays written by tests/hostcheck/run.py (the scene description's arrays
//   and the dLdI image); steps.txt: one step per line, executed in order:
//     translate <mesh> <dx> <dy> <dz>   move the mesh's vertices (host vertex update + rebuild)
//     primal <W> <H> <spp> <frame> <maxBounces> <lds>
//     adjoint <W> <H> <spp> <frame> <maxBounces> <lds> <hash>
//     geom <W> <H> <spp> <frame> <aux> <log10kappa> <log10kappaScreen> <mesh> <hash> <lds> [out=<file>] [nee=0] [neeheld=1]
// (out: the mesh's position then normal gradients, hash rows summed, as float64; nee=0: no NEE;
// neeheld=1: the held NEE direction of useSolidAngleNEEWarp)
// A primal step may end with cap=<n>: the lane's stack then holds n entries in total (fewer than
// the tree needs), to check that overflow is reported without any out-of-bounds access.
// Prints one line per step: rays, the deepest stack entry used, the stack capacity and the
// trace error bits; exits non-zero when a step sets a trace error bit.
#include "../../falcor_amd/csrc/fhp_war_geom.h"
#include "../../falcor_amd/csrc/fhp_war_path.h"
#include "../../falcor_amd/csrc/scene_build.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

using namespace fhp;

namespace {

std::map<std::string, std::vector<uint8_t>> readArrays(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    char magic[4];
    f.read(magic, 4);
    if (std::memcmp(magic, "FHC1", 4) != 0) throw std::runtime_error("bad scene file");
    uint32_t n = 0;
    f.read(reinterpret_cast<char*>(&n), 4);
    std::map<std::string, std::vector<uint8_t>> out;
    for (uint32_t i = 0; i < n; i++) {
        char name[16];
        uint64_t bytes = 0;
        f.read(name, 16);
        f.read(reinterpret_cast<char*>(&bytes), 8);
        std::vector<uint8_t> v(bytes);
        f.read(reinterpret_cast<char*>(v.data()), std::streamsize(bytes));
        if (!f) throw std::runtime_error("truncated scene file");
        out[std::string(name, strnlen(name, 16))] = std::move(v);
    }
    return out;
}

// Exact-size heap copies of the arrays a SceneDev points at (no slack for the sanitizer to miss).
struct Arena {
    std::vector<std::unique_ptr<uint8_t[]>> blocks;
    template <typename T, typename S> const T* dup(const std::vector<S>& v) {
        const size_t bytes = v.size() * sizeof(S);
        if (bytes % sizeof(T)) throw std::runtime_error("array size is not a whole number of records");
        std::unique_ptr<uint8_t[]> b(new uint8_t[bytes]);
        if (bytes) std::memcpy(b.get(), v.data(), bytes);
        blocks.push_back(std::move(b));
        return reinterpret_cast<const T*>(blocks.back().get());
    }
};

// The binding buildAndUpload (fhp_api.cpp) gives the kernels, from host copies.
SceneDev bindScene(const HostScene& h, Arena& ar) {
    SceneDev dv{};
    dv.bvh = ar.dup<float4>(h.bvh);
    dv.triPos = ar.dup<float4>(h.triPos);
    dv.idx = ar.dup<uint32_t>(h.idx);
    dv.vtx = ar.dup<float4>(packVertexRecords(h));
    dv.triMat = ar.dup<uint32_t>(h.triMat);
    dv.mats = ar.dup<float>(h.mats);
    dv.texels = ar.dup<float>(h.texels);
    dv.matsS = dv.mats;
    dv.texelsS = dv.texels;
    dv.texDesc = ar.dup<int4>(h.texDesc);
    dv.lights = ar.dup<float4>(h.lights);
    dv.activeLights = ar.dup<uint32_t>(h.activeLights);
    dv.triToLight = ar.dup<uint32_t>(h.triToLight);
    dv.lbvh = ar.dup<uint4>(h.lbvhPacked);
    dv.lbvhS = reinterpret_cast<const float4*>(reinterpret_cast<const uint32_t*>(dv.lbvh) + 8 * size_t(h.nLbvh));
    dv.lbvhTri = ar.dup<uint32_t>(h.lbvhTri);
    dv.lbvhMask = ar.dup<uint2>(h.lbvhMask);
    dv.cam.pos = mk3(h.cam[0], h.cam[1], h.cam[2]);
    dv.cam.U = mk3(h.camU[0], h.camU[1], h.camU[2]);
    dv.cam.V = mk3(h.camV[0], h.camV[1], h.camV[2]);
    dv.cam.W = mk3(h.camW[0], h.camW[1], h.camW[2]);
    dv.cam.nearZ = h.cam[12];
    dv.cam.farZ = h.cam[13];
    dv.meshTriOff = h.meshTriContiguous ? ar.dup<uint32_t>(h.meshTriOff) : nullptr;
    dv.nMeshes = uint32_t(h.meshOff.size() - 1);
    dv.hit = HitInfoLayout{h.hitCompressed, h.hitTypeBits, h.hitInstBits, h.hitPrimBits};
    dv.nTris = h.nT;
    dv.hasTransmission = hasTransmission(h) ? 1u : 0u;
    dv.nLights = h.nLights;
    dv.nActive = uint32_t(h.activeLights.size());
    dv.nLbvh = h.nLbvh;
    return dv;
}

// One lane's traversal stack: LDS part of `lds` entries and the overflow part the library sizes
// from the tree (stackNeed + 1 - lds entries), each its own exact-size allocation.
struct LaneStack {
    std::unique_ptr<uint32_t[]> lds, ovf;
    uint32_t err = 0;
    TraceStack st{};
    LaneStack(uint32_t ldsDepth, uint32_t stackNeed, uint32_t cap = 0) {
        uint32_t need = stackNeed + 1;  // as buildAndUpload sizes it
        if (cap) {
            need = cap;
            ldsDepth = std::min(ldsDepth, cap);
        }
        const uint32_t ovfDepth = need > ldsDepth ? need - ldsDepth : 0u;
        lds.reset(new uint32_t[ldsDepth]);
        if (ovfDepth) ovf.reset(new uint32_t[ovfDepth]);
        st = TraceStack{lds.get(), 1u, ldsDepth, ovf.get(), 1u, ovfDepth, &err};
    }
};

struct Checker {
    FhpSceneDesc desc{};
    HostScene h;
    std::map<std::string, std::vector<uint8_t>> arrays;
    std::vector<float> pos;  // current vertex positions (the description points here)

    template <typename T> const T* arr(const char* name, size_t* count = nullptr) {
        auto it = arrays.find(name);
        if (it == arrays.end()) throw std::runtime_error(std::string("missing array ") + name);
        if (count) *count = it->second.size() / sizeof(T);
        return it->second.empty() ? nullptr : reinterpret_cast<const T*>(it->second.data());
    }

    void build() {
        std::string msg;
        h = HostScene{};
        if (sceneFromDesc(desc, h, msg)) throw std::runtime_error("scene description: " + msg);
        buildBVH4(h, 4);
        buildLights(h);
        buildLightBVH(h);
    }

    explicit Checker(const char* path) : arrays(readArrays(path)) {
        size_t nv = 0, nt = 0, nm = 0, ntex = 0, ntx = 0, nmesh = 0;
        const float* p = arr<float>("positions", &nv);
        pos.assign(p, p + nv);
        desc.num_vertices = uint32_t(nv / 3);
        desc.h_positions = pos.data();
        desc.h_normals = arr<float>("normals");
        desc.h_texcoords = arr<float>("texcoords");
        desc.h_indices = arr<uint32_t>("indices", &nt);
        desc.num_triangles = uint32_t(nt / 3);
        desc.h_tri_material = arr<uint32_t>("tri_material");
        desc.h_materials = arr<float>("materials", &nm);
        desc.num_materials = uint32_t(nm / M_STRIDE);
        desc.h_tex_desc = arr<int32_t>("tex_desc", &ntex);
        desc.num_textures = uint32_t(ntex / 4);
        desc.h_texels = arr<float>("texels", &ntx);
        desc.num_texels_floats = ntx;
        desc.h_camera = arr<float>("camera");
        desc.h_mesh_vertex_offsets = arr<uint32_t>("mesh_offsets", &nmesh);
        desc.num_meshes = uint32_t(nmesh - 1);
        build();
    }

    void translate(uint32_t mesh, float dx, float dy, float dz) {
        if (mesh + 1 >= h.meshOff.size()) throw std::runtime_error("mesh id out of range");
        for (uint32_t v = h.meshOff[mesh]; v < h.meshOff[mesh + 1]; v++) {
            pos[3 * v] += dx;
            pos[3 * v + 1] += dy;
            pos[3 * v + 2] += dz;
        }
        build();  // fhp_scene_update_vertices: host vertex update, then buildAndUpload's rebuild
    }
};

void report(const char* what, uint64_t paths, unsigned long long rays, const LaneStack& ls, uint32_t peak) {
    std::printf("%s paths=%llu rays=%llu stack_peak=%u stack_capacity=%u trace_errors=%u\n", what,
                (unsigned long long)paths, rays, peak, ls.st.capacity(), ls.err);
    std::fflush(stdout);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: hostcheck <scene.bin> <steps.txt>\n");
        return 2;
    }
    try {
        Checker ck(argv[1]);
        size_t ndl = 0;
        const float* dLdI = ck.arr<float>("dLdI", &ndl);
        std::ifstream steps(argv[2]);
        std::string line;
        int bad = 0;
        while (std::getline(steps, line)) {
            std::istringstream in(line);
            std::string op;
            if (!(in >> op)) continue;
            if (op == "translate") {
                uint32_t mesh;
                float dx, dy, dz;
                in >> mesh >> dx >> dy >> dz;
                ck.translate(mesh, dx, dy, dz);
                std::printf("translate mesh=%u stack_need=%u bvh_nodes=%u\n", mesh, ck.h.stackNeed, ck.h.bvhNodes);
                continue;
            }
            Arena ar;
            const SceneDev s = bindScene(ck.h, ar);
            if (op == "primal" || op == "adjoint") {
                uint32_t W, H, spp, frame, B, lds, hash = 1;
                in >> W >> H >> spp >> frame >> B >> lds;
                const bool adj = op == "adjoint";
                if (adj) in >> hash;
                if (!in) throw std::runtime_error("bad step: " + line);
                uint32_t cap = 0;
                bool useLog = false;
                for (std::string opt; in >> opt;) {
                    if (!adj && opt.rfind("cap=", 0) == 0) cap = uint32_t(std::stoul(opt.substr(4)));
                    else if (adj && opt == "log=1") useLog = true;
                    else throw std::runtime_error("bad step option: " + opt);
                }
                if (adj && ndl != 3 * size_t(W) * H) throw std::runtime_error("dLdI size does not match the frame");
                std::vector<float4> out(adj ? 0 : size_t(W) * H * spp);
                std::vector<float> grad(size_t(hash) * ck.h.nM * 20, 0.f);
                std::vector<float> gradTex(ck.h.texels.size(), 0.f);
                WarArgs a{};
                a.W = W; a.H = H; a.spp = spp; a.frame = frame * spp; a.maxBounces = B;
                a.useBSDFSampling = a.useNEE = a.useMIS = a.useAntithetic = 1;
                a.out = out.data();
                a.dLdI = dLdI;
                a.gradMat = grad.data();
                a.hashSize = hash;
                a.matGradDim = ck.h.nM * 20;
                a.gradTex = gradTex.empty() ? nullptr : gradTex.data();
                LaneStack ls(lds, ck.h.stackNeed, cap);
                // The adjoint's hit log of one thread (the only thread here): exact size, so an
                // entry beyond its depth is an out-of-bounds access.
                std::unique_ptr<uint32_t[]> logBuf(useLog ? new uint32_t[B + 2]() : nullptr);
                const HitLog hl = useLog ? HitLog{logBuf.get(), 1u, B + 2} : noLog();
                unsigned long long nc = 0, nv = 0;
                uint32_t peak = 0;
                const uint64_t n = uint64_t(W) * H * spp;
                for (uint32_t id = 0; id < n; id++) {
                    if (adj && s.hasTransmission) warItem<true, false, true>(s, a, id, ls.st, nc, nv, hl);
                    else if (adj) warItem<true, false>(s, a, id, ls.st, nc, nv, hl);
                    else warItem<false, false>(s, a, id, ls.st, nc, nv);
                    peak = std::max(peak, ls.st.peak);
                }
                report(op.c_str(), n, nc + nv, ls, peak);
                if (adj) {
                    // Order-independent fingerprint of the gradients (bit patterns summed).
                    uint64_t fp = 0;
                    for (float v : grad) fp += uint64_t(f2u(v)) * 2654435761u;
                    for (float v : gradTex) fp += uint64_t(f2u(v)) * 2246822519u;
                    std::printf("adjoint grad_fingerprint=%016llx log=%d\n", (unsigned long long)fp, int(useLog));
                }
                bad |= ls.err != 0;
            } else if (op == "geom") {
                uint32_t W, H, spp, frame, aux, mesh, hash, lds;
                float lk, lks;
                in >> W >> H >> spp >> frame >> aux >> lk >> lks >> mesh >> hash >> lds;
                if (!in) throw std::runtime_error("bad step: " + line);
                std::string gradOut;
                uint32_t useNEE = 1, neeHeld = 0, bounces = 0;
                bool useLog = false, general = false;  // general: the recorded reverse sweep also at 0 bounces
                for (std::string opt; in >> opt;) {
                    if (opt.rfind("out=", 0) == 0) gradOut = opt.substr(4);
                    else if (opt == "nee=0") useNEE = 0;
                    else if (opt == "neeheld=1") neeHeld = 1;
                    else if (opt == "log=1") useLog = true;
                    else if (opt.rfind("b=", 0) == 0) bounces = uint32_t(std::stoul(opt.substr(2)));
                    else if (opt == "general=1") general = true;
                    else throw std::runtime_error("bad step option: " + opt);
                }
                if (ndl != 3 * size_t(W) * H) throw std::runtime_error("dLdI size does not match the frame");
                if (mesh + 1 >= ck.h.meshOff.size()) throw std::runtime_error("mesh id out of range");
                const uint32_t v0 = ck.h.meshOff[mesh], nvtx = ck.h.meshOff[mesh + 1] - v0;
                // Exact-size gradient rows, as fhp_grads_set_mesh allocates them.
                std::unique_ptr<float[]> gp(new float[size_t(hash) * 3 * nvtx]()), gn(new float[size_t(hash) * 3 * nvtx]());
                GeomArgs a{};
                a.W = W; a.H = H; a.spp = spp; a.frame = frame * spp;
                a.useBSDFSampling = a.useNEE = a.useMIS = a.useAntithetic = 1;
                a.useNEE = useNEE;
                a.neeHeld = neeHeld;
                a.aux = aux;
                a.kappa = float(std::pow(10.0, double(lk)));
                a.kappaScreen = float(std::pow(10.0, double(lks)));
                a.dLdI = dLdI;
                a.gradPos = gp.get();
                a.gradNrm = gn.get();
                a.hashSize = hash;
                a.meshV0 = v0;
                a.meshVerts = nvtx;
                LaneStack ls(lds, ck.h.stackNeed);
                unsigned long long nc = 0;
                uint32_t peak = 0;
                const uint64_t n = uint64_t(W) * H * spp;
                // The aux-hit log of the one thread here: exact size (3 words per aux ray).
                std::unique_ptr<uint32_t[]> logBuf(useLog ? new uint32_t[3 * size_t(aux)]() : nullptr);
                const AuxLog lg = useLog ? AuxLog{logBuf.get(), 1u} : noAuxLog();
                // The vertex records of the one thread here, exact size.
                a.maxBounces = bounces;
                std::unique_ptr<uint32_t[]> vtxBuf(new uint32_t[size_t(kGeomVertexWords) * (bounces + 2)]());
                const VertexLog vl{vtxBuf.get(), 1u, bounces + 2};
                for (uint32_t id = 0; id < n; id++) {
                    if (bounces == 0 && !general) geomItem<true>(s, a, id, ls.st, nc, lg, vl);
                    else geomItem<false>(s, a, id, ls.st, nc, lg, vl);
                    peak = std::max(peak, ls.st.peak);
                }
                {
                    uint64_t fp = 0;
                    for (size_t i = 0; i < size_t(hash) * 3 * nvtx; i++)
                        fp += uint64_t(f2u(gp[i])) * 2654435761u + uint64_t(f2u(gn[i])) * 2246822519u;
                    std::printf("geom grad_fingerprint=%016llx log=%d\n", (unsigned long long)fp, int(useLog));
                }
                double sum = 0.0;
                bool finite = true;
                for (size_t i = 0; i < size_t(hash) * 3 * nvtx; i++) {
                    sum += std::fabs(double(gp[i]));
                    finite &= std::isfinite(gp[i]) && std::isfinite(gn[i]);
                }
                report("geom", n, nc, ls, peak);
                std::printf("geom grad_abs_sum=%.6e finite=%d\n", sum, int(finite));
                if (!gradOut.empty()) {
                    // Position then normal gradients of the mesh's vertices, hash rows summed (float64).
                    std::vector<double> g(6 * size_t(nvtx), 0.0);
                    for (uint32_t r = 0; r < hash; r++)
                        for (size_t i = 0; i < 3 * size_t(nvtx); i++) {
                            g[i] += gp[size_t(r) * 3 * nvtx + i];
                            g[3 * size_t(nvtx) + i] += gn[size_t(r) * 3 * nvtx + i];
                        }
                    std::ofstream f(gradOut, std::ios::binary);
                    f.write(reinterpret_cast<const char*>(g.data()), std::streamsize(g.size() * sizeof(double)));
                }
                bad |= ls.err != 0;
            } else {
                throw std::runtime_error("unknown step: " + line);
            }
        }
        return bad ? 1 : 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "hostcheck: %s\n", e.what());
        return 2;
    }
}
