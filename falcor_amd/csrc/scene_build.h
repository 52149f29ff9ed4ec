// Host-side scene build: BVH4 over triangles, light collection and light BVH.
// These replace the driver-built BLAS/TLAS (Scene.cpp:3376, :3936), LightCollection::build
// (LightCollection.cpp:220) and LightBVHBuilder::build (LightBVHBuilder.cpp:228).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fhp.h"

namespace fhp {

struct HostScene {
    // Input (world space).
    uint32_t nV = 0, nT = 0, nM = 0, nTex = 0;
    std::vector<float> pos, nrm, uv;  // 3, 3, 2 floats per vertex
    std::vector<uint32_t> idx, triMat;
    std::vector<uint32_t> meshOff;  // mesh m owns vertices [meshOff[m], meshOff[m + 1])
    std::vector<float> mats;
    std::vector<int32_t> texDesc;
    std::vector<float> texels;
    float cam[14] = {};  // pos, target, up, fovY, aspect, focalDistance, nearZ, farZ
    // Camera basis (Camera.cpp:187-193).
    float camU[3], camV[3], camW[3];
    // View-projection matrix without jitter, row-major (Camera.cpp: viewMat = matrixFromLookAt,
    // projMat = perspective with depth in [0, 1], MatrixMath.h:553-566, :699-725).
    float viewProj[16];

    // Geometry instances (meshes) for PackedHitInfo: first triangle of each mesh (meshes + 1
    // entries), valid when every mesh's triangles are contiguous; bit allocation (HitInfo.cpp).
    std::vector<uint32_t> meshTriOff;
    bool meshTriContiguous = false;
    uint32_t hitCompressed = 0, hitTypeBits = 0, hitInstBits = 0, hitPrimBits = 0;

    // BVH4 (quantised, 4 float4 per node: packBVH4Node) and triangles in leaf order (12 floats
    // per triangle).
    std::vector<float> bvh;
    std::vector<float> triPos;
    uint32_t bvhNodes = 0, bvhDepth = 0, stackNeed = 0, bvhLeaves = 0;
    double bvhSah = 0.0;
    float bvhAbsPad = 0.f;  // absolute part of the outward box padding (scene extent * 1e-6)
    // Refit schedule: node ids grouped by depth (levelOff[d] .. levelOff[d + 1]), root first.
    std::vector<uint32_t> levelNodes, levelOff;

    // Lights: 20 floats per light, see SceneDev::lights.
    std::vector<float> lights;
    std::vector<uint32_t> activeLights, triToLight;
    // Light BVH: the packed nodes the device samples (8 words per node, packLightNode), their
    // decoded attributes as 16 floats per node ((origin, flux) (extent, cosCone) (coneDir, 0)
    // (right child | leaf bit, first triangle, triangle count, 0)), leaf triangle lists and
    // per-light traversal bitmasks.
    std::vector<uint32_t> lbvhPacked;  // packed nodes (8 words each), then their sampling records (12 each)
    std::vector<float> lbvh;
    std::vector<uint32_t> lbvhTri;
    std::vector<uint32_t> lbvhMask;  // 2 words per light
    uint32_t nLights = 0, nLbvh = 0;
    // Light BVH refit schedule (LightBVH::updateNodeIndices, LightBVH.cpp:219-250): internal
    // nodes of depth d at lrefitNodes[lrefitOff[d] .. lrefitOff[d + 1]) for d < lrefitDepth, then
    // every leaf (any depth) at [lrefitOff[lrefitDepth] .. lrefitOff[lrefitDepth + 1]).
    std::vector<uint32_t> lrefitNodes, lrefitOff;
    uint32_t lrefitDepth = 0;
    // Rest pose of every vertex (positions, normals): the vertices as created or as last set by a
    // vertex update; a node transform maps a mesh's rest pose to its current pose.
    std::vector<float> restPos, restNrm;
};

// fhp_scene_create's validation and host copy of a scene description: FHP_OK, or FHP_EINVAL /
// FHP_EBUILD with the message in err. Computes the camera basis and the PackedHitInfo layout.
int sceneFromDesc(const FhpSceneDesc& d, HostScene& h, std::string& err);
// Vertex records as the kernels read them: (p.xyz, n.x) (n.yz, u, v), 8 floats per vertex.
std::vector<float> packVertexRecords(const HostScene& h);
// True when some StandardMaterial has specular or diffuse transmission (such scenes run the
// adjoint instances that differentiate the transmission lobes).
bool hasTransmission(const HostScene& h);
void computeCamera(HostScene& s);
// Mesh triangle ranges and the PackedHitInfo bit allocation; throws when the header does not fit
// in 64 bits (HitInfo.cpp:83-87).
void computeHitInfo(HostScene& s, bool useCompression);
void buildBVH4(HostScene& s, uint32_t maxLeaf = 4);
void buildLights(HostScene& s);
void refreshLightEmission(HostScene& s);
void buildLightBVH(HostScene& s);
// Rewrites the decoded host mirror (lbvh) from the packed nodes (after a device refit).
void decodeLightNodes(HostScene& s);

}  // namespace fhp
