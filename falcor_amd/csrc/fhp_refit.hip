// BVH refit after a mesh's vertices move (the reference refits BLASes by default on
// setMeshVertices: Scene.cpp:4314-4336, Scene.h:1366 UpdateMode::Refit). The tree topology and
// the leaf order stay; the updated vertices are written into the vertex records and into the
// leaf-order triangle records, then every node's four child boxes are recomputed level by level
// from the deepest level up, with the build's outward padding.
#include "fhp_internal.h"

namespace fhp {
namespace {

constexpr int kRefitBlock = 256;

__global__ void kUpdateVertices(float4* __restrict__ vtx, uint32_t v0, uint32_t n, const float* __restrict__ pos,
                                const float* __restrict__ nrm) {
    const uint32_t i = blockIdx.x * kRefitBlock + threadIdx.x;
    if (i >= n) return;
    const size_t v = size_t(v0) + i;
    float4 a = vtx[2 * v], b = vtx[2 * v + 1];  // (p.xyz, n.x) (n.yz, uv)
    a.x = pos[3 * size_t(i)];
    a.y = pos[3 * size_t(i) + 1];
    a.z = pos[3 * size_t(i) + 2];
    if (nrm) {
        a.w = nrm[3 * size_t(i)];
        b.x = nrm[3 * size_t(i) + 1];
        b.y = nrm[3 * size_t(i) + 2];
    }
    vtx[2 * v] = a;
    vtx[2 * v + 1] = b;
}

// Leaf-order triangle records (v0, id) (v1, 0) (v2, 0) of the triangles indexing vertices
// [v0, v1) (a mesh's triangles index only its own vertices).
__global__ void kUpdateTriPos(float4* __restrict__ triPos, uint32_t nT, const uint32_t* __restrict__ idx,
                              const float4* __restrict__ vtx, uint32_t v0, uint32_t v1) {
    const uint32_t i = blockIdx.x * kRefitBlock + threadIdx.x;
    if (i >= nT) return;
    const float idw = triPos[3 * size_t(i)].w;
    const uint32_t t = __float_as_uint(idw);
    const uint32_t* tv = idx + 3 * size_t(t);
    if (tv[0] < v0 || tv[0] >= v1) return;
    for (int j = 0; j < 3; j++) {
        const float4 p = vtx[2 * size_t(tv[j])];
        triPos[3 * size_t(i) + j] = make_float4(p.x, p.y, p.z, j == 0 ? idw : 0.f);
    }
}

__device__ __forceinline__ float padLo(float v, float absPad) { return v - (fabsf(v) * 1e-6f + absPad); }
__device__ __forceinline__ float padHi(float v, float absPad) { return v + (fabsf(v) * 1e-6f + absPad); }

// One depth level: node layout per node (32 floats) = lo/hi per axis for children 0..3
// ([8a + i], [8a + 4 + i]) then the child references ([24 + i]).
__global__ void kRefitLevel(float* __restrict__ bvh, const uint32_t* __restrict__ nodes, uint32_t count,
                            const float4* __restrict__ triPos, float absPad) {
    const uint32_t k = blockIdx.x * kRefitBlock + threadIdx.x;
    if (k >= count) return;
    float* q = bvh + 32 * size_t(nodes[k]);
    for (int i = 0; i < 4; i++) {
        const uint32_t c = __float_as_uint(q[24 + i]);
        if (c == kInvalid) continue;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (childIsLeaf(c)) {
            const uint32_t first = leafFirst(c), cnt = leafCount(c);
            for (uint32_t s = first; s < first + cnt; s++)
                for (int j = 0; j < 3; j++) {
                    const float4 p = triPos[3 * size_t(s) + j];
                    lo[0] = fminf(lo[0], p.x); hi[0] = fmaxf(hi[0], p.x);
                    lo[1] = fminf(lo[1], p.y); hi[1] = fmaxf(hi[1], p.y);
                    lo[2] = fminf(lo[2], p.z); hi[2] = fmaxf(hi[2], p.z);
                }
        } else {
            const float* cq = bvh + 32 * size_t(c);
            for (int ci = 0; ci < 4; ci++) {
                if (__float_as_uint(cq[24 + ci]) == kInvalid) continue;
                for (int a = 0; a < 3; a++) {
                    lo[a] = fminf(lo[a], cq[8 * a + ci]);
                    hi[a] = fmaxf(hi[a], cq[8 * a + 4 + ci]);
                }
            }
        }
        for (int a = 0; a < 3; a++) {
            q[8 * a + i] = padLo(lo[a], absPad);
            q[8 * a + 4 + i] = padHi(hi[a], absPad);
        }
    }
}

uint32_t blocksFor(uint64_t n) { return uint32_t((n + kRefitBlock - 1) / kRefitBlock); }

}  // namespace

hipError_t launchRefit(const RefitArgs& r, hipStream_t st) {
    if (r.nVerts) {
        hipLaunchKernelGGL(kUpdateVertices, dim3(blocksFor(r.nVerts)), dim3(kRefitBlock), 0, st, r.vtx, r.v0, r.nVerts, r.pos,
                           r.nrm);
        hipLaunchKernelGGL(kUpdateTriPos, dim3(blocksFor(r.nTris)), dim3(kRefitBlock), 0, st, r.triPos, r.nTris, r.idx, r.vtx,
                           r.v0, r.v0 + r.nVerts);
    }
    for (int d = int(r.nLevels) - 1; d >= 0; d--) {
        const uint32_t b = r.levelOff[d], e = r.levelOff[d + 1];
        if (e > b)
            hipLaunchKernelGGL(kRefitLevel, dim3(blocksFor(e - b)), dim3(kRefitBlock), 0, st, r.bvh, r.levelNodes + b, e - b,
                               r.triPos, r.absPad);
    }
    return hipGetLastError();
}

}  // namespace fhp
