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

// One depth level: each node's four child boxes (a leaf child's from its triangles, an internal
// child's as the union of the boxes its own record decodes to), padded as the build pads them,
// then repacked into the node's quantised record (packBVH4Node). Children are one level deeper,
// so their records are already refitted.
__global__ void kRefitLevel(float* __restrict__ bvh, const uint32_t* __restrict__ nodes, uint32_t count,
                            const float4* __restrict__ triPos, float absPad) {
    const uint32_t k = blockIdx.x * kRefitBlock + threadIdx.x;
    if (k >= count) return;
    uint32_t* q = reinterpret_cast<uint32_t*>(bvh) + kBVHNodeWords * size_t(nodes[k]);
    uint32_t ref[4];
    float clo[4][3], chi[4][3];
    for (int i = 0; i < 4; i++) {
        const uint32_t c = q[12 + i];
        ref[i] = c;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (c != kInvalid && childIsLeaf(c)) {
            const uint32_t first = leafFirst(c), cnt = leafCount(c);
            for (uint32_t s = first; s < first + cnt; s++)
                for (int j = 0; j < 3; j++) {
                    const float4 p = triPos[3 * size_t(s) + j];
                    lo[0] = fminf(lo[0], p.x); hi[0] = fmaxf(hi[0], p.x);
                    lo[1] = fminf(lo[1], p.y); hi[1] = fmaxf(hi[1], p.y);
                    lo[2] = fminf(lo[2], p.z); hi[2] = fmaxf(hi[2], p.z);
                }
        } else if (c != kInvalid) {
            const uint32_t* cq = reinterpret_cast<const uint32_t*>(bvh) + kBVHNodeWords * size_t(c);
            for (int ci = 0; ci < 4; ci++) {
                if (cq[12 + ci] == kInvalid) continue;
                for (int a = 0; a < 3; a++) {
                    const float o = __uint_as_float(cq[a]), st = __uint_as_float(cq[3 + a]);
                    lo[a] = fminf(lo[a], bvhPlane(o, st, (cq[6 + 2 * a] >> (8 * ci)) & 0xffu));
                    hi[a] = fmaxf(hi[a], bvhPlane(o, st, (cq[7 + 2 * a] >> (8 * ci)) & 0xffu));
                }
            }
        }
        for (int a = 0; a < 3; a++) {
            clo[i][a] = c != kInvalid ? padLo(lo[a], absPad) : 0.f;
            chi[i][a] = c != kInvalid ? padHi(hi[a], absPad) : 0.f;
        }
    }
    uint32_t w[kBVHNodeWords];
    packBVH4Node(w, clo, chi, ref);
    for (uint32_t j = 0; j < kBVHNodeWords; j++) q[j] = w[j];
}

// Node transform: rest-pose records (p.xyz, n.x) (n.yz, uv) -> (M p, N n); uvs kept. The normal
// stays unnormalised, as the reference transforms interpolated normals by the inverse transpose
// and normalises after interpolation (Scene.slang computeVertexData).
struct Affine { float m[12]; };
struct Linear3 { float m[9]; };
__global__ void kTransformVertices(float4* __restrict__ vtx, const float4* __restrict__ rest, uint32_t v0, uint32_t n,
                                   Affine M, Linear3 N) {
    const uint32_t i = blockIdx.x * kRefitBlock + threadIdx.x;
    if (i >= n) return;
    const size_t v = size_t(v0) + i;
    const float4 a = rest[2 * v], b = rest[2 * v + 1];
    const float px = a.x, py = a.y, pz = a.z, nx = a.w, ny = b.x, nz = b.y;
    float4 oa, ob;
    oa.x = M.m[0] * px + M.m[1] * py + M.m[2] * pz + M.m[3];
    oa.y = M.m[4] * px + M.m[5] * py + M.m[6] * pz + M.m[7];
    oa.z = M.m[8] * px + M.m[9] * py + M.m[10] * pz + M.m[11];
    oa.w = N.m[0] * nx + N.m[1] * ny + N.m[2] * nz;
    ob.x = N.m[3] * nx + N.m[4] * ny + N.m[5] * nz;
    ob.y = N.m[6] * nx + N.m[7] * ny + N.m[8] * nz;
    ob.z = b.z;
    ob.w = b.w;
    vtx[2 * v] = oa;
    vtx[2 * v + 1] = ob;
}

// UpdateTriangleVertices.cs.slang:42-63 with Scene::computeFaceNormalAndAreaW (Scene.slang:392-407):
// light record (p0, area) (p1, flux) (p2, triangle) (normal, 0) (Le, texture); flux is kept.
__global__ void kUpdateLights(float* __restrict__ lights, uint32_t n, const float4* __restrict__ vtx,
                              const uint32_t* __restrict__ idx) {
    const uint32_t i = blockIdx.x * kRefitBlock + threadIdx.x;
    if (i >= n) return;
    float* r = lights + 20 * size_t(i);
    const uint32_t t = __float_as_uint(r[11]);
    f3 p[3];
    for (int j = 0; j < 3; j++) {
        const float4 q = vtx[2 * size_t(idx[3 * size_t(t) + j])];
        p[j] = mk3(q.x, q.y, q.z);
    }
    const f3 N = cross(p[1] - p[0], p[2] - p[0]);
    const float area = 0.5f * length(N);
    const f3 nn = normalize(N);
    for (int j = 0; j < 3; j++) {
        r[4 * j] = p[j].x;
        r[4 * j + 1] = p[j].y;
        r[4 * j + 2] = p[j].z;
    }
    r[3] = area;
    r[12] = nn.x;
    r[13] = nn.y;
    r[14] = nn.z;
}

constexpr float kInvalidCone = -1.f;  // LightBVHTypes.slang:43 (kInvalidCosConeAngle)
// Packed light-BVH nodes (packLightNode): the refit reads and writes the packed layout, as
// LightBVHRefit.cs.slang does through PackedNode; internal nodes see their children's decoded
// (quantised) attributes.
__device__ __forceinline__ LNodeAttr readLightNode(const uint32_t* lbvh, uint32_t n) {
    const uint4* q = reinterpret_cast<const uint4*>(lbvh) + 2 * size_t(n);
    return unpackLightNode(q[0], q[1]);
}
// Packs node n and rewrites its sampling record (after the nNodes packed nodes) from the packed
// values, as the scene builder derives it.
__device__ __forceinline__ void writeLightNode(uint32_t* lbvh, uint32_t nNodes, uint32_t n, const LNodeAttr& a) {
    uint32_t w[8];
    packLightNode(w, lbvh[8 * size_t(n)], a);
    uint4* q = reinterpret_cast<uint4*>(lbvh) + 2 * size_t(n);
    const uint4 q0 = make_uint4(w[0], w[1], w[2], w[3]), q1 = make_uint4(w[4], w[5], w[6], w[7]);
    q[0] = q0;
    q[1] = q1;
    float rec[kLNodeSampleWords];
    lightNodeSampleRecord(rec, w[0], unpackLightNode(q0, q1));
    float4* r = reinterpret_cast<float4*>(lbvh + 8 * size_t(nNodes)) + 3 * size_t(n);
    r[0] = make_float4(rec[0], rec[1], rec[2], rec[3]);
    r[1] = make_float4(rec[4], rec[5], rec[6], rec[7]);
    r[2] = make_float4(rec[8], rec[9], rec[10], rec[11]);
}
__device__ __forceinline__ void setAABB(LNodeAttr& a, f3 lo, f3 hi) {
    a.origin = (hi + lo) * 0.5f;
    a.extent = (hi - lo) * 0.5f;
}

// LightBVHRefit.cs.slang:46-93 (updateLeafNodes).
__global__ void kLightRefitLeaves(uint32_t* __restrict__ lbvh, uint32_t nNodes, const uint32_t* __restrict__ nodes, uint32_t count,
                                  const float* __restrict__ lights, const uint32_t* __restrict__ lbvhTri) {
    const uint32_t k = blockIdx.x * kRefitBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t n = nodes[k];
    const uint32_t w0 = lbvh[8 * size_t(n)];
    const uint32_t first = w0 & kLNodeOffsetMask, cnt = (w0 >> kLNodeCountShift) & kLNodeCountMask;
    LNodeAttr a = readLightNode(lbvh, n);
    f3 lo = mk3(kFltMax), hi = mk3(-kFltMax), sum = mk3(0.f);
    for (uint32_t i = 0; i < cnt; i++) {
        const float* r = lights + 20 * size_t(lbvhTri[first + i]);
        for (int j = 0; j < 3; j++) {
            const f3 p = mk3(r[4 * j], r[4 * j + 1], r[4 * j + 2]);
            lo = mk3(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z));
            hi = mk3(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z));
        }
        sum = sum + mk3(r[12], r[13], r[14]);
    }
    setAABB(a, lo, hi);
    const float len = length(sum);
    const f3 dir = sum / len;
    float cosCone = kInvalidCone;
    if (len >= kFltMin) {
        cosCone = 1.f;
        for (uint32_t i = 0; i < cnt; i++) {
            const float* r = lights + 20 * size_t(lbvhTri[first + i]);
            cosCone = fminf(cosCone, dot(dir, mk3(r[12], r[13], r[14])));
        }
        cosCone = fmaxf(cosCone, -1.f);
    }
    a.cosCone = cosCone;
    a.coneDir = dir;
    writeLightNode(lbvh, nNodes, n, a);
}

__device__ __forceinline__ float sinFromCos(float c) { return sqrtf(fmaxf(0.f, 1.f - c * c)); }

// LightBVHRefit.cs.slang:101-170 (updateInternalNodes): box union, and the normal cone that holds
// both children's cones (each child's spread rotated away from the merged direction).
__global__ void kLightRefitInternal(uint32_t* __restrict__ lbvh, uint32_t nNodes, const uint32_t* __restrict__ nodes,
                                    uint32_t count) {
    const uint32_t k = blockIdx.x * kRefitBlock + threadIdx.x;
    if (k >= count) return;
    const uint32_t ni = nodes[k];
    LNodeAttr a = readLightNode(lbvh, ni);
    const LNodeAttr L = readLightNode(lbvh, ni + 1), R = readLightNode(lbvh, lbvh[8 * size_t(ni)]);
    const f3 lo = mk3(fminf(L.origin.x - L.extent.x, R.origin.x - R.extent.x), fminf(L.origin.y - L.extent.y, R.origin.y - R.extent.y),
                      fminf(L.origin.z - L.extent.z, R.origin.z - R.extent.z));
    const f3 hi = mk3(fmaxf(L.origin.x + L.extent.x, R.origin.x + R.extent.x), fmaxf(L.origin.y + L.extent.y, R.origin.y + R.extent.y),
                      fmaxf(L.origin.z + L.extent.z, R.origin.z + R.extent.z));
    setAABB(a, lo, hi);
    const f3 ld = L.coneDir, rd = R.coneDir;
    const float lc = L.cosCone, rc = R.cosCone;
    const f3 sum = ld + rd;
    const float len = length(sum);
    const f3 dir = sum / len;
    float cosCone = kInvalidCone;
    if (len >= kFltMin && lc != kInvalidCone && rc != kInvalidCone) {
        const float cosL = dot(dir, ld), sinL = sinFromCos(cosL);
        const float cosR = dot(dir, rd), sinR = sinFromCos(cosR);
        const float sinLc = sinFromCos(lc), sinRc = sinFromCos(rc);
        const float sinLt = sinLc * cosL + sinL * lc;
        const float sinRt = sinRc * cosR + sinR * rc;
        if (sinLt > 0.f && sinRt > 0.f) {
            const float cosLt = lc * cosL - sinLc * sinL;
            const float cosRt = rc * cosR - sinRc * sinR;
            cosCone = fmaxf(fminf(cosLt, cosRt), -1.f);
        }
    }
    a.cosCone = cosCone;
    a.coneDir = dir;
    writeLightNode(lbvh, nNodes, ni, a);
}

uint32_t blocksFor(uint64_t n) { return uint32_t((n + kRefitBlock - 1) / kRefitBlock); }

}  // namespace

hipError_t launchRefit(const RefitArgs& r, hipStream_t st) {
    if (r.nVerts) {
        if (r.pos)
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

hipError_t launchNodeTransform(const NodeXformArgs& a, hipStream_t st) {
    if (a.nVerts) {
        Affine M;
        Linear3 N;
        for (int i = 0; i < 12; i++) M.m[i] = a.M[i];
        for (int i = 0; i < 9; i++) N.m[i] = a.N[i];
        hipLaunchKernelGGL(kTransformVertices, dim3(blocksFor(a.nVerts)), dim3(kRefitBlock), 0, st, a.vtx, a.rest, a.v0,
                           a.nVerts, M, N);
    }
    return hipGetLastError();
}

hipError_t launchLightRefit(const LightRefitArgs& a, hipStream_t st) {
    if (a.nLights)
        hipLaunchKernelGGL(kUpdateLights, dim3(blocksFor(a.nLights)), dim3(kRefitBlock), 0, st, a.lights, a.nLights, a.vtx,
                           a.idx);
    const uint32_t lb = a.off[a.depth], le = a.off[a.depth + 1];
    if (le > lb)
        hipLaunchKernelGGL(kLightRefitLeaves, dim3(blocksFor(le - lb)), dim3(kRefitBlock), 0, st, a.lbvh, a.nNodes, a.nodes + lb,
                           le - lb, a.lights, a.lbvhTri);
    for (int d = int(a.depth) - 1; d >= 0; d--) {
        const uint32_t b = a.off[d], e = a.off[d + 1];
        if (e > b)
            hipLaunchKernelGGL(kLightRefitInternal, dim3(blocksFor(e - b)), dim3(kRefitBlock), 0, st, a.lbvh, a.nNodes, a.nodes + b,
                               e - b);
    }
    return hipGetLastError();
}

}  // namespace fhp
