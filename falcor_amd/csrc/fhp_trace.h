// BVH4 traversal for gfx950: closest-hit and any-hit queries.
//
// The stack lives in LDS, one 32-bit slot per entry laid out [entry][lane] so the 64 lanes of a
// wave touch 64 consecutive dwords (no bank conflicts). Entries beyond the LDS depth spill to a
// per-thread overflow array in HBM, so traversal is exact for any tree depth.
//
// Closest hit follows the reference's TraceRay contract (MinimalPathTracer.rt.slang:233-243,
// RaytracingInline.slang:86-250) with a deterministic tie rule: among hits with tmin <= t <= tmax
// the smallest t wins, equal t resolves to the lowest triangle id. Triangles use the watertight
// test, so the answer is the same whatever the tree shape.
#pragma once
#include "fhp_internal.h"

namespace fhp {

struct HitRec {
    uint32_t tri;
    float t, b1, b2;
};

// Address-space-qualified pointers keep stack traffic on ds_read/ds_write (LDS) and
// global_load/global_store (overflow); through generic pointers the compiler merges the two
// paths into flat accesses, which are slower for LDS.
#ifdef FHP_HOST_CHECK
typedef uint32_t LdsU32;
typedef uint32_t GlobalU32;
#else
typedef __attribute__((address_space(3))) uint32_t LdsU32;
typedef __attribute__((address_space(1))) uint32_t GlobalU32;
#endif

// A traversal stack of capacity() = ldsDepth + ovfDepth entries. The scene builder sizes it to
// the tree's exact worst case; the overflow part is bounds checked anyway (the LDS part needs no
// check: its index test is the branch that selects it). A push beyond the capacity is dropped and
// sets the scene's trace error word, and a pop beyond it returns the leaf of triangle 0, so a
// stale bound fails loudly (the ray may miss hits) instead of reading or writing out of bounds.
struct TraceStack {
    LdsU32* lds;         // this lane's slot 0; entry i at lds[i * ldsStride]
    uint32_t ldsStride;  // threads per block
    uint32_t ldsDepth;
    GlobalU32* ovf;      // this thread's overflow slot 0; entry i at ovf[i * ovfStride]
    uint32_t ovfStride;
    uint32_t ovfDepth;
    uint32_t* err;       // trace error word (bit 0: stack overflow), or null
#ifdef FHP_HOST_CHECK
    uint32_t peak = 0;   // entries in use at the deepest push
#endif
    FHP_DEVI void put(uint32_t i, uint32_t v) {
#ifdef FHP_HOST_CHECK
        if (i + 1 > peak) peak = i + 1;
#endif
        if (i < ldsDepth) {
            lds[i * ldsStride] = v;
        } else if (i - ldsDepth < ovfDepth) {
            ovf[(i - ldsDepth) * ovfStride] = v;
        } else {
#ifdef FHP_HOST_CHECK
            if (err) *err |= kTraceErrStackOverflow;
#else
            if (err) atomicOr(err, kTraceErrStackOverflow);
#endif
        }
    }
    // Entry i, known to be an LDS entry (i < ldsDepth).
    FHP_DEVI void putLds(uint32_t i, uint32_t v) { lds[i * ldsStride] = v; }
    // Records a stack height reached by putLds writes (the host check build's peak).
    FHP_DEVI void notePeak(uint32_t top) {
#ifdef FHP_HOST_CHECK
        if (top > peak) peak = top;
#else
        (void)top;
#endif
    }
    FHP_DEVI uint32_t get(uint32_t i) const {
        if (i < ldsDepth) return lds[i * ldsStride];
        return i - ldsDepth < ovfDepth ? ovf[(i - ldsDepth) * ovfStride] : kOverflowLeaf;
    }
    FHP_DEVI uint32_t capacity() const { return ldsDepth + ovfDepth; }
};

// Stack of one lane of a persistent traversal grid: LDS entries [entry][lane] in the block's
// array, overflow entries [entry][global thread] in the scene's overflow buffer.
template <int B, int D> FHP_DEVI TraceStack makeStack(uint32_t* lds, const TraceScratch& ts, uint32_t block, uint32_t lane) {
    const uint32_t gtid = block * B + lane;
    return TraceStack{(LdsU32*)(lds + lane), uint32_t(B), uint32_t(D), (GlobalU32*)(ts.ovf + gtid), ts.gridBlocks * B,
                      ts.ovfDepth, ts.err};
}

// The arrays traversal reads: a SceneDev, or this pair when the scene varies per ray.
struct GeomRef {
    const float4* __restrict__ bvh;
    const float4* __restrict__ triPos;
};

struct TraceCounters {
    uint64_t nodes, tris;
};

// Slab test of the 4 children of one node (packBVH4Node's layout: the children's planes decoded
// from 8-bit codes). Returns a 4-bit hit mask and entry distances. Conservative: the decoded boxes
// contain the padded child boxes and the exit distance is scaled by (1 + 4 ulp).
FHP_DEVI uint32_t testNode4(const float4* __restrict__ node, f3 o, f3 inv, float tmin, float tmax,
                                              float tn[4], uint4& child) {
    const float4 n0 = node[0], n1 = node[1], n2 = node[2], cb = node[3];
    child = make_uint4(f2u(cb.x), f2u(cb.y), f2u(cb.z), f2u(cb.w));
    // Planes go straight to ray distance: (o + q s - org) inv = q (s inv) + (o - org) inv. The
    // rounding this adds is far below the build's box padding, so the test stays conservative.
    const float bx0 = (n0.x - o.x) * inv.x, by0 = (n0.y - o.y) * inv.y, bz0 = (n0.z - o.z) * inv.z;
    const float kx = n0.w * inv.x, ky = n1.x * inv.y, kz = n1.y * inv.z;
    // The distance is monotone in the code (fmaf is), non-decreasing where k >= 0: the near plane
    // of every child on an axis is its low plane there and its high plane elsewhere, so the codes
    // are sorted into near and far once per node instead of a min and a max per child and axis.
    const uint32_t qlx = f2u(n1.z), qhx = f2u(n1.w), qly = f2u(n2.x), qhy = f2u(n2.y), qlz = f2u(n2.z), qhz = f2u(n2.w);
    const bool px = kx >= 0.f, py = ky >= 0.f, pz = kz >= 0.f;
    const uint32_t nqx = px ? qlx : qhx, fqx = px ? qhx : qlx, nqy = py ? qly : qhy, fqy = py ? qhy : qly;
    const uint32_t nqz = pz ? qlz : qhz, fqz = pz ? qhz : qlz;
    const uint32_t ch[4] = {child.x, child.y, child.z, child.w};
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t sh = 8u * uint32_t(i);
        const float nx = fmaf(float((nqx >> sh) & 0xffu), kx, bx0), fx = fmaf(float((fqx >> sh) & 0xffu), kx, bx0);
        const float ny = fmaf(float((nqy >> sh) & 0xffu), ky, by0), fy = fmaf(float((fqy >> sh) & 0xffu), ky, by0);
        const float nz = fmaf(float((nqz >> sh) & 0xffu), kz, bz0), fz = fmaf(float((fqz >> sh) & 0xffu), kz, bz0);
        float n = fmaxf(fmaxf(nx, ny), fmaxf(nz, tmin));
        float f = fminf(fminf(fx, fy), fminf(fz, tmax));
        f = f * 1.00000048f;
        tn[i] = n;
        if (n <= f && ch[i] != kInvalid) mask |= 1u << i;
    }
    return mask;
}

// Compare-exchange on (distance, child) pairs: afterwards ta <= tb.
FHP_DEVI void sortPair(float& ta, uint32_t& ca, float& tb, uint32_t& cb) {
    const bool sw = tb < ta;
    const float t = sw ? tb : ta;
    const uint32_t c = sw ? cb : ca;
    tb = sw ? ta : tb;
    cb = sw ? ca : cb;
    ta = t;
    ca = c;
}

FHP_HD f3 safeInv(f3 d) {
    // Direction components below 1e-20 in magnitude are nudged so 1/d stays finite; boxes are
    // padded, so the slab test stays conservative.
    float x = fabsf(d.x) < 1e-20f ? copysignf(1e-20f, d.x) : d.x;
    float y = fabsf(d.y) < 1e-20f ? copysignf(1e-20f, d.y) : d.y;
    float z = fabsf(d.z) < 1e-20f ? copysignf(1e-20f, d.z) : d.z;
    return mk3(1.f / x, 1.f / y, 1.f / z);
}

// Resumable traversal of one ray. Each `step` does one unit of work: at an internal node its
// children are tested and the nearest hit child becomes current (the others are pushed); at a
// leaf one triangle is tested, and after the leaf's last one the next node is popped. One
// triangle per step keeps a wave's lanes in leaves and at internal nodes advancing together
// (a step costs one node test or one triangle test) instead of every lane waiting for the
// largest leaf among them. `step` returns false once the ray is finished (stack empty, or the
// first hit of an any-hit ray).
struct TravState {
    f3 o, d, inv;
    float tmin;
    HitRec hit;
    uint32_t cur, sp;
    uint32_t k;  // next triangle of the current leaf (0 unless cur is a leaf being tested)
    bool any;
    FHP_DEVI void start(f3 org, f3 dir, float t0, float t1, bool anyHit) {
        o = org;
        d = dir;
        inv = safeInv(dir);
        tmin = t0;
        hit.tri = kInvalid;
        hit.t = t1;
        hit.b1 = hit.b2 = 0.f;
        cur = 0;  // root node (always internal)
        sp = 0;
        k = 0;
        any = anyHit;
    }
    template <bool COUNT, typename SceneT>
    FHP_DEVI bool step(const SceneT& s, TraceStack& st, TraceCounters* cnt) {
        if (!childIsLeaf(cur)) {
            const float4* node = s.bvh + 4 * size_t(cur);
            float tn[4];
            uint4 child;
            if (COUNT) cnt->nodes++;
            const uint32_t mask = testNode4(node, o, inv, tmin, hit.t, tn, child);
            if (mask) {
                // Order hit children by entry distance with a 4-input sorting network (missed
                // children sort last at +inf): the nearest becomes current, the others are pushed
                // farthest first. All indices are static, so everything stays in registers.
                float t0 = (mask & 1u) ? tn[0] : kInf, t1 = (mask & 2u) ? tn[1] : kInf;
                float t2 = (mask & 4u) ? tn[2] : kInf, t3 = (mask & 8u) ? tn[3] : kInf;
                uint32_t c0 = child.x, c1 = child.y, c2 = child.z, c3 = child.w;
                sortPair(t0, c0, t1, c1);
                sortPair(t2, c2, t3, c3);
                sortPair(t0, c0, t2, c2);
                sortPair(t1, c1, t3, c3);
                sortPair(t1, c1, t2, c2);
                // The scene builder sizes the stack to the tree's exact worst case
                // (sum of (children - 1) along any root-to-leaf path). The k - 1 pushed entries
                // are p0, p1, c1 cut to their first k - 1; while three entries fit in LDS all
                // three are written without branches (entries above the new top are free).
                const uint32_t k = popc32(mask);
                if (k > 1u) {
                    const uint32_t p0 = k > 3u ? c3 : (k > 2u ? c2 : c1), p1 = k > 3u ? c2 : c1;
                    if (sp + 3u <= st.ldsDepth) {
                        st.putLds(sp, p0);
                        st.putLds(sp + 1u, p1);
                        st.putLds(sp + 2u, c1);
                        st.notePeak(sp + k - 1u);
                    } else {
                        st.put(sp, p0);
                        if (k > 2u) st.put(sp + 1u, p1);
                        if (k > 3u) st.put(sp + 2u, c1);
                    }
                    sp += k - 1u;
                }
                cur = c0;
                return true;
            }
        } else {
            const float4* tp = s.triPos + 3 * size_t(leafFirst(cur) + k);
            const float4 a = tp[0], b = tp[1], c = tp[2];
            const uint32_t id = f2u(a.w);
            float t, b1, b2;
            if (COUNT) cnt->tris++;
            if (watertight(o, d, mk3(a.x, a.y, a.z), mk3(b.x, b.y, b.z), mk3(c.x, c.y, c.z), t, b1, b2)) {
                if (t >= tmin && t <= hit.t && (t < hit.t || hit.tri == kInvalid || id < hit.tri)) {
                    hit.tri = id; hit.t = t; hit.b1 = b1; hit.b2 = b2;
                    if (any) return false;
                }
            }
            if (++k < leafCount(cur)) return true;
            k = 0;
        }
        if (sp == 0) return false;
        cur = st.get(--sp);
        return true;
    }
};

// Closest hit (ANY = false) or first hit (ANY = true) in [tmin, tmax]. Returns true on a hit.
// `anyRt` selects first-hit per lane at run time (kernels that trace both kinds of ray).
template <bool ANY, bool COUNT, typename SceneT>
FHP_DEVI bool traceBVH4(const SceneT& s, f3 o, f3 d, float tmin, float tmax, TraceStack& st,
                                          HitRec& hit, TraceCounters* cnt = nullptr, bool anyRt = false) {
    TravState r;
    r.start(o, d, tmin, tmax, ANY || anyRt);
    while (r.step<COUNT>(s, st, cnt)) {
    }
    hit = r.hit;
    return hit.tri != kInvalid;
}

}  // namespace fhp
