// WARDiffPathTracer geometry gradients (per-path code, shared by the kernel and the host check
// build): warped-area reparameterization (useWAR = 1) at any maxBounces (the reference's shape
// optimization runs at maxBounces = 0, scripts/inv-rendering/shape_optimization/shape_example.py:23).
//
// The path (WARDiffPathTracer.rt.slang:75-143, PTUtils.slang:253-341) is: a screen sample ps with
// Gaussian jitter; the camera ray hits x1 (AttachToRay, DiffSceneQuery.slang:131-219); each vertex
// x_k adds its (MIS-weighted) emission, then, below the bounce limit, NEE toward a light sample
// (computeNEE, PTUtils.slang:211-249) and a BSDF-sampled ray that hits x_k+1, attached to that ray.
// With useWAR the screen sample and every NEE and scatter direction are warped
// (WarpedAreaReparam.slang:199-332): each gets a warp V (zero in value) and a Jacobian
// determinant detJ (one in value) whose parameter derivatives carry the boundary terms. Sampled
// directions, light samples, pdfs and MIS weights are detached as in the reference.
//
// With S = dLdI . radiance, thp0 = detJ_screen G(ps + V_screen - pc) / G^ and thp_k the product
// of the scatter factors detJ_j f_j / pdf_j before x_k,
//   S = thp0 sum_k thp_k [E_k + w_N (a2s/a2s^) (f_N/f_N^) Q_N,k],
// so the adjoint is a reverse sweep over the vertices: BSDF partials (isotropic BSDF: a function
// of N.wi, N.wo and wi.wo), the NEE geometry factor, the warps, the interpolated shading normal,
// the barycentrics and the ray/plane intersection of each vertex, then the screen warp.
//
// Warp adjoint. With aux directions a_i (vMF around the base direction, antithetic pairs), aux
// hit directions d_i (AttachToGeometry: fixed barycentrics, so d_i moves with the aux triangle's
// vertices and the ray origin), harmonic weights w_i and their direction derivatives g_i (the
// custom forward derivative of computeHarmonicWeight), m = sum w_i d_i / sum w_i:
//   dV    = N(m) dm,                 N(m) = (I - u u^T) / |m|, u = m / |m|
//   detJ' = trace of the warp Jacobian's parameter derivative
//         = d/dtheta [ sum_i g_i^T N(m) d_i / sum w_i ]   (the m G^T term vanishes as N(m) m = 0).
// The screen warp is the same construction projected to screen space (computeCameraRayScreenPos),
// which adds the projection's gradient P and Hessian. The closed forms below were checked
// against numerical differentiation of the defining expressions.
#pragma once
#include "fhp_internal.h"
#include "fhp_trace.h"

namespace fhp {
namespace {


// ------------------------------------------------------------------ dual numbers over (ci, co, cio)
struct DF {
    float v, d0, d1, d2;
};
FHP_DEVI DF dconst(float c) { return DF{c, 0.f, 0.f, 0.f}; }
FHP_DEVI DF operator+(DF a, DF b) { return DF{a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2}; }
FHP_DEVI DF operator-(DF a, DF b) { return DF{a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2}; }
FHP_DEVI DF operator*(DF a, DF b) {
    return DF{a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2};
}
FHP_DEVI DF operator*(DF a, float s) { return DF{a.v * s, a.d0 * s, a.d1 * s, a.d2 * s}; }
FHP_DEVI DF operator*(float s, DF a) { return a * s; }
FHP_DEVI DF operator+(DF a, float s) { return DF{a.v + s, a.d0, a.d1, a.d2}; }
FHP_DEVI DF operator+(float s, DF a) { return a + s; }
FHP_DEVI DF operator-(float s, DF a) { return DF{s - a.v, -a.d0, -a.d1, -a.d2}; }
FHP_DEVI DF operator/(DF a, DF b) {
    const float iv = 1.f / b.v, q = a.v * iv;
    return DF{q, (a.d0 - q * b.d0) * iv, (a.d1 - q * b.d1) * iv, (a.d2 - q * b.d2) * iv};
}
FHP_DEVI DF dsqrt(DF a) {
    const float r = sqrtf(a.v), k = r > 0.f ? 0.5f / r : 0.f;
    return DF{r, a.d0 * k, a.d1 * k, a.d2 * k};
}
FHP_DEVI DF dpow5(DF a) {
    const float a2 = a.v * a.v, a4 = a2 * a2, k = 5.f * a4;
    return DF{a4 * a.v, a.d0 * k, a.d1 * k, a.d2 * k};
}
FHP_DEVI DF dlambda(float a2, DF c) {
    // IsotropicGGX.slang:163 evalLambdaGGX: (-1 + sqrt(1 + a2 tan^2)) / 2, tan^2 = (1 - c^2) / c^2
    const DF c2 = c * c;
    const DF t2 = (1.f - c2) / c2;
    return (dsqrt(1.f + a2 * t2) + (-1.f)) * 0.5f;
}

// phi = adj . f(wi, wo) and its partials with respect to ci = N.wi, co = N.wo, cio = wi.wo for
// the reflection BSDFs of StandardMaterial (Frostbite diffuse + GGX, StandardBSDF.slang:228-240)
// and PBRTDiffuseMaterial (Lambert). Returns 0 where the BSDF is zero or invalid.
FHP_DEV float bsdfGeomGrad(const Shading& sd, const Material& m, f3 wo, f3 adj, float& gci, float& gco, float& gcio) {
    gci = gco = gcio = 0.f;
    const f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(wo);
    if (m.type == MAT_PBRT_DIFFUSE) {
        if (!validR(sd, wiL, woL, wo)) return 0.f;
        const float k = dot(adj, m.diffuse) * kInvPi;
        gco = k;
        return k * woL.z;
    }
    if (!validRT(sd, wiL, woL, wo)) return 0.f;
    Bsdf b;
    b.init(wiL, m);
    if (fminf_(wiL.z, woL.z) < kMinCosTheta) return 0.f;
    const DF ci{wiL.z, 1.f, 0.f, 0.f}, co{woL.z, 0.f, 1.f, 0.f}, cio{dot(sd.V, wo), 0.f, 0.f, 1.f};
    DF phi = dconst(0.f);
    if (b.pDiff > 0.f) {
        // (wo.h)^2 = (1 + wi.wo) / 2 for h = normalize(wi + wo)
        const float r = b.rough;
        const DF fd90 = (0.f + (0.5f - 0.f) * r) + (2.f * r) * ((1.f + cio) * 0.5f);
        const DF wiS = 1.f + (fd90 + (-1.f)) * dpow5(1.f - ci);
        const DF woS = 1.f + (fd90 + (-1.f)) * dpow5(1.f - co);
        const float eF = 1.f + (1.f / 1.51f - 1.f) * r;
        phi = phi + (wiS * woS * co) * (dot(adj, b.diff) * eF * kInvPi);
    }
    if (b.pSpec > 0.f && b.alpha > 0.f) {
        const float a2 = b.alpha * b.alpha;
        const DF len = dsqrt(2.f + 2.f * cio);
        const DF hz = (ci + co) / len, wiH = (1.f + cio) / len;
        const DF dd = hz * hz * (a2 - 1.f) + 1.f;
        const DF D = dconst(a2 * kInvPi) / (dd * dd);
        const DF G = dconst(1.f) / (1.f + dlambda(a2, ci) + dlambda(a2, co));
        const DF sfac = D * G * 0.25f / ci;
        const DF w5 = dpow5(1.f - wiH);
        const float as = dot(adj, b.spec), a1 = adj.x + adj.y + adj.z;
        phi = phi + sfac * (1.f - w5) * as + sfac * w5 * a1;
    }
    gci = phi.d0;
    gco = phi.d1;
    gcio = phi.d2;
    return phi.v;
}

// ------------------------------------------------------------------ auxiliary rays
// Orthonormal frame of a direction (ShadingFrame(N, perp_stark(N)), ShadingFrame.slang:45-51).
struct Frame3 {
    f3 T, B, N;
};
FHP_DEVI Frame3 dirFrame(f3 w) {
    Frame3 f;
    f.N = normalize(w);
    const f3 t = perpStark(f.N);
    f.T = normalize(t - f.N * dot(t, f.N));
    f.B = cross(f.N, f.T);
    return f;
}

// Regenerates the aux directions of one traceAsymptoticWeightedMeanIntersection call
// (WarpedAreaReparam.slang:111-193) from a copy of the path's sample generator: aux 2j and 2j + 1
// share the uniform pair (rnd[2j], rnd[2j + 1]), the odd one reflected about the base direction;
// without antithetic sampling the N rnd values are drawn and discarded and each aux draws a pair.
struct AuxGen {
    Rng rng;
    Frame3 f;
    float kappa;
    bool anti;
    float u0, u1;
    FHP_DEV void start(Rng r, f3 w, float k, bool antithetic, uint32_t n) {
        rng = r;
        f = dirFrame(w);
        kappa = k;
        anti = antithetic;
        u0 = u1 = 0.f;
        if (!anti)
            for (uint32_t i = 0; i < n; i++) rng.u1();
    }
    FHP_DEV f3 next(uint32_t i, float& uy) {
        if (!anti || (i & 1u) == 0u) {
            u0 = rng.u1();
            u1 = rng.u1();
        }
        uy = u1;
        // sampleVonMisesFisher (MathHelpers.slang:403-411)
        const float sy = fmaxf_(1.f - u1, 1e-6f);
        const float phi = kTwoPi * u0;
        const float cosT = 1.f + logf(expf(-2.f * kappa) * (1.f - sy) + sy) / kappa;
        const float sinT = sqrtf(fmaxf_(1.f - cosT * cosT, 0.f));
        const f3 dl = mk3(cosf(phi) * sinT, sinf(phi) * sinT, cosT);
        f3 a = normalize(f.T * dl.x + f.B * dl.y + f.N * dl.z);
        if (anti && (i & 1u)) a = a * -1.f + f.N * (2.f * dot(f.N, a));  // reflect(-a, N)
        return a;
    }
};

struct AuxHit {
    bool hit;
    uint32_t tri;
    float b[3];
    f3 p, d;       // hit point and normalized direction from the origin (or the aux direction)
    float w, c;    // harmonic weight and its direction-derivative coefficient
};

// Aux-ray hits of one warp's first loop over its aux rays, read back by its second loop, which
// then traces nothing: entry i is (triangle or kInvalid, b1, b2) at p[(3 i + k) * stride]. The
// second loop regenerates the same aux directions from the same origin, so these are the answers
// a second traversal would give.
struct AuxLog {
    GlobalU32* p;     // this thread's word 0, or null (both loops trace)
    uint32_t stride;  // threads of the persistent grid
    FHP_DEVI void put(uint32_t i, bool hit, const HitRec& h) const {
        p[(3 * i) * stride] = hit ? h.tri : kInvalid;
        p[(3 * i + 1) * stride] = f2u(h.b1);
        p[(3 * i + 2) * stride] = f2u(h.b2);
    }
    FHP_DEVI bool get(uint32_t i, HitRec& h) const {
        h.tri = p[(3 * i) * stride];
        h.b1 = u2f(p[(3 * i + 1) * stride]);
        h.b2 = u2f(p[(3 * i + 2) * stride]);
        return h.tri != kInvalid;
    }
};
FHP_DEVI AuxLog noAuxLog() { return AuxLog{nullptr, 0u}; }
enum AuxMode { kAuxTrace, kAuxRecord, kAuxReplay };

// Traces aux ray i of a warp (AttachToGeometry) and evaluates computeHarmonicWeight and the
// coefficient of its custom forward derivative (WarpedAreaReparam.slang:50-98). With a log, the
// first loop records the traversal's answer (kAuxRecord) and the second reads it (kAuxReplay).
FHP_DEV AuxHit auxTrace(const SceneDev& s, TraceStack& st, f3 o, f3 a, float uy, float kappa, const AuxLog& lg,
                        uint32_t i, AuxMode mode, unsigned long long& nc) {
    AuxHit r;
    HitRec h;
    if (mode == kAuxReplay && lg.p) {
        r.hit = lg.get(i, h);
    } else {
        nc++;
        r.hit = traceBVH4<false, false>(s, o, a, 0.f, kFltMax, st, h);
        if (mode == kAuxRecord && lg.p) lg.put(i, r.hit, h);
    }
    float boundary = 0.f;
    r.d = a;
    r.tri = kInvalid;
    if (r.hit) {
        r.tri = h.tri;
        r.b[1] = h.b1;
        r.b[2] = h.b2;
        r.b[0] = 1.f - h.b1 - h.b2;
        f3 p[3], n[3];
        float tu, tv;
        for (int k = 0; k < 3; k++) loadVertex(s, s.idx[3 * h.tri + k], p[k], n[k], tu, tv);
        r.p = p[0] * r.b[0] + p[1] * r.b[1] + p[2] * r.b[2];
        const f3 nrm = n[0] * r.b[0] + n[1] * r.b[1] + n[2] * r.b[2];  // computeShadingFrame: not normalized
        const f3 v = r.p - o;
        const float l = length(v);
        r.d = l > 0.f ? v / l : mk3(0.f);
        const float dn = dot(nrm, a);
        boundary = dn * dn;
    }
    const float sy = fmaxf_(1.f - uy, 1e-6f);
    const float invVMF = 1.f / ((1.f - sy) * expf(-2.f * kappa) + sy);
    const float wDen = invVMF - 1.f + boundary;
    const float rcp = wDen > 1e-4f ? 1.f / wDen : 0.f;
    r.w = rcp * rcp * rcp * invVMF;
    r.c = invVMF * r.w * rcp * kappa * 3.f;
    return r;
}

// Position / normal gradient of the three vertices of triangle tri, if it belongs to the mesh.
FHP_DEV void addVertexGrad(const SceneDev& s, float* grad, const GeomArgs& a, uint32_t row, uint32_t tri, const float b[3],
                              f3 g) {
    if (!(g.x != 0.f || g.y != 0.f || g.z != 0.f)) return;
    for (int k = 0; k < 3; k++) {
        const uint32_t v = s.idx[3 * tri + k];
        if (v < a.meshV0 || v - a.meshV0 >= a.meshVerts) continue;
        float* q = grad + size_t(row) * 3 * a.meshVerts + 3 * size_t(v - a.meshV0);
        const f3 gk = g * b[k];
        if (gk.x != 0.f) gradAdd(q + 0, gk.x);
        if (gk.y != 0.f) gradAdd(q + 1, gk.y);
        if (gk.z != 0.f) gradAdd(q + 2, gk.z);
    }
}
FHP_DEV void addVertexGrads3(const SceneDev& s, float* grad, const GeomArgs& a, uint32_t row, uint32_t tri, const f3 g[3]) {
    for (int k = 0; k < 3; k++) {
        const uint32_t v = s.idx[3 * tri + k];
        if (v < a.meshV0 || v - a.meshV0 >= a.meshVerts) continue;
        float* q = grad + size_t(row) * 3 * a.meshVerts + 3 * size_t(v - a.meshV0);
        if (g[k].x != 0.f) gradAdd(q + 0, g[k].x);
        if (g[k].y != 0.f) gradAdd(q + 1, g[k].y);
        if (g[k].z != 0.f) gradAdd(q + 2, g[k].z);
    }
}

// The harmonic weight's direction derivative dW = c dot(-w x (w x d), a) is clamped to
// [-1e10, 1e10] (fwd_computeHarmonicWeight, WarpedAreaReparam.slang:93). For the tangents
// d = e_k of a ray warp it is the k-th component of g = c (a (w.w) - w (w.a)); a clamped component
// is constant in the parameters.
constexpr float kDWClamp = 1e10f;
FHP_DEVI float clampDW(float v) { return fminf_(fmaxf_(v, -kDWClamp), kDWClamp); }
FHP_DEVI f3 clampDW3(f3 g) { return mk3(clampDW(g.x), clampDW(g.y), clampDW(g.z)); }

// Adjoint of reparameterizeRay (WarpedAreaReparam.slang:226-263) for a ray (o, w): given dS/dV
// (gV) and dS/d detJ (gDet), adds the aux-triangle vertex gradients and returns dS/do; with
// wantOmega also dS/dw through the harmonic weights' direction derivatives.
FHP_DEV void warpAdjoint(const SceneDev& s, const GeomArgs& a, TraceStack& st, const AuxLog& lg, f3 o, f3 w, Rng rng,
                         f3 gV, float gDet, bool wantOmega, uint32_t row, f3& gO, f3& gW, unsigned long long& nc) {
    gO = mk3(0.f);
    gW = mk3(0.f);
    const float ww = dot(w, w);
    AuxGen gen;
    gen.start(rng, w, a.kappa, a.useAntithetic != 0, a.aux);
    float Sw = 0.f;
    f3 Sd = mk3(0.f);
    f3 Mc0 = mk3(0.f), Mc1 = mk3(0.f), Mc2 = mk3(0.f);  // M = sum d g^T, stored by columns
    for (uint32_t i = 0; i < a.aux; i++) {
        float uy;
        const f3 ai = gen.next(i, uy);
        const AuxHit h = auxTrace(s, st, o, ai, uy, a.kappa, lg, i, kAuxRecord, nc);
        const f3 g = clampDW3((ai * ww - w * dot(w, ai)) * h.c);
        Sw += h.w;
        Sd = Sd + h.d * h.w;
        Mc0 = Mc0 + h.d * g.x;
        Mc1 = Mc1 + h.d * g.y;
        Mc2 = Mc2 + h.d * g.z;
    }
    const float Swc = fmaxf_(Sw, 1e-8f);
    const f3 m = Sd / Swc;
    const float r = length(m);
    if (!(r > 0.f)) return;
    const f3 u = m / r;
    auto Nm = [&](f3 v) { return (v - u * dot(u, v)) / r; };
    const f3 Mu = Mc0 * u.x + Mc1 * u.y + Mc2 * u.z;      // M u
    const f3 MTu = mk3(dot(Mc0, u), dot(Mc1, u), dot(Mc2, u));  // M^T u
    const float trM = Mc0.x + Mc1.y + Mc2.z;
    const float uMu = dot(u, Mu);
    const f3 qV = Nm(gV);
    const f3 common = Nm(Mu + MTu) * (-1.f / r) - u * ((trM - uMu) / (r * r));
    gen.start(rng, w, a.kappa, a.useAntithetic != 0, a.aux);
    for (uint32_t i = 0; i < a.aux; i++) {
        float uy;
        const f3 ai = gen.next(i, uy);
        const AuxHit h = auxTrace(s, st, o, ai, uy, a.kappa, lg, i, kAuxReplay, nc);
        const f3 gu = (ai * ww - w * dot(w, ai)) * h.c;
        const f3 g = clampDW3(gu);
        const float sk = h.w / Swc;
        const f3 dT = Nm(g) + common * sk;
        const f3 gd = qV * sk + dT * (gDet / Swc);
        if (h.hit) {
            const f3 v = h.p - o;
            const float l = length(v);
            if (l > 0.f) {
                const f3 gp = (gd - h.d * dot(h.d, gd)) / l;
                gO = gO - gp;
                addVertexGrad(s, a.gradPos, a, row, h.tri, h.b, gp);
            }
        }
        if (wantOmega) {
            // d/dw of g . qk over the components of g that are not clamped.
            f3 qk = Nm(h.d);
            qk = mk3(fabsf(gu.x) <= kDWClamp ? qk.x : 0.f, fabsf(gu.y) <= kDWClamp ? qk.y : 0.f,
                     fabsf(gu.z) <= kDWClamp ? qk.z : 0.f);
            gW = gW + (w * (2.f * dot(ai, qk)) - ai * dot(w, qk) - qk * dot(w, ai)) * (gDet / Swc * h.c);
        }
    }
}

// Camera ray direction for a screen position (computeCameraRayDirection, DiffSceneQuery.slang:252-261).
FHP_DEVI f3 camRaw(const CameraDev& c, float sx, float sy, uint32_t W, uint32_t H) {
    const float p_x = sx / float(W), p_y = sy / float(H);
    const float ndc_x = 2.f * p_x + -1.f, ndc_y = -2.f * p_y + 1.f;
    return c.U * ndc_x + c.V * ndc_y + c.W;
}

// The camera of view `view` of a batched launch (the scene's camera without batched views).
FHP_DEVI CameraDev geomCamera(const SceneDev& s, const GeomArgs& a, uint32_t view) {
    return a.views ? a.views[view].cam : s.cam;
}

// Adjoint of reparameterizeScreenSample (WarpedAreaReparam.slang:269-332) at screen sample ps:
// gVs = dS/dV (screen units), gDet = dS/d detJ. Adds aux-triangle vertex gradients.
// Inlined into its callers: as a call it costs the kernels a register save and restore around it.
FHP_DEVI void screenWarpAdjoint(const SceneDev& s, const GeomArgs& a, uint32_t view, TraceStack& st, const AuxLog& lg,
                               float psx, float psy, Rng rng, float gVx, float gVy, float gDet, uint32_t row,
                               unsigned long long& nc) {
    const CameraDev cam = geomCamera(s, a, view);
    const f3 o = cam.pos;
    const f3 r0 = camRaw(cam, psx, psy, a.W, a.H);
    const float rl = length(r0);
    if (!(rl > 0.f)) return;
    const f3 w = r0 / rl;
    auto Nw = [&](f3 v) { return (v - w * dot(w, v)) / rl; };
    const f3 cx = Nw(cam.U * (2.f / float(a.W))), cy = Nw(cam.V * (-2.f / float(a.H)));  // columns of d w / d ps
    AuxGen gen;
    gen.start(rng, w, a.kappaScreen, a.useAntithetic != 0, a.aux);
    float Sw = 0.f;
    f3 Sd = mk3(0.f), Mh0 = mk3(0.f), Mh1 = mk3(0.f);  // M_h = sum d h^T (3 x 2)
    for (uint32_t i = 0; i < a.aux; i++) {
        float uy;
        const f3 ai = gen.next(i, uy);
        const AuxHit h = auxTrace(s, st, o, ai, uy, a.kappaScreen, lg, i, kAuxRecord, nc);
        const f3 g = (ai - w * dot(w, ai)) * h.c;
        Sw += h.w;
        Sd = Sd + h.d * h.w;
        Mh0 = Mh0 + h.d * clampDW(dot(cx, g));
        Mh1 = Mh1 + h.d * clampDW(dot(cy, g));
    }
    const float Swc = fmaxf_(Sw, 1e-8f);
    const f3 m = Sd / Swc;
    const float r = length(m);
    if (!(r > 0.f)) return;
    const f3 u = m / r;
    auto Nm = [&](f3 v) { return (v - u * dot(u, v)) / r; };
    // Screen projection proj_c(u) = beta_c (A_c . u) / (W . u) (computeCameraRayScreenPos with an
    // orthogonal camera basis), its gradient rows P_c and Hessians H_c.
    const f3 Wc = cam.W;
    const float WW = dot(Wc, Wc);
    const float bx = float(a.W) * WW / (2.f * dot(cam.U, cam.U));
    const float by = -float(a.H) * WW / (2.f * dot(cam.V, cam.V));
    const float Wu = dot(Wc, u);
    const f3 Px = (cam.U / Wu - Wc * (dot(cam.U, u) / (Wu * Wu))) * bx;
    const f3 Py = (cam.V / Wu - Wc * (dot(cam.V, u) / (Wu * Wu))) * by;
    auto Hx = [&](f3 v) {
        return ((cam.U * dot(Wc, v) + Wc * dot(cam.U, v)) * (-1.f / (Wu * Wu)) + Wc * (2.f * dot(cam.U, u) * dot(Wc, v) / (Wu * Wu * Wu))) * bx;
    };
    auto Hy = [&](f3 v) {
        return ((cam.V * dot(Wc, v) + Wc * dot(cam.V, v)) * (-1.f / (Wu * Wu)) + Wc * (2.f * dot(cam.V, u) * dot(Wc, v) / (Wu * Wu * Wu))) * by;
    };
    const f3 NM0 = Nm(Mh0), NM1 = Nm(Mh1);
    const f3 PtMtu = Px * dot(Mh0, u) + Py * dot(Mh1, u);
    const f3 MhPu = Mh0 * dot(Px, u) + Mh1 * dot(Py, u);
    const float trPMh = dot(Px, Mh0) + dot(Py, Mh1);
    const float uMhPu = dot(u, MhPu);
    const f3 hv = Nm(Hx(NM0) + Hy(NM1));
    const f3 qV = Nm(Px * gVx + Py * gVy);
    const f3 common = Nm(PtMtu + MhPu) * (-1.f / r) - u * ((trPMh - uMhPu) / (r * r)) + hv;
    gen.start(rng, w, a.kappaScreen, a.useAntithetic != 0, a.aux);
    for (uint32_t i = 0; i < a.aux; i++) {
        float uy;
        const f3 ai = gen.next(i, uy);
        const AuxHit h = auxTrace(s, st, o, ai, uy, a.kappaScreen, lg, i, kAuxReplay, nc);
        if (!h.hit) continue;
        const f3 g = (ai - w * dot(w, ai)) * h.c;
        const float hx = clampDW(dot(cx, g)), hy = clampDW(dot(cy, g));
        const float sk = h.w / Swc;
        const f3 dT = Nm(Px * hx + Py * hy) + common * sk;
        const f3 gd = qV * sk + dT * (gDet / Swc);
        const f3 v = h.p - o;
        const float l = length(v);
        if (l > 0.f) addVertexGrad(s, a.gradPos, a, row, h.tri, h.b, (gd - h.d * dot(h.d, gd)) / l);
    }
}

// Gaussian pixel jitter (WARDiffPathTracer.rt.slang:60-63).
FHP_DEVI void gaussianJitter(float ux, float uy, float sigma, float& gx, float& gy) {
    float r = sqrtf(-2.f * logf(ux));
    gx = (r * cosf(kTwoPi * uy)) * sigma;
    gy = (r * sinf(kTwoPi * uy)) * sigma;
}

// One path at maxBounces = 0 (direct illumination, the reference's shape optimization setting):
// the primal replay of tracePaths / tracePath (x1 with its emission, NEE and a BSDF-sampled ray to
// x2's emission) and its reverse sweep in one pass, keeping x1's state in registers. The same
// estimator as geomPath without the vertex records (x2 adds emission only, which the parameters
// do not move).
FHP_DEV void geomPathDirect(const SceneDev& s, const GeomArgs& a, uint32_t view, uint32_t px, uint32_t py, uint32_t seed, f3 dL,
                      uint32_t row, TraceStack& st, const AuxLog& lg, unsigned long long& nc) {
    Rng rng = rngSeed(px, py, seed / 2);
    for (int i = 0; i < 6; i++) rng.u1();
    const float pcx = float(px) + 0.5f, pcy = float(py) + 0.5f;
    float ux = rng.u1(), uy = rng.u1();
    float jx, jy;
    gaussianJitter(ux, uy, 0.5f, jx, jy);
    if (a.useAntithetic && seed % 2 == 0) { jx = -jx; jy = -jy; }
    const float psx = pcx + jx, psy = pcy + jy;
    const Rng rngScreen = rng;  // reparameterizeScreenSample takes the generator by value
    const CameraDev cam0 = geomCamera(s, a, view);
    const f3 r0 = camRaw(cam0, psx, psy, a.W, a.H);
    const float rl = length(r0);
    const f3 w0 = rl > 0.f ? r0 / rl : mk3(0.f);
    const f3 o = cam0.pos;
    HitRec h;
    nc++;
    if (!traceBVH4<false, false>(s, o, w0, 0.f, kFltMax, st, h)) return;

    // AttachToRay (DiffSceneQuery.slang:192-219): ray/plane distance, area barycentrics.
    const uint32_t tri = h.tri;
    f3 v[3], nv[3];
    float tu, tv;
    for (int k = 0; k < 3; k++) loadVertex(s, s.idx[3 * tri + k], v[k], nv[k], tu, tv);
    const f3 e1 = v[1] - v[0], e2 = v[2] - v[0];
    const f3 pvec = cross(w0, e2);
    const float det = dot(e1, pvec);
    const f3 qvec = cross(o - v[0], e1);
    const float t = dot(e2, qvec) / det;
    const f3 p = o + w0 * t;
    const f3 ep0 = v[0] - p, ep1 = v[1] - p, ep2 = v[2] - p;
    const float A0 = length(cross(ep1, ep2)), A1 = length(cross(ep2, ep0)), A2 = length(cross(ep0, ep1));
    const float At = A0 + A1 + A2;
    const float bary[3] = {A0 / At, A1 / At, A2 / At};
    const f3 x1 = v[0] * bary[0] + v[1] * bary[1] + v[2] * bary[2];
    const f3 normalW = nv[0] * bary[0] + nv[1] * bary[1] + nv[2] * bary[2];
    Shading sd = shadingFromHit(s, tri, bary[1], bary[2], w0);
    sd.V = -w0;
    setFrame(sd, normalW);
    const Shading sd0 = sd;
    const Material m = loadMaterial(s, s.mats, s.texels, sd0, sd, nullptr);
    const float nl = length(normalW);
    const f3 nh = normalW / nl;
    const float sgn = dot(sd.N, nh) < 0.f ? -1.f : 1.f;  // double-sided back-face flip

    // Primal contributions and the adjoint of x1's shading (S = Q_E + Q_N + Q_S at the base point).
    const float QE = anyPos(m.emission) ? dot(dL, m.emission) : 0.f;
    float QN = 0.f, QS = 0.f;
    f3 gX = mk3(0.f), gNs = mk3(0.f), gWi = mk3(0.f);
    const f3 faceO = sd.frontFacing ? sd.faceN : -sd.faceN;
    if (a.useNEE) {
        LightSample ls;
        if (sampleLight(s, LS_UNIFORM, x1, normalW, rng, ls)) {
            const f3 Li = ls.pdf > 0.f ? ls.Le / ls.pdf : mk3(0.f);
            const f3 lightPos = rayOriginOffset(ls.posW, ls.normalW);
            const f3 origin = rayOriginOffset(x1, normalW);
            const f3 toL = lightPos - origin;
            const float distance = length(toL);
            const f3 ldir = normalize(toL);
            HitRec sh;
            if (anyPos(Li) && !traceBVH4<true, false>(s, origin, ldir, 0.f, distance, st, sh)) {
                f3 wo = ls.posW - x1;
                const float dist = length(wo);
                wo = wo / dist;
                float misN = 1.f;
                if (a.useMIS) {
                    float sp;
                    if (a.useBSDFSampling || m.type == MAT_PBRT_DIFFUSE) sp = matPdf(sd, m, ldir);
                    else {
                        f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(ldir);
                        sp = (validRT(sd, wiL, woL, ldir) && fminf_(wiL.z, woL.z) >= kMinCosTheta) ? woL.z * kInvPi : 0.f;
                    }
                    misN = misPower(ls.pdf, sp);
                }
                float gci, gco, gcio;
                const float phi = bsdfGeomGrad(sd, m, wo, dL * Li * misN, gci, gco, gcio);
                const float cosL = dot(ls.normalW, -wo);
                if (phi != 0.f && a.neeHeld) {
                    // useSolidAngleNEEWarp: the sampled direction wo is held (a solid-angle sample)
                    // and only the warp moves it, d = wo + V; no area factor. The warp then follows
                    // the light point seen from the moving x1 and detJ carries its solid-angle
                    // change, which is the derivative of the area-measure integrand.
                    QN = phi;
                    gNs = gNs + sd.V * gci + wo * gco;
                    gWi = gWi + sd.N * gci + wo * gcio;
                    const f3 gF = sd.N * gco + sd.V * gcio;
                    f3 gO, gOm;
                    warpAdjoint(s, a, st, lg, rayOriginOffset(x1, faceO), wo, rng, gF, QN, false, row, gO, gOm, nc);
                    gX = gX + gO;
                } else if (phi != 0.f && cosL > 0.f) {
                    // computeNEE (PTUtils.slang:211-249): Q_N = phi(x1, N, d) (a2s / a2s^) detJ with
                    // wo = normalize(y - x1) toward the held light point y, d = wo + V and
                    // a2s = |y - x1|^2 / dot(n_y, -d); the warp's rays start at x1's view-side
                    // offset and its Jacobian's harmonic weights differentiate along wo.
                    QN = phi;
                    gNs = gNs + sd.V * gci + wo * gco;
                    gWi = gWi + sd.N * gci + wo * gcio;
                    const f3 gD = sd.N * gco + sd.V * gcio + ls.normalW * (phi / cosL);  // dS/dd
                    f3 gO, gOm;
                    warpAdjoint(s, a, st, lg, rayOriginOffset(x1, faceO), wo, rng, gD, QN, true, row, gO, gOm, nc);
                    const f3 gWo = gD + gOm;
                    gX = gX + gO - (gWo - wo * dot(wo, gWo)) / dist - wo * (2.f * phi / dist);
                }
            }
        }
    }
    Scatter sc;
    if (matSample(sd, m, rng, sc, a.useBSDFSampling != 0) && sc.pdf > 0.f) {
        const f3 woS = sc.wo;
        const f3 scO = rayOriginOffset(x1, faceO);
        const f3 fS = matEval(sd, m, woS);
        HitRec h2;
        nc++;
        if (anyPos(fS / sc.pdf) && (!a.useNEE || a.useMIS) && traceBVH4<false, false>(s, scO, woS, 0.f, kFltMax, st, h2)) {
            Shading sd2 = shadingFromHit(s, h2.tri, h2.b1, h2.b2, woS);
            const Shading sd20 = sd2;
            const Material m2 = loadMaterial(s, s.mats, s.texels, sd20, sd2, nullptr);
            if (anyPos(m2.emission)) {
                float mis = 1.f;
                if (a.useNEE && a.useMIS) {
                    const float lp = lightPdf(s, LS_UNIFORM, scO, sd.orientedFaceN(), s.triToLight[h2.tri], sd2.posW,
                                              sd2.orientedFaceN());
                    mis = misPower(sc.pdf, lp);
                }
                float gci, gco, gcio;
                const float phi = bsdfGeomGrad(sd, m, woS, dL * m2.emission * (mis / sc.pdf), gci, gco, gcio);
                if (phi != 0.f) {
                    QS = phi;
                    gNs = gNs + sd.V * gci + woS * gco;
                    gWi = gWi + sd.N * gci + woS * gcio;
                    const f3 gDir = sd.N * gco + sd.V * gcio;
                    f3 gO, gOm;
                    warpAdjoint(s, a, st, lg, scO, woS, rng, gDir, QS, false, row, gO, gOm, nc);
                    gX = gX + gO;
                }
            }
        }
    }

    // Shading normal N = sgn normalize(sum b_k n_k) -> vertex normals and barycentrics.
    const f3 gNw = (gNs - nh * dot(nh, gNs)) * (sgn / nl);
    const float gB[3] = {dot(gNw, nv[0]), dot(gNw, nv[1]), dot(gNw, nv[2])};
    {
        const f3 gn[3] = {gNw * bary[0], gNw * bary[1], gNw * bary[2]};
        addVertexGrads3(s, a.gradNrm, a, row, tri, gn);
    }
    // Barycentrics of x1 (b1, b2 solve the 2x2 Gram system of the edges; b0 = 1 - b1 - b2).
    f3 gV[3] = {mk3(0.f), mk3(0.f), mk3(0.f)};
    {
        const float g1 = gB[1] - gB[0], g2 = gB[2] - gB[0];
        const f3 q = x1 - v[0];
        const float a11 = dot(e1, e1), a12 = dot(e1, e2), a22 = dot(e2, e2);
        const float dg = a11 * a22 - a12 * a12;
        if (dg != 0.f) {
            const float gr1 = (a22 * g1 - a12 * g2) / dg, gr2 = (-a12 * g1 + a11 * g2) / dg;
            const float b1 = bary[1], b2 = bary[2];
            const float gA11 = -gr1 * b1, gA12 = -(gr1 * b2 + gr2 * b1), gA22 = -gr2 * b2;
            const f3 gE1 = q * gr1 + e1 * (2.f * gA11) + e2 * gA12;
            const f3 gE2 = q * gr2 + e1 * gA12 + e2 * (2.f * gA22);
            const f3 gQ = e1 * gr1 + e2 * gr2;
            gX = gX + gQ;
            gV[0] = gV[0] - gQ - gE1 - gE2;
            gV[1] = gV[1] + gE1;
            gV[2] = gV[2] + gE2;
        }
    }
    // x1 = o + t w0 with t = nf . (v0 - o) / (nf . w0).
    f3 gW0 = gWi * -1.f;  // wi = -w0
    {
        const f3 nf = cross(e1, e2);
        const float den = dot(nf, w0);
        if (den != 0.f) {
            const float gT = dot(gX, w0);
            gW0 = gW0 + gX * t - nf * (gT * t / den);
            gV[0] = gV[0] + nf * (gT / den);
            const f3 gNf = (v[0] - x1) * (gT / den);
            const f3 gE1 = cross(e2, gNf), gE2 = cross(gNf, e1);
            gV[1] = gV[1] + gE1;
            gV[2] = gV[2] + gE2;
            gV[0] = gV[0] - gE1 - gE2;
        }
    }
    addVertexGrads3(s, a.gradPos, a, row, tri, gV);

    // Screen warp: thp0 = detJ * G(ps + V - pc) / G^ and w0 = camDir(ps + V).
    const float Qt = QE + QN + QS;
    const float inv2s2 = 1.f / (0.5f * 0.5f);
    float gVx = -Qt * (psx - pcx) * inv2s2, gVy = -Qt * (psy - pcy) * inv2s2;
    {
        const f3 q = (gW0 - w0 * dot(w0, gW0)) / rl;
        const CameraDev cam1 = geomCamera(s, a, view);
        gVx += dot(cam1.U, q) * (2.f / float(a.W));
        gVy += dot(cam1.V, q) * (-2.f / float(a.H));
    }
    if (Qt != 0.f || gVx != 0.f || gVy != 0.f) screenWarpAdjoint(s, a, view, st, lg, psx, psy, rngScreen, gVx, gVy, Qt, row, nc);
}

// Per-path vertex records of the forward pass, read back in reverse order by the adjoint sweep.
// Record k (words k * kVtxWords + w at p[(k * kVtxWords + w) * stride]) holds: w0 the triangle
// hit by the incoming ray (bit 31: the vertex's light sample is visible), w1-3 the incoming ray's
// origin and w4-6 its direction (held values), w7-10 the sample generator's state at the vertex,
// w11-13 the path throughput at the vertex and w14-16 the (MIS-weighted) emission the vertex adds
// to the path radiance, so the reverse sweep need not shade a vertex that only adds emission.
constexpr uint32_t kVtxWords = kGeomVertexWords;
constexpr uint32_t kVtxLit = 0x80000000u;
FHP_DEVI bool anyNonZero(f3 v) { return v.x != 0.f || v.y != 0.f || v.z != 0.f; }
struct VertexLog {
    GlobalU32* p;     // this thread's word 0
    uint32_t stride;  // threads of the persistent grid
    uint32_t depth;   // records per thread (maxBounces + 2)
    FHP_DEVI void put(uint32_t k, uint32_t w, uint32_t v) const { p[(k * kVtxWords + w) * stride] = v; }
    FHP_DEVI uint32_t get(uint32_t k, uint32_t w) const { return p[(k * kVtxWords + w) * stride]; }
    FHP_DEVI void put3(uint32_t k, uint32_t w, f3 v) const {
        put(k, w, f2u(v.x));
        put(k, w + 1, f2u(v.y));
        put(k, w + 2, f2u(v.z));
    }
    FHP_DEVI f3 get3(uint32_t k, uint32_t w) const { return mk3(u2f(get(k, w)), u2f(get(k, w + 1)), u2f(get(k, w + 2))); }
};

// AttachToRay (DiffSceneQuery.slang:131-219): the ray/plane distance t of triangle tri, the area
// barycentrics of the point on the ray, the attached point x and the interpolated (unnormalized)
// vertex normal.
struct Attached {
    f3 v[3], n[3], e1, e2, x, nrm;
    float t, b[3];
};
FHP_DEVI void attachRay(const SceneDev& s, uint32_t tri, f3 o, f3 d, Attached& A) {
    float tu, tv;
    for (int k = 0; k < 3; k++) loadVertex(s, s.idx[3 * tri + k], A.v[k], A.n[k], tu, tv);
    A.e1 = A.v[1] - A.v[0];
    A.e2 = A.v[2] - A.v[0];
    const f3 pvec = cross(d, A.e2);
    const float det = dot(A.e1, pvec);
    const f3 qvec = cross(o - A.v[0], A.e1);
    A.t = dot(A.e2, qvec) / det;
    const f3 p = o + d * A.t;
    const f3 ep0 = A.v[0] - p, ep1 = A.v[1] - p, ep2 = A.v[2] - p;
    const float A0 = length(cross(ep1, ep2)), A1 = length(cross(ep2, ep0)), A2 = length(cross(ep0, ep1));
    const float At = A0 + A1 + A2;
    A.b[0] = A0 / At;
    A.b[1] = A1 / At;
    A.b[2] = A2 / At;
    A.x = A.v[0] * A.b[0] + A.v[1] * A.b[1] + A.v[2] * A.b[2];
    A.nrm = A.n[0] * A.b[0] + A.n[1] * A.b[1] + A.n[2] * A.b[2];
}

// NEE at a vertex (generateEmissiveSample + computeNEE, PTUtils.slang:177-249): the MIS weight of
// a visible light sample toward wo (the held light point seen from x).
FHP_DEVI float neeMis(const SceneDev& s, const GeomArgs& a, const Shading& sd, const Material& m, const LightSample& ls,
                      f3 ldir) {
    if (!a.useMIS) return 1.f;
    float sp;
    if (a.useBSDFSampling || m.type == MAT_PBRT_DIFFUSE) sp = matPdf(sd, m, ldir);
    else {
        f3 wiL = sd.toLocal(sd.V), woL = sd.toLocal(ldir);
        sp = (validRT(sd, wiL, woL, ldir) && fminf_(wiL.z, woL.z) >= kMinCosTheta) ? woL.z * kInvPi : 0.f;
    }
    return misPower(ls.pdf, sp);
}

// One path. The forward pass replays the primal (tracePaths / tracePath with handleHit,
// PTUtils.slang:255-341) and records its vertices. The reverse sweep visits them from the last
// to the first. At vertex k it takes R, the radiance the later vertices add, and the adjoints of
// vertex k's outgoing ray origin and direction, handed back by vertex k + 1. It then:
//   * differentiates the NEE term and the scatter factor f_k (dS/df_k = dLdI R / f_k, the path
//     replay identity), with their warps (dS/d detJ = the radiance the warped ray carries);
//   * routes the shading normal through the vertex normals and barycentrics;
//   * differentiates AttachToRay into the triangle's vertices and the incoming ray's origin and
//     direction, which it hands to vertex k - 1. The origin follows the previous vertex with an
//     identity derivative (computeRayOrigin, GeometryHelpers.slang:70-75). The direction is the
//     previous scatter sample plus that vertex's warp.
// The camera ray's direction adjoint ends in the screen warp.
FHP_DEV void geomPath(const SceneDev& s, const GeomArgs& a, uint32_t view, uint32_t px, uint32_t py, uint32_t seed, f3 dL,
                      uint32_t row, TraceStack& st, const AuxLog& lg, const VertexLog& vl, unsigned long long& nc) {
    Rng rng = rngSeed(px, py, seed / 2);
    for (int i = 0; i < 6; i++) rng.u1();
    const float pcx = float(px) + 0.5f, pcy = float(py) + 0.5f;
    float ux = rng.u1(), uy = rng.u1();
    float jx, jy;
    gaussianJitter(ux, uy, 0.5f, jx, jy);
    if (a.useAntithetic && seed % 2 == 0) { jx = -jx; jy = -jy; }
    const float psx = pcx + jx, psy = pcy + jy;
    const Rng rngScreen = rng;  // reparameterizeScreenSample takes the generator by value
    const CameraDev cam0 = geomCamera(s, a, view);
    const f3 r0 = camRaw(cam0, psx, psy, a.W, a.H);
    const float rl = length(r0);
    const f3 w0 = rl > 0.f ? r0 / rl : mk3(0.f);

    // ---------------------------------------------------------------- forward: record the vertices
    f3 rayO = cam0.pos, rayD = w0;
    HitRec h;
    nc++;
    if (!traceBVH4<false, false>(s, rayO, rayD, 0.f, kFltMax, st, h)) return;
    f3 thp = mk3(1.f), prevN = mk3(0.f);
    float prevPdf = 0.f;
    uint32_t nv = 0;
    for (uint32_t len = 0; len < vl.depth; len++) {
        // The light sample and its visibility ray first: only the attached point and normal are
        // live across that traversal (the shading data and material are loaded after it).
        f3 x, nrm;
        {
            Attached A;
            attachRay(s, h.tri, rayO, rayD, A);
            x = A.x;
            nrm = A.nrm;
        }
        vl.put3(len, 1, rayO);
        vl.put3(len, 4, rayD);
        vl.put(len, 7, rng.s0);
        vl.put(len, 8, rng.s1);
        vl.put(len, 9, rng.s2);
        vl.put(len, 10, rng.s3);
        vl.put3(len, 11, thp);
        nv = len + 1;
        const bool stop = len >= a.maxBounces + 1;
        bool lit = false;
        if (!stop && a.useNEE) {
            LightSample ls;
            if (sampleLight(s, LS_UNIFORM, x, nrm, rng, ls)) {
                const f3 Li = ls.pdf > 0.f ? ls.Le / ls.pdf : mk3(0.f);
                const f3 origin = rayOriginOffset(x, nrm);
                const f3 toL = rayOriginOffset(ls.posW, ls.normalW) - origin;
                HitRec sh;
                lit = anyPos(Li) && !traceBVH4<true, false>(s, origin, normalize(toL), 0.f, length(toL), st, sh);
            }
        }
        vl.put(len, 0, h.tri | (lit ? kVtxLit : 0u));
        Shading sd;
        {
            Attached A;
            attachRay(s, h.tri, rayO, rayD, A);
            sd = shadingFromHit(s, h.tri, A.b[1], A.b[2], rayD);
        }
        sd.V = -rayD;
        setFrame(sd, nrm);
        const Shading sd0 = sd;
        const Material m = loadMaterial(s, s.mats, s.texels, sd0, sd, nullptr);
        f3 emitted = mk3(0.f);
        const bool isPrimary = len == 0;
        if ((isPrimary || !a.useNEE || a.useMIS) && anyPos(m.emission)) {
            float misE = 1.f;
            if (!isPrimary && a.useNEE && a.useMIS)
                misE = misPower(prevPdf, lightPdf(s, LS_UNIFORM, rayO, prevN, s.triToLight[h.tri], sd.posW, sd.orientedFaceN()));
            emitted = thp * m.emission * misE;
        }
        vl.put3(len, 14, emitted);
        if (stop) break;
        Scatter sc;
        if (!(matSample(sd, m, rng, sc, a.useBSDFSampling != 0) && sc.pdf > 0.f)) break;
        const f3 fS = matEval(sd, m, sc.wo);
        thp = thp * (fS / sc.pdf);
        prevN = sd.orientedFaceN();
        prevPdf = sc.pdf;
        if (!anyPos(thp)) break;
        rayO = rayOriginOffset(x, sd.frontFacing ? sd.faceN : -sd.faceN);
        rayD = sc.wo;
        nc++;
        if (!traceBVH4<false, false>(s, rayO, rayD, 0.f, kFltMax, st, h)) break;
    }

    // ---------------------------------------------------------------- reverse sweep
    f3 R = mk3(0.f);                     // radiance added after the current vertex
    f3 gOo = mk3(0.f), gOd = mk3(0.f);  // adjoints of the current vertex's outgoing ray origin / direction
    for (int k = int(nv) - 1; k >= 0; k--) {
        const uint32_t w0k = vl.get(uint32_t(k), 0);
        const uint32_t tri = w0k & ~kVtxLit;
        const bool lit = (w0k & kVtxLit) != 0u;
        const f3 o = vl.get3(uint32_t(k), 1), d = vl.get3(uint32_t(k), 4);
        Rng r{vl.get(uint32_t(k), 7), vl.get(uint32_t(k), 8), vl.get(uint32_t(k), 9), vl.get(uint32_t(k), 10)};
        f3 contrib = vl.get3(uint32_t(k), 14);
        if (uint32_t(k) >= a.maxBounces + 1) {
            // A terminating vertex only adds its emission, which the parameters do not move.
            R = R + contrib;
            gOo = gOd = mk3(0.f);
            continue;
        }
        const f3 thpk = vl.get3(uint32_t(k), 11);
        // Local terms first (BSDF partials of the NEE and scatter factors, the shading normal's
        // routing into vertex normals and barycentrics), then the warps, then the attachment of
        // x to its incoming ray: the shading data and material are dead before the warps' aux
        // traversals, which keeps the kernel's spill traffic down.
        f3 gX = gOo, gWi = mk3(0.f), gTri[3] = {mk3(0.f), mk3(0.f), mk3(0.f)};
        f3 x, faceO;
        // NEE warp (origin x's view-side offset, base direction wo; the reference form also takes
        // the direction derivative of its harmonic weights).
        bool neeWarp = false, neeRef = false;
        f3 neeWo = mk3(0.f), neeG = mk3(0.f);
        float neePhi = 0.f, neeDist = 1.f;
        Rng rN = r;
        // Scatter warp (same origin, base direction the BSDF sample).
        bool scWarp = false;
        f3 scWo = mk3(0.f), scG = mk3(0.f);
        float scDet = 0.f;
        Rng rS = r;
        {
            Attached A;
            attachRay(s, tri, o, d, A);
            x = A.x;
            Shading sd = shadingFromHit(s, tri, A.b[1], A.b[2], d);
            sd.V = -d;
            setFrame(sd, A.nrm);
            const Shading sd0 = sd;
            const Material m = loadMaterial(s, s.mats, s.texels, sd0, sd, nullptr);
            faceO = sd.frontFacing ? sd.faceN : -sd.faceN;
            f3 gNs = mk3(0.f);
            if (a.useNEE) {
                LightSample ls;
                if (sampleLight(s, LS_UNIFORM, A.x, A.nrm, r, ls) && lit) {
                    const f3 Li = ls.pdf > 0.f ? ls.Le / ls.pdf : mk3(0.f);
                    const f3 origin = rayOriginOffset(A.x, A.nrm);
                    const f3 ldir = normalize(rayOriginOffset(ls.posW, ls.normalW) - origin);
                    f3 wo = ls.posW - A.x;
                    const float dist = length(wo);
                    wo = wo / dist;
                    const float misN = neeMis(s, a, sd, m, ls, ldir);
                    contrib = contrib + thpk * (Li * matEval(sd, m, wo) * misN);
                    float gci, gco, gcio;
                    const float phi = bsdfGeomGrad(sd, m, wo, dL * thpk * Li * misN, gci, gco, gcio);
                    const float cosL = dot(ls.normalW, -wo);
                    if (phi != 0.f && (a.neeHeld || cosL > 0.f)) {
                        // useSolidAngleNEEWarp: the sampled direction wo is held (a solid-angle
                        // sample) and only the warp moves it, d = wo + V; no area factor. The
                        // warp then follows the light point seen from the moving vertex and detJ
                        // carries its solid-angle change, the derivative of the area-measure
                        // integrand. Otherwise computeNEE (PTUtils.slang:211-249): Q_N =
                        // phi(x, N, d) (a2s / a2s^) detJ with wo = normalize(y - x) toward the
                        // held light point y, d = wo + V and a2s = |y - x|^2 / dot(n_y, -d).
                        gNs = gNs + sd.V * gci + wo * gco;
                        gWi = gWi + sd.N * gci + wo * gcio;
                        neeWarp = true;
                        neeRef = !a.neeHeld;
                        neeWo = wo;
                        neePhi = phi;
                        neeDist = dist;
                        neeG = sd.N * gco + sd.V * gcio + (neeRef ? ls.normalW * (phi / cosL) : mk3(0.f));
                        rN = r;
                    }
                }
            }
            Scatter sc;
            if (matSample(sd, m, r, sc, a.useBSDFSampling != 0) && sc.pdf > 0.f) {
                // The scatter factor f_k and the warp of the outgoing ray: dS/df_k = dLdI R / f_k,
                // dS/d detJ_k = dLdI . R, and the later vertices' adjoint of the outgoing direction.
                const f3 woS = sc.wo;
                const f3 fS = matEval(sd, m, woS);
                const f3 adj = mk3(fS.x != 0.f ? dL.x * R.x / fS.x : 0.f, fS.y != 0.f ? dL.y * R.y / fS.y : 0.f,
                                   fS.z != 0.f ? dL.z * R.z / fS.z : 0.f);
                float gci, gco, gcio;
                const float phi = bsdfGeomGrad(sd, m, woS, adj, gci, gco, gcio);
                gNs = gNs + sd.V * gci + woS * gco;
                gWi = gWi + sd.N * gci + woS * gcio;
                scG = sd.N * gco + sd.V * gcio + gOd;
                scDet = dot(dL, R);
                scWo = woS;
                scWarp = phi != 0.f || scDet != 0.f || anyNonZero(scG);
                rS = r;
            }
            // Shading normal N = sgn normalize(sum b_k n_k) -> vertex normals and barycentrics.
            const float nl = length(A.nrm);
            const f3 nh = A.nrm / nl;
            const float sgn = dot(sd.N, nh) < 0.f ? -1.f : 1.f;  // double-sided back-face flip
            const f3 gNw = (gNs - nh * dot(nh, gNs)) * (sgn / nl);
            const float gB[3] = {dot(gNw, A.n[0]), dot(gNw, A.n[1]), dot(gNw, A.n[2])};
            {
                const f3 gn[3] = {gNw * A.b[0], gNw * A.b[1], gNw * A.b[2]};
                addVertexGrads3(s, a.gradNrm, a, row, tri, gn);
            }
            // Barycentrics of x (b1, b2 solve the 2x2 Gram system of the edges; b0 = 1 - b1 - b2).
            const float g1 = gB[1] - gB[0], g2 = gB[2] - gB[0];
            const f3 q = A.x - A.v[0];
            const float a11 = dot(A.e1, A.e1), a12 = dot(A.e1, A.e2), a22 = dot(A.e2, A.e2);
            const float dg = a11 * a22 - a12 * a12;
            if (dg != 0.f) {
                const float gr1 = (a22 * g1 - a12 * g2) / dg, gr2 = (-a12 * g1 + a11 * g2) / dg;
                const float b1 = A.b[1], b2 = A.b[2];
                const float gA11 = -gr1 * b1, gA12 = -(gr1 * b2 + gr2 * b1), gA22 = -gr2 * b2;
                const f3 gE1 = q * gr1 + A.e1 * (2.f * gA11) + A.e2 * gA12;
                const f3 gE2 = q * gr2 + A.e1 * gA12 + A.e2 * (2.f * gA22);
                const f3 gQ = A.e1 * gr1 + A.e2 * gr2;
                gX = gX + gQ;
                gTri[0] = gTri[0] - gQ - gE1 - gE2;
                gTri[1] = gTri[1] + gE1;
                gTri[2] = gTri[2] + gE2;
            }
        }
        R = R + contrib;
        const f3 wOrigin = rayOriginOffset(x, faceO);
        if (neeWarp) {
            f3 gO, gOm;
            warpAdjoint(s, a, st, lg, wOrigin, neeWo, rN, neeG, neePhi, neeRef, row, gO, gOm, nc);
            gX = gX + gO;
            if (neeRef) {
                const f3 gWo = neeG + gOm;
                gX = gX - (gWo - neeWo * dot(neeWo, gWo)) / neeDist - neeWo * (2.f * neePhi / neeDist);
            }
        }
        if (scWarp) {
            f3 gO, gOm;
            warpAdjoint(s, a, st, lg, wOrigin, scWo, rS, scG, scDet, false, row, gO, gOm, nc);
            gX = gX + gO;
        }
        // x = o + t d with t = nf . (v0 - o) / (nf . d): into the triangle's vertices, the incoming
        // direction (also the view direction wi = -d of the shading) and the incoming origin.
        f3 gD = gWi * -1.f, gO = mk3(0.f);
        {
            Attached A;
            attachRay(s, tri, o, d, A);
            const f3 nf = cross(A.e1, A.e2);
            const float den = dot(nf, d);
            if (den != 0.f) {
                const float gT = dot(gX, d);
                gD = gD + gX * A.t - nf * (gT * A.t / den);
                gO = gX - nf * (gT / den);
                gTri[0] = gTri[0] + nf * (gT / den);
                const f3 gNf = (A.v[0] - A.x) * (gT / den);
                const f3 gE1 = cross(A.e2, gNf), gE2 = cross(gNf, A.e1);
                gTri[1] = gTri[1] + gE1;
                gTri[2] = gTri[2] + gE2;
                gTri[0] = gTri[0] - gE1 - gE2;
            }
        }
        addVertexGrads3(s, a.gradPos, a, row, tri, gTri);
        gOo = gO;
        gOd = gD;
    }

    // Screen warp: thp0 = detJ * G(ps + V - pc) / G^ and w0 = camDir(ps + V); gOd is now the
    // camera direction's adjoint and R the path's radiance.
    const float Qt = dot(dL, R);
    const float inv2s2 = 1.f / (0.5f * 0.5f);
    float gVx = -Qt * (psx - pcx) * inv2s2, gVy = -Qt * (psy - pcy) * inv2s2;
    {
        const f3 q = (gOd - w0 * dot(w0, gOd)) / rl;
        gVx += dot(cam0.U, q) * (2.f / float(a.W));
        gVy += dot(cam0.V, q) * (-2.f / float(a.H));
    }
    if (Qt != 0.f || gVx != 0.f || gVy != 0.f) screenWarpAdjoint(s, a, view, st, lg, psx, psy, rngScreen, gVx, gVy, Qt, row, nc);
}

// One item of the kernel's grid loop: path id = (view * W * H + pixel) * spp + sample.
// DIRECT: the maxBounces = 0 specialisation (geomPathDirect), otherwise the general reverse sweep.
template <bool DIRECT>
FHP_DEVI void geomItem(const SceneDev& s, const GeomArgs& a, uint32_t id, TraceStack& st, unsigned long long& nc,
                       const AuxLog& lg, const VertexLog& vl) {
    uint32_t view = 0, local = id, frame = a.frame;
    const float* dLdI = a.dLdI;
    if (a.views) {
        const uint32_t per = a.W * a.H * a.spp;
        view = id / per;
        local = id - view * per;
        frame = a.views[view].frame;
        dLdI = a.views[view].dLdI;
    }
    const uint32_t pix = local / a.spp, k = local % a.spp;
    f3 dL = mk3(1.f, 0.f, 0.f);  // dLdI null: the BackwardDiffDebug mode's dL/dI = (1, 0, 0)
    if (dLdI) {
        const float* dl = dLdI + 3 * size_t(pix);
        dL = mk3(dl[0], dl[1], dl[2]);
    }
    if (dL.x == 0.f && dL.y == 0.f && dL.z == 0.f) return;
    if (DIRECT) geomPathDirect(s, a, view, pix % a.W, pix / a.W, frame + k, dL, jenkins(pix) % a.hashSize, st, lg, nc);
    else geomPath(s, a, view, pix % a.W, pix / a.W, frame + k, dL, jenkins(pix) % a.hashSize, st, lg, vl, nc);
}

}  // namespace
}  // namespace fhp
