/* fhp: MI355X-native path-tracing core behind Falcor's render-pass interface.
 *
 * C ABI of libfhp.so. Plain pointers and sizes only. "d_" pointers are HIP device pointers
 * (for example torch-ROCm tensor data); "h_" pointers are host memory. Every call returns 0 on
 * success or a negative FHP_E* code; fhp_last_error() gives the message of the last failure
 * on the calling thread. Handles are not thread-safe; work is issued on the stream passed in
 * (0 = the legacy default stream) and no call synchronises the host unless stated.
 *
 * Each entry point names the reference interface it replaces (paths relative to the Falcor
 * source tree).
 */
#ifndef FHP_H
#define FHP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FHP_OK 0
#define FHP_EINVAL -1   /* bad argument (null pointer, size mismatch, unknown pass kind) */
#define FHP_EHIP -2     /* a HIP runtime call failed */
#define FHP_ESTATE -3   /* handle in the wrong state (no scene, buffers missing) */
#define FHP_EBUILD -4   /* scene build failed (empty scene, limits exceeded) */

typedef struct FhpScene_* FhpScene;
typedef struct FhpPass_* FhpPass;
typedef struct FhpGrads_* FhpGrads;
typedef struct FhpSession_* FhpSession;
typedef void* FhpStream; /* hipStream_t */

/* Library version (major * 10000 + minor * 100 + patch). */
int fhp_get_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* fhp_last_error(void);
/* Message of the last warning (for example an unknown property), "" if none. */
const char* fhp_last_warning(void);

/* ---------------------------------------------------------------- session
 * A device and a stream for callers without their own (torch callers pass torch's current
 * stream instead). Selects device_ordinal (hipSetDevice) and creates a non-blocking stream of
 * default priority; every FhpStream argument accepts it. Replaces the Device / RenderContext a
 * Falcor host hands to its passes (Source/Falcor/Core/API/Device.h, RenderContext.h). */
int fhp_session_create(int device_ordinal, FhpSession* out);
int fhp_session_stream(FhpSession session, FhpStream* out);
int fhp_session_destroy(FhpSession session);

/* ---------------------------------------------------------------- scene
 * Scene description: world-space triangle geometry (instances flattened by the caller).
 * Materials are 20 floats each: [type (0 StandardMaterial, 1 PBRTDiffuseMaterial),
 * baseColor rgb, roughness, metallic, ior, emissive rgb, emissiveFactor, doubleSided,
 * baseColor texture index (-1 none), emissive texture index + 1 (0 none: the emissive colour is
 * used; MaterialSystem.slang:201-213), specularTransmission, diffuseTransmission, transmission
 * colour rgb, thinSurface] (BasicMaterialData.slang:50-66; a material with transmission is
 * double-sided, BasicMaterial.cpp:662-673). Textures are RGB fp32, texDesc = 4 ints
 * (float offset into texels, width, height, 0). Camera = 14 floats [position, target, up,
 * fovY (radians), aspect, focalDistance, nearZ, farZ].
 * Replaces Scene::create + SceneBuilder (Source/Falcor/Scene/Scene.cpp, SceneBuilder.cpp),
 * the BLAS/TLAS builds (Scene.cpp:3376, :3936), LightCollection::build
 * (Source/Falcor/Scene/Lights/LightCollection.cpp:220) and LightBVHBuilder::build
 * (Source/Falcor/Rendering/Lights/LightBVHBuilder.cpp:228). */
typedef struct FhpSceneDesc {
    uint32_t num_vertices;
    const float* h_positions; /* 3 * num_vertices */
    const float* h_normals;   /* 3 * num_vertices */
    const float* h_texcoords; /* 2 * num_vertices */
    uint32_t num_triangles;
    const uint32_t* h_indices;     /* 3 * num_triangles */
    const uint32_t* h_tri_material; /* num_triangles */
    uint32_t num_materials;
    const float* h_materials; /* 20 * num_materials */
    uint32_t num_textures;
    const int32_t* h_tex_desc; /* 4 * num_textures */
    uint64_t num_texels_floats;
    const float* h_texels;
    const float* h_camera; /* 14 floats */
    /* Meshes (optional): mesh m owns vertices [offsets[m], offsets[m + 1]) and the triangles
     * indexing them (SceneBuilder's mesh ids, Scene.cpp:4314-4336). 0 = one mesh. */
    uint32_t num_meshes;
    const uint32_t* h_mesh_vertex_offsets; /* num_meshes + 1 */
    /* FHP_SCENE_* flags: FHP_SCENE_USE_COMPRESSED_HIT_INFO selects the 64-bit PackedHitInfo when
     * the header fits (SceneBuilder::Flags::UseCompressedHitInfo, SceneBuilder.h:78). */
    uint32_t flags;
} FhpSceneDesc;
#define FHP_SCENE_USE_COMPRESSED_HIT_INFO 0x8000u

typedef struct FhpSceneInfo {
    uint32_t num_triangles, num_vertices, num_materials, num_textures;
    uint32_t bvh_nodes, bvh_leaves, bvh_depth, stack_need;
    uint32_t num_lights, num_active_lights, light_bvh_nodes, num_meshes;
    double bvh_sah_cost;
    uint64_t device_bytes;
    float camera_basis[9]; /* U, V, W as computed by Camera.cpp:187-193 */
    /* PackedHitInfo format (HitInfo::init, Source/Falcor/Scene/HitInfo.cpp:52-106): 1 = 64-bit
     * records (2 words per pixel), 0 = 128-bit (4 words); bits of the type, instance id (mesh)
     * and primitive index (triangle within the mesh). meshes_contiguous = 1 when each mesh's
     * triangles are consecutive, which V-buffer encoding and decoding need. */
    uint32_t hit_info_compressed, hit_info_type_bits, hit_info_instance_bits, hit_info_primitive_bits;
    uint32_t meshes_contiguous;
} FhpSceneInfo;

int fhp_scene_create(const FhpSceneDesc* desc, FhpScene* out);
/* Scene mutators that take host arrays (fhp_scene_destroy, fhp_scene_update_vertices,
 * fhp_bvh_build, fhp_scene_set_materials, fhp_scene_set_material_params, fhp_scene_set_env_map,
 * fhp_scene_set_analytic_lights) are synchronous: they wait for all queued work of the device,
 * on every stream, before they overwrite or free buffers that kernels read. */
int fhp_scene_destroy(FhpScene scene);
int fhp_scene_info(FhpScene scene, FhpSceneInfo* out);
/* Replace the scene's camera (14 floats, the layout of FhpSceneDesc.h_camera: position, target,
 * up, vertical FOV in radians, aspect, focal distance, near and far plane). Later launches use
 * it; the camera reaches the kernels by value, so work already queued is unaffected. Replaces
 * the camera updates of Scene::getCamera() (Camera::setPosition / setTarget / setUpVector /
 * setFocalLength / setAspectRatio / setDepthRange, Camera.cpp:112-193). */
int fhp_scene_set_camera(FhpScene scene, const float* h_camera);
/* Trace error bits accumulated by the scene's traversal kernels since the last reset (bit 0:
 * a ray needed more stack entries than the tree's worst case; that ray ended without a hit).
 * Zero when the scene's acceleration structure and its launch bounds agree. Synchronises the
 * stream; reset = 1 clears the bits. No counterpart in the reference (a driver-built
 * acceleration structure reports no such error). */
int fhp_scene_trace_errors(FhpScene scene, uint32_t* h_bits, int reset, FhpStream stream);
/* Replace the vertex positions (and normals, if h_normals is not null) of mesh mesh_id from
 * host memory (3 floats per vertex of the mesh), then rebuild the BVH, the light collection and
 * the light BVH. Replaces Scene::setMeshVertices / set_mesh_vertices (Scene.cpp:4314-4336) and
 * the BLAS rebuild it triggers. */
int fhp_scene_update_vertices(FhpScene scene, uint32_t mesh_id, const float* h_positions, const float* h_normals);
/* The same update from device memory. rebuild = 0 refits the BVH on the GPU (the reference's
 * default BLAS update mode, Scene.h:1366 UpdateMode::Refit): vertex and triangle records are
 * rewritten and every node box recomputed bottom-up, keeping the tree's topology. rebuild = 1,
 * or a mesh with emissive triangles (whose lights move), rebuilds the BVH, the light collection
 * and the light BVH as fhp_scene_update_vertices does. Copies the vertices to host memory
 * (synchronises the stream). */
int fhp_scene_update_vertices_device(FhpScene scene, uint32_t mesh_id, const float* d_positions, const float* d_normals,
                                     int rebuild, FhpStream stream);
/* Move mesh mesh_id as an animated scene-graph node: its vertices become M * rest position and
 * M^-T * rest normal, where the rest pose is the mesh as created or as last set by a vertex
 * update and h_matrix is the 3x4 row-major affine M (determinant > 0; the transform is not
 * cumulative). On the GPU: the BVH is refitted, every emissive triangle's vertices, face normal
 * and area are recomputed (flux kept), and the light BVH is refitted, leaves first, then internal
 * nodes from the deepest level up. Synchronises the stream (host mirrors).
 * Replaces Scene::updateNodeTransform (Scene.cpp:4282-4289) and the animation update it starts:
 * LightCollection::update / updateTrianglePositions (LightCollection.cpp:78-114, :550-574,
 * UpdateTriangleVertices.cs.slang) and LightBVHSampler::update -> LightBVH::refit
 * (LightBVHSampler.cpp:36-75, LightBVH.cpp:49-89, LightBVHRefit.cs.slang). */
int fhp_scene_update_node_transform(FhpScene scene, uint32_t mesh_id, const float* h_matrix, FhpStream stream);
/* Current vertex positions and normals of mesh mesh_id (3 floats per vertex, host memory; either
 * pointer may be null). Replaces Scene::getMeshVerticesAndIndices (Scene.cpp:4291-4312). */
int fhp_scene_get_mesh_vertices(FhpScene scene, uint32_t mesh_id, float* h_positions, float* h_normals);
/* Rebuild the BVH4 (leaves of at most max_leaf_triangles triangles, 1..8; 0 = 4), the light
 * collection and the light BVH from the scene's current vertices. Replaces the BLAS/TLAS builds
 * (Scene::buildBlas / buildTlas, Scene.cpp:3376, :3936-4040); creation and vertex updates build
 * the same structures by themselves. */
int fhp_bvh_build(FhpScene scene, uint32_t max_leaf_triangles);
/* Serialized StandardMaterial parameters of materials h_ids[0..n): 20 floats each in
 * StandardMaterialParamLayout (StandardMaterialParamLayout.slang:32-41: baseColor 0-2,
 * metallic 3, roughness 4, ior 5, transmissionColor 6-8, diffuseTransmission 9,
 * specularTransmission 10, emissiveColor 11-13, emissiveFactor 14, 15-19 zero). Other material
 * types are refused. Replaces Scene get_material_params / Material::serializeParams
 * (Scene.cpp:4426-4447, StandardMaterial.cpp:234-238). */
int fhp_scene_get_material_params(FhpScene scene, uint32_t n, const uint32_t* h_ids, float* h_params);
/* The inverse: each value is clamped to its layout range (MaterialParamLayout.slangh:40-45:
 * [1e-4, 1 - 1e-4] by default, roughness >= 0.05, ior [1e-3, 1e3], transmissions >= 0,
 * emissiveFactor [1e-8, 1e8]) and rounded to float16 where the reference stores float16
 * (colours, metallic, roughness, ior, transmissions); type, double-sided, textures and thin
 * surface are kept. Replaces Scene set_material_params (Scene.cpp:4453-4476). */
int fhp_scene_set_material_params(FhpScene scene, uint32_t n, const uint32_t* h_ids, const float* h_params);
/* Copy all material records into host memory (20 * num_materials floats). */
int fhp_scene_get_materials(FhpScene scene, float* h_materials);
/* Replace all material records (host memory, 20 * num_materials floats).
 * Replaces Scene::setMaterialParams (Source/Falcor/Scene/Scene.cpp:4620-4643). */
int fhp_scene_set_materials(FhpScene scene, const float* h_materials);
/* Copy texels from device memory into the scene (same layout as the description). */
int fhp_scene_set_texels(FhpScene scene, const float* d_texels, uint64_t num_floats, FhpStream stream);
/* Finite-difference support: enable = 1 snapshots the current materials and texels as the
 * sampling parameters of WARDiffPathTracer (lobe selection, direction sampling, pdfs and MIS),
 * so later parameter changes only move BSDF values and emission, the quantities the adjoint
 * differentiates (sampling is detached in StandardBSDF.slang:155-178 and PTUtils.slang:136-249).
 * enable = 0 returns to normal rendering. No counterpart in the reference. */
int fhp_scene_freeze_sampling(FhpScene scene, int enable, FhpStream stream);
/* Environment map (EnvMap, Source/Falcor/Scene/Lights/EnvMap.cpp / EnvMap.slang): a lat-long
 * RGB fp32 map of height * width texels (h_texels, row 0 at the top: v = acos(y) / pi), h_params =
 * 7 floats [intensity, tint rgb, rotation in degrees xyz], flags 1 = light the scene (rays that
 * miss it pick up its radiance, MinimalPathTracer.rt.slang:475-486), 2 = background of primary
 * misses (:462-465), i.e. Scene::useEnvLight / useEnvBackground. width = 0 removes the map. */
int fhp_scene_set_env_map(FhpScene scene, const float* h_texels, uint32_t width, uint32_t height, const float* h_params,
                          uint32_t flags);
/* EnvMapSampler (Source/Falcor/Rendering/Lights/EnvMapSampler.cpp:39-115, EnvMapSampler.slang):
 * fhp_scene_set_env_map builds the hierarchical importance map (512 x 512 equal-area octahedral
 * luminance, 64 samples per texel, 2 x 2 means down to 1 x 1). get_env_importance writes its
 * dimension (0: no map, or one whose luminance integrates to zero) and, when h_levels is not null,
 * all levels coarse to fine ((4 dim^2 - 1) / 3 floats, the r x r level at (r^2 - 1) / 3). */
int fhp_scene_get_env_importance(FhpScene scene, uint32_t* dim, float* h_levels);
/* EnvMapSampler::sample for n pairs d_u (2 floats each, in [0,1)): d_out 8 floats per sample
 * (world direction, pdf per solid angle, Le, valid). */
int fhp_env_sample(FhpScene scene, const float* d_u, uint32_t n, float* d_out, FhpStream stream);
/* EnvMapSampler::evalPdf for n world directions d_dirs (3 floats each): d_pdf n floats. */
int fhp_env_eval_pdf(FhpScene scene, const float* d_dirs, uint32_t n, float* d_pdf, FhpStream stream);
/* Analytic lights (Scene::addLight with PointLight / DirectionalLight, Source/Falcor/Scene/Lights/
 * Light.cpp), n records of 16 floats: [type (0 point, 1 directional), intensity rgb, position xyz,
 * openingAngle, direction xyz, penumbraAngle, 0, 0, 0, 0]. Directions are normalised, the opening
 * angle clamped to [0, pi] and the penumbra to [0, opening] (Light.cpp:174-223). MinimalPathTracer
 * samples one of them uniformly at every vertex with a shadow ray (evalDirectAnalytic,
 * MinimalPathTracer.rt.slang:253-286; LightHelpers.slang:190-258). n = 0 removes them. */
int fhp_scene_set_analytic_lights(FhpScene scene, uint32_t n, const float* h_lights);
/* Light collection export (for checks): num_lights records of 20 floats [p0 xyz, area, p1 xyz,
 * flux, p2 xyz, triangle id (bits), normal xyz, 0, Le rgb (a textured emitter: emissiveFactor),
 * emissive texture + 1 (0: constant)]. The flux of a textured emitter comes from the
 * coverage-weighted texel average of EmissiveIntegrator.3d.slang / FinalizeIntegration.cs.slang. */
int fhp_scene_lights(FhpScene scene, float* h_lights);
/* Light BVH export (for checks): nodes 16 floats each, leaf light lists, per-light bitmasks
 * (2 words each). Sizes from fhp_scene_info. */
int fhp_scene_light_bvh(FhpScene scene, float* h_nodes, uint32_t* h_tri, uint32_t* h_masks);

/* ---------------------------------------------------------------- ray queries
 * Rays: 8 floats each (origin xyz, tmin, direction xyz, tmax). Hits: 4 words each
 * (triangle id, t, b1, b2) with b1, b2 the barycentric weights of vertices 1 and 2
 * (HitInfo.slang:76-79); a miss is (0xffffffff, 0, 0, 0). Closest hit replaces TraceRay with ray flags 0
 * (Source/RenderPasses/MinimalPathTracer/MinimalPathTracer.rt.slang:233-243) and
 * SceneRayQuery::traceRay (Source/Falcor/Scene/RaytracingInline.slang:86-250); any hit
 * replaces the RAY_FLAG_ACCEPT_FIRST_HIT_AND_END_SEARCH shadow query (:205-226) and
 * traceSceneVisibilityRay (RaytracingInline.slang:253-347): occluded[i] = 1 if any triangle
 * is hit in [tmin, tmax]. */
int fhp_trace_closest(FhpScene scene, const float* d_rays, uint32_t num_rays, uint32_t* d_hits, FhpStream stream);
int fhp_trace_any(FhpScene scene, const float* d_rays, uint32_t num_rays, uint32_t* d_occluded, FhpStream stream);

/* ---------------------------------------------------------------- render passes
 * kind: "MinimalPathTracer", "WARDiffPathTracer", "AccumulatePass" or "VBufferRT". Properties are (key, value) pairs with
 * numeric values; enumerations are passed as their integer value. Unknown keys are ignored
 * with a warning (fhp_last_warning), as the reference's parseProperties does.
 * Replaces registerPlugin/RenderPass::create (MinimalPathTracer.cpp:32-35, :66-89;
 * WARDiffPathTracer.cpp:31-35, :72-144). */
int fhp_pass_create(const char* kind, int num_props, const char* const* keys, const double* values, FhpPass* out);
int fhp_pass_destroy(FhpPass pass);
int fhp_pass_set_property(FhpPass pass, const char* key, double value);
int fhp_pass_get_property(FhpPass pass, const char* key, double* value);

/* One frame of the pass into d_color (height * width * 4 floats, RGBA, row-major).
 * spp samples per pixel are rendered in one launch sequence and averaged as
 * AccumulatePass Single mode does (Accumulate.cs.slang:76-114); sample k of frame f uses
 * sample number f * spp + k (MinimalPathTracer) or frame seed f * spp + k (WARDiffPathTracer).
 * Replaces RenderPass::execute (MinimalPathTracer.cpp:110-199, WARDiffPathTracer.cpp:221-...).
 * stats (optional, 8 words): closest-hit rays, visibility rays, then BVH nodes visited and
 * triangles tested by closest-hit and by visibility queries (these four only when the pass
 * property "countTraversal" is 1; MinimalPathTracer only), then the closest-hit and visibility
 * rays of the first two words that the bounce loop's tail launch traced (MinimalPathTracer
 * property "tailPaths"; zero otherwise). */
int fhp_pass_execute(FhpPass pass, FhpScene scene, uint32_t width, uint32_t height, uint32_t spp, uint32_t frame,
                     float* d_color, uint64_t* h_stats, FhpStream stream);

/* The fhp_pass_execute frame of a MinimalPathTracer pass without any device allocation: no
 * counterpart in the reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The frame's wavefront state lives in a
 * static device array of the library (about 124 MB per device, held until the process ends), so
 * a frame holds at most 524288 paths (width * height * spp) and 524288 pixels; larger frames are
 * refused with FHP_EINVAL. Output, seeds and properties are those of fhp_pass_execute (the bounce
 * loop is launched directly, not through a captured graph); no statistics, no traversal counting,
 * no kernel timing. Calls on one device share that device's array: calls that overlap in time on
 * the same device (any streams) race and must be ordered by the caller. */
int fhp_pass_execute_noalloc(FhpPass pass, FhpScene scene, uint32_t width, uint32_t height, uint32_t spp, uint32_t frame,
                             float* d_color, FhpStream stream);
/* A batch of views rendered in one launch sequence. MinimalPathTracer: every view's paths share
 * the wavefront queues and the persistent traversal grid, each ray traversing its own view's
 * scene. WARDiffPathTracer: the views share one scene, width, height and spp (frames and cameras
 * differ) and run in one primal launch (Primal and BackwardDiff modes, and ForwardDiffDebug
 * without a debug variable; BackwardDiffDebug clears the outputs; gradient images are refused).
 * h_camera is 14 floats as in FhpSceneDesc, or NULL for the scene's camera; each view is
 * resolved into its d_color (height * width * 4 floats) exactly as fhp_pass_execute would render
 * it. No counterpart in the reference (one graph execution per view there,
 * RenderGraph::execute, Source/Falcor/RenderGraph/RenderGraph.cpp). MinimalPathTracer: at most
 * 65535 views; with useGraph the bounce loop is captured once per batch layout (view count,
 * path and pixel totals, traversal launch) and replayed with each call's view table, which is
 * staged through pinned memory, so the call does not wait for the batch (only for the previous
 * call's table copy). WARDiffPathTracer synchronises the stream before returning. stats
 * (optional) as fhp_pass_execute's, summed over the views. */
typedef struct FhpView {
    FhpScene scene;
    const float* h_camera;
    uint32_t width, height, spp, frame;
    float* d_color;
} FhpView;
int fhp_pass_execute_batch(FhpPass pass, uint32_t num_views, const FhpView* views, uint64_t* h_stats, FhpStream stream);

/* Accumulated kernel times (ms) and launch counts by kernel kind, recorded with HIP events on
 * the launch stream when the pass property "timeKernels" is 1: [0] closest-hit traversal,
 * [1] shadow traversal, [2] shading, [3] WAR primal, [4] WAR adjoint, [5] other (the bounce loop's
 * tail launch).
 * reset != 0 clears the accumulators. Replaces the GPU timer events of FALCOR_PROFILE
 * (Source/Falcor/Utils/Timing/Profiler.h:305). */
int fhp_pass_kernel_times(FhpPass pass, double* ms, uint64_t* launches, int reset);

/* ---------------------------------------------------------------- gradients
 * Replaces SceneGradients (Source/Falcor/DiffRendering/SceneGradients.cpp:42-206).
 * Material gradients: 20 floats per material (StandardMaterialParamLayout.slang:32-41),
 * accumulated in hash_size replicas selected by jenkinsHash(pixelID) % hash_size
 * (GradientIOWrapper.slang:39-42) and summed by fhp_grads_aggregate. Texture gradients
 * (per-texel baseColor, an extension) share the texel layout of the scene. */
int fhp_grads_create(FhpScene scene, uint32_t material_hash_size, int with_texture_grads, FhpGrads* out);
int fhp_grads_destroy(FhpGrads grads);
int fhp_grads_clear(FhpGrads grads, FhpStream stream);
int fhp_grads_aggregate(FhpGrads grads, FhpStream stream);
/* Device pointers and sizes (in floats) of the aggregated material gradients and of the
 * texture gradients (NULL / 0 when disabled). */
int fhp_grads_get(FhpGrads grads, float** d_material, uint64_t* material_floats, float** d_texture,
                  uint64_t* texture_floats);
/* Enables the MeshPosition and MeshNormal gradients of mesh mesh_id (3 floats per vertex of
 * the mesh each, same hash size as the material gradients); written by fhp_pass_backward when
 * the pass property useWAR is 1. Replaces GradConfig(MeshPosition / MeshNormal, dim = 3 *
 * vertex count) of SceneGradients::create (SceneGradients.cpp:42-102, shape_example.py:92-111)
 * together with WARDiffPathTracer::setMeshToOptimize (WARDiffPathTracer.cpp:697-715). */
int fhp_grads_set_mesh(FhpGrads grads, uint32_t mesh_id);
/* Copies aggregated gradients into d_dst: which = 0 material, 1 texture, 2 mesh position,
 * 3 mesh normal (num_floats must equal the gradient's size).
 * Replaces SceneGradients::getGradsBuffer + Buffer::copy_to_torch (common.py:103-110). */
int fhp_grads_copy(FhpGrads grads, int which, float* d_dst, uint64_t num_floats, FhpStream stream);

/* Adjoint of WARDiffPathTracer's image: accumulates d(sum_pixels dLdI . I)/d(params) into
 * grads by path replay (recomputes each path from its seed; no path storage). d_dLdI is
 * height * width * 3 floats. Frame/sample numbering as in fhp_pass_execute. Needs diffMode
 * BackwardDiff (1), or BackwardDiffDebug (3), where dLdI = (1, 0, 0) at every pixel and d_dLdI
 * is ignored (WARDiffPathTracer.rt.slang:201-211); other modes are refused.
 * With useWAR = 1 the vertex gradients of the mesh enabled on grads are accumulated too, for
 * maxBounces up to 8 and scenes without transmissive materials (other cases are refused).
 * Extension property useHitLog (default 1): the replay reads the first pass's traversal answers
 * from a per-pass hit log instead of tracing again; 0 traces again (same gradients).
 * Replaces rayGen's bwd_diff(tracePaths) branch (WARDiffPathTracer.rt.slang:161-180) and
 * the host loop of render_grad (scripts/inv-rendering/common.py:90-114). */
int fhp_pass_backward(FhpPass pass, FhpScene scene, uint32_t width, uint32_t height, uint32_t spp, uint32_t frame,
                      const float* d_dLdI, FhpGrads grads, uint64_t* h_stats, FhpStream stream);
/* fhp_pass_backward for a batch of views of one scene and one frame shape in one launch
 * sequence (the material adjoint and, with useWAR, the vertex gradients): view v contributes the
 * gradient of its own loss with dL/dI d_dLdI (height * width * 3 floats; ignored in
 * BackwardDiffDebug mode), seeds frame * spp + k (or fixedSeed) and camera h_camera (14 floats as
 * in FhpSceneDesc, or NULL for the scene's camera); all views accumulate into the same gradient
 * rows. Equal to the sum of fhp_pass_backward over the views up to the order of the float
 * atomics. No counterpart in the reference (one runBackward per view,
 * WARDiffPathTracer.cpp:221-...). Synchronises the stream before its launches (the view table is
 * staged from host memory). */
typedef struct FhpDiffView {
    const float* h_camera;
    uint32_t frame;
    const float* d_dLdI;
} FhpDiffView;
int fhp_pass_backward_batch(FhpPass pass, FhpScene scene, uint32_t num_views, const FhpDiffView* views, uint32_t width,
                            uint32_t height, uint32_t spp, FhpGrads grads, uint64_t* h_stats, FhpStream stream);

/* AccumulatePass: a pass of kind "AccumulatePass" (fhp_pass_create) with the reference's
 * properties enabled (1), autoReset (1), precisionMode (0 Double, 1 Single [default],
 * 2 SingleCompensated), maxFrameCount (0 = no limit; above 0 the Single and Double modes
 * average with moving-average updates), overflowMode (0 Stop, 1 Reset, 2 EMA), and the
 * deprecated alias enableAccumulation. The output is RGBA32F at the input's size (the
 * reference's outputFormat / outputSize / fixedOutputSize are not properties here).
 * fhp_accumulate adds one frame d_input (height * width RGBA texels: input_format 0 float32,
 * 1 uint32, 2 int32) into the running average and writes it to d_output (height * width * 4
 * floats). refresh != 0 reports a change that restarts accumulation when autoReset is 1 (the
 * render-pass refresh flags or scene updates of AccumulatePass.cpp:129-155). With
 * overflowMode Stop and maxFrameCount frames accumulated, d_output is left untouched.
 * Replaces AccumulatePass::execute / accumulate / reset (AccumulatePass.cpp:127-292, :375-378)
 * and the accumulate* kernels (Accumulate.cs.slang:76-190). */
int fhp_accumulate(FhpPass pass, const void* d_input, int input_format, uint32_t width, uint32_t height, float* d_output,
                   uint32_t refresh, FhpStream stream);
int fhp_accumulate_reset(FhpPass pass);
/* Frames accumulated so far (mFrameCount, shown by AccumulatePass::renderUI, :346). */
int fhp_accumulate_frame_count(FhpPass pass, uint32_t* count);

/* ImageCompare: error of image b against image a (height * width RGBA floats each, device),
 * metric 0 mse, 1 rmse (relative), 2 mae, 3 mape, over RGB or RGBA (alpha != 0): the mean over
 * pixels of each pixel's mean channel error, returned in *h_error. d_error_map (optional,
 * height * width floats) receives the per-pixel errors; d_heat_map (optional, needs the error
 * map, height * width * 4 floats) their colour-ramp heat map. Synchronises the stream.
 * Replaces the metrics, compare and generateHeatMap of Source/Tools/ImageCompare/ImageCompare.cpp
 * (:191-270); file loading and the threshold test are the caller's (falcor_amd.image_compare). */
int fhp_image_compare(int metric, const float* d_a, const float* d_b, uint32_t width, uint32_t height, int alpha,
                      float* d_error_map, float* d_heat_map, double* h_error, FhpStream stream);

/* VBufferRT: a pass of kind "VBufferRT" (properties samplePattern [0 Center, 1 DirectX, 2 Halton,
 * 3 Stratified], sampleCount, useAlphaTest, adjustShadingNormals, forceCullMode, cull [forced
 * Front/Back culling is refused], useTraceRayInline, useDOF, and the traversal launch
 * properties). Each fhp_vbuffer_execute call is one frame: it takes the next camera jitter
 * sample of the pattern (scaled by 1 / frame size; returned in h_jitter[2] when not null), traces
 * the pinhole camera's jittered primary ray of every pixel and writes d_vbuffer: one PackedHitInfo per pixel
 * (2 or 4 words, hit_info_compressed of fhp_scene_info; a miss is all zeros), and when not null
 * d_depth (NDC depth z/w of the hit, 1 on a miss), d_viewW (4 floats: -ray direction, 0) and
 * d_mask (1 hit, 0 miss). The instance id of a hit is its mesh, the primitive index its triangle
 * within the mesh. Replaces VBufferRT::execute and its ray generation / hit / miss programs
 * (Source/RenderPasses/GBuffer/VBuffer/VBufferRT.cpp, VBufferRT.rt.slang, VBufferRT.slang:56-143)
 * with TriangleHit::pack (Source/Falcor/Scene/HitInfo.slang:94-121, :457-473). */
int fhp_vbuffer_execute(FhpPass pass, FhpScene scene, uint32_t width, uint32_t height, uint32_t* d_vbuffer, float* d_depth,
                        float* d_viewW, float* d_mask, float* h_jitter, FhpStream stream);
/* Binds (d_data != NULL) or unbinds an input channel of a MinimalPathTracer for later
 * fhp_pass_execute calls: "vbuffer" (PackedHitInfo per pixel in the scene's format, as written
 * by fhp_vbuffer_execute; its hits replace the primary traversal) or "viewW" (4 floats per pixel;
 * the primary direction is -viewW). The buffers must hold width * height records at execution.
 * Records that do not name a triangle of the scene read as misses. Replaces the "vbuffer" and
 * "viewW" inputs of MinimalPathTracer (MinimalPathTracer.cpp:48-53, MinimalPathTracer.rt.slang). */
/* Bind an output channel of a WARDiffPathTracer pass (WARDiffPathTracer.cpp:46-51): "dColor",
 * height * width * 4 floats (RGBA), written by fhp_pass_execute in the gradient-image debug modes
 * (NULL unbinds). ForwardDiffDebug writes (d red / d theta . v, 0, 0, 1) per pixel, averaged over
 * the frame's samples, for the debug variable set below; samples whose derivative exceeds the
 * property assertThreshold (1e9) in magnitude contribute zero (WARDiffPathTracer.rt.slang:182-199);
 * with no debug variable it is (0, 0, 0, 1); BackwardDiffDebug clears it (and color). */
int fhp_pass_set_output(FhpPass pass, const char* channel, void* d_data);
/* The gradient-image debug variable of a WARDiffPathTracer pass (setDiffDebugParams,
 * WARDiffPathTracer.cpp:689-695; DiffDebugParams, SharedTypes.slang:98-104): var_type 0 None,
 * 1 Material (parameter entry starting at `offset` of material `id`'s 20-float layout, direction
 * h_grad[0..3]), 2 GeometryTranslation (mesh `id` translated along h_grad[0..2]). The property
 * diffVarName (1 CBOX_BUNNY_MATERIAL, 2 CBOX_BUNNY_TRANSLATION) sets the reference's presets
 * (WARDiffPathTracer.cpp:84-93). ForwardDiffDebug images are produced for Material variables;
 * a GeometryTranslation variable is refused there. */
int fhp_pass_set_diff_debug(FhpPass pass, int var_type, uint32_t id, uint32_t offset, const float* h_grad);
int fhp_pass_set_input(FhpPass pass, const char* channel, const void* d_data);

#ifdef __cplusplus
}
#endif
#endif /* FHP_H */
