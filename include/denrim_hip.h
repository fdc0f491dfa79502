/*
 * denrim_hip.h -- C ABI of the MI355X (gfx950) backend of the progressive
 * path tracer.
 *
 * This is the boundary a Swift `hip` acceleration mode binds through a
 * module map. Every entry point replaces one call of the
 * reference's public session API or its internal acceleration plug point;
 * the reference interface each one stands in for is cited beside it (paths
 * relative to Sources/DenrimRendererKit/ of markusmoenig/DenrimRendererKit).
 *
 * Conventions
 *   - Plain C types only; all arrays are caller-owned host memory that the
 *     call copies (nothing is retained after it returns).
 *   - Scene records use the reference's GPU layouts byte for byte
 *     (Metal/GPUTypes.swift): GPUTriangle 192 B, GPUMaterial 272 B,
 *     GPULightRecord 32 B, GPUEnvironmentSample 8 B, GPUTextureDescriptor
 *     16 B, texels RGBA32F, GPUCamera 80 B. They are exactly the arrays of
 *     AccelerationBuild (Acceleration/AccelerationBuild.swift:4-44).
 *   - Every function returns a dr_status; on failure a thread-local message
 *     is available from dr_last_error().
 *   - A session is single-caller, like RenderSession (no internal locks).
 */
#ifndef DENRIM_HIP_H
#define DENRIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* DenrimRendererError cases (Renderer/DenrimRenderer.swift:5-26) plus the
 * argument and memory errors a C caller can hit. */
typedef enum {
    DR_OK = 0,
    DR_ERR_NO_DEVICE = 1,              /* .noMetalDevice */
    DR_ERR_MISSING_SHADER_FUNCTION = 2,/* .missingShaderFunction */
    DR_ERR_INVALID_SCENE = 3,          /* .invalidScene */
    DR_ERR_COMMAND_BUFFER_FAILED = 4,  /* .commandBufferFailed */
    DR_ERR_PNG_EXPORT_FAILED = 5,      /* .pngExportFailed */
    DR_ERR_INVALID_ARGUMENT = 6,
    DR_ERR_OUT_OF_MEMORY = 7
} dr_status;

/* RenderOutput (Renderer/RenderOutput.swift:4-25) in case order, plus the
 * internal triangle-ID plane used for parity checks (the global materialized
 * triangle index + 1, stored as uint32 bits in all four channels). */
typedef enum {
    DR_OUTPUT_BEAUTY = 0,
    DR_OUTPUT_DEPTH = 1,
    DR_OUTPUT_NORMAL = 2,
    DR_OUTPUT_ALBEDO = 3,
    DR_OUTPUT_MATERIAL_ID = 4,
    DR_OUTPUT_OBJECT_ID = 5,
    DR_OUTPUT_MOTION_VECTOR = 6,
    DR_OUTPUT_TRIANGLE_ID = 7
} dr_output;

typedef struct dr_device dr_device;
typedef struct dr_scene dr_scene;
typedef struct dr_session dr_session;

/* GPUCamera (Metal/GPUTypes.swift:4-10): origin (w = 1 orthographic),
 * lowerLeft, horizontal, vertical, lens (x = aperture radius, y = focus). */
typedef struct {
    float origin[4];
    float lower_left[4];
    float horizontal[4];
    float vertical[4];
    float lens[4];
} dr_gpu_camera;

/* The arrays of AccelerationBuild for the triangle path
 * (Acceleration/AccelerationBuild.swift:4-44) plus the environment fields
 * the session copies into GPURenderConstants (Renderer/RenderSession.swift:908-947). */
typedef struct {
    const void *triangles;            uint32_t triangle_count;           /* GPUTriangle[] */
    const void *materials;            uint32_t material_count;           /* GPUMaterial[] */
    const void *lights;               uint32_t light_count;              /* GPULightRecord[] */
    const void *environment_samples;  uint32_t environment_sample_count; /* GPUEnvironmentSample[] */
    const void *texture_descriptors;  uint32_t texture_count;            /* GPUTextureDescriptor[] */
    const void *texture_pixels;       uint64_t texture_pixel_count;      /* SIMD4<Float>[] */
    uint32_t environment_texture_index_plus_one;
    float environment_intensity;
    float environment_rotation_y;
    float environment_max_radiance;
    uint32_t max_leaf_triangles;      /* 0 = default (4) */
    /* Signed-distance volumes: the volume arrays of AccelerationBuild
     * (AccelerationBuild.swift:5-20) in the reference layouts
     * (Metal/GPUTypes.swift: GPUVolumeDescriptor 240 B, dense
     * GPUVolumeSample 80 B, GPUVolumeBrickDescriptor 128 B,
     * GPUVolumeBrickSample 16 B, GPUVolumeBrickGrid 64 B, grid + macro-grid
     * indices u32). Brick distances are stored as FP16 on the device when
     * every one of them is FP16-exact, else as FP32; scenes with volumes may
     * hold at most 65536 materials. */
    const void *volumes;              uint32_t volume_count;
    const void *volume_samples;       uint32_t volume_sample_count;
    const void *volume_bricks;        uint32_t volume_brick_count;
    const void *volume_brick_samples; uint32_t volume_brick_sample_count;
    const void *volume_brick_grids;   uint32_t volume_brick_grid_count;
    const uint32_t *volume_brick_grid_indices; uint32_t volume_brick_grid_index_count;
} dr_scene_desc;

/* RenderSettings (Renderer/RenderSettings.swift:29-101) as the session
 * resolves it, plus the sample seed and batching knob of this backend. */
typedef struct {
    uint32_t width, height;
    uint32_t max_bounces;             /* clamped to >= 1 like RenderSession.swift:916 */
    uint32_t quality;                 /* 0 preview, 1 interactive, 2 final: constants.renderQuality (RenderSettings.swift:106-115) */
    float sample_radiance_clamp;      /* < 0: the quality default (10 / 24 / 64) */
    uint32_t transparent_background;
    uint32_t shows_environment_background;
    float background_color[3];
    uint32_t denoiser_enabled;        /* AOVs from the pixel-centre ray (PathTrace.metal:5531-5557) */
    uint32_t collects_stats;          /* collectsSDFTraversalStats: query / node / triangle counters */
    uint32_t frame_seed;              /* 0x1234ABCD in the reference (RenderSession.swift:917) */
    uint32_t samples_per_launch;      /* samples of one pixel in flight together; 0 = auto: a session
                                       * keeps about 64 M paths in flight (at most 64 samples), a
                                       * server request about 4 M (at most its sample count) */
    /* DenoiseSettings (Denoising/Denoiser.swift:20-75): the beauty output is
     * denoised on read when `denoiser` is not DR_DENOISER_NONE (which also
     * switches the AOVs to the pixel-centre ray, RenderSession.swift:941). */
    uint32_t denoiser;                /* dr_denoiser */
    uint32_t denoise_radius;          /* clamped to 1...2 (Denoiser.swift:144) */
    uint32_t denoise_iterations;      /* clamped to 1...5 (Denoiser.swift:122) */
    float denoise_normal_sigma, denoise_depth_sigma, denoise_albedo_sigma, denoise_color_sigma;
    uint32_t intersection;            /* dr_intersection */
    uint32_t environment_sampling;    /* dr_environment_sampling */
} dr_settings;

/* How NEE picks an HDRI texel. The reference searches the CDF
 * (Resources/Shaders/PathTrace.metal:3929-3950); the alias table (Walker,
 * Vose) draws the same texel distribution in O(1) with one more random
 * number per sample, so its images differ sample by sample and agree in
 * the limit (judged on converged relMSE). */
typedef enum {
    DR_ENV_SAMPLING_CDF = 0,
    DR_ENV_SAMPLING_ALIAS = 1,
} dr_environment_sampling;

/* Ray/triangle test. The reference's is Moller-Trumbore
 * (Resources/Shaders/PathTrace.metal:1801-1852), which can let a ray slip
 * through an edge shared by two triangles; the watertight test (Woop,
 * Benthin and Wald, JCGT 2013) cannot. Both accept t > 0.0005 and give the
 * same u, v convention, so shading is unchanged. */
typedef enum {
    DR_INTERSECTION_MOLLER_TRUMBORE = 0,
    DR_INTERSECTION_WATERTIGHT = 1,
} dr_intersection;

/* RenderDenoiser (Denoising/Denoiser.swift:8-18) plus this backend's
 * learned U-Net. Apple SVGF (a closed Apple library) is refused with
 * DR_ERR_INVALID_SCENE as on platforms without MPS (RenderSession.swift:575-576). */
typedef enum {
    DR_DENOISER_NONE = 0,
    DR_DENOISER_SIMPLE_SPATIAL = 1,   /* simpleSpatialDenoiseKernel (Resources/Shaders/Denoise.metal:39-115) */
    DR_DENOISER_APPLE_SVGF = 2,
    DR_DENOISER_LEARNED = 3,          /* the U-Net of a dr_unet set with dr_hip_session_set_denoiser_network */
} dr_denoiser;

/* Counters read by dr_hip_session_stats (the reference's primary, bounce
 * and shadow scene-query counters, PathTrace.metal:201-203, plus traversal
 * work). */
typedef enum {
    DR_STAT_PRIMARY_QUERIES = 0,
    DR_STAT_BOUNCE_QUERIES = 1,
    DR_STAT_SHADOW_QUERIES = 2,
    DR_STAT_NODE_VISITS = 3,          /* extension queries: node fetches (4 child boxes each) */
    DR_STAT_TRIANGLE_TESTS = 4,
    DR_STAT_SHADOW_NODE_VISITS = 5,   /* NEE shadow queries */
    DR_STAT_SHADOW_TRIANGLE_TESTS = 6,
    DR_STAT_SUBSURFACE_QUERIES = 7,   /* random-walk queries, also in BOUNCE_QUERIES */
    /* 8-17: the SDF counters sdfCounterDense* / sdfCounterSparse*
     * (PathTrace.metal:191-200, slot k at 8 + k) for extension, subsurface
     * and AOV queries. */
    DR_STAT_DENSE_VOLUME_TESTS = 8,
    DR_STAT_DENSE_MARCH_STEPS = 9,
    DR_STAT_SPARSE_GRID_CELLS_VISITED = 10,
    DR_STAT_SPARSE_GRID_MACRO_SKIPS = 11,
    DR_STAT_SPARSE_BRICK_TESTS = 12,
    DR_STAT_SPARSE_BRICK_INVALID = 13,
    DR_STAT_SPARSE_BRICK_RANGE_CULLS = 14,
    DR_STAT_SPARSE_BRICK_MARCHES = 15,
    DR_STAT_SPARSE_BRICK_MARCH_STEPS = 16,
    DR_STAT_SPARSE_BRICK_HITS = 17,
    /* 18-27: the same ten counters for NEE shadow queries only; the
     * reference's counters are the sums of both groups. */
    DR_STAT_SHADOW_SDF_BASE = 18,
    DR_STAT_SUBSURFACE_NODE_VISITS = 28, /* subsurface-walk queries */
    DR_STAT_SUBSURFACE_TRIANGLE_TESTS = 29,
    /* 30-32: the part of the shadow counters (2, 5, 6) traced by the
     * subsurface-walk kernel's launch (shadow rays and walks share it). */
    DR_STAT_SHADOW_QUERIES_IN_WALK_KERNEL = 30,
    DR_STAT_SHADOW_NODE_VISITS_IN_WALK_KERNEL = 31,
    DR_STAT_SHADOW_TRIANGLE_TESTS_IN_WALK_KERNEL = 32,
    /* 33-35: the part traced by the extension kernel's launch (the previous
     * bounce's shadow rays of paths whose subsurface walk failed). */
    DR_STAT_SHADOW_QUERIES_IN_EXTEND_KERNEL = 33,
    DR_STAT_SHADOW_NODE_VISITS_IN_EXTEND_KERNEL = 34,
    DR_STAT_SHADOW_TRIANGLE_TESTS_IN_EXTEND_KERNEL = 35,
    /* every scene query (extension, shadow-chain step, subsurface-walk
     * segment), counted in every run, counting or not */
    DR_STAT_RAY_QUERIES = 36,
    DR_STAT_COUNT = 37
} dr_stat;

/* SDF primitive for the volume builders (SDFPrimitive / CompiledPrimitive,
 * Scene/SDFModel.swift:4-70, 664-713; the torus is an addition). The shape
 * is evaluated at world_to_primitive * p - center and scaled by
 * distance_scale when has_transform is set. */
enum { DR_SDF_SPHERE = 0, DR_SDF_TORUS = 1, DR_SDF_BOX = 2, DR_SDF_CYLINDER = 3 };
enum { DR_SDF_UNION = 0, DR_SDF_SUBTRACT = 1 };
typedef struct {
    uint32_t kind;                 /* DR_SDF_* */
    uint32_t material;
    float center[3];               /* shape origin in primitive space */
    float radius;                  /* sphere radius / torus major radius / cylinder radius */
    float axis[3];                 /* torus axis */
    float minor_radius;            /* torus tube radius */
    float half_extents[3];         /* box */
    float corner_radius;           /* box */
    float half_height;             /* cylinder (along local y) */
    float smooth_radius;           /* smoothUnionRadius; < 0: the build settings' radius */
    uint32_t operation;            /* DR_SDF_UNION or DR_SDF_SUBTRACT */
    uint32_t has_transform;
    float world_to_primitive[16];  /* column-major inverse of the primitive transform */
    float distance_scale;          /* smallest axis scale of the transform (SDFModel.swift:708-713) */
    float lipschitz;               /* bound on |grad d| of the scaled distance (1 without shear) */
} dr_sdf_primitive;

/* DistanceVolumeBuildSettings / SparseDistanceVolumeBuildSettings
 * (Scene/SDFModel.swift:256-275). */
typedef struct {
    uint32_t dims[3];       /* sample grid */
    uint32_t brick_size;    /* sparse: core samples per brick side */
    uint32_t apron;         /* sparse: stored samples beyond the core per side (the reference uses 2) */
    float narrow_band;      /* sparse: keep bricks whose samples reach |d| <= band */
    float smooth_radius;    /* default smooth-union radius */
    float bounds_min[3], bounds_max[3];
    uint32_t quantize;      /* 1: FP16 distances, blends on a 1/65535 grid, exact skipping */
} dr_sdf_build_settings;

typedef struct dr_sdf_volume dr_sdf_volume;

/* Render server (no reference counterpart: the reference renders one
 * session at a time on the caller's thread). */
typedef struct dr_server dr_server;
typedef struct {
    uint64_t memory_budget_bytes;  /* session memory of requests on the GPU at once */
    uint32_t max_active;           /* requests on the GPU at once */
    uint32_t samples_per_batch;    /* samples a request encodes per scheduling step; 0 = all */
    uint32_t use_graphs;           /* replay batch graphs (dr_hip_session_set_graphs) */
    uint32_t admission;            /* DR_ADMIT_ARRIVAL or DR_ADMIT_LEAST_WORK */
} dr_server_settings;
/* Order in which queued requests are admitted: arrival order, or least work
 * first (width x height x samples, ties in arrival order), which lets short
 * previews pass long renders in the queue. */
enum { DR_ADMIT_ARRIVAL = 0, DR_ADMIT_LEAST_WORK = 1 };
typedef enum { DR_REQ_QUEUED = 0, DR_REQ_RUNNING = 1, DR_REQ_DONE = 2, DR_REQ_FAILED = 3 } dr_request_state;

/* Thread-local message of the last failing call ("" if none). */
const char *dr_last_error(void);

/* DenrimRenderer.init(device:) (Renderer/DenrimRenderer.swift:52-63):
 * binds HIP device `ordinal` and checks that the gfx950 kernels load. */
dr_status dr_hip_device_create(int ordinal, dr_device **out);
dr_status dr_hip_device_destroy(dr_device *device);

/* AccelerationBackend.build + buffer upload
 * (Acceleration/AccelerationBackend.swift:7-67, Renderer/RenderSession.swift:371-520):
 * copies the records into HBM and builds this backend's traversal BVH. */
dr_status dr_hip_scene_upload(dr_device *device, const dr_scene_desc *desc, dr_scene **out);
dr_status dr_hip_scene_release(dr_scene *scene);
/* Build facts: node count, leaf count, max depth, build seconds (as float). */
dr_status dr_hip_scene_info(const dr_scene *scene, uint32_t *node_count, uint32_t *leaf_count,
                            uint32_t *max_depth, double *build_seconds);

/* DenrimRenderer.makeSession(scene:settings:accelerationMode:)
 * (Renderer/DenrimRenderer.swift:253-292) for accelerationMode "hip".
 * `previous_camera` may be NULL (then the current camera). */
dr_status dr_hip_session_create(dr_scene *scene, const dr_settings *settings, const dr_gpu_camera *camera,
                                const dr_gpu_camera *previous_camera, dr_session **out);
dr_status dr_hip_session_destroy(dr_session *session);

/* RenderSession.render(samples:) (Renderer/RenderSession.swift:1010-1017):
 * renders `samples` more progressive samples and waits for them. */
dr_status dr_hip_session_render(dr_session *session, uint32_t samples);

/* RenderSession.encodeNextSample(into:) (Renderer/RenderSession.swift:854-860):
 * enqueues `samples` samples on the session stream without waiting. */
dr_status dr_hip_session_encode(dr_session *session, uint32_t samples);
/* commandBuffer.waitUntilCompleted() (Renderer/RenderSession.swift:841-847). */
dr_status dr_hip_session_synchronize(dr_session *session);

/* RenderSession.renderNextTile(_:completesSample:) (Renderer/RenderSession.swift:876-889). */
dr_status dr_hip_session_render_tile(dr_session *session, uint32_t x, uint32_t y, uint32_t width, uint32_t height,
                                     int completes_sample);

/* RenderSession.resetAccumulation() (Renderer/RenderSession.swift:582-588). */
dr_status dr_hip_session_reset(dr_session *session);

/* RenderSession.updateCamera(_:previousCamera:) (Renderer/RenderSession.swift:595-608);
 * `previous_camera` NULL means the camera active before the call. */
dr_status dr_hip_session_update_camera(dr_session *session, const dr_gpu_camera *camera,
                                       const dr_gpu_camera *previous_camera);

/* RenderSession.sampleCount (Renderer/RenderSession.swift:227). */
dr_status dr_hip_session_sample_count(const dr_session *session, uint32_t *out);

/* RenderSession.pixels(for:) (Renderer/RenderSession.swift:1037-1039):
 * RGBA32F, row-major, row 0 at the top; `floats` must be >= width*height*4. */
dr_status dr_hip_session_read_output(dr_session *session, dr_output kind, float *rgba, size_t floats);

/* RenderSession.sdfTraversalStats() (Renderer/RenderSession.swift:807-831):
 * counters accumulated since creation or the last reset_stats. */
dr_status dr_hip_session_stats(dr_session *session, uint64_t *counters, size_t count);
dr_status dr_hip_session_reset_stats(dr_session *session);

/* Closest triangle hit of `count` rays (8 floats each: origin xyz, unused,
 * direction xyz, unused) through the scene's traversal BVH with the chosen
 * dr_intersection test: t (+inf on a miss) and the triangle index in the
 * scene's triangle array (~0 on a miss); the lowest index wins ties.
 * Blocking. The reference's ray query outside a render is CPUIntersector
 * (Acceleration/CPUIntersector.swift:5-43, brute force on the CPU). */
dr_status dr_hip_scene_intersect(dr_scene *scene, const float *rays, uint32_t count, uint32_t intersection, float *t,
                                 uint32_t *ids);

/* The same ray queries on caller device buffers, allocating nothing. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). rays (count x 8 floats), t and ids are
 * device memory of the scene's device (else DR_ERR_INVALID_ARGUMENT); the
 * queries are enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and the call returns without waiting. The traversal stack spills into a
 * static 16 MiB array per device, held until the process ends: the rays are
 * traced in launches of as many lanes as that array holds at the scene's BVH
 * depth, and a scene whose spill for one block of 128 rays exceeds it is
 * refused. Calls that overlap on one device race on that array. */
dr_status dr_hip_scene_intersect_device(dr_scene *scene, const float *rays, uint32_t count, uint32_t intersection,
                                        float *t, uint32_t *ids, void *stream);

/* Rebinds an existing session to another scene, settings and camera,
 * reusing its device memory (no allocation): the session then behaves like
 * a newly created one. Fails with DR_ERR_OUT_OF_MEMORY when the new
 * settings need larger buffers than the session holds. No reference
 * counterpart: it lets a server recycle sessions without hipMalloc/hipFree. */
dr_status dr_hip_session_retarget(dr_session *session, dr_scene *scene, const dr_settings *settings,
                                  const dr_gpu_camera *camera, const dr_gpu_camera *previous_camera);

/* Device memory a session with these settings allocates (path state and
 * output planes): the admission size of a served request. */
dr_status dr_hip_session_estimate_bytes(const dr_scene *scene, const dr_settings *settings, uint64_t *bytes);

/* Batch graphs (no reference counterpart: the reference encodes Metal
 * command buffers per sample, RenderSession.swift:891-1007): when enabled,
 * each full-frame batch size is captured once as a HIP graph (all stage
 * launches of one sample batch) and replayed; results are identical to the
 * launch-by-launch path. Invalidated by update_camera. */
dr_status dr_hip_session_set_graphs(dr_session *session, int enabled);

/* Enqueues the configured denoiser on the session's current beauty
 * (RenderSession.ensureDenoisedBeautyIsCurrent, RenderSession.swift:1115-1168)
 * without reading it back; dr_hip_session_read_output(DR_OUTPUT_BEAUTY)
 * calls it when the denoised plane is not current. */
dr_status dr_hip_session_denoise(dr_session *session);

/* Learned denoiser: a 5-level U-Net (3x3 convolutions, channels 32-128,
 * 633,600 weights + biases) on BF16 MFMA with FP32 accumulation, for inference
 * and training. Input per pixel: demodulated log beauty, albedo, normal,
 * log depth and 1/spp; output: demodulated log radiance. No reference
 * counterpart (the reference's denoisers are the a-trous filter and Apple
 * SVGF). Parameters are FP32 masters in layer order, each layer W[9][cout][cin]
 * then b[cout]. */
typedef struct dr_unet dr_unet;
dr_status dr_hip_unet_create(dr_device *device, uint32_t seed, dr_unet **out);
dr_status dr_hip_unet_destroy(dr_unet *net);
dr_status dr_hip_unet_param_count(const dr_unet *net, uint64_t *count);
dr_status dr_hip_unet_get_params(dr_unet *net, float *params, size_t count);
dr_status dr_hip_unet_set_params(dr_unet *net, const float *params, size_t count);
/* A training batch from host arrays: features n*h*w*16 (rounded to BF16),
 * targets n*h*w*4 (demodulated log radiance in channels 0-2); h, w multiples of 16. */
dr_status dr_hip_unet_set_batch(dr_unet *net, const float *features, const float *targets, uint32_t n, uint32_t h,
                                uint32_t w);
/* Forward + backward on the current batch (mean L1 loss over 3 channels);
 * gradients are kept for dr_hip_unet_get_grads, parameters unchanged. */
dr_status dr_hip_unet_forward_backward(dr_unet *net, float *loss);
dr_status dr_hip_unet_get_grads(dr_unet *net, float *grads, size_t count);
dr_status dr_hip_unet_get_output(dr_unet *net, float *out, size_t floats);  /* n*h*w*16 of the last forward */
/* Forward + backward + Adam (beta 0.9 / 0.999, eps 1e-8) on the current batch. */
dr_status dr_hip_unet_train_step(dr_unet *net, float learning_rate, float *loss);
/* Device-resident training set built from sessions: each example is a noisy
 * session's beauty/albedo/normal/depth and a target session's beauty (same
 * size h x w, camera and scene); training draws batches from it on the device. */
dr_status dr_hip_unet_dataset_reserve(dr_unet *net, uint32_t capacity, uint32_t h, uint32_t w);
dr_status dr_hip_unet_dataset_add(dr_unet *net, dr_session *noisy, dr_session *target, uint32_t *index);
dr_status dr_hip_unet_train_dataset(dr_unet *net, uint32_t steps, uint32_t batch, float learning_rate,
                                    float final_learning_rate, uint32_t seed, float *losses);
/* Network used by a session whose settings select DR_DENOISER_LEARNED. */
dr_status dr_hip_session_set_denoiser_network(dr_session *session, dr_unet *net);

/* Non-blocking completion check of the work encoded so far
 * (MTLCommandBuffer status in the reference). */
dr_status dr_hip_session_query(dr_session *session, int *done);

/* Device-resident render server with continuous batching: requests
 * (resident scene, settings, camera, sample count) are admitted by memory
 * budget and active limit while others run, each on its own stream and a
 * recycled session (graph-replayed batches when a request has several); a
 * request's pixels equal those of a session rendering it alone. Results
 * stay on the device until read (dr_hip_server_result) and released; a
 * released request's session returns to the server's pool. The budget
 * bounds all session memory the server holds, pooled sessions included.
 * The serving loop of the material-preview stream; the reference has no
 * counterpart. A request holds a reference to its scene until released
 * (dr_hip_scene_release defers the free). */
dr_status dr_hip_server_create(dr_device *device, const dr_server_settings *settings, dr_server **out);
dr_status dr_hip_server_destroy(dr_server *server);  /* finishes admitted and queued requests first */
dr_status dr_hip_server_submit(dr_server *server, dr_scene *scene, const dr_settings *settings,
                               const dr_gpu_camera *camera, uint32_t samples, uint64_t *ticket);
dr_status dr_hip_server_poll(dr_server *server, uint64_t ticket, dr_request_state *state);
dr_status dr_hip_server_wait(dr_server *server, uint64_t ticket, uint32_t timeout_ms, dr_request_state *state);
dr_status dr_hip_server_result(dr_server *server, uint64_t ticket, dr_output kind, float *rgba, size_t floats);
dr_status dr_hip_server_timing(dr_server *server, uint64_t ticket, double *queued_ms, double *run_ms);
dr_status dr_hip_server_release(dr_server *server, uint64_t ticket);
dr_status dr_hip_server_usage(dr_server *server, uint64_t *used_bytes, uint32_t *active, uint32_t *queued);
/* Network for served requests whose settings select DR_DENOISER_LEARNED;
 * denoising runs on the device when a request's samples are done. */
dr_status dr_hip_server_set_denoiser_network(dr_server *server, dr_unet *net);

/* GPU time of the session's work, measured with HIP events on the session
 * stream: `total_ms` of the last render/encode call; per-stage sums since the
 * last reset_stats (raygen, extend, shade, shadow, resolve, finalize,
 * subsurface walk, shade resume) when stage timing is enabled. */
dr_status dr_hip_session_set_stage_timing(dr_session *session, int enabled);
dr_status dr_hip_session_timing(dr_session *session, double *total_ms, double *stage_ms, uint64_t *stage_launches,
                                size_t stages);

/* DistanceVolumeBuilder.buildSparse (Scene/SDFModel.swift:256-355) on the
 * host: samples the smooth union of `prims` on the virtual grid and keeps the
 * narrow-band bricks (distances rounded to FP16). */
dr_status dr_sdf_build_sparse(const dr_sdf_primitive *prims, uint32_t count, const dr_sdf_build_settings *settings,
                              dr_sdf_volume **out);
dr_status dr_sdf_volume_info(const dr_sdf_volume *volume, uint32_t *bricks, uint64_t *samples, uint64_t *candidates);
/* DistanceVolumeBuilder.build (Scene/SDFModel.swift:220-254): every sample
 * of the dense grid (x fastest) into caller arrays of dims.x*dims.y*dims.z. */
dr_status dr_sdf_build_dense(const dr_sdf_primitive *prims, uint32_t count, const dr_sdf_build_settings *settings,
                             float *distance, uint32_t *material_a, uint32_t *material_b, float *blend);
/* Per brick (z, y, x order): stored origin / dims and core origin / dims
 * (3 x int32 each); per stored sample (x fastest): distance, material A / B,
 * blend. Any pointer may be NULL. */
dr_status dr_sdf_volume_export(const dr_sdf_volume *volume, int32_t *stored_origin, int32_t *stored_dims,
                               int32_t *core_origin, int32_t *core_dims, float *distance, uint32_t *material_a,
                               uint32_t *material_b, float *blend);
dr_status dr_sdf_volume_free(dr_sdf_volume *volume);

#ifdef __cplusplus
}
#endif

#endif /* DENRIM_HIP_H */
