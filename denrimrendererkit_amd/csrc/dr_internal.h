// dr_internal.h -- the C ABI's opaque handles as the library sees them
// (shared by the session, denoiser and server sources; not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/denrim_hip.h"
#include "dr_bvh.h"
#include "dr_kernels.h"
#include "dr_scene.h"

// Host<->device copy that is synchronous for the caller but runs on
// `stream`, never on the legacy null stream: another thread may be capturing
// a batch graph on its own stream, and the null stream's implicit
// synchronization with other streams is refused while a capture is open.
inline hipError_t drCopy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t stream) {
    if (bytes == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream);
    return e == hipSuccess ? hipStreamSynchronize(stream) : e;
}

struct dr_device {
    int ordinal = 0;
    int cuCount = 0;
    std::string arch;
    // Device address of this device's static traversal-stack scratch
    // (dr_hip_scene_intersect_device), taken once at device creation.
    uint32_t *intersectScratch = nullptr;
};

struct dr_scene {
    dr_device *device = nullptr;
    DrSceneDev dev{};
    DrNode4 *nodes = nullptr;
    uint32_t stackMax = 2;  // traversal-stack entries a lane can need in this scene's BVH
    float4 *triPos = nullptr;
    DrTriangle *triangles = nullptr;
    DrMaterial *materials = nullptr;
    uint32_t *matKey = nullptr;
    DrLight *lights = nullptr;
    float2 *env = nullptr;
    uint32_t *envGuide = nullptr;
    uint2 *envAlias = nullptr;
    uint4 *texDesc = nullptr;
    float4 *texels = nullptr;
    DrVolume *volumes = nullptr;
    DrVolumeSample *denseSamples = nullptr;
    DrBrick *bricks = nullptr;
    uint16_t *dist16 = nullptr;
    float *dist32 = nullptr;
    uint4 *corner16 = nullptr;
    uint2 *sampleMaterial = nullptr;
    DrBrickGrid *grids = nullptr;
    uint32_t *gridIndices = nullptr;
    dr::BvhBuildStats stats;
    int sessions = 0;
    bool released = false;
};

struct dr_session {
    dr_scene *scene = nullptr;
    DrSceneDev sceneDev{};  // the scene's device view with this session's quality level
    dr_settings settings{};
    DrCamera camera{}, previousCamera{};
    uint32_t sampleCount = 0;
    uint32_t samplesPerLaunch = 1;
    uint32_t nSlots = 0;
    uint32_t traceGrid = 0;
    hipStream_t stream = nullptr;
    DrPaths paths{};
    DrOutputs out{};
    void *pathBlock = nullptr;   // one allocation for all path-state arrays
    void *outBlock = nullptr;    // one allocation for all output planes
    hipEvent_t evBegin = nullptr, evEnd = nullptr;
    std::vector<hipEvent_t> stageEvents;
    bool stageTiming = false;
    double stageMs[DR_STAGE_COUNT] = {0};
    uint64_t stageLaunches[DR_STAGE_COUNT] = {0};
    double lastMs = 0.0;
    // Batch graphs: full-frame batches captured once per batch size and
    // replayed; the batch's first sample index lives in sampleBase.
    bool useGraphs = false;
    uint32_t *sampleBase = nullptr;
    size_t pathCapacity = 0, outCapacity = 0;  // bytes of the two blocks
    std::vector<std::pair<uint32_t, hipGraphExec_t>> graphs;
    // Denoised beauty (RenderSession.swift:1115-1168): a plane and a
    // ping-pong temporary, valid until the next sample or reset.
    float4 *denoised = nullptr, *denoiseTemp = nullptr;
    bool denoisedCurrent = false;
    dr_unet *network = nullptr;  // learned denoiser, when settings.denoiser is DR_DENOISER_LEARNED
    void *unetWork = nullptr;    // the network's activations for this session's image
    size_t unetWorkBytes = 0;
};

// One lock orders graph capture (a batch graph captured on a session stream,
// possibly on the render server's scheduler thread) against the calls that
// synchronize the device or change its allocations on other threads:
// hipMalloc / hipFree, stream, event and graph-exec creation and
// destruction. It is recursive because the locked calls nest (a session
// destroy drops its graphs). Held only around those calls and the capture
// itself, never around kernel execution.
std::recursive_mutex &drDeviceMutex();
#define DR_DEVICE_LOCK std::lock_guard<std::recursive_mutex> drDeviceLock_(drDeviceMutex())

// Sets the thread's dr_last_error() message and returns s.
dr_status drFail(dr_status s, const std::string &msg);
// Samples in flight per batch of a session that leaves samples_per_launch
// open (0): about 64 M path slots, 1 to 64 samples.
uint32_t drAutoSamplesPerLaunch(uint64_t pixels);
// Device bytes of the U-Net's activations for one width x height image.
size_t drUnetWorkBytes(uint32_t width, uint32_t height);
// Enqueues the U-Net on the session's beauty into session->denoised.
dr_status drUnetDenoiseSession(dr_unet *net, dr_session *session, void **work, size_t *workBytes);
// Allocates the session's U-Net workspace for its current size now (no-op
// when it already holds enough), so that a later denoise allocates nothing.
dr_status drSessionReserveUnet(dr_session *session);

