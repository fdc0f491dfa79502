// dr_server.cpp -- device-resident render server with continuous batching.
//
// Requests (a scene already resident in HBM, settings, camera, sample
// count) are submitted from any thread and admitted by one scheduler thread
// as soon as the memory budget and the active-request limit allow, also
// while earlier requests are still on the GPU. Each admitted request
// renders on its own session (own stream, path state and output planes),
// taken from a pool of finished sessions and rebound to the request
// (dr_hip_session_retarget: no allocation, no free, so no device-wide
// synchronization in steady state); requests with several sample batches
// replay them as HIP graphs. Requests therefore overlap on the device, and
// a request's pixels are the same alone and in any mix: nothing is shared
// between sessions but the read-only scene. A finished request keeps its
// output planes in its session until released; when the budget has no room
// for a queued request, the oldest finished ones are copied to host memory
// and their sessions recycled.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dr_internal.h"

int drDeviceOrdinal(const dr_device *device);  // dr_api.cpp
void drSceneRetain(dr_scene *scene);            // dr_api.cpp: scene references
void drSceneDrop(dr_scene *scene);

namespace {

using Clock = std::chrono::steady_clock;

struct Shell {  // a session and the bytes it holds
    dr_session *session = nullptr;
    uint64_t bytes = 0;
};

struct Request {
    uint64_t ticket = 0;
    dr_scene *scene = nullptr;
    dr_settings settings{};
    dr_gpu_camera camera{};
    uint32_t samples = 0;
    uint64_t bytes = 0;  // a new session's size for this request
    dr_request_state state = DR_REQ_QUEUED;
    std::string error;
    Shell shell;
    std::vector<float> host;  // output planes copied off the device (8 x pixels x 4)
    int readers = 0;          // result() calls reading the session's planes
    uint32_t done = 0, inflight = 0;
    bool finishing = false;   // samples done, denoiser enqueued
    Clock::time_point tSubmit, tStart, tDone;
};

double ms(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

}  // namespace

struct dr_server {
    dr_device *device = nullptr;
    dr_server_settings cfg{};
    std::mutex mu;
    std::condition_variable cv;      // scheduler wake-up
    std::condition_variable doneCv;  // request completion
    std::deque<Request *> pending;
    std::vector<Request *> active;
    std::vector<Shell> pool;         // finished sessions ready for reuse
    std::map<uint64_t, std::unique_ptr<Request>> all;
    uint64_t nextTicket = 1;
    uint64_t allocated = 0;          // bytes of every session the server holds
    dr_unet *network = nullptr;      // learned denoiser for requests that select it
    bool stop = false;
    std::thread worker;
    int ordinal = 0;
    // Fault injection for tests of the failure path: when set (environment
    // DENRIM_HIP_SERVER_FAIL_ENCODE = n > 0 at server creation), every n-th
    // encode call is reported as failed without touching the device.
    uint32_t failEncodeEvery = 0, encodeCalls = 0;

    void run();
    void admit();
    bool bind(Request *r);
    bool spillOne();
    void fail(Request *r, const std::string &err);
};

void dr_server::fail(Request *r, const std::string &err) {
    r->state = DR_REQ_FAILED;
    r->error = err;
    r->tDone = Clock::now();
    doneCv.notify_all();
}

// Copies the outputs of the oldest finished request that still holds a
// session to host memory and returns the session to the pool. Called with mu.
bool dr_server::spillOne() {
    for (auto &kv : all) {
        Request *r = kv.second.get();
        if (r->state != DR_REQ_DONE || !r->shell.session || r->readers) continue;
        const size_t plane = size_t(r->settings.width) * r->settings.height * 4;
        r->host.assign(plane * 8, 0.f);
        for (int k = 0; k < 8; ++k)
            if (dr_hip_session_read_output(r->shell.session, dr_output(k), r->host.data() + k * plane, plane) != DR_OK) {
                r->host.clear();
                return false;
            }
        pool.push_back(r->shell);
        r->shell = Shell{};
        return true;
    }
    return false;
}

// Gives r a session: a pooled one that fits, else a new one within the
// budget, making room by dropping pooled sessions and spilling finished
// requests to the host. Returns false while there is no room. Called with mu.
// A request with the learned denoiser takes only a session that already
// holds a large enough U-Net workspace, and a new session for it gets that
// workspace at once: the budget counts it (the request's bytes include it)
// and the denoise on the scheduler thread then allocates nothing.
bool dr_server::bind(Request *r) {
    const bool learned = r->settings.denoiser == DR_DENOISER_LEARNED;
    const size_t unetBytes = learned ? drUnetWorkBytes(r->settings.width, r->settings.height) : 0u;
    for (;;) {
        for (size_t i = 0; i < pool.size(); ++i) {
            if (learned && pool[i].session->unetWorkBytes < unetBytes) continue;
            if (dr_hip_session_retarget(pool[i].session, r->scene, &r->settings, &r->camera, &r->camera) == DR_OK) {
                r->shell = pool[i];
                pool.erase(pool.begin() + long(i));
                return true;
            }
        }
        while (allocated + r->bytes > cfg.memory_budget_bytes && !pool.empty()) {
            auto big = std::max_element(pool.begin(), pool.end(),
                                        [](const Shell &a, const Shell &b) { return a.bytes < b.bytes; });
            dr_hip_session_destroy(big->session);
            allocated -= big->bytes;
            pool.erase(big);
        }
        if (allocated + r->bytes <= cfg.memory_budget_bytes) break;
        if (!spillOne()) return false;
    }
    dr_session *s = nullptr;
    if (dr_hip_session_create(r->scene, &r->settings, &r->camera, &r->camera, &s) != DR_OK) {
        fail(r, dr_last_error());
        return true;
    }
    if (learned && drSessionReserveUnet(s) != DR_OK) {
        std::string err = dr_last_error();
        dr_hip_session_destroy(s);
        fail(r, err);
        return true;
    }
    r->shell = Shell{s, r->bytes};
    allocated += r->bytes;
    return true;
}

// Admits queued requests (in arrival order, or least work first) while the
// budget and the active limit allow; a request larger than the whole budget
// fails. Called with mu.
void dr_server::admit() {
    while (!pending.empty() && active.size() < cfg.max_active) {
        auto next = pending.begin();
        if (cfg.admission == DR_ADMIT_LEAST_WORK) {
            auto work = [](const Request *q) {
                return uint64_t(q->settings.width) * q->settings.height * std::max(1u, q->samples);
            };
            for (auto it = pending.begin(); it != pending.end(); ++it)
                if (work(*it) < work(*next)) next = it;
        }
        Request *r = *next;
        if (r->bytes > cfg.memory_budget_bytes) {
            pending.erase(next);
            fail(r, "request needs more memory than the server budget");
            continue;
        }
        if (!bind(r)) break;
        pending.erase(next);
        if (r->state == DR_REQ_FAILED) continue;
        const uint32_t k = std::max(1u, r->settings.samples_per_launch);
        const bool graphs = cfg.use_graphs && (r->samples + k - 1) / k >= 4u;
        dr_hip_session_set_graphs(r->shell.session, graphs ? 1 : 0);
        dr_hip_session_set_denoiser_network(r->shell.session, network);
        r->state = DR_REQ_RUNNING;
        r->tStart = Clock::now();
        active.push_back(r);
    }
}

// Scheduler: admit, keep one encode call in flight per active request,
// retire finished requests, sleep briefly when nothing progressed.
void dr_server::run() {
    (void)hipSetDevice(ordinal);
    while (true) {
        std::vector<Request *> act;
        {
            std::unique_lock<std::mutex> lk(mu);
            admit();
            if (stop && active.empty() && pending.empty()) break;
            act = active;
        }
        bool progressed = false;
        for (Request *&r : act) {
            if (r->inflight == 0 && r->done < r->samples) {
                uint32_t k = std::min(r->samples - r->done, cfg.samples_per_batch);
                const bool injected = failEncodeEvery && ++encodeCalls % failEncodeEvery == 0;
                if (injected || dr_hip_session_encode(r->shell.session, k) != DR_OK) {
                    std::lock_guard<std::mutex> lk(mu);
                    active.erase(std::find(active.begin(), active.end(), r));
                    fail(r, injected ? std::string("injected encode failure") : std::string(dr_last_error()));
                    // Once FAILED is published a client may release (and
                    // free) the request: drop it from this snapshot.
                    r = nullptr;
                    progressed = true;
                    continue;
                }
                r->inflight = k;
                progressed = true;
            }
        }
        for (Request *r : act) {
            if (!r || (r->inflight == 0 && !r->finishing) || r->state != DR_REQ_RUNNING) continue;
            int done = 0;
            dr_status q = dr_hip_session_query(r->shell.session, &done);
            if (q == DR_OK && !done) continue;
            progressed = true;
            if (q == DR_OK && !r->finishing) {
                r->done += r->inflight;
                r->inflight = 0;
                if (r->done < r->samples) continue;
                if (r->settings.denoiser != DR_DENOISER_NONE) {
                    // Denoise on the device before the request is done, so a
                    // result is a plain copy (RenderSession.swift:1115-1168).
                    q = dr_hip_session_denoise(r->shell.session);
                    if (q == DR_OK) {
                        r->finishing = true;
                        continue;
                    }
                }
            }
            std::lock_guard<std::mutex> lk(mu);
            active.erase(std::find(active.begin(), active.end(), r));
            if (q != DR_OK) {
                fail(r, dr_last_error());
                continue;
            }
            r->state = DR_REQ_DONE;
            r->tDone = Clock::now();
            doneCv.notify_all();
        }
        if (!progressed) {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait_for(lk, std::chrono::microseconds(50));
        }
    }
}

extern "C" {

dr_status dr_hip_server_create(dr_device *device, const dr_server_settings *settings, dr_server **out) {
    if (!device || !settings || !out) return DR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (settings->max_active == 0 || settings->memory_budget_bytes == 0 || settings->admission > DR_ADMIT_LEAST_WORK)
        return DR_ERR_INVALID_ARGUMENT;
    auto *s = new dr_server;
    s->device = device;
    s->cfg = *settings;
    if (s->cfg.samples_per_batch == 0) s->cfg.samples_per_batch = 0xffffffffu;
    s->ordinal = drDeviceOrdinal(device);
    if (const char *f = std::getenv("DENRIM_HIP_SERVER_FAIL_ENCODE")) s->failEncodeEvery = uint32_t(std::strtoul(f, nullptr, 10));
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return DR_OK;
}

dr_status dr_hip_server_destroy(dr_server *s) {
    if (!s) return DR_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    s->worker.join();
    for (auto &kv : s->all) {
        if (kv.second->shell.session) dr_hip_session_destroy(kv.second->shell.session);
        drSceneDrop(kv.second->scene);
    }
    for (auto &sh : s->pool) dr_hip_session_destroy(sh.session);
    delete s;
    return DR_OK;
}

dr_status dr_hip_server_submit(dr_server *s, dr_scene *scene, const dr_settings *settings,
                               const dr_gpu_camera *camera, uint32_t samples, uint64_t *ticket) {
    if (!s || !scene || !settings || !camera || !ticket) return DR_ERR_INVALID_ARGUMENT;
    if (settings->width == 0 || settings->height == 0) return DR_ERR_INVALID_SCENE;
    auto r = std::make_unique<Request>();
    r->scene = scene;
    r->settings = *settings;
    r->camera = *camera;
    r->samples = samples;
    // Up to about 4 M paths per batch (1.4 GB of path state): small requests
    // run in one batch, and a 1280x720 request in batches of 4 samples. The
    // requests on the GPU share it, so a batch needs fewer paths than a
    // session rendering alone (on the 48-request material-preview stream, 4
    // active, least work first: 27 images/s at 1 M, 37 at 4 M, 35 at 16 M).
    if (r->settings.samples_per_launch == 0) {
        uint64_t px = uint64_t(settings->width) * settings->height;
        r->settings.samples_per_launch = uint32_t(
            std::max<uint64_t>(1, std::min<uint64_t>(samples ? samples : 1, (4u << 20) / std::max<uint64_t>(px, 1))));
    }
    dr_status rc = dr_hip_session_estimate_bytes(scene, &r->settings, &r->bytes);
    if (rc != DR_OK) return rc;
    r->tSubmit = Clock::now();
    drSceneRetain(scene);  // the request holds the scene until released
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->nextTicket++;
    *ticket = r->ticket;
    Request *p = r.get();
    s->all[r->ticket] = std::move(r);
    s->pending.push_back(p);
    s->cv.notify_all();
    return DR_OK;
}

dr_status dr_hip_server_poll(dr_server *s, uint64_t ticket, dr_request_state *state) {
    if (!s || !state) return DR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return DR_ERR_INVALID_ARGUMENT;
    *state = it->second->state;
    return DR_OK;
}

dr_status dr_hip_server_wait(dr_server *s, uint64_t ticket, uint32_t timeout_ms, dr_request_state *state) {
    if (!s) return DR_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return DR_ERR_INVALID_ARGUMENT;
    Request *r = it->second.get();
    s->doneCv.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                       [r] { return r->state == DR_REQ_DONE || r->state == DR_REQ_FAILED; });
    if (state) *state = r->state;
    return DR_OK;
}

dr_status dr_hip_server_result(dr_server *s, uint64_t ticket, dr_output kind, float *rgba, size_t floats) {
    if (!s || !rgba) return DR_ERR_INVALID_ARGUMENT;
    if (int(kind) < 0 || int(kind) > 7) return DR_ERR_INVALID_ARGUMENT;
    Request *r = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->all.find(ticket);
        if (it == s->all.end()) return DR_ERR_INVALID_ARGUMENT;
        r = it->second.get();
        if (r->state == DR_REQ_FAILED) return DR_ERR_COMMAND_BUFFER_FAILED;
        if (r->state != DR_REQ_DONE) return DR_ERR_INVALID_ARGUMENT;
        if (!r->shell.session) {  // spilled to the host
            const size_t plane = size_t(r->settings.width) * r->settings.height * 4;
            if (floats < plane) return DR_ERR_INVALID_ARGUMENT;
            std::memcpy(rgba, r->host.data() + size_t(kind) * plane, plane * sizeof(float));
            return DR_OK;
        }
        r->readers++;  // keeps the scheduler from recycling the session meanwhile
    }
    dr_status rc = dr_hip_session_read_output(r->shell.session, kind, rgba, floats);
    std::lock_guard<std::mutex> lk(s->mu);
    r->readers--;
    return rc;
}

dr_status dr_hip_server_timing(dr_server *s, uint64_t ticket, double *queued_ms, double *run_ms) {
    if (!s) return DR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return DR_ERR_INVALID_ARGUMENT;
    Request *r = it->second.get();
    if (r->state != DR_REQ_DONE && r->state != DR_REQ_FAILED) return DR_ERR_INVALID_ARGUMENT;
    const bool ran = r->state == DR_REQ_DONE;
    if (queued_ms) *queued_ms = ran ? ms(r->tSubmit, r->tStart) : ms(r->tSubmit, r->tDone);
    if (run_ms) *run_ms = ran ? ms(r->tStart, r->tDone) : 0.0;
    return DR_OK;
}

dr_status dr_hip_server_release(dr_server *s, uint64_t ticket) {
    if (!s) return DR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return DR_ERR_INVALID_ARGUMENT;
    Request *r = it->second.get();
    if (r->state != DR_REQ_DONE && r->state != DR_REQ_FAILED) return DR_ERR_INVALID_ARGUMENT;
    if (r->readers) return DR_ERR_INVALID_ARGUMENT;  // a result() call is reading it
    if (r->shell.session) s->pool.push_back(r->shell);
    drSceneDrop(r->scene);
    s->all.erase(it);
    s->cv.notify_all();
    return DR_OK;
}

dr_status dr_hip_server_set_denoiser_network(dr_server *s, dr_unet *net) {
    if (!s) return DR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    s->network = net;
    return DR_OK;
}

dr_status dr_hip_server_usage(dr_server *s, uint64_t *used_bytes, uint32_t *active, uint32_t *queued) {
    if (!s) return DR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    if (used_bytes) *used_bytes = s->allocated;
    if (active) *active = uint32_t(s->active.size());
    if (queued) *queued = uint32_t(s->pending.size());
    return DR_OK;
}

}  // extern "C"
