// server.cpp -- render-request serving over the C ABI: resident scenes, per-request views with
// their ReSTIR reservoirs kept on the device, admission by device-memory footprint, and either
// continuous batching (every step renders one frame of every admitted request in one merged
// launch set, hdr_render_batch; requests join and leave between steps) or the plain per-request
// path (one request at a time, all its frames through hdr_render_frame).
//
// No counterpart in the reference: its renderer serves one Hydra render pass at a time
// (renderer.cpp:390-478). The estimator is unchanged: a request renders frames 0..spp-1 of its
// view exactly as hdr_render_frame would, so its image is the same alone and in any mix.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/hdrgpu.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Device bytes a view of W x H pixels can hold for a pipeline (the buffers hdr_render_frame
// allocates: reservoirs, accumulation, colour, G-buffer; the deferred path's colour pair, flags
// and shadow queue; the wavefront path's path records and queues; spatial reuse's post-temporal
// reservoirs and per-pixel record; RNG spill slots when the path can draw past output 311).
uint64_t view_footprint(int w, int h, const hdr_pipeline_desc &p) {
    const uint64_t P = uint64_t(w) * uint64_t(h);
    uint64_t per = sizeof(hdr_reservoir) + 16 + 16 + 2 * 64;  // G-buffer: two halves
    // RIS with skipVisibility: the deferred path (depth 1; colour pair, flags, shadow queue) or,
    // at any depth, the wavefront path (path record, bounce hit, ray / shadow / contribution
    // queues, two live lists); the scene decides which, so the larger one is counted
    if (p.direct_light == HDR_DIRECT_RIS && p.skip_visibility) per += 144 + 64 + 32 + 32 + 16 + 8;
    // spatial reuse: post-temporal reservoirs, per-pixel record, colour pair and flags
    if (p.spatial_count > 0) per += sizeof(hdr_reservoir) + 32 + 32 + 4;
    // draw bound of the megakernel path (host_scene / hdr_render_frame use the same count)
    const int M = p.candidate_count;
    const int nNee = std::max(1, M / 2), nBsdf = std::max(0, M - nNee);
    const int perVertex =
        1 + 9 + 1 + (p.direct_light == HDR_DIRECT_MIS ? 3 : nNee * 3 + std::max(0, nBsdf - 1) * 9 + M + 1);
    if (5 + p.max_depth * perVertex >= 312) per += 2 * 312 * 8;
    return P * per + (uint64_t(1) << 20);
}

struct Request {
    uint64_t id = 0;
    const hdr_scene *scene = nullptr;
    int w = 0, h = 0, spp = 0, done = 0;
    hdr_frame_desc frame{};
    hdr_pipeline_desc pipe{};
    uint64_t footprint = 0;
    hdr_view *view = nullptr;
    Clock::time_point submitted, admitted, finished;
};

struct PooledView {
    hdr_view *view = nullptr;
    int w = 0, h = 0;
    uint64_t footprint = 0;
};

}  // namespace

struct hdr_server {
    hdr_context *ctx = nullptr;
    hdr_server_desc desc{};
    hdr_batch *batch = nullptr;
    uint64_t next_id = 1;
    uint64_t in_use = 0;  // footprint of active, finished-but-uncollected and pooled views
    uint64_t steps = 0;
    std::deque<Request> queued;
    std::vector<Request> active;
    std::deque<Request> finished;
    std::vector<PooledView> pool;
    std::string err;
};

namespace {

hdr_status release_view(hdr_server *s, hdr_view *v, int w, int h, uint64_t footprint) {
    s->pool.push_back(PooledView{v, w, h, footprint});
    return HDR_OK;
}

// A view for the request: a pooled one of the same size, else a new one if the budget allows
// (idle pooled views of other sizes are freed first, oldest first).
hdr_status acquire_view(hdr_server *s, Request &r, bool &ok) {
    ok = false;
    for (size_t i = 0; i < s->pool.size(); ++i) {
        PooledView &pv = s->pool[i];
        if (pv.w == r.w && pv.h == r.h && pv.footprint >= r.footprint) {
            r.view = pv.view;
            r.footprint = pv.footprint;
            s->pool.erase(s->pool.begin() + long(i));
            if (hdr_status st = hdr_view_reset(r.view)) return st;
            ok = true;
            return HDR_OK;
        }
    }
    while (s->in_use + r.footprint > s->desc.hbm_budget_bytes && !s->pool.empty()) {
        PooledView pv = s->pool.front();
        s->pool.erase(s->pool.begin());
        s->in_use -= pv.footprint;
        if (hdr_status st = hdr_view_destroy(pv.view)) return st;
    }
    if (s->in_use + r.footprint > s->desc.hbm_budget_bytes) return HDR_OK;  // wait for memory
    if (hdr_status st = hdr_view_create(s->ctx, r.w, r.h, &r.view)) return st;
    s->in_use += r.footprint;
    ok = true;
    return HDR_OK;
}

hdr_status admit(hdr_server *s) {
    const size_t cap = s->desc.batched ? size_t(s->desc.max_active) : 1;
    while (!s->queued.empty() && s->active.size() < cap) {
        Request &r = s->queued.front();
        bool ok = false;
        if (hdr_status st = acquire_view(s, r, ok)) return st;
        if (!ok) break;  // FIFO: the head waits for memory, nothing overtakes it
        r.admitted = Clock::now();
        s->active.push_back(r);
        s->queued.pop_front();
    }
    return HDR_OK;
}

hdr_frame_desc frame_of(const Request &r) {
    hdr_frame_desc f = r.frame;
    f.frame_index = r.frame.frame_index + r.done;
    return f;
}

}  // namespace

extern "C" {

hdr_status hdr_server_create(hdr_context *ctx, const hdr_server_desc *desc, hdr_server **out) {
    if (!ctx || !desc || !out) return HDR_ERR_INVALID_ARGUMENT;
    if (desc->struct_size != sizeof(hdr_server_desc)) return HDR_ERR_INVALID_ARGUMENT;
    if (desc->max_active < 1 || desc->max_active > 255 || desc->hbm_budget_bytes == 0)
        return HDR_ERR_INVALID_ARGUMENT;
    std::unique_ptr<hdr_server> s(new (std::nothrow) hdr_server());
    if (!s) return HDR_ERR_OUT_OF_MEMORY;
    s->ctx = ctx;
    s->desc = *desc;
    if (desc->batched)
        if (hdr_status st = hdr_batch_create(ctx, &s->batch)) return st;
    *out = s.release();
    return HDR_OK;
}

hdr_status hdr_server_destroy(hdr_server *s) {
    if (!s) return HDR_OK;
    if (s->batch) (void)hdr_batch_synchronize(s->batch);
    for (Request &r : s->active)
        if (r.view) (void)hdr_view_destroy(r.view);
    for (Request &r : s->finished)
        if (r.view) (void)hdr_view_destroy(r.view);
    for (PooledView &pv : s->pool) (void)hdr_view_destroy(pv.view);
    if (s->batch) (void)hdr_batch_destroy(s->batch);
    delete s;
    return HDR_OK;
}

hdr_status hdr_server_submit(hdr_server *s, const hdr_request_desc *req, uint64_t *id) {
    if (!s || !req || !id) return HDR_ERR_INVALID_ARGUMENT;
    if (req->struct_size != sizeof(hdr_request_desc) || !req->scene || req->width <= 0 || req->height <= 0 ||
        req->spp < 1 || req->frame.struct_size != sizeof(hdr_frame_desc) ||
        req->pipe.struct_size != sizeof(hdr_pipeline_desc))
        return HDR_ERR_INVALID_ARGUMENT;
    if (req->pipe.enable_firefly_filter && s->desc.batched) return HDR_ERR_UNSUPPORTED;
    Request r;
    r.id = s->next_id++;
    r.scene = req->scene;
    r.w = req->width;
    r.h = req->height;
    r.spp = req->spp;
    r.frame = req->frame;
    r.pipe = req->pipe;
    r.footprint = view_footprint(r.w, r.h, r.pipe);
    if (r.footprint > s->desc.hbm_budget_bytes) return HDR_ERR_OUT_OF_MEMORY;  // could never be admitted
    r.submitted = Clock::now();
    s->queued.push_back(r);
    *id = r.id;
    return HDR_OK;
}

hdr_status hdr_server_step(hdr_server *s, int32_t *n_finished) {
    if (!s) return HDR_ERR_INVALID_ARGUMENT;
    if (n_finished) *n_finished = 0;
    if (hdr_status st = admit(s)) return st;
    if (s->active.empty()) return HDR_OK;
    if (s->desc.batched) {
        const size_t n = s->active.size();
        std::vector<hdr_view *> views(n);
        std::vector<const hdr_scene *> scenes(n);
        std::vector<hdr_frame_desc> frames(n);
        std::vector<hdr_pipeline_desc> pipes(n);
        for (size_t i = 0; i < n; ++i) {
            views[i] = s->active[i].view;
            scenes[i] = s->active[i].scene;
            frames[i] = frame_of(s->active[i]);
            pipes[i] = s->active[i].pipe;
        }
        if (hdr_status st = hdr_render_batch(s->batch, int32_t(n), views.data(), scenes.data(), frames.data(),
                                             pipes.data()))
            return st;
        bool finishing = false;
        for (Request &r : s->active) finishing = ++r.done >= r.spp || finishing;
        // steps run ahead on the device; the host waits only when a request completes here (its
        // completion time is taken after the wait, and its image is read after it)
        if (finishing)
            if (hdr_status st = hdr_batch_synchronize(s->batch)) return st;
    } else {
        Request &r = s->active.front();  // plain path: the whole request, frame by frame
        while (r.done < r.spp) {
            const hdr_frame_desc f = frame_of(r);
            if (hdr_status st = hdr_render_frame(r.view, r.scene, &f, &r.pipe, nullptr, nullptr)) return st;
            ++r.done;
        }
        if (hdr_status st = hdr_view_synchronize(r.view)) return st;
    }
    ++s->steps;
    const Clock::time_point now = Clock::now();
    int32_t k = 0;
    for (size_t i = 0; i < s->active.size();) {
        if (s->active[i].done >= s->active[i].spp) {
            s->active[i].finished = now;
            s->finished.push_back(s->active[i]);
            s->active.erase(s->active.begin() + long(i));
            ++k;
        } else {
            ++i;
        }
    }
    if (n_finished) *n_finished = k;
    return HDR_OK;
}

hdr_status hdr_server_peek_result(hdr_server *s, hdr_request_result *out, int32_t *available) {
    if (!s || !out || !available) return HDR_ERR_INVALID_ARGUMENT;
    *available = s->finished.empty() ? 0 : 1;
    if (!*available) return HDR_OK;
    const Request &r = s->finished.front();
    out->id = r.id;
    out->width = r.w;
    out->height = r.h;
    out->spp = r.spp;
    out->queued_ms = ms_since(r.submitted, r.admitted);
    out->latency_ms = ms_since(r.submitted, r.finished);
    return HDR_OK;
}

hdr_status hdr_server_take_result(hdr_server *s, uint64_t id, float *rgba) {
    if (!s) return HDR_ERR_INVALID_ARGUMENT;
    for (auto it = s->finished.begin(); it != s->finished.end(); ++it) {
        if (it->id != id) continue;
        if (rgba)
            if (hdr_status st = hdr_view_read_color(it->view, rgba)) return st;
        if (hdr_status st = release_view(s, it->view, it->w, it->h, it->footprint)) return st;
        s->finished.erase(it);
        if (hdr_status st = admit(s)) return st;
        return HDR_OK;
    }
    return HDR_ERR_INVALID_ARGUMENT;
}

hdr_status hdr_server_get_stats(hdr_server *s, hdr_server_stats *out) {
    if (!s || !out) return HDR_ERR_INVALID_ARGUMENT;
    out->queued = int32_t(s->queued.size());
    out->active = int32_t(s->active.size());
    out->finished = int32_t(s->finished.size());
    out->pooled_views = int32_t(s->pool.size());
    out->hbm_in_use = s->in_use;
    out->steps = s->steps;
    return HDR_OK;
}

}  // extern "C"
