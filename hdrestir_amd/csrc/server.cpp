// server.cpp -- render-request serving over the C ABI: resident scenes, per-request views with
// their ReSTIR reservoirs kept on the device, admission by device-memory footprint, and either
// continuous batching (every step renders one frame of every admitted request in one merged
// launch set, hdr_render_batch; requests join and leave between steps) or the plain per-request
// path (one request at a time, all its frames through hdr_render_frame). The server is stepped by
// the caller, or runs on a host thread of its own (own_thread).
//
// No counterpart in the reference: its renderer serves one Hydra render pass at a time
// (renderer.cpp:390-478). The estimator is unchanged: a request renders frames 0..spp-1 of its
// view exactly as hdr_render_frame would, so its image is the same alone and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hdrgpu.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// A queued request may be overtaken by later ones that fit the free memory at most this many times;
// then admission waits for it (no starvation of large requests).
constexpr int kMaxOvertakes = 8;

struct Request {
    uint64_t id = 0;
    const hdr_scene *scene = nullptr;
    int w = 0, h = 0, spp = 0, done = 0;
    hdr_frame_desc frame{};
    hdr_pipeline_desc pipe{};
    uint64_t footprint = 0;
    hdr_view *view = nullptr;
    int overtaken = 0;
    uint64_t last_step = 0;  // batch step that carries its last frame
    hdr_status status = HDR_OK;
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
    uint64_t in_use = 0;  // footprint of active, draining, finished-but-uncollected and pooled views
    uint64_t steps = 0;
    std::deque<Request> queued;
    std::vector<Request> active;    // frames still to render
    std::vector<Request> draining;  // every frame queued, the device not yet done
    std::deque<Request> finished;
    std::vector<PooledView> pool;
    // own_thread: the worker and what guards the state above (in step mode it is taken too, so the
    // calls are also safe from several threads)
    std::mutex mu;
    std::condition_variable work, results;
    std::thread worker;
    bool stop = false;
    hdr_status worker_status = HDR_OK;
};

namespace {

void release_view(hdr_server *s, hdr_view *v, int w, int h, uint64_t footprint) {
    s->pool.push_back(PooledView{v, w, h, footprint});
}

// A view for the request: a pooled one of the same size, else a new one if the budget allows
// (idle pooled views of other sizes are freed first, oldest first). ok = false: not now.
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

// Admits queued requests in arrival order while slots are free; a request that does not fit the
// free memory is passed over (and counted) unless it has been passed over kMaxOvertakes times.
hdr_status admit(hdr_server *s) {
    const size_t cap = s->desc.batched ? size_t(s->desc.max_active) : 1;
    for (auto it = s->queued.begin(); it != s->queued.end() && s->active.size() < cap;) {
        bool ok = false;
        if (hdr_status st = acquire_view(s, *it, ok)) return st;
        if (!ok) {
            if (++it->overtaken > kMaxOvertakes) break;  // this one goes next: nothing overtakes it
            ++it;
            continue;
        }
        it->admitted = Clock::now();
        s->active.push_back(*it);
        it = s->queued.erase(it);
    }
    return HDR_OK;
}

hdr_frame_desc frame_of(const Request &r) {
    hdr_frame_desc f = r.frame;
    f.frame_index = r.frame.frame_index + r.done;
    return f;
}

void finish(hdr_server *s, Request &r, hdr_status st, Clock::time_point when) {
    r.status = st;
    r.finished = when;
    s->finished.push_back(r);
}

// Moves draining requests whose last step the device has finished (all of them with wait) to the
// finished list; returns how many.
hdr_status collect(hdr_server *s, bool wait, int32_t &moved) {
    moved = 0;
    if (s->draining.empty()) return HDR_OK;
    uint64_t issued = 0, completed = 0;
    if (wait) {
        if (hdr_status st = hdr_batch_synchronize(s->batch)) return st;
    }
    if (hdr_status st = hdr_batch_poll(s->batch, &issued, &completed)) return st;
    const Clock::time_point now = Clock::now();
    for (size_t i = 0; i < s->draining.size();) {
        if (s->draining[i].last_step <= completed) {
            finish(s, s->draining[i], HDR_OK, now);
            s->draining.erase(s->draining.begin() + long(i));
            ++moved;
        } else {
            ++i;
        }
    }
    return HDR_OK;
}

// One step (the state lock held): admit, render one frame of every active request (batched) or one
// whole request (plain), retire what the device has finished.
hdr_status step_locked(hdr_server *s, int32_t &n_finished) {
    n_finished = 0;
    if (hdr_status st = admit(s)) return st;
    if (s->desc.batched) {
        if (!s->active.empty()) {
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
            const hdr_status bs =
                hdr_render_batch(s->batch, int32_t(n), views.data(), scenes.data(), frames.data(), pipes.data());
            if (bs != HDR_OK) {
                // render each request's frame alone (the same frame): a request whose frame fails
                // finishes with that error, the others go on
                const Clock::time_point now = Clock::now();
                for (size_t i = 0; i < s->active.size();) {
                    Request &r = s->active[i];
                    const hdr_frame_desc f = frame_of(r);
                    const hdr_status st = hdr_render_frame(r.view, r.scene, &f, &r.pipe, nullptr, nullptr);
                    if (st != HDR_OK) {
                        finish(s, r, st, now);
                        s->active.erase(s->active.begin() + long(i));
                        ++n_finished;
                        continue;
                    }
                    ++r.done;
                    ++i;
                }
            } else {
                uint64_t issued = 0, completed = 0;
                if (hdr_status st = hdr_batch_poll(s->batch, &issued, &completed)) return st;
                for (Request &r : s->active) {
                    ++r.done;
                    r.last_step = issued;
                }
            }
            for (size_t i = 0; i < s->active.size();) {
                if (s->active[i].done >= s->active[i].spp) {
                    s->draining.push_back(s->active[i]);
                    s->active.erase(s->active.begin() + long(i));
                } else {
                    ++i;
                }
            }
            ++s->steps;
        }
        // steps run ahead on the device; the host waits only when nothing is left to queue
        int32_t moved = 0;
        if (hdr_status st = collect(s, s->active.empty(), moved)) return st;
        n_finished += moved;
    } else if (!s->active.empty()) {
        Request &r = s->active.front();  // plain path: the whole request, frame by frame
        hdr_status st = HDR_OK;
        while (r.done < r.spp && st == HDR_OK) {
            const hdr_frame_desc f = frame_of(r);
            st = hdr_render_frame(r.view, r.scene, &f, &r.pipe, nullptr, nullptr);
            if (st == HDR_OK) ++r.done;
        }
        if (st == HDR_OK) st = hdr_view_synchronize(r.view);
        finish(s, r, st, Clock::now());
        s->active.erase(s->active.begin());
        ++s->steps;
        n_finished = 1;
    }
    return HDR_OK;
}

void worker_loop(hdr_server *s) {
    std::unique_lock<std::mutex> lock(s->mu);
    while (!s->stop) {
        if (s->queued.empty() && s->active.empty() && s->draining.empty()) {
            s->work.wait(lock);
            continue;
        }
        int32_t n = 0;
        const hdr_status st = step_locked(s, n);
        if (st != HDR_OK) {  // a device error: reported to the waiting callers, the worker stops
            s->worker_status = st;
            s->results.notify_all();
            s->work.wait(lock, [s] { return s->stop; });
            break;
        }
        if (n > 0) s->results.notify_all();
        if (s->active.empty() && s->draining.empty() && !s->queued.empty()) {
            // everything queued waits for memory held by uncollected results
            s->work.wait_for(lock, std::chrono::milliseconds(2));
        } else if (s->desc.batched) {
            lock.unlock();  // let submissions and collections in between steps
            std::this_thread::yield();
            lock.lock();
        }
    }
}

}  // namespace

extern "C" {

hdr_status hdr_server_create(hdr_context *ctx, const hdr_server_desc *desc, hdr_server **out) {
    if (!ctx || !desc || !out) return HDR_ERR_INVALID_ARGUMENT;
    if (desc->struct_size != sizeof(hdr_server_desc)) return HDR_ERR_INVALID_ARGUMENT;
    if (desc->max_active < 1 || desc->max_active > 255 || desc->hbm_budget_bytes == 0)
        return HDR_ERR_INVALID_ARGUMENT;
    if (desc->own_thread != 0 && desc->own_thread != 1) return HDR_ERR_INVALID_ARGUMENT;
    std::unique_ptr<hdr_server> s(new (std::nothrow) hdr_server());
    if (!s) return HDR_ERR_OUT_OF_MEMORY;
    s->ctx = ctx;
    s->desc = *desc;
    if (desc->batched)
        if (hdr_status st = hdr_batch_create(ctx, &s->batch)) return st;
    if (desc->own_thread) {
        try {
            s->worker = std::thread(worker_loop, s.get());
        } catch (...) {
            if (s->batch) (void)hdr_batch_destroy(s->batch);
            return HDR_ERR_OUT_OF_MEMORY;
        }
    }
    *out = s.release();
    return HDR_OK;
}

hdr_status hdr_server_destroy(hdr_server *s) {
    if (!s) return HDR_OK;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lock(s->mu);
            s->stop = true;
        }
        s->work.notify_all();
        s->worker.join();
    }
    if (s->batch) (void)hdr_batch_synchronize(s->batch);
    for (Request &r : s->active)
        if (r.view) (void)hdr_view_destroy(r.view);
    for (Request &r : s->draining)
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
        req->spp < 1 || req->frame.struct_size != sizeof(hdr_frame_desc))
        return HDR_ERR_INVALID_ARGUMENT;
    if (req->pipe.enable_firefly_filter && s->desc.batched) return HDR_ERR_UNSUPPORTED;
    Request r;
    // the pipeline is validated on its scene here, as hdr_render_frame would (INVALID_ARGUMENT,
    // UNSUPPORTED), and the view's device bytes for it are what admission charges
    if (hdr_status st = hdr_view_plan_bytes(req->scene, req->width, req->height, &req->pipe, &r.footprint))
        return st;
    if (r.footprint > s->desc.hbm_budget_bytes) return HDR_ERR_OUT_OF_MEMORY;  // could never be admitted
    r.scene = req->scene;
    r.w = req->width;
    r.h = req->height;
    r.spp = req->spp;
    r.frame = req->frame;
    r.pipe = req->pipe;
    r.submitted = Clock::now();
    {
        std::lock_guard<std::mutex> lock(s->mu);
        r.id = s->next_id++;
        s->queued.push_back(r);
    }
    s->work.notify_one();
    *id = r.id;
    return HDR_OK;
}

hdr_status hdr_server_step(hdr_server *s, int32_t *n_finished) {
    if (!s) return HDR_ERR_INVALID_ARGUMENT;
    if (n_finished) *n_finished = 0;
    if (s->desc.own_thread) return HDR_ERR_UNSUPPORTED;  // the server's own thread steps it
    std::lock_guard<std::mutex> lock(s->mu);
    int32_t n = 0;
    const hdr_status st = step_locked(s, n);
    if (n_finished) *n_finished = n;
    return st;
}

static void fill_result(const Request &r, hdr_request_result *out) {
    out->id = r.id;
    out->width = r.w;
    out->height = r.h;
    out->spp = r.spp;
    out->status = r.status;
    out->queued_ms = ms_since(r.submitted, r.admitted);
    out->latency_ms = ms_since(r.submitted, r.finished);
}

hdr_status hdr_server_peek_result(hdr_server *s, hdr_request_result *out, int32_t *available) {
    if (!s || !out || !available) return HDR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->mu);
    *available = s->finished.empty() ? 0 : 1;
    if (*available) fill_result(s->finished.front(), out);
    return HDR_OK;
}

hdr_status hdr_server_wait_result(hdr_server *s, int32_t timeout_ms, hdr_request_result *out, int32_t *available) {
    if (!s || !out || !available || timeout_ms < 0) return HDR_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(s->mu);
    if (s->desc.own_thread)
        s->results.wait_for(lock, std::chrono::milliseconds(timeout_ms),
                            [s] { return !s->finished.empty() || s->worker_status != HDR_OK; });
    if (s->worker_status != HDR_OK) return s->worker_status;
    *available = s->finished.empty() ? 0 : 1;
    if (*available) fill_result(s->finished.front(), out);
    return HDR_OK;
}

hdr_status hdr_server_take_result(hdr_server *s, uint64_t id, float *rgba) {
    if (!s) return HDR_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(s->mu);
    for (auto it = s->finished.begin(); it != s->finished.end(); ++it) {
        if (it->id != id) continue;
        Request r = *it;
        s->finished.erase(it);
        // the view is no longer the server's to step: read it without holding the lock
        lock.unlock();
        hdr_status st = HDR_OK;
        if (rgba && r.status == HDR_OK) st = hdr_view_read_color(r.view, rgba);
        lock.lock();
        release_view(s, r.view, r.w, r.h, r.footprint);
        if (!s->desc.own_thread)
            if (hdr_status ast = admit(s)) return ast;
        lock.unlock();
        s->work.notify_one();
        return st;
    }
    return HDR_ERR_INVALID_ARGUMENT;
}

hdr_status hdr_server_get_stats(hdr_server *s, hdr_server_stats *out) {
    if (!s || !out) return HDR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->mu);
    out->queued = int32_t(s->queued.size());
    out->active = int32_t(s->active.size() + s->draining.size());
    out->finished = int32_t(s->finished.size());
    out->pooled_views = int32_t(s->pool.size());
    out->hbm_in_use = s->in_use;
    out->steps = s->steps;
    return HDR_OK;
}

}  // extern "C"
