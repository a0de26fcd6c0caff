// server.cpp - the serving form: scenes admitted by a device-memory budget,
// their worlds resident in HBM.  Scenes that fit a batch and share its
// schedule are stepped together, one launch per kernel for all of them
// (ribo_batch_*), by scheduler thread 0, the batch rebuilt when a scene joins
// or finishes; the others are fed chunks of up to 64 steps on their own
// streams as the streams drain.
//
// Host code only: it drives worlds through the C ABI of world.hip and
// batch.hip, whose kernels compute for a scene in a batch what they compute
// for it alone.

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ribo_batch.h"
#include "../../include/ribo_server.h"

namespace {

struct Scene {
    uint64_t ticket = 0;
    RiboScene req{};
    std::vector<std::vector<uint8_t>> inputs;  // owned copies of the request's buffers
    uint64_t footprint = 0;
    RiboWorld* world = nullptr;
    bool charged = false;  // footprint counted in RiboServer::used
    uint32_t left = 0;
    RiboSceneStatus st{};
    std::string error;  // the library's message when the scene failed
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct RiboServer {
    uint64_t budget = 0;
    uint32_t max_running = 0;
    uint64_t used = 0;
    uint32_t admitted = 0;  // scenes charged against the budget and not yet finished
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Scene*> queue;
    std::unordered_map<uint64_t, Scene*> scenes;
    uint64_t next_ticket = 1;
    bool stop = false;
    // scheduler threads: each steps the worlds admitted to it, so the host
    // work of recording and launching steps (graph launches cost host time
    // per node) proceeds for several worlds at once
    std::vector<std::thread> threads;
    std::vector<std::vector<Scene*>> inbox;  // admitted scenes not yet picked up, per thread
    std::vector<uint32_t> load;              // scenes each thread is running or about to start
    std::chrono::steady_clock::time_point t0;
    bool batching = true;                    // batchable scenes go to thread 0, stepped in batches
};

namespace {

// Creates and loads an admitted scene's world (scheduler thread, no lock).
int start_scene(Scene* sc) {
    RiboWorld* w = nullptr;
    int rc = ribo_world_create(&sc->req.cfg, sc->req.seed, &w);
    if (rc) return rc;
    // the server steps many scenes at once: one stream per scene
    if ((rc = ribo_world_set_pipeline(w, 0)) || (rc = ribo_world_set_params(w, &sc->req.params)) ||
        (rc = ribo_world_set_settings(w, &sc->req.settings))) {
        ribo_world_destroy(w);
        return rc;
    }
    // gamma first: the derived terrain (wall mask, slope, multigrid masks)
    // follows it
    std::vector<int> order{RIBO_BUF_GAMMA};
    for (int b = 0; b < RIBO_BUF_COUNT; ++b)
        if (b != RIBO_BUF_GAMMA) order.push_back(b);
    for (int b : order) {
        if (sc->inputs[b].empty()) continue;
        if ((rc = ribo_world_upload(w, (RiboBuffer)b, sc->inputs[b].data(), sc->inputs[b].size()))) {
            ribo_world_destroy(w);
            return rc;
        }
    }
    sc->inputs.clear();
    sc->inputs.shrink_to_fit();
    sc->world = w;
    return RIBO_OK;
}

// First in, first admitted, while the budget and the running cap allow; an
// admitted scene goes to the least loaded thread.  Called with the lock held.
void admit(RiboServer* s) {
    while (!s->queue.empty()) {
        Scene* sc = s->queue.front();
        if (s->used + sc->footprint > s->budget) break;
        if (s->max_running && s->admitted >= s->max_running) break;
        s->queue.pop_front();
        s->used += sc->footprint;
        sc->charged = true;
        ++s->admitted;
        sc->st.state = RIBO_SCENE_RUNNING;  // from here on only its thread touches its world
        const size_t k = s->batching ? 0 : (size_t)(std::min_element(s->load.begin(), s->load.end()) - s->load.begin());
        ++s->load[k];
        s->inbox[k].push_back(sc);
    }
}

// Scenes stepped together: the same schedule (fluid_substep, transport_every
// and the phase of the epoch in it) and solver (tolerance mode, precision,
// V-cycles), and the same ledger setting.
using GroupKey = std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>;
struct Group {
    std::vector<Scene*> scenes;  // scenes with steps left
    RiboBatch* batch = nullptr;
    bool dirty = true;           // the batch does not hold exactly `scenes`
};

GroupKey group_key(const Scene* sc) {
    const RiboWorldSettings& st = sc->req.settings;
    const RiboSimParams& p = sc->req.params;
    const uint32_t sub = std::min(std::max(st.fluid_substep, 1u), 4u), te = std::max(p.transport_every, 1u);
    uint32_t a = sub, b = te;
    while (b) {
        const uint32_t t = a % b;
        a = b;
        b = t;
    }
    const uint32_t period = sub / a * te;
    const uint32_t done = sc->req.steps - sc->left;
    const bool tol = st.mg_tol > 0.0f;
    return GroupKey{sub, te, (p.epoch + done) % period, tol ? 1u : 0u, tol ? st.precision : 0u,
                    tol ? 0u : std::max(st.mg_cycles, 1u), st.ledger_check ? 1u : 0u};
}

void fail_scene(RiboServer* s, Scene* sc, int rc, size_t me) {
    std::lock_guard<std::mutex> lk(s->mu);
    sc->st.state = RIBO_SCENE_FAILED;
    sc->st.error = rc;
    sc->error = ribo_last_error();
    --s->admitted;  // its memory stays charged until release
    --s->load[me];
}

void scheduler(RiboServer* s, size_t me) {
    std::vector<Scene*> running, starting, draining;
    std::map<GroupKey, Group> groups;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop) break;
            admit(s);
            starting.swap(s->inbox[me]);
            if (starting.empty() && running.empty() && groups.empty() && draining.empty()) {
                s->cv.wait_for(lk, std::chrono::milliseconds(2));
                continue;
            }
        }
        for (Scene* sc : starting) {
            const int rc = start_scene(sc);
            std::lock_guard<std::mutex> lk(s->mu);
            if (rc) {
                sc->st.state = RIBO_SCENE_FAILED;
                sc->st.error = rc;
                sc->error = ribo_last_error();
                s->used -= sc->footprint;
                sc->charged = false;
                --s->admitted;
                --s->load[me];
            } else {
                sc->st.t_admit = seconds_since(s->t0);
                if (s->batching && sc->left > 0 && ribo_batch_accepts(sc->world) == 1) {
                    Group& g = groups[group_key(sc)];
                    g.scenes.push_back(sc);
                    g.dirty = true;
                } else {
                    running.push_back(sc);
                }
            }
        }
        starting.clear();
        bool fed = false;
        // the batches: every scene of a group steps as many steps as the
        // group's scene with the fewest left (at most 64)
        for (auto it = groups.begin(); it != groups.end();) {
            Group& g = it->second;
            if (g.dirty && g.batch) {
                ribo_batch_destroy(g.batch);
                g.batch = nullptr;
            }
            if (g.scenes.empty()) {
                it = groups.erase(it);
                continue;
            }
            if (!g.batch) {
                std::vector<RiboWorld*> ws;
                for (Scene* sc : g.scenes) ws.push_back(sc->world);
                if (ribo_batch_create(ws.data(), (uint32_t)ws.size(), &g.batch) != RIBO_OK) {
                    // not batchable together after all: their own streams
                    g.batch = nullptr;
                    for (Scene* sc : g.scenes) running.push_back(sc);
                    it = groups.erase(it);
                    continue;
                }
                g.dirty = false;
            }
            uint32_t chunk = RIBO_MAX_STEPS_PER_CALL;
            for (Scene* sc : g.scenes) chunk = std::min(chunk, sc->left);
            const int n = ribo_batch_encode_steps(g.batch, chunk);
            if (n <= 0) {
                const int rc = n < 0 ? n : RIBO_ERR_ARG;
                for (Scene* sc : g.scenes) fail_scene(s, sc, rc, me);
                ribo_batch_destroy(g.batch);
                it = groups.erase(it);
                s->cv.notify_all();
                continue;
            }
            fed = true;
            std::vector<Scene*> keep;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                for (Scene* sc : g.scenes) {
                    sc->left -= (uint32_t)n;
                    sc->st.steps_done += (uint32_t)n;
                    if (sc->left > 0) keep.push_back(sc);
                    else draining.push_back(sc);
                }
            }
            if (keep.size() != g.scenes.size()) g.dirty = true;
            g.scenes.swap(keep);
            ++it;
        }
        // scenes whose last batched steps are queued: done once they ran
        for (size_t i = 0; i < draining.size();) {
            Scene* sc = draining[i];
            const int idle = ribo_world_idle(sc->world);
            if (idle == 0) {
                ++i;
                continue;
            }
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (idle < 0) {
                    sc->st.state = RIBO_SCENE_FAILED;
                    sc->st.error = idle;
                    sc->error = ribo_last_error();
                } else {
                    sc->st.state = RIBO_SCENE_DONE;
                    sc->st.t_done = seconds_since(s->t0);
                }
                --s->admitted;
                --s->load[me];
            }
            s->cv.notify_all();
            draining.erase(draining.begin() + (ptrdiff_t)i);
            fed = true;
        }
        for (size_t i = 0; i < running.size();) {
            Scene* sc = running[i];
            const int idle = ribo_world_idle(sc->world);
            int rc = idle < 0 ? idle : RIBO_OK;
            if (idle == 1 && sc->left > 0) {
                const int n = ribo_world_encode_steps(sc->world, std::min(sc->left, RIBO_MAX_STEPS_PER_CALL));
                if (n > 0) {
                    sc->left -= (uint32_t)n;
                    std::lock_guard<std::mutex> lk(s->mu);
                    sc->st.steps_done += (uint32_t)n;
                    fed = true;
                    ++i;
                    continue;
                }
                rc = n < 0 ? n : RIBO_ERR_ARG;
            }
            if (idle == 0 && rc == RIBO_OK) {
                ++i;
                continue;
            }
            // finished (every step done on the device) or failed
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (rc) {
                    sc->st.state = RIBO_SCENE_FAILED;
                    sc->st.error = rc;
                    sc->error = ribo_last_error();
                } else {
                    sc->st.state = RIBO_SCENE_DONE;
                    sc->st.t_done = seconds_since(s->t0);
                }
                --s->admitted;  // its memory stays charged until release
                --s->load[me];
            }
            s->cv.notify_all();  // a running slot is free
            running.erase(running.begin() + (ptrdiff_t)i);
            fed = true;
        }
        if (!fed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    for (auto& kv : groups)
        if (kv.second.batch) ribo_batch_destroy(kv.second.batch);
}

}  // namespace

extern "C" {

int ribo_server_create(uint64_t hbm_budget, uint32_t max_running, uint32_t threads, RiboServer** out) {
    if (!out || hbm_budget == 0) return RIBO_ERR_ARG;
    if (threads == 0) threads = std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2));
    threads = std::min(threads, 64u);
    RiboServer* s = new RiboServer();
    s->budget = hbm_budget;
    s->max_running = max_running;
    s->t0 = std::chrono::steady_clock::now();
    s->inbox.resize(threads);
    s->load.assign(threads, 0u);
    for (uint32_t k = 0; k < threads; ++k) s->threads.emplace_back(scheduler, s, (size_t)k);
    *out = s;
    return RIBO_OK;
}

int ribo_server_set_batching(RiboServer* s, int on) {
    if (!s || (on != 0 && on != 1)) return RIBO_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->batching = on != 0;
    return RIBO_OK;
}

int ribo_server_destroy(RiboServer* s) {
    if (!s) return RIBO_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    for (auto& t : s->threads) t.join();
    for (auto& kv : s->scenes) {
        if (kv.second->world) ribo_world_destroy(kv.second->world);
        delete kv.second;
    }
    delete s;
    return RIBO_OK;
}

int ribo_server_submit(RiboServer* s, const RiboScene* scene, uint64_t* ticket) {
    if (!s || !scene || !ticket) return RIBO_ERR_ARG;
    Scene* sc = new Scene();
    sc->req = *scene;
    sc->inputs.resize(RIBO_BUF_COUNT);
    for (int b = 0; b < RIBO_BUF_COUNT; ++b) {
        sc->req.input[b] = nullptr;
        if (!scene->input[b]) continue;
        if (scene->input_bytes[b] != ribo_buffer_bytes(&scene->cfg, (RiboBuffer)b)) {
            delete sc;
            return RIBO_ERR_ARG;
        }
        const uint8_t* p = (const uint8_t*)scene->input[b];
        sc->inputs[b].assign(p, p + scene->input_bytes[b]);
    }
    sc->footprint = ribo_world_footprint(&scene->cfg, &scene->settings);
    sc->left = scene->steps;
    if (sc->footprint == 0) {
        delete sc;
        return RIBO_ERR_ARG;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (sc->footprint > s->budget) {
            delete sc;
            return RIBO_ERR_OOM;
        }
        sc->ticket = s->next_ticket++;
        sc->st.state = RIBO_SCENE_QUEUED;
        sc->st.t_submit = seconds_since(s->t0);
        sc->st.footprint = sc->footprint;
        s->scenes[sc->ticket] = sc;
        s->queue.push_back(sc);
        *ticket = sc->ticket;
    }
    s->cv.notify_all();
    return RIBO_OK;
}

int ribo_server_status(RiboServer* s, uint64_t ticket, RiboSceneStatus* out) {
    if (!s || !out) return RIBO_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->scenes.find(ticket);
    if (it == s->scenes.end()) return RIBO_ERR_ARG;
    *out = it->second->st;
    return RIBO_OK;
}

int ribo_server_error(RiboServer* s, uint64_t ticket, char* buf, size_t cap) {
    if (!s || !buf || cap == 0) return RIBO_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->scenes.find(ticket);
    if (it == s->scenes.end()) return RIBO_ERR_ARG;
    const std::string& e = it->second->error;
    const size_t n = std::min(cap - 1, e.size());
    memcpy(buf, e.data(), n);
    buf[n] = 0;
    return RIBO_OK;
}

int ribo_server_download(RiboServer* s, uint64_t ticket, RiboBuffer which, void* dst, size_t bytes) {
    if (!s || !dst) return RIBO_ERR_ARG;
    RiboWorld* w = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->scenes.find(ticket);
        if (it == s->scenes.end() || it->second->st.state != RIBO_SCENE_DONE) return RIBO_ERR_ARG;
        w = it->second->world;
    }
    // a finished world is no longer touched by the scheduler
    return ribo_world_download(w, which, dst, bytes);
}

int ribo_server_release(RiboServer* s, uint64_t ticket) {
    if (!s) return RIBO_ERR_ARG;
    Scene* sc = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->scenes.find(ticket);
        if (it == s->scenes.end()) return RIBO_ERR_ARG;
        sc = it->second;
        if (sc->st.state == RIBO_SCENE_RUNNING) return RIBO_ERR_ARG;
        if (sc->st.state == RIBO_SCENE_QUEUED) {
            auto q = std::find(s->queue.begin(), s->queue.end(), sc);
            if (q != s->queue.end()) s->queue.erase(q);
        }
        if (sc->charged) s->used -= sc->footprint;
        s->scenes.erase(it);
    }
    if (sc->world) ribo_world_destroy(sc->world);
    delete sc;
    s->cv.notify_all();
    return RIBO_OK;
}

int ribo_server_stats(RiboServer* s, uint32_t* queued, uint32_t* running, uint32_t* finished, uint64_t* used) {
    if (!s) return RIBO_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    uint32_t q = 0, r = 0, f = 0;
    for (auto& kv : s->scenes) {
        const int st = kv.second->st.state;
        if (st == RIBO_SCENE_QUEUED) ++q;
        else if (st == RIBO_SCENE_RUNNING) ++r;
        else ++f;
    }
    if (queued) *queued = q;
    if (running) *running = r;
    if (finished) *finished = f;
    if (used) *used = s->used;
    return RIBO_OK;
}

}  // extern "C"
