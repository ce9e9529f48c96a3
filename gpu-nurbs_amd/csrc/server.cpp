// server.cpp -- continuous-batching inversion server (gn_server_*).
//
// One dispatcher thread and one thread per step slot.  Submitted requests wait in arrival order
// until their result buffers fit the HBM arena (admission); the dispatcher hands every admitted
// request that is not yet running to the next free slot.  The slot's thread builds the step (its
// request table and tile lists, written in class order straight into pinned staging), uploads it
// with one copy, launches the class groups and curves on the slot's streams, and polls the
// groups' events and the step's event (a class group's requests complete when that group's
// kernels have, curves with the step); up to `streams` steps are in flight at once, so requests arriving while a step
// runs are gathered into the next one as soon as a slot frees, and one slot's launch work never
// delays another slot's completion.  Every point is computed by the same per-point iteration
// whatever else shares its step, so a request's results are the same alone and in a mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include <time.h>

#include "arena.hpp"
#include "gn_internal.hpp"

using gn::fail;

namespace {

constexpr int kMaxStepReqs = 2048;     // requests per step
constexpr int kMaxStepTiles = 32768;   // tiles per step (surfaces + curves)
constexpr int kCurveTile = 256;        // points per curve tile (one workgroup, a point per lane)

enum ReqState { kWaiting, kAdmitted, kRunning, kDone, kFailed };

struct Req {
    int64_t id = 0;
    int obj = 0;
    bool curve = false;
    size_t n = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;
    size_t off = 0, bytes = 0;  // arena range
    double *u = nullptr, *v = nullptr;
    uint8_t *status = nullptr;
    int64_t submit_ns = 0, admit_ns = 0, launch_ns = 0, done_ns = 0;
    ReqState state = kWaiting;
};

size_t a256(size_t b) { return gn::Arena::round(b); }

size_t result_bytes(bool curve, size_t n) {
    return curve ? a256(8 * n) + a256(n) : 2 * a256(8 * n) + a256(n);
}

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    gn::SideStreams side;  // the step's class launches run concurrently on these
    char *h_stage = nullptr;  // pinned, coherent
    void *h_stage_dev = nullptr;  // h_stage's device-side address
    char *d_stage = nullptr;
    std::vector<Req *> reqs;  // the step's unfinished requests (set by the dispatcher under the lock)
    std::vector<std::pair<int, hipEvent_t>> group_done;  // class groups of the step on side streams
    bool busy = false;        // assigned and not finished (under the lock)
    bool go = false;          // a new assignment for the slot's thread (under the lock)
    std::condition_variable cv;
    std::thread th;
};

constexpr size_t kStageReqs = sizeof(GnReq) * kMaxStepReqs;
constexpr size_t kStageBytes = kStageReqs + sizeof(GnTile) * kMaxStepTiles + 256;

}  // namespace

int64_t gn_now_ns_impl() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (int64_t)ts.tv_sec * 1000000000ll + ts.tv_nsec;
}

struct gn_server {
    gn_batch *batch = nullptr;
    int nsurf = 0, ncurve = 0;
    std::vector<int> cls;          // degree class per surface
    GnBatchObj *d_objs = nullptr;  // the batch's object table
    int tile = 1024, bpc = 0, waves = 1;
    GnCurveDev *d_curves = nullptr;
    int max_samples = 0;
    gn_invert_opts o{};
    int device = 0;
    size_t budget = 0, resident = 0;
    char *d_arena = nullptr;
    gn::Arena arena;
    std::deque<Slot> slots;  // not moved after creation (threads hold references)

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> waiting, admitted;
    std::unordered_map<int64_t, Req *> reqs;
    int64_t next_id = 1;
    bool stop = false, dirty = false, quit = false;
    gn_server_stats stats{};
    std::thread worker;

    void admit_locked();
    void run();
    void run_slot(Slot &sl);
    int launch_step(Slot &sl);
    void finish_group(Slot &sl, int grp);
    void finish(Slot &sl, bool ok);
};

// Helpers exported by batch_api.cpp for the server (batch internals).
namespace gn {
int batch_view(const gn_batch *b, const GnBatchObj **objs, int *count, std::vector<int> *cls,
               int *tile, int *bpc, size_t *bytes, bool *grids);
}

void gn_server::admit_locked() {
    while (!waiting.empty()) {
        Req *r = waiting.front();
        size_t off = 0;
        if (!arena.alloc(r->bytes, &off)) break;  // arrival order: wait for releases
        r->off = off;
        char *base = d_arena + off;
        r->u = (double *)base;
        if (r->curve) {
            r->v = nullptr;
            r->status = (uint8_t *)(base + a256(8 * r->n));
        } else {
            r->v = (double *)(base + a256(8 * r->n));
            r->status = (uint8_t *)(base + 2 * a256(8 * r->n));
        }
        r->state = kAdmitted;
        r->admit_ns = gn_now_ns_impl();
        waiting.pop_front();
        admitted.push_back(r);
    }
    if (arena.peak() > stats.arena_peak) stats.arena_peak = arena.peak();
}

// Builds the step image in the slot's pinned staging and launches it.  Runs on the slot's thread
// without the lock; sl.reqs were handed over by the dispatcher.
int gn_server::launch_step(Slot &sl) {
    const int nreq = (int)sl.reqs.size();
    size_t spts = 0, cpts = 0;
    for (Req *r : sl.reqs) (r->curve ? cpts : spts) += r->n;
    // Surface tile size per degree class (each class is its own launch): about two tiles per
    // resident wave, at most the batch's tile; all grown by a common factor if the step would
    // exceed kMaxStepTiles (sum of ceil(n/T) <= pts/T + nreq).
    const size_t room = (size_t)(kMaxStepTiles - 2 * nreq);
    size_t per[3] = {0, 0, 0}, tsz[3];
    for (Req *r : sl.reqs)
        if (!r->curve) per[gn::class_group(cls[r->obj])] += r->n;
    size_t est = 0;
    for (int g = 0; g < 3; g++) {
        tsz[g] = (size_t)gn::class_tile(per[g], waves, tile);
        est += per[g] / tsz[g];
    }
    for (size_t grow = 2; est > room / 2; grow *= 2) {
        est = 0;
        for (int g = 0; g < 3; g++) {
            tsz[g] = (size_t)gn::class_tile(per[g], waves, tile) * grow;
            est += per[g] / tsz[g];
        }
    }
    size_t tc = kCurveTile;
    if (cpts / tc > room / 2) tc = (cpts + room / 2 - 1) / (room / 2);
    // Tiles in class_order (group, then class; stable in request order) by a counting pass:
    // count tiles per order key, then write each request's tiles at its key's running offset.
    constexpr int kKeys = 3 * 32;
    size_t cnt[kKeys + 1] = {};  // [kKeys]: curves, after every surface tile
    auto key_of = [&](const Req *r) { return r->curve ? kKeys : gn::class_order(cls[r->obj]); };
    auto tile_of = [&](const Req *r) { return r->curve ? tc : tsz[gn::class_group(cls[r->obj])]; };
    for (const Req *r : sl.reqs) {
        const int key = key_of(r);
        if (key < 0 || key > kKeys) return fail(GN_EDEVICE, "degree class out of range");
        const size_t T = tile_of(r);
        cnt[key] += (r->n + T - 1) / T;
    }
    size_t pos[kKeys + 1], total = 0;
    for (int k = 0; k <= kKeys; k++) {
        pos[k] = total;
        total += cnt[k];
    }
    if (total > (size_t)kMaxStepTiles) return fail(GN_EDEVICE, "step tile list overflow");
    const size_t ns = pos[kKeys], nc = cnt[kKeys];
    // staging image: request table, then the tiles (256-byte aligned); one copy of what is used
    GnReq *rq = (GnReq *)sl.h_stage;
    const size_t toff = a256(sizeof(GnReq) * nreq);
    GnTile *tl = (GnTile *)(sl.h_stage + toff);
    for (int k = 0; k < nreq; k++) {
        Req *r = sl.reqs[k];
        rq[k] = GnReq{r->x, r->y, r->z, r->u, r->v, r->status, nullptr, nullptr};
        const size_t T = tile_of(r);
        size_t &p = pos[key_of(r)];
        const int obj = r->curve ? r->obj - nsurf : r->obj;
        for (size_t off = 0; off < r->n; off += T)
            tl[p++] = GnTile{obj, (int32_t)std::min(T, r->n - off), (int64_t)off, k, 0};
    }
    const size_t bytes = toff + sizeof(GnTile) * (ns + nc);
    // an ordinary kernel carries the image: hipMemcpyAsync of it has held the calling thread (and
    // the other slot's launches) for 12-15 ms now and then under load
    int crc = gn::launch_stage_copy(sl.h_stage_dev, sl.d_stage, bytes, sl.st);
    if (crc) return crc;
    const GnReq *d_rq = (const GnReq *)sl.d_stage;
    const GnTile *d_tl = (const GnTile *)(sl.d_stage + toff);
    // surface groups (seeds and precision phases) on the slot's side streams, curves on the
    // slot's stream beside them; all joined back before the step's event
    std::function<int(hipStream_t)> curves;
    if (nc)
        curves = [&](hipStream_t cs) {
            return gn::launch_curve_batch(d_curves, d_tl + ns, (int)nc, d_rq, max_samples, o, cs);
        };
    int rc = GN_OK;
    sl.group_done.clear();
    if (ns)
        rc = gn::launch_batch_invert(d_objs, d_tl, gn::group_ranges(tl, ns, cls), d_rq, o, bpc,
                                     sl.st, &sl.side, curves, &sl.group_done);
    else if (nc)
        rc = curves(sl.st);
    if (rc) return rc;
    GN_HIP(hipEventRecord(sl.ev, sl.st));
    return GN_OK;
}

// Completes the step's surface requests of class group grp, whose work has finished while the
// rest of the step runs on: they leave sl.reqs (a client may release them from now on; the
// step's remaining kernels touch only the other requests' buffers).
void gn_server::finish_group(Slot &sl, int grp) {
    const int64_t now = gn_now_ns_impl();
    {
        std::lock_guard<std::mutex> lk(mu);
        size_t keep = 0;
        for (Req *r : sl.reqs) {
            if (r->curve || gn::class_group(cls[r->obj]) != grp) {
                sl.reqs[keep++] = r;
                continue;
            }
            r->done_ns = now;
            r->state = kDone;
            stats.completed++;
            stats.points += (int64_t)r->n;
        }
        sl.reqs.resize(keep);
    }
    cv_done.notify_all();
}

void gn_server::finish(Slot &sl, bool ok) {
    const int64_t now = gn_now_ns_impl();
    {
        std::lock_guard<std::mutex> lk(mu);
        for (Req *r : sl.reqs) {
            r->done_ns = now;
            r->state = ok ? kDone : kFailed;
            if (ok) {
                stats.completed++;
                stats.points += (int64_t)r->n;
            }
        }
        sl.reqs.clear();
        sl.busy = false;
        dirty = true;
    }
    cv_done.notify_all();
    cv_work.notify_one();
}

// A slot's thread: launches each step handed to it, then polls the step's event.
void gn_server::run_slot(Slot &sl) {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        sl.cv.wait(lk, [&] { return sl.go || quit; });
        if (!sl.go) return;  // quit with nothing assigned
        sl.go = false;
        lk.unlock();
        const int64_t t_launch = gn_now_ns_impl();
        bool ok = launch_step(sl) == GN_OK;
        {
            const int64_t dt = gn_now_ns_impl() - t_launch;
            std::lock_guard<std::mutex> g(mu);
            stats.launch_ns_total += dt;
            stats.launch_ns_max = std::max(stats.launch_ns_max, dt);
        }
        if (!ok) {
            // the launch joins every side stream it forked into sl.st even when it fails; the
            // side streams are waited for as well, so no kernel of the failed step still writes
            // into result ranges that may be released and handed to new requests
            (void)hipStreamSynchronize(sl.st);
            for (int k = 0; k < gn::kSideStreams; k++) (void)hipStreamSynchronize(sl.side.s[k]);
        } else {
            // each class group's requests complete when that group's work has; the rest (and
            // everything after an error) with the step's event
            hipError_t q;
            size_t open = sl.group_done.size();
            while ((q = hipEventQuery(sl.ev)) == hipErrorNotReady) {
                for (auto &g : sl.group_done) {
                    if (!g.second) continue;
                    const hipError_t e = hipEventQuery(g.second);
                    if (e == hipErrorNotReady) continue;
                    if (e == hipSuccess) finish_group(sl, g.first);
                    g.second = nullptr;
                    open--;
                }
                if (!open) {
                    while ((q = hipEventQuery(sl.ev)) == hipErrorNotReady) std::this_thread::yield();
                    break;
                }
                std::this_thread::yield();
            }
            ok = q == hipSuccess;
        }
        finish(sl, ok);
        lk.lock();
    }
}

// The dispatcher: admission and hand-over of admitted requests to free slots.
void gn_server::run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        dirty = false;
        admit_locked();
        bool any_busy = false;
        for (Slot &sl : slots) any_busy |= sl.busy;
        if (stop && !any_busy) {  // nothing in flight: fail what never ran, stop the slots, exit
            for (auto *dq : {&waiting, &admitted})
                for (Req *r : *dq) {
                    r->state = kFailed;
                    r->done_ns = gn_now_ns_impl();
                }
            waiting.clear();
            admitted.clear();
            quit = true;
            for (Slot &sl : slots) sl.cv.notify_one();
            lk.unlock();
            cv_done.notify_all();
            return;
        }
        if (!stop && !admitted.empty()) {
            Slot *free_slot = nullptr;
            for (Slot &sl : slots)
                if (!sl.busy) {
                    free_slot = &sl;
                    break;
                }
            if (free_slot) {
                const int64_t now = gn_now_ns_impl();
                while (!admitted.empty() && (int)free_slot->reqs.size() < kMaxStepReqs) {
                    Req *r = admitted.front();
                    r->state = kRunning;
                    r->launch_ns = now;
                    free_slot->reqs.push_back(r);
                    admitted.pop_front();
                }
                free_slot->busy = true;
                free_slot->go = true;
                stats.steps++;
                stats.max_step_requests =
                    std::max<int64_t>(stats.max_step_requests, (int64_t)free_slot->reqs.size());
                free_slot->cv.notify_one();
                continue;
            }
        }
        // Sleeps, releasing mu, until a submit or a finished step marks the state dirty.  While
        // stopping with steps in flight the wait must release mu too: the slots' threads need it
        // to finish their steps (a predicate that held on `stop` alone returned at once, so this
        // loop kept mu and the slots blocked in finish: gn_server_destroy never returned).
        cv_work.wait_for(lk, std::chrono::milliseconds(2), [this] {
            if (dirty) return true;
            if (!stop) return false;
            for (const Slot &sl : slots)
                if (sl.busy) return false;
            return true;
        });
    }
}

extern "C" {

int64_t gn_now_ns(void) { return gn_now_ns_impl(); }

int gn_server_create(gn_batch *batch, gn_curve *const *curves, int ncurves,
                     const gn_invert_opts *opts, size_t hbm_budget, int streams,
                     gn_server **out) {
    if (!out) return fail(GN_EINVAL, "out is null");
    *out = nullptr;
    if (ncurves < 0 || (ncurves > 0 && !curves)) return fail(GN_EINVAL, "bad curve list");
    if (!batch && ncurves == 0) return fail(GN_EINVAL, "a server needs objects");
    if (streams < 1 || streams > 8) return fail(GN_EINVAL, "streams must be 1..8");
    gn_invert_opts o;
    if (opts) o = *opts;
    else gn_invert_opts_default(&o);
    if (o.precision != GN_FP64 && o.precision != GN_MIXED)
        return fail(GN_EINVAL, "precision must be GN_FP64 or GN_MIXED");
    if (o.seed_mode != GN_SEED_GREEDY_LEAF && o.seed_mode != GN_SEED_NEAREST_LEAF &&
        o.seed_mode != GN_SEED_GREEDY_STEP && o.seed_mode != GN_SEED_GRID_STEP)
        return fail(GN_EINVAL, "server seeds: GN_SEED_GREEDY_LEAF, GN_SEED_GREEDY_STEP, "
                               "GN_SEED_GRID_STEP or GN_SEED_NEAREST_LEAF");
    if (o.max_iter < 0 || o.fp32_iters < 0 || o.fp64_tail < 0 || o.max_iter > 10000 ||
        o.fp32_iters > 10000 || o.fp64_tail > 10000)
        return fail(GN_EINVAL, "iteration counts must be 0..10000");
    for (int i = 0; i < ncurves; i++)
        if (!curves[i]) return fail(GN_EINVAL, "null curve in the list");
    int rc = gn::ensure_runtime();
    if (rc) return rc;
    auto *s = new gn_server();
    s->o = o;
    s->budget = hbm_budget;
    GN_HIP(hipGetDevice(&s->device));
    size_t batch_bytes = 0;
    if (batch) {
        const GnBatchObj *objs = nullptr;
        bool grids = false;
        if ((rc = gn::batch_view(batch, &objs, &s->nsurf, &s->cls, &s->tile, &s->bpc, &batch_bytes,
                                 &grids))) {
            delete s;
            return rc;
        }
        if (o.seed_mode == GN_SEED_GRID_STEP && !grids) {
            delete s;
            return fail(GN_ESTATE, "grid seeds need every surface's seed grid (gn_build_seed_grid) "
                                   "built before gn_batch_create");
        }
        s->batch = batch;
        s->d_objs = const_cast<GnBatchObj *>(objs);
        s->waves = std::max(1, gn::batch_waves(s->bpc));
    }
    s->ncurve = ncurves;
    size_t curve_bytes = 0;
    std::vector<GnCurveDev> cd(ncurves);
    for (int i = 0; i < ncurves; i++) {
        cd[i] = curves[i]->dev();
        curve_bytes += curves[i]->bytes;
        s->max_samples = std::max(s->max_samples, cd[i].ns * gn::kCurveSamples);
    }
    s->resident = batch_bytes + curve_bytes + sizeof(GnCurveDev) * ncurves +
                  (size_t)streams * kStageBytes;
    if (hbm_budget < s->resident + (1u << 20)) {
        delete s;
        return fail(GN_EBUDGET, "objects leave less than 1 MiB of the HBM budget for results");
    }
    const size_t arena = (hbm_budget - s->resident) & ~(size_t)255;
    hipError_t e = hipMalloc((void **)&s->d_arena, arena);
    if (e == hipSuccess && ncurves) {
        e = hipMalloc((void **)&s->d_curves, sizeof(GnCurveDev) * ncurves);
        if (e == hipSuccess)
            e = hipMemcpy(s->d_curves, cd.data(), sizeof(GnCurveDev) * ncurves, hipMemcpyHostToDevice);
    }
    s->arena.reset(arena);
    s->stats.arena_bytes = arena;
    for (int k = 0; k < streams; k++) s->slots.emplace_back();
    for (Slot &sl : s->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void **)&sl.h_stage, kStageBytes, hipHostMallocCoherent);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&sl.h_stage_dev, sl.h_stage, 0);
        if (e == hipSuccess) e = hipMalloc((void **)&sl.d_stage, kStageBytes);
        if (e == hipSuccess && gn::side_streams_create(&sl.side) != GN_OK) e = hipErrorOutOfMemory;
    }
    if (e != hipSuccess) {
        for (Slot &sl : s->slots) {
            if (sl.st) (void)hipStreamDestroy(sl.st);
            if (sl.ev) (void)hipEventDestroy(sl.ev);
            if (sl.h_stage) (void)hipHostFree(sl.h_stage);
            if (sl.d_stage) (void)hipFree(sl.d_stage);
            gn::side_streams_destroy(&sl.side);
        }
        (void)hipFree(s->d_arena);
        (void)hipFree(s->d_curves);
        delete s;
        return gn::hip_fail(e, "server set-up");
    }
    for (Slot &sl : s->slots) {
        Slot *p = &sl;
        sl.th = std::thread([s, p] { s->run_slot(*p); });
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return GN_OK;
}

int gn_server_submit(gn_server *s, int obj, const double *x, const double *y, const double *z,
                     size_t n, int64_t *id) {
    if (!s || !id) return fail(GN_EINVAL, "null argument");
    if (obj < 0 || obj >= s->nsurf + s->ncurve) return fail(GN_ENOTFOUND, "unknown object id");
    if (n && (!x || !y || !z)) return fail(GN_EINVAL, "null point array");
    if (n > ((size_t)1 << 31)) return fail(GN_EINVAL, "at most 2^31 points per request");
    auto *r = new Req();
    r->obj = obj;
    r->curve = obj >= s->nsurf;
    r->n = n;
    r->x = x;
    r->y = y;
    r->z = z;
    r->bytes = result_bytes(r->curve, n);
    if (r->bytes > s->arena.capacity()) {
        delete r;
        return fail(GN_EBUDGET, "request results exceed the server's HBM arena");
    }
    r->submit_ns = gn_now_ns_impl();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete r;
            return fail(GN_ESTATE, "server is shutting down");
        }
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->waiting.push_back(r);
        s->stats.submitted++;
        s->dirty = true;
        *id = r->id;
    }
    s->cv_work.notify_one();
    return GN_OK;
}

static int report(const Req *r, gn_request_info *info) {
    if (r->state == kFailed) return fail(GN_EDEVICE, "request failed (step launch error or server shut down)");
    if (r->state != kDone) return GN_EAGAIN;
    if (info) {
        info->submit_ns = r->submit_ns;
        info->admit_ns = r->admit_ns;
        info->launch_ns = r->launch_ns;
        info->done_ns = r->done_ns;
        info->u = r->u;
        info->v = r->v;
        info->status = r->status;
        info->n = r->n;
        info->obj = r->obj;
    }
    return GN_OK;
}

int gn_server_poll(gn_server *s, int64_t id, gn_request_info *info) {
    if (!s) return fail(GN_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(GN_ENOTFOUND, "unknown request id");
    return report(it->second, info);
}

int gn_server_wait(gn_server *s, int64_t id, int timeout_ms, gn_request_info *info) {
    if (!s) return fail(GN_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(GN_ENOTFOUND, "unknown request id");
    Req *r = it->second;
    auto ready = [r] { return r->state == kDone || r->state == kFailed; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    return report(r, info);
}

int gn_server_release(gn_server *s, int64_t id) {
    if (!s) return fail(GN_EINVAL, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return fail(GN_ENOTFOUND, "unknown request id");
        Req *r = it->second;
        if (r->state != kDone && r->state != kFailed)
            return fail(GN_ESTATE, "request is still queued or running");
        if (r->u) s->arena.release(r->off, r->bytes);
        s->reqs.erase(it);
        delete r;
        s->dirty = true;
    }
    s->cv_work.notify_one();
    return GN_OK;
}

int gn_server_stats_get(gn_server *s, gn_server_stats *stats) {
    if (!s || !stats) return fail(GN_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *stats = s->stats;
    return GN_OK;
}

int gn_server_destroy(gn_server *s) {
    if (!s) return GN_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();  // returns once no step is in flight
    for (Slot &sl : s->slots)
        if (sl.th.joinable()) sl.th.join();
    for (Slot &sl : s->slots) {
        (void)hipStreamSynchronize(sl.st);
        (void)hipStreamDestroy(sl.st);
        (void)hipEventDestroy(sl.ev);
        (void)hipHostFree(sl.h_stage);
        (void)hipFree(sl.d_stage);
        gn::side_streams_destroy(&sl.side);
    }
    (void)hipFree(s->d_arena);
    (void)hipFree(s->d_curves);
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
    return GN_OK;
}

}  // extern "C"
