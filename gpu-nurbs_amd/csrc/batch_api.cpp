// batch_api.cpp -- C ABI of the batched inversion: packing many surfaces into one device
// allocation, the tile list, the tile-shape autotuner and hipGraph capture of a batch launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "gn_internal.hpp"

using gn::fail;

struct gn_batch {
    int count = 0;
    std::vector<int> cls;
    GnBatchObj *d_objs = nullptr;
    char *d_pack = nullptr;
    size_t bytes = 0;
    int tile = 1024;  // points per tile
    int bpc = 0;      // workgroups per CU cap (0: occupancy limit)
    bool grids = false;  // every surface brought its seed grid (GN_SEED_GRID_STEP allowed)
};

struct gn_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    char *d_buf = nullptr;  // tile list + request table of the captured call
    // the claim-counter slots baked into the graph: its own, so no eager launch is ever given
    // them (each replay leaves them zeroed for the next)
    unsigned long long *d_ctr = nullptr;
};

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

int degree_class(int p, int q) {
    if (p >= 2 && p <= 5 && q >= 2 && q <= 5) return (p - 2) * 4 + (q - 2);
    return 16;
}

int check_segments(const gn_batch *b, const gn_segment *segs, int nseg, size_t n) {
    if (nseg < 0 || (nseg > 0 && !segs)) return fail(GN_EINVAL, "bad segment list");
    for (int i = 0; i < nseg; i++) {
        const gn_segment &s = segs[i];
        if (s.obj < 0 || s.obj >= b->count) return fail(GN_EINVAL, "segment object out of range");
        if (s.count < 0 || s.start < 0 || (size_t)s.start + (size_t)s.count > n)
            return fail(GN_EINVAL, "segment outside the point arrays");
    }
    return GN_OK;
}

// Tiles of every segment, sorted by class group, then class; each group gets its own tile size
// (at most `tile`), sized for the group's own launch.
std::vector<GnTile> build_tiles(const gn_batch *b, const gn_segment *segs, int nseg, int tile) {
    size_t per[3] = {0, 0, 0};
    for (int i = 0; i < nseg; i++) per[gn::class_group(b->cls[segs[i].obj])] += (size_t)segs[i].count;
    const int waves = gn::batch_waves(b->bpc);
    int tsz[3];
    for (int g = 0; g < 3; g++) tsz[g] = gn::class_tile(per[g], waves, tile);
    std::vector<GnTile> t;
    for (int i = 0; i < nseg; i++) {
        const int64_t T = tsz[gn::class_group(b->cls[segs[i].obj])];
        for (int64_t off = 0; off < segs[i].count; off += T) {
            const int64_t c = std::min<int64_t>(T, segs[i].count - off);
            t.push_back(GnTile{segs[i].obj, (int32_t)c, segs[i].start + off, 0, 0});
        }
    }
    std::stable_sort(t.begin(), t.end(), [b](const GnTile &x, const GnTile &y) {
        return gn::class_order(b->cls[x.obj]) < gn::class_order(b->cls[y.obj]);
    });
    return t;
}

// Device image of one batched call: the tile list, then (256-byte aligned) a one-entry request
// table holding the call's point arrays.
struct CallImage {
    std::vector<char> bytes;
    size_t ntiles = 0, req_off = 0;
    std::vector<gn::GroupRange> ranges;
};

CallImage call_image(const gn_batch *b, const std::vector<GnTile> &t, const GnReq &r) {
    CallImage c;
    c.ntiles = t.size();
    c.ranges = gn::group_ranges(t.data(), t.size(), b->cls);
    c.req_off = align256(sizeof(GnTile) * t.size());
    c.bytes.resize(c.req_off + sizeof(GnReq));
    if (!t.empty()) std::memcpy(c.bytes.data(), t.data(), sizeof(GnTile) * t.size());
    std::memcpy(c.bytes.data() + c.req_off, &r, sizeof(GnReq));
    return c;
}

void free_image_host(void *p) { delete static_cast<CallImage *>(p); }

// The call image goes to a stream-ordered device buffer; the host copy is released by a host
// callback in stream order, after the upload has consumed it.
int run_batch(gn_batch *b, const gn_segment *segs, int nseg, const double *x, const double *y,
              const double *z, size_t n, const gn_invert_opts &o, double *u, double *v,
              uint8_t *status, uint8_t *it32, uint8_t *it64, int tile, int bpc, hipStream_t st) {
    const GnReq req{x, y, z, u, v, status, it32, it64};
    auto *img = new CallImage(call_image(b, build_tiles(b, segs, nseg, tile), req));
    if (img->ntiles == 0) {
        delete img;
        return GN_OK;
    }
    char *d = nullptr;
    hipError_t e = hipMallocAsync((void **)&d, img->bytes.size(), st);
    if (e != hipSuccess) {
        delete img;
        return gn::hip_fail(e, "hipMallocAsync(tiles)");
    }
    const size_t req_off = img->req_off;
    const std::vector<gn::GroupRange> ranges = img->ranges;
    e = hipMemcpyAsync(d, img->bytes.data(), img->bytes.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipLaunchHostFunc(st, free_image_host, img);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        delete img;
        (void)hipFreeAsync(d, st);
        return gn::hip_fail(e, "tile upload");
    }
    (void)n;
    int rc = gn::launch_batch_invert(b->d_objs, (const GnTile *)d, ranges,
                                     (const GnReq *)(d + req_off), o, bpc, st);
    GN_HIP(hipFreeAsync(d, st));
    return rc;
}

int check_opts(const gn_batch *b, const gn_invert_opts &o) {
    if (o.precision != GN_FP64 && o.precision != GN_MIXED)
        return fail(GN_EINVAL, "precision must be GN_FP64 or GN_MIXED");
    if (o.seed_mode == GN_SEED_REF_FIRST_HIT)
        return fail(GN_EINVAL, "first-hit seeding is a single-surface reference mode");
    if (o.seed_mode == GN_SEED_GLOBAL)
        return fail(GN_EINVAL, "the global projection is a single-surface mode");
    if (o.seed_mode == GN_SEED_GRID_STEP && !b->grids)
        return fail(GN_ESTATE, "grid seeds need every surface's seed grid (gn_build_seed_grid) "
                               "built before gn_batch_create");
    if (o.seed_mode < GN_SEED_NEAREST_LEAF || o.seed_mode > GN_SEED_GRID_STEP)
        return fail(GN_EINVAL, "bad seed_mode");
    if (o.max_iter < 0 || o.fp32_iters < 0 || o.fp64_tail < 0 || o.max_iter > 10000 ||
        o.fp32_iters > 10000 || o.fp64_tail > 10000)
        return fail(GN_EINVAL, "iteration counts must be 0..10000");
    return GN_OK;
}

std::string signature(const gn_batch *b, const gn_segment *segs, int nseg, size_t n) {
    std::vector<double> per(17, 0.0);
    for (int i = 0; i < nseg; i++) per[b->cls[segs[i].obj]] += (double)segs[i].count;
    std::ostringstream os;
    os << "objs=" << b->count << ",n2=" << (n ? (int)std::log2((double)n) : 0);
    for (int c = 0; c < 17; c++) os << "," << (per[c] > 0 ? (int)std::log2(per[c]) : -1);
    return os.str();
}

}  // namespace

namespace gn {
// Read-only view of a batch for the inversion server.
int batch_view(const gn_batch *b, const GnBatchObj **objs, int *count, std::vector<int> *cls,
               int *tile, int *bpc, size_t *bytes, bool *grids) {
    if (!b) return fail(GN_EINVAL, "null batch");
    *grids = b->grids;
    *objs = b->d_objs;
    *count = b->count;
    *cls = b->cls;
    *tile = b->tile;
    *bpc = b->bpc;
    *bytes = b->bytes;
    return GN_OK;
}
}  // namespace gn

extern "C" {

int gn_batch_create(gn_surface *const *surfaces, int count, gn_stream stream, gn_batch **out) {
    if (!out) return fail(GN_EINVAL, "out is null");
    *out = nullptr;
    if (count <= 0 || !surfaces) return fail(GN_EINVAL, "empty surface list");
    hipStream_t st = (hipStream_t)stream;
    // layout of the pack
    struct Part {
        const void *src;
        size_t bytes;
        size_t off;
    };
    std::vector<std::vector<Part>> parts(count);
    size_t total = 0;
    for (int i = 0; i < count; i++) {
        const gn_surface *s = surfaces[i];
        if (!s) return fail(GN_EINVAL, "null surface in the list");
        if (!s->tree_levels) return fail(GN_ESTATE, "every surface needs its tree (gn_build_tree)");
        const int D = 3 * (s->tree_levels - 1);
        const size_t qnodes = (((size_t)1 << (2 * (D + 1))) - 1) / 3, leaves = (size_t)1 << (2 * D);
        const Part ps[] = {
            {s->d_ku, sizeof(double) * s->ku.size(), 0},
            {s->d_kv, sizeof(double) * s->kv.size(), 0},
            {s->d_ihu, sizeof(double) * (s->nu - s->p), 0},
            {s->d_ihv, sizeof(double) * (s->nv - s->q), 0},
            {s->d_tab64, sizeof(double) * s->tab_entries, 0},
            {s->d_tab32, sizeof(float) * s->tab_entries, 0},
            {s->d_sbox32, sizeof(float) * qnodes * 6, 0},
            {s->d_scentre, sizeof(double) * leaves * 3, 0},
            {s->d_sseed, sizeof(double) * leaves * 2, 0},
            {s->d_sref, sizeof(int32_t) * leaves, 0},
            {s->d_sleaf32, sizeof(float4) * leaves, 0},
            {s->d_srec, sizeof(float4) * leaves * kRecVec, 0},
            {s->d_grid, s->grid_bytes, 0},  // none (0 bytes) without a seed grid
        };
        for (Part p : ps) {
            p.off = total;
            total += align256(p.bytes);
            parts[i].push_back(p);
        }
    }
    gn_batch *b = new gn_batch();
    b->count = count;
    b->cls.resize(count);
    b->grids = true;
    for (int i = 0; i < count; i++) b->grids = b->grids && surfaces[i]->d_grid != nullptr;
    hipError_t e = hipMalloc((void **)&b->d_pack, total);
    if (e != hipSuccess) {
        delete b;
        return gn::hip_fail(e, "hipMalloc(batch pack)");
    }
    e = hipMalloc((void **)&b->d_objs, sizeof(GnBatchObj) * count);
    if (e != hipSuccess) {
        (void)hipFree(b->d_pack);
        delete b;
        return gn::hip_fail(e, "hipMalloc(batch objects)");
    }
    b->bytes = total + sizeof(GnBatchObj) * count;
    std::vector<GnBatchObj> objs(count);
    for (int i = 0; i < count && e == hipSuccess; i++) {
        const gn_surface *s = surfaces[i];
        for (const Part &p : parts[i])
            if (e == hipSuccess && p.bytes)
                e = hipMemcpyAsync(b->d_pack + p.off, p.src, p.bytes, hipMemcpyDeviceToDevice, st);
        char *base = b->d_pack;
        GnBatchObj &ob = objs[i];
        ob.s = s->dev();
        ob.s.ku = (const double *)(base + parts[i][0].off);
        ob.s.kv = (const double *)(base + parts[i][1].off);
        ob.s.ihu = (const double *)(base + parts[i][2].off);
        ob.s.ihv = (const double *)(base + parts[i][3].off);
        ob.s.tab64 = (const double *)(base + parts[i][4].off);
        ob.s.tab32 = (const float *)(base + parts[i][5].off);
        ob.sd = s->search();
        ob.sd.box = nullptr;
        ob.sd.box32 = (const float *)(base + parts[i][6].off);
        ob.sd.centre = (const double *)(base + parts[i][7].off);
        ob.sd.seed_uv = (const double *)(base + parts[i][8].off);
        ob.sd.ref_index = (const int32_t *)(base + parts[i][9].off);
        ob.sd.leaf32 = (const float4 *)(base + parts[i][10].off);
        ob.sd.rec = (const float4 *)(base + parts[i][11].off);
        ob.sd.cone32 = nullptr;  // normal cones serve the global projection only
        ob.sd.grid = s->d_grid ? (const void *)(base + parts[i][12].off) : nullptr;
        ob.cls = degree_class(s->p, s->q);
        ob.pad_ = 0;
        b->cls[i] = ob.cls;
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_objs, objs.data(), sizeof(GnBatchObj) * count,
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipFree(b->d_pack);
        (void)hipFree(b->d_objs);
        delete b;
        return gn::hip_fail(e, "batch packing copies");
    }
    *out = b;
    return GN_OK;
}

int gn_batch_destroy(gn_batch *b) {
    if (!b) return GN_OK;
    (void)hipDeviceSynchronize();
    (void)hipFree(b->d_pack);
    (void)hipFree(b->d_objs);
    delete b;
    return GN_OK;
}

size_t gn_batch_device_bytes(const gn_batch *b) { return b ? b->bytes : 0; }

int gn_batch_set_shape(gn_batch *b, int tile_points, int blocks_per_cu) {
    if (!b) return fail(GN_EINVAL, "null batch");
    if (tile_points < 1 || tile_points > (1 << 24) || blocks_per_cu < 0 || blocks_per_cu > 64)
        return fail(GN_EINVAL, "tile_points 1..2^24, blocks_per_cu 0..64");
    b->tile = tile_points;
    b->bpc = blocks_per_cu;
    return GN_OK;
}

int gn_batch_get_shape(const gn_batch *b, int *tile_points, int *blocks_per_cu) {
    if (!b) return fail(GN_EINVAL, "null batch");
    if (tile_points) *tile_points = b->tile;
    if (blocks_per_cu) *blocks_per_cu = b->bpc;
    return GN_OK;
}

int gn_batch_invert(gn_batch *b, const gn_segment *segs, int nseg, const double *x,
                    const double *y, const double *z, size_t n, const gn_invert_opts *opts,
                    double *u, double *v, uint8_t *status, gn_stream stream) {
    if (!b) return fail(GN_EINVAL, "null batch");
    gn_invert_opts o;
    if (opts) o = *opts;
    else gn_invert_opts_default(&o);
    int rc = check_opts(b, o);
    if (rc) return rc;
    if ((rc = check_segments(b, segs, nseg, n))) return rc;
    if (n && (!x || !y || !z || !u || !v || !status)) return fail(GN_EINVAL, "null buffer");
    return run_batch(b, segs, nseg, x, y, z, n, o, u, v, status, nullptr, nullptr, b->tile, b->bpc,
                     (hipStream_t)stream);
}

int gn_batch_tune(gn_batch *b, const gn_segment *segs, int nseg, const double *x, const double *y,
                  const double *z, size_t n, const gn_invert_opts *opts, double *u, double *v,
                  uint8_t *status, const char *cache_path, gn_stream stream) {
    if (!b) return fail(GN_EINVAL, "null batch");
    gn_invert_opts o;
    if (opts) o = *opts;
    else gn_invert_opts_default(&o);
    int rc = check_opts(b, o);
    if (rc) return rc;
    if ((rc = check_segments(b, segs, nseg, n))) return rc;
    if (n && (!x || !y || !z || !u || !v || !status)) return fail(GN_EINVAL, "null buffer");
    const std::string sig = signature(b, segs, nseg, n);
    if (cache_path) {  // cached shape for this signature?
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string s;
            int t, c;
            if (ls >> s >> t >> c && s == sig && t > 0 && c >= 0) {
                b->tile = t;
                b->bpc = c;
                return GN_OK;
            }
        }
    }
    hipStream_t st = (hipStream_t)stream;
    // given seeds live in u and v, which every run overwrites: keep a copy and put it back
    // before each run (outside the timed region)
    double *seeds = nullptr;
    const bool given = o.seed_mode == GN_SEED_GIVEN && n > 0;
    if (given) {
        GN_HIP(hipMallocAsync((void **)&seeds, sizeof(double) * 2 * n, st));
        GN_HIP(hipMemcpyAsync(seeds, u, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        GN_HIP(hipMemcpyAsync(seeds + n, v, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    }
    hipEvent_t e0, e1;
    GN_HIP(hipEventCreate(&e0));
    GN_HIP(hipEventCreate(&e1));
    const int tiles[] = {128, 256, 512, 1024, 4096};
    const int caps[] = {0, 1};
    float best = 1e30f;
    int bt = b->tile, bc = b->bpc;
    for (int t : tiles) {
        for (int c : caps) {
            float sum = 0;
            for (int r = 0; r < 3 && !rc; r++) {  // first run warms up, not counted
                if (given) {
                    GN_HIP(hipMemcpyAsync(u, seeds, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
                    GN_HIP(hipMemcpyAsync(v, seeds + n, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
                }
                GN_HIP(hipEventRecord(e0, st));
                rc = run_batch(b, segs, nseg, x, y, z, n, o, u, v, status, nullptr, nullptr, t, c, st);
                GN_HIP(hipEventRecord(e1, st));
                GN_HIP(hipEventSynchronize(e1));
                float ms = 0;
                GN_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (r) sum += ms;
            }
            if (rc) break;
            if (sum > 0 && sum < best) {
                best = sum;
                bt = t;
                bc = c;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (given) GN_HIP(hipFreeAsync(seeds, st));
    if (rc) return rc;
    b->tile = bt;
    b->bpc = bc;
    if (cache_path) {
        std::ofstream outf(cache_path, std::ios::app);
        outf << sig << " " << bt << " " << bc << "\n";
    }
    return GN_OK;
}

int gn_batch_graph_create(gn_batch *b, const gn_segment *segs, int nseg, const double *x,
                          const double *y, const double *z, size_t n, const gn_invert_opts *opts,
                          double *u, double *v, uint8_t *status, gn_stream stream,
                          gn_graph **out) {
    if (!b || !out) return fail(GN_EINVAL, "null argument");
    *out = nullptr;
    if (!stream) return fail(GN_EINVAL, "graph capture needs a created (non-null) stream");
    gn_invert_opts o;
    if (opts) o = *opts;
    else gn_invert_opts_default(&o);
    int rc = check_opts(b, o);
    if (rc) return rc;
    if ((rc = check_segments(b, segs, nseg, n))) return rc;
    if (n && (!x || !y || !z || !u || !v || !status)) return fail(GN_EINVAL, "null buffer");
    if ((rc = gn::ensure_runtime())) return rc;
    if ((rc = gn::batch_prepare())) return rc;
    const GnReq req{x, y, z, u, v, status, nullptr, nullptr};
    const CallImage img = call_image(b, build_tiles(b, segs, nseg, b->tile), req);
    gn_graph *g = new gn_graph();
    hipStream_t st = (hipStream_t)stream;
    // one slot per persistent launch: one per class group and precision phase
    gn::SlotPool pool;
    pool.nslots = (int)img.ranges.size() * (o.precision == GN_MIXED ? 2 : 1);
    const size_t ctr_bytes = sizeof(unsigned long long) * gn::slot_words() * (pool.nslots ? pool.nslots : 1);
    hipError_t e = hipMalloc((void **)&g->d_buf, img.bytes.size());
    if (e == hipSuccess)
        e = hipMemcpy(g->d_buf, img.bytes.data(), img.bytes.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&g->d_ctr, ctr_bytes);
    if (e == hipSuccess) e = hipMemset(g->d_ctr, 0, ctr_bytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(g->d_buf);
        (void)hipFree(g->d_ctr);
        delete g;
        return gn::hip_fail(e, "graph tile list / counter slots");
    }
    pool.ctr = g->d_ctr;
    e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipFree(g->d_buf);
        (void)hipFree(g->d_ctr);
        delete g;
        return gn::hip_fail(e, "hipStreamBeginCapture");
    }
    {
        gn::PoolScope scope(&pool);
        rc = gn::launch_batch_invert(b->d_objs, (const GnTile *)g->d_buf, img.ranges,
                                     (const GnReq *)(g->d_buf + img.req_off), o, b->bpc, st);
    }
    e = hipStreamEndCapture(st, &g->graph);
    if (!rc && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (rc || e != hipSuccess) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        (void)hipFree(g->d_buf);
        (void)hipFree(g->d_ctr);
        delete g;
        return rc ? rc : gn::hip_fail(e, "graph capture / instantiate");
    }
    *out = g;
    return GN_OK;
}

int gn_graph_launch(gn_graph *g, gn_stream stream) {
    if (!g) return fail(GN_EINVAL, "null graph");
    GN_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return GN_OK;
}

int gn_graph_destroy(gn_graph *g) {
    if (!g) return GN_OK;
    (void)hipDeviceSynchronize();
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    (void)hipFree(g->d_buf);
    (void)hipFree(g->d_ctr);
    delete g;
    return GN_OK;
}

}  // extern "C"
