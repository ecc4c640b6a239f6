/*
 * gm_server.cpp -- persistent inversion service (gmilos.h, "serving loop").
 *
 * One worker thread drives one engine per registered setup.  An engine owns a
 * device-resident pool of LM state slots (the same structure-of-arrays state
 * as gm_invert, plus the observed and last synthesised spectra per slot) and
 * one stream.  The worker repeats:
 *
 *   retire   observe finished rounds (event query), hand completed tiles to
 *            the results queue, free their slot ranges and HBM footprint;
 *   cancel   drop cancelled tiles from the queue; make cancelled tiles in
 *            flight leave the LM loop at the next round;
 *   admit    examine queued tiles in arrival order and admit each one that
 *            fits -- the summed footprint of in-flight tiles stays within the
 *            budget and its engine has a free contiguous slot range -- so a
 *            tile that does not fit yet (a large one, or one whose engine is
 *            full) does not hold back later tiles of other sizes and setups;
 *            once the oldest waiting tile has waited aging_ms, nothing passes
 *            it and it enters as soon as room frees.  Admission copies the
 *            observations into the pool, sets the start state, evaluates the
 *            first normal equations and appends the slots to the engine's
 *            active list;
 *   round    one LM round over the engine's whole active list (solve, synth
 *            + chi^2, accept/reject, normal equations of accepted pixels),
 *            then copy pixels that left the loop into their tiles' outputs;
 *            the last pixel of a tile marks the tile complete.
 *
 * Up to two rounds per engine are queued ahead of the host so the device
 * does not idle while the worker observes the previous one.  A pixel's LM
 * trajectory depends only on its own data, so results are identical to
 * gm_invert's whatever else shares the launches.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "gm_internal.h"
#include "gmilos.h"

namespace {

constexpr int kDepth = 2;   /* rounds queued per engine */

struct Pending {
    uint64_t ticket;
    int32_t setup;
    const float *obs;
    int64_t npix;
    const double *init;
    gm_out out;
    double t_submit;
};

struct TileSlot {
    bool used = false;
    bool cancelled = false;
    uint64_t ticket = 0;
    int64_t slot0 = 0, npix = 0, bytes = 0;
    double t_submit = 0, t_admit = 0;
};

struct Range {
    int64_t lo, n;
};

struct Engine {
    gm_ctx *ctx = nullptr;
    const gm_dsetup *dS = nullptr;
    gm_lanecfg lc{};
    int nl = 0, mixed = 0, solver = 0, init_mode = 0, cordic10 = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[kDepth] = {nullptr, nullptr};
    int32_t *completed_host[kDepth] = {nullptr, nullptr};   /* pinned */
    int queued = 0, head = 0;   /* rounds in flight; ring position of the oldest */
    int64_t cap = 0, bytes_per_px = 0;
    void *ws = nullptr;
    gm_state st{};
    float *obs = nullptr, *synth = nullptr;
    int32_t *slot_tile = nullptr, *newlist = nullptr, *done = nullptr, *cnt = nullptr;
    int32_t *completed_dev = nullptr;
    gm_tile_dev *tiles_dev = nullptr, *tiles_host = nullptr;   /* host mirror pinned */
    int max_tiles = 0;
    std::vector<TileSlot> tiles;
    std::vector<int> free_tiles;
    std::vector<Range> free_ranges;   /* sorted by lo, coalesced */
    int cur = 0;
    int64_t inflight_px = 0;
    int ntiles = 0;
};

bool alloc_range(Engine &e, int64_t n, int64_t &lo)
{
    for (size_t i = 0; i < e.free_ranges.size(); ++i) {
        Range &r = e.free_ranges[i];
        if (r.n >= n) {
            lo = r.lo;
            r.lo += n;
            r.n -= n;
            if (r.n == 0) e.free_ranges.erase(e.free_ranges.begin() + (long)i);
            return true;
        }
    }
    return false;
}

void free_range(Engine &e, int64_t lo, int64_t n)
{
    auto it = std::lower_bound(e.free_ranges.begin(), e.free_ranges.end(), lo,
                               [](const Range &r, int64_t v) { return r.lo < v; });
    it = e.free_ranges.insert(it, Range{lo, n});
    size_t i = (size_t)(it - e.free_ranges.begin());
    if (i + 1 < e.free_ranges.size() && e.free_ranges[i].lo + e.free_ranges[i].n == e.free_ranges[i + 1].lo) {
        e.free_ranges[i].n += e.free_ranges[i + 1].n;
        e.free_ranges.erase(e.free_ranges.begin() + (long)i + 1);
    }
    if (i > 0 && e.free_ranges[i - 1].lo + e.free_ranges[i - 1].n == e.free_ranges[i].lo) {
        e.free_ranges[i - 1].n += e.free_ranges[i].n;
        e.free_ranges.erase(e.free_ranges.begin() + (long)i);
    }
}

} // namespace

struct gm_server {
    gm_server_config cfg{};
    std::vector<Engine *> engines;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Pending> pending;
    std::deque<gm_tile_result> results;
    std::vector<uint64_t> cancel_req;   /* in-flight tickets to cancel */
    uint64_t next_ticket = 1;
    bool stop = false;
    gm_status error = GM_OK;
    std::thread worker;
    std::chrono::steady_clock::time_point t0;
    /* under mu */
    int64_t inflight_bytes = 0, peak_bytes = 0, inflight_tiles = 0;
    int64_t tiles_done = 0, pixels_done = 0, rounds = 0;
};

static double now_ms(const gm_server *s)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - s->t0).count();
}

extern "C" double gm_server_clock_ms(const gm_server *s) { return s ? now_ms(s) : 0.0; }

static void destroy_engine(Engine *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (int k = 0; k < kDepth; ++k) {
        if (e->ev[k]) (void)hipEventDestroy(e->ev[k]);
        if (e->completed_host[k]) (void)hipHostFree(e->completed_host[k]);
    }
    if (e->tiles_host) (void)hipHostFree(e->tiles_host);
    if (e->ws) (void)hipFree(e->ws);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->ctx) gm_ctx_destroy(e->ctx);
    delete e;
}

/* Pool layout: LM state SoA, observed and synthesised spectra, slot->tile
 * map, two scratch lists, counters, tile table, completed list. */
static gm_status build_engine(gm_server *srv, const gm_setup *setup, Engine **out)
{
    Engine *e = new (std::nothrow) Engine();
    if (!e) return GM_ERR_NOMEM;
    gm_status st = gm_ctx_create(setup, &e->ctx);
    if (st != GM_OK) { delete e; return st; }
    e->dS = gm_ctx_device_setup(e->ctx);
    e->lc = gm_ctx_lanes(e->ctx);
    const gm_setup &ps = gm_ctx_public_setup(e->ctx);
    e->nl = ps.nlambda;
    e->mixed = ps.precision == GM_PREC_MIXED;
    e->solver = ps.solver;
    e->cordic10 = gm_cordic10(ps, e->lc.np);
    e->init_mode = ps.init_mode;
    e->max_tiles = srv->cfg.max_tiles;
    const int64_t nl4 = 4 * (int64_t)e->nl;
    e->bytes_per_px = (int64_t)(sizeof(double) * (2 * GM_MODEL_REC + 5) +
                                sizeof(float) * (GM_AB_REC11 + 2 * GM_PRE_REC) +
                                sizeof(int32_t) * (2 + 4 + 3) + sizeof(float) * 2 * nl4);
    int64_t cap = srv->cfg.pool_pixels;
    if (cap <= 0) cap = std::min<int64_t>(srv->cfg.hbm_budget_bytes / e->bytes_per_px, (int64_t)1 << 22);
    cap = ((cap + 255) / 256) * 256;
    if (cap < 256 || cap > 0x7fffffff) { destroy_engine(e); return GM_ERR_INVALID; }
    e->cap = cap;
    auto rnd = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t bytes = 0;
    bytes += rnd(sizeof(double) * GM_MODEL_REC * cap) * 2;
    bytes += rnd(sizeof(float) * GM_AB_REC11 * cap);
    bytes += rnd(sizeof(float) * GM_PRE_REC * cap) * 2;
    bytes += rnd(sizeof(double) * cap) * 5 + rnd(sizeof(int32_t) * cap) * 2;
    bytes += rnd(sizeof(int32_t) * cap) * 4;                       /* lists      */
    bytes += rnd(sizeof(float) * nl4 * cap) * 2;                   /* obs, synth */
    bytes += rnd(sizeof(int32_t) * cap) * 3;                       /* slot_tile, newlist, done */
    bytes += rnd(sizeof(int32_t) * 8);
    bytes += rnd(sizeof(gm_tile_dev) * e->max_tiles) + rnd(sizeof(int32_t) * (e->max_tiles + 1));
    if (hipMalloc(&e->ws, bytes) != hipSuccess) { e->ws = nullptr; destroy_engine(e); return GM_ERR_NOMEM; }
    char *p = (char *)e->ws;
    auto take = [&](size_t b) { char *r = p; p += rnd(b); return (void *)r; };
    gm_state &s = e->st;
    s.cap = cap;
    s.best = (double *)take(sizeof(double) * GM_MODEL_REC * cap);
    s.trial = (double *)take(sizeof(double) * GM_MODEL_REC * cap);
    const bool p11 = e->lc.np == 11;
    s.ab_rec = p11 ? GM_AB_REC11 : GM_AB_REC9;
    s.alpha = (float *)take(sizeof(float) * GM_AB_REC11 * cap);
    s.beta = s.alpha + (p11 ? GM_AB_BETA11 : GM_AB_BETA9);
    s.pre_best = (float *)take(sizeof(float) * GM_PRE_REC * cap);
    s.pre_trial = (float *)take(sizeof(float) * GM_PRE_REC * cap);
    s.chi_trial = (double *)take(sizeof(double) * cap);
    s.ochisqr = (double *)take(sizeof(double) * cap);
    s.chisqr0 = (double *)take(sizeof(double) * cap);
    s.flambda = (double *)take(sizeof(double) * cap);
    s.factor = (double *)take(sizeof(double) * cap);
    s.iter = (int32_t *)take(sizeof(int32_t) * cap);
    s.nacc = (int32_t *)take(sizeof(int32_t) * cap);
    for (int i = 0; i < 4; ++i) s.list[i] = (int32_t *)take(sizeof(int32_t) * cap);
    e->obs = (float *)take(sizeof(float) * nl4 * cap);
    e->synth = (float *)take(sizeof(float) * nl4 * cap);
    e->slot_tile = (int32_t *)take(sizeof(int32_t) * cap);
    e->newlist = (int32_t *)take(sizeof(int32_t) * cap);
    e->done = (int32_t *)take(sizeof(int32_t) * cap);
    e->cnt = (int32_t *)take(sizeof(int32_t) * 8);
    s.counts = e->cnt;
    s.hist = nullptr;
    e->tiles_dev = (gm_tile_dev *)take(sizeof(gm_tile_dev) * e->max_tiles);
    e->completed_dev = (int32_t *)take(sizeof(int32_t) * (e->max_tiles + 1));
    if (hipHostMalloc((void **)&e->tiles_host, sizeof(gm_tile_dev) * e->max_tiles) != hipSuccess) {
        e->tiles_host = nullptr;
        destroy_engine(e);
        return GM_ERR_NOMEM;
    }
    for (int k = 0; k < kDepth; ++k) {
        if (hipHostMalloc((void **)&e->completed_host[k], sizeof(int32_t) * (e->max_tiles + 1)) != hipSuccess) {
            e->completed_host[k] = nullptr;
            destroy_engine(e);
            return GM_ERR_NOMEM;
        }
        if (hipEventCreateWithFlags(&e->ev[k], hipEventDisableTiming) != hipSuccess) {
            e->ev[k] = nullptr;
            destroy_engine(e);
            return GM_ERR_HIP;
        }
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        e->stream = nullptr;
        destroy_engine(e);
        return GM_ERR_HIP;
    }
    /* counters: both active lists empty */
    gm_launch_fill_i32(e->cnt, 8, 0, e->stream);
    if (hipStreamSynchronize(e->stream) != hipSuccess) { destroy_engine(e); return GM_ERR_HIP; }
    e->tiles.resize(e->max_tiles);
    for (int t = e->max_tiles - 1; t >= 0; --t) e->free_tiles.push_back(t);
    e->free_ranges.push_back(Range{0, cap});
    *out = e;
    return GM_OK;
}

/* Stream-ordered admission of one tile into engine e. */
static void admit(Engine &e, const Pending &pd, int t, int64_t slot0)
{
    hipStream_t s = e.stream;
    const int64_t nl4 = 4 * (int64_t)e.nl;
    gm_tile_dev &th = e.tiles_host[t];
    th.model = pd.out.model;
    th.chisqr = pd.out.chisqr;
    th.iters = pd.out.iters;
    th.naccs = pd.out.naccs;
    th.synth = pd.out.synth;
    th.slot0 = slot0;
    th.remaining = (int32_t)pd.npix;
    th.pad = 0;
    (void)hipMemcpyAsync(e.tiles_dev + t, &th, sizeof(gm_tile_dev), hipMemcpyHostToDevice, s);
    (void)hipMemcpyAsync(e.obs + slot0 * nl4, pd.obs, sizeof(float) * nl4 * pd.npix,
                         hipMemcpyDeviceToDevice, s);
    gm_launch_fill_i32(e.slot_tile + slot0, pd.npix, (int32_t)t, s);
    gm_launch_init_state(e.dS, e.st, e.obs, pd.init, e.init_mode, e.nl, slot0, pd.npix, e.newlist,
                         s);
    gm_launch_set_count(&e.cnt[4], (int32_t)pd.npix, s);
    gm_launch_rf(e.lc, e.dS, e.st.best, e.st.pre_best, -GM_MODEL_REC, e.obs, e.newlist, &e.cnt[4], 0, pd.npix, e.st.alpha,
                 e.st.beta, -e.st.ab_rec, e.st.ochisqr, e.synth, nullptr, e.mixed, s);
    gm_launch_init_chi(e.st, slot0, pd.npix, s);
    gm_launch_append_range(e.st.list[e.cur], &e.cnt[e.cur], slot0, pd.npix, s);
    e.inflight_px += pd.npix;
    e.ntiles += 1;
}

/* k_decide's workgroups in the serving loop: 256 threads (1) or the
 * inversion's 1024 (0). */
#ifndef GM_SERVE_DECIDE_SMALL
#define GM_SERVE_DECIDE_SMALL 1
#endif

/* One LM round over the engine's active list (cf. gm_invert). */
static void launch_round(Engine &e)
{
    hipStream_t s = e.stream;
    gm_state st = e.st;
    int32_t *cnt = e.cnt;
    const int cur = e.cur, nxt = 1 - cur;
    const int64_t grid = std::max<int64_t>(e.inflight_px, 1);
    const int slot = (e.head + e.queued) % kDepth;
    gm_launch_set_count(&cnt[5], 0, s);
    gm_launch_set_count(e.completed_dev, 0, s);
    gm_launch_solve(e.dS, st, st.list[cur], &cnt[cur], grid, e.mixed, e.solver, &cnt[nxt], &cnt[2],
                    nullptr, st.list[3], &cnt[3], s, e.lc.np, e.cordic10);
    gm_launch_synth(e.lc, e.dS, st.trial, st.pre_trial, -GM_MODEL_REC, e.obs, st.list[cur], &cnt[cur], 0, grid,
                    st.chi_trial, e.synth, e.mixed, s);
    gm_launch_decide(e.dS, st, st.list[cur], &cnt[cur], grid, st.list[nxt], &cnt[nxt], st.list[2],
                     &cnt[2], &cnt[3], e.done, &cnt[5], e.mixed, s, 1, nullptr, GM_SERVE_DECIDE_SMALL);
    gm_launch_rf(e.lc, e.dS, st.trial, st.pre_trial, -GM_MODEL_REC, e.obs, st.list[2], &cnt[2], 0, grid,
                 st.alpha, st.beta, -st.ab_rec, nullptr, nullptr, nullptr, e.mixed, s, 0, st.best,
                 st.pre_best);
    gm_launch_retire(st, e.done, &cnt[5], grid, e.synth, 4 * e.nl, e.slot_tile, e.tiles_dev,
                     e.completed_dev, s);
    (void)hipMemcpyAsync(e.completed_host[slot], e.completed_dev, sizeof(int32_t) * (e.max_tiles + 1),
                         hipMemcpyDeviceToHost, s);
    (void)hipEventRecord(e.ev[slot], s);
    e.cur = nxt;
    e.queued += 1;
}

static void fail_all(gm_server *srv, gm_status err)
{
    /* under mu */
    srv->error = err;
    const double t = now_ms(srv);
    for (const Pending &pd : srv->pending) {
        gm_tile_result r{pd.ticket, pd.setup, (int32_t)err, pd.npix, pd.t_submit, t, t};
        srv->results.push_back(r);
    }
    srv->pending.clear();
    for (size_t ei = 0; ei < srv->engines.size(); ++ei)
        for (TileSlot &ts : srv->engines[ei]->tiles) {
            if (!ts.used) continue;
            gm_tile_result r{ts.ticket, (int32_t)ei, (int32_t)err, ts.npix, ts.t_submit, ts.t_admit, t};
            srv->results.push_back(r);
            ts.used = false;
        }
    srv->inflight_tiles = 0;
    srv->inflight_bytes = 0;
    srv->cv_done.notify_all();
}

static void worker_main(gm_server *srv)
{
    if (hipSetDevice(srv->cfg.device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(srv->mu);
        fail_all(srv, GM_ERR_HIP);
        return;
    }
    for (;;) {
        bool progressed = false;
        std::vector<Engine *> engines;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            engines = srv->engines;
        }
        /* retire */
        for (size_t ei = 0; ei < engines.size(); ++ei) {
            Engine &e = *engines[ei];
            while (e.queued > 0) {
                const hipError_t q = hipEventQuery(e.ev[e.head]);
                if (q == hipErrorNotReady) break;
                if (q != hipSuccess) {
                    std::lock_guard<std::mutex> lk(srv->mu);
                    fail_all(srv, GM_ERR_HIP);
                    return;
                }
                const int32_t *c = e.completed_host[e.head];
                const int n = std::min<int>(c[0], e.max_tiles);
                const double t = now_ms(srv);
                std::lock_guard<std::mutex> lk(srv->mu);
                for (int k = 0; k < n; ++k) {
                    const int tid = c[1 + k];
                    if (tid < 0 || tid >= e.max_tiles || !e.tiles[tid].used) continue;
                    TileSlot &ts = e.tiles[tid];
                    gm_tile_result r{ts.ticket, (int32_t)ei, ts.cancelled ? GM_CANCELLED : GM_OK,
                                     ts.npix, ts.t_submit, ts.t_admit, t};
                    srv->results.push_back(r);
                    srv->inflight_bytes -= ts.bytes;
                    srv->inflight_tiles -= 1;
                    srv->tiles_done += 1;
                    srv->pixels_done += ts.npix;
                    free_range(e, ts.slot0, ts.npix);
                    e.inflight_px -= ts.npix;
                    e.ntiles -= 1;
                    ts.used = false;
                    ts.cancelled = false;
                    e.free_tiles.push_back(tid);
                }
                if (n) srv->cv_done.notify_all();
                e.head = (e.head + 1) % kDepth;
                e.queued -= 1;
                progressed = true;
            }
        }
        /* cancel: in-flight tiles leave the LM loop at their next decision
         * (iteration count past the cap); stream order puts this after the
         * rounds already queued and before any reuse of the slots */
        {
            std::vector<std::pair<Engine *, TileSlot>> todo;
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                for (uint64_t tk : srv->cancel_req)
                    for (Engine *ep : engines)
                        for (TileSlot &ts : ep->tiles)
                            if (ts.used && ts.ticket == tk && !ts.cancelled) {
                                ts.cancelled = true;
                                todo.push_back({ep, ts});
                            }
                srv->cancel_req.clear();
            }
            for (auto &c : todo) {
                gm_launch_fill_i32(c.first->st.iter + c.second.slot0, c.second.npix,
                                   0x3fffffff, c.first->stream);
                progressed = true;
            }
        }
        /* admit: every queued tile that fits, in arrival order, within the
         * first `scan` of the queue; a tile that does not fit is passed by
         * later ones until it has waited aging_ms */
        for (;;) {
            Pending pd;
            int t = -1;
            int64_t slot0 = 0;
            Engine *ep = nullptr;
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                if (srv->pending.empty()) break;
                const double tnow = now_ms(srv);
                const int scan = srv->cfg.scan;
                size_t pick = srv->pending.size();
                for (size_t q = 0; q < srv->pending.size() && (int)q < scan; ++q) {
                    const Pending &c = srv->pending[q];
                    Engine &e = *srv->engines[c.setup];
                    const int64_t bytes = c.npix * e.bytes_per_px;
                    const bool fits = (srv->inflight_bytes == 0 ||
                                       srv->inflight_bytes + bytes <= srv->cfg.hbm_budget_bytes) &&
                                      !e.free_tiles.empty() && alloc_range(e, c.npix, slot0);
                    if (fits) { pick = q; break; }
                    /* strict order, or the oldest tile has waited long enough:
                     * nothing passes it */
                    if (srv->cfg.aging_ms < 0 || tnow - c.t_submit >= srv->cfg.aging_ms) break;
                }
                if (pick == srv->pending.size()) break;
                pd = srv->pending[pick];
                Engine &e = *srv->engines[pd.setup];
                const int64_t bytes = pd.npix * e.bytes_per_px;
                t = e.free_tiles.back();
                e.free_tiles.pop_back();
                TileSlot &ts = e.tiles[t];
                ts.used = true;
                ts.cancelled = false;
                ts.ticket = pd.ticket;
                ts.slot0 = slot0;
                ts.npix = pd.npix;
                ts.bytes = bytes;
                ts.t_submit = pd.t_submit;
                ts.t_admit = tnow;
                srv->inflight_bytes += bytes;
                srv->peak_bytes = std::max(srv->peak_bytes, srv->inflight_bytes);
                srv->inflight_tiles += 1;
                srv->pending.erase(srv->pending.begin() + (long)pick);
                ep = &e;
            }
            admit(*ep, pd, t, slot0);
            progressed = true;
        }
        /* rounds */
        for (Engine *ep : engines) {
            Engine &e = *ep;
            if (e.ntiles > 0 && e.queued < kDepth) {
                launch_round(e);
                std::lock_guard<std::mutex> lk(srv->mu);
                srv->rounds += 1;
                progressed = true;
            }
        }
        if (hipGetLastError() != hipSuccess) {
            std::lock_guard<std::mutex> lk(srv->mu);
            fail_all(srv, GM_ERR_HIP);
            return;
        }
        if (progressed) continue;
        bool busy = false;
        for (Engine *ep : engines) busy = busy || ep->queued > 0;
        if (busy) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            continue;
        }
        /* idle: a queued tile always fits an idle engine (submit bounds its
         * size by the pool), so the next pass admits whatever arrived */
        std::unique_lock<std::mutex> lk(srv->mu);
        if (srv->stop && srv->pending.empty() && srv->inflight_tiles == 0) return;
        if (!srv->pending.empty() || !srv->cancel_req.empty()) continue;
        srv->cv_work.wait_for(lk, std::chrono::milliseconds(5));
    }
}

extern "C" gm_status gm_server_create(const gm_server_config *cfg, gm_server **out)
{
    if (!out) return GM_ERR_INVALID;
    gm_server *s = new (std::nothrow) gm_server();
    if (!s) return GM_ERR_NOMEM;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.hbm_budget_bytes <= 0) s->cfg.hbm_budget_bytes = (int64_t)8 << 30;
    if (s->cfg.max_tiles <= 0) s->cfg.max_tiles = 1024;
    if (s->cfg.aging_ms == 0) s->cfg.aging_ms = 1000;
    if (s->cfg.scan <= 0) s->cfg.scan = 64;
    if (s->cfg.pool_pixels < 0 || s->cfg.max_tiles > (1 << 20)) { delete s; return GM_ERR_INVALID; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || s->cfg.device < 0 || s->cfg.device >= ndev) {
        delete s;
        return GM_ERR_HIP;
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread(worker_main, s);
    *out = s;
    return GM_OK;
}

extern "C" gm_status gm_server_add_setup(gm_server *s, const gm_setup *setup, int32_t *id)
{
    if (!s || !setup || !id) return GM_ERR_INVALID;
    if (hipSetDevice(s->cfg.device) != hipSuccess) return GM_ERR_HIP;
    Engine *e = nullptr;
    gm_status st = build_engine(s, setup, &e);
    if (st != GM_OK) return st;
    std::lock_guard<std::mutex> lk(s->mu);
    s->engines.push_back(e);
    *id = (int32_t)s->engines.size() - 1;
    return GM_OK;
}

extern "C" gm_status gm_server_submit(gm_server *s, int32_t setup_id, const float *obs, int64_t npix,
                                      const double *init, const gm_out *out, uint64_t *ticket)
{
    if (!s || !out || npix <= 0 || !obs) return GM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return GM_ERR_INVALID;
    if (s->error != GM_OK) return s->error;
    if (setup_id < 0 || setup_id >= (int32_t)s->engines.size()) return GM_ERR_INVALID;
    const Engine &e = *s->engines[setup_id];
    if (npix > e.cap) return GM_ERR_INVALID;
    if (e.init_mode == GM_INIT_PER_PIXEL && !init) return GM_ERR_INVALID;
    Pending pd{s->next_ticket++, setup_id, obs, npix, init, *out, now_ms(s)};
    s->pending.push_back(pd);
    if (ticket) *ticket = pd.ticket;
    s->cv_work.notify_one();
    return GM_OK;
}

extern "C" gm_status gm_server_cancel(gm_server *s, uint64_t ticket)
{
    if (!s) return GM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto it = s->pending.begin(); it != s->pending.end(); ++it)
        if (it->ticket == ticket) {
            const double t = now_ms(s);
            gm_tile_result r{it->ticket, it->setup, (int32_t)GM_CANCELLED, it->npix, it->t_submit,
                             t, t};
            s->pending.erase(it);
            s->results.push_back(r);
            s->cv_done.notify_all();
            return GM_OK;
        }
    for (Engine *e : s->engines)
        for (const TileSlot &ts : e->tiles)
            if (ts.used && ts.ticket == ticket) {
                if (ts.cancelled) return GM_ERR_NOT_FOUND;
                for (uint64_t q : s->cancel_req)
                    if (q == ticket) return GM_ERR_NOT_FOUND;
                s->cancel_req.push_back(ticket);
                s->cv_work.notify_one();
                return GM_OK;
            }
    return GM_ERR_NOT_FOUND;
}

extern "C" gm_status gm_server_poll(gm_server *s, gm_tile_result *res, int32_t max, int32_t *n,
                                    int32_t wait_ms)
{
    if (!s || !n || max < 0 || (max > 0 && !res)) return GM_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] {
        return !s->results.empty() || (s->pending.empty() && s->inflight_tiles == 0);
    };
    if (s->results.empty() && wait_ms != 0) {
        if (wait_ms < 0)
            s->cv_done.wait(lk, ready);
        else
            s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
    }
    int k = 0;
    while (k < max && !s->results.empty()) {
        res[k++] = s->results.front();
        s->results.pop_front();
    }
    *n = k;
    return GM_OK;
}

extern "C" gm_status gm_server_get_stats(gm_server *s, gm_server_stats *o)
{
    if (!s || !o) return GM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    o->tiles_done = s->tiles_done;
    o->pixels_done = s->pixels_done;
    o->lm_rounds = s->rounds;
    o->pending_tiles = (int64_t)s->pending.size();
    o->inflight_tiles = s->inflight_tiles;
    o->inflight_bytes = s->inflight_bytes;
    o->peak_inflight_bytes = s->peak_bytes;
    return s->error;
}

extern "C" gm_status gm_server_destroy(gm_server *s)
{
    if (!s) return GM_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->cfg.device);
    for (Engine *e : s->engines) destroy_engine(e);
    delete s;
    return GM_OK;
}
