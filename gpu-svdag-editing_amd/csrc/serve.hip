// Serving: continuous batching of edit requests with slab-headroom
// admission, and per-frame primary traces that overlap the edit batches.
//
// Threads.  The caller submits requests (any thread).  The batch thread
// takes every waiting request, admits the FIFO prefix whose estimated slab
// demand fits the arenas' headroom, and runs it through run_edits (fused
// waves by default) on the context's edit stream; requests that arrive
// meanwhile wait for the next batch.  The frame thread traces the latest
// published root on its own stream every frame period.  Nodes are immutable
// once inserted, so a trace may run while a batch inserts; only arena
// reallocation and garbage collection must not overlap a trace, and both
// go through FrameQuiesce.
//
// Admission.  A request's demand is estimated from its bounds: an upper
// bound on the edit-SVO items of its surface, 6 (span/4)^2 * 4/3.  Each
// table converts items to slabs with its own measured rate (slabs the table
// gained per edit-SVO item over earlier batches), and a prefix of requests
// is admitted while every table's demand fits its headroom (free slabs
// minus a reserve fraction of its capacity).  When the head request does
// not fit, the batch thread collects garbage (if allowed) and then grows the
// arenas; the edit pipeline's own grow-and-retry remains the last line.
#include "dag_internal.h"
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>

namespace svdag {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket;
    EditorDev ed;
    double estItems;
};

// Per table: slabs in use and headroom (free - reserve * capacity).
struct TableRoom {
    std::vector<uint64_t> used;
    std::vector<double> room;
};

struct Server {
    svdag_server_config cfg {};
    std::vector<svdag_camera> cameras;
    Clock::time_point t0;
    uint64_t seqSpine0 = 0, handedBack0 = 0; // the context's counters when the server started
    std::mutex qMu;
    std::condition_variable qCv;      // new request or stop
    std::condition_variable doneCv;   // a batch finished
    std::deque<Request> queue;
    std::vector<svdag_request_info> info; // by ticket
    std::vector<uint8_t> done;
    bool stopping = false;
    bool framesStop = false;
    std::thread batchThread, frameThread;
    svdag_server_stats stats {};
    std::atomic<uint32_t> published { kSentinel };
    std::vector<double> slabsPerItem; // per table (batch thread only)
    AllocState* hAlloc = nullptr;     // pinned: every table's allocator state (batch thread only)
    AllocState* dAlloc = nullptr;     // the same, gathered on the device
    size_t hAllocCap = 0;
    // The headroom read after the last batch, valid for the next admission
    // while the root, the collections run and the tables' capacity are as
    // they were then (batch thread only).
    TableRoom lastRoom;
    bool lastRoomOk = false;
    uint32_t lastRoomRoot = 0;
    uint64_t lastRoomGc = 0, lastRoomCap = 0;
    // Frame state (frame thread only, plus frameMu for launches).
    hipEvent_t fStart = nullptr, fEnd = nullptr;
    void* dFrame = nullptr;
    bool frameInFlight = false;
    double traceMsSum = 0.0;
    uint64_t traceMsCount = 0;
    uint32_t lastFrameRoot = kSentinel;
    std::atomic<int> error { SVDAG_OK };
};

static double seconds_since(const Clock::time_point& t0)
{
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

static double estimate_items(const svdag_ctx* ctx, const EditorDev& ed)
{
    const Bounds b = edit_bounds(ctx, ed);
    double span = 0.0;
    for (int a = 0; a < 3; ++a)
        span = std::max(span, (double)(b.hi[a] - b.lo[a] + 1));
    return 6.0 * (span / 4.0) * (span / 4.0) * 4.0 / 3.0 + 64.0;
}

static int read_memory(svdag_ctx* ctx, svdag_memory& m)
{
    m = svdag_memory {};
    m.min_free_fraction = 1.0;
    for (const HostTable& t : ctx->ts.tables) {
        AllocState a {};
        SVDAG_TRY(table_read_alloc(t, a, ctx->stream));
        const uint64_t fresh = std::min<uint64_t>(a.fresh, t.slabCap);
        const uint64_t freeSlabs = (t.slabCap - fresh) + (uint64_t)std::max(a.freeCount, 0);
        m.slab_capacity += t.slabCap;
        m.slabs_free += freeSlabs;
        m.slabs_used += t.slabCap - freeSlabs;
        m.bytes_reserved += (uint64_t)t.slabCap * (kLineWords + 1) * 4; // lines and free links
        m.min_free_fraction = std::min(m.min_free_fraction, (double)freeSlabs / (double)t.slabCap);
    }
    return SVDAG_OK;
}

// Every table's allocator state in one pinned buffer: all copies queued,
// then one synchronisation (a copy and a wait per table cost about as much
// as a small batch's kernels).
// Every table's allocator state into one array: one launch and one copy
// back instead of a copy per table.
__global__ void alloc_states_gather(DagDev dag, AllocState* out)
{
    if (threadIdx.x < dag.numTables)
        out[threadIdx.x] = *dag.tables[threadIdx.x].alloc;
}

static int headroom(svdag_ctx* ctx, Server* srv, double reserveFraction, TableRoom& r)
{
    const size_t nt = ctx->ts.tables.size();
    r.used.assign(nt, 0);
    r.room.assign(nt, 0.0);
    if (!srv->hAlloc) {
        SVDAG_HIP(hipHostMalloc((void**)&srv->hAlloc, std::max<size_t>(nt, 1) * sizeof(AllocState)));
        SVDAG_HIP(hipMalloc((void**)&srv->dAlloc, std::max<size_t>(nt, 1) * sizeof(AllocState)));
        srv->hAllocCap = nt;
    }
    if (srv->hAllocCap < nt || nt > 64 || ctx->dag.numTables != nt) {
        set_error("serving: more tables than the allocator-state buffer holds");
        return SVDAG_E_INVALID;
    }
    alloc_states_gather<<<1, 64, 0, ctx->stream>>>(ctx->dag, srv->dAlloc);
    SVDAG_HIP(hipMemcpyAsync(srv->hAlloc, srv->dAlloc, nt * sizeof(AllocState), hipMemcpyDeviceToHost, ctx->stream));
    SVDAG_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < nt; ++i) {
        const HostTable& t = ctx->ts.tables[i];
        const AllocState a = srv->hAlloc[i];
        const uint64_t fresh = std::min<uint64_t>(a.fresh, t.slabCap);
        const uint64_t freeSlabs = (t.slabCap - fresh) + (uint64_t)std::max(a.freeCount, 0);
        r.used[i] = t.slabCap - freeSlabs;
        r.room[i] = (double)freeSlabs - reserveFraction * (double)t.slabCap;
    }
    return SVDAG_OK;
}

// Whether `items` edit-SVO items fit every table's headroom.
static bool fits(const Server* srv, const TableRoom& r, double items)
{
    for (size_t i = 0; i < r.room.size(); ++i)
        if (items * srv->slabsPerItem[i] > r.room[i])
            return false;
    return true;
}

static void frame_loop(svdag_ctx* ctx, Server* srv)
{
    const auto period = std::chrono::microseconds(srv->cfg.frame_period_us ? srv->cfg.frame_period_us : 16667);
    auto next = Clock::now();
    uint64_t f = 0;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(srv->qMu);
            if (srv->doneCv.wait_until(lk, next, [&] { return srv->framesStop; }))
                break;
        }
        next += period;
        std::lock_guard<std::mutex> lk(ctx->frameMu);
        if (srv->frameInFlight) {
            if (hipEventQuery(srv->fEnd) == hipErrorNotReady) {
                srv->stats.frames_skipped++;
                continue;
            }
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, srv->fStart, srv->fEnd) == hipSuccess) {
                srv->traceMsSum += ms;
                srv->traceMsCount++;
            }
            srv->frameInFlight = false;
        }
        const uint32_t root = srv->published.load();
        const svdag_camera& cam = srv->cameras[f % srv->cameras.size()];
        if (hipEventRecord(srv->fStart, ctx->frameStream) != hipSuccess
            || launch_trace(ctx, cam, srv->cfg.frame_width, srv->cfg.frame_height, root, ctx->frameStream,
                            (svdag_hit*)srv->dFrame)
                != SVDAG_OK
            || hipEventRecord(srv->fEnd, ctx->frameStream) != hipSuccess) {
            srv->error = SVDAG_E_DEVICE;
            break;
        }
        srv->frameInFlight = true;
        srv->lastFrameRoot = root;
        srv->stats.frames++;
        ++f;
    }
    std::lock_guard<std::mutex> lk(ctx->frameMu);
    if (srv->frameInFlight) {
        (void)hipEventSynchronize(srv->fEnd);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, srv->fStart, srv->fEnd) == hipSuccess) {
            srv->traceMsSum += ms;
            srv->traceMsCount++;
        }
        srv->frameInFlight = false;
    }
}

static void batch_loop(svdag_ctx* ctx, Server* srv)
{
    const uint32_t maxBatch = srv->cfg.max_batch ? srv->cfg.max_batch : 256u;
    const double reserve = srv->cfg.reserve_fraction > 0.0 ? srv->cfg.reserve_fraction : 0.02;
    std::vector<Request> batch;
    std::vector<EditorDev> eds;
    std::vector<svdag_edit_stats> st;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(srv->qMu);
            srv->qCv.wait(lk, [&] { return !srv->queue.empty() || srv->stopping; });
            if (srv->queue.empty())
                break;
            batch.clear();
            while (!srv->queue.empty() && batch.size() < maxBatch) {
                batch.push_back(srv->queue.front());
                srv->queue.pop_front();
            }
        }
        std::unique_lock<std::mutex> ctxLock(ctx->mu);
        // Admission against every table's headroom.
        TableRoom room;
        uint64_t capSum = 0;
        for (const HostTable& t : ctx->ts.tables)
            capSum += t.slabCap;
        const bool reuse = srv->lastRoomOk && srv->lastRoomRoot == ctx->root && srv->lastRoomGc == ctx->gcSerial
                           && srv->lastRoomCap == capSum;
        srv->lastRoomOk = false;
        if (srv->slabsPerItem.size() != ctx->ts.tables.size()) {
            // Before any measurement: one slot per item in every table.
            srv->slabsPerItem.resize(ctx->ts.tables.size());
            for (size_t i = 0; i < ctx->ts.tables.size(); ++i)
                srv->slabsPerItem[i] = 1.0 / (double)line_slots(ctx->ts.tables[i].itemWords);
        }
        if (reuse)
            room = srv->lastRoom;
        else
            (void)headroom(ctx, srv, reserve, room);
        const double need = batch.front().estItems;
        if (!fits(srv, room, need) && srv->cfg.allow_gc) {
            ctx->undo.clear();
            ctx->redo.clear();
            svdag_gc_stats gs {};
            FrameQuiesce q(ctx);
            if (gc_run(ctx, &gs) == SVDAG_OK)
                srv->stats.gcs++;
            (void)headroom(ctx, srv, reserve, room);
        }
        if (!fits(srv, room, need)) {
            FrameQuiesce q(ctx);
            bool grew = false;
            for (size_t i = 0; i < ctx->ts.tables.size(); ++i) {
                HostTable& t = ctx->ts.tables[i];
                if (need * srv->slabsPerItem[i] <= room.room[i])
                    continue;
                const uint64_t target = std::min<uint64_t>((uint64_t)t.slabCap * 2, ctx->slabLimit[i]);
                if (target > t.slabCap && table_grow(t, (uint32_t)target, ctx->stream) == SVDAG_OK)
                    grew = true;
            }
            if (grew && ctx_refresh_tables(ctx) == SVDAG_OK)
                srv->stats.grows++;
            (void)headroom(ctx, srv, reserve, room);
        }
        size_t admit = 1; // the head request always runs (the pipeline grows on OOM)
        double sum = need;
        while (admit < batch.size() && fits(srv, room, sum + batch[admit].estItems)) {
            sum += batch[admit].estItems;
            ++admit;
        }
        if (admit < batch.size()) {
            std::lock_guard<std::mutex> lk(srv->qMu);
            for (size_t i = batch.size(); i-- > admit;)
                srv->queue.push_front(batch[i]);
            srv->stats.deferred += batch.size() - admit;
            batch.resize(admit);
        }
        const double admitS = seconds_since(srv->t0);
        eds.resize(batch.size());
        st.assign(batch.size(), svdag_edit_stats {});
        for (size_t i = 0; i < batch.size(); ++i)
            eds[i] = batch[i].ed;
        const auto b0 = Clock::now();
        const int rc = run_edits(ctx, eds.data(), (uint32_t)eds.size(), st.data(), nullptr);
        srv->stats.edit_busy_s += std::chrono::duration<double>(Clock::now() - b0).count();
        if (rc == SVDAG_OK) {
            for (size_t i = 0; i < batch.size(); ++i)
                if (ctx->undo.empty() || ctx->undo.back() != st[i].root_before) {
                    ctx->redo.clear();
                    ctx->undo.push_back(st[i].root_before);
                }
            uint64_t items = 0;
            for (const auto& s : st)
                for (uint32_t L = 0; L < 32; ++L)
                    items += s.svo_items[L];
            TableRoom after;
            if (items > 0 && headroom(ctx, srv, reserve, after) == SVDAG_OK) {
                uint64_t capNow = 0;
                for (const HostTable& t : ctx->ts.tables)
                    capNow += t.slabCap;
                srv->lastRoom = after;
                srv->lastRoomOk = true;
                srv->lastRoomRoot = ctx->root;
                srv->lastRoomGc = ctx->gcSerial;
                srv->lastRoomCap = capNow;
                // Slabs each table gained per item; a table that gained none
                // keeps a small floor so a sudden demand still counts.
                for (size_t i = 0; i < after.used.size() && i < room.used.size(); ++i) {
                    const double gained = after.used[i] >= room.used[i] ? (double)(after.used[i] - room.used[i]) : 0.0;
                    const double r = gained / (double)items;
                    srv->slabsPerItem[i] = std::max(1e-3, 0.5 * srv->slabsPerItem[i] + 0.5 * r);
                }
            }
        }
        srv->published.store(ctx->root);
        ctxLock.unlock();
        const double doneS = seconds_since(srv->t0);
        {
            std::lock_guard<std::mutex> lk(srv->qMu);
            for (size_t i = 0; i < batch.size(); ++i) {
                svdag_request_info& o = srv->info[batch[i].ticket];
                o.status = rc;
                o.applied = rc == SVDAG_OK ? st[i].applied : 0u;
                o.root_after = rc == SVDAG_OK ? st[i].root_after : kSentinel;
                o.batch = (uint32_t)srv->stats.batches;
                o.admit_s = admitS;
                o.done_s = doneS;
                srv->done[batch[i].ticket] = 1;
            }
            srv->stats.batches++;
            srv->stats.max_batch_seen = std::max<uint32_t>(srv->stats.max_batch_seen, (uint32_t)batch.size());
            if (rc != SVDAG_OK)
                srv->error = rc;
        }
        srv->doneCv.notify_all();
    }
}

static void server_release(svdag_ctx* ctx, Server* srv)
{
    if (srv->hAlloc)
        (void)hipHostFree(srv->hAlloc);
    if (srv->dAlloc)
        (void)hipFree(srv->dAlloc);
    if (srv->fStart)
        (void)hipEventDestroy(srv->fStart);
    if (srv->fEnd)
        (void)hipEventDestroy(srv->fEnd);
    if (srv->dFrame)
        (void)hipFree(srv->dFrame);
    if (ctx->frameStream) {
        (void)hipStreamSynchronize(ctx->frameStream);
        (void)hipStreamDestroy(ctx->frameStream);
        ctx->frameStream = nullptr;
    }
    delete srv;
}

} // namespace svdag

using namespace svdag;

extern "C" {

int svdag_memory_usage(svdag_ctx* ctx, svdag_memory* out)
{
    if (!ctx || !out)
        return SVDAG_E_INVALID;
    std::lock_guard<std::mutex> lock(ctx->mu);
    return read_memory(ctx, *out);
}

int svdag_server_start(svdag_ctx* ctx, const svdag_server_config* cfg)
{
    if (!ctx || !cfg || (cfg->frame_width && (!cfg->cameras || cfg->num_cameras == 0 || cfg->frame_height == 0))
        || (uint64_t)cfg->frame_width * cfg->frame_height > (1ull << 26) || cfg->reserve_fraction < 0.0
        || cfg->reserve_fraction >= 1.0) {
        set_error("svdag_server_start: bad arguments");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->server) {
        set_error("svdag_server_start: a server is already running on this context");
        return SVDAG_E_INVALID;
    }
    if (ctx->dag.levels > 20 && cfg->frame_width) {
        set_error("svdag_server_start: frames need levels <= 20");
        return SVDAG_E_INVALID;
    }
    Server* srv = new Server();
    srv->cfg = *cfg;
    if (cfg->frame_width)
        srv->cameras.assign(cfg->cameras, cfg->cameras + cfg->num_cameras);
    srv->cfg.cameras = nullptr;
    srv->published.store(ctx->root);
    int rc = SVDAG_OK;
    if (cfg->frame_width) {
        if (hipStreamCreateWithFlags(&ctx->frameStream, hipStreamNonBlocking) != hipSuccess
            || hipEventCreate(&srv->fStart) != hipSuccess || hipEventCreate(&srv->fEnd) != hipSuccess
            || hipMalloc(&srv->dFrame, (size_t)cfg->frame_width * cfg->frame_height * sizeof(svdag_hit)) != hipSuccess)
            rc = SVDAG_E_DEVICE;
    }
    if (rc != SVDAG_OK) {
        set_error("svdag_server_start: device resources");
        server_release(ctx, srv);
        return rc;
    }
    srv->t0 = Clock::now();
    srv->seqSpine0 = ctx->seqSpineEdits;
    srv->handedBack0 = ctx->handedBackEdits;
    ctx->server = srv;
    srv->batchThread = std::thread(batch_loop, ctx, srv);
    if (cfg->frame_width)
        srv->frameThread = std::thread(frame_loop, ctx, srv);
    return SVDAG_OK;
}

int svdag_server_submit(svdag_ctx* ctx, const svdag_edit_desc* edit, uint64_t* ticket)
{
    if (!ctx || !edit || !ticket || !ctx->server) {
        set_error("svdag_server_submit: bad arguments or no server");
        return SVDAG_E_INVALID;
    }
    Server* srv = ctx->server;
    Request r {};
    SVDAG_TRY(make_editor(ctx, *edit, r.ed));
    r.estItems = estimate_items(ctx, r.ed);
    {
        std::lock_guard<std::mutex> lk(srv->qMu);
        if (srv->stopping) {
            set_error("svdag_server_submit: server is stopping");
            return SVDAG_E_INVALID;
        }
        r.ticket = srv->info.size();
        svdag_request_info info {};
        info.submit_s = seconds_since(srv->t0);
        info.root_after = kSentinel;
        srv->info.push_back(info);
        srv->done.push_back(0);
        srv->queue.push_back(r);
        srv->stats.requests++;
    }
    srv->qCv.notify_one();
    *ticket = r.ticket;
    return SVDAG_OK;
}

int svdag_server_wait(svdag_ctx* ctx, uint64_t ticket, svdag_request_info* out)
{
    if (!ctx || !ctx->server || !out) {
        set_error("svdag_server_wait: bad arguments or no server");
        return SVDAG_E_INVALID;
    }
    Server* srv = ctx->server;
    std::unique_lock<std::mutex> lk(srv->qMu);
    if (ticket >= srv->info.size()) {
        set_error("svdag_server_wait: unknown ticket");
        return SVDAG_E_INVALID;
    }
    srv->doneCv.wait(lk, [&] { return srv->done[ticket] != 0; });
    *out = srv->info[ticket];
    return SVDAG_OK;
}

int svdag_server_stop(svdag_ctx* ctx, svdag_server_stats* stats)
{
    if (!ctx || !ctx->server) {
        set_error("svdag_server_stop: no server");
        return SVDAG_E_INVALID;
    }
    Server* srv = ctx->server;
    {
        std::lock_guard<std::mutex> lk(srv->qMu);
        srv->stopping = true;
    }
    srv->qCv.notify_all();
    srv->batchThread.join();
    {
        std::lock_guard<std::mutex> lk(srv->qMu);
        srv->framesStop = true;
    }
    srv->doneCv.notify_all();
    if (srv->frameThread.joinable())
        srv->frameThread.join();
    srv->stats.wall_s = seconds_since(srv->t0);
    srv->stats.trace_ms_avg = srv->traceMsCount ? srv->traceMsSum / (double)srv->traceMsCount : 0.0;
    srv->stats.last_frame_root = srv->lastFrameRoot;
    srv->stats.sequential_spine_edits = ctx->seqSpineEdits - srv->seqSpine0;
    srv->stats.handed_back_edits = ctx->handedBackEdits - srv->handedBack0;
    if (stats)
        *stats = srv->stats;
    const int err = srv->error.load();
    // Keep the last frame for svdag_server_last_frame.
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (ctx->dLastFrame)
        (void)hipFree(ctx->dLastFrame);
    ctx->dLastFrame = srv->dFrame;
    ctx->lastFrameW = srv->cfg.frame_width;
    ctx->lastFrameH = srv->cfg.frame_height;
    ctx->lastFrameRoot = srv->lastFrameRoot;
    srv->dFrame = nullptr;
    ctx->server = nullptr;
    server_release(ctx, srv);
    if (err != SVDAG_OK) {
        set_error("svdag_server_stop: a batch or frame failed");
        return err;
    }
    return SVDAG_OK;
}

int svdag_server_last_frame(svdag_ctx* ctx, svdag_hit* out, uint32_t* root)
{
    if (!ctx || !out || ctx->server) {
        set_error("svdag_server_last_frame: bad arguments or server still running");
        return SVDAG_E_INVALID;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!ctx->lastFrameW || !ctx->dLastFrame || ctx->lastFrameRoot == kSentinel) {
        set_error("svdag_server_last_frame: no frame was traced");
        return SVDAG_E_EMPTY;
    }
    SVDAG_HIP(hipMemcpy(out, ctx->dLastFrame, (size_t)ctx->lastFrameW * ctx->lastFrameH * sizeof(svdag_hit),
                        hipMemcpyDeviceToHost));
    if (root)
        *root = ctx->lastFrameRoot;
    return SVDAG_OK;
}

int svdag_serve_replay(svdag_ctx* ctx, const svdag_server_config* cfg, const svdag_edit_desc* edits,
                       const double* arrival_s, uint32_t n, svdag_request_info* out, svdag_server_stats* stats)
{
    if (!ctx || !cfg || (n && (!edits || !arrival_s || !out))) {
        set_error("svdag_serve_replay: bad arguments");
        return SVDAG_E_INVALID;
    }
    for (uint32_t i = 0; i < n; ++i) {
        EditorDev tmp;
        SVDAG_TRY(make_editor(ctx, edits[i], tmp));
        if (!(arrival_s[i] >= 0.0) || (i && arrival_s[i] < arrival_s[i - 1])) {
            set_error("svdag_serve_replay: arrival times must be >= 0 and non-decreasing");
            return SVDAG_E_INVALID;
        }
    }
    SVDAG_TRY(svdag_server_start(ctx, cfg));
    Server* srv = ctx->server;
    std::vector<uint64_t> tickets(n);
    int rc = SVDAG_OK;
    for (uint32_t i = 0; i < n && rc == SVDAG_OK; ++i) {
        std::this_thread::sleep_until(srv->t0 + std::chrono::duration_cast<Clock::duration>(
                                                    std::chrono::duration<double>(arrival_s[i])));
        rc = svdag_server_submit(ctx, &edits[i], &tickets[i]);
    }
    for (uint32_t i = 0; i < n && rc == SVDAG_OK; ++i)
        rc = svdag_server_wait(ctx, tickets[i], &out[i]);
    const int rc2 = svdag_server_stop(ctx, stats);
    return rc != SVDAG_OK ? rc : rc2;
}

} // extern "C"
