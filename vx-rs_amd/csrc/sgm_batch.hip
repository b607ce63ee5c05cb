// Batched SGM: many pairs of uneven size and disparity range replayed from
// one hipGraph.
//
// The reference runs one pair per command buffer (vx-vision/src/kernels/
// sgm.rs:135-261) and its Pipeline batches encodes of different kernels into
// one command buffer (vx-vision/src/pipeline.rs:12-80).  Two forms:
//
// Persistent (default).  Pairs, largest first, are grouped into chunks whose
// workspaces fit a byte budget; chunks run one after the other in one shared
// workspace.  A chunk is four steps, each one launch over all of its pairs:
// the census of every image, the cost volume of every pair, the path
// aggregation of every direction of every pair, and the sum + winner-take-all
// of every pair (plus the LR check of the pairs that ask for it).  The
// aggregation is one persistent launch per lane-layout class and volume
// format present (launched side by side): its waves take (pass, scanline
// group) items by ticket from one list per XCD, ordered pair by pair, so a
// single chunk fills the chip with waves from many pairs instead of the 1.7
// waves per SIMD one 720p pair provides.  Split scanlines are not used (they
// shorten one pair's critical path, which a full chip does not need).
//
// Streams.  The per-pair pipeline of every pair (the plain path's exact
// launches, including split scanlines) is captured over several streams:
// pairs are ordered by work and each dealt to the least loaded stream; the
// streams run side by side, and pairs on one stream run in turn and share
// that stream's workspace.
//
// Both forms run the plain path's kernels or their multi-pair wrappers around
// the same per-tile, per-scanline and per-pixel device code, with each pair's
// own schedule and volume format, so every map is bit-identical to the plain
// path's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sgm_internal.h"
#include "sgm_pipeline.h"
#include "vxh/sgm.h"
#include "vxh/sgm_batch.h"

namespace {

thread_local std::string g_batch_error;

vxh_status bfail(vxh_status st, const std::string &m) {
    g_batch_error = m;
    return st;
}

constexpr uint32_t kDefaultStreams = 4;
constexpr uint32_t kMaxStreams = 16;
// Default chunk budget: this share (in tenths) of the device memory free when
// the batch is created, at least kMinBudget (kFallbackBudget if the query
// fails).  Fewer, larger chunks run faster: the 64-pair mixture (133.5 GiB of
// workspace) took 41.9 / 39.9 / 39.4 / 38.9 ms per replay at budgets of 64 /
// 96 / 128 / 160 GiB (3 / 2 / 2 / 1 chunks).
constexpr uint64_t kBudgetTenths = 6;
constexpr uint64_t kMinBudget = (uint64_t)1 << 30;
constexpr uint64_t kFallbackBudget = (uint64_t)64 << 30;
constexpr uint32_t kFormPersistent = 0, kFormStreams = 1, kFormSweep = 2;

// Scratch of one pair of the sweep form, byte offsets from its base: the
// census of both images, the cost volume (u8, pix*Dp + d), the forward sums
// (u16, same layout), the waves' rings, and for the LR check the summed
// volume and the winners.
struct SweepLayout {
    size_t lc, rc, cost, sf, ring, sum, best, total;
};
SweepLayout sweep_layout(uint32_t w, uint32_t h, uint32_t D, bool lr) {
    using namespace vxh;
    SweepLayout l;
    const size_t n = (size_t)w * h;
    const size_t Dp = (size_t)volume_stride((int)D);
    l.lc = 0;
    l.rc = l.lc + ws_align(n * 8);
    l.cost = l.rc + ws_align(n * 8);
    l.sf = l.cost + ws_align(n * Dp);
    l.ring = l.sf + ws_align(n * Dp * 2);
    l.sum = l.ring + ws_align((size_t)sweep_waves((int)h, (int)D) * sweep_ring_slots((int)w) *
                              sweep_slot_bytes((int)D));
    l.best = l.sum + (lr ? ws_align(n * Dp * 2) : 0);
    l.total = l.best + (lr ? ws_align(n * 2) : 0);
    return l;
}

// Index of (L, K) among the layouts of class c, -1 if none.
int layout_index(int c, uint32_t L, uint32_t K) {
    for (int i = 0; i < vxh::kClassCount[c]; ++i)
        if ((uint32_t)vxh::kClassLayouts[c][i][0] == L &&
            (uint32_t)vxh::kClassLayouts[c][i][1] == K)
            return i;
    return -1;
}

struct PairPlan {
    vxh::Schedule sched;
    int mode;
    size_t ws_bytes;
    uint32_t stream;   // streams form: stream; persistent form: chunk
    size_t ws_off;     // persistent form: offset in the shared workspace
};

// One chunk of the persistent form: its pairs and the device tables of its
// launches (pointers into the batch's table allocation).
struct Chunk {
    std::vector<uint32_t> pairs;
    size_t ws_bytes = 0;
    const vxh::CensusJob *census = nullptr;
    int ncensus = 0;
    uint32_t census_blocks = 0;
    const vxh::CostJob *cost = nullptr;
    int ncost = 0;
    uint32_t cost_blocks = 0;
    // One aggregation launch per lane-layout class and volume format, each
    // followed on its own stream by the sums + winners of its pairs.
    struct Group {
        int cls, mode;
        vxh::BatchPaths args;
        int blocks;
        const vxh::WtaJob *wta;
        int nwta;
        uint32_t wta_blocks;
        int wta_kinds;   // as for launch_wta_multi
    };
    std::vector<Group> groups;
    // Sweep form: per lane-layout class present, the forward and backward
    // sweep launches.
    struct Sweep {
        int cls;
        vxh::SweepArgs fwd, bwd;
        int blocks_fwd, blocks_bwd;
    };
    std::vector<Sweep> sweeps;
};

int sweep_occupancy(int cls, bool bwd) {
    return cls == 0 ? vxh::sweep_occupancy_c0(bwd)
                    : cls == 1 ? vxh::sweep_occupancy_c1(bwd) : vxh::sweep_occupancy_c2(bwd);
}

hipError_t launch_sweep(int cls, bool bwd, const vxh::SweepArgs &a, int blocks, hipStream_t s) {
    return cls == 0 ? vxh::launch_sweep_c0(bwd, a, blocks, s)
                    : cls == 1 ? vxh::launch_sweep_c1(bwd, a, blocks, s)
                               : vxh::launch_sweep_c2(bwd, a, blocks, s);
}

int class_occupancy(int cls, int mode) {
    return cls == 0 ? vxh::batch_paths_occupancy_c0(mode)
                    : cls == 1 ? vxh::batch_paths_occupancy_c1(mode)
                               : vxh::batch_paths_occupancy_c2(mode);
}

hipError_t launch_class(int cls, int mode, const vxh::BatchPaths &b, int blocks, hipStream_t s) {
    return cls == 0 ? vxh::launch_batch_paths_c0(mode, b, blocks, s)
                    : cls == 1 ? vxh::launch_batch_paths_c1(mode, b, blocks, s)
                               : vxh::launch_batch_paths_c2(mode, b, blocks, s);
}

// Appends `bytes` (aligned to 256) to a host staging image of the device
// tables; returns its offset.
size_t stage(std::vector<uint8_t> &img, const void *src, size_t bytes) {
    const size_t off = vxh::ws_align(img.size());
    img.resize(off + vxh::ws_align(bytes));
    std::memcpy(img.data() + off, src, bytes);
    return off;
}

}  // namespace

struct vxh_sgm_batch {
    int device = 0;
    uint32_t form = kFormPersistent;
    std::vector<vxh_sgm_pair_desc> pairs;
    std::vector<PairPlan> plan;
    std::vector<void *> ws;              // streams: one per stream; persistent: one shared
    void *tables = nullptr;              // persistent / sweeps: job tables, items, counters
    uint32_t *counters = nullptr;        // ticket counters (sweeps: and progress words)
    size_t ncounters = 0;                // words from `counters` cleared at every replay
    std::vector<Chunk> chunks;
    uint32_t *d_error = nullptr;         // OR of every pair's error word
    hipStream_t stream = nullptr;        // the batch's own stream (capture origin)
    std::vector<hipStream_t> sub;        // streams the pairs (or aggregation groups) run on
    std::vector<hipEvent_t> events;      // fork + one join per sub stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t launches = 0;
};

namespace {

// Builds the chunks of the persistent form, allocates the shared workspace
// (b->ws[0]) and the device tables of every chunk's launches (b->tables: job
// tables, passes, item lists, ticket counters), and copies the tables.
hipError_t plan_persistent(vxh_sgm_batch *b, const std::vector<uint32_t> &order,
                           uint64_t budget, uint32_t blocks_per_cu) {
    using namespace vxh;
    // Chunks: pairs in order, a new chunk when the budget would be exceeded.
    Chunk cur;
    for (uint32_t i : order) {
        const size_t wbytes = ws_align(b->plan[i].ws_bytes);
        if (!cur.pairs.empty() && cur.ws_bytes + wbytes > budget) {
            b->chunks.push_back(cur);
            cur = Chunk();
        }
        b->plan[i].ws_off = cur.ws_bytes;
        b->plan[i].stream = (uint32_t)b->chunks.size();
        cur.pairs.push_back(i);
        cur.ws_bytes += wbytes;
    }
    b->chunks.push_back(cur);
    size_t ws_total = 256;
    for (const Chunk &c : b->chunks) ws_total = std::max(ws_total, c.ws_bytes);
    void *ws = nullptr;
    hipError_t e = hipMalloc(&ws, ws_total);
    if (e != hipSuccess) return e;
    b->ws.push_back(ws);

    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) return e;
    const int cus = prop.multiProcessorCount;

    // Host image of the tables; offsets into it become device pointers once
    // it is allocated.
    std::vector<uint8_t> img;
    struct Off {
        size_t census, cost;
        std::vector<size_t> passes, items, wta;
    };
    std::vector<Off> off(b->chunks.size());
    size_t ncounters = 0;
    for (size_t ci = 0; ci < b->chunks.size(); ++ci) {
        Chunk &c = b->chunks[ci];
        std::vector<CensusJob> cj;
        std::vector<CostJob> kj;
        uint32_t cb = 0, kb = 0;
        struct G {
            int cls, mode;
            std::vector<BatchPass> passes;
            std::vector<BatchItem> items[8];
            std::vector<WtaJob> wta;   // blocks numbered within the group
            uint32_t wb = 0;
        };
        std::vector<G> groups;
        for (uint32_t i : c.pairs) {
            const vxh_sgm_pair_desc &q = b->pairs[i];
            const PairPlan &p = b->plan[i];
            const int w = (int)q.width, h = (int)q.height, D = (int)q.cfg.num_disparities;
            const int Dp = volume_stride(D);
            const WsLayout L = workspace_layout(q.width, q.height, (uint32_t)D,
                                                (int)q.cfg.num_paths, p.mode,
                                                q.cfg.lr_check != 0);
            char *base = (char *)ws + p.ws_off;
            const bool u8 = p.mode == kOutU8;
            // Census of both images; the left one also clears the flags and
            // the synchronisation words, as run_pipeline does.
            CensusJob l{q.d_left, (uint2 *)(base + L.lc),
                        u8 ? (uint2 *)(base + L.flags) : nullptr, (uint32_t *)(base + L.sync),
                        sync_words(q.width, q.height), w, h, (int)q.cfg.census_radius_x,
                        (int)q.cfg.census_radius_y, (w + 63) / 64, cb};
            cb += (uint32_t)(l.nbx * ((h + kCensusTileH - 1) / kCensusTileH));
            CensusJob r = l;
            r.img = q.d_right;
            r.out = (uint2 *)(base + L.rc);
            r.zero_flags = nullptr;
            r.zero_words = nullptr;
            r.nwords = 0;
            r.first_block = cb;
            cb += (uint32_t)(r.nbx * ((h + kCensusTileH - 1) / kCensusTileH));
            cj.push_back(l);
            cj.push_back(r);
            const uint32_t p1 = (uint32_t)q.cfg.p1 & 0xFFFFu, p2 = (uint32_t)q.cfg.p2 & 0xFFFFu;
            CostJob k{(const uint2 *)(base + L.lc), (const uint2 *)(base + L.rc),
                      (uint8_t *)(base + L.cost), w, h, D, Dp, cost_bias(p1),
                      (w + 255) / 256, kb};
            kb += (uint32_t)(k.nbx * h);
            kj.push_back(k);
            // The group of the pair's layout class and volume format: its
            // direction passes (longest scanlines first) and its winner job.
            const int cls = layout_class(D);
            size_t gi = groups.size();
            for (size_t x = 0; x < groups.size(); ++x)
                if (groups[x].cls == cls && groups[x].mode == p.mode) gi = x;
            if (gi == groups.size()) groups.push_back(G{cls, p.mode, {}, {}, {}, 0});
            G &g = groups[gi];
            const bool lr = q.cfg.lr_check != 0;
            // LR check: the right view's keys in the left census's area (the
            // winner launch comes after the costs).
            WtaJob t{(const uint8_t *)(base + L.vol), L.vol_bytes,
                     u8 ? (const uint8_t *)(base + L.hi) : nullptr,
                     u8 ? (const uint2 *)(base + L.flags) : nullptr, q.d_out,
                     lr ? (int16_t *)(base + L.best) : nullptr,
                     lr ? (uint32_t *)(base + L.lc) : nullptr, (int)q.cfg.num_paths, p.mode,
                     w * h, w, D, Dp, g.wb};
            g.wb += wta_job_blocks(t);
            g.wta.push_back(t);
            PassParams pp{};
            pp.cost = (const uint8_t *)(base + L.cost);
            pp.flags = u8 ? (uint8_t *)(base + L.flags) : nullptr;
            pp.w = w;
            pp.h = h;
            pp.D = D;
            pp.Dp = Dp;
            pp.p1x2 = p1 | (p1 << 16);
            pp.p2x2 = p2 | (p2 << 16);
            const uint32_t xq = p1 - cost_bias(p1);
            pp.xqx2 = xq | (xq << 16);
            pp.mode = p.mode;
            pp.wait_polls = kSplitWaitPolls;
            int dirs[kMaxPasses];
            const int P = (int)q.cfg.num_paths;
            for (int j = 0; j < P; ++j) dirs[j] = j;
            std::stable_sort(dirs, dirs + P, [&](int a, int z) {
                return scan_length(a, w, h) > scan_length(z, w, h);
            });
            for (int j = 0; j < P; ++j) {
                const int dir = dirs[j];
                BatchPass bp{pp, 0};
                bp.p.vol = (uint8_t *)(base + L.vol + (size_t)dir * L.vol_bytes);
                bp.p.hi = u8 ? (uint8_t *)(base + L.hi + (size_t)dir * L.vol_bytes) : nullptr;
                bp.p.dir = dir;
                bp.p.dx = kDirX[dir];
                bp.p.dy = kDirY[dir];
                bp.p.nscan = num_scanlines(dir, w, h);
                bp.layout = bp.p.dy == 0 ? p.sched.horiz : p.sched.other;
                const int per_wave = 128 / kClassLayouts[cls][bp.layout][0];
                const uint32_t nw = (uint32_t)((bp.p.nscan + per_wave - 1) / per_wave);
                const uint32_t pi = (uint32_t)g.passes.size();
                g.passes.push_back(bp);
                for (uint32_t xk = 0; xk < 8; ++xk)
                    for (uint32_t v = nw * xk / 8; v < nw * (xk + 1) / 8; ++v)
                        g.items[xk].push_back(BatchItem{pi, v});
            }
        }
        c.ncensus = (int)cj.size();
        c.census_blocks = cb;
        c.ncost = (int)kj.size();
        c.cost_blocks = kb;
        off[ci].census = stage(img, cj.data(), cj.size() * sizeof(CensusJob));
        off[ci].cost = stage(img, kj.data(), kj.size() * sizeof(CostJob));
        for (G &g : groups) {
            int kinds = 0;
            for (const WtaJob &t : g.wta) kinds |= wta_job_kind(t);
            Chunk::Group cg{g.cls, g.mode, {}, 0, nullptr, (int)g.wta.size(), g.wb, kinds};
            std::vector<BatchItem> all;
            uint32_t most = 0;
            for (int xk = 0; xk < 8; ++xk) {
                cg.args.item_off[xk] = (uint32_t)all.size();
                cg.args.item_cnt[xk] = (uint32_t)g.items[xk].size();
                most = std::max(most, cg.args.item_cnt[xk]);
                all.insert(all.end(), g.items[xk].begin(), g.items[xk].end());
            }
            off[ci].passes.push_back(
                stage(img, g.passes.data(), g.passes.size() * sizeof(BatchPass)));
            off[ci].items.push_back(stage(img, all.data(), all.size() * sizeof(BatchItem)));
            off[ci].wta.push_back(stage(img, g.wta.data(), g.wta.size() * sizeof(WtaJob)));
            // Persistent grid: as many blocks as the chip holds at once (or
            // blocks_per_cu per CU), no more than the items need, and a
            // multiple of 8 so every XCD's list has waves.
            int per_cu = std::max(1, class_occupancy(g.cls, g.mode));
            if (blocks_per_cu) per_cu = std::min(per_cu, (int)blocks_per_cu);
            const int need = 8 * (int)((most + 3) / 4);
            cg.blocks = std::max(8, std::min(need, (per_cu * cus + 7) / 8 * 8));
            cg.args.counters = nullptr;
            c.groups.push_back(cg);
            ncounters += 8;
        }
    }
    const size_t counters_off = ws_align(img.size());
    img.resize(counters_off + ws_align(std::max<size_t>(ncounters, 1) * 4), 0);
    if ((e = hipMalloc(&b->tables, img.size())) != hipSuccess) return e;
    char *tb = (char *)b->tables;
    b->counters = (uint32_t *)(tb + counters_off);
    b->ncounters = ncounters;
    size_t next_counter = 0;
    for (size_t ci = 0; ci < b->chunks.size(); ++ci) {
        Chunk &c = b->chunks[ci];
        c.census = (const CensusJob *)(tb + off[ci].census);
        c.cost = (const CostJob *)(tb + off[ci].cost);
        for (size_t g = 0; g < c.groups.size(); ++g) {
            BatchPaths &a = c.groups[g].args;
            a.passes = (const BatchPass *)(tb + off[ci].passes[g]);
            a.items = (const BatchItem *)(tb + off[ci].items[g]);
            c.groups[g].wta = (const WtaJob *)(tb + off[ci].wta[g]);
            a.counters = b->counters + next_counter;
            next_counter += 8;
        }
    }
    return hipMemcpy(b->tables, img.data(), img.size(), hipMemcpyHostToDevice);
}

// Records the persistent form's launches on b->stream (being captured): the
// counters and the error word cleared, then per chunk census, costs, the
// aggregation groups side by side on the sub streams, each followed on its
// stream by the sums + winners of its pairs (a group that finishes early
// takes its winners while the others still aggregate), and the LR checks.
hipError_t capture_persistent(vxh_sgm_batch *b) {
    using namespace vxh;
    hipError_t e = launch_zero(b->d_error, 1, b->stream);
    if (e == hipSuccess && b->ncounters) e = launch_zero(b->counters, b->ncounters, b->stream);
    b->launches = 2;
    for (const Chunk &c : b->chunks) {
        if (e == hipSuccess) e = launch_census_multi(c.census, c.ncensus, c.census_blocks, b->stream);
        if (e == hipSuccess) e = launch_cost_multi(c.cost, c.ncost, c.cost_blocks, b->stream);
        b->launches += 2;
        // The right view's keys of the pairs with the LR check start at zero.
        for (uint32_t i : c.pairs) {
            const vxh_sgm_pair_desc &q = b->pairs[i];
            if (!q.cfg.lr_check || e != hipSuccess) continue;
            const WsLayout L = workspace_layout(q.width, q.height, q.cfg.num_disparities,
                                                (int)q.cfg.num_paths, b->plan[i].mode, true);
            char *base = (char *)b->ws[0] + b->plan[i].ws_off;
            e = launch_zero((uint32_t *)(base + L.lc), (size_t)q.width * q.height, b->stream);
            b->launches += 1;
        }
        if (e == hipSuccess && c.groups.size() == 1) {
            const Chunk::Group &g = c.groups[0];
            e = launch_class(g.cls, g.mode, g.args, g.blocks, b->stream);
            if (e == hipSuccess) e = launch_wta_multi(g.wta, g.nwta, g.wta_blocks, g.wta_kinds, g.mode,
                                                      b->stream);
            b->launches += 1 + __builtin_popcount((unsigned)g.wta_kinds);
        } else if (e == hipSuccess) {
            // Groups side by side: fork onto the sub streams, join back.
            e = hipEventRecord(b->events[0], b->stream);
            for (size_t g = 0; g < c.groups.size() && e == hipSuccess; ++g) {
                hipStream_t s = b->sub[g];
                e = hipStreamWaitEvent(s, b->events[0], 0);
                const Chunk::Group &cg = c.groups[g];
                if (e == hipSuccess) e = launch_class(cg.cls, cg.mode, cg.args, cg.blocks, s);
                if (e == hipSuccess) e = launch_wta_multi(cg.wta, cg.nwta, cg.wta_blocks, cg.wta_kinds,
                                                          cg.mode, s);
                if (e == hipSuccess) e = hipEventRecord(b->events[1 + g], s);
                if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->events[1 + g], 0);
                b->launches += 1 + __builtin_popcount((unsigned)cg.wta_kinds);
            }
        }
        for (uint32_t i : c.pairs) {
            const vxh_sgm_pair_desc &q = b->pairs[i];
            if (!q.cfg.lr_check || e != hipSuccess) continue;
            const PairPlan &p = b->plan[i];
            const WsLayout L = workspace_layout(q.width, q.height, q.cfg.num_disparities,
                                                (int)q.cfg.num_paths, p.mode, true);
            char *base = (char *)b->ws[0] + p.ws_off;
            e = launch_lr_apply((const int16_t *)(base + L.best), (const uint32_t *)(base + L.lc),
                                (int)q.width, (int)q.height, (int)q.cfg.lr_tolerance, q.d_out,
                                b->stream);
            b->launches += 1;
        }
    }
    return e;
}

// The sweep form's plan: chunks as the persistent form; per chunk the census
// and cost job tables, and per lane-layout class the forward and backward
// job tables and one item list per XCD.  Each class's pairs are dealt to the
// XCD lists by work (largest first onto the least loaded list), and a list
// holds its pairs' waves in wave order (wave 0 of every pair, then wave 1,
// ...), so a wave's predecessor always holds an earlier ticket of the same
// list and is running before it waits on it.  Counters and progress words
// follow each other at the end of the tables and are cleared per replay.
hipError_t plan_sweep(vxh_sgm_batch *b, const std::vector<uint32_t> &order, uint64_t budget) {
    using namespace vxh;
    Chunk cur;
    for (uint32_t i : order) {
        const size_t wbytes = ws_align(b->plan[i].ws_bytes);
        if (!cur.pairs.empty() && cur.ws_bytes + wbytes > budget) {
            b->chunks.push_back(cur);
            cur = Chunk();
        }
        b->plan[i].ws_off = cur.ws_bytes;
        b->plan[i].stream = (uint32_t)b->chunks.size();
        cur.pairs.push_back(i);
        cur.ws_bytes += wbytes;
    }
    b->chunks.push_back(cur);
    size_t ws_total = 256;
    for (const Chunk &c : b->chunks) ws_total = std::max(ws_total, c.ws_bytes);
    void *ws = nullptr;
    hipError_t e = hipMalloc(&ws, ws_total);
    if (e != hipSuccess) return e;
    b->ws.push_back(ws);

    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) return e;
    const int cus = prop.multiProcessorCount;

    std::vector<uint8_t> img;
    // A zeroed ring slot (the row above every first wave): 64 lanes.
    const size_t fresh_off = stage(img, std::vector<uint8_t>(64 * kSweepRec, 0).data(),
                                   64 * kSweepRec);
    struct Off {
        size_t census, cost;
        std::vector<size_t> fwd, bwd, items;
    };
    std::vector<Off> off(b->chunks.size());
    // Counter / progress words: per chunk and class 16 counters, then per
    // pair 2 * nwaves progress words (forward, backward).
    struct Words {
        size_t counters;
        std::vector<size_t> prog;   // per job of the class
    };
    std::vector<std::vector<Words>> words(b->chunks.size());
    size_t nwords = 0;
    for (size_t ci = 0; ci < b->chunks.size(); ++ci) {
        Chunk &c = b->chunks[ci];
        std::vector<CensusJob> cj;
        std::vector<CostJob> kj;
        uint32_t cb = 0, kb = 0;
        std::vector<SweepJob> jobs[kLayoutClasses];
        for (uint32_t i : c.pairs) {
            const vxh_sgm_pair_desc &q = b->pairs[i];
            const PairPlan &p = b->plan[i];
            const int w = (int)q.width, h = (int)q.height, D = (int)q.cfg.num_disparities;
            const int Dp = volume_stride(D);
            const bool lr = q.cfg.lr_check != 0;
            const SweepLayout L = sweep_layout(q.width, q.height, (uint32_t)D, lr);
            char *base = (char *)ws + p.ws_off;
            CensusJob l{q.d_left, (uint2 *)(base + L.lc), nullptr, nullptr, 0, w, h,
                        (int)q.cfg.census_radius_x, (int)q.cfg.census_radius_y, (w + 63) / 64,
                        cb};
            cb += (uint32_t)(l.nbx * ((h + kCensusTileH - 1) / kCensusTileH));
            CensusJob r = l;
            r.img = q.d_right;
            r.out = (uint2 *)(base + L.rc);
            r.first_block = cb;
            cb += (uint32_t)(r.nbx * ((h + kCensusTileH - 1) / kCensusTileH));
            cj.push_back(l);
            cj.push_back(r);
            const uint32_t p1 = (uint32_t)q.cfg.p1 & 0xFFFFu, p2 = (uint32_t)q.cfg.p2 & 0xFFFFu;
            CostJob k{(const uint2 *)(base + L.lc), (const uint2 *)(base + L.rc),
                      (uint8_t *)(base + L.cost), w, h, D, Dp, cost_bias(p1),
                      (w + 255) / 256, kb};
            kb += (uint32_t)(k.nbx * h);
            kj.push_back(k);
            SweepJob j{};
            j.cost = (const uint8_t *)(base + L.cost);
            j.sf = (uint16_t *)(base + L.sf);
            j.ring = (uint8_t *)(base + L.ring);
            j.prog = nullptr;   // set once the tables are placed
            j.out = q.d_out;
            j.best = lr ? (int16_t *)(base + L.best) : nullptr;
            j.sum = lr ? (uint16_t *)(base + L.sum) : nullptr;
            j.w = w;
            j.h = h;
            j.D = D;
            j.Dp = Dp;
            j.p1x2 = p1 | (p1 << 16);
            j.p2x2 = p2 | (p2 << 16);
            const uint32_t xq = p1 - cost_bias(p1);
            j.xqx2 = xq | (xq << 16);
            j.nwaves = sweep_waves(h, D);
            j.slot_bytes = (uint32_t)sweep_slot_bytes(D);
            const int cls = layout_class(D);
            jobs[cls].push_back(j);
        }
        c.ncensus = (int)cj.size();
        c.census_blocks = cb;
        c.ncost = (int)kj.size();
        c.cost_blocks = kb;
        off[ci].census = stage(img, cj.data(), cj.size() * sizeof(CensusJob));
        off[ci].cost = stage(img, kj.data(), kj.size() * sizeof(CostJob));
        for (int cls = 0; cls < kLayoutClasses; ++cls) {
            if (jobs[cls].empty()) continue;
            // Deal the pairs (largest first: the chunk's order) to the XCDs.
            const size_t nj = jobs[cls].size();
            std::vector<int> xcd_of(nj);
            uint64_t load[8] = {};
            for (size_t x = 0; x < nj; ++x) {
                const SweepJob &j = jobs[cls][x];
                const int k = (int)(std::min_element(load, load + 8) - load);
                xcd_of[x] = k;
                load[k] += (uint64_t)j.w * j.h * j.D;
            }
            Chunk::Sweep sw{};
            sw.cls = cls;
            std::vector<SweepItem> all;
            int most_waves = 0;
            for (const SweepJob &j : jobs[cls]) most_waves = std::max(most_waves, j.nwaves);
            uint32_t most = 0;
            for (int k = 0; k < 8; ++k) {
                sw.fwd.item_off[k] = (uint32_t)all.size();
                for (int v = 0; v < most_waves; ++v)
                    for (size_t x = 0; x < nj; ++x)
                        if (xcd_of[x] == k && v < jobs[cls][x].nwaves)
                            all.push_back(SweepItem{(uint32_t)x, (uint32_t)v});
                sw.fwd.item_cnt[k] = (uint32_t)all.size() - sw.fwd.item_off[k];
                most = std::max(most, sw.fwd.item_cnt[k]);
            }
            Words wd;
            wd.counters = nwords;
            nwords += 16;
            for (const SweepJob &j : jobs[cls]) {
                wd.prog.push_back(nwords);
                nwords += 2 * (size_t)j.nwaves;
            }
            words[ci].push_back(wd);
            off[ci].fwd.push_back(stage(img, jobs[cls].data(), nj * sizeof(SweepJob)));
            off[ci].bwd.push_back(stage(img, jobs[cls].data(), nj * sizeof(SweepJob)));
            off[ci].items.push_back(stage(img, all.data(), all.size() * sizeof(SweepItem)));
            // Persistent grid: as many one-wave blocks as the chip holds, no
            // more than the largest list needs on every XCD.
            const int need = 8 * (int)most;
            sw.blocks_fwd = std::max(8, std::min(need, std::max(1, sweep_occupancy(cls, false)) * cus));
            sw.blocks_bwd = std::max(8, std::min(need, std::max(1, sweep_occupancy(cls, true)) * cus));
            for (int k = 0; k < 8; ++k) {
                sw.bwd.item_off[k] = sw.fwd.item_off[k];
                sw.bwd.item_cnt[k] = sw.fwd.item_cnt[k];
            }
            c.sweeps.push_back(sw);
        }
    }
    const size_t counters_off = ws_align(img.size());
    img.resize(counters_off + ws_align(std::max<size_t>(nwords, 1) * 4), 0);
    // The progress pointers are known now: patch the staged job tables.
    char *tb = nullptr;
    if ((e = hipMalloc(&b->tables, img.size())) != hipSuccess) return e;
    tb = (char *)b->tables;
    b->counters = (uint32_t *)(tb + counters_off);
    b->ncounters = nwords;
    for (size_t ci = 0; ci < b->chunks.size(); ++ci) {
        Chunk &c = b->chunks[ci];
        c.census = (const CensusJob *)(tb + off[ci].census);
        c.cost = (const CostJob *)(tb + off[ci].cost);
        for (size_t x = 0; x < c.sweeps.size(); ++x) {
            Chunk::Sweep &sw = c.sweeps[x];
            const Words &wd = words[ci][x];
            SweepJob *hf = (SweepJob *)(img.data() + off[ci].fwd[x]);
            SweepJob *hb = (SweepJob *)(img.data() + off[ci].bwd[x]);
            for (size_t j = 0; j < wd.prog.size(); ++j) {
                hf[j].prog = b->counters + wd.prog[j];
                hb[j].prog = b->counters + wd.prog[j] + hf[j].nwaves;
            }
            for (SweepArgs *a : {&sw.fwd, &sw.bwd}) {
                a->jobs = (const SweepJob *)(tb + (a == &sw.fwd ? off[ci].fwd[x] : off[ci].bwd[x]));
                a->items = (const SweepItem *)(tb + off[ci].items[x]);
                a->counters = b->counters + wd.counters + (a == &sw.fwd ? 0 : 8);
                a->error = b->d_error;
                a->fresh = (const uint8_t *)(tb + fresh_off);
                a->wait_polls = kSweepWaitPolls;
                a->zero = 0;
            }
        }
    }
    return hipMemcpy(b->tables, img.data(), img.size(), hipMemcpyHostToDevice);
}

// Records the sweep form's launches on b->stream (being captured): the error
// word, counters and progress words cleared, then per chunk census, costs,
// each class's forward then backward sweep on its own sub stream (classes
// side by side), and the LR checks.
hipError_t capture_sweep(vxh_sgm_batch *b) {
    using namespace vxh;
    hipError_t e = launch_zero(b->d_error, 1, b->stream);
    if (e == hipSuccess && b->ncounters) e = launch_zero(b->counters, b->ncounters, b->stream);
    b->launches = 2;
    for (const Chunk &c : b->chunks) {
        if (e == hipSuccess) e = launch_census_multi(c.census, c.ncensus, c.census_blocks, b->stream);
        if (e == hipSuccess) e = launch_cost_multi(c.cost, c.ncost, c.cost_blocks, b->stream);
        b->launches += 2;
        if (e == hipSuccess) e = hipEventRecord(b->events[0], b->stream);
        for (size_t x = 0; x < c.sweeps.size() && e == hipSuccess; ++x) {
            const Chunk::Sweep &sw = c.sweeps[x];
            hipStream_t s = b->sub[x];
            e = hipStreamWaitEvent(s, b->events[0], 0);
            if (e == hipSuccess) e = launch_sweep(sw.cls, false, sw.fwd, sw.blocks_fwd, s);
            if (e == hipSuccess) e = launch_sweep(sw.cls, true, sw.bwd, sw.blocks_bwd, s);
            if (e == hipSuccess) e = hipEventRecord(b->events[1 + x], s);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->events[1 + x], 0);
            b->launches += 2;
        }
        for (uint32_t i : c.pairs) {
            const vxh_sgm_pair_desc &q = b->pairs[i];
            if (!q.cfg.lr_check || e != hipSuccess) continue;
            const SweepLayout L = sweep_layout(q.width, q.height, q.cfg.num_disparities, true);
            char *base = (char *)b->ws[0] + b->plan[i].ws_off;
            const int D = (int)q.cfg.num_disparities;
            e = launch_lr_check((const uint16_t *)(base + L.sum), (int)q.width, (int)q.height, D,
                                volume_stride(D), (int)q.cfg.lr_tolerance,
                                (const int16_t *)(base + L.best), q.d_out,
                                (int16_t *)(base + L.lc), b->stream);
            b->launches += 1;
        }
    }
    return e;
}

// Records the streams form's launches on b->stream (being captured).
vxh_status capture_streams(vxh_sgm_batch *b, const std::vector<uint32_t> &order,
                           hipError_t *err) {
    using namespace vxh;
    const uint32_t streams = (uint32_t)b->sub.size();
    hipError_t e = launch_zero(b->d_error, 1, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->events[0], b->stream);
    for (uint32_t s = 0; s < streams && e == hipSuccess; ++s)
        e = hipStreamWaitEvent(b->sub[s], b->events[0], 0);
    vxh_status st = VXH_OK;
    for (size_t k = 0; k < order.size() && e == hipSuccess && st == VXH_OK; ++k) {
        const uint32_t i = order[k];
        const vxh_sgm_pair_desc &q = b->pairs[i];
        const PairPlan &p = b->plan[i];
        char *ws = (char *)b->ws[p.stream];
        st = run_pipeline(q.d_left, q.d_right, q.width, q.height, &q.cfg, q.d_out, ws, p.mode,
                          p.sched, b->sub[p.stream]);
        if (st != VXH_OK) break;
        e = launch_or_word(b->d_error, (const uint32_t *)(ws + kErrorWordOffset),
                           b->sub[p.stream]);
        // census, costs, paths, winners, error word; LR: key clear and check.
        b->launches += 5 + (q.cfg.lr_check ? 2 : 0);
    }
    for (uint32_t s = 0; s < streams && e == hipSuccess && st == VXH_OK; ++s) {
        e = hipEventRecord(b->events[1 + s], b->sub[s]);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->events[1 + s], 0);
    }
    *err = e;
    return st;
}

}  // namespace

extern "C" {

const char *vxh_sgm_batch_last_error(void) { return g_batch_error.c_str(); }

vxh_status vxh_sgm_batch_create(int device, const vxh_sgm_pair_desc *pairs, uint32_t n,
                                vxh_sgm_batch **out) {
    return vxh_sgm_batch_create_ex(device, pairs, n, nullptr, out);
}

vxh_status vxh_sgm_batch_create_ex(int device, const vxh_sgm_pair_desc *pairs, uint32_t n,
                                   const vxh_sgm_batch_options *opt, vxh_sgm_batch **out) {
    using namespace vxh;
    if (!out || (!pairs && n)) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    uint32_t streams = kDefaultStreams, form = kFormPersistent, blocks_per_cu = 0;
    uint64_t budget = 0;
    if (opt) {
        if (opt->streams > kMaxStreams || opt->form > kFormSweep)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch options");
        if (opt->streams) streams = opt->streams;
        form = opt->form;
        blocks_per_cu = opt->blocks_per_cu;
        if (opt->budget_bytes) budget = opt->budget_bytes;
    }
    if (n == 0 || n > 65535) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch size");
    if (hipSetDevice(device) != hipSuccess)
        return bfail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    streams = std::min(streams, n);
    if (!budget) {
        size_t free_b = 0, total_b = 0;
        budget = hipMemGetInfo(&free_b, &total_b) == hipSuccess
                     ? std::max<uint64_t>(free_b / 10 * kBudgetTenths, kMinBudget)
                     : kFallbackBudget;
    }

    std::vector<PairPlan> plan(n);
    for (uint32_t i = 0; i < n; ++i) {
        const vxh_sgm_pair_desc &q = pairs[i];
        if (vxh_sgm_config_validate(&q.cfg, q.width, q.height) != VXH_OK)
            return bfail(VXH_ERR_INVALID_CONFIG,
                         std::string("pair ") + std::to_string(i) + ": " + vxh_last_error());
        if (!q.d_left || !q.d_right || !q.d_out)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null pair buffer");
        if (q.volume_format > 2 || q.split > 2)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: pair options");
        if (form == kFormSweep && q.cfg.num_paths != 8)
            return bfail(VXH_ERR_INVALID_CONFIG,
                         "invalid config: the sweep form needs num_paths 8 (pair " +
                             std::to_string(i) + ")");
        const int D = (int)q.cfg.num_disparities;
        const int cls = layout_class(D);
        Schedule sc = default_schedule(D);
        if (q.lanes || q.per_lane) sc.other = layout_index(cls, q.lanes, q.per_lane);
        if (q.lanes_h || q.per_lane_h) sc.horiz = layout_index(cls, q.lanes_h, q.per_lane_h);
        if (sc.other < 0 || sc.horiz < 0)
            return bfail(VXH_ERR_INVALID_CONFIG,
                         "invalid config: unsupported lane layout (pair " + std::to_string(i) +
                             ", D=" + std::to_string(D) + ")");
        sc.split = form == kFormStreams ? (int)q.split : 0;
        plan[i].sched = sc;
        plan[i].mode = q.volume_format == 1 ? kOutU8
                       : q.volume_format == 2 ? kOutU16
                                              : auto_volume_mode(q.cfg.p2);
        plan[i].ws_bytes = form == kFormSweep
                               ? sweep_layout(q.width, q.height, (uint32_t)D,
                                              q.cfg.lr_check != 0).total
                               : workspace_layout(q.width, q.height, (uint32_t)D,
                                                  (int)q.cfg.num_paths, plan[i].mode,
                                                  q.cfg.lr_check != 0)
                                     .total;
        plan[i].ws_off = 0;
    }
    // Largest pairs first.
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    auto work = [&](uint32_t i) {
        return (size_t)pairs[i].width * pairs[i].height * pairs[i].cfg.num_disparities;
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t c) { return work(a) > work(c); });

    vxh_sgm_batch *b = new vxh_sgm_batch();
    b->device = device;
    b->form = form;
    b->pairs.assign(pairs, pairs + n);
    b->plan = plan;
    auto gpu_fail = [&](const char *what, hipError_t e) {
        vxh_sgm_batch_destroy(b);
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            return bfail(VXH_ERR_BUFFER_ALLOC, std::string("buffer allocation failed: ") + what);
        return bfail(VXH_ERR_GPU, std::string("GPU error: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
    // Sub streams: the streams form's pair streams, the persistent form's
    // side-by-side aggregation groups (at most the three layout classes times
    // two volume formats), the sweep form's classes.
    const uint32_t nsub = form == kFormStreams ? streams : 6;
    if (form == kFormStreams) {
        // Pairs, largest first, each to the stream with the least work so far
        // (W*H*D*paths); one workspace per stream, sized for the largest pair
        // dealt to it.
        std::vector<size_t> ws_bytes(streams, 0);
        std::vector<uint64_t> work(streams, 0);
        for (uint32_t k = 0; k < n; ++k) {
            PairPlan &p = b->plan[order[k]];
            const vxh_sgm_pair_desc &q = pairs[order[k]];
            p.stream = (uint32_t)(std::min_element(work.begin(), work.end()) - work.begin());
            work[p.stream] += (uint64_t)q.width * q.height * q.cfg.num_disparities * q.cfg.num_paths;
            ws_bytes[p.stream] = std::max(ws_bytes[p.stream], p.ws_bytes);
        }
        for (uint32_t s = 0; s < streams && e == hipSuccess; ++s) {
            void *p = nullptr;
            e = hipMalloc(&p, ws_bytes[s]);
            if (e == hipSuccess) b->ws.push_back(p);
        }
        if (e != hipSuccess) return gpu_fail("batch workspaces", e);
    }
    // The sweep form's tables point at the error word.
    if ((e = hipMalloc((void **)&b->d_error, 256)) != hipSuccess)
        return gpu_fail("batch error word", e);
    if (form == kFormPersistent) {
        if ((e = plan_persistent(b, order, budget, blocks_per_cu)) != hipSuccess)
            return gpu_fail("batch plan", e);
    } else if (form == kFormSweep) {
        if ((e = plan_sweep(b, order, budget)) != hipSuccess) return gpu_fail("batch plan", e);
    }
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return gpu_fail("stream", e);
    for (uint32_t s = 0; s < nsub; ++s) {
        hipStream_t st = nullptr;
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
            return gpu_fail("stream", e);
        b->sub.push_back(st);
    }
    for (uint32_t s = 0; s <= nsub; ++s) {
        hipEvent_t ev = nullptr;
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
            return gpu_fail("event", e);
        b->events.push_back(ev);
    }

    e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return gpu_fail("graph capture", e);
    vxh_status st = VXH_OK;
    std::string what;
    if (form == kFormStreams) {
        st = capture_streams(b, order, &e);
        if (st != VXH_OK) what = vxh_last_error();
    } else if (form == kFormSweep) {
        e = capture_sweep(b);
    } else {
        e = capture_persistent(b);
    }
    const hipError_t e2 = hipStreamEndCapture(b->stream, &b->graph);
    if (st != VXH_OK) {
        vxh_sgm_batch_destroy(b);
        return bfail(st, "batch capture: " + what);
    }
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return gpu_fail("batch graph", e);
    *out = b;
    return VXH_OK;
}

vxh_status vxh_sgm_batch_launch(vxh_sgm_batch *b, void *stream) {
    if (!b) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null batch");
    if (hipSetDevice(b->device) != hipSuccess)
        return bfail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    hipError_t e = hipGraphLaunch(b->exec, stream ? (hipStream_t)stream : b->stream);
    if (e != hipSuccess)
        return bfail(VXH_ERR_GPU, std::string("GPU error: graph launch: ") + hipGetErrorString(e));
    return VXH_OK;
}

vxh_status vxh_sgm_batch_run(vxh_sgm_batch *b) {
    vxh_status st = vxh_sgm_batch_launch(b, nullptr);
    if (st != VXH_OK) return st;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess)
        return bfail(VXH_ERR_GPU, std::string("GPU error: batch sync: ") + hipGetErrorString(e));
    return vxh_sgm_batch_check(b);
}

vxh_status vxh_sgm_batch_check(vxh_sgm_batch *b) {
    if (!b) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null batch");
    uint32_t err = 0;
    hipError_t e = hipMemcpy(&err, b->d_error, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return bfail(VXH_ERR_GPU, std::string("GPU error: batch check: ") + hipGetErrorString(e));
    if (err) return bfail(VXH_ERR_GPU, "GPU error: a bounded wait (split scanline or sweep wave) timed out");
    return VXH_OK;
}

vxh_status vxh_sgm_batch_stats_get(const vxh_sgm_batch *b, uint32_t pair,
                                   vxh_sgm_batch_stats *out) {
    if (!b || !out || pair >= (uint32_t)b->pairs.size())
        return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch stats");
    const vxh_sgm_pair_desc &q = b->pairs[pair];
    const PairPlan &p = b->plan[pair];
    const int c = vxh::layout_class((int)q.cfg.num_disparities);
    out->stream = b->form == kFormStreams ? p.stream : 0;
    out->lanes = (uint32_t)vxh::kClassLayouts[c][p.sched.other][0];
    out->per_lane = (uint32_t)vxh::kClassLayouts[c][p.sched.other][1];
    out->lanes_h = (uint32_t)vxh::kClassLayouts[c][p.sched.horiz][0];
    out->per_lane_h = (uint32_t)vxh::kClassLayouts[c][p.sched.horiz][1];
    out->volume_format = p.mode == vxh::kOutU8 ? 1u : 2u;
    out->workspace_bytes = p.ws_bytes;
    out->chunk = b->form == kFormStreams ? 0 : p.stream;
    out->form = b->form;
    return VXH_OK;
}

uint32_t vxh_sgm_batch_items(const vxh_sgm_batch *b) { return b ? b->launches : 0; }

uint32_t vxh_sgm_batch_chunks(const vxh_sgm_batch *b) {
    return b ? (b->form == kFormStreams ? 1u : (uint32_t)b->chunks.size()) : 0u;
}

void vxh_sgm_batch_destroy(vxh_sgm_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipStream_t s : b->sub) (void)hipStreamSynchronize(s);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t ev : b->events) (void)hipEventDestroy(ev);
    for (hipStream_t s : b->sub) (void)hipStreamDestroy(s);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (void *p : b->ws) (void)hipFree(p);
    if (b->tables) (void)hipFree(b->tables);
    if (b->d_error) (void)hipFree(b->d_error);
    delete b;
}

}  // extern "C"
