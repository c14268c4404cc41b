/*
 * hpm_batch.cpp — RunPatchMatch of many independent problems as one batch
 * (no reference counterpart; the reference runs problems one at a time,
 * HPM.cu:1906-1957).
 *
 * All problems advance through the same schedule (init, 3 x red/black
 * propagation, finalize, red/black filter); each stage is one launch of a
 * grid of resident blocks that walk a table of tiles covering every problem
 * (hpm_launch_batch), in contiguous ranges or pulling work items from
 * per-XCD queues. The 11 launches are captured once per
 * batch shape in a hipGraph and replayed; the problem table (pointers,
 * sizes, RNG call index) is uploaded before every replay; up to 8 shapes
 * keep their graphs (a pass schedule alternates between the shapes of its
 * passes). The schedule (static ranges over a grid of 2-16 blocks per CU,
 * or work queues with items of 1, 2 or 4 tiles) is chosen by timing the
 * candidates on the first call of a shape, with the
 * problems' state saved before and restored after every trial, so tuning
 * leaves the results untouched. Per pixel the work is that of
 * hpm_run_patchmatch, so every problem's maps equal an unbatched run's.
 */
#include <hip/hip_runtime.h>

#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "hpm_batch.h"
#include "hpm_internal.h"

/* one batch shape (problem count and sizes): its device tables and the
 * captured schedule, whose launches point into these tables */
struct hpm_batch_shape {
    std::vector<int> key;
    hpm_dev_problem* d_pbs = nullptr;      /* [2n]: init counters, then propagation counters */
    unsigned long long* d_neval = nullptr; /* [2n] */
    hpm_tile* d_tiles = nullptr;           /* full-grid tiles, then colour-grid tiles */
    unsigned long long* d_queue = nullptr; /* 8 u32 work-queue counters per launch, zeroed per run */
    hipGraphExec_t exec = nullptr;
    int grid = 0;
    int gran = 0;                          /* 0: static ranges; else tiles per work item */
    bool persistent = false;               /* one cooperative launch for the whole schedule */
    int n_full = 0, n_colour = 0;
    uint64_t last_use = 0;
};

struct hpm_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::string err;
    std::string log;
    std::vector<hpm_batch_shape> shapes;
    uint64_t clock = 0;
    float* d_save = nullptr;          /* state saved around tuning trials */
    size_t cap_save = 0;
    hpm_dev_problem* h_pbs = nullptr; /* pinned */
    size_t cap_h_pbs = 0;
    int fixed_grid = 0, fixed_items = -1; /* hpm_batch_set_schedule; -1: autotune */
    bool fixed_persistent = false;
};

namespace {

int bfail(hpm_batch* b, int code, const std::string& msg)
{
    b->err = msg;
    return code;
}

#define BHIP(b, expr)                                                                                              \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess) {                                                                                    \
            return bfail((b), HPM_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));                    \
        }                                                                                                          \
    } while (0)

template <typename T>
int grow(hpm_batch* b, T** p, size_t* cap, size_t count)
{
    if (*cap >= count) {
        return HPM_OK;
    }
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    BHIP(b, hipMalloc((void**)p, count * sizeof(T)));
    *cap = count;
    return HPM_OK;
}

/* tiles of the full grid (x in [0, W), y in [0, H); 64x4) and of the colour
 * grid (x = 2i + parity, rows below min(H, 2 yy_limit); HPM_COLOUR_TILE_X x
 * HPM_COLOUR_TILE_Y) of every problem */
void make_tiles(const hpm_dev_problem* pbs, int n, std::vector<hpm_tile>* full, std::vector<hpm_tile>* colour)
{
    full->clear();
    colour->clear();
    for (int i = 0; i < n; ++i) {
        const hpm_dev_problem& pb = pbs[i];
        for (int by = 0; by < (pb.H + 3) / 4; ++by) {
            for (int bx = 0; bx < (pb.W + 63) / 64; ++bx) {
                full->push_back({i, bx, by, 0});
            }
        }
        const int half_w = (pb.W + 1) / 2;
        const int rows = std::min(pb.H, 2 * pb.yy_limit);
        for (int by = 0; by < (rows + HPM_COLOUR_TILE_Y - 1) / HPM_COLOUR_TILE_Y; ++by) {
            for (int bx = 0; bx < (half_w + HPM_COLOUR_TILE_X - 1) / HPM_COLOUR_TILE_X; ++bx) {
                colour->push_back({i, bx, by, 0});
            }
        }
    }
}

const size_t kShapes = 8;

void free_shape(hpm_batch_shape* sh)
{
    if (sh->exec) (void)hipGraphExecDestroy(sh->exec);
    if (sh->d_pbs) (void)hipFree(sh->d_pbs);
    if (sh->d_neval) (void)hipFree(sh->d_neval);
    if (sh->d_tiles) (void)hipFree(sh->d_tiles);
    if (sh->d_queue) (void)hipFree(sh->d_queue);
    *sh = hpm_batch_shape();
}

const int kLaunches = 11;

/* one schedule: 11 stage launches, or the zeroing kernels and one
 * persistent cooperative launch */
int enqueue(hpm_batch* b, const hpm_batch_shape& sh, int n, int grid, int gran, bool persistent)
{
    if (persistent) {
        BHIP(b, (hipError_t)hpm_launch_zero_u64(sh.d_neval, 2 * n, b->stream));
        BHIP(b, (hipError_t)hpm_launch_zero_u64(sh.d_queue, 4 * kLaunches, b->stream));
        BHIP(b, (hipError_t)hpm_launch_batch_persistent(sh.d_pbs, n, sh.d_tiles, sh.n_full, sh.n_colour, gran,
                                                        (unsigned int*)sh.d_queue, grid, b->stream));
        return HPM_OK;
    }
    const hpm_tile* tf = sh.d_tiles;
    const hpm_tile* tc = sh.d_tiles + sh.n_full;
    const hpm_dev_problem* p_init = sh.d_pbs;
    const hpm_dev_problem* p_prop = sh.d_pbs + n;
    unsigned int* q = (unsigned int*)sh.d_queue;
    int li = 0;
    auto stage = [&](int st, const hpm_dev_problem* pbs, const hpm_tile* tiles, int nt, int it, int colour) {
        return (hipError_t)hpm_launch_batch(st, pbs, tiles, nt, grid, it, colour, gran, q + 8 * li++, b->stream);
    };
    BHIP(b, (hipError_t)hpm_launch_zero_u64(sh.d_neval, 2 * n, b->stream));
    BHIP(b, (hipError_t)hpm_launch_zero_u64(sh.d_queue, 4 * kLaunches, b->stream));
    BHIP(b, stage(HPM_STAGE_INIT, p_init, tf, sh.n_full, 0, 0));
    for (int it = 0; it < 3; ++it) { /* hard-coded 3 (HPM.cu:1933) */
        for (int colour = 0; colour < 2; ++colour) {
            BHIP(b, stage(HPM_STAGE_PROPAGATE, p_prop, tc, sh.n_colour, it, colour));
        }
    }
    BHIP(b, stage(HPM_STAGE_FINALIZE, p_prop, tf, sh.n_full, 0, 0));
    BHIP(b, stage(HPM_STAGE_FILTER, p_prop, tc, sh.n_colour, 0, 0));
    BHIP(b, stage(HPM_STAGE_FILTER, p_prop, tc, sh.n_colour, 0, 1));
    return HPM_OK;
}

/* hypotheses, costs and view masks of every problem <-> d_save */
int save_state(hpm_batch* b, hpm_handle* const* hs, int n, bool restore)
{
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        const size_t np = (size_t)hs[i]->cams[0].width * hs[i]->cams[0].height;
        float* s = b->d_save + off;
        if (restore) {
            BHIP(b, hipMemcpyAsync(hs[i]->d_hyp, s, np * 16, hipMemcpyDeviceToDevice, b->stream));
            BHIP(b, hipMemcpyAsync(hs[i]->d_cost, s + 4 * np, np * 4, hipMemcpyDeviceToDevice, b->stream));
            BHIP(b, hipMemcpyAsync(hs[i]->d_sel, s + 5 * np, np * 4, hipMemcpyDeviceToDevice, b->stream));
            /* the restored hypotheses are not the ones the cost cache holds */
            if (hs[i]->d_vcost) {
                BHIP(b, (hipError_t)hpm_launch_fill_u32((uint32_t*)hs[i]->d_vcost, (int64_t)np, 0xffffffffu,
                                                        b->stream));
            }
        } else {
            BHIP(b, hipMemcpyAsync(s, hs[i]->d_hyp, np * 16, hipMemcpyDeviceToDevice, b->stream));
            BHIP(b, hipMemcpyAsync(s + 4 * np, hs[i]->d_cost, np * 4, hipMemcpyDeviceToDevice, b->stream));
            BHIP(b, hipMemcpyAsync(s + 5 * np, hs[i]->d_sel, np * 4, hipMemcpyDeviceToDevice, b->stream));
        }
        off += 6 * np;
    }
    return HPM_OK;
}

} // namespace

void hpm_make_tiles(const hpm_dev_problem* pbs, int n, std::vector<hpm_tile>* full, std::vector<hpm_tile>* colour)
{
    make_tiles(pbs, n, full, colour);
}

extern "C" hpm_batch* hpm_batch_create(int device)
{
    hpm_batch* b = new hpm_batch();
    b->device = device < 0 ? 0 : device;
    if (hipSetDevice(b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess) {
        hpm_batch_destroy(b);
        return nullptr;
    }
    return b;
}

extern "C" void hpm_batch_destroy(hpm_batch* b)
{
    if (!b) {
        return;
    }
    (void)hipSetDevice(b->device);
    if (b->stream) {
        (void)hipStreamSynchronize(b->stream);
    }
    for (hpm_batch_shape& sh : b->shapes) {
        free_shape(&sh);
    }
    if (b->d_save) (void)hipFree(b->d_save);
    if (b->h_pbs) (void)hipHostFree(b->h_pbs);
    for (hipEvent_t e : b->ev) {
        if (e) (void)hipEventDestroy(e);
    }
    if (b->stream) {
        (void)hipStreamDestroy(b->stream);
    }
    delete b;
}

extern "C" int hpm_batch_set_schedule(hpm_batch* b, int grid, int items, int persistent)
{
    if (!b || !(items == -1 || items == 0 || items == 1 || items == 2 || items == 4) || grid < 0 ||
        (persistent && items != 1 && items != 2)) {
        return HPM_EINVAL;
    }
    b->fixed_grid = grid;
    b->fixed_items = items;
    b->fixed_persistent = persistent != 0;
    for (hpm_batch_shape& sh : b->shapes) {
        free_shape(&sh);
    }
    b->shapes.clear();
    return HPM_OK;
}

extern "C" const char* hpm_batch_last_error(const hpm_batch* b)
{
    return b ? b->err.c_str() : "null batch";
}

extern "C" const char* hpm_batch_log(const hpm_batch* b)
{
    return b ? b->log.c_str() : "";
}

extern "C" int hpm_run_batch(hpm_batch* b, hpm_handle* const* hs, int n, int want_readback, hpm_batch_stats* stats)
{
    if (!b || !hs || n < 1) {
        return HPM_EINVAL;
    }
    auto t_host = std::chrono::steady_clock::now();
    BHIP(b, hipSetDevice(b->device));
    std::vector<hpm_dev_problem> pbs((size_t)n);
    std::vector<int> key;
    key.push_back(n);
    size_t save_floats = 0;
    for (int i = 0; i < n; ++i) {
        hpm_handle* h = hs[i];
        if (!h) {
            return bfail(b, HPM_EINVAL, "null handle in the batch");
        }
        if (h->device != b->device) {
            return bfail(b, HPM_EINVAL, "handles of a batch must be on the batch's device");
        }
        const int rc = hpm_ensure_ready(h);
        if (rc != HPM_OK) {
            return bfail(b, rc, "problem " + std::to_string(i) + ": " + h->err);
        }
        if (h->params.acmm) {
            return bfail(b, HPM_EINVAL, "problem " + std::to_string(i) + ": the batched form runs HPM-MVS problems");
        }
        hpm_fill_problem(h, &pbs[(size_t)i]);
        key.push_back(pbs[(size_t)i].W);
        key.push_back(pbs[(size_t)i].H);
        save_floats += (size_t)pbs[(size_t)i].W * pbs[(size_t)i].H * 6;
    }
    /* the shape's entry: a new one gets its tables, the oldest is evicted */
    hpm_batch_shape* sh = nullptr;
    for (hpm_batch_shape& x : b->shapes) {
        if (x.key == key) {
            sh = &x;
        }
    }
    const bool reused = sh != nullptr;
    if (!sh) {
        if (b->shapes.size() >= kShapes) {
            auto old = std::min_element(b->shapes.begin(), b->shapes.end(),
                                        [](const hpm_batch_shape& a, const hpm_batch_shape& c) {
                                            return a.last_use < c.last_use;
                                        });
            free_shape(&*old);
            b->shapes.erase(old);
        }
        b->shapes.emplace_back();
        sh = &b->shapes.back();
        sh->key = key;
        BHIP(b, hipMalloc((void**)&sh->d_pbs, 2 * (size_t)n * sizeof(hpm_dev_problem)));
        BHIP(b, hipMalloc((void**)&sh->d_neval, 2 * (size_t)n * sizeof(unsigned long long)));
        BHIP(b, hipMalloc((void**)&sh->d_queue, 4 * (size_t)kLaunches * sizeof(unsigned long long)));
    }
    sh->last_use = ++b->clock;
    int rc = HPM_OK;
    /* problem tables: init counts into 2i, propagation into 2i + 1 */
    if (b->cap_h_pbs < 2 * (size_t)n) {
        if (b->h_pbs) (void)hipHostFree(b->h_pbs);
        b->h_pbs = nullptr;
        BHIP(b, hipHostMalloc((void**)&b->h_pbs, 2 * (size_t)n * sizeof(hpm_dev_problem), hipHostMallocDefault));
        b->cap_h_pbs = 2 * (size_t)n;
    }
    for (int i = 0; i < n; ++i) {
        pbs[(size_t)i].n_eval = sh->d_neval + 2 * i;
        b->h_pbs[i] = pbs[(size_t)i];
        b->h_pbs[n + i] = pbs[(size_t)i];
        b->h_pbs[n + i].n_eval = sh->d_neval + 2 * i + 1;
    }
    BHIP(b, hipMemcpyAsync(sh->d_pbs, b->h_pbs, 2 * (size_t)n * sizeof(hpm_dev_problem), hipMemcpyHostToDevice,
                           b->stream));
    /* every handle's earlier work (uploads, reloads) precedes the batch */
    for (int i = 0; i < n; ++i) {
        BHIP(b, hipStreamSynchronize(hs[i]->stream));
    }

    if (!reused) {
        std::vector<hpm_tile> full, colour;
        make_tiles(pbs.data(), n, &full, &colour);
        sh->n_full = (int)full.size();
        sh->n_colour = (int)colour.size();
        BHIP(b, hipMalloc((void**)&sh->d_tiles, (full.size() + colour.size()) * sizeof(hpm_tile)));
        BHIP(b, hipMemcpyAsync(sh->d_tiles, full.data(), full.size() * sizeof(hpm_tile), hipMemcpyHostToDevice,
                               b->stream));
        BHIP(b, hipMemcpyAsync(sh->d_tiles + full.size(), colour.data(), colour.size() * sizeof(hpm_tile),
                               hipMemcpyHostToDevice, b->stream));
        BHIP(b, hipStreamSynchronize(b->stream));
        /* autotuner over the schedule: static ranges for grids of 2-16
         * blocks per CU (2 blocks of 256 lanes fit a CU, LDS-bound; more
         * blocks balance uneven tiles through the hardware scheduler), or
         * one resident grid pulling work items of 1, 2 or 4 tiles from the
         * per-XCD queues */
        hipDeviceProp_t prop;
        BHIP(b, hipGetDeviceProperties(&prop, b->device));
        const int cus = std::max(1, prop.multiProcessorCount);
        struct Cand {
            int grid, gran;
            bool persistent;
        };
        int pgrid[3] = {0, 0, 0};
        for (int g = 1; g <= 2; ++g) {
            if (hpm_batch_persistent_grid(b->device, g, &pgrid[g]) != 0) {
                pgrid[g] = 0;
            }
        }
        std::vector<Cand> cands = {{2 * cus, 0, false}, {4 * cus, 0, false}, {8 * cus, 0, false},
                                   {16 * cus, 0, false}, {2 * cus, 1, false}, {2 * cus, 2, false},
                                   {2 * cus, 4, false}};
        for (int g = 1; g <= 2; ++g) {
            if (pgrid[g] > 0) {
                cands.push_back({pgrid[g], g, true});
            }
        }
        rc = grow(b, &b->d_save, &b->cap_save, save_floats);
        if (rc != HPM_OK) return rc;
        rc = save_state(b, hs, n, false);
        if (rc != HPM_OK) return rc;
        char line[256];
        snprintf(line, sizeof(line), "batch of %d problems, %d full / %d colour tiles, %d CUs:", n, sh->n_full,
                 sh->n_colour, cus);
        std::string log = line;
        float best_ms = -1.0f;
        Cand best = cands[0];
        const bool fixed = b->fixed_items >= 0;
        if (fixed) {
            if (b->fixed_persistent && pgrid[b->fixed_items] <= 0) {
                return bfail(b, HPM_EDEVICE, "the persistent schedule cannot be launched on this device");
            }
            best = b->fixed_persistent ? Cand{pgrid[b->fixed_items], b->fixed_items, true}
                                       : Cand{b->fixed_grid > 0 ? b->fixed_grid : 2 * cus, b->fixed_items, false};
            log += " fixed schedule;";
        }
        for (const Cand& c : cands) {
            if (fixed) {
                break;
            }
            rc = save_state(b, hs, n, true);
            if (rc != HPM_OK) return rc;
            BHIP(b, hipEventRecord(b->ev[0], b->stream));
            rc = enqueue(b, *sh, n, c.grid, c.gran, c.persistent);
            if (rc != HPM_OK) return rc;
            BHIP(b, hipEventRecord(b->ev[1], b->stream));
            BHIP(b, hipEventSynchronize(b->ev[1]));
            float ms = 0.0f;
            BHIP(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
            if (c.persistent) {
                snprintf(line, sizeof(line), " persistent launch, items of %d tiles, grid %d: %.3f ms;", c.gran,
                         c.grid, ms);
            } else if (c.gran) {
                snprintf(line, sizeof(line), " queue items of %d tiles, grid %d: %.3f ms;", c.gran, c.grid, ms);
            } else {
                snprintf(line, sizeof(line), " static grid %d: %.3f ms;", c.grid, ms);
            }
            log += line;
            if (best_ms < 0.0f || ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        rc = save_state(b, hs, n, true);
        if (rc != HPM_OK) return rc;
        snprintf(line, sizeof(line), " chosen %s grid %d items %d\n",
                 best.persistent ? "persistent" : (best.gran ? "queue" : "static"), best.grid, best.gran);
        log += line;
        b->log += log;
        sh->grid = best.grid;
        sh->gran = best.gran;
        sh->persistent = best.persistent;
    }
    if (!reused && !sh->persistent) {
        /* capture the schedule (the persistent one is a single launch) */
        hipGraph_t graph = nullptr;
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        rc = enqueue(b, *sh, n, sh->grid, sh->gran, false);
        hipError_t ce = hipStreamEndCapture(b->stream, &graph);
        if (rc != HPM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        BHIP(b, ce);
        hipError_t ie = hipGraphInstantiate(&sh->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        BHIP(b, ie);
    }
    BHIP(b, hipEventRecord(b->ev[0], b->stream));
    if (sh->persistent) {
        rc = enqueue(b, *sh, n, sh->grid, sh->gran, true);
        if (rc != HPM_OK) return rc;
    } else {
        BHIP(b, hipGraphLaunch(sh->exec, b->stream));
    }
    BHIP(b, hipEventRecord(b->ev[1], b->stream));
    std::vector<unsigned long long> nev(2 * (size_t)n, 0ull);
    BHIP(b, hipMemcpyAsync(nev.data(), sh->d_neval, nev.size() * 8, hipMemcpyDeviceToHost, b->stream));
    if (want_readback) {
        for (int i = 0; i < n; ++i) {
            hpm_handle* h = hs[i];
            const size_t np = (size_t)h->cams[0].width * h->cams[0].height;
            h->h_hyp.resize(np * 4);
            h->h_cost.resize(np);
            BHIP(b, hipMemcpyAsync(h->h_hyp.data(), h->d_hyp, np * 16, hipMemcpyDeviceToHost, b->stream));
            BHIP(b, hipMemcpyAsync(h->h_cost.data(), h->d_cost, np * 4, hipMemcpyDeviceToHost, b->stream));
        }
    }
    BHIP(b, hipStreamSynchronize(b->stream));
    if (want_readback) {
        for (int i = 0; i < n; ++i) {
            hs[i]->mirror_filled = true;
        }
    }
    float ms = 0.0f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    uint64_t total = 0;
    for (int i = 0; i < n; ++i) {
        hpm_handle* h = hs[i];
        h->stats.n_eval = nev[2 * (size_t)i] + nev[2 * (size_t)i + 1];
        h->stats.n_eval_propagate = nev[2 * (size_t)i + 1];
        h->stats.launches = 11;
        h->stats.ms_total = ms; /* the whole batch */
        h->stats.ms_init = h->stats.ms_propagate = h->stats.ms_finalize_filter = 0.0f;
        h->call_index += 1;
        total += h->stats.n_eval;
    }
    if (stats) {
        stats->ms_device = ms;
        stats->ms_host = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_host).count();
        stats->n_eval = total;
        stats->problems = n;
        stats->tiles_full = sh->n_full;
        stats->tiles_colour = sh->n_colour;
        stats->grid = sh->grid;
        stats->items = sh->gran;
        stats->persistent = sh->persistent ? 1 : 0;
        stats->graph_reused = reused ? 1 : 0;
    }
    return HPM_OK;
}
