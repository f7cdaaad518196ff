// spk_ensemble.hip -- many independent boxes advanced together.
//
// Each box is a solver of its own (grid, physics, coefficients, precision of
// the bulk, its own dt), so an ensemble of uneven boxes (64^3 and 128^3 MHD,
// 512x512x2 Boussinesq-MHD with RKL2 conduction) is a list of solvers.
//
// Boxes of one shape -- the same grid, physics, precision and back end, with
// diffusion coefficients (nu, nu_m, nu_th, chi) free per box -- form a group,
// and a group is advanced as one batch: every kernel of its step is launched
// once for all of its boxes, with the box in blockIdx.z, each box's arrays
// looked up in a device table (BoxRow) and its coefficients overriding the
// launch's.  A 64^3 transform fills a fraction of the 256 CUs; 48 of them in
// one launch fill the chip.  A group may be cut into several batches
// ("split"), each on its own HIP stream ("lane"), so that batches of different
// groups and halves of one group overlap.  Split 0 launches box by box over
// the lanes, as one solver each.
//
// Two steps of every box are captured once into one HIP graph (fork from an
// origin stream into the lanes, join back) and each replay is a single graph
// launch; a second graph of the same steps continues from the inverse x pass
// the previous replay left (the fresh one projects and transforms first).  A captured step needs no host synchronization: dt stays on the
// device and RKL2 cycles are planned on the device (spk_solver_set_rkl_device),
// every box with its own stage count (stages past a box's count return at once).
#include "spk_internal.hpp"

#include <chrono>
#include <vector>

namespace {

struct BoxBatch {
    std::vector<int> boxes;         // indices into spk_ensemble_s::box; boxes[0] leads
    int lane = 0;
    spk::BoxRow *rows = nullptr;    // device table, one row per box
    const char *base[spk::BOX_SLOTS] = {};  // the leader's arrays
    size_t size[spk::BOX_SLOTS] = {};
};

}  // namespace

struct spk_ensemble_s {
    int device = 0;
    int bulk = SPK_F64;
    int nlanes = 0;
    int use_graph = 1;
    int split = 1;                    // 0: box by box; k >= 1: each group in k batches
    std::vector<spk_solver> box;
    std::vector<spk_ctx> bctx;        // one context per box; its stream is the box's lane
    std::vector<std::vector<int>> groups;
    std::vector<BoxBatch> batches;
    std::vector<hipStream_t> lanes;   // MAX_LANES streams, the first nlanes in use
    std::vector<hipEvent_t> ev_join;  // one per lane
    hipStream_t origin = nullptr;     // capture origin and graph launch stream
    hipEvent_t ev_fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // the same two steps continuing from the inverse x pass the previous replay
    // left in W (spk_main_loop_resume), where every box's steps leave one
    hipGraph_t graph_warm = nullptr;
    hipGraphExec_t exec_warm = nullptr;
    std::vector<int> parity;          // host step parity of each box when the graph was captured
};

using namespace spk;

namespace {

constexpr int MAX_LANES = 16;

void drop_graph(spk_ensemble e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->exec_warm) (void)hipGraphExecDestroy(e->exec_warm);
    if (e->graph_warm) (void)hipGraphDestroy(e->graph_warm);
    e->exec = nullptr;
    e->graph = nullptr;
    e->exec_warm = nullptr;
    e->graph_warm = nullptr;
}

// two steps of one solver: a fresh start, or continuing from W
int two_steps(spk_solver s, bool resume) { return resume ? spk_main_loop_resume(s, 2) : spk_main_loop(s, 2); }

void drop_batches(spk_ensemble e) {
    for (BoxBatch &b : e->batches)
        if (b.rows) (void)hipFree(b.rows);
    e->batches.clear();
}

// origin waits for every lane (work issued on a box's own stream before a
// replay), or every lane waits for origin (calls on a box after a replay)
int join_lanes_into_origin(spk_ensemble e) {
    for (int l = 0; l < e->nlanes; l++) {
        SPK_HIP(hipEventRecord(e->ev_join[l], e->lanes[l]));
        SPK_HIP(hipStreamWaitEvent(e->origin, e->ev_join[l], 0));
    }
    return SPK_OK;
}
int fork_origin_into_lanes(spk_ensemble e) {
    SPK_HIP(hipEventRecord(e->ev_fork, e->origin));
    for (int l = 0; l < e->nlanes; l++) SPK_HIP(hipStreamWaitEvent(e->lanes[l], e->ev_fork, 0));
    return SPK_OK;
}

// Two steps of the boxes of one batch: one launch per kernel for all of them,
// through the leader's context.  Boxes whose host step differs from the
// leader's (one was stepped on its own) run box by box instead.
int batch_two_steps(spk_ensemble e, const BoxBatch &b, bool resume) {
    spk_solver lead = e->box[b.boxes[0]];
    bool aligned = true;
    for (int i : b.boxes) aligned = aligned && solver_host_step(e->box[i]) == solver_host_step(lead);
    if (!aligned || b.boxes.size() == 1) {
        for (int i : b.boxes) SPK_TRY(two_steps(e->box[i], resume));
        return SPK_OK;
    }
    spk_ctx c = solver_ctx(lead);
    c->rows = b.rows;
    c->nbox = (int)b.boxes.size();
    for (int k = 0; k < BOX_SLOTS; k++) {
        c->base[k] = b.base[k];
        c->size[k] = b.size[k];
    }
    const int rc = two_steps(lead, resume);
    c->rows = nullptr;
    c->nbox = 1;
    SPK_TRY(rc);
    for (size_t k = 1; k < b.boxes.size(); k++) solver_advance_host(e->box[b.boxes[k]], 2);
    return SPK_OK;
}

// Two steps of every box: batch by batch, or box by box over the lanes.
int issue_two_steps(spk_ensemble e, bool resume) {
    if (e->split == 0) {
        for (spk_solver s : e->box) SPK_TRY(two_steps(s, resume));
        return SPK_OK;
    }
    for (const BoxBatch &b : e->batches) SPK_TRY(batch_two_steps(e, b, resume));
    return SPK_OK;
}

// every box's fused steps leave the next step's inverse x pass in W
bool all_continue(spk_ensemble e) {
    for (spk_solver s : e->box)
        if (!solver_continues(s)) return false;
    return true;
}
bool all_w_ok(spk_ensemble e) {
    for (spk_solver s : e->box)
        if (!solver_w_ok(s)) return false;
    return true;
}

int capture_one(spk_ensemble e, bool resume, hipGraph_t *graph, hipGraphExec_t *exec) {
    SPK_HIP(hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal));
    int rc = fork_origin_into_lanes(e);
    if (rc == SPK_OK) rc = issue_two_steps(e, resume);
    if (rc == SPK_OK) rc = join_lanes_into_origin(e);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(e->origin, &g);
    // the captured host code advanced each box's host step count; no kernel ran
    for (spk_solver s : e->box) solver_advance_host(s, -2);
    if (rc != SPK_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_error("spk_ensemble: graph capture failed: %s", hipGetErrorString(ec));
        return SPK_ERR_HIP;
    }
    *graph = g;
    const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        *exec = nullptr;
        set_error("spk_ensemble: graph instantiation failed: %s", hipGetErrorString(ei));
        return SPK_ERR_HIP;
    }
    return SPK_OK;
}

// The fresh-start graph, and where every box continues from step to step the
// graph that continues from the previous replay's W.  Capturing runs the step
// logic on the host only, so each box's record of W is kept across it.
int capture(spk_ensemble e) {
    drop_graph(e);
    std::vector<char> wok(e->box.size());
    for (size_t i = 0; i < e->box.size(); i++) wok[i] = solver_w_ok(e->box[i]);
    int rc = capture_one(e, false, &e->graph, &e->exec);
    if (rc == SPK_OK && all_continue(e)) {
        for (spk_solver s : e->box) solver_set_w_ok(s, true);
        rc = capture_one(e, true, &e->graph_warm, &e->exec_warm);
    }
    for (size_t i = 0; i < e->box.size(); i++) solver_set_w_ok(e->box[i], wok[i] != 0);
    if (rc != SPK_OK) {
        drop_graph(e);
        return rc;
    }
    e->parity.resize(e->box.size());
    for (size_t i = 0; i < e->box.size(); i++) e->parity[i] = solver_host_step(e->box[i]) & 1;
    return SPK_OK;
}

int sync_all(spk_ensemble e) {
    for (hipStream_t st : e->lanes) SPK_HIP(hipStreamSynchronize(st));
    SPK_HIP(hipStreamSynchronize(e->origin));
    return SPK_OK;
}

// Box by box over nlanes streams (box i on lane i % nlanes).
int assign_lanes(spk_ensemble e, int nlanes) {
    SPK_REQUIRE(nlanes >= 1 && nlanes <= MAX_LANES, SPK_ERR_ARG, "spk_ensemble: %d lanes (1..%d)", nlanes,
                MAX_LANES);
    SPK_TRY(sync_all(e));
    drop_graph(e);
    drop_batches(e);
    e->split = 0;
    e->nlanes = nlanes;
    for (size_t i = 0; i < e->box.size(); i++) e->bctx[i]->stream = e->lanes[i % nlanes];
    return SPK_OK;
}

// Each group cut into `split` batches of consecutive boxes; batch b on lane
// b % MAX_LANES, and every box's own calls on its batch's lane.
int assign_batches(spk_ensemble e, int split) {
    SPK_REQUIRE(split >= 1, SPK_ERR_ARG, "spk_ensemble: split %d (at least 1)", split);
    SPK_TRY(sync_all(e));
    drop_graph(e);
    drop_batches(e);
    e->split = split;
    for (const std::vector<int> &g : e->groups) {
        const int n = (int)g.size();
        const int k = split < n ? split : n;
        for (int j = 0; j < k; j++) {
            BoxBatch b;
            for (int i = (int)((long long)n * j / k); i < (int)((long long)n * (j + 1) / k); i++) b.boxes.push_back(g[i]);
            b.lane = (int)e->batches.size() % MAX_LANES;
            e->batches.push_back(b);
        }
    }
    e->nlanes = (int)e->batches.size() < MAX_LANES ? (int)e->batches.size() : MAX_LANES;
    for (BoxBatch &b : e->batches) {
        std::vector<BoxRow> h(b.boxes.size());
        for (size_t k = 0; k < b.boxes.size(); k++) {
            const int i = b.boxes[k];
            size_t sz[BOX_SLOTS];
            solver_box_row(e->box[i], &h[k], sz);
            if (k == 0)
                for (int q = 0; q < BOX_SLOTS; q++) {
                    b.base[q] = h[0].p[q];
                    b.size[q] = sz[q];
                }
            e->bctx[i]->stream = e->lanes[b.lane];
        }
        SPK_HIP(hipMalloc((void **)&b.rows, sizeof(BoxRow) * h.size()));
        SPK_HIP(hipMemcpy(b.rows, h.data(), sizeof(BoxRow) * h.size(), hipMemcpyHostToDevice));
    }
    return SPK_OK;
}

void destroy_all(spk_ensemble e) {
    for (hipStream_t st : e->lanes) (void)hipStreamSynchronize(st);
    if (e->origin) (void)hipStreamSynchronize(e->origin);
    drop_graph(e);
    drop_batches(e);
    for (spk_solver s : e->box) spk_solver_destroy(s);
    for (spk_ctx c : e->bctx) spk_ctx_destroy(c);
    for (hipEvent_t ev : e->ev_join) (void)hipEventDestroy(ev);
    if (e->ev_fork) (void)hipEventDestroy(e->ev_fork);
    for (hipStream_t st : e->lanes) (void)hipStreamDestroy(st);
    if (e->origin) (void)hipStreamDestroy(e->origin);
}

int create_impl(spk_ensemble e, int device, int nboxes, const spk_params *params, int bulk, int nlanes) {
    SPK_HIP(hipSetDevice(device));
    e->device = device;
    e->bulk = bulk;
    SPK_HIP(hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking));
    SPK_HIP(hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming));
    for (int l = 0; l < MAX_LANES; l++) {
        hipStream_t st;
        SPK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        e->lanes.push_back(st);
        hipEvent_t ev;
        SPK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        e->ev_join.push_back(ev);
    }
    for (int i = 0; i < nboxes; i++) {
        const spk_params &p = params[i];
        SPK_REQUIRE(!p.shearing, SPK_ERR_ARG,
                    "spk_ensemble: box %d is a shearing box (its remap clock is kept on the host)", i);
        spk_ctx c = nullptr;
        SPK_TRY(spk_ctx_create(device, e->lanes[i % nlanes], &c));
        e->bctx.push_back(c);
        spk_solver s = nullptr;
        SPK_TRY(spk_solver_create(c, &p, bulk, &s));
        e->box.push_back(s);
        if (p.supertimestepping && p.sts_mask) SPK_TRY(spk_solver_set_rkl_device(s, 1));
        // the first box of each shape leads its group; the others take its back end
        int g = 0;
        while (g < (int)e->groups.size() && !solver_same_shape(e->box[e->groups[g][0]], s)) g++;
        if (g == (int)e->groups.size()) {
            e->groups.push_back({i});
        } else {
            SPK_TRY(solver_follow_back_end(s, e->box[e->groups[g][0]]));
            e->groups[g].push_back(i);
        }
    }
    e->nlanes = nlanes;
    return assign_batches(e, 1);
}

}  // namespace

extern "C" int spk_ensemble_create(int device, int nboxes, const spk_params *params, int bulk, int lanes,
                                   spk_ensemble *out) {
    SPK_REQUIRE(out && params && nboxes > 0, SPK_ERR_ARG, "spk_ensemble_create: bad arguments");
    SPK_REQUIRE(bulk == SPK_F64 || bulk == SPK_F32, SPK_ERR_ARG, "spk_ensemble_create: bulk precision %d", bulk);
    const int nl = lanes > 0 ? lanes : (nboxes < 4 ? nboxes : 4);
    SPK_REQUIRE(nl <= MAX_LANES, SPK_ERR_ARG, "spk_ensemble_create: %d lanes (at most %d)", nl, MAX_LANES);
    spk_ensemble e = new spk_ensemble_s();
    const int rc = create_impl(e, device, nboxes, params, bulk, nl);
    if (rc != SPK_OK) {
        destroy_all(e);
        delete e;
        return rc;
    }
    *out = e;
    return SPK_OK;
}

extern "C" int spk_ensemble_destroy(spk_ensemble e) {
    if (!e) return SPK_OK;
    destroy_all(e);
    delete e;
    return SPK_OK;
}

extern "C" int spk_ensemble_size(spk_ensemble e, int *nboxes, int *lanes) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_size: null ensemble");
    if (nboxes) *nboxes = (int)e->box.size();
    if (lanes) *lanes = e->nlanes;
    return SPK_OK;
}

extern "C" int spk_ensemble_layout(spk_ensemble e, int *ngroups, int *split, int *nbatches) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_layout: null ensemble");
    if (ngroups) *ngroups = (int)e->groups.size();
    if (split) *split = e->split;
    if (nbatches) *nbatches = (int)e->batches.size();
    return SPK_OK;
}

extern "C" int spk_ensemble_box(spk_ensemble e, int i, spk_solver *out) {
    SPK_REQUIRE(e && out && i >= 0 && i < (int)e->box.size(), SPK_ERR_ARG, "spk_ensemble_box: bad arguments");
    *out = e->box[i];
    return SPK_OK;
}

extern "C" int spk_ensemble_set_graph(spk_ensemble e, int on) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_set_graph: null ensemble");
    e->use_graph = on ? 1 : 0;
    if (!on) drop_graph(e);
    return SPK_OK;
}

extern "C" int spk_ensemble_set_lanes(spk_ensemble e, int lanes) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_set_lanes: null ensemble");
    return assign_lanes(e, lanes);
}

extern "C" int spk_ensemble_set_split(spk_ensemble e, int split) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_set_split: null ensemble");
    SPK_REQUIRE(split >= 0, SPK_ERR_ARG, "spk_ensemble_set_split: split %d", split);
    if (split == 0) return assign_lanes(e, e->nlanes > 0 ? e->nlanes : 1);
    return assign_batches(e, split);
}

extern "C" int spk_ensemble_run(spk_ensemble e, int nsteps) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_run: null ensemble");
    SPK_REQUIRE(nsteps >= 0 && nsteps % 2 == 0, SPK_ERR_ARG,
                "spk_ensemble_run: %d steps (an even count: a replay advances every box by two)", nsteps);
    if (nsteps == 0) return SPK_OK;
    // every box continues from the inverse x pass its previous steps left in W
    // (no box was touched since), or the first two steps start fresh
    const bool cont = all_continue(e);
    const bool warm = cont && all_w_ok(e);
    if (!e->use_graph) {
        for (int r = 0; r < nsteps / 2; r++) SPK_TRY(issue_two_steps(e, r > 0 ? cont : warm));
        for (spk_solver s : e->box) solver_set_w_ok(s, cont);
        return SPK_OK;
    }
    bool stale = e->exec == nullptr;
    for (size_t i = 0; !stale && i < e->box.size(); i++)
        stale = (solver_host_step(e->box[i]) & 1) != e->parity[i];
    if (stale) SPK_TRY(capture(e));
    SPK_TRY(join_lanes_into_origin(e));
    for (int r = 0; r < nsteps / 2; r++) {
        const bool w = e->exec_warm != nullptr && (r > 0 || warm);
        SPK_HIP(hipGraphLaunch(w ? e->exec_warm : e->exec, e->origin));
    }
    SPK_TRY(fork_origin_into_lanes(e));
    for (spk_solver s : e->box) {
        solver_advance_host(s, nsteps);
        solver_set_w_ok(s, e->exec_warm != nullptr);
    }
    return SPK_OK;
}

extern "C" int spk_ensemble_sync(spk_ensemble e) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_sync: null ensemble");
    SPK_HIP(hipStreamSynchronize(e->origin));
    for (int l = 0; l < e->nlanes; l++) SPK_HIP(hipStreamSynchronize(e->lanes[l]));
    return SPK_OK;
}

// Autotune of the launch layout: in batched mode (split >= 1) over splits 1, 2
// and 4, box by box (split 0) over 1, 2, 4, 8 and 16 lanes.  Each candidate
// runs `probe` (even) steps of every box, timed after one warm-up replay; the
// state and clock of every box are saved before and restored after each
// candidate, so tuning leaves the boxes as it found them.  The fastest is kept.
extern "C" int spk_ensemble_tune(spk_ensemble e, int probe, int *best_out, double *best_ms) {
    SPK_REQUIRE(e && probe > 0 && probe % 2 == 0, SPK_ERR_ARG, "spk_ensemble_tune: bad arguments");
    SPK_TRY(spk_ensemble_sync(e));
    const bool batched = e->split > 0;
    const int nb = (int)e->box.size();
    std::vector<void *> backup(nb, nullptr);
    std::vector<size_t> bytes(nb);
    std::vector<double> clock(8 * (size_t)nb);
    std::vector<int> step0(nb);
    int rc = SPK_OK;
    for (int i = 0; i < nb && rc == SPK_OK; i++) {
        int vars[10];
        size_t ntc;
        void *u;
        rc = spk_solver_info(e->box[i], vars, &ntc);
        if (rc == SPK_OK) rc = spk_solver_buffers(e->box[i], &u, nullptr, nullptr);
        if (rc == SPK_OK) rc = spk_solver_get_time(e->box[i], &clock[8 * (size_t)i]);
        if (rc != SPK_OK) break;
        step0[i] = solver_host_step(e->box[i]);
        bytes[i] = ntc * 16 * (size_t)vars[9];
        if (hipMalloc(&backup[i], bytes[i]) != hipSuccess) {
            backup[i] = nullptr;
            set_error("spk_ensemble_tune: no room to save box %d", i);
            rc = SPK_ERR_ALLOC;
            break;
        }
        if (hipMemcpy(backup[i], u, bytes[i], hipMemcpyDeviceToDevice) != hipSuccess) rc = SPK_ERR_HIP;
    }
    auto restore = [&]() -> int {
        SPK_TRY(spk_ensemble_sync(e));
        for (int i = 0; i < nb; i++) {
            void *u;
            SPK_TRY(spk_solver_buffers(e->box[i], &u, nullptr, nullptr));
            SPK_HIP(hipMemcpy(u, backup[i], bytes[i], hipMemcpyDeviceToDevice));
            const double *c = &clock[8 * (size_t)i];
            SPK_TRY(spk_solver_set_time(e->box[i], c[0], (int)c[7], c[6]));
            solver_advance_host(e->box[i], step0[i] - solver_host_step(e->box[i]));
        }
        return SPK_OK;
    };
    int best = batched ? e->split : e->nlanes;
    double tbest = 1e300;
    const std::vector<int> cand = batched ? std::vector<int>{1, 2, 4} : std::vector<int>{1, 2, 4, 8, 16};
    for (int L : cand) {
        if (rc != SPK_OK) break;
        if (!batched && L > nb && L != 1) continue;
        rc = batched ? assign_batches(e, L) : assign_lanes(e, L);
        if (rc == SPK_OK) rc = spk_ensemble_run(e, 2);
        if (rc == SPK_OK) rc = spk_ensemble_sync(e);
        if (rc != SPK_OK) break;
        const auto t0 = std::chrono::steady_clock::now();
        rc = spk_ensemble_run(e, probe);
        if (rc == SPK_OK) rc = spk_ensemble_sync(e);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (rc == SPK_OK && ms < tbest) {
            tbest = ms;
            best = L;
        }
        if (rc == SPK_OK) rc = restore();
    }
    if (rc == SPK_OK) rc = batched ? assign_batches(e, best) : assign_lanes(e, best);
    for (void *b : backup)
        if (b) (void)hipFree(b);
    if (rc != SPK_OK) return rc;
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = tbest;
    return SPK_OK;
}
