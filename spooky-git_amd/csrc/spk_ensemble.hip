// spk_ensemble.hip -- many independent boxes advanced together.
//
// Each box is a solver of its own (grid, physics, coefficients, precision of
// the bulk, its own dt), so an ensemble of uneven boxes (64^3 and 128^3 MHD,
// 512x512x2 Boussinesq-MHD with RKL2 conduction) is a list of solvers.  The
// boxes are spread over HIP streams ("lanes"): a 64^3 transform fills a
// fraction of the 256 CUs, so several boxes must run at once to fill the chip.
// Two steps of every box are captured once into one HIP graph (fork from an
// origin stream into the lanes, join back) and each replay is a single graph
// launch.  A captured step needs no host synchronization: dt stays on the
// device and RKL2 cycles are planned on the device (spk_solver_set_rkl_device).
#include "spk_internal.hpp"

#include <chrono>
#include <vector>

struct spk_ensemble_s {
    int device = 0;
    int bulk = SPK_F64;
    int nlanes = 0;
    int use_graph = 1;
    std::vector<spk_solver> box;
    std::vector<spk_ctx> bctx;        // one context per box; its stream is the box's lane
    std::vector<hipStream_t> lanes;   // MAX_LANES streams, the first nlanes in use
    std::vector<hipEvent_t> ev_join;  // one per lane
    hipStream_t origin = nullptr;     // capture origin and graph launch stream
    hipEvent_t ev_fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> parity;          // host step parity of each box when the graph was captured
};

using namespace spk;

namespace {

constexpr int MAX_LANES = 16;

void drop_graph(spk_ensemble e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
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

// Two steps of every box, boxes interleaved over the lanes (box i on lane i % nlanes).
int issue_two_steps(spk_ensemble e) {
    for (spk_solver s : e->box) SPK_TRY(spk_main_loop(s, 2));
    return SPK_OK;
}

int capture(spk_ensemble e) {
    drop_graph(e);
    SPK_HIP(hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal));
    int rc = fork_origin_into_lanes(e);
    if (rc == SPK_OK) rc = issue_two_steps(e);
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
    e->graph = g;
    const hipError_t ei = hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        e->exec = nullptr;
        drop_graph(e);
        set_error("spk_ensemble: graph instantiation failed: %s", hipGetErrorString(ei));
        return SPK_ERR_HIP;
    }
    e->parity.resize(e->box.size());
    for (size_t i = 0; i < e->box.size(); i++) e->parity[i] = solver_host_step(e->box[i]) & 1;
    return SPK_OK;
}

int assign_lanes(spk_ensemble e, int nlanes) {
    SPK_REQUIRE(nlanes >= 1 && nlanes <= MAX_LANES, SPK_ERR_ARG, "spk_ensemble: %d lanes (1..%d)", nlanes,
                MAX_LANES);
    // finish what is queued before boxes change streams
    for (int l = 0; l < MAX_LANES; l++) SPK_HIP(hipStreamSynchronize(e->lanes[l]));
    SPK_HIP(hipStreamSynchronize(e->origin));
    drop_graph(e);
    e->nlanes = nlanes;
    for (size_t i = 0; i < e->box.size(); i++) e->bctx[i]->stream = e->lanes[i % nlanes];
    return SPK_OK;
}

void destroy_all(spk_ensemble e) {
    for (hipStream_t st : e->lanes) (void)hipStreamSynchronize(st);
    if (e->origin) (void)hipStreamSynchronize(e->origin);
    drop_graph(e);
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
    }
    e->nlanes = nlanes;
    return SPK_OK;
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

extern "C" int spk_ensemble_run(spk_ensemble e, int nsteps) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_run: null ensemble");
    SPK_REQUIRE(nsteps >= 0 && nsteps % 2 == 0, SPK_ERR_ARG,
                "spk_ensemble_run: %d steps (an even count: a replay advances every box by two)", nsteps);
    if (nsteps == 0) return SPK_OK;
    if (!e->use_graph) {
        for (int r = 0; r < nsteps / 2; r++) SPK_TRY(issue_two_steps(e));
        return SPK_OK;
    }
    bool stale = e->exec == nullptr;
    for (size_t i = 0; !stale && i < e->box.size(); i++)
        stale = (solver_host_step(e->box[i]) & 1) != e->parity[i];
    if (stale) SPK_TRY(capture(e));
    SPK_TRY(join_lanes_into_origin(e));
    for (int r = 0; r < nsteps / 2; r++) SPK_HIP(hipGraphLaunch(e->exec, e->origin));
    SPK_TRY(fork_origin_into_lanes(e));
    for (spk_solver s : e->box) solver_advance_host(s, nsteps);
    return SPK_OK;
}

extern "C" int spk_ensemble_sync(spk_ensemble e) {
    SPK_REQUIRE(e, SPK_ERR_ARG, "spk_ensemble_sync: null ensemble");
    SPK_HIP(hipStreamSynchronize(e->origin));
    for (int l = 0; l < e->nlanes; l++) SPK_HIP(hipStreamSynchronize(e->lanes[l]));
    return SPK_OK;
}

// Lane-count autotune: each candidate runs `probe` (even) steps of every box,
// timed after one warm-up replay; the state and clock of every box are saved
// before and restored after each candidate, so tuning leaves the boxes as it
// found them.  The fastest count is kept.
extern "C" int spk_ensemble_tune(spk_ensemble e, int probe, int *best_lanes, double *best_ms) {
    SPK_REQUIRE(e && probe > 0 && probe % 2 == 0, SPK_ERR_ARG, "spk_ensemble_tune: bad arguments");
    SPK_TRY(spk_ensemble_sync(e));
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
    int best = e->nlanes;
    double tbest = 1e300;
    const int cand[] = {1, 2, 4, 8, 16};
    for (int L : cand) {
        if (rc != SPK_OK) break;
        if (L > nb && L != 1) continue;
        rc = assign_lanes(e, L);
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
    if (rc == SPK_OK) rc = assign_lanes(e, best);
    for (void *b : backup)
        if (b) (void)hipFree(b);
    if (rc != SPK_OK) return rc;
    if (best_lanes) *best_lanes = best;
    if (best_ms) *best_ms = tbest;
    return SPK_OK;
}
