// dem_graph.inl -- multi-step runs: the reference's pure-DEM loop body
// (advanceParticles, computeForces without diagnostics, finishParticleStep;
// cuda_simulation_integrator.cpp:263-269) repeated on the device with as
// little host work per step as possible.  Included at the end of dem_hip.hip.
//
// A step launches ~40 small kernels; for small scenes the host launch cost
// dominates.  The step sequence is therefore captured once into a HIP graph
// and replayed.  Kernel arguments that change from step to step repeat with
// a period: the seen-mask parity (2) and the reordering double buffer (two
// reorders bring the state back to the same buffer), so a graph of
// P = 2 * max(1, reorder interval) steps captured at step s0 replays for any
// s0' = s0 (mod P).  The first force evaluation (dt = 0) and runs with bodies
// (their triangle grid needs a host-side count) stay eager.
//
// Steps with candidate lists read a displacement flag back every step and
// run eagerly.  The optional autotuner times candidate lists against the full
// scan with several reorder intervals, each with several per-contact grid
// sizes, on the run's own steps (every choice leaves every result
// bit-identical), and keeps the fastest.

struct GraphCache {
    struct Entry {
        int phase, cur, period, pairBlocks, reorder;
        const double* hist;  // the history buffer in use when captured
        hipGraphExec_t exec;
    };
    std::vector<Entry> entries;
    bool tuned = false;
};

void graph_release(GraphCache* g)
{
    if (!g) return;
    for (auto& e : g->entries)
        if (e.exec) hipGraphExecDestroy(e.exec);
    delete g;
}

namespace {

// One step of a multi-step run; `more`: another step of the run follows, so
// the law kernel may finish this step and advance the next one itself.
int eager_step(dem_solver* s, double dt, bool more)
{
    TRY(dem_advance_particles(s, dt, nullptr));
    s->nextDt = more ? dt : 0.0;
    const int rc = dem_compute_forces(s, nullptr, nullptr, 0);
    s->nextDt = 0.0;
    TRY(rc);
    return dem_finish_particle_step(s, dt, 0.5 * dt);
}

int graph_period(const dem_solver* s) { return 2 * std::max(1, s->reorderInterval); }

// Captures `period` steps starting at the current step; on return the
// solver's host-side counters are those after the captured steps (the caller
// launches the graph once to execute them).
int capture_steps(dem_solver* s, double dt, int period, hipGraphExec_t* out)
{
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "begin graph capture");
    s->capturing = true;
    int rc = DEM_OK;
    for (int k = 0; k < period && rc == DEM_OK; ++k) rc = eager_step(s, dt, k + 1 < period);
    s->capturing = false;
    const hipError_t e = hipStreamEndCapture(s->stream, &graph);
    TRY(rc);
    if (e != hipSuccess) return set_error(DEM_E_DEVICE, "end graph capture: %s", hipGetErrorString(e));
    const hipError_t ie = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) return set_error(DEM_E_DEVICE, "instantiate step graph: %s", hipGetErrorString(ie));
    return DEM_OK;
}

}  // namespace

namespace {

// Runs `count` steps: graph replays when the solver is in full-scan mode
// without bodies, eager launches otherwise.  Returns the steps run in *done.
int run_span(dem_solver* s, double dt, int count, bool allowGraphs, int* done)
{
    GraphCache& G = *s->graphs;
    int k = 0;
    auto step1 = [&](bool more) {
        TRY(eager_step(s, dt, more));
        ++s->nEagerSteps;
        ++k;
        return DEM_OK;
    };
    // the first evaluation after initialisation or a load runs eagerly; a
    // captured sequence starts with a finish pending (its first kernel is the
    // fused finish + advance), as every replay does
    while (k < count && (s->stepIndex == 0 || !s->pendingFinish)) TRY(step1(false));
    const bool graphs = allowGraphs && s->nb == 0 && !s->verlet;
    if (graphs) {
        const int P = graph_period(s);
        while (count - k >= P) {
            const int phase = s->stepIndex % P;
            GraphCache::Entry* hit = nullptr;
            for (auto& e : G.entries)
                if (e.phase == phase && e.cur == s->cur && e.period == P && e.pairBlocks == s->pairBlocks &&
                    e.reorder == s->reorderInterval && e.hist == s->H.ppVal)
                    hit = &e;
            if (hit) {
                // the captured steps reorder an even number of times: the buffer is unchanged
                HIP_TRY(hipGraphLaunch(hit->exec, s->stream), "launch step graph");
                s->stepIndex += P;
                s->scalarsReset = false;  // the replayed evaluations wrote the counters
                // every captured evaluation sorts (graphs run the full scan only)
                s->nEvaluations += P;
                s->nSorts += P;
                s->lastSorted = true;
            } else {
                GraphCache::Entry e{phase, s->cur, P, s->pairBlocks, s->reorderInterval, s->H.ppVal, nullptr};
                TRY(capture_steps(s, dt, P, &e.exec));
                HIP_TRY(hipGraphLaunch(e.exec, s->stream), "launch step graph");
                s->scalarsReset = false;
                G.entries.push_back(e);
            }
            k += P;
            s->nGraphSteps += P;
        }
    }
    while (k < count) TRY(step1(k + 1 < count));
    *done += k;
    // replayed steps read nothing back: the span ends with the range count of
    // all its evaluations
    HIP_TRY(hipMemcpyAsync(&s->hScalars->grid_overflow, &s->sc->grid_overflow, sizeof(int), hipMemcpyDeviceToHost,
                           s->stream),
            "read range count");
    HIP_TRY(hipStreamSynchronize(s->stream), "read range count");
    if (range_seen(s)) return range_error(s);
    return DEM_OK;
}

}  // namespace

extern "C" {

int dem_run_steps(dem_solver* s, int steps, double dt, int flags, int* steps_done)
{
    if (!s || !s->initialized) return set_error(DEM_E_INVALID, "CUDA DEM solver is not initialized");
    if (steps < 0) return set_error(DEM_E_INVALID, "negative step count");
    if (steps_done) *steps_done = 0;
    if (!s->graphs) s->graphs = new GraphCache();
    GraphCache& G = *s->graphs;
    const bool allowGraphs = (flags & DEM_RUN_NO_GRAPH) == 0;
    int done = 0;
    if ((flags & DEM_RUN_AUTOTUNE) && !G.tuned && steps >= 96) {
        // candidates: candidate lists (when available) or the full scan with a
        // reorder interval, each with three contact-grid sizes; each timed on
        // a few steps of the run itself, graphs included
        struct Cfg { bool verlet; int reorder, grid; };
        std::vector<Cfg> cfgs;
        const int grids[3] = {2048, 8192, 32768};
        const bool canLists = s->refC != nullptr && s->nb == 0 && !s->verletBroken;
        // clumps keep their rows: reorder interval 0 only
        const std::vector<int> reorders = s->clump ? std::vector<int>{0} : std::vector<int>{1, 4, 8};
        for (int g : grids) {
            if (canLists) cfgs.push_back({true, s->reorderInterval, g});
            for (int r : reorders) cfgs.push_back({false, r, g});
        }
        const int span = 4;
        if (steps >= (int)cfgs.size() * (span + 4) + 8) {
            float best = 1e30f;
            Cfg pick{s->verlet, s->reorderInterval, s->pairBlocks};
            for (const Cfg& c : cfgs) {
                if (c.verlet != s->verlet) s->vlistValid = false;
                s->verlet = c.verlet && canLists;
                s->reorderInterval = c.reorder;
                s->pairBlocks = c.grid;
                TRY(run_span(s, dt, 4, allowGraphs, &done));  // captures graphs, settles the mode
                TRY(flush_finish(s));
                HIP_TRY(hipEventRecord(s->ev[0], s->stream), "event");
                TRY(run_span(s, dt, span, allowGraphs, &done));
                TRY(flush_finish(s));
                HIP_TRY(hipEventRecord(s->ev[5], s->stream), "event");
                HIP_TRY(hipEventSynchronize(s->ev[5]), "event");
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, s->ev[0], s->ev[5]), "event");
                if (s->verlet == c.verlet && ms < best) {
                    best = ms;
                    pick = c;
                }
            }
            if (pick.verlet != s->verlet) s->vlistValid = false;
            s->verlet = pick.verlet && canLists && !s->verletBroken;
            s->reorderInterval = pick.reorder;
            s->pairBlocks = pick.grid;
            G.tuned = true;
        }
    }
    TRY(run_span(s, dt, steps - done, allowGraphs, &done));
    if (steps_done) *steps_done = done;
    return DEM_OK;
}

}  // extern "C"
