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
// The optional autotuner times the candidate reorder intervals and
// per-contact grid sizes on the run's own steps (both leave every result
// bit-identical) and keeps the fastest.

struct GraphCache {
    struct Entry {
        int phase, cur, period, pairBlocks, reorder;
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

int eager_step(dem_solver* s, double dt)
{
    TRY(dem_advance_particles(s, dt, nullptr));
    TRY(dem_compute_forces(s, nullptr, nullptr, 0));
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
    for (int k = 0; k < period && rc == DEM_OK; ++k) rc = eager_step(s, dt);
    s->capturing = false;
    const hipError_t e = hipStreamEndCapture(s->stream, &graph);
    TRY(rc);
    if (e != hipSuccess) return set_error(DEM_E_DEVICE, "end graph capture: %s", hipGetErrorString(e));
    const hipError_t ie = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ie != hipSuccess) return set_error(DEM_E_DEVICE, "instantiate step graph: %s", hipGetErrorString(ie));
    return DEM_OK;
}

int timed_steps(dem_solver* s, double dt, int steps, float* ms)
{
    TRY(eager_step(s, dt));  // settle the pending finish
    HIP_TRY(hipEventRecord(s->ev[0], s->stream), "event");
    for (int k = 0; k < steps; ++k) TRY(eager_step(s, dt));
    TRY(flush_finish(s));
    HIP_TRY(hipEventRecord(s->ev[5], s->stream), "event");
    HIP_TRY(hipEventSynchronize(s->ev[5]), "event");
    HIP_TRY(hipEventElapsedTime(ms, s->ev[0], s->ev[5]), "event");
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
    int done = 0;
    auto step1 = [&]() {
        TRY(eager_step(s, dt));
        ++done;
        return DEM_OK;
    };
    // the first evaluation after initialisation or a load runs eagerly, as do bodies
    const bool graphs = (flags & DEM_RUN_NO_GRAPH) == 0 && s->nb == 0;
    while (done < steps && (s->stepIndex == 0 || !s->pendingFinish)) TRY(step1());
    if ((flags & DEM_RUN_AUTOTUNE) && !G.tuned && steps - done >= 48) {
        const int intervals[4] = {1, 2, 4, 8};
        const int grids[3] = {2048, 8192, 32768};
        float best = 1e30f;
        int bi = s->reorderInterval, bg = s->pairBlocks;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 3; ++b) {
                s->reorderInterval = intervals[a];
                s->pairBlocks = grids[b];
                float ms = 0.0f;
                TRY(timed_steps(s, dt, 2, &ms));
                done += 3;
                if (ms < best) { best = ms; bi = intervals[a]; bg = grids[b]; }
            }
        s->reorderInterval = bi;
        s->pairBlocks = bg;
        G.tuned = true;
    }
    if (!graphs) {
        while (done < steps) TRY(step1());
        if (steps_done) *steps_done = done;
        return DEM_OK;
    }
    // a captured sequence starts with a finish pending (its first kernel is the
    // fused finish + advance), as every replay does
    while (done < steps && !s->pendingFinish) TRY(step1());
    const int P = graph_period(s);
    while (steps - done >= P) {
        const int phase = s->stepIndex % P;
        GraphCache::Entry* hit = nullptr;
        for (auto& e : G.entries)
            if (e.phase == phase && e.cur == s->cur && e.period == P && e.pairBlocks == s->pairBlocks &&
                e.reorder == s->reorderInterval)
                hit = &e;
        if (hit) {
            // the captured steps reorder an even number of times: the buffer is unchanged
            HIP_TRY(hipGraphLaunch(hit->exec, s->stream), "launch step graph");
            s->stepIndex += P;
        } else {
            GraphCache::Entry e{phase, s->cur, P, s->pairBlocks, s->reorderInterval, nullptr};
            TRY(capture_steps(s, dt, P, &e.exec));
            HIP_TRY(hipGraphLaunch(e.exec, s->stream), "launch step graph");
            G.entries.push_back(e);
        }
        done += P;
    }
    while (done < steps) TRY(step1());
    if (steps_done) *steps_done = done;
    return DEM_OK;
}

}  // extern "C"
