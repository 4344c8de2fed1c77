// sph_capi.hip -- context management and the extern "C" boundary (sphmhd.h).
//
// The host side only sequences launches on the context's stream; every
// per-particle and per-pair operation runs in the kernels of sph_nnps.hip,
// sph_pairs.hip and sph_stage.hip.  An evaluation enqueues the neighbour
// search, speculative density passes and the force pass without waiting,
// then waits once (the reference also blocks per iteration:
// equations.py:91-94 pulls equation_has_converged).  Inside sphmhd_step the
// enqueue-only part of every step after the first is captured once as a
// hipGraph and replayed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "sph_internal.h"

using namespace sph;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

// the library's error record for the other translation units
int sph::report_error(int code, const std::string& msg) { return fail(code, msg); }

namespace {

#define HIPCHK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return fail(SPHMHD_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(T** p, int64_t count) {
    *p = nullptr;
    if (count <= 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)count);
    if (e != hipSuccess) return fail(SPHMHD_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemset(*p, 0, sizeof(T) * (size_t)count);
    if (e != hipSuccess) return fail(SPHMHD_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return 0;
}

}  // namespace

// HIP-event timer of one phase: events are recorded around launches and
// read after the next synchronisation of the stream.
struct PhaseTimer {
    hipEvent_t a = nullptr, b = nullptr;
    bool pending = false;
    double ms = 0.0;
    int64_t count = 0;
};

// the enqueue-only part of a step, in timed phases: stages, neighbour
// search, density passes, prep + lists, force pass, then the tail (warning
// collection and the control-block copy)
enum { P_STAGE = 0, P_NNPS, P_DENSITY, P_LISTS, P_FORCE, kSegs };

struct sphmhd_ctx {
    Dev d{};
    bool timing = false;
    PhaseTimer tm[5];  // force, density, nnps, lists, stage
    std::vector<Inst> hinst;
    std::vector<int64_t> offsets;
    bool bound = false;
    int max_iter_all = 1;
    int min_iter_all = 1;
    bool lean_now = false;  // the device's Ctl::lean (see sphmhd_step)
    bool persist_fixed = false;  // launch modes set by the caller or SPHMHD_PERSIST, not tuned
    int spec_passes = 2;  // density passes launched before the first wait
    int spec_fixed = 0;   // SPHMHD_SPEC: a fixed count instead (results do not depend on it)
    int spec_margin = 0;       // one extra pass for a while after a redo
    int evals_since_redo = 0;
    // Host mirror of the control block (pinned, mapped): slots 0 and 1 are
    // written by the kernel that ends an evaluation (slot seq % 2), slot 2 by
    // synchronous reads; hcur is the record the host logic reads.
    CtlView* hctl = nullptr;
    CtlView* hcur = nullptr;
    int32_t seq_next = 0;  // seq the next enqueued evaluation end will carry
    hipEvent_t sev = nullptr;  // end of the step being waited for (pipelined runs)
    sphmhd_stats stats{};
    // captured steps, one per speculative pass count; generation bumps on
    // bind, list growth and timing changes invalidate them
    // With timing on, the step is captured as one graph per timed phase
    // (launched between HIP events, which are not captured); otherwise as
    // one graph.
    struct StepGraph {
        hipGraphExec_t seg[kSegs] = {};
        int nseg = 0;
        int spec = -1;
        uint64_t gen = 0;
        bool timed = false;
    };
    std::vector<StepGraph> graphs;
    uint64_t graph_gen = 1;
    int prune_mode = -1;  // list kernel: -1 by the last evaluation's grids, 0 plain, 1 pruned
    bool stage_began = false;  // the stage launch of this front began the evaluation
    bool use_graphs = true;
    bool capturing = false;
    // Multi-step runs go on a stream of the library's own (a capture cannot
    // use the legacy default stream the caller may pass), ordered after the
    // caller's stream by an event and synchronised before returning.
    hipStream_t gstream = nullptr;
    hipEvent_t gev = nullptr;
    // tangent-linear step
    TanDev tg;
    bool tbound = false;
    // adjoint step: bar rows and work arrays, and the recorded rollout
    // (checkpoints of the state every `every` steps, each with its clocks)
    struct Adj {
        bool bound = false;
        AdjDev a{};
        bool recorded = false;
        bool warm = false;
        int every = 1;
        int64_t steps = 0;
        double* init = nullptr;      // SPHMHD_NFIELDS x n + 2 x ninst: the state before the rollout
        std::vector<double*> ckpt;   // per segment: kNState x n + 2 x ninst
        double* final_clk = nullptr; // 2 x ninst: the clock after the last step
        double* seg = nullptr;       // every x kNState x n: the states of one segment's steps
        double* clk = nullptr;       // (every + 1) x 2 x ninst: their clocks, then the clock after
        int seg_cap = 0;
        double* ain_live = nullptr;  // the entry rows of evaluations run on the live context
        // each step of a segment after its evaluation: what its reverse reads
        // (property rows, the evaluation's entry, force pack, cells, lists,
        // instances), so a step is reversed without computing it again
        struct Snap {
            double* rows = nullptr;   // kSnapRows x n, then the entry rows (5 x n)
            double* fpack = nullptr;  // kFRecMax x n
            int32_t* ints = nullptr;  // perm, keys, nbcnt (3 x n)
            int32_t* nb = nullptr;
            int64_t nb_cap = 0;
            int32_t* cells = nullptr; // cstart, cend of the used cells
            int64_t cells_cap = 0;
            Inst* inst = nullptr;
            int stride = 0;
            int64_t ncells = 0;
        };
        std::vector<Snap> snaps;
    } adj;
    // launch-shape tuner: the first kTuneSteps steps of sphmhd_step rotate
    // the tuned launches through the configurations (workgroup size, which is
    // also the tile of particles a work item covers, and plain or persistent
    // launch; the results do not depend on either) and time each launch class
    // with HIP events
    struct Tuner {
        int state = 0;  // 0: not started, 1: running, 2: tuned, 3: fixed (caller or SPHMHD_TUNE=0)
        int step = 0;
        double best_ms[kLaunchClasses][6];   // fastest sample per class and configuration
        double worst_ms[kLaunchClasses][6];  // slowest sample
        bool saved_timing = false;
        double saved_ms[5];
        int64_t saved_count[5];
    } tune;
};

static const int kShapes[3] = {64, 128, 256};
// configuration c: workgroup size kShapes[c % 3], persistent when c >= 3;
// only the first three (plain) when the launch mode is fixed
constexpr int kTuneConfigs = 6;
constexpr int kTuneRounds = 2;
constexpr int kTuneSteps = kTuneConfigs * kTuneRounds;
static int tune_configs(const sphmhd_ctx* c);

enum { T_FORCE = 0, T_DENSITY, T_NNPS, T_LISTS, T_STAGE };

static int tune_configs(const sphmhd_ctx* c) { return c->persist_fixed ? 3 : kTuneConfigs; }

// Inside a stream capture the records become event nodes of the graph
// (hipEventRecordExternal), re-recorded at every replay.
// no records while a graph is being captured: timed graphs are cut at the
// phase boundaries and the events go between their launches
static void tbegin(sphmhd_ctx* c, int k) {
    if (c->timing && !c->capturing && !c->tm[k].pending) (void)hipEventRecord(c->tm[k].a, c->d.stream);
}

static void tend(sphmhd_ctx* c, int k) {
    if (c->timing && !c->capturing && !c->tm[k].pending) {
        (void)hipEventRecord(c->tm[k].b, c->d.stream);
        c->tm[k].pending = true;
    }
}

// fold finished event pairs into the sums (call after a stream sync)
static void tcollect(sphmhd_ctx* c) {
    if (!c->timing) return;
    for (PhaseTimer& t : c->tm) {
        if (!t.pending) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            t.ms += ms;
            t.count += 1;
        }
        t.pending = false;
    }
    // a failed query only drops that sample: keep it out of the sticky error
    // state that launch checks read
    (void)hipGetLastError();
}

// ---- launch-shape tuner (see sphmhd_set_launch) ----
static int tune_events(sphmhd_ctx* c) {
    for (PhaseTimer& t : c->tm)
        if (!t.a) {
            HIPCHK(hipEventCreate(&t.a));
            HIPCHK(hipEventCreate(&t.b));
        }
    return 0;
}

// before a tuning step: this step's shape, fresh sums (the caller's saved)
static int tune_begin_step(sphmhd_ctx* c) {
    auto& T = c->tune;
    if (T.state == 0) {
        int rc = tune_events(c);
        if (rc) return rc;
        T.state = 1;
        T.step = 0;
        for (auto& row : T.best_ms)
            for (double& v : row) v = 1.0e300;
        for (auto& row : T.worst_ms)
            for (double& v : row) v = 0.0;
        T.saved_timing = c->timing;
        for (int k = 0; k < 5; ++k) {
            T.saved_ms[k] = c->tm[k].ms;
            T.saved_count[k] = c->tm[k].count;
        }
    }
    const int cfg = T.step % tune_configs(c);
    for (int q = 0; q < kLaunchClasses; ++q) {
        c->d.blk[q] = kShapes[cfg % 3];
        if (!c->persist_fixed) c->d.persist[q] = cfg >= 3 ? 1 : 0;
    }
    c->timing = true;
    for (PhaseTimer& t : c->tm) {
        t.ms = 0.0;
        t.count = 0;
        t.pending = false;
    }
    return 0;
}

// after it (the evaluation's wait has collected the events): keep the
// fastest sample per shape and class; after the last one pick the shapes
// (a smaller workgroup must be 3% faster to replace a larger one) and give
// the caller back its timing mode and sums
static void tune_end_step(sphmhd_ctx* c) {
    auto& T = c->tune;
    const int ncfg = tune_configs(c);
    const int s = T.step % ncfg;
    const int timer[kLaunchClasses] = {T_DENSITY, T_LISTS, T_FORCE};
    for (int q = 0; q < kLaunchClasses; ++q) {
        const PhaseTimer& t = c->tm[timer[q]];
        if (t.count > 0) {
            T.best_ms[q][s] = std::min(T.best_ms[q][s], t.ms);
            T.worst_ms[q][s] = std::max(T.worst_ms[q][s], t.ms);
        }
    }
    if (++T.step < ncfg * kTuneRounds) return;
    // the default unless another configuration's slowest sample beats the
    // default's fastest by 5% (the early steps the samples come from differ
    // from each other by a few percent).  The default is 256 plain, and 256
    // persistent in 3-D: a 3-D state that develops a spread of h (the blast's
    // cavity and shell) loads the pair kernels unevenly, which persistent
    // workgroups pulling tiles even out (blast 128^3: 6.8 -> 6.2 ms per step),
    // while the early steps the tuner times show no difference yet
    const int def = (c->d.dim == 3 && ncfg == kTuneConfigs) ? 5 : 2;
    for (int q = 0; q < kLaunchClasses; ++q) {
        int best = def;
        for (int k = 0; k < ncfg; ++k)
            if (k != def && T.worst_ms[q][k] < 0.95 * (best == def ? T.best_ms[q][def] : T.worst_ms[q][best]))
                best = k;
        c->d.blk[q] = kShapes[best % 3];
        if (!c->persist_fixed) c->d.persist[q] = best >= 3 ? 1 : 0;
    }
    T.state = 2;
    c->timing = T.saved_timing;
    for (int k = 0; k < 5; ++k) {
        c->tm[k].ms = T.saved_ms[k];
        c->tm[k].count = T.saved_count[k];
        c->tm[k].pending = false;
    }
    c->graph_gen += 1;  // captured steps hold the old shapes
}

// The list kernel follows the last evaluation's grids: the pruned one where
// some grid was split (both give the same lists; the pruning state costs the
// unsplit case occupancy).  Captured steps hold the kernel, so a change
// re-captures them once the queue has drained.
static int sync_prune(sphmhd_ctx* c) {
    const bool p = c->prune_mode >= 0 ? c->prune_mode != 0 : c->hcur->any_split != 0;
    if (p == c->d.prune_lists) return 0;
    HIPCHK(hipStreamSynchronize(c->d.stream));
    c->d.prune_lists = p;
    c->graph_gen += 1;
    return 0;
}

static int read_ctl(sphmhd_ctx* c) {
    launch_ctl_export(c->d);
    c->hcur = &c->hctl[2];
    HIPCHK(hipStreamSynchronize(c->d.stream));
    tcollect(c);
    return sync_prune(c);
}

static int pull_inst(sphmhd_ctx* c) {
    HIPCHK(hipMemcpyAsync(c->hinst.data(), c->d.inst, sizeof(Inst) * c->hinst.size(), hipMemcpyDeviceToHost,
                          c->d.stream));
    HIPCHK(hipStreamSynchronize(c->d.stream));
    return 0;
}

static int push_inst(sphmhd_ctx* c) {
    HIPCHK(hipMemcpyAsync(c->d.inst, c->hinst.data(), sizeof(Inst) * c->hinst.size(), hipMemcpyHostToDevice,
                          c->d.stream));
    HIPCHK(hipStreamSynchronize(c->d.stream));
    return 0;
}

static int grow_nb(sphmhd_ctx* c, int need) {
    int stride = ((need + 31) / 32) * 32;
    HIPCHK(hipStreamSynchronize(c->d.stream));
    HIPCHK(hipFree(c->d.nb));
    c->d.nb = nullptr;
    int rc = dalloc(&c->d.nb, (int64_t)stride * c->d.n);
    if (rc) return rc;
    c->d.nb_stride = stride;
    c->graph_gen += 1;
    return 0;
}

// neighbour lists for the current {x,y,z,h} pack; grows the slots and
// rebuilds when a particle has more neighbours than slots
static int build_lists(sphmhd_ctx* c) {
    for (int attempt = 0; attempt < 2; ++attempt) {
        launch_nb_build(c->d, true);
        int rc = read_ctl(c);
        if (rc) return rc;
        if (c->hcur->max_nb <= c->d.nb_stride) return 0;
        rc = grow_nb(c, c->hcur->max_nb);
        if (rc) return rc;
        launch_reset_nb_stats(c->d);
    }
    return fail(SPHMHD_ECAP, "neighbour list did not fit after growing");
}

// begun: the evaluation's start (eval_begin_body with reset_dt), the h_max
// reduction and the grid set-up already ran in the fused stage kernel;
// pack: the sort also writes the density pack
static int nnps_rebuild(sphmhd_ctx* c, bool reset_clock_reductions, bool begun = false, bool pack = false) {
    if (!begun) {
        launch_eval_begin(c->d, reset_clock_reductions);
        launch_hmax_grid(c->d);
    }
    launch_keys_sort_ranges(c->d, pack);
    HIPCHK(hipGetLastError());
    return 0;
}

// prep (G1, G2, G4 and the force pack), neighbour lists, fused force pass
// and the warning collection; first: the first pass of an evaluation, which
// keeps the previous accelerations a redone pass needs
static int force_pass(sphmhd_ctx* c, bool first) {
    Dev& d = c->d;
    tbegin(c, T_LISTS);
    launch_prep_force(d, first);
    launch_nb_build(d, true);
    tend(c, T_LISTS);
    tbegin(c, T_FORCE);
    launch_force(d);  // ends with the evaluation's end (warning bits, host mirror)
    tend(c, T_FORCE);
    c->seq_next += 1;
    HIPCHK(hipGetLastError());
    return 0;
}

// One evaluation with one host wait.  The density passes are launched
// speculatively (as many as recent evaluations needed, at least
// min_iterations; a pass launched after an instance has finished is a no-op
// for it) and the force pass follows at once (eval_front, enqueue only).  The
// host then reads the iteration flags and the list overflow flag together
// (eval_finish).  Rarely the iteration needed more passes, or a list did not
// fit its slots: the remaining passes run (one wait each) or the lists grow,
// and the force pass is redone from the kept previous accelerations.  After a
// redo one extra speculative pass is launched for the next 16 evaluations.
// With fused density passes the first pass ends the iteration (passes:
// false), so one is launched; the tangent's passes are launched one by one.
static int next_spec(const sphmhd_ctx* c, bool passes = false) {
    if (c->spec_fixed > 0) return c->spec_fixed;
    if (c->d.dens_fused && !passes) return 1;
    return std::max(c->spec_passes + c->spec_margin, 1);
}

static const int kPhaseTimer[kSegs] = {T_STAGE, T_NNPS, T_DENSITY, T_LISTS, T_FORCE};

// the launches of one phase of the enqueue-only step front
static int front_phase(sphmhd_ctx* c, int ph, bool first, int spec) {
    Dev& d = c->d;
    switch (ph) {
    case P_STAGE:
        // both kernels end by beginning the evaluation and setting up its grid
        if (first) launch_init_stage1(d);
        else launch_stage2_init_stage1(d);  // also commits the clock
        c->stage_began = true;
        break;
    case P_NNPS: {
        const int rc = nnps_rebuild(c, true, !first || c->stage_began, true);
        c->stage_began = false;
        if (rc) return rc;
        break;
    }
    case P_DENSITY:
        for (int q = 0; q < spec; ++q) {
            d.pass = q;
            launch_density(d);  // ends with the group control
        }
        break;
    case P_LISTS:
        launch_prep_force(d, true);
        launch_nb_build(d, true);
        break;
    default:
        launch_force(d);  // ends with the evaluation's end (warning bits, host mirror)
        break;
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// phases [from, kSegs) with their timers
static int front_range(sphmhd_ctx* c, int from, bool first, int spec) {
    for (int ph = from; ph < kSegs; ++ph) {
        const int k = kPhaseTimer[ph];
        if (k >= 0) tbegin(c, k);
        const int rc = front_phase(c, ph, first, spec);
        if (rc) return rc;
        if (k >= 0) tend(c, k);
    }
    return 0;
}

// a standalone evaluation: no fused stage kernel ran before it, so the
// evaluation's start is launched (first = true)
static int eval_front(sphmhd_ctx* c, int spec) {
    const int rc = front_range(c, P_NNPS, true, spec);
    c->seq_next += 1;  // the force pass ends the evaluation
    return rc;
}

// the record of the evaluation that ended with sequence number sq
static int pick_record(sphmhd_ctx* c, int32_t sq) {
    c->hcur = &c->hctl[sq & 1];
    if (c->hcur->seq != sq) return fail(SPHMHD_ESTATE, "control record out of sequence");
    return sync_prune(c);
}

// After the evaluation's record is in c->hcur: the rest of the density
// passes and the force pass again when the iteration needed more passes or a
// list did not fit (the record's abort flag, which also made every kernel
// enqueued after it exit at entry), then the statistics and warnings.
static int eval_process(sphmhd_ctx* c, int spec) {
    Dev& d = c->d;
    int rc = 0;
    int passes = spec;
    bool redo = false;
    if (c->hcur->abort) launch_clear_abort(d);
    if (!c->hcur->all_done) {
        // the prep overwrote the {x,y,z,m} density pack with {x,y,z,h}
        launch_pack_density(d);
        while (!c->hcur->all_done) {
            if (passes > c->max_iter_all + 1) return fail(SPHMHD_ESTATE, "density iteration control did not stop");
            tbegin(c, T_DENSITY);
            d.pass = passes++;
            launch_density(d);
            tend(c, T_DENSITY);
            HIPCHK(hipGetLastError());
            if ((rc = read_ctl(c))) return rc;
        }
        redo = true;
        c->spec_margin = 1;
        c->evals_since_redo = 0;
    } else if (c->spec_margin && ++c->evals_since_redo >= 16) {
        c->spec_margin = 0;
    }
    c->spec_passes = std::max(c->hcur->iters_max, c->min_iter_all);
    c->stats.density_iterations += c->hcur->iters_max;
    c->stats.last_iterations_max = c->hcur->iters_max;
    for (int grows = 0;;) {
        if (c->hcur->overflow) {
            if (++grows > 2) return fail(SPHMHD_ECAP, "neighbour list did not fit after growing");
            if ((rc = grow_nb(c, c->hcur->max_nb))) return rc;
            redo = true;
        }
        if (!redo) break;
        launch_clear_abort(d);
        launch_reset_nb_stats(d);
        launch_reset_force_outputs(d);
        if ((rc = force_pass(c, false))) return rc;
        if ((rc = read_ctl(c))) return rc;
        redo = false;
        c->stats.force_redos += 1;
    }
    // a redone pass of the step that reaches tf raised the flag again
    if (c->hcur == &c->hctl[2] && c->hcur->abort) launch_clear_abort(d);
    c->stats.max_neighbours = c->hcur->max_nb;
    c->stats.pairs_force = c->hcur->pairs;
    c->stats.cells = c->hcur->total_cells;
    c->stats.evals += 1;
    int st = c->hcur->status;
    if (st & 2) return SPHMHD_WNAN;
    if (st & 1) return SPHMHD_WNOCONV;
    return 0;
}

// the one host wait of an evaluation enqueued last
static int eval_finish(sphmhd_ctx* c, int spec) {
    HIPCHK(hipStreamSynchronize(c->d.stream));
    tcollect(c);
    const int rc = pick_record(c, c->seq_next - 1);
    if (rc) return rc;
    return eval_process(c, spec);
}

static int do_eval(sphmhd_ctx* c) {
    const int spec = next_spec(c);
    int rc = eval_front(c, spec);
    if (rc) return rc;
    return eval_finish(c, spec);
}

// enqueue-only part of one step: the stages (the first step of a run starts
// with initialize + stage1, later ones with the fused stage2 + initialize +
// stage1 and the clock) and the evaluation front
static int step_front(sphmhd_ctx* c, bool first, int spec) {
    const int rc = front_range(c, P_STAGE, first, spec);
    c->seq_next += 1;  // the force pass ends the evaluation
    return rc;
}

static void release(sphmhd_ctx::StepGraph& g) {
    for (hipGraphExec_t& e : g.seg)
        if (e) {
            (void)hipGraphExecDestroy(e);
            e = nullptr;
        }
    g.nseg = 0;
}

static void drop_graphs(sphmhd_ctx* c) {
    for (auto& g : c->graphs) release(g);
    c->graphs.clear();
}

// captures phases [a, b) of a later step into one executable graph
static int capture_phases(sphmhd_ctx* c, int a, int b, int spec, hipGraphExec_t* out) {
    *out = nullptr;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamBeginCapture(c->d.stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = 0;
    for (int ph = a; ph < b && rc == 0; ++ph) rc = front_phase(c, ph, false, spec);
    c->capturing = false;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->d.stream, &graph);
    const hipError_t ei = (rc == 0 && e == hipSuccess) ? hipGraphInstantiate(out, graph, nullptr, nullptr, 0)
                                                       : hipErrorUnknown;
    if (graph) (void)hipGraphDestroy(graph);
    if (rc || e != hipSuccess || ei != hipSuccess) {
        std::fprintf(stderr, "sphmhd: step graph capture failed (%s; end capture: %s; instantiate: %s); "
                             "using plain launches\n",
                     rc ? g_err.c_str() : "launches ok", hipGetErrorString(e), hipGetErrorString(ei));
        (void)hipGetLastError();
        if (*out) (void)hipGraphExecDestroy(*out);
        *out = nullptr;
        return 1;
    }
    return 0;
}

// replays the captured front of a later step, capturing it first when no
// graph for this pass count, timing mode and generation exists; falls back
// to plain launches (and stops using graphs) if a capture fails
static int step_front_graph(sphmhd_ctx* c, int spec) {
    sphmhd_ctx::StepGraph* g = nullptr;
    for (auto& e : c->graphs)
        if (e.spec == spec && e.gen == c->graph_gen && e.timed == c->timing) g = &e;
    if (!g) {
        for (auto& e : c->graphs)
            if (e.gen != c->graph_gen) release(e);
        c->graphs.erase(std::remove_if(c->graphs.begin(), c->graphs.end(),
                                       [](const sphmhd_ctx::StepGraph& e) { return e.nseg == 0; }),
                        c->graphs.end());
        sphmhd_ctx::StepGraph ng;
        ng.spec = spec;
        ng.gen = c->graph_gen;
        ng.timed = c->timing;
        ng.nseg = ng.timed ? kSegs : 1;
        for (int q = 0; q < ng.nseg; ++q) {
            const int a = ng.timed ? q : 0, b = ng.timed ? q + 1 : kSegs;
            if (capture_phases(c, a, b, spec, &ng.seg[q])) {
                release(ng);
                c->use_graphs = false;
                return step_front(c, false, spec);
            }
        }
        c->graphs.push_back(ng);
        g = &c->graphs.back();
    }
    for (int q = 0; q < g->nseg; ++q) {
        const int k = g->timed ? kPhaseTimer[q] : -1;
        if (k >= 0) tbegin(c, k);
        HIPCHK(hipGraphLaunch(g->seg[q], c->d.stream));
        if (k >= 0) tend(c, k);
    }
    c->stats.graph_replays += 1;
    c->seq_next += 1;  // the captured force pass ends the evaluation
    return 0;
}

static int check_ready(sphmhd_ctx* c) {
    if (!c) return fail(SPHMHD_EINVAL, "null context");
    if (!c->bound) return fail(SPHMHD_ESTATE, "no fields bound");
    return 0;
}

int sph::ctx_view(sphmhd_ctx* c, Dev** d, int64_t* longest) {
    int rc = check_ready(c);
    if (rc) return rc;
    *d = &c->d;
    int64_t m = 0;
    for (size_t k = 0; k + 1 < c->offsets.size(); ++k) m = std::max(m, c->offsets[k + 1] - c->offsets[k]);
    *longest = m;
    return 0;
}

static int count_unconverged(sphmhd_ctx* c, int64_t* out) {
    launch_count_unconverged(c->d);
    int rc = read_ctl(c);
    if (rc) return rc;
    *out = c->hcur->count;
    return 0;
}

// ---------------------------------------------------------------------------
// adjoint step: host sequencing
// ---------------------------------------------------------------------------

static void adj_free(sphmhd_ctx* c) {
    sphmhd_ctx::Adj& A = c->adj;
    AdjDev& a = A.a;
    a.ain = A.ain_live;
    A.ain_live = nullptr;
    for (void* p : {(void*)a.acc, (void*)a.argmax, (void*)a.par, (void*)a.dtrow, (void*)a.ain, (void*)a.pass,
                    (void*)a.npass, (void*)a.bsave, (void*)a.ai, (void*)A.init, (void*)A.final_clk, (void*)A.seg,
                    (void*)A.clk})
        if (p) (void)hipFree(p);
    for (double* p : A.ckpt)
        if (p) (void)hipFree(p);
    A.ckpt.clear();
    for (auto& sn : A.snaps)
        for (void* p : {(void*)sn.rows, (void*)sn.fpack, (void*)sn.ints, (void*)sn.nb, (void*)sn.cells, (void*)sn.inst})
            if (p) (void)hipFree(p);
    A.snaps.clear();
    double* keep[SPHMHD_NFIELDS];
    std::memcpy(keep, a.b, sizeof keep);
    a = AdjDev{};
    std::memcpy(a.b, keep, sizeof keep);
    A.init = A.final_clk = A.seg = A.clk = nullptr;
    A.seg_cap = 0;
    A.recorded = false;
}

static RowList rows_state() {
    RowList L;
    L.count = kNState;
    for (int r = 0; r < kNState; ++r) L.f[r] = kStateFields[r];
    return L;
}

static RowList rows_all() {
    RowList L;
    L.count = SPHMHD_NFIELDS;
    for (int r = 0; r < SPHMHD_NFIELDS; ++r) L.f[r] = r;
    return L;
}

// the property rows the reverse kernels of a step read after its evaluation
static const int kSnapFields[] = {
    SPHMHD_F_X, SPHMHD_F_Y, SPHMHD_F_Z, SPHMHD_F_H, SPHMHD_F_RHO, SPHMHD_F_M, SPHMHD_F_BX, SPHMHD_F_BY,
    SPHMHD_F_BZ, SPHMHD_F_U, SPHMHD_F_V, SPHMHD_F_W, SPHMHD_F_EPRED, SPHMHD_F_AU, SPHMHD_F_AV, SPHMHD_F_AW,
    SPHMHD_F_ABEVOLX, SPHMHD_F_ABEVOLY, SPHMHD_F_ABEVOLZ, SPHMHD_F_UPRED, SPHMHD_F_VPRED, SPHMHD_F_WPRED,
    SPHMHD_F_PSIPRED, SPHMHD_F_OMEGA, SPHMHD_F_H0, SPHMHD_F_CS, SPHMHD_F_BEVOLX, SPHMHD_F_BEVOLY,
    SPHMHD_F_BEVOLZ, SPHMHD_F_BEVOLXPRED, SPHMHD_F_BEVOLYPRED, SPHMHD_F_BEVOLZPRED, SPHMHD_F_APSI,
    SPHMHD_F_ALPHA10, SPHMHD_F_ALPHA1, SPHMHD_F_AE, SPHMHD_F_DT_ADAPT};
constexpr int kSnapRows = (int)(sizeof(kSnapFields) / sizeof(int));

static RowList rows_snap() {
    RowList L;
    L.count = kSnapRows;
    for (int r = 0; r < kSnapRows; ++r) L.f[r] = kSnapFields[r];
    return L;
}

template <class T>
static int grow(T** p, int64_t* cap, int64_t need) {
    if (*p && *cap >= need) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    int rc = dalloc(p, need);
    if (rc) return rc;
    *cap = need;
    return 0;
}

// the step just evaluated, kept for its reverse (slot j of the segment);
// the entry rows were written there before the evaluation
static int adj_snap_save(sphmhd_ctx* c, sphmhd_ctx::Adj::Snap& sn) {
    Dev& d = c->d;
    const int64_t n = d.n;
    int rc;
    launch_a_rows(d, rows_snap(), sn.rows, false);
    HIPCHK(hipMemcpyAsync(sn.fpack, d.fpack, sizeof(double) * kFRecMax * n, hipMemcpyDeviceToDevice, d.stream));
    HIPCHK(hipMemcpyAsync(sn.ints, d.perm, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, d.stream));
    HIPCHK(hipMemcpyAsync(sn.ints + n, d.keys, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, d.stream));
    HIPCHK(hipMemcpyAsync(sn.ints + 2 * n, d.nbcnt, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, d.stream));
    const int64_t nbn = (int64_t)d.nb_stride * n;
    if ((rc = grow(&sn.nb, &sn.nb_cap, nbn))) return rc;
    HIPCHK(hipMemcpyAsync(sn.nb, d.nb, sizeof(int32_t) * nbn, hipMemcpyDeviceToDevice, d.stream));
    sn.stride = d.nb_stride;
    // the whole cell table: the recomputation enqueues evaluations without
    // waiting for their control records, so the host does not know this
    // evaluation's cell count
    const int64_t tc = std::max<int64_t>(d.cell_cap, 1);
    if ((rc = grow(&sn.cells, &sn.cells_cap, 2 * tc))) return rc;
    HIPCHK(hipMemcpyAsync(sn.cells, d.cstart, sizeof(int32_t) * tc, hipMemcpyDeviceToDevice, d.stream));
    HIPCHK(hipMemcpyAsync(sn.cells + tc, d.cend, sizeof(int32_t) * tc, hipMemcpyDeviceToDevice, d.stream));
    sn.ncells = tc;
    HIPCHK(hipMemcpyAsync(sn.inst, d.inst, sizeof(Inst) * d.ninst, hipMemcpyDeviceToDevice, d.stream));
    return 0;
}

// the context as it was after that step's evaluation, for the reverse
// kernels (the live rows elsewhere are never read by them)
static Dev adj_snap_view(sphmhd_ctx* c, const sphmhd_ctx::Adj::Snap& sn) {
    Dev v = c->d;
    const int64_t n = v.n;
    for (int r = 0; r < kSnapRows; ++r)
        if (v.fields.f[kSnapFields[r]]) v.fields.f[kSnapFields[r]] = sn.rows + (int64_t)r * n;
    v.fpack = sn.fpack;
    v.perm = sn.ints;
    v.keys = (uint32_t*)(sn.ints + n);
    v.nbcnt = sn.ints + 2 * n;
    v.nb = sn.nb;
    v.nb_stride = sn.stride;
    v.cstart = sn.cells;
    v.cend = sn.cells + sn.ncells;
    v.inst = sn.inst;
    return v;
}

static int adj_check(sphmhd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!c->adj.bound) return fail(SPHMHD_ESTATE, "no adjoint rows bound");
    return 0;
}

// the work arrays of the reverse sweep (once per context)
static int adj_alloc_work(sphmhd_ctx* c) {
    Dev& d = c->d;
    AdjDev& a = c->adj.a;
    if (a.acc) return 0;
    const int64_t n = d.n;
    a.passmax = std::min(std::max(c->max_iter_all, 1), kAdjPassMax);
    int rc;
    if ((rc = dalloc(&a.acc, (int64_t)kAdjAcc * n))) return rc;
    if ((rc = dalloc(&a.argmax, n))) return rc;
    if ((rc = dalloc(&a.par, (int64_t)SPHMHD_NTPARAMS * n))) return rc;
    if ((rc = dalloc(&a.dtrow, n))) return rc;
    if ((rc = dalloc(&a.ain, 5 * n))) return rc;
    c->adj.ain_live = a.ain;
    if ((rc = dalloc(&a.pass, (int64_t)5 * a.passmax * n))) return rc;
    if ((rc = dalloc(&a.npass, n))) return rc;
    if ((rc = dalloc(&a.bsave, 3 * n))) return rc;
    if ((rc = dalloc(&a.ai, d.ninst))) return rc;
    if ((rc = dalloc(&c->adj.final_clk, 2 * (int64_t)d.ninst))) return rc;
    return 0;
}

// one PEC step launch by launch: initialize, stage1, the evaluation, stage2,
// the clock (equal to a step of sphmhd_step bit for bit)
static int adj_plain_step(sphmhd_ctx* c) {
    Dev& d = c->d;
    launch_initialize(d);
    launch_stage1(d);
    int rc = do_eval(c);
    if (rc < 0) return rc;
    launch_stage2(d);
    launch_clock(d);
    HIPCHK(hipGetLastError());
    return 0;
}

// The warm start of sphmhd_warm_start up to a point, from the recorded
// initial state: upto 0 = the first sync, 1 = the first evaluation, 2 = the
// field restored and the second sync, 3 = the second evaluation, 4 = the
// third; the entry of the last evaluation run is saved for its reverse.
static int adj_warm_forward(sphmhd_ctx* c, int upto) {
    Dev& d = c->d;
    AdjDev& a = c->adj.a;
    double** f = d.fields.f;
    const int64_t n = d.n;
    launch_a_rows(d, rows_all(), c->adj.init, true);
    launch_a_clock(d, c->adj.init + (int64_t)SPHMHD_NFIELDS * n, true);
    launch_copy(d, d.scratch3, f[SPHMHD_F_BX], n);
    launch_copy(d, d.scratch3 + n, f[SPHMHD_F_BY], n);
    launch_copy(d, d.scratch3 + 2 * n, f[SPHMHD_F_BZ], n);
    launch_copy(d, f[SPHMHD_F_H0], f[SPHMHD_F_H], n);
    launch_sync_initial(d);
    int rc;
    for (int stage = 1; stage <= upto; ++stage) {
        if (stage == 2) {
            launch_copy(d, f[SPHMHD_F_BX], d.scratch3, n);
            launch_copy(d, f[SPHMHD_F_BY], d.scratch3 + n, n);
            launch_copy(d, f[SPHMHD_F_BZ], d.scratch3 + 2 * n, n);
            launch_copy(d, f[SPHMHD_F_H0], f[SPHMHD_F_H], n);
            launch_sync_initial(d);
            continue;
        }
        if (stage == upto) launch_a_save_entry(d, a);
        if ((rc = do_eval(c)) < 0) return rc;
    }
    HIPCHK(hipGetLastError());
    return 0;
}

static int adj_warm_reverse(sphmhd_ctx* c, const double* clk_next) {
    Dev& d = c->d;
    AdjDev& a = c->adj.a;
    int rc;
    // the third and second evaluations (the first dt is the third's pending minimum)
    for (int upto = 4; upto >= 3; --upto) {
        if ((rc = adj_warm_forward(c, upto))) return rc;
        launch_a_eval_begin(d, a, clk_next, upto == 4 ? 1 : 2);
        launch_a_eval(d, a);
    }
    // the field restored and the second sync
    if ((rc = adj_warm_forward(c, 2))) return rc;
    launch_a_sync(d, a);
    launch_a_bsave(d, a, true);
    // the first evaluation
    if ((rc = adj_warm_forward(c, 1))) return rc;
    launch_a_eval_begin(d, a, clk_next, 2);
    launch_a_eval(d, a);
    // the first sync, and the saved field
    if ((rc = adj_warm_forward(c, 0))) return rc;
    launch_a_sync(d, a);
    launch_a_bsave(d, a, false);
    HIPCHK(hipGetLastError());
    return 0;
}

extern "C" {

int sphmhd_abi_version(void) { return SPHMHD_ABI_VERSION; }

const char* sphmhd_last_error(void) { return g_err.c_str(); }

int sphmhd_create(const sphmhd_params* params, int32_t n_instances, const int64_t* offsets, void* hip_stream,
                  sphmhd_ctx** out) {
    try {
        if (!out) return fail(SPHMHD_EINVAL, "out is null");
        *out = nullptr;
        if (!params || !offsets || n_instances < 1) return fail(SPHMHD_EINVAL, "need at least one instance");
        if (offsets[0] != 0) return fail(SPHMHD_EINVAL, "offsets[0] must be 0");
        const sphmhd_params& p0 = params[0];
        if (p0.dim != 2 && p0.dim != 3) return fail(SPHMHD_EINVAL, "dim must be 2 or 3");
        if (p0.kernel < SPHMHD_CUBIC || p0.kernel > SPHMHD_SUPER_GAUSSIAN) return fail(SPHMHD_EINVAL, "unknown kernel");
        for (int k = 0; k < n_instances; ++k) {
            const sphmhd_params& p = params[k];
            if (offsets[k + 1] < offsets[k]) return fail(SPHMHD_EINVAL, "offsets must be ascending");
            if (p.dim != p0.dim || p.kernel != p0.kernel || p.periodic != p0.periodic)
                return fail(SPHMHD_EINVAL, "dim, kernel and periodic must agree across instances");
            if (!(p.gamma > 1.0)) return fail(SPHMHD_EINVAL, "gamma must exceed 1 (integrator.py:8)");
            if (!(p.mu0 > 0.0) || !(p.hfact > 0.0)) return fail(SPHMHD_EINVAL, "mu0 and hfact must be positive");
            if (p.max_iterations < 1) return fail(SPHMHD_EINVAL, "max_iterations must be >= 1");
            for (int a = 0; a < p.dim; ++a)
                if (!(p.hi[a] > p.lo[a])) return fail(SPHMHD_EINVAL, "domain bounds need hi > lo");
        }
        const int64_t n = offsets[n_instances];
        if (n < 1 || n > (int64_t)0x3FFFFFFF) return fail(SPHMHD_EINVAL, "particle count out of range");

        sphmhd_ctx* c = new sphmhd_ctx();
        c->offsets.assign(offsets, offsets + n_instances + 1);
        Dev& d = c->d;
        d.n = n;
        d.ninst = n_instances;
        d.dim = p0.dim;
        d.periodic = p0.periodic;
        d.kernel = p0.kernel;
        d.stream = (hipStream_t)hip_stream;
        if (const char* g = std::getenv("SPHMHD_GRAPHS")) c->use_graphs = std::strcmp(g, "0") != 0;
        for (int q = 0; q < kLaunchClasses; ++q) d.blk[q] = kBlock;
        if (const char* g = std::getenv("SPHMHD_SPEC")) c->spec_fixed = std::atoi(g);
        // SPHMHD_PRUNE_LISTS=1 / 0: always / never the pruned list kernel
        if (const char* g = std::getenv("SPHMHD_PRUNE_LISTS")) c->prune_mode = std::strcmp(g, "0") != 0;
        // SPHMHD_DENSITY_TRIM=0: the 3-D first density pass scans whole stencil rows (same results)
        if (const char* g = std::getenv("SPHMHD_DENSITY_TRIM")) d.dens_trim = std::strcmp(g, "0") != 0;
        // later density passes inside the first launch in 2-D, as launches of
        // their own in 3-D (there the fused passes gained nothing on the
        // blast); SPHMHD_DENSITY_FUSED=1 / 0 chooses either way
        d.dens_fused = p0.dim == 2;
        if (const char* g = std::getenv("SPHMHD_DENSITY_FUSED")) d.dens_fused = std::strcmp(g, "0") != 0;
        d.prune_lists = c->prune_mode == 1;
        if (const char* g = std::getenv("SPHMHD_TUNE"))
            if (std::strcmp(g, "0") == 0) c->tune.state = 3;
        // SPHMHD_PERSIST=1 / 0: every tuned launch persistent / plain (not tuned)
        if (const char* g = std::getenv("SPHMHD_PERSIST")) {
            c->persist_fixed = true;
            for (int q = 0; q < kLaunchClasses; ++q) d.persist[q] = std::strcmp(g, "0") != 0;
        }
        int64_t cap = 0;
        for (int k = 0; k < n_instances; ++k) {
            int64_t nk = offsets[k + 1] - offsets[k];
            cap += kCellsPerParticle * nk > 1 ? kCellsPerParticle * nk : 1;
        }
        d.cell_cap = cap;
        d.nb_stride = p0.dim == 2 ? 64 : 96;

        c->hinst.resize(n_instances);
        for (int k = 0; k < n_instances; ++k) {
            const sphmhd_params& p = params[k];
            Inst& s = c->hinst[k];
            std::memset(&s, 0, sizeof(Inst));
            s.off0 = offsets[k];
            s.off1 = offsets[k + 1];
            s.dim = p.dim;
            s.kernel = p.kernel;
            s.density_iterate = p.density_iterate;
            s.min_iter = p.min_iterations;
            s.max_iter = p.max_iterations;
            s.periodic = p.periodic;
            s.adaptive = p.adaptive_timestep;
            for (int a = 0; a < 3; ++a) {
                s.lo[a] = p.lo[a];
                s.hi[a] = p.hi[a];
                s.L[a] = a < p.dim ? p.hi[a] - p.lo[a] : 0.0;
                s.halfL[a] = 0.5 * s.L[a];
            }
            s.gamma = p.gamma;
            s.gamma1 = p.gamma - 1.0;
            s.mu0 = p.mu0;
            s.hfact = p.hfact;
            s.av = p.artificial_viscosity;
            s.ac = p.artificial_thermal_conductivity;
            s.ar = p.artificial_magnetic_dissipation;
            s.chf = p.cleaning_speed_factor;
            s.sigma = p.cleaning_damping_factor;
            s.tsf = p.timestep_factor;
            s.alpha_min = p.alpha_min;
            s.alpha_max = p.alpha_max;
            s.tf = p.tf;
            s.cleaning_active = p.cleaning_speed_factor > 0.0 ? 1.0 : 0.0;
            s.c_cour = p.timestep_factor * 0.3;
            s.c_force = p.timestep_factor * 0.25;
            s.dtp_bits = 0x7FF0000000000000ull;
            c->max_iter_all = std::max(c->max_iter_all, p.density_iterate ? p.max_iterations : 1);
            c->min_iter_all = std::max(c->min_iter_all, p.density_iterate ? p.min_iterations : 1);
        }

        std::vector<int32_t> inst_of(n);
        for (int k = 0; k < n_instances; ++k)
            for (int64_t i = offsets[k]; i < offsets[k + 1]; ++i) inst_of[i] = k;

        int rc = 0;
        auto cleanup = [&](int code) {
            sphmhd_destroy(c);
            return code;
        };
        if ((rc = dalloc(&d.inst, n_instances))) return cleanup(rc);
        if ((rc = dalloc(&d.ctl, 1))) return cleanup(rc);
        if ((rc = dalloc(&d.arrive, kArriveInts))) return cleanup(rc);
        if ((rc = dalloc(&d.queue, kQueueInts))) return cleanup(rc);
        if ((rc = dalloc(&d.hblk, kRed * ((n + kBlock - 1) / kBlock)))) return cleanup(rc);
        if ((rc = dalloc(&d.kblk, (n + kBlock - 1) / kBlock))) return cleanup(rc);
        if ((rc = dalloc(&d.inst_of, n))) return cleanup(rc);
        if ((rc = dalloc(&d.keys, n))) return cleanup(rc);
        if ((rc = dalloc(&d.keys_alt, n))) return cleanup(rc);
        if ((rc = dalloc(&d.perm, n))) return cleanup(rc);
        if ((rc = dalloc(&d.perm_alt, n))) return cleanup(rc);
        if ((rc = dalloc(&d.ccount, cap))) return cleanup(rc);
        if ((rc = dalloc(&d.cblock, (cap + kScanCells - 1) / kScanCells))) return cleanup(rc);
        if ((rc = dalloc(&d.bigc, n / (kSmallBucket + 1) + 1))) return cleanup(rc);
        if ((rc = dalloc(&d.cstart, cap))) return cleanup(rc);
        if ((rc = dalloc(&d.cend, cap))) return cleanup(rc);
        if ((rc = dalloc(&d.hcell, 4 * cap))) return cleanup(rc);
        // one 3-D instance: the h histogram its cell size is chosen from
        if (n_instances == 1 && p0.dim == 3) {
            if ((rc = dalloc(&d.hist, (int64_t)kHistSlots * kHistCopies))) return cleanup(rc);
            hipError_t eh = hipMemcpy(&d.ctl->hist, &d.hist, sizeof(d.hist), hipMemcpyHostToDevice);
            if (eh != hipSuccess) return cleanup(fail(SPHMHD_EHIP, hipGetErrorString(eh)));
        }
        if ((rc = dalloc(&d.cls, cap))) return cleanup(rc);
        if ((rc = dalloc(&d.cle, cap))) return cleanup(rc);
        if ((rc = dalloc(&d.dpack, 4 * n))) return cleanup(rc);
        if ((rc = dalloc(&d.fpack, (int64_t)kFRecMax * n))) return cleanup(rc);
        if ((rc = dalloc(&d.nb, (int64_t)d.nb_stride * n))) return cleanup(rc);
        if ((rc = dalloc(&d.nbcnt, n))) return cleanup(rc);
        if ((rc = dalloc(&d.act[0], n))) return cleanup(rc);
        if ((rc = dalloc(&d.act[1], n))) return cleanup(rc);
        if ((rc = dalloc(&d.scratch3, 3 * n))) return cleanup(rc);
        if ((rc = dalloc(&d.aprev, 3 * n))) return cleanup(rc);
        if ((rc = dalloc(&d.totals, (int64_t)n_instances * SPHMHD_NTOTALS))) return cleanup(rc);
        // this device's copy of the allocation-free totals' static scratch
        if ((rc = scratch_setup(&d.tot_scratch))) return cleanup(rc);
        // the control block's host mirror: pinned, mapped and coherent, so the
        // kernel that ends an evaluation writes it directly (no copy)
        hipError_t e = hipHostMalloc((void**)&c->hctl, 3 * sizeof(CtlView), hipHostMallocMapped | hipHostMallocCoherent);
        if (e != hipSuccess) return cleanup(fail(SPHMHD_EHIP, "hipHostMalloc failed"));
        e = hipHostGetDevicePointer((void**)&d.hctl_dev, c->hctl, 0);
        if (e != hipSuccess) return cleanup(fail(SPHMHD_EHIP, "hipHostGetDevicePointer failed"));
        std::memset(c->hctl, 0, 3 * sizeof(CtlView));
        c->hcur = &c->hctl[2];
        e = hipMemcpy(d.inst, c->hinst.data(), sizeof(Inst) * n_instances, hipMemcpyHostToDevice);
        if (e != hipSuccess) return cleanup(fail(SPHMHD_EHIP, hipGetErrorString(e)));
        e = hipMemcpy(d.inst_of, inst_of.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess) return cleanup(fail(SPHMHD_EHIP, hipGetErrorString(e)));
        // the cell sort visits particles in the previous sort's order: start
        // from the storage order
        for (int64_t i = 0; i < n; ++i) inst_of[i] = (int32_t)i;
        e = hipMemcpy(d.perm, inst_of.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess) return cleanup(fail(SPHMHD_EHIP, hipGetErrorString(e)));
        *out = c;
        return 0;
    } catch (const std::bad_alloc&) {
        return fail(SPHMHD_ECAP, "host allocation failed");
    } catch (...) {
        return fail(SPHMHD_EINVAL, "unexpected exception");
    }
}

int sphmhd_destroy(sphmhd_ctx* c) {
    if (!c) return 0;
    Dev& d = c->d;
    (void)hipStreamSynchronize(d.stream);
    drop_graphs(c);
    for (void* p : {(void*)c->tg.itan, (void*)c->tg.aprev, (void*)c->tg.bsave, (void*)c->tg.totals})
        if (p) (void)hipFree(p);
    adj_free(c);
    if (c->gstream) (void)hipStreamDestroy(c->gstream);
    if (c->gev) (void)hipEventDestroy(c->gev);
    if (c->sev) (void)hipEventDestroy(c->sev);
    void* ptrs[] = {d.inst, d.ctl, d.arrive, d.queue, d.hblk, d.kblk, d.inst_of, d.keys, d.keys_alt, d.perm, d.perm_alt, d.ccount, d.cblock, d.bigc,
                    d.cstart, d.cend, d.hcell, d.cls, d.cle, d.dpack, d.fpack, d.nb, d.nbcnt, d.scratch3, d.aprev, d.totals, d.act[0],
                    d.act[1], d.hist};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->hctl) (void)hipHostFree(c->hctl);
    for (PhaseTimer& t : c->tm) {
        if (t.a) (void)hipEventDestroy(t.a);
        if (t.b) (void)hipEventDestroy(t.b);
    }
    delete c;
    return 0;
}

int sphmhd_bind(sphmhd_ctx* c, double* const* fields, int32_t* itype) {
    if (!c || !fields || !itype) return fail(SPHMHD_EINVAL, "null argument");
    for (int f = 0; f < SPHMHD_NFIELDS; ++f) {
        bool optional = c->d.dim == 2 && (f == SPHMHD_F_Z || f == SPHMHD_F_Z0);
        if (!fields[f] && !optional) return fail(SPHMHD_EINVAL, "field " + std::to_string(f) + " is null");
        c->d.fields.f[f] = fields[f];
    }
    c->d.fields.itype = itype;
    c->bound = true;
    c->graph_gen += 1;
    return 0;
}

int sphmhd_bind_strain(sphmhd_ctx* c, double* const* dvdx) {
    if (!c) return fail(SPHMHD_EINVAL, "null context");
    for (int q = 0; q < 9; ++q) c->d.strain.d[q] = dvdx ? dvdx[q] : nullptr;
    if (dvdx)
        for (int q = 0; q < 9; ++q)
            if (!dvdx[q]) {
                for (int r = 0; r < 9; ++r) c->d.strain.d[r] = nullptr;
                return fail(SPHMHD_EINVAL, "strain row " + std::to_string(q) + " is null");
            }
    c->graph_gen += 1;  // captured steps hold the force pass's arguments
    return 0;
}

int sphmhd_eval(sphmhd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    try {
        return do_eval(c);
    } catch (...) {
        return fail(SPHMHD_EINVAL, "unexpected exception");
    }
}

int sphmhd_pec_initialize(sphmhd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    launch_initialize(c->d);
    HIPCHK(hipGetLastError());
    return 0;
}

int sphmhd_pec_stage1(sphmhd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    launch_stage1(c->d);
    HIPCHK(hipGetLastError());
    return 0;
}

int sphmhd_pec_stage2(sphmhd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    launch_stage2(c->d);
    HIPCHK(hipGetLastError());
    return 0;
}

int sphmhd_advance_clock(sphmhd_ctx* c) {
    if (!c) return fail(SPHMHD_EINVAL, "null context");
    launch_clock(c->d);
    HIPCHK(hipGetLastError());
    return 0;
}

// runs the steps of one sphmhd_step call on the library's stream when graphs
// are in use, and restores the caller's stream afterwards
struct StreamSwap {
    sphmhd_ctx* c;
    hipStream_t user;
    bool on = false;
    ~StreamSwap() {
        if (on) {
            (void)hipStreamSynchronize(c->d.stream);
            c->d.stream = user;
        }
    }
};

// the force pass of the evaluation just finished again, from the kept
// previous accelerations (same results, and the rows of the current lean
// setting written)
static int force_redo(sphmhd_ctx* c) {
    Dev& d = c->d;
    launch_clear_abort(d);
    launch_reset_nb_stats(d);
    launch_reset_force_outputs(d);
    int rc = force_pass(c, false);
    if (rc) return rc;
    if ((rc = read_ctl(c))) return rc;
    if (c->hcur->abort) launch_clear_abort(d);
    c->stats.force_redos += 1;
    return 0;
}

// the host's copy of the device's lean flag; set() launches only on change
// and returns the value set; clears the flag when the step call ends
struct LeanGuard {
    sphmhd_ctx* c;
    bool set(bool v) {
        if (v != c->lean_now) {
            launch_set_lean(c->d, v ? 1 : 0);
            c->lean_now = v;
        }
        return v;
    }
    ~LeanGuard() { set(false); }
};

int sphmhd_step(sphmhd_ctx* c, int64_t nsteps) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (nsteps < 0) return fail(SPHMHD_EINVAL, "nsteps < 0");
    if (nsteps == 0) return 0;
    int worst = 0;
    try {
        StreamSwap swap{c, c->d.stream};
        if (c->use_graphs && nsteps > 1) {
            if (!c->gstream) {
                HIPCHK(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
                HIPCHK(hipEventCreateWithFlags(&c->gev, hipEventDisableTiming));
            }
            HIPCHK(hipEventRecord(c->gev, c->d.stream));
            HIPCHK(hipStreamWaitEvent(c->gstream, c->gev, 0));
            c->d.stream = c->gstream;
            swap.on = true;
        }
        // With graphs (and no timing or tuning) step s + 1 is enqueued before
        // the host waits for step s, so the GPU does not idle through the
        // host's turn.  If step s needs more density passes or longer lists,
        // its end sets the abort flag and every kernel of step s + 1 exits at
        // entry: the host then finishes step s and enqueues step s + 1 again.
        // Every step but the last of the call runs lean: its force pass skips
        // the diagnostic rows no later kernel reads, so the fields at return
        // are those of plain single steps.  The flag is set before a step's
        // launches are enqueued and cleared on every way out.
        LeanGuard lean{c};
        bool ahead = false;  // step s was enqueued during step s - 1
        int ahead_spec = 0;
        bool step_lean = false, ahead_lean = false;  // the flag step s and step s + 1 were enqueued with
        for (int64_t s = 0; s < nsteps; ++s) {
            // tuning steps are launched one by one with events around the phases
            const bool tuning = c->tune.state <= 1;
            if (tuning && (rc = tune_begin_step(c))) return rc;
            int spec = ahead_spec;
            step_lean = ahead_lean;
            if (!ahead) {
                spec = next_spec(c);
                step_lean = lean.set(s + 1 < nsteps);
                rc = (!tuning && s > 0 && swap.on && c->use_graphs) ? step_front_graph(c, spec)
                                                                     : step_front(c, s == 0, spec);
                if (rc < 0) return rc;
            }
            ahead = false;
            if (!tuning && swap.on && c->use_graphs && !c->timing && s + 1 < nsteps) {
                if (!c->sev) HIPCHK(hipEventCreateWithFlags(&c->sev, hipEventDisableTiming));
                HIPCHK(hipEventRecord(c->sev, c->d.stream));  // end of step s
                ahead_spec = next_spec(c);
                const int64_t replays = c->stats.graph_replays;
                ahead_lean = lean.set(s + 2 < nsteps);
                rc = step_front_graph(c, ahead_spec);  // replayed, or launched plainly if a capture failed
                if (rc < 0) return rc;
                ahead = true;
                HIPCHK(hipEventSynchronize(c->sev));
                if ((rc = pick_record(c, c->seq_next - 2))) return rc;
                if (c->hcur->abort) {
                    // step s + 1 exited at entry: let it drain, take back its count
                    HIPCHK(hipStreamSynchronize(c->d.stream));
                    c->seq_next -= 1;
                    c->stats.graph_replays = replays;
                    ahead = false;
                }
                rc = eval_process(c, spec);
            } else {
                rc = eval_finish(c, spec);
            }
            if (rc < 0) return rc;
            if (rc > worst) worst = rc;
            if (tuning) tune_end_step(c);
            c->stats.steps += 1;
            // this step brings every clock to tf: the run ends with it (the
            // step enqueued after it, if any, exited at entry); a step that
            // ran lean repeats its force pass to leave the diagnostic rows
            if (c->hcur->at_tf) {
                if (step_lean) {
                    lean.set(false);
                    if ((rc = force_redo(c)) < 0) return rc;
                }
                break;
            }
        }
        launch_stage2_clock(c->d);  // stage2 of the last step and the clock commit
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->d.stream));
    } catch (...) {
        return fail(SPHMHD_EINVAL, "unexpected exception");
    }
    return worst;
}

int sphmhd_sync_initial_state(sphmhd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    launch_sync_initial(c->d);
    HIPCHK(hipGetLastError());
    return 0;
}

int sphmhd_warm_start(sphmhd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    Dev& d = c->d;
    double** f = d.fields.f;
    const int64_t n = d.n;
    launch_copy(d, d.scratch3, f[SPHMHD_F_BX], n);
    launch_copy(d, d.scratch3 + n, f[SPHMHD_F_BY], n);
    launch_copy(d, d.scratch3 + 2 * n, f[SPHMHD_F_BZ], n);
    launch_copy(d, f[SPHMHD_F_H0], f[SPHMHD_F_H], n);
    launch_sync_initial(d);
    rc = do_eval(c);
    if (rc < 0) return rc;
    int64_t bad = 0;
    if ((rc = count_unconverged(c, &bad))) return rc;
    if (bad) return fail(SPHMHD_ENOTCONV, "warm start: " + std::to_string(bad) + " particles did not converge");
    launch_copy(d, f[SPHMHD_F_BX], d.scratch3, n);
    launch_copy(d, f[SPHMHD_F_BY], d.scratch3 + n, n);
    launch_copy(d, f[SPHMHD_F_BZ], d.scratch3 + 2 * n, n);
    launch_copy(d, f[SPHMHD_F_H0], f[SPHMHD_F_H], n);
    launch_sync_initial(d);
    for (int r = 0; r < 2; ++r) {
        rc = do_eval(c);
        if (rc < 0) return rc;
    }
    if ((rc = count_unconverged(c, &bad))) return rc;
    if (bad) return fail(SPHMHD_ENOTCONV, "warm start: " + std::to_string(bad) + " particles did not converge");
    launch_warm_dt(d);
    HIPCHK(hipGetLastError());
    if ((rc = read_ctl(c))) return rc;
    if (c->hcur->status & 4) return fail(SPHMHD_EINVAL, "warm start: dt_cfl <= 0 (application.py:139)");
    return 0;
}

int sphmhd_get_clock(sphmhd_ctx* c, double* t, double* dt) {
    if (!c || !t || !dt) return fail(SPHMHD_EINVAL, "null argument");
    int rc = pull_inst(c);
    if (rc) return rc;
    for (size_t k = 0; k < c->hinst.size(); ++k) {
        t[k] = c->hinst[k].t;
        dt[k] = c->hinst[k].dt;
    }
    return 0;
}

int sphmhd_set_clock(sphmhd_ctx* c, const double* t, const double* dt) {
    if (!c || !t || !dt) return fail(SPHMHD_EINVAL, "null argument");
    int rc = pull_inst(c);
    if (rc) return rc;
    for (size_t k = 0; k < c->hinst.size(); ++k) {
        c->hinst[k].t = t[k];
        c->hinst[k].dt = dt[k];
    }
    return push_inst(c);
}

int sphmhd_get_pending_dt(sphmhd_ctx* c, double* dt) {
    if (!c || !dt) return fail(SPHMHD_EINVAL, "null argument");
    int rc = pull_inst(c);
    if (rc) return rc;
    for (size_t k = 0; k < c->hinst.size(); ++k) {
        unsigned long long b = c->hinst[k].dtp_bits;
        std::memcpy(&dt[k], &b, sizeof(double));
    }
    return 0;
}

int sphmhd_totals(sphmhd_ctx* c, double* out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out) return fail(SPHMHD_EINVAL, "null argument");
    launch_totals(c->d);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, c->d.totals, sizeof(double) * c->d.ninst * SPHMHD_NTOTALS, hipMemcpyDeviceToHost,
                          c->d.stream));
    HIPCHK(hipStreamSynchronize(c->d.stream));
    return 0;
}

int sphmhd_cells(sphmhd_ctx* c, uint32_t* keys, int32_t* perm) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!keys || !perm) return fail(SPHMHD_EINVAL, "null argument");
    if ((rc = nnps_rebuild(c, false))) return rc;
    HIPCHK(hipMemcpyAsync(keys, c->d.keys, sizeof(uint32_t) * c->d.n, hipMemcpyDeviceToHost, c->d.stream));
    HIPCHK(hipMemcpyAsync(perm, c->d.perm, sizeof(int32_t) * c->d.n, hipMemcpyDeviceToHost, c->d.stream));
    HIPCHK(hipStreamSynchronize(c->d.stream));
    return 0;
}

int sphmhd_neighbours(sphmhd_ctx* c, int64_t* offs, int32_t* idx, int64_t cap, int64_t* total) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!offs || !total || (cap > 0 && !idx)) return fail(SPHMHD_EINVAL, "null argument");
    try {
        if ((rc = nnps_rebuild(c, false))) return rc;
        if ((rc = read_ctl(c))) return rc;  // this grid picks the list kernel
        launch_pack_xh(c->d);
        if ((rc = build_lists(c))) return rc;
        const int64_t n = c->d.n;
        const int stride = c->d.nb_stride;
        std::vector<int32_t> nb((size_t)stride * n), cnt(n), perm(n);
        HIPCHK(hipMemcpyAsync(nb.data(), c->d.nb, sizeof(int32_t) * nb.size(), hipMemcpyDeviceToHost, c->d.stream));
        HIPCHK(hipMemcpyAsync(cnt.data(), c->d.nbcnt, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->d.stream));
        HIPCHK(hipMemcpyAsync(perm.data(), c->d.perm, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->d.stream));
        HIPCHK(hipStreamSynchronize(c->d.stream));
        std::vector<std::vector<int32_t>> lists(n);
        for (int64_t t = 0; t < n; ++t) {
            std::vector<int32_t>& l = lists[perm[t]];
            l.resize(cnt[t]);
            for (int q = 0; q < cnt[t]; ++q) l[q] = perm[nb[(size_t)q * n + t]];
            std::sort(l.begin(), l.end());
        }
        int64_t pos = 0;
        for (int64_t i = 0; i < n; ++i) {
            offs[i] = pos;
            for (int32_t j : lists[i]) {
                if (pos < cap) idx[pos] = j;
                ++pos;
            }
        }
        offs[n] = pos;
        *total = pos;
        return pos > cap ? fail(SPHMHD_ECAP, "idx buffer too small") : 0;
    } catch (const std::bad_alloc&) {
        return fail(SPHMHD_ECAP, "host allocation failed");
    }
}

int sphmhd_get_stats(sphmhd_ctx* c, sphmhd_stats* out) {
    if (!c || !out) return fail(SPHMHD_EINVAL, "null argument");
    *out = c->stats;
    out->block_density = c->d.blk[LC_DENSITY];
    out->block_lists = c->d.blk[LC_LISTS];
    out->block_force = c->d.blk[LC_FORCE];
    out->launch_tuned = c->tune.state == 2 ? 1 : 0;
    out->persist_density = c->d.persist[LC_DENSITY];
    out->persist_lists = c->d.persist[LC_LISTS];
    out->persist_force = c->d.persist[LC_FORCE];
    return 0;
}

int sphmhd_set_persistent(sphmhd_ctx* c, const int32_t* modes) {
    if (!c) return fail(SPHMHD_EINVAL, "null context");
    if (c->tune.state == 1) return fail(SPHMHD_ESTATE, "launch shapes are being tuned");
    if (modes) {
        for (int q = 0; q < kLaunchClasses; ++q)
            if (modes[q] != 0 && modes[q] != 1) return fail(SPHMHD_EINVAL, "launch modes must be 0 or 1");
        for (int q = 0; q < kLaunchClasses; ++q) c->d.persist[q] = modes[q];
        c->persist_fixed = true;
    } else {
        c->persist_fixed = false;
        c->tune.state = 0;  // tuned again with the workgroup sizes
    }
    c->graph_gen += 1;
    return 0;
}

int sphmhd_set_launch(sphmhd_ctx* c, const int32_t* blocks) {
    if (!c) return fail(SPHMHD_EINVAL, "null context");
    if (c->tune.state == 1) return fail(SPHMHD_ESTATE, "launch shapes are being tuned");
    if (blocks) {
        for (int q = 0; q < kLaunchClasses; ++q)
            if (blocks[q] != 64 && blocks[q] != 128 && blocks[q] != 256)
                return fail(SPHMHD_EINVAL, "workgroup sizes must be 64, 128 or 256");
        for (int q = 0; q < kLaunchClasses; ++q) c->d.blk[q] = blocks[q];
        c->tune.state = 3;
    } else {
        c->tune.state = 0;
    }
    c->graph_gen += 1;
    return 0;
}

int sphmhd_get_launch(sphmhd_ctx* c, int32_t* blocks, int32_t* tuned) {
    if (!c || !blocks) return fail(SPHMHD_EINVAL, "null argument");
    for (int q = 0; q < kLaunchClasses; ++q) blocks[q] = c->d.blk[q];
    if (tuned) *tuned = c->tune.state == 2 ? 1 : 0;
    return 0;
}

// ---------------------------------------------------------------------------
// tangent-linear step
// ---------------------------------------------------------------------------

static int t_check(sphmhd_ctx* c) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!c->tbound) return fail(SPHMHD_ESTATE, "no tangent rows bound");
    return 0;
}

// one evaluation over values and tangents: the primal neighbour search and
// list build on the values, dual density passes (one wait each), dual prep,
// the list (grown if needed), the dual force pass and the dt tangent
// the tangent prep, lists and force pass (with the dt tangent and the
// evaluation's end); keep_prev: a redo, which reads the first prep's copy of
// the previous accelerations
static int t_force_pass(sphmhd_ctx* c, bool keep_prev) {
    Dev& d = c->d;
    TanDev& g = c->tg;
    tbegin(c, T_LISTS);
    launch_t_prep(d, g, keep_prev);
    launch_reset_nb_stats(d);
    launch_nb_build(d);
    tend(c, T_LISTS);
    tbegin(c, T_FORCE);
    launch_t_force(d, g);
    tend(c, T_FORCE);
    launch_eval_end(d);
    c->seq_next += 1;
    HIPCHK(hipGetLastError());
    return 0;
}

// One wait per evaluation, as the plain path: speculative density passes
// and a speculative force pass, then the remaining passes and a redone force
// pass when the iteration needed more passes or a list had to grow.
static int t_eval(sphmhd_ctx* c) {
    Dev& d = c->d;
    TanDev& g = c->tg;
    int rc = nnps_rebuild(c, true);
    if (rc) return rc;
    launch_t_eval_begin(d, g);
    launch_pack_density(d);
    const int spec = next_spec(c, true);
    int passes = 0;
    tbegin(c, T_DENSITY);
    for (; passes < spec; ++passes) {
        d.pass = passes;
        launch_t_density(d, g);
        launch_g0_control(d);
    }
    tend(c, T_DENSITY);
    if ((rc = t_force_pass(c, false))) return rc;
    if ((rc = read_ctl(c))) return rc;
    bool redo = false;
    if (c->hcur->abort) launch_clear_abort(d);
    if (!c->hcur->all_done) {
        launch_pack_density(d);  // the prep overwrote the density pack with {x,y,z,h}
        while (!c->hcur->all_done) {
            if (passes > c->max_iter_all + 1) return fail(SPHMHD_ESTATE, "density iteration control did not stop");
            d.pass = passes++;
            launch_t_density(d, g);
            launch_g0_control(d);
            HIPCHK(hipGetLastError());
            if ((rc = read_ctl(c))) return rc;
        }
        redo = true;
        c->spec_margin = 1;
        c->evals_since_redo = 0;
    } else if (c->spec_margin && ++c->evals_since_redo >= 16) {
        c->spec_margin = 0;
    }
    c->spec_passes = std::max(c->hcur->iters_max, c->min_iter_all);
    c->stats.density_iterations += c->hcur->iters_max;
    c->stats.last_iterations_max = c->hcur->iters_max;
    for (int grows = 0;;) {
        if (c->hcur->overflow) {
            if (++grows > 2) return fail(SPHMHD_ECAP, "neighbour list did not fit after growing");
            if ((rc = grow_nb(c, c->hcur->max_nb))) return rc;
            redo = true;
        }
        if (!redo) break;
        launch_clear_abort(d);
        launch_reset_force_outputs(d);
        launch_t_eval_begin(d, g);
        if ((rc = t_force_pass(c, true))) return rc;
        if ((rc = read_ctl(c))) return rc;
        redo = false;
        c->stats.force_redos += 1;
    }
    if (c->hcur->abort) launch_clear_abort(d);  // a redone pass of the step that reaches tf
    c->stats.max_neighbours = c->hcur->max_nb;
    c->stats.pairs_force = c->hcur->pairs;
    c->stats.cells = c->hcur->total_cells;
    c->stats.evals += 1;
    const int st = c->hcur->status;
    if (st & 2) return SPHMHD_WNAN;
    if (st & 1) return SPHMHD_WNOCONV;
    return 0;
}

int sphmhd_tangent_bind(sphmhd_ctx* c, double* const* tangent_fields) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!tangent_fields) return fail(SPHMHD_EINVAL, "null argument");
    Dev& d = c->d;
    TanDev& g = c->tg;
    for (int q = 0; q < kTanDirs; ++q)
        for (int f = 0; f < SPHMHD_NFIELDS; ++f) {
            double* p = tangent_fields[q * SPHMHD_NFIELDS + f];
            if (!p && d.fields.f[f]) return fail(SPHMHD_EINVAL, "tangent row " + std::to_string(f) + " is null");
            g.tf.f[q][f] = p;
        }
    if (!g.itan) {
        if ((rc = dalloc(&g.itan, d.ninst))) return rc;
        if ((rc = dalloc(&g.aprev, (int64_t)3 * (1 + kTanDirs) * d.n))) return rc;
        if ((rc = dalloc(&g.bsave, (int64_t)3 * (1 + kTanDirs) * d.n))) return rc;
        if ((rc = dalloc(&g.totals, (int64_t)d.ninst * SPHMHD_NTOTALS * (1 + kTanDirs)))) return rc;
    }
    c->tbound = true;
    return 0;
}

int sphmhd_tangent_seed(sphmhd_ctx* c, const double* param_tangents) {
    int rc = t_check(c);
    if (rc) return rc;
    if (!param_tangents) return fail(SPHMHD_EINVAL, "null argument");
    std::vector<InstTan> h(c->d.ninst);
    for (int k = 0; k < c->d.ninst; ++k) {
        InstTan& it = h[k];
        std::memset(&it, 0, sizeof(it));
        it.argmin = INT32_MAX;
        for (int p = 0; p < SPHMHD_NTPARAMS; ++p)
            for (int q = 0; q < kTanDirs; ++q)
                it.par[p][q] = param_tangents[((int64_t)k * SPHMHD_NTPARAMS + p) * kTanDirs + q];
    }
    HIPCHK(hipMemcpyAsync(c->tg.itan, h.data(), sizeof(InstTan) * h.size(), hipMemcpyHostToDevice, c->d.stream));
    HIPCHK(hipStreamSynchronize(c->d.stream));
    return 0;
}

int sphmhd_tangent_warm_start(sphmhd_ctx* c) {
    int rc = t_check(c);
    if (rc) return rc;
    Dev& d = c->d;
    TanDev& g = c->tg;
    const int64_t blk = (int64_t)(1 + kTanDirs) * d.n;
    const int bf[3] = {SPHMHD_F_BX, SPHMHD_F_BY, SPHMHD_F_BZ};
    for (int q = 0; q < 3; ++q) launch_t_save(d, g, bf[q], g.bsave + q * blk, false);
    launch_t_copy(d, g, SPHMHD_F_H0, SPHMHD_F_H);
    launch_t_sync_initial(d, g);
    rc = t_eval(c);
    if (rc < 0) return rc;
    int64_t bad = 0;
    if ((rc = count_unconverged(c, &bad))) return rc;
    if (bad) return fail(SPHMHD_ENOTCONV, "warm start: " + std::to_string(bad) + " particles did not converge");
    for (int q = 0; q < 3; ++q) launch_t_save(d, g, bf[q], g.bsave + q * blk, true);
    launch_t_copy(d, g, SPHMHD_F_H0, SPHMHD_F_H);
    launch_t_sync_initial(d, g);
    for (int r = 0; r < 2; ++r) {
        rc = t_eval(c);
        if (rc < 0) return rc;
    }
    if ((rc = count_unconverged(c, &bad))) return rc;
    if (bad) return fail(SPHMHD_ENOTCONV, "warm start: " + std::to_string(bad) + " particles did not converge");
    launch_warm_dt(d);
    launch_t_warm_dt(d, g);
    HIPCHK(hipGetLastError());
    if ((rc = read_ctl(c))) return rc;
    if (c->hcur->status & 4) return fail(SPHMHD_EINVAL, "warm start: dt_cfl <= 0 (application.py:139)");
    return 0;
}

int sphmhd_tangent_step(sphmhd_ctx* c, int64_t nsteps) {
    int rc = t_check(c);
    if (rc) return rc;
    if (nsteps < 0) return fail(SPHMHD_EINVAL, "nsteps < 0");
    int worst = 0;
    for (int64_t s = 0; s < nsteps; ++s) {
        // a context that only takes tangent steps tunes its launch shapes here
        const bool tuning = c->tune.state <= 1;
        if (tuning && (rc = tune_begin_step(c))) return rc;
        launch_t_initialize(c->d, c->tg);
        launch_t_stage1(c->d, c->tg);
        rc = t_eval(c);
        if (rc < 0) return rc;
        if (rc > worst) worst = rc;
        if (tuning) tune_end_step(c);
        launch_t_stage2(c->d, c->tg);
        launch_t_clock(c->d, c->tg);
        HIPCHK(hipGetLastError());
        c->stats.steps += 1;
        if (c->hcur->at_tf) break;  // the clocks reached tf
    }
    HIPCHK(hipStreamSynchronize(c->d.stream));
    return worst;
}

int sphmhd_tangent_totals(sphmhd_ctx* c, double* out) {
    int rc = t_check(c);
    if (rc) return rc;
    if (!out) return fail(SPHMHD_EINVAL, "null argument");
    launch_t_totals(c->d, c->tg, c->tg.totals);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, c->tg.totals, sizeof(double) * c->d.ninst * SPHMHD_NTOTALS * (1 + kTanDirs),
                          hipMemcpyDeviceToHost, c->d.stream));
    HIPCHK(hipStreamSynchronize(c->d.stream));
    return 0;
}

int sphmhd_tangent_get_dt(sphmhd_ctx* c, double* dt_tangent) {
    int rc = t_check(c);
    if (rc) return rc;
    if (!dt_tangent) return fail(SPHMHD_EINVAL, "null argument");
    std::vector<InstTan> h(c->d.ninst);
    HIPCHK(hipMemcpyAsync(h.data(), c->tg.itan, sizeof(InstTan) * h.size(), hipMemcpyDeviceToHost, c->d.stream));
    HIPCHK(hipStreamSynchronize(c->d.stream));
    for (int k = 0; k < c->d.ninst; ++k)
        for (int q = 0; q < kTanDirs; ++q) dt_tangent[(int64_t)k * kTanDirs + q] = h[k].dt[q];
    return 0;
}

// ---------------------------------------------------------------------------
// adjoint (reverse-mode) step
// ---------------------------------------------------------------------------

int sphmhd_adjoint_bind(sphmhd_ctx* c, double* const* bar_fields) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!bar_fields) return fail(SPHMHD_EINVAL, "null argument");
    for (int f = 0; f < SPHMHD_NFIELDS; ++f)
        if (!bar_fields[f] && c->d.fields.f[f])
            return fail(SPHMHD_EINVAL, "adjoint row " + std::to_string(f) + " is null");
    for (int f = 0; f < SPHMHD_NFIELDS; ++f) c->adj.a.b[f] = c->d.fields.f[f] ? bar_fields[f] : nullptr;
    c->adj.bound = true;
    return 0;
}

int sphmhd_adjoint_record(sphmhd_ctx* c, int32_t warm, int64_t nsteps, int32_t every) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (nsteps < 0 || every < 1) return fail(SPHMHD_EINVAL, "nsteps < 0 or every < 1");
    Dev& d = c->d;
    sphmhd_ctx::Adj& A = c->adj;
    const int64_t n = d.n;
    const int64_t sn = (int64_t)kNState * n + 2 * d.ninst;
    if ((rc = adj_alloc_work(c))) return rc;
    for (double* p : A.ckpt)
        if (p) (void)hipFree(p);
    A.ckpt.clear();
    A.recorded = false;
    if (!A.init && (rc = dalloc(&A.init, (int64_t)SPHMHD_NFIELDS * n + 2 * d.ninst))) return rc;
    if (A.seg_cap < every) {
        if (A.seg) (void)hipFree(A.seg);
        if (A.clk) (void)hipFree(A.clk);
        A.seg = A.clk = nullptr;
        if ((rc = dalloc(&A.seg, (int64_t)every * kNState * n))) return rc;
        if ((rc = dalloc(&A.clk, (int64_t)(every + 1) * 2 * d.ninst))) return rc;
        A.seg_cap = every;
    }
    while ((int)A.snaps.size() < every) {
        sphmhd_ctx::Adj::Snap sn;
        if ((rc = dalloc(&sn.rows, (int64_t)(kSnapRows + 5) * n))) return rc;
        if ((rc = dalloc(&sn.fpack, (int64_t)kFRecMax * n))) return rc;
        if ((rc = dalloc(&sn.ints, 3 * n))) return rc;
        if ((rc = dalloc(&sn.inst, d.ninst))) return rc;
        A.snaps.push_back(sn);
    }
    launch_a_rows(d, rows_all(), A.init, false);
    launch_a_clock(d, A.init + (int64_t)SPHMHD_NFIELDS * n, false);
    int worst = 0;
    if (warm) {
        rc = sphmhd_warm_start(c);
        if (rc < 0) return rc;
        worst = std::max(worst, rc);
    }
    int64_t taken = 0;
    while (taken < nsteps) {
        double* ck = nullptr;
        if ((rc = dalloc(&ck, sn))) return rc;
        A.ckpt.push_back(ck);
        launch_a_rows(d, rows_state(), ck, false);
        launch_a_clock(d, ck + (int64_t)kNState * n, false);
        const int64_t len = std::min<int64_t>(every, nsteps - taken);
        const int64_t before = c->stats.steps;
        rc = sphmhd_step(c, len);
        if (rc < 0) return rc;
        worst = std::max(worst, rc);
        const int64_t got = c->stats.steps - before;
        taken += got;
        if (got < len) break;  // the clocks reached tf
    }
    launch_a_clock(d, A.final_clk, false);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(d.stream));
    A.warm = warm != 0;
    A.every = every;
    A.steps = taken;
    A.recorded = true;
    return worst;
}

int sphmhd_adjoint_steps(sphmhd_ctx* c, int64_t* steps) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!steps) return fail(SPHMHD_EINVAL, "null argument");
    *steps = c->adj.recorded ? c->adj.steps : -1;
    return 0;
}

int sphmhd_adjoint_seed_totals(sphmhd_ctx* c, const double* weights) {
    int rc = adj_check(c);
    if (rc) return rc;
    if (!weights) return fail(SPHMHD_EINVAL, "null argument");
    if (weights[SPHMHD_T_DIVB] != 0.0) return fail(SPHMHD_EINVAL, "the divB total has no derivative: weight 0");
    Weights w;
    for (int q = 0; q < SPHMHD_NTOTALS; ++q) w.w[q] = weights[q];
    launch_a_seed_totals(c->d, c->adj.a, w);
    HIPCHK(hipGetLastError());
    return 0;
}

int sphmhd_adjoint_reverse(sphmhd_ctx* c, double* param_bars) {
    int rc = adj_check(c);
    if (rc) return rc;
    sphmhd_ctx::Adj& A = c->adj;
    if (!A.recorded) return fail(SPHMHD_ESTATE, "no recorded rollout");
    Dev& d = c->d;
    AdjDev& a = A.a;
    const int64_t n = d.n;
    const int ninst = d.ninst;
    HIPCHK(hipMemsetAsync(a.ai, 0, sizeof(AdjInst) * ninst, d.stream));
    launch_a_zero(d, a.par, (int64_t)SPHMHD_NTPARAMS * n);
    const int64_t nseg = (int64_t)A.ckpt.size();
    const int64_t sn = (int64_t)kNState * n;
    for (int64_t s = nseg - 1; s >= 0; --s) {
        const int64_t len = std::min<int64_t>(A.every, A.steps - s * A.every);
        // The segment again from its checkpoint, keeping each step's state,
        // clock and evaluation.  The forward run took these very steps, so
        // the recomputed evaluations are enqueued without a host wait each,
        // with the density passes the recent evaluations needed; one that
        // needed more passes or longer lists raises the abort flag, every
        // kernel after it exits at entry, and the segment is then recomputed
        // with the wait after each evaluation.
        for (int attempt = 0; attempt < 2; ++attempt) {
            const bool wait_each = attempt > 0;
            launch_a_rows(d, rows_state(), A.ckpt[s], true);
            launch_a_clock(d, A.ckpt[s] + sn, true);
            const int spec = next_spec(c);
            for (int64_t j = 0; j < len; ++j) {
                sphmhd_ctx::Adj::Snap& sp = A.snaps[j];
                launch_a_rows(d, rows_state(), A.seg + j * sn, false);
                launch_a_clock(d, A.clk + 2 * ninst * j, false);
                launch_initialize(d);
                launch_stage1(d);
                a.ain = sp.rows + (int64_t)kSnapRows * n;
                launch_a_save_entry(d, a);
                if (wait_each) {
                    if ((rc = do_eval(c)) < 0) return rc;
                } else if ((rc = eval_front(c, spec))) {
                    return rc;
                }
                if ((rc = adj_snap_save(c, sp))) return rc;
                launch_stage2(d);
                launch_clock(d);
                HIPCHK(hipGetLastError());
            }
            if (wait_each) break;
            if ((rc = read_ctl(c))) return rc;
            // evaluations that exited at entry ended nothing: take the count
            // of ended evaluations from the device
            c->seq_next = c->hcur->seq;
            if (!c->hcur->abort) break;
            launch_clear_abort(d);
        }
        launch_a_clock(d, A.clk + 2 * ninst * len, false);
        // its steps backwards, each from its kept evaluation
        for (int64_t j = len - 1; j >= 0; --j) {
            const sphmhd_ctx::Adj::Snap& sp = A.snaps[j];
            Dev v = adj_snap_view(c, sp);
            a.ain = sp.rows + (int64_t)kSnapRows * n;
            a.snap = A.seg + j * sn;
            launch_a_eval_begin(v, a, A.clk + 2 * ninst * (j + 1), 0);
            launch_a_stage2(v, a);
            launch_a_eval(v, a);
            launch_a_stage1_init(v, a);
            launch_a_reduce_dt(v, a);
            HIPCHK(hipGetLastError());
        }
    }
    a.ain = A.ain_live;
    if (A.warm) {
        const double* clk_next = nseg ? A.ckpt[0] + sn : A.final_clk;
        if ((rc = adj_warm_reverse(c, clk_next))) return rc;
    }
    launch_a_reduce_par(d, a);
    // leave the fields at the recorded initial state (the bars are its gradient)
    launch_a_rows(d, rows_all(), A.init, true);
    launch_a_clock(d, A.init + (int64_t)SPHMHD_NFIELDS * n, true);
    HIPCHK(hipGetLastError());
    std::vector<AdjInst> h(ninst);
    HIPCHK(hipMemcpyAsync(h.data(), a.ai, sizeof(AdjInst) * ninst, hipMemcpyDeviceToHost, d.stream));
    HIPCHK(hipStreamSynchronize(d.stream));
    for (int k = 0; k < ninst; ++k) {
        if (h[k].status & 2)
            return fail(SPHMHD_ECAP, "adjoint: a density iteration ran more passes than the reverse sweep holds");
        if (h[k].status & 1)
            return fail(SPHMHD_ESTATE, "adjoint: the recomputed density iteration differs from the forward one");
        if (param_bars)
            for (int p = 0; p < SPHMHD_NTPARAMS; ++p) param_bars[(int64_t)k * SPHMHD_NTPARAMS + p] = h[k].par[p];
    }
    return 0;
}

int sphmhd_synchronize(sphmhd_ctx* c) {
    if (!c) return fail(SPHMHD_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->d.stream));
    tcollect(c);
    return 0;
}

int sphmhd_set_timing(sphmhd_ctx* c, int32_t enable) {
    if (!c) return fail(SPHMHD_EINVAL, "null context");
    HIPCHK(hipStreamSynchronize(c->d.stream));
    for (PhaseTimer& t : c->tm) {
        if (enable && !t.a) {
            HIPCHK(hipEventCreate(&t.a));
            HIPCHK(hipEventCreate(&t.b));
        }
        t.ms = 0.0;
        t.count = 0;
        t.pending = false;
    }
    c->timing = enable != 0;
    c->graph_gen += 1;  // captured steps record events only when timing was on
    return 0;
}

int sphmhd_get_timing(sphmhd_ctx* c, sphmhd_timing* out) {
    if (!c || !out) return fail(SPHMHD_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(c->d.stream));
    tcollect(c);
    out->force_ms = c->tm[T_FORCE].ms;
    out->density_ms = c->tm[T_DENSITY].ms;
    out->nnps_ms = c->tm[T_NNPS].ms;
    out->lists_ms = c->tm[T_LISTS].ms;
    out->stage_ms = c->tm[T_STAGE].ms;
    out->force_launches = c->tm[T_FORCE].count;
    out->density_launches = c->tm[T_DENSITY].count;
    out->nnps_builds = c->tm[T_NNPS].count;
    out->list_builds = c->tm[T_LISTS].count;
    out->stage_launches = c->tm[T_STAGE].count;
    return 0;
}

}  // extern "C"
