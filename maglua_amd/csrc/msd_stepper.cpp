// Fused step drivers of libmsd.so: Heun (PredictorCorrector single pass), the RK
// stage set, learned windows, graphs, the autotuner and per-launch timing.
#include "msd_host.h"

// Whether two argument sets launch the same kernels with the same shapes, so that a
// graph captured for one serves the other once its step nodes take the new values.
static bool same_launch(const StepArgs& a, const StepArgs& b)
{
    return std::memcmp(&a.d, &b.d, sizeof a.d) == 0 && a.has_dip == b.has_dip && (a.Slo == nullptr) == (b.Slo == nullptr) &&
           a.ty == b.ty && std::memcmp(&a.dp, &b.dp, sizeof a.dp) == 0;
}

int msd_stepper::refresh()
{
    StepArgs a;
    std::memset(&a, 0, sizeof a);     // padding too: run_kind compares whole copies
    a.d = s->d;
    a.dt = (float)s->dt;
    a.S = s->S; a.P = P; a.C = C;
    a.Slo = compensated ? s->Slo : nullptr;
    a.Clo = compensated ? Clo : nullptr;
    if (ex) { if (int rc = ex->compile()) return rc; a.ex = ex->fused_dev(); }
    if (an) { if (int rc = an->compile()) return rc; a.an = an->dev; }
    if (th) { if (int rc = th->dev(a.th)) return rc; }
    a.lp = s->llgdev(&llg);
    a.has_dip = dip ? 1 : 0;
    if (dip) {
        if (int rc = dip->prepare()) return rc;
        a.dp = dip->dev;
        a.ty = dip->dev.ty;
    } else if (lr3) {
        if (int rc = lr3->prepare()) return rc;
        a.ty = lr3->ty;
    } else {
        a.ty = 1;
        for (int c : {8, 4, 2}) if (s->d.ny % c == 0) { a.ty = c; break; }
    }
    if (uni) {
        if (int rc = uni->compile()) return rc;
        a.uni = UniDev{uni->start, uni->dops, (float)uni->scale};
    }
    if (!af.empty()) {
        // uniform fields fold into one vector; per-site ones into one [3][n] sum
        double B[3] = {0, 0, 0};
        bool het = false;
        for (msd_appliedfield* f : af) {
            if (f->het) het = true;
            else for (int c = 0; c < 3; ++c) B[c] += f->B[c] * f->scale;
        }
        a.has_app = 1;
        a.Bapp = make_float3((float)B[0], (float)B[1], (float)B[2]);
        if (het) {
            const size_t n3 = (size_t)3 * s->d.n;
            if (!Happ)
                if (int rc = dalloc(&Happ, n3)) return rc;
            hipError_t e = launch_fill(Happ, 0.0f, n3, G.stream);
            for (msd_appliedfield* f : af) {
                if (!f->het) continue;
                if (int rc = f->upload()) return rc;
                if (e == hipSuccess) e = launch_axpy(Happ, (float)f->scale, f->dh, n3, G.stream);
            }
            if (e != hipSuccess) return fail(MSD_EDEVICE, "stepper: applied fields: %s", hipGetErrorString(e));
            a.Happ = Happ;
        }
    }
    a.draw = draw;
    args = a;
    if (lr3) {
        const std::vector<size_t> lds = {lr3->lds_rows, lr3->lds_y, lr3->lds_z};
        const uint64_t gen0 = b3.gen;
        if (int rc = b3.update({rep()}, {args.ty}, lds)) return rc;
        if (b3.gen != gen0) drop_graph();   // the captured launches held the old work tables
    }
    return MSD_OK;
}

int msd_stepper::launch_step(hipStream_t st)
{
    hipError_t e = hipSuccess;
    if (kind() == SK_CONV3) {
        e = b3.evaluate(R3_PRED, compensated, st);
        if (e == hipSuccess) e = b3.evaluate(R3_CORR, compensated, st);
    } else {
        const void* fp = nullptr;
        const void* fc = nullptr;
        if (args.has_dip) e = launch_cols(args.dp, st);
        if (e == hipSuccess) e = launch_step_pred(args, st, &fp);
        if (e == hipSuccess && args.has_dip) e = launch_cols(args.dp, st);
        if (e == hipSuccess) e = launch_step_corr(args, st, &fc);
        log_step(fp, 0);
        log_step(fc, 0);
    }
    if (e != hipSuccess) return fail(MSD_EDEVICE, "step launch: %s", hipGetErrorString(e));
    return MSD_OK;
}

// One Heun step with the surrogate's dipole field: H_dip(X) = H_ref + dH(X - m_ref)
// for the predictor (X = S) and the corrector (X = P); the fused local kernels
// add both parts to exchange + anisotropy (+ thermal).
int msd_stepper::launch_learned_step(hipStream_t st, bool window_start)
{
    const StepArgs a = learned_args(false);
    const float sin = (float)sur->s_in;
    hipError_t e = hipSuccess;
    const void* fn = nullptr;
    if (window_start) {
        // S is m_ref at a window's first step: the network's input is zero and so is
        // its output (linear layers without bias), so the predictor takes H_ref alone
        e = launch_step_pred(learned_args(true), st, &fn);
        log_step(fn, 1);
    } else {
        e = launch_cnn_input(s->d, s->S, Mref, sin, sur->X0, st);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "learned step: %s", hipGetErrorString(e));
        if (int rc = sur->forward(Hext, nullptr, st)) return rc;
        e = launch_step_pred(a, st, &fn);
        log_step(fn, 2);
    }
    if (e == hipSuccess) e = launch_cnn_input(s->d, P, Mref, sin, sur->X0, st);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "learned step: %s", hipGetErrorString(e));
    if (int rc = sur->forward(Hext, nullptr, st)) return rc;
    e = launch_step_corr(a, st, &fn);
    log_step(fn, 2);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "learned step: %s", hipGetErrorString(e));
    return MSD_OK;
}

extern "C" int msd_stepper_create(msd_system* s, msd_exchange* ex, msd_anisotropy* ani, msd_dipole* dip, msd_thermal* th,
                                  msd_llg* llg, msd_stepper** out)
{
    NONNULL(s); NONNULL(out);
    *out = nullptr;
    const Dims& d = s->d;
    if ((ex && !same_dims(d, ex->d.nx, ex->d.ny, ex->d.nz)) || (ani && !same_dims(d, ani->d.nx, ani->d.ny, ani->d.nz)) ||
        (dip && !same_dims(d, dip->d.nx, dip->d.ny, dip->d.nz)) || (th && !same_dims(d, th->d.nx, th->d.ny, th->d.nz)))
        return fail(MSD_EMISMATCH, "stepper: operator and SpinSystem sizes differ");
    std::unique_ptr<msd_stepper> st(new msd_stepper);
    st->s = s; st->ex = ex; st->an = ani; st->dip = dip; st->th = th;
    if (llg) st->llg = *llg;
    if (int rc = dalloc(&st->P, (size_t)3 * d.n)) return rc;
    if (int rc = dalloc(&st->C, (size_t)3 * d.n)) return rc;
    if (int rc = dalloc(&st->Clo, (size_t)3 * d.n)) return rc;
    if (int rc = dalloc(&st->draw, 1)) return rc;
    hipError_t e = launch_fill(st->P, 0.0f, (size_t)3 * d.n, G.stream);
    if (e == hipSuccess) e = launch_fill(st->C, 0.0f, (size_t)3 * d.n, G.stream);
    if (e == hipSuccess) e = launch_fill(st->Clo, 0.0f, (size_t)3 * d.n, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "stepper: %s", hipGetErrorString(e));
    if (int rc = st->refresh()) return rc;
    *out = st.release();
    return MSD_OK;
}

extern "C" int msd_stepper_destroy(msd_stepper* st) { delete st; return MSD_OK; }
extern "C" int msd_stepper_set_graph(msd_stepper* st, int yes) { NONNULL(st); st->use_graph = yes != 0; return MSD_OK; }
extern "C" int msd_stepper_set_compensated(msd_stepper* st, int yes)
{
    NONNULL(st);
    st->compensated = yes != 0;
    st->drop_graph();
    return MSD_OK;
}

extern "C" int msd_stepper_set_longrange3d(msd_stepper* st, msd_longrange3d* l)
{
    NONNULL(st);
    if (l && !same_dims(st->s->d, l->d.nx, l->d.ny, l->d.nz)) return fail(MSD_EMISMATCH, "stepper: LongRange3D size differs");
    if (l && st->dip) return fail(MSD_EUNSUPPORTED, "stepper: a Dipole and a LongRange3D in one step driver");
    st->lr3 = l;
    st->drop_graph();
    st->b3.release();
    st->spec_version = 0;
    return st->refresh();
}

// state checks shared by single steppers and ensembles before a run of steps
int member_begin(msd_stepper* st)
{
    if (int rc = st->refresh()) return rc;
    if (st->compensated && !st->s->lo_valid) {
        hipError_t e = launch_fill(st->s->Slo, 0.0f, (size_t)3 * st->s->d.n, G.stream);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "stepper: %s", hipGetErrorString(e));
        st->s->lo_valid = true;
    }
    // device draw counter: the predictor increments it before the corrector reads it
    if (st->th) {
        const unsigned long long v = (unsigned long long)(st->th->draw - 1);
        HIPCHK(hipMemcpyAsync(st->draw, &v, sizeof v, hipMemcpyHostToDevice, G.stream));
        HIPCHK(hipStreamSynchronize(G.stream));
    }
    return MSD_OK;
}

bool spectrum_stale(const msd_stepper* st)
{
    return st->spec_version != st->s->version || (st->dip && st->dip->epoch != st->dip_epoch) ||
           (st->lr3 && st->lr3->epoch != st->lr3_epoch);
}

static int stepper_begin(msd_stepper* st)
{
    if (int rc = member_begin(st)) return rc;
    // spin spectrum of the current state for the first field evaluation
    if (spectrum_stale(st)) {
        hipError_t e = hipSuccess;
        if (st->kind() == SK_LAYERED) e = launch_rows_fwd(st->args.dp, st->s->S, G.stream);
        else if (st->kind() == SK_CONV3)
            e = st->b3.rows(R3_FWD, false, G.stream);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "stepper: %s", hipGetErrorString(e));
    }
    return MSD_OK;
}

void stepper_end(msd_stepper* st, int nsteps)
{
    st->s->time += nsteps * st->s->dt;
    if (!st->compensated) st->s->lo_valid = false;
    if (st->th) st->th->draw += (uint64_t)nsteps;
    st->s->version++;
    st->spec_version = st->s->version;
    // the steps left this system's spin spectrum in the operator's shared buffers:
    // claim them, so that another driver sharing the operator transforms again
    if (st->dip) st->dip_epoch = ++st->dip->epoch;
    if (st->lr3) st->lr3_epoch = ++st->lr3->epoch;
}

// Capture the steps of graph slot `slot` into a graph: 0 one exact step, 1 a learned
// window's first step, 2 a later learned step, 3 STEP_SPAN exact steps (the span graph)
static int capture_step(msd_stepper* st, int slot, hipGraph_t* graph, hipGraphExec_t* exec)
{
    const int kind = slot == 3 ? 0 : slot;
    st->cap_log.clear();
    st->gnodes[slot].clear();
    st->capturing = true;
    HIPCHK(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
    int rc = MSD_OK;
    for (int q = 0; q < (slot == 3 ? STEP_SPAN : 1) && !rc; ++q)
        rc = kind ? st->launch_learned_step(G.stream, kind == 1) : st->launch_step(G.stream);
    st->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(G.stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(MSD_EDEVICE, "graph capture: %s", hipGetErrorString(e));
    *graph = g;
    HIPCHK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    // the fused step kernels' nodes, matched by kernel: they can take new values later
    size_t n = 0;
    HIPCHK(hipGraphGetNodes(g, nullptr, &n));
    std::vector<hipGraphNode_t> nodes(n);
    if (n) HIPCHK(hipGraphGetNodes(g, nodes.data(), &n));
    for (hipGraphNode_t node : nodes) {
        hipGraphNodeType t;
        if (hipGraphNodeGetType(node, &t) != hipSuccess || t != hipGraphNodeTypeKernel) continue;
        hipKernelNodeParams p{};
        if (hipGraphKernelNodeGetParams(node, &p) != hipSuccess) continue;
        for (const auto& c : st->cap_log)
            if (c.first == p.func) { st->gnodes[slot].push_back({node, c.second}); break; }
    }
    // the 3D form's kernels read their arguments from device tables (Batch3), which
    // refresh() rewrites: its graphs have no step nodes to update
    st->gnodes_ok[slot] = st->gnodes[slot].size() == st->cap_log.size();
    return MSD_OK;
}

// New argument values into the step nodes of an instantiated graph; false when the
// graph's step nodes were not all found (the caller captures again)
static bool update_step_nodes(msd_stepper* st, int slot, hipGraphExec_t exec)
{
    if (!st->gnodes_ok[slot]) return false;
    for (const msd_stepper::StepNode& sn : st->gnodes[slot]) {
        StepArgs v = st->variant_args(sn.variant);
        hipKernelNodeParams p{};
        if (hipGraphKernelNodeGetParams(sn.node, &p) != hipSuccess) return false;
        void* kp[1] = {&v};
        p.kernelParams = kp;
        p.extra = nullptr;
        if (hipGraphExecKernelNodeSetParams(exec, sn.node, &p) != hipSuccess) return false;
    }
    return true;
}

// The instantiated graph of slot `slot` (as capture_step) for the current arguments:
// kept, its step nodes given new values in place, or captured again.
static int slot_graph(msd_stepper* st, int slot, hipGraphExec_t* out)
{
    const int kind = slot == 3 ? 0 : slot;
    hipGraphExec_t* exec = slot == 3 ? &st->mexec : kind ? &st->lexec[kind - 1] : &st->exec;
    hipGraph_t* graph = slot == 3 ? &st->mgraph : kind ? &st->lgraph[kind - 1] : &st->graph;
    StepArgs* gargs = slot == 3 ? &st->mgraph_args : kind ? &st->lgraph_args[kind - 1] : &st->graph_args;
    const bool scales_moved =
        kind && (st->lgraph_sin[kind - 1] != st->sur->s_in || st->lgraph_sout[kind - 1] != st->sur->s_out);
    bool fresh = *exec && same_launch(st->args, *gargs) && !scales_moved;
    // same kernels and shapes with other values: the step nodes take them in place
    if (fresh && std::memcmp(&st->args, gargs, sizeof(StepArgs)) != 0) {
        fresh = update_step_nodes(st, slot, *exec);
        if (fresh) *gargs = st->args;
    }
    if (!fresh) {
        if (*exec) (void)hipGraphExecDestroy(*exec);
        if (*graph) (void)hipGraphDestroy(*graph);
        *exec = nullptr; *graph = nullptr;
        if (int rc = capture_step(st, slot, graph, exec)) return rc;
        st->graph_builds++;
        *gargs = st->args;
        if (kind) { st->lgraph_sin[kind - 1] = st->sur->s_in; st->lgraph_sout[kind - 1] = st->sur->s_out; }
    }
    *out = *exec;
    return MSD_OK;
}

// nsteps of one kind (0 exact, 1 a learned window's first step, 2 later learned
// steps), graph-replayed when enabled; exact runs of at least STEP_SPAN steps replay
// the span graph and the remainder one step at a time
static int run_kind(msd_stepper* st, int nsteps, int kind)
{
    if (nsteps <= 0) return MSD_OK;
    if (st->use_graph) {
        int single = nsteps;
        if (kind == 0 && nsteps >= STEP_SPAN) {
            hipGraphExec_t span = nullptr;
            if (int rc = slot_graph(st, 3, &span)) return rc;
            for (int i = 0; i < nsteps / STEP_SPAN; ++i) HIPCHK(hipGraphLaunch(span, G.stream));
            single = nsteps % STEP_SPAN;
        }
        if (single > 0) {
            hipGraphExec_t one = nullptr;
            if (int rc = slot_graph(st, kind, &one)) return rc;
            for (int i = 0; i < single; ++i) HIPCHK(hipGraphLaunch(one, G.stream));
        }
    } else {
        for (int i = 0; i < nsteps; ++i)
            if (int rc = kind ? st->launch_learned_step(G.stream, kind == 1) : st->launch_step(G.stream)) return rc;
    }
    return MSD_OK;
}

// nsteps exact (learned = false) steps, or a learned window of nsteps steps
static int run_steps(msd_stepper* st, int nsteps, bool learned)
{
    if (!learned) return run_kind(st, nsteps, 0);
    if (int rc = run_kind(st, 1, 1)) return rc;
    return run_kind(st, nsteps - 1, 2);
}

// Windows of K steps on the surrogate field, each checked when it ends.  A window
// starts from a state whose exact field (layered Dipole or LongRange3D) is known:
// that field becomes H_ref and the state m_ref, and the state (hi and lo words) is
// kept.  After the window's learned steps the exact field of the state reached is
// compared with the surrogate's prediction for it, max|pred - exact| / max|exact|.
// Within tol the window stands, and its exact field starts the next window.
// Otherwise the kept state and the thermal draw counter are restored and the
// window runs again as exact steps, so every step kept is either checked or exact.
// tol = 0 makes every window exact: the trajectory is the exact driver's.
static int stepper_heun_learned(msd_stepper* st, int nsteps)
{
    if (int rc = member_begin(st)) return rc;
    if (int rc = st->sur->pack()) return rc;
    const size_t n3 = (size_t)3 * st->s->d.n;
    const ExactField exact{st->dip, st->lr3};
    const uint64_t draw0 = st->th ? st->th->draw : 0;
    bool fresh = false;      // Hx holds the exact field of the current state
    int done = 0;
    while (done < nsteps) {
        if (!fresh)
            if (int rc = exact.of(st->s->S, st->Hx)) return rc;
        HIPCHK(hipMemcpyAsync(st->Href, st->Hx, n3 * sizeof(float), hipMemcpyDeviceToDevice, G.stream));
        HIPCHK(hipMemcpyAsync(st->Mref, st->s->S, n3 * sizeof(float), hipMemcpyDeviceToDevice, G.stream));
        HIPCHK(hipMemcpyAsync(st->Ssnap, st->s->Slo, n3 * sizeof(float), hipMemcpyDeviceToDevice, G.stream));
        st->have_ref = true;
        const int L = std::min(st->sur_K, nsteps - done);
        if (int rc = run_steps(st, L, true)) return rc;
        // the check: the exact field of the state the window reached
        if (int rc = exact.of(st->s->S, st->Hx)) return rc;
        hipError_t e = launch_cnn_input(st->s->d, st->s->S, st->Mref, (float)st->sur->s_in, st->sur->X0, G.stream);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "surrogate check: %s", hipGetErrorString(e));
        if (int rc = st->sur->forward(st->Hext, st->Href, G.stream)) return rc;
        double md, mb;
        if (int rc = field_maxdiff(st->Hext, st->Hx, n3, &md, &mb)) return rc;
        const double res = mb > 0 ? md / mb : md;
        st->sur_last_res = res;
        st->sur_max_res = std::max(st->sur_max_res, res);
        st->sur_windows++;
        fresh = st->sur_tol > 0 && res <= st->sur_tol;
        if (!fresh) {
            // restore the window's start (m_ref holds its hi words) and run it exactly
            st->sur_fallbacks++;
            st->sur_exact_windows++;
            HIPCHK(hipMemcpyAsync(st->s->S, st->Mref, n3 * sizeof(float), hipMemcpyDeviceToDevice, G.stream));
            HIPCHK(hipMemcpyAsync(st->s->Slo, st->Ssnap, n3 * sizeof(float), hipMemcpyDeviceToDevice, G.stream));
            if (st->th) {
                const unsigned long long v = (unsigned long long)(draw0 + (uint64_t)done - 1);
                HIPCHK(hipMemcpyAsync(st->draw, &v, sizeof v, hipMemcpyHostToDevice, G.stream));
                HIPCHK(hipStreamSynchronize(G.stream));
            }
            // the exact steps start from the spin spectrum of the restored state
            e = st->lr3 ? st->b3.rows(R3_FWD, false, G.stream) : launch_rows_fwd(st->args.dp, st->s->S, G.stream);
            if (e != hipSuccess) return fail(MSD_EDEVICE, "surrogate fallback: %s", hipGetErrorString(e));
            if (int rc = run_steps(st, L, false)) return rc;
        }
        done += L;
    }
    stepper_end(st, nsteps);
    st->spec_version = 0;    // learned steps leave no spin spectrum behind
    return MSD_OK;
}

extern "C" int msd_stepper_heun(msd_stepper* st, int nsteps)
{
    NONNULL(st);
    if (nsteps < 0) return fail(MSD_EINVAL, "heun: nsteps %d", nsteps);
    if (nsteps == 0) return MSD_OK;
    if (st->sur) return stepper_heun_learned(st, nsteps);
    if (int rc = stepper_begin(st)) return rc;
    if (int rc = run_steps(st, nsteps, false)) return rc;
    stepper_end(st, nsteps);
    return MSD_OK;
}

// Butcher tables of common/scripts/RungeKutta.lua:57-160 (published coefficients; the
// rows, c column omitted, define the integrator a script selects by name)
static bool butcher_rows(const std::string& k, std::vector<std::vector<double>>& t)
{
    if (k == "RK1") t = {{1}};
    else if (k == "RK2") t = {{0.5}, {0, 1}};
    else if (k == "K3") t = {{0.5}, {-1, 2}, {1.0 / 6, 2.0 / 3, 1.0 / 6}};
    else if (k == "RK4") t = {{0.5}, {0, 0.5}, {0, 0, 1}, {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6}};
    else if (k == "RK3") t = {{1.0 / 3}, {0, 2.0 / 3}, {0.25, 0, 0.75}};
    else if (k == "RK4_38") t = {{1.0 / 3}, {-1.0 / 3, 1}, {1, -1, 1}, {0.125, 0.375, 0.375, 0.125}};
    else if (k == "RK6")
        t = {{1.0 / 3}, {0, 2.0 / 3}, {1.0 / 12, 4.0 / 12, -1.0 / 12}, {-1.0 / 16, 18.0 / 16, -3.0 / 16, -6.0 / 16},
             {0, 9.0 / 8, -3.0 / 8, -6.0 / 8, 4.0 / 8}, {9.0 / 44, -36.0 / 44, 63.0 / 44, 72.0 / 44, 0, -64.0 / 44},
             {11.0 / 120, 0, 81.0 / 120, 81.0 / 120, -32.0 / 120, -32.0 / 120, 11.0 / 120}};
    else if (k == "BS3") t = {{0.5}, {0, 0.75}, {2.0 / 9, 1.0 / 3, 4.0 / 9}, {7.0 / 24, 0.25, 1.0 / 3, 0.125}};
    else return false;
    return true;
}

// One step of make_rk_step_function (RungeKutta.lua:185-214) with the fused stage
// kernels: per stage, the layered column pass of the stage spins (with a Dipole)
// and one row kernel (field, Butcher row, thermal kick at the last).
static int rk_launch_step(msd_stepper* st, const std::vector<std::vector<double>>& bt, hipStream_t s)
{
    const int ns = (int)bt.size();
    const size_t n = st->s->d.n;
    hipError_t e = hipSuccess;
    if (st->th) e = launch_draw_advance(st->draw, s);
    for (int i = 0; i < ns && e == hipSuccess; ++i) {
        if (st->args.has_dip) e = launch_cols(st->args.dp, s);
        if (e != hipSuccess) break;
        RKStage r{};
        r.X = i == 0 ? st->s->S : st->rk_k[i];
        r.Hout = st->rk_h[i];
        // the last row writes S, but one stage that also reads S's neighbours (RK1)
        // writes a copy first
        r.out = i + 1 < ns ? st->rk_k[i + 1] : (ns == 1 ? st->rk_tmp : st->s->S);
        for (int j = 0; j < i; ++j) {
            r.K[j] = j == 0 ? st->s->S : st->rk_k[j];
            r.H[j] = st->rk_h[j];
        }
        r.terms = i + 1;
        for (int j = 0; j <= i; ++j) r.f[j] = (float)bt[i][j];
        r.last = i + 1 == ns;
        e = launch_rk_stage(st->args, r, s);
    }
    if (e == hipSuccess && ns == 1) e = hipMemcpyAsync(st->s->S, st->rk_tmp, 3 * n * sizeof(float), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "RK step: %s", hipGetErrorString(e));
    return MSD_OK;
}

extern "C" int msd_stepper_rk(msd_stepper* st, const char* kind, int nsteps)
{
    NONNULL(st); NONNULL(kind);
    std::vector<std::vector<double>> bt;
    if (!butcher_rows(kind, bt))
        return fail(MSD_EINVAL, "RK type must be one of \"RK1\", \"RK2\", \"K3\", \"RK3\", \"RK4\", \"RK4_38\", \"RK6\", \"BS3\"");
    if (nsteps < 0) return fail(MSD_EINVAL, "rk: nsteps %d", nsteps);
    if (st->lr3) return fail(MSD_EUNSUPPORTED, "rk: the fused Runge-Kutta stages cover the local and layered forms");
    if (st->sur) return fail(MSD_EUNSUPPORTED, "rk: the learned dipole field runs with the Heun step");
    if (nsteps == 0) return MSD_OK;
    const int ns = (int)bt.size();
    const size_t n3 = (size_t)3 * st->s->d.n;
    while ((int)st->rk_k.size() < ns) {
        float* b = nullptr;
        if (int rc = dalloc(&b, n3)) return rc;
        st->rk_k.push_back(b);
    }
    while ((int)st->rk_h.size() < ns) {
        float* b = nullptr;
        if (int rc = dalloc(&b, n3)) return rc;
        st->rk_h.push_back(b);
    }
    if (!st->rk_tmp)
        if (int rc = dalloc(&st->rk_tmp, n3)) return rc;
    if (int rc = stepper_begin(st)) return rc;
    if (st->use_graph) {
        if (!st->rk_exec || st->rk_kind != kind || std::memcmp(&st->rk_args, &st->args, sizeof(StepArgs)) != 0) {
            if (st->rk_exec) (void)hipGraphExecDestroy(st->rk_exec);
            if (st->rk_graph) (void)hipGraphDestroy(st->rk_graph);
            st->rk_exec = nullptr; st->rk_graph = nullptr;
            HIPCHK(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
            int rc = rk_launch_step(st, bt, G.stream);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(G.stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return fail(MSD_EDEVICE, "RK graph capture: %s", hipGetErrorString(e));
            st->rk_graph = g;
            HIPCHK(hipGraphInstantiate(&st->rk_exec, g, nullptr, nullptr, 0));
            st->graph_builds++;
            st->rk_kind = kind;
            st->rk_args = st->args;
        }
        for (int i = 0; i < nsteps; ++i) HIPCHK(hipGraphLaunch(st->rk_exec, G.stream));
    } else {
        for (int i = 0; i < nsteps; ++i)
            if (int rc = rk_launch_step(st, bt, G.stream)) return rc;
    }
    // time as the script's sub-steps advance it: the last row's terms from S, each
    // nonzero one adding its coefficient times dt (a zero first term copies, no time)
    const double dt = st->s->dt;
    for (int k = 0; k < nsteps; ++k) {
        const std::vector<double>& f = bt.back();
        if (f[0] * dt != 0.0) st->s->time = st->s->time + f[0] * dt;
        for (size_t j = 1; j < f.size(); ++j)
            if (f[j] != 0.0) st->s->time = st->s->time + f[j] * dt;
    }
    // the state is plain FP32 words now: the compensated low words no longer apply
    st->s->lo_valid = false;
    if (st->th) st->th->draw += (uint64_t)nsteps;
    st->s->version++;
    st->spec_version = st->s->version;
    if (st->dip) st->dip_epoch = ++st->dip->epoch;
    return MSD_OK;
}

// One driver's autotuner trials: its state kept (StateKeep) and a pair of events.
struct TrialState : StateKeep {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    explicit TrialState(msd_stepper* s) : StateKeep(s) {}
    ~TrialState()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    int save()
    {
        if (int rc = StateKeep::save()) return rc;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return fail(MSD_EDEVICE, "autotune: events");
        return MSD_OK;
    }
    // ms per step of `steps` graph-replayed steps of the current configuration (after
    // a capture and a warm step), the state restored afterwards
    int time(int steps, double* per)
    {
        if (int rc = stepper_begin(st)) return rc;
        if (int rc = run_steps(st, 2, false)) return rc;
        if (hipEventRecord(e0, G.stream) != hipSuccess) return fail(MSD_EDEVICE, "autotune: event");
        if (int rc = run_steps(st, steps, false)) return rc;
        if (hipEventRecord(e1, G.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return fail(MSD_EDEVICE, "autotune: event");
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *per = ms / steps;
        return restore();
    }
};

// Autotuner of the layered step: times `steps` graph-replayed steps for every
// compiled row-tile shape (rows per tile, threads per workgroup) of the lattice
// and keeps the fastest.  The spin state, its low words, time and thermal draw
// are restored after every trial, so tuning leaves the trajectory untouched.
extern "C" int msd_stepper_autotune(msd_stepper* st, int steps, int* best_ty, int* best_tpb, int* best_split,
                                    double* best_ms)
{
    NONNULL(st); NONNULL(best_ty); NONNULL(best_tpb); NONNULL(best_split); NONNULL(best_ms);
    if (st->kind() != SK_LAYERED) return fail(MSD_EUNSUPPORTED, "autotune: needs a layered Dipole step driver");
    if (steps < 1) return fail(MSD_EINVAL, "autotune: steps %d", steps);
    msd_dipole* dip = st->dip;
    int tys[16], tpbs[16];
    const int nc = row_shapes_for(dip->d.nx, dip->d.ny, tys, tpbs, 16);
    *best_ty = dip->dev.ty; *best_tpb = dip->dev.tpb; *best_split = dip->dev.split; *best_ms = -1.0;
    if (nc == 0) return MSD_OK;           // runtime-shaped kernels only: nothing to choose
    TrialState ts(st);
    int rc = ts.save();
    // every trial rebuilds the step (and the tensor words when the column form changes)
    auto use = [&](int ty, int tpb, bool sp) -> int {
        if (int r = layered_shape(dip, ty, tpb, true, sp, "autotune")) return r;
        st->drop_graph();
        dip->epoch++;
        return dip->prepare();
    };
    const DipoleDev orig = dip->dev;
    for (int c = 0; c < 2 * nc && !rc; ++c) {
        const int ty = tys[c / 2], tpb = tpbs[c / 2];
        const bool sp = (c & 1) == 0;
        const size_t need = (size_t)((3 * ty + 1) / 2 + 1) * dip->d.nx * sizeof(float2);
        if (need > G.max_lds) continue;
        if (layered_shape(dip, ty, tpb, true, sp, "autotune") != MSD_OK) { dip->dev = orig; continue; }
        if (sp && dip->dev.split == 0) continue;    // no split form for this shape: its full-column trial covers it
        if ((rc = use(ty, tpb, sp))) break;
        double per = 0.0;
        if ((rc = ts.time(steps, &per))) break;
        if (*best_ms < 0 || per < *best_ms) {
            *best_ms = per; *best_ty = ty; *best_tpb = tpb; *best_split = dip->dev.split;
        }
    }
    int rc2 = MSD_OK;
    if (!rc && *best_ms > 0) rc2 = use(*best_ty, *best_tpb, *best_split != 0);
    else {
        dip->dev = orig;
        st->drop_graph();
        dip->epoch++;
        rc2 = dip->prepare();
    }
    if (!rc) rc = rc2;
    (void)hipStreamSynchronize(G.stream);
    if (!rc) rc = st->refresh();
    st->spec_version = 0;
    return rc;
}

// Autotuner of the LongRange3D step: times `steps` graph-replayed steps for every
// pair of compiled kx widths of the y pass and of the z pass form and keeps the
// fastest, the state restored after every trial.
extern "C" int msd_stepper_autotune_passes(msd_stepper* st, int steps, int* best_wy, int* best_wz, double* best_ms)
{
    NONNULL(st); NONNULL(best_wy); NONNULL(best_wz); NONNULL(best_ms);
    if (st->kind() != SK_CONV3) return fail(MSD_EUNSUPPORTED, "autotune: needs a LongRange3D step driver");
    if (steps < 1) return fail(MSD_EINVAL, "autotune: steps %d", steps);
    msd_longrange3d* l = st->lr3;
    int wys[8], wzs[8];
    int ny = l->yt ? ypass_widths(l->yt, wys, 8) : 0;
    int nz = l->zt ? zconv_widths(l->zt, wzs, 8) : 0;
    if (ny == 0) { wys[0] = 0; ny = 1; }   // runtime-length pass: its width stays
    if (nz == 0) { wzs[0] = 0; nz = 1; }
    const int wy0 = l->wy, wz0 = l->wz;
    *best_wy = wy0; *best_wz = wz0; *best_ms = -1.0;
    if (ny * nz <= 1) return MSD_OK;
    TrialState ts(st);
    int rc = ts.save();
    for (int c = 0; c < ny * nz && !rc; ++c) {
        const int wy = wys[c / nz], wz = wzs[c % nz];
        if (lr3_set_widths(l, wy, wz) != MSD_OK) continue;    // over the LDS: not a candidate
        double per = 0.0;
        if ((rc = ts.time(steps, &per))) break;
        if (*best_ms < 0 || per < *best_ms) { *best_ms = per; *best_wy = l->wy; *best_wz = l->wz; }
    }
    const int rc2 = (!rc && *best_ms > 0) ? lr3_set_widths(l, *best_wy, *best_wz) : lr3_set_widths(l, wy0, wz0);
    if (!rc) rc = rc2;
    (void)hipStreamSynchronize(G.stream);
    if (!rc) rc = st->refresh();
    st->spec_version = 0;
    return rc;
}

extern "C" int msd_stepper_set_uniaxial(msd_stepper* st, msd_uniaxial* u)
{
    NONNULL(st);
    if (u && !same_dims(st->s->d, u->d.nx, u->d.ny, u->d.nz)) return fail(MSD_EMISMATCH, "stepper: Anisotropy.Uniaxial size differs");
    st->uni = u;
    return MSD_OK;
}

extern "C" int msd_stepper_add_applied_field(msd_stepper* st, msd_appliedfield* a)
{
    NONNULL(st);
    if (!a) {
        st->af.clear();
        return MSD_OK;
    }
    if (!same_dims(st->s->d, a->d.nx, a->d.ny, a->d.nz)) return fail(MSD_EMISMATCH, "stepper: AppliedField size differs");
    for (msd_appliedfield* f : st->af)
        if (f == a) return MSD_OK;
    st->af.push_back(a);
    return MSD_OK;
}

extern "C" int msd_stepper_set_surrogate(msd_stepper* st, msd_surrogate* sur, int K, double tol)
{
    NONNULL(st);
    if (!sur) {
        st->sur = nullptr;
        st->have_ref = false;
        st->spec_version = 0;
        return MSD_OK;
    }
    if (!st->dip && !st->lr3) return fail(MSD_EUNSUPPORTED, "stepper: the surrogate replaces a Dipole or LongRange3D field");
    if (!same_dims(st->s->d, sur->d.nx, sur->d.ny, sur->d.nz)) return fail(MSD_EMISMATCH, "stepper: Surrogate size differs");
    if (K < 1) return fail(MSD_EINVAL, "stepper: refresh interval %d", K);
    if (!(tol >= 0)) return fail(MSD_EINVAL, "stepper: residual tolerance must be >= 0");
    const size_t n3 = (size_t)3 * st->s->d.n;
    for (float** p : {&st->Href, &st->Mref, &st->Hext, &st->Hx, &st->Ssnap})
        if (!*p)
            if (int rc = dalloc(p, n3)) return rc;
    st->sur = sur;
    st->sur_K = K;
    st->sur_tol = tol;
    st->have_ref = false;
    st->sur_windows = st->sur_exact_windows = st->sur_fallbacks = 0;
    st->sur_last_res = st->sur_max_res = 0.0;
    st->drop_learned_graphs();
    return MSD_OK;
}

extern "C" int msd_stepper_graph_builds(const msd_stepper* st, long* n)
{
    NONNULL(st); NONNULL(n);
    *n = st->graph_builds;
    return MSD_OK;
}

extern "C" int msd_stepper_surrogate_stats(const msd_stepper* st, long* windows, long* exact_windows, long* fallbacks,
                                           double* last_residual, double* max_residual)
{
    NONNULL(st); NONNULL(windows); NONNULL(exact_windows); NONNULL(fallbacks); NONNULL(last_residual); NONNULL(max_residual);
    *windows = st->sur_windows;
    *exact_windows = st->sur_exact_windows;
    *fallbacks = st->sur_fallbacks;
    *last_residual = st->sur_last_res;
    *max_residual = st->sur_max_res;
    return MSD_OK;
}

// one launch of a step's sequence, by code (for the event-timed profile)
enum { L_COLS, L_PRED2, L_CORR2, L_Y_FWD, L_ZCONV, L_Y_INV, L_PRED3, L_CORR3 };
static const char* launch_name(int code)
{
    static const char* names[] = {"k_cols", "k_step_pred", "k_step_corr", "k3_ypass_fwd",
                                  "k3_zconv", "k3_ypass_inv", "k3_rows_pred", "k3_rows_corr"};
    return names[code];
}

static hipError_t launch_code(const msd_stepper* st, int code, hipStream_t s)
{
    const Batch3& b = st->b3;
    switch (code) {
        case L_COLS: return launch_cols(st->args.dp, s);
        case L_PRED2: return launch_step_pred(st->args, s);
        case L_CORR2: return launch_step_corr(st->args, s);
        case L_Y_FWD: return b.ypass(-1, s);
        case L_ZCONV: return b.zpass(s);
        case L_Y_INV: return b.ypass(1, s);
        case L_PRED3: return b.rows(R3_PRED, st->compensated, s);
        default: return b.rows(R3_CORR, st->compensated, s);
    }
}

extern "C" int msd_stepper_profile(msd_stepper* st, int nsteps, int cap, char names[][32], double* avg_ms, int* count)
{
    NONNULL(st); NONNULL(names); NONNULL(avg_ms); NONNULL(count);
    if (nsteps <= 0) return fail(MSD_EINVAL, "profile: nsteps %d", nsteps);
    if (int rc = stepper_begin(st)) return rc;
    std::vector<int> seq;
    if (st->kind() == SK_CONV3) seq = {L_Y_FWD, L_ZCONV, L_Y_INV, L_PRED3, L_Y_FWD, L_ZCONV, L_Y_INV, L_CORR3};
    else if (st->kind() == SK_LAYERED) seq = {L_COLS, L_PRED2, L_COLS, L_CORR2};
    else seq = {L_PRED2, L_CORR2};
    const int per = (int)seq.size();
    if (st->kind() != SK_CONV3) {
        // Each launch of one step repeated nsteps times back to back between two
        // events, so the averages carry no per-launch event gaps.  The layered and
        // local kernels are idempotent in their inputs (columns: Mh -> Hh;
        // predictor: S, Hh -> P, C, Mh; corrector: P, C, Hh -> S, Mh), so the
        // repeats leave the state exactly one step ahead once the predictor's
        // draw-counter increments are taken back to one.
        std::vector<hipEvent_t> ev(2 * (size_t)per);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        int rc = MSD_OK;
        for (int k = 0; k < per && !rc; ++k) {
            hipError_t e = hipEventRecord(ev[2 * k], G.stream);
            for (int i = 0; i < nsteps && e == hipSuccess; ++i) e = launch_code(st, seq[k], G.stream);
            if (e == hipSuccess) e = hipEventRecord(ev[2 * k + 1], G.stream);
            if (e == hipSuccess && seq[k] == L_PRED2 && st->th) {
                const unsigned long long v = (unsigned long long)st->th->draw;
                e = hipMemcpyAsync(st->draw, &v, sizeof v, hipMemcpyHostToDevice, G.stream);
                if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
            }
            if (e != hipSuccess) rc = fail(MSD_EDEVICE, "profile: %s", hipGetErrorString(e));
        }
        if (!rc) {
            HIPCHK(hipStreamSynchronize(G.stream));
            std::vector<std::string> nm;
            std::vector<double> tot;
            std::vector<int> cnt;
            for (int k = 0; k < per; ++k) {
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
                const char* name = launch_name(seq[k]);
                size_t u = 0;
                while (u < nm.size() && nm[u] != name) ++u;
                if (u == nm.size()) { nm.push_back(name); tot.push_back(0); cnt.push_back(0); }
                tot[u] += ms; cnt[u] += nsteps;
            }
            *count = std::min<int>((int)nm.size(), cap);
            for (int u = 0; u < *count; ++u) {
                std::snprintf(names[u], 32, "%s", nm[u].c_str());
                avg_ms[u] = tot[u] / cnt[u];
            }
            stepper_end(st, 1);
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        return rc;
    }
    std::vector<hipEvent_t> ev((size_t)per * nsteps + 1);
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    int rc = MSD_OK;
    size_t q = 0;
    HIPCHK(hipEventRecord(ev[q++], G.stream));
    for (int i = 0; i < nsteps && !rc; ++i) {
        hipError_t e = hipSuccess;
        for (int k = 0; k < per && e == hipSuccess; ++k) {
            e = launch_code(st, seq[k], G.stream);
            if (e == hipSuccess) e = hipEventRecord(ev[q++], G.stream);
        }
        if (e != hipSuccess) rc = fail(MSD_EDEVICE, "profile: %s", hipGetErrorString(e));
    }
    if (!rc) {
        HIPCHK(hipStreamSynchronize(G.stream));
        std::vector<std::string> nm;
        std::vector<double> tot;
        std::vector<int> cnt;
        for (size_t j = 1; j < q; ++j) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[j - 1], ev[j]));
            const char* name = launch_name(seq[(j - 1) % per]);
            size_t u = 0;
            while (u < nm.size() && nm[u] != name) ++u;
            if (u == nm.size()) { nm.push_back(name); tot.push_back(0); cnt.push_back(0); }
            tot[u] += ms; cnt[u]++;
        }
        *count = std::min<int>((int)nm.size(), cap);
        for (int u = 0; u < *count; ++u) {
            std::snprintf(names[u], 32, "%s", nm[u].c_str());
            avg_ms[u] = tot[u] / cnt[u];
        }
        stepper_end(st, nsteps);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}
