// Ensembles of libmsd.so: independent replicas advanced by one set of launches
// per pass, and their per-pass timing.
#include "msd_host.h"

extern "C" int msd_ensemble_create(msd_stepper* const* members, int count, msd_ensemble** out)
{
    NONNULL(members); NONNULL(out);
    *out = nullptr;
    if (count < 1) return fail(MSD_EINVAL, "ensemble: %d members", count);
    std::unique_ptr<msd_ensemble> e(new msd_ensemble);
    for (int i = 0; i < count; ++i) {
        msd_stepper* s = members[i];
        if (!s) return fail(MSD_EINVAL, "ensemble: member %d is NULL", i);
        for (int j = 0; j < i; ++j) {
            if (members[j] == s || members[j]->s == s->s) return fail(MSD_EINVAL, "ensemble: member %d repeats a SpinSystem", i);
            if (s->lr3 && members[j]->lr3 == s->lr3)
                return fail(MSD_EINVAL, "ensemble: members %d and %d share a LongRange3D (its spectrum buffer)", j, i);
            if (s->dip && members[j]->dip == s->dip)
                return fail(MSD_EINVAL, "ensemble: members %d and %d share a Dipole (its spectrum buffers)", j, i);
            // one Thermal would give both replicas the same noise and advance its draw twice a step
            if (s->th && members[j]->th == s->th)
                return fail(MSD_EINVAL, "ensemble: members %d and %d share a Thermal (its noise stream)", j, i);
        }
        e->m.push_back(s);
    }
    if (int rc = e->refresh()) return rc;
    *out = e.release();
    return MSD_OK;
}

extern "C" int msd_ensemble_destroy(msd_ensemble* e) { delete e; return MSD_OK; }
extern "C" int msd_ensemble_set_graph(msd_ensemble* e, int yes) { NONNULL(e); e->use_graph = yes != 0; return MSD_OK; }
extern "C" int msd_ensemble_set_compensated(msd_ensemble* e, int yes)
{
    NONNULL(e);
    e->compensated = yes != 0;
    e->drop_graph();
    return MSD_OK;
}

// Event-timed passes of an ensemble step, for nsteps steps of progress (direct
// launches, no graph): the LongRange3D members' y, z and inverse y passes and row
// passes, and the film groups' column and row passes, each pass group between two
// events on the library stream.  Returns pass name -> average ms per step's launch.
extern "C" int msd_ensemble_profile(msd_ensemble* e, int nsteps, int cap, char names[][32], double* avg_ms, int* count)
{
    NONNULL(e); NONNULL(names); NONNULL(avg_ms); NONNULL(count);
    if (nsteps <= 0) return fail(MSD_EINVAL, "ensemble profile: nsteps %d", nsteps);
    if (int rc = e->refresh()) return rc;
    bool stale = false;
    for (msd_stepper* s : e->m) {
        if (int rc = member_begin(s)) return rc;
        stale |= s->lr3 && spectrum_stale(s);
        if (s->kind() == SK_LAYERED && spectrum_stale(s)) {
            hipError_t err = launch_rows_fwd(s->args.dp, s->s->S, G.stream);
            if (err != hipSuccess) return fail(MSD_EDEVICE, "ensemble profile: %s", hipGetErrorString(err));
        }
    }
    if (stale) {
        hipError_t err = e->b.rows(R3_FWD, false, G.stream);
        if (err != hipSuccess) return fail(MSD_EDEVICE, "ensemble profile: %s", hipGetErrorString(err));
    }
    struct Pass {
        const char* name;
        std::function<hipError_t(hipStream_t)> run;
    };
    std::vector<Pass> seq;
    for (int half = 0; half < 2; ++half) {
        const bool pred = half == 0;
        if (e->b.any_lr) {
            seq.push_back({"k3_ypass_fwd", [e](hipStream_t st) { return e->b.ypass(-1, st); }});
            seq.push_back({"k3_zconv", [e](hipStream_t st) { return e->b.zpass(st); }});
            seq.push_back({"k3_ypass_inv", [e](hipStream_t st) { return e->b.ypass(1, st); }});
        }
        if (!e->b.rg.empty())
            seq.push_back({pred ? "k3_rows_pred" : "k3_rows_corr",
                           [e, pred](hipStream_t st) { return e->b.rows(pred ? R3_PRED : R3_CORR, e->compensated, st); }});
        if (!e->films.empty())
            seq.push_back({pred ? "films_pred" : "films_corr", [e, pred](hipStream_t st) {
                               hipError_t r = hipSuccess;
                               for (const FilmGroup& f : e->films)
                                   if (r == hipSuccess) r = f.step(pred, st);
                               return r;
                           }});
    }
    const size_t per = seq.size();
    std::vector<hipEvent_t> ev(per * (size_t)nsteps + 1);
    for (auto& x : ev) HIPCHK(hipEventCreate(&x));
    int rc = MSD_OK;
    size_t q = 0;
    HIPCHK(hipEventRecord(ev[q++], G.stream));
    for (int i = 0; i < nsteps && !rc; ++i)
        for (size_t k = 0; k < per && !rc; ++k) {
            hipError_t err = seq[k].run(G.stream);
            if (err == hipSuccess) err = hipEventRecord(ev[q++], G.stream);
            if (err != hipSuccess) rc = fail(MSD_EDEVICE, "ensemble profile: %s", hipGetErrorString(err));
        }
    if (!rc) {
        HIPCHK(hipStreamSynchronize(G.stream));
        std::vector<std::string> nm;
        std::vector<double> tot;
        std::vector<int> cnt;
        for (size_t j = 1; j < q; ++j) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[j - 1], ev[j]));
            const std::string name = seq[(j - 1) % per].name;
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
        for (msd_stepper* s : e->m) stepper_end(s, nsteps);
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    return rc;
}

extern "C" int msd_ensemble_heun(msd_ensemble* e, int nsteps)
{
    NONNULL(e);
    if (nsteps < 0) return fail(MSD_EINVAL, "ensemble heun: nsteps %d", nsteps);
    if (nsteps == 0) return MSD_OK;
    if (int rc = e->refresh()) return rc;
    bool stale = false;
    for (msd_stepper* s : e->m) {
        if (int rc = member_begin(s)) return rc;
        stale |= s->lr3 && spectrum_stale(s);
        // a film's first column pass reads the spin spectrum of its current state
        if (s->kind() == SK_LAYERED && spectrum_stale(s)) {
            hipError_t err = launch_rows_fwd(s->args.dp, s->s->S, G.stream);
            if (err != hipSuccess) return fail(MSD_EDEVICE, "ensemble: %s", hipGetErrorString(err));
        }
    }
    if (stale) {
        hipError_t err = e->b.rows(R3_FWD, false, G.stream);
        if (err != hipSuccess) return fail(MSD_EDEVICE, "ensemble: %s", hipGetErrorString(err));
    }
    auto one_step = [&](hipStream_t st) -> hipError_t {
        hipError_t err = e->evaluate(true, st);
        if (err == hipSuccess) err = e->evaluate(false, st);
        return err;
    };
    if (e->use_graph) {
        // a graph of `reps` steps, captured once and kept until the members change
        auto ready = [&](hipGraph_t* graph, hipGraphExec_t* exec, int reps) -> int {
            if (*exec) return MSD_OK;
            HIPCHK(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
            hipError_t err = hipSuccess;
            for (int q = 0; q < reps && err == hipSuccess; ++q) err = one_step(G.stream);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(G.stream, &g);
            if (err != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(MSD_EDEVICE, "ensemble graph: %s", hipGetErrorString(err != hipSuccess ? err : e2));
            }
            *graph = g;
            HIPCHK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
            return MSD_OK;
        };
        // runs of at least STEP_SPAN steps replay the span graph, the rest one step each
        if (nsteps >= STEP_SPAN) {
            if (int rc = ready(&e->mgraph, &e->mexec, STEP_SPAN)) return rc;
            for (int i = 0; i < nsteps / STEP_SPAN; ++i) HIPCHK(hipGraphLaunch(e->mexec, G.stream));
        }
        const int single = nsteps >= STEP_SPAN ? nsteps % STEP_SPAN : nsteps;
        if (single > 0) {
            if (int rc = ready(&e->graph, &e->exec, 1)) return rc;
            for (int i = 0; i < single; ++i) HIPCHK(hipGraphLaunch(e->exec, G.stream));
        }
    } else {
        for (int i = 0; i < nsteps; ++i) {
            hipError_t err = one_step(G.stream);
            if (err != hipSuccess) return fail(MSD_EDEVICE, "ensemble step: %s", hipGetErrorString(err));
        }
    }
    for (msd_stepper* s : e->m) stepper_end(s, nsteps);
    return MSD_OK;
}

// Autotuner of an ensemble's LongRange3D passes, timed in the ensemble's own launches
// (a narrow block that fills the GPU for one replica can lose to a wide one when the
// replicas share a launch).  Members are grouped by lattice (size and boundaries);
// one sweep over the groups tries every compiled (y, z) width pair on a group's
// members with the other groups' widths held, keeping the fastest by `steps`
// graph-replayed steps.  Every member's state is restored after each trial.
extern "C" int msd_ensemble_autotune_passes(msd_ensemble* e, int steps, double* best_ms)
{
    NONNULL(e); NONNULL(best_ms);
    if (steps < 1) return fail(MSD_EINVAL, "ensemble autotune: steps %d", steps);
    *best_ms = -1.0;
    struct Lattice {
        Dims d;
        bool open;
        std::vector<msd_longrange3d*> lr;
    };
    std::vector<Lattice> groups;
    for (msd_stepper* s : e->m) {
        if (!s->lr3) continue;
        size_t gi = 0;
        while (gi < groups.size() && !(same_dims(groups[gi].d, s->lr3->d.nx, s->lr3->d.ny, s->lr3->d.nz) &&
                                       groups[gi].open == s->lr3->open)) ++gi;
        if (gi == groups.size()) groups.push_back(Lattice{s->lr3->d, s->lr3->open, {}});
        groups[gi].lr.push_back(s->lr3);
    }
    if (groups.empty()) return MSD_OK;
    std::vector<std::unique_ptr<StateKeep>> keep;
    for (msd_stepper* s : e->m) {
        keep.emplace_back(new StateKeep(s));
        if (int rc = keep.back()->save()) return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        return fail(MSD_EDEVICE, "ensemble autotune: events");
    }
    auto time_once = [&](double* per) -> int {
        if (int rc = msd_ensemble_heun(e, 2)) return rc;          // capture + warm
        if (hipEventRecord(e0, G.stream) != hipSuccess) return fail(MSD_EDEVICE, "ensemble autotune: event");
        if (int rc = msd_ensemble_heun(e, steps)) return rc;
        if (hipEventRecord(e1, G.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
            return fail(MSD_EDEVICE, "ensemble autotune: event");
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *per = ms / steps;
        for (auto& k : keep)
            if (int rc = k->restore()) return rc;
        return MSD_OK;
    };
    auto set_group = [](const Lattice& g, int wy, int wz) -> int {
        for (msd_longrange3d* l : g.lr)
            if (int rc = lr3_set_widths(l, wy, wz)) return rc;
        return MSD_OK;
    };
    double best = 0.0;
    int rc = time_once(&best);
    for (const Lattice& g : groups) {
        if (rc) break;
        msd_longrange3d* l0 = g.lr[0];
        int wys[8], wzs[8];
        int ny = l0->yt ? ypass_widths(l0->yt, wys, 8) : 0;
        int nz = l0->zt ? zconv_widths(l0->zt, wzs, 8) : 0;
        if (ny == 0) { wys[0] = 0; ny = 1; }
        if (nz == 0) { wzs[0] = 0; nz = 1; }
        int bwy = l0->wy, bwz = l0->wz;
        for (int c = 0; c < ny * nz && !rc; ++c) {
            const int wy = wys[c / nz], wz = wzs[c % nz];
            if ((wy == 0 || wy == bwy) && wz == bwz) continue;     // the current choice is timed
            if (set_group(g, wy, wz) != MSD_OK) continue;          // over the LDS: not a candidate
            double per = 0.0;
            if ((rc = time_once(&per))) break;
            if (per < best) { best = per; bwy = l0->wy; bwz = l0->wz; }
        }
        if (!rc) rc = set_group(g, bwy, bwz);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc) {
        *best_ms = best;
        rc = e->refresh();
    }
    return rc;
}
