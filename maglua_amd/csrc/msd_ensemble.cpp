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
        if (!e->exec) {
            HIPCHK(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
            hipError_t err = one_step(G.stream);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(G.stream, &g);
            if (err != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(MSD_EDEVICE, "ensemble graph: %s", hipGetErrorString(err != hipSuccess ? err : e2));
            }
            e->graph = g;
            HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        }
        for (int i = 0; i < nsteps; ++i) HIPCHK(hipGraphLaunch(e->exec, G.stream));
    } else {
        for (int i = 0; i < nsteps; ++i) {
            hipError_t err = one_step(G.stream);
            if (err != hipSuccess) return fail(MSD_EDEVICE, "ensemble step: %s", hipGetErrorString(err));
        }
    }
    for (msd_stepper* s : e->m) stepper_end(s, nsteps);
    return MSD_OK;
}
