// Field operators and LLG of libmsd.so: Exchange, Anisotropy, Anisotropy.Uniaxial,
// AppliedField, Thermal and LLG.Cartesian / LLG.Quaternion on the operator API.
#include "msd_host.h"

// CSR by target in path insertion order, plus the class table when the
// couplings reduce to at most 256 distinct per-site stencils.
int msd_exchange::compile()
{
    if (!dirty) return MSD_OK;
    release();
    const int n = d.n;
    const size_t np = paths.size();
    std::vector<int> cnt(n + 1, 0);
    for (const Path& p : paths) cnt[p.to + 1]++;
    for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> hcol(np);
    std::vector<float> hval(np);
    std::vector<int> fillp(cnt.begin(), cnt.end() - 1);
    for (const Path& p : paths) {
        const int k = fillp[p.to]++;
        hcol[k] = p.from;
        hval[k] = (float)(p.J * scale);
    }
    if (int rc = dalloc(&row, n + 1)) return rc;
    if (int rc = dalloc(&col, np)) return rc;
    if (int rc = dalloc(&val, np)) return rc;
    HIPCHK(hipMemcpyAsync(row, cnt.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, G.stream));
    if (np) {
        HIPCHK(hipMemcpyAsync(col, hcol.data(), sizeof(int) * np, hipMemcpyHostToDevice, G.stream));
        HIPCHK(hipMemcpyAsync(val, hval.data(), sizeof(float) * np, hipMemcpyHostToDevice, G.stream));
    }
    // class table
    lut_ok = true;
    std::map<std::string, int> ids;
    std::vector<uint8_t> hcls(n);
    std::vector<int> starts(1, 0);
    std::vector<ExEntry> ents;
    std::string key;
    for (int t = 0; t < n && lut_ok; ++t) {
        key.clear();
        std::vector<ExEntry> mine;
        for (int k = cnt[t]; k < cnt[t + 1]; ++k) {
            ExEntry e;
            e.dl = hcol[k] - t;
            e.J = hval[k];
            mine.push_back(e);
            key.append(reinterpret_cast<const char*>(&e), sizeof e);
        }
        auto it = ids.find(key);
        int id;
        if (it == ids.end()) {
            id = (int)ids.size();
            if (id > 255) { lut_ok = false; break; }
            ids.emplace(key, id);
            ents.insert(ents.end(), mine.begin(), mine.end());
            starts.push_back((int)ents.size());
        } else {
            id = it->second;
        }
        hcls[t] = (uint8_t)id;
    }
    if (lut_ok) {
        ncls = (int)ids.size();
        nent = (int)ents.size();
        if (ids.size() > 1) {
            if (int rc = dalloc(&cls, n)) return rc;
            HIPCHK(hipMemcpyAsync(cls, hcls.data(), n, hipMemcpyHostToDevice, G.stream));
        }
        if (int rc = dalloc(&cls_start, starts.size())) return rc;
        if (int rc = dalloc(&ent, std::max<size_t>(ents.size(), 1))) return rc;
        HIPCHK(hipMemcpyAsync(cls_start, starts.data(), sizeof(int) * starts.size(), hipMemcpyHostToDevice, G.stream));
        if (!ents.empty())
            HIPCHK(hipMemcpyAsync(ent, ents.data(), sizeof(ExEntry) * ents.size(), hipMemcpyHostToDevice, G.stream));
    }
    HIPCHK(hipStreamSynchronize(G.stream));
    dirty = false;
    return MSD_OK;
}

extern "C" int msd_exchange_create(int nx, int ny, int nz, msd_exchange** out)
{
    NONNULL(out);
    if (nx < 1 || ny < 1 || nz < 1) return fail(MSD_EINVAL, "Exchange.new: bad size");
    if (int rc = ensure_init()) return rc;
    msd_exchange* e = new msd_exchange;
    e->d = Dims{nx, ny, nz, nx * ny * nz};
    *out = e;
    return MSD_OK;
}

extern "C" int msd_exchange_destroy(msd_exchange* e) { delete e; return MSD_OK; }
extern "C" int msd_exchange_set_pbc(msd_exchange* e, int px, int py, int pz)
{
    NONNULL(e);
    e->pbc[0] = px != 0; e->pbc[1] = py != 0; e->pbc[2] = pz != 0;
    return MSD_OK;
}
extern "C" int msd_exchange_set_normalize(msd_exchange* e, int yes) { NONNULL(e); e->normalize = yes != 0; e->dirty = true; return MSD_OK; }
extern "C" int msd_exchange_set_scale(msd_exchange* e, double s) { NONNULL(e); e->scale = s; e->dirty = true; return MSD_OK; }

extern "C" int msd_exchange_add_path(msd_exchange* e, int from, int to, double J)
{
    NONNULL(e);
    if (from < 0 || from >= e->d.n || to < 0 || to >= e->d.n)
        return fail(MSD_EINVAL, "addPath: site outside 0..%d", e->d.n - 1);
    if (J != 0) { e->paths.push_back({from, to, J}); e->dirty = true; }
    return MSD_OK;
}

extern "C" int msd_exchange_add(msd_exchange* e, const int a[3], const int b[3], double J, int* added)
{
    NONNULL(e); NONNULL(a); NONNULL(b);
    int p[3] = {a[0], a[1], a[2]}, q[3] = {b[0], b[1], b[2]};
    const int n[3] = {e->d.nx, e->d.ny, e->d.nz};
    if (added) *added = 0;
    for (int i = 0; i < 3; ++i) {
        if (e->pbc[i]) {
            p[i] = ((p[i] % n[i]) + n[i]) % n[i];
            q[i] = ((q[i] % n[i]) + n[i]) % n[i];
        }
        if (p[i] < 0 || p[i] >= n[i] || q[i] < 0 || q[i] >= n[i]) return MSD_OK;
    }
    const int f = p[0] + n[0] * (p[1] + n[1] * p[2]);
    const int t = q[0] + n[0] * (q[1] + n[1] * q[2]);
    if (added) *added = 1;
    return msd_exchange_add_path(e, f, t, J);
}

extern "C" int msd_exchange_add_paths(msd_exchange* e, const int* from, const int* to, const double* J, int count)
{
    NONNULL(e);
    if (count < 0) return fail(MSD_EINVAL, "addPaths: count %d", count);
    if (count && (!from || !to || !J)) return fail(MSD_EINVAL, "addPaths: NULL array");
    for (int i = 0; i < count; ++i)
        if (from[i] < 0 || from[i] >= e->d.n || to[i] < 0 || to[i] >= e->d.n)
            return fail(MSD_EINVAL, "addPaths: path %d has a site outside 0..%d", i, e->d.n - 1);
    e->paths.reserve(e->paths.size() + count);
    for (int i = 0; i < count; ++i)
        if (J[i] != 0) e->paths.push_back({from[i], to[i], J[i]});
    e->dirty = true;
    return MSD_OK;
}

// Exchange::encode (cpu/exchange/spinoperationexchange.cpp:52-72): version 1, the
// normalisation flag, pbc and the pathways in the order they were added
extern "C" int msd_exchange_encode(msd_exchange* e, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects)
{
    NONNULL(e); NONNULL(len);
    if (first_index < 0) return fail(MSD_EINVAL, "Exchange.encode: object index %d < 0", first_index);
    WireOut w;
    w.b.reserve(64 + 16 * e->paths.size());
    wire_operator(w, first_index, e->d, e->scale);
    w.ch(1);
    w.i32(e->normalize ? 1 : 0);
    for (int k = 0; k < 3; ++k) w.i32(e->pbc[k]);
    w.i32((int32_t)e->paths.size());
    for (const auto& p : e->paths) { w.i32(p.from); w.i32(p.to); w.f64(p.J); }
    if (n_objects) *n_objects = 1;
    return wire_finish(w, "Exchange.encode", buf, cap, len);
}

// Exchange::decode (spinoperationexchange.cpp:74-113): a version-0 stream has no
// normalisation flag and means unnormalised exchange
extern "C" int msd_exchange_decode(const void* buf, int64_t len, int first_index, msd_exchange** out, int64_t* used, int* n_objects)
{
    NONNULL(buf); NONNULL(out);
    *out = nullptr;
    if (len < 0) return fail(MSD_EINVAL, "Exchange.decode: length %lld", (long long)len);
    WireIn r{(const uint8_t*)buf, (size_t)len};
    int nx, ny, nz;
    double scale;
    if (int rc = unwire_operator(r, "Exchange.decode", first_index, &nx, &ny, &nz, &scale)) return rc;
    char version;
    int32_t norm = 0, pbc[3], num;
    if (!r.ch(&version)) return fail(MSD_EINVAL, "Exchange.decode: stream too short");
    if (version < 0) return fail(MSD_EINVAL, "Exchange.decode: unknown version %d", (int)version);
    if (version >= 1 && !r.i32(&norm)) return fail(MSD_EINVAL, "Exchange.decode: stream too short");
    if (!r.i32(&pbc[0]) || !r.i32(&pbc[1]) || !r.i32(&pbc[2]) || !r.i32(&num))
        return fail(MSD_EINVAL, "Exchange.decode: stream ends inside the header");
    if (num < 0 || (size_t)num > (r.n - r.at) / 16) return fail(MSD_EINVAL, "Exchange.decode: %d pathways do not fit the stream", num);
    msd_exchange* e = nullptr;
    if (int rc = msd_exchange_create(nx, ny, nz, &e)) return rc;
    std::unique_ptr<msd_exchange> own(e);
    e->scale = scale;
    e->normalize = norm != 0;
    for (int k = 0; k < 3; ++k) e->pbc[k] = pbc[k];
    e->paths.reserve(num);
    for (int i = 0; i < num; ++i) {
        int32_t f, t;
        double J;
        r.i32(&f); r.i32(&t); r.f64(&J);
        if (f < 0 || f >= e->d.n || t < 0 || t >= e->d.n)
            return fail(MSD_EINVAL, "Exchange.decode: pathway %d has a site outside 0..%d", i, e->d.n - 1);
        e->paths.push_back({f, t, J});
    }
    e->dirty = true;
    if (used) *used = (int64_t)r.at;
    if (n_objects) *n_objects = 1;
    *out = own.release();
    return MSD_OK;
}

extern "C" int msd_exchange_num_paths(const msd_exchange* e, int* n) { NONNULL(e); NONNULL(n); *n = (int)e->paths.size(); return MSD_OK; }

extern "C" int msd_exchange_apply(msd_exchange* e, msd_system* s)
{
    NONNULL(e); NONNULL(s);
    if (!same_dims(s->d, e->d.nx, e->d.ny, e->d.nz)) return fail(MSD_EMISMATCH, "Exchange:apply: size mismatch");
    if (int rc = e->compile()) return rc;
    int k;
    if (int rc = s->mark("Exchange", &k)) return rc;
    if (e->paths.empty()) return MSD_OK;
    hipError_t err = launch_exchange_op(s->d, e->csr_dev(), s->S, s->slots[k].h, G.stream);
    if (err != hipSuccess) return fail(MSD_EDEVICE, "Exchange:apply: %s", hipGetErrorString(err));
    return MSD_OK;
}


int msd_anisotropy::compile()
{
    if (!dirty) return MSD_OK;
    (void)hipStreamSynchronize(G.stream);
    dfree(axk); dfree(mat);
    dev = AnisoDev{};
    const int n = d.n;
    if (ops.empty()) { dirty = false; return MSD_OK; }
    std::vector<int> per(n, 0);
    for (const Op& o : ops) per[o.site]++;
    bool single = true, full = true;
    for (int i = 0; i < n; ++i) { single &= per[i] <= 1; full &= per[i] == 1; }
    bool uniform = full;
    for (size_t k = 1; uniform && k < ops.size(); ++k)
        uniform = ops[k].ax == ops[0].ax && ops[k].ay == ops[0].ay && ops[k].az == ops[0].az && ops[k].K == ops[0].K;
    if (uniform) {
        const Op& o = ops[0];
        const double K = o.K * scale;
        const double m[6] = {K * o.ax * o.ax, K * o.ax * o.ay, K * o.ax * o.az, K * o.ay * o.ay, K * o.ay * o.az, K * o.az * o.az};
        dev.mode = 1;
        for (int k = 0; k < 6; ++k) dev.u[k] = (float)m[k];
    } else if (single) {
        std::vector<float4> h(n, make_float4(0.f, 0.f, 1.f, 0.f));
        for (const Op& o : ops) h[o.site] = make_float4((float)o.ax, (float)o.ay, (float)o.az, (float)(o.K * scale));
        if (int rc = dalloc(&axk, n)) return rc;
        HIPCHK(hipMemcpyAsync(axk, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice, G.stream));
        HIPCHK(hipStreamSynchronize(G.stream));
        dev.mode = 2;
        dev.axk = axk;
    } else {
        std::vector<double> h((size_t)6 * n, 0.0);
        for (const Op& o : ops) {
            const double K = o.K * scale;
            const double m[6] = {K * o.ax * o.ax, K * o.ax * o.ay, K * o.ax * o.az, K * o.ay * o.ay, K * o.ay * o.az, K * o.az * o.az};
            for (int k = 0; k < 6; ++k) h[(size_t)k * n + o.site] += m[k];
        }
        if (int rc = dalloc(&mat, (size_t)6 * n)) return rc;
        if (int rc = upload_f32(mat, h.data(), (size_t)6 * n)) return rc;
        dev.mode = 3;
        dev.mat = mat;
    }
    dirty = false;
    return MSD_OK;
}

extern "C" int msd_anisotropy_create(int nx, int ny, int nz, msd_anisotropy** out)
{
    NONNULL(out);
    if (nx < 1 || ny < 1 || nz < 1) return fail(MSD_EINVAL, "Anisotropy.new: bad size");
    if (int rc = ensure_init()) return rc;
    msd_anisotropy* a = new msd_anisotropy;
    a->d = Dims{nx, ny, nz, nx * ny * nz};
    *out = a;
    return MSD_OK;
}
extern "C" int msd_anisotropy_destroy(msd_anisotropy* a) { delete a; return MSD_OK; }
extern "C" int msd_anisotropy_set_scale(msd_anisotropy* a, double s) { NONNULL(a); a->scale = s; a->dirty = true; return MSD_OK; }

extern "C" int msd_anisotropy_add(msd_anisotropy* a, int site, const double axis[3], double K, int* added)
{
    NONNULL(a); NONNULL(axis);
    if (added) *added = 0;
    if (site < 0 || site >= a->d.n) return fail(MSD_EINVAL, "Anisotropy:add: site %d outside 0..%d", site, a->d.n - 1);
    const double len = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (!(len > 0)) return MSD_OK;
    a->ops.push_back({site, axis[0] / len, axis[1] / len, axis[2] / len, K});
    a->dirty = true;
    if (added) *added = 1;
    return MSD_OK;
}

extern "C" int msd_anisotropy_add_many(msd_anisotropy* a, const int* site, const double* axes, const double* K, int count)
{
    NONNULL(a);
    if (count < 0) return fail(MSD_EINVAL, "Anisotropy:add: count %d", count);
    if (count && (!site || !axes || !K)) return fail(MSD_EINVAL, "Anisotropy:add: NULL array");
    for (int i = 0; i < count; ++i) {
        int added;
        if (int rc = msd_anisotropy_add(a, site[i], axes + 3 * (size_t)i, K[i], &added)) return rc;
    }
    return MSD_OK;
}

// Anisotropy::encode (cpu/anisotropy/spinoperationanisotropy.cpp:191-208): the ops in the
// order they were added, each site, unit axis and strength
extern "C" int msd_anisotropy_encode(msd_anisotropy* a, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects)
{
    NONNULL(a); NONNULL(len);
    if (first_index < 0) return fail(MSD_EINVAL, "Anisotropy.encode: object index %d < 0", first_index);
    WireOut w;
    w.b.reserve(64 + 36 * a->ops.size());
    wire_operator(w, first_index, a->d, a->scale);
    w.ch(0);
    w.i32((int32_t)a->ops.size());
    for (const auto& o : a->ops) { w.i32(o.site); w.f64(o.ax); w.f64(o.ay); w.f64(o.az); w.f64(o.K); }
    if (n_objects) *n_objects = 1;
    return wire_finish(w, "Anisotropy.encode", buf, cap, len);
}

// Anisotropy::decode (spinoperationanisotropy.cpp:210-240): each op goes through add,
// which normalises the axis again as addAnisotropy does
extern "C" int msd_anisotropy_decode(const void* buf, int64_t len, int first_index, msd_anisotropy** out, int64_t* used, int* n_objects)
{
    NONNULL(buf); NONNULL(out);
    *out = nullptr;
    if (len < 0) return fail(MSD_EINVAL, "Anisotropy.decode: length %lld", (long long)len);
    WireIn r{(const uint8_t*)buf, (size_t)len};
    int nx, ny, nz;
    double scale;
    if (int rc = unwire_operator(r, "Anisotropy.decode", first_index, &nx, &ny, &nz, &scale)) return rc;
    char version;
    int32_t num;
    if (!r.ch(&version) || !r.i32(&num)) return fail(MSD_EINVAL, "Anisotropy.decode: stream too short");
    if (version != 0) return fail(MSD_EINVAL, "Anisotropy.decode: unknown version %d", (int)version);
    if (num < 0 || (size_t)num > (r.n - r.at) / 36) return fail(MSD_EINVAL, "Anisotropy.decode: %d ops do not fit the stream", num);
    msd_anisotropy* a = nullptr;
    if (int rc = msd_anisotropy_create(nx, ny, nz, &a)) return rc;
    std::unique_ptr<msd_anisotropy> own(a);
    a->scale = scale;
    for (int i = 0; i < num; ++i) {
        int32_t site;
        double ax[3], K;
        r.i32(&site); r.f64(&ax[0]); r.f64(&ax[1]); r.f64(&ax[2]); r.f64(&K);
        // stored as read, as addAnisotropy does (:172-189): only the Lua add normalises
        if (site < 0 || site >= a->d.n) return fail(MSD_EINVAL, "Anisotropy.decode: op %d names site %d", i, site);
        a->ops.push_back({site, ax[0], ax[1], ax[2], K});
        a->dirty = true;
    }
    if (used) *used = (int64_t)r.at;
    if (n_objects) *n_objects = 1;
    *out = own.release();
    return MSD_OK;
}

extern "C" int msd_anisotropy_apply(msd_anisotropy* a, msd_system* s)
{
    NONNULL(a); NONNULL(s);
    if (!same_dims(s->d, a->d.nx, a->d.ny, a->d.nz)) return fail(MSD_EMISMATCH, "Anisotropy:apply: size mismatch");
    if (int rc = a->compile()) return rc;
    int k;
    if (int rc = s->mark("Anisotropy", &k)) return rc;
    hipError_t e = launch_anisotropy_op(s->d, a->dev, s->S, s->slots[k].h, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Anisotropy:apply: %s", hipGetErrorString(e));
    return MSD_OK;
}


extern "C" int msd_uniaxial_create(int nx, int ny, int nz, msd_uniaxial** out)
{
    NONNULL(out);
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1) return fail(MSD_EINVAL, "Anisotropy.Uniaxial.new: bad size");
    if (int rc = ensure_init()) return rc;
    msd_uniaxial* u = new msd_uniaxial;
    u->d = Dims{nx, ny, nz, nx * ny * nz};
    *out = u;
    return MSD_OK;
}
extern "C" int msd_uniaxial_destroy(msd_uniaxial* u) { delete u; return MSD_OK; }
extern "C" int msd_uniaxial_set_scale(msd_uniaxial* u, double s) { NONNULL(u); u->scale = s; return MSD_OK; }
extern "C" int msd_uniaxial_add(msd_uniaxial* u, int site, const double* axis, double K1, double K2, int* added)
{
    NONNULL(u); NONNULL(axis);
    if (site < 0 || site >= u->d.n) return fail(MSD_EINVAL, "Anisotropy.Uniaxial:add: site %d is not part of the operator", site);
    const double len = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    if (added) *added = 0;
    if (!(len > 0)) return MSD_OK;    // l_add: a zero axis adds nothing
    u->ops.push_back(msd_uniaxial::op{site, {axis[0] / len, axis[1] / len, axis[2] / len}, K1, K2});
    u->dirty = true;
    if (added) *added = 1;
    return MSD_OK;
}
extern "C" int msd_uniaxial_num_axes(const msd_uniaxial* u, int* n)
{
    NONNULL(u); NONNULL(n);
    *n = (int)u->ops.size();
    return MSD_OK;
}
/* i-th op (0-based, order of addition): site, axis[3], K1, K2 */
extern "C" int msd_uniaxial_axis_at(const msd_uniaxial* u, int i, int* site, double* axis, double* K1, double* K2)
{
    NONNULL(u); NONNULL(site); NONNULL(axis); NONNULL(K1); NONNULL(K2);
    if (i < 0 || i >= (int)u->ops.size()) return fail(MSD_EINVAL, "Anisotropy.Uniaxial:axisAt: index %d out of range", i + 1);
    const msd_uniaxial::op& o = u->ops[i];
    *site = o.site;
    for (int c = 0; c < 3; ++c) axis[c] = o.n[c];
    *K1 = o.K1; *K2 = o.K2;
    return MSD_OK;
}
extern "C" int msd_uniaxial_apply(msd_uniaxial* u, msd_system* s)
{
    NONNULL(u); NONNULL(s);
    if (!same_dims(s->d, u->d.nx, u->d.ny, u->d.nz)) return fail(MSD_EMISMATCH, "Anisotropy.Uniaxial:apply: size mismatch");
    if (int rc = u->compile()) return rc;
    int k;
    if (int rc = s->mark("UniaxialAnisotropy", &k)) return rc;
    hipError_t e = launch_uniaxial_op(s->d, u->start, u->dops, (float)u->scale, s->S, s->slots[k].h, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Anisotropy.Uniaxial:apply: %s", hipGetErrorString(e));
    return MSD_OK;
}


extern "C" int msd_appliedfield_create(int nx, int ny, int nz, int heterogeneous, msd_appliedfield** out)
{
    NONNULL(out);
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1) return fail(MSD_EINVAL, "AppliedField.new: bad size");
    if (int rc = ensure_init()) return rc;
    std::unique_ptr<msd_appliedfield> a(new msd_appliedfield);
    a->d = Dims{nx, ny, nz, nx * ny * nz};
    a->het = heterogeneous != 0;
    if (a->het) {
        a->h.assign((size_t)3 * a->d.n, 0.0f);
        a->dirty = true;
    }
    *out = a.release();
    return MSD_OK;
}
extern "C" int msd_appliedfield_destroy(msd_appliedfield* a) { delete a; return MSD_OK; }
extern "C" int msd_appliedfield_set(msd_appliedfield* a, double x, double y, double z)
{
    NONNULL(a);
    if (a->het) return fail(MSD_EINVAL, "AppliedField.Heterogeneous: the field is set per site");
    a->B[0] = x; a->B[1] = y; a->B[2] = z;
    return MSD_OK;
}
extern "C" int msd_appliedfield_get(const msd_appliedfield* a, double* xyz)
{
    NONNULL(a); NONNULL(xyz);
    for (int c = 0; c < 3; ++c) xyz[c] = a->B[c];
    return MSD_OK;
}
extern "C" int msd_appliedfield_set_scale(msd_appliedfield* a, double s) { NONNULL(a); a->scale = s; return MSD_OK; }
extern "C" int msd_appliedfield_set_site(msd_appliedfield* a, int site, double x, double y, double z)
{
    NONNULL(a);
    if (!a->het) return fail(MSD_EINVAL, "AppliedField: the field is uniform (AppliedField.Heterogeneous sets sites)");
    if (site < 0 || site >= a->d.n) return fail(MSD_EINVAL, "AppliedField.Heterogeneous: site %d outside the lattice", site);
    const size_t n = a->d.n;
    a->h[site] = (float)x; a->h[n + site] = (float)y; a->h[2 * n + site] = (float)z;
    a->dirty = true;
    return MSD_OK;
}
extern "C" int msd_appliedfield_get_site(const msd_appliedfield* a, int site, double* xyz)
{
    NONNULL(a); NONNULL(xyz);
    if (!a->het) return fail(MSD_EINVAL, "AppliedField: the field is uniform");
    if (site < 0 || site >= a->d.n) return fail(MSD_EINVAL, "AppliedField.Heterogeneous: site %d outside the lattice", site);
    for (int c = 0; c < 3; ++c) xyz[c] = a->h[(size_t)c * a->d.n + site];
    return MSD_OK;
}
extern "C" int msd_appliedfield_apply(msd_appliedfield* a, msd_system* s)
{
    NONNULL(a); NONNULL(s);
    if (!same_dims(s->d, a->d.nx, a->d.ny, a->d.nz)) return fail(MSD_EMISMATCH, "AppliedField:apply: size mismatch");
    const size_t n = s->d.n;
    int k;
    hipError_t e = hipSuccess;
    if (a->het) {
        if (int rc = a->upload()) return rc;
        if (int rc = s->mark("Zeeman_Heterogeneous", &k)) return rc;
        e = launch_axpy(s->slots[k].h, (float)a->scale, a->dh, 3 * n, G.stream);
    } else {
        if (int rc = s->mark("Zeeman", &k)) return rc;
        for (int c = 0; c < 3 && e == hipSuccess; ++c)
            e = launch_fill(s->slots[k].h + c * n, (float)(a->B[c] * a->scale), n, G.stream);
    }
    if (e != hipSuccess) return fail(MSD_EDEVICE, "AppliedField:apply: %s", hipGetErrorString(e));
    return MSD_OK;
}

// =================================================================== Thermal
// A seed from /dev/urandom, else from the clock: what an unseeded (or seed 0) MagLua RNG
// takes (RNG::seed, common/random/random.cpp:29-62).  Never 0.
static uint64_t urandom_seed()
{
    uint64_t v = 0;
    if (FILE* f = fopen("/dev/urandom", "rb")) {
        if (fread(&v, sizeof v, 1, f) != 1) v = 0;
        fclose(f);
    }
    if (!v) v = (uint64_t)time(nullptr) * 0x9E3779B97F4A7C15ull;
    return v ? v : 1;
}


extern "C" int msd_thermal_create(int nx, int ny, int nz, msd_thermal** out)
{
    NONNULL(out);
    if (nx < 1 || ny < 1 || nz < 1) return fail(MSD_EINVAL, "Thermal.new: bad size");
    if (int rc = ensure_init()) return rc;
    msd_thermal* t = new msd_thermal;
    t->d = Dims{nx, ny, nz, nx * ny * nz};
    t->site.assign(t->d.n, 1.0);
    t->seed = urandom_seed();
    *out = t;
    return MSD_OK;
}
extern "C" int msd_thermal_destroy(msd_thermal* t) { delete t; return MSD_OK; }
extern "C" int msd_thermal_set_temperature(msd_thermal* t, double T) { NONNULL(t); t->T = T; return MSD_OK; }
extern "C" int msd_thermal_set_scale(msd_thermal* t, double s) { NONNULL(t); t->scale = s; return MSD_OK; }
extern "C" int msd_thermal_scale_site(msd_thermal* t, int site, double v)
{
    NONNULL(t);
    if (site < 0 || site >= t->d.n) return fail(MSD_EINVAL, "Thermal:scaleSite: site outside lattice");
    t->site[site] = v;
    t->site_dirty = true;
    return MSD_OK;
}
extern "C" int msd_thermal_set_seed(msd_thermal* t, uint64_t seed, uint64_t draw) { NONNULL(t); t->seed = seed; t->draw = draw; return MSD_OK; }
extern "C" int msd_thermal_get_draw(const msd_thermal* t, uint64_t* draw) { NONNULL(t); NONNULL(draw); *draw = t->draw; return MSD_OK; }
extern "C" int msd_thermal_get_seed(const msd_thermal* t, uint64_t* seed) { NONNULL(t); NONNULL(seed); *seed = t->seed; return MSD_OK; }


// AppliedField::encode (cpu/appliedfield/spinoperationappliedfield.cpp:33-42): the
// operator preamble, version 0 and B.  AppliedField_Heterogeneous::encode
// (spinoperationappliedfield_heterogeneous.cpp:39-48): the preamble, version 0 and the
// three per-site component arrays, each written by its own encode (object preamble, no
// Lua type tag), widened from the FP32 values the build holds.
extern "C" int msd_appliedfield_encode(msd_appliedfield* a, int first_index, void* buf, int64_t cap, int64_t* len,
                                       int* n_objects)
{
    NONNULL(a); NONNULL(len);
    if (first_index < 0) return fail(MSD_EINVAL, "AppliedField.encode: object index %d < 0", first_index);
    WireOut w;
    wire_operator(w, first_index, a->d, a->scale);
    w.ch(0);
    if (!a->het) {
        for (int c = 0; c < 3; ++c) w.f64(a->B[c]);
        if (n_objects) *n_objects = 1;
    } else {
        std::vector<double> v(a->d.n);
        for (int c = 0; c < 3; ++c) {
            for (int i = 0; i < a->d.n; ++i) v[i] = a->h[(size_t)c * a->d.n + i];
            wire_darray(w, first_index + 1 + c, a->d, v.data(), false);
        }
        if (n_objects) *n_objects = 4;
    }
    return wire_finish(w, "AppliedField.encode", buf, cap, len);
}

// AppliedField::decode (:44-58) and AppliedField_Heterogeneous::decode (:50-72); the
// latter reads its arrays with the array's decode from the version char on, so, as with
// Thermal, it cannot read the preambles its encode wrote: this reads what encode writes.
extern "C" int msd_appliedfield_decode(const void* buf, int64_t len, int first_index, int heterogeneous,
                                       msd_appliedfield** out, int64_t* used, int* n_objects)
{
    NONNULL(buf); NONNULL(out);
    *out = nullptr;
    const char* who = heterogeneous ? "AppliedField.Heterogeneous.decode" : "AppliedField.decode";
    if (len < 0) return fail(MSD_EINVAL, "%s: length %lld", who, (long long)len);
    WireIn r{(const uint8_t*)buf, (size_t)len};
    int nx, ny, nz;
    double scale;
    char version;
    if (int rc = unwire_operator(r, who, first_index, &nx, &ny, &nz, &scale)) return rc;
    if (!r.ch(&version)) return fail(MSD_EINVAL, "%s: stream too short", who);
    if (version != 0) return fail(MSD_EINVAL, "%s: unknown version %d", who, (int)version);
    msd_appliedfield* a = nullptr;
    if (int rc = msd_appliedfield_create(nx, ny, nz, heterogeneous, &a)) return rc;
    std::unique_ptr<msd_appliedfield> own(a);
    a->scale = scale;
    if (!heterogeneous) {
        for (int c = 0; c < 3; ++c)
            if (!r.f64(&a->B[c])) return fail(MSD_EINVAL, "%s: stream ends inside the field", who);
    } else {
        std::vector<double> v;
        for (int c = 0; c < 3; ++c) {
            if (int rc = unwire_darray(r, who, first_index + 1 + c, a->d, v, false)) return rc;
            for (int i = 0; i < a->d.n; ++i) a->h[(size_t)c * a->d.n + i] = (float)v[i];
        }
        a->dirty = true;
    }
    if (used) *used = (int64_t)r.at;
    if (n_objects) *n_objects = heterogeneous ? 4 : 1;
    *out = own.release();
    return MSD_OK;
}

// AnisotropyUniaxial::encode (cpu/anisotropy/spinoperationanisotropy_uniaxial.cpp:181-199):
// the operator preamble, version 0, the op count and each op (site, axis, K1, K2) in order
// of addition.
extern "C" int msd_uniaxial_encode(msd_uniaxial* u, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects)
{
    NONNULL(u); NONNULL(len);
    if (first_index < 0) return fail(MSD_EINVAL, "Anisotropy.Uniaxial.encode: object index %d < 0", first_index);
    WireOut w;
    wire_operator(w, first_index, u->d, u->scale);
    w.ch(0);
    w.i32((int32_t)u->ops.size());
    for (const msd_uniaxial::op& o : u->ops) {
        w.i32(o.site);
        for (int c = 0; c < 3; ++c) w.f64(o.n[c]);
        w.f64(o.K1);
        w.f64(o.K2);
    }
    if (n_objects) *n_objects = 1;
    return wire_finish(w, "Anisotropy.Uniaxial.encode", buf, cap, len);
}

// AnisotropyUniaxial::decode (:201-234): each op through addAnisotropy (:158-179), which
// divides the axis by its squared length when that is positive (a unit axis is kept)
extern "C" int msd_uniaxial_decode(const void* buf, int64_t len, int first_index, msd_uniaxial** out, int64_t* used,
                                   int* n_objects)
{
    NONNULL(buf); NONNULL(out);
    *out = nullptr;
    const char* who = "Anisotropy.Uniaxial.decode";
    if (len < 0) return fail(MSD_EINVAL, "%s: length %lld", who, (long long)len);
    WireIn r{(const uint8_t*)buf, (size_t)len};
    int nx, ny, nz;
    double scale;
    char version;
    int32_t num;
    if (int rc = unwire_operator(r, who, first_index, &nx, &ny, &nz, &scale)) return rc;
    if (!r.ch(&version)) return fail(MSD_EINVAL, "%s: stream too short", who);
    if (version != 0) return fail(MSD_EINVAL, "%s: unknown version %d", who, (int)version);
    if (!r.i32(&num)) return fail(MSD_EINVAL, "%s: stream ends before the op count", who);
    if (num < 0 || (size_t)num > (r.n - r.at) / 44) return fail(MSD_EINVAL, "%s: %d ops do not fit the stream", who, num);
    msd_uniaxial* u = nullptr;
    if (int rc = msd_uniaxial_create(nx, ny, nz, &u)) return rc;
    std::unique_ptr<msd_uniaxial> own(u);
    u->scale = scale;
    for (int k = 0; k < num; ++k) {
        int32_t site;
        double a[3], K1, K2;
        if (!r.i32(&site) || !r.f64(&a[0]) || !r.f64(&a[1]) || !r.f64(&a[2]) || !r.f64(&K1) || !r.f64(&K2))
            return fail(MSD_EINVAL, "%s: stream ends inside op %d", who, k);
        if (site < 0 || site >= u->d.n) return fail(MSD_EINVAL, "%s: op %d at site %d outside the lattice", who, k, site);
        const double l2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        if (l2 > 0) for (double& c : a) c /= l2;
        u->ops.push_back(msd_uniaxial::op{site, {a[0], a[1], a[2]}, K1, K2});
    }
    u->dirty = true;
    if (used) *used = (int64_t)r.at;
    if (n_objects) *n_objects = 1;
    *out = own.release();
    return MSD_OK;
}

// Thermal::encode (cpu/thermal/spinoperationthermal.cpp:51-61): the operator preamble,
// version 0, the temperature and the per-site scale array, written by the array's own
// encode (object preamble, no Lua type tag).  The noise stream (seed and draw counter)
// is not part of the reference's stream: a decoded Thermal takes a fresh seed.
extern "C" int msd_thermal_encode(msd_thermal* t, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects)
{
    NONNULL(t); NONNULL(len);
    if (first_index < 0) return fail(MSD_EINVAL, "Thermal.encode: object index %d < 0", first_index);
    WireOut w;
    wire_operator(w, first_index, t->d, t->scale);
    w.ch(0);
    w.f64(t->T);
    std::vector<double> ones;
    const double* site = t->site.data();
    if (t->site.size() != (size_t)t->d.n) {
        ones.assign(t->d.n, 1.0);
        site = ones.data();
    }
    wire_darray(w, first_index + 1, t->d, site, false);
    if (n_objects) *n_objects = 2;
    return wire_finish(w, "Thermal.encode", buf, cap, len);
}

// Thermal::decode (spinoperationthermal.cpp:63-80) reads the scale array with the
// array's decode, which starts at the version char, so it cannot read the object
// preamble its own encode wrote; this decode reads the stream the encode writes.
extern "C" int msd_thermal_decode(const void* buf, int64_t len, int first_index, msd_thermal** out, int64_t* used,
                                  int* n_objects)
{
    NONNULL(buf); NONNULL(out);
    *out = nullptr;
    if (len < 0) return fail(MSD_EINVAL, "Thermal.decode: length %lld", (long long)len);
    WireIn r{(const uint8_t*)buf, (size_t)len};
    int nx, ny, nz;
    double scale, T;
    char version;
    if (int rc = unwire_operator(r, "Thermal.decode", first_index, &nx, &ny, &nz, &scale)) return rc;
    if (!r.ch(&version)) return fail(MSD_EINVAL, "Thermal.decode: stream too short");
    if (version != 0) return fail(MSD_EINVAL, "Thermal.decode: unknown version %d", (int)version);
    if (!r.f64(&T)) return fail(MSD_EINVAL, "Thermal.decode: stream ends before the temperature");
    msd_thermal* t = nullptr;
    if (int rc = msd_thermal_create(nx, ny, nz, &t)) return rc;
    std::unique_ptr<msd_thermal> own(t);
    std::vector<double> site;
    if (int rc = unwire_darray(r, "Thermal.decode", first_index + 1, t->d, site, false)) return rc;
    t->T = T;
    t->scale = scale;
    t->site = site;
    t->site_dirty = true;
    if (used) *used = (int64_t)r.at;
    if (n_objects) *n_objects = 2;
    *out = own.release();
    return MSD_OK;
}

// LLG::encode (cpu/llg/llg.cpp:50-60), the same for LLG.Cartesian and LLG.Quaternion:
// the object preamble, version 1, thermalOnlyFirstTerm, disableRenormalization and the
// precession and damping scales (no SpinOperation part).
extern "C" int msd_llg_encode(msd_llg* l, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects)
{
    NONNULL(l); NONNULL(len);
    if (first_index < 0) return fail(MSD_EINVAL, "LLG.encode: object index %d < 0", first_index);
    WireOut w;
    w.ch(WIRE_NEW);
    w.i32(first_index);
    w.ch(1);
    w.i32(l->thermal_first ? 1 : 0);
    w.i32(l->no_renorm ? 1 : 0);
    w.f64(l->ps);
    w.f64(l->ds);
    if (n_objects) *n_objects = 1;
    return wire_finish(w, "LLG.encode", buf, cap, len);
}

// LLG::decode (llg.cpp:62-88): version 0 has no scales (both 1)
extern "C" int msd_llg_decode(const void* buf, int64_t len, int first_index, int quaternion, msd_llg** out, int64_t* used,
                              int* n_objects)
{
    NONNULL(buf); NONNULL(out);
    *out = nullptr;
    if (len < 0) return fail(MSD_EINVAL, "LLG.decode: length %lld", (long long)len);
    WireIn r{(const uint8_t*)buf, (size_t)len};
    char magic, version;
    int32_t pos, tf, nr;
    if (!r.ch(&magic) || !r.i32(&pos)) return fail(MSD_EINVAL, "LLG.decode: stream too short");
    if (magic == WIRE_OLD) return fail(MSD_EINVAL, "LLG.decode: object %d was decoded earlier in the stream", pos);
    if (magic != WIRE_NEW) return fail(MSD_EINVAL, "LLG.decode: malformed stream (magic %d)", (int)magic);
    if (pos != first_index) return fail(MSD_EINVAL, "LLG.decode: object index %d, expected %d", pos, first_index);
    if (!r.ch(&version)) return fail(MSD_EINVAL, "LLG.decode: stream too short");
    if (version != 0 && version != 1) return fail(MSD_EINVAL, "LLG.decode: unknown version %d", (int)version);
    if (!r.i32(&tf) || !r.i32(&nr)) return fail(MSD_EINVAL, "LLG.decode: stream ends inside the flags");
    double ps = 1.0, ds = 1.0;
    if (version == 1 && (!r.f64(&ps) || !r.f64(&ds))) return fail(MSD_EINVAL, "LLG.decode: stream ends inside the scales");
    msd_llg* l = nullptr;
    if (int rc = quaternion ? msd_llg_create_quaternion(&l) : msd_llg_create(&l)) return rc;
    l->thermal_first = tf != 0;
    l->no_renorm = nr != 0;
    l->ps = ps;
    l->ds = ds;
    if (used) *used = (int64_t)r.at;
    if (n_objects) *n_objects = 1;
    *out = l;
    return MSD_OK;
}

LLGDev msd_system::llgdev(const msd_llg* l) const
{
    LLGDev p{};
    p.thermal_first = l ? l->thermal_first : 1;
    p.renorm = l ? !l->no_renorm : 1;
    p.quaternion = l ? l->quaternion : 0;
    p.ps = l ? (float)l->ps : 1.0f;
    p.ds = l ? (float)l->ds : 1.0f;
    p.alpha = (float)alpha;
    p.gamma = (float)gamma;
    p.site_alpha = site_alpha;
    p.site_gamma = site_gamma;
    return p;
}

extern "C" int msd_thermal_apply(msd_thermal* t, msd_system* s)
{
    NONNULL(t); NONNULL(s);
    if (!same_dims(s->d, t->d.nx, t->d.ny, t->d.nz)) return fail(MSD_EMISMATCH, "Thermal:apply: size mismatch");
    ThermalDev td;
    if (int rc = t->dev(td)) return rc;
    int k;
    if (int rc = s->mark("Thermal", &k)) return rc;
    hipError_t e = launch_thermal_op(s->d, td, s->llgdev(nullptr), (float)s->dt, t->draw, s->S, s->slots[k].h, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Thermal:apply: %s", hipGetErrorString(e));
    t->draw++;
    return MSD_OK;
}

// ======================================================================= LLG
extern "C" int msd_llg_create(msd_llg** out)
{
    NONNULL(out);
    if (int rc = ensure_init()) return rc;
    *out = new msd_llg;
    return MSD_OK;
}
extern "C" int msd_llg_create_quaternion(msd_llg** out)
{
    NONNULL(out);
    if (int rc = ensure_init()) return rc;
    msd_llg* l = new msd_llg;
    l->quaternion = true;
    *out = l;
    return MSD_OK;
}
extern "C" int msd_llg_destroy(msd_llg* l) { delete l; return MSD_OK; }
extern "C" int msd_llg_set_thermal_only_first_term(msd_llg* l, int yes) { NONNULL(l); l->thermal_first = yes != 0; return MSD_OK; }
extern "C" int msd_llg_set_disable_renormalization(msd_llg* l, int yes) { NONNULL(l); l->no_renorm = yes != 0; return MSD_OK; }
extern "C" int msd_llg_set_scales(msd_llg* l, double ps, double ds) { NONNULL(l); l->ps = ps; l->ds = ds; return MSD_OK; }

// LLG::apply argument semantics (cpu/llg/llg.cpp:100-171) over LLGCartesian::apply
extern "C" int msd_llg_apply(msd_llg* l, msd_system* from, double scale, msd_system* dmdt, msd_system* to, int advance)
{
    NONNULL(l); NONNULL(from);
    if (!dmdt) dmdt = from;
    if (!to) to = from;
    if (!same_dims(from->d, dmdt->d.nx, dmdt->d.ny, dmdt->d.nz) || !same_dims(from->d, to->d.nx, to->d.ny, to->d.nz))
        return fail(MSD_EMISMATCH, "LLG:apply: size mismatch");
    const double dt = scale * dmdt->dt;
    if (l->quaternion && dmdt->find("Total") < 0) {
        // nothing to rotate about: only the time moves (llgquat.cpp:131-137)
        if (advance) to->time = from->time + dt;
        return MSD_OK;
    }
    int T;
    if (int rc = dmdt->slot("Total", &T)) return rc;
    const int TH = dmdt->find("Thermal");
    hipError_t e = launch_llg_op(from->d, dmdt->llgdev(l), (float)dt, from->S, dmdt->S, dmdt->slots[T].h,
                                 TH >= 0 ? dmdt->slots[TH].h : nullptr, to->S, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "LLG:apply: %s", hipGetErrorString(e));
    // LLG.Cartesian leaves the time alone for dt = 0 (llgcartesian.cpp:132-137); LLG.Quaternion sets it
    if ((dt != 0 || l->quaternion) && advance) to->time = from->time + dt;
    to->lo_valid = false;
    to->version++;
    return MSD_OK;
}
