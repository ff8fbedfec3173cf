// Layered dipole operators of libmsd.so: Dipole (LongRange with the lattice-sum
// tensor), LongRange2D and its tensor builders (Dipole2D, Magnetostatics2D).
#include "msd_host.h"

namespace msd {
bool make_plan(int L, FFTPlan& p)
{
    p.L = L;
    p.nst = 0;
    int m = L;
    const int order[5] = {8, 4, 2, 3, 5};
    for (int r : order)
        while (m % r == 0 && p.nst < 16) { p.rad[p.nst++] = r; m /= r; }
    return m == 1 && L >= 1;
}
}  // namespace msd

int twiddles(int L, float2** d32, double2** d64)
{
    std::vector<float2> a(L);
    std::vector<double2> b(L);
    for (int k = 0; k < L; ++k) {
        const double ang = -2.0 * M_PI * (double)k / (double)L;
        b[k] = make_double2(std::cos(ang), std::sin(ang));
        a[k] = make_float2((float)b[k].x, (float)b[k].y);
    }
    if (int rc = dalloc(d32, L)) return rc;
    if (int rc = dalloc(d64, L)) return rc;
    HIPCHK(hipMemcpyAsync(*d32, a.data(), sizeof(float2) * L, hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipMemcpyAsync(*d64, b.data(), sizeof(double2) * L, hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    return MSD_OK;
}

// Row tiles of ty rows and tpb threads, and the column form that goes with them.
// The split y transform (DipoleDev::split) is taken when `want_split` is set, the row
// shape is compiled, ty is 2, 4 or 8, ny / ty is a compiled split length and its
// column workgroup fits the LDS, and the operator is the lattice-sum form; the other
// cases transform full columns.  Compile-time shapes transform in place (one line
// buffer plus the twiddle table), others use two ping-pong buffers.
int layered_shape(msd_dipole* p, int ty, int tpb, bool row_ct, bool want_split, const char* what)
{
    DipoleDev& dv = p->dev;
    const int nx = p->d.nx, ny = p->d.ny, nz = p->d.nz;
    dv.ty = ty;
    dv.tpb = tpb;
    const size_t rbuf = (size_t)((3 * ty + 1) / 2) * nx;
    dv.lds_rows = (row_ct ? rbuf + nx : 2 * rbuf + nx) * sizeof(float2);
    dv.split = 0;
    dv.ly = ny;
    if (want_split && row_ct && !p->pair && (ty == 2 || ty == 4 || ty == 8) && ny % ty == 0 &&
        split_length_compiled(ny / ty, ty) && split_cols_lds(ny / ty, nz) <= G.max_lds) {
        dv.split = ty;
        dv.ly = ny / ty;
        dv.lds_cols = split_cols_lds(dv.ly, nz);
        // the row passes stage the tile's 3 (nx/2+1) runs of ty words after the lines
        dv.lds_rows += (size_t)3 * (nx / 2 + 1) * ty * sizeof(float2);
    } else {
        // compile-time column lengths also stage the workgroup's tensor words (3 x nz x ny floats)
        const bool pow2_len_y = ny >= 32 && ny <= 1024 && (ny & (ny - 1)) == 0;
        const size_t cbuf = (size_t)3 * nz * ny;
        dv.lds_cols = pow2_len_y ? (cbuf + ny) * sizeof(float2) + (size_t)3 * nz * ny * sizeof(float)
                                 : 2 * cbuf * sizeof(float2);
    }
    if (dv.lds_rows > G.max_lds || dv.lds_cols > G.max_lds)
        return fail(MSD_EUNSUPPORTED, "%s.new: %dx%dx%d needs %zu/%zu bytes of LDS per workgroup (device limit %zu)", what,
                    nx, ny, nz, dv.lds_rows, dv.lds_cols, G.max_lds);
    return MSD_OK;
}

static int layered_create(int nx, int ny, int nz, bool pair, const char* what, msd_dipole** out)
{
    NONNULL(out);
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1) return fail(MSD_EINVAL, "%s.new: bad size", what);
    if (nx % 2) return fail(MSD_EUNSUPPORTED, "%s.new: nx=%d must be even (real-to-complex transform)", what, nx);
    if (nz > MSD_MAX_LAYERS)
        return fail(MSD_EUNSUPPORTED, "%s.new: nz=%d > %d layers is not supported by the layered kernel", what, nz,
                    MSD_MAX_LAYERS);
    FFTPlan px, py;
    if (!make_plan(nx, px) || !make_plan(ny, py))
        return fail(MSD_EUNSUPPORTED, "%s.new: %dx%d needs prime factors in {2,3,5}", what, nx, ny);
    if (int rc = ensure_init()) return rc;
    std::unique_ptr<msd_dipole> p(new msd_dipole);
    p->d = Dims{nx, ny, nz, nx * ny * nz};
    p->nxh = nx / 2 + 1;
    DipoleDev& dv = p->dev;
    dv.nx = nx; dv.ny = ny; dv.nz = nz; dv.nxh = p->nxh;
    dv.px = px; dv.py = py;
    // Row tiles: the default compile-time shape for nx (kernels_dipole.hip's
    // MSD_ROW_SHAPES) whose rows per tile divide ny; MSD_ROW_TILE="ty:tpb[:split]" picks
    // another compiled shape for tuning runs (split 0: full-column transforms).  Without one, runtime-shaped
    // kernels take the largest of 8,4,2,1 rows dividing ny whose two LDS
    // buffers fit 64 KiB.
    int ty = 0, tpb = 0, split = 0;
    row_shape_default(nx, ny, &ty, &tpb);
    if (const char* env = getenv("MSD_ROW_TILE")) {
        int t = 0, b = 0, sp = 1;
        const int got = sscanf(env, "%d:%d:%d", &t, &b, &sp);
        if (got >= 2 && t > 0 && ny % t == 0 && row_shape_compiled(nx, t, b)) { ty = t; tpb = b; split = got == 3 ? sp : 0; }
    }
    const bool row_ct = ty > 0;
    if (!row_ct) {
        ty = 1;
        tpb = 1024;
        for (int c : {8, 4, 2}) {
            const size_t need = (2 * (size_t)((3 * c + 1) / 2) + 1) * nx * sizeof(float2);
            if (ny % c == 0 && need <= 65536) { ty = c; break; }
        }
    }
    if (pair) p->pair = true;
    if (int rc = layered_shape(p.get(), ty, tpb, row_ct, split, what)) return rc;
    const size_t spec = (size_t)3 * nz * p->nxh * ny;
    if (int rc = dalloc(&p->Mh, spec)) return rc;
    if (int rc = dalloc(&p->Hh, spec)) return rc;
    if (pair) {
        p->slot = "LongRange2D";
        p->P.assign((size_t)9 * nz * nz * nx * ny, 0.0);
        p->sms.assign(nz, 1.0);
        p->dfs.assign(nz, 1.0);
        if (int rc = dalloc(&p->Tp, (size_t)9 * nz * nz * p->nxh * ny)) return rc;
        if (int rc = dalloc(&p->lsc, (size_t)2 * MSD_MAX_LAYERS)) return rc;
        dv.Tp = p->Tp;
        dv.lsc = p->lsc;
    } else {
        if (int rc = dalloc(&p->Tw, (size_t)6 * nz * p->nxh * ny)) return rc;
    }
    if (int rc = twiddles(nx, &p->twx, &p->twx64)) return rc;
    if (int rc = twiddles(ny, &p->twy, &p->twy64)) return rc;
    dv.Mh = p->Mh; dv.Hh = p->Hh; dv.Tw = p->Tw; dv.twx = p->twx; dv.twy = p->twy;
    *out = p.release();
    return MSD_OK;
}

extern "C" int msd_dipole_create(int nx, int ny, int nz, msd_dipole** out)
{
    return layered_create(nx, ny, nz, false, "Dipole", out);
}

extern "C" int msd_dipole_destroy(msd_dipole* d) { delete d; return MSD_OK; }
// the lattice-sum Dipole's own entry points refuse a LongRange2D object
#define LATTICE_ONLY(d, what)                                                                            \
    if ((d)->pair) return fail(MSD_EUNSUPPORTED, "%s: not defined for a %s operator", what, (d)->slot.c_str())

// LongRange2D's setStrength raises "deprecated" (cpu/longrange2d/longrange2d_luafuncs.lua:11-13)
extern "C" int msd_dipole_set_strength(msd_dipole* d, double g)
{
    NONNULL(d);
    if (d->pair)
        return fail(MSD_EUNSUPPORTED, "setStrength is deprecated. Please use setSourceMomentScale and setDestinationFieldScale.");
    d->g = g;
    return MSD_OK;
}
extern "C" int msd_dipole_set_scale(msd_dipole* d, double s) { NONNULL(d); d->scale = s; return MSD_OK; }

extern "C" int msd_dipole_set_truncation(msd_dipole* d, const int gmax[4])
{
    NONNULL(d); NONNULL(gmax);
    LATTICE_ONLY(d, "setTruncation");
    // -1: math.huge (the reference's marker, dipolesupport.cpp:511-519); its sum is the
    // square of images |X|, |Y| <= 4999 (kernels_tensor.hip:k_tensor_huge)
    if (gmax[0] < 1 && gmax[0] != -1) return fail(MSD_EINVAL, "setTruncation: radial limit %d", gmax[0]);
    for (int i = 0; i < 4; ++i) d->gmax[i] = gmax[i];
    d->loaded = false; d->custom = false; d->spectrum = false;
    return MSD_OK;
}

extern "C" int msd_dipole_set_unit_cell(msd_dipole* d, const double ABC[9])
{
    NONNULL(d); NONNULL(ABC);
    LATTICE_ONLY(d, "setUnitCell");
    for (int i = 0; i < 9; ++i) d->ABC[i] = ABC[i];
    d->loaded = false; d->custom = false; d->spectrum = false;
    return MSD_OK;
}

// Dipole::encode (cpu/dipole/spinoperationdipole.cpp:35-41) around LongRange::encode
// (cpu/longrange/spinoperationlongrange.cpp:55-68): the truncation, strength and unit
// cell from which the tensor is rebuilt; a matrix set by setMatrix is not in the stream
extern "C" int msd_dipole_encode(msd_dipole* d, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects)
{
    NONNULL(d); NONNULL(len);
    LATTICE_ONLY(d, "Dipole.encode");
    if (first_index < 0) return fail(MSD_EINVAL, "Dipole.encode: object index %d < 0", first_index);
    WireOut w;
    wire_operator(w, first_index, d->d, d->scale);
    w.ch(0);
    for (int k = 0; k < 4; ++k) w.i32(d->gmax[k]);
    w.f64(d->g);
    for (int k = 0; k < 9; ++k) w.f64(d->ABC[k]);
    w.ch(0);
    if (n_objects) *n_objects = 1;
    return wire_finish(w, "Dipole.encode", buf, cap, len);
}

// Dipole::decode (spinoperationdipole.cpp:43-59) around LongRange::decode (:70-92)
extern "C" int msd_dipole_decode(const void* buf, int64_t len, int first_index, msd_dipole** out, int64_t* used, int* n_objects)
{
    NONNULL(buf); NONNULL(out);
    *out = nullptr;
    if (len < 0) return fail(MSD_EINVAL, "Dipole.decode: length %lld", (long long)len);
    WireIn r{(const uint8_t*)buf, (size_t)len};
    int nx, ny, nz;
    double scale, g, ABC[9];
    if (int rc = unwire_operator(r, "Dipole.decode", first_index, &nx, &ny, &nz, &scale)) return rc;
    char v1, v2;
    int32_t gmax[4];
    if (!r.ch(&v1)) return fail(MSD_EINVAL, "Dipole.decode: stream too short");
    if (v1 != 0) return fail(MSD_EINVAL, "Dipole.decode: unknown LongRange version %d", (int)v1);
    for (int k = 0; k < 4; ++k)
        if (!r.i32(&gmax[k])) return fail(MSD_EINVAL, "Dipole.decode: stream ends inside the truncation");
    if (!r.f64(&g)) return fail(MSD_EINVAL, "Dipole.decode: stream too short");
    for (int k = 0; k < 9; ++k)
        if (!r.f64(&ABC[k])) return fail(MSD_EINVAL, "Dipole.decode: stream ends inside the unit cell");
    if (!r.ch(&v2)) return fail(MSD_EINVAL, "Dipole.decode: stream too short");
    if (v2 != 0) return fail(MSD_EINVAL, "Dipole.decode: unknown version %d", (int)v2);
    msd_dipole* d = nullptr;
    if (int rc = msd_dipole_create(nx, ny, nz, &d)) return rc;
    std::unique_ptr<msd_dipole> own(d);
    if (int rc = msd_dipole_set_truncation(d, gmax)) return rc;
    if (int rc = msd_dipole_set_unit_cell(d, ABC)) return rc;
    d->g = g;
    d->scale = scale;
    if (used) *used = (int64_t)r.at;
    if (n_objects) *n_objects = 1;
    *out = own.release();
    return MSD_OK;
}

static int dipole_load(msd_dipole* d)
{
    if (d->loaded) return MSD_OK;
    const size_t m = (size_t)6 * (2 * d->d.nz - 1) * d->d.nx * d->d.ny;
    if (!d->T64)
        if (int rc = dalloc(&d->T64, m)) return rc;
    // gmax[0] = -1 marks setTruncation(math.huge) (dipolesupport.cpp:511-519, 756)
    hipError_t e = d->gmax[0] == -1 ? launch_tensor_huge(d->d.nx, d->d.ny, d->d.nz, d->ABC, d->T64, G.stream)
                                    : launch_tensor_sum(d->d.nx, d->d.ny, d->d.nz, d->gmax, d->ABC, d->T64, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "dipole tensor: %s", hipGetErrorString(e));
    d->loaded = true;
    d->spectrum = false;
    return MSD_OK;
}

// LongRange::getAB offset mapping (spinoperationlongrange.cpp:143-175)
static int dipole_offset(msd_dipole* d, int which, int ox, int oy, int oz, size_t* idx)
{
    LATTICE_ONLY(d, "get/setMatrix(which, ox, oy, oz)");
    if (which < 0 || which > 5) return fail(MSD_EINVAL, "matrix index %d (0..5 = XX XY XZ YY YZ ZZ)", which);
    const int nx = d->d.nx, ny = d->d.ny, nz = d->d.nz;
    ox = ((ox % nx) + nx) % nx;
    oy = ((oy % ny) + ny) % ny;
    if (oz < -(nz - 1) || oz > nz - 1) return fail(MSD_EINVAL, "layer offset %d outside -(nz-1)..nz-1", oz);
    *idx = (size_t)which * (2 * nz - 1) * nx * ny + ox + (size_t)nx * (oy + (size_t)ny * (oz + nz - 1));
    return MSD_OK;
}

extern "C" int msd_dipole_get_matrix(msd_dipole* d, int which, int ox, int oy, int oz, double* v)
{
    NONNULL(d); NONNULL(v);
    size_t idx;
    if (int rc = dipole_offset(d, which, ox, oy, oz, &idx)) return rc;
    if (int rc = dipole_load(d)) return rc;
    HIPCHK(hipMemcpyAsync(v, d->T64 + idx, sizeof(double), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    return MSD_OK;
}

extern "C" int msd_dipole_set_matrix(msd_dipole* d, int which, int ox, int oy, int oz, double v)
{
    NONNULL(d);
    size_t idx;
    if (int rc = dipole_offset(d, which, ox, oy, oz, &idx)) return rc;
    if (int rc = dipole_load(d)) return rc;
    HIPCHK(hipMemcpyAsync(d->T64 + idx, &v, sizeof(double), hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    d->custom = true;
    d->spectrum = false;
    return MSD_OK;
}

// The layered kernel stores one word per component and uses the lattice-sum
// parities: XX YY ZZ XY even in the layer offset, XZ YZ odd; in-plane, XX YY ZZ
// even-even, XY odd-odd, XZ odd in x, YZ odd in y.  Custom tensors must keep them.
static int check_symmetry(msd_dipole* d)
{
    const int nx = d->d.nx, ny = d->d.ny, nz = d->d.nz, nl = 2 * nz - 1;
    const size_t plane = (size_t)nl * nx * ny;
    std::vector<double> t((size_t)6 * plane);
    HIPCHK(hipMemcpyAsync(t.data(), d->T64, sizeof(double) * t.size(), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    const int zsg[6] = {1, 1, -1, 1, -1, 1}, xsg[6] = {1, -1, -1, 1, 1, 1}, ysg[6] = {1, -1, 1, 1, -1, 1};
    double big = 0;
    for (double v : t) big = std::max(big, std::fabs(v));
    const double tol = 1e-12 * big;
    for (int a = 0; a < 6; ++a)
        for (int c = 0; c < nl; ++c)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const double v = t[a * plane + x + (size_t)nx * (y + (size_t)ny * c)];
                    const int cz = nl - 1 - c, xm = (nx - x) % nx, ym = (ny - y) % ny;
                    const double vz = t[a * plane + x + (size_t)nx * (y + (size_t)ny * cz)];
                    const double vx = t[a * plane + xm + (size_t)nx * (y + (size_t)ny * c)];
                    const double vy = t[a * plane + x + (size_t)nx * (ym + (size_t)ny * c)];
                    if (std::fabs(v - zsg[a] * vz) > tol || std::fabs(v - xsg[a] * vx) > tol || std::fabs(v - ysg[a] * vy) > tol)
                        return fail(MSD_EUNSUPPORTED,
                                    "Dipole: custom tensor component %d lacks the lattice-sum parity at (%d,%d,%d)", a, x, y,
                                    c - nz + 1);
                }
    return MSD_OK;
}

// LongRange2D::compile (cpu/longrange2d/spinoperationlongrange2d.cpp:396-446): FP64 2D
// transforms of every tensor plane, prescaled by 1/(nx ny), kept as FP32 words; the
// per-layer scales (applied at apply time, :463-526) go to a small device table
static int pair_prepare(msd_dipole* d)
{
    const int nx = d->d.nx, ny = d->d.ny, nz = d->d.nz;
    if (d->pair_dirty) {
        const int planes = 9 * nz * nz;
        double* T = nullptr;
        double2* U = nullptr;
        if (int rc = dalloc(&T, (size_t)planes * nx * ny)) return rc;
        if (int rc = dalloc(&U, (size_t)planes * ny * d->nxh)) { dfree(T); return rc; }
        hipError_t e = hipMemcpyAsync(T, d->P.data(), sizeof(double) * d->P.size(), hipMemcpyHostToDevice, G.stream);
        if (e == hipSuccess)
            e = launch_plane_spectrum(planes, nx, ny, T, U, d->Tp, 1.0 / ((double)nx * (double)ny), d->twx64, d->twy64,
                                      G.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
        dfree(T);
        dfree(U);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "%s tensor spectrum: %s", d->slot.c_str(), hipGetErrorString(e));
        d->pair_dirty = false;
    }
    float want[2 * MSD_MAX_LAYERS] = {};
    for (int l = 0; l < nz; ++l) {
        want[l] = (float)d->sms[l];
        want[MSD_MAX_LAYERS + l] = (float)(d->dfs[l] * d->scale);
    }
    if (!d->lsc_valid || std::memcmp(want, d->lsc_host, sizeof want) != 0) {
        std::memcpy(d->lsc_host, want, sizeof want);
        HIPCHK(hipMemcpyAsync(d->lsc, d->lsc_host, sizeof want, hipMemcpyHostToDevice, G.stream));
        HIPCHK(hipStreamSynchronize(G.stream));
        d->lsc_valid = true;
    }
    return MSD_OK;
}

int msd_dipole::prepare()
{
    if (pair) return pair_prepare(this);
    if (int rc = dipole_load(this)) return rc;
    if (!spectrum) {
        if (custom)
            if (int rc = check_symmetry(this)) return rc;
        const size_t wn = (size_t)6 * d.nz * nxh * d.ny;
        if (!U64)
            if (int rc = dalloc(&U64, wn)) return rc;
        if (!W64)
            if (int rc = dalloc(&W64, wn)) return rc;
        hipError_t e = launch_tensor_spectrum(d.nx, d.ny, d.nz, T64, U64, W64, twx64, twy64, G.stream);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "dipole spectrum: %s", hipGetErrorString(e));
        spectrum = true;
        words_mult = NAN;
    }
    const double mult = g * scale;
    if (!(words_mult == mult) || words_split != dev.split) {
        hipError_t e = launch_tensor_words(W64, Tw, mult, d.nz, nxh, d.ny, dev.split, G.stream);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "dipole words: %s", hipGetErrorString(e));
        words_mult = mult;
        words_split = dev.split;
    }
    return MSD_OK;
}

extern "C" int msd_dipole_schedule(const msd_dipole* d, int* rows_per_tile, int* threads, int* split)
{
    NONNULL(d); NONNULL(rows_per_tile); NONNULL(threads); NONNULL(split);
    *rows_per_tile = d->dev.ty;
    *threads = d->dev.tpb;
    *split = d->dev.split;
    return MSD_OK;
}

extern "C" int msd_dipole_apply(msd_dipole* d, msd_system* s)
{
    NONNULL(d); NONNULL(s);
    if (!same_dims(s->d, d->d.nx, d->d.ny, d->d.nz)) return fail(MSD_EMISMATCH, "%s:apply: size mismatch", d->slot.c_str());
    if (int rc = d->prepare()) return rc;
    int k;
    if (int rc = s->mark(d->slot.c_str(), &k)) return rc;
    d->epoch++;
    hipError_t e = launch_rows_fwd(d->dev, s->S, G.stream);
    if (e == hipSuccess) e = launch_cols(d->dev, G.stream);
    if (e == hipSuccess) e = launch_rows_inv(d->dev, s->slots[k].h, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "%s:apply: %s", d->slot.c_str(), hipGetErrorString(e));
    return MSD_OK;
}

// ------------------------------------------------------------ LongRange2D
// cpu/longrange2d/spinoperationlongrange2d.cpp: tensors per ordered layer pair, source-moment
// and destination-field scales per layer, the operator scale over all of it.

extern "C" int msd_lr2d_create(int nx, int ny, int nz, const char* slot, msd_dipole** out)
{
    if (int rc = layered_create(nx, ny, nz, true, slot && *slot ? slot : "LongRange2D", out)) return rc;
    if (slot && *slot) (*out)->slot = slot;
    return MSD_OK;
}

static int pair_plane(msd_dipole* d, int dest, int src, int ab, size_t* off, const char* what)
{
    if (!d->pair) return fail(MSD_EUNSUPPORTED, "%s: not a LongRange2D operator", what);
    const int nz = d->d.nz;
    if (dest < 0 || dest >= nz || src < 0 || src >= nz)
        return fail(MSD_EINVAL, "%s: layers (%d, %d) outside 1..%d", what, dest + 1, src + 1, nz);
    if (ab < 0 || ab > 8) return fail(MSD_EINVAL, "%s: tensor index %d (0..8 = XX XY XZ YX YY YZ ZX ZY ZZ)", what, ab);
    *off = (((size_t)ab * nz + dest) * nz + src) * (size_t)d->d.nx * d->d.ny;
    return MSD_OK;
}

// LongRange2D::setLAB / l_setTensorArray (spinoperationlongrange2d.cpp:198-224,594-602)
extern "C" int msd_lr2d_set_tensor(msd_dipole* d, int dest, int src, int ab, const double* a)
{
    NONNULL(d); NONNULL(a);
    size_t off;
    if (int rc = pair_plane(d, dest, src, ab, &off, "setTensorArray")) return rc;
    std::memcpy(d->P.data() + off, a, sizeof(double) * d->d.nx * d->d.ny);
    d->pair_dirty = true;
    return MSD_OK;
}

// LongRange2D::getLAB / l_getTensorArray (:168-196,585-592)
extern "C" int msd_lr2d_get_tensor(msd_dipole* d, int dest, int src, int ab, double* a)
{
    NONNULL(d); NONNULL(a);
    size_t off;
    if (int rc = pair_plane(d, dest, src, ab, &off, "tensorArray")) return rc;
    std::memcpy(a, d->P.data() + off, sizeof(double) * d->d.nx * d->d.ny);
    return MSD_OK;
}

// l_ssms / l_sdfs (:533-568): one value per layer
extern "C" int msd_lr2d_set_layer_scales(msd_dipole* d, const double* src_moment, const double* dest_field)
{
    NONNULL(d);
    if (!d->pair) return fail(MSD_EUNSUPPORTED, "layer scales: not a LongRange2D operator");
    for (int l = 0; l < d->d.nz; ++l) {
        if (src_moment) d->sms[l] = src_moment[l];
        if (dest_field) d->dfs[l] = dest_field[l];
    }
    return MSD_OK;
}

extern "C" int msd_lr2d_get_layer_scales(msd_dipole* d, double* src_moment, double* dest_field)
{
    NONNULL(d);
    if (!d->pair) return fail(MSD_EUNSUPPORTED, "layer scales: not a LongRange2D operator");
    for (int l = 0; l < d->d.nz; ++l) {
        if (src_moment) src_moment[l] = d->sms[l];
        if (dest_field) dest_field[l] = d->dfs[l];
    }
    return MSD_OK;
}

// Dipole2D point tensor (cpu/dipole2d/demag.cpp:3-62): N_ab(r) = 3 r_a r_b / r^5 - delta_ab / r^3,
// 0 at r < 1e-80.  Component order XX XY XZ YX YY YZ ZX ZY ZZ.
static void dipole2d_point(double rx, double ry, double rz, double N[9])
{
    const double r = std::sqrt(rx * rx + ry * ry + rz * rz);
    if (r < 1e-80) {
        for (int k = 0; k < 9; ++k) N[k] = 0;
        return;
    }
    const double ir = 1.0 / r, ir3 = ir * ir * ir, ir5 = ir3 * ir * ir;
    const double v[3] = {rx, ry, rz};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) N[3 * a + b] = -1.0 * ((a == b ? ir3 : 0.0) - 3.0 * v[a] * v[b] * ir5);
}

// Dipole2D makeData (cpu/dipole2d/dip2d_luafuncs.lua:537-668).  For every label pair (dest,
// src) with |src - dest| <= truncZ and every X in [-truncX, truncX], Y in [-truncY, truncY],
// N(p_src - p_dest) is added at (X mod nx, Y mod ny) of T[dest][src], p = idx2pos(...) (:537-557):
// X A1 + Y B1' + (C_1 + ... + C_{layer-1}) + offset[layer], with A1, B1 the first layer's
// cell vectors.  The reference evaluates p_src and p_dest at the same (X, Y), so the
// in-plane terms cancel and every cell of the box gets the layers' vertical coupling;
// and its B1' is (B1y, B1y, B1z).  physical = 0 keeps both (the reference's tensors);
// physical = 1 places the destination at the origin, p_dest = idx2pos(0, 0, dest), and
// uses B1 as given, which makes T[dest][src](X, Y) the coupling of a source at the origin of
// layer `dest` to a site at in-plane cell (X, Y) of layer `src` (the pair the reference's
// apply reads as [source][destination]).  ABC: [nz][3][3] (A, B, C rows), offset: [nz][3].
extern "C" int msd_lr2d_build_dipole2d(msd_dipole* d, const double* ABC, const double* offset, const int trunc[4],
                                       int physical)
{
    NONNULL(d); NONNULL(ABC); NONNULL(offset); NONNULL(trunc);
    if (!d->pair) return fail(MSD_EUNSUPPORTED, "Dipole2D makeData: not a LongRange2D operator");
    const int nx = d->d.nx, ny = d->d.ny, nz = d->d.nz, tX = trunc[1], tY = trunc[2], tZ = trunc[3];
    if (tX < 0 || tY < 0 || tZ < 0) return fail(MSD_EINVAL, "Dipole2D: negative truncation (%d, %d, %d)", tX, tY, tZ);
    if ((double)(2 * tX + 1) * (2 * tY + 1) * nz * nz > 4e9)
        return fail(MSD_EUNSUPPORTED, "Dipole2D: truncation box (%d, %d) too large", tX, tY);
    auto pos = [&](int X, int Y, int layer, double p[3]) {
        const double* A = ABC;          // first layer's A and B
        const double* B = ABC + 3;
        double dz[3] = {0, 0, 0};
        for (int Z = 0; Z < layer; ++Z)
            for (int c = 0; c < 3; ++c) dz[c] += ABC[9 * Z + 6 + c];
        const double bj[3] = {physical ? B[0] : B[1], B[1], B[2]};
        for (int c = 0; c < 3; ++c) p[c] = X * A[c] + Y * bj[c] + dz[c] + offset[3 * layer + c];
    };
    const size_t plane = (size_t)nx * ny;
    std::fill(d->P.begin(), d->P.end(), 0.0);
    for (int dest = 0; dest < nz; ++dest)
        for (int src = 0; src < nz; ++src) {
            if (std::abs(src - dest) > tZ) continue;
            for (int X = -tX; X <= tX; ++X)
                for (int Y = -tY; Y <= tY; ++Y) {
                    double pd[3], ps[3], N[9];
                    if (physical) pos(0, 0, dest, pd);
                    else pos(X, Y, dest, pd);
                    pos(X, Y, src, ps);
                    dipole2d_point(ps[0] - pd[0], ps[1] - pd[1], ps[2] - pd[2], N);
                    const int x = ((X % nx) + nx) % nx, y = ((Y % ny) + ny) % ny;
                    for (int ab = 0; ab < 9; ++ab)
                        d->P[(((size_t)ab * nz + dest) * nz + src) * plane + x + (size_t)nx * y] += N[ab];
                }
        }
    d->pair_dirty = true;
    return MSD_OK;
}

// ---------------------------------------------------------- Magnetostatics2D
// The demagnetising tensor between two rectangular prisms (Newell, Williams and
// Dunlop, J. Geophys. Res. 98 (1993) 9551; Donahue's form for prisms of unequal
// size): with Newell's f (diagonal) and g (off-diagonal) potentials,
//   N_xx(r; s, d) = D_x D_y D_z f(r) / (4 pi V_d),  D g(u) = g(u) - g(u + d) + g(u + d - s) - g(u - s)
// along each axis, s the source prism, d the destination prism, r the vector from the
// source's origin corner to the destination's.  N is the destination average of the
// field of the uniformly magnetised source, per unit magnetisation, with the sign that
// makes a cube's self term +1/3.  The other components permute the axes.
namespace {
double nw_asinh(double num, double den) { return den > 0.0 ? std::asinh(num / den) : 0.0; }
double nw_atan(double num, double den) { return (num == 0.0 && den == 0.0) ? 0.0 : std::atan2(num, den); }

// f(x, y, z) = y/2 (z^2 - x^2) asinh(y / sqrt(x^2 + z^2)) + z/2 (y^2 - x^2) asinh(z / sqrt(x^2 + y^2))
//              - x y z atan(y z / (x R)) + (2 x^2 - y^2 - z^2) R / 6, even in every argument
double newell_f(double X, double Y, double Z)
{
    const double x = std::fabs(X), y = std::fabs(Y), z = std::fabs(Z);
    const double x2 = x * x, y2 = y * y, z2 = z * z, R = std::sqrt(x2 + y2 + z2);
    return 0.5 * y * (z2 - x2) * nw_asinh(y, std::sqrt(x2 + z2)) + 0.5 * z * (y2 - x2) * nw_asinh(z, std::sqrt(x2 + y2)) -
           x * y * z * nw_atan(y * z, x * R) + (2.0 * x2 - y2 - z2) * R / 6.0;
}

// g(x, y, z) = x y z asinh(z / sqrt(x^2 + y^2)) + y/6 (3 z^2 - y^2) asinh(x / sqrt(y^2 + z^2))
//            + x/6 (3 z^2 - x^2) asinh(y / sqrt(x^2 + z^2)) - z^3/6 atan(x y / (z R))
//            - z y^2/2 atan(x z / (y R)) - z x^2/2 atan(y z / (x R)) - x y R / 3,
// odd in x and in y, even in z
double newell_g(double X, double Y, double Z)
{
    const double x = std::fabs(X), y = std::fabs(Y), z = std::fabs(Z);
    const double x2 = x * x, y2 = y * y, z2 = z * z, R = std::sqrt(x2 + y2 + z2);
    const double v = x * y * z * nw_asinh(z, std::sqrt(x2 + y2)) + y * (3.0 * z2 - y2) / 6.0 * nw_asinh(x, std::sqrt(y2 + z2)) +
                     x * (3.0 * z2 - x2) / 6.0 * nw_asinh(y, std::sqrt(x2 + z2)) - z2 * z / 6.0 * nw_atan(x * y, z * R) -
                     0.5 * z * y2 * nw_atan(x * z, y * R) - 0.5 * z * x2 * nw_atan(y * z, x * R) - x * y * R / 3.0;
    return ((X < 0) != (Y < 0)) ? -v : v;
}

// D_x D_y D_z fn over the corner offsets {0, d, d - s, -s}, signs + - + -
template <double (*FN)(double, double, double)>
double prism_difference(double X, double Y, double Z, const double s[3], const double d[3])
{
    const double off[3][4] = {{0.0, d[0], d[0] - s[0], -s[0]}, {0.0, d[1], d[1] - s[1], -s[1]}, {0.0, d[2], d[2] - s[2], -s[2]}};
    double acc = 0.0;
    for (int a = 0; a < 4; ++a) {
        double ay = 0.0;
        for (int b = 0; b < 4; ++b) {
            double az = 0.0;
            for (int c = 0; c < 4; ++c) az += ((c & 1) ? -1.0 : 1.0) * FN(X + off[0][a], Y + off[1][b], Z + off[2][c]);
            ay += ((b & 1) ? -1.0 : 1.0) * az;
        }
        acc += ((a & 1) ? -1.0 : 1.0) * ay;
    }
    return acc / (4.0 * M_PI * d[0] * d[1] * d[2]);
}

// the nine components N[ab] (XX XY XZ YX YY YZ ZX ZY ZZ) at r between source s and destination d
void prism_tensor(const double r[3], const double s[3], const double d[3], double N[9])
{
    auto perm = [](const double* v, int i, int j, int k, double* o) { o[0] = v[i]; o[1] = v[j]; o[2] = v[k]; };
    double ps[3], pd[3];
    N[0] = prism_difference<newell_f>(r[0], r[1], r[2], s, d);
    perm(s, 1, 0, 2, ps); perm(d, 1, 0, 2, pd);
    N[4] = prism_difference<newell_f>(r[1], r[0], r[2], ps, pd);
    perm(s, 2, 1, 0, ps); perm(d, 2, 1, 0, pd);
    N[8] = prism_difference<newell_f>(r[2], r[1], r[0], ps, pd);
    N[1] = N[3] = prism_difference<newell_g>(r[0], r[1], r[2], s, d);
    perm(s, 0, 2, 1, ps); perm(d, 0, 2, 1, pd);
    N[2] = N[6] = prism_difference<newell_g>(r[0], r[2], r[1], ps, pd);
    perm(s, 1, 2, 0, ps); perm(d, 1, 2, 0, pd);
    N[5] = N[7] = prism_difference<newell_g>(r[1], r[2], r[0], ps, pd);
}
}  // namespace

// Magnetostatics2D makeData (cpu/magnetostatics2d/mag2d_luafuncs.lua:598-750): for source
// layer z1 and destination z2 with |z1 - z2| <= truncZ, images X, Y in [-trunc, trunc]
// within truncX, truncY, r = X A + Y B + c(z2) - c(z1) (A, B of the first layer, c(z) the
// sum of the C vectors of the layers below z), -N(r; grain z1, grain z2) is added at
// (X mod nx, Y mod ny) of the pair stored under the labels [z1][z2].
extern "C" int msd_lr2d_build_magnetostatics2d(msd_dipole* d, const double* ABC, const double* grain, const int trunc[4])
{
    NONNULL(d); NONNULL(ABC); NONNULL(grain); NONNULL(trunc);
    if (!d->pair) return fail(MSD_EUNSUPPORTED, "Magnetostatics2D makeData: not a LongRange2D operator");
    const int nx = d->d.nx, ny = d->d.ny, nz = d->d.nz, t = trunc[0], tX = trunc[1], tY = trunc[2], tZ = trunc[3];
    if (t < 0 || tX < 0 || tY < 0 || tZ < 0)
        return fail(MSD_EINVAL, "Magnetostatics2D: negative truncation (%d, %d, %d, %d)", t, tX, tY, tZ);
    if ((double)(2 * t + 1) * (2 * t + 1) * nz * nz > 1e8)
        return fail(MSD_EUNSUPPORTED, "Magnetostatics2D: truncation %d too large", t);
    for (int k = 0; k < 3 * nz; ++k)
        if (!(grain[k] > 0.0)) return fail(MSD_EINVAL, "Magnetostatics2D: grain sizes must be positive");
    const double* A = ABC;
    const double* B = ABC + 3;
    const size_t plane = (size_t)nx * ny;
    std::fill(d->P.begin(), d->P.end(), 0.0);
    for (int z1 = 0; z1 < nz; ++z1)
        for (int z2 = 0; z2 < nz; ++z2) {
            if (std::abs(z1 - z2) > tZ) continue;
            double c[3] = {0.0, 0.0, 0.0};
            for (int j = std::min(z1, z2); j < std::max(z1, z2); ++j)
                for (int k = 0; k < 3; ++k) c[k] += (z2 > z1 ? 1.0 : -1.0) * ABC[9 * j + 6 + k];
            for (int X = -t; X <= t; ++X) {
                if (std::abs(X) > tX) continue;
                for (int Y = -t; Y <= t; ++Y) {
                    if (std::abs(Y) > tY) continue;
                    const double r[3] = {A[0] * X + B[0] * Y + c[0], A[1] * X + B[1] * Y + c[1], A[2] * X + B[2] * Y + c[2]};
                    double N[9];
                    prism_tensor(r, grain + 3 * z1, grain + 3 * z2, N);
                    const size_t cell = (size_t)(((X % nx) + nx) % nx) + (size_t)nx * (((Y % ny) + ny) % ny);
                    for (int ab = 0; ab < 9; ++ab) d->P[(((size_t)ab * nz + z1) * nz + z2) * plane + cell] -= N[ab];
                }
            }
        }
    d->pair_dirty = true;
    return MSD_OK;
}

// dipole field of arbitrary device spins s3 ([3][n]) into h ([3][n]): the layered
// transform of the Dipole operator without a SpinSystem
int dipole_field_of(msd_dipole* dip, const float* s3, float* h)
{
    if (int rc = dip->prepare()) return rc;
    dip->epoch++;
    hipError_t e = launch_rows_fwd(dip->dev, s3, G.stream);
    if (e == hipSuccess) e = launch_cols(dip->dev, G.stream);
    if (e == hipSuccess) e = launch_rows_inv(dip->dev, h, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "dipole field: %s", hipGetErrorString(e));
    return MSD_OK;
}
