// LongRange3D of libmsd.so (periodic and zero-padded 3D convolution) and the pass
// tables that run batches of 3D and local-field members.
#include "msd_host.h"

static int choose_width(int L, int lines_per, size_t budget)
{
    for (int w : {16, 8, 4, 2})
        if (2 * (size_t)w * lines_per * L * sizeof(float2) <= budget) return w;
    return 1;
}

extern "C" int msd_longrange3d_create(int nx, int ny, int nz, int open, msd_longrange3d** out)
{
    NONNULL(out);
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1) return fail(MSD_EINVAL, "LongRange3D.new: bad size");
    if (int rc = ensure_init()) return rc;
    std::unique_ptr<msd_longrange3d> l(new msd_longrange3d);
    l->d = Dims{nx, ny, nz, nx * ny * nz};
    l->open = open != 0;
    l->fx = open ? 2 * nx : nx; l->fy = open ? 2 * ny : ny; l->fz = open ? 2 * nz : nz;
    if (l->fx % 2) return fail(MSD_EUNSUPPORTED, "LongRange3D.new: the x transform length %d must be even", l->fx);
    if (!make_plan(l->fx, l->px) || !make_plan(l->fy, l->py) || !make_plan(l->fz, l->pz))
        return fail(MSD_EUNSUPPORTED, "LongRange3D.new: %dx%dx%d transform needs prime factors in {2,3,5}", l->fx, l->fy, l->fz);
    l->fxh = l->fx / 2 + 1;
    l->fxp = (l->fxh + 15) / 16 * 16;   // 128-byte W rows: the y and z passes' kx runs stay within cache lines
    // 64 KiB tiles: measured faster at 256^3 open than the full LDS (occupancy)
    const size_t budget = 65536;
    l->ty = 1;
    for (int c : {8, 4, 2})
        if ((2 * (size_t)((3 * c + 1) / 2)) * l->fx * sizeof(float2) <= budget) { l->ty = c; break; }
    l->wy = choose_width(l->fy, 1, budget);
    l->wz = choose_width(l->fz, 3, budget);
    l->lds_rows = 2 * (size_t)((3 * l->ty + 1) / 2) * l->fx * sizeof(float2);
    l->lds_y = 2 * (size_t)l->wy * l->fy * sizeof(float2);
    l->lds_z = 2 * (size_t)3 * l->wz * l->fz * sizeof(float2);
    // the compile-time y and z passes (in-place transforms, wider kx blocks) where
    // they have instances
    if (const int th = rows_template_height(l->fx)) {
        // the compile-time row pass transforms in place: one buffer of its lines
        // and the x twiddles
        const size_t lr = (size_t)((3 * th + 1) / 2 + 1) * l->fx * sizeof(float2);
        if (th <= ny && lr <= G.max_lds) { l->rt = l->fx; l->ty = th; l->lds_rows = lr; }
    }
    if (const int wt = ypass_template_width(l->fy)) {
        const size_t ly = (size_t)wt * (l->fy + CONV3_LPAD) * sizeof(float2);
        if (ly <= G.max_lds) { l->yt = l->fy; l->wy = wt; l->lds_y = ly; }
    }
    if (const int wt = zconv_template_width(l->fz)) {
        const size_t lz = (size_t)3 * wt * (l->fz + CONV3_LPAD) * sizeof(float2);
        if (lz <= G.max_lds) { l->zt = l->fz; l->wz = wt; l->lds_z = lz; }
    }
    // open boundaries: the split pass of two half-length transforms (lines of nz
    // plus the half-length and shift twiddles)
    if (const int wt = l->open ? zconv_open_width(l->fz) : 0) {
        const size_t lz = ((size_t)3 * wt * (nz + CONV3_LPAD) + 2 * (size_t)nz) * sizeof(float2);
        if (lz <= G.max_lds) { l->zt = l->fz + 1; l->wz = wt; l->lds_z = lz; }
    }
    if (l->lds_rows > G.max_lds || l->lds_y > G.max_lds || l->lds_z > G.max_lds)
        return fail(MSD_EUNSUPPORTED, "LongRange3D.new: %dx%dx%d transform exceeds the LDS per workgroup", l->fx, l->fy, l->fz);
    if (int rc = twiddles(l->fx, &l->twx, &l->twx64)) return rc;
    if (int rc = twiddles(l->fy, &l->twy, &l->twy64)) return rc;
    if (int rc = twiddles(l->fz, &l->twz, &l->twz64)) return rc;
    if (int rc = dalloc(&l->W, (size_t)3 * nz * l->fy * l->fxp)) return rc;
    if (int rc = dalloc(&l->Tw, (size_t)6 * l->fz * l->fy * l->fxh)) return rc;
    *out = l.release();
    return MSD_OK;
}

extern "C" int msd_longrange3d_destroy(msd_longrange3d* l) { delete l; return MSD_OK; }

// LDS bytes of the compile-time y pass and z pass form at kx width w
static size_t lds_ypass(const msd_longrange3d* l, int w) { return (size_t)w * (l->fy + CONV3_LPAD) * sizeof(float2); }
static size_t lds_zpass(const msd_longrange3d* l, int w)
{
    if (l->zt & 1) return ((size_t)3 * w * (l->d.nz + CONV3_LPAD) + 2 * (size_t)l->d.nz) * sizeof(float2);
    return (size_t)3 * w * (l->fz + CONV3_LPAD) * sizeof(float2);
}

// Whether width w is compiled for the pass form and fits the LDS
static bool width_ok(const msd_longrange3d* l, bool ypass, int w)
{
    int ws[8];
    const int n = ypass ? ypass_widths(l->yt, ws, 8) : zconv_widths(l->zt, ws, 8);
    for (int i = 0; i < n; ++i)
        if (ws[i] == w) return (ypass ? lds_ypass(l, w) : lds_zpass(l, w)) <= G.max_lds;
    return false;
}

int lr3_set_widths(msd_longrange3d* l, int wy, int wz)
{
    if (wy == l->wy) wy = 0;    // the width in use stays (also a runtime-length pass's own)
    if (wz == l->wz) wz = 0;
    if (wy && (!l->yt || !width_ok(l, true, wy)))
        return fail(MSD_EUNSUPPORTED, "LongRange3D: no y pass of width %d at y length %d", wy, l->fy);
    if (wz && (!l->zt || !width_ok(l, false, wz)))
        return fail(MSD_EUNSUPPORTED, "LongRange3D: no z pass of width %d at z length %d", wz, l->fz);
    if (!wy && !wz) return MSD_OK;
    if (wy) { l->wy = wy; l->lds_y = lds_ypass(l, wy); }
    if (wz) { l->wz = wz; l->lds_z = lds_zpass(l, wz); }
    // the operator API's work items follow the widths: rebuilt at the next field
    (void)hipStreamSynchronize(G.stream);
    dfree(l->orep); dfree(l->oitems_r); dfree(l->oitems_y); dfree(l->oitems_z);
    l->orep = nullptr; l->oitems_r = l->oitems_y = l->oitems_z = nullptr;
    return MSD_OK;
}

extern "C" int msd_longrange3d_set_pass_widths(msd_longrange3d* l, int wy, int wz)
{
    NONNULL(l);
    if (wy < 0 || wz < 0) return fail(MSD_EINVAL, "LongRange3D: pass widths %d, %d", wy, wz);
    return lr3_set_widths(l, wy, wz);
}

extern "C" int msd_longrange3d_pass_widths(msd_longrange3d* l, int* wy, int* wz, int* y_choices, int* ny,
                                          int* z_choices, int* nz)
{
    NONNULL(l); NONNULL(wy); NONNULL(wz);
    *wy = l->wy;
    *wz = l->wz;
    // the compiled widths that also fit the LDS
    auto list = [&](bool ypass, int* out, int* n) {
        int ws[8];
        const int m = (ypass ? l->yt : l->zt) ? (ypass ? ypass_widths(l->yt, ws, 8) : zconv_widths(l->zt, ws, 8)) : 0;
        int k = 0;
        for (int i = 0; i < m; ++i)
            if (width_ok(l, ypass, ws[i])) { if (out) out[k] = ws[i]; ++k; }
        if (n) *n = k;
    };
    list(true, y_choices, ny);
    list(false, z_choices, nz);
    return MSD_OK;
}
extern "C" int msd_longrange3d_set_strength(msd_longrange3d* l, double g) { NONNULL(l); l->g = g; return MSD_OK; }
extern "C" int msd_longrange3d_set_scale(msd_longrange3d* l, double s) { NONNULL(l); l->scale = s; return MSD_OK; }
extern "C" int msd_longrange3d_set_radius(msd_longrange3d* l, int R)
{
    NONNULL(l);
    if (R < 1) return fail(MSD_EINVAL, "LongRange3D: lattice-sum radius %d", R);
    l->R = R;
    l->loaded = false; l->custom = false; l->spectrum = false;
    return MSD_OK;
}

static int lr3_load(msd_longrange3d* l)
{
    if (l->loaded) return MSD_OK;
    const size_t m = (size_t)6 * l->fx * l->fy * l->fz;
    if (!l->T64)
        if (int rc = dalloc(&l->T64, m)) return rc;
    hipError_t e = launch3_tensor(l->d.nx, l->d.ny, l->d.nz, l->fx, l->fy, l->fz, l->open ? 0 : l->R, l->T64, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "LongRange3D tensor: %s", hipGetErrorString(e));
    l->loaded = true;
    l->spectrum = false;
    return MSD_OK;
}

static int lr3_offset(msd_longrange3d* l, int which, int ox, int oy, int oz, size_t* idx)
{
    if (which < 0 || which > 5) return fail(MSD_EINVAL, "matrix index %d (0..5 = XX XY XZ YY YZ ZZ)", which);
    ox = ((ox % l->fx) + l->fx) % l->fx;
    oy = ((oy % l->fy) + l->fy) % l->fy;
    oz = ((oz % l->fz) + l->fz) % l->fz;
    *idx = (size_t)which * l->fx * l->fy * l->fz + ox + (size_t)l->fx * (oy + (size_t)l->fy * oz);
    return MSD_OK;
}

extern "C" int msd_longrange3d_get_matrix(msd_longrange3d* l, int which, int ox, int oy, int oz, double* v)
{
    NONNULL(l); NONNULL(v);
    size_t idx;
    if (int rc = lr3_offset(l, which, ox, oy, oz, &idx)) return rc;
    if (int rc = lr3_load(l)) return rc;
    HIPCHK(hipMemcpyAsync(v, l->T64 + idx, sizeof(double), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    return MSD_OK;
}

extern "C" int msd_longrange3d_set_matrix(msd_longrange3d* l, int which, int ox, int oy, int oz, double v)
{
    NONNULL(l);
    size_t idx;
    if (int rc = lr3_offset(l, which, ox, oy, oz, &idx)) return rc;
    if (int rc = lr3_load(l)) return rc;
    HIPCHK(hipMemcpyAsync(l->T64 + idx, &v, sizeof(double), hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    l->custom = true;
    l->spectrum = false;
    return MSD_OK;
}

// the stored spectrum is the real part: the tensor must be even, T(-r) = T(r)
static int lr3_check_even(msd_longrange3d* l)
{
    const size_t plane = (size_t)l->fx * l->fy * l->fz;
    std::vector<double> t(6 * plane);
    HIPCHK(hipMemcpyAsync(t.data(), l->T64, sizeof(double) * t.size(), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    double big = 0;
    for (double v : t) big = std::max(big, std::fabs(v));
    for (int a = 0; a < 6; ++a)
        for (int z = 0; z < l->fz; ++z)
            for (int y = 0; y < l->fy; ++y)
                for (int x = 0; x < l->fx; ++x) {
                    const size_t i = a * plane + x + (size_t)l->fx * (y + (size_t)l->fy * z);
                    const size_t j = a * plane + (l->fx - x) % l->fx + (size_t)l->fx * ((l->fy - y) % l->fy + (size_t)l->fy * ((l->fz - z) % l->fz));
                    if (std::fabs(t[i] - t[j]) > 1e-12 * big)
                        return fail(MSD_EUNSUPPORTED, "LongRange3D: custom tensor component %d is not even at (%d,%d,%d)", a, x, y, z);
                }
    return MSD_OK;
}

// Whether every component has its parity along one axis (1 = y, 2 = z) alone:
// T(x, -y, z) = +T (xx xz yy zz) or -T (xy yz), T(x, y, -z) = +T (xx xy yy zz) or
// -T (xz yz).  The spectrum then has the same parity in that k and only its first
// half is stored.  The built-in tensors have both (reflection symmetry of the
// lattice sums).
static int lr3_parity(msd_longrange3d* l, int axis, bool* has)
{
    const size_t plane = (size_t)l->fx * l->fy * l->fz;
    std::vector<double> t(6 * plane);
    HIPCHK(hipMemcpyAsync(t.data(), l->T64, sizeof(double) * t.size(), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    double big = 0;
    for (double v : t) big = std::max(big, std::fabs(v));
    *has = true;
    for (int a = 0; a < 6 && *has; ++a) {
        const bool odd = axis == 1 ? (a == 1 || a == 4) : (a == 2 || a == 4);
        const double sg = odd ? -1.0 : 1.0;
        for (int z = 0; z < l->fz && *has; ++z)
            for (int y = 0; y < l->fy && *has; ++y) {
                const int zm = axis == 2 ? (l->fz - z) % l->fz : z, ym = axis == 1 ? (l->fy - y) % l->fy : y;
                const double* u = &t[a * plane + ((size_t)z * l->fy + y) * l->fx];
                const double* v = &t[a * plane + ((size_t)zm * l->fy + ym) * l->fx];
                for (int x = 0; x < l->fx; ++x)
                    if (std::fabs(u[x] - sg * v[x]) > 1e-12 * big) { *has = false; break; }
            }
    }
    return MSD_OK;
}

int msd_longrange3d::prepare()
{
    if (int rc = lr3_load(this)) return rc;
    if (!spectrum) {
        bool zpar = true, ypar = true;
        if (custom) {
            if (int rc = lr3_check_even(this)) return rc;
            if (int rc = lr3_parity(this, 2, &zpar)) return rc;
            if (int rc = lr3_parity(this, 1, &ypar)) return rc;
        }
        fzs = zpar ? fz / 2 + 1 : fz;
        fys = ypar ? fy / 2 + 1 : fy;
        words_mult = NAN;
        spectrum = true;
    }
    const double mult = g * scale / ((double)fx * fy * fz);
    if (!(words_mult == mult)) {
        const size_t wn = (size_t)6 * fz * fy * fxh;
        double2 *U1 = nullptr, *U2 = nullptr;
        if (int rc = dalloc(&U1, wn)) return rc;
        if (int rc = dalloc(&U2, wn)) { dfree(U1); return rc; }
        hipError_t e = launch3_spectrum(fx, fy, fz, fys, fzs, T64, U1, U2, Tw, mult, twx64, twy64, twz64, G.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
        dfree(U1); dfree(U2);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "LongRange3D spectrum: %s", hipGetErrorString(e));
        words_mult = mult;
    }
    return MSD_OK;
}

// work items of one replica appended to the tables
void items3(int rep, const Dims& d, const Conv3Dev* cv, int ty, std::vector<int4>& rows, std::vector<int4>& ys,
                   std::vector<int4>& zs)
{   // zs: this replica's z-pass items (its own z-pass form, cv->zt)
    for (int z = 0; z < d.nz; ++z)
        for (int y0 = 0; y0 < d.ny; y0 += ty) rows.push_back(make_int4(rep, z, y0, 0));
    if (!cv) return;
    for (int l = 0; l < 3 * d.nz; ++l)
        for (int k = 0; k < cv->fxh; k += cv->wy) ys.push_back(make_int4(rep, l, k, 0));
    // rows ky and fy - ky side by side: with the tensor stored for ky <= fy/2 they
    // read the same words, which the second then finds in its XCD's L2
    auto zrow = [&](int ky) {
        for (int k = 0; k < cv->fxh; k += cv->wz) zs.push_back(make_int4(rep, ky, k, 0));
    };
    zrow(0);
    for (int j = 1; 2 * j <= cv->fy; ++j) {
        zrow(j);
        if (cv->fy - j != j) zrow(cv->fy - j);
    }
}

// Workgroups go to the 8 XCDs round-robin by index.  Reorder a y or z pass's items
// so that each XCD takes one contiguous stretch of the natural order (balanced to
// within one item): neighbouring kx blocks, which share the cache lines at their
// edges, then meet in one L2.  (Row tiles keep the natural order: measured faster.)
std::vector<int4> xcd_order(const std::vector<int4>& v)
{
    constexpr size_t parts = 8;
    const size_t n = v.size();
    std::vector<int4> out;
    out.reserve(n);
    for (size_t j = 0; out.size() < n; ++j)
        for (size_t p = 0; p < parts; ++p) {
            const size_t b = p * n / parts, e = (p + 1) * n / parts;
            if (b + j < e) out.push_back(v[b + j]);
        }
    return out;
}


// Exact LongRange3D field of the spins S ([3][n]) into H ([3][n], overwritten): the
// five passes with the single-replica descriptor.  The working spectrum W is left
// holding the field's spectrum.
int lr3_field_of(msd_longrange3d* l, const float* S, float* H)
{
    if (int rc = l->prepare()) return rc;
    l->epoch++;
    Rep3 r{};
    r.a.d = l->d;
    r.cv = l->conv();
    r.cv.sin = S;
    r.cv.hout = H;
    r.has_lr = 1;
    if (!l->orep) {
        std::vector<int4> rows, ys, zs;
        items3(0, l->d, &r.cv, l->ty, rows, ys, zs);
        if (int rc = upload_vec(&l->oitems_r, rows)) return rc;
        if (int rc = upload_vec(&l->oitems_y, xcd_order(ys))) return rc;
        if (int rc = upload_vec(&l->oitems_z, xcd_order(zs))) return rc;
        l->onr = (int)rows.size(); l->ony = (int)ys.size(); l->onz = (int)zs.size();
        if (int rc = dalloc(&l->orep, 1)) return rc;
    }
    HIPCHK(hipMemcpyAsync(l->orep, &r, sizeof r, hipMemcpyHostToDevice, G.stream));
    hipError_t e = launch3_rows(R3_FWD, false, l->orep, l->oitems_r, l->onr, l->lds_rows, l->rt, G.stream);
    if (e == hipSuccess) e = launch3_ypass(-1, l->orep, l->oitems_y, l->ony, l->lds_y, l->yt, l->wy, G.stream);
    if (e == hipSuccess) e = launch3_zconv(l->orep, l->oitems_z, l->onz, l->lds_z, l->zt, l->wz, G.stream);
    if (e == hipSuccess) e = launch3_ypass(1, l->orep, l->oitems_y, l->ony, l->lds_y, l->yt, l->wy, G.stream);
    if (e == hipSuccess) e = launch3_rows(R3_INV, false, l->orep, l->oitems_r, l->onr, l->lds_rows, l->rt, G.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G.stream);   // the descriptor lives on the host stack
    if (e != hipSuccess) return fail(MSD_EDEVICE, "LongRange3D field: %s", hipGetErrorString(e));
    return MSD_OK;
}

extern "C" int msd_longrange3d_apply(msd_longrange3d* l, msd_system* s)
{
    NONNULL(l); NONNULL(s);
    if (!same_dims(s->d, l->d.nx, l->d.ny, l->d.nz)) return fail(MSD_EMISMATCH, "LongRange3D:apply: size mismatch");
    if (int rc = l->prepare()) return rc;
    int k;
    if (int rc = s->mark("LongRange3D", &k)) return rc;
    return lr3_field_of(l, s->S, s->slots[k].h);
}
