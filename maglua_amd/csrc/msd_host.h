// Host-side objects of libmsd.so shared by the C-ABI sources: the MagLua objects
// (SpinSystem, operators, LLG), the step drivers, the surrogate and ensembles, the
// library globals and the helpers every source uses.  Host state is FP64; device
// state FP32 structure-of-arrays.
#pragma once

#include "msd_internal.h"
#include "../../include/msd_abi.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>
#include <strings.h>

using namespace msd;

// the last error message of this thread (msd_last_error) and its setter
extern thread_local std::string g_err;
int fail(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(MSD_EDEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                         \
    } while (0)

#define NONNULL(p) \
    do { if (!(p)) return fail(MSD_EINVAL, "%s: NULL handle or pointer '%s'", __func__, #p); } while (0)

// ---------------------------------------------------------------- shared helpers
struct msd_system;
struct msd_exchange;
struct msd_anisotropy;
struct msd_uniaxial;
struct msd_appliedfield;
struct msd_thermal;
struct msd_llg;
struct msd_dipole;
struct msd_longrange3d;
struct msd_surrogate;
struct msd_stepper;
struct msd_ensemble;
struct WireOut;
struct WireIn;
struct ExactField;
struct Batch3;

int ensure_init();
int upload_f32(float* d, const double* h, size_t n);
int download_f32(double* h, const float* d, size_t n);
bool same_dims(const Dims& a, int nx, int ny, int nz);
int copy_dev(float* dst, const float* src, size_t n);
int wire_finish(const WireOut& w, const char* who, void* buf, int64_t cap, int64_t* len);
void wire_operator(WireOut& w, int index, const Dims& d, double scale);
int unwire_operator(WireIn& r, const char* who, int index, int* nx, int* ny, int* nz, double* scale);
// one Array.Double object of the lattice's size: Array<double>::encode, after the Lua
// type and type id when tagged (encodeT<dArray>; an owner calling the array's encode
// directly writes it untagged)
void wire_darray(WireOut& w, int index, const Dims& d, const double* v, bool tagged);
int unwire_darray(WireIn& r, const char* who, int index, const Dims& d, std::vector<double>& v, bool tagged);
int twiddles(int L, float2** d32, double2** d64);
int layered_shape(msd_dipole* p, int ty, int tpb, bool row_ct, bool want_split, const char* what);
int dipole_field_of(msd_dipole* dip, const float* s3, float* h);
void items3(int rep, const Dims& d, const Conv3Dev* cv, int ty, std::vector<int4>& rows, std::vector<int4>& ys, std::vector<int4>& zs);
std::vector<int4> xcd_order(const std::vector<int4>& v);
int lr3_field_of(msd_longrange3d* l, const float* S, float* H);
int lr3_set_widths(msd_longrange3d* l, int wy, int wz);   // 0 keeps a width
int field_maxdiff(const float* a, const float* b, size_t n3, double* maxdiff, double* maxb);
int member_begin(msd_stepper* st);
bool spectrum_stale(const msd_stepper* st);
void stepper_end(msd_stepper* st, int nsteps);

// ===================================================================== device
struct Globals {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    float2* part = nullptr;           // partial maxima of the field check [512]
    size_t max_lds = 65536;
};
extern Globals G;

template <class T>
inline int dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) return fail(MSD_ENOMEM, "hipMalloc(%zu bytes): %s", sizeof(T) * count, hipGetErrorString(e));
    return MSD_OK;
}

template <class T>
inline void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// ================================================================ SpinSystem
struct Slot {
    std::string name;
    float* h = nullptr;   // [3][n]
    bool used = false;
};

struct msd_system {
    Dims d{};
    float* S = nullptr;           // [4][n] x y z ms
    float* Slo = nullptr;         // [3][n] low words of the compensated (hi + lo) spin state
    bool lo_valid = false;        // Slo holds the residual of the current spins
    double alpha = 1, gamma = 1, dt = 1, time = 0;
    float* site_alpha = nullptr;
    float* site_gamma = nullptr;
    std::vector<Slot> slots;
    double* part = nullptr;       // reduction partials
    double* out = nullptr;        // reduction result
    uint64_t version = 1;         // bumped whenever spins change outside the step driver
    ~msd_system()
    {
        dfree(S); dfree(Slo); dfree(site_alpha); dfree(site_gamma); dfree(part); dfree(out);
        for (Slot& s : slots) dfree(s.h);
    }
    int find(const char* nm) const
    {
        for (size_t i = 0; i < slots.size(); ++i)
            if (strcasecmp(slots[i].name.c_str(), nm) == 0) return (int)i;
        return -1;
    }
    // SpinSystem::register_slot_name + ensureSlotExists: zeroed on creation
    int slot(const char* nm, int* idx)
    {
        int k = find(nm);
        if (k < 0) {
            Slot s;
            s.name = nm;
            if (int rc = dalloc(&s.h, (size_t)3 * d.n)) return rc;
            hipError_t e = launch_fill(s.h, 0.0f, (size_t)3 * d.n, G.stream);
            if (e != hipSuccess) { dfree(s.h); return fail(MSD_EDEVICE, "slot fill: %s", hipGetErrorString(e)); }
            slots.push_back(s);
            k = (int)slots.size() - 1;
        }
        *idx = k;
        return MSD_OK;
    }
    int mark(const char* nm, int* idx)
    {
        if (int rc = slot(nm, idx)) return rc;
        slots[*idx].used = true;
        return MSD_OK;
    }
    LLGDev llgdev(const msd_llg* l) const;
};

// An object of the checkpoint stream opens with a magic char (111: new to the stream,
// 222: written earlier) and its index in the stream's object list (luabaseobject.h:63-72).
enum : char { WIRE_NEW = 111, WIRE_OLD = (char)222 };
enum : int { WIRE_LUA_TUSERDATA = 7 };
struct WireOut {
    std::vector<uint8_t> b;
    void put(const void* p, size_t n) { const uint8_t* q = (const uint8_t*)p; b.insert(b.end(), q, q + n); }
    void ch(char c) { put(&c, 1); }
    void i32(int32_t v) { put(&v, 4); }
    void f64(double v) { put(&v, 8); }
};

struct WireIn {
    const uint8_t* p;
    size_t n, at = 0;
    bool get(void* d, size_t k)
    {
        if (k > n - at) return false;
        std::memcpy(d, p + at, k);
        at += k;
        return true;
    }
    bool ch(char* c) { return get(c, 1); }
    bool i32(int32_t* v) { return get(v, 4); }
    bool f64(double* v) { return get(v, 8); }
};

// ================================================================== Exchange
struct msd_exchange {
    Dims d{};
    int pbc[3] = {1, 1, 1};
    bool normalize = true;
    double scale = 1;
    struct Path { int from, to; double J; };
    std::vector<Path> paths;
    bool dirty = true;
    // device forms
    int *row = nullptr, *col = nullptr;
    float* val = nullptr;
    uint8_t* cls = nullptr;
    int* cls_start = nullptr;
    ExEntry* ent = nullptr;
    bool lut_ok = false;
    int ncls = 0, nent = 0;
    ~msd_exchange() { release(); }
    void release()
    {
        (void)hipStreamSynchronize(G.stream);
        dfree(row); dfree(col); dfree(val); dfree(cls); dfree(cls_start); dfree(ent);
    }
    int compile();
    ExchangeDev csr_dev() const
    {
        ExchangeDev e{};
        e.mode = paths.empty() ? 0 : 2;
        e.normalize = normalize;
        e.row = row; e.col = col; e.val = val;
        return e;
    }
    ExchangeDev fused_dev() const
    {
        ExchangeDev e = csr_dev();
        // the fused kernels stage class tables in LDS: larger ones take the CSR form
        if (lut_ok && !paths.empty() && ncls + 1 <= EX_LDS_CLS && nent <= EX_LDS_ENT) {
            e.mode = 1; e.cls = cls; e.cls_start = cls_start; e.ent = ent;
            e.ncls = ncls; e.nent = nent;
        }
        return e;
    }
};

// ================================================================ Anisotropy
struct msd_anisotropy {
    Dims d{};
    double scale = 1;
    struct Op { int site; double ax, ay, az, K; };
    std::vector<Op> ops;
    bool dirty = true;
    float4* axk = nullptr;
    float* mat = nullptr;
    AnisoDev dev{};
    ~msd_anisotropy() { (void)hipStreamSynchronize(G.stream); dfree(axk); dfree(mat); }
    int compile();
};

// ======================================================= Anisotropy.Uniaxial
// AnisotropyUniaxial (cpu/anisotropy/spinoperationanisotropy_uniaxial.cpp): ops of
// (site, unit axis, K1, K2) appended by add (:158-179), applied per site in order of
// addition; the device form is a site-sorted CSR of the ops (stable, so each site
// keeps its order), rebuilt when ops change.
struct msd_uniaxial {
    Dims d{};
    double scale = 1;
    struct op { int site; double n[3]; double K1, K2; };
    std::vector<op> ops;
    int* start = nullptr;     // [n+1]
    float* dops = nullptr;    // [nops][5]
    bool dirty = true;
    ~msd_uniaxial()
    {
        (void)hipStreamSynchronize(G.stream);
        dfree(start); dfree(dops);
    }
    int compile()
    {
        if (!dirty) return MSD_OK;
        const int n = d.n;
        std::vector<int> cnt(n + 1, 0);
        for (const op& o : ops) cnt[o.site + 1]++;
        for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
        std::vector<float> packed(5 * std::max<size_t>(ops.size(), 1), 0.0f);
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (const op& o : ops) {
            float* p = &packed[5 * (size_t)fill[o.site]++];
            p[0] = (float)o.n[0]; p[1] = (float)o.n[1]; p[2] = (float)o.n[2]; p[3] = (float)o.K1; p[4] = (float)o.K2;
        }
        dfree(start); dfree(dops);
        if (int rc = dalloc(&start, (size_t)n + 1)) return rc;
        if (int rc = dalloc(&dops, packed.size())) return rc;
        HIPCHK(hipMemcpyAsync(start, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice, G.stream));
        HIPCHK(hipMemcpyAsync(dops, packed.data(), sizeof(float) * packed.size(), hipMemcpyHostToDevice, G.stream));
        HIPCHK(hipStreamSynchronize(G.stream));
        dirty = false;
        return MSD_OK;
    }
};

// ============================================================= AppliedField
// AppliedField (cpu/appliedfield/spinoperationappliedfield.cpp): its Lua apply
// (appliedfield_luafuncs.lua:46-57) overwrites slot "Zeeman" with B*scale at every
// site.  AppliedField.Heterogeneous::apply
// (spinoperationappliedfield_heterogeneous.cpp:88-99) adds scale*h(site) to slot
// "Zeeman_Heterogeneous".  The step driver folds either into its site kernels.
struct msd_appliedfield {
    Dims d{};
    bool het = false;
    double B[3] = {0, 0, 0};
    double scale = 1;
    std::vector<float> h;    // [3][n], heterogeneous form
    float* dh = nullptr;     // device copy of h
    bool dirty = false;
    ~msd_appliedfield()
    {
        (void)hipStreamSynchronize(G.stream);
        dfree(dh);
    }
    int upload()
    {
        if (!het || !dirty) return MSD_OK;
        if (!dh)
            if (int rc = dalloc(&dh, h.size())) return rc;
        HIPCHK(hipMemcpyAsync(dh, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, G.stream));
        HIPCHK(hipStreamSynchronize(G.stream));
        dirty = false;
        return MSD_OK;
    }
};

// =================================================================== Thermal
struct msd_thermal {
    Dims d{};
    double T = 0, scale = 1;
    std::vector<double> site;
    bool site_dirty = false;
    float* dsite = nullptr;
    uint64_t seed = 0, draw = 0;
    ~msd_thermal() { (void)hipStreamSynchronize(G.stream); dfree(dsite); }
    int dev(ThermalDev& t)
    {
        if (site_dirty) {
            bool ones = true;
            for (double v : site) ones &= v == 1.0;
            if (ones) { (void)hipStreamSynchronize(G.stream); dfree(dsite); }
            else {
                if (!dsite)
                    if (int rc = dalloc(&dsite, d.n)) return rc;
                if (int rc = upload_f32(dsite, site.data(), d.n)) return rc;
            }
            site_dirty = false;
        }
        t.on = 1;
        t.T_scale = (float)(T * scale);
        t.site = dsite;
        t.seed = seed;
        return MSD_OK;
    }
};

// ======================================================================= LLG
struct msd_llg {
    bool thermal_first = true;
    bool no_renorm = false;
    bool quaternion = false;    // LLG.Quaternion
    double ps = 1, ds = 1;
};

// ==================================================================== Dipole
struct msd_dipole {
    Dims d{};
    int nxh = 0;
    double g = 1, scale = 1;
    int gmax[4] = {1000, 1000, 1000, 1000};
    double ABC[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    bool loaded = false, custom = false, spectrum = false;
    // bumped by every op-API transform through this object's spectrum buffers and by
    // every step driver's steps (which record the value), so a step driver sharing
    // them knows whether its cached spin spectrum is still there
    uint64_t epoch = 0;
    double words_mult = NAN;
    int words_split = -1;     // dev.split the words were laid out for
    double* T64 = nullptr;    // [6][2nz-1][ny][nx]
    double2* U64 = nullptr;
    double* W64 = nullptr;    // [6][nz][nxh][ny]
    double2 *twx64 = nullptr, *twy64 = nullptr;
    float* Tw = nullptr;
    float2 *Mh = nullptr, *Hh = nullptr, *twx = nullptr, *twy = nullptr;
    DipoleDev dev{};
    // Per-layer-pair form (LongRange2D and the Dipole2D builder): the same row passes,
    // a column pass with a full complex tensor per ordered layer pair.
    bool pair = false;
    std::string slot = "Dipole";
    std::vector<double> P;    // real space [9][nz][nz][ny][nx]: XX..ZZ, labels [first][second]
    bool pair_dirty = true;   // P changed since the words were made (LongRange2D compileRequired)
    std::vector<double> sms, dfs;     // source-moment / destination-field scale per layer
    float lsc_host[2 * MSD_MAX_LAYERS] = {};
    bool lsc_valid = false;
    float2* Tp = nullptr;     // [9][nz][nz][nxh][ny] words
    float* lsc = nullptr;     // [2][MSD_MAX_LAYERS]
    ~msd_dipole()
    {
        (void)hipStreamSynchronize(G.stream);
        dfree(T64); dfree(U64); dfree(W64); dfree(twx64); dfree(twy64);
        dfree(Tw); dfree(Mh); dfree(Hh); dfree(twx); dfree(twy);
        dfree(Tp); dfree(lsc);
    }
    int prepare();
};

// =============================================================== LongRange3D
// LongRange3D (cpu/longrange3d/spinoperationlongrange3d.cpp:331-432) with the
// dipole tensor built in FP64 on the device: the periodic lattice sum to radius R
// (the 3D analogue of Dipole's truncated sum), or, with open boundaries, the
// bare dipole kernel on the doubled FFT grid.
struct msd_longrange3d {
    Dims d{};
    bool open = false;
    int fx = 0, fy = 0, fz = 0, fxh = 0;
    double g = 1, scale = 1;
    int R = 1000;
    bool loaded = false, custom = false, spectrum = false;
    double words_mult = NAN;
    double* T64 = nullptr;   // [6][fz][fy][fx]
    float* Tw = nullptr;     // [6][fys][fxh][fzs], fys <= fy, fzs <= fz
    float2* W = nullptr;     // working spectrum [3][nz][fy][fxp]
    float2 *twx = nullptr, *twy = nullptr, *twz = nullptr;
    double2 *twx64 = nullptr, *twy64 = nullptr, *twz64 = nullptr;
    FFTPlan px{}, py{}, pz{};
    int ty = 1, wy = 1, wz = 1, zt = 0, yt = 0, rt = 0, fzs = 0, fys = 0, fxp = 0;
    size_t lds_rows = 0, lds_y = 0, lds_z = 0;
    // single-replica descriptor and work items for the operator API
    Rep3* orep = nullptr;
    int4 *oitems_r = nullptr, *oitems_y = nullptr, *oitems_z = nullptr;
    int onr = 0, ony = 0, onz = 0;
    uint64_t epoch = 0;      // bumped whenever W changes hands: field evaluations, a driver's steps
    ~msd_longrange3d()
    {
        (void)hipStreamSynchronize(G.stream);
        dfree(T64); dfree(Tw); dfree(W); dfree(twx); dfree(twy); dfree(twz);
        dfree(twx64); dfree(twy64); dfree(twz64); dfree(orep); dfree(oitems_r); dfree(oitems_y); dfree(oitems_z);
    }
    int prepare();
    Conv3Dev conv() const
    {
        Conv3Dev c{};
        c.fx = fx; c.fy = fy; c.fz = fz; c.fxh = fxh; c.fxp = fxp;
        c.ty = ty; c.wy = wy; c.wz = wz; c.zt = zt; c.yt = yt; c.rt = rt; c.fzs = fzs; c.fys = fys;
        c.px = px; c.py = py; c.pz = pz;
        c.twx = twx; c.twy = twy; c.twz = twz;
        c.W = W; c.Tw = Tw;
        return c;
    }
};

template <class T>
inline int upload_vec(T** dst, const std::vector<T>& v)
{
    dfree(*dst);
    if (v.empty()) return MSD_OK;
    if (int rc = dalloc(dst, v.size())) return rc;
    HIPCHK(hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, G.stream));
    return MSD_OK;
}

// ------------------------------------------------------------ batch of replicas
// Device descriptors and work tables of replicas stepped by the same launches.
struct Batch3 {
    // the row, y and z passes run once per pass form among the members (the
    // compile-time pass for each length that has one, the runtime pass for the
    // rest), so every member is transformed by the same kernels as when it runs alone
    struct ZGroup {
        int zt = 0;
        int w = 0;           // kx width of the y / z pass form (0 for row groups)
        int4* items = nullptr;
        int n = 0;
        size_t lds = 0;
    };
    Rep3* reps = nullptr;
    int nrep = 0;
    std::vector<ZGroup> rg, yg, zg;      // row, y-pass and z-pass groups (key: Conv3Dev::rt / (yt, wy) / (zt, wz))
    bool any_lr = false;
    uint64_t gen = 0;                    // bumped whenever the tables are rebuilt (graphs holding them are stale)
    std::vector<Rep3> host;
    ~Batch3() { release(); }
    void release()
    {
        (void)hipStreamSynchronize(G.stream);
        dfree(reps);
        for (ZGroup& z : rg) dfree(z.items);
        for (ZGroup& z : yg) dfree(z.items);
        for (ZGroup& z : zg) dfree(z.items);
        rg.clear(); yg.clear(); zg.clear();
        host.clear();
        nrep = 0;
    }
    // whether member b launches the same pass forms over the same work items as a
    static bool same_passes(const Rep3& a, const Rep3& b)
    {
        return a.has_lr == b.has_lr && a.a.d.nx == b.a.d.nx && a.a.d.ny == b.a.d.ny && a.a.d.nz == b.a.d.nz &&
               a.a.ty == b.a.ty && a.cv.rt == b.cv.rt && a.cv.yt == b.cv.yt && a.cv.wy == b.cv.wy &&
               a.cv.zt == b.cv.zt && a.cv.wz == b.cv.wz && a.cv.fxh == b.cv.fxh && a.cv.fy == b.cv.fy;
    }
    // rebuild the tables when the member set or a member's pass forms changed;
    // refresh descriptors always
    int update(const std::vector<Rep3>& r, const std::vector<int>& tys, const std::vector<size_t>& lds3)
    {
        bool same_shape = host.size() == r.size() && reps != nullptr;
        for (size_t i = 0; same_shape && i < r.size(); ++i) same_shape = same_passes(host[i], r[i]);
        if (!same_shape) {
            release();
            ++gen;
            std::vector<std::vector<int4>> R, Y, Z;
            std::vector<ZGroup> rgroups, ygroups, zgroups;
            any_lr = false;
            // append a member's pass items to the group of its pass form
            auto add = [](std::vector<ZGroup>& groups, std::vector<std::vector<int4>>& lists, int key, int w,
                          size_t lds, const std::vector<int4>& items) {
                if (items.empty()) return;
                size_t gi = 0;
                while (gi < groups.size() && (groups[gi].zt != key || groups[gi].w != w)) ++gi;
                if (gi == groups.size()) {
                    groups.push_back(ZGroup{});
                    groups.back().zt = key;
                    groups.back().w = w;
                    lists.emplace_back();
                }
                groups[gi].lds = std::max(groups[gi].lds, lds);
                lists[gi].insert(lists[gi].end(), items.begin(), items.end());
            };
            for (size_t i = 0; i < r.size(); ++i) {
                std::vector<int4> ri, yi, zi;
                items3((int)i, r[i].a.d, r[i].has_lr ? &r[i].cv : nullptr, tys[i], ri, yi, zi);
                any_lr |= r[i].has_lr != 0;
                // the runtime-length passes read each member's width from its descriptor
                add(rgroups, R, r[i].has_lr ? r[i].cv.rt : 0, 0, lds3[3 * i], ri);
                add(ygroups, Y, r[i].cv.yt, r[i].cv.yt ? r[i].cv.wy : 0, lds3[3 * i + 1], yi);
                add(zgroups, Z, r[i].cv.zt, r[i].cv.zt ? r[i].cv.wz : 0, lds3[3 * i + 2], zi);
            }
            for (size_t gi = 0; gi < rgroups.size(); ++gi) {
                if (int rc = upload_vec(&rgroups[gi].items, R[gi])) return rc;
                rgroups[gi].n = (int)R[gi].size();
            }
            for (size_t gi = 0; gi < ygroups.size(); ++gi) {
                if (int rc = upload_vec(&ygroups[gi].items, xcd_order(Y[gi]))) return rc;
                ygroups[gi].n = (int)Y[gi].size();
            }
            for (size_t gi = 0; gi < zgroups.size(); ++gi) {
                if (int rc = upload_vec(&zgroups[gi].items, xcd_order(Z[gi]))) return rc;
                zgroups[gi].n = (int)Z[gi].size();
            }
            rg = rgroups;
            yg = ygroups;
            zg = zgroups;
            if (int rc = dalloc(&reps, r.size())) return rc;
            nrep = (int)r.size();
        }
        if (!r.empty() && (!same_shape || std::memcmp(host.data(), r.data(), sizeof(Rep3) * r.size()) != 0)) {
            host = r;
            HIPCHK(hipMemcpyAsync(reps, host.data(), sizeof(Rep3) * host.size(), hipMemcpyHostToDevice, G.stream));
            HIPCHK(hipStreamSynchronize(G.stream));
        }
        return MSD_OK;
    }
    hipError_t zpass(hipStream_t st) const
    {
        hipError_t e = hipSuccess;
        for (const ZGroup& z : zg)
            if (e == hipSuccess) e = launch3_zconv(reps, z.items, z.n, z.lds, z.zt, z.w, st);
        return e;
    }
    hipError_t ypass(int sign, hipStream_t st) const
    {
        hipError_t e = hipSuccess;
        for (const ZGroup& y : yg)
            if (e == hipSuccess) e = launch3_ypass(sign, reps, y.items, y.n, y.lds, y.zt, y.w, st);
        return e;
    }
    hipError_t rows(int row_mode, bool comp, hipStream_t st) const
    {
        hipError_t e = hipSuccess;
        for (const ZGroup& g : rg)
            if (e == hipSuccess) e = launch3_rows(row_mode, comp, reps, g.items, g.n, g.lds, g.zt, st);
        return e;
    }
    hipError_t evaluate(int row_mode, bool comp, hipStream_t st) const
    {
        hipError_t e = hipSuccess;
        if (any_lr) {
            e = ypass(-1, st);
            if (e == hipSuccess) e = zpass(st);
            if (e == hipSuccess) e = ypass(1, st);
        }
        if (e == hipSuccess) e = rows(row_mode, comp, st);
        return e;
    }
};

// ============================================================ step driver
// Exact steps per span graph (msd_stepper::mgraph, msd_ensemble::mgraph): a graph
// launch costs about 4.6 us between steps on MI355X, so 256x256x8 film steps of 64 us
// replayed one graph per step run 7% slower than spans of 16.
constexpr int STEP_SPAN = 16;

// The step closure of PredictorCorrector.lua with one corrector pass, as fused
// kernels.  Three forms: local fields only, layered dipole (Dipole), or the 3D
// convolution (LongRange3D); a stepper of the 3D form is a batch of one.
enum StepKind { SK_LOCAL = 0, SK_LAYERED = 1, SK_CONV3 = 2 };

// ============================================================ surrogate
// Dipolar-field surrogate: three 5x5x3 convolution layers (3 -> 16 -> 16 -> 3
// channels, BF16 on MFMA, FP32 accumulate) mapping s_in (m - m_ref) to the
// change of the layered dipole field, dH = s_out Y3.  FP32 master weights
// [layer][tap][out][in] (16 x 16 per tap, zero beyond a layer's channels) and
// their Adam moments live on the device; the packed MFMA fragments of every
// layer and of its adjoint are rebuilt after each update.
struct msd_surrogate {
    Dims d{};
    CnnDims cd{};
    int grid = 0;                                                        // weight-gradient workgroups
    CnnWalk walk[3]{};                                                   // row-walk items per shape
    int rgrid[3] = {0, 0, 0};                                            // row-walk workgroups per shape
    static constexpr int NT = 75, WSZ = 75 * 256, PACK = 45 * 64 * 8;
    float *W = nullptr, *Mo = nullptr, *Vo = nullptr, *gW = nullptr;     // [3][WSZ]
    uint16_t *Af = nullptr, *At = nullptr;                               // [3][PACK]
    uint16_t *X0 = nullptr, *X1 = nullptr, *X2 = nullptr;                // activations [n][4], [n][16], [n][16]
    uint16_t *G1 = nullptr, *G2 = nullptr, *G3 = nullptr;                // back-propagated [n][16], [n][16], [n][4]
    uint16_t* zeros = nullptr;                                           // 16 zero bytes
    float *Y3 = nullptr, *T = nullptr, *D = nullptr;                     // [3][n]
    float* partial = nullptr;                                            // [grid][WSZ]
    double* lpart = nullptr;
    int nlpart = 0;
    double s_in = 1.0, s_out = 1.0, sigma2 = 1.0;
    int adam_t = 0;
    bool packed = false;
    int cus = 0;
    ~msd_surrogate()
    {
        (void)hipStreamSynchronize(G.stream);
        dfree(W); dfree(Mo); dfree(Vo); dfree(gW); dfree(Af); dfree(At);
        dfree(X0); dfree(X1); dfree(X2); dfree(G1); dfree(G2); dfree(G3);
        dfree(Y3); dfree(T); dfree(D); dfree(partial); dfree(lpart); dfree(zeros);
    }
    static int cin(int l) { return l == 0 ? 3 : 16; }
    static int cout(int l) { return l == 2 ? 3 : 16; }
    // Row-walk items: 128 x ty sites of one layer, or 64 x ty sites of every layer
    // of a film of 1, 2, 4 or 8 layers (periodic x and y only).  A workgroup walks
    // ty + 4 input rows per item; per layer shape pick ty (at most 32) for the
    // fewest input rows per resident workgroup.
    void plan_walks()
    {
        const int nx = d.nx, ny = d.ny, nz = d.nz;
        const bool film = cnn_film_supported(nz) && !cd.open;
        for (int v = 0; v < 3; ++v) {
            CnnWalk& wk = walk[v];
            wk.film = film ? 1 : 0;
            wk.rtiles_x = film ? (nx + 63) / 64 : (nx + 127) / 128;
            const int zitems = film ? 1 : nz;
            double best = 0.0;
            const long slots = (long)cnn_rows_occupancy(v, film ? nz : 0) * cus;
            for (int ty = std::min(ny, cnn_rows_ty_max()); ty >= 1; --ty) {
                const long items = (long)wk.rtiles_x * zitems * ((ny + ty - 1) / ty);
                const double cost = (double)((items + slots - 1) / slots) * (ty + 4);
                if (best == 0.0 || cost < best) { best = cost; wk.ty = ty; }
            }
            wk.rtiles_y = (ny + wk.ty - 1) / wk.ty;
            wk.rtiles = wk.rtiles_x * wk.rtiles_y * zitems;
            rgrid[v] = (int)std::min<long>(wk.rtiles, slots);
        }
    }
    // row-walk shape of each layer and of the adjoints that back-propagation runs
    static constexpr int FWD_SHAPE[3] = {CV_IN3, CV_F16, CV_OUT3};
    static constexpr int ADJ_SHAPE[3] = {-1, CV_F16, CV_IN3};
    int pack()
    {
        if (packed) return MSD_OK;
        for (int l = 0; l < 3; ++l) {
            hipError_t e = launch_cnn_pack(W + (size_t)l * WSZ, 0, FWD_SHAPE[l], Af + (size_t)l * PACK, G.stream);
            if (e == hipSuccess && ADJ_SHAPE[l] >= 0)
                e = launch_cnn_pack(W + (size_t)l * WSZ, 1, ADJ_SHAPE[l], At + (size_t)l * PACK, G.stream);
            if (e != hipSuccess) return fail(MSD_EDEVICE, "surrogate: %s", hipGetErrorString(e));
        }
        packed = true;
        return MSD_OK;
    }
    hipError_t rows(int shape, const uint16_t* in, const uint16_t* A, uint16_t* out_bf, float* out_f, float scale,
                    hipStream_t st) const
    {
        return launch_cnn_rows(shape, cd, walk[shape], rgrid[shape], in, zeros, A, out_bf, out_f, scale, st);
    }
    // X0 -> dH = s_out * L3(L2(L1(X0))) (FP32 [3][n] into out), plus href if given
    int forward(float* out, const float* href, hipStream_t st)
    {
        if (int rc = pack()) return rc;
        hipError_t e = rows(CV_IN3, X0, Af, X1, nullptr, 1.0f, st);
        if (e == hipSuccess) e = rows(CV_F16, X1, Af + PACK, X2, nullptr, 1.0f, st);
        if (e == hipSuccess) e = rows(CV_OUT3, X2, Af + 2 * PACK, nullptr, out, (float)s_out, st);
        if (e == hipSuccess && href) e = launch_cnn_add((size_t)3 * d.n, out, href, out, st);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "surrogate forward: %s", hipGetErrorString(e));
        return MSD_OK;
    }
};

// The exact field operator a surrogate stands in for: a layered Dipole or a
// LongRange3D (exactly one is set).
struct ExactField {
    msd_dipole* dip;
    msd_longrange3d* lr3;
    int of(const float* S, float* H) const { return dip ? dipole_field_of(dip, S, H) : lr3_field_of(lr3, S, H); }
    bool fits(const Dims& d) const
    {
        return dip ? same_dims(d, dip->d.nx, dip->d.ny, dip->d.nz) : same_dims(d, lr3->d.nx, lr3->d.ny, lr3->d.nz);
    }
};

struct msd_stepper {
    msd_system* s = nullptr;
    msd_exchange* ex = nullptr;
    msd_anisotropy* an = nullptr;
    msd_dipole* dip = nullptr;
    msd_longrange3d* lr3 = nullptr;
    msd_thermal* th = nullptr;
    msd_llg llg;
    float *P = nullptr, *C = nullptr, *Clo = nullptr;
    unsigned long long* draw = nullptr;
    bool use_graph = true;
    bool compensated = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    StepArgs args{};
    StepArgs graph_args{};
    // The fused step kernels of each graph ([0] exact, [1] a learned window's first
    // step, [2] later learned steps) and which arguments they take (0 exact, 1 a
    // learned window's first predictor, 2 the other learned passes): when only values
    // change (temp:set(T), setTimeStep) the nodes take them in place, no new capture.
    struct StepNode {
        hipGraphNode_t node;
        int variant;
    };
    std::vector<StepNode> gnodes[4];                    // [3]: the span graph's
    bool gnodes_ok[4] = {false, false, false, false};   // every step launch of the graph found
    std::vector<std::pair<const void*, int>> cap_log;   // step kernels launched during a capture
    bool capturing = false;
    long graph_builds = 0;               // graphs captured and instantiated (diagnostics)
    // Runge-Kutta form (msd_stepper_rk): stage spins k_1.., stage fields, its graph
    std::vector<float*> rk_k, rk_h;
    float* rk_tmp = nullptr;
    hipGraph_t rk_graph = nullptr;
    hipGraphExec_t rk_exec = nullptr;
    std::string rk_kind;
    StepArgs rk_args{};
    Batch3 b3;
    uint64_t spec_version = 0;
    uint64_t dip_epoch = 0, lr3_epoch = 0;
    // learned dipole field (msd_stepper_set_surrogate)
    msd_surrogate* sur = nullptr;
    int sur_K = 10;
    double sur_tol = 1e-3;
    float *Href = nullptr, *Mref = nullptr, *Hext = nullptr, *Hx = nullptr;   // [3][n]
    float* Ssnap = nullptr;                   // [3][n] low words of a window's starting state
    std::vector<msd_appliedfield*> af;                                        // applied fields
    msd_uniaxial* uni = nullptr;                                              // Anisotropy.Uniaxial
    float* Happ = nullptr;                                                    // [3][n] their per-site sum
    bool have_ref = false;
    long sur_windows = 0, sur_exact_windows = 0, sur_fallbacks = 0;
    double sur_last_res = 0.0, sur_max_res = 0.0;
    // span graph: STEP_SPAN exact steps captured in one graph, replayed for runs of
    // at least that many steps (one graph launch's overhead per span, not per step)
    hipGraph_t mgraph = nullptr;
    hipGraphExec_t mexec = nullptr;
    StepArgs mgraph_args{};
    // learned-step graphs: [0] a window's first step, [1] the others
    hipGraph_t lgraph[2] = {nullptr, nullptr};
    hipGraphExec_t lexec[2] = {nullptr, nullptr};
    StepArgs lgraph_args[2]{};
    double lgraph_sin[2] = {0.0, 0.0}, lgraph_sout[2] = {0.0, 0.0};   // surrogate scales baked into them
    ~msd_stepper()
    {
        (void)hipStreamSynchronize(G.stream);
        drop_graph();
        dfree(P); dfree(C); dfree(Clo); dfree(draw);
        for (float*& b : rk_k) dfree(b);
        for (float*& b : rk_h) dfree(b);
        dfree(rk_tmp);
        dfree(Href); dfree(Mref); dfree(Hext); dfree(Hx); dfree(Ssnap); dfree(Happ);
    }
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr; graph = nullptr;
        gnodes[0].clear();
        gnodes_ok[0] = false;
        if (mexec) (void)hipGraphExecDestroy(mexec);
        if (mgraph) (void)hipGraphDestroy(mgraph);
        mexec = nullptr; mgraph = nullptr;
        gnodes[3].clear();
        gnodes_ok[3] = false;
        drop_learned_graphs();
        if (rk_exec) (void)hipGraphExecDestroy(rk_exec);
        if (rk_graph) (void)hipGraphDestroy(rk_graph);
        rk_exec = nullptr; rk_graph = nullptr;
    }
    void drop_learned_graphs()
    {
        for (int k = 0; k < 2; ++k) {
            if (lexec[k]) (void)hipGraphExecDestroy(lexec[k]);
            if (lgraph[k]) (void)hipGraphDestroy(lgraph[k]);
            lexec[k] = nullptr; lgraph[k] = nullptr;
            gnodes[k + 1].clear();
            gnodes_ok[k + 1] = false;
        }
    }
    int kind() const { return lr3 ? SK_CONV3 : (dip ? SK_LAYERED : SK_LOCAL); }
    int refresh();
    Rep3 rep() const
    {
        Rep3 r{};
        r.a = args;
        if (lr3) {
            r.cv = lr3->conv();
            r.has_lr = 1;
        } else {
            r.cv.fx = s->d.nx;
            r.cv.ty = args.ty;
        }
        r.cv.sin = s->S;
        return r;
    }
    int launch_step(hipStream_t st);
    int launch_learned_step(hipStream_t st, bool window_start);
    // arguments of the learned passes: no dipole term in the fused kernels, the
    // surrogate's field (Hext, absent at a window's first predictor) and H_ref added
    StepArgs learned_args(bool first) const
    {
        StepArgs a = args;
        a.has_dip = 0;
        a.Hext = first ? nullptr : Hext;
        a.Href = Href;
        return a;
    }
    StepArgs variant_args(int v) const { return v == 0 ? args : learned_args(v == 1); }
    void log_step(const void* fn, int variant)
    {
        if (capturing) cap_log.push_back({fn, variant});
    }
};

// ================================================================= ensemble
// Independent replicas (uneven sizes, own temperatures and seeds) advanced by
// one set of launches per step: every member is a step driver of the local or
// the 3D form; its descriptor is one entry of the batch.
// Films (layered-Dipole members) of one shape -- dimensions, row tiles, column pass and
// state form -- step together: each pass is one launch over all of them, the films'
// arguments and dipole descriptors read from device tables.
struct FilmGroup {
    std::vector<msd_stepper*> mem;
    std::vector<StepArgs> hargs;
    std::vector<DipoleDev> hdps;
    StepArgs* dargs = nullptr;
    DipoleDev* ddps = nullptr;
    static bool same_shape(const StepArgs& a, const StepArgs& b)
    {
        return std::memcmp(&a.d, &b.d, sizeof a.d) == 0 && a.ty == b.ty && (a.Slo == nullptr) == (b.Slo == nullptr) &&
               (a.dp.Tp == nullptr) == (b.dp.Tp == nullptr) && a.dp.tpb == b.dp.tpb && a.dp.ty == b.dp.ty && a.dp.lds_rows == b.dp.lds_rows &&
               a.dp.lds_cols == b.dp.lds_cols && a.dp.split == b.dp.split && a.dp.ly == b.dp.ly &&
               std::memcmp(&a.dp.px, &b.dp.px, sizeof a.dp.px) == 0 && std::memcmp(&a.dp.py, &b.dp.py, sizeof a.dp.py) == 0;
    }
    void release() { dfree(dargs); dfree(ddps); }
    // the tables into device memory when a value changed (not inside a capture)
    int upload()
    {
        std::vector<StepArgs> a;
        std::vector<DipoleDev> d;
        for (msd_stepper* s : mem) { a.push_back(s->args); d.push_back(s->args.dp); }
        if (!dargs) {
            if (int rc = dalloc(&dargs, mem.size())) return rc;
            if (int rc = dalloc(&ddps, mem.size())) return rc;
        }
        if (a.size() != hargs.size() || std::memcmp(a.data(), hargs.data(), sizeof(StepArgs) * a.size()) != 0 ||
            std::memcmp(d.data(), hdps.data(), sizeof(DipoleDev) * d.size()) != 0) {
            hargs = a;
            hdps = d;
            HIPCHK(hipMemcpyAsync(dargs, hargs.data(), sizeof(StepArgs) * hargs.size(), hipMemcpyHostToDevice, G.stream));
            HIPCHK(hipMemcpyAsync(ddps, hdps.data(), sizeof(DipoleDev) * hdps.size(), hipMemcpyHostToDevice, G.stream));
            HIPCHK(hipStreamSynchronize(G.stream));
        }
        return MSD_OK;
    }
    hipError_t step(bool pred, hipStream_t st) const
    {
        const int n = (int)mem.size();
        hipError_t e = launch_cols_batch(hargs[0].dp, ddps, n, st);
        if (e == hipSuccess) e = pred ? launch_step_pred_batch(hargs[0], dargs, n, st) : launch_step_corr_batch(hargs[0], dargs, n, st);
        return e;
    }
};

// A driver's state (spins, low words, time, thermal draw) kept across autotuner
// trials and put back after each, so that tuning leaves the trajectory untouched.
struct StateKeep {
    msd_stepper* st;
    float *S0 = nullptr, *L0 = nullptr;
    double time0 = 0.0;
    uint64_t draw0 = 0;
    bool lo0 = false;
    explicit StateKeep(msd_stepper* s) : st(s) {}
    StateKeep(const StateKeep&) = delete;
    StateKeep& operator=(const StateKeep&) = delete;
    ~StateKeep()
    {
        (void)hipStreamSynchronize(G.stream);
        dfree(S0); dfree(L0);
    }
    int save()
    {
        const size_t n = st->s->d.n;
        if (int rc = dalloc(&S0, 4 * n)) return rc;
        if (int rc = dalloc(&L0, 3 * n)) return rc;
        time0 = st->s->time;
        draw0 = st->th ? st->th->draw : 0;
        lo0 = st->s->lo_valid;
        if (int rc = copy_dev(S0, st->s->S, 4 * n)) return rc;
        return copy_dev(L0, st->s->Slo, 3 * n);
    }
    int restore()
    {
        const size_t n = st->s->d.n;
        if (int r = copy_dev(st->s->S, S0, 4 * n)) return r;
        if (int r = copy_dev(st->s->Slo, L0, 3 * n)) return r;
        st->s->time = time0;
        st->s->lo_valid = lo0;
        if (st->th) st->th->draw = draw0;
        st->s->version++;
        return MSD_OK;
    }
};

struct msd_ensemble {
    std::vector<msd_stepper*> m;
    bool compensated = true;
    bool use_graph = true;
    Batch3 b;                        // local-field and LongRange3D members
    std::vector<FilmGroup> films;    // layered-Dipole members, by shape
    hipGraph_t graph = nullptr;      // one step
    hipGraphExec_t exec = nullptr;
    hipGraph_t mgraph = nullptr;     // STEP_SPAN steps
    hipGraphExec_t mexec = nullptr;
    ~msd_ensemble()
    {
        (void)hipStreamSynchronize(G.stream);
        drop_graph();
        for (FilmGroup& f : films) f.release();
    }
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr; graph = nullptr;
        if (mexec) (void)hipGraphExecDestroy(mexec);
        if (mgraph) (void)hipGraphDestroy(mgraph);
        mexec = nullptr; mgraph = nullptr;
    }
    int refresh()
    {
        std::vector<Rep3> r;
        std::vector<int> tys;
        std::vector<size_t> lds;
        std::vector<msd_stepper*> layered;
        for (msd_stepper* s : m) {
            s->compensated = compensated;
            if (int rc = s->refresh()) return rc;
            if (s->kind() == SK_LAYERED) { layered.push_back(s); continue; }
            r.push_back(s->rep());
            tys.push_back(s->args.ty);
            if (s->lr3) { lds.push_back(s->lr3->lds_rows); lds.push_back(s->lr3->lds_y); lds.push_back(s->lr3->lds_z); }
            else { lds.push_back(0); lds.push_back(0); lds.push_back(0); }
        }
        // films by shape; a shape change (compensated state, autotuned tiles) regroups
        std::vector<std::vector<msd_stepper*>> groups;
        for (msd_stepper* s : layered) {
            size_t gi = 0;
            while (gi < groups.size() && !FilmGroup::same_shape(groups[gi][0]->args, s->args)) ++gi;
            if (gi == groups.size()) groups.emplace_back();
            groups[gi].push_back(s);
        }
        bool regroup = groups.size() != films.size();
        for (size_t gi = 0; !regroup && gi < groups.size(); ++gi) regroup = groups[gi] != films[gi].mem;
        if (regroup) {
            drop_graph();
            for (FilmGroup& f : films) f.release();
            films.assign(groups.size(), FilmGroup{});
            for (size_t gi = 0; gi < groups.size(); ++gi) films[gi].mem = groups[gi];
        }
        for (FilmGroup& f : films)
            if (int rc = f.upload()) return rc;
        const uint64_t gen0 = b.gen;
        if (int rc = b.update(r, tys, lds)) return rc;
        if (b.gen != gen0) drop_graph();    // the captured launches held the old work tables
        return MSD_OK;
    }
    hipError_t evaluate(bool pred, hipStream_t st) const
    {
        hipError_t e = b.evaluate(pred ? R3_PRED : R3_CORR, compensated, st);
        for (const FilmGroup& f : films)
            if (e == hipSuccess) e = f.step(pred, st);
        return e;
    }
};
