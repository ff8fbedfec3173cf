// C ABI of libmsd.so: MagLua's hot-path objects (SpinSystem, Exchange,
// Anisotropy, Thermal, Dipole, LLG.Cartesian) held on one MI355X and a fused
// Heun step driver.  Host-side state is FP64; device state FP32 SoA.
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

// ===================================================================== errors
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(MSD_EDEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                         \
    } while (0)

#define NONNULL(p) \
    do { if (!(p)) return fail(MSD_EINVAL, "%s: NULL handle or pointer '%s'", __func__, #p); } while (0)

// ===================================================================== device
static struct {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    float2* part = nullptr;           // partial maxima of the field check [512]
    size_t max_lds = 65536;
} G;

static int ensure_init()
{
    if (G.ready) return MSD_OK;
    return msd_init(0);
}

// ============================================================ allocation-free
// Scratch memory of msd_net_moment_noalloc: the FP64 block partials of the moment
// reduction (6 per block, at most 1024 blocks) followed by the 6 folded sums.  Each
// GPU holds its own copy for the life of the process.  In host code `scratch` is a
// host-side stand-in: its device address, per device, comes from hipGetSymbolAddress
// in msd_init and is the only form passed to kernels.
static constexpr int NOALLOC_PART_BLOCKS = 1024;
static constexpr int NOALLOC_DOUBLES = 6 * NOALLOC_PART_BLOCKS + 8;
static constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
static double* g_scratch_dev[NOALLOC_MAX_DEVICES];

template <class T>
static int dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) return fail(MSD_ENOMEM, "hipMalloc(%zu bytes): %s", sizeof(T) * count, hipGetErrorString(e));
    return MSD_OK;
}

template <class T>
static void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// upload FP64 host data as FP32 device data
static int upload_f32(float* d, const double* h, size_t n)
{
    std::vector<float> t(n);
    for (size_t i = 0; i < n; ++i) t[i] = (float)h[i];
    HIPCHK(hipMemcpyAsync(d, t.data(), n * sizeof(float), hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    return MSD_OK;
}

static int download_f32(double* h, const float* d, size_t n)
{
    std::vector<float> t(n);
    HIPCHK(hipMemcpyAsync(t.data(), d, n * sizeof(float), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    for (size_t i = 0; i < n; ++i) h[i] = t[i];
    return MSD_OK;
}

extern "C" const char* msd_last_error(void) { return g_err.c_str(); }
extern "C" int msd_version(void) { return 1; }

extern "C" int msd_init(int device)
{
    if (G.ready) {
        if (device != G.device) return fail(MSD_EINVAL, "msd_init: already initialised on device %d", G.device);
        return MSD_OK;
    }
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(MSD_EINVAL, "msd_init: device %d not present (%d devices)", device, count);
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    // dynamic LDS budget per workgroup: the device's LDS less 8 KiB kept for the
    // kernels' static tables (exchange classes); plans are checked against it
    G.max_lds = prop.sharedMemPerBlock > 16384 ? prop.sharedMemPerBlock - 8192 : prop.sharedMemPerBlock;
    if (configure_lds(G.max_lds) != hipSuccess || configure_lds3(G.max_lds) != hipSuccess)
        G.max_lds = std::min<size_t>(G.max_lds, 65536 - 8192);
    (void)configure_fft64(65536);
    if (device >= NOALLOC_MAX_DEVICES) return fail(MSD_EINVAL, "msd_init: device %d beyond %d", device, NOALLOC_MAX_DEVICES);
    void* sp = nullptr;   // this device's copy of the allocation-free entry point's scratch
    HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = static_cast<double*>(sp);
    HIPCHK(hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking));
    G.device = device;
    G.ready = true;
    return MSD_OK;
}

extern "C" int msd_synchronize(void)
{
    if (!G.ready) return MSD_OK;
    HIPCHK(hipStreamSynchronize(G.stream));
    return MSD_OK;
}

extern "C" int msd_philox_device(const uint32_t* ctr, int count, const uint32_t key[2], uint32_t* out)
{
    NONNULL(ctr); NONNULL(key); NONNULL(out);
    if (count <= 0) return fail(MSD_EINVAL, "msd_philox_device: count %d", count);
    if (int rc = ensure_init()) return rc;
    uint32_t *dc = nullptr, *dout = nullptr;
    if (int rc = dalloc(&dc, (size_t)4 * count)) return rc;
    if (int rc = dalloc(&dout, (size_t)4 * count)) { dfree(dc); return rc; }
    hipError_t e = hipMemcpyAsync(dc, ctr, sizeof(uint32_t) * 4 * count, hipMemcpyHostToDevice, G.stream);
    if (e == hipSuccess) e = launch_philox(dc, count, key[0], key[1], dout, G.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(uint32_t) * 4 * count, hipMemcpyDeviceToHost, G.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
    dfree(dc); dfree(dout);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "msd_philox_device: %s", hipGetErrorString(e));
    return MSD_OK;
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

static bool same_dims(const Dims& a, int nx, int ny, int nz) { return a.nx == nx && a.ny == ny && a.nz == nz; }

extern "C" int msd_system_create(int nx, int ny, int nz, msd_system** out)
{
    NONNULL(out);
    *out = nullptr;
    if (nx < 1 || ny < 1 || nz < 1) return fail(MSD_EINVAL, "SpinSystem.new: bad size %dx%dx%d", nx, ny, nz);
    // the step kernels address per-site arrays (up to 16 bytes a site) with 32-bit byte offsets
    if ((double)nx * ny * nz > (double)(1u << 28))
        return fail(MSD_EUNSUPPORTED, "SpinSystem.new: %dx%dx%d exceeds 2^28 sites", nx, ny, nz);
    if (int rc = ensure_init()) return rc;
    std::unique_ptr<msd_system> s(new msd_system);
    s->d = Dims{nx, ny, nz, nx * ny * nz};
    if (int rc = dalloc(&s->S, (size_t)4 * s->d.n)) return rc;
    if (int rc = dalloc(&s->Slo, (size_t)3 * s->d.n)) return rc;
    if (int rc = dalloc(&s->part, (size_t)8 * 1024)) return rc;
    if (int rc = dalloc(&s->out, 8)) return rc;
    hipError_t e = launch_fill(s->S, 0.0f, (size_t)4 * s->d.n, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "SpinSystem.new: %s", hipGetErrorString(e));
    int k;
    if (int rc = s->slot("Total", &k)) return rc;
    HIPCHK(hipStreamSynchronize(G.stream));
    *out = s.release();
    return MSD_OK;
}

extern "C" int msd_system_destroy(msd_system* s)
{
    if (!s) return MSD_OK;
    (void)hipStreamSynchronize(G.stream);
    delete s;
    return MSD_OK;
}

extern "C" int msd_system_dims(const msd_system* s, int* nx, int* ny, int* nz)
{
    NONNULL(s);
    if (nx) *nx = s->d.nx;
    if (ny) *ny = s->d.ny;
    if (nz) *nz = s->d.nz;
    return MSD_OK;
}

static int copy_dev(float* dst, const float* src, size_t n)
{
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, G.stream));
    return MSD_OK;
}

extern "C" int msd_system_copy(const msd_system* s, msd_system** out)
{
    NONNULL(s); NONNULL(out);
    msd_system* c = nullptr;
    if (int rc = msd_system_create(s->d.nx, s->d.ny, s->d.nz, &c)) return rc;
    const size_t n = s->d.n;
    int rc = copy_dev(c->S, s->S, 4 * n);
    if (!rc && s->lo_valid) rc = copy_dev(c->Slo, s->Slo, 3 * n);
    c->lo_valid = s->lo_valid;
    c->alpha = s->alpha; c->gamma = s->gamma; c->dt = s->dt; c->time = s->time;
    if (!rc && s->site_alpha) { rc = dalloc(&c->site_alpha, n); if (!rc) rc = copy_dev(c->site_alpha, s->site_alpha, n); }
    if (!rc && s->site_gamma) { rc = dalloc(&c->site_gamma, n); if (!rc) rc = copy_dev(c->site_gamma, s->site_gamma, n); }
    for (size_t k = 0; !rc && k < s->slots.size(); ++k) {
        int idx;
        rc = c->slot(s->slots[k].name.c_str(), &idx);
        if (!rc) { c->slots[idx].used = s->slots[k].used; rc = copy_dev(c->slots[idx].h, s->slots[k].h, 3 * n); }
    }
    if (rc) { delete c; return rc; }
    *out = c;
    return MSD_OK;
}

extern "C" int msd_system_copy_spins_to(const msd_system* src, msd_system* dst)
{
    NONNULL(src); NONNULL(dst);
    if (!same_dims(src->d, dst->d.nx, dst->d.ny, dst->d.nz)) return fail(MSD_EMISMATCH, "copySpinsTo: size mismatch");
    if (src == dst) return MSD_OK;
    if (int rc = copy_dev(dst->S, src->S, (size_t)4 * src->d.n)) return rc;
    if (src->lo_valid)
        if (int rc = copy_dev(dst->Slo, src->Slo, (size_t)3 * src->d.n)) return rc;
    dst->lo_valid = src->lo_valid;
    dst->version++;
    return MSD_OK;
}

extern "C" int msd_system_set_spins(msd_system* s, const double* x, const double* y, const double* z, const double* ms)
{
    NONNULL(s); NONNULL(x); NONNULL(y); NONNULL(z);
    const size_t n = s->d.n;
    std::vector<double> h(4 * n);
    for (size_t i = 0; i < n; ++i) {
        h[i] = x[i]; h[n + i] = y[i]; h[2 * n + i] = z[i];
        h[3 * n + i] = ms ? ms[i] : std::sqrt(x[i] * x[i] + y[i] * y[i] + z[i] * z[i]);
    }
    if (int rc = upload_f32(s->S, h.data(), 4 * n)) return rc;
    // residuals below FP32 for the compensated state of the step driver
    std::vector<double> lo(3 * n);
    for (size_t i = 0; i < 3 * n; ++i) lo[i] = h[i] - (double)(float)h[i];
    if (int rc = upload_f32(s->Slo, lo.data(), 3 * n)) return rc;
    s->lo_valid = true;
    s->version++;
    return MSD_OK;
}

extern "C" int msd_system_get_spins(msd_system* s, double* x, double* y, double* z, double* ms)
{
    NONNULL(s);
    const size_t n = s->d.n;
    std::vector<double> h(4 * n);
    if (int rc = download_f32(h.data(), s->S, 4 * n)) return rc;
    if (s->lo_valid) {
        std::vector<double> lo(3 * n);
        if (int rc = download_f32(lo.data(), s->Slo, 3 * n)) return rc;
        for (size_t i = 0; i < 3 * n; ++i) h[i] += lo[i];
    }
    double* o[4] = {x, y, z, ms};
    for (int c = 0; c < 4; ++c)
        if (o[c]) std::memcpy(o[c], h.data() + c * n, n * sizeof(double));
    return MSD_OK;
}

// l_setspin (cpu/core/spinsystem.cpp:1312-1361) then SpinSystem::set (:1051-1059)
extern "C" int msd_system_set_spin(msd_system* s, int site, double sx, double sy, double sz, double m)
{
    NONNULL(s);
    if (site < 0 || site >= s->d.n) return fail(MSD_EINVAL, "setSpin: site %d outside 0..%d", site, s->d.n - 1);
    if (m >= 0) {
        const double rr = sx * sx + sy * sy + sz * sz;
        if (rr > 0) { const double f = m / std::sqrt(rr); sx *= f; sy *= f; sz *= f; }
        else { sx = 0; sy = 0; sz = m; }
    }
    const float v[4] = {(float)sx, (float)sy, (float)sz, (float)std::sqrt(sx * sx + sy * sy + sz * sz)};
    const size_t n = s->d.n;
    const float lo[3] = {(float)(sx - (double)v[0]), (float)(sy - (double)v[1]), (float)(sz - (double)v[2])};
    for (int c = 0; c < 4; ++c)
        HIPCHK(hipMemcpyAsync(s->S + c * n + site, &v[c], sizeof(float), hipMemcpyHostToDevice, G.stream));
    if (s->lo_valid)
        for (int c = 0; c < 3; ++c)
            HIPCHK(hipMemcpyAsync(s->Slo + c * n + site, &lo[c], sizeof(float), hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    s->version++;
    return MSD_OK;
}

extern "C" int msd_system_get_spin(msd_system* s, int site, double out[4])
{
    NONNULL(s); NONNULL(out);
    if (site < 0 || site >= s->d.n) return fail(MSD_EINVAL, "spin: site %d outside 0..%d", site, s->d.n - 1);
    float v[4];
    const size_t n = s->d.n;
    for (int c = 0; c < 4; ++c)
        HIPCHK(hipMemcpyAsync(&v[c], s->S + c * n + site, sizeof(float), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    for (int c = 0; c < 4; ++c) out[c] = v[c];
    return MSD_OK;
}

extern "C" int msd_system_random_spins(msd_system* s, uint64_t seed)
{
    NONNULL(s);
    hipError_t e = launch_random_spins(s->d, seed, s->S, s->Slo, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "random_spins: %s", hipGetErrorString(e));
    s->lo_valid = true;
    s->version++;
    return MSD_OK;
}

extern "C" int msd_system_set_param(msd_system* s, int which, double v)
{
    NONNULL(s);
    switch (which) {
        case MSD_PARAM_ALPHA: s->alpha = v; dfree(s->site_alpha); break;   // setAlpha clears site values
        case MSD_PARAM_GAMMA: s->gamma = v; dfree(s->site_gamma); break;
        case MSD_PARAM_DT: s->dt = v; break;
        case MSD_PARAM_TIME: s->time = v; break;
        default: return fail(MSD_EINVAL, "set_param: unknown parameter %d", which);
    }
    return MSD_OK;
}

extern "C" int msd_system_get_param(const msd_system* s, int which, double* v)
{
    NONNULL(s); NONNULL(v);
    switch (which) {
        case MSD_PARAM_ALPHA: *v = s->alpha; break;
        case MSD_PARAM_GAMMA: *v = s->gamma; break;
        case MSD_PARAM_DT: *v = s->dt; break;
        case MSD_PARAM_TIME: *v = s->time; break;
        default: return fail(MSD_EINVAL, "get_param: unknown parameter %d", which);
    }
    return MSD_OK;
}

static int set_site_array(msd_system* s, float** dst, const double* a)
{
    if (!a) { (void)hipStreamSynchronize(G.stream); dfree(*dst); return MSD_OK; }
    if (!*dst)
        if (int rc = dalloc(dst, s->d.n)) return rc;
    return upload_f32(*dst, a, s->d.n);
}

extern "C" int msd_system_set_site_alpha(msd_system* s, const double* a) { NONNULL(s); return set_site_array(s, &s->site_alpha, a); }
extern "C" int msd_system_set_site_gamma(msd_system* s, const double* g) { NONNULL(s); return set_site_array(s, &s->site_gamma, g); }

// ------------------------------------------------------------------ checkpoint byte stream
// MagLua's object encoding (common/luabaseobject/luabaseobject.cpp:140-170): int32 and
// IEEE double in host (little-endian) order, chars as one byte.  An object opens with a
// magic char (111: new to the stream, 222: written earlier) and its index in the stream's
// object list (luabaseobject.h:63-72).
enum : char { WIRE_NEW = 111, WIRE_OLD = (char)222 };
enum : int { WIRE_LUA_TUSERDATA = 7 };

// MurmurHash3_x86_32 (Austin Appleby's published, public-domain hash), seed 1000: MagLua's
// type ids (common/luabaseobject/factory.cpp:123-131).  MagLua's copy rotates the tail
// word by 16 where the published hash rotates by 15 (MurmurHash3.cpp:143), so names whose
// length is not a multiple of 4 ("SpinSystem") hash differently; the ids follow MagLua.
static uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

static uint32_t murmur3_x86_32(const uint8_t* p, size_t len, uint32_t seed)
{
    const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
    auto mix = [&](uint32_t k, int r) { k *= c1; k = rotl32(k, r); return k * c2; };
    uint32_t h = seed;
    size_t i = 0;
    for (; i + 4 <= len; i += 4) {
        uint32_t k;
        std::memcpy(&k, p + i, 4);
        h ^= mix(k, 15);
        h = rotl32(h, 13) * 5 + 0xe6546b64u;
    }
    uint32_t k = 0;
    for (size_t j = len - i; j > 0; --j) k = (k << 8) | p[i + j - 1];
    if (len - i) h ^= mix(k, 16);
    h ^= (uint32_t)len;
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

static int32_t type_id(const char* name) { return (int32_t)murmur3_x86_32((const uint8_t*)name, strlen(name), 1000u); }

extern "C" int msd_hash32(const char* name, int32_t* out)
{
    NONNULL(name); NONNULL(out);
    *out = type_id(name);
    return MSD_OK;
}

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

static int wire_finish(const WireOut& w, const char* who, void* buf, int64_t cap, int64_t* len)
{
    *len = (int64_t)w.b.size();
    if (!buf) return MSD_OK;
    if (cap < *len) return fail(MSD_EINVAL, "%s: buffer of %lld bytes, %lld needed", who, (long long)cap, (long long)*len);
    std::memcpy(buf, w.b.data(), w.b.size());
    return MSD_OK;
}

// ENCODE_PREAMBLE of a new object, then SpinOperation::encode (cpu/core/spinoperation.cpp:67-75)
static void wire_operator(WireOut& w, int index, const Dims& d, double scale)
{
    w.ch(WIRE_NEW);
    w.i32(index);
    w.ch(0);
    w.i32(d.nx); w.i32(d.ny); w.i32(d.nz);
    w.f64(scale);
}

// the preamble as decodeLuaBaseObject reads it (luabaseobject.cpp:303-350), then
// SpinOperation::decode (spinoperation.cpp:77-93)
static int unwire_operator(WireIn& r, const char* who, int index, int* nx, int* ny, int* nz, double* scale)
{
    char magic, version;
    int32_t pos;
    if (!r.ch(&magic) || !r.i32(&pos)) return fail(MSD_EINVAL, "%s: stream too short", who);
    if (magic == WIRE_OLD) return fail(MSD_EINVAL, "%s: object %d was decoded earlier in the stream", who, pos);
    if (magic != WIRE_NEW) return fail(MSD_EINVAL, "%s: malformed stream (magic %d)", who, (int)magic);
    if (pos != index) return fail(MSD_EINVAL, "%s: object index %d, expected %d", who, pos, index);
    if (!r.ch(&version) || !r.i32(nx) || !r.i32(ny) || !r.i32(nz) || !r.f64(scale))
        return fail(MSD_EINVAL, "%s: stream ends inside the operator header", who);
    if (version != 0) return fail(MSD_EINVAL, "%s: unknown operator version %d", who, (int)version);
    return MSD_OK;
}

// encodeT<dArray> (luabaseobject.h:101-106) then Array<double>::encode
// (cpu/array/array_core_cpu.h:200-224): one value when all sites hold the same value,
// else the raw doubles.
static void wire_darray(WireOut& w, int index, const Dims& d, const double* v)
{
    w.i32(WIRE_LUA_TUSERDATA);
    w.i32(type_id("Array.Double"));
    w.ch(WIRE_NEW);
    w.i32(index);
    w.ch(0);
    w.i32(d.nx); w.i32(d.ny); w.i32(d.nz);
    bool same = true;
    for (int i = 1; i < d.n && same; ++i) same = v[i] == v[0];
    w.i32(same ? 1 : 0);
    if (same) w.f64(v[0]);
    else w.put(v, (size_t)d.n * sizeof(double));
}

// SpinSystem::encode (cpu/core/spinsystem.cpp:669-760).  Spins are the FP64 values of the
// compensated device state (what msd_system_get_spins returns); site alpha/gamma and ms are
// their FP32 device values widened.  No site carries extra Lua data, so the trailing
// count is 0 (the "partial data" form with no entries).
extern "C" int msd_system_encode(msd_system* s, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects)
{
    NONNULL(s); NONNULL(len);
    if (first_index < 0) return fail(MSD_EINVAL, "SpinSystem.encode: object index %d < 0", first_index);
    const size_t n = s->d.n;
    std::vector<double> sp(4 * n);
    if (int rc = msd_system_get_spins(s, sp.data(), sp.data() + n, sp.data() + 2 * n, sp.data() + 3 * n)) return rc;
    WireOut w;
    w.b.reserve(242 + 48 * n);
    int idx = first_index;
    w.ch(WIRE_NEW);
    w.i32(idx++);
    w.ch(0);
    w.i32(s->d.nx); w.i32(s->d.ny); w.i32(s->d.nz);
    w.f64(s->alpha); w.f64(s->dt); w.f64(s->gamma); w.f64(s->time);
    for (int c = 0; c < 4; ++c) wire_darray(w, idx++, s->d, sp.data() + c * n);
    std::vector<double> site(n);
    for (float* a : {s->site_alpha, s->site_gamma}) {
        w.i32(a ? 1 : 0);
        if (!a) continue;
        if (int rc = download_f32(site.data(), a, n)) return rc;
        wire_darray(w, idx++, s->d, site.data());
    }
    w.i32(0);
    *len = (int64_t)w.b.size();
    if (n_objects) *n_objects = idx - first_index;
    if (!buf) return MSD_OK;
    if (cap < *len) return fail(MSD_EINVAL, "SpinSystem.encode: buffer of %lld bytes, %lld needed", (long long)cap, (long long)*len);
    std::memcpy(buf, w.b.data(), w.b.size());
    return MSD_OK;
}

// decodeT<dArray> (luabaseobject.cpp:303-350) then Array<double>::decode (array_core_cpu.h:226-252)
// Reads one Array.Double of the system's size into v, which is sized only once the
// array's header has been checked and, for an explicit array, the stream holds its
// 8n bytes: a truncated or corrupt stream fails before any large allocation.
static int unwire_darray(WireIn& r, int index, const Dims& d, std::vector<double>& v)
{
    int32_t tag, type, pos, nx, ny, nz, flag;
    char magic, version;
    if (!r.i32(&tag) || !r.i32(&type) || !r.ch(&magic) || !r.i32(&pos))
        return fail(MSD_EINVAL, "SpinSystem.decode: stream ends inside an array header");
    if (tag != WIRE_LUA_TUSERDATA || type != type_id("Array.Double"))
        return fail(MSD_EINVAL, "SpinSystem.decode: expected an Array.Double (tag %d, type %d)", tag, type);
    if (magic == WIRE_OLD)
        return fail(MSD_EUNSUPPORTED, "SpinSystem.decode: array %d refers to an earlier object of the stream", pos);
    if (magic != WIRE_NEW) return fail(MSD_EINVAL, "SpinSystem.decode: malformed stream (magic %d)", (int)magic);
    if (pos != index) return fail(MSD_EINVAL, "SpinSystem.decode: object index %d, expected %d", pos, index);
    if (!r.ch(&version) || !r.i32(&nx) || !r.i32(&ny) || !r.i32(&nz) || !r.i32(&flag))
        return fail(MSD_EINVAL, "SpinSystem.decode: stream ends inside an array");
    if (version != 0) return fail(MSD_EINVAL, "SpinSystem.decode: array version %d", (int)version);
    if (!same_dims(d, nx, ny, nz))
        return fail(MSD_EINVAL, "SpinSystem.decode: array of %dx%dx%d in a %dx%dx%d system", nx, ny, nz, d.nx, d.ny, d.nz);
    if (flag == 1) {
        double c;
        if (!r.f64(&c)) return fail(MSD_EINVAL, "SpinSystem.decode: stream ends inside an array");
        v.assign(d.n, c);
        return MSD_OK;
    }
    if (flag != 0) return fail(MSD_EINVAL, "SpinSystem.decode: array flag %d", flag);
    const size_t bytes = (size_t)d.n * sizeof(double);
    if (bytes > r.n - r.at) return fail(MSD_EINVAL, "SpinSystem.decode: stream ends inside an array");
    v.resize(d.n);
    if (!r.get(v.data(), bytes)) return fail(MSD_EINVAL, "SpinSystem.decode: stream ends inside an array");
    return MSD_OK;
}

// SpinSystem::decode (cpu/core/spinsystem.cpp:764-846), entered at the object's magic
// char so that it reads what msd_system_encode writes.
extern "C" int msd_system_decode(const void* buf, int64_t len, int first_index, msd_system** out, int64_t* used, int* n_objects)
{
    NONNULL(buf); NONNULL(out);
    *out = nullptr;
    if (len < 0) return fail(MSD_EINVAL, "SpinSystem.decode: length %lld", (long long)len);
    WireIn r{(const uint8_t*)buf, (size_t)len};
    char magic, version;
    int32_t pos, nx, ny, nz;
    double alpha, dt, gamma, time;
    if (!r.ch(&magic) || !r.i32(&pos)) return fail(MSD_EINVAL, "SpinSystem.decode: stream too short");
    if (magic == WIRE_OLD)
        return fail(MSD_EINVAL, "SpinSystem.decode: object %d was decoded earlier in the stream", pos);
    if (magic != WIRE_NEW) return fail(MSD_EINVAL, "SpinSystem.decode: malformed stream (magic %d)", (int)magic);
    if (pos != first_index) return fail(MSD_EINVAL, "SpinSystem.decode: object index %d, expected %d", pos, first_index);
    if (!r.ch(&version)) return fail(MSD_EINVAL, "SpinSystem.decode: stream too short");
    if (version != 0) return fail(MSD_EINVAL, "SpinSystem.decode: unknown version %d", (int)version);
    if (!r.i32(&nx) || !r.i32(&ny) || !r.i32(&nz) || !r.f64(&alpha) || !r.f64(&dt) || !r.f64(&gamma) || !r.f64(&time))
        return fail(MSD_EINVAL, "SpinSystem.decode: stream ends inside the header");
    if (nx < 1 || ny < 1 || nz < 1 || (double)nx * ny * nz > (double)(1u << 28))
        return fail(MSD_EINVAL, "SpinSystem.decode: lattice %dx%dx%d", nx, ny, nz);
    const Dims d{nx, ny, nz, nx * ny * nz};
    const size_t n = d.n;
    int idx = first_index + 1;
    std::vector<double> sp[4];   // x y z ms
    for (int c = 0; c < 4; ++c)
        if (int rc = unwire_darray(r, idx++, d, sp[c])) return rc;
    std::vector<double> site[2];
    for (int k = 0; k < 2; ++k) {
        int32_t has;
        if (!r.i32(&has)) return fail(MSD_EINVAL, "SpinSystem.decode: stream ends before the site arrays");
        if (!has) continue;
        if (int rc = unwire_darray(r, idx++, d, site[k])) return rc;
    }
    int32_t extra;
    if (!r.i32(&extra)) return fail(MSD_EINVAL, "SpinSystem.decode: stream ends before the extra-data count");
    if (extra > 0) return fail(MSD_EUNSUPPORTED, "SpinSystem.decode: %d sites carry extra Lua data", extra);
    if (extra < 0)   // every site written explicitly: accepted when each is nil
        for (size_t i = 0; i < n; ++i) {
            int32_t t;
            if (!r.i32(&t)) return fail(MSD_EINVAL, "SpinSystem.decode: stream ends inside the extra data");
            if (t != 0) return fail(MSD_EUNSUPPORTED, "SpinSystem.decode: site %zu carries extra Lua data", i);
        }
    msd_system* s = nullptr;
    if (int rc = msd_system_create(nx, ny, nz, &s)) return rc;
    std::unique_ptr<msd_system> own(s);
    s->alpha = alpha; s->dt = dt; s->gamma = gamma; s->time = time;
    if (int rc = msd_system_set_spins(s, sp[0].data(), sp[1].data(), sp[2].data(), sp[3].data())) return rc;
    if (!site[0].empty())
        if (int rc = set_site_array(s, &s->site_alpha, site[0].data())) return rc;
    if (!site[1].empty())
        if (int rc = set_site_array(s, &s->site_gamma, site[1].data())) return rc;
    HIPCHK(hipStreamSynchronize(G.stream));
    if (used) *used = (int64_t)r.at;
    if (n_objects) *n_objects = idx - first_index;
    *out = own.release();
    return MSD_OK;
}

extern "C" int msd_system_reset_fields(msd_system* s)
{
    NONNULL(s);
    for (Slot& sl : s->slots) {
        sl.used = false;
        hipError_t e = launch_fill(sl.h, 0.0f, (size_t)3 * s->d.n, G.stream);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "resetFields: %s", hipGetErrorString(e));
    }
    return MSD_OK;
}

extern "C" int msd_system_sum_fields(msd_system* s)
{
    NONNULL(s);
    int T;
    if (int rc = s->slot("Total", &T)) return rc;
    std::vector<const float*> src;
    for (size_t k = 0; k < s->slots.size(); ++k)
        if ((int)k != T && s->slots[k].used) src.push_back(s->slots[k].h);
    if (src.size() > 16) return fail(MSD_EUNSUPPORTED, "sumFields: more than 16 used slots");
    hipError_t e = launch_sum_fields(s->slots[T].h, src.data(), (int)src.size(), (size_t)3 * s->d.n, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "sumFields: %s", hipGetErrorString(e));
    return MSD_OK;
}

extern "C" int msd_system_get_field(msd_system* s, const char* slot, double* hx, double* hy, double* hz)
{
    NONNULL(s); NONNULL(slot);
    const int k = s->find(slot);
    if (k < 0) return fail(MSD_EINVAL, "fieldArray: unknown slot '%s'", slot);
    const size_t n = s->d.n;
    std::vector<double> h(3 * n);
    if (int rc = download_f32(h.data(), s->slots[k].h, 3 * n)) return rc;
    if (hx) std::memcpy(hx, h.data(), n * sizeof(double));
    if (hy) std::memcpy(hy, h.data() + n, n * sizeof(double));
    if (hz) std::memcpy(hz, h.data() + 2 * n, n * sizeof(double));
    return MSD_OK;
}

extern "C" int msd_system_slot_used(const msd_system* s, const char* slot, int* used)
{
    NONNULL(s); NONNULL(slot); NONNULL(used);
    const int k = s->find(slot);
    *used = (k >= 0 && s->slots[k].used) ? 1 : 0;
    return MSD_OK;
}

static int fetch_out(msd_system* s, double* dst, int count)
{
    HIPCHK(hipMemcpyAsync(dst, s->out, sizeof(double) * count, hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    return MSD_OK;
}

extern "C" int msd_system_net_moment(msd_system* s, double out[8])
{
    NONNULL(s); NONNULL(out);
    hipError_t e = launch_reduce_moment(s->d, s->S, s->lo_valid ? s->Slo : nullptr, s->part, s->out, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "netMoment: %s", hipGetErrorString(e));
    double v[6];
    if (int rc = fetch_out(s, v, 6)) return rc;
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    out[3] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    out[4] = v[3]; out[5] = v[4]; out[6] = v[5];
    out[7] = std::sqrt(v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
    return MSD_OK;
}

extern "C" int msd_net_moment_noalloc(const float* spins, const float* lo, int n, double out[8])
{
    NONNULL(spins); NONNULL(out);
    // msd_init creates the stream and looks up the scratch: neither may happen here
    if (!G.ready) return fail(MSD_EINVAL, "msd_net_moment_noalloc: msd_init has not been called");
    if (n <= 0 || n > MSD_NOALLOC_MAX_SITES)
        return fail(MSD_EINVAL, "msd_net_moment_noalloc: n = %d outside 1..%d", n, MSD_NOALLOC_MAX_SITES);
    double* part = g_scratch_dev[G.device];
    if (!part) return fail(MSD_EDEVICE, "msd_net_moment_noalloc: no scratch address for device %d", G.device);
    const Dims d{n, 1, 1, n};   // the reductions see only the site count
    const int nb = reduce_blocks(d);
    if (nb > NOALLOC_PART_BLOCKS) return fail(MSD_EINVAL, "msd_net_moment_noalloc: %d partial blocks exceed %d", nb, NOALLOC_PART_BLOCKS);
    double* res = part + 6 * NOALLOC_PART_BLOCKS;
    // every partial is written by k_moment before k_fold reads it: nothing to clear
    hipError_t e = launch_reduce_moment(d, spins, lo, part, res, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "msd_net_moment_noalloc: %s", hipGetErrorString(e));
    double v[6];
    HIPCHK(hipMemcpyAsync(v, res, sizeof v, hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    out[3] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    out[4] = v[3]; out[5] = v[4]; out[6] = v[5];
    out[7] = std::sqrt(v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
    return MSD_OK;
}

extern "C" int msd_system_field_dot(msd_system* s, const char* slot, double* out)
{
    NONNULL(s); NONNULL(slot); NONNULL(out);
    const int k = s->find(slot);
    if (k < 0) return fail(MSD_EINVAL, "field_dot: unknown slot '%s'", slot);
    hipError_t e = launch_reduce_dot(s->d, s->slots[k].h, s->S, s->lo_valid ? s->Slo : nullptr, s->part, s->out, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "field_dot: %s", hipGetErrorString(e));
    return fetch_out(s, out, 1);
}

extern "C" int msd_system_diff(msd_system* a, msd_system* b, double out[4])
{
    NONNULL(a); NONNULL(b); NONNULL(out);
    if (!same_dims(a->d, b->d.nx, b->d.ny, b->d.nz)) return fail(MSD_EMISMATCH, "diff: size mismatch");
    hipError_t e = launch_reduce_diff(a->d, a->S, a->lo_valid ? a->Slo : nullptr, b->S, b->lo_valid ? b->Slo : nullptr, a->part,
                                      a->out, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "diff: %s", hipGetErrorString(e));
    if (int rc = fetch_out(a, out, 3)) return rc;
    out[3] = std::sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
    return MSD_OK;
}

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
        if (lut_ok && !paths.empty()) {
            e.mode = 1; e.cls = cls; e.cls_start = cls_start; e.ent = ent;
            e.ncls = ncls; e.nent = nent;
        }
        return e;
    }
};

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

struct msd_llg {
    bool thermal_first = true;
    bool no_renorm = false;
    bool quaternion = false;    // LLG.Quaternion
    double ps = 1, ds = 1;
};

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

static int twiddles(int L, float2** d32, double2** d64)
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
static int layered_shape(msd_dipole* p, int ty, int tpb, bool row_ct, bool want_split, const char* what)
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
static int dipole_field_of(msd_dipole* dip, const float* s3, float* h)
{
    if (int rc = dip->prepare()) return rc;
    dip->epoch++;
    hipError_t e = launch_rows_fwd(dip->dev, s3, G.stream);
    if (e == hipSuccess) e = launch_cols(dip->dev, G.stream);
    if (e == hipSuccess) e = launch_rows_inv(dip->dev, h, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "dipole field: %s", hipGetErrorString(e));
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
    float* Tw = nullptr;     // [6][fy][fxh][fzs], fzs <= fz
    float2* W = nullptr;     // working spectrum [3][nz][fy][fxp]
    float2 *twx = nullptr, *twy = nullptr, *twz = nullptr;
    double2 *twx64 = nullptr, *twy64 = nullptr, *twz64 = nullptr;
    FFTPlan px{}, py{}, pz{};
    int ty = 1, wy = 1, wz = 1, zt = 0, yt = 0, rt = 0, fzs = 0, fxp = 0;
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
        c.ty = ty; c.wy = wy; c.wz = wz; c.zt = zt; c.yt = yt; c.rt = rt; c.fzs = fzs;
        c.px = px; c.py = py; c.pz = pz;
        c.twx = twx; c.twy = twy; c.twz = twz;
        c.W = W; c.Tw = Tw;
        return c;
    }
};

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

// Whether every component has its parity in z alone, T(x, y, -z) = +T (xx xy yy zz)
// or -T (xz yz): then the spectrum has the same parity in kz and only kz <= fz/2
// is stored.  The built-in tensors have it (reflection symmetry of the lattice sums).
static int lr3_z_parity(msd_longrange3d* l, bool* has)
{
    const size_t plane = (size_t)l->fx * l->fy * l->fz;
    std::vector<double> t(6 * plane);
    HIPCHK(hipMemcpyAsync(t.data(), l->T64, sizeof(double) * t.size(), hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    double big = 0;
    for (double v : t) big = std::max(big, std::fabs(v));
    *has = true;
    for (int a = 0; a < 6 && *has; ++a) {
        const double sg = (a == 2 || a == 4) ? -1.0 : 1.0;
        for (int z = 0; z < l->fz && *has; ++z)
            for (size_t xy = 0; xy < (size_t)l->fx * l->fy; ++xy) {
                const double u = t[a * plane + (size_t)z * l->fx * l->fy + xy];
                const double v = t[a * plane + (size_t)((l->fz - z) % l->fz) * l->fx * l->fy + xy];
                if (std::fabs(u - sg * v) > 1e-12 * big) { *has = false; break; }
            }
    }
    return MSD_OK;
}

int msd_longrange3d::prepare()
{
    if (int rc = lr3_load(this)) return rc;
    if (!spectrum) {
        bool zpar = true;
        if (custom) {
            if (int rc = lr3_check_even(this)) return rc;
            if (int rc = lr3_z_parity(this, &zpar)) return rc;
        }
        fzs = zpar ? fz / 2 + 1 : fz;
        words_mult = NAN;
        spectrum = true;
    }
    const double mult = g * scale / ((double)fx * fy * fz);
    if (!(words_mult == mult)) {
        const size_t wn = (size_t)6 * fz * fy * fxh;
        double2 *U1 = nullptr, *U2 = nullptr;
        if (int rc = dalloc(&U1, wn)) return rc;
        if (int rc = dalloc(&U2, wn)) { dfree(U1); return rc; }
        hipError_t e = launch3_spectrum(fx, fy, fz, fzs, T64, U1, U2, Tw, mult, twx64, twy64, twz64, G.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
        dfree(U1); dfree(U2);
        if (e != hipSuccess) return fail(MSD_EDEVICE, "LongRange3D spectrum: %s", hipGetErrorString(e));
        words_mult = mult;
    }
    return MSD_OK;
}

// work items of one replica appended to the tables
static void items3(int rep, const Dims& d, const Conv3Dev* cv, int ty, std::vector<int4>& rows, std::vector<int4>& ys,
                   std::vector<int4>& zs)
{   // zs: this replica's z-pass items (its own z-pass form, cv->zt)
    for (int z = 0; z < d.nz; ++z)
        for (int y0 = 0; y0 < d.ny; y0 += ty) rows.push_back(make_int4(rep, z, y0, 0));
    if (!cv) return;
    for (int l = 0; l < 3 * d.nz; ++l)
        for (int k = 0; k < cv->fxh; k += cv->wy) ys.push_back(make_int4(rep, l, k, 0));
    for (int ky = 0; ky < cv->fy; ++ky)
        for (int k = 0; k < cv->fxh; k += cv->wz) zs.push_back(make_int4(rep, ky, k, 0));
}

// Workgroups go to the 8 XCDs round-robin by index.  Reorder a y or z pass's items
// so that each XCD takes one contiguous stretch of the natural order (balanced to
// within one item): neighbouring kx blocks, which share the cache lines at their
// edges, then meet in one L2.  (Row tiles keep the natural order: measured faster.)
static std::vector<int4> xcd_order(const std::vector<int4>& v)
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

template <class T>
static int upload_vec(T** dst, const std::vector<T>& v)
{
    dfree(*dst);
    if (v.empty()) return MSD_OK;
    if (int rc = dalloc(dst, v.size())) return rc;
    HIPCHK(hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, G.stream));
    return MSD_OK;
}

// Exact LongRange3D field of the spins S ([3][n]) into H ([3][n], overwritten): the
// five passes with the single-replica descriptor.  The working spectrum W is left
// holding the field's spectrum.
static int lr3_field_of(msd_longrange3d* l, const float* S, float* H)
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
    if (e == hipSuccess) e = launch3_ypass(-1, l->orep, l->oitems_y, l->ony, l->lds_y, l->yt, G.stream);
    if (e == hipSuccess) e = launch3_zconv(l->orep, l->oitems_z, l->onz, l->lds_z, l->zt, G.stream);
    if (e == hipSuccess) e = launch3_ypass(1, l->orep, l->oitems_y, l->ony, l->lds_y, l->yt, G.stream);
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

// ------------------------------------------------------------ batch of replicas
// Device descriptors and work tables of replicas stepped by the same launches.
struct Batch3 {
    // the row, y and z passes run once per pass form among the members (the
    // compile-time pass for each length that has one, the runtime pass for the
    // rest), so every member is transformed by the same kernels as when it runs alone
    struct ZGroup {
        int zt = 0;
        int4* items = nullptr;
        int n = 0;
        size_t lds = 0;
    };
    Rep3* reps = nullptr;
    int nrep = 0;
    std::vector<ZGroup> rg, yg, zg;      // row, y-pass and z-pass groups (key: Conv3Dev::rt / yt / zt)
    bool any_lr = false;
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
    // rebuild the tables when the member set changed; refresh descriptors always
    int update(const std::vector<Rep3>& r, const std::vector<int>& tys, const std::vector<size_t>& lds3)
    {
        const bool same_shape = host.size() == r.size() && reps != nullptr;
        if (!same_shape) {
            release();
            std::vector<std::vector<int4>> R, Y, Z;
            std::vector<ZGroup> rgroups, ygroups, zgroups;
            any_lr = false;
            // append a member's pass items to the group of its pass form
            auto add = [](std::vector<ZGroup>& groups, std::vector<std::vector<int4>>& lists, int key, size_t lds,
                          const std::vector<int4>& items) {
                if (items.empty()) return;
                size_t gi = 0;
                while (gi < groups.size() && groups[gi].zt != key) ++gi;
                if (gi == groups.size()) {
                    groups.push_back(ZGroup{});
                    groups.back().zt = key;
                    lists.emplace_back();
                }
                groups[gi].lds = std::max(groups[gi].lds, lds);
                lists[gi].insert(lists[gi].end(), items.begin(), items.end());
            };
            for (size_t i = 0; i < r.size(); ++i) {
                std::vector<int4> ri, yi, zi;
                items3((int)i, r[i].a.d, r[i].has_lr ? &r[i].cv : nullptr, tys[i], ri, yi, zi);
                any_lr |= r[i].has_lr != 0;
                add(rgroups, R, r[i].has_lr ? r[i].cv.rt : 0, lds3[3 * i], ri);
                add(ygroups, Y, r[i].cv.yt, lds3[3 * i + 1], yi);
                add(zgroups, Z, r[i].cv.zt, lds3[3 * i + 2], zi);
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
            if (e == hipSuccess) e = launch3_zconv(reps, z.items, z.n, z.lds, z.zt, st);
        return e;
    }
    hipError_t ypass(int sign, hipStream_t st) const
    {
        hipError_t e = hipSuccess;
        for (const ZGroup& y : yg)
            if (e == hipSuccess) e = launch3_ypass(sign, reps, y.items, y.n, y.lds, y.zt, st);
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

static int surrogate_dims_ok(int nx, int ny, int nz)
{
    if (nx < 32 || nx % 32) return fail(MSD_EUNSUPPORTED, "Surrogate.new: nx=%d must be a multiple of 32", nx);
    if (ny < 1 || nz < 1) return fail(MSD_EINVAL, "Surrogate.new: bad size");
    if ((double)nx * ny * nz > (double)(1u << 27)) return fail(MSD_EUNSUPPORTED, "Surrogate.new: more than 2^27 sites");
    return MSD_OK;
}

extern "C" int msd_surrogate_set_open(msd_surrogate* s, int open)
{
    NONNULL(s);
    s->cd.open = open ? 1 : 0;
    s->plan_walks();
    return MSD_OK;
}

extern "C" int msd_surrogate_create(int nx, int ny, int nz, msd_surrogate** out)
{
    NONNULL(out);
    *out = nullptr;
    if (int rc = surrogate_dims_ok(nx, ny, nz)) return rc;
    if (int rc = ensure_init()) return rc;
    std::unique_ptr<msd_surrogate> s(new msd_surrogate);
    s->d = Dims{nx, ny, nz, nx * ny * nz};
    CnnDims& c = s->cd;
    c.nx = nx; c.ny = ny; c.nz = nz; c.n = s->d.n;
    c.tiles_x = nx / 32; c.tiles_y = (ny + 7) / 8; c.ntiles = c.tiles_x * c.tiles_y * nz;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, G.device));
    s->cus = prop.multiProcessorCount;
    s->grid = std::min(c.ntiles, 2 * s->cus);
    s->plan_walks();
    const size_t n = s->d.n, W3 = (size_t)3 * msd_surrogate::WSZ;
    for (float** p : {&s->W, &s->Mo, &s->Vo, &s->gW})
        if (int rc = dalloc(p, W3)) return rc;
    if (int rc = dalloc(&s->Af, (size_t)3 * msd_surrogate::PACK)) return rc;
    if (int rc = dalloc(&s->At, (size_t)3 * msd_surrogate::PACK)) return rc;
    for (uint16_t** p : {&s->X1, &s->X2, &s->G1, &s->G2})
        if (int rc = dalloc(p, 16 * n)) return rc;
    for (uint16_t** p : {&s->X0, &s->G3})
        if (int rc = dalloc(p, 4 * n)) return rc;
    if (int rc = dalloc(&s->zeros, 8)) return rc;
    for (float** p : {&s->Y3, &s->T, &s->D})
        if (int rc = dalloc(p, 3 * n)) return rc;
    if (int rc = dalloc(&s->partial, (size_t)s->grid * msd_surrogate::WSZ)) return rc;
    s->nlpart = (int)((n + 255) / 256);
    if (int rc = dalloc(&s->lpart, (size_t)s->nlpart)) return rc;
    hipError_t e = hipSuccess;
    for (float* p : {s->W, s->Mo, s->Vo, s->gW})
        if (e == hipSuccess) e = launch_fill(p, 0.0f, W3, G.stream);
    if (e == hipSuccess) e = launch_fill(reinterpret_cast<float*>(s->zeros), 0.0f, 4, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.new: %s", hipGetErrorString(e));
    HIPCHK(hipStreamSynchronize(G.stream));
    *out = s.release();
    return MSD_OK;
}

extern "C" int msd_surrogate_destroy(msd_surrogate* s) { delete s; return MSD_OK; }

extern "C" int msd_surrogate_set_weights(msd_surrogate* s, int layer, const double* w)
{
    NONNULL(s); NONNULL(w);
    if (layer < 0 || layer > 2) return fail(MSD_EINVAL, "Surrogate.setWeights: layer %d", layer);
    const int co = msd_surrogate::cout(layer), ci = msd_surrogate::cin(layer);
    std::vector<float> h(msd_surrogate::WSZ, 0.0f);
    for (int t = 0; t < msd_surrogate::NT; ++t)
        for (int o = 0; o < co; ++o)
            for (int i = 0; i < ci; ++i) h[(t * 16 + o) * 16 + i] = (float)w[(t * co + o) * ci + i];
    HIPCHK(hipMemcpyAsync(s->W + (size_t)layer * msd_surrogate::WSZ, h.data(), h.size() * sizeof(float),
                          hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    s->packed = false;
    return MSD_OK;
}

static int surrogate_get(const float* dsrc, int layer, double* w)
{
    const int co = msd_surrogate::cout(layer), ci = msd_surrogate::cin(layer);
    std::vector<float> h(msd_surrogate::WSZ);
    HIPCHK(hipMemcpyAsync(h.data(), dsrc + (size_t)layer * msd_surrogate::WSZ, h.size() * sizeof(float),
                          hipMemcpyDeviceToHost, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    for (int t = 0; t < msd_surrogate::NT; ++t)
        for (int o = 0; o < co; ++o)
            for (int i = 0; i < ci; ++i) w[(t * co + o) * ci + i] = h[(t * 16 + o) * 16 + i];
    return MSD_OK;
}

extern "C" int msd_surrogate_get_weights(const msd_surrogate* s, int layer, double* w)
{
    NONNULL(s); NONNULL(w);
    if (layer < 0 || layer > 2) return fail(MSD_EINVAL, "Surrogate.weights: layer %d", layer);
    return surrogate_get(s->W, layer, w);
}

extern "C" int msd_surrogate_get_grads(const msd_surrogate* s, int layer, double* g)
{
    NONNULL(s); NONNULL(g);
    if (layer < 0 || layer > 2) return fail(MSD_EINVAL, "Surrogate.grads: layer %d", layer);
    return surrogate_get(s->gW, layer, g);
}

extern "C" int msd_surrogate_set_scales(msd_surrogate* s, double s_in, double s_out, double sigma2)
{
    NONNULL(s);
    if (!(s_in > 0) || !(s_out > 0) || !(sigma2 > 0)) return fail(MSD_EINVAL, "Surrogate.setScales: scales must be > 0");
    s->s_in = s_in; s->s_out = s_out; s->sigma2 = sigma2;
    return MSD_OK;
}

extern "C" int msd_surrogate_get_scales(const msd_surrogate* s, double* s_in, double* s_out, double* sigma2)
{
    NONNULL(s); NONNULL(s_in); NONNULL(s_out); NONNULL(sigma2);
    *s_in = s->s_in; *s_out = s->s_out; *sigma2 = s->sigma2;
    return MSD_OK;
}

extern "C" int msd_surrogate_set_input(msd_surrogate* s, const double* dm)
{
    NONNULL(s); NONNULL(dm);
    if (int rc = upload_f32(s->D, dm, (size_t)3 * s->d.n)) return rc;
    hipError_t e = launch_cnn_input(s->d, s->D, nullptr, (float)s->s_in, s->X0, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.setInput: %s", hipGetErrorString(e));
    return MSD_OK;
}

extern "C" int msd_surrogate_set_input_systems(msd_surrogate* s, const msd_system* m, const msd_system* m_ref)
{
    NONNULL(s); NONNULL(m); NONNULL(m_ref);
    if (!same_dims(s->d, m->d.nx, m->d.ny, m->d.nz) || !same_dims(s->d, m_ref->d.nx, m_ref->d.ny, m_ref->d.nz))
        return fail(MSD_EMISMATCH, "Surrogate.setInput: SpinSystem size differs");
    hipError_t e = launch_cnn_input(s->d, m->S, m_ref->S, (float)s->s_in, s->X0, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.setInput: %s", hipGetErrorString(e));
    return MSD_OK;
}

extern "C" int msd_surrogate_set_target(msd_surrogate* s, const double* t)
{
    NONNULL(s); NONNULL(t);
    return upload_f32(s->T, t, (size_t)3 * s->d.n);
}

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

static int surrogate_set_target(msd_surrogate* s, ExactField f, const msd_system* m, const msd_system* m_ref)
{
    if (!same_dims(s->d, m->d.nx, m->d.ny, m->d.nz) || !same_dims(s->d, m_ref->d.nx, m_ref->d.ny, m_ref->d.nz) ||
        !f.fits(s->d))
        return fail(MSD_EMISMATCH, "Surrogate.setTarget: SpinSystem or field operator size differs");
    hipError_t e = launch_cnn_sub((size_t)3 * s->d.n, m->S, m_ref->S, s->D, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.setTarget: %s", hipGetErrorString(e));
    return f.of(s->D, s->T);   // the field is linear in the spins
}

extern "C" int msd_surrogate_set_target_dipole(msd_surrogate* s, msd_dipole* dip, const msd_system* m,
                                               const msd_system* m_ref)
{
    NONNULL(s); NONNULL(dip); NONNULL(m); NONNULL(m_ref);
    return surrogate_set_target(s, ExactField{dip, nullptr}, m, m_ref);
}

extern "C" int msd_surrogate_set_target_longrange3d(msd_surrogate* s, msd_longrange3d* lr, const msd_system* m,
                                                    const msd_system* m_ref)
{
    NONNULL(s); NONNULL(lr); NONNULL(m); NONNULL(m_ref);
    return surrogate_set_target(s, ExactField{nullptr, lr}, m, m_ref);
}

extern "C" int msd_surrogate_forward(msd_surrogate* s, double* out)
{
    NONNULL(s);
    if (int rc = s->forward(s->Y3, nullptr, G.stream)) return rc;
    if (out) return download_f32(out, s->Y3, (size_t)3 * s->d.n);
    return MSD_OK;
}

// forward, loss, back-propagation and the three weight gradients
static int surrogate_loss_grads(msd_surrogate* s, double* loss)
{
    const size_t n = s->d.n;
    if (int rc = s->forward(s->Y3, nullptr, G.stream)) return rc;
    const float gscale = (float)(2.0 / (3.0 * (double)n * s->sigma2));
    const int P = msd_surrogate::PACK, WS = msd_surrogate::WSZ;
    // the forward output already carries s_out: dL/dY3 = s_out dL/d(dH)
    hipError_t e = launch_cnn_loss(s->d, s->Y3, s->T, 1.0f, gscale * (float)s->s_out, s->G3, s->lpart, G.stream);
    if (e == hipSuccess) e = s->rows(CV_IN3, s->G3, s->At + 2 * P, s->G2, nullptr, 1.0f, G.stream);
    if (e == hipSuccess) e = s->rows(CV_F16, s->G2, s->At + P, s->G1, nullptr, 1.0f, G.stream);
    if (e == hipSuccess) e = launch_cnn_wgrad(s->cd, s->grid, s->X2, 16, s->G3, 4, s->partial, s->gW + 2 * WS, G.stream);
    if (e == hipSuccess) e = launch_cnn_wgrad(s->cd, s->grid, s->X1, 16, s->G2, 16, s->partial, s->gW + WS, G.stream);
    if (e == hipSuccess) e = launch_cnn_wgrad(s->cd, s->grid, s->X0, 4, s->G1, 16, s->partial, s->gW, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "surrogate backward: %s", hipGetErrorString(e));
    if (loss) {
        std::vector<double> h(s->nlpart);
        HIPCHK(hipMemcpyAsync(h.data(), s->lpart, sizeof(double) * h.size(), hipMemcpyDeviceToHost, G.stream));
        HIPCHK(hipStreamSynchronize(G.stream));
        double acc = 0.0;
        for (double v : h) acc += v;
        *loss = acc / (3.0 * (double)n * s->sigma2);
    }
    return MSD_OK;
}

extern "C" int msd_surrogate_loss_grads(msd_surrogate* s, double* loss)
{
    NONNULL(s);
    return surrogate_loss_grads(s, loss);
}

extern "C" int msd_surrogate_adam(msd_surrogate* s, double lr)
{
    NONNULL(s);
    if (!(lr > 0)) return fail(MSD_EINVAL, "Surrogate.adam: lr must be > 0");
    const double b1 = 0.9, b2 = 0.999;
    s->adam_t++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, s->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, s->adam_t)));
    hipError_t e = launch_cnn_adam(s->W, s->gW, s->Mo, s->Vo, 3 * msd_surrogate::WSZ, (float)lr, (float)b1, (float)b2,
                                   c1, c2, 1e-8f, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.adam: %s", hipGetErrorString(e));
    s->packed = false;
    return MSD_OK;
}

extern "C" int msd_surrogate_reset_optimizer(msd_surrogate* s)
{
    NONNULL(s);
    hipError_t e = launch_fill(s->Mo, 0.0f, (size_t)3 * msd_surrogate::WSZ, G.stream);
    if (e == hipSuccess) e = launch_fill(s->Vo, 0.0f, (size_t)3 * msd_surrogate::WSZ, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.resetOptimizer: %s", hipGetErrorString(e));
    s->adam_t = 0;
    return MSD_OK;
}

extern "C" int msd_surrogate_train_step(msd_surrogate* s, double lr, double* loss)
{
    NONNULL(s);
    if (int rc = surrogate_loss_grads(s, loss)) return rc;
    return msd_surrogate_adam(s, lr);
}

// max |a - b| and max |b| over [3][n] device fields
static int field_maxdiff(const float* a, const float* b, size_t n3, double* maxdiff, double* maxb)
{
    const int blocks = 512;
    if (!G.part)    // kept for the process: the check runs at every learned window
        if (int rc = dalloc(&G.part, blocks)) return rc;
    hipError_t e = launch_cnn_maxdiff(n3, a, b, G.part, blocks, G.stream);
    std::vector<float2> h(blocks);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), G.part, sizeof(float2) * blocks, hipMemcpyDeviceToHost, G.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "field check: %s", hipGetErrorString(e));
    double md = 0.0, mb = 0.0;
    for (const float2& v : h) { md = std::max(md, (double)v.x); mb = std::max(mb, (double)v.y); }
    *maxdiff = md;
    *maxb = mb;
    return MSD_OK;
}

// Surrogate prediction of the dipole field at m (m_ref, H_ref its reference
// state and exact field) against a field given on the host: returns
// max|pred - exact| / max|exact| (the residual the step driver checks).
static int surrogate_field_error(msd_surrogate* s, ExactField f, const msd_system* m, const msd_system* m_ref,
                                 double* rel_err, double* frozen_err)
{
    if (!same_dims(s->d, m->d.nx, m->d.ny, m->d.nz) || !same_dims(s->d, m_ref->d.nx, m_ref->d.ny, m_ref->d.nz) ||
        !f.fits(s->d))
        return fail(MSD_EMISMATCH, "Surrogate.fieldError: SpinSystem or field operator size differs");
    const size_t n3 = (size_t)3 * s->d.n;
    // exact H(m) into T, exact H(m_ref) into D, prediction D + CNN(m - m_ref) into Y3
    if (int rc = f.of(m->S, s->T)) return rc;
    if (int rc = f.of(m_ref->S, s->D)) return rc;
    hipError_t e = launch_cnn_input(s->d, m->S, m_ref->S, (float)s->s_in, s->X0, G.stream);
    if (e != hipSuccess) return fail(MSD_EDEVICE, "Surrogate.fieldError: %s", hipGetErrorString(e));
    if (int rc = s->forward(s->Y3, s->D, G.stream)) return rc;
    double md, mb;
    if (int rc = field_maxdiff(s->Y3, s->T, n3, &md, &mb)) return rc;
    *rel_err = mb > 0 ? md / mb : md;
    if (frozen_err) {
        if (int rc = field_maxdiff(s->D, s->T, n3, &md, &mb)) return rc;
        *frozen_err = mb > 0 ? md / mb : md;
    }
    return MSD_OK;
}

extern "C" int msd_surrogate_field_error(msd_surrogate* s, msd_dipole* dip, const msd_system* m, const msd_system* m_ref,
                                         double* rel_err, double* frozen_err)
{
    NONNULL(s); NONNULL(dip); NONNULL(m); NONNULL(m_ref); NONNULL(rel_err);
    return surrogate_field_error(s, ExactField{dip, nullptr}, m, m_ref, rel_err, frozen_err);
}

extern "C" int msd_surrogate_field_error_longrange3d(msd_surrogate* s, msd_longrange3d* lr, const msd_system* m,
                                                     const msd_system* m_ref, double* rel_err, double* frozen_err)
{
    NONNULL(s); NONNULL(lr); NONNULL(m); NONNULL(m_ref); NONNULL(rel_err);
    return surrogate_field_error(s, ExactField{nullptr, lr}, m, m_ref, rel_err, frozen_err);
}

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
    std::vector<StepNode> gnodes[3];
    bool gnodes_ok[3] = {false, false, false};          // every step launch of the graph found
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
        if (int rc = b3.update({rep()}, {args.ty}, lds)) return rc;
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
    if (llg && llg->quaternion)
        return fail(MSD_EUNSUPPORTED, "stepper: the fused driver integrates LLG.Cartesian; drive LLG.Quaternion "
                                      "through the operator API (make_pc_step_function)");
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
static int member_begin(msd_stepper* st)
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

static bool spectrum_stale(const msd_stepper* st)
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

static void stepper_end(msd_stepper* st, int nsteps)
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

// capture one step into a graph: kind 0 exact, 1 a learned window's first step, 2 a
// later learned step
static int capture_step(msd_stepper* st, int kind, hipGraph_t* graph, hipGraphExec_t* exec)
{
    st->cap_log.clear();
    st->gnodes[kind].clear();
    st->capturing = true;
    HIPCHK(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
    int rc = kind ? st->launch_learned_step(G.stream, kind == 1) : st->launch_step(G.stream);
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
            if (c.first == p.func) { st->gnodes[kind].push_back({node, c.second}); break; }
    }
    // the 3D form's kernels read their arguments from device tables (Batch3), which
    // refresh() rewrites: its graphs have no step nodes to update
    st->gnodes_ok[kind] = st->gnodes[kind].size() == st->cap_log.size();
    return MSD_OK;
}

// New argument values into the step nodes of an instantiated graph; false when the
// graph's step nodes were not all found (the caller captures again)
static bool update_step_nodes(msd_stepper* st, int kind, hipGraphExec_t exec)
{
    if (!st->gnodes_ok[kind]) return false;
    for (const msd_stepper::StepNode& sn : st->gnodes[kind]) {
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

// nsteps of one kind (0 exact, 1 a learned window's first step, 2 later learned
// steps), graph-replayed when enabled
static int run_kind(msd_stepper* st, int nsteps, int kind)
{
    if (nsteps <= 0) return MSD_OK;
    if (st->use_graph) {
        hipGraphExec_t* exec = kind ? &st->lexec[kind - 1] : &st->exec;
        hipGraph_t* graph = kind ? &st->lgraph[kind - 1] : &st->graph;
        StepArgs* gargs = kind ? &st->lgraph_args[kind - 1] : &st->graph_args;
        const bool scales_moved =
            kind && (st->lgraph_sin[kind - 1] != st->sur->s_in || st->lgraph_sout[kind - 1] != st->sur->s_out);
        bool fresh = *exec && same_launch(st->args, *gargs) && !scales_moved;
        // same kernels and shapes with other values: the step nodes take them in place
        if (fresh && std::memcmp(&st->args, gargs, sizeof(StepArgs)) != 0) {
            fresh = update_step_nodes(st, kind, *exec);
            if (fresh) *gargs = st->args;
        }
        if (!fresh) {
            if (*exec) (void)hipGraphExecDestroy(*exec);
            if (*graph) (void)hipGraphDestroy(*graph);
            *exec = nullptr; *graph = nullptr;
            if (int rc = capture_step(st, kind, graph, exec)) return rc;
            st->graph_builds++;
            *gargs = st->args;
            if (kind) { st->lgraph_sin[kind - 1] = st->sur->s_in; st->lgraph_sout[kind - 1] = st->sur->s_out; }
        }
        for (int i = 0; i < nsteps; ++i) HIPCHK(hipGraphLaunch(*exec, G.stream));
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
    const size_t n = st->s->d.n;
    float *S0 = nullptr, *L0 = nullptr;
    if (int rc = dalloc(&S0, 4 * n)) return rc;
    if (int rc = dalloc(&L0, 3 * n)) { dfree(S0); return rc; }
    const double time0 = st->s->time;
    const uint64_t draw0 = st->th ? st->th->draw : 0;
    const bool lo0 = st->s->lo_valid;
    int rc = copy_dev(S0, st->s->S, 4 * n);
    if (!rc) rc = copy_dev(L0, st->s->Slo, 3 * n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = fail(MSD_EDEVICE, "autotune: events");
    auto restore = [&]() -> int {
        if (int r = copy_dev(st->s->S, S0, 4 * n)) return r;
        if (int r = copy_dev(st->s->Slo, L0, 3 * n)) return r;
        st->s->time = time0;
        st->s->lo_valid = lo0;
        if (st->th) st->th->draw = draw0;
        st->s->version++;
        return MSD_OK;
    };
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
        if ((rc = stepper_begin(st))) break;
        if ((rc = run_steps(st, 2, false))) break;          // capture + warm
        if (hipEventRecord(e0, G.stream) != hipSuccess) { rc = fail(MSD_EDEVICE, "autotune: event"); break; }
        if ((rc = run_steps(st, steps, false))) break;
        if (hipEventRecord(e1, G.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = fail(MSD_EDEVICE, "autotune: event");
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double per = ms / steps;
        if (*best_ms < 0 || per < *best_ms) {
            *best_ms = per; *best_ty = ty; *best_tpb = tpb; *best_split = dip->dev.split;
        }
        if ((rc = restore())) break;
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
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(G.stream);
    dfree(S0); dfree(L0);
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

struct msd_ensemble {
    std::vector<msd_stepper*> m;
    bool compensated = true;
    bool use_graph = true;
    Batch3 b;                        // local-field and LongRange3D members
    std::vector<FilmGroup> films;    // layered-Dipole members, by shape
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
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
        return b.update(r, tys, lds);
    }
    hipError_t evaluate(bool pred, hipStream_t st) const
    {
        hipError_t e = b.evaluate(pred ? R3_PRED : R3_CORR, compensated, st);
        for (const FilmGroup& f : films)
            if (e == hipSuccess) e = f.step(pred, st);
        return e;
    }
};

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
