// C ABI of libmsd.so, core: errors, device set-up, the allocation-free entry point,
// SpinSystem and MagLua's checkpoint byte stream (type ids, wire helpers).  The
// operators, step drivers, surrogate and ensembles live in msd_ops.cpp,
// msd_dipole.cpp, msd_conv3.cpp, msd_stepper.cpp, msd_surrogate.cpp and
// msd_ensemble.cpp; their shared objects in msd_host.h.
#include "msd_host.h"

// ===================================================================== errors
thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// ===================================================================== device
Globals G;

int ensure_init()
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



// upload FP64 host data as FP32 device data
int upload_f32(float* d, const double* h, size_t n)
{
    std::vector<float> t(n);
    for (size_t i = 0; i < n; ++i) t[i] = (float)h[i];
    HIPCHK(hipMemcpyAsync(d, t.data(), n * sizeof(float), hipMemcpyHostToDevice, G.stream));
    HIPCHK(hipStreamSynchronize(G.stream));
    return MSD_OK;
}

int download_f32(double* h, const float* d, size_t n)
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



bool same_dims(const Dims& a, int nx, int ny, int nz) { return a.nx == nx && a.ny == ny && a.nz == nz; }

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

int copy_dev(float* dst, const float* src, size_t n)
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
// IEEE double in host (little-endian) order, chars as one byte (WireOut / WireIn and the
// magic chars in msd_host.h).

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



int wire_finish(const WireOut& w, const char* who, void* buf, int64_t cap, int64_t* len)
{
    *len = (int64_t)w.b.size();
    if (!buf) return MSD_OK;
    if (cap < *len) return fail(MSD_EINVAL, "%s: buffer of %lld bytes, %lld needed", who, (long long)cap, (long long)*len);
    std::memcpy(buf, w.b.data(), w.b.size());
    return MSD_OK;
}

// ENCODE_PREAMBLE of a new object, then SpinOperation::encode (cpu/core/spinoperation.cpp:67-75)
void wire_operator(WireOut& w, int index, const Dims& d, double scale)
{
    w.ch(WIRE_NEW);
    w.i32(index);
    w.ch(0);
    w.i32(d.nx); w.i32(d.ny); w.i32(d.nz);
    w.f64(scale);
}

// the preamble as decodeLuaBaseObject reads it (luabaseobject.cpp:303-350), then
// SpinOperation::decode (spinoperation.cpp:77-93)
int unwire_operator(WireIn& r, const char* who, int index, int* nx, int* ny, int* nz, double* scale)
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

// encodeT<dArray> (luabaseobject.h:101-106: the Lua type and the type id) when tagged,
// then Array<double>::encode (cpu/array/array_core_cpu.h:200-224): one value when all
// sites hold the same value, else the raw doubles.
void wire_darray(WireOut& w, int index, const Dims& d, const double* v, bool tagged)
{
    if (tagged) {
        w.i32(WIRE_LUA_TUSERDATA);
        w.i32(type_id("Array.Double"));
    }
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
    for (int c = 0; c < 4; ++c) wire_darray(w, idx++, s->d, sp.data() + c * n, true);
    std::vector<double> site(n);
    for (float* a : {s->site_alpha, s->site_gamma}) {
        w.i32(a ? 1 : 0);
        if (!a) continue;
        if (int rc = download_f32(site.data(), a, n)) return rc;
        wire_darray(w, idx++, s->d, site.data(), true);
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
int unwire_darray(WireIn& r, const char* who, int index, const Dims& d, std::vector<double>& v, bool tagged)
{
    int32_t tag = WIRE_LUA_TUSERDATA, type = type_id("Array.Double"), pos, nx, ny, nz, flag;
    char magic, version;
    if (tagged && (!r.i32(&tag) || !r.i32(&type))) return fail(MSD_EINVAL, "%s: stream ends inside an array header", who);
    if (!r.ch(&magic) || !r.i32(&pos)) return fail(MSD_EINVAL, "%s: stream ends inside an array header", who);
    if (tag != WIRE_LUA_TUSERDATA || type != type_id("Array.Double"))
        return fail(MSD_EINVAL, "%s: expected an Array.Double (tag %d, type %d)", who, tag, type);
    if (magic == WIRE_OLD)
        return fail(MSD_EUNSUPPORTED, "%s: array %d refers to an earlier object of the stream", who, pos);
    if (magic != WIRE_NEW) return fail(MSD_EINVAL, "%s: malformed stream (magic %d)", who, (int)magic);
    if (pos != index) return fail(MSD_EINVAL, "%s: object index %d, expected %d", who, pos, index);
    if (!r.ch(&version) || !r.i32(&nx) || !r.i32(&ny) || !r.i32(&nz) || !r.i32(&flag))
        return fail(MSD_EINVAL, "%s: stream ends inside an array", who);
    if (version != 0) return fail(MSD_EINVAL, "%s: array version %d", who, (int)version);
    if (!same_dims(d, nx, ny, nz))
        return fail(MSD_EINVAL, "%s: array of %dx%dx%d in a %dx%dx%d system", who, nx, ny, nz, d.nx, d.ny, d.nz);
    if (flag == 1) {
        double c;
        if (!r.f64(&c)) return fail(MSD_EINVAL, "%s: stream ends inside an array", who);
        v.assign(d.n, c);
        return MSD_OK;
    }
    if (flag != 0) return fail(MSD_EINVAL, "%s: array flag %d", who, flag);
    const size_t bytes = (size_t)d.n * sizeof(double);
    if (bytes > r.n - r.at) return fail(MSD_EINVAL, "%s: stream ends inside an array", who);
    v.resize(d.n);
    if (!r.get(v.data(), bytes)) return fail(MSD_EINVAL, "%s: stream ends inside an array", who);
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
        if (int rc = unwire_darray(r, "SpinSystem.decode", idx++, d, sp[c], true)) return rc;
    std::vector<double> site[2];
    for (int k = 0; k < 2; ++k) {
        int32_t has;
        if (!r.i32(&has)) return fail(MSD_EINVAL, "SpinSystem.decode: stream ends before the site arrays");
        if (!has) continue;
        if (int rc = unwire_darray(r, "SpinSystem.decode", idx++, d, site[k], true)) return rc;
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
