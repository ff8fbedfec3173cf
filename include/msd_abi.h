/*
 * msd_abi.h -- C ABI of libmsd.so, the MI355X (gfx950) spin-dynamics hot path
 * behind MagLua's operator API.
 *
 * Conventions
 *   - Every entry point returns int: MSD_OK (0) or a negative MSD_E* code; the
 *     message of the last failure on the calling thread is msd_last_error().
 *     This replaces MagLua's "bool + errormsg, then luaL_error" convention
 *     (src/modules/cpu/core/spinoperation.cpp:436-442).
 *   - Objects are opaque handles with create/destroy (MagLua: refcounted
 *     LuaBaseObject, common/luabaseobject/luabaseobject.h:191-215).
 *   - Sites and coordinates are 0-based; a Lua binding subtracts 1 as MagLua's
 *     own wrappers do (cpu/exchange/spinoperationexchange.cpp:384-387).
 *   - Site index = x + nx*(y + ny*z) (cpu/core/spinsystem.cpp:1125-1142).
 *   - Host arrays are FP64 and SoA; device storage is FP32 SoA.  Calls are
 *     asynchronous on the library's stream except those that return host data.
 *   - Host calls on one handle must be serialised by the caller (MagLua runs a
 *     single Lua state on one thread).
 */
#ifndef MSD_ABI_H
#define MSD_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSD_OK            0
#define MSD_EINVAL       -1   /* bad argument (size, index, name, NULL) */
#define MSD_EDEVICE      -2   /* HIP runtime failure */
#define MSD_EUNSUPPORTED -3   /* valid MagLua input this build does not handle */
#define MSD_EMISMATCH    -4   /* objects of different lattice sizes */
#define MSD_ENOMEM       -5

typedef struct msd_system     msd_system;
typedef struct msd_exchange   msd_exchange;
typedef struct msd_anisotropy msd_anisotropy;
typedef struct msd_thermal    msd_thermal;
typedef struct msd_appliedfield msd_appliedfield;
typedef struct msd_uniaxial   msd_uniaxial;
typedef struct msd_dipole     msd_dipole;
typedef struct msd_llg        msd_llg;
typedef struct msd_stepper    msd_stepper;

/* ------------------------------------------------------------------ library */
const char* msd_last_error(void);
int msd_version(void);
/* Select the device and create the library stream (idempotent). */
int msd_init(int device);
/* Wait for all work queued on the library stream. */
int msd_synchronize(void);
/* Philox4x32-10 on the device, for known-answer checks: out[4*i..] for ctr[4*i..]. */
int msd_philox_device(const uint32_t* ctr, int count, const uint32_t key[2], uint32_t* out);

/* -------------------------------------------------------------- SpinSystem */
/* SpinSystem.new(nx,ny,nz): cpu/core/spinsystem.cpp:29-41,502-566 (alpha=gamma=dt=1, time=0) */
int msd_system_create(int nx, int ny, int nz, msd_system** out);
int msd_system_destroy(msd_system* s);
int msd_system_dims(const msd_system* s, int* nx, int* ny, int* nz);
/* ss:copy() -- spins, parameters and field slots (spinsystem.cpp copy) */
int msd_system_copy(const msd_system* s, msd_system** out);
/* ss:copySpinsTo(dst) -- spins and moments only */
int msd_system_copy_spins_to(const msd_system* src, msd_system* dst);
/* Bulk SpinSystem::set (spinsystem.cpp:1051-1059); ms may be NULL (then ms = |s|). */
int msd_system_set_spins(msd_system* s, const double* x, const double* y, const double* z, const double* ms);
/* x,y,z,ms may individually be NULL. */
int msd_system_get_spins(msd_system* s, double* x, double* y, double* z, double* ms);
/* ss:setSpin({x,y,z},{sx,sy,sz}[,m]): m < 0 means m = |s|; the stored spin has length m. */
int msd_system_set_spin(msd_system* s, int site, double sx, double sy, double sz, double m);
/* ss:spin(site) -> sx, sy, sz, m */
int msd_system_get_spin(msd_system* s, int site, double out[4]);
/* Unit spins uniform on the sphere from a counter-based stream keyed by seed
 * (the build's seeded initial state; MagLua scripts use math.random). */
int msd_system_random_spins(msd_system* s, uint64_t seed);

#define MSD_PARAM_ALPHA 0
#define MSD_PARAM_GAMMA 1
#define MSD_PARAM_DT    2
#define MSD_PARAM_TIME  3
/* setAlpha / setGamma / setTimeStep / setTime and their getters */
int msd_system_set_param(msd_system* s, int which, double v);
int msd_system_get_param(const msd_system* s, int which, double* v);
/* per-site alpha/gamma arrays (SpinSystem site_alpha/site_gamma); NULL clears */
int msd_system_set_site_alpha(msd_system* s, const double* a);
int msd_system_set_site_gamma(msd_system* s, const double* g);

/* Checkpoint byte stream.  MagLua's object encoding: int32 and IEEE double in
 * little-endian order, objects opened by a magic char (111 new, 222 repeat) and
 * their index in the stream's object list (common/luabaseobject/luabaseobject.cpp:140-220).
 *
 * hash32 (common/luabaseobject/factory.cpp:123-131): MurmurHash3_x86_32 of a type
 * name with seed 1000, MagLua's type id ("SpinSystem", "Array.Double"). */
int msd_hash32(const char* name, int32_t* out);
/* SpinSystem::encode (cpu/core/spinsystem.cpp:669-760): the bytes the system appends to
 * a checkpoint buffer whose object list already holds first_index objects (the system
 * takes index first_index, its arrays the next ones; *n_objects counts them all).
 * buf NULL: only *len (and *n_objects) are returned; else cap must be >= *len, and
 * 242 + 48*n bytes always suffice.  Spins are the FP64 values of the compensated device
 * state; ms and the site alpha/gamma arrays are their FP32 device values. */
int msd_system_encode(msd_system* s, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects);
/* SpinSystem::decode (cpu/core/spinsystem.cpp:764-846), read from the object's magic
 * char (what msd_system_encode writes; the Lua type tag and type id before it belong to
 * the container).  Creates *out; *used = bytes read.  Per-site extra Lua data and arrays
 * that repeat an earlier object of the stream are MSD_EUNSUPPORTED. */
int msd_system_decode(const void* buf, int64_t len, int first_index, msd_system** out, int64_t* used, int* n_objects);
/* Exchange::encode/decode (cpu/exchange/spinoperationexchange.cpp:52-113, with
 * SpinOperation::encode/decode, cpu/core/spinoperation.cpp:67-93): scale, normalisation
 * flag, pbc and the pathways in the order they were added; one object.  Same buffer and
 * index conventions as msd_system_encode/decode. */
int msd_exchange_encode(msd_exchange* e, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects);
int msd_exchange_decode(const void* buf, int64_t len, int first_index, msd_exchange** out, int64_t* used, int* n_objects);
/* Anisotropy::encode/decode (cpu/anisotropy/spinoperationanisotropy.cpp:191-240): scale
 * and the ops (site, unit axis, K) in the order they were added; decode stores each op
 * as read, as addAnisotropy does (:172-189), so a round trip keeps every bit. */
int msd_anisotropy_encode(msd_anisotropy* a, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects);
int msd_anisotropy_decode(const void* buf, int64_t len, int first_index, msd_anisotropy** out, int64_t* used, int* n_objects);
/* Dipole::encode/decode (cpu/dipole/spinoperationdipole.cpp:35-59 around
 * cpu/longrange/spinoperationlongrange.cpp:55-92): scale, truncation, strength and unit
 * cell; the tensor is rebuilt from them (a matrix from setMatrix is not in the stream,
 * as in MagLua). */
int msd_dipole_encode(msd_dipole* d, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects);
int msd_dipole_decode(const void* buf, int64_t len, int first_index, msd_dipole** out, int64_t* used, int* n_objects);

/* ss:resetFields() -- SpinSystem::zeroFields, spinsystem.cpp:968-980 */
int msd_system_reset_fields(msd_system* s);
/* ss:sumFields() -- SpinSystem::sumFields, spinsystem.cpp:853-871 */
int msd_system_sum_fields(msd_system* s);
/* ss:fieldArrayX/Y/Z(slot) copied to the host; MSD_EINVAL if the slot is unknown */
int msd_system_get_field(msd_system* s, const char* slot, double* hx, double* hy, double* hz);
/* 1 if the named slot is registered and marked used since the last reset */
int msd_system_slot_used(const msd_system* s, const char* slot, int* used);
/* ss:netMoment(): sums (not means) -- cpu/core/spinsystem_luafuncs.lua:17-54.
 * out = {Sx, Sy, Sz, |S|, Sx2, Sy2, Sz2, |S2|}, FP64 accumulation on the device. */
int msd_system_net_moment(msd_system* s, double out[8]);
/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture set-up, tight serving loops, harnesses that forbid
 * hipMalloc): msd_system_net_moment over spins the caller holds in device memory,
 * allocating nothing.  spins = device FP32 [3][n] (x, then y, then z), lo = optional
 * device FP32 [3][n] low words added to them (NULL: none), out = host {Sx, Sy, Sz, |S|,
 * Sx2, Sy2, Sz2, |S2|} as msd_system_net_moment.  The partial sums live in a fixed
 * device array of 6*1024+8 doubles (49,216 bytes) per GPU, reserved for the process;
 * n is limited to MSD_NOALLOC_MAX_SITES.  msd_init must have been called.  Calls
 * share that array: calls that overlap on one GPU (several host threads) race. */
#define MSD_NOALLOC_MAX_SITES (1 << 30)
int msd_net_moment_noalloc(const float* spins, const float* lo, int n, double out[8]);
/* sum_i h_slot(i) . s(i), FP64 accumulation (the scripts' fieldArray:dot(spinArray) sums) */
int msd_system_field_dot(msd_system* s, const char* slot, double* out);
/* ss:diff(other): {sum|dx|, sum|dy|, sum|dz|, |(those three)|} (SpinSystem::diff, spinsystem.cpp:301-307) */
int msd_system_diff(msd_system* a, msd_system* b, double out[4]);

/* ---------------------------------------------------------------- Exchange */
/* Exchange.new: cpu/exchange/spinoperationexchange.cpp:28-39 (pbc 1,1,1, normalised) */
int msd_exchange_create(int nx, int ny, int nz, msd_exchange** out);
int msd_exchange_destroy(msd_exchange* e);
int msd_exchange_set_pbc(msd_exchange* e, int px, int py, int pz);
/* ex:normalizeMoments(bool) */
int msd_exchange_set_normalize(msd_exchange* e, int yes);
int msd_exchange_set_scale(msd_exchange* e, double s);
/* ex:add({a},{b},J): l_addpath, spinoperationexchange.cpp:368-429.  Sets *added to 0
 * when a coordinate lies outside a non-periodic dimension (the path is dropped). */
int msd_exchange_add(msd_exchange* e, const int a[3], const int b[3], double J, int* added);
/* Exchange::addPath(from, to, J), spinoperationexchange.cpp:252-307 (J == 0 ignored) */
int msd_exchange_add_path(msd_exchange* e, int from_site, int to_site, double J);
/* Bulk form of addPath for count paths (a loop of ex:addPath in a script). */
int msd_exchange_add_paths(msd_exchange* e, const int* from_site, const int* to_site, const double* J, int count);
int msd_exchange_num_paths(const msd_exchange* e, int* n);
/* ex:apply(ss): h[to] += J s[from] / (m[from] m[to]) into slot "Exchange" */
int msd_exchange_apply(msd_exchange* e, msd_system* s);

/* -------------------------------------------------------------- Anisotropy */
int msd_anisotropy_create(int nx, int ny, int nz, msd_anisotropy** out);
int msd_anisotropy_destroy(msd_anisotropy* a);
int msd_anisotropy_set_scale(msd_anisotropy* a, double s);
/* ani:add({site}, {axis}, K): cpu/anisotropy/spinoperationanisotropy.cpp:367-413.
 * The axis is normalised; a zero axis adds nothing (*added = 0). */
int msd_anisotropy_add(msd_anisotropy* a, int site, const double axis[3], double K, int* added);
/* Bulk form of add for count sites: axes[3*i..3*i+2], K[i] (a loop of ani:add). */
int msd_anisotropy_add_many(msd_anisotropy* a, const int* site, const double* axes, const double* K, int count);
/* ani:apply(ss): Anisotropy::apply, spinoperationanisotropy.cpp:247-284, slot "Anisotropy" */
int msd_anisotropy_apply(msd_anisotropy* a, msd_system* s);

/* ----------------------------------------------------- Anisotropy.Uniaxial
 * Anisotropy.Uniaxial.new(n) (cpu/anisotropy/spinoperationanisotropy_uniaxial.cpp):
 * add(site, axis, K1, K2) appends an op with the axis normalised, a zero axis
 * adding nothing (l_add :387-431; its Lua binding passes K1 as K2 when a fourth
 * number is given, which the host mirror reproduces); numberOfAxes / axisAt
 * (0-based here); apply (:248-298) writes slot "UniaxialAnisotropy" with
 * h = n (2 K1 d + 4 K2 d^3) scale, d = sx nx + sy ny + sz nx as the reference's
 * line 287 computes it, s = spin / ms.  Sites are 0-based here. */
int msd_uniaxial_create(int nx, int ny, int nz, msd_uniaxial** out);
int msd_uniaxial_destroy(msd_uniaxial* u);
int msd_uniaxial_set_scale(msd_uniaxial* u, double scale);
int msd_uniaxial_add(msd_uniaxial* u, int site, const double* axis, double K1, double K2, int* added);
int msd_uniaxial_num_axes(const msd_uniaxial* u, int* n);
int msd_uniaxial_axis_at(const msd_uniaxial* u, int i, int* site, double* axis, double* K1, double* K2);
int msd_uniaxial_apply(msd_uniaxial* u, msd_system* s);
/* AnisotropyUniaxial::encode / decode (cpu/anisotropy/spinoperationanisotropy_uniaxial.cpp:
 * 181-234): ops in order of addition; decode passes each axis through addAnisotropy's
 * division by its squared length (:172-174). */
int msd_uniaxial_encode(msd_uniaxial* u, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects);
int msd_uniaxial_decode(const void* buf, int64_t len, int first_index, msd_uniaxial** out, int64_t* used, int* n_objects);

/* ------------------------------------------------------------ AppliedField
 * AppliedField.new(n) (cpu/appliedfield/spinoperationappliedfield.cpp:18-25; x y z
 * setX setY setZ :127-190, set/add/get in appliedfield_luafuncs.lua:7-40) and,
 * heterogeneous = 1, AppliedField.Heterogeneous
 * (spinoperationappliedfield_heterogeneous.cpp; per-site set/get,
 * appliedfield_heterogeneous_luafuncs.lua:7-31).  Sites are 0-based here. */
int msd_appliedfield_create(int nx, int ny, int nz, int heterogeneous, msd_appliedfield** out);
int msd_appliedfield_destroy(msd_appliedfield* a);
int msd_appliedfield_set(msd_appliedfield* a, double x, double y, double z);
int msd_appliedfield_get(const msd_appliedfield* a, double* xyz);
int msd_appliedfield_set_scale(msd_appliedfield* a, double scale);
int msd_appliedfield_set_site(msd_appliedfield* a, int site, double x, double y, double z);
int msd_appliedfield_get_site(const msd_appliedfield* a, int site, double* xyz);
/* af:apply(ss): uniform, the Lua apply (appliedfield_luafuncs.lua:46-57) sets slot
 * "Zeeman" to B*scale; heterogeneous (spinoperationappliedfield_heterogeneous.cpp:88-99)
 * adds scale*h to slot "Zeeman_Heterogeneous" */
int msd_appliedfield_apply(msd_appliedfield* a, msd_system* s);
/* AppliedField::encode / decode (cpu/appliedfield/spinoperationappliedfield.cpp:33-58):
 * B; AppliedField_Heterogeneous (spinoperationappliedfield_heterogeneous.cpp:39-72): three
 * per-site arrays written by their own encode (four objects).  heterogeneous selects the
 * decoded form. */
int msd_appliedfield_encode(msd_appliedfield* a, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects);
int msd_appliedfield_decode(const void* buf, int64_t len, int first_index, int heterogeneous, msd_appliedfield** out,
                            int64_t* used, int* n_objects);

/* ----------------------------------------------------------------- Thermal */
int msd_thermal_create(int nx, int ny, int nz, msd_thermal** out);
int msd_thermal_destroy(msd_thermal* t);
/* temp:set(T) */
int msd_thermal_set_temperature(msd_thermal* t, double T);
int msd_thermal_set_scale(msd_thermal* t, double s);
/* temp:scaleSite({site}, v) */
int msd_thermal_scale_site(msd_thermal* t, int site, double v);
/* Counter-based noise stream: (seed, draw, site) -> 3 normals.  apply() uses the
 * draw counter and then increments it (MagLua: Random.* RNG objects).  A new Thermal
 * is keyed by a seed from /dev/urandom, as an unseeded MagLua RNG is
 * (common/random/random.cpp:29-62); set_seed gives a reproducible stream. */
int msd_thermal_set_seed(msd_thermal* t, uint64_t seed, uint64_t draw);
/* Thermal::encode / Thermal::decode (cpu/thermal/spinoperationthermal.cpp:51-80): the
 * operator preamble, version 0, the temperature and the per-site scale array (two
 * objects: the Thermal and its Array.Double).  The noise stream is not in the
 * reference's stream: a decoded Thermal has a seed from /dev/urandom, as a new one. */
int msd_thermal_encode(msd_thermal* t, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects);
int msd_thermal_decode(const void* buf, int64_t len, int first_index, msd_thermal** out, int64_t* used, int* n_objects);
int msd_thermal_get_draw(const msd_thermal* t, uint64_t* draw);
int msd_thermal_get_seed(const msd_thermal* t, uint64_t* seed);
/* temp:apply(ss): Thermal::apply, cpu/thermal/spinoperationthermal.cpp:135-191, slot "Thermal" */
int msd_thermal_apply(msd_thermal* t, msd_system* s);

/* ------------------------------------------------------- Dipole (LongRange) */
/* Dipole.new: layered 2D FFT convolution, cpu/longrange/spinoperationlongrange.cpp:26-46.
 * Needs nx even and nx, ny with prime factors in {2,3,5}. */
int msd_dipole_create(int nx, int ny, int nz, msd_dipole** out);
int msd_dipole_destroy(msd_dipole* d);
/* dip:setStrength(g) */
int msd_dipole_set_strength(msd_dipole* d, double g);
int msd_dipole_set_scale(msd_dipole* d, double s);
/* dip:setTruncation: gmax = {radial, x, y, z}; the x image limit reads gmax[2] and
 * the y limit gmax[3] as in getWAB_range (cpu/dipole/dipolesupport.cpp:71-127).
 * gmax[0] = -1 is setTruncation(math.huge) (the reference's marker,
 * dipolesupport.cpp:511-519): the sum over the images |X|, |Y| <= 4999 that the
 * reference's extrapolation loop ends with (:756-840), ~1e8 image terms per tensor
 * element in FP64 on the device. */
int msd_dipole_set_truncation(msd_dipole* d, const int gmax[4]);
/* dip:setUnitCell(A,B,C): ABC row-major {A;B;C} */
int msd_dipole_set_unit_cell(msd_dipole* d, const double ABC[9]);
/* dip:getMatrix/setMatrix("XX".."ZZ", {ox,oy,oz}, v): which 0..5 = XX XY XZ YY YZ ZZ,
 * oz in [-(nz-1), nz-1] (LongRange::getAB/setAB, spinoperationlongrange.cpp:153-210). */
int msd_dipole_get_matrix(msd_dipole* d, int which, int ox, int oy, int oz, double* v);
int msd_dipole_set_matrix(msd_dipole* d, int which, int ox, int oy, int oz, double v);
/* dip:apply(ss): LongRange::apply, spinoperationlongrange.cpp:354-430, slot "Dipole".
 * Also LongRange2D:apply (cpu/longrange2d/spinoperationlongrange2d.cpp:448-531) for an
 * object made by msd_lr2d_create, into its own slot. */
int msd_dipole_apply(msd_dipole* d, msd_system* s);
/* No counterpart in the reference: the transform schedule in use -- rows per row
 * tile, threads per row workgroup, and split = rows per tile when the row kernels
 * take the first radix-ty stage of the y transform (0: full-length columns). */
int msd_dipole_schedule(const msd_dipole* d, int* rows_per_tile, int* threads, int* split);

/* ------------------------------------------------------------- LongRange2D */
/* LongRange2D.new / Dipole2D.new (cpu/longrange2d/spinoperationlongrange2d.cpp:26-48,
 * cpu/dipole2d/spinoperationdipole2d.cpp:19-40): a layered convolution with a full real
 * tensor T[a][b][dest][src](x, y) for every ordered pair of layers (not only a function of
 * the layer offset), per-layer source-moment and destination-field scales (default 1)
 * and the operator scale (msd_dipole_set_scale).  The field of layer d is
 *   H_a[d] = scale dfs[d] sum_s sms[s] sum_b T_ab[s][d] (*) m_b[s],
 * (*) the periodic in-plane convolution: the reference's apply reads the pair stored under
 * [s][d] for destination d and source s (spinoperationlongrange2d.cpp:463-526), and so does
 * this one.  The object is an msd_dipole: msd_dipole_apply, msd_dipole_set_scale,
 * msd_dipole_destroy, the step drivers and the surrogate accept it; the lattice-sum entry
 * points (strength, truncation, unit cell, get/set_matrix, encode) refuse it.  slot names
 * the field slot ("LongRange2D" if null or empty; Dipole2D uses "Dipole2D").  Same size
 * limits as msd_dipole_create. */
int msd_lr2d_create(int nx, int ny, int nz, const char* slot, msd_dipole** out);
/* lr:setTensorArray(dest, src, AB, array) / lr:tensorArray(dest, src, AB)
 * (spinoperationlongrange2d.cpp:168-224,585-602): dest, src 0-based layers, ab 0..8 =
 * XX XY XZ YX YY YZ ZX ZY ZZ, a = nx*ny doubles, x fastest.  A set takes effect at the
 * next apply or step (the reference's setCompileRequired). */
int msd_lr2d_set_tensor(msd_dipole* d, int dest, int src, int ab, const double* a);
int msd_lr2d_get_tensor(msd_dipole* d, int dest, int src, int ab, double* a);
/* setSourceMomentScale / setDestinationFieldScale for all nz layers at once (:533-568);
 * a null array leaves that set unchanged. */
int msd_lr2d_set_layer_scales(msd_dipole* d, const double* src_moment, const double* dest_field);
int msd_lr2d_get_layer_scales(msd_dipole* d, double* src_moment, double* dest_field);
/* Dipole2D makeData (cpu/dipole2d/dip2d_luafuncs.lua:537-668, point tensor
 * cpu/dipole2d/demag.cpp:3-62): fills every T[dest][src] with the truncated box sum of
 * the dipole tensor.  ABC = [nz][3][3] cell vectors (A, B, C) per layer, offset = [nz][3],
 * trunc = {radial, x, y, z} (the radial value is not used, as in the reference).
 * physical = 0 reproduces the reference's tensors: it evaluates source and destination
 * at the same in-plane cell, so only the layers' vertical coupling enters, repeated over
 * the box.  physical = 1 places the destination at the origin (the truncated in-plane
 * dipole sum).  The file cache of the reference (GAB_*.lua) is not kept. */
int msd_lr2d_build_dipole2d(msd_dipole* d, const double* ABC, const double* offset, const int trunc[4], int physical);
/* Magnetostatics2D makeData (cpu/magnetostatics2d/mag2d_luafuncs.lua:598-750) into a
 * LongRange2D operator: ABC [nz][3][3] unit cells, grain [nz][3] prism sizes, trunc =
 * {truncation, X, Y, Z}; the prism-prism tensors of cpu/magnetostatics2d/demag.cpp, -N
 * summed over the images of each cell, stored under the labels [source][destination]. */
int msd_lr2d_build_magnetostatics2d(msd_dipole* d, const double* ABC, const double* grain, const int trunc[4]);

/* ------------------------------------------------------------- LongRange3D */
/* LongRange3D.new: 3D FFT convolution, cpu/longrange3d/spinoperationlongrange3d.cpp:331-432.
 * open = 0: periodic in x, y, z; the tensor defaults to the dipole lattice sum over
 * periodic images within radius R (default 1000, as Dipole's truncation).
 * open = 1: open boundaries by zero padding to (2nx, 2ny, 2nz); the tensor is the
 * bare dipole kernel.  Transform lengths need prime factors in {2,3,5}. */
typedef struct msd_longrange3d msd_longrange3d;
int msd_longrange3d_create(int nx, int ny, int nz, int open, msd_longrange3d** out);
int msd_longrange3d_destroy(msd_longrange3d* l);
int msd_longrange3d_set_strength(msd_longrange3d* l, double g);
int msd_longrange3d_set_scale(msd_longrange3d* l, double s);
int msd_longrange3d_set_radius(msd_longrange3d* l, int R);
/* tensor element which = 0..5 (XX XY XZ YY YZ ZZ; YX = XY etc.) at FFT-grid offset
 * (ox,oy,oz), taken modulo the grid (LongRange3D setMatrix/getMatrix).  A custom
 * tensor must stay even, T(-r) = T(r). */
int msd_longrange3d_get_matrix(msd_longrange3d* l, int which, int ox, int oy, int oz, double* v);
int msd_longrange3d_set_matrix(msd_longrange3d* l, int which, int ox, int oy, int oz, double v);
/* lr:apply(ss): slot "LongRange3D" */
int msd_longrange3d_apply(msd_longrange3d* l, msd_system* s);
/* Pass schedule (no counterpart in the reference: its FFTs are FFTW plans).  The
 * compile-time y pass and z pass forms each have a few kx widths per workgroup;
 * wy / wz = 0 keeps a width.  A width with no compiled instance at this transform
 * length, or over the LDS per workgroup, is MSD_EUNSUPPORTED.  The field does not
 * depend on the widths beyond FP32 rounding.  msd_longrange3d_pass_widths returns
 * the widths in use and the ones each pass form offers, default first, in
 * y_choices[ny] / z_choices[nz] (at most 8 each; ny or nz = 0: the runtime-length
 * pass, whose width is fixed); the choice arrays and counts may be NULL. */
int msd_longrange3d_set_pass_widths(msd_longrange3d* l, int wy, int wz);
int msd_longrange3d_pass_widths(msd_longrange3d* l, int* wy, int* wz, int* y_choices, int* ny, int* z_choices,
                                int* nz);

/* --------------------------------------------------------------------- LLG */
/* LLG.Cartesian: cpu/llg/llgcartesian.cpp:90-278 (thermalOnlyFirstTerm = true) */
int msd_llg_create(msd_llg** out);
/* LLG.Quaternion: cpu/llg/llgquat.cpp:127-373 (rotation of S about the damped
 * field).  Shares the flags and msd_llg_apply.  The fused Heun step and the fused
 * Runge-Kutta stages take it (each turn added as the increment R S - S, which the
 * compensated state keeps). */
int msd_llg_create_quaternion(msd_llg** out);
int msd_llg_destroy(msd_llg* l);
int msd_llg_set_thermal_only_first_term(msd_llg* l, int yes);
int msd_llg_set_disable_renormalization(msd_llg* l, int yes);
int msd_llg_set_scales(msd_llg* l, double precess_scale, double damp_scale);
/* LLG::encode / LLG::decode (cpu/llg/llg.cpp:50-88), for LLG.Cartesian and
 * LLG.Quaternion alike (quaternion selects the decoded type): version 1 with
 * thermalOnlyFirstTerm, disableRenormalization and the two scales; version 0 streams
 * (no scales) decode with scales 1. */
int msd_llg_encode(msd_llg* l, int first_index, void* buf, int64_t cap, int64_t* len, int* n_objects);
int msd_llg_decode(const void* buf, int64_t len, int first_index, int quaternion, msd_llg** out, int64_t* used,
                   int* n_objects);
/* llg:apply(from [,scale] [,dmdt] [,to] [,advance]): cpu/llg/llg.cpp:100-171.
 * dmdt and to may be NULL (= from). */
int msd_llg_apply(msd_llg* l, msd_system* from, double scale, msd_system* dmdt, msd_system* to, int advance);

/* ----------------------------------------------------- fused step driver */
/* The step closure of common/scripts/PredictorCorrector.lua:54-109 with one
 * corrector pass (Heun) and the thermal kick of :98-103, over the field sum
 * Exchange + Anisotropy + Dipole, as fused device kernels.  Any operator may be
 * NULL.  The operators' configuration is captured at create time. */
int msd_stepper_create(msd_system* s, msd_exchange* ex, msd_anisotropy* ani, msd_dipole* dip,
                       msd_thermal* th, msd_llg* llg, msd_stepper** out);
int msd_stepper_destroy(msd_stepper* st);
/* Replay steps through a captured HIP graph (default 1) */
int msd_stepper_set_graph(msd_stepper* st, int yes);
/* Carry the spin state as FP32 hi + lo words with error-free updates (default 1),
 * or as plain FP32 (0).  Field arithmetic is FP32 either way. */
int msd_stepper_set_compensated(msd_stepper* st, int yes);
/* Use a LongRange3D as the long-range field of this step driver (NULL removes it);
 * not combinable with a Dipole. */
int msd_stepper_set_longrange3d(msd_stepper* st, msd_longrange3d* l);
/* Add an applied field to the fused step's field sum (a = NULL removes all);
 * changes made to it later are picked up by the next msd_stepper_heun */
int msd_stepper_add_applied_field(msd_stepper* st, msd_appliedfield* a);
/* Anisotropy.Uniaxial in the fused step's field sum (u = NULL removes it); ops added
 * later are picked up by the next msd_stepper_heun */
int msd_stepper_set_uniaxial(msd_stepper* st, msd_uniaxial* u);
/* Advance nsteps Heun steps; asynchronous.  The fused kernels read the step's values
 * (temperature, dt, alpha, gamma, applied fields, ...) from a device copy that is
 * stored on the stream when they change, so a script that calls temp:set(T) or
 * ss:setTimeStep(dt) between steps (doc/manual/BasicMicromagneticSimulation.lua)
 * replays the same graph; only a change of kernels or shapes captures a new one.
 * A call of nsteps >= 16 replays a graph of 16 steps nsteps/16 times and the one-step
 * graph for the rest (one graph launch per 16 steps). */
int msd_stepper_heun(msd_stepper* st, int nsteps);
/* make_rk_step_function(ss, kind, calcField, llg, temp) steps (common/scripts/RungeKutta.lua:
 * 57-214) with fused stage kernels: kind names a Butcher table of the script (RK1, RK2,
 * K3, RK3, RK4, RK4_38, RK6, BS3); each stage is the field at the stage spins (and with
 * a Dipole its column pass) and the Butcher row as chained renormalised LLG sub-steps
 * (Cartesian Euler steps, or LLG.Quaternion turns by c dt), then the thermal kick.
 * Local and layered-Dipole drivers; the state is carried as plain FP32 (the
 * compensated low words are dropped). */
int msd_stepper_rk(msd_stepper* st, const char* kind, int nsteps);
/* Number of step graphs captured and instantiated so far (no counterpart in the
 * reference; diagnostics of the graph reuse above). */
int msd_stepper_graph_builds(const msd_stepper* st, long* n);
/* Kernel-level timing on the library stream, names[i]/avg_ms[i] for up to cap
 * kernels.  Layered and local forms: each launch of one Heun step is repeated
 * nsteps times back to back between two HIP events (the kernels are idempotent
 * in their inputs), leaving the state one step ahead.  LongRange3D form: nsteps
 * more steps with events around every launch. */
int msd_stepper_profile(msd_stepper* st, int nsteps, int cap, char names[][32], double* avg_ms, int* count);

/* -------------------------------------------------------------- ensembles */
/* Independent replicas -- SpinSystems of uneven sizes with their own operators,
 * temperatures and seeds, each described by a step driver of the local or
 * LongRange3D form -- advanced together: every pass of a Heun step is one launch
 * over all replicas.  No reference counterpart: MagLua loops
 * SpinOperation::apply over same-sized systems (cpu/core/spinoperation.cpp:288-295).
 * Members must outlive the ensemble. */
/* Autotuner (no counterpart in the reference): times `steps` steps of the layered
 * driver for every compiled row-tile shape (rows per tile, threads per workgroup)
 * in both y-transform schedules -- split (the row kernels take the first radix-ty
 * stage across their tile's rows, best_split = ty) and full columns (best_split = 0)
 * -- and keeps the fastest; the spin state, time and thermal draw are restored.
 * Returns the chosen shape and its ms per step (-1 when only runtime shapes exist). */
int msd_stepper_autotune(msd_stepper* st, int steps, int* best_ty, int* best_tpb, int* best_split, double* best_ms);
/* Autotuner of a LongRange3D step driver (no counterpart in the reference): times
 * `steps` steps for every pair of compiled y-pass and z-pass kx widths of its
 * LongRange3D and keeps the fastest (msd_longrange3d_set_pass_widths); the state is
 * restored.  Returns the widths and their ms per step (-1 when there is no choice). */
int msd_stepper_autotune_passes(msd_stepper* st, int steps, int* best_wy, int* best_wz, double* best_ms);

/* ---------------------------------------------------------------- surrogate
 * No counterpart in the reference: a learned predictor of the dipole field's
 * change between exact refreshes (the field itself is the layered
 * LongRange::apply, cpu/longrange/spinoperationlongrange.cpp:354-430, or the 3D
 * LongRange3D::apply, cpu/longrange3d/spinoperationlongrange3d.cpp:369-432).
 * Three 5x5x3 convolution layers, 3 -> 16 -> 16 -> 3 channels, no bias, identity
 * activation, periodic in x and y (zero-padded there too after
 * msd_surrogate_set_open(s, 1)) and zero-padded in z, BF16 MFMA with FP32
 * accumulation:
 *   dH = s_out * L3(L2(L1(s_in * (m - m_ref)))).
 * Weights of layer l are [75 taps][cout][cin] doubles, tap = (dz+1)*25 + (dy+2)*5
 * + (dx+2) for offsets dz in -1..1, dy and dx in -2..2.  nx must be a multiple
 * of 32.  Training minimises mean((dH - target)^2) / sigma2 with Adam. */
typedef struct msd_surrogate msd_surrogate;
int msd_surrogate_create(int nx, int ny, int nz, msd_surrogate** out);
int msd_surrogate_destroy(msd_surrogate* s);
/* open = 1: zero outside the lattice in x and y (a lattice with open boundaries);
 * 0 (the default): periodic in x and y */
int msd_surrogate_set_open(msd_surrogate* s, int open);
int msd_surrogate_set_weights(msd_surrogate* s, int layer, const double* w);
int msd_surrogate_get_weights(const msd_surrogate* s, int layer, double* w);
int msd_surrogate_set_scales(msd_surrogate* s, double s_in, double s_out, double sigma2);
int msd_surrogate_get_scales(const msd_surrogate* s, double* s_in, double* s_out, double* sigma2);
/* input dm given on the host ([3][n] SoA), or as m - m_ref of two systems */
int msd_surrogate_set_input(msd_surrogate* s, const double* dm);
int msd_surrogate_set_input_systems(msd_surrogate* s, const msd_system* m, const msd_system* m_ref);
/* training target: given on the host ([3][n]), or the exact Dipole / LongRange3D
 * field of m - m_ref */
int msd_surrogate_set_target(msd_surrogate* s, const double* t);
int msd_surrogate_set_target_dipole(msd_surrogate* s, msd_dipole* d, const msd_system* m, const msd_system* m_ref);
int msd_surrogate_set_target_longrange3d(msd_surrogate* s, msd_longrange3d* l, const msd_system* m,
                                         const msd_system* m_ref);
/* dH of the current input; out ([3][n]) may be NULL */
int msd_surrogate_forward(msd_surrogate* s, double* out);
/* loss of the current input/target and the weight gradients (msd_surrogate_get_grads) */
int msd_surrogate_loss_grads(msd_surrogate* s, double* loss);
int msd_surrogate_get_grads(const msd_surrogate* s, int layer, double* g);
int msd_surrogate_adam(msd_surrogate* s, double lr);
int msd_surrogate_reset_optimizer(msd_surrogate* s);
int msd_surrogate_train_step(msd_surrogate* s, double lr, double* loss);
/* max|H_ref + dH(m - m_ref) - H(m)| / max|H(m)| with the exact Dipole (or
 * LongRange3D) fields, and (frozen_err, may be NULL) the same for H_ref alone */
int msd_surrogate_field_error(msd_surrogate* s, msd_dipole* d, const msd_system* m, const msd_system* m_ref,
                              double* rel_err, double* frozen_err);
int msd_surrogate_field_error_longrange3d(msd_surrogate* s, msd_longrange3d* l, const msd_system* m,
                                          const msd_system* m_ref, double* rel_err, double* frozen_err);
/* Learned dipole field in the step driver (a Dipole or LongRange3D driver): windows of
 * K steps on H_ref + dH(m - m_ref), H_ref the exact field of the window's starting
 * state m_ref.  Each window is checked when it ends: the exact field of the state it
 * reached against the surrogate's prediction, max|pred - exact| / max|exact|.  A
 * window within tol stands (its exact field starts the next one); otherwise its
 * starting state and thermal draw are restored and it runs again as exact steps.
 * tol = 0 makes every window exact.  sur = NULL returns to exact steps. */
int msd_stepper_set_surrogate(msd_stepper* st, msd_surrogate* sur, int K, double tol);
int msd_stepper_surrogate_stats(const msd_stepper* st, long* windows, long* exact_windows, long* fallbacks,
                                double* last_residual, double* max_residual);

typedef struct msd_ensemble msd_ensemble;
/* Members must not share a SpinSystem, a LongRange3D or a Thermal (MSD_EINVAL). */
int msd_ensemble_create(msd_stepper* const* members, int count, msd_ensemble** out);
int msd_ensemble_destroy(msd_ensemble* e);
int msd_ensemble_set_graph(msd_ensemble* e, int yes);
int msd_ensemble_set_compensated(msd_ensemble* e, int yes);
int msd_ensemble_heun(msd_ensemble* e, int nsteps);
/* nsteps more steps with HIP events around every pass group of a step (the
 * LongRange3D members' y / z / inverse-y / row passes, the film groups' passes);
 * names[i] and avg_ms[i] per pass group, *count of them (at most cap). */
int msd_ensemble_profile(msd_ensemble* e, int nsteps, int cap, char names[][32], double* avg_ms, int* count);
/* Autotuner of an ensemble's LongRange3D pass widths (no counterpart in the
 * reference), timed in the ensemble's own launches: members grouped by lattice (size
 * and boundaries); one sweep over the groups tries every compiled (y, z) kx width
 * pair on a group with the others held and keeps the fastest by `steps` steps.  The
 * members' states are restored; the widths chosen are each LongRange3D's
 * (msd_longrange3d_pass_widths).  Returns the ms per step reached (-1: no LongRange3D
 * member). */
int msd_ensemble_autotune_passes(msd_ensemble* e, int steps, double* best_ms);

#ifdef __cplusplus
}
#endif
#endif
