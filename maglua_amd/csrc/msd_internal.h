// Internal declarations shared by the C-ABI host layer (msd_abi.cpp) and the
// gfx950 kernels (*.hip).  Device storage is FP32 structure-of-arrays; site
// index = x + nx*(y + ny*z).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

namespace msd {

// ------------------------------------------------------------ FFT plan
// Mixed-radix Stockham plan for one line length L: radices applied in order.
struct FFTPlan {
    int L;
    int nst;
    int rad[16];
};
bool make_plan(int L, FFTPlan& p);

// ------------------------------------------------------------ exchange
// Per-site class table: a class is the ordered list of couplings (dl, J) a
// target site receives, with the source at linear index target + dl.  Sites on
// a periodic face see wrapped offsets and so fall in classes of their own.
struct ExEntry {
    int dl;
    float J;
};
// Class tables of at most this many classes (plus one) and entries are what the
// fused step kernels hold in LDS; the fused form uses a class table only then
// (msd_exchange::fused_dev), so in those kernels mode 1 always reads LDS.
constexpr int EX_LDS_CLS = 64, EX_LDS_ENT = 512;
struct ExchangeDev {
    int mode;               // 0 none, 1 class table (fused kernels: staged in LDS), 2 CSR by target
    int normalize;
    const uint8_t* cls;     // [n] or nullptr (all class 0)
    const int* cls_start;   // [ncls+1]
    const ExEntry* ent;     // class entries
    int ncls, nent;         // class table sizes (mode 1)
    const int* row;         // CSR [n+1]
    const int* col;         // CSR from-site
    const float* val;       // CSR J*scale
};

// ---------------------------------------------------------- anisotropy
// h = 2 A s / ms^2 with A = sum_ops K n n^T (scale folded in).
struct AnisoDev {
    int mode;               // 0 none, 1 uniform matrix, 2 per-site axis+K, 3 per-site matrix
    float u[6];             // uniform: xx xy xz yy yz zz
    const float4* axk;      // mode 2: (nx, ny, nz, K)
    const float* mat;       // mode 3: [6][n]
};

// ------------------------------------------------------------- thermal
struct ThermalDev {
    int on;
    float T_scale;          // temperature * global scale
    const float* site;      // per-site scale or nullptr (1)
    uint64_t seed;
};

// ----------------------------------------------------------------- LLG
struct LLGDev {
    int thermal_first;
    int renorm;
    int quaternion;         // LLG.Quaternion (rotation) instead of LLG.Cartesian (Euler)
    float ps, ds;           // precession / damping scales
    float alpha, gamma;     // uniform values
    const float* site_alpha;
    const float* site_gamma;
};

// ------------------------------------------------------ layered dipole
constexpr int MSD_MAX_LAYERS = 8;   // layers of the layered column kernels
struct DipoleDev {
    int nx, ny, nz, nxh;
    int ty;                 // rows per row-tile
    int tpb;                // threads per row-kernel workgroup
    FFTPlan px, py;
    const float2* twx;      // exp(-2 pi i k / nx), k < nx
    const float2* twy;      // exp(-2 pi i k / ny), k < ny
    // Split y transform (split = R > 0, ny = R * ly): a row tile holds the R rows
    // b + ly*a (a < R) of one layer, b < ly.  The row kernels do the length-R DFT
    // across the tile's rows and the twiddle exp(-+2 pi i b k1 / ny); the column
    // kernel does length-ly transforms for each (kx, k1), ky = k1 + R*k2.  Between
    // the passes the spectra hold, per (comp, layer, kx), [b][k1] with k1 fastest.
    // split = 0: consecutive rows per tile, [y] per (comp, layer, kx), full-length
    // column transforms.
    int split, ly;
    float2* Mh;             // spin spectrum [3][nz][nxh][ny]
    float2* Hh;             // field spectrum [3][nz][nxh][ny]
    // tensor words (XX XY XZ YY YZ ZZ; XZ,YZ imaginary): [6][nz][nxh][ny], or with a
    // split transform [nxh][R][6][nz][ly] (each column workgroup's words contiguous)
    const float* Tw;
    // per-layer-pair form (LongRange2D and its builders); null for the lattice-sum Dipole
    const float2* Tp;       // complex words [9][nz][nz][nxh][ny], 1/(nx ny) folded in
    const float* lsc;       // [2][MSD_MAX_LAYERS]: source-moment scales, destination-field scales
    size_t lds_rows;        // dynamic LDS bytes of a row kernel
    size_t lds_cols;        // dynamic LDS bytes of a column kernel
};

struct Dims {
    int nx, ny, nz, n;
};

// ------------------------------------------------------------ launches
// All launches go to stream st and return hipGetLastError().
hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st);
hipError_t launch_axpy(float* y, float a, const float* x, size_t n, hipStream_t st);   // y += a x
hipError_t launch_sum_fields(float* total, const float* const* src, int nsrc, size_t n3, hipStream_t st);
hipError_t launch_exchange_op(const Dims& d, const ExchangeDev& ex, const float* s, float* h, hipStream_t st);
hipError_t launch_anisotropy_op(const Dims& d, const AnisoDev& an, const float* s, float* h, hipStream_t st);
hipError_t launch_uniaxial_op(const Dims& d, const int* start, const float* ops, float scale, const float* s, float* h,
                              hipStream_t st);
hipError_t launch_thermal_op(const Dims& d, const ThermalDev& th, const LLGDev& lp, float dt, uint64_t draw,
                             const float* s, float* h, hipStream_t st);
hipError_t launch_llg_op(const Dims& d, const LLGDev& lp, float dt, const float* from, const float* dmdt,
                         const float* H, const float* Hth, float* to, hipStream_t st);
hipError_t launch_random_spins(const Dims& d, uint64_t seed, float* s, float* lo, hipStream_t st);
hipError_t launch_philox(const uint32_t* ctr, int count, uint32_t k0, uint32_t k1, uint32_t* out, hipStream_t st);
hipError_t launch_thermal_normals(const Dims& d, uint64_t seed, uint64_t draw, float* out, hipStream_t st);

// FP64 reductions over spins (hi words plus optional low words lo):
// moment -> out[6] sums; dot(h,s) -> out[0]; diff -> out[3]
hipError_t launch_reduce_moment(const Dims& d, const float* s, const float* lo, double* part, double* out, hipStream_t st);
hipError_t launch_reduce_dot(const Dims& d, const float* h, const float* s, const float* lo, double* part, double* out,
                             hipStream_t st);
hipError_t launch_reduce_diff(const Dims& d, const float* a, const float* alo, const float* b, const float* blo,
                              double* part, double* out, hipStream_t st);
int reduce_blocks(const Dims& d);

// dipole: op-API pieces
hipError_t launch_rows_fwd(const DipoleDev& dp, const float* s, hipStream_t st);
hipError_t launch_cols(const DipoleDev& dp, hipStream_t st);
hipError_t launch_rows_inv(const DipoleDev& dp, float* h, hipStream_t st);

// dipole tensor precompute in FP64
hipError_t launch_tensor_sum(int nx, int ny, int nz, const int gmax[4], const double ABC[9], double* T, hipStream_t st);
// setTruncation(math.huge): the square image sum |X|, |Y| <= 4999 the reference ends with
hipError_t launch_tensor_huge(int nx, int ny, int nz, const double ABC[9], double* T, hipStream_t st);
// T: [6][2nz-1][ny][nx] real space; U scratch [6][nz][ny][nx/2+1] complex; W out [6][nz][nx/2+1][ny] words
// twx/twy: FP64 tables (cos, -sin)(2 pi k / L)
hipError_t launch_tensor_spectrum(int nx, int ny, int nz, const double* T, double2* U, double* W,
                                  const double2* twx, const double2* twy, hipStream_t st);
// FP32 words Tw = mult * W; split = R > 0 lays them out [nxh][R][6][nz][ny/R] (ky = k1 + R k2)
hipError_t launch_tensor_words(const double* W, float* Tw, double mult, int nz, int nxh, int ny, int split,
                               hipStream_t st);
// planes of nx x ny real FP64 values T [planes][ny][nx] -> R2C spectra out [planes][nx/2+1][ny]
// as FP32 complex, times mult; U scratch [planes][ny][nx/2+1]
hipError_t launch_plane_spectrum(int planes, int nx, int ny, const double* T, double2* U, float2* out, double mult,
                                 const double2* twx, const double2* twy, hipStream_t st);

// fused Heun step
// Anisotropy.Uniaxial ops as a site-sorted CSR: start [n+1], ops [nops][5] = nx ny nz K1 K2
struct UniDev {
    const int* start;
    const float* ops;       // nullptr: no uniaxial anisotropy
    float scale;
};
struct StepArgs {
    Dims d;
    float dt;
    float* S;               // [4][n]  x y z ms (updated in place by the corrector)
    float* P;               // [3][n]  predictor spins
    float* C;               // [3][n]  half-step spins
    float* Slo;             // [3][n]  low words of S (compensated state) or nullptr
    float* Clo;             // [3][n]  low words of C
    ExchangeDev ex;
    AnisoDev an;
    ThermalDev th;
    LLGDev lp;
    int has_dip;
    DipoleDev dp;
    unsigned long long* draw;   // device counter of the thermal draw
    int ty;                 // rows per tile when there is no dipole
    const float* Hext;      // [3][n] field added at every site (the surrogate's dipole field) or nullptr
    const float* Href;      // [3][n] second such field (the surrogate's reference field) or nullptr
    UniDev uni;             // Anisotropy.Uniaxial
    int has_app;            // applied fields (AppliedField): uniform Bapp plus per-site Happ
    float3 Bapp;
    const float* Happ;      // [3][n] or nullptr
};
// fn (optional) receives the kernel launched, so a graph node of it can take new
// argument values without a new capture
hipError_t launch_step_pred(const StepArgs& a, hipStream_t st, const void** fn = nullptr);
hipError_t launch_step_corr(const StepArgs& a, hipStream_t st, const void** fn = nullptr);
// One stage of a Runge-Kutta step (RungeKutta.lua's istep after calcField): the total
// field at the stage spins X, then the Butcher row as chained renormalised Euler
// sub-steps from S over the stages j <= i, then (last stage) the thermal kick.
struct RKStage {
    const float* X;         // stage spins k_i [3][n] (S at stage 0)
    float* Hout;            // the stage's total field [3][n], kept for later rows, or nullptr
    float* out;             // the row's result [3][n]: k_(i+1), or the new S at the last stage
    const float* K[8];      // spins of k_j, j < i (K[0] = S)
    const float* H[8];      // total fields of k_j, j < i
    float f[8];             // the Butcher row: coefficients of j = 0 .. terms-1
    int terms;              // i + 1
    int last;               // 1: the last stage (thermal kick after the row)
};
hipError_t launch_rk_stage(const StepArgs& a, const RKStage& r, hipStream_t st);
// *draw += 1: the thermal draw of a Runge-Kutta step, read by its last stage
hipError_t launch_draw_advance(unsigned long long* draw, hipStream_t st);
// films of one shape (a0's) in one launch per pass: arguments / dipole descriptors of
// nrep films in device memory
hipError_t launch_step_pred_batch(const StepArgs& a0, const StepArgs* as, int nrep, hipStream_t st);
hipError_t launch_step_corr_batch(const StepArgs& a0, const StepArgs* as, int nrep, hipStream_t st);
hipError_t launch_cols_batch(const DipoleDev& d0, const DipoleDev* dps, int nrep, hipStream_t st);

// ------------------------------------------------ 3D convolution (LongRange3D)
// FFT grid (fx, fy, fz) equals the lattice when periodic and doubles it when
// open (zero padding).  The working spectrum W is [3][nz][fy][fxp], kx < fx/2+1
// used and rows padded to 128 bytes: planes z >= nz of the transform are zero and
// never stored.
// LDS lines that a pass fills by transposing copies are padded by this many
// float2, which puts neighbouring lines on different banks.
constexpr int CONV3_LPAD = 2;
struct Conv3Dev {
    int fx, fy, fz, fxh;
    int fxp;                // row pitch of W in float2: fxh rounded up to 16 (128-byte rows)
    int ty;                 // rows per row tile
    int wy, wz;             // kx block widths of the y and z passes
    int zt;                 // z length of the compile-time z pass (0: the runtime-length pass; fz + 1: open split pass)
    int yt;                 // y length of the compile-time y pass (0: the runtime-length pass)
    int rt;                 // x length of the compile-time row pass (0: the runtime-length pass)
    int fzs;                // kz planes of the tensor words (fz/2 + 1 when the tensor has z parity, else fz)
    int fys;                // ky rows of the tensor words (fy/2 + 1 when the tensor has y parity, else fy)
    FFTPlan px, py, pz;
    const float2* twx;
    const float2* twy;
    const float2* twz;
    float2* W;              // working spectrum [3][nz][fy][fxp], kx < fxh used
    const float* Tw;        // tensor words [6][fys][fxh][fzs] (real spectra), g/(fx fy fz) folded in
    float* hout;            // op-API target slot [3][n] (row kernel mode INV)
    const float* sin;       // op-API source spins (row kernel mode FWD)
};
// One replica: its step arguments and its convolution.
struct Rep3 {
    StepArgs a;
    Conv3Dev cv;
    int has_lr;
};
// row kernel modes: forward transform of cv.sin, inverse transform into cv.hout,
// and the fused predictor / corrector of the Heun step (the predictor of the
// item with z = y0 = 0 advances the replica's thermal draw counter)
enum Row3Mode { R3_FWD = 0, R3_INV = 1, R3_PRED = 2, R3_CORR = 3 };
// work items: int4 (replica, p, q, unused); rows: p = z, q = y0; y pass: p = c*nz+z,
// q = kx0; z pass: p = ky, q = kx0
hipError_t launch3_rows(int mode, bool comp, const Rep3* reps, const int4* items, int nitems, size_t lds, int rt,
                        hipStream_t st);
int rows_template_height(int fx);   // tile height of the compile-time row pass of length fx, 0 if none
// yt, wy: every replica's compile-time y length and kx width (Conv3Dev::yt, wy), or
// yt = 0 for the runtime-length pass
hipError_t launch3_ypass(int sign, const Rep3* reps, const int4* items, int nitems, size_t lds, int yt, int wy,
                         hipStream_t st);
// default kx width of the compile-time y pass for a y length (0 when it has no instance)
int ypass_template_width(int fy);
// kx widths compiled for the y pass of length fy (default first); returns how many
int ypass_widths(int fy, int* w, int cap);
int zconv_open_width(int fz);       // default kx per workgroup of the open split z pass of length fz, 0 if none
// zt, wz: every replica's compile-time z pass form and kx width (Conv3Dev::zt, wz),
// or zt = 0 for the runtime-length pass
hipError_t launch3_zconv(const Rep3* reps, const int4* items, int nitems, size_t lds, int zt, int wz, hipStream_t st);
// default kx width of the full compile-time z pass for a z length (0 when it has no instance)
int zconv_template_width(int fz);
// kx widths compiled for the z pass form zt (fz, or fz + 1 for the open split pass),
// default first; returns how many
int zconv_widths(int zt, int* w, int cap);
// tensor: periodic lattice sum (R > 0) or open kernel (R == 0) on the FFT grid, FP64
hipError_t launch3_tensor(int nx, int ny, int nz, int fx, int fy, int fz, int R, double* T, hipStream_t st);
// real spectrum words of T: U1/U2 complex FP64 scratch [6][fz][fy][fxh]; tw* FP64 tables
hipError_t launch3_spectrum(int fx, int fy, int fz, int fys, int fzs, const double* T, double2* U1, double2* U2, float* Tw,
                            double mult, const double2* twx, const double2* twy, const double2* twz, hipStream_t st);
hipError_t configure_lds3(size_t max_bytes);
hipError_t configure_fft64(size_t bytes);
const char* kernel_name(int id);
// allow row/column kernels the device's full LDS per workgroup
hipError_t configure_lds(size_t max_bytes);

// ------------------------------------------------ dipolar-field surrogate network
// Weight-gradient tiles: 32 x 8 sites of one layer (nx must be a multiple of 32).
struct CnnDims {
    int nx, ny, nz, n;
    int tiles_x, tiles_y, ntiles;
    int open;               // 0: periodic in x and y; 1: zero outside the lattice in x and y too
};
// Row-walk layer items: 128 sites along x, ty output rows, one layer; or (film) 64
// sites along x, ty output rows, every layer of a film of 1, 2, 4 or 8 layers.
struct CnnWalk {
    int ty, rtiles_x, rtiles_y, rtiles;
    int film;
};
// layer shapes of the row walk: 16 -> 16, 3 -> 16 and 16 -> 3 channels
enum { CV_F16 = 0, CV_IN3 = 1, CV_OUT3 = 2 };
// BF16 activations are channels-last: 16 channels (32 bytes) per site, or 4
// (8 bytes, channel 3 zero) for the network input and the output gradient.
// zeros: 16 zero bytes of device memory (the 3-channel input's layers outside the lattice);
// CV_OUT3 writes FP32 [3][n] scale * Y into out_f, the others BF16 [n][16] into out_bf
hipError_t launch_cnn_rows(int variant, const CnnDims& d, const CnnWalk& wk, int grid, const uint16_t* in,
                           const uint16_t* zeros, const uint16_t* apack, uint16_t* out_bf, float* out_f,
                           float scale, hipStream_t st);
hipError_t launch_cnn_wgrad(const CnnDims& d, int grid, const uint16_t* X, int xs, const uint16_t* G, int gs,
                            float* partial, float* grad, hipStream_t st);
hipError_t launch_cnn_pack(const float* W, int transposed, int variant, uint16_t* apack, hipStream_t st);
hipError_t launch_cnn_input(const Dims& d, const float* m, const float* mref, float s_in, uint16_t* X0, hipStream_t st);
hipError_t launch_cnn_loss(const Dims& d, const float* Y3, const float* T, float s_out, float gscale, uint16_t* G3,
                           double* part, hipStream_t st);
hipError_t launch_cnn_adam(float* W, const float* g, float* m, float* v, int count, float lr, float b1, float b2,
                           float c1, float c2, float eps, hipStream_t st);
hipError_t launch_cnn_sub(size_t n3, const float* a, const float* b, float* D, hipStream_t st);
hipError_t launch_cnn_add(size_t n3, const float* a, const float* b, float* S, hipStream_t st);
hipError_t launch_cnn_maxdiff(size_t n3, const float* a, const float* b, float2* part, int blocks, hipStream_t st);
int cnn_rows_ty_max();
bool cnn_film_supported(int nz);
int cnn_rows_occupancy(int variant, int film_nz);
// Row-tile shapes with compile-time instances: whether (nx, ty, tpb) is one, and
// the default (ty, tpb) for an nx x ny layer (ty = 0 when none fits).
bool row_shape_compiled(int nx, int ty, int tpb);
// every compiled (ty, tpb) for an nx x ny layer, default first; returns the count
int row_shapes_for(int nx, int ny, int* ty, int* tpb, int cap);
void row_shape_default(int nx, int ny, int* ty, int* tpb);
// split column kernel: compiled sub-column lengths and its dynamic LDS bytes
bool split_length_compiled(int ly, int R);
size_t split_cols_lds(int ly, int nz);

}  // namespace msd
