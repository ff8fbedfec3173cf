// 3D FFT convolution (LongRange3D::apply, cpu/longrange3d/spinoperationlongrange3d.cpp:369-432)
// on gfx950, for one system or a batch of independent replicas of uneven sizes
// advanced by the same launches.
//
// Every launch walks a table of work items (replica, p, q); a replica's sizes,
// plans, arrays and physics come from its Rep3 descriptor, so replicas of
// different lattice sizes share one launch.  Per field evaluation:
//   rows  : x transforms (real <-> half complex, two real rows per complex line),
//           fused with the Heun site updates of the step driver;
//   ypass : y transforms of kx blocks (coalesced runs of w*8 bytes);
//   zconv : forward z transform of the three components, the per-k 3x3 tensor
//           product (real spectra: the tensor is even), inverse z transform.
// With open boundaries the FFT grid is twice the lattice: input rows/planes
// beyond the lattice are zero and output beyond it is never stored.
#include "msd_device.h"

#include <algorithm>

namespace msd {

static constexpr int R3_TPB = 512;
static constexpr int Y3_TPB = 256;
static constexpr int Z3_TPB = 512;

__device__ __forceinline__ float& rowval3(float2* buf, int q, int x, int fx)
{
    return reinterpret_cast<float*>(buf + (size_t)(q >> 1) * fx + x)[q & 1];
}

__device__ __forceinline__ size_t widx(const Conv3Dev& cv, int nz, int c, int z, int y, int k)
{
    return (((size_t)c * nz + z) * cv.fy + y) * cv.fxp + k;
}

// The six tensor words (xx xy xz yy yz zz) of spectrum point (kx, ky, kz), stored
// [6][fys][fxh][fzs] with kz fastest.  With fzs < fz only kz <= fz/2 is stored and
// kz beyond reads fz - kz: the tensor is even in z except xz and yz, which are odd.
// Likewise with fys < fy only ky <= fy/2 is stored, ky beyond reads fy - ky, and xy
// and yz, odd in y, change sign.
struct TensorWords {
    float xx, xy, xz, yy, yz, zz;
};
__device__ __forceinline__ TensorWords tensor_words(const Conv3Dev& cv, int ky, int kx, int kz)
{
    const bool mir = kz >= cv.fzs, ymir = ky >= cv.fys;
    const size_t plane = (size_t)cv.fys * cv.fxh * cv.fzs;
    const size_t t = ((size_t)(ymir ? cv.fy - ky : ky) * cv.fxh + kx) * cv.fzs + (mir ? cv.fz - kz : kz);
    TensorWords w{gload(cv.Tw, t), gload(cv.Tw, plane + t), gload(cv.Tw, 2 * plane + t), gload(cv.Tw, 3 * plane + t),
                  gload(cv.Tw, 4 * plane + t), gload(cv.Tw, 5 * plane + t)};
    if (mir) { w.xz = -w.xz; w.yz = -w.yz; }
    if (ymir) { w.xy = -w.xy; w.yz = -w.yz; }
    return w;
}

// Y = T X at one spectrum point for the three components held at px, py, pz
__device__ __forceinline__ void tensor_apply(const TensorWords& w, float2& px, float2& py, float2& pz)
{
    const float2 mx = px, my = py, mz = pz;
    px = make_float2(w.xx * mx.x + w.xy * my.x + w.xz * mz.x, w.xx * mx.y + w.xy * my.y + w.xz * mz.y);
    py = make_float2(w.xy * mx.x + w.yy * my.x + w.yz * mz.x, w.xy * mx.y + w.yy * my.y + w.yz * mz.y);
    pz = make_float2(w.xz * mx.x + w.yz * my.x + w.zz * mz.x, w.xz * mx.y + w.yz * my.y + w.zz * mz.y);
}

// ------------------------------------------------------------------- rows
// A row tile is ty rows y0.. of plane z.  Its 3*ty real rows (component-major,
// q = c*ty + r) are packed two to a complex line of length fx: line l holds rows
// 2l (real part) and 2l+1 (imaginary part).  Tiles read and write the spectrum
// W[c][z][y][kx] row by row, kx fastest across the threads (contiguous runs).

// Copy the exchange class table into LDS (the fused form has one only when it fits,
// msd_exchange::fused_dev).
template <int MODE>
__device__ __forceinline__ ExchangeDev rows_exchange(const ExchangeDev& ex, int* s_start, ExEntry* s_ent)
{
    ExchangeDev e = ex;
    if (MODE == R3_PRED || MODE == R3_CORR) {
        if (ex.mode == 1) {
            for (int i = threadIdx.x; i <= ex.ncls; i += blockDim.x) s_start[i] = ex.cls_start[i];
            for (int i = threadIdx.x; i < ex.nent; i += blockDim.x) s_ent[i] = ex.ent[i];
            e.cls_start = s_start;
            e.ent = s_ent;
        }
    }
    return e;
}

// The two spectrum values of packed line l at kx = kk (0 beyond the tile's rows).
__device__ __forceinline__ float4 rows_load_pair(const Conv3Dev& cv, int nz, int z, int y0, int ty, int l, int kk)
{
    const int q0 = 2 * l, q1 = q0 + 1;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (q0 < 3 * ty) {
        const int c0 = q0 / ty, r0 = q0 - c0 * ty;
        const float2 a = gload(cv.W, widx(cv, nz, c0, z, y0 + r0, kk));
        v.x = a.x; v.y = a.y;
        if (q1 < 3 * ty) {
            const int c1 = q1 / ty, r1 = q1 - c1 * ty;
            const float2 b = gload(cv.W, widx(cv, nz, c1, z, y0 + r1, kk));
            v.z = b.x; v.w = b.y;
        }
    }
    return v;
}

// Packed line value at k from the pair read at kk = min(k, fx - k): the half
// spectra of two real rows A, B combine as A + iB, with k > fx/2 the conjugates.
__device__ __forceinline__ float2 rows_pack(float4 v, bool mirror)
{
    if (mirror) { v.y = -v.y; v.w = -v.w; }
    return make_float2(v.x - v.w, v.y + v.z);
}

// Split transformed line l at k (Zk) and fx - k (Zm) into the half spectra of
// its two rows and store them.
__device__ __forceinline__ void rows_store_pair(const Conv3Dev& cv, int nz, int z, int y0, int ty, int l, int k, float2 Zk,
                                                float2 Zm)
{
    const int q0 = 2 * l, q1 = q0 + 1;
    const int c0 = q0 / ty, r0 = q0 - c0 * ty;
    gstore(cv.W, widx(cv, nz, c0, z, y0 + r0, k), make_float2(0.5f * (Zk.x + Zm.x), 0.5f * (Zk.y - Zm.y)));
    if (q1 < 3 * ty) {
        const int c1 = q1 / ty, r1 = q1 - c1 * ty;
        gstore(cv.W, widx(cv, nz, c1, z, y0 + r1, k), make_float2(0.5f * (Zk.y + Zm.y), 0.5f * (Zm.x - Zk.x)));
    }
}

// Site (x, y0 + r, z) of the tile: read its dipolar field from the transformed
// rows R (with a convolution), then per MODE store the field (INV), or take the
// spins (FWD) or the Heun predictor / corrector update, and put them in R for the
// forward transform.  The padded tail x in [nx, fx) of every row is zero.
template <int MODE, bool COMP>
__device__ __forceinline__ void rows_site(const Rep3& rp, const ExchangeDev& exl, float2* R, int fx, int ty, int r, int x,
                                          int y0, int z, uint64_t draw)
{
    const StepArgs& a = rp.a;
    const Conv3Dev& cv = rp.cv;
    const bool lr = rp.has_lr != 0;
    const int nx = a.d.nx, n = a.d.n;
    float3 v = make_float3(0.f, 0.f, 0.f);
    if (x < nx) {
        const int i = x + nx * (y0 + r + a.d.ny * z);
        if (MODE == R3_FWD) {
            v = make_float3(gload(cv.sin, i), gload(cv.sin, (size_t)n + i), gload(cv.sin, 2 * (size_t)n + i));
        } else {
            float3 Hd = make_float3(0.f, 0.f, 0.f);
            if (lr) Hd = make_float3(rowval3(R, r, x, fx), rowval3(R, ty + r, x, fx), rowval3(R, 2 * ty + r, x, fx));
            if (MODE == R3_INV) {
                gstore(cv.hout, i, Hd.x); gstore(cv.hout, (size_t)n + i, Hd.y); gstore(cv.hout, 2 * (size_t)n + i, Hd.z);
            } else if (MODE == R3_PRED) {
                v = site_pred<COMP>(a, exl, i, Hd);
            } else {
                v = site_corr<COMP>(a, exl, i, Hd, draw);
            }
        }
    }
    if (lr && MODE != R3_INV) {
        rowval3(R, r, x, fx) = v.x;
        rowval3(R, ty + r, x, fx) = v.y;
        rowval3(R, 2 * ty + r, x, fx) = v.z;
    }
}

template <int MODE, bool COMP>
__global__ void __launch_bounds__(R3_TPB) k3_rows(const Rep3* __restrict__ reps, const int4* __restrict__ items)
{
    extern __shared__ float2 lds[];
    __shared__ int s_start[EX_LDS_CLS];
    __shared__ ExEntry s_ent[EX_LDS_ENT];
    const int4 it = items[blockIdx.x];
    const Rep3& rp = reps[it.x];
    const StepArgs& a = rp.a;
    const Conv3Dev& cv = rp.cv;
    const int ny = a.d.ny, nz = a.d.nz;
    const int fx = cv.fx, fxh = cv.fxh, half = fx >> 1;
    const int z = it.y, y0 = it.z;
    const int ty = min(cv.ty, ny - y0);
    const int lines = (3 * ty + 1) >> 1;
    const bool lr = rp.has_lr != 0;
    float2* A = lds;
    float2* B = lds + (size_t)lines * fx;
    float2* R = A;

    const ExchangeDev exl = rows_exchange<MODE>(a.ex, s_start, s_ent);
    // inverse x transform of the field spectrum rows
    if (lr && MODE != R3_FWD) {
        for (int idx = threadIdx.x; idx < lines * fx; idx += R3_TPB) {
            const int l = idx / fx, k = idx - l * fx;
            const bool mirror = k > half;
            A[idx] = rows_pack(rows_load_pair(cv, nz, z, y0, ty, l, mirror ? fx - k : k), mirror);
        }
        __syncthreads();
        R = lds_fft<1>(A, B, lines, cv.px, cv.twx);
    } else {
        __syncthreads();
    }
    if (MODE == R3_PRED && a.th.on && z == 0 && y0 == 0 && threadIdx.x == 0) *a.draw += 1ull;
    const uint64_t draw = (MODE == R3_CORR && a.th.on) ? (uint64_t)*a.draw : 0ull;

    for (int li = threadIdx.x; li < ty * fx; li += R3_TPB) {
        const int r = li / fx, x = li - r * fx;
        rows_site<MODE, COMP>(rp, exl, R, fx, ty, r, x, y0, z, draw);
    }
    if (!lr || MODE == R3_INV) return;
    // the unpaired imaginary half of the last line (odd 3*ty) must be zero
    if ((3 * ty) & 1)
        for (int x = threadIdx.x; x < fx; x += R3_TPB)
            reinterpret_cast<float*>(R + (size_t)(lines - 1) * fx + x)[1] = 0.0f;
    __syncthreads();
    const float2* F = lds_fft<-1>(R, R == A ? B : A, lines, cv.px, cv.twx);
    for (int idx = threadIdx.x; idx < lines * fxh; idx += R3_TPB) {
        const int l = idx / fxh, k = idx - l * fxh;
        rows_store_pair(cv, nz, z, y0, ty, l, k, F[(size_t)l * fx + k], F[(size_t)l * fx + (k ? fx - k : 0)]);
    }
}

// Compile-time x length FX (32..512 powers of two, and 96) and tile height TY: both x
// transforms run in place (fft_ip) with the twiddles staged in LDS, the spectrum
// rows are loaded in batches, and the site loop has a fixed trip count.  A tile
// at the lattice's y edge may hold fewer rows; its unused lines are zero.
template <int FX>
__host__ __device__ constexpr int rows_template_ty()
{
    return FX >= 512 ? 4 : FX >= 256 ? 8 : FX >= 96 ? 16 : 32;
}
// threads per workgroup (measured: 256 at x length 512, +1.5% per 256^3 open step;
// 256 at the ensemble's lengths 32..128 and 96 cost 10%, so those keep 512)
template <int FX>
__host__ __device__ constexpr int rows_template_threads()
{
    return FX >= 512 ? 256 : R3_TPB;
}

template <int MODE, bool COMP, int FX>
__global__ void __launch_bounds__(rows_template_threads<FX>()) k3_rows_t(const Rep3* __restrict__ reps, const int4* __restrict__ items)
{
    constexpr int TY = rows_template_ty<FX>(), LINES = (3 * TY + 1) / 2, FXH = FX / 2 + 1, HALF = FX / 2;
    constexpr int NT = rows_template_threads<FX>();
    extern __shared__ float2 lds[];                     // [LINES][FX] lines, then [FX] twiddles
    __shared__ int s_start[EX_LDS_CLS];
    __shared__ ExEntry s_ent[EX_LDS_ENT];
    const int4 it = items[blockIdx.x];
    const Rep3& rp = reps[it.x];
    const StepArgs& a = rp.a;
    const Conv3Dev& cv = rp.cv;
    const int nz = a.d.nz;
    const int z = it.y, y0 = it.z;
    const int ty = min(TY, a.d.ny - y0);
    const int lines = (3 * ty + 1) >> 1;
    const bool lr = rp.has_lr != 0;
    float2* A = lds;
    float2* s_tw = lds + LINES * FX;

    const ExchangeDev exl = rows_exchange<MODE>(a.ex, s_start, s_ent);
    if (lr)
        for (int i = threadIdx.x; i < FX; i += NT) s_tw[i] = gload(cv.twx, i);
    if (lr && MODE != R3_FWD) {
        batched_copy<LINES * FX, NT, 8>(
            [&](int idx) {
                const int l = idx / FX, k = idx % FX;
                return l < lines ? rows_load_pair(cv, nz, z, y0, ty, l, k > HALF ? FX - k : k)
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
            },
            [&](int idx, float4 v) { A[idx] = rows_pack(v, idx % FX > HALF); });
        __syncthreads();
        fft_ip<FX, LINES, NT, 1>(A, s_tw);
    } else {
        __syncthreads();
    }
    if (MODE == R3_PRED && a.th.on && z == 0 && y0 == 0 && threadIdx.x == 0) *a.draw += 1ull;
    const uint64_t draw = (MODE == R3_CORR && a.th.on) ? (uint64_t)*a.draw : 0ull;

    constexpr int SITES = TY * FX;
#pragma unroll 2
    for (int li = threadIdx.x; li < SITES; li += NT) {
        const int r = li / FX, x = li % FX;
        if (r < ty) rows_site<MODE, COMP>(rp, exl, A, FX, ty, r, x, y0, z, draw);
    }
    if (!lr || MODE == R3_INV) return;
    // the unpaired imaginary half of the last line (odd 3*ty) must be zero
    if ((3 * ty) & 1)
        for (int x = threadIdx.x; x < FX; x += NT)
            reinterpret_cast<float*>(A + (size_t)(lines - 1) * FX + x)[1] = 0.0f;
    __syncthreads();
    fft_ip<FX, LINES, NT, -1>(A, s_tw);
    for (int idx = threadIdx.x; idx < lines * FXH; idx += NT) {
        const int l = idx / FXH, k = idx - l * FXH;
        rows_store_pair(cv, nz, z, y0, ty, l, k, A[l * FX + k], A[l * FX + (k ? FX - k : 0)]);
    }
}

// ----------------------------------------------------------------- y pass
// lines along y of kx block [kx0, kx0+w) in plane (c, z); forward reads rows
// y < ny (zero beyond) and writes all fy, inverse reads fy and writes y < ny.
template <int SIGN>
__global__ void __launch_bounds__(Y3_TPB) k3_ypass(const Rep3* __restrict__ reps, const int4* __restrict__ items)
{
    extern __shared__ float2 lds[];
    const int4 it = items[blockIdx.x];
    const Rep3& rp = reps[it.x];
    const Conv3Dev& cv = rp.cv;
    const int ny = rp.a.d.ny;
    const int fy = cv.fy, fxh = cv.fxh, fxp = cv.fxp;
    const int kx0 = it.z;
    const int w = min(cv.wy, fxh - kx0);
    float2* A = lds;
    float2* B = lds + (size_t)w * fy;
    float2* base = cv.W + (size_t)it.y * fy * fxp + kx0;
    const int yin = SIGN < 0 ? ny : fy;
    for (int idx = threadIdx.x; idx < w * fy; idx += Y3_TPB) {
        const int y = idx / w, j = idx - y * w;
        A[(size_t)j * fy + y] = y < yin ? gload(base, (size_t)y * fxp + j) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2* R = lds_fft<SIGN>(A, B, w, cv.py, cv.twy);
    const int yout = SIGN < 0 ? fy : ny;
    for (int idx = threadIdx.x; idx < w * yout; idx += Y3_TPB) {
        const int y = idx / w, j = idx - y * w;
        gstore(base, (size_t)y * fxp + j, R[(size_t)j * fy + y]);
    }
}

// Compile-time y length FY (32..512 powers of two, and 96), W kx columns per
// workgroup (8 by default, 4 and 16 for the autotuner), transformed in place
// (fft_ip) in half the runtime pass's LDS.  At W = 8 every y row is read and written
// as 64-byte runs; with W's 128-byte rows and XCD-local items the two blocks of a
// line meet in one L2 (measured faster than W = 16 at 256^3 open).
template <int SIGN, int FY, int W>
__global__ void __launch_bounds__(Y3_TPB) k3_ypass_t(const Rep3* __restrict__ reps, const int4* __restrict__ items)
{
    extern __shared__ float2 lds[];
    const int4 it = items[blockIdx.x];
    const Rep3& rp = reps[it.x];
    const Conv3Dev& cv = rp.cv;
    const int ny = rp.a.d.ny;
    const int fxp = cv.fxp;
    const int kx0 = it.z;
    const int w = min(W, cv.fxh - kx0);
    constexpr int LS = FY + CONV3_LPAD;
    float2* A = lds;                                     // [j][LS]
    float2* base = cv.W + (size_t)it.y * FY * fxp + kx0;
    const int yin = SIGN < 0 ? ny : FY;
    batched_copy<W * FY, Y3_TPB, 8>(
        [&](int idx) {
            const int y = idx / W, j = idx - y * W;
            return (y < yin && j < w) ? gload(base, (size_t)y * fxp + j) : make_float2(0.f, 0.f);
        },
        [&](int idx, float2 v) {
            const int y = idx / W, j = idx - y * W;
            A[j * LS + y] = v;
        });
    __syncthreads();
    fft_ip<FY, W, Y3_TPB, SIGN, LS>(A, gview(cv.twy));
    const int yout = SIGN < 0 ? FY : ny;
    for (int idx = threadIdx.x; idx < W * yout; idx += Y3_TPB) {
        const int y = idx / W, j = idx - y * W;
        if (j < w) gstore(base, (size_t)y * fxp + j, A[j * LS + y]);
    }
}

// kx widths of the compile-time y pass, the default (measured fastest at 256^3
// open) first; the step autotuner times the others
static constexpr int YPASS_WIDTHS[] = {8, 16, 4};

template <int SIGN, int FY>
static const void* ypass_w(int w)
{
    switch (w) {
    case 4: return (const void*)k3_ypass_t<SIGN, FY, 4>;
    case 8: return (const void*)k3_ypass_t<SIGN, FY, 8>;
    case 16: return (const void*)k3_ypass_t<SIGN, FY, 16>;
    default: return nullptr;
    }
}

template <int SIGN>
static const void* ypass_template(int fy, int w)
{
    switch (fy) {
    case 32: return ypass_w<SIGN, 32>(w);
    case 64: return ypass_w<SIGN, 64>(w);
    case 96: return ypass_w<SIGN, 96>(w);
    case 128: return ypass_w<SIGN, 128>(w);
    case 256: return ypass_w<SIGN, 256>(w);
    case 512: return ypass_w<SIGN, 512>(w);
    default: return nullptr;
    }
}

int ypass_template_width(int fy) { return ypass_template<1>(fy, YPASS_WIDTHS[0]) ? YPASS_WIDTHS[0] : 0; }

int ypass_widths(int fy, int* w, int cap)
{
    int n = 0;
    for (int c : YPASS_WIDTHS)
        if (ypass_template<1>(fy, c) && n < cap) w[n++] = c;
    return n;
}

// ----------------------------------------------------------------- z pass
__global__ void __launch_bounds__(Z3_TPB) k3_zconv(const Rep3* __restrict__ reps, const int4* __restrict__ items)
{
    extern __shared__ float2 lds[];
    const int4 it = items[blockIdx.x];
    const Rep3& rp = reps[it.x];
    const Conv3Dev& cv = rp.cv;
    const int nz = rp.a.d.nz;
    const int fz = cv.fz, fxh = cv.fxh;
    const int ky = it.y, kx0 = it.z;
    const int w = min(cv.wz, fxh - kx0);
    const int lines = 3 * w;
    float2* A = lds;
    float2* B = lds + (size_t)lines * fz;
    for (int idx = threadIdx.x; idx < lines * fz; idx += Z3_TPB) {
        const int zz = idx / lines, l = idx - zz * lines;
        const int c = l / w, j = l - c * w;
        A[(size_t)l * fz + zz] = zz < nz ? gload(cv.W, widx(cv, nz, c, zz, ky, kx0 + j)) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    float2* R = lds_fft<-1>(A, B, lines, cv.pz, cv.twz);
    // kz fastest: the tensor words and the LDS lines are both read in runs
    for (int idx = threadIdx.x; idx < w * fz; idx += Z3_TPB) {
        const int j = idx / fz, kz = idx - j * fz;
        tensor_apply(tensor_words(cv, ky, kx0 + j, kz), R[(size_t)(0 * w + j) * fz + kz], R[(size_t)(1 * w + j) * fz + kz],
                     R[(size_t)(2 * w + j) * fz + kz]);
    }
    __syncthreads();
    const float2* O = lds_fft<1>(R, R == A ? B : A, lines, cv.pz, cv.twz);
    for (int idx = threadIdx.x; idx < lines * nz; idx += Z3_TPB) {
        const int zz = idx / lines, l = idx - zz * lines;
        const int c = l / w, j = l - c * w;
        gstore(cv.W, widx(cv, nz, c, zz, ky, kx0 + j), O[(size_t)l * fz + zz]);
    }
}

// Compile-time z length FZ (16..512 powers of two, and 24), W kx columns per workgroup:
// the three components' z lines (3 W x FZ) transformed in place in one LDS buffer
// (fft_ip), so W is 4x the runtime pass's at the same LDS and each z plane is read
// and written as runs of W consecutive kx (64-128 bytes instead of 16).
template <int FZ, int W>
__global__ void __launch_bounds__(Z3_TPB) k3_zconv_t(const Rep3* __restrict__ reps, const int4* __restrict__ items)
{
    extern __shared__ float2 lds[];
    constexpr int lines = 3 * W, LS = FZ + CONV3_LPAD;
    const int4 it = items[blockIdx.x];
    const Rep3& rp = reps[it.x];
    const Conv3Dev& cv = rp.cv;
    const int nz = rp.a.d.nz;
    const int ky = it.y, kx0 = it.z;
    const int w = min(W, cv.fxh - kx0);
    float2* A = lds;                                     // [c W + j][LS]
    // element idx = zz lines + l: consecutive threads walk the kx of one plane
    batched_copy<lines * FZ, Z3_TPB, 8>(
        [&](int idx) {
            const int zz = idx / lines, l = idx - zz * lines;
            const int c = l / W, j = l - c * W;
            return (zz < nz && j < w) ? gload(cv.W, widx(cv, nz, c, zz, ky, kx0 + j)) : make_float2(0.f, 0.f);
        },
        [&](int idx, float2 v) {
            const int zz = idx / lines, l = idx - zz * lines;
            A[l * LS + zz] = v;
        });
    __syncthreads();
    fft_ip<FZ, lines, Z3_TPB, -1, LS>(A, gview(cv.twz));
    // kz fastest: the tensor words and the LDS lines are both read in runs
    for (int idx = threadIdx.x; idx < W * FZ; idx += Z3_TPB) {
        const int j = idx / FZ, kz = idx % FZ;
        if (j < w)
            tensor_apply(tensor_words(cv, ky, kx0 + j, kz), A[(0 * W + j) * LS + kz], A[(1 * W + j) * LS + kz],
                         A[(2 * W + j) * LS + kz]);
    }
    __syncthreads();
    fft_ip<FZ, lines, Z3_TPB, 1, LS>(A, gview(cv.twz));
    for (int idx = threadIdx.x; idx < lines * nz; idx += Z3_TPB) {
        const int zz = idx / lines, l = idx - zz * lines;
        const int c = l / W, j = l - c * W;
        if (j < w) gstore(cv.W, widx(cv, nz, c, zz, ky, kx0 + j), A[l * LS + zz]);
    }
}

// Open boundaries, fz = 2 NZ: input planes z >= NZ are zero and only output planes
// z < NZ are kept.  With w = exp(-2 pi i / fz) the length-fz transform splits by
// the parity of kz into two length-NZ transforms,
//   X(2m) = DFT_NZ(x)(m),   X(2m+1) = DFT_NZ(w^z x)(m),
//   y(z) = IDFT_NZ(T X_even)(z) + w^-z IDFT_NZ(T X_odd)(z),   z < NZ,
// so the LDS lines are NZ long, half those of the full pass, and a CU holds twice
// the workgroups.  The halves run one after the other; a thread keeps its share of
// the input, the even half's result and the odd half's tensor words in registers.
template <int NZ, int W, int NT>
__global__ void __launch_bounds__(NT) k3_zconv_o(const Rep3* __restrict__ reps, const int4* __restrict__ items)
{
    extern __shared__ float2 lds[];
    constexpr int lines = 3 * W, LS = NZ + CONV3_LPAD;
    constexpr int TOT = lines * NZ, PER = (TOT + NT - 1) / NT;
    constexpr int MUL = W * NZ, MPER = (MUL + NT - 1) / NT;
    const int4 it = items[blockIdx.x];
    const Rep3& rp = reps[it.x];
    const Conv3Dev& cv = rp.cv;
    const int ky = it.y, kx0 = it.z;
    const int w = min(W, cv.fxh - kx0);
    float2* A = lds;                                     // [c W + j][LS]
    float2* twh = lds + lines * LS;                      // exp(-2 pi i k / NZ), k < NZ
    float2* wz = twh + NZ;                               // w^z, z < NZ
    for (int i = threadIdx.x; i < NZ; i += NT) {
        twh[i] = gload(cv.twz, 2 * i);
        wz[i] = gload(cv.twz, i);
    }
    // element idx = z lines + l: consecutive threads walk the kx of one plane
    float2 x[PER], y[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int idx = (int)threadIdx.x + p * NT;
        const int zz = idx / lines, l = idx - zz * lines;
        const int c = l / W, j = l - c * W;
        x[p] = ((TOT % NT == 0 || idx < TOT) && j < w) ? gload(cv.W, widx(cv, NZ, c, zz, ky, kx0 + j)) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int idx = (int)threadIdx.x + p * NT;
        const int zz = idx / lines, l = idx - zz * lines;
        if (TOT % NT == 0 || idx < TOT) A[l * LS + zz] = x[p];
    }
    __syncthreads();
    fft_ip<NZ, lines, NT, -1, LS>(A, twh);
    // even kz; the odd kz words are read in the same pass and kept
    TensorWords odd[MPER];
#pragma unroll
    for (int p = 0; p < MPER; ++p) {
        const int idx = (int)threadIdx.x + p * NT;
        const int j = idx / NZ, m = idx & (NZ - 1);
        if ((MUL % NT == 0 || idx < MUL) && j < w) {
            odd[p] = tensor_words(cv, ky, kx0 + j, 2 * m + 1);
            tensor_apply(tensor_words(cv, ky, kx0 + j, 2 * m), A[(0 * W + j) * LS + m], A[(1 * W + j) * LS + m],
                         A[(2 * W + j) * LS + m]);
        }
    }
    __syncthreads();
    fft_ip<NZ, lines, NT, 1, LS>(A, twh);
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int idx = (int)threadIdx.x + p * NT;
        const int zz = idx / lines, l = idx - zz * lines;
        if (TOT % NT == 0 || idx < TOT) y[p] = A[l * LS + zz];
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int idx = (int)threadIdx.x + p * NT;
        const int zz = idx / lines, l = idx - zz * lines;
        if (TOT % NT == 0 || idx < TOT) A[l * LS + zz] = cmul(x[p], wz[zz]);
    }
    __syncthreads();
    fft_ip<NZ, lines, NT, -1, LS>(A, twh);
#pragma unroll
    for (int p = 0; p < MPER; ++p) {
        const int idx = (int)threadIdx.x + p * NT;
        const int j = idx / NZ, m = idx & (NZ - 1);
        if ((MUL % NT == 0 || idx < MUL) && j < w)
            tensor_apply(odd[p], A[(0 * W + j) * LS + m], A[(1 * W + j) * LS + m], A[(2 * W + j) * LS + m]);
    }
    __syncthreads();
    fft_ip<NZ, lines, NT, 1, LS>(A, twh);
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int idx = (int)threadIdx.x + p * NT;
        const int zz = idx / lines, l = idx - zz * lines;
        const int c = l / W, j = l - c * W;
        if ((TOT % NT == 0 || idx < TOT) && j < w) {
            const float2 o = cmulc(A[l * LS + zz], wz[zz]);
            gstore(cv.W, widx(cv, NZ, c, zz, ky, kx0 + j), make_float2(y[p].x + o.x, y[p].y + o.y));
        }
    }
}

// kx columns and threads per workgroup of the open split pass of half length nz
// (measured at nz = 128 and 256: narrow blocks, two to four workgroups per CU whose
// memory phases overlap each other's transforms; a thread holds 6 of each line set).
// The default width comes first; the step autotuner times the others.
static constexpr int zopen_width(int nz) { return nz >= 128 ? 4 : 16; }
// 8 columns at nz >= 128 take twice the threads (the same 6 elements of each line
// set per thread): runs of 64 bytes per z plane instead of 32 at the same registers
static constexpr int zopen_threads(int nz, int w) { return (nz == 128 ? 256 : Z3_TPB) * (nz >= 128 && w >= 8 ? 2 : 1); }
template <int NZ>
static const void* zopen_kernel(int w)
{
    if constexpr (NZ >= 128) {
        switch (w) {
        case 4: return (const void*)k3_zconv_o<NZ, 4, zopen_threads(NZ, 4)>;
        case 2: return (const void*)k3_zconv_o<NZ, 2, zopen_threads(NZ, 2)>;
        case 8: return (const void*)k3_zconv_o<NZ, 8, zopen_threads(NZ, 8)>;
        default: return nullptr;
        }
    }
    else {
        switch (w) {
        case 16: return (const void*)k3_zconv_o<NZ, 16, zopen_threads(NZ, 16)>;
        case 8: return (const void*)k3_zconv_o<NZ, 8, zopen_threads(NZ, 8)>;
        default: return nullptr;
        }
    }
}

// full compile-time pass of length FZ: 16 kx columns (8 at 512, whose 16 exceed the
// LDS) by default, and the narrower widths
template <int FZ>
static const void* zfull_kernel(int w)
{
    switch (w) {
    case 16:
        if constexpr (FZ < 512) return (const void*)k3_zconv_t<FZ, 16>;
        else return nullptr;
    case 8: return (const void*)k3_zconv_t<FZ, 8>;
    case 4: return (const void*)k3_zconv_t<FZ, 4>;
    default: return nullptr;
    }
}

// zt: fz for the full compile-time pass, fz + 1 for the open split pass
static const void* zconv_template(int zt, int w)
{
    switch (zt) {
    case 16: return zfull_kernel<16>(w);
    case 24: return zfull_kernel<24>(w);
    case 32: return zfull_kernel<32>(w);
    case 64: return zfull_kernel<64>(w);
    case 128: return zfull_kernel<128>(w);
    case 256: return zfull_kernel<256>(w);
    case 512: return zfull_kernel<512>(w);
    case 33: return zopen_kernel<16>(w);
    case 65: return zopen_kernel<32>(w);
    case 129: return zopen_kernel<64>(w);
    case 257: return zopen_kernel<128>(w);
    case 513: return zopen_kernel<256>(w);
    default: return nullptr;
    }
}

int zconv_template_width(int fz)
{
    for (int w : {16, 8})
        if (zconv_template(fz, w)) return w;
    return 0;
}

int zconv_open_width(int fz)
{
    const int nz = fz / 2;
    return (fz % 2 == 0 && zconv_template(fz + 1, zopen_width(nz))) ? zopen_width(nz) : 0;
}

int zconv_widths(int zt, int* w, int cap)
{
    int n = 0;
    const int first = (zt & 1) ? zopen_width(zt / 2) : zconv_template_width(zt);
    if (first && n < cap) w[n++] = first;
    for (int c : {16, 8, 4, 2})
        if (c != first && zconv_template(zt, c) && n < cap) w[n++] = c;
    return n;
}

// --------------------------------------------------------------- launches
int rows_template_height(int fx)
{
    switch (fx) {
    case 32: return rows_template_ty<32>();
    case 64: return rows_template_ty<64>();
    case 96: return rows_template_ty<96>();
    case 128: return rows_template_ty<128>();
    case 256: return rows_template_ty<256>();
    case 512: return rows_template_ty<512>();
    default: return 0;
    }
}

template <int MODE, bool COMP>
static const void* rows_kernel(int rt)
{
    switch (rt) {
    case 0: return (const void*)k3_rows<MODE, COMP>;
    case 32: return (const void*)k3_rows_t<MODE, COMP, 32>;
    case 64: return (const void*)k3_rows_t<MODE, COMP, 64>;
    case 96: return (const void*)k3_rows_t<MODE, COMP, 96>;
    case 128: return (const void*)k3_rows_t<MODE, COMP, 128>;
    case 256: return (const void*)k3_rows_t<MODE, COMP, 256>;
    case 512: return (const void*)k3_rows_t<MODE, COMP, 512>;
    default: return nullptr;
    }
}

static const void* rows_kernel(int mode, bool comp, int rt)
{
    switch (mode) {
    case R3_FWD: return rows_kernel<R3_FWD, false>(rt);
    case R3_INV: return rows_kernel<R3_INV, false>(rt);
    case R3_PRED: return comp ? rows_kernel<R3_PRED, true>(rt) : rows_kernel<R3_PRED, false>(rt);
    default: return comp ? rows_kernel<R3_CORR, true>(rt) : rows_kernel<R3_CORR, false>(rt);
    }
}

hipError_t launch3_rows(int mode, bool comp, const Rep3* reps, const int4* items, int nitems, size_t lds, int rt,
                        hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
    const void* f = rows_kernel(mode, comp, rt);
    if (!f) return hipErrorInvalidValue;
    const int nt = rt >= 512 ? rows_template_threads<512>() : R3_TPB;
    void* args[] = {(void*)&reps, (void*)&items};
    return hipLaunchKernel(f, dim3(nitems), dim3(nt), args, lds, st);
}

hipError_t launch3_ypass(int sign, const Rep3* reps, const int4* items, int nitems, size_t lds, int yt, int wy,
                         hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
    if (yt) {
        const void* f = sign < 0 ? ypass_template<-1>(yt, wy) : ypass_template<1>(yt, wy);
        if (!f) return hipErrorInvalidValue;
        void* args[] = {(void*)&reps, (void*)&items};
        return hipLaunchKernel(f, dim3(nitems), dim3(Y3_TPB), args, lds, st);
    }
    if (sign < 0) k3_ypass<-1><<<nitems, Y3_TPB, lds, st>>>(reps, items);
    else k3_ypass<1><<<nitems, Y3_TPB, lds, st>>>(reps, items);
    return hipGetLastError();
}

hipError_t launch3_zconv(const Rep3* reps, const int4* items, int nitems, size_t lds, int zt, int wz, hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
    const void* f = zt ? zconv_template(zt, wz) : (const void*)k3_zconv;
    if (!f) return hipErrorInvalidValue;
    const int nt = (zt & 1) ? zopen_threads((zt - 1) / 2, wz) : Z3_TPB;   // open split: zt = 2 nz + 1
    void* args[] = {(void*)&reps, (void*)&items};
    return hipLaunchKernel(f, dim3(nitems), dim3(nt), args, lds, st);
}

hipError_t configure_lds3(size_t max_bytes)
{
    const int v = (int)max_bytes;
    const void* fns[] = {(const void*)k3_ypass<-1>, (const void*)k3_ypass<1>, (const void*)k3_zconv};
    hipError_t first = hipSuccess;
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, v);
        if (e != hipSuccess && first == hipSuccess) first = e;
    }
    for (int rt : {0, 32, 64, 96, 128, 256, 512})
        for (int mode : {R3_FWD, R3_INV, R3_PRED, R3_CORR})
            for (bool comp : {false, true}) {
                hipError_t e = hipFuncSetAttribute(rows_kernel(mode, comp, rt), hipFuncAttributeMaxDynamicSharedMemorySize, v);
                if (e != hipSuccess && first == hipSuccess) first = e;
            }
    for (int f : {16, 24, 32, 64, 96, 128, 256, 512})
        for (int w : {2, 4, 8, 16})
            for (const void* k : {zconv_template(f, w), zconv_template(f + 1, w), ypass_template<-1>(f, w),
                                  ypass_template<1>(f, w)}) {
                if (!k) continue;    // no instance of that pass at this length and width
                hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, v);
                if (e != hipSuccess && first == hipSuccess) first = e;
            }
    (void)hipGetLastError();
    return first;
}

// ------------------------------------------------------- tensor (FP64)
__device__ __forceinline__ void gamma6(double rx, double ry, double rz, double* g)
{
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    if (r < 1e-40) { for (int i = 0; i < 6; ++i) g[i] = 0; return; }
    const double ir = 1.0 / r, ir3 = ir * ir * ir, ir5 = ir3 * ir * ir;
    g[0] = -1.0 * (ir3 - 3.0 * rx * rx * ir5);
    g[1] = -1.0 * (-3.0 * rx * ry * ir5);
    g[2] = -1.0 * (-3.0 * rx * rz * ir5);
    g[3] = -1.0 * (ir3 - 3.0 * ry * ry * ir5);
    g[4] = -1.0 * (-3.0 * ry * rz * ir5);
    g[5] = -1.0 * (ir3 - 3.0 * rz * rz * ir5);
}

__device__ __forceinline__ long long fdiv3(long long a, long long b)
{
    long long q = a / b;
    if ((a % b) && ((a < 0) != (b < 0))) --q;
    return q;
}

// R > 0: periodic lattice sum over images within radius R (grid = lattice);
// R == 0: open boundaries, Gamma at the wrapped displacement on the doubled grid.
__global__ void k3_tensor(int nx, int ny, int nz, int fx, int fy, int fz, int R, double* __restrict__ T)
{
    const size_t total = (size_t)fx * fy * fz;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % fx);
        const int j = (int)((idx / fx) % fy);
        const int k = (int)(idx / ((size_t)fx * fy));
        double acc[6] = {0, 0, 0, 0, 0, 0};
        if (R > 0) {
            const long long LR = R;
            const long long x0 = -fdiv3(LR + i, nx), x1 = fdiv3(LR - i, nx);
            const long long y0 = -fdiv3(LR + j, ny), y1 = fdiv3(LR - j, ny);
            const long long z0 = -fdiv3(LR + k, nz), z1 = fdiv3(LR - k, nz);
            const double r2max = (double)R * R;
            for (long long X = x0; X <= x1; ++X)
                for (long long Y = y0; Y <= y1; ++Y)
                    for (long long Z = z0; Z <= z1; ++Z) {
                        const double rx = (double)(X * nx + i), ry = (double)(Y * ny + j), rz = (double)(Z * nz + k);
                        if (rx * rx + ry * ry + rz * rz > r2max) continue;
                        double g[6];
                        gamma6(rx, ry, rz, g);
                        for (int a = 0; a < 6; ++a) acc[a] += g[a];
                    }
        } else if (i != nx && j != ny && k != nz) {
            gamma6(i < nx ? i : i - fx, j < ny ? j : j - fy, k < nz ? k : k - fz, acc);
        }
        for (int a = 0; a < 6; ++a) T[(size_t)a * total + idx] = acc[a];
    }
}

hipError_t launch3_tensor(int nx, int ny, int nz, int fx, int fy, int fz, int R, double* T, hipStream_t st)
{
    const size_t total = (size_t)fx * fy * fz;
    const int g = (int)std::min<size_t>((total + 127) / 128, 65536);
    k3_tensor<<<g, 128, 0, st>>>(nx, ny, nz, fx, fy, fz, R, T);
    return hipGetLastError();
}

// x DFT (real -> half spectrum): U1[a][z][y][kx]
__global__ void k3_dftx(int fx, int fy, int fz, const double* __restrict__ T, double2* __restrict__ U, const double2* __restrict__ tw)
{
    const int fxh = fx / 2 + 1;
    const size_t total = (size_t)6 * fz * fy * fxh;
    const size_t plane = (size_t)fx * fy * fz;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(idx % fxh);
        const size_t rowi = idx / fxh;               // a*fz*fy + z*fy + y
        const int a = (int)(rowi / ((size_t)fz * fy));
        const size_t zy = rowi - (size_t)a * fz * fy;
        const double* row = T + (size_t)a * plane + zy * fx;
        double re = 0, im = 0;
        int m = 0;
        for (int x = 0; x < fx; ++x) {
            const double2 w = tw[m];
            re += row[x] * w.x;
            im += row[x] * w.y;
            m += kx; if (m >= fx) m -= fx;
        }
        U[idx] = make_double2(re, im);
    }
}

// complex DFT along y (stride fxh) or z (stride fy*fxh) of [6][fz][fy][fxh]
template <int AXIS>
__global__ void k3_dftyz(int fy, int fz, int fxh, const double2* __restrict__ U, double2* __restrict__ V,
                         const double2* __restrict__ tw)
{
    const size_t total = (size_t)6 * fz * fy * fxh;
    const int L = AXIS == 1 ? fy : fz;
    const size_t stride = AXIS == 1 ? (size_t)fxh : (size_t)fy * fxh;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(idx % fxh);
        const int y = (int)((idx / fxh) % fy);
        const int z = (int)((idx / ((size_t)fxh * fy)) % fz);
        const size_t a = idx / ((size_t)fxh * fy * fz);
        const int kk = AXIS == 1 ? y : z;
        const size_t base = a * (size_t)fz * fy * fxh + (AXIS == 1 ? ((size_t)z * fy * fxh + kx) : ((size_t)y * fxh + kx));
        double re = 0, im = 0;
        int m = 0;
        for (int t = 0; t < L; ++t) {
            const double2 u = U[base + (size_t)t * stride];
            const double2 w = tw[m];
            re += u.x * w.x - u.y * w.y;
            im += u.x * w.y + u.y * w.x;
            m += kk; if (m >= L) m -= L;
        }
        V[idx] = make_double2(re, im);
    }
}

// ---------------------------------------------------------------- FP64 FFT
// Mixed-radix Stockham transforms of FP64 lines in LDS for the tensor spectrum
// (forward sign, the table tw[k] = exp(-2 pi i k / L) of the axis).  A workgroup
// takes `lpb` lines at a time.  Line maps over the [6][fz][fy][*] arrays:
//   kind 0: real rows T[a][z][y][0..fx) -> half spectra U[a][z][y][0..fxh);
//   kind 1: columns along y of [6][fz][fy][fxh]; kind 2: columns along z.
struct Lines64 {
    int kind, fx, fy, fz, fxh;
};

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ size_t line_elem(const Lines64& m, int line, int t)
{
    if (m.kind == 1) {
        const int a = line / (m.fz * m.fxh), r = line - a * m.fz * m.fxh, z = r / m.fxh, kx = r - z * m.fxh;
        return (((size_t)a * m.fz + z) * m.fy + t) * m.fxh + kx;
    }
    const int a = line / (m.fy * m.fxh), r = line - a * m.fy * m.fxh, y = r / m.fxh, kx = r - y * m.fxh;
    return (((size_t)a * m.fz + t) * m.fy + y) * m.fxh + kx;
}

__global__ void k3_fft64(Lines64 m, FFTPlan p, const double2* __restrict__ tw, const double* __restrict__ Treal,
                         const double2* __restrict__ in, double2* __restrict__ out, int nlines, int lpb)
{
    extern __shared__ double2 z64[];
    const int L = p.L;
    for (int lb = blockIdx.x * lpb; lb < nlines; lb += gridDim.x * lpb) {
        const int nl = min(lpb, nlines - lb);
        double2* A = z64;
        double2* Bf = z64 + (size_t)lpb * L;
        // load: rows with t fastest (contiguous), columns with the line fastest (adjacent kx)
        for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
            int ln, t;
            if (m.kind == 0) { ln = e / L; t = e - ln * L; }
            else { t = e / nl; ln = e - t * nl; }
            const int line = lb + ln;
            double2 v;
            if (m.kind == 0) v = make_double2(Treal[(size_t)line * m.fx + t], 0.0);
            else v = in[line_elem(m, line, t)];
            A[ln * L + t] = v;
        }
        __syncthreads();
        int Ns = 1;
        double2* src = A;
        double2* dst = Bf;
        for (int s = 0; s < p.nst; ++s) {
            const int R = p.rad[s], nb = L / R, tstep = L / (Ns * R), rstep = L / R;
            for (int b = threadIdx.x; b < nl * nb; b += blockDim.x) {
                const int ln = b / nb, j = b - ln * nb, k = j % Ns;
                double2 v[8], V[8];
                for (int r = 0; r < R; ++r) {
                    v[r] = src[ln * L + j + r * nb];
                    if (Ns > 1 && r) v[r] = zmul(v[r], tw[k * r * tstep]);
                }
                for (int q = 0; q < R; ++q) {
                    double2 acc = v[0];
                    for (int r = 1; r < R; ++r) {
                        const double2 w = tw[((r * q) % R) * rstep];
                        acc.x += v[r].x * w.x - v[r].y * w.y;
                        acc.y += v[r].x * w.y + v[r].y * w.x;
                    }
                    V[q] = acc;
                }
                double2* d = dst + ln * L + (j / Ns) * Ns * R + k;
                for (int q = 0; q < R; ++q) d[q * Ns] = V[q];
            }
            __syncthreads();
            double2* t2 = src; src = dst; dst = t2;
            Ns *= R;
        }
        for (int e = threadIdx.x; e < nl * L; e += blockDim.x) {
            int ln, t;
            if (m.kind == 0) { ln = e / L; t = e - ln * L; }
            else { t = e / nl; ln = e - t * nl; }
            const int line = lb + ln;
            if (m.kind == 0) { if (t < m.fxh) out[(size_t)line * m.fxh + t] = src[ln * L + t]; }
            else out[line_elem(m, line, t)] = src[ln * L + t];
        }
        __syncthreads();
    }
}

// FP32 tensor words [6][fys][fxh][fzs] (kz fastest, kz < fzs, ky < fys) from the FP64
// spectrum V [6][fz][fy][fxh]
__global__ void k3_words(const double2* __restrict__ V, float* __restrict__ Tw, double mult, int fy, int fys, int fz,
                         int fxh, int fzs)
{
    const size_t n = (size_t)6 * fys * fxh * fzs;
    for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < n; o += (size_t)gridDim.x * blockDim.x) {
        const int kz = (int)(o % fzs);
        const size_t r = o / fzs;
        const int kx = (int)(r % fxh);
        const size_t r2 = r / fxh;
        const int ky = (int)(r2 % fys);
        const int a = (int)(r2 / fys);
        Tw[o] = (float)(V[(((size_t)a * fz + kz) * fy + ky) * fxh + kx].x * mult);
    }
}

// One axis of the FP64 tensor transform as LDS Stockham passes; lines per
// workgroup sized to 64 KiB of LDS (two buffers).
static hipError_t fft64_axis(int kind, int fx, int fy, int fz, const FFTPlan& p, const double2* tw, const double* T,
                             const double2* in, double2* out, hipStream_t st)
{
    const int fxh = fx / 2 + 1;
    const Lines64 m{kind, fx, fy, fz, fxh};
    const int L = p.L;
    const int nlines = kind == 0 ? 6 * fz * fy : (kind == 1 ? 6 * fz * fxh : 6 * fy * fxh);
    const int lpb = std::max(1, std::min(64, (int)(65536 / (2 * (size_t)L * sizeof(double2)))));
    const size_t lds = (size_t)2 * lpb * L * sizeof(double2);
    const int g = std::min((nlines + lpb - 1) / lpb, 4096);
    k3_fft64<<<g, 256, lds, st>>>(m, p, tw, T, in, out, nlines, lpb);
    return hipGetLastError();
}

hipError_t configure_fft64(size_t bytes)
{
    return hipFuncSetAttribute((const void*)k3_fft64, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch3_spectrum(int fx, int fy, int fz, int fys, int fzs, const double* T, double2* U1, double2* U2, float* Tw,
                            double mult, const double2* twx, const double2* twy, const double2* twz, hipStream_t st)
{
    const int fxh = fx / 2 + 1;
    const size_t tot = (size_t)6 * fz * fy * fxh;
    const int g = (int)std::min<size_t>((tot + 255) / 256, 65536);
    FFTPlan px, py, pz;
    // transforms need each length to factor into 2, 3 and 5 (radix <= 5 here); otherwise direct DFTs
    auto plan64 = [](int L, FFTPlan& q) {
        q.L = L;
        q.nst = 0;
        int mm = L;
        for (int r : {4, 2, 3, 5})
            while (mm % r == 0 && q.nst < 16) { q.rad[q.nst++] = r; mm /= r; }
        return mm == 1;
    };
    if (plan64(fx, px) && plan64(fy, py) && plan64(fz, pz) && fx <= 4096 && fy <= 4096 && fz <= 4096) {
        hipError_t e = fft64_axis(0, fx, fy, fz, px, twx, T, nullptr, U1, st);
        if (e == hipSuccess) e = fft64_axis(1, fx, fy, fz, py, twy, nullptr, U1, U2, st);
        if (e == hipSuccess) e = fft64_axis(2, fx, fy, fz, pz, twz, nullptr, U2, U1, st);
        if (e != hipSuccess) return e;
    } else {
        k3_dftx<<<g, 256, 0, st>>>(fx, fy, fz, T, U1, twx);
        k3_dftyz<1><<<g, 256, 0, st>>>(fy, fz, fxh, U1, U2, twy);
        k3_dftyz<2><<<g, 256, 0, st>>>(fy, fz, fxh, U2, U1, twz);
    }
    k3_words<<<g, 256, 0, st>>>(U1, Tw, mult, fy, fys, fz, fxh, fzs);
    return hipGetLastError();
}

}  // namespace msd
