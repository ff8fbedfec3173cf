// Layered dipole convolution (LongRange::apply, cpu/longrange/spinoperationlongrange.cpp:354-430)
// and the fused Heun step kernels for gfx950.
//
// Spectra live in HBM in "column" layout [comp][layer][kx][ky] (ky contiguous):
//   row kernels  : one workgroup per tile of `ty` consecutive y-rows of one layer,
//                  real<->complex transforms along x in LDS (two real rows packed
//                  into one complex line), ty*8-byte contiguous runs per kx;
//   column kernel: one workgroup per kx; the 3*nz lines along y are transformed in
//                  LDS, multiplied by the per-k 3nz x 3nz layer/component tensor
//                  in registers, and transformed back -- one HBM read and one
//                  write of each spectrum per field evaluation.
// The tensor is stored as one FP32 word per component per k per layer offset
// 0..nz-1: XX XY YY ZZ spectra are real, XZ YZ imaginary, and XZ/YZ are odd in
// the layer offset (source - target, spinoperationlongrange.cpp:377), so a
// negative offset flips their sign.  1/(nx*ny), g and the global scale are folded in.
//
// Every kernel exists as a compile-time-shaped instance for power-of-two line
// lengths 32..1024 (NX/NY > 0: unrolled stages, LDS twiddles) and as a runtime-
// shaped instance (NX/NY == 0) for the other mixed-radix sizes.
#include "msd_device.h"

#include <type_traits>

namespace msd {

// threads per workgroup of the runtime-shaped row kernels (compile-time shapes
// carry their own count)
static constexpr int ROW_TPB = 1024;
// Threads per column workgroup: 256 for columns of up to 256 (one thread per ky, so
// every thread takes part in the per-k product; measured +0.9% whole-step at
// 256x256x8), 512 for longer columns (measured +1.1% at 512x512x4) and runtime lengths.
__host__ __device__ constexpr int col_tpb(int ny) { return ny > 0 && ny <= 256 ? 256 : 512; }
// Waves per SIMD the fused step kernels are compiled to allow: 4, a budget of 128
// VGPRs.  Workgroups of 512 threads then fit two to a CU (without the hint the
// corrector at 512 threads took more registers, one workgroup per CU and two rounds
// of tiles: 28 against 20 us at 256x256x8); at 1024 threads a budget of 64 spills.
__host__ __device__ constexpr int row_waves(int) { return 4; }
// exchange class tables up to this size are staged in LDS by the fused kernels

// Workgroups on the same XCD (the dispatcher deals bid % 8 to XCD bid % 8) get
// contiguous tiles: XCD r takes the r-th run of nb/8 (+1 for the first nb % 8)
// consecutive tiles.  A bijection of [0, nb) for any nb.
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    const int r = bid & 7, q = bid >> 3, base = nb >> 3, rem = nb & 7;
    return r * base + (r < rem ? r : rem) + q;
}

// two real rows per complex line: real row q -> line q/2, component q&1
__device__ __forceinline__ float& rowval(float2* buf, int q, int x, int nx)
{
    return reinterpret_cast<float*>(buf + (size_t)(q >> 1) * nx + x)[q & 1];
}

__host__ __device__ constexpr int row_lines(int ty) { return (3 * ty + 1) >> 1; }

// Build the packed complex lines Z = A + iB from the half spectra of rows
// (2l, 2l+1) read from `spec` (column layout), using A[nx-k] = conj(A[k]).
// Compile-time shapes (NX > 0) issue the loads in batches.
template <int NT, int NX = 0, int TY = 0>
__device__ __forceinline__ void rows_pack_spectrum(const float2* __restrict__ spec, int nx, int ny, int nz, int ty,
                                                   int z, int y0, float2* __restrict__ lds)
{
    const int nxh = nx / 2 + 1, lines = row_lines(ty), total = lines * nx, half = nx >> 1;
    auto ld = [&](int idx) {
        const int l = idx % lines;
        const int k = idx / lines;
        const bool mirror = k > half;
        const int kk = mirror ? nx - k : k;
        const int q0 = 2 * l, q1 = q0 + 1;
        const int c0 = q0 / ty, r0 = q0 - c0 * ty;
        float2 A = gload(spec, ((size_t)(c0 * nz + z) * nxh + kk) * ny + y0 + r0);
        float2 B = make_float2(0.f, 0.f);
        if (q1 < 3 * ty) {
            const int c1 = q1 / ty, r1 = q1 - c1 * ty;
            B = gload(spec, ((size_t)(c1 * nz + z) * nxh + kk) * ny + y0 + r1);
        }
        return make_float4(A.x, A.y, B.x, B.y);
    };
    auto st = [&](int idx, float4 v) {
        const int l = idx % lines;
        const int k = idx / lines;
        if (k > half) { v.y = -v.y; v.w = -v.w; }
        lds[(size_t)l * nx + k] = make_float2(v.x - v.w, v.y + v.z);
    };
    if constexpr (NX > 0) {
        batched_copy<row_lines(TY) * NX, NT, 8>(ld, st);
    } else {
        for (int idx = threadIdx.x; idx < total; idx += NT) st(idx, ld(idx));
    }
}

// Split the transformed packed lines into the half spectra of both rows and
// store them in column layout.
template <int NT>
__device__ __forceinline__ void rows_unpack_spectrum(const float2* __restrict__ lds, int nx, int ny, int nz, int ty, int z,
                                                     int y0, float2* __restrict__ spec)
{
    const int nxh = nx / 2 + 1, lines = row_lines(ty), total = lines * nxh;
    for (int idx = threadIdx.x; idx < total; idx += NT) {
        const int l = idx % lines;
        const int k = idx / lines;
        const float2 Zk = lds[(size_t)l * nx + k];
        const float2 Zm = lds[(size_t)l * nx + (k ? nx - k : 0)];
        const int q0 = 2 * l, q1 = q0 + 1;
        const int c0 = q0 / ty, r0 = q0 - c0 * ty;
        gstore(spec, ((size_t)(c0 * nz + z) * nxh + k) * ny + y0 + r0, make_float2(0.5f * (Zk.x + Zm.x), 0.5f * (Zk.y - Zm.y)));
        if (q1 < 3 * ty) {
            const int c1 = q1 / ty, r1 = q1 - c1 * ty;
            gstore(spec, ((size_t)(c1 * nz + z) * nxh + k) * ny + y0 + r1,
                   make_float2(0.5f * (Zk.y + Zm.y), 0.5f * (Zm.x - Zk.x)));
        }
    }
}

// Row r of a row tile lies at y = y0 + ys * r of layer z: consecutive rows
// (y0 = yb ty, ys = 1), or with a split y transform the rows b + ly a (y0 = b = yb,
// ys = ly).
struct RowTile {
    int z, yb, y0, ys;
};

__device__ __forceinline__ RowTile row_tile(const DipoleDev& dp, int tile, int nz, int ty)
{
    RowTile t;
    t.yb = tile / nz;
    t.z = tile - t.yb * nz;
    if (dp.split) { t.y0 = t.yb; t.ys = dp.ly; }
    else { t.y0 = t.yb * ty; t.ys = 1; }
    return t;
}

template <int R, int SIGN> __device__ __forceinline__ void dft_r(float2* v)
{
    if constexpr (R == 2) dft2<SIGN>(v);
    else if constexpr (R == 4) dft4<SIGN>(v);
    else dft8<SIGN>(v);
}

__host__ __device__ constexpr bool split_rows(int ty) { return ty == 2 || ty == 4 || ty == 8; }

// Staging area of the split row passes: the tile's 3 (nx/2+1) runs of TY words,
// after the packed lines and the x twiddles.  Global reads and writes of the runs
// go through it as whole float4s spread over all threads (coalesced), and one
// thread per (c, kx) works on its run in LDS.
template <int NX, int TY>
__device__ __forceinline__ float4* split_stage(float2* lds)
{
    return reinterpret_cast<float4*>(lds + (size_t)row_lines(TY) * NX + NX);
}

// Split y transform, inverse side: for every (component c, kx <= nx/2) the TY words
// Z_b(k1) of this tile (b = yb) become the tile's TY rows at kx,
//   x(b + ly a) = sum_k1 exp(+2 pi i a k1 / TY) exp(+2 pi i b k1 / ny) Z_b(k1),
// packed two rows to a complex line as rows_pack_spectrum does.
template <int NT, int NX, int TY>
__device__ __forceinline__ void rows_pack_split(const DipoleDev& dp, const float2* __restrict__ spec, int nz,
                                                const RowTile& t, float2* __restrict__ lds)
{
    constexpr int NXH = NX / 2 + 1, HALF = NX / 2, ITEMS = 3 * NXH, CH = TY / 2;
    const int ny = dp.ny;
    float4* stg = split_stage<NX, TY>(lds);
    const float4* __restrict__ src = reinterpret_cast<const float4*>(spec);
    batched_copy<ITEMS * CH, NT, 4>(
        [&](int idx) {
            const int item = idx / CH, h = idx - item * CH;
            const int c = item / NXH, kx = item - c * NXH;
            return gload(src, ((((size_t)(c * nz + t.z) * NXH + kx) * ny + t.yb * TY) >> 1) + h);
        },
        [&](int idx, float4 v) { stg[idx] = v; });
    float2 tw[TY];
#pragma unroll
    for (int k = 1; k < TY; ++k) tw[k] = gload(dp.twy, t.yb * k);
    __syncthreads();
    for (int idx = threadIdx.x; idx < ITEMS; idx += NT) {
        const int c = idx / NXH, kx = idx - c * NXH;
        float2 v[TY];
#pragma unroll
        for (int h = 0; h < CH; ++h) {
            const float4 q = stg[idx * CH + h];
            v[2 * h] = make_float2(q.x, q.y);
            v[2 * h + 1] = make_float2(q.z, q.w);
        }
#pragma unroll
        for (int k = 1; k < TY; ++k) v[k] = cmulc(v[k], tw[k]);
        dft_r<TY, 1>(v);
#pragma unroll
        for (int j = 0; j < TY / 2; ++j) {
            const float2 A = v[2 * j], B = v[2 * j + 1];
            float2* line = lds + (size_t)(c * (TY / 2) + j) * NX;
            line[kx] = make_float2(A.x - B.y, A.y + B.x);
            if (kx > 0 && kx < HALF) line[NX - kx] = make_float2(A.x + B.y, B.x - A.y);
        }
    }
}

// Split y transform, forward side: the tile's TY transformed rows at kx become
//   Z_b(k1) = exp(-2 pi i b k1 / ny) sum_a exp(-2 pi i a k1 / TY) X_{b + ly a}(kx),
// stored as one run of TY words per (c, kx).  Caller synchronised before.
template <int NT, int NX, int TY>
__device__ __forceinline__ void rows_unpack_split(const DipoleDev& dp, const float2* __restrict__ lds, int nz,
                                                  const RowTile& t, float2* __restrict__ spec)
{
    constexpr int NXH = NX / 2 + 1, ITEMS = 3 * NXH, CH = TY / 2;
    const int ny = dp.ny;
    float4* stg = split_stage<NX, TY>(const_cast<float2*>(lds));
    float2 tw[TY];
#pragma unroll
    for (int k = 1; k < TY; ++k) tw[k] = gload(dp.twy, t.yb * k);
    for (int idx = threadIdx.x; idx < ITEMS; idx += NT) {
        const int c = idx / NXH, kx = idx - c * NXH;
        float2 v[TY];
#pragma unroll
        for (int j = 0; j < TY / 2; ++j) {
            const float2* line = lds + (size_t)(c * (TY / 2) + j) * NX;
            const float2 Zk = line[kx], Zm = line[kx ? NX - kx : 0];
            v[2 * j] = make_float2(0.5f * (Zk.x + Zm.x), 0.5f * (Zk.y - Zm.y));
            v[2 * j + 1] = make_float2(0.5f * (Zk.y + Zm.y), 0.5f * (Zm.x - Zk.x));
        }
        dft_r<TY, -1>(v);
#pragma unroll
        for (int k = 1; k < TY; ++k) v[k] = cmul(v[k], tw[k]);
#pragma unroll
        for (int h = 0; h < CH; ++h) stg[idx * CH + h] = make_float4(v[2 * h].x, v[2 * h].y, v[2 * h + 1].x, v[2 * h + 1].y);
    }
    __syncthreads();
    float4* __restrict__ dst = reinterpret_cast<float4*>(spec);
    for (int idx = threadIdx.x; idx < ITEMS * CH; idx += NT) {
        const int item = idx / CH, h = idx - item * CH;
        const int c = item / NXH, kx = item - c * NXH;
        gstore(dst, ((((size_t)(c * nz + t.z) * NXH + kx) * ny + t.yb * TY) >> 1) + h, stg[idx]);
    }
}

// The field spectrum into the tile's packed LDS lines (before the inverse x transform)
template <int NT, int NX, int TY>
__device__ __forceinline__ void rows_in(const DipoleDev& dp, const float2* __restrict__ spec, int nx, int ny, int nz,
                                        int ty, const RowTile& t, float2* __restrict__ lds)
{
    if constexpr (NX > 0 && split_rows(TY)) {
        if (dp.split) { rows_pack_split<NT, NX, TY>(dp, spec, nz, t, lds); return; }
    }
    rows_pack_spectrum<NT, NX, TY>(spec, nx, ny, nz, ty, t.z, t.y0, lds);
}

// The tile's forward-transformed LDS lines into the spin spectrum
template <int NT, int NX, int TY>
__device__ __forceinline__ void rows_out(const DipoleDev& dp, const float2* __restrict__ lds, int nx, int ny, int nz,
                                         int ty, const RowTile& t, float2* __restrict__ spec)
{
    if constexpr (NX > 0 && split_rows(TY)) {
        if (dp.split) { rows_unpack_split<NT, NX, TY>(dp, lds, nz, t, spec); return; }
    }
    rows_unpack_spectrum<NT>(lds, nx, ny, nz, ty, t.z, t.y0, spec);
}

// Copy a small exchange class table into LDS so the per-site lookups cost no
// global round trip; larger tables stay in global memory.
__device__ __forceinline__ ExchangeDev stage_exchange(const ExchangeDev& ex, int* s_start, ExEntry* s_ent)
{
    ExchangeDev e = ex;
    if (ex.mode == 1) {     // always fits (msd_exchange::fused_dev)
        for (int i = threadIdx.x; i <= ex.ncls; i += blockDim.x) s_start[i] = ex.cls_start[i];
        for (int i = threadIdx.x; i < ex.nent; i += blockDim.x) s_ent[i] = ex.ent[i];
        e.cls_start = s_start;
        e.ent = s_ent;
    }
    return e;
}

// stage the x twiddles in LDS (compile-time shapes only)
template <int NX>
__device__ __forceinline__ const float2* row_twiddles(const DipoleDev& dp, float2* lds_tw)
{
    if constexpr (NX > 0) {
        for (int i = threadIdx.x; i < NX; i += blockDim.x) lds_tw[i] = gload(dp.twx, i);
        return lds_tw;
    } else {
        return dp.twx;
    }
}

template <int NX, int TY, int TPB, int SIGN>
__device__ __forceinline__ float2* row_fft(float2* A, float2* B, const DipoleDev& dp, const float2* tw)
{
    if constexpr (NX > 0) {
        fft_ip<NX, row_lines(TY), TPB, SIGN>(A, tw);   // in place: B is not used
        return A;
    }
    else return lds_fft<SIGN>(A, B, row_lines(dp.ty), dp.px, tw);
}

// ------------------------------------------------------------ op-API rows
// SpinSystem::fft (cpu/core/spinsystem.cpp:936-966) along x: spins -> Mh
template <int NX, int TY, int TPB>
__global__ void __launch_bounds__(TPB) k_rows_fwd(DipoleDev dp, const float* __restrict__ s)
{
    extern __shared__ float2 lds[];
    const int nx = NX ? NX : dp.nx, ty = TY ? TY : dp.ty, lines = row_lines(ty), ny = dp.ny;
    // layer fastest: an XCD's run of tiles is a band of rows through every layer
    const RowTile t = row_tile(dp, xcd_remap(blockIdx.x, gridDim.x), dp.nz, ty);
    const int n = nx * ny * dp.nz;
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    const float2* tw = row_twiddles<NX>(dp, A + (size_t)lines * nx);
    for (int li = threadIdx.x; li < lines * nx; li += TPB) A[li] = make_float2(0.f, 0.f);
    __syncthreads();
    auto ld = [&](int li) {
        const int q = li / nx, x = li - q * nx;
        const int c = q / ty, r = q - c * ty;
        return s[(size_t)c * n + x + (size_t)nx * (t.y0 + t.ys * r + (size_t)ny * t.z)];
    };
    auto st = [&](int li, float v) { const int q = li / nx; rowval(A, q, li - q * nx, nx) = v; };
    if constexpr (NX > 0) batched_copy<3 * TY * NX, TPB, 8>(ld, st);
    else for (int li = threadIdx.x; li < 3 * ty * nx; li += TPB) st(li, ld(li));
    __syncthreads();
    const float2* R = row_fft<NX, TY, TPB, -1>(A, B, dp, tw);
    rows_out<TPB, NX, TY>(dp, R, nx, ny, dp.nz, ty, t, dp.Mh);
}

// inverse along x: Hh -> real field rows, written to slot h
template <int NX, int TY, int TPB>
__global__ void __launch_bounds__(TPB) k_rows_inv(DipoleDev dp, float* __restrict__ h)
{
    extern __shared__ float2 lds[];
    const int nx = NX ? NX : dp.nx, ty = TY ? TY : dp.ty, lines = row_lines(ty), ny = dp.ny;
    // layer fastest: an XCD's run of tiles is a band of rows through every layer
    const RowTile t = row_tile(dp, xcd_remap(blockIdx.x, gridDim.x), dp.nz, ty);
    const int n = nx * ny * dp.nz;
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    const float2* tw = row_twiddles<NX>(dp, A + (size_t)lines * nx);
    rows_in<TPB, NX, TY>(dp, dp.Hh, nx, ny, dp.nz, ty, t, A);
    __syncthreads();
    float2* R = row_fft<NX, TY, TPB, 1>(A, B, dp, tw);
    for (int li = threadIdx.x; li < 3 * ty * nx; li += TPB) {
        const int q = li / nx, x = li - q * nx;
        const int c = q / ty, r = q - c * ty;
        h[(size_t)c * n + x + (size_t)nx * (t.y0 + t.ys * r + (size_t)ny * t.z)] = rowval(R, q, x, nx);
    }
}

// ------------------------------------------------------------- column pass
// For one kx and one output component a: forward along y of all 3nz input lines,
// the per-k layer/component tensor product for component a, inverse along y of
// its nz lines.  Three workgroups per kx (one per output component) give the
// launch 3(nx/2+1) workgroups; the input column is read three times, from L2.
// wg: the workgroup's place among the 3 (nx/2+1) of its film (XCD-remapped, so the
// three workgroups of one kx run on one XCD and two of them find the column in
// that XCD's L2)
template <int NY, int NZ>
__device__ __forceinline__ void cols_body(const DipoleDev& dp, int wg)
{
    constexpr int COL_TPB = col_tpb(NY);
    extern __shared__ float2 lds[];
    const int kx = wg / 3;
    const int oa = wg - 3 * kx;
    const int ny = NY ? NY : dp.ny, nxh = dp.nxh;
    constexpr int lines = 3 * NZ;
    // LDS: lines (one buffer in place, two ping-pong for runtime shapes) and the y twiddles
    float2* A = lds;
    float2* B = lds + (size_t)lines * ny;
    const float2* tw = dp.twy;
    // row oa of the symmetric tensor: components (oa,0) (oa,1) (oa,2) and which are odd in z
    const int cidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    const int c0 = cidx[oa][0], c1 = cidx[oa][1], c2 = cidx[oa][2];
    const bool odd0 = (c0 == 2 || c0 == 4), odd1 = (c1 == 2 || c1 == 4), odd2 = (c2 == 2 || c2 == 4);
    // Compile-time lengths: the twiddles and this workgroup's tensor words
    // (components c0 c1 c2, offsets 0..NZ-1, every ky) are staged in LDS after the
    // lines, loaded in the same batch as the column so one memory latency covers all.
    float* wl = nullptr;
    if constexpr (NY > 0) {
        float2* twl = A + (size_t)lines * NY;   // in-place transforms: no second buffer
        wl = reinterpret_cast<float*>(twl + NY);
        constexpr int TWI = (NY + COL_TPB - 1) / COL_TPB;
        constexpr int MI = (lines * NY + COL_TPB - 1) / COL_TPB;
        constexpr int WI = (3 * NZ * NY + COL_TPB - 1) / COL_TPB;
        auto ldm = [&](int idx) { const int l = idx / NY, y = idx - l * NY; return gload(dp.Mh, ((size_t)l * nxh + kx) * NY + y); };
        auto ldw = [&](int idx) {
            const int l = idx / NY, y = idx - l * NY;     // l = b*NZ + o
            const int b = l / NZ, o = l - b * NZ;
            const int comp = b == 0 ? c0 : (b == 1 ? c1 : c2);
            return gload(dp.Tw, ((size_t)(comp * NZ + o) * nxh + kx) * NY + y);
        };
        float2 twv[TWI];
#pragma unroll
        for (int k = 0; k < TWI; ++k)
            if (NY % COL_TPB == 0 || (int)threadIdx.x + k * COL_TPB < NY) twv[k] = gload(dp.twy, threadIdx.x + k * COL_TPB);
        if constexpr (MI + WI <= 32) {
            float2 mv[MI];
            float wv[WI];
#pragma unroll
            for (int k = 0; k < MI; ++k) {
                const int idx = (int)threadIdx.x + k * COL_TPB;
                if ((lines * NY) % COL_TPB == 0 || idx < lines * NY) mv[k] = ldm(idx);
            }
#pragma unroll
            for (int k = 0; k < WI; ++k) {
                const int idx = (int)threadIdx.x + k * COL_TPB;
                if ((3 * NZ * NY) % COL_TPB == 0 || idx < 3 * NZ * NY) wv[k] = ldw(idx);
            }
#pragma unroll
            for (int k = 0; k < MI; ++k) {
                const int idx = (int)threadIdx.x + k * COL_TPB;
                if ((lines * NY) % COL_TPB == 0 || idx < lines * NY) A[idx] = mv[k];
            }
#pragma unroll
            for (int k = 0; k < WI; ++k) {
                const int idx = (int)threadIdx.x + k * COL_TPB;
                if ((3 * NZ * NY) % COL_TPB == 0 || idx < 3 * NZ * NY) wl[idx] = wv[k];
            }
        } else {
            batched_copy<lines * NY, COL_TPB, 16>(ldm, [&](int idx, float2 v) { A[idx] = v; });
            batched_copy<3 * NZ * NY, COL_TPB, 16>(ldw, [&](int idx, float v) { wl[idx] = v; });
        }
#pragma unroll
        for (int k = 0; k < TWI; ++k)
            if (NY % COL_TPB == 0 || (int)threadIdx.x + k * COL_TPB < NY) twl[threadIdx.x + k * COL_TPB] = twv[k];
        tw = twl;
    } else {
        for (int idx = threadIdx.x; idx < lines * ny; idx += COL_TPB) {
            const int l = idx / ny, y = idx - l * ny;
            A[idx] = gload(dp.Mh, ((size_t)l * nxh + kx) * ny + y);
        }
    }
    __syncthreads();
    float2* R;
    if constexpr (NY > 0) {
        fft_ip<NY, lines, COL_TPB, -1>(A, tw);
        R = A;
    }
    else R = lds_fft<-1>(A, B, lines, dp.py, tw);
    for (int ky = threadIdx.x; ky < ny; ky += COL_TPB) {
        float w[3][NZ];
        if constexpr (NY > 0) {
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int o = 0; o < NZ; ++o) w[b][o] = wl[(b * NZ + o) * NY + ky];
        } else {
#pragma unroll
            for (int o = 0; o < NZ; ++o) {
                w[0][o] = gload(dp.Tw, ((size_t)(c0 * NZ + o) * nxh + kx) * ny + ky);
                w[1][o] = gload(dp.Tw, ((size_t)(c1 * NZ + o) * nxh + kx) * ny + ky);
                w[2][o] = gload(dp.Tw, ((size_t)(c2 * NZ + o) * nxh + kx) * ny + ky);
            }
        }
        // source layer outer, target layer inner: one source's three components
        // live at a time next to the NZ accumulators
        float2 acc[NZ];
#pragma unroll
        for (int t = 0; t < NZ; ++t) acc[t] = make_float2(0.f, 0.f);
#pragma unroll
        for (int s = 0; s < NZ; ++s) {
            float2 m[3];
#pragma unroll
            for (int b = 0; b < 3; ++b) m[b] = R[(b * NZ + s) * ny + ky];
#pragma unroll
            for (int t = 0; t < NZ; ++t) {
                const int o = s >= t ? s - t : t - s;
                const float sg = s < t ? -1.0f : 1.0f;
                // real components multiply, imaginary (XZ, YZ) ones give i*w*m = w*(-m.y, m.x)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const bool odd = b == 0 ? odd0 : (b == 1 ? odd1 : odd2);
                    if (odd) {
                        const float wv = sg * w[b][o];
                        acc[t].x -= wv * m[b].y;
                        acc[t].y += wv * m[b].x;
                    } else {
                        const float wv = w[b][o];
                        acc[t].x += wv * m[b].x;
                        acc[t].y += wv * m[b].y;
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NZ; ++t) R[(oa * NZ + t) * ny + ky] = acc[t];
    }
    __syncthreads();
    float2* Rs = R + (size_t)oa * NZ * ny;
    float2* Os = (R == A ? B : A) + (size_t)oa * NZ * ny;
    float2* O;
    if constexpr (NY > 0) {
        fft_ip<NY, NZ, COL_TPB, 1>(Rs, tw);
        O = Rs;
    }
    else O = lds_fft<1>(Rs, Os, NZ, dp.py, tw);
    for (int idx = threadIdx.x; idx < NZ * ny; idx += COL_TPB) {
        const int l = idx / ny, y = idx - l * ny;
        gstore(dp.Hh, ((size_t)(oa * NZ + l) * nxh + kx) * ny + y, O[idx]);
    }
}

template <int NY, int NZ>
__global__ void __launch_bounds__(col_tpb(NY)) k_cols(DipoleDev dp)
{
    cols_body<NY, NZ>(dp, xcd_remap(blockIdx.x, gridDim.x));
}

// films of one shape in one launch: `per` workgroups each, film r's at [r per, (r+1) per)
template <int NY, int NZ>
__global__ void __launch_bounds__(col_tpb(NY)) k_cols_batch(const DipoleDev* __restrict__ dps, int per)
{
    const int w = xcd_remap(blockIdx.x, gridDim.x), r = w / per;
    cols_body<NY, NZ>(dps[r], w - r * per);
}

// ------------------------------------------------------------ split column pass
// With a split y transform (dp.split = R, ny = R ly) the column work of one kx falls
// into R independent problems, k1 = 0..R-1: the 3 nz lines b = 0..ly-1 (stride R in
// the spectrum) transform forward to k2 (ky = k1 + R k2), take the per-k
// layer/component tensor product for all three output components, and transform
// back.  Workgroup (kx, k1): nxh R workgroups, no transform done twice.
//
// LDS holds the lines position-major, element b of line l at A[b PS + l] with an odd
// pitch PS >= 3 nz: the transform stages give consecutive lanes consecutive lines and
// the product gives them consecutive k2, so neither meets a bank conflict.
//
// Product lanes: wave w forms output component w / (ly / 64) (one wave per component
// below ly = 64, lanes from ly on idle), so the branch on the component's odd words
// (XZ, YZ: the third word of x and y, the first two of z) is wave-uniform.  From
// ly = 64 on, with an even nz, the destination layers are split in two halves over
// twice the lanes (split_halves): twice the waves per workgroup take the latency of
// the loads and stores, and each lane's product is half as long.
__host__ __device__ constexpr int split_pitch(int nz) { return (3 * nz) | 1; }
__host__ __device__ constexpr int split_halves(int L, int nz) { return (L >= 64 && nz % 2 == 0 && 3 * L <= 512) ? 2 : 1; }
__host__ __device__ constexpr int split_tpb(int L, int nz) { return (L < 64 ? 192 : 3 * L) * split_halves(L, nz); }

// One in-place Stockham stage over LN position-major lines (pitch PS), NT threads
// with the line index fastest across lanes.
template <int L, int LN, int PS, int NT, int SIGN, int NS, int R>
__device__ __forceinline__ void split_stage(float2* __restrict__ A, const float2* __restrict__ tw)
{
    constexpr int nb = L / R, total = LN * nb, iters = (total + NT - 1) / NT, tstep = L / (NS * R);
    float2 v[iters][R];
#pragma unroll
    for (int it = 0; it < iters; ++it) {
        const int b = (int)threadIdx.x + it * NT;
        if (total % NT == 0 || b < total) {
            const int l = b % LN, j = b / LN;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = A[(j + r * nb) * PS + l];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < iters; ++it) {
        const int b = (int)threadIdx.x + it * NT;
        if (total % NT == 0 || b < total) {
            const int l = b % LN, j = b / LN, k = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 w = tw[k * r * tstep];
                    v[it][r] = SIGN < 0 ? cmul(v[it][r], w) : cmulc(v[it][r], w);
                }
            }
            dft_r<R, SIGN>(v[it]);
            const int o = (j / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) A[(o + q * NS) * PS + l] = v[it][q];
        }
    }
    __syncthreads();
}

template <int L, int LN, int PS, int NT, int SIGN, int NS>
struct SplitFFT {
    static __device__ __forceinline__ void run(float2* A, const float2* tw)
    {
        constexpr int R = (L / NS) % 8 == 0 ? 8 : ((L / NS) % 4 == 0 ? 4 : 2);
        split_stage<L, LN, PS, NT, SIGN, NS, R>(A, tw);
        SplitFFT<L, LN, PS, NT, SIGN, NS * R>::run(A, tw);
    }
};
template <int L, int LN, int PS, int NT, int SIGN>
struct SplitFFT<L, LN, PS, NT, SIGN, L> {
    static __device__ __forceinline__ void run(float2*, const float2*) {}
};

template <int L, int NZ>
__device__ __forceinline__ void split_cols_body(const DipoleDev& dp, int wg)
{
    constexpr int NT = split_tpb(L, NZ), LN = 3 * NZ, PS = split_pitch(NZ);
    constexpr int HV = split_halves(L, NZ), TN = NZ / HV;    // destination layers per lane
    extern __shared__ float2 lds[];
    const int R = dp.split, ny = dp.ny, nxh = dp.nxh;
    const int kx = wg / R, k1 = wg - kx * R;
    float2* A = lds;                                  // A[b PS + l]
    float2* twl = A + L * PS;                         // exp(-2 pi i k / L)
    float* wl = reinterpret_cast<float*>(twl + L);    // words [6][NZ][L]
    const size_t lst = (size_t)nxh * ny;              // spectrum stride between lines
    const float2* __restrict__ src = dp.Mh + (size_t)kx * ny + k1;
    {
        constexpr int MN = LN * L, MI = (MN + NT - 1) / NT;
        constexpr int WN = 6 * NZ * L / 4, WI = (WN + NT - 1) / NT;
        const float4* __restrict__ wsrc = reinterpret_cast<const float4*>(dp.Tw + (size_t)wg * (6 * NZ * L));
        float2 mv[MI];
        float4 wv[WI];
        float2 tv = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k < MI; ++k) {
            const int idx = (int)threadIdx.x + k * NT;
            if (MN % NT == 0 || idx < MN) {
                const int l = idx / L, b = idx - l * L;
                mv[k] = gload(src, l * lst + (size_t)b * R);
            }
        }
#pragma unroll
        for (int k = 0; k < WI; ++k) {
            const int idx = (int)threadIdx.x + k * NT;
            if (WN % NT == 0 || idx < WN) wv[k] = gload(wsrc, idx);
        }
        if ((int)threadIdx.x < L) tv = gload(dp.twy, threadIdx.x * R);
#pragma unroll
        for (int k = 0; k < MI; ++k) {
            const int idx = (int)threadIdx.x + k * NT;
            if (MN % NT == 0 || idx < MN) {
                const int l = idx / L, b = idx - l * L;
                A[b * PS + l] = mv[k];
            }
        }
#pragma unroll
        for (int k = 0; k < WI; ++k) {
            const int idx = (int)threadIdx.x + k * NT;
            if (WN % NT == 0 || idx < WN) reinterpret_cast<float4*>(wl)[idx] = wv[k];
        }
        if ((int)threadIdx.x < L) twl[threadIdx.x] = tv;
    }
    __syncthreads();
    SplitFFT<L, LN, PS, NT, -1, 1>::run(A, twl);
    {
        // lane -> (output component oa, k2); zpart: the wave of component z
        int oa, k2, t0 = 0;
        bool active;
        if constexpr (L < 64) {
            oa = (int)threadIdx.x >> 6;
            k2 = (int)threadIdx.x & 63;
            active = k2 < L;
        } else {
            const int h = (int)threadIdx.x / (3 * L), rem = (int)threadIdx.x - h * 3 * L;
            oa = rem / L;
            k2 = rem - oa * L;
            t0 = h * TN;                // wave-uniform: 3 L is a multiple of 64
            active = true;
        }
        const bool zpart = __builtin_amdgcn_readfirstlane(oa) == 2;
        const int c0 = oa == 0 ? 0 : (oa == 1 ? 1 : 2);
        const int c1 = oa == 0 ? 1 : (oa == 1 ? 3 : 4);
        const int c2 = oa == 0 ? 2 : (oa == 1 ? 4 : 5);
        float2 acc[TN];
#pragma unroll
        for (int t = 0; t < TN; ++t) acc[t] = make_float2(0.f, 0.f);
        // destination layers T0 .. T0 + TN - 1 (compile-time, so every sign folds)
        auto product = [&](auto t0c) {
            constexpr int T0 = decltype(t0c)::value;
            float w[3][NZ];
#pragma unroll
            for (int o = 0; o < NZ; ++o) {
                w[0][o] = wl[(c0 * NZ + o) * L + k2];
                w[1][o] = wl[(c1 * NZ + o) * L + k2];
                w[2][o] = wl[(c2 * NZ + o) * L + k2];
            }
            const float2* Ak = A + k2 * PS;
            // odd words: (x, y) outputs take XZ / YZ third; z takes them first and second
            auto mac = [&](float2& a, float wv, float2 m, bool odd, float sg) {
                if (odd) { const float u = sg * wv; a.x -= u * m.y; a.y += u * m.x; }
                else { a.x += wv * m.x; a.y += wv * m.y; }
            };
#pragma unroll
            for (int s = 0; s < NZ; ++s) {
                float2 m[3];
#pragma unroll
                for (int b = 0; b < 3; ++b) m[b] = Ak[b * NZ + s];
#pragma unroll
                for (int tt = 0; tt < TN; ++tt) {
                    const int t = T0 + tt;
                    const int o = s >= t ? s - t : t - s;
                    const float sg = s < t ? -1.0f : 1.0f;
                    if (zpart) {
                        mac(acc[tt], w[0][o], m[0], true, sg);
                        mac(acc[tt], w[1][o], m[1], true, sg);
                        mac(acc[tt], w[2][o], m[2], false, sg);
                    } else {
                        mac(acc[tt], w[0][o], m[0], false, sg);
                        mac(acc[tt], w[1][o], m[1], false, sg);
                        mac(acc[tt], w[2][o], m[2], true, sg);
                    }
                }
            }
        };
        if (active) {
            if constexpr (HV == 1) product(std::integral_constant<int, 0>{});
            else if (t0 == 0) product(std::integral_constant<int, 0>{});
            else product(std::integral_constant<int, TN>{});
        }
        __syncthreads();    // every lane has read its sources: the products go in place
        if (active) {
#pragma unroll
            for (int tt = 0; tt < TN; ++tt) A[k2 * PS + oa * NZ + t0 + tt] = acc[tt];
        }
    }
    __syncthreads();
    SplitFFT<L, LN, PS, NT, 1, 1>::run(A, twl);
    float2* __restrict__ dst = dp.Hh + (size_t)kx * ny + k1;
    for (int idx = threadIdx.x; idx < LN * L; idx += NT) {
        const int l = idx / L, b = idx - l * L;
        gstore(dst, l * lst + (size_t)b * R, A[b * PS + l]);
    }
}

template <int L, int NZ>
__global__ void __launch_bounds__(split_tpb(L, NZ)) k_cols_split(DipoleDev dp)
{
    split_cols_body<L, NZ>(dp, xcd_remap(blockIdx.x, gridDim.x));
}

// films of one shape in one launch: `per` = nxh R workgroups each
template <int L, int NZ>
__global__ void __launch_bounds__(split_tpb(L, NZ)) k_cols_split_batch(const DipoleDev* __restrict__ dps, int per)
{
    const int w = xcd_remap(blockIdx.x, gridDim.x), r = w / per;
    split_cols_body<L, NZ>(dps[r], w - r * per);
}

// ------------------------------------------------------------ per-layer-pair column pass
// LongRange2D::apply (cpu/longrange2d/spinoperationlongrange2d.cpp:448-531): a full complex
// tensor for every ordered layer pair and every component,
//   H^_a[d] = dsc[d] sum_s ssc[s] sum_b T^_ab[s][d] M^_b[s],
// where the reference reads the pair stored under the labels [s][d] (its tensorArray(s, d))
// for destination layer d and source layer s; the words are stored [9][nz][nz][nxh][ny]
// (XX XY XZ YX YY YZ ZX ZY ZZ, first label, second label) with 1/(nx ny) folded in.
// ssc / dsc are the per-layer source-moment and destination-field scales, the operator's
// own scale folded into dsc.  Workgroup (kx, a) transforms all 3 nz columns of kx,
// forms output component a for every layer and transforms those nz columns back.
template <int NY, int NZ>
__device__ __forceinline__ void pair_cols_body(const DipoleDev& dp, int wg)
{
    constexpr int COL_TPB = col_tpb(NY);
    extern __shared__ float2 lds[];
    const int kx = wg / 3;
    const int oa = wg - 3 * kx;
    const int ny = NY ? NY : dp.ny, nxh = dp.nxh;
    constexpr int lines = 3 * NZ;
    float2* A = lds;
    float2* B = lds + (size_t)lines * ny;
    const float2* tw = dp.twy;
    if constexpr (NY > 0) {
        float2* twl = A + (size_t)lines * NY;   // in place: the twiddles follow the lines
        for (int k = threadIdx.x; k < NY; k += COL_TPB) twl[k] = gload(dp.twy, k);
        batched_copy<lines * NY, COL_TPB, 16>(
            [&](int idx) { const int l = idx / NY, y = idx - l * NY; return gload(dp.Mh, ((size_t)l * nxh + kx) * NY + y); },
            [&](int idx, float2 v) { A[idx] = v; });
        tw = twl;
    } else {
        for (int idx = threadIdx.x; idx < lines * ny; idx += COL_TPB) {
            const int l = idx / ny, y = idx - l * ny;
            A[idx] = gload(dp.Mh, ((size_t)l * nxh + kx) * ny + y);
        }
    }
    __syncthreads();
    float2* R;
    if constexpr (NY > 0) {
        fft_ip<NY, lines, COL_TPB, -1>(A, tw);
        R = A;
    }
    else R = lds_fft<-1>(A, B, lines, dp.py, tw);
    float ssc[NZ], dsc[NZ];
#pragma unroll
    for (int l = 0; l < NZ; ++l) {
        ssc[l] = dp.lsc[l];
        dsc[l] = dp.lsc[MSD_MAX_LAYERS + l];
    }
    const size_t plane = (size_t)nxh * ny;
    for (int ky = threadIdx.x; ky < ny; ky += COL_TPB) {
        const float2* T = dp.Tp + (size_t)kx * ny + ky;
        float2 acc[NZ];
#pragma unroll
        for (int d = 0; d < NZ; ++d) acc[d] = make_float2(0.f, 0.f);
#pragma unroll
        for (int s = 0; s < NZ; ++s) {
            float2 m[3];
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const float2 v = R[(b * NZ + s) * ny + ky];
                m[b] = make_float2(ssc[s] * v.x, ssc[s] * v.y);
            }
#pragma unroll
            for (int d = 0; d < NZ; ++d)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const float2 t = T[((size_t)((oa * 3 + b) * NZ + s) * NZ + d) * plane];
                    acc[d].x += t.x * m[b].x - t.y * m[b].y;
                    acc[d].y += t.x * m[b].y + t.y * m[b].x;
                }
        }
#pragma unroll
        for (int d = 0; d < NZ; ++d) R[(oa * NZ + d) * ny + ky] = make_float2(dsc[d] * acc[d].x, dsc[d] * acc[d].y);
    }
    __syncthreads();
    float2* Rs = R + (size_t)oa * NZ * ny;
    float2* Os = (R == A ? B : A) + (size_t)oa * NZ * ny;
    float2* O;
    if constexpr (NY > 0) {
        fft_ip<NY, NZ, COL_TPB, 1>(Rs, tw);
        O = Rs;
    }
    else O = lds_fft<1>(Rs, Os, NZ, dp.py, tw);
    for (int idx = threadIdx.x; idx < NZ * ny; idx += COL_TPB) {
        const int l = idx / ny, y = idx - l * ny;
        gstore(dp.Hh, ((size_t)(oa * NZ + l) * nxh + kx) * ny + y, O[idx]);
    }
}

template <int NY, int NZ>
__global__ void __launch_bounds__(col_tpb(NY)) k_cols_pair(DipoleDev dp)
{
    pair_cols_body<NY, NZ>(dp, xcd_remap(blockIdx.x, gridDim.x));
}

// films of one shape with per-layer-pair tensors in one launch (layout as k_cols_batch)
template <int NY, int NZ>
__global__ void __launch_bounds__(col_tpb(NY)) k_cols_pair_batch(const DipoleDev* __restrict__ dps, int per)
{
    const int w = xcd_remap(blockIdx.x, gridDim.x), r = w / per;
    pair_cols_body<NY, NZ>(dps[r], w - r * per);
}

// ------------------------------------------------------------ fused Heun
// Site loop of the fused kernels.  Compile-time shapes give each thread SPT
// sites, software-pipelined: site 0's loads are issued before the inverse x
// transform (their latency hides behind it), and site k+1's before site k is
// finished.  Runtime shapes gather and finish each site in turn.
template <int NX, int TY, int TPB, bool DIP, bool PRED, bool COMP>
__device__ __forceinline__ void fused_sites(const StepArgs& a, const ExchangeDev& exl, float2* A, float2* B, float2*& R,
                                            const float2* tw, int nx, int ny, int ty, const RowTile& t, uint64_t draw)
{
    const DipoleDev& dp = a.dp;
    auto site_of = [&](int li, int& r, int& x) {
        r = li / nx;
        x = li - r * nx;
        return x + nx * (t.y0 + t.ys * r + ny * t.z);
    };
    auto gather = [&](int i, SiteGather& g) {
        if constexpr (PRED) pred_gather<COMP>(a, exl, i, g);
        else corr_gather<COMP>(a, exl, i, g);
    };
    auto finish = [&](int li, const SiteGather& g) {
        int r, x;
        const int i = site_of(li, r, x);
        float3 Hd = make_float3(0.f, 0.f, 0.f);
        if constexpr (DIP) Hd = make_float3(rowval(R, r, x, nx), rowval(R, ty + r, x, nx), rowval(R, 2 * ty + r, x, nx));
        float3 v;
        if constexpr (PRED) v = pred_finish<COMP>(a, exl, i, Hd, g);
        else v = corr_finish<COMP>(a, exl, i, Hd, draw, g);
        if constexpr (DIP) {
            rowval(R, r, x, nx) = v.x;
            rowval(R, ty + r, x, nx) = v.y;
            rowval(R, 2 * ty + r, x, nx) = v.z;
        }
    };
    if constexpr (NX > 0) {
        constexpr int SITES = TY * NX;
        constexpr int SPT = (SITES + TPB - 1) / TPB;
        SiteGather g[2];
        int r, x;
        const int li0 = (int)threadIdx.x;
        if (SITES % TPB == 0 || li0 < SITES) gather(site_of(li0, r, x), g[0]);
        if constexpr (DIP) R = row_fft<NX, TY, TPB, 1>(A, B, dp, tw);
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int li = (int)threadIdx.x + k * TPB;
            const int ln = li + TPB;
            if (k + 1 < SPT && (SITES % TPB == 0 || ln < SITES)) gather(site_of(ln, r, x), g[(k + 1) & 1]);
            if (SITES % TPB == 0 || li < SITES) finish(li, g[k & 1]);
        }
    } else {
        if constexpr (DIP) R = row_fft<NX, TY, TPB, 1>(A, B, dp, tw);
        for (int li = threadIdx.x; li < ty * nx; li += TPB) {
            int r, x;
            SiteGather g;
            gather(site_of(li, r, x), g);
            finish(li, g);
        }
    }
}

// Predictor half of the step closure in common/scripts/PredictorCorrector.lua:77-84:
//   H(S) = Exchange + Anisotropy + Dipole,  dM = LLG(S, H(S)),
//   P = N(S + dt dM),  C = N(S + dt/2 dM),
// then the x-transform of P for the next column pass.  COMP carries S and C as
// compensated FP32 pairs (hi + lo).
// tile: the workgroup's tile among its film's nz (ny/ty), XCD-remapped layer fastest
// (an XCD's run of tiles is a band of rows through every layer, so the z
// neighbours of its sites are mostly in its own L2)
template <int NX, int TY, int TPB, bool DIP, bool COMP>
__device__ __forceinline__ void step_pred_body(const StepArgs& a, int tile)
{
    extern __shared__ float2 lds[];
    const DipoleDev& dp = a.dp;
    const int nx = NX ? NX : a.d.nx, ty = TY ? TY : a.ty, lines = row_lines(ty);
    const int ny = a.d.ny;
    const RowTile t = row_tile(dp, tile, a.d.nz, ty);
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    float2* R = A;
    const float2* tw = nullptr;
    __shared__ int s_start[EX_LDS_CLS];
    __shared__ ExEntry s_ent[EX_LDS_ENT];
    const ExchangeDev exl = stage_exchange(a.ex, s_start, s_ent);
    if constexpr (DIP) {
        tw = row_twiddles<NX>(dp, A + (size_t)lines * nx);
        rows_in<TPB, NX, TY>(dp, dp.Hh, nx, ny, a.d.nz, ty, t, A);
    }
    __syncthreads();
    if (a.th.on && tile == 0 && threadIdx.x == 0) *a.draw += 1ull;
    fused_sites<NX, TY, TPB, DIP, true, COMP>(a, exl, A, B, R, tw, nx, ny, ty, t, 0ull);
    if constexpr (DIP) {
        __syncthreads();
        const float2* F = row_fft<NX, TY, TPB, -1>(R, R == A ? B : A, dp, tw);
        rows_out<TPB, NX, TY>(dp, F, nx, ny, a.d.nz, ty, t, dp.Mh);
    }
}

// Corrector half (PredictorCorrector.lua:81-85) and the thermal kick (:98-103):
//   S' = N(C + dt/2 LLG(P, H(P))),  S'' = N(S' + dt LLG(S', H_th)) when Thermal is present,
// then the x-transform of S'' for the next step's first column pass.
// tile: the workgroup's tile among its film's nz (ny/ty), XCD-remapped layer fastest
// (an XCD's run of tiles is a band of rows through every layer, so the z
// neighbours of its sites are mostly in its own L2)
template <int NX, int TY, int TPB, bool DIP, bool COMP>
__device__ __forceinline__ void step_corr_body(const StepArgs& a, int tile)
{
    extern __shared__ float2 lds[];
    const DipoleDev& dp = a.dp;
    const int nx = NX ? NX : a.d.nx, ty = TY ? TY : a.ty, lines = row_lines(ty);
    const int ny = a.d.ny;
    const RowTile t = row_tile(dp, tile, a.d.nz, ty);
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    float2* R = A;
    const float2* tw = nullptr;
    __shared__ int s_start[EX_LDS_CLS];
    __shared__ ExEntry s_ent[EX_LDS_ENT];
    const ExchangeDev exl = stage_exchange(a.ex, s_start, s_ent);
    if constexpr (DIP) {
        tw = row_twiddles<NX>(dp, A + (size_t)lines * nx);
        rows_in<TPB, NX, TY>(dp, dp.Hh, nx, ny, a.d.nz, ty, t, A);
    }
    __syncthreads();
    const uint64_t draw = a.th.on ? (uint64_t)*a.draw : 0ull;
    fused_sites<NX, TY, TPB, DIP, false, COMP>(a, exl, A, B, R, tw, nx, ny, ty, t, draw);
    if constexpr (DIP) {
        __syncthreads();
        const float2* F = row_fft<NX, TY, TPB, -1>(R, R == A ? B : A, dp, tw);
        rows_out<TPB, NX, TY>(dp, F, nx, ny, a.d.nz, ty, t, dp.Mh);
    }
}

template <int NX, int TY, int TPB, bool DIP, bool COMP>
__global__ void __launch_bounds__(TPB, row_waves(TPB)) k_step_pred(StepArgs a)
{
    step_pred_body<NX, TY, TPB, DIP, COMP>(a, xcd_remap(blockIdx.x, gridDim.x));
}

template <int NX, int TY, int TPB, bool DIP, bool COMP>
__global__ void __launch_bounds__(TPB, row_waves(TPB)) k_step_corr(StepArgs a)
{
    step_corr_body<NX, TY, TPB, DIP, COMP>(a, xcd_remap(blockIdx.x, gridDim.x));
}

// ------------------------------------------------------ Runge-Kutta stages
// Stage i of make_rk_step_function (common/scripts/RungeKutta.lua:185-214) on one row
// tile: the inverse x transform of the stage's dipole spectrum, the total field at
// the stage spins X (exchange, anisotropy, dipole, applied fields), then the row
// out = N(...N(N(S + f0 dt F(S)) + f1 dt F(k_1))... + f_i dt F(X)), each term an
// LLGCartesian sub-step (llgcartesian.cpp:90-278: a zero coefficient copies, every
// sub-step renormalises to |S|), the thermal kick at the last stage, and the forward
// x transform of the result for the next stage's column pass.
template <int NX, int TY, int TPB, bool DIP>
__global__ void __launch_bounds__(TPB) k_rk_stage(StepArgs a, RKStage r)
{
    extern __shared__ float2 lds[];
    const DipoleDev& dp = a.dp;
    const int nx = NX ? NX : a.d.nx, ty = TY ? TY : a.ty, lines = row_lines(ty);
    const int ny = a.d.ny;
    const RowTile t = row_tile(dp, xcd_remap(blockIdx.x, gridDim.x), a.d.nz, ty);
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    float2* R = A;
    const float2* tw = nullptr;
    __shared__ int s_start[EX_LDS_CLS];
    __shared__ ExEntry s_ent[EX_LDS_ENT];
    const ExchangeDev exl = stage_exchange(a.ex, s_start, s_ent);
    if constexpr (DIP) {
        tw = row_twiddles<NX>(dp, A + (size_t)lines * nx);
        rows_in<TPB, NX, TY>(dp, dp.Hh, nx, ny, a.d.nz, ty, t, A);
    }
    __syncthreads();
    if constexpr (DIP) R = row_fft<NX, TY, TPB, 1>(A, B, dp, tw);
    const uint64_t draw = (r.last && a.th.on) ? (uint64_t)*a.draw : 0ull;
    const uint32_t n = (uint32_t)a.d.n;
    const float* __restrict__ msa = a.S + 3 * (size_t)n;
    const float dt = a.dt;
    const float3 zero = make_float3(0.f, 0.f, 0.f);
    for (int li = threadIdx.x; li < ty * nx; li += TPB) {
        const int rr = li / nx, x = li - rr * nx;
        const uint32_t i = (uint32_t)(x + nx * (t.y0 + t.ys * rr + ny * t.z));
        const float ms = ldg(msa, i);
        const float3 X = make_float3(ldg(r.X, i), ldg(r.X + n, i), ldg(r.X + 2 * (size_t)n, i));
        ExGather g;
        exchange_gather(exl, n, i, r.X, msa, g);
        float3 H = exchange_finish(exl, n, i, r.X, msa, ms, g);
        const float3 ha = aniso_field(a.an, (int)i, (int)n, X, ms);
        H.x += ha.x; H.y += ha.y; H.z += ha.z;
        if constexpr (DIP) {
            H.x += rowval(R, rr, x, nx); H.y += rowval(R, ty + rr, x, nx); H.z += rowval(R, 2 * ty + rr, x, nx);
        }
        if (a.uni.ops) { const float3 hu = uniaxial_field(a.uni, (int)i, X, ms); H.x += hu.x; H.y += hu.y; H.z += hu.z; }
        if (a.has_app) {
            H.x += a.Bapp.x; H.y += a.Bapp.y; H.z += a.Bapp.z;
            if (a.Happ) { H.x += ldg(a.Happ, i); H.y += ldg(a.Happ + n, i); H.z += ldg(a.Happ + 2 * (size_t)n, i); }
        }
        if (r.Hout) { stg(r.Hout, i, H.x); stg(r.Hout + n, i, H.y); stg(r.Hout + 2 * (size_t)n, i, H.z); }
        const float al = a.lp.site_alpha ? ldg(a.lp.site_alpha, i) : a.lp.alpha;
        const float ga = a.lp.site_gamma ? ldg(a.lp.site_gamma, i) : a.lp.gamma;
        float3 d = make_float3(ldg(a.S, i), ldg(a.S + n, i), ldg(a.S + 2 * (size_t)n, i));
        for (int j = 0; j < r.terms; ++j) {
            const float c = r.f[j];
            if (j > 0 && c == 0.0f) continue;        // the script skips zero coefficients after the first
            if (c != 0.0f && ms > 0.0f) {
                const bool cur = j == r.terms - 1;
                const float3 M = cur ? X : make_float3(ldg(r.K[j], i), ldg(r.K[j] + n, i), ldg(r.K[j] + 2 * (size_t)n, i));
                const float3 Hj = cur ? H : make_float3(ldg(r.H[j], i), ldg(r.H[j] + n, i), ldg(r.H[j] + 2 * (size_t)n, i));
                const float h = c * dt;
                if (a.lp.quaternion) {
                    // LLG:apply(d, c, K_j, d): d turned about the axis of K_j's field by c dt
                    float qw;
                    float3 qv;
                    quat_rot(a.lp, d, M, Hj, Hj, al, ga, ms, h, qw, qv);
                    const float3 dd = quat_delta(qw, qv, d);
                    d = make_float3(d.x + dd.x, d.y + dd.y, d.z + dd.z);
                } else {
                    const float3 dM = a.lp.thermal_first ? llg_rhs(M, Hj, Hj, al, ga, ms, a.lp.ps, a.lp.ds)
                                                         : llg_rhs(M, Hj, Hj, al, ga, ms, 1.0f, 1.0f);
                    d = make_float3(d.x + h * dM.x, d.y + h * dM.y, d.z + h * dM.z);
                }
            }
            if (a.lp.renorm) d = renorm(d, ms);
        }
        if (r.last && a.th.on) {
            const float sg = thermal_sigma(a.th, a.lp, (int)i, ms, dt);
            if (ms > 0.0f) {
                float3 Ht = zero;
                if (sg != 0.0f) {
                    const float3 gn = thermal_normals(a.th.seed, draw, (int)i);
                    Ht = make_float3(sg * gn.x, sg * gn.y, sg * gn.z);
                }
                if (a.lp.quaternion) {
                    float qw;
                    float3 qv;
                    quat_rot(a.lp, d, d, Ht, a.lp.thermal_first ? zero : Ht, al, ga, ms, dt, qw, qv);
                    const float3 dd = quat_delta(qw, qv, d);
                    d = make_float3(d.x + dd.x, d.y + dd.y, d.z + dd.z);
                } else {
                    const float3 dM = a.lp.thermal_first ? llg_rhs(d, Ht, zero, al, ga, ms, a.lp.ps, a.lp.ds)
                                                         : llg_rhs(d, Ht, Ht, al, ga, ms, 1.0f, 1.0f);
                    d = make_float3(d.x + dt * dM.x, d.y + dt * dM.y, d.z + dt * dM.z);
                }
            }
            if (a.lp.renorm) d = renorm(d, ms);
        }
        stg(r.out, i, d.x); stg(r.out + n, i, d.y); stg(r.out + 2 * (size_t)n, i, d.z);
        if constexpr (DIP) {
            rowval(R, rr, x, nx) = d.x;
            rowval(R, ty + rr, x, nx) = d.y;
            rowval(R, 2 * ty + rr, x, nx) = d.z;
        }
    }
    if constexpr (DIP) {
        __syncthreads();
        const float2* F = row_fft<NX, TY, TPB, -1>(R, R == A ? B : A, dp, tw);
        rows_out<TPB, NX, TY>(dp, F, nx, ny, a.d.nz, ty, t, dp.Mh);
    }
}

__global__ void k_draw_advance(unsigned long long* draw)
{
    if (threadIdx.x == 0) *draw += 1ull;
}

// films of one shape in one launch: film r's arguments at as[r], its `per` tiles at
// [r per, (r+1) per) of the remapped workgroup index
template <int NX, int TY, int TPB, bool COMP>
__global__ void __launch_bounds__(TPB, row_waves(TPB)) k_step_pred_batch(const StepArgs* __restrict__ as, int per)
{
    const int w = xcd_remap(blockIdx.x, gridDim.x), r = w / per;
    step_pred_body<NX, TY, TPB, true, COMP>(as[r], w - r * per);
}

template <int NX, int TY, int TPB, bool COMP>
__global__ void __launch_bounds__(TPB, row_waves(TPB)) k_step_corr_batch(const StepArgs* __restrict__ as, int per)
{
    const int w = xcd_remap(blockIdx.x, gridDim.x), r = w / per;
    step_corr_body<NX, TY, TPB, true, COMP>(as[r], w - r * per);
}

// ------------------------------------------------------------- dispatch
// Compile-time row shapes (nx, rows per tile, threads per workgroup); the first
// listed for an nx is its default, the others are alternatives a tuning run can
// select.  Column lengths likewise.
#define MSD_ROW_SHAPES(X)                                                           \
    X(32, 8, 256) X(64, 8, 256) X(128, 8, 512) X(256, 8, 1024) X(256, 4, 512)       \
    X(256, 8, 512) X(256, 2, 256) X(256, 4, 1024) X(512, 4, 1024) X(512, 2, 512)    \
    X(1024, 2, 1024)
#define MSD_COL_LENGTHS(X) X(32) X(64) X(128) X(256) X(512) X(1024)
// sub-column lengths ly = ny / R of the split column kernel
#define MSD_SPLIT_LENGTHS(X) X(8) X(16) X(32) X(64) X(128) X(256)

// whether the split column kernel is compiled for sub-columns of ly (R rows per tile)
bool split_length_compiled(int ly, int R)
{
    (void)R;
#define MSD_SL(L) if (ly == L) return true;
    MSD_SPLIT_LENGTHS(MSD_SL)
#undef MSD_SL
    return false;
}

// lines (ly x odd pitch), twiddles and the workgroup's 6 nz ly words
size_t split_cols_lds(int ly, int nz) { return (size_t)ly * split_pitch(nz) * 8 + (size_t)8 * ly + (size_t)24 * nz * ly; }

static int row_tiles(int ny, int nz, int ty) { return nz * (ny / ty); }

bool row_shape_compiled(int nx, int ty, int tpb)
{
#define MSD_HAS(NXV, TYV, TPBV) if (nx == NXV && ty == TYV && tpb == TPBV) return true;
    MSD_ROW_SHAPES(MSD_HAS)
#undef MSD_HAS
    return false;
}

int row_shapes_for(int nx, int ny, int* ty, int* tpb, int cap)
{
    int k = 0;
#define MSD_LIST(NXV, TYV, TPBV)                                                    \
    if (nx == NXV && ny % TYV == 0 && k < cap) { ty[k] = TYV; tpb[k] = TPBV; ++k; }
    MSD_ROW_SHAPES(MSD_LIST)
#undef MSD_LIST
    return k;
}

void row_shape_default(int nx, int ny, int* ty, int* tpb)
{
    *ty = 0;
#define MSD_DEF(NXV, TYV, TPBV)                                                     \
    if (*ty == 0 && nx == NXV && ny % TYV == 0) { *ty = TYV; *tpb = TPBV; }
    MSD_ROW_SHAPES(MSD_DEF)
#undef MSD_DEF
}

hipError_t launch_rows_fwd(const DipoleDev& dp, const float* s, hipStream_t st)
{
    const int g = row_tiles(dp.ny, dp.nz, dp.ty);
#define MSD_FWD(NXV, TYV, TPBV)                                                     \
    if (dp.nx == NXV && dp.ty == TYV && dp.tpb == TPBV) {                           \
        k_rows_fwd<NXV, TYV, TPBV><<<g, TPBV, dp.lds_rows, st>>>(dp, s);            \
        return hipGetLastError();                                                   \
    }
    MSD_ROW_SHAPES(MSD_FWD)
#undef MSD_FWD
    k_rows_fwd<0, 0, ROW_TPB><<<g, ROW_TPB, dp.lds_rows, st>>>(dp, s);
    return hipGetLastError();
}

hipError_t launch_rows_inv(const DipoleDev& dp, float* h, hipStream_t st)
{
    const int g = row_tiles(dp.ny, dp.nz, dp.ty);
#define MSD_INV(NXV, TYV, TPBV)                                                     \
    if (dp.nx == NXV && dp.ty == TYV && dp.tpb == TPBV) {                           \
        k_rows_inv<NXV, TYV, TPBV><<<g, TPBV, dp.lds_rows, st>>>(dp, h);            \
        return hipGetLastError();                                                   \
    }
    MSD_ROW_SHAPES(MSD_INV)
#undef MSD_INV
    k_rows_inv<0, 0, ROW_TPB><<<g, ROW_TPB, dp.lds_rows, st>>>(dp, h);
    return hipGetLastError();
}

template <int NZ>
static void launch_cols_nz(const DipoleDev& dp, hipStream_t st)
{
#define MSD_COL(NYV)                                                                \
    if (dp.ny == NYV) {                                                             \
        k_cols<NYV, NZ><<<3 * dp.nxh, col_tpb(NYV), dp.lds_cols, st>>>(dp);             \
        return;                                                                     \
    }
    MSD_COL_LENGTHS(MSD_COL)
#undef MSD_COL
    k_cols<0, NZ><<<3 * dp.nxh, col_tpb(0), dp.lds_cols, st>>>(dp);
}

template <int NZ>
static void launch_cols_pair_nz(const DipoleDev& dp, hipStream_t st)
{
#define MSD_COL(NYV)                                                                \
    if (dp.ny == NYV) {                                                             \
        k_cols_pair<NYV, NZ><<<3 * dp.nxh, col_tpb(NYV), dp.lds_cols, st>>>(dp);   \
        return;                                                                     \
    }
    MSD_COL_LENGTHS(MSD_COL)
#undef MSD_COL
    k_cols_pair<0, NZ><<<3 * dp.nxh, col_tpb(0), dp.lds_cols, st>>>(dp);
}

static hipError_t launch_cols_pair(const DipoleDev& dp, hipStream_t st)
{
    switch (dp.nz) {
        case 1: launch_cols_pair_nz<1>(dp, st); break;
        case 2: launch_cols_pair_nz<2>(dp, st); break;
        case 3: launch_cols_pair_nz<3>(dp, st); break;
        case 4: launch_cols_pair_nz<4>(dp, st); break;
        case 5: launch_cols_pair_nz<5>(dp, st); break;
        case 6: launch_cols_pair_nz<6>(dp, st); break;
        case 7: launch_cols_pair_nz<7>(dp, st); break;
        case 8: launch_cols_pair_nz<8>(dp, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int NZ>
static void launch_cols_split_nz(const DipoleDev& dp, hipStream_t st)
{
#define MSD_CS(L)                                                                   \
    if (dp.ly == L) {                                                               \
        k_cols_split<L, NZ><<<dp.nxh * dp.split, split_tpb(L, NZ), dp.lds_cols, st>>>(dp); \
        return;                                                                     \
    }
    MSD_SPLIT_LENGTHS(MSD_CS)
#undef MSD_CS
}

static hipError_t launch_cols_split(const DipoleDev& dp, hipStream_t st)
{
    if (!split_length_compiled(dp.ly, dp.split)) return hipErrorInvalidValue;
    switch (dp.nz) {
        case 1: launch_cols_split_nz<1>(dp, st); break;
        case 2: launch_cols_split_nz<2>(dp, st); break;
        case 3: launch_cols_split_nz<3>(dp, st); break;
        case 4: launch_cols_split_nz<4>(dp, st); break;
        case 5: launch_cols_split_nz<5>(dp, st); break;
        case 6: launch_cols_split_nz<6>(dp, st); break;
        case 7: launch_cols_split_nz<7>(dp, st); break;
        case 8: launch_cols_split_nz<8>(dp, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_cols(const DipoleDev& dp, hipStream_t st)
{
    if (dp.split) return launch_cols_split(dp, st);
    if (dp.Tp) return launch_cols_pair(dp, st);
    switch (dp.nz) {
        case 1: launch_cols_nz<1>(dp, st); break;
        case 2: launch_cols_nz<2>(dp, st); break;
        case 3: launch_cols_nz<3>(dp, st); break;
        case 4: launch_cols_nz<4>(dp, st); break;
        case 5: launch_cols_nz<5>(dp, st); break;
        case 6: launch_cols_nz<6>(dp, st); break;
        case 7: launch_cols_nz<7>(dp, st); break;
        case 8: launch_cols_nz<8>(dp, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// fn (optional) receives the kernel launched: a graph node of it can later take new
// argument values (hipGraphExecKernelNodeSetParams) without a new capture
template <bool PRED, int NX, int TY, int TPB, bool DIP, bool COMP>
static void launch_step_t(const StepArgs& a, int g, size_t lds, hipStream_t st, const void** fn)
{
    if (PRED) {
        k_step_pred<NX, TY, TPB, DIP, COMP><<<g, TPB, lds, st>>>(a);
        if (fn) *fn = (const void*)k_step_pred<NX, TY, TPB, DIP, COMP>;
    } else {
        k_step_corr<NX, TY, TPB, DIP, COMP><<<g, TPB, lds, st>>>(a);
        if (fn) *fn = (const void*)k_step_corr<NX, TY, TPB, DIP, COMP>;
    }
}

template <bool PRED>
static hipError_t launch_step(const StepArgs& a, hipStream_t st, const void** fn)
{
    const int g = row_tiles(a.d.ny, a.d.nz, a.ty);
    const bool comp = a.Slo != nullptr;
    if (!a.has_dip) {
        if (comp) launch_step_t<PRED, 0, 0, ROW_TPB, false, true>(a, g, 0, st, fn);
        else launch_step_t<PRED, 0, 0, ROW_TPB, false, false>(a, g, 0, st, fn);
        return hipGetLastError();
    }
    const size_t lds = a.dp.lds_rows;
#define MSD_STEP(NXV, TYV, TPBV)                                                    \
    if (a.d.nx == NXV && a.ty == TYV && a.dp.tpb == TPBV) {                         \
        if (comp) launch_step_t<PRED, NXV, TYV, TPBV, true, true>(a, g, lds, st, fn);  \
        else launch_step_t<PRED, NXV, TYV, TPBV, true, false>(a, g, lds, st, fn);      \
        return hipGetLastError();                                                   \
    }
    MSD_ROW_SHAPES(MSD_STEP)
#undef MSD_STEP
    if (comp) launch_step_t<PRED, 0, 0, ROW_TPB, true, true>(a, g, lds, st, fn);
    else launch_step_t<PRED, 0, 0, ROW_TPB, true, false>(a, g, lds, st, fn);
    return hipGetLastError();
}

hipError_t launch_step_pred(const StepArgs& a, hipStream_t st, const void** fn) { return launch_step<true>(a, st, fn); }
hipError_t launch_step_corr(const StepArgs& a, hipStream_t st, const void** fn) { return launch_step<false>(a, st, fn); }

// nrep films shaped like a0 (dimensions, row tiles, column length, state form),
// arguments as[] in device memory
template <bool PRED>
static hipError_t launch_step_batch(const StepArgs& a0, const StepArgs* as, int nrep, hipStream_t st)
{
    const int per = row_tiles(a0.d.ny, a0.d.nz, a0.ty), g = per * nrep;
    const bool comp = a0.Slo != nullptr;
    const size_t lds = a0.dp.lds_rows;
#define MSD_STEPB(NXV, TYV, TPBV)                                                   \
    if (a0.d.nx == NXV && a0.ty == TYV && a0.dp.tpb == TPBV) {                      \
        if (PRED) {                                                                 \
            if (comp) k_step_pred_batch<NXV, TYV, TPBV, true><<<g, TPBV, lds, st>>>(as, per);   \
            else k_step_pred_batch<NXV, TYV, TPBV, false><<<g, TPBV, lds, st>>>(as, per);       \
        } else {                                                                    \
            if (comp) k_step_corr_batch<NXV, TYV, TPBV, true><<<g, TPBV, lds, st>>>(as, per);   \
            else k_step_corr_batch<NXV, TYV, TPBV, false><<<g, TPBV, lds, st>>>(as, per);       \
        }                                                                           \
        return hipGetLastError();                                                   \
    }
    MSD_ROW_SHAPES(MSD_STEPB)
#undef MSD_STEPB
    if (PRED) {
        if (comp) k_step_pred_batch<0, 0, ROW_TPB, true><<<g, ROW_TPB, lds, st>>>(as, per);
        else k_step_pred_batch<0, 0, ROW_TPB, false><<<g, ROW_TPB, lds, st>>>(as, per);
    } else {
        if (comp) k_step_corr_batch<0, 0, ROW_TPB, true><<<g, ROW_TPB, lds, st>>>(as, per);
        else k_step_corr_batch<0, 0, ROW_TPB, false><<<g, ROW_TPB, lds, st>>>(as, per);
    }
    return hipGetLastError();
}

hipError_t launch_step_pred_batch(const StepArgs& a0, const StepArgs* as, int nrep, hipStream_t st)
{
    return launch_step_batch<true>(a0, as, nrep, st);
}
hipError_t launch_step_corr_batch(const StepArgs& a0, const StepArgs* as, int nrep, hipStream_t st)
{
    return launch_step_batch<false>(a0, as, nrep, st);
}

hipError_t launch_rk_stage(const StepArgs& a, const RKStage& r, hipStream_t st)
{
    const int g = row_tiles(a.d.ny, a.d.nz, a.ty);
    if (!a.has_dip) {
        k_rk_stage<0, 0, ROW_TPB, false><<<g, ROW_TPB, 0, st>>>(a, r);
        return hipGetLastError();
    }
    const size_t lds = a.dp.lds_rows;
#define MSD_RK(NXV, TYV, TPBV)                                                      \
    if (a.d.nx == NXV && a.ty == TYV && a.dp.tpb == TPBV) {                         \
        k_rk_stage<NXV, TYV, TPBV, true><<<g, TPBV, lds, st>>>(a, r);               \
        return hipGetLastError();                                                   \
    }
    MSD_ROW_SHAPES(MSD_RK)
#undef MSD_RK
    k_rk_stage<0, 0, ROW_TPB, true><<<g, ROW_TPB, lds, st>>>(a, r);
    return hipGetLastError();
}

hipError_t launch_draw_advance(unsigned long long* draw, hipStream_t st)
{
    k_draw_advance<<<1, 64, 0, st>>>(draw);
    return hipGetLastError();
}

// every member of a batch has the form of d0 (lattice-sum words or per-pair tensors)
template <int NZ>
static void launch_cols_batch_nz(const DipoleDev& d0, const DipoleDev* dps, int nrep, hipStream_t st)
{
    const int per = 3 * d0.nxh;
#define MSD_COLB(NYV)                                                               \
    if (d0.ny == NYV) {                                                             \
        if (d0.Tp) k_cols_pair_batch<NYV, NZ><<<per * nrep, col_tpb(NYV), d0.lds_cols, st>>>(dps, per); \
        else k_cols_batch<NYV, NZ><<<per * nrep, col_tpb(NYV), d0.lds_cols, st>>>(dps, per); \
        return;                                                                     \
    }
    MSD_COL_LENGTHS(MSD_COLB)
#undef MSD_COLB
    if (d0.Tp) k_cols_pair_batch<0, NZ><<<per * nrep, col_tpb(0), d0.lds_cols, st>>>(dps, per);
    else k_cols_batch<0, NZ><<<per * nrep, col_tpb(0), d0.lds_cols, st>>>(dps, per);
}

template <int NZ>
static void launch_cols_split_batch_nz(const DipoleDev& d0, const DipoleDev* dps, int nrep, hipStream_t st)
{
#define MSD_CSB(L)                                                                  \
    if (d0.ly == L) {                                                               \
        const int per = d0.nxh * d0.split;                                          \
        k_cols_split_batch<L, NZ><<<per * nrep, split_tpb(L, NZ), d0.lds_cols, st>>>(dps, per); \
        return;                                                                     \
    }
    MSD_SPLIT_LENGTHS(MSD_CSB)
#undef MSD_CSB
}

hipError_t launch_cols_batch(const DipoleDev& d0, const DipoleDev* dps, int nrep, hipStream_t st)
{
    if (d0.split) {
        if (!split_length_compiled(d0.ly, d0.split)) return hipErrorInvalidValue;
        switch (d0.nz) {
            case 1: launch_cols_split_batch_nz<1>(d0, dps, nrep, st); break;
            case 2: launch_cols_split_batch_nz<2>(d0, dps, nrep, st); break;
            case 3: launch_cols_split_batch_nz<3>(d0, dps, nrep, st); break;
            case 4: launch_cols_split_batch_nz<4>(d0, dps, nrep, st); break;
            case 5: launch_cols_split_batch_nz<5>(d0, dps, nrep, st); break;
            case 6: launch_cols_split_batch_nz<6>(d0, dps, nrep, st); break;
            case 7: launch_cols_split_batch_nz<7>(d0, dps, nrep, st); break;
            case 8: launch_cols_split_batch_nz<8>(d0, dps, nrep, st); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (d0.nz) {
        case 1: launch_cols_batch_nz<1>(d0, dps, nrep, st); break;
        case 2: launch_cols_batch_nz<2>(d0, dps, nrep, st); break;
        case 3: launch_cols_batch_nz<3>(d0, dps, nrep, st); break;
        case 4: launch_cols_batch_nz<4>(d0, dps, nrep, st); break;
        case 5: launch_cols_batch_nz<5>(d0, dps, nrep, st); break;
        case 6: launch_cols_batch_nz<6>(d0, dps, nrep, st); break;
        case 7: launch_cols_batch_nz<7>(d0, dps, nrep, st); break;
        case 8: launch_cols_batch_nz<8>(d0, dps, nrep, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

const char* kernel_name(int id)
{
    static const char* names[] = {"k_cols", "k_step_pred", "k_step_corr", "k_rows_fwd"};
    return (id >= 0 && id < 4) ? names[id] : "?";
}

// Allow every LDS-staged kernel the device's full LDS per workgroup.
static void set_lds(const void* f, int v, hipError_t& first)
{
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, v);
    if (e != hipSuccess && first == hipSuccess) first = e;
}

template <int NZ>
static void cols_attr(int v, hipError_t& first)
{
#define MSD_CA(NYV) set_lds((const void*)k_cols<NYV, NZ>, v, first); set_lds((const void*)k_cols_batch<NYV, NZ>, v, first);
    MSD_COL_LENGTHS(MSD_CA)
#undef MSD_CA
    set_lds((const void*)k_cols<0, NZ>, v, first);
    set_lds((const void*)k_cols_batch<0, NZ>, v, first);
#define MSD_CP(NYV) set_lds((const void*)k_cols_pair<NYV, NZ>, v, first); set_lds((const void*)k_cols_pair_batch<NYV, NZ>, v, first);
    MSD_COL_LENGTHS(MSD_CP)
#undef MSD_CP
    set_lds((const void*)k_cols_pair<0, NZ>, v, first);
    set_lds((const void*)k_cols_pair_batch<0, NZ>, v, first);
#define MSD_CSA(L) set_lds((const void*)k_cols_split<L, NZ>, v, first); set_lds((const void*)k_cols_split_batch<L, NZ>, v, first);
    MSD_SPLIT_LENGTHS(MSD_CSA)
#undef MSD_CSA
}

hipError_t configure_lds(size_t max_bytes)
{
    const int v = (int)max_bytes;
    hipError_t first = hipSuccess;
#define MSD_RA(NXV, TYV, TPBV)                                                      \
    set_lds((const void*)k_rows_fwd<NXV, TYV, TPBV>, v, first);                     \
    set_lds((const void*)k_rows_inv<NXV, TYV, TPBV>, v, first);                     \
    set_lds((const void*)k_step_pred<NXV, TYV, TPBV, true, true>, v, first);        \
    set_lds((const void*)k_step_pred<NXV, TYV, TPBV, true, false>, v, first);       \
    set_lds((const void*)k_step_corr<NXV, TYV, TPBV, true, true>, v, first);        \
    set_lds((const void*)k_step_corr<NXV, TYV, TPBV, true, false>, v, first);   \
    set_lds((const void*)k_step_pred_batch<NXV, TYV, TPBV, true>, v, first);      \
    set_lds((const void*)k_step_pred_batch<NXV, TYV, TPBV, false>, v, first);     \
    set_lds((const void*)k_step_corr_batch<NXV, TYV, TPBV, true>, v, first);      \
    set_lds((const void*)k_step_corr_batch<NXV, TYV, TPBV, false>, v, first);     \
    set_lds((const void*)k_rk_stage<NXV, TYV, TPBV, true>, v, first);
    MSD_ROW_SHAPES(MSD_RA)
    MSD_RA(0, 0, ROW_TPB)
#undef MSD_RA
    cols_attr<1>(v, first); cols_attr<2>(v, first); cols_attr<3>(v, first); cols_attr<4>(v, first);
    cols_attr<5>(v, first); cols_attr<6>(v, first); cols_attr<7>(v, first); cols_attr<8>(v, first);
    (void)hipGetLastError();
    return first;
}

}  // namespace msd
