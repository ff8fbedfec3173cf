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
#include "msd_device.h"

namespace msd {

static constexpr int ROW_TPB = 256;
static constexpr int COL_TPB = 256;

// workgroups on the same XCD (hardware assigns bid % 8) get contiguous tiles
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    if (nb & 7) return bid;
    return (bid & 7) * (nb >> 3) + (bid >> 3);
}

// two real rows per complex line: real row q -> line q/2, component q&1
__device__ __forceinline__ float& rowval(float2* buf, int q, int x, int nx)
{
    return reinterpret_cast<float*>(buf + (size_t)(q >> 1) * nx + x)[q & 1];
}

__device__ __forceinline__ int row_lines(int ty) { return (3 * ty + 1) >> 1; }

// Build the packed complex lines Z = A + iB from the half spectra of rows
// (2l, 2l+1) read from `spec` (column layout), using A[nx-k] = conj(A[k]).
__device__ void rows_pack_spectrum(const DipoleDev& dp, const float2* __restrict__ spec, int z, int y0, int ty,
                                   float2* __restrict__ lds)
{
    const int nx = dp.nx, lines = row_lines(ty), total = lines * nx, half = nx >> 1;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int l = idx % lines;
        const int k = idx / lines;
        const bool mirror = k > half;
        const int kk = mirror ? nx - k : k;
        const int q0 = 2 * l, q1 = q0 + 1;
        const int c0 = q0 / ty, r0 = q0 - c0 * ty;
        float2 A = spec[((size_t)(c0 * dp.nz + z) * dp.nxh + kk) * dp.ny + y0 + r0];
        float2 B = make_float2(0.f, 0.f);
        if (q1 < 3 * ty) {
            const int c1 = q1 / ty, r1 = q1 - c1 * ty;
            B = spec[((size_t)(c1 * dp.nz + z) * dp.nxh + kk) * dp.ny + y0 + r1];
        }
        if (mirror) { A.y = -A.y; B.y = -B.y; }
        lds[(size_t)l * nx + k] = make_float2(A.x - B.y, A.y + B.x);
    }
}

// Split the transformed packed lines into the half spectra of both rows and
// store them in column layout.
__device__ void rows_unpack_spectrum(const DipoleDev& dp, const float2* __restrict__ lds, int z, int y0, int ty,
                                     float2* __restrict__ spec)
{
    const int nx = dp.nx, lines = row_lines(ty), total = lines * dp.nxh;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int l = idx % lines;
        const int k = idx / lines;
        const float2 Zk = lds[(size_t)l * nx + k];
        const float2 Zm = lds[(size_t)l * nx + (k ? nx - k : 0)];
        const int q0 = 2 * l, q1 = q0 + 1;
        const int c0 = q0 / ty, r0 = q0 - c0 * ty;
        spec[((size_t)(c0 * dp.nz + z) * dp.nxh + k) * dp.ny + y0 + r0] =
            make_float2(0.5f * (Zk.x + Zm.x), 0.5f * (Zk.y - Zm.y));
        if (q1 < 3 * ty) {
            const int c1 = q1 / ty, r1 = q1 - c1 * ty;
            spec[((size_t)(c1 * dp.nz + z) * dp.nxh + k) * dp.ny + y0 + r1] =
                make_float2(0.5f * (Zk.y + Zm.y), 0.5f * (Zm.x - Zk.x));
        }
    }
}

// ------------------------------------------------------------ op-API rows
// SpinSystem::fft (cpu/core/spinsystem.cpp:936-966) along x: spins -> Mh
__global__ void __launch_bounds__(ROW_TPB) k_rows_fwd(DipoleDev dp, const float* __restrict__ s)
{
    extern __shared__ float2 lds[];
    const int ty = dp.ty, tpl = dp.ny / ty;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int z = tile / tpl, y0 = (tile - z * tpl) * ty;
    const int nx = dp.nx, n = dp.nx * dp.ny * dp.nz, lines = row_lines(ty);
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    for (int li = threadIdx.x; li < lines * nx; li += blockDim.x) A[li] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int li = threadIdx.x; li < 3 * ty * nx; li += blockDim.x) {
        const int q = li / nx, x = li - q * nx;
        const int c = q / ty, r = q - c * ty;
        rowval(A, q, x, nx) = s[(size_t)c * n + x + (size_t)nx * (y0 + r + (size_t)dp.ny * z)];
    }
    __syncthreads();
    const float2* R = lds_fft<-1>(A, B, lines, dp.px, dp.twx);
    rows_unpack_spectrum(dp, R, z, y0, ty, dp.Mh);
}

// inverse along x: Hh -> real field rows, written to slot h
__global__ void __launch_bounds__(ROW_TPB) k_rows_inv(DipoleDev dp, float* __restrict__ h)
{
    extern __shared__ float2 lds[];
    const int ty = dp.ty, tpl = dp.ny / ty;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int z = tile / tpl, y0 = (tile - z * tpl) * ty;
    const int nx = dp.nx, n = dp.nx * dp.ny * dp.nz, lines = row_lines(ty);
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    rows_pack_spectrum(dp, dp.Hh, z, y0, ty, A);
    __syncthreads();
    float2* R = lds_fft<1>(A, B, lines, dp.px, dp.twx);
    for (int li = threadIdx.x; li < 3 * ty * nx; li += blockDim.x) {
        const int q = li / nx, x = li - q * nx;
        const int c = q / ty, r = q - c * ty;
        h[(size_t)c * n + x + (size_t)nx * (y0 + r + (size_t)dp.ny * z)] = rowval(R, q, x, nx);
    }
}

// ------------------------------------------------------------- column pass
// For one kx: forward along y, per-k layer/component tensor product, inverse along y.
template <int NZ>
__global__ void __launch_bounds__(COL_TPB) k_cols(DipoleDev dp)
{
    extern __shared__ float2 lds[];
    const int kx = blockIdx.x;
    const int ny = dp.ny, nxh = dp.nxh, lines = 3 * NZ;
    float2* A = lds;
    float2* B = lds + (size_t)lines * ny;
    for (int idx = threadIdx.x; idx < lines * ny; idx += blockDim.x) {
        const int l = idx / ny, y = idx - l * ny;
        A[idx] = dp.Mh[((size_t)l * nxh + kx) * ny + y];
    }
    __syncthreads();
    float2* R = lds_fft<-1>(A, B, lines, dp.py, dp.twy);
    for (int ky = threadIdx.x; ky < ny; ky += blockDim.x) {
        float2 m[3][NZ];
        float w[6][NZ];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int s = 0; s < NZ; ++s) m[c][s] = R[(c * NZ + s) * ny + ky];
#pragma unroll
        for (int a = 0; a < 6; ++a)
#pragma unroll
            for (int o = 0; o < NZ; ++o) w[a][o] = dp.Tw[((size_t)(a * NZ + o) * nxh + kx) * ny + ky];
#pragma unroll
        for (int t = 0; t < NZ; ++t) {
            float2 hx = make_float2(0.f, 0.f), hy = hx, hz = hx;
#pragma unroll
            for (int s = 0; s < NZ; ++s) {
                const int o = s >= t ? s - t : t - s;
                const float sg = s < t ? -1.0f : 1.0f;
                const float xx = w[0][o], xy = w[1][o], xz = sg * w[2][o];
                const float yy = w[3][o], yz = sg * w[4][o], zz = w[5][o];
                const float2 mx = m[0][s], my = m[1][s], mz = m[2][s];
                // i*w*m = w*(-m.y, m.x)
                hx.x += xx * mx.x + xy * my.x - xz * mz.y;
                hx.y += xx * mx.y + xy * my.y + xz * mz.x;
                hy.x += xy * mx.x + yy * my.x - yz * mz.y;
                hy.y += xy * mx.y + yy * my.y + yz * mz.x;
                hz.x += -xz * mx.y - yz * my.y + zz * mz.x;
                hz.y += xz * mx.x + yz * my.x + zz * mz.y;
            }
            R[(0 * NZ + t) * ny + ky] = hx;
            R[(1 * NZ + t) * ny + ky] = hy;
            R[(2 * NZ + t) * ny + ky] = hz;
        }
    }
    __syncthreads();
    float2* O = lds_fft<1>(R, R == A ? B : A, lines, dp.py, dp.twy);
    for (int idx = threadIdx.x; idx < lines * ny; idx += blockDim.x) {
        const int l = idx / ny, y = idx - l * ny;
        dp.Hh[((size_t)l * nxh + kx) * ny + y] = O[idx];
    }
}

static int row_tiles(const DipoleDev& dp) { return dp.nz * (dp.ny / dp.ty); }

hipError_t launch_rows_fwd(const DipoleDev& dp, const float* s, hipStream_t st)
{
    k_rows_fwd<<<row_tiles(dp), ROW_TPB, dp.lds_rows, st>>>(dp, s);
    return hipGetLastError();
}

hipError_t launch_rows_inv(const DipoleDev& dp, float* h, hipStream_t st)
{
    k_rows_inv<<<row_tiles(dp), ROW_TPB, dp.lds_rows, st>>>(dp, h);
    return hipGetLastError();
}

hipError_t launch_cols(const DipoleDev& dp, hipStream_t st)
{
    switch (dp.nz) {
#define MSD_COLS(N) case N: k_cols<N><<<dp.nxh, COL_TPB, dp.lds_cols, st>>>(dp); break;
        MSD_COLS(1) MSD_COLS(2) MSD_COLS(3) MSD_COLS(4) MSD_COLS(5) MSD_COLS(6) MSD_COLS(7) MSD_COLS(8)
#undef MSD_COLS
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ fused Heun
// Predictor half of the step closure in common/scripts/PredictorCorrector.lua:77-84:
//   H(S) = Exchange + Anisotropy + Dipole,  dM = LLG(S, H(S)),
//   P = N(S + dt dM),  C = N(S + dt/2 dM),
// then the x-transform of P for the next column pass.
template <bool DIP>
__global__ void __launch_bounds__(ROW_TPB) k_step_pred(StepArgs a)
{
    extern __shared__ float2 lds[];
    const DipoleDev& dp = a.dp;
    const int nx = a.d.nx, ny = a.d.ny, n = a.d.n;
    const int ty = a.ty, tpl = ny / ty;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int z = tile / tpl, y0 = (tile - z * tpl) * ty;
    const int lines = row_lines(ty);
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    float2* R = A;
    if (DIP) {
        rows_pack_spectrum(dp, dp.Hh, z, y0, ty, A);
        __syncthreads();
        R = lds_fft<1>(A, B, lines, dp.px, dp.twx);
    }
    if (a.th.on && blockIdx.x == 0 && threadIdx.x == 0) *a.draw += 1ull;
    const float* __restrict__ S = a.S;
    const float* __restrict__ ms_a = a.S + 3 * (size_t)n;
    const float dt = a.dt, hdt = 0.5f * a.dt;
    for (int li = threadIdx.x; li < ty * nx; li += blockDim.x) {
        const int r = li / nx, x = li - r * nx, y = y0 + r;
        const int i = x + nx * (y + ny * z);
        const float ms = ms_a[i];
        const float3 s = make_float3(S[i], S[n + i], S[2 * n + i]);
        float3 H = exchange_field(a.ex, a.d, i, x, y, z, S, ms_a);
        const float3 ha = aniso_field(a.an, i, n, s, ms);
        H.x += ha.x; H.y += ha.y; H.z += ha.z;
        if (DIP) {
            H.x += rowval(R, r, x, nx);
            H.y += rowval(R, ty + r, x, nx);
            H.z += rowval(R, 2 * ty + r, x, nx);
        }
        float3 p = s, c = s;
        if (ms > 0.0f) {
            const float al = a.lp.site_alpha ? a.lp.site_alpha[i] : a.lp.alpha;
            const float ga = a.lp.site_gamma ? a.lp.site_gamma[i] : a.lp.gamma;
            const float3 dM = a.lp.thermal_first ? llg_rhs(s, H, H, al, ga, ms, a.lp.ps, a.lp.ds)
                                                 : llg_rhs(s, H, H, al, ga, ms, 1.0f, 1.0f);
            p = make_float3(s.x + dt * dM.x, s.y + dt * dM.y, s.z + dt * dM.z);
            c = make_float3(s.x + hdt * dM.x, s.y + hdt * dM.y, s.z + hdt * dM.z);
        }
        if (a.lp.renorm) { p = renorm(p, ms); c = renorm(c, ms); }
        a.P[i] = p.x; a.P[n + i] = p.y; a.P[2 * n + i] = p.z;
        a.C[i] = c.x; a.C[n + i] = c.y; a.C[2 * n + i] = c.z;
        if (DIP) {
            rowval(R, r, x, nx) = p.x;
            rowval(R, ty + r, x, nx) = p.y;
            rowval(R, 2 * ty + r, x, nx) = p.z;
        }
    }
    if (DIP) {
        __syncthreads();
        const float2* F = lds_fft<-1>(R, R == A ? B : A, lines, dp.px, dp.twx);
        rows_unpack_spectrum(dp, F, z, y0, ty, dp.Mh);
    }
}

// Corrector half (PredictorCorrector.lua:81-85) and the thermal kick (:98-103):
//   S' = N(C + dt/2 LLG(P, H(P))),  S'' = N(S' + dt LLG(S', H_th)) when Thermal is present,
// then the x-transform of S'' for the next step's first column pass.
template <bool DIP>
__global__ void __launch_bounds__(ROW_TPB) k_step_corr(StepArgs a)
{
    extern __shared__ float2 lds[];
    const DipoleDev& dp = a.dp;
    const int nx = a.d.nx, ny = a.d.ny, n = a.d.n;
    const int ty = a.ty, tpl = ny / ty;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int z = tile / tpl, y0 = (tile - z * tpl) * ty;
    const int lines = row_lines(ty);
    float2* A = lds;
    float2* B = lds + (size_t)lines * nx;
    float2* R = A;
    if (DIP) {
        rows_pack_spectrum(dp, dp.Hh, z, y0, ty, A);
        __syncthreads();
        R = lds_fft<1>(A, B, lines, dp.px, dp.twx);
    }
    const uint64_t draw = a.th.on ? (uint64_t)*a.draw : 0ull;
    float* __restrict__ S = a.S;
    const float* __restrict__ ms_a = a.S + 3 * (size_t)n;
    const float* __restrict__ P = a.P;
    const float dt = a.dt, hdt = 0.5f * a.dt;
    for (int li = threadIdx.x; li < ty * nx; li += blockDim.x) {
        const int r = li / nx, x = li - r * nx, y = y0 + r;
        const int i = x + nx * (y + ny * z);
        const float ms = ms_a[i];
        const float3 p = make_float3(P[i], P[n + i], P[2 * n + i]);
        float3 H = exchange_field(a.ex, a.d, i, x, y, z, P, ms_a);
        const float3 ha = aniso_field(a.an, i, n, p, ms);
        H.x += ha.x; H.y += ha.y; H.z += ha.z;
        if (DIP) {
            H.x += rowval(R, r, x, nx);
            H.y += rowval(R, ty + r, x, nx);
            H.z += rowval(R, 2 * ty + r, x, nx);
        }
        float3 v = make_float3(a.C[i], a.C[n + i], a.C[2 * n + i]);
        const float al = a.lp.site_alpha ? a.lp.site_alpha[i] : a.lp.alpha;
        const float ga = a.lp.site_gamma ? a.lp.site_gamma[i] : a.lp.gamma;
        if (ms > 0.0f) {
            const float3 dM = a.lp.thermal_first ? llg_rhs(p, H, H, al, ga, ms, a.lp.ps, a.lp.ds)
                                                 : llg_rhs(p, H, H, al, ga, ms, 1.0f, 1.0f);
            v = make_float3(v.x + hdt * dM.x, v.y + hdt * dM.y, v.z + hdt * dM.z);
        }
        if (a.lp.renorm) v = renorm(v, ms);
        if (a.th.on) {
            const float sg = thermal_sigma(a.th, a.lp, i, ms, dt);
            if (ms > 0.0f) {
                float3 Ht = make_float3(0.f, 0.f, 0.f);
                if (sg != 0.0f) {
                    const float3 g = thermal_normals(a.th.seed, draw, i);
                    Ht = make_float3(sg * g.x, sg * g.y, sg * g.z);
                }
                const float3 zero = make_float3(0.f, 0.f, 0.f);
                const float3 dM = a.lp.thermal_first ? llg_rhs(v, Ht, zero, al, ga, ms, a.lp.ps, a.lp.ds)
                                                     : llg_rhs(v, Ht, Ht, al, ga, ms, 1.0f, 1.0f);
                v = make_float3(v.x + dt * dM.x, v.y + dt * dM.y, v.z + dt * dM.z);
            }
            if (a.lp.renorm) v = renorm(v, ms);
        }
        S[i] = v.x; S[n + i] = v.y; S[2 * n + i] = v.z;
        if (DIP) {
            rowval(R, r, x, nx) = v.x;
            rowval(R, ty + r, x, nx) = v.y;
            rowval(R, 2 * ty + r, x, nx) = v.z;
        }
    }
    if (DIP) {
        __syncthreads();
        const float2* F = lds_fft<-1>(R, R == A ? B : A, lines, dp.px, dp.twx);
        rows_unpack_spectrum(dp, F, z, y0, ty, dp.Mh);
    }
}

hipError_t launch_step_pred(const StepArgs& a, hipStream_t st)
{
    const int tiles = a.d.nz * (a.d.ny / a.ty);
    if (a.has_dip) k_step_pred<true><<<tiles, ROW_TPB, a.dp.lds_rows, st>>>(a);
    else k_step_pred<false><<<tiles, ROW_TPB, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_step_corr(const StepArgs& a, hipStream_t st)
{
    const int tiles = a.d.nz * (a.d.ny / a.ty);
    if (a.has_dip) k_step_corr<true><<<tiles, ROW_TPB, a.dp.lds_rows, st>>>(a);
    else k_step_corr<false><<<tiles, ROW_TPB, 0, st>>>(a);
    return hipGetLastError();
}

const char* kernel_name(int id)
{
    static const char* names[] = {"k_cols", "k_step_pred", "k_step_corr", "k_rows_fwd"};
    return (id >= 0 && id < 4) ? names[id] : "?";
}

hipError_t configure_lds(size_t max_bytes)
{
    const int v = (int)max_bytes;
    const void* fns[] = {(const void*)k_rows_fwd, (const void*)k_rows_inv,
                         (const void*)k_cols<1>, (const void*)k_cols<2>, (const void*)k_cols<3>, (const void*)k_cols<4>,
                         (const void*)k_cols<5>, (const void*)k_cols<6>, (const void*)k_cols<7>, (const void*)k_cols<8>,
                         (const void*)k_step_pred<true>, (const void*)k_step_corr<true>};
    hipError_t first = hipSuccess;
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, v);
        if (e != hipSuccess && first == hipSuccess) first = e;
    }
    (void)hipGetLastError();
    return first;
}

// --------------------------------------------------- tensor precompute (FP64)
struct Mat9 { double a[9]; };

__device__ __forceinline__ long long fdiv_ll(long long a, long long b)
{
    long long q = a / b;
    if ((a % b) && ((a < 0) != (b < 0))) --q;
    return q;
}

// getWAB / getWAB_range (cpu/dipole/dipolesupport.cpp:71-250) for smin = 0:
// images X,Y in [-(gmax0-1), gmax0-1], ascending, |xx| <= min(gmax2,gmax0),
// |yy| <= min(gmax3,gmax0), xx^2+yy^2+zz^2 <= gmax0^2; no z cut beyond the radial one.
__global__ void k_tensor_sum(int nx, int ny, int nz, int4 gm, Mat9 M, double* __restrict__ T)
{
    const int nl = 2 * nz - 1;
    const size_t total = (size_t)nx * ny * nl, plane = total;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % nx);
        const int j = (int)((idx / nx) % ny);
        const int c = (int)(idx / ((size_t)nx * ny));
        const long long zz = c - nz + 1;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        if (llabs(zz) <= gm.x) {
            const long long smax = gm.x;
            const long long limx = min(gm.z, gm.x), limy = min(gm.w, gm.x);
            const long long x0 = max(-(smax - 1), -fdiv_ll(limx + i, nx));
            const long long x1 = min(smax - 1, fdiv_ll(limx - i, nx));
            const long long y0 = max(-(smax - 1), -fdiv_ll(limy + j, ny));
            const long long y1 = min(smax - 1, fdiv_ll(limy - j, ny));
            const double r2max = (double)gm.x * (double)gm.x;
            for (long long X = x0; X <= x1; ++X) {
                const long long xx = X * nx + i;
                for (long long Y = y0; Y <= y1; ++Y) {
                    const long long yy = Y * ny + j;
                    if ((double)xx * xx + (double)yy * yy + (double)zz * zz <= r2max) {
                        const double rx = (double)xx * M.a[0] + (double)yy * M.a[3] + (double)zz * M.a[6];
                        const double ry = (double)xx * M.a[1] + (double)yy * M.a[4] + (double)zz * M.a[7];
                        const double rz = (double)xx * M.a[2] + (double)yy * M.a[5] + (double)zz * M.a[8];
                        const double r = sqrt(rx * rx + ry * ry + rz * rz);
                        if (r < 1e-40) continue;
                        const double ir = 1.0 / r, ir3 = ir * ir * ir, ir5 = ir3 * ir * ir;
                        acc[0] += -1.0 * (ir3 - 3.0 * rx * rx * ir5);
                        acc[1] += -1.0 * (-3.0 * rx * ry * ir5);
                        acc[2] += -1.0 * (-3.0 * rx * rz * ir5);
                        acc[3] += -1.0 * (ir3 - 3.0 * ry * ry * ir5);
                        acc[4] += -1.0 * (-3.0 * ry * rz * ir5);
                        acc[5] += -1.0 * (ir3 - 3.0 * rz * rz * ir5);
                    }
                }
            }
        }
        for (int a = 0; a < 6; ++a) T[(size_t)a * plane + idx] = acc[a];
    }
}

hipError_t launch_tensor_sum(int nx, int ny, int nz, const int gmax[4], const double ABC[9], double* T, hipStream_t st)
{
    Mat9 M;
    for (int k = 0; k < 9; ++k) M.a[k] = ABC[k];
    const size_t total = (size_t)nx * ny * (2 * nz - 1);
    const int g = (int)std::min<size_t>((total + 127) / 128, 65536);
    k_tensor_sum<<<g, 128, 0, st>>>(nx, ny, nz, make_int4(gmax[0], gmax[1], gmax[2], gmax[3]), M, T);
    return hipGetLastError();
}

// x-DFT of layers z >= 0: U[a][zz][y][kx] = sum_x T[a][nz-1+zz][y][x] e^{-2 pi i kx x / nx}
__global__ void k_tensor_dftx(int nx, int ny, int nz, const double* __restrict__ T, double2* __restrict__ U,
                              const double2* __restrict__ tw)
{
    const int nxh = nx / 2 + 1, nl = 2 * nz - 1;
    const size_t total = (size_t)6 * nz * ny * nxh;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(idx % nxh);
        const int y = (int)((idx / nxh) % ny);
        const int zz = (int)((idx / ((size_t)nxh * ny)) % nz);
        const int a = (int)(idx / ((size_t)nxh * ny * nz));
        const double* row = T + (((size_t)a * nl + (nz - 1 + zz)) * ny + y) * nx;
        double re = 0, im = 0;
        int m = 0;
        for (int x = 0; x < nx; ++x) {
            const double2 w = tw[m];
            re += row[x] * w.x;
            im += row[x] * w.y;
            m += kx; if (m >= nx) m -= nx;
        }
        U[idx] = make_double2(re, im);
    }
}

// y-DFT and word extraction: W[a][zz][kx][ky] = Re or Im of sum_y U e^{-2 pi i ky y / ny}, / (nx ny)
__global__ void k_tensor_dfty(int nx, int ny, int nz, const double2* __restrict__ U, double* __restrict__ W,
                              const double2* __restrict__ tw)
{
    const int nxh = nx / 2 + 1;
    const size_t total = (size_t)6 * nz * nxh * ny;
    const double inv = 1.0 / ((double)nx * (double)ny);
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int ky = (int)(idx % ny);
        const int kx = (int)((idx / ny) % nxh);
        const int zz = (int)((idx / ((size_t)ny * nxh)) % nz);
        const int a = (int)(idx / ((size_t)ny * nxh * nz));
        const double2* col = U + ((size_t)a * nz + zz) * ny * nxh + kx;
        double re = 0, im = 0;
        int m = 0;
        for (int y = 0; y < ny; ++y) {
            const double2 u = col[(size_t)y * nxh];
            const double2 w = tw[m];
            re += u.x * w.x - u.y * w.y;
            im += u.x * w.y + u.y * w.x;
            m += ky; if (m >= ny) m -= ny;
        }
        W[idx] = ((a == 2 || a == 4) ? im : re) * inv;
    }
}

hipError_t launch_tensor_spectrum(int nx, int ny, int nz, const double* T, double2* U, double* W,
                                  const double2* twx, const double2* twy, hipStream_t st)
{
    const size_t tot = (size_t)6 * nz * ny * (nx / 2 + 1);
    const int g = (int)std::min<size_t>((tot + 255) / 256, 65536);
    k_tensor_dftx<<<g, 256, 0, st>>>(nx, ny, nz, T, U, twx);
    k_tensor_dfty<<<g, 256, 0, st>>>(nx, ny, nz, U, W, twy);
    return hipGetLastError();
}

__global__ void k_tensor_words(const double* __restrict__ W, float* __restrict__ Tw, double mult, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        Tw[i] = (float)(W[i] * mult);
}

hipError_t launch_tensor_words(const double* W, float* Tw, double mult, size_t n, hipStream_t st)
{
    const int g = (int)std::min<size_t>((n + 255) / 256, 65536);
    k_tensor_words<<<g, 256, 0, st>>>(W, Tw, mult, n);
    return hipGetLastError();
}

}  // namespace msd
