// Dipole tensor precompute in FP64 on the device: the periodic in-plane lattice
// sum of cpu/dipole/dipolesupport.cpp and its 2D spectrum (LongRange::updateData,
// cpu/longrange/spinoperationlongrange.cpp:245-309), reduced to one FP32 word per
// component per k.
#include "msd_device.h"

#include <algorithm>

namespace msd {

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

// setTruncation(math.huge) (cpu/dipole/dipolesupport.cpp:756-840): the reference sums
// rings of image lattices 5 wide and asks extrapolate.lua for the limit after each,
// but reads the script's success flag from a padding nil (lua_pcall(L, 1, 2, 0) with
// a one-value script, then lua_toboolean(L, -1)), so every element runs its 1000
// rings and keeps the sum over the images |X|, |Y| <= 4999 with no radial or
// in-plane cut.  One workgroup per element (in-plane offset i, j, layer offset c):
// thread t sums the image rows X = -4999 + t, t + HUGE_TPB, ..., each over all Y, and
// the partial sums are added in a fixed tree, so the result is reproducible.
constexpr int HUGE_TPB = 256, HUGE_HALF = 4999;

__global__ void __launch_bounds__(HUGE_TPB) k_tensor_huge(int nx, int ny, int nz, Mat9 M, size_t first,
                                                          double* __restrict__ T)
{
    __shared__ double part[6][HUGE_TPB];
    const int nl = 2 * nz - 1;
    const size_t plane = (size_t)nx * ny * nl;
    const size_t idx = first + blockIdx.x;
    const int i = (int)(idx % nx);
    const int j = (int)((idx / nx) % ny);
    const long long zz = (long long)(idx / ((size_t)nx * ny)) - nz + 1;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int X = -HUGE_HALF + (int)threadIdx.x; X <= HUGE_HALF; X += HUGE_TPB) {
        const double xx = (double)((long long)X * nx + i);
        for (int Y = -HUGE_HALF; Y <= HUGE_HALF; ++Y) {
            const double yy = (double)((long long)Y * ny + j);
            const double rx = xx * M.a[0] + yy * M.a[3] + (double)zz * M.a[6];
            const double ry = xx * M.a[1] + yy * M.a[4] + (double)zz * M.a[7];
            const double rz = xx * M.a[2] + yy * M.a[5] + (double)zz * M.a[8];
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
    for (int a = 0; a < 6; ++a) part[a][threadIdx.x] = acc[a];
    __syncthreads();
    for (int w = HUGE_TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 6; ++a) part[a][threadIdx.x] += part[a][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) T[(size_t)threadIdx.x * plane + idx] = part[threadIdx.x][0];
}

// launched in chunks of elements (each element is ~1e8 image terms, ~50 ms of a CU)
hipError_t launch_tensor_huge(int nx, int ny, int nz, const double ABC[9], double* T, hipStream_t st)
{
    Mat9 M;
    for (int k = 0; k < 9; ++k) M.a[k] = ABC[k];
    const size_t total = (size_t)nx * ny * (2 * nz - 1), chunk = 4096;
    for (size_t first = 0; first < total; first += chunk) {
        const size_t n = std::min(chunk, total - first);
        k_tensor_huge<<<(unsigned)n, HUGE_TPB, 0, st>>>(nx, ny, nz, M, first, T);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
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

// Per-layer-pair tensors (LongRange2D::compile, cpu/longrange2d/spinoperationlongrange2d.cpp:396-446):
// every nx x ny plane of real FP64 values gets its 2D forward transform (sign -1, unnormalised),
// prescaled by mult = 1/(nx ny).  Only the R2C half kx <= nx/2 is kept: the fields are real.
__global__ void k_plane_dftx(int planes, int nx, int ny, const double* __restrict__ T, double2* __restrict__ U,
                             const double2* __restrict__ tw)
{
    const int nxh = nx / 2 + 1;
    const size_t total = (size_t)planes * ny * nxh;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(idx % nxh);
        const size_t py = idx / nxh;                 // plane * ny + y
        const double* row = T + py * nx;
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

__global__ void k_plane_dfty(int planes, int nx, int ny, const double2* __restrict__ U, float2* __restrict__ out,
                             double mult, const double2* __restrict__ tw)
{
    const int nxh = nx / 2 + 1;
    const size_t total = (size_t)planes * nxh * ny;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total; idx += (size_t)gridDim.x * blockDim.x) {
        const int ky = (int)(idx % ny);
        const int kx = (int)((idx / ny) % nxh);
        const size_t p = idx / ((size_t)ny * nxh);
        const double2* col = U + p * ny * nxh + kx;
        double re = 0, im = 0;
        int m = 0;
        for (int y = 0; y < ny; ++y) {
            const double2 u = col[(size_t)y * nxh];
            const double2 w = tw[m];
            re += u.x * w.x - u.y * w.y;
            im += u.x * w.y + u.y * w.x;
            m += ky; if (m >= ny) m -= ny;
        }
        out[idx] = make_float2((float)(re * mult), (float)(im * mult));
    }
}

hipError_t launch_plane_spectrum(int planes, int nx, int ny, const double* T, double2* U, float2* out, double mult,
                                 const double2* twx, const double2* twy, hipStream_t st)
{
    const size_t tot = (size_t)planes * ny * (nx / 2 + 1);
    const int g = (int)std::min<size_t>((tot + 255) / 256, 65536);
    k_plane_dftx<<<g, 256, 0, st>>>(planes, nx, ny, T, U, twx);
    k_plane_dfty<<<g, 256, 0, st>>>(planes, nx, ny, U, out, mult, twy);
    return hipGetLastError();
}

// W [6][nz][nxh][ny] -> Tw.  With a split y transform (R > 0) output word
// (((kx R + k1) 6 + a) nz + o) ly + k2 takes W[a][o][kx][k1 + R k2]: the words of one
// column workgroup (kx, k1) are one contiguous run.
__global__ void k_tensor_words(const double* __restrict__ W, float* __restrict__ Tw, double mult, int nz, int nxh, int ny,
                               int R)
{
    const size_t n = (size_t)6 * nz * nxh * ny;
    const int ly = R > 0 ? ny / R : ny;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        size_t src = i;
        if (R > 0) {
            const int k2 = (int)(i % ly);
            const int o = (int)((i / ly) % nz);
            const int a = (int)((i / ((size_t)ly * nz)) % 6);
            const size_t w = i / ((size_t)ly * nz * 6);          // kx * R + k1
            const int kx = (int)(w / R), k1 = (int)(w % R);
            src = (((size_t)a * nz + o) * nxh + kx) * ny + k1 + (size_t)R * k2;
        }
        Tw[i] = (float)(W[src] * mult);
    }
}

hipError_t launch_tensor_words(const double* W, float* Tw, double mult, int nz, int nxh, int ny, int split,
                               hipStream_t st)
{
    const size_t n = (size_t)6 * nz * nxh * ny;
    const int g = (int)std::min<size_t>((n + 255) / 256, 65536);
    k_tensor_words<<<g, 256, 0, st>>>(W, Tw, mult, nz, nxh, ny, split);
    return hipGetLastError();
}

}  // namespace msd
