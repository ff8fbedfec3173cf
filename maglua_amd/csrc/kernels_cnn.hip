// Dipolar-field surrogate network on gfx950: three 5x5x3 convolution layers
// (3 -> 16 -> 16 -> 3 channels, no bias, identity activation), BF16 operands on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  The network has no
// counterpart in the reference; it predicts the change of the layered dipole
// field (LongRange::apply, cpu/longrange/spinoperationlongrange.cpp:354-430)
// between exact refreshes.
//
// Activations are channels-last BF16 with 16 channels per site (32 bytes; the
// 3-channel input and output are zero-padded).  A layer is
//   Y[s][o] = sum_t sum_i W[t][o][i] X[s + off(t)][i],
//   off(t) = (dx, dy, dz) = (t%5 - 2, (t/5)%5 - 2, t/25 - 1),
// periodic in x and y, zero outside the lattice in z.  One MFMA covers two taps
// (K = 2 taps x 16 input channels), so 38 MFMA steps cover the 75 taps (the
// 76th tap's weights are zero).
//
// MFMA operand layout (16x16x32, checked on the device): lane l holds
// A[l%16][8(l/16) .. +7], B[8(l/16) .. +7][l%16], and D[4(l/16) .. +3][l%16].
#include "msd_device.h"

namespace msd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static constexpr int CNN_TPB = 256;                  // 4 waves
static constexpr int CTX = 32, CTY = 4;              // output tile: 32 x, 4 rows (one per wave), one layer
static constexpr int HX = CTX + 4, HY = CTY + 4, HZ = 3;
static constexpr int HALO = HX * HY * HZ;            // input sites staged per tile
static constexpr int NSTEP = 38;                     // MFMA steps per layer
static constexpr int NTAP = 75;

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);                 // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ void tap_off(int t, int& dx, int& dy, int& dz)
{
    dz = t / 25 - 1;
    dy = (t / 5) % 5 - 2;
    dx = t % 5 - 2;
}

__device__ __forceinline__ void tile_of(const CnnDims& d, int tile, int& x0, int& y0, int& z)
{
    const int tx = tile % d.tiles_x;
    const int r = tile / d.tiles_x;
    const int ty = r % d.tiles_y;
    z = r / d.tiles_y;
    x0 = tx * CTX;
    y0 = ty * CTY;
}

__device__ __forceinline__ int wrap(int v, int n)
{
    v += v < 0 ? n : 0;
    return v - (v >= n ? n : 0);
}

// ------------------------------------------------------------ layer
// One persistent workgroup per CU pair: the layer's packed weights are staged
// in LDS once, then each output tile's input halo is staged and every wave runs
// the 38 MFMA steps for its row's two 16-site tiles.
//   mode 0: BF16 channels-last output (hidden layers);
//   mode 1: FP32 SoA output[c][n] = href[c][n] (if given) + scale * Y[c], c < 3.
__global__ void __launch_bounds__(CNN_TPB) k_cnn_layer(CnnDims d, const uint4* __restrict__ in,
                                                    const uint4* __restrict__ apack, int mode,
                                                    uint4* __restrict__ out_bf, float* __restrict__ out_f,
                                                    const float* __restrict__ href, float scale)
{
    extern __shared__ uint4 cnn_lds[];
    uint4* sA = cnn_lds;                             // [NSTEP][64]
    uint4* sX = cnn_lds + NSTEP * 64;                // [HALO][2]
    for (int i = threadIdx.x; i < NSTEP * 64; i += CNN_TPB) sA[i] = apack[i];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, col = l & 15, g = l >> 4;
    const uint32_t n = (uint32_t)d.n;
    for (int tile = blockIdx.x; tile < d.ntiles; tile += gridDim.x) {
        int x0, y0, z;
        tile_of(d, tile, x0, y0, z);
        __syncthreads();
        // halo: HALO sites x 2 uint4, loaded in batches of 7 per thread
        constexpr int HI = (HALO * 2 + CNN_TPB - 1) / CNN_TPB;
        uint4 v[HI];
#pragma unroll
        for (int k = 0; k < HI; ++k) {
            const int i = (int)threadIdx.x + k * CNN_TPB;
            v[k] = make_uint4(0u, 0u, 0u, 0u);
            if (i < HALO * 2) {
                const int site = i >> 1, half = i & 1;
                const int hx = site % HX, r = site / HX, hy = r % HY, hz = r / HY;
                const int zz = z + hz - 1;
                if (zz >= 0 && zz < d.nz) {
                    const int xx = wrap(x0 + hx - 2, d.nx), yy = wrap(y0 + hy - 2, d.ny);
                    v[k] = ldg(in, (uint32_t)(xx + d.nx * (yy + d.ny * zz)) * 2u + (uint32_t)half);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < HI; ++k) {
            const int i = (int)threadIdx.x + k * CNN_TPB;
            if (i < HALO * 2) sX[i] = v[k];
        }
        __syncthreads();
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        for (int p = 0; p < NSTEP; ++p) {
            const bf16x8 a = as_bf8(sA[p * 64 + l]);
            int t = 2 * p + (g >> 1);
            t = t < NTAP ? t : NTAP - 1;                 // the zero-weight 76th tap reads tap 75's data
            int dx, dy, dz;
            tap_off(t, dx, dy, dz);
            const int base = ((1 + dz) * HY + (w + 2 + dy)) * HX + 2 + dx + col;
            const bf16x8 b0 = as_bf8(sX[base * 2 + (g & 1)]);
            const bf16x8 b1 = as_bf8(sX[(base + 16) * 2 + (g & 1)]);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
        }
        const int y = y0 + w;
        if (y < d.ny) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const f32x4 acc = j ? acc1 : acc0;
                const uint32_t site = (uint32_t)(x0 + 16 * j + col + d.nx * (y + d.ny * z));
                if (mode == 0) {
                    const uint32_t lo = (uint32_t)f2bf(acc[0]) | ((uint32_t)f2bf(acc[1]) << 16);
                    const uint32_t hi = (uint32_t)f2bf(acc[2]) | ((uint32_t)f2bf(acc[3]) << 16);
                    // channels 4g..4g+3 of the site: 8 bytes at 32 site + 8 g
                    stg(reinterpret_cast<uint2*>(out_bf), site * 4u + (uint32_t)g, make_uint2(lo, hi));
                } else if (g == 0) {
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        const float h0 = href ? ldg(href + (size_t)r * n, site) : 0.0f;
                        stg(out_f + (size_t)r * n, site, h0 + scale * acc[r]);
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------ weight gradient
// dW[t][o][i] = sum_s G[s][o] X[s + off(t)][i] as MFMAs with the sites as K:
// A[o][k] = G[site k][o], B[k][i] = X[site k + off(t)][i].  The output tile's G
// and the input halo are staged channel-major (site runs contiguous); wave w
// accumulates taps w, w+4, ... across every tile it visits and writes one
// partial per workgroup.
// 8 consecutive BF16 values from LDS at element offset e (any parity).
__device__ __forceinline__ bf16x8 lds_run8(const uint16_t* base, int e)
{
    const uint32_t* p = reinterpret_cast<const uint32_t*>(base) + (e >> 1);
    uint4 r;
    if (e & 1) {
        const uint32_t w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
        r = make_uint4(__builtin_amdgcn_alignbit(w1, w0, 16), __builtin_amdgcn_alignbit(w2, w1, 16),
                       __builtin_amdgcn_alignbit(w3, w2, 16), __builtin_amdgcn_alignbit(w4, w3, 16));
    } else {
        r = make_uint4(p[0], p[1], p[2], p[3]);
    }
    return as_bf8(r);
}

static constexpr int WG_TAPS = (NTAP + 3) / 4;       // taps per wave (19)

__global__ void __launch_bounds__(CNN_TPB) k_cnn_wgrad(CnnDims d, const uint16_t* __restrict__ X,
                                                    const uint16_t* __restrict__ Gr, float* __restrict__ partial)
{
    __shared__ uint16_t sG[16 * CTY * CTX];          // [o][row][x]
    __shared__ uint16_t sXc[16 * HZ * HY * HX];      // [i][hz][hy][hx]
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, col = l & 15, g = l >> 4;
    f32x4 acc[WG_TAPS];
#pragma unroll
    for (int k = 0; k < WG_TAPS; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tile = blockIdx.x; tile < d.ntiles; tile += gridDim.x) {
        int x0, y0, z;
        tile_of(d, tile, x0, y0, z);
        __syncthreads();
        // G tile: 128 sites x 16 channels, transposed into [o][row][x]
        for (int i = threadIdx.x; i < CTY * CTX * 16; i += CNN_TPB) {
            const int o = i & 15, s = i >> 4, x = s % CTX, row = s / CTX;
            const int y = y0 + row;
            uint16_t v = 0;
            if (y < d.ny) v = ldg(Gr, (uint32_t)(x0 + x + d.nx * (y + d.ny * z)) * 16u + (uint32_t)o);
            sG[(o * CTY + row) * CTX + x] = v;
        }
        // X halo transposed into [i][hz][hy][hx]
        for (int i = threadIdx.x; i < HALO * 16; i += CNN_TPB) {
            const int c = i & 15, s = i >> 4;
            const int hx = s % HX, r = s / HX, hy = r % HY, hz = r / HY;
            const int zz = z + hz - 1;
            uint16_t v = 0;
            if (zz >= 0 && zz < d.nz) {
                const int xx = wrap(x0 + hx - 2, d.nx), yy = wrap(y0 + hy - 2, d.ny);
                v = ldg(X, (uint32_t)(xx + d.nx * (yy + d.ny * zz)) * 16u + (uint32_t)c);
            }
            sXc[((c * HZ + hz) * HY + hy) * HX + hx] = v;
        }
        __syncthreads();
        bf16x8 a[CTY];
#pragma unroll
        for (int row = 0; row < CTY; ++row)
            a[row] = as_bf8(*reinterpret_cast<const uint4*>(sG + ((col * CTY + row) * CTX + 8 * g)));
#pragma unroll
        for (int k = 0; k < WG_TAPS; ++k) {
            const int t = w + 4 * k;
            if (t < NTAP) {
                int dx, dy, dz;
                tap_off(t, dx, dy, dz);
#pragma unroll
                for (int row = 0; row < CTY; ++row) {
                    const int e = ((col * HZ + 1 + dz) * HY + (row + 2 + dy)) * HX + 2 + dx + 8 * g;
                    acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[row], lds_run8(sXc, e), acc[k], 0, 0, 0);
                }
            }
        }
    }
    // D[o][i]: lane holds i = col, o = 4g + r
#pragma unroll
    for (int k = 0; k < WG_TAPS; ++k) {
        const int t = w + 4 * k;
        if (t < NTAP) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                partial[(((size_t)blockIdx.x * NTAP + t) * 16 + 4 * g + r) * 16 + col] = acc[k][r];
        }
    }
}

// sum the per-workgroup partials: grad[t][o][i] (16 x 16 per tap)
__global__ void k_cnn_wreduce(const float* __restrict__ partial, int nparts, float* __restrict__ grad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NTAP * 256) return;
    float s = 0.0f;
    for (int b = 0; b < nparts; ++b) s += partial[(size_t)b * NTAP * 256 + e];
    grad[e] = s;
}

// ------------------------------------------------------------ weight packing
// master FP32 W[t][o][i] (16 x 16 per tap, zero beyond the layer's channels) ->
// MFMA A fragments [38][64][8] BF16 for the layer (transposed = 0) or for its
// adjoint (transposed = 1: A'[t'] = W[74 - t']^T).
__global__ void k_cnn_pack_weights(const float* __restrict__ W, int transposed, uint16_t* __restrict__ apack)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NSTEP * 64 * 8) return;
    const int j = e & 7, lane = (e >> 3) & 63, p = e >> 9;
    const int g = lane >> 4, row = lane & 15, kk = 8 * (g & 1) + j;
    const int t = 2 * p + (g >> 1);
    float v = 0.0f;
    if (t < NTAP) {
        if (!transposed) v = W[(t * 16 + row) * 16 + kk];
        else v = W[((NTAP - 1 - t) * 16 + kk) * 16 + row];
    }
    apack[e] = f2bf(v);
}

// ------------------------------------------------------------ input, loss
// X0[s] = bf16(s_in (m - m_ref)) in channels 0..2, zero in 3..15
__global__ void k_cnn_input(Dims d, const float* __restrict__ m, const float* __restrict__ mref, float s_in,
                            uint4* __restrict__ X0)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = (uint32_t)d.n;
    if (i >= n) return;
    float v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float a = ldg(m + (size_t)c * n, i);
        const float b = mref ? ldg(mref + (size_t)c * n, i) : 0.0f;
        v[c] = s_in * (a - b);
    }
    const uint32_t w0 = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
    const uint32_t w1 = (uint32_t)f2bf(v[2]);
    stg(X0, 2u * i, make_uint4(w0, w1, 0u, 0u));
    stg(X0, 2u * i + 1u, make_uint4(0u, 0u, 0u, 0u));
}

// R = scale_out Y3 - T; loss partials sum R^2 (FP64); G3 = bf16(2 scale_out R / (3 n sigma2))
__global__ void k_cnn_loss(Dims d, const float* __restrict__ Y3, const float* __restrict__ T, float s_out,
                           float gscale, uint4* __restrict__ G3, double* __restrict__ part)
{
    __shared__ double red[256];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = (uint32_t)d.n;
    double acc = 0.0;
    if (i < n) {
        float gv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float r = s_out * ldg(Y3 + (size_t)c * n, i) - ldg(T + (size_t)c * n, i);
            acc += (double)r * (double)r;
            gv[c] = gscale * r;
        }
        const uint32_t w0 = (uint32_t)f2bf(gv[0]) | ((uint32_t)f2bf(gv[1]) << 16);
        const uint32_t w1 = (uint32_t)f2bf(gv[2]);
        stg(G3, 2u * i, make_uint4(w0, w1, 0u, 0u));
        stg(G3, 2u * i + 1u, make_uint4(0u, 0u, 0u, 0u));
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ------------------------------------------------------------ Adam
// Kingma & Ba (2015): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
// W -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)
__global__ void k_cnn_adam(float* __restrict__ W, const float* __restrict__ gr, float* __restrict__ m,
                           float* __restrict__ v, int count, float lr, float b1, float b2, float c1, float c2, float eps)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const float g = gr[e];
    const float mm = b1 * m[e] + (1.0f - b1) * g;
    const float vv = b2 * v[e] + (1.0f - b2) * g * g;
    m[e] = mm;
    v[e] = vv;
    W[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

// ------------------------------------------------------------ field helpers
// D = a - b over the three components ([3][n])
__global__ void k_cnn_sub(uint32_t n3, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ D)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) D[i] = a[i] - b[i];
}

// per-block (max |a - b|, max |b|) over n3 values
__global__ void k_cnn_maxdiff(uint32_t n3, const float* __restrict__ a, const float* __restrict__ b, float2* __restrict__ part)
{
    __shared__ float2 red[256];
    float md = 0.0f, mb = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += gridDim.x * blockDim.x) {
        const float bv = b[i];
        md = fmaxf(md, fabsf(a[i] - bv));
        mb = fmaxf(mb, fabsf(bv));
    }
    red[threadIdx.x] = make_float2(md, mb);
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[threadIdx.x].x = fmaxf(red[threadIdx.x].x, red[threadIdx.x + s].x);
            red[threadIdx.x].y = fmaxf(red[threadIdx.x].y, red[threadIdx.x + s].y);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// ------------------------------------------------------------ launches
hipError_t launch_cnn_sub(size_t n3, const float* a, const float* b, float* D, hipStream_t st)
{
    k_cnn_sub<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>((uint32_t)n3, a, b, D);
    return hipGetLastError();
}

hipError_t launch_cnn_maxdiff(size_t n3, const float* a, const float* b, float2* part, int blocks, hipStream_t st)
{
    k_cnn_maxdiff<<<blocks, 256, 0, st>>>((uint32_t)n3, a, b, part);
    return hipGetLastError();
}

hipError_t launch_cnn_layer(const CnnDims& d, int grid, const uint16_t* in, const uint16_t* apack, int mode,
                            uint16_t* out_bf, float* out_f, const float* href, float scale, hipStream_t st)
{
    const size_t lds = (size_t)(NSTEP * 64 + HALO * 2) * sizeof(uint4);
    k_cnn_layer<<<grid, CNN_TPB, lds, st>>>(d, reinterpret_cast<const uint4*>(in), reinterpret_cast<const uint4*>(apack),
                                           mode, reinterpret_cast<uint4*>(out_bf), out_f, href, scale);
    return hipGetLastError();
}

hipError_t launch_cnn_wgrad(const CnnDims& d, int grid, const uint16_t* X, const uint16_t* Gr, float* partial,
                            float* grad, hipStream_t st)
{
    k_cnn_wgrad<<<grid, CNN_TPB, 0, st>>>(d, X, Gr, partial);
    k_cnn_wreduce<<<(NTAP * 256 + 255) / 256, 256, 0, st>>>(partial, grid, grad);
    return hipGetLastError();
}

hipError_t launch_cnn_pack_weights(const float* W, int transposed, uint16_t* apack, hipStream_t st)
{
    k_cnn_pack_weights<<<(NSTEP * 64 * 8 + 255) / 256, 256, 0, st>>>(W, transposed, apack);
    return hipGetLastError();
}

hipError_t launch_cnn_input(const Dims& d, const float* m, const float* mref, float s_in, uint16_t* X0, hipStream_t st)
{
    k_cnn_input<<<(d.n + 255) / 256, 256, 0, st>>>(d, m, mref, s_in, reinterpret_cast<uint4*>(X0));
    return hipGetLastError();
}

hipError_t launch_cnn_loss(const Dims& d, const float* Y3, const float* T, float s_out, float gscale, uint16_t* G3,
                           double* part, hipStream_t st)
{
    k_cnn_loss<<<(d.n + 255) / 256, 256, 0, st>>>(d, Y3, T, s_out, gscale, reinterpret_cast<uint4*>(G3), part);
    return hipGetLastError();
}

hipError_t launch_cnn_adam(float* W, const float* g, float* m, float* v, int count, float lr, float b1, float b2,
                           float c1, float c2, float eps, hipStream_t st)
{
    k_cnn_adam<<<(count + 255) / 256, 256, 0, st>>>(W, g, m, v, count, lr, b1, b2, c1, c2, eps);
    return hipGetLastError();
}

int cnn_layer_lds_bytes() { return (NSTEP * 64 + HALO * 2) * (int)sizeof(uint4); }

hipError_t configure_lds_cnn(size_t bytes)
{
    return hipFuncSetAttribute((const void*)k_cnn_layer, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace msd
