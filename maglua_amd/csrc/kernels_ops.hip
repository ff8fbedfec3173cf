// Per-operator kernels behind the MagLua operator API (one launch per
// op:apply), FP64 reductions, the seeded spin generator and Philox checks.
// All arrays are FP32 SoA: spins [x|y|z|ms] (4n), fields [hx|hy|hz] (3n).
#include "msd_device.h"

namespace msd {

static constexpr int TPB = 256;

static inline int grid_for(size_t n, int tpb = TPB)
{
    size_t g = (n + tpb - 1) / tpb;
    if (g > (1u << 30)) g = 1u << 30;
    return (int)(g ? g : 1);
}

// --------------------------------------------------------------- utilities
__global__ void k_fill(float* p, float v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st)
{
    k_fill<<<grid_for(n), TPB, 0, st>>>(p, v, n);
    return hipGetLastError();
}

__global__ void k_axpy(float* y, float a, const float* __restrict__ x, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] += a * x[i];
}

hipError_t launch_axpy(float* y, float a, const float* x, size_t n, hipStream_t st)
{
    k_axpy<<<grid_for(n), TPB, 0, st>>>(y, a, x, n);
    return hipGetLastError();
}

// Anisotropy.Uniaxial (cpu/anisotropy/spinoperationanisotropy_uniaxial.cpp:248-298):
// one thread per site walks that site's ops in the order they were added,
//   d = sx nx + sy ny + sz nx (the reference's line 287 pairs sz with nx),
//   h += n (2 K1 d + 4 K2 d^3) scale,  s = spin / ms;
// every site's slot value is written (zero without ops).  ops: [nops][5] = nx ny nz K1 K2.
__global__ void k_uniaxial_op(Dims d, const int* __restrict__ start, const float* __restrict__ ops, float scale,
                              const float* __restrict__ s, float* __restrict__ h)
{
    const size_t n = d.n;
    const UniDev u{start, ops, scale};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float3 v = uniaxial_field(u, (int)i, make_float3(s[i], s[n + i], s[2 * n + i]), s[3 * n + i]);
        h[i] = v.x; h[n + i] = v.y; h[2 * n + i] = v.z;
    }
}

hipError_t launch_uniaxial_op(const Dims& d, const int* start, const float* ops, float scale, const float* s, float* h,
                              hipStream_t st)
{
    k_uniaxial_op<<<grid_for(d.n), TPB, 0, st>>>(d, start, ops, scale, s, h);
    return hipGetLastError();
}

// SpinSystem::sumFields: total = sum of the used slots, in slot order
struct SrcList { const float* p[16]; };
__global__ void k_sum_fields(float* total, SrcList src, int nsrc, size_t n3)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        float acc = 0.0f;
        for (int k = 0; k < nsrc; ++k) acc = acc + src.p[k][i];
        total[i] = acc;
    }
}

hipError_t launch_sum_fields(float* total, const float* const* src, int nsrc, size_t n3, hipStream_t st)
{
    if (nsrc < 0 || nsrc > 16) return hipErrorInvalidValue;
    SrcList l;
    for (int k = 0; k < nsrc; ++k) l.p[k] = src[k];
    k_sum_fields<<<grid_for(n3), TPB, 0, st>>>(total, l, nsrc, n3);
    return hipGetLastError();
}

// ---------------------------------------------------------------- exchange
// Exchange::apply (cpu/exchange/spinoperationexchange.cpp:134-213), gathered by
// target over a CSR list kept in path insertion order; accumulates into h.
__global__ void k_exchange_op(Dims d, ExchangeDev ex, const float* __restrict__ s, float* __restrict__ h)
{
    const int n = d.n;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const float mt = s[3 * n + t];
        float hx = h[t], hy = h[n + t], hz = h[2 * n + t];
        for (int j = ex.row[t]; j < ex.row[t + 1]; ++j) {
            const int f = ex.col[j];
            const float mf = s[3 * n + f];
            if (mf == 0.0f) continue;
            const float J = ex.val[j];
            const float den = ex.normalize ? mf * mt : mf;
            hx += s[f] * J / den;
            hy += s[n + f] * J / den;
            hz += s[2 * n + f] * J / den;
        }
        h[t] = hx; h[n + t] = hy; h[2 * n + t] = hz;
    }
}

hipError_t launch_exchange_op(const Dims& d, const ExchangeDev& ex, const float* s, float* h, hipStream_t st)
{
    k_exchange_op<<<grid_for(d.n), TPB, 0, st>>>(d, ex, s, h);
    return hipGetLastError();
}

// -------------------------------------------------------------- anisotropy
// Anisotropy::apply (cpu/anisotropy/spinoperationanisotropy.cpp:247-284): overwrites the slot
__global__ void k_anisotropy_op(Dims d, AnisoDev an, const float* __restrict__ s, float* __restrict__ h)
{
    const int n = d.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float3 v = aniso_field(an, i, n, make_float3(s[i], s[n + i], s[2 * n + i]), s[3 * n + i]);
        h[i] = v.x; h[n + i] = v.y; h[2 * n + i] = v.z;
    }
}

hipError_t launch_anisotropy_op(const Dims& d, const AnisoDev& an, const float* s, float* h, hipStream_t st)
{
    k_anisotropy_op<<<grid_for(d.n), TPB, 0, st>>>(d, an, s, h);
    return hipGetLastError();
}

// ----------------------------------------------------------------- thermal
// Thermal::apply (cpu/thermal/spinoperationthermal.cpp:135-191): h = sigma * N(0,1)^3
__global__ void k_thermal_op(Dims d, ThermalDev th, LLGDev lp, float dt, uint64_t draw, const float* __restrict__ s,
                             float* __restrict__ h)
{
    const int n = d.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float sg = thermal_sigma(th, lp, i, s[3 * n + i], dt);
        float3 g = make_float3(0.f, 0.f, 0.f);
        if (sg != 0.0f) g = thermal_normals(th.seed, draw, i);
        h[i] = sg * g.x; h[n + i] = sg * g.y; h[2 * n + i] = sg * g.z;
    }
}

hipError_t launch_thermal_op(const Dims& d, const ThermalDev& th, const LLGDev& lp, float dt, uint64_t draw,
                             const float* s, float* h, hipStream_t st)
{
    k_thermal_op<<<grid_for(d.n), TPB, 0, st>>>(d, th, lp, dt, draw, s, h);
    return hipGetLastError();
}

__global__ void k_thermal_normals(Dims d, uint64_t seed, uint64_t draw, float* out)
{
    const int n = d.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float3 g = thermal_normals(seed, draw, i);
        out[i] = g.x; out[n + i] = g.y; out[2 * n + i] = g.z;
    }
}

hipError_t launch_thermal_normals(const Dims& d, uint64_t seed, uint64_t draw, float* out, hipStream_t st)
{
    k_thermal_normals<<<grid_for(d.n), TPB, 0, st>>>(d, seed, draw, out);
    return hipGetLastError();
}

// --------------------------------------------------------------------- LLG
// LLGQuaternion::apply at one site (cpu/llg/llgquat.cpp:127-373): S turned by the
// unit quaternion of quat_rot, as the product q (0,S) q*.
__device__ __forceinline__ float3 quat_step(const LLGDev& lp, float3 S, float3 M, float3 H, float3 h, float al, float ga,
                                            float ms, float dt)
{
    float qw;
    float3 qv;
    quat_rot(lp, S, M, H, h, al, ga, ms, dt, qw, qv);
    // t = q (0,S); result = vector part of t q*
    const float tw = -(qv.x * S.x + qv.y * S.y + qv.z * S.z);
    const float3 qs = cross3(qv, S);
    const float3 tv = make_float3(qw * S.x + qs.x, qw * S.y + qs.y, qw * S.z + qs.z);
    const float3 c = cross3(tv, qv);
    return make_float3(-tw * qv.x + qw * tv.x - c.x, -tw * qv.y + qw * tv.y - c.y, -tw * qv.z + qw * tv.z - c.z);
}

// LLGCartesian::apply (cpu/llg/llgcartesian.cpp:90-278) or, with lp.quaternion,
// LLGQuaternion::apply.  from/dmdt/to may alias; every site is read before it is
// written by the same thread.
__global__ void k_llg_op(Dims d, LLGDev lp, float dt, const float* from, const float* dmdt, const float* H,
                         const float* Hth, float* to)
{
    const int n = d.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float ms = from[3 * n + i];
        float3 v;
        if (lp.quaternion) {
            if (ms > 0.0f) {
                const float3 S = make_float3(from[i], from[n + i], from[2 * n + i]);
                const float3 M = make_float3(dmdt[i], dmdt[n + i], dmdt[2 * n + i]);
                const float3 Hv = make_float3(H[i], H[n + i], H[2 * n + i]);
                float3 hv = Hv;
                if (Hth) hv = make_float3(Hv.x - Hth[i], Hv.y - Hth[n + i], Hv.z - Hth[2 * n + i]);
                const float al = lp.site_alpha ? lp.site_alpha[i] : lp.alpha;
                const float ga = lp.site_gamma ? lp.site_gamma[i] : lp.gamma;
                v = quat_step(lp, S, M, Hv, hv, al, ga, ms, dt);
            } else {
                v = make_float3(to[i], to[n + i], to[2 * n + i]);
            }
            to[3 * n + i] = ms;
        } else if (dt == 0.0f) {
            v = make_float3(from[i], from[n + i], from[2 * n + i]);
        } else {
            if (ms > 0.0f) {
                const float3 M = make_float3(dmdt[i], dmdt[n + i], dmdt[2 * n + i]);
                const float3 Hv = make_float3(H[i], H[n + i], H[2 * n + i]);
                float3 hv = Hv;
                if (lp.thermal_first && Hth) hv = make_float3(Hv.x - Hth[i], Hv.y - Hth[n + i], Hv.z - Hth[2 * n + i]);
                const float al = lp.site_alpha ? lp.site_alpha[i] : lp.alpha;
                const float ga = lp.site_gamma ? lp.site_gamma[i] : lp.gamma;
                const float3 dM = lp.thermal_first ? llg_rhs(M, Hv, hv, al, ga, ms, lp.ps, lp.ds)
                                                   : llg_rhs(M, Hv, Hv, al, ga, ms, 1.0f, 1.0f);
                v = make_float3(from[i] + dt * dM.x, from[n + i] + dt * dM.y, from[2 * n + i] + dt * dM.z);
            } else {
                v = make_float3(to[i], to[n + i], to[2 * n + i]);
            }
            to[3 * n + i] = ms;
        }
        if (lp.renorm) v = renorm(v, ms);
        to[i] = v.x; to[n + i] = v.y; to[2 * n + i] = v.z;
    }
}

hipError_t launch_llg_op(const Dims& d, const LLGDev& lp, float dt, const float* from, const float* dmdt,
                         const float* H, const float* Hth, float* to, hipStream_t st)
{
    k_llg_op<<<grid_for(d.n), TPB, 0, st>>>(d, lp, dt, from, dmdt, H, Hth, to);
    return hipGetLastError();
}

// ---------------------------------------------------- seeded initial spins
// Unit spins uniform on the sphere: z = 2u1-1, phi = 2 pi u2 from 53-bit
// uniforms of Philox4x32-10 keyed by seed, counter (site, 0, ~0, 'SPIN').
__global__ void k_random_spins(Dims d, uint64_t seed, float* s, float* lo)
{
    const int n = d.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint4 r = philox4x32_10(make_uint4((uint32_t)i, 0u, 0xFFFFFFFFu, 0x5350494Eu),
                                      make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
        const double u1 = ((double)(r.x >> 5) * 67108864.0 + (double)(r.y >> 6) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((double)(r.z >> 5) * 67108864.0 + (double)(r.w >> 6) + 0.5) * (1.0 / 9007199254740992.0);
        const double zc = 2.0 * u1 - 1.0;
        const double rho = sqrt(fmax(0.0, 1.0 - zc * zc));
        double sp, cp;
        sincospi(2.0 * u2, &sp, &cp);
        const double x = rho * cp, y = rho * sp;
        s[i] = (float)x; s[n + i] = (float)y; s[2 * n + i] = (float)zc;
        lo[i] = (float)(x - (double)(float)x);
        lo[n + i] = (float)(y - (double)(float)y);
        lo[2 * n + i] = (float)(zc - (double)(float)zc);
        s[3 * n + i] = (float)sqrt(x * x + y * y + zc * zc);
    }
}

hipError_t launch_random_spins(const Dims& d, uint64_t seed, float* s, float* lo, hipStream_t st)
{
    k_random_spins<<<grid_for(d.n), TPB, 0, st>>>(d, seed, s, lo);
    return hipGetLastError();
}

__global__ void k_philox(const uint32_t* ctr, int count, uint32_t k0, uint32_t k1, uint32_t* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint4 r = philox4x32_10(make_uint4(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]), make_uint2(k0, k1));
    out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
}

hipError_t launch_philox(const uint32_t* ctr, int count, uint32_t k0, uint32_t k1, uint32_t* out, hipStream_t st)
{
    k_philox<<<grid_for(count), TPB, 0, st>>>(ctr, count, k0, k1, out);
    return hipGetLastError();
}

// ------------------------------------------------------- FP64 reductions
// Two-stage, fixed-order reductions (deterministic): per-block partials, then
// one block folds the partials.
static constexpr int RED_TPB = 256;

int reduce_blocks(const Dims& d)
{
    int g = (d.n + RED_TPB * 8 - 1) / (RED_TPB * 8);
    return g < 1 ? 1 : (g > 1024 ? 1024 : g);
}

template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* part)
{
    __shared__ double sh[K][RED_TPB];
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = RED_TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < K; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) part[(size_t)blockIdx.x * K + k] = sh[k][0];
    }
}

template <int K>
__global__ void k_fold(const double* part, int nb, double* out)
{
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += part[(size_t)b * K + k];
    block_sum<K>(v, out);
}

// spin component c of site i with its optional low word (compensated state)
__device__ __forceinline__ double spin_at(const float* s, const float* lo, int c, int n, int i)
{
    const double v = s[(size_t)c * n + i];
    return lo ? v + (double)lo[(size_t)c * n + i] : v;
}

__global__ void k_moment(Dims d, const float* __restrict__ s, const float* __restrict__ lo, double* part)
{
    double v[6] = {0, 0, 0, 0, 0, 0};
    const int n = d.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double x = spin_at(s, lo, c, n, i);
            v[c] += x;
            v[3 + c] += x * x;
        }
    }
    block_sum<6>(v, part);
}

__global__ void k_dot(Dims d, const float* __restrict__ h, const float* __restrict__ s, const float* __restrict__ lo,
                      double* part)
{
    double v[1] = {0};
    const int n = d.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[0] += (double)h[(size_t)c * n + i] * spin_at(s, lo, c, n, i);
    block_sum<1>(v, part);
}

__global__ void k_diff(Dims d, const float* __restrict__ a, const float* __restrict__ alo, const float* __restrict__ b,
                       const float* __restrict__ blo, double* part)
{
    double v[3] = {0, 0, 0};
    const int n = d.n;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] += fabs(spin_at(a, alo, c, n, i) - spin_at(b, blo, c, n, i));
    block_sum<3>(v, part);
}

hipError_t launch_reduce_moment(const Dims& d, const float* s, const float* lo, double* part, double* out, hipStream_t st)
{
    const int nb = reduce_blocks(d);
    k_moment<<<nb, RED_TPB, 0, st>>>(d, s, lo, part);
    k_fold<6><<<1, RED_TPB, 0, st>>>(part, nb, out);
    return hipGetLastError();
}

hipError_t launch_reduce_dot(const Dims& d, const float* h, const float* s, const float* lo, double* part, double* out,
                             hipStream_t st)
{
    const int nb = reduce_blocks(d);
    k_dot<<<nb, RED_TPB, 0, st>>>(d, h, s, lo, part);
    k_fold<1><<<1, RED_TPB, 0, st>>>(part, nb, out);
    return hipGetLastError();
}

hipError_t launch_reduce_diff(const Dims& d, const float* a, const float* alo, const float* b, const float* blo,
                              double* part, double* out, hipStream_t st)
{
    const int nb = reduce_blocks(d);
    k_diff<<<nb, RED_TPB, 0, st>>>(d, a, alo, b, blo, part);
    k_fold<3><<<1, RED_TPB, 0, st>>>(part, nb, out);
    return hipGetLastError();
}

}  // namespace msd
