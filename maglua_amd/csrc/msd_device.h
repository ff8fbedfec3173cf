// Device-side building blocks for the gfx950 kernels: counter-based normals,
// mixed-radix LDS FFT stages, LLG arithmetic.
#pragma once

#include "msd_internal.h"

namespace msd {

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of two 32x32->64
// multiplies with Weyl key increments.  Each product is one v_mad_u64_u32 (both
// halves at once) rather than a separate low and high multiply.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k.x += 0x9E3779B9u; k.y += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k.x, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k.y, (uint32_t)p0);
    }
    return c;
}

// top 24 bits -> (0,1), exact in FP32
__device__ __forceinline__ float u24f(uint32_t a) { return ((float)(a >> 8) + 0.5f) * (1.0f / 16777216.0f); }

// Three standard normals for (seed, draw, site): Box-Muller on the four words,
// r = sqrt(-2 ln u0), (r cos 2 pi u1, r sin 2 pi u1), r' cos 2 pi u3 with r' from u2.
// The transcendentals are the single hardware instructions: v_log_f32 (log2),
// v_sqrt_f32, and v_sin_f32 / v_cos_f32, whose argument is in revolutions.
__device__ __forceinline__ float3 thermal_normals(uint64_t seed, uint64_t draw, int site)
{
    const uint4 r = philox4x32_10(make_uint4((uint32_t)site, 0u, (uint32_t)draw, (uint32_t)(draw >> 32)),
                                  make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
    const float m2ln2 = -1.38629436111989061883f;   // -2 ln 2: -2 ln u = -2 ln2 log2 u
    const float r0 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u24f(r.x)));
    const float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u24f(r.z)));
    const float a0 = u24f(r.y), a1 = u24f(r.w);
    return make_float3(r0 * __builtin_amdgcn_cosf(a0), r0 * __builtin_amdgcn_sinf(a0), r1 * __builtin_amdgcn_cosf(a1));
}

// ------------------------------------------------------ address spaces
// Address-space views.  A pointer that a kernel reads out of a device table (an
// ensemble's argument records, a batch's replica descriptors, an LDS-staged class
// table) reaches the compiler as a generic pointer, and an access through it is a
// flat instruction: flat loads count against both the vector-memory and the LDS
// wait counters, so every LDS wait then also waits for the global loads in flight.
// Accesses through gview / lview are global_* and ds_* instructions.
#define MSD_GLOBAL __attribute__((address_space(1)))
#define MSD_LDS __attribute__((address_space(3)))
template <typename T> __device__ __forceinline__ const MSD_GLOBAL T* gview(const T* p) { return (const MSD_GLOBAL T*)p; }
template <typename T> __device__ __forceinline__ MSD_GLOBAL T* gview(T* p) { return (MSD_GLOBAL T*)p; }
template <typename T> __device__ __forceinline__ const MSD_LDS T* lview(const T* p) { return (const MSD_LDS T*)p; }

// Element types move through these views as builtin scalars or clang vectors (the
// HIP vector types' members take generic references only).
template <typename T> struct RawOf {
    using type = T;
    static __device__ __forceinline__ T get(type v) { return v; }
    static __device__ __forceinline__ type put(T v) { return v; }
};
template <typename U, unsigned N> struct RawOf<HIP_vector_type<U, N>> {
    using T = HIP_vector_type<U, N>;
    using type = typename HIP_vector_base<U, N>::Native_vec_;
    static __device__ __forceinline__ T get(type v) { T t; __builtin_memcpy(&t, &v, sizeof t); return t; }
    static __device__ __forceinline__ type put(T v) { type r; __builtin_memcpy(&r, &v, sizeof r); return r; }
};

// element i of a global array, through the global view
template <typename T>
__device__ __forceinline__ T gload(const T* p, size_t i)
{
    using R = typename RawOf<T>::type;
    return RawOf<T>::get(reinterpret_cast<const MSD_GLOBAL R*>(gview(p))[i]);
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, size_t i, T v)
{
    using R = typename RawOf<T>::type;
    reinterpret_cast<MSD_GLOBAL R*>(gview(p))[i] = RawOf<T>::put(v);
}


// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) { return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }  // a*conj(b)
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by SIGN*i
template <int SIGN> __device__ __forceinline__ float2 mul_si(float2 a)
{
    return SIGN > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// In-register DFT of size R, V[q] = sum_r v[r] exp(SIGN 2 pi i r q / R).
template <int SIGN> __device__ __forceinline__ void dft2(float2* v)
{
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
}
template <int SIGN> __device__ __forceinline__ void dft4(float2* v)
{
    const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const float2 t2 = cadd(v[1], v[3]), t3 = mul_si<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
}
template <int SIGN> __device__ __forceinline__ void dft8(float2* v)
{
    const float h = 0.70710678118654752440f;
    float2 a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) { a[r] = cadd(v[r], v[r + 4]); b[r] = csub(v[r], v[r + 4]); }
    // b[r] *= exp(SIGN 2 pi i r / 8)
    b[1] = make_float2(h * (b[1].x - SIGN * b[1].y), h * (b[1].y + SIGN * b[1].x));
    b[2] = mul_si<SIGN>(b[2]);
    b[3] = make_float2(h * (-b[3].x - SIGN * b[3].y), h * (-b[3].y + SIGN * b[3].x));
    dft4<SIGN>(a);
    dft4<SIGN>(b);
#pragma unroll
    for (int m = 0; m < 4; ++m) { v[2 * m] = a[m]; v[2 * m + 1] = b[m]; }
}
template <int SIGN> __device__ __forceinline__ void dft3(float2* v)
{
    const float c = -0.5f, s = 0.86602540378443864676f;
    const float2 sum = cadd(v[1], v[2]);
    const float2 dif = mul_si<SIGN>(csub(v[1], v[2]));
    const float2 base = make_float2(v[0].x + c * sum.x, v[0].y + c * sum.y);
    v[0] = cadd(v[0], sum);
    v[1] = make_float2(base.x + s * dif.x, base.y + s * dif.y);
    v[2] = make_float2(base.x - s * dif.x, base.y - s * dif.y);
}
template <int SIGN> __device__ __forceinline__ void dft5(float2* v)
{
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const float2 s14 = cadd(v[1], v[4]), d14 = csub(v[1], v[4]);
    const float2 s23 = cadd(v[2], v[3]), d23 = csub(v[2], v[3]);
    const float2 a1 = make_float2(v[0].x + c1 * s14.x + c2 * s23.x, v[0].y + c1 * s14.y + c2 * s23.y);
    const float2 a2 = make_float2(v[0].x + c2 * s14.x + c1 * s23.x, v[0].y + c2 * s14.y + c1 * s23.y);
    const float2 b1 = mul_si<SIGN>(make_float2(s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y));
    const float2 b2 = mul_si<SIGN>(make_float2(s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y));
    v[0] = make_float2(v[0].x + s14.x + s23.x, v[0].y + s14.y + s23.y);
    v[1] = cadd(a1, b1); v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2); v[3] = csub(a2, b2);
}

// One Stockham stage of radix R over `count` lines of length L held
// contiguously in LDS (src -> dst).  tw[k] = exp(-2 pi i k / L).
template <int R, int SIGN>
__device__ __forceinline__ void stockham_stage(const float2* __restrict__ src, float2* __restrict__ dst, int count, int L,
                                               int Ns, const float2* __restrict__ tw, int tid, int nthr)
{
    const int nb = L / R;
    const int total = count * nb;
    const int tstep = L / (Ns * R);
    for (int b = tid; b < total; b += nthr) {
        const int seq = b / nb;
        const int j = b - seq * nb;
        const int k = j % Ns;
        const float2* s = src + seq * L + j;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = s[r * nb];
        if (k) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const float2 w = tw[k * r * tstep];
                v[r] = SIGN < 0 ? cmul(v[r], w) : cmulc(v[r], w);
            }
        }
        if (R == 2) dft2<SIGN>(v);
        else if (R == 3) dft3<SIGN>(v);
        else if (R == 4) dft4<SIGN>(v);
        else if (R == 5) dft5<SIGN>(v);
        else dft8<SIGN>(v);
        float2* d = dst + seq * L + (j / Ns) * Ns * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) d[q * Ns] = v[q];
    }
}

// Full transform of `count` lines; returns the buffer holding the result.
// Caller must have synchronised after filling `a`; returns after a barrier.
template <int SIGN>
__device__ float2* lds_fft(float2* a, float2* b, int count, const FFTPlan& p, const float2* __restrict__ tw)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    int Ns = 1;
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.rad[s];
        switch (R) {
            case 2: stockham_stage<2, SIGN>(a, b, count, p.L, Ns, tw, tid, nthr); break;
            case 3: stockham_stage<3, SIGN>(a, b, count, p.L, Ns, tw, tid, nthr); break;
            case 4: stockham_stage<4, SIGN>(a, b, count, p.L, Ns, tw, tid, nthr); break;
            case 5: stockham_stage<5, SIGN>(a, b, count, p.L, Ns, tw, tid, nthr); break;
            default: stockham_stage<8, SIGN>(a, b, count, p.L, Ns, tw, tid, nthr); break;
        }
        Ns *= R;
        __syncthreads();
        float2* t = a; a = b; b = t;
    }
    return a;
}

// ------------------------------------------- compile-time Stockham transforms
// Same stages as lds_fft with the line length, line count and thread count known
// at compile time: index arithmetic folds to shifts and masks, loops unroll, and
// twiddles come from a table already staged in LDS.
__host__ __device__ constexpr int next_radix(int rem)
{
    return rem % 8 == 0 ? 8 : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 3 == 0 ? 3 : 5;
}

template <int L, int COUNT, int NT, int SIGN, int NS, int R>
__device__ __forceinline__ void stage_ct(const float2* __restrict__ src, float2* __restrict__ dst,
                                         const float2* __restrict__ tw)
{
    constexpr int nb = L / R;
    constexpr int total = COUNT * nb;
    constexpr int iters = (total + NT - 1) / NT;
    constexpr int tstep = L / (NS * R);
#pragma unroll
    for (int it = 0; it < iters; ++it) {
        const int b = (int)threadIdx.x + it * NT;
        if (total % NT == 0 || b < total) {
            const int seq = b / nb;
            const int j = b - seq * nb;
            const int k = j % NS;
            const float2* s = src + seq * L + j;
            float2 v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = s[r * nb];
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 w = tw[k * r * tstep];
                    v[r] = SIGN < 0 ? cmul(v[r], w) : cmulc(v[r], w);
                }
            }
            if (R == 2) dft2<SIGN>(v);
            else if (R == 3) dft3<SIGN>(v);
            else if (R == 4) dft4<SIGN>(v);
            else if (R == 5) dft5<SIGN>(v);
            else dft8<SIGN>(v);
            float2* d = dst + seq * L + (j / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) d[q * NS] = v[q];
        }
    }
}

template <int L, int COUNT, int NT, int SIGN, int NS>
struct FFTct {
    static __device__ __forceinline__ float2* run(float2* a, float2* b, const float2* tw)
    {
        constexpr int R = next_radix(L / NS);
        stage_ct<L, COUNT, NT, SIGN, NS, R>(a, b, tw);
        __syncthreads();
        return FFTct<L, COUNT, NT, SIGN, NS * R>::run(b, a, tw);
    }
};
template <int L, int COUNT, int NT, int SIGN>
struct FFTct<L, COUNT, NT, SIGN, L> {
    static __device__ __forceinline__ float2* run(float2* a, float2*, const float2*) { return a; }
};

// COUNT lines of length L in LDS (a), result buffer returned; twl = LDS twiddles
template <int L, int COUNT, int NT, int SIGN>
__device__ __forceinline__ float2* fft_ct(float2* a, float2* b, const float2* twl)
{
    return FFTct<L, COUNT, NT, SIGN, 1>::run(a, b, twl);
}

// In-place form: every thread holds its butterflies' inputs in registers across a
// barrier, so one LDS buffer serves all stages (half the LDS of the ping-pong form).
//
// Between stages the line is kept in a swizzled order, element a at
// a ^ ((a >> 3) & 31) (a permutation of each power-of-two line of 32 or more).
// Unswizzled, the first stage's writes (a thread's R outputs at R consecutive
// places) and the second's (stride NS = 8) land lanes on a few LDS banks
// (8-way and 4-way conflicts); the swizzle spreads both over all banks and
// keeps the stride-1 reads conflict-free.  The first stage reads and the last
// stage writes the natural order, so callers never see the swizzle.
template <int L>
__device__ __forceinline__ int lds_swz(int a)
{
    if constexpr (L >= 32 && (L & (L - 1)) == 0) return a ^ ((a >> 3) & 31);
    else return a;
}

// twiddle i from LDS (a generic pointer into it) or from global memory (its view)
__device__ __forceinline__ float2 twv(const float2* t, int i) { return t[i]; }
__device__ __forceinline__ float2 twv(const MSD_GLOBAL float2* t, int i)
{
    return RawOf<float2>::get(reinterpret_cast<const MSD_GLOBAL RawOf<float2>::type*>(t)[i]);
}

template <int L, int COUNT, int NT, int SIGN, int NS, int R, bool SWIN, bool SWOUT, int LS, typename TWP>
__device__ __forceinline__ void stage_ip(float2* __restrict__ buf, TWP tw)
{
    constexpr int nb = L / R;
    constexpr int total = COUNT * nb;
    constexpr int iters = (total + NT - 1) / NT;
    constexpr int tstep = L / (NS * R);
    float2 v[iters][R];
#pragma unroll
    for (int it = 0; it < iters; ++it) {
        const int b = (int)threadIdx.x + it * NT;
        if (total % NT == 0 || b < total) {
            const int seq = b / nb;
            const int j = b - seq * nb;
            const float2* s = buf + seq * LS;
#pragma unroll
            for (int r = 0; r < R; ++r) v[it][r] = s[SWIN ? lds_swz<L>(j + r * nb) : j + r * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < iters; ++it) {
        const int b = (int)threadIdx.x + it * NT;
        if (total % NT == 0 || b < total) {
            const int seq = b / nb;
            const int j = b - seq * nb;
            const int k = j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 w = twv(tw, k * r * tstep);
                    v[it][r] = SIGN < 0 ? cmul(v[it][r], w) : cmulc(v[it][r], w);
                }
            }
            if (R == 2) dft2<SIGN>(v[it]);
            else if (R == 3) dft3<SIGN>(v[it]);
            else if (R == 4) dft4<SIGN>(v[it]);
            else if (R == 5) dft5<SIGN>(v[it]);
            else dft8<SIGN>(v[it]);
            float2* d = buf + seq * LS;
            const int o = (j / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) d[SWOUT ? lds_swz<L>(o + q * NS) : o + q * NS] = v[it][q];
        }
    }
    __syncthreads();
}

template <int L, int COUNT, int NT, int SIGN, int NS, int LS>
struct FFTip {
    template <typename TWP>
    static __device__ __forceinline__ void run(float2* a, TWP tw)
    {
        constexpr int R = next_radix(L / NS);
        stage_ip<L, COUNT, NT, SIGN, NS, R, NS != 1, NS * R != L, LS>(a, tw);
        FFTip<L, COUNT, NT, SIGN, NS * R, LS>::run(a, tw);
    }
};
template <int L, int COUNT, int NT, int SIGN, int LS>
struct FFTip<L, COUNT, NT, SIGN, L, LS> {
    template <typename TWP>
    static __device__ __forceinline__ void run(float2*, TWP) {}
};

// COUNT lines of length L transformed in place in LDS, line s at a + s*LS (a
// stride LS > L staggers the lines' banks for transposing copies); caller
// synchronised before.  twl: twiddles in LDS, or the global view of a table.
template <int L, int COUNT, int NT, int SIGN, int LS = L, typename TWP = const float2*>
__device__ __forceinline__ void fft_ip(float2* a, TWP twl)
{
    FFTip<L, COUNT, NT, SIGN, 1, LS>::run(a, twl);
}

// ------------------------------------------------------- staged copies
// This thread's share (idx = threadIdx.x + k*NT) of TOT items: ld(idx) for CH
// items at a time, every load of a chunk issued before the first st(idx, v).
// A loop that loads, waits and stores pays one memory latency per trip.
template <int TOT, int NT, int CH, typename LD, typename ST>
__device__ __forceinline__ void batched_copy(LD ld, ST st)
{
    constexpr int IT = (TOT + NT - 1) / NT;
    using V = decltype(ld(0));
#pragma unroll
    for (int c0 = 0; c0 < IT; c0 += CH) {
        V v[CH];
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int it = c0 + k, idx = (int)threadIdx.x + it * NT;
            if (it < IT && (TOT % NT == 0 || idx < TOT)) v[k] = ld(idx);
        }
#pragma unroll
        for (int k = 0; k < CH; ++k) {
            const int it = c0 + k, idx = (int)threadIdx.x + it * NT;
            if (it < IT && (TOT % NT == 0 || idx < TOT)) st(idx, v[k]);
        }
    }
}

// ------------------------------------------------------ site addressing
// Per-site arrays are addressed as a uniform base plus a 32-bit byte offset
// (the global_load saddr form: one VALU op per address instead of a 64-bit
// multiply-add).  Every array the step kernels index this way holds fewer than
// 2^30 elements (the system size limit checked at creation).
template <typename T>
__device__ __forceinline__ T ldg(const T* __restrict__ base, uint32_t i)
{
    using R = typename RawOf<T>::type;
    return RawOf<T>::get(*reinterpret_cast<const MSD_GLOBAL R*>(reinterpret_cast<const MSD_GLOBAL char*>(gview(base)) +
                                                                (uint64_t)(i * (uint32_t)sizeof(T))));
}
template <typename T>
__device__ __forceinline__ void stg(T* __restrict__ base, uint32_t i, T v)
{
    using R = typename RawOf<T>::type;
    *reinterpret_cast<MSD_GLOBAL R*>(reinterpret_cast<MSD_GLOBAL char*>(gview(base)) + (uint64_t)(i * (uint32_t)sizeof(T))) =
        RawOf<T>::put(v);
}

// Hardware reciprocal and reciprocal square root (v_rcp_f32 / v_rsq_f32, 1 ulp);
// the correctly rounded IEEE division sequence costs ~10 VALU ops per use.
__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsqf(float x) { return __builtin_amdgcn_rsqf(x); }

// ---------------------------------------------------------------- LLG
__device__ __forceinline__ float3 cross3(float3 a, float3 b)
{
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// dM = ps*(-gaa) M x H - ds*(alpha/ms) gaa M x (M x h)   (LLGCartesian with thermalOnlyFirstTerm)
__device__ __forceinline__ float3 llg_rhs(float3 M, float3 H, float3 h, float alpha, float gamma, float ms, float ps, float ds)
{
    const float gaa = gamma * rcpf(1.0f + alpha * alpha);
    const float3 Mh = cross3(M, h);
    const float3 MMh = cross3(M, Mh);
    const float3 MH = cross3(M, H);
    const float damp = ds * alpha * rcpf(ms) * gaa;
    const float prec = ps * -gaa;
    return make_float3(prec * MH.x - damp * MMh.x, prec * MH.y - damp * MMh.y, prec * MH.z - damp * MMh.z);
}

// LLGQuaternion::apply (cpu/llg/llgquat.cpp:127-373) at one site: S turns about
// A = ps H + ds a (M x h)/ms (thermalOnlyFirstTerm; h = H - H_thermal) or
// A = H + a (S x H)/ms by the angle g |A| dt / (1 + a^2).  quat_rot gives the unit
// quaternion (qw, qv) of that turn: (1, 0) when A vanishes.
__device__ __forceinline__ void quat_rot(const LLGDev& lp, float3 S, float3 M, float3 H, float3 h, float al, float ga,
                                         float ms, float dt, float& qw, float3& qv)
{
    float3 A;
    if (lp.thermal_first) {
        const float3 Mh = cross3(M, h);
        const float f = lp.ds * al / ms;
        A = make_float3(lp.ps * H.x + f * Mh.x, lp.ps * H.y + f * Mh.y, lp.ps * H.z + f * Mh.z);
    } else {
        const float3 SH = cross3(S, H);
        const float f = al / ms;
        A = make_float3(H.x + f * SH.x, H.y + f * SH.y, H.z + f * SH.z);
    }
    const float len = sqrtf(A.x * A.x + A.y * A.y + A.z * A.z);
    if (!(len > 0.0f)) { qw = 1.0f; qv = make_float3(0.f, 0.f, 0.f); return; }
    float sn;
    sincosf(0.5f * ga * len / (1.0f + al * al) * dt, &sn, &qw);
    const float k = sn / len;
    qv = make_float3(k * A.x, k * A.y, k * A.z);
}

// The turn (qw, qv) applied to S, as the increment R S - S = 2 qw (qv x S) +
// 2 qv x (qv x S) (Rodrigues), so that a compensated state takes it as one more
// addend, as it takes dt dM of LLG.Cartesian.
__device__ __forceinline__ float3 quat_delta(float qw, float3 qv, float3 S)
{
    const float3 t = cross3(qv, S);
    const float3 u = cross3(qv, t);
    return make_float3(2.0f * (qw * t.x + u.x), 2.0f * (qw * t.y + u.y), 2.0f * (qw * t.z + u.z));
}

// -------------------------------------------- compensated (hi + lo) state
// Error-free transformations (Knuth TwoSum, FMA-based product) keep the spin
// state to ~1e-14 while every operation stays FP32.
__device__ __forceinline__ void df_add(float& hi, float& lo, float b)
{
#pragma clang fp contract(off)
    const float s = hi + b;
    const float bb = s - hi;
    const float e = ((hi - (s - bb)) + (b - bb)) + lo;
    hi = s + e;
    lo = e - (hi - s);
}

__device__ __forceinline__ void df_scale(float& hi, float& lo, float f)
{
#pragma clang fp contract(off)
    const float p = hi * f;
    const float e = fmaf(lo, f, fmaf(hi, f, -p));
    hi = p + e;
    lo = e - (hi - p);
}

// rescale (hi + lo) to length ms; the length is set to FP32 accuracy, the
// direction is kept to the accuracy of the compensated pair
__device__ __forceinline__ void df_renorm(float3& hi, float3& lo, float ms)
{
    const float mm = hi.x * hi.x + hi.y * hi.y + hi.z * hi.z + 2.0f * (hi.x * lo.x + hi.y * lo.y + hi.z * lo.z);
    if (mm != 0.0f) {
        const float inv = ms * rsqf(mm);
        df_scale(hi.x, lo.x, inv);
        df_scale(hi.y, lo.y, inv);
        df_scale(hi.z, lo.z, inv);
    }
}

__device__ __forceinline__ float3 renorm(float3 v, float ms)
{
    const float mm = v.x * v.x + v.y * v.y + v.z * v.z;
    if (mm != 0.0f) {
        const float inv = ms * rsqf(mm);
        v.x *= inv; v.y *= inv; v.z *= inv;
    }
    return v;
}

// ------------------------------------------------------------ fields
// Anisotropy field (cpu/anisotropy/spinoperationanisotropy.cpp:247-284) as
// h = 2 A s / ms^2 with A = sum over the site's ops of K n n^T.
// Anisotropy.Uniaxial at site i (cpu/anisotropy/spinoperationanisotropy_uniaxial.cpp:248-298):
// the site's ops in order of addition, h += n (2 K1 d + 4 K2 d^3) scale with
// d = sx nx + sy ny + sz nx (the reference's line 287 pairs sz with nx), s = spin / ms
__device__ __forceinline__ float3 uniaxial_field(const UniDev& u, int i, float3 s, float ms)
{
    float3 h = make_float3(0.f, 0.f, 0.f);
    if (!u.ops || !(ms > 0.0f)) return h;
    const float inv = 1.0f / ms;
    const float sx = s.x * inv, sy = s.y * inv, sz = s.z * inv;
    for (int j = u.start[i]; j < u.start[i + 1]; ++j) {
        const float* o = u.ops + 5 * (size_t)j;
        const float dd = sx * o[0] + sy * o[1] + sz * o[0];
        const float mag = (2.0f * o[3] * dd + 4.0f * o[4] * dd * dd * dd) * u.scale;
        h.x += o[0] * mag; h.y += o[1] * mag; h.z += o[2] * mag;
    }
    return h;
}

__device__ __forceinline__ float3 aniso_field(const AnisoDev& an, int i, int n, float3 s, float ms)
{
    if (!(ms > 0.0f) || an.mode == 0) return make_float3(0.f, 0.f, 0.f);
    float3 h;
    const float f = 2.0f * rcpf(ms * ms);
    if (an.mode == 2) {
        const float4 a = ldg(an.axk, (uint32_t)i);
        const float v = f * a.w * (s.x * a.x + s.y * a.y + s.z * a.z);
        h = make_float3(a.x * v, a.y * v, a.z * v);
    } else {
        float m[6];
        if (an.mode == 1) {
#pragma unroll
            for (int k = 0; k < 6; ++k) m[k] = an.u[k];
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) m[k] = ldg(an.mat + (size_t)k * n, (uint32_t)i);
        }
        h = make_float3(f * (m[0] * s.x + m[1] * s.y + m[2] * s.z),
                        f * (m[1] * s.x + m[3] * s.y + m[4] * s.z),
                        f * (m[2] * s.x + m[4] * s.y + m[5] * s.z));
    }
    return h;
}

// Thermal field amplitude (cpu/thermal/spinoperationthermal.cpp:165-178)
__device__ __forceinline__ float thermal_sigma(const ThermalDev& th, const LLGDev& lp, int i, float ms, float dt)
{
    const float al = lp.site_alpha ? ldg(lp.site_alpha, (uint32_t)i) : lp.alpha;
    const float ga = lp.site_gamma ? ldg(lp.site_gamma, (uint32_t)i) : lp.gamma;
    if (ms == 0.0f || th.T_scale == 0.0f || !(ga > 0.0f)) return 0.0f;
    const float sc = th.site ? ldg(th.site, (uint32_t)i) : 1.0f;
    return sqrtf((2.0f * al * th.T_scale * sc) * rcpf(ms * dt * ga));
}

// Exchange field at target i (cpu/exchange/spinoperationexchange.cpp:134-213):
// sum_j J_j s_f / m_f over the site's couplings, then / m_i when normalising.
// Class-table entries carry the source as a linear offset f - i.  The field is
// built in two parts so a caller can issue one site's loads (exchange_gather)
// before it finishes another (exchange_finish): the first 8 couplings of a
// class are loaded by the gather, further couplings and CSR rows by the finish.
struct ExGather {
    float vx[8], vy[8], vz[8], mf[8];
    int e0, cnt;
};

__device__ __forceinline__ void exchange_gather(const ExchangeDev& ex, uint32_t n, uint32_t i, const float* __restrict__ s,
                                                const float* __restrict__ m, ExGather& g)
{
    g.e0 = 0;
    g.cnt = 0;
    if (ex.mode != 1) return;
    // mode 1 is the fused kernels' LDS-staged class table
    const MSD_LDS int* cs = lview(ex.cls_start);
    const MSD_LDS ExEntry* en = lview(ex.ent);
    const int c = ex.cls ? (int)ldg(ex.cls, i) : 0;
    g.e0 = cs[c];
    g.cnt = cs[c + 1] - g.e0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        g.mf[u] = 0.0f;
        if (u < g.cnt) {
            const uint32_t f = i + (uint32_t)en[g.e0 + u].dl;
            g.mf[u] = ldg(m, f);
            g.vx[u] = ldg(s, f); g.vy[u] = ldg(s + n, f); g.vz[u] = ldg(s + 2 * (size_t)n, f);
        }
    }
}

__device__ __forceinline__ float3 exchange_finish(const ExchangeDev& ex, uint32_t n, uint32_t i, const float* __restrict__ s,
                                                  const float* __restrict__ m, float mi, const ExGather& g)
{
    const float* __restrict__ sx = s;
    const float* __restrict__ sy = s + n;
    const float* __restrict__ sz = s + 2 * (size_t)n;
    float3 h = make_float3(0.f, 0.f, 0.f);
    if (ex.mode == 1) {
        const MSD_LDS ExEntry* en = lview(ex.ent);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (g.mf[u] != 0.0f) {
                const float w = en[g.e0 + u].J * rcpf(g.mf[u]);
                h.x += g.vx[u] * w; h.y += g.vy[u] * w; h.z += g.vz[u] * w;
            }
        }
        for (int e = g.e0 + 8; e < g.e0 + g.cnt; ++e) {
            const int dl = en[e].dl;
            const float J = en[e].J;
            const uint32_t f = i + (uint32_t)dl;
            const float mfe = ldg(m, f);
            if (mfe != 0.0f) {
                const float w = J * rcpf(mfe);
                h.x += ldg(sx, f) * w; h.y += ldg(sy, f) * w; h.z += ldg(sz, f) * w;
            }
        }
    } else if (ex.mode == 2) {
        const int j1 = ldg(ex.row, i + 1u);
        for (int j = ldg(ex.row, i); j < j1; ++j) {
            const uint32_t f = (uint32_t)ldg(ex.col, (uint32_t)j);
            const float mfe = ldg(m, f);
            if (mfe != 0.0f) {
                const float w = ldg(ex.val, (uint32_t)j) * rcpf(mfe);
                h.x += ldg(sx, f) * w; h.y += ldg(sy, f) * w; h.z += ldg(sz, f) * w;
            }
        }
    }
    if (ex.normalize && ex.mode) {
        const float inv = rcpf(mi);
        h.x *= inv; h.y *= inv; h.z *= inv;
    }
    return h;
}

// ------------------------------------------------------ Heun site updates
// Each half of the step at a site is split the same way: *_gather issues the
// site's loads into a SiteGather, *_finish computes and stores.
struct SiteGather {
    float3 v;       // predictor: S; corrector: P
    float3 c;       // corrector: C
    float3 lo;      // predictor: S_lo; corrector: C_lo
    float ms;
    ExGather ex;
};

// Predictor at site i (PredictorCorrector.lua:77-79 with LLGCartesian::apply):
// H = exchange + anisotropy + Hd at S; P = N(S + dt dM), C = N(S + dt/2 dM).
template <bool COMP>
__device__ __forceinline__ void pred_gather(const StepArgs& a, const ExchangeDev& ex, int i, SiteGather& g)
{
    const uint32_t n = (uint32_t)a.d.n, u = (uint32_t)i;
    const float* __restrict__ S = a.S;
    g.ms = ldg(S + 3 * (size_t)n, u);
    g.v = make_float3(ldg(S, u), ldg(S + n, u), ldg(S + 2 * (size_t)n, u));
    if (COMP) g.lo = make_float3(ldg(a.Slo, u), ldg(a.Slo + n, u), ldg(a.Slo + 2 * (size_t)n, u));
    exchange_gather(ex, n, u, S, S + 3 * (size_t)n, g.ex);
}

// Writes P, C (and C's low word when COMP); returns P for the next transform.
template <bool COMP>
__device__ __forceinline__ float3 pred_finish(const StepArgs& a, const ExchangeDev& ex, int i, float3 Hd, const SiteGather& g)
{
    const uint32_t n = (uint32_t)a.d.n, u = (uint32_t)i;
    const float ms = g.ms;
    const float3 s = g.v;
    float3 clo = COMP ? g.lo : make_float3(0.f, 0.f, 0.f);
    float3 H = exchange_finish(ex, n, u, a.S, a.S + 3 * (size_t)n, ms, g.ex);
    const float3 ha = aniso_field(a.an, i, (int)n, s, ms);
    H.x += ha.x + Hd.x; H.y += ha.y + Hd.y; H.z += ha.z + Hd.z;
    if (a.uni.ops) { const float3 hu = uniaxial_field(a.uni, i, s, ms); H.x += hu.x; H.y += hu.y; H.z += hu.z; }
    if (a.Hext) { H.x += ldg(a.Hext, u); H.y += ldg(a.Hext + n, u); H.z += ldg(a.Hext + 2 * (size_t)n, u); }
    if (a.Href) { H.x += ldg(a.Href, u); H.y += ldg(a.Href + n, u); H.z += ldg(a.Href + 2 * (size_t)n, u); }
    if (a.has_app) {
        H.x += a.Bapp.x; H.y += a.Bapp.y; H.z += a.Bapp.z;
        if (a.Happ) { H.x += ldg(a.Happ, u); H.y += ldg(a.Happ + n, u); H.z += ldg(a.Happ + 2 * (size_t)n, u); }
    }
    const float dt = a.dt, hdt = 0.5f * a.dt;
    float3 p = s, c = s;
    if (ms > 0.0f) {
        const float al = a.lp.site_alpha ? ldg(a.lp.site_alpha, u) : a.lp.alpha;
        const float ga = a.lp.site_gamma ? ldg(a.lp.site_gamma, u) : a.lp.gamma;
        float3 dp, dc;      // P - S and C - S
        if (a.lp.quaternion) {
            float qw;
            float3 qv;
            quat_rot(a.lp, s, s, H, H, al, ga, ms, dt, qw, qv);
            dp = quat_delta(qw, qv, s);
            quat_rot(a.lp, s, s, H, H, al, ga, ms, hdt, qw, qv);
            dc = quat_delta(qw, qv, s);
        } else {
            const float3 dM = a.lp.thermal_first ? llg_rhs(s, H, H, al, ga, ms, a.lp.ps, a.lp.ds)
                                                 : llg_rhs(s, H, H, al, ga, ms, 1.0f, 1.0f);
            dp = make_float3(dt * dM.x, dt * dM.y, dt * dM.z);
            dc = make_float3(hdt * dM.x, hdt * dM.y, hdt * dM.z);
        }
        p = make_float3(s.x + dp.x, s.y + dp.y, s.z + dp.z);
        if (COMP) {
            df_add(c.x, clo.x, dc.x);
            df_add(c.y, clo.y, dc.y);
            df_add(c.z, clo.z, dc.z);
        } else {
            c = make_float3(s.x + dc.x, s.y + dc.y, s.z + dc.z);
        }
    }
    if (a.lp.renorm) {
        p = renorm(p, ms);
        if (COMP) df_renorm(c, clo, ms);
        else c = renorm(c, ms);
    }
    stg(a.P, u, p.x); stg(a.P + n, u, p.y); stg(a.P + 2 * (size_t)n, u, p.z);
    stg(a.C, u, c.x); stg(a.C + n, u, c.y); stg(a.C + 2 * (size_t)n, u, c.z);
    if (COMP) { stg(a.Clo, u, clo.x); stg(a.Clo + n, u, clo.y); stg(a.Clo + 2 * (size_t)n, u, clo.z); }
    return p;
}

template <bool COMP>
__device__ __forceinline__ float3 site_pred(const StepArgs& a, const ExchangeDev& ex, int i, float3 Hd)
{
    SiteGather g;
    pred_gather<COMP>(a, ex, i, g);
    return pred_finish<COMP>(a, ex, i, Hd, g);
}

// Corrector and thermal kick at site i (PredictorCorrector.lua:81-85,98-103):
// S' = N(C + dt/2 dM(P, H(P))), then S'' = N(S' + dt LLG(S', H_th)) when Thermal
// is on.
template <bool COMP>
__device__ __forceinline__ void corr_gather(const StepArgs& a, const ExchangeDev& ex, int i, SiteGather& g)
{
    const uint32_t n = (uint32_t)a.d.n, u = (uint32_t)i;
    const float* __restrict__ P = a.P;
    g.ms = ldg(a.S + 3 * (size_t)n, u);
    g.v = make_float3(ldg(P, u), ldg(P + n, u), ldg(P + 2 * (size_t)n, u));
    g.c = make_float3(ldg(a.C, u), ldg(a.C + n, u), ldg(a.C + 2 * (size_t)n, u));
    if (COMP) g.lo = make_float3(ldg(a.Clo, u), ldg(a.Clo + n, u), ldg(a.Clo + 2 * (size_t)n, u));
    exchange_gather(ex, n, u, P, a.S + 3 * (size_t)n, g.ex);
}

// Writes S (and its low word when COMP); returns S''.
template <bool COMP>
__device__ __forceinline__ float3 corr_finish(const StepArgs& a, const ExchangeDev& ex, int i, float3 Hd, uint64_t draw,
                                              const SiteGather& g)
{
    const uint32_t n = (uint32_t)a.d.n, u = (uint32_t)i;
    const float ms = g.ms;
    const float3 p = g.v;
    float3 v = g.c;
    float3 vlo = COMP ? g.lo : make_float3(0.f, 0.f, 0.f);
    float3 H = exchange_finish(ex, n, u, a.P, a.S + 3 * (size_t)n, ms, g.ex);
    const float3 ha = aniso_field(a.an, i, (int)n, p, ms);
    H.x += ha.x + Hd.x; H.y += ha.y + Hd.y; H.z += ha.z + Hd.z;
    if (a.uni.ops) { const float3 hu = uniaxial_field(a.uni, i, p, ms); H.x += hu.x; H.y += hu.y; H.z += hu.z; }
    if (a.Hext) { H.x += ldg(a.Hext, u); H.y += ldg(a.Hext + n, u); H.z += ldg(a.Hext + 2 * (size_t)n, u); }
    if (a.Href) { H.x += ldg(a.Href, u); H.y += ldg(a.Href + n, u); H.z += ldg(a.Href + 2 * (size_t)n, u); }
    if (a.has_app) {
        H.x += a.Bapp.x; H.y += a.Bapp.y; H.z += a.Bapp.z;
        if (a.Happ) { H.x += ldg(a.Happ, u); H.y += ldg(a.Happ + n, u); H.z += ldg(a.Happ + 2 * (size_t)n, u); }
    }
    const float dt = a.dt, hdt = 0.5f * a.dt;
    const float al = a.lp.site_alpha ? ldg(a.lp.site_alpha, u) : a.lp.alpha;
    const float ga = a.lp.site_gamma ? ldg(a.lp.site_gamma, u) : a.lp.gamma;
    if (ms > 0.0f) {
        float3 dv;          // S' - C
        if (a.lp.quaternion) {
            // C turned about the axis of P's field (LLG:apply(C, dt/2, P, C))
            float qw;
            float3 qv;
            quat_rot(a.lp, v, p, H, H, al, ga, ms, hdt, qw, qv);
            dv = quat_delta(qw, qv, v);
        } else {
            const float3 dM = a.lp.thermal_first ? llg_rhs(p, H, H, al, ga, ms, a.lp.ps, a.lp.ds)
                                                 : llg_rhs(p, H, H, al, ga, ms, 1.0f, 1.0f);
            dv = make_float3(hdt * dM.x, hdt * dM.y, hdt * dM.z);
        }
        if (COMP) {
            df_add(v.x, vlo.x, dv.x);
            df_add(v.y, vlo.y, dv.y);
            df_add(v.z, vlo.z, dv.z);
        } else {
            v = make_float3(v.x + dv.x, v.y + dv.y, v.z + dv.z);
        }
    }
    if (a.lp.renorm) {
        if (COMP) df_renorm(v, vlo, ms);
        else v = renorm(v, ms);
    }
    if (a.th.on) {
        const float sg = thermal_sigma(a.th, a.lp, i, ms, dt);
        if (ms > 0.0f) {
            float3 Ht = make_float3(0.f, 0.f, 0.f);
            if (sg != 0.0f) {
                const float3 gn = thermal_normals(a.th.seed, draw, i);
                Ht = make_float3(sg * gn.x, sg * gn.y, sg * gn.z);
            }
            const float3 zero = make_float3(0.f, 0.f, 0.f);
            float3 dv;      // the kick: only the thermal field acts (h = H - H_thermal = 0)
            if (a.lp.quaternion) {
                float qw;
                float3 qv;
                quat_rot(a.lp, v, v, Ht, a.lp.thermal_first ? zero : Ht, al, ga, ms, dt, qw, qv);
                dv = quat_delta(qw, qv, v);
            } else {
                const float3 dM = a.lp.thermal_first ? llg_rhs(v, Ht, zero, al, ga, ms, a.lp.ps, a.lp.ds)
                                                     : llg_rhs(v, Ht, Ht, al, ga, ms, 1.0f, 1.0f);
                dv = make_float3(dt * dM.x, dt * dM.y, dt * dM.z);
            }
            if (COMP) {
                df_add(v.x, vlo.x, dv.x);
                df_add(v.y, vlo.y, dv.y);
                df_add(v.z, vlo.z, dv.z);
            } else {
                v = make_float3(v.x + dv.x, v.y + dv.y, v.z + dv.z);
            }
        }
        if (a.lp.renorm) {
            if (COMP) df_renorm(v, vlo, ms);
            else v = renorm(v, ms);
        }
    }
    stg(a.S, u, v.x); stg(a.S + n, u, v.y); stg(a.S + 2 * (size_t)n, u, v.z);
    if (COMP) { stg(a.Slo, u, vlo.x); stg(a.Slo + n, u, vlo.y); stg(a.Slo + 2 * (size_t)n, u, vlo.z); }
    return v;
}

template <bool COMP>
__device__ __forceinline__ float3 site_corr(const StepArgs& a, const ExchangeDev& ex, int i, float3 Hd, uint64_t draw)
{
    SiteGather g;
    corr_gather<COMP>(a, ex, i, g);
    return corr_finish<COMP>(a, ex, i, Hd, draw, g);
}

}  // namespace msd
