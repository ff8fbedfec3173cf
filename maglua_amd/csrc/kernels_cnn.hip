// Dipolar-field surrogate network on gfx950: three 5x5x3 convolution layers
// (3 -> 16 -> 16 -> 3 channels, no bias, identity activation), BF16 operands on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  The network has no
// counterpart in the reference; it predicts the change of the layered dipole
// field (LongRange::apply, cpu/longrange/spinoperationlongrange.cpp:354-430)
// between exact refreshes.
//
// Activations are channels-last BF16: 16 channels (32 bytes) per site for the
// hidden layers, 4 (8 bytes, the 4th zero) for the 3-channel network input and
// output gradient.  A layer is
//   Y[s][o] = sum_t sum_i W[t][o][i] X[s + off(t)][i],
//   off(t) = (dx, dy, dz) = (t%5 - 2, (t/5)%5 - 2, t/25 - 1),
// periodic in x and y (or, with CnnDims::open, zero outside the lattice in x and y
// too), zero outside the lattice in z.
//
// Row walk.  A wave owns 16 XF output sites along x of one layer z and walks the
// input rows y' of its tile in order.  Input row y' feeds output rows y' - dy,
// dy = -2..2, so every B fragment read from LDS feeds five MFMAs (one per dy, into
// five rotating accumulators) and the A fragments (the weights) live in registers
// for the whole kernel.  Per 16 sites and input row:
//   CV_F16  (16 -> 16): K = 2 dx x 16 channels, 9 k-steps (dz, dx pair) x 5 dy;
//   CV_IN3  ( 3 -> 16): K = 8 (dx, dz) taps x 4 channels, 2 k-steps x 5 dy;
//   CV_OUT3 (16 ->  3): M = 15 (dy, channel) rows, 9 k-steps, one accumulator
//                       whose rows are added into an LDS buffer of output rows.
// That is 64 MFMAs per 16 sites for the network (114 with 16-channel padding of
// every layer and one MFMA per B fragment).
//
// MFMA operand layout (16x16x32, checked on the device): lane l holds
// A[l%16][8(l/16) .. +7], B[8(l/16) .. +7][l%16], and D[4(l/16) .. +3][l%16].
#include "msd_device.h"

#include <type_traits>

namespace msd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

static constexpr int CNN_TPB = 256;                  // 4 waves
static constexpr int CTX = 32, CTY = 8;              // weight-gradient tile: 32 x, 8 rows, one layer
static constexpr int HX = CTX + 4, HY = CTY + 4, HZ = 3;
static constexpr int HALO = HX * HY * HZ;            // input sites staged per weight-gradient tile
static constexpr int NTAP = 75;
static constexpr int RXF = 2;                        // 16-site fragments per wave in the row walk
static constexpr int RWX = 64 * RXF;                 // sites along x per row-walk workgroup
static constexpr int RXW = RWX + 4;                  // staged sites per input row and layer

template <int V> struct CnnShape {
    static constexpr int NF = V == CV_F16 ? 45 : V == CV_IN3 ? 10 : 9;   // A fragments
    static constexpr bool NARROW = V == CV_IN3;                          // 8-byte input sites
};

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);                 // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// channels 4g..4g+3 of an accumulator as 4 BF16 (v_cvt_pk_bf16_f32: round to nearest
// even, the rounding of f2bf for finite values)
__device__ __forceinline__ uint2 pack_bf16x4(const f32x4& a)
{
    const bf16x2 lo = __builtin_convertvector((f32x2){a[0], a[1]}, bf16x2);
    const bf16x2 hi = __builtin_convertvector((f32x2){a[2], a[3]}, bf16x2);
    return make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
}

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

// ------------------------------------------------------------ layer (row walk)
// Persistent: workgroups take items (RWX sites along x, wk.ty output rows, one
// layer) in turn and walk the input rows y0-2 .. y0+ny_out+1 of the item.  Each
// input row goes straight from global memory into one of five LDS slots
// (global_load_lds_dwordx4, no registers held for rows in flight), WALK_RPF rows
// ahead of the row being multiplied (row j + 4 reuses the slot of row j - 1, free
// once every wave passed the barrier closing row j - 1).  The slots are separate arrays selected at
// compile time, so the compiler sees that the LDS reads of the current row do not
// alias the loads in flight; the barrier closing a row waits (vmcnt) only for the
// loads of the next row: every wave issues at least WALK_GUAR(V) loads per row,
// and the loads of later rows are younger.
//
// Every row issues all five dy products: an accumulator slot is cleared when its
// output row completes, so the products of the first rows that belong to rows
// before y0 never reach a stored row, and the last rows' products for rows past
// the tile are dropped.  The dz layers outside the lattice are skipped at compile
// time (four specialisations of the walk) where the taps of a k-step share one
// layer (CV_F16, CV_OUT3); the 3-channel input (CV_IN3, taps of two layers per
// k-step) stages zeros for them instead, read from a zeroed device buffer.
//   CV_F16, CV_IN3: BF16 [n][16] output;
//   CV_OUT3: FP32 SoA output[c][n] = scale * Y[c].  The product of input row j
//   has rows (ky, c), ky = dy + 2: lane group g (rows 4g..4g+3) holds ky = g for
//   c = 0..2 in its registers 0..2 and, for g < 3, (ky = 4, c = g) in register 3.
//   The products of the last five rows stay in registers; output row o = j - 4
//   takes ky = g from row o + g in each lane group (a select over the ring), and
//   the four lane groups are summed with two cross-lane exchanges.  (The reference
//   field a prediction adds to this is added by its consumer: a global load whose
//   result is used in the same row would hold the walk.)
static constexpr int WALK_RPF = 4;                   // rows in flight (five slots >= WALK_RPF + 1)
typedef __attribute__((address_space(1))) void gvoid_t;
typedef __attribute__((address_space(3))) void lvoid_t;

template <int V> struct WalkShape {
    static constexpr bool NARROW = V == CV_IN3;
    // uint4 per staged input row: 3 layers x RXW sites, 8 or 32 bytes a site
    static constexpr int ROWE = NARROW ? 3 * RXW / 2 : 3 * RXW * 2;
    static constexpr int SLOT = ROWE;
    // loads every wave issues per staged row (elements 256 k + tid, k < GUAR, exist in all waves)
    static constexpr int GUAR = NARROW ? 1 : 3;
    static_assert(NARROW ? (ROWE > 3 * 64 && ROWE <= 256) : (ROWE >= 768 && ROWE <= 1024), "staging split");
};

// x summed over the four 16-lane rows of the wave (lanes l, l^16, l^32, l^48), in
// every lane: v_permlane16_swap / v_permlane32_swap with both operands x exchange
// rows 2i <-> 2i+1 (then halves) between the two results, so their sum pairs them
// (VALU; a ds_bpermute would make each step an LDS round trip)
__device__ __forceinline__ float lane_rows_sum(float x)
{
    const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x), false, false);
    const float y = __uint_as_float(p[0]) + __uint_as_float(p[1]);
    const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(y), __float_as_uint(y), false, false);
    return __uint_as_float(q[0]) + __uint_as_float(q[1]);
}

// s_waitcnt vmcnt(N) leaving the LDS and export counters alone (gfx9 encoding)
template <int N> __device__ __forceinline__ void wait_vm()
{
    static_assert(N >= 0 && N < 64, "vmcnt");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// At the end of row j (RPF rows in flight), wait for the loads of row j + 1 and
// no others.  vmcnt counts loads and stores and completes them in issue order, so
// the wait may leave outstanding exactly the operations issued after row j + 1's
// loads:
//   - the loads of rows j + 2 .. min(j + RPF, nrows - 1), >= GUAR each;
//   - the output stores of rows j + 1 - RPF .. j (each row issues its next-row
//     loads before its stores), NST store instructions in a row whose output row
//     exists (4 <= row < nout + 4) when the wave stores at all.
// Rows in the middle of the walk take the full count; the first and last rows
// wait with the loads alone, a lower bound.  GUAR = 0 (a wave that issues no load
// in some row) waits for everything.
template <int RPF, int GUAR, int NST>
__device__ __forceinline__ void wait_next_row(int j, int nrows, int nout, bool stores)
{
    if (GUAR == 0) { wait_vm<0>(); return; }
    const int later = min(j + RPF, nrows - 1) - (j + 1);
    if (stores && later == RPF - 1 && j - RPF + 1 >= 4 && j < nout + 4) {
        wait_vm<(RPF - 1) * GUAR + RPF * NST>();
        return;
    }
    if (RPF >= 4 && later >= 3) wait_vm<3 * GUAR>();
    else if (later >= 2) wait_vm<2 * GUAR>();
    else if (later >= 1) wait_vm<GUAR>();
    else wait_vm<0>();
}

template <int N> using IC = std::integral_constant<int, N>;

template <int V>
__global__ void __launch_bounds__(CNN_TPB) k_cnn_walk(CnnDims d, CnnWalk wk, const uint4* __restrict__ in,
                                                   const uint4* __restrict__ zeros, const uint4* __restrict__ apack,
                                                   uint2* __restrict__ out_bf, float* __restrict__ out_f, float scale)
{
    using Sh = CnnShape<V>;
    using WS = WalkShape<V>;
    constexpr int XF = RXF, XW = RXW, WX = RWX, RPF = WALK_RPF;
    constexpr int ROWE = WS::ROWE, SLOT = WS::SLOT, GUAR = WS::GUAR;
    constexpr int NK = (ROWE + CNN_TPB - 1) / CNN_TPB;  // staging loads per thread and row
    constexpr bool NARROW = WS::NARROW;
    __shared__ uint4 st0[SLOT], st1[SLOT], st2[SLOT], st3[SLOT], st4[SLOT];
    auto slot = [&](auto sl) -> uint4* {
        constexpr int S = decltype(sl)::value;
        return S == 0 ? st0 : S == 1 ? st1 : S == 2 ? st2 : S == 3 ? st3 : st4;
    };
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, col = l & 15, g = l >> 4;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const uint32_t n = (uint32_t)d.n;
    bf16x8 wa[Sh::NF];
#pragma unroll
    for (int i = 0; i < Sh::NF; ++i) wa[i] = as_bf8(apack[i * 64 + l]);
    // per-lane B offsets inside a staged row: uint2 (CV_IN3) or uint4 elements of fragment 0
    const int cb = 16 * XF * w + col;
    int boff[4];
    if (NARROW) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {                    // k = 2 ks + q: tap slot 8 ks + 2 g + q
            const int c = min(8 * (k >> 1) + 2 * g + (k & 1), 14);
            boff[k] = (c / 5) * XW + c % 5 + cb;
        }
    } else {
        boff[0] = (cb + (g >> 1)) * 2 + (g & 1);        // dx pairs (-2,-1) and (0,1): + 4 pair
        boff[1] = (cb + 4) * 2 + (g & 1);               // pair (2, -): dx = 2 for both halves
        boff[2] = boff[3] = 0;
    }

    for (int item = blockIdx.x; item < wk.rtiles; item += gridDim.x) {
        const int xt = item % wk.rtiles_x;
        const int rr = item / wk.rtiles_x;
        const int z = rr / wk.rtiles_y;
        const int y0 = (rr - z * wk.rtiles_y) * wk.ty;
        const int x0 = xt * WX;
        const int nout = min(wk.ty, ny - y0);
        const int nrows = nout + 4;
        const bool wv = x0 + 16 * XF * w < nx;           // this wave's sites exist (nx is a multiple of 16 XF)
        // staged element e = tid + 256 k of a row, in uint4 units of the row of `in`;
        // -1: a layer outside the lattice (CV_IN3 reads zeros; the others read the
        // nearest layer, whose taps they skip)
        int gsrc[NK];
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            const int e = tid + k * CNN_TPB;
            if (NARROW) {                                // two 8-byte sites per element
                const int dzi = e / (XW / 2), q = e - dzi * (XW / 2);
                const int zz = z + dzi - 1, xx = x0 - 2 + 2 * q;   // nx even: a pair is inside or outside
                const bool in_x = !d.open || (unsigned)xx < (unsigned)nx;
                gsrc[k] = (zz >= 0 && zz < nz && in_x) ? ((zz * ny) * nx + wrap(xx, nx)) / 2 : -1;
            } else {
                const int site = e >> 1, dzi = site / XW, xs = site - dzi * XW;
                const int zz = min(max(z + dzi - 1, 0), nz - 1), xx = x0 + xs - 2;
                const bool in_x = !d.open || (unsigned)xx < (unsigned)nx;
                gsrc[k] = in_x ? ((zz * ny) * nx + wrap(xx, nx)) * 2 + (e & 1) : -1;
            }
        }
        const int rstride = NARROW ? nx / 2 : nx * 2;    // uint4 per layer row of `in`
        auto load_row = [&](auto sl, int j) __attribute__((always_inline)) {
            uint4* dst = slot(sl) + 64 * w;
            const int yy = y0 - 2 + j;
            const bool in_y = !d.open || (unsigned)yy < (unsigned)ny;
            const int yo = wrap(yy, ny) * rstride;
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const int e = tid + k * CNN_TPB;
                if (e < ROWE) {                          // lanes past the row idle; every wave keeps >= GUAR loads
                    const uint4* src = (gsrc[k] >= 0 && in_y) ? in + gsrc[k] + yo : zeros;
                    __builtin_amdgcn_global_load_lds((gvoid_t*)src, (lvoid_t*)(dst + k * CNN_TPB), 16, 0, 0);
                }
            }
        };
        // CV_F16, CV_IN3: output row o accumulates in acc[o % 5]; CV_OUT3: acc[j % 5]
        // holds the product of input row j
        f32x4 acc[5][XF];
#pragma unroll
        for (int s = 0; s < 5; ++s)
#pragma unroll
            for (int f = 0; f < XF; ++f) acc[s][f] = f32x4{0.f, 0.f, 0.f, 0.f};

        auto walk = [&](auto dzm) __attribute__((always_inline)) {
            constexpr int DZM = decltype(dzm)::value;    // bit dzi: layer z + dzi - 1 exists
            // the previous item's rows are consumed (every wave passed its last barrier)
            load_row(IC<0>{}, 0);
            if (1 < nrows) load_row(IC<1>{}, 1);
            if (2 < nrows) load_row(IC<2>{}, 2);
            if (3 < nrows) load_row(IC<3>{}, 3);
            if (3 < nrows) wait_vm<3 * GUAR>();          // row 0 (rows 1..3 are younger)
            else wait_vm<0>();
            __builtin_amdgcn_s_barrier();
            auto row = [&](auto ph, int j) __attribute__((always_inline)) {
                constexpr int PH = decltype(ph)::value;
                if (j >= nrows) return;
                // Phase 0 follows the loop back-edge, where the compiler no longer
                // knows which slots the loads in flight target and waits for all of
                // them before the first LDS read: its row load goes after that read.
                constexpr int DZ0 = (DZM & 1) ? 0 : 1;
                const bool issue = j + RPF < nrows;
                if (PH != 0 && issue) load_row(IC<(PH + RPF) % 5>{}, j + RPF);
                const uint4* S = slot(IC<PH>{});
                bool pending = PH == 0 && issue;
                if (wv) {
                    if (V == CV_IN3) {
                        const uint2* S2 = reinterpret_cast<const uint2*>(S);
#pragma unroll
                        for (int ks = 0; ks < 2; ++ks) {
                            bf16x8 b[XF];
#pragma unroll
                            for (int f = 0; f < XF; ++f) {
                                const uint2 lo = S2[boff[2 * ks] + 16 * f];
                                const uint2 hi = S2[boff[2 * ks + 1] + 16 * f];
                                b[f] = as_bf8(make_uint4(lo.x, lo.y, hi.x, hi.y));
                            }
                            if (ks == 0 && pending) { load_row(IC<RPF % 5>{}, j + RPF); pending = false; }
#pragma unroll
                            for (int dyi = 0; dyi < 5; ++dyi)
#pragma unroll
                                for (int f = 0; f < XF; ++f) {
                                    f32x4& a = acc[(PH - dyi + 10) % 5][f];
                                    a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[dyi * 2 + ks], b[f], a, 0, 0, 0);
                                }
                        }
                    } else {
                        f32x4 (&c3)[XF] = acc[PH];
                        if (V == CV_OUT3) {
#pragma unroll
                            for (int f = 0; f < XF; ++f) c3[f] = f32x4{0.f, 0.f, 0.f, 0.f};
                        }
#pragma unroll
                        for (int dzi = 0; dzi < 3; ++dzi) {
                            if (!((DZM >> dzi) & 1)) continue;
#pragma unroll
                            for (int pair = 0; pair < 3; ++pair) {
                                bf16x8 b[XF];
#pragma unroll
                                for (int f = 0; f < XF; ++f)
                                    b[f] = as_bf8(S[dzi * XW * 2 + 32 * f + (pair < 2 ? boff[0] + 4 * pair : boff[1])]);
                                if (dzi == DZ0 && pair == 0 && pending) { load_row(IC<RPF % 5>{}, j + RPF); pending = false; }
                                if (V == CV_F16) {
#pragma unroll
                                    for (int dyi = 0; dyi < 5; ++dyi)
#pragma unroll
                                        for (int f = 0; f < XF; ++f) {
                                            f32x4& a = acc[(PH - dyi + 10) % 5][f];
                                            a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                                wa[(dyi * 3 + dzi) * 3 + pair], b[f], a, 0, 0, 0);
                                        }
                                } else {
#pragma unroll
                                    for (int f = 0; f < XF; ++f)
                                        c3[f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[dzi * 3 + pair], b[f], c3[f],
                                                                                        0, 0, 0);
                                }
                            }
                        }
                        if (V == CV_OUT3) {
                            const int o = j - 4;         // complete after this row
                            if (o >= 0 && o < nout) {
                                const uint32_t site0 = (uint32_t)(x0 + cb + nx * (y0 + o + ny * z));
#pragma unroll
                                for (int f = 0; f < XF; ++f) {
                                    // ky = g from input row o + g (ring slot (PH + 1 + g) % 5)
                                    const f32x4& r0 = acc[(PH + 1) % 5][f];
                                    const f32x4& r1 = acc[(PH + 2) % 5][f];
                                    const f32x4& r2 = acc[(PH + 3) % 5][f];
                                    const f32x4& r3 = acc[(PH + 4) % 5][f];
                                    float q[3];
#pragma unroll
                                    for (int c = 0; c < 3; ++c) {
                                        q[c] = g == 0 ? r0[c] : g == 1 ? r1[c] : g == 2 ? r2[c] : r3[c];
                                        if (c == g) q[c] += acc[PH][f][3];     // (ky = 4, c = g), this row
                                    }
#pragma unroll
                                    for (int c = 0; c < 3; ++c) q[c] = lane_rows_sum(q[c]);
                                    if (g < 3)
                                        stg(out_f + (size_t)g * n, site0 + 16u * f,
                                            scale * (g == 0 ? q[0] : g == 1 ? q[1] : q[2]));
                                }
                            }
                        }
                    }
                    if (V != CV_OUT3) {
                        const int o = j - 4;             // complete after this row
                        if (o >= 0 && o < nout) {
                            const uint32_t site0 = (uint32_t)(x0 + cb + nx * (y0 + o + ny * z));
#pragma unroll
                            for (int f = 0; f < XF; ++f)   // channels 4g..4g+3: 8 bytes at 32 site + 8 g
                                stg(out_bf, (site0 + 16u * f) * 4u + (uint32_t)g, pack_bf16x4(acc[(PH + 1) % 5][f]));
                        }
#pragma unroll
                        for (int f = 0; f < XF; ++f) acc[(PH + 1) % 5][f] = f32x4{0.f, 0.f, 0.f, 0.f};
                    }
                }
                if (pending) load_row(IC<RPF % 5>{}, j + RPF);
                wait_next_row<RPF, GUAR, XF>(j, nrows, nout, wv);
                __builtin_amdgcn_s_barrier();
            };
            for (int j0 = 0; j0 < nrows; j0 += 5) {
                row(IC<0>{}, j0);
                row(IC<1>{}, j0 + 1);
                row(IC<2>{}, j0 + 2);
                row(IC<3>{}, j0 + 3);
                row(IC<4>{}, j0 + 4);
            }
        };
        if (V == CV_IN3) {
            walk(IC<7>{});
        } else {
            const bool lo = z > 0, hi = z + 1 < nz;
            if (lo && hi) walk(IC<7>{});
            else if (hi) walk(IC<6>{});
            else if (lo) walk(IC<3>{});
            else walk(IC<2>{});
        }
    }
}

// ------------------------------------------------------------ film walk (every layer)
// The row walk of a film of NZ <= 8 layers with every layer in one item: a
// workgroup owns FWX = 64 x-sites (16 per wave) and ty rows of all NZ layers.  An
// input row of every layer is staged once (the per-layer walk stages each row for
// the three output layers that read it), and each B fragment feeds every output
// layer it reaches (z = zi - dz) times the five dy: up to 15 MFMAs per LDS read.
// Staging, slots, waits and phases are those of k_cnn_walk.  The 3-channel input
// stages a zero layer below and above the film (written once, never loaded over)
// because its k-steps mix layers; the other shapes skip taps outside the film at
// compile time.
static constexpr int FWX = 64, FXW = FWX + 4;

template <int V, int NZ> struct FilmShape {
    static constexpr bool NARROW = V == CV_IN3;
    static constexpr int NL = NARROW ? NZ + 2 : NZ;        // staged layers
    static constexpr int SITE_B = NARROW ? 8 : 32;         // bytes per staged site
    static constexpr int EB = NARROW ? 4 : 16;             // bytes per load element (2 per site)
    static constexpr int SLOT16 = NL * FXW * SITE_B / 16;  // uint4 per slot
    static constexpr int LOADE = NZ * FXW * 2;             // load elements per row
    static constexpr int NK = (LOADE + CNN_TPB - 1) / CNN_TPB;
    static constexpr int GUAR = LOADE / CNN_TPB;           // rounds in which every wave loads
    static constexpr int ZOFF = NARROW ? FXW * SITE_B / EB : 0;   // first loaded element of a slot
};

template <int V, int NZ>
__global__ void __launch_bounds__(CNN_TPB) k_cnn_film(CnnDims d, CnnWalk wk, const void* __restrict__ in_,
                                                   const uint4* __restrict__ apack, uint2* __restrict__ out_bf,
                                                   float* __restrict__ out_f, float scale)
{
    using Sh = CnnShape<V>;
    using FS = FilmShape<V, NZ>;
    constexpr int RPF = WALK_RPF, GUAR = FS::GUAR, SLOT16 = FS::SLOT16, EB = FS::EB;
    constexpr bool NARROW = FS::NARROW;
    __shared__ uint4 st0[SLOT16], st1[SLOT16], st2[SLOT16], st3[SLOT16], st4[SLOT16];
    auto slot = [&](auto sl) -> uint4* {
        constexpr int S = decltype(sl)::value;
        return S == 0 ? st0 : S == 1 ? st1 : S == 2 ? st2 : S == 3 ? st3 : st4;
    };
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, col = l & 15, g = l >> 4;
    const int nx = d.nx, ny = d.ny;
    const uint32_t n = (uint32_t)d.n;
    const char* in = reinterpret_cast<const char*>(in_);
    if (NARROW) {                                        // the zero layers of every slot
        for (int i = tid; i < FXW / 2; i += CNN_TPB) {
            const uint4 zv = make_uint4(0u, 0u, 0u, 0u);
            st0[i] = zv; st1[i] = zv; st2[i] = zv; st3[i] = zv; st4[i] = zv;
            const int t = (NZ + 1) * FXW / 2 + i;
            st0[t] = zv; st1[t] = zv; st2[t] = zv; st3[t] = zv; st4[t] = zv;
        }
    }
    bf16x8 wa[Sh::NF];
#pragma unroll
    for (int i = 0; i < Sh::NF; ++i) wa[i] = as_bf8(apack[i * 64 + l]);
    const int cb = 16 * w + col;
    int boff[4];
    if (NARROW) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {                    // k = 2 ks + q: tap slot 8 ks + 2 g + q
            const int c = min(8 * (k >> 1) + 2 * g + (k & 1), 14);
            boff[k] = (c / 5) * FXW + c % 5 + cb;       // uint2 units, + z FXW for output layer z
        }
    } else {
        boff[0] = (cb + (g >> 1)) * 2 + (g & 1);        // uint4 units, + 2 FXW zi
        boff[1] = (cb + 4) * 2 + (g & 1);
        boff[2] = boff[3] = 0;
    }
    __syncthreads();

    for (int item = blockIdx.x; item < wk.rtiles; item += gridDim.x) {
        const int xt = item % wk.rtiles_x;
        const int y0 = (item / wk.rtiles_x) * wk.ty;
        const int x0 = xt * FWX;
        const int nout = min(wk.ty, ny - y0);
        const int nrows = nout + 4;
        const bool wv = x0 + 16 * w < nx;
        // load element e = tid + 256 k of a row: (layer, site, half) -> element of `in`
        int gsrc[FS::NK];
#pragma unroll
        for (int k = 0; k < FS::NK; ++k) {
            const int e = min(tid + k * CNN_TPB, FS::LOADE - 1), site = e >> 1;
            const int zl = site / FXW, xs = site - zl * FXW;
            gsrc[k] = ((zl * ny) * nx + wrap(x0 + xs - 2, nx)) * 2 + (e & 1);
        }
        auto load_row = [&](auto sl, int j) __attribute__((always_inline)) {
            char* dst = reinterpret_cast<char*>(slot(sl)) + (FS::ZOFF + 64 * w) * EB;
            const char* src = in + (size_t)wrap(y0 - 2 + j, ny) * nx * 2 * EB;
#pragma unroll
            for (int k = 0; k < FS::NK; ++k)
                if (tid + k * CNN_TPB < FS::LOADE) {
                    gvoid_t* gp = (gvoid_t*)(src + (size_t)gsrc[k] * EB);
                    lvoid_t* lp = (lvoid_t*)(dst + k * CNN_TPB * EB);
                    // literal sizes: the host pass checks the builtin without gfx950's 16-byte form
                    if constexpr (FS::EB == 16) __builtin_amdgcn_global_load_lds(gp, lp, 16, 0, 0);
                    else __builtin_amdgcn_global_load_lds(gp, lp, 4, 0, 0);
                }
        };
        // CV_F16, CV_IN3: output row o of layer z accumulates in acc[z][o % 5];
        // CV_OUT3: acc[z][j % 5] holds the product of input row j for output layer z
        f32x4 acc[NZ][5];
#pragma unroll
        for (int z = 0; z < NZ; ++z)
#pragma unroll
            for (int s2 = 0; s2 < 5; ++s2) acc[z][s2] = f32x4{0.f, 0.f, 0.f, 0.f};

        load_row(IC<0>{}, 0);
        if (1 < nrows) load_row(IC<1>{}, 1);
        if (2 < nrows) load_row(IC<2>{}, 2);
        if (3 < nrows) load_row(IC<3>{}, 3);
        if (GUAR > 0 && 3 < nrows) wait_vm<3 * GUAR>();  // row 0 (rows 1..3 are younger)
        else wait_vm<0>();
        __builtin_amdgcn_s_barrier();
        auto row = [&](auto ph, int j) __attribute__((always_inline)) {
            constexpr int PH = decltype(ph)::value;
            if (j >= nrows) return;
            const bool issue = j + RPF < nrows;
            if (PH != 0 && issue) load_row(IC<(PH + RPF) % 5>{}, j + RPF);
            bool pending = PH == 0 && issue;            // phase 0 loads after its first LDS read
            const uint4* S = slot(IC<PH>{});
            if (wv) {
                if (V == CV_IN3) {
                    const uint2* S2 = reinterpret_cast<const uint2*>(S);
#pragma unroll
                    for (int z = 0; z < NZ; ++z)
#pragma unroll
                        for (int ks = 0; ks < 2; ++ks) {
                            const uint2 lo = S2[boff[2 * ks] + z * FXW];
                            const uint2 hi = S2[boff[2 * ks + 1] + z * FXW];
                            const bf16x8 b = as_bf8(make_uint4(lo.x, lo.y, hi.x, hi.y));
                            if (z == 0 && ks == 0 && pending) { load_row(IC<RPF % 5>{}, j + RPF); pending = false; }
#pragma unroll
                            for (int dyi = 0; dyi < 5; ++dyi) {
                                f32x4& a = acc[z][(PH - dyi + 10) % 5];
                                a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[dyi * 2 + ks], b, a, 0, 0, 0);
                            }
                        }
                } else {
                    if (V == CV_OUT3) {
#pragma unroll
                        for (int z = 0; z < NZ; ++z) acc[z][PH] = f32x4{0.f, 0.f, 0.f, 0.f};
                    }
                    // B of k-step ks = 3 zi + pair, read one k-step ahead of its MFMAs
                    auto bread = [&](int ks) {
                        const int zi = ks / 3, pair = ks - 3 * zi;
                        return as_bf8(S[zi * FXW * 2 + (pair < 2 ? boff[0] + 4 * pair : boff[1])]);
                    };
                    bf16x8 bnext = bread(0);
#pragma unroll
                    for (int zi = 0; zi < NZ; ++zi)
#pragma unroll
                        for (int pair = 0; pair < 3; ++pair) {
                            const bf16x8 b = bnext;
                            if (3 * zi + pair + 1 < 3 * NZ) bnext = bread(3 * zi + pair + 1);
                            if (zi == 0 && pair == 0 && pending) { load_row(IC<RPF % 5>{}, j + RPF); pending = false; }
#pragma unroll
                            for (int dzi = 0; dzi < 3; ++dzi) {
                                const int z = zi - dzi + 1;      // output layer reached through dz = dzi - 1
                                if (z < 0 || z >= NZ) continue;
                                if (V == CV_F16) {
#pragma unroll
                                    for (int dyi = 0; dyi < 5; ++dyi) {
                                        f32x4& a = acc[z][(PH - dyi + 10) % 5];
                                        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[(dyi * 3 + dzi) * 3 + pair], b,
                                                                                    a, 0, 0, 0);
                                    }
                                } else {
                                    f32x4& a = acc[z][PH];
                                    a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[dzi * 3 + pair], b, a, 0, 0, 0);
                                }
                            }
                        }
                }
                const int o = j - 4;                     // complete after this row
                if (V != CV_OUT3) {
                    if (o >= 0 && o < nout) {
#pragma unroll
                        for (int z = 0; z < NZ; ++z) {
                            const uint32_t site = (uint32_t)(x0 + cb + nx * (y0 + o + ny * z));
                            stg(out_bf, site * 4u + (uint32_t)g, pack_bf16x4(acc[z][(PH + 1) % 5]));
                        }
                    }
#pragma unroll
                    for (int z = 0; z < NZ; ++z) acc[z][(PH + 1) % 5] = f32x4{0.f, 0.f, 0.f, 0.f};
                } else if (o >= 0 && o < nout) {
#pragma unroll
                    for (int z = 0; z < NZ; ++z) {
                        // ky = g from input row o + g (ring slot (PH + 1 + g) % 5)
                        const f32x4& r0 = acc[z][(PH + 1) % 5];
                        const f32x4& r1 = acc[z][(PH + 2) % 5];
                        const f32x4& r2 = acc[z][(PH + 3) % 5];
                        const f32x4& r3 = acc[z][(PH + 4) % 5];
                        float q[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            q[c] = g == 0 ? r0[c] : g == 1 ? r1[c] : g == 2 ? r2[c] : r3[c];
                            if (c == g) q[c] += acc[z][PH][3];   // (ky = 4, c = g), this row
                        }
#pragma unroll
                        for (int c = 0; c < 3; ++c) q[c] = lane_rows_sum(q[c]);
                        const uint32_t site = (uint32_t)(x0 + cb + nx * (y0 + o + ny * z));
                        if (g < 3) stg(out_f + (size_t)g * n, site, scale * (g == 0 ? q[0] : g == 1 ? q[1] : q[2]));
                    }
                }
            }
            if (pending) load_row(IC<RPF % 5>{}, j + RPF);
            wait_next_row<RPF, GUAR, NZ>(j, nrows, nout, wv);
            __builtin_amdgcn_s_barrier();
        };
        for (int j0 = 0; j0 < nrows; j0 += 5) {
            row(IC<0>{}, j0);
            row(IC<1>{}, j0 + 1);
            row(IC<2>{}, j0 + 2);
            row(IC<3>{}, j0 + 3);
            row(IC<4>{}, j0 + 4);
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

// X and G hold xs and gs channels per site (4 or 16); the tile's channels past
// them are zero.
__global__ void __launch_bounds__(CNN_TPB) k_cnn_wgrad(CnnDims d, const uint16_t* __restrict__ X, int xs,
                                                    const uint16_t* __restrict__ Gr, int gs,
                                                    float* __restrict__ partial)
{
    __shared__ uint16_t sG[16 * CTY * CTX];          // [o][row][x]
    __shared__ uint16_t sXc[16 * HZ * HY * HX];      // [i][hz][hy][hx]
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, col = l & 15, g = l >> 4;
    f32x4 acc[WG_TAPS];
#pragma unroll
    for (int k = 0; k < WG_TAPS; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    // channels past gs / xs (4-channel tensors) stay zero for every tile
    for (int i = threadIdx.x + gs * CTY * CTX; i < 16 * CTY * CTX; i += CNN_TPB) sG[i] = 0;
    for (int i = threadIdx.x + xs * HZ * HY * HX; i < 16 * HZ * HY * HX; i += CNN_TPB) sXc[i] = 0;
    for (int tile = blockIdx.x; tile < d.ntiles; tile += gridDim.x) {
        int x0, y0, z;
        tile_of(d, tile, x0, y0, z);
        __syncthreads();
        // G tile: CTY rows x 32 sites, transposed into [o][row][x]; 8-byte loads of 4
        // channels (gs / 4 per site)
        const int gq = gs >> 2, xq = xs >> 2;
        for (int i = threadIdx.x; i < CTY * CTX * gq; i += CNN_TPB) {
            const int q = i % gq, s = i / gq, x = s % CTX, row = s / CTX;
            const int y = y0 + row;
            uint2 v = make_uint2(0u, 0u);
            if (y < d.ny) v = ldg(reinterpret_cast<const uint2*>(Gr), (uint32_t)(x0 + x + d.nx * (y + d.ny * z)) * gq + q);
            const int b0 = ((4 * q) * CTY + row) * CTX + x, st = CTY * CTX;
            sG[b0] = (uint16_t)v.x; sG[b0 + st] = (uint16_t)(v.x >> 16);
            sG[b0 + 2 * st] = (uint16_t)v.y; sG[b0 + 3 * st] = (uint16_t)(v.y >> 16);
        }
        // X halo transposed into [i][hz][hy][hx]
        for (int i = threadIdx.x; i < HALO * xq; i += CNN_TPB) {
            const int q = i % xq, s = i / xq;
            const int hx = s % HX, r = s / HX, hy = r % HY, hz = r / HY;
            const int zz = z + hz - 1;
            uint2 v = make_uint2(0u, 0u);
            const int xr = x0 + hx - 2, yr = y0 + hy - 2;
            const bool in_xy = !d.open || ((unsigned)xr < (unsigned)d.nx && (unsigned)yr < (unsigned)d.ny);
            if (zz >= 0 && zz < d.nz && in_xy) {
                const int xx = wrap(xr, d.nx), yy = wrap(yr, d.ny);
                v = ldg(reinterpret_cast<const uint2*>(X), (uint32_t)(xx + d.nx * (yy + d.ny * zz)) * xq + q);
            }
            const int b0 = (((4 * q) * HZ + hz) * HY + hy) * HX + hx, st = HZ * HY * HX;
            sXc[b0] = (uint16_t)v.x; sXc[b0 + st] = (uint16_t)(v.x >> 16);
            sXc[b0 + 2 * st] = (uint16_t)v.y; sXc[b0 + 3 * st] = (uint16_t)(v.y >> 16);
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

// sum the per-workgroup partials: grad[t][o][i] (16 x 16 per tap).  A block of
// 16 x 16 threads owns 16 consecutive entries; its 16 thread rows each sum a
// strided slice of the partials (coalesced 64-byte rows) and LDS combines them.
__global__ void __launch_bounds__(256) k_cnn_wreduce(const float* __restrict__ partial, int nparts,
                                                     float* __restrict__ grad)
{
    __shared__ float red[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int e = blockIdx.x * 16 + tx;
    float acc = 0.0f;
    if (e < NTAP * 256)
        for (int b = ty; b < nparts; b += 16) acc += partial[(size_t)b * NTAP * 256 + e];
    red[ty][tx] = acc;
    __syncthreads();
    if (ty == 0 && e < NTAP * 256) {
        float s2 = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; ++k) s2 += red[k][tx];
        grad[e] = s2;
    }
}

// ------------------------------------------------------------ weight packing
// master FP32 W[t][o][i] (16 x 16 per tap, zero beyond the layer's channels) ->
// the row walk's A fragments [NF][64 lanes][8] BF16 for the layer (transposed = 0)
// or for its adjoint (transposed = 1: W'[t] = W[74 - t]^T, the taps mirrored).
// Fragment order: CV_F16 (dy, dz, dx pair); CV_IN3 (dy, k-step); CV_OUT3 (dz, dx pair),
// whose rows are (ky = dy + 2, channel) pairs laid out for the row walk's epilogue.
__global__ void k_cnn_pack(const float* __restrict__ W, int transposed, int variant, uint16_t* __restrict__ apack)
{
    const int nf = variant == CV_F16 ? 45 : variant == CV_IN3 ? 10 : 9;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nf * 64 * 8) return;
    const int j = e & 7, lane = (e >> 3) & 63, fi = e >> 9;
    const int r16 = lane & 15, g = lane >> 4;
    int t = -1, o = 0, i = 0;
    if (variant == CV_F16) {
        const int dyi = fi / 9, dzi = (fi / 3) % 3, dxi = 2 * (fi % 3) + (g >> 1);
        if (dxi <= 4) { t = dzi * 25 + dyi * 5 + dxi; o = r16; i = 8 * (g & 1) + j; }
    } else if (variant == CV_IN3) {
        const int dyi = fi / 2, c = 8 * (fi % 2) + 2 * g + (j >> 2), ch = j & 3;
        if (c < 15 && ch < 3) { t = (c / 5) * 25 + dyi * 5 + c % 5; o = r16; i = ch; }
    } else {
        // row m = 4 gm + rm: (ky, c) = (gm, rm) for rm < 3, (4, gm) for rm = 3 and gm < 3
        const int dzi = fi / 3, dxi = 2 * (fi % 3) + (g >> 1), gm = r16 >> 2, rm = r16 & 3;
        const int ky = rm < 3 ? gm : 4, c = rm < 3 ? rm : gm;
        if ((rm < 3 || gm < 3) && dxi <= 4) { t = dzi * 25 + ky * 5 + dxi; o = c; i = 8 * (g & 1) + j; }
    }
    float v = 0.0f;
    if (t >= 0) v = transposed ? W[((NTAP - 1 - t) * 16 + i) * 16 + o] : W[(t * 16 + o) * 16 + i];
    apack[e] = f2bf(v);
}

// ------------------------------------------------------------ input, loss
// X0[s] = bf16(s_in (m - m_ref)) in channels 0..2, zero in 3 (4 channels per site)
__global__ void k_cnn_input(Dims d, const float* __restrict__ m, const float* __restrict__ mref, float s_in,
                            uint2* __restrict__ X0)
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
    stg(X0, i, make_uint2(w0, w1));
}

// R = scale_out Y3 - T; loss partials sum R^2 (FP64); G3 = bf16(2 scale_out R / (3 n sigma2)),
// 4 channels per site
__global__ void k_cnn_loss(Dims d, const float* __restrict__ Y3, const float* __restrict__ T, float s_out,
                           float gscale, uint2* __restrict__ G3, double* __restrict__ part)
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
        stg(G3, i, make_uint2(w0, w1));
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

// S = a + b over the three components ([3][n])
__global__ void k_cnn_add(uint32_t n3, const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ S)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) S[i] = a[i] + b[i];
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

hipError_t launch_cnn_add(size_t n3, const float* a, const float* b, float* S, hipStream_t st)
{
    k_cnn_add<<<(unsigned)((n3 + 255) / 256), 256, 0, st>>>((uint32_t)n3, a, b, S);
    return hipGetLastError();
}

hipError_t launch_cnn_maxdiff(size_t n3, const float* a, const float* b, float2* part, int blocks, hipStream_t st)
{
    k_cnn_maxdiff<<<blocks, 256, 0, st>>>((uint32_t)n3, a, b, part);
    return hipGetLastError();
}

template <int NZ> static const void* film_kernel(int variant)
{
    return variant == CV_F16 ? (const void*)k_cnn_film<CV_F16, NZ>
         : variant == CV_IN3 ? (const void*)k_cnn_film<CV_IN3, NZ> : (const void*)k_cnn_film<CV_OUT3, NZ>;
}

bool cnn_film_supported(int nz) { return nz == 1 || nz == 2 || nz == 4 || nz == 8; }

static const void* walk_kernel(int variant, int film_nz)
{
    switch (film_nz) {
    case 1: return film_kernel<1>(variant);
    case 2: return film_kernel<2>(variant);
    case 4: return film_kernel<4>(variant);
    case 8: return film_kernel<8>(variant);
    default: break;
    }
    return variant == CV_F16 ? (const void*)k_cnn_walk<CV_F16>
         : variant == CV_IN3 ? (const void*)k_cnn_walk<CV_IN3> : (const void*)k_cnn_walk<CV_OUT3>;
}

template <int NZ>
static void launch_film(int variant, const CnnDims& d, const CnnWalk& wk, int grid, const void* in, const uint4* A,
                        uint2* O, float* out_f, float scale, hipStream_t st)
{
    if (variant == CV_F16) k_cnn_film<CV_F16, NZ><<<grid, CNN_TPB, 0, st>>>(d, wk, in, A, O, out_f, scale);
    else if (variant == CV_IN3) k_cnn_film<CV_IN3, NZ><<<grid, CNN_TPB, 0, st>>>(d, wk, in, A, O, out_f, scale);
    else k_cnn_film<CV_OUT3, NZ><<<grid, CNN_TPB, 0, st>>>(d, wk, in, A, O, out_f, scale);
}

hipError_t launch_cnn_rows(int variant, const CnnDims& d, const CnnWalk& wk, int grid, const uint16_t* in,
                           const uint16_t* zeros, const uint16_t* apack, uint16_t* out_bf, float* out_f,
                           float scale, hipStream_t st)
{
    const uint4* I = reinterpret_cast<const uint4*>(in);
    const uint4* Z = reinterpret_cast<const uint4*>(zeros);
    const uint4* A = reinterpret_cast<const uint4*>(apack);
    uint2* O = reinterpret_cast<uint2*>(out_bf);
    if (wk.film) {
        switch (d.nz) {
        case 1: launch_film<1>(variant, d, wk, grid, in, A, O, out_f, scale, st); break;
        case 2: launch_film<2>(variant, d, wk, grid, in, A, O, out_f, scale, st); break;
        case 4: launch_film<4>(variant, d, wk, grid, in, A, O, out_f, scale, st); break;
        case 8: launch_film<8>(variant, d, wk, grid, in, A, O, out_f, scale, st); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    if (variant == CV_F16) k_cnn_walk<CV_F16><<<grid, CNN_TPB, 0, st>>>(d, wk, I, Z, A, O, out_f, scale);
    else if (variant == CV_IN3) k_cnn_walk<CV_IN3><<<grid, CNN_TPB, 0, st>>>(d, wk, I, Z, A, O, out_f, scale);
    else k_cnn_walk<CV_OUT3><<<grid, CNN_TPB, 0, st>>>(d, wk, I, Z, A, O, out_f, scale);
    return hipGetLastError();
}

hipError_t launch_cnn_wgrad(const CnnDims& d, int grid, const uint16_t* X, int xs, const uint16_t* Gr, int gs,
                            float* partial, float* grad, hipStream_t st)
{
    k_cnn_wgrad<<<grid, CNN_TPB, 0, st>>>(d, X, xs, Gr, gs, partial);
    k_cnn_wreduce<<<(NTAP * 256 + 15) / 16, 256, 0, st>>>(partial, grid, grad);
    return hipGetLastError();
}

hipError_t launch_cnn_pack(const float* W, int transposed, int variant, uint16_t* apack, hipStream_t st)
{
    k_cnn_pack<<<(45 * 64 * 8 + 255) / 256, 256, 0, st>>>(W, transposed, variant, apack);
    return hipGetLastError();
}

hipError_t launch_cnn_input(const Dims& d, const float* m, const float* mref, float s_in, uint16_t* X0, hipStream_t st)
{
    k_cnn_input<<<(d.n + 255) / 256, 256, 0, st>>>(d, m, mref, s_in, reinterpret_cast<uint2*>(X0));
    return hipGetLastError();
}

hipError_t launch_cnn_loss(const Dims& d, const float* Y3, const float* T, float s_out, float gscale, uint16_t* G3,
                           double* part, hipStream_t st)
{
    k_cnn_loss<<<(d.n + 255) / 256, 256, 0, st>>>(d, Y3, T, s_out, gscale, reinterpret_cast<uint2*>(G3), part);
    return hipGetLastError();
}

hipError_t launch_cnn_adam(float* W, const float* g, float* m, float* v, int count, float lr, float b1, float b2,
                           float c1, float c2, float eps, hipStream_t st)
{
    k_cnn_adam<<<(count + 255) / 256, 256, 0, st>>>(W, g, m, v, count, lr, b1, b2, c1, c2, eps);
    return hipGetLastError();
}

int cnn_rows_ty_max() { return 32; }

// resident row-walk workgroups per CU of each layer shape (film_nz: the film walk of
// that many layers, 0: the per-layer walk)
int cnn_rows_occupancy(int variant, int film_nz)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, walk_kernel(variant, film_nz), CNN_TPB, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace msd
