// fft_engine.h -- batched complex FFT of TC columns held in LDS, for gfx950.
//
// A tile holds TC independent columns of N complex points, point-major
// (lds[n * LDR + c], LDR >= TC; an odd LDR avoids bank conflicts).  All NT threads
// of the work-group run a self-sorting Stockham FFT through the radix plan of N:
// each pass loads R points per work item from LDS into registers, applies the
// twiddles, runs a radix-R butterfly in registers and writes the R outputs back in
// Stockham order.  Output is in natural order; transforms are unnormalised with
// forward sign -1, inverse +1.  Twiddles come from a table tw[m] = exp(-2 pi i m/N).
#pragma once
#include "ddns_internal.h"

namespace ddns {

// ---------------------------------------------------------------- tile transfers
// A tile is R rows of W adjacent complex values in HBM (row r at src + r * rstride)
// and R * LDR values in LDS (row-major, or transposed when TRANS: lds[c * LDR + r]).
// Every lane moves 16 bytes per access and the trip count is a compile-time
// constant, so a lane issues all of its loads before the first wait.
template <typename C> struct vec16;
template <> struct vec16<float2> {
    using V = float4;
    static constexpr int VW = 2;
    __device__ static void split(const V& v, float2* o) { o[0] = make_float2(v.x, v.y); o[1] = make_float2(v.z, v.w); }
    __device__ static V join(const float2* c) { return make_float4(c[0].x, c[0].y, c[1].x, c[1].y); }
};
template <> struct vec16<double2> {
    using V = double2;
    static constexpr int VW = 1;
    __device__ static void split(const V& v, double2* o) { o[0] = v; }
    __device__ static V join(const double2* c) { return c[0]; }
};

// the registers one lane holds for a tile in flight
template <typename C, int R, int W, int NT>
struct TileRegs {
    using VT = vec16<C>;
    static constexpr int VW = VT::VW, PER_ROW = W / VW, TOT = R * PER_ROW, Q = TOT / NT;
    static_assert(W % VW == 0 && TOT % NT == 0, "tile must split evenly over the work-group");
    typename VT::V v[Q];
};

// issue the loads of a tile into registers (returns without waiting)
template <typename C, int R, int W, int NT>
__device__ __forceinline__ void tile_fetch(TileRegs<C, R, W, NT>& t, const C* __restrict__ src, size_t rstride, int tid)
{
    using TR = TileRegs<C, R, W, NT>;
#pragma unroll
    for (int q = 0; q < TR::Q; ++q) {
        const int e = tid + q * NT, r = e / TR::PER_ROW, c = (e % TR::PER_ROW) * TR::VW;
        t.v[q] = *(const typename TR::VT::V*)(src + (size_t)r * rstride + c);
    }
}

// write a fetched tile to LDS
template <typename C, int R, int W, int LDR, int NT, bool TRANS = false>
__device__ __forceinline__ void tile_put(C* __restrict__ lds, const TileRegs<C, R, W, NT>& t, int tid)
{
    using TR = TileRegs<C, R, W, NT>;
#pragma unroll
    for (int q = 0; q < TR::Q; ++q) {
        const int e = tid + q * NT, r = e / TR::PER_ROW, c = (e % TR::PER_ROW) * TR::VW;
        C v[TR::VW];
        TR::VT::split(t.v[q], v);
#pragma unroll
        for (int k = 0; k < TR::VW; ++k) lds[TRANS ? (c + k) * LDR + r : r * LDR + c + k] = v[k];
    }
}

template <typename C, int R, int W, int LDR, int NT, bool TRANS = false>
__device__ __forceinline__ void tile_load(C* __restrict__ lds, const C* __restrict__ src, size_t rstride, int tid)
{
    TileRegs<C, R, W, NT> t;
    tile_fetch<C, R, W, NT>(t, src, rstride, tid);
    tile_put<C, R, W, LDR, NT, TRANS>(lds, t, tid);
}

template <typename C, int R, int W, int LDR, int NT>
__device__ __forceinline__ void tile_store(C* __restrict__ dst, const C* __restrict__ lds, size_t rstride, int tid,
                                           decltype(C{}.x) scale = 1)
{
    using VT = vec16<C>;
    constexpr int VW = VT::VW, PER_ROW = W / VW, TOT = R * PER_ROW;
    static_assert(W % VW == 0 && TOT % NT == 0, "tile must split evenly over the work-group");
#pragma unroll
    for (int q = 0; q < TOT / NT; ++q) {
        const int e = tid + q * NT, r = e / PER_ROW, c = (e % PER_ROW) * VW;
        C v[VW];
#pragma unroll
        for (int k = 0; k < VW; ++k) v[k] = cscale(lds[r * LDR + c + k], scale);
        *(typename VT::V*)(dst + (size_t)r * rstride + c) = VT::join(v);
    }
}

// ---------------------------------------------------------------- register DFTs
template <typename C, bool INV>
__device__ __forceinline__ void dft2(C* v)
{
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename C, bool INV>
__device__ __forceinline__ void dft4(C* v)
{
    C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]), t3 = cmul_mi<C, INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <typename C, bool INV>
__device__ __forceinline__ void dft8(C* v)
{
    using T = decltype(C{}.x);
    const T h = (T)0.70710678118654752440;
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    dft4<C, INV>(e);
    dft4<C, INV>(o);
    // o[k] *= exp(-+ 2 pi i k / 8)
    o[1] = INV ? mk<C>(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
               : mk<C>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    o[2] = cmul_mi<C, INV>(o[2]);
    o[3] = INV ? mk<C>(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
               : mk<C>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

template <typename C, bool INV>
__device__ __forceinline__ void dft16(C* v)
{
    using T = decltype(C{}.x);
    C e[8], o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { e[k] = v[2 * k]; o[k] = v[2 * k + 1]; }
    dft8<C, INV>(e);
    dft8<C, INV>(o);
    const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173;
    const T h = (T)0.70710678118654752440;
    const T sg = INV ? (T)1 : (T)-1;
    // w_k = cos(2 pi k/16) + sg i sin(2 pi k/16)
    const C w1 = mk<C>(c1, sg * s1), w2 = mk<C>(h, sg * h), w3 = mk<C>(s1, sg * c1);
    const C w5 = mk<C>(-s1, sg * c1), w6 = mk<C>(-h, sg * h), w7 = mk<C>(-c1, sg * s1);
    o[1] = cmul(o[1], w1);
    o[2] = cmul(o[2], w2);
    o[3] = cmul(o[3], w3);
    o[4] = cmul_mi<C, INV>(o[4]);
    o[5] = cmul(o[5], w5);
    o[6] = cmul(o[6], w6);
    o[7] = cmul(o[7], w7);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = cadd(e[k], o[k]);
        v[k + 8] = csub(e[k], o[k]);
    }
}

template <int R, typename C, bool INV>
__device__ __forceinline__ void dft(C* v)
{
    if constexpr (R == 2) dft2<C, INV>(v);
    else if constexpr (R == 4) dft4<C, INV>(v);
    else if constexpr (R == 8) dft8<C, INV>(v);
    else if constexpr (R == 16) dft16<C, INV>(v);
}

// ---------------------------------------------------------------- radix plans
template <int N> struct Plan;
template <> struct Plan<4>    { static constexpr int P = 1, R0 = 4,  R1 = 1,  R2 = 1; };
template <> struct Plan<8>    { static constexpr int P = 1, R0 = 8,  R1 = 1,  R2 = 1; };
template <> struct Plan<16>   { static constexpr int P = 1, R0 = 16, R1 = 1,  R2 = 1; };
template <> struct Plan<32>   { static constexpr int P = 2, R0 = 8,  R1 = 4,  R2 = 1; };
template <> struct Plan<64>   { static constexpr int P = 2, R0 = 8,  R1 = 8,  R2 = 1; };
template <> struct Plan<128>  { static constexpr int P = 2, R0 = 16, R1 = 8,  R2 = 1; };
template <> struct Plan<256>  { static constexpr int P = 2, R0 = 16, R1 = 16, R2 = 1; };
template <> struct Plan<512>  { static constexpr int P = 3, R0 = 8,  R1 = 8,  R2 = 8; };
template <> struct Plan<1024> { static constexpr int P = 3, R0 = 16, R1 = 16, R2 = 4; };

// One Stockham pass of stride NS (product of the radices already applied) over K
// tiles at once (tile k at lds + k * TS): each work item runs one radix-R butterfly
// per tile, so a lane has K independent butterflies in flight.  Caller guarantees the
// tiles are complete (barrier) before the call; the pass ends with a barrier.  tw may
// point to LDS or to global memory.
template <typename T, int N, int R, int NS, bool INV, int TC, int LDR, int NT, int K = 1, int TS = 0>
__device__ __forceinline__ void fft_pass(cpx<T>* __restrict__ lds, const cpx<T>* __restrict__ tw, int tid)
{
    using C = cpx<T>;
    constexpr int NJ = N / R;
    constexpr int ITEMS = TC * NJ;
    constexpr int Q = (ITEMS + NT - 1) / NT;
    C v[K][Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int it = tid + NT * q;
        if (ITEMS % NT == 0 || it < ITEMS) {
            const int c = it % TC, j = it / TC;
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int r = 0; r < R; ++r) v[k][q][r] = lds[k * TS + (j + r * NJ) * LDR + c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int it = tid + NT * q;
        if (ITEMS % NT == 0 || it < ITEMS) {
            const int c = it % TC, j = it / TC;
            const int m = j % NS;
            if constexpr (NS > 1) {
                constexpr int STEP = N / (NS * R);
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    C w = tw[(m * r * STEP) & (N - 1)];
                    if (INV) w.y = -w.y;
#pragma unroll
                    for (int k = 0; k < K; ++k) v[k][q][r] = cmul(v[k][q][r], w);
                }
            }
            const int base = (j / NS) * NS * R + m;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                dft<R, C, INV>(v[k][q]);
#pragma unroll
                for (int r = 0; r < R; ++r) lds[k * TS + (base + r * NS) * LDR + c] = v[k][q][r];
            }
        }
    }
    __syncthreads();
}

// FFT of every column of K tiles (tile k at lds + k * TS) through the plan of N.
template <typename T, int N, bool INV, int TC, int LDR, int NT, int K = 1, int TS = 0>
__device__ __forceinline__ void fft_tile(cpx<T>* __restrict__ lds, const cpx<T>* __restrict__ tw, int tid)
{
    using P = Plan<N>;
    fft_pass<T, N, P::R0, 1, INV, TC, LDR, NT, K, TS>(lds, tw, tid);
    if constexpr (P::P > 1) fft_pass<T, N, P::R1, P::R0, INV, TC, LDR, NT, K, TS>(lds, tw, tid);
    if constexpr (P::P > 2) fft_pass<T, N, P::R2, P::R0 * P::R1, INV, TC, LDR, NT, K, TS>(lds, tw, tid);
}

// the twiddle table exp(-2 pi i m / N), m < N, copied to LDS by the whole work-group
template <typename C, int N, int NT>
__device__ __forceinline__ void stage_twiddles(C* __restrict__ tws, const C* __restrict__ tw, int tid)
{
#pragma unroll
    for (int m = tid; m < N; m += NT) tws[m] = tw[m];
}

}  // namespace ddns
