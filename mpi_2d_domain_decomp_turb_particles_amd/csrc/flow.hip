// flow.hip -- spectral Navier-Stokes step on one MI355X.
//
// Reference: turb.f90:506-691 (curl, inverse FFTs, u x w, forward FFTs, dealiasing,
// AB2/Crank-Nicolson update, projection, forcing, symmetrize), mpifft.f90:81-198
// (the 3-D real FFT as x / y / z line transforms with a 1/N scale after each
// forward stage), flowlib.f90:180-562 (dealiasing, supforDET, symmetrize,
// supforSTH).
//
// Layout: every complex field is [s][r][kx], kx fastest with kx < N/2, so the x
// stage reads contiguous lines and the y and z stages read tiles of TC adjacent kx
// columns (one 128-256 byte segment per row).  The step runs as five launches:
//   Ky(inverse, 6 fields, kx = 0 column symmetrised)  ->  Kx(6 C2R, u x w, 3 R2C,
//   physical u for the droplets)  ->  Ky(forward, 3 fields)  ->  Kpre (one block:
//   deterministic-forcing shell energies or the OU forcing table)  ->  Kz(forward z
//   FFT, update, curl, inverse z FFT of u and w)  ->  Ksym (kx = 0 plane of u-hat).
// The z stage therefore closes step n and opens step n + 1 in one pass over HBM.
#include "fft_engine.h"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

namespace ddns {

constexpr int NT = 256;

enum { KZ_FULL = 0, KZ_CURL_INV = 1, KZ_PLAIN_FWD = 2, KZ_PLAIN_INV = 3 };

// tile widths (kx columns per work-group) for the y and z stages
template <typename T, int N> constexpr int ky_tc() { return (sizeof(T) == 4) ? (N >= 1024 ? 8 : 16) : (N >= 1024 ? 4 : 8); }
// The y stage runs at one of three tile widths, chosen per grid and precision by timing
// (flow_tune): half, once and twice the default, where the tile splits evenly over the
// work-group's 16-byte accesses and fits LDS with the twiddles.  The width changes only
// which columns a work-group transforms together, never the arithmetic of a column, so
// every choice gives the same bits.
template <typename T, int N, int TC> constexpr bool ky_fits()
{
    constexpr int VW = sizeof(T) == 4 ? 2 : 1;
    return TC >= VW && TC <= N / 2 && TC % VW == 0 && (N * (TC / VW)) % 256 == 0 &&
           (N * (TC >= 8 ? TC + 1 : TC) + N) * (int)sizeof(cpx<T>) <= 160 * 1024;
}
template <typename T, int N> constexpr int ky_cand(int k) { return k == 0 ? ky_tc<T, N>() / 2 : (k == 1 ? ky_tc<T, N>() : 2 * ky_tc<T, N>()); }
// z-stage tile widths for N <= 256, overridable at build time for tuning sweeps
#ifndef DDNS_KZ_TC_F32
#define DDNS_KZ_TC_F32 16
#endif
#ifndef DDNS_KZ_TC_F64
#define DDNS_KZ_TC_F64 8
#endif
// persistent z stage, overridable at build time for tuning sweeps: the update
// operands of DDNS_KZ_EPTP_BYTES / sizeof(complex) elements per thread are loaded
// before the forward transforms, the rest of the update loop is unrolled
// DDNS_KZ_UNROLL times, and with DDNS_KZ_NEXTB the next tile's input is loaded during
// this tile's inverse transforms (single precision only)
#ifndef DDNS_KZ_UNROLL
#define DDNS_KZ_UNROLL 4
#endif
#ifndef DDNS_KZ_NEXTB
#define DDNS_KZ_NEXTB 1
#endif
#ifndef DDNS_KZ_EPTP_BYTES
#define DDNS_KZ_EPTP_BYTES 32
#endif
template <typename T, int N> constexpr int kz_tc()
{
    if constexpr (sizeof(T) == 4) return N < 256 ? 16 : (N == 256 ? DDNS_KZ_TC_F32 : (N == 512 ? 8 : 4));
    else return N <= 256 ? DDNS_KZ_TC_F64 : (N == 512 ? 4 : 2);
}

// Work-groups are dispatched to the 8 XCDs round-robin by linear id, and each XCD has
// its own L2.  When a tile row is narrower than a 128-byte line, the P tiles that share
// a line get linear ids L, L+8, ..., so they land on the same XCD and meet in its L2.
// Returns the tile index of linear work-group id L (total a multiple of 8 P).
template <int P>
__device__ __forceinline__ int xcd_tile(int L)
{
    if constexpr (P <= 1) {
        return L;
    } else {
        const int g = L / (8 * P), r = L % (8 * P);
        return g * (8 * P) + (r % 8) * P + r / 8;
    }
}
template <int TC> constexpr int ldr_of() { return (TC >= 8) ? TC + 1 : TC; }
// One-tile (plain) z transforms for N >= 512 take rows of 128 bytes: one padded tile of
// N rows fits in LDS at that width even where the four tiles of the fused stage do not
template <typename T, int N> constexpr int kz_tc_plain()
{
    if constexpr (N < 512) return kz_tc<T, N>();
    else return sizeof(T) == 4 ? 16 : 8;
}
// Where the fused z stage's tiles hold rows shorter than 128 bytes (N >= 512) the stage
// runs as separate kernels instead: three plain forward transforms, the pointwise update,
// the curl, six plain inverse transforms.  They move 39 fields instead of 21 but at the
// rate of full rows (36 with the curl formed in the update): faster by 9 % (FP32) and
// 15 % (FP64) at 512^3, 38 % at 1024^3.
// DDNS_Z_SPLIT=1 or 0 in the environment selects one path at any N, to check them
// against each other.
// threads of a z work-group: the single-tile transforms of N = 1024 take 512, so that a CU,
// which holds one such tile in LDS, has twice the loads in flight (8 % faster there; at
// N = 512 two 256-thread tiles fit in LDS and do better)
template <typename T, int N, int MODE> constexpr int kz_threads()
{
    return ((MODE == KZ_PLAIN_FWD || MODE == KZ_PLAIN_INV) && N >= 1024) ? 512 : NT;
}
template <typename T, int N> constexpr bool kz_split_default() { return kz_tc<T, N>() * (int)sizeof(cpx<T>) < 128; }
template <typename T, int N> constexpr int kz_ldr()
{
    constexpr int TC = kz_tc<T, N>();
    return ((4 * N * (TC + 1) + N) * (int)sizeof(cpx<T>) <= 160 * 1024) ? TC + 1 : TC;   // 4 tiles + twiddles
}

__device__ __forceinline__ int wave_index(int i, int N) { return i <= N / 2 ? i : i - N; }

// Is mode (kx, ky, kz) inside the dealiasing sphere (turb.f90:405, 648-650)?  update_mode's
// own shell index, in its precision: outside it the update sets both the new u-hat and
// the stored nonlinear term to zero (the mean mode counts as inside: its u-hat is zero
// but its nonlinear term is kept).
template <typename T>
__device__ __forceinline__ bool mode_in_sphere(int kx, int ky, int kz, int N, int nek)
{
    const int kyw = wave_index(ky, N), kzw = wave_index(kz, N);
    const T fx = (T)kx, fy = (T)kyw, fz = (T)kzw;
    T k2 = fx * fx + fy * fy + fz * fz;
    if (kx == 0 && kyw == 0 && kzw == 0) k2 = (T)1.e-5;
    return (int)(sqrt(k2) + (T)0.5) <= nek;
}

// Batched realisations: realisation r of a field starts r * rstride complex values after
// realisation 0 (rstride = 0 for one realisation), its low-mode table r * LOW_STRIDE
// doubles after, and the physical velocity r * pstride reals after.  With forcing < 0
// every realisation takes its forcing kind and amplitudes from its own low-mode table
// (LOW_PAR onwards).
constexpr int LOW_STRIDE = 1024;
constexpr int LOW_PAR = 1000;       // [0] forcing kind, [1..2] force(1:2), [3] OU time scale, [4] OU variance
// [LOW_ZOUT] = 1 once a step has written the state's modes beyond the dealiasing sphere,
// which every update sets to zero: the z stage then neither loads nor stores them (u-hat
// and the AB2 history stay zero there).  Cleared by every upload of the state; set by the
// step's k_sym, which runs after the z stage.
constexpr int LOW_ZOUT = LOW_PAR + 5;
constexpr int STH_STATE_STRIDE = 64;
constexpr int STH_B_STRIDE = 512;

__device__ __forceinline__ int real_forcing(int forcing, const double* lowr)
{
    return forcing >= 0 ? forcing : (int)lowr[LOW_PAR];
}

struct KzArgs {
    const void* B[3];
    void* U[3];
    void* O[3];
    void* A[6];
    const void* tw;
    const double* low;
    double dt, rnu;
    int nek, forcing;
    double scale;      // plain modes: output scale
    size_t rstride;    // complex values between realisations of one field
};

struct KyArgs {
    void* F[6];
    const void* tw;
    double scale;
    int nf, R;         // fields per realisation, realisations: grid z runs over nf * R
    size_t rstride;
};

struct KxArgs {
    const void* A[6];
    void* B[3];
    void* uphys;
    const void* tw;    // N entries (post/pre-processing twiddles)
    const void* twh;   // N/2 entries (half-length FFT)
    size_t rstride;    // grid y runs over the realisations
    size_t pstride;    // reals between realisations of the physical velocity
};

// ------------------------------------------------------------------ spectral update
// One mode of turb.f90:662-687 for the three components held in n[3] (forward FFT
// output, already scaled) and the state u[3], o[3]; returns the new u-hat in v[3]
// and the dealiased nonlinear term in nd[3].  ff: supforDET factors; fa: the
// supforSTH table a(6,5,5) per component.
template <typename T>
__device__ __forceinline__ void update_mode(int kx, int ky, int kz, int N, int nek, int forcing,
                                            T dt, T rnu, const double* __restrict__ low,
                                            const cpx<T>* n, const cpx<T>* u, const cpx<T>* o,
                                            cpx<T>* v, cpx<T>* nd)
{
    using C = cpx<T>;
    const int kyw = wave_index(ky, N), kzw = wave_index(kz, N);
    const T fx = (T)kx, fy = (T)kyw, fz = (T)kzw;
    T k2 = fx * fx + fy * fy + fz * fz;
    const bool zero_mode = (kx == 0 && kyw == 0 && kzw == 0);
    if (zero_mode) k2 = (T)1.e-5;                                   // flowlib.f90:36-38
    const int ik2 = (int)(sqrt(k2) + (T)0.5);                       // turb.f90:405
    const bool keep = ik2 <= nek && !zero_mode;
    const T dt_h = (T)0.5 * dt, dtr = (T)1.5 * dt;                   // turb.f90:483-484
    const T ac = rnu * dt_h * k2;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        nd[c] = (ik2 <= nek) ? n[c] : mk<C>((T)0, (T)0);             // dealiasing, 648-650
        v[c] = csub(cadd(cscale(u[c], (T)1 - ac), cscale(nd[c], dtr)), cscale(o[c], dt_h));
    }
    if (forcing == 2 && kx < 3) {                                    // supforSTH add, flowlib.f90:529-559
        const int js = ky <= 2 ? ky : (ky >= N - 2 ? ky - (N - 5) : -1);
        const int ks = kz <= 2 ? kz : (kz >= N - 2 ? kz - (N - 5) : -1);
        if (js >= 0 && ks >= 0) {
            const double* fa = low + 8;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int o0 = 2 * kx + 6 * (js + 5 * ks);
                v[c].x = v[c].x + dt * (T)fa[150 * c + o0];
                v[c].y = v[c].y + dt * (T)fa[150 * c + o0 + 1];
            }
        }
    }
    // projection, 669-672
    const C t = cscale(cadd(cadd(cscale(v[0], fx), cscale(v[1], fy)), cscale(v[2], fz)), (T)1 / k2);
    v[0] = csub(v[0], cscale(t, fx));
    v[1] = csub(v[1], cscale(t, fy));
    v[2] = csub(v[2], cscale(t, fz));
    const T den = (T)1 + ac;                                         // 674-676
    T f = (T)1 / den;
    if (forcing == 1 && (ik2 == 1 || ik2 == 2)) f = f * (T)low[ik2 - 1];   // supforDET
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        v[c] = keep ? cscale(v[c], f) : mk<C>((T)0, (T)0);
    }
}

// ------------------------------------------------------------------ z stage
// One tile per work-group: curl and inverse z transforms of u-hat (KZ_CURL_INV, the
// first step) or a plain forward / inverse z transform of one field.
template <typename T, int N, int MODE>
__global__ void __launch_bounds__((kz_threads<T, N, MODE>())) k_z(KzArgs a)
{
    constexpr int NTK = kz_threads<T, N, MODE>();
    using C = cpx<T>;
    constexpr int M = N / 2;
    constexpr bool PLAIN = MODE == KZ_PLAIN_FWD || MODE == KZ_PLAIN_INV;
    constexpr int TC = PLAIN ? kz_tc_plain<T, N>() : kz_tc<T, N>();
    constexpr int LDR = PLAIN ? ldr_of<TC>() : kz_ldr<T, N>();
    constexpr int NTILE = MODE == KZ_CURL_INV ? 4 : 1;
    __shared__ C lds[NTILE][N * LDR];
    __shared__ C tws[N];
    const int tid = threadIdx.x;
    constexpr int NBX = M / TC;
    constexpr int P = (TC * (int)sizeof(C) >= 128) ? 1 : 128 / (TC * (int)sizeof(C));
    const int tile = xcd_tile<(NBX % P == 0) ? P : 1>(blockIdx.x);
    const int kx0 = (tile % NBX) * TC;
    const int ky = tile / NBX;
    stage_twiddles<C, N, NTK>(tws, (const C*)a.tw, tid);
    const C* tw = tws;
    const size_t rowstride = (size_t)N * M;
    const size_t base = (size_t)blockIdx.y * a.rstride + (size_t)ky * M + kx0;

    if constexpr (MODE == KZ_PLAIN_FWD || MODE == KZ_PLAIN_INV) {
        const C* src = (const C*)a.B[0];
        C* dst = (C*)a.U[0];
        tile_load<C, N, TC, LDR, NTK>(lds[0], src + base, rowstride, tid);
        __syncthreads();
        fft_tile<T, N, MODE == KZ_PLAIN_INV, TC, LDR, NTK>(lds[0], tw, tid);
        tile_store<C, N, TC, LDR, NTK>(dst + base, lds[0], rowstride, tid, (T)a.scale);
        return;
    } else {
#pragma unroll
        for (int q = 0; q < 3; ++q) tile_load<C, N, TC, LDR, NTK>(lds[q], (const C*)a.U[q] + base, rowstride, tid);
        __syncthreads();
        // vorticity w = i k x u (turb.f90:506-517), one component at a time in tile 3
#pragma unroll 1
        for (int q = 0; q < 3; ++q) {
            for (int e = tid; e < TC * N; e += NTK) {
                const int n = e / TC, c = e % TC;
                const T fx = (T)(kx0 + c), fy = (T)wave_index(ky, N), fz = (T)wave_index(n, N);
                const C ux = lds[0][n * LDR + c], uy = lds[1][n * LDR + c], uz = lds[2][n * LDR + c];
                C w;
                if (q == 0) w = csub(cscale(uz, fy), cscale(uy, fz));
                else if (q == 1) w = csub(cscale(ux, fz), cscale(uz, fx));
                else w = csub(cscale(uy, fx), cscale(ux, fy));
                lds[3][n * LDR + c] = mk<C>(-w.y, w.x);
            }
            __syncthreads();
            fft_tile<T, N, true, TC, LDR, NTK>(lds[3], tw, tid);
            tile_store<C, N, TC, LDR, NTK>((C*)a.A[3 + q] + base, lds[3], rowstride, tid);
            __syncthreads();
        }
        fft_tile<T, N, true, TC, LDR, NTK, 3, N * LDR>(lds[0], tw, tid);
#pragma unroll
        for (int q = 0; q < 3; ++q) tile_store<C, N, TC, LDR, NTK>((C*)a.A[q] + base, lds[q], rowstride, tid);
    }
}

// The split z stage's pointwise parts.  k_z_update: the update of k_z_full (dealias,
// AB2/CN, forcing, projection, CN divide) on the z-transformed N-hat already scaled by
// 1/N, and the curl of the new u-hat.  k_z_curl: the curl alone, for a new state.
// The curl is w-hat = i k x u-hat (turb.f90:506-517), into A[3..5].
template <typename T>
__device__ __forceinline__ void curl_to(const KzArgs& a, size_t t, int kx, int ky, int kz, int N, const cpx<T>* u)
{
    using C = cpx<T>;
    const T fx = (T)kx, fy = (T)wave_index(ky, N), fz = (T)wave_index(kz, N);
    const C w0 = csub(cscale(u[2], fy), cscale(u[1], fz));
    const C w1 = csub(cscale(u[0], fz), cscale(u[2], fx));
    const C w2 = csub(cscale(u[1], fx), cscale(u[0], fy));
    ((C*)a.A[3])[t] = mk<C>(-w0.y, w0.x);
    ((C*)a.A[4])[t] = mk<C>(-w1.y, w1.x);
    ((C*)a.A[5])[t] = mk<C>(-w2.y, w2.x);
}
template <typename T, int N>
__global__ void __launch_bounds__(256) k_z_update(KzArgs a)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * M) return;
    const int kx = (int)(t % M), ky = (int)((t / M) % N), kz = (int)(t / ((size_t)N * M));
    // beyond the dealiasing sphere the state is zero once a step has written it
    const bool in = a.low[LOW_ZOUT] == 0.0 || mode_in_sphere<T>(kx, ky, kz, N, a.nek);
    C nn[3], u3[3], o3[3], vv[3], nd[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        nn[q] = ((const C*)a.B[q])[t];
        u3[q] = in ? ((const C*)a.U[q])[t] : mk<C>((T)0, (T)0);
        o3[q] = in ? ((const C*)a.O[q])[t] : mk<C>((T)0, (T)0);
    }
    update_mode<T>(kx, ky, kz, N, a.nek, a.forcing, (T)a.dt, (T)a.rnu, a.low, nn, u3, o3, vv, nd);
    if (in) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            ((C*)a.U[q])[t] = vv[q];
            ((C*)a.O[q])[t] = nd[q];
        }
    }
    curl_to<T>(a, t, kx, ky, kz, N, vv);
}

template <typename T, int N>
__global__ void __launch_bounds__(256) k_z_curl(KzArgs a)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * M) return;
    const int kx = (int)(t % M), ky = (int)((t / M) % N), kz = (int)(t / ((size_t)N * M));
    const C u[3] = {((const C*)a.U[0])[t], ((const C*)a.U[1])[t], ((const C*)a.U[2])[t]};
    curl_to<T>(a, t, kx, ky, kz, N, u);
}

// The full z stage as a persistent kernel: one work-group per CU walks the tiles
// tile = blockIdx.x, blockIdx.x + gridDim.x, ...  The LDS tiles allow only one
// work-group per CU, so instead of a second work-group hiding HBM latency, a tile's
// update operands (U-hat, O-hat) are loaded into registers before its forward
// transforms, and the next tile's input B is loaded into registers during this
// tile's curl and inverse transforms.  The curl and the inverse transforms are those
// of k_z<KZ_CURL_INV>.
#ifndef DDNS_KZ_NT
#define DDNS_KZ_NT 256
#endif
// threads of the persistent z-stage work-group: DDNS_KZ_NT where the tile splits
// evenly over them (16-byte moves, one update element per thread and pass), else NT
template <typename T, int N> constexpr int kz_nt()
{
    constexpr int tot = N * kz_tc<T, N>() * (int)sizeof(cpx<T>) / 16;
    return (tot % DDNS_KZ_NT == 0 && (N * kz_tc<T, N>()) % DDNS_KZ_NT == 0) ? DDNS_KZ_NT : NT;
}

// The persistent z stage runs with kz_nt threads or twice as many, chosen per grid and
// precision by timing (flow_tune), where the tile splits evenly over the threads.  More
// threads keep more of a CU's loads in flight (a CU holds one work-group: its four
// tiles fill LDS) but leave each thread half the registers; the elements a thread
// updates and the columns it transforms change, their arithmetic does not.  (Four times
// as many threads leave 128 registers per thread, and the kernel spills heavily there:
// that size is not built.)
template <typename T, int N, int NTZ> constexpr bool kz_nt_fits()
{
    constexpr int tot = N * kz_tc<T, N>() * (int)sizeof(cpx<T>) / 16;
    return NTZ <= 512 && tot % NTZ == 0 && (N * kz_tc<T, N>()) % NTZ == 0;
}
template <typename T, int N> constexpr int kz_nt_cand(int k) { return kz_nt<T, N>() << k; }

template <typename T, int N, int NTZ>
__global__ void __launch_bounds__(NTZ) k_z_full(KzArgs a, int ntiles)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    constexpr int TC = kz_tc<T, N>();
    constexpr int LDR = kz_ldr<T, N>();
    constexpr int NBX = M / TC;
    constexpr int EPT = TC * N / NTZ;          // update elements per thread
    static_assert((TC * N) % NTZ == 0, "the update must split evenly over the work-group");
    // register budget: the three-tile transforms already hold 48 complex values per
    // lane, so only part of the update operands is loaded ahead, and the next tile's
    // input only in single precision
    constexpr int EPTP = DDNS_KZ_EPTP_BYTES > 0 ? std::max(1, std::min(EPT, (int)(DDNS_KZ_EPTP_BYTES / sizeof(C)))) : EPT;
    constexpr bool NEXT_B = DDNS_KZ_NEXTB && sizeof(T) == 4;
    __shared__ C lds[4][N * LDR];
    __shared__ C tws[N];
    const int tid = threadIdx.x;
    stage_twiddles<C, N, NTZ>(tws, (const C*)a.tw, tid);
    const C* tw = tws;
    const size_t rowstride = (size_t)N * M;
    // ntiles counts the tiles of every realisation: tile = r * TPR + local tile
    constexpr int TPR = NBX * N;
    auto tile_base = [&](int tl) {
        const int r = tl / TPR, lt = tl - r * TPR;
        return (size_t)r * a.rstride + (size_t)(lt / NBX) * M + (lt % NBX) * TC;
    };
    using TR = TileRegs<C, N, TC, NTZ>;
    TR nb[3];
    int tile = blockIdx.x;
    if (tile < ntiles) {
        const size_t base = tile_base(tile);
#pragma unroll
        for (int q = 0; q < 3; ++q) tile_fetch<C, N, TC, NTZ>(nb[q], (const C*)a.B[q] + base, rowstride, tid);
    }
    for (; tile < ntiles; tile += gridDim.x) {
        const int rr = tile / TPR, lt = tile - rr * TPR;
        const int kx0 = (lt % NBX) * TC, ky = lt / NBX;
        const size_t base = (size_t)rr * a.rstride + (size_t)ky * M + kx0;
        const double* lowr = a.low + (size_t)rr * LOW_STRIDE;
        const int forcing = real_forcing(a.forcing, lowr);
        // beyond the dealiasing sphere the state is zero once a step has written it
        const bool zout = lowr[LOW_ZOUT] != 0.0;
        if (!NEXT_B && tile != (int)blockIdx.x) {
#pragma unroll
            for (int q = 0; q < 3; ++q) tile_fetch<C, N, TC, NTZ>(nb[q], (const C*)a.B[q] + base, rowstride, tid);
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) tile_put<C, N, TC, LDR, NTZ>(lds[q], nb[q], tid);
        // the first update operands, in flight during the forward transforms
        C uu[EPTP][3], oo[EPTP][3];
#pragma unroll
        for (int k = 0; k < EPTP; ++k) {
            const int e = tid + k * NTZ, n = e / TC, c = e % TC;
            const size_t g = base + (size_t)n * rowstride + c;
            const bool ld = !zout || mode_in_sphere<T>(kx0 + c, ky, n, N, a.nek);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                uu[k][q] = ld ? ((const C*)a.U[q])[g] : mk<C>((T)0, (T)0);
                oo[k][q] = ld ? ((const C*)a.O[q])[g] : mk<C>((T)0, (T)0);
            }
        }
        __syncthreads();
        fft_tile<T, N, false, TC, LDR, NTZ, 3, N * LDR>(lds[0], tw, tid);     // the three components together
        const T inv_n = (T)1 / (T)N;
        auto update = [&](int k, const C* u3, const C* o3) {
            const int e = tid + k * NTZ, n = e / TC, c = e % TC;
            const size_t g = base + (size_t)n * rowstride + c;
            C nn[3], vv[3], nd[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) nn[q] = cscale(lds[q][n * LDR + c], inv_n);
            update_mode<T>(kx0 + c, ky, n, N, a.nek, forcing, (T)a.dt, (T)a.rnu, lowr, nn, u3, o3, vv, nd);
            const bool wr = !zout || mode_in_sphere<T>(kx0 + c, ky, n, N, a.nek);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (wr) {
                    ((C*)a.U[q])[g] = vv[q];
                    ((C*)a.O[q])[g] = nd[q];
                }
                lds[q][n * LDR + c] = vv[q];
            }
        };
#pragma unroll
        for (int k = 0; k < EPTP; ++k) update(k, uu[k], oo[k]);
#pragma unroll DDNS_KZ_UNROLL
        for (int k = EPTP; k < EPT; ++k) {
            const int e = tid + k * NTZ, n = e / TC, c = e % TC;
            const size_t g = base + (size_t)n * rowstride + c;
            const bool ld = !zout || mode_in_sphere<T>(kx0 + c, ky, n, N, a.nek);
            C u3[3], o3[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                u3[q] = ld ? ((const C*)a.U[q])[g] : mk<C>((T)0, (T)0);
                o3[q] = ld ? ((const C*)a.O[q])[g] : mk<C>((T)0, (T)0);
            }
            update(k, u3, o3);
        }
        // the next tile's input, in flight during the curl and the inverse transforms
        // (the barrier keeps the compiler from hoisting these loads over the update)
        asm volatile("" ::: "memory");
        const int nt = tile + (int)gridDim.x;
        if (NEXT_B && nt < ntiles) {
            const size_t nbase = tile_base(nt);
#pragma unroll
            for (int q = 0; q < 3; ++q) tile_fetch<C, N, TC, NTZ>(nb[q], (const C*)a.B[q] + nbase, rowstride, tid);
        }
        __syncthreads();
        // vorticity w = i k x u (turb.f90:506-517), one component at a time in tile 3
#pragma unroll 1
        for (int q = 0; q < 3; ++q) {
            for (int e = tid; e < TC * N; e += NTZ) {
                const int n = e / TC, c = e % TC;
                const T fx = (T)(kx0 + c), fy = (T)wave_index(ky, N), fz = (T)wave_index(n, N);
                const C ux = lds[0][n * LDR + c], uy = lds[1][n * LDR + c], uz = lds[2][n * LDR + c];
                C w;
                if (q == 0) w = csub(cscale(uz, fy), cscale(uy, fz));
                else if (q == 1) w = csub(cscale(ux, fz), cscale(uz, fx));
                else w = csub(cscale(uy, fx), cscale(ux, fy));
                lds[3][n * LDR + c] = mk<C>(-w.y, w.x);
            }
            __syncthreads();
            fft_tile<T, N, true, TC, LDR, NTZ>(lds[3], tw, tid);
            tile_store<C, N, TC, LDR, NTZ>((C*)a.A[3 + q] + base, lds[3], rowstride, tid);
            __syncthreads();
        }
        fft_tile<T, N, true, TC, LDR, NTZ, 3, N * LDR>(lds[0], tw, tid);
#pragma unroll
        for (int q = 0; q < 3; ++q) tile_store<C, N, TC, LDR, NTZ>((C*)a.A[q] + base, lds[q], rowstride, tid);
        __syncthreads();   // the stores have read the tiles before the next tile is put
    }
}

// ------------------------------------------------------------------ y stage
template <typename T, int N, bool INV, bool SYM, int TC>
__global__ void __launch_bounds__(NT) k_y(KyArgs a)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    constexpr int LDR = ldr_of<TC>();
    __shared__ C lds[N * LDR];
    __shared__ C tws[N];
    const int tid = threadIdx.x;
    // tiles whose rows share a 128-byte line (P of them, side by side in kx) go to linear
    // ids L, L+8, ..., i.e. to one XCD, where the second finds the line in L2
    constexpr int P = (TC * (int)sizeof(C) >= 128) ? 1 : 128 / (TC * (int)sizeof(C));
    int kxb = blockIdx.x, z = blockIdx.y, fi = blockIdx.z;
    if constexpr (P > 1) {
        if ((M / TC) % P == 0 && (gridDim.y * gridDim.z) % 8 == 0) {
            const int t = xcd_tile<P>(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
            kxb = t % gridDim.x;
            z = (t / gridDim.x) % gridDim.y;
            fi = t / (gridDim.x * gridDim.y);
        }
    }
    const int kx0 = kxb * TC;
    stage_twiddles<C, N, NT>(tws, (const C*)a.tw, tid);
    const int rr = fi / a.nf;
    C* f = (C*)a.F[fi - rr * a.nf] + (size_t)rr * a.rstride;
    const size_t base = (size_t)z * N * M + kx0;
    tile_load<C, N, TC, LDR, NT>(lds, f + base, M, tid);
    __syncthreads();
    if constexpr (SYM) {
        // kx = 0 column: (c(ky) + conj c(-ky)) / 2, the z-transformed form of the
        // kx = 0 plane symmetrisation of flowlib.f90:264-329
        if (kx0 == 0) {
            for (int k = tid; k <= N / 2; k += NT) {
                const int km = (N - k) & (N - 1);
                const C p = lds[k * LDR], m = lds[km * LDR];
                const C s = mk<C>((T)0.5 * (p.x + m.x), (T)0.5 * (p.y - m.y));
                lds[k * LDR] = s;
                lds[km * LDR] = cconj(s);
            }
            __syncthreads();
        }
    }
    fft_tile<T, N, INV, TC, LDR, NT>(lds, tws, tid);
    tile_store<C, N, TC, LDR, NT>(f + base, lds, M, tid, INV ? (T)1 : (T)a.scale);
}

// ------------------------------------------------------------------ x stage
// Half-length packing of a real line of N points as N/2 complex points.
template <int N> constexpr int kx_lines() { return 2048 / N; }   // 4 complex points per thread
#ifndef DDNS_KXF_PTS
#define DDNS_KXF_PTS 2048
#endif
// lines of one k_x_fused work-group: DDNS_KXF_PTS / N (QP = DDNS_KXF_PTS / 512 complex
// points per thread and field)
template <int N> constexpr int kxf_lines() { return DDNS_KXF_PTS / N >= 1 ? DDNS_KXF_PTS / N : 1; }

// C2R pre-processing of the tile (in place, pairs k and M-k of every line).
template <typename T, int N, int TL, int LDR>
__device__ __forceinline__ void c2r_pack(cpx<T>* lds, const cpx<T>* tw, int tid)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    for (int e = tid; e < TL * (M / 2 + 1); e += NT) {
        const int line = e / (M / 2 + 1), k = e % (M / 2 + 1);
        if (k == 0) {
            const C x0 = lds[line];
            lds[line] = mk<C>(x0.x, x0.x);               // X(N/2) = 0, Im X(0) ignored
        } else {
            const int k2 = M - k;
            const C xa = lds[k * LDR + line], xb = lds[k2 * LDR + line];
            C w = tw[k];                                  // exp(-2 pi i k / N)
            w.y = -w.y;                                   // exp(+2 pi i k / N)
            const C sa = cadd(xa, cconj(xb)), da = csub(xa, cconj(xb));
            const C za = cadd(sa, cmul(mk<C>(-w.y, w.x), da));       // + i w (xa - conj xb)
            if (k2 != k) {
                const C wb = mk<C>(-w.x, w.y);            // exp(+2 pi i (M-k)/N) = -conj(w)
                const C sb = cadd(xb, cconj(xa)), db = csub(xb, cconj(xa));
                const C zb = cadd(sb, cmul(mk<C>(-wb.y, wb.x), db));
                lds[k2 * LDR + line] = zb;
            }
            lds[k * LDR + line] = za;
        }
    }
}

// R2C post-processing: from Z = FFT_{N/2}(x even + i x odd) to X(k), k < N/2, scaled.
template <typename T, int N, int TL, int LDR>
__device__ __forceinline__ void r2c_unpack_store(const cpx<T>* lds, const cpx<T>* tw, cpx<T>* dst,
                                                 size_t line0, T scale, int tid)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    for (int e = tid; e < TL * (M / 2 + 1); e += NT) {
        const int line = e / (M / 2 + 1), k = e % (M / 2 + 1);
        C* out = dst + (line0 + line) * M;
        if (k == 0) {
            const C z0 = lds[line];
            out[0] = mk<C>((z0.x + z0.y) * scale, (T)0);
        } else {
            const int k2 = M - k;
            const C za = lds[k * LDR + line], zb = lds[k2 * LDR + line];
            const C w = tw[k];                            // exp(-2 pi i k / N)
            const C e1 = cadd(za, cconj(zb)), d1 = csub(za, cconj(zb));
            // X = e1/2 - (i/2) w d1
            const C wd = cmul(w, d1);
            const C xa = mk<C>((T)0.5 * (e1.x + wd.y), (T)0.5 * (e1.y - wd.x));
            out[k] = cscale(xa, scale);
            if (k2 != k) {
                const C wb = mk<C>(-w.x, w.y);            // exp(-2 pi i (M-k)/N) = -conj(w)
                const C e2 = cadd(zb, cconj(za)), d2 = csub(zb, cconj(za));
                const C wd2 = cmul(wb, d2);
                const C xb = mk<C>((T)0.5 * (e2.x + wd2.y), (T)0.5 * (e2.y - wd2.x));
                out[k2] = cscale(xb, scale);
            }
        }
    }
}

// two neighbouring nodes (x = 2m, 2m + 1) of the physical velocity, three components
// each, 6 values in a row: three 2-vector stores (p is 2-vector aligned)
__device__ __forceinline__ void store_node_pair(float* p, float a0, float b0, float c0, float a1, float b1, float c1)
{
    ((float2*)p)[0] = make_float2(a0, b0);
    ((float2*)p)[1] = make_float2(c0, a1);
    ((float2*)p)[2] = make_float2(b1, c1);
}
__device__ __forceinline__ void store_node_pair(double* p, double a0, double b0, double c0, double a1, double b1,
                                                double c1)
{
    ((double2*)p)[0] = make_double2(a0, b0);
    ((double2*)p)[1] = make_double2(c0, a1);
    ((double2*)p)[2] = make_double2(b1, c1);
}

// Fused x stage: six C2R (u, w), physical u for the droplets, u x w, three R2C.
// The C2R transforms run two fields at a time and the R2C transforms three at a time
// through the same FFT passes, so every pass has two or three butterflies per lane;
// the next pair's loads are in flight while the current pair is transformed.
template <typename T, int N>
__global__ void __launch_bounds__(NT) k_x_fused(KxArgs a)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    constexpr int TL = kxf_lines<N>();
    constexpr int LDR = TL + 1;
    constexpr int TS = M * LDR;
    constexpr int QP = TL * M / NT;
    static_assert((TL * M) % NT == 0, "the lines must split evenly over the work-group");
    __shared__ C lds[3 * TS];
    __shared__ C tws[N], twhs[M];
    const int tid = threadIdx.x;
    const size_t line0 = (size_t)blockIdx.x * TL;
    const size_t roff = (size_t)blockIdx.y * a.rstride;     // this realisation's fields
    stage_twiddles<C, N, NT>(tws, (const C*)a.tw, tid);
    stage_twiddles<C, M, NT>(twhs, (const C*)a.twh, tid);
    const C* tw = tws;
    const C* twh = twhs;
    C ph[6][QP];
    TileRegs<C, TL, M, NT> next0, next1;
    tile_fetch<C, TL, M, NT>(next0, (const C*)a.A[0] + roff + line0 * M, M, tid);
    tile_fetch<C, TL, M, NT>(next1, (const C*)a.A[1] + roff + line0 * M, M, tid);
#pragma unroll
    for (int f = 0; f < 6; f += 2) {
        tile_put<C, TL, M, LDR, NT, true>(lds, next0, tid);
        tile_put<C, TL, M, LDR, NT, true>(lds + TS, next1, tid);
        if (f < 4) {
            tile_fetch<C, TL, M, NT>(next0, (const C*)a.A[f + 2] + roff + line0 * M, M, tid);
            tile_fetch<C, TL, M, NT>(next1, (const C*)a.A[f + 3] + roff + line0 * M, M, tid);
        }
        __syncthreads();
        c2r_pack<T, N, TL, LDR>(lds, tw, tid);
        c2r_pack<T, N, TL, LDR>(lds + TS, tw, tid);
        __syncthreads();
        fft_tile<T, M, true, TL, LDR, NT, 2, TS>(lds, twh, tid);
#pragma unroll
        for (int q = 0; q < QP; ++q) {
            const int e = tid + NT * q;
            const int line = e / M, m = e % M;
            ph[f][q] = lds[m * LDR + line];
            ph[f + 1][q] = lds[TS + m * LDR + line];
        }
        __syncthreads();
    }
    // physical velocity for the droplet gather: [line][x][3], the two nodes of a complex
    // value in three vector stores
    {
        T* up = (T*)a.uphys + (size_t)blockIdx.y * a.pstride + line0 * (size_t)N * 3;
#pragma unroll
        for (int q = 0; q < QP; ++q) {
            const int e = tid + NT * q;
            const int line = e / M, m = e % M;
            T* p = up + ((size_t)line * N + 2 * m) * 3;
            store_node_pair(p, ph[0][q].x, ph[1][q].x, ph[2][q].x, ph[0][q].y, ph[1][q].y, ph[2][q].y);
        }
    }
    // nonlinear term u x w (turb.f90:634-637) in natural component order, into the
    // three tiles
#pragma unroll
    for (int q = 0; q < QP; ++q) {
        const int e = tid + NT * q;
        const int line = e / M, m = e % M;
        const C ux = ph[0][q], uy = ph[1][q], uz = ph[2][q];
        const C wx = ph[3][q], wy = ph[4][q], wz = ph[5][q];
        lds[m * LDR + line] = mk<C>(uy.x * wz.x - uz.x * wy.x, uy.y * wz.y - uz.y * wy.y);
        lds[TS + m * LDR + line] = mk<C>(uz.x * wx.x - ux.x * wz.x, uz.y * wx.y - ux.y * wz.y);
        lds[2 * TS + m * LDR + line] = mk<C>(ux.x * wy.x - uy.x * wx.x, ux.y * wy.y - uy.y * wx.y);
    }
    __syncthreads();
    fft_tile<T, M, false, TL, LDR, NT, 3, TS>(lds, twh, tid);
    const T scale = (T)1 / (T)N;
#pragma unroll
    for (int c = 0; c < 3; ++c) r2c_unpack_store<T, N, TL, LDR>(lds + c * TS, tw, (C*)a.B[c] + roff, line0, scale, tid);
}

// Stand-alone x transforms of one field (kernel-level parity of mpifft3DRC/CR).
template <typename T, int N, bool INV>
__global__ void __launch_bounds__(NT) k_x_plain(const void* src, void* dst, const void* tw_, const void* twh_)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    constexpr int TL = kx_lines<N>();
    constexpr int LDR = TL + 1;
    __shared__ C lds[M * LDR];
    const int tid = threadIdx.x;
    const size_t line0 = (size_t)blockIdx.x * TL;
    const C* tw = (const C*)tw_;
    const C* twh = (const C*)twh_;
    // both directions see the line as M complex values (real pairs for the C2R output / R2C input)
    tile_load<C, TL, M, LDR, NT, true>(lds, (const C*)src + line0 * M, M, tid);
    __syncthreads();
    if constexpr (INV) {
        c2r_pack<T, N, TL, LDR>(lds, tw, tid);
        __syncthreads();
        fft_tile<T, M, true, TL, LDR, NT>(lds, twh, tid);
        C* d = (C*)dst + line0 * M;
        for (int e = tid; e < TL * M; e += NT) {
            const int line = e / M, m = e % M;
            d[(size_t)line * M + m] = lds[m * LDR + line];
        }
    } else {
        fft_tile<T, M, false, TL, LDR, NT>(lds, twh, tid);
        r2c_unpack_store<T, N, TL, LDR>(lds, tw, (C*)dst, line0, (T)1 / (T)N, tid);
    }
}

// ------------------------------------------------------------------ low-mode pre-kernel
// DET (forcing = 1): the new value of every mode with |k| < 2.5 is computed from a
// direct z-DFT of the y-transformed nonlinear term, then the shell-1 and shell-2
// energies (kx = 0 halved, flowlib.f90:207-217) give ff = sqrt(force/E); low[0..1].
// STH (forcing = 2): ranff white noise, OU step, projection and Hermitian fix-up
// (flowlib.f90:399-510) into low[8 .. 8+450).
constexpr int LOW_DFT = 512;   // offset of the supforDET z-DFT sums in FlowBuffers::low

struct PreArgs {
    const void* B[3];
    const void* U[3];
    const void* O[3];
    const void* tw;
    double* low;
    int* sth_state;
    double* sth_b;
    double dt, rnu, force0, force1, ou_tf, ou_var, pi2;
    int nek, forcing;  // forcing < 0: each realisation's kind and amplitudes from its table
    size_t rstride;    // complex values between realisations of one field
};

__device__ double dev_ranff(int* idum, int* iy, int* iv)
{
    const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB_ = 32;
    const int NDIV = 1 + (IM - 1) / NTAB_;
    const double AM = 1.0 / IM, EPS = 1.2e-7, RNMX = 1.0 - EPS;
    int j, k;
    if (*idum <= 0 || *iy == 0) {
        *idum = (-*idum > 1) ? -*idum : 1;
        // 8 warm-up draws, then the table from iv[31] down to iv[0]: the same
        // sequence as one loop j = 40..1 storing at j <= 32, written so that no
        // store index is ever outside the table
        for (j = 0; j < 8; ++j) {
            k = *idum / IQ;
            *idum = IA * (*idum - k * IQ) - IR * k;
            if (*idum < 0) *idum += IM;
        }
        for (j = NTAB_ - 1; j >= 0; --j) {
            k = *idum / IQ;
            *idum = IA * (*idum - k * IQ) - IR * k;
            if (*idum < 0) *idum += IM;
            iv[j] = *idum;
        }
        *iy = iv[0];
    }
    k = *idum / IQ;
    *idum = IA * (*idum - k * IQ) - IR * k;
    if (*idum < 0) *idum += IM;
    j = 1 + *iy / NDIV;
    *iy = iv[j - 1];
    iv[j - 1] = *idum;
    const double r = AM * (double)(*iy);
    return r < RNMX ? r : RNMX;
}

// supforDET's 225 low-mode sums (75 modes kx in 0..2, ky, kz in -2..2, times 3
// components) of the forward z transform of N-hat, by direct DFT: one work-group per
// sum, lanes over z, a fixed-order reduction; the result, scaled by 1/N like the
// forward FFT, goes to low[LOW_DFT ...] for k_pre.
template <typename T, int N>
__global__ void __launch_bounds__(NT) k_pre_dft(PreArgs a)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const int id = blockIdx.x, comp = id / 75, mi = id % 75;
    const int kx = mi % 3, kyw = (mi / 3) % 5 - 2, kzw = mi / 15 - 2;
    const int ky = (kyw + N) % N, kzi = (kzw + N) % N;
    double* low = a.low + (size_t)blockIdx.y * LOW_STRIDE;     // realisation blockIdx.y
    if (real_forcing(a.forcing, low) != 1) return;
    const C* b = (const C*)a.B[comp] + (size_t)blockIdx.y * a.rstride;
    const C* tw = (const C*)a.tw;
    double sr = 0.0, si = 0.0;
    for (int z = threadIdx.x; z < N; z += NT) {
        const C v = b[((size_t)z * N + ky) * M + kx];
        const C w = tw[(kzi * z) & (N - 1)];
        sr += (double)v.x * (double)w.x - (double)v.y * (double)w.y;
        si += (double)v.x * (double)w.y + (double)v.y * (double)w.x;
    }
    __shared__ double wr[NT / 64], wi[NT / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o, 64);
        si += __shfl_xor(si, o, 64);
    }
    if ((threadIdx.x & 63) == 0) { wr[threadIdx.x >> 6] = sr; wi[threadIdx.x >> 6] = si; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, i = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) { r += wr[w]; i += wi[w]; }
        low[LOW_DFT + id] = r / N;
        low[LOW_DFT + 225 + id] = i / N;
    }
}

template <typename T, int N>
__global__ void __launch_bounds__(NT) k_pre(PreArgs a)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const int tid = threadIdx.x;
    const int rr = blockIdx.x;                                       // realisation
    double* low = a.low + (size_t)rr * LOW_STRIDE;
    const int forcing = real_forcing(a.forcing, low);
    const bool own = a.forcing < 0;
    if (forcing == 2) {
        if (tid != 0) return;
        int* st = a.sth_state + (size_t)rr * STH_STATE_STRIDE;
        int idum = st[0], iy = st[1];
        // shuffle table in LDS: a private array indexed by a data-dependent j is
        // lowered to indexed register moves, which gives no protection at all
        // against a stray index
        __shared__ int iv[32];
        for (int i = 0; i < 32; ++i) iv[i] = st[2 + i];
        const double dt = a.dt, tf = own ? low[LOW_PAR + 3] : a.ou_tf, var = own ? low[LOW_PAR + 4] : a.ou_var;
        const double pi2 = a.pi2;
        double fr[3][150];
        for (int mc = 0; mc < 3; ++mc)                                   // flowlib.f90:405-419
            for (int k = 0; k < 5; ++k)
                for (int j = 0; j < 5; ++j)
                    for (int i = 0; i < 6; i += 2) {
                        double cont1 = dev_ranff(&idum, &iy, iv);
                        if (cont1 < 1.e-6) cont1 = 1.e-6;
                        const double rmag = sqrt(-4.0 * var * dt * log(cont1) / tf);
                        const double rph = dev_ranff(&idum, &iy, iv);
                        fr[mc][i + 6 * (j + 5 * k)] = rmag * cos(pi2 * rph);
                        fr[mc][i + 1 + 6 * (j + 5 * k)] = rmag * sin(pi2 * rph);
                    }
        for (int mc = 0; mc < 3; ++mc) {                                 // 424-440
            for (int k = 1; k <= 5; ++k) {
                const int kk = (k == 1) ? 1 : 7 - k;
                for (int j = 2; j <= 3; ++j) {
                    const int jj = 7 - j;
                    fr[mc][0 + 6 * ((jj - 1) + 5 * (kk - 1))] = fr[mc][0 + 6 * ((j - 1) + 5 * (k - 1))];
                    fr[mc][1 + 6 * ((jj - 1) + 5 * (kk - 1))] = -fr[mc][1 + 6 * ((j - 1) + 5 * (k - 1))];
                }
            }
            for (int k = 2; k <= 3; ++k) {
                const int kk = 7 - k;
                fr[mc][0 + 6 * (5 * (kk - 1))] = fr[mc][0 + 6 * (5 * (k - 1))];
                fr[mc][1 + 6 * (5 * (kk - 1))] = -fr[mc][1 + 6 * (5 * (k - 1))];
            }
            fr[mc][0] = 0.0;
            fr[mc][1] = 0.0;
        }
        double* b = a.sth_b + (size_t)rr * STH_B_STRIDE;
        double* fa = low + 8;
        const double smallx = 1.0e-18, xkf = sqrt(8.0);
        for (int k = 0; k < 5; ++k)                                      // 447-478
            for (int j = 0; j < 5; ++j)
                for (int i = 0; i < 6; ++i) {
                    const int o = i + 6 * (j + 5 * k);
                    const double kx_ = (double)(i / 2);
                    const double ky_ = (double)(j <= 2 ? j : j - 5);
                    const double kz_ = (double)(k <= 2 ? k : k - 5);
                    const double radm = sqrt(kx_ * kx_ + ky_ * ky_ + kz_ * kz_);
                    if (radm < xkf) {
                        b[o] = b[o] * (1.0 - dt / tf) + fr[0][o];
                        b[150 + o] = b[150 + o] * (1.0 - dt / tf) + fr[1][o];
                        b[300 + o] = b[300 + o] * (1.0 - dt / tf) + fr[2][o];
                        const double t1 = kx_ * b[o], t2 = ky_ * b[150 + o], t3 = kz_ * b[300 + o];
                        double divb = (t1 + t2 + t3) / ((radm + smallx) * (radm + smallx));
                        if (radm == 0.) divb = 0.0;
                        fa[o] = b[o] - kx_ * divb;
                        fa[150 + o] = b[150 + o] - ky_ * divb;
                        fa[300 + o] = b[300 + o] - kz_ * divb;
                    } else {
                        fa[o] = 0.0; fa[150 + o] = 0.0; fa[300 + o] = 0.0;
                    }
                }
        for (int mc = 0; mc < 3; ++mc) { fa[150 * mc] = 0.0; fa[150 * mc + 1] = 0.0; }   // 481-483
        for (int k = 1; k <= 5; ++k) {                                   // 488-500
            const int kk = (k == 1) ? 1 : 7 - k;
            for (int j = 2; j <= 3; ++j) {
                const int jj = 7 - j;
                for (int mc = 0; mc < 3; ++mc) {
                    fa[150 * mc + 6 * ((jj - 1) + 5 * (kk - 1))] = fa[150 * mc + 6 * ((j - 1) + 5 * (k - 1))];
                    fa[150 * mc + 1 + 6 * ((jj - 1) + 5 * (kk - 1))] = -fa[150 * mc + 1 + 6 * ((j - 1) + 5 * (k - 1))];
                }
            }
        }
        for (int k = 2; k <= 3; ++k) {                                   // 502-510
            const int kk = 7 - k;
            for (int mc = 0; mc < 3; ++mc) {
                fa[150 * mc + 6 * (5 * (kk - 1))] = fa[150 * mc + 6 * (5 * (k - 1))];
                fa[150 * mc + 1 + 6 * (5 * (kk - 1))] = -fa[150 * mc + 1 + 6 * (5 * (k - 1))];
            }
        }
        st[0] = idum; st[1] = iy;
        for (int i = 0; i < 32; ++i) st[2 + i] = iv[i];
        return;
    }
    if (forcing != 1) return;
    // 75 low modes: kx in 0..2, ky, kz in -2..2; 225 (mode, component) DFT sums
    __shared__ double nre[225], nim[225];
    __shared__ double emode[75];
    if (tid < 225) {                      // the z-DFT sums of k_pre_dft
        nre[tid] = low[LOW_DFT + tid];
        nim[tid] = low[LOW_DFT + 225 + tid];
    }
    __syncthreads();
    if (tid < 75) {
        const int kx = tid % 3, kyw = (tid / 3) % 5 - 2, kzw = tid / 15 - 2;
        const int ky = (kyw + N) % N, kz = (kzw + N) % N;
        const size_t g = ((size_t)kz * N + ky) * M + kx;
        C nn[3], uu[3], oo[3], vv[3], nd[3];
        for (int q = 0; q < 3; ++q) {
            nn[q] = mk<C>((T)nre[q * 75 + tid], (T)nim[q * 75 + tid]);
            uu[q] = ((const C*)a.U[q])[(size_t)rr * a.rstride + g];
            oo[q] = ((const C*)a.O[q])[(size_t)rr * a.rstride + g];
        }
        // the update without the forcing factor (ff = 1 here)
        double one[2] = {1.0, 1.0};
        update_mode<T>(kx, ky, kz, N, a.nek, 0, (T)a.dt, (T)a.rnu, one, nn, uu, oo, vv, nd);
        const double k2 = (double)(kx * kx + kyw * kyw + kzw * kzw);
        const int ik2 = (int)(sqrt(k2 == 0.0 ? 1.e-5 : k2) + 0.5);
        double e = 0.0;
        if (ik2 == 1 || ik2 == 2) {
            for (int q = 0; q < 3; ++q) e += (double)vv[q].x * (double)vv[q].x + (double)vv[q].y * (double)vv[q].y;
            if (kx == 0) e *= 0.5;
            e = (ik2 == 1) ? e : -e;      // sign tags the shell
        }
        emode[tid] = e;
    }
    __syncthreads();
    if (tid == 0) {
        double e1 = 0.0, e2 = 0.0;
        for (int i = 0; i < 75; ++i) {
            if (emode[i] > 0) e1 += emode[i];
            else e2 -= emode[i];
        }
        low[0] = sqrt((own ? low[LOW_PAR + 1] : a.force0) / e1);
        low[1] = sqrt((own ? low[LOW_PAR + 2] : a.force1) / e2);
        low[2] = e1;
        low[3] = e2;
    }
}

// ------------------------------------------------------------------ kx = 0 plane of u-hat
template <typename T, int N>
__global__ void k_sym(void* u0, void* u1, void* u2, size_t rstride, double* low)
{
    // the z stage before this kernel wrote every mode beyond the dealiasing sphere
    if (blockIdx.x == 0 && threadIdx.x == 0) low[(size_t)blockIdx.y * LOW_STRIDE + LOW_ZOUT] = 1.0;
    using C = cpx<T>;
    constexpr int M = N / 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N * N) return;
    const int kz = t / N, ky = t % N;
    const int mz = (N - kz) & (N - 1), my = (N - ky) & (N - 1);
    if (mz * N + my < t) return;           // the partner with the smaller index does the pair
    void* fs[3] = {u0, u1, u2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        C* f = (C*)fs[q] + (size_t)blockIdx.y * rstride;     // realisation blockIdx.y
        const size_t gp = ((size_t)kz * N + ky) * M, gm = ((size_t)mz * N + my) * M;
        const C p = f[gp], m = f[gm];
        const C s = mk<C>((T)0.5 * (p.x + m.x), (T)0.5 * (p.y - m.y));
        f[gp] = s;
        f[gm] = cconj(s);
    }
}

// ------------------------------------------------------------------ initial field
// The random start of initial_flow (flowlib.f90:53-144) on the device.  The reference
// draws from Intel's IFPORT rand, which is not available; here every mode draws its own
// pair of uniforms from a counter-based generator, SplitMix64 (Steele, Lea and Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014): output i of the
// stream with key s is the SplitMix64 finaliser of s + (i + 1) * 0x9E3779B97F4A7C15.
__device__ __forceinline__ double init_uniform(unsigned long long key, unsigned long long i)
{
    unsigned long long z = key + (i + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;                      // [0, 1)
}

// Box-Muller pair of component c at mode (kz, ky, kx), as flowlib.f90:158-172 forms it
// from two uniforms (each held above 1e-10)
template <int N>
__device__ __forceinline__ double2 init_gauss(unsigned long long key, int c, int kz, int ky, int kx)
{
    const unsigned long long m = (((unsigned long long)c * N + kz) * N + ky) * (N / 2) + kx;
    double t1 = init_uniform(key, 2 * m), t2 = init_uniform(key, 2 * m + 1);
    if (t1 <= 1.e-10) t1 = 1.e-10;
    if (t2 <= 1.e-10) t2 = 1.e-10;
    const double a = sqrt(-2.0 * log(t1)), ph = 2.0 * 3.14159265358979323846 * t2;
    return make_double2(a * cos(ph), a * sin(ph));
}

// One thread per mode: the three Gaussian components, the Hermitian kx = 0 plane
// ((g(k) + conj g(-k)) / 2, computed from both draws so the two halves agree), the
// Nyquist planes and the mean set to zero, the projection onto k.u = 0 (flowlib.f90:
// 94-101); then twice the shell's mode count, 2 per mode with kx > 0 and 1 on the kx = 0
// plane (flowlib.f90:107-118).  The shell energies are summed afterwards by the E(k)
// reduction (kx = 0 modes at half weight), whose fixed order makes the field the same
// bits on every run.
template <typename T, int N>
__global__ void k_init_modes(void* u0, void* u1, void* u2, unsigned long long key, int nek, unsigned long long* cnt2)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * M) return;
    const int kx = (int)(t % M), ky = (int)((t / M) % N), kz = (int)(t / ((size_t)N * M));
    const int kyw = wave_index(ky, N), kzw = wave_index(kz, N);
    const bool zero = ky == N / 2 || kz == N / 2 || (kx == 0 && ky == 0 && kz == 0);
    double2 g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        g[c] = init_gauss<N>(key, c, kz, ky, kx);
        if (kx == 0) {
            const double2 h = init_gauss<N>(key, c, (N - kz) & (N - 1), (N - ky) & (N - 1), 0);
            g[c] = make_double2(0.5 * (g[c].x + h.x), 0.5 * (g[c].y - h.y));
        }
        if (zero) g[c] = make_double2(0.0, 0.0);
    }
    const double k[3] = {(double)kx, (double)kyw, (double)kzw};
    double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
    if (kx == 0 && ky == 0 && kz == 0) k2 = 1.e-5;
    const double pr = (k[0] * g[0].x + k[1] * g[1].x + k[2] * g[2].x) / k2;
    const double pi = (k[0] * g[0].y + k[1] * g[1].y + k[2] * g[2].y) / k2;
    void* u[3] = {u0, u1, u2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double x = g[c].x - k[c] * pr, y = g[c].y - k[c] * pi;
        ((C*)u[c])[t] = mk<C>((T)x, (T)y);
    }
    const int ik2 = (int)(sqrt(k2) + 0.5);
    if (ik2 < 1 || ik2 > nek) return;
    // twice the shell's mode count (kx = 0 modes count one half): integer sums, the same
    // in any order; the shell energies come from the fixed-tree E(k) reduction
    atomicAdd(cnt2 + ik2 - 1, kx == 0 ? 1ull : 2ull);
}

// the shell factors (flowlib.f90:119-137), and zero above nek (139-141)
template <typename T, int N>
__global__ void k_init_scale(void* u0, void* u1, void* u2, int nek, const double* fac)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * M) return;
    const int kx = (int)(t % M), kyw = wave_index((int)((t / M) % N), N), kzw = wave_index((int)(t / ((size_t)N * M)), N);
    double k2 = (double)(kx * kx + kyw * kyw + kzw * kzw);
    if (k2 == 0.0) k2 = 1.e-5;
    const int ik2 = (int)(sqrt(k2) + 0.5);
    if (ik2 < 1) return;
    const double f = ik2 <= nek ? fac[ik2 - 1] : 0.0;
    void* u[3] = {u0, u1, u2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        C v = ((C*)u[c])[t];
        v = mk<C>((T)((double)v.x * f), (T)((double)v.y * f));
        ((C*)u[c])[t] = v;
    }
}

// ================================================================== host side
static size_t field_elems(int N) { return (size_t)N * N * (N / 2); }

int flow_alloc(FlowBuffers& fb, int N, int precision, int R)
{
    const size_t cs = (precision == DDNS_FP64) ? sizeof(double2) : sizeof(float2);
    const size_t rs = (precision == DDNS_FP64) ? sizeof(double) : sizeof(float);
    if (R < 1) R = 1;
    fb.R = R;
    fb.rstride = R > 1 ? field_elems(N) : 0;
    fb.pstride = R > 1 ? (size_t)N * N * N * 3 : 0;
    const size_t fe = field_elems(N) * R;
    for (int i = 0; i < 3; ++i) {
        DDNS_HIP(hipMalloc(&fb.U[i], fe * cs));
        DDNS_HIP(hipMalloc(&fb.O[i], fe * cs));
        DDNS_HIP(hipMalloc(&fb.B[i], fe * cs));
        DDNS_HIP(hipMemset(fb.U[i], 0, fe * cs));
        DDNS_HIP(hipMemset(fb.O[i], 0, fe * cs));
    }
    for (int i = 0; i < 6; ++i) DDNS_HIP(hipMalloc(&fb.A[i], fe * cs));
    DDNS_HIP(hipMalloc(&fb.uphys, (size_t)N * N * N * 3 * rs * R));
    DDNS_HIP(hipMemset(fb.uphys, 0, (size_t)N * N * N * 3 * rs * R));
    // twiddles exp(-2 pi i m / L), computed in double and rounded once
    for (int pass = 0; pass < 2; ++pass) {
        const int L = pass == 0 ? N : N / 2;
        std::vector<double> d(2 * (size_t)L);
        for (int m = 0; m < L; ++m) {
            const double ang = -2.0 * M_PI * (double)m / (double)L;
            d[2 * m] = std::cos(ang);
            d[2 * m + 1] = std::sin(ang);
        }
        void** dst = pass == 0 ? &fb.tw : &fb.twh;
        if (precision == DDNS_FP64) {
            DDNS_HIP(hipMalloc(dst, d.size() * sizeof(double)));
            DDNS_HIP(hipMemcpy(*dst, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(d.begin(), d.end());
            DDNS_HIP(hipMalloc(dst, f.size() * sizeof(float)));
            DDNS_HIP(hipMemcpy(*dst, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    DDNS_HIP(hipMalloc(&fb.low, (size_t)LOW_STRIDE * R * sizeof(double)));
    DDNS_HIP(hipMemset(fb.low, 0, (size_t)LOW_STRIDE * R * sizeof(double)));
    DDNS_HIP(hipMalloc(&fb.sth_state, (size_t)STH_STATE_STRIDE * R * sizeof(int)));
    DDNS_HIP(hipMemset(fb.sth_state, 0, (size_t)STH_STATE_STRIDE * R * sizeof(int)));
    DDNS_HIP(hipMalloc(&fb.sth_b, (size_t)STH_B_STRIDE * R * sizeof(double)));
    DDNS_HIP(hipMemset(fb.sth_b, 0, (size_t)STH_B_STRIDE * R * sizeof(double)));
    DDNS_HIP(hipMalloc(&fb.spec_part, ((size_t)SPEC_BLOCKS + 1) * SPEC_NEK_MAX * sizeof(double)));
    return 0;
}

FlowBuffers flow_view(const FlowBuffers& fb, int r, int precision)
{
    const size_t cs = (precision == DDNS_FP64) ? sizeof(double2) : sizeof(float2);
    const size_t rs = (precision == DDNS_FP64) ? sizeof(double) : sizeof(float);
    FlowBuffers v{};
    auto off = [&](void* p) { return p ? (void*)((char*)p + (size_t)r * fb.rstride * cs) : nullptr; };
    for (int i = 0; i < 3; ++i) { v.U[i] = off(fb.U[i]); v.O[i] = off(fb.O[i]); v.B[i] = off(fb.B[i]); }
    for (int i = 0; i < 6; ++i) v.A[i] = off(fb.A[i]);
    v.uphys = (char*)fb.uphys + (size_t)r * fb.pstride * rs;
    v.tw = fb.tw; v.twh = fb.twh;
    v.low = fb.low + (size_t)r * LOW_STRIDE;
    v.sth_state = fb.sth_state + (size_t)r * STH_STATE_STRIDE;
    v.sth_b = fb.sth_b + (size_t)r * STH_B_STRIDE;
    v.ky_tc = fb.ky_tc; v.kz_nt = fb.kz_nt;
    return v;
}

int flow_set_realisation(const FlowBuffers& fb, int r, int forcing, double force0, double force1, double ou_tf,
                         double ou_var)
{
    const double par[5] = {(double)forcing, force0, force1, ou_tf, ou_var};
    DDNS_HIP(hipMemcpy(fb.low + (size_t)r * LOW_STRIDE + LOW_PAR, par, sizeof(par), hipMemcpyHostToDevice));
    return 0;
}

void flow_free(FlowBuffers& fb)
{
    for (int i = 0; i < 3; ++i) { (void)hipFree(fb.U[i]); (void)hipFree(fb.O[i]); (void)hipFree(fb.B[i]); }
    for (int i = 0; i < 6; ++i) (void)hipFree(fb.A[i]);
    (void)hipFree(fb.uphys); (void)hipFree(fb.tw); (void)hipFree(fb.twh);
    (void)hipFree(fb.low); (void)hipFree(fb.sth_state); (void)hipFree(fb.sth_b);
    for (int q = 0; q < 8; ++q) if (fb.dg[q]) (void)hipFree(fb.dg[q]);
    if (fb.dg_part) (void)hipFree(fb.dg_part);
    if (fb.spec_part) (void)hipFree(fb.spec_part);
    fb = FlowBuffers{};
}

template <typename T, int N>
static int launch_z(int mode, const FlowBuffers& fb, const FlowConst& fc, hipStream_t s,
                    const void* src = nullptr, void* dst = nullptr, double scale = 1.0)
{
    KzArgs a{};
    for (int i = 0; i < 3; ++i) { a.B[i] = fb.B[i]; a.U[i] = fb.U[i]; a.O[i] = fb.O[i]; }
    for (int i = 0; i < 6; ++i) a.A[i] = fb.A[i];
    a.tw = fb.tw; a.low = fb.low; a.dt = fc.dt; a.rnu = fc.rnu; a.nek = fc.nek; a.forcing = fc.forcing;
    a.scale = scale;
    a.rstride = fb.rstride;
    const int R = fb.R;
    if (mode == KZ_PLAIN_FWD || mode == KZ_PLAIN_INV) { a.B[0] = src; a.U[0] = dst; a.rstride = 0; }
    constexpr int TC = kz_tc<T, N>();
    dim3 grid((N / 2 / TC) * N);
    const char* env = (mode == KZ_FULL || mode == KZ_CURL_INV) ? std::getenv("DDNS_Z_SPLIT") : nullptr;
    // a batch (R > 1) always runs the fused z stage: the split kernels take one realisation
    const bool split = R == 1 && (env && (env[0] == '0' || env[0] == '1') ? env[0] == '1' : kz_split_default<T, N>());
    if (split && (mode == KZ_FULL || mode == KZ_CURL_INV)) {
        const unsigned gp = (unsigned)(N / 2 / kz_tc_plain<T, N>() * N);
        const unsigned ge = (unsigned)(((size_t)N * N * (N / 2) + 255) / 256);
        KzArgs p = a;
        int st;
        if (mode == KZ_FULL) {
            p.scale = 1.0 / N;
            for (int q = 0; q < 3; ++q) {                  // forward z of N-hat, in place, x 1/N
                p.B[0] = a.B[q]; p.U[0] = (void*)a.B[q];
                hipLaunchKernelGGL((k_z<T, N, KZ_PLAIN_FWD>), dim3(gp), dim3(kz_threads<T, N, KZ_PLAIN_FWD>()), 0, s, p);
            }
            hipLaunchKernelGGL((k_z_update<T, N>), dim3(ge), dim3(256), 0, s, a);      // and the curl
        } else {
            hipLaunchKernelGGL((k_z_curl<T, N>), dim3(ge), dim3(256), 0, s, a);
        }
        p.scale = 1.0;
        for (int q = 0; q < 6; ++q) {                      // inverse z of u-hat and w-hat
            p.B[0] = q < 3 ? (const void*)a.U[q] : (const void*)a.A[q];
            p.U[0] = a.A[q];
            hipLaunchKernelGGL((k_z<T, N, KZ_PLAIN_INV>), dim3(gp), dim3(kz_threads<T, N, KZ_PLAIN_INV>()), 0, s, p);
        }
        st = 0;
        DDNS_HIP(hipGetLastError());
        return st;
    }
    if (mode == KZ_PLAIN_FWD || mode == KZ_PLAIN_INV) grid = dim3((N / 2 / kz_tc_plain<T, N>()) * N);
    if (mode == KZ_FULL) {
        // persistent: one work-group per CU (the LDS tiles allow no more), of fb.kz_nt
        // threads (0: the default)
        int dev = 0, ncu = 0;
        DDNS_HIP(hipGetDevice(&dev));
        DDNS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int ntiles = (int)grid.x * R;
        const dim3 g(std::max(1, std::min(ncu, ntiles)));
        constexpr int N0 = kz_nt_cand<T, N>(0), N1 = kz_nt_cand<T, N>(1), N2 = kz_nt_cand<T, N>(2);
        if constexpr (kz_nt_fits<T, N, N1>()) {
            if (fb.kz_nt == N1) {
                hipLaunchKernelGGL((k_z_full<T, N, N1>), g, dim3(N1), 0, s, a, ntiles);
                DDNS_HIP(hipGetLastError());
                return 0;
            }
        }
        if constexpr (kz_nt_fits<T, N, N2>()) {
            if (fb.kz_nt == N2) {
                hipLaunchKernelGGL((k_z_full<T, N, N2>), g, dim3(N2), 0, s, a, ntiles);
                DDNS_HIP(hipGetLastError());
                return 0;
            }
        }
        hipLaunchKernelGGL((k_z_full<T, N, N0>), g, dim3(N0), 0, s, a, ntiles);
        DDNS_HIP(hipGetLastError());
        return 0;
    }
    if (mode == KZ_CURL_INV) grid.y = R;
    switch (mode) {
    case KZ_CURL_INV: hipLaunchKernelGGL((k_z<T, N, KZ_CURL_INV>), grid, dim3(NT), 0, s, a); break;
    case KZ_PLAIN_FWD: hipLaunchKernelGGL((k_z<T, N, KZ_PLAIN_FWD>), grid, dim3(kz_threads<T, N, KZ_PLAIN_FWD>()), 0, s, a); break;
    default: hipLaunchKernelGGL((k_z<T, N, KZ_PLAIN_INV>), grid, dim3(kz_threads<T, N, KZ_PLAIN_INV>()), 0, s, a); break;
    }
    DDNS_HIP(hipGetLastError());
    return 0;
}

template <typename T, int N, int TC>
static void launch_y_tc(bool inv, bool sym, const KyArgs& a, int nf, hipStream_t s)
{
    dim3 grid(N / 2 / TC, N, nf * a.R);
    if (inv && sym) hipLaunchKernelGGL((k_y<T, N, true, true, TC>), grid, dim3(NT), 0, s, a);
    else if (inv) hipLaunchKernelGGL((k_y<T, N, true, false, TC>), grid, dim3(NT), 0, s, a);
    else hipLaunchKernelGGL((k_y<T, N, false, false, TC>), grid, dim3(NT), 0, s, a);
}

// candidate k of the y-stage tile widths (0 = half, 1 = default, 2 = twice the default)
template <typename T, int N, int K>
static bool launch_y_cand(bool inv, bool sym, const KyArgs& a, int nf, hipStream_t s)
{
    constexpr int TC = ky_cand<T, N>(K);
    if constexpr (ky_fits<T, N, TC>()) {
        launch_y_tc<T, N, TC>(inv, sym, a, nf, s);
        return true;
    } else {
        return false;
    }
}

// the y stage at the tile width fb.ky_tc (0: the default)
template <typename T, int N>
static int launch_y(bool inv, bool sym, void* const* f, int nf, const FlowBuffers& fb, hipStream_t s)
{
    KyArgs a{};
    for (int i = 0; i < nf; ++i) a.F[i] = f[i];
    a.tw = fb.tw;
    a.scale = 1.0 / N;
    a.nf = nf;
    a.R = fb.R;
    a.rstride = fb.rstride;
    const int tc = fb.ky_tc;
    bool done = false;
    if (tc == ky_cand<T, N>(0)) done = launch_y_cand<T, N, 0>(inv, sym, a, nf, s);
    else if (tc == ky_cand<T, N>(2)) done = launch_y_cand<T, N, 2>(inv, sym, a, nf, s);
    if (!done) launch_y_tc<T, N, ky_tc<T, N>()>(inv, sym, a, nf, s);
    DDNS_HIP(hipGetLastError());
    return 0;
}

template <typename T, int N>
static int launch_x(const FlowBuffers& fb, hipStream_t s)
{
    KxArgs a{};
    for (int i = 0; i < 6; ++i) a.A[i] = fb.A[i];
    for (int i = 0; i < 3; ++i) a.B[i] = fb.B[i];
    a.uphys = fb.uphys; a.tw = fb.tw; a.twh = fb.twh;
    a.rstride = fb.rstride; a.pstride = fb.pstride;
    dim3 grid((unsigned)((size_t)N * N / kxf_lines<N>()), fb.R);
    hipLaunchKernelGGL((k_x_fused<T, N>), grid, dim3(NT), 0, s, a);
    DDNS_HIP(hipGetLastError());
    return 0;
}

template <typename T, int N>
static int launch_pre(const FlowBuffers& fb, const FlowConst& fc, hipStream_t s)
{
    if (fc.forcing != 1 && fc.forcing != 2 && fc.forcing >= 0) return 0;
    PreArgs a{};
    for (int i = 0; i < 3; ++i) { a.B[i] = fb.B[i]; a.U[i] = fb.U[i]; a.O[i] = fb.O[i]; }
    a.tw = fb.tw; a.low = fb.low; a.sth_state = fb.sth_state; a.sth_b = fb.sth_b;
    a.dt = fc.dt; a.rnu = fc.rnu; a.force0 = fc.force0; a.force1 = fc.force1;
    a.ou_tf = fc.ou_tf; a.ou_var = fc.ou_var; a.pi2 = fc.pi2; a.nek = fc.nek; a.forcing = fc.forcing;
    a.rstride = fb.rstride;
    if (fc.forcing == 1 || fc.forcing < 0) hipLaunchKernelGGL((k_pre_dft<T, N>), dim3(225, fb.R), dim3(NT), 0, s, a);
    hipLaunchKernelGGL((k_pre<T, N>), dim3(fb.R), dim3(NT), 0, s, a);
    DDNS_HIP(hipGetLastError());
    return 0;
}

template <typename T, int N>
static int launch_sym(const FlowBuffers& fb, hipStream_t s)
{
    const int tot = N * N;
    hipLaunchKernelGGL((k_sym<T, N>), dim3((tot + 255) / 256, fb.R), dim3(256), 0, s, fb.U[0], fb.U[1], fb.U[2],
                       fb.rstride, fb.low);
    DDNS_HIP(hipGetLastError());
    return 0;
}

template <typename T, int N>
static int step_impl(const FlowBuffers& fb, const FlowConst& fc, hipStream_t s, hipEvent_t* ev, int part,
                     unsigned evflags)
{
    // FLOW_INV: inverse y + fused x (produces the physical velocity);
    // FLOW_FWD: forward y + pre + z + sym (closes the step, opens the next), or its two
    // pieces FLOW_FWD_Y (forward y + pre) and FLOW_FWD_Z (z + sym).  Events: 0-2 around
    // the first half, 3 forward y, 4 pre, 5 end of pre, 8 start of z, 6 end of z, 7 end.
    // evflags = hipEventRecordExternal while the stream is being captured into a graph:
    // the timing events then become event-record nodes of the graph
    if (part == FLOW_INV) {
        if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[0], s, evflags));
        int st = launch_y<T, N>(true, true, fb.A, 6, fb, s);
        if (st) return st;
        if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[1], s, evflags));
        st = launch_x<T, N>(fb, s);
        if (st) return st;
        if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[2], s, evflags));
        return 0;
    }
    int st = 0;
    if (part != FLOW_FWD_Z) {
        if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[3], s, evflags));
        void* bb[3] = {fb.B[0], fb.B[1], fb.B[2]};
        st = launch_y<T, N>(false, false, bb, 3, fb, s);
        if (st) return st;
        if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[4], s, evflags));
        st = launch_pre<T, N>(fb, fc, s);
        if (st) return st;
        if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[5], s, evflags));
        if (part == FLOW_FWD_Y) return 0;
    }
    if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[8], s, evflags));
    st = launch_z<T, N>(KZ_FULL, fb, fc, s);
    if (st) return st;
    if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[6], s, evflags));
    st = launch_sym<T, N>(fb, s);
    if (st) return st;
    if (ev) DDNS_HIP(hipEventRecordWithFlags(ev[7], s, evflags));
    return 0;
}

#define DDNS_DISPATCH_N(N_, CALL)                         \
    switch (N_) {                                          \
    case 32: { constexpr int NN = 32; CALL; } break;       \
    case 64: { constexpr int NN = 64; CALL; } break;       \
    case 128: { constexpr int NN = 128; CALL; } break;     \
    case 256: { constexpr int NN = 256; CALL; } break;     \
    case 512: { constexpr int NN = 512; CALL; } break;     \
    case 1024: { constexpr int NN = 1024; CALL; } break;   \
    default: set_error("unsupported grid size"); return DDNS_ERR_ARG; \
    }

int flow_prologue(const FlowBuffers& fb, const FlowConst& fc, int precision, hipStream_t s)
{
    int st = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(fc.N, st = (launch_z<double, NN>(KZ_CURL_INV, fb, fc, s))) }
    else { DDNS_DISPATCH_N(fc.N, st = (launch_z<float, NN>(KZ_CURL_INV, fb, fc, s))) }
    return st;
}

int flow_step(const FlowBuffers& fb, const FlowConst& fc, int precision, int part, hipStream_t s, hipEvent_t* ev,
              bool capturing)
{
    const unsigned evflags = capturing ? hipEventRecordExternal : hipEventRecordDefault;
    int st = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(fc.N, st = (step_impl<double, NN>(fb, fc, s, ev, part, evflags))) }
    else { DDNS_DISPATCH_N(fc.N, st = (step_impl<float, NN>(fb, fc, s, ev, part, evflags))) }
    return st;
}

// ---------------------------------------------------------------- host conversion
template <typename T>
static int upload_spectral(void* dev, int N, const double* v, hipStream_t s)
{
    using C = cpx<T>;
    const int L = N + 2, M = N / 2;
    std::vector<C> h(field_elems(N));
    for (int kz = 0; kz < N; ++kz)
        for (int ky = 0; ky < N; ++ky)
            for (int kx = 0; kx < M; ++kx) {
                const size_t o = (size_t)kz + (size_t)L * ((size_t)(2 * kx) + (size_t)L * ky);
                h[((size_t)kz * N + ky) * M + kx] = mk<C>((T)v[o], (T)v[o + L]);
            }
    DDNS_HIP(hipMemcpyAsync(dev, h.data(), h.size() * sizeof(C), hipMemcpyHostToDevice, s));
    DDNS_HIP(hipStreamSynchronize(s));
    return 0;
}

template <typename T>
static int download_spectral(const void* dev, int N, double* v, hipStream_t s)
{
    using C = cpx<T>;
    const int L = N + 2, M = N / 2;
    std::vector<C> h(field_elems(N));
    DDNS_HIP(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(C), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    std::memset(v, 0, sizeof(double) * (size_t)L * L * N);
    for (int kz = 0; kz < N; ++kz)
        for (int ky = 0; ky < N; ++ky)
            for (int kx = 0; kx < M; ++kx) {
                const size_t o = (size_t)kz + (size_t)L * ((size_t)(2 * kx) + (size_t)L * ky);
                const C c = h[((size_t)kz * N + ky) * M + kx];
                v[o] = (double)c.x;
                v[o + L] = (double)c.y;
            }
    return 0;
}

// the state's modes beyond the dealiasing sphere may be nonzero again (an upload, a new
// field): the next z stage loads and stores them all
static int mark_dense(const FlowBuffers& fb)
{
    static const double zero = 0.0;
    DDNS_HIP(hipMemcpy(fb.low + LOW_ZOUT, &zero, sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

int flow_upload_state(const FlowBuffers& fb, int N, int precision, const double* const src[6], hipStream_t s)
{
    DDNS_HIP(hipStreamSynchronize(s));
    const int sm = mark_dense(fb);
    if (sm) return sm;
    for (int f = 0; f < 6; ++f) {
        if (!src[f]) continue;
        void* dev = f < 3 ? fb.U[f] : fb.O[f - 3];
        const int st = precision == DDNS_FP64 ? upload_spectral<double>(dev, N, src[f], s)
                                              : upload_spectral<float>(dev, N, src[f], s);
        if (st) return st;
    }
    return 0;
}

int flow_download_state(const FlowBuffers& fb, int N, int precision, double* const dst[6], hipStream_t s)
{
    for (int f = 0; f < 6; ++f) {
        if (!dst[f]) continue;
        const void* dev = f < 3 ? fb.U[f] : fb.O[f - 3];
        const int st = precision == DDNS_FP64 ? download_spectral<double>(dev, N, dst[f], s)
                                              : download_spectral<float>(dev, N, dst[f], s);
        if (st) return st;
    }
    return 0;
}

template <typename T>
static int download_phys_t(const void* dev, int N, double* ux, double* uy, double* uz, hipStream_t s)
{
    const size_t n3 = (size_t)N * N * N;
    std::vector<T> h(n3 * 3);
    DDNS_HIP(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(T), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    double* d[3] = {ux, uy, uz};
    for (int c = 0; c < 3; ++c)
        if (d[c])
            for (size_t i = 0; i < n3; ++i) d[c][i] = (double)h[3 * i + c];
    return 0;
}

int flow_download_phys(const FlowBuffers& fb, int N, int precision, double* ux, double* uy, double* uz, hipStream_t s)
{
    return precision == DDNS_FP64 ? download_phys_t<double>(fb.uphys, N, ux, uy, uz, s)
                                  : download_phys_t<float>(fb.uphys, N, ux, uy, uz, s);
}

// ---------------------------------------------------------------- initial field
template <typename T, int N>
static int init_impl(FlowBuffers& fb, int nek, int kind, int seed, double k0, double urms, double force0,
                     hipStream_t s)
{
    const size_t tot = field_elems(N);
    const unsigned g = (unsigned)((tot + 255) / 256);
    double* d = nullptr;                       // twice the mode counts (integers), then the factors (2 nek)
    DDNS_HIP(hipMalloc(&d, 2 * (size_t)nek * sizeof(double)));
    DDNS_HIP(hipMemsetAsync(d, 0, (size_t)nek * sizeof(double), s));
    const unsigned long long key = (unsigned long long)(long long)seed;
    hipLaunchKernelGGL((k_init_modes<T, N>), dim3(g), dim3(256), 0, s, fb.U[0], fb.U[1], fb.U[2], key, nek,
                       (unsigned long long*)d);
    DDNS_HIP(hipGetLastError());
    std::vector<unsigned long long> c2(nek);
    std::vector<double> h(2 * (size_t)nek), fac(nek, 1.0);
    DDNS_HIP(hipMemcpyAsync(c2.data(), d, c2.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    // shell energies of the unscaled field (kx = 0 modes at half weight), by the same
    // fixed-order reduction as E(k): the same bits on every run
    int sst = flow_spectrum(fb, N, nek, sizeof(T) == 8 ? DDNS_FP64 : DDNS_FP32_MIXED, h.data(), s);
    if (sst) { (void)hipFree(d); return sst; }
    for (int i = 0; i < nek; ++i) h[nek + i] = 0.5 * (double)c2[i];
    const double* e_t = h.data();
    const double* enumv = h.data() + nek;
    const double pi = 4. * std::atan(1.);
    if (kind == DDNS_SPEC_K4_GAUSS) {
        // E(k) proportional to (k/k0)^4 exp(-2 (k/k0)^2), total 3/2 urms^2
        std::vector<double> target(nek);
        double sum = 0.0;
        for (int i = 1; i <= nek; ++i) {
            const double q = (double)i / k0;
            target[i - 1] = q * q * q * q * std::exp(-2.0 * q * q);
            sum += target[i - 1];
        }
        for (int i = 1; i <= nek; ++i) {
            const double tg = target[i - 1] * 1.5 * urms * urms / sum;
            fac[i - 1] = e_t[i - 1] > 0 ? std::sqrt(tg / e_t[i - 1]) : 0.0;
        }
    } else {
        // initial_flow's own scaling: shell 1 at force(1) per shell volume, k^(-5/3) above
        const double ccc = force0 * 4. / 3. * pi * (1.5 * 1.5 * 1.5 - 0.5 * 0.5 * 0.5) / enumv[0];
        for (int i = 1; i <= nek; ++i) {
            const double vol = 4. / 3. * pi * ((i + 0.5) * (i + 0.5) * (i + 0.5) - (i - 0.5) * (i - 0.5) * (i - 0.5));
            fac[i - 1] = e_t[i - 1] > 0 ? std::sqrt(ccc / e_t[i - 1] * enumv[i - 1] / vol) / std::pow((double)i, 0.833333)
                                        : 0.0;
        }
    }
    DDNS_HIP(hipMemcpyAsync(d + nek, fac.data(), nek * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_init_scale<T, N>), dim3(g), dim3(256), 0, s, fb.U[0], fb.U[1], fb.U[2], nek,
                       (const double*)(d + nek));
    DDNS_HIP(hipGetLastError());
    // the history starts equal to the velocity (turb.f90:452-454)
    for (int c = 0; c < 3; ++c)
        DDNS_HIP(hipMemcpyAsync(fb.O[c], fb.U[c], tot * sizeof(cpx<T>), hipMemcpyDeviceToDevice, s));
    DDNS_HIP(hipStreamSynchronize(s));
    DDNS_HIP(hipFree(d));
    return 0;
}

int flow_init(FlowBuffers& fb, int N, int nek, int precision, int kind, int seed, double k0, double urms, double force0,
              hipStream_t s)
{
    int st = 0;
    DDNS_HIP(hipStreamSynchronize(s));
    st = mark_dense(fb);
    if (st) return st;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(N, st = (init_impl<double, NN>(fb, nek, kind, seed, k0, urms, force0, s))) }
    else { DDNS_DISPATCH_N(N, st = (init_impl<float, NN>(fb, nek, kind, seed, k0, urms, force0, s))) }
    return st;
}

// ---------------------------------------------------------------- stand-alone FFTs
template <typename T, int N>
static int r2c_impl(const FlowBuffers& fb, double* v, hipStream_t s)
{
    using C = cpx<T>;
    const int L = N + 2;
    const size_t n3 = (size_t)N * N * N;
    std::vector<T> h(n3);
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x)
                h[((size_t)z * N + y) * N + x] = (T)v[(size_t)x + (size_t)L * ((size_t)y + (size_t)L * z)];
    // A[0] holds the real input (same byte size as one complex field), B[0] the output
    DDNS_HIP(hipMemcpyAsync(fb.A[0], h.data(), n3 * sizeof(T), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_x_plain<T, N, false>), dim3((unsigned)((size_t)N * N / kx_lines<N>())), dim3(NT), 0, s,
                       fb.A[0], fb.B[0], fb.tw, fb.twh);
    DDNS_HIP(hipGetLastError());
    void* f[1] = {fb.B[0]};
    int st = launch_y<T, N>(false, false, f, 1, fb, s);
    if (st) return st;
    FlowConst fc{};
    fc.N = N;
    st = launch_z<T, N>(KZ_PLAIN_FWD, fb, fc, s, fb.B[0], fb.A[1], 1.0 / N);
    if (st) return st;
    return download_spectral<T>(fb.A[1], N, v, s);
}

template <typename T, int N>
static int c2r_impl(const FlowBuffers& fb, double* v, hipStream_t s)
{
    const int L = N + 2;
    const size_t n3 = (size_t)N * N * N;
    int st = upload_spectral<T>(fb.A[1], N, v, s);
    if (st) return st;
    FlowConst fc{};
    fc.N = N;
    st = launch_z<T, N>(KZ_PLAIN_INV, fb, fc, s, fb.A[1], fb.B[0], 1.0);
    if (st) return st;
    void* f[1] = {fb.B[0]};
    st = launch_y<T, N>(true, false, f, 1, fb, s);
    if (st) return st;
    hipLaunchKernelGGL((k_x_plain<T, N, true>), dim3((unsigned)((size_t)N * N / kx_lines<N>())), dim3(NT), 0, s,
                       fb.B[0], fb.A[0], fb.tw, fb.twh);
    DDNS_HIP(hipGetLastError());
    std::vector<T> h(n3);
    DDNS_HIP(hipMemcpyAsync(h.data(), fb.A[0], n3 * sizeof(T), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    std::memset(v, 0, sizeof(double) * (size_t)L * L * N);
    for (int z = 0; z < N; ++z)
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x)
                v[(size_t)x + (size_t)L * ((size_t)y + (size_t)L * z)] = (double)h[((size_t)z * N + y) * N + x];
    return 0;
}

int fft_r2c_host(const FlowBuffers& fb, int N, int precision, double* v, hipStream_t s)
{
    int st = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(N, st = (r2c_impl<double, NN>(fb, v, s))) }
    else { DDNS_DISPATCH_N(N, st = (r2c_impl<float, NN>(fb, v, s))) }
    return st;
}

int fft_c2r_host(const FlowBuffers& fb, int N, int precision, double* v, hipStream_t s)
{
    int st = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(N, st = (c2r_impl<double, NN>(fb, v, s))) }
    else { DDNS_DISPATCH_N(N, st = (c2r_impl<float, NN>(fb, v, s))) }
    return st;
}

// ---------------------------------------------------------------- diagnostics
// POSTPROC (postprocessing.f90:76-272) and TRANSPOSTPROC (275-315) on the current
// u-hat.  Spectral fields are formed by the kernels below, taken to physical space by
// the plain z / y / x transforms into scratch fields, and reduced with fixed grids.

// dst = i k_c src (c = 0, 1, 2: x, y, z), the derivative of component src along c
// (postprocessing.f90:97-111)
template <typename T, int N>
__global__ void k_mul_ik(const void* src, void* dst, int c)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * M) return;
    const int kx = (int)(t % M), ky = (int)((t / M) % N), kz = (int)(t / ((size_t)N * M));
    const T k = (T)(c == 0 ? kx : (c == 1 ? wave_index(ky, N) : wave_index(kz, N)));
    const C v = ((const C*)src)[t];
    ((C*)dst)[t] = mk<C>(-(k * v.y), k * v.x);
}

// dst = i (k x u)_c for c < 3 (the vorticity, turb.f90:506-517), or i k . u for c = 3
// (the divergence, postprocessing.f90:219-223)
template <typename T, int N>
__global__ void k_curl_div(const void* u0, const void* u1, const void* u2, void* dst, int c)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * M) return;
    const int kx = (int)(t % M), ky = (int)((t / M) % N), kz = (int)(t / ((size_t)N * M));
    const T fx = (T)kx, fy = (T)wave_index(ky, N), fz = (T)wave_index(kz, N);
    const C a = ((const C*)u0)[t], b = ((const C*)u1)[t], d = ((const C*)u2)[t];
    C w;
    if (c == 0) w = csub(cscale(d, fy), cscale(b, fz));
    else if (c == 1) w = csub(cscale(a, fz), cscale(d, fx));
    else if (c == 2) w = csub(cscale(b, fx), cscale(a, fy));
    else w = cadd(cadd(cscale(a, fx), cscale(b, fy)), cscale(d, fz));
    ((C*)dst)[t] = mk<C>(-w.y, w.x);
}

// in place: (w0, w1, w2) <- u x w, the nonlinear term in physical space (turb.f90:634-637)
template <typename T>
__global__ void k_cross(const T* __restrict__ u0, const T* __restrict__ u1, const T* __restrict__ u2,
                        T* w0, T* w1, T* w2, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T ux = u0[i], uy = u1[i], uz = u2[i], wx = w0[i], wy = w1[i], wz = w2[i];
    w0[i] = uy * wz - uz * wy;
    w1[i] = uz * wx - ux * wz;
    w2[i] = ux * wy - uy * wx;
}

constexpr int DG_BLOCKS = 256;   // fixed grid of the diagnostic reductions

// per work-group sums over a real field: mode 0: f^2, f^3, f^4 of one field;
// mode 1: the max of |f0| + |f1| + |f2|; mode 2: sum |g|, |g|^2 and |g| f0 with
// |g| = sqrt(f1^2 + f2^2 + f3^2) (vorticity magnitude against the divergence f0);
// mode 3: f, f^2 of one field
template <typename T>
__global__ void __launch_bounds__(256) k_dg_reduce(const T* f0, const T* f1, const T* f2, const T* f3, size_t n,
                                                  int mode, double* part)
{
    __shared__ double sh[3][256];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (mode == 0) {
            const double v = (double)f0[i], v2 = v * v;
            a0 += v2; a1 += v2 * v; a2 += v2 * v2;
        } else if (mode == 1) {
            const double v = fabs((double)f0[i]) + fabs((double)f1[i]) + fabs((double)f2[i]);
            a0 = fmax(a0, v);
        } else if (mode == 3) {
            const double v = (double)f0[i];
            a0 += v; a1 += v * v;
        } else {
            const double w1 = (double)f1[i], w2 = (double)f2[i], w3 = (double)f3[i];
            const double g = sqrt(w1 * w1 + w2 * w2 + w3 * w3);
            a0 += g; a1 += g * g; a2 += g * (double)f0[i];
        }
    }
    sh[0][threadIdx.x] = a0; sh[1][threadIdx.x] = a1; sh[2][threadIdx.x] = a2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            if (mode == 1) sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
            else {
                sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
                sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
                sh[2][threadIdx.x] += sh[2][threadIdx.x + o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sh[0][0];
        part[gridDim.x + blockIdx.x] = sh[1][0];
        part[2 * gridDim.x + blockIdx.x] = sh[2][0];
    }
}

// shell sums over |k| shells 1..nek (kx = 0 plane halved, postprocessing.f90:143-170,
// 285-313): e = |u|^2, d = 2 nu k^2 |u|^2, tr = Re(conj(u) . F) with F the nonlinear term
template <typename T, int N>
__global__ void k_shells(const void* u0, const void* u1, const void* u2, const void* F0, const void* F1,
                         const void* F2, int nek, double rnu, double* e, double* d, double* tr)
{
    using C = cpx<T>;
    constexpr int M = N / 2;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)N * N * M) return;
    const int kx = (int)(t % M), ky = (int)((t / M) % N), kz = (int)(t / ((size_t)N * M));
    const int kyw = wave_index(ky, N), kzw = wave_index(kz, N);
    double k2 = (double)(kx * kx + kyw * kyw + kzw * kzw);
    if (k2 == 0.0) k2 = 1.e-5;
    const int ik2 = (int)(sqrt(k2) + 0.5);
    if (ik2 < 1 || ik2 > nek) return;
    const C a = ((const C*)u0)[t], b = ((const C*)u1)[t], c = ((const C*)u2)[t];
    const double h = kx == 0 ? 0.5 : 1.0;
    const double ee = h * ((double)a.x * a.x + (double)a.y * a.y + (double)b.x * b.x + (double)b.y * b.y +
                           (double)c.x * c.x + (double)c.y * c.y);
    atomicAdd(e + ik2 - 1, ee);
    atomicAdd(d + ik2 - 1, 2.0 * rnu * ee * k2);
    if (F0) {
        const C p = ((const C*)F0)[t], q = ((const C*)F1)[t], r = ((const C*)F2)[t];
        const double tt = (double)a.x * p.x + (double)a.y * p.y + (double)b.x * q.x + (double)b.y * q.y +
                          (double)c.x * r.x + (double)c.y * r.y;
        atomicAdd(tr + ik2 - 1, h * tt);
    }
}

template <typename T, int N>
static int c2r_dev(const FlowBuffers& fb, const void* src, void* tmp, void* dst, hipStream_t s)
{
    FlowConst fc{};
    fc.N = N;
    int st = launch_z<T, N>(KZ_PLAIN_INV, fb, fc, s, src, tmp, 1.0);
    if (st) return st;
    void* f[1] = {tmp};
    st = launch_y<T, N>(true, false, f, 1, fb, s);
    if (st) return st;
    hipLaunchKernelGGL((k_x_plain<T, N, true>), dim3((unsigned)((size_t)N * N / kx_lines<N>())), dim3(NT), 0, s, tmp, dst,
                       fb.tw, fb.twh);
    DDNS_HIP(hipGetLastError());
    return 0;
}

template <typename T, int N>
static int r2c_dev(const FlowBuffers& fb, const void* src, void* tmp, void* dst, hipStream_t s)
{
    hipLaunchKernelGGL((k_x_plain<T, N, false>), dim3((unsigned)((size_t)N * N / kx_lines<N>())), dim3(NT), 0, s, src, tmp,
                       fb.tw, fb.twh);
    DDNS_HIP(hipGetLastError());
    void* f[1] = {tmp};
    int st = launch_y<T, N>(false, false, f, 1, fb, s);
    if (st) return st;
    FlowConst fc{};
    fc.N = N;
    return launch_z<T, N>(KZ_PLAIN_FWD, fb, fc, s, tmp, dst, 1.0 / N);
}

// partials of a k_dg_reduce launch summed (or maxed) on the host in block order
template <typename T>
static int dg_reduce(const T* f0, const T* f1, const T* f2, const T* f3, size_t n, int mode, double* part, double out[3],
                     hipStream_t s)
{
    hipLaunchKernelGGL((k_dg_reduce<T>), dim3(DG_BLOCKS), dim3(256), 0, s, f0, f1, f2, f3, n, mode, part);
    DDNS_HIP(hipGetLastError());
    double h[3 * DG_BLOCKS];
    DDNS_HIP(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < 3; ++q) {
        double a = 0.0;
        for (int b = 0; b < DG_BLOCKS; ++b) a = mode == 1 ? fmax(a, h[q * DG_BLOCKS + b]) : a + h[q * DG_BLOCKS + b];
        out[q] = a;
    }
    return 0;
}

template <typename T, int N>
static int diag_impl(FlowBuffers& fb, int nek, double rnu, int want, double* ek, double* dk, double* tk, double* mom,
                     hipStream_t s)
{
    const size_t n3 = (size_t)N * N * N, nc = field_elems(N);
    const size_t fbytes = n3 * sizeof(T);                    // one real field = one complex field
    if (!fb.dg[0]) {
        for (int q = 0; q < 8; ++q) DDNS_HIP(hipMalloc(&fb.dg[q], fbytes));
        // reduction partials (3 DG_BLOCKS) or the 3 nek shell sums, nek < N / 2
        DDNS_HIP(hipMalloc(&fb.dg_part, 3 * (size_t)std::max(DG_BLOCKS, N / 2) * sizeof(double)));
    }
    T* R[6];
    for (int q = 0; q < 6; ++q) R[q] = (T*)fb.dg[q];
    void* ctmp = fb.dg[6];
    void* cout = fb.dg[7];
    const unsigned gc = (unsigned)((nc + 255) / 256), gr = (unsigned)((n3 + 255) / 256);
    int st;
    // physical u and w
    for (int c = 0; c < 3; ++c) {
        st = c2r_dev<T, N>(fb, fb.U[c], ctmp, R[c], s);
        if (st) return st;
        hipLaunchKernelGGL((k_curl_div<T, N>), dim3(gc), dim3(256), 0, s, fb.U[0], fb.U[1], fb.U[2], cout, c);
        st = c2r_dev<T, N>(fb, cout, ctmp, R[3 + c], s);
        if (st) return st;
    }
    if (want & 1) {
        // max |u| + |v| + |w| (the CFL number, postprocessing.f90:204-207)
        double o[3];
        st = dg_reduce<T>(R[0], R[1], R[2], nullptr, n3, 1, fb.dg_part, o, s);
        if (st) return st;
        mom[0] = o[0];
        // longitudinal velocity derivatives: sums of d^2, d^3, d^4 over the three
        // components (postprocessing.f90:97-137), into the ctmp / cout scratch
        double s2 = 0.0, s3 = 0.0, s4 = 0.0;
        for (int c = 0; c < 3; ++c) {
            hipLaunchKernelGGL((k_mul_ik<T, N>), dim3(gc), dim3(256), 0, s, fb.U[c], cout, c);
            // the derivative field replaces nothing needed later: use the second half of R
            // only after the vorticity moments (below) have been taken
            st = c2r_dev<T, N>(fb, cout, ctmp, cout, s);
            if (st) return st;
            st = dg_reduce<T>((const T*)cout, nullptr, nullptr, nullptr, n3, 0, fb.dg_part, o, s);
            if (st) return st;
            s2 += o[0]; s3 += o[1]; s4 += o[2];
        }
        mom[1] = s2; mom[2] = s3; mom[3] = s4;
    }
    if (want & 2) {
        // divergence and vorticity statistics (postprocessing.f90:217-270)
        hipLaunchKernelGGL((k_curl_div<T, N>), dim3(gc), dim3(256), 0, s, fb.U[0], fb.U[1], fb.U[2], cout, 3);
        st = c2r_dev<T, N>(fb, cout, ctmp, cout, s);
        if (st) return st;
        double o[3], g[3];
        st = dg_reduce<T>((const T*)cout, nullptr, nullptr, nullptr, n3, 3, fb.dg_part, o, s);
        if (st) return st;
        st = dg_reduce<T>((const T*)cout, R[3], R[4], R[5], n3, 2, fb.dg_part, g, s);
        if (st) return st;
        mom[4] = o[0]; mom[5] = o[1];                  // sum div, sum div^2
        mom[6] = g[0]; mom[7] = g[1]; mom[8] = g[2];   // sum |w|, sum |w|^2, sum |w| div
    }
    // spectra; the transfer needs the nonlinear term F = FFT(u x w) (turb.f90:634-650)
    double* dd = fb.dg_part;           // 3 nek shell sums
    DDNS_HIP(hipMemsetAsync(dd, 0, 3 * (size_t)nek * sizeof(double), s));
    const void* F[3] = {nullptr, nullptr, nullptr};
    if (tk) {
        hipLaunchKernelGGL((k_cross<T>), dim3(gr), dim3(256), 0, s, R[0], R[1], R[2], R[3], R[4], R[5], n3);
        DDNS_HIP(hipGetLastError());
        for (int c = 0; c < 3; ++c) {
            st = r2c_dev<T, N>(fb, R[3 + c], ctmp, R[c], s);        // F_c over u_c, no longer needed
            if (st) return st;
            F[c] = R[c];
        }
    }
    hipLaunchKernelGGL((k_shells<T, N>), dim3(gc), dim3(256), 0, s, fb.U[0], fb.U[1], fb.U[2], F[0], F[1], F[2], nek, rnu,
                       dd, dd + nek, dd + 2 * nek);
    DDNS_HIP(hipGetLastError());
    std::vector<double> h(3 * (size_t)nek);
    DDNS_HIP(hipMemcpyAsync(h.data(), dd, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nek; ++i) {
        if (ek) ek[i] = h[i];
        if (dk) dk[i] = h[nek + i];
        if (tk) tk[i] = h[2 * nek + i];
    }
    return 0;
}

int flow_diagnostics(FlowBuffers& fb, int N, int nek, double rnu, int precision, int want, double* ek, double* dk,
                     double* tk, double* mom, hipStream_t s)
{
    if (nek > N / 2) { set_error("too many shells for the diagnostics"); return DDNS_ERR_ARG; }
    int st = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(N, st = (diag_impl<double, NN>(fb, nek, rnu, want, ek, dk, tk, mom, s))) }
    else { DDNS_DISPATCH_N(N, st = (diag_impl<float, NN>(fb, nek, rnu, want, ek, dk, tk, mom, s))) }
    return st;
}

// ---------------------------------------------------------------- per-kernel timing
template <typename T, int N>
static int profile_impl(const FlowBuffers& fb, const FlowConst& fc, int k, double* ms, int nk, hipStream_t s)
{
    hipEvent_t ev[FLOW_EVENTS];
    for (int i = 0; i < FLOW_EVENTS; ++i) DDNS_HIP(hipEventCreate(&ev[i]));
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int it = 0; it < k; ++it) {
        int st = step_impl<T, N>(fb, fc, s, ev, FLOW_INV, hipEventRecordDefault);
        if (st) return st;
        st = step_impl<T, N>(fb, fc, s, ev, FLOW_FWD, hipEventRecordDefault);
        if (st) return st;
        DDNS_HIP(hipEventSynchronize(ev[7]));
        float t;
        DDNS_HIP(hipEventElapsedTime(&t, ev[0], ev[1])); acc[0] += t;   // y inverse
        DDNS_HIP(hipEventElapsedTime(&t, ev[1], ev[2])); acc[1] += t;   // x fused
        DDNS_HIP(hipEventElapsedTime(&t, ev[3], ev[4])); acc[2] += t;   // y forward
        DDNS_HIP(hipEventElapsedTime(&t, ev[4], ev[5])); acc[3] += t;   // pre
        DDNS_HIP(hipEventElapsedTime(&t, ev[8], ev[6])); acc[4] += t;   // z stage
        DDNS_HIP(hipEventElapsedTime(&t, ev[6], ev[7])); acc[5] += t;   // sym
        DDNS_HIP(hipEventElapsedTime(&t, ev[0], ev[7])); acc[6] += t;   // whole
    }
    for (int i = 0; i < nk && i < 7; ++i) ms[i] = acc[i] / k;
    for (int i = 0; i < FLOW_EVENTS; ++i) (void)hipEventDestroy(ev[i]);
    return 0;
}

// plain streaming copy, 16 B per lane per iteration: the HBM reference kernel
__global__ void __launch_bounds__(256) k_copy_calibration(const float4* __restrict__ src, float4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int flow_profile_copy(const FlowBuffers& fb, int N, int precision, int reps, double* ms, double* bytes, hipStream_t s)
{
    const size_t cs = (precision == DDNS_FP64) ? sizeof(double2) : sizeof(float2);
    const size_t n4 = field_elems(N) * cs / sizeof(float4);
    hipEvent_t e0, e1;
    DDNS_HIP(hipEventCreate(&e0));
    DDNS_HIP(hipEventCreate(&e1));
    const unsigned grid = 256 * 32;
    hipLaunchKernelGGL(k_copy_calibration, dim3(grid), dim3(256), 0, s, (const float4*)fb.A[0], (float4*)fb.A[1], n4);
    DDNS_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy_calibration, dim3(grid), dim3(256), 0, s, (const float4*)fb.A[0], (float4*)fb.A[1], n4);
    DDNS_HIP(hipGetLastError());
    DDNS_HIP(hipEventRecord(e1, s));
    DDNS_HIP(hipEventSynchronize(e1));
    float t = 0;
    DDNS_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / reps;
    *bytes = 2.0 * (double)n4 * sizeof(float4);
    return 0;
}

// ---------------------------------------------------------------- tile-shape tuner
// The role the FFTW_MEASURE plans play in the reference (mpifft.f90:36-56): at set-up,
// before any state is loaded, each candidate shape is timed on the handle's own buffers
// (cleared first), one untimed run and then the mean of three, and the fastest is kept:
//   - the y-stage tile width: the inverse stage over six fields and the forward stage
//     over three, as a step runs them;
//   - the fused z stage's work-group size (where the step uses the fused z stage).
// The choices are cached per grid and precision for the life of the process.
// DDNS_KY_TC / DDNS_KZ_NT in the environment force a shape (one that is not a candidate,
// or does not fit, means the default) and skip the timing; DDNS_TUNE=0 keeps the defaults.
// No choice changes the arithmetic of a column or of a mode, so every one gives the
// same bits.
template <typename T, int N>
static bool ky_cand_fits(int k)
{
    switch (k) {
    case 0: return ky_fits<T, N, ky_cand<T, N>(0)>();
    case 1: return ky_fits<T, N, ky_cand<T, N>(1)>();
    default: return ky_fits<T, N, ky_cand<T, N>(2)>();
    }
}

template <typename T, int N>
static bool kz_nt_cand_fits(int k)
{
    switch (k) {
    case 0: return kz_nt_fits<T, N, kz_nt_cand<T, N>(0)>();
    case 1: return kz_nt_fits<T, N, kz_nt_cand<T, N>(1)>();
    default: return kz_nt_fits<T, N, kz_nt_cand<T, N>(2)>();
    }
}

struct TuneRecord {
    int ky_tc, kz_nt;
    float ky_ms[3], kz_ms[3];
};

// the fused z stage is the one a step runs (not the separate kernels of the split path)
template <typename T, int N>
static bool z_fused_in_use()
{
    const char* env = std::getenv("DDNS_Z_SPLIT");
    return env && (env[0] == '0' || env[0] == '1') ? env[0] == '0' : !kz_split_default<T, N>();
}

// times `reps` runs of launch after one untimed run (ms per run), 0 on failure
template <typename F>
static float time_runs(F&& launch, int reps, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int* st)
{
    for (int rep = 0; rep <= reps && !*st; ++rep) {
        if (rep == 1 && hipEventRecord(e0, s) != hipSuccess) *st = DDNS_ERR_HIP;
        if (!*st) *st = launch();
    }
    float ms = 0.f;
    if (!*st && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                 hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        *st = DDNS_ERR_HIP;
    return *st ? 0.f : ms / (float)reps;
}

template <typename T, int N>
static int tune_impl(FlowBuffers& fb, const FlowConst& fc, hipStream_t s)
{
    static std::mutex mu;
    static std::map<int, TuneRecord> cache;          // key: N * 4096 + R, per precision (one map per T)
    fb.ky_tc = ky_tc<T, N>();
    fb.kz_nt = kz_nt<T, N>();
    for (int k = 0; k < 3; ++k) fb.ky_ms[k] = fb.kz_ms[k] = 0.f;
    bool forced = false;
    if (const char* f = std::getenv("DDNS_KY_TC")) {
        const int want = std::atoi(f);
        for (int k = 0; k < 3; ++k)
            if (want == ky_cand<T, N>(k) && ky_cand_fits<T, N>(k)) fb.ky_tc = want;
        forced = true;
    }
    if (const char* f = std::getenv("DDNS_KZ_NT")) {
        const int want = std::atoi(f);
        for (int k = 0; k < 3; ++k)
            if (want == kz_nt_cand<T, N>(k) && kz_nt_cand_fits<T, N>(k)) fb.kz_nt = want;
        forced = true;
    }
    if (forced) return 0;
    if (const char* t = std::getenv("DDNS_TUNE"))
        if (t[0] == '0') return 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(N * 4096 + fb.R);        // per grid and number of realisations
        if (it != cache.end()) {
            fb.ky_tc = it->second.ky_tc;
            fb.kz_nt = it->second.kz_nt;
            for (int k = 0; k < 3; ++k) {
                fb.ky_ms[k] = it->second.ky_ms[k];
                fb.kz_ms[k] = it->second.kz_ms[k];
            }
            return 0;
        }
    }
    const size_t bytes = field_elems(N) * sizeof(cpx<T>) * fb.R;
    for (int q = 0; q < 6; ++q) DDNS_HIP(hipMemsetAsync(fb.A[q], 0, bytes, s));
    for (int q = 0; q < 3; ++q) DDNS_HIP(hipMemsetAsync(fb.B[q], 0, bytes, s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DDNS_HIP(hipEventCreate(&e0));
    DDNS_HIP(hipEventCreate(&e1));
    for (int q = 0; q < 3; ++q) {
        DDNS_HIP(hipMemsetAsync(fb.U[q], 0, bytes, s));
        DDNS_HIP(hipMemsetAsync(fb.O[q], 0, bytes, s));
    }
    DDNS_HIP(hipMemsetAsync(fb.low, 0, (size_t)LOW_STRIDE * fb.R * sizeof(double), s));
    float best = 0.f;
    int best_tc = ky_tc<T, N>(), best_nt = kz_nt<T, N>();
    int st = 0;
    for (int k = 0; k < 3 && !st; ++k) {
        if (!ky_cand_fits<T, N>(k)) continue;
        fb.ky_tc = ky_cand<T, N>(k);
        fb.ky_ms[k] = time_runs([&]() {
            int r = launch_y<T, N>(true, true, fb.A, 6, fb, s);
            return r ? r : launch_y<T, N>(false, false, fb.B, 3, fb, s);
        }, 3, s, e0, e1, &st);
        if (!st && (best == 0.f || fb.ky_ms[k] < best)) {
            best = fb.ky_ms[k];
            best_tc = fb.ky_tc;
        }
    }
    if (z_fused_in_use<T, N>()) {
        best = 0.f;
        for (int k = 0; k < 3 && !st; ++k) {
            if (!kz_nt_cand_fits<T, N>(k)) continue;
            fb.kz_nt = kz_nt_cand<T, N>(k);
            fb.kz_ms[k] = time_runs([&]() { return launch_z<T, N>(KZ_FULL, fb, fc, s); }, 3, s, e0, e1, &st);
            if (!st && (best == 0.f || fb.kz_ms[k] < best)) {
                best = fb.kz_ms[k];
                best_nt = fb.kz_nt;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) {
        set_error("tile-width tuner: a HIP call failed");
        return st;
    }
    // the state the timing runs wrote is cleared again: a new handle starts at rest
    for (int q = 0; q < 3; ++q) {
        DDNS_HIP(hipMemsetAsync(fb.U[q], 0, bytes, s));
        DDNS_HIP(hipMemsetAsync(fb.O[q], 0, bytes, s));
    }
    DDNS_HIP(hipMemsetAsync(fb.low, 0, (size_t)LOW_STRIDE * fb.R * sizeof(double), s));
    DDNS_HIP(hipStreamSynchronize(s));
    fb.ky_tc = best_tc;
    fb.kz_nt = best_nt;
    std::lock_guard<std::mutex> lock(mu);
    TuneRecord r{};
    r.ky_tc = best_tc;
    r.kz_nt = best_nt;
    for (int k = 0; k < 3; ++k) {
        r.ky_ms[k] = fb.ky_ms[k];
        r.kz_ms[k] = fb.kz_ms[k];
    }
    cache[N * 4096 + fb.R] = r;
    return 0;
}

int flow_tune(FlowBuffers& fb, const FlowConst& fc, int precision, hipStream_t s)
{
    int st = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(fc.N, st = (tune_impl<double, NN>(fb, fc, s))) }
    else { DDNS_DISPATCH_N(fc.N, st = (tune_impl<float, NN>(fb, fc, s))) }
    return st;
}

int flow_default_kz_nt(int N, int precision)
{
    int nt = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(N, nt = (kz_nt<double, NN>())) }
    else { DDNS_DISPATCH_N(N, nt = (kz_nt<float, NN>())) }
    return nt;
}

int flow_default_ky_tc(int N, int precision)
{
    int tc = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(N, tc = (ky_tc<double, NN>())) }
    else { DDNS_DISPATCH_N(N, tc = (ky_tc<float, NN>())) }
    return tc;
}

int flow_profile(const FlowBuffers& fb, const FlowConst& fc, int precision, int k, double* ms, int nk, hipStream_t s)
{
    int st = 0;
    if (precision == DDNS_FP64) { DDNS_DISPATCH_N(fc.N, st = (profile_impl<double, NN>(fb, fc, k, ms, nk, s))) }
    else { DDNS_DISPATCH_N(fc.N, st = (profile_impl<float, NN>(fb, fc, k, ms, nk, s))) }
    return st;
}

}  // namespace ddns
