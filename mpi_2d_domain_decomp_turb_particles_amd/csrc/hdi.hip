// hdi.hip -- hydrodynamic interactions between droplets.
//
// Reference: NEI_LIST_PRECOMP partlib.f90:3218-3371 (neighbours within HDI_trunc
// radii of the *source* droplet, Stokes-sphere coefficients M = gamma r r + beta I),
// AFUN_PRECOMP 3375-3510 (A x = x + M x and b = M (V - U)), HDI_GMRES 2795-3005
// (unrestarted GMRES: modified Gram-Schmidt Arnoldi, explicit Givens factor Qn,
// back substitution, tolerance relaxed at iteration maxiter/2 and from maxiter-2,
// warm start from the previous disturbance velocity) and EUC_NORM 2756-2791.
//
// The system is solved in cell-sorted droplet order: row t is the droplet at sorted
// position t, so the neighbours a row gathers sit in the few cells around it and
// their vector entries are close in memory.  The solution is scattered back to the
// reference's droplet order at the end.  Neighbour lists are ELL, slot-major: slot s
// of row t holds the neighbour's sorted position and the 6 distinct entries of its
// symmetric 3x3 block, at s * n + t, so a wavefront reads each slot as one run.  One
// pass over the coarse cells fills them; the slot width grows when a row needs more.
// Krylov vectors use the reference's blocked order x(1:n), x(n+1:2n), x(2n+1:3n).
// In FP32-mixed mode the matrix and the Krylov basis are float, every inner product
// and norm accumulates in double, and the small least-squares problem is double.
// Reductions use a fixed grid and a fixed summation tree, and the last work-group to
// finish does the final sum, so results do not depend on timing.
#include "ddns_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ddns {

constexpr int HT = 256;
constexpr int RED_BLOCKS = 256;    // reduction grids: at most one work-group per CU
constexpr int ELL_STEP = 8;        // the ELL width grows in multiples of this


// deterministic wavefront sum: lane 0 holds the butterfly result
__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// ------------------------------------------------------------------ neighbour lists
// Row t (sorted position): every droplet p != t in the 27 cells around t with
// |x_t - x_p| < trunc * r_p, in cell order, as partlib.f90:3280-3330 visits them.
constexpr int NT_NEI = HT;

template <typename T>
__global__ void __launch_bounds__(NT_NEI) k_nei(const double4* __restrict__ pk, const int* __restrict__ cell_start,
                                                int n, int G, double hc, double pi2, double trunc, int width,
                                                int* __restrict__ cnt, int* __restrict__ idx, T* __restrict__ mm,
                                                int* flags)
{
    const int t = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int k = 0;
    if (t < n) {
        const double4 me = pk[t];
        const double xi = me.x, yi = me.y, zi = me.z, ri = me.w;
        int cc[3];
        const double xs[3] = {xi, yi, zi};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            int q = (int)(xs[c] / hc);
            cc[c] = q < 0 ? 0 : (q >= G ? G - 1 : q);
        }
        const size_t ld = (size_t)n;
        bool overlap = false;
        for_each_candidate(cell_start, G, cc, pk, (const int*)nullptr, [&](int p, int kc, const double4& o, int) {
            if (p == t) return;
            int wx, wy, wz;
            neighbour_shift(kc, G, cc, &wx, &wy, &wz);
            const double d0 = xi - (o.x + (double)wx * pi2);
            const double d1 = yi - (o.y + (double)wy * pi2);
            const double d2 = zi - (o.z + (double)wz * pi2);
            const double dnij = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            const double rj = o.w;
            if (dnij < ri + rj) overlap = true;                             // 3294-3299
            if (!(dnij < trunc * rj)) return;                               // 3302-3305
            if (k < width) {
                double fact1 = rj / dnij;                                   // 3312-3324
                const double fact2 = fact1 * fact1 * fact1;
                fact1 = 0.75 * fact1;
                const double gamma = (fact1 - 0.75 * fact2) / (dnij * dnij);
                const double beta = fact1 + 0.25 * fact2;
                idx[(size_t)k * ld + t] = p;
                T* m = mm + (size_t)(6 * k) * ld + t;
                m[0] = (T)(gamma * (d0 * d0) + beta);
                m[ld] = (T)(gamma * d0 * d1);
                m[2 * ld] = (T)(gamma * d0 * d2);
                m[3 * ld] = (T)(gamma * (d1 * d1) + beta);
                m[4 * ld] = (T)(gamma * d1 * d2);
                m[5 * ld] = (T)(gamma * (d2 * d2) + beta);
            }
            ++k;
        });
        if (overlap) atomicOr(flags, 1);
        if (k > DDNS_MAXNEI) atomicOr(flags, 2);
        // a row with more neighbours than the lists hold keeps what fits: the products that
        // may run before the host has read the width flag stay inside the lists, and the
        // host fills them again, wider, before it uses any result
        cnt[t] = k < width ? k : width;
    }
    // the widest row of the wavefront, for the width check
    int km = k;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) km = max(km, __shfl_xor(km, o, 64));
    if ((threadIdx.x & 63) == 0 && km > width) atomicMax(flags + 4, km);
}

// y = x + M x (AFUN_PRECOMP).  MODE 0: x = x0 in double, y0d = A x0 and b = M w.
// MODE 1: x is a basis vector.  MODE 2: x = v / h for the previous work vector v; the
// row's own entries of x are also stored as the next basis vector (q = v / h,
// partlib.f90:2897).  In modes 1 and 2 the Arnoldi step runs on M, y = M x: A = I + M
// spans the same Krylov spaces and its Hessenberg matrix is M's plus the identity
// (added to H(it, it) with the first projection), so the projections no longer cancel
// the large identity part of A x against q_it.
template <typename T, int MODE>
__device__ __forceinline__ void k_spmv_body(const int* __restrict__ cnt, const int* __restrict__ idx,
                                             const T* __restrict__ mm, int n, const T* __restrict__ x,
                                             const double* __restrict__ hx, T* __restrict__ qout, T* __restrict__ y,
                                             const double* __restrict__ x0d, double* __restrict__ y0d,
                                             const double* __restrict__ w, double* __restrict__ b,
                                             const double* __restrict__ stop, int bx, int nbx)
{
    if (stop && *stop != 0.0) return;   // converged in an earlier iteration of this batch
    const int t = xcd_chunk(bx, nbx) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t N = (size_t)n;
    const double h = MODE == 2 ? *hx : 1.0;
    auto xv = [&](size_t q) -> double {
        if (MODE == 0) return x0d[q];
        if (MODE == 1) return (double)x[q];
        return (double)(T)((double)x[q] / h);
    };
    const int kk = cnt[t];
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    for (int s = 0; s < kk; ++s) {
        const size_t q = (size_t)s * N + t;
        const int j = idx[q];
        const T* m = mm + (size_t)(6 * s) * N + t;
        const double m0 = m[0], m1 = m[N], m2 = m[2 * N], m3 = m[3 * N], m4 = m[4 * N], m5 = m[5 * N];
        const double v0 = xv(j), v1 = xv(N + j), v2 = xv(2 * N + j);
        r0 = r0 + m0 * v0 + m1 * v1 + m2 * v2;
        r1 = r1 + m1 * v0 + m3 * v1 + m4 * v2;
        r2 = r2 + m2 * v0 + m4 * v1 + m5 * v2;
        if (MODE == 0) {
            const double q0 = w[j], q1 = w[N + j], q2 = w[2 * N + j];
            b0 = b0 + m0 * q0 + m1 * q1 + m2 * q2;
            b1 = b1 + m1 * q0 + m3 * q1 + m4 * q2;
            b2 = b2 + m2 * q0 + m4 * q1 + m5 * q2;
        }
    }
    const double s0 = xv(t), s1 = xv(N + t), s2 = xv(2 * N + t);
    if (MODE == 0) {
        y0d[t] = r0 + s0; y0d[N + t] = r1 + s1; y0d[2 * N + t] = r2 + s2;
        b[t] = b0; b[N + t] = b1; b[2 * N + t] = b2;
    } else {
        y[t] = (T)r0; y[N + t] = (T)r1; y[2 * N + t] = (T)r2;   // M x: the identity is added to H
        if (MODE == 2) { qout[t] = (T)s0; qout[N + t] = (T)s1; qout[2 * N + t] = (T)s2; }
    }
}
template <typename T, int MODE>
__global__ void __launch_bounds__(HT) k_spmv(const int* __restrict__ cnt, const int* __restrict__ idx,
                                             const T* __restrict__ mm, int n, const T* __restrict__ x,
                                             const double* __restrict__ hx, T* __restrict__ qout, T* __restrict__ y,
                                             const double* __restrict__ x0d, double* __restrict__ y0d,
                                             const double* __restrict__ w, double* __restrict__ b,
                                             const double* __restrict__ stop)
{
    k_spmv_body<T, MODE>(cnt, idx, mm, n, x, hx, qout, y, x0d, y0d, w, b, stop, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------ vector kernels
// r0 = b - A x0 into the first Krylov column, with the work-group partial sums of b.b
// and r0.r0 (part[blk], part[RED_BLOCKS + blk]) on a fixed grid
template <typename T>
__global__ void __launch_bounds__(HT) k_residual0(const double* __restrict__ b, const double* __restrict__ ax0,
                                                  T* __restrict__ q, long long n, double* __restrict__ part)
{
    __shared__ double sb[HT], sq[HT];
    double ab = 0.0, aq = 0.0;
    for (long long i = (long long)blockIdx.x * HT + threadIdx.x; i < n; i += (long long)gridDim.x * HT) {
        const double bi = b[i];
        const T r = (T)(bi - ax0[i]);
        q[i] = r;
        ab += bi * bi;
        aq += (double)r * (double)r;
    }
    sb[threadIdx.x] = ab;
    sq[threadIdx.x] = aq;
    __syncthreads();
    for (int o = HT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sb[threadIdx.x] += sb[threadIdx.x + o];
            sq[threadIdx.x] += sq[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sb[0];
        part[RED_BLOCKS + blockIdx.x] = sq[0];
    }
}

// The small problem cleared (ntot values), then norm_b = |b| and norm_rhs = |r0| from the
// np partials of k_residual0 (fixed order), the tolerance, and the stop flag when the
// initial residual is below 1e-5 (2866-2868)
__global__ void __launch_bounds__(HT) k_gm_start(double* sm, int ntot, double tol, const double* __restrict__ part,
                                                 int np)
{
    __shared__ double sb[HT], sq[HT];
    for (int i = threadIdx.x; i < ntot; i += HT) sm[i] = 0.0;
    double ab = 0.0, aq = 0.0;
    for (int i = threadIdx.x; i < np; i += HT) {
        ab += part[i];
        aq += part[RED_BLOCKS + i];
    }
    sb[threadIdx.x] = ab;
    sq[threadIdx.x] = aq;
    __syncthreads();
    for (int o = HT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            sb[threadIdx.x] += sb[threadIdx.x + o];
            sq[threadIdx.x] += sq[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {                // after the barriers of the tree: the clearing is done
        sm[0] = sqrt(sb[0]);
        sm[1] = sqrt(sq[0]);
        sm[3] = tol;
        if (sm[1] < 1e-5) sm[4] = -1.0;
    }
}

template <typename T>
__global__ void k_scale(T* q, const double* den, long long n)
{
    const long long i = (long long)blockIdx.x * HT + threadIdx.x;
    if (i < n) q[i] = (T)((double)q[i] / *den);
}

// ---- Arnoldi step: classical Gram-Schmidt applied twice, the second pass's update
// only where the first pass lost more than 1 - 1/sqrt(2) of the vector's norm (the
// DGKS criterion: otherwise its corrections are below rounding).  Each projection kernel
// runs on a fixed grid of P <= RED_BLOCKS work-groups and writes one partial sum per
// work-group and column (a fixed tree inside the work-group); the kernel that needs
// the sums adds the P partials itself, in a fixed order, at its start.  Every result
// is independent of scheduling.
constexpr int RT = 1024;          // threads per work-group of the projection kernels
constexpr int RW = RT / 64;       // wavefronts per work-group
constexpr int MD_J = 8;           // Krylov columns per pass of k_mdot

// Work-group sums of acc[0..J): red[jj] valid for all threads on return.
template <int J>
__device__ __forceinline__ void block_sums(double (&acc)[J], double (*sh)[J], double* red)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
        const double v = wave_sum(acc[jj]);
        if (lane == 0) sh[wv][jj] = v;
    }
    __syncthreads();
    if (threadIdx.x < J) {
        double v = sh[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < RW; ++w) v += sh[w][threadIdx.x];
        red[threadIdx.x] = v;
    }
    __syncthreads();
}

// hs[j] = sum over b < P of part[j * P + b] for j < k (lane-strided in order, then a
// fixed wavefront tree), the same in every work-group.  Work-group 0 also sets
// (add = 0, with A's identity added to H(k, k)) or increments (add = 1) the Hessenberg
// column hcol.
__device__ void sum_partials(const double* __restrict__ part, int P, int k, double* hs, double* hcol, int add, int bx)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int j = wv; j < k; j += RW) {
        double acc = 0.0;
        for (int b = lane; b < P; b += 64) acc += part[(size_t)j * P + b];
        acc = wave_sum(acc);
        if (lane == 0) {
            hs[j] = acc;
            if (bx == 0) hcol[j] = add ? hcol[j] + acc : acc + (j == k - 1 ? 1.0 : 0.0);
        }
    }
    __syncthreads();
}

// Krylov columns start every ld elements (ld a multiple of 64), so each lane moves
// 16 bytes per access: 4 floats or 2 doubles.  Element e of a vector sits at
// Q + j * ld + e for e < n3; the last, partial vector is done element by element.
template <typename T> struct kvec;
template <> struct kvec<float> { using V = float4; static constexpr int W = 4; };
template <> struct kvec<double> { using V = double2; static constexpr int W = 2; };

template <typename T>
__device__ __forceinline__ void vld(const T* p, T (&x)[kvec<T>::W])
{
    using V = typename kvec<T>::V;
    *(V*)x = *(const V*)p;
}
template <typename T>
__device__ __forceinline__ void vst(T* p, const T (&x)[kvec<T>::W])
{
    using V = typename kvec<T>::V;
    *(V*)p = *(const V*)x;
}

// partials of Q_j . v for j < k, and of v . v as column k
template <typename T>
__device__ __forceinline__ void k_mdot_body(const T* __restrict__ Q, const T* __restrict__ v, int k, long long n3,
                                             long long ld, double* __restrict__ part, const double* __restrict__ stop, int bx, int nbx)
{
    if (*stop != 0.0) return;
    constexpr int W = kvec<T>::W;
    __shared__ double sh[RW][MD_J + 1];
    __shared__ double red[MD_J + 1];
    const int tid = threadIdx.x;
    const long long nv = (n3 + W - 1) / W;
    for (int j0 = 0; j0 < k; j0 += MD_J) {
        double acc[MD_J + 1];
#pragma unroll
        for (int jj = 0; jj <= MD_J; ++jj) acc[jj] = 0.0;
        for (long long iv = (long long)bx * RT + tid; iv < nv; iv += (long long)nbx * RT) {
            const long long i = iv * W;
            if (i + W <= n3) {
                T x[W];
                vld(v + i, x);
                if (j0 == 0) {
#pragma unroll
                    for (int e = 0; e < W; ++e) acc[MD_J] += (double)x[e] * (double)x[e];
                }
#pragma unroll
                for (int jj = 0; jj < MD_J; ++jj)
                    if (j0 + jj < k) {
                        T q[W];
                        vld(Q + (size_t)(j0 + jj) * ld + i, q);
#pragma unroll
                        for (int e = 0; e < W; ++e) acc[jj] += (double)q[e] * (double)x[e];
                    }
            } else {
                for (long long e = i; e < n3; ++e) {
                    if (j0 == 0) acc[MD_J] += (double)v[e] * (double)v[e];
#pragma unroll
                    for (int jj = 0; jj < MD_J; ++jj)
                        if (j0 + jj < k) acc[jj] += (double)Q[(size_t)(j0 + jj) * ld + e] * (double)v[e];
                }
            }
        }
        block_sums<MD_J + 1>(acc, sh, red);
        if (tid < MD_J && j0 + tid < k) part[(size_t)(j0 + tid) * nbx + bx] = red[tid];
        if (j0 == 0 && tid == 0) part[(size_t)k * nbx + bx] = red[MD_J];
    }
}
template <typename T>
__global__ void __launch_bounds__(RT) k_mdot(const T* __restrict__ Q, const T* __restrict__ v, int k, long long n3,
                                             long long ld, double* __restrict__ part, const double* __restrict__ stop)
{
    k_mdot_body<T>(Q, v, k, n3, ld, part, stop, blockIdx.x, gridDim.x);
}

// First pass's update fused with the second pass's projections, for k <= KC columns:
// h = sum of part_in (first pass), v -= sum_j h_j Q_j, partials of Q_j . v into part_out
// and of v . v (the updated v) as its column k
template <typename T, int KC>
__device__ __forceinline__ void k_maxpy_mdot_body(T* __restrict__ v, const T* __restrict__ Q, int k, long long n3,
                                                   long long ld, const double* __restrict__ part_in,
                                                   double* __restrict__ hcol, double* __restrict__ part_out,
                                                   const double* __restrict__ stop, int bx, int nbx)
{
    if (*stop != 0.0) return;
    constexpr int W = kvec<T>::W;
    __shared__ double hs[KC];
    __shared__ double sh[RW][KC + 1];
    __shared__ double red[KC + 1];
    const int tid = threadIdx.x;
    sum_partials(part_in, nbx, k, hs, hcol, 0, bx);
    double acc[KC + 1];
#pragma unroll
    for (int jj = 0; jj <= KC; ++jj) acc[jj] = 0.0;
    const long long nv = (n3 + W - 1) / W;
    for (long long iv = (long long)bx * RT + tid; iv < nv; iv += (long long)nbx * RT) {
        const long long i = iv * W;
        if (i + W <= n3) {
            T q[KC][W];
            T xv[W];
            vld(v + i, xv);
            double x[W];
#pragma unroll
            for (int e = 0; e < W; ++e) x[e] = (double)xv[e];
#pragma unroll
            for (int jj = 0; jj < KC; ++jj)
                if (jj < k) {
                    vld(Q + (size_t)jj * ld + i, q[jj]);
#pragma unroll
                    for (int e = 0; e < W; ++e) x[e] -= hs[jj] * (double)q[jj][e];
                }
#pragma unroll
            for (int e = 0; e < W; ++e) {
                xv[e] = (T)x[e];
                acc[KC] += (double)xv[e] * (double)xv[e];
            }
            vst(v + i, xv);
#pragma unroll
            for (int jj = 0; jj < KC; ++jj)
                if (jj < k) {
#pragma unroll
                    for (int e = 0; e < W; ++e) acc[jj] += (double)q[jj][e] * (double)xv[e];
                }
        } else {
            for (long long e = i; e < n3; ++e) {
                double x = (double)v[e];
                for (int jj = 0; jj < k; ++jj) x -= hs[jj] * (double)Q[(size_t)jj * ld + e];
                const T xt = (T)x;
                v[e] = xt;
                acc[KC] += (double)xt * (double)xt;
#pragma unroll
                for (int jj = 0; jj < KC; ++jj)
                    if (jj < k) acc[jj] += (double)Q[(size_t)jj * ld + e] * (double)xt;
            }
        }
    }
    block_sums<KC + 1>(acc, sh, red);
    if (tid < k) part_out[(size_t)tid * nbx + bx] = red[tid];
    if (tid == 0) part_out[(size_t)k * nbx + bx] = red[KC];
}
template <typename T, int KC>
__global__ void __launch_bounds__(RT) k_maxpy_mdot(T* __restrict__ v, const T* __restrict__ Q, int k, long long n3,
                                                   long long ld, const double* __restrict__ part_in,
                                                   double* __restrict__ hcol, double* __restrict__ part_out,
                                                   const double* __restrict__ stop)
{
    k_maxpy_mdot_body<T, KC>(v, Q, k, n3, ld, part_in, hcol, part_out, stop, blockIdx.x, gridDim.x);
}

// sum over b < P of part[b] (lane-strided in order, then the wavefront tree), the same
// in every work-group; all threads get it
__device__ __forceinline__ double sum_one(const double* __restrict__ part, int P, double* sh1)
{
    if (threadIdx.x < 64) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < P; b += 64) acc += part[b];
        acc = wave_sum(acc);
        if (threadIdx.x == 0) *sh1 = acc;
    }
    __syncthreads();
    return *sh1;
}

// The product of an Arnoldi step fused with the first pass's projections, for k <= KC
// columns: y = x + M x as k_spmv<T, MODE> (MODE 1: x = Q_0; MODE 2: x = v / h, stored as
// Q_(k-1)), and the partials of Q_j . y for j < k and of y . y (column k).  The grid is
// the projection kernels' (P <= RED_BLOCKS work-groups); work-group b takes a contiguous
// run of rows, the runs placed so that neighbouring ones share an XCD.
template <typename T, int MODE, int KC>
__device__ __forceinline__ void k_spmv_mdot_body(const int* __restrict__ cnt, const int* __restrict__ idx,
                                                  const T* __restrict__ mm, int n, const T* __restrict__ x,
                                                  const double* __restrict__ hx, T* __restrict__ qout,
                                                  T* __restrict__ y, const T* __restrict__ Q, long long ld, int k,
                                                  double* __restrict__ part, const double* __restrict__ stop, int bx, int nbx)
{
    if (*stop != 0.0) return;   // converged in an earlier iteration of this batch
    __shared__ double sh[RW][KC + 1];
    __shared__ double red[KC + 1];
    const size_t N = (size_t)n;
    const double h = MODE == 2 ? *hx : 1.0;
    auto xv = [&](size_t q) -> double {
        if (MODE == 1) return (double)x[q];
        return (double)(T)((double)x[q] / h);
    };
    double acc[KC + 1];
#pragma unroll
    for (int jj = 0; jj <= KC; ++jj) acc[jj] = 0.0;
    const int lb = xcd_chunk(bx, nbx);
    const int per = (n + (int)nbx - 1) / (int)nbx;
    const int t1 = min(n, (lb + 1) * per);
    for (int t = lb * per + threadIdx.x; t < t1; t += RT) {
        const int kk = cnt[t];
        double r0 = 0.0, r1 = 0.0, r2 = 0.0;
        for (int sl = 0; sl < kk; ++sl) {
            const size_t q = (size_t)sl * N + t;
            const int j = idx[q];
            const T* m = mm + (size_t)(6 * sl) * N + t;
            const double m0 = m[0], m1 = m[N], m2 = m[2 * N], m3 = m[3 * N], m4 = m[4 * N], m5 = m[5 * N];
            const double v0 = xv(j), v1 = xv(N + j), v2 = xv(2 * N + j);
            r0 = r0 + m0 * v0 + m1 * v1 + m2 * v2;
            r1 = r1 + m1 * v0 + m3 * v1 + m4 * v2;
            r2 = r2 + m2 * v0 + m4 * v1 + m5 * v2;
        }
        const double s0 = xv(t), s1 = xv(N + t), s2 = xv(2 * N + t);
        const T y0 = (T)r0, y1 = (T)r1, y2 = (T)r2;   // M x: the identity is added to H
        y[t] = y0; y[N + t] = y1; y[2 * N + t] = y2;
        if (MODE == 2) { qout[t] = (T)s0; qout[N + t] = (T)s1; qout[2 * N + t] = (T)s2; }
        const double d0 = (double)y0, d1 = (double)y1, d2 = (double)y2;
        acc[KC] += d0 * d0 + d1 * d1 + d2 * d2;
#pragma unroll
        for (int jj = 0; jj < KC; ++jj)
            if (jj < k) {
                double q0, q1, q2;
                if (jj == k - 1) {      // this row of x itself (Q_0 in MODE 1, the new column in MODE 2)
                    q0 = (double)(T)s0; q1 = (double)(T)s1; q2 = (double)(T)s2;
                } else {
                    const T* qc = Q + (size_t)jj * ld;
                    q0 = (double)qc[t]; q1 = (double)qc[N + t]; q2 = (double)qc[2 * N + t];
                }
                acc[jj] += q0 * d0 + q1 * d1 + q2 * d2;
            }
    }
    block_sums<KC + 1>(acc, sh, red);
    if ((int)threadIdx.x < k) part[(size_t)threadIdx.x * nbx + bx] = red[threadIdx.x];
    if (threadIdx.x == 0) part[(size_t)k * nbx + bx] = red[KC];
}
template <typename T, int MODE, int KC>
__global__ void __launch_bounds__(RT) k_spmv_mdot(const int* __restrict__ cnt, const int* __restrict__ idx,
                                                  const T* __restrict__ mm, int n, const T* __restrict__ x,
                                                  const double* __restrict__ hx, T* __restrict__ qout,
                                                  T* __restrict__ y, const T* __restrict__ Q, long long ld, int k,
                                                  double* __restrict__ part, const double* __restrict__ stop)
{
    k_spmv_mdot_body<T, MODE, KC>(cnt, idx, mm, n, x, hx, qout, y, Q, ld, k, part, stop, blockIdx.x, gridDim.x);
}

// h = sum of part_in, v -= sum_j h[j] Q_j; with normpart, also the partials of |v|^2.
// With gate (the second pass's update): gate[0 .. P) and gate[P .. 2P) hold the partials
// of |w|^2 before the first pass and |v|^2 after it; when the first pass kept at least
// 1/sqrt(2) of the norm (|v|^2 >= |w|^2 / 2, the DGKS criterion) the second pass's
// corrections are below rounding and its update is skipped: v, hcol stay, and normpart
// takes the partials of |v|^2 already summed.
template <typename T>
__device__ __forceinline__ void k_maxpy_body(T* __restrict__ v, const T* __restrict__ Q, int k, long long n3,
                                              long long ld, const double* __restrict__ part_in,
                                              double* __restrict__ hcol, int add, double* __restrict__ normpart,
                                              const double* __restrict__ stop, const double* __restrict__ gate_w,
                                              const double* __restrict__ gate_v, int bx, int nbx)
{
    if (*stop != 0.0) return;
    constexpr int W = kvec<T>::W;
    __shared__ double hs[64];
    __shared__ double sh[RW][1];
    __shared__ double red[1];
    __shared__ double g1[2];
    const int tid = threadIdx.x;
    if (gate_w) {
        const double nw = sum_one(gate_w, nbx, &g1[0]);
        const double nv2 = sum_one(gate_v, nbx, &g1[1]);
        if (nv2 >= 0.5 * nw) {
            if (normpart && tid == 0) normpart[bx] = gate_v[bx];
            return;
        }
    }
    sum_partials(part_in, nbx, k, hs, hcol, add, bx);
    double nacc[1] = {0.0};
    const long long nv = (n3 + W - 1) / W;
    for (long long iv = (long long)bx * RT + tid; iv < nv; iv += (long long)nbx * RT) {
        const long long i = iv * W;
        if (i + W <= n3) {
            T xv[W];
            vld(v + i, xv);
            double x[W];
#pragma unroll
            for (int e = 0; e < W; ++e) x[e] = (double)xv[e];
#pragma unroll 4
            for (int j = 0; j < k; ++j) {
                T q[W];
                vld(Q + (size_t)j * ld + i, q);
                const double hj = hs[j];
#pragma unroll
                for (int e = 0; e < W; ++e) x[e] -= hj * (double)q[e];
            }
#pragma unroll
            for (int e = 0; e < W; ++e) {
                xv[e] = (T)x[e];
                nacc[0] += (double)xv[e] * (double)xv[e];
            }
            vst(v + i, xv);
        } else {
            for (long long e = i; e < n3; ++e) {
                double x = (double)v[e];
                for (int j = 0; j < k; ++j) x -= hs[j] * (double)Q[(size_t)j * ld + e];
                const T xt = (T)x;
                v[e] = xt;
                nacc[0] += (double)xt * (double)xt;
            }
        }
    }
    if (!normpart) return;
    block_sums<1>(nacc, sh, red);
    if (tid == 0) normpart[bx] = red[0];
}
template <typename T>
__global__ void __launch_bounds__(RT) k_maxpy(T* __restrict__ v, const T* __restrict__ Q, int k, long long n3,
                                              long long ld, const double* __restrict__ part_in,
                                              double* __restrict__ hcol, int add, double* __restrict__ normpart,
                                              const double* __restrict__ stop, const double* __restrict__ gate_w,
                                              const double* __restrict__ gate_v)
{
    k_maxpy_body<T>(v, Q, k, n3, ld, part_in, hcol, add, normpart, stop, gate_w, gate_v, blockIdx.x, gridDim.x);
}

// pertvel = x0 + Q y (2951-2956), scattered back to the reference's droplet order.  The
// number of Krylov columns is the small problem's count of the iterations that ran
// (*itp), read on the device, so the kernel can be queued before the host has seen it.
template <typename T>
__global__ void k_solution(const double* x0, const T* Q, long long ld, const double* y, const double* itp,
                           const int* sorted, double* pert, int n, int cap)
{
    // one droplet per thread: its three components go out as one 24-byte record
    const int t = blockIdx.x * HT + threadIdx.x;
    if (t >= n) return;
    const int iters = (int)*itp;
    double out[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t i = (size_t)c * n + t;
        double acc = 0.0;
        for (int j = 0; j < iters; ++j) acc = acc + (double)Q[(size_t)j * ld + i] * y[j];
        out[c] = x0[i] + acc;
    }
    double* dst = pert + 3 * (size_t)sorted[t];
#pragma unroll
    for (int c = 0; c < 3; ++c) dst[c] = out[c];
}

// Small least-squares update of HDI_GMRES (2904-2971), one wavefront.
// sm layout: [0] norm_b [1] norm_rhs [2] res [3] maxerr [4] done [5] iter
//            H at 16 + (r-1) + M1*(c-1), Rn after it, Qn after it, bn, y, v1, v2, resv
struct GmSmall {
    int maxiter;
    __host__ __device__ int M1() const { return maxiter + 1; }
    __host__ __device__ size_t H0() const { return 16; }
    __host__ __device__ size_t R0() const { return H0() + (size_t)M1() * maxiter; }
    __host__ __device__ size_t Q0() const { return R0() + (size_t)M1() * M1(); }
    __host__ __device__ size_t B0() const { return Q0() + (size_t)M1() * M1(); }
    __host__ __device__ size_t Y0() const { return B0() + M1(); }
    __host__ __device__ size_t V1() const { return Y0() + M1(); }
    __host__ __device__ size_t V2() const { return V1() + M1(); }
    __host__ __device__ size_t RV() const { return V2() + M1(); }
    __host__ __device__ size_t total() const { return RV() + M1(); }
};

constexpr int LT = 64;         // k_ls: one wavefront; maxiter + 1 <= LT
constexpr int LS_MAX = 8192;   // GmSmall::total() for maxiter <= 50, the reference's value

__device__ void ls_solve(double* sm, GmSmall L, int iter);

__device__ __forceinline__ void copy_range(double* dst, const double* src, size_t a, size_t b)
{
    for (size_t i = a + threadIdx.x; i < b; i += LT) dst[i] = src[i];
}

// The five ranges of the small problem that an iteration reads, copied to LDS with
// the loads of up to 8 * LT values issued together: one memory round trip for the
// usual iteration counts instead of one per range and pass.
__device__ __forceinline__ void copy_in(double* sm, const double* __restrict__ gsm, GmSmall L, int iter)
{
    constexpr int KB = 8;
    const size_t M1 = L.M1();
    const size_t n0 = 16, n1 = M1 * iter, n2 = M1 * iter, n3 = M1 * (iter + 1), n4 = L.total() - L.B0();
    const size_t tot = n0 + n1 + n2 + n3 + n4;
    for (size_t b = 0; b < tot; b += (size_t)KB * LT) {
        double r[KB];
        size_t at[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
            size_t j = b + (size_t)k * LT + threadIdx.x;
            at[k] = ~(size_t)0;
            if (j < tot) {
                size_t a;
                if (j < n0) a = j;
                else if ((j -= n0) < n1) a = L.H0() + j;
                else if ((j -= n1) < n2) a = L.R0() + j;
                else if ((j -= n2) < n3) a = L.Q0() + j;
                else a = L.B0() + (j - n3);
                at[k] = a;
                r[k] = gsm[a];
            }
        }
#pragma unroll
        for (int k = 0; k < KB; ++k)
            if (at[k] != ~(size_t)0) sm[at[k]] = r[k];
    }
}

__device__ __forceinline__ void k_ls_body(double* gsm, GmSmall L, int iter, const double* __restrict__ normpart,
                                          int P, int bx, int nbx)
{
    // the part of the small problem this iteration touches is copied to LDS, solved
    // there, and the changed part copied back
    __shared__ double sm[LS_MAX];
    const size_t M1 = L.M1();
    // H(iter+1, iter) = |v| from the partials of the last projection kernel; their
    // loads are issued together with those of the LDS copy
    static_assert(RED_BLOCKS % LT == 0, "partials per lane");
    double np[RED_BLOCKS / LT];
#pragma unroll
    for (int k = 0; k < RED_BLOCKS / LT; ++k) {
        const int b = threadIdx.x + k * LT;
        np[k] = b < P ? normpart[b] : 0.0;
    }
    const double done = gsm[4];
    copy_in(sm, gsm, L, iter);
    if (done != 0.0) return;     // converged in an earlier iteration of this batch
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < RED_BLOCKS / LT; ++k) acc += np[k];
    acc = wave_sum(acc);
    __syncthreads();
    if (threadIdx.x == 0) {
        const size_t e = L.H0() + iter + M1 * (iter - 1);
        sm[e] = gsm[e] = sqrt(acc);
    }
    __syncthreads();
    ls_solve(sm, L, iter);
    __syncthreads();
    copy_range(gsm, sm, 0, 16);
    copy_range(gsm, sm, L.R0(), L.R0() + M1 * iter);
    copy_range(gsm, sm, L.Q0(), L.Q0() + M1 * (iter + 1));
    copy_range(gsm, sm, L.B0(), L.total());
}

__global__ void __launch_bounds__(LT) k_ls(double* gsm, GmSmall L, int iter, const double* __restrict__ normpart,
                                          int P)
{
    k_ls_body(gsm, L, iter, normpart, P, blockIdx.x, gridDim.x);
}

// All LT lanes call it.  Lane r owns row r + 1 of the triangular factor and of the
// residual; the Givens rotation and the back substitution run over the lanes.
__device__ void ls_solve(double* sm, GmSmall L, int iter)
{
    const int lane = threadIdx.x;
    const int M1 = L.M1();
    double* H = sm + L.H0();
    double* Rn = sm + L.R0();
    double* Qn = sm + L.Q0();
    double* bn = sm + L.B0();
    double* y = sm + L.Y0();
    double* rv = sm + L.RV();
#define HH(r, c) H[((r) - 1) + (size_t)M1 * ((c) - 1)]
#define RN(r, c) Rn[((r) - 1) + (size_t)M1 * ((c) - 1)]
#define QN(r, c) Qn[((r) - 1) + (size_t)M1 * ((c) - 1)]
    const double norm_rhs = sm[1];
    if (iter == 1) {
        if (lane == 0) {
            const double nrm = sqrt(HH(1, 1) * HH(1, 1) + HH(2, 1) * HH(2, 1));
            const double q11 = HH(1, 1) / nrm, q21 = HH(2, 1) / nrm;
            QN(1, 1) = q11; QN(2, 1) = q21;
            QN(1, 2) = q21; QN(2, 2) = -q11;
            RN(1, 1) = sqrt(HH(1, 1) * HH(1, 1) + HH(2, 1) * HH(2, 1));
            bn[0] = q11 * norm_rhs;
            y[0] = bn[0] / RN(1, 1);
        }
        __syncthreads();
    } else {
        // new column of R: Qn(1:iter, 1:iter) H(1:iter, iter)
        if (lane < iter) {
            double acc = 0.0;
            for (int c = 1; c <= iter; ++c) acc = acc + QN(lane + 1, c) * HH(c, iter);
            RN(lane + 1, iter) = acc;
        }
        __syncthreads();
        const double aa = RN(iter, iter), bb = HH(iter + 1, iter);
        const double dd = sqrt(aa * aa + bb * bb);
        const double cc = aa / dd, ss = bb / dd;
        const double qi = lane < iter ? QN(iter, lane + 1) : 0.0;
        __syncthreads();
        // rows iter and iter+1 of Qn after the rotation
        if (lane < iter) { QN(iter, lane + 1) = cc * qi; QN(iter + 1, lane + 1) = -ss * qi; }
        if (lane == iter) { QN(iter, iter + 1) = ss; QN(iter + 1, iter + 1) = cc; }
        if (lane == 0) RN(iter, iter) = dd;
        __syncthreads();
        if (lane == 0) {
            bn[iter - 2] = QN(iter - 1, 1) * norm_rhs;
            bn[iter - 1] = QN(iter, 1) * norm_rhs;
        }
        __syncthreads();
        // back substitution by columns: lane r accumulates row r + 1
        double acc = 0.0;
        for (int c = iter; c >= 1; --c) {
            if (lane == c - 1) y[c - 1] = (bn[c - 1] - acc) / RN(c, c);
            __syncthreads();
            const double yc = y[c - 1];
            if (lane < c - 1) acc = acc + yc * RN(lane + 1, c);
        }
        __syncthreads();
    }
    // residual H y - |r0| e1 (2958-2966)
    if (lane <= iter) {
        double acc = 0.0;
        for (int c = 1; c <= iter; ++c) acc = acc + HH(lane + 1, c) * y[c - 1];
        rv[lane] = acc;
    }
    __syncthreads();
    if (lane == 0) {
        rv[0] = rv[0] - norm_rhs;
        double x_c = 0.0;
        for (int r = 0; r <= iter; ++r) x_c = x_c + rv[r] * rv[r];
        const double res = sqrt(x_c) / sm[0];
        sm[2] = res;
        sm[5] = iter;
        double maxerr = sm[3];
        int done = 0;
        if (res < maxerr) done = 1;
        else {
            if (iter == L.maxiter / 2) maxerr = maxerr * 10;
            if (iter >= L.maxiter - 2) maxerr = maxerr * 10;
        }
        sm[3] = maxerr;
        sm[4] = done;
    }
#undef HH
#undef RN
#undef QN
}

// ------------------------------------------------------------------ driver
// NEI_LIST_PRECOMP (partlib.f90:3218-3371) queued without a host wait: the flags it
// raises are read with the solve's first convergence check (lists_outcome).
template <typename T>
static int queue_lists(PartBuffers& pb, const PartConst& pc, hipStream_t s)
{
    const int n = pc.n;
    if (!pb.nei_idx) {
        pb.ell_width = 2 * ELL_STEP;
        DDNS_HIP(hipMalloc(&pb.nei_idx, (size_t)pb.ell_width * pb.cap * sizeof(int)));
        DDNS_HIP(hipMalloc(&pb.nei_m, (size_t)6 * pb.ell_width * pb.cap * sizeof(T)));
    }
    DDNS_HIP(hipMemsetAsync(pb.flags, 0, 8 * sizeof(int), s));
    hipLaunchKernelGGL((k_nei<T>), dim3((n + NT_NEI - 1) / NT_NEI), dim3(NT_NEI), 0, s, (const double4*)pb.pk, pb.cell_start,
                       n, pb.G, pb.hc, pc.pi2, pc.hdi_trunc, pb.ell_width, pb.nei_cnt, pb.nei_idx, (T*)pb.nei_m,
                       pb.flags);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// The flags of queue_lists, once read: an error (overlapping droplets, more than 100
// neighbours), 1 when some row had more neighbours than the lists hold (they have been
// widened; fill them again), else 0.
template <typename T>
static int lists_outcome(PartBuffers& pb, const int* f)
{
    if (f[0] & 1) { set_error("inside NEI_LIST_PRECOMP: overlapping is not allowed"); return DDNS_ERR_OVERLAP; }
    if (f[0] & 2) { set_error("more than 100 HI neighbours for a droplet"); return DDNS_ERR_NEIGHBOUR_OVERFLOW; }
    if (f[4] <= pb.ell_width) return 0;
    if (pb.nei_idx) { (void)hipFree(pb.nei_idx); (void)hipFree(pb.nei_m); pb.nei_idx = nullptr; pb.nei_m = nullptr; }
    pb.ell_width = std::min(DDNS_MAXNEI, (f[4] + ELL_STEP - 1) / ELL_STEP * ELL_STEP);
    DDNS_HIP(hipMalloc(&pb.nei_idx, (size_t)pb.ell_width * pb.cap * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.nei_m, (size_t)6 * pb.ell_width * pb.cap * sizeof(T)));
    return 1;
}

// The solve's start (2820-2868) queued on s without a host wait: the neighbour lists,
// b = M w and A x0, r0 = b - A x0 into the first Krylov column with the norms, the
// small problem cleared, the tolerance and the stop flag, and q_1 = r0 / |r0|.
template <typename T>
static int hdi_begin(PartBuffers& pb, const PartConst& pc, int maxiter, double tol, hipStream_t s)
{
    const int n = pc.n;
    const long long n3 = 3LL * n;
    if (maxiter + 1 > LT) { set_error("GMRES maxiter too large (at most 50)"); return DDNS_ERR_ARG; }
    GmSmall L{maxiter};
    if (L.total() > (size_t)LS_MAX) { set_error("GMRES maxiter too large (at most 50)"); return DDNS_ERR_ARG; }
    // Krylov storage
    if (!pb.gm_small) {
        DDNS_HIP(hipMalloc(&pb.gm_small, 16384 * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_red, (size_t)(128 + 2) * RED_BLOCKS * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_b, 3 * (size_t)pb.cap * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_v, 3 * (size_t)pb.cap * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_Q, (size_t)(maxiter + 2) * ((3LL * pb.cap + 63) / 64 * 64) * sizeof(T)));
    }
    double* sm = pb.gm_small;
    double* ax0 = (double*)pb.gm_v;
    T* Q = (T*)pb.gm_Q;
    const unsigned nb3 = (unsigned)((n3 + HT - 1) / HT);
    const unsigned nbr = (unsigned)((n + HT - 1) / HT);
    const int sq = queue_lists<T>(pb, pc, s);
    if (sq) return sq;
    // x0 (pertvel) and w = V - U in cell order were written by the interpolation
    hipLaunchKernelGGL((k_spmv<T, 0>), dim3(nbr), dim3(HT), 0, s, (const int*)pb.nei_cnt, (const int*)pb.nei_idx,
                       (const T*)pb.nei_m, n, (const T*)nullptr, (const double*)nullptr, (T*)nullptr, (T*)nullptr,
                       (const double*)pb.gm_x0, ax0, (const double*)pb.gm_w, pb.gm_b, (const double*)nullptr);
    const int nr0 = (int)std::min<long long>(RED_BLOCKS, (n3 + HT - 1) / HT);
    hipLaunchKernelGGL((k_residual0<T>), dim3(nr0), dim3(HT), 0, s, pb.gm_b, ax0, Q, n3, pb.gm_red);
    // the tolerance, and the stop flag when the initial residual is below 1e-5
    // (2866-2868): then no iteration runs and pertvel keeps its warm-start value
    hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(HT), 0, s, sm, (int)L.total(), tol, (const double*)pb.gm_red, nr0);
    hipLaunchKernelGGL((k_scale<T>), dim3(nb3), dim3(HT), 0, s, Q, sm + 1, n3);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// work-groups of the projection kernels of a system of n droplets
static int hdi_nred(int n)
{
    const long long n3 = 3LL * n;
    return (int)std::min<long long>(RED_BLOCKS, (n3 + 4LL * RT - 1) / (4LL * RT));
}

template <typename T>
static int hdi_impl(PartBuffers& pb, const PartConst& pc, int maxiter, double tol, int* iters_out, double* res_out,
                    hipStream_t s)
{
    const int n = pc.n;
    const long long n3 = 3LL * n;
    *iters_out = 0;
    *res_out = 0.0;
    if (n == 0) return 0;
    GmSmall L{maxiter};
    const long long ld = (n3 + 63) / 64 * 64;     // column stride of the Krylov vectors
    const unsigned nbr = (unsigned)((n + HT - 1) / HT);
    const unsigned nred = (unsigned)hdi_nred(n);   // reduction grid
    // The neighbour lists are filled and the solve started without a host wait; the
    // first convergence check also reads the lists' flags, and when a row had more
    // neighbours than the lists hold they are widened and the solve starts over (its
    // inputs x0 and w are unchanged, and pertvel is written only at the end).
    int iter = 0;
    double st6[6] = {0, 0, 0, 0, 0, 0};
    double* sm = nullptr;
    T* Q = nullptr;
    for (int lists = 1; lists > 0;) {
        // the Krylov storage is allocated by the first hdi_begin of the handle: every
        // pointer into it is taken after the call
        const int sq = hdi_begin<T>(pb, pc, maxiter, tol, s);
        if (sq) return sq;
        lists = 0;
        sm = pb.gm_small;
        Q = (T*)pb.gm_Q;
        double* part1 = pb.gm_red;                          // first-pass projections
        double* part2 = part1 + (size_t)64 * RED_BLOCKS;    // second-pass projections
        double* normpart = part2 + (size_t)64 * RED_BLOCKS;
        T* work[2] = {Q + (size_t)maxiter * ld, Q + (size_t)(maxiter + 1) * ld};
        // Iterations are launched in batches; every kernel of an iteration returns at
        // once when the least-squares update of an earlier one has set the converged
        // flag sm[4], so the host checks it once per batch.  The first batch runs as many
        // iterations as the last solve needed (consecutive steps converge alike), the
        // later ones three at a time.
        constexpr int BATCH = 3;
        const double* stop = sm + 4;
        double* Hd = sm + L.H0();
        const int M1 = L.M1();
        iter = 0;
        for (double& v : st6) v = 0.0;
        bool first = true;
        while (iter < maxiter && st6[4] == 0.0) {
            const int last = std::min(maxiter, iter + (iter == 0 ? std::max(BATCH, pb.gm_last_iters) : BATCH));
            for (int it = iter + 1; it <= last; ++it) {
                T* v = work[it & 1];
                // v = A q_it, q_it = v_(it-1) / H(it, it-1) formed inside the product; up to
                // 12 columns the first pass's projections are summed in the same kernel
                const T* xin = it == 1 ? (const T*)Q : (const T*)work[(it - 1) & 1];
                const double* hx = it == 1 ? (const double*)nullptr : (const double*)(Hd + (size_t)M1 * (it - 2) + (it - 1));
                T* qo = it == 1 ? (T*)nullptr : Q + (size_t)(it - 1) * ld;
                const int* nc = (const int*)pb.nei_cnt;
                const int* ni = (const int*)pb.nei_idx;
                const T* nm = (const T*)pb.nei_m;
                if (it == 1)
                    hipLaunchKernelGGL((k_spmv_mdot<T, 1, 8>), dim3(nred), dim3(RT), 0, s, nc, ni, nm, n, xin, hx, qo, v,
                                       (const T*)Q, ld, it, part1, stop);
                else if (it <= 8)
                    hipLaunchKernelGGL((k_spmv_mdot<T, 2, 8>), dim3(nred), dim3(RT), 0, s, nc, ni, nm, n, xin, hx, qo, v,
                                       (const T*)Q, ld, it, part1, stop);
                else if (it <= 12)
                    hipLaunchKernelGGL((k_spmv_mdot<T, 2, 12>), dim3(nred), dim3(RT), 0, s, nc, ni, nm, n, xin, hx, qo,
                                       v, (const T*)Q, ld, it, part1, stop);
                else {
                    hipLaunchKernelGGL((k_spmv<T, 2>), dim3(nbr), dim3(HT), 0, s, nc, ni, nm, n, xin, hx, qo, v,
                                       (const double*)nullptr, (double*)nullptr, (const double*)nullptr, (double*)nullptr,
                                       stop);
                    hipLaunchKernelGGL((k_mdot<T>), dim3(nred), dim3(RT), 0, s, (const T*)Q, (const T*)v, it, n3, ld, part1,
                                       stop);
                }
                // orthogonalise v against q_1..q_it: classical Gram-Schmidt, twice (the
                // reference's modified Gram-Schmidt, 2886-2894, in two block passes), the
                // second pass's update only where the DGKS criterion asks for it
                double* hcol = Hd + (size_t)M1 * (it - 1);
                if (it <= 8)
                    hipLaunchKernelGGL((k_maxpy_mdot<T, 8>), dim3(nred), dim3(RT), 0, s, v, (const T*)Q, it, n3, ld,
                                       (const double*)part1, hcol, part2, stop);
                else if (it <= 12)
                    hipLaunchKernelGGL((k_maxpy_mdot<T, 12>), dim3(nred), dim3(RT), 0, s, v, (const T*)Q, it, n3, ld,
                                       (const double*)part1, hcol, part2, stop);
                else {
                    hipLaunchKernelGGL((k_maxpy<T>), dim3(nred), dim3(RT), 0, s, v, (const T*)Q, it, n3, ld,
                                       (const double*)part1, hcol, 0, (double*)nullptr, stop, (const double*)nullptr,
                                       (const double*)nullptr);
                    hipLaunchKernelGGL((k_mdot<T>), dim3(nred), dim3(RT), 0, s, (const T*)Q, (const T*)v, it, n3, ld, part2,
                                       stop);
                }
                // the second pass's update, skipped where the first kept 1/sqrt(2) of |v|
                hipLaunchKernelGGL((k_maxpy<T>), dim3(nred), dim3(RT), 0, s, v, (const T*)Q, it, n3, ld, (const double*)part2,
                                   hcol, 1, normpart, stop, (const double*)(part1 + (size_t)it * nred),
                                   (const double*)(part2 + (size_t)it * nred));
                hipLaunchKernelGGL(k_ls, dim3(1), dim3(LT), 0, s, sm, L, it, (const double*)normpart, (int)nred);
            }
            DDNS_HIP(hipGetLastError());
            iter = last;
            // the first check also reads the neighbour lists' flags (pinned memory: the two
            // copies are queued together and waited for once)
            if (!pb.gm_host) DDNS_HIP(hipHostMalloc((void**)&pb.gm_host, 16 * sizeof(double), hipHostMallocDefault));
            int f[8];
            int* fh = (int*)(pb.gm_host + 8);
            if (first) DDNS_HIP(hipMemcpyAsync(fh, pb.flags, sizeof(f), hipMemcpyDeviceToHost, s));
            DDNS_HIP(hipMemcpyAsync(pb.gm_host, sm, sizeof(st6), hipMemcpyDeviceToHost, s));
            if (!pb.sync_ev) DDNS_HIP(hipEventCreateWithFlags(&pb.sync_ev, hipEventDisableTiming));
            DDNS_HIP(hipEventRecord(pb.sync_ev, s));
            // the solution as far as this batch got (pertvel = x0 where no iteration ran),
            // queued behind the copies: when the batch converged, the usual case, it runs
            // while the host reads the check and queues what follows; a later batch is
            // followed by its own
            hipLaunchKernelGGL((k_solution<T>), dim3(nbr), dim3(HT), 0, s, pb.gm_x0, (const T*)Q, ld, sm + L.Y0(),
                               (const double*)(sm + 5), (const int*)pb.sorted, pb.pert, n, pb.cap);
            DDNS_HIP(spin_wait(pb.sync_ev));
            std::memcpy(st6, pb.gm_host, sizeof(st6));
            if (first) std::memcpy(f, fh, sizeof(f));
            if (first) {
                first = false;
                lists = lists_outcome<T>(pb, f);
                if (lists < 0) return lists;
                if (lists > 0) break;            // fill the widened lists and solve again
            }
        }
    }
    if (st6[4] < 0.0) return 0;     // initial residual below 1e-5: no iteration
    const bool done = st6[4] != 0.0;
    const int used = (int)st6[5];   // the last iteration that ran
    pb.gm_last_iters = used;
    DDNS_HIP(hipGetLastError());
    // the solution was queued behind the last check; the residual and its bound were
    // final then: the solution kernel reads the small problem but does not change it
    *iters_out = used;
    *res_out = st6[2];
    if (!done && st6[2] > st6[3]) return DDNS_WARN_GMRES;                   // 2980-2982
    return 0;
}

int part_hdi(PartBuffers& pb, const PartConst& pc, int maxiter, double tol, int* iters, double* res, hipStream_t s)
{
    if (pc.precision == DDNS_FP64) return hdi_impl<double>(pb, pc, maxiter, tol, iters, res, s);
    return hdi_impl<float>(pb, pc, maxiter, tol, iters, res, s);
}

int part_hdi_begin(PartBuffers& pb, const PartConst& pc, int maxiter, double tol, hipStream_t s)
{
    if (pc.n == 0) return 0;
    if (pc.precision == DDNS_FP64) return hdi_begin<double>(pb, pc, maxiter, tol, s);
    return hdi_begin<float>(pb, pc, maxiter, tol, s);
}

// ------------------------------------------------------------------ batched solve
// The GMRES iterations of several independent systems (one per realisation of a batch)
// run together: each kernel of an iteration is one launch whose work-group row y works
// on system y.  Every system keeps its own Krylov basis, small problem, stop flag and
// reduction partials, and its reductions run on the grid of nred work-groups its own
// solve would use, so each system gets the same bits as when solved alone; a converged
// system's work-groups return at once, as in a solve of its own.
template <typename T>
struct GmDesc {
    const int* cnt;
    const int* idx;
    const T* mm;
    const int* flags;          // the neighbour lists' flags (read at the first check)
    T* Q;                      // maxiter Krylov columns, then the two work vectors
    double* sm;                // the small problem (GmSmall layout)
    double* part1;             // first-pass projection partials
    double* part2;             // second-pass projection partials
    double* normpart;          // |v|^2 partials
    long long n3, ld;
    int n, nred, nbr;          // nred = 0: the system takes no part (any more)
};

template <typename T>
struct GmIter {
    const T* xin;
    const double* hx;
    T* qo;
    T* v;
    double* hcol;
    const double* stop;
};

// the vectors and Hessenberg entries of iteration it, as hdi_impl forms them
template <typename T>
__device__ __forceinline__ GmIter<T> gm_iter(const GmDesc<T>& g, GmSmall L, int it)
{
    GmIter<T> r;
    T* w0 = g.Q + (size_t)L.maxiter * g.ld;
    T* w1 = w0 + g.ld;
    double* Hd = g.sm + L.H0();
    const size_t M1 = L.M1();
    r.v = (it & 1) ? w1 : w0;
    r.xin = it == 1 ? g.Q : (((it - 1) & 1) ? w1 : w0);
    r.hx = it == 1 ? nullptr : Hd + M1 * (it - 2) + (it - 1);
    r.qo = it == 1 ? nullptr : g.Q + (size_t)(it - 1) * g.ld;
    r.hcol = Hd + M1 * (it - 1);
    r.stop = g.sm + 4;
    return r;
}

template <typename T, int MODE, int KC>
__global__ void __launch_bounds__(RT) kb_spmv_mdot(const GmDesc<T>* __restrict__ d, GmSmall L, int it)
{
    const GmDesc<T>& g = d[blockIdx.y];
    if ((int)blockIdx.x >= g.nred) return;
    const GmIter<T> q = gm_iter(g, L, it);
    k_spmv_mdot_body<T, MODE, KC>(g.cnt, g.idx, g.mm, g.n, q.xin, q.hx, q.qo, q.v, g.Q, g.ld, it, g.part1, q.stop,
                                  blockIdx.x, g.nred);
}

template <typename T>
__global__ void __launch_bounds__(HT) kb_spmv2(const GmDesc<T>* __restrict__ d, GmSmall L, int it)
{
    const GmDesc<T>& g = d[blockIdx.y];
    if (g.nred == 0 || (int)blockIdx.x >= g.nbr) return;
    const GmIter<T> q = gm_iter(g, L, it);
    k_spmv_body<T, 2>(g.cnt, g.idx, g.mm, g.n, q.xin, q.hx, q.qo, q.v, nullptr, nullptr, nullptr, nullptr, q.stop,
                      blockIdx.x, g.nbr);
}

template <typename T>
__global__ void __launch_bounds__(RT) kb_mdot(const GmDesc<T>* __restrict__ d, GmSmall L, int it, int second)
{
    const GmDesc<T>& g = d[blockIdx.y];
    if ((int)blockIdx.x >= g.nred) return;
    const GmIter<T> q = gm_iter(g, L, it);
    k_mdot_body<T>(g.Q, q.v, it, g.n3, g.ld, second ? g.part2 : g.part1, q.stop, blockIdx.x, g.nred);
}

template <typename T, int KC>
__global__ void __launch_bounds__(RT) kb_maxpy_mdot(const GmDesc<T>* __restrict__ d, GmSmall L, int it)
{
    const GmDesc<T>& g = d[blockIdx.y];
    if ((int)blockIdx.x >= g.nred) return;
    const GmIter<T> q = gm_iter(g, L, it);
    k_maxpy_mdot_body<T, KC>(q.v, g.Q, it, g.n3, g.ld, g.part1, q.hcol, g.part2, q.stop, blockIdx.x, g.nred);
}

// second = 0: the first pass's update (k > 12); 1: the second pass's DGKS-gated update
template <typename T>
__global__ void __launch_bounds__(RT) kb_maxpy(const GmDesc<T>* __restrict__ d, GmSmall L, int it, int second)
{
    const GmDesc<T>& g = d[blockIdx.y];
    if ((int)blockIdx.x >= g.nred) return;
    const GmIter<T> q = gm_iter(g, L, it);
    if (second)
        k_maxpy_body<T>(q.v, g.Q, it, g.n3, g.ld, g.part2, q.hcol, 1, g.normpart, q.stop,
                        g.part1 + (size_t)it * g.nred, g.part2 + (size_t)it * g.nred, blockIdx.x, g.nred);
    else
        k_maxpy_body<T>(q.v, g.Q, it, g.n3, g.ld, g.part1, q.hcol, 0, nullptr, q.stop, nullptr, nullptr, blockIdx.x,
                        g.nred);
}

template <typename T>
__global__ void __launch_bounds__(LT) kb_ls(const GmDesc<T>* __restrict__ d, GmSmall L, int it)
{
    const GmDesc<T>& g = d[blockIdx.x];
    if (g.nred == 0) return;
    k_ls_body(g.sm, L, it, g.normpart, g.nred, 0, 1);
}

// each system's small-problem head sm[0..5] (and its lists' flags) into 16 doubles
template <typename T>
__global__ void kb_heads(const GmDesc<T>* __restrict__ d, double* out, int with_flags)
{
    const GmDesc<T>& g = d[blockIdx.x];
    double* o = out + 16 * (size_t)blockIdx.x;
    const int i = threadIdx.x;
    if (!g.sm) {                       // no droplets, no solve: reads as "no iteration"
        if (i < 6) o[i] = i == 4 ? -1.0 : 0.0;
        if (i < 8) ((int*)(o + 8))[i] = 0;
        return;
    }
    if (i < 6) o[i] = g.sm[i];
    if (with_flags && i < 8) ((int*)(o + 8))[i] = g.flags[i];
}

void part_hdi_batch_free(HdiBatch& hb)
{
    if (hb.table) (void)hipFree(hb.table);
    if (hb.heads) (void)hipFree(hb.heads);
    if (hb.table_host) (void)hipHostFree(hb.table_host);
    if (hb.heads_host) (void)hipHostFree(hb.heads_host);
    hb = HdiBatch{};
}

template <typename T>
static int hdi_batch_impl(HdiBatch& hb, PartBuffers* const* pbs, const PartConst* const* pcs, int R, int maxiter,
                          int* iters, double* res, int* status, int* redo, hipStream_t s)
{
    if (R <= 0) return 0;
    GmSmall L{maxiter};
    if (hb.cap < R) {
        part_hdi_batch_free(hb);
        DDNS_HIP(hipMalloc(&hb.table, (size_t)R * sizeof(GmDesc<T>)));
        DDNS_HIP(hipMalloc(&hb.heads, (size_t)R * 16 * sizeof(double)));
        DDNS_HIP(hipHostMalloc(&hb.table_host, (size_t)R * sizeof(GmDesc<double>), hipHostMallocDefault));
        DDNS_HIP(hipHostMalloc((void**)&hb.heads_host, (size_t)R * 16 * sizeof(double), hipHostMallocDefault));
        hb.cap = R;
    }
    GmDesc<T>* th = (GmDesc<T>*)hb.table_host;
    int nredmax = 0, nbrmax = 0, first = 3;
    for (int r = 0; r < R; ++r) {
        PartBuffers& pb = *pbs[r];
        const int n = pcs[r]->n;
        iters[r] = 0; res[r] = 0.0; status[r] = 0; redo[r] = 0;
        GmDesc<T> g{};
        if (n == 0 || !pb.gm_small) {  // nothing to solve (part_hdi_begin queued nothing)
            th[r] = g;
            continue;
        }
        g.cnt = pb.nei_cnt; g.idx = pb.nei_idx; g.mm = (const T*)pb.nei_m; g.flags = pb.flags;
        g.Q = (T*)pb.gm_Q; g.sm = pb.gm_small;
        g.part1 = pb.gm_red; g.part2 = g.part1 + (size_t)64 * RED_BLOCKS; g.normpart = g.part2 + (size_t)64 * RED_BLOCKS;
        g.n3 = 3LL * n; g.ld = (g.n3 + 63) / 64 * 64; g.n = n;
        g.nred = hdi_nred(n); g.nbr = (n + HT - 1) / HT;
        th[r] = g;
        nredmax = std::max(nredmax, g.nred);
        nbrmax = std::max(nbrmax, g.nbr);
        first = std::max(first, pb.gm_last_iters);
    }
    const GmDesc<T>* td = (const GmDesc<T>*)hb.table;
    DDNS_HIP(hipMemcpyAsync(hb.table, th, (size_t)R * sizeof(GmDesc<T>), hipMemcpyHostToDevice, s));
    constexpr int BATCH = 3;
    std::vector<double> st6((size_t)R * 6, 0.0);
    std::vector<char> done(R, 0);
    int iter = 0;
    bool firstcheck = true;
    while (iter < maxiter) {
        bool any = false;
        for (int r = 0; r < R; ++r) any = any || (!done[r] && !redo[r]);
        if (!any) break;
        const int last = std::min(maxiter, iter + (iter == 0 ? first : BATCH));
        const dim3 g(nredmax, R);
        for (int it = iter + 1; it <= last; ++it) {
            if (it == 1) hipLaunchKernelGGL((kb_spmv_mdot<T, 1, 8>), g, dim3(RT), 0, s, td, L, it);
            else if (it <= 8) hipLaunchKernelGGL((kb_spmv_mdot<T, 2, 8>), g, dim3(RT), 0, s, td, L, it);
            else if (it <= 12) hipLaunchKernelGGL((kb_spmv_mdot<T, 2, 12>), g, dim3(RT), 0, s, td, L, it);
            else {
                hipLaunchKernelGGL((kb_spmv2<T>), dim3(nbrmax, R), dim3(HT), 0, s, td, L, it);
                hipLaunchKernelGGL((kb_mdot<T>), g, dim3(RT), 0, s, td, L, it, 0);
            }
            if (it <= 8) hipLaunchKernelGGL((kb_maxpy_mdot<T, 8>), g, dim3(RT), 0, s, td, L, it);
            else if (it <= 12) hipLaunchKernelGGL((kb_maxpy_mdot<T, 12>), g, dim3(RT), 0, s, td, L, it);
            else {
                hipLaunchKernelGGL((kb_maxpy<T>), g, dim3(RT), 0, s, td, L, it, 0);
                hipLaunchKernelGGL((kb_mdot<T>), g, dim3(RT), 0, s, td, L, it, 1);
            }
            hipLaunchKernelGGL((kb_maxpy<T>), g, dim3(RT), 0, s, td, L, it, 1);
            hipLaunchKernelGGL((kb_ls<T>), dim3(R), dim3(LT), 0, s, td, L, it);
        }
        DDNS_HIP(hipGetLastError());
        iter = last;
        // one check for every system: the heads (and, the first time, the lists' flags)
        hipLaunchKernelGGL((kb_heads<T>), dim3(R), dim3(64), 0, s, td, hb.heads, firstcheck ? 1 : 0);
        DDNS_HIP(hipMemcpyAsync(hb.heads_host, hb.heads, (size_t)R * 16 * sizeof(double), hipMemcpyDeviceToHost, s));
        DDNS_HIP(hipStreamSynchronize(s));
        bool retable = false;
        for (int r = 0; r < R; ++r) {
            if (redo[r] || !th[r].sm) {
                done[r] = 1;
                st6[6 * (size_t)r + 4] = -1.0;
                continue;
            }
            const double* h = hb.heads_host + 16 * (size_t)r;
            for (int k = 0; k < 6; ++k) st6[6 * (size_t)r + k] = h[k];
            if (firstcheck) {
                int f[8];
                std::memcpy(f, h + 8, sizeof(f));
                // the system leaves the merged launches before its lists are reallocated
                if (f[0] & 3 || f[4] > pbs[r]->ell_width) {
                    th[r].nred = 0;
                    th[r].nbr = 0;
                    retable = true;
                }
                const int lo = lists_outcome<T>(*pbs[r], f);
                if (lo < 0) {
                    status[r] = lo;
                    DDNS_HIP(hipMemcpyAsync(hb.table, th, (size_t)R * sizeof(GmDesc<T>), hipMemcpyHostToDevice, s));
                    DDNS_HIP(hipStreamSynchronize(s));
                    return lo;
                }
                if (lo > 0) redo[r] = 1;        // widened: solved again on its own
            }
            done[r] = st6[6 * (size_t)r + 4] != 0.0;
        }
        if (retable) DDNS_HIP(hipMemcpyAsync(hb.table, th, (size_t)R * sizeof(GmDesc<T>), hipMemcpyHostToDevice, s));
        firstcheck = false;
    }
    for (int r = 0; r < R; ++r) {
        if (redo[r]) continue;
        PartBuffers& pb = *pbs[r];
        const double* h = &st6[6 * (size_t)r];
        if (h[4] < 0.0) continue;      // initial residual below 1e-5: no iteration
        const bool ok = h[4] != 0.0;
        const int used = (int)h[5];
        pb.gm_last_iters = used;
        const int n = pcs[r]->n;
        const long long ld = (3LL * n + 63) / 64 * 64;
        hipLaunchKernelGGL((k_solution<T>), dim3((n + HT - 1) / HT), dim3(HT), 0, s, pb.gm_x0, (const T*)pb.gm_Q, ld,
                           pb.gm_small + L.Y0(), (const double*)(pb.gm_small + 5), (const int*)pb.sorted, pb.pert, n,
                           pb.cap);
        iters[r] = used;
        res[r] = h[2];
        if (!ok && h[2] > h[3]) status[r] = DDNS_WARN_GMRES;                // 2980-2982
    }
    DDNS_HIP(hipGetLastError());
    return 0;
}

int part_hdi_batch(HdiBatch& hb, PartBuffers* const* pbs, const PartConst* const* pcs, int R, int maxiter,
                   int* iters, double* res, int* status, int* redo, hipStream_t s)
{
    if (R <= 0) return 0;
    if (pcs[0]->precision == DDNS_FP64) return hdi_batch_impl<double>(hb, pbs, pcs, R, maxiter, iters, res, status, redo, s);
    return hdi_batch_impl<float>(hb, pbs, pcs, R, maxiter, iters, res, status, redo, s);
}

}  // namespace ddns
