// hdi.hip -- hydrodynamic interactions between droplets.
//
// Reference: NEI_LIST_PRECOMP partlib.f90:3218-3371 (neighbours within HDI_trunc
// radii of the *source* droplet, Stokes-sphere coefficients M = gamma r r + beta I),
// AFUN_PRECOMP 3375-3510 (A x = x + M x and b = M (V - U)), HDI_GMRES 2795-3005
// (unrestarted GMRES: modified Gram-Schmidt Arnoldi, explicit Givens factor Qn,
// back substitution, tolerance relaxed at iteration maxiter/2 and from maxiter-2,
// warm start from the previous disturbance velocity) and EUC_NORM 2756-2791.
//
// Layout: the neighbour lists are CSR rows (column index + the 6 distinct entries
// of the symmetric 3x3 block), built in two passes over the coarse cell grid.
// Krylov vectors use the reference's blocked order x(1:n), x(n+1:2n), x(2n+1:3n).
// In FP32-mixed mode the matrix and the Krylov basis are float, every inner product
// and norm accumulates in double, and the small least-squares problem is double.
// Reductions are two-pass with a fixed grid, so results do not depend on timing.
#include "ddns_internal.h"

#include <algorithm>
#include <cmath>

namespace ddns {

constexpr int HT = 256;
constexpr int RED_BLOCKS = 2048;   // fixed reduction grid: 8 work-groups per CU

int exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t s);

// ------------------------------------------------------------------ neighbour lists
template <typename T>
__global__ void __launch_bounds__(HT) k_nei(int pass, const double4* __restrict__ pk,
                                            const int* __restrict__ cell_start, const int* __restrict__ sorted,
                                            int n, int cap, int G, double hc, double pi2, double trunc,
                                            int* cnt, const int* ptr, int* idx, T* mm, int* flags)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = sorted[t];
    const double4 me = pk[t];
    const double xi = me.x, yi = me.y, zi = me.z, ri = me.w;
    int cc[3];
    const double xs[3] = {xi, yi, zi};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int q = (int)(xs[c] / hc);
        cc[c] = q < 0 ? 0 : (q >= G ? G - 1 : q);
    }
    int k = 0;
    int base = pass ? ptr[i] : 0;
    for (int ox = -1; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
                int qx = cc[0] + ox, qy = cc[1] + oy, qz = cc[2] + oz;
                const int wx = qx < 0 ? -1 : (qx >= G ? 1 : 0);
                const int wy = qy < 0 ? -1 : (qy >= G ? 1 : 0);
                const int wz = qz < 0 ? -1 : (qz >= G ? 1 : 0);
                qx -= wx * G; qy -= wy * G; qz -= wz * G;
                const int cell = cell_code(qx, qy, qz);
                const int e = cell_start[cell + 1];
                for (int p = cell_start[cell]; p < e; ++p) {
                    if (p == t) continue;
                    const double4 o = pk[p];
                    const double d0 = xi - (o.x + (double)wx * pi2);
                    const double d1 = yi - (o.y + (double)wy * pi2);
                    const double d2 = zi - (o.z + (double)wz * pi2);
                    const double dnij = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                    const double rj = o.w;
                    if (dnij < ri + rj) atomicOr(flags, 1);                 // 3294-3299
                    if (!(dnij < trunc * rj)) continue;                     // 3302-3305
                    if (pass) {
                        if (k < 100) {
                            double fact1 = rj / dnij;                       // 3312-3324
                            const double fact2 = fact1 * fact1 * fact1;
                            fact1 = 0.75 * fact1;
                            const double gamma = (fact1 - 0.75 * fact2) / (dnij * dnij);
                            const double beta = fact1 + 0.25 * fact2;
                            const size_t q = (size_t)base + k;
                            idx[q] = sorted[p];
                            mm[6 * q + 0] = (T)(gamma * (d0 * d0) + beta);
                            mm[6 * q + 1] = (T)(gamma * d0 * d1);
                            mm[6 * q + 2] = (T)(gamma * d0 * d2);
                            mm[6 * q + 3] = (T)(gamma * (d1 * d1) + beta);
                            mm[6 * q + 4] = (T)(gamma * d1 * d2);
                            mm[6 * q + 5] = (T)(gamma * (d2 * d2) + beta);
                        }
                    }
                    ++k;
                }
            }
    if (!pass) {
        if (k > DDNS_MAXNEI) { atomicOr(flags, 2); k = DDNS_MAXNEI; }
        cnt[i] = k;
    }
}

// y = x + M x (AFUN_PRECOMP); on first use also b = M (V - U).  x, y blocked.
template <typename T>
__global__ void __launch_bounds__(HT) k_spmv(const int* __restrict__ ptr, const int* __restrict__ idx,
                                             const T* __restrict__ mm, const T* __restrict__ x, T* __restrict__ y,
                                             int n, const double* x0d, double* y0d,
                                             const double* vel, const double* vp, double* b, int cap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = ptr[i], e = ptr[i + 1];
    double r0 = 0.0, r1 = 0.0, r2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    const size_t N = (size_t)n, C = (size_t)cap;
    for (int q = s; q < e; ++q) {
        const int j = idx[q];
        const double m0 = mm[6 * (size_t)q + 0], m1 = mm[6 * (size_t)q + 1], m2 = mm[6 * (size_t)q + 2];
        const double m3 = mm[6 * (size_t)q + 3], m4 = mm[6 * (size_t)q + 4], m5 = mm[6 * (size_t)q + 5];
        double v0, v1, v2;
        if (x0d) { v0 = x0d[j]; v1 = x0d[N + j]; v2 = x0d[2 * N + j]; }
        else { v0 = (double)x[j]; v1 = (double)x[N + j]; v2 = (double)x[2 * N + j]; }
        r0 = r0 + m0 * v0 + m1 * v1 + m2 * v2;
        r1 = r1 + m1 * v0 + m3 * v1 + m4 * v2;
        r2 = r2 + m2 * v0 + m4 * v1 + m5 * v2;
        if (b) {
            const double q0 = vel[j] - vp[j], q1 = vel[C + j] - vp[C + j], q2 = vel[2 * C + j] - vp[2 * C + j];
            b0 = b0 + m0 * q0 + m1 * q1 + m2 * q2;
            b1 = b1 + m1 * q0 + m3 * q1 + m4 * q2;
            b2 = b2 + m2 * q0 + m4 * q1 + m5 * q2;
        }
    }
    if (x0d) {
        y0d[i] = r0 + x0d[i];
        y0d[N + i] = r1 + x0d[N + i];
        y0d[2 * N + i] = r2 + x0d[2 * N + i];
    } else {
        y[i] = (T)(r0 + (double)x[i]);
        y[N + i] = (T)(r1 + (double)x[N + i]);
        y[2 * N + i] = (T)(r2 + (double)x[2 * N + i]);
    }
    if (b) { b[i] = b0; b[N + i] = b1; b[2 * N + i] = b2; }
}

// ------------------------------------------------------------------ reductions
template <typename TA, typename TB>
__global__ void __launch_bounds__(HT) k_dot_part(const TA* a, const TB* b, long long n, double* part)
{
    __shared__ double s[HT];
    double acc = 0.0;
#pragma unroll 4
    for (long long i = (long long)blockIdx.x * HT + threadIdx.x; i < n; i += (long long)gridDim.x * HT)
        acc += (double)a[i] * (double)b[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = HT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// out = sum(part) (mode 0), sqrt(sum) (mode 1)
__global__ void __launch_bounds__(HT) k_dot_final(const double* part, int np, double* out, int mode)
{
    __shared__ double s[HT];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += HT) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = HT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = mode ? sqrt(s[0]) : s[0];
}

template <typename TA, typename TB>
static int dot(const TA* a, const TB* b, long long n, double* part, double* out, int mode, hipStream_t s)
{
    const int nb = (int)std::min<long long>(RED_BLOCKS, (n + HT - 1) / HT);
    hipLaunchKernelGGL((k_dot_part<TA, TB>), dim3(nb > 0 ? nb : 1), dim3(HT), 0, s, a, b, n, part);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(HT), 0, s, part, nb > 0 ? nb : 1, out, mode);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ vector kernels
template <typename T>
__global__ void k_residual0(const double* b, const double* ax0, T* q, long long n)
{
    const long long i = (long long)blockIdx.x * HT + threadIdx.x;
    if (i < n) q[i] = (T)(b[i] - ax0[i]);
}

template <typename T>
__global__ void k_scale(T* q, const double* den, long long n)
{
    const long long i = (long long)blockIdx.x * HT + threadIdx.x;
    if (i < n) q[i] = (T)((double)q[i] / *den);
}

// v = v - h * q, h read from device memory (MGS, partlib.f90:2892)
template <typename T>
__global__ void k_axpy_dev(T* v, const T* q, const double* h, long long n)
{
    const long long i = (long long)blockIdx.x * HT + threadIdx.x;
    if (i < n) v[i] = (T)((double)v[i] - *h * (double)q[i]);
}

// q_out = v / h
template <typename T>
__global__ void k_div_dev(T* qo, const T* v, const double* h, long long n)
{
    const long long i = (long long)blockIdx.x * HT + threadIdx.x;
    if (i < n) qo[i] = (T)((double)v[i] / *h);
}

// ---- block-Gram-Schmidt kernels (Arnoldi step with classical Gram-Schmidt applied
// twice).  Every reduction uses a fixed grid of RED_BLOCKS work-groups and a fixed
// summation tree, so the result does not depend on scheduling.
constexpr int MD_J = 8;   // Krylov columns per pass of k_mdot

// part[j * gridDim.x + block] = sum over this block's share of Q_j . v, j < k
template <typename T>
__global__ void __launch_bounds__(HT) k_mdot(const T* __restrict__ Q, const T* __restrict__ v, int k, long long n3,
                                             double* __restrict__ part)
{
    __shared__ double sh[MD_J][HT];
    const int tid = threadIdx.x;
    for (int j0 = 0; j0 < k; j0 += MD_J) {
        double acc[MD_J];
#pragma unroll
        for (int jj = 0; jj < MD_J; ++jj) acc[jj] = 0.0;
#pragma unroll 2
        for (long long i = (long long)blockIdx.x * HT + tid; i < n3; i += (long long)gridDim.x * HT) {
            const double vi = (double)v[i];
#pragma unroll
            for (int jj = 0; jj < MD_J; ++jj)
                if (j0 + jj < k) acc[jj] += (double)Q[(size_t)(j0 + jj) * n3 + i] * vi;
        }
#pragma unroll
        for (int jj = 0; jj < MD_J; ++jj) sh[jj][tid] = acc[jj];
        __syncthreads();
        for (int o = HT / 2; o > 0; o >>= 1) {
            if (tid < o) {
#pragma unroll
                for (int jj = 0; jj < MD_J; ++jj) sh[jj][tid] += sh[jj][tid + o];
            }
            __syncthreads();
        }
        if (tid < MD_J && j0 + tid < k) part[(size_t)(j0 + tid) * gridDim.x + blockIdx.x] = sh[tid][0];
        __syncthreads();
    }
}

// one work-group per column j: h[j] = sum of its partials; the Hessenberg column
// entry is set (first pass) or incremented (second pass)
__global__ void __launch_bounds__(HT) k_mfinal(const double* __restrict__ part, int np, double* h, double* hcol, int add)
{
    __shared__ double sh[HT];
    const int j = blockIdx.x, tid = threadIdx.x;
    double acc = 0.0;
    for (int i = tid; i < np; i += HT) acc += part[(size_t)j * np + i];
    sh[tid] = acc;
    __syncthreads();
    for (int o = HT / 2; o > 0; o >>= 1) {
        if (tid < o) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    if (tid == 0) {
        h[j] = sh[0];
        hcol[j] = add ? hcol[j] + sh[0] : sh[0];
    }
}

// v -= sum_j h[j] Q_j; with normpart, also the block partials of |v|^2
template <typename T>
__global__ void __launch_bounds__(HT) k_maxpy(T* __restrict__ v, const T* __restrict__ Q, const double* __restrict__ h,
                                              int k, long long n3, double* normpart)
{
    __shared__ double hs[128];
    __shared__ double sh[HT];
    const int tid = threadIdx.x;
    for (int j = tid; j < k; j += HT) hs[j] = h[j];
    __syncthreads();
    double nacc = 0.0;
#pragma unroll 2
    for (long long i = (long long)blockIdx.x * HT + tid; i < n3; i += (long long)gridDim.x * HT) {
        double x = (double)v[i];
#pragma unroll 4
        for (int j = 0; j < k; ++j) x -= hs[j] * (double)Q[(size_t)j * n3 + i];
        const T xt = (T)x;
        v[i] = xt;
        nacc += (double)xt * (double)xt;
    }
    if (!normpart) return;
    sh[tid] = nacc;
    __syncthreads();
    for (int o = HT / 2; o > 0; o >>= 1) {
        if (tid < o) sh[tid] += sh[tid + o];
        __syncthreads();
    }
    if (tid == 0) normpart[blockIdx.x] = sh[0];
}

// pertvel = x0 + Q y (2951-2956)
template <typename T>
__global__ void k_solution(const double* x0, const T* Q, const double* y, int iters, double* pert, int n, int cap)
{
    const long long i = (long long)blockIdx.x * HT + threadIdx.x;
    const long long n3 = 3LL * n;
    if (i >= n3) return;
    double acc = 0.0;
    for (int j = 0; j < iters; ++j) acc = acc + (double)Q[(size_t)j * n3 + i] * y[j];
    const int c = (int)(i / n), r = (int)(i % n);
    pert[(size_t)c * cap + r] = x0[i] + acc;
}

__global__ void k_pert_to_x0(const double* pert, double* x0, int n, int cap)
{
    const long long i = (long long)blockIdx.x * HT + threadIdx.x;
    if (i >= 3LL * n) return;
    const int c = (int)(i / n), r = (int)(i % n);
    x0[i] = pert[(size_t)c * cap + r];
}

// Small least-squares update of HDI_GMRES (2904-2971), one lane.
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

__device__ void ls_solve(double* sm, GmSmall L, int iter, int* done_host);
constexpr int LS_MAX = 8192;   // GmSmall::total() for maxiter <= 50, the reference's value

__global__ void __launch_bounds__(HT) k_ls(double* gsm, GmSmall L, int iter, int* done_host)
{
    // the whole small problem (at most LS_MAX doubles, 64 KB) is copied to LDS, solved
    // by one lane at LDS latency, and copied back
    __shared__ double sm[LS_MAX];
    const size_t tot = L.total();
    for (size_t i = threadIdx.x; i < tot; i += blockDim.x) sm[i] = gsm[i];
    __syncthreads();
    if (threadIdx.x == 0) ls_solve(sm, L, iter, done_host);
    __syncthreads();
    for (size_t i = threadIdx.x; i < tot; i += blockDim.x) gsm[i] = sm[i];
}

__device__ void ls_solve(double* sm, GmSmall L, int iter, int* done_host)
{
    const int M1 = L.M1();
    double* H = sm + L.H0();
    double* Rn = sm + L.R0();
    double* Qn = sm + L.Q0();
    double* bn = sm + L.B0();
    double* y = sm + L.Y0();
    double* v1 = sm + L.V1();
    double* v2 = sm + L.V2();
    double* rv = sm + L.RV();
#define HH(r, c) H[((r) - 1) + (size_t)M1 * ((c) - 1)]
#define RN(r, c) Rn[((r) - 1) + (size_t)M1 * ((c) - 1)]
#define QN(r, c) Qn[((r) - 1) + (size_t)M1 * ((c) - 1)]
    const double norm_rhs = sm[1];
    if (iter == 1) {
        const double nrm = sqrt(HH(1, 1) * HH(1, 1) + HH(2, 1) * HH(2, 1));
        const double q11 = HH(1, 1) / nrm, q21 = HH(2, 1) / nrm;
        QN(1, 1) = q11; QN(2, 1) = q21;
        QN(1, 2) = q21; QN(2, 2) = -q11;
        RN(1, 1) = sqrt(HH(1, 1) * HH(1, 1) + HH(2, 1) * HH(2, 1));
        bn[0] = q11 * norm_rhs;
        y[0] = bn[0] / RN(1, 1);
    } else {
        for (int r = 1; r <= iter; ++r) {
            double acc = 0.0;
            for (int c = 1; c <= iter; ++c) acc = acc + QN(r, c) * HH(c, iter);
            RN(r, iter) = acc;
        }
        const double aa = RN(iter, iter), bb = HH(iter + 1, iter);
        const double dd = sqrt(aa * aa + bb * bb);
        const double cc = aa / dd, ss = bb / dd;
        RN(iter, iter) = dd;
        for (int c = 1; c <= iter; ++c) { v1[c - 1] = cc * QN(iter, c); v2[c - 1] = -ss * QN(iter, c); }
        v1[iter] = ss; v2[iter] = cc;
        for (int c = 1; c <= iter + 1; ++c) { QN(iter, c) = v1[c - 1]; QN(iter + 1, c) = v2[c - 1]; }
        bn[iter - 2] = QN(iter - 1, 1) * norm_rhs;
        bn[iter - 1] = QN(iter, 1) * norm_rhs;
        y[iter - 1] = bn[iter - 1] / RN(iter, iter);
        for (int r = iter - 1; r >= 1; --r) {
            double d = 0.0;
            for (int c = r + 1; c <= iter; ++c) d = d + y[c - 1] * RN(r, c);
            y[r - 1] = (bn[r - 1] - d) / RN(r, r);
        }
    }
    for (int r = 1; r <= iter + 1; ++r) {
        double acc = 0.0;
        for (int c = 1; c <= iter; ++c) acc = acc + HH(r, c) * y[c - 1];
        rv[r - 1] = acc;
    }
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
    *done_host = done;
#undef HH
#undef RN
#undef QN
}

// ------------------------------------------------------------------ driver
template <typename T>
static int hdi_impl(PartBuffers& pb, const PartConst& pc, int maxiter, double tol, int* iters_out, double* res_out,
                    hipStream_t s)
{
    const int n = pc.n;
    const long long n3 = 3LL * n;
    *iters_out = 0;
    *res_out = 0.0;
    if (n == 0) return 0;
    DDNS_HIP(hipMemsetAsync(pb.flags, 0, 4 * sizeof(int), s));
    // neighbour lists, two passes
    hipLaunchKernelGGL((k_nei<T>), dim3((n + HT - 1) / HT), dim3(HT), 0, s, 0, (const double4*)pb.pk, pb.cell_start, pb.sorted,
                       n, pb.cap, pb.G, pb.hc, pc.pi2, pc.hdi_trunc, pb.nei_cnt, (const int*)nullptr, (int*)nullptr,
                       (T*)nullptr, pb.flags);
    int st = exclusive_scan(pb.nei_cnt, pb.nei_ptr, n, (int*)pb.scan_tmp, s);
    if (st) return st;
    int nnz = 0, flags = 0;
    DDNS_HIP(hipMemcpyAsync(&nnz, pb.nei_ptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipMemcpyAsync(&flags, pb.flags, sizeof(int), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    if (flags & 1) { set_error("inside NEI_LIST_PRECOMP: overlapping is not allowed"); return DDNS_ERR_OVERLAP; }
    if (flags & 2) { set_error("more than 100 HI neighbours for a droplet"); return DDNS_ERR_NEIGHBOUR_OVERFLOW; }
    if (nnz > pb.nnz_cap) {
        if (pb.nei_idx) { (void)hipFree(pb.nei_idx); (void)hipFree(pb.nei_m); }
        pb.nnz_cap = (long long)nnz + nnz / 4 + 1024;
        DDNS_HIP(hipMalloc(&pb.nei_idx, (size_t)pb.nnz_cap * sizeof(int)));
        DDNS_HIP(hipMalloc(&pb.nei_m, (size_t)pb.nnz_cap * 6 * sizeof(double)));
    }
    hipLaunchKernelGGL((k_nei<T>), dim3((n + HT - 1) / HT), dim3(HT), 0, s, 1, (const double4*)pb.pk, pb.cell_start, pb.sorted,
                       n, pb.cap, pb.G, pb.hc, pc.pi2, pc.hdi_trunc, pb.nei_cnt, (const int*)pb.nei_ptr, pb.nei_idx,
                       (T*)pb.nei_m, pb.flags);
    DDNS_HIP(hipGetLastError());
    // Krylov storage
    GmSmall L{maxiter};
    if (!pb.gm_small) {
        DDNS_HIP(hipMalloc(&pb.gm_small, 16384 * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_red, (size_t)(128 + 2) * RED_BLOCKS * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_x0, 3 * (size_t)pb.cap * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_b, 3 * (size_t)pb.cap * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_v, 3 * (size_t)pb.cap * sizeof(double)));
        DDNS_HIP(hipMalloc(&pb.gm_Q, (size_t)(maxiter + 1) * 3 * pb.cap * sizeof(T)));
    }
    if (L.total() > (size_t)LS_MAX) { set_error("GMRES maxiter too large (at most 50)"); return DDNS_ERR_ARG; }
    double* normpart = pb.gm_red + (size_t)128 * RED_BLOCKS;
    double* hbuf = normpart + RED_BLOCKS;
    double* sm = pb.gm_small;
    double* ax0 = (double*)pb.gm_v;       // reuse as double scratch for A x0
    T* Q = (T*)pb.gm_Q;
    T* v = Q + (size_t)maxiter * n3;      // the (maxiter+1)-th column doubles as the work vector
    DDNS_HIP(hipMemsetAsync(sm, 0, L.total() * sizeof(double), s));
    const unsigned nb3 = (unsigned)((n3 + HT - 1) / HT);
    hipLaunchKernelGGL(k_pert_to_x0, dim3(nb3), dim3(HT), 0, s, pb.pert, pb.gm_x0, n, pb.cap);
    hipLaunchKernelGGL((k_spmv<T>), dim3((n + HT - 1) / HT), dim3(HT), 0, s, pb.nei_ptr, pb.nei_idx, (const T*)pb.nei_m,
                       (const T*)nullptr, (T*)nullptr, n, pb.gm_x0, ax0, pb.vel, pb.vp, pb.gm_b, pb.cap);
    hipLaunchKernelGGL((k_residual0<T>), dim3(nb3), dim3(HT), 0, s, pb.gm_b, ax0, Q, n3);
    st = dot(pb.gm_b, pb.gm_b, n3, pb.gm_red, sm + 0, 1, s);
    if (st) return st;
    st = dot(Q, Q, n3, pb.gm_red, sm + 1, 1, s);
    if (st) return st;
    double h2[2];
    DDNS_HIP(hipMemcpyAsync(h2, sm, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    if (h2[1] < 1e-5) return 0;                                            // 2866-2868
    {
        double me = tol;
        DDNS_HIP(hipMemcpyAsync(sm + 3, &me, sizeof(double), hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL((k_scale<T>), dim3(nb3), dim3(HT), 0, s, Q, sm + 1, n3);
    int* done_dev = pb.flags + 8;
    int iter = 0, done = 0;
    double* Hd = sm + L.H0();
    const int M1 = L.M1();
    for (iter = 1; iter <= maxiter; ++iter) {
        T* qi = Q + (size_t)(iter - 1) * n3;
        hipLaunchKernelGGL((k_spmv<T>), dim3((n + HT - 1) / HT), dim3(HT), 0, s, pb.nei_ptr, pb.nei_idx,
                           (const T*)pb.nei_m, (const T*)qi, v, n, (const double*)nullptr, (double*)nullptr,
                           (const double*)nullptr, (const double*)nullptr, (double*)nullptr, pb.cap);
        // orthogonalise v against q_1..q_iter: classical Gram-Schmidt, twice (the
        // reference's modified Gram-Schmidt, 2886-2894, in two block passes)
        double* hcol = Hd + (size_t)M1 * (iter - 1);
        double* hn = hcol + iter;
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL((k_mdot<T>), dim3(RED_BLOCKS), dim3(HT), 0, s, (const T*)Q, (const T*)v, iter, n3, pb.gm_red);
            hipLaunchKernelGGL(k_mfinal, dim3(iter), dim3(HT), 0, s, (const double*)pb.gm_red, RED_BLOCKS, hbuf, hcol, pass);
            hipLaunchKernelGGL((k_maxpy<T>), dim3(RED_BLOCKS), dim3(HT), 0, s, v, (const T*)Q, (const double*)hbuf, iter, n3,
                               pass ? normpart : (double*)nullptr);
        }
        hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(HT), 0, s, (const double*)normpart, RED_BLOCKS, hn, 1);
        if (iter < maxiter) hipLaunchKernelGGL((k_div_dev<T>), dim3(nb3), dim3(HT), 0, s, Q + (size_t)iter * n3, v, hn, n3);
        hipLaunchKernelGGL(k_ls, dim3(1), dim3(HT), 0, s, sm, L, iter, done_dev);
        DDNS_HIP(hipGetLastError());
        DDNS_HIP(hipMemcpyAsync(&done, done_dev, sizeof(int), hipMemcpyDeviceToHost, s));
        DDNS_HIP(hipStreamSynchronize(s));
        if (done) break;
    }
    const int used = std::min(iter, maxiter);
    hipLaunchKernelGGL((k_solution<T>), dim3(nb3), dim3(HT), 0, s, pb.gm_x0, (const T*)Q, sm + L.Y0(), used, pb.pert, n, pb.cap);
    DDNS_HIP(hipGetLastError());
    double tail[6];
    DDNS_HIP(hipMemcpyAsync(tail, sm, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    *iters_out = used;
    *res_out = tail[2];
    if (!done && tail[2] > tail[3]) return DDNS_WARN_GMRES;                 // 2980-2982
    return 0;
}

int part_hdi(PartBuffers& pb, const PartConst& pc, int maxiter, double tol, int* iters, double* res, hipStream_t s)
{
    if (pc.precision == DDNS_FP64) return hdi_impl<double>(pb, pc, maxiter, tol, iters, res, s);
    return hdi_impl<float>(pb, pc, maxiter, tol, iters, res, s);
}

}  // namespace ddns
