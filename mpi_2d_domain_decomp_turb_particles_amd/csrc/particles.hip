// particles.hip -- Lagrangian droplet step on one MI355X (everything but the pair
// searches, which live in collide.hip, and the HI solve in hdi.hip).
//
// Reference: PREPARE_INTERP / INTERPOLATE partlib.f90:305-520 (6-point Lagrange
// gather), PART_HISTORY_FVP/FDP 553-625, PART_ADVANCE 629-726, UPDATE_PART_POSITION
// 730-757, PERIODIC wrap 1142-1151, RELOCATE 1984-2012.  The reference's linked
// cell list (HEADLIST 2171-2444) is replaced by a counting sort of droplets into a
// coarse periodic grid; periodic images are formed on the fly by the pair kernels.
//
// Droplet state is structure-of-arrays in double, slot i = the reference's array
// index i+1, so every index-order rule of the reference (ih2 > ih1, the order of the
// removal list and of the relocation draws) is evaluated on slot numbers.
#include "ddns_internal.h"

#include <algorithm>
#include <cmath>

namespace ddns {

constexpr int PT = 256;
constexpr int CELL_GMAX = 256;   // finest pair-search grid per dimension

static inline unsigned nblk(long long n, int t = PT) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------------------------ scans
// Exclusive scan of n ints in two launches: the totals of tiles of E * SCAN_T ints, then
// each tile scanned with the sum of the totals before it (every work-group adds those
// itself, at most a few thousand ints from L2).  No work-group waits for another.
// Work-groups are small (256 threads, no LDS tile): a scan runs beside the flow kernels,
// and a work-group of 1024 threads was seen to wait for a whole flow kernel to drain.
constexpr int SCAN_T = 256;
constexpr int SCAN_W = SCAN_T / 64;

// the work-group's sum of x, returned to every thread
__device__ __forceinline__ int block_sum_int(int x, int* sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = x;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int w = 0; w < SCAN_W; ++w) t += sh[w];
    __syncthreads();
    return t;
}

// thread tid's E consecutive ints of the tile (zero past n)
template <int E>
__device__ __forceinline__ void scan_load(const int* in, long long i0, int n, int (&v)[E])
{
    if (i0 + E <= n) {
#pragma unroll
        for (int q = 0; q < E; q += 4) {
            const int4 x = *(const int4*)(in + i0 + q);
            v[q] = x.x; v[q + 1] = x.y; v[q + 2] = x.z; v[q + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < E; ++q) v[q] = i0 + q < n ? in[i0 + q] : 0;
    }
}

template <int E>
__global__ void __launch_bounds__(SCAN_T) k_scan_tiles(const int* __restrict__ in, int* __restrict__ bsum, int n)
{
    __shared__ int sh[SCAN_W];
    int v[E];
    scan_load<E>(in, ((long long)blockIdx.x * SCAN_T + threadIdx.x) * E, n, v);
    int tot = 0;
#pragma unroll
    for (int q = 0; q < E; ++q) tot += v[q];
    tot = block_sum_int(tot, sh);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// in may alias out: a thread reads its ints before it writes them.  zin (= in, not
// aliasing out, or null): the input is cleared as it is read.
template <int E>
__global__ void __launch_bounds__(SCAN_T) k_scan_apply(const int* in, int* out, const int* __restrict__ bsum, int n,
                                                       int* zin)
{
    __shared__ int sh[SCAN_W];
    __shared__ int wsum[SCAN_W];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int pre = 0;
    for (int b = tid; b < (int)blockIdx.x; b += SCAN_T) pre += bsum[b];
    pre = block_sum_int(pre, sh);
    const long long i0 = ((long long)blockIdx.x * SCAN_T + tid) * E;
    int v[E];
    scan_load<E>(in, i0, n, v);
    if (zin) {
        if (i0 + E <= n) {
#pragma unroll
            for (int q = 0; q < E; q += 4) *(int4*)(zin + i0 + q) = make_int4(0, 0, 0, 0);
        } else {
            for (int q = 0; q < E; ++q)
                if (i0 + q < n) zin[i0 + q] = 0;
        }
    }
    int tot = 0;
#pragma unroll
    for (int q = 0; q < E; ++q) tot += v[q];
    int incl = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int run = pre + incl - tot;
#pragma unroll
    for (int w = 0; w < SCAN_W; ++w)
        if (w < wv) run += wsum[w];
    int r[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
        r[q] = run;
        run += v[q];
    }
    if (i0 + E <= n) {
#pragma unroll
        for (int q = 0; q < E; q += 4) *(int4*)(out + i0 + q) = make_int4(r[q], r[q + 1], r[q + 2], r[q + 3]);
    } else {
        for (int q = 0; q < E; ++q)
            if (i0 + q < n) out[i0 + q] = r[q];
    }
    if (blockIdx.x == gridDim.x - 1 && tid == SCAN_T - 1) out[n] = run;
}

template <int E>
static int scan_launch(const int* in, int* out, int n, int* tmp, int* zin, hipStream_t s)
{
    const unsigned nb = (unsigned)(((long long)n + (long long)E * SCAN_T - 1) / ((long long)E * SCAN_T));
    hipLaunchKernelGGL(k_scan_tiles<E>, dim3(nb), dim3(SCAN_T), 0, s, in, tmp, n);
    hipLaunchKernelGGL(k_scan_apply<E>, dim3(nb), dim3(SCAN_T), 0, s, in, out, (const int*)tmp, n, zin);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// out[0..n-1] = exclusive scan of in, out[n] = total (in may alias out); tmp holds
// scan_tmp_ints(n) ints; in and out 16-byte aligned.  clear_in (in not aliasing out):
// in[0..n-1] is left zero.
int exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t s, bool clear_in)
{
    int* zin = clear_in ? const_cast<int*>(in) : nullptr;
    if (n <= 0) {
        hipLaunchKernelGGL(k_scan_apply<4>, dim3(1), dim3(SCAN_T), 0, s, in, out, (const int*)tmp, 0, zin);
        DDNS_HIP(hipGetLastError());
        return 0;
    }
    if ((long long)n <= 16LL * SCAN_T * 1024) return scan_launch<16>(in, out, n, tmp, zin, s);
    return scan_launch<32>(in, out, n, tmp, zin, s);
}

// ints of scratch exclusive_scan needs for n elements
size_t scan_tmp_ints(long long n)
{
    size_t t = 0;
    long long m = n;
    do {
        m = (m + 1023) / 1024;
        t += (size_t)m + 1;
    } while (m > 1);
    return t + 16;
}

// ------------------------------------------------------------------ allocation
int part_alloc(PartBuffers& pb, int cap, int nset, int G)
{
    pb.cap = cap;
    const size_t c = (size_t)cap;
    DDNS_HIP(hipMalloc(&pb.pos, 3 * c * 8));
    DDNS_HIP(hipMalloc(&pb.vel, 3 * c * 8));
    DDNS_HIP(hipMalloc(&pb.rad, c * 8));
    DDNS_HIP(hipMalloc(&pb.gid, c * 8));
    DDNS_HIP(hipMalloc(&pb.fvp, 9 * c * 8));
    DDNS_HIP(hipMalloc(&pb.fdp, 12 * c * 8));
    DDNS_HIP(hipMalloc(&pb.dmove, 3 * c * 8));
    DDNS_HIP(hipMalloc(&pb.vp, 3 * c * 8));
    DDNS_HIP(hipMalloc(&pb.pert, 3 * c * 8));
    DDNS_HIP(hipMemset(pb.pert, 0, 3 * c * 8));
    DDNS_HIP(hipMemset(pb.dmove, 0, 3 * c * 8));
    DDNS_HIP(hipMemset(pb.vp, 0, 3 * c * 8));
    DDNS_HIP(hipMalloc(&pb.cell_of, c * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.cell_rank, c * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.sorted, c * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.pk, c * sizeof(double4)));
    DDNS_HIP(hipMalloc(&pb.mult, (c + 1) * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.draw_off, (c + 1) * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.nei_cnt, (c + 1) * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.flags, 64 * sizeof(int)));
    DDNS_HIP(hipMemset(pb.flags, 0, 64 * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.tested, 4 * sizeof(long long)));
    const size_t nb = (size_t)nset * nset * DDNS_NBIN;
    DDNS_HIP(hipMalloc(&pb.nirbin, nb * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.bins, 7 * nb * sizeof(double)));
    DDNS_HIP(hipMalloc(&pb.counts, 512 * sizeof(int)));
    pb.pair_cap = 1 << 16;
    DDNS_HIP(hipMalloc(&pb.pairs, (size_t)pb.pair_cap * DDNS_PAIR_REC * sizeof(long long)));
    pb.draw_cap = 1 << 16;
    DDNS_HIP(hipMalloc(&pb.draws, (size_t)pb.draw_cap * sizeof(double)));
    DDNS_HIP(hipMalloc(&pb.rng_state, 64 * sizeof(int)));
    DDNS_HIP(hipMemset(pb.rng_state, 0, 64 * sizeof(int)));
    DDNS_HIP(hipMalloc(&pb.red, 8192 * sizeof(double)));
    pb.scan_tmp_bytes = scan_tmp_ints(std::max((long long)G * G * G + 1, (long long)c + 1)) * sizeof(int);
    DDNS_HIP(hipMalloc(&pb.scan_tmp, pb.scan_tmp_bytes));
    pb.G = 0;
    return 0;
}

void part_free(PartBuffers& pb)
{
    void* ps[] = {pb.pos, pb.vel, pb.rad, pb.gid, pb.fvp, pb.fdp, pb.dmove, pb.vp, pb.pert,
                  pb.cell_count, pb.cell_start, pb.cell_of, pb.cell_rank, pb.sorted, pb.pk, pb.scan_tmp,
                  pb.nei_cnt, pb.nei_idx, pb.nei_m, pb.gm_Q, pb.gm_v, pb.gm_x0, pb.gm_b, pb.gm_w,
                  pb.gm_red, pb.gm_small, pb.mult, pb.cand, pb.hit, pb.flags, pb.tested, pb.nirbin, pb.bins, pb.counts,
                  pb.pairs, pb.contact, pb.draws, pb.draw_off, pb.rng_state, pb.red};
    for (void* p : ps) if (p) (void)hipFree(p);
    void* sp[] = {pb.spare.cell_count, pb.spare.cell_start, pb.spare.cell_of, pb.spare.sorted, pb.spare.pk};
    for (void* p : sp) if (p) (void)hipFree(p);
    if (pb.sync_ev) (void)hipEventDestroy(pb.sync_ev);
    if (pb.gm_host) (void)hipHostFree(pb.gm_host);
    if (pb.found_host) (void)hipHostFree(pb.found_host);
    if (pb.counts_host) (void)hipHostFree(pb.counts_host);
    pb = PartBuffers{};
}

// ------------------------------------------------------------------ interpolation
// Lagrange weights for nodes node-2 .. node+3 at fractional position z (partlib.f90:333-338)
__device__ __forceinline__ void lagrange6(double z, double* bg)
{
    const double z2 = z * z, z3 = z2 * z, z4 = z3 * z, z5 = z4 * z;
    bg[0] = (6.0 * z - 5.0 * z2 - 5.0 * z3 + 5.0 * z4 - z5) / 120.0;
    bg[1] = (-12.0 * z + 16.0 * z2 - z3 - 4.0 * z4 + z5) / 24.0;
    bg[2] = (12.0 - 4.0 * z - 15.0 * z2 + 5.0 * z3 + 3.0 * z4 - z5) / 12.0;
    bg[3] = (12.0 * z + 8.0 * z2 - 7.0 * z3 - 2.0 * z4 + z5) / 12.0;
    bg[4] = (-6.0 * z - z2 + 7.0 * z3 + z4 - z5) / 24.0;
    bg[5] = (4.0 * z - 5.0 * z3 + z5) / 120.0;
}

// The physical velocity is stored node-major, [z][y][x][3] (u, v, w), so one node is one
// 12-byte (float) or three 8-byte (double) loads, and the six nodes of a stencil row are
// one contiguous run of 72 (144) bytes.
__device__ __forceinline__ void load_node(const float* p, double& a, double& b, double& c)
{
    const float3 v = *(const float3*)p;
    a = v.x; b = v.y; c = v.z;
}
__device__ __forceinline__ void load_node(const double* p, double& a, double& b, double& c)
{
    a = p[0]; b = p[1]; c = p[2];
}

// Six lanes per droplet, one per x offset of the 6 x 6 x 6 stencil: for each of the
// 36 (y, z) stencil rows the six lanes read the row's six nodes as one contiguous run,
// so a wavefront load touches about 10 cache lines instead of one line per lane.  Ten
// droplets share a wavefront (lanes 60..63 idle), taken in cell order (sorted[]) so
// that they read overlapping rows.  Each lane sums its 36 nodes with the weights
// bg(x) bg(y) bg(z) of partlib.f90:509; the six partial sums are added in x order.
constexpr int IG = 6;              // lanes per droplet
constexpr int IPW = 64 / IG;       // droplets per wavefront

// In cell order (order != null) it writes the result in cell order, us[c * n + t], and
// k_interp_scatter takes it to the droplets: the gather then shares L2 with nothing but
// the field, instead of with three droplet arrays read and one written a line per
// droplet each (which halved its L2 hit rate).  Without a cell order it writes vp.
template <typename T>
__global__ void __launch_bounds__(PT) k_interp(const T* __restrict__ u, int N, double pi2,
                                               const double* __restrict__ pos, const int* __restrict__ order,
                                               const double4* __restrict__ pk, double* __restrict__ vp, int n, int cap,
                                               double* __restrict__ us)
{
    const int lane = threadIdx.x & 63;
    const int wave = (xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) >> 6;
    const int g = lane / IG, a = lane - g * IG;
    const int t = wave * IPW + g;
    const bool active = g < IPW && t < n;
    const int i = active ? (order ? order[t] : t) : 0;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (active) {
        const double dx = pi2 / (double)N;
        double wx[6], wy[6], wz[6];
        double y0, y1, y2;
        if (pk) {                         // the packed record of the cell order holds the same values
            const double4 r = pk[t];
            y0 = r.x; y1 = r.y; y2 = r.z;
        } else {
            y0 = pos[i]; y1 = pos[(size_t)cap + i]; y2 = pos[2 * (size_t)cap + i];
        }
        const int nx = (int)floor(y0 / dx), ny = (int)floor(y1 / dx), nz = (int)floor(y2 / dx);
        lagrange6((y0 - (double)nx * dx) / dx, wx);
        lagrange6((y1 - (double)ny * dx) / dx, wy);
        lagrange6((y2 - (double)nz * dx) / dx, wz);
        double wxa = wx[0];
#pragma unroll
        for (int m = 1; m < 6; ++m) wxa = a == m ? wx[m] : wxa;
        const size_t ix = (size_t)(((nx - 2 + a) % N + N) % N) * 3;
        size_t rowy[6], rowz[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            rowy[m] = (size_t)(((ny - 2 + m) % N + N) % N) * N;
            rowz[m] = (size_t)(((nz - 2 + m) % N + N) % N) * N * N;
        }
#pragma unroll
        for (int d = 0; d < 6; ++d) {
#pragma unroll
            for (int b2 = 0; b2 < 6; ++b2) {
                double u0, u1, u2;
                load_node(u + (rowz[d] + rowy[b2]) * 3 + ix, u0, u1, u2);
                const double w = wxa * wy[b2] * wz[d];
                acc0 += u0 * w;
                acc1 += u1 * w;
                acc2 += u2 * w;
            }
        }
    }
    // lane a == 0 of each group adds the group's six partial sums in x order
    const int g0 = g * IG;
    double s0 = acc0, s1 = acc1, s2 = acc2;
#pragma unroll
    for (int m = 1; m < IG; ++m) {
        const int src = g0 + m < 64 ? g0 + m : lane;
        const double p0 = __shfl(acc0, src, 64), p1 = __shfl(acc1, src, 64), p2 = __shfl(acc2, src, 64);
        s0 += p0; s1 += p1; s2 += p2;
    }
    if (active && a == 0) {
        if (us) {
            const size_t M = (size_t)n;
            us[t] = s0;
            us[M + t] = s1;
            us[2 * M + t] = s2;
        } else {
            const size_t C = (size_t)cap;
            vp[i] = s0;
            vp[C + i] = s1;
            vp[2 * C + i] = s2;
        }
    }
}

// The gathered velocities, in cell order in us, to the droplets (vp in the reference's
// order).  With HI on it also writes, in cell order, the inputs of the HI solve: the
// warm start x0 = pertvel and the velocity difference w = V - U, so the solve reads
// them contiguously; w may be us itself (each element is read before it is written).
__global__ void __launch_bounds__(PT) k_interp_scatter(const int* __restrict__ order, const double* us, double* vp,
                                                       int n, int cap, const double* __restrict__ vel,
                                                       const double* __restrict__ pert, double* __restrict__ x0,
                                                       double* w)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = order[t];
    const size_t C = (size_t)cap, M = (size_t)n;
    double s3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) s3[c] = us[c * M + t];
#pragma unroll
    for (int c = 0; c < 3; ++c) vp[c * C + i] = s3[c];
    if (x0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x0[c * M + t] = pert[3 * (size_t)i + c];
            w[c * M + t] = vel[c * C + i] - s3[c];
        }
    }
}

int part_interp(PartBuffers& pb, const PartConst& pc, const void* uphys, hipStream_t s)
{
    if (pc.n == 0) return 0;
    const int* order = pb.G > 0 ? pb.sorted : nullptr;
    const double4* pk = order ? pb.pk : nullptr;
    // the HI solve works in the cell order of this grid
    if (pc.hdi && !order) { set_error("HI needs the cell grid before the interpolation"); return DDNS_ERR_ARG; }
    if (order && !pb.gm_w) DDNS_HIP(hipMalloc(&pb.gm_w, 3 * (size_t)pb.cap * sizeof(double)));
    if (pc.hdi && !pb.gm_x0) DDNS_HIP(hipMalloc(&pb.gm_x0, 3 * (size_t)pb.cap * sizeof(double)));
    double* us = order ? pb.gm_w : nullptr;       // cell-order result, then w in place
    const unsigned nb = (unsigned)(((long long)pc.n + IPW * (PT / 64) - 1) / (IPW * (PT / 64)));
    if (pc.precision == DDNS_FP64)
        hipLaunchKernelGGL(k_interp<double>, dim3(nb), dim3(PT), 0, s, (const double*)uphys, pc.N, pc.pi2,
                           pb.pos, order, pk, pb.vp, pc.n, pb.cap, us);
    else
        hipLaunchKernelGGL(k_interp<float>, dim3(nb), dim3(PT), 0, s, (const float*)uphys, pc.N, pc.pi2,
                           pb.pos, order, pk, pb.vp, pc.n, pb.cap, us);
    if (order)
        hipLaunchKernelGGL(k_interp_scatter, dim3(nblk(pc.n)), dim3(PT), 0, s, order, (const double*)us, pb.vp, pc.n,
                           pb.cap, (const double*)pb.vel, (const double*)pb.pert, pc.hdi ? pb.gm_x0 : nullptr,
                           pc.hdi ? pb.gm_w : nullptr);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------ histories
// PART_HISTORY_FVP, partlib.f90:562-592
__global__ void k_hist_fvp(double* vel, const double* rad, double* fvp, int n, int cap, int newp, double Sv, int v0,
                           int v1, int v2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (newp) {
        const double a = rad[i];
        vel[i] = Sv * (a * a);
        vel[(size_t)cap + i] = 0.0;
        vel[2 * (size_t)cap + i] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double v = vel[(size_t)c * cap + i];
        fvp[(size_t)(v0 + c) * cap + i] = v;
        if (newp) { fvp[(size_t)(v1 + c) * cap + i] = v; fvp[(size_t)(v2 + c) * cap + i] = v; }
    }
}

int part_history_fvp(PartBuffers& pb, const PartConst& pc, hipStream_t s)
{
    if (pc.n == 0) return 0;
    hipLaunchKernelGGL(k_hist_fvp, dim3(nblk(pc.n)), dim3(PT), 0, s, pb.vel, pb.rad, pb.fvp, pc.n, pb.cap,
                       pc.new_particle, pc.Sv_const, pb.vcol(0), pb.vcol(1), pb.vcol(2));
    DDNS_HIP(hipGetLastError());
    pb.v0_stale = false;
    return 0;
}

// ------------------------------------------------------------------ advance
// PART_HISTORY_FDP (599-625) + PART_ADVANCE (629-726): implicit Adams-Moulton
// velocity with Adams-Bashforth fluid history, Stokes or Schiller-Naumann drag,
// gravity along +x, displacement dmove.
__global__ void k_advance(double* vel, const double* rad, double* fvp, double* fdp, const double* vp,
                          const double* pert, double* dmove, int n, int cap, int newp, int hdi,
                          int ndrag, double dt, double rnu, double Sv, double tau2, int4 vc, int4 dc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t C = (size_t)cap;
    double f0[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        f0[c] = hdi ? vp[c * C + i] + pert[3 * (size_t)i + c] : vp[c * C + i];
        fdp[(dc.x + c) * C + i] = f0[c];
        if (newp) { fdp[(dc.y + c) * C + i] = f0[c]; fdp[(dc.z + c) * C + i] = f0[c]; fdp[(dc.w + c) * C + i] = f0[c]; }
    }
    const double a = rad[i];
    double FSN = 1.;
    if (ndrag == 1) {
        double d[3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
            d[c] = hdi ? vel[c * C + i] - vp[c * C + i] - pert[3 * (size_t)i + c] : vel[c * C + i] - vp[c * C + i];
        const double Vrel = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        FSN = 1. + 0.15 * pow(Vrel * 2. * a / rnu, 0.687);
    }
    const double taup = tau2 * (a * a);
    const double dtau = dt / (24.0 * taup) * FSN;
    const double wp_vel = Sv * (a * a);
    const double dtc = dt / 24.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double unitv = c == 0 ? 1.0 : 0.0;
        const double fv0 = fvp[(vc.x + c) * C + i], fv1 = fvp[(vc.y + c) * C + i], fv2 = fvp[(vc.z + c) * C + i];
        const double tempv = 19.0 * fv0 - 5.0 * fv1 + fv2;
        const double tempd = 55.0 * f0[c] - 59.0 * fdp[(dc.y + c) * C + i] + 37.0 * fdp[(dc.z + c) * C + i] -
                             9.0 * fdp[(dc.w + c) * C + i];
        const double vnew = (vel[c * C + i] + dt * unitv * wp_vel / taup * FSN - dtau * tempv + dtau * tempd) / (1.0 + 9.0 * dtau);
        vel[c * C + i] = vnew;
        dmove[c * C + i] = dtc * (9.0 * vnew + 19.0 * fv0 - 5.0 * fv1 + fv2);
    }
}

int part_advance(PartBuffers& pb, const PartConst& pc, hipStream_t s)
{
    if (pc.n == 0) return 0;
    const int4 vc = make_int4(pb.vcol(0), pb.vcol(1), pb.vcol(2), 0);
    const int4 dc = make_int4(pb.dcol(0), pb.dcol(1), pb.dcol(2), pb.dcol(3));
    hipLaunchKernelGGL(k_advance, dim3(nblk(pc.n)), dim3(PT), 0, s, pb.vel, pb.rad, pb.fvp, pb.fdp, pb.vp, pb.pert,
                       pb.dmove, pc.n, pb.cap, pc.new_particle, pc.hdi, pc.ndrag, pc.dt, pc.rnu, pc.Sv_const, pc.tau2,
                       vc, dc);
    DDNS_HIP(hipGetLastError());
    pb.d0_stale = false;
    return 0;
}

// ------------------------------------------------------------------ position update, relocation, wrap
// UPDATE_PART_POSITION (741-751)
// PERIODIC wrap (1142-1151) of one coordinate
__device__ __forceinline__ double wrap_periodic(double y, double pi2)
{
    if (y >= pi2) y = y - pi2;
    if (y < 0.0) y = y + pi2;
    return y;
}

// UPDATE_PART_POSITION (730-757): the position, with the PERIODIC wrap of 608 applied at
// once (the relocation between them wraps the positions it draws itself); the history
// shift turns the rings (part_update_and_wrap)
__global__ void k_update_pos(double* pos, const double* dmove, int n, int cap, double pi2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t C = (size_t)cap;
#pragma unroll
    for (int c = 0; c < 3; ++c) pos[c * C + i] = wrap_periodic(pos[c * C + i] + dmove[c * C + i], pi2);
}


// ranff2, accesories.f90:107-147, sequential on one lane: the relocation draws of
// RELOCATE / OVERLAP_DET come from one stream in the reference's order.
__device__ double dev_ranff2(int* idum, int* iv, int* iy, int* idum2)
{
    const int IM1 = 2147483563, IM2 = 2147483399, IMM1 = IM1 - 1;
    const int IA1 = 40014, IA2 = 40692, IQ1 = 53668, IQ2 = 52774, IR1 = 12211, IR2 = 3791;
    const int NDIV = 1 + IMM1 / 32;
    const double AM = 1.0 / IM1, EPS = 1.2e-7, RNMX = 1.0 - EPS;
    int j, k;
    if (*idum <= 0) {
        *idum = (-*idum > 1) ? -*idum : 1;
        *idum2 = *idum;
        // 8 warm-up draws, then iv[31] .. iv[0]: the loop j = 40..1 with a store at
        // j <= 32, split so that no store index is ever outside the table
        for (j = 0; j < 8; ++j) {
            k = *idum / IQ1;
            *idum = IA1 * (*idum - k * IQ1) - k * IR1;
            if (*idum < 0) *idum += IM1;
        }
        for (j = 31; j >= 0; --j) {
            k = *idum / IQ1;
            *idum = IA1 * (*idum - k * IQ1) - k * IR1;
            if (*idum < 0) *idum += IM1;
            iv[j] = *idum;
        }
        *iy = iv[0];
    }
    k = *idum / IQ1;
    *idum = IA1 * (*idum - k * IQ1) - k * IR1;
    if (*idum < 0) *idum += IM1;
    k = *idum2 / IQ2;
    *idum2 = IA2 * (*idum2 - k * IQ2) - k * IR2;
    if (*idum2 < 0) *idum2 += IM2;
    j = 1 + *iy / NDIV;
    *iy = iv[j - 1] - *idum2;
    iv[j - 1] = *idum;
    if (*iy < 1) *iy += IMM1;
    const double r = AM * (double)(*iy);
    return r < RNMX ? r : RNMX;
}

// state layout: [0] idum, [1] iy, [2] idum2, [3..34] iv
// At most cap draws are made: a removal list that needs more sets *err (the host then
// reports a capacity error) and is cut short.
__global__ void k_draws(int* st, double* draws, const int* total, int cap, int* err)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int idum = st[0], iy = st[1], idum2 = st[2];
    // shuffle table in LDS, not in a private array that the compiler turns into
    // indexed register moves
    __shared__ int iv[32];
    for (int q = 0; q < 32; ++q) iv[q] = st[3 + q];
    int nd = 3 * (*total);
    if (nd > cap) {
        atomicOr(err, 1);
        nd = cap - cap % 3;
    }
    for (int q = 0; q < nd; ++q) draws[q] = dev_ranff2(&idum, iv, &iy, &idum2);
    st[0] = idum; st[1] = iy; st[2] = idum2;
    for (int q = 0; q < 32; ++q) st[3 + q] = iv[q];
}

// A droplet that appears m times in the sorted removal list consumes 3m draws at
// offset off; the last triple is the position it keeps (RELOCATE, 1994-2006).
__global__ void k_apply_reloc(double* pos, const int* mult, const int* off, const double* draws, int n, int cap, double pi2,
                              long long draw_cap, int wrap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = mult[i];
    if (m <= 0) return;
    if (3 * ((long long)off[i] + m) > draw_cap) return;      // cut short by k_draws (capacity error)
    const double* d = draws + 3 * (size_t)(off[i] + m - 1);
    double y[3] = {d[0] * pi2, d[1] * pi2 + 0.0 * pi2, d[2] * pi2 + 0.0 * pi2};
    if (wrap)
        for (int c = 0; c < 3; ++c) y[c] = wrap_periodic(y[c], pi2);
    pos[i] = y[0];
    pos[(size_t)cap + i] = y[1];
    pos[2 * (size_t)cap + i] = y[2];
}

__global__ void k_set_rng(int* st, int seed)
{
    if (threadIdx.x != 0) return;
    st[0] = seed; st[1] = 0; st[2] = 0;
    for (int q = 0; q < 32; ++q) st[3 + q] = 0;
}

int part_set_rng(const PartBuffers& pb, int iseedr, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_rng, dim3(1), dim3(1), 0, s, pb.rng_state, iseedr);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// relocate every droplet with mult > 0; returns the number of draws triples via *total_host
int relocate_by_mult(PartBuffers& pb, const PartConst& pc, int* total_host, hipStream_t s)
{
    int st = exclusive_scan(pb.mult, pb.draw_off, pc.n, (int*)pb.scan_tmp, s);
    if (st) return st;
    int total = 0;
    DDNS_HIP(hipMemcpyAsync(&total, pb.draw_off + pc.n, sizeof(int), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    *total_host = total;
    if (total == 0) return 0;
    invalidate_cells(pb);                 // relocated droplets move
    if (3 * total > pb.draw_cap) {
        DDNS_HIP(hipFree(pb.draws));
        pb.draw_cap = 3 * total + 1024;
        DDNS_HIP(hipMalloc(&pb.draws, (size_t)pb.draw_cap * sizeof(double)));
    }
    hipLaunchKernelGGL(k_draws, dim3(1), dim3(1), 0, s, pb.rng_state, pb.draws, pb.draw_off + pc.n, pb.draw_cap,
                       pb.flags + 48);
    hipLaunchKernelGGL(k_apply_reloc, dim3(nblk(pc.n)), dim3(PT), 0, s, pb.pos, pb.mult, pb.draw_off, pb.draws,
                       pc.n, pb.cap, pc.pi2, (long long)pb.draw_cap, 0);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// The relocation of the collision step (RELOCATE, partlib.f90:1984-2012) without a host
// round trip: the draws are made for whatever the removal list holds, into a buffer of
// three draws per droplet slot (each droplet relocated once; a step whose removal list
// needs more sets the capacity flag flags[48], which the step reports as an error).
// Droplets with no removal entry keep their positions.
int relocate_async(PartBuffers& pb, const PartConst& pc, hipStream_t s, bool wrap)
{
    int st = exclusive_scan(pb.mult, pb.draw_off, pc.n, (int*)pb.scan_tmp, s);
    if (st) return st;
    invalidate_cells(pb);                 // relocated droplets move
    const int need = 3 * std::max(pb.cap, 1 << 16);
    if (pb.draw_cap < need) {
        DDNS_HIP(hipFree(pb.draws));
        pb.draw_cap = need;
        DDNS_HIP(hipMalloc(&pb.draws, (size_t)pb.draw_cap * sizeof(double)));
    }
    hipLaunchKernelGGL(k_draws, dim3(1), dim3(1), 0, s, pb.rng_state, pb.draws, pb.draw_off + pc.n, pb.draw_cap,
                       pb.flags + 48);
    hipLaunchKernelGGL(k_apply_reloc, dim3(nblk(pc.n)), dim3(PT), 0, s, pb.pos, pb.mult, pb.draw_off, pb.draws,
                       pc.n, pb.cap, pc.pi2, (long long)pb.draw_cap, wrap ? 1 : 0);
    DDNS_HIP(hipGetLastError());
    return 0;
}

int part_update_and_wrap(PartBuffers& pb, const PartConst& pc, bool relocate, hipStream_t s)
{
    if (pc.n == 0) return 0;
    hipLaunchKernelGGL(k_update_pos, dim3(nblk(pc.n)), dim3(PT), 0, s, pb.pos, pb.dmove, pc.n, pb.cap, pc.pi2);
    DDNS_HIP(hipGetLastError());
    // fvp(2) = fvp(1), fvp(1) = fvp(0); fdp(3..1) = fdp(2..0): level t+1 takes level t's columns
    pb.hv_v = (pb.hv_v + 2) % 3;
    pb.hv_d = (pb.hv_d + 3) % 4;
    pb.v0_stale = true;
    pb.d0_stale = true;
    if (relocate) {
        int st = relocate_async(pb, pc, s, true);
        if (st) return st;
    }
    return 0;
}

// ------------------------------------------------------------------ coarse cell grid
// A periodic grid of G^3 cells of width hc = 2pi/G >= the largest pair distance the
// caller needs, so a 27-cell neighbourhood holds every candidate.  Counting sort by
// cell, then each cell's slice is ordered by slot number (deterministic).
// cell of droplet i and its place among the droplets of that cell (the order in which
// the atomics were served; k_cell_order makes the final order deterministic)
__global__ void k_cell_key(const double* pos, int* cell_of, int* rank, int* cnt, int n, int cap, int G, double hc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int q = (int)(pos[(size_t)c * cap + i] / hc);
        q = q < 0 ? 0 : (q >= G ? G - 1 : q);
        c3[c] = q;
    }
    const int cell = cell_code(c3[0], c3[1], c3[2]);
    cell_of[i] = cell;
    rank[i] = atomicAdd(cnt + cell, 1);
}

// Droplet i (read in slot order, so the reads coalesce) takes the next free place of
// its cell and writes its index and its packed (x, y, z, radius) record there: the
// pair searches then read each candidate cell as one contiguous run.
__global__ void k_cell_fill(const int* __restrict__ cell_of, const int* __restrict__ rank, const int* __restrict__ start,
                            int* __restrict__ sorted, const double* __restrict__ pos, const double* __restrict__ rad,
                            double4* __restrict__ pk, int n, int cap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell_of[i];
    const int p = start[c] + rank[i];
    sorted[p] = i;
    pk[p] = make_double4(pos[i], pos[(size_t)cap + i], pos[2 * (size_t)cap + i], rad[i]);
}

// the order inside a cell is by droplet index, so it does not depend on the order in
// which the fill's atomics were served; the packed records move with their indices
__global__ void k_cell_order(const int* __restrict__ start, int* __restrict__ sorted, double4* __restrict__ pk, int ncell)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int b = start[c], e = start[c + 1];
    for (int i = b + 1; i < e; ++i) {
        const int v = sorted[i];
        const double4 r = pk[i];
        int j = i - 1;
        while (j >= b && sorted[j] > v) { sorted[j + 1] = sorted[j]; pk[j + 1] = pk[j]; --j; }
        sorted[j + 1] = v;
        pk[j + 1] = r;
    }
}

void invalidate_cells(PartBuffers& pb)
{
    pb.cells_valid = false;
    pb.spare.valid = false;
}

static void swap_cells(PartBuffers& pb)
{
    PartBuffers::CellSet& o = pb.spare;
    std::swap(pb.G, o.G);
    std::swap(pb.hc, o.hc);
    std::swap(pb.cell_count, o.cell_count);
    std::swap(pb.cell_count_zero, o.count_zero);
    std::swap(pb.cell_start, o.cell_start);
    std::swap(pb.cell_of, o.cell_of);
    std::swap(pb.sorted, o.sorted);
    std::swap(pb.pk, o.pk);
    std::swap(pb.cells_valid, o.valid);
}

static int alloc_spare(PartBuffers& pb)
{
    PartBuffers::CellSet& o = pb.spare;
    if (o.cell_count) return 0;
    const long long nmax = (long long)CELL_GMAX * CELL_GMAX * CELL_GMAX + 1;
    const size_t c = (size_t)pb.cap;
    DDNS_HIP(hipMalloc(&o.cell_count, nmax * sizeof(int)));
    DDNS_HIP(hipMalloc(&o.cell_start, nmax * sizeof(int)));
    DDNS_HIP(hipMalloc(&o.cell_of, c * sizeof(int)));
    DDNS_HIP(hipMalloc(&o.sorted, c * sizeof(int)));
    DDNS_HIP(hipMalloc(&o.pk, c * sizeof(double4)));
    o.valid = false;
    return 0;
}

int part_build_cells(PartBuffers& pb, const PartConst& pc, double hmin, hipStream_t s)
{
    constexpr int GMAX = CELL_GMAX;
    int G = (int)floor(pc.pi2 / hmin);
    if (G > GMAX) G = GMAX;
    if (G < 3) { set_error("domain too small for the pair search grid (need 3 cells of the interaction radius)"); return DDNS_ERR_DOMAIN; }
    const long long ncell = 1LL << (3 * cell_bits(G));      // Morton numbers of the G^3 cells (G <= 256)
    if (!pb.cell_count) {
        // sized once for the finest grid, so switching between the collision and the
        // overlap grids never reallocates
        const long long nmax = (long long)GMAX * GMAX * GMAX + 1;
        DDNS_HIP(hipMalloc(&pb.cell_count, nmax * sizeof(int)));
        DDNS_HIP(hipMalloc(&pb.cell_start, nmax * sizeof(int)));
        const size_t need = scan_tmp_ints(std::max(nmax, (long long)pb.cap + 1)) * sizeof(int);
        if (need > pb.scan_tmp_bytes) {
            (void)hipFree(pb.scan_tmp);
            DDNS_HIP(hipMalloc(&pb.scan_tmp, need));
            pb.scan_tmp_bytes = need;
        }
    }
    if (pb.cells_valid && pb.G == G) return 0;                     // already built on these positions
    if (pb.spare.valid && pb.spare.G == G) {                      // kept from an earlier build
        swap_cells(pb);
        return 0;
    }
    if (pb.cells_valid) {
        // keep the current grid (another width) for a later search: build into the spare
        int st = alloc_spare(pb);
        if (st) return st;
        swap_cells(pb);
    }
    pb.cells_valid = false;
    pb.G = G;
    pb.hc = pc.pi2 / G;
    // the counts are left zero by the scan of the last build into this buffer
    if (pb.cell_count_zero < ncell + 1) {
        DDNS_HIP(hipMemsetAsync(pb.cell_count, 0, (ncell + 1) * sizeof(int), s));
        pb.cell_count_zero = ncell + 1;
    }
    const long long zero = pb.cell_count_zero;
    pb.cell_count_zero = 0;
    if (pc.n > 0)
        hipLaunchKernelGGL(k_cell_key, dim3(nblk(pc.n)), dim3(PT), 0, s, pb.pos, pb.cell_of, pb.cell_rank, pb.cell_count, pc.n,
                           pb.cap, G, pb.hc);
    int st = exclusive_scan(pb.cell_count, pb.cell_start, (int)ncell, (int*)pb.scan_tmp, s, true);
    if (st) return st;
    pb.cell_count_zero = zero;
    if (pc.n > 0) {
        hipLaunchKernelGGL(k_cell_fill, dim3(nblk(pc.n)), dim3(PT), 0, s, pb.cell_of, pb.cell_rank, pb.cell_start,
                           pb.sorted, pb.pos, pb.rad, pb.pk, pc.n, pb.cap);
        hipLaunchKernelGGL(k_cell_order, dim3(nblk(ncell)), dim3(PT), 0, s, pb.cell_start, pb.sorted, pb.pk, (int)ncell);
    }
    DDNS_HIP(hipGetLastError());
    pb.cells_valid = true;
    return 0;
}

}  // namespace ddns
