// spectrum.hip -- the energy spectrum E(k) of u-hat (postprocessing.f90:151-170):
// E(i) = sum over the modes with int(|k| + 0.5) = i of |u|^2 + |v|^2 + |w|^2, the
// kx = 0 plane counted half, for the shells i = 1..nek.
//
// The sum is deterministic: every wavefront owns a fixed set of (kz, ky) rows and adds
// its modes in a fixed order, so the result does not depend on scheduling.  Along one
// row the shell number never decreases with kx, so the 64 lanes of a row chunk form
// runs of equal shell number; a segmented inclusive scan over the lanes sums each run,
// and only the last lane of a run adds it to the wavefront's own LDS shell table (no
// two lanes touch one entry).  The four wavefronts' tables are added in order into one
// partial row per work-group, and a second kernel adds the SPEC_BLOCKS partial rows of
// each shell in order.  Products and sums are in double in both precision modes.
//
// ddns_get_spectrum runs them with the partial rows in a buffer of the handle;
// ddns_get_spectrum_device with the partial rows in a file-scope device array.
#include "ddns_internal.h"

#include <mutex>

namespace ddns {

constexpr int SPEC_T = 256;                        // threads per work-group: 4 wavefronts
constexpr int SPEC_W = SPEC_T / 64;

__device__ __forceinline__ int spec_wave_index(int i, int N) { return i <= N / 2 ? i : i - N; }

// Partial shell sums: part[s * SPEC_BLOCKS + b] for shell s + 1 and work-group b.
template <typename T>
__global__ void __launch_bounds__(SPEC_T) k_shell_partials(const void* __restrict__ u0, const void* __restrict__ u1,
                                                           const void* __restrict__ u2, int N, int nek,
                                                           double* __restrict__ part)
{
    using C = cpx<T>;
    __shared__ double acc[SPEC_W][SPEC_NEK_MAX];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < SPEC_W * SPEC_NEK_MAX; i += SPEC_T) (&acc[0][0])[i] = 0.0;
    __syncthreads();
    const int M = N / 2;
    const long long rows = (long long)N * N;
    const C* f[3] = {(const C*)u0, (const C*)u1, (const C*)u2};
    for (long long r = (long long)blockIdx.x * SPEC_W + w; r < rows; r += (long long)gridDim.x * SPEC_W) {
        const int ky = (int)(r % N), kz = (int)(r / N);
        const int kyw = spec_wave_index(ky, N), kzw = spec_wave_index(kz, N);
        const int kyz2 = kyw * kyw + kzw * kzw;
        for (int c0 = 0; c0 < M; c0 += 64) {
            const int kx = c0 + lane;
            double e = 0.0;
            int sh = nek + 1;                     // beyond the last shell: dropped
            if (kx < M) {
                double k2 = (double)(kx * kx + kyz2);
                if (k2 == 0.0) k2 = 1.e-5;
                sh = (int)(sqrt(k2) + 0.5);
                const size_t t = (size_t)r * M + kx;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const C a = f[q][t];
                    e += (double)a.x * (double)a.x + (double)a.y * (double)a.y;
                }
                if (kx == 0) e *= 0.5;
                if (sh > nek) sh = nek + 1;
            }
            // segmented inclusive scan: sh is non-decreasing over the lanes
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const double ev = __shfl_up(e, o, 64);
                const int sv = __shfl_up(sh, o, 64);
                if (lane >= o && sv == sh) e += ev;
            }
            const int snext = __shfl_down(sh, 1, 64);
            const bool tail = lane == 63 || snext != sh;
            if (tail && sh >= 1 && sh <= nek) acc[w][sh - 1] += e;
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < nek; s += SPEC_T) {
        double v = acc[0][s];
#pragma unroll
        for (int q = 1; q < SPEC_W; ++q) v += acc[q][s];
        part[(size_t)s * SPEC_BLOCKS + blockIdx.x] = v;
    }
}

// ek[s] = the SPEC_BLOCKS partials of shell s + 1, added lane-strided then by a fixed
// wavefront tree: one wavefront per shell
__global__ void __launch_bounds__(SPEC_T) k_shell_final(const double* __restrict__ part, int nek, double* __restrict__ ek)
{
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * SPEC_W + (threadIdx.x >> 6);
    if (s >= nek) return;
    double v = 0.0;
    for (int b = lane; b < SPEC_BLOCKS; b += 64) v += part[(size_t)s * SPEC_BLOCKS + b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) ek[s] = v;
}

// Both kernels on stream s: part holds SPEC_BLOCKS * nek doubles, ek_dev nek doubles.
static int launch_shells(const FlowBuffers& fb, int N, int nek, int precision, double* part, double* ek_dev,
                         hipStream_t s)
{
    if (precision == DDNS_FP64)
        hipLaunchKernelGGL(k_shell_partials<double>, dim3(SPEC_BLOCKS), dim3(SPEC_T), 0, s, (const void*)fb.U[0],
                           (const void*)fb.U[1], (const void*)fb.U[2], N, nek, part);
    else
        hipLaunchKernelGGL(k_shell_partials<float>, dim3(SPEC_BLOCKS), dim3(SPEC_T), 0, s, (const void*)fb.U[0],
                           (const void*)fb.U[1], (const void*)fb.U[2], N, nek, part);
    DDNS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_shell_final, dim3((unsigned)((nek + SPEC_W - 1) / SPEC_W)), dim3(SPEC_T), 0, s,
                       (const double*)part, nek, ek_dev);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// ddns_get_spectrum: partial rows and the result in the handle's buffer fb.spec_part
// (SPEC_BLOCKS * SPEC_NEK_MAX + SPEC_NEK_MAX doubles), then copied to the host.
int flow_spectrum(const FlowBuffers& fb, int N, int nek, int precision, double* ek, hipStream_t s)
{
    if (nek < 1 || nek > SPEC_NEK_MAX) { set_error("spectrum: nek out of range"); return DDNS_ERR_ARG; }
    double* res = fb.spec_part + (size_t)SPEC_BLOCKS * SPEC_NEK_MAX;
    const int st = launch_shells(fb, N, nek, precision, fb.spec_part, res, s);
    if (st) return st;
    DDNS_HIP(hipMemcpyAsync(ek, res, sizeof(double) * nek, hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    return 0;
}

// ---------------------------------------------------------------- allocation-free variant
// The partial rows of ddns_get_spectrum_device live in this array: SPEC_BLOCKS rows of
// SPEC_NEK_MAX shells, 1 MiB, enough for every grid a handle accepts (nek <= 510 at
// N = 1024).  Each GPU holds its own copy from the first use until the process ends.
// In host code the name is only a stand-in; kernels get the device address that
// hipGetSymbolAddress returns for the current device (g_scratch below).  Every call on
// one GPU shares the array, so calls that overlap on one GPU race.  Nothing needs
// clearing between calls: the partial kernel's SPEC_BLOCKS work-groups write every
// entry (shell s < nek, work-group b) that the final kernel then reads.
constexpr int SCRATCH_DEVICES = 64;
static __device__ __align__(256) double scratch[SPEC_BLOCKS * SPEC_NEK_MAX];
static double* g_scratch[SCRATCH_DEVICES] = {};
static std::mutex g_scratch_mu;

// Resolves the device address of this device's copy of scratch, once per device, with
// the device current.  ddns_create calls it for the device the handle lives on.
int spectrum_scratch_init()
{
    int dev = 0;
    DDNS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_DEVICES) {
        set_error("spectrum scratch: device index beyond 63");
        return DDNS_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[dev]) return 0;
    void* p = nullptr;
    DDNS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        set_error("spectrum scratch: hipGetSymbolAddress gave a null address");
        return DDNS_ERR_HIP;
    }
    g_scratch[dev] = (double*)p;
    return 0;
}

// ddns_get_spectrum_device: the same two kernels as flow_spectrum on stream s, the
// partial rows in scratch and the nek sums straight into the caller's device array
// ek_dev.  No allocation, no copy, no wait.
int flow_spectrum_device(const FlowBuffers& fb, int N, int nek, int precision, int dev, double* ek_dev, hipStream_t s)
{
    if (nek < 1 || nek > SPEC_NEK_MAX) {
        set_error("ddns_get_spectrum_device: nek = " + std::to_string(nek) + " exceeds the static scratch of " +
                  std::to_string(SPEC_NEK_MAX) + " shells");
        return DDNS_ERR_ARG;
    }
    if (dev < 0 || dev >= SCRATCH_DEVICES) {
        set_error("ddns_get_spectrum_device: device index beyond 63");
        return DDNS_ERR_ARG;
    }
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        part = g_scratch[dev];
    }
    if (!part) {
        set_error("ddns_get_spectrum_device: the spectrum scratch of this device was never resolved");
        return DDNS_ERR_HIP;
    }
    return launch_shells(fb, N, nek, precision, part, ek_dev, s);
}

}  // namespace ddns
