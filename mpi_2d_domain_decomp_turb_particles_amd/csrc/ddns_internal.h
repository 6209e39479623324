// ddns_internal.h -- internal declarations shared by the HIP translation units of
// libddns.so: precision traits, complex helpers, device buffers of a handle and the
// host-side launch wrappers of every kernel family.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ddns.h"

#define DDNS_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) return ddns::hip_fail(e_, #call, __FILE__, __LINE__); \
    } while (0)

namespace ddns {

// One detected collision pair on the device, in 64-bit words:
//   [0] canonical key (min gid << 32 | max gid << 2 | type)   [1] ih1 slot   [2] ih2 slot
//   [3] packed cell offsets   [4] ih2 image direction (0: the droplet itself, 5..8:
//   HEADLIST halo directions)   [5] radius of ih1, [6] radius of ih2 (bit patterns)
//   [7] the reference's visiting order of the pair (ih1, cell loop, linked list)
constexpr int DDNS_PAIR_REC = 8;

// One pair in contact (rcolla < d <= rcollb), in doubles: [0] ih1 slot and [1] the
// cell-loop / linked-list order of ih2 (bit patterns of 64-bit integers), then the
// reference's buffer(1:12, ij): index1, index2, va(3), vb(3), unit offset (3), Wr
constexpr int DDNS_CONTACT_REC = 14;

// Coarse-cell number of cell (x, y, z), 0 <= x, y, z < 256: the Morton (Z-curve)
// interleave of the three indices, x in the lowest bit.  Droplets sorted by it form
// compact 3-D blocks, so the lanes of a wavefront touch overlapping field stencils and
// neighbouring cells.  The numbers of a G^3 grid lie below 8^cell_bits(G).
__host__ __device__ __forceinline__ unsigned cell_spread(unsigned v)
{
    v &= 0xffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__host__ __device__ __forceinline__ int cell_code(int x, int y, int z)
{
    return (int)(cell_spread((unsigned)x) | (cell_spread((unsigned)y) << 1) | (cell_spread((unsigned)z) << 2));
}
// Work-group b of a launch runs on XCD b % 8, and each of the 8 XCDs has its own L2.
// This gives the work-groups of one XCD a contiguous range of the nb logical blocks,
// so droplets that are close in cell order, which read the same field nodes and
// neighbour cells, share an L2.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_chunk(int b, int nb)
{
    const int x = b & 7, r = b >> 3, q = nb >> 3, m = nb & 7;
    return x * q + (x < m ? x : m) + r;
}

// Offsets (-1 or 0 up to 0 or +1) of the neighbour cells to visit along each axis.
struct CellSpan {
    int lo[3] = {-1, -1, -1};
    int hi[3] = {1, 1, 1};
};

// The neighbour cells that can hold a point within distance reach of x, which lies in
// cell cc of width hc: along each axis the lower (upper) neighbour is needed only when
// x is closer than reach to the lower (upper) face.  The margin of 1e-9 hc covers the
// rounding of the cell index; clamped and periodic cells stay covered because the
// test only ever adds cells.
__device__ __forceinline__ CellSpan reach_span(const double x[3], const int cc[3], double hc, double reach)
{
    const double m = reach + 1e-9 * hc;
    CellSpan s;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        s.lo[c] = (x[c] - (double)cc[c] * hc < m) ? -1 : 0;
        s.hi[c] = ((double)(cc[c] + 1) * hc - x[c] < m) ? 1 : 0;
    }
    return s;
}

// The candidates of the 27 neighbour cells (or of the part of them that span selects)
// in 9 rows: for each (y, z) offset the cells x-1, x, x+1 as one sequence.  One row's
// three ranges are held at a time: the next row's ranges load while the current row runs,
// and inside a row the next record loads while f works on the current one.  Holding all
// 27 ranges (and unrolling the nine rows) took about 50 more registers, and a walk that
// waits on about one record load per candidate gains more from the wavefronts those
// registers allow per SIMD.  f(p, kc, rec, ib) gets the sorted position, the cell number
// k = 9 (dx+1) + 3 (dy+1) + (dz+1) of neighbour_shift, the record pk[p] and, when sorted
// is given, the droplet index sorted[p].
template <typename F>
__device__ __forceinline__ void for_each_candidate(const int* __restrict__ cell_start, int G, const int cc[3],
                                                   const double4* __restrict__ pk, const int* __restrict__ sorted,
                                                   F&& f, const CellSpan& span = CellSpan())
{
    auto wrap = [G](int q) { return q < 0 ? q + G : (q >= G ? q - G : q); };
    // [s, e) of the cells x-1, x, x+1 of row r = 3 my + mz; empty where span leaves it out
    auto load_row = [&](int r, int& s0, int& e0, int& s1, int& e1, int& s2, int& e2) {
        const int my = r / 3, mz = r - 3 * my;
        s0 = e0 = s1 = e1 = s2 = e2 = 0;
        if (my - 1 < span.lo[1] || my - 1 > span.hi[1] || mz - 1 < span.lo[2] || mz - 1 > span.hi[2]) return;
        const int qy = wrap(cc[1] + my - 1), qz = wrap(cc[2] + mz - 1);
        if (span.lo[0] < 0) {
            const int c = cell_code(wrap(cc[0] - 1), qy, qz);
            s0 = cell_start[c]; e0 = cell_start[c + 1];
        }
        {
            const int c = cell_code(cc[0], qy, qz);
            s1 = cell_start[c]; e1 = cell_start[c + 1];
        }
        if (span.hi[0] > 0) {
            const int c = cell_code(wrap(cc[0] + 1), qy, qz);
            s2 = cell_start[c]; e2 = cell_start[c + 1];
        }
    };
    int s0, e0, s1, e1, s2, e2;
    load_row(0, s0, e0, s1, e1, s2, e2);
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
        int n0 = 0, m0 = 0, n1 = 0, m1 = 0, n2 = 0, m2 = 0;
        if (r + 1 < 9) load_row(r + 1, n0, m0, n1, m1, n2, m2);
        int m = 0, p = s0, end = e0;
        if (p >= end) { m = 1; p = s1; end = e1; }
        if (p >= end) { m = 2; p = s2; end = e2; }
        if (p < end) {
            double4 next = pk[p];
            int inext = sorted ? sorted[p] : 0;
#pragma unroll 1
            while (true) {
                const int pc = p, kc = 9 * m + r;
                const double4 rec = next;
                const int ib = inext;
                ++p;
                while (p >= end && m < 2) {
                    ++m;
                    p = m == 1 ? s1 : s2;
                    end = m == 1 ? e1 : e2;
                }
                const bool more = p < end;
                if (more) {
                    next = pk[p];
                    if (sorted) inext = sorted[p];
                }
                f(pc, kc, rec, ib);
                if (!more) break;
            }
        }
        s0 = n0; e0 = m0; s1 = n1; e1 = m1; s2 = n2; e2 = m2;
    }
}

// periodic shift (-1, 0, +1) of neighbour k along each axis for a lane in cell cc
__device__ __forceinline__ void neighbour_shift(int k, int G, const int cc[3], int* wx, int* wy, int* wz)
{
    const int qx = cc[0] + k / 9 - 1, qy = cc[1] + (k / 3) % 3 - 1, qz = cc[2] + k % 3 - 1;
    *wx = qx < 0 ? -1 : (qx >= G ? 1 : 0);
    *wy = qy < 0 ? -1 : (qy >= G ? 1 : 0);
    *wz = qz < 0 ? -1 : (qz >= G ? 1 : 0);
}

inline int cell_bits(int G)
{
    int b = 0;
    while ((1 << b) < G) ++b;
    return b;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line);
struct PartBuffers;
// droplet positions or slots changed: neither cell grid describes them any more
void invalidate_cells(PartBuffers& pb);
void set_error(const std::string& msg);

// ---------------------------------------------------------------- complex helpers
template <typename T> struct cpx_t;
template <> struct cpx_t<float>  { using type = float2; };
template <> struct cpx_t<double> { using type = double2; };
template <typename T> using cpx = typename cpx_t<T>::type;

template <typename C> __host__ __device__ __forceinline__ C mk(decltype(C{}.x) a, decltype(C{}.x) b) { C r; r.x = a; r.y = b; return r; }
template <typename C> __host__ __device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __host__ __device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __host__ __device__ __forceinline__ C cmul(C a, C b) { return mk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename C> __host__ __device__ __forceinline__ C cconj(C a) { return mk<C>(a.x, -a.y); }
template <typename C> __host__ __device__ __forceinline__ C cscale(C a, decltype(C{}.x) s) { return mk<C>(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <typename C, bool INV> __host__ __device__ __forceinline__ C cmul_mi(C a) { return INV ? mk<C>(-a.y, a.x) : mk<C>(a.y, -a.x); }

// ---------------------------------------------------------------- flow layout
// Spectral and intermediate fields: complex arrays [s][r][kx] with kx fastest,
// kx in [0, N/2) (the kx = N/2 plane is identically zero and not stored), r the
// y-or-ky index and s the z-or-kz index.  One field = N*N*(N/2) complex values.
// Physical velocity for the droplets: real [z][y][x][3] interleaved.
struct FlowBuffers {
    void* U[3] = {nullptr, nullptr, nullptr};  // state u-hat [kz][ky][kx]
    void* O[3] = {nullptr, nullptr, nullptr};  // AB2 history [kz][ky][kx]
    void* A[6] = {};                            // inverse pipeline: u(3), w(3)
    void* B[3] = {};                            // forward pipeline: (u x w)
    void* uphys = nullptr;                      // [z][y][x][3]: u, v, w
    void* tw = nullptr;                         // N twiddles exp(-2 pi i m / N)
    void* twh = nullptr;                        // N/2 twiddles for the half-length x FFT
    double* low = nullptr;                      // pre-kernel scratch (1024): ff[2], energies, forcing table, DFT sums
    int* sth_state = nullptr;                   // ranff state on the device: iseedf, iyf, ivf[32]
    double* sth_b = nullptr;                    // b1r..b3r (450)
    void* dg[8] = {};                           // diagnostics scratch, 8 fields, allocated on first use
    double* dg_part = nullptr;                  // diagnostics reduction partials
    double* spec_part = nullptr;                // E(k): SPEC_BLOCKS partial rows of SPEC_NEK_MAX shells, then the result
    int ky_tc = 0;                              // y-stage tile width chosen by flow_tune (0: the default)
    float ky_ms[3] = {0, 0, 0};                 // the tuner's time per y stage pair of each candidate (0: not valid)
    int kz_nt = 0;                              // threads of the fused z stage chosen by flow_tune (0: the default)
    float kz_ms[3] = {0, 0, 0};                 // the tuner's time per fused z stage of each candidate (0: not valid)
    // Independent realisations held side by side (a batch): every field above holds R
    // realisations, realisation r at r * rstride complex values (uphys: r * pstride reals,
    // low: r * 1024 doubles, sth_state: r * 64 ints, sth_b: r * 512 doubles), and every
    // flow launch covers all of them.  R = 1, rstride = pstride = 0 for one realisation.
    int R = 1;
    size_t rstride = 0, pstride = 0;
};
// The buffers of realisation r of a batch as a one-realisation FlowBuffers (pointers
// into the batch's arrays; nothing owned), for the transfers and diagnostics.
FlowBuffers flow_view(const FlowBuffers& fb, int r, int precision);

// E(k) shell sums (spectrum.hip): work-groups of the partial-sum kernel and the most
// shells any grid has (nek = 510 at N = 1024)
constexpr int SPEC_BLOCKS = 256;
constexpr int SPEC_NEK_MAX = 512;

struct FlowConst {
    int N;
    int nek;
    int forcing;
    double dt, rnu;
    double force0, force1;
    double ou_tf, ou_var;
    double pi2;
};

// launch wrappers (flow.hip)
int flow_alloc(FlowBuffers& fb, int N, int precision, int R = 1);
// the forcing kind and amplitudes of realisation r into its low-mode table (batches run
// with FlowConst::forcing = -1 and read them there)
int flow_set_realisation(const FlowBuffers& fb, int r, int forcing, double force0, double force1, double ou_tf,
                         double ou_var);
void flow_free(FlowBuffers& fb);
int flow_prologue(const FlowBuffers& fb, const FlowConst& fc, int precision, hipStream_t s);
// Parts of the flow step: the inverse stages (first half), the whole forward half, or the
// forward half in two pieces (forward y + pre, then z + sym) so that a wait can sit
// between them.  capturing: the stream is being captured into a graph (the events
// become graph nodes).
enum FlowPart { FLOW_INV = 0, FLOW_FWD = 1, FLOW_FWD_Y = 2, FLOW_FWD_Z = 3 };
int flow_step(const FlowBuffers& fb, const FlowConst& fc, int precision, int part, hipStream_t s,
              hipEvent_t* ev /* FLOW_EVENTS events or nullptr */, bool capturing = false);
constexpr int FLOW_EVENTS = 9;
int flow_upload_state(const FlowBuffers& fb, int N, int precision, const double* const src[6], hipStream_t s);
int flow_download_state(const FlowBuffers& fb, int N, int precision, double* const dst[6], hipStream_t s);
int flow_download_phys(const FlowBuffers& fb, int N, int precision, double* ux, double* uy, double* uz, hipStream_t s);
// POSTPROC / TRANSPOSTPROC sums of the current u-hat.  want: 1 = CFL maximum and the
// velocity-derivative moments, 2 = divergence / vorticity sums.  ek, dk, tk (nek each,
// any may be null): energy, dissipation and transfer spectra.  mom: [0] max |u|+|v|+|w|,
// [1..3] sums of d^2, d^3, d^4 over the three longitudinal derivatives, [4..5] sums of
// div, div^2, [6..8] sums of |w|, |w|^2, |w| div
int flow_init(FlowBuffers& fb, int N, int nek, int precision, int kind, int seed, double k0, double urms, double force0,
              hipStream_t s);
int flow_diagnostics(FlowBuffers& fb, int N, int nek, double rnu, int precision, int want, double* ek, double* dk,
                     double* tk, double* mom, hipStream_t s);
int flow_spectrum(const FlowBuffers& fb, int N, int nek, int precision, double* ek, hipStream_t s);
// allocation-free E(k) (spectrum.hip): the partial rows in a file-scope device array whose
// address is resolved once per device by spectrum_scratch_init
int spectrum_scratch_init();
int flow_spectrum_device(const FlowBuffers& fb, int N, int nek, int precision, int dev, double* ek_dev, hipStream_t s);
int fft_r2c_host(const FlowBuffers& fb, int N, int precision, double* v, hipStream_t s);
int fft_c2r_host(const FlowBuffers& fb, int N, int precision, double* v, hipStream_t s);
int flow_profile(const FlowBuffers& fb, const FlowConst& fc, int precision, int k, double* ms, int nk, hipStream_t s);
// times the candidate y-stage tile widths and fused z-stage work-group sizes on fb's
// buffers (before any state is loaded) and keeps the fastest in fb.ky_tc, fb.kz_nt
// (cached per grid and precision); the defaults of a grid
int flow_tune(FlowBuffers& fb, const FlowConst& fc, int precision, hipStream_t s);
int flow_default_ky_tc(int N, int precision);
int flow_default_kz_nt(int N, int precision);
int flow_profile_copy(const FlowBuffers& fb, int N, int precision, int reps, double* ms, double* bytes, hipStream_t s);

// ---------------------------------------------------------------- droplets
struct PartBuffers {
    int cap = 0;
    // state, SoA FP64, slot order = the reference's droplet index order
    double* pos = nullptr;      // 3*cap
    double* vel = nullptr;      // 3*cap   yp(4:6)
    double* rad = nullptr;      // cap     yp(7)
    double* gid = nullptr;      // cap     yp(8)
    double* fvp = nullptr;      // 9*cap   fvp(:,0:2) as [t][c][cap]
    double* fdp = nullptr;      // 12*cap  fdp(:,0:3) as [t][c][cap]
    double* dmove = nullptr;    // 3*cap
    double* vp = nullptr;       // 3*cap
    double* pert = nullptr;     // 3*cap   pertvel(3, n), droplet-major as the reference stores it, so one
                                //         droplet's three components share a cache line
    // cell grid
    int G = 0;                  // coarse cells per dimension
    double hc = 0;
    int* cell_count = nullptr;  // G^3 + 1
    int* cell_start = nullptr;  // G^3 + 1
    int* cell_of = nullptr;     // cap
    int* sorted = nullptr;      // cap
    double4* pk = nullptr;      // cap: (x, y, z, radius) of sorted[p], refreshed with the cells
    int* cell_rank = nullptr;   // cap: place of a droplet among those of its cell (build scratch)
    bool cells_valid = false;   // the grid above was built on the current positions
    hipEvent_t sync_ev = nullptr;   // waits of the droplet block (wait_stream)
    double* gm_host = nullptr;      // pinned: GMRES's small-problem head [0..5] and the list flags [8..11]
    // History levels as rings: level t of fvp (3 levels) is column block (t + hv_v) % 3,
    // of fdp (4 levels) (t + hv_d) % 4, so the step's history shift (level t+1 = level t)
    // only turns the rings.  After a shift level 0 holds the oldest level until it is
    // written again (fvp by the history step, fdp by the advance); until then it reads as
    // level 1, which the shift copied it to (v0_stale, d0_stale).
    int hv_v = 0, hv_d = 0;
    bool v0_stale = false, d0_stale = false;
    int vcol(int t) const { return 3 * ((t + hv_v) % 3); }   // first column of fvp level t
    int dcol(int t) const { return 3 * ((t + hv_d) % 4); }   // first column of fdp level t
    long long cell_count_zero = 0;  // cell_count[0 .. this) is zero (the last build's scan clears it)
    // the step's first overlap sweep is queued and its count not yet read; the validity
    // of the grids before its relocation, restored when it relocated nothing
    bool overlap_queued = false;
    bool overlap_cells_kept = false, overlap_spare_kept = false;
    // a second grid, kept from the last build of another width while it stays valid:
    // a search at that width then swaps it in instead of rebuilding
    struct CellSet {
        int G = 0;
        double hc = 0;
        int *cell_count = nullptr, *cell_start = nullptr, *cell_of = nullptr, *sorted = nullptr;
        double4* pk = nullptr;
        bool valid = false;
        long long count_zero = 0;
    } spare;
    void* scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    // HI, in cell-sorted droplet order: ELL neighbour lists, slot-major
    int* nei_cnt = nullptr;     // cap + 1
    int* nei_idx = nullptr;     // ell_width * cap: sorted position of neighbour s of row t at s * n + t
    void* nei_m = nullptr;      // 6 * ell_width * cap (float or double): entry c at (6 s + c) * n + t
    int ell_width = 0;
    void* gm_Q = nullptr;       // (maxiter+2) * 3 cap: Krylov basis, then two work vectors
    void* gm_v = nullptr;       // 3 cap
    double* gm_x0 = nullptr;    // 3 cap
    double* gm_b = nullptr;     // 3 cap
    double* gm_w = nullptr;     // 3 cap
    double* gm_red = nullptr;   // reduction partials
    double* gm_small = nullptr; // Hessenberg etc
    int gm_last_iters = 0;      // GMRES iterations of the last solve: the size of the next first batch
    // collisions
    int* mult = nullptr;        // cap: removal-list multiplicity
    void* cand = nullptr;       // cand_cap int2: pairs handed from the collision search to the pair kernel
    int cand_cap = 0;
    int last_ncand = 0;         // the last search's count: the size of the next pair-kernel grid
    int* hit = nullptr;         // cap: droplet already counted as ih1 of a collision
    int* flags = nullptr;       // device flags / counters (64 ints)
    long long* tested = nullptr;
    int* nirbin = nullptr;
    double* bins = nullptr;     // 7 * nset*nset*NBIN
    int* counts = nullptr;      // icc[3], colltype[100], nir[100]
    long long* pairs = nullptr; // pair_cap * DDNS_PAIR_REC: one record per detected pair
    int pair_cap = 0;
    double* contact = nullptr;  // contact_cap * DDNS_CONTACT_REC: one record per pair in contact
    int contact_cap = 0;
    int* found_host = nullptr;  // pinned: candidate and contact counts of the last collision search
    int* counts_host = nullptr; // pinned: the step's counts (512 ints) and tested pairs (4 long long)
    double* draws = nullptr;    // relocation draws
    int draw_cap = 0;
    int* draw_off = nullptr;    // cap + 1
    int* rng_state = nullptr;   // ranff2 state: idum, iy, idum2, iv[32]
    double* red = nullptr;      // generic reduction scratch
};

struct PartConst {
    int N;            // flow grid (reference cell grid: ncdx = N)
    int n;            // droplets
    int nset;
    double radius[DDNS_MAXSET];
    double radmax;
    double dt, rnu, st;
    double Sv_const, tau2;
    double pi2;
    double d_coll, d_hdi, hdi_trunc;
    int nx_coll, nx_max, imax_stat;
    int hdi, nonoverlap, ndrag;
    int new_particle;
    int precision;
};

int part_alloc(PartBuffers& pb, int cap, int nset, int G);
// Waits for event e by polling it: after a blocking wait the host thread ran again tens
// of microseconds after the GPU had finished, and the stream waited that long for the
// launches that follow.
inline hipError_t spin_wait(hipEvent_t e)
{
    for (;;) {
        const hipError_t r = hipEventQuery(e);
        if (r != hipErrorNotReady) return r;
    }
}
// all work queued on s so far, waited for by polling (pb.sync_ev, made on first use)
inline hipError_t wait_stream(PartBuffers& pb, hipStream_t s)
{
    if (!pb.sync_ev) {
        const hipError_t r = hipEventCreateWithFlags(&pb.sync_ev, hipEventDisableTiming);
        if (r != hipSuccess) return r;
    }
    const hipError_t r = hipEventRecord(pb.sync_ev, s);
    return r != hipSuccess ? r : spin_wait(pb.sync_ev);
}
void part_free(PartBuffers& pb);
// with HI on, also writes the HI solve's x0 and w in cell order (after part_history_fvp)
int part_interp(PartBuffers& pb, const PartConst& pc, const void* uphys, hipStream_t s);
int part_history_fvp(PartBuffers& pb, const PartConst& pc, hipStream_t s);
int part_build_cells(PartBuffers& pb, const PartConst& pc, double hmin, hipStream_t s);
int part_hdi(PartBuffers& pb, const PartConst& pc, int maxiter, double tol, int* iters, double* res, hipStream_t s);
// The start of a solve (neighbour lists, right-hand side, first Krylov vector) queued on s.
int part_hdi_begin(PartBuffers& pb, const PartConst& pc, int maxiter, double tol, hipStream_t s);
// The iterations of R started solves at once on stream s (every kernel one launch over
// the R systems), then each solution scattered to its droplets.  iters, res, status (0,
// DDNS_WARN_GMRES, or an error of the neighbour lists) per system; redo[r] = 1 where the
// lists had to be widened: that system is to be solved again on its own.
struct HdiBatch {
    void* table = nullptr;          // device: one descriptor per system
    double* heads = nullptr;        // device: 16 doubles per system
    void* table_host = nullptr;     // pinned staging of the descriptors
    double* heads_host = nullptr;   // pinned copy of the heads
    int cap = 0;
};
int part_hdi_batch(HdiBatch& hb, PartBuffers* const* pbs, const PartConst* const* pcs, int R, int maxiter,
                   int* iters, double* res, int* status, int* redo, hipStream_t s);
void part_hdi_batch_free(HdiBatch& hb);
int part_advance(PartBuffers& pb, const PartConst& pc, hipStream_t s);
int part_collide(PartBuffers& pb, const PartConst& pc, hipStream_t s);
// the collision search queued without a host wait, and its completion (waits for the
// search's counts, then queues the pair kernel)
int part_collide_search(PartBuffers& pb, const PartConst& pc, hipStream_t s);
// part_collide_finish / part_collide: some droplet's reach (its statistics window or
// its collision window) exceeds the cell width, so the walk may have missed partners:
// the caller builds wider cells and searches again
constexpr int DDNS_COLL_REGRID = 2;
int part_collide_finish(PartBuffers& pb, const PartConst& pc, hipStream_t s);
// queue the step's counts into pinned memory; read them after the stream was waited for
int part_queue_counts(PartBuffers& pb, hipStream_t s);
// exclusive scan of n ints into out[0..n-1], the total into out[n] (particles.hip);
// clear_in leaves in[0..n-1] zero (in must not alias out then)
int exclusive_scan(const int* in, int* out, int n, int* tmp, hipStream_t s, bool clear_in = false);
int part_take_counts(const PartBuffers& pb, const PartConst& pc, ddns_collision_counts* cc);
int part_update_and_wrap(PartBuffers& pb, const PartConst& pc, bool relocate, hipStream_t s);
// wrap: the drawn positions get the PERIODIC wrap (the step's wrap follows its relocation)
int relocate_async(PartBuffers& pb, const PartConst& pc, hipStream_t s, bool wrap = false);
int part_overlap(PartBuffers& pb, const PartConst& pc, int* sweeps, int* relocated, hipStream_t s);
// the step's overlap sweep: the first sweep queued without a host wait; the rest, if it
// relocated droplets, after the caller has waited for the stream
// next_width > 0: the grid of that width (the next step's first) is built too, on the
// positions the sweep saw, and kept if the sweep moves nothing
int part_overlap_queue(PartBuffers& pb, const PartConst& pc, hipStream_t s, double next_width = 0.0);
int part_overlap_finish(PartBuffers& pb, const PartConst& pc, int* sweeps, int* relocated, hipStream_t s);
int part_read_counts(const PartBuffers& pb, const PartConst& pc, ddns_collision_counts* cc, hipStream_t s);
int part_read_bins(const PartBuffers& pb, const PartConst& pc, int* nirbin, double* const bins[7], hipStream_t s);
int part_read_pairs(PartBuffers& pb, const PartConst& pc, int64_t* triples, int cap, int* npairs, hipStream_t s);
int part_read_contacts(const PartBuffers& pb, double* rec, int cap, int* n, hipStream_t s);
int part_coalesce(PartBuffers& pb, const PartConst& pc, int* removed, hipStream_t s);
int part_set_rng(const PartBuffers& pb, int iseedr, hipStream_t s);

}  // namespace ddns
