// collide.hip -- droplet pair searches: COLLISION_DET (partlib.f90:1166-1846) with
// its RDF / relative-velocity statistics, and the overlap sweep of OVERLAP_DET
// (2047-2167).  Compiled with -ffp-contract=off: every double operation of the
// pair tests is the same IEEE operation, in the same order, as the reference, so
// in FP64 the detected pairs and their types are bit-identical to it.
//
// Search: one lane per droplet scans the 27 neighbouring cells of the coarse grid
// (cell width >= the largest distance that can matter).  For each candidate the
// lane decides whether the reference visits this pair from this droplet and with
// which periodic image, replaying the reference's rules exactly:
//   - the reference scans cells around ih1 and takes only ih2 > ih1 (1325), where
//     halo images of other droplets carry indices above every local droplet;
//   - y/z periodicity comes from images made by HEADLIST within `dist` of a face
//     or edge, shifted by +-2pi (2212-2349); the double-count filter (1309-1310)
//     keeps only the images in the upper z halo and in the upper y halo inside the
//     z range, i.e. image shifts (sy,sz) in {(+1,0),(+1,+1),(0,+1),(-1,+1)};
//   - x periodicity is a cell wrap with xb = x + sx*2pi (1312-1321, 1330);
//   - the cell indices, windows and clamps are the reference's (1288-1306, 2407-2429).
// The droplet that plays ih1 is therefore the one that sees its partner through an
// upper image, or the lower slot number for a pair without y/z image.
#include <algorithm>
#include <cstring>

#include "ddns_internal.h"

namespace ddns {

constexpr int CT = 128;

struct PairGeom {
    int N;            // reference cell grid ncdx = ncdy = ncdz
    double h;         // hx = hy = hz = 2 pi
    int imax;         // scan half-width (d_stat or d_overl cells)
    int range;        // |iy2|, |iz2| admissible halo range (ny_coll or jmax)
    int nx_max;       // HEADLIST clamp
    double dist;      // HEADLIST image distance
    int G;            // coarse grid
    double hc;
};

__device__ __forceinline__ int ref_cell_local(double x, double h, int N)
{
    int c = 1 + (int)(x / h * (double)N);             // 1288-1294
    return c > N ? N : c;
}

__device__ __forceinline__ int ref_cell_halo(double y, double h, int N, int nmax)
{
    const double d = y - 0.0 * h;                      // 2409-2429
    int c = d > 0 ? 1 + (int)(d / h * (double)N) : -(int)(fabs(d) / h * (double)N);
    if (c > N + nmax) c = N + nmax;
    if (c < 1 - nmax) c = 1 - nmax;
    return c;
}

// Does the reference evaluate (ih1 = a, ih2 = image of b with shifts (wx,wy,wz))?
// On success returns the partner coordinates and the reference cell offsets.
__device__ __forceinline__ bool ref_visit(const PairGeom& g, int ia, int ib, double xa, double ya, double za,
                                         double xb0, double yb0, double zb0, int wx, int wy, int wz,
                                         double* xb, double* yb, double* zb, int* di, int* dj, int* dk)
{
    const double h = g.h;
    const bool image = (wy != 0 || wz != 0);
    if (image) {
        if (!(wz == 1 || (wz == 0 && wy == 1))) return false;
        const double d_e = fabs(1.0 * h - yb0), d_w = fabs(0.0 * h - yb0);
        const double d_s = fabs(0.0 * h - zb0);
        bool in;
        if (wy == 1 && wz == 0) in = d_w <= g.dist;                              // direction 5
        else if (wy == 1 && wz == 1) in = sqrt(d_w * d_w + d_s * d_s) <= g.dist;  // 6
        else if (wy == 0 && wz == 1) in = d_s <= g.dist;                          // 7
        else in = sqrt(d_e * d_e + d_s * d_s) <= g.dist;                          // 8
        if (!in) return false;
    } else if (ib <= ia) {
        return false;
    }
    const double pi2 = h;
    const double ys = wy == 0 ? yb0 : (wy > 0 ? yb0 + pi2 : yb0 - pi2);
    const double zs = wz == 0 ? zb0 : (wz > 0 ? zb0 + pi2 : zb0 - pi2);
    const int N = g.N;
    const int ixa = ref_cell_local(xa, h, N), iya = ref_cell_local(ya, h, N), iza = ref_cell_local(za, h, N);
    int ixb = 1 + (int)(xb0 / h * (double)N);
    if (ixb > N) ixb = N;
    const int iyb = ref_cell_halo(ys, h, N, g.nx_max), izb = ref_cell_halo(zs, h, N, g.nx_max);
    const int j = iyb - iya, k = izb - iza;
    if (j < -g.imax || j > g.imax || k < -g.imax || k > g.imax) return false;
    if (iyb < 1 - g.range || iyb > N + g.range || izb < 1 - g.range || izb > N + g.range) return false;
    if (izb < 1) return false;
    if (izb <= N && iyb < 1) return false;
    const int i0 = ixb - ixa;
    int i, sx;
    if (i0 >= -g.imax && i0 <= g.imax) { i = i0; sx = 0; }
    else if (i0 + N >= -g.imax && i0 + N <= g.imax) { i = i0 + N; sx = 1; }
    else if (i0 - N >= -g.imax && i0 - N <= g.imax) { i = i0 - N; sx = -1; }
    else return false;
    if (sx != wx) return false;
    *xb = xb0 + (double)sx * h;
    *yb = ys;
    *zb = zs;
    *di = i; *dj = j; *dk = k;
    return true;
}

__device__ __forceinline__ void findindex(const double* radius, int nset, double r1, double r2, int* i1, int* i2)
{
    double diff1 = 1000000., diff2 = 1000000.;
    int a = 1, b = 1;
    for (int il = 1; il <= nset; ++il) {                // 1857-1882
        const double d1 = fabs(radius[il - 1] - r1), d2 = fabs(radius[il - 1] - r2);
        if (d1 <= diff1) { diff1 = d1; a = il; }
        if (d2 <= diff2) { diff2 = d2; b = il; }
    }
    if (b < a) { const int t = b; b = a; a = t; }
    *i1 = a; *i2 = b;
}

struct CollArgs {
    const double *pos, *vel, *rad, *gid, *fvp0, *dmove, *vp;
    const double4* pk;     // (x, y, z, radius) in cell order
    const int *cell_start, *sorted;
    int n, cap, nset;
    double radius[DDNS_MAXSET];
    double dt, st;
    int hdi;
    PairGeom g;
    // outputs
    int* nirbin;
    double* bins;          // wr, wx, wy, wz, ur, theta, wrtan (each nset*nset*NBIN)
    int* counts;           // [0..2] icc, [8..107] colltype, [108..207] nir, [256] npairs
    long long* tested;     // [0] evaluated, [1] within 10 rcoll
    long long* pairs;      // DDNS_PAIR_REC per pair, see k_collide
    int pair_cap;
    int* mult;
    const unsigned long long* dmax;   // bits of the largest |dmove| of this step, and of the largest radius
    int2* cand;            // pairs handed from the search to k_collide_pairs
    int cand_cap;
    int* ncand;
    int* hit;              // per droplet: ih1 of a collision already counted
    double* contact;       // DDNS_CONTACT_REC per pair in contact, count at counts[258]
    int contact_cap;
    int* ncontact;         // pairs in contact found by the search (sizes contact)
    int* far;              // set when a droplet's reach exceeds the cell width (the walk would miss partners)
};

// The largest displacement length |dmove| of the step (out[0], when dmove is given) and
// the largest droplet radius (out[1]), as the bits of non-negative doubles (whose order
// is the numeric order): the exact pre-filter of the collision time window and the
// reach of the candidate walks.  The radius is the droplets' own, which coalescence
// can grow beyond the largest size class.
__global__ void __launch_bounds__(256) k_extent_max(const double* __restrict__ dmove, const double* __restrict__ rad,
                                                    int n, int cap, unsigned long long* out)
{
    // grid-stride over a small fixed grid, then one atomic per work-group (one per
    // wavefront would serialise thousands of atomics on the single address)
    __shared__ double wm[2][4];
    const size_t C = (size_t)cap;
    double m = 0.0, mr = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (dmove) {
            const double d0 = dmove[i], d1 = dmove[C + i], d2 = dmove[2 * C + i];
            m = fmax(m, sqrt(d0 * d0 + d1 * d1 + d2 * d2));
        }
        mr = fmax(mr, rad[i]);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        m = fmax(m, __shfl_xor(m, o, 64));
        mr = fmax(mr, __shfl_xor(mr, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        wm[0][threadIdx.x >> 6] = m;
        wm[1][threadIdx.x >> 6] = mr;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(wm[0][0], wm[0][1]), fmax(wm[0][2], wm[0][3]));
        mr = fmax(fmax(wm[1][0], wm[1][1]), fmax(wm[1][2], wm[1][3]));
        if (m > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
        if (mr > 0.0) atomicMax(out + 1, (unsigned long long)__double_as_longlong(mr));
    }
}

// The reference's work on one pair (ih1 = droplet ia, ih2 = droplet ib or one of its
// images, at distance dnij0 with offset xd = xa - xb): the statistics of
// partlib.f90:1358-1441 and the collision tests and record of 1445-1562.
__device__ __forceinline__ void collide_pair(const CollArgs& a, int ia, const double (&xa)[3], double ra, double dmove_a,
                                             int ib, const double4& cb, int wy, int wz, const double (&xb)[3],
                                             const double (&xd)[3], double dnij0, int di, int dj, int dk,
                                             long long& instat, int& ncollip1, double* bins, int* nirbin)
{
    const size_t C = (size_t)a.cap;
    const int nbin = DDNS_NBIN, ns = a.nset;
    const double rb = cb.w;
    const double rcoll = ra + rb;
    const double rcolla = a.hdi ? rcoll : rcoll * (1. - a.st / 100.);
    const double rcollb = rcoll * (1. + a.st / 100.);
    if (dnij0 > rcolla && dnij0 <= 10 * rcoll) {      // 1358-1441
        ++instat;
        int i1 = 1, i2 = 1;
        if (dnij0 > rcoll) {
            double va[3], vb[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                va[c] = a.fvp0[c * C + ia];
                vb[c] = a.fvp0[c * C + ib];
            }
            double xh[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) xh[c] = xd[c] / dnij0;
            const double Vr1 = va[0] * xh[0] + va[1] * xh[1] + va[2] * xh[2];
            const double Vr2 = vb[0] * xh[0] + vb[1] * xh[1] + vb[2] * xh[2];
            const double Wr = Vr1 - Vr2;
            int ir = (int)((dnij0 - rcoll) / ((9. / (double)nbin) * rcoll));
            if (ir > nbin - 1) ir = nbin - 1;
            if (ir < 0) ir = 0;
            findindex(a.radius, ns, ra, rb, &i1, &i2);
            const int bi = (i1 - 1) + ns * ((i2 - 1) + ns * ir);
            const size_t nb = (size_t)ns * ns * nbin;
            atomicAdd(nirbin + bi, 1);
            if (i2 == i1) {
                atomicAdd(bins + 1 * nb + bi, va[0] - vb[0]);
                atomicAdd(bins + 2 * nb + bi, va[1] - vb[1]);
                atomicAdd(bins + 3 * nb + bi, va[2] - vb[2]);
            }
            if (ra > rb) {
                atomicAdd(bins + 1 * nb + bi, va[0] - vb[0]);
                atomicAdd(bins + 2 * nb + bi, va[1] - vb[1]);
                atomicAdd(bins + 3 * nb + bi, va[2] - vb[2]);
            } else if (ra < rb) {
                atomicAdd(bins + 1 * nb + bi, vb[0] - va[0]);
                atomicAdd(bins + 2 * nb + bi, vb[1] - va[1]);
                atomicAdd(bins + 3 * nb + bi, vb[2] - va[2]);
            }
            atomicAdd(bins + 0 * nb + bi, Wr);
            const double du0 = a.vp[ia] - a.vp[ib], du1 = a.vp[C + ia] - a.vp[C + ib], du2 = a.vp[2 * C + ia] - a.vp[2 * C + ib];
            atomicAdd(bins + 4 * nb + bi, du0 * xh[0] + du1 * xh[1] + du2 * xh[2]);
            const double aaa = (va[0] - vb[0]) * xh[0] + (va[1] - vb[1]) * xh[1] + (va[2] - vb[2]) * xh[2];
            const double bbb = sqrt((va[0] - vb[0]) * (va[0] - vb[0]) + (va[1] - vb[1]) * (va[1] - vb[1]) +
                                    (va[2] - vb[2]) * (va[2] - vb[2]));
            if (bbb != 0.) {
                const double q = fabs(aaa / bbb);
                const double vangle = acos(q);
                atomicAdd(bins + 5 * nb + bi, vangle);
                // sin(acos q) = sqrt(1 - q^2) for q in [0, 1]: the same value
                // as the reference's sin(vangle) (partlib.f90:1400) to rounding,
                // without the register cost of a general double sine
                const double sb = bbb * sqrt(fmax(0.0, 1.0 - q * q));
                atomicAdd(bins + 6 * nb + bi, sb * sb / dnij0);
            }
        }
        if (dnij0 <= rcollb) {                        // 1411-1437
            if (dnij0 <= rcoll) findindex(a.radius, ns, ra, rb, &i1, &i2);
            atomicAdd(a.counts + 108 + (i1 - 1) + DDNS_MAXSET * (i2 - 1), 1);
            const int slot = atomicAdd(a.counts + 258, 1);
            if (slot < a.contact_cap) {
                // the record buffer(1:12, ij), with the visiting order of the pair in front
                double va[3], vb[3], xh[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    va[c] = a.fvp0[c * C + ia];
                    vb[c] = a.fvp0[c * C + ib];
                    xh[c] = xd[c] / dnij0;
                }
                const double Vr1 = va[0] * xh[0] + va[1] * xh[1] + va[2] * xh[2];
                const double Vr2 = vb[0] * xh[0] + vb[1] * xh[1] + vb[2] * xh[2];
                const int dir = (wy == 0 && wz == 0) ? 0 : (wz == 0 ? 5 : (wy == 1 ? 6 : (wy == 0 ? 7 : 8)));
                const long long ypn = (long long)dir * a.cap + ib;
                const long long cell = ((long long)(di + 128) << 16) | ((dj + 128) << 8) | (dk + 128);
                double* r = a.contact + (size_t)DDNS_CONTACT_REC * slot;
                r[0] = __longlong_as_double((long long)ia);
                r[1] = __longlong_as_double((cell << 40) | (((1LL << 40) - 1) - ypn));
                r[2] = (double)i1;
                r[3] = (double)i2;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    r[4 + c] = va[c];
                    r[7 + c] = vb[c];
                    r[10 + c] = xh[c];
                }
                r[13] = Vr1 - Vr2;
            }
        }
    }
    if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) return;   // 1445
    // -------- coarse check and types I/II/III, 1451-1516
    double bx1[3], xbn[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        bx1[c] = a.dmove[c * C + ib];
        xbn[c] = xb[c] + bx1[c];
    }
    const double dmove_b = sqrt(bx1[0] * bx1[0] + bx1[1] * bx1[1] + bx1[2] * bx1[2]);
    const double tmove = rcoll + dmove_a + dmove_b;
    if (dnij0 > tmove) return;
    double ax1[3], xan[3], xe[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        ax1[c] = a.dmove[c * C + ia];
        xan[c] = xa[c] + ax1[c];
        xe[c] = xan[c] - xbn[c];
    }
    const double dnij1 = sqrt(xe[0] * xe[0] + xe[1] * xe[1] + xe[2] * xe[2]);
    if (dnij0 < rcoll && dnij1 > rcoll) return;
    int type = 0;
    if (dnij0 > rcoll && dnij1 <= rcoll) {
        type = 1;
    } else {
        double vbn[3], vb[3], va[3], vna[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            vbn[c] = a.vel[c * C + ib];
            vb[c] = a.fvp0[c * C + ib];
            va[c] = a.fvp0[c * C + ia];
            vna[c] = a.vel[c * C + ia];
        }
        const double dt = a.dt;
        #pragma unroll 1
        for (int icheck = 1; icheck <= 19 && type == 0; ++icheck) {
            const double tfac = 0.05 * (double)icheck;
            const double tfac1 = dt * tfac, tfac2 = tfac * tfac, tfac3 = tfac2 * tfac;
            double xt[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double xta = xa[c] + va[c] * tfac1 + (3. * ax1[c] - (2. * va[c] + vna[c]) * dt) * tfac2 +
                                   ((vna[c] + va[c]) * dt - 2. * ax1[c]) * tfac3;
                const double xtb = xb[c] + vb[c] * tfac1 + (3. * bx1[c] - (2. * vb[c] + vbn[c]) * dt) * tfac2 +
                                   ((vbn[c] + vb[c]) * dt - 2. * bx1[c]) * tfac3;
                xt[c] = xta - xtb;
            }
            const double dnij = sqrt(xt[0] * xt[0] + xt[1] * xt[1] + xt[2] * xt[2]);
            if (dnij0 > rcoll && dnij < rcoll) type = 2;
            else if (dnij1 <= rcoll && dnij > rcoll) type = 3;
        }
    }
    if (type == 0) return;
    // -------- count and record, 1518-1562
    ++ncollip1;
    atomicAdd(a.counts + (type - 1), 1);
    int i1, i2;
    findindex(a.radius, ns, ra, rb, &i1, &i2);
    atomicAdd(a.counts + 8 + (i1 - 1) + DDNS_MAXSET * (i2 - 1), 1);
    atomicAdd(a.mult + ib, 1);
    const int slot = atomicAdd(a.counts + 256, 1);
    if (slot < a.pair_cap) {
        // the collist record (1555-1562) plus where the reference meets
        // the pair: ih1, the cell offset (i, j, k) of its loop and the
        // partner's ypn index (halo images follow the locals, by HEADLIST
        // direction 5..8, then by droplet index)
        const long long ga = (long long)a.gid[ia], gb = (long long)a.gid[ib];
        const long long lo = ga < gb ? ga : gb, hi = ga < gb ? gb : ga;
        const int dir = (wy == 0 && wz == 0) ? 0 : (wz == 0 ? 5 : (wy == 1 ? 6 : (wy == 0 ? 7 : 8)));
        const long long ypn = (long long)dir * a.cap + ib;
        const int off = ((di + 1) * 3 + (dj + 1)) * 3 + (dk + 1);
        long long* r = a.pairs + DDNS_PAIR_REC * (size_t)slot;
        r[0] = (lo << 32) | (hi << 2) | type;
        r[1] = ia;
        r[2] = ib;
        r[3] = ((long long)(di + 8) << 16) | ((dj + 8) << 8) | (dk + 8);
        r[4] = dir;
        r[5] = __double_as_longlong(ra);
        r[6] = __double_as_longlong(rb);
        // visiting order: ih1 ascending, then the (i, j, k) loop, then the
        // linked list of the cell, which runs from high to low ypn index
        r[7] = ((long long)ia << 32) | ((long long)off << 27) | (((1LL << 27) - 1) - ypn);
    }
}

// Pair search, the first half of COLLISION_DET (partlib.f90:1166-1846): one lane per
// droplet ih1 walks the candidates, replays the reference's visiting rules and
// counts the pairs it tests.  Only the few pairs that can reach the statistics or
// the collision tests are handed on (cell-order position of ih1, of the partner and
// the neighbour-cell number), so the heavy, rarely taken branches do not stall
// whole wavefronts here.
// The handed-on pairs of a wavefront are staged in its own LDS buffer of STAGE entries
// and written out with one atomic on the global count per flush: one atomic per walk
// step with an emitting lane (a droplet in a dense realisation has tens of pairs in
// the statistics window) serialises the search on that count.
constexpr int STAGE = 512;

struct CandStage {
    int2* buf;              // STAGE entries of this wavefront (LDS)
    volatile int* cnt;      // entries held (LDS); written by one lane, read by all
};

// Hand the staged entries to the global list: the lanes active here (mask exec) copy
// them to the slots one atomic reserved.  Entries past cand_cap are dropped and
// counted; the host sees the overflow and runs the search again with a larger list.
__device__ __forceinline__ void stage_flush(const CollArgs& a, const CandStage& st, int c, int lane,
                                            unsigned long long exec)
{
    const int leader = __ffsll((long long)exec) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(a.ncand, c);
    base = __shfl(base, leader, 64);
    const int nact = __popcll(exec), rank = __popcll(exec & ((1ull << lane) - 1));
    for (int i = rank; i < c; i += nact)
        if (base + i < a.cand_cap) a.cand[base + i] = st.buf[i];
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void collide_search(const CollArgs& a, int t, int lane, int& tested, int& touching,
                                               const CandStage& st)
{
    const int ia = a.sorted[t];
    const size_t C = (size_t)a.cap;
    const PairGeom& g = a.g;
    // position and radius from the droplet's own packed record (the same doubles as
    // pos and rad, read contiguously in cell order instead of gathered)
    const double4 me = a.pk[t];
    const double xa[3] = {me.x, me.y, me.z};
    const double ra = me.w;
    double dmove_a;
    {
        const double d0 = a.dmove[ia], d1 = a.dmove[C + ia], d2 = a.dmove[2 * C + ia];
        dmove_a = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    }
    int cc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int q = (int)(xa[c] / g.hc);
        cc[c] = q < 0 ? 0 : (q >= g.G ? g.G - 1 : q);
    }
    // rcoll + dmove_a + dmove_b <= rcoll + dmove_a + dmax: a pair farther apart than the
    // right side cannot pass the collision time-window test
    const double dmax = __longlong_as_double((long long)a.dmax[0]);
    // a partner farther than 10 (ra + rbig) is outside the statistics and farther than
    // ra + rbig + dmove_a + dmax outside the collision window, rbig the largest radius:
    // only the cells within the larger of the two are walked
    const double rbig = __longlong_as_double((long long)a.dmax[1]);
    const double reach = fmax(10. * (ra + rbig), ra + rbig + dmove_a + dmax);
    const double reach2 = reach * reach * (1. + 1e-9);
    // the walk covers the neighbour cells only: a reach beyond the cell width is reported,
    // and the search is repeated on wider cells
    if (reach + 1e-9 * g.hc > g.hc) atomicOr(a.far, 1);
    for_each_candidate(a.cell_start, g.G, cc, a.pk, a.sorted, [&](int p, int kc, const double4& cb, int ib) {
        if (p == t) return;
        int wx, wy, wz;
        neighbour_shift(kc, g.G, cc, &wx, &wy, &wz);
        ++tested;
        // the distance to the image the visiting rules would take (shift w * 2 pi): a
        // partner beyond the reach can neither be counted nor collide, so the rules are
        // replayed for the near ones only
        {
            const double e0 = xa[0] - (cb.x + (double)wx * g.h), e1 = xa[1] - (cb.y + (double)wy * g.h),
                         e2 = xa[2] - (cb.z + (double)wz * g.h);
            if (e0 * e0 + e1 * e1 + e2 * e2 > reach2) return;
        }
        double xb[3];
        int di, dj, dk;
        if (!ref_visit(g, ia, ib, xa[0], xa[1], xa[2], cb.x, cb.y, cb.z,
                       wx, wy, wz, &xb[0], &xb[1], &xb[2], &di, &dj, &dk))
            return;
        double xd[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) xd[c] = xa[c] - xb[c];
        const double dnij0 = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
        const double rcoll = ra + cb.w;
        const double rcolla = a.hdi ? rcoll : rcoll * (1. - a.st / 100.);
        const bool stat = dnij0 > rcolla && dnij0 <= 10 * rcoll;                         // 1358
        touching += (stat && dnij0 <= rcoll * (1. + a.st / 100.)) ? 1 : 0;               // 1411
        const bool near = !(di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) &&   // 1445
                          !(dnij0 > rcoll + dmove_a + dmax);
        const bool emit = stat || near;
        // the emitting lanes take consecutive entries of the wavefront's stage
        const unsigned long long m = __ballot(emit);
        if (m) {
            const int k = __popcll(m);
            int c = *st.cnt;
            if (c + k > STAGE) {
                stage_flush(a, st, c, lane, __ballot(1));
                c = 0;
            }
            if (emit) st.buf[c + __popcll(m & ((1ull << lane) - 1))] = make_int2(t, (int)((unsigned)p | ((unsigned)kc << 27)));
            __builtin_amdgcn_wave_barrier();
            if (lane == __ffsll((long long)m) - 1) *st.cnt = c + k;
            __builtin_amdgcn_wave_barrier();
        }
    }, reach_span(xa, cc, g.hc, reach));
}

__global__ void __launch_bounds__(CT) k_collide(CollArgs a)
{
    const int t = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    __shared__ int2 stage_buf[CT / 64][STAGE];
    __shared__ int stage_cnt[CT / 64];
    const CandStage st{stage_buf[threadIdx.x >> 6], stage_cnt + (threadIdx.x >> 6)};
    if (lane == 0) *st.cnt = 0;
    __builtin_amdgcn_wave_barrier();
    int tested = 0, touching = 0;
    if (t < a.n) collide_search(a, t, lane, tested, touching, st);
    // the whole wavefront is back together: hand on what its stage still holds
    {
        const int c = *st.cnt;
        if (c) stage_flush(a, st, c, lane, __ballot(1));
    }
    // one atomic per wavefront for the counters (one per lane would serialise a
    // million atomics on two addresses)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        tested += __shfl_xor(tested, o, 64);
        touching += __shfl_xor(touching, o, 64);
    }
    if (lane == 0) {
        if (tested) atomicAdd((unsigned long long*)a.tested, (unsigned long long)tested);
        if (touching) atomicAdd(a.ncontact, touching);
    }
}

// Second half: one thread per handed-on pair redoes the pair's geometry exactly as the
// search did and runs the statistics and collision tests.
// The kernel is queued right behind the search and reads the search's count on the
// device (the host checks it for overflow afterwards); a grid-stride loop over the pairs.
// With LB the statistics bins (nset^2 NBIN of them, at most LB_MAXBIN) are summed in
// LDS per work-group and added to the global bins once at the end: a realisation of one
// or two droplet sizes puts all its pairs on 180 or 720 bins, and one global atomic per
// pair and bin value serialises the kernel on them.
constexpr int LB_MAXBIN = 4 * DDNS_NBIN;

template <bool LB>
__global__ void __launch_bounds__(256) k_collide_pairs(CollArgs a)
{
    extern __shared__ double lds_bins[];
    const int nb = a.nset * a.nset * DDNS_NBIN;
    double* bins = a.bins;
    int* nirbin = a.nirbin;
    if (LB) {
        bins = lds_bins;
        nirbin = (int*)(lds_bins + 7 * nb);
        for (int i = threadIdx.x; i < 7 * nb; i += blockDim.x) bins[i] = 0.0;
        for (int i = threadIdx.x; i < nb; i += blockDim.x) nirbin[i] = 0;
        __syncthreads();
    }
    const int ncand = min(*a.ncand, a.cand_cap);
    long long instat = 0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < ncand; r += gridDim.x * blockDim.x) {
        const int2 rec = a.cand[r];
        const int t = rec.x, p = rec.y & ((1 << 27) - 1), kc = (int)((unsigned)rec.y >> 27);   // kc < 27: 5 bits
        const int ia = a.sorted[t];
        const size_t C = (size_t)a.cap;
        const PairGeom& g = a.g;
        double xa[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) xa[c] = a.pos[c * C + ia];
        const double ra = a.rad[ia];
        double dmove_a;
        {
            const double d0 = a.dmove[ia], d1 = a.dmove[C + ia], d2 = a.dmove[2 * C + ia];
            dmove_a = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        }
        int cc[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            int q = (int)(xa[c] / g.hc);
            cc[c] = q < 0 ? 0 : (q >= g.G ? g.G - 1 : q);
        }
        int wx, wy, wz;
        neighbour_shift(kc, g.G, cc, &wx, &wy, &wz);
        const int ib = a.sorted[p];
        const double4 cb = a.pk[p];
        double xb[3];
        int di, dj, dk;
        if (ref_visit(g, ia, ib, xa[0], xa[1], xa[2], cb.x, cb.y, cb.z, wx, wy, wz, &xb[0], &xb[1], &xb[2], &di, &dj, &dk)) {
            double xd[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) xd[c] = xa[c] - xb[c];
            const double dnij0 = sqrt(xd[0] * xd[0] + xd[1] * xd[1] + xd[2] * xd[2]);
            int ncollip1 = 0;
            collide_pair(a, ia, xa, ra, dmove_a, ib, cb, wy, wz, xb, xd, dnij0, di, dj, dk, instat, ncollip1,
                         bins, nirbin);
            // ih1 counts once among the droplets that collided (1578-1583)
            if (ncollip1 != 0 && atomicExch(a.hit + ia, 1) == 0) {
                atomicAdd(a.mult + ia, 1);
                atomicAdd(a.counts + 257, 1);
            }
        }
    }
    // pairs in the statistics window: one atomic per wavefront
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) instat += __shfl_xor(instat, o, 64);
    if ((threadIdx.x & 63) == 0 && instat) atomicAdd((unsigned long long*)(a.tested + 1), (unsigned long long)instat);
    if (LB) {
        __syncthreads();
        // a bin with no pair in this work-group adds nothing (its count is 0)
        for (int i = threadIdx.x; i < nb; i += blockDim.x) {
            if (!nirbin[i]) continue;
            atomicAdd(a.nirbin + i, nirbin[i]);
#pragma unroll
            for (int v = 0; v < 7; ++v) {
                const double x = bins[v * nb + i];
                if (x != 0.0) atomicAdd(a.bins + (size_t)v * nb + i, x);
            }
        }
    }
}

// OVERLAP_DET sweep (2076-2157): count, per droplet as ih1, the overlapping pairs
// the reference finds; each one costs three relocation draws.
struct OverArgs {
    const double *pos, *rad;
    const double4* pk;
    const int *cell_start, *sorted;
    int n, cap;
    PairGeom g;
    const unsigned long long* rbig;   // bits of the largest radius: a partner closer than ra + rbig can overlap
    int* cnt;
    int* total;
};

__global__ void __launch_bounds__(CT) k_overlap(OverArgs a)
{
    const int t = xcd_chunk(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const int ia = a.sorted[t];
    const PairGeom& g = a.g;
    const double4 ca = a.pk[t];
    const double xa = ca.x, ya = ca.y, za = ca.z;
    const double ra = ca.w;
    int cc[3];
    const double xs[3] = {xa, ya, za};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        int q = (int)(xs[c] / g.hc);
        cc[c] = q < 0 ? 0 : (q >= g.G ? g.G - 1 : q);
    }
    int cnt = 0;
    // a light kernel: the row-wise walk needs no LDS, which keeps its occupancy high;
    // only the cells within ra + rbig of the droplet can hold an overlapping partner,
    // and the visiting rules are replayed for the partners within that distance only
    const double reach = ra + __longlong_as_double((long long)*a.rbig);
    const double reach2 = reach * reach * (1. + 1e-9);
    for_each_candidate(a.cell_start, g.G, cc, a.pk, a.sorted, [&](int p, int kc, const double4& cb, int ib) {
        if (p == t) return;
        int wx, wy, wz;
        neighbour_shift(kc, g.G, cc, &wx, &wy, &wz);
        {
            const double e0 = xa - (cb.x + (double)wx * g.h), e1 = ya - (cb.y + (double)wy * g.h),
                         e2 = za - (cb.z + (double)wz * g.h);
            if (e0 * e0 + e1 * e1 + e2 * e2 > reach2) return;
        }
        double xb, yb, zb;
        int di, dj, dk;
        if (!ref_visit(g, ia, ib, xa, ya, za, cb.x, cb.y, cb.z, wx, wy, wz, &xb, &yb, &zb, &di, &dj, &dk)) return;
        if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) return;   // 2091
        const double d0 = xa - xb, d1 = ya - yb, d2 = za - zb;
        const double dnij = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        const double rcoll = ra + cb.w;
        if (dnij <= rcoll) ++cnt;
    }, reach_span(xs, cc, g.hc, reach));
    a.cnt[ia] = cnt;
    if (cnt) atomicAdd(a.total, cnt);
}

// ------------------------------------------------------------------ host side
int relocate_by_mult(PartBuffers& pb, const PartConst& pc, int* total_host, hipStream_t s);

static PairGeom make_geom(const PartBuffers& pb, const PartConst& pc, double dstat, double dist, int range)
{
    PairGeom g;
    g.N = pc.N;
    g.h = pc.pi2;
    g.imax = 1 + (int)(dstat * (double)pc.N / pc.pi2);
    g.range = range;
    g.nx_max = pc.nx_max;
    g.dist = dist;
    g.G = pb.G;
    g.hc = pb.hc;
    return g;
}

// Several device spans (4-byte multiples) cleared by one launch instead of one fill
// per span: the collision statistics start from zero every step.
struct ZeroSpans {
    unsigned* p[8];
    size_t words[8];
    int n;
};

__global__ void __launch_bounds__(256) k_zero_spans(ZeroSpans z)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int k = 0; k < z.n; ++k)
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < z.words[k]; i += stride) z.p[k][i] = 0u;
}

// The collision search and the pair kernel behind it are queued by part_collide_search
// (no host wait), so the caller can queue other streams' work while they run;
// part_collide_finish then waits for them and reads the search's counts: when the
// candidates or the contact records overflowed their buffers, it widens them and runs
// both again (the search clears everything the pair kernel writes).
int part_collide_search(PartBuffers& pb, const PartConst& pc, hipStream_t s);
int part_collide_finish(PartBuffers& pb, const PartConst& pc, hipStream_t s);

int part_collide(PartBuffers& pb, const PartConst& pc, hipStream_t s)
{
    const int st = part_collide_search(pb, pc, s);
    return st ? st : part_collide_finish(pb, pc, s);
}

static CollArgs collide_args(PartBuffers& pb, const PartConst& pc)
{
    CollArgs a{};
    a.pos = pb.pos; a.vel = pb.vel; a.rad = pb.rad; a.gid = pb.gid; a.fvp0 = pb.fvp + (size_t)pb.vcol(0) * pb.cap; a.dmove = pb.dmove; a.vp = pb.vp;
    a.cell_start = pb.cell_start; a.sorted = pb.sorted; a.pk = pb.pk;
    a.n = pc.n; a.cap = pb.cap; a.nset = pc.nset;
    for (int i = 0; i < DDNS_MAXSET; ++i) a.radius[i] = pc.radius[i];
    a.dt = pc.dt; a.st = pc.st; a.hdi = pc.hdi;
    const double dstat = 20.0 * pc.radmax;                                 // 1264
    a.g = make_geom(pb, pc, dstat, pc.hdi ? pc.d_hdi : pc.d_coll, pc.nx_coll);
    a.nirbin = pb.nirbin; a.bins = pb.bins; a.counts = pb.counts; a.tested = pb.tested;
    a.pairs = pb.pairs; a.pair_cap = pb.pair_cap; a.mult = pb.mult;
    a.dmax = (unsigned long long*)(pb.flags + 32);
    a.hit = pb.hit;
    a.ncand = pb.flags + 40;
    a.ncontact = pb.flags + 41;
    a.far = pb.flags + 42;
    a.cand = (int2*)pb.cand;
    a.cand_cap = pb.cand_cap;
    a.contact = pb.contact;
    a.contact_cap = pb.contact_cap;
    return a;
}

int part_collide_search(PartBuffers& pb, const PartConst& pc, hipStream_t s)
{
    const size_t nb = (size_t)pc.nset * pc.nset * DDNS_NBIN;
    if (!pb.hit) DDNS_HIP(hipMalloc(&pb.hit, (size_t)pb.cap * sizeof(int)));
    {
        // bins, counters, the removal multiplicities, the per-droplet hit marks and the
        // search's flags ([32..35] extents, [40..41] pair counts)
        ZeroSpans z{};
        auto add = [&](void* p, size_t bytes) { z.p[z.n] = (unsigned*)p; z.words[z.n] = bytes / 4; ++z.n; };
        add(pb.nirbin, nb * sizeof(int));
        add(pb.bins, 7 * nb * sizeof(double));
        add(pb.counts, 512 * sizeof(int));
        add(pb.tested, 4 * sizeof(long long));
        add(pb.mult, (size_t)(pc.n + 1) * sizeof(int));
        add(pb.hit, (size_t)pc.n * sizeof(int));
        add(pb.flags + 32, 11 * sizeof(int));
        add(pb.flags + 48, sizeof(int));               // relocation-draw overflow of this step
        hipLaunchKernelGGL(k_zero_spans, dim3(512), dim3(256), 0, s, z);
        DDNS_HIP(hipGetLastError());
    }
    if (pc.n == 0) return 0;
    if (!pb.cand) {
        pb.cand_cap = std::max(1 << 16, 2 * pb.cap);
        DDNS_HIP(hipMalloc(&pb.cand, (size_t)pb.cand_cap * sizeof(int2)));
    }
    if (!pb.found_host) DDNS_HIP(hipHostMalloc((void**)&pb.found_host, 4 * sizeof(int), hipHostMallocDefault));
    if (!pb.contact) {
        pb.contact_cap = 4096;
        DDNS_HIP(hipMalloc(&pb.contact, (size_t)pb.contact_cap * DDNS_CONTACT_REC * sizeof(double)));
    }
    CollArgs a = collide_args(pb, pc);
    hipLaunchKernelGGL(k_extent_max, dim3(std::min(256, (pc.n + 255) / 256)), dim3(256), 0, s, (const double*)pb.dmove,
                       (const double*)pb.rad, pc.n, pb.cap, (unsigned long long*)(pb.flags + 32));
    hipLaunchKernelGGL(k_collide, dim3((pc.n + CT - 1) / CT), dim3(CT), 0, s, a);
    // handed-on pairs and pairs in contact, read by part_collide_finish
    DDNS_HIP(hipMemcpyAsync(pb.found_host, a.ncand, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
    // the pair kernel, sized by the last step's count (a grid-stride loop covers more)
    const int nwg = std::max(256, std::min(4096, (int)((pb.last_ncand + pb.last_ncand / 4 + 255) / 256)));
    if ((int)nb <= LB_MAXBIN)
        // bins in LDS: two work-groups per CU, each folding its sums into the global bins once
        hipLaunchKernelGGL(k_collide_pairs<true>, dim3(std::min(nwg, 512)), dim3(256), nb * (7 * sizeof(double) + sizeof(int)),
                           s, a);
    else
        hipLaunchKernelGGL(k_collide_pairs<false>, dim3(nwg), dim3(256), 0, s, a);
    DDNS_HIP(hipGetLastError());
    return 0;
}

int part_collide_finish(PartBuffers& pb, const PartConst& pc, hipStream_t s)
{
    if (pc.n == 0) return 0;
    for (;;) {
        DDNS_HIP(wait_stream(pb, s));
        const int ncand = pb.found_host[0], ncontact = pb.found_host[1];
        pb.last_ncand = ncand;
        if (pb.found_host[2]) return DDNS_COLL_REGRID;     // cells too narrow for a droplet's reach
        if (ncand <= pb.cand_cap && ncontact <= pb.contact_cap) return 0;
        // more pairs to hand on, or more contact records, than the buffers hold: widen
        // them and run the search and the pair kernel again
        if (ncand > pb.cand_cap) {
            DDNS_HIP(hipFree(pb.cand));
            pb.cand_cap = ncand + ncand / 2;
            DDNS_HIP(hipMalloc(&pb.cand, (size_t)pb.cand_cap * sizeof(int2)));
        }
        if (ncontact > pb.contact_cap) {
            if (pb.contact) DDNS_HIP(hipFree(pb.contact));
            pb.contact_cap = std::max(4096, ncontact + ncontact / 2);
            DDNS_HIP(hipMalloc(&pb.contact, (size_t)pb.contact_cap * DDNS_CONTACT_REC * sizeof(double)));
        }
        const int st = part_collide_search(pb, pc, s);
        if (st) return st;
    }
}

// The last detection's contact records, buffer(1:12, ij) of partlib.f90:1431-1437, in the
// reference's visiting order: ih1 ascending, then the (i, j, k) cell loop, then the linked
// list of the cell from high to low ypn index.
int part_read_contacts(const PartBuffers& pb, double* rec, int cap, int* n, hipStream_t s)
{
    int cnt = 0;
    DDNS_HIP(hipMemcpyAsync(&cnt, pb.counts + 258, sizeof(int), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    *n = cnt;
    if (cnt > pb.contact_cap) { set_error("contact record overflow"); return DDNS_ERR_CAPACITY; }
    if (!rec || cap <= 0 || cnt == 0) return 0;
    std::vector<double> raw((size_t)cnt * DDNS_CONTACT_REC);
    DDNS_HIP(hipMemcpyAsync(raw.data(), pb.contact, raw.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    std::vector<int> order(cnt);
    for (int i = 0; i < cnt; ++i) order[i] = i;
    auto key = [&](int i, int w) { long long k; std::memcpy(&k, &raw[(size_t)i * DDNS_CONTACT_REC + w], 8); return k; };
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        const long long x0 = key(x, 0), y0 = key(y, 0);
        return x0 != y0 ? x0 < y0 : key(x, 1) < key(y, 1);
    });
    const int m = std::min(cnt, cap);
    for (int i = 0; i < m; ++i)
        std::memcpy(rec + (size_t)12 * i, &raw[(size_t)order[i] * DDNS_CONTACT_REC + 2], 12 * sizeof(double));
    return 0;
}

// the largest radius into rbig[1] (relocation moves droplets, it does not resize them)
static int overlap_setup(PartBuffers& pb, const PartConst& pc, hipStream_t s)
{
    unsigned long long* rbig = (unsigned long long*)(pb.flags + 36);
    DDNS_HIP(hipMemsetAsync(rbig, 0, 2 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_extent_max, dim3(std::min(256, (pc.n + 255) / 256)), dim3(256), 0, s, (const double*)nullptr,
                       (const double*)pb.rad, pc.n, pb.cap, rbig);
    return 0;
}

// One sweep's search: the overlap grid, and the per-droplet overlap counts in mult
static int overlap_search(PartBuffers& pb, const PartConst& pc, hipStream_t s)
{
    const double d_overl = 2.0 * pc.radmax;                                // 2063
    // two droplets in adjacent reference cells (the window of partlib.f90:2091) are
    // less than 2 dx apart per component, so cells of width >= 2 dx keep every such
    // pair within the 27-cell neighbourhood
    int st = part_build_cells(pb, pc, std::max(d_overl, 2.0 * pc.pi2 / pc.N), s);
    if (st) return st;
    OverArgs a{};
    a.pos = pb.pos; a.rad = pb.rad; a.pk = pb.pk; a.cell_start = pb.cell_start; a.sorted = pb.sorted;
    a.n = pc.n; a.cap = pb.cap;
    const int jmax = 1 + (int)(d_overl * (double)pc.N / pc.pi2);
    a.g = make_geom(pb, pc, d_overl, d_overl, jmax);
    a.rbig = (unsigned long long*)(pb.flags + 36) + 1;
    a.cnt = pb.mult;
    a.total = pb.flags + 1;
    DDNS_HIP(hipMemsetAsync(pb.flags + 1, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_overlap, dim3((pc.n + CT - 1) / CT), dim3(CT), 0, s, a);
    DDNS_HIP(hipGetLastError());
    return 0;
}

// sweeps from the current positions until one relocates nothing, each relocation waited for
static int overlap_loop(PartBuffers& pb, const PartConst& pc, int* sweeps, int* relocated, hipStream_t s)
{
    for (int it = *sweeps; it < 10000; ++it) {
        int st = overlap_search(pb, pc, s);
        if (st) return st;
        int total = 0;
        st = relocate_by_mult(pb, pc, &total, s);
        if (st) return st;
        if (total == 0) return 0;
        *sweeps += 1;
        *relocated += total;
    }
    set_error("OVERLAP_DET did not terminate after 10000 sweeps");
    return DDNS_ERR_OVERLAP;
}

int part_overlap(PartBuffers& pb, const PartConst& pc, int* sweeps, int* relocated, hipStream_t s)
{
    *sweeps = 0;
    *relocated = 0;
    if (pc.n == 0) return 0;
    int st = overlap_setup(pb, pc, s);
    if (st) return st;
    return overlap_loop(pb, pc, sweeps, relocated, s);
}

// The step's sweep without a host wait where it finds nothing to relocate (the usual
// case): the first sweep and its relocation are queued, the draws made for whatever the
// count is (as in the collision step's relocation: none for a count of 0), and the count
// is copied to pinned memory for the step's final wait.
int part_overlap_queue(PartBuffers& pb, const PartConst& pc, hipStream_t s, double next_width)
{
    pb.overlap_queued = false;
    if (pc.n == 0) return 0;
    int st = overlap_setup(pb, pc, s);
    if (st) return st;
    st = overlap_search(pb, pc, s);
    if (st) return st;
    // the next step's grid on the same positions, while the flow's z stage runs (a
    // relocation by this sweep invalidates it with the sweep's own grid)
    if (next_width > 0.0) {
        st = part_build_cells(pb, pc, next_width, s);
        if (st) return st;
    }
    const bool kept = pb.cells_valid, spare_kept = pb.spare.valid;
    st = relocate_async(pb, pc, s);          // invalidates the grids: positions may move
    if (st) return st;
    pb.overlap_cells_kept = kept;
    pb.overlap_spare_kept = spare_kept;
    if (!pb.counts_host)
        DDNS_HIP(hipHostMalloc((void**)&pb.counts_host, (512 + 8 + 4) * sizeof(int), hipHostMallocDefault));
    DDNS_HIP(hipMemcpyAsync(pb.counts_host + 521, pb.draw_off + pc.n, sizeof(int), hipMemcpyDeviceToHost, s));
    pb.overlap_queued = true;
    return 0;
}

// After the caller's wait: a first sweep that relocated nothing leaves the positions and
// the grids as they were; one that did is followed by the sweeps of part_overlap.
int part_overlap_finish(PartBuffers& pb, const PartConst& pc, int* sweeps, int* relocated, hipStream_t s)
{
    *sweeps = 0;
    *relocated = 0;
    if (!pb.overlap_queued) return 0;
    pb.overlap_queued = false;
    const int total = pb.counts_host[521];
    if (total == 0) {
        pb.cells_valid = pb.overlap_cells_kept;
        pb.spare.valid = pb.overlap_spare_kept;
        return 0;
    }
    *sweeps = 1;
    *relocated = total;
    return overlap_loop(pb, pc, sweeps, relocated, s);
}

// canonical pair list: sort the packed keys (min gid, max gid, type)
__global__ void k_sort_keys(long long* keys, int n)
{
    // single work-group odd-even merge via bitonic network in LDS (n <= 4096)
    __shared__ long long s[4096];
    int P = 1;
    while (P < n) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x) s[i] = i < n ? keys[i] : 0x7fffffffffffffffLL;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const long long x = s[i], y = s[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { s[i] = y; s[l] = x; }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < n; i += blockDim.x) keys[i] = s[i];
}

__global__ void k_bitonic_step(long long* s, int P, int j, int k)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int l = i ^ j;
    if (l > i) {
        const long long x = s[i], y = s[l];
        const bool up = (i & k) == 0;
        if ((x > y) == up) { s[i] = y; s[l] = x; }
    }
}

__global__ void k_gather_keys(const long long* pairs, long long* keys, int n, int P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    keys[i] = i < n ? pairs[DDNS_PAIR_REC * (size_t)i] : 0x7fffffffffffffffLL;
}

int part_read_pairs(PartBuffers& pb, const PartConst& pc, int64_t* triples, int cap, int* npairs, hipStream_t s)
{
    int np = 0;
    DDNS_HIP(hipMemcpyAsync(&np, pb.counts + 256, sizeof(int), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    *npairs = np;
    if (np > pb.pair_cap) { set_error("collision pair list overflow"); return DDNS_ERR_CAPACITY; }
    if (np == 0 || !triples || cap <= 0) return 0;
    int P = 1;
    while (P < np) P <<= 1;
    long long* keys = nullptr;
    DDNS_HIP(hipMalloc(&keys, (size_t)P * sizeof(long long)));
    hipLaunchKernelGGL(k_gather_keys, dim3((P + 255) / 256), dim3(256), 0, s, pb.pairs, keys, np, P);
    if (P <= 4096) {
        hipLaunchKernelGGL(k_sort_keys, dim3(1), dim3(1024), 0, s, keys, P);
    } else {
        for (int k = 2; k <= P; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1)
                hipLaunchKernelGGL(k_bitonic_step, dim3((P + 255) / 256), dim3(256), 0, s, keys, P, j, k);
    }
    DDNS_HIP(hipGetLastError());
    std::vector<long long> h(np);
    DDNS_HIP(hipMemcpyAsync(h.data(), keys, (size_t)np * sizeof(long long), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    DDNS_HIP(hipFree(keys));
    const int m = np < cap ? np : cap;
    for (int i = 0; i < m; ++i) {
        triples[3 * i + 0] = h[i] >> 32;
        triples[3 * i + 1] = (h[i] >> 2) & 0x3fffffff;
        triples[3 * i + 2] = h[i] & 3;
    }
    return 0;
}

// ------------------------------------------------------------------ coalescence
// (key, value) sort of up to 4096 entries in one work-group
__global__ void __launch_bounds__(1024) k_sort_kv_block(long long* keys, int* vals, int P)
{
    __shared__ long long sk[4096];
    __shared__ int sv[4096];
    for (int i = threadIdx.x; i < P; i += blockDim.x) { sk[i] = keys[i]; sv[i] = vals[i]; }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const long long x = sk[i], y = sk[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { sk[i] = y; sk[l] = x; const int t = sv[i]; sv[i] = sv[l]; sv[l] = t; }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < P; i += blockDim.x) { keys[i] = sk[i]; vals[i] = sv[i]; }
}

__global__ void k_bitonic_kv_step(long long* keys, int* vals, int P, int j, int k)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int l = i ^ j;
    if (l > i) {
        const long long x = keys[i], y = keys[l];
        const bool up = (i & k) == 0;
        if ((x > y) == up) { keys[i] = y; keys[l] = x; const int t = vals[i]; vals[i] = vals[l]; vals[l] = t; }
    }
}

__global__ void k_gather_order(const long long* pairs, long long* keys, int* vals, int n, int P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    keys[i] = i < n ? pairs[DDNS_PAIR_REC * (size_t)i + 7] : 0x7fffffffffffffffLL;
    vals[i] = i < n ? i : -1;
}

struct CoalArgs {
    double *pos, *vel, *rad, *gid, *vp, *dmove, *fvp, *fdp;
    const long long* pairs;
    const int* order;      // record slots in the reference's visiting order
    int npairs, n, cap;
    double Sv;
    int* removed;
};

// COALESCE, partlib.f90:3642-3720, applied pair by pair in the reference's order by
// one wavefront: ih1 takes the volume-equivalent radius and the terminal velocity;
// a partner that is a droplet of this domain (not a halo image) is overwritten by
// the current last droplet, whose slot is cleared.  Indices are the ones recorded at
// detection, as in the reference, so a droplet moved by an earlier pair is addressed
// by its old slot.  35 state columns are copied, one per lane.
__global__ void __launch_bounds__(64) k_coalesce(CoalArgs a)
{
    const int lane = threadIdx.x;
    const size_t C = (size_t)a.cap;
    int jj = 0;
    for (int q = 0; q < a.npairs; ++q) {
        const long long* r = a.pairs + DDNS_PAIR_REC * (size_t)a.order[q];
        const int p1 = (int)r[1], p2 = (int)r[2], dir = (int)r[4];
        if (lane == 0) {
            const double r1 = __longlong_as_double(r[5]), r2 = __longlong_as_double(r[6]);
            const double an = pow(pow(r1, 3.0) + pow(r2, 3.0), 1.0 / 3.0);
            a.rad[p1] = an;
            a.vel[p1] = a.Sv * (an * an);
            a.vel[C + p1] = 0.0;
            a.vel[2 * C + p1] = 0.0;
        }
        __threadfence();
        __syncthreads();
        if (dir == 0) {
            if (jj >= a.n) { jj = -1; break; }       // more removals than droplets
            ++jj;
            const size_t last = (size_t)(a.n - jj);
            double* col = nullptr;
            if (lane < 3) col = a.pos + lane * C;
            else if (lane < 6) col = a.vel + (lane - 3) * C;
            else if (lane == 6) col = a.rad;
            else if (lane == 7) col = a.gid;
            else if (lane < 11) col = a.vp + (lane - 8) * C;
            else if (lane < 14) col = a.dmove + (lane - 11) * C;
            else if (lane < 23) col = a.fvp + (lane - 14) * C;
            else if (lane < 35) col = a.fdp + (lane - 23) * C;
            if (col) {
                col[p2] = col[last];
                col[last] = 0.0;
            }
            __threadfence();
            __syncthreads();
        }
    }
    if (lane == 0) *a.removed = jj;
}

int part_coalesce(PartBuffers& pb, const PartConst& pc, int* removed, hipStream_t s)
{
    invalidate_cells(pb);
    *removed = 0;
    int np = 0;
    DDNS_HIP(hipMemcpyAsync(&np, pb.counts + 256, sizeof(int), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    if (np == 0) return 0;
    if (np > pb.pair_cap) { set_error("collision pair list overflow"); return DDNS_ERR_CAPACITY; }
    int P = 1;
    while (P < np) P <<= 1;
    long long* keys = nullptr;
    int* vals = nullptr;
    DDNS_HIP(hipMalloc(&keys, (size_t)P * sizeof(long long)));
    DDNS_HIP(hipMalloc(&vals, (size_t)P * sizeof(int)));
    hipLaunchKernelGGL(k_gather_order, dim3((P + 255) / 256), dim3(256), 0, s, pb.pairs, keys, vals, np, P);
    if (P <= 4096) {
        hipLaunchKernelGGL(k_sort_kv_block, dim3(1), dim3(1024), 0, s, keys, vals, P);
    } else {
        for (int k = 2; k <= P; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1)
                hipLaunchKernelGGL(k_bitonic_kv_step, dim3((P + 255) / 256), dim3(256), 0, s, keys, vals, P, j, k);
    }
    CoalArgs a{};
    a.pos = pb.pos; a.vel = pb.vel; a.rad = pb.rad; a.gid = pb.gid; a.vp = pb.vp; a.dmove = pb.dmove;
    a.fvp = pb.fvp; a.fdp = pb.fdp; a.pairs = pb.pairs; a.order = vals; a.npairs = np; a.n = pc.n;
    a.cap = pb.cap; a.Sv = pc.Sv_const; a.removed = pb.flags + 3;
    hipLaunchKernelGGL(k_coalesce, dim3(1), dim3(64), 0, s, a);
    DDNS_HIP(hipGetLastError());
    int rm = 0;
    DDNS_HIP(hipMemcpyAsync(&rm, pb.flags + 3, sizeof(int), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    DDNS_HIP(hipFree(keys));
    DDNS_HIP(hipFree(vals));
    if (rm < 0) { set_error("coalescence would remove every droplet"); return DDNS_ERR_DROPLET_LOSS; }
    *removed = rm;
    return 0;
}

int part_read_counts(const PartBuffers& pb, const PartConst& pc, ddns_collision_counts* cc, hipStream_t s)
{
    int h[512];
    long long t[4];
    DDNS_HIP(hipMemcpyAsync(h, pb.counts, sizeof(h), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipMemcpyAsync(t, pb.tested, sizeof(t), hipMemcpyDeviceToHost, s));
    int ncoll = 0;
    DDNS_HIP(hipStreamSynchronize(s));
    cc->icc[0] = h[0]; cc->icc[1] = h[1]; cc->icc[2] = h[2];
    for (int i = 0; i < DDNS_MAXSET * DDNS_MAXSET; ++i) { cc->colltype[i] = h[8 + i]; cc->nir[i] = h[108 + i]; }
    cc->npairs = h[256];
    cc->ncoll = h[256] + h[257];
    cc->nset = pc.nset;
    cc->pairs_tested = t[0];
    cc->pairs_in_dstat = t[1];
    (void)ncoll;
    return 0;
}

// The counts of part_read_counts through pinned memory: queued on the stream, read once
// the caller has waited for it (no extra round trip).  A relocation that ran out of draws
// (flags[48]) is reported here as a capacity error.
int part_queue_counts(PartBuffers& pb, hipStream_t s)
{
    if (!pb.counts_host)
        DDNS_HIP(hipHostMalloc((void**)&pb.counts_host, (512 + 8 + 4) * sizeof(int), hipHostMallocDefault));
    DDNS_HIP(hipMemcpyAsync(pb.counts_host, pb.counts, 512 * sizeof(int), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipMemcpyAsync(pb.counts_host + 512, pb.tested, 4 * sizeof(long long), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipMemcpyAsync(pb.counts_host + 520, pb.flags + 48, sizeof(int), hipMemcpyDeviceToHost, s));
    return 0;
}

int part_take_counts(const PartBuffers& pb, const PartConst& pc, ddns_collision_counts* cc)
{
    const int* h = pb.counts_host;
    long long t[4];
    std::memcpy(t, h + 512, sizeof(t));
    cc->icc[0] = h[0]; cc->icc[1] = h[1]; cc->icc[2] = h[2];
    for (int i = 0; i < DDNS_MAXSET * DDNS_MAXSET; ++i) { cc->colltype[i] = h[8 + i]; cc->nir[i] = h[108 + i]; }
    cc->npairs = h[256];
    cc->ncoll = h[256] + h[257];
    cc->nset = pc.nset;
    cc->pairs_tested = t[0];
    cc->pairs_in_dstat = t[1];
    if (h[520]) {
        set_error("relocation: the removal list needs more draws than three per droplet slot");
        return DDNS_ERR_CAPACITY;
    }
    return 0;
}

int part_read_bins(const PartBuffers& pb, const PartConst& pc, int* nirbin, double* const bins[7], hipStream_t s)
{
    const size_t nb = (size_t)pc.nset * pc.nset * DDNS_NBIN;
    if (nirbin) DDNS_HIP(hipMemcpyAsync(nirbin, pb.nirbin, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    for (int k = 0; k < 7; ++k)
        if (bins[k]) DDNS_HIP(hipMemcpyAsync(bins[k], pb.bins + k * nb, nb * sizeof(double), hipMemcpyDeviceToHost, s));
    DDNS_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // namespace ddns
