// ddns_api.hip -- the C ABI of libddns.so (declared in include/ddns.h).
//
// A handle owns one periodic domain on the current GPU and one HIP stream.  The
// step follows the reference driver loop turb.f90:492-707: the spectral step is
// split around the droplet block exactly where the reference places it (the
// droplets see the physical velocity of the inverse FFTs of this step).
#include "ddns_internal.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace ddns {

static thread_local std::string g_err = "no error";

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    g_err = buf;
    return DDNS_ERR_HIP;
}

int relocate_by_mult(PartBuffers& pb, const PartConst& pc, int* total_host, hipStream_t s);

}  // namespace ddns

using namespace ddns;

struct ddns_handle {
    ddns_params p;
    int N = 0, nek = 0, prec = 0;
    int dev = 0;                 // the GPU the handle's buffers live on (current at ddns_create)
    hipStream_t stream = nullptr;
    // the droplet block of a step runs here, beside the second half of the flow step on
    // stream (the droplets read only the physical velocity of the first half)
    hipStream_t dstream = nullptr;
    bool flow_rest_queued = false;   // the flow's second half of this step is on stream
    // z_gate: the second half is queued in two pieces, and its z stage waits for the
    // droplet block's collision-pair launch (DDNS_Z_GATE=0 at ddns_create queues it in
    // one piece); z_pending: the first piece is queued, the z stage not yet
    bool z_gate = true;
    bool z_pending = false;
    FlowBuffers fb;
    FlowConst fc{};
    bool a_valid = false;        // inverse pipeline holds the start of the next step
    bool phys_valid = false;
    int istep = 0;
    double time = 0;
    // droplets
    PartBuffers pb;
    PartConst pc{};
    int n = 0;
    int new_particle = 0;
    ddns_collision_counts cc{};
    ddns_timing timing{};
    hipEvent_t ev[16] = {};
    // flow stage boundaries (0 first half start, 1 end, 2 second half start, 3 end) and
    // kernel boundaries, one set per parity of the step: the next step's first half may
    // be queued before this step's times are read
    hipEvent_t flev[2][4] = {};
    hipEvent_t fev[2][FLOW_EVENTS] = {};
    // the parts of the flow step (FlowPart) as instantiated graphs, one per part and event
    // set (step parity), captured by the first step that runs them; graph_mode 0 launches
    // the kernels one by one (DDNS_GRAPH=0 in the environment at ddns_create)
    hipGraphExec_t gflow[4][2] = {};
    int graph_mode = 1;
    int par = 0;                 // parity of the step in progress
    bool want_next = false;      // a further step follows in this call
    bool next_queued = false;    // its first half is already on the flow stream
    // A member of a batch (ddns_batch): fb is a view of one realisation of the batch's
    // flow buffers (not owned), the batch runs the flow, and the member's step is only its
    // droplet block, whose gather waits for the batch's event phys_ev
    bool member = false;
    hipEvent_t phys_ev = nullptr;
    // the collision search once found a droplet whose reach exceeded its cells: the cells
    // stay at least two grid spacings wide from then on
    bool coll_wide = false;
    std::string err;
};

// ------------------------------------------------------------------ helpers
static double pi2_ref() { return 4. * std::atan(1.) * 2.0; }                 // turb.f90:296-297

// PART_INIT scaling, partlib.f90:104-150
static void droplet_scaling(const ddns_params& p, PartConst& pc)
{
    const double etk_CLOUD = std::sqrt(std::sqrt(p.viscosity_CLOUD * p.viscosity_CLOUD * p.viscosity_CLOUD / p.ediss_CLOUD));
    const double tauk_CLOUD = std::sqrt(p.viscosity_CLOUD / p.ediss_CLOUD);
    const double vk_CLOUD = etk_CLOUD / tauk_CLOUD;
    const double vk_DNS = p.etk_DNS / p.tauk_DNS;
    const double vk_ratio = vk_CLOUD / vk_DNS;
    const double etk_ratio = etk_CLOUD / p.etk_DNS;
    const double tauk_ratio = tauk_CLOUD / p.tauk_DNS;
    pc.radmax = 0.0;
    for (int il = 0; il < DDNS_MAXSET; ++il) pc.radius[il] = 0.0;
    for (int il = 0; il < p.nset; ++il) {
        pc.radius[il] = p.radius[il] * (p.etk_DNS / etk_CLOUD) / 10000.0;
        if (il == 0 || pc.radius[il] > pc.radmax) pc.radmax = pc.radius[il];
    }
    const double tau_CLOUD_const = 2. * p.rho_water * (etk_ratio * etk_ratio) / (9. * p.rho_air * p.viscosity_CLOUD);
    pc.Sv_const = tau_CLOUD_const * p.gravity / vk_ratio;
    pc.tau2 = tau_CLOUD_const / tauk_ratio;
    pc.pi2 = pi2_ref();
    pc.d_coll = 20.0 * pc.radmax;
    pc.nx_coll = 1 + (int)(pc.d_coll * (double)pc.N / pc.pi2);
    pc.nx_max = pc.nx_coll;
    pc.d_hdi = 0.0;
    if (p.hdi_included) {
        pc.d_hdi = p.hdi_trunc * pc.radmax;
        const int nx_hdi = 1 + (int)(pc.d_hdi * (double)pc.N / pc.pi2);
        pc.nx_max = std::max(pc.nx_max, nx_hdi);
    }
    pc.imax_stat = pc.nx_coll;
}

static int fail(ddns_handle* h, int code, const std::string& msg)
{
    set_error(msg);
    if (h) h->err = msg;
    return code;
}

static int keep(ddns_handle* h, int st)
{
    if (st < 0 && h) h->err = g_err;
    return st;
}

// ranff (accesories.f90:38-68) on the host, the portable generator of the setup
// routines (in place of IFPORT rand in flowlib.f90:163-164).
static double host_ranff(int* idum, int* iy, int* iv)
{
    const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB = 32;
    const int NDIV = 1 + (IM - 1) / NTAB;
    const double AM = 1.0 / IM, EPS = 1.2e-7, RNMX = 1.0 - EPS;
    int j, k;
    if (*idum <= 0 || *iy == 0) {
        *idum = (-*idum > 1) ? -*idum : 1;
        for (j = NTAB + 8; j >= 1; --j) {
            k = *idum / IQ;
            *idum = IA * (*idum - k * IQ) - IR * k;
            if (*idum < 0) *idum += IM;
            if (j <= NTAB) iv[j - 1] = *idum;
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

// ------------------------------------------------------------------ parameters
extern "C" int ddns_params_default(ddns_params* p)
{
    if (!p) return DDNS_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->nx = p->ny = p->nz = 64;
    p->nprocY = 1;
    p->nstep = 10;
    p->dt = 0.001;
    p->rnu = 0.026;
    p->time = 0.0;
    p->new_flow = 1;
    p->idp = 0;
    p->iseed_m0 = 123;
    p->forcing = 1;
    p->force[0] = 1.242477;
    p->force[1] = 0.391356;
    p->iseedf = 232300;
    p->ieout = 10;
    p->ispec = 1000;
    p->divstat = 1;
    p->rho_water = 1.0;
    p->rho_air = 0.001;
    p->etk_DNS = 0.0755;
    p->tauk_DNS = 0.2195;
    p->st = 2.0;
    p->ediss_CLOUD = 400.0;
    p->viscosity_CLOUD = 0.17;
    p->gravity = 980.0;
    p->new_particle = 1;
    p->nonoverlap = 0;
    p->hdi_included = 0;
    p->hdi_trunc = 50.0;
    p->npart = 0;
    p->ndrag = 0;
    p->nset = 1;
    p->radius[0] = 20.0;
    p->precision = DDNS_FP64;
    p->nstep_grow = 40000;
    p->ou_tf = 0.038;
    p->ou_var = 447.31;
    p->gmres_maxiter = 50;
    p->capacity = 0;
    p->gmres_tol = 1.0e-6;
    p->relocate_seed = 0;
    p->istep0 = 0;
    return DDNS_OK;
}

static bool parse_logical(const std::string& s, int* out)
{
    std::string t;
    for (char c : s) if (!std::isspace((unsigned char)c)) t += (char)std::toupper((unsigned char)c);
    if (t.find("TRUE") != std::string::npos || t == "T" || t == ".T.") { *out = 1; return true; }
    if (t.find("FALSE") != std::string::npos || t == "F" || t == ".F.") { *out = 0; return true; }
    return false;
}

// Positional reader: value lines at the positions turb.f90:103-164 reads them; every
// value line is followed by one label line.  Commas separate list values.
extern "C" int ddns_read_parameter_file(const char* path, ddns_params* p)
{
    if (!path || !p) return DDNS_ERR_ARG;
    std::ifstream f(path);
    if (!f) { set_error(std::string("cannot open ") + path); return DDNS_ERR_IO; }
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(f, line)) lines.push_back(line);
    ddns_params_default(p);
    size_t li = 0;
    auto next_value = [&](std::string& out) -> bool {
        if (li >= lines.size()) return false;
        out = lines[li++];
        for (char& c : out) if (c == ',') c = ' ';
        return true;
    };
    auto skip_label = [&]() { if (li < lines.size()) ++li; };
    std::string v;
    auto rd_int = [&](int* x) { if (!next_value(v)) return false; std::istringstream is(v); return (bool)(is >> *x); };
    auto rd_dbl = [&](double* x) { if (!next_value(v)) return false; std::istringstream is(v); std::string tok; if (!(is >> tok)) return false; for (char& c : tok) if (c == 'd' || c == 'D') c = 'e'; *x = std::strtod(tok.c_str(), nullptr); return true; };
    auto rd_log = [&](int* x) { if (!next_value(v)) return false; return parse_logical(v, x); };
    bool ok = true;
    {
        if (!next_value(v)) ok = false;
        std::istringstream is(v);
        ok = ok && (bool)(is >> p->nx >> p->ny >> p->nz);
    }
    skip_label(); ok = ok && rd_int(&p->nprocY);
    skip_label(); ok = ok && rd_int(&p->nstep);
    skip_label(); ok = ok && rd_dbl(&p->dt);
    skip_label(); ok = ok && rd_dbl(&p->rnu);
    skip_label(); ok = ok && rd_dbl(&p->time);
    skip_label(); ok = ok && rd_log(&p->new_flow);
    skip_label(); ok = ok && rd_int(&p->idp);
    skip_label(); ok = ok && rd_int(&p->iseed_m0);
    skip_label(); ok = ok && rd_int(&p->forcing);
    skip_label();
    if (ok && next_value(v)) {
        std::istringstream is(v);
        ok = (bool)(is >> p->force[0] >> p->force[1]);
    } else ok = false;
    skip_label(); ok = ok && rd_int(&p->iseedf);
    skip_label(); ok = ok && rd_int(&p->ieout);
    skip_label(); ok = ok && rd_int(&p->ispec);
    skip_label(); ok = ok && rd_log(&p->divstat);
    skip_label(); ok = ok && rd_dbl(&p->rho_water);
    skip_label(); ok = ok && rd_dbl(&p->rho_air);
    skip_label();
    if (ok && next_value(v)) {
        std::istringstream is(v);
        ok = (bool)(is >> p->etk_DNS >> p->tauk_DNS);
    } else ok = false;
    skip_label(); ok = ok && rd_dbl(&p->st);
    skip_label(); ok = ok && rd_dbl(&p->ediss_CLOUD);
    skip_label(); ok = ok && rd_dbl(&p->viscosity_CLOUD);
    skip_label(); ok = ok && rd_dbl(&p->gravity);
    skip_label(); ok = ok && rd_log(&p->new_particle);
    skip_label(); ok = ok && rd_log(&p->nonoverlap);
    skip_label(); ok = ok && rd_log(&p->hdi_included);
    skip_label(); ok = ok && rd_dbl(&p->hdi_trunc);
    skip_label(); ok = ok && rd_int(&p->npart);
    skip_label(); ok = ok && rd_int(&p->ndrag);
    skip_label(); ok = ok && rd_int(&p->nset);
    skip_label();
    if (ok && (p->nset < 1 || p->nset > DDNS_MAXSET)) ok = false;
    for (int i = 0; ok && i < p->nset; ++i) ok = rd_dbl(&p->radius[i]);
    if (!ok) { set_error(std::string("malformed parameter file ") + path); return DDNS_ERR_IO; }
    if (p->hdi_included) p->nonoverlap = 1;                                 // turb.f90:171
    return DDNS_OK;
}

// ------------------------------------------------------------------ lifetime
// A handle with its own flow buffers (view == nullptr), or a batch member whose flow
// buffers are the given view of the batch's (nothing allocated or tuned for the flow).
static void destroy_impl(ddns_handle* h);
static int create_impl(const ddns_params* pin, ddns_handle** out, const FlowBuffers* view)
{
    if (!pin || !out) return fail(nullptr, DDNS_ERR_ARG, "null argument");
    *out = nullptr;
    const int N = pin->nx;
    if (N != pin->ny || N != pin->nz || N < 32 || N > 1024 || (N & (N - 1)))
        return fail(nullptr, DDNS_ERR_ARG, "grid must be cubic, a power of two, 32..1024");
    if (pin->precision != DDNS_FP64 && pin->precision != DDNS_FP32_MIXED)
        return fail(nullptr, DDNS_ERR_ARG, "precision must be DDNS_FP64 or DDNS_FP32_MIXED");
    if (pin->nset < 1 || pin->nset > DDNS_MAXSET) return fail(nullptr, DDNS_ERR_ARG, "nset must be 1..10");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, DDNS_ERR_NO_DEVICE, "no GPU visible");
    ddns_handle* h = new ddns_handle();
    h->p = *pin;
    if (h->p.hdi_included) h->p.nonoverlap = 1;
    if (h->p.gmres_maxiter <= 0) h->p.gmres_maxiter = 50;
    if (h->p.gmres_tol <= 0) h->p.gmres_tol = 1.0e-6;
    if (h->p.ou_tf <= 0) h->p.ou_tf = 0.038;
    if (h->p.ou_var <= 0) h->p.ou_var = 447.31;
    if (h->p.nstep_grow <= 0) h->p.nstep_grow = 40000;
    h->N = N;
    h->nek = (int)(N / 2 - 1.5);
    h->prec = pin->precision;
    h->istep = pin->istep0;
    h->time = pin->time;
    h->new_particle = pin->new_particle;
    int st = 0;
    if (hipGetDevice(&h->dev) != hipSuccess) { delete h; return fail(nullptr, DDNS_ERR_HIP, "hipGetDevice"); }
    if (const char* g = std::getenv("DDNS_GRAPH")) h->graph_mode = g[0] != '0';
    if (const char* g = std::getenv("DDNS_Z_GATE")) h->z_gate = g[0] != '0';
    // the device address of the allocation-free spectrum's scratch, once per device
    st = spectrum_scratch_init();
    if (st) { delete h; return st; }
    if (hipStreamCreate(&h->stream) != hipSuccess) { delete h; return fail(nullptr, DDNS_ERR_HIP, "stream"); }
    if (hipStreamCreate(&h->dstream) != hipSuccess) { destroy_impl(h); return fail(nullptr, DDNS_ERR_HIP, "stream"); }
    if (view) {
        h->fb = *view;
        h->member = true;
        // the member's own scratch of the spectrum (the diagnostics allocate theirs on first use)
        if (hipMalloc(&h->fb.spec_part, ((size_t)SPEC_BLOCKS + 1) * SPEC_NEK_MAX * sizeof(double)) != hipSuccess) {
            h->fb.spec_part = nullptr;
            destroy_impl(h);
            return fail(nullptr, DDNS_ERR_HIP, "hipMalloc");
        }
    } else {
        st = flow_alloc(h->fb, N, h->prec);
        if (st) { destroy_impl(h); return st; }
    }
    h->fc.N = N; h->fc.nek = h->nek; h->fc.forcing = h->p.forcing; h->fc.dt = h->p.dt; h->fc.rnu = h->p.rnu;
    h->fc.force0 = h->p.force[0]; h->fc.force1 = h->p.force[1];
    h->fc.ou_tf = h->p.ou_tf; h->fc.ou_var = h->p.ou_var; h->fc.pi2 = pi2_ref();
    if (!view) {
        st = flow_tune(h->fb, h->fc, h->prec, h->stream);  // tile shapes, as FFTW_MEASURE plans (mpifft.f90:36-56)
        if (st) { destroy_impl(h); return st; }
    }
    // stochastic forcing starts from rest with iseedf = -iseedf (turb.f90:441-447)
    {
        int stt[34] = {0};
        stt[0] = -h->p.iseedf;
        if (hipMemcpy(h->fb.sth_state, stt, sizeof(stt), hipMemcpyHostToDevice) != hipSuccess) { destroy_impl(h); return fail(nullptr, DDNS_ERR_HIP, "copy"); }
    }
    // droplets
    h->pc.N = N;
    h->pc.nset = h->p.nset;
    h->pc.dt = h->p.dt; h->pc.rnu = h->p.rnu; h->pc.st = h->p.st;
    h->pc.hdi = h->p.hdi_included; h->pc.nonoverlap = h->p.nonoverlap; h->pc.ndrag = h->p.ndrag;
    h->pc.hdi_trunc = h->p.hdi_trunc; h->pc.precision = h->prec;
    droplet_scaling(h->p, h->pc);
    const int cap = std::max(1, h->p.capacity > 0 ? h->p.capacity : h->p.npart);
    if (h->p.npart > 0) {
        const double pi2 = h->pc.pi2;
        if (pi2 <= h->pc.d_coll || pi2 <= 2.0 * h->pc.radmax || (h->p.hdi_included && pi2 <= h->pc.d_hdi) ||
            2 * h->pc.nx_max + 1 >= N) {
            destroy_impl(h);
            return fail(nullptr, DDNS_ERR_DOMAIN, "cell/halo too thin for the collision or HI radius (partlib.f90:181-220)");
        }
    }
    st = part_alloc(h->pb, cap, h->p.nset, 1);
    if (st) { destroy_impl(h); return st; }
    st = part_set_rng(h->pb, h->p.relocate_seed, h->stream);
    if (st) { destroy_impl(h); return st; }
    for (int i = 0; i < 16; ++i)
        if (hipEventCreate(&h->ev[i]) != hipSuccess) { destroy_impl(h); return fail(nullptr, DDNS_ERR_HIP, "event"); }
    for (int q = 0; q < 2; ++q) {
        for (int i = 0; i < FLOW_EVENTS; ++i)
            if (hipEventCreate(&h->fev[q][i]) != hipSuccess) { destroy_impl(h); return fail(nullptr, DDNS_ERR_HIP, "event"); }
        for (int i = 0; i < 4; ++i)
            if (hipEventCreate(&h->flev[q][i]) != hipSuccess) { destroy_impl(h); return fail(nullptr, DDNS_ERR_HIP, "event"); }
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) { destroy_impl(h); return fail(nullptr, DDNS_ERR_HIP, "sync"); }
    *out = h;
    return DDNS_OK;
}

extern "C" int ddns_create(const ddns_params* pin, ddns_handle** out) { return create_impl(pin, out, nullptr); }

extern "C" int ddns_destroy(ddns_handle* h)
{
    if (!h) return DDNS_OK;
    if (h->member) return fail(h, DDNS_ERR_ARG, "a batch member is destroyed with its batch");
    destroy_impl(h);
    return DDNS_OK;
}

static void destroy_impl(ddns_handle* h)
{
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->dstream) (void)hipStreamSynchronize(h->dstream);
    if (!h->member) {
        flow_free(h->fb);
    } else {                     // a view: only its own scratch
        for (int q = 0; q < 8; ++q) if (h->fb.dg[q]) (void)hipFree(h->fb.dg[q]);
        if (h->fb.dg_part) (void)hipFree(h->fb.dg_part);
        if (h->fb.spec_part) (void)hipFree(h->fb.spec_part);
    }
    part_free(h->pb);
    for (int i = 0; i < 16; ++i) if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
    for (int q = 0; q < 2; ++q) {
        for (int i = 0; i < FLOW_EVENTS; ++i) if (h->fev[q][i]) (void)hipEventDestroy(h->fev[q][i]);
        for (int i = 0; i < 4; ++i) if (h->flev[q][i]) (void)hipEventDestroy(h->flev[q][i]);
        for (int a = 0; a < 4; ++a) if (h->gflow[a][q]) (void)hipGraphExecDestroy(h->gflow[a][q]);
    }
    if (h->dstream) (void)hipStreamDestroy(h->dstream);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

extern "C" const char* ddns_last_error(const ddns_handle* h)
{
    if (h && !h->err.empty()) return h->err.c_str();
    return g_err.c_str();
}

extern "C" int ddns_nek(const ddns_handle* h) { return h ? h->nek : 0; }

// ------------------------------------------------------------------ flow state
extern "C" int ddns_set_flow(ddns_handle* h, const double* vx, const double* vy, const double* vz,
                             const double* ox, const double* oy, const double* oz)
{
    if (!h) return DDNS_ERR_ARG;
    const double* src[6] = {vx, vy, vz, ox, oy, oz};
    const int st = flow_upload_state(h->fb, h->N, h->prec, src, h->stream);
    h->a_valid = false;
    h->phys_valid = false;
    return keep(h, st);
}

extern "C" int ddns_get_flow(ddns_handle* h, double* vx, double* vy, double* vz, double* ox, double* oy, double* oz)
{
    if (!h) return DDNS_ERR_ARG;
    double* dst[6] = {vx, vy, vz, ox, oy, oz};
    return keep(h, flow_download_state(h->fb, h->N, h->prec, dst, h->stream));
}

extern "C" int ddns_set_forcing_state(ddns_handle* h, const double* b, int iseedf, int iyf, const int* ivf)
{
    if (!h) return DDNS_ERR_ARG;
    int stt[34] = {0};
    stt[0] = iseedf; stt[1] = iyf;
    for (int i = 0; i < 32; ++i) stt[2 + i] = ivf ? ivf[i] : 0;
    DDNS_HIP(hipMemcpy(h->fb.sth_state, stt, sizeof(stt), hipMemcpyHostToDevice));
    if (b) DDNS_HIP(hipMemcpy(h->fb.sth_b, b, 450 * sizeof(double), hipMemcpyHostToDevice));
    return DDNS_OK;
}

extern "C" int ddns_get_forcing_state(ddns_handle* h, double* b, int* iseedf, int* iyf, int* ivf)
{
    if (!h) return DDNS_ERR_ARG;
    int stt[34];
    DDNS_HIP(hipStreamSynchronize(h->stream));
    DDNS_HIP(hipMemcpy(stt, h->fb.sth_state, sizeof(stt), hipMemcpyDeviceToHost));
    if (b) DDNS_HIP(hipMemcpy(b, h->fb.sth_b, 450 * sizeof(double), hipMemcpyDeviceToHost));
    if (iseedf) *iseedf = stt[0];
    if (iyf) *iyf = stt[1];
    if (ivf) for (int i = 0; i < 32; ++i) ivf[i] = stt[2 + i];
    return DDNS_OK;
}

// initial_flow (flowlib.f90:53-144) on the device: random solenoidal modes from a
// counter-based generator keyed by seed, scaled shell by shell; kind < 0 is the flow at
// rest of the stochastic forcing (turb.f90:441-449).  The history starts equal to the
// velocity (turb.f90:452-454).
extern "C" int ddns_init_flow(ddns_handle* h, int kind, int seed, double k0, double urms)
{
    if (!h) return DDNS_ERR_ARG;
    if (kind >= 0 && kind != DDNS_SPEC_K4_GAUSS && kind != DDNS_SPEC_KOLMOGOROV)
        return fail(h, DDNS_ERR_ARG, "unknown initial spectrum");
    if (kind == DDNS_SPEC_K4_GAUSS && !(k0 > 0.0)) return fail(h, DDNS_ERR_ARG, "k0 must be positive");
    int st;
    if (kind < 0) {
        const size_t bytes = (size_t)h->N * h->N * (h->N / 2) * (h->prec == DDNS_FP64 ? sizeof(double2) : sizeof(float2));
        for (int c = 0; c < 3; ++c) {
            DDNS_HIP(hipMemsetAsync(h->fb.U[c], 0, bytes, h->stream));
            DDNS_HIP(hipMemsetAsync(h->fb.O[c], 0, bytes, h->stream));
        }
        DDNS_HIP(hipStreamSynchronize(h->stream));
        st = 0;
    } else {
        st = flow_init(h->fb, h->N, h->nek, h->prec, kind, seed, k0, urms, h->p.force[0], h->stream);
    }
    h->a_valid = false;
    h->phys_valid = false;
    return keep(h, st);
}

extern "C" int ddns_get_physical_velocity(ddns_handle* h, double* ux, double* uy, double* uz)
{
    if (!h) return DDNS_ERR_ARG;
    if (!h->phys_valid) return fail(h, DDNS_ERR_ARG, "no step has run since the flow was set");
    return keep(h, flow_download_phys(h->fb, h->N, h->prec, ux, uy, uz, h->stream));
}

extern "C" int ddns_get_spectrum(ddns_handle* h, double* ek)
{
    if (!h || !ek) return DDNS_ERR_ARG;
    return keep(h, flow_spectrum(h->fb, h->N, h->nek, h->prec, ek, h->stream));
}

extern "C" int ddns_get_spectrum_device(ddns_handle* h, double* ek_dev, void* stream)
{
    if (!h || !ek_dev) return DDNS_ERR_ARG;
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    return keep(h, flow_spectrum_device(h->fb, h->N, h->nek, h->prec, h->dev, ek_dev, s));
}

extern "C" int ddns_get_spectra(ddns_handle* h, double* ek, double* dk, double* tk)
{
    if (!h) return DDNS_ERR_ARG;
    double mom[9];
    return keep(h, flow_diagnostics(h->fb, h->N, h->nek, h->p.rnu, h->prec, 0, ek, dk, tk, mom, h->stream));
}

// POSTPROC, postprocessing.f90:97-270: the sums come from the device, the rest is the
// reference's arithmetic on them
extern "C" int ddns_get_monitors(ddns_handle* h, int divstat, ddns_monitors* m)
{
    if (!h || !m) return DDNS_ERR_ARG;
    const int nek = h->nek, N = h->N;
    std::vector<double> ek(nek), dk(nek);
    double mom[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int st = flow_diagnostics(h->fb, N, nek, h->p.rnu, h->prec, 1 | (divstat ? 2 : 0), ek.data(), dk.data(),
                                    nullptr, mom, h->stream);
    if (st) return keep(h, st);
    const double n3 = (double)N * N * N, rnu = h->p.rnu;
    *m = ddns_monitors{};
    const double a2 = mom[1] / 3.0, a3 = mom[2] / 3.0, a4 = mom[3] / 3.0;          // 115-137
    m->vskew = a3 / n3 / std::pow(a2 / n3, 1.5);
    m->vflat = a4 / n3 / std::pow(a2 / n3, 2.0);
    double qq = 0.0, dissp = 0.0, xintls = 0.0;
    for (int i = 1; i <= nek; ++i) {                                                 // 160-170
        qq += ek[i - 1];
        dissp += dk[i - 1];
        xintls += ek[i - 1] / (double)i;
    }
    const double pi2 = 8.0 * std::atan(1.0);
    m->qq = qq;
    m->dissp = dissp;
    m->eta = std::pow(rnu * rnu * rnu / dissp, 0.25);                               // 190-202
    m->tk = std::sqrt(rnu / dissp);
    m->uprime = std::sqrt(2.0 * qq / 3.0);
    m->tmse = std::sqrt(15.0 * rnu * m->uprime * m->uprime / dissp);
    m->ree = m->uprime * m->tmse / rnu;
    m->xl = m->uprime * m->uprime * m->uprime / dissp;
    m->et = m->xl / m->uprime;
    const double akmax = (double)(N / 2) - 1.5;                                      // turb.f90:303
    m->resol = akmax * m->eta;
    m->xintls = xintls * (pi2 / 2.0) / 2.0 / (m->uprime * m->uprime);
    m->pmax = mom[0];
    m->cfl = mom[0] * h->p.dt / (pi2 / (double)N);                                   // 204-206
    if (divstat) {                                                                   // 229-266
        m->vdiv1 = mom[4] / n3;
        m->vdiv2 = std::sqrt((mom[5] - m->vdiv1 * m->vdiv1) / n3);
        m->vort1 = mom[6] / n3;
        m->vort2 = std::sqrt((mom[7] - m->vort1 * m->vort1) / n3);
        m->divvt = (mom[8] - m->vdiv1 * m->vort1) / (2.0 * n3);
    }
    return DDNS_OK;
}

// ------------------------------------------------------------------ droplets
extern "C" int ddns_set_droplets(ddns_handle* h, int n, const double* yp, const double* fvp, const double* fdp,
                                 const double* pertvel)
{
    if (!h || n < 0) return DDNS_ERR_ARG;
    if (n > h->pb.cap) return fail(h, DDNS_ERR_CAPACITY, "more droplets than capacity");
    const int cap = h->pb.cap;
    std::vector<double> pos(3 * (size_t)cap, 0.0), vel(3 * (size_t)cap, 0.0), rad(cap, 0.0), gid(cap, 0.0);
    std::vector<double> fv(9 * (size_t)cap, 0.0), fd(12 * (size_t)cap, 0.0), pv(3 * (size_t)cap, 0.0);
    for (int i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            pos[(size_t)c * cap + i] = yp[8 * (size_t)i + c];
            vel[(size_t)c * cap + i] = yp[8 * (size_t)i + 3 + c];
            if (pertvel) pv[3 * (size_t)i + c] = pertvel[3 * (size_t)i + c];
            for (int t = 0; t < 3; ++t) if (fvp) fv[(size_t)(3 * t + c) * cap + i] = fvp[9 * (size_t)i + 3 * t + c];
            for (int t = 0; t < 4; ++t) if (fdp) fd[(size_t)(3 * t + c) * cap + i] = fdp[12 * (size_t)i + 3 * t + c];
        }
        rad[i] = yp[8 * (size_t)i + 6];
        gid[i] = yp[8 * (size_t)i + 7];
    }
    DDNS_HIP(hipMemcpy(h->pb.pos, pos.data(), pos.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.vel, vel.data(), vel.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.rad, rad.data(), rad.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.gid, gid.data(), gid.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.fvp, fv.data(), fv.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.fdp, fd.data(), fd.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.pert, pv.data(), pv.size() * 8, hipMemcpyHostToDevice));
    h->pb.hv_v = h->pb.hv_d = 0;                 // the levels were written in order
    h->pb.v0_stale = h->pb.d0_stale = false;
    invalidate_cells(h->pb);
    h->n = n;
    h->pc.n = n;
    return DDNS_OK;
}

extern "C" int ddns_get_droplets(ddns_handle* h, int* n, double* yp, double* fvp, double* fdp, double* pertvel,
                                 double* vp, double* dmove)
{
    if (!h) return DDNS_ERR_ARG;
    DDNS_HIP(hipStreamSynchronize(h->stream));
    const int cap = h->pb.cap, m = h->n;
    if (n) *n = m;
    std::vector<double> pos(3 * (size_t)cap), vel(3 * (size_t)cap), rad(cap), gid(cap);
    std::vector<double> fv(9 * (size_t)cap), fd(12 * (size_t)cap), pv(3 * (size_t)cap), vv(3 * (size_t)cap), dm(3 * (size_t)cap);
    DDNS_HIP(hipMemcpy(pos.data(), h->pb.pos, pos.size() * 8, hipMemcpyDeviceToHost));
    DDNS_HIP(hipMemcpy(vel.data(), h->pb.vel, vel.size() * 8, hipMemcpyDeviceToHost));
    DDNS_HIP(hipMemcpy(rad.data(), h->pb.rad, rad.size() * 8, hipMemcpyDeviceToHost));
    DDNS_HIP(hipMemcpy(gid.data(), h->pb.gid, gid.size() * 8, hipMemcpyDeviceToHost));
    DDNS_HIP(hipMemcpy(fv.data(), h->pb.fvp, fv.size() * 8, hipMemcpyDeviceToHost));
    DDNS_HIP(hipMemcpy(fd.data(), h->pb.fdp, fd.size() * 8, hipMemcpyDeviceToHost));
    DDNS_HIP(hipMemcpy(pv.data(), h->pb.pert, pv.size() * 8, hipMemcpyDeviceToHost));
    DDNS_HIP(hipMemcpy(vv.data(), h->pb.vp, vv.size() * 8, hipMemcpyDeviceToHost));
    DDNS_HIP(hipMemcpy(dm.data(), h->pb.dmove, dm.size() * 8, hipMemcpyDeviceToHost));
    // history level t in its ring column (level 0 reads as level 1 after a shift)
    int vc[3], dc[4];
    for (int t = 0; t < 3; ++t) vc[t] = h->pb.vcol(t == 0 && h->pb.v0_stale ? 1 : t);
    for (int t = 0; t < 4; ++t) dc[t] = h->pb.dcol(t == 0 && h->pb.d0_stale ? 1 : t);
    for (int i = 0; i < m; ++i) {
        for (int c = 0; c < 3; ++c) {
            if (yp) { yp[8 * (size_t)i + c] = pos[(size_t)c * cap + i]; yp[8 * (size_t)i + 3 + c] = vel[(size_t)c * cap + i]; }
            if (pertvel) pertvel[3 * (size_t)i + c] = pv[3 * (size_t)i + c];
            if (vp) vp[3 * (size_t)i + c] = vv[(size_t)c * cap + i];
            if (dmove) dmove[3 * (size_t)i + c] = dm[(size_t)c * cap + i];
            for (int t = 0; t < 3; ++t) if (fvp) fvp[9 * (size_t)i + 3 * t + c] = fv[(size_t)(vc[t] + c) * cap + i];
            for (int t = 0; t < 4; ++t) if (fdp) fdp[12 * (size_t)i + 3 * t + c] = fd[(size_t)(dc[t] + c) * cap + i];
        }
        if (yp) { yp[8 * (size_t)i + 6] = rad[i]; yp[8 * (size_t)i + 7] = gid[i]; }
    }
    return DDNS_OK;
}

extern "C" int ddns_overlap_fix(ddns_handle* h)
{
    if (!h) return DDNS_ERR_ARG;
    int sw = 0, rl = 0;
    const int st = part_overlap(h->pb, h->pc, &sw, &rl, h->stream);
    h->cc.noverlap_sweeps = sw;
    h->cc.noverlap_relocated = rl;
    return keep(h, st);
}

// PART_INIT positions and radii (partlib.f90:227-281), then PERIODIC + OVERLAP_DET.
extern "C" int ddns_init_droplets(ddns_handle* h, int seed, int radii_kind, double median_um, double sigma_ln,
                                  double rmin_um, double rmax_um)
{
    if (!h) return DDNS_ERR_ARG;
    const int npart = h->p.npart;
    if (npart > h->pb.cap) return fail(h, DDNS_ERR_CAPACITY, "npart exceeds capacity");
    const double pi2 = h->pc.pi2;
    const double etk_CLOUD = std::sqrt(std::sqrt(h->p.viscosity_CLOUD * h->p.viscosity_CLOUD * h->p.viscosity_CLOUD / h->p.ediss_CLOUD));
    const double scale = (h->p.etk_DNS / etk_CLOUD) / 10000.0;
    std::vector<double> yp(8 * (size_t)std::max(npart, 1), 0.0);
    int idum = -seed, iy = 0, iv[32] = {0};
    for (int ih = 0; ih < npart; ++ih) {
        const double l1 = host_ranff(&idum, &iy, iv), l2 = host_ranff(&idum, &iy, iv), l3 = host_ranff(&idum, &iy, iv);
        yp[8 * (size_t)ih + 0] = pi2 * l1;
        yp[8 * (size_t)ih + 1] = pi2 * l2;
        yp[8 * (size_t)ih + 2] = pi2 * l3;
        yp[8 * (size_t)ih + 7] = (double)(ih + 1);
    }
    if (radii_kind == DDNS_RADII_LOGNORMAL) {
        int is2 = -(seed + 1), iy2 = 0, iv2[32] = {0};
        for (int ih = 0; ih < npart; ih += 2) {
            const double u1 = host_ranff(&is2, &iy2, iv2), u2 = host_ranff(&is2, &iy2, iv2);
            const double rr = std::sqrt(-2.0 * std::log(u1)), th = 2.0 * M_PI * u2;
            const double g[2] = {rr * std::cos(th), rr * std::sin(th)};
            for (int q = 0; q < 2 && ih + q < npart; ++q) {
                double r = median_um * std::exp(sigma_ln * g[q]);
                r = std::min(std::max(r, rmin_um), rmax_um);
                yp[8 * (size_t)(ih + q) + 6] = r * scale;
            }
        }
    } else {
        for (int ih = 0; ih < npart; ++ih) {
            const int i1 = (int)std::floor((double)ih * (double)h->p.nset / (double)npart) + 1;
            yp[8 * (size_t)ih + 6] = h->pc.radius[i1 - 1];
        }
    }
    for (int ih = 0; ih < npart; ++ih)                                        // PERIODIC, turb.f90:427
        for (int c = 0; c < 3; ++c) {
            double& y = yp[8 * (size_t)ih + c];
            if (y >= pi2) y = y - pi2;
            if (y < 0.0) y = y + pi2;
        }
    int st = ddns_set_droplets(h, npart, yp.data(), nullptr, nullptr, nullptr);
    if (st) return st;
    if (h->p.nonoverlap) return ddns_overlap_fix(h);                          // turb.f90:429
    return DDNS_OK;
}

// ------------------------------------------------------------------ the step
// Cell widths of the pair searches.  A collision-search walk visits the neighbour cells,
// so a cell must be as wide as a droplet's reach: the statistics window 10 (a + a_max) <=
// 20 a_max, or the collision window 2 a_max plus both displacements of the step.  Two grid
// spacings cover any step's displacement, and so does the statistics window at the usual
// step sizes: where the droplets are dense the cells are only as wide as 20 a_max, with at
// most about one (Morton-numbered) cell per four droplets so that the cell scan stays
// small; the search reports any droplet whose reach exceeds the cells (collide_regrid).
// HI needs its truncation radius, usually much wider, so it gets its own, coarser grid.
static double coll_cell_width(const ddns_handle* h)
{
    const double dx = h->pc.pi2 / h->N;
    const double wide = std::max(20.0 * h->pc.radmax, 2.0 * dx);
    if (h->coll_wide) return wide;
    static const long long per = [] {
        const char* e = std::getenv("DDNS_COLL_CELLS");     // Morton cells allowed per droplet
        const long long v = e ? std::atoll(e) : 4;
        return v > 0 ? v : 4;
    }();
    int b = 0;
    while (b < 8 && (1LL << (3 * (b + 1))) <= per * std::max(1, h->n)) ++b;
    const double fine = std::max(20.0 * h->pc.radmax * (1.0 + 1e-6), h->pc.pi2 / (double)(1 << b));
    return std::min(wide, fine);
}

static double cell_width(const ddns_handle* h)
{
    double w = coll_cell_width(h);
    if (h->p.hdi_included) w = std::max(w, h->pc.d_hdi);
    return w;
}

// Part `part` of the flow step (FlowPart: the inverse stages, the forward stages and
// update, or one of the two pieces of the latter) with the timing events of parity q.
// In graph mode the part is captured once into a graph (its kernels and their events)
// and replayed with one launch: the flow step takes no decision on the host and its
// buffers never move, so every step replays the same graph.  A failed capture turns
// graph mode off for the handle and the kernels run one by one.
static int run_flow_half(ddns_handle* h, int part, int q)
{
    hipStream_t s = h->stream;
    if (h->graph_mode) {
        hipGraphExec_t& gx = h->gflow[part][q];
        if (!gx) {
            hipGraph_t g = nullptr;
            int st = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess ? 0 : DDNS_ERR_HIP;
            if (!st) {
                st = flow_step(h->fb, h->fc, h->prec, part, s, h->fev[q], true);
                const hipError_t e = hipStreamEndCapture(s, &g);
                if (!st && e != hipSuccess) st = DDNS_ERR_HIP;
            }
            if (!st && hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) != hipSuccess) {
                gx = nullptr;
                st = DDNS_ERR_HIP;
            }
            if (g) (void)hipGraphDestroy(g);
            if (st) {
                (void)hipGetLastError();
                h->graph_mode = 0;
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 2; ++b)
                        if (h->gflow[a][b]) { (void)hipGraphExecDestroy(h->gflow[a][b]); h->gflow[a][b] = nullptr; }
            }
        }
        if (gx) {
            DDNS_HIP(hipGraphLaunch(gx, s));
            return 0;
        }
    }
    return flow_step(h->fb, h->fc, h->prec, part, s, h->fev[q]);
}

// The first half of a step's flow (the curl and inverse z transforms when the state
// is new, turb.f90:506-526; the inverse y and x stages and the physical velocity,
// 527-539) on the flow stream, with the event set of parity q.
static int queue_flow_first(ddns_handle* h, int q)
{
    hipStream_t s = h->stream;
    DDNS_HIP(hipEventRecord(h->flev[q][0], s));
    if (!h->a_valid) {
        const int st = flow_prologue(h->fb, h->fc, h->prec, s);
        if (st) return st;
        h->a_valid = true;
    }
    const int st = run_flow_half(h, FLOW_INV, q);
    if (st) return st;
    DDNS_HIP(hipEventRecord(h->flev[q][1], s));
    return 0;
}

// The forward transforms and the update of the step (turb.f90:644-691) on the flow
// stream.  The droplet block reads only the physical velocity of the first half and the
// flow reads nothing of the droplets, so this may run beside any part of the block; it
// is queued to start when the HI solve is done (event after): the collision search,
// the pair work and the overlap sweep that follow are latency-bound and leave HBM to it,
// while the GMRES kernels before stream the Krylov basis themselves.  With z_gate only
// the forward y stage and pre are queued here, and queue_flow_z adds the z stage later.
static int queue_flow_next(ddns_handle* h, hipEvent_t after);
static int queue_flow_z(ddns_handle* h, hipEvent_t after);

static int queue_flow_rest(ddns_handle* h, hipEvent_t after)
{
    if (h->flow_rest_queued) return h->z_pending ? queue_flow_z(h, nullptr) : 0;
    hipStream_t s = h->stream;
    const int q = h->par;
    if (after) DDNS_HIP(hipStreamWaitEvent(s, after, 0));
    DDNS_HIP(hipEventRecord(h->flev[q][2], s));
    h->flow_rest_queued = true;
    if (h->z_gate && after) {
        h->z_pending = true;
        return run_flow_half(h, FLOW_FWD_Y, q);
    }
    int st = run_flow_half(h, FLOW_FWD, q);
    if (st) return st;
    DDNS_HIP(hipEventRecord(h->flev[q][3], s));
    return queue_flow_next(h, after);
}

// The z stage and the symmetry pass of a second half queued in two pieces, after event
// `after` of the droplet stream (the collision-pair launch); the droplet block queues
// the next step's first half after it with queue_flow_next.  A pair kernel queued while
// the z stage ran made no progress until the z stage ended (35 us alone, 400-460 us
// then, in a quarter of the steps at 256^3 with 1M droplets); forcing the order cost the
// flow less than that on average.
static int queue_flow_z(ddns_handle* h, hipEvent_t after)
{
    if (!h->z_pending) return 0;
    hipStream_t s = h->stream;
    const int q = h->par;
    if (after) DDNS_HIP(hipStreamWaitEvent(s, after, 0));
    int st = run_flow_half(h, FLOW_FWD_Z, q);
    if (st) return st;
    DDNS_HIP(hipEventRecord(h->flev[q][3], s));
    h->z_pending = false;
    return 0;
}

// The next step's first half can follow at once: it overwrites the physical velocity,
// and the gather that reads it comes before the event the second half waited for.  It
// then runs beside the rest of this droplet block.
static int queue_flow_next(ddns_handle* h, hipEvent_t after)
{
    if (h->want_next && after) {
        const int st = queue_flow_first(h, 1 - h->par);
        if (st) return st;
        h->next_queued = true;
    }
    return 0;
}

// The droplet block (turb.f90:546-623) in three pieces: droplet_first up to the gather
// (binning, history, interpolation), the HI solve, droplet_rest from the advance on.
// A batch runs the HI solves of all its realisations together between the two.
static int droplet_first(ddns_handle* h)
{
    PartConst& pc = h->pc;
    pc.n = h->n;
    pc.new_particle = h->new_particle;
    hipStream_t s = h->dstream;
    int st;
    DDNS_HIP(hipEventRecord(h->ev[8], s));
    st = part_build_cells(h->pb, pc, cell_width(h), s);                       // HEADLIST (binning)
    if (st) return st;
    DDNS_HIP(hipEventRecord(h->ev[9], s));
    // PART_HISTORY_FVP (564) reads no interpolated velocity, so it runs first: the
    // interpolation then also gathers the HI inputs from the final droplet velocities
    st = part_history_fvp(h->pb, pc, s);                                      // 564
    if (st) return st;
    // the binning and the history ran beside the flow's first half; the gather needs
    // the physical velocity it writes
    DDNS_HIP(hipStreamWaitEvent(s, h->member ? h->phys_ev : h->flev[h->par][1], 0));
    st = part_interp(h->pb, pc, h->fb.uphys, s);                              // 553-556
    if (st) return st;
    DDNS_HIP(hipEventRecord(h->ev[10], s));
    h->cc = ddns_collision_counts{};
    return 0;
}

static int droplet_rest(ddns_handle* h, float* ms);

// The collision search found a droplet whose reach exceeds the width of its cells (a
// displacement, or a radius grown by coalescence, beyond what the width allowed): the
// cells are made wider and the search runs again on the same positions.
static int collide_regrid(ddns_handle* h, hipStream_t s)
{
    for (int tries = 0; tries < 3; ++tries) {
        h->coll_wide = true;
        const double w = std::max(coll_cell_width(h), 2.0 * h->pb.hc);
        int st = part_build_cells(h->pb, h->pc, w, s);
        if (st) return st;
        st = part_collide_search(h->pb, h->pc, s);
        if (st) return st;
        st = part_collide_finish(h->pb, h->pc, s);
        if (st != DDNS_COLL_REGRID) return st;
    }
    return fail(h, DDNS_ERR_DOMAIN, "a droplet's collision reach exceeds the cells of the pair search");
}

static int droplet_block(ddns_handle* h, float* ms)
{
    int status = 0;
    int st = droplet_first(h);
    if (st) return st;
    if (h->p.hdi_included) {                                                  // 571-572
        int it = 0; double res = 0;
        st = part_hdi(h->pb, h->pc, h->p.gmres_maxiter, h->p.gmres_tol, &it, &res, h->dstream);
        if (st < 0) return st;
        if (st > 0) status = st;
        h->cc.gmres_iters = it;
        h->cc.gmres_residual = res;
    }
    st = droplet_rest(h, ms);
    return st < 0 ? st : status;
}

static int droplet_rest(ddns_handle* h, float* ms)
{
    PartConst& pc = h->pc;
    hipStream_t s = h->dstream;
    int status = 0, st;
    DDNS_HIP(hipEventRecord(h->ev[11], s));
    // The droplet stream's next work is queued before the flow's launches, so it never
    // waits for the host: the advance, the collision grid, the collision search and the
    // pair kernel.  The flow's forward y stage and pre start when the HI solve is done
    // (event 11), its z stage when the pair kernel has run (event 6); both are queued while
    // the search runs, ahead of the host's check of the search's counts.
    st = part_advance(h->pb, pc, s);                                          // 574-576
    if (st) return st;
    h->new_particle = 0;
    pc.new_particle = 0;
    DDNS_HIP(hipEventRecord(h->ev[12], s));
    if (h->p.hdi_included && std::min(256, (int)std::floor(pc.pi2 / coll_cell_width(h))) > h->pb.G) {
        // finer grid for the collision search; the grid of the last overlap sweep has
        // this width and the same positions, and is swapped in when it is still valid
        st = part_build_cells(h->pb, pc, coll_cell_width(h), s);
        if (st) return st;
    }
    st = part_collide_search(h->pb, pc, s);                                   // 584, search and pairs
    if (st) return st;
    if (h->z_gate) DDNS_HIP(hipEventRecord(h->ev[6], s));                     // the pair kernel is queued
    if (!h->member) {                              // a batch queues its flow itself
        st = queue_flow_rest(h, h->ev[11]);                                   // after the HI solve
        if (st) return st;
    }
    const bool gated = h->z_pending;
    if (gated) {
        st = queue_flow_z(h, h->ev[6]);
        if (st) return st;
    }
    st = part_collide_finish(h->pb, pc, s);                                   // overflow check
    if (st == DDNS_COLL_REGRID) st = collide_regrid(h, s);
    if (st) return st;
    const bool reloc = h->p.nonoverlap && h->istep < h->p.nstep_grow;         // 590
    invalidate_cells(h->pb);
    st = part_update_and_wrap(h->pb, pc, reloc, s);                           // 586, 590, 608
    if (st) return st;
    if (h->p.nonoverlap && h->istep >= h->p.nstep_grow) {                     // 591: COALESCE
        // the wrap is per droplet and coalescence only copies whole droplets, so
        // the two commute; the removed droplets leave the end of the arrays
        int removed = 0;
        st = part_coalesce(h->pb, pc, &removed, s);
        if (st) return st;
        h->n -= removed;
        pc.n = h->n;
        if (h->n < h->p.npart / 2) return fail(h, DDNS_ERR_DROPLET_LOSS, "npart_t < npart/2 (turb.f90:600-604)");
    }
    DDNS_HIP(hipEventRecord(h->ev[13], s));
    // OVERLAP_DET (610): the first sweep is queued; its count is read with the step's
    // counts, and only a sweep that relocated droplets is followed by more
    if (h->p.nonoverlap) {
        st = part_overlap_queue(h->pb, pc, s, cell_width(h));
        if (st) return st;
    }
    st = part_queue_counts(h->pb, s);
    if (st) return st;
    DDNS_HIP(hipEventRecord(h->ev[14], s));
    // the next step's first half, queued on the flow stream behind the z stage once the
    // droplet block's work is queued: a graph launch holds the host for tens of us, which
    // the droplet stream now spends on the overlap sweep instead of idling
    if (gated) {
        st = queue_flow_next(h, h->ev[6]);
        if (st) return st;
    }
    DDNS_HIP(spin_wait(h->ev[14]));
    if (h->p.nonoverlap) {
        int sw = 0, rl = 0;
        st = part_overlap_finish(h->pb, pc, &sw, &rl, s);
        if (st) return st;
        h->cc.noverlap_sweeps = sw;
        h->cc.noverlap_relocated = rl;
        if (sw > 0) {                    // the further sweeps belong to the overlap phase
            DDNS_HIP(hipEventRecord(h->ev[14], s));
            DDNS_HIP(spin_wait(h->ev[14]));
        }
    }
    float t;
    DDNS_HIP(hipEventElapsedTime(&t, h->ev[8], h->ev[9])); ms[0] += t;
    DDNS_HIP(hipEventElapsedTime(&t, h->ev[9], h->ev[10])); ms[1] += t;
    DDNS_HIP(hipEventElapsedTime(&t, h->ev[10], h->ev[11])); ms[2] += t;
    DDNS_HIP(hipEventElapsedTime(&t, h->ev[11], h->ev[12])); ms[3] += t;
    DDNS_HIP(hipEventElapsedTime(&t, h->ev[12], h->ev[13])); ms[4] += t;
    DDNS_HIP(hipEventElapsedTime(&t, h->ev[13], h->ev[14])); ms[5] += t;
    ddns_collision_counts c2{};
    st = part_take_counts(h->pb, pc, &c2);
    if (st) return st;
    for (int k = 0; k < 3; ++k) h->cc.icc[k] = c2.icc[k];
    std::memcpy(h->cc.colltype, c2.colltype, sizeof(c2.colltype));
    std::memcpy(h->cc.nir, c2.nir, sizeof(c2.nir));
    h->cc.npairs = c2.npairs;
    h->cc.ncoll = c2.ncoll;
    h->cc.nset = pc.nset;
    h->cc.pairs_tested = c2.pairs_tested;
    h->cc.pairs_in_dstat = c2.pairs_in_dstat;
    return status;
}

// A step that stops part-way leaves the inverse pipeline A partly transformed in place
// (the inverse y stage overwrites its input) and may leave the next step's first half
// queued.  Wait for both streams and have the next call rebuild A from u-hat.  The flow
// state is then at the end of the failed step when its update had been queued, else at
// its start; time and istep count the failed step either way.
static int step_fail(ddns_handle* h, int st)
{
    (void)hipStreamSynchronize(h->dstream);
    (void)hipStreamSynchronize(h->stream);
    h->a_valid = false;
    h->next_queued = false;
    h->flow_rest_queued = false;
    h->z_pending = false;
    return keep(h, st);
}

extern "C" int ddns_step(ddns_handle* h, int nsteps)
{
    if (!h || nsteps < 0) return DDNS_ERR_ARG;
    if (h->member) return fail(h, DDNS_ERR_ARG, "a batch member steps with its batch (ddns_batch_step)");
    hipStream_t s = h->stream;
    int status = 0, st;
    float pms[6] = {0, 0, 0, 0, 0, 0};
    double flow_ms = 0;
    double kms[6] = {0, 0, 0, 0, 0, 0};
    DDNS_HIP(hipEventRecord(h->ev[15], s));
    h->next_queued = false;
    const bool flow_only_graph = h->graph_mode && !(h->p.npart > 0 && h->n > 0);
    for (int k = 0; k < nsteps; ++k) {
        h->istep += 1;                                                        // turb.f90:492
        // inverse y, x stage (the first half may have been queued during the last step)
        if (!h->next_queued) {
            st = queue_flow_first(h, h->par);
            if (st) return step_fail(h, st);
        }
        h->next_queued = false;
        h->phys_valid = true;
        // the droplet block (546-623) on its own stream; it queues the forward transforms
        // and the update (644-691) on this one when it reaches its latency-bound part, and
        // then the first half of the next step
        h->flow_rest_queued = false;
        h->want_next = k + 1 < nsteps;
        if (h->p.npart > 0 && h->n > 0) {
            st = droplet_block(h, pms);
            if (st < 0) return step_fail(h, st);
            if (st > 0 && status == 0) status = st;
            if (h->n < h->p.npart / 2)
                return step_fail(h, fail(h, DDNS_ERR_DROPLET_LOSS, "npart_t < npart/2 (turb.f90:600-604)"));
        }
        st = queue_flow_rest(h, nullptr);
        if (st) return step_fail(h, st);
        const int q = h->par;
        // A flow-only step replayed from graphs is not waited for: the call waits once, at
        // its end, and reports its whole time as flow time, without the per-kernel split.
        if (!flow_only_graph) {
            DDNS_HIP(hipEventSynchronize(h->flev[q][3]));
            const int pairs[6][2] = {{0, 1}, {1, 2}, {3, 4}, {4, 5}, {8, 6}, {6, 7}};
            for (int i = 0; i < 6; ++i) {
                float t;
                DDNS_HIP(hipEventElapsedTime(&t, h->fev[q][pairs[i][0]], h->fev[q][pairs[i][1]]));
                kms[i] += t;
            }
            float t1, t2;
            DDNS_HIP(hipEventElapsedTime(&t1, h->flev[q][0], h->flev[q][1]));
            DDNS_HIP(hipEventElapsedTime(&t2, h->flev[q][2], h->flev[q][3]));
            flow_ms += t1 + t2;
        }
        h->par = 1 - q;
        h->time += h->p.dt;                                                   // 693
    }
    DDNS_HIP(hipEventRecord(h->ev[7], s));
    DDNS_HIP(hipEventSynchronize(h->ev[7]));
    float tot = 0;
    DDNS_HIP(hipEventElapsedTime(&tot, h->ev[15], h->ev[7]));
    h->timing = ddns_timing{};
    h->timing.flow_ms = flow_only_graph ? tot : flow_ms;
    h->timing.cells_ms = pms[0];
    h->timing.interp_ms = pms[1];
    h->timing.hdi_ms = pms[2];
    h->timing.advance_ms = pms[3];
    h->timing.colldet_ms = pms[4];
    h->timing.overlap_ms = pms[5];
    h->timing.total_ms = tot;
    h->timing.nsteps = nsteps;
    for (int q = 0; q < 6; ++q) h->timing.flow_kernel_ms[q] = kms[q];
    return status;
}

extern "C" int ddns_get_collision_counts(ddns_handle* h, ddns_collision_counts* out)
{
    if (!h || !out) return DDNS_ERR_ARG;
    *out = h->cc;
    return DDNS_OK;
}

extern "C" int ddns_get_collision_bins(ddns_handle* h, int* nirbin, double* wrbin, double* wxbin, double* wybin,
                                       double* wzbin, double* urbin, double* thetabin, double* wrtanbin)
{
    if (!h) return DDNS_ERR_ARG;
    double* const b[7] = {wrbin, wxbin, wybin, wzbin, urbin, thetabin, wrtanbin};
    return keep(h, part_read_bins(h->pb, h->pc, nirbin, b, h->stream));
}

extern "C" int ddns_get_collision_pairs(ddns_handle* h, int64_t* triples, int cap, int* npairs)
{
    if (!h || !npairs) return DDNS_ERR_ARG;
    return keep(h, part_read_pairs(h->pb, h->pc, triples, cap, npairs, h->stream));
}

extern "C" int ddns_get_contact_records(ddns_handle* h, double* rec, int cap, int* n)
{
    if (!h || !n || cap < 0) return DDNS_ERR_ARG;
    return keep(h, part_read_contacts(h->pb, rec, cap, n, h->stream));
}

extern "C" int ddns_get_tuning(ddns_handle* h, int* shapes, double* ms)
{
    if (!h) return DDNS_ERR_ARG;
    if (shapes) {
        shapes[0] = h->fb.ky_tc;
        shapes[1] = flow_default_ky_tc(h->N, h->prec);
        shapes[2] = h->fb.kz_nt;
        shapes[3] = flow_default_kz_nt(h->N, h->prec);
    }
    if (ms)
        for (int k = 0; k < 3; ++k) {
            ms[k] = h->fb.ky_ms[k];
            ms[3 + k] = h->fb.kz_ms[k];
        }
    return DDNS_OK;
}

extern "C" int ddns_get_timing(ddns_handle* h, ddns_timing* out)
{
    if (!h || !out) return DDNS_ERR_ARG;
    *out = h->timing;
    return DDNS_OK;
}

// ------------------------------------------------------------------ kernel-level entry points
extern "C" int ddns_fft_r2c(ddns_handle* h, double* v)
{
    if (!h || !v) return DDNS_ERR_ARG;
    h->a_valid = false;
    return keep(h, fft_r2c_host(h->fb, h->N, h->prec, v, h->stream));
}

extern "C" int ddns_fft_c2r(ddns_handle* h, double* v)
{
    if (!h || !v) return DDNS_ERR_ARG;
    h->a_valid = false;
    return keep(h, fft_c2r_host(h->fb, h->N, h->prec, v, h->stream));
}

extern "C" int ddns_collision_detect(ddns_handle* h, int n, const double* yp, const double* dmove,
                                     const double* fvp0, const double* vnew, const double* vp)
{
    if (!h || n < 0 || !yp || !dmove || !fvp0 || !vnew || !vp) return DDNS_ERR_ARG;
    if (n > h->pb.cap) return fail(h, DDNS_ERR_CAPACITY, "more droplets than capacity");
    int st = ddns_set_droplets(h, n, yp, nullptr, nullptr, nullptr);
    if (st) return st;
    const int cap = h->pb.cap;
    std::vector<double> a(3 * (size_t)cap, 0.0), b(3 * (size_t)cap, 0.0), c(3 * (size_t)cap, 0.0), d(3 * (size_t)cap, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            a[(size_t)k * cap + i] = dmove[3 * (size_t)i + k];
            b[(size_t)k * cap + i] = fvp0[3 * (size_t)i + k];
            c[(size_t)k * cap + i] = vnew[3 * (size_t)i + k];
            d[(size_t)k * cap + i] = vp[3 * (size_t)i + k];
        }
    DDNS_HIP(hipMemcpy(h->pb.dmove, a.data(), a.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.fvp, b.data(), b.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.vel, c.data(), c.size() * 8, hipMemcpyHostToDevice));
    DDNS_HIP(hipMemcpy(h->pb.vp, d.data(), d.size() * 8, hipMemcpyHostToDevice));
    h->pc.n = n;
    st = part_build_cells(h->pb, h->pc, cell_width(h), h->stream);
    if (st) return keep(h, st);
    st = part_collide(h->pb, h->pc, h->stream);
    if (st == DDNS_COLL_REGRID) st = collide_regrid(h, h->stream);
    if (st) return keep(h, st);
    ddns_collision_counts c2{};
    st = part_read_counts(h->pb, h->pc, &c2, h->stream);
    if (st) return keep(h, st);
    h->cc = c2;
    return DDNS_OK;
}

extern "C" int ddns_profile_flow(ddns_handle* h, int k, double* ms_per_kernel, int nkernels)
{
    if (!h || k <= 0 || !ms_per_kernel) return DDNS_ERR_ARG;
    if (!h->a_valid) {
        const int st = flow_prologue(h->fb, h->fc, h->prec, h->stream);
        if (st) return keep(h, st);
        h->a_valid = true;
    }
    return keep(h, flow_profile(h->fb, h->fc, h->prec, k, ms_per_kernel, nkernels, h->stream));
}

extern "C" int ddns_profile_copy(ddns_handle* h, int reps, double* ms_per_launch, double* bytes_per_launch)
{
    if (!h || reps <= 0 || !ms_per_launch || !bytes_per_launch) return DDNS_ERR_ARG;
    // the copy overwrites the inverse pipeline, which the next step would start from
    h->a_valid = false;
    return keep(h, flow_profile_copy(h->fb, h->N, h->prec, reps, ms_per_launch, bytes_per_launch, h->stream));
}

// ------------------------------------------------------------------ batches
// R independent realisations of the same grid, precision, time step and viscosity (the
// reference runs each as its own job; turb.f90:492-707 is the step of one).  The batch
// holds the flow fields of all of them side by side and runs every flow stage as one
// launch over all realisations, replayed from two HIP graphs per step.  Each realisation
// is a member handle: a view of its flow fields plus its own droplets, forcing kind and
// random streams, reached through ddns_batch_member for state transfers and diagnostics.
// A step queues the batch's inverse stages, then its forward stages and update (they
// touch nothing the droplets read), and runs the members' droplet blocks on worker
// threads, each on its own streams, beside the flow.
struct ddns_batch {
    int R = 0, N = 0, prec = 0, nek = 0;
    std::vector<ddns_handle*> m;
    FlowBuffers fb;
    FlowConst fc{};
    hipStream_t stream = nullptr;
    hipEvent_t flev[4] = {};               // inverse half start, end; forward half start, end
    hipEvent_t fev[FLOW_EVENTS] = {};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipGraphExec_t gflow[2] = {};          // FLOW_INV, FLOW_FWD
    int graph_mode = 1;
    int threads = 8;                       // worker threads of the droplet blocks
    // the HI solves of all realisations as one batched GMRES (DDNS_BATCH_HI=0: each
    // realisation solves its own inside its droplet block)
    bool merged_hi = true;
    hipStream_t hstream = nullptr;
    hipEvent_t hi_ev = nullptr;
    HdiBatch hb;
    ddns_timing timing{};
    std::string err;
};

static int bfail(ddns_batch* b, int code, const std::string& msg)
{
    set_error(msg);
    if (b) b->err = msg;
    return code;
}

extern "C" int ddns_batch_destroy(ddns_batch* b)
{
    if (!b) return DDNS_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (ddns_handle* h : b->m)
        if (h) destroy_impl(h);
    flow_free(b->fb);
    for (auto& e : b->flev) if (e) (void)hipEventDestroy(e);
    for (auto& e : b->fev) if (e) (void)hipEventDestroy(e);
    if (b->t0) (void)hipEventDestroy(b->t0);
    if (b->t1) (void)hipEventDestroy(b->t1);
    for (auto& g : b->gflow) if (g) (void)hipGraphExecDestroy(g);
    if (b->hstream) (void)hipStreamSynchronize(b->hstream);
    part_hdi_batch_free(b->hb);
    if (b->hi_ev) (void)hipEventDestroy(b->hi_ev);
    if (b->hstream) (void)hipStreamDestroy(b->hstream);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return DDNS_OK;
}

extern "C" int ddns_batch_create(const ddns_params* params, int nreal, ddns_batch** out)
{
    if (!params || !out || nreal < 1) return bfail(nullptr, DDNS_ERR_ARG, "null argument or no realisation");
    *out = nullptr;
    const ddns_params& p0 = params[0];
    const int N = p0.nx;
    if (N > 256) return bfail(nullptr, DDNS_ERR_ARG, "batched realisations need N <= 256 (the fused z stage)");
    for (int r = 1; r < nreal; ++r) {
        const ddns_params& p = params[r];
        if (p.nx != N || p.ny != p0.ny || p.nz != p0.nz || p.precision != p0.precision || p.dt != p0.dt ||
            p.rnu != p0.rnu)
            return bfail(nullptr, DDNS_ERR_ARG, "the realisations of a batch share grid, precision, dt and viscosity");
    }
    ddns_batch* b = new ddns_batch();
    b->R = nreal;
    b->N = N;
    b->prec = p0.precision;
    b->nek = (int)(N / 2 - 1.5);
    if (const char* g = std::getenv("DDNS_GRAPH")) b->graph_mode = g[0] != '0';
    if (const char* t = std::getenv("DDNS_BATCH_THREADS")) b->threads = std::max(1, std::atoi(t));
    if (const char* t = std::getenv("DDNS_BATCH_HI")) b->merged_hi = t[0] != '0';
    int st = 0;
    if (hipStreamCreate(&b->stream) != hipSuccess || hipStreamCreate(&b->hstream) != hipSuccess ||
        hipEventCreateWithFlags(&b->hi_ev, hipEventDisableTiming) != hipSuccess) {
        ddns_batch_destroy(b);
        return bfail(nullptr, DDNS_ERR_HIP, "stream");
    }
    for (auto& e : b->flev) if (hipEventCreate(&e) != hipSuccess) st = DDNS_ERR_HIP;
    for (auto& e : b->fev) if (hipEventCreate(&e) != hipSuccess) st = DDNS_ERR_HIP;
    if (hipEventCreate(&b->t0) != hipSuccess || hipEventCreate(&b->t1) != hipSuccess) st = DDNS_ERR_HIP;
    if (st) { ddns_batch_destroy(b); return bfail(nullptr, st, "event"); }
    // the shape checks of ddns_create run on the first member below; the flow buffers of
    // all realisations, then the tile shapes timed on them
    if (N != p0.ny || N != p0.nz || N < 32 || (N & (N - 1))) { ddns_batch_destroy(b); return bfail(nullptr, DDNS_ERR_ARG, "grid must be cubic, a power of two, 32..256"); }
    if (b->prec != DDNS_FP64 && b->prec != DDNS_FP32_MIXED) { ddns_batch_destroy(b); return bfail(nullptr, DDNS_ERR_ARG, "precision must be DDNS_FP64 or DDNS_FP32_MIXED"); }
    st = flow_alloc(b->fb, N, b->prec, nreal);
    if (st) { ddns_batch_destroy(b); return st; }
    b->fc.N = N; b->fc.nek = b->nek; b->fc.forcing = -1; b->fc.dt = p0.dt; b->fc.rnu = p0.rnu;
    b->fc.pi2 = pi2_ref();
    st = flow_tune(b->fb, b->fc, b->prec, b->stream);
    if (st) { ddns_batch_destroy(b); return st; }
    b->m.assign(nreal, nullptr);
    for (int r = 0; r < nreal; ++r) {
        const FlowBuffers v = flow_view(b->fb, r, b->prec);
        st = create_impl(&params[r], &b->m[r], &v);
        if (st) { ddns_batch_destroy(b); return st; }
        const ddns_handle* h = b->m[r];
        st = flow_set_realisation(b->fb, r, h->p.forcing, h->p.force[0], h->p.force[1], h->p.ou_tf, h->p.ou_var);
        if (st) { ddns_batch_destroy(b); return st; }
    }
    *out = b;
    return DDNS_OK;
}

extern "C" ddns_handle* ddns_batch_member(ddns_batch* b, int r)
{
    if (!b || r < 0 || r >= b->R) return nullptr;
    return b->m[r];
}

extern "C" int ddns_batch_size(const ddns_batch* b) { return b ? b->R : 0; }

extern "C" const char* ddns_batch_last_error(const ddns_batch* b)
{
    if (b && !b->err.empty()) return b->err.c_str();
    return g_err.c_str();
}

// one half of the batch's flow step (FLOW_INV or FLOW_FWD), from its graph when it has one
static int batch_flow_half(ddns_batch* b, int part)
{
    hipStream_t s = b->stream;
    const int gi = part == FLOW_INV ? 0 : 1;
    if (b->graph_mode) {
        hipGraphExec_t& gx = b->gflow[gi];
        if (!gx) {
            hipGraph_t g = nullptr;
            int st = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess ? 0 : DDNS_ERR_HIP;
            if (!st) {
                st = flow_step(b->fb, b->fc, b->prec, part, s, b->fev, true);
                const hipError_t e = hipStreamEndCapture(s, &g);
                if (!st && e != hipSuccess) st = DDNS_ERR_HIP;
            }
            if (!st && hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) != hipSuccess) {
                gx = nullptr;
                st = DDNS_ERR_HIP;
            }
            if (g) (void)hipGraphDestroy(g);
            if (st) {
                (void)hipGetLastError();
                b->graph_mode = 0;
                for (auto& x : b->gflow) if (x) { (void)hipGraphExecDestroy(x); x = nullptr; }
            }
        }
        if (gx) {
            DDNS_HIP(hipGraphLaunch(gx, s));
            return 0;
        }
    }
    return flow_step(b->fb, b->fc, b->prec, part, s, b->fev);
}

// after a failed step: every stream idle, and the next step rebuilds the inverse pipeline
static int batch_fail(ddns_batch* b, int st)
{
    (void)hipStreamSynchronize(b->stream);
    for (ddns_handle* h : b->m) {
        (void)hipStreamSynchronize(h->dstream);
        (void)hipStreamSynchronize(h->stream);
        h->a_valid = false;
    }
    if (st != DDNS_OK && b->err.empty()) b->err = g_err;
    return st;
}

extern "C" int ddns_batch_step(ddns_batch* b, int nsteps)
{
    if (!b || nsteps < 0) return DDNS_ERR_ARG;
    b->err.clear();
    hipStream_t s = b->stream;
    const int R = b->R;
    int status = 0, st;
    double flow_ms = 0, kms[6] = {0, 0, 0, 0, 0, 0};
    std::vector<std::array<float, 6>> pms(R);
    for (auto& a : pms) a.fill(0.f);
    std::vector<int> rc(R, 0);
    DDNS_HIP(hipEventRecord(b->t0, s));
    for (int k = 0; k < nsteps; ++k) {
        bool prologue = false;
        for (ddns_handle* h : b->m) {
            h->istep += 1;                                                    // turb.f90:492
            if (!h->a_valid) prologue = true;
        }
        // the inverse pipeline of every realisation (rebuilt from u-hat when any member's
        // state was set since the last step), the inverse y and x stages and the physical
        // velocity; then the forward stages and the update, which read nothing the
        // droplet blocks write and write nothing they read
        DDNS_HIP(hipEventRecord(b->flev[0], s));
        if (prologue) {
            st = flow_prologue(b->fb, b->fc, b->prec, s);
            if (st) return batch_fail(b, st);
        }
        st = batch_flow_half(b, FLOW_INV);
        if (st) return batch_fail(b, st);
        DDNS_HIP(hipEventRecord(b->flev[1], s));
        DDNS_HIP(hipEventRecord(b->flev[2], s));
        st = batch_flow_half(b, FLOW_FWD);
        if (st) return batch_fail(b, st);
        DDNS_HIP(hipEventRecord(b->flev[3], s));
        for (ddns_handle* h : b->m) {
            h->a_valid = true;
            h->phys_valid = true;
            h->phys_ev = b->flev[1];
        }
        // the droplet blocks (546-623).  With the merged HI solve: each realisation's
        // binning, history, gather and HI start are queued from this thread (no host
        // waits), the GMRES iterations of all of them run as one batched solve, and the
        // rest of each block runs on the worker threads
        std::fill(rc.begin(), rc.end(), 0);
        std::vector<char> active(R, 0), merged(R, 0);
        for (int r = 0; r < R; ++r) {
            ddns_handle* h = b->m[r];
            h->cc = ddns_collision_counts{};
            active[r] = h->p.npart > 0 && h->n > 0;
        }
        if (b->merged_hi) {
            int maxiter = -1;
            std::vector<PartBuffers*> pbs;
            std::vector<const PartConst*> pcs;
            std::vector<int> idx;
            for (int r = 0; r < R; ++r) {
                if (!active[r]) continue;
                ddns_handle* h = b->m[r];
                st = droplet_first(h);
                if (st) { b->err = "realisation " + std::to_string(r) + ": " + ddns_last_error(h); return batch_fail(b, st); }
                if (!h->p.hdi_included) continue;
                if (maxiter < 0) maxiter = h->p.gmres_maxiter;
                if (h->p.gmres_maxiter != maxiter) continue;        // solved on its own below
                st = part_hdi_begin(h->pb, h->pc, maxiter, h->p.gmres_tol, h->dstream);
                if (st) { b->err = "realisation " + std::to_string(r) + ": " + g_err; return batch_fail(b, st); }
                DDNS_HIP(hipEventRecord(h->ev[4], h->dstream));
                DDNS_HIP(hipStreamWaitEvent(b->hstream, h->ev[4], 0));
                merged[r] = 1;
                pbs.push_back(&h->pb);
                pcs.push_back(&h->pc);
                idx.push_back(r);
            }
            const int nh = (int)idx.size();
            std::vector<int> its(nh), sts(nh), redo(nh);
            std::vector<double> res(nh);
            if (nh > 0) {
                st = part_hdi_batch(b->hb, pbs.data(), pcs.data(), nh, maxiter, its.data(), res.data(), sts.data(),
                                    redo.data(), b->hstream);
                if (st) {
                    for (int i = 0; i < nh; ++i)
                        if (sts[i] < 0) b->err = "realisation " + std::to_string(idx[i]) + ": " + g_err;
                    return batch_fail(b, st);
                }
                DDNS_HIP(hipEventRecord(b->hi_ev, b->hstream));
            }
            for (int i = 0; i < nh; ++i) {
                ddns_handle* h = b->m[idx[i]];
                DDNS_HIP(hipStreamWaitEvent(h->dstream, b->hi_ev, 0));
                if (redo[i]) {                 // lists widened: the whole solve again, alone
                    int it = 0; double rs = 0;
                    const int q = part_hdi(h->pb, h->pc, h->p.gmres_maxiter, h->p.gmres_tol, &it, &rs, h->dstream);
                    if (q < 0) { b->err = "realisation " + std::to_string(idx[i]) + ": " + g_err; return batch_fail(b, q); }
                    its[i] = it; res[i] = rs; sts[i] = q;
                }
                h->cc.gmres_iters = its[i];
                h->cc.gmres_residual = res[i];
                if (sts[i] > 0) rc[idx[i]] = sts[i];
            }
        }
        auto work = [&](int w, int nw) {
            for (int r = w; r < R; r += nw) {
                ddns_handle* h = b->m[r];
                if (!active[r]) continue;
                int q;
                if (!b->merged_hi) {
                    q = droplet_block(h, pms[r].data());
                } else {
                    q = 0;
                    if (h->p.hdi_included && !merged[r]) {
                        int it = 0; double rs = 0;
                        q = part_hdi(h->pb, h->pc, h->p.gmres_maxiter, h->p.gmres_tol, &it, &rs, h->dstream);
                        h->cc.gmres_iters = it;
                        h->cc.gmres_residual = rs;
                    }
                    const int warn = q > 0 ? q : rc[r];
                    if (q >= 0) q = droplet_rest(h, pms[r].data());
                    if (q >= 0) q = warn;
                }
                if (q >= 0 && h->n < h->p.npart / 2)
                    q = fail(h, DDNS_ERR_DROPLET_LOSS, "npart_t < npart/2 (turb.f90:600-604)");
                rc[r] = q;
            }
        };
        const int nw = std::min(b->threads, R);
        if (nw <= 1) {
            work(0, 1);
        } else {
            std::vector<std::thread> pool;
            for (int w = 0; w < nw; ++w) pool.emplace_back(work, w, nw);
            for (auto& t : pool) t.join();
        }
        for (int r = 0; r < R; ++r) {
            if (rc[r] < 0) {
                b->err = "realisation " + std::to_string(r) + ": " + ddns_last_error(b->m[r]);
                return batch_fail(b, rc[r]);
            }
            if (rc[r] > 0 && status == 0) status = rc[r];
        }
        // the next step's inverse half overwrites the physical velocity: it waits for
        // every droplet stream
        for (ddns_handle* h : b->m) {
            DDNS_HIP(hipEventRecord(h->ev[5], h->dstream));
            DDNS_HIP(hipStreamWaitEvent(s, h->ev[5], 0));
        }
        DDNS_HIP(hipEventSynchronize(b->flev[3]));
        const int pairs[6][2] = {{0, 1}, {1, 2}, {3, 4}, {4, 5}, {8, 6}, {6, 7}};
        for (int i = 0; i < 6; ++i) {
            float t;
            DDNS_HIP(hipEventElapsedTime(&t, b->fev[pairs[i][0]], b->fev[pairs[i][1]]));
            kms[i] += t;
        }
        float t1, t2;
        DDNS_HIP(hipEventElapsedTime(&t1, b->flev[0], b->flev[1]));
        DDNS_HIP(hipEventElapsedTime(&t2, b->flev[2], b->flev[3]));
        flow_ms += t1 + t2;
        for (ddns_handle* h : b->m) h->time += h->p.dt;                       // 693
    }
    DDNS_HIP(hipEventRecord(b->t1, s));
    DDNS_HIP(hipEventSynchronize(b->t1));
    float tot = 0;
    DDNS_HIP(hipEventElapsedTime(&tot, b->t0, b->t1));
    b->timing = ddns_timing{};
    b->timing.flow_ms = flow_ms;
    b->timing.total_ms = tot;
    b->timing.nsteps = nsteps;
    for (int q = 0; q < 6; ++q) b->timing.flow_kernel_ms[q] = kms[q];
    for (int r = 0; r < R; ++r) {
        ddns_handle* h = b->m[r];
        h->timing = ddns_timing{};
        h->timing.cells_ms = pms[r][0];
        h->timing.interp_ms = pms[r][1];
        h->timing.hdi_ms = pms[r][2];
        h->timing.advance_ms = pms[r][3];
        h->timing.colldet_ms = pms[r][4];
        h->timing.overlap_ms = pms[r][5];
        h->timing.nsteps = nsteps;
        // the batch-wide sums: droplet phases added over the realisations
        b->timing.cells_ms += pms[r][0];
        b->timing.interp_ms += pms[r][1];
        b->timing.hdi_ms += pms[r][2];
        b->timing.advance_ms += pms[r][3];
        b->timing.colldet_ms += pms[r][4];
        b->timing.overlap_ms += pms[r][5];
    }
    return status;
}

extern "C" int ddns_batch_get_timing(ddns_batch* b, ddns_timing* out)
{
    if (!b || !out) return DDNS_ERR_ARG;
    *out = b->timing;
    return DDNS_OK;
}

extern "C" int ddns_batch_get_tuning(ddns_batch* b, int* shapes, double* ms)
{
    if (!b) return DDNS_ERR_ARG;
    if (shapes) {
        shapes[0] = b->fb.ky_tc;
        shapes[1] = flow_default_ky_tc(b->N, b->prec);
        shapes[2] = b->fb.kz_nt;
        shapes[3] = flow_default_kz_nt(b->N, b->prec);
    }
    if (ms)
        for (int k = 0; k < 3; ++k) {
            ms[k] = b->fb.ky_ms[k];
            ms[3 + k] = b->fb.kz_ms[k];
        }
    return DDNS_OK;
}
