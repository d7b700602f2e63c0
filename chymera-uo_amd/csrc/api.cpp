// C-ABI of libchypot (include/chypot.h): context, SETBDY/BDYGEN/POT3 entry
// points and their host-pointer forms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "chypot_impl.h"

namespace chy {

static thread_local std::string g_last_error;

void set_device_error(hipError_t e, const char *where)
{
    g_last_error = std::string(where) + ": " + hipGetErrorString(e);
}

// a captured group's key: kGroupKeyHead argument words, then per field its
// grid, density, PHI, COEF and KLOCAT
constexpr size_t kGroupKeyHead = 8, kGroupKeyPerField = 5;
constexpr size_t kGroupGraphCache = 16;
// a captured batch's key: kBatchKeyHead words, then per item its grid,
// arrays, arguments, KLOCAT in and whether it wants diagnostics
constexpr size_t kBatchKeyHead = 6, kBatchKeyPerItem = 11;
constexpr size_t kBatchGraphCache = 4;

static int dev_rc(int hip_rc, const char *where)
{
    if (hip_rc == (int)hipSuccess) return CHY_OK;
    set_device_error((hipError_t)hip_rc, where);
    return CHY_EDEVICE;
}

void set_field(Fields &F, int f, const chy_grid *w, const void *rho, double *phi, double *coef, int klocat_in)
{
    F.rho[f] = rho;
    F.phi[f] = phi;
    F.coef[f] = coef;
    F.spec[f] = w->d.spec;
    F.yh[f] = w->d.yh;
    F.mom[f] = w->d.mom;
    F.chunk[f] = w->d.chunk;
    F.massio[f] = w->d.massio;
    F.dscal[f] = w->d.dscal;
    F.ticket[f] = w->d.ticket;
    F.zsamp[f] = w->d.zsamp;
    F.rstat[f] = w->d.rstat;
    F.klocat_in[f] = klocat_in;
}

Fields one_field(const chy_grid *g, const void *rho, double *phi, double *coef, int klocat_in)
{
    Fields F{};
    F.nf = 1;
    set_field(F, 0, g, rho, phi, coef, klocat_in);
    return F;
}

// Grids prepared from identical descriptions have identical device tables, so
// one launch can run fields of both (each on its own grid's workspace).
static bool same_tables(const chy_grid *a, const chy_grid *b)
{
    return a->J == b->J && a->K == b->K && a->L == b->L && a->isym == b->isym && a->dtheta == b->dtheta &&
           a->pi == b->pi && a->grav == b->grav && a->ncu == b->ncu && a->r == b->r && a->z == b->z &&
           a->rhf == b->rhf && a->zhf == b->zhf && a->mode_solver == b->mode_solver;
}

static void free_grid_memory(chy_grid *g)
{
    for (void *p : g->allocations) (void)hipFree(p);
    g->allocations.clear();
    if (g->d.h_rho) (void)hipFree(g->d.h_rho);
    if (g->d.h_phi) (void)hipFree(g->d.h_phi);
    if (g->d.h_coef) (void)hipFree(g->d.h_coef);
    g->d.h_rho = g->d.h_phi = g->d.h_coef = nullptr;
}

struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev)
    {
        (void)hipGetDevice(&prev);
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceScope()
    {
        int cur = -1;
        (void)hipGetDevice(&cur);
        if (prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

static size_t field_elems(const chy_grid *g) { return (size_t)(g->J + 2) * (g->K + 2) * g->L; }
static size_t coef_elems(const chy_grid *g) { return (size_t)(g->J + 2) * (g->K + 2) * (g->L / 2) * 2; }

static int check_pot3_args(const chy_grid *g, int &npoint, int coef_mode, const void *coef)
{
    if (npoint < 1) return CHY_EINVAL;
    if (npoint > 10) npoint = 10;  // pot3.f:328-333
    const int num = (npoint - npoint % 2) / 2;
    // ZAXPHI reads PHI(J,...) for J up to NUM+1 (NUM+2 for odd NPOINT)
    if (num + 1 + (npoint % 2) > g->J + 1) return CHY_EINVAL;
    if (coef_mode < CHY_COEF_NONE || coef_mode > CHY_COEF_FULL) return CHY_EINVAL;
    if (coef_mode != CHY_COEF_NONE && !coef) return CHY_EINVAL;
    return CHY_OK;
}

// ---- stage profiler ----
hipEvent_t Profiler::take()
{
    hipEvent_t e = nullptr;
    if (!pool.empty()) {
        e = pool.back();
        pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

void Profiler::reserve(size_t n)
{
    while (pool.size() < n) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        pool.push_back(e);
    }
}

void Profiler::release()
{
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
}

// A failed record leaves the stage untimed (and clears the failure, which
// would otherwise surface at the next launch check).
void Profiler::begin(int stage, hipStream_t s)
{
    if (!on) return;
    hipEvent_t e = take();
    if (!e) return;
    if (hipEventRecord(e, s) != hipSuccess) {
        (void)hipGetLastError();
        pool.push_back(e);
        return;
    }
    open_stage = stage;
    open_ev = e;
}

void Profiler::end(hipStream_t s)
{
    if (!on || open_stage < 0) return;
    hipEvent_t e = take();
    if (!e) return;
    if (hipEventRecord(e, s) != hipSuccess) {
        (void)hipGetLastError();
        pool.push_back(e);
        pool.push_back(open_ev);
    } else {
        pending.push_back(Rec{open_stage, open_ev, e});
    }
    open_stage = -1;
    open_ev = nullptr;
}

int Profiler::collect()
{
    int rc = CHY_OK;
    for (Rec &r : pending) {
        float t = 0.0f;
        hipError_t e = hipEventSynchronize(r.b);
        if (e == hipSuccess) e = hipEventElapsedTime(&t, r.a, r.b);
        if (e == hipSuccess) {
            ms[r.stage] += t;
            count[r.stage] += 1;
        } else {
            rc = dev_rc((int)e, "stage timing");
        }
        pool.push_back(r.a);
        pool.push_back(r.b);
    }
    pending.clear();
    return rc;
}

void Profiler::reset()
{
    collect();
    for (int i = 0; i < CHY_NSTAGES; ++i) { ms[i] = 0.0; count[i] = 0; }
}

static bool bdygen_args_ok(int maxtrm, int isym, double redge)
{
    return maxtrm >= 0 && maxtrm <= kMaxTrm && std::abs(isym) == 2 && redge >= 0.0;
}

// Waits for a graph's last launch (its done event), then destroys the event.
// Should the wait fail, the whole device is waited for instead, and the
// failure is cleared: HIP keeps a failed call's error for the next
// hipGetLastError on the thread, where a later launch check (ours or the
// caller's) would take it for its own.
static void wait_done(hipEvent_t done)
{
    if (!done) return;
    if (hipEventSynchronize(done) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipDeviceSynchronize();
        (void)hipGetLastError();
    }
    (void)hipEventDestroy(done);
}

// waits for the entry's last launch, then frees it
static void retire_graph(chy_grid::GraphEntry &e)
{
    wait_done(e.done);
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    e.done = nullptr;
    e.exec = nullptr;
}

static void retire_group_graph(chy_ctx::GroupGraph &e)
{
    wait_done(e.done);
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    e.done = nullptr;
    e.exec = nullptr;
}

// drops the cached graphs of v whose keys (head words, then per item a
// record whose first word is the grid) name grid g (all with g null)
static void drop_graphs_of(std::vector<chy_ctx::GroupGraph> &v, size_t head, size_t per, const chy_grid *g)
{
    for (size_t i = 0; i < v.size();) {
        bool hit = g == nullptr;
        for (size_t k = head; !hit && k < v[i].key.size(); k += per) hit = v[i].key[k] == (uintptr_t)g;
        if (hit) {
            retire_group_graph(v[i]);
            v.erase(v.begin() + (long)i);
        } else {
            ++i;
        }
    }
}

// drops the captured group and batch graphs that launch work on grid g (all
// with g null)
static void drop_group_graphs(chy_ctx *c, const chy_grid *g)
{
    drop_graphs_of(c->group_graphs, kGroupKeyHead, kGroupKeyPerField, g);
    drop_graphs_of(c->batch_graphs, kBatchKeyHead, kBatchKeyPerItem, g);
}

// true while s records into a graph (a capture of ours or of the caller's):
// work issued then must be plain launches, not graph replays or captures
static bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &st) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return st != hipStreamCaptureStatusNone;
}

// BDYGEN alone: ring moments from a transform of the interior lanes (no
// spectrum kept), then the radial sums and the boundary values.
static int run_bdygen(chy_grid *g, int maxtrm, int isym, double redge, const double *rho, double *phi)
{
    if (!bdygen_args_ok(maxtrm, isym, redge)) return CHY_EINVAL;
    hipStream_t s = g->ctx->stream;
    Profiler *p = g->ctx->prof.on ? &g->ctx->prof : nullptr;
    if (p) p->begin(CHY_STAGE_FFT_FWD, s);
    int rc = dev_rc(launch_fft_forward(g, rho, phi, nullptr, CHY_COEF_NONE, LANES_INTERIOR, true, false, s),
                    "moment transform launch");
    if (p) p->end(s);
    if (rc) return rc;
    if (p) p->begin(CHY_STAGE_BDYGEN, s);
    rc = dev_rc(launch_bdygen_from_moments(g, maxtrm, isym, redge, phi, false, nullptr, CHY_COEF_NONE, s),
                "bdygen launch");
    if (p) p->end(s);
    return rc;
}

// the solve after the forward transform: mode solves, inverse transform, ghosts, ZAXPHI
static int run_pot3_tail(chy_grid *g, hipStream_t s, int npoint, double *phi, int klocat_in, bool rhs_edges)
{
    Profiler *p = g->ctx->prof.on ? &g->ctx->prof : nullptr;
    int rc = dev_rc(launch_mode_solve(g, rhs_edges, s, p), "mode solve launch");
    if (rc) return rc;
    if (p) p->begin(CHY_STAGE_FFT_INV, s);
    rc = dev_rc(launch_fft_inverse(g, phi, s), "inverse FFT launch");
    if (p) p->end(s);
    if (rc) return rc;
    if (p) p->begin(CHY_STAGE_GHOST, s);
    rc = dev_rc(launch_zaxphi(g, npoint, klocat_in, phi, zaxphi_from_samples(g), s), "ZAXPHI launch");
    if (p) p->end(s);
    return rc;
}

// BDYGEN + POT3 with the density read once: the interior transform writes
// both the spectrum and the ring moments; BDYGEN then writes the boundary
// lanes' spectrum directly (a Fourier series of order <= MAXTRM).  For
// LMAX/2 <= 10 the boundary values go through PHI and a transform instead.
static int run_solve_eager(chy_grid *g, hipStream_t s, int maxtrm, int isym, double redge, int npoint,
                           const double *rho, double *phi, double *coef, int coef_mode, int klocat_in)
{
    Profiler *p = g->ctx->prof.on ? &g->ctx->prof : nullptr;
    const bool spectral = bdygen_spectral_ok(g);
    if (p) p->begin(CHY_STAGE_FFT_FWD, s);
    int rc = dev_rc(launch_fft_forward(g, rho, phi, coef, coef_mode, LANES_INTERIOR, true, true, s),
                    "forward FFT launch");
    if (p) p->end(s);
    if (rc) return rc;
    if (p) p->begin(CHY_STAGE_BDYGEN, s);
    rc = dev_rc(launch_bdygen_from_moments(g, maxtrm, isym, redge, phi, spectral, coef, coef_mode, s),
                "bdygen launch");
    if (p) p->end(s);
    if (rc) return rc;
    if (!spectral) {
        if (p) p->begin(CHY_STAGE_FFT_BDY, s);
        rc = dev_rc(launch_fft_forward(g, rho, phi, coef, coef_mode, LANES_BOUNDARY, false, true, s),
                    "boundary FFT launch");
        if (p) p->end(s);
        if (rc) return rc;
    }
    return run_pot3_tail(g, s, npoint, phi, klocat_in, !spectral);
}

// The same sequence replayed from a HIP graph (captured on first use for a
// given set of arguments) when stage timing is off: one launch per solve.
static int run_solve(chy_grid *g, hipStream_t s, int maxtrm, int isym, double redge, int npoint,
                     const double *rho, double *phi, double *coef, int coef_mode, int klocat_in)
{
    if (!bdygen_args_ok(maxtrm, isym, redge)) return CHY_EINVAL;
    chy_ctx *c = g->ctx;
    if (c->prof.on || !c->graphs || capturing(s))
        return run_solve_eager(g, s, maxtrm, isym, redge, npoint, rho, phi, coef, coef_mode, klocat_in);
    SolveKey key{rho, phi, coef, coef_mode, maxtrm, isym, npoint, klocat_in, redge, s};
    hipGraphExec_t exec = nullptr;
    for (size_t i = 0; i < g->graphs.size(); ++i) {
        if (g->graphs[i].key == key) {
            chy_grid::GraphEntry hit = g->graphs[i];
            g->graphs.erase(g->graphs.begin() + (long)i);
            g->graphs.push_back(hit);
            exec = hit.exec;
            break;
        }
    }
    if (!exec) {
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return dev_rc((int)e, "hipStreamBeginCapture");
        int rc = run_solve_eager(g, s, maxtrm, isym, redge, npoint, rho, phi, coef, coef_mode, klocat_in);
        e = hipStreamEndCapture(s, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return dev_rc((int)e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return dev_rc((int)e, "hipGraphInstantiate");
        hipEvent_t done = nullptr;
        e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipGraphExecDestroy(exec);
            return dev_rc((int)e, "hipEventCreate");
        }
        if ((int)g->graphs.size() >= kGraphCache) retire_graph(g->graphs.front());
        if ((int)g->graphs.size() >= kGraphCache) g->graphs.erase(g->graphs.begin());
        g->graphs.push_back(chy_grid::GraphEntry{key, exec, done});
    }
    int rc = dev_rc((int)hipGraphLaunch(exec, s), "hipGraphLaunch");
    if (rc == CHY_OK) rc = dev_rc((int)hipEventRecord(g->graphs.back().done, s), "hipEventRecord");
    return rc;
}

// BDYGEN + POT3 of every field in F (spectral BDYGEN), one launch per kernel
// on grid g's tables: the sequence of run_solve_eager, field by field the
// same arithmetic, so each field's result equals its own chy_solve's.
static int run_solve_fields(chy_grid *g, const Fields &F, hipStream_t s, int maxtrm, int isym, double redge,
                            int npoint, int coef_mode)
{
    Profiler *p = g->ctx->prof.on ? &g->ctx->prof : nullptr;
    if (p) p->begin(CHY_STAGE_FFT_FWD, s);
    int rc = dev_rc(launch_fft_forward_f(g, F, coef_mode, LANES_INTERIOR, true, true, s), "forward FFT launch");
    if (p) p->end(s);
    if (rc) return rc;
    if (p) p->begin(CHY_STAGE_BDYGEN, s);
    rc = dev_rc(launch_bdygen_from_moments_f(g, F, maxtrm, isym, redge, true, coef_mode, s), "bdygen launch");
    if (p) p->end(s);
    if (rc) return rc;
    rc = dev_rc(launch_mode_solve_f(g, F, false, s, p), "mode solve launch");
    if (rc) return rc;
    if (p) p->begin(CHY_STAGE_FFT_INV, s);
    rc = dev_rc(launch_fft_inverse_f(g, F, s), "inverse FFT launch");
    if (p) p->end(s);
    if (rc) return rc;
    if (p) p->begin(CHY_STAGE_GHOST, s);
    rc = dev_rc(launch_zaxphi_f(g, F, npoint, zaxphi_from_samples(g), s), "ZAXPHI launch");
    if (p) p->end(s);
    return rc;
}

static int run_pot3(chy_grid *g, int npoint, const double *rho, double *phi, double *coef, int coef_mode,
                    int klocat_in)
{
    hipStream_t s = g->ctx->stream;
    Profiler *p = g->ctx->prof.on ? &g->ctx->prof : nullptr;
    if (p) p->begin(CHY_STAGE_FFT_FWD, s);
    int rc = dev_rc(launch_fft_forward(g, rho, phi, coef, coef_mode, LANES_ALL, false, true, s),
                    "forward FFT launch");
    if (p) p->end(s);
    if (rc) return rc;
    return run_pot3_tail(g, s, npoint, phi, klocat_in, true);
}

// Buffers of the refined solve, allocated on first use.
static int ensure_refine(chy_grid *g)
{
    auto get = [&](void **p, size_t bytes) -> int {
        if (*p) return CHY_OK;
        if (hipMalloc(p, bytes) != hipSuccess) return CHY_ENOMEM;
        if (hipMemset(*p, 0, bytes) != hipSuccess) return CHY_EDEVICE;
        g->allocations.push_back(*p);
        g->device_bytes += bytes;
        return CHY_OK;
    };
    if (!refine_supported(g)) return CHY_EUNSUPPORTED;
    int rc = get((void **)&g->d.spec32, sizeof(float) * (size_t)g->L * g->K * g->J);
    if (!rc) rc = get((void **)&g->d.yh32, sizeof(float) * (size_t)g->L * g->PP * g->JP);
    if (!rc) rc = get((void **)&g->d.rstat, sizeof(unsigned long long) * 2 * kMaxRefine);
    // the r factors of the one-kernel-per-plane form, once a grid may use it
    if (!rc && g->refine_fused > 0 && !g->d.fac32 && refine_planes_ok(g, 2)) {
        rc = get((void **)&g->d.fac32, sizeof(float) * 2 * (size_t)(g->N + 1) * g->PP * g->JP);
        if (!rc) rc = dev_rc(launch_thomas_factors(g, g->ctx->stream), "r-solve factor launch");
    }
    return rc;
}

// The right-hand side spectrum of POT3's mode systems for every field of F:
// the interior lanes' transform (with BDYGEN's ring moments and COEF), then
// BDYGEN -- in the spectrum when LMAX/2 > MAXTRM's 10, else through PHI, a
// transform of the boundary lanes and the boundary terms of Y.
static int rhs_spectrum_f(chy_grid *g, const Fields &F, hipStream_t s, int maxtrm, int isym, double redge,
                          int coef_mode)
{
    Profiler *p = g->ctx->prof.on ? &g->ctx->prof : nullptr;
    const bool spectral = bdygen_spectral_ok(g);
    if (p) p->begin(CHY_STAGE_FFT_FWD, s);
    int rc = dev_rc(launch_fft_forward_f(g, F, coef_mode, LANES_INTERIOR, true, true, s), "forward FFT launch");
    if (p) p->end(s);
    if (p) p->begin(CHY_STAGE_BDYGEN, s);
    if (!rc) rc = dev_rc(launch_bdygen_from_moments_f(g, F, maxtrm, isym, redge, spectral, coef_mode, s),
                         "bdygen launch");
    if (p) p->end(s);
    if (!rc && !spectral) {
        if (p) p->begin(CHY_STAGE_FFT_BDY, s);
        rc = dev_rc(launch_fft_forward_f(g, F, coef_mode, LANES_BOUNDARY, false, true, s), "boundary FFT launch");
        if (!rc) rc = dev_rc(launch_rhs_edges_f(g, F, s), "boundary terms launch");
        if (p) p->end(s);
    }
    return rc;
}

// Refined solve of every field of F (grids[f] owns field f's workspace; g's
// tables), by iterative refinement of the mode systems in the azimuthal
// spectrum, where the 7-point operator is block diagonal by mode:
//   B = the right-hand side spectrum (binary64, as the plain solve forms it);
//   X = 0; for each step: R = B - A X (binary64, per mode, rounded to
//   binary32), D = A^-1 R by the binary32 mode solves (binary64 r solves for
//   orders m <= CHY_TUNE_REFINE_M64), X += D (binary64);
//   then the inverse transform of X (binary64) into PHI with POT3's ghosts,
//   and ZAXPHI.
// The first step's R is B itself.  X lives in the grid's binary64 eigen
// coordinate buffer (unused by this path): in binary32 for two steps, whose
// only stored iterate is the first, the binary32 sum of the FP32 matrix cores
// (exact, half the bytes for its store and both reads); R and the eigen
// coordinates in binary32 buffers.  with_stats: per step max |D| and max |X| into d.rstat.
static int run_refined_fields(chy_grid *g, const std::vector<chy_grid *> &grids, const Fields &F, hipStream_t s,
                              int maxtrm, int isym, double redge, int npoint, int coef_mode, int steps,
                              bool with_stats)
{
    if (!refine_supported(g)) return CHY_EUNSUPPORTED;
    int rc = rhs_spectrum_f(g, F, s, maxtrm, isym, redge, coef_mode);
    if (rc) return rc;
    if (with_stats) rc = dev_rc(launch_zero_u64_f(F, 2 * kMaxRefine, s), "statistics clear");
    Fields Fr = F, Ff0 = F, Ff = F, Fi = F;
    for (int f = 0; f < F.nf; ++f) {
        const DeviceGrid &d = grids[f]->d;
        Fr.aux[f] = d.yh;       // residual: B (spec) and X in, R out
        Fr.yh[f] = d.spec32;
        Ff0.yh[f] = d.yh32;     // first step: B (binary64) in
        Ff.spec[f] = d.spec32;  // later steps: R (binary32) in
        Ff.yh[f] = d.yh32;
        Fi.yh[f] = d.yh32;      // inverse: into X, the last step into spec
        Fi.aux[f] = d.yh;
    }
    Profiler *p = g->ctx->prof.on ? &g->ctx->prof : nullptr;
    // the fused per-plane kernel where it applies (it collects no per-step
    // statistics: those calls take the kernel-per-stage path)
    if (F.nf <= g->refine_fused && g->d.fac32 && !with_stats && refine_planes_ok(g, steps)) {
        Fields Fp = F;
        for (int f = 0; f < F.nf; ++f) Fp.aux[f] = grids[f]->d.yh32;
        if (p) p->begin(CHY_STAGE_PLANES, s);
        if (!rc) rc = dev_rc(launch_refine_planes_f(g, Fp, steps, g->refine_m64, s), "refined mode planes launch");
        if (p) p->end(s);
        steps = 0;  // the mode work is done
    }
    // two steps: the first iterate, a binary32 value, is kept in binary32
    const bool x32 = steps == 2;
    for (int it = 0; it < steps && !rc; ++it) {
        const bool last = it == steps - 1;
        const int epi = steps == 1 ? EPI_SPEC
                        : it == 0  ? (x32 ? EPI_XSET32 : EPI_XSET)
                        : last     ? (x32 ? EPI_XFIN32 : EPI_XFIN)
                                   : EPI_XADD;
        if (it > 0) {
            if (p) p->begin(CHY_STAGE_RESID, s);
            rc = dev_rc(launch_spec_residual_f(g, Fr, x32, s), "spectral residual launch");
            if (p) p->end(s);
        }
        if (!rc)
            rc = dev_rc(launch_mode_solve32_f(g, it == 0 ? Ff0 : Ff, Fi, g->refine_m64, it == 0, epi,
                                              with_stats ? it : -1, s, p),
                        "correction mode solve launch");
    }
    if (p) p->begin(CHY_STAGE_FFT_INV, s);
    if (!rc) rc = dev_rc(launch_fft_inverse_f(g, F, s), "inverse FFT launch");
    if (p) p->end(s);
    if (p) p->begin(CHY_STAGE_GHOST, s);
    if (!rc) rc = dev_rc(launch_zaxphi_f(g, F, npoint, zaxphi_from_samples(g), s), "ZAXPHI launch");
    if (p) p->end(s);
    return rc;
}

static int run_refined(chy_grid *g, hipStream_t s, int maxtrm, int isym, double redge, int npoint,
                       const double *rho, double *phi, double *coef, int coef_mode, int steps, int klocat_in,
                       bool with_stats)
{
    int rc = ensure_refine(g);
    if (rc) return rc;
    const Fields F = one_field(g, rho, phi, coef, klocat_in);
    return run_refined_fields(g, std::vector<chy_grid *>{g}, F, s, maxtrm, isym, redge, npoint, coef_mode, steps,
                              with_stats);
}

static int read_diag(chy_grid *g, chy_diag *diag, bool tmass, bool zax)
{
    double h[3];
    hipError_t e = hipMemcpyAsync(h, g->d.dscal, sizeof(h), hipMemcpyDeviceToHost, g->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
    if (e != hipSuccess) {
        set_device_error(e, "diagnostics readback");
        return CHY_EDEVICE;
    }
    if (tmass) diag->tmass = h[0];
    if (zax) {
        diag->phichk = h[1];
        diag->klocat = (int)h[2];
    }
    return CHY_OK;
}

static int ensure_staging(chy_grid *g, bool coef)
{
    const size_t fb = sizeof(double) * field_elems(g);
    if (!g->d.h_rho && hipMalloc(&g->d.h_rho, fb) != hipSuccess) return CHY_ENOMEM;
    if (!g->d.h_phi && hipMalloc(&g->d.h_phi, fb) != hipSuccess) return CHY_ENOMEM;
    if (coef && !g->d.h_coef && hipMalloc(&g->d.h_coef, sizeof(double) * coef_elems(g)) != hipSuccess)
        return CHY_ENOMEM;
    return CHY_OK;
}

static void release_workers(chy_ctx *c)
{
    for (hipStream_t w : c->workers) (void)hipStreamDestroy(w);
    for (hipEvent_t e : c->worker_ev) (void)hipEventDestroy(e);
    c->workers.clear();
    c->worker_ev.clear();
}

static int ensure_workers(chy_ctx *c)
{
    if (!c->fork_ev && hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess)
        return CHY_EDEVICE;
    while ((int)c->workers.size() < c->nworkers) {
        hipStream_t w = nullptr;
        hipEvent_t e = nullptr;
        hipError_t err = hipStreamCreateWithFlags(&w, hipStreamNonBlocking);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (err != hipSuccess) {
            if (w) (void)hipStreamDestroy(w);
            return dev_rc((int)err, "worker stream");
        }
        c->workers.push_back(w);
        c->worker_ev.push_back(e);
    }
    return CHY_OK;
}

}  // namespace chy

using namespace chy;

extern "C" {

const char *chy_strerror(int code)
{
    switch (code) {
    case CHY_OK: return "success";
    case CHY_EINVAL: return "invalid argument or grid description";
    case CHY_EUNSUPPORTED: return "unsupported grid (KMAX must be 2^k >= 4, LMAX even and <= 1024 with LMAX/2 23-smooth, JMAX <= 4096, |ISYM| = 2)";
    case CHY_ENOMEM: return "out of memory";
    case CHY_EDEVICE: return "HIP runtime error";
    case CHY_ESTATE: return "call out of order";
    case CHY_ENUMERIC: return "z operator not symmetrisable or eigen-solver did not converge";
    case CHY_ELIMIT: return "input larger than the fixed scratch of a no-allocation entry point";
    default: return "unknown error";
    }
}

const char *chy_last_device_error(void) { return g_last_error.c_str(); }

int chy_ctx_create(int device, chy_ctx **out)
{
    if (!out) return CHY_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return CHY_EINVAL;
    chy_ctx *c = new (std::nothrow) chy_ctx;
    if (!c) return CHY_ENOMEM;
    c->device = device;
    DeviceScope ds(device);
    hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_device_error(e, "hipStreamCreate");
        delete c;
        return CHY_EDEVICE;
    }
    c->stream = c->own;
    e = (hipError_t)force_scratch_address(device, &c->force_scratch);
    if (e != hipSuccess) {
        set_device_error(e, "hipGetSymbolAddress");
        (void)hipStreamDestroy(c->own);
        delete c;
        return CHY_EDEVICE;
    }
    *out = c;
    return CHY_OK;
}

int chy_ctx_destroy(chy_ctx *ctx)
{
    if (!ctx) return CHY_EINVAL;
    DeviceScope ds(ctx->device);
    ctx->prof.collect();
    ctx->prof.release();
    drop_group_graphs(ctx, nullptr);
    release_workers(ctx);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->d_diag) (void)hipFree(ctx->d_diag);
    if (ctx->h_diag) (void)hipHostFree(ctx->h_diag);
    for (const chy_ctx::TrigTable &t : ctx->trig) (void)hipFree(t.d);
    if (ctx->force_part) (void)hipFree(ctx->force_part);
    if (ctx->force_out) (void)hipFree(ctx->force_out);
    if (ctx->force_h) (void)hipHostFree(ctx->force_h);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
    (void)hipGetLastError();  // a failed release must not surface at the caller's next launch check
    return CHY_OK;
}

// The worker streams already made stay until the context is destroyed (a
// smaller count leaves the rest idle): the captured graphs' done events and
// the stage timer's events refer to the streams they were recorded on.
int chy_ctx_workers(chy_ctx *ctx, int n)
{
    if (!ctx || n < 1 || n > CHY_MAX_WORKERS) return CHY_EINVAL;
    ctx->nworkers = n;
    return CHY_OK;
}

// chy_solve_batch (steps = 0) and chy_solve_batch_refined (steps > 0)
// One group of a batch (fields on identically described grids, one launch per
// kernel), replayed from a captured HIP graph when the same group with the
// same arrays and arguments ran before on the same stream (graphs on, stage
// timing off); the capture holds kernel launches only.
static int run_group(chy_ctx *ctx, const chy_solve_item &lead, const std::vector<chy_grid *> &grids,
                     const Fields &F, hipStream_t s, int np, int steps)
{
    auto eager = [&]() {
        if (steps > 0)
            return run_refined_fields(lead.grid, grids, F, s, lead.maxtrm, lead.isym, lead.redge, np,
                                      lead.coef_mode, steps, false);
        return run_solve_fields(lead.grid, F, s, lead.maxtrm, lead.isym, lead.redge, np, lead.coef_mode);
    };
    if (ctx->prof.on || !ctx->graphs || capturing(s)) return eager();
    std::vector<uintptr_t> key(kGroupKeyHead + kGroupKeyPerField * (size_t)F.nf);
    uint64_t rbits;
    std::memcpy(&rbits, &lead.redge, sizeof(rbits));
    key[0] = (uintptr_t)steps; key[1] = (uintptr_t)lead.maxtrm; key[2] = (uintptr_t)(lead.isym + 8);
    key[3] = (uintptr_t)np; key[4] = (uintptr_t)lead.coef_mode; key[5] = (uintptr_t)rbits;
    key[6] = (uintptr_t)s; key[7] = (uintptr_t)F.nf;
    for (int f = 0; f < F.nf; ++f) {
        uintptr_t *k = key.data() + kGroupKeyHead + kGroupKeyPerField * f;
        k[0] = (uintptr_t)grids[f]; k[1] = (uintptr_t)F.rho[f]; k[2] = (uintptr_t)F.phi[f];
        k[3] = (uintptr_t)F.coef[f]; k[4] = (uintptr_t)(uint32_t)F.klocat_in[f];
    }
    auto &cache = ctx->group_graphs;
    size_t hit = cache.size();
    for (size_t i = 0; i < cache.size(); ++i)
        if (cache[i].key == key) hit = i;
    if (hit == cache.size()) {
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return dev_rc((int)e, "hipStreamBeginCapture");
        int rc = eager();
        e = hipStreamEndCapture(s, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return dev_rc((int)e, "hipStreamEndCapture");
        chy_ctx::GroupGraph g;
        g.key = std::move(key);
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return dev_rc((int)e, "hipGraphInstantiate");
        e = hipEventCreateWithFlags(&g.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipGraphExecDestroy(g.exec);
            return dev_rc((int)e, "hipEventCreate");
        }
        if (cache.size() >= kGroupGraphCache) {
            retire_group_graph(cache.front());
            cache.erase(cache.begin());
        }
        cache.push_back(std::move(g));
        hit = cache.size() - 1;
    } else {
        std::rotate(cache.begin() + (long)hit, cache.begin() + (long)hit + 1, cache.end());
        hit = cache.size() - 1;
    }
    int rc = dev_rc((int)hipGraphLaunch(cache[hit].exec, s), "hipGraphLaunch");
    if (rc == CHY_OK) rc = dev_rc((int)hipEventRecord(cache[hit].done, s), "hipEventRecord");
    return rc;
}

// The fork from the context's stream, the groups on the worker streams and
// the join of one batch (solve_batch).
static int issue_batch(chy_ctx *ctx, const chy_solve_item *items, int n, int steps)
{
    int rc = CHY_OK;
    // Rounds: item i is in round r when r earlier items use its grid object
    // (they share its workspace).  Within a round, items on grids with the
    // same tables and the same BDYGEN/POT3 parameters form groups of up to
    // kMaxBatch fields that run as one launch per kernel.  Every grid object
    // has a worker stream (order of first appearance); a group runs on its
    // first item's worker and waits only for the workers that last used its
    // other grids, so work on unrelated grids never waits.
    std::vector<int> round(n, 0);
    int nrounds = 0;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < i; ++j)
            if (items[j].grid == items[i].grid) ++round[i];
        nrounds = std::max(nrounds, round[i] + 1);
    }
    auto batchable = [](const chy_solve_item &a, const chy_solve_item &b) {
        return same_tables(a.grid, b.grid) && bdygen_spectral_ok(a.grid) && a.maxtrm == b.maxtrm &&
               a.isym == b.isym && a.redge == b.redge && std::min(a.npoint, 10) == std::min(b.npoint, 10) &&
               a.coef_mode == b.coef_mode;
    };
    const int nw = ctx->nworkers;
    std::vector<chy_grid *> seen;      // grid objects in order of first appearance (worker = index % nw)
    std::vector<int> last_w;           // worker of each seen grid's latest use, -1 before it
    auto grid_index = [&](chy_grid *g) {
        size_t k = 0;
        while (k < seen.size() && seen[k] != g) ++k;
        if (k == seen.size()) {
            seen.push_back(g);
            last_w.push_back(-1);
        }
        return (int)k;
    };
    for (int i = 0; i < n; ++i) (void)grid_index(items[i].grid);
    const int used = std::min((int)seen.size(), nw);
    rc = dev_rc((int)hipEventRecord(ctx->fork_ev, ctx->stream), "hipEventRecord");
    for (int w = 0; w < used && rc == CHY_OK; ++w)
        rc = dev_rc((int)hipStreamWaitEvent(ctx->workers[w], ctx->fork_ev, 0), "hipStreamWaitEvent");
    for (int r = 0; r < nrounds && rc == CHY_OK; ++r) {
        std::vector<std::vector<int>> groups;
        for (int i = 0; i < n; ++i) {
            if (round[i] != r) continue;
            size_t k = 0;
            while (k < groups.size() &&
                   !(groups[k].size() < (size_t)kMaxBatch && batchable(items[groups[k][0]], items[i])))
                ++k;
            if (k == groups.size()) groups.emplace_back();
            groups[k].push_back(i);
        }
        for (size_t k = 0; k < groups.size() && rc == CHY_OK; ++k) {
            const std::vector<int> &grp = groups[k];
            const chy_solve_item &lead = items[grp[0]];
            const int w = grid_index(lead.grid) % nw;
            hipStream_t s = ctx->workers[w];
            // order after the latest use of every member grid on another worker
            std::vector<int> waited;
            for (int i : grp) {
                const int v = last_w[grid_index(items[i].grid)];
                if (v >= 0 && v != w && std::find(waited.begin(), waited.end(), v) == waited.end()) {
                    rc = dev_rc((int)hipStreamWaitEvent(s, ctx->worker_ev[v], 0), "hipStreamWaitEvent");
                    waited.push_back(v);
                }
            }
            const int np = std::min(lead.npoint, 10);
            if (rc == CHY_OK && grp.size() == 1 && steps == 0) {
                rc = run_solve(lead.grid, s, lead.maxtrm, lead.isym, lead.redge, np, lead.rho_d, lead.phi_d,
                               lead.coef_d, lead.coef_mode, lead.diag ? lead.diag->klocat : 0);
            } else if (rc == CHY_OK) {
                Fields F{};
                F.nf = (int)grp.size();
                std::vector<chy_grid *> grids(F.nf);
                for (int f = 0; f < F.nf; ++f) {
                    const chy_solve_item &it = items[grp[f]];
                    set_field(F, f, it.grid, it.rho_d, it.phi_d, it.coef_d, it.diag ? it.diag->klocat : 0);
                    grids[f] = it.grid;
                }
                rc = run_group(ctx, lead, grids, F, s, np, steps);
            }
            for (size_t f = 0; f < grp.size() && rc == CHY_OK; ++f) {
                const chy_solve_item &it = items[grp[f]];
                if (it.diag)
                    rc = dev_rc((int)hipMemcpyAsync(ctx->d_diag + 3 * grp[f], it.grid->d.dscal, 3 * sizeof(double),
                                                    hipMemcpyDeviceToDevice, s),
                                "diagnostics copy");
            }
            // this worker's progress so far, for later groups that reuse these grids elsewhere
            if (rc == CHY_OK) rc = dev_rc((int)hipEventRecord(ctx->worker_ev[w], s), "hipEventRecord");
            for (int i : grp) last_w[grid_index(items[i].grid)] = w;
        }
    }
    // join: the context's stream waits for every worker it forked
    for (int w = 0; w < used; ++w) {
        int r2 = dev_rc((int)hipEventRecord(ctx->worker_ev[w], ctx->workers[w]), "hipEventRecord");
        if (r2 == CHY_OK) r2 = dev_rc((int)hipStreamWaitEvent(ctx->stream, ctx->worker_ev[w], 0),
                                      "hipStreamWaitEvent");
        if (rc == CHY_OK) rc = r2;
    }
    return rc;
}

// A batch replayed from a HIP graph of the whole fork, groups and join when
// the same batch (grids, arrays, arguments, worker count, stream, diagnostics
// staging) ran before; captured on first use.  Work on different grids stays
// on different branches of the graph, as on the worker streams.
static int run_batch_graph(chy_ctx *ctx, const chy_solve_item *items, int n, int steps, bool any_diag)
{
    std::vector<uintptr_t> key(kBatchKeyHead + kBatchKeyPerItem * (size_t)n);
    key[0] = (uintptr_t)steps; key[1] = (uintptr_t)n; key[2] = (uintptr_t)ctx->nworkers;
    key[3] = (uintptr_t)ctx->stream; key[4] = (uintptr_t)any_diag; key[5] = (uintptr_t)ctx->d_diag;
    for (int i = 0; i < n; ++i) {
        const chy_solve_item &it = items[i];
        uint64_t rbits;
        std::memcpy(&rbits, &it.redge, sizeof(rbits));
        uintptr_t *k = key.data() + kBatchKeyHead + kBatchKeyPerItem * (size_t)i;
        k[0] = (uintptr_t)it.grid; k[1] = (uintptr_t)it.rho_d; k[2] = (uintptr_t)it.phi_d;
        k[3] = (uintptr_t)it.coef_d; k[4] = (uintptr_t)it.coef_mode; k[5] = (uintptr_t)it.maxtrm;
        k[6] = (uintptr_t)(it.isym + 8); k[7] = (uintptr_t)it.npoint; k[8] = (uintptr_t)rbits;
        k[9] = (uintptr_t)(uint32_t)(it.diag ? it.diag->klocat : 0); k[10] = (uintptr_t)(it.diag != nullptr);
    }
    auto &cache = ctx->batch_graphs;
    size_t hit = cache.size();
    for (size_t i = 0; i < cache.size(); ++i)
        if (cache[i].key == key) hit = i;
    if (hit == cache.size()) {
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return dev_rc((int)e, "hipStreamBeginCapture");
        int rc = issue_batch(ctx, items, n, steps);
        e = hipStreamEndCapture(ctx->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return dev_rc((int)e, "hipStreamEndCapture");
        chy_ctx::GroupGraph b;
        b.key = std::move(key);
        e = hipGraphInstantiate(&b.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) return dev_rc((int)e, "hipGraphInstantiate");
        e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipGraphExecDestroy(b.exec);
            return dev_rc((int)e, "hipEventCreate");
        }
        if (cache.size() >= kBatchGraphCache) {
            retire_group_graph(cache.front());
            cache.erase(cache.begin());
        }
        cache.push_back(std::move(b));
        hit = cache.size() - 1;
    } else {
        std::rotate(cache.begin() + (long)hit, cache.begin() + (long)hit + 1, cache.end());
        hit = cache.size() - 1;
    }
    int rc = dev_rc((int)hipGraphLaunch(cache[hit].exec, ctx->stream), "hipGraphLaunch");
    if (rc == CHY_OK) rc = dev_rc((int)hipEventRecord(cache[hit].done, ctx->stream), "hipEventRecord");
    return rc;
}

static int solve_batch(chy_ctx *ctx, const chy_solve_item *items, int n, int steps)
{
    if (!ctx || n < 0 || (n > 0 && !items)) return CHY_EINVAL;
    if (steps < 0 || steps > kMaxRefine) return CHY_EINVAL;
    if (n == 0) return CHY_OK;
    // validate everything before anything is launched
    bool any_diag = false;
    for (int i = 0; i < n; ++i) {
        const chy_solve_item &it = items[i];
        if (!it.grid) return CHY_ESTATE;
        if (it.grid->ctx != ctx || !it.rho_d || !it.phi_d) return CHY_EINVAL;
        int np = it.npoint;
        int rc = check_pot3_args(it.grid, np, it.coef_mode, it.coef_d);
        if (rc) return rc;
        if (!bdygen_args_ok(it.maxtrm, it.isym, it.redge)) return CHY_EINVAL;
        any_diag = any_diag || it.diag;
    }
    DeviceScope ds(ctx->device);
    int rc = ensure_workers(ctx);
    if (rc) return rc;
    for (int i = 0; i < n && steps > 0 && rc == CHY_OK; ++i) rc = ensure_refine(items[i].grid);
    if (rc) return rc;
    if (any_diag && ctx->diag_cap < n) {
        drop_graphs_of(ctx->batch_graphs, kBatchKeyHead, kBatchKeyPerItem, nullptr);  // they copy into the old staging
        if (ctx->d_diag) (void)hipFree(ctx->d_diag);
        if (ctx->h_diag) (void)hipHostFree(ctx->h_diag);
        ctx->d_diag = ctx->h_diag = nullptr;
        ctx->diag_cap = 0;
        if (hipMalloc(&ctx->d_diag, sizeof(double) * 3 * n) != hipSuccess) return CHY_ENOMEM;
        if (hipHostMalloc(&ctx->h_diag, sizeof(double) * 3 * n) != hipSuccess) return CHY_ENOMEM;
        ctx->diag_cap = n;
    }
    // one graph launch for the whole batch unless stage timing is on, graphs
    // are off or the caller is capturing the context's stream
    if (ctx->graphs && !ctx->prof.on && !capturing(ctx->stream))
        rc = run_batch_graph(ctx, items, n, steps, any_diag);
    else
        rc = issue_batch(ctx, items, n, steps);
    if (rc || !any_diag) return rc;
    hipError_t e = hipMemcpyAsync(ctx->h_diag, ctx->d_diag, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return dev_rc((int)e, "diagnostics readback");
    for (int i = 0; i < n; ++i) {
        if (!items[i].diag) continue;
        items[i].diag->tmass = ctx->h_diag[3 * i];
        items[i].diag->phichk = ctx->h_diag[3 * i + 1];
        items[i].diag->klocat = (int)ctx->h_diag[3 * i + 2];
    }
    return CHY_OK;
}

int chy_solve_batch(chy_ctx *ctx, const chy_solve_item *items, int n) { return solve_batch(ctx, items, n, 0); }

int chy_solve_batch_refined(chy_ctx *ctx, const chy_solve_item *items, int n, int steps)
{
    if (steps < 1) return CHY_EINVAL;
    return solve_batch(ctx, items, n, steps);
}

int chy_ctx_profile(chy_ctx *ctx, int enable)
{
    if (!ctx) return CHY_EINVAL;
    DeviceScope ds(ctx->device);
    ctx->prof.reset();
    ctx->prof.on = enable != 0;
    // events for a few hundred solves' stages, created here rather than
    // between launches
    if (ctx->prof.on) ctx->prof.reserve(2048);
    return CHY_OK;
}

int chy_ctx_stage_times(chy_ctx *ctx, double *ms, int *counts, int n)
{
    if (!ctx || n < 0 || (n > 0 && (!ms || !counts))) return CHY_EINVAL;
    DeviceScope ds(ctx->device);
    int rc = ctx->prof.collect();
    for (int i = 0; i < n && i < CHY_NSTAGES; ++i) {
        ms[i] = ctx->prof.ms[i];
        counts[i] = ctx->prof.count[i];
    }
    return rc;
}

int chy_ctx_graphs(chy_ctx *ctx, int enable)
{
    if (!ctx) return CHY_EINVAL;
    ctx->graphs = enable != 0;
    return CHY_OK;
}

int chy_ctx_set_stream(chy_ctx *ctx, void *stream)
{
    if (!ctx) return CHY_EINVAL;
    ctx->stream = stream ? (hipStream_t)stream : ctx->own;
    return CHY_OK;
}

int chy_ctx_sync(chy_ctx *ctx)
{
    if (!ctx) return CHY_EINVAL;
    DeviceScope ds(ctx->device);
    return dev_rc((int)hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int chy_setbdy(chy_ctx *ctx, const chy_grid_desc *desc, int isym, chy_grid **out)
{
    if (!ctx || !desc || !out) return CHY_EINVAL;
    *out = nullptr;
    chy_grid *g = new (std::nothrow) chy_grid;
    if (!g) return CHY_ENOMEM;
    g->ctx = ctx;
    DeviceScope ds(ctx->device);
    int rc = build_grid(g, desc, isym);
    if (rc != CHY_OK) {
        free_grid_memory(g);
        delete g;
        return rc;
    }
    *out = g;
    return CHY_OK;
}

int chy_grid_set_mode_solver(chy_grid *g, int kind)
{
    if (!g || (kind != CHY_MODES_CR && kind != CHY_MODES_EIGEN)) return CHY_EINVAL;
    if (kind == CHY_MODES_CR && !cr_supported(g)) return CHY_EUNSUPPORTED;
    if (kind == g->mode_solver) return CHY_OK;
    DeviceScope ds(g->ctx->device);
    for (auto &e : g->graphs) retire_graph(e);
    g->graphs.clear();
    drop_group_graphs(g->ctx, g);
    g->mode_solver = kind;
    return CHY_OK;
}

int chy_grid_mode_solver(const chy_grid *g) { return g ? g->mode_solver : -1; }

int chy_grid_set_tune(chy_grid *g, int param, int value)
{
    if (!g) return CHY_EINVAL;
    int *slot = nullptr;
    if (param == CHY_TUNE_CR_LANES && cr_lanes_ok(g, value)) slot = &g->cr_lanes;
    if (param == CHY_TUNE_REFINE_M64 && value >= -1 && value <= g->N) slot = &g->refine_m64;
    if (param == CHY_TUNE_REFINE_FUSED && value >= 0 && value <= kMaxBatch) slot = &g->refine_fused;
    if (!slot) return CHY_EINVAL;
    if (*slot == value) return CHY_OK;
    DeviceScope ds(g->ctx->device);
    for (auto &e : g->graphs) retire_graph(e);
    g->graphs.clear();
    drop_group_graphs(g->ctx, g);
    *slot = value;
    return CHY_OK;
}

int chy_grid_tune(const chy_grid *g, int param)
{
    if (!g) return -1;
    if (param == CHY_TUNE_CR_LANES) return g->cr_lanes;
    if (param == CHY_TUNE_REFINE_M64) return g->refine_m64;
    if (param == CHY_TUNE_REFINE_FUSED) return g->refine_fused;
    return -1;
}

int chy_grid_destroy(chy_grid *g)
{
    if (!g) return CHY_EINVAL;
    DeviceScope ds(g->ctx->device);
    if (hipStreamSynchronize(g->ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipDeviceSynchronize();
    }
    for (auto &e : g->graphs) retire_graph(e);
    g->graphs.clear();
    drop_group_graphs(g->ctx, g);
    free_grid_memory(g);
    delete g;
    (void)hipGetLastError();  // a failed release must not surface at the caller's next launch check
    return CHY_OK;
}

size_t chy_grid_device_bytes(const chy_grid *g) { return g ? g->device_bytes : 0; }

int chy_bdygen(chy_grid *g, int maxtrm, int isym, double redge, const double *rho_d, double *phi_d,
               chy_diag *diag)
{
    if (!g) return CHY_ESTATE;  // BDYGEN without SETBDY (boundary.f:311)
    if (!rho_d || !phi_d) return CHY_EINVAL;
    DeviceScope ds(g->ctx->device);
    int rc = run_bdygen(g, maxtrm, isym, redge, rho_d, phi_d);
    if (rc == CHY_OK && diag) rc = read_diag(g, diag, true, false);
    return rc;
}

int chy_pot3(chy_grid *g, int npoint, const double *rho_d, double *phi_d, double *coef_d, int coef_mode,
             chy_diag *diag)
{
    if (!g) return CHY_ESTATE;
    if (!rho_d || !phi_d) return CHY_EINVAL;
    int rc = check_pot3_args(g, npoint, coef_mode, coef_d);
    if (rc) return rc;
    DeviceScope ds(g->ctx->device);
    rc = run_pot3(g, npoint, rho_d, phi_d, coef_d, coef_mode, diag ? diag->klocat : 0);
    if (rc == CHY_OK && diag) rc = read_diag(g, diag, false, true);
    return rc;
}

int chy_solve(chy_grid *g, int maxtrm, int isym, double redge, int npoint, const double *rho_d,
              double *phi_d, double *coef_d, int coef_mode, chy_diag *diag)
{
    if (!g) return CHY_ESTATE;
    if (!rho_d || !phi_d) return CHY_EINVAL;
    int rc = check_pot3_args(g, npoint, coef_mode, coef_d);
    if (rc) return rc;
    DeviceScope ds(g->ctx->device);
    rc = run_solve(g, g->ctx->stream, maxtrm, isym, redge, npoint, rho_d, phi_d, coef_d, coef_mode,
                   diag ? diag->klocat : 0);
    if (rc == CHY_OK && diag) rc = read_diag(g, diag, true, true);
    return rc;
}

int chy_bdygen_host(chy_grid *g, int maxtrm, int isym, double redge, const double *rho_h, double *phi_h,
                    chy_diag *diag)
{
    if (!g) return CHY_ESTATE;
    if (!rho_h || !phi_h) return CHY_EINVAL;
    DeviceScope ds(g->ctx->device);
    int rc = ensure_staging(g, false);
    if (rc) return rc;
    hipStream_t s = g->ctx->stream;
    const size_t fb = sizeof(double) * field_elems(g);
    hipError_t e = hipMemcpyAsync(g->d.h_rho, rho_h, fb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(g->d.h_phi, phi_h, fb, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return dev_rc((int)e, "host to device copy");
    rc = run_bdygen(g, maxtrm, isym, redge, g->d.h_rho, g->d.h_phi);
    if (rc) return rc;
    e = hipMemcpyAsync(phi_h, g->d.h_phi, fb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return dev_rc((int)e, "device to host copy");
    return diag ? read_diag(g, diag, true, false) : CHY_OK;
}

int chy_pot3_host(chy_grid *g, int npoint, const double *rho_h, double *phi_h, double *coef_h, int coef_mode,
                  chy_diag *diag)
{
    if (!g) return CHY_ESTATE;
    if (!rho_h || !phi_h) return CHY_EINVAL;
    int rc = check_pot3_args(g, npoint, coef_mode, coef_h);
    if (rc) return rc;
    DeviceScope ds(g->ctx->device);
    rc = ensure_staging(g, coef_mode != CHY_COEF_NONE);
    if (rc) return rc;
    hipStream_t s = g->ctx->stream;
    const size_t fb = sizeof(double) * field_elems(g);
    const size_t cb = sizeof(double) * coef_elems(g);
    hipError_t e = hipMemcpyAsync(g->d.h_rho, rho_h, fb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(g->d.h_phi, phi_h, fb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && coef_mode != CHY_COEF_NONE)
        e = hipMemcpyAsync(g->d.h_coef, coef_h, cb, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return dev_rc((int)e, "host to device copy");
    rc = run_pot3(g, npoint, g->d.h_rho, g->d.h_phi, g->d.h_coef, coef_mode, diag ? diag->klocat : 0);
    if (rc) return rc;
    e = hipMemcpyAsync(phi_h, g->d.h_phi, fb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && coef_mode != CHY_COEF_NONE)
        e = hipMemcpyAsync(coef_h, g->d.h_coef, cb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return dev_rc((int)e, "device to host copy");
    return diag ? read_diag(g, diag, false, true) : CHY_OK;
}

int chy_solve_refined(chy_grid *g, int maxtrm, int isym, double redge, int npoint, const double *rho_d,
                      double *phi_d, double *coef_d, int coef_mode, int steps, chy_diag *diag, double *rel_corr)
{
    if (!g) return CHY_ESTATE;
    if (!rho_d || !phi_d || steps < 1 || steps > kMaxRefine) return CHY_EINVAL;
    int rc = check_pot3_args(g, npoint, coef_mode, coef_d);
    if (rc) return rc;
    if (!bdygen_args_ok(maxtrm, isym, redge)) return CHY_EINVAL;
    DeviceScope ds(g->ctx->device);
    hipStream_t s = g->ctx->stream;
    rc = run_refined(g, s, maxtrm, isym, redge, npoint, rho_d, phi_d, coef_d, coef_mode, steps,
                     diag ? diag->klocat : 0, rel_corr != nullptr);
    if (rc == CHY_OK && rel_corr) {
        unsigned long long h[2 * kMaxRefine];
        hipError_t e = hipMemcpyAsync(h, g->d.rstat, sizeof(h), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return dev_rc((int)e, "refinement statistics readback");
        rel_corr[0] = 1.0;  // the first step's correction is the whole solution
        for (int it = 1; it < steps; ++it) {
            // max |D| of the step over max |X| after it (azimuthal spectrum)
            double dm, pm;
            std::memcpy(&dm, &h[2 * it], sizeof(double));
            std::memcpy(&pm, &h[2 * it + 1], sizeof(double));
            rel_corr[it] = pm > 0.0 ? dm / pm : 0.0;
        }
    }
    if (rc == CHY_OK && diag) rc = read_diag(g, diag, true, true);
    return rc;
}

int chy_zaxphi(chy_grid *g, int npoint, double *phi_d, chy_diag *diag)
{
    if (!g) return CHY_ESTATE;
    if (!phi_d) return CHY_EINVAL;
    int rc = check_pot3_args(g, npoint, CHY_COEF_NONE, nullptr);
    if (rc) return rc;
    DeviceScope ds(g->ctx->device);
    hipStream_t s = g->ctx->stream;
    Profiler *p = g->ctx->prof.on ? &g->ctx->prof : nullptr;
    if (p) p->begin(CHY_STAGE_GHOST, s);
    rc = dev_rc(launch_zaxphi(g, npoint, diag ? diag->klocat : 0, phi_d, false, s), "ZAXPHI launch");
    if (p) p->end(s);
    if (rc == CHY_OK && diag) rc = read_diag(g, diag, false, true);
    return rc;
}

int chy_zaxphi_host(chy_grid *g, int npoint, double *phi_h, chy_diag *diag)
{
    if (!g) return CHY_ESTATE;
    if (!phi_h) return CHY_EINVAL;
    int rc = check_pot3_args(g, npoint, CHY_COEF_NONE, nullptr);
    if (rc) return rc;
    DeviceScope ds(g->ctx->device);
    rc = ensure_staging(g, false);
    if (rc) return rc;
    hipStream_t s = g->ctx->stream;
    const size_t fb = sizeof(double) * field_elems(g);
    hipError_t e = hipMemcpyAsync(g->d.h_phi, phi_h, fb, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return dev_rc((int)e, "host to device copy");
    rc = dev_rc(launch_zaxphi(g, npoint, diag ? diag->klocat : 0, g->d.h_phi, false, s), "ZAXPHI launch");
    if (rc) return rc;
    e = hipMemcpyAsync(phi_h, g->d.h_phi, fb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return dev_rc((int)e, "device to host copy");
    return diag ? read_diag(g, diag, false, true) : CHY_OK;
}

int chy_source_grav(chy_ctx *ctx, const chy_source_desc *d, const double *r_d, const double *rho_d,
                    const double *omega_d, const double *phi_d, const double *p_d, double *s_d,
                    double *t_d, double *a_d)
{
    if (!ctx) return CHY_ESTATE;
    if (!d || !r_d || !rho_d || !omega_d || !phi_d || !p_d || !s_d || !t_d || !a_d) return CHY_EINVAL;
    if (d->jmax < 1 || d->kmax < 1 || d->lmax < 1 || d->pot_jmax < d->jmax || d->pot_kmax < d->kmax)
        return CHY_EINVAL;
    if (d->jstart < 2 || d->jstart > d->jmax + 1) return CHY_EINVAL;
    if (!(d->dr > 0.0) || !(d->dz > 0.0) || !(d->dtheta > 0.0)) return CHY_EINVAL;
    DeviceScope ds(ctx->device);
    return dev_rc(launch_source_grav(d, r_d, rho_d, omega_d, phi_d, p_d, s_d, t_d, a_d, ctx->stream),
                  "SOURCE launch");
}

// (cos, sin) of the azimuth of plane L = 1..LMAX, cached per context.  kind 0:
// (L-1+1/2) DTHETA as ExternalPot.f:46-48 forms it; kind 1: L DPHI - DPHI/2 as
// wiggle.F:115 does.  Built on the host with the C library's cos/sin (the
// first call for a key waits for the upload).
static int trig_table(chy_ctx *ctx, int kind, int lmax, double dtheta, const double **out)
{
    for (const chy_ctx::TrigTable &t : ctx->trig)
        if (t.kind == kind && t.lmax == lmax && t.dtheta == dtheta) {
            *out = t.d;
            return CHY_OK;
        }
    std::vector<double> h(2 * (size_t)lmax);
    for (int l = 1; l <= lmax; ++l) {
        const double a = kind == 0 ? ((double)(l - 1) + 0.5) * dtheta : (double)l * dtheta - 0.5 * dtheta;
        h[2 * (l - 1)] = std::cos(a);
        h[2 * (l - 1) + 1] = std::sin(a);
    }
    double *dev = nullptr;
    if (hipMalloc(&dev, h.size() * sizeof(double)) != hipSuccess) return CHY_ENOMEM;
    int rc = dev_rc((int)hipMemcpy(dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice),
                    "cos/sin table upload");
    if (rc) {
        (void)hipFree(dev);
        return rc;
    }
    ctx->trig.push_back(chy_ctx::TrigTable{kind, lmax, dtheta, dev});
    *out = dev;
    return CHY_OK;
}

int chy_find_force(chy_ctx *ctx, const chy_force_desc *d, const double *rhf_d, const double *zhf_d,
                   const double *rho_d, double x, double y, double *f_h)
{
    if (!ctx) return CHY_ESTATE;
    if (!d || !rhf_d || !zhf_d || !rho_d || !f_h) return CHY_EINVAL;
    if (d->jmax < 1 || d->kmax < 1 || d->lmax < 1 || d->jmin < 4 || d->jmin > d->jmax + 1) return CHY_EINVAL;
    if (!(d->dr > 0.0) || !(d->dz > 0.0) || !(d->dphi > 0.0)) return CHY_EINVAL;
    DeviceScope ds(ctx->device);
    const double *cs = nullptr;
    int rc = trig_table(ctx, 1, d->lmax, d->dphi, &cs);
    if (rc) return rc;
    const int nb = force_blocks(d);
    if (nb > ctx->force_cap) {
        if (ctx->force_part) (void)hipFree(ctx->force_part);
        ctx->force_part = nullptr;
        ctx->force_cap = 0;
        if (hipMalloc(&ctx->force_part, sizeof(double) * 2 * (size_t)nb) != hipSuccess) return CHY_ENOMEM;
        ctx->force_cap = nb;
    }
    if (!ctx->force_out) {
        if (hipMalloc(&ctx->force_out, sizeof(double) * 2) != hipSuccess) return CHY_ENOMEM;
        if (hipHostMalloc(&ctx->force_h, sizeof(double) * 2) != hipSuccess) {
            (void)hipFree(ctx->force_out);
            ctx->force_out = nullptr;
            return CHY_ENOMEM;
        }
    }
    hipStream_t s = ctx->stream;
    rc = dev_rc(launch_find_force(d, rhf_d, zhf_d, cs, rho_d, x, y, ctx->force_part, ctx->force_out, s),
                "find_force launch");
    if (!rc)
        rc = dev_rc((int)hipMemcpyAsync(ctx->force_h, ctx->force_out, 2 * sizeof(double), hipMemcpyDeviceToHost, s),
                    "force readback");
    if (!rc) rc = dev_rc((int)hipStreamSynchronize(s), "hipStreamSynchronize");
    if (!rc) {
        f_h[0] = ctx->force_h[0];
        f_h[1] = ctx->force_h[1];
    }
    return rc;
}

int chy_find_force_noalloc(chy_ctx *ctx, const chy_force_desc *d, const double *rhf_d, const double *zhf_d,
                           const double *rho_d, double x, double y, double *f_d)
{
    if (!ctx) return CHY_ESTATE;
    if (!d || !rhf_d || !zhf_d || !rho_d || !f_d) return CHY_EINVAL;
    if (d->jmax < 1 || d->kmax < 1 || d->lmax < 1 || d->jmin < 4 || d->jmin > d->jmax + 1) return CHY_EINVAL;
    if (!(d->dr > 0.0) || !(d->dz > 0.0) || !(d->dphi > 0.0)) return CHY_EINVAL;
    const long nb = (long)force_blocks(d);
    if (nb > kForceNoAllocBlocks || d->lmax > kForceNoAllocL) {
        g_last_error = "chy_find_force_noalloc: " + std::to_string(nb) + " block partials and LMAX " +
                       std::to_string(d->lmax) + " exceed its scratch (" + std::to_string(kForceNoAllocBlocks) +
                       " partials, LMAX " + std::to_string(kForceNoAllocL) + ")";
        return CHY_ELIMIT;
    }
    DeviceScope ds(ctx->device);
    return dev_rc(launch_find_force_noalloc(d, rhf_d, zhf_d, rho_d, x, y, ctx->force_scratch, f_d, ctx->stream),
                  "find_force launch");
}

int chy_external_pot(chy_ctx *ctx, const chy_extpot_desc *d, const double *rhf_d, const double *zhf_d,
                     double *phi_d, double *starphi_d)
{
    if (!ctx) return CHY_ESTATE;
    if (!d || !rhf_d || !zhf_d || !phi_d || !starphi_d) return CHY_EINVAL;
    if (d->jmax < 1 || d->kmax < 1 || d->lmax < 1 || d->pot_jmax < d->jmax || d->pot_kmax < d->kmax)
        return CHY_EINVAL;
    if (!(d->dtheta > 0.0)) return CHY_EINVAL;
    DeviceScope ds(ctx->device);
    const double *cs = nullptr;
    int rc = trig_table(ctx, 0, d->lmax, d->dtheta, &cs);
    if (rc) return rc;
    // the star's position as ExternalPot.f:18-19 forms it
    const double ix = d->rpstar * std::cos(d->phi_star), iy = d->rpstar * std::sin(d->phi_star);
    return dev_rc(launch_external_pot(d, rhf_d, zhf_d, cs, phi_d, starphi_d, ix, iy, ctx->stream),
                  "ExternalPot launch");
}

int chy_phac(chy_grid *g, double *phac_h, int n)
{
    if (!g) return CHY_ESTATE;
    if (!phac_h || n < 0 || n > g->K) return CHY_EINVAL;
    if (n == 0) return CHY_OK;
    DeviceScope ds(g->ctx->device);
    hipStream_t s = g->ctx->stream;
    // PHAC(K), K = 2..KMAX+1, lives at dscal[4 + K]
    hipError_t e = hipMemcpyAsync(phac_h, g->d.dscal + 6, sizeof(double) * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return dev_rc((int)e, "PHAC readback");
}

}  // extern "C"
