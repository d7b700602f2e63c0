// Internal declarations of libchypot: per-grid state, device tables and the
// kernel launchers.  Not part of the C-ABI (include/chypot.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "chypot.h"

namespace chy {

// Largest multipole order BDYGEN supports (MAXTRM <= 10, pot/boundary.f:94-96)
// and the number of (l,m) terms with l+m even up to it (the ISYM=2 set,
// pot/boundary.f:345-354,412-500).
constexpr int kMaxTrm = 10;
constexpr int kMaxTerms = 36;
// Ring moments per ring: SUMASS, CM(1..10), SM(1..10) (pot/boundary.f:526-538).
constexpr int kMoments = 21;
// (l, m) of the terms, l-major with m ascending (the order TERM(I) is summed
// in, boundary.f:821-828); terms_table() carries the same order.
constexpr int kTermL[kMaxTerms] = {0, 1, 2, 2, 3, 3, 4, 4, 4, 5, 5, 5, 6, 6, 6, 6, 7, 7,
                                   7, 7, 8, 8, 8, 8, 8, 9, 9, 9, 9, 9, 10, 10, 10, 10, 10, 10};
constexpr int kTermM[kMaxTerms] = {0, 1, 0, 2, 1, 3, 0, 2, 4, 1, 3, 5, 0, 2, 4, 6, 1, 3,
                                   5, 7, 0, 2, 4, 6, 8, 1, 3, 5, 7, 9, 0, 2, 4, 6, 8, 10};
// Rings per boundary chunk (the unit of the deterministic radial reduction).
constexpr int kChunkRings = 64;
// Captured solve graphs kept per grid (one per distinct argument set).
constexpr int kGraphCache = 8;
// Most refinement steps of chy_solve_refined.
constexpr int kMaxRefine = 8;

struct Term {
    int l, m;
    double dlm;       // boundary factor DLM (pot/boundary.f:694-772), REAL*8 TW constants
    int np;           // number of polynomial coefficients
    double poly[6];   // BLM(c,s) = s^m * sum_i poly[i] c^(2i) (pot/boundary.f:204-253)
};
const Term *terms_table(int *count);  // the kMaxTerms terms, l-major, m ascending

// Cyclic reduction of the per-mode systems (pot/blktri.f:60-260).  A program
// is one PRDCT product, in one of two forms:
//   chain (has_pre = 1, ops CR_SCALE): w = sgn x, y = pre w, then for each op
//     solve (T_m - rt I) w = y along r and y = f w; the result is y;
//   partial fractions (ops CR_PF): the result is sum over ops of
//     f (T_m - rt I)^-1 (sgn x), the solves independent of each other.
enum { CR_SCALE = 2, CR_PF = 3 };
struct CrOp {
    double rt, f;
    int post, pad;
};
struct CrProg {
    int off, nsolve, has_pre, pad;
    double sgn, pre;
};
// One row a phase updates (1-based z row) and its programs (-1: none):
//   CR_PHASE_ODD  (reduction, rows at odd multiples of h): P(row) = pm(Y(row))
//   CR_PHASE_EVEN (reduction, rows at even multiples of h):
//                 Y(row) = (pa(P(row-h)) + pc(P(row+h))) - Y(row)
//   CR_PHASE_BACK (back substitution, odd multiples of h):
//                 Y(row) = pm((Y(row) - pa(Y(row-h))) - pc(Y(row+h)))
enum { CR_PHASE_ODD = 0, CR_PHASE_EVEN = 1, CR_PHASE_BACK = 2 };
struct CrRow {
    int row, pa, pc, pm;
};
struct CrPhase {
    int kind, h, row0, nrow;
};
struct CrPlan {
    int kx = 0;
    long nsolve = 0;               // r solves per mode
    std::vector<CrPhase> phases;   // in execution order
    std::vector<CrRow> rows;
    std::vector<CrProg> progs;
    std::vector<CrOp> ops;
};
// Root tables and programs of the z operator tridiag(an, bn, cn) of order n
// (an[0] = cn[n-1] = 0), n = 2^k - 1.
int build_cr_plan(CrPlan &plan, int n, const double *an, const double *bn, const double *cn);

// Device-side grid tables and workspace.  All pointers are device memory.
struct DeviceGrid {
    // FFT
    double2 *twL = nullptr;      // (cos, sin)(2 pi e / L), e = 0..L-1
    double2 *tw64 = nullptr;     // (cos, sin)(2 pi e / 64), e = 0..63 (in-register DFTs)
    double2 *twh = nullptr;      // (cos, sin)(pi m / L), m = 0..10 (cell-centre phase of the ring moments)
    // z eigenbasis (BLKTRI's m-independent z operator, diagonalised)
    double *wf = nullptr;        // [NZ][PP]   Y -> eigen coordinates:  Yh[p] = sum_k wf[k][p] Y[k]
    double *wi = nullptr;        // [NZ][PP]   eigen coordinates -> X:  X[k] = sum_p wi[p][k] Z[p]
    double *lam = nullptr;       // [PP]       eigenvalues (padded entries 1)
    float *wf32 = nullptr;       // [PP][PP]   wf rounded to binary32, zero past NZ (rows k, columns p)
    float *wi32 = nullptr;       // [PP][PP]   wi rounded to binary32, zero past NZ (rows p, columns k)
    float *fac32 = nullptr;      // [N+1][JP][PP] x 2: LU factors of T_m + lambda_p (the one-kernel-per-plane
                                 // refined form; made by the first refined call after the form is enabled)
    // r operator (pot3.f:172-194,226-230): AM(1)=0, CM(M)=0, BM per azimuthal
    // order m; rows padded to JP (rows >= M: AM = CM = 0, BM = 1)
    double *am = nullptr;        // [JP]
    double *cm = nullptr;        // [JP]
    double *bm = nullptr;        // [(N+1)][JP]
    double *rinv = nullptr;      // [(N+1)][NZ][JP] reciprocal Thomas pivots of T_m + lambda_p (JMAX-1 > 1024 only)
    // the z operator for the refinement's spectral residuals: AN (AN(1)=0),
    // CN (CN(NZ)=0; CNMAX separate), BN
    double *rz_an = nullptr;     // [NZ]
    double *rz_cn = nullptr;     // [NZ]
    double *rz_bn = nullptr;     // [NZ] BN (BN(1) = -CN(1); the spectral residual)
    // cyclic reduction programs (CrPlan, uploaded)
    CrOp *cr_ops = nullptr;
    CrProg *cr_progs = nullptr;
    CrRow *cr_rows = nullptr;
    // refinement buffers (allocated on first use): the binary32 right-hand
    // side / residual spectrum and eigen coordinates, the step statistics
    float *spec32 = nullptr;     // [L][KMAX][JMAX]
    float *yh32 = nullptr;       // [L][PP][JP]
    unsigned long long *rstat = nullptr;  // [2 * kMaxRefine]: max |dPHI|, max |PHI| per step
    // ZAXPHI abscissae (pot3.f:342-348)
    double *zx = nullptr;        // [11]
    // BDYGEN
    int *ring_of_pos = nullptr;  // [nring]   radius-sorted position -> ring index (jj-2)+(J-1)(kk-2)
    double *tin = nullptr;       // [nring][nt]  BLM(theta) r^l at sorted position
    double *tout = nullptr;      // [nring][nt]  BLM(theta) r^-(l+1)  (= tin * r^-(2l+1))
    double *rsph = nullptr;      // [nring]   spherical radius at sorted position
    double *dv = nullptr;        // [nring]   DV by ring index (boundary.f:516)
    int *chunk_start = nullptr;  // [nchunk+1] chunk boundaries in sorted positions
    int *jk_cut = nullptr;       // [jkmax]   chunks 0..cut-1 are inside boundary point jk
    double *tb = nullptr;        // [jkmax][nt] DLM*BLM(theta_b) (boundary.f:670-772)
    double *rb = nullptr;        // [jkmax]
    int *jpos = nullptr, *kpos = nullptr;
    double *cosm = nullptr;      // [L][10] cos(m psi_L), psi_L=(L-1/2) DTHETA (boundary.f:31-37)
    double *sinm = nullptr;      // [L][10]
    // workspace
    double *spec = nullptr;      // [L][KMAX][JMAX]  azimuthal spectrum, lanes J=2..JMAX+1, K=2..KMAX+1
    double *yh = nullptr;        // [L][PP][JP]      z-eigen coordinates of the mode RHS / solution
    double *mom = nullptr;       // [nring][kMoments] by ring index (the interior lane index)
    double *chunk = nullptr;     // [2][2*kMaxTerms][nchunk]  inside / outside chunk sums
    double *massio = nullptr;    // [nchunk+1][2][2*kMaxTerms] exclusive prefix (in) / suffix (out)
    double *dscal = nullptr;     // [4 + KMAX+2]: tmass, phichk, klocat, _, phac[K]
    unsigned *ticket = nullptr;  // ZAXPHI blocks finished (0 between launches)
    double *zsamp = nullptr;     // [KMAX][kZaxSamples][L]  PHI(J=2..6,K,L) as the inverse FFT wrote it
    // staging for the host-pointer entry points (allocated on first use)
    double *h_rho = nullptr, *h_phi = nullptr, *h_coef = nullptr;
};

}  // namespace chy

namespace chy {
// HIP-event brackets around solver stages (enabled by chy_ctx_profile).
struct Profiler {
    bool on = false;
    struct Rec { int stage; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;  // events ready for reuse (collect() returns them here)
    double ms[CHY_NSTAGES] = {};
    int count[CHY_NSTAGES] = {};
    int open_stage = -1;
    hipEvent_t open_ev = nullptr;
    void begin(int stage, hipStream_t s);
    void end(hipStream_t s);
    int collect();  // wait for pending records and accumulate
    void reset();
    void reserve(size_t n);  // fill the pool to n events
    void release();          // destroy pooled events
    hipEvent_t take();
};
}  // namespace chy

struct chy_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    bool graphs = true;  // replay chy_solve from captured HIP graphs
    chy::Profiler prof;
    // chy_solve_batch: worker streams (default priority), fork/join events,
    // per-item diagnostics staging
    int nworkers = 4;
    std::vector<hipStream_t> workers;
    std::vector<hipEvent_t> worker_ev;
    hipEvent_t fork_ev = nullptr;
    double *d_diag = nullptr, *h_diag = nullptr;
    int diag_cap = 0;
    // chy_external_pot: (cos psi_L, sin psi_L) tables per (LMAX, DTHETA)
    struct TrigTable {
        int kind;  // 0: psi_L = (L-1+1/2) DTHETA (ExternalPot); 1: L DPHI - DPHI/2 (find_force)
        int lmax;
        double dtheta;
        double *d;
    };
    std::vector<TrigTable> trig;
    // chy_find_force: block partials and the (fx, fy) result
    double *force_part = nullptr, *force_out = nullptr, *force_h = nullptr;
    int force_cap = 0;
    // chy_find_force_noalloc: the device's fixed scratch (force.hip)
    double *force_scratch = nullptr;
    // chy_solve_batch: captured graphs of multi-field groups, keyed on the
    // group's grids, arrays and arguments, most recently used last
    struct GroupGraph {
        std::vector<uintptr_t> key;
        hipGraphExec_t exec = nullptr;
        hipEvent_t done = nullptr;
    };
    std::vector<GroupGraph> group_graphs;
    // chy_solve_batch: captured graphs of whole batches (fork, groups, join),
    // keyed on every item, the worker count, the stream and the diagnostics
    // staging, most recently used last
    std::vector<GroupGraph> batch_graphs;
};

namespace chy {
// Arguments a captured solve graph was recorded with.
struct SolveKey {
    const void *rho;
    void *phi, *coef;
    int coef_mode, maxtrm, isym, npoint, klocat_in;
    double redge;
    hipStream_t stream;
    bool operator==(const SolveKey &o) const
    {
        return rho == o.rho && phi == o.phi && coef == o.coef && coef_mode == o.coef_mode &&
               maxtrm == o.maxtrm && isym == o.isym && npoint == o.npoint && klocat_in == o.klocat_in &&
               redge == o.redge && stream == o.stream;
    }
};
}  // namespace chy

struct chy_grid {
    chy_ctx *ctx = nullptr;
    int J = 0, K = 0, L = 0;     // JMAX, KMAX, LMAX
    int N = 0;                   // LMAX/2
    int M = 0, NZ = 0;           // JMAX-1, KMAX-1: interior r and z unknowns of a mode
    int PP = 0;                  // NZ padded to a multiple of 64
    int JP = 0;                  // M padded to a multiple of 64
    int isym = 2;
    double dtheta = 0, pi = 0, grav = 0;
    double cmmax = 0, cnmax = 0, ammin = 0, pig4 = 0;
    int nring = 0, nchunk = 0, jkmax = 0;
    int ncu = 1;                 // compute units of the device (persistent launches)
    std::vector<double> r, z, rhf, zhf;
    chy::CrPlan cr;              // per-mode cyclic reduction (host copy of the phases)
    int mode_solver = CHY_MODES_CR;  // CHY_MODES_CR or CHY_MODES_EIGEN
    int cr_lanes = 64;               // lanes per r system in the cyclic reduction (64 or 32)
    int refine_m64 = 4;              // refinement: orders m <= this solved in binary64
    int refine_fused = 0;            // refinement: one kernel per mode plane for launches of up to this many fields
    size_t device_bytes = 0;
    std::vector<void *> allocations;
    chy::DeviceGrid d;
    // captured chy_solve graphs, most recently used last (at most kGraphCache);
    // `done` is recorded after every launch so an entry can be retired safely
    struct GraphEntry {
        chy::SolveKey key;
        hipGraphExec_t exec;
        hipEvent_t done;
    };
    std::vector<GraphEntry> graphs;
};

namespace chy {

// host-side preparation (plan.cpp)
int build_grid(chy_grid *g, const chy_grid_desc *desc, int isym);
int symmetric_tridiag_eigen(int n, std::vector<long double> &d, std::vector<long double> &e,
                            std::vector<long double> &q);

// Which (J,K) lanes a forward transform covers: all J=2..JMAX+1 x K=2..KMAX+1,
// the interior J<=JMAX, K<=KMAX (density only), or the boundary row/column
// (BDYGEN's values).
enum { LANES_ALL = 0, LANES_INTERIOR = 1, LANES_BOUNDARY = 2 };

// Per-field pointers of a launch over nf fields that share one grid's tables
// (identical grid descriptions): blockIdx.y, or the tile index of a
// persistent kernel, picks field f.  Each field has the workspace of the grid
// object it was prepared on, so fields on different grid objects run in one
// launch without sharing scratch.  Single-field launches use nf = 1.
// ZAXPHI reads PHI(J,K,L) for J = 2..NPOINT/2+2 <= 6 (NPOINT <= 10)
constexpr int kZaxSamples = 5;
constexpr int kMaxBatch = 32;
struct Fields {
    int nf;
    const void *rho[kMaxBatch];  // forward input (binary64 density, or the binary32 residual)
    double *phi[kMaxBatch];      // PHI: forward input on boundary lanes, inverse and BDYGEN output
    double *coef[kMaxBatch];     // COEF (may be null)
    void *spec[kMaxBatch];       // spectrum workspace (binary64, or binary32 for refinement)
    void *yh[kMaxBatch];         // z-eigen workspace
    double *mom[kMaxBatch];      // ring moments
    double *chunk[kMaxBatch];    // BDYGEN chunk sums
    double *massio[kMaxBatch];   // BDYGEN prefix/suffix sums
    double *dscal[kMaxBatch];    // diagnostics
    unsigned *ticket[kMaxBatch]; // ZAXPHI ticket counter
    double *zsamp[kMaxBatch];    // ZAXPHI's samples PHI(J=2..6,K,L), copied by the inverse FFT
    unsigned long long *rstat[kMaxBatch];  // refinement statistics (max |dPHI|, max |PHI| per step)
    void *aux[kMaxBatch];        // a third per-field array where a kernel needs one (refinement: the
                                 // binary64 solution spectrum)
    int klocat_in[kMaxBatch];    // KLOCAT before the call
};
// one field on grid g's own workspace
Fields one_field(const chy_grid *g, const void *rho, double *phi, double *coef, int klocat_in);
// field f of a batch: user arrays, and the workspace of grid w
void set_field(Fields &F, int f, const chy_grid *w, const void *rho, double *phi, double *coef, int klocat_in);

// kernel launchers (one .hip file each); all return hipError_t as int.  The
// *_f forms run every field of F; the others one field on g's workspace.
// Forward azimuthal transform; optionally writes BDYGEN's ring moments of the
// interior lanes (d.mom) and/or the spectrum (d.spec, with COEF).
int launch_fft_forward(const chy_grid *g, const double *rho, const double *phi, double *coef,
                       int coef_mode, int lanes, bool moments, bool spectrum, hipStream_t s);
int launch_fft_inverse(const chy_grid *g, double *phi, hipStream_t s);
int launch_fft_forward_f(const chy_grid *g, const Fields &F, int coef_mode, int lanes, bool moments,
                         bool spectrum, hipStream_t s);
int launch_fft_inverse_f(const chy_grid *g, const Fields &F, hipStream_t s);
// LMAX the transforms take: even, at most 1024, LMAX/2 with no prime factor
// above 23 (powers of two by the specialised kernels, the rest mixed-radix)
bool fft_size_ok(int L);
// the reciprocal pivots of the r solves by the Thomas algorithm (JMAX-1 > 1024)
int launch_thomas_pivots(const chy_grid *g, hipStream_t s);

// Mode solves on d.spec; rhs_edges applies the boundary terms of Y first
// (skipped when the spectral BDYGEN has applied them).
int launch_mode_solve(const chy_grid *g, bool rhs_edges, hipStream_t s, Profiler *prof);
// BLKTRI's cyclic reduction on every field's spectrum (cr.hip), in place, the
// boundary terms of Y already applied; cr_supported(g): the grid's r solves
// fit the kernel (JP/64 in {1, 2, 4, 8}, z rows 2^k - 1).
bool cr_supported(const chy_grid *g);
bool cr_lanes_ok(const chy_grid *g, int lanes);
int launch_cr_solve_f(const chy_grid *g, const Fields &F, hipStream_t s);
int launch_mode_solve_f(const chy_grid *g, const Fields &F, bool rhs_edges, hipStream_t s, Profiler *prof);
// The refinement's correction solves (spectral refinement of the mode
// systems): right-hand side F.spec (binary64 when rhs64, else binary32,
// [q][k][j] pitch J), eigen coordinates F.yh (binary32); the z transforms in
// binary32 on the FP32 matrix cores, the r solves in binary32 except for
// orders m <= m_fp64, which are solved in binary64; the correction X goes to
// the solution spectrum F.aux per epi (EPI_XSET: X, EPI_XADD: += X,
// EPI_XFIN: F.spec = F.aux + X; F.aux binary64, or binary32 for EPI_XSET32 and
// EPI_XFIN32: the first iterate is the FP32 matrix cores' sum, a binary32
// value, so a 2-step solve keeps it exactly in half the bytes).
// refine_supported(g): KMAX-1 <= 255.
// Ff carries the forward transform's arrays, Fi the inverse's.  stat_step >= 0:
// the inverse also collects max |X| and max |F.aux + X| into
// F.rstat + 2 * stat_step.
enum { EPI_SPEC = 0, EPI_XSET = 1, EPI_XADD = 2, EPI_XFIN = 3, EPI_XSET32 = 4, EPI_XFIN32 = 5 };
int refine_supported(const chy_grid *g);
// The whole mode work of a 1- or 2-step refined solve in one kernel per mode
// plane (binary32 plane in LDS): reads the binary64 right-hand side spectrum
// F.spec, writes the solution into it; F.aux: a binary32 scratch of
// L x PP x JP per field (two steps).  refine_planes_ok: the plane fits the
// kernel (KMAX <= 129, JMAX <= 257) and steps is 1 or 2.
bool refine_planes_ok(const chy_grid *g, int steps);
// fills d.fac32 (allocated by the caller) on stream s
int launch_thomas_factors(const chy_grid *g, hipStream_t s);
int launch_refine_planes_f(const chy_grid *g, const Fields &F, int steps, int m_fp64, hipStream_t s);
struct Profiler;
int launch_mode_solve32_f(const chy_grid *g, const Fields &Ff, const Fields &Fi, int m_fp64, bool rhs64, int epi,
                          int stat_step, hipStream_t s, Profiler *prof);
// Zero the first n words of every field's F.rstat (a fill kernel: callers may
// capture the sequence into a graph).
int launch_zero_u64_f(const Fields &F, int n, hipStream_t s);
// The boundary terms of Y on the spectrum (pot3.f:212-224), for every field.
int launch_rhs_edges_f(const chy_grid *g, const Fields &F, hipStream_t s);
// The mode systems' binary64 residual B - A X in the spectrum (F.spec = B,
// F.aux = X, binary32 when x32 else binary64) rounded to binary32 into F.yh
// ([q][k][j], pitch J).
int launch_spec_residual_f(const chy_grid *g, const Fields &F, bool x32, hipStream_t s);

// ZAXPHI with PHAC/PHICHK/KLOCAT (POT3's ghost cells come with the inverse FFT).
// samples: read PHI(J=2..6,K,L) from the copy the inverse FFT just made
// (only where zaxphi_from_samples(g)), else from PHI itself.
bool zaxphi_from_samples(const chy_grid *g);
int launch_zaxphi(const chy_grid *g, int npoint, int klocat_in, double *phi, bool samples, hipStream_t s);
int launch_zaxphi_f(const chy_grid *g, const Fields &F, int npoint, bool samples, hipStream_t s);
// BDYGEN from the ring moments in d.mom: radial split sums, then either the
// boundary values and ghosts in PHI, or (spectral) the boundary lanes'
// azimuthal spectrum written straight into d.spec with their COEF entries
// and the boundary terms of the mode right-hand side.
int launch_bdygen_from_moments(const chy_grid *g, int maxtrm, int isym, double redge, double *phi,
                               bool spectral, double *coef, int coef_mode, hipStream_t s);
int launch_bdygen_from_moments_f(const chy_grid *g, const Fields &F, int maxtrm, int isym, double redge,
                                 bool spectral, int coef_mode, hipStream_t s);
bool bdygen_spectral_ok(const chy_grid *g);
// The star potential of pot/ExternalPot.f (arguments checked by the caller; cs
// holds (cos psi_L, sin psi_L) for L = 1..LMAX, ix/iy the star's position).
int launch_external_pot(const chy_extpot_desc *d, const double *rhf, const double *zhf, const double *cs,
                        double *phi, double *starphi, double ix, double iy, hipStream_t st);
// find_force of pot/wiggle.F: block partials into part (force_blocks(d) pairs),
// the total (fx, fy) into out.
int force_blocks(const chy_force_desc *d);
// chy_find_force_noalloc: largest partial count and LMAX its fixed scratch
// holds; the scratch's device address on a device (taken once per device);
// the launch on that scratch (cos/sin table built on the device).
constexpr int kForceNoAllocBlocks = 65536;
constexpr int kForceNoAllocL = 4096;
int force_scratch_address(int device, double **out);
int launch_find_force_noalloc(const chy_force_desc *d, const double *rhf, const double *zhf, const double *rho,
                              double x, double y, double *scr, double *out, hipStream_t st);
int launch_find_force(const chy_force_desc *d, const double *rhf, const double *zhf, const double *cs,
                      const double *rho, double x, double y, double *part, double *out, hipStream_t st);
// The momentum sources of radhydro/source.f:41-79 (arguments checked by the caller).
int launch_source_grav(const chy_source_desc *d, const double *r, const double *rho, const double *omega,
                       const double *phi, const double *p, double *s, double *t, double *a, hipStream_t st);

void set_device_error(hipError_t e, const char *where);

}  // namespace chy
