// Per-mode generalised cyclic reduction on the GPU: the solve BLKTRI does for
// every azimuthal mode (pot/blktri.f:60-260, called at pot/pot3.f:233-239),
// with BLKTRI's own root tables (crplan.cpp).
//
// A mode's unknowns are X(j, i): j = 1..M along r (JMAX-1 rows), i = 1..N
// along z (KMAX-1 = 2^k - 1 rows).  The reduction works on whole z rows: each
// step is a product of r-tridiagonal solves (T_m - rt I)^-1 and combinations
// applied to one r vector (one PRDCT, pot/blktri.f:177-260).  Here
//   - one wavefront carries one r vector of one mode pair: the cosine and sine
//     parts of order m share T_m (BM depends on m only), so every solve has
//     two right-hand sides; lane t holds rows tR .. tR+R-1 in registers;
//   - each r solve is a partitioned elimination inside the lane followed by
//     parallel cyclic reduction of the lanes' interface rows over the 64
//     lanes (wavefront shuffles);
//   - a PRDCT is a flat op list prepared on the host (the root order and the
//     choice of each multiplier depend on the roots alone), so all lanes, and
//     all modes, run the same control flow;
//   - one launch per reduction/back-substitution phase; every (row, mode
//     pair) of a phase is independent.  The reduction's intermediate vectors
//     P(row) live in the grid's second workspace buffer (yh), rows of JP.
// The spectrum spec[q][k][j] (J contiguous) is the mode right-hand side on
// entry (with POT3's boundary terms applied) and the mode solution on exit,
// as for the other mode solver.
#include <hip/hip_runtime.h>

#include "chypot_impl.h"
#include "lanes.h"

namespace chy {

namespace {

struct CrArgs {
    const double *am, *cm, *bm;  // r operator: [JP], [JP], [(N+1)][JP] (rows >= M padded: 0, 0, 1)
    const CrOp *ops;
    const CrProg *progs;
    const CrRow *rows;
    int J, K, M, N, PP, JP;
};

struct Red {
    double al, be, ga, d0, d1, inv;  // row of the interface system and 1/be
};

// 1/x to binary64 accuracy: the hardware estimate and two Newton steps
// (each squares the relative error of the estimate).
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// r, or the identity row where !keep, by selects: a branch here makes the
// wavefront wait for every shuffle in flight before it issues the next ones
__device__ __forceinline__ Red identity_unless(bool keep, const Red &r)
{
    return Red{keep ? r.al : 0.0, keep ? r.be : 1.0, keep ? r.ga : 0.0,
               keep ? r.d0 : 0.0, keep ? r.d1 : 0.0, keep ? r.inv : 1.0};
}

// neighbour s lanes below / above within a group of G lanes; identity rows
// outside the group (distance 1 by DPP moves)
template <int G>
__device__ __forceinline__ Red red_up(const Red &r, int s, int gl)
{
    Red o;
    if (s == 1) {
        o = Red{lane_prev(r.al), lane_prev(r.be), lane_prev(r.ga), lane_prev(r.d0), lane_prev(r.d1),
                lane_prev(r.inv)};
    } else {
        o.al = __shfl_up(r.al, s, G);
        o.be = __shfl_up(r.be, s, G);
        o.ga = __shfl_up(r.ga, s, G);
        o.d0 = __shfl_up(r.d0, s, G);
        o.d1 = __shfl_up(r.d1, s, G);
        o.inv = __shfl_up(r.inv, s, G);
    }
    return identity_unless(gl >= s, o);
}

template <int G>
__device__ __forceinline__ Red red_down(const Red &r, int s, int gl)
{
    Red o;
    if (s == 1) {
        o = Red{lane_next(r.al), lane_next(r.be), lane_next(r.ga), lane_next(r.d0), lane_next(r.d1),
                lane_next(r.inv)};
    } else {
        o.al = __shfl_down(r.al, s, G);
        o.be = __shfl_down(r.be, s, G);
        o.ga = __shfl_down(r.ga, s, G);
        o.d0 = __shfl_down(r.d0, s, G);
        o.d1 = __shfl_down(r.d1, s, G);
        o.inv = __shfl_down(r.inv, s, G);
    }
    return identity_unless(gl + s < G, o);
}

// (T_m - rt I) x = d for two right-hand sides over a group of G lanes (R*G
// rows).  gl: lane within the group; ca, cc: the lane's sub/super-diagonal
// rows, bd: its diagonal rows of T_m, live[r]: 1 for rows < M, 0 for padding
// rows (identity rows: diagonal 1, x = d).
template <int R, int G>
__device__ __forceinline__ void tri_solve(const double (&ca)[R], const double (&bd)[R], const double (&live)[R],
                                          const double (&cc)[R], double rt, int gl, const double (&d0)[R],
                                          const double (&d1)[R], double (&x0)[R], double (&x1)[R])
{
    double cb[R];
#pragma unroll
    for (int r = 0; r < R; ++r) cb[r] = bd[r] - live[r] * rt;

    // rows f..f+R-2 in terms of XL = x_{f-1} and XR = x_{f+R-1}:
    // x_i = D_i + P_i XL + Q_i XR
    double D0[R], D1[R], P[R], Q[R];
    Red red;
    if constexpr (R == 1) {
        red = Red{ca[0], cb[0], cc[0], d0[0], d1[0], 0.0};
    } else {
        double g[R], w0[R], w1[R], wp[R], wq[R];
#pragma unroll
        for (int i = 0; i < R - 1; ++i) {
            const double den = i == 0 ? cb[0] : cb[i] - ca[i] * g[i - 1];
            const double inv = rcp_nr(den);
            g[i] = cc[i] * inv;
            const double rp = i == 0 ? -ca[0] : 0.0;
            const double rq = i == R - 2 ? -cc[R - 2] : 0.0;
            w0[i] = (d0[i] - (i ? ca[i] * w0[i - 1] : 0.0)) * inv;
            w1[i] = (d1[i] - (i ? ca[i] * w1[i - 1] : 0.0)) * inv;
            wp[i] = (rp - (i ? ca[i] * wp[i - 1] : 0.0)) * inv;
            wq[i] = (rq - (i ? ca[i] * wq[i - 1] : 0.0)) * inv;
        }
        D0[R - 2] = w0[R - 2]; D1[R - 2] = w1[R - 2]; P[R - 2] = wp[R - 2]; Q[R - 2] = wq[R - 2];
#pragma unroll
        for (int i = R - 3; i >= 0; --i) {
            D0[i] = w0[i] - g[i] * D0[i + 1];
            D1[i] = w1[i] - g[i] * D1[i + 1];
            P[i] = wp[i] - g[i] * P[i + 1];
            Q[i] = wq[i] - g[i] * Q[i + 1];
        }
        const double nD0 = lane_next(D0[0]), nD1 = lane_next(D1[0]);
        const double nP = lane_next(P[0]), nQ = lane_next(Q[0]);
        const bool last = gl == G - 1;
        const double al = ca[R - 1], be = cb[R - 1], c = last ? 0.0 : cc[R - 1];
        red.al = al * P[R - 2];
        red.be = be + al * Q[R - 2] + c * (last ? 0.0 : nP);
        red.ga = c * (last ? 0.0 : nQ);
        red.d0 = d0[R - 1] - al * D0[R - 2] - c * (last ? 0.0 : nD0);
        red.d1 = d1[R - 1] - al * D1[R - 2] - c * (last ? 0.0 : nD1);
    }
    red.inv = rcp_nr(red.be);

    // parallel cyclic reduction of the interface rows over the group; each
    // lane's 1/be travels with its row
#pragma unroll
    for (int s = 1; s < G; s <<= 1) {
        const Red lo = red_up<G>(red, s, gl);
        const Red hi = red_down<G>(red, s, gl);
        const double k1 = red.al * lo.inv;
        const double k2 = red.ga * hi.inv;
        Red nr;
        nr.al = -lo.al * k1;
        nr.ga = -hi.ga * k2;
        nr.be = red.be - lo.ga * k1 - hi.al * k2;
        nr.d0 = red.d0 - lo.d0 * k1 - hi.d0 * k2;
        nr.d1 = red.d1 - lo.d1 * k1 - hi.d1 * k2;
        nr.inv = rcp_nr(nr.be);
        red = nr;
    }
    const double xr0 = red.d0 * red.inv, xr1 = red.d1 * red.inv;
    x0[R - 1] = xr0;
    x1[R - 1] = xr1;
    if constexpr (R > 1) {
        double xl0 = lane_prev(xr0), xl1 = lane_prev(xr1);
        if (gl == 0) { xl0 = 0.0; xl1 = 0.0; }
#pragma unroll
        for (int i = 0; i < R - 1; ++i) {
            x0[i] = D0[i] + P[i] * xl0 + Q[i] * xr0;
            x1[i] = D1[i] + P[i] * xl1 + Q[i] * xr1;
        }
    }
}

// The r-operator rows one lane owns for order m.
template <int R>
struct Rows {
    double ca[R], cb[R], cc[R], live[R];
    __device__ void load(const CrArgs &a, int m, int f)
    {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ca[r] = a.am[f + r];
            cc[r] = a.cm[f + r];
            cb[r] = a.bm[(size_t)m * a.JP + f + r];
            live[r] = f + r < a.M ? 1.0 : 0.0;
        }
    }
};

// A chain program (a product of scalar factors, pot/blktri.f:197-236 with
// NA > 0) applied in place to (y0, y1).  The next op's scalars are fetched
// while the current solve runs.
template <int R, int G>
__device__ void run_chain(const CrArgs &a, const Rows<R> &c, int pi, int gl, double (&y0)[R], double (&y1)[R])
{
    const CrProg p = a.progs[pi];
    double w0[R], w1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        w0[r] = p.sgn * y0[r];
        w1[r] = p.sgn * y1[r];
        y0[r] = p.pre * w0[r];
        y1[r] = p.pre * w1[r];
    }
    if (p.nsolve <= 0) return;
    const CrOp *op = a.ops + p.off;
    double rt = op[0].rt, f = op[0].f;
    for (int i = 0; i < p.nsolve; ++i) {
        const int ni = i + 1 < p.nsolve ? i + 1 : i;
        const double rt_n = op[ni].rt, f_n = op[ni].f;
        tri_solve<R, G>(c.ca, c.cb, c.live, c.cc, rt, gl, y0, y1, w0, w1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            y0[r] = f * w0[r];
            y1[r] = f * w1[r];
        }
        rt = rt_n;
        f = f_n;
    }
}

// This wavefront's share of a partial-fraction program: the sum over ops
// i = first, first + stride, ... of f_i (T - rt_i)^-1 (sgn x), into (s0, s1).
template <int R, int G>
__device__ void run_pf(const CrArgs &a, const Rows<R> &c, int pi, int gl, int first, int stride,
                       const double (&x0)[R], const double (&x1)[R], double (&s0)[R], double (&s1)[R])
{
    const CrProg p = a.progs[pi];
    double xs0[R], xs1[R], w0[R], w1[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        xs0[r] = p.sgn * x0[r];
        xs1[r] = p.sgn * x1[r];
        s0[r] = 0.0;
        s1[r] = 0.0;
    }
    if (first >= p.nsolve) return;
    const CrOp *op = a.ops + p.off;
    double rt = op[first].rt, f = op[first].f;
    for (int i = first; i < p.nsolve; i += stride) {
        const int ni = i + stride < p.nsolve ? i + stride : i;
        const double rt_n = op[ni].rt, f_n = op[ni].f;
        tri_solve<R, G>(c.ca, c.cb, c.live, c.cc, rt, gl, xs0, xs1, w0, w1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            s0[r] += f * w0[r];
            s1[r] += f * w1[r];
        }
        rt = rt_n;
        f = f_n;
    }
}

template <int R>
__device__ __forceinline__ void load_vec(const double *p, int f, int M, double (&d)[R])
{
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = f + r < M ? p[f + r] : 0.0;
}

template <int R>
__device__ __forceinline__ void store_vec(double *p, int f, int M, const double (&d)[R])
{
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (f + r < M) p[f + r] = d[r];
}

template <int R>
__device__ __forceinline__ void zero_vec(double (&d)[R])
{
#pragma unroll
    for (int r = 0; r < R; ++r) d[r] = 0.0;
}

// One launch per phase; one block of W wavefronts per task (row of the
// phase, group of orders).  A wavefront holds 64/G groups of G lanes, each
// group one order m (cosine and sine parts), so a task covers 64/G orders:
// block b takes row b / NW, NW = ceil((N+1) / (64/G)), and orders
// (b % NW) * 64/G + lane / G; a group past m = N computes on zeros and
// stores nothing.  blockIdx.y is the field.  Within a task:
//   - the two chains of a row (W1 from the row below, W2 from the row above)
//     run on wavefronts 0 and 1 (both on wavefront 0 when W = 1);
//   - the partial-fraction product's solves are dealt round-robin to the W
//     wavefronts; their partial sums meet in LDS and wavefront 0 adds them
//     in a fixed order ((0 + 1) + (2 + 3)), so results do not depend on timing.
// LDS slots hold one wavefront's two vectors, element (r, rhs, lane) at
// ((slot * R + r) * 2 + rhs) * 64 + lane (consecutive lanes, no bank clash).
template <int R, int G, int W, int KIND>
__global__ __launch_bounds__(64 * W) void k_cr_group(CrArgs a, const Fields F, CrPhase ph)
{
    constexpr int SPW = 64 / G;
    constexpr int NSLOT = W + 2;  // W partial sums, then W1 and W2
    __shared__ double lds[W > 1 ? NSLOT * R * 2 * 64 : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int gl = lane % G;
    const int nw = (a.N + 1 + SPW - 1) / SPW;
    const int task = blockIdx.x;
    if (task >= ph.nrow * nw) return;  // the whole block leaves together
    const int mr = (task % nw) * SPW + lane / G;
    const bool one = mr <= a.N;
    const int m = one ? mr : a.N;
    const bool two = one && m >= 1 && m <= a.N - 1;
    const CrRow row = a.rows[ph.row0 + task / nw];
    double *spec = static_cast<double *>(F.spec[blockIdx.y]);
    double *pv = static_cast<double *>(F.yh[blockIdx.y]);
    const size_t qc = (size_t)m, qs = (size_t)(a.N + m);
    const int f = gl * R, M = a.M;
    // Y(., i) and P(., i) of mode q, i = 1..N
    auto ycol = [&](size_t q, int i) { return spec + (q * a.K + (size_t)(i - 1)) * a.J; };
    auto pcol = [&](size_t q, int i) { return pv + (q * a.PP + (size_t)(i - 1)) * a.JP; };
    auto load2 = [&](bool fromp, int i, double (&v0)[R], double (&v1)[R]) {
        if (one) load_vec<R>(fromp ? pcol(qc, i) : ycol(qc, i), f, M, v0); else zero_vec<R>(v0);
        if (two) load_vec<R>(fromp ? pcol(qs, i) : ycol(qs, i), f, M, v1); else zero_vec<R>(v1);
    };
    auto store2 = [&](bool top, int i, const double (&v0)[R], const double (&v1)[R]) {
        if (one) store_vec<R>(top ? pcol(qc, i) : ycol(qc, i), f, M, v0);
        if (two) store_vec<R>(top ? pcol(qs, i) : ycol(qs, i), f, M, v1);
    };
    auto put = [&](int slot, const double (&v0)[R], const double (&v1)[R]) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            lds[((slot * R + r) * 2 + 0) * 64 + lane] = v0[r];
            lds[((slot * R + r) * 2 + 1) * 64 + lane] = v1[r];
        }
    };
    auto get = [&](int slot, int r, int h) { return lds[((slot * R + r) * 2 + h) * 64 + lane]; };

    Rows<R> c;
    c.load(a, m, f);
    double y0[R], y1[R], s0[R], s1[R];

    // sum of the W partial sums (this wavefront's own in s0/s1) on wavefront 0
    auto reduce = [&]() -> bool {
        if constexpr (W > 1) {
            put(wave, s0, s1);
            __syncthreads();
            if (wave != 0) return false;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                for (int h = 0; h < 2; ++h) {
                    double acc;
                    if constexpr (W == 2) acc = get(0, r, h) + get(1, r, h);
                    else acc = (get(0, r, h) + get(1, r, h)) + (get(2, r, h) + get(3, r, h));
                    (h ? s1 : s0)[r] = acc;
                }
            }
        }
        return true;
    };

    if constexpr (KIND == CR_PHASE_ODD) {
        // P(o) = pm(Y(o))  (pot/blktri.f:86-87, 106-107)
        load2(false, row.row, y0, y1);
        run_pf<R, G>(a, c, row.pm, gl, wave, W, y0, y1, s0, s1);
        if (reduce()) store2(true, row.row, s0, s1);
        return;
    }
    if constexpr (KIND == CR_PHASE_EVEN) {
        // Y(e) = (W1 + W2) - Y(e), W1 from P(e-h), W2 from P(e+h)  (pot/blktri.f:104-116)
        if (W == 1 || wave == 0) {
            load2(true, row.row - ph.h, s0, s1);
            run_chain<R, G>(a, c, row.pa, gl, s0, s1);
        }
        if (W == 1 || wave == 1) {
            load2(true, row.row + ph.h, y0, y1);
            run_chain<R, G>(a, c, row.pc, gl, y0, y1);
        }
        if constexpr (W > 1) {
            if (wave == 1) put(W + 1, y0, y1);
            __syncthreads();
            if (wave != 0) return;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                y0[r] = get(W + 1, r, 0);
                y1[r] = get(W + 1, r, 1);
            }
        }
        double t0[R], t1[R];
        load2(false, row.row, t0, t1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            y0[r] = (s0[r] + y0[r]) - t0[r];
            y1[r] = (s1[r] + y1[r]) - t1[r];
        }
        store2(false, row.row, y0, y1);
        return;
    }
    if constexpr (KIND == CR_PHASE_BACK) {
        // back substitution: Y(i) = pm((Y(i) - W1) - W2)  (pot/blktri.f:137-167)
        double w10[R], w11[R], w20[R], w21[R];
        zero_vec<R>(w10); zero_vec<R>(w11); zero_vec<R>(w20); zero_vec<R>(w21);
        if (row.pa >= 0 && (W == 1 || wave == 0)) {
            load2(false, row.row - ph.h, w10, w11);
            run_chain<R, G>(a, c, row.pa, gl, w10, w11);
        }
        if (row.pc >= 0 && (W == 1 || wave == 1)) {
            load2(false, row.row + ph.h, w20, w21);
            run_chain<R, G>(a, c, row.pc, gl, w20, w21);
        }
        if constexpr (W > 1) {
            if (wave == 0) put(W, w10, w11);
            if (wave == 1) put(W + 1, w20, w21);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                w10[r] = get(W, r, 0); w11[r] = get(W, r, 1);
                w20[r] = get(W + 1, r, 0); w21[r] = get(W + 1, r, 1);
            }
        }
        load2(false, row.row, y0, y1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            y0[r] = (y0[r] - w10[r]) - w20[r];
            y1[r] = (y1[r] - w11[r]) - w21[r];
        }
        run_pf<R, G>(a, c, row.pm, gl, wave, W, y0, y1, s0, s1);
        if (reduce()) store2(false, row.row, s0, s1);
    }
}

// Wavefronts per task for a phase: the partial-fraction products of level h
// have 2h - 1 solves and the chains h - 1, so short ones take one wavefront
// (many tasks run side by side) and long ones four.
int phase_waves(const CrPhase &ph)
{
    if (ph.h < 4) return 1;
    return ph.kind == CR_PHASE_EVEN ? 2 : 4;
}

// one phase, its kernel specialised on the phase kind (each kind keeps only
// its own vectors live)
template <int R, int G, int KIND>
void launch_phase(const CrPhase &ph, dim3 grid, const CrArgs &a, const Fields &F, hipStream_t s)
{
    switch (phase_waves(ph)) {
    case 1: hipLaunchKernelGGL((k_cr_group<R, G, 1, KIND>), grid, dim3(64), 0, s, a, F, ph); break;
    case 2: hipLaunchKernelGGL((k_cr_group<R, G, 2, KIND>), grid, dim3(128), 0, s, a, F, ph); break;
    default: hipLaunchKernelGGL((k_cr_group<R, G, 4, KIND>), grid, dim3(256), 0, s, a, F, ph); break;
    }
}

template <int R, int G>
int launch_phases(const chy_grid *g, const CrArgs &a, const Fields &F, hipStream_t s)
{
    constexpr int SPW = 64 / G;
    const int nw = (g->N + 1 + SPW - 1) / SPW;
    for (const CrPhase &ph : g->cr.phases) {
        if (ph.nrow <= 0) continue;
        const dim3 grid((unsigned)(ph.nrow * nw), F.nf);
        switch (ph.kind) {
        case CR_PHASE_ODD: launch_phase<R, G, CR_PHASE_ODD>(ph, grid, a, F, s); break;
        case CR_PHASE_EVEN: launch_phase<R, G, CR_PHASE_EVEN>(ph, grid, a, F, s); break;
        default: launch_phase<R, G, CR_PHASE_BACK>(ph, grid, a, F, s); break;
        }
    }
    return (int)hipGetLastError();
}

}  // namespace

bool cr_supported(const chy_grid *g)
{
    const int r = g->JP / 64;
    return g->cr.kx >= 2 && (r == 1 || r == 2 || r == 4 || r == 8);
}

// lanes per r system: 64, or 32 (two orders per wavefront) where JP/32 <= 8
bool cr_lanes_ok(const chy_grid *g, int lanes)
{
    if (lanes == 64) return cr_supported(g);
    if (lanes == 32) return cr_supported(g) && g->JP / 32 <= 8;
    return false;
}

int launch_cr_solve_f(const chy_grid *g, const Fields &F, hipStream_t s)
{
    CrArgs a{};
    a.am = g->d.am;
    a.cm = g->d.cm;
    a.bm = g->d.bm;
    a.ops = g->d.cr_ops;
    a.progs = g->d.cr_progs;
    a.rows = g->d.cr_rows;
    a.J = g->J; a.K = g->K; a.M = g->M; a.N = g->N; a.PP = g->PP; a.JP = g->JP;
    if (g->cr_lanes == 32) {
        switch (g->JP / 32) {
        case 2: return launch_phases<2, 32>(g, a, F, s);
        case 4: return launch_phases<4, 32>(g, a, F, s);
        case 8: return launch_phases<8, 32>(g, a, F, s);
        default: break;
        }
    }
    switch (g->JP / 64) {
    case 1: return launch_phases<1, 64>(g, a, F, s);
    case 2: return launch_phases<2, 64>(g, a, F, s);
    case 4: return launch_phases<4, 64>(g, a, F, s);
    case 8: return launch_phases<8, 64>(g, a, F, s);
    default: return (int)hipErrorInvalidValue;
    }
}

}  // namespace chy
