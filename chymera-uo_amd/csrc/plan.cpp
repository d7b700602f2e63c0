// Host-side preparation of a grid: everything the potential solve needs that
// depends on the grid alone, computed once (the work SETBDY does for the
// boundary, pot/boundary.f:2-82, plus the grid-only parts of POT3 and BLKTRI),
// then uploaded and kept resident on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "chypot_impl.h"

namespace chy {

// ---------------------------------------------------------------------------
// Multipole terms with l+m even, l <= 10 (the ISYM=2 set: F1=0 removes the
// odd-l m=0 terms and the ISYMA=2 branches skip l+m odd, boundary.f:345-354,
// 412-500).  BLM polynomials follow the statement functions B20..B108
// (boundary.f:204-253) written as s^m * P(c^2); DLM as boundary.f:694-772
// with the TW constants as REAL*8 values (the decimal TW15..TW18 of
// boundary.f:184-185 included, as a real-8 build uses them).
static constexpr double TW5 = 0.03125, TW6 = 0.015625, TW7 = 7.8125e-3, TW8 = 3.90625e-3,
                        TW9 = 1.953125e-3, TW10 = 9.765625e-04, TW11 = 4.8828125e-4,
                        TW12 = 2.44140625e-4, TW13 = 1.220703125e-4, TW14 = 6.103515625e-5,
                        TW15 = 3.051757813e-5, TW16 = 1.525878907e-5, TW17 = 7.629394535e-6,
                        TW18 = 3.814697266e-6;

static const Term kTerms[kMaxTerms] = {
    {0, 0, 1.0, 1, {1}},
    {1, 1, 0.5, 1, {1}},
    {2, 0, 0.25, 2, {-1, 3}},
    {2, 2, 0.375, 1, {1}},
    {3, 1, 0.1875, 2, {-1, 5}},
    {3, 3, 0.3125, 1, {1}},
    {4, 0, TW6, 3, {3, -30, 35}},
    {4, 2, 5.0 * TW5, 2, {-1, 7}},
    {4, 4, 35.0 * TW7, 1, {1}},
    {5, 1, 15.0 * TW7, 3, {1, -14, 21}},
    {5, 3, 35.0 * TW8, 2, {-1, 9}},
    {5, 5, 63.0 * TW8, 1, {1}},
    {6, 0, TW8, 4, {-5, 105, -315, 231}},
    {6, 2, 105.0 * TW10, 3, {1, -18, 33}},
    {6, 4, 63.0 * TW9, 2, {-1, 11}},
    {6, 6, 231.0 * TW10, 1, {1}},
    {7, 1, 7.0 * TW11, 4, {-5, 135, -495, 429}},
    {7, 3, 21.0 * TW11, 3, {3, -66, 143}},
    {7, 5, 231.0 * TW11, 2, {-1, 13}},
    {7, 7, 429.0 * TW11, 1, {1}},
    {8, 0, TW14, 5, {35, -1260, 6930, -12012, 6435}},
    {8, 2, 315.0 * TW12, 4, {-1, 33, -143, 143}},
    {8, 4, 693.0 * TW13, 3, {1, -26, 65}},
    {8, 6, 429.0 * TW12, 2, {-1, 15}},
    {8, 8, 6435.0 * TW15, 1, {1}},
    {9, 1, 45.0 * TW15, 5, {7, -308, 2002, -4004, 2431}},
    {9, 3, 1155.0 * TW14, 4, {-1, 39, -195, 221}},
    {9, 5, 1287.0 * TW14, 3, {1, -30, 85}},
    {9, 7, 6435.0 * TW16, 2, {-1, 17}},
    {9, 9, 12155.0 * TW16, 1, {1}},
    {10, 0, TW16, 6, {-63, 3465, -30030, 90090, -109395, 46189}},
    {10, 2, 165.0 * TW17, 5, {7, -364, 2730, -6188, 4199}},
    {10, 4, 2145.0 * TW15, 4, {-1, 45, -255, 323}},
    {10, 6, 2145.0 * TW18, 3, {3, -102, 323}},
    {10, 8, 12155.0 * TW17, 2, {-1, 19}},
    {10, 10, 46189.0 * TW18, 1, {1}},
};

const Term *terms_table(int *count)
{
    if (count) *count = kMaxTerms;
    return kTerms;
}

static double blm(const Term &t, double c, double s)
{
    double c2 = c * c, p = 0.0, cc = 1.0;
    for (int i = 0; i < t.np; ++i) {
        p += t.poly[i] * cc;
        cc *= c2;
    }
    double sm = 1.0;
    for (int i = 0; i < t.m; ++i) sm *= s;
    return p * sm;
}

static double ipow(double x, int n)
{
    double r = 1.0;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
}

// ---------------------------------------------------------------------------
// Symmetric tridiagonal eigensystem by the implicit-shift QL iteration with
// Wilkinson shifts, accumulating the rotations.  d: diagonal in, eigenvalues
// out; e[1..n-1]: off-diagonal (destroyed); q: n*n, q[k*n+p] = component k of
// eigenvector p.  Run in long double so that the small eigenvalues -- the
// smooth z modes, whose solves amplify any error most -- carry full binary64
// relative accuracy after rounding.
int symmetric_tridiag_eigen(int n, std::vector<long double> &d, std::vector<long double> &e,
                            std::vector<long double> &q)
{
    q.assign((size_t)n * n, 0.0L);
    for (int i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0L;
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    if (n > 0) e[n - 1] = 0.0L;
    const long double eps = std::numeric_limits<long double>::epsilon();
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                long double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 80) return CHY_ENUMERIC;
            long double g = (d[l + 1] - d[l]) / (2.0L * e[l]);
            long double r = std::hypot(g, 1.0L);
            g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
            long double s = 1.0L, c = 1.0L, p = 0.0L;
            int i = m - 1;
            bool early = false;
            for (; i >= l; --i) {
                long double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0L) {
                    d[i + 1] -= p;
                    e[m] = 0.0L;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0L * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int k = 0; k < n; ++k) {
                    long double zf = q[(size_t)k * n + i + 1];
                    q[(size_t)k * n + i + 1] = s * q[(size_t)k * n + i] + c * zf;
                    q[(size_t)k * n + i] = c * q[(size_t)k * n + i] - s * zf;
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0L;
        }
    }
    return CHY_OK;
}

// ---------------------------------------------------------------------------

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
static int round_up(int v, int m) { return (v + m - 1) / m * m; }

template <typename T>
static int upload(chy_grid *g, T **dst, const std::vector<T> &src)
{
    size_t bytes = sizeof(T) * std::max<size_t>(src.size(), 1);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        set_device_error(e, "hipMalloc");
        return CHY_ENOMEM;
    }
    g->allocations.push_back(p);
    g->device_bytes += bytes;
    if (!src.empty()) {
        e = hipMemcpy(p, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_device_error(e, "hipMemcpy");
            return CHY_EDEVICE;
        }
    }
    *dst = static_cast<T *>(p);
    return CHY_OK;
}

template <typename T>
static int reserve(chy_grid *g, T **dst, size_t count)
{
    size_t bytes = sizeof(T) * std::max<size_t>(count, 1);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        set_device_error(e, "hipMalloc");
        return CHY_ENOMEM;
    }
    e = hipMemset(p, 0, bytes);
    if (e != hipSuccess) {
        set_device_error(e, "hipMemset");
        return CHY_EDEVICE;
    }
    g->allocations.push_back(p);
    g->device_bytes += bytes;
    *dst = static_cast<T *>(p);
    return CHY_OK;
}

#define CHY_TRY(x)                   \
    do {                             \
        int rc_ = (x);               \
        if (rc_ != CHY_OK) return rc_; \
    } while (0)

int build_grid(chy_grid *g, const chy_grid_desc *desc, int isym)
{
    const int J = desc->jmax, K = desc->kmax, L = desc->lmax;
    if (!desc->r || !desc->z || !desc->rhf || !desc->zhf) return CHY_EINVAL;
    if (J < 3 || K < 2 || L < 2) return CHY_EINVAL;
    // BLKTRI needs N = KMAX-1 = 2^k-1 with k >= 2 (pot/blktri.f:19-31)
    if (!is_pow2(K) || K < 4) return CHY_EUNSUPPORTED;
    // LMAX/2 may have prime factors up to 23, as in pot/fft.f:111-126
    if (!fft_size_ok(L)) return CHY_EUNSUPPORTED;
    if (J > 4096) return CHY_EUNSUPPORTED;
    if (std::abs(isym) != 2) return CHY_EUNSUPPORTED;
    if (!(desc->dtheta > 0.0) || !std::isfinite(desc->pi) || !std::isfinite(desc->grav))
        return CHY_EINVAL;

    g->J = J; g->K = K; g->L = L;
    g->N = L / 2;
    g->M = J - 1;
    g->NZ = K - 1;
    g->PP = round_up(g->NZ, 64);
    {
        // the r solve gives each of a wavefront's 64 lanes R consecutive rows,
        // R a power of two
        int r = 1;
        while (64 * r < g->M) r *= 2;
        g->JP = 64 * r;
    }
    g->isym = isym;
    g->dtheta = desc->dtheta; g->pi = desc->pi; g->grav = desc->grav;
    {
        int ncu = 0;
        hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, g->ctx->device);
        if (e != hipSuccess) {
            set_device_error(e, "hipDeviceGetAttribute");
            return CHY_EDEVICE;
        }
        g->ncu = ncu > 0 ? ncu : 1;
    }
    g->r.assign(desc->r, desc->r + J + 2);
    g->z.assign(desc->z, desc->z + K + 2);
    g->rhf.assign(desc->rhf, desc->rhf + J + 2);
    g->zhf.assign(desc->zhf, desc->zhf + K + 2);
    for (int j = 1; j < J + 2; ++j)
        if (!(g->rhf[j] > g->rhf[j - 1])) return CHY_EINVAL;
    for (int k = 1; k < K + 2; ++k)
        if (!(g->zhf[k] > g->zhf[k - 1])) return CHY_EINVAL;
    if (!(g->rhf[1] > 0.0) || !(g->zhf[1] > 0.0)) return CHY_EINVAL;

    const int M = g->M, NZ = g->NZ, N = g->N, PP = g->PP, JP = g->JP;
    // 1-based views, as the reference indexes them
    const double *rhf = g->rhf.data() - 1, *zhf = g->zhf.data() - 1;

    // ---- mode-solve coefficients, pot3.f:158-194 (binary64, same order) ----
    std::vector<double> rd2(J + 1), rd3(J + 1), denr(J + 1), zd2(K + 1), denz(K + 1);
    std::vector<double> am(M), cm(M), c(M), an(NZ), bn(NZ), cn(NZ);
    double dthet2 = 1.0 / (g->dtheta * g->dtheta);
    rd2[1] = 0.5 / rhf[2];
    rd3[1] = 1.0 / rhf[1];
    for (int j = 2; j <= J; ++j) {
        denr[j] = 1.0 / (rhf[j + 1] - rhf[j - 1]);
        rd2[j] = 1.0 / (rhf[j + 1] - rhf[j]);
        rd3[j] = 1.0 / rhf[j];
    }
    zd2[1] = 0.5 / zhf[2];
    for (int k = 2; k <= K; ++k) {
        denz[k] = 1.0 / (zhf[k + 1] - zhf[k - 1]);
        zd2[k] = 1.0 / (zhf[k + 1] - zhf[k]);
    }
    for (int j = 2; j <= J; ++j) {
        cm[j - 2] = 2.0 * (rd2[j] + 0.5 * rd3[j]) * denr[j];
        am[j - 2] = 2.0 * (rd2[j - 1] - 0.5 * rd3[j]) * denr[j];
        c[j - 2] = rd3[j] * rd3[j] * dthet2;
    }
    for (int k = 2; k <= K; ++k) {
        cn[k - 2] = 2.0 * zd2[k] * denz[k];
        an[k - 2] = 2.0 * zd2[k - 1] * denz[k];
        bn[k - 2] = -cn[k - 2] - an[k - 2];
    }
    g->cmmax = cm[M - 1]; cm[M - 1] = 0.0;
    g->ammin = am[0];     am[0] = 0.0;
    g->cnmax = cn[NZ - 1]; cn[NZ - 1] = 0.0;
    bn[0] = -cn[0];       an[0] = 0.0;
    g->pig4 = 4.0 * g->pi * g->grav;

    // BM for every azimuthal order m = 0..N (pot3.f:204-230)
    std::vector<double> bmt((size_t)(N + 1) * M);
    for (int m = 0; m <= N; ++m) {
        double powrm = (m % 2) ? -1.0 : 1.0;
        double xlam = std::cos(m * g->dtheta);
        double *bm = &bmt[(size_t)m * M];
        bm[0] = -cm[0] + (powrm - 1.0) * g->ammin + 2.0 * (xlam - 1.0) * c[0];
        for (int j = 2; j <= J - 2; ++j) bm[j - 1] = -cm[j - 1] - am[j - 1] + 2.0 * (xlam - 1.0) * c[j - 1];
        bm[M - 1] = -g->cmmax - am[M - 1] + 2.0 * (xlam - 1.0) * c[M - 1];
    }

    // ---- z operator eigenbasis (replaces BLKTRI's root tables, blktri.f:264-344) ----
    // T_z = tridiag(AN, BN, CN); D T_z D^-1 is symmetric with
    // d_k = d_{k-1} sqrt(CN_{k-1}/AN_k), e_k = sqrt(AN_k CN_{k-1}).
    std::vector<long double> dsc(NZ), dg(NZ), ev(NZ + 1, 0.0L), q;
    dsc[0] = 1.0L;
    for (int k = 1; k < NZ; ++k) {
        if (!(an[k] * cn[k - 1] > 0.0)) return CHY_ENUMERIC;
        dsc[k] = dsc[k - 1] * std::sqrt((long double)cn[k - 1] / (long double)an[k]);
        ev[k] = std::sqrt((long double)an[k] * (long double)cn[k - 1]);
    }
    for (int k = 0; k < NZ; ++k) dg[k] = bn[k];
    CHY_TRY(symmetric_tridiag_eigen(NZ, dg, ev, q));
    std::vector<double> wf((size_t)NZ * PP, 0.0), wi((size_t)NZ * PP, 0.0), lam(PP, 1.0);
    for (int p = 0; p < NZ; ++p) lam[p] = (double)dg[p];
    for (int k = 0; k < NZ; ++k)
        for (int p = 0; p < NZ; ++p) {
            long double qv = q[(size_t)k * NZ + p];
            wf[(size_t)k * PP + p] = (double)(dsc[k] * qv);
            wi[(size_t)p * PP + k] = (double)(qv / dsc[k]);  // stored [p][k]: reduction index major
        }

    // ---- BLKTRI's root tables and the cyclic reduction's programs (crplan.cpp) ----
    CHY_TRY(build_cr_plan(g->cr, NZ, an.data(), bn.data(), cn.data()));
    g->mode_solver = cr_supported(g) ? CHY_MODES_CR : CHY_MODES_EIGEN;
    // two orders per wavefront where the r system fits 32 lanes: measured
    // faster or equal on every grid class tried (chypot/tuned_mi355x.json)
    g->cr_lanes = cr_lanes_ok(g, 32) ? 32 : 64;

    // ---- FFT twiddles (exact angles; the reference uses recurrences, fft.f:225-232) ----
    std::vector<double2> twL(L), tw64(64);
    for (int e = 0; e < L; ++e) {
        double ang = 2.0 * 3.14159265358979323846 * (double)e / (double)L;
        twL[e] = make_double2(std::cos(ang), std::sin(ang));
    }
    for (int e = 0; e < 64; ++e) {
        double ang = 2.0 * 3.14159265358979323846 * (double)e / 64.0;
        tw64[e] = make_double2(std::cos(ang), std::sin(ang));
    }
    std::vector<double2> twh(kMaxTrm + 1);
    for (int m = 0; m <= kMaxTrm; ++m) {
        double ang = 3.14159265358979323846 * (double)m / (double)L;
        twh[m] = make_double2(std::cos(ang), std::sin(ang));
    }

    // ---- ZAXPHI abscissae source: RHF(1..11) (pot3.f:342-348) ----
    std::vector<double> zx(11, 0.0);
    for (int i = 0; i < 11 && i < J + 2; ++i) zx[i] = g->rhf[i];

    // ---- BDYGEN tables ----
    const int nring = (J - 1) * (K - 1);
    g->nring = nring;
    std::vector<double> rsp(nring), dvr(nring), c1r(nring), s1r(nring);
    for (int kk = 2; kk <= K; ++kk) {
        double zz = zhf[kk], z2 = zz * zz, zdel = g->z[kk] - g->z[kk - 1];
        for (int jj = 2; jj <= J; ++jj) {
            int ir = (jj - 2) + (J - 1) * (kk - 2);
            double rr = rhf[jj], r2 = rr * rr, rs = std::sqrt(r2 + z2);
            rsp[ir] = rs;
            c1r[ir] = zz / rs;
            s1r[ir] = rr / rs;
            // DV = 0.5*DTHETA*ZDEL*(R(JJ+1)**2-R(JJ)**2)*F2, F2=2 for ISYM=2 (boundary.f:516)
            dvr[ir] = 0.5 * g->dtheta * zdel * (g->r[jj] * g->r[jj] - g->r[jj - 1] * g->r[jj - 1]) * 2.0;
        }
    }
    std::vector<int> order(nring);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rsp[a] < rsp[b]; });
    std::vector<double> rs_sorted(nring), tin((size_t)nring * kMaxTerms), tout((size_t)nring * kMaxTerms);
    for (int pos = 0; pos < nring; ++pos) {
        int ir = order[pos];
        double rs = rsp[ir];
        rs_sorted[pos] = rs;
        for (int t = 0; t < kMaxTerms; ++t) {
            const Term &tt = kTerms[t];
            double term = blm(tt, c1r[ir], s1r[ir]) * ipow(rs, tt.l);  // TERM(I), boundary.f:412-500
            double qh = 1.0 / ipow(rs, 2 * tt.l + 1);                 // RSPHER**(-(2L+1)), boundary.f:587
            tin[(size_t)pos * kMaxTerms + t] = term;
            tout[(size_t)pos * kMaxTerms + t] = term * qh;
        }
    }
    // boundary points of SETBDY for ISYM=2 (boundary.f:50-79)
    std::vector<int> jpos, kpos;
    std::vector<double> rb;
    {
        double zk2 = zhf[K + 1] * zhf[K + 1];
        for (int j = 2; j <= J + 1; ++j) {
            rb.push_back(std::sqrt(zk2 + rhf[j] * rhf[j]));
            jpos.push_back(j);
            kpos.push_back(K + 1);
        }
        double rj2 = rhf[J + 1] * rhf[J + 1];
        for (int k = 2; k <= K; ++k) {
            rb.push_back(std::sqrt(rj2 + zhf[k] * zhf[k]));
            jpos.push_back(J + 1);
            kpos.push_back(k);
        }
    }
    const int jkmax = (int)rb.size();
    g->jkmax = jkmax;
    // inside test RSPHER <= RBDY(JK) (boundary.f:561): rings [0, cut) of the sorted list
    std::vector<int> cutpos(jkmax);
    for (int jk = 0; jk < jkmax; ++jk)
        cutpos[jk] = (int)(std::upper_bound(rs_sorted.begin(), rs_sorted.end(), rb[jk]) - rs_sorted.begin());
    std::vector<int> bounds(cutpos);
    bounds.push_back(0);
    bounds.push_back(nring);
    std::sort(bounds.begin(), bounds.end());
    bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
    std::vector<int> chunk_start;
    for (size_t s = 0; s + 1 < bounds.size(); ++s)
        for (int p = bounds[s]; p < bounds[s + 1]; p += kChunkRings) chunk_start.push_back(p);
    chunk_start.push_back(nring);
    const int nchunk = (int)chunk_start.size() - 1;
    g->nchunk = nchunk;
    std::vector<int> jk_cut(jkmax);
    for (int jk = 0; jk < jkmax; ++jk)
        jk_cut[jk] = (int)(std::lower_bound(chunk_start.begin(), chunk_start.end(), cutpos[jk]) - chunk_start.begin());
    std::vector<double> tb((size_t)jkmax * kMaxTerms);
    for (int jk = 0; jk < jkmax; ++jk) {
        double rs = rb[jk], c1 = zhf[kpos[jk]] / rs, s1 = rhf[jpos[jk]] / rs;
        for (int t = 0; t < kMaxTerms; ++t) tb[(size_t)jk * kMaxTerms + t] = kTerms[t].dlm * blm(kTerms[t], c1, s1);
    }
    std::vector<double> cosm((size_t)L * 10), sinm((size_t)L * 10);
    for (int l = 1; l <= L; ++l) {
        double psi = (l - 0.5) * g->dtheta;
        for (int m = 1; m <= 10; ++m) {
            cosm[(size_t)(l - 1) * 10 + m - 1] = std::cos(m * psi);
            sinm[(size_t)(l - 1) * 10 + m - 1] = std::sin(m * psi);
        }
    }

    // ---- upload ----
    DeviceGrid &d = g->d;
    CHY_TRY(upload(g, &d.twL, twL));
    CHY_TRY(upload(g, &d.tw64, tw64));
    CHY_TRY(upload(g, &d.twh, twh));
    CHY_TRY(upload(g, &d.wf, wf));
    CHY_TRY(upload(g, &d.wi, wi));
    CHY_TRY(upload(g, &d.lam, lam));
    {
        // binary32 copies of the weight tables for the refined solve's fused
        // plane kernel, PP x PP with zero rows past NZ
        std::vector<float> wf32((size_t)PP * PP, 0.0f), wi32((size_t)PP * PP, 0.0f);
        for (int r = 0; r < NZ; ++r)
            for (int c = 0; c < PP; ++c) {
                wf32[(size_t)r * PP + c] = (float)wf[(size_t)r * PP + c];
                wi32[(size_t)r * PP + c] = (float)wi[(size_t)r * PP + c];
            }
        CHY_TRY(upload(g, &d.wf32, wf32));
        CHY_TRY(upload(g, &d.wi32, wi32));
    }
    {
        // r coefficients padded to JP rows per order (rows >= M: AM = CM = 0,
        // BM = 1) so that the r solves read whole aligned rows
        const int JP = g->JP;
        std::vector<double> amp(JP, 0.0), cmp(JP, 0.0), bmp((size_t)(N + 1) * JP, 1.0);
        std::copy(am.begin(), am.end(), amp.begin());
        std::copy(cm.begin(), cm.end(), cmp.begin());
        for (int m = 0; m <= N; ++m)
            std::copy(bmt.begin() + (size_t)m * M, bmt.begin() + (size_t)(m + 1) * M, bmp.begin() + (size_t)m * JP);
        CHY_TRY(upload(g, &d.am, amp));
        CHY_TRY(upload(g, &d.cm, cmp));
        CHY_TRY(upload(g, &d.bm, bmp));
    }
    CHY_TRY(upload(g, &d.zx, zx));
    CHY_TRY(upload(g, &d.cr_ops, g->cr.ops));
    CHY_TRY(upload(g, &d.cr_progs, g->cr.progs));
    CHY_TRY(upload(g, &d.cr_rows, g->cr.rows));
    // the z operator of the same system for the refinement's spectral residuals
    CHY_TRY(upload(g, &d.rz_an, an));
    CHY_TRY(upload(g, &d.rz_cn, cn));
    CHY_TRY(upload(g, &d.rz_bn, bn));
    CHY_TRY(upload(g, &d.ring_of_pos, order));
    CHY_TRY(upload(g, &d.tin, tin));
    CHY_TRY(upload(g, &d.tout, tout));
    CHY_TRY(upload(g, &d.rsph, rs_sorted));
    CHY_TRY(upload(g, &d.dv, dvr));
    CHY_TRY(upload(g, &d.chunk_start, chunk_start));
    CHY_TRY(upload(g, &d.jk_cut, jk_cut));
    CHY_TRY(upload(g, &d.tb, tb));
    CHY_TRY(upload(g, &d.rb, rb));
    CHY_TRY(upload(g, &d.jpos, jpos));
    CHY_TRY(upload(g, &d.kpos, kpos));
    CHY_TRY(upload(g, &d.cosm, cosm));
    CHY_TRY(upload(g, &d.sinm, sinm));
    CHY_TRY(reserve(g, &d.spec, (size_t)L * K * J));
    CHY_TRY(reserve(g, &d.yh, (size_t)L * PP * JP));
    CHY_TRY(reserve(g, &d.mom, (size_t)nring * kMoments));
    CHY_TRY(reserve(g, &d.chunk, (size_t)nchunk * 4 * kMaxTerms));
    CHY_TRY(reserve(g, &d.massio, (size_t)(nchunk + 1) * 4 * kMaxTerms));
    CHY_TRY(reserve(g, &d.dscal, (size_t)4 + K + 2));
    CHY_TRY(reserve(g, &d.ticket, 1));
    CHY_TRY(reserve(g, &d.zsamp, (size_t)K * kZaxSamples * L));
    if (JP > 1024) {
        // r systems longer than a wavefront holds: the Thomas algorithm, its
        // pivots tabulated once here (on the context's stream, waited for)
        CHY_TRY(reserve(g, &d.rinv, (size_t)(N + 1) * NZ * JP));
        hipError_t e = (hipError_t)launch_thomas_pivots(g, g->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
        if (e != hipSuccess) {
            set_device_error(e, "r-solve pivots");
            return CHY_EDEVICE;
        }
    }
    return CHY_OK;
}

}  // namespace chy
