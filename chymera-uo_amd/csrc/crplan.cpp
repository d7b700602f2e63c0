// Host-side plan of the per-mode cyclic reduction (pot/blktri.f): the root
// tables BLKTRI's initialisation computes (COMPB, pot/blktri.f:264-344, with
// the eigenvalues of TQLRAT, :348-423) and, for every row the reduction and
// back-substitution touch, the r solves and combinations of PRDCT
// (:177-260) as flat op lists: the products of scalar factors as BLKTRI
// chains them, the products with multiplier roots in partial-fraction form
// (independent solves, summed).
//
// The roots are formed with the reference's own arithmetic and convergence
// rule.  That matters for parity: the reference's FP64 result differs from the
// exact solution of the discrete system by up to a few 1e-12 (256x128x256),
// and that difference comes from these root values, not from the rounding of
// the r solves (any elimination order with the same roots reproduces it to
// ~1e-14).
#include <cmath>
#include <vector>

#include "chypot_impl.h"

namespace chy {

namespace {

// 100 x the first power of ten below binary64 resolution, formed by repeated
// division as pot/blktri.f:271-277 forms its EPSLON.
double cr_machep()
{
    volatile double e = 1.0, one_plus;
    do {
        e = e / 10.0;
        one_plus = 1.0 + e;
    } while (one_plus > 1.0);
    return 100.0 * e;
}

// Eigenvalues of the symmetric tridiagonal with diagonal d[0..n-1] and
// squared couplings e2[0..n-2] (e2[i] couples rows i and i+1): Reinsch's
// rational QL iteration (EISPACK TQLRAT), with the convergence test and
// ordering of pot/blktri.f:348-423: the result is sorted by increasing
// magnitude.  Returns 0, or the 1-based row that failed to converge.
int rational_ql(int n, double *d, double *e2, double machep)
{
    if (n == 1) return 0;
    e2[n - 1] = 0.0;
    double f = 0.0, b = 0.0, c = 0.0;
    for (int l = 0; l < n; ++l) {
        double h = machep * (std::fabs(d[l]) + std::sqrt(e2[l]));
        if (!(b > h)) {
            b = h;
            c = b * b;
        }
        int m = l;
        while (!(e2[m] <= c)) ++m;
        if (m != l) {
            for (int it = 0;; ++it) {
                if (it == 30) return l + 1;
                double s = std::sqrt(e2[l]), g = d[l];
                double p = (d[l + 1] - g) / (2.0 * s);
                double r = std::sqrt(p * p + 1.0);
                d[l] = s / (p + std::copysign(r, p));
                h = g - d[l];
                for (int i = l + 1; i < n; ++i) d[i] = d[i] - h;
                f = f + h;
                g = d[m];
                if (g == 0.0) g = b;
                h = g;
                s = 0.0;
                for (int i = m - 1; i >= l; --i) {
                    p = g * h;
                    r = p + e2[i];
                    e2[i + 1] = s * r;
                    s = e2[i] / r;
                    d[i + 1] = h + s * (h + d[i]);
                    g = d[i] - e2[i] / g;
                    if (g == 0.0) g = b;
                    h = g * p / r;
                }
                e2[l] = s * g;
                d[l] = h;
                if (h == 0.0) break;
                if (std::fabs(e2[l]) <= std::fabs(c / h)) break;
                e2[l] = h * e2[l];
                if (e2[l] == 0.0) break;
            }
        }
        double p = d[l] + f;
        int i = l;
        while (i > 0 && !(std::fabs(p) >= std::fabs(d[i - 1]))) {
            d[i] = d[i - 1];
            --i;
        }
        d[i] = p;
    }
    return 0;
}

struct Builder {
    int n = 0, kx = 0;
    // lev[lv][b * (2^(lv+1)-1) + i]: roots of block b of level lv (rows
    // 1 + b 2^lv .. b 2^lv + 2^(lv+1) - 1), negated eigenvalues of the z
    // sub-tridiagonal, by increasing magnitude; level 0 = -BN(row)
    std::vector<std::vector<double>> lev;
    std::vector<CrOp> ops;
    std::vector<CrProg> progs;

    const double *blk(int lv, int b) const { return lev[lv].data() + (size_t)b * ((2 << lv) - 1); }

    // Program of a product with divisors and multiplier roots (na = 0), by
    // partial fractions: prod_j (T - c_j) / prod_i (T - b_i) = sum_i alpha_i
    // (T - b_i)^-1 with alpha_i = prod_j (b_i - c_j) / prod_{k != i} (b_i - b_k).
    // The multipliers are the eigenvalues of the two halves of the divisor's
    // block, which interlace with the divisor roots (Cauchy), so every alpha_i
    // is >= 0 and the sum does not cancel: its rounding is that of one more
    // elimination order (~1e-14 from BLKTRI's product order, which applies the
    // same roots one after another).  The solves are independent.  Each alpha
    // is formed in long double with every numerator factor paired with a
    // denominator factor, so no partial product overflows.
    int prog_mult(int nd, const double *div, int n1, const double *m1r, int n2, const double *m2r, int sgn)
    {
        CrProg p{};
        p.off = (int)ops.size();
        p.sgn = (nd > 0 && sgn <= 0) ? -1.0 : 1.0;
        p.has_pre = 0;
        std::vector<long double> c;
        for (int j = 0; j < n1; ++j) c.push_back(m1r[j]);
        for (int j = 0; j < n2; ++j) c.push_back(m2r[j]);
        for (int i = nd - 1; i >= 0; --i) {
            long double al = 1.0L;
            int kk = 0;
            for (size_t j = 0; j < c.size(); ++j) {
                while (kk == i) ++kk;
                al *= ((long double)div[i] - c[j]) / ((long double)div[i] - (long double)div[kk]);
                ++kk;
            }
            for (; kk < nd; ++kk)
                if (kk != i) al /= ((long double)div[i] - (long double)div[kk]);
            ops.push_back(CrOp{div[i], (double)al, CR_PF, 0});
        }
        p.nsolve = (int)ops.size() - p.off;
        progs.push_back(p);
        return (int)progs.size() - 1;
    }

    // program of a product of scalar factors aa[na-1..0] and nd = na-1 divisors
    int prog_scale(int nd, const double *div, int na, const double *aa, int sgn)
    {
        CrProg p{};
        p.off = (int)ops.size();
        p.sgn = (nd > 0 && sgn <= 0) ? -1.0 : 1.0;
        p.has_pre = 1;
        p.pre = aa[na - 1];
        int id = nd, ia = na - 1;
        while (id > 0) {
            CrOp op{div[--id], aa[--ia], CR_SCALE, 0};
            ops.push_back(op);
        }
        p.nsolve = (int)ops.size() - p.off;
        progs.push_back(p);
        return (int)progs.size() - 1;
    }
};

}  // namespace

int build_cr_plan(CrPlan &plan, int n, const double *an, const double *bn, const double *cn)
{
    Builder B;
    int kx = 0;
    while ((1 << kx) - 1 < n) ++kx;
    if ((1 << kx) - 1 != n || kx < 2) return CHY_EUNSUPPORTED;
    B.n = n;
    B.kx = kx;
    // squared symmetrised couplings, formed as the reference forms them:
    // the square of sqrt(AN(i) CN(i-1)) (pot/blktri.f:283-293, :360-362)
    std::vector<double> cpl2(n, 0.0);
    for (int i = 1; i < n; ++i) {
        double arg = an[i] * cn[i - 1];
        if (arg < 0.0) return CHY_ENUMERIC;
        double e = std::sqrt(arg);
        cpl2[i] = e * e;
    }
    const double machep = cr_machep();
    B.lev.resize(kx);
    B.lev[0].resize(n);
    for (int i = 0; i < n; ++i) B.lev[0][i] = -bn[i];
    for (int lv = 1; lv < kx; ++lv) {
        const int step = 1 << lv, size = 2 * step - 1, nblk = (1 << (kx - lv)) - 1;
        B.lev[lv].resize((size_t)size * nblk);
        std::vector<double> d(size), e2(size);
        for (int b = 0; b < nblk; ++b) {
            const int r0 = b * step;
            for (int i = 0; i < size; ++i) d[i] = bn[r0 + i];
            for (int i = 0; i + 1 < size; ++i) e2[i] = cpl2[r0 + i + 1];
            if (rational_ql(size, d.data(), e2.data(), machep) != 0) return CHY_ENUMERIC;
            for (int i = 0; i < size; ++i) B.lev[lv][(size_t)b * size + i] = -d[i];
        }
    }
    // CN shifted one row down, as BLKTRI hands it to its reduction
    // (pot/blktri.f:46-51): cnp[i] couples row i to row i-1's equation
    std::vector<double> cnp(n);
    cnp[0] = cn[n - 1];
    for (int i = 1; i < n; ++i) cnp[i] = cn[i - 1];

    plan.kx = kx;
    plan.phases.clear();
    plan.rows.clear();
    auto begin_phase = [&](int kind, int h) {
        CrPhase ph{};
        ph.kind = kind;
        ph.h = h;
        ph.row0 = (int)plan.rows.size();
        plan.phases.push_back(ph);
    };
    auto end_phase = [&]() { plan.phases.back().nrow = (int)plan.rows.size() - plan.phases.back().row0; };
    // reduction (pot/blktri.f:72-120)
    for (int lv = 0; lv < kx - 1; ++lv) {
        const int h = 1 << lv, iim = h - 1, iiz = 2 * h - 1, sg = (lv & 1) ? -1 : 1;
        begin_phase(CR_PHASE_ODD, h);
        for (int o = h; o <= n; o += 2 * h) {
            const int b = o / h - 1;
            CrRow r{o, -1, -1, -1};
            r.pm = B.prog_mult(iiz, B.blk(lv, b), iim, lv ? B.blk(lv - 1, 2 * b) : nullptr, iim,
                               lv ? B.blk(lv - 1, 2 * b + 2) : nullptr, sg);
            plan.rows.push_back(r);
        }
        end_phase();
        begin_phase(CR_PHASE_EVEN, h);
        for (int e = 2 * h; e + h <= n; e += 2 * h) {
            const int s = e / (2 * h);
            CrRow r{e, -1, -1, -1};
            r.pa = B.prog_scale(iim, lv ? B.blk(lv - 1, 4 * s - 2) : nullptr, h, an + (e - h), -sg);
            r.pc = B.prog_scale(iim, lv ? B.blk(lv - 1, 4 * s) : nullptr, h, cnp.data() + e, -sg);
            plan.rows.push_back(r);
        }
        end_phase();
    }
    // back substitution (pot/blktri.f:122-170)
    for (int lv = kx - 1; lv >= 0; --lv) {
        const int h = 1 << lv, iim = h - 1, iiz = 2 * h - 1, sg = (lv & 1) ? -1 : 1;
        const int nrow = (1 << (kx - lv)) - 1;
        begin_phase(CR_PHASE_BACK, h);
        for (int jj = 1; jj <= nrow; jj += 2) {
            const int i = jj * h;
            CrRow r{i, -1, -1, -1};
            if (jj > 1)
                r.pa = B.prog_scale(iim, lv ? B.blk(lv - 1, 2 * jj - 2) : nullptr, h, an + (i - h), -sg);
            if (jj < nrow) r.pc = B.prog_scale(iim, lv ? B.blk(lv - 1, 2 * jj) : nullptr, h, cnp.data() + i, -sg);
            r.pm = B.prog_mult(iiz, B.blk(lv, jj - 1), iim, lv ? B.blk(lv - 1, 2 * jj - 2) : nullptr, iim,
                               lv ? B.blk(lv - 1, 2 * jj) : nullptr, sg);
            plan.rows.push_back(r);
        }
        end_phase();
    }
    plan.ops = std::move(B.ops);
    plan.progs = std::move(B.progs);
    plan.nsolve = (long)plan.ops.size();
    return CHY_OK;
}

}  // namespace chy
