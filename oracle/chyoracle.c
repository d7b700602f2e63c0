/*
 * chyoracle.c -- CPU restatement of CHYMERA's self-gravity potential solve:
 * SETBDY + BDYGEN (pot/boundary.f) and POT3 + ZAXPHI (pot/pot3.f), with the
 * azimuthal real FFT of pot/fft.f and the block-tridiagonal mode solves of
 * pot/blktri.f.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py call it through ctypes.  It is
 * never linked into, or called by, the product library.
 *
 * What it follows (file:line in the reference):
 *   - layout PHI(JMAX+2,KMAX+2,LMAX), J fastest      headers/globals.h:55-58
 *   - forward FFT + REALTR packing, COEF diagnostics  pot/pot3.f:90-149
 *       (the FFT itself is an exact mixed-radix DFT; sign e^{+i theta} as
 *        FFT(ISN=+1) at pot/fft.f:217-218, REALTR(+1) scaling pot/fft.f:4-14)
 *   - mode coefficients AM,BM,CM,AN,BN,CN,Y         pot/pot3.f:158-231
 *   - per-mode solve of AN X(.,k-1)+(BN(k)+T_m)X(.,k)+CN X(.,k+1)=Y, the system
 *       BLKTRI solves (pot/blktri.f:3-56), two ways:
 *       (a) an exact direct method: diagonalise the m-independent z operator
 *           (symmetrised, implicit-shift QL), then one r-tridiagonal (Thomas)
 *           solve per z eigenvalue;
 *       (b) BLKTRI's own generalised cyclic reduction (pot/blktri.f:60-260)
 *           with its root tables (COMPB/TQLRAT, :264-423), in its order of
 *           operations: bitwise the reference's BLKTRI.
 *       Both pinned against BLKTRI golden vectors.
 *   - inverse FFT + ghosts                          pot/pot3.f:256-298
 *   - ZAXPHI Aitken extrapolation to r=0           pot/pot3.f:311-399
 *   - SETBDY boundary list and cos/sin tables       pot/boundary.f:28-79
 *   - BDYGEN ring moments, inside/outside split,
 *       boundary evaluation and ISYM=2 ghosts       pot/boundary.f:325-894
 * Only |ISYM|=2 (equatorial symmetry, the only geometry POT3's Neumann
 * midplane supports) is implemented.
 *
 * Working precision.  The default build is IEEE binary64 (the checker of the
 * FP64 product path).  Building with -DORA_EXTENDED runs the same algorithm in
 * x87 80-bit arithmetic on the same FP64 inputs and the same FP64 operator
 * coefficients (they are always formed in binary64, as the reference forms
 * them), which measures how far any FP64 solver -- the reference's BLKTRI
 * included -- is from the exact solution of the discrete system.
 */
#include <float.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

#ifdef ORA_EXTENDED
typedef long double real;
typedef long double _Complex cplx;
#define REAL_EPS LDBL_EPSILON
#define PI_R 3.14159265358979323846264338327950288L
#else
typedef double real;
typedef double _Complex cplx;
#define REAL_EPS DBL_EPSILON
#define PI_R 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* grid                                                                      */

typedef struct {
    int jmax, kmax, lmax;
    const double *r, *z, *rhf, *zhf; /* lengths jmax+2, kmax+2; Fortran index i is [i-1] */
    double dtheta, pi, grav;
} ora_grid;

#define JS(g) ((size_t)(g)->jmax + 2)
#define KS(g) ((size_t)(g)->kmax + 2)
/* Fortran PHI(J,K,L) with 1-based indices */
#define P3(g, j, k, l) (((size_t)(j) - 1) + JS(g) * (((size_t)(k) - 1) + KS(g) * ((size_t)(l) - 1)))

/* ------------------------------------------------------------------------ */
/* exact DFT (recursive mixed radix, smallest prime factor first)            */

static int smallest_factor(int n)
{
    for (int p = 2; p * p <= n; ++p)
        if (n % p == 0) return p;
    return n;
}

/* out[k] = sum_{j<n} in[j*s] * tw[(j*k*ts) mod N0], tw[e] = exp(sign*2*pi*i*e/N0) */
static void dft_rec(int n, const cplx *in, int s, cplx *out, const cplx *tw, int ts, int N0,
                    cplx *scratch)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n), q = n / p;
    for (int r = 0; r < p; ++r)
        dft_rec(q, in + (size_t)r * s, s * p, out + (size_t)r * q, tw, ts * p, N0, scratch);
    for (int k = 0; k < q; ++k)
        for (int t = 0; t < p; ++t) {
            cplx acc = 0;
            for (int r = 0; r < p; ++r) {
                long e = ((long)r * (k + (long)q * t)) % n;
                acc += tw[(e * ts) % N0] * out[(size_t)r * q + k];
            }
            scratch[k + (size_t)q * t] = acc;
        }
    memcpy(out, scratch, sizeof(cplx) * (size_t)n);
}

typedef struct { int n; cplx *twp, *twm, *buf, *out, *scr; } dft_plan;

static int plan_init(dft_plan *pl, int n)
{
    pl->n = n;
    pl->twp = malloc(sizeof(cplx) * n);
    pl->twm = malloc(sizeof(cplx) * n);
    pl->buf = malloc(sizeof(cplx) * n);
    pl->out = malloc(sizeof(cplx) * n);
    pl->scr = malloc(sizeof(cplx) * n);
    if (!pl->twp || !pl->twm || !pl->buf || !pl->out || !pl->scr) return -1;
    for (int e = 0; e < n; ++e) {
        real ang = 2 * PI_R * (real)e / (real)n;
        pl->twp[e] = cos(ang) + I * sin(ang);
        pl->twm[e] = cos(ang) - I * sin(ang);
    }
    return 0;
}

static void plan_free(dft_plan *pl)
{
    free(pl->twp); free(pl->twm); free(pl->buf); free(pl->out); free(pl->scr);
}

/* Raw forward real transform of L=2N samples, as FFT(+1) then REALTR(+1):
 * A[m] = 2 sum_n x_n cos(2 pi m n/L), B[m] = 2 sum_n x_n sin(2 pi m n/L),
 * m = 0..N (pot/fft.f:4-14). */
static void rfft_fwd(dft_plan *pl, const real *x, long xs, real *A, real *B)
{
    int L = pl->n, N = L / 2;
    for (int n = 0; n < L; ++n) pl->buf[n] = x[(long)n * xs];
    dft_rec(L, pl->buf, 1, pl->out, pl->twp, 1, L, pl->scr);
    for (int m = 0; m <= N; ++m) {
        A[m] = 2 * creal(pl->out[m]);
        B[m] = 2 * cimag(pl->out[m]);
    }
}

/* Raw inverse, as REALTR(-1) then FFT(-1) with B[0]=B[N]=0 (pot/pot3.f:260-271):
 * out_n = A_0 + 2 sum_{m=1}^{N-1} (A_m cos + B_m sin)(2 pi m n/L) + A_N cos(pi n). */
static void rfft_inv(dft_plan *pl, const real *A, const real *B, real *out)
{
    int L = pl->n, N = L / 2;
    for (int m = 0; m < L; ++m) pl->buf[m] = 0;
    pl->buf[0] = A[0];
    pl->buf[N] += A[N];
    for (int m = 1; m < N; ++m) {
        pl->buf[m] = A[m] + I * B[m];
        pl->buf[L - m] = A[m] - I * B[m];
    }
    dft_rec(L, pl->buf, 1, pl->out, pl->twm, 1, L, pl->scr);
    for (int n = 0; n < L; ++n) out[n] = creal(pl->out[n]);
}

int ora_rfft_fwd(int L, const double *x, double *A, double *B)
{
    dft_plan pl;
    if (L < 2 || (L & 1)) return 1;
    if (plan_init(&pl, L)) return 3;
    real *xr = malloc(sizeof(real) * L), *Ar = malloc(sizeof(real) * (L / 2 + 1));
    real *Br = malloc(sizeof(real) * (L / 2 + 1));
    for (int i = 0; i < L; ++i) xr[i] = x[i];
    rfft_fwd(&pl, xr, 1, Ar, Br);
    for (int i = 0; i <= L / 2; ++i) { A[i] = (double)Ar[i]; B[i] = (double)Br[i]; }
    free(xr); free(Ar); free(Br);
    plan_free(&pl);
    return 0;
}

int ora_rfft_inv(int L, const double *A, const double *B, double *out)
{
    dft_plan pl;
    if (L < 2 || (L & 1)) return 1;
    if (plan_init(&pl, L)) return 3;
    real *Ar = malloc(sizeof(real) * (L / 2 + 1)), *Br = malloc(sizeof(real) * (L / 2 + 1));
    real *o = malloc(sizeof(real) * L);
    for (int i = 0; i <= L / 2; ++i) { Ar[i] = A[i]; Br[i] = B[i]; }
    rfft_inv(&pl, Ar, Br, o);
    for (int i = 0; i < L; ++i) out[i] = (double)o[i];
    free(Ar); free(Br); free(o);
    plan_free(&pl);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* symmetric tridiagonal eigensystem: implicit-shift QL with eigenvectors.   */
/* d[0..n-1] diagonal (out: eigenvalues), e[1..n-1] off-diagonal (destroyed),*/
/* z n*n row-major, z[k*n+p] = component k of eigenvector p.                 */

/* z may be NULL (eigenvalues only).  Instantiated for the working precision
 * and for x87 80-bit (the exact-root variant of the cyclic reduction below). */
#define DEFINE_TRIDIAG_QL(NAME, T, EPS)                                                      \
    static int NAME(int n, T *d, T *e, T *z)                                                 \
    {                                                                                        \
        if (z) {                                                                             \
            for (int i = 0; i < n * n; ++i) z[i] = 0;                                        \
            for (int i = 0; i < n; ++i) z[i * n + i] = 1;                                    \
        }                                                                                    \
        for (int i = 1; i < n; ++i) e[i - 1] = e[i];                                         \
        if (n > 0) e[n - 1] = 0;                                                             \
        for (int l = 0; l < n; ++l) {                                                        \
            int iter = 0, m;                                                                 \
            for (;;) {                                                                       \
                for (m = l; m < n - 1; ++m) {                                                \
                    T dd = fabs(d[m]) + fabs(d[m + 1]);                                      \
                    if (fabs(e[m]) <= (EPS) * dd) break;                                     \
                }                                                                            \
                if (m == l) break;                                                           \
                if (++iter > 60) return 1;                                                   \
                T g = (d[l + 1] - d[l]) / (2 * e[l]);                                        \
                T r = hypot(g, (T)1);                                                        \
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));                \
                T s = 1, c = 1, p = 0;                                                       \
                int i;                                                                       \
                for (i = m - 1; i >= l; --i) {                                               \
                    T f = s * e[i], b = c * e[i];                                            \
                    r = hypot(f, g);                                                         \
                    e[i + 1] = r;                                                            \
                    if (r == 0) { d[i + 1] -= p; e[m] = 0; break; }                          \
                    s = f / r; c = g / r;                                                    \
                    g = d[i + 1] - p;                                                        \
                    r = (d[i] - g) * s + 2 * c * b;                                          \
                    p = s * r;                                                               \
                    d[i + 1] = g + p;                                                        \
                    g = c * r - b;                                                           \
                    if (z)                                                                   \
                        for (int k = 0; k < n; ++k) {                                        \
                            T zf = z[k * n + i + 1];                                         \
                            z[k * n + i + 1] = s * z[k * n + i] + c * zf;                    \
                            z[k * n + i] = c * z[k * n + i] - s * zf;                        \
                        }                                                                    \
                }                                                                            \
                if (r == 0 && i >= l) continue;                                              \
                d[l] -= p; e[l] = g; e[m] = 0;                                               \
            }                                                                                \
        }                                                                                    \
        return 0;                                                                            \
    }

DEFINE_TRIDIAG_QL(tridiag_ql, real, REAL_EPS)
DEFINE_TRIDIAG_QL(tridiag_ql_ld, long double, LDBL_EPSILON)

/* ------------------------------------------------------------------------ */
/* per-mode solve of the BLKTRI system (pot/blktri.f:3-56 contract):
 *   AN(k) X(i,k-1) + AM(i) X(i-1,k) + (BN(k)+BM(i)) X(i,k) + CM(i) X(i+1,k)
 *   + CN(k) X(i,k+1) = Y(i,k),  i=1..M, k=1..N,  AN(1)=CN(N)=AM(1)=CM(M)=0.   */

typedef struct {
    int n;       /* z unknowns */
    real *lam;   /* eigenvalues of the z operator */
    real *fwd;   /* fwd[k*n+p] = d_k Q[k][p]   (Y -> eigen coordinates) */
    real *inv;   /* inv[k*n+p] = Q[k][p] / d_k (eigen coordinates -> X) */
} zbasis;

static void zbasis_free(zbasis *zb) { free(zb->lam); free(zb->fwd); free(zb->inv); }

static int zbasis_init(zbasis *zb, int n, const real *an, const real *bn, const real *cn)
{
    zb->n = n;
    zb->lam = malloc(sizeof(real) * n);
    zb->fwd = malloc(sizeof(real) * n * n);
    zb->inv = malloc(sizeof(real) * n * n);
    real *e = malloc(sizeof(real) * (n + 1)), *dsc = malloc(sizeof(real) * n);
    real *q = malloc(sizeof(real) * n * n);
    if (!zb->lam || !zb->fwd || !zb->inv || !e || !dsc || !q) { free(e); free(dsc); free(q); return 3; }
    /* D T D^-1 symmetric: d_k = d_{k-1} sqrt(CN_{k-1}/AN_k), e_k = sqrt(AN_k CN_{k-1}) */
    dsc[0] = 1;
    for (int k = 1; k < n; ++k) {
        if (!(an[k] * cn[k - 1] > 0)) { free(e); free(dsc); free(q); return 4; }
        dsc[k] = dsc[k - 1] * sqrt(cn[k - 1] / an[k]);
        e[k] = sqrt(an[k] * cn[k - 1]);
    }
    for (int k = 0; k < n; ++k) zb->lam[k] = bn[k];
    int rc = tridiag_ql(n, zb->lam, e, q);
    for (int k = 0; k < n; ++k)
        for (int p = 0; p < n; ++p) {
            zb->fwd[k * n + p] = dsc[k] * q[k * n + p];
            zb->inv[k * n + p] = q[k * n + p] / dsc[k];
        }
    free(q); free(e); free(dsc);
    return rc ? 4 : 0;
}

/* one r-tridiagonal (Thomas) solve: AM(i) x(i-1) + (BM(i)+lam) x(i) + CM(i) x(i+1) = rhs */
static void thomas(int m, const real *am, const real *bm, const real *cm, real lam, real *x,
                   real *cp)
{
    real den = bm[0] + lam;
    cp[0] = cm[0] / den;
    x[0] = x[0] / den;
    for (int i = 1; i < m; ++i) {
        den = bm[i] + lam - am[i] * cp[i - 1];
        cp[i] = cm[i] / den;
        x[i] = (x[i] - am[i] * x[i - 1]) / den;
    }
    for (int i = m - 2; i >= 0; --i) x[i] -= cp[i] * x[i + 1];
}

/* y is Y(M,N) column-major (i fastest), overwritten by X */
static void mode_solve(const zbasis *zb, int m, const real *am, const real *bm, const real *cm,
                       real *y, real *work)
{
    int n = zb->n;
    real *yh = work; /* m*n */
    real *cp = work + (size_t)m * n;
    for (int p = 0; p < n; ++p) {
        real *col = yh + (size_t)p * m;
        for (int i = 0; i < m; ++i) col[i] = 0;
        for (int k = 0; k < n; ++k) {
            real w = zb->fwd[k * n + p];
            const real *yk = y + (size_t)k * m;
            for (int i = 0; i < m; ++i) col[i] += yk[i] * w;
        }
        thomas(m, am, bm, cm, zb->lam[p], col, cp);
    }
    for (int k = 0; k < n; ++k) {
        real *yk = y + (size_t)k * m;
        for (int i = 0; i < m; ++i) yk[i] = 0;
        for (int p = 0; p < n; ++p) {
            real w = zb->inv[k * n + p];
            const real *col = yh + (size_t)p * m;
            for (int i = 0; i < m; ++i) yk[i] += col[i] * w;
        }
    }
}

static real *to_real(const double *a, size_t n)
{
    real *o = malloc(sizeof(real) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) o[i] = a[i];
    return o;
}

int ora_mode_solve(int m, int n, const double *an, const double *bn, const double *cn,
                   const double *am, const double *bm, const double *cm, double *y)
{
    zbasis zb;
    if (m < 2 || n < 1) return 1;
    real *ran = to_real(an, n), *rbn = to_real(bn, n), *rcn = to_real(cn, n);
    real *ram = to_real(am, m), *rbm = to_real(bm, m), *rcm = to_real(cm, m);
    real *ry = to_real(y, (size_t)m * n);
    int rc = zbasis_init(&zb, n, ran, rbn, rcn);
    if (!rc) {
        real *work = malloc(sizeof(real) * ((size_t)m * n + m));
        mode_solve(&zb, m, ram, rbm, rcm, ry, work);
        for (size_t i = 0; i < (size_t)m * n; ++i) y[i] = (double)ry[i];
        free(work);
    }
    zbasis_free(&zb);
    free(ran); free(rbn); free(rcn); free(ram); free(rbm); free(rcm); free(ry);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Generalised cyclic reduction of the same per-mode system, in the order    */
/* the reference's BLKTRI applies it (pot/blktri.f:60-260 BLKTR1 and PRDCT,  */
/* root tables of :264-344 COMPB and :348-423 TQLRAT).  N = 2^kx - 1 z rows,  */
/* AN(1) = CN(N) = 0 (BLKTRI's NP=1 case, the only one POT3 uses).            */
/*                                                                            */
/* Row i (1-based) of level lv (lv = 0..kx-1) eliminates the rows between    */
/* multiples of 2^lv.  Block b of level lv is the z sub-tridiagonal on rows   */
/* 1 + b 2^lv .. b 2^lv + 2^(lv+1) - 1 (diagonal BN, coupling AN(i) CN(i-1)); */
/* its roots are the negated eigenvalues, ordered by increasing magnitude.   */
/* Level 0 blocks are single rows with root -BN(i).                          */

typedef struct {
    int n, kx;
    double **lev;   /* lev[lv] + b * (2^(lv+1) - 1): roots of block b of level lv */
} cr_roots;

/* How the roots and the r solves are formed (the variants measure which part
 * of BLKTRI's rounding its result carries):
 *   roots: 0 = rational QL with the reference's tolerance (100 x the first
 *              power of ten below binary64 resolution), as TQLRAT;
 *          1 = eigenvalues of the same sub-tridiagonals by implicit QL in
 *              x87 80-bit arithmetic, rounded once;
 *   rsolve: 0 = elimination from the last r row up, substitution down (the
 *               order of pot/blktri.f:208-231);
 *           1 = elimination from the first row down, substitution up.      */
static int cr_roots_kind = 0, cr_rsolve_kind = 0;

void ora_cr_variant(int roots, int rsolve) { cr_roots_kind = roots; cr_rsolve_kind = rsolve; }

/* Mode solver of ora_pot3: 0 = z eigenbasis (default; in the 80-bit build the
 * exact reference point), 1 = the cyclic reduction below (binary64 in both
 * builds: BLKTRI's own arithmetic). */
static int ora_modes = 0;

void ora_mode_solver(int kind) { ora_modes = kind; }

/* the reference's convergence scale: 100 x the first 10^-k with 1 + 10^-k == 1 */
static double cr_machep(void)
{
    volatile double e = 1, one_plus;
    do {
        e = e / 10.;
        one_plus = 1. + e;
    } while (one_plus > 1.);
    return 100. * e;
}

/* Eigenvalues of the symmetric tridiagonal with diagonal d[0..n-1] and squared
 * off-diagonals e2[0..n-2] (e2[i] couples rows i and i+1) by the rational QL
 * iteration without square roots of the off-diagonals (Reinsch's algorithm, as
 * EISPACK TQLRAT; pot/blktri.f:348-423).  On return d holds the eigenvalues in
 * order of increasing magnitude.  e2 is destroyed (length n). */
static int rational_ql(int n, double *d, double *e2, double machep)
{
    if (n == 1) return 0;
    e2[n - 1] = 0;
    double f = 0, b = 0, c = 0;
    for (int l = 0; l < n; ++l) {
        double h = machep * (fabs(d[l]) + sqrt(e2[l]));
        if (!(b > h)) { b = h; c = b * b; }
        int m = l;
        while (!(e2[m] <= c)) ++m;
        if (m != l) {
            for (int it = 0;; ++it) {
                if (it == 30) return l + 1;
                double s = sqrt(e2[l]), g = d[l];
                double p = (d[l + 1] - g) / (2.0 * s);
                double r = sqrt(p * p + 1.0);
                d[l] = s / (p + copysign(r, p));
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
                if (fabs(e2[l]) <= fabs(c / h)) break;
                e2[l] = h * e2[l];
                if (e2[l] == 0.0) break;
            }
        }
        double p = d[l] + f;
        int i = l;
        while (i > 0 && !(fabs(p) >= fabs(d[i - 1]))) { d[i] = d[i - 1]; --i; }
        d[i] = p;
    }
    return 0;
}

static void cr_roots_free(cr_roots *t)
{
    if (t->lev)
        for (int lv = 0; lv < t->kx; ++lv) free(t->lev[lv]);
    free(t->lev);
    t->lev = NULL;
}

/* an, bn, cn: the z coefficients (0-based, an[0] = cn[n-1] = 0) */
static int cr_roots_init(cr_roots *t, int n, const double *an, const double *bn, const double *cn)
{
    int kx = 0;
    while ((1 << kx) - 1 < n) ++kx;
    if ((1 << kx) - 1 != n || kx < 2) return 2;
    t->n = n; t->kx = kx;
    t->lev = calloc(kx, sizeof(double *));
    /* squared symmetrised couplings, as COMPB forms them: sqrt(AN(i) CN(i-1)) squared */
    double *cpl2 = malloc(sizeof(double) * n);
    for (int i = 1; i < n; ++i) {
        double arg = an[i] * cn[i - 1];
        if (arg < 0) { free(cpl2); cr_roots_free(t); return 4; }
        double e = sqrt(arg);
        cpl2[i] = e * e;
    }
    double machep = cr_machep();
    t->lev[0] = malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) t->lev[0][i] = -bn[i];
    for (int lv = 1; lv < kx; ++lv) {
        int step = 1 << lv, size = 2 * step - 1, nblk = (1 << (kx - lv)) - 1;
        t->lev[lv] = malloc(sizeof(double) * size * nblk);
        double *d = malloc(sizeof(double) * size), *e2 = malloc(sizeof(double) * size);
        for (int b = 0; b < nblk; ++b) {
            int r0 = b * step; /* 0-based first row */
            for (int i = 0; i < size; ++i) d[i] = bn[r0 + i];
            for (int i = 0; i + 1 < size; ++i) e2[i] = cpl2[r0 + i + 1];
            int rc = 0;
            if (cr_roots_kind == 0) {
                rc = rational_ql(size, d, e2, machep);
            } else {
                long double *dd = malloc(sizeof(long double) * size);
                long double *ee = malloc(sizeof(long double) * (size + 1));
                long double *zz = malloc(sizeof(long double) * size * size);
                for (int i = 0; i < size; ++i) dd[i] = bn[r0 + i];
                for (int i = 1; i < size; ++i) ee[i] = sqrtl((long double)an[r0 + i] * cn[r0 + i - 1]);
                /* eigenvalues only are needed; the 80-bit QL of tridiag_ql is reused */
                rc = tridiag_ql_ld(size, dd, ee, zz);
                for (int i = 0; i < size; ++i) d[i] = (double)dd[i];
                free(dd); free(ee); free(zz);
                /* order of increasing magnitude, as the reference applies them */
                for (int i = 1; i < size; ++i) {
                    double v = d[i];
                    int k = i;
                    while (k > 0 && fabs(v) < fabs(d[k - 1])) { d[k] = d[k - 1]; --k; }
                    d[k] = v;
                }
            }
            if (rc) { free(d); free(e2); cr_roots_free(t); return 4; }
            for (int i = 0; i < size; ++i) t->lev[lv][b * size + i] = -d[i];
        }
        free(d); free(e2);
    }
    free(cpl2);
    return 0;
}

static const double *cr_blk(const cr_roots *t, int lv, int b)
{
    return t->lev[lv] + (size_t)b * ((2 << lv) - 1);
}

/* (T - rt I) w = y along r, T = tridiag(am, bm, cm) of order m */
static void cr_rsolve(int m, const double *am, const double *bm, const double *cm, double rt,
                      const double *y, double *w, double *d)
{
    if (cr_rsolve_kind == 0) {
        d[m - 1] = am[m - 1] / (bm[m - 1] - rt);
        w[m - 1] = y[m - 1] / (bm[m - 1] - rt);
        for (int i = m - 2; i >= 1; --i) {
            double den = bm[i] - rt - cm[i] * d[i + 1];
            d[i] = am[i] / den;
            w[i] = (y[i] - cm[i] * w[i + 1]) / den;
        }
        double den = bm[0] - rt - cm[0] * d[1];
        w[0] = (den == 0) ? 1. : (y[0] - cm[0] * w[1]) / den;
        for (int i = 1; i < m; ++i) w[i] = w[i] - d[i] * w[i - 1];
    } else {
        d[0] = cm[0] / (bm[0] - rt);
        w[0] = y[0] / (bm[0] - rt);
        for (int i = 1; i < m - 1; ++i) {
            double den = bm[i] - rt - am[i] * d[i - 1];
            d[i] = cm[i] / den;
            w[i] = (y[i] - am[i] * w[i - 1]) / den;
        }
        double den = bm[m - 1] - rt - am[m - 1] * d[m - 2];
        w[m - 1] = (y[m - 1] - am[m - 1] * w[m - 2]) / den;
        for (int i = m - 2; i >= 0; --i) w[i] = w[i] - d[i] * w[i + 1];
    }
}

/* y = sgn [prod (T - mult)] [prod (T - div)^-1] x with scalar factors aa
 * interleaved (pot/blktri.f:177-260).  Divisors are taken from the end of
 * div; after each division one multiplier root is folded in as
 * y + (rt - mult) w = (T - mult) w unless the next divisor is closer to it.
 * With na > 0 (no multipliers) each division is preceded by a factor
 * aa[ia], ia = na-1 down to 0.  work: 2m doubles. */
static void cr_product(int m, const double *am, const double *bm, const double *cm,
                       int nd, const double *div, int n1, const double *mult1, int n2,
                       const double *mult2, int na, const double *aa, const double *x,
                       double *y, double *w, double *work, int sgn)
{
    double *dw = work;
    if (nd > 0 && sgn <= 0)
        for (int i = 0; i < m; ++i) { w[i] = -x[i]; y[i] = w[i]; }
    else
        for (int i = 0; i < m; ++i) { w[i] = x[i]; y[i] = x[i]; }
    int id = nd, ia = na, m1 = n1, m2 = n2, deferred = 0;
    for (;;) {
        if (ia > 0) {
            double f = aa[--ia];
            for (int i = 0; i < m; ++i) y[i] = f * w[i];
        }
        if (id <= 0) return;
        double rt = div[--id];
        if (id == 0) deferred = 1;
        cr_rsolve(m, am, bm, cm, rt, y, w, dw);
        if (na > 0) continue;
        const double *mr = NULL;
        if (m1 > 0 && (m2 <= 0 || !(fabs(mult1[m1 - 1]) <= fabs(mult2[m2 - 1])))) {
            if (deferred || !(fabs(mult1[m1 - 1] - div[id - 1]) < fabs(mult1[m1 - 1] - rt)))
                mr = &mult1[--m1];
        } else if (m2 > 0) {
            if (deferred || !(fabs(mult2[m2 - 1] - div[id - 1]) < fabs(mult2[m2 - 1] - rt)))
                mr = &mult2[--m2];
        }
        if (!mr) {
            for (int i = 0; i < m; ++i) y[i] = w[i];
            deferred = 1;
            continue;
        }
        rt = rt - *mr;
        for (int i = 0; i < m; ++i) y[i] = y[i] + rt * w[i];
    }
}

/* The product of a multiplier program (na = 0) by partial fractions: with
 * divisor roots b_i and multiplier roots c_j (one fewer), prod(T - c_j) /
 * prod(T - b_i) = sum_i alpha_i (T - b_i)^-1, alpha_i = prod_j (b_i - c_j) /
 * prod_{k != i} (b_i - b_k).  BLKTRI's multipliers are the eigenvalues of the
 * two halves of the divisor's block, which interlace with the divisor roots
 * (Cauchy), so every alpha_i >= 0 and the sum has no cancellation.  The
 * solves are independent (the form a parallel machine takes).  Variant 2 of
 * cr_rsolve_kind's companion switch cr_pf. */
static int cr_pf = 0;

void ora_cr_partial_fractions(int on) { cr_pf = on; }

static void cr_product_pf(int m, const double *am, const double *bm, const double *cm, int nd,
                          const double *div, int n1, const double *mult1, int n2, const double *mult2,
                          const double *x, double *y, double *w, double *work, int sgn)
{
    long double *c = malloc(sizeof(long double) * (n1 + n2 + 1));
    int nc = 0;
    for (int j = 0; j < n1; ++j) c[nc++] = mult1[j];
    for (int j = 0; j < n2; ++j) c[nc++] = mult2[j];
    double *xs = malloc(sizeof(double) * m);
    for (int r = 0; r < m; ++r) { xs[r] = (nd > 0 && sgn <= 0) ? -x[r] : x[r]; y[r] = 0; }
    for (int i = nd - 1; i >= 0; --i) {
        /* alpha_i, each numerator factor paired with a denominator factor */
        long double al = 1.0L;
        int kk = 0;
        for (int j = 0; j < nc; ++j) {
            while (kk == i) ++kk;
            al *= ((long double)div[i] - c[j]) / ((long double)div[i] - div[kk]);
            ++kk;
        }
        for (; kk < nd; ++kk)
            if (kk != i) al /= ((long double)div[i] - div[kk]);
        cr_rsolve(m, am, bm, cm, div[i], xs, w, work);
        double a = (double)al;
        for (int r = 0; r < m; ++r) y[r] += a * w[r];
    }
    free(c); free(xs);
}

/* y is Y(M,N) column-major (r fastest), overwritten by X.  cnp is CN shifted
 * one row down (cnp[i] = cn[i-1], cnp[0] = cn[n-1] = 0), the form BLKTRI
 * hands its reduction (pot/blktri.f:46-51). */
static void cr_solve(const cr_roots *t, int m, const double *an, const double *cnp,
                     const double *am, const double *bm, const double *cm, double *y)
{
    int n = t->n, kx = t->kx;
    double *p = malloc(sizeof(double) * (size_t)m * (n + 1));
    double *w1 = malloc(sizeof(double) * m), *w2 = malloc(sizeof(double) * m);
    double *ww = malloc(sizeof(double) * m), *wk = malloc(sizeof(double) * 2 * m);
#define YC(i) (y + (size_t)((i) - 1) * m)
#define PC(i) (p + (size_t)(i) * m)
    /* reduction (pot/blktri.f:72-120) */
    for (int lv = 0; lv < kx - 1; ++lv) {
        int h = 1 << lv, iim = h - 1, iiz = 2 * h - 1;
        int sg = (lv & 1) ? -1 : 1;
        for (int o = h; o <= n; o += 2 * h) {
            int b = o / h - 1;
            if (cr_pf)
                cr_product_pf(m, am, bm, cm, iiz, cr_blk(t, lv, b), iim, lv ? cr_blk(t, lv - 1, 2 * b) : NULL,
                              iim, lv ? cr_blk(t, lv - 1, 2 * b + 2) : NULL, YC(o), PC(o), ww, wk, sg);
            else
                cr_product(m, am, bm, cm, iiz, cr_blk(t, lv, b), iim, lv ? cr_blk(t, lv - 1, 2 * b) : NULL,
                           iim, lv ? cr_blk(t, lv - 1, 2 * b + 2) : NULL, 0, NULL, YC(o), PC(o), ww, wk, sg);
        }
        for (int e = 2 * h; e + h <= n; e += 2 * h) {
            int s = e / (2 * h);
            cr_product(m, am, bm, cm, iim, lv ? cr_blk(t, lv - 1, 4 * s - 2) : NULL, 0, NULL, 0, NULL,
                       h, an + (e - h), PC(e - h), w1, ww, wk, -sg);
            cr_product(m, am, bm, cm, iim, lv ? cr_blk(t, lv - 1, 4 * s) : NULL, 0, NULL, 0, NULL,
                       h, cnp + e, PC(e + h), w2, ww, wk, -sg);
            double *ye = YC(e);
            for (int i = 0; i < m; ++i) ye[i] = w1[i] + w2[i] - ye[i];
        }
    }
    /* back substitution (pot/blktri.f:122-170) */
    for (int lv = kx - 1; lv >= 0; --lv) {
        int h = 1 << lv, iim = h - 1, iiz = 2 * h - 1;
        int sg = (lv & 1) ? -1 : 1, nrow = (1 << (kx - lv)) - 1;
        for (int jj = 1; jj <= nrow; jj += 2) {
            int i = jj * h;
            if (jj > 1)
                cr_product(m, am, bm, cm, iim, lv ? cr_blk(t, lv - 1, 2 * jj - 2) : NULL, 0, NULL, 0,
                           NULL, h, an + (i - h), YC(i - h), w1, ww, wk, -sg);
            else
                for (int r = 0; r < m; ++r) w1[r] = 0;
            if (jj < nrow)
                cr_product(m, am, bm, cm, iim, lv ? cr_blk(t, lv - 1, 2 * jj) : NULL, 0, NULL, 0, NULL,
                           h, cnp + i, YC(i + h), w2, ww, wk, -sg);
            else
                for (int r = 0; r < m; ++r) w2[r] = 0;
            double *yi = YC(i);
            for (int r = 0; r < m; ++r) w1[r] = yi[r] - w1[r] - w2[r];
            if (cr_pf)
                cr_product_pf(m, am, bm, cm, iiz, cr_blk(t, lv, jj - 1), iim,
                              lv ? cr_blk(t, lv - 1, 2 * jj - 2) : NULL, iim, lv ? cr_blk(t, lv - 1, 2 * jj) : NULL,
                              w1, yi, ww, wk, sg);
            else
                cr_product(m, am, bm, cm, iiz, cr_blk(t, lv, jj - 1), iim,
                           lv ? cr_blk(t, lv - 1, 2 * jj - 2) : NULL, iim, lv ? cr_blk(t, lv - 1, 2 * jj) : NULL,
                           0, NULL, w1, yi, ww, wk, sg);
        }
    }
#undef YC
#undef PC
    free(p); free(w1); free(w2); free(ww); free(wk);
}

/* BLKTRI's system solved by the cyclic reduction above (binary64 always) */
int ora_cr_solve(int m, int n, const double *an, const double *bn, const double *cn,
                 const double *am, const double *bm, const double *cm, double *y)
{
    cr_roots t = {0};
    if (m < 2 || n < 3) return 1;
    int rc = cr_roots_init(&t, n, an, bn, cn);
    if (rc) return rc;
    double *cnp = malloc(sizeof(double) * n);
    cnp[0] = cn[n - 1];
    for (int i = 1; i < n; ++i) cnp[i] = cn[i - 1];
    cr_solve(&t, m, an, cnp, am, bm, cm, y);
    free(cnp);
    cr_roots_free(&t);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* coefficient arrays of pot/pot3.f:158-194, always formed in binary64 as the */
/* reference forms them (0-based: index i <-> Fortran i+1)                   */

typedef struct {
    real *am, *cm, *c, *an, *bn, *cn;
    real cmmax, ammin, cnmax;
} pot3_coef;

static void pot3_coef_init(const ora_grid *g, pot3_coef *pc)
{
    int J = g->jmax, K = g->kmax, M = J - 1, N = K - 1;
    const double *rhf = g->rhf - 1, *zhf = g->zhf - 1; /* 1-based views */
    double *rd2 = calloc(J + 1, sizeof(double)), *rd3 = calloc(J + 1, sizeof(double));
    double *denr = calloc(J + 1, sizeof(double)), *zd2 = calloc(K + 1, sizeof(double));
    double *denz = calloc(K + 1, sizeof(double));
    double *am = malloc(sizeof(double) * M), *cm = malloc(sizeof(double) * M);
    double *c = malloc(sizeof(double) * M), *an = malloc(sizeof(double) * N);
    double *bn = malloc(sizeof(double) * N), *cn = malloc(sizeof(double) * N);
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
    pc->cmmax = cm[M - 1]; cm[M - 1] = 0.0;
    pc->ammin = am[0];     am[0] = 0.0;
    pc->cnmax = cn[N - 1]; cn[N - 1] = 0.0;
    bn[0] = -cn[0];        an[0] = 0.0;
    pc->am = to_real(am, M); pc->cm = to_real(cm, M); pc->c = to_real(c, M);
    pc->an = to_real(an, N); pc->bn = to_real(bn, N); pc->cn = to_real(cn, N);
    free(am); free(cm); free(c); free(an); free(bn); free(cn);
    free(rd2); free(rd3); free(denr); free(zd2); free(denz);
}

static void pot3_coef_free(pot3_coef *pc)
{
    free(pc->am); free(pc->cm); free(pc->c); free(pc->an); free(pc->bn); free(pc->cn);
}

/* ------------------------------------------------------------------------ */
/* ZAXPHI (pot/pot3.f:311-399): NPOINT-point Aitken extrapolation to r=0     */

static void zaxphi(const ora_grid *g, int npoint, real *phi, double *phichk, int *klocat)
{
    int lm = g->lmax / 2, jmax2 = g->jmax + 2;
    real x[11], p2[11];
    if (npoint > 10) npoint = 10;
    int ichk = npoint % 2, nn = npoint - ichk, num = nn / 2, ispecl = num + num + 1, jsp = num + 1;
    for (int j = 2; j <= jsp; ++j) {
        x[num + (j - 1)] = g->rhf[j - 1];
        x[num - (j - 2)] = -g->rhf[j - 1];
    }
    if (ichk == 1) x[ispecl] = g->rhf[num + 1];
    real *phac = malloc(sizeof(real) * (g->kmax + 2));
    real *ph = malloc(sizeof(real) * (lm > 0 ? lm : 1));
    for (int k = 2; k <= g->kmax + 1; ++k) {
        real phmax = 0;
        for (int l = 1; l <= lm; ++l) {
            int lp = l + lm;
            for (int j = 2; j <= jsp; ++j) {
                p2[num + (j - 1)] = phi[P3(g, j, k, l)];
                p2[num - (j - 2)] = phi[P3(g, j, k, lp)];
            }
            if (ichk == 1) p2[ispecl] = phi[P3(g, num + 2, k, l)];
            for (int i = 1; i <= npoint - 1; ++i) {
                real p1 = p2[i], xinv = 1 / x[i], xr = 0 * xinv;
                for (int j = i + 1; j <= npoint; ++j) {
                    real xratio = x[j] * xinv, pb = p2[j];
                    p2[j] = (p1 * (xratio - xr) + pb * (xr - 1)) / (xratio - 1);
                }
            }
            ph[l - 1] = p2[npoint];
            if (fabs(ph[l - 1]) > fabs(phmax)) phmax = ph[l - 1];
            phi[P3(g, jmax2, k, l)] = ph[l - 1];
            phi[P3(g, jmax2, k, lp)] = ph[l - 1];
        }
        real err = 0;
        for (int l = 1; l <= lm; ++l) {
            real er = 1 - ph[l - 1] / phmax;
            if (fabs(er) > fabs(err)) err = er;
        }
        phac[k] = err;
    }
    real err = 0;
    for (int k = 2; k <= g->kmax + 1; ++k) {
        real xx = fabs(phac[k]);
        if (xx > fabs(err)) { err = phac[k]; *klocat = k; }
    }
    *phichk = (double)err;
    free(phac); free(ph);
}

/* ------------------------------------------------------------------------ */
/* POT3 (pot/pot3.f:15-306)                                                  */
/* phi: in/out PHI(JMAX+2,KMAX+2,LMAX) holding BDYGEN's boundary values;
 * rho: RHO(...); coef: COEF(JMAX+2,KMAX+2,LMAX/2,2) in/out (may be NULL);
 * phichk/klocat: /INSIDE/ diagnostics (klocat in/out). */
int ora_pot3(int jmax, int kmax, int lmax, const double *r, const double *z, const double *rhf,
             const double *zhf, double dtheta, double pi, double grav, int npoint,
             const double *rho, double *phi_io, double *coef, double *phichk, int *klocat)
{
    ora_grid G = {jmax, kmax, lmax, r, z, rhf, zhf, dtheta, pi, grav};
    const ora_grid *g = &G;
    if (jmax < 3 || kmax < 2 || lmax < 2 || (lmax & 1) || npoint < 1) return 1;
    int N = lmax / 2, M = jmax - 1, NZ = kmax - 1;
    real sf1 = (real)0.5 / (real)N, sf2 = 0.5;
    size_t cplane = JS(g) * KS(g), ntot = cplane * (size_t)lmax;
    real *phi = to_real(phi_io, ntot);

    /* rho -> phi on J=2..JMAX, K=2..KMAX (pot3.f:71-81) */
    for (int l = 1; l <= lmax; ++l)
        for (int k = 2; k <= kmax; ++k)
            for (int j = 2; j <= jmax; ++j) phi[P3(g, j, k, l)] = rho[P3(g, j, k, l)];

    /* forward transform of every lane J=2..JMAX+1, K=2..KMAX+1 (pot3.f:92-147) */
#pragma omp parallel
    {
        dft_plan pl;
        real *A = malloc(sizeof(real) * (N + 1)), *B = malloc(sizeof(real) * (N + 1));
        plan_init(&pl, lmax);
#pragma omp for collapse(2) schedule(static)
        for (int k = 2; k <= kmax + 1; ++k)
            for (int j = 2; j <= jmax + 1; ++j) {
                rfft_fwd(&pl, phi + P3(g, j, k, 1), (long)cplane, A, B);
                if (coef) {
                    real x1 = fabs(A[0]);
                    if (x1 != 0) {
                        for (int m = 1; m <= N - 1; ++m)
                            coef[P3(g, j, k, 1) + cplane * (size_t)(m - 1)] =
                                (double)(2 * sqrt(A[m] * A[m] + B[m] * B[m]) / x1);
                        coef[P3(g, j, k, 1) + cplane * (size_t)(N - 1)] = (double)(fabs(A[N]) / x1);
                    }
                    size_t ph0 = cplane * (size_t)N;
                    for (int m = 1; m <= N - 1; ++m) {
                        real philag;
                        if (A[m] > 0) philag = atan(B[m] / A[m]);
                        else if (A[m] < 0) philag = atan(B[m] / A[m]) + (real)pi;
                        else if (B[m] > 0) philag = (real)0.5 * (real)pi;
                        else if (B[m] < 0) philag = (real)-0.5 * (real)pi;
                        else philag = 4 * (real)pi;
                        coef[P3(g, j, k, 1) + ph0 + cplane * (size_t)(m - 1)] =
                            (double)(philag * 180 / (real)pi);
                    }
                    coef[P3(g, j, k, 1) + ph0 + cplane * (size_t)(N - 1)] = (double)A[0];
                }
                for (int l = 1; l <= N + 1; ++l) {
                    phi[P3(g, j, k, l)] = A[l - 1] * sf1;
                    if (l < N) phi[P3(g, j, k, l + N + 1)] = B[l] * sf1;
                }
            }
        plan_free(&pl);
        free(A); free(B);
    }

    /* mode solves (pot3.f:158-248) */
    pot3_coef pc;
    pot3_coef_init(g, &pc);
    zbasis zb;
    int rc = zbasis_init(&zb, NZ, pc.an, pc.bn, pc.cn);
    if (rc) { zbasis_free(&zb); pot3_coef_free(&pc); free(phi); return rc; }
    /* or BLKTRI's cyclic reduction (binary64, the reference's arithmetic) */
    cr_roots crt = {0};
    double *cr_an = NULL, *cr_cnp = NULL, *cr_am = NULL, *cr_cm = NULL;
    if (ora_modes == 1) {
        cr_an = malloc(sizeof(double) * NZ); cr_cnp = malloc(sizeof(double) * NZ);
        double *bn = malloc(sizeof(double) * NZ), *cn = malloc(sizeof(double) * NZ);
        cr_am = malloc(sizeof(double) * M); cr_cm = malloc(sizeof(double) * M);
        for (int k = 0; k < NZ; ++k) { cr_an[k] = (double)pc.an[k]; bn[k] = (double)pc.bn[k]; cn[k] = (double)pc.cn[k]; }
        for (int j = 0; j < M; ++j) { cr_am[j] = (double)pc.am[j]; cr_cm[j] = (double)pc.cm[j]; }
        cr_cnp[0] = cn[NZ - 1];
        for (int k = 1; k < NZ; ++k) cr_cnp[k] = cn[k - 1];
        rc = cr_roots_init(&crt, NZ, cr_an, bn, cn);
        free(bn); free(cn);
        if (rc) {
            free(cr_an); free(cr_cnp); free(cr_am); free(cr_cm);
            zbasis_free(&zb); pot3_coef_free(&pc); free(phi);
            return rc;
        }
    }
    real pig4 = 4 * (real)pi * (real)grav;
#pragma omp parallel
    {
        real *y = malloc(sizeof(real) * (size_t)M * NZ);
        real *bm = malloc(sizeof(real) * M);
        real *work = malloc(sizeof(real) * ((size_t)M * NZ + M));
#pragma omp for schedule(dynamic, 1)
        for (int l = 1; l <= lmax; ++l) {
            int mm = (l <= N + 1) ? l - 1 : l - (N + 1);
            real powrm = (mm % 2) ? -1 : 1;
            real xlam = (double)cos((double)mm * dtheta);
            for (int k = 2; k <= kmax; ++k)
                for (int j = 2; j <= jmax; ++j)
                    y[(size_t)(j - 2) + (size_t)M * (k - 2)] = pig4 * phi[P3(g, j, k, l)];
            for (int j = 2; j <= jmax; ++j)
                y[(size_t)(j - 2) + (size_t)M * (kmax - 2)] -= pc.cnmax * phi[P3(g, j, kmax + 1, l)];
            for (int k = 2; k <= kmax; ++k)
                y[(size_t)(jmax - 2) + (size_t)M * (k - 2)] -= pc.cmmax * phi[P3(g, jmax + 1, k, l)];
            bm[0] = -pc.cm[0] + (powrm - 1) * pc.ammin + 2 * (xlam - 1) * pc.c[0];
            for (int j = 2; j <= jmax - 2; ++j)
                bm[j - 1] = -pc.cm[j - 1] - pc.am[j - 1] + 2 * (xlam - 1) * pc.c[j - 1];
            bm[jmax - 2] = -pc.cmmax - pc.am[jmax - 2] + 2 * (xlam - 1) * pc.c[jmax - 2];
            if (ora_modes == 1) {
                double *yd = malloc(sizeof(double) * (size_t)M * NZ), *bmd = malloc(sizeof(double) * M);
                for (size_t i = 0; i < (size_t)M * NZ; ++i) yd[i] = (double)y[i];
                for (int j = 0; j < M; ++j) bmd[j] = (double)bm[j];
                cr_solve(&crt, M, cr_an, cr_cnp, cr_am, bmd, cr_cm, yd);
                for (size_t i = 0; i < (size_t)M * NZ; ++i) y[i] = yd[i];
                free(yd); free(bmd);
            } else {
                mode_solve(&zb, M, pc.am, bm, pc.cm, y, work);
            }
            for (int k = 2; k <= kmax; ++k)
                for (int j = 2; j <= jmax; ++j)
                    phi[P3(g, j, k, l)] = y[(size_t)(j - 2) + (size_t)M * (k - 2)];
        }
        free(y); free(bm); free(work);
    }
    zbasis_free(&zb);
    pot3_coef_free(&pc);
    if (ora_modes == 1) {
        cr_roots_free(&crt);
        free(cr_an); free(cr_cnp); free(cr_am); free(cr_cm);
    }

    /* inverse transform (pot3.f:256-283) */
#pragma omp parallel
    {
        dft_plan pl;
        real *A = calloc(N + 1, sizeof(real)), *B = calloc(N + 1, sizeof(real));
        real *out = malloc(sizeof(real) * lmax);
        plan_init(&pl, lmax);
#pragma omp for collapse(2) schedule(static)
        for (int k = 2; k <= kmax + 1; ++k)
            for (int j = 2; j <= jmax + 1; ++j) {
                for (int l = 1; l <= N + 1; ++l) A[l - 1] = phi[P3(g, j, k, l)];
                B[0] = 0; B[N] = 0;
                for (int l = 1; l < N; ++l) B[l] = phi[P3(g, j, k, l + N + 1)];
                rfft_inv(&pl, A, B, out);
                for (int l = 1; l <= lmax; ++l) phi[P3(g, j, k, l)] = sf2 * out[l - 1];
            }
        plan_free(&pl);
        free(A); free(B); free(out);
    }

    /* ghosts (pot3.f:287-298): axis J=1 is the value across the axis, K=1 mirrors K=2 */
    for (int l = 1; l <= N; ++l) {
        for (int k = 2; k <= kmax + 1; ++k) {
            phi[P3(g, 1, k, l)] = phi[P3(g, 2, k, N + l)];
            phi[P3(g, 1, k, N + l)] = phi[P3(g, 2, k, l)];
        }
        for (int j = 1; j <= jmax + 1; ++j) {
            phi[P3(g, j, 1, l)] = phi[P3(g, j, 2, l)];
            phi[P3(g, j, 1, N + l)] = phi[P3(g, j, 2, N + l)];
        }
    }
    zaxphi(g, npoint, phi, phichk, klocat);
    for (size_t i = 0; i < ntot; ++i) phi_io[i] = (double)phi[i];
    free(phi);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* SETBDY + BDYGEN (pot/boundary.f)                                          */

/* One (l,m) multipole term with l+m even (the ISYM=2 set, boundary.f:412-500).
 * blm(c,s) = s^m * sum_i poly[i] c^(2i) is the reference's BLM statement
 * function (boundary.f:204-253); dlm is the boundary factor of
 * boundary.f:694-772 formed from the reference's TW constants as REAL*8. */
typedef struct { int l, m; double dlm; int np; double poly[6]; } mterm;

#define TW5 0.03125
#define TW6 0.015625
#define TW7 7.8125e-3
#define TW8 3.90625e-3
#define TW9 1.953125e-3
#define TW10 9.765625e-04
#define TW11 4.8828125e-4
#define TW12 2.44140625e-4
#define TW13 1.220703125e-4
#define TW14 6.103515625e-5
#define TW15 3.051757813e-5
#define TW16 1.525878907e-5
#define TW17 7.629394535e-6
#define TW18 3.814697266e-6

static const mterm TERMS[] = {
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
#define NTERMS ((int)(sizeof(TERMS) / sizeof(TERMS[0])))

/* exported for the table checks in tests/: l, m, dlm, np, poly[6] per term */
int ora_terms(int *l, int *m, double *dlm, int *np, double *poly)
{
    for (int t = 0; t < NTERMS; ++t) {
        l[t] = TERMS[t].l; m[t] = TERMS[t].m; dlm[t] = TERMS[t].dlm; np[t] = TERMS[t].np;
        for (int i = 0; i < 6; ++i) poly[t * 6 + i] = TERMS[t].poly[i];
    }
    return NTERMS;
}

static real blm(const mterm *t, real c, real s)
{
    real c2 = c * c, p = 0, cc = 1;
    for (int i = 0; i < t->np; ++i) { p += (real)t->poly[i] * cc; cc *= c2; }
    real sm = 1;
    for (int i = 0; i < t->m; ++i) sm *= s;
    return p * sm;
}

static real ipow(real x, int n)
{
    real r = 1;
    for (int i = 0; i < n; ++i) r *= x;
    return r;
}

/* Boundary point list of SETBDY for ISYM=2 (boundary.f:50-79): top edge
 * K=KMAX+1, J=2..JMAX+1, then the outer edge J=JMAX+1, K=2..KMAX. */
static int boundary_points(const ora_grid *g, int *jpos, int *kpos, real *rbdy)
{
    int i = 0;
    real zk2 = (real)g->zhf[g->kmax] * g->zhf[g->kmax];
    for (int j = 2; j <= g->jmax + 1; ++j, ++i) {
        rbdy[i] = sqrt(zk2 + (real)g->rhf[j - 1] * g->rhf[j - 1]);
        jpos[i] = j; kpos[i] = g->kmax + 1;
    }
    real rj2 = (real)g->rhf[g->jmax] * g->rhf[g->jmax];
    for (int k = 2; k <= g->kmax; ++k, ++i) {
        rbdy[i] = sqrt(rj2 + (real)g->zhf[k - 1] * g->zhf[k - 1]);
        jpos[i] = g->jmax + 1; kpos[i] = k;
    }
    return i;
}

int ora_bdygen(int jmax, int kmax, int lmax, const double *r, const double *z, const double *rhf,
               const double *zhf, double dtheta, double pi, double grav, int maxtrm, int isym,
               double redge, const double *rho, double *phi, double *tmass_out)
{
    ora_grid G = {jmax, kmax, lmax, r, z, rhf, zhf, dtheta, pi, grav};
    const ora_grid *g = &G;
    if (abs(isym) != 2 || maxtrm < 0 || maxtrm > 10 || lmax < 2 || (lmax & 1)) return 1;
    int nb = jmax + kmax - 1;
    int *jpos = malloc(sizeof(int) * nb), *kpos = malloc(sizeof(int) * nb);
    real *rbdy = malloc(sizeof(real) * nb);
    int jkmax = boundary_points(g, jpos, kpos, rbdy);
    int nt = 0;
    while (nt < NTERMS && TERMS[nt].l <= maxtrm) ++nt;

    /* cos/sin tables at cell-centre angles (L-1/2)*DTHETA (boundary.f:31-37) */
    real *cosm = malloc(sizeof(real) * lmax * 10), *sinm = malloc(sizeof(real) * lmax * 10);
    for (int l = 1; l <= lmax; ++l) {
        real psi = ((real)l - (real)0.5) * (real)dtheta;
        for (int m = 1; m <= 10; ++m) {
            cosm[(l - 1) * 10 + m - 1] = cos((real)m * psi);
            sinm[(l - 1) * 10 + m - 1] = sin((real)m * psi);
        }
    }

    /* ring moments (boundary.f:366-554): MRING per ring and term */
    int nring = (jmax - 1) * (kmax - 1);
    real *mr1 = calloc((size_t)nring * nt, sizeof(real));
    real *mr2 = calloc((size_t)nring * nt, sizeof(real));
    real *mo1 = calloc((size_t)nring * nt, sizeof(real));
    real *mo2 = calloc((size_t)nring * nt, sizeof(real));
    real *rsp = malloc(sizeof(real) * nring);
    char *used = calloc(nring, 1);
    real tmass = 0;
    const real f2 = 2;
    for (int kk = 2; kk <= kmax; ++kk) {
        real zz = zhf[kk - 1], z2 = zz * zz, zdel = (real)z[kk] - (real)z[kk - 1];
        for (int jj = 2; jj <= jmax; ++jj) {
            int ir = (jj - 2) + (jmax - 1) * (kk - 2);
            real rr = rhf[jj - 1], r2 = rr * rr, rspher = sqrt(r2 + z2);
            rsp[ir] = rspher;
            if (redge > 0.0 && rspher > redge) continue;
            used[ir] = 1;
            real c1 = zz / rspher, s1 = rr / rspher;
            real dv = (real)0.5 * (real)dtheta * zdel * ((real)r[jj] * r[jj] - (real)r[jj - 1] * r[jj - 1]) * f2;
            real cm[11] = {0}, sm[11] = {0}, sumass = 0;
            for (int ll = 1; ll <= lmax; ++ll) {
                real h = rho[P3(g, jj, kk, ll)] * dv;
                sumass += h;
                for (int m = 1; m <= maxtrm; ++m) {
                    cm[m] += cosm[(ll - 1) * 10 + m - 1] * h;
                    sm[m] += sinm[(ll - 1) * 10 + m - 1] * h;
                }
            }
            for (int t = 0; t < nt; ++t) {
                const mterm *tt = &TERMS[t];
                real term = blm(tt, c1, s1) * ipow(rspher, tt->l);
                size_t o = (size_t)ir * nt + t;
                if (tt->m == 0) { mr1[o] = term * sumass; mr2[o] = 0; }
                else { mr1[o] = term * cm[tt->m]; mr2[o] = term * sm[tt->m]; }
                /* outside a boundary point the ring contributes MRING*r^-(2l+1) (boundary.f:585-593) */
                real qh = 1 / ipow(rspher, 2 * tt->l + 1);
                mo1[o] = mr1[o] * qh;
                mo2[o] = mr2[o] * qh;
            }
            tmass += sumass; /* MRING(1,1) = TERM(1)*SUMASS, TERM(1)=1 */
        }
    }

    /* inside/outside accumulation per boundary point (boundary.f:556-604),
     * rings visited in the reference's KK-outer, JJ-inner order */
    real *min1 = calloc((size_t)jkmax * nt, sizeof(real)), *min2 = calloc((size_t)jkmax * nt, sizeof(real));
    real *mou1 = calloc((size_t)jkmax * nt, sizeof(real)), *mou2 = calloc((size_t)jkmax * nt, sizeof(real));
    int njk = isym > 0 ? jkmax : 1;
#pragma omp parallel for schedule(dynamic, 4)
    for (int jk = 0; jk < njk; ++jk) {
        for (int ir = 0; ir < nring; ++ir) {
            if (!used[ir]) continue;
            if (isym < 0 || rsp[ir] <= rbdy[jk]) {
                for (int t = 0; t < nt; ++t) {
                    min1[(size_t)jk * nt + t] += mr1[(size_t)ir * nt + t];
                    min2[(size_t)jk * nt + t] += mr2[(size_t)ir * nt + t];
                }
            } else {
                for (int t = 0; t < nt; ++t) {
                    mou1[(size_t)jk * nt + t] += mo1[(size_t)ir * nt + t];
                    mou2[(size_t)jk * nt + t] += mo2[(size_t)ir * nt + t];
                }
            }
        }
    }

    /* boundary potential (boundary.f:659-830) */
#pragma omp parallel for schedule(static)
    for (int jk = 0; jk < jkmax; ++jk) {
        int jj = jpos[jk], kk = kpos[jk];
        real rs = rbdy[jk], c1 = zhf[kk - 1] / rs, s1 = rhf[jj - 1] / rs;
        real t1[64], t2[64];
        size_t src = isym > 0 ? (size_t)jk : 0;
        for (int t = 0; t < nt; ++t) {
            const mterm *tt = &TERMS[t];
            real term = (real)tt->dlm * blm(tt, c1, s1);
            real rin = 1 / ipow(rs, tt->l + 1), rout = ipow(rs, tt->l);
            real b = term * rin, h = term * rout;
            t1[t] = b * min1[src * nt + t] + h * mou1[src * nt + t];
            t2[t] = b * min2[src * nt + t] + h * mou2[src * nt + t];
        }
        for (int ll = 1; ll <= lmax; ++ll) {
            real sum = 0;
            for (int t = 0; t < nt; ++t) {
                int m = TERMS[t].m;
                if (m == 0) sum += t1[t];
                else sum += 2 * (cosm[(ll - 1) * 10 + m - 1] * t1[t] + sinm[(ll - 1) * 10 + m - 1] * t2[t]);
            }
            phi[P3(g, jj, kk, ll)] = (double)(-(real)grav * sum);
        }
    }

    /* ISYM=2 symmetry ghosts (boundary.f:876-893) */
    int lhaf = lmax / 2;
    for (int ll = 1; ll <= lmax; ++ll) phi[P3(g, jmax + 1, 1, ll)] = phi[P3(g, jmax + 1, 2, ll)];
    for (int ll = 1; ll <= lhaf; ++ll) {
        int lp = ll + lhaf;
        phi[P3(g, 1, kmax + 1, ll)] = phi[P3(g, 2, kmax + 1, lp)];
        phi[P3(g, 1, kmax + 1, lp)] = phi[P3(g, 2, kmax + 1, ll)];
    }
    *tmass_out = (double)tmass;

    free(jpos); free(kpos); free(rbdy); free(cosm); free(sinm); free(mr1); free(mr2);
    free(mo1); free(mo2); free(rsp); free(used); free(min1); free(min2); free(mou1); free(mou2);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Discrete 7-point operator of pot/pot3.f:158-231 applied to a physical
 * field: res = A phi - 4 pi G rho on J=2..JMAX, K=2..KMAX, using phi's own
 * J=JMAX+1 and K=KMAX+1 values; the axis (AM(1)=0) and the K=1 mirror
 * (AN(1)=0, BN(1)=-CN(1)) close the stencil, periodic in L.  scale (may be
 * NULL) receives sum |terms| per cell, the size of the rounding floor. */
int ora_residual(int jmax, int kmax, int lmax, const double *rhf, const double *zhf, double dtheta,
                 double pi, double grav, const double *phi, const double *rho, double *res,
                 double *scale)
{
    ora_grid G = {jmax, kmax, lmax, NULL, NULL, rhf, zhf, dtheta, pi, grav};
    const ora_grid *g = &G;
    pot3_coef pc;
    pot3_coef_init(g, &pc);
    real pig4 = 4 * (real)pi * (real)grav;
    for (int l = 1; l <= lmax; ++l) {
        int lm = l == 1 ? lmax : l - 1, lp = l == lmax ? 1 : l + 1;
        for (int k = 2; k <= kmax; ++k)
            for (int j = 2; j <= jmax; ++j) {
                real p0 = phi[P3(g, j, k, l)];
                real cmj = j == jmax ? pc.cmmax : pc.cm[j - 2];
                real cnk = k == kmax ? pc.cnmax : pc.cn[k - 2];
                real t[6];
                t[0] = cmj * (phi[P3(g, j + 1, k, l)] - p0);
                t[1] = pc.am[j - 2] * (phi[P3(g, j - 1, k, l)] - p0);
                t[2] = cnk * (phi[P3(g, j, k + 1, l)] - p0);
                t[3] = pc.an[k - 2] * (phi[P3(g, j, k - 1, l)] - p0);
                t[4] = pc.c[j - 2] * (phi[P3(g, j, k, lp)] - 2 * p0 + phi[P3(g, j, k, lm)]);
                t[5] = -pig4 * rho[P3(g, j, k, l)];
                real a = 0, s = 0;
                for (int i = 0; i < 6; ++i) { a += t[i]; s += fabs(t[i]); }
                res[P3(g, j, k, l)] = (double)a;
                if (scale) scale[P3(g, j, k, l)] = (double)s;
            }
    }
    pot3_coef_free(&pc);
    return 0;
}

/*
 * Momentum sources that read the potential: the first loop of SOURCE
 * (radhydro/source.f:41-79, BINARY=0) with the loop bounds of source.f:41-48
 * (L periodic, K=2..KMAX+1), S from J=JSTART+1 (source.f:51), T and A from
 * J=JSTART (source.f:63,70), JSTART = max(JMIN, 2) (source.f:31-35).
 * Always binary64 (also in the extended build), each operation in the order
 * the Fortran expressions give: parenthesised sums first, then products and
 * quotients left to right (source.f:52-56, 64-66, 71-72).  x86-64 without FMA
 * rounds every operation separately, as the reference build does.
 * Arrays are Fortran-ordered; rho and phi have the potential grid's leading
 * dimensions (pjmax+2, pkmax+2), the others the hydro grid's (jmax+2, kmax+2).
 */
int ora_source_grav(int jmax, int kmax, int lmax, int pjmax, int pkmax, int jstart, double delt,
                    double dr, double dz, double dtheta, const double *r, const double *rho,
                    const double *omega, const double *phi, const double *p, double *s, double *t,
                    double *a)
{
    if (jmax < 1 || kmax < 1 || lmax < 1 || pjmax < jmax || pkmax < kmax || jstart < 2 || jstart > jmax + 1)
        return 1;
    const long hj = jmax + 2, hk = kmax + 2, pj = pjmax + 2, pk = pkmax + 2;
#define HX(J, K, L) (((long)(L) - 1) * hk * hj + ((long)(K) - 1) * hj + (long)(J) - 1)
#define PX(J, K, L) (((long)(L) - 1) * pk * pj + ((long)(K) - 1) * pj + (long)(J) - 1)
    for (int L = 1; L <= lmax; ++L) {
        int LM = L - 1, LP = L + 1;
        if (L == 1) LM = lmax;
        if (L == lmax) LP = 1;
        for (int K = 2; K <= kmax + 1; ++K) {
            for (int J = jstart + 1; J <= jmax + 1; ++J) {
                double rr = rho[PX(J, K, L)] + rho[PX(J - 1, K, L)];
                double oo = omega[HX(J, K, L)] + omega[HX(J - 1, K, L)];
                double t1 = r[J - 1] * rr * (oo * oo) * 0.25 * 0.5;
                double t2 = rr * (phi[PX(J, K, L)] - phi[PX(J - 1, K, L)]) / (2.0 * dr);
                double t3 = (p[HX(J, K, L)] - p[HX(J - 1, K, L)]) / dr;
                s[HX(J, K, L)] = s[HX(J, K, L)] + delt * (t1 - t2 - t3);
            }
            for (int J = jstart; J <= jmax + 1; ++J) {
                double t1 = 0.5 * (rho[PX(J, K, L)] + rho[PX(J, K - 1, L)]) *
                            (phi[PX(J, K, L)] - phi[PX(J, K - 1, L)]) / dz;
                double t2 = (p[HX(J, K, L)] - p[HX(J, K - 1, L)]) / dz;
                t[HX(J, K, L)] = t[HX(J, K, L)] - delt * (t1 + t2);
            }
            for (int J = jstart; J <= jmax + 1; ++J) {
                double x = rho[PX(J, K, L)] * (phi[PX(J, K, LP)] - phi[PX(J, K, LM)]) + p[HX(J, K, LP)] -
                           p[HX(J, K, LM)];
                a[HX(J, K, L)] = a[HX(J, K, L)] - delt * x / (2.0 * dtheta);
            }
        }
    }
#undef HX
#undef PX
    return 0;
}

/*
 * The central star's potential added to PHI: ExternalPot (pot/ExternalPot.f:1-88)
 * with EXTERNAL_POT=1, INDIRECT=0, BINARY=0 (the default build's flags).  The
 * star sits at (RPSTAR cos PHI_STAR, RPSTAR sin PHI_STAR) (ExternalPot.f:18-19);
 * for J=1..JMAX+2, K=1..KMAX+2, L=1..LMAX (ExternalPot.f:43-51) STARPHI gets
 * -(MASS_STAR+TMASSACC)/sqrt(RHF^2 + ZHF^2 + x*^2 + y*^2 - 2 RHF (cos psi x* +
 * sin psi y*)) with psi = (L-1+1/2) DTHETA, and PHI += STARPHI.  Binary64,
 * operations in the Fortran expression's order.  phi has the potential grid's
 * leading dimensions (pjmax+2, pkmax+2), starphi the hydro grid's.
 */
int ora_external_pot(int jmax, int kmax, int lmax, int pjmax, int pkmax, double dtheta, double mass_star,
                     double tmassacc, double rpstar, double phi_star, const double *rhf, const double *zhf,
                     double *phi, double *starphi)
{
    if (jmax < 1 || kmax < 1 || lmax < 1 || pjmax < jmax || pkmax < kmax) return 1;
    const long hj = jmax + 2, hk = kmax + 2, pj = pjmax + 2, pk = pkmax + 2;
    const double ix = rpstar * cos(phi_star), iy = rpstar * sin(phi_star);
    for (int L = 1; L <= lmax; ++L) {
        const double psi = ((double)(L - 1) + 0.5) * dtheta;
        const double c = cos(psi), s = sin(psi);
        for (int K = 1; K <= kmax + 2; ++K)
            for (int J = 1; J <= jmax + 2; ++J) {
                const double rh = rhf[J - 1], zh = zhf[K - 1];
                const double sp = -(mass_star + tmassacc) /
                                  sqrt(rh * rh + zh * zh + ix * ix + iy * iy - 2.0 * rh * (c * ix + s * iy));
                starphi[(L - 1) * hk * hj + (K - 1) * hj + (J - 1)] = sp;
                const long q = (L - 1) * pk * pj + (K - 1) * pj + (J - 1);
                phi[q] = phi[q] + sp;
            }
    }
    return 0;
}

/*
 * The disk's pull on the star: find_force (pot/wiggle.F:99-135), summed in the
 * loop order of one OpenMP thread (L, then K, then J).  rho has the leading
 * dimensions (JMAX+2, KMAX+2) find_force declares (wiggle.F:103).  Each term
 * is formed in binary64 as the reference forms it; the sums are kept in the
 * working precision (binary64, or 80-bit in the extended build, where they are
 * close to the exact sums of the binary64 terms).  f gets (fx, fy); fabs_sum
 * (may be NULL) gets the sums of |terms|, the scale of the rounding
 * differences between summation orders.
 */
int ora_find_force(int jmax, int kmax, int lmax, int jmin, double dr, double dz, double dphi, double gsoft,
                   const double *rhf, const double *zhf, const double *rho, double x, double y, double *f,
                   double *fabs_sum)
{
    if (jmax < 1 || kmax < 1 || lmax < 1 || jmin < 4 || jmin > jmax + 1) return 1;
    const long hj = jmax + 2, hk = kmax + 2;
    const double vol = dr * dz * dphi * 2.0;                      /* wiggle.F:107 */
    real fx = 0, fy = 0, ax = 0, ay = 0;
    for (int L = 1; L <= lmax; ++L) {
        const double lphi = (double)L * dphi - 0.5 * dphi;         /* wiggle.F:115 */
        for (int K = 2; K <= kmax + 1; ++K) {
            const double zz = zhf[K - 1];
            for (int J = jmin - 3; J <= jmax + 1; ++J) {
                const double xx = rhf[J - 1] * cos(lphi) - x;
                const double yy = rhf[J - 1] * sin(lphi) - y;
                const double tmass = rho[(L - 1) * hk * hj + (K - 1) * hj + (J - 1)] * rhf[J - 1] * vol;
                const double s = sqrt(xx * xx + yy * yy + zz * zz + gsoft * gsoft);
                const double dd = 1.0 / (s * s * s);                /* wiggle.F:123 */
                const double tx = tmass * dd * xx, ty = tmass * dd * yy;
                fx = fx + tx;
                fy = fy + ty;
                ax += fabs(tx);
                ay += fabs(ty);
            }
        }
    }
    f[0] = (double)fx;
    f[1] = (double)fy;
    if (fabs_sum) {
        fabs_sum[0] = (double)ax;
        fabs_sum[1] = (double)ay;
    }
    return 0;
}
