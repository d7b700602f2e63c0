// Residual of the discrete potential equation for iterative refinement.
//
// The mode systems of pot/pot3.f:158-231 are, in physical space, the 7-point
// operator
//   (A PHI)(j,k,l) = CM(j) (PHI(j+1)-PHI) + AM(j) (PHI(j-1)-PHI)
//                  + CN(k) (PHI(k+1)-PHI) + AN(k) (PHI(k-1)-PHI)
//                  + C(j) (PHI(l+1) - 2 PHI + PHI(l-1))
//                  + AMMIN (PHI(J=2,k,l+N) - PHI(J=2,k,l))          [j = J=2 only]
// with AM(J=2) = 0, AN(K=2) = 0 (the midplane mirror), CM(JMAX) = CMMAX and
// CN(KMAX) = CNMAX multiplying the boundary values PHI(JMAX+1,.) and
// PHI(.,KMAX+1); its per-mode form is BM_m = -CM - AM + 2(cos m DTHETA - 1) C
// plus ((-1)^m - 1) AMMIN on the first row (pot3.f:226-230).  The residual
// r = 4 pi G RHO - A PHI is computed in binary64 and stored in binary32: it is
// the right-hand side of the correction solve, whose boundary values are 0.
//
// One thread per interior column (J,K) and a run of LC consecutive L, with a
// sliding window over L; neighbouring J are neighbouring threads (their loads
// share cache lines), the run of L keeps the grid at ~1000 blocks.
#include "chypot_impl.h"

namespace chy {
namespace {

constexpr int LC = 32;  // L values per thread (fewer when LMAX < 32)

struct ResArgs {
    const double *phi, *rho;
    float *res;
    const double *am, *cm, *c, *an, *cn;
    unsigned long long *stat;
    double ammin, cmmax, cnmax, pig4;
    int J, K, L, JS, KS, M, NZ;
    int first, clear_stat, lc;
};

__global__ __launch_bounds__(256) void k_residual(ResArgs a)
{
    if (a.clear_stat && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x < 2 * kMaxRefine)
        a.stat[threadIdx.x] = 0ull;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;  // J-2, 0..M-1
    const int k = blockIdx.y;                              // K-2, 0..NZ-1
    if (j >= a.M) return;
    const int l0 = blockIdx.z * a.lc;
    const int N = a.L / 2;
    const size_t plane = (size_t)a.JS * a.KS;
    const size_t base = (size_t)(j + 1) + (size_t)a.JS * (size_t)(k + 1);
    const bool jlast = j == a.M - 1, klast = k == a.NZ - 1;
    const double cmj = jlast ? a.cmmax : a.cm[j];
    const double amj = a.am[j];  // 0 on the first row
    const double cj = a.c[j];
    const double cnk = klast ? a.cnmax : a.cn[k];
    const double ank = a.an[k];  // 0 on the first row
    const bool axis = j == 0 && a.ammin != 0.0;
    // PHI at a neighbour: boundary cells (J=JMAX+1 or K=KMAX+1) always count;
    // interior cells only after the first step (before it they may hold
    // anything, so they are not used at all)
    auto at = [&](size_t i, bool boundary) -> double {
        const double v = a.phi[i];
        return (boundary || !a.first) ? v : 0.0;
    };
    const int lm = (l0 + a.L - 1) % a.L;
    double pm = at(base + plane * lm, false), p0 = at(base + plane * l0, false);
    for (int t = 0; t < a.lc; ++t) {
        const int l = l0 + t;
        const int lp = (l + 1) % a.L;
        const size_t i = base + plane * l;
        const double pp = at(base + plane * lp, false);
        const double pj1 = at(i + 1, jlast);
        const double pj0 = j > 0 ? at(i - 1, false) : 0.0;
        const double pk1 = at(i + a.JS, klast);
        const double pk0 = k > 0 ? at(i - a.JS, false) : 0.0;
        double ap = cmj * (pj1 - p0);
        if (j > 0) ap += amj * (pj0 - p0);
        ap += cnk * (pk1 - p0);
        if (k > 0) ap += ank * (pk0 - p0);
        ap += cj * (pp - 2.0 * p0 + pm);
        if (axis) ap += a.ammin * (at(base + plane * ((l + N) % a.L), false) - p0);
        a.res[i] = (float)(a.pig4 * a.rho[i] - ap);
        pm = p0;
        p0 = pp;
    }
}

}  // namespace

int launch_residual(const chy_grid *g, const double *rho, const double *phi, bool first, bool clear_stat,
                    hipStream_t s)
{
    ResArgs a{};
    a.phi = phi;
    a.rho = rho;
    a.res = g->d.res32;
    a.am = g->d.am;
    a.cm = g->d.cm;
    a.c = g->d.rc_c;
    a.an = g->d.rz_an;
    a.cn = g->d.rz_cn;
    a.stat = g->d.rstat;
    a.ammin = g->ammin; a.cmmax = g->cmmax; a.cnmax = g->cnmax; a.pig4 = g->pig4;
    a.J = g->J; a.K = g->K; a.L = g->L; a.JS = g->J + 2; a.KS = g->K + 2;
    a.M = g->M; a.NZ = g->NZ;
    a.first = first ? 1 : 0;
    a.clear_stat = clear_stat ? 1 : 0;
    a.lc = g->L < LC ? g->L : LC;  // LMAX is a power of two
    const dim3 grid((g->M + 255) / 256, g->NZ, g->L / a.lc);
    hipLaunchKernelGGL(k_residual, grid, dim3(256), 0, s, a);
    return (int)hipGetLastError();
}

}  // namespace chy
