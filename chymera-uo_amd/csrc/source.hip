// Momentum sources of one hydro step that read the potential: the first
// loop of SOURCE (radhydro/source.f:41-79, BINARY=0), which adds to the
// radial (S), vertical (T) and angular (A) momentum densities
//
//   S(J,K,L) += DELT*( R(J)*(RHO(J)+RHO(J-1))*(OMEGA(J)+OMEGA(J-1))**2 /8
//                      - (RHO(J)+RHO(J-1))*(PHI(J)-PHI(J-1))/(2 DR)
//                      - (P(J)-P(J-1))/DR )                 J = JSTART+1..JMAX+1
//   T(J,K,L) -= DELT*( (RHO(K)+RHO(K-1))/2*(PHI(K)-PHI(K-1))/DZ
//                      + (P(K)-P(K-1))/DZ )                 J = JSTART..JMAX+1
//   A(J,K,L) -= DELT*( RHO*(PHI(L+1)-PHI(L-1)) + P(L+1)-P(L-1) )/(2 DTHETA)
//
// for K = 2..KMAX+1 and L = 1..LMAX with L periodic.  RHO and PHI live on the
// potential grid (POT3JMAX+2, POT3KMAX+2, LMAX); S, T, A, OMEGA and P on the
// hydro grid (JMAX+2, KMAX+2, LMAX).  Every product, quotient and sum is taken
// in the order the Fortran expressions give, one rounding each (no fused
// multiply-add), so the result is the reference's to the bit.
//
// One thread per (J, K, L), J fastest: a wavefront reads 64 consecutive J of
// each array.  Blocks of 64 J x 4 K walk one L plane after another, so the
// plane L-1 that a block needs was read by the block with the same (J, K)
// tile of the previous plane -- on the same XCD, since a plane has a
// multiple of 8 blocks whenever KMAX is a multiple of 32 -- and comes from L2.
#include "chypot_impl.h"

namespace chy {
namespace {

struct SourceArgs {
    const double *r, *rho, *omega, *phi, *p;
    double *s, *t, *a;
    int j2, k2;    // hydro pitches JMAX+2, KMAX+2
    int pj2, pk2;  // potential-grid pitches POT3JMAX+2, POT3KMAX+2
    int jstart, jmax1, kmax, lmax;
    double delt, dr, dz, dtheta;
};

__global__ __launch_bounds__(256) void k_source_grav(SourceArgs a)
{
#pragma clang fp contract(off)
    const int J = a.jstart + (int)(blockIdx.x * 64 + threadIdx.x);  // Fortran indices
    const int K = 2 + (int)(blockIdx.y * 4 + threadIdx.y);
    const int L = 1 + (int)blockIdx.z;
    if (J > a.jmax1 || K > a.kmax + 1) return;
    const int LM = L == 1 ? a.lmax : L - 1;
    const int LP = L == a.lmax ? 1 : L + 1;
    const size_t hp = (size_t)a.j2 * a.k2, pp = (size_t)a.pj2 * a.pk2;
    const size_t h = (size_t)(J - 1) + (size_t)a.j2 * (K - 1) + hp * (L - 1);
    const size_t q = (size_t)(J - 1) + (size_t)a.pj2 * (K - 1) + pp * (L - 1);
    const double rho = a.rho[q], phi = a.phi[q], p = a.p[h];

    if (J >= a.jstart + 1) {
        const double rs = rho + a.rho[q - 1];
        const double om = a.omega[h] + a.omega[h - 1];
        const double cen = a.r[J - 1] * rs * (om * om) * 0.25 * 0.5;
        const double grav = rs * (phi - a.phi[q - 1]) / (2.0 * a.dr);
        const double pres = (p - a.p[h - 1]) / a.dr;
        a.s[h] = a.s[h] + a.delt * (cen - grav - pres);
    }
    {
        const double grav = 0.5 * (rho + a.rho[q - a.pj2]) * (phi - a.phi[q - a.pj2]) / a.dz;
        const double pres = (p - a.p[h - a.j2]) / a.dz;
        a.t[h] = a.t[h] - a.delt * (grav + pres);
    }
    {
        const size_t qp = q + pp * (size_t)(LP - L), qm = q - pp * (size_t)(L - LM);
        const size_t hpl = h + hp * (size_t)(LP - L), hml = h - hp * (size_t)(L - LM);
        const double x = rho * (a.phi[qp] - a.phi[qm]) + a.p[hpl] - a.p[hml];
        a.a[h] = a.a[h] - a.delt * x / (2.0 * a.dtheta);
    }
}

}  // namespace

int launch_source_grav(const chy_source_desc *d, const double *r, const double *rho, const double *omega,
                       const double *phi, const double *p, double *s, double *t, double *a, hipStream_t st)
{
    SourceArgs x{};
    x.r = r; x.rho = rho; x.omega = omega; x.phi = phi; x.p = p;
    x.s = s; x.t = t; x.a = a;
    x.j2 = d->jmax + 2; x.k2 = d->kmax + 2;
    x.pj2 = d->pot_jmax + 2; x.pk2 = d->pot_kmax + 2;
    x.jstart = d->jstart; x.jmax1 = d->jmax + 1; x.kmax = d->kmax; x.lmax = d->lmax;
    x.delt = d->delt; x.dr = d->dr; x.dz = d->dz; x.dtheta = d->dtheta;
    const int nj = x.jmax1 - x.jstart + 1;
    const dim3 grid((nj + 63) / 64, (d->kmax + 3) / 4, d->lmax), block(64, 4);
    hipLaunchKernelGGL(k_source_grav, grid, block, 0, st, x);
    return (int)hipGetLastError();
}

}  // namespace chy
