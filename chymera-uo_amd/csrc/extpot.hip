// The central star's potential added to PHI after the self-gravity solve:
// ExternalPot (pot/ExternalPot.f:1-88) as the default build runs it
// (EXTERNAL_POT=1, INDIRECT=0, BINARY=0; the hydro step calls it right after
// POT3 and the mass sum, radhydro/3dhyd-main.f:494-511).  With the star at
// (x*, y*) = (RPSTAR cos PHI_STAR, RPSTAR sin PHI_STAR) (ExternalPot.f:18-19),
// for J = 1..JMAX+2, K = 1..KMAX+2, L = 1..LMAX (ExternalPot.f:43-51):
//
//   STARPHI(J,K,L) = -(MASS_STAR+TMASSACC) / sqrt( RHF(J)^2 + ZHF(K)^2 + x*^2
//                     + y*^2 - 2 RHF(J) (cos(psi_L) x* + sin(psi_L) y*) )
//   PHI(J,K,L)    += STARPHI(J,K,L),        psi_L = (L-1+1/2) DTHETA
//
// STARPHI has the hydro grid's layout, PHI the potential grid's.  Sums and
// products are taken in the Fortran expression's order with one rounding each
// (no fused multiply-add); cos/sin of psi_L come from a table the host fills
// with the C library's functions, and sqrt and the quotient are correctly
// rounded, so the results are the host's to the bit.
//
// An HBM-bound pass that reads PHI and writes PHI and STARPHI (24 bytes per
// cell).  Each plane L of the hydro grid, (JMAX+2)*(KMAX+2) cells, is one
// contiguous run; a block takes EP_CELLS consecutive cells of one plane, and
// thread t its cells t, t+256, ..., so every load and store instruction of a
// wavefront covers 512 contiguous bytes.  All of a thread's PHI loads are
// issued before the first is used.
#include "chypot_impl.h"

namespace chy {
namespace {

struct ExtArgs {
    const double *rhf, *zhf, *cs;  // RHF(1..), ZHF(1..), (cos psi_L, sin psi_L) pairs
    double *phi, *starphi;
    int j2, k2, pj2, pk2, lmax;    // hydro JMAX+2, KMAX+2; potential grid POT3JMAX+2, POT3KMAX+2
    double mass, ix, iy;           // MASS_STAR+TMASSACC, x*, y*
};

constexpr int EP_PER = 4;               // cells per thread
constexpr int EP_CELLS = 256 * EP_PER;  // cells per block

__global__ __launch_bounds__(256) void k_external_pot(ExtArgs a)
{
#pragma clang fp contract(off)
    const int L = 1 + (int)blockIdx.y;
    const int plane = a.j2 * a.k2;
    const int e0 = (int)blockIdx.x * EP_CELLS + (int)threadIdx.x;
    const double c = a.cs[2 * (L - 1)], s = a.cs[2 * (L - 1) + 1];
    const size_t hl = (size_t)plane * (L - 1), ql = (size_t)a.pj2 * a.pk2 * (L - 1);
    size_t q[EP_PER];
    double old[EP_PER];
    int jm[EP_PER], km[EP_PER];
#pragma unroll
    for (int i = 0; i < EP_PER; ++i) {
        const int e = e0 + 256 * i;
        const int ec = e < plane ? e : plane - 1;  // clamped: in bounds, stored by nobody
        km[i] = ec / a.j2;
        jm[i] = ec - km[i] * a.j2;
        q[i] = ql + (size_t)jm[i] + (size_t)a.pj2 * km[i];
        old[i] = a.phi[q[i]];
    }
#pragma unroll
    for (int i = 0; i < EP_PER; ++i) {
        const int e = e0 + 256 * i;
        const double rh = a.rhf[jm[i]], zh = a.zhf[km[i]];
        const double d2 = rh * rh + zh * zh + a.ix * a.ix + a.iy * a.iy - 2.0 * rh * (c * a.ix + s * a.iy);
        const double sp = -a.mass / sqrt(d2);
        if (e < plane) {
            a.starphi[hl + (size_t)e] = sp;
            a.phi[q[i]] = old[i] + sp;
        }
    }
}

}  // namespace

int launch_external_pot(const chy_extpot_desc *d, const double *rhf, const double *zhf, const double *cs,
                        double *phi, double *starphi, double ix, double iy, hipStream_t st)
{
    ExtArgs x{};
    x.rhf = rhf; x.zhf = zhf; x.cs = cs; x.phi = phi; x.starphi = starphi;
    x.j2 = d->jmax + 2; x.k2 = d->kmax + 2; x.pj2 = d->pot_jmax + 2; x.pk2 = d->pot_kmax + 2;
    x.lmax = d->lmax;
    x.mass = d->mass_star + d->tmassacc;
    x.ix = ix; x.iy = iy;
    const dim3 grid((x.j2 * x.k2 + EP_CELLS - 1) / EP_CELLS, d->lmax);
    hipLaunchKernelGGL(k_external_pot, grid, dim3(256), 0, st, x);
    return (int)hipGetLastError();
}

}  // namespace chy
