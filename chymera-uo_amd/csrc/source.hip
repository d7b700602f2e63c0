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
// One thread per (J, L) and run of SRC_KC rows K, J fastest: a wavefront
// reads 64 consecutive J of each array.  The thread walks its rows upward and
// carries RHO, PHI and P of row K-1 in registers; a block covers SRC_LB
// consecutive planes, so the L+-1 neighbours of all but its two edge planes
// are elements its own wavefronts read in the same step, and the J-1
// neighbour is the adjacent lane's.
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

// Raw buffer stores (resource word 3 of the gfx9 family): a lane whose
// offset is kDropOffset, beyond the range, has its store dropped, so the
// conditional S store needs no branch.
constexpr int kRawRange = 0x7fffffff;
constexpr int kRawWord3 = 0x00020000;
constexpr unsigned kDropOffset = 0x80000000u;
typedef unsigned int ux2 __attribute__((ext_vector_type(2)));

constexpr int SRC_KC = 16;  // K rows per thread
constexpr int SRC_LB = 8;   // L planes per block

__global__ __launch_bounds__(64 * SRC_LB) void k_source_grav(SourceArgs a, int nbx, int nby, int nblocks)
{
#pragma clang fp contract(off)
    // Blocks are handed to the 8 XCDs in turn (block b to XCD b % 8); tile
    // t = (b % 8) * (nblocks / 8) + b / 8 puts consecutive tiles -- the J
    // neighbours, then the L neighbours, whose rows share cache lines and
    // planes -- on one XCD and so behind one L2.
    const int b = (int)blockIdx.x;
    const int tile = nblocks % 8 == 0 ? (b % 8) * (nblocks / 8) + b / 8 : b;
    const int bx = tile % nbx, by = (tile / nbx) % nby, bz = tile / (nbx * nby);
    const int J = a.jstart + bx * 64 + (int)threadIdx.x;  // Fortran indices
    const int L = 1 + by * SRC_LB + (int)threadIdx.y;
    const int K0 = 2 + bz * SRC_KC;
    if (J > a.jmax1 || L > a.lmax) return;
    const int K1 = min(K0 + SRC_KC - 1, a.kmax + 1);
    const int LM = L == 1 ? a.lmax : L - 1;
    const int LP = L == a.lmax ? 1 : L + 1;
    const size_t hp = (size_t)a.j2 * a.k2, pp = (size_t)a.pj2 * a.pk2;
    const size_t hl = (size_t)(J - 1) + hp * (L - 1);  // this (J, L) at K = 1
    const size_t ql = (size_t)(J - 1) + pp * (L - 1);
    const ptrdiff_t hdp = (ptrdiff_t)hp * (LP - L), hdm = (ptrdiff_t)hp * (L - LM);
    const ptrdiff_t qdp = (ptrdiff_t)pp * (LP - L), qdm = (ptrdiff_t)pp * (L - LM);
    const bool sj = J >= a.jstart + 1;
    const double twodr = 2.0 * a.dr, twodth = 2.0 * a.dtheta;
    // the arrays do not overlap: every load of a row can be issued before
    // its stores
    const double *__restrict__ rr = a.r;
    const double *__restrict__ rho_ = a.rho;
    const double *__restrict__ om_ = a.omega;
    const double *__restrict__ phi_ = a.phi;
    const double *__restrict__ p_ = a.p;
    double *__restrict__ s_ = a.s;
    double *__restrict__ t_ = a.t;
    double *__restrict__ a_ = a.a;

    const double rj = rr[J - 1];
    // S stores through a buffer over this wavefront's plane (a wavefront is
    // 64 J of one plane L); the offset stays below the plane size
    const int Lw = __builtin_amdgcn_readfirstlane(L);
    const __amdgpu_buffer_rsrc_t srs = __builtin_amdgcn_make_buffer_rsrc(s_ + hp * (size_t)(Lw - 1), 0, kRawRange,
                                                                        kRawWord3);
    // Everything one row K reads; the next row's values are loaded before
    // this row's stores are issued, so the loads of consecutive rows overlap
    // (the last row loads itself again instead of a row beyond the run).
    struct Row {
        double rho, phi, p, rhoj, phij, pj, om0, omj, phip, phim, ppl, pml, s0, t0, a0;
    };
    auto load = [&](int K) {
        const size_t h = hl + (size_t)a.j2 * (K - 1), q = ql + (size_t)a.pj2 * (K - 1);
        Row w;
        w.rho = rho_[q]; w.phi = phi_[q]; w.p = p_[h];
        w.rhoj = rho_[q - 1]; w.phij = phi_[q - 1]; w.pj = p_[h - 1];  // in bounds for J = JSTART too
        w.om0 = om_[h]; w.omj = om_[h - 1];
        w.phip = phi_[q + qdp]; w.phim = phi_[q - qdm]; w.ppl = p_[h + hdp]; w.pml = p_[h - hdm];
        w.s0 = s_[h]; w.t0 = t_[h]; w.a0 = a_[h];
        return w;
    };
    // RHO, PHI and P of row K-1, carried from one row to the next
    double rhom = rho_[ql + (size_t)a.pj2 * (K0 - 2)], phm = phi_[ql + (size_t)a.pj2 * (K0 - 2)];
    double pm = p_[hl + (size_t)a.j2 * (K0 - 2)];
    auto step = [&](int K, const Row &c) {
        const size_t h = hl + (size_t)a.j2 * (K - 1);
        {
            const double rs = c.rho + c.rhoj;
            const double om = c.om0 + c.omj;
            const double cen = rj * rs * (om * om) * 0.25 * 0.5;
            const double grav = rs * (c.phi - c.phij) / twodr;
            const double pres = (c.p - c.pj) / a.dr;
            const unsigned vo = sj ? (unsigned)((h - hp * (size_t)(L - 1)) * sizeof(double)) : kDropOffset;
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(ux2, c.s0 + a.delt * (cen - grav - pres)), srs,
                                                  vo, 0, 0);
        }
        {
            const double grav = 0.5 * (c.rho + rhom) * (c.phi - phm) / a.dz;
            const double pres = (c.p - pm) / a.dz;
            t_[h] = c.t0 - a.delt * (grav + pres);
        }
        {
            const double x = c.rho * (c.phip - c.phim) + c.ppl - c.pml;
            a_[h] = c.a0 - a.delt * x / twodth;
        }
        rhom = c.rho; phm = c.phi; pm = c.p;
    };
    // two register sets in turn: the row after next loads while a row is
    // computed and stored
    Row c0 = load(K0);
    for (int K = K0;; K += 2) {
        const Row c1 = load(min(K + 1, K1));
        step(K, c0);
        if (K + 1 > K1) break;
        c0 = load(min(K + 2, K1));
        step(K + 1, c1);
        if (K + 2 > K1) break;
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
    const int nbx = (nj + 63) / 64, nby = (d->lmax + SRC_LB - 1) / SRC_LB, nbz = (d->kmax + SRC_KC - 1) / SRC_KC;
    const int nblocks = nbx * nby * nbz;
    hipLaunchKernelGGL(k_source_grav, dim3(nblocks), dim3(64, SRC_LB), 0, st, x, nbx, nby, nblocks);
    return (int)hipGetLastError();
}

}  // namespace chy
