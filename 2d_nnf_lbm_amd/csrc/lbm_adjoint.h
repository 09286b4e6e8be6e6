// Discrete adjoint of the fused time step for one disk: gradient of a
// particle-velocity loss with respect to the yield stress, the plastic
// viscosity and the particle density, with checkpointed recomputation.
//
// State of step k (the canonical arrays after the call that ends at step k):
// post-stream populations f_k, node rates om_k and the particle record P_k
// (which holds the hydrodynamic force F_k computed from f_k and P_k).  One
// transition k -> k+1 is
//   f*_{k+1}, om_{k+1} = collide(f_k, om_k, phi(P_k), up(P_k))   per node
//   P_{k+1} = update(P_k, F_k, T_k), F_k = sum_nodes fc, T_k = sum (r x fc)
//   f_{k+1} = stream(f*_{k+1})
// With walls at rest the stream and bounce-back move every population to
// exactly one place, so its transpose is a gather through the inverse map.
#pragma once
#include "lbm_adjoint_local.h"
#include "lbm_kernels.h"

namespace lbm {

// particle adjoint slots: X, Y, Vx, Vy, W, Vold_x, Vold_y, Vold_w
constexpr int AX = 0, AY = 1, AVX = 2, AVY = 3, AW = 4, AOX = 5, AOY = 6, AOW = 7, ANP = 8;
// per-node partial sums: X, Y, Vx, Vy, W, tau_y, mu0
constexpr int AR = 7;

// Encoded source indices through the fused pull-stream and bounce-back: with
// walls at rest every streamed value is a copy, so f_k(x) receives exactly
// the id of the population it came from.
__global__ void k_adj_ids(DevP P, double *ids) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    for (int q = 0; q < 9; ++q) ids[q * P.plane + n] = (double)(q * P.nn + n);
}

__global__ void k_adj_srcmap(DevP P, const double *ids, const double *erho, int *src) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    int j = (int)(n / P.W), i = (int)(n - (long long)j * P.W);
    double v[9];
    load_fk64<StF64>(P, ids, erho, n, i, j, v);
    for (int k = 0; k < 9; ++k) src[k * P.plane + n] = (int)v[k];
}

// Transpose of the particle update (F6:299-312) at P_k: from the adjoint of
// P_{k+1} (mu_next) to the adjoint of P_k through the update itself
// (mu_cur), the adjoint of the hydrodynamic force and torque (fbar), and the
// particle-density adjoint (theta[2]).  m = rho_p pi r^2 and I = m r^2 / 2
// follow rho_p.  One thread.
__global__ void k_adj_particle(DevP P, const double *rec, const double *mu_next, double *mu_cur,
                               double *fbar, double *theta) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int m_ = P.maxp;
    const double rho0 = P.rho0;
    const double rhop = rec[PRHOP * m_], mp = rec[PMP * m_], ipm = rec[PIP * m_];
    const double v[3] = {rec[PVX * m_], rec[PVY * m_], rec[PW * m_]};
    const double vo[3] = {rec[PVOX * m_], rec[PVOY * m_], rec[PVOW * m_]};
    const double frc[3] = {rec[PFX * m_], rec[PFY * m_], rec[PTQ * m_]};
    const double fcol[2] = {rec[PFCX * m_], rec[PFCY * m_]};
    // adjoint reaching the new velocity: directly and through X' = X + (V' + V)/2
    const double a[3] = {mu_next[AVX] + 0.5 * mu_next[AX], mu_next[AVY] + 0.5 * mu_next[AY],
                         mu_next[AW]};
    const double mass[3] = {mp, mp, ipm};
    double rb = 0.0;
    for (int c = 0; c < 3; ++c) {
        // V' = V + F/m + (rho0/rho_p)(V - Vold) [+ (1 - rho0/rho_p) g] [+ Fcol/m]
        const double vdir = (c < 2) ? 0.5 * mu_next[c] : 0.0;  // X' = X + (V' + V)/2
        const int vs = (c == 0) ? AVX : (c == 1 ? AVY : AW);
        const int os = (c == 0) ? AOX : (c == 1 ? AOY : AOW);
        mu_cur[vs] = a[c] * (1.0 + rho0 / rhop) + vdir + mu_next[os];
        mu_cur[os] = -a[c] * rho0 / rhop;
        fbar[c] = a[c] / mass[c];
        double d = -frc[c] / (mass[c] * rhop) - rho0 * (v[c] - vo[c]) / (rhop * rhop);
        if (c == 1) d += rho0 * P.gy / (rhop * rhop);
        if (c < 2) d -= fcol[c] / (mass[c] * rhop);
        rb += a[c] * d;
    }
    mu_cur[AX] = mu_next[AX];
    mu_cur[AY] = mu_next[AY];
    theta[2] += rb;
}

// Per-node adjoint of transition k.  lf_next, lw_next: adjoints of f_{k+1}
// (9 planes) and om_{k+1}; dst: inverse stream map; fk, omk: state k; rec:
// P_k; fbar: adjoints of F_k and T_k.  Writes the adjoints of f_k and om_k
// and one row of AR partial sums per block.
__global__ __launch_bounds__(NT) void k_adj_node(DevP P, adj::LocalP L, const double *lf_next,
                                                 const double *lw_next, const int *dst,
                                                 const double *fk, const double *omk,
                                                 const double *rec, const double *fbar,
                                                 double *lf_cur, double *lw_cur,
                                                 double *partial) {
    __shared__ double lds[AR * NT];
    const long long n = (long long)blockIdx.x * NT + threadIdx.x;
    double acc[AR] = {0, 0, 0, 0, 0, 0, 0};
    if (n < P.nn) {
        const int m_ = P.maxp;
        const int j = (int)(n / P.W), i = (int)(n - (long long)j * P.W);
        const double xc = rec[PX * m_], yc = rec[PY * m_];
        const double vx = rec[PVX * m_], vy = rec[PVY * m_], w = rec[PW * m_];
        const double dm = rec[PDIAM * m_], zt = rec[PZETA * m_];
        const double dx = (double)i - xc, dy = (double)j - yc;
        const double phi = phi_at(xc, yc, dm, zt, i, j);
        const double upx = vx - w * dy, upy = vy + w * dx;
        double f[9], lfs[9];
        for (int q = 0; q < 9; ++q) {
            f[q] = fk[q * P.plane + n];
            lfs[q] = lf_next[dst[q * P.plane + n]];
        }
        const double om = P.rheology != 0 ? omk[n] : P.omega;
        adj::Tape t;
        double fs[9], om2, fcx, fcy;
        adj::node_forward(L, f, om, phi, upx, upy, fs, om2, fcx, fcy, t);
        // force and torque adjoints seen by this node's contribution
        const double lfx = fbar[0] - fbar[2] * dy, lfy = fbar[1] + fbar[2] * dx;
        const double lw = P.rheology != 0 ? lw_next[n] : 0.0;
        double fb[9], omb, phib, upxb, upyb, tyb = 0.0, mub = 0.0;
        adj::node_backward(L, t, lfs, lw, lfx, lfy, fb, omb, phib, upxb, upyb, tyb, mub);
        for (int q = 0; q < 9; ++q) lf_cur[q * P.plane + n] = fb[q];
        if (P.rheology != 0) lw_cur[n] = omb;
        else mub += omb * (-3.0 * P.omega * P.omega);  // OMEGA = 1/(3 mu0 + 1/2)
        // phi(d), d = r - |x - X|: dphi/dX = phi'(d) dx / dist
        const double dist = sqrt(dx * dx + dy * dy);
        const double di = 0.5 * dm - dist;
        double dphi = 0.0;
        if (fabs(di) < 0.5 * zt && dist > 0.0)
            dphi = 0.5 * cos(PI_F * di / zt) * (PI_F / zt);
        const double gx = dist > 0.0 ? dphi * dx / dist : 0.0;
        const double gy = dist > 0.0 ? dphi * dy / dist : 0.0;
        // up = (Vx - W dy, Vy + W dx); torque T = sum dx fcy - dy fcx
        acc[0] = phib * gx - upyb * w - fbar[2] * fcy;
        acc[1] = phib * gy + upxb * w + fbar[2] * fcx;
        acc[2] = upxb;
        acc[3] = upyb;
        acc[4] = -upxb * dy + upyb * dx;
        acc[5] = tyb;
        acc[6] = mub;
    }
    block_sum<AR>(acc, lds);
    if (threadIdx.x == 0)
        for (int c = 0; c < AR; ++c) partial[(size_t)blockIdx.x * AR + c] = acc[c];
}

// Ordered sum of the block partials into the particle adjoint of step k and
// the parameter adjoints; adds the loss weight to the Vy slot.  One block.
__global__ __launch_bounds__(NT) void k_adj_reduce(const double *partial, int nb, double *mu_cur,
                                                   double *theta, double loss_w) {
    __shared__ double lds[AR * NT];
    double acc[AR] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += NT)
        for (int c = 0; c < AR; ++c) acc[c] += partial[(size_t)b * AR + c];
    block_sum<AR>(acc, lds);
    if (threadIdx.x == 0) {
        mu_cur[AX] += acc[0];
        mu_cur[AY] += acc[1];
        mu_cur[AVX] += acc[2];
        mu_cur[AVY] += acc[3] + loss_w;
        mu_cur[AW] += acc[4];
        theta[0] += acc[5];
        theta[1] += acc[6];
    }
}

// The initial node rates are OMEGA = 1/(3 mu0 + 1/2) (F6:101-102): their
// adjoint adds sum(lw) * dOMEGA/dmu0 to the mu0 adjoint.  One block.
__global__ __launch_bounds__(NT) void k_adj_omega0(DevP P, const double *lw, double *theta) {
    __shared__ double lds[NT];
    double acc[1] = {0.0};
    for (long long n = threadIdx.x; n < P.nn; n += NT) acc[0] += lw[n];
    block_sum<1>(acc, lds);
    if (threadIdx.x == 0) theta[1] += acc[0] * (-3.0 * P.omega * P.omega);
}

}  // namespace lbm
