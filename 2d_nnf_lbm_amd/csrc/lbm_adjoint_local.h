// Hand-written reverse-mode derivative of the per-node part of one time step
// (macroscopic fields, smoothed-profile forcing and velocity correction, the
// hydrodynamic-force contribution, and the MRT / Papanastasiou collision),
// in FP64.  Plain C++ with host/device qualifiers, so the same code runs in
// the adjoint kernels and in a host test against finite differences.
//
// Forward, for one node with post-stream populations f and rate om:
//   rho = sum f, u = (1/rho) j                       (F6:1405-1408)
//   G = phi (up - u), uc = u + G/2                   (F6:485-495, F6:1111-1125)
//   fc = rho phi (u - up)                            (F6:544-553, per node)
//   m = M f, meq(rho, uc), mf(rho, G, uc)            (F6:878-886, F6:964-989)
//   m* = s meq + (1-s) m + (1-s/2) mf, f* = M^-1 m*  (F6:890-900)
//   Bingham: strain rate from s (m - meq) (F6:903-916), Papanastasiou
//   viscosity (F6:939-945), om' = 1/(3 nu/rho + 1/2) (F6:956-957).
// The rate vector s depends on om (s7 = s8 = om; s4 = s6 = 8(2-om)/(8-om)
// when q_mode is set, F6:1009-1011).  Differentiable inputs: f, om, phi, up,
// tau_y, mu0.
#pragma once
#include <math.h>

#ifndef LBM_HD
#if defined(__HIPCC__)
#define LBM_HD __host__ __device__
#else
#define LBM_HD
#endif
#endif

namespace lbm {
namespace adj {

struct LocalP {
    int rheology, q_mode;
    double s0, s1, s2, s3, s5, sq;
    double mu0, tau_y, m_pap, gamma_min;
    double d_mrt;  // fl32(1/36) widened (F6:873)
};

// Intermediates of the forward pass that the backward pass reads.
struct Tape {
    double f[9], om, phi, upx, upy;
    double rho, r, jx, jy, ux, uy, gx, gy, ucx, ucy;
    double s[9], m[9], meq[9], mf[9], mne[9];
    double fx, fy;
    // Bingham
    double e[9], d11, d22, d12, ds11, ds22, ds12, gam, ex, nu_a, om2;
    int small_gamma;
};

// M f: rows of the moment transform (F6:878-886).
LBM_HD inline void moments(const double f[9], double m[9]) {
    m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    m[1] = -(4.0 * f[0]) - f[1] - f[2] - f[3] - f[4] + 2.0 * (f[5] + f[6] + f[7] + f[8]);
    m[2] = 4.0 * f[0] - 2.0 * (f[1] + f[2] + f[3] + f[4]) + (f[5] + f[6] + f[7] + f[8]);
    m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    m[4] = -(2.0 * (f[1] - f[3])) + f[5] - f[6] - f[7] + f[8];
    m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    m[6] = -(2.0 * (f[2] - f[4])) + f[5] + f[6] - f[7] - f[8];
    m[7] = f[1] - f[2] + f[3] - f[4];
    m[8] = f[5] - f[6] + f[7] - f[8];
}

// Transpose of M applied to a moment adjoint: fbar += M^T mbar.
LBM_HD inline void moments_T(const double mb[9], double fb[9]) {
    fb[0] += mb[0] - 4.0 * mb[1] + 4.0 * mb[2];
    fb[1] += mb[0] - mb[1] - 2.0 * mb[2] + mb[3] - 2.0 * mb[4] + mb[7];
    fb[2] += mb[0] - mb[1] - 2.0 * mb[2] + mb[5] - 2.0 * mb[6] - mb[7];
    fb[3] += mb[0] - mb[1] - 2.0 * mb[2] - mb[3] + 2.0 * mb[4] + mb[7];
    fb[4] += mb[0] - mb[1] - 2.0 * mb[2] - mb[5] + 2.0 * mb[6] - mb[7];
    fb[5] += mb[0] + 2.0 * mb[1] + mb[2] + mb[3] + mb[4] + mb[5] + mb[6] + mb[8];
    fb[6] += mb[0] + 2.0 * mb[1] + mb[2] - mb[3] - mb[4] + mb[5] + mb[6] - mb[8];
    fb[7] += mb[0] + 2.0 * mb[1] + mb[2] - mb[3] - mb[4] - mb[5] - mb[6] + mb[8];
    fb[8] += mb[0] + 2.0 * mb[1] + mb[2] + mb[3] + mb[4] - mb[5] - mb[6] - mb[8];
}

// Coefficients of the inverse transform f = d * C m (F6:892-900), and of the
// strain moments f1 = C (s * mne) (F6:903-911), which use the same matrix.
LBM_HD inline double cinv(int q, int k) {
    const double C[9][9] = {
        {4, -4, 4, 0, 0, 0, 0, 0, 0},
        {4, -1, -2, 6, -6, 0, 0, 9, 0},
        {4, -1, -2, 0, 0, 6, -6, -9, 0},
        {4, -1, -2, -6, 6, 0, 0, 9, 0},
        {4, -1, -2, 0, 0, -6, 6, -9, 0},
        {4, 2, 1, 6, 3, 6, 3, 0, 9},
        {4, 2, 1, -6, -3, 6, 3, 0, -9},
        {4, 2, 1, -6, -3, -6, -3, 0, 9},
        {4, 2, 1, 6, 3, -6, -3, 0, -9}};
    return C[q][k];
}

LBM_HD inline void rates(const LocalP &P, double om, double s[9]) {
    double q = P.q_mode ? (8.0 * (2.0 - om)) / (8.0 - om) : P.sq;
    s[0] = P.s0; s[1] = P.s1; s[2] = P.s2; s[3] = P.s3; s[4] = q;
    s[5] = P.s5; s[6] = q; s[7] = om; s[8] = om;
}

// Forward pass.  Outputs: fs (post-collision), om2 (new node rate; om for
// Newtonian runs), fcx, fcy (hydrodynamic-force contribution of the node).
LBM_HD inline void node_forward(const LocalP &P, const double f[9], double om, double phi,
                                double upx, double upy, double fs[9], double &om2,
                                double &fcx, double &fcy, Tape &t) {
    for (int k = 0; k < 9; ++k) t.f[k] = f[k];
    t.om = om; t.phi = phi; t.upx = upx; t.upy = upy;
    t.rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    t.r = 1.0 / t.rho;
    t.jx = f[1] + f[5] + f[8] - (f[6] + f[3] + f[7]);
    t.jy = f[5] + f[2] + f[6] - (f[7] + f[4] + f[8]);
    t.ux = t.r * t.jx;
    t.uy = t.r * t.jy;
    t.gx = phi * (upx - t.ux);
    t.gy = phi * (upy - t.uy);
    t.ucx = t.ux + 0.5 * t.gx;
    t.ucy = t.uy + 0.5 * t.gy;
    fcx = t.rho * phi * (t.ux - upx);
    fcy = t.rho * phi * (t.uy - upy);
    rates(P, om, t.s);
    moments(f, t.m);
    const double rho = t.rho, ux = t.ucx, uy = t.ucy;
    t.fx = rho * t.gx;
    t.fy = rho * t.gy;
    t.mf[0] = 0.0;
    t.mf[1] = 6.0 * (t.fx * ux + t.fy * uy);
    t.mf[2] = -(6.0 * (t.fx * ux + t.fy * uy));
    t.mf[3] = t.fx; t.mf[4] = -t.fx; t.mf[5] = t.fy; t.mf[6] = -t.fy;
    t.mf[7] = 2.0 * (t.fx * ux - t.fy * uy);
    t.mf[8] = t.fy * ux + t.fx * uy;
    const double u2 = ux * ux + uy * uy;
    t.meq[0] = rho;
    t.meq[1] = rho * (-2.0 + 3.0 * u2);
    t.meq[2] = rho * (1.0 - 3.0 * u2);
    t.meq[3] = rho * ux; t.meq[4] = -(rho * ux);
    t.meq[5] = rho * uy; t.meq[6] = -(rho * uy);
    t.meq[7] = rho * (ux * ux - uy * uy);
    t.meq[8] = rho * (ux * uy);
    double ms[9];
    for (int k = 0; k < 9; ++k) {
        t.mne[k] = t.m[k] - t.meq[k];
        ms[k] = t.s[k] * t.meq[k] + (1.0 - t.s[k]) * t.m[k] + (1.0 - 0.5 * t.s[k]) * t.mf[k];
    }
    for (int q = 0; q < 9; ++q) {
        double a = 0.0;
        for (int k = 0; k < 9; ++k) a += cinv(q, k) * ms[k];
        fs[q] = P.d_mrt * a;
    }
    om2 = om;
    t.om2 = om;
    if (P.rheology != 1) return;
    // Chai strain rate (F6:903-924) and the Papanastasiou viscosity
    for (int k = 0; k < 9; ++k) t.e[k] = t.s[k] * t.mne[k];
    double f1[9];
    for (int q = 1; q < 9; ++q) {
        double a = 0.0;
        for (int k = 0; k < 9; ++k) a += cinv(q, k) * t.e[k];
        f1[q] = a;
    }
    const double z = -1.5 * P.d_mrt;
    const double gu = t.gx * ux + t.gy * uy;
    t.d11 = z * (f1[1] + f1[5] + f1[6] + f1[3] + f1[7] + f1[8]) +
            0.75 * ((t.s[7] - t.s[1]) * gu - t.s[7] * (2.0 * t.gx * ux));
    t.d22 = z * (f1[5] + f1[2] + f1[6] + f1[7] + f1[4] + f1[8]) +
            0.75 * ((t.s[7] - t.s[1]) * gu - t.s[7] * (2.0 * t.gy * uy));
    t.d12 = z * (f1[5] - f1[6] + f1[7] - f1[8]) + 0.75 * (-(t.s[7] * (t.gx * uy + t.gy * ux)));
    t.ds11 = t.d11 * t.r; t.ds22 = t.d22 * t.r; t.ds12 = t.d12 * t.r;
    t.gam = sqrt(2.0 * fabs(t.ds11 * t.ds11 + t.ds22 * t.ds22 + 2.0 * (t.ds12 * t.ds12)));
    t.small_gamma = t.gam < P.gamma_min;
    if (t.small_gamma) {
        t.ex = 0.0;
        t.nu_a = P.mu0 + P.m_pap * P.tau_y;
    } else {
        t.ex = exp(-(P.m_pap * t.gam));
        t.nu_a = P.mu0 + (P.tau_y / t.gam) * (1.0 - t.ex);
    }
    t.om2 = 1.0 / (3.0 * (t.nu_a * t.r) + 0.5);
    om2 = t.om2;
}

// Backward pass: given the adjoints of the outputs (fsb of fs, om2b of om2,
// fcxb, fcyb), returns the adjoints of the inputs: fb (9), omb, phib, upxb,
// upyb, and adds the parameter adjoints to *tyb (tau_y) and *mub (mu0).
LBM_HD inline void node_backward(const LocalP &P, const Tape &t, const double fsb[9],
                                 double om2b, double fcxb, double fcyb, double fb[9],
                                 double &omb, double &phib, double &upxb, double &upyb,
                                 double &tyb, double &mub) {
    double mb[9] = {0}, meqb[9] = {0}, mfb[9] = {0}, sb[9] = {0}, mneb[9] = {0};
    double rhob = 0.0, rb = 0.0, ucxb = 0.0, ucyb = 0.0, gxb = 0.0, gyb = 0.0;
    double uxb = 0.0, uyb = 0.0;
    phib = 0.0; upxb = 0.0; upyb = 0.0;
    const double ux = t.ucx, uy = t.ucy, rho = t.rho;
    // ---- Bingham rate
    if (P.rheology == 1 && om2b != 0.0) {
        // om2 = 1 / (3 nu r + 1/2)
        const double w2 = t.om2 * t.om2;
        double nub = om2b * (-w2 * 3.0 * t.r);
        rb += om2b * (-w2 * 3.0 * t.nu_a);
        double gamb = 0.0;
        mub += nub;
        if (t.small_gamma) {
            tyb += nub * P.m_pap;
        } else {
            tyb += nub * (1.0 - t.ex) / t.gam;
            gamb = nub * P.tau_y * (-(1.0 - t.ex) / (t.gam * t.gam) + P.m_pap * t.ex / t.gam);
        }
        // gam = sqrt(2 Q), Q = ds11^2 + ds22^2 + 2 ds12^2  => dgam/dQ = 1/gam
        double ds11b = 0.0, ds22b = 0.0, ds12b = 0.0;
        if (t.gam > 0.0) {
            const double qb = gamb / t.gam;
            ds11b = qb * 2.0 * t.ds11;
            ds22b = qb * 2.0 * t.ds22;
            ds12b = qb * 4.0 * t.ds12;
        }
        double d11b = ds11b * t.r, d22b = ds22b * t.r, d12b = ds12b * t.r;
        rb += ds11b * t.d11 + ds22b * t.d22 + ds12b * t.d12;
        // d11, d22, d12 from f1 (strain moments), s7, s1, G and uc
        const double z = -1.5 * P.d_mrt;
        double f1b[9] = {0};
        f1b[1] += z * d11b; f1b[5] += z * d11b; f1b[6] += z * d11b;
        f1b[3] += z * d11b; f1b[7] += z * d11b; f1b[8] += z * d11b;
        f1b[5] += z * d22b; f1b[2] += z * d22b; f1b[6] += z * d22b;
        f1b[7] += z * d22b; f1b[4] += z * d22b; f1b[8] += z * d22b;
        f1b[5] += z * d12b; f1b[6] -= z * d12b; f1b[7] += z * d12b; f1b[8] -= z * d12b;
        const double gu = t.gx * ux + t.gy * uy;
        const double ds71 = t.s[7] - t.s[1];
        // d11 extra: 0.75 (ds71 gu - s7 2 gx ux)
        double gub = 0.75 * ds71 * (d11b + d22b);
        sb[7] += 0.75 * gu * (d11b + d22b);
        sb[1] -= 0.75 * gu * (d11b + d22b);
        sb[7] += -0.75 * (2.0 * t.gx * ux) * d11b - 0.75 * (2.0 * t.gy * uy) * d22b;
        gxb += -0.75 * t.s[7] * 2.0 * ux * d11b;
        ucxb += -0.75 * t.s[7] * 2.0 * t.gx * d11b;
        gyb += -0.75 * t.s[7] * 2.0 * uy * d22b;
        ucyb += -0.75 * t.s[7] * 2.0 * t.gy * d22b;
        // d12 extra: -0.75 s7 (gx uy + gy ux)
        sb[7] += -0.75 * (t.gx * uy + t.gy * ux) * d12b;
        gxb += -0.75 * t.s[7] * uy * d12b;
        ucyb += -0.75 * t.s[7] * t.gx * d12b;
        gyb += -0.75 * t.s[7] * ux * d12b;
        ucxb += -0.75 * t.s[7] * t.gy * d12b;
        // gu = gx ux + gy uy
        gxb += gub * ux; ucxb += gub * t.gx;
        gyb += gub * uy; ucyb += gub * t.gy;
        // f1[q] = sum_k C[q][k] e_k, e_k = s_k mne_k
        for (int k = 0; k < 9; ++k) {
            double eb = 0.0;
            for (int q = 1; q < 9; ++q) eb += cinv(q, k) * f1b[q];
            sb[k] += eb * t.mne[k];
            mneb[k] += eb * t.s[k];
        }
    }
    // ---- relaxation and inverse transform
    for (int k = 0; k < 9; ++k) {
        double msb = 0.0;
        for (int q = 0; q < 9; ++q) msb += cinv(q, k) * fsb[q];
        msb *= P.d_mrt;
        mb[k] += (1.0 - t.s[k]) * msb;
        meqb[k] += t.s[k] * msb;
        mfb[k] += (1.0 - 0.5 * t.s[k]) * msb;
        sb[k] += msb * (t.meq[k] - t.m[k] - 0.5 * t.mf[k]);
    }
    for (int k = 0; k < 9; ++k) {
        mb[k] += mneb[k];
        meqb[k] -= mneb[k];
    }
    // ---- equilibrium moments meq(rho, uc)
    const double u2 = ux * ux + uy * uy;
    rhob += meqb[0] + meqb[1] * (-2.0 + 3.0 * u2) + meqb[2] * (1.0 - 3.0 * u2) +
            (meqb[3] - meqb[4]) * ux + (meqb[5] - meqb[6]) * uy +
            meqb[7] * (ux * ux - uy * uy) + meqb[8] * (ux * uy);
    const double u2b = (meqb[1] - meqb[2]) * 3.0 * rho;
    ucxb += u2b * 2.0 * ux + (meqb[3] - meqb[4]) * rho + meqb[7] * rho * 2.0 * ux +
            meqb[8] * rho * uy;
    ucyb += u2b * 2.0 * uy + (meqb[5] - meqb[6]) * rho - meqb[7] * rho * 2.0 * uy +
            meqb[8] * rho * ux;
    // ---- forcing moments mf(fx, fy, uc)
    const double a12 = 6.0 * (mfb[1] - mfb[2]);
    double fxb = a12 * ux + (mfb[3] - mfb[4]) + mfb[7] * 2.0 * ux + mfb[8] * uy;
    double fyb = a12 * uy + (mfb[5] - mfb[6]) - mfb[7] * 2.0 * uy + mfb[8] * ux;
    ucxb += a12 * t.fx + mfb[7] * 2.0 * t.fx + mfb[8] * t.fy;
    ucyb += a12 * t.fy - mfb[7] * 2.0 * t.fy + mfb[8] * t.fx;
    rhob += fxb * t.gx + fyb * t.gy;
    gxb += fxb * rho;
    gyb += fyb * rho;
    // ---- rates s(om); Newtonian runs pass om through as om2
    omb = sb[7] + sb[8] + (P.rheology == 1 ? 0.0 : om2b);
    if (P.q_mode) {
        const double dq = -48.0 / ((8.0 - t.om) * (8.0 - t.om));
        omb += (sb[4] + sb[6]) * dq;
    }
    // ---- m = M f
    for (int k = 0; k < 9; ++k) fb[k] = 0.0;
    moments_T(mb, fb);
    // ---- uc = u + G/2; G = phi (up - u); fc = rho phi (u - up)
    uxb += ucxb; uyb += ucyb;
    gxb += 0.5 * ucxb; gyb += 0.5 * ucyb;
    phib += gxb * (t.upx - t.ux) + gyb * (t.upy - t.uy);
    upxb += t.phi * gxb; upyb += t.phi * gyb;
    uxb -= t.phi * gxb; uyb -= t.phi * gyb;
    rhob += fcxb * t.phi * (t.ux - t.upx) + fcyb * t.phi * (t.uy - t.upy);
    phib += fcxb * rho * (t.ux - t.upx) + fcyb * rho * (t.uy - t.upy);
    uxb += fcxb * rho * t.phi; uyb += fcyb * rho * t.phi;
    upxb -= fcxb * rho * t.phi; upyb -= fcyb * rho * t.phi;
    // ---- u = r j, r = 1/rho, rho = sum f
    rb += uxb * t.jx + uyb * t.jy;
    const double jxb = uxb * t.r, jyb = uyb * t.r;
    rhob += -rb * t.r * t.r;
    for (int k = 0; k < 9; ++k) fb[k] += rhob;
    fb[1] += jxb; fb[5] += jxb; fb[8] += jxb; fb[6] -= jxb; fb[3] -= jxb; fb[7] -= jxb;
    fb[5] += jyb; fb[2] += jyb; fb[6] += jyb; fb[7] -= jyb; fb[4] -= jyb; fb[8] -= jyb;
}

}  // namespace adj
}  // namespace lbm
