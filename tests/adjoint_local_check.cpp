// Host check of the hand-written per-node adjoint (lbm_adjoint_local.h)
// against central finite differences of its own forward pass.  Prints one
// line per case: the largest relative mismatch over all 15 input adjoints.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../2d_nnf_lbm_amd/csrc/lbm_adjoint_local.h"

using namespace lbm::adj;

static double urand(unsigned long long &s) {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
}

struct In {
    double f[9], om, phi, upx, upy;
};

// scalar objective: lambda . outputs
static double objective(const LocalP &P, const In &x, const double lf[9], double lom,
                        double lfx, double lfy) {
    double fs[9], om2, fcx, fcy;
    Tape t;
    node_forward(P, x.f, x.om, x.phi, x.upx, x.upy, fs, om2, fcx, fcy, t);
    double v = lom * om2 + lfx * fcx + lfy * fcy;
    for (int q = 0; q < 9; ++q) v += lf[q] * fs[q];
    return v;
}

int main(int argc, char **argv) {
    int ncases = argc > 1 ? atoi(argv[1]) : 24;
    const double hstate = argc > 2 ? atof(argv[2]) : 1e-8;
    unsigned long long seed = 12345;
    const double w[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                         1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
    const int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1}, cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    double worst = 0.0;
    for (int c = 0; c < ncases; ++c) {
        LocalP P;
        P.rheology = (c % 4 == 3) ? 0 : 1;
        P.q_mode = (c % 3 == 2) ? 0 : 1;
        P.s0 = P.q_mode ? 0.0 : 1.0; P.s1 = P.q_mode ? 1.1 : 1.4; P.s2 = P.q_mode ? 1.25 : 1.4;
        P.s3 = P.q_mode ? 0.0 : 1.0; P.s5 = P.q_mode ? 0.0 : 1.0; P.sq = 1.2;
        P.mu0 = 0.1 * (1.0 + 0.2 * urand(seed));
        P.tau_y = 2e-6 * (1.0 + urand(seed));
        P.m_pap = 1e5;
        P.gamma_min = 1e-15;
        P.d_mrt = (double)(1.0f / 36.0f);
        In x;
        // near-equilibrium populations with a shear perturbation of varied size
        double ux = 0.01 * (urand(seed) - 0.5), uy = 0.01 * (urand(seed) - 0.5);
        double amp = pow(10.0, -3.0 - 3.0 * urand(seed));
        for (int q = 0; q < 9; ++q) {
            double cu = cx[q] * ux + cy[q] * uy;
            x.f[q] = w[q] * (1.0 + 3 * cu + 4.5 * cu * cu - 1.5 * (ux * ux + uy * uy)) +
                     amp * (urand(seed) - 0.5) * w[q];
        }
        x.om = 1.0 / (3.0 * P.mu0 + 0.5) * (0.8 + 0.4 * urand(seed));
        x.phi = (c % 5 == 0) ? 0.0 : urand(seed);
        x.upx = 0.005 * (urand(seed) - 0.5);
        x.upy = 0.005 * (urand(seed) - 0.5);
        double lf[9], lom = urand(seed) - 0.5, lfx = urand(seed) - 0.5, lfy = urand(seed) - 0.5;
        for (int q = 0; q < 9; ++q) lf[q] = urand(seed) - 0.5;
        // analytic
        double fs[9], om2, fcx, fcy;
        Tape t;
        node_forward(P, x.f, x.om, x.phi, x.upx, x.upy, fs, om2, fcx, fcy, t);
        double fb[9], omb, phib, upxb, upyb, tyb = 0.0, mub = 0.0;
        node_backward(P, t, lf, lom, lfx, lfy, fb, omb, phib, upxb, upyb, tyb, mub);
        double ana[15];
        for (int q = 0; q < 9; ++q) ana[q] = fb[q];
        ana[9] = omb; ana[10] = phib; ana[11] = upxb; ana[12] = upyb; ana[13] = tyb; ana[14] = mub;
        // finite differences
        double scale = 0.0, err = 0.0;
        for (int i = 0; i < 15; ++i) {
            // fourth-order central difference; state inputs step by an
            // absolute hstate (the strain rate resolves 1/m_pap), the two
            // parameters by 1e-4 of their value
            double F[4];
            const double off[4] = {2.0, 1.0, -1.0, -2.0};
            double h = 0.0;
            for (int a = 0; a < 4; ++a) {
                In xp = x;
                LocalP Pp = P;
                double *v;
                if (i < 9) v = &xp.f[i];
                else if (i == 9) v = &xp.om;
                else if (i == 10) v = &xp.phi;
                else if (i == 11) v = &xp.upx;
                else if (i == 12) v = &xp.upy;
                else if (i == 13) v = &Pp.tau_y;
                else v = &Pp.mu0;
                h = (i >= 13) ? 1e-4 * fabs(*v) : hstate;
                *v += off[a] * h;
                F[a] = objective(Pp, xp, lf, lom, lfx, lfy);
            }
            double fd = (-F[0] + 8.0 * F[1] - 8.0 * F[2] + F[3]) / (12.0 * h);
            scale = fmax(scale, fabs(fd));
            err = fmax(err, fabs(fd - ana[i]));
        }
        double rel = err / (scale > 0 ? scale : 1.0);
        worst = fmax(worst, rel);
        printf("case %d rheology %d q_mode %d gamma %.3e rel_err %.3e\n", c, P.rheology, P.q_mode,
               t.gam, rel);
    }
    printf("worst %.3e\n", worst);
    return 0;
}
