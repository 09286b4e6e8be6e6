/*
 * CPU oracle: serial restatement of the reference time step.  TEST
 * INFRASTRUCTURE ONLY (see lbm_oracle.h).  Each function names the reference
 * routine and lines it follows.  Build with -O2 -ffp-contract=off.
 */
#include "lbm_oracle.h"
#include <math.h>
#include <stddef.h>
#include <string.h>

#define F32(x) ((double)(float)(x))
#define NODES(p) ((size_t)((p)->nx + 1) * (size_t)((p)->ny + 1))
#define NID(p, i, j) ((size_t)(j) * (size_t)((p)->nx + 1) + (size_t)(i))
#define PF(st, fld, ip) ((st)->part[(size_t)(fld) * (size_t)(st)->np + (size_t)(ip)])

static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};  /* F6:13 */
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};  /* F6:14 */

/* COMPUTEFEQNEW, F6:1130-1148. */
void orc_feq(double ux, double uy, double rho, double *feq)
{
    const double w0 = F32(2.0f / 9.0f), w1 = F32(1.0f / 18.0f), w2 = F32(1.0f / 36.0f);
    double u2 = ux * ux + uy * uy;
    feq[0] = w0 * rho * (2.0 - 3.0 * u2);
    feq[1] = w1 * rho * (2.0 + 6.0 * ux + 9.0 * ux * ux - 3.0 * u2);
    feq[2] = w1 * rho * (2.0 + 6.0 * uy + 9.0 * uy * uy - 3.0 * u2);
    feq[3] = w1 * rho * (2.0 - 6.0 * ux + 9.0 * ux * ux - 3.0 * u2);
    feq[4] = w1 * rho * (2.0 - 6.0 * uy + 9.0 * uy * uy - 3.0 * u2);
    feq[5] = w2 * rho * (1.0 + 3.0 * (ux + uy) + 9.0 * ux * uy + 3.0 * u2);
    feq[6] = w2 * rho * (1.0 - 3.0 * (ux - uy) - 9.0 * ux * uy + 3.0 * u2);
    feq[7] = w2 * rho * (1.0 - 3.0 * (ux + uy) + 9.0 * ux * uy + 3.0 * u2);
    feq[8] = w2 * rho * (1.0 + 3.0 * (ux - uy) - 9.0 * ux * uy + 3.0 * u2);
}

/* ComputeRelaxation_MRT, F6:1002-1011: the rate vector for a given omega. */
void orc_relaxation(const orc_params *p, double omega, double *s)
{
    double q = p->q_mode ? (8.0 * (2.0 - omega)) / (8.0 - omega) : p->sq;
    s[0] = p->s0; s[1] = p->s1; s[2] = p->s2; s[3] = p->s3; s[4] = q;
    s[5] = p->s5; s[6] = q; s[7] = omega; s[8] = omega;
}

/* INITIALIZE, F6:563-580 (plus OMEGA_BN=OMEGA, F6:102, and the initial
 * ComputeRelaxation_MRT(OMEGA), F6:199). */
void orc_initialize(const orc_params *p, orc_state *st)
{
    size_t n = NODES(p);
    for (size_t id = 0; id < n; ++id) {
        st->rho[id] = p->rho0;
        st->ux[id] = 0.0;
        st->uy[id] = 0.0;
        orc_feq(st->ux[id], st->uy[id], st->rho[id], st->f + 9 * id);
        st->g[2 * id] = 0.0;
        st->g[2 * id + 1] = 0.0;
        st->omega_bn[id] = p->omega;
    }
    orc_relaxation(p, p->omega, st->s_mrt);
}

/* ComputeSourceEquilib_MRT, F6:964-989 (G_X = G_Y = 0, F6:65-66). */
static void source_equilib(double ux, double uy, double rho, double gx, double gy,
                           double *mf, double *meq)
{
    double fx = rho * (gx + 0.0);
    double fy = rho * (gy + 0.0);
    mf[0] = 0.0;
    mf[1] = 6.0 * (fx * ux + fy * uy);
    mf[2] = -(6.0 * (fx * ux + fy * uy));
    mf[3] = fx;
    mf[4] = -fx;
    mf[5] = fy;
    mf[6] = -fy;
    mf[7] = 2.0 * (fx * ux - fy * uy);
    mf[8] = fy * ux + fx * uy;
    double u2 = ux * ux + uy * uy;
    meq[0] = rho;
    meq[1] = rho * (-2.0 + 3.0 * u2);
    meq[2] = rho * (1.0 - 3.0 * u2);
    meq[3] = rho * (ux);
    meq[4] = rho * (-ux);
    meq[5] = rho * (uy);
    meq[6] = rho * (-uy);
    meq[7] = rho * (ux * ux - uy * uy);
    meq[8] = rho * (ux * uy);
}

/* Collision_MRT, F6:867-962.  rate_mode 0 keeps the module-global rate that
 * the previous node set (F6:958); rate_mode 1 first loads the node's own
 * previous-step rate (the one-line oracle patch at the top of the node loop). */
void orc_collide(const orc_params *p, orc_state *st)
{
    const double d = F32(1.0f / 36.0f);          /* d_mrt, F6:873 */
    double *s = st->s_mrt;
    for (int j = 0; j <= p->ny; ++j) {
        for (int i = 0; i <= p->nx; ++i) {
            size_t id = NID(p, i, j);
            double *f = st->f + 9 * id;
            if (p->rate_mode == 1) orc_relaxation(p, st->omega_bn[id], s);
            double m[9], mf[9], meq[9], mne[9], f1[9];
            m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
            m[1] = -(4.0 * f[0]) - f[1] - f[2] - f[3] - f[4] + 2.0 * (f[5] + f[6] + f[7] + f[8]);
            m[2] = 4.0 * f[0] - 2.0 * (f[1] + f[2] + f[3] + f[4]) + (f[5] + f[6] + f[7] + f[8]);
            m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
            m[4] = -(2.0 * (f[1] - f[3])) + f[5] - f[6] - f[7] + f[8];
            m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
            m[6] = -(2.0 * (f[2] - f[4])) + f[5] + f[6] - f[7] - f[8];
            m[7] = f[1] - f[2] + f[3] - f[4];
            m[8] = f[5] - f[6] + f[7] - f[8];
            double ux = st->ux[id], uy = st->uy[id], rho = st->rho[id];
            double gx = st->g[2 * id], gy = st->g[2 * id + 1];
            source_equilib(ux, uy, rho, gx, gy, mf, meq);
            for (int k = 0; k < 9; ++k) mne[k] = m[k] - meq[k];
            for (int k = 0; k < 9; ++k)
                m[k] = s[k] * meq[k] + (1.0 - s[k]) * m[k] + (1.0 - 0.5 * s[k]) * mf[k];
            f[0] = d * (4.0 * (m[0] - m[1] + m[2]));
            f[1] = d * (4.0 * m[0] - m[1] - 2.0 * m[2] + 6.0 * m[3] - 6.0 * m[4] + 9.0 * m[7]);
            f[2] = d * (4.0 * m[0] - m[1] - 2.0 * m[2] + 6.0 * m[5] - 6.0 * m[6] - 9.0 * m[7]);
            f[3] = d * (4.0 * m[0] - m[1] - 2.0 * m[2] - 6.0 * m[3] + 6.0 * m[4] + 9.0 * m[7]);
            f[4] = d * (4.0 * m[0] - m[1] - 2.0 * m[2] - 6.0 * m[5] + 6.0 * m[6] - 9.0 * m[7]);
            f[5] = d * (4.0 * m[0] + 2.0 * m[1] + m[2] + 6.0 * m[3] + 3.0 * m[4] + 6.0 * m[5] + 3.0 * m[6] + 9.0 * m[8]);
            f[6] = d * (4.0 * m[0] + 2.0 * m[1] + m[2] - 6.0 * m[3] - 3.0 * m[4] + 6.0 * m[5] + 3.0 * m[6] - 9.0 * m[8]);
            f[7] = d * (4.0 * m[0] + 2.0 * m[1] + m[2] - 6.0 * m[3] - 3.0 * m[4] - 6.0 * m[5] - 3.0 * m[6] + 9.0 * m[8]);
            f[8] = d * (4.0 * m[0] + 2.0 * m[1] + m[2] + 6.0 * m[3] + 3.0 * m[4] - 6.0 * m[5] - 3.0 * m[6] - 9.0 * m[8]);

            /* Chai strain rate, F6:903-916 (uses the same rates s). */
            f1[1] = 4.0 * s[0] * mne[0] - 1.0 * s[1] * mne[1] - 2.0 * s[2] * mne[2] + 6.0 * s[3] * mne[3] - 6.0 * s[4] * mne[4] + 9.0 * s[7] * mne[7];
            f1[2] = 4.0 * s[0] * mne[0] - 1.0 * s[1] * mne[1] - 2.0 * s[2] * mne[2] + 6.0 * s[5] * mne[5] - 6.0 * s[6] * mne[6] - 9.0 * s[7] * mne[7];
            f1[3] = 4.0 * s[0] * mne[0] - 1.0 * s[1] * mne[1] - 2.0 * s[2] * mne[2] - 6.0 * s[3] * mne[3] + 6.0 * s[4] * mne[4] + 9.0 * s[7] * mne[7];
            f1[4] = 4.0 * s[0] * mne[0] - 1.0 * s[1] * mne[1] - 2.0 * s[2] * mne[2] - 6.0 * s[5] * mne[5] + 6.0 * s[6] * mne[6] - 9.0 * s[7] * mne[7];
            f1[5] = 4.0 * s[0] * mne[0] + 2.0 * s[1] * mne[1] + 1.0 * s[2] * mne[2] + 6.0 * s[3] * mne[3] + 3.0 * s[4] * mne[4] + 6.0 * s[5] * mne[5] + 3.0 * s[6] * mne[6] + 9.0 * s[8] * mne[8];
            f1[6] = 4.0 * s[0] * mne[0] + 2.0 * s[1] * mne[1] + 1.0 * s[2] * mne[2] - 6.0 * s[3] * mne[3] - 3.0 * s[4] * mne[4] + 6.0 * s[5] * mne[5] + 3.0 * s[6] * mne[6] - 9.0 * s[8] * mne[8];
            f1[7] = 4.0 * s[0] * mne[0] + 2.0 * s[1] * mne[1] + 1.0 * s[2] * mne[2] - 6.0 * s[3] * mne[3] - 3.0 * s[4] * mne[4] - 6.0 * s[5] * mne[5] - 3.0 * s[6] * mne[6] + 9.0 * s[8] * mne[8];
            f1[8] = 4.0 * s[0] * mne[0] + 2.0 * s[1] * mne[1] + 1.0 * s[2] * mne[2] + 6.0 * s[3] * mne[3] + 3.0 * s[4] * mne[4] - 6.0 * s[5] * mne[5] - 3.0 * s[6] * mne[6] - 9.0 * s[8] * mne[8];
            double zarib = -1.5 * d;
            double d11 = zarib * (f1[1] + f1[5] + f1[6] + f1[3] + f1[7] + f1[8])
                       + 0.75 * ((s[7] - s[1]) * (gx * ux + gy * uy) - s[7] * (2.0 * gx * ux));
            double d22 = zarib * (f1[5] + f1[2] + f1[6] + f1[7] + f1[4] + f1[8])
                       + 0.75 * ((s[7] - s[1]) * (gx * ux + gy * uy) - s[7] * (2.0 * gy * uy));
            double d12 = zarib * (f1[5] - f1[6] + f1[7] - f1[8])
                       + 0.75 * (-(s[7] * (gx * uy + gy * ux)));
            double *str = st->stress + 3 * id;
            str[0] = p->nu * 2.0 * d11;
            str[1] = p->nu * 2.0 * d22;
            str[2] = p->nu * 2.0 * d12;
            if (p->rheology != 0) {
                /* Papanastasiou-Bingham (F6:921-958) or, rheology 2, the
                 * Papanastasiou-Casson lines F6 carries commented out
                 * (F6:941, F6:944; F3:986-989). */
                double inv_rho = 1.0 / rho;
                d11 = d11 * inv_rho; d22 = d22 * inv_rho; d12 = d12 * inv_rho;
                double gam = sqrt(2.0 * fabs(d11 * d11 + d22 * d22 + 2.0 * (d12 * d12)));
                double nu_a;
                if (p->rheology == 2) {
                    if (gam < p->gamma_min) {
                        nu_a = sqrt(p->mu0) + sqrt(p->m_pap * p->tau_y);
                    } else {
                        double c = sqrt(p->mu0) +
                                   sqrt(p->tau_y / gam) * (1.0 - exp(-sqrt(p->m_pap * gam)));
                        nu_a = c * c;
                    }
                } else if (gam < p->gamma_min)
                    nu_a = p->mu0 + p->m_pap * p->tau_y;
                else
                    nu_a = p->mu0 + (p->tau_y / gam) * (1.0 - exp(-(p->m_pap * gam)));
                double t11 = nu_a * 2.0 * d11, t22 = nu_a * 2.0 * d22, t12 = nu_a * 2.0 * d12;
                str[0] = t11; str[1] = t22; str[2] = t12;
                double tmag = sqrt(0.5 * fabs(t11 * t11 + t22 * t22 + 2.0 * (t12 * t12)));
                st->image[id] = (tmag <= p->tau_y) ? 5 : 0;
                nu_a = nu_a * inv_rho;
                st->omega_bn[id] = 1.0 / (3.0 * nu_a + 0.5);
                orc_relaxation(p, st->omega_bn[id], s);
            }
        }
    }
}

/* F_POST = F, F6:329. */
void orc_copy_post(const orc_params *p, orc_state *st)
{
    memcpy(st->fpost, st->f, 9 * NODES(p) * sizeof(double));
}

/* STREAMING, F6:1154-1190: pull from in-domain sources only. */
void orc_stream(const orc_params *p, orc_state *st)
{
    for (int j = 0; j <= p->ny; ++j)
        for (int i = 0; i <= p->nx; ++i)
            for (int k = 0; k < 9; ++k) {
                int io = i - CX[k], jo = j - CY[k];
                if (io >= 0 && io <= p->nx && jo >= 0 && jo <= p->ny)
                    st->f[9 * NID(p, i, j) + k] = st->fpost[9 * NID(p, io, jo) + k];
            }
}

/* BOUNDARYCONDITION: halfway bounce-back with wall velocities on the edges
 * (F6:1195-1390) or, boundary 1, the Zou-He edges of F1:1222-1417; the
 * corners bounce back in both.  The bounce-back rules use the previous
 * step's density.  Boundary 2 is F4's rule (F4:1205-1402): every edge and
 * corner node takes the equilibrium of the previous step's RHO, UX, UY at
 * its interior neighbour, the diagonal one at the corners. */
void orc_boundary(const orc_params *p, orc_state *st, int32_t kk)
{
    (void)kk;
    const double inv6 = F32(1.0f / 6.0f), clb2 = 1.0;
    const int nx = p->nx, ny = p->ny;
    double *F = st->f, *FP = st->fpost, *R = st->rho;
    if (p->boundary == 2) {
        /* F4:1218-1219 bottom and top, F4:1277-1278 left and right, then
         * the corners F4:1368-1373; every read is of an interior node, so
         * the order of the writes does not matter */
        const double *U = st->ux, *V = st->uy;
#define EQ(i, j, si, sj) \
    orc_feq(U[NID(p, si, sj)], V[NID(p, si, sj)], R[NID(p, si, sj)], F + 9 * NID(p, i, j))
        for (int i = 1; i <= nx - 1; ++i) {
            EQ(i, 0, i, 1);
            EQ(i, ny, i, ny - 1);
        }
        for (int j = 1; j <= ny - 1; ++j) {
            EQ(0, j, 1, j);
            EQ(nx, j, nx - 1, j);
        }
        EQ(0, 0, 1, 1);
        EQ(0, ny, 1, ny - 1);
        EQ(nx, ny, nx - 1, ny - 1);
        EQ(nx, 0, nx - 1, 1);
#undef EQ
        return;
    }
    double uxS = p->wall[0], uyS = p->wall[1], uxR = p->wall[2], uyR = p->wall[3];
    double uxN = p->wall[4], uyN = p->wall[5], uxL = p->wall[6], uyL = p->wall[7];
#define FA(k, i, j) F[9 * NID(p, i, j) + (k)]
#define FPA(k, i, j) FP[9 * NID(p, i, j) + (k)]
#define RA(i, j) R[NID(p, i, j)]
    if (p->boundary == 1) {
        /* Zou-He edges, F1:1246-1258 (top, bottom) and F1:1319-1341 (left,
         * right), from the node's own streamed populations. */
        for (int i = 1; i <= nx - 1; ++i) {
            double rhon = (FA(0, i, ny) + FA(1, i, ny) + FA(3, i, ny) +
                           2. * (FA(2, i, ny) + FA(6, i, ny) + FA(5, i, ny))) / (1.0 + uyN);
            FA(4, i, ny) = FA(2, i, ny) - 4. * inv6 * rhon * uyN;
            FA(8, i, ny) = FA(6, i, ny) - 0.5 * (FA(1, i, ny) - FA(3, i, ny)) + 0.5 * rhon * uxN -
                           inv6 * rhon * uyN;
            FA(7, i, ny) = FA(5, i, ny) + 0.5 * (FA(1, i, ny) - FA(3, i, ny)) - 0.5 * rhon * uxN -
                           inv6 * rhon * uyN;
            double rhos = (FA(0, i, 0) + FA(1, i, 0) + FA(3, i, 0) +
                           2. * (FA(4, i, 0) + FA(7, i, 0) + FA(8, i, 0))) / (1.0 - uyS);
            FA(2, i, 0) = FA(4, i, 0) + 4.0 * inv6 * rhos * uyS;
            FA(5, i, 0) = FA(7, i, 0) - 0.5 * (FA(1, i, 0) - FA(3, i, 0)) + 0.5 * rhos * uxS +
                          inv6 * rhos * uyS;
            FA(6, i, 0) = FA(8, i, 0) + 0.5 * (FA(1, i, 0) - FA(3, i, 0)) - 0.5 * rhos * uxS +
                          inv6 * rhos * uyS;
        }
        for (int j = 1; j <= ny - 1; ++j) {
            double rl = (FA(0, 0, j) + FA(2, 0, j) + FA(4, 0, j) +
                         2.0 * (FA(3, 0, j) + FA(6, 0, j) + FA(7, 0, j))) / (1.0 - uxL);
            FA(5, 0, j) = FA(7, 0, j) - 0.5 * (FA(2, 0, j) - FA(4, 0, j)) + inv6 * rl * uxL +
                          0.5 * rl * uyL;
            FA(8, 0, j) = FA(6, 0, j) + 0.5 * (FA(2, 0, j) - FA(4, 0, j)) + inv6 * rl * uxL -
                          0.5 * rl * uyL;
            FA(1, 0, j) = FA(3, 0, j) + 4.0 * inv6 * rl * uxL;
            double rr = (FA(0, nx, j) + FA(2, nx, j) + FA(4, nx, j) +
                         2.0 * (FA(1, nx, j) + FA(5, nx, j) + FA(8, nx, j))) / (1.0 + uxR);
            FA(6, nx, j) = FA(8, nx, j) - 0.5 * (FA(2, nx, j) - FA(4, nx, j)) - inv6 * rr * uxR +
                           0.5 * rr * uyR;
            FA(7, nx, j) = FA(5, nx, j) + 0.5 * (FA(2, nx, j) - FA(4, nx, j)) - inv6 * rr * uxR -
                           0.5 * rr * uyR;
            FA(3, nx, j) = FA(1, nx, j) - 4.0 * inv6 * rr * uxR;
        }
    } else {
    for (int i = 1; i <= nx - 1; ++i) {
        FA(8, i, ny) = FPA(6, i, ny) - inv6 * clb2 * RA(i, ny) * (-uxN + uyN);
        FA(4, i, ny) = FPA(2, i, ny) - 4.0 * inv6 * clb2 * RA(i, ny) * (uyN);
        FA(7, i, ny) = FPA(5, i, ny) - inv6 * clb2 * RA(i, ny) * (uxN + uyN);
        FA(5, i, 0) = FPA(7, i, 0) - inv6 * clb2 * RA(i, 0) * (-uxS - uyS);
        FA(2, i, 0) = FPA(4, i, 0) - 4.0 * inv6 * clb2 * RA(i, 0) * (-uyS);
        FA(6, i, 0) = FPA(8, i, 0) - inv6 * clb2 * RA(i, 0) * (uxS - uyS);
    }
    for (int j = 1; j <= ny - 1; ++j) {
        FA(6, nx, j) = FPA(8, nx, j) - inv6 * clb2 * RA(nx, j) * (uxR - uyR);
        FA(3, nx, j) = FPA(1, nx, j) - 4.0 * inv6 * clb2 * RA(nx, j) * (uxR);
        FA(7, nx, j) = FPA(5, nx, j) - inv6 * clb2 * RA(nx, j) * (uxR + uyR);
        FA(5, 0, j) = FPA(7, 0, j) - inv6 * clb2 * RA(0, j) * (-uxL - uyL);
        FA(1, 0, j) = FPA(3, 0, j) - 4.0 * inv6 * clb2 * RA(0, j) * (-uxL);
        FA(8, 0, j) = FPA(6, 0, j) - inv6 * clb2 * RA(0, j) * (-uxL + uyL);
    }
    }
    /* corners: bounce-back with the corner's density (F6:1364-1386, the
     * same in F1:1390-1414) */
    double uxw = p->corner[0], uyw = p->corner[1], rw;
    rw = RA(0, 0);
    FA(1, 0, 0) = FPA(3, 0, 0) - 4.0 * inv6 * clb2 * rw * (-uxw);
    FA(5, 0, 0) = FPA(7, 0, 0) - inv6 * clb2 * rw * (-uxw - uyw);
    FA(2, 0, 0) = FPA(4, 0, 0) - 4.0 * inv6 * clb2 * rw * (-uyw);
    rw = RA(0, ny);
    FA(1, 0, ny) = FPA(3, 0, ny) - 4.0 * inv6 * clb2 * rw * (-uxw);
    FA(8, 0, ny) = FPA(6, 0, ny) - inv6 * clb2 * rw * (-uxw + uyw);
    FA(4, 0, ny) = FPA(2, 0, ny) - 4.0 * inv6 * clb2 * rw * (uyw);
    rw = RA(nx, ny);
    FA(3, nx, ny) = FPA(1, nx, ny) - 4.0 * inv6 * clb2 * rw * (uxw);
    FA(7, nx, ny) = FPA(5, nx, ny) - inv6 * clb2 * rw * (uxw + uyw);
    FA(4, nx, ny) = FPA(2, nx, ny) - 4.0 * inv6 * clb2 * rw * (uyw);
    rw = RA(nx, 0);
    FA(2, nx, 0) = FPA(4, nx, 0) - 4.0 * inv6 * clb2 * rw * (-uyw);
    FA(6, nx, 0) = FPA(8, nx, 0) - inv6 * clb2 * rw * (uxw - uyw);
    FA(3, nx, 0) = FPA(1, nx, 0) - 4.0 * inv6 * clb2 * rw * (uxw);
#undef FA
#undef FPA
#undef RA
}

/* MACROSCOPICVARIABLES, F6:1394-1412. */
void orc_macros(const orc_params *p, orc_state *st)
{
    for (int j = 0; j <= p->ny; ++j)
        for (int i = 0; i <= p->nx; ++i) {
            size_t id = NID(p, i, j);
            const double *f = st->f + 9 * id;
            double r = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
            st->rho[id] = r;
            double rinv = 1.0 / r;
            st->ux[id] = rinv * (f[1] + f[5] + f[8] - (f[6] + f[3] + f[7]));
            st->uy[id] = rinv * (f[5] + f[2] + f[6] - (f[7] + f[4] + f[8]));
        }
}

/* ComputeConcenterationFunction, F6:430-473. */
void orc_phi(const orc_params *p, orc_state *st)
{
    const double PI = F32(3.14159265359f);
    size_t n = NODES(p);
    for (size_t id = 0; id < n; ++id) st->image[id] = 5;
    for (int ip = 0; ip < st->np; ++ip) {
        double xpc = PF(st, ORC_PF_X, ip), ypc = PF(st, ORC_PF_Y, ip);
        double rp = 0.5 * PF(st, ORC_PF_DIAM, ip);
        double tp = PF(st, ORC_PF_ANG, ip);
        double zeta = PF(st, ORC_PF_ZETA, ip);
        double xpr = xpc + 0.75 * rp * cos(tp);
        double ypr = ypc + 0.75 * rp * sin(tp);
        double *phi = st->phi + (size_t)ip * n;
        for (int j = 0; j <= p->ny; ++j)
            for (int i = 0; i <= p->nx; ++i) {
                size_t id = NID(p, i, j);
                double di = rp - sqrt(((double)i - xpc) * ((double)i - xpc) +
                                      ((double)j - ypc) * ((double)j - ypc));
                if (di <= -(0.5 * zeta))
                    phi[id] = 0.0;
                else if (fabs(di) < 0.5 * zeta)
                    phi[id] = 0.5 * (sin(PI * di / zeta) + 1.0);
                else if (di >= 0.5 * zeta)
                    phi[id] = 1.0;
                if (phi[id] >= 0.5) {
                    st->image[id] = 10;
                    double dp = 25.0 - (((double)i - xpr) * ((double)i - xpr) +
                                        ((double)j - ypr) * ((double)j - ypr));
                    if (dp >= 0) st->image[id] = 20;
                }
            }
    }
}

/* ComputeForceField_SPM, F6:477-503: the highest particle index with phi>0
 * overwrites G. */
void orc_forcing(const orc_params *p, orc_state *st)
{
    size_t n = NODES(p);
    for (int j = 0; j <= p->ny; ++j)
        for (int i = 0; i <= p->nx; ++i) {
            size_t id = NID(p, i, j);
            st->g[2 * id] = 0.0;
            st->g[2 * id + 1] = 0.0;
            for (int ip = 0; ip < st->np; ++ip) {
                double ph = st->phi[(size_t)ip * n + id];
                if (ph > 0.0) {
                    double uxp = PF(st, ORC_PF_VX, ip) - PF(st, ORC_PF_W, ip) * ((double)j - PF(st, ORC_PF_Y, ip));
                    double uyp = PF(st, ORC_PF_VY, ip) + PF(st, ORC_PF_W, ip) * ((double)i - PF(st, ORC_PF_X, ip));
                    st->g[2 * id] = ph * (uxp - st->ux[id]);
                    st->g[2 * id + 1] = ph * (uyp - st->uy[id]);
                }
            }
        }
}

/* ComputeForce_HYDRO, F6:534-558: per particle, j outer / i inner. */
void orc_hydro(const orc_params *p, orc_state *st)
{
    size_t n = NODES(p);
    for (int ip = 0; ip < st->np; ++ip) {
        double fp0 = 0.0, fp1 = 0.0, fp2 = 0.0;
        double vx = PF(st, ORC_PF_VX, ip), vy = PF(st, ORC_PF_VY, ip), w = PF(st, ORC_PF_W, ip);
        double xc = PF(st, ORC_PF_X, ip), yc = PF(st, ORC_PF_Y, ip);
        const double *phi = st->phi + (size_t)ip * n;
        for (int j = 0; j <= p->ny; ++j)
            for (int i = 0; i <= p->nx; ++i) {
                size_t id = NID(p, i, j);
                double ph = phi[id];
                if (ph > 0.0) {
                    double uxp = vx - w * ((double)j - yc);
                    double uyp = vy + w * ((double)i - xc);
                    double r = st->rho[id], ux = st->ux[id], uy = st->uy[id];
                    fp0 = fp0 + r * ph * (ux - uxp);
                    fp1 = fp1 + r * ph * (uy - uyp);
                    fp2 = fp2 + r * ph * (((double)i - xc) * (uy - uyp) - ((double)j - yc) * (ux - uxp));
                }
            }
        PF(st, ORC_PF_FX, ip) = fp0;
        PF(st, ORC_PF_FY, ip) = fp1;
        PF(st, ORC_PF_TQ, ip) = fp2;
    }
}

/* VelocityCorrector, F6:1111-1125 (DT = 1, G_X = G_Y = 0). */
void orc_correct(const orc_params *p, orc_state *st)
{
    size_t n = NODES(p);
    for (size_t id = 0; id < n; ++id) {
        st->ux[id] = st->ux[id] + 0.5 * 1.0 * (st->g[2 * id] + 0.0);
        st->uy[id] = st->uy[id] + 0.5 * 1.0 * (st->g[2 * id + 1] + 0.0);
    }
}

/* The reference's (..)**8 and (..)**14 (F4:595-596, F4:616-617) with the
 * roundings of its compiled form: **8 as seven multiplications left to
 * right, **14 by the binary method (x^2, x^4, x^8 by squaring, then
 * (x^2 x^4) x^8).  Of the orders tried (binary, left to right, from x^7
 * squared, reversed products) only this pair reproduces every F_COL of the
 * compiled Lennard-Jones routine. */
static double pow8(double x)
{
    return ((((((x * x) * x) * x) * x) * x) * x) * x;
}
static double pow14(double x)
{
    double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
    return (x2 * x4) * x8;
}

/* Pair term of the pair (ip, jp), ip < jp, with the stale wall distance dpw
 * for the Glowinski branch; returns 0 (and zero force) when the pair is out
 * of range d > r_i + r_j + zeta.
 *   contact 0, Glowinski: F3:746-770 (F6:700-720 for one pair).
 *   contact 1, Wan-Turek: F4:632-645 / F1:654-667: eps_p (x_i - x_j) times
 *     (r_i + r_j - d) inside contact, (r_i + r_j + zeta - d)^2 in the gap.
 *   contact 2, Lennard-Jones: F4:590-599 / F1:612-621 with R_P the radius of
 *     particle 1 (F4:588): 2.4 R_P^2 (2 q^14 - q^8) (x_i - x_j) / (4 R_P^2),
 *     q = 2 R_P / d. */
static int pair_term(const orc_params *p, const orc_state *st, int ip, int jp, double dpw,
                     double *f0, double *f1)
{
    const double zeta = p->zeta_c, cij = p->cij, eps_p = p->eps_p, e_p = p->e_p;
    double xi = PF(st, ORC_PF_X, ip), yi = PF(st, ORC_PF_Y, ip);
    double xj = PF(st, ORC_PF_X, jp), yj = PF(st, ORC_PF_Y, jp);
    double di = PF(st, ORC_PF_DIAM, ip), dj = PF(st, ORC_PF_DIAM, jp);
    double dpp = sqrt((xi - xj) * (xi - xj) + (yi - yj) * (yi - yj));
    double rsum = 0.5 * di + 0.5 * dj;
    *f0 = 0.0; *f1 = 0.0;
    if (dpp > rsum + zeta) return 0;
    if (p->contact == 1) {
        if (dpp < rsum) {
            *f0 = eps_p * (xi - xj) * (rsum - dpp);
            *f1 = eps_p * (yi - yj) * (rsum - dpp);
        } else {
            double t = rsum + zeta - dpp;
            *f0 = eps_p * (xi - xj) * (t * t);
            *f1 = eps_p * (yi - yj) * (t * t);
        }
        return 1;
    }
    if (p->contact == 2) {
        double rp = 0.5 * PF(st, ORC_PF_DIAM, 0);
        double q = 2.0 * rp / dpp;
        double a = F32(2.4f) * (rp * rp) * (2.0 * pow14(q) - pow8(q));
        *f0 = a * (xi - xj) / (4.0 * rp * rp);
        *f1 = a * (yi - yj) / (4.0 * rp * rp);
        return 1;
    }
    double t = (dpp - 0.5 * di - 0.5 * dj - zeta) / zeta;
    if (dpw > rsum) {
        *f0 = eps_p * cij * (t * t) * (xi - xj) / dpp;
        *f1 = eps_p * cij * (t * t) * (yi - yj) / dpp;
    } else {
        *f0 = (eps_p * cij * (t * t) + e_p * cij * (rsum - dpp) / zeta) * (xi - xj) / dpp;
        *f1 = (eps_p * cij * (t * t) + e_p * cij * (rsum - dpp) / zeta) * (yi - yj) / dpp;
    }
    return 1;
}

/* Force of wall iw (1 bottom, 2 right, 3 top, 4 left) on a disk at (x, y),
 * diameter dm, from the wall image R_W; returns the distance to it.
 *   contact 0, Glowinski: images at half a diameter outside the domain
 *     (F3:773-804 / F6:726-759).
 *   contact 1, Wan-Turek: mirror images about the wall edges E
 *     (F4:648-676 / F1:670-698): eps_w (c - R_W) (dm - d) inside,
 *     (dm + zeta - d)^2 within zeta.
 *   contact 2, Lennard-Jones: the images of F4:600-621 / F1:622-643 on the
 *     nx x ny box ((x, 0), (nx, y), (x, 2 ny - y), (-x, y)) and
 *     2.4 R_P^2 (2 q^14 - q^8) (c - R_W) / R_P^2, q = R_P / d. */
static double wall_term(const orc_params *p, const orc_state *st, double x, double y, double dm,
                        int iw, double *f0, double *f1)
{
    const double zeta = p->zeta_c, cij = p->cij, eps_w = p->eps_w, e_w = p->e_w;
    const double *E = p->wall_edge;
    double rw0, rw1;
    if (p->contact == 1) {
        if (iw == 1) { rw0 = x; rw1 = 2.0 * E[0] - y; }
        else if (iw == 2) { rw0 = 2.0 * E[1] - x; rw1 = y; }
        else if (iw == 3) { rw0 = x; rw1 = 2.0 * E[2] - y; }
        else { rw0 = 2.0 * E[3] - x; rw1 = y; }
    } else if (p->contact == 2) {
        if (iw == 1) { rw0 = x; rw1 = 0.0; }
        else if (iw == 2) { rw0 = (double)p->nx; rw1 = y; }
        else if (iw == 3) { rw0 = x; rw1 = 2.0 * (double)p->ny - y; }
        else { rw0 = -x; rw1 = y; }
    } else {
        if (iw == 1) { rw0 = x; rw1 = -(0.5 * dm); }
        else if (iw == 2) { rw0 = (double)p->nx + 0.5 * dm; rw1 = y; }
        else if (iw == 3) { rw0 = x; rw1 = (double)p->ny + 0.5 * dm; }
        else { rw0 = -(0.5 * dm); rw1 = y; }
    }
    double d = sqrt((x - rw0) * (x - rw0) + (y - rw1) * (y - rw1));
    *f0 = 0.0; *f1 = 0.0;
    if (d > dm + zeta) return d;
    if (p->contact == 1) {
        if (d < dm) {
            *f0 = eps_w * (x - rw0) * (dm - d);
            *f1 = eps_w * (y - rw1) * (dm - d);
        } else {
            double t = dm + zeta - d;
            *f0 = eps_w * (x - rw0) * (t * t);
            *f1 = eps_w * (y - rw1) * (t * t);
        }
    } else if (p->contact == 2) {
        double rp = 0.5 * PF(st, ORC_PF_DIAM, 0);
        double q = rp / d;
        double a = F32(2.4f) * (rp * rp) * (2.0 * pow14(q) - pow8(q));
        *f0 = a * (x - rw0) / (rp * rp);
        *f1 = a * (y - rw1) / (rp * rp);
    } else if (d > dm) {
        double t = (d - dm - zeta) / zeta;
        *f0 = eps_w * cij * (t * t) * (x - rw0) / d;
        *f1 = eps_w * cij * (t * t) * (y - rw1) / d;
    } else {
        double t = (d - dm - zeta) / zeta;
        *f0 = (eps_w * cij * (t * t) + e_w * cij * (dm - d) / zeta) * (x - rw0) / d;
        *f1 = (eps_w * cij * (t * t) + e_w * cij * (dm - d) / zeta) * (y - rw1) / d;
    }
    return d;
}

/* ComputeCollisionForce: N-body pair loop of F3:739-770 (accumulating, IP
 * ascending then JP ascending) and the four walls in the order bottom,
 * right, top, left, with the pair and wall terms of the selected model
 * (pair_term, wall_term).  The reference's Wan-Turek and Lennard-Jones
 * blocks handle the one pair (1, 2) by assignment (F4:632-645); for two
 * disks the accumulation gives the same bits, and for more it sums every
 * pair as F3's loop does.  The Glowinski pair branch reads D_PW as left by
 * the previous call's last wall evaluation (F3:760, F6:711); dpw_stale
 * carries it.  Returns the number of pairs with d <= r_i + r_j + zeta
 * (written i<j, ascending, into pairs[2*n] when there is room). */
int32_t orc_contact(const orc_params *p, orc_state *st, int32_t *pairs,
                    int32_t max_pairs)
{
    const int np = st->np;
    double dpw = st->dpw_stale;
    int32_t npairs = 0;
    for (int ip = 0; ip < np; ++ip) {
        PF(st, ORC_PF_FCX, ip) = 0.0;
        PF(st, ORC_PF_FCY, ip) = 0.0;
    }
    for (int ip = 0; ip < np - 1; ++ip) {
        for (int jp = ip + 1; jp < np; ++jp) {
            double fpp0, fpp1;
            if (pair_term(p, st, ip, jp, dpw, &fpp0, &fpp1)) {
                if (pairs && npairs < max_pairs) {
                    pairs[2 * npairs] = ip;
                    pairs[2 * npairs + 1] = jp;
                }
                ++npairs;
            }
            PF(st, ORC_PF_FCX, ip) = PF(st, ORC_PF_FCX, ip) + fpp0;
            PF(st, ORC_PF_FCY, ip) = PF(st, ORC_PF_FCY, ip) + fpp1;
            PF(st, ORC_PF_FCX, jp) = PF(st, ORC_PF_FCX, jp) - fpp0;
            PF(st, ORC_PF_FCY, jp) = PF(st, ORC_PF_FCY, jp) - fpp1;
        }
    }
    for (int ip = 0; ip < np; ++ip) {
        double x = PF(st, ORC_PF_X, ip), y = PF(st, ORC_PF_Y, ip), dm = PF(st, ORC_PF_DIAM, ip);
        for (int iw = 1; iw <= 4; ++iw) {
            double fpw0, fpw1;
            dpw = wall_term(p, st, x, y, dm, iw, &fpw0, &fpw1);
            PF(st, ORC_PF_FCX, ip) = PF(st, ORC_PF_FCX, ip) + fpw0;
            PF(st, ORC_PF_FCY, ip) = PF(st, ORC_PF_FCY, ip) + fpw1;
        }
    }
    st->dpw_stale = dpw;
    return npairs;
}

/* Particle update, inline in MAIN, F6:299-312 / F3:312-328; with motion = 1
 * the prescribed oscillation of F4:221-226 for loop iteration kk, with the
 * added-mass term of F4:256-258. */
void orc_update(const orc_params *p, orc_state *st, int32_t kk)
{
    const double rho0 = p->rho0;
    if (p->motion == 1) {
        const double pi = F32(3.14159265359f);
        for (int ip = 0; ip < st->np; ++ip) {
            double rhop = PF(st, ORC_PF_RHOP, ip), mp = PF(st, ORC_PF_MP, ip);
            double ipm = PF(st, ORC_PF_IP, ip);
            double v0 = PF(st, ORC_PF_VX, ip), v1 = PF(st, ORC_PF_VY, ip), v2 = PF(st, ORC_PF_W, ip);
            PF(st, ORC_PF_VOX, ip) = v0; PF(st, ORC_PF_VOY, ip) = v1; PF(st, ORC_PF_VOW, ip) = v2;
            double n0 = -p->osc_u * cos(2.0 * pi * (double)kk / p->osc_period);
            double n1 = 0.0, n2 = 0.0;
            PF(st, ORC_PF_VX, ip) = n0; PF(st, ORC_PF_VY, ip) = n1; PF(st, ORC_PF_W, ip) = n2;
            PF(st, ORC_PF_X, ip) = PF(st, ORC_PF_X, ip) + 0.5 * (n0 + v0);
            PF(st, ORC_PF_Y, ip) = PF(st, ORC_PF_Y, ip) + 0.5 * (n1 + v1);
            PF(st, ORC_PF_FHX, ip) = (rho0 / rhop) * mp * (n0 - v0);
            PF(st, ORC_PF_FHY, ip) = (rho0 / rhop) * mp * (n1 - v1);
            PF(st, ORC_PF_FHW, ip) = (rho0 / rhop) * ipm * (n2 - v2);
        }
        return;
    }
    for (int ip = 0; ip < st->np; ++ip) {
        double rhop = PF(st, ORC_PF_RHOP, ip), mp = PF(st, ORC_PF_MP, ip), ipm = PF(st, ORC_PF_IP, ip);
        double v0 = PF(st, ORC_PF_VX, ip), v1 = PF(st, ORC_PF_VY, ip), v2 = PF(st, ORC_PF_W, ip);
        double fh0 = (rho0 / rhop) * mp * (v0 - PF(st, ORC_PF_VOX, ip));
        double fh1 = (rho0 / rhop) * mp * (v1 - PF(st, ORC_PF_VOY, ip));
        double fh2 = (rho0 / rhop) * ipm * (v2 - PF(st, ORC_PF_VOW, ip));
        PF(st, ORC_PF_FHX, ip) = fh0; PF(st, ORC_PF_FHY, ip) = fh1; PF(st, ORC_PF_FHW, ip) = fh2;
        PF(st, ORC_PF_VOX, ip) = v0; PF(st, ORC_PF_VOY, ip) = v1; PF(st, ORC_PF_VOW, ip) = v2;
        double n0 = v0 + PF(st, ORC_PF_FX, ip) / mp + fh0 / mp + PF(st, ORC_PF_FCX, ip) / mp;
        double n1 = v1 + PF(st, ORC_PF_FY, ip) / mp + fh1 / mp + (1.0 - rho0 / rhop) * p->gy
                  + PF(st, ORC_PF_FCY, ip) / mp;
        double n2 = v2 + PF(st, ORC_PF_TQ, ip) / ipm + fh2 / ipm;
        PF(st, ORC_PF_VX, ip) = n0; PF(st, ORC_PF_VY, ip) = n1; PF(st, ORC_PF_W, ip) = n2;
        PF(st, ORC_PF_X, ip) = PF(st, ORC_PF_X, ip) + 0.5 * (n0 + v0);
        PF(st, ORC_PF_Y, ip) = PF(st, ORC_PF_Y, ip) + 0.5 * (n1 + v1);
        if (p->update_angle)
            PF(st, ORC_PF_ANG, ip) = PF(st, ORC_PF_ANG, ip) + 0.5 * (n2 + v2);
    }
}

/* SUM(RHO) in column-major order (F6:348). */
double orc_total_mass(const orc_params *p, const orc_state *st)
{
    double s = 0.0;
    size_t n = NODES(p);
    for (size_t id = 0; id < n; ++id) s += st->rho[id];
    return s;
}

/* Nodes flagged yielded (tau_mag > tau_y) by the last collision, F6:949-954. */
int64_t orc_yielded_nodes(const orc_params *p, const orc_state *st)
{
    int64_t c = 0;
    size_t n = NODES(p);
    for (size_t id = 0; id < n; ++id) c += (st->image[id] == 0);
    return c;
}

/* MAIN time loop, F6:291-349. */
void orc_run(const orc_params *p, orc_state *st, int32_t kk0, int32_t nsteps,
             double *hist)
{
    const size_t rec = (size_t)st->np * ORC_H_PCOUNT + 2;
    for (int32_t s = 0; s < nsteps; ++s) {
        int32_t kk = kk0 + s + 1;
        orc_contact(p, st, NULL, 0);
        orc_update(p, st, kk);
        orc_phi(p, st);
        orc_collide(p, st);
        orc_copy_post(p, st);
        orc_stream(p, st);
        orc_boundary(p, st, kk);
        orc_macros(p, st);
        orc_forcing(p, st);
        orc_hydro(p, st);
        orc_correct(p, st);
        if (hist) {
            double *h = hist + (size_t)s * rec;
            for (int ip = 0; ip < st->np; ++ip) {
                double *hp = h + (size_t)ip * ORC_H_PCOUNT;
                hp[ORC_H_X] = PF(st, ORC_PF_X, ip);
                hp[ORC_H_Y] = PF(st, ORC_PF_Y, ip);
                hp[ORC_H_ANG] = PF(st, ORC_PF_ANG, ip);
                hp[ORC_H_VX] = PF(st, ORC_PF_VX, ip);
                hp[ORC_H_VY] = PF(st, ORC_PF_VY, ip);
                hp[ORC_H_W] = PF(st, ORC_PF_W, ip);
                hp[ORC_H_FX] = PF(st, ORC_PF_FX, ip);
                hp[ORC_H_FY] = PF(st, ORC_PF_FY, ip);
                hp[ORC_H_TQ] = PF(st, ORC_PF_TQ, ip);
                hp[ORC_H_FHX] = PF(st, ORC_PF_FHX, ip);
                hp[ORC_H_FHY] = PF(st, ORC_PF_FHY, ip);
                hp[ORC_H_FHW] = PF(st, ORC_PF_FHW, ip);
            }
            h[(size_t)st->np * ORC_H_PCOUNT] = orc_total_mass(p, st);
            h[(size_t)st->np * ORC_H_PCOUNT + 1] = (double)orc_yielded_nodes(p, st);
        }
    }
}
