/*
 * CPU oracle: a serial C restatement of the D2Q9 MRT / smoothed-profile time
 * step of taha-re/2D_NNF_LBM (Fortran, 6_FreeFall_YieldStress/
 * D2Q9_BinghamSediment_Validated_20191127.f90, "F6" below; the N-body contact
 * loop follows 3_MultiParticle_Sedimentation/D2Q9_RTParticle_20191003.f90,
 * "F3").
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only as the checker.
 *
 * Conventions kept from the reference so results are bit-comparable:
 *  - nodes 0..nx by 0..ny (F6:8-9); populations stored as F(0:8,0:NX,0:NY) in
 *    Fortran column-major order: q fastest, then i, then j;
 *  - every float32 literal of the reference is reproduced as (double)(float)
 *    (F6:873 d_mrt, F6:1202 Inv_Six, F6:1135-1137 weights, F6:3 PI, ...);
 *  - sweeps run j outer, i inner, and every sum keeps the reference's
 *    left-to-right order.  Build with -ffp-contract=off.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Particle record fields, stored SoA: part[field * np + p]. */
enum {
    ORC_PF_X = 0, ORC_PF_Y, ORC_PF_ANG, ORC_PF_VX, ORC_PF_VY, ORC_PF_W,
    ORC_PF_DIAM, ORC_PF_FX, ORC_PF_FY, ORC_PF_TQ, ORC_PF_FCX, ORC_PF_FCY,
    ORC_PF_RHOP, ORC_PF_MP, ORC_PF_IP, ORC_PF_ZETA,
    ORC_PF_VOX, ORC_PF_VOY, ORC_PF_VOW, ORC_PF_FHX, ORC_PF_FHY, ORC_PF_FHW,
    ORC_PF_COUNT
};

/* Per-step history record (written after the step's velocity correction). */
enum {
    ORC_H_X = 0, ORC_H_Y, ORC_H_ANG, ORC_H_VX, ORC_H_VY, ORC_H_W,
    ORC_H_FX, ORC_H_FY, ORC_H_TQ, ORC_H_FHX, ORC_H_FHY, ORC_H_FHW,
    ORC_H_PCOUNT
};

typedef struct orc_params {
    int32_t nx, ny;          /* cells; nodes 0..nx, 0..ny (F6:8-9) */
    int32_t rheology;        /* 0 Newtonian, 1 Papanastasiou-Bingham (F6:4),
                                2 Papanastasiou-Casson (F6:941, F6:944) */
    int32_t rate_mode;       /* 0 lagged module-global rate (F6:890, F6:958),
                                1 each node's own previous-step rate */
    int32_t q_mode;          /* 1: s4=s6=8(2-w)/(8-w) (F6:1009); 0: sq */
    int32_t update_angle;    /* 1: F6:312 integrates the angle; 0: F3:327 */
    int32_t motion;          /* 0 free (F6:299-312); 1 prescribed x-oscillation (F4:221-226) */
    int32_t boundary;        /* 0 halfway bounce-back (F6:1195-1390), 1 Zou-He edges (F1:1222-1417),
                                2 equilibrium edges and corners (F4:1205-1402) */
    double omega;            /* OMEGA = 1/(3 nu + 0.5) (F6:101) */
    double s0, s1, s2, s3, s5, sq;  /* remaining MRT rates (F6:1009-1011) */
    double nu;               /* NU, Newtonian stress factor (F6:918) */
    double mu0, tau_y, m_pap, gamma_min;  /* F6:939-945 */
    double rho0;             /* RHO0 (F6:60) */
    double gy;               /* A_G, gravity in y (F6:28, F6:308) */
    double wall[8];          /* uxS,uyS, uxR,uyR, uxN,uyN, uxL,uyL (F6:1204-1264) */
    double corner[2];        /* ux_w, uy_w at the four corners (F6:1365) */
    double zeta_c, cij, eps_p, eps_w, e_p, e_w;  /* contact (F6:691-693) */
    double osc_u, osc_period;  /* U_C and T_C of the prescribed velocity (F4:224) */
    int32_t contact;         /* 0 Glowinski (F3:739-804), 1 Wan-Turek (F4:624-677),
                                2 Lennard-Jones (F4:587-622) */
    int32_t pad_;
    double wall_edge[4];     /* Wan-Turek walls WALL%EDGE(1:4): bottom y, right x, top y,
                                left x (F4:182-185) */
} orc_params;

typedef struct orc_state {
    double *f, *fpost;       /* 9 * nodes */
    double *rho, *ux, *uy;   /* nodes */
    double *g;               /* 2 * nodes, G(1:2,I,J) */
    double *omega_bn;        /* nodes */
    double *stress;          /* 3 * nodes */
    int32_t *image;          /* DOMAIN_IMAGE, nodes */
    double *phi;             /* np * nodes */
    double *part;            /* ORC_PF_COUNT * np */
    double s_mrt[9];         /* module-global OMEGA_MRT (F6:58) */
    double dpw_stale;        /* D_PW left by the previous contact call */
    int32_t np;
} orc_state;

void orc_feq(double ux, double uy, double rho, double *feq);
void orc_relaxation(const orc_params *p, double omega, double *s);
void orc_initialize(const orc_params *p, orc_state *st);
void orc_collide(const orc_params *p, orc_state *st);
void orc_copy_post(const orc_params *p, orc_state *st);
void orc_stream(const orc_params *p, orc_state *st);
void orc_boundary(const orc_params *p, orc_state *st, int32_t kk);
void orc_macros(const orc_params *p, orc_state *st);
void orc_phi(const orc_params *p, orc_state *st);
void orc_forcing(const orc_params *p, orc_state *st);
void orc_hydro(const orc_params *p, orc_state *st);
void orc_correct(const orc_params *p, orc_state *st);
int32_t orc_contact(const orc_params *p, orc_state *st, int32_t *pairs,
                    int32_t max_pairs);
void orc_update(const orc_params *p, orc_state *st, int32_t kk);
double orc_total_mass(const orc_params *p, const orc_state *st);
int64_t orc_yielded_nodes(const orc_params *p, const orc_state *st);
/* Runs steps kk0+1 .. kk0+nsteps in the reference order (F6:291-370).
 * hist (optional): per step, ORC_H_PCOUNT doubles per particle, then the total
 * mass and the yielded-node count. */
void orc_run(const orc_params *p, orc_state *st, int32_t kk0, int32_t nsteps,
             double *hist);

#ifdef __cplusplus
}
#endif
#endif
