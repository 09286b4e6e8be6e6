/*
 * lbm_amd.h -- C ABI of the MI355X (gfx950) D2Q9 MRT / smoothed-profile time
 * step.  It is the drop-in boundary for the per-step subroutines of
 * taha-re/2D_NNF_LBM, whose Fortran driver calls them through implicit
 * interfaces with arrays passed by address.  Reference routines are cited as
 * F6:line (6_FreeFall_YieldStress/D2Q9_BinghamSediment_Validated_20191127.f90)
 * and F3:line (3_MultiParticle_Sedimentation/D2Q9_RTParticle_20191003.f90).
 *
 * Conventions
 *  - Grid: nx, ny cells; nodes 0..nx by 0..ny (F6:8-9).
 *  - Host arrays use the reference layout: F(0:8,0:NX,0:NY) column-major
 *    (q fastest, then i, then j); node fields (RHO, UX, ...) i fastest;
 *    G(1:2,I,J) and STRESS(1:3,I,J) component fastest.
 *  - Particles cross as one SoA block of LBM_PF_COUNT * n doubles:
 *    field f of particle p at part[f*n + p] (fields below).  No full-grid
 *    solid fraction crosses the boundary.
 *  - The library owns all device memory.  Host arrays are caller-owned and are
 *    read or written only inside the set/get call that receives them.
 *  - Every call returns 0 on success or a negative LBM_ERR_* code; the text of
 *    the last error is in lbm_last_error().  Nothing calls exit().
 *  - One handle per host thread; calls are serialised on the handle's stream.
 */
#ifndef LBM_AMD_H
#define LBM_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_PARAM (-1)       /* bad parameter, e.g. tau < 0.5 (F6:115-118) */
#define LBM_ERR_HIP (-2)         /* HIP runtime failure */
#define LBM_ERR_UNSUPPORTED (-3) /* combination not provided on the GPU path */
#define LBM_ERR_CAPACITY (-4)    /* particle, pair or history capacity exceeded */
#define LBM_ERR_NAN (-5)         /* non-finite value detected by a check */
#define LBM_ERR_IO (-6)          /* a restart file could not be read or written */

#define LBM_NEWTONIAN 0
#define LBM_BINGHAM 1            /* Papanastasiou-regularised Bingham (F6:939-945) */
#define LBM_CASSON 2             /* Papanastasiou-regularised Casson (F6:941, F6:944; F3:986-989) */

#define LBM_MOTION_FREE 0        /* Newton-Euler update of the particles (F6:299-312) */
#define LBM_MOTION_OSC_X 1       /* prescribed x-oscillation V_x = -U cos(2 pi KK / T) (F4:221-226) */

#define LBM_BOUNCE_BACK 0        /* halfway bounce-back with moving walls on every edge (F6:1195-1390) */
#define LBM_ZOU_HE 1             /* Zou-He velocity edges, bounce-back corners (F1:1222-1417) */
#define LBM_EQUILIBRIUM 2        /* equilibrium of the interior neighbour on edges and corners (F4:1205-1402) */
#define LBM_CONTACT_GLOWINSKI 0  /* Glowinski springs, N-body (F3:739-804; F6:690-760) */
#define LBM_CONTACT_WAN_TUREK 1  /* Wan-Turek, mirror walls about wall_edge (F4:624-677; F1:646-699) */
#define LBM_CONTACT_LJ 2         /* Lennard-Jones (F4:587-622; F1:609-644) */

#define LBM_RATE_LAGGED 0        /* module-global rate set by the previous node (F6:958) */
#define LBM_RATE_LOCAL 1         /* each node's own previous-step rate */

#define LBM_FP64 0               /* populations stored and collided in FP64 */
#define LBM_FP32 1               /* FP32 deviations f - w_i*rho0, FP32 collide */
#define LBM_FP16 2               /* FP16 deviations (scaled by 2^10), FP32 collide */
#define LBM_BF16 3               /* BF16 deviations, FP32 collide */

/* Particle record fields (TYPE PARTICLE, F6:69-81, plus VP_OLD and F_H_IN, F6:96). */
enum {
    LBM_PF_X = 0, LBM_PF_Y, LBM_PF_ANG, LBM_PF_VX, LBM_PF_VY, LBM_PF_W,
    LBM_PF_DIAM, LBM_PF_FX, LBM_PF_FY, LBM_PF_TQ, LBM_PF_FCX, LBM_PF_FCY,
    LBM_PF_RHOP, LBM_PF_MP, LBM_PF_IP, LBM_PF_ZETA,
    LBM_PF_VOX, LBM_PF_VOY, LBM_PF_VOW, LBM_PF_FHX, LBM_PF_FHY, LBM_PF_FHW,
    LBM_PF_COUNT
};

/* Per-step history record per particle (written by lbm_step). */
enum {
    LBM_H_X = 0, LBM_H_Y, LBM_H_ANG, LBM_H_VX, LBM_H_VY, LBM_H_W,
    LBM_H_FX, LBM_H_FY, LBM_H_TQ, LBM_H_FHX, LBM_H_FHY, LBM_H_FHW,
    LBM_H_PCOUNT
};

/* Run parameters; the compile-time PARAMETERs of MODULE PARAM (F6:2-83)
 * become run-time values. */
typedef struct lbm_params {
    int32_t nx, ny;            /* F6:8-9 */
    int32_t rheology;          /* LBM_NEWTONIAN / LBM_BINGHAM / LBM_CASSON (F6:4) */
    int32_t rate_mode;         /* LBM_RATE_*; the GPU path implements LOCAL */
    int32_t q_mode;            /* 1: s4=s6=8(2-w)/(8-w) (F6:1009); 0: sq */
    int32_t update_angle;      /* 1: integrate the angle (F6:312); 0: F3:327 */
    int32_t storage;           /* LBM_FP64 / LBM_FP32 / LBM_FP16 / LBM_BF16 */
    int32_t max_particles;     /* capacity of the particle arrays */
    int32_t motion;            /* LBM_MOTION_FREE / LBM_MOTION_OSC_X */
    int32_t boundary;          /* LBM_BOUNCE_BACK / LBM_ZOU_HE (edges; corners bounce back) / LBM_EQUILIBRIUM */
    double omega;              /* OMEGA = 1/(3 nu + 0.5) (F6:101) */
    double s0, s1, s2, s3, s5, sq;  /* other MRT rates (F6:1009-1011) */
    double nu;                 /* NU, Newtonian stress factor (F6:918) */
    double mu0, tau_y, m_pap, gamma_min;  /* Papanastasiou (F6:939-945) */
    double rho0;               /* RHO0 (F6:60) */
    double gy;                 /* A_G, gravity along y (F6:28, F6:308) */
    double wall[8];            /* uxS,uyS, uxR,uyR, uxN,uyN, uxL,uyL (F6:1204-1264) */
    double corner[2];          /* ux_w, uy_w at the four corners (F6:1365) */
    double zeta_c, cij, eps_p, eps_w, e_p, e_w;  /* contact (F6:691-693, F3:739-742) */
    double osc_u, osc_period;  /* U_C, T_C of LBM_MOTION_OSC_X (F4:224) */
    int32_t contact;           /* LBM_CONTACT_*; zeta_c and eps_p, eps_w are the model's
                                  ZITA, EPS_P, EPS_W (Wan-Turek: F4:625, F1:647; L-J: F4:588) */
    int32_t reserved;
    double wall_edge[4];       /* Wan-Turek walls WALL%EDGE(1:4): bottom y, right x, top y,
                                  left x (F4:182-185); the domain box is 0, nx, ny, 0 */
} lbm_params;

typedef struct lbm_handle lbm_handle;

/* Lifecycle.  lbm_create validates the parameters (tau >= 0.5, F6:115-118),
 * allocates device state and leaves it zeroed; call lbm_initialize next. */
int lbm_create(const lbm_params *p, lbm_handle **out);
int lbm_destroy(lbm_handle *h);
/* Text of the last error on h; with h == NULL, the reason the last
 * lbm_create on this thread failed. */
const char *lbm_last_error(const lbm_handle *h);
int lbm_sync(lbm_handle *h);

/* INITIALIZE, F6:563-580 (f = feq(rho0, 0), G = 0), with OMEGA_BN = OMEGA
 * (F6:102) and the initial relaxation vector (F6:199).  Also restarts the run
 * state: the stochastic-rounding sequence of the reduced formats and the
 * stale wall distance of the contact force, so a run after lbm_initialize
 * does not depend on what the handle ran before. */
int lbm_initialize(lbm_handle *h);

/* State transfer (reference layout on the host, SoA on the device). */
int lbm_set_f(lbm_handle *h, const double *f);        /* F(0:8,0:NX,0:NY) */
int lbm_get_f(lbm_handle *h, double *f);
int lbm_set_fpost(lbm_handle *h, const double *f);    /* F_POST */
int lbm_get_fpost(lbm_handle *h, double *f);
/* RHO, UX, UY, G(1:2,:,:), OMEGA_BN; any pointer may be NULL to skip it. */
int lbm_set_fields(lbm_handle *h, const double *rho, const double *ux,
                   const double *uy, const double *g, const double *omega_bn);
int lbm_get_fields(lbm_handle *h, double *rho, double *ux, double *uy,
                   double *g, double *omega_bn);
int lbm_get_stress(lbm_handle *h, double *stress);    /* STRESS(1:3,0:NX,0:NY) */
int lbm_get_image(lbm_handle *h, int32_t *image);     /* DOMAIN_IMAGE */
int lbm_set_particles(lbm_handle *h, int32_t n, const double *part);
int lbm_get_particles(lbm_handle *h, double *part);
/* Solid fraction of particle ip on the full grid (PHI(0:NX,0:NY) of F6:79),
 * as computed by the last lbm_spm_phi. */
int lbm_get_phi(lbm_handle *h, int32_t ip, double *phi);

/* Checkpoint and restart.
 * lbm_restart_write replaces WRITE_OUTPUT (F6:377-405): it writes the text
 * state files 0_F_output.txt, 0_UX/0_UY/0_RHO/0_Gspm/0_OmegaBN/
 * 0_DomainImage_output.txt (one record per node, I outer, J inner),
 * 0_Particle_output.txt (CENTER, VELOCITY, FORCE per disk),
 * 0_OmegaMRT_output.txt (the relaxation vector OMEGA_MRT(0:8) that the last
 * node of the sweep leaves, F6:958) and the empty 0_Wall_output.txt into
 * directory dir, numbers with 17 significant digits.
 * lbm_restart_read replaces READ_INPUT (F6:174-198): it reads the same files
 * (not the image, as the reference) in any list-directed number format; the
 * particles must be set first (the file replaces CENTER, VELOCITY, FORCE,
 * the other fields stay, as the reference sets them from its parameters),
 * and entry 7 of OMEGA_MRT becomes the module-global rate of
 * LBM_RATE_LAGGED collisions.
 * Both return LBM_ERR_IO with the file name in lbm_last_error on failure. */
int lbm_restart_write(lbm_handle *h, const char *dir);
int lbm_restart_read(lbm_handle *h, const char *dir);

/* Binary snapshot of the whole run state in host memory: F, F_POST, RHO,
 * UX, UY, G, OMEGA_BN, STRESS, DOMAIN_IMAGE, the particle records, the stale
 * wall distance of the contact force, the module-global rate and the
 * position of the stochastic-rounding sequence; with reduced-precision
 * storage also the deviations h = f - w_k rho0 of F exactly (FP32: the
 * FP64 value w_k rho0 + h rounds the smallest h).  Loading it into a handle
 * created with the same parameters continues the run exactly: the next
 * lbm_step gives the same results bit for bit as it would have given on the
 * saved handle (a handle with another grid or storage format is refused).
 * No counterpart in the reference (its restart files are the text ones
 * above); for callers that checkpoint in memory. */
int lbm_snapshot_size(const lbm_handle *h, uint64_t *bytes);
int lbm_snapshot_save(lbm_handle *h, void *buf, uint64_t bytes);
int lbm_snapshot_load(lbm_handle *h, const void *buf, uint64_t bytes);

/* Drop-in per-subroutine entries (FP64 state, reference semantics). */
/* Collision_MRT, F6:867-962.  With a yield-stress rheology (LBM_BINGHAM,
 * LBM_CASSON) and LBM_RATE_LAGGED it keeps
 * the reference's module-global rate (each node collides with the rate the
 * node before it set, F6:958), which is a sequential chain over the grid:
 * one GPU thread walks it, so use it for drop-in fidelity on small grids. */
int lbm_collide(lbm_handle *h);
int lbm_copy_post(lbm_handle *h);         /* F_POST = F, F6:329 */
int lbm_stream(lbm_handle *h);            /* STREAMING, F6:1154-1190 */
int lbm_boundary(lbm_handle *h, int32_t kk);  /* BOUNDARYCONDITION, F6:1195-1390 */
int lbm_macros(lbm_handle *h);            /* MACROSCOPICVARIABLES, F6:1394-1412 */
int lbm_spm_phi(lbm_handle *h);           /* ComputeConcenterationFunction, F6:430-473 */
int lbm_spm_forcing(lbm_handle *h);       /* ComputeForceField_SPM, F6:477-503 */
int lbm_hydro_force(lbm_handle *h);       /* ComputeForce_HYDRO, F6:534-558 */
int lbm_correct(lbm_handle *h);           /* VelocityCorrector, F6:1111-1125 */
/* ComputeCollisionForce (F6:584-761; N-body loop F3:632-806).  pairs (may be
 * NULL) receives the in-range pairs (i<j, ascending, 0-based) as 2*n ints. */
int lbm_contact_force(lbm_handle *h, int32_t *pairs, int32_t max_pairs,
                      int32_t *npairs);
/* Particle update of loop iteration kk: inline in MAIN (F6:299-312), or the
 * prescribed motion of F4:221-226 and its added-mass term (F4:256-258). */
int lbm_particle_update(lbm_handle *h, int32_t kk);
int lbm_total_mass(lbm_handle *h, double *mass);  /* SUM(RHO), F6:348 */

/* Fused time steps kk0+1 .. kk0+nsteps (the loop body F6:293-349), all on the
 * device, in the storage precision of the handle.  hist (may be NULL):
 * nsteps rows of (n_particles*LBM_H_PCOUNT + 2) doubles -- the particle
 * records, then SUM(RHO) and the number of yielded nodes.
 * The state stays resident on the device between calls: a call that
 * continues the previous one (kk0 equal to its last step, and nothing set
 * or read through the per-subroutine entries in between) runs exactly
 * nsteps step launches; the reference's loop-end arrays (F, F_POST, RHO,
 * UX, UY, G, OMEGA_BN, STRESS, DOMAIN_IMAGE, the particle records) are
 * written when a get/set call, a per-subroutine entry or a snapshot needs
 * them.  Results do not depend on how the steps are split into calls. */
int lbm_step(lbm_handle *h, int32_t kk0, int32_t nsteps, double *hist);
/* Pairs found by the contact evaluation of the last step of lbm_step
 * (the contact force of its loop-end state). */
int lbm_last_pairs(lbm_handle *h, int32_t *pairs, int32_t max_pairs,
                   int32_t *npairs);
/* Like lbm_step (no history), timed on the handle's stream with HIP events:
 * ms_total covers every launch of the call; ms_main covers its nsteps step
 * launches, which run back to back with no other work between them (a call
 * that starts from the loop-end arrays first runs one prologue launch,
 * outside ms_main), so ms_main/nsteps is the average duration of one launch
 * of the main kernel. */
int lbm_step_timed(lbm_handle *h, int32_t kk0, int32_t nsteps,
                   double *ms_total, double *ms_main);
/* lbm_step with the history, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * history rows (layout as lbm_step's hist, which must not be NULL and may be
 * host or device memory) are staged in a static 4 MiB device array that the
 * library holds on each GPU until the process ends, instead of the handle's
 * history buffers, which lbm_step grows with hipMalloc.  A call needs
 * nsteps * (24 * n_particles + 66) * 8 bytes of it; a larger call fails with
 * LBM_ERR_CAPACITY before it runs anything.  Calls on one GPU share the
 * array: calls that overlap in time on the same GPU (other threads, other
 * handles' streams) race and must be serialised by the caller. */
int lbm_step_hist_static(lbm_handle *h, int32_t kk0, int32_t nsteps, double *hist);

/* ---------------------------------------------------------------- adjoint
 * Sensitivity of the settling velocity of one disk (the reference has no
 * adjoint; this entry serves parameter studies).  Runs nsteps fused steps
 * from the handle's current state and returns
 *   loss = (1/window) * sum over the last `window` steps of V_y,
 *   grad = d loss / d (tau_y, mu0, rho_p),
 * by the discrete adjoint of the step (FP64 arithmetic) with a checkpoint
 * every `chk` steps and recomputation of each segment, so device memory is
 * O(nsteps/chk + chk) states.  mu0 enters as the Papanastasiou plastic
 * viscosity and through OMEGA = 1/(3 mu0 + 1/2), the initial node rate;
 * rho_p moves the mass m = rho_p pi r^2 and inertia m r^2 / 2 with it.
 * Requires FP64 or FP32 storage, one particle, free motion and walls at rest;
 * returns LBM_ERR_UNSUPPORTED if the particle reaches contact range.  The
 * handle ends in the state after nsteps.  mem_bytes (may be NULL) receives
 * the device memory the adjoint allocated. */
int lbm_adjoint_velocity(lbm_handle *h, int32_t nsteps, int32_t window, int32_t chk,
                         double *loss, double *grad, double *mem_bytes);

/* ------------------------------------------------------------------ batch
 * Many independent simulations (grid sizes, rheology, walls and particle
 * motion may differ; one storage format) stepped together: one launch per
 * time step covers every grid.  The reference runs one simulation per
 * program; these entries have no counterpart there and serve ensembles. */
typedef struct lbm_batch lbm_batch;
/* params: n records; each simulation gets its own handle (lbm_batch_sim),
 * through which particles are set and state is read with the calls above. */
int lbm_batch_create(int32_t n, const lbm_params *params, lbm_batch **out);
int lbm_batch_destroy(lbm_batch *b);
const char *lbm_batch_last_error(const lbm_batch *b);
int lbm_batch_size(const lbm_batch *b, int32_t *n);
lbm_handle *lbm_batch_sim(lbm_batch *b, int32_t i);
int lbm_batch_initialize(lbm_batch *b);
/* nsteps fused steps (kk0+1 .. kk0+nsteps) of every simulation, as lbm_step
 * per simulation; hist (may be NULL) receives each simulation's history rows
 * (nsteps rows of n_particles*LBM_H_PCOUNT + 2) one simulation after another. */
int lbm_batch_step(lbm_batch *b, int32_t kk0, int32_t nsteps, double *hist);
int lbm_batch_step_timed(lbm_batch *b, int32_t kk0, int32_t nsteps, double *ms_total);
/* chunk > 0 (even): steps 1 .. nsteps-2 of a call are replayed in chunks of
 * `chunk` launches from one captured hipGraph; 0: every launch eager. */
int lbm_batch_set_graph(lbm_batch *b, int32_t chunk);
/* 256-node tiles handled by one 256-thread block, per simulation (1..64). */
int lbm_batch_set_tiles(lbm_batch *b, const int32_t *tiles_per_block);
int lbm_batch_get_tiles(const lbm_batch *b, int32_t *tiles_per_block);
/* Autotuner: times whole-batch steps (max(steps, 8) of them, less a 2-step
 * call, so each call's start and end launches cancel) with 1, 2, 4 and 8
 * tiles per block for every simulation, keeps the fastest, then tries the
 * other three for each size bucket (floor(log2(nodes))) alone, keeping a
 * change that is 1 % faster.  The timed steps advance the simulations:
 * initialise and set the particles again afterwards.  tiles_per_block (may
 * be NULL) gets the choice. */
int lbm_batch_tune(lbm_batch *b, int32_t steps, int32_t *tiles_per_block);

#ifdef __cplusplus
}
#endif
#endif
