// Kernels of the D2Q9 MRT / smoothed-profile time step for gfx950.
//
// Drop-in kernels (one per reference subroutine) work on the FP64 state in
// the reference's meaning; the fused kernels run the whole time step in one
// pass over the grid.  Layout in HBM: structure of arrays, one plane per
// population (plane stride P.plane elements), node index j*(nx+1)+i.
#pragma once
#include "lbm_device.h"

namespace lbm {

// Tiles are handed to the 8 XCDs in contiguous ranges so that rows shared
// between neighbouring tiles meet in the same L2 (blocks are dispatched to
// XCDs round-robin by block index).  A bijection of [0, T).
__device__ __forceinline__ long long xcd_tile(long long b, long long T) {
    long long x = b & 7, per = T >> 3, rem = T & 7;
    long long start = x * per + (x < rem ? x : rem);
    return start + (b >> 3);
}

// =============================================================== drop-in
__global__ void k_init(DevP P, double *f, double *rho, double *ux, double *uy, double *g,
                       double *obn) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    double r = P.rho0, u = 0.0, v = 0.0;
    double fe[9];
    feq64(u, v, r, fe);  // COMPUTEFEQNEW at rest (INITIALIZE, F6:555-572)
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k * P.plane + n] = fe[k];
    rho[n] = r; ux[n] = u; uy[n] = v;
    g[n] = 0.0; g[P.plane + n] = 0.0;
    obn[n] = P.omega;
}

// Collision_MRT (F6:867-962) with each node's own rate.
__global__ void k_collide(DevP P, double *f, const double *rho, const double *ux,
                          const double *uy, const double *g, double *obn, double *stress,
                          int *image) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = f[k * P.plane + n];
    double om = P.rheology != 0 ? obn[n] : P.omega;
    CollideOut o;
    collide64(P, v, rho[n], ux[n], uy[n], g[n], g[P.plane + n], om, o);
#pragma unroll
    for (int k = 0; k < 9; ++k) f[k * P.plane + n] = v[k];
    stress[n] = o.stress[0];
    stress[P.plane + n] = o.stress[1];
    stress[2 * P.plane + n] = o.stress[2];
    if (P.rheology != 0) {
        image[n] = o.yielded ? 0 : 5;
        obn[n] = o.omega_bn;
    }
}

// Collision_MRT with the reference's module-global rate (F6:890, F6:958):
// each node collides with the rate the node before it set, in the reference's
// node order (J outer, I inner), which makes the pass one sequential chain.
// One thread walks the grid; *lag carries the rate from call to call (the
// module variable OMEGA_MRT, F6:58).  A fidelity mode for drop-in runs of
// the unmodified reference; the fused step uses each node's own rate.
__global__ void k_collide_lagged(DevP P, double *f, const double *rho, const double *ux,
                                 const double *uy, const double *g, double *obn,
                                 double *stress, int *image, double *lag) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double om = *lag;
    for (long long n = 0; n < P.nn; ++n) {
        double v[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = f[k * P.plane + n];
        CollideOut o;
        collide64(P, v, rho[n], ux[n], uy[n], g[n], g[P.plane + n], om, o);
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k * P.plane + n] = v[k];
        stress[n] = o.stress[0];
        stress[P.plane + n] = o.stress[1];
        stress[2 * P.plane + n] = o.stress[2];
        image[n] = o.yielded ? 0 : 5;
        obn[n] = o.omega_bn;
        om = o.omega_bn;
    }
    *lag = om;
}

// STREAMING, F6:1164-1188.
__global__ void k_stream(DevP P, double *f, const double *fp) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    int j = (int)(n / P.W), i = (int)(n - (long long)j * P.W);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        int si = i - cxk(k), sj = j - cyk(k);
        if ((unsigned)si <= (unsigned)P.nx && (unsigned)sj <= (unsigned)P.ny)
            f[k * P.plane + n] = fp[k * P.plane + n - cxk(k) - (long long)cyk(k) * P.W];
    }
}

__device__ __forceinline__ void bnd_node(const DevP &P, int b, int &i, int &j) {
    if (b < P.W) { i = b; j = 0; }
    else if (b < 2 * P.W) { i = b - P.W; j = P.ny; }
    else if (b < 2 * P.W + P.H - 2) { i = 0; j = b - 2 * P.W + 1; }
    else { i = P.nx; j = b - 2 * P.W - (P.H - 2) + 1; }
}

// BOUNDARYCONDITION, F6:1195-1390 (or the Zou-He edges of F1:1222-1417, or
// the equilibrium edges and corners of F4:1205-1402):
// one thread per boundary node.  The rules at each node are those of
// load_fk64 with the streamed values kept for the populations the rules do
// not touch.
__global__ void k_boundary(DevP P, double *f, const double *fp, const double *rho,
                           const double *ux, const double *uy) {
    int b = blockIdx.x * NT + threadIdx.x;
    if (b >= P.nbnd) return;
    int i, j;
    bnd_node(P, b, i, j);
    long long n = (long long)j * P.W + i;
    if (P.boundary == 2) {  // F4's equilibrium of the interior neighbour
        int si = i < 1 ? 1 : (i > P.nx - 1 ? P.nx - 1 : i);
        int sj = j < 1 ? 1 : (j > P.ny - 1 ? P.ny - 1 : j);
        long long s = (long long)sj * P.W + si;
        double v[9];
        feq64(ux[s], uy[s], rho[s], v);
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k * P.plane + n] = v[k];
        return;
    }
    const bool corner = (i == 0 || i == P.nx) && (j == 0 || j == P.ny);
    if (P.boundary == 1 && !corner) {
        double v[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = f[k * P.plane + n];
        zou_he_edge<double>(P, i, j, 0.0, v);
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k * P.plane + n] = v[k];
        return;
    }
    const double clb2 = 1.0;
    double rw = rho[n];
    auto fo = [&](int k) { return fp[k * P.plane + n]; };
    auto F = [&](int k) -> double & { return f[k * P.plane + n]; };
    const double *wl = P.wall;
    if (j == P.ny && i >= 1 && i <= P.nx - 1) {
        double uxN = wl[4], uyN = wl[5];
        F(8) = fo(6) - INV6 * clb2 * rw * (-uxN + uyN);
        F(4) = fo(2) - 4.0 * INV6 * clb2 * rw * (uyN);
        F(7) = fo(5) - INV6 * clb2 * rw * (uxN + uyN);
    } else if (j == 0 && i >= 1 && i <= P.nx - 1) {
        double uxS = wl[0], uyS = wl[1];
        F(5) = fo(7) - INV6 * clb2 * rw * (-uxS - uyS);
        F(2) = fo(4) - 4.0 * INV6 * clb2 * rw * (-uyS);
        F(6) = fo(8) - INV6 * clb2 * rw * (uxS - uyS);
    } else if (i == P.nx && j >= 1 && j <= P.ny - 1) {
        double uxR = wl[2], uyR = wl[3];
        F(6) = fo(8) - INV6 * clb2 * rw * (uxR - uyR);
        F(3) = fo(1) - 4.0 * INV6 * clb2 * rw * (uxR);
        F(7) = fo(5) - INV6 * clb2 * rw * (uxR + uyR);
    } else if (i == 0 && j >= 1 && j <= P.ny - 1) {
        double uxL = wl[6], uyL = wl[7];
        F(5) = fo(7) - INV6 * clb2 * rw * (-uxL - uyL);
        F(1) = fo(3) - 4.0 * INV6 * clb2 * rw * (-uxL);
        F(8) = fo(6) - INV6 * clb2 * rw * (-uxL + uyL);
    } else {
        double uxw = P.corner[0], uyw = P.corner[1];
        if (i == 0 && j == 0) {
            F(1) = fo(3) - 4.0 * INV6 * clb2 * rw * (-uxw);
            F(5) = fo(7) - INV6 * clb2 * rw * (-uxw - uyw);
            F(2) = fo(4) - 4.0 * INV6 * clb2 * rw * (-uyw);
        } else if (i == 0) {
            F(1) = fo(3) - 4.0 * INV6 * clb2 * rw * (-uxw);
            F(8) = fo(6) - INV6 * clb2 * rw * (-uxw + uyw);
            F(4) = fo(2) - 4.0 * INV6 * clb2 * rw * (uyw);
        } else if (j == P.ny) {
            F(3) = fo(1) - 4.0 * INV6 * clb2 * rw * (uxw);
            F(7) = fo(5) - INV6 * clb2 * rw * (uxw + uyw);
            F(4) = fo(2) - 4.0 * INV6 * clb2 * rw * (uyw);
        } else {
            F(2) = fo(4) - 4.0 * INV6 * clb2 * rw * (-uyw);
            F(6) = fo(8) - INV6 * clb2 * rw * (uxw - uyw);
            F(3) = fo(1) - 4.0 * INV6 * clb2 * rw * (uxw);
        }
    }
}

// MACROSCOPICVARIABLES, F6:1394-1412.
__global__ void k_macros(DevP P, const double *f, double *rho, double *ux, double *uy) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = f[k * P.plane + n];
    double r, u, w;
    macros64(v, r, u, w);
    rho[n] = r; ux[n] = u; uy[n] = w;
}

// Solid-fraction patches (ComputeConcenterationFunction, F6:449-459): phi of
// particle p on a PS x PS patch whose origin covers every node that can have
// phi > 0.  Nodes outside the domain hold 0.
__device__ __forceinline__ void patch_origin(const double *part, int maxp, int p, int &i0,
                                             int &j0) {
    double reach = 0.5 * part[PDIAM * maxp + p] + 0.5 * part[PZETA * maxp + p] + 1.0;
    i0 = (int)floor(part[PX * maxp + p] - reach);
    j0 = (int)floor(part[PY * maxp + p] - reach);
}

__global__ void k_phi_patch(DevP P, const double *part, double *phi, int *porig, int PS) {
    int p = blockIdx.y;
    int a = blockIdx.x * NT + threadIdx.x;
    if (a >= PS * PS) return;
    int i0, j0;
    patch_origin(part, P.maxp, p, i0, j0);
    if (a == 0) { porig[2 * p] = i0; porig[2 * p + 1] = j0; }
    int i = i0 + a % PS, j = j0 + a / PS;
    double ph = 0.0;
    if (i >= 0 && i <= P.nx && j >= 0 && j <= P.ny)
        ph = phi_at(part[PX * P.maxp + p], part[PY * P.maxp + p], part[PDIAM * P.maxp + p],
                    part[PZETA * P.maxp + p], i, j);
    phi[(size_t)p * PS * PS + a] = ph;
}

__device__ __forceinline__ double patch_phi(const double *phi, const int *porig, int PS, int p,
                                            int i, int j) {
    int a = i - porig[2 * p], b = j - porig[2 * p + 1];
    if (a < 0 || b < 0 || a >= PS || b >= PS) return 0.0;
    return phi[(size_t)p * PS * PS + (size_t)b * PS + a];
}

// DOMAIN_IMAGE of ComputeConcenterationFunction (F6:438, F6:460-467).
__global__ void k_image(DevP P, const double *part, const double *phi, const int *porig,
                        int PS, int *image) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    int j = (int)(n / P.W), i = (int)(n - (long long)j * P.W);
    int img = 5;
    for (int p = 0; p < P.np; ++p) {
        double ph = patch_phi(phi, porig, PS, p, i, j);
        if (ph >= 0.5) {
            img = 10;
            double rp = 0.5 * part[PDIAM * P.maxp + p], tp = part[PANG * P.maxp + p];
            double xpr = part[PX * P.maxp + p] + 0.75 * rp * cos(tp);
            double ypr = part[PY * P.maxp + p] + 0.75 * rp * sin(tp);
            double dp = 25.0 - (((double)i - xpr) * ((double)i - xpr) +
                                ((double)j - ypr) * ((double)j - ypr));
            if (dp >= 0) img = 20;
        }
    }
    image[n] = img;
}

// ComputeForceField_SPM, F6:485-495.
__global__ void k_forcing(DevP P, const double *part, const double *phi, const int *porig,
                          int PS, const double *ux, const double *uy, double *g) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    int j = (int)(n / P.W), i = (int)(n - (long long)j * P.W);
    double gx = 0.0, gy = 0.0;
    for (int p = 0; p < P.np; ++p) {
        double ph = patch_phi(phi, porig, PS, p, i, j);
        if (ph > 0.0) {
            double w = part[PW * P.maxp + p];
            double uxp = part[PVX * P.maxp + p] - w * ((double)j - part[PY * P.maxp + p]);
            double uyp = part[PVY * P.maxp + p] + w * ((double)i - part[PX * P.maxp + p]);
            gx = ph * (uxp - ux[n]);
            gy = ph * (uyp - uy[n]);
        }
    }
    g[n] = gx;
    g[P.plane + n] = gy;
}

// ComputeForce_HYDRO, F6:542-556: one thread per particle summing its patch
// in the reference's j-outer / i-inner order (bit-identical to the full-grid
// sum, since nodes with phi = 0 add nothing).
__global__ void k_hydro_seq(DevP P, double *part, const double *phi, const int *porig, int PS,
                            const double *rho, const double *ux, const double *uy) {
    int p = blockIdx.x * NT + threadIdx.x;
    if (p >= P.np) return;
    const int m = P.maxp;
    double vx = part[PVX * m + p], vy = part[PVY * m + p], w = part[PW * m + p];
    double xc = part[PX * m + p], yc = part[PY * m + p];
    int i0 = porig[2 * p], j0 = porig[2 * p + 1];
    double fp0 = 0.0, fp1 = 0.0, fp2 = 0.0;
    for (int b = 0; b < PS; ++b) {
        int j = j0 + b;
        if (j < 0 || j > P.ny) continue;
        for (int a = 0; a < PS; ++a) {
            int i = i0 + a;
            if (i < 0 || i > P.nx) continue;
            double ph = phi[(size_t)p * PS * PS + (size_t)b * PS + a];
            if (ph > 0.0) {
                long long n = (long long)j * P.W + i;
                double uxp = vx - w * ((double)j - yc);
                double uyp = vy + w * ((double)i - xc);
                double r = rho[n], u = ux[n], v = uy[n];
                fp0 = fp0 + r * ph * (u - uxp);
                fp1 = fp1 + r * ph * (v - uyp);
                fp2 = fp2 + r * ph * (((double)i - xc) * (v - uyp) - ((double)j - yc) * (u - uxp));
            }
        }
    }
    part[PFX * m + p] = fp0;
    part[PFY * m + p] = fp1;
    part[PTQ * m + p] = fp2;
}

// VelocityCorrector, F6:1118-1124.
__global__ void k_correct(DevP P, double *ux, double *uy, const double *g) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    ux[n] = ux[n] + 0.5 * 1.0 * (g[n] + 0.0);
    uy[n] = uy[n] + 0.5 * 1.0 * (g[P.plane + n] + 0.0);
}

// ComputeCollisionForce: one thread per particle (F3:739-804).
__global__ void k_contact(DevP P, double *part, const double *dpw_in, double *dpw_out,
                          int *pairs, int *paircnt) {
    int p = blockIdx.x * NT + threadIdx.x;
    if (p >= P.np) return;
    double fcx, fcy, dl;
    const int m = P.maxp;
    int cnt = contact_one(P, part + PX * m, part + PY * m, part + PDIAM * m, P.np, p, *dpw_in, fcx,
                          fcy, pairs + p * MAXNB, &dl);
    paircnt[p] = cnt;
    // all threads have read the centres before any writes the contact force:
    // the force fields are not read by contact_one
    part[PFCX * P.maxp + p] = fcx;
    part[PFCY * P.maxp + p] = fcy;
    if (p == P.np - 1) *dpw_out = dl;
}

// Particle update, F6:299-312: one thread per particle.
__global__ void k_update(DevP P, double *part, int kk) {
    int p = blockIdx.x * NT + threadIdx.x;
    if (p >= P.np) return;
    double rec[PCOUNT];
    for (int f = 0; f < PCOUNT; ++f) rec[f] = part[f * P.maxp + p];
    update_one(P, rec, kk);
    for (int f = 0; f < PCOUNT; ++f) part[f * P.maxp + p] = rec[f];
}

// SUM(RHO): per-block partial sums, then one block sums the partials in order.
__global__ void k_sum_partial(const double *x, long long n, double *part) {
    __shared__ double lds[NT];
    long long i = (long long)blockIdx.x * NT + threadIdx.x;
    double v[1] = {i < n ? x[i] : 0.0};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}
__global__ void k_sum_final(const double *part, int nb, double *out) {
    __shared__ double lds[NT];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) acc += part[b];
    double v[1] = {acc};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) *out = v[0];
}

// AoS (component fastest) <-> SoA planes.
__global__ void k_aos_to_soa(long long nn, int nc, const double *aos, double *soa,
                             long long plane) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= nn) return;
    for (int c = 0; c < nc; ++c) soa[c * plane + n] = aos[n * nc + c];
}
__global__ void k_soa_to_aos(long long nn, int nc, const double *soa, double *aos,
                             long long plane) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= nn) return;
    for (int c = 0; c < nc; ++c) aos[n * nc + c] = soa[c * plane + n];
}

// ================================================================= fused
struct StepArgs {
    DevP P;
    const void *src;              // post-collision populations f*_k (storage format)
    void *dst;                    // f*_{k+1} (storage format) or, final step, f_k (FP64)
    double *cf, *cfp;             // FP64 state arrays F, F_POST
    double *crho, *cux, *cuy, *cg, *cobn, *cstress;
    int *cimage;
    float *rate32;                // node rates of the reduced formats
    // edge arrays: previous-step density on the boundary nodes or, with
    // equilibrium edges, density and corrected velocity of each boundary
    // node's interior source (3 x nbnd)
    const void *erho_src;
    void *erho_dst;
    const double *pin;            // particle state k
    double *pout;                 // particle state k+1 (final step: same as pin)
    const double *dpw_in;
    double *dpw_out;
    int *pairs, *paircnt;
    double *hist_part;            // history row of step k, or null
    unsigned long long *slots_mass;   // NSLOT fixed-point mass sums of step k, or null
    unsigned long long *slots_yield;  // NSLOT yielded-node counts of the collide's step
    int n_pblocks;
    int kk_upd;                   // loop iteration of the particle update in this launch
    long long ntiles;
    int aux;                      // write STRESS / DOMAIN_IMAGE
    // per-tile candidate lists (more than NP_LOOP particles): read for the
    // state the tiles use (generation gen_in), written for the next state
    const unsigned long long *tl_hdr_in;
    const int *tl_ids_in;
    unsigned long long *tl_hdr_out;
    int *tl_ids_out;
    unsigned gen_in, gen_out;
    int *err;                     // bit 0: a tile list overflowed
    // the force on one particle is summed by nsub blocks; the last of them
    // (counter pcount) adds the partials pforce in order and advances it
    int nsub;
    double *pforce;               // [maxp][nsub][3]
    int *pcount;                  // [maxp]
    // FP32 smoothed-profile records (reduced-precision runs; null for FP64)
    const float *spf_in;          // state k
    float *spf_out;               // state k+1
    // per-call bases from which step_fields() derives the per-step fields
    double *hist_base;            // history rows of the call, or null
    unsigned long long *slots_base;  // monitor slots of the call, or null
    int rowp;                     // history doubles per step (particles)
    int kk0;                      // step index before the call
    unsigned g0;                  // tile-list generation at the start of the call
    int tpb;                      // 256-node tiles per block (batched launches)
    // monitor partials: each tile block stores its fixed-point mass and its
    // yielded count (plain stores, [mass | yield] x ntiles, two buffers by
    // launch parity); a reducer block of the next launch sums the previous
    // launch's partials into that launch's monitor slots
    unsigned long long *mpart[2];
    unsigned long long *mpart_out;        // this launch's partials, or null
    const unsigned long long *mpart_in;   // the previous launch's, or null
    unsigned long long *red_mass, *red_yield;  // where their sums go (null: none)
    int nred;                     // reducer block slots in the block layout (0 or 1)
};

// Per-step fields of launch s of a call of nsteps fused steps (s = 0: the
// prologue; s = nsteps: the final step): tile-list generations, the loop
// iteration of the particle update, the history row and monitor slots, and
// whether the launch writes the output-only fields.
__host__ __device__ __forceinline__ void step_fields(StepArgs &a, int s, int nsteps) {
    const bool rec = a.slots_base != nullptr;
    const bool fin = s == nsteps;
    a.gen_in = a.g0 + (unsigned)s;
    a.gen_out = a.gen_in + 1u;
    a.kk_upd = a.kk0 + s + 1;
    if (s == 0) {
        a.hist_part = nullptr;
        a.slots_mass = nullptr;
        a.slots_yield = rec ? a.slots_base + NSLOT : nullptr;
        a.aux = nsteps == 1;
    } else {
        a.hist_part = rec ? a.hist_base + (size_t)(s - 1) * (size_t)a.rowp : nullptr;
        a.slots_mass = rec ? a.slots_base + (size_t)(s - 1) * 2 * NSLOT : nullptr;
        a.slots_yield = (rec && !fin) ? a.slots_base + (size_t)s * 2 * NSLOT + NSLOT : nullptr;
        a.aux = (s == nsteps - 1) || fin;
    }
    // the previous launch (s - 1) wrote mass partials from its second launch
    // on and yield partials unless it was the last (it never is here)
    a.mpart_out = rec ? a.mpart[s & 1] : nullptr;
    a.mpart_in = (rec && s > 0) ? a.mpart[(s - 1) & 1] : nullptr;
    a.red_mass = (rec && s > 1) ? a.slots_base + (size_t)(s - 2) * 2 * NSLOT : nullptr;
    a.red_yield = (rec && s > 0) ? a.slots_base + (size_t)(s - 1) * 2 * NSLOT + NSLOT : nullptr;
}

// FP32 variant of tile_has_particles for the reduced-precision path.
__device__ __forceinline__ bool tile_has_particles_f32(const StepArgs &a, long long tile) {
    const DevP &P = a.P;
    if (P.np == 0) return false;
    if (P.np > NP_LOOP) return true;
    long long n0 = tile * NT, n1 = min(n0 + NT - 1, P.nn - 1);
    float jlo = (float)row_of(P, (unsigned)n0), jhi = (float)row_of(P, (unsigned)n1);
    for (int p = 0; p < P.np; ++p) {
        const float *s = a.spf_in + p * SPF;
        float y = s[SY] + s[SYF], r = s[SREACH];
        if (y + r >= jlo && y - r <= jhi) return true;
    }
    return false;
}

// G at a node of a tile, FP32 (reduced-precision path).
__device__ __forceinline__ void spm_tile_node_f32(const StepArgs &a, long long tile, bool any,
                                                  int i, int j, float ux, float uy, float &gx,
                                                  float &gy) {
    const DevP &P = a.P;
    gx = 0.0f;
    gy = 0.0f;
    if (!any) return;
    if (P.np <= NP_LOOP) {
        spm_node_f32(a.spf_in, nullptr, P.np, i, j, ux, uy, gx, gy);
        return;
    }
    unsigned long long h = a.tl_hdr_in[tile];
    if ((unsigned)(h >> 32) != a.gen_in) return;
    int cnt = min((int)(h & 0xffffffffu), MAXC);
    spm_node_f32(a.spf_in, a.tl_ids_in + tile * MAXC, cnt, i, j, ux, uy, gx, gy);
}

// Can any particle reach the rows of this tile?  (Uniform over the block;
// with few particles it lets most tiles skip the smoothed-profile work.)
__device__ __forceinline__ bool tile_has_particles(const StepArgs &a, long long tile) {
    const DevP &P = a.P;
    if (P.np == 0) return false;
    if (P.np > NP_LOOP) return true;
    long long n0 = tile * NT, n1 = min(n0 + NT - 1, P.nn - 1);
    double jlo = (double)row_of(P, (unsigned)n0), jhi = (double)row_of(P, (unsigned)n1);
    const int m = P.maxp;
    for (int p = 0; p < P.np; ++p) {
        double reach = 0.5 * a.pin[PDIAM * m + p] + 0.5 * a.pin[PZETA * m + p] + 1.0;
        double y = a.pin[PY * m + p];
        if (y + reach >= jlo && y - reach <= jhi) return true;
    }
    return false;
}

// G at a node of a tile: all particles tested directly when there are few,
// else the tile's candidate list for the current generation.
__device__ __forceinline__ void spm_tile_node(const StepArgs &a, long long tile, bool any, int i,
                                              int j, double ux, double uy, double &gx,
                                              double &gy) {
    const DevP &P = a.P;
    gx = 0.0;
    gy = 0.0;
    if (!any) return;
    if (P.np <= NP_LOOP) {
        spm_node(a.pin, P.maxp, P.np, i, j, ux, uy, gx, gy);
        return;
    }
    unsigned long long h = a.tl_hdr_in[tile];
    if ((unsigned)(h >> 32) != a.gen_in) return;
    int cnt = min((int)(h & 0xffffffffu), MAXC);
    spm_node_cand(a.pin, P.maxp, a.tl_ids_in + tile * MAXC, cnt, i, j, ux, uy, gx, gy);
}

// Particle block (all NT threads): contact force of particle p over all
// partners, then thread 0 runs the Newton-Euler update from state k (pin) into
// state k+1 (pout); with many particles the block then registers p in the
// candidate lists of the tiles it reaches.
__device__ __forceinline__ void particle_advance(const StepArgs &a, int p) {
    const DevP &P = a.P;
    const int m = P.maxp;
    __shared__ double srec[PCOUNT];
    __shared__ double sdpw;
    __shared__ ContactLDS cl;
    if (threadIdx.x < PCOUNT) srec[threadIdx.x] = a.pin[threadIdx.x * m + p];
    if (threadIdx.x == 0) sdpw = *a.dpw_in;
    __syncthreads();
    double fcx = 0.0, fcy = 0.0, dl = 0.0;
    int cnt = contact_block(P, a.pin + PX * m, a.pin + PY * m, a.pin + PDIAM * m, P.np, p, sdpw,
                            cl, fcx, fcy, a.pairs + p * MAXNB, &dl);
    if (threadIdx.x == 0) {
        double rec[PCOUNT];
        for (int f = 0; f < PCOUNT; ++f) rec[f] = srec[f];
        a.paircnt[p] = cnt;
        rec[PFCX] = fcx; rec[PFCY] = fcy;
        update_one(P, rec, a.kk_upd);
        for (int f = 0; f < PCOUNT; ++f) { a.pout[f * m + p] = rec[f]; srec[f] = rec[f]; }
        if (a.spf_out) make_spf(rec, a.spf_out + p * SPF);
        if (p == P.np - 1) *a.dpw_out = dl;
    }
    __syncthreads();
    if (P.np > NP_LOOP) register_tiles(P, srec, p, a.gen_out, a.tl_hdr_out, a.tl_ids_out, a.err);
}

// Hydrodynamic force and torque on particle p at step k (ComputeForce_HYDRO,
// F6:542-556) over the nodes that can have phi > 0, from populations
// rebuilt by the same pull-stream and bounce-back as the grid pass; a
// fixed-order block reduction keeps it deterministic.  Block `sub` of `nsub`
// takes every nsub-th group of NT nodes of the reach box.
template <class St, bool DEV, int BND>
__device__ void particle_force(const StepArgs &a, int p, int sub, int nsub, double (&F)[3]) {
    __shared__ double lds[3 * NT];
    const DevP &P = a.P;
    const int m = P.maxp;
    const double *pin = a.pin;
    double xc = 0.0, yc = 0.0, vx = 0.0, vy = 0.0, w = 0.0, dm = 0.0, zt = 0.0;
    float sv[SPF];
    int i0, i1, j0, j1;
    if constexpr (DEV) {
        // The FP32 profile record alone (one 48-byte load) gives the reach
        // box and everything the loop needs.  A node with phi > 0 lies more
        // than one lattice unit inside the box, so the FP32 rounding of its
        // edges cannot drop a contributing node.
        const float *s = a.spf_in + p * SPF;
#pragma unroll
        for (int k = 0; k < SPF; ++k) sv[k] = s[k];
        const int X = (int)sv[SX], Y = (int)sv[SY];
        i0 = max(0, X + (int)ceilf(sv[SXF] - sv[SREACH]));
        i1 = min(P.nx, X + (int)floorf(sv[SXF] + sv[SREACH]));
        j0 = max(0, Y + (int)ceilf(sv[SYF] - sv[SREACH]));
        j1 = min(P.ny, Y + (int)floorf(sv[SYF] + sv[SREACH]));
    } else {
        xc = pin[PX * m + p]; yc = pin[PY * m + p];
        vx = pin[PVX * m + p]; vy = pin[PVY * m + p]; w = pin[PW * m + p];
        dm = pin[PDIAM * m + p]; zt = pin[PZETA * m + p];
        double reach = 0.5 * dm + 0.5 * zt + 1.0;
        i0 = max(0, (int)ceil(xc - reach)); i1 = min(P.nx, (int)floor(xc + reach));
        j0 = max(0, (int)ceil(yc - reach)); j1 = min(P.ny, (int)floor(yc + reach));
    }
    int bw = i1 - i0 + 1, bh = j1 - j0 + 1;
    double acc[3] = {0.0, 0.0, 0.0};
    if (bw > 0 && bh > 0) {
        const typename St::T *src = (const typename St::T *)a.src;
        for (int t = sub * NT + threadIdx.x; t < bw * bh; t += nsub * NT) {
            int i = i0 + t % bw, j = j0 + t / bw;
            long long node = (long long)j * P.W + i;
            double r, u, v;
            if constexpr (DEV) {
                // same FP32 profile and rigid-body velocity as the grid pass
                float dx, dy;
                float phf = phi_f32(sv, i, j, dx, dy);
                if (!(phf > 0.0f)) continue;
                float h[9], rd, rf, uf, vf;
                load_fk_dev<St, BND>(P, src, (const float *)a.erho_src, node, i, j, h);
                macros_dev(h, P.frho0, rd, rf, uf, vf);
                float uxpf = sv[SVX] - sv[SW] * dy, uypf = sv[SVY] + sv[SW] * dx;
                double ph = phf;
                r = rf;
                double du = (double)(uf - uxpf), dv = (double)(vf - uypf);
                acc[0] = acc[0] + r * ph * du;
                acc[1] = acc[1] + r * ph * dv;
                acc[2] = acc[2] + r * ph * ((double)dx * dv - (double)dy * du);
                continue;
            } else {
                double ph = phi_at(xc, yc, dm, zt, i, j);
                if (!(ph > 0.0)) continue;
                double f[9];
                load_fk64<St, BND>(P, src, (const double *)a.erho_src, node, i, j, f);
                macros64(f, r, u, v);
                double uxp = vx - w * ((double)j - yc);
                double uyp = vy + w * ((double)i - xc);
                acc[0] = acc[0] + r * ph * (u - uxp);
                acc[1] = acc[1] + r * ph * (v - uyp);
                acc[2] = acc[2] + r * ph * (((double)i - xc) * (v - uyp) - ((double)j - yc) * (u - uxp));
            }
        }
    }
    block_sum<3, !DEV>(acc, lds);
    F[0] = acc[0]; F[1] = acc[1]; F[2] = acc[2];
}

// rec: this particle's record staged in LDS (field f at rec[f]), or null to
// read it from global memory (a round trip inside the particle chain).
__device__ __forceinline__ void hist_write(const StepArgs &a, int p, const double F[3],
                                           const double *rec) {
    if (!a.hist_part || threadIdx.x != 0) return;
    const int m = a.P.maxp;
    auto fld = [&](int f) { return rec ? rec[f] : a.pin[f * m + p]; };
    double *h = a.hist_part + (size_t)p * HPC;
    h[0] = fld(PX); h[1] = fld(PY); h[2] = fld(PANG);
    h[3] = fld(PVX); h[4] = fld(PVY); h[5] = fld(PW);
    h[6] = F[0]; h[7] = F[1]; h[8] = F[2];
    h[9] = fld(PFHX); h[10] = fld(PFHY); h[11] = fld(PFHW);
}

// Block totals of the monitors: FP64 mass as a fixed-point integer (so the
// per-step total is independent of the order in which it is summed) and the
// yielded-node count, stored as this tile's partials.  (Atomic adds of every
// block into a few shared counters serialised at the coherence point and
// doubled the FP16 step time of a 642k-node grid.)
template <bool DEV>
__device__ __forceinline__ void tile_monitors(const StepArgs &a, long long tile, double mass,
                                              int yielded, bool want_mass, bool want_yield) {
    __shared__ double lds[2 * NT];
    double v[2] = {mass, (double)yielded};
    block_sum<2, !DEV>(v, lds);
    if (threadIdx.x == 0) {
        if (want_mass)
            a.mpart_out[tile] = (unsigned long long)(long long)llrint(v[0] * MASS_SCALE);
        if (want_yield) a.mpart_out[a.ntiles + tile] = (unsigned long long)(long long)v[1];
    }
}

// Sum of n partials by one block: wrap-around integer addition, so any order
// gives the same total.
__device__ __forceinline__ unsigned long long block_sum_u64(const unsigned long long *x,
                                                            long long n) {
    __shared__ unsigned long long red[NT];
    unsigned long long acc = 0;
    for (long long t = threadIdx.x; t < n; t += NT) acc += x[t];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const unsigned long long r = red[0];
    __syncthreads();
    return r;
}

// Reducer block: the previous launch's monitor partials into its slots (the
// first slot of each; the others stay zero).
__device__ __forceinline__ void monitor_reduce(const unsigned long long *part, long long ntiles,
                                               unsigned long long *mass_slot,
                                               unsigned long long *yield_slot) {
    if (!part) return;
    if (mass_slot) {
        const unsigned long long m = block_sum_u64(part, ntiles);
        if (threadIdx.x == 0) *mass_slot = m;
    }
    if (yield_slot) {
        const unsigned long long y = block_sum_u64(part + ntiles, ntiles);
        if (threadIdx.x == 0) *yield_slot = y;
    }
}

// The partials of a call's last launch, reduced after it: one grid
// (arguments by value), or every simulation of a batch (block b: args[b]).
__global__ __launch_bounds__(NT) void k_monitor_final(StepArgs a) {
    monitor_reduce(a.mpart_out, a.ntiles, a.slots_mass, a.slots_yield);
}
__global__ __launch_bounds__(NT) void k_batch_monitor_final(const StepArgs *__restrict__ args) {
    const StepArgs &a = args[blockIdx.x];
    monitor_reduce(a.mpart_out, a.ntiles, a.slots_mass, a.slots_yield);
}

// Equilibrium edges: interior node (i, j) hands its density (FP64) or
// density deviation (reduced formats), and corrected velocity, to the edge
// arrays of the boundary nodes it is the source of, for the next launch.
template <class T>
__device__ __forceinline__ void eq_put(const StepArgs &a, const DevP &P, int i, int j, T r,
                                       T ux, T uy) {
    T *e = (T *)a.erho_dst;
    eq_sources(P, i, j, [&](int b) {
        e[b] = r; e[P.nbnd + b] = ux; e[2 * P.nbnd + b] = uy;
    });
}

// Prologue of a fused run: collide the FP64 state of step k (F, RHO, corrected
// UX/UY, G, OMEGA_BN) into f*_{k+1} in the storage format; particle blocks
// take the particles from state k to k+1.
template <class St, bool DEV>
__device__ __forceinline__ void prologue_tile(const StepArgs &a, long long tile) {
    const DevP &P = a.P;
    long long n = tile * NT + threadIdx.x;
    int yielded = 0;
    if (n < P.nn) {
        int j = row_of(P, (unsigned)n), i = (int)(n - (long long)j * P.W);
        typename St::T *dst = (typename St::T *)a.dst;
        double r = a.crho[n], ux = a.cux[n], uy = a.cuy[n];
        double gx = a.cg[n], gy = a.cg[P.plane + n];
        int b = (i == 0 || j == 0 || i == P.nx || j == P.ny) ? bnd_index(P, i, j) : -1;
        if constexpr (DEV) {
            float h[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                double fk = a.cf[k * P.plane + n];
                h[k] = (float)(fk - (double)wdev(k) * P.rho0);
            }
            // density deviation from the same populations
            float rd = ((h[0] + h[1]) + (h[2] + h[3])) + ((h[4] + h[5]) + (h[6] + h[7])) + h[8];
            float om = (float)(P.rheology != 0 ? a.cobn[n] : P.omega), om2, str[3];
            if (a.aux || a.slots_yield)
                collide_dev<true>(P, h, rd, (float)r, (float)ux, (float)uy, (float)gx, (float)gy,
                                  om, om2, str, yielded);
            else
                collide_dev<false>(P, h, rd, (float)r, (float)ux, (float)uy, (float)gx, (float)gy,
                                   om, om2, str, yielded);
            uint32_t nh = mix32((uint32_t)n ^ mix32(a.gen_out));
            const auto drs = make_rsrc(dst, P.buf_bytes);
#pragma unroll
            for (int k = 0; k < 9; ++k)
                St::bst(drs, h[k], (unsigned)n * (unsigned)sizeof(typename St::T),
                        (unsigned)(k * P.plane * (long long)sizeof(typename St::T)), sr_word(nh, k));
            if (P.rheology != 0) a.rate32[n] = om2;
            if (a.aux) {
                a.cstress[n] = str[0]; a.cstress[P.plane + n] = str[1];
                a.cstress[2 * P.plane + n] = str[2];
                if (P.rheology != 0) a.cimage[n] = yielded ? 0 : 5;
            }
            if (P.boundary == 2) {
                if (b < 0) eq_put(a, P, i, j, (float)(r - P.rho0), (float)ux, (float)uy);
            } else if (b >= 0) {
                ((float *)a.erho_dst)[b] = (float)r;
            }
        } else {
            double f[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) f[k] = a.cf[k * P.plane + n];
            CollideOut o;
            double om = P.rheology != 0 ? a.cobn[n] : P.omega;
            collide64(P, f, r, ux, uy, gx, gy, om, o);
#pragma unroll
            for (int k = 0; k < 9; ++k) St::st(dst + k * P.plane + n, f[k], 0u);
            if (P.rheology != 0) a.cobn[n] = o.omega_bn;
            yielded = o.yielded;
            if (a.aux) {
                a.cstress[n] = o.stress[0]; a.cstress[P.plane + n] = o.stress[1];
                a.cstress[2 * P.plane + n] = o.stress[2];
                if (P.rheology != 0) a.cimage[n] = o.yielded ? 0 : 5;
            }
            if (P.boundary == 2) {
                if (b < 0) eq_put(a, P, i, j, r, ux, uy);
            } else if (b >= 0) {
                ((double *)a.erho_dst)[b] = r;
            }
        }
    }
    if (a.slots_yield) tile_monitors<DEV>(a, tile, 0.0, yielded, false, true);
}

template <class St, bool DEV>
__global__ __launch_bounds__(NT) void k_prologue(StepArgs a) {
    if ((int)blockIdx.x < a.n_pblocks) {
        if ((int)blockIdx.x % a.nsub == 0) particle_advance(a, (int)blockIdx.x / a.nsub);
        return;
    }
    if ((int)blockIdx.x < a.n_pblocks + a.nred) return;  // nothing before the prologue
    prologue_tile<St, DEV>(a, xcd_tile(blockIdx.x - a.n_pblocks - a.nred, a.ntiles));
}

// DOMAIN_IMAGE from the particles alone (F6:438, F6:460-467).
__device__ __forceinline__ int image_node(const DevP &P, const double *part, int i, int j) {
    int img = 5;
    const int m = P.maxp;
    for (int p = 0; p < P.np; ++p) {
        if (!in_reach(part, m, p, i, j)) continue;
        double ph = phi_at(part[PX * m + p], part[PY * m + p], part[PDIAM * m + p],
                           part[PZETA * m + p], i, j);
        if (ph >= 0.5) {
            img = 10;
            double rp = 0.5 * part[PDIAM * m + p], tp = part[PANG * m + p];
            double xpr = part[PX * m + p] + 0.75 * rp * cos(tp);
            double ypr = part[PY * m + p] + 0.75 * rp * sin(tp);
            double dp = 25.0 - (((double)i - xpr) * ((double)i - xpr) +
                                ((double)j - ypr) * ((double)j - ypr));
            if (dp >= 0) img = 20;
        }
    }
    return img;
}

// One fused time step k over the grid: pull-stream + bounce-back (f_k),
// density and predicted velocity, smoothed-profile forcing G_k, velocity
// correction, then (FINAL = false) the MRT / Papanastasiou collision of step
// k+1 stored as f*_{k+1}, or (FINAL = true) the FP64 state of step k.
// Particle blocks compute the hydrodynamic force of step k and advance the
// particles to state k+1.
template <class St, bool DEV, bool FINAL, int BND>
__device__ void particle_block(const StepArgs &a, int pb) {
    const DevP &P = a.P;
    {
        int p = pb / a.nsub, sub = pb - p * a.nsub;
        const int m = P.maxp;
        // Stage the inputs of the update in LDS: this particle's record, the
        // contact inputs of all particles (NP_LDS <= NT: one per thread) and
        // the stale wall distance.  In the reduced-precision steps they are
        // loaded into registers before the force gather and written to LDS
        // after it, so their round trip overlaps the gather's instead of
        // preceding it (a load feeding an LDS store is waited for at the
        // store); the thread that finishes the update reads LDS instead of
        // global memory.
        static_assert(NP_LDS <= NT, "one staged particle per thread");
        __shared__ double srec[PCOUNT];
        __shared__ double scon[3 * NP_LDS];
        __shared__ double sdpw;
        const bool staged = !FINAL && P.np <= NP_LDS;
        double r_rec = 0.0, r_x = 0.0, r_y = 0.0, r_d = 0.0, r_dpw = 0.0;
        if (!FINAL) {
            const int t = threadIdx.x;
            if (t < PCOUNT) r_rec = a.pin[t * m + p];
            if (staged && t < P.np) {
                r_x = a.pin[PX * m + t]; r_y = a.pin[PY * m + t]; r_d = a.pin[PDIAM * m + t];
            }
            if (t == 0) r_dpw = *a.dpw_in;
        }
        auto stage = [&]() {
            const int t = threadIdx.x;
            if (t < PCOUNT) srec[t] = r_rec;
            if (staged && t < P.np) {
                scon[t] = r_x; scon[NP_LDS + t] = r_y; scon[2 * NP_LDS + t] = r_d;
            }
            if (t == 0) sdpw = r_dpw;
        };
        // The FP64 step is bound by its tiles' HBM traffic, not by this
        // chain: it stages before the gather (the barriers of the gather's
        // block sum publish the LDS) and keeps the registers for the gather
        // (staging after it raised its allocation from 98 to 106 VGPRs and
        // its step time by 3 %).
        if constexpr (!DEV)
            if (!FINAL) stage();
        double F[3];
        particle_force<St, DEV, BND>(a, p, sub, a.nsub, F);  // ends with block barriers
        if constexpr (DEV)
            if (!FINAL) {
                stage();
                if (a.nsub <= 1) __syncthreads();  // else the barrier after the partial sum
            }
        if (a.nsub > 1) {
            // partial of this block; the last block of the particle adds all
            // partials in sub order and carries on with the update.
            // The partials are written and read with agent-scope atomic
            // accesses, which bypass the per-XCD L2, and the counter add is
            // issued only after those stores have completed (vmcnt drained).
            // A __threadfence() here would instead write back every dirty
            // line of this XCD's L2 (the tile stores of the whole launch) on
            // every particle block; that cost ~40 us per step at 504 disks.
            __shared__ int last;
            if (threadIdx.x == 0) {
                double *pf = a.pforce + ((size_t)p * a.nsub + sub) * 3;
                for (int c = 0; c < 3; ++c)
                    __hip_atomic_store(pf + c, F[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                last = __hip_atomic_fetch_add(a.pcount + p, 1, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT) == a.nsub - 1;
                if (last) {
                    const double *q = a.pforce + (size_t)p * a.nsub * 3;
                    F[0] = 0.0; F[1] = 0.0; F[2] = 0.0;
                    for (int s = 0; s < a.nsub; ++s)
                        for (int c = 0; c < 3; ++c)
                            F[c] += __hip_atomic_load(q + 3 * s + c, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(a.pcount + p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            __syncthreads();
            if (!last) return;
        }
        hist_write(a, p, F, FINAL ? nullptr : srec);  // srec is staged and published here
        if (threadIdx.x == 0) {
            if (FINAL) {
                a.pout[PFX * m + p] = F[0]; a.pout[PFY * m + p] = F[1]; a.pout[PTQ * m + p] = F[2];
            }
        }
        if (!FINAL) {
            // contact force over all partners (one wavefront up to 64
            // disks, the whole block beyond), then the update, which uses
            // FORCE_k: written into the outgoing record first
            __shared__ ContactLDS cl;
            double fcx = 0.0, fcy = 0.0, dl = 0.0;
            int cnt = 0;
            if (P.np <= 64) {
                // one wavefront suffices (the staged copy holds every disk)
                if (threadIdx.x < 64)
                    cnt = contact_wave(P, scon, scon + NP_LDS, scon + 2 * NP_LDS, P.np, p, sdpw,
                                       fcx, fcy, a.pairs + p * MAXNB, &dl);
            } else {
                cnt = staged
                    ? contact_block(P, scon, scon + NP_LDS, scon + 2 * NP_LDS, P.np, p, sdpw, cl,
                                    fcx, fcy, a.pairs + p * MAXNB, &dl)
                    : contact_block(P, a.pin + PX * m, a.pin + PY * m, a.pin + PDIAM * m, P.np,
                                    p, sdpw, cl, fcx, fcy, a.pairs + p * MAXNB, &dl);
            }
            if (threadIdx.x < 64) {
                // the update by wavefront 0; FORCE_k and the contact force
                // are final in thread 0 (lane 0): broadcast them
                const double val =
                    update_wave(P, srec, __shfl(F[0], 0), __shfl(F[1], 0), __shfl(F[2], 0),
                                __shfl(fcx, 0), __shfl(fcy, 0), a.kk_upd);
                if (threadIdx.x < PCOUNT) {
                    a.pout[threadIdx.x * m + p] = val;
                    srec[threadIdx.x] = val;
                }
                if (threadIdx.x == 0) {
                    a.paircnt[p] = cnt;
                    if (p == P.np - 1) *a.dpw_out = dl;
                }
            }
            __syncthreads();
            if (threadIdx.x == 0 && a.spf_out) make_spf(srec, a.spf_out + p * SPF);
            if (P.np > NP_LOOP)
                register_tiles(P, srec, p, a.gen_out, a.tl_hdr_out, a.tl_ids_out, a.err);
        }
    }
}

template <class St, bool DEV, bool FINAL, int BND>
__device__ __forceinline__ void step_tile(const StepArgs &a, long long tile) {
    const DevP &P = a.P;
    long long n = tile * NT + threadIdx.x;
    double mass = 0.0;
    int yielded = 0;
    if (n < P.nn) {
        int j = row_of(P, (unsigned)n), i = (int)(n - (long long)j * P.W);
        const typename St::T *src = (const typename St::T *)a.src;
        int b = (i == 0 || j == 0 || i == P.nx || j == P.ny) ? bnd_index(P, i, j) : -1;
        if constexpr (DEV) {
            float h[9], rd, rho, ux, uy;
            load_fk_dev<St, BND>(P, src, (const float *)a.erho_src, n, i, j, h);
            float om = (!FINAL && P.rheology != 0) ? a.rate32[n] : P.fomega;
            // particle test after the population loads are in flight
            const bool any = tile_has_particles_f32(a, tile);
            macros_dev(h, P.frho0, rd, rho, ux, uy);
            float gx, gy;
            spm_tile_node_f32(a, tile, any, i, j, ux, uy, gx, gy);
            float ucx = ux + 0.5f * gx, ucy = uy + 0.5f * gy;
            mass = (double)P.rho0 + (double)rd;
            if (FINAL) {
                const typename St::T *s0 = src;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    double w0 = (double)wdev(k) * P.rho0;
                    a.cf[k * P.plane + n] = w0 + (double)h[k];
                    a.cfp[k * P.plane + n] = w0 + (double)St::ld(s0 + k * P.plane + n);
                }
                a.crho[n] = mass; a.cux[n] = ucx; a.cuy[n] = ucy;
                a.cg[n] = gx; a.cg[P.plane + n] = gy;
                if (P.rheology != 0) a.cobn[n] = a.rate32[n];
                if (a.aux && P.rheology == 0) a.cimage[n] = image_node(P, a.pin, i, j);
            } else {
                typename St::T *dst = (typename St::T *)a.dst;
                float om2, str[3];
                if (a.aux || a.slots_yield)
                    collide_dev<true>(P, h, rd, rho, ucx, ucy, gx, gy, om, om2, str, yielded);
                else
                    collide_dev<false>(P, h, rd, rho, ucx, ucy, gx, gy, om, om2, str, yielded);
                uint32_t nh = mix32((uint32_t)n ^ mix32(a.gen_out));
                const auto drs = make_rsrc(dst, P.buf_bytes);
#pragma unroll
                for (int k = 0; k < 9; ++k)
                    St::bst(drs, h[k], (unsigned)n * (unsigned)sizeof(typename St::T),
                            (unsigned)(k * P.plane * (long long)sizeof(typename St::T)),
                            sr_word(nh, k));
                if (P.rheology != 0) a.rate32[n] = om2;
                if (a.aux) {
                    a.cstress[n] = str[0]; a.cstress[P.plane + n] = str[1];
                    a.cstress[2 * P.plane + n] = str[2];
                    if (P.rheology != 0) a.cimage[n] = yielded ? 0 : 5;
                }
                if ((BND >= 0 ? BND : P.boundary) == 2) {
                    if (b < 0) eq_put(a, P, i, j, rd, ucx, ucy);
                } else if (b >= 0) {
                    ((float *)a.erho_dst)[b] = rho;
                }
            }
        } else {
            double f[9], rho, ux, uy;
            load_fk64<St, BND>(P, src, (const double *)a.erho_src, n, i, j, f);
            const bool any = tile_has_particles(a, tile);
            macros64(f, rho, ux, uy);
            double gx, gy;
            spm_tile_node(a, tile, any, i, j, ux, uy, gx, gy);
            double ucx = ux + 0.5 * 1.0 * (gx + 0.0), ucy = uy + 0.5 * 1.0 * (gy + 0.0);
            mass = rho;
            if (FINAL) {
                double *cfk = (double *)a.dst;
#pragma unroll
                for (int k = 0; k < 9; ++k) cfk[k * P.plane + n] = f[k];
                a.crho[n] = rho; a.cux[n] = ucx; a.cuy[n] = ucy;
                a.cg[n] = gx; a.cg[P.plane + n] = gy;
                if (a.aux && P.rheology == 0) a.cimage[n] = image_node(P, a.pin, i, j);
            } else {
                typename St::T *dst = (typename St::T *)a.dst;
                CollideOut o;
                double om = P.rheology != 0 ? a.cobn[n] : P.omega;
                collide64(P, f, rho, ucx, ucy, gx, gy, om, o);
#pragma unroll
                for (int k = 0; k < 9; ++k) St::st(dst + k * P.plane + n, f[k], 0u);
                if (P.rheology != 0) a.cobn[n] = o.omega_bn;
                yielded = o.yielded;
                if (a.aux) {
                    a.cstress[n] = o.stress[0]; a.cstress[P.plane + n] = o.stress[1];
                    a.cstress[2 * P.plane + n] = o.stress[2];
                    if (P.rheology != 0) a.cimage[n] = o.yielded ? 0 : 5;
                }
                if ((BND >= 0 ? BND : P.boundary) == 2) {
                    if (b < 0) eq_put(a, P, i, j, rho, ucx, ucy);
                } else if (b >= 0) {
                    ((double *)a.erho_dst)[b] = rho;
                }
            }
        }
    }
    if (a.slots_mass || (!FINAL && a.slots_yield))
        tile_monitors<DEV>(a, tile, mass, yielded, a.slots_mass != nullptr,
                      !FINAL && a.slots_yield != nullptr);
}

template <class St, bool DEV, bool FINAL, int BND>
__global__ __launch_bounds__(NT) void k_step(StepArgs a) {
    if ((int)blockIdx.x < a.n_pblocks) {
        particle_block<St, DEV, FINAL, BND>(a, (int)blockIdx.x);
        return;
    }
    if ((int)blockIdx.x < a.n_pblocks + a.nred) {
        monitor_reduce(a.mpart_in, a.ntiles, a.red_mass, a.red_yield);
        return;
    }
    step_tile<St, DEV, FINAL, BND>(a, xcd_tile(blockIdx.x - a.n_pblocks - a.nred, a.ntiles));
}

// ================================================================= batch
// Many independent grids in one launch.  bmap[blockIdx.x] = (simulation,
// block within it); each simulation's arguments for this launch's step
// parity are read in place from args[] (k_batch_fields writes their per-step
// fields just before the launch, from the step index s = s_const + *s_dev, so
// one captured graph serves every chunk of steps).  A tile block handles tpb
// consecutive 256-node tiles.
__global__ void k_batch_fields(StepArgs *args, int n, const int *s_dev, int s_const,
                               int nsteps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    step_fields(args[i], s_const + (s_dev ? *s_dev : 0), nsteps);
}

template <class St, bool DEV>
__global__ __launch_bounds__(NT) void k_batch_prologue(const StepArgs *__restrict__ args,
                                                       const int2 *__restrict__ bmap) {
    const int2 m = bmap[blockIdx.x];
    const StepArgs &a = args[m.x];
    if (m.y < a.n_pblocks) {
        if (m.y % a.nsub == 0) particle_advance(a, m.y / a.nsub);
        return;
    }
    if (m.y < a.n_pblocks + a.nred) return;  // nothing before the prologue
    const long long tb = m.y - a.n_pblocks - a.nred;
    for (int r = 0; r < a.tpb; ++r) {
        const long long t = tb * a.tpb + r;
        if (t < a.ntiles) prologue_tile<St, DEV>(a, t);
    }
}

template <class St, bool DEV, bool FINAL, int BND>
__global__ __launch_bounds__(NT) void k_batch_step(const StepArgs *__restrict__ args,
                                                   const int2 *__restrict__ bmap) {
    const int2 m = bmap[blockIdx.x];
    const StepArgs &a = args[m.x];
    if (m.y < a.n_pblocks) {
        particle_block<St, DEV, FINAL, BND>(a, m.y);
        return;
    }
    if (m.y < a.n_pblocks + a.nred) {
        monitor_reduce(a.mpart_in, a.ntiles, a.red_mass, a.red_yield);
        return;
    }
    const long long tb = m.y - a.n_pblocks - a.nred;
    for (int r = 0; r < a.tpb; ++r) {
        const long long t = tb * a.tpb + r;
        if (t < a.ntiles) step_tile<St, DEV, FINAL, BND>(a, t);
    }
}

// Step counter of captured chunks: one thread adds the chunk length.
__global__ void k_counter_add(int *c, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c += v;
}
__global__ void k_counter_set(int *c, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = v;
}

}  // namespace lbm
