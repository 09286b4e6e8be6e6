// Kernels of the D2Q9 MRT / smoothed-profile time step for gfx950.
//
// Drop-in kernels (one per reference subroutine) work on the FP64 state in
// the reference's meaning; the fused kernels run the whole time step in one
// pass over the grid.  Layout in HBM: structure of arrays, one plane per
// population (plane stride P.plane elements), node index j*(nx+1)+i.
#pragma once
#include <cstring>

#include "lbm_device.h"

namespace lbm {

// Tiles are handed to the 8 XCDs in contiguous ranges so that rows shared
// between neighbouring tiles meet in the same L2.  Blocks are dispatched to
// the XCDs round-robin by block index (block g runs on XCD g % 8).  For the
// m blocks base .. base + m - 1 of a launch: the position of block g in the
// order that gives XCD 0 the first n_0 positions, XCD 1 the next n_1, ...
// (n_x: how many of those blocks XCD x runs), each XCD's blocks in dispatch
// order.  A bijection of [0, m).  With m = 8 q + rem, the offsets from base
// that XCD y runs are those congruent to (y - base) mod 8, q + 1 of them when
// that residue is below rem and q otherwise; the XCDs with the extra block
// are base mod 8 .. base mod 8 + rem - 1 (cyclically), bits of the mask M,
// so XCD x starts at x q + popcount(M & ((1 << x) - 1)).
__host__ __device__ __forceinline__ int xcd_rank(int g, int base, int m) {
    const unsigned x = (unsigned)g & 7u, b8 = (unsigned)base & 7u;
    const unsigned low = (1u << ((unsigned)m & 7u)) - 1u;
    const unsigned M = ((low << b8) | (low >> (8u - b8))) & 0xffu;
    const int start = (int)x * (m >> 3) + __builtin_popcount(M & ((1u << x) - 1u));
    return start + ((g - base - (int)((x - b8) & 7u)) >> 3);
}

// The arguments of a launch, read in place from the kernel argument segment
// through a pointer the compiler cannot see through.  A kernel that names
// its by-value argument struct directly gets all the fields it uses loaded
// at entry (they are invariant), which for the fused step (1 KiB of
// arguments, two block roles) keeps more scalar registers live than there
// are and spills them to vector registers; read through this pointer, each
// field is loaded where it is used.  The struct must be the kernel's only
// explicit argument (it then starts the segment).
template <class T>
__device__ __forceinline__ const T &launch_args() {
    typedef const __attribute__((address_space(4))) T CT;
    CT *p = (CT *)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return *(const T *)p;
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
// Deviations of the canonical F in storage units (scale S), as a prologue
// derives them from F.
__global__ void k_hexact(DevP P, const double *f, float S, float *hx) {
    const long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= P.nn) return;
    for (int k = 0; k < 9; ++k)
        hx[k * P.plane + n] = (float)((f[k * P.plane + n] - (double)wdev(k) * P.rho0) * (double)S);
}

__global__ void k_soa_to_aos(long long nn, int nc, const double *soa, double *aos,
                             long long plane) {
    long long n = (long long)blockIdx.x * NT + threadIdx.x;
    if (n >= nn) return;
    for (int c = 0; c < nc; ++c) aos[n * nc + c] = soa[c * plane + n];
}

// ================================================================= fused
// One launch per time step.  The launches of a fused run form a chain
// c = 0, 1, 2, ...; chain launch c >= 1 runs loop iteration k (its call's
// kk0 + s):
//  * tile blocks pull-stream f*_k into f_k with the boundary rules, form
//    rho and u, the SPM forcing G from particle state k (complete in memory
//    since the launch before), correct u and collide into f*_{k+1}, the
//    collision of iteration k+1;
//  * particle helper blocks (dispatched before the tiles) each gather one
//    chunk of a particle's reach box: f_k, rho and u at its nodes exactly as
//    the owning tile forms them, and the chunk's share of the hydrodynamic
//    force and torque (F6:542-556) as tagged words; the block that completes
//    a particle's count adds the chunks in a fixed order into FORCE_k,
//    computes the contact force on state k (F3:739-804) and runs the
//    Newton-Euler update to state k+1 (F6:299-313), which the next launch
//    reads.
// Chain launch 0 (the prologue) collides the canonical FP64 state into the
// storage format while its particle blocks advance state kk0 to kk0+1 with
// the record's FORCE.  A FINAL launch takes the inputs of a chain launch c
// and, instead of colliding, writes the reference's loop-end state of
// iteration k into the canonical FP64 arrays (F, F_POST, RHO, UX, UY, G,
// OMEGA_BN, STRESS, DOMAIN_IMAGE and FORCE_k in the particle record).  The
// storage populations, node rates, edge arrays and particle state are
// double-buffered by chain parity, so a later call continues the chain
// where the last launch left it, and the inputs of that last launch stay
// intact for a FINAL launch that materialises its iteration on demand.
// The fields a tile block needs before its first population load, copied
// into the first 128 bytes of the launch arguments (finish_head) so that a
// wavefront reads them with two scalar loads at entry instead of a chain of
// dependent ones: the block layout (helper blocks, first tile block, tiles),
// the grid geometry of the row division and the interior test, the node-rate
// and population inputs with their byte sizes, the stochastic-rounding
// position, and the byte offsets of the nine pulled populations of an
// interior node relative to the node's own.
struct StepHead {
    int n_pblocks, tile0, ntiles, np;
    unsigned nn, W, w_magic, w_shift;
    int nx, ny, margin, rheology;   // margin: 1, or 2 with equilibrium edges
    unsigned srh, buf_bytes, rate_bytes, plane_bytes;  // srh: mix32(sr), node_hash_h
    const void *src, *rate_in;
    unsigned pull[9];
    unsigned spare[3];
};
static_assert(sizeof(StepHead) == 128, "two 16-dword scalar loads");

// What a reduced-precision tile's collision and stores read besides the
// header: the FP32 collision parameters of DevP under the same names (so
// collide_dev reads either), the output flags (bit 0: the collision forms
// stress and yield flags, aux or want_yield; bit 1: aux) and the two
// destination buffers.  A single-simulation tile reads it with two scalar
// loads before its collision, so the collision waits on no other scalar
// load.
struct CollideHead {
    int q_mode, rheology;
    float cD1a, cD1b, cD2a, cD2b, cC3, cC5, cDq1, cDq2, cS1, cS2;
    float fnu, fgamma_min2, fmu0, ftau_y, fm_pap, fsqrt_mu0, fcas_lim;
    unsigned out;
    void *dst, *rate_out;
};
static_assert(sizeof(CollideHead) == 96, "a 16-dword and an 8-dword scalar load");

struct StepArgs {
    StepHead hd;                  // copies of fields below (finish_head)
    CollideHead ch;               // likewise
    DevP P;
    const void *src;              // post-collision populations f*_k (storage format)
    void *dst;                    // f*_{k+1} (storage format)
    double *cf, *cfp;             // canonical FP64 state arrays F, F_POST
    double *crho, *cux, *cuy, *cg, *cobn, *cstress;
    int *cimage;
    const void *rate_in;          // node rates of iteration k (FP32 or FP64 as the storage)
    void *rate_out;               // node rates set by the collision of iteration k+1
    // STRESS and DOMAIN_IMAGE of the collisions, by iteration parity: the
    // last two collisions of every call write them (xstress, ximage), a
    // FINAL launch copies iteration k's into the canonical arrays
    double *astress[2];
    int *aimage[2];
    double *xstress;
    int *ximage;
    const double *astress_in;
    const int *aimage_in;
    // edge arrays: previous-step density on the boundary nodes or, with
    // equilibrium edges, density and corrected velocity of each boundary
    // node's interior source (3 x nbnd)
    const void *erho_src;
    void *erho_dst;
    const double *pin;            // particle state k (SoA records)
    double *pout;                 // particle state k+1 (FINAL: same as pin)
    const float *spf_in;          // FP32 profile records of state k (reduced formats)
    float *spf_out;               // those of state k+1
    // reduced formats: deviations of the canonical F in storage units, exact
    // (FINAL writes them beside F; a prologue starts from them when valid)
    float *hx;
    const float *hx_in;
    const int *pbox_in;           // reach boxes of state k [maxp][4]: i0, i1, j0, j1
    int *pbox_out;                // those of state k+1
    unsigned long long *fslot;    // tagged chunk force partials [maxp][kpart][SLOTW]
    int kpart;                    // helper blocks (box chunks of NT nodes) per particle
    const double *dpw_in;
    double *dpw_out;
    int *pairs, *paircnt;         // contact pairs found by this launch
    double *hist_part;            // history row of step k, or null
    unsigned long long *slots_mass;   // NSLOT fixed-point mass sums of step k, or null
    unsigned long long *slots_yield;  // NSLOT yielded-node counts of the collide's step, or null
    int want_yield;               // the tiles store yielded-count partials
    int n_pblocks;
    int kk_upd;                   // loop iteration of the particle update in this launch
    long long ntiles;
    int aux;                      // write STRESS / DOMAIN_IMAGE (FINAL: copy them)
    int fin;                      // a FINAL launch
    // per-wavefront candidate lists (more than NP_LOOP particles): read for this
    // launch (generation list_gen), written for the next one (list_gen + 1)
    const unsigned long long *tl_hdr_in;
    const int *tl_ids_in;
    unsigned long long *tl_hdr_out;
    int *tl_ids_out;
    unsigned tag;                 // this launch's tag of the force partial words (unique)
    unsigned list_gen;            // generation of the tile lists this launch reads
    unsigned sr;                  // stochastic-rounding sequence position of its collide
    int *err;                     // LBM_EF_* bits
    // per-call bases from which step_fields() derives the per-step fields
    double *hist_base;            // history rows of the call, or null
    unsigned long long *slots_base;  // monitor slots of the call, or null
    int rowp;                     // history doubles per step (particles)
    int kk0;                      // step index before the call
    unsigned tag0, lg0, sr0;      // tag, list generation, rounding position at s = 0
    int c0par;                    // chain parity of s = 0
    int tpb;                      // 256-node tiles per block (batched launches)
    // monitor partials: each tile block stores its fixed-point mass and its
    // yielded count (plain stores, [mass | yield] x ntiles, two buffers by
    // chain parity); a reducer block of the next launch sums the previous
    // launch's partials into that launch's monitor slots
    unsigned long long *mpart[2];
    unsigned long long *mpart_out;        // this launch's partials, or null
    const unsigned long long *mpart_in;   // the previous launch's, or null
    unsigned long long *red_mass, *red_yield;  // where their sums go (null: none)
    int nred;                     // reducer block slots in the block layout (0 or 1)
};

// The header of a launch's arguments, read with two 16-dword scalar loads.
__device__ __forceinline__ StepHead load_head(const StepArgs &a) {
    typedef unsigned V16 __attribute__((ext_vector_type(16)));
    const __attribute__((address_space(4))) V16 *p =
        (const __attribute__((address_space(4))) V16 *)&a.hd;
    struct {
        V16 lo, hi;
    } v = {p[0], p[1]};
    return __builtin_bit_cast(StepHead, v);
}

__device__ __forceinline__ CollideHead load_collide(const StepArgs &a) {
    typedef unsigned V16 __attribute__((ext_vector_type(16)));
    typedef unsigned V8 __attribute__((ext_vector_type(8)));
    typedef const __attribute__((address_space(4))) char CC;
    CC *p = (CC *)&a.ch;
    const V16 lo = *(const __attribute__((address_space(4))) V16 *)p;
    const V8 hi = *(const __attribute__((address_space(4))) V8 *)(p + 64);
    unsigned w[24];
#pragma unroll
    for (int k = 0; k < 16; ++k) w[k] = lo[k];
#pragma unroll
    for (int k = 0; k < 8; ++k) w[16 + k] = hi[k];
    return __builtin_bit_cast(CollideHead, w);
}

// j = n / W from the header (row_of).
__device__ __forceinline__ int row_of(const StepHead &h, unsigned n) {
    unsigned t = __umulhi(h.w_magic, n);
    return (int)((t + ((n - t) >> 1)) >> (h.w_shift - 1u));
}

// error bits of the fused path
enum {
    LBM_EF_LIST = 1,      // a wavefront candidate list overflowed
    LBM_EF_NAN = 2,       // a non-finite density or particle state
    LBM_EF_WAIT = 4,      // force partials did not appear within the wait bound
    LBM_EF_SLOT = 16,     // a reach box holds more nodes than its helper blocks cover
};

// Per-step fields of launch s of a call of nsteps steps (s = 0: the
// prologue; a.fin: a FINAL launch, which materialises iteration kk0 + s):
// the tag, list generation, rounding position, loop iteration of the
// particle update, the history row and monitor slots, and whether the
// launch writes the output-only fields.  History row r holds iteration
// kk0 + r + 1: its mass comes from launch r + 1, its yielded count from the
// collision in launch r (the yielded count of the last launch's collision
// belongs to the next call's first row).
__host__ __device__ __forceinline__ void step_fields(StepArgs &a, int s, int nsteps) {
    const bool rec = a.slots_base != nullptr;
    a.tag = a.tag0 + (unsigned)s;
    a.list_gen = a.lg0 + (unsigned)s;
    a.sr = a.sr0 + (unsigned)s;
    a.kk_upd = a.kk0 + s + 1;
    const size_t row2 = 2 * NSLOT;
    if (a.fin) {
        a.hist_part = (rec && s > 0) ? a.hist_base + (size_t)(s - 1) * (size_t)a.rowp : nullptr;
        a.slots_mass = (rec && s > 0) ? a.slots_base + (size_t)(s - 1) * row2 : nullptr;
        a.slots_yield = nullptr;
        a.want_yield = 0;
        a.aux = 1;
    } else if (s == 0) {
        a.hist_part = nullptr;
        a.slots_mass = nullptr;
        a.slots_yield = rec ? a.slots_base + NSLOT : nullptr;
        a.want_yield = rec;
        a.aux = nsteps <= 1;
    } else {
        a.hist_part = a.hist_base ? a.hist_base + (size_t)(s - 1) * (size_t)a.rowp : nullptr;
        a.slots_mass = rec ? a.slots_base + (size_t)(s - 1) * row2 : nullptr;
        a.slots_yield = (rec && s < nsteps) ? a.slots_base + (size_t)s * row2 + NSLOT : nullptr;
        a.want_yield = rec;
        a.aux = s >= nsteps - 1;
    }
    const int ax = (a.kk0 + s + 1) & 1;  // parity of the iteration this launch collides
    a.xstress = a.astress[ax];
    a.ximage = a.aimage[ax];
    a.astress_in = a.astress[ax ^ 1];
    a.aimage_in = a.aimage[ax ^ 1];
    const int par = (a.c0par + s) & 1;
    a.mpart_out = rec ? a.mpart[par] : nullptr;
    a.mpart_in = (rec && s > 0) ? a.mpart[par ^ 1] : nullptr;
    a.red_mass = (rec && s > 1) ? a.slots_base + (size_t)(s - 2) * row2 : nullptr;
    a.red_yield = (rec && s > 0) ? a.slots_base + (size_t)(s - 1) * row2 + NSLOT : nullptr;
    a.hd.srh = mix32(a.sr);
    a.ch.out = ((a.aux || a.want_yield) ? 1u : 0u) | (a.aux ? 2u : 0u);
}

// Fill the header from the other fields (host, after the last change to
// them); elem: bytes per stored population.
inline void finish_head(StepArgs &a, unsigned elem) {
    const DevP &P = a.P;
    StepHead &h = a.hd;
    std::memset(&h, 0, sizeof h);
    h.n_pblocks = a.n_pblocks;
    h.tile0 = a.n_pblocks + a.nred;
    h.ntiles = (int)a.ntiles;
    h.np = P.np;
    h.nn = (unsigned)P.nn;
    h.W = (unsigned)P.W;
    h.w_magic = P.w_magic;
    h.w_shift = (unsigned)P.w_shift;
    h.nx = P.nx;
    h.ny = P.ny;
    h.margin = P.boundary == 2 ? 2 : 1;
    h.rheology = P.rheology;
    h.srh = mix32(a.sr);
    h.buf_bytes = P.buf_bytes;
    h.rate_bytes = (unsigned)(P.plane * (elem == 8 ? 8 : 4));
    h.plane_bytes = (unsigned)(P.plane * elem);
    h.src = a.src;
    h.rate_in = a.rate_in;
    const int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    const int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    for (int k = 0; k < 9; ++k)
        h.pull[k] = (unsigned)((k * P.plane - cx[k] - (long long)cy[k] * P.W) * (long long)elem);
    CollideHead &c = a.ch;
    std::memset(&c, 0, sizeof c);
    c.q_mode = P.q_mode;
    c.rheology = P.rheology;
    c.cD1a = P.cD1a; c.cD1b = P.cD1b; c.cD2a = P.cD2a; c.cD2b = P.cD2b;
    c.cC3 = P.cC3; c.cC5 = P.cC5; c.cDq1 = P.cDq1; c.cDq2 = P.cDq2;
    c.cS1 = P.cS1; c.cS2 = P.cS2;
    c.fnu = P.fnu; c.fgamma_min2 = P.fgamma_min2; c.fmu0 = P.fmu0; c.ftau_y = P.ftau_y;
    c.fm_pap = P.fm_pap; c.fsqrt_mu0 = P.fsqrt_mu0; c.fcas_lim = P.fcas_lim;
    c.out = ((a.aux || a.want_yield) ? 1u : 0u) | (a.aux ? 2u : 0u);
    c.dst = a.dst;
    c.rate_out = a.rate_out;
}

__device__ __forceinline__ void set_err(const StepArgs &a, int bits) { atomicOr(a.err, bits); }

// A 64-bit value every lane holds alike, as a wavefront-uniform one.
__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// history row of one particle: record fields of the state, its FORCE, F_H_IN
__device__ __forceinline__ void hist_row(double *h, const double *rec, const double F[3]) {
    h[0] = rec[PX]; h[1] = rec[PY]; h[2] = rec[PANG];
    h[3] = rec[PVX]; h[4] = rec[PVY]; h[5] = rec[PW];
    h[6] = F[0]; h[7] = F[1]; h[8] = F[2];
    h[9] = rec[PFHX]; h[10] = rec[PFHY]; h[11] = rec[PFHW];
}

// The reach box of a state stored for the tiles of the next launch
// (i0, i1, j0, j1; empty when i0 > i1).
__device__ __forceinline__ void put_box(int *dst, const Box &b) {
    dst[0] = b.i0; dst[1] = b.i1; dst[2] = b.j0; dst[3] = b.j1;
}

// Prologue particle block: contact force on state kk0, the update with the
// record's FORCE (the reference's next loop iteration, F6:299-313), the new
// state's box, profile record and tile registration.
template <bool DEV>
__device__ void particle_advance(const StepArgs &a, int p) {
    const DevP &P = a.P;
    const int m = P.maxp;
    __shared__ double srec[PCOUNT];
    __shared__ double scon[3 * NP_LDS];
    __shared__ double sdpw;
    __shared__ ContactLDS cl;
    const int t = threadIdx.x;
    const bool staged = P.np <= NP_LDS;
    const int tr = min(t, PCOUNT - 1), tc = min(t, max(P.np - 1, 0));
    const double r_rec = a.pin[tr * m + p];
    const double r_x = a.pin[PX * m + tc], r_y = a.pin[PY * m + tc], r_d = a.pin[PDIAM * m + tc];
    const double r_dpw = *a.dpw_in;
    if (t < PCOUNT) srec[t] = r_rec;
    if (staged && t < P.np) {
        scon[t] = r_x; scon[NP_LDS + t] = r_y; scon[2 * NP_LDS + t] = r_d;
    }
    if (t == 0) sdpw = r_dpw;
    __syncthreads();
    double fcx = 0.0, fcy = 0.0, dl = 0.0;
    int cnt = 0;
    if (P.np <= 64) {
        if (t < 64)
            cnt = contact_wave(P, scon, scon + NP_LDS, scon + 2 * NP_LDS, P.np, p, sdpw, fcx, fcy,
                               a.pairs + p * MAXNB, &dl);
    } else {
        cnt = staged
            ? contact_block(P, scon, scon + NP_LDS, scon + 2 * NP_LDS, P.np, p, sdpw, cl, fcx, fcy,
                            a.pairs + p * MAXNB, &dl)
            : contact_block(P, a.pin + PX * m, a.pin + PY * m, a.pin + PDIAM * m, P.np, p, sdpw,
                            cl, fcx, fcy, a.pairs + p * MAXNB, &dl);
    }
    __shared__ double snew[PCOUNT];
    if (t < 64) {
        const double val = update_wave(P, srec, srec[PFX], srec[PFY], srec[PTQ], __shfl(fcx, 0),
                                       __shfl(fcy, 0), a.kk_upd);
        if (t < PCOUNT) {
            a.pout[t * m + p] = val;
            snew[t] = val;
        }
        if (t == 0) {
            a.paircnt[p] = cnt;
            if (p == P.np - 1) *a.dpw_out = dl;
        }
    }
    __syncthreads();
    const Box nb = reach_box(P, snew[PX], snew[PY], snew[PDIAM], snew[PZETA]);
    if (t == 0) {
        put_box(a.pbox_out + 4 * p, nb);
        if (a.spf_out) make_spf(snew, a.spf_out + p * SPF);
        bool bad = false;
        for (int f = 0; f < PCOUNT; ++f) bad = bad || !__builtin_isfinite(snew[f]);
        if (bad) set_err(a, LBM_EF_NAN);
    }
    if (P.np > NP_LOOP)
        register_tiles(P, nb, p, a.list_gen + 1u, a.tl_hdr_out, a.tl_ids_out, a.err);
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

// Sums of n partials of x and of y (y may be null) by one block:
// wrap-around integer addition, so any order gives the same totals.  Each
// thread issues eight loads of each array before it adds any: the partials
// come from the L2, and a loop that waited for each load in turn made this
// one block the last to finish a recording launch over a large grid
// (16,401 partials: 64 round trips per thread, about 13 us).
__device__ __forceinline__ void block_sum2_u64(const unsigned long long *x,
                                               const unsigned long long *y, long long n,
                                               unsigned long long &sx, unsigned long long &sy) {
    __shared__ unsigned long long red[2][NT];
    constexpr int U = 8;
    unsigned long long ax = 0, ay = 0;
    long long t = threadIdx.x;
    for (; t + (long long)(U - 1) * NT < n; t += (long long)U * NT) {
        unsigned long long vx[U], vy[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            vx[u] = x[t + (long long)u * NT];
            vy[u] = y ? y[t + (long long)u * NT] : 0ull;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ax += vx[u];
            ay += vy[u];
        }
    }
    for (; t < n; t += NT) {
        ax += x[t];
        if (y) ay += y[t];
    }
    red[0][threadIdx.x] = ax;
    red[1][threadIdx.x] = ay;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    sx = red[0][0];
    sy = red[1][0];
    __syncthreads();
}

// Reducer block: the previous launch's monitor partials into its slots (the
// first slot of each; the others stay zero).
__device__ __forceinline__ void monitor_reduce(const unsigned long long *part, long long ntiles,
                                               unsigned long long *mass_slot,
                                               unsigned long long *yield_slot) {
    if (!part) return;
    unsigned long long m, y;
    block_sum2_u64(part, yield_slot ? part + ntiles : nullptr, ntiles, m, y);
    if (threadIdx.x == 0) {
        if (mass_slot) *mass_slot = m;
        if (yield_slot) *yield_slot = y;
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

// Stochastic-rounding words of node n: one 32-bit hash of the node and the
// launch, spread over the nine populations by distinct odd 24-bit multipliers
// (24 x 24-bit products issue at full rate).
__device__ __forceinline__ uint32_t node_hash(unsigned n, unsigned sr) {
    return mix32(n ^ mix32(sr));
}
// The same from the header's mix32(sr).
__device__ __forceinline__ uint32_t node_hash_h(unsigned n, const StepHead &h) {
    return mix32(n ^ h.srh);
}

// Store of the nine post-collision values of node n in storage units.
template <class St>
__device__ __forceinline__ void store_dev(const DevP &P, void *dst, unsigned n, const float H[9],
                                          uint32_t nh) {
    const auto drs = make_rsrc(dst, P.buf_bytes);
#pragma unroll
    for (int k = 0; k < 9; ++k)
        St::bst(drs, H[k], n * (unsigned)sizeof(typename St::T),
                (unsigned)(k * P.plane * (long long)sizeof(typename St::T)), sr_word(nh, k));
}

// The same from the header's byte sizes (a tile's stores).
template <class St>
__device__ __forceinline__ void store_dev_h(const StepHead &h, void *dst, unsigned n,
                                            const float H[9], uint32_t nh) {
    const auto drs = make_rsrc(dst, h.buf_bytes);
#pragma unroll
    for (int k = 0; k < 9; ++k)
        St::bst(drs, H[k], n * (unsigned)sizeof(typename St::T), (unsigned)k * h.plane_bytes,
                sr_word(nh, k));
}

// Prologue of a fused run: collide the FP64 state of step k (F, RHO, corrected
// UX/UY, G, OMEGA_BN) into f*_{k+1} in the storage format.
template <class St, bool DEV>
__device__ __forceinline__ void prologue_tile(const StepArgs &a, long long tile) {
    const DevP &P = a.P;
    long long n = tile * NT + threadIdx.x;
    int yielded = 0;
    if (n < P.nn) {
        int j = row_of(P, (unsigned)n), i = (int)(n - (long long)j * P.W);
        double r = a.crho[n], ux = a.cux[n], uy = a.cuy[n];
        double gx = a.cg[n], gy = a.cg[P.plane + n];
        int b = (i == 0 || j == 0 || i == P.nx || j == P.ny) ? bnd_index(P, i, j) : -1;
        if constexpr (DEV) {
            const float S = St::S;
            float H[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                double fk = a.cf[k * P.plane + n];
                H[k] = a.hx_in ? a.hx_in[k * P.plane + n]
                               : (float)((fk - (double)wdev(k) * P.rho0) * (double)S);
            }
            const DevMom mo = dev_moments(H);
            float om = (float)(P.rheology != 0 ? a.cobn[n] : P.omega), om2, str[3];
            const float rhof = (float)r;
            const float rinvS = frcp(rhof) * St::invS;
            if (a.aux || a.want_yield)
                collide_dev<true, true>(P, H, mo, S, rhof, rinvS, (float)ux, (float)uy, (float)gx,
                                        (float)gy, om, om2, str, yielded);
            else
                collide_dev<true, false>(P, H, mo, S, rhof, rinvS, (float)ux, (float)uy,
                                         (float)gx, (float)gy, om, om2, str, yielded);
            store_dev<St>(P, a.dst, n, H, node_hash(n, a.sr));
            if (P.rheology != 0)
                __builtin_amdgcn_raw_buffer_store_b32(
                    __builtin_bit_cast(unsigned, om2),
                    make_rsrc(a.rate_out, (unsigned)(P.plane * 4)), (int)(n * 4u), 0,
                    ST_THROUGH);
            if (a.aux) {
                a.xstress[n] = str[0]; a.xstress[P.plane + n] = str[1];
                a.xstress[2 * P.plane + n] = str[2];
                if (P.rheology != 0) a.ximage[n] = yielded ? 0 : 5;
            }
            if (P.boundary == 2) {
                if (b < 0) eq_put(a, P, i, j, (float)(r - P.rho0), (float)ux, (float)uy);
            } else if (b >= 0) {
                ((float *)a.erho_dst)[b] = (float)r;
            }
        } else {
            typename St::T *dst = (typename St::T *)a.dst;
            double f[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) f[k] = a.cf[k * P.plane + n];
            CollideOut o;
            double om = P.rheology != 0 ? a.cobn[n] : P.omega;
            collide64_fused(P, f, r, ux, uy, gx, gy, om, o);
#pragma unroll
            for (int k = 0; k < 9; ++k) St::st(dst + k * P.plane + n, f[k], 0u);
            if (P.rheology != 0) ((double *)a.rate_out)[n] = o.omega_bn;
            yielded = o.yielded;
            if (a.aux) {
                a.xstress[n] = o.stress[0]; a.xstress[P.plane + n] = o.stress[1];
                a.xstress[2 * P.plane + n] = o.stress[2];
                if (P.rheology != 0) a.ximage[n] = o.yielded ? 0 : 5;
            }
            if (P.boundary == 2) {
                if (b < 0) eq_put(a, P, i, j, r, ux, uy);
            } else if (b >= 0) {
                ((double *)a.erho_dst)[b] = r;
            }
        }
    }
    if (a.want_yield) tile_monitors<DEV>(a, tile, 0.0, yielded, false, true);
}

template <class St, bool DEV>
__global__ __launch_bounds__(NT) void k_prologue(StepArgs args) {
    const StepArgs &a = launch_args<StepArgs>();
    (void)args;
    if ((int)blockIdx.x < a.n_pblocks) {
        // the first helper slot of each particle advances it
        if ((int)blockIdx.x % a.kpart == 0) particle_advance<DEV>(a, (int)blockIdx.x / a.kpart);
        return;
    }
    if ((int)blockIdx.x < a.n_pblocks + a.nred) return;  // nothing before the prologue
    prologue_tile<St, DEV>(a, xcd_rank((int)blockIdx.x, (int)(a.n_pblocks + a.nred), (int)a.ntiles));
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

// ---------------------------------------------------------------- tiles
// The nodes of the calling thread's wavefront in tile t: its first node, its
// last node (clamped to the grid's last node) and their rows.  Uniform.  A
// wavefront wholly past the last node gets a span whose first row lies
// below its last (it meets no box).
__device__ __forceinline__ TileSpan wave_span(const StepHead &h, int tile) {
    TileSpan s;
    s.n0 = (unsigned)tile * NT +
           (unsigned)__builtin_amdgcn_readfirstlane((int)(threadIdx.x & ~63u));
    s.n1 = min(s.n0 + 63u, h.nn - 1u);
    s.jl = row_of(h, s.n0);
    s.jh = row_of(h, s.n1);
    return s;
}

// Reach boxes of state k for the per-wavefront candidate test (1 ..
// NP_LOOP particles): lane p holds particle p's box (i0, i1, j0, j1); lanes
// past the last particle read zeros (the buffer's bound).  Issued before the
// population loads, so waiting for it does not wait for them.  (Written by
// the previous launch.)
struct BoxLane {
    int i0, i1, j0, j1;
};
__device__ __forceinline__ BoxLane load_box_lane(const StepArgs &a, const StepHead &h) {
    BoxLane b{1, 0, 1, 0};
    if (h.np > 0 && h.np <= NP_LOOP) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(
            make_rsrc(a.pbox_in, (unsigned)h.np * 16u), (int)((threadIdx.x & 63u) * 16u), 0, 0);
        b = BoxLane{(int)v[0], (int)v[1], (int)v[2], (int)v[3]};
    }
    return b;
}

// Which particles can reach the calling wavefront's nodes in this launch?
// Up to NP_LOOP particles: a bit per particle whose reach box of state k
// meets the nodes, either per wavefront with one lane per particle (LANES:
// vector work and one compare mask, boxes from load_box_lane) or per tile
// with scalar loads and scalar arithmetic; beyond NP_LOOP: the
// wavefront's candidate list of this launch (header: generation << 32 |
// count).  A
// candidate whose box misses a node contributes nothing there, so the
// wavefront's set and its tile's give the same forcing.  (These buffers
// were written by the previous launch, not this one.)
template <bool LANES>
__device__ __forceinline__ int tile_candidates(const StepArgs &a, const StepHead &h, int t,
                                               const TileSpan &sp, const BoxLane &b,
                                               unsigned &mask, const int *&ids) {
    mask = 0u;
    ids = nullptr;
    const int np = h.np;
    if (np <= 0) return 0;
    if (np <= NP_LOOP) {
        if (LANES) {
            const bool hit = span_meets_box_lane((int)h.W, sp, b.i0, b.i1, b.j0, b.j1,
                                                 (int)(threadIdx.x & 63u) < np);
            mask = (unsigned)__builtin_amdgcn_ballot_w64(hit);
        } else {
            TileSpan ts;
            ts.n0 = (unsigned)t * NT;
            ts.n1 = min(ts.n0 + (unsigned)(NT - 1), h.nn - 1u);
            ts.jl = row_of(h, ts.n0);
            ts.jh = row_of(h, ts.n1);
            unsigned m = 0u;
            for (int p = 0; p < np; ++p) {
                const int *w = a.pbox_in + 4 * p;
                if (tile_meets_box((int)h.W, ts, cload(w), cload(w + 1), cload(w + 2),
                                   cload(w + 3)))
                    m |= 1u << p;
            }
            mask = m;
        }
        return __builtin_popcount(mask);
    }
    const unsigned u = sp.n0 / (unsigned)LIST_UNIT;  // the wavefront's list unit
    const unsigned long long hdr = cload(a.tl_hdr_in + u);
    if ((unsigned)(hdr >> 32) != a.list_gen) return 0;
    ids = a.tl_ids_in + (size_t)u * MAXC;
    return min((int)(hdr & 0xffffffffu), MAXC);
}

// A value every lane holds alike, as a wavefront-uniform one.
__device__ __forceinline__ int uniform_i32(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Particle inputs of a tile's SPM pass for one candidate particle p: its
// reach box of state k and, reduced formats, its FP32 profile record, or,
// FP64, the state fields the profile and the rigid-body velocity need.
// Read with scalar loads (p is uniform; these buffers are not written during
// the launch).
template <bool DEV> struct Cand;
template <> struct Cand<true> {
    int p;
    Box b;
    Prof s;
};
template <> struct Cand<false> {
    int p;
    Box b;
    double xc, yc, vx, vy, w, dm, zt;
};
template <bool DEV>
__device__ __forceinline__ Cand<DEV> load_cand(const StepArgs &a, int p) {
    Cand<DEV> c;
    c.p = p;
    const int *w = a.pbox_in + 4 * p;
    c.b.i0 = cload(w); c.b.i1 = cload(w + 1); c.b.j0 = cload(w + 2); c.b.j1 = cload(w + 3);
    if constexpr (DEV) {
        const float *q = a.spf_in + p * SPF;
        c.s.x = cload(q + SX); c.s.xf = cload(q + SXF);
        c.s.y = cload(q + SY); c.s.yf = cload(q + SYF);
        c.s.vx = cload(q + SVX); c.s.vy = cload(q + SVY);
        c.s.w = cload(q + SW); c.s.rp = cload(q + SRP);
        c.s.hz = cload(q + SHZ); c.s.pz = cload(q + SPZ);
    } else {
        const int m = a.P.maxp;
        c.xc = cload(a.pin + PX * m + p); c.yc = cload(a.pin + PY * m + p);
        c.vx = cload(a.pin + PVX * m + p); c.vy = cload(a.pin + PVY * m + p);
        c.w = cload(a.pin + PW * m + p);
        c.dm = cload(a.pin + PDIAM * m + p); c.zt = cload(a.pin + PZETA * m + p);
    }
    return c;
}

// SPM forcing G at a tile's node (F6:485-495): of the candidate particles
// whose phi at (i, j) is positive, the one with the highest index sets
// G = phi (u_p - u), u_p the rigid-body velocity at the node, from state k.
// u: uncorrected velocity of the node (F6:333-340); valid: the thread holds
// a node.
template <bool DEV, class Real>
__device__ __forceinline__ void tile_forcing(const StepArgs &a, int cnt, unsigned mask,
                                             const int *ids, int i, int j, bool valid, Real u,
                                             Real v, Real &gx, Real &gy) {
    gx = (Real)0;
    gy = (Real)0;
    int best = -1;
    unsigned rest = mask;
    for (int c = 0; c < cnt; ++c) {
        const int p = ids ? cload(ids + c) : __builtin_ctz(rest);
        rest &= rest - 1u;
        const Cand<DEV> cd = load_cand<DEV>(a, p);
        if constexpr (DEV) {
            float dx, dy;
            const float ph = valid ? phi_prof(cd.s, i, j, dx, dy) : 0.0f;
            if (ph > 0.0f && p > best) {
                gx = ph * ((cd.s.vx - cd.s.w * dy) - u);
                gy = ph * ((cd.s.vy + cd.s.w * dx) - v);
                best = p;
            }
        } else {
            const double ph = valid ? phi_at<true>(cd.xc, cd.yc, cd.dm, cd.zt, i, j) : 0.0;
            if (ph > 0.0 && p > best) {
                gx = ph * ((cd.vx - cd.w * ((double)j - cd.yc)) - u);
                gy = ph * ((cd.vy + cd.w * ((double)i - cd.xc)) - v);
                best = p;
            }
        }
    }
}

// Left and right edge nodes of a PULL_SIDES wavefront (rows 1 .. ny-1, so
// no corner), after the interior pull of every lane: with Zou-He edges the
// rule replaces exactly the populations whose sources lie outside and reads
// only pulled ones, so it applies to the loaded values as they are; with
// bounce-back the replaced populations (left: 1, 5, 8; right: 3, 6, 7) are
// the node's own opposite ones (3, 7, 6; 1, 8, 5) less the wall terms of
// bounce_back_node.  The same values as load_fk_lane_dev / load_fk_lane64;
// only the boundary lanes execute the loads.
template <class St, int BND, class V>
__device__ __forceinline__ bool side_fetch(const DevP &P, const StepHead &h, unsigned n, int i,
                                           V v[9]) {
    const bool l = i == 0, r = i == P.nx;
    if (!(l || r)) return false;
    if ((BND >= 0 ? BND : P.boundary) == 1) return true;
    // three loads with the plane chosen per lane, then selects at fixed
    // indices (a store per branch at different indices would put v in
    // scratch memory)
    const auto rs = make_rsrc(h.src, h.buf_bytes);
    const unsigned vo = n * (unsigned)sizeof(typename St::T), pb = h.plane_bytes;
    const V x = St::bld(rs, vo + (l ? 3u : 1u) * pb, 0u);  // left: f*_3 -> 1; right: f*_1 -> 3
    const V y = St::bld(rs, vo + (l ? 7u : 8u) * pb, 0u);  // left: f*_7 -> 5; right: f*_8 -> 6
    const V z = St::bld(rs, vo + (l ? 6u : 5u) * pb, 0u);  // left: f*_6 -> 8; right: f*_5 -> 7
    v[1] = l ? x : v[1];
    v[3] = r ? x : v[3];
    v[5] = l ? y : v[5];
    v[6] = r ? y : v[6];
    v[8] = l ? z : v[8];
    v[7] = r ? z : v[7];
    return true;
}
template <class St, int BND>
__device__ __forceinline__ void side_edges_dev(const DevP &P, const StepHead &h,
                                               const float *erho, unsigned n, int i, int j,
                                               float H[9]) {
    if (!side_fetch<St, BND>(P, h, n, i, H)) return;
    const float S = St::S;
    if ((BND >= 0 ? BND : P.boundary) == 1) {
        zou_he_edge<float>(P, i, j, S * P.frho0, H);
        return;
    }
    const float rw = erho[bnd_index(P, i, j)];
    const float a = (float)INV6 * S, b = (float)(4.0 * INV6) * S;
    bounce_back_node<float>(P, i, j, a * rw, b * rw, H);
}
template <int BND>
__device__ __forceinline__ void side_edges64(const DevP &P, const StepHead &h,
                                             const double *erho, unsigned n, int i, int j,
                                             double v[9]) {
    if (!side_fetch<StF64, BND>(P, h, n, i, v)) return;
    if ((BND >= 0 ? BND : P.boundary) == 1) {
        zou_he_edge<double>(P, i, j, 0.0, v);
        return;
    }
    const double clb2 = 1.0;
    const double rw = erho[bnd_index(P, i, j)];
    bounce_back_node<double>(P, i, j, INV6 * clb2 * rw, 4.0 * INV6 * clb2 * rw, v);
}

// How a wavefront pulls its populations (uniform over the wavefront):
// PULL_INTERIOR: every node is an interior node away from the edge rules
// (for equilibrium edges, also away from the nodes that feed the edge
// arrays); one vector offset per node.  PULL_SIDES: every row the wavefront
// touches lies strictly inside the domain, so its only boundary nodes are on
// the left and right edges (bounce-back or Zou-He edges): the interior pull
// for every lane, then those lanes alone replace the populations whose
// sources lie outside (side_edges).  PULL_GENERAL: anything else (the
// bottom and top rows, equilibrium edges, the grid's last wavefront): the
// per-lane pull of load_fk_lane_dev.  jw: the wavefront's first row.
enum { PULL_INTERIOR = 0, PULL_SIDES = 1, PULL_GENERAL = 2 };
template <int BND>
__device__ __forceinline__ int wave_class(const StepHead &h, const TileSpan &sp, int &jw) {
    jw = 0;
    if (sp.n0 + 63u >= h.nn) return PULL_GENERAL;
    const int e = BND == 2 ? 2 : (BND >= 0 ? 1 : h.margin);
    const int j0 = sp.jl, j1 = sp.jh;
    const int i0 = (int)(sp.n0 - (unsigned)j0 * h.W);
    const int i1 = (int)(sp.n1 - (unsigned)j1 * h.W);
    jw = j0;
    if (j0 == j1 && j0 >= e && j0 <= h.ny - e && i0 >= e && i1 <= h.nx - e) return PULL_INTERIOR;
    if (e == 1 && j0 >= 1 && j1 <= h.ny - 1) return PULL_SIDES;
    return PULL_GENERAL;
}

// The nine pulled populations of an interior node from the header's
// offsets: one vector offset for the node, a uniform offset per population.
template <class St, class V>
__device__ __forceinline__ void load_interior_h(const StepHead &h, unsigned n, V v[9]) {
    const auto rs = make_rsrc(h.src, h.buf_bytes);
    const unsigned vo = n * (unsigned)sizeof(typename St::T);
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = St::bld(rs, vo, h.pull[k]);
}
// The same as the stored bits (reduced formats), converted later with
// St::cvt: the loads are then not waited for until the first use of a value.
template <class St>
__device__ __forceinline__ void load_interior_raw(const StepHead &h, unsigned n,
                                                  typename St::R v[9]) {
    const auto rs = make_rsrc(h.src, h.buf_bytes);
    const unsigned vo = n * (unsigned)sizeof(typename St::T);
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = St::braw(rs, vo, h.pull[k]);
}

// Density and uncorrected velocity of a node from its moments in storage
// units (reduced formats); the tiles and the particle helper blocks form
// them alike.
__device__ __forceinline__ void dev_velocity(const DevP &P, const DevMom &mo, float invS,
                                             float &rho, float &rinvS, float &ux, float &uy) {
    rho = P.frho0 + mo.rd * invS;
    rinvS = frcp(rho) * invS;
    ux = mo.jx * rinvS;
    uy = mo.jy * rinvS;
}

// FINAL launch, node n: STRESS and the yield-flag DOMAIN_IMAGE of
// iteration k from the copies its collision wrote (by iteration parity), or,
// Newtonian, DOMAIN_IMAGE from the particles (F6:438, F6:460-467).
__device__ __forceinline__ void final_aux(const StepArgs &a, unsigned n, int i, int j) {
    const DevP &P = a.P;
#pragma unroll
    for (int c = 0; c < 3; ++c) a.cstress[c * P.plane + n] = a.astress_in[c * P.plane + n];
    a.cimage[n] = P.rheology != 0 ? a.aimage_in[n] : image_node(P, a.pin, i, j);
}

// Outputs of a tile's node, reduced-precision storage: (FINAL = false) the
// collision of step k+1 stored as f*_{k+1}, the rate, edge arrays and
// auxiliary fields, or (FINAL = true) the FP64 state of step k; then the
// monitor partials.  FORCED: G may be non-zero.
template <class St, bool FINAL, int BND, bool FORCED, class CP>
__device__ __forceinline__ void tile_dev_out(const StepArgs &a, const StepHead &h,
                                             const CP &cp, unsigned out, void *dst,
                                             void *rate_out, int tile,
                                             unsigned n, bool valid, bool edges, int i, int j,
                                             float (&H)[9],
                                             const DevMom &mo, float rho, float rinvS, float ucx,
                                             float ucy, float gx, float gy, float om) {
    const DevP &P = a.P;
    const float S = St::S, invS = St::invS;
    int yielded = 0;
    if (valid) {
        if (FINAL) {
            const typename St::T *s0 = (const typename St::T *)a.src;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                double w0 = (double)wdev(k) * P.rho0;
                a.cf[k * P.plane + n] = w0 + (double)H[k] * (double)invS;
                a.cfp[k * P.plane + n] = w0 + (double)St::ld(s0 + k * P.plane + n) * (double)invS;
                if (a.hx) a.hx[k * P.plane + n] = H[k];
            }
            a.crho[n] = (double)P.rho0 + (double)mo.rd * (double)invS;
            a.cux[n] = ucx; a.cuy[n] = ucy;
            a.cg[n] = gx; a.cg[P.plane + n] = gy;
            if (P.rheology != 0) a.cobn[n] = ((const float *)a.rate_in)[n];
            final_aux(a, n, i, j);
        } else {
            float om2, str[3];
            if (out & 1u)
                collide_dev<FORCED, true>(cp, H, mo, S, rho, rinvS, ucx, ucy, gx, gy, om, om2, str,
                                          yielded);
            else
                collide_dev<FORCED, false>(cp, H, mo, S, rho, rinvS, ucx, ucy, gx, gy, om, om2,
                                           str, yielded);
            store_dev_h<St>(h, dst, n, H, node_hash_h(n, h));
            if (cp.rheology != 0)
                __builtin_amdgcn_raw_buffer_store_b32(
                    __builtin_bit_cast(unsigned, om2), make_rsrc(rate_out, h.rate_bytes),
                    (int)(n * 4u), 0, ST_THROUGH);
            if (out & 2u) {
                a.xstress[n] = str[0]; a.xstress[P.plane + n] = str[1];
                a.xstress[2 * P.plane + n] = str[2];
                if (P.rheology != 0) a.ximage[n] = yielded ? 0 : 5;
            }
            if (edges) {
                if ((BND >= 0 ? BND : P.boundary) == 2) {
                    if (!(i == 0 || j == 0 || i == P.nx || j == P.ny))
                        eq_put(a, P, i, j, mo.rd * invS, ucx, ucy);
                } else if (i == 0 || j == 0 || i == P.nx || j == P.ny) {
                    ((float *)a.erho_dst)[bnd_index(P, i, j)] = rho;
                }
            }
        }
    }
    if (a.slots_mass || (!FINAL && a.want_yield)) {
        const double mass = valid ? (double)P.rho0 + (double)mo.rd * (double)invS : 0.0;
        tile_monitors<true>(a, tile, mass, yielded, a.slots_mass != nullptr,
                            !FINAL && a.want_yield);
    }
}

// One fused time step k over a tile, reduced-precision storage (FP32
// collide): pull-stream + boundary rules (f_k, by the wavefront's pull
// class), density and velocity, SPM forcing G_k from the particles that can
// reach the wavefront, velocity correction, then the outputs
// (tile_dev_out).  The node rate (and, SINGLE, the particles' reach boxes)
// are requested first; the candidate test runs while the population loads
// are in flight.  SINGLE: a single-simulation launch (the batch kernel, which
// keeps more of each simulation's arguments in scalar registers, tests the
// candidates per tile with scalar work and reads the collision parameters
// from DevP).
template <class St, bool FINAL, int BND, bool SINGLE>
__device__ __forceinline__ void tile_dev(const StepArgs &a, const StepHead &h, int tile) {
    const DevP &P = a.P;
    const unsigned n = (unsigned)tile * NT + threadIdx.x;
    const bool valid = n < h.nn;
    // lanes past the last node read the zero padding of the rate plane
    const float om = (!FINAL && h.rheology != 0)
                         ? __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                               make_rsrc(h.rate_in, h.rate_bytes), (int)(n * 4u), 0, 0))
                         : P.fomega;
    const TileSpan sp = wave_span(h, tile);
    const BoxLane bx = SINGLE ? load_box_lane(a, h) : BoxLane{1, 0, 1, 0};
    CollideHead ch;
    if constexpr (SINGLE) ch = load_collide(a);
    int jw;
    const int kind = wave_class<BND>(h, sp, jw);
    int i, j;
    float H[9];
    unsigned mask;
    const int *ids;
    int cnt = 0;
    // SINGLE: each branch tests the candidates after issuing its own loads,
    // so the test waits for the reach boxes alone; the batch looks them up
    // (scalar) after the branches, while the loads are in flight
    if (kind == PULL_INTERIOR) {
        // the stored bits, converted after the candidate test
        typename St::R R[9];
        j = jw;
        i = (int)(n - (unsigned)jw * h.W);
        load_interior_raw<St>(h, n, R);
        cnt = SINGLE ? tile_candidates<true>(a, h, tile, sp, bx, mask, ids)
                     : tile_candidates<false>(a, h, tile, sp, bx, mask, ids);
#pragma unroll
        for (int k = 0; k < 9; ++k) H[k] = St::cvt(R[k]);
    } else if (kind == PULL_SIDES) {
        j = row_of(h, n);
        i = (int)(n - (unsigned)j * h.W);
        load_interior_h<St>(h, n, H);
        if constexpr (SINGLE) cnt = tile_candidates<true>(a, h, tile, sp, bx, mask, ids);
        side_edges_dev<St, BND>(P, h, (const float *)a.erho_src, n, i, j, H);
    } else {
        const unsigned nc = valid ? n : h.nn - 1u;
        j = row_of(h, nc);
        i = (int)(nc - (unsigned)j * h.W);
        load_fk_lane_dev<St, BND>(P, h.src, (const float *)a.erho_src, nc, i, j, H);
        if constexpr (SINGLE) cnt = tile_candidates<true>(a, h, tile, sp, bx, mask, ids);
    }
    if constexpr (!SINGLE) {
        if (kind != PULL_INTERIOR) cnt = tile_candidates<false>(a, h, tile, sp, bx, mask, ids);
    }
    const bool edges = kind != PULL_INTERIOR;
    const DevMom mo = dev_moments(H);
    float rho, rinvS, ux, uy;
    dev_velocity(P, mo, St::invS, rho, rinvS, ux, uy);
    if (valid && !__builtin_isfinite(rho)) set_err(a, LBM_EF_NAN);
    if constexpr (SINGLE) {
        if (cnt > 0) {
            float gx, gy;
            tile_forcing<true, float>(a, cnt, mask, ids, i, j, valid, ux, uy, gx, gy);
            tile_dev_out<St, FINAL, BND, true>(a, h, ch, ch.out, ch.dst, ch.rate_out, tile, n,
                                               valid, edges, i, j, H, mo, rho, rinvS,
                                               ux + 0.5f * gx, uy + 0.5f * gy, gx, gy, om);
        } else {
            tile_dev_out<St, FINAL, BND, false>(a, h, ch, ch.out, ch.dst, ch.rate_out, tile, n,
                                                valid, edges, i, j, H, mo, rho, rinvS, ux, uy,
                                                0.0f, 0.0f, om);
        }
    } else {
        if (cnt > 0) {
            float gx, gy;
            tile_forcing<true, float>(a, cnt, mask, ids, i, j, valid, ux, uy, gx, gy);
            tile_dev_out<St, FINAL, BND, true>(a, h, P, a.ch.out, a.dst, a.rate_out, tile, n,
                                               valid, edges, i, j, H, mo, rho, rinvS,
                                               ux + 0.5f * gx, uy + 0.5f * gy, gx, gy, om);
        } else {
            tile_dev_out<St, FINAL, BND, false>(a, h, P, a.ch.out, a.dst, a.rate_out, tile, n,
                                                valid, edges, i, j, H, mo, rho, rinvS, ux, uy,
                                                0.0f, 0.0f, om);
        }
    }
}

// The same step with FP64 storage, in the reference's operation order.
template <class St, bool FINAL, int BND>
__device__ __forceinline__ void tile_f64(const StepArgs &a, const StepHead &h, int tile) {
    const DevP &P = a.P;
    const unsigned n = (unsigned)tile * NT + threadIdx.x;
    const bool valid = n < h.nn;
    const typename St::T *src = (const typename St::T *)h.src;
    // lanes past the last node read the zero padding of the rate plane
    const double om = (!FINAL && h.rheology != 0)
                          ? StF64::bld(make_rsrc(h.rate_in, h.rate_bytes), n * 8u, 0u)
                          : P.omega;
    const TileSpan sp = wave_span(h, tile);
    int jw;
    const int kind = wave_class<BND>(h, sp, jw);
    int i, j;
    double f[9], rho, ux, uy;
    if (kind == PULL_INTERIOR) {
        j = jw;
        i = (int)(n - (unsigned)jw * h.W);
        load_interior_h<StF64>(h, n, f);
    } else if (kind == PULL_SIDES) {
        j = row_of(h, n);
        i = (int)(n - (unsigned)j * h.W);
        load_interior_h<StF64>(h, n, f);
        side_edges64<BND>(P, h, (const double *)a.erho_src, n, i, j, f);
    } else {
        const unsigned nc = valid ? n : h.nn - 1u;
        j = row_of(h, nc);
        i = (int)(nc - (unsigned)j * h.W);
        load_fk_lane64<BND>(P, h.src, (const double *)a.erho_src, nc, i, j, f);
    }
    const bool edges = kind != PULL_INTERIOR;
    unsigned mask;
    const int *ids;
    const int cnt = tile_candidates<false>(a, h, tile, sp, BoxLane{1, 0, 1, 0}, mask, ids);
    macros64(f, rho, ux, uy);
    if (valid && !__builtin_isfinite(rho)) set_err(a, LBM_EF_NAN);
    double gx = 0.0, gy = 0.0;
    if (cnt > 0) tile_forcing<false, double>(a, cnt, mask, ids, i, j, valid, ux, uy, gx, gy);
    const double ucx = ux + 0.5 * 1.0 * (gx + 0.0), ucy = uy + 0.5 * 1.0 * (gy + 0.0);
    int yielded = 0;
    if (valid) {
        if (FINAL) {
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                a.cf[k * P.plane + n] = f[k];
                a.cfp[k * P.plane + n] = St::ld(src + k * P.plane + n);
            }
            a.crho[n] = rho; a.cux[n] = ucx; a.cuy[n] = ucy;
            a.cg[n] = gx; a.cg[P.plane + n] = gy;
            if (P.rheology != 0) a.cobn[n] = ((const double *)a.rate_in)[n];
            final_aux(a, n, i, j);
        } else {
            CollideOut o;
            collide64_fused(P, f, rho, ucx, ucy, gx, gy, om, o);
            const auto drs = make_rsrc(a.dst, P.buf_bytes);
#pragma unroll
            for (int k = 0; k < 9; ++k)
                StF64::bst(drs, f[k], n * 8u, (unsigned)(k * P.plane * 8LL));
            if (P.rheology != 0)
                StF64::bst(make_rsrc(a.rate_out, (unsigned)(P.plane * 8)), o.omega_bn, n * 8u, 0u);
            yielded = o.yielded;
            if (a.aux) {
                a.xstress[n] = o.stress[0]; a.xstress[P.plane + n] = o.stress[1];
                a.xstress[2 * P.plane + n] = o.stress[2];
                if (P.rheology != 0) a.ximage[n] = o.yielded ? 0 : 5;
            }
            if (edges) {
                const int b =
                    (i == 0 || j == 0 || i == P.nx || j == P.ny) ? bnd_index(P, i, j) : -1;
                if ((BND >= 0 ? BND : P.boundary) == 2) {
                    if (b < 0) eq_put(a, P, i, j, rho, ucx, ucy);
                } else if (b >= 0) {
                    ((double *)a.erho_dst)[b] = rho;
                }
            }
        }
    }
    if (a.slots_mass || (!FINAL && a.want_yield))
        tile_monitors<false>(a, tile, valid ? rho : 0.0, yielded, a.slots_mass != nullptr,
                             !FINAL && a.want_yield);
}

template <class St, bool DEV, bool FINAL, int BND, bool SINGLE = true>
__device__ __forceinline__ void step_tile(const StepArgs &a, const StepHead &h, int tile) {
    if constexpr (DEV) tile_dev<St, FINAL, BND, SINGLE>(a, h, tile);
    else tile_f64<St, FINAL, BND>(a, h, tile);
}

// ---------------------------------------------------------------- particles
// Chunk c of particle p's reach box (box nodes c NT .. c NT + NT - 1 in
// row-major order, one per thread): the chunk's share of the hydrodynamic
// force and torque on p (ComputeForce_HYDRO, F6:542-556).  At each node with
// phi > 0, f_k is pulled from f*_k with the boundary rules and the node's
// density and uncorrected velocity are formed exactly as the tile that owns
// the node forms them; the node adds rho phi (u - u_p) and its moment arm
// cross that.  All NT threads; the block total is valid in thread 0 (a fixed
// DPP tree per wavefront, then the four wavefronts in order).
template <class St, bool DEV, int BND>
__device__ __forceinline__ void chunk_force(const StepArgs &a, int p, int c, double F[3]) {
    static_assert(NT == 256, "four wavefronts per block");
    const DevP &P = a.P;
    const Cand<DEV> cd = load_cand<DEV>(a, p);
    const Box &b = cd.b;
    const int bw = b.i1 - b.i0 + 1, bh = b.j1 - b.j0 + 1;
    const int q = c * NT + (int)threadIdx.x;
    const bool valid = bw > 0 && bh > 0 && q < bw * bh;
    if (c == 0 && threadIdx.x == 0 && bw > 0 && bh > 0 && bw * bh > a.kpart * NT)
        set_err(a, LBM_EF_SLOT);
    const int i = valid ? b.i0 + q % bw : 0, j = valid ? b.j0 + q / bw : 0;
    const long long n = (long long)j * P.W + i;
    const typename St::T *src = (const typename St::T *)a.src;
    double acc[3] = {0.0, 0.0, 0.0};
    if constexpr (DEV) {
        float dx, dy;
        const float ph = valid ? phi_prof(cd.s, i, j, dx, dy) : 0.0f;
        if (ph > 0.0f) {
            float H[9];
            load_fk_lane_dev<St, BND>(P, a.src, (const float *)a.erho_src, (unsigned)n, i, j, H);
            const DevMom mo = dev_moments(H);
            float rho, rinvS, ux, uy;
            dev_velocity(P, mo, St::invS, rho, rinvS, ux, uy);
            const float uxp = cd.s.vx - cd.s.w * dy, uyp = cd.s.vy + cd.s.w * dx;
            const double rp = (double)rho * (double)ph;
            const double du = (double)(ux - uxp), dv = (double)(uy - uyp);
            acc[0] = rp * du;
            acc[1] = rp * dv;
            acc[2] = rp * ((double)dx * dv - (double)dy * du);
        }
    } else {
        const double ph = valid ? phi_at<true>(cd.xc, cd.yc, cd.dm, cd.zt, i, j) : 0.0;
        if (ph > 0.0) {
            double f[9], r, u, v;
            load_fk_lane64<BND>(P, a.src, (const double *)a.erho_src, (unsigned)n, i, j, f);
            macros64(f, r, u, v);
            const double uxp = cd.vx - cd.w * ((double)j - cd.yc);
            const double uyp = cd.vy + cd.w * ((double)i - cd.xc);
            acc[0] = r * ph * (u - uxp);
            acc[1] = r * ph * (v - uyp);
            acc[2] = r * ph * (((double)i - cd.xc) * (v - uyp) - ((double)j - cd.yc) * (u - uxp));
        }
    }
    __shared__ double red[NT / 64][3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double x = wave_sum_f64(acc[k]);
        if (lane == 63) red[wv][k] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) F[k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
}

// FORCE and TORQUE on particle p from the tagged chunk partials 0 .. K-1
// (K = a.kpart), by one wavefront: lane l takes chunks l, l + 64, ... in
// order, waiting (bounded) until their words carry this launch's tag; a
// fixed shuffle tree then adds the lanes.  The other chunks' blocks run at
// the same time as the caller (see particle_helper), so the wait is about
// as long as the slowest chunk's gather.  Valid in every lane; false when
// the wait ran out.
__device__ __forceinline__ bool sum_force_slots(const StepArgs &a, int p, double F[3]) {
    const int lane = threadIdx.x & 63;
    const unsigned long long *q = a.fslot + (size_t)p * a.kpart * SLOTW;
    double acc[3] = {0.0, 0.0, 0.0};
    bool ok = true;
    for (int r = lane; r < a.kpart; r += 64) {
        const unsigned long long *w = q + (size_t)r * SLOTW;
        unsigned long long v[SLOTW];
        int it = 0;
        for (;;) {
            bool ready = true;
#pragma unroll
            for (int k = 0; k < SLOTW; ++k) {
                v[k] = __hip_atomic_load(w + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ready = ready && (unsigned)(v[k] >> 32) == a.tag;
            }
            if (ready) break;
            if (++it >= TAG_SPIN) { ok = false; break; }
            __builtin_amdgcn_s_sleep(2);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
            acc[c] = acc[c] + __builtin_bit_cast(double, (v[2 * c] << 32) | (v[2 * c + 1] & 0xffffffffull));
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double x = acc[c];
#pragma unroll
        for (int sft = 32; sft > 0; sft >>= 1) x = x + __shfl_xor(x, sft, 64);
        F[c] = x;
    }
    return __all(ok);
}

// Helper block c of particle p in launch s >= 1 (all NT threads): the chunk
// partial of the force (chunk_force), stored as tagged words in slot (p, c).
// Chunk 0 then finishes the particle: FORCE_k from the K partials in chunk
// order (waiting, bounded, for the other chunks' words), the contact force
// on state k (F3:739-804), the history row of step k, and the Newton-Euler
// update to state k+1 (F6:299-313) with its reach box, profile record and,
// for many particles, tile registration; the final launch only writes
// FORCE_k into the record.  The wait is safe: the helper blocks are the
// lowest block indices of the launch, dispatched in order before any tile
// block, so the K - 1 blocks chunk 0 waits for have been given CU slots (a
// launch has far fewer helper blocks than the GPU has slots).  Waiting on
// the words themselves costs one memory round trip after the last chunk
// stores them; a done-counter whose last incrementer finishes would cost
// two (the counter's device-scope atomic, then the words).  The particle's
// record and the contact inputs are requested before the chunk's
// populations, so the finishing block has them at hand.
template <class St, bool DEV, bool FINAL, int BND>
__device__ __forceinline__ void particle_helper(const StepArgs &a, int p, int c) {
    const DevP &P = a.P;
    const int m = P.maxp;
    static_assert(NP_LDS <= NT, "one staged particle per thread");
    __builtin_amdgcn_s_setprio(3);  // the particle chain is the launch's critical path
    const int t = threadIdx.x;
    const int tr = min(t, PCOUNT - 1), tc = min(t, max(P.np - 1, 0));
    const double r_rec = a.pin[tr * m + p];
    double r_x = 0.0, r_y = 0.0, r_d = 0.0, r_dpw = 0.0;
    if (!FINAL) {
        r_x = a.pin[PX * m + tc]; r_y = a.pin[PY * m + tc]; r_d = a.pin[PDIAM * m + tc];
        r_dpw = *a.dpw_in;
    }
    double Fc[3];
    chunk_force<St, DEV, BND>(a, p, c, Fc);
    if (t == 0) {
        unsigned long long *o = a.fslot + ((size_t)p * a.kpart + c) * SLOTW;
#pragma unroll
        for (int k = 0; k < 3; ++k) put_tagged_f64(o + 2 * k, a.tag, Fc[k]);
    }
    if (c != 0) return;
    // ---- chunk 0: finish the particle
    __shared__ double srec[PCOUNT];
    __shared__ double scon[3 * NP_LDS];
    __shared__ double sdpw;
    __shared__ ContactLDS cl;
    __shared__ double snew[PCOUNT];
    const bool staged = P.np <= NP_LDS;
    if (t < PCOUNT) srec[t] = r_rec;
    if (!FINAL) {
        if (staged && t < P.np) {
            scon[t] = r_x; scon[NP_LDS + t] = r_y; scon[2 * NP_LDS + t] = r_d;
        }
        if (t == 0) sdpw = r_dpw;
    }
    __syncthreads();
    double fcx = 0.0, fcy = 0.0, dl = 0.0;
    int cnt = 0;
    if (!FINAL) {
        if (P.np <= 64) {
            if (t < 64)
                cnt = contact_wave(P, scon, scon + NP_LDS, scon + 2 * NP_LDS, P.np, p, sdpw, fcx,
                                   fcy, a.pairs + p * MAXNB, &dl);
        } else {
            cnt = staged
                ? contact_block(P, scon, scon + NP_LDS, scon + 2 * NP_LDS, P.np, p, sdpw, cl, fcx,
                                fcy, a.pairs + p * MAXNB, &dl)
                : contact_block(P, a.pin + PX * m, a.pin + PY * m, a.pin + PDIAM * m, P.np, p,
                                sdpw, cl, fcx, fcy, a.pairs + p * MAXNB, &dl);
        }
    }
    if (t < 64) {
        double F[3];
        if (!sum_force_slots(a, p, F) && t == 0) set_err(a, LBM_EF_WAIT);
        if (a.hist_part && t == 0) hist_row(a.hist_part + (size_t)p * HPC, srec, F);
        if (FINAL) {
            if (t == 0) {
                a.pout[PFX * m + p] = F[0]; a.pout[PFY * m + p] = F[1]; a.pout[PTQ * m + p] = F[2];
                if (!__builtin_isfinite(F[0] + F[1] + F[2])) set_err(a, LBM_EF_NAN);
            }
        } else {
            // the update by wavefront 0; the contact force is final in lane 0
            const double val = update_wave(P, srec, F[0], F[1], F[2], __shfl(fcx, 0),
                                           __shfl(fcy, 0), a.kk_upd);
            if (t < PCOUNT) {
                a.pout[t * m + p] = val;
                snew[t] = val;
            }
            if (t == 0) {
                a.paircnt[p] = cnt;
                if (p == P.np - 1) *a.dpw_out = dl;
            }
        }
    }
    if (FINAL) return;
    __syncthreads();
    const Box nb = reach_box(P, snew[PX], snew[PY], snew[PDIAM], snew[PZETA]);
    if (t == 0) {
        put_box(a.pbox_out + 4 * p, nb);
        if (a.spf_out) make_spf(snew, a.spf_out + p * SPF);
        bool bad = false;
        for (int f = 0; f < PCOUNT; ++f) bad = bad || !__builtin_isfinite(snew[f]);
        if (bad) set_err(a, LBM_EF_NAN);
    }
    if (P.np > NP_LOOP)
        register_tiles(P, nb, p, a.list_gen + 1u, a.tl_hdr_out, a.tl_ids_out, a.err);
}

// Resident blocks per CU the step kernels are compiled for, which with
// 4-wavefront blocks is wavefronts per SIMD (the register budget: 512 VGPRs
// per SIMD lane, 800 SGPRs per SIMD): the FP64 step 5
// without spills (6, with one spilled value per node, measured slower); the
// reduced formats 8 (60 VGPRs, the collide header's scalar registers kept
// within the 8-wavefront budget; left to the compiler they took 104 scalar
// registers and 7 wavefronts).  The batch kernel is left to the compiler
// (bounded at 8 it spills 126 scalar registers and runs 40 % slower).
#define STEP_BLOCKS_PER_CU(DEV, FINAL) ((FINAL) ? 1 : (DEV) ? 8 : 5)

// Block layout of a step launch: the particle helper blocks (particle-major,
// K per particle), the reducer slot, then the tiles in one contiguous range
// per XCD.
template <class St, bool DEV, bool FINAL, int BND>
__global__ __launch_bounds__(NT, STEP_BLOCKS_PER_CU(DEV, FINAL)) void k_step(StepArgs args) {
    const StepArgs &a = launch_args<StepArgs>();
    (void)args;
    const StepHead h = load_head(a);
    const int b = (int)blockIdx.x;
    if (b < h.n_pblocks) {
        particle_helper<St, DEV, FINAL, BND>(a, b / a.kpart, b % a.kpart);
        return;
    }
    if (b < h.tile0) {
        monitor_reduce(a.mpart_in, a.ntiles, a.red_mass, a.red_yield);
        return;
    }
    step_tile<St, DEV, FINAL, BND>(a, h, xcd_rank(b, h.tile0, h.ntiles));
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
        if (m.y % a.kpart == 0) particle_advance<DEV>(a, m.y / a.kpart);
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
    const StepHead h = load_head(a);
    if (m.y < a.n_pblocks) {
        particle_helper<St, DEV, FINAL, BND>(a, m.y / a.kpart, m.y % a.kpart);
        return;
    }
    if (m.y < a.n_pblocks + a.nred) {
        monitor_reduce(a.mpart_in, a.ntiles, a.red_mass, a.red_yield);
        return;
    }
    const long long tb = m.y - a.n_pblocks - a.nred;
    for (int r = 0; r < a.tpb; ++r) {
        const long long t = tb * a.tpb + r;
        if (t < a.ntiles) step_tile<St, DEV, FINAL, BND, false>(a, h, (int)t);
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
