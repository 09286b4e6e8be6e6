// Device-side building blocks of the D2Q9 MRT / smoothed-profile time step
// for gfx950.  Reference routines are cited as F6:line
// (6_FreeFall_YieldStress/D2Q9_BinghamSediment_Validated_20191127.f90) and
// F3:line (3_MultiParticle_Sedimentation/D2Q9_RTParticle_20191003.f90).
//
// Two per-node pipelines share the index logic:
//  * FP64, reference-exact: every expression keeps the reference's operation
//    order and float32-rounded literals (the library is built with
//    -ffp-contract=off), so it reproduces the reference up to the last bit of
//    sin/exp;
//  * FP32 deviation form: populations are handled as h_i = f_i - w_i*rho0 and
//    every moment is taken relative to the rest state, so FP32 arithmetic
//    keeps the small quantities at full relative precision.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

constexpr int NT = 256;        // threads per block (4 wavefronts of 64)
constexpr int NSLOT = 32;      // fixed-point accumulation slots per step
constexpr double MASS_SCALE = 68719476736.0;  // 2^36: fixed point of mass sums
constexpr int MAXNB = 32;      // contact partners recorded per particle

// float32-rounded literals of the reference, widened to double
constexpr double D_MRT = (double)(1.0f / 36.0f);    // F6:873
constexpr double INV6 = (double)(1.0f / 6.0f);      // F6:1202
constexpr double PI_F = (double)3.14159265359f;     // F6:3
constexpr double FEQ_W0 = (double)(2.0f / 9.0f);    // F6:1135
constexpr double FEQ_W1 = (double)(1.0f / 18.0f);   // F6:1136
constexpr double FEQ_W2 = (double)(1.0f / 36.0f);   // F6:1137
constexpr double LJ_A = (double)2.4f;               // F4:595, default-REAL literal

// Lattice weights of the rest state used by the deviation form.
constexpr float WF0 = 4.0f / 9.0f, WF1 = 1.0f / 9.0f, WF2 = 1.0f / 36.0f;

// Particle record fields, SoA with stride maxp (LBM_PF_* of lbm_amd.h).
enum { PX = 0, PY, PANG, PVX, PVY, PW, PDIAM, PFX, PFY, PTQ, PFCX, PFCY,
       PRHOP, PMP, PIP, PZETA, PVOX, PVOY, PVOW, PFHX, PFHY, PFHW, PCOUNT };
constexpr int HPC = 12;        // history doubles per particle

struct DevP {
    int nx, ny, W, H;
    long long nn;        // nodes (W*H)
    long long plane;     // SoA plane stride in elements
    int rheology, q_mode, update_angle, np;  // rheology: LBM_NEWTONIAN/BINGHAM/CASSON
    int motion;          // 0 free, 1 prescribed x-oscillation
    int boundary;        // 0 halfway bounce-back, 1 Zou-He edges, 2 equilibrium edges
    double osc_u, osc_period;
    int maxp;            // particle SoA stride
    int nbnd;            // boundary nodes
    double omega, s0, s1, s2, s3, s5, sq, nu, mu0, tau_y, m_pap, gamma_min, rho0, gy;
    double wall[8], corner[2];
    const double *wtab;  // device table of the wall velocities (WT_* pairs), FP64
    const float *wtabf;  // the same in FP32
    double zeta_c, cij, eps_p, eps_w, e_p, e_w;
    int contact;         // 0 Glowinski, 1 Wan-Turek, 2 Lennard-Jones
    double wall_edge[4]; // Wan-Turek walls: bottom y, right x, top y, left x
    unsigned w_magic;    // row = node / W by multiply-high (w_shift >= 1)
    int w_shift;
    // FP32 copies of the parameters used by the reduced-precision path
    float fomega, fs0, fs1, fs2, fs3, fs5, fsq, fnu, fmu0, ftau_y, fm_pap, fgamma_min, frho0;
    // relaxation factors of the FP32 collide: -s_k / N_k and (1 - s_k/2) c_k / N_k
    // for the uniform rates (forcing-moment constants c_k folded in), the
    // fixed-q factors, and the strain-rate factors -s_1/4 and -0.75 s_1
    float cD1a, cD1b, cD2a, cD2b, cD3a, cD3b, cD5a, cD5b, cDq1, cDq2, cS1, cS2;
    // forcing factors of D_3 / N_3 and D_5 / N_5 per unit FX, FY (with
    // n3 = -FX/2): -cD3a/2 + cD3b and -cD5a/2 + cD5b
    float cC3, cC5;
    float fgamma_min2;  // gamma_min^2: the small-rate branch compares gamma^2
    float fsqrt_mu0, fcas_lim;  // Casson: sqrt(mu0), sqrt(mu0) + sqrt(m tau_y)
    unsigned buf_bytes;  // bytes of one 9-plane storage buffer (< 2^32)
};

// Raw buffer access to the storage planes: one buffer descriptor per array,
// a per-node 32-bit vector offset and a uniform per-population offset, so the
// nine loads and stores share one address computation.  Offsets beyond
// num_records read as zero and drop writes.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             0x00020000);
}

// A load through the constant address space: data that no block writes
// during the launch, at a wavefront-uniform address, is read with scalar
// loads that need no wait for the vector memory pipe.
template <class T>
__device__ __forceinline__ T cload(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}

// Cache policy of the step's population and rate stores: SC1 (coherent at
// device scope), which writes each line through the XCD's L2 as it is
// stored.  With the default write-back policy the lines a launch writes
// stay dirty in the L2s and are written back when the kernel ends, after
// the last tile's arithmetic; written through, they drain while the tiles
// compute (measured on MI355X: a 300-FMA-per-node FP16 pull pass over
// 642,001 nodes takes 8.0 us instead of 10.7 us, a plain copy is
// unchanged).
constexpr int ST_THROUGH = 16;

// j = n / W for 0 <= n < 2^31 by the round-up multiply-high method (the
// host computes w_magic, w_shift and checks them).
__device__ __forceinline__ int row_of(const DevP &P, unsigned n) {
    unsigned t = __umulhi(P.w_magic, n);
    return (int)((t + ((n - t) >> 1)) >> (P.w_shift - 1));
}

__device__ __forceinline__ int cxk(int k) {
    return (k == 1 || k == 5 || k == 8) ? 1 : ((k == 3 || k == 6 || k == 7) ? -1 : 0);
}
__device__ __forceinline__ int cyk(int k) {
    return (k == 2 || k == 5 || k == 6) ? 1 : ((k == 4 || k == 7 || k == 8) ? -1 : 0);
}

// Index of boundary node (i, j) in the edge arrays: bottom row, top row,
// left column (j=1..ny-1), right column (j=1..ny-1); -1 for interior nodes.
__device__ __forceinline__ int bnd_index(const DevP &P, int i, int j) {
    if (j == 0) return i;
    if (j == P.ny) return P.W + i;
    if (i == 0) return 2 * P.W + (j - 1);
    if (i == P.nx) return 2 * P.W + (P.H - 2) + (j - 1);
    return -1;
}

// ---------------------------------------------------------------- storage
// Storage formats of the fused path.  FP64 stores f itself; the reduced
// formats store h = f - w_i*rho0 (times 2^10 for FP16 to stay clear of
// subnormals), so load/store are exact inverses up to the format's rounding.
// 32-bit integer hash (xorshift-multiply finaliser) for stochastic rounding.
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// st(p, v, r): r is a per-value random word for stochastic rounding (FP16,
// BF16), whose high bits are used.
__device__ __forceinline__ uint32_t sr_word(uint32_t node_hash, int k) {
    // distinct odd 24-bit multipliers spread one node hash over the 9
    // populations; the 24 x 24-bit product (low 32 bits) issues at full
    // rate, a 32-bit integer multiply at quarter rate
    const uint32_t K[9] = {0x3779B1u, 0xEBCA77u, 0xB2AE3Du, 0xD4EB2Fu, 0x5667B1u,
                           0xA2646Bu, 0x7046C5u, 0x5A4F09u, 0x5A3C2Bu};
    return __umul24(node_hash, K[k]);
}
// The conversions below store only the low half of their result; the other
// half passes through from the input register itself (cvt_sr writes one half
// of its destination, which is tied to that input), so no register is
// cleared for it.
//
// The reduced formats exchange values in storage units: S h with the
// format's scale S (2^10 for FP16, 1 otherwise).  The deviation-form collide
// is linear in h, so it runs in those units and the scale is folded into a
// few per-node factors instead of one multiply per population on each load
// and store.
struct StF64 {
    using T = double;
    __device__ static double ld(const T *p) { return *p; }
    __device__ static void st(T *p, double v, uint32_t) { *p = v; }
    __device__ static double bld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0));
    }
    __device__ static void bst(__amdgpu_buffer_rsrc_t r, double v, unsigned vo, unsigned so) {
        using W = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(W, v), r, (int)vo, (int)so,
                                              ST_THROUGH);
    }
};
struct StF32 {
    using T = float;
    using R = unsigned;  // a stored value's bits (braw), to cvt later
    static constexpr float S = 1.0f, invS = 1.0f;
    __device__ static R braw(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        return __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0);
    }
    __device__ static float cvt(R b) { return __builtin_bit_cast(float, b); }
    __device__ static float ld(const T *p) { return *p; }
    __device__ static void st(T *p, float v, uint32_t) { *p = v; }
    __device__ static float bld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
    }
    __device__ static void bst(__amdgpu_buffer_rsrc_t r, float v, unsigned vo, unsigned so,
                               uint32_t) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vo,
                                              (int)so, ST_THROUGH);
    }
};
struct StF16 {
    using T = uint16_t;
    // one 32-bit register per value (two 16-bit values packed into one
    // register would be waited for where they are packed)
    using R = unsigned;
    __device__ static R braw(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        return __builtin_amdgcn_raw_buffer_load_b16(r, (int)vo, (int)so, 0);
    }
    __device__ static float cvt(R b) {
        return (float)__builtin_bit_cast(_Float16, (unsigned short)b);
    }
    // h is stored as 2^10 h, clear of the FP16 subnormals
    static constexpr float S = 1024.0f, invS = 1.0f / 1024.0f;
    __device__ static float ld(const T *p) { return (float)__builtin_bit_cast(_Float16, *p); }
    // stochastic rounding (gfx950 v_cvt_sr_f16_f32): in the normal FP16 range
    // the top 13 bits of r are added below the kept mantissa bits before
    // truncation, which makes the stored value unbiased
    __device__ static void st(T *p, float v, uint32_t r) {
        h2 w = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(h2, v), v, r, false);
        *p = __builtin_bit_cast(uint16_t, w[0]);
    }
    __device__ static float bld(__amdgpu_buffer_rsrc_t rs, unsigned vo, unsigned so) {
        unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(rs, (int)vo, (int)so, 0);
        return (float)__builtin_bit_cast(_Float16, b);
    }
    __device__ static void bst(__amdgpu_buffer_rsrc_t rs, float v, unsigned vo, unsigned so,
                               uint32_t r) {
        h2 w = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(h2, v), v, r, false);
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, w[0]), rs,
                                              (int)vo, (int)so, ST_THROUGH);
    }
    typedef __attribute__((ext_vector_type(2))) _Float16 h2;
};
// BF16 keeps 8 significant bits.  Round-to-nearest of a quasi-steady field
// repeats the same rounding error every step, which acts as a persistent
// source term; stochastic rounding (add 16 random low bits, truncate) makes
// the stored value unbiased.
struct StBF16 {
    using T = uint16_t;
    using R = unsigned;
    __device__ static R braw(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so) {
        return __builtin_amdgcn_raw_buffer_load_b16(r, (int)vo, (int)so, 0);
    }
    __device__ static float cvt(R b) { return __builtin_bit_cast(float, b << 16); }
    static constexpr float S = 1.0f, invS = 1.0f;
    __device__ static float ld(const T *p) {
        return __builtin_bit_cast(float, (uint32_t)(*p) << 16);
    }
    // gfx950 v_cvt_sr_bf16_f32: adds the top 16 bits of r below the kept
    // mantissa bits and truncates
    __device__ static void st(T *p, float v, uint32_t r) {
        b2 w = __builtin_amdgcn_cvt_sr_bf16_f32((b2){}, v, r, false);
        *p = __builtin_bit_cast(uint16_t, w[0]);
    }
    __device__ static float bld(__amdgpu_buffer_rsrc_t rs, unsigned vo, unsigned so) {
        unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(rs, (int)vo, (int)so, 0);
        return __builtin_bit_cast(float, (uint32_t)b << 16);
    }
    __device__ static void bst(__amdgpu_buffer_rsrc_t rs, float v, unsigned vo, unsigned so,
                               uint32_t r) {
        b2 w = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(b2, v), v, r, false);
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, w[0]), rs,
                                              (int)vo, (int)so, ST_THROUGH);
    }
    typedef __attribute__((ext_vector_type(2))) __bf16 b2;
};

__device__ __forceinline__ float wdev(int k) {
    return k == 0 ? WF0 : (k < 5 ? WF1 : WF2);
}

// --------------------------------------------------------- FP64 pipeline
// MRT relaxation vector for a node rate (ComputeRelaxation_MRT, F6:1002-1011).
__device__ __forceinline__ void rates64(const DevP &P, double om, double s[9]) {
    double q = P.q_mode ? (8.0 * (2.0 - om)) / (8.0 - om) : P.sq;
    s[0] = P.s0; s[1] = P.s1; s[2] = P.s2; s[3] = P.s3; s[4] = q;
    s[5] = P.s5; s[6] = q; s[7] = om; s[8] = om;
}

// Wall velocity of a boundary node from the handle's device table (pairs
// (u_x, u_y) of the corners, top, bottom, right and left walls, in FP64 and
// FP32).  The edge rules read it only on boundary nodes: a choice among the
// wall fields of the parameter block would either be a run-time index into
// it (which the compiler copies to scratch memory) or eight loads that it
// hoists to the kernel entry, where they occupy scalar registers for the
// whole kernel and push the collide's parameters into spills.
template <class T> __device__ __forceinline__ const T *wall_table(const DevP &P);
template <> __device__ __forceinline__ const double *wall_table<double>(const DevP &P) {
    return P.wtab;
}
template <> __device__ __forceinline__ const float *wall_table<float>(const DevP &P) {
    return P.wtabf;
}
enum { WT_CORNER = 0, WT_TOP, WT_BOTTOM, WT_RIGHT, WT_LEFT };

// Wall velocity (u, v) of edge class wc for a lane: the whole table is read
// with scalar loads (it is uniform and tiny) and the lane's pair selected,
// so that no lane waits on a load whose address depends on its node.
template <class T>
__device__ __forceinline__ void wall_velocity(const DevP &P, int wc, T &u, T &v) {
    const T *t = wall_table<T>(P);
    u = cload(t + 2 * WT_CORNER);
    v = cload(t + 2 * WT_CORNER + 1);
#pragma unroll
    for (int q = WT_TOP; q <= WT_LEFT; ++q) {
        const T tu = cload(t + 2 * q), tv = cload(t + 2 * q + 1);
        u = wc == q ? tu : u;
        v = wc == q ? tv : v;
    }
}

// Zou-He edge rule (F1:1246-1258 top and bottom, F1:1319-1341 left and
// right) at edge node (i, j), corners excluded: the three unknown
// populations from the node's own streamed ones and the wall velocity
// (wall_table: P.wall's uxS,uyS, uxR,uyR, uxN,uyN, uxL,uyL), in F1's
// operation order.
// v holds f (rho_base = 0) or deviations f - w_i rho0 (rho_base = rho0:
// the density sums of the rule pick up exactly rho0 from the weights).
//
// The four edges share one form, evaluated once with per-edge operands
// selected by value (a branch per edge, each writing its own three
// elements, would make the compiler merge those stores into one with a
// run-time index and keep the population array in scratch memory):
//   r     = (rho_base + ((v0 + a1) + a2 + 2 ((b1 + b2) + b3))) / (1 + s_n u_n)
//   axis  = b1 - s_n ((4/6) r) u_n
//   diag1 = ((c1 - D) + s11 T1) + s12 T2,  diag2 = ((c2 + D) + s21 T1) + s22 T2
// with D = 0.5 (a1 - a2), T1 = (k1 r) ux, T2 = (k2 r) uy; x + (-y) rounds as
// x - y, so every edge keeps F1's expression bit for bit.
//   edge    a1,a2  b1,b2,b3  s_n u_n  c1->k  c2->k  k1,k2      s11 s12 s21 s22
//   top     v1,v3  v2,v6,v5  +  uy    v6->8  v5->7  1/2, 1/6   +   -   -   -
//   bottom  v1,v3  v4,v7,v8  -  uy    v7->5  v8->6  1/2, 1/6   +   +   -   +
//   left    v2,v4  v3,v6,v7  -  ux    v7->5  v6->8  1/6, 1/2   +   +   +   -
//   right   v2,v4  v1,v5,v8  +  ux    v8->6  v5->7  1/6, 1/2   -   +   -   -
// (axis -> 4, 2, 1, 3 for top, bottom, left, right.)
template <class T>
__device__ __forceinline__ void zou_he_edge(const DevP &P, int i, int j, T rho_base, T v[9]) {
    const T inv6 = (T)INV6, half = (T)0.5;
    const bool top = j == P.ny, bot = !top && j == 0, lft = !top && !bot && i == 0;
    const bool rgt = !top && !bot && !lft;
    const bool vert = top || bot;  // top or bottom edge: normal along y
    const int wc = top ? WT_TOP : bot ? WT_BOTTOM : lft ? WT_LEFT : WT_RIGHT;
    T ux, uy;
    wall_velocity<T>(P, wc, ux, uy);
    const T a1 = vert ? v[1] : v[2], a2 = vert ? v[3] : v[4];
    const T b1 = top ? v[2] : bot ? v[4] : lft ? v[3] : v[1];
    const T b2 = top ? v[6] : bot ? v[7] : lft ? v[6] : v[5];
    const T b3 = top ? v[5] : bot ? v[8] : lft ? v[7] : v[8];
    const T c1 = top ? v[6] : (bot || lft) ? v[7] : v[8];
    const T c2 = top ? v[5] : bot ? v[8] : lft ? v[6] : v[5];
    const T un = vert ? uy : ux;
    const bool pos = top || rgt;  // s_n = +1
    const T sn_un = pos ? un : -un;
    const T r = (rho_base + (v[0] + a1 + a2 + (T)2. * (b1 + b2 + b3))) / ((T)1.0 + sn_un);
    const T ka = (T)4. * inv6 * r * un;
    const T axis = pos ? b1 - ka : b1 + ka;
    const T d = half * (a1 - a2);
    const T t1 = (vert ? half : inv6) * r * ux;
    const T t2 = (vert ? inv6 : half) * r * uy;
    const T s11 = rgt ? -t1 : t1, s12 = top ? -t2 : t2;
    const T s21 = lft ? t1 : -t1, s22 = (bot ? t2 : -t2);
    const T diag1 = c1 - d + s11 + s12;
    const T diag2 = c2 + d + s21 + s22;
    v[1] = lft ? axis : v[1];
    v[2] = bot ? axis : v[2];
    v[3] = rgt ? axis : v[3];
    v[4] = top ? axis : v[4];
    v[5] = (bot || lft) ? diag1 : v[5];
    v[6] = bot ? diag2 : (rgt ? diag1 : v[6]);
    v[7] = (top || rgt) ? diag2 : v[7];
    v[8] = top ? diag1 : (lft ? diag2 : v[8]);
}

// COMPUTEFEQNEW (F6:1130-1148, F4:1140-1158): the equilibrium with the
// float-rounded weights, in the reference's operation order.
__device__ __forceinline__ void feq64(double u, double v, double r, double fe[9]) {
    double u2 = u * u + v * v;
    fe[0] = FEQ_W0 * r * (2.0 - 3.0 * u2);
    fe[1] = FEQ_W1 * r * (2.0 + 6.0 * u + 9.0 * u * u - 3.0 * u2);
    fe[2] = FEQ_W1 * r * (2.0 + 6.0 * v + 9.0 * v * v - 3.0 * u2);
    fe[3] = FEQ_W1 * r * (2.0 - 6.0 * u + 9.0 * u * u - 3.0 * u2);
    fe[4] = FEQ_W1 * r * (2.0 - 6.0 * v + 9.0 * v * v - 3.0 * u2);
    fe[5] = FEQ_W2 * r * (1.0 + 3.0 * (u + v) + 9.0 * u * v + 3.0 * u2);
    fe[6] = FEQ_W2 * r * (1.0 - 3.0 * (u - v) - 9.0 * u * v + 3.0 * u2);
    fe[7] = FEQ_W2 * r * (1.0 - 3.0 * (u + v) + 9.0 * u * v + 3.0 * u2);
    fe[8] = FEQ_W2 * r * (1.0 + 3.0 * (u - v) - 9.0 * u * v + 3.0 * u2);
}

// The same equilibrium as deviations from w_k rho0 in FP32, for density
// deviation rd = rho - rho0: W_k (rd p_k + rho0 (p_k - b_k)), with p_k the
// bracket of feq64, b_k its value at rest (2 or 1) and W_k the float-rounded
// weights (W_k b_k equals the float weight w_k exactly: doubling commutes
// with rounding), so no term of size rho0 is formed and cancelled.
__device__ __forceinline__ void feq_dev(float rho0, float u, float v, float rd, float h[9]) {
    const float W0 = 2.0f / 9.0f, W1 = 1.0f / 18.0f, W2 = 1.0f / 36.0f;
    const float u2 = u * u + v * v;
    const float d[9] = {-3.0f * u2,
                        6.0f * u + 9.0f * u * u - 3.0f * u2,
                        6.0f * v + 9.0f * v * v - 3.0f * u2,
                        -6.0f * u + 9.0f * u * u - 3.0f * u2,
                        -6.0f * v + 9.0f * v * v - 3.0f * u2,
                        3.0f * (u + v) + 9.0f * u * v + 3.0f * u2,
                        -3.0f * (u - v) - 9.0f * u * v + 3.0f * u2,
                        -3.0f * (u + v) + 9.0f * u * v + 3.0f * u2,
                        3.0f * (u - v) - 9.0f * u * v + 3.0f * u2};
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float Wk = k == 0 ? W0 : (k < 5 ? W1 : W2);
        const float bk = k < 5 ? 2.0f : 1.0f;
        h[k] = Wk * (rd * (bk + d[k]) + rho0 * d[k]);
    }
}

// Equilibrium edges and corners (F4:1218-1220, F4:1277-1279, F4:1368-1373):
// every population of boundary node (i, j) is the equilibrium of the
// previous step's density and corrected velocity at interior node
// (clamp(i, 1, nx-1), clamp(j, 1, ny-1)).  The fused path carries those
// three fields per boundary node (edge arrays, SoA [rho | ux | uy] of nbnd
// entries each); eq_sources calls put(b) for every boundary node b whose
// source is interior node (i, j): up to three (an interior corner feeds
// both edges and the domain corner).
template <class Put>
__device__ __forceinline__ void eq_sources(const DevP &P, int i, int j, Put put) {
    const bool l = i == 1, r = i == P.nx - 1, b = j == 1, t = j == P.ny - 1;
    if (!(l || r || b || t)) return;
    const int bi[3] = {i, l ? 0 : -1, r ? P.nx : -1};
    const int bj[3] = {j, b ? 0 : -1, t ? P.ny : -1};
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y)
            if ((x | y) && bi[x] >= 0 && bj[y] >= 0) put(bnd_index(P, bi[x], bj[y]));
}

// Population k of boundary node (i, j) (l, r, b, t: on the left, right,
// bottom, top edge) is replaced by the bounce-back rule (F6:1206-1294 edges,
// F6:1364-1386 corners): its source lies outside the domain and, at a
// corner, it points into the domain (the two tangential diagonals keep the
// node's own value).  opp(k): the population of opposite velocity.
__device__ __forceinline__ bool bb_replaced(int k, bool l, bool r, bool b, bool t) {
    switch (k) {
    case 1: return l;
    case 2: return b;
    case 3: return r;
    case 4: return t;
    case 5: return (b && !r) || (l && !t);
    case 6: return (b && !l) || (r && !t);
    case 7: return (t && !l) || (r && !b);
    case 8: return (t && !r) || (l && !b);
    default: return false;
    }
}
__device__ __forceinline__ constexpr int opp_k(int k) {
    return k == 0 ? 0 : (k <= 4 ? (k + 1) % 4 + 1 : (k - 3) % 4 + 5);
}

// Halfway bounce-back with moving walls at boundary node (i, j), in one
// branch-free form: each replaced population k (bb_replaced) becomes
// fo_opp(k) - c rw (c_opp(k) . u_w), with c rw precomputed as
// a46 = ((4/6) clb^2) rw along the axes and a6 = ((1/6) clb^2) rw on the
// diagonals, and u_w the node's wall velocity (the corner velocity at the
// four corners).  On entry v[k] already holds fo_opp(k), the node's own
// post-collision population opp(k), for every replaced k.  For every edge
// and corner the operands and their order are those of the reference's
// expressions ((-x) - y, x - y, x + y, (-x) + y), so the result is bitwise
// the same.  (A branch per edge would write different elements of v in each
// branch, which the compiler merges into one store at a run-time index and
// so keeps v in scratch memory.)
template <class T>
__device__ __forceinline__ void bounce_back_node(const DevP &P, int i, int j, T a6, T a46,
                                                 T v[9]) {
    const bool l = i == 0, r = i == P.nx, b = j == 0, t = j == P.ny;
    const bool cor = (l || r) && (b || t);
    const int wc = cor ? WT_CORNER : t ? WT_TOP : b ? WT_BOTTOM : r ? WT_RIGHT : WT_LEFT;
    T uw, vw;
    wall_velocity<T>(P, wc, uw, vw);
    v[1] = l ? v[1] - a46 * (-uw) : v[1];
    v[2] = b ? v[2] - a46 * (-vw) : v[2];
    v[3] = r ? v[3] - a46 * (uw) : v[3];
    v[4] = t ? v[4] - a46 * (vw) : v[4];
    v[5] = bb_replaced(5, l, r, b, t) ? v[5] - a6 * (-uw - vw) : v[5];
    v[6] = bb_replaced(6, l, r, b, t) ? v[6] - a6 * (uw - vw) : v[6];
    v[7] = bb_replaced(7, l, r, b, t) ? v[7] - a6 * (uw + vw) : v[7];
    v[8] = bb_replaced(8, l, r, b, t) ? v[8] - a6 * (-uw + vw) : v[8];
}

// Pull-stream with in-domain sources (F6:1164-1188) and the boundary rules
// for node (i, j): v receives f_k.  src: post-collision populations (SoA
// planes), erho: previous-step density on the boundary nodes (equilibrium
// edges: the edge arrays of eq_sources).  BND: the boundary mode when it is
// known at compile time (the fused kernels are instantiated per mode, so a
// kernel carries only its own edge rule), -1 to read P.boundary.
template <class St, int BND = -1>
__device__ __forceinline__ void load_fk64(const DevP &P, const typename St::T *src,
                                          const double *erho, long long node, int i,
                                          int j, double v[9]) {
    const int bnd = BND >= 0 ? BND : P.boundary;
    if (i > 0 && j > 0 && i < P.nx && j < P.ny) {  // interior: every source in the domain
#pragma unroll
        for (int k = 0; k < 9; ++k)
            v[k] = St::ld(src + k * P.plane + node - cxk(k) - (long long)cyk(k) * P.W);
        return;
    }
    if (bnd == 2) {  // boundary node: the equilibrium of its interior source
        const int b = bnd_index(P, i, j);
        feq64(erho[P.nbnd + b], erho[2 * P.nbnd + b], erho[b], v);
        return;
    }
    // a source outside the domain reads the node itself: population opp(k)
    // where the bounce-back rule replaces k (bb_replaced), k itself otherwise
    const bool el = i == 0, er = i == P.nx, eb = j == 0, et = j == P.ny;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        int si = i - cxk(k), sj = j - cyk(k);
        bool in = (unsigned)si <= (unsigned)P.nx && (unsigned)sj <= (unsigned)P.ny;
        long long s = in ? node - cxk(k) - (long long)cyk(k) * P.W : node;
        const int kk = (!in && bb_replaced(k, el, er, eb, et)) ? opp_k(k) : k;
        v[k] = St::ld(src + kk * P.plane + s);
    }
    if (i == 0 || j == 0 || i == P.nx || j == P.ny) {
        const bool corner = (i == 0 || i == P.nx) && (j == 0 || j == P.ny);
        if (bnd == 1 && !corner) {
            zou_he_edge<double>(P, i, j, 0.0, v);
            return;
        }
        const double clb2 = 1.0;
        const double rw = erho[bnd_index(P, i, j)];
        bounce_back_node<double>(P, i, j, INV6 * clb2 * rw, 4.0 * INV6 * clb2 * rw, v);
    }
}

// Raw buffer loads of the nine pulled FP64 populations of an interior node
// (every source in the domain): one vector offset for the node, a uniform
// offset per population.
__device__ __forceinline__ void load_interior64(const DevP &P, const double *src, long long node,
                                                double v[9]) {
    const auto rs = make_rsrc(src, P.buf_bytes);
    const unsigned vo = (unsigned)node * 8u;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const unsigned so = (unsigned)((k * P.plane - cxk(k) - (long long)cyk(k) * P.W) * 8LL);
        v[k] = StF64::bld(rs, vo, so);
    }
}

// MACROSCOPICVARIABLES, F6:1405-1408.
__device__ __forceinline__ void macros64(const double v[9], double &rho, double &ux,
                                         double &uy) {
    rho = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7] + v[8];
    double rinv = 1.0 / rho;
    ux = rinv * (v[1] + v[5] + v[8] - (v[6] + v[3] + v[7]));
    uy = rinv * (v[5] + v[2] + v[6] - (v[7] + v[4] + v[8]));
}

// sin x for |x| <= pi/2 from its Taylor series through x^21 (the first term
// left out, x^23/23!, is below 1e-18 |x| there), evaluated by Horner's rule
// in x^2 with fused multiply-adds: within about an ulp of sin, with a small
// fraction of the registers of the general-argument sin.
__device__ __forceinline__ double sin_halfpi(double x) {
    const double x2 = x * x;
    // sin x = x + x^3 (c1 + c2 x^2 + ... + c10 x^18), c_k = (-1)^k / (2k+1)!
    double p = 1.0 / 51090942171709440000.0;                 // c10 = 1/21!
    p = __builtin_fma(p, x2, -1.0 / 121645100408832000.0);  // c9 = -1/19!
    p = __builtin_fma(p, x2, 1.0 / 355687428096000.0);      // c8 = 1/17!
    p = __builtin_fma(p, x2, -1.0 / 1307674368000.0);       // c7 = -1/15!
    p = __builtin_fma(p, x2, 1.0 / 6227020800.0);           // c6 = 1/13!
    p = __builtin_fma(p, x2, -1.0 / 39916800.0);            // c5 = -1/11!
    p = __builtin_fma(p, x2, 1.0 / 362880.0);               // c4 = 1/9!
    p = __builtin_fma(p, x2, -1.0 / 5040.0);                // c3 = -1/7!
    p = __builtin_fma(p, x2, 1.0 / 120.0);                  // c2 = 1/5!
    p = __builtin_fma(p, x2, -1.0 / 6.0);                   // c1 = -1/3!
    return __builtin_fma(x * x2, p, x);
}

// Solid fraction of one disk at node (i, j), F6:451-459.  POLY: the interface
// sine by sin_halfpi (its argument is within pi/2 there), as the fused FP64
// step evaluates it; otherwise the library sine.
template <bool POLY = false>
__device__ __forceinline__ double phi_at(double xc, double yc, double diam, double zeta,
                                         int i, int j) {
    double rp = 0.5 * diam;
    double dx = (double)i - xc, dy = (double)j - yc;
    double di = rp - sqrt(dx * dx + dy * dy);
    double ph = 0.0;
    if (di <= -(0.5 * zeta))
        ph = 0.0;
    else if (fabs(di) < 0.5 * zeta)
        ph = 0.5 * ((POLY ? sin_halfpi(PI_F * di / zeta) : sin(PI_F * di / zeta)) + 1.0);
    else if (di >= 0.5 * zeta)
        ph = 1.0;
    return ph;
}

// Conservative test: can particle p have phi > 0 at node (i, j)?  phi > 0
// needs distance < r + zeta/2; one lattice unit of margin covers rounding.
__device__ __forceinline__ bool in_reach(const double *part, int maxp, int p, int i, int j) {
    double reach = 0.5 * part[PDIAM * maxp + p] + 0.5 * part[PZETA * maxp + p] + 1.0;
    return fabs((double)i - part[PX * maxp + p]) <= reach &&
           fabs((double)j - part[PY * maxp + p]) <= reach;
}

// ComputeForceField_SPM at one node (F6:485-495): the highest particle index
// with phi > 0 sets G.  Returns true when some particle covers the node.
__device__ __forceinline__ bool spm_node(const double *part, int maxp, int np, int i, int j,
                                         double ux, double uy, double &gx, double &gy) {
    gx = 0.0;
    gy = 0.0;
    bool any = false;
    for (int p = 0; p < np; ++p) {
        if (!in_reach(part, maxp, p, i, j)) continue;
        double xc = part[PX * maxp + p], yc = part[PY * maxp + p];
        double ph = phi_at(xc, yc, part[PDIAM * maxp + p], part[PZETA * maxp + p], i, j);
        if (ph > 0.0) {
            double w = part[PW * maxp + p];
            double uxp = part[PVX * maxp + p] - w * ((double)j - yc);
            double uyp = part[PVY * maxp + p] + w * ((double)i - xc);
            gx = ph * (uxp - ux);
            gy = ph * (uyp - uy);
            any = true;
        }
    }
    return any;
}

// Same rule for a short candidate list (many-particle runs): the largest
// index with phi > 0 sets G, whatever the order of the list.
__device__ __forceinline__ bool spm_node_cand(const double *part, int maxp, const int *ids,
                                              int cnt, int i, int j, double ux, double uy,
                                              double &gx, double &gy) {
    gx = 0.0;
    gy = 0.0;
    int best = -1;
    double bphi = 0.0;
    for (int c = 0; c < cnt; ++c) {
        int p = ids[c];
        if (p <= best || !in_reach(part, maxp, p, i, j)) continue;
        double ph = phi_at(part[PX * maxp + p], part[PY * maxp + p], part[PDIAM * maxp + p],
                           part[PZETA * maxp + p], i, j);
        if (ph > 0.0) { best = p; bphi = ph; }
    }
    if (best < 0) return false;
    double xc = part[PX * maxp + best], yc = part[PY * maxp + best], w = part[PW * maxp + best];
    double uxp = part[PVX * maxp + best] - w * ((double)j - yc);
    double uyp = part[PVY * maxp + best] + w * ((double)i - xc);
    gx = bphi * (uxp - ux);
    gy = bphi * (uyp - uy);
    return true;
}

// ------------------------------------------------ particle state in a launch
// Reach box of a disk: the nodes that can have phi > 0 (distance below
// r + zeta/2, plus one lattice unit that covers rounding), clipped to the
// grid; empty when i0 > i1 or j0 > j1.
struct Box {
    int i0, i1, j0, j1;
};
__device__ __forceinline__ Box reach_box(const DevP &P, double x, double y, double diam,
                                         double zeta) {
    const double reach = 0.5 * diam + 0.5 * zeta + 1.0;
    Box b;
    b.i0 = max(0, (int)ceil(x - reach)); b.i1 = min(P.nx, (int)floor(x + reach));
    b.j0 = max(0, (int)ceil(y - reach)); b.j1 = min(P.ny, (int)floor(y + reach));
    return b;
}

// Does tile t hold a node of the box [i0, i1] x [j0, j1] (bounds may lie
// outside the grid)?  Uniform over the block.
// Nodes n0..n1 of a tile (rows jl..jh) and the box i0..i1 x j0..j1 share a
// node: a full row between the tile's first and last row meets any box
// column range; otherwise the (at most two) partial rows in the box's row
// range are tested.
struct TileSpan {
    unsigned n0, n1;
    int jl, jh;
};
__device__ __forceinline__ bool tile_meets_box(int W, const TileSpan &t, int i0, int i1, int j0,
                                               int j1) {
    const int ja = max(t.jl, j0), jb = min(t.jh, j1);
    if (ja > jb || i0 > i1) return false;
    if (jb - ja >= 2) return true;
    auto seg = [&](int j) {
        const int lo = j == t.jl ? (int)(t.n0 - (unsigned)j * (unsigned)W) : 0;
        const int hi = j == t.jh ? (int)(t.n1 - (unsigned)j * (unsigned)W) : W - 1;
        return max(lo, i0) <= min(hi, i1);
    };
    return seg(ja) || seg(jb);
}
// The same test with a box that differs between lanes (one particle per
// lane), branch-free: each clause is an integer that is <= 0 when it holds
// (rows share a node, the box has a column, a full row is shared or one of
// the two partial rows meets the box's columns), so the test is a
// min/max/select chain ending in one compare.  `live` false: no node.
__device__ __forceinline__ bool span_meets_box_lane(int W, const TileSpan &t, int i0, int i1,
                                                    int j0, int j1, bool live) {
    const int ja = max(t.jl, j0), jb = min(t.jh, j1);
    const int d = jb - ja;
    const int iF = (int)(t.n0 - (unsigned)t.jl * (unsigned)W);
    const int iL = (int)(t.n1 - (unsigned)t.jh * (unsigned)W);
    const int lo_a = ja == t.jl ? iF : 0, hi_a = ja == t.jh ? iL : W - 1;
    const int lo_b = jb == t.jl ? iF : 0, hi_b = jb == t.jh ? iL : W - 1;
    const int mx_a = max(lo_a, i0) - min(hi_a, i1);
    const int mx_b = max(lo_b, i0) - min(hi_b, i1);
    const int any = min(min(2 - d, mx_a), mx_b);
    const int v = max(max(-d, i0 - i1), any);
    return (live ? v : 1) <= 0;
}

// Force partials cross between blocks of one launch as generation-tagged
// 64-bit words: (tag << 32) | 32-bit payload, written and read with
// agent-scope atomic accesses.  A reader polls until each word it needs
// carries the launch's tag; a word is a single atomic location holding both
// the tag and its value, so a word with the right tag holds the right value
// and no ordering between different words is needed.  Tags never repeat on
// a handle.  A partial (force x, y, torque) is 6 words: hi and lo halves.
constexpr int SLOTW = 6;
constexpr int TAG_SPIN = 1 << 16;  // bounded wait for tagged words (tens of ms)

// Sum of x over the 64 lanes of a wavefront by DPP lane moves, in a fixed
// order: pairs, quads, half rows, rows (each lane of a group ends with the
// group's sum), then row 1 += row 0 and row 3 += row 2, then rows 2 and 3
// += the sum of rows 0 and 1.  The total is valid in lane 63.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWS, 0xf, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWS, 0xf, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double wave_sum_f64(double x) {
    x = x + dpp_f64<0xB1, 0xF>(x);   // quad_perm [1, 0, 3, 2]
    x = x + dpp_f64<0x4E, 0xF>(x);   // quad_perm [2, 3, 0, 1]
    x = x + dpp_f64<0x141, 0xF>(x);  // row_half_mirror
    x = x + dpp_f64<0x140, 0xF>(x);  // row_mirror
    x = x + dpp_f64<0x142, 0xA>(x);  // row_bcast:15 into rows 1 and 3
    x = x + dpp_f64<0x143, 0xC>(x);  // row_bcast:31 into rows 2 and 3
    return x;
}


__device__ __forceinline__ void put_tagged_f64(unsigned long long *w, unsigned tag, double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const unsigned long long t = (unsigned long long)tag << 32;
    __hip_atomic_store(w, t | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(w + 1, t | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// FP32 profile record of a disk for the reduced-precision grid pass: the
// centre split into an integer and a fractional part (node offsets i - X
// stay exact up to the rounding of the fraction), velocity, spin, radius,
// half width, pi/zeta.
enum { SX = 0, SXF, SY, SYF, SVX, SVY, SW, SRP, SHZ, SPZ, SPAD0, SPAD1, SPF };
__device__ __forceinline__ void make_spf(const double *rec /* PCOUNT */, float *out) {
    const double x = rec[PX], y = rec[PY];
    const double xi = floor(x), yi = floor(y);
    out[SX] = (float)xi; out[SXF] = (float)(x - xi);
    out[SY] = (float)yi; out[SYF] = (float)(y - yi);
    out[SVX] = (float)rec[PVX]; out[SVY] = (float)rec[PVY]; out[SW] = (float)rec[PW];
    out[SRP] = (float)(0.5 * rec[PDIAM]);
    out[SHZ] = (float)(0.5 * rec[PZETA]);
    out[SPZ] = (float)(PI_F / rec[PZETA]);
    out[SPAD0] = 0.0f; out[SPAD1] = 0.0f;
}
struct Prof {
    float x, xf, y, yf, vx, vy, w, rp, hz, pz;
};

// phi of one particle at node (i, j) in FP32 (the piecewise profile of
// phi_at, F6:451-459) from the profile fields; dx, dy receive the node
// offset from the centre.
__device__ __forceinline__ float phi_prof(const Prof &s, int i, int j, float &dx, float &dy) {
    dx = ((float)i - s.x) - s.xf;
    dy = ((float)j - s.y) - s.yf;
    const float di = s.rp - __builtin_amdgcn_sqrtf(dx * dx + dy * dy);
    if (di <= -s.hz) return 0.0f;
    if (fabsf(di) < s.hz) return 0.5f * (__sinf(s.pz * di) + 1.0f);
    return 1.0f;
}

constexpr int NP_LOOP = 16;    // above this many particles, tiles use candidate lists
constexpr int MAXC = 64;       // candidate particles per list unit
constexpr int LIST_UNIT = 64;  // nodes per candidate list: one wavefront's nodes
constexpr int NP_LDS = 128;    // up to this many particles, contact inputs are staged in LDS

// Registers particle p in the candidate list of every list unit (the
// LIST_UNIT consecutive nodes of one wavefront) that its reach box b
// touches, for generation gen (the tag of the launch whose tiles read the
// lists).  Each unit header packs (generation << 32 | count); the first
// registration of a new generation resets the count (lock-free CAS loop,
// one contender per touching particle).  Called by all NT threads of a
// block: thread r takes box rows j0+r, j0+r+NT, ...; a unit that row j
// shares with row j-1 is left to row j-1, so each unit is registered once
// per particle.  (A box narrower than a unit touches one or two units per
// row, as many as it would 256-node tiles, and a wavefront then meets only
// the particles that reach its own nodes.)
__device__ __forceinline__ void register_tiles(const DevP &P, Box b, int p, unsigned gen,
                                               unsigned long long *hdr, int *ids, int *err) {
    if (b.i0 > b.i1 || b.j0 > b.j1) return;
    const int i0 = b.i0, i1 = b.i1, j0 = b.j0, j1 = b.j1;
    for (int j = j0 + (int)threadIdx.x; j <= j1; j += NT) {
        long long tlo = ((long long)j * P.W + i0) / LIST_UNIT;
        const long long thi = ((long long)j * P.W + i1) / LIST_UNIT;
        if (j > j0) {
            long long prev = ((long long)(j - 1) * P.W + i1) / LIST_UNIT;
            if (prev >= tlo) tlo = prev + 1;
        }
        for (long long t = tlo; t <= thi; ++t) {
            unsigned long long old = hdr[t], assumed, nv;
            unsigned slot;
            do {
                assumed = old;
                unsigned g = (unsigned)(assumed >> 32);
                slot = (g == gen) ? (unsigned)(assumed & 0xffffffffu) : 0u;
                nv = ((unsigned long long)gen << 32) | (unsigned long long)(slot + 1);
                old = atomicCAS(hdr + t, assumed, nv);
            } while (old != assumed);
            if (slot < (unsigned)MAXC) ids[t * MAXC + slot] = p;
            else atomicOr(err, 1);
        }
    }
}

struct CollideOut {
    double omega_bn;     // new node rate (Bingham)
    double stress[3];
    int yielded;         // tau_mag > tau_y
};

__device__ __forceinline__ void rheology64(const DevP &P, double rho, double om, double d11,
                                           double d22, double d12, CollideOut &o);

// Collision_MRT at one node, F6:878-958, with the node's own rate om
// (Newtonian runs pass OMEGA).  f is overwritten by the post-collision values.
// REF: the Chai strain rate in the reference's operation order (F6:903-916:
// nine strain moments f1, then their sums); otherwise from the three column
// sums those sums reduce to -- with e_k = s_k (m_k - meq_k),
//   f1_1+f1_5+f1_6+f1_3+f1_7+f1_8 = 24 e0 + 6 e1 + 18 e7,
//   f1_5+f1_2+f1_6+f1_7+f1_4+f1_8 = 24 e0 + 6 e1 - 18 e7,
//   f1_5-f1_6+f1_7-f1_8 = 36 e8 --
// which keeps four non-equilibrium moments live instead of eighteen values
// (the fused step's register budget) and rounds differently in the last bit.
template <bool REF = true>
__device__ __forceinline__ void collide64(const DevP &P, double f[9], double rho, double ux,
                                          double uy, double gx, double gy, double om,
                                          CollideOut &o) {
    double s[9];
    rates64(P, om, s);
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
    // ComputeSourceEquilib_MRT, F6:964-989
    double fx = rho * (gx + 0.0), fy = rho * (gy + 0.0);
    mf[0] = 0.0;
    mf[1] = 6.0 * (fx * ux + fy * uy);
    mf[2] = -(6.0 * (fx * ux + fy * uy));
    mf[3] = fx; mf[4] = -fx; mf[5] = fy; mf[6] = -fy;
    mf[7] = 2.0 * (fx * ux - fy * uy);
    mf[8] = fy * ux + fx * uy;
    double u2 = ux * ux + uy * uy;
    meq[0] = rho;
    meq[1] = rho * (-2.0 + 3.0 * u2);
    meq[2] = rho * (1.0 - 3.0 * u2);
    meq[3] = rho * (ux); meq[4] = rho * (-ux);
    meq[5] = rho * (uy); meq[6] = rho * (-uy);
    meq[7] = rho * (ux * ux - uy * uy);
    meq[8] = rho * (ux * uy);
    double e0 = 0.0, e1 = 0.0, e7 = 0.0, e8 = 0.0;
    if constexpr (REF) {
#pragma unroll
        for (int k = 0; k < 9; ++k) mne[k] = m[k] - meq[k];
    } else {
        e0 = s[0] * (m[0] - meq[0]);
        e1 = s[1] * (m[1] - meq[1]);
        e7 = s[7] * (m[7] - meq[7]);
        e8 = s[8] * (m[8] - meq[8]);
    }
#pragma unroll
    for (int k = 0; k < 9; ++k)
        m[k] = s[k] * meq[k] + (1.0 - s[k]) * m[k] + (1.0 - 0.5 * s[k]) * mf[k];
    const double d = D_MRT;
    f[0] = d * (4.0 * (m[0] - m[1] + m[2]));
    f[1] = d * (4.0 * m[0] - m[1] - 2.0 * m[2] + 6.0 * m[3] - 6.0 * m[4] + 9.0 * m[7]);
    f[2] = d * (4.0 * m[0] - m[1] - 2.0 * m[2] + 6.0 * m[5] - 6.0 * m[6] - 9.0 * m[7]);
    f[3] = d * (4.0 * m[0] - m[1] - 2.0 * m[2] - 6.0 * m[3] + 6.0 * m[4] + 9.0 * m[7]);
    f[4] = d * (4.0 * m[0] - m[1] - 2.0 * m[2] - 6.0 * m[5] + 6.0 * m[6] - 9.0 * m[7]);
    f[5] = d * (4.0 * m[0] + 2.0 * m[1] + m[2] + 6.0 * m[3] + 3.0 * m[4] + 6.0 * m[5] + 3.0 * m[6] + 9.0 * m[8]);
    f[6] = d * (4.0 * m[0] + 2.0 * m[1] + m[2] - 6.0 * m[3] - 3.0 * m[4] + 6.0 * m[5] + 3.0 * m[6] - 9.0 * m[8]);
    f[7] = d * (4.0 * m[0] + 2.0 * m[1] + m[2] - 6.0 * m[3] - 3.0 * m[4] - 6.0 * m[5] - 3.0 * m[6] + 9.0 * m[8]);
    f[8] = d * (4.0 * m[0] + 2.0 * m[1] + m[2] + 6.0 * m[3] + 3.0 * m[4] - 6.0 * m[5] - 3.0 * m[6] - 9.0 * m[8]);
    // Chai strain rate, F6:903-916
    double c11, c22, c12;
    if constexpr (REF) {
    f1[1] = 4.0 * s[0] * mne[0] - 1.0 * s[1] * mne[1] - 2.0 * s[2] * mne[2] + 6.0 * s[3] * mne[3] - 6.0 * s[4] * mne[4] + 9.0 * s[7] * mne[7];
    f1[2] = 4.0 * s[0] * mne[0] - 1.0 * s[1] * mne[1] - 2.0 * s[2] * mne[2] + 6.0 * s[5] * mne[5] - 6.0 * s[6] * mne[6] - 9.0 * s[7] * mne[7];
    f1[3] = 4.0 * s[0] * mne[0] - 1.0 * s[1] * mne[1] - 2.0 * s[2] * mne[2] - 6.0 * s[3] * mne[3] + 6.0 * s[4] * mne[4] + 9.0 * s[7] * mne[7];
    f1[4] = 4.0 * s[0] * mne[0] - 1.0 * s[1] * mne[1] - 2.0 * s[2] * mne[2] - 6.0 * s[5] * mne[5] + 6.0 * s[6] * mne[6] - 9.0 * s[7] * mne[7];
    f1[5] = 4.0 * s[0] * mne[0] + 2.0 * s[1] * mne[1] + 1.0 * s[2] * mne[2] + 6.0 * s[3] * mne[3] + 3.0 * s[4] * mne[4] + 6.0 * s[5] * mne[5] + 3.0 * s[6] * mne[6] + 9.0 * s[8] * mne[8];
    f1[6] = 4.0 * s[0] * mne[0] + 2.0 * s[1] * mne[1] + 1.0 * s[2] * mne[2] - 6.0 * s[3] * mne[3] - 3.0 * s[4] * mne[4] + 6.0 * s[5] * mne[5] + 3.0 * s[6] * mne[6] - 9.0 * s[8] * mne[8];
    f1[7] = 4.0 * s[0] * mne[0] + 2.0 * s[1] * mne[1] + 1.0 * s[2] * mne[2] - 6.0 * s[3] * mne[3] - 3.0 * s[4] * mne[4] - 6.0 * s[5] * mne[5] - 3.0 * s[6] * mne[6] + 9.0 * s[8] * mne[8];
    f1[8] = 4.0 * s[0] * mne[0] + 2.0 * s[1] * mne[1] + 1.0 * s[2] * mne[2] + 6.0 * s[3] * mne[3] + 3.0 * s[4] * mne[4] - 6.0 * s[5] * mne[5] - 3.0 * s[6] * mne[6] - 9.0 * s[8] * mne[8];
    c11 = f1[1] + f1[5] + f1[6] + f1[3] + f1[7] + f1[8];
    c22 = f1[5] + f1[2] + f1[6] + f1[7] + f1[4] + f1[8];
    c12 = f1[5] - f1[6] + f1[7] - f1[8];
    } else {
        const double base = 24.0 * e0 + 6.0 * e1;
        c11 = base + 18.0 * e7;
        c22 = base - 18.0 * e7;
        c12 = 36.0 * e8;
    }
    double zarib = -1.5 * d;
    double d11 = zarib * c11
               + 0.75 * ((s[7] - s[1]) * (gx * ux + gy * uy) - s[7] * (2.0 * gx * ux));
    double d22 = zarib * c22
               + 0.75 * ((s[7] - s[1]) * (gx * ux + gy * uy) - s[7] * (2.0 * gy * uy));
    double d12 = zarib * c12
               + 0.75 * (-(s[7] * (gx * uy + gy * ux)));
    rheology64(P, rho, om, d11, d22, d12, o);
}

// Stress, yield flag and the node's new rate from the strain rate d11, d22,
// d12 (before the 1/rho of F6:921): Newtonian stress 2 nu d (F6:918), or the
// Papanastasiou-Bingham viscosity (F6:921-957) or, rheology 2, the
// Papanastasiou-Casson lines F6 carries commented out (F6:941, F6:944; the
// model of F3:986-989, whose zero-shear limit is written without the square).
__device__ __forceinline__ void rheology64(const DevP &P, double rho, double om, double d11,
                                           double d22, double d12, CollideOut &o) {
    o.stress[0] = P.nu * 2.0 * d11;
    o.stress[1] = P.nu * 2.0 * d22;
    o.stress[2] = P.nu * 2.0 * d12;
    o.omega_bn = om;
    o.yielded = 0;
    if (P.rheology != 0) {
        double inv_rho = 1.0 / rho;
        d11 = d11 * inv_rho; d22 = d22 * inv_rho; d12 = d12 * inv_rho;
        double gam = sqrt(2.0 * fabs(d11 * d11 + d22 * d22 + 2.0 * (d12 * d12)));
        double nu_a;
        if (P.rheology == 2) {
            if (gam < P.gamma_min) {
                nu_a = sqrt(P.mu0) + sqrt(P.m_pap * P.tau_y);
            } else {
                double c = sqrt(P.mu0) + sqrt(P.tau_y / gam) * (1.0 - exp(-sqrt(P.m_pap * gam)));
                nu_a = c * c;
            }
        } else if (gam < P.gamma_min)
            nu_a = P.mu0 + P.m_pap * P.tau_y;
        else
            nu_a = P.mu0 + (P.tau_y / gam) * (1.0 - exp(-(P.m_pap * gam)));
        double t11 = nu_a * 2.0 * d11, t22 = nu_a * 2.0 * d22, t12 = nu_a * 2.0 * d12;
        o.stress[0] = t11; o.stress[1] = t22; o.stress[2] = t12;
        double tmag = sqrt(0.5 * fabs(t11 * t11 + t22 * t22 + 2.0 * (t12 * t12)));
        o.yielded = (tmag <= P.tau_y) ? 0 : 1;
        nu_a = nu_a * inv_rho;
        o.omega_bn = 1.0 / (3.0 * nu_a + 0.5);
    }
}

// The same collision for the fused FP64 step, written as a change of the
// populations so that only f and the non-equilibrium moments are live (the
// moment vectors of the reference form keep four times as many values in
// registers).  With n_k = m_k - meq_k and the relaxation change
// D_k = m*_k - m_k = -s_k n_k + (1 - s_k/2) mf_k (F6:889-890), the
// reference's back transform f* = d A m* (A = 36 M^-1, the integer rows of
// F6:892-900, d = fl32(1/36)) equals (36 d) f + d A D: the float-literal
// factor 36 d of the reference is kept, the rounding differs in the last
// bits.  The strain rate comes from the column sums of collide64<false>.
__device__ __forceinline__ void collide64_fused(const DevP &P, double f[9], double rho,
                                                double ux, double uy, double gx, double gy,
                                                double om, CollideOut &o) {
    double s[9];
    rates64(P, om, s);
    // moments (F6:878-886) as non-equilibrium parts (F6:964-989)
    const double u2 = ux * ux + uy * uy;
    const double n0 = (f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8]) - rho;
    const double n1 = (-(4.0 * f[0]) - f[1] - f[2] - f[3] - f[4] + 2.0 * (f[5] + f[6] + f[7] + f[8]))
                    - rho * (-2.0 + 3.0 * u2);
    const double n2 = (4.0 * f[0] - 2.0 * (f[1] + f[2] + f[3] + f[4]) + (f[5] + f[6] + f[7] + f[8]))
                    - rho * (1.0 - 3.0 * u2);
    const double j3 = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    const double j5 = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    const double n3 = j3 - rho * ux, n5 = j5 - rho * uy;
    const double n4 = (-(2.0 * (f[1] - f[3])) + f[5] - f[6] - f[7] + f[8]) + rho * ux;
    const double n6 = (-(2.0 * (f[2] - f[4])) + f[5] + f[6] - f[7] - f[8]) + rho * uy;
    const double n7 = (f[1] - f[2] + f[3] - f[4]) - rho * (ux * ux - uy * uy);
    const double n8 = (f[5] - f[6] + f[7] - f[8]) - rho * (ux * uy);
    // forcing moments mf (F6:964-989)
    const double fx = rho * (gx + 0.0), fy = rho * (gy + 0.0);
    const double fu = 6.0 * (fx * ux + fy * uy);
    // D_k = -s_k n_k + (1 - s_k/2) mf_k
    const double D0 = -s[0] * n0;
    const double D1 = -s[1] * n1 + (1.0 - 0.5 * s[1]) * fu;
    const double D2 = -s[2] * n2 - (1.0 - 0.5 * s[2]) * fu;
    const double D3 = -s[3] * n3 + (1.0 - 0.5 * s[3]) * fx;
    const double D4 = -s[4] * n4 - (1.0 - 0.5 * s[4]) * fx;
    const double D5 = -s[5] * n5 + (1.0 - 0.5 * s[5]) * fy;
    const double D6 = -s[6] * n6 - (1.0 - 0.5 * s[6]) * fy;
    const double D7 = -s[7] * n7 + (1.0 - 0.5 * s[7]) * (2.0 * (fx * ux - fy * uy));
    const double D8 = -s[8] * n8 + (1.0 - 0.5 * s[8]) * (fy * ux + fx * uy);
    // f* = (36 d) f + d A D
    const double d = D_MRT, c = 36.0 * D_MRT;
    const double a0 = 4.0 * D0, b1 = a0 - D1 - 2.0 * D2, b5 = a0 + 2.0 * D1 + D2;
    const double x3 = 6.0 * D3 - 6.0 * D4, x5 = 6.0 * D5 - 6.0 * D6, x7 = 9.0 * D7;
    const double y3 = 6.0 * D3 + 3.0 * D4, y5 = 6.0 * D5 + 3.0 * D6, y8 = 9.0 * D8;
    f[0] = c * f[0] + d * (4.0 * (D0 - D1 + D2));
    f[1] = c * f[1] + d * (b1 + x3 + x7);
    f[2] = c * f[2] + d * (b1 + x5 - x7);
    f[3] = c * f[3] + d * (b1 - x3 + x7);
    f[4] = c * f[4] + d * (b1 - x5 - x7);
    f[5] = c * f[5] + d * (b5 + y3 + y5 + y8);
    f[6] = c * f[6] + d * (b5 - y3 + y5 - y8);
    f[7] = c * f[7] + d * (b5 - y3 - y5 + y8);
    f[8] = c * f[8] + d * (b5 + y3 - y5 - y8);
    // Chai strain rate (F6:903-916) from the column sums
    const double e0 = s[0] * n0, e1 = s[1] * n1, e7 = s[7] * n7, e8 = s[8] * n8;
    const double base = 24.0 * e0 + 6.0 * e1;
    const double zarib = -1.5 * d;
    double d11 = zarib * (base + 18.0 * e7)
               + 0.75 * ((s[7] - s[1]) * (gx * ux + gy * uy) - s[7] * (2.0 * gx * ux));
    double d22 = zarib * (base - 18.0 * e7)
               + 0.75 * ((s[7] - s[1]) * (gx * ux + gy * uy) - s[7] * (2.0 * gy * uy));
    double d12 = zarib * (36.0 * e8) + 0.75 * (-(s[7] * (gx * uy + gy * ux)));
    rheology64(P, rho, om, d11, d22, d12, o);
}

// ------------------------------------------------ FP32 deviation pipeline
// Populations are handled as H_k = S h_k = S (f_k - w_k rho0) in the storage
// units of the format (scale St::S).

// Raw buffer loads of the nine pulled populations of an interior node (every
// source in the domain): one vector offset for the node, a uniform offset per
// population.
template <class St>
__device__ __forceinline__ void load_interior_dev(const DevP &P, const typename St::T *src,
                                                  long long node, float H[9]) {
    const auto rs = make_rsrc(src, P.buf_bytes);
    const unsigned vo = (unsigned)node * (unsigned)sizeof(typename St::T);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const unsigned so = (unsigned)((k * P.plane - cxk(k) - (long long)cyk(k) * P.W) *
                                       (long long)sizeof(typename St::T));
        H[k] = St::bld(rs, vo, so);
    }
}

// Pull-stream of the nine populations of node (i, j) of any lane
// (F6:1164-1188; the fused path's edge wavefronts and the particle helpers):
// raw buffer loads with one 32-bit offset per population (the fused path's
// storage buffers are below 4 GiB), so that a wavefront whose lanes mix
// interior and boundary nodes issues one load per population and no 64-bit
// address arithmetic.  A source outside the domain reads the node itself:
// population opp(k) where the bounce-back rule replaces k (bb_replaced), k
// itself otherwise.
template <class St, class V>
__device__ __forceinline__ void pull_lane(const DevP &P, const void *src, unsigned node, int i,
                                          int j, V v[9]) {
    const auto rs = make_rsrc(src, P.buf_bytes);
    const bool el = i == 0, er = i == P.nx, eb = j == 0, et = j == P.ny;
    const unsigned sz = (unsigned)sizeof(typename St::T);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int si = i - cxk(k), sj = j - cyk(k);
        const bool in = (unsigned)si <= (unsigned)P.nx && (unsigned)sj <= (unsigned)P.ny;
        const unsigned s = in ? node - (unsigned)(cxk(k) + cyk(k) * P.W) : node;
        // plane offsets of k and opp(k): uniform, no per-lane multiply
        const unsigned pk = (unsigned)k * (unsigned)P.plane;
        const unsigned po = (unsigned)opp_k(k) * (unsigned)P.plane;
        const unsigned e = (!in && bb_replaced(k, el, er, eb, et)) ? po : pk;
        v[k] = St::bld(rs, (s + e) * sz, 0u);
    }
}

// f_k of node (i, j) in storage units for any lane, reduced formats:
// pull_lane and the boundary rules (F6:1206-1294, or the Zou-He /
// equilibrium edges) on deviations in storage units, with the edge arrays
// read by buffer loads issued beside the populations.  Opposite directions
// share a weight, so the bounce-back rule keeps its form; every edge rule is
// linear in (rho0, h), so it is applied in storage units with S rho0 and S
// times the wall terms.
template <class St, int BND = -1>
__device__ __forceinline__ void load_fk_lane_dev(const DevP &P, const void *src,
                                                 const float *erho, unsigned node, int i,
                                                 int j, float H[9]) {
    const int bnd = BND >= 0 ? BND : P.boundary;
    const float S = St::S;
    const bool edge = i == 0 || j == 0 || i == P.nx || j == P.ny;
    const auto er = make_rsrc(erho, (unsigned)(3 * P.nbnd * 4));
    const unsigned b4 = edge ? (unsigned)bnd_index(P, i, j) * 4u : 0u;
    float e0, e1 = 0.0f, e2 = 0.0f;
    pull_lane<St>(P, src, node, i, j, H);
    e0 = StF32::bld(er, b4, 0u);
    if (bnd == 2) {
        e1 = StF32::bld(er, b4, (unsigned)(P.nbnd * 4));
        e2 = StF32::bld(er, b4, (unsigned)(2 * P.nbnd * 4));
    }
    if (!edge) return;
    if (bnd == 2) {  // boundary node: the equilibrium of its interior source
        feq_dev(S * P.frho0, e1, e2, S * e0, H);
        return;
    }
    const bool corner = (i == 0 || i == P.nx) && (j == 0 || j == P.ny);
    if (bnd == 1 && !corner) {
        zou_he_edge<float>(P, i, j, S * P.frho0, H);
        return;
    }
    const float a = (float)INV6 * S, b = (float)(4.0 * INV6) * S;
    bounce_back_node<float>(P, i, j, a * e0, b * e0, H);
}

// The same for FP64 storage (the values and edge rules of load_fk64).
template <int BND = -1>
__device__ __forceinline__ void load_fk_lane64(const DevP &P, const void *src,
                                               const double *erho, unsigned node, int i, int j,
                                               double v[9]) {
    const int bnd = BND >= 0 ? BND : P.boundary;
    const bool edge = i == 0 || j == 0 || i == P.nx || j == P.ny;
    const auto er = make_rsrc(erho, (unsigned)(3 * P.nbnd * 8));
    const unsigned b8 = edge ? (unsigned)bnd_index(P, i, j) * 8u : 0u;
    double e0, e1 = 0.0, e2 = 0.0;
    pull_lane<StF64>(P, src, node, i, j, v);
    e0 = StF64::bld(er, b8, 0u);
    if (bnd == 2) {
        e1 = StF64::bld(er, b8, (unsigned)(P.nbnd * 8));
        e2 = StF64::bld(er, b8, (unsigned)(2 * P.nbnd * 8));
    }
    if (!edge) return;
    if (bnd == 2) {
        feq64(e1, e2, e0, v);
        return;
    }
    const bool corner = (i == 0 || i == P.nx) && (j == 0 || j == P.ny);
    if (bnd == 1 && !corner) {
        zou_he_edge<double>(P, i, j, 0.0, v);
        return;
    }
    const double clb2 = 1.0;
    bounce_back_node<double>(P, i, j, INV6 * clb2 * e0, 4.0 * INV6 * clb2 * e0, v);
}

// Moments of the deviations in storage units (F6:878-886): density deviation
// rd and momentum jx, jy (MACROSCOPICVARIABLES, F6:1405-1408; the sums of
// c_k w_k vanish), and the non-conserved moments the collide needs (m3 = jx
// and m5 = jy).  Computed once for the macroscopic fields and the collide.
struct DevMom {
    float rd, jx, jy, m1, m2, m4, m6, m7, m8;
};
__device__ __forceinline__ DevMom dev_moments(const float H[9]) {
    DevMom m;
    const float s13 = H[1] + H[3], s24 = H[2] + H[4];
    const float s14 = s13 + s24;
    const float c56 = H[5] + H[6], c78 = H[7] + H[8], s58 = c56 + c78;
    const float a13 = H[1] - H[3], a24 = H[2] - H[4];
    const float b56 = H[5] - H[6], b78 = H[7] - H[8];
    const float d1 = b56 - b78, d2 = c56 - c78;
    m.rd = (H[0] + s14) + s58;
    m.jx = a13 + d1;
    m.jy = a24 + d2;
    m.m1 = (-4.0f * H[0] - s14) + 2.0f * s58;
    m.m2 = (4.0f * H[0] - 2.0f * s14) + s58;
    m.m4 = -2.0f * a13 + d1;
    m.m6 = -2.0f * a24 + d2;
    m.m7 = s13 - s24;
    m.m8 = b56 + b78;
    return m;
}

// Hardware reciprocal, square root and base-2 exponential (about 1 ulp):
// the FP32 path trades IEEE-exact division for issue rate.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fexp(float x) {
    return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
}

// Collision_MRT in deviation form, FP32, storage units (same physics as
// collide64, F6:867-958).  rho: density, rinvS = 1 / (S rho); ucx, ucy: the
// corrected velocity (F6:1118-1124); gx, gy: the SPM forcing G (FORCED
// only; zero otherwise); om: the node's rate.  OUT: also the stress tensor
// and the yield flag (F6:945-954), output only.
//
// Written as a change of the populations: with n_k = m_k - meq_k the
// post-collision moments differ from m_k by D_k = -s_k n_k + (1 - s_k/2) mf_k
// (D_0 = 0: n_0 = mf_0 = 0), and since the rows of the moment matrix M are
// orthogonal, M^-1 = M^T diag(1/N) with N = (9, 36, 36, 6, 12, 6, 12, 4, 4),
// so h* = h + M^T (D / N).  The uniform factors s_k / N_k and
// (1 - s_k/2) / N_k come precomputed in DevP (cD*).  The momentum S rho u_c
// of the corrected velocity is jx + FX/2 with FX = S rho gx, so n3 = -FX/2
// and n5 = -FY/2 (zero without forcing) are formed without cancellation.
template <bool FORCED, bool OUT, class CP = DevP>
__device__ __forceinline__ void collide_dev(const CP &P, float H[9], const DevMom &mo,
                                            float S, float rho, float rinvS, float ucx,
                                            float ucy, float gx, float gy, float om,
                                            float &om_out, float stress[3], int &yielded) {
#pragma clang fp contract(on)  // fused within one expression: the same in every kernel
    float FX = 0.0f, FY = 0.0f, rux = mo.jx, ruy = mo.jy;
    if constexpr (FORCED) {
        const float rS = S * rho;
        FX = rS * gx;
        FY = rS * gy;
        rux = mo.jx + 0.5f * FX;
        ruy = mo.jy + 0.5f * FY;
    }
    // non-equilibrium parts (equilibrium relative to the rest state)
    const float ru2 = rux * ucx + ruy * ucy;
    const float n1 = (mo.m1 + 2.0f * mo.rd) - 3.0f * ru2;
    const float n2 = (mo.m2 - mo.rd) + 3.0f * ru2;
    const float n4 = mo.m4 + rux, n6 = mo.m6 + ruy;
    const float n7 = mo.m7 - (rux * ucx - ruy * ucy), n8 = mo.m8 - rux * ucy;
    // D_k / N_k; the q and om factors vary per node
    float qa, qb = 0.0f;
    if (P.q_mode) {
        const float q = (8.0f * (2.0f - om)) * frcp(8.0f - om);
        qa = q * (-1.0f / 12.0f);
        if constexpr (FORCED) qb = q * (1.0f / 24.0f) - (1.0f / 12.0f);
    } else {
        qa = P.cDq1;
        if constexpr (FORCED) qb = P.cDq2;
    }
    const float oa = om * -0.25f;
    float A = n1 * P.cD1a, B = n2 * P.cD2a, C = 0.0f, E = 0.0f;
    float Q = n4 * qa, G = n6 * qa, R7 = n7 * oa, R8 = n8 * oa;
    if constexpr (FORCED) {
        const float fu = FX * ucx + FY * ucy;
        const float mf7 = FX * ucx - FY * ucy, mf8 = FY * ucx + FX * ucy;
        A = A + fu * P.cD1b;
        B = B + fu * P.cD2b;
        C = FX * P.cC3;
        E = FY * P.cC5;
        Q = Q + FX * qb;
        G = G + FY * qb;
        R7 = R7 + mf7 * (0.5f - 0.25f * om);
        R8 = R8 + mf8 * (0.25f - 0.125f * om);
    }
    // h* = h + M^T (D / N)
    const float ab1 = -A - 2.0f * B, ab5 = 2.0f * A + B;
    const float x = C - 2.0f * Q, y = E - 2.0f * G, u = C + Q, v = E + G;
    const float t13 = ab1 + R7, t24 = ab1 - R7, t57 = ab5 + R8, t68 = ab5 - R8;
    const float uv = u + v, vu = v - u;
    H[0] = H[0] + 4.0f * (B - A);
    H[1] = H[1] + (t13 + x);
    H[3] = H[3] + (t13 - x);
    H[2] = H[2] + (t24 + y);
    H[4] = H[4] + (t24 - y);
    H[5] = H[5] + (t57 + uv);
    H[7] = H[7] + (t57 - uv);
    H[6] = H[6] + (t68 + vu);
    H[8] = H[8] + (t68 - vu);
    om_out = om;
    yielded = 0;
    if (!OUT && P.rheology == 0) return;
    // Chai strain rate (F6:903-924).  The strain moments f1 = C (s * n)
    // enter d11, d22, d12 only through three column sums of C, which leave
    //   d11 = -(e0 + e1/4 + 3 e7/4), d22 = -(e0 + e1/4 - 3 e7/4), d12 = -3 e8/2
    // (e_k = s_k n_k; e0 = 0) plus the forcing terms (which the reference
    // adds before its 1/rho), here in storage units.
    const float base = n1 * P.cS1;  // -e1 / 4
    const float e7 = om * n7;
    float D11 = base - 0.75f * e7, D22 = base + 0.75f * e7, D12 = -1.5f * (om * n8);
    if constexpr (FORCED) {
        const float gxS = S * gx, gyS = S * gy;
        const float gs = (0.75f * om + P.cS2) * (gxS * ucx + gyS * ucy);  // 0.75 (s7 - s1) gu
        D11 = D11 + (gs - 1.5f * (om * (gxS * ucx)));
        D22 = D22 + (gs - 1.5f * (om * (gyS * ucy)));
        D12 = D12 - 0.75f * (om * (gxS * ucy + gyS * ucx));
    }
    if (OUT && P.rheology == 0) {
        const float c = 2.0f * P.fnu * (1.0f / S);
        stress[0] = c * D11; stress[1] = c * D22; stress[2] = c * D12;
    }
    if (P.rheology != 0) {
        // gamma^2 = 2 |d:d| / rho^2 with d = D / S; one reciprocal square root
        // gives both gamma and 1/gamma
        const float g2 = (2.0f * (D11 * D11 + D22 * D22 + 2.0f * (D12 * D12))) * (rinvS * rinvS);
        const float rg = frsq(g2), gam = g2 * rg;
        const bool small = !(g2 >= P.fgamma_min2);
        const float mu0 = P.fmu0, ty = P.ftau_y, mp = P.fm_pap;
        float nu_a;
        if (P.rheology == 2) {
            // Papanastasiou-Casson (F6:941, F6:944)
            const float c = P.fsqrt_mu0 + fsqrt(ty * rg) * (1.0f - fexp(-fsqrt(mp * gam)));
            nu_a = small ? P.fcas_lim : c * c;
        } else {
            nu_a = small ? mu0 + mp * ty : mu0 + (ty * rg) * (1.0f - fexp(-(mp * gam)));
        }
        const float rinv = S * rinvS;
        if (OUT) {
            // tau = 2 nu_a d / rho, |tau| = sqrt(tau:tau / 2) = nu_a gamma
            const float c = 2.0f * nu_a * rinvS;
            stress[0] = c * D11; stress[1] = c * D22; stress[2] = c * D12;
            yielded = (small ? 0.0f : nu_a * gam) <= ty ? 0 : 1;
        }
        om_out = frcp(3.0f * (nu_a * rinv) + 0.5f);
    }
}

// --------------------------------------------------------- block helpers
// Fixed-order tree reduction of NV doubles over the block (deterministic):
// at stride s = 128, 64, ..., 1 element t < s adds element t + s.  With
// WAVE_TAIL the two strides that cross wavefronts go through LDS and the
// last six run inside wavefront 0 with shuffles (lane t reads lane t + s):
// the same additions in the same order with four block barriers instead of
// nine.  Which form a kernel uses is chosen per instantiation by
// measurement: the shuffles changed the register allocation of the whole
// fused kernel, for the better in FP64 (-4 % per step) and for the worse in
// the reduced formats (+3 %).
template <int NV, bool WAVE_TAIL = false>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds /* NV*NT */) {
    static_assert(NT == 256, "the tree below is written for 256 threads");
    const int t = threadIdx.x;
#pragma unroll
    for (int a = 0; a < NV; ++a) lds[a * NT + t] = v[a];
    __syncthreads();
    if constexpr (WAVE_TAIL) {
        if (t < 128) {
#pragma unroll
            for (int a = 0; a < NV; ++a) lds[a * NT + t] += lds[a * NT + t + 128];
        }
        __syncthreads();
        if (t < 64) {
#pragma unroll
            for (int a = 0; a < NV; ++a) {
                double x = lds[a * NT + t] + lds[a * NT + t + 64];
#pragma unroll
                for (int s = 32; s > 0; s >>= 1) x += __shfl_down(x, s, 64);
                if (t == 0) lds[a * NT] = x;
            }
        }
        __syncthreads();
    } else {
        for (int s = NT / 2; s > 0; s >>= 1) {
            if (t < s) {
#pragma unroll
                for (int a = 0; a < NV; ++a) lds[a * NT + t] += lds[a * NT + t + s];
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < NV; ++a) v[a] = lds[a * NT];
    __syncthreads();
}

// The reference's (..)**8 and (..)**14 (F4:595-596, F4:616-617) with the
// roundings of its compiled form: **8 as seven multiplications left to
// right, **14 by the binary method ((x^2 x^4) x^8).
__device__ __forceinline__ double pow8_lj(double x) {
    return ((((((x * x) * x) * x) * x) * x) * x) * x;
}
__device__ __forceinline__ double pow14_lj(double x) {
    double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4;
    return (x2 * x4) * x8;
}

// Pair force of the pair (ip, jp), ip < jp, for the contact model P.contact.
// Returns false (and a zero force) when the pair is out of range
// d > r_i + r_j + zeta.
//   0 Glowinski (F3:746-770): the stale wall distance dpw chooses the branch
//     (F3:760).
//   1 Wan-Turek (F4:632-645, F1:654-667): eps_p (x_i - x_j) (r_i + r_j - d)
//     in contact, eps_p (x_i - x_j) (r_i + r_j + zeta - d)^2 in the gap.
//   2 Lennard-Jones (F4:590-599, F1:612-621): 2.4 R_P^2 (2 q^14 - q^8)
//     (x_i - x_j) / (4 R_P^2), q = 2 R_P / d, R_P the radius of particle 1
//     (F4:588).
__device__ __forceinline__ bool pair_force(const DevP &P, double dpw, const double *X,
                                           const double *Y, const double *D, int ip, int jp,
                                           double &f0, double &f1) {
    const double zeta = P.zeta_c, cij = P.cij;
    double xi = X[ip], yi = Y[ip];
    double xj = X[jp], yj = Y[jp];
    double di = D[ip], dj = D[jp];
    double dpp = sqrt((xi - xj) * (xi - xj) + (yi - yj) * (yi - yj));
    double rsum = 0.5 * di + 0.5 * dj;
    f0 = 0.0; f1 = 0.0;
    if (dpp > rsum + zeta) return false;
    if (P.contact == 1) {
        if (dpp < rsum) {
            f0 = P.eps_p * (xi - xj) * (rsum - dpp);
            f1 = P.eps_p * (yi - yj) * (rsum - dpp);
        } else {
            double t = rsum + zeta - dpp;
            f0 = P.eps_p * (xi - xj) * (t * t);
            f1 = P.eps_p * (yi - yj) * (t * t);
        }
        return true;
    }
    if (P.contact == 2) {
        double rp = 0.5 * D[0];
        double q = 2.0 * rp / dpp;
        double a = LJ_A * (rp * rp) * (2.0 * pow14_lj(q) - pow8_lj(q));
        f0 = a * (xi - xj) / (4.0 * rp * rp);
        f1 = a * (yi - yj) / (4.0 * rp * rp);
        return true;
    }
    double t = (dpp - 0.5 * di - 0.5 * dj - zeta) / zeta;
    if (dpw > rsum) {
        f0 = P.eps_p * cij * (t * t) * (xi - xj) / dpp;
        f1 = P.eps_p * cij * (t * t) * (yi - yj) / dpp;
    } else {
        f0 = (P.eps_p * cij * (t * t) + P.e_p * cij * (rsum - dpp) / zeta) * (xi - xj) / dpp;
        f1 = (P.eps_p * cij * (t * t) + P.e_p * cij * (rsum - dpp) / zeta) * (yi - yj) / dpp;
    }
    return true;
}

// Force of wall iw (1 bottom, 2 right, 3 top, 4 left) on a disk at (x, y),
// diameter dm, from the wall image R_W; returns the distance to it.  rp1:
// radius of particle 1 (Lennard-Jones).
//   0 Glowinski (F3:773-804): images half a diameter outside the domain.
//   1 Wan-Turek (F4:648-676, F1:670-698): mirror images about the wall edges
//     P.wall_edge (WALL%EDGE, F4:182-185); eps_w (c - R_W) (dm - d) inside,
//     eps_w (c - R_W) (dm + zeta - d)^2 within zeta.
//   2 Lennard-Jones (F4:600-621, F1:622-643): images (x, 0), (nx, y),
//     (x, 2 ny - y), (-x, y); 2.4 R_P^2 (2 q^14 - q^8) (c - R_W) / R_P^2,
//     q = R_P / d.
__device__ __forceinline__ double wall_one(const DevP &P, double x, double y, double dm,
                                           double rp1, int iw, double &f0, double &f1) {
    const double zeta = P.zeta_c, cij = P.cij;
    double rw0, rw1;
    if (P.contact == 1) {
        const double *E = P.wall_edge;
        if (iw == 1) { rw0 = x; rw1 = 2.0 * E[0] - y; }
        else if (iw == 2) { rw0 = 2.0 * E[1] - x; rw1 = y; }
        else if (iw == 3) { rw0 = x; rw1 = 2.0 * E[2] - y; }
        else { rw0 = 2.0 * E[3] - x; rw1 = y; }
    } else if (P.contact == 2) {
        if (iw == 1) { rw0 = x; rw1 = 0.0; }
        else if (iw == 2) { rw0 = (double)P.nx; rw1 = y; }
        else if (iw == 3) { rw0 = x; rw1 = 2.0 * (double)P.ny - y; }
        else { rw0 = -x; rw1 = y; }
    } else {
        if (iw == 1) { rw0 = x; rw1 = -(0.5 * dm); }
        else if (iw == 2) { rw0 = (double)P.nx + 0.5 * dm; rw1 = y; }
        else if (iw == 3) { rw0 = x; rw1 = (double)P.ny + 0.5 * dm; }
        else { rw0 = -(0.5 * dm); rw1 = y; }
    }
    const double d = sqrt((x - rw0) * (x - rw0) + (y - rw1) * (y - rw1));
    f0 = 0.0; f1 = 0.0;
    if (d > dm + zeta) return d;
    if (P.contact == 1) {
        if (d < dm) {
            f0 = P.eps_w * (x - rw0) * (dm - d);
            f1 = P.eps_w * (y - rw1) * (dm - d);
        } else {
            double t = dm + zeta - d;
            f0 = P.eps_w * (x - rw0) * (t * t);
            f1 = P.eps_w * (y - rw1) * (t * t);
        }
    } else if (P.contact == 2) {
        double q = rp1 / d;
        double a = LJ_A * (rp1 * rp1) * (2.0 * pow14_lj(q) - pow8_lj(q));
        f0 = a * (x - rw0) / (rp1 * rp1);
        f1 = a * (y - rw1) / (rp1 * rp1);
    } else if (d > dm) {
        double t = (d - dm - zeta) / zeta;
        f0 = P.eps_w * cij * (t * t) * (x - rw0) / d;
        f1 = P.eps_w * cij * (t * t) * (y - rw1) / d;
    } else {
        double t = (d - dm - zeta) / zeta;
        f0 = (P.eps_w * cij * (t * t) + P.e_w * cij * (dm - d) / zeta) * (x - rw0) / d;
        f1 = (P.eps_w * cij * (t * t) + P.e_w * cij * (dm - d) / zeta) * (y - rw1) / d;
    }
    return d;
}

// Wall forces on a disk at (x, y), diameter dm, added in the wall order of
// F3:773-804 (bottom, right, top, left); returns the last wall distance,
// which becomes the next call's stale D_PW.
__device__ __forceinline__ double wall_forces(const DevP &P, double x, double y, double dm,
                                              double rp1, double &fcx, double &fcy) {
    double d = 0.0;
    for (int iw = 1; iw <= 4; ++iw) {
        double f0, f1;
        d = wall_one(P, x, y, dm, rp1, iw, f0, f1);
        fcx = fcx + f0; fcy = fcy + f1;
    }
    return d;
}

// Contact force on particle p (model P.contact; N-body pair order of
// F3:746-770, walls of F3:773-804), evaluated by one thread.  The pair branch
// reads dpw, the D_PW left by the previous call (F3:760).  Writes in-range
// partners j>p into nb[0..MAXNB) and returns their count; *dpw_last receives
// this particle's last wall distance.  X, Y, D: centres and diameters of all
// particles (global memory or an LDS copy).
__device__ __forceinline__ int contact_one(const DevP &P, const double *X, const double *Y,
                                           const double *D, int np, int p, double dpw,
                                           double &fcx, double &fcy, int *nb, double *dpw_last) {
    fcx = 0.0; fcy = 0.0;
    int cnt = 0;
    // accumulation order seen by particle p in the reference loop: pairs
    // (j, p) for j < p subtract, then pairs (p, j) for j > p add
    for (int q = 0; q < np; ++q) {
        if (q == p) continue;
        double f0, f1;
        bool hit = (q < p) ? pair_force(P, dpw, X, Y, D, q, p, f0, f1)
                           : pair_force(P, dpw, X, Y, D, p, q, f0, f1);
        if (!hit) continue;
        if (q < p) { fcx = fcx - f0; fcy = fcy - f1; }
        else {
            fcx = fcx + f0; fcy = fcy + f1;
            if (cnt < MAXNB) nb[cnt] = q;
            ++cnt;
        }
    }
    *dpw_last = wall_forces(P, X[p], Y[p], D[p], 0.5 * D[0], fcx, fcy);
    return cnt;
}

// Block-wide form of contact_one (all NT threads of the block call it; the
// results are valid in thread 0).  Each thread tests one partner per round of
// NT; the in-range partners are compacted into LDS in ascending index order
// and thread 0 adds them in that order, which is the order contact_one (and
// the reference loop) uses: out-of-range partners contribute +0.0, which
// leaves the sums unchanged.  More than CCAP in-range partners fall back to
// the serial loop.
constexpr int CCAP = 64;
struct ContactLDS {
    double f0[CCAP], f1[CCAP];
    int q[CCAP];
    int wcount[NT / 64];
    int total;
};

__device__ __forceinline__ int contact_block(const DevP &P, const double *X, const double *Y,
                                             const double *D, int np, int p, double dpw,
                                             ContactLDS &L, double &fcx, double &fcy, int *nb,
                                             double *dpw_last) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) L.total = 0;
    __syncthreads();
    for (int base = 0; base < np; base += NT) {
        const int q = base + t;
        double f0 = 0.0, f1 = 0.0;
        bool hit = false;
        if (q < np && q != p)
            hit = (q < p) ? pair_force(P, dpw, X, Y, D, q, p, f0, f1)
                          : pair_force(P, dpw, X, Y, D, p, q, f0, f1);
        const unsigned long long m = __ballot(hit);
        if (lane == 0) L.wcount[w] = __popcll(m);
        __syncthreads();
        int off = L.total;
        for (int v = 0; v < w; ++v) off += L.wcount[v];
        if (hit) {
            int pos = off + __popcll(m & ((1ULL << lane) - 1ULL));
            if (pos < CCAP) { L.q[pos] = q; L.f0[pos] = f0; L.f1[pos] = f1; }
        }
        __syncthreads();
        if (t == 0) {
            int s = 0;
            for (int v = 0; v < NT / 64; ++v) s += L.wcount[v];
            L.total += s;
        }
        __syncthreads();
    }
    int cnt = 0;
    if (t == 0) {
        if (L.total > CCAP) return contact_one(P, X, Y, D, np, p, dpw, fcx, fcy, nb, dpw_last);
        fcx = 0.0; fcy = 0.0;
        for (int k = 0; k < L.total; ++k) {
            if (L.q[k] < p) { fcx = fcx - L.f0[k]; fcy = fcy - L.f1[k]; }
            else {
                fcx = fcx + L.f0[k]; fcy = fcy + L.f1[k];
                if (cnt < MAXNB) nb[cnt] = L.q[k];
                ++cnt;
            }
        }
        *dpw_last = wall_forces(P, X[p], Y[p], D[p], 0.5 * D[0], fcx, fcy);
    }
    return cnt;
}

// contact_block for at most 64 particles, by one wavefront (threads 0..63,
// no block barriers): lane q tests partner q, the in-range partners are
// visited in ascending index order through the ballot mask, and the four
// walls are evaluated by four lanes; every sum keeps the order of
// contact_one.  Results are valid in lane 0 (they are the same in every lane).
__device__ __forceinline__ int contact_wave(const DevP &P, const double *X, const double *Y,
                                            const double *D, int np, int p, double dpw,
                                            double &fcx, double &fcy, int *nb, double *dpw_last) {
    const int lane = threadIdx.x & 63;
    double f0 = 0.0, f1 = 0.0;
    bool hit = false;
    if (lane < np && lane != p)
        hit = (lane < p) ? pair_force(P, dpw, X, Y, D, lane, p, f0, f1)
                         : pair_force(P, dpw, X, Y, D, p, lane, f0, f1);
    unsigned long long m = __ballot(hit);
    double w0 = 0.0, w1 = 0.0, wd = 0.0;
    if (lane < 4) wd = wall_one(P, X[p], Y[p], D[p], 0.5 * D[0], lane + 1, w0, w1);
    fcx = 0.0; fcy = 0.0;
    int cnt = 0;
    while (m) {
        const int k = __ffsll((long long)m) - 1;
        m &= m - 1;
        const double g0 = __shfl(f0, k), g1 = __shfl(f1, k);
        if (k < p) { fcx = fcx - g0; fcy = fcy - g1; }
        else {
            fcx = fcx + g0; fcy = fcy + g1;
            if (lane == 0 && nb && cnt < MAXNB) nb[cnt] = k;
            ++cnt;
        }
    }
    for (int w = 0; w < 4; ++w) {
        fcx = fcx + __shfl(w0, w);
        fcy = fcy + __shfl(w1, w);
    }
    const double dl = __shfl(wd, 3);
    if (lane == 0) *dpw_last = dl;
    return cnt;
}

// update_one by one wavefront: lane c < 3 updates component c (x, y,
// rotation) with the operations of update_one in the same order, so the
// result is bitwise the same.  rec: the record of state k (read by every
// lane); F0..F2, fcx, fcy: FORCE/TORQUE and the contact force, the same in
// every lane.  Lane t < PCOUNT returns field t of the new record.
__device__ __forceinline__ double update_wave(const DevP &P, const double *rec, double F0,
                                              double F1, double F2, double fcx, double fcy,
                                              int kk) {
    const int lane = threadIdx.x & 63;
    const int c = lane < 3 ? lane : 0;
    const double rho0 = P.rho0, rhop = rec[PRHOP];
    const double mass = c < 2 ? rec[PMP] : rec[PIP];
    const double v = rec[PVX + c], vo = rec[PVOX + c];
    const double fk = c == 0 ? F0 : (c == 1 ? F1 : F2);
    const double fc = c == 0 ? fcx : fcy;
    double n, fh;
    if (P.motion == 1) {
        // prescribed oscillation (F4:221-226), added mass (F4:256-258)
        n = c == 0 ? -P.osc_u * cos(2.0 * PI_F * (double)kk / P.osc_period) : 0.0;
        fh = (rho0 / rhop) * mass * (n - v);
    } else {
        fh = (rho0 / rhop) * mass * (v - vo);
        n = v + fk / mass + fh / mass;
        if (c == 1) n = n + (1.0 - rho0 / rhop) * P.gy;
        if (c < 2) n = n + fc / mass;
    }
    const double pos = rec[PX + c] + 0.5 * (n + v);
    const int t = lane;
    auto from = [&](double x, int base) { return __shfl(x, min(max(t - base, 0), 2)); };
    const double g_pos = from(pos, PX), g_n = from(n, PVX), g_vo = from(v, PVOX),
                 g_fh = from(fh, PFHX);
    double val = t < PCOUNT ? rec[t] : 0.0;
    if (t == PX || t == PY || (t == PANG && P.motion != 1 && P.update_angle)) val = g_pos;
    if (t >= PVX && t <= PW) val = g_n;
    if (t >= PVOX && t <= PVOW) val = g_vo;
    if (t >= PFHX && t <= PFHW) val = g_fh;
    if (t == PFX) val = F0;
    if (t == PFY) val = F1;
    if (t == PTQ) val = F2;
    if (t == PFCX) val = fcx;
    if (t == PFCY) val = fcy;
    return val;
}

// Update of one particle record in place for loop iteration kk: Newton-Euler
// (F6:302-312) or the prescribed motion (F4:221-226).
__device__ __forceinline__ void update_one(const DevP &P, double *rec /* PCOUNT, stride 1 */,
                                           int kk) {
    const double rho0 = P.rho0;
    if (P.motion == 1) {
        // prescribed oscillation of loop iteration kk (F4:221-226) and the
        // added-mass term from the velocity change (F4:256-258)
        double rhop = rec[PRHOP], mp = rec[PMP], ipm = rec[PIP];
        double v0 = rec[PVX], v1 = rec[PVY], v2 = rec[PW];
        rec[PVOX] = v0; rec[PVOY] = v1; rec[PVOW] = v2;
        double n0 = -P.osc_u * cos(2.0 * PI_F * (double)kk / P.osc_period);
        double n1 = 0.0, n2 = 0.0;
        rec[PVX] = n0; rec[PVY] = n1; rec[PW] = n2;
        rec[PX] = rec[PX] + 0.5 * (n0 + v0);
        rec[PY] = rec[PY] + 0.5 * (n1 + v1);
        rec[PFHX] = (rho0 / rhop) * mp * (n0 - v0);
        rec[PFHY] = (rho0 / rhop) * mp * (n1 - v1);
        rec[PFHW] = (rho0 / rhop) * ipm * (n2 - v2);
        return;
    }
    double rhop = rec[PRHOP], mp = rec[PMP], ipm = rec[PIP];
    double v0 = rec[PVX], v1 = rec[PVY], v2 = rec[PW];
    double fh0 = (rho0 / rhop) * mp * (v0 - rec[PVOX]);
    double fh1 = (rho0 / rhop) * mp * (v1 - rec[PVOY]);
    double fh2 = (rho0 / rhop) * ipm * (v2 - rec[PVOW]);
    rec[PFHX] = fh0; rec[PFHY] = fh1; rec[PFHW] = fh2;
    rec[PVOX] = v0; rec[PVOY] = v1; rec[PVOW] = v2;
    double n0 = v0 + rec[PFX] / mp + fh0 / mp + rec[PFCX] / mp;
    double n1 = v1 + rec[PFY] / mp + fh1 / mp + (1.0 - rho0 / rhop) * P.gy + rec[PFCY] / mp;
    double n2 = v2 + rec[PTQ] / ipm + fh2 / ipm;
    rec[PVX] = n0; rec[PVY] = n1; rec[PW] = n2;
    rec[PX] = rec[PX] + 0.5 * (n0 + v0);
    rec[PY] = rec[PY] + 0.5 * (n1 + v1);
    if (P.update_angle) rec[PANG] = rec[PANG] + 0.5 * (n2 + v2);
}

}  // namespace lbm
