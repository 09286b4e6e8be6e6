// C ABI of the MI355X D2Q9 MRT / smoothed-profile time step (include/lbm_amd.h).
// Host side: parameter checks, device allocation, layout transposes and the
// launch sequence of the fused step.  Kernels are in lbm_kernels.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm_amd.h"
#include "lbm_kernels.h"
#include "lbm_adjoint.h"

using namespace lbm;

static_assert(LBM_PF_COUNT == PCOUNT, "particle record layout");
static_assert(LBM_H_PCOUNT == HPC, "history record layout");

struct lbm_handle {
    lbm_params prm;
    DevP P;
    hipStream_t stream = nullptr;
    std::string err;
    // FP64 state (SoA planes)
    double *f = nullptr, *fp = nullptr, *rho = nullptr, *ux = nullptr, *uy = nullptr;
    double *g = nullptr, *obn = nullptr, *stress = nullptr;
    int *image = nullptr;
    // fused chain: storage populations (every format, FP64 included) and
    // node rates (FP32 or FP64 as the storage), by chain parity
    void *sA = nullptr, *sB = nullptr;
    void *rate[2] = {nullptr, nullptr};
    // STRESS and DOMAIN_IMAGE of the fused collisions, by iteration parity
    double *astress[2] = {nullptr, nullptr};
    int *aimage[2] = {nullptr, nullptr};
    // contact pairs found by the fused launches, by chain parity
    int *cpairs[2] = {nullptr, nullptr}, *cpaircnt[2] = {nullptr, nullptr};
    int last_pairs = 0;  // which of them holds the last contact of the loop-end state
    // reduced formats: the deviations of the canonical F in storage units as
    // the FINAL launch computed them (FP32; w_k rho0 + h in FP64 rounds the
    // smallest h), valid until the canonical F is written otherwise; a new
    // chain's prologue starts from them, so a restart is exact
    float *hexact = nullptr;
    bool hexact_ok = false;
    void *erho[2] = {nullptr, nullptr};
    void *wtab = nullptr;  // wall velocities per edge class: 10 doubles then 10 floats
    unsigned long long *mpart[2] = {nullptr, nullptr};  // monitor partials by launch parity
    // particles
    double *pbuf[2] = {nullptr, nullptr};
    int pcur = 0;
    double *dpw[2] = {nullptr, nullptr};
    int dcur = 0;
    int *pairs = nullptr, *paircnt = nullptr;
    double *phi = nullptr;
    int *porig = nullptr;
    int PS = 0;
    // scratch
    double *staging = nullptr;
    double *partial = nullptr;
    double *scalar = nullptr;
    double *hist_part = nullptr;
    unsigned long long *hist_slots = nullptr;
    size_t hist_rows = 0;
    // history buffers the next recording fused call writes: the two above,
    // or regions of the device's static scratch array
    double *hist_part_use = nullptr;
    unsigned long long *hist_slots_use = nullptr;
    int dev = 0;  // device the handle's memory lives on
    // timing events: whole sequence (ev0..ev1), back-to-back main launches (evA..evB)
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evA = nullptr, evB = nullptr;
    // per-wavefront candidate lists of many-particle runs
    unsigned long long *tl_hdr[2] = {nullptr, nullptr};
    int *tl_ids[2] = {nullptr, nullptr};
    int *err_flag = nullptr;
    double *lag = nullptr;  // module-global rate of the lagged collide
    // fused path: reach boxes and FP32 profile records of the particle
    // states (same parity as pbuf), tagged force partials, slots per particle
    int *pbox[2] = {nullptr, nullptr};
    float *spf[2] = {nullptr, nullptr};
    unsigned long long *fslot = nullptr;  // chunk force partials [maxp][kpart][SLOTW]
    size_t fslot_bytes = 0;
    int kpart = 1;                        // helper blocks per particle
    // Word tags of the force partials never repeat on a handle (wtag_next);
    // tile-list generations run along a chain (lg_next: beyond every one
    // used so far); sr_next is the stochastic-rounding position of the
    // collision of iteration res_kk + 1, restarted by lbm_initialize and
    // carried by snapshots.
    unsigned wtag_next = 1;
    unsigned lg_next = 1;
    unsigned sr_next = 1;
    // Resident continuation of a fused run: the chain's last launch (index
    // res_c, list generation res_lg, particle and wall-distance bases) ran
    // loop iteration res_kk and collided iteration res_kk + 1; its inputs are
    // intact, so the next fused call continues the chain, and the canonical
    // FP64 arrays are materialised from them when something reads them
    // (canon: they hold iteration res_kk).
    bool res = false;
    bool canon = true;
    // the fused path addresses each 9-plane population buffer with 32-bit
    // buffer offsets: below 4 GiB (FP64: about 59.6 million nodes)
    bool fused_ok = true;
    bool res_yield = false;  // the last launch stored yielded-count partials
    int res_kk = 0, res_c = 0, res_pbase = 0, res_dbase = 0;
    unsigned res_lg = 0;
};

namespace {

int fail(lbm_handle *h, int code, const std::string &msg) {
    if (h) h->err = msg;
    return code;
}

#define HIPCHECK(h, call)                                                               \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail((h), LBM_ERR_HIP,                                               \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while (0)

inline unsigned nblk(long long n) { return (unsigned)((n + NT - 1) / NT); }

size_t storage_bytes(int storage) {
    switch (storage) {
        case LBM_FP64: return 8;
        case LBM_FP32: return 4;
        default: return 2;
    }
}

// message of the last failed lbm_create on this thread (no handle exists)
thread_local std::string t_create_err;

int bad_create(const char *msg) {
    t_create_err = msg;
    return LBM_ERR_PARAM;
}

int check_launch(lbm_handle *h, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, LBM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return LBM_OK;
}

// The canonical FP64 arrays of a handle whose fused run is resident are
// materialised (a FINAL launch) before anything reads them; a call that
// changes them also ends the resident continuation, so the next fused call
// starts a new chain from them (defined with the fused path below).
int ensure_canonical(lbm_handle *h);
inline void drop_resident(lbm_handle *h) {
    h->res = false;
    h->canon = true;
}
#define LBM_READS_STATE(h)                     \
    do {                                       \
        const int rc_ = ensure_canonical(h);   \
        if (rc_) return rc_;                   \
    } while (0)
#define LBM_WRITES_STATE(h)                    \
    do {                                       \
        const int rc_ = ensure_canonical(h);   \
        if (rc_) return rc_;                   \
        drop_resident(h);                      \
        (h)->hexact_ok = false;                \
    } while (0)

// ------------------------------------------------------------ static scratch
// Device memory of lbm_step_hist_static: the monitor slots, the particle
// history rows and the formatted history rows of one recording call, in a
// file-scope array that each GPU holds until the process ends, so that the
// call allocates nothing.  In host code the array's name is not a device
// address: the address of the current device's copy comes from
// hipGetSymbolAddress, looked up once per device (scratch_setup, from
// lbm_create with the handle's device current) and kept per device.
constexpr size_t SCRATCH_WORDS = (size_t)1 << 19;  // 4 MiB of 8-byte words
constexpr int SCRATCH_MAX_DEV = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[SCRATCH_MAX_DEV] = {};
hipError_t g_scratch_err[SCRATCH_MAX_DEV] = {};
bool g_scratch_done[SCRATCH_MAX_DEV] = {};

}  // namespace

static __device__ __align__(256) unsigned long long scratch[SCRATCH_WORDS];

namespace {

void scratch_setup(int dev) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEV) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[dev]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    g_scratch_err[dev] = e;
    g_scratch_ptr[dev] = e == hipSuccess ? ptr : nullptr;
    g_scratch_done[dev] = true;
}

// This device's scratch address (null, with the reason in *why, when the
// lookup failed or never ran).
void *scratch_address(int dev, std::string *why) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEV) {
        *why = "device index " + std::to_string(dev) + " beyond the scratch table";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[dev]) {
        *why = "scratch address never looked up";
        return nullptr;
    }
    if (!g_scratch_ptr[dev])
        *why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(g_scratch_err[dev]);
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

const char *lbm_last_error(const lbm_handle *h) {
    if (h) return h->err.c_str();
    return t_create_err.empty() ? "null handle" : t_create_err.c_str();
}

int lbm_create(const lbm_params *p, lbm_handle **out) {
    t_create_err.clear();
    if (!p || !out) return bad_create("null params or output pointer");
    *out = nullptr;
    if (p->nx < 2 || p->ny < 2) return bad_create("nx and ny must be at least 2");
    // F6:115-118: tau = 1/omega must not be below 0.5
    if (!(p->omega > 0.0) || 1.0 / p->omega < 0.5)
        return bad_create("relaxation time 1/omega below 0.5 (F6:115-118)");
    if (p->rheology != LBM_NEWTONIAN && p->rheology != LBM_BINGHAM && p->rheology != LBM_CASSON)
        return bad_create("unknown rheology");
    if (p->boundary != LBM_BOUNCE_BACK && p->boundary != LBM_ZOU_HE &&
        p->boundary != LBM_EQUILIBRIUM)
        return bad_create("unknown boundary mode");
    if (p->contact != LBM_CONTACT_GLOWINSKI && p->contact != LBM_CONTACT_WAN_TUREK &&
        p->contact != LBM_CONTACT_LJ)
        return bad_create("unknown contact model");
    if (p->storage < LBM_FP64 || p->storage > LBM_BF16) return bad_create("unknown storage");
    if (p->max_particles < 0 || p->max_particles > 65535)
        return bad_create("max_particles outside 0..65535");
    if (p->motion != LBM_MOTION_FREE && p->motion != LBM_MOTION_OSC_X)
        return bad_create("unknown particle motion");
    if (p->motion == LBM_MOTION_OSC_X && !(p->osc_period > 0.0))
        return bad_create("oscillation period must be positive");
    lbm_handle *h = new lbm_handle();
    h->prm = *p;
    DevP &P = h->P;
    std::memset(&P, 0, sizeof(P));
    P.nx = p->nx; P.ny = p->ny; P.W = p->nx + 1; P.H = p->ny + 1;
    P.nn = (long long)P.W * P.H;
    P.plane = (P.nn + 255) / 256 * 256;
    P.rheology = p->rheology; P.q_mode = p->q_mode; P.update_angle = p->update_angle;
    P.motion = p->motion; P.osc_u = p->osc_u; P.osc_period = p->osc_period;
    P.boundary = p->boundary;
    P.np = 0;
    P.maxp = std::max(1, p->max_particles);
    P.nbnd = 2 * P.W + 2 * (P.H - 2);
    P.omega = p->omega; P.s0 = p->s0; P.s1 = p->s1; P.s2 = p->s2; P.s3 = p->s3;
    P.s5 = p->s5; P.sq = p->sq; P.nu = p->nu; P.mu0 = p->mu0; P.tau_y = p->tau_y;
    P.m_pap = p->m_pap; P.gamma_min = p->gamma_min; P.rho0 = p->rho0; P.gy = p->gy;
    for (int k = 0; k < 8; ++k) P.wall[k] = p->wall[k];
    P.corner[0] = p->corner[0]; P.corner[1] = p->corner[1];
    P.zeta_c = p->zeta_c; P.cij = p->cij; P.eps_p = p->eps_p; P.eps_w = p->eps_w;
    P.e_p = p->e_p; P.e_w = p->e_w;
    P.contact = p->contact;
    for (int k = 0; k < 4; ++k) P.wall_edge[k] = p->wall_edge[k];
    P.fomega = (float)P.omega; P.fs0 = (float)P.s0; P.fs1 = (float)P.s1; P.fs2 = (float)P.s2;
    P.fs3 = (float)P.s3; P.fs5 = (float)P.s5; P.fsq = (float)P.sq; P.fnu = (float)P.nu;
    P.fmu0 = (float)P.mu0; P.ftau_y = (float)P.tau_y; P.fm_pap = (float)P.m_pap;
    P.fgamma_min = (float)P.gamma_min; P.frho0 = (float)P.rho0;
    P.fsqrt_mu0 = (float)std::sqrt(P.mu0);
    P.fcas_lim = (float)(std::sqrt(P.mu0) + std::sqrt(P.m_pap * P.tau_y));
    // FP32 collide factors (D_k / N_k = n_k cDka + mf_k cDkb): mf_1 = 6 fu,
    // mf_2 = -6 fu, mf_3 = fx, mf_4 = -fx, mf_5 = fy, mf_6 = -fy
    P.cD1a = (float)(-P.s1 / 36.0); P.cD1b = (float)((1.0 - 0.5 * P.s1) * 6.0 / 36.0);
    P.cD2a = (float)(-P.s2 / 36.0); P.cD2b = (float)(-(1.0 - 0.5 * P.s2) * 6.0 / 36.0);
    P.cD3a = (float)(-P.s3 / 6.0); P.cD3b = (float)((1.0 - 0.5 * P.s3) / 6.0);
    P.cD5a = (float)(-P.s5 / 6.0); P.cD5b = (float)((1.0 - 0.5 * P.s5) / 6.0);
    P.cDq1 = (float)(-P.sq / 12.0); P.cDq2 = (float)(-(1.0 - 0.5 * P.sq) / 12.0);
    P.cS1 = (float)(-0.25 * P.s1); P.cS2 = (float)(-0.75 * P.s1);
    P.cC3 = (float)(0.5 * P.s3 / 6.0 + (1.0 - 0.5 * P.s3) / 6.0);
    P.cC5 = (float)(0.5 * P.s5 / 6.0 + (1.0 - 0.5 * P.s5) / 6.0);
    P.fgamma_min2 = (float)(P.gamma_min * P.gamma_min);
    {
        unsigned long long bb = 9ULL * (unsigned long long)P.plane *
                                (p->storage == LBM_FP64 ? 8ULL : p->storage == LBM_FP32 ? 4ULL : 2ULL);
        if (p->storage != LBM_FP64 && bb >= (1ULL << 32)) {
            delete h;
            return bad_create("reduced-format population buffer of 4 GiB or more");
        }
        P.buf_bytes = (unsigned)std::min<unsigned long long>(bb, 0xFFFFFFFFULL);
        h->fused_ok = bb < (1ULL << 32);
    }
    // magic numbers for row = node / W (Granlund-Montgomery round-up method)
    if (P.nn >= (1LL << 31)) { delete h; return bad_create("2^31 or more nodes"); }
    {
        int l = 0;
        while ((1ULL << l) < (unsigned long long)P.W) ++l;
        P.w_shift = l;
        P.w_magic = (unsigned)(((1ULL << 32) * ((1ULL << l) - (unsigned long long)P.W)) /
                               (unsigned long long)P.W + 1ULL);
        auto div = [&](unsigned n) {
            unsigned t = (unsigned)(((unsigned long long)P.w_magic * n) >> 32);
            return (t + ((n - t) >> 1)) >> (l - 1);
        };
        const unsigned probe[] = {0u, 1u, (unsigned)P.W - 1, (unsigned)P.W, (unsigned)P.W + 1,
                                  (unsigned)(P.nn - 1), (unsigned)(P.nn / 2), 2147483647u};
        for (unsigned n : probe)
            if (div(n) != n / (unsigned)P.W) { delete h; return bad_create("row division check failed"); }
    }

    auto al = [&](void **ptr, size_t bytes) -> int {
        hipError_t e = hipMalloc(ptr, bytes);
        if (e != hipSuccess) return fail(h, LBM_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        e = hipMemset(*ptr, 0, bytes);
        if (e != hipSuccess) return fail(h, LBM_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
        return LBM_OK;
    };
    size_t pl = (size_t)P.plane;
    int rc = LBM_OK;
    hipError_t es = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (es != hipSuccess) { *out = h; return fail(h, LBM_ERR_HIP, "hipStreamCreate failed"); }
    if (hipGetDevice(&h->dev) != hipSuccess) { *out = h; return fail(h, LBM_ERR_HIP, "hipGetDevice failed"); }
    scratch_setup(h->dev);  // this device's static scratch address, once per device
    if ((rc = al((void **)&h->f, 9 * pl * 8)) || (rc = al((void **)&h->fp, 9 * pl * 8)) ||
        (rc = al((void **)&h->rho, pl * 8)) || (rc = al((void **)&h->ux, pl * 8)) ||
        (rc = al((void **)&h->uy, pl * 8)) || (rc = al((void **)&h->g, 2 * pl * 8)) ||
        (rc = al((void **)&h->obn, pl * 8)) || (rc = al((void **)&h->stress, 3 * pl * 8)) ||
        (rc = al((void **)&h->image, pl * 4)) || (rc = al((void **)&h->staging, 9 * pl * 8)) ||
        (rc = al((void **)&h->erho[0], (size_t)P.nbnd * 3 * 8)) ||
        (rc = al((void **)&h->erho[1], (size_t)P.nbnd * 3 * 8)) ||
        (rc = al((void **)&h->pbuf[0], (size_t)PCOUNT * P.maxp * 8)) ||
        (rc = al((void **)&h->pbuf[1], (size_t)PCOUNT * P.maxp * 8)) ||
        (rc = al((void **)&h->dpw[0], 8)) || (rc = al((void **)&h->dpw[1], 8)) ||
        (rc = al((void **)&h->pairs, (size_t)P.maxp * MAXNB * 4)) ||
        (rc = al((void **)&h->paircnt, (size_t)P.maxp * 4)) ||
        (rc = al((void **)&h->porig, (size_t)P.maxp * 2 * 4)) ||
        (rc = al((void **)&h->partial, (size_t)nblk(P.nn) * 8)) ||
        (rc = al((void **)&h->scalar, 8)) ||
        (rc = al((void **)&h->err_flag, 4)) || (rc = al((void **)&h->lag, 8)) ||
        (rc = al((void **)&h->pbox[0], (size_t)P.maxp * 4 * 4)) ||
        (rc = al((void **)&h->pbox[1], (size_t)P.maxp * 4 * 4)) ||
        (rc = al((void **)&h->spf[0], (size_t)P.maxp * SPF * 4)) ||
        (rc = al((void **)&h->spf[1], (size_t)P.maxp * SPF * 4)) ||
        (rc = al(&h->wtab, 128)) ||
        (rc = al((void **)&h->mpart[0], (size_t)nblk(P.nn) * 2 * 8)) ||
        (rc = al((void **)&h->mpart[1], (size_t)nblk(P.nn) * 2 * 8))) {
        *out = h;
        return rc;
    }
    {
        // (u_x, u_y) per edge class WT_CORNER, WT_TOP, WT_BOTTOM, WT_RIGHT,
        // WT_LEFT; the FP32 copies are the same values rounded once
        const double *w = p->wall;
        const double wd[10] = {p->corner[0], p->corner[1], w[4], w[5], w[0], w[1],
                               w[2], w[3], w[6], w[7]};
        float wf[10];
        for (int k = 0; k < 10; ++k) wf[k] = (float)wd[k];
        char blob[128] = {0};
        std::memcpy(blob, wd, sizeof wd);
        std::memcpy(blob + 80, wf, sizeof wf);
        if (hipMemcpy(h->wtab, blob, sizeof blob, hipMemcpyHostToDevice) != hipSuccess) {
            *out = h;
            return fail(h, LBM_ERR_HIP, "hipMemcpy of the wall table failed");
        }
        h->P.wtab = (const double *)h->wtab;
        h->P.wtabf = (const float *)((const char *)h->wtab + 80);
    }
    if (p->max_particles > NP_LOOP) {
        // one list per wavefront of every tile block
        size_t nt = (size_t)nblk(P.nn) * (NT / LIST_UNIT);
        for (int b = 0; b < 2; ++b)
            if ((rc = al((void **)&h->tl_hdr[b], nt * 8)) ||
                (rc = al((void **)&h->tl_ids[b], nt * MAXC * 4))) {
                *out = h;
                return rc;
            }
    }
    if (h->fused_ok) {
        const size_t sb = storage_bytes(p->storage);
        const size_t rb = p->storage == LBM_FP64 ? 8 : 4;
        if ((rc = al(&h->sA, 9 * pl * sb)) || (rc = al(&h->sB, 9 * pl * sb)) ||
            (rc = al(&h->rate[0], pl * rb)) || (rc = al(&h->rate[1], pl * rb)) ||
            (rc = al((void **)&h->astress[0], 3 * pl * 8)) ||
            (rc = al((void **)&h->astress[1], 3 * pl * 8)) ||
            (rc = al((void **)&h->aimage[0], pl * 4)) || (rc = al((void **)&h->aimage[1], pl * 4)) ||
            (rc = al((void **)&h->cpairs[0], (size_t)P.maxp * MAXNB * 4)) ||
            (rc = al((void **)&h->cpairs[1], (size_t)P.maxp * MAXNB * 4)) ||
            (rc = al((void **)&h->cpaircnt[0], (size_t)P.maxp * 4)) ||
            (rc = al((void **)&h->cpaircnt[1], (size_t)P.maxp * 4)) ||
            (p->storage != LBM_FP64 && (rc = al((void **)&h->hexact, 9 * pl * 4)))) {
            *out = h;
            return rc;
        }
    }
    double inf = INFINITY;
    (void)hipMemcpy(h->dpw[0], &inf, 8, hipMemcpyHostToDevice);
    (void)hipMemcpy(h->dpw[1], &inf, 8, hipMemcpyHostToDevice);
    (void)hipEventCreate(&h->ev0);
    (void)hipEventCreate(&h->ev1);
    (void)hipEventCreate(&h->evA);
    (void)hipEventCreate(&h->evB);
    *out = h;
    return LBM_OK;
}

int lbm_destroy(lbm_handle *h) {
    if (!h) return LBM_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void *bufs[] = {h->f, h->fp, h->rho, h->ux, h->uy, h->g, h->obn, h->stress, h->image,
                    h->sA, h->sB, h->rate[0], h->rate[1], h->astress[0], h->astress[1],
                    h->aimage[0], h->aimage[1], h->cpairs[0], h->cpairs[1], h->cpaircnt[0],
                    h->cpaircnt[1], h->hexact, h->erho[0], h->erho[1], h->pbuf[0], h->pbuf[1],
                    h->dpw[0], h->dpw[1], h->pairs, h->paircnt, h->phi, h->porig,
                    h->staging, h->partial, h->scalar, h->hist_part, h->hist_slots,
                    h->tl_hdr[0], h->tl_hdr[1], h->tl_ids[0], h->tl_ids[1], h->err_flag,
                    h->pbox[0], h->pbox[1], h->spf[0], h->spf[1], h->fslot, h->lag, h->wtab,
                    h->mpart[0], h->mpart[1]};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->evA) (void)hipEventDestroy(h->evA);
    if (h->evB) (void)hipEventDestroy(h->evB);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return LBM_OK;
}

int lbm_sync(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_initialize(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    const DevP &P = h->P;
    hipLaunchKernelGGL(k_init, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P, h->f, h->rho,
                       h->ux, h->uy, h->g, h->obn);
    int rc = check_launch(h, "k_init");
    if (rc) return rc;
    // run state back to its start: the stochastic-rounding sequence restarts
    // (tags and list generations go on: they only have to be unique), no
    // fused run is resident, and the stale wall distance of the contact
    // force is unset
    drop_resident(h);
    h->hexact_ok = false;
    h->sr_next = 1;
    for (int b = 0; b < 2; ++b)
        if (h->tl_hdr[b])
            HIPCHECK(h, hipMemsetAsync(h->tl_hdr[b], 0,
                                       (size_t)nblk(P.nn) * (NT / LIST_UNIT) * 8, h->stream));
    double inf = INFINITY;
    HIPCHECK(h, hipMemcpyAsync(h->dpw[0], &inf, 8, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, hipMemcpyAsync(h->dpw[1], &inf, 8, hipMemcpyHostToDevice, h->stream));
    // the module-global rate starts from OMEGA (F6:199)
    HIPCHECK(h, hipMemcpyAsync(h->lag, &P.omega, 8, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    h->dcur = 0;
    return LBM_OK;
}

static int upload_planes(lbm_handle *h, const double *host, int nc, double *dev) {
    const DevP &P = h->P;
    HIPCHECK(h, hipMemcpyAsync(h->staging, host, (size_t)P.nn * nc * 8, hipMemcpyHostToDevice,
                               h->stream));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P.nn, nc,
                       h->staging, dev, P.plane);
    int rc = check_launch(h, "k_aos_to_soa");
    if (rc) return rc;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

static int download_planes(lbm_handle *h, const double *dev, int nc, double *host) {
    const DevP &P = h->P;
    hipLaunchKernelGGL(k_soa_to_aos, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P.nn, nc, dev,
                       h->staging, P.plane);
    int rc = check_launch(h, "k_soa_to_aos");
    if (rc) return rc;
    HIPCHECK(h, hipMemcpyAsync(host, h->staging, (size_t)P.nn * nc * 8, hipMemcpyDeviceToHost,
                               h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_set_f(lbm_handle *h, const double *f) {
    if (!h || !f) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    return upload_planes(h, f, 9, h->f);
}
int lbm_get_f(lbm_handle *h, double *f) {
    if (!h || !f) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    return download_planes(h, h->f, 9, f);
}
int lbm_set_fpost(lbm_handle *h, const double *f) {
    if (!h || !f) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    return upload_planes(h, f, 9, h->fp);
}
int lbm_get_fpost(lbm_handle *h, double *f) {
    if (!h || !f) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    return download_planes(h, h->fp, 9, f);
}

int lbm_set_fields(lbm_handle *h, const double *rho, const double *ux, const double *uy,
                   const double *g, const double *omega_bn) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    int rc;
    if (rho && (rc = upload_planes(h, rho, 1, h->rho))) return rc;
    if (ux && (rc = upload_planes(h, ux, 1, h->ux))) return rc;
    if (uy && (rc = upload_planes(h, uy, 1, h->uy))) return rc;
    if (g && (rc = upload_planes(h, g, 2, h->g))) return rc;
    if (omega_bn && (rc = upload_planes(h, omega_bn, 1, h->obn))) return rc;
    return LBM_OK;
}

int lbm_get_fields(lbm_handle *h, double *rho, double *ux, double *uy, double *g,
                   double *omega_bn) {
    if (!h) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    int rc;
    if (rho && (rc = download_planes(h, h->rho, 1, rho))) return rc;
    if (ux && (rc = download_planes(h, h->ux, 1, ux))) return rc;
    if (uy && (rc = download_planes(h, h->uy, 1, uy))) return rc;
    if (g && (rc = download_planes(h, h->g, 2, g))) return rc;
    if (omega_bn && (rc = download_planes(h, h->obn, 1, omega_bn))) return rc;
    return LBM_OK;
}

int lbm_get_stress(lbm_handle *h, double *stress) {
    if (!h || !stress) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    return download_planes(h, h->stress, 3, stress);
}

int lbm_get_image(lbm_handle *h, int32_t *image) {
    if (!h || !image) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    HIPCHECK(h, hipMemcpyAsync(image, h->image, (size_t)h->P.nn * 4, hipMemcpyDeviceToHost,
                               h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_set_particles(lbm_handle *h, int32_t n, const double *part) {
    if (!h || n < 0 || (n > 0 && !part)) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    if (n > h->prm.max_particles)
        return fail(h, LBM_ERR_CAPACITY, "more particles than max_particles");
    DevP &P = h->P;
    std::vector<double> buf((size_t)PCOUNT * P.maxp, 0.0);
    double reach = 0.0;
    for (int f = 0; f < PCOUNT; ++f)
        for (int p = 0; p < n; ++p) buf[(size_t)f * P.maxp + p] = part[(size_t)f * n + p];
    for (int p = 0; p < n; ++p) {
        double r = 0.5 * part[(size_t)LBM_PF_DIAM * n + p] + 0.5 * part[(size_t)LBM_PF_ZETA * n + p] + 1.0;
        if (!(r > 0.0) || !std::isfinite(r)) return fail(h, LBM_ERR_PARAM, "bad particle diameter");
        reach = std::max(reach, r);
    }
    P.np = n;
    HIPCHECK(h, hipMemcpyAsync(h->pbuf[h->pcur], buf.data(), buf.size() * 8,
                               hipMemcpyHostToDevice, h->stream));
    int PS = (int)std::ceil(2.0 * reach) + 2;
    // helper blocks per particle: NT-node chunks of the largest reach box
    // (at most `side` nodes across, clipped to the grid)
    {
        const long long side = 2LL * (long long)std::ceil(reach) + 1;
        const long long nodes = std::min<long long>(side, P.W) * std::min<long long>(side, P.H);
        h->kpart = n > 0 ? (int)((nodes + NT - 1) / NT) : 1;
        const size_t bytes = (size_t)P.maxp * h->kpart * SLOTW * 8;
        if (n > 0 && bytes > h->fslot_bytes) {
            if (h->fslot) (void)hipFree(h->fslot);
            h->fslot = nullptr;
            h->fslot_bytes = 0;
            HIPCHECK(h, hipMalloc((void **)&h->fslot, bytes));
            HIPCHECK(h, hipMemsetAsync(h->fslot, 0, bytes, h->stream));
            h->fslot_bytes = bytes;
        }
    }
    if (n > 0 && PS > h->PS) {
        if (h->phi) (void)hipFree(h->phi);
        h->phi = nullptr;
        HIPCHECK(h, hipMalloc((void **)&h->phi, (size_t)P.maxp * PS * PS * 8));
        HIPCHECK(h, hipMemsetAsync(h->phi, 0, (size_t)P.maxp * PS * PS * 8, h->stream));
        h->PS = PS;
    }
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_get_particles(lbm_handle *h, double *part) {
    if (!h || !part) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    const DevP &P = h->P;
    std::vector<double> buf((size_t)PCOUNT * P.maxp);
    HIPCHECK(h, hipMemcpyAsync(buf.data(), h->pbuf[h->pcur], buf.size() * 8,
                               hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    for (int f = 0; f < PCOUNT; ++f)
        for (int p = 0; p < P.np; ++p) part[(size_t)f * P.np + p] = buf[(size_t)f * P.maxp + p];
    return LBM_OK;
}

int lbm_get_phi(lbm_handle *h, int32_t ip, double *phi) {
    if (!h || !phi || ip < 0 || ip >= h->P.np || !h->phi) return LBM_ERR_PARAM;
    const DevP &P = h->P;
    int PS = h->PS;
    std::vector<double> patch((size_t)PS * PS);
    int orig[2];
    HIPCHECK(h, hipMemcpyAsync(patch.data(), h->phi + (size_t)ip * PS * PS, patch.size() * 8,
                               hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipMemcpyAsync(orig, h->porig + 2 * ip, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    std::fill(phi, phi + P.nn, 0.0);
    for (int b = 0; b < PS; ++b)
        for (int a = 0; a < PS; ++a) {
            int i = orig[0] + a, j = orig[1] + b;
            if (i < 0 || i > P.nx || j < 0 || j > P.ny) continue;
            phi[(size_t)j * P.W + i] = patch[(size_t)b * PS + a];
        }
    return LBM_OK;
}

}  // extern "C"

// ------------------------------------------------------ checkpoint / restart
// Text state files of the reference (WRITE_OUTPUT F6:377-405, READ_INPUT
// F6:174-198): one record per node, I outer and J inner, list-directed
// numbers; the particle file holds CENTER(2), VELOCITY(3), FORCE(3) per disk
// and 0_OmegaMRT_output.txt the relaxation vector OMEGA_MRT(0:8).  Numbers are written
// with 17 significant digits (exact round trip) and read as a stream of
// tokens, as Fortran list-directed input does (D exponents accepted).
namespace {

const char *const RS_F = "0_F_output.txt", *const RS_UX = "0_UX_output.txt",
                  *const RS_UY = "0_UY_output.txt", *const RS_RHO = "0_RHO_output.txt",
                  *const RS_G = "0_Gspm_output.txt", *const RS_OBN = "0_OmegaBN_output.txt",
                  *const RS_PART = "0_Particle_output.txt", *const RS_OM = "0_OmegaMRT_output.txt",
                  *const RS_WALL = "0_Wall_output.txt", *const RS_IMG = "0_DomainImage_output.txt";

std::string join_path(const char *dir, const char *name) {
    std::string d(dir);
    if (!d.empty() && d.back() != '/') d += '/';
    return d + name;
}

// Node-major reference layout (component fastest, i fastest among nodes)
// -> one text record per node in I-outer / J-inner order.
bool write_nodes(const std::string &path, const DevP &P, int nc, const double *a) {
    FILE *fp = std::fopen(path.c_str(), "w");
    if (!fp) return false;
    for (int i = 0; i <= P.nx; ++i)
        for (int j = 0; j <= P.ny; ++j) {
            const double *v = a + ((size_t)j * P.W + i) * nc;
            for (int c = 0; c < nc; ++c) std::fprintf(fp, " %.17g", v[c]);
            std::fputc('\n', fp);
        }
    return std::fclose(fp) == 0;
}

// Reads exactly n numbers; false if the file is missing, short or malformed.
bool read_numbers(const std::string &path, size_t n, std::vector<double> &out,
                  std::string &why) {
    FILE *fp = std::fopen(path.c_str(), "r");
    if (!fp) { why = "cannot open " + path; return false; }
    out.clear();
    out.reserve(n);
    char tok[128];
    while (out.size() < n && std::fscanf(fp, "%127s", tok) == 1) {
        for (char *c = tok; *c; ++c)
            if (*c == 'D' || *c == 'd') *c = 'E';
        char *end = nullptr;
        double v = std::strtod(tok, &end);
        if (end == tok || *end != '\0') {
            std::fclose(fp);
            why = path + ": not a number: " + tok;
            return false;
        }
        out.push_back(v);
    }
    std::fclose(fp);
    if (out.size() < n) {
        why = path + ": " + std::to_string(out.size()) + " numbers, expected " + std::to_string(n);
        return false;
    }
    return true;
}

// I-outer / J-inner records -> node-major reference layout.
void nodes_from_records(const DevP &P, int nc, const std::vector<double> &rec, double *a) {
    size_t r = 0;
    for (int i = 0; i <= P.nx; ++i)
        for (int j = 0; j <= P.ny; ++j)
            for (int c = 0; c < nc; ++c) a[((size_t)j * P.W + i) * nc + c] = rec[r++];
}

// fields of the particle file, in the order of F6:393/F6:192
const int RS_PFIELDS[8] = {LBM_PF_X, LBM_PF_Y, LBM_PF_VX, LBM_PF_VY, LBM_PF_W,
                           LBM_PF_FX, LBM_PF_FY, LBM_PF_TQ};

struct SnapHeader {
    char magic[8];
    uint32_t version, gen_base;
    int32_t nx, ny, maxp, np, storage, rheology;
    uint64_t bytes;
};
const char SNAP_MAGIC[8] = {'L', 'B', 'M', 'S', 'N', 'A', 'P', '1'};

// F, F_POST (9 each), RHO, UX, UY, G (2), OMEGA_BN, STRESS (3): 27 doubles
// per node at these plane offsets, then DOMAIN_IMAGE as int32, the particle
// records, D_PW and the module-global rate
constexpr int SN_F = 0, SN_FP = 9, SN_RHO = 18, SN_UX = 19, SN_UY = 20, SN_G = 21, SN_OBN = 23,
              SN_STRESS = 24, SN_PLANES = 27;

// Reduced formats also hold the exact deviations of F (9 FP32 planes) after
// DOMAIN_IMAGE, so a loaded snapshot continues bit for bit.
uint64_t snapshot_bytes(long long nn, int np, int storage) {
    return sizeof(SnapHeader) + (uint64_t)nn * SN_PLANES * 8 + (uint64_t)nn * 4 +
           (storage != LBM_FP64 ? (uint64_t)nn * 9 * 4 : 0) + (uint64_t)PCOUNT * np * 8 + 2 * 8;
}

}  // namespace

extern "C" {

int lbm_restart_write(lbm_handle *h, const char *dir) {
    if (!h || !dir) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    const DevP &P = h->P;
    std::vector<double> f((size_t)P.nn * 9), rho(P.nn), ux(P.nn), uy(P.nn), g((size_t)P.nn * 2),
        obn(P.nn), part((size_t)PCOUNT * std::max(1, P.np));
    std::vector<int32_t> img(P.nn);
    int rc;
    if ((rc = lbm_get_f(h, f.data())) || (rc = lbm_get_fields(h, rho.data(), ux.data(), uy.data(),
                                                                g.data(), obn.data())) ||
        (rc = lbm_get_image(h, img.data())) || (rc = lbm_get_particles(h, part.data())))
        return rc;
    // OMEGA_MRT(0:8): the relaxation vector the last node of the J-outer /
    // I-inner sweep leaves behind (F6:958, F6:1009-1011)
    const double om = h->prm.rheology != LBM_NEWTONIAN ? obn[(size_t)P.ny * P.W + P.nx] : P.omega;
    const double q = P.q_mode ? (8.0 * (2.0 - om)) / (8.0 - om) : P.sq;
    const double omv[9] = {P.s0, P.s1, P.s2, P.s3, q, P.s5, q, om, om};
    bool ok = write_nodes(join_path(dir, RS_F), P, 9, f.data()) &&
              write_nodes(join_path(dir, RS_UX), P, 1, ux.data()) &&
              write_nodes(join_path(dir, RS_UY), P, 1, uy.data()) &&
              write_nodes(join_path(dir, RS_RHO), P, 1, rho.data()) &&
              write_nodes(join_path(dir, RS_G), P, 2, g.data()) &&
              write_nodes(join_path(dir, RS_OBN), P, 1, obn.data());
    if (ok) {
        // DOMAIN_IMAGE is an integer array: written as integers
        FILE *fp = std::fopen(join_path(dir, RS_IMG).c_str(), "w");
        ok = fp != nullptr;
        for (int i = 0; ok && i <= P.nx; ++i)
            for (int j = 0; j <= P.ny; ++j)
                std::fprintf(fp, " %d\n", (int)img[(size_t)j * P.W + i]);
        if (fp) ok = (std::fclose(fp) == 0) && ok;
    }
    if (ok) {
        FILE *fp = std::fopen(join_path(dir, RS_PART).c_str(), "w");
        ok = fp != nullptr;
        for (int p = 0; ok && p < P.np; ++p) {
            for (int c = 0; c < 8; ++c)
                std::fprintf(fp, " %.17g", part[(size_t)RS_PFIELDS[c] * P.np + p]);
            std::fputc('\n', fp);
        }
        if (fp) ok = (std::fclose(fp) == 0) && ok;
    }
    if (ok) {
        FILE *fp = std::fopen(join_path(dir, RS_OM).c_str(), "w");
        ok = fp != nullptr;
        for (int k = 0; ok && k < 9; ++k) std::fprintf(fp, " %.17g", omv[k]);
        if (fp) { std::fputc('\n', fp); ok = (std::fclose(fp) == 0) && ok; }
    }
    if (ok) {
        // opened by WRITE_OUTPUT (F6:387) with nothing written to it
        FILE *fp = std::fopen(join_path(dir, RS_WALL).c_str(), "w");
        ok = fp != nullptr && std::fclose(fp) == 0;
    }
    if (!ok) return fail(h, LBM_ERR_IO, std::string("cannot write restart files in ") + dir);
    return LBM_OK;
}

int lbm_restart_read(lbm_handle *h, const char *dir) {
    if (!h || !dir) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    const size_t nn = (size_t)P.nn;
    std::vector<double> rec, f(nn * 9), rho(nn), ux(nn), uy(nn), g(nn * 2), obn(nn);
    std::string why;
    struct Item { const char *name; int nc; double *dst; } items[] = {
        {RS_F, 9, f.data()}, {RS_UX, 1, ux.data()}, {RS_UY, 1, uy.data()},
        {RS_RHO, 1, rho.data()}, {RS_G, 2, g.data()}, {RS_OBN, 1, obn.data()}};
    for (const Item &it : items) {
        if (!read_numbers(join_path(dir, it.name), nn * it.nc, rec, why))
            return fail(h, LBM_ERR_IO, why);
        nodes_from_records(P, it.nc, rec, it.dst);
    }
    std::vector<double> prec, lag;
    if (!read_numbers(join_path(dir, RS_PART), (size_t)8 * P.np, prec, why) ||
        !read_numbers(join_path(dir, RS_OM), 9, lag, why))
        return fail(h, LBM_ERR_IO, why);
    // CENTER, VELOCITY and FORCE replace those fields of the particle
    // record; the rest (diameter, density, mass, ...) is the record already
    // set, as the reference sets them from its parameters before reading
    std::vector<double> part((size_t)PCOUNT * std::max(1, P.np));
    int rc = P.np > 0 ? lbm_get_particles(h, part.data()) : LBM_OK;
    if (rc) return rc;
    for (int p = 0; p < P.np; ++p)
        for (int c = 0; c < 8; ++c) part[(size_t)RS_PFIELDS[c] * P.np + p] = prec[(size_t)8 * p + c];
    if ((rc = lbm_set_f(h, f.data())) ||
        (rc = lbm_set_fields(h, rho.data(), ux.data(), uy.data(), g.data(), obn.data())) ||
        (P.np > 0 && (rc = lbm_set_particles(h, P.np, part.data()))))
        return rc;
    // the module-global rate is entry 7 (omega) of OMEGA_MRT(0:8)
    HIPCHECK(h, hipMemcpyAsync(h->lag, lag.data() + 7, 8, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_snapshot_size(const lbm_handle *h, uint64_t *bytes) {
    if (!h || !bytes) return LBM_ERR_PARAM;
    *bytes = snapshot_bytes(h->P.nn, h->P.np, h->prm.storage);
    return LBM_OK;
}

int lbm_snapshot_save(lbm_handle *h, void *buf, uint64_t bytes) {
    if (!h || !buf) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    const DevP &P = h->P;
    const uint64_t need = snapshot_bytes(P.nn, P.np, h->prm.storage);
    if (bytes < need) return fail(h, LBM_ERR_CAPACITY, "snapshot buffer too small");
    SnapHeader hd;
    std::memset(&hd, 0, sizeof(hd));
    std::memcpy(hd.magic, SNAP_MAGIC, 8);
    hd.version = 2; hd.gen_base = h->sr_next;
    hd.nx = P.nx; hd.ny = P.ny; hd.maxp = P.maxp; hd.np = P.np;
    hd.storage = h->prm.storage; hd.rheology = h->prm.rheology; hd.bytes = need;
    char *o = (char *)buf;
    std::memcpy(o, &hd, sizeof(hd));
    double *d = (double *)(o + sizeof(hd));
    const size_t nn = (size_t)P.nn;
    int rc;
    if ((rc = lbm_get_f(h, d + SN_F * nn)) || (rc = lbm_get_fpost(h, d + SN_FP * nn)) ||
        (rc = lbm_get_fields(h, d + SN_RHO * nn, d + SN_UX * nn, d + SN_UY * nn, d + SN_G * nn,
                             d + SN_OBN * nn)) ||
        (rc = lbm_get_stress(h, d + SN_STRESS * nn)))
        return rc;
    int32_t *img = (int32_t *)(d + SN_PLANES * nn);
    if ((rc = lbm_get_image(h, img))) return rc;
    char *next = (char *)(img + nn);
    if (h->prm.storage != LBM_FP64) {
        if (!h->hexact_ok) {
            // canonical F set from outside: its deviations as a prologue derives them
            const float S = h->prm.storage == LBM_FP16 ? StF16::S : 1.0f;
            hipLaunchKernelGGL(k_hexact, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P, h->f, S,
                               h->hexact);
            if ((rc = check_launch(h, "k_hexact"))) return rc;
            h->hexact_ok = true;
        }
        float *hx = (float *)next;
        for (int k = 0; k < 9; ++k)
            HIPCHECK(h, hipMemcpyAsync(hx + k * nn, h->hexact + (size_t)k * P.plane, nn * 4,
                                       hipMemcpyDeviceToHost, h->stream));
        next += nn * 9 * 4;
    }
    double *tail = (double *)next;
    if (P.np > 0 && (rc = lbm_get_particles(h, tail))) return rc;
    tail += (size_t)PCOUNT * P.np;
    HIPCHECK(h, hipMemcpyAsync(tail, h->dpw[h->dcur], 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipMemcpyAsync(tail + 1, h->lag, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

int lbm_snapshot_load(lbm_handle *h, const void *buf, uint64_t bytes) {
    if (!h || !buf) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    SnapHeader hd;
    if (bytes < sizeof(hd)) return fail(h, LBM_ERR_PARAM, "snapshot too short");
    std::memcpy(&hd, buf, sizeof(hd));
    if (std::memcmp(hd.magic, SNAP_MAGIC, 8) != 0 || hd.version != 2)
        return fail(h, LBM_ERR_PARAM, "not an lbm snapshot");
    if (hd.storage != h->prm.storage)
        return fail(h, LBM_ERR_PARAM, "snapshot storage format differs from the handle's");
    if (hd.nx != P.nx || hd.ny != P.ny)
        return fail(h, LBM_ERR_PARAM, "snapshot grid differs from the handle's");
    if (hd.np > h->prm.max_particles)
        return fail(h, LBM_ERR_CAPACITY, "snapshot has more particles than max_particles");
    const size_t nn = (size_t)P.nn;
    const uint64_t need = hd.np >= 0 ? snapshot_bytes(P.nn, hd.np, hd.storage) : 0;
    if (hd.np < 0 || hd.bytes != need || bytes < need)
        return fail(h, LBM_ERR_PARAM, "snapshot size mismatch");
    const double *d = (const double *)((const char *)buf + sizeof(hd));
    int rc;
    if ((rc = lbm_set_f(h, d + SN_F * nn)) || (rc = lbm_set_fpost(h, d + SN_FP * nn)) ||
        (rc = lbm_set_fields(h, d + SN_RHO * nn, d + SN_UX * nn, d + SN_UY * nn, d + SN_G * nn,
                             d + SN_OBN * nn)) ||
        (rc = upload_planes(h, d + SN_STRESS * nn, 3, h->stress)))
        return rc;
    const int32_t *img = (const int32_t *)(d + SN_PLANES * nn);
    HIPCHECK(h, hipMemcpyAsync(h->image, img, nn * 4, hipMemcpyHostToDevice, h->stream));
    const char *next = (const char *)(img + nn);
    const float *hx = nullptr;
    if (h->prm.storage != LBM_FP64) {
        hx = (const float *)next;
        next += nn * 9 * 4;
    }
    const double *tail = (const double *)next;
    if ((rc = lbm_set_particles(h, hd.np, tail))) return rc;
    if (hx) {
        for (int k = 0; k < 9; ++k)
            HIPCHECK(h, hipMemcpyAsync(h->hexact + (size_t)k * P.plane, hx + k * nn, nn * 4,
                                       hipMemcpyHostToDevice, h->stream));
    }
    tail += (size_t)PCOUNT * hd.np;
    HIPCHECK(h, hipMemcpyAsync(h->dpw[h->dcur], tail, 8, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, hipMemcpyAsync(h->lag, tail + 1, 8, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    // the stochastic-rounding sequence continues where the saved run was
    // (launch tags go on), from the exact deviations
    h->sr_next = hd.gen_base;
    h->hexact_ok = hx != nullptr;
    return LBM_OK;
}

// --------------------------------------------------------------- drop-in
int lbm_collide(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    if (h->prm.rheology != LBM_NEWTONIAN && h->prm.rate_mode == LBM_RATE_LAGGED) {
        hipLaunchKernelGGL(k_collide_lagged, dim3(1), dim3(64), 0, h->stream, P, h->f, h->rho,
                           h->ux, h->uy, h->g, h->obn, h->stress, h->image, h->lag);
        return check_launch(h, "k_collide_lagged");
    }
    hipLaunchKernelGGL(k_collide, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P, h->f, h->rho,
                       h->ux, h->uy, h->g, h->obn, h->stress, h->image);
    return check_launch(h, "k_collide");
}

int lbm_copy_post(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    HIPCHECK(h, hipMemcpyAsync(h->fp, h->f, (size_t)9 * h->P.plane * 8, hipMemcpyDeviceToDevice,
                               h->stream));
    return LBM_OK;
}

int lbm_stream(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    hipLaunchKernelGGL(k_stream, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P, h->f, h->fp);
    return check_launch(h, "k_stream");
}

int lbm_boundary(lbm_handle *h, int32_t kk) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    (void)kk;
    const DevP &P = h->P;
    hipLaunchKernelGGL(k_boundary, dim3(nblk(P.nbnd)), dim3(NT), 0, h->stream, P, h->f, h->fp,
                       h->rho, h->ux, h->uy);
    return check_launch(h, "k_boundary");
}

int lbm_macros(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    hipLaunchKernelGGL(k_macros, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P, h->f, h->rho,
                       h->ux, h->uy);
    return check_launch(h, "k_macros");
}

int lbm_spm_phi(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    double *part = h->pbuf[h->pcur];
    if (P.np > 0) {
        hipLaunchKernelGGL(k_phi_patch, dim3(nblk((long long)h->PS * h->PS), P.np), dim3(NT), 0,
                           h->stream, P, part, h->phi, h->porig, h->PS);
        int rc = check_launch(h, "k_phi_patch");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_image, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P, part, h->phi,
                       h->porig, h->PS, h->image);
    return check_launch(h, "k_image");
}

int lbm_spm_forcing(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    hipLaunchKernelGGL(k_forcing, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P,
                       h->pbuf[h->pcur], h->phi, h->porig, h->PS, h->ux, h->uy, h->g);
    return check_launch(h, "k_forcing");
}

int lbm_hydro_force(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    if (P.np == 0) return LBM_OK;
    hipLaunchKernelGGL(k_hydro_seq, dim3(nblk(P.np)), dim3(NT), 0, h->stream, P,
                       h->pbuf[h->pcur], h->phi, h->porig, h->PS, h->rho, h->ux, h->uy);
    return check_launch(h, "k_hydro_seq");
}

int lbm_correct(lbm_handle *h) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    hipLaunchKernelGGL(k_correct, dim3(nblk(P.nn)), dim3(NT), 0, h->stream, P, h->ux, h->uy,
                       h->g);
    return check_launch(h, "k_correct");
}

// The pair list in (dcnt, dpairs): partner count and partners per particle.
static int gather_pairs(lbm_handle *h, const int *dcnt, const int *dpairs, int32_t *pairs,
                        int32_t max_pairs, int32_t *npairs) {
    const DevP &P = h->P;
    std::vector<int> cnt(P.np), nb((size_t)P.np * MAXNB);
    if (P.np > 0) {
        HIPCHECK(h, hipMemcpyAsync(cnt.data(), dcnt, (size_t)P.np * 4,
                                   hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(h, hipMemcpyAsync(nb.data(), dpairs, nb.size() * 4, hipMemcpyDeviceToHost,
                                   h->stream));
    }
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    int total = 0;
    bool over = false;
    for (int p = 0; p < P.np; ++p) {
        if (cnt[p] > MAXNB) over = true;
        int c = std::min(cnt[p], MAXNB);
        for (int q = 0; q < c; ++q) {
            if (pairs && total < max_pairs) {
                pairs[2 * total] = p;
                pairs[2 * total + 1] = nb[(size_t)p * MAXNB + q];
            }
            ++total;
        }
    }
    if (npairs) *npairs = total;
    if (over) return fail(h, LBM_ERR_CAPACITY, "more contact partners than recorded per particle");
    if (pairs && total > max_pairs) return fail(h, LBM_ERR_CAPACITY, "pair buffer too small");
    return LBM_OK;
}

int lbm_contact_force(lbm_handle *h, int32_t *pairs, int32_t max_pairs, int32_t *npairs) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    if (P.np > 0) {
        hipLaunchKernelGGL(k_contact, dim3(nblk(P.np)), dim3(NT), 0, h->stream, P,
                           h->pbuf[h->pcur], h->dpw[h->dcur], h->dpw[h->dcur ^ 1], h->pairs,
                           h->paircnt);
        int rc = check_launch(h, "k_contact");
        if (rc) return rc;
        h->dcur ^= 1;
    }
    if (pairs || npairs) return gather_pairs(h, h->paircnt, h->pairs, pairs, max_pairs, npairs);
    return LBM_OK;
}

int lbm_last_pairs(lbm_handle *h, int32_t *pairs, int32_t max_pairs, int32_t *npairs) {
    if (!h) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    return gather_pairs(h, h->cpaircnt[h->last_pairs], h->cpairs[h->last_pairs], pairs,
                        max_pairs, npairs);
}

int lbm_particle_update(lbm_handle *h, int32_t kk) {
    if (!h) return LBM_ERR_PARAM;
    LBM_WRITES_STATE(h);
    const DevP &P = h->P;
    if (P.np == 0) return LBM_OK;
    hipLaunchKernelGGL(k_update, dim3(nblk(P.np)), dim3(NT), 0, h->stream, P,
                       h->pbuf[h->pcur], (int)kk);
    return check_launch(h, "k_update");
}

int lbm_total_mass(lbm_handle *h, double *mass) {
    if (!h || !mass) return LBM_ERR_PARAM;
    LBM_READS_STATE(h);
    const DevP &P = h->P;
    unsigned nb = nblk(P.nn);
    hipLaunchKernelGGL(k_sum_partial, dim3(nb), dim3(NT), 0, h->stream, h->rho, P.nn, h->partial);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(NT), 0, h->stream, h->partial, (int)nb,
                       h->scalar);
    int rc = check_launch(h, "k_sum");
    if (rc) return rc;
    HIPCHECK(h, hipMemcpyAsync(mass, h->scalar, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

}  // extern "C"

// ----------------------------------------------------------------- fused
namespace {

// Where a call's launches sit in a chain of fused launches: chain index,
// tag, list generation and rounding position of its launch s = 0, and the
// particle and wall-distance buffers the chain started from.  A call that
// starts a chain runs the prologue as s = 0; a call that continues one
// starts at s = 1 (its s = 0 is the last launch of the call before).
struct Chain {
    int c0 = 0;
    unsigned tag0 = 0, lg0 = 0, sr0 = 0;
    int pbase = 0, dbase = 0;
};

// Arguments of launch s (chain launch c0 + s) of a fused call of nsteps
// steps on handle h; fin: a FINAL launch with the inputs of that chain
// launch.  step_fields() sets the per-step fields.
template <bool DEV>
StepArgs make_args(const lbm_handle *h, int32_t kk0, bool record, int s, bool fin,
                   const Chain &ch) {
    const DevP &P = h->P;
    StepArgs a;
    std::memset(&a, 0, sizeof(a));
    a.P = P;
    a.cf = h->f; a.cfp = h->fp; a.crho = h->rho; a.cux = h->ux; a.cuy = h->uy; a.cg = h->g;
    a.cobn = h->obn; a.cstress = h->stress; a.cimage = h->image;
    for (int b = 0; b < 2; ++b) {
        a.astress[b] = h->astress[b];
        a.aimage[b] = h->aimage[b];
        a.mpart[b] = h->mpart[b];
    }
    const int c = ch.c0 + s;
    a.pairs = h->cpairs[c & 1]; a.paircnt = h->cpaircnt[c & 1];
    a.n_pblocks = P.np * h->kpart;
    a.kpart = h->kpart;
    a.ntiles = (P.nn + NT - 1) / NT;
    a.err = h->err_flag;
    a.hist_base = record ? h->hist_part_use : nullptr;
    a.slots_base = record ? h->hist_slots_use : nullptr;
    a.rowp = P.np * HPC;
    a.kk0 = kk0;
    a.tag0 = ch.tag0; a.lg0 = ch.lg0; a.sr0 = ch.sr0;
    a.c0par = ch.c0 & 1;
    a.fin = fin;
    a.tpb = 1;
    a.nred = (record && s > 0) ? 1 : 0;  // a reducer block when there is something to reduce
    // tile lists: read by this launch's tiles (its generation), written for
    // the next launch
    const unsigned g = ch.lg0 + (unsigned)s;
    a.tl_hdr_in = h->tl_hdr[g & 1]; a.tl_ids_in = h->tl_ids[g & 1];
    a.tl_hdr_out = h->tl_hdr[(g + 1) & 1]; a.tl_ids_out = h->tl_ids[(g + 1) & 1];
    a.fslot = h->fslot;
    void *buf[2] = {h->sA, h->sB};
    if (fin) a.hx = h->hexact;
    if (c == 0) {
        a.hx_in = h->hexact_ok ? h->hexact : nullptr;
        // prologue: collide step kk0's canonical state into f*_{kk0+1};
        // particles to kk0+1
        a.dst = buf[0];
        a.erho_dst = h->erho[0];
        a.rate_out = h->rate[0];
        a.pin = h->pbuf[ch.pbase]; a.pout = h->pbuf[ch.pbase ^ 1];
        a.spf_out = DEV ? h->spf[ch.pbase ^ 1] : nullptr;
        a.pbox_in = h->pbox[ch.pbase]; a.pbox_out = h->pbox[ch.pbase ^ 1];
        a.dpw_in = h->dpw[ch.dbase]; a.dpw_out = h->dpw[ch.dbase ^ 1];
        return a;
    }
    a.src = buf[(c - 1) & 1];
    a.dst = buf[c & 1];
    a.erho_src = h->erho[(c - 1) & 1];
    a.erho_dst = h->erho[c & 1];
    a.rate_in = h->rate[(c - 1) & 1];
    a.rate_out = h->rate[c & 1];
    const int pc = ch.pbase ^ (c & 1);
    a.pin = h->pbuf[pc];
    a.pout = fin ? h->pbuf[pc] : h->pbuf[pc ^ 1];
    a.spf_in = DEV ? h->spf[pc] : nullptr;
    a.spf_out = (DEV && !fin) ? h->spf[pc ^ 1] : nullptr;
    a.pbox_in = h->pbox[pc];
    a.pbox_out = h->pbox[pc ^ 1];
    const int dc = ch.dbase ^ (c & 1);
    a.dpw_in = h->dpw[dc]; a.dpw_out = h->dpw[dc ^ 1];
    return a;
}

// A chain that starts from the canonical state (prologue at s = 0).
Chain new_chain(const lbm_handle *h) {
    Chain ch;
    ch.c0 = 0;
    ch.tag0 = h->wtag_next;
    ch.lg0 = h->lg_next;
    ch.sr0 = h->sr_next;
    ch.pbase = h->pcur;
    ch.dbase = h->dcur;
    return ch;
}

// The chain of the resident run, continued (s = 1 is the next launch).
Chain resident_chain(const lbm_handle *h) {
    Chain ch;
    ch.c0 = h->res_c;
    ch.tag0 = h->wtag_next - 1u;
    ch.lg0 = h->res_lg;
    ch.sr0 = h->sr_next;
    ch.pbase = h->res_pbase;
    ch.dbase = h->res_dbase;
    return ch;
}

// Handle state after a call of nsteps launches on chain ch ended with
// chain launch c0 + nsteps.  fin: that launch was a FINAL one (the canonical
// arrays hold its iteration and nothing stays resident), else the run is
// resident.  Either way the next collision (of iteration kk0 + nsteps + 1)
// takes the rounding position of launch nsteps + 1 minus one, so calls
// split anywhere give the results of one call.
void finish_chain(lbm_handle *h, int32_t kk0, int32_t nsteps, bool record, bool fin,
                  bool prologue, const Chain &ch) {
    const int m = ch.c0 + nsteps;
    h->wtag_next += (unsigned)nsteps + (prologue ? 1u : 0u);
    h->sr_next = ch.sr0 + (unsigned)nsteps;
    h->res_lg = ch.lg0 + (unsigned)nsteps;
    h->lg_next = h->res_lg + 2u;
    h->res_kk = kk0 + nsteps;
    h->res_c = m;
    h->res_pbase = ch.pbase;
    h->res_dbase = ch.dbase;
    h->res_yield = record;
    if (fin) {
        h->res = false;
        h->canon = true;
        h->hexact_ok = h->hexact != nullptr;
        h->pcur = ch.pbase ^ (m & 1);
        h->dcur = ch.dbase ^ (m & 1);
        h->last_pairs = (m - 1) & 1;
    } else {
        h->res = true;
        h->canon = false;
    }
}

// Error bits reported by the kernels of the last call (cleared when read).
int check_flags(lbm_handle *h) {
    int ef = 0;
    HIPCHECK(h, hipMemcpyAsync(&ef, h->err_flag, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (!ef) return LBM_OK;
    (void)hipMemsetAsync(h->err_flag, 0, 4, h->stream);
    (void)hipStreamSynchronize(h->stream);
    if (ef & LBM_EF_NAN)
        return fail(h, LBM_ERR_NAN, "non-finite density or particle state: the run diverged");
    if (ef & LBM_EF_LIST)
        return fail(h, LBM_ERR_CAPACITY, "more particles reach one grid tile than its candidate list holds");
    if (ef & LBM_EF_SLOT)
        return fail(h, LBM_ERR_CAPACITY, "a reach box covers more tiles than its force partial slots");
    return fail(h, LBM_ERR_HIP, "force partials of a particle did not appear within the wait bound");
}

// The fused step kernel of one boundary mode (each instantiation carries only
// its own edge rule).
template <class St, bool DEV, bool FINAL>
void launch_step(int boundary, dim3 grid, dim3 block, hipStream_t stream, const StepArgs &a) {
    if (boundary == LBM_ZOU_HE)
        hipLaunchKernelGGL((k_step<St, DEV, FINAL, LBM_ZOU_HE>), grid, block, 0, stream, a);
    else if (boundary == LBM_EQUILIBRIUM)
        hipLaunchKernelGGL((k_step<St, DEV, FINAL, LBM_EQUILIBRIUM>), grid, block, 0, stream, a);
    else
        hipLaunchKernelGGL((k_step<St, DEV, FINAL, LBM_BOUNCE_BACK>), grid, block, 0, stream, a);
}

// nsteps fused steps kk0+1 .. kk0+nsteps: the resident run continues when
// it ended at kk0 (and, for a recording call, stored the yielded counts of
// its last collision); otherwise a new chain starts from the canonical
// state with the prologue.  The last launch is an ordinary one: the run
// stays resident and the canonical arrays are materialised on demand.
template <class St, bool DEV>
int run_fused(lbm_handle *h, int32_t kk0, int32_t nsteps, bool record, int timed,
              double *ms_total, double *ms_main) {
    const DevP &P = h->P;
    const long long ntiles = (P.nn + NT - 1) / NT;
    const int n_pblocks = P.np * h->kpart;
    const bool cont = h->res && h->res_kk == kk0 && (!record || h->res_yield);
    int rc;
    if (!cont) {
        if ((rc = ensure_canonical(h))) return rc;
        drop_resident(h);
    }
    const Chain ch = cont ? resident_chain(h) : new_chain(h);
    dim3 grid((unsigned)(n_pblocks + ntiles)), block(NT);
    // the step launches of a recording call carry one reducer block
    dim3 sgrid((unsigned)(n_pblocks + (record ? 1 : 0) + ntiles));

    if (timed) (void)hipEventRecord(h->ev0, h->stream);
    StepArgs a;
    if (!cont) {
        a = make_args<DEV>(h, kk0, record, 0, false, ch);
        step_fields(a, 0, nsteps);
        finish_head(a, (unsigned)storage_bytes(h->prm.storage));
        hipLaunchKernelGGL((k_prologue<St, DEV>), grid, block, 0, h->stream, a);
        if ((rc = check_launch(h, "k_prologue"))) return rc;
    }
    // the step launches run back to back between evA and evB
    if (timed) (void)hipEventRecord(h->evA, h->stream);
    for (int s = 1; s <= nsteps; ++s) {
        a = make_args<DEV>(h, kk0, record, s, false, ch);
        step_fields(a, s, nsteps);
        finish_head(a, (unsigned)storage_bytes(h->prm.storage));
        launch_step<St, DEV, false>(P.boundary, sgrid, block, h->stream, a);
        if ((rc = check_launch(h, "k_step"))) return rc;
    }
    if (timed) (void)hipEventRecord(h->evB, h->stream);
    if (record) {
        // the mass partials of the last launch
        hipLaunchKernelGGL(k_monitor_final, dim3(1), block, 0, h->stream, a);
        if ((rc = check_launch(h, "k_monitor_final"))) return rc;
    }
    if (timed) (void)hipEventRecord(h->ev1, h->stream);
    finish_chain(h, kk0, nsteps, record, false, !cont, ch);
    if ((rc = check_flags(h))) return rc;
    if (timed) {
        HIPCHECK(h, hipEventSynchronize(h->ev1));
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, h->ev0, h->ev1);
        *ms_total = t;
        if (ms_main) {
            float tm = 0.0f;
            (void)hipEventElapsedTime(&tm, h->evA, h->evB);
            *ms_main = tm;
        }
    }
    return LBM_OK;
}

// The FINAL launch that writes the loop-end state of iteration res_kk from
// the inputs of the resident run's last launch.
template <class St, bool DEV>
int materialise(lbm_handle *h) {
    const DevP &P = h->P;
    const long long ntiles = (P.nn + NT - 1) / NT;
    Chain ch;
    ch.c0 = h->res_c - 1;
    ch.tag0 = h->wtag_next - 1u;
    ch.lg0 = h->res_lg - 1u;
    ch.pbase = h->res_pbase;
    ch.dbase = h->res_dbase;
    StepArgs a = make_args<DEV>(h, h->res_kk - 1, false, 1, true, ch);
    step_fields(a, 1, 1);
    finish_head(a, (unsigned)storage_bytes(h->prm.storage));
    launch_step<St, DEV, true>(P.boundary, dim3((unsigned)(P.np * h->kpart + ntiles)), dim3(NT),
                               h->stream, a);
    int rc = check_launch(h, "k_step (final)");
    if (rc) return rc;
    h->wtag_next += 1u;
    h->canon = true;
    h->hexact_ok = h->hexact != nullptr;
    h->pcur = h->res_pbase ^ (h->res_c & 1);
    h->dcur = h->res_dbase ^ (h->res_c & 1);
    h->last_pairs = (h->res_c - 1) & 1;
    return check_flags(h);
}

int fused_dispatch(lbm_handle *h, int32_t kk0, int32_t nsteps, bool record, int timed,
                   double *ms_total, double *ms_main) {
    switch (h->prm.storage) {
        case LBM_FP64: return run_fused<StF64, false>(h, kk0, nsteps, record, timed, ms_total, ms_main);
        case LBM_FP32: return run_fused<StF32, true>(h, kk0, nsteps, record, timed, ms_total, ms_main);
        case LBM_FP16: return run_fused<StF16, true>(h, kk0, nsteps, record, timed, ms_total, ms_main);
        default: return run_fused<StBF16, true>(h, kk0, nsteps, record, timed, ms_total, ms_main);
    }
}

int ensure_canonical(lbm_handle *h) {
    if (!h->res || h->canon) return LBM_OK;
    switch (h->prm.storage) {
        case LBM_FP64: return materialise<StF64, false>(h);
        case LBM_FP32: return materialise<StF32, true>(h);
        case LBM_FP16: return materialise<StF16, true>(h);
        default: return materialise<StBF16, true>(h);
    }
}

int check_fused_mode(lbm_handle *h) {
    if (!h->fused_ok)
        return fail(h, LBM_ERR_UNSUPPORTED,
                    "the fused path addresses each population buffer below 4 GiB (FP64: about "
                    "59.6 million nodes); use the per-subroutine entries for this grid");
    if (h->prm.rheology != LBM_NEWTONIAN && h->prm.rate_mode == LBM_RATE_LAGGED)
        return fail(h, LBM_ERR_UNSUPPORTED,
                    "the lagged module-global rate is a sequential recurrence; the GPU path "
                    "implements rate_mode LBM_RATE_LOCAL");
    return LBM_OK;
}

// Device history buffers for nsteps rows, monitor slots cleared.
int prepare_history(lbm_handle *h, int32_t nsteps) {
    const size_t rowp = (size_t)h->P.np * HPC;
    if ((size_t)nsteps > h->hist_rows || !h->hist_part) {
        if (h->hist_part) (void)hipFree(h->hist_part);
        if (h->hist_slots) (void)hipFree(h->hist_slots);
        h->hist_part = nullptr;
        h->hist_slots = nullptr;
        size_t rows = std::max<size_t>(nsteps, 16);
        HIPCHECK(h, hipMalloc((void **)&h->hist_part, std::max<size_t>(rows * rowp, 1) * 8));
        HIPCHECK(h, hipMalloc((void **)&h->hist_slots, rows * 2 * NSLOT * 8));
        h->hist_rows = rows;
    }
    HIPCHECK(h, hipMemsetAsync(h->hist_slots, 0, (size_t)nsteps * 2 * NSLOT * 8, h->stream));
    h->hist_part_use = h->hist_part;
    h->hist_slots_use = h->hist_slots;
    return LBM_OK;
}

// History rows of the last recorded call: per step the particle records,
// then SUM(RHO) and the yielded-node count.
int collect_history(lbm_handle *h, int32_t nsteps, double *hist) {
    const size_t rowp = (size_t)h->P.np * HPC;
    std::vector<double> hp(std::max<size_t>((size_t)nsteps * rowp, 1));
    std::vector<unsigned long long> hs((size_t)nsteps * 2 * NSLOT);
    if (rowp)
        HIPCHECK(h, hipMemcpyAsync(hp.data(), h->hist_part, (size_t)nsteps * rowp * 8,
                                   hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipMemcpyAsync(hs.data(), h->hist_slots, hs.size() * 8, hipMemcpyDeviceToHost,
                               h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    const size_t row = rowp + 2;
    for (int s = 0; s < nsteps; ++s) {
        double *out = hist + (size_t)s * row;
        for (size_t c = 0; c < rowp; ++c) out[c] = hp[(size_t)s * rowp + c];
        long long msum = 0, ysum = 0;
        for (int q = 0; q < NSLOT; ++q) {
            msum += (long long)hs[(size_t)s * 2 * NSLOT + q];
            ysum += (long long)hs[(size_t)s * 2 * NSLOT + NSLOT + q];
        }
        out[rowp] = (double)msum / MASS_SCALE;
        out[rowp + 1] = (double)ysum;
    }
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_step(lbm_handle *h, int32_t kk0, int32_t nsteps, double *hist) {
    if (!h || nsteps < 0) return LBM_ERR_PARAM;
    if (nsteps == 0) return LBM_OK;
    int rc = check_fused_mode(h);
    if (rc) return rc;
    const bool record = hist != nullptr;
    if (record && (rc = prepare_history(h, nsteps))) return rc;
    if ((rc = fused_dispatch(h, kk0, nsteps, record, 0, nullptr, nullptr))) return rc;
    if (record) return collect_history(h, nsteps, hist);
    return LBM_OK;
}

int lbm_step_timed(lbm_handle *h, int32_t kk0, int32_t nsteps, double *ms_total,
                   double *ms_main) {
    if (!h || nsteps <= 0 || !ms_total) return LBM_ERR_PARAM;
    int rc0 = check_fused_mode(h);
    if (rc0) return rc0;
    return fused_dispatch(h, kk0, nsteps, false, 1, ms_total, ms_main);
}

}  // extern "C"

namespace {

__global__ void k_fill_u64(unsigned long long *x, long long n, unsigned long long v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// History rows in lbm_step's layout: per step the particle rows, then the
// mass (sum of the NSLOT fixed-point slots / 2^36) and the yielded count.
__global__ void k_hist_format(const double *part, const unsigned long long *slots, int nsteps,
                              long long rowp, double *out) {
    const long long row = rowp + 2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nsteps * row) return;
    const long long s = i / row, c = i - s * row;
    if (c < rowp) {
        out[i] = part[s * rowp + c];
        return;
    }
    const unsigned long long *q = slots + s * 2 * NSLOT + (c == rowp ? 0 : NSLOT);
    unsigned long long sum = 0;
    for (int k = 0; k < NSLOT; ++k) sum += q[k];
    out[i] = c == rowp ? (double)(long long)sum / MASS_SCALE : (double)(long long)sum;
}

}  // namespace

extern "C" {

int lbm_step_hist_static(lbm_handle *h, int32_t kk0, int32_t nsteps, double *hist) {
    if (!h || nsteps < 0 || !hist) return LBM_ERR_PARAM;
    if (nsteps == 0) return LBM_OK;
    const size_t rowp = (size_t)h->P.np * HPC, row = rowp + 2;
    const size_t nslots = (size_t)nsteps * 2 * NSLOT;
    const size_t words = nslots + (size_t)nsteps * (rowp + row);
    if (words > SCRATCH_WORDS)
        return fail(h, LBM_ERR_CAPACITY,
                    "lbm_step_hist_static: " + std::to_string(nsteps) + " steps of " +
                        std::to_string(h->P.np) + " particles need " +
                        std::to_string(words * 8) +
                        " bytes of scratch (steps x (24 x particles + 66) x 8); the static "
                        "scratch holds " + std::to_string(SCRATCH_WORDS * 8) + " bytes");
    int rc = check_fused_mode(h);
    if (rc) return rc;
    std::string why;
    unsigned long long *base = (unsigned long long *)scratch_address(h->dev, &why);
    if (!base) return fail(h, LBM_ERR_HIP, "lbm_step_hist_static: " + why);
    unsigned long long *slots = base;
    double *part = (double *)(base + nslots);
    double *fmt = part + (size_t)nsteps * rowp;
    // the slots start at zero (a fill kernel, not a memset)
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk((long long)nslots)), dim3(NT), 0, h->stream, slots,
                       (long long)nslots, 0ull);
    if ((rc = check_launch(h, "k_fill_u64"))) return rc;
    h->hist_part_use = part;
    h->hist_slots_use = slots;
    rc = fused_dispatch(h, kk0, nsteps, true, 0, nullptr, nullptr);
    h->hist_part_use = nullptr;
    h->hist_slots_use = nullptr;
    if (rc) return rc;
    hipLaunchKernelGGL(k_hist_format, dim3(nblk((long long)nsteps * (long long)row)), dim3(NT), 0,
                       h->stream, part, slots, nsteps, (long long)rowp, fmt);
    if ((rc = check_launch(h, "k_hist_format"))) return rc;
    HIPCHECK(h, hipMemcpyAsync(hist, fmt, (size_t)nsteps * row * 8, hipMemcpyDefault, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

}  // extern "C"

// ================================================================= batch
// Independent simulations of different sizes and parameters stepped together:
// one launch per time step covers every grid (block map bmap), the per-step
// arguments of each simulation come from a small device table, and the
// steps between the first and the last two of a call can be replayed from a
// captured hipGraph.
struct lbm_batch {
    std::vector<lbm_handle *> sims;
    int storage = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<int> tpb;                 // 256-node tiles per block, per simulation
    int2 *bmap = nullptr;                 // block -> (simulation, block within it)
    size_t bmap_cap = 0, nblocks = 0;
    bool map_dirty = true;
    std::vector<long long> map_key;       // (particles, helper blocks per particle, tiles per block) the map was built for
    StepArgs *dargs = nullptr;            // [4][n]: prologue, odd steps, even steps, final
    int *sdev = nullptr;                  // first step index of a replayed chunk
    int chunk = 0;                        // graph chunk length (even), 0: no graph
    hipGraphExec_t gexec = nullptr;
    int gexec_chunk = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool all_bounce_back = true;          // every simulation has LBM_BOUNCE_BACK walls
    long long max_tiles = 0;              // 256-node tiles of the largest grid
};

namespace {

constexpr int BIG_NSTEPS = 0x3fffffff;  // chunk launches are never the last two steps

int bfail(lbm_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}

#define BHIPCHECK(b, call)                                                               \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return bfail((b), LBM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

void drop_graph(lbm_batch *b) {
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->gexec_chunk = 0;
}

// Block map of the simulations in `which`: every particle block first (they
// start early and carry the particle update), then the tile blocks, placed so
// that each of the 8 XCDs runs a contiguous range of (simulation, tile) order.
int build_map(lbm_batch *b, const std::vector<int> &which) {
    std::vector<int2> map;
    for (int i : which) {
        const lbm_handle *h = b->sims[i];
        const int npb = h->P.np * h->kpart;
        for (int k = 0; k < npb; ++k) map.push_back(make_int2(i, k));
    }
    std::vector<int2> tiles;
    for (int i : which) {
        const lbm_handle *h = b->sims[i];
        const int npb = h->P.np * h->kpart;
        const long long ntiles = (h->P.nn + NT - 1) / NT;
        const long long tb = (ntiles + b->tpb[i] - 1) / b->tpb[i];
        // block npb of each simulation is its monitor reducer (idle unless recording)
        tiles.push_back(make_int2(i, npb));
        for (long long k = 0; k < tb; ++k) tiles.push_back(make_int2(i, npb + 1 + (int)k));
    }
    // particle helper blocks first, then the tile blocks in one contiguous
    // range per XCD
    const long long T = (long long)tiles.size();
    const size_t base = map.size();
    map.resize(base + tiles.size());
    for (long long g = 0; g < T; ++g)
        map[base + g] = tiles[xcd_rank((int)(base + g), (int)base, (int)T)];
    if (map.size() > b->bmap_cap) {
        if (b->bmap) (void)hipFree(b->bmap);
        b->bmap = nullptr;
        BHIPCHECK(b, hipMalloc((void **)&b->bmap, map.size() * sizeof(int2)));
        b->bmap_cap = map.size();
    }
    BHIPCHECK(b, hipMemcpy(b->bmap, map.data(), map.size() * sizeof(int2), hipMemcpyHostToDevice));
    b->nblocks = map.size();
    drop_graph(b);
    return LBM_OK;
}

template <class St, bool DEV>
int batch_launches(lbm_batch *b, int32_t nsteps) {
    const int n = (int)b->sims.size();
    const dim3 grid((unsigned)b->nblocks), block(NT), fgrid((unsigned)((n + 63) / 64)), fblock(64);
    StepArgs *A0 = b->dargs, *Aodd = b->dargs + n, *Aeven = b->dargs + 2 * n,
             *Afin = b->dargs + 3 * n;
    auto step_args = [&](int s) { return (s & 1) ? Aodd : Aeven; };
    // per-step fields, then the launch that reads them
    auto launch = [&](StepArgs *A, const int *sdev, int s_const, int ns, bool fin) {
        hipLaunchKernelGGL(k_batch_fields, fgrid, fblock, 0, b->stream, A, n, sdev, s_const, ns);
        // every simulation with bounce-back walls: the kernel without the
        // other edge rules; otherwise each simulation's mode is read at run time
        const bool bb = b->all_bounce_back;
        if (fin && bb)
            hipLaunchKernelGGL((k_batch_step<St, DEV, true, LBM_BOUNCE_BACK>), grid, block, 0,
                               b->stream, (const StepArgs *)A, (const int2 *)b->bmap);
        else if (fin)
            hipLaunchKernelGGL((k_batch_step<St, DEV, true, -1>), grid, block, 0, b->stream,
                               (const StepArgs *)A, (const int2 *)b->bmap);
        else if (bb)
            hipLaunchKernelGGL((k_batch_step<St, DEV, false, LBM_BOUNCE_BACK>), grid, block, 0,
                               b->stream, (const StepArgs *)A, (const int2 *)b->bmap);
        else
            hipLaunchKernelGGL((k_batch_step<St, DEV, false, -1>), grid, block, 0, b->stream,
                               (const StepArgs *)A, (const int2 *)b->bmap);
    };
    hipLaunchKernelGGL(k_batch_fields, fgrid, fblock, 0, b->stream, A0, n, (const int *)nullptr, 0,
                       nsteps);
    hipLaunchKernelGGL((k_batch_prologue<St, DEV>), grid, block, 0, b->stream,
                       (const StepArgs *)A0, (const int2 *)b->bmap);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, LBM_ERR_HIP, std::string("k_batch_prologue: ") + hipGetErrorString(e));
    int s = 1;
    const int K = b->chunk;
    if (K > 0 && nsteps - 2 >= K) {
        const int nchunks = (nsteps - 2) / K;
        if (!b->gexec || b->gexec_chunk != K) {
            drop_graph(b);
            hipGraph_t g = nullptr;
            BHIPCHECK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            for (int j = 0; j < K; ++j) launch(step_args(1 + j), b->sdev, 1 + j, BIG_NSTEPS, false);
            hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, b->stream, b->sdev, K);
            BHIPCHECK(b, hipStreamEndCapture(b->stream, &g));
            hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess)
                return bfail(b, LBM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            b->gexec_chunk = K;
        }
        hipLaunchKernelGGL(k_counter_set, dim3(1), dim3(64), 0, b->stream, b->sdev, 0);
        for (int c = 0; c < nchunks; ++c) BHIPCHECK(b, hipGraphLaunch(b->gexec, b->stream));
        s = 1 + nchunks * K;
    }
    for (; s < nsteps; ++s) launch(step_args(s), nullptr, s, nsteps, false);
    launch(Afin, nullptr, nsteps, nsteps, true);
    // the last launch's monitor partials (a no-op per simulation unless recording)
    hipLaunchKernelGGL(k_batch_monitor_final, dim3((unsigned)n), block, 0, b->stream,
                       (const StepArgs *)Afin);
    e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, LBM_ERR_HIP, std::string("k_batch_step: ") + hipGetErrorString(e));
    return LBM_OK;
}

// nsteps fused steps of the simulations in `which` (every simulation when
// empty); hist (optional) receives each simulation's history rows in turn.
int batch_run(lbm_batch *b, int32_t kk0, int32_t nsteps, double *hist, bool timed,
              double *ms_total, const std::vector<int> &which_in) {
    if (nsteps <= 0) return LBM_OK;
    std::vector<int> which = which_in;
    if (which.empty())
        for (size_t i = 0; i < b->sims.size(); ++i) which.push_back((int)i);
    const bool all = which.size() == b->sims.size();
    const bool record = hist != nullptr;
    const size_t n = b->sims.size();
    std::vector<Chain> chains(n);
    for (int i : which) {
        lbm_handle *h = b->sims[i];
        int rc = check_fused_mode(h);
        if (rc) return bfail(b, rc, "simulation " + std::to_string(i) + ": " + h->err);
        // a batch call starts a chain from each simulation's canonical state
        // and ends it with a FINAL launch
        if ((rc = ensure_canonical(h)))
            return bfail(b, rc, "simulation " + std::to_string(i) + ": " + h->err);
        drop_resident(h);
        chains[i] = new_chain(h);
        if (record && (rc = prepare_history(h, nsteps)))
            return bfail(b, rc, "simulation " + std::to_string(i) + ": " + h->err);
        BHIPCHECK(b, hipStreamSynchronize(h->stream));
    }
    // argument tables
    std::vector<StepArgs> args(4 * n);
    std::memset(args.data(), 0, args.size() * sizeof(StepArgs));
    for (int i : which) {
        const lbm_handle *h = b->sims[i];
        const bool dev = b->storage != LBM_FP64;
        auto mk = [&](int s, bool fin) {
            StepArgs a = dev ? make_args<true>(h, kk0, record, s, fin, chains[i])
                             : make_args<false>(h, kk0, record, s, fin, chains[i]);
            a.tpb = b->tpb[i];
            a.nred = 1;  // the block map always holds the reducer block
            finish_head(a, (unsigned)storage_bytes(h->prm.storage));
            return a;
        };
        args[i] = mk(0, false);
        args[n + i] = mk(1, false);
        args[2 * n + i] = mk(2, false);
        args[3 * n + i] = mk(nsteps, true);
    }
    BHIPCHECK(b, hipMemcpyAsync(b->dargs, args.data(), args.size() * sizeof(StepArgs),
                                hipMemcpyHostToDevice, b->stream));
    // the map depends on each simulation's particle count, helper blocks per
    // particle and tiling
    std::vector<long long> key;
    for (size_t i = 0; i < n; ++i) {
        key.push_back(b->sims[i]->P.np);
        key.push_back(((long long)b->sims[i]->kpart << 32) | (unsigned)b->tpb[i]);
    }
    if (b->map_dirty || !all || key != b->map_key) {
        int rc = build_map(b, which);
        if (rc) return rc;
        b->map_dirty = !all;
        b->map_key = key;
    }
    if (timed) (void)hipEventRecord(b->ev0, b->stream);
    int rc;
    switch (b->storage) {
        case LBM_FP64: rc = batch_launches<StF64, false>(b, nsteps); break;
        case LBM_FP32: rc = batch_launches<StF32, true>(b, nsteps); break;
        case LBM_FP16: rc = batch_launches<StF16, true>(b, nsteps); break;
        default: rc = batch_launches<StBF16, true>(b, nsteps); break;
    }
    if (rc) return rc;
    if (timed) (void)hipEventRecord(b->ev1, b->stream);
    BHIPCHECK(b, hipStreamSynchronize(b->stream));
    // every simulation's error bits are read (and cleared); the first error
    // is reported
    int first = LBM_OK;
    for (int i : which) {
        lbm_handle *h = b->sims[i];
        finish_chain(h, kk0, nsteps, record, true, true, chains[i]);
        if ((rc = check_flags(h)) && first == LBM_OK)
            first = bfail(b, rc, "simulation " + std::to_string(i) + ": " + h->err);
    }
    if (timed) {
        float t = 0.0f;
        (void)hipEventElapsedTime(&t, b->ev0, b->ev1);
        *ms_total = t;
    }
    if (first) return first;
    if (record) {
        double *out = hist;
        for (int i : which) {
            lbm_handle *h = b->sims[i];
            if ((rc = collect_history(h, nsteps, out)))
                return bfail(b, rc, "simulation " + std::to_string(i) + ": " + h->err);
            out += (size_t)nsteps * ((size_t)h->P.np * HPC + 2);
        }
    }
    return LBM_OK;
}

}  // namespace

extern "C" {

const char *lbm_batch_last_error(const lbm_batch *b) { return b ? b->err.c_str() : "null batch"; }

int lbm_batch_create(int32_t n, const lbm_params *params, lbm_batch **out) {
    if (!out || !params || n <= 0) return LBM_ERR_PARAM;
    *out = nullptr;
    for (int32_t i = 1; i < n; ++i)
        if (params[i].storage != params[0].storage) {
            t_create_err = "all simulations of a batch use one storage format";
            return LBM_ERR_PARAM;
        }
    lbm_batch *b = new lbm_batch();
    b->storage = params[0].storage;
    for (int32_t i = 0; i < n; ++i) {
        lbm_handle *h = nullptr;
        int rc = lbm_create(&params[i], &h);
        if (rc) {
            t_create_err = "simulation " + std::to_string(i) + ": " +
                           (h ? h->err : t_create_err);
            if (h) lbm_destroy(h);
            lbm_batch_destroy(b);
            return rc;
        }
        b->sims.push_back(h);
        if (params[i].boundary != LBM_BOUNCE_BACK) b->all_bounce_back = false;
        b->max_tiles = std::max(b->max_tiles, (h->P.nn + NT - 1) / NT);
    }
    b->tpb.assign(n, 1);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&b->dargs, 4 * (size_t)n * sizeof(StepArgs)) != hipSuccess ||
        hipMalloc((void **)&b->sdev, sizeof(int)) != hipSuccess) {
        t_create_err = "batch stream or argument table allocation failed";
        lbm_batch_destroy(b);
        return LBM_ERR_HIP;
    }
    (void)hipEventCreate(&b->ev0);
    (void)hipEventCreate(&b->ev1);
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch *b) {
    if (!b) return LBM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (lbm_handle *h : b->sims) lbm_destroy(h);
    if (b->bmap) (void)hipFree(b->bmap);
    if (b->dargs) (void)hipFree(b->dargs);
    if (b->sdev) (void)hipFree(b->sdev);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LBM_OK;
}

int lbm_batch_size(const lbm_batch *b, int32_t *n) {
    if (!b || !n) return LBM_ERR_PARAM;
    *n = (int32_t)b->sims.size();
    return LBM_OK;
}

lbm_handle *lbm_batch_sim(lbm_batch *b, int32_t i) {
    if (!b || i < 0 || i >= (int32_t)b->sims.size()) return nullptr;
    return b->sims[i];
}

int lbm_batch_initialize(lbm_batch *b) {
    if (!b) return LBM_ERR_PARAM;
    for (size_t i = 0; i < b->sims.size(); ++i) {
        int rc = lbm_initialize(b->sims[i]);
        if (rc) return bfail(b, rc, "simulation " + std::to_string(i) + ": " + b->sims[i]->err);
    }
    return LBM_OK;
}

int lbm_batch_step(lbm_batch *b, int32_t kk0, int32_t nsteps, double *hist) {
    if (!b || nsteps < 0) return LBM_ERR_PARAM;
    return batch_run(b, kk0, nsteps, hist, false, nullptr, {});
}

int lbm_batch_step_timed(lbm_batch *b, int32_t kk0, int32_t nsteps, double *ms_total) {
    if (!b || nsteps <= 0 || !ms_total) return LBM_ERR_PARAM;
    return batch_run(b, kk0, nsteps, nullptr, true, ms_total, {});
}

int lbm_batch_set_graph(lbm_batch *b, int32_t chunk) {
    if (!b || chunk < 0 || (chunk & 1)) return LBM_ERR_PARAM;
    if (chunk != b->chunk) drop_graph(b);
    b->chunk = chunk;
    return LBM_OK;
}

int lbm_batch_set_tiles(lbm_batch *b, const int32_t *tiles_per_block) {
    if (!b || !tiles_per_block) return LBM_ERR_PARAM;
    for (size_t i = 0; i < b->sims.size(); ++i)
        if (tiles_per_block[i] < 1 || tiles_per_block[i] > 64) return LBM_ERR_PARAM;
    for (size_t i = 0; i < b->sims.size(); ++i) b->tpb[i] = tiles_per_block[i];
    b->map_dirty = true;
    return LBM_OK;
}

int lbm_batch_get_tiles(const lbm_batch *b, int32_t *tiles_per_block) {
    if (!b || !tiles_per_block) return LBM_ERR_PARAM;
    for (size_t i = 0; i < b->sims.size(); ++i) tiles_per_block[i] = b->tpb[i];
    return LBM_OK;
}

int lbm_batch_tune(lbm_batch *b, int32_t steps, int32_t *tiles_per_block) {
    if (!b || steps < 2) return LBM_ERR_PARAM;
    const size_t n = b->sims.size();
    // size buckets: floor(log2(nodes)) of each grid
    std::vector<int> bucket(n);
    for (size_t i = 0; i < n; ++i) {
        long long nn = b->sims[i]->P.nn;
        int k = 0;
        while ((1LL << (k + 1)) <= nn) ++k;
        bucket[i] = k;
    }
    std::vector<int> keys(bucket);
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const std::vector<int> all;  // every simulation, in one launch per step
    const int ns = std::max(steps, 8);
    // time of one step of the whole batch under the current tiling: a call
    // of 2 + ns steps less a call of 2 steps, so the call's prologue and final
    // launch cancel; the runs only time the tilings, so a state that diverges
    // under them (the tuner may run on an uninitialised batch) is not an
    // error here
    auto per_step = [&](double &ms) {
        double t_short = 0.0, t_long = 0.0;
        int rc = batch_run(b, 0, 2, nullptr, false, nullptr, all);  // settle
        if (!rc || rc == LBM_ERR_NAN) rc = batch_run(b, 0, 2, nullptr, true, &t_short, all);
        if (!rc || rc == LBM_ERR_NAN) rc = batch_run(b, 0, 2 + ns, nullptr, true, &t_long, all);
        if (rc && rc != LBM_ERR_NAN) return rc;
        ms = (t_long - t_short) / ns;
        return (int)LBM_OK;
    };
    // the best uniform tiling, then one pass over the size buckets trying
    // the other tilings for that bucket alone (kept when 1 % faster)
    const int cand[] = {1, 2, 4, 8};
    double best = 1e300;
    int best_t = 1;
    for (int t : cand) {
        for (size_t i = 0; i < n; ++i) b->tpb[i] = t;
        double ms = 0.0;
        int rc = per_step(ms);
        if (rc) return rc;
        if (ms < best) { best = ms; best_t = t; }
    }
    for (size_t i = 0; i < n; ++i) b->tpb[i] = best_t;
    for (int key : keys) {
        int keep = best_t;
        for (int t : cand) {
            if (t == best_t) continue;
            for (size_t i = 0; i < n; ++i)
                if (bucket[i] == key) b->tpb[i] = t;
            double ms = 0.0;
            int rc = per_step(ms);
            if (rc) return rc;
            if (ms < 0.99 * best) { best = ms; keep = t; }
        }
        for (size_t i = 0; i < n; ++i)
            if (bucket[i] == key) b->tpb[i] = keep;
    }
    b->map_dirty = true;
    drop_graph(b);
    if (tiles_per_block)
        for (size_t i = 0; i < n; ++i) tiles_per_block[i] = b->tpb[i];
    return LBM_OK;
}

}  // extern "C"


// ================================================================ adjoint
// Gradient of L = (1/W) sum_{k=N-W+1..N} V_y(P_k) for one disk with respect
// to (tau_y, mu0, rho_p) by the discrete adjoint of the fused step
// (lbm_adjoint.h).  The forward pass keeps a checkpoint every `chk` steps;
// the backward pass restores each checkpoint in turn, recomputes its
// segment one step per call and walks the segment backwards.
namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

int adj_alloc(lbm_handle *h, DevBuf &b, size_t bytes, size_t &total) {
    HIPCHECK(h, hipMalloc(&b.p, std::max<size_t>(bytes, 8)));
    total += bytes;
    return LBM_OK;
}

// canonical state needed to resume: F, RHO, UX, UY, G(2), OMEGA_BN planes
constexpr int CK_PLANES = 15;

struct Checkpoint {
    size_t off = 0;  // into the checkpoint buffer (doubles)
    int pcur = 0, dcur = 0;
    unsigned gen = 0;
};

int save_state(lbm_handle *h, double *dst, Checkpoint &c) {
    LBM_READS_STATE(h);
    const size_t pl = (size_t)h->P.plane;
    const size_t pc = (size_t)PCOUNT * h->P.maxp;
    double *d = dst;
    HIPCHECK(h, hipMemcpyAsync(d, h->f, 9 * pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += 9 * pl;
    HIPCHECK(h, hipMemcpyAsync(d, h->rho, pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += pl;
    HIPCHECK(h, hipMemcpyAsync(d, h->ux, pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += pl;
    HIPCHECK(h, hipMemcpyAsync(d, h->uy, pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += pl;
    HIPCHECK(h, hipMemcpyAsync(d, h->g, 2 * pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += 2 * pl;
    HIPCHECK(h, hipMemcpyAsync(d, h->obn, pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += pl;
    HIPCHECK(h, hipMemcpyAsync(d, h->pbuf[h->pcur], pc * 8, hipMemcpyDeviceToDevice, h->stream)); d += pc;
    HIPCHECK(h, hipMemcpyAsync(d, h->dpw[h->dcur], 8, hipMemcpyDeviceToDevice, h->stream));
    c.pcur = h->pcur; c.dcur = h->dcur; c.gen = h->sr_next;
    return LBM_OK;
}

int restore_state(lbm_handle *h, const double *src, const Checkpoint &c) {
    const size_t pl = (size_t)h->P.plane;
    const size_t pc = (size_t)PCOUNT * h->P.maxp;
    drop_resident(h);
    h->hexact_ok = false;  // F is restored from the checkpoint
    h->pcur = c.pcur; h->dcur = c.dcur; h->sr_next = c.gen;
    const double *d = src;
    HIPCHECK(h, hipMemcpyAsync(h->f, d, 9 * pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += 9 * pl;
    HIPCHECK(h, hipMemcpyAsync(h->rho, d, pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += pl;
    HIPCHECK(h, hipMemcpyAsync(h->ux, d, pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += pl;
    HIPCHECK(h, hipMemcpyAsync(h->uy, d, pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += pl;
    HIPCHECK(h, hipMemcpyAsync(h->g, d, 2 * pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += 2 * pl;
    HIPCHECK(h, hipMemcpyAsync(h->obn, d, pl * 8, hipMemcpyDeviceToDevice, h->stream)); d += pl;
    HIPCHECK(h, hipMemcpyAsync(h->pbuf[h->pcur], d, pc * 8, hipMemcpyDeviceToDevice, h->stream)); d += pc;
    HIPCHECK(h, hipMemcpyAsync(h->dpw[h->dcur], d, 8, hipMemcpyDeviceToDevice, h->stream));
    return LBM_OK;
}

}  // namespace

extern "C" int lbm_adjoint_velocity(lbm_handle *h, int32_t nsteps, int32_t window, int32_t chk,
                                    double *loss, double *grad, double *mem_bytes) {
    if (!h || nsteps < 1 || window < 1 || window > nsteps || chk < 1 || !loss || !grad)
        return LBM_ERR_PARAM;
    const DevP &P = h->P;
    const lbm_params &pr = h->prm;
    if (pr.storage != LBM_FP64 && pr.storage != LBM_FP32)
        return fail(h, LBM_ERR_UNSUPPORTED, "the adjoint recomputes from FP64 or FP32 storage only");
    if (P.np != 1) return fail(h, LBM_ERR_UNSUPPORTED, "the adjoint follows exactly one particle");
    if (pr.rheology == LBM_CASSON)
        return fail(h, LBM_ERR_UNSUPPORTED, "the adjoint differentiates the Newtonian and "
                                            "Papanastasiou-Bingham rates only");
    if (P.motion != LBM_MOTION_FREE)
        return fail(h, LBM_ERR_UNSUPPORTED, "the adjoint needs the free (Newton-Euler) particle update");
    for (int k = 0; k < 8; ++k)
        if (P.wall[k] != 0.0) return fail(h, LBM_ERR_UNSUPPORTED, "the adjoint needs walls at rest");
    if (P.corner[0] != 0.0 || P.corner[1] != 0.0)
        return fail(h, LBM_ERR_UNSUPPORTED, "the adjoint needs walls at rest");
    if (pr.boundary != LBM_BOUNCE_BACK)
        return fail(h, LBM_ERR_UNSUPPORTED, "the adjoint needs the bounce-back walls");
    if (pr.contact != LBM_CONTACT_GLOWINSKI)
        return fail(h, LBM_ERR_UNSUPPORTED, "the adjoint needs the Glowinski contact model");
    int rc = check_fused_mode(h);
    if (rc) return rc;

    const size_t pl = (size_t)P.plane;
    const size_t pc = (size_t)PCOUNT * P.maxp;
    const size_t ck_doubles = CK_PLANES * pl + pc + 1;
    const size_t snap_doubles = 10 * pl + pc;  // f_k (9), om_k, P_k
    const int nck = (nsteps + chk - 1) / chk;
    const unsigned nb = nblk(P.nn);
    size_t total = 0;
    DevBuf ckb, segb, lf[2], lw[2], mu[2], fbar, theta, partial, dstmap, ids, zero;
    if ((rc = adj_alloc(h, ckb, (size_t)(nck + 1) * ck_doubles * 8, total)) ||
        (rc = adj_alloc(h, segb, (size_t)chk * snap_doubles * 8, total)) ||
        (rc = adj_alloc(h, lf[0], 9 * pl * 8, total)) || (rc = adj_alloc(h, lf[1], 9 * pl * 8, total)) ||
        (rc = adj_alloc(h, lw[0], pl * 8, total)) || (rc = adj_alloc(h, lw[1], pl * 8, total)) ||
        (rc = adj_alloc(h, mu[0], ANP * 8, total)) || (rc = adj_alloc(h, mu[1], ANP * 8, total)) ||
        (rc = adj_alloc(h, fbar, 3 * 8, total)) || (rc = adj_alloc(h, theta, 3 * 8, total)) ||
        (rc = adj_alloc(h, partial, (size_t)nb * AR * 8, total)) ||
        (rc = adj_alloc(h, dstmap, 9 * pl * 4, total)) || (rc = adj_alloc(h, ids, 9 * pl * 8, total)) ||
        (rc = adj_alloc(h, zero, (size_t)P.nbnd * 8, total)))
        return rc;
    double *ck = (double *)ckb.p, *seg = (double *)segb.p;

    // inverse stream map (a permutation when the walls are at rest)
    HIPCHECK(h, hipMemsetAsync(zero.p, 0, (size_t)P.nbnd * 8, h->stream));
    HIPCHECK(h, hipMemsetAsync(ids.p, 0, 9 * pl * 8, h->stream));
    hipLaunchKernelGGL(k_adj_ids, dim3(nb), dim3(NT), 0, h->stream, P, (double *)ids.p);
    hipLaunchKernelGGL(k_adj_srcmap, dim3(nb), dim3(NT), 0, h->stream, P, (const double *)ids.p,
                       (const double *)zero.p, (int *)dstmap.p);
    if ((rc = check_launch(h, "k_adj_srcmap"))) return rc;
    {
        std::vector<int> src(9 * pl), dst(9 * pl, -1);
        HIPCHECK(h, hipMemcpyAsync(src.data(), dstmap.p, 9 * pl * 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(h, hipStreamSynchronize(h->stream));
        for (int k = 0; k < 9; ++k)
            for (long long x = 0; x < P.nn; ++x) {
                long long sid = src[(size_t)k * pl + x];
                long long q = sid / P.nn, y = sid - q * P.nn;
                size_t at = (size_t)q * pl + (size_t)y;
                if (q < 0 || q > 8 || dst[at] != -1)
                    return fail(h, LBM_ERR_UNSUPPORTED, "stream map is not a permutation");
                dst[at] = (int)((size_t)k * pl + (size_t)x);
            }
        // padding entries point at a zero adjoint slot of their own plane
        for (int q = 0; q < 9; ++q)
            for (size_t y = (size_t)P.nn; y < pl; ++y) dst[(size_t)q * pl + y] = (int)((size_t)q * pl + y);
        HIPCHECK(h, hipMemcpy(dstmap.p, dst.data(), 9 * pl * 4, hipMemcpyHostToDevice));
    }

    // forward pass with checkpoints; the loss from the history rows
    std::vector<Checkpoint> cks(nck + 1);
    const size_t rowp = (size_t)HPC + 2;
    std::vector<double> hist((size_t)chk * rowp);
    double lsum = 0.0;
    for (int c = 0; c < nck; ++c) {
        const int k0 = c * chk, n = std::min(chk, nsteps - k0);
        cks[c].off = (size_t)c * ck_doubles;
        if ((rc = save_state(h, ck + cks[c].off, cks[c]))) return rc;
        if ((rc = lbm_step(h, k0, n, hist.data()))) return rc;
        for (int s = 0; s < n; ++s) {
            const int k = k0 + s + 1;
            if (k > nsteps - window) lsum += hist[(size_t)s * rowp + LBM_H_VY];
        }
    }
    cks[nck].off = (size_t)nck * ck_doubles;
    if ((rc = save_state(h, ck + cks[nck].off, cks[nck]))) return rc;
    *loss = lsum / window;

    // backward pass
    adj::LocalP L;
    L.rheology = P.rheology; L.q_mode = P.q_mode;
    L.s0 = P.s0; L.s1 = P.s1; L.s2 = P.s2; L.s3 = P.s3; L.s5 = P.s5; L.sq = P.sq;
    L.mu0 = P.mu0; L.tau_y = P.tau_y; L.m_pap = P.m_pap; L.gamma_min = P.gamma_min;
    L.d_mrt = D_MRT;
    for (int b = 0; b < 2; ++b) {
        HIPCHECK(h, hipMemsetAsync(lf[b].p, 0, 9 * pl * 8, h->stream));
        HIPCHECK(h, hipMemsetAsync(lw[b].p, 0, pl * 8, h->stream));
        HIPCHECK(h, hipMemsetAsync(mu[b].p, 0, ANP * 8, h->stream));
    }
    HIPCHECK(h, hipMemsetAsync(theta.p, 0, 3 * 8, h->stream));
    {
        // the loss term of step N seeds the particle adjoint
        double seed[ANP] = {0, 0, 0, 1.0 / window, 0, 0, 0, 0};
        HIPCHECK(h, hipMemcpy(mu[0].p, seed, sizeof(seed), hipMemcpyHostToDevice));
    }
    int cur = 0;  // buffers [cur] hold the adjoint of the later state
    std::vector<double> fcol(2 * P.maxp);
    for (int c = nck - 1; c >= 0; --c) {
        const int k0 = c * chk, k1 = std::min(k0 + chk, nsteps);
        if ((rc = restore_state(h, ck + cks[c].off, cks[c]))) return rc;
        for (int k = k0; k < k1; ++k) {
            if (k > k0 && (rc = lbm_step(h, k - 1, 1, nullptr))) return rc;
            if ((rc = ensure_canonical(h))) return rc;
            double *sn = seg + (size_t)(k - k0) * snap_doubles;
            HIPCHECK(h, hipMemcpyAsync(sn, h->f, 9 * pl * 8, hipMemcpyDeviceToDevice, h->stream));
            HIPCHECK(h, hipMemcpyAsync(sn + 9 * pl, h->obn, pl * 8, hipMemcpyDeviceToDevice, h->stream));
            HIPCHECK(h, hipMemcpyAsync(sn + 10 * pl, h->pbuf[h->pcur], pc * 8, hipMemcpyDeviceToDevice,
                                       h->stream));
            // the contact force is not differentiated: it must stay zero
            HIPCHECK(h, hipMemcpyAsync(fcol.data(), h->pbuf[h->pcur] + (size_t)PFCX * P.maxp,
                                       2 * P.maxp * 8, hipMemcpyDeviceToHost, h->stream));
            HIPCHECK(h, hipStreamSynchronize(h->stream));
            if (fcol[0] != 0.0 || fcol[P.maxp] != 0.0)
                return fail(h, LBM_ERR_UNSUPPORTED,
                            "the particle came within contact range; the adjoint treats the "
                            "contact force as absent");
        }
        for (int k = k1 - 1; k >= k0; --k) {
            const double *sn = seg + (size_t)(k - k0) * snap_doubles;
            const double *rec = sn + 10 * pl;
            const int nx = cur ^ 1;
            hipLaunchKernelGGL(k_adj_particle, dim3(1), dim3(64), 0, h->stream, P, rec,
                               (const double *)mu[cur].p, (double *)mu[nx].p, (double *)fbar.p,
                               (double *)theta.p);
            hipLaunchKernelGGL(k_adj_node, dim3(nb), dim3(NT), 0, h->stream, P, L,
                               (const double *)lf[cur].p, (const double *)lw[cur].p,
                               (const int *)dstmap.p, sn, sn + 9 * pl, rec,
                               (const double *)fbar.p, (double *)lf[nx].p, (double *)lw[nx].p,
                               (double *)partial.p);
            const double lw_k = (k >= 1 && k > nsteps - window) ? 1.0 / window : 0.0;
            hipLaunchKernelGGL(k_adj_reduce, dim3(1), dim3(NT), 0, h->stream,
                               (const double *)partial.p, (int)nb, (double *)mu[nx].p,
                               (double *)theta.p, lw_k);
            if ((rc = check_launch(h, "adjoint step"))) return rc;
            cur = nx;
        }
    }
    if (P.rheology == LBM_BINGHAM)
        hipLaunchKernelGGL(k_adj_omega0, dim3(1), dim3(NT), 0, h->stream, P,
                           (const double *)lw[cur].p, (double *)theta.p);
    HIPCHECK(h, hipMemcpyAsync(grad, theta.p, 3 * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    // leave the handle in the state after nsteps
    if ((rc = restore_state(h, ck + cks[nck].off, cks[nck]))) return rc;
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    if (mem_bytes) *mem_bytes = (double)total;
    return LBM_OK;
}
