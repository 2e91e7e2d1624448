// api.hip -- C ABI of the solver (declared in include/mcg.h): device memory,
// launch sequencing and state exchange.  One translation unit; the kernels
// come from kernels.hpp (two-pass step, service kernels) and fused.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <functional>
#include <mutex>
#include <vector>

#include "../../include/mcg.h"
#include "adjoint.hpp"
#include "device.hpp"
#include "fused.hpp"
#include "kernels.hpp"
#include "probe.hpp"

using namespace mcg;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail(MCG_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

#define LAUNCH_CHECK(name)                                                                            \
    do {                                                                                              \
        hipError_t e_ = hipGetLastError();                                                            \
        if (e_ != hipSuccess) return fail(MCG_E_HIP, std::string("launch ") + name + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kInitJet = 0, kInitBubble = 1, kInitBubbleTanh = 2, kInitSpheres = 3;

// launch KERNEL<storage> for the run-time storage mode
#define MCG_DISPATCH_ST(st, KERNEL, ...)                                                 \
    do {                                                                                 \
        switch (st) {                                                                    \
            case MCG_STORE_FP16: hipLaunchKernelGGL((KERNEL<MCG_STORE_FP16>), __VA_ARGS__); break; \
            case MCG_STORE_BF16: hipLaunchKernelGGL((KERNEL<MCG_STORE_BF16>), __VA_ARGS__); break; \
            default: hipLaunchKernelGGL((KERNEL<MCG_STORE_FP32>), __VA_ARGS__); break;   \
        }                                                                                \
    } while (0)

// Partial sums of mcg_global_sums_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends.  Its address is resolved once per
// device (in mcg_create, with that device current) and kept in
// g_scratch_ptr[device]; host code passes that pointer, never the array's
// name, which on the host is only a stand-in.  Every call on a GPU shares the
// array: calls that overlap on one GPU race.  Each call overwrites every
// partial it reads, so nothing needs clearing between calls.
constexpr size_t kScratchDoubles = 8192;  // 64 KiB: kRedQ partials of up to 1170 workgroups
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// resolve this device's copy of the scratch array (the current device)
int scratch_resolve(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(MCG_E_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return MCG_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(MCG_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double *>(ptr);
    return MCG_OK;
}

}  // namespace

// Timed choice among launch shapes of the step kernel: `passes` passes over
// every candidate, then kRefinePasses more over the kRefineTop fastest (the
// best shapes of a grid lie within about 1% of each other, a margin two
// samples do not resolve), keeping the fastest time seen.  Each timed launch
// runs between two HIP events and is waited for.
struct ShapeTuner {
    static constexpr int kRefineTop = 4, kRefinePasses = 3;
    std::vector<mcg::FusedGrid> cand;
    int passes = 2;
    std::vector<float> ms;
    std::vector<size_t> top;
    size_t next = 0;
    bool on = true, done = false;
    mcg::FusedGrid chosen{};
    bool active() const { return on && !done && cand.size() > 1; }
    void reset(bool enable) {
        ms.clear();
        top.clear();
        next = 0;
        done = false;
        on = enable;
    }
    size_t first() const { return cand.size() * size_t(passes); }
    size_t total() const { return first() + std::min(cand.size(), size_t(kRefineTop)) * size_t(kRefinePasses); }
    // the candidate of the next timed launch
    size_t pick() {
        const size_t n = cand.size();
        if (ms.size() != n) ms.assign(n, 1e30f);
        if (next == first()) {
            std::vector<size_t> order(n);
            for (size_t i = 0; i < n; ++i) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ms[a] < ms[b]; });
            order.resize(std::min(n, size_t(kRefineTop)));
            top = order;
        }
        return next < first() ? next % n : top[(next - first()) % top.size()];
    }
    void record(size_t k, float t) {
        ms[k] = std::min(ms[k], t);
        if (++next >= total()) {
            chosen = cand[size_t(std::min_element(ms.begin(), ms.end()) - ms.begin())];
            done = true;
        }
    }
};

struct mcg_solver {
    mcg_case c{};
    DevParams P{};
    int storage = MCG_STORE_FP32;
    int kernel = MCG_KERNEL_FUSED;
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cu = 256;
    size_t fused_lds = 65536;  // LDS a fused-kernel workgroup may use (statistics histogram + psi ring)

    void *f[2] = {nullptr, nullptr};
    float *bin[2] = {nullptr, nullptr};
    float *bout[2] = {nullptr, nullptr};
    float *gB[2] = {nullptr, nullptr};
    float *grr = nullptr, *grb = nullptr;
    float *mom = nullptr;
    double *stats = nullptr;
    unsigned long long *stats_fx = nullptr;  // one step's sums in fixed point, zero between steps
    unsigned long long *count = nullptr;
    double *red_partials = nullptr;
    double *red_out = nullptr;
    int red_blocks = 0;
    float *staging = nullptr;
    size_t staging_bytes = 0;

    // launch shape of the fused step: the static choice, then the fastest of the
    // candidate shapes timed on the first plain steps (uniform z segmentations
    // and banded schedules).  Statistics steps run another instance with its
    // own best shapes (8-row tiles, the histogram in LDS) and time uniform
    // segmentations of their own on the first statistics steps; until then
    // they take the plain steps' shape.
    FusedGrid grid{};
    ShapeTuner tune_plain, tune_stats;
    FusedUnit *units = nullptr;  // unit tables of the banded candidates (one allocation)
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
    // the shape a statistics step launches with
    FusedGrid stats_grid() const { return tune_stats.done ? tune_stats.chosen : fused_grid_for_instance(P, grid, kStatsTY); }

    int cur = 0;
    int64_t step = 0;
    bool initialized = false;
    bool moments_valid = false;
    bool inlet_reset_pending = false;
    bool started = false;
    int64_t start_step = 0, samples = 0;
    InitGeo geo{};
    int init_kind = -1;
    size_t bytes = 0;

    // adjoint: checkpoints of the recorded forward, the states of one segment,
    // the adjoint fields and the FP64 parameter sums
    bool stats_frozen = false;  // recomputation must not touch the statistics
    float *rec_mom = nullptr;   // if set, one_step stores the moments of its source state here
    float *inlet_du = nullptr;  // jet cases: inlet velocity offset per column (3 x NX*NZ) or null
    // the time-windowed u_y functional sum_{n=first..last} sum_x w(x) u_y(x, n)
    // (FP64 accumulator and workgroup partials), accumulated by ordinary steps
    double *obs_acc = nullptr, *obs_part = nullptr;
    bool adj_recorded = false;
    int adj_nsteps = 0, adj_every = 0;
    int64_t adj_step0 = 0;
    // recorded states: adj_ckpt holds the state buffers ("slots"); adj_seg_slot
    // maps segment k (steps k*every ..) to the slot holding its first state, or
    // -1.  adj_limit caps the slots (0: one per segment); below one per segment
    // the sweep recomputes the missing states (binomial schedule).
    std::vector<void *> adj_ckpt;
    std::vector<int> adj_seg_slot;
    int adj_limit = 0;
    int64_t adj_recomputed = 0;  // forward steps recomputed by the last backward sweep
    void *adj_pp[2] = {nullptr, nullptr};  // alternating states of the segment recomputation
    float *adj_mom = nullptr;  // moments of the states of one segment
    float *adj_gbs = nullptr;  // jet cases: outlet states of the states of one segment
    float *adj_lam[2] = {nullptr, nullptr};
    float *adj_lm = nullptr, *adj_lF = nullptr, *adj_ps = nullptr;
    double *adj_acc = nullptr, *adj_part = nullptr;
    double adj_ms = 0.0;  // device time of the last backward sweep (recomputation + adjoint)
    // jet cases: adjoint of the outlet state, dJ/d(inlet velocity) per column,
    // the u_y seed and the steps it applies to
    float *adj_lg[2] = {nullptr, nullptr};
    double *adj_din = nullptr;
    float *adj_psf = nullptr;  // psi over the full grid
    float *adj_wuy = nullptr;
    int64_t adj_uy_first = 0, adj_uy_last = -1;  // solver step numbers
    double adj_inlet[2] = {0.0, 0.0};  // dJ/djet_velocity, dJ/dperturbation_amp of the last sweep
    double adj_grad[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // dJ/d{A, beta, gravity, jet_velocity, perturbation_amp}

    void adj_release() {
        for (void *p : adj_ckpt) (void)hipFree(p);
        adj_ckpt.clear();
        adj_seg_slot.clear();
        for (void *p : {adj_pp[0], adj_pp[1], (void *)adj_mom, (void *)adj_gbs, (void *)adj_lam[0], (void *)adj_lam[1], (void *)adj_lm, (void *)adj_lF,
                        (void *)adj_ps, (void *)adj_acc, (void *)adj_part, (void *)adj_lg[0], (void *)adj_lg[1],
                        (void *)adj_din, (void *)adj_psf})
            if (p) (void)hipFree(p);
        adj_mom = adj_gbs = nullptr;
        adj_pp[0] = adj_pp[1] = nullptr;
        adj_lam[0] = adj_lam[1] = nullptr;
        adj_lm = adj_lF = adj_ps = nullptr;
        adj_acc = adj_part = nullptr;
        adj_lg[0] = adj_lg[1] = nullptr;
        adj_din = nullptr;
        adj_psf = nullptr;
        adj_recorded = false;
    }
    // the u_y functional (seed of the adjoint, observed by ordinary steps)
    void uy_release() {
        for (void *p : {(void *)adj_wuy, (void *)obs_acc, (void *)obs_part})
            if (p) (void)hipFree(p);
        adj_wuy = nullptr;
        obs_acc = obs_part = nullptr;
        adj_uy_first = 0;
        adj_uy_last = -1;
    }

    size_t N() const { return size_t(c.NX) * c.NY * c.NZ; }
    size_t nxz() const { return size_t(c.NX) * c.NZ; }
    size_t pop_bytes() const { return size_t(P.rowstride) * size_t(P.IY) * size_t(P.IZ) * sizeof(float); }
    size_t bnd_bytes() const { return size_t(18) * nxz() * sizeof(float); }
    size_t gB_bytes() const { return size_t(5) * nxz() * sizeof(float); }
    size_t mom_set_bytes() const { return al256(size_t(11) * P.plane * sizeof(float)); }
    size_t gB_set_bytes() const { return al256(gB_bytes()); }
    // one recorded forward state: the populations and, for jet cases, the
    // inlet and outlet source arrays and the outlet state (256-byte aligned parts)
    static size_t al256(size_t b) { return (b + 255) & ~size_t(255); }
    bool jet() const { return c.y_mode == MCG_Y_JET; }
    size_t state_bytes() const { return al256(pop_bytes()) + (jet() ? 2 * al256(bnd_bytes()) + al256(gB_bytes()) : 0); }
    float *st_bin(void *st) const { return (float *)((char *)st + al256(pop_bytes())); }
    float *st_bout(void *st) const { return (float *)((char *)st + al256(pop_bytes()) + al256(bnd_bytes())); }
    float *st_gB(void *st) const { return (float *)((char *)st + al256(pop_bytes()) + 2 * al256(bnd_bytes())); }
    size_t bins() const { return size_t(c.NY) * P.NR; }

    dim3 grid_interior() const { return dim3(unsigned(P.PX / 64), unsigned((P.IY + 3) / 4), unsigned(P.IZ)); }
    dim3 grid_full() const { return dim3(unsigned((c.NX + 63) / 64), unsigned((c.NY + 3) / 4), unsigned(c.NZ)); }

    DevBufs bufs(bool do_stats) const {
        DevBufs B;
        B.f_src = f[cur];
        B.f_dst = f[cur ^ 1];
        B.in_src = bin[cur];
        B.in_dst = bin[cur ^ 1];
        B.out_src = bout[cur];
        B.out_dst = bout[cur ^ 1];
        B.gB_src = gB[cur];
        B.gB_dst = gB[cur ^ 1];
        B.ghost_rr = grr;
        B.ghost_rb = grb;
        B.mom = mom;
        B.mom_out = rec_mom;
        B.inlet_du = inlet_du;
        B.stats_fx = stats_fx;
        B.count = count;
        B.do_stats = do_stats ? 1 : 0;
        B.stats_ring = fused_stats_ring(P, fused_lds) ? 1 : 0;
        B.step = (unsigned long long)(step);
        return B;
    }

    int ensure_staging(size_t nb) {
        if (staging_bytes >= nb) return MCG_OK;
        if (staging) (void)hipFree(staging);
        staging = nullptr;
        staging_bytes = 0;
        if (hipMalloc(&staging, nb) != hipSuccess) return fail(MCG_E_NOMEM, "staging allocation failed");
        staging_bytes = nb;
        return MCG_OK;
    }
    void release_staging() {
        if (staging) (void)hipFree(staging);
        staging = nullptr;
        staging_bytes = 0;
    }
};

namespace {

// derived constants of constants.cuh:59-76 for a case
// words of padding after each (y,z) row block of the populations: concurrent
// workgroups then start their rows at addresses spread over the memory channels
// (64 words: +7% FP32 MLUPS at 256x512x256)
#ifndef MCG_ROW_PAD
#define MCG_ROW_PAD 64
#endif
constexpr uint32_t kRowPad = MCG_ROW_PAD;

void fill_params(const mcg_case &c, int storage, DevParams &P) {
    P.NX = c.NX;
    P.NY = c.NY;
    P.NZ = c.NZ;
    P.PX = ((c.NX - 2) + 63) / 64 * 64;
    P.IY = c.NY - 2;
    P.IZ = c.NZ - 2;
    P.plane = uint32_t(size_t(P.PX) * P.IY * P.IZ);
    P.rowstride = uint32_t(P.PX * (storage == MCG_STORE_FP32 ? 2 * Q : Q) + kRowPad);
    P.planeel = P.rowstride * uint32_t(P.IY);
    P.y_mode = c.y_mode;
    P.wrap_gradient = c.wrap_gradient;
    P.visc_mode = c.viscosity_mode;
    P.sponge_yend = c.NY - 2;
    P.sponge_ystart = P.sponge_yend - c.sponge_cells + 1;
    const double cs2 = 1.0 / 3.0;
    P.nu = float((c.taub - 0.5) * cs2);
    P.sponge_K = float(c.sponge_K);
    P.taur = float(c.taur);
    P.taub = float(c.taub);
    P.omega_b = 1.0f / float(c.taub);
    P.A_fixed = c.A_override > 0 ? float(c.A_override) : 0.0f;
    P.sig4 = float(c.sigma) / (4.0f * mcg::cs4);
    P.beta = float(c.beta_recolor);
    P.jet_x0 = float(c.jet_x0);
    P.jet_z0 = float(c.jet_z0);
    P.jet_R = float(c.jet_radius);
    P.jet_U = float(c.jet_velocity);
    P.rhor0 = float(c.rhor0);
    P.rhob0 = float(c.rhob0);
    P.NR = c.NX / 2;
    P.storage = storage;
    P.inlet_profile = c.inlet_profile;
    P.inlet_w = float(c.interface_width);
    P.pert_amp = float(c.perturbation_amp);
    P.pert_seed = uint32_t(c.perturbation_seed);
    P.wall = c.wall_xz;
    P.ghalf = float(0.5 * c.gravity);
    P.grav = float(c.gravity);
}

int check_solver(const mcg_solver *s) {
    if (!s) return fail(MCG_E_INVALID, "null solver");
    return MCG_OK;
}

int check_ready(const mcg_solver *s) {
    if (!s) return fail(MCG_E_INVALID, "null solver");
    if (!s->initialized) return fail(MCG_E_STATE, "solver not initialised (call mcg_init_jet / mcg_init_bubble first)");
    return MCG_OK;
}

int launch_moments(mcg_solver *s, int which) {
    MCG_DISPATCH_ST(s->storage, k_moments, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, s->f[which], s->bin[which],
                    s->bout[which], s->mom);
    LAUNCH_CHECK("k_moments");
    return MCG_OK;
}

// interior moments of the last step (the fused kernel does not store them)
int ensure_moments(mcg_solver *s) {
    if (s->moments_valid) return MCG_OK;
    const int rc = launch_moments(s, s->cur ^ 1);
    if (rc) return rc;
    s->moments_valid = true;
    return MCG_OK;
}

int run_reduce(mcg_solver *s, double out[kRedQ]) {
    const size_t total = size_t(s->c.NX - 2) * (s->c.NY - 2) * (s->c.NZ - 2);
    const int npb = int((total + s->red_blocks - 1) / s->red_blocks);
    hipLaunchKernelGGL(k_reduce_partials, dim3(s->red_blocks), dim3(256), 0, s->stream, s->P, s->mom, s->red_partials, npb);
    LAUNCH_CHECK("k_reduce_partials");
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s->stream, s->red_partials, s->red_blocks, s->red_out);
    LAUNCH_CHECK("k_reduce_final");
    HIP_TRY(hipMemcpyAsync(out, s->red_out, sizeof(double) * kRedQ, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCG_OK;
}

int do_init(mcg_solver *s, int kind) {
    if (!s) return fail(MCG_E_INVALID, "null solver");
    const mcg_case &c = s->c;
    s->geo.g[0] = float(c.bubble_x0);
    s->geo.g[1] = float(c.bubble_y0);
    s->geo.g[2] = float(c.bubble_z0);
    s->geo.g[3] = float(c.bubble_radius);
    s->geo.g[4] = float(c.interface_width);
    if ((kind == kInitBubbleTanh || kind == kInitSpheres) && !(c.interface_width > 0))
        return fail(MCG_E_INVALID, "interface_width must be > 0 for the tanh bubble and spheres");
    s->cur = 0;
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipMemsetAsync(s->bin[b], 0, s->bnd_bytes(), s->stream));
        HIP_TRY(hipMemsetAsync(s->bout[b], 0, s->bnd_bytes(), s->stream));
    }
    HIP_TRY(hipMemsetAsync(s->mom, 0, size_t(11) * s->P.plane * sizeof(float), s->stream));
    MCG_DISPATCH_ST(s->storage, k_init_pull, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, kind, s->geo, s->f[0], s->bin[0],
                       s->bout[0]);
    LAUNCH_CHECK("k_init_pull");
    hipLaunchKernelGGL(k_init_ghost, s->grid_full(), dim3(64, 4, 1), 0, s->stream, s->P, kind, s->geo, s->grr, s->grb,
                       s->gB[0], s->gB[1]);
    LAUNCH_CHECK("k_init_ghost");
    HIP_TRY(hipMemcpyAsync(s->bin[1], s->bin[0], s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->bout[1], s->bout[0], s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->stats, 0, sizeof(double) * 7 * s->bins(), s->stream));
    HIP_TRY(hipMemsetAsync(s->stats_fx, 0, sizeof(unsigned long long) * 7 * s->bins(), s->stream));
    HIP_TRY(hipMemsetAsync(s->count, 0, sizeof(unsigned long long) * s->bins(), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->step = 0;
    s->init_kind = kind;
    s->initialized = true;
    s->moments_valid = true;  // all-zero interior moments before the first step (memory.cuh:85-94)
    s->inlet_reset_pending = (c.y_mode == MCG_Y_JET);
    s->started = false;
    s->start_step = 0;
    s->samples = 0;
    return MCG_OK;
}

// One timed launch of the tuner's next candidate.
int timed_launch(mcg_solver *s, const DevBufs &B, ShapeTuner &T) {
    for (hipEvent_t &e : s->tune_ev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    const size_t k = T.pick();
    HIP_TRY(hipEventRecord(s->tune_ev[0], s->stream));
    const int rc = launch_fused(s->P, B, s->stream, T.cand[k]);
    if (rc != 0) return fail(MCG_E_HIP, std::string("fused step launch: ") + hipGetErrorString(hipError_t(rc)));
    HIP_TRY(hipEventRecord(s->tune_ev[1], s->stream));
    HIP_TRY(hipEventSynchronize(s->tune_ev[1]));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, s->tune_ev[0], s->tune_ev[1]));
    T.record(k, ms);
    return MCG_OK;
}

// One fused step.  Until their launch shapes are chosen, plain steps (no
// statistics, no recorded moments) and statistics steps time their
// candidates in turn (ShapeTuner); every other step uses the current choice.
int launch_step(mcg_solver *s, const DevBufs &B) {
    if (B.do_stats && s->tune_stats.active()) return timed_launch(s, B, s->tune_stats);
    if (!B.do_stats && !B.mom_out && s->tune_plain.active()) {
        const int rc = timed_launch(s, B, s->tune_plain);
        if (s->tune_plain.done) s->grid = s->tune_plain.chosen;
        return rc;
    }
    const int rc = launch_fused(s->P, B, s->stream, B.do_stats ? s->stats_grid() : s->grid);
    if (rc != 0) return fail(MCG_E_HIP, std::string("fused step launch: ") + hipGetErrorString(hipError_t(rc)));
    return MCG_OK;
}

int one_step(mcg_solver *s) {
    // two-pass kernel: the moments pass itself writes the recorded moments
    struct MomSwap {
        mcg_solver *s;
        float *keep;
        ~MomSwap() {
            if (keep) s->mom = keep;
        }
    } swap_guard{s, nullptr};
    if (s->rec_mom && s->kernel == MCG_KERNEL_TWOPASS) {
        swap_guard.keep = s->mom;
        s->mom = s->rec_mom;
    }
    // update_radial_statistics_start, meanF/meanFieldsRT.cuh:173-189
    if (!s->stats_frozen && !s->started && s->step >= s->c.NSTATS_START_STEP) {
        s->started = true;
        s->start_step = s->step;
        s->samples = 0;
    }
    const int64_t every = s->c.NSTATS_SAMPLE > 0 ? s->c.NSTATS_SAMPLE : 1;
    const bool do_stats = !s->stats_frozen && s->started && ((s->step - s->start_step) % every == 0);
    // the u_y functional observes the moments pass of this step
    const bool observe = s->adj_wuy && !s->stats_frozen && s->step >= s->adj_uy_first && s->step <= s->adj_uy_last;
    const DevBufs B = s->bufs(do_stats);
    if (s->kernel == MCG_KERNEL_TWOPASS) {
        int rc = launch_moments(s, s->cur);
        if (rc) return rc;
        if (do_stats) {
            const dim3 g(unsigned(s->P.IY), unsigned((s->P.IZ + kRadialZChunk - 1) / kRadialZChunk));
            hipLaunchKernelGGL(k_radial, g, dim3(256), size_t(8) * s->P.NR * sizeof(unsigned long long), s->stream, s->P,
                               s->mom, s->stats_fx, s->count);
            LAUNCH_CHECK("k_radial");
        }
        MCG_DISPATCH_ST(s->storage, k_collide, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, B);
        LAUNCH_CHECK("k_collide");
        s->moments_valid = true;
    } else {
        DevBufs Bf = B;
        bool have_moments = false;
        if (do_stats && (observe || !fused_has_stats(s->P, s->fused_lds))) {
            // statistics from the moments of this step, LDS-privatised histogram
            int rc = launch_moments(s, s->cur);
            if (rc) return rc;
            const dim3 g(unsigned(s->P.IY), unsigned((s->P.IZ + kRadialZChunk - 1) / kRadialZChunk));
            hipLaunchKernelGGL(k_radial, g, dim3(256), size_t(8) * s->P.NR * sizeof(unsigned long long), s->stream, s->P,
                               s->mom, s->stats_fx, s->count);
            LAUNCH_CHECK("k_radial");
            Bf.do_stats = 0;
            have_moments = true;
        } else if (observe) {
            Bf.mom_out = s->mom;  // the step stores the moments of its source state
            have_moments = true;
        }
        if (int rc = launch_step(s, Bf)) return rc;
        s->moments_valid = have_moments;
    }
    if (observe) {
        const dim3 g = s->grid_interior();
        hipLaunchKernelGGL(k_uy_partial, g, dim3(64, 4, 1), 0, s->stream, s->P, s->mom, s->adj_wuy, s->obs_part);
        LAUNCH_CHECK("k_uy_partial");
        hipLaunchKernelGGL(k_adj_sum<2>, dim3(1), dim3(kSumThreads), 0, s->stream, s->obs_part, size_t(g.x) * g.y * g.z,
                           s->obs_acc);
        LAUNCH_CHECK("k_adj_sum");
    }
    if (do_stats) {
        // this step's fixed-point sums into the FP64 sums, in a fixed order
        hipLaunchKernelGGL(k_stats_fold, dim3(unsigned((s->bins() + 255) / 256)), dim3(256), 0, s->stream, s->bins(),
                           s->stats_fx, s->stats);
        LAUNCH_CHECK("k_stats_fold");
        ++s->samples;  // never during adjoint recomputation (stats_frozen)
    }
    if (s->inlet_reset_pending) {
        hipLaunchKernelGGL(k_inlet_ghost_reset, dim3(unsigned((s->c.NX - 2 + 255) / 256), unsigned(s->c.NZ - 2)), dim3(256), 0,
                           s->stream, s->P, s->grr, s->grb);
        LAUNCH_CHECK("k_inlet_ghost_reset");
        s->inlet_reset_pending = false;
    }
    s->cur ^= 1;
    ++s->step;
    return MCG_OK;
}

}  // namespace

extern "C" {

const char *mcg_last_error(void) { return g_err.c_str(); }
const char *mcg_version(void) { return "mcg-mi355x 0.1.0 (gfx950)"; }

int mcg_case_reference_jet(mcg_case *c) {
    if (!c) return fail(MCG_E_INVALID, "null case");
    std::memset(c, 0, sizeof(*c));
    // constants.cuh:8-76
    c->NX = 128;
    c->NZ = 128;
    c->NY = 400;
    c->NSTEP = 200000;
    c->NOUTPUT = 2000;
    c->NSTATS_START_STEP = 6000;
    c->NSTATS_SAMPLE = 1;
    c->write_vti_output = 1;
    c->sponge_cells = 34;
    c->y_mode = MCG_Y_JET;
    c->wrap_gradient = 0;
    c->viscosity_mode = MCG_VISC_SINGLE;
    c->inlet_profile = MCG_INLET_SHARP;
    c->jet_radius = 7.0;
    c->jet_x0 = double(c->NX - 1) / 2.0;
    c->jet_z0 = double(c->NZ - 1) / 2.0;
    c->jet_velocity = 0.05;
    c->bubble_radius = 10.0;
    c->bubble_x0 = double(c->NX - 1) / 2.0;
    c->bubble_y0 = double(c->NY - 1) / 2.0;
    c->bubble_z0 = double(c->NZ - 1) / 2.0;
    c->rhor0 = 1.0;
    c->rhob0 = 1.0;
    c->Re = 5000.0;
    const double nu = (2.0 * c->jet_radius * c->jet_velocity) / c->Re;
    c->taur = 0.5 + nu / (1.0 / 3.0);
    c->taub = 0.5 + nu / (1.0 / 3.0);
    c->We = 2000.0;
    c->sigma = (c->rhob0 * c->jet_velocity * c->jet_velocity * 2.0 * c->jet_radius) / c->We;
    c->beta_recolor = 0.80;
    c->A_override = 0.0;
    c->sponge_K = 0.0;
    c->interface_width = 2.0;
    c->perturbation_amp = 0.0;
    return MCG_OK;
}

size_t mcg_case_size(void) { return sizeof(mcg_case); }

int mcg_case_validate(const mcg_case *c) {
    if (!c) return fail(MCG_E_INVALID, "null case");
    if (c->NX < 4 || c->NY < 3 || c->NZ < 4) return fail(MCG_E_INVALID, "grid too small (NX,NZ >= 4, NY >= 3)");
    // (also keeps a radial bin under about 8 NX = 16384 samples per step, far inside
    // the 2^20 the fixed-point statistics sums hold)
    if (c->NX > 2048) return fail(MCG_E_INVALID, "NX > 2048 (radial bins NX/2 must fit the LDS histogram)");
    const size_t px = size_t((c->NX - 2 + 63) / 64 * 64);
    // population addresses are a z-plane base plus a 32-bit byte offset within
    // the plane, moments and node-layout fields 32-bit element offsets
    if ((px * 54 + kRowPad) * size_t(c->NY - 2) * sizeof(float) >= (size_t(1) << 32))
        return fail(MCG_E_INVALID, "grid too large: one z-plane of populations must stay below 4 GiB");
    if (px * size_t(c->NY - 2) * size_t(c->NZ - 2) >= (size_t(1) << 32))
        return fail(MCG_E_INVALID, "grid too large: more than 2^32 padded interior nodes");
    if (c->y_mode != MCG_Y_JET && c->y_mode != MCG_Y_PERIODIC) return fail(MCG_E_INVALID, "unknown y_mode");
    if (c->viscosity_mode != MCG_VISC_SINGLE && c->viscosity_mode != MCG_VISC_INTERFACE)
        return fail(MCG_E_INVALID, "unknown viscosity_mode");
    if (c->inlet_profile != MCG_INLET_SHARP && c->inlet_profile != MCG_INLET_TANH)
        return fail(MCG_E_INVALID, "unknown inlet_profile");
    if (c->inlet_profile == MCG_INLET_TANH && !(c->interface_width > 0))
        return fail(MCG_E_INVALID, "interface_width must be > 0 for the tanh inlet");
    if (!(c->perturbation_amp >= 0.0)) return fail(MCG_E_INVALID, "perturbation_amp must be >= 0");
    if (!(c->taub > 0.5) || !(c->taur > 0.5)) return fail(MCG_E_INVALID, "relaxation times must exceed 0.5");
    if (c->sponge_cells < 2 && c->sponge_K != 0.0) return fail(MCG_E_INVALID, "sponge_cells must be >= 2");
    if (c->rhor0 <= 0 || c->rhob0 <= 0) return fail(MCG_E_INVALID, "reference densities must be > 0");
    if (c->wall_xz != 0 && c->wall_xz != 1) return fail(MCG_E_INVALID, "wall_xz must be 0 or 1");
    if (c->wall_xz && c->y_mode != MCG_Y_PERIODIC) return fail(MCG_E_UNSUPPORTED, "x/z walls need y_mode periodic");
    if (!std::isfinite(c->gravity)) return fail(MCG_E_INVALID, "gravity must be finite");
    return MCG_OK;
}

int mcg_create(const mcg_case *c, int storage, void *stream, mcg_solver **out) {
    if (!out) return fail(MCG_E_INVALID, "null output pointer");
    *out = nullptr;
    int rc = mcg_case_validate(c);
    if (rc) return rc;
    if (storage != MCG_STORE_FP32 && storage != MCG_STORE_FP16 && storage != MCG_STORE_BF16)
        return fail(MCG_E_INVALID, "unknown storage mode");
    mcg_solver *s = new (std::nothrow) mcg_solver();
    if (!s) return fail(MCG_E_NOMEM, "host allocation failed");
    s->c = *c;
    s->storage = storage;
    s->stream = static_cast<hipStream_t>(stream);
    fill_params(*c, storage, s->P);
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return fail(MCG_E_HIP, "hipGetDevice failed (no GPU?)");
    }
    if (hipDeviceGetAttribute(&s->num_cu, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess || s->num_cu <= 0)
        s->num_cu = 256;
    if (int rc = scratch_resolve(s->device)) {
        delete s;
        return rc;
    }
    s->grid = fused_grid(s->P, s->num_cu);
    s->tune_plain.cand = fused_candidates(s->P);
    // statistics steps: uniform segmentations of 8-row tiles (at 256x512x256
    // the FP32 statistics step took 3.44 ms with 13 segments, 3.55 with 4 and
    // 3.77 with the 2-band schedule the plain step kept; one pass each, as
    // they come one step in NSTATS_SAMPLE)
    s->tune_stats.cand = fused_candidates_ty(s->P, kStatsTY, 32);
    s->tune_stats.passes = 1;
    {
        // the banded schedules join the timed candidates
        const std::vector<BandedSchedule> banded = banded_candidates(s->P, s->num_cu);
        size_t total = 0;
        for (const BandedSchedule &b : banded) total += b.units.size();
        if (total) {
            std::vector<FusedUnit> all;
            all.reserve(total);
            for (const BandedSchedule &b : banded) all.insert(all.end(), b.units.begin(), b.units.end());
            if (hipMalloc((void **)&s->units, total * sizeof(FusedUnit)) != hipSuccess ||
                hipMemcpy(s->units, all.data(), total * sizeof(FusedUnit), hipMemcpyHostToDevice) != hipSuccess) {
                if (s->units) (void)hipFree(s->units);
                delete s;
                return fail(MCG_E_NOMEM, "launch schedule tables: device allocation or upload failed");
            }
            size_t off = 0;
            for (const BandedSchedule &b : banded) {
                FusedGrid G = fused_grid_for(s->P, b.bands, fused_plain_ty(s->P.storage));
                G.nzs = b.bands;
                G.nwg = int(b.units.size());
                G.units = s->units + off;
                G.deal = int(b.deal);
                off += b.units.size();
                s->tune_plain.cand.push_back(G);
            }
        }
    }
    {
        // a workgroup's LDS above the default 64 KiB (the histogram beside the
        // psi and moment rings) needs the kernels' dynamic-LDS limit raised
        int lds = 0;
        const size_t need = fused_stats_lds(s->P) + fused_static_lds(false);
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) == hipSuccess &&
            size_t(lds) >= need) {
            int rc = 0;
            switch (storage) {
                case MCG_STORE_FP16: rc = fused_allow_lds<MCG_STORE_FP16>(fused_stats_lds(s->P)); break;
                case MCG_STORE_BF16: rc = fused_allow_lds<MCG_STORE_BF16>(fused_stats_lds(s->P)); break;
                default: rc = fused_allow_lds<MCG_STORE_FP32>(fused_stats_lds(s->P)); break;
            }
            if (rc == 0) s->fused_lds = size_t(lds);
            (void)hipGetLastError();
        }
    }
    auto alloc = [&](void **p, size_t nb) -> bool {
        if (hipMalloc(p, nb) != hipSuccess) return false;
        s->bytes += nb;
        return true;
    };
    bool ok = true;
    for (int b = 0; b < 2 && ok; ++b) {
        ok = ok && alloc((void **)&s->f[b], s->pop_bytes());
        ok = ok && alloc((void **)&s->bin[b], s->bnd_bytes());
        ok = ok && alloc((void **)&s->bout[b], s->bnd_bytes());
        ok = ok && alloc((void **)&s->gB[b], s->gB_bytes());
    }
    ok = ok && alloc((void **)&s->grr, s->N() * sizeof(float));
    ok = ok && alloc((void **)&s->grb, s->N() * sizeof(float));
    ok = ok && alloc((void **)&s->mom, size_t(11) * s->P.plane * sizeof(float));
    ok = ok && alloc((void **)&s->stats, sizeof(double) * 7 * s->bins());
    ok = ok && alloc((void **)&s->stats_fx, sizeof(unsigned long long) * 7 * s->bins());
    ok = ok && alloc((void **)&s->count, sizeof(unsigned long long) * s->bins());
    s->red_blocks = 1024;
    ok = ok && alloc((void **)&s->red_partials, sizeof(double) * kRedQ * s->red_blocks);
    ok = ok && alloc((void **)&s->red_out, sizeof(double) * kRedQ);
    if (!ok) {
        mcg_destroy(s);
        return fail(MCG_E_NOMEM, "device allocation failed");
    }
    // the fixed-point statistics sums are zero between steps from here on, also
    // for a solver that starts from a checkpoint rather than an initial state
    if (hipMemsetAsync(s->stats_fx, 0, sizeof(unsigned long long) * 7 * s->bins(), s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
        mcg_destroy(s);
        return fail(MCG_E_HIP, "hipMemset of the statistics sums failed");
    }
    *out = s;
    return MCG_OK;
}

int mcg_destroy(mcg_solver *s) {
    if (!s) return MCG_OK;
    for (int b = 0; b < 2; ++b) {
        for (void *p : {s->f[b], (void *)s->bin[b], (void *)s->bout[b], (void *)s->gB[b]})
            if (p) (void)hipFree(p);
    }
    for (void *p : {(void *)s->grr, (void *)s->grb, (void *)s->mom, (void *)s->stats, (void *)s->stats_fx,
                    (void *)s->count, (void *)s->red_partials, (void *)s->red_out})
        if (p) (void)hipFree(p);
    for (hipEvent_t e : s->tune_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->units) (void)hipFree(s->units);
    s->release_staging();
    s->adj_release();
    s->uy_release();
    if (s->inlet_du) (void)hipFree(s->inlet_du);
    delete s;
    return MCG_OK;
}

int mcg_set_kernel(mcg_solver *s, int kernel) {
    if (int rc = check_solver(s)) return rc;
    if (kernel != MCG_KERNEL_FUSED && kernel != MCG_KERNEL_TWOPASS) return fail(MCG_E_INVALID, "unknown kernel");
    s->kernel = kernel;
    return MCG_OK;
}

int mcg_set_launch_segments(mcg_solver *s, int segments) {
    if (int rc = check_solver(s)) return rc;
    if (segments < 0) {
        int k = 0;
        for (const FusedGrid &G : s->tune_plain.cand) {
            if (G.units && ++k == -segments) {
                s->tune_plain.reset(false);
                s->tune_stats.reset(false);
                s->grid = G;
                return MCG_OK;
            }
        }
        return fail(MCG_E_INVALID, "no banded launch schedule " + std::to_string(-segments) + " (" + std::to_string(k) +
                                       " for this grid)");
    }
    s->tune_plain.reset(segments == 0);
    s->tune_stats.reset(segments == 0);
    s->grid = fused_grid(s->P, s->num_cu);
    if (segments > 0) {
        // the uniform candidate with the segment count nearest to the one asked for
        const FusedGrid *best = &s->tune_plain.cand.front();
        for (const FusedGrid &G : s->tune_plain.cand)
            if (!G.units && std::abs(G.nzs - segments) < std::abs(best->nzs - segments)) best = &G;
        s->grid = *best;
    }
    return MCG_OK;
}

int mcg_launch_info(const mcg_solver *s, int32_t *segments, int32_t *workgroups, int32_t *tuned, int32_t *banded,
                    int32_t *shape) {
    if (int rc = check_solver(s)) return rc;
    if (segments) *segments = s->grid.nzs;
    if (workgroups) *workgroups = s->grid.nwg;
    if (tuned) *tuned = s->tune_plain.done ? 1 : 0;
    if (banded) *banded = s->grid.units ? 1 + s->grid.deal : 0;
    if (shape) {
        // the mcg_set_launch_segments argument that fixes this shape
        *shape = s->grid.nzs;
        int k = 0;
        for (const FusedGrid &G : s->tune_plain.cand)
            if (G.units && (++k, G.units == s->grid.units)) *shape = -k;
    }
    return MCG_OK;
}

int mcg_stats_launch_info(const mcg_solver *s, int32_t *segments, int32_t *workgroups, int32_t *tuned) {
    if (int rc = check_solver(s)) return rc;
    const FusedGrid G = s->stats_grid();
    if (segments) *segments = G.nzs;
    if (workgroups) *workgroups = G.nwg;
    if (tuned) *tuned = s->tune_stats.done ? 1 : 0;
    return MCG_OK;
}

int mcg_init_jet(mcg_solver *s) { return do_init(s, kInitJet); }
int mcg_init_bubble(mcg_solver *s) { return do_init(s, kInitBubble); }
int mcg_init_bubble_tanh(mcg_solver *s) { return do_init(s, kInitBubbleTanh); }

int mcg_sphere_list(uint32_t seed, int n, int NX, int NY, int NZ, double rmin, double rmax, double *out) {
    if (!out || n < 1 || NX < 3 || NY < 3 || NZ < 3 || !(rmin > 0) || rmax < rmin)
        return fail(MCG_E_INVALID, "sphere list: need n >= 1, a grid of at least 3^3 and 0 < rmin <= rmax");
    const uint64_t base = splitmix64_host(uint64_t(seed));
    auto u = [](uint64_t key, int j) { return double(splitmix64_host(key + uint64_t(j)) >> 40) / 16777216.0; };
    for (int k = 0; k < n; ++k) {
        const uint64_t key = splitmix64_host(base ^ uint64_t(k));
        out[4 * k + 0] = 1.0 + u(key, 1) * (NX - 3);
        out[4 * k + 1] = 1.0 + u(key, 2) * (NY - 3);
        out[4 * k + 2] = 1.0 + u(key, 3) * (NZ - 3);
        out[4 * k + 3] = rmin + u(key, 4) * (rmax - rmin);
    }
    return MCG_OK;
}

int mcg_init_spheres(mcg_solver *s, int n, uint32_t seed, double rmin, double rmax) {
    if (int rc = check_solver(s)) return rc;
    if (n > 65536) return fail(MCG_E_INVALID, "at most 65536 spheres");
    std::vector<double> d(4 * size_t(std::max(n, 1)));
    if (int rc = mcg_sphere_list(seed, n, s->c.NX, s->c.NY, s->c.NZ, rmin, rmax, d.data())) return rc;
    std::vector<float> h(d.begin(), d.end());
    float *dev = nullptr;
    HIP_TRY(hipMalloc((void **)&dev, h.size() * sizeof(float)));
    int rc = MCG_OK;
    if (hipMemcpy(dev, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        rc = fail(MCG_E_HIP, "sphere list upload failed");
    } else {
        s->geo.sph = dev;
        s->geo.nsph = n;
        rc = do_init(s, kInitSpheres);
    }
    s->geo.sph = nullptr;
    s->geo.nsph = 0;
    (void)hipFree(dev);
    return rc;
}

int mcg_step(mcg_solver *s, int nsteps) {
    if (int rc = check_ready(s)) return rc;
    if (nsteps < 0) return fail(MCG_E_INVALID, "nsteps < 0");
    for (int k = 0; k < nsteps; ++k) {
        const int rc = one_step(s);
        if (rc) return rc;
    }
    return MCG_OK;
}

int mcg_synchronize(mcg_solver *s) {
    if (int rc = check_solver(s)) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCG_OK;
}

int mcg_get_step(const mcg_solver *s, int64_t *step) {
    if (int rc = check_solver(s)) return rc;
    if (!step) return fail(MCG_E_INVALID, "null pointer");
    *step = s->step;
    return MCG_OK;
}

int mcg_stats_info(const mcg_solver *s, int32_t *started, int64_t *start_step, int64_t *samples) {
    if (int rc = check_solver(s)) return rc;
    if (started) *started = s->started ? 1 : 0;
    if (start_step) *start_step = s->start_step;
    if (samples) *samples = s->samples;
    return MCG_OK;
}

int mcg_stats_reset(mcg_solver *s) {
    if (int rc = check_ready(s)) return rc;
    HIP_TRY(hipMemsetAsync(s->stats, 0, sizeof(double) * 7 * s->bins(), s->stream));
    HIP_TRY(hipMemsetAsync(s->stats_fx, 0, sizeof(unsigned long long) * 7 * s->bins(), s->stream));
    HIP_TRY(hipMemsetAsync(s->count, 0, sizeof(unsigned long long) * s->bins(), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->started = false;
    s->start_step = 0;
    s->samples = 0;
    return MCG_OK;
}

int mcg_copy_radial(mcg_solver *s, double *sum_uy, double *sum_uy2, double *sum_ur, double *sum_ur2, double *sum_uruy,
                    uint64_t *count, double *sum_phi, double *sum_phi2) {
    if (int rc = check_ready(s)) return rc;
    const size_t nb = s->bins();
    double *dst[7] = {sum_uy, sum_uy2, sum_ur, sum_ur2, sum_uruy, sum_phi, sum_phi2};
    for (int k = 0; k < 7; ++k)
        if (dst[k]) HIP_TRY(hipMemcpyAsync(dst[k], s->stats + k * nb, nb * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    if (count) HIP_TRY(hipMemcpyAsync(count, s->count, nb * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCG_OK;
}

int mcg_tke(mcg_solver *s, double *tke_total) {
    if (int rc = check_ready(s)) return rc;
    if (!tke_total) return fail(MCG_E_INVALID, "null pointer");
    if (int rc = ensure_moments(s)) return rc;
    double r[kRedQ];
    if (int rc = run_reduce(s, r)) return rc;
    *tke_total = r[0];
    return MCG_OK;
}

int mcg_global_sums(mcg_solver *s, double out[5]) {
    if (int rc = check_ready(s)) return rc;
    if (!out) return fail(MCG_E_INVALID, "null pointer");
    if (int rc = ensure_moments(s)) return rc;
    double r[kRedQ];
    if (int rc = run_reduce(s, r)) return rc;
    for (int k = 0; k < 5; ++k) out[k] = r[1 + k];
    return MCG_OK;
}

int mcg_global_sums_noalloc(mcg_solver *s, double *out_dev) {
    if (int rc = check_ready(s)) return rc;
    if (!out_dev) return fail(MCG_E_INVALID, "null output pointer");
    const size_t need = size_t(kRedQ) * size_t(s->red_blocks);
    if (need > kScratchDoubles)
        return fail(MCG_E_INVALID, "global sums: " + std::to_string(need) + " partial sums exceed the scratch array of " +
                                       std::to_string(kScratchDoubles) + " doubles");
    double *part = (s->device >= 0 && s->device < kMaxDevices) ? g_scratch_ptr[s->device] : nullptr;
    if (!part) return fail(MCG_E_STATE, "scratch array of this device was not resolved");
    if (int rc = ensure_moments(s)) return rc;
    const size_t total = size_t(s->c.NX - 2) * (s->c.NY - 2) * (s->c.NZ - 2);
    const int npb = int((total + s->red_blocks - 1) / s->red_blocks);
    hipLaunchKernelGGL(k_reduce_partials, dim3(s->red_blocks), dim3(256), 0, s->stream, s->P, s->mom, part, npb);
    LAUNCH_CHECK("k_reduce_partials");
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s->stream, part, s->red_blocks, out_dev);
    LAUNCH_CHECK("k_reduce_final");
    return MCG_OK;
}

int mcg_check_health(mcg_solver *s) {
    if (int rc = check_ready(s)) return rc;
    if (s->step == 0) return MCG_OK;
    if (int rc = ensure_moments(s)) return rc;
    double r[kRedQ];
    if (int rc = run_reduce(s, r)) return rc;
    if (r[6] > 0) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "watchdog: %.0f interior nodes with non-finite or non-positive density/velocity at step %lld",
                      r[6], (long long)s->step);
        return fail(MCG_E_NUMERIC, buf);
    }
    return MCG_OK;
}

int mcg_copy_fields(mcg_solver *s, float *rhor, float *rhob, float *ux, float *uy, float *uz) {
    if (int rc = check_ready(s)) return rc;
    if (int rc = ensure_moments(s)) return rc;
    const size_t N = s->N();
    if (int rc = s->ensure_staging(5 * N * sizeof(float))) return rc;
    hipLaunchKernelGGL(k_export_fields, s->grid_full(), dim3(64, 4, 1), 0, s->stream, s->P, s->mom, 0, s->gB[s->cur], s->grr,
                       s->grb, s->staging);
    LAUNCH_CHECK("k_export_fields");
    float *dst[5] = {rhor, rhob, ux, uy, uz};
    for (int k = 0; k < 5; ++k)
        if (dst[k]) HIP_TRY(hipMemcpyAsync(dst[k], s->staging + k * N, N * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCG_OK;
}

int mcg_copy_populations(mcg_solver *s, float *fr, float *fb) {
    if (int rc = check_ready(s)) return rc;
    const size_t n = size_t(Q) * s->N();
    if (int rc = s->ensure_staging(2 * n * sizeof(float))) return rc;
    HIP_TRY(hipMemsetAsync(s->staging, 0, 2 * n * sizeof(float), s->stream));
    MCG_DISPATCH_ST(s->storage, k_gather_pops, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, s->f[s->cur], s->bin[s->cur],
                       s->bout[s->cur], s->staging, s->staging + n);
    LAUNCH_CHECK("k_gather_pops");
    if (fr) HIP_TRY(hipMemcpyAsync(fr, s->staging, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (fb) HIP_TRY(hipMemcpyAsync(fb, s->staging + n, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCG_OK;
}

int mcg_set_populations(mcg_solver *s, const float *fr, const float *fb) {
    if (int rc = check_ready(s)) return rc;
    if (!fr || !fb) return fail(MCG_E_INVALID, "null pointer");
    const size_t n = size_t(Q) * s->N();
    if (int rc = s->ensure_staging(2 * n * sizeof(float))) return rc;
    HIP_TRY(hipMemcpyAsync(s->staging, fr, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->staging + n, fb, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
    MCG_DISPATCH_ST(s->storage, k_scatter_pops, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, s->f[s->cur], s->bin[s->cur],
                       s->bout[s->cur], s->staging, s->staging + n);
    LAUNCH_CHECK("k_scatter_pops");
    // entries the boundary never rewrites must agree in both copies
    HIP_TRY(hipMemcpyAsync(s->bin[s->cur ^ 1], s->bin[s->cur], s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->bout[s->cur ^ 1], s->bout[s->cur], s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCG_OK;
}

// checkpoint image: header, populations (current), boundary buffers (current),
// outlet ghost state (current), ghost densities, moments, statistics
struct CkptHeader {
    char magic[8];
    int32_t NX, NY, NZ, storage;
    int64_t step, start_step, samples;
    int32_t started, inlet_reset_pending, moments_valid, reserved;
};

static size_t ckpt_bytes(const mcg_solver *s) {
    return sizeof(CkptHeader) + s->pop_bytes() + 2 * s->bnd_bytes() + s->gB_bytes() + 2 * s->N() * sizeof(float) +
           size_t(11) * s->P.plane * sizeof(float) + sizeof(double) * 7 * s->bins() + sizeof(uint64_t) * s->bins();
}

int mcg_checkpoint_size(const mcg_solver *s, uint64_t *nbytes) {
    if (int rc = check_solver(s)) return rc;
    if (!nbytes) return fail(MCG_E_INVALID, "null pointer");
    *nbytes = ckpt_bytes(s);
    return MCG_OK;
}

int mcg_checkpoint_save(mcg_solver *s, void *host_buf, uint64_t nbytes) {
    if (int rc = check_ready(s)) return rc;
    if (!host_buf || nbytes < ckpt_bytes(s)) return fail(MCG_E_INVALID, "checkpoint buffer too small");
    char *p = static_cast<char *>(host_buf);
    CkptHeader h{};
    std::memcpy(h.magic, "MCGCKPT1", 8);
    h.NX = s->c.NX; h.NY = s->c.NY; h.NZ = s->c.NZ; h.storage = s->storage;
    h.step = s->step; h.start_step = s->start_step; h.samples = s->samples;
    h.started = s->started; h.inlet_reset_pending = s->inlet_reset_pending; h.moments_valid = s->moments_valid;
    if (int rc = ensure_moments(s)) return rc;
    h.moments_valid = 1;
    std::memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    auto d2h = [&](const void *src, size_t nb) -> hipError_t {
        hipError_t e = hipMemcpyAsync(p, src, nb, hipMemcpyDeviceToHost, s->stream);
        p += nb;
        return e;
    };
    HIP_TRY(d2h(s->f[s->cur], s->pop_bytes()));
    HIP_TRY(d2h(s->bin[s->cur], s->bnd_bytes()));
    HIP_TRY(d2h(s->bout[s->cur], s->bnd_bytes()));
    HIP_TRY(d2h(s->gB[s->cur], s->gB_bytes()));
    HIP_TRY(d2h(s->grr, s->N() * sizeof(float)));
    HIP_TRY(d2h(s->grb, s->N() * sizeof(float)));
    HIP_TRY(d2h(s->mom, size_t(11) * s->P.plane * sizeof(float)));
    HIP_TRY(d2h(s->stats, sizeof(double) * 7 * s->bins()));
    HIP_TRY(d2h(s->count, sizeof(uint64_t) * s->bins()));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCG_OK;
}

int mcg_checkpoint_load(mcg_solver *s, const void *host_buf, uint64_t nbytes) {
    if (int rc = check_solver(s)) return rc;
    if (!host_buf || nbytes < ckpt_bytes(s)) return fail(MCG_E_INVALID, "checkpoint buffer too small");
    const char *p = static_cast<const char *>(host_buf);
    CkptHeader h;
    std::memcpy(&h, p, sizeof(h));
    if (std::memcmp(h.magic, "MCGCKPT1", 8) != 0) return fail(MCG_E_INVALID, "not a checkpoint image");
    if (h.NX != s->c.NX || h.NY != s->c.NY || h.NZ != s->c.NZ || h.storage != s->storage)
        return fail(MCG_E_INVALID, "checkpoint was taken for another grid or storage mode");
    p += sizeof(h);
    s->cur = 0;
    auto h2d = [&](void *dst, size_t nb) -> hipError_t {
        hipError_t e = hipMemcpyAsync(dst, p, nb, hipMemcpyHostToDevice, s->stream);
        p += nb;
        return e;
    };
    HIP_TRY(h2d(s->f[0], s->pop_bytes()));
    HIP_TRY(h2d(s->bin[0], s->bnd_bytes()));
    HIP_TRY(h2d(s->bout[0], s->bnd_bytes()));
    HIP_TRY(h2d(s->gB[0], s->gB_bytes()));
    HIP_TRY(h2d(s->grr, s->N() * sizeof(float)));
    HIP_TRY(h2d(s->grb, s->N() * sizeof(float)));
    HIP_TRY(h2d(s->mom, size_t(11) * s->P.plane * sizeof(float)));
    HIP_TRY(h2d(s->stats, sizeof(double) * 7 * s->bins()));
    HIP_TRY(h2d(s->count, sizeof(uint64_t) * s->bins()));
    HIP_TRY(hipMemcpyAsync(s->bin[1], s->bin[0], s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->bout[1], s->bout[0], s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->gB[1], s->gB[0], s->gB_bytes(), hipMemcpyDeviceToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->step = h.step;
    s->start_step = h.start_step;
    s->samples = h.samples;
    s->started = h.started != 0;
    s->inlet_reset_pending = h.inlet_reset_pending != 0;
    s->moments_valid = true;
    s->initialized = true;
    return MCG_OK;
}

int mcg_time_steps(mcg_solver *s, int nsteps, double *ms) {
    if (int rc = check_ready(s)) return rc;
    if (!ms) return fail(MCG_E_INVALID, "null pointer");
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, s->stream));
    int rc = mcg_step(s, nsteps);
    if (rc) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return rc;
    }
    HIP_TRY(hipEventRecord(b, s->stream));
    HIP_TRY(hipEventSynchronize(b));
    float t = 0.0f;
    HIP_TRY(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = double(t);
    return MCG_OK;
}

int mcg_write_vti(mcg_solver *s, const char *path) {
    if (int rc = check_ready(s)) return rc;
    if (!path) return fail(MCG_E_INVALID, "null path");
    const size_t N = s->N();
    std::vector<float> h(5 * N);
    if (int rc = mcg_copy_fields(s, h.data(), h.data() + N, h.data() + 2 * N, h.data() + 3 * N, h.data() + 4 * N)) return rc;
    const int rc = mcg_write_vti_arrays(path, s->c.NX, s->c.NY, s->c.NZ, h.data(), h.data() + N, h.data() + 2 * N,
                                        h.data() + 3 * N, h.data() + 4 * N);
    if (rc) return fail(rc, mcg_io_last_error());
    return MCG_OK;
}

int mcg_write_vti_step(mcg_solver *s, const char *vti_dir, int64_t step) {
    if (!vti_dir) return fail(MCG_E_INVALID, "null directory");
    char name[64];
    std::snprintf(name, sizeof(name), "/lbm_%08lld.vti", (long long)step);
    return mcg_write_vti(s, (std::string(vti_dir) + name).c_str());
}

int mcg_write_radial(mcg_solver *s, const char *mean_dir) {
    if (int rc = check_ready(s)) return rc;
    const size_t nb = s->bins();
    std::vector<double> h(7 * nb);
    std::vector<uint64_t> cnt(nb);
    if (s->started && s->samples > 0) {
        if (int rc = mcg_copy_radial(s, h.data(), h.data() + nb, h.data() + 2 * nb, h.data() + 3 * nb, h.data() + 4 * nb,
                                     cnt.data(), h.data() + 5 * nb, h.data() + 6 * nb))
            return rc;
    }
    const int rc = mcg_write_radial_files(&s->c, mean_dir, h.data(), h.data() + nb, h.data() + 2 * nb, h.data() + 3 * nb,
                                          h.data() + 4 * nb, cnt.data(), h.data() + 5 * nb, h.data() + 6 * nb,
                                          s->started ? 1 : 0, s->start_step, s->samples);
    if (rc) return fail(rc, mcg_io_last_error());
    return MCG_OK;
}

int mcg_hbm_probe(int kind, int vec_bytes, int nontemporal, int blocks_per_cu, uint64_t nbytes, int reps, double *gbs) {
    if (!gbs || reps < 1 || blocks_per_cu < 1 || nbytes < (uint64_t(1) << 20) || (nbytes & 255) ||
        (kind != kProbeRead && kind != kProbeWrite && kind != kProbeCopy) ||
        (vec_bytes != 4 && vec_bytes != 8 && vec_bytes != 16))
        return fail(MCG_E_INVALID, "hbm probe: kind 0..2, vec_bytes 4/8/16, blocks_per_cu >= 1, reps >= 1, "
                                   "nbytes >= 1 MiB and a multiple of 256");
    int dev = 0, ncu = 256;
    HIP_TRY(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
    void *a = nullptr, *b = nullptr, *sink = nullptr;
    if (hipMalloc(&a, nbytes) != hipSuccess || (kind == kProbeCopy && hipMalloc(&b, nbytes) != hipSuccess) ||
        hipMalloc(&sink, 256) != hipSuccess) {
        for (void *p : {a, b, sink})
            if (p) (void)hipFree(p);
        return fail(MCG_E_NOMEM, "hbm probe: device allocation failed");
    }
    const uint32_t *src = static_cast<const uint32_t *>(a);
    uint32_t *dst = static_cast<uint32_t *>(kind == kProbeCopy ? b : a);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MCG_OK;
    float best = 0.0f;
    auto run = [&]() {
        switch (kind) {
            case kProbeRead: launch_probe_vw<kProbeRead>(vec_bytes, nontemporal != 0, blocks_per_cu * ncu, src, dst, nbytes,
                                                         static_cast<uint32_t *>(sink)); break;
            case kProbeWrite: launch_probe_vw<kProbeWrite>(vec_bytes, nontemporal != 0, blocks_per_cu * ncu, src, dst, nbytes,
                                                           static_cast<uint32_t *>(sink)); break;
            default: launch_probe_vw<kProbeCopy>(vec_bytes, nontemporal != 0, blocks_per_cu * ncu, src, dst, nbytes,
                                                 static_cast<uint32_t *>(sink)); break;
        }
    };
    if (hipMemset(a, 0x5A, nbytes) != hipSuccess || (b && hipMemset(b, 0, nbytes) != hipSuccess) ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(MCG_E_HIP, "hbm probe: set-up failed");
    } else {
        run();  // warm-up
        for (int r = 0; r < reps && rc == MCG_OK; ++r) {
            float ms = 0.0f;
            if (hipEventRecord(e0, nullptr) != hipSuccess) rc = fail(MCG_E_HIP, "hbm probe: event record failed");
            run();
            if (rc == MCG_OK && (hipGetLastError() != hipSuccess || hipEventRecord(e1, nullptr) != hipSuccess ||
                                 hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = fail(MCG_E_HIP, "hbm probe: launch or timing failed");
            const double moved = double(nbytes) * (kind == kProbeCopy ? 2.0 : 1.0);
            if (rc == MCG_OK && ms > 0.0f) best = std::max(best, float(moved / (double(ms) * 1e-3) / 1e9));
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (void *p : {a, b, sink})
        if (p) (void)hipFree(p);
    if (rc == MCG_OK) *gbs = double(best);
    return rc;
}

int mcg_device_bytes(const mcg_solver *s, uint64_t *nbytes) {
    if (int rc = check_solver(s)) return rc;
    if (!nbytes) return fail(MCG_E_INVALID, "null pointer");
    *nbytes = s->bytes;
    return MCG_OK;
}

// ---------------------------------------------------------------- adjoint

static int adj_supported(const mcg_case &c, int storage) {
    if (c.y_mode == MCG_Y_JET) {
        if (c.wrap_gradient) return fail(MCG_E_UNSUPPORTED, "the jet adjoint covers the reference gradient (wrap_gradient 0)");
        if (c.viscosity_mode != MCG_VISC_SINGLE) return fail(MCG_E_UNSUPPORTED, "the jet adjoint covers the single relaxation rate");
        if (c.gravity != 0.0) return fail(MCG_E_UNSUPPORTED, "the jet adjoint does not cover body forces");
    } else if (!c.wrap_gradient) {
        return fail(MCG_E_UNSUPPORTED, "the periodic adjoint needs wrap_gradient 1 (the gradient wraps the periodic axes)");
    }
    if (c.sponge_K != 0.0) return fail(MCG_E_UNSUPPORTED, "the adjoint assumes no sponge (sponge_K = 0)");
    if (c.viscosity_mode == MCG_VISC_INTERFACE && !(c.A_override > 0))
        return fail(MCG_E_UNSUPPORTED, "the adjoint with per-node viscosity needs A given directly (A_override > 0)");
    if (storage != MCG_STORE_FP32) return fail(MCG_E_UNSUPPORTED, "the adjoint runs on FP32 populations");
    return MCG_OK;
}

// Binomial checkpointing (Griewank and Walther, "Algorithm 799: revolve",
// ACM TOMS 26(1), 2000).  With f free state slots and t repetitions at most
// beta(f, t) = C(f + t, f) segments can be reversed: the first f slots split
// the range so that the right part needs f - 1 slots and t repetitions and the
// left part f slots and t - 1.  Returns the split point m of [a, b) for f >= 1
// free slots: the smallest t with beta(f, t) >= b - a, then the right part as
// large as beta(f - 1, t) allows.
static double revolve_beta(int f, int t) {
    double v = 1.0;
    for (int i = 1; i <= f; ++i) v = v * double(t + i) / double(i);
    return v;
}

static int revolve_split(int a, int b, int f) {
    const int L = b - a;
    int t = 0;
    while (revolve_beta(f, t) < double(L)) ++t;
    const double right = revolve_beta(f - 1, t);
    const int left = std::max(1, L - int(std::min(right, double(L))));
    return a + std::min(left, L - 1);
}

// copy the solver's current state (populations and, for jet cases, the
// boundary arrays) into a recorded state buffer
static int save_state(mcg_solver *s, void *st) {
    HIP_TRY(hipMemcpyAsync(st, s->f[s->cur], s->pop_bytes(), hipMemcpyDeviceToDevice, s->stream));
    if (s->jet()) {
        HIP_TRY(hipMemcpyAsync(s->st_bin(st), s->bin[s->cur], s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->st_bout(st), s->bout[s->cur], s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->st_gB(st), s->gB[s->cur], s->gB_bytes(), hipMemcpyDeviceToDevice, s->stream));
    }
    return MCG_OK;
}

int mcg_adjoint_record(mcg_solver *s, int nsteps, int checkpoint_every) {
    if (int rc = check_ready(s)) return rc;
    if (int rc = adj_supported(s->c, s->storage)) return rc;
    if (nsteps < 1 || checkpoint_every < 1) return fail(MCG_E_INVALID, "nsteps and checkpoint_every must be >= 1");
    // the recomputation reads the ghost densities as constants; the first
    // step's inlet reset keeps them only after the jet initialisation
    if (s->inlet_reset_pending && s->init_kind != kInitJet)
        return fail(MCG_E_UNSUPPORTED, "the adjoint of the first jet step needs the jet initialisation (take one step first)");
    s->adj_release();
    const int nseg = (nsteps + checkpoint_every - 1) / checkpoint_every;
    const int nslot = s->adj_limit > 0 ? std::min(nseg, s->adj_limit) : nseg;
    for (int k = 0; k < nslot; ++k) {
        void *p = nullptr;
        if (hipMalloc(&p, s->state_bytes()) != hipSuccess) {
            s->adj_release();
            return fail(MCG_E_NOMEM, "adjoint checkpoint allocation failed");
        }
        s->adj_ckpt.push_back(p);
    }
    // the states stored now: every segment, or the first descent of the
    // sweep's schedule (segment 0, then each split point toward the end)
    s->adj_seg_slot.assign(size_t(nseg), -1);
    if (nslot == nseg) {
        for (int k = 0; k < nseg; ++k) s->adj_seg_slot[size_t(k)] = k;
    } else {
        int used = 0, a = 0;
        s->adj_seg_slot[0] = used++;
        while (used < nslot && nseg - a > 1) {
            a = revolve_split(a, nseg, nslot - used);
            s->adj_seg_slot[size_t(a)] = used++;
        }
    }
    s->adj_step0 = s->step;
    for (int k = 0; k < nsteps; ++k) {
        if (k % checkpoint_every == 0) {
            const int slot = s->adj_seg_slot[size_t(k / checkpoint_every)];
            if (slot >= 0)
                if (int rc = save_state(s, s->adj_ckpt[size_t(slot)])) return rc;
        }
        const int rc = one_step(s);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->adj_nsteps = nsteps;
    s->adj_every = checkpoint_every;
    s->adj_recorded = true;
    return MCG_OK;
}

int mcg_adjoint_set_uy_seed(mcg_solver *s, const float *w_uy, int64_t first, int64_t last) {
    if (int rc = check_solver(s)) return rc;
    if (!w_uy) {
        s->uy_release();
        return MCG_OK;
    }
    if (!s->jet()) return fail(MCG_E_UNSUPPORTED, "the u_y functional is defined for jet cases");
    if (first < 0 || last < first) return fail(MCG_E_INVALID, "u_y window must satisfy 0 <= first <= last");
    const dim3 g = s->grid_interior();
    auto get = [&](void **p, size_t nb) -> bool { return *p || hipMalloc(p, nb) == hipSuccess; };
    if (!get((void **)&s->adj_wuy, s->N() * sizeof(float)) || !get((void **)&s->obs_acc, 2 * sizeof(double)) ||
        !get((void **)&s->obs_part, 2 * sizeof(double) * size_t(g.x) * g.y * g.z)) {
        s->uy_release();
        return fail(MCG_E_NOMEM, "u_y functional allocation failed");
    }
    HIP_TRY(hipMemcpyAsync(s->adj_wuy, w_uy, s->N() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->obs_acc, 0, 2 * sizeof(double), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->adj_uy_first = first;
    s->adj_uy_last = last;
    return MCG_OK;
}

int mcg_uy_functional(mcg_solver *s, double *value) {
    if (int rc = check_solver(s)) return rc;
    if (!value) return fail(MCG_E_INVALID, "null pointer");
    if (!s->obs_acc) return fail(MCG_E_STATE, "no u_y functional (call mcg_adjoint_set_uy_seed first)");
    double acc[2];
    HIP_TRY(hipMemcpyAsync(acc, s->obs_acc, sizeof(acc), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    *value = acc[0];
    return MCG_OK;
}

int mcg_set_inlet_offset(mcg_solver *s, const float *du) {
    if (int rc = check_solver(s)) return rc;
    if (!du) {
        if (s->inlet_du) (void)hipFree(s->inlet_du);
        s->inlet_du = nullptr;
        return MCG_OK;
    }
    if (!s->jet()) return fail(MCG_E_UNSUPPORTED, "the inlet offset is defined for jet cases");
    if (!s->inlet_du && hipMalloc((void **)&s->inlet_du, 3 * s->nxz() * sizeof(float)) != hipSuccess)
        return fail(MCG_E_NOMEM, "inlet offset allocation failed");
    HIP_TRY(hipMemcpyAsync(s->inlet_du, du, 3 * s->nxz() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return MCG_OK;
}

// One state of a recomputed segment: populations and, for jet cases, the
// inlet/outlet source arrays and the outlet state.
struct StateRef {
    void *f;
    float *bin, *bout, *gB;
};

// One step from state src into dst, storing the moments of src into rec (11
// planes).  The solver's own buffers are put back afterwards; the step counter
// advances.  Jet cases: the stale entries of the inlet/outlet arrays (sources
// the boundaries never rewrite) are the same in every state of one recorded
// forward, and dst must already hold them.
static int step_from(mcg_solver *s, const StateRef &src, const StateRef &dst, float *rec) {
    void *f0 = s->f[0], *f1 = s->f[1];
    float *bi[2] = {s->bin[0], s->bin[1]}, *bo[2] = {s->bout[0], s->bout[1]}, *gb[2] = {s->gB[0], s->gB[1]};
    const int cur = s->cur;
    if (s->jet()) {
        s->bin[0] = src.bin;
        s->bin[1] = dst.bin;
        s->bout[0] = src.bout;
        s->bout[1] = dst.bout;
        s->gB[0] = src.gB;
        s->gB[1] = dst.gB;
    }
    s->f[0] = src.f;
    s->f[1] = dst.f;
    s->cur = 0;
    s->rec_mom = rec;
    const bool mv = s->moments_valid;  // the solver's moment planes are not touched
    const int rc = one_step(s);
    s->moments_valid = mv;
    s->rec_mom = nullptr;
    s->f[0] = f0;
    s->f[1] = f1;
    for (int b = 0; b < 2; ++b) {
        s->bin[b] = bi[b];
        s->bout[b] = bo[b];
        s->gB[b] = gb[b];
    }
    s->cur = cur;
    return rc;
}

static int adj_alloc(mcg_solver *s) {
    auto get = [&](void **p, size_t nb) -> bool { return *p || hipMalloc(p, nb) == hipSuccess; };
    const size_t seg_len = size_t(std::min(s->adj_every, s->adj_nsteps));  // states per segment
    bool ok = true;
    for (int b = 0; b < 2; ++b) {
        ok = ok && get((void **)&s->adj_lam[b], s->pop_bytes());
        ok = ok && get(&s->adj_pp[b], s->state_bytes());
        if (s->jet()) ok = ok && get((void **)&s->adj_lg[b], s->gB_bytes());
    }
    if (s->jet()) ok = ok && get((void **)&s->adj_din, 3 * s->nxz() * sizeof(double));
    if (s->jet() || s->c.wall_xz) {
        ok = ok && get((void **)&s->adj_psf, s->N() * sizeof(float));
        ok = ok && get((void **)&s->adj_gbs, s->gB_set_bytes() * seg_len);
    } else {
        ok = ok && get((void **)&s->adj_ps, size_t(s->P.plane) * sizeof(float));
    }
    ok = ok && get((void **)&s->adj_lm, size_t(11) * s->P.plane * sizeof(float));
    ok = ok && get((void **)&s->adj_lF, size_t(3) * s->P.plane * sizeof(float));
    ok = ok && get((void **)&s->adj_acc, 5 * sizeof(double));
    ok = ok && get((void **)&s->adj_mom, s->mom_set_bytes() * seg_len);
    const dim3 g = s->grid_interior();
    ok = ok && get((void **)&s->adj_part, 3 * sizeof(double) * size_t(g.x) * g.y * g.z);
    return ok ? MCG_OK : fail(MCG_E_NOMEM, "adjoint buffer allocation failed");
}

// adjoint of recorded step k (its moments mom; jet cases: its outlet state
// gBn): lam_in -> lam_out; jet cases also carry the outlet-state adjoint
// lg_in -> lg_out and add the inlet gradients
static int adjoint_step(mcg_solver *s, int k, const float *gBn, const float *mom, float *lam_in, float *lam_out,
                        float *lg_in, float *lg_out) {
    const dim3 g = s->grid_interior();
    const size_t nb = size_t(g.x) * g.y * g.z;
    AdjJet J{};
    // node partials {A, beta, gravity} -> acc[0..2]; jet column partials {U, amp} -> acc[3..4]
    if (s->jet()) {
        hipLaunchKernelGGL(k_psi_full, s->grid_full(), dim3(64, 4, 1), 0, s->stream, s->P, mom, gBn, s->grr, s->grb,
                           s->adj_psf);
        LAUNCH_CHECK("k_psi_full");
        J.psf = s->adj_psf;
        const int64_t n = s->adj_step0 + k;
        J.w_uy = (s->adj_wuy && n >= s->adj_uy_first && n <= s->adj_uy_last) ? s->adj_wuy : nullptr;
        hipLaunchKernelGGL(k_adj_node<kAdjJet>, g, dim3(64, 4, 1), 0, s->stream, s->P, mom, s->adj_psf, lam_in, s->adj_lm,
                           s->adj_lF, s->adj_part, J);
        LAUNCH_CHECK("k_adj_node");
        hipLaunchKernelGGL(k_adj_sum<3>, dim3(1), dim3(kSumThreads), 0, s->stream, s->adj_part, nb, s->adj_acc);
        LAUNCH_CHECK("k_adj_sum");
        const int cols = (s->c.NX - 2) * (s->c.NZ - 2);
        const dim3 gc(unsigned((cols + 255) / 256));
        hipLaunchKernelGGL(k_adj_cols, gc, dim3(256), 0, s->stream, s->P, mom, lam_in, s->adj_lF, s->adj_lm, gBn, lg_in,
                           lg_out, s->adj_din, s->adj_part, (unsigned long long)(s->adj_step0 + k), s->inlet_du);
        LAUNCH_CHECK("k_adj_cols");
        hipLaunchKernelGGL(k_adj_sum<2>, dim3(1), dim3(kSumThreads), 0, s->stream, s->adj_part, size_t(gc.x), s->adj_acc + 3);
        LAUNCH_CHECK("k_adj_sum");
        hipLaunchKernelGGL(k_adj_pop<kAdjJet>, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, mom, s->adj_lm,
                           s->adj_lF, lam_in, lam_out);
    } else if (s->c.wall_xz) {
        hipLaunchKernelGGL(k_psi_full, s->grid_full(), dim3(64, 4, 1), 0, s->stream, s->P, mom, nullptr, s->grr, s->grb,
                           s->adj_psf);
        LAUNCH_CHECK("k_psi_full");
        J.psf = s->adj_psf;
        hipLaunchKernelGGL(k_adj_node<kAdjWall>, g, dim3(64, 4, 1), 0, s->stream, s->P, mom, s->adj_psf, lam_in,
                           s->adj_lm, s->adj_lF, s->adj_part, J);
        LAUNCH_CHECK("k_adj_node");
        hipLaunchKernelGGL(k_adj_sum<3>, dim3(1), dim3(kSumThreads), 0, s->stream, s->adj_part, nb, s->adj_acc);
        LAUNCH_CHECK("k_adj_sum");
        hipLaunchKernelGGL(k_adj_pop<kAdjWall>, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, mom, s->adj_lm,
                           s->adj_lF, lam_in, lam_out);
    } else {
        hipLaunchKernelGGL(k_psi_plane, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, mom, s->adj_ps);
        LAUNCH_CHECK("k_psi_plane");
        hipLaunchKernelGGL(k_adj_node<kAdjPeriodic>, g, dim3(64, 4, 1), 0, s->stream, s->P, mom, s->adj_ps, lam_in,
                           s->adj_lm, s->adj_lF, s->adj_part, J);
        LAUNCH_CHECK("k_adj_node");
        hipLaunchKernelGGL(k_adj_sum<3>, dim3(1), dim3(kSumThreads), 0, s->stream, s->adj_part, nb, s->adj_acc);
        LAUNCH_CHECK("k_adj_sum");
        hipLaunchKernelGGL(k_adj_pop<kAdjPeriodic>, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, mom,
                           s->adj_lm, s->adj_lF, lam_in, lam_out);
    }
    LAUNCH_CHECK("k_adj_pop");
    return MCG_OK;
}

int mcg_adjoint_backward(mcg_solver *s, const float *lam_fr, const float *lam_fb, float *out_lam_fr, float *out_lam_fb,
                         double *dparams) {
    if (int rc = check_ready(s)) return rc;
    if (!s->adj_recorded) return fail(MCG_E_STATE, "no recorded forward (call mcg_adjoint_record first)");
    if (s->step != s->adj_step0 + s->adj_nsteps) return fail(MCG_E_STATE, "the solver advanced after mcg_adjoint_record");
    if (!lam_fr || !lam_fb) return fail(MCG_E_INVALID, "null adjoint seed");
    if (int rc = adj_alloc(s)) return rc;
    const size_t n = size_t(Q) * s->N();
    if (int rc = s->ensure_staging(2 * n * sizeof(float))) return rc;
    // The current state (buffer cur0) is never written.  The other buffer
    // (the previous state, which the next step overwrites) takes the output of
    // each segment's last recomputed step, whose only use is the moments of
    // the segment's last state; the solver's moments of the last step (the
    // fields it exports) are taken from it first.
    if (int rc = ensure_moments(s)) return rc;
    const int cur0 = s->cur;
    HIP_TRY(hipMemcpyAsync(s->staging, lam_fr, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemcpyAsync(s->staging + n, lam_fb, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->adj_lam[0], 0, s->pop_bytes(), s->stream));
    hipLaunchKernelGGL(k_rows_from_ref, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, s->staging, s->staging + n,
                       s->adj_lam[0]);
    LAUNCH_CHECK("k_rows_from_ref");
    HIP_TRY(hipMemsetAsync(s->adj_acc, 0, 5 * sizeof(double), s->stream));
    if (s->jet()) {
        // stale inlet/outlet entries, identical in every recorded state
        for (void *p : s->adj_pp) {
            HIP_TRY(hipMemcpyAsync(s->st_bin(p), s->st_bin(s->adj_ckpt[0]), s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
            HIP_TRY(hipMemcpyAsync(s->st_bout(p), s->st_bout(s->adj_ckpt[0]), s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
        }
        // the objective does not read the final outlet state
        HIP_TRY(hipMemsetAsync(s->adj_lg[0], 0, s->gB_bytes(), s->stream));
        HIP_TRY(hipMemsetAsync(s->adj_lg[1], 0, s->gB_bytes(), s->stream));
        HIP_TRY(hipMemsetAsync(s->adj_din, 0, 3 * s->nxz() * sizeof(double), s->stream));
    }
    s->stats_frozen = true;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->stream));
    int a = 0;
    int rc = MCG_OK;
    const int nseg = int(s->adj_seg_slot.size()), K = s->adj_every;
    const size_t mset = s->mom_set_bytes() / sizeof(float), gset = s->gB_set_bytes() / sizeof(float);
    const bool jet = s->jet();
    s->adj_recomputed = 0;
    auto state = [&](void *p) { return StateRef{p, s->st_bin(p), s->st_bout(p), s->st_gB(p)}; };
    // Reverse segment seg from its first state in buffer sb: recompute the
    // segment through the two alternating buffers (the one that is not sb
    // first), each step recording the moments (and, jet cases, the outlet
    // state) of its source state, all the adjoint reads; the last step's
    // output goes to the solver's other buffer.  Then the adjoint steps.
    auto reverse_segment = [&](int seg, void *sb) -> int {
        const int n0 = seg * K, n1 = std::min(n0 + K, s->adj_nsteps);
        auto moms = [&](int k) -> float * { return s->adj_mom + size_t(k - n0) * mset; };
        auto gbs = [&](int k) -> float * { return jet ? s->adj_gbs + size_t(k - n0) * gset : nullptr; };
        StateRef src{sb, s->st_bin(sb), s->st_bout(sb), gbs(n0)};
        if (jet) HIP_TRY(hipMemcpyAsync(gbs(n0), s->st_gB(sb), s->gB_bytes(), hipMemcpyDeviceToDevice, s->stream));
        s->step = s->adj_step0 + n0;
        int x = sb == s->adj_pp[0] ? 1 : 0;
        for (int k = n0; k < n1; ++k) {
            StateRef dst;
            if (k + 1 < n1) {
                void *p = s->adj_pp[x];
                x ^= 1;
                dst = StateRef{p, s->st_bin(p), s->st_bout(p), gbs(k + 1)};
            } else {
                dst = StateRef{s->f[cur0 ^ 1], s->bin[cur0 ^ 1], s->bout[cur0 ^ 1], s->gB[cur0 ^ 1]};
            }
            if (int r = step_from(s, src, dst, moms(k))) return r;
            src = dst;
        }
        s->adj_recomputed += n1 - n0;
        for (int k = n1 - 1; k >= n0; --k) {
            if (int r = adjoint_step(s, k, gbs(k), moms(k), s->adj_lam[a], s->adj_lam[a ^ 1], s->adj_lg[a], s->adj_lg[a ^ 1]))
                return r;
            a ^= 1;
        }
        return MCG_OK;
    };
    // Advance from the first state of segment sa (buffer sb) to the first
    // state of segment sm, into buffer to (null: whichever alternating
    // buffer the last step writes, returned in *out).
    auto advance = [&](int sa, void *sb, int sm, void *to, void **out) -> int {
        StateRef src = state(sb);
        int x = 0;
        const int n_end = sm * K;
        for (int n = sa * K; n < n_end; ++n) {
            void *p = (to && n + 1 == n_end) ? to : s->adj_pp[x];
            x ^= 1;
            const StateRef dst = state(p);
            s->step = s->adj_step0 + n;
            if (int r = step_from(s, src, dst, nullptr)) return r;
            src = dst;
        }
        s->adj_recomputed += n_end - sa * K;
        if (out) *out = src.f;
        return MCG_OK;
    };
    std::vector<int> free_slots;
    std::vector<char> rewritten(s->adj_ckpt.size(), 0);  // slots whose recorded state the sweep replaced
    {
        std::vector<char> used(s->adj_ckpt.size(), 0);
        for (int v : s->adj_seg_slot)
            if (v >= 0) used[size_t(v)] = 1;
        for (size_t v = 0; v < used.size(); ++v)
            if (!used[v]) {
                free_slots.push_back(int(v));
                if (jet) {  // stale inlet/outlet entries, as in every recorded state
                    void *p = s->adj_ckpt[v];
                    HIP_TRY(hipMemcpyAsync(s->st_bin(p), s->st_bin(s->adj_ckpt[0]), s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
                    HIP_TRY(hipMemcpyAsync(s->st_bout(p), s->st_bout(s->adj_ckpt[0]), s->bnd_bytes(), hipMemcpyDeviceToDevice, s->stream));
                }
            }
    }
    // Reverse segments [sa, sb): the first state of sa is stored.  Slots for
    // this range: the free ones and those holding stored states inside it
    // (the record's first descent), so the splits repeat the record's.
    std::function<int(int, int)> rev = [&](int sa, int sb) -> int {
        while (sb - sa > 1) {
            int f = int(free_slots.size());
            for (int j = sa + 1; j < sb; ++j) f += s->adj_seg_slot[size_t(j)] >= 0;
            void *base = s->adj_ckpt[size_t(s->adj_seg_slot[size_t(sa)])];
            if (f == 0) {  // no slot: recompute every segment's first state from sa
                for (int j = sb - 1; j > sa; --j) {
                    void *p = nullptr;
                    if (int r = advance(sa, base, j, nullptr, &p)) return r;
                    if (int r = reverse_segment(j, p)) return r;
                }
                break;
            }
            const int m = revolve_split(sa, sb, f);
            if (s->adj_seg_slot[size_t(m)] < 0) {
                const int slot = free_slots.back();
                free_slots.pop_back();
                if (int r = advance(sa, base, m, s->adj_ckpt[size_t(slot)], nullptr)) return r;
                s->adj_seg_slot[size_t(m)] = slot;
                rewritten[size_t(slot)] = 1;
            }
            if (int r = rev(m, sb)) return r;
            free_slots.push_back(s->adj_seg_slot[size_t(m)]);
            s->adj_seg_slot[size_t(m)] = -1;
            sb = m;
        }
        return reverse_segment(sa, s->adj_ckpt[size_t(s->adj_seg_slot[size_t(sa)])]);
    };
    const std::vector<int> seg_slot0 = s->adj_seg_slot;
    rc = rev(0, nseg);
    // recorded states the sweep did not overwrite stay usable for another sweep
    s->adj_seg_slot = seg_slot0;
    for (int &v : s->adj_seg_slot)
        if (v >= 0 && rewritten[size_t(v)]) v = -1;
    s->stats_frozen = false;
    (void)hipEventRecord(e1, s->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->adj_ms = double(ms);
    s->cur = cur0;
    s->step = s->adj_step0 + s->adj_nsteps;
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(s->staging, 0, 2 * n * sizeof(float), s->stream));
    hipLaunchKernelGGL(k_rows_to_ref, s->grid_interior(), dim3(64, 4, 1), 0, s->stream, s->P, s->adj_lam[a], s->staging,
                       s->staging + n);
    LAUNCH_CHECK("k_rows_to_ref");
    if (out_lam_fr) HIP_TRY(hipMemcpyAsync(out_lam_fr, s->staging, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (out_lam_fb) HIP_TRY(hipMemcpyAsync(out_lam_fb, s->staging + n, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    HIP_TRY(hipMemcpyAsync(acc, s->adj_acc, sizeof(acc), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (int v = 0; v < 5; ++v) s->adj_grad[v] = acc[v];
    s->adj_inlet[0] = acc[3];
    s->adj_inlet[1] = acc[4];
    if (dparams) {
        dparams[0] = acc[0];
        dparams[1] = acc[1];
    }
    return MCG_OK;
}

int mcg_adjoint_set_checkpoint_limit(mcg_solver *s, int max_states) {
    if (int rc = check_solver(s)) return rc;
    if (max_states < 0) return fail(MCG_E_INVALID, "max_states must be >= 0 (0: one state per segment)");
    s->adj_limit = max_states;
    return MCG_OK;
}

int mcg_adjoint_recomputed_steps(const mcg_solver *s, int64_t *nsteps) {
    if (int rc = check_solver(s)) return rc;
    if (!nsteps) return fail(MCG_E_INVALID, "null pointer");
    *nsteps = s->adj_recomputed;
    return MCG_OK;
}

int mcg_adjoint_gradients(const mcg_solver *s, double *out, int n) {
    if (int rc = check_solver(s)) return rc;
    if (!out || n < 0 || n > 5) return fail(MCG_E_INVALID, "need 0 <= n <= 5 and an output array");
    for (int v = 0; v < n; ++v) out[v] = s->adj_grad[v];
    return MCG_OK;
}

int mcg_adjoint_inlet_gradient(const mcg_solver *s, double *dparams, double *dinlet) {
    if (int rc = check_solver(s)) return rc;
    if (!s->jet()) return fail(MCG_E_UNSUPPORTED, "the inlet gradient is defined for jet cases");
    if (!s->adj_din) return fail(MCG_E_STATE, "no adjoint sweep (call mcg_adjoint_backward first)");
    if (dparams) {
        dparams[0] = s->adj_inlet[0];
        dparams[1] = s->adj_inlet[1];
    }
    if (dinlet) HIP_TRY(hipMemcpy(dinlet, s->adj_din, 3 * s->nxz() * sizeof(double), hipMemcpyDeviceToHost));
    return MCG_OK;
}

int mcg_adjoint_last_ms(const mcg_solver *s, double *ms) {
    if (int rc = check_solver(s)) return rc;
    if (!ms) return fail(MCG_E_INVALID, "null pointer");
    *ms = s->adj_ms;
    return MCG_OK;
}

int mcg_adjoint_release(mcg_solver *s) {
    if (int rc = check_solver(s)) return rc;
    s->adj_release();
    return MCG_OK;
}

int mcg_cpu_adjoint(const mcg_case *c, int nsteps, const double *fr0, const double *fb0, const double *fields0,
                    const double *lam_fr, const double *lam_fb, double *out_fr, double *out_fb, double *out_lam_fr,
                    double *out_lam_fb, double *dparams) {
    if (!c || !fr0 || !fb0) return fail(MCG_E_INVALID, "null pointer");
    if (int rc = mcg_case_validate(c)) return rc;
    if (c->y_mode != MCG_Y_PERIODIC) return fail(MCG_E_UNSUPPORTED, "jet cases: use mcg_cpu_adjoint_jet");
    if (int rc = adj_supported(*c, MCG_STORE_FP32)) return rc;
    if (c->wall_xz && !fields0) return fail(MCG_E_INVALID, "walls need fields0 (the ghost densities)");
    if (nsteps < 0) return fail(MCG_E_INVALID, "nsteps < 0");
    CpuAdjoint A;
    A.NX = c->NX;
    A.NY = c->NY;
    A.NZ = c->NZ;
    A.N = size_t(c->NX) * c->NY * c->NZ;
    A.omega = 1.0 / c->taub;
    A.A = c->A_override > 0.0 ? c->A_override : c->sigma / (4.0 / 9.0) / c->taub;
    A.beta = c->beta_recolor;
    A.wall = c->wall_xz != 0;
    A.ext.iface = c->viscosity_mode == MCG_VISC_INTERFACE;
    A.ext.taur = c->taur;
    A.ext.taub = c->taub;
    A.ext.grav = c->gravity;
    if (A.wall) {
        A.gpsi.resize(A.N);
        for (size_t g = 0; g < A.N; ++g) A.gpsi[g] = (fields0[g] - fields0[A.N + g]) / (fields0[g] + fields0[A.N + g]);
    }
    const size_t n = size_t(Q) * A.N;
    std::vector<std::vector<double>> R(size_t(nsteps) + 1, std::vector<double>(n)), B(size_t(nsteps) + 1, std::vector<double>(n));
    std::copy(fr0, fr0 + n, R[0].begin());
    std::copy(fb0, fb0 + n, B[0].begin());
    for (int k = 0; k < nsteps; ++k) A.step(R[k].data(), B[k].data(), R[k + 1].data(), B[k + 1].data());
    if (out_fr) std::copy(R[nsteps].begin(), R[nsteps].end(), out_fr);
    if (out_fb) std::copy(B[nsteps].begin(), B[nsteps].end(), out_fb);
    if (!lam_fr || !lam_fb) return MCG_OK;
    std::vector<double> lr(lam_fr, lam_fr + n), lb(lam_fb, lam_fb + n), tr(n), tb(n);
    double dA = 0.0, dbeta = 0.0, dgrav = 0.0;
    for (int k = nsteps - 1; k >= 0; --k) {
        A.adjoint(R[k].data(), B[k].data(), lr.data(), lb.data(), tr.data(), tb.data(), dA, dbeta, dgrav);
        lr.swap(tr);
        lb.swap(tb);
    }
    if (out_lam_fr) std::copy(lr.begin(), lr.end(), out_lam_fr);
    if (out_lam_fb) std::copy(lb.begin(), lb.end(), out_lam_fb);
    if (dparams) {
        dparams[0] = dA;
        dparams[1] = dbeta;
        dparams[2] = dgrav;
    }
    return MCG_OK;
}

int mcg_cpu_adjoint_jet(const mcg_case *c, int nsteps, const double *fr0, const double *fb0, const double *fields0,
                        const double *inlet_du, const double *lam_fr, const double *lam_fb, const double *w_uy, int uy_first, int uy_last,
                        double *out_fr, double *out_fb, double *out_lam_fr, double *out_lam_fb, double *dparams,
                        double *dinlet) {
    if (!c || !fr0 || !fb0 || !fields0) return fail(MCG_E_INVALID, "null pointer");
    if (int rc = mcg_case_validate(c)) return rc;
    if (c->y_mode != MCG_Y_JET || c->wrap_gradient) return fail(MCG_E_UNSUPPORTED, "the jet adjoint covers y_mode jet with the reference gradient");
    if (c->gravity != 0.0) return fail(MCG_E_UNSUPPORTED, "the jet adjoint does not cover body forces");
    if (c->viscosity_mode != MCG_VISC_SINGLE || c->sponge_K != 0.0)
        return fail(MCG_E_UNSUPPORTED, "the CPU adjoint assumes a uniform relaxation rate");
    if (nsteps < 0) return fail(MCG_E_INVALID, "nsteps < 0");
    CpuJetAdjoint J;
    J.c = *c;
    J.NX = c->NX;
    J.NY = c->NY;
    J.NZ = c->NZ;
    J.N = size_t(c->NX) * c->NY * c->NZ;
    J.omega = 1.0 / c->taub;
    J.A = c->A_override > 0.0 ? c->A_override : c->sigma / (4.0 / 9.0) / c->taub;
    J.beta = c->beta_recolor;
    const size_t N = J.N, n = size_t(Q) * N, nxz = size_t(c->NX) * c->NZ;
    if (inlet_du) J.du.assign(inlet_du, inlet_du + 3 * nxz);
    J.gpsi.resize(N);
    for (size_t g = 0; g < N; ++g) J.gpsi[g] = (fields0[g] - fields0[N + g]) / (fields0[g] + fields0[N + g]);
    std::vector<std::vector<double>> R(size_t(nsteps) + 1, std::vector<double>(n)), B(size_t(nsteps) + 1, std::vector<double>(n)),
        G(size_t(nsteps) + 1, std::vector<double>(5 * nxz, 0.0));
    std::copy(fr0, fr0 + n, R[0].begin());
    std::copy(fb0, fb0 + n, B[0].begin());
    for (int z = 1; z < c->NZ - 1; ++z)
        for (int x = 1; x < c->NX - 1; ++x)
            for (int k = 0; k < 5; ++k)
                G[0][k * nxz + size_t(z) * c->NX + x] = fields0[k * N + J.id(x, c->NY - 1, z)];
    for (int k = 0; k < nsteps; ++k) J.step(k, R[k].data(), B[k].data(), G[k].data(), R[k + 1].data(), B[k + 1].data(), G[k + 1].data());
    if (out_fr) std::copy(R[nsteps].begin(), R[nsteps].end(), out_fr);
    if (out_fb) std::copy(B[nsteps].begin(), B[nsteps].end(), out_fb);
    if (!lam_fr || !lam_fb) return MCG_OK;
    std::vector<double> lr(lam_fr, lam_fr + n), lb(lam_fb, lam_fb + n), tr(n), tb(n), lg(5 * nxz, 0.0), tg(5 * nxz);
    std::vector<double> din(3 * nxz, 0.0);
    double dp[4] = {0, 0, 0, 0};
    for (int k = nsteps - 1; k >= 0; --k) {
        const double *w = (w_uy && k >= uy_first && k <= uy_last) ? w_uy : nullptr;
        J.adjoint(k, R[k].data(), B[k].data(), G[k].data(), lr.data(), lb.data(), lg.data(), tr.data(), tb.data(), tg.data(), dp,
                  din.data(), w);
        lr.swap(tr);
        lb.swap(tb);
        lg.swap(tg);
    }
    if (out_lam_fr) std::copy(lr.begin(), lr.end(), out_lam_fr);
    if (out_lam_fb) std::copy(lb.begin(), lb.end(), out_lam_fb);
    if (dparams) std::copy(dp, dp + 4, dparams);
    if (dinlet) std::copy(din.begin(), din.end(), dinlet);
    return MCG_OK;
}

}  // extern "C"
