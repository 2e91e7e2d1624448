/*
 * mcg.h -- C ABI of the MI355X D3Q27 color-gradient lattice Boltzmann solver.
 *
 * The reference (Felipe-Aristi/Multicomponent_Color_Gradient) has no plugin or
 * FFI layer: its operator API is the set of inline host launchers in
 * launch.cuh:13-119 plus the allocation helpers in memory.cuh:58-133 and
 * meanF/memoryMeanFields.cuh:57-126, driven by main.cu:17-120.  Every entry
 * point below names the reference function(s) it replaces.
 *
 * Conventions
 *   - Every function returns MCG_OK (0) or a negative MCG_E* status; the
 *     library never exits the process (the reference's CUDA_CHECK calls
 *     std::exit, utilities/cudaUtilities.cuh:9-19).  mcg_last_error() returns
 *     a thread-local message for the last failure.
 *   - Grid indexing is the reference's: cell id = x + NX*(y + NY*z)
 *     (utilities/indexing.cuh:10-13), population id = i*Ncells + id
 *     (utilities/indexing.cuh:15-19), D3Q27 directions in the order of
 *     stencil_ct.cuh:9-16.  Host buffers passed to the copy functions use
 *     this layout and cover the full NX*NY*NZ grid, ghosts included.
 *   - One solver handle = one device, one HIP stream (hipStream_t passed as
 *     void*, NULL = the null stream).  A handle is not thread safe.
 */
#ifndef MCG_H
#define MCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCG_OK 0
#define MCG_E_INVALID (-1)     /* bad argument or case                         */
#define MCG_E_HIP (-2)         /* HIP runtime error                            */
#define MCG_E_NOMEM (-3)       /* device or host allocation failed             */
#define MCG_E_STATE (-4)       /* call out of order (e.g. step before init)    */
#define MCG_E_UNSUPPORTED (-5) /* option combination not implemented           */
#define MCG_E_NUMERIC (-6)     /* watchdog: NaN/Inf or non-positive density    */
#define MCG_E_IO (-7)          /* file could not be written                    */

/* y-direction treatment */
#define MCG_Y_JET 0      /* inlet at y=0, convective outlet at y=NY-1 (jet/boundary_conditions.cuh) */
#define MCG_Y_PERIODIC 1 /* periodic in y with the reference's wrap (utilities/bounds.cuh:36-47)      */

/* collision relaxation time */
#define MCG_VISC_SINGLE 0    /* omega = 1/taub everywhere (lbm.cuh:211, forces.cuh:102)             */
#define MCG_VISC_INTERFACE 1 /* tau = (1+psi)/2 taur + (1-psi)/2 taub per node (forces.cuh:77-85)   */

/* inlet velocity profile (MCG_Y_JET only) */
#define MCG_INLET_SHARP 0 /* top-hat disk, jet/boundary_conditions.cuh:33-47                         */
#define MCG_INLET_TANH 1  /* tanh-edged top-hat plus seeded transverse perturbation (extension)       */

/* population storage */
#define MCG_STORE_FP32 0
#define MCG_STORE_FP16 1 /* fp16 offsets from the lattice weights, fp32 arithmetic */
#define MCG_STORE_BF16 2 /* bf16 offsets from the lattice weights, fp32 arithmetic */
/* With 16-bit storage the steps store every population with stochastic
 * rounding: unbiased, with random bits from a hash of (node, step), so a run
 * is reproducible bit for bit.  mcg_set_populations rounds to nearest. */

/* step kernel variant */
#define MCG_KERNEL_FUSED 0   /* one fused launch per step (default)                    */
#define MCG_KERNEL_TWOPASS 1 /* moments pass + collide pass (the reference's data flow) */

/*
 * Case description: the run-time form of the compile-time constants of
 * constants.cuh:8-76.  Names follow constants.cuh.  Derived quantities
 * (nu, omega, sigma_over_4cs4) are computed from these fields exactly as
 * constants.cuh:59-76 does; mcg_case_reference_jet() fills the reference's
 * values.
 */
typedef struct mcg_case {
    int32_t NX, NY, NZ;          /* constants.cuh:14-16 (ghost layers included)  */
    int32_t NSTEP;               /* constants.cuh:8                              */
    int32_t NOUTPUT;             /* constants.cuh:9                              */
    int32_t NSTATS_START_STEP;   /* constants.cuh:10                             */
    int32_t NSTATS_SAMPLE;       /* sampling period of the radial statistics;
                                    the reference samples every step (main.cu:58-62,
                                    meanF/meanFieldsRT.cuh:246 writes 1)         */
    int32_t write_vti_output;    /* constants.cuh:11                             */
    int32_t sponge_cells;        /* constants.cuh:17                             */
    int32_t y_mode;              /* MCG_Y_*                                      */
    int32_t wrap_gradient;       /* 0: colour gradient reads ghost nodes as they
                                    are (forces.cuh:31-33); 1: periodic wrap     */
    int32_t viscosity_mode;      /* MCG_VISC_*                                   */
    int32_t inlet_profile;       /* MCG_INLET_*                                  */
    int32_t perturbation_seed;   /* seed of the inlet perturbation (MCG_INLET_TANH) */
    int32_t wall_xz;             /* 1: no-slip walls (half-way bounce-back) on the x and z
                                    ghost layers instead of periodic x/z (y_mode periodic
                                    only); the colour gradient reads the frozen ghost
                                    densities there (an extension)                  */
    int32_t reserved_i[2];

    double jet_radius, jet_x0, jet_z0, jet_velocity;       /* constants.cuh:32-35 */
    double bubble_radius, bubble_x0, bubble_y0, bubble_z0; /* constants.cuh:38-41 */
    double rhor0, rhob0;         /* constants.cuh:56-57                          */
    double Re, We;               /* constants.cuh:59,71 (labels of output folders) */
    double taur, taub;           /* constants.cuh:62-64                          */
    double sigma;                /* constants.cuh:73                             */
    double beta_recolor;         /* constants.cuh:74                             */
    double A_override;           /* >0: surface-tension amplitude A used as given;
                                    <=0: A = sigma/(4 cs^4 tau) (forces.cuh:87-90) */
    double sponge_K;             /* utilities/mathUtilities.cuh:50 (reference 0)  */
    double interface_width;      /* tanh width for mcg_init_bubble_tanh / inlet   */
    double perturbation_amp;     /* inlet transverse perturbation, fraction of U  */
    double gravity;              /* body force F = (0, -gravity rho_r, 0) on the red
                                    (heavy) component (an extension): u = (j + F/2)/rho
                                    and a source w_i [(1 - omega/2)(9 (c_i.u) c_i - 3 u)
                                    + 1.5 c_i].F in the collision, so that the
                                    post-collision momentum is j + F               */
    double reserved_d[3];
} mcg_case;

typedef struct mcg_solver mcg_solver;

/* ---- case helpers ------------------------------------------------------ */

/* Fill *c with the reference case of constants.cuh:8-76 (128x400x128 jet,
 * Re 5000, We 2000). */
int mcg_case_reference_jet(mcg_case *c);

/* Validate a case; returns MCG_OK or MCG_E_INVALID / MCG_E_UNSUPPORTED. */
int mcg_case_validate(const mcg_case *c);

/* sizeof(mcg_case), for bindings that mirror the struct. */
size_t mcg_case_size(void);

/* ---- lifetime ------------------------------------------------------------ */

/* Allocate the device state for case *c on the current HIP device.
 * Replaces allocate_device_memory (memory.cuh:58-99), allocate_mean_fields_device
 * (meanF/memoryMeanFields.cuh:57-78) and make_cudaConfig (utilities/cudaConfig.cuh:22-31).
 * storage: MCG_STORE_*.  stream: hipStream_t or NULL. */
int mcg_create(const mcg_case *c, int storage, void *stream, mcg_solver **out);

/* Free everything.  Replaces free_device_memory (memory.cuh:113-133) and
 * free_mean_fields_device (meanF/memoryMeanFields.cuh:103-115). */
int mcg_destroy(mcg_solver *s);

/* Select the step kernel (MCG_KERNEL_*).  Both give identical results. */
int mcg_set_kernel(mcg_solver *s, int kernel);

/* Launch shape of the fused step (an extension; the reference launches fixed
 * block shapes, utilities/cudaConfig.cuh:22-31).  The step kernel cuts z into
 * segments of equal length (uniform shapes) or, per XCD, into z-bands of
 * decreasing thickness (banded schedules).  segments = 0 (the default): the
 * solver's first plain steps (no statistics) try the candidate shapes in turn,
 * each timed with HIP events and waited for, two passes over all of them and
 * three more over the four fastest, and keep the fastest.  segments > 0: the uniform shape nearest to that many segments,
 * fixed, with no timing.  segments = -k: the k-th banded schedule (1 = the
 * shortest estimated span), fixed; MCG_E_INVALID if there is none.  The
 * results do not depend on the launch shape. */
int mcg_set_launch_segments(mcg_solver *s, int segments);

/* The current launch shape: z segments (or the z-bands of a banded
 * schedule), workgroups per step, whether the timing has chosen it, and
 * whether it is banded (0: uniform; 1, 2, 3: banded with the tile-planes
 * dealt to the XCDs in contiguous shares, in chunk pairs k and k+8, or in
 * chunk pairs k and 15-k), and the mcg_set_launch_segments argument that
 * fixes this shape (n segments, or -k for the k-th banded schedule).  Any
 * pointer may be NULL. */
int mcg_launch_info(const mcg_solver *s, int32_t *segments, int32_t *workgroups, int32_t *tuned, int32_t *banded,
                    int32_t *shape);

/* The launch shape of statistics steps (an extension, as above): they run
 * another instance of the step kernel (8-row tiles), whose best shapes differ
 * from the plain step's.  With segments = 0 the first statistics steps time
 * uniform z segmentations (one pass each, three more over the four fastest)
 * and keep the fastest; until then, and with a fixed shape, they take the
 * plain step's shape re-tiled.  Reported: z segments, workgroups, chosen by
 * the timing or not.  Any pointer may be NULL. */
int mcg_stats_launch_info(const mcg_solver *s, int32_t *segments, int32_t *workgroups, int32_t *tuned);

/* ---- case API: initial conditions ------------------------------------- */

/* Jet initial state: launch_jetDensity + launch_Injet (launch.cuh:19-29,
 * main.cu:39-40; jet/jetIn.cuh:13-62).  Resets the step counter and the
 * statistics. */
int mcg_init_jet(mcg_solver *s);

/* Sharp bubble initial state: launch_bubble (launch.cuh:13-17,
 * bubble/bubbleIn.cuh:10-39). */
int mcg_init_bubble(mcg_solver *s);

/* Extension: bubble with a tanh interface, blue fraction
 * (1 + tanh((R - r)/interface_width))/2, populations w_i * rho_k. */
int mcg_init_bubble_tanh(mcg_solver *s);

/* Initialise at rest with a field of n tanh spheres (an extension, the
 * interface-rich workload): blue fraction = max over the spheres k of
 * (1 + tanh((R_k - r_k)/interface_width))/2 at every node (ghosts included),
 * rho_r = (1 - frac) rhor0, rho_b = frac rhob0, f_i = w_i rho.  Centres and
 * radii come from mcg_sphere_list(seed, n, NX, NY, NZ, rmin, rmax). */
int mcg_init_spheres(mcg_solver *s, int n, uint32_t seed, double rmin, double rmax);

/* The spheres of mcg_init_spheres (host only): out[4k..4k+3] = x, y, z, R of
 * sphere k, uniform in [1, N-2] per axis and in [rmin, rmax], from SplitMix64:
 * key_k = sm(sm(seed) ^ k), u_j = (sm(key_k + j) >> 40) / 2^24 for j = 1..4. */
int mcg_sphere_list(uint32_t seed, int n, int NX, int NY, int NZ, double rmin, double rmax, double *out);

/* ---- the hot path ---------------------------------------------------- */

/* Advance nsteps time steps.  One step is the body of the reference loop
 * main.cu:54-66: launch_Macros -> [launch_accumulate_radial_moments once
 * step >= NSTATS_START_STEP, every NSTATS_SAMPLE steps] -> launch_collistream
 * -> launch_inlet_bc -> launch_neumann_bc (launch.cuh:31-119).  Asynchronous
 * on the solver's stream. */
int mcg_step(mcg_solver *s, int nsteps);

/* Wait for the solver's stream. */
int mcg_synchronize(mcg_solver *s);

/* Number of steps taken since the last init. */
int mcg_get_step(const mcg_solver *s, int64_t *step);

/* ---- statistics and diagnostics ------------------------------------- */

/* Radial statistics state (MeanFieldsState, meanF/memoryMeanFields.cuh:39-45). */
int mcg_stats_info(const mcg_solver *s, int32_t *started, int64_t *start_step, int64_t *samples);

/* Zero the radial accumulators and restart the start-step logic. */
int mcg_stats_reset(mcg_solver *s);

/* Copy the radial accumulators, each NY*NR_BINS (row y, column r-bin),
 * NR_BINS = NX/2.  Replaces the D2H copies of write_radial_profile_outputs
 * (meanF/meanFieldsRT.cuh:305-328).  sum_phi / sum_phi2 (phase field psi and
 * its square) are an extension; any pointer may be NULL to skip it. */
int mcg_copy_radial(mcg_solver *s, double *sum_uy, double *sum_uy2, double *sum_ur,
                    double *sum_ur2, double *sum_uruy, uint64_t *count,
                    double *sum_phi, double *sum_phi2);

/* Total kinetic energy sum over interior nodes of 0.5|u|^2 (FP64).
 * Replaces compute_total_tke (kernels.cu:136-152, launch.cuh:39-49). */
int mcg_tke(mcg_solver *s, double *tke_total);

/* FP64 global sums over interior nodes: out[0]=sum rho_r, out[1]=sum rho_b,
 * out[2..4]=sum rho*u.  Extension (the reference has none). */
int mcg_global_sums(mcg_solver *s, double out[5]);

/* The same FP64 sums as mcg_tke, mcg_global_sums and the watchdog, written
 * to device memory on the solver's stream without allocating anything and
 * without synchronising: out_dev (7 doubles of device memory) receives
 * {sum 0.5|u|^2, sum rho_r, sum rho_b, sum rho u_x, sum rho u_y, sum rho u_z,
 * number of interior nodes the watchdog rejects}.  The partial sums live in a
 * library-owned device array of 8192 doubles per GPU (64 KiB, held until the
 * process ends); a solver needing more partials is rejected with
 * MCG_E_INVALID.  Calls that overlap on one GPU race on that array.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). */
int mcg_global_sums_noalloc(mcg_solver *s, double *out_dev);

/* Watchdog: returns MCG_E_NUMERIC if any interior density is NaN/Inf or <= 0. */
int mcg_check_health(mcg_solver *s);

/* ---- state exchange --------------------------------------------------- */

/* Copy rho_r, rho_b, u_x, u_y, u_z of the full grid (NX*NY*NZ each, ghosts
 * included) as they stand after the last step: interior nodes hold the
 * moments computed at the start of that step, ghost nodes the boundary values.
 * Replaces copy_out_D2H (memory.cuh:135-142). */
int mcg_copy_fields(mcg_solver *s, float *rhor, float *rhob, float *ux, float *uy, float *uz);

/* Copy the populations f_r, f_b (27*NX*NY*NZ each, reference layout) that
 * the next step will read; ghost-node slots are written as 0. */
int mcg_copy_populations(mcg_solver *s, float *fr, float *fb);

/* Overwrite the interior populations (reference layout, ghost slots ignored).
 * Ghost densities and boundary state are left as the last init set them. */
int mcg_set_populations(mcg_solver *s, const float *fr, const float *fb);

/* ---- checkpoint / restart (extension) ------------------------------- */

/* Size in bytes of a checkpoint image. */
int mcg_checkpoint_size(const mcg_solver *s, uint64_t *nbytes);
/* Copy the full device state (populations, boundary state, step counter,
 * statistics) into host memory / restore it.  Bitwise round trip. */
int mcg_checkpoint_save(mcg_solver *s, void *host_buf, uint64_t nbytes);
int mcg_checkpoint_load(mcg_solver *s, const void *host_buf, uint64_t nbytes);

/* ---- adjoint (extension: the reference has no adjoint) ---------------- */
/* Discrete adjoint of the step, FP32 populations, sponge_K 0, for the
 * periodic-y family (y_mode periodic, wrap_gradient 1; also with x/z walls,
 * the body force and the per-node viscosity, which needs A_override > 0)
 * and for jet cases (y_mode jet, wrap_gradient 0, single relaxation rate: the
 * inlet and the convective outlet are differentiated and the outlet state is
 * carried between steps).
 * The state is the populations in the reference layout (what
 * mcg_copy_populations returns: f_i(x) after streaming); the parameters are A
 * (the perturbation amplitude, A_override or sigma/(4 cs^4 taub)) and beta,
 * and for jet cases jet_velocity, perturbation_amp and the inlet velocity of
 * every jet column.  Bulk/interface switches, the |F| <= 1e-4 cut and the
 * outlet's clamp are held at the forward state. */

/* Run nsteps forward steps from the current state and keep a copy of the
 * state (populations; jet cases also the inlet/outlet arrays) every
 * checkpoint_every = K steps.  Device memory: ceil(nsteps/K) population sets
 * for the checkpoints, and during mcg_adjoint_backward 4 population sets plus
 * min(K, nsteps) moment sets of 11 floats per node (11/54 of a population
 * set each).  Jet cases: right after an initialisation the first step must
 * follow mcg_init_jet. */
int mcg_adjoint_record(mcg_solver *s, int nsteps, int checkpoint_every);

/* Cap the states mcg_adjoint_record keeps (max_states >= 1; 0, the default,
 * keeps one per segment of K steps).  Below one per segment the record keeps
 * the first state and the split points of the sweep's first descent, and
 * mcg_adjoint_backward recomputes the missing segment states with a binomial
 * schedule (Griewank and Walther's revolve): memory max_states + 2 state sets
 * and K moment sets for any nsteps, at the cost of recomputed forward steps
 * (mcg_adjoint_recomputed_steps).  The gradients are bitwise those of the
 * unlimited record.  Applies to the next mcg_adjoint_record.  No counterpart
 * in the reference (an extension). */
int mcg_adjoint_set_checkpoint_limit(mcg_solver *s, int max_states);

/* Forward steps recomputed by the last mcg_adjoint_backward: nsteps with one
 * state per segment, more under mcg_adjoint_set_checkpoint_limit.  No
 * counterpart in the reference (an extension). */
int mcg_adjoint_recomputed_steps(const mcg_solver *s, int64_t *nsteps);

/* Jet cases: define the u_y functional J_uy = sum over the moments passes of
 * steps first..last (solver step numbers: the pass of the step that takes
 * step n to n+1 is pass n) of sum_x w_uy(x) u_y(x) (w_uy: NX*NY*NZ floats, x
 * fastest; NULL clears it).  Ordinary steps from then on accumulate its value
 * in FP64 (mcg_uy_functional), and mcg_adjoint_backward adds it to the
 * objective: J = <lam, f_final> + J_uy.  Kept until cleared or the solver is
 * destroyed; setting it restarts the sum.  No counterpart in the reference
 * (an extension: the objective of the inverse jet problem). */
int mcg_adjoint_set_uy_seed(mcg_solver *s, const float *w_uy, int64_t first, int64_t last);

/* The value of the u_y functional accumulated by the steps taken since
 * mcg_adjoint_set_uy_seed (FP64; synchronises the stream).  No counterpart in
 * the reference (an extension). */
int mcg_uy_functional(mcg_solver *s, double *value);

/* Jet cases: add du to the inlet velocity of every jet column at every step
 * (du: 3 x NX*NZ floats, component-major u_x, u_y, u_z planes, x fastest;
 * entries of non-jet columns are ignored; NULL removes it).  The control of
 * the inverse jet problem: mcg_adjoint_inlet_gradient returns dJ/d(du).  A
 * solver setting, not part of a checkpoint.  No counterpart in the reference
 * (an extension of jet/boundary_conditions.cuh:45-47). */
int mcg_set_inlet_offset(mcg_solver *s, const float *du);

/* Jet cases, after mcg_adjoint_backward: dparams = {dJ/djet_velocity,
 * dJ/dperturbation_amp} (FP64 sums) and dinlet = dJ/de for the inlet velocity
 * e of each jet column summed over the recorded steps (3 x NX*NZ doubles,
 * component-major, zero outside the jet).  Either may be NULL.  No counterpart
 * in the reference (an extension). */
int mcg_adjoint_inlet_gradient(const mcg_solver *s, double *dparams, double *dinlet);

/* Given lam = dJ/df at the final state (reference layout, 27 x NX*NY*NZ per
 * colour; ghost slots ignored), return dJ/df at the recorded initial state
 * (ghost slots 0) and dparams = {dJ/dA, dJ/dbeta} summed in FP64.  Each
 * segment is recomputed from its checkpoint; statistics are not touched and
 * the final state is restored. */
int mcg_adjoint_backward(mcg_solver *s, const float *lam_fr, const float *lam_fb, float *out_lam_fr, float *out_lam_fb,
                         double *dparams);

/* The FP64 parameter gradients of the last mcg_adjoint_backward sweep: the
 * first n (<= 5) of {dJ/dA, dJ/dbeta, dJ/dgravity, dJ/djet_velocity,
 * dJ/dperturbation_amp} (the last two are zero outside jet cases).  No
 * counterpart in the reference (an extension). */
int mcg_adjoint_gradients(const mcg_solver *s, double *out, int n);

/* Device milliseconds of the last mcg_adjoint_backward sweep (segment
 * recomputation and adjoint kernels, HIP events on the solver's stream;
 * excludes the host transfers of the seed and the result). */
int mcg_adjoint_last_ms(const mcg_solver *s, double *ms);

/* Free the adjoint's checkpoints and buffers. */
int mcg_adjoint_release(mcg_solver *s);

/* The same forward and adjoint on the CPU in FP64 (host code, no GPU), for
 * y_mode periodic (also with x/z walls, per-node viscosity and the body
 * force): nsteps steps from (fr0, fb0), final state into (out_fr, out_fb),
 * and, when lam_fr/lam_fb are given, the adjoint at the initial state and
 * dparams = {dJ/dA, dJ/dbeta, dJ/dgravity}.  fields0 (5 x NX*NY*NZ: rho_r,
 * rho_b, u of the initial state, ghosts included) gives the frozen ghost psi
 * of the walls; it may be NULL without walls. */
int mcg_cpu_adjoint(const mcg_case *c, int nsteps, const double *fr0, const double *fb0, const double *fields0,
                    const double *lam_fr, const double *lam_fb, double *out_fr, double *out_fb, double *out_lam_fr,
                    double *out_lam_fb, double *dparams);

/* The FP64 CPU forward and adjoint for jet cases (y_mode jet, reference
 * colour gradient, single relaxation rate): the inlet and the convective
 * outlet are differentiated, the outlet state (rho_r, rho_b, u on the plane
 * NY-1) is carried between steps.  fields0 = rho_r, rho_b, u_x, u_y, u_z of the
 * initial state over the full grid (5 x NX*NY*NZ, ghosts included: they give
 * the constant ghost psi and the initial outlet state); the steps are numbered
 * from 0.  inlet_du (3 x NX*NZ, may be NULL) is added to the inlet velocity
 * as by mcg_set_inlet_offset.  J = <lam, f_final> + sum over the moments
 * passes of steps uy_first..uy_last of sum_x w_uy(x) u_y(x) (w_uy may be NULL).  dparams = {dJ/dA, dJ/dbeta, dJ/djet_velocity,
 * dJ/dperturbation_amp}; dinlet = dJ/de for the inlet velocity e of each jet
 * column summed over steps (3 x NX*NZ, component-major).  No counterpart in the
 * reference (an extension). */
int mcg_cpu_adjoint_jet(const mcg_case *c, int nsteps, const double *fr0, const double *fb0, const double *fields0,
                        const double *inlet_du, const double *lam_fr, const double *lam_fb, const double *w_uy, int uy_first, int uy_last,
                        double *out_fr, double *out_fb, double *out_lam_fr, double *out_lam_fb, double *dparams,
                        double *dinlet);

/* ---- measurement ------------------------------------------------------ */

/* Run nsteps steps bracketed by HIP events on the solver's stream and return
 * the elapsed milliseconds (pure step launches, no host sync inside). */
int mcg_time_steps(mcg_solver *s, int nsteps, double *ms);

/* The device's achievable HBM bandwidth for a plain stream over nbytes of
 * device memory (allocated and freed by the call, on the current device and
 * the null stream): kind 0 reads, 1 writes, 2 copies to a second buffer; each
 * lane moves vec_bytes (4, 8 or 16) per access with 4 accesses in flight, over
 * blocks_per_cu x (number of CUs) workgroups of 256 threads; nontemporal != 0
 * streams the stores past L2.  *gbs = the best of reps timed launches after a
 * warm-up, in GB/s of bytes moved (copies count read + write).  The ceiling
 * the step kernels' traffic is compared with.  No counterpart in the
 * reference (an extension). */
int mcg_hbm_probe(int kind, int vec_bytes, int nontemporal, int blocks_per_cu, uint64_t nbytes, int reps, double *gbs);

/* Bytes of device memory held by the solver. */
int mcg_device_bytes(const mcg_solver *s, uint64_t *nbytes);

/* ---- output files (the reference's layout; host code, no GPU needed) --- */

/* Create <root>/JET_VTK/Re<Re>_We<We>/{mean_profiles,vti_data} and return the
 * three paths (make_case_folder_name / ensure_case_directories,
 * meanF/meanFieldsRT.cuh:36-69).  Any output pointer may be NULL. */
int mcg_output_dirs(const mcg_case *c, const char *root, char *case_dir, char *mean_dir, char *vti_dir, size_t cap);

/* case_info.txt (write_case_metadata, meanF/meanFieldsRT.cuh:93-118). */
int mcg_write_case_info(const mcg_case *c, const char *case_dir);

/* VTK XML ImageData file of rho_r, rho_b, rho, u over the full grid from host
 * arrays (write_vti, io/save_data.cuh:137-200). */
int mcg_write_vti_arrays(const char *path, int32_t NX, int32_t NY, int32_t NZ, const float *rhor, const float *rhob,
                         const float *ux, const float *uy, const float *uz);

/* radial_sum_*.bin, radial_count.bin and radial_profile_metadata.txt from
 * host arrays (write_radial_profile_outputs / write_radial_profile_metadata,
 * meanF/meanFieldsRT.cuh:229-364).  sum_phi / sum_phi2 may be NULL. */
int mcg_write_radial_files(const mcg_case *c, const char *mean_dir, const double *sum_uy, const double *sum_uy2,
                           const double *sum_ur, const double *sum_ur2, const double *sum_uruy, const uint64_t *count,
                           const double *sum_phi, const double *sum_phi2, int32_t started, int64_t start_step,
                           int64_t samples);

/* Running TKE average and delta (update_tke_average kernels.cu:154-166,
 * update_tke_state meanF/meanFieldsRT.cuh:191-227) in FP32.
 * state = {tke_avg, tke_avg_prev, first_sample_flag}; initialise to {0,0,1}.
 * out = {tke_total, tke_avg, |delta|} as appended to the tke_*.bin files. */
int mcg_tke_update(float state[3], float tke_total, int64_t step, int32_t noutput, float out[3]);

/* Truncate / append one sample to tke_total.bin, tke_avg.bin, tke_diff.bin
 * (open_mean_fields_files / write_tke_outputs, meanF/meanFieldsRT.cuh:73-91,366-375). */
int mcg_tke_files_reset(const char *mean_dir);
int mcg_append_tke(const char *mean_dir, const float vals[3]);

/* Solver-level writers: copy the state and call the writers above
 * (write_vti_step_device io/save_data.cuh:203-215 writes lbm_%08d.vti). */
int mcg_write_vti(mcg_solver *s, const char *path);
int mcg_write_vti_step(mcg_solver *s, const char *vti_dir, int64_t step);
int mcg_write_radial(mcg_solver *s, const char *mean_dir);

const char *mcg_io_last_error(void);

/* ---- misc ------------------------------------------------------------- */

const char *mcg_last_error(void);
const char *mcg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MCG_H */
