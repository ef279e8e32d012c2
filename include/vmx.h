/*
 * vmx.h — C ABI of the MI355X spin-dynamics hot path (libvmx.so).
 *
 * The ABI is what VAMPIRE's accelerator boundary (namespace gpu, hdr/gpu.hpp:13-58)
 * needs from a device backend.  gpu::initialize() (src/gpu/initialize.cpp:25) reads the
 * atoms::, mp::, exchange:: and anisotropy:: globals and hands them to vmx_create();
 * gpu::llg_heun() (src/gpu/llg_heun.cpp:23) becomes vmx_set_conditions() + vmx_llg_heun();
 * gpu::stats::update/get/reset (src/gpu/statistics.cpp:25,45,59) become vmx_stats_*;
 * gpu::config::synchronise / transfer_spin_positions_from_gpu_to_cpu (src/gpu/config.cpp:24,
 * src/gpu/transfer.cpp:21) become vmx_get_spins().  The C++ shim that does this marshalling
 * is vampire_amd/shim/gpu_shim.cpp.
 *
 * Conventions
 *  - Every function returns VMX_OK (0) or a non-zero status; vmx_last_error() gives the text
 *    (thread-local).  The shim maps a non-zero status to err::vexit(), as the reference does
 *    for backend errors (src/gpu/initialize.cpp:33-40, src/cuda/cuda_utils.hpp:32-43).
 *  - All pointers are host pointers unless a name ends in _dev.  Nothing passed in is retained
 *    after the call returns: the library copies what it needs into device memory it owns
 *    (allocated in vmx_create, released in vmx_destroy).
 *  - One context is used from one host thread at a time.  Work is queued on the context's
 *    stream; functions that return data to the host synchronise that stream.
 *  - Units follow the reference: fields in tesla, exchange J_ij/mu_s (T), anisotropy K/mu_s
 *    (T), reduced time step dt = dt_SI * gamma (src/main/initialise_variables.cpp:355-356).
 */
#ifndef VMX_H
#define VMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
enum {
    VMX_OK = 0,
    VMX_ERR_ARG = 1,     /* invalid argument (size, index out of range, null pointer)          */
    VMX_ERR_HIP = 2,     /* the HIP runtime reported an error                                    */
    VMX_ERR_STATE = 3,   /* call not valid in the context's current state                        */
    VMX_ERR_NOMEM = 4    /* device or host allocation failed                                     */
};

/* ---- exchange representations: values of exchange::exchange_t (hdr/exchange_types.hpp:24-27) */
enum {
    VMX_EXCHANGE_ISOTROPIC = 0,   /* one J per interaction id (zval_t, hdr/exchange.hpp:24-34)  */
    VMX_EXCHANGE_VECTORIAL = 1,   /* diagonal Jxx,Jyy,Jzz (zvec_t, hdr/exchange.hpp:39-49)       */
    VMX_EXCHANGE_TENSORIAL = 2    /* full 3x3 row-major (zten_t, hdr/exchange.hpp:54-67); DMI   */
};

/* ---- arithmetic of the spin and field arrays ------------------------------------------ */
enum {
    VMX_FP64 = 0,   /* spins, fields and integration in double                              */
    VMX_FP32 = 1    /* spins, fields and integration in float; statistics and energies are
                       always accumulated in double                                          */
};

/* ---- integrator kernel organisation ---------------------------------------------------- */
enum {
    VMX_KERNEL_PLAIN = 0,  /* five kernels per Heun step with field, noise and dS arrays in
                              HBM, the dataflow of the reference step (src/simulate/LLGHeun.cpp
                              :160-282, src/cuda/llg_heun.cu:121-208)                        */
    VMX_KERNEL_FUSED = 1   /* two kernels per step: predictor, corrector; spin fields live in
                              registers only; dS and the external field pass from predictor to
                              corrector through HBM as options choose (recompute_euler,
                              keep_external_field); the corrector recomputes the rest        */
};

/* ---- anisotropy terms (flags of anisotropy::internal, src/anisotropy/fields.cpp:31-68) ----- */
#define VMX_ANISOTROPY_UNIAXIAL_SECOND_ORDER 0x1u  /* uniaxial_second_order.cpp:51-98          */
#define VMX_ANISOTROPY_CUBIC_FOURTH_ORDER    0x2u  /* cubic_fourth_order.cpp:34-74             */
#define VMX_ANISOTROPY_CUBIC_FOURTH_ROTATION 0x4u  /* cubic_fourth_order_rotation.cpp:34-100   */

/* ---- Hamiltonian terms of the external field (src/simulate/fields.cpp:120-169) ---------- */
#define VMX_FIELD_THERMAL       0x1u  /* sim::hamiltonian_simulation_flags[3] (fields.cpp:153)  */
#define VMX_FIELD_APPLIED       0x2u  /* sim::hamiltonian_simulation_flags[2] (fields.cpp:156)  */
#define VMX_FIELD_LOCAL_APPLIED 0x4u  /* sim::local_applied_field (fields.cpp:191)              */

/* ---- per-material constants after mp::set_derived_parameters ---------------------------- */
typedef struct vmx_material {
    double alpha;                        /* Gilbert damping                          material.hpp  */
    double gamma_rel;                    /* gyromagnetic ratio relative to gamma_SI               */
    double mu_s_SI;                      /* moment (J/T)                                         */
    double one_oneplusalpha_sq;          /* -gamma_rel/(1+alpha^2)   initialise_variables.cpp:423 */
    double alpha_oneplusalpha_sq;        /* alpha*one_oneplusalpha_sq initialise_variables.cpp:424 */
    double H_th_sigma;                   /* sqrt(2 alpha kB/(mu_s gamma_rel dt))             :431 */
    double temperature_rescaling_alpha;  /* T<Tc: Tc (T/Tc)^alpha   (src/simulate/fields.cpp:269) */
    double temperature_rescaling_Tc;
    double ku2;                          /* Ku/mu_s (T)        anisotropy/initialize.cpp:59-62    */
    double ku_vector[3];                 /* uniaxial easy axis                                   */
    double kc4;                          /* Kc/mu_s (T)        anisotropy/initialize.cpp:74-77    */
    double kc_vector1[3];                /* orthonormal cubic basis (initialize.cpp:102-143)      */
    double kc_vector2[3];
    double kc_vector3[3];
    double applied_field_strength;       /* local applied field (fields.cpp:191-199)              */
    double applied_field_unit_vector[3];
} vmx_material;

/* ---- the system handed over at initialisation ---------------------------------------- */
typedef struct vmx_system {
    int32_t num_atoms;                 /* atoms::num_atoms                                       */
    int32_t num_materials;             /* mp::num_materials                                      */
    const int32_t* type;               /* [num_atoms] atoms::type_array                           */
    const double* moment;              /* [num_atoms] atoms::m_spin_array (mu_B), statistics      */
    const int32_t* nbr_start;          /* [num_atoms] atoms::neighbour_list_start_index, inclusive */
    const int32_t* nbr_end;            /* [num_atoms] atoms::neighbour_list_end_index, inclusive   */
    int64_t num_nbr;                   /* length of the two arrays below                          */
    const int32_t* nbr;                /* atoms::neighbour_list_array                             */
    const int32_t* nbr_iid;            /* atoms::neighbour_interaction_type_array                 */
    int32_t exchange_type;             /* VMX_EXCHANGE_*  (exchange::get_exchange_type())         */
    int64_t num_interactions;          /* entries in the exchange table                           */
    const double* exchange;            /* [num_interactions][1|3|9]: i/v/t_exchange_list          */
    const vmx_material* materials;     /* [num_materials]                                         */
    uint32_t anisotropy_flags;         /* VMX_ANISOTROPY_*                                       */
    int32_t num_replicas;              /* >= 1; independent systems integrated together           */
    const int32_t* replica_offset;     /* [num_replicas+1] atom ranges, or NULL when num_replicas=1 */
} vmx_system;

typedef struct vmx_options {
    int32_t precision;    /* VMX_FP64 or VMX_FP32                                              */
    int32_t kernel;       /* VMX_KERNEL_PLAIN or VMX_KERNEL_FUSED                              */
    int32_t block_size;   /* threads per workgroup, 64..512 in multiples of 64; 0 = 256        */
    int32_t device;       /* HIP device ordinal; -1 = current device                            */
    uint64_t seed;        /* Philox key for thermal noise and Monte Carlo moves                 */
    int32_t topology;     /* VMX_TOPOLOGY_*: device layout of the neighbour list                */
    int32_t recompute_euler; /* fused FP64 kernels: 0 = the predictor keeps the Euler derivative
                             dS for the corrector, as the reference keeps its euler arrays; 1 =
                             the corrector recomputes it (one neighbour gather more, 6 fewer
                             reals of memory traffic per atom).  Results are identical.  FP32
                             keeps nothing: its corrector rebuilds S + dt/2 dS from S and S'. */
    int32_t atom_order;   /* VMX_ATOM_ORDER_*: how the device stores atoms                     */
    int32_t keep_external_field; /* fused kernels: 0 = the corrector regenerates the external
                             (thermal + applied) field from the same Philox counter; 1 = the
                             predictor keeps it for the corrector, as the reference keeps its
                             field arrays (6 more reals of traffic per atom).  Same results.   */
    int32_t fp64_normals; /* VMX_FP32 only, a measurement mode: 1 = each step's thermal normals
                             are drawn by the FP64 generator (the normals of a VMX_FP64 context
                             with the same seed) and rounded to float, one more kernel per step,
                             so FP32 and FP64 runs differ by their arithmetic alone.  0 = the
                             FP32 generator inside the step kernels (the fast path).  Graph
                             replay refuses the mode.                                        */
} vmx_options;

/* Atom storage order on the device.  Invisible to the caller: every array passed in or out is in
 * the caller's order, and Philox counters use the caller's atom index. */
enum {
    VMX_ATOM_ORDER_AUTO = 0,   /* basis-major when the rows repeat with a lattice period        */
    VMX_ATOM_ORDER_CALLER = 1  /* the caller's order                                            */
};

/* Device layout of the neighbour list.  Every layout keeps each row's neighbour order, so all
 * give identical results; they differ only in bytes moved per row. */
enum {
    VMX_TOPOLOGY_AUTO = 0,     /* stencil if the rows fit a small table, else ELL, else CSR      */
    VMX_TOPOLOGY_CSR = 1,      /* row offsets + neighbour ids, as handed over                    */
    VMX_TOPOLOGY_ELL = 2,      /* neighbour ids transposed to [slot][atom] (coalesced)            */
    VMX_TOPOLOGY_STENCIL = 3   /* per-atom id of a (j - i offsets, J ids) stencil: rows of a
                                  periodic lattice cost 2 bytes                                   */
};

typedef struct vmx_ctx vmx_ctx;

/* Layout chosen for a context (VMX_TOPOLOGY_CSR/ELL/STENCIL) and the number of stencils. */
int vmx_get_topology(vmx_ctx* ctx, int32_t* topology, int32_t* num_stencils);

/* Lattice period of the internal basis-major atom order (1: the caller's order). */
int vmx_get_atom_period(vmx_ctx* ctx, int32_t* period);

/* Text of the last error on this thread ("" if none). */
const char* vmx_last_error(void);

/* Library version string. */
const char* vmx_version(void);

/* Number of visible HIP devices (0 when there is no GPU or no driver). */
int vmx_device_count(int32_t* count);

/* Builds a context: validates the system, compresses the neighbour list and exchange table,
 * uploads them and allocates all device state.  Replaces vcuda::initialize
 * (src/cuda/initialize.cu:43-151) behind gpu::initialize (src/gpu/initialize.cpp:25-43). */
int vmx_create(const vmx_system* sys, const vmx_options* opt, vmx_ctx** out);

/* Releases all device memory of the context (gpu::finalize, src/gpu/finalize.cpp:23-32). */
int vmx_destroy(vmx_ctx* ctx);

/* Copies host spins (atoms::{x,y,z}_spin_array) to the device. */
int vmx_set_spins(vmx_ctx* ctx, const double* sx, const double* sy, const double* sz);

/* Copies device spins to the host (gpu::config::synchronise, src/gpu/config.cpp:24;
 * gpu::transfer_spin_positions_from_gpu_to_cpu, src/gpu/transfer.cpp:21). */
int vmx_get_spins(vmx_ctx* ctx, double* sx, double* sy, double* sz);

/* Reduced integration step mp::dt and mp::half_dt (initialise_variables.cpp:355-356). */
int vmx_set_time_step(vmx_ctx* ctx, double dt, double half_dt);

/* Conditions read by the reference at every step (src/cuda/external_fields.cu:52-56):
 * temperature sim::temperature, applied field H = sim::H_vec * sim::H_applied (T), and the
 * VMX_FIELD_* flags.  replica = -1 sets every replica. */
int vmx_set_conditions(vmx_ctx* ctx, int32_t replica, double temperature, const double H[3],
                       uint32_t field_flags);

/* The thin-film demagnetising field of calculate_applied_fields (src/simulate/fields.cpp:219-239,
 * enabled by sim:demagnetisation-factor, sim::ext_demag): a uniform field (T) added to the step's
 * applied field after the global and local ones, and not to the applied-field energy
 * (src/simulate/energy.cpp:94-98).  The reference forms it from the latest system magnetisation
 * statistic; the caller passes the field itself.  Applies while VMX_FIELD_APPLIED is set; zero
 * removes it.  replica = -1 sets every replica.  Not held by checkpoints. */
int vmx_set_demag_field(vmx_ctx* ctx, int32_t replica, const double H[3]);

/* Advances every replica by nsteps stochastic LLG-Heun steps (sim::LLG_Heun,
 * src/simulate/LLGHeun.cpp:160-282; gpu::llg_heun, src/gpu/llg_heun.cpp:23).
 * noise == NULL: thermal noise is generated on the device by Philox-4x32-10 keyed by
 *   (seed, replica-local atom index, step, stream).
 * noise != NULL: host array [nsteps][3][num_atoms] of standard normal deviates, consumed in
 *   the reference's order (x for all atoms, then y, then z: src/simulate/fields.cpp:274-276);
 *   used to feed the device exactly the stream a CPU run consumed. */
int vmx_llg_heun(vmx_ctx* ctx, int32_t nsteps, const double* noise);

/* nsteps LLG midpoint steps (sim::LLG_Midpoint, src/simulate/LLGMidpoint.cpp:87-186; integrator
 * 2 of sim::integrate_serial, src/simulate/sim.cpp:616-621): fields at S, Cayley rotation about
 * F = H + alpha S x H to the unnormalised midpoint, spin field at the midpoint with the same
 * external field, rotation of S.  Noise, step counter and dipole refresh as vmx_llg_heun.  The
 * reference has no device version (LLGMidpoint.cpp:212-216). */
int vmx_llg_midpoint(vmx_ctx* ctx, int32_t nsteps, const double* noise);

/* Records nsteps Heun steps of the fused kernels into a HIP graph once and replays it
 * nrepeat times (launch-bound small systems and replica batches).  Noise is Philox. */
int vmx_llg_heun_graph(vmx_ctx* ctx, int32_t nsteps, int32_t nrepeat);

/* Number of Heun steps taken so far (the Philox step counter). */
int vmx_get_step(vmx_ctx* ctx, int64_t* step);
int vmx_set_step(vmx_ctx* ctx, int64_t step);

/* Number of Metropolis sweeps taken so far (the Philox counter of vmx_mc_sweep).  With
 * vmx_set_step, the device counterpart of the random-generator state that the reference's
 * checkpoint restores (src/utility/checkpoint.cpp:166). */
int vmx_get_sweep(vmx_ctx* ctx, int64_t* sweep);
int vmx_set_sweep(vmx_ctx* ctx, int64_t sweep);

/* Checkerboard Metropolis: nsweeps sweeps; each sweep visits every atom once, one graph colour
 * at a time (colours computed at the first call), with the adaptive Gaussian cone move and
 * acceptance of montecarlo::mc_step (src/montecarlo/mc.cpp:29-127, mc_moves.cpp:71-87,158-162).
 * Temperatures and fields come from vmx_set_conditions. */
int vmx_mc_sweep(vmx_ctx* ctx, int32_t nsweeps);

/* Adaptive cone width per replica (montecarlo::internal::adaptive_sigma) and the move and
 * reject totals (sim::mc_statistics_moves / _reject). */
int vmx_mc_get_state(vmx_ctx* ctx, double* sigma, double* moves, double* rejects);
int vmx_mc_set_sigma(vmx_ctx* ctx, int32_t replica, double sigma);
int vmx_mc_num_colours(vmx_ctx* ctx, int32_t* ncolours);

/* ---- constrained Monte Carlo (montecarlo::cmc_step, src/montecarlo/cmc.cpp:282-497) ----------
 * Pair moves that keep the magnetisation along a constraint direction (integrator 3 of
 * sim::integrate_serial, sim.cpp:623-628): spin 1 takes the Monte Carlo move, spin 2 cancels the
 * change of the transverse moment in the constraint frame, and the pair is accepted with
 * exp(-dE/kT) (Mz'/Mz)^2 |s2z/s2z'| when Mz' > 0.  The reference picks both sites at random; the
 * device pairs atoms of one graph colour (which do not interact) through a random permutation per
 * round, two rounds per colour, so a sweep makes N pair trials as the reference's does, and the
 * acceptance uses the magnetisation at the start of the round.  The reference initialises the
 * spins along the constraint (CMCinit, cmc.cpp:204-278): the caller sets them (vmx_set_spins). */
enum {
    VMX_CMC_MOVE_CONE = 0,    /* Gaussian cone of the replica's Metropolis width (vmx_mc_set_sigma,
                                 60 by default): montecarlo's adaptive move, whose width cmc_step
                                 never adapts (mc_moves.cpp:158-162)                            */
    VMX_CMC_MOVE_ANGLE = 1    /* the angle move with width (kB T/mu_B)^0.2 0.08 per material, 0.02
                                 below 1 K (cmc.cpp:320-325)                                     */
};

/* Constraint direction: polar angle phi and azimuth theta in degrees (sim::constraint_phi,
 * sim::constraint_theta), the frame of cmc::polar_rot_matrix (cmc.cpp:43-138).  replica = -1
 * sets every replica. */
int vmx_cmc_set_constraint(vmx_ctx* ctx, int32_t replica, double phi, double theta);

/* VMX_CMC_MOVE_CONE (default) or VMX_CMC_MOVE_ANGLE. */
int vmx_cmc_set_move(vmx_ctx* ctx, int32_t move);

/* nsweeps constrained sweeps at the conditions of vmx_set_conditions; advances the sweep counter
 * (the Philox counter shared with vmx_mc_sweep).  Refused while a dipole field is set. */
int vmx_cmc_sweep(vmx_ctx* ctx, int32_t nsweeps);

/* Outcome totals per replica, [num_replicas][3]: accepted pairs, energy rejections, sphere
 * rejections (montecarlo::cmc::mc_success, energy_reject, sphere_reject). */
int vmx_cmc_get_counts(vmx_ctx* ctx, double* counts);

/* ---- statistics (src/statistics, gpu::stats) ------------------------------------------ */
/* Registers a magnetisation statistic with mask[num_atoms] in [0, mask_size)
 * (magnetization_statistic_t::set_mask, src/statistics/magnetization.cpp:53-113); the last
 * mask id collects non-magnetic atoms and is not normalised.  Returns a handle. */
int vmx_stats_add_magnetization(vmx_ctx* ctx, int32_t mask_size, const int32_t* mask,
                                int32_t* handle);

/* Registers an energy statistic (energy_statistic_t, src/statistics/energy.cpp:52-258). */
int vmx_stats_add_energy(vmx_ctx* ctx, int32_t mask_size, const int32_t* mask,
                         int32_t* handle);

/* Computes every registered statistic from the current device spins and adds it to the
 * running means (stats::update, src/statistics/statistics.cpp:26-63). Accumulation is FP64. */
int vmx_stats_update(vmx_ctx* ctx);

/* vmx_stats_update without any allocation: no counterpart in the reference.  For callers that
 * cannot allocate device memory (HIP graph capture, serving loops, harnesses that forbid it).
 * Same kernels and results; the per-workgroup partial sums go to a fixed 4 MiB device array of
 * the library (one per GPU, kept until the process ends) instead of a buffer the library grows.
 * A registered statistic of mask_size > 16 bins needs min(ceil(num_atoms/256), 512) x mask_size
 * x 32 bytes of partials; beyond 4 MiB (only masks of more than 256 bins reach it) the call is
 * refused with VMX_ERR_ARG before any work is queued.  Calls on one GPU share the
 * array: calls that overlap on one GPU (several host threads or streams) race. */
int vmx_stats_update_noalloc(vmx_ctx* ctx);

/* Latest magnetisation [4*mask_size] (m_x,m_y,m_z unit vector, |M|/Ms per mask id: CPU
 * semantics of magnetization.cpp:157-168), running sum [4*mask_size] and update counter
 * (what gpu::stats::get hands to set_magnetization, magnetization.cpp:194-208). */
int vmx_stats_get_magnetization(vmx_ctx* ctx, int32_t handle, double* magnetization,
                                double* mean_sum, int64_t* counter);

/* Latest energies [5][mask_size] in order total, exchange, anisotropy, applied-field,
 * magnetostatic (tesla x mu_B, energy.cpp:141-258), running sums [5][mask_size], counter. */
int vmx_stats_get_energy(vmx_ctx* ctx, int32_t handle, double* energy, double* mean_sum,
                         int64_t* counter);

/* Zeroes running sums and counters of every statistic (gpu::stats::reset). */
int vmx_stats_reset(vmx_ctx* ctx);

/* ---- diagnostics ---------------------------------------------------------------------- */
/* Spin-dependent field H_spin(S) (exchange + anisotropy; fields.cpp:67-118) and external field
 * (thermal from `noise` [3][N] or Philox at the current step, plus applied; fields.cpp
 * :120-289) of the current spins, copied to host [3][num_atoms] arrays (NULL to skip). */
int vmx_compute_fields(vmx_ctx* ctx, const double* noise, double* h_spin, double* h_ext);

/* The Philox standard normals the device uses for thermal noise at `step`, [3][num_atoms]. */
int vmx_thermal_noise(vmx_ctx* ctx, int64_t step, double* noise);

/* Blocks until all queued work of the context has finished. */
int vmx_synchronize(vmx_ctx* ctx);

/* Average device time (ms) of the last `vmx_llg_heun` call's kernels per Heun step, measured
 * with HIP events on the context's stream (0 if not measured: a call of one step records no
 * events, since they would cost it more than its own launches; calls of two or more steps and
 * graph replays are timed). */
int vmx_last_step_time_ms(vmx_ctx* ctx, double* ms);

/* Runs nsteps Heun steps (Philox noise) with HIP events around every kernel launch and returns
 * the average device time (ms) of each kernel of the step in `avg_ms` (fused: predictor,
 * corrector; plain: spin fields of S, external fields, predictor, spin fields of S', corrector)
 * and their number in `nkernels`.  `avg_ms` must hold 5 values. */
int vmx_time_kernels(vmx_ctx* ctx, int32_t nsteps, double* avg_ms, int32_t* nkernels);

/* ---- differentiation: deterministic trajectory and its adjoint ---------------------------- */
/* The T = 0 Heun step of sim::LLG_Heun (src/simulate/LLGHeun.cpp:160-282) with explicit
 * parameters, for a context of one replica, one material, isotropic uniform exchange on a
 * symmetric neighbour graph.  The field is H_i = J sum_j S_j + 2 ku2 (S_i.e) e + H + h_i
 * (e = the material's uniaxial axis, h_i an optional per-site field). */
typedef struct vmx_fit_params {
    double J;          /* exchange J_ij/mu_s of every bond (T)                              */
    double ku2;        /* uniaxial Ku/mu_s (T)                                              */
    double alpha;      /* Gilbert damping; gamma_rel is the material's                       */
    double H[3];       /* uniform applied field (T)                                         */
} vmx_fit_params;

/* Runs nsteps steps from spins0 [3][N] and writes <m_z> after every record_every steps to
 * mz [nsteps/record_every].  Keeps a checkpoint every checkpoint_every steps for the adjoint.
 * h: per-site field [3][N] or NULL.  nsteps must be a multiple of record_every. */
int vmx_fit_forward(vmx_ctx* ctx, const vmx_fit_params* p, const double* h, const double* spins0,
                    int32_t nsteps, int32_t record_every, int32_t checkpoint_every, double* mz);

/* Discrete adjoint of the last vmx_fit_forward: given dL/d<m_z> [nsteps/record_every], writes
 * grad = {dL/dJ, dL/dku2, dL/dalpha} and, if dL_dh is not NULL, dL/dh_i [3][N].  Segments between
 * checkpoints are recomputed forward, then swept backward (hand-written adjoint kernels). */
int vmx_fit_backward(vmx_ctx* ctx, const double* dL_dmz, double* grad, double* dL_dh);

/* Chooses the workgroup size (64, 128, 256, 512) and whether workgroups are remapped to give
 * each XCD one contiguous stretch of the system, by timing a few Heun steps of each candidate
 * on this system (the better of two trials of steps_per_trial steps); the result applies to
 * later calls.  The spins and step counter are restored
 * afterwards.  Writes the chosen block size.  (Replaces the fixed gpu::num_threads launch
 * width of the reference backend, hdr/gpu.hpp:22.) */
int vmx_autotune(vmx_ctx* ctx, int32_t steps_per_trial, int32_t* chosen_block_size);

/* ---- checkpoints: resume a run with the same spins, Philox counters and Metropolis state ------
 * Replaces, for the device state, sim::save_checkpoint / load_checkpoint (src/utility/
 * checkpoint.cpp:27-207), which write the host spin arrays and the MT19937 state; here the random
 * state is the Heun step and Metropolis sweep counters with the seed (Philox keys), and the spins
 * come from the device.  The buffer is self-describing; the caller stores it as it likes.
 * vmx_checkpoint_load needs a context created for the same system, precision and seed. */
int vmx_checkpoint_bytes(vmx_ctx* ctx, uint64_t* bytes);
int vmx_checkpoint_save(vmx_ctx* ctx, void* buf, uint64_t bytes);
int vmx_checkpoint_load(vmx_ctx* ctx, const void* buf, uint64_t bytes);

/* The launch configuration in use: threads per workgroup and 1 when workgroups are remapped
 * per XCD.  No counterpart in the reference; reported by benchmarks next to their timings. */
int vmx_get_launch_config(vmx_ctx* ctx, int32_t* block_size, int32_t* xcd_remap);

/* ---- macrocell dipole field (tensor solver) ----------------------------------------------------
 * The cells and the cell-cell tensor as the reference builds them at set-up (dipole::initialize,
 * src/dipole/initialize.cpp:31-150, tensor.cpp, inter.cpp, intra.cpp); the device then forms the
 * field itself: cell moments (cells::mag, src/cells/mag.cpp:35-77), cell fields
 * (dipole::internal::update_field, src/dipole/update.cpp:37-105) and atom fields
 * (src/dipole/field.cpp:106-126), added to H_ext after the thermal and applied fields
 * (fields.cpp:166, 291-310).  The reference's device backend leaves this to the host and copies
 * spins out and fields in at every update (field.cpp:61, 80). */
typedef struct vmx_dipole {
    int32_t num_cells;             /* cells::num_cells                                          */
    int32_t num_local_cells;       /* cells with magnetic atoms: rows of the tensor              */
    const int32_t* local_cell;     /* [num_local_cells] cells::cell_id_array                     */
    const int32_t* atom_cell;      /* [num_atoms] cells::atom_cell_id_array (caller's order)      */
    const int32_t* magnetic;       /* [num_materials] 0 for mp::material[m].non_magnetic; NULL:
                                      every material magnetic                                    */
    const double* volume;          /* [num_cells] cells::volume_array (A^3), self-demag term      */
    const double* tensor;          /* [6][num_local_cells][num_cells]: rij_tensor_xx, _xy, _xz,
                                      _yy, _yz, _zz [local cell][cell] (dipole/internal.hpp)      */
    int32_t update_rate;           /* dipole::update_rate: steps (Heun steps or Metropolis
                                      sweeps) between refreshes                                  */
} vmx_dipole;

/* Sets the dipole solver and computes the field of the current spins (time 0, as
 * dipole::initialize does, initialize.cpp:143); vmx_llg_heun and vmx_mc_sweep then refresh it
 * after every step or sweep whose count since this call is a multiple of update_rate
 * (dipole::calculate_field after sim::increment_time, field.cpp:43-71, sim.cpp:191-200).  The
 * Metropolis energy and the energy statistics take the magnetostatic term -mu0Hd.S
 * (simulate/energy.cpp:123-125, 171; statistics/energy.cpp:198-208).  NULL removes the field.
 * One replica only; graph replay, checkpoints and the fitting step refuse while a dipole field
 * is set. */
int vmx_set_dipole(vmx_ctx* ctx, const vmx_dipole* dipole);

/* Recomputes the field from the current spins now (e.g. after vmx_set_spins). */
int vmx_update_dipole(vmx_ctx* ctx);

/* Per-atom fields of the last refresh, caller's order (any pointer may be NULL): B
 * (dipole::atom_dipolar_field_array_*, the field in H_ext) and mu0 Hd
 * (dipole::atom_mu0demag_field_array_*, the field of the magnetostatic energy). */
int vmx_get_dipole_field(vmx_ctx* ctx, double* bx, double* by, double* bz, double* hx, double* hy,
                         double* hz);

/* A per-atom field (tesla, caller's order) added to H_ext where the dipole field goes, for callers
 * that compute the dipole field themselves: gpu::transfer_dipole_fields_from_cpu_to_gpu
 * (src/gpu/transfer.cpp:39, called from dipole::calculate_field, field.cpp:80) uploads
 * dipole::atom_dipolar_field_array_* with it.  Three NULLs remove it.  It carries no demag field,
 * so Metropolis sweeps and device energy statistics refuse while it is set. */
int vmx_set_atom_field(vmx_ctx* ctx, const double* hx, const double* hy, const double* hz);

#ifdef __cplusplus
}
#endif

#endif /* VMX_H */
