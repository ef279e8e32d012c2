// gpu_shim.cpp — VAMPIRE's accelerator boundary (namespace gpu, hdr/gpu.hpp:33-56) implemented
// on the C ABI of libvmx.so (include/vmx.h).
//
// In a VAMPIRE tree this file replaces src/gpu/initialize.cpp, llg_heun.cpp, finalize.cpp,
// statistics.cpp, transfer.cpp and config.cpp (the #ifdef CUDA / OPENCL dispatchers); data.cpp
// (the flags) and interface.cpp (the gpu: input keys) stay.  Nothing else in VAMPIRE changes:
// sim::integrate_serial already calls gpu::llg_heun() when gpu::acceleration is set
// (src/simulate/sim.cpp:600) and stats::update/reset call gpu::stats::update/reset
// (src/statistics/statistics.cpp:35-37,71-72).  Build: compile with the VAMPIRE headers
// (-I hdr -I src) and link with -lvmx.
//
// Behaviour kept from the reference backend:
//  * gpu::initialize copies the system from the atoms::, mp::, exchange:: and anisotropy::
//    globals once (src/cuda/initialize.cu:43-151); the device spins are authoritative afterwards.
//  * gpu::llg_heun reads sim::temperature, sim::H_vec * sim::H_applied and the Hamiltonian flags
//    at every call (src/cuda/external_fields.cu:52-56) and takes one Heun step.
//  * gpu::stats::get replaces the host statistics with the device ones: reset the averages, then
//    set_magnetization(current, mean sums, counter) (src/cuda/statistics.cu:252-284); the energy
//    statistics, which the reference backend leaves to the CPU, go the same way through
//    set_*_energy and update_mean_counter (src/statistics/energy.cpp:308-390).
//  * terms the device step does not compute are refused in gpu::initialize rather than left out.
//  * errors end the run through err::vexit(), as a failed CUDA call does (cuda_utils.hpp:32-43).
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "atoms.hpp"
#include "create.hpp"
#include "dipole.hpp"
#include "montecarlo/internal.hpp"
#include "errors.hpp"
#include "exchange.hpp"
#include "gpu.hpp"
#include "material.hpp"
#include "random.hpp"
#include "sim.hpp"
#include "stats.hpp"
#include "vio.hpp"
#include "anisotropy/internal.hpp"

#include "simulate/internal.hpp"
#include "spintorque/internal.hpp"

#include "vmx.h"

namespace {

vmx_ctx* g_ctx = nullptr;

struct MagHandle {
    stats::magnetization_statistic_t* stat;
    int mask_size;
    int handle;
};
std::vector<MagHandle> g_mags;

struct EnergyHandle {
    stats::energy_statistic_t* stat;
    int mask_size;
    int handle;
};
std::vector<EnergyHandle> g_energies;   // energy statistics computed on the device

// Thin-film demagnetising field (sim::ext_demag, src/simulate/fields.cpp:219-239): a uniform
// field -mu0 D_k m_k / V from the latest system magnetisation statistic, added to the applied
// field of the step (not to the applied-field energy).  g_sys_m holds the m_l the reference reads
// with stats::system_magnetization.get_magnetization(); in device-statistics mode it is refreshed
// from the device at every statistics update.
double g_sys_m[3] = {0.0, 0.0, 0.0};
double g_demag_sent[3] = {0.0, 0.0, 0.0};

void check(int rc, const char* what) {
    if (rc != VMX_OK) {
        std::cerr << "Error: GPU backend call " << what << " failed: " << vmx_last_error() << std::endl;
        zlog << zTs() << "Error: GPU backend call " << what << " failed: " << vmx_last_error() << std::endl;
        err::vexit();
    }
}

uint32_t field_flags() {
    uint32_t f = 0;
    if (sim::hamiltonian_simulation_flags[3] == 1) f |= VMX_FIELD_THERMAL;
    if (sim::hamiltonian_simulation_flags[2] == 1) f |= VMX_FIELD_APPLIED;
    if (sim::local_applied_field) f |= VMX_FIELD_LOCAL_APPLIED;
    return f;
}

void push_demag() {
    if (!sim::ext_demag) return;
    const double mu_0 = -4.0 * M_PI * 1.0e-7 /
                        (cs::system_dimensions[0] * cs::system_dimensions[1] * cs::system_dimensions[2] * 1.0e-30);
    const double HD[3] = {mu_0 * sim::demag_factor[0] * g_sys_m[0], mu_0 * sim::demag_factor[1] * g_sys_m[1],
                          mu_0 * sim::demag_factor[2] * g_sys_m[2]};
    if (HD[0] == g_demag_sent[0] && HD[1] == g_demag_sent[1] && HD[2] == g_demag_sent[2]) return;
    check(vmx_set_demag_field(g_ctx, -1, HD), "vmx_set_demag_field");
    for (int k = 0; k < 3; ++k) g_demag_sent[k] = HD[k];
}

void push_conditions() {
    const double H[3] = {sim::H_vec[0] * sim::H_applied, sim::H_vec[1] * sim::H_applied,
                         sim::H_vec[2] * sim::H_applied};
    push_demag();
    check(vmx_set_conditions(g_ctx, -1, sim::temperature, H, field_flags()), "vmx_set_conditions");
}

// m_l of the thin-film demag field from the host statistic (host-statistics mode, or before
// the first device update)
void host_system_m() {
    if (!sim::ext_demag) return;
    const std::vector<double> m = ::stats::system_magnetization.get_magnetization();
    for (int k = 0; k < 3; ++k) g_sys_m[k] = m.size() > 2 ? m[k] : 0.0;
}

// Constraint angles last sent to the device (none yet: -1e300) and the device outcome totals
// already added to the montecarlo::cmc counters.
double g_cmc_sent[2] = {-1.0e300, -1.0e300};
double g_cmc_counted[3] = {0.0, 0.0, 0.0};

// montecarlo::cmc counters take the device's outcome totals; the reference prints them at the end
// of sim::run (sim.cpp:505-515), so they are brought up to date at every statistics update.
void sync_cmc_counts() {
    if (!g_ctx || g_cmc_sent[0] == -1.0e300) return;
    double c[3];
    check(vmx_cmc_get_counts(g_ctx, c), "vmx_cmc_get_counts");
    const double d0 = c[0] - g_cmc_counted[0], d1 = c[1] - g_cmc_counted[1], d2 = c[2] - g_cmc_counted[2];
    montecarlo::cmc::mc_success += d0;
    montecarlo::cmc::energy_reject += d1;
    montecarlo::cmc::sphere_reject += d2;
    montecarlo::cmc::mc_total += d0 + d1 + d2;
    for (int k = 0; k < 3; ++k) g_cmc_counted[k] = c[k];
}

void refuse(bool cond, const char* what) {
    if (!cond) return;
    std::cerr << "Error: the GPU backend does not compute " << what << "; run this input with "
              << "gpu:acceleration=off" << std::endl;
    zlog << zTs() << "Error: the GPU backend does not compute " << what << std::endl;
    err::vexit();
}

void add_magnetization(stats::magnetization_statistic_t& st, bool enabled) {
    if (!enabled || !st.is_initialized()) return;
    std::vector<int> mask;
    std::vector<double> saturation;
    st.get_mask(mask, saturation);
    const int mask_size = (int)saturation.size();
    int h = -1;
    check(vmx_stats_add_magnetization(g_ctx, mask_size, mask.data(), &h), "vmx_stats_add_magnetization");
    g_mags.push_back(MagHandle{&st, mask_size, h});
}

void add_energy(stats::energy_statistic_t& st, bool enabled) {
    if (!enabled || !st.is_initialized()) return;
    std::vector<int> mask;
    std::vector<double> normalisation;
    st.get_mask(mask, normalisation);
    const int mask_size = (int)normalisation.size();
    int h = -1;
    check(vmx_stats_add_energy(g_ctx, mask_size, mask.data(), &h), "vmx_stats_add_energy");
    g_energies.push_back(EnergyHandle{&st, mask_size, h});
}

// Host-derived statistics, from the host magnetisation and energy objects as the CPU path forms
// them (src/statistics/statistics.cpp:45-55)
void derived_statistics() {
    if (::stats::calculate_system_specific_heat)
        ::stats::system_specific_heat.calculate(::stats::system_energy.get_total_energy());
    if (::stats::calculate_material_specific_heat)
        ::stats::material_specific_heat.calculate(::stats::material_energy.get_total_energy());
    if (::stats::calculate_material_standard_deviation)
        ::stats::material_standard_deviation.update(::stats::system_magnetization.get_magnetization());
    if (::stats::calculate_system_susceptibility)
        ::stats::system_susceptibility.calculate(::stats::system_magnetization.get_magnetization());
    if (::stats::calculate_material_susceptibility)
        ::stats::material_susceptibility.calculate(::stats::material_magnetization.get_magnetization());
}

bool need_derived() {
    return ::stats::calculate_system_specific_heat || ::stats::calculate_material_specific_heat ||
           ::stats::calculate_material_standard_deviation || ::stats::calculate_system_susceptibility ||
           ::stats::calculate_material_susceptibility;
}

}  // namespace

namespace gpu {

void initialize() {
    const int n = atoms::num_atoms;
    const int nmat = mp::num_materials;
    namespace ai = anisotropy::internal;

    // exchange table in the representation the neighbour list refers to
    const int et = (int)exchange::get_exchange_type();
    std::vector<double> table;
    if (et == 0) {
        for (const zval_t& z : atoms::i_exchange_list) table.push_back(z.Jij);
    } else if (et == 1) {
        for (const zvec_t& z : atoms::v_exchange_list)
            for (int k = 0; k < 3; ++k) table.push_back(z.Jij[k]);
    } else {
        for (const zten_t& z : atoms::t_exchange_list)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) table.push_back(z.Jij[a][b]);
    }
    const int K = et == 0 ? 1 : (et == 1 ? 3 : 9);

    std::vector<vmx_material> mats(nmat);
    for (int m = 0; m < nmat; ++m) {
        const mp::materials_t& a = mp::material[m];
        vmx_material& v = mats[m];
        v.alpha = a.alpha;
        v.gamma_rel = a.gamma_rel;
        v.mu_s_SI = a.mu_s_SI;
        v.one_oneplusalpha_sq = a.one_oneplusalpha_sq;
        v.alpha_oneplusalpha_sq = a.alpha_oneplusalpha_sq;
        v.H_th_sigma = a.H_th_sigma;
        v.temperature_rescaling_alpha = a.temperature_rescaling_alpha;
        v.temperature_rescaling_Tc = a.temperature_rescaling_Tc;
        v.ku2 = ai::enable_uniaxial_second_order ? ai::ku2[m] : 0.0;
        v.ku_vector[0] = ai::ku_vector[m].x;
        v.ku_vector[1] = ai::ku_vector[m].y;
        v.ku_vector[2] = ai::ku_vector[m].z;
        v.kc4 = (ai::enable_cubic_fourth_order || ai::enable_cubic_fourth_order_rotation) ? ai::kc4[m] : 0.0;
        for (int k = 0; k < 3; ++k) {
            v.kc_vector1[k] = ai::mp[m].kc_vector1[k];
            v.kc_vector2[k] = ai::mp[m].kc_vector2[k];
            v.kc_vector3[k] = ai::mp[m].kc_vector3[k];
            v.applied_field_unit_vector[k] = a.applied_field_unit_vector[k];
        }
        v.applied_field_strength = a.applied_field_strength;
    }
    uint32_t ani = 0;
    if (ai::enable_uniaxial_second_order) ani |= VMX_ANISOTROPY_UNIAXIAL_SECOND_ORDER;
    if (ai::enable_cubic_fourth_order) ani |= VMX_ANISOTROPY_CUBIC_FOURTH_ORDER;
    if (ai::enable_cubic_fourth_order_rotation) ani |= VMX_ANISOTROPY_CUBIC_FOURTH_ROTATION;
    if (ai::enable_uniaxial_fourth_order || ai::enable_uniaxial_sixth_order || ai::enable_cubic_sixth_order ||
        ai::enable_neel_anisotropy || ai::enable_lattice_anisotropy || exchange::biquadratic) {
        std::cerr << "Error: the GPU backend supports second-order uniaxial and fourth-order cubic anisotropy "
                     "and bilinear exchange only" << std::endl;
        err::vexit();
    }
    // terms of calculate_spin_fields / calculate_external_fields (src/simulate/fields.cpp:111-114,
    // 137-166, 264) that the device step leaves out: a run with them would silently differ
    refuse(sim::lagrange_multiplier, "the Lagrange-multiplier field (sim:constraint-angle / lagrange programs)");
    refuse(sim::internal::enable_spin_torque_fields, "spin-torque fields");
    refuse(st::internal::enabled, "the spin-transfer-torque field (spin-torque module)");
    refuse(sim::enable_fmr, "the FMR field");
    refuse(sim::program == 7, "the HAMR fields (sim:program=hamr-simulation)");
    refuse(sim::program == 13, "localised thermal fields (sim:program=localised-temperature-pulse)");
    refuse(sim::local_temperature, "material-specific temperatures (material:temperature)");
    vmx_system s{};
    s.num_atoms = n;
    s.num_materials = nmat;
    s.type = atoms::type_array.data();
    s.moment = atoms::m_spin_array.data();
    s.nbr_start = atoms::neighbour_list_start_index.data();
    s.nbr_end = atoms::neighbour_list_end_index.data();
    s.num_nbr = (int64_t)atoms::neighbour_list_array.size();
    s.nbr = atoms::neighbour_list_array.data();
    s.nbr_iid = atoms::neighbour_interaction_type_array.data();
    s.exchange_type = et;
    s.num_interactions = (int64_t)(table.size() / K);
    s.exchange = table.data();
    s.materials = mats.data();
    s.anisotropy_flags = ani;
    s.num_replicas = 1;
    s.replica_offset = nullptr;

    vmx_options o{};
    // Arithmetic of the device step.  The reference backend picks it when it is compiled
    // (CUDA_DP / CUDA_SP / CUDA_MP, src/cuda/makefile:36-38); the drop-in reads it at run time from
    // VMX_PRECISION: double (default) or single / mixed (spins and fields in float, statistics and
    // energies accumulated in double).
    o.precision = VMX_FP64;
    if (const char* p = std::getenv("VMX_PRECISION")) {
        const std::string v(p);
        if (v == "single" || v == "mixed" || v == "fp32") o.precision = VMX_FP32;
        else if (v != "double" && v != "fp64") {
            std::cerr << "Error: VMX_PRECISION must be double, single or mixed (got " << v << ")" << std::endl;
            err::vexit();
        }
    }
    o.kernel = VMX_KERNEL_FUSED;
    // gpu:num-threads becomes the workgroup size when it is one the kernels accept
    o.block_size = (gpu::num_threads >= 64 && gpu::num_threads <= 512 && gpu::num_threads % 64 == 0)
                       ? gpu::num_threads : 0;
    o.device = gpu::device;
    o.seed = (uint64_t)(uint32_t)mtrandom::integration_seed;
    o.topology = VMX_TOPOLOGY_AUTO;

    check(vmx_create(&s, &o, &g_ctx), "vmx_create");
    check(vmx_set_time_step(g_ctx, mp::dt, mp::half_dt), "vmx_set_time_step");
    // The Philox counters (Heun and midpoint steps, Metropolis sweeps) start at sim::time.  A
    // fresh run starts at 0; with sim:load-checkpoint=continue, load_checkpoint has restored
    // sim::time before this call (src/utility/checkpoint.cpp:176-178, called at
    // src/simulate/sim.cpp:246), so the resumed run continues the noise stream of the run that
    // wrote the file, as the reference does by restoring its MT19937 state (checkpoint.cpp:166).
    check(vmx_set_step(g_ctx, (int64_t)sim::time), "vmx_set_step");
    check(vmx_set_sweep(g_ctx, (int64_t)sim::time), "vmx_set_sweep");
    check(vmx_set_spins(g_ctx, atoms::x_spin_array.data(), atoms::y_spin_array.data(), atoms::z_spin_array.data()),
          "vmx_set_spins");
    check(vmx_mc_set_sigma(g_ctx, -1, montecarlo::internal::adaptive_sigma), "vmx_mc_set_sigma");
    g_cmc_sent[0] = g_cmc_sent[1] = -1.0e300;
    g_cmc_counted[0] = g_cmc_counted[1] = g_cmc_counted[2] = 0.0;
    // VMX_AUTOTUNE=1: choose the workgroup size and XCD mapping by timing a few steps of this
    // system (spins and step counter are restored afterwards)
    if (const char* a = std::getenv("VMX_AUTOTUNE")) {
        if (std::string(a) == "1") {
            int chosen = 0;
            check(vmx_autotune(g_ctx, 10, &chosen), "vmx_autotune");
            zlog << zTs() << "GPU backend autotuned: " << chosen << " threads per workgroup" << std::endl;
        }
    }
    if (sim::ext_demag && !::stats::system_magnetization.is_initialized()) {
        std::cerr << "Error: the thin-film demagnetising field needs the system magnetisation statistic" << std::endl;
        err::vexit();
    }
    g_demag_sent[0] = g_demag_sent[1] = g_demag_sent[2] = 0.0;
    host_system_m();
    push_conditions();
    // dipole::initialize has computed the time-0 field on the host already (sim.cpp:270-295)
    transfer_dipole_fields_from_cpu_to_gpu();

    // statistics computed on the device unless the input asks for CPU statistics
    g_mags.clear();
    g_energies.clear();
    if (!gpu::cpu_stats) {
        add_magnetization(::stats::system_magnetization, ::stats::calculate_system_magnetization);
        add_magnetization(::stats::material_magnetization, ::stats::calculate_material_magnetization);
        add_magnetization(::stats::height_magnetization, ::stats::calculate_height_magnetization);
        add_magnetization(::stats::material_height_magnetization,
                          ::stats::calculate_material_height_magnetization);
        // energies on the device too, unless the host dipole solver's field is in H_ext: that field
        // carries no demag field for the magnetostatic energy, so those runs form energies on the
        // host from synchronised spins
        if (!dipole::activated) {
            add_energy(::stats::system_energy, ::stats::calculate_system_energy);
            add_energy(::stats::material_energy, ::stats::calculate_material_energy);
        }
    }
    gpu::initialized = true;
    zlog << zTs() << "GPU backend initialised: " << vmx_version() << ", " << n << " atoms" << std::endl;
}

void llg_heun() {
    push_conditions();
    check(vmx_llg_heun(g_ctx, 1, nullptr), "vmx_llg_heun");
}

// ---- integrators the reference runs on the host only ------------------------------------------
// sim::integrate_serial has no accelerator dispatch for integrator 1 (Monte Carlo) or 2 (LLG
// midpoint) (src/simulate/sim.cpp:608-621).  A VAMPIRE build that wants them on the device
// declares these two in hdr/gpu.hpp and dispatches beside the Heun step (sim.cpp:600):
//     if(gpu::acceleration) gpu::mc_step();      else montecarlo::mc_step(...);
//     if(gpu::acceleration) gpu::llg_midpoint(); else sim::LLG_Midpoint();
// One call is one sweep of N trial moves (colour-ordered on the device, random-site on the host,
// montecarlo::mc_step, src/montecarlo/mc.cpp:29-127) or one midpoint step.
void mc_step() {
    if (montecarlo::internal::algorithm != montecarlo::internal::adaptive) {
        std::cerr << "Error: the GPU backend implements the adaptive-cone Monte Carlo move only" << std::endl;
        err::vexit();
    }
    push_conditions();
    check(vmx_mc_sweep(g_ctx, 1), "vmx_mc_sweep");
}

// Integrator 3 (sim.cpp:623-628) on the device: one call is one constrained sweep of N pair
// trials (montecarlo::cmc_step, src/montecarlo/cmc.cpp:282-497), colour-paired on the device.
//     if(gpu::acceleration) gpu::cmc_step(); else montecarlo::cmc_step();
// The first call after cmc::is_initialised was cleared runs CMCinit on the host spins (it sets or
// rotates them to the constraint direction, cmc.cpp:205-273) and uploads them; when the constraint
// angles have moved (cmc_anisotropy moves them between temperature loops and re-runs CMCinit on
// the host) the host spins and the new angles go to the device.  The move is
// the reference's montecarlo::internal::mc_move: the adaptive cone at the device's Metropolis
// width (the same width gpu::mc_step adapts) or the temperature-scaled angle move.
void cmc_step() {
    namespace mi = montecarlo::internal;
    if (mi::algorithm != mi::adaptive && mi::algorithm != mi::angle) {
        std::cerr << "Error: the GPU backend implements the adaptive and angle moves of constrained Monte Carlo only"
                  << std::endl;
        err::vexit();
    }
    refuse(dipole::activated, "the dipole field in constrained Monte Carlo");
    const bool moved = sim::constraint_phi != g_cmc_sent[0] || sim::constraint_theta != g_cmc_sent[1];
    bool upload = moved;   // the programs that move the constraint re-run CMCinit on the host spins
                           // first (src/program/cmc_anisotropy.cpp:118): those hold the new start
    if (!montecarlo::cmc::is_initialised) {
        transfer_spin_positions_from_gpu_to_cpu();
        montecarlo::CMCinit();
        upload = true;
    }
    if (upload)
        check(vmx_set_spins(g_ctx, atoms::x_spin_array.data(), atoms::y_spin_array.data(),
                            atoms::z_spin_array.data()),
              "vmx_set_spins");
    if (moved) {
        check(vmx_cmc_set_constraint(g_ctx, -1, sim::constraint_phi, sim::constraint_theta),
              "vmx_cmc_set_constraint");
        g_cmc_sent[0] = sim::constraint_phi;
        g_cmc_sent[1] = sim::constraint_theta;
    }
    check(vmx_cmc_set_move(g_ctx, mi::algorithm == mi::angle ? VMX_CMC_MOVE_ANGLE : VMX_CMC_MOVE_CONE),
          "vmx_cmc_set_move");
    push_conditions();
    check(vmx_cmc_sweep(g_ctx, 1), "vmx_cmc_sweep");
    // with sim:constraint-rotation the next CMCinit rotates the current spins (cmc.cpp:240-262),
    // so the host arrays follow the device after every sweep
    if (sim::constraint_rotation) transfer_spin_positions_from_gpu_to_cpu();
}

void llg_midpoint() {
    push_conditions();
    check(vmx_llg_midpoint(g_ctx, 1, nullptr), "vmx_llg_midpoint");
}

void finalize() {
    if (g_ctx) {
        sync_cmc_counts();
        transfer_spin_positions_from_gpu_to_cpu();
        check(vmx_destroy(g_ctx), "vmx_destroy");
        g_ctx = nullptr;
    }
}

// Without an initialised backend these calls do nothing, like the reference dispatchers in a
// build without CUDA or OpenCL (config output calls synchronise unconditionally).
void transfer_spin_positions_from_gpu_to_cpu() {
    if (!g_ctx) return;
    check(vmx_get_spins(g_ctx, atoms::x_spin_array.data(), atoms::y_spin_array.data(), atoms::z_spin_array.data()),
          "vmx_get_spins");
}

// dipole::calculate_field computes the field on the host from spins copied out by
// transfer_spin_positions_from_gpu_to_cpu and hands it over here (src/dipole/field.cpp:61-80);
// the device adds it to H_ext after the thermal and applied fields, as fields.cpp:291-310 does.
void transfer_dipole_fields_from_cpu_to_gpu() {
    if (!g_ctx || !dipole::activated) return;
    check(vmx_set_atom_field(g_ctx, dipole::atom_dipolar_field_array_x.data(), dipole::atom_dipolar_field_array_y.data(),
                             dipole::atom_dipolar_field_array_z.data()),
          "vmx_set_atom_field");
}

namespace config {
void synchronise() { transfer_spin_positions_from_gpu_to_cpu(); }
}  // namespace config

namespace stats {

// Energies on the host from synchronised spins: only when the host dipole solver's field is in
// H_ext (device energies refuse it, see gpu::initialize) or with host statistics.
void host_energies() {
    if (!g_energies.empty()) return;
    if (!::stats::calculate_system_energy && !::stats::calculate_material_energy) return;
    if (!gpu::cpu_stats) transfer_spin_positions_from_gpu_to_cpu();
    if (::stats::calculate_system_energy)
        ::stats::system_energy.calculate(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array,
                                         atoms::m_spin_array, atoms::type_array, sim::temperature);
    if (::stats::calculate_material_energy)
        ::stats::material_energy.calculate(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array,
                                           atoms::m_spin_array, atoms::type_array, sim::temperature);
}

// The host statistic objects take the device's latest values and running sums: for each
// magnetisation reset_magnetization_averages then set_magnetization(current, sums, counter)
// (src/cuda/statistics.cu:252-284); for each energy reset_averages, set_*_energy(current, sums)
// for the five terms and update_mean_counter(counter) (src/statistics/energy.cpp:308-390).
void sync_from_device() {
    for (const MagHandle& h : g_mags) {
        std::vector<double> cur(4 * h.mask_size), mean(4 * h.mask_size);
        int64_t counter = 0;
        check(vmx_stats_get_magnetization(g_ctx, h.handle, cur.data(), mean.data(), &counter),
              "vmx_stats_get_magnetization");
        h.stat->reset_magnetization_averages();
        h.stat->set_magnetization(cur, mean, (long)counter);
    }
    for (const EnergyHandle& h : g_energies) {
        const int ms = h.mask_size;
        std::vector<double> cur(5 * (size_t)ms), mean(5 * (size_t)ms);
        int64_t counter = 0;
        check(vmx_stats_get_energy(g_ctx, h.handle, cur.data(), mean.data(), &counter), "vmx_stats_get_energy");
        std::vector<double> c[5], m[5];
        for (int t = 0; t < 5; ++t) {
            c[t].assign(cur.begin() + (size_t)t * ms, cur.begin() + (size_t)(t + 1) * ms);
            m[t].assign(mean.begin() + (size_t)t * ms, mean.begin() + (size_t)(t + 1) * ms);
        }
        h.stat->reset_averages();
        h.stat->set_total_energy(c[0], m[0]);
        h.stat->set_exchange_energy(c[1], m[1]);
        h.stat->set_anisotropy_energy(c[2], m[2]);
        h.stat->set_applied_field_energy(c[3], m[3]);
        h.stat->set_magnetostatic_energy(c[4], m[4]);
        h.stat->update_mean_counter((long)counter);
    }
}

// stats::update under acceleration (src/statistics/statistics.cpp:35-37).  Host statistics: the
// four magnetisation statistics and the energies from freshly copied spins, as
// src/cuda/statistics.cu:34-48 does with cpu_stats.  Device statistics: one device update.  Either
// way the statistics the CPU path derives from them (specific heat, susceptibility, standard
// deviation; statistics.cpp:45-55) follow, and the thin-film demag field takes the new system
// magnetisation.
void update() {
    if (!g_ctx) return;
    if (gpu::cpu_stats) {
        transfer_spin_positions_from_gpu_to_cpu();
        host_energies();
        const auto& X = atoms::x_spin_array;
        const auto& Y = atoms::y_spin_array;
        const auto& Z = atoms::z_spin_array;
        const auto& M = atoms::m_spin_array;
        if (::stats::calculate_system_magnetization)
            ::stats::system_magnetization.calculate_magnetization(X, Y, Z, M);
        if (::stats::calculate_material_magnetization)
            ::stats::material_magnetization.calculate_magnetization(X, Y, Z, M);
        if (::stats::calculate_height_magnetization)
            ::stats::height_magnetization.calculate_magnetization(X, Y, Z, M);
        if (::stats::calculate_material_height_magnetization)
            ::stats::material_height_magnetization.calculate_magnetization(X, Y, Z, M);
    } else {
        push_conditions();
        check(vmx_stats_update(g_ctx), "vmx_stats_update");
        host_energies();
        if (need_derived() || sim::ext_demag) sync_from_device();
    }
    derived_statistics();
    host_system_m();
    sync_cmc_counts();
}

// Before output (src/vio/datalog.cpp:281).  With device statistics the host objects take the
// device values.  A continuous checkpoint written in this output step (datalog.cpp:487) reads
// the host spin arrays, so they are synchronised first.
void get() {
    if (!g_ctx) return;
    if (!gpu::cpu_stats) sync_from_device();
    if (sim::save_checkpoint_flag && sim::save_checkpoint_continuous_flag &&
        sim::time % sim::save_checkpoint_rate == 0)
        transfer_spin_positions_from_gpu_to_cpu();
}

// With host statistics the averages live in the host classes, so they are the ones cleared
// (src/cuda/statistics.cu:115-121); with device statistics the device sums are.  Host energies
// and the derived statistics are cleared as the CPU path clears them (statistics.cpp:74-93).
void reset() {
    if (!g_ctx) return;
    if (gpu::cpu_stats) {
        if (::stats::calculate_system_magnetization) ::stats::system_magnetization.reset_magnetization_averages();
        if (::stats::calculate_material_magnetization) ::stats::material_magnetization.reset_magnetization_averages();
        if (::stats::calculate_height_magnetization) ::stats::height_magnetization.reset_magnetization_averages();
        if (::stats::calculate_material_height_magnetization)
            ::stats::material_height_magnetization.reset_magnetization_averages();
    } else {
        check(vmx_stats_reset(g_ctx), "vmx_stats_reset");
    }
    if (::stats::calculate_system_energy) ::stats::system_energy.reset_averages();
    if (::stats::calculate_material_energy) ::stats::material_energy.reset_averages();
    if (::stats::calculate_material_standard_deviation) ::stats::material_standard_deviation.reset_averages();
    if (::stats::calculate_system_specific_heat) ::stats::system_specific_heat.reset_averages();
    if (::stats::calculate_material_specific_heat) ::stats::material_specific_heat.reset_averages();
    if (::stats::calculate_system_susceptibility) ::stats::system_susceptibility.reset_averages();
    if (::stats::calculate_material_susceptibility) ::stats::material_susceptibility.reset_averages();
}

}  // namespace stats

}  // namespace gpu
