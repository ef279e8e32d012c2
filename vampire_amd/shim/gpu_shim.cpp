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
//    set_magnetization(current, mean sums, counter) (src/cuda/statistics.cu:252-284).
//  * errors end the run through err::vexit(), as a failed CUDA call does (cuda_utils.hpp:32-43).
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

#include "atoms.hpp"
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

#include "vmx.h"

namespace {

vmx_ctx* g_ctx = nullptr;

struct MagHandle {
    stats::magnetization_statistic_t* stat;
    int mask_size;
    int handle;
};
std::vector<MagHandle> g_mags;

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

void push_conditions() {
    const double H[3] = {sim::H_vec[0] * sim::H_applied, sim::H_vec[1] * sim::H_applied,
                         sim::H_vec[2] * sim::H_applied};
    check(vmx_set_conditions(g_ctx, -1, sim::temperature, H, field_flags()), "vmx_set_conditions");
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
    o.precision = VMX_FP64;
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
    push_conditions();
    // dipole::initialize has computed the time-0 field on the host already (sim.cpp:270-295)
    transfer_dipole_fields_from_cpu_to_gpu();

    // statistics computed on the device unless the input asks for CPU statistics
    g_mags.clear();
    if (!gpu::cpu_stats) {
        add_magnetization(::stats::system_magnetization, ::stats::calculate_system_magnetization);
        add_magnetization(::stats::material_magnetization, ::stats::calculate_material_magnetization);
        add_magnetization(::stats::height_magnetization, ::stats::calculate_height_magnetization);
        add_magnetization(::stats::material_height_magnetization,
                          ::stats::calculate_material_height_magnetization);
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

void llg_midpoint() {
    push_conditions();
    check(vmx_llg_midpoint(g_ctx, 1, nullptr), "vmx_llg_midpoint");
}

void finalize() {
    if (g_ctx) {
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

// Energy statistics are computed on the host from synchronised spins: energy_statistic_t takes
// device values only without their update counter (hdr/stats.hpp:108-112), so its means could not
// be formed from device sums.
void host_energies() {
    if (!::stats::calculate_system_energy && !::stats::calculate_material_energy) return;
    if (!gpu::cpu_stats) transfer_spin_positions_from_gpu_to_cpu();
    if (::stats::calculate_system_energy)
        ::stats::system_energy.calculate(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array,
                                         atoms::m_spin_array, atoms::type_array, sim::temperature);
    if (::stats::calculate_material_energy)
        ::stats::material_energy.calculate(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array,
                                           atoms::m_spin_array, atoms::type_array, sim::temperature);
}

void update() {
    if (!g_ctx) return;
    if (gpu::cpu_stats) {
        // statistics on the host from freshly copied spins (src/cuda/statistics.cu:34-48)
        transfer_spin_positions_from_gpu_to_cpu();
        if (::stats::calculate_system_magnetization)
            ::stats::system_magnetization.calculate_magnetization(atoms::x_spin_array, atoms::y_spin_array,
                                                                  atoms::z_spin_array, atoms::m_spin_array);
        if (::stats::calculate_material_magnetization)
            ::stats::material_magnetization.calculate_magnetization(atoms::x_spin_array, atoms::y_spin_array,
                                                                    atoms::z_spin_array, atoms::m_spin_array);
        host_energies();
        return;
    }
    push_conditions();
    check(vmx_stats_update(g_ctx), "vmx_stats_update");
    host_energies();
}

void get() {
    if (!g_ctx || gpu::cpu_stats) return;
    for (const MagHandle& h : g_mags) {
        std::vector<double> cur(4 * h.mask_size), mean(4 * h.mask_size);
        int64_t counter = 0;
        check(vmx_stats_get_magnetization(g_ctx, h.handle, cur.data(), mean.data(), &counter),
              "vmx_stats_get_magnetization");
        h.stat->reset_magnetization_averages();
        h.stat->set_magnetization(cur, mean, (long)counter);
    }
}

// With host statistics the averages live in the host classes, so they are the ones cleared
// (src/cuda/statistics.cu:115-121). The host energy averages are cleared in both modes, as the CPU
// path does (src/statistics/statistics.cpp:74-77), since this backend accumulates them on the host.
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
}

}  // namespace stats

}  // namespace gpu
