// ref_harness.cpp — drives the reference VAMPIRE CPU code (compiled from its own sources by
// oracle/ref/Makefile) to produce golden vectors for the oracle.  Development container only.
//
//   ref_harness <rundir> dump                       system after cs::create + sim::run set-up
//   ref_harness <rundir> llg <nsteps> <every> [mid|demag]
//                                                   sim::LLG_Heun (sim::LLG_Midpoint with `mid`) with
//                                                   injected noise; `demag`: statistics updated before
//                                                   every step, so the thin-film demag field follows the
//                                                   spins, and that field written per step (demag.bin)
//   ref_harness <rundir> mc <nsweeps> <T> <every>   montecarlo::mc_step, reference RNG
//   ref_harness <rundir> cmc <nsweeps> <T> <every> <phi> <theta>
//                                                   montecarlo::cmc_step (constrained), reference RNG
//   ref_harness <rundir> dipole                     macrocell dipole set-up and fields (dipole.bin)
//   ref_harness_vmx <rundir> gpullg <nsteps> <every> <heun|mid> <seed>
//   ref_harness_vmx <rundir> gpumc <nsweeps> <T> <every>
//   ref_harness_vmx <rundir> gpucmc <nsweeps> <T> <every> <phi> <theta> [<phi2> <theta2> <sweep>]
//                                                   the cmc command through the gpu:: drop-in
//                                                   (built with -DVMX_SHIM, linked with the shim)
//
// <rundir> holds `input` (and the material file it names).  Inputs read from <rundir>:
//   spins_in.bin  optional: 3*N doubles (x block, y block, z block) replacing the initial spins
//   noise.bin     llg: nsteps*3*N doubles served by mtrandom::gaussian in call order
// Outputs written to <rundir>: system.bin, fields.bin, traj.bin, stats.bin, mc.bin, meta.txt.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>
#include <unistd.h>

#include "anisotropy.hpp"
#include "atoms.hpp"
#include "cells.hpp"
#include "dipole.hpp"
#include "dipole/internal.hpp"
#include "create.hpp"
#include "exchange.hpp"
#include "material.hpp"
#include "montecarlo.hpp"
#include "random.hpp"
#include "sim.hpp"
#include "stats.hpp"
#include "vio.hpp"
#include "vmpi.hpp"
#ifdef VMX_SHIM
#include "gpu.hpp"
namespace gpu {
// defined by the drop-in shim; the reference's gpu.hpp declares no such calls
void mc_step();
void llg_midpoint();
void cmc_step();
}
#endif

// fields.cpp entry points (declared there, not in a header)
int calculate_spin_fields(const int, const int);
int calculate_external_fields(const int, const int);

namespace mtrandom {
double mt_ziggurat_gaussian();  // the reference Ziggurat, renamed at compile time
}

namespace {
const double* g_noise = nullptr;
size_t g_noise_len = 0, g_noise_pos = 0;
}  // namespace

namespace mtrandom {
// Serves the injected stream while one is installed, the reference generator otherwise.
double gaussian() {
    if (g_noise) {
        if (g_noise_pos >= g_noise_len) {
            std::cerr << "ref_harness: injected noise exhausted" << std::endl;
            std::exit(3);
        }
        return g_noise[g_noise_pos++];
    }
    return mt_ziggurat_gaussian();
}
}  // namespace mtrandom

namespace {

template <typename T>
void put(std::ofstream& f, const std::vector<T>& v) {
    f.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

std::vector<double> read_doubles(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<double> v;
    if (!f) return v;
    f.seekg(0, std::ios::end);
    const size_t bytes = f.tellg();
    f.seekg(0);
    v.resize(bytes / sizeof(double));
    f.read(reinterpret_cast<char*>(v.data()), bytes);
    return v;
}

void setup() {
    mp::initialise("input");
    cs::create();
    // the set-up part of sim::run (src/simulate/sim.cpp:236-267)
    sim::initialize(mp::num_materials);
    montecarlo::initialize();
    anisotropy::initialize(atoms::num_atoms, atoms::type_array, mp::mu_s_array);
    mtrandom::grnd.seed(vmpi::parallel_rng_seed(mtrandom::integration_seed));
    std::vector<bool> nm(mp::num_materials, false);
    stats::initialize(atoms::num_atoms, mp::num_materials, atoms::m_spin_array, atoms::type_array,
                      atoms::category_array, nm);
    std::vector<double> sp = read_doubles("spins_in.bin");
    const int n = atoms::num_atoms;
    if (sp.size() == 3 * (size_t)n) {
        for (int i = 0; i < n; ++i) {
            atoms::x_spin_array[i] = sp[i];
            atoms::y_spin_array[i] = sp[n + i];
            atoms::z_spin_array[i] = sp[2 * n + i];
        }
    }
    // the dipole set-up of sim::run (src/simulate/sim.cpp:270-292); returns at once unless the
    // input activates the dipole field, and computes the field of the spins above at time 0
    dipole::initialize(cells::num_atoms_in_unit_cell, cells::num_cells, cells::num_local_cells,
                       cells::macro_cell_size, cells::local_cell_array, cells::num_atoms_in_cell,
                       cells::num_atoms_in_cell_global, cells::index_atoms_array, cells::volume_array,
                       cells::pos_and_mom_array, cells::atom_in_cell_coords_array_x,
                       cells::atom_in_cell_coords_array_y, cells::atom_in_cell_coords_array_z,
                       atoms::type_array, cells::atom_cell_id_array, atoms::x_coord_array,
                       atoms::y_coord_array, atoms::z_coord_array, atoms::x_spin_array,
                       atoms::y_spin_array, atoms::z_spin_array, atoms::m_spin_array, atoms::num_atoms);
}

// Macrocell dipole data after a field update: sizes in meta_dipole.txt; dipole.bin holds
// local_cell[nl] (int32), atom_cell[n] (int32), num_atoms_in_cell[nc] (int32), volume[nc],
// the six tensors [6][nl][nc], cells::mag[3][nc], the cell B and mu0Hd fields [3][nc] each and the
// atom B and mu0Hd fields [3][n] each (doubles).
void dump_dipole(std::ofstream& f) {
    const int nc = dipole::internal::cells_num_cells, nl = dipole::internal::cells_num_local_cells;
    std::ofstream meta("meta_dipole.txt");
    meta.precision(17);
    meta << "num_cells " << nc << "\nnum_local_cells " << nl << "\nupdate_rate " << dipole::update_rate
         << "\nmacro_cell_size " << cells::macro_cell_size << "\n";
    std::vector<int> lc(cells::cell_id_array.begin(), cells::cell_id_array.begin() + nl);
    put(f, lc);
    put(f, dipole::internal::atom_cell_id_array);
    std::vector<int> cnt(dipole::internal::cells_num_atoms_in_cell.begin(),
                         dipole::internal::cells_num_atoms_in_cell.begin() + nc);
    put(f, cnt);
    std::vector<double> vol(dipole::internal::cells_volume_array.begin(),
                            dipole::internal::cells_volume_array.begin() + nc);
    put(f, vol);
    for (auto* t : {&dipole::internal::rij_tensor_xx, &dipole::internal::rij_tensor_xy, &dipole::internal::rij_tensor_xz,
                    &dipole::internal::rij_tensor_yy, &dipole::internal::rij_tensor_yz, &dipole::internal::rij_tensor_zz})
        for (int l = 0; l < nl; ++l) put(f, (*t)[l]);
    put(f, cells::mag_array_x); put(f, cells::mag_array_y); put(f, cells::mag_array_z);
    put(f, dipole::cells_field_array_x); put(f, dipole::cells_field_array_y); put(f, dipole::cells_field_array_z);
    put(f, dipole::cells_mu0Hd_field_array_x); put(f, dipole::cells_mu0Hd_field_array_y);
    put(f, dipole::cells_mu0Hd_field_array_z);
    put(f, dipole::atom_dipolar_field_array_x); put(f, dipole::atom_dipolar_field_array_y);
    put(f, dipole::atom_dipolar_field_array_z);
    put(f, dipole::atom_mu0demag_field_array_x); put(f, dipole::atom_mu0demag_field_array_y);
    put(f, dipole::atom_mu0demag_field_array_z);
}

void dump_system() {
    const int n = atoms::num_atoms;
    std::ofstream meta("meta.txt");
    meta.precision(17);
    meta << "num_atoms " << n << "\n";
    meta << "num_nbr " << atoms::neighbour_list_array.size() << "\n";
    meta << "exchange_type " << exchange::get_exchange_type() << "\n";
    meta << "num_materials " << mp::num_materials << "\n";
    meta << "dt " << mp::dt << "\nhalf_dt " << mp::half_dt << "\n";
    for (int m = 0; m < mp::num_materials; ++m) {
        const mp::materials_t& a = mp::material[m];
        meta << "material " << m << " alpha " << a.alpha << " gamma_rel " << a.gamma_rel << " mu_s_SI " << a.mu_s_SI
             << " one_oneplusalpha_sq " << a.one_oneplusalpha_sq << " alpha_oneplusalpha_sq "
             << a.alpha_oneplusalpha_sq << " H_th_sigma " << a.H_th_sigma << " ku2 "
             << anisotropy::get_ku2(m) << " mu_s_muB " << mp::mu_s_array[m] << "\n";
    }
    std::ofstream f("system.bin", std::ios::binary);
    put(f, atoms::type_array);
    put(f, atoms::neighbour_list_start_index);
    put(f, atoms::neighbour_list_end_index);
    put(f, atoms::neighbour_list_array);
    put(f, atoms::neighbour_interaction_type_array);
    put(f, atoms::m_spin_array);
    put(f, atoms::x_coord_array);
    put(f, atoms::y_coord_array);
    put(f, atoms::z_coord_array);
    put(f, atoms::x_spin_array);
    put(f, atoms::y_spin_array);
    put(f, atoms::z_spin_array);
    const int et = exchange::get_exchange_type();
    std::vector<double> J;
    if (et == 0) for (auto& z : atoms::i_exchange_list) J.push_back(z.Jij);
    if (et == 1) for (auto& z : atoms::v_exchange_list) for (int k = 0; k < 3; ++k) J.push_back(z.Jij[k]);
    if (et == 2) for (auto& z : atoms::t_exchange_list) for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) J.push_back(z.Jij[a][b]);
    meta << "num_interactions " << J.size() / (et == 0 ? 1 : (et == 1 ? 3 : 9)) << "\n";
    put(f, J);
}

void dump_stats(std::ofstream& f) {
    stats::update(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array, atoms::m_spin_array,
                  atoms::type_array, sim::temperature);
    std::vector<double> v;
    const std::vector<double>& sm = stats::system_magnetization.get_magnetization();
    v.insert(v.end(), sm.begin(), sm.end());
    const std::vector<double>& mm = stats::material_magnetization.get_magnetization();
    v.insert(v.end(), mm.begin(), mm.end());
    for (auto* e : {&stats::system_energy, &stats::material_energy}) {
        if (!e->is_initialized()) continue;
        typedef const std::vector<double>& (stats::energy_statistic_t::*getter_t)();
        const getter_t getters[5] = {&stats::energy_statistic_t::get_total_energy,
                                     &stats::energy_statistic_t::get_exchange_energy,
                                     &stats::energy_statistic_t::get_anisotropy_energy,
                                     &stats::energy_statistic_t::get_applied_field_energy,
                                     &stats::energy_statistic_t::get_magnetostatic_energy};
        for (int g = 0; g < 5; ++g) {
            const std::vector<double>& ev = (e->*getters[g])();
            v.insert(v.end(), ev.begin(), ev.end());
        }
    }
    put(f, v);
}

void dump_fields(std::ofstream& f) {
    calculate_spin_fields(0, atoms::num_atoms);
    calculate_external_fields(0, atoms::num_atoms);
    put(f, atoms::x_total_spin_field_array);
    put(f, atoms::y_total_spin_field_array);
    put(f, atoms::z_total_spin_field_array);
    put(f, atoms::x_total_external_field_array);
    put(f, atoms::y_total_external_field_array);
    put(f, atoms::z_total_external_field_array);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: ref_harness <rundir> dump|llg|mc ..." << std::endl;
        return 2;
    }
    if (chdir(argv[1]) != 0) { std::perror("chdir"); return 2; }
    vmpi::initialise(argc, argv);
    vout::zLogTsInit(std::string("ref_harness"));
    const std::string cmd = argv[2];
    setup();
    const int n = atoms::num_atoms;
    if (cmd == "dump") {
        dump_system();
        std::ofstream st("stats.bin", std::ios::binary);
        dump_stats(st);
        return 0;
    }
    if (cmd == "llg") {
        const int nsteps = std::atoi(argv[3]);
        const int every = std::atoi(argv[4]);
        const bool midpoint = argc > 5 && std::string(argv[5]) == "mid";
        const bool demag = argc > 5 && std::string(argv[5]) == "demag";
        if (demag && !sim::ext_demag) {
            std::cerr << "the input does not set sim:demagnetisation-factor" << std::endl;
            return 2;
        }
        // the field calculate_applied_fields adds (src/simulate/fields.cpp:219-239) for the current
        // system magnetisation statistic, after bringing that statistic up to date
        std::ofstream demag_file;
        if (demag) demag_file.open("demag.bin", std::ios::binary);
        auto demag_update = [&]() {
            if (!demag) return;
            stats::update(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array, atoms::m_spin_array,
                          atoms::type_array, sim::temperature);
            const std::vector<double> m_l = stats::system_magnetization.get_magnetization();
            const double mu_0 = -4.0 * M_PI * 1.0e-7 /
                                (cs::system_dimensions[0] * cs::system_dimensions[1] * cs::system_dimensions[2] * 1.0e-30);
            const std::vector<double> HD = {mu_0 * sim::demag_factor[0] * m_l[0], mu_0 * sim::demag_factor[1] * m_l[1],
                                            mu_0 * sim::demag_factor[2] * m_l[2]};
            put(demag_file, HD);
        };
        std::vector<double> noise = read_doubles("noise.bin");
        // one extra 3N block at the front feeds the field dump before the first step
        if (noise.size() != (size_t)(nsteps + 1) * 3 * n) {
            std::cerr << "noise.bin must hold (nsteps+1)*3*N doubles" << std::endl;
            return 2;
        }
        g_noise = noise.data();
        g_noise_len = noise.size();
        demag_update();
        {
            std::ofstream ff("fields.bin", std::ios::binary);
            dump_fields(ff);
        }
        std::ofstream tf("traj.bin", std::ios::binary);
        for (int s = 1; s <= nsteps; ++s) {
            if (s > 1) demag_update();
            if (midpoint) sim::LLG_Midpoint();
            else sim::LLG_Heun();
            // sim::increment_time refreshes the dipole field after each step (sim.cpp:191-200)
            dipole::calculate_field(s, atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array);
            if (s % every == 0) {
                put(tf, atoms::x_spin_array);
                put(tf, atoms::y_spin_array);
                put(tf, atoms::z_spin_array);
            }
        }
        if (g_noise_pos != g_noise_len) {
            std::cerr << "noise consumed " << g_noise_pos << " of " << g_noise_len << std::endl;
            return 4;
        }
        g_noise = nullptr;
        std::ofstream st("stats.bin", std::ios::binary);
        dump_stats(st);
        return 0;
    }
    if (cmd == "mc") {
        const int nsweeps = std::atoi(argv[3]);
        sim::temperature = std::atof(argv[4]);
        const int every = std::atoi(argv[5]);
        std::ofstream mf("mc.bin", std::ios::binary);
        for (int s = 1; s <= nsweeps; ++s) {
            montecarlo::mc_step(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array, atoms::num_atoms,
                                atoms::type_array);
            if (s % every == 0) {
                stats::system_magnetization.calculate_magnetization(atoms::x_spin_array, atoms::y_spin_array,
                                                                    atoms::z_spin_array, atoms::m_spin_array);
                const std::vector<double>& m = stats::system_magnetization.get_magnetization();
                put(mf, m);
            }
        }
        return 0;
    }
    if (cmd == "cmc") {
        // constrained Monte Carlo at constraint (phi, theta): the first cmc_step runs CMCinit, which
        // sets every spin along the constraint (src/montecarlo/cmc.cpp:204-278); every `every`
        // sweeps the system magnetisation (4 values) and the total energy per atom are recorded
        const int nsweeps = std::atoi(argv[3]);
        sim::temperature = std::atof(argv[4]);
        const int every = std::atoi(argv[5]);
        sim::constraint_phi = std::atof(argv[6]);
        sim::constraint_theta = std::atof(argv[7]);
        std::ofstream cf("cmc.bin", std::ios::binary);
        for (int s = 1; s <= nsweeps; ++s) {
            montecarlo::cmc_step();
            if (s % every == 0) {
                stats::system_magnetization.calculate_magnetization(atoms::x_spin_array, atoms::y_spin_array,
                                                                    atoms::z_spin_array, atoms::m_spin_array);
                std::vector<double> row = stats::system_magnetization.get_magnetization();
                row.resize(4);
                stats::system_energy.calculate(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array,
                                               atoms::m_spin_array, atoms::type_array, sim::temperature);
                row.push_back(stats::system_energy.get_total_energy()[0]);
                put(cf, row);
            }
        }
        const std::vector<double> counts = {montecarlo::cmc::mc_success, montecarlo::cmc::energy_reject,
                                            montecarlo::cmc::sphere_reject, montecarlo::cmc::mc_total};
        put(cf, counts);
        return 0;
    }
#ifdef VMX_SHIM
    if (cmd == "gpullg") {
        // <nsteps> <every> <heun|mid> <seed>: the llg command through the gpu:: drop-in.
        // gpu::initialize takes mtrandom::integration_seed as the Philox key and sim::time as the
        // step counter, so seed = the noise seed and sim::time = 1 make the device draw the
        // stream that the llg command's noise.bin holds for steps 1..nsteps
        const int nsteps = std::atoi(argv[3]);
        const int every = std::atoi(argv[4]);
        const bool midpoint = std::string(argv[5]) == "mid";
        mtrandom::integration_seed = std::atoi(argv[6]);
        sim::time = 1;
        gpu::initialize();
        std::ofstream tf("traj.bin", std::ios::binary);
        for (int s = 1; s <= nsteps; ++s) {
            if (midpoint) gpu::llg_midpoint();
            else gpu::llg_heun();
            if (s % every == 0) {
                gpu::transfer_spin_positions_from_gpu_to_cpu();
                put(tf, atoms::x_spin_array);
                put(tf, atoms::y_spin_array);
                put(tf, atoms::z_spin_array);
            }
        }
        gpu::finalize();
        return 0;
    }
    if (cmd == "gpumc") {
        // the mc command through the drop-in: gpu::mc_step per sweep, the host spin arrays
        // refreshed before each sample
        const int nsweeps = std::atoi(argv[3]);
        sim::temperature = std::atof(argv[4]);
        const int every = std::atoi(argv[5]);
        gpu::initialize();
        std::ofstream mf("mc.bin", std::ios::binary);
        for (int s = 1; s <= nsweeps; ++s) {
            gpu::mc_step();
            if (s % every == 0) {
                gpu::transfer_spin_positions_from_gpu_to_cpu();
                stats::system_magnetization.calculate_magnetization(atoms::x_spin_array, atoms::y_spin_array,
                                                                    atoms::z_spin_array, atoms::m_spin_array);
                put(mf, stats::system_magnetization.get_magnetization());
            }
        }
        gpu::finalize();
        return 0;
    }
    if (cmd == "gpucmc") {
        // the cmc command through the drop-in boundary: gpu::initialize, then gpu::cmc_step per
        // sweep (the device runs the sweep; its first call runs CMCinit on the host spins), the
        // host spin arrays refreshed by gpu::transfer_spin_positions_from_gpu_to_cpu before each
        // sample; gpu::finalize brings the montecarlo::cmc counters up to date
        const int nsweeps = std::atoi(argv[3]);
        sim::temperature = std::atof(argv[4]);
        const int every = std::atoi(argv[5]);
        sim::constraint_phi = std::atof(argv[6]);
        sim::constraint_theta = std::atof(argv[7]);
        // optional <phi2> <theta2> <sweep>: at that sweep the constraint moves and CMCinit runs on
        // the host, as src/program/cmc_anisotropy.cpp:114-118,163-165 does between its loops
        const bool move = argc > 10;
        const int move_at = move ? std::atoi(argv[10]) : 0;
        gpu::initialize();
        std::ofstream cf("cmc.bin", std::ios::binary);
        for (int s = 1; s <= nsweeps; ++s) {
            if (move && s == move_at) {
                sim::constraint_phi = std::atof(argv[8]);
                sim::constraint_theta = std::atof(argv[9]);
                sim::constraint_phi_changed = true;
                sim::constraint_theta_changed = true;
                montecarlo::CMCinit();
            }
            gpu::cmc_step();
            if (s % every == 0) {
                gpu::transfer_spin_positions_from_gpu_to_cpu();
                stats::system_magnetization.calculate_magnetization(atoms::x_spin_array, atoms::y_spin_array,
                                                                    atoms::z_spin_array, atoms::m_spin_array);
                std::vector<double> row = stats::system_magnetization.get_magnetization();
                row.resize(4);
                stats::system_energy.calculate(atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array,
                                               atoms::m_spin_array, atoms::type_array, sim::temperature);
                row.push_back(stats::system_energy.get_total_energy()[0]);
                put(cf, row);
            }
        }
        gpu::finalize();
        const std::vector<double> counts = {montecarlo::cmc::mc_success, montecarlo::cmc::energy_reject,
                                            montecarlo::cmc::sphere_reject, montecarlo::cmc::mc_total};
        put(cf, counts);
        return 0;
    }
#endif
    if (cmd == "dipole") {
        if (!dipole::activated) {
            std::cerr << "the input does not activate the dipole field" << std::endl;
            return 2;
        }
        // the time-0 field of spins_in.bin, then the field of spins2.bin at the next update time
        std::ofstream df("dipole.bin", std::ios::binary);
        dump_dipole(df);
        std::vector<double> sp2 = read_doubles("spins2.bin");
        if (sp2.size() != 3 * (size_t)n) {
            std::cerr << "spins2.bin must hold 3*N doubles" << std::endl;
            return 2;
        }
        for (int i = 0; i < n; ++i) {
            atoms::x_spin_array[i] = sp2[i];
            atoms::y_spin_array[i] = sp2[n + i];
            atoms::z_spin_array[i] = sp2[2 * n + i];
        }
        dipole::calculate_field(dipole::update_rate, atoms::x_spin_array, atoms::y_spin_array, atoms::z_spin_array);
        std::ofstream df2("dipole2.bin", std::ios::binary);
        dump_dipole(df2);
        return 0;
    }
    std::cerr << "unknown command " << cmd << std::endl;
    return 2;
}
