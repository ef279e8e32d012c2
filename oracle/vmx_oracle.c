/*
 * vmx_oracle.c — CPU oracle for the LLG-Heun / Metropolis hot path.
 *
 * TEST INFRASTRUCTURE.  This file restates, in plain C and in the reference's own loop and
 * operation order, the VAMPIRE CPU functions the device path replaces.  It is used only by
 * tests/ (parity checks), __graft_entry__.smoke() (the reference result of its small run) and
 * bench.py (the timed CPU baseline).  The product path (vampire_amd/, libvmx.so) never calls it.
 *
 * Built by __graft_entry__.build() as oracle/liboracle.so:
 *   gcc -O3 -std=c99 -ffp-contract=off -fopenmp -fPIC -shared vmx_oracle.c -lm
 * (no FMA contraction: every product and sum rounds where the reference's C++ rounds).
 *
 * Reference functions restated (file:line in the VAMPIRE tree):
 *   or_spin_fields        sim::calculate_spin_fields        src/simulate/fields.cpp:67-118
 *                         exchange::internal::exchange_fields src/exchange/exchange_fields.cpp:29-169
 *                         anisotropy::fields               src/anisotropy/fields.cpp:31-68,
 *                           uniaxial_second_order.cpp:51-98, cubic_fourth_order.cpp:34-74,
 *                           cubic_fourth_order_rotation.cpp:34-100
 *   or_external_fields    sim::calculate_external_fields   src/simulate/fields.cpp:120-169
 *                         calculate_thermal_fields         fields.cpp:246-289
 *                         calculate_applied_fields         fields.cpp:171-244
 *   or_llg_heun           sim::LLG_Heun                    src/simulate/LLGHeun.cpp:160-282
 *   or_llg_midpoint       sim::LLG_Midpoint                src/simulate/LLGMidpoint.cpp:87-186
 *   or_magnetization      magnetization_statistic_t::calculate_magnetization
 *                                                          src/statistics/magnetization.cpp:134-180
 *   or_energy             energy_statistic_t::calculate    src/statistics/energy.cpp:141-258
 *   or_spin_energy        sim::calculate_spin_energy       src/simulate/energy.cpp:147-174
 *                         exchange::single_spin_energy     src/exchange/energy.cpp:27-140
 *                         anisotropy::single_spin_energy   src/anisotropy/energy.cpp:31-50
 *   or_mc_random_site     montecarlo::mc_step              src/montecarlo/mc.cpp:29-127,
 *                                                          mc_moves.cpp:71-87,158-162
 *   or_mc_checkerboard    the colour-ordered variant of mc_step run by the device
 *   or_cmc_random_pairs   montecarlo::cmc_step (constrained Monte Carlo)
 *                                                          src/montecarlo/cmc.cpp:282-497,
 *                                                          cmc::polar_rot_matrix cmc.cpp:43-138
 *   or_cmc_checkerboard   the colour-paired variant of cmc_step run by the device
 *   or_dipole_update      cells::mag                       src/cells/mag.cpp:35-77
 *                         dipole::internal::update_field   src/dipole/update.cpp:37-105 (tensor solver)
 *                         atom fields from cell fields     src/dipole/field.cpp:106-126
 *                         the refresh every update_rate steps, src/dipole/field.cpp:43-71 after
 *                         sim::increment_time (src/simulate/sim.cpp:191-200); the field enters
 *                         H_ext last (fields.cpp:166, 291-310), the demag field the single-spin
 *                         and statistic energies (src/simulate/energy.cpp:123-125, 171;
 *                         src/statistics/energy.cpp:198-208)
 *
 * Pins (tests/test_oracle.py; every function above has one, none is left unpinned):
 *   or_spin_fields, or_external_fields, or_llg_heun, or_llg_midpoint, or_dipole_update: bit-identical
 *     to the reference's own runs of its input decks (golden vectors written by oracle/ref's harness,
 *     built from the reference's sources), plus closed forms (damped precession, pair sums,
 *     -dE/dS of each anisotropy) and second-order convergence;
 *   or_magnetization, or_energy: the reference's statistic objects on the same decks (<= 1e-12);
 *   or_philox4x32_10: the Random123 known-answer vectors;
 *   or_mc_random_site, or_mc_checkerboard, or_cmc_*: the single-spin Boltzmann density and the
 *     reference's own mc_step / cmc_step runs (M(T), E(T), outcome fractions, within error bars);
 *     the constraint frame against its closed form.
 *
 * Noise: the reference draws MT19937 + Ziggurat normals (src/random/random.cpp:219-244), x for all
 * atoms, then y, then z (fields.cpp:274-276).  The oracle takes either an explicit noise array in
 * that order or generates the device's documented stream: Philox-4x32-10 (Salmon et al., SC'11)
 * keyed by the seed, counter (atom, step_lo, step_hi, stream), Box-Muller on word pairs.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    double alpha, gamma_rel, mu_s_SI;
    double one_oneplusalpha_sq, alpha_oneplusalpha_sq, H_th_sigma;
    double temperature_rescaling_alpha, temperature_rescaling_Tc;
    double ku2, ku_vector[3];
    double kc4, kc_vector1[3], kc_vector2[3], kc_vector3[3];
    double applied_field_strength, applied_field_unit_vector[3];
} or_material;

/* Macrocell dipole field, tensor solver: inputs as the reference builds them at set-up, and the
 * field state, refreshed by or_dipole_update. */
typedef struct {
    int ncells, nlocal;
    const int* local_cell;       /* [nlocal] cells::cell_id_array: the cell of each tensor row */
    const int* atom_cell;        /* [n] cells::atom_cell_id_array */
    const int* magnetic;         /* [nmat] 1 unless mp::material[m].non_magnetic */
    const double* volume;        /* [ncells] cells::volume_array */
    const double* tensor;        /* [6][nlocal][ncells] rij_tensor_{xx,xy,xz,yy,yz,zz}[lc][j] */
    int update_rate;             /* dipole::update_rate */
    long long time;              /* Heun steps / sweeps since the field was set up */
    double* atom_B;              /* [3][n] dipole::atom_dipolar_field_array_{x,y,z} */
    double* atom_Hd;             /* [3][n] dipole::atom_mu0demag_field_array_{x,y,z} */
    double* cell_mag;            /* [3][ncells] cells::mag_array_{x,y,z} */
    double* cell_B;              /* [3][ncells] dipole::cells_field_array_{x,y,z} */
    double* cell_Hd;             /* [3][ncells] dipole::cells_mu0Hd_field_array_{x,y,z} */
} or_dipole;

typedef struct {
    int n, nmat;
    const int* type;
    const double* moment;
    const int* nbr_start;        /* inclusive */
    const int* nbr_end;          /* inclusive */
    const int* nbr;
    const int* iid;
    int exchange_type;           /* 0 isotropic, 1 vectorial, 2 tensorial */
    const double* exchange;      /* [num_interactions][1|3|9] */
    const or_material* mat;
    unsigned ani;                /* 1 uniaxial2, 2 cubic4, 4 cubic4 rotated */
    or_dipole* dip;              /* NULL: no dipole field */
} or_sys;

typedef struct {
    double temperature;
    double H[3];                 /* sim::H_vec * sim::H_applied */
    unsigned fld;                /* 1 thermal, 2 applied, 4 local applied */
    double dt, half_dt;
    double Hd[3];                /* thin-film demag field of calculate_applied_fields (sim::ext_demag) */
} or_cond;

/* ---- Philox-4x32-10 and Box-Muller ------------------------------------------------------ */
static uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = mulhi32(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        const uint32_t hi1 = mulhi32(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox_at(uint64_t seed, uint32_t local, uint64_t step, uint32_t stream, uint32_t out[4]) {
    const uint32_t ctr[4] = {local, (uint32_t)step, (uint32_t)(step >> 32), stream};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    or_philox4x32_10(ctr, key, out);
}

/* three normals from one block: (x0,x1) and (x2,x3) Box-Muller, uniforms (k+0.5)/2^32 */
void or_normals3(uint64_t seed, int local, uint64_t step, uint32_t stream, double g[3]) {
    uint32_t u[4];
    philox_at(seed, (uint32_t)local, step, stream, u);
    const double s = 2.3283064365386963e-10, twopi = 6.283185307179586;
    const double u1 = ((double)u[0] + 0.5) * s, u2 = ((double)u[1] + 0.5) * s;
    const double u3 = ((double)u[2] + 0.5) * s, u4 = ((double)u[3] + 0.5) * s;
    const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    g[0] = r1 * cos(twopi * u2);
    g[1] = r1 * sin(twopi * u2);
    g[2] = r2 * cos(twopi * u4);
}

/* [3][n] thermal normals of one step in the reference's consumption layout */
void or_thermal_noise(uint64_t seed, int n, uint64_t step, double* out) {
    for (int i = 0; i < n; ++i) {
        double g[3];
        or_normals3(seed, i, step, 0u, g);
        out[i] = g[0];
        out[(size_t)n + i] = g[1];
        out[2 * (size_t)n + i] = g[2];
    }
}

/* ---- fields ------------------------------------------------------------------------------- */
static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

void or_spin_fields(const or_sys* s, const double* sx, const double* sy, const double* sz,
                    double* hx_out, double* hy_out, double* hz_out) {
    const int n = s->n;
    /* fields.cpp:79-81 — zero */
    for (int atom = 0; atom < n; ++atom) { hx_out[atom] = 0.0; hy_out[atom] = 0.0; hz_out[atom] = 0.0; }
    /* exchange_fields.cpp:52-162 */
#pragma omp parallel for schedule(static)
    for (int atom = 0; atom < n; ++atom) {
        double hx = 0.0, hy = 0.0, hz = 0.0;
        const int start = s->nbr_start[atom];
        const int end = s->nbr_end[atom] + 1;
        for (int nn = start; nn < end; ++nn) {
            const int natom = s->nbr[nn];
            const int iid = s->iid[nn];
            if (s->exchange_type == 0) {
                const double Jij = s->exchange[iid];
                hx += Jij * sx[natom];
                hy += Jij * sy[natom];
                hz += Jij * sz[natom];
            } else if (s->exchange_type == 1) {
                const double* J = s->exchange + 3 * (size_t)iid;
                hx += J[0] * sx[natom];
                hy += J[1] * sy[natom];
                hz += J[2] * sz[natom];
            } else {
                const double* J = s->exchange + 9 * (size_t)iid;
                const double S[3] = {sx[natom], sy[natom], sz[natom]};
                hx += (J[0] * S[0] + J[1] * S[1] + J[2] * S[2]);
                hy += (J[3] * S[0] + J[4] * S[1] + J[5] * S[2]);
                hz += (J[6] * S[0] + J[7] * S[1] + J[8] * S[2]);
            }
        }
        hx_out[atom] += hx;
        hy_out[atom] += hy;
        hz_out[atom] += hz;
    }
    /* anisotropy terms, one pass each, in the order of anisotropy/fields.cpp:44-59 */
    if (s->ani & 1u) {
        const double scale = 2.0;
        for (int atom = 0; atom < n; ++atom) {
            const or_material* m = &s->mat[s->type[atom]];
            const double ex = m->ku_vector[0], ey = m->ku_vector[1], ez = m->ku_vector[2];
            const double sdote = (sx[atom] * ex + sy[atom] * ey + sz[atom] * ez);
            const double k2 = scale * m->ku2 * sdote;
            hx_out[atom] += ex * k2;
            hy_out[atom] += ey * k2;
            hz_out[atom] += ez * k2;
        }
    }
    if (s->ani & 2u) {
        const double scale = 0.5 * 4.0;
        for (int atom = 0; atom < n; ++atom) {
            const or_material* m = &s->mat[s->type[atom]];
            const double k4 = scale * m->kc4;
            const double x = sx[atom], y = sy[atom], z = sz[atom];
            hx_out[atom] += x * x * x * k4;
            hy_out[atom] += y * y * y * k4;
            hz_out[atom] += z * z * z * k4;
        }
    }
    if (s->ani & 4u) {
        const double scale = 0.5 * 4.0;
        for (int atom = 0; atom < n; ++atom) {
            const or_material* m = &s->mat[s->type[atom]];
            const double* e1 = m->kc_vector1;
            const double* e2 = m->kc_vector2;
            const double* e3 = m->kc_vector3;
            const double k4 = scale * m->kc4;
            const double x = sx[atom], y = sy[atom], z = sz[atom];
            const double d1 = x * e1[0] + y * e1[1] + z * e1[2];
            const double d2 = x * e2[0] + y * e2[1] + z * e2[2];
            const double d3 = x * e3[0] + y * e3[1] + z * e3[2];
            const double d1c = d1 * d1 * d1, d2c = d2 * d2 * d2, d3c = d3 * d3 * d3;
            hx_out[atom] += k4 * (d1c * e1[0] + d2c * e2[0] + d3c * e3[0]);
            hy_out[atom] += k4 * (d1c * e1[1] + d2c * e2[1] + d3c * e3[1]);
            hz_out[atom] += k4 * (d1c * e1[2] + d2c * e2[2] + d3c * e3[2]);
        }
    }
}

/* per-material thermal width sigma_prefactor (fields.cpp:258-272) and applied field (:181-210) */
static void material_conditions(const or_sys* s, const or_cond* c, double* sigma, double* happ) {
    for (int m = 0; m < s->nmat; ++m) {
        const or_material* mt = &s->mat[m];
        const double temperature = c->temperature;
        const double alpha = mt->temperature_rescaling_alpha;
        const double Tc = mt->temperature_rescaling_Tc;
        const double rescaled_temperature = temperature < Tc ? Tc * pow(temperature / Tc, alpha) : temperature;
        const double sqrt_T = sqrt(rescaled_temperature);
        sigma[m] = sqrt_T * mt->H_th_sigma;
        for (int k = 0; k < 3; ++k) {
            if (c->fld & 4u)
                happ[3 * m + k] = c->H[k] + mt->applied_field_strength * mt->applied_field_unit_vector[k];
            else
                happ[3 * m + k] = c->H[k];
        }
    }
}

void or_external_fields(const or_sys* s, const or_cond* c, const double* noise, uint64_t seed, uint64_t step,
                        double* ex, double* ey, double* ez) {
    const int n = s->n;
    double* sigma = (double*)malloc(sizeof(double) * (size_t)s->nmat);
    double* happ = (double*)malloc(sizeof(double) * 3 * (size_t)s->nmat);
    material_conditions(s, c, sigma, happ);
    for (int atom = 0; atom < n; ++atom) { ex[atom] = 0.0; ey[atom] = 0.0; ez[atom] = 0.0; }
    if (c->fld & 1u) {
        if (noise) {
            for (int atom = 0; atom < n; ++atom) {
                ex[atom] = noise[atom];
                ey[atom] = noise[(size_t)n + atom];
                ez[atom] = noise[2 * (size_t)n + atom];
            }
        } else {
#pragma omp parallel for schedule(static)
            for (int atom = 0; atom < n; ++atom) {
                double g[3];
                or_normals3(seed, atom, step, 0u, g);
                ex[atom] = g[0]; ey[atom] = g[1]; ez[atom] = g[2];
            }
        }
        for (int atom = 0; atom < n; ++atom) {
            const double H_th_sigma = sigma[s->type[atom]];
            ex[atom] *= H_th_sigma;
            ey[atom] *= H_th_sigma;
            ez[atom] *= H_th_sigma;
        }
    }
    if (c->fld & 2u) {
        for (int atom = 0; atom < n; ++atom) {
            const int m = s->type[atom];
            ex[atom] += happ[3 * m + 0];
            ey[atom] += happ[3 * m + 1];
            ez[atom] += happ[3 * m + 2];
        }
        /* thin-film demag field, after the global and local fields (fields.cpp:219-239) */
        if (c->Hd[0] != 0.0 || c->Hd[1] != 0.0 || c->Hd[2] != 0.0) {
            for (int atom = 0; atom < n; ++atom) {
                ex[atom] += c->Hd[0];
                ey[atom] += c->Hd[1];
                ez[atom] += c->Hd[2];
            }
        }
    }
    if (s->dip) {                                               /* fields.cpp:291-310 */
        const double* B = s->dip->atom_B;
        for (int atom = 0; atom < n; ++atom) {
            ex[atom] += B[atom];
            ey[atom] += B[(size_t)n + atom];
            ez[atom] += B[2 * (size_t)n + atom];
        }
    }
    free(sigma);
    free(happ);
}

/* ---- macrocell dipole field (tensor solver) -------------------------------------------------- */
void or_dipole_update(const or_sys* s, const double* sx, const double* sy, const double* sz) {
    or_dipole* d = s->dip;
    const int n = s->n, nc = d->ncells;
    double* mx = d->cell_mag;
    double* my = mx + nc;
    double* mz = my + nc;
    int* count = (int*)calloc((size_t)nc, sizeof(int));
    for (int i = 0; i < nc; ++i) { mx[i] = 0.0; my[i] = 0.0; mz[i] = 0.0; }
    for (int i = 0; i < n; ++i) {                               /* cells/mag.cpp:53-66 */
        const int cell = d->atom_cell[i];
        const int type = s->type[i];
        const double mus = s->mat[type].mu_s_SI;
        if (d->magnetic[type]) {
            mx[cell] += sx[i] * mus;
            my[cell] += sy[i] * mus;
            mz[cell] += sz[i] * mus;
            count[cell]++;                                      /* cells/initialize.cpp:209-224 */
        }
    }
    const double imuB = 1.0 / 9.27400915e-24;
    const size_t ts = (size_t)d->nlocal * nc;
    double* bx = d->cell_B;
    double* by = bx + nc;
    double* bz = by + nc;
    double* hx = d->cell_Hd;
    double* hy = hx + nc;
    double* hz = hy + nc;
    for (int lc = 0; lc < d->nlocal; lc++) {                    /* dipole/update.cpp:52-93 */
        const int i = d->local_cell[lc];
        if (count[i] > 0) {
            const double eightPI_three_cell_volume = 8.0 * M_PI / (3.0 * d->volume[i]);
            const double self_demag = eightPI_three_cell_volume;
            const double mx_i = mx[i] * imuB;
            const double my_i = my[i] * imuB;
            const double mz_i = mz[i] * imuB;
            bx[i] = self_demag * mx_i * 0.0;
            by[i] = self_demag * my_i * 0.0;
            bz[i] = self_demag * mz_i * 0.0;
            hx[i] = -0.5 * self_demag * mx_i;
            hy[i] = -0.5 * self_demag * my_i;
            hz[i] = -0.5 * self_demag * mz_i;
            const double* txx = d->tensor + (size_t)lc * nc;
            const double* txy = txx + ts;
            const double* txz = txy + ts;
            const double* tyy = txz + ts;
            const double* tyz = tyy + ts;
            const double* tzz = tyz + ts;
            for (int j = 0; j < nc; j++) {
                if (count[j] > 0) {
                    const double mxj = mx[j] * imuB;
                    const double myj = my[j] * imuB;
                    const double mzj = mz[j] * imuB;
                    bx[i] += (mxj * txx[j] + myj * txy[j] + mzj * txz[j]);
                    by[i] += (mxj * txy[j] + myj * tyy[j] + mzj * tyz[j]);
                    bz[i] += (mxj * txz[j] + myj * tyz[j] + mzj * tzz[j]);
                    hx[i] += (mxj * txx[j] + myj * txy[j] + mzj * txz[j]);
                    hy[i] += (mxj * txy[j] + myj * tyy[j] + mzj * tyz[j]);
                    hz[i] += (mxj * txz[j] + myj * tyz[j] + mzj * tzz[j]);
                }
            }
            bx[i] = bx[i] * 9.27400915e-01;
            by[i] = by[i] * 9.27400915e-01;
            bz[i] = bz[i] * 9.27400915e-01;
            hx[i] = hx[i] * 9.27400915e-01;
            hy[i] = hy[i] * 9.27400915e-01;
            hz[i] = hz[i] * 9.27400915e-01;
        }
    }
    for (int atom = 0; atom < n; atom++) {                      /* dipole/field.cpp:113-125 */
        const int cell = d->atom_cell[atom];
        if (count[cell] > 0 && d->magnetic[s->type[atom]]) {
            d->atom_B[atom] = bx[cell];
            d->atom_B[(size_t)n + atom] = by[cell];
            d->atom_B[2 * (size_t)n + atom] = bz[cell];
            d->atom_Hd[atom] = hx[cell];
            d->atom_Hd[(size_t)n + atom] = hy[cell];
            d->atom_Hd[2 * (size_t)n + atom] = hz[cell];
        }
    }
    free(count);
}

/* After a Heun step or a Monte Carlo sweep: the refresh of dipole::calculate_field when the time
 * reaches a multiple of the update rate (src/dipole/field.cpp:52-58). */
static void dipole_tick(const or_sys* s, const double* sx, const double* sy, const double* sz) {
    if (!s->dip) return;
    s->dip->time++;
    if (s->dip->time % s->dip->update_rate == 0) or_dipole_update(s, sx, sy, sz);
}

/* ---- integrator: sim::LLG_Heun ------------------------------------------------------------ */
typedef struct {
    double *x0, *y0, *z0, *xs, *ys, *zs, *xe, *ye, *ze, *xh, *yh, *zh;
    double *hsx, *hsy, *hsz, *hex, *hey, *hez;
} heun_arrays;

static int alloc_arrays(heun_arrays* a, int n) {
    double** p[] = {&a->x0, &a->y0, &a->z0, &a->xs, &a->ys, &a->zs, &a->xe, &a->ye, &a->ze, &a->xh, &a->yh,
                    &a->zh, &a->hsx, &a->hsy, &a->hsz, &a->hex, &a->hey, &a->hez};
    for (size_t k = 0; k < sizeof p / sizeof p[0]; ++k) {
        *p[k] = (double*)calloc((size_t)n, sizeof(double));
        if (!*p[k]) return -1;
    }
    return 0;
}

static void free_arrays(heun_arrays* a) {
    double* p[] = {a->x0, a->y0, a->z0, a->xs, a->ys, a->zs, a->xe, a->ye, a->ze, a->xh, a->yh,
                   a->zh, a->hsx, a->hsy, a->hsz, a->hex, a->hey, a->hez};
    for (size_t k = 0; k < sizeof p / sizeof p[0]; ++k) free(p[k]);
}

static void heun_step(const or_sys* s, const or_cond* c, double* sx, double* sy, double* sz, const double* noise,
                      uint64_t seed, uint64_t step, heun_arrays* a) {
    const int num_atoms = s->n;
    for (int atom = 0; atom < num_atoms; atom++) {             /* LLGHeun.cpp:177-181 */
        a->x0[atom] = sx[atom]; a->y0[atom] = sy[atom]; a->z0[atom] = sz[atom];
    }
    or_spin_fields(s, sx, sy, sz, a->hsx, a->hsy, a->hsz);      /* :184 */
    or_external_fields(s, c, noise, seed, step, a->hex, a->hey, a->hez);   /* :185 */
#pragma omp parallel for schedule(static)
    for (int atom = 0; atom < num_atoms; atom++) {             /* :188-226 */
        const or_material* m = &s->mat[s->type[atom]];
        const double one_oneplusalpha_sq = m->one_oneplusalpha_sq;
        const double alpha_oneplusalpha_sq = m->alpha_oneplusalpha_sq;
        const double S[3] = {sx[atom], sy[atom], sz[atom]};
        const double H[3] = {a->hsx[atom] + a->hex[atom], a->hsy[atom] + a->hey[atom], a->hsz[atom] + a->hez[atom]};
        double xyz[3], S_new[3];
        xyz[0] = (one_oneplusalpha_sq) * (S[1] * H[2] - S[2] * H[1]) +
                 (alpha_oneplusalpha_sq) * (S[1] * (S[0] * H[1] - S[1] * H[0]) - S[2] * (S[2] * H[0] - S[0] * H[2]));
        xyz[1] = (one_oneplusalpha_sq) * (S[2] * H[0] - S[0] * H[2]) +
                 (alpha_oneplusalpha_sq) * (S[2] * (S[1] * H[2] - S[2] * H[1]) - S[0] * (S[0] * H[1] - S[1] * H[0]));
        xyz[2] = (one_oneplusalpha_sq) * (S[0] * H[1] - S[1] * H[0]) +
                 (alpha_oneplusalpha_sq) * (S[0] * (S[2] * H[0] - S[0] * H[2]) - S[1] * (S[1] * H[2] - S[2] * H[1]));
        a->xe[atom] = xyz[0]; a->ye[atom] = xyz[1]; a->ze[atom] = xyz[2];
        S_new[0] = S[0] + xyz[0] * c->dt;
        S_new[1] = S[1] + xyz[1] * c->dt;
        S_new[2] = S[2] + xyz[2] * c->dt;
        const double mod_S = 1.0 / sqrt(S_new[0] * S_new[0] + S_new[1] * S_new[1] + S_new[2] * S_new[2]);
        a->xs[atom] = S_new[0] * mod_S;
        a->ys[atom] = S_new[1] * mod_S;
        a->zs[atom] = S_new[2] * mod_S;
    }
    for (int atom = 0; atom < num_atoms; atom++) {             /* :229-233 */
        sx[atom] = a->xs[atom]; sy[atom] = a->ys[atom]; sz[atom] = a->zs[atom];
    }
    or_spin_fields(s, sx, sy, sz, a->hsx, a->hsy, a->hsz);      /* :236 */
#pragma omp parallel for schedule(static)
    for (int atom = 0; atom < num_atoms; atom++) {             /* :239-260 */
        const or_material* m = &s->mat[s->type[atom]];
        const double one_oneplusalpha_sq = m->one_oneplusalpha_sq;
        const double alpha_oneplusalpha_sq = m->alpha_oneplusalpha_sq;
        const double S[3] = {sx[atom], sy[atom], sz[atom]};
        const double H[3] = {a->hsx[atom] + a->hex[atom], a->hsy[atom] + a->hey[atom], a->hsz[atom] + a->hez[atom]};
        a->xh[atom] = (one_oneplusalpha_sq) * (S[1] * H[2] - S[2] * H[1]) +
                      (alpha_oneplusalpha_sq) * (S[1] * (S[0] * H[1] - S[1] * H[0]) - S[2] * (S[2] * H[0] - S[0] * H[2]));
        a->yh[atom] = (one_oneplusalpha_sq) * (S[2] * H[0] - S[0] * H[2]) +
                      (alpha_oneplusalpha_sq) * (S[2] * (S[1] * H[2] - S[2] * H[1]) - S[0] * (S[0] * H[1] - S[1] * H[0]));
        a->zh[atom] = (one_oneplusalpha_sq) * (S[0] * H[1] - S[1] * H[0]) +
                      (alpha_oneplusalpha_sq) * (S[0] * (S[2] * H[0] - S[0] * H[2]) - S[1] * (S[1] * H[2] - S[2] * H[1]));
    }
#pragma omp parallel for schedule(static)
    for (int atom = 0; atom < num_atoms; atom++) {             /* :263-279 */
        double S_new[3];
        S_new[0] = a->x0[atom] + c->half_dt * (a->xe[atom] + a->xh[atom]);
        S_new[1] = a->y0[atom] + c->half_dt * (a->ye[atom] + a->yh[atom]);
        S_new[2] = a->z0[atom] + c->half_dt * (a->ze[atom] + a->zh[atom]);
        const double mod_S = 1.0 / sqrt(S_new[0] * S_new[0] + S_new[1] * S_new[1] + S_new[2] * S_new[2]);
        sx[atom] = S_new[0] * mod_S;
        sy[atom] = S_new[1] * mod_S;
        sz[atom] = S_new[2] * mod_S;
    }
}

/* nsteps Heun steps; noise [nsteps][3][n] or NULL for the Philox stream at steps step0.. */
int or_llg_heun(const or_sys* s, const or_cond* c, double* sx, double* sy, double* sz, int nsteps,
                const double* noise, uint64_t seed, uint64_t step0, int nthreads) {
    heun_arrays a;
    memset(&a, 0, sizeof a);
    if (alloc_arrays(&a, s->n)) { free_arrays(&a); return -1; }
    set_threads(nthreads);
    for (int k = 0; k < nsteps; ++k) {
        heun_step(s, c, sx, sy, sz, noise ? noise + (size_t)k * 3 * s->n : NULL, seed, step0 + (uint64_t)k, &a);
        dipole_tick(s, sx, sy, sz);
    }
    free_arrays(&a);
    return 0;
}

/* ---- integrator: sim::LLG_Midpoint ------------------------------------------------------- */
static void midpoint_step(const or_sys* s, const or_cond* c, double* sx, double* sy, double* sz,
                          const double* noise, uint64_t seed, uint64_t step, heun_arrays* a) {
    const int num_atoms = s->n;
    for (int atom = 0; atom < num_atoms; atom++) {             /* LLGMidpoint.cpp:97-101 */
        a->x0[atom] = sx[atom]; a->y0[atom] = sy[atom]; a->z0[atom] = sz[atom];
    }
    or_spin_fields(s, sx, sy, sz, a->hsx, a->hsy, a->hsz);      /* :104 */
    or_external_fields(s, c, noise, seed, step, a->hex, a->hey, a->hez);   /* :105 */
#pragma omp parallel for schedule(static)
    for (int atom = 0; atom < num_atoms; atom++) {             /* :108-141 */
        const or_material* m = &s->mat[s->type[atom]];
        const double alpha = m->alpha;
        const double beta = -1.0 * c->dt * m->one_oneplusalpha_sq * 0.5;
        const double beta2 = beta * beta;
        const double S[3] = {sx[atom], sy[atom], sz[atom]};
        const double H[3] = {a->hsx[atom] + a->hex[atom], a->hsy[atom] + a->hey[atom], a->hsz[atom] + a->hez[atom]};
        const double F[3] = {H[0] + alpha * (S[1] * H[2] - S[2] * H[1]),
                             H[1] + alpha * (S[2] * H[0] - S[0] * H[2]),
                             H[2] + alpha * (S[0] * H[1] - S[1] * H[0])};
        const double FdotF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
        const double beta2FdotS = beta2 * (F[0] * S[0] + F[1] * S[1] + F[2] * S[2]);
        const double one_o_one_plus_beta2FdotF = 1.0 / (1.0 + beta2 * FdotF);
        const double one_minus_beta2FdotF = 1.0 - beta2 * FdotF;
        a->xs[atom] = (S[0] + one_o_one_plus_beta2FdotF * (S[0] * one_minus_beta2FdotF +
                                                          2.0 * (beta * (F[1] * S[2] - F[2] * S[1]) + F[0] * beta2FdotS))) * 0.5;
        a->ys[atom] = (S[1] + one_o_one_plus_beta2FdotF * (S[1] * one_minus_beta2FdotF +
                                                          2.0 * (beta * (F[2] * S[0] - F[0] * S[2]) + F[1] * beta2FdotS))) * 0.5;
        a->zs[atom] = (S[2] + one_o_one_plus_beta2FdotF * (S[2] * one_minus_beta2FdotF +
                                                          2.0 * (beta * (F[0] * S[1] - F[1] * S[0]) + F[2] * beta2FdotS))) * 0.5;
    }
    for (int atom = 0; atom < num_atoms; atom++) {             /* :144-148 */
        sx[atom] = a->xs[atom]; sy[atom] = a->ys[atom]; sz[atom] = a->zs[atom];
    }
    or_spin_fields(s, sx, sy, sz, a->hsx, a->hsy, a->hsz);      /* :151 */
#pragma omp parallel for schedule(static)
    for (int atom = 0; atom < num_atoms; atom++) {             /* :154-183 */
        const or_material* m = &s->mat[s->type[atom]];
        const double alpha = m->alpha;
        const double beta = -1.0 * c->dt * m->one_oneplusalpha_sq * 0.5;
        const double beta2 = beta * beta;
        const double M[3] = {sx[atom], sy[atom], sz[atom]};
        const double S[3] = {a->x0[atom], a->y0[atom], a->z0[atom]};
        const double H[3] = {a->hsx[atom] + a->hex[atom], a->hsy[atom] + a->hey[atom], a->hsz[atom] + a->hez[atom]};
        const double F[3] = {H[0] + alpha * (M[1] * H[2] - M[2] * H[1]),
                             H[1] + alpha * (M[2] * H[0] - M[0] * H[2]),
                             H[2] + alpha * (M[0] * H[1] - M[1] * H[0])};
        const double FdotF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
        const double beta2FdotS = beta2 * (F[0] * S[0] + F[1] * S[1] + F[2] * S[2]);
        const double one_o_one_plus_beta2FdotF = 1.0 / (1.0 + beta2 * FdotF);
        const double one_minus_beta2FdotF = 1.0 - beta2 * FdotF;
        sx[atom] = one_o_one_plus_beta2FdotF * (S[0] * one_minus_beta2FdotF + 2.0 * (beta * (F[1] * S[2] - F[2] * S[1]) + F[0] * beta2FdotS));
        sy[atom] = one_o_one_plus_beta2FdotF * (S[1] * one_minus_beta2FdotF + 2.0 * (beta * (F[2] * S[0] - F[0] * S[2]) + F[1] * beta2FdotS));
        sz[atom] = one_o_one_plus_beta2FdotF * (S[2] * one_minus_beta2FdotF + 2.0 * (beta * (F[0] * S[1] - F[1] * S[0]) + F[2] * beta2FdotS));
    }
}

/* nsteps midpoint steps; noise [nsteps][3][n] or NULL for the Philox stream at steps step0.. */
int or_llg_midpoint(const or_sys* s, const or_cond* c, double* sx, double* sy, double* sz, int nsteps,
                    const double* noise, uint64_t seed, uint64_t step0, int nthreads) {
    heun_arrays a;
    memset(&a, 0, sizeof a);
    if (alloc_arrays(&a, s->n)) { free_arrays(&a); return -1; }
    set_threads(nthreads);
    for (int k = 0; k < nsteps; ++k) {
        midpoint_step(s, c, sx, sy, sz, noise ? noise + (size_t)k * 3 * s->n : NULL, seed, step0 + (uint64_t)k, &a);
        dipole_tick(s, sx, sy, sz);
    }
    free_arrays(&a);
    return 0;
}

/* ---- statistics --------------------------------------------------------------------------- */
void or_magnetization(int n, const double* sx, const double* sy, const double* sz, const double* mm,
                      const int* mask, int in_mask_size, double* magnetization) {
    const int mask_size = in_mask_size - 1;
    int* count = (int*)calloc((size_t)in_mask_size, sizeof(int));
    for (int k = 0; k < 4 * in_mask_size; ++k) magnetization[k] = 0.0;
    for (int atom = 0; atom < n; ++atom) {
        const int mask_id = mask[atom];
        magnetization[4 * mask_id + 0] += sx[atom] * mm[atom];
        magnetization[4 * mask_id + 1] += sy[atom] * mm[atom];
        magnetization[4 * mask_id + 2] += sz[atom] * mm[atom];
        magnetization[4 * mask_id + 3] += mm[atom];
        count[mask_id]++;
    }
    for (int mask_id = 0; mask_id < mask_size; ++mask_id) {
        const double msat = magnetization[4 * mask_id + 3];
        const double magm = sqrt(magnetization[4 * mask_id + 0] * magnetization[4 * mask_id + 0] +
                                 magnetization[4 * mask_id + 1] * magnetization[4 * mask_id + 1] +
                                 magnetization[4 * mask_id + 2] * magnetization[4 * mask_id + 2]);
        magnetization[4 * mask_id + 0] = magnetization[4 * mask_id + 0] / magm;
        magnetization[4 * mask_id + 1] = magnetization[4 * mask_id + 1] / magm;
        magnetization[4 * mask_id + 2] = magnetization[4 * mask_id + 2] / magm;
        magnetization[4 * mask_id + 3] = magm / msat;
    }
    for (int mask_id = 0; mask_id < mask_size; ++mask_id)
        if (count[mask_id] == 0)
            for (int k = 0; k < 4; ++k) magnetization[4 * mask_id + k] = 0.0;
    free(count);
}

static double exchange_energy(const or_sys* s, const double* sx_a, const double* sy_a, const double* sz_a,
                              int atom, double sx, double sy, double sz) {
    double energy = 0.0;
    for (int nn = s->nbr_start[atom]; nn <= s->nbr_end[atom]; ++nn) {
        const int natom = s->nbr[nn];
        const int iid = s->iid[nn];
        if (s->exchange_type == 0) {
            const double Jij = s->exchange[iid];
            energy -= Jij * (sx_a[natom] * sx + sy_a[natom] * sy + sz_a[natom] * sz);
        } else if (s->exchange_type == 1) {
            const double* J = s->exchange + 3 * (size_t)iid;
            energy -= (J[0] * sx_a[natom] * sx + J[1] * sy_a[natom] * sy + J[2] * sz_a[natom] * sz);
        } else {
            const double* J = s->exchange + 9 * (size_t)iid;
            const double S[3] = {sx_a[natom], sy_a[natom], sz_a[natom]};
            energy -= (J[0] * S[0] * sx + J[1] * S[1] * sx + J[2] * S[2] * sx +
                       J[3] * S[0] * sy + J[4] * S[1] * sy + J[5] * S[2] * sy +
                       J[6] * S[0] * sz + J[7] * S[1] * sz + J[8] * S[2] * sz);
        }
    }
    return energy;
}

static double anisotropy_energy(const or_sys* s, int mat, double sx, double sy, double sz) {
    double energy = 0.0;
    const or_material* m = &s->mat[mat];
    if (s->ani & 1u) {
        const double sdote = (sx * m->ku_vector[0] + sy * m->ku_vector[1] + sz * m->ku_vector[2]);
        energy += -m->ku2 * (sdote * sdote);
    }
    if (s->ani & 2u) {
        const double sx4 = sx * sx * sx * sx, sy4 = sy * sy * sy * sy, sz4 = sz * sz * sz * sz;
        energy += -0.5 * m->kc4 * (sx4 + sy4 + sz4);
    }
    if (s->ani & 4u) {
        const double* e1 = m->kc_vector1;
        const double* e2 = m->kc_vector2;
        const double* e3 = m->kc_vector3;
        const double d1 = sx * e1[0] + sy * e1[1] + sz * e1[2];
        const double d2 = sx * e2[0] + sy * e2[1] + sz * e2[2];
        const double d3 = sx * e3[0] + sy * e3[1] + sz * e3[2];
        const double a = d1 * d1 * d1 * d1, b = d2 * d2 * d2 * d2, cc = d3 * d3 * d3 * d3;
        energy += -0.5 * m->kc4 * (a + b + cc);
    }
    return energy;
}

static double applied_energy(const or_cond* c, double sx, double sy, double sz) {
    /* -H_applied (H_vec . S) with H = H_vec*H_applied supplied directly */
    return -(c->H[0] * sx + c->H[1] * sy + c->H[2] * sz);
}

double or_spin_energy(const or_sys* s, const or_cond* c, const double* sx_a, const double* sy_a,
                      const double* sz_a, int atom, double sx, double sy, double sz) {
    double energy = 0.0;
    energy += exchange_energy(s, sx_a, sy_a, sz_a, atom, sx, sy, sz);
    energy += anisotropy_energy(s, s->type[atom], sx, sy, sz);
    energy += applied_energy(c, sx, sy, sz);
    if (s->dip) {                                               /* simulate/energy.cpp:123-125, 171 */
        const double* H = s->dip->atom_Hd;
        const size_t n = (size_t)s->n;
        energy += -1.0 * (H[atom] * sx + H[n + atom] * sy + H[2 * n + atom] * sz);
    }
    return energy;
}

/* energies [5][in_mask_size]: total, exchange, anisotropy, applied, magnetostatic */
void or_energy(const or_sys* s, const or_cond* c, const double* sx, const double* sy, const double* sz,
               const int* mask, int in_mask_size, double* out) {
    const int n = s->n, mask_size = in_mask_size - 1;
    double* tot = out;
    double* exc = out + in_mask_size;
    double* ani = out + 2 * in_mask_size;
    double* app = out + 3 * in_mask_size;
    double* mag = out + 4 * in_mask_size;
    int* count = (int*)calloc((size_t)in_mask_size, sizeof(int));
    for (int k = 0; k < 5 * in_mask_size; ++k) out[k] = 0.0;
    for (int atom = 0; atom < n; ++atom) {
        exc[mask[atom]] += exchange_energy(s, sx, sy, sz, atom, sx[atom], sy[atom], sz[atom]) * s->moment[atom];
        count[mask[atom]]++;
    }
    for (int id = 0; id < mask_size; ++id) exc[id] = 0.5 * exc[id];
    for (int atom = 0; atom < n; ++atom)
        ani[mask[atom]] += anisotropy_energy(s, s->type[atom], sx[atom], sy[atom], sz[atom]) * s->moment[atom];
    for (int atom = 0; atom < n; ++atom)
        app[mask[atom]] += applied_energy(c, sx[atom], sy[atom], sz[atom]) * s->moment[atom];
    if (s->dip) {                                               /* statistics/energy.cpp:198-203 */
        const double* H = s->dip->atom_Hd;
        for (int atom = 0; atom < n; ++atom)
            mag[mask[atom]] += -1.0 * (H[atom] * sx[atom] + H[(size_t)n + atom] * sy[atom] +
                                       H[2 * (size_t)n + atom] * sz[atom]) * s->moment[atom];
    }
    for (int id = 0; id < mask_size; ++id) mag[id] = 0.5 * mag[id];
    for (int id = 0; id < mask_size; ++id) tot[id] = exc[id] + ani[id] + app[id] + mag[id];
    for (int id = 0; id < mask_size; ++id)
        if (count[id] == 0) { tot[id] = 0.0; exc[id] = 0.0; ani[id] = 0.0; app[id] = 0.0; mag[id] = 0.0; }
    free(count);
}

/* ---- Metropolis ------------------------------------------------------------------------------- */
static double kBTBohr_of(const or_material* m, double T) {
    const double alpha = m->temperature_rescaling_alpha;
    const double Tc = m->temperature_rescaling_Tc;
    const double rescaled_temperature = T < Tc ? Tc * pow(T / Tc, alpha) : T;
    return 9.27400915e-24 / (rescaled_temperature * 1.3806503e-23);
}

static int metropolis_move(const or_sys* s, const or_cond* c, double* sx, double* sy, double* sz, int atom,
                           const double g[3], double angle, double u) {
    const int imaterial = s->type[atom];
    const double Sold[3] = {sx[atom], sy[atom], sz[atom]};
    double Snew[3] = {Sold[0] + g[0] * angle, Sold[1] + g[1] * angle, Sold[2] + g[2] * angle};
    const double r = 1.0 / sqrt(Snew[0] * Snew[0] + Snew[1] * Snew[1] + Snew[2] * Snew[2]);
    Snew[0] *= r; Snew[1] *= r; Snew[2] *= r;
    const double Eold = or_spin_energy(s, c, sx, sy, sz, atom, Sold[0], Sold[1], Sold[2]);
    const double Enew = or_spin_energy(s, c, sx, sy, sz, atom, Snew[0], Snew[1], Snew[2]);
    const double DE = (Enew - Eold) * s->mat[imaterial].mu_s_SI * 1.07828231e23;
    int accept = DE < 0;
    if (!accept) accept = exp(-DE * kBTBohr_of(&s->mat[imaterial], c->temperature)) >= u;
    if (accept) { sx[atom] = Snew[0]; sy[atom] = Snew[1]; sz[atom] = Snew[2]; }
    return accept;
}

static void adapt(double* sigma, double moves, double rej) {
    const double last_rejection_rate = rej / moves;
    const double factor = 0.5 / last_rejection_rate;
    *sigma *= factor;
    if (*sigma > 60.0 || *sigma < 1e-5) *sigma = 60.0;
}

/* Colour-ordered sweeps: every atom once per sweep, colours in order; move normals from
 * Philox stream 1 and acceptance uniform (word 0 / 2^32) from stream 2 at (atom, sweep). */
void or_mc_checkerboard(const or_sys* s, const or_cond* c, double* sx, double* sy, double* sz,
                        const int* colour_list, const int* colour_start, int ncolours, int nsweeps,
                        uint64_t seed, uint64_t sweep0, double* sigma, double* tot_moves, double* tot_rej) {
    for (int w = 0; w < nsweeps; ++w) {
        const uint64_t sweep = sweep0 + (uint64_t)w;
        double rej = 0.0;
        for (int k = 0; k < ncolours; ++k) {
            for (int t = colour_start[k]; t < colour_start[k + 1]; ++t) {
                const int atom = colour_list[t];
                double g[3];
                or_normals3(seed, atom, sweep, 1u, g);
                uint32_t ua[4];
                philox_at(seed, (uint32_t)atom, sweep, 2u, ua);
                const double u = (double)ua[0] * 2.3283064365386963e-10;
                if (!metropolis_move(s, c, sx, sy, sz, atom, g, *sigma, u)) rej += 1.0;
            }
        }
        *tot_moves += (double)s->n;
        *tot_rej += rej;
        adapt(sigma, (double)s->n, rej);
        dipole_tick(s, sx, sy, sz);
    }
}

/* montecarlo::mc_step: N trials per sweep at random sites (with replacement).  The uniform
 * stream is Philox at (trial, sweep, stream 3..4): site = int(N u) from word 0, acceptance from
 * word 1; move normals from stream 5. */
void or_mc_random_site(const or_sys* s, const or_cond* c, double* sx, double* sy, double* sz, int nsweeps,
                       uint64_t seed, uint64_t sweep0, double* sigma, double* tot_moves, double* tot_rej) {
    const int nmoves = s->n;
    for (int w = 0; w < nsweeps; ++w) {
        const uint64_t sweep = sweep0 + (uint64_t)w;
        double rej = 0.0;
        for (int i = 0; i < nmoves; ++i) {
            uint32_t uu[4];
            philox_at(seed, (uint32_t)i, sweep, 3u, uu);
            const int atom = (int)(nmoves * ((double)uu[0] * 2.3283064365386963e-10));
            double g[3];
            or_normals3(seed, i, sweep, 5u, g);
            const double u = (double)uu[1] * 2.3283064365386963e-10;
            if (!metropolis_move(s, c, sx, sy, sz, atom, g, *sigma, u)) rej += 1.0;
        }
        *tot_moves += (double)nmoves;
        *tot_rej += rej;
        adapt(sigma, (double)nmoves, rej);
        dipole_tick(s, sx, sy, sz);
    }
}

/* ---- constrained Monte Carlo (montecarlo::cmc_step, src/montecarlo/cmc.cpp:282-497) ----------
 * Pair moves that keep the magnetisation along a constraint direction: spin 1 takes a Monte Carlo
 * move, spin 2 compensates the transverse magnetisation in the constraint frame, and the pair is
 * accepted with exp(-dE) (Mz'/Mz)^2 |s2z / s2z'| if Mz' > 0 (Asselin et al., PRB 82, 054415). */
typedef struct {
    double P[9], Pt[9], v[3];   /* cmc::polar_matrix, its transpose, cmc::polar_vector */
    int move;                   /* 0: cone of width sigma (montecarlo's adaptive move, whose width
                                   cmc_step never adapts); 1: the angle move with the width of
                                   cmc.cpp:320-325 per material */
    double sigma;
} or_cmc;

/* cmc::polar_rot_matrix (cmc.cpp:43-138): P = Y(phi) Z(theta) with vmath::matmul's sums from 0 in
 * k order, Pt its transpose, v = (0,0,1) P.  phi, theta in degrees. */
void or_cmc_frame(double phi, double theta, double P[9], double Pt[9], double v[3]) {
    const double pi = 3.14159265358979323846264338327;
    const double dy = (phi / 360.0) * 2.0 * pi, dz = (theta / 360.0) * 2.0 * pi;
    const double sin_y = sin(dy), cos_y = cos(dy), sin_z = sin(dz), cos_z = cos(dz);
    const double Y[9] = {cos_y, 0.0, -sin_y, 0.0, 1.0, 0.0, sin_y, 0.0, cos_y};
    const double Z[9] = {cos_z, sin_z, 0.0, -sin_z, cos_z, 0.0, 0.0, 0.0, 1.0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double c = 0.0;
            for (int k = 0; k < 3; ++k) c += Y[3 * i + k] * Z[3 * k + j];
            P[3 * i + j] = c;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Pt[3 * i + j] = P[3 * j + i];
    const double ref[3] = {0.0, 0.0, 1.0};
    for (int j = 0; j < 3; ++j) {
        double c = 0.0;
        for (int k = 0; k < 3; ++k) c += ref[k] * P[3 * k + j];
        v[j] = c;
    }
}

/* width of the angle move per material (cmc.cpp:320-325) */
static double cmc_angle_width(const or_material* m, double T) {
    const double alpha = m->temperature_rescaling_alpha, Tc = m->temperature_rescaling_Tc;
    const double rescaled = T < Tc ? Tc * pow(T / Tc, alpha) : T;
    const double kbt = 9.27400915e-24 / (rescaled * 1.3806503e-23);
    return rescaled < 1.0 ? 0.02 : pow(1.0 / kbt, 0.2) * 0.08;
}

static void rotate3(const double* M, const double* x, double* out) {   /* cmc.cpp:363-365 order */
    out[0] = M[0] * x[0] + M[1] * x[1] + M[2] * x[2];
    out[1] = M[3] * x[0] + M[4] * x[1] + M[5] * x[2];
    out[2] = M[6] * x[0] + M[7] * x[1] + M[8] * x[2];
}

/* One pair trial of cmc_step (cmc.cpp:355-494) with move normals g and acceptance uniform u.
 * M is the magnetisation the acceptance uses; with update_M it takes the accepted change, as
 * M_other does in the reference.  Returns 0 accepted, 1 energy reject, 2 sphere reject. */
static int cmc_pair(const or_sys* s, const or_cond* c, const or_cmc* k, double* sx, double* sy, double* sz,
                    int a1, int a2, const double g[3], double u, double M[3], int update_M) {
    const int imat1 = s->type[a1];
    const double mu1 = s->moment[a1];
    const double delta = k->move == 1 ? cmc_angle_width(&s->mat[imat1], c->temperature) : k->sigma;
    const double s1i[3] = {sx[a1], sy[a1], sz[a1]};
    double s1i_mvd[3], s1f_mvd[3], s2i_mvd[3], s2f_mvd[3];
    rotate3(k->P, s1i, s1i_mvd);
    double s1f[3] = {s1i[0] + g[0] * delta, s1i[1] + g[1] * delta, s1i[2] + g[2] * delta};   /* mc_angle */
    const double r = 1.0 / sqrt(s1f[0] * s1f[0] + s1f[1] * s1f[1] + s1f[2] * s1f[2]);
    s1f[0] *= r; s1f[1] *= r; s1f[2] *= r;
    rotate3(k->P, s1f, s1f_mvd);
    double Eold = or_spin_energy(s, c, sx, sy, sz, a1, sx[a1], sy[a1], sz[a1]);
    sx[a1] = s1f[0]; sy[a1] = s1f[1]; sz[a1] = s1f[2];
    double Enew = or_spin_energy(s, c, sx, sy, sz, a1, sx[a1], sy[a1], sz[a1]);
    const double delta_energy1 = (Enew - Eold) * s->mat[imat1].mu_s_SI * 1.07828231e23;
    const int imat2 = s->type[a2];
    const double mu2 = s->moment[a2];
    const double s2i[3] = {sx[a2], sy[a2], sz[a2]};
    rotate3(k->P, s2i, s2i_mvd);
    const double mom_ratio = (mu1 / mu2);
    s2f_mvd[0] = mom_ratio * (s1i_mvd[0] - s1f_mvd[0]) + s2i_mvd[0];
    s2f_mvd[1] = mom_ratio * (s1i_mvd[1] - s1f_mvd[1]) + s2i_mvd[1];
    if (((s2f_mvd[0] * s2f_mvd[0] + s2f_mvd[1] * s2f_mvd[1]) < 1.0) && (a1 != a2)) {
        s2f_mvd[2] = (s2i_mvd[2] < 0.0 ? -1.0 : 1.0) * sqrt(1.0 - s2f_mvd[0] * s2f_mvd[0] - s2f_mvd[1] * s2f_mvd[1]);
        double s2f[3];
        rotate3(k->Pt, s2f_mvd, s2f);
        Eold = or_spin_energy(s, c, sx, sy, sz, a2, sx[a2], sy[a2], sz[a2]);
        sx[a2] = s2f[0]; sy[a2] = s2f[1]; sz[a2] = s2f[2];
        Enew = or_spin_energy(s, c, sx, sy, sz, a2, sx[a2], sy[a2], sz[a2]);
        const double delta_energy2 = (Enew - Eold) * s->mat[imat2].mu_s_SI * 1.07828231e23;
        const double delta_energy21 = delta_energy1 * kBTBohr_of(&s->mat[imat1], c->temperature) +
                                      delta_energy2 * kBTBohr_of(&s->mat[imat2], c->temperature);
        const double* v = k->v;
        const double Mz_old = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
        const double Mz_new = (M[0] + mu1 * s1f[0] + mu2 * s2f[0] - mu1 * s1i[0] - mu2 * s2i[0]) * v[0] +
                              (M[1] + mu1 * s1f[1] + mu2 * s2f[1] - mu1 * s1i[1] - mu2 * s2i[1]) * v[1] +
                              (M[2] + mu1 * s1f[2] + mu2 * s2f[2] - mu1 * s1i[2] - mu2 * s2i[2]) * v[2];
        const double probability =
            exp(-delta_energy21) * ((Mz_new / Mz_old) * (Mz_new / Mz_old)) * fabs(s2i_mvd[2] / s2f_mvd[2]);
        if ((probability >= u) && (Mz_new > 0.0)) {
            if (update_M)
                for (int q = 0; q < 3; ++q) M[q] = M[q] + mu1 * s1f[q] + mu2 * s2f[q] - mu1 * s1i[q] - mu2 * s2i[q];
            return 0;
        }
        sx[a1] = s1i[0]; sy[a1] = s1i[1]; sz[a1] = s1i[2];
        sx[a2] = s2i[0]; sy[a2] = s2i[1]; sz[a2] = s2i[2];
        return 1;
    }
    sx[a1] = s1i[0]; sy[a1] = s1i[1]; sz[a1] = s1i[2];
    return 2;
}

static void cmc_moment(const or_sys* s, const double* sx, const double* sy, const double* sz, double M[3]) {
    M[0] = 0.0; M[1] = 0.0; M[2] = 0.0;                           /* cmc.cpp:346-351 */
    for (int atom = 0; atom < s->n; ++atom) {
        const double mu = s->moment[atom];
        M[0] = M[0] + sx[atom] * mu;
        M[1] = M[1] + sy[atom] * mu;
        M[2] = M[2] + sz[atom] * mu;
    }
}

/* cmc_step as the reference runs it: per sweep M from all spins, then N trials, each at two
 * random sites (with replacement); uniforms from Philox (trial, sweep, stream 6): site 1 from
 * word 0, site 2 from word 1, acceptance from word 2; move normals from stream 7.
 * counts[3] += accepted, energy rejects, sphere rejects. */
void or_cmc_random_pairs(const or_sys* s, const or_cond* c, const or_cmc* k, double* sx, double* sy, double* sz,
                         int nsweeps, uint64_t seed, uint64_t sweep0, double* counts) {
    const int n = s->n;
    for (int w = 0; w < nsweeps; ++w) {
        const uint64_t sweep = sweep0 + (uint64_t)w;
        double M[3];
        cmc_moment(s, sx, sy, sz, M);
        for (int t = 0; t < n; ++t) {
            uint32_t uu[4];
            philox_at(seed, (uint32_t)t, sweep, 6u, uu);
            const int a1 = (int)(n * ((double)uu[0] * 2.3283064365386963e-10));
            const int a2 = (int)(n * ((double)uu[1] * 2.3283064365386963e-10));
            double g[3];
            or_normals3(seed, t, sweep, 7u, g);
            counts[cmc_pair(s, c, k, sx, sy, sz, a1, a2, g, (double)uu[2] * 2.3283064365386963e-10, M, 1)] += 1.0;
        }
    }
}

/* The permutation of a round of the colour-paired sweep: from Philox (colour | round << 16, sweep,
 * stream 8), b = word 1 mod m and a = 1 + word 0 mod (m - 1), moved up (1 after m - 1) to the first
 * value coprime with m. */
void or_cmc_permutation(uint64_t seed, int colour, int round, uint64_t sweep, int m, uint64_t* a, uint64_t* b) {
    uint32_t w[4];
    philox_at(seed, (uint32_t)colour | ((uint32_t)round << 16), sweep, 8u, w);
    *b = (uint64_t)(w[1] % (uint32_t)m);
    uint32_t x = m > 1 ? 1u + w[0] % (uint32_t)(m - 1) : 1u;
    for (;;) {
        uint32_t p = x, q = (uint32_t)m;
        while (q) { const uint32_t t = p % q; p = q; q = t; }
        if (p == 1u) break;
        x = x + 1u >= (uint32_t)m ? 1u : x + 1u;
    }
    *a = x;
}

/* The colour-paired sweep the device runs.  Atoms of one colour do not interact, so pairs taken
 * within a colour have independent energy changes.  Per sweep, per colour k (list L_k, caller
 * order, m entries) and round r = 0, 1: the acceptance uses the magnetisation M of the spins at
 * the start of the round; half = m/2 pairs, pair p joins entries pi(2p) (spin 1) and pi(2p + 1)
 * (spin 2) of the random affine permutation pi(t) = (a t + b) mod m drawn for the round
 * (or_cmc_permutation): an entry's partner lies a entries away, and a changes from round to
 * round.  Move normals: Philox (spin 1's
 * index, sweep, stream 9 + r); acceptance: word 0 of (spin 1, sweep, stream 11 + r).
 * counts[3] += accepted, energy rejects, sphere rejects. */
void or_cmc_checkerboard(const or_sys* s, const or_cond* c, const or_cmc* k, double* sx, double* sy, double* sz,
                         const int* colour_list, const int* colour_start, int ncolours, int nsweeps, uint64_t seed,
                         uint64_t sweep0, double* counts) {
    for (int w = 0; w < nsweeps; ++w) {
        const uint64_t sweep = sweep0 + (uint64_t)w;
        for (int col = 0; col < ncolours; ++col) {
            const int* L = colour_list + colour_start[col];
            const int m = colour_start[col + 1] - colour_start[col];
            if (m < 2) continue;
            const int half = m / 2;
            for (int r = 0; r < 2; ++r) {
                uint64_t a, b;
                or_cmc_permutation(seed, col, r, sweep, m, &a, &b);
                double M[3];
                cmc_moment(s, sx, sy, sz, M);
                for (int p = 0; p < half; ++p) {
                    const int e1 = (int)((a * (uint64_t)(2 * p) + b) % (uint64_t)m);
                    const int e2 = (int)((a * (uint64_t)(2 * p + 1) + b) % (uint64_t)m);
                    const int a1 = L[e1], a2 = L[e2];
                    double g[3];
                    or_normals3(seed, a1, sweep, 9u + (uint32_t)r, g);
                    uint32_t ua[4];
                    philox_at(seed, (uint32_t)a1, sweep, 11u + (uint32_t)r, ua);
                    counts[cmc_pair(s, c, k, sx, sy, sz, a1, a2, g, (double)ua[0] * 2.3283064365386963e-10, M, 0)] += 1.0;
                }
            }
        }
    }
}

/* Greedy colouring in atom order over the symmetrised neighbour graph (the device's colouring).
 * colour[n] out; returns the number of colours. */
int or_greedy_colour(const or_sys* s, int* colour) {
    const int n = s->n;
    int nnz = 0;
    for (int i = 0; i < n; ++i) nnz += s->nbr_end[i] - s->nbr_start[i] + 1;
    int* rrow = (int*)calloc((size_t)n + 1, sizeof(int));
    int* rcol = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
    int* fill = (int*)malloc(sizeof(int) * (size_t)n);
    int* seen = (int*)malloc(sizeof(int) * 1024);
    for (int i = 0; i < n; ++i)
        for (int e = s->nbr_start[i]; e <= s->nbr_end[i]; ++e) rrow[s->nbr[e] + 1]++;
    for (int i = 0; i < n; ++i) rrow[i + 1] += rrow[i];
    for (int i = 0; i < n; ++i) fill[i] = rrow[i];
    for (int i = 0; i < n; ++i)
        for (int e = s->nbr_start[i]; e <= s->nbr_end[i]; ++e) rcol[fill[s->nbr[e]]++] = i;
    for (int k = 0; k < 1024; ++k) seen[k] = -1;
    int nc = 0;
    for (int i = 0; i < n; ++i) colour[i] = -1;
    for (int i = 0; i < n; ++i) {
        for (int e = s->nbr_start[i]; e <= s->nbr_end[i]; ++e) {
            const int j = s->nbr[e];
            if (j != i && colour[j] >= 0) seen[colour[j]] = i;
        }
        for (int e = rrow[i]; e < rrow[i + 1]; ++e) {
            const int j = rcol[e];
            if (j != i && colour[j] >= 0) seen[colour[j]] = i;
        }
        int k = 0;
        while (k < 1024 && seen[k] == i) ++k;
        colour[i] = k;
        if (k + 1 > nc) nc = k + 1;
    }
    free(rrow); free(rcol); free(fill); free(seen);
    return nc;
}
