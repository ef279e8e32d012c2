// vmx_internal.hpp — launchers exported by the precision units (vmx_f64.hip, vmx_f32.hip) to
// the context code (vmx_api.hip).  Each launcher picks the kernel instantiation for the
// exchange type and J-storage mode of the context and queues it on `stream`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vmx_device.hpp"

namespace vmx {

struct Launch {
    int grid;        // hardware workgroups (padded to a multiple of 8 when XCD remapping)
    int block;       // threads per workgroup
    hipStream_t stream;
};

template <typename R>
struct Ops {
    // fused Heun step: predictor writes P from S; corrector writes O from S and P.  The Philox
    // step of a launch is *s.step_base + step_off.
    // D: Euler derivative (FP64) or S + dt/2 dS (FP32) kept by the predictor for the corrector;
    // E: external field kept the same way.  All-null arrays make the corrector recompute them.
    static hipError_t fused_predictor(const Launch& L, int et, int jm, const SysDev<R>& s, Spin3<R> S,
                                      Spin3<R> P, Spin3<R> D, Spin3<R> E, const double* noise, uint32_t step_off);
    static hipError_t fused_corrector(const Launch& L, int et, int jm, const SysDev<R>& s, Spin3<R> S,
                                      Spin3<R> P, Spin3<R> O, Spin3<R> D, Spin3<R> E, const double* noise,
                                      uint32_t step_off);
    // LLG midpoint step: predictor writes the midpoint P from S (E: H_ext kept, or null);
    // corrector writes O from S and P
    static hipError_t mid_predictor(const Launch& L, int et, int jm, const SysDev<R>& s, Spin3<R> S, Spin3<R> P,
                                    Spin3<R> E, const double* noise, uint32_t step_off);
    static hipError_t mid_corrector(const Launch& L, int et, int jm, const SysDev<R>& s, Spin3<R> S, Spin3<R> P,
                                    Spin3<R> O, Spin3<R> E, const double* noise, uint32_t step_off);
    // plain Heun step
    static hipError_t spin_fields(const Launch& L, int et, int jm, const SysDev<R>& s, Spin3<R> S, Spin3<R> F);
    static hipError_t ext_fields(const Launch& L, const SysDev<R>& s, Spin3<R> E, const double* noise,
                                 uint32_t step_off);
    static hipError_t plain_predictor(const Launch& L, const SysDev<R>& s, Spin3<R> S, Spin3<R> F,
                                      Spin3<R> E, Spin3<R> D, Spin3<R> P);
    static hipError_t plain_corrector(const Launch& L, const SysDev<R>& s, Spin3<R> S, Spin3<R> P,
                                      Spin3<R> F, Spin3<R> E, Spin3<R> D);
    // statistics partial sums (grid-stride, `grid` workgroups of 256)
    static hipError_t mag_partial(int grid, hipStream_t st, int n, Spin3<R> S, const int* mask,
                                  Moments mo, int nbins, double* partial);
    static hipError_t energy_partial(int grid, hipStream_t st, int et, int jm, const SysDev<R>& s, Spin3<R> S,
                                     const int* mask, Moments mo, const int* atom_rep, int nbins,
                                     double* partial);
    // Monte Carlo: one launch of a sweep, workgroups wg0 .. wg0 + nwg - 1; workgroup w takes
    // entries [w * 256 * McAtoms, ...) of atom (and pat, the stencil ids) and wg[w] = {replica,
    // replica's first atom}
    // Monte Carlo, colour order: workgroup w takes the atoms [wg[w].x, wg[w].z) of replica wg[w].y
    static hipError_t mc_colour_ordered(hipStream_t st, int et, int jm, const SysDev<R>& s, Spin3<R> S,
                                        const int* lidx, const int4* wg, int wg0, int nwg, const double* sigma,
                                        uint64_t sweep, double* rejects);
    // spins to colour order (dst[p] = src[perm[p]]) or back (dst[perm[p]] = src[p])
    static hipError_t mc_permute(hipStream_t st, int n, const int* perm, Spin3<R> src, Spin3<R> dst,
                                 bool to_colour);
    static hipError_t mc_colour(hipStream_t st, int et, int jm, const SysDev<R>& s, Spin3<R> S, const int* atom,
                                const uint16_t* pat, const int2* wg, int wg0, int nwg, const double* sigma,
                                uint64_t sweep, double* rejects);
    // constrained Monte Carlo: one round of one colour (tab[ntab] workgroups of up to 256 pairs)
    // lidx: each atom's caller index within its replica when the spins are in colour order (else null)
    static hipError_t cmc_pairs(hipStream_t st, int et, int jm, const SysDev<R>& s, Spin3<R> S, const double* mm,
                                const int* list, const int* lidx, const int4* tab, int ntab, const int4* perm,
                                const CmcFrame<R>* frame,
                                const double* M, const double* sigma, int move, int round, uint64_t sweep,
                                double* counts, double* dM);
    // the permutations of every (colour, round, replica) of a sweep: perm[(2 k + round) nrep + r] =
    // {a, b, a^-1 mod m, 0} for the run length m = mlen[k nrep + r] (m < 2: {1, 0, 1, 0})
    static hipError_t cmc_perm(hipStream_t st, const uint2* rep_key, int nrep, int ncolours, const int* mlen,
                               uint64_t sweep, int4* perm);
    // each replica's sum(mu s) (double [nrep][3]): per-workgroup partials over the one-atom block
    // table, then one workgroup per replica over its partials (rep_blk[r] .. rep_blk[r+1])
    // (gather: spins in colour order, atom i's at gather[i]; else null)
    static hipError_t replica_moment(hipStream_t st, const SysDev<R>& s, Spin3<R> S, Moments mo,
                                     const int* gather, const int* rep_blk, int nrep, double* partial, double* M);
    // M[r] += the sum of the cmc_pairs partials dM[range[r] .. range[r+1]], in a fixed order
    static hipError_t moment_add(hipStream_t st, const int* range, int nrep, const double* dM, double* M);
    // deterministic fitting step and its adjoint (isotropic uniform exchange); grids of 256
    static hipError_t fit_predictor(hipStream_t st, int am, const SysDev<R>& s, const FitPrm<R>& p, Spin3<R> h,
                                    Spin3<R> S, Spin3<R> P);
    static hipError_t fit_corrector(hipStream_t st, int am, const SysDev<R>& s, const FitPrm<R>& p, Spin3<R> h,
                                    Spin3<R> S, Spin3<R> P, Spin3<R> O);
    static hipError_t adj_b1(hipStream_t st, int am, const SysDev<R>& s, const FitPrm<R>& p, Spin3<R> h,
                             Spin3<R> S, Spin3<R> P, Spin3<R> lam, Spin3<R> lH1, Spin3<R> lPd, Spin3<R> lT,
                             Spin3<R> H0, Spin3<R> sumS, R* nhat, double* acc);
    static hipError_t adj_b2(hipStream_t st, int am, const SysDev<R>& s, const FitPrm<R>& p, Spin3<R> S,
                             Spin3<R> P, Spin3<R> lH1, Spin3<R> lPd, Spin3<R> lT, Spin3<R> H0, Spin3<R> sumS,
                             const R* nhat, Spin3<R> lH0, Spin3<R> lSp, Spin3<R> dh, double* acc);
    static hipError_t adj_b3(hipStream_t st, int am, const SysDev<R>& s, const FitPrm<R>& p, Spin3<R> lH0,
                             Spin3<R> lSp, Spin3<R> lam, R inj);
    static hipError_t fit_mz(hipStream_t st, int n, Spin3<R> S, double* out4);
    // Philox normals of `step` as double [3][n]
    static hipError_t thermal_noise(const Launch& L, const SysDev<R>& s, uint64_t step, double* out);
    // conversions between the double host/staging layout and R device arrays
    static hipError_t from_double(hipStream_t st, int n, const double* in, R* out);
    static hipError_t to_double(hipStream_t st, int n, const R* in, double* out);
};

extern template struct Ops<double>;
extern template struct Ops<float>;

}  // namespace vmx
