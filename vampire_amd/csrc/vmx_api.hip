// vmx_api.hip — the C ABI of include/vmx.h: context set-up, host-side validation and
// compression of the system, stepping, statistics, Monte Carlo, graphs and autotuning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vmx.h"
#include "vmx_internal.hpp"

using namespace vmx;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define VMX_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(VMX_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                               \
    } while (0)

#define VMX_TRY(expr)                  \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != VMX_OK) return rc_; \
    } while (0)

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t b) {
        if (p && bytes >= b) return VMX_OK;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        if (b == 0) b = 16;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) { p = nullptr; return fail(VMX_ERR_NOMEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e)); }
        bytes = b;
        return VMX_OK;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Copies a host vector into b (allocated to fit, at least one element).
template <typename T>
int upload(DevBuf& b, const std::vector<T>& v) {
    VMX_TRY(b.alloc(std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) VMX_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return VMX_OK;
}

// Pinned staging ring for small host->device updates (conditions, sigma): a slot is reused only
// after the copy that read it has completed.
struct PinnedRing {
    static constexpr int kSlots = 32;
    void* slot[kSlots] = {};
    hipEvent_t done[kSlots] = {};
    size_t bytes = 0;
    int next = 0;
    ~PinnedRing() {
        for (int k = 0; k < kSlots; ++k) {
            if (slot[k]) (void)hipHostFree(slot[k]);
            if (done[k]) (void)hipEventDestroy(done[k]);
        }
    }
    int init(size_t b) {
        bytes = b ? b : 16;
        for (int k = 0; k < kSlots; ++k) {
            VMX_HIP(hipHostMalloc(&slot[k], bytes, hipHostMallocDefault));
            VMX_HIP(hipEventCreateWithFlags(&done[k], hipEventDisableTiming));
            VMX_HIP(hipEventRecord(done[k], nullptr));
        }
        return VMX_OK;
    }
    int copy(void* dst, const void* src, size_t b, hipStream_t st) {
        if (b > bytes) return fail(VMX_ERR_ARG, "staging copy of %zu bytes exceeds %zu", b, bytes);
        const int k = next;
        next = (next + 1) % kSlots;
        VMX_HIP(hipEventSynchronize(done[k]));
        std::memcpy(slot[k], src, b);
        VMX_HIP(hipMemcpyAsync(dst, slot[k], b, hipMemcpyHostToDevice, st));
        VMX_HIP(hipEventRecord(done[k], st));
        return VMX_OK;
    }
};

// Small host->device updates (a few condition records) written by one workgroup whose kernel
// arguments carry the bytes: they travel in the dispatch packet and the write is ordered on the
// stream like the copy it replaces, without a DMA transfer between kernel launches.
constexpr int kArgWords = 256;                    // 1 KiB of kernel arguments
struct ArgWords { uint32_t w[kArgWords]; };
__global__ void __launch_bounds__(kArgWords) k_put_words(uint32_t* __restrict__ dst, ArgWords v, int n) {
    const int k = (int)threadIdx.x;
    if (k < n) dst[k] = v.w[k];
}

// ---- scratch of the statistics update that allocates nothing -----------------------------------
// vmx_stats_update_noalloc keeps its per-workgroup partial sums in this file-scope device array
// instead of a buffer the library allocates.  Every GPU the library runs on holds its own copy
// until the process ends; calls on one GPU share it (the last call's partials stay in it), so
// calls that overlap on one GPU race.  Host code never passes the array's name to a kernel: its
// address on a device comes from hipGetSymbolAddress, taken once per device when the first
// context is created there (scratch_on_current_device).
constexpr size_t kStatsScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) double scratch[kStatsScratchBytes / sizeof(double)];

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;   // device ordinal -> that device's copy of `scratch`

int scratch_on_current_device(double** out) {
    int dev = 0;
    VMX_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        VMX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(VMX_ERR_HIP, "hipGetSymbolAddress returned no address for the statistics scratch");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return VMX_OK;
}

struct MagStat {
    int nbins;
    DevBuf mask, cur, mean;
    std::vector<int> empty;  // mask ids (< nbins-1) without atoms
    DevBuf empty_dev;
    DevBuf counter;          // int64 on device
};

struct EnergyStat {
    int nbins;
    DevBuf mask, cur, mean, empty_dev, counter;
    std::vector<int> empty;
};

__global__ void k_advance_step(unsigned long long* base, unsigned int n) { *base += n; }
__global__ void k_set_step(unsigned long long* base, unsigned long long v) { *base = v; }

// adds dL/dS_z (the same for every atom) to the adjoint of a recorded state
template <typename R>
__global__ void k_lam_inject(int n, R* __restrict__ lz, const R* __restrict__ inj) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const R v = *inj;
    if (i < n && v != R(0)) lz[i] += v;
}

template <typename R>
__global__ void k_fill(size_t n, R* __restrict__ a, R v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// Reduces the workgroup partials in block order, then normalises (magnetization.cpp:150-178).
// Sum over the nblk per-workgroup partials [nblk][W] of every value k < W, handed to put(k, sum).
// With fewer values than threads, each value's partials are split over blockDim / W threads whose
// sub-sums are then added in thread order: the whole block works and the order is fixed, so the
// result is the same on every run.  With many values, one thread per value.
template <typename F>
__device__ void reduce_partials(int nblk, const double* __restrict__ partial, int W, F&& put) {
    __shared__ double red[256];                   // the finish kernels run one block of 256 threads
    const int per = (int)blockDim.x / W;
    if (per >= 2) {
        const int k = (int)threadIdx.x % W, j = (int)threadIdx.x / W;
        double a = 0.0;
        if (j < per) {
#pragma unroll 8
            for (int b = j; b < nblk; b += per) a += partial[(size_t)b * W + k];
        }
        red[threadIdx.x] = a;
        __syncthreads();
        if ((int)threadIdx.x < W) {
            double t = 0.0;
            for (int q = 0; q < per; ++q) t += red[q * W + (int)threadIdx.x];
            put((int)threadIdx.x, t);
        }
    } else {
        for (int k = threadIdx.x; k < W; k += blockDim.x) {
            double acc = 0.0;
            for (int b = 0; b < nblk; ++b) acc += partial[(size_t)b * W + k];
            put(k, acc);
        }
    }
}

__global__ void k_mag_finish(int nblk, const double* __restrict__ partial, int nbins, double* cur,
                             double* mean, const int* __restrict__ empty, int nempty,
                             long long* counter) {
    const int W = 4 * nbins;
    reduce_partials(nblk, partial, W, [&](int k, double v) { cur[k] = v; });
    __threadfence_block();
    __syncthreads();
    for (int id = threadIdx.x; id < nbins - 1; id += blockDim.x) {
        const double msat = cur[4 * id + 3];
        const double magm = sqrt(cur[4 * id + 0] * cur[4 * id + 0] + cur[4 * id + 1] * cur[4 * id + 1] +
                                 cur[4 * id + 2] * cur[4 * id + 2]);
        cur[4 * id + 0] = cur[4 * id + 0] / magm;
        cur[4 * id + 1] = cur[4 * id + 1] / magm;
        cur[4 * id + 2] = cur[4 * id + 2] / magm;
        cur[4 * id + 3] = magm / msat;
    }
    __threadfence_block();
    __syncthreads();
    for (int e = threadIdx.x; e < nempty; e += blockDim.x) {
        const int id = empty[e];
        cur[4 * id + 0] = 0.0; cur[4 * id + 1] = 0.0; cur[4 * id + 2] = 0.0; cur[4 * id + 3] = 0.0;
    }
    __threadfence_block();
    __syncthreads();
    for (int k = threadIdx.x; k < W; k += blockDim.x) mean[k] += cur[k];
    if (threadIdx.x == 0) *counter += 1;
}

// Energies [5][nbins]: total, exchange, anisotropy, applied, magnetostatic
// (src/statistics/energy.cpp:141-258; the magnetostatic term is zero without a device dipole field).
__global__ void k_energy_finish(int nblk, const double* __restrict__ partial, int nbins, double* cur,
                                double* mean, const int* __restrict__ empty, int nempty,
                                long long* counter) {
    const int W = 4 * nbins;
    reduce_partials(nblk, partial, W, [&](int k, double v) {
        const int id = k / 4, term = k % 4;     // term: exchange, anisotropy, applied, magnetostatic
        cur[(size_t)(1 + term) * nbins + id] = v;
    });
    for (int id = threadIdx.x; id < nbins; id += blockDim.x) cur[id] = 0.0;
    __threadfence_block();
    __syncthreads();
    const int ms = nbins - 1;
    for (int id = threadIdx.x; id < ms; id += blockDim.x) {
        cur[(size_t)1 * nbins + id] = 0.5 * cur[(size_t)1 * nbins + id];
        cur[(size_t)4 * nbins + id] = 0.5 * cur[(size_t)4 * nbins + id];
        cur[id] = cur[(size_t)1 * nbins + id] + cur[(size_t)2 * nbins + id] + cur[(size_t)3 * nbins + id] +
                  cur[(size_t)4 * nbins + id];
        for (int t = 0; t < 5; ++t) mean[(size_t)t * nbins + id] += cur[(size_t)t * nbins + id];
    }
    __threadfence_block();
    __syncthreads();
    for (int e = threadIdx.x; e < nempty; e += blockDim.x) {
        const int id = empty[e];
        for (int t = 0; t < 5; ++t) cur[(size_t)t * nbins + id] = 0.0;
    }
    if (threadIdx.x == 0) *counter += 1;
}

// Adaptive cone width after a sweep (src/montecarlo/mc.cpp:112-120) and running totals.  A
// replica's rejections arrive in kRejectSlots counters (integers held in doubles: exact in any
// order), which are summed and cleared here: one wavefront per replica, a slot per lane.
static_assert(kRejectSlots == 64, "one rejection slot per lane of a wavefront");
__global__ void __launch_bounds__(256) k_mc_adapt(int nrep, const int* __restrict__ rep_off, double* sigma,
                                                  double* rejects, double* tot_moves, double* tot_rej) {
    const int r = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);   // uniform in a wavefront
    if (r >= nrep) return;
    const int lane = (int)(threadIdx.x & 63u);
    double* slot = rejects + ((size_t)r * kRejectSlots + lane) * kRejectStride;
    double rej = *slot;
    *slot = 0.0;
    for (int off = 32; off > 0; off >>= 1) rej += __shfl_xor(rej, off);
    if (lane != 0) return;
    const double moves = (double)(rep_off[r + 1] - rep_off[r]);
    const double last_rejection_rate = rej / moves;
    const double factor = 0.5 / last_rejection_rate;
    double s = sigma[r] * factor;
    if (s > 60.0 || s < 1e-5) s = 60.0;
    sigma[r] = s;
    tot_moves[r] += moves;
    tot_rej[r] += rej;
}

// ---- macrocell dipole field, tensor solver ------------------------------------------------
// Cell moments, cells::mag (src/cells/mag.cpp:53-66): one thread per cell sums its magnetic
// atoms in ascending caller index, the order of the reference's atom loop.
template <typename R>
__global__ void k_dip_cell_mag(int nc, const int* __restrict__ start, const int* __restrict__ atoms,
                               const uint8_t* __restrict__ type, const double* __restrict__ mus,
                               const R* __restrict__ sx, const R* __restrict__ sy, const R* __restrict__ sz,
                               double* __restrict__ mag) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double mx = 0.0, my = 0.0, mz = 0.0;
    // K atoms at a time: their indices, then their types and spins, then the moments are loaded
    // before any is added, so each batch waits three memory round trips instead of each atom
    constexpr int K = 16;
    const int end = start[c + 1];
    int a = start[c];
    for (; a + K <= end; a += K) {
        int i[K], ty[K];
        double m[K], vx[K], vy[K], vz[K];
#pragma unroll
        for (int u = 0; u < K; ++u) i[u] = atoms[a + u];
#pragma unroll
        for (int u = 0; u < K; ++u) {
            ty[u] = type[i[u]];
            vx[u] = (double)sx[i[u]]; vy[u] = (double)sy[i[u]]; vz[u] = (double)sz[i[u]];
        }
#pragma unroll
        for (int u = 0; u < K; ++u) m[u] = mus[ty[u]];
#pragma unroll
        for (int u = 0; u < K; ++u) {
            mx += vx[u] * m[u];
            my += vy[u] * m[u];
            mz += vz[u] * m[u];
        }
    }
    for (; a < end; ++a) {
        const int i = atoms[a];
        const double m = mus[type[i]];
        mx += (double)sx[i] * m;
        my += (double)sy[i] * m;
        mz += (double)sz[i] * m;
    }
    mag[c] = mx; mag[(size_t)nc + c] = my; mag[2 * (size_t)nc + c] = mz;
}

// Cell fields, dipole::internal::update_field (src/dipole/update.cpp:52-93): one thread per field
// component of each cell with magnetic atoms, the j loop and every sum in the reference's order.
// T holds the tensor as [xx, xy, xz, yy, yz, zz][source cell j][row lc] so that a wavefront reads
// consecutive rows.
__global__ void __launch_bounds__(64) k_dip_cell_field(int nl, int nc, const int* __restrict__ local_cell,
                                 const int* __restrict__ count, const double* __restrict__ volume,
                                 const double* __restrict__ mag, const double* __restrict__ T,
                                 double* __restrict__ B, double* __restrict__ Hd) {
    // One thread per (local cell, field component r): component r of B and Hd sums tensor row r
    // only, so the three rows run in separate threads (three times the wavefronts) and every sum
    // keeps its terms and order.
    const int lc = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = (int)blockIdx.y;
    if (lc >= nl) return;
    const int i = local_cell[lc];
    if (count[i] <= 0) return;
    const double imuB = 1.0 / 9.27400915e-24;
    const size_t ncs = (size_t)nc, ts = (size_t)nc * nl;
    const double eightPI_three_cell_volume = 8.0 * M_PI / (3.0 * volume[i]);
    const double self_demag = eightPI_three_cell_volume;
    const double m_i = mag[r * ncs + i] * imuB;
    double bf = self_demag * m_i * 0.0;
    double hf = -0.5 * self_demag * m_i;
    // row r of [xx xy xz; xy yy yz; xz yz zz] in the [xx, xy, xz, yy, yz, zz] storage
    const int c0 = r == 0 ? 0 : (r == 1 ? 1 : 2);
    const int c1 = r == 0 ? 1 : (r == 1 ? 3 : 4);
    const int c2 = r == 0 ? 2 : (r == 1 ? 4 : 5);
    const double* __restrict__ T0 = T + c0 * ts;
    const double* __restrict__ T1 = T + c1 * ts;
    const double* __restrict__ T2 = T + c2 * ts;
    // Source cells in order (the reference's summation order), K at a time: the loads of all K
    // are issued before the first is added, so the loop waits for memory once per K cells
    // instead of once per cell.  Empty cells are loaded and skipped, as the reference skips them.
    auto add = [&](int cnt, double gx, double gy, double gz, double t0, double t1, double t2) {
        if (cnt > 0) {
            const double mx = gx * imuB, my = gy * imuB, mz = gz * imuB;
            bf += (mx * t0 + my * t1 + mz * t2);
            hf += (mx * t0 + my * t1 + mz * t2);
        }
    };
    constexpr int K = 16;
    int j = 0;
    for (; j + K <= nc; j += K) {
        int cnt[K];
        double gx[K], gy[K], gz[K], t0[K], t1[K], t2[K];
#pragma unroll
        for (int u = 0; u < K; ++u) {
            cnt[u] = count[j + u];
            gx[u] = mag[j + u]; gy[u] = mag[ncs + j + u]; gz[u] = mag[2 * ncs + j + u];
            const size_t o = (size_t)(j + u) * nl + lc;
            t0[u] = T0[o]; t1[u] = T1[o]; t2[u] = T2[o];
        }
#pragma unroll
        for (int u = 0; u < K; ++u) add(cnt[u], gx[u], gy[u], gz[u], t0[u], t1[u], t2[u]);
    }
    for (; j < nc; ++j) {
        const size_t o = (size_t)j * nl + lc;
        add(count[j], mag[j], mag[ncs + j], mag[2 * ncs + j], T0[o], T1[o], T2[o]);
    }
    B[r * ncs + i] = bf * 9.27400915e-01;
    Hd[r * ncs + i] = hf * 9.27400915e-01;
}

// Atom fields from cell fields (src/dipole/field.cpp:113-125): atoms of non-magnetic materials
// (cell -1) keep their zero field.
template <typename R>
__global__ void k_dip_unroll(int n, int nc, const int* __restrict__ acell, const double* __restrict__ B,
                             const double* __restrict__ Hd, R* __restrict__ out, double* __restrict__ hd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = acell[i];
    if (c < 0) return;
    out[i] = (R)B[c];
    out[(size_t)n + i] = (R)B[(size_t)nc + c];
    out[2 * (size_t)n + i] = (R)B[2 * (size_t)nc + c];
    hd[i] = Hd[c];
    hd[(size_t)n + i] = Hd[(size_t)nc + c];
    hd[2 * (size_t)n + i] = Hd[2 * (size_t)nc + c];
}

}  // namespace

// ============================================================================================
struct vmx_ctx {
    int device = -1;
    hipStream_t stream = nullptr;
    int precision = VMX_FP64;
    int kernel = VMX_KERNEL_FUSED;
    int block = 256;
    int xcd_remap = 1;
    uint64_t seed = 0;

    int n = 0, nmat = 0, nrep = 1;
    int64_t nnz = 0;
    int et = 0, jm = 0, K = 1;
    int tp = TP_CSR, zmax = 0, npat = 0, wsel = 0;   // zmax: row stride W for ELL / stencil
    int full_rows = 0;                               // ELL / stencil: every row has W neighbours
    // internal atom order: with lattice period q > 1 atoms are stored basis-major inside each
    // replica; pi maps a caller index to the internal one, inv back (both empty when q == 1)
    int q = 1;
    std::vector<int> pi, inv;
    // k consecutive blocks of n values, caller order -> internal order and back
    void to_internal(const double* in, double* out, int k) const {
        for (int b = 0; b < k; ++b)
            for (int i = 0; i < n; ++i) out[(size_t)b * n + (q > 1 ? pi[i] : i)] = in[(size_t)b * n + i];
    }
    void to_caller(const double* in, double* out, int k) const {
        for (int b = 0; b < k; ++b)
            for (int i = 0; i < n; ++i) out[(size_t)b * n + i] = in[(size_t)b * n + (q > 1 ? pi[i] : i)];
    }
    int am() const { return am_of(tp, jm, wsel); }
    std::vector<int> rowh_, colh_;   // host CSR (graph colouring for Monte Carlo)
    std::vector<int> type_h;         // material ids, caller order
    uint32_t ani = 0, fld = 0;
    double dt = 0.0, half_dt = 0.0;
    uint64_t step = 0;       // Heun / midpoint steps taken (the Philox step counter)
    // The value *step_base holds once the work queued so far has run.  Eager launches pass their
    // step as an offset from it (step - dev_step + k), so they need no device-side advance; a
    // graph replay reads and advances the device counter itself, so it is brought to `step`
    // first (sync_step_base).
    uint64_t dev_step = 0;
    int sync_step_base() {
        if (dev_step == step) return VMX_OK;
        hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, stream, step_base.as<unsigned long long>(),
                           (unsigned long long)step);
        VMX_HIP(hipGetLastError());
        dev_step = step;
        return VMX_OK;
    }
    // the offset of an eager call's first step from *step_base, with room for nsteps more
    int eager_step_off(int nsteps, uint32_t& off) {
        if (step - dev_step + (uint64_t)nsteps >= (1ull << 31)) VMX_TRY(sync_step_base());
        off = (uint32_t)(step - dev_step);
        return VMX_OK;
    }
    uint64_t sweep = 0;

    std::vector<vmx_material> mats;
    std::vector<int> rep_off_h;
    std::vector<double> T_h;      // per replica
    std::vector<double> H_h;      // per replica [3]
    std::vector<double> Hd_h;     // per replica [3]: thin-film demag field, LLG applied field only
    bool demag_on() const {
        for (double v : Hd_h) if (v != 0.0) return true;
        return false;
    }

    // device state
    DevBuf row, col, jidx, jtab, type, mat, cond, blk, blk_fused, rep_off, rep_key, step_base;
    std::vector<uint2> rep_key_h;    // the replicas' Philox keys (seed + r)
    DevBuf deg, pat, pat_off, pat_jid, pat_deg;
    DevBuf rep_m;                 // [nrep] replica size / q
    DevBuf rep_blk;               // [nrep+1] each replica's first entry of the one-atom block table
    DevBuf spin[3];               // three buffers of 3N R each (S, P, O roles rotate)
    int s_idx = 0, o_idx = 2;     // current spin buffer and the corrector's output buffer
    // The fused correctors read their own atom's S and gather only S' (or the midpoint), so they
    // can write S_next over S, and a step touches two spin arrays instead of three; the FP64
    // corrector that recomputes dS gathers S and cannot.
    bool corrector_in_place() const { return precision == VMX_FP32 || store_euler; }
    DevBuf plainF, plainE, plainD;
    int store_euler = 1;          // fused FP64 step: the predictor keeps dS (FP32 keeps nothing)
    DevBuf euler_buf;             // 3N R
    int store_field = 0;          // fused step: predictor keeps H_ext for the corrector
    DevBuf field_buf;             // 3N R
    DevBuf mm, atom_rep;
    DevBuf mm_tab;                   // [nmat] each material's moment, when every atom has its material's
    // the moments the statistics read: the per-material table when it describes every atom
    Moments moments() const {
        return Moments{mm.as<double>(), type.as<uint8_t>(), mm_tab.p ? mm_tab.as<double>() : nullptr};
    }
    DevBuf staging;               // 3N doubles for host<->device spin copies and noise
    DevBuf noise_dev;             // 3N doubles
    // FP32 contexts: draw the thermal normals with the FP64 kernels' Box-Muller into noise_dev
    // (one more launch per step) and feed them to the FP32 step as injected noise, so that FP32
    // and FP64 runs consume the same normals (to float rounding)
    int fp64_normals = 0;
    std::vector<double> noise_host;   // one step of caller noise in internal order
    // a per-atom int array in internal order
    std::vector<int> internal_ints(const int32_t* a) const {
        std::vector<int> out(a, a + n);
        if (q > 1)
            for (int i = 0; i < n; ++i) out[pi[i]] = a[i];
        return out;
    }
    DevBuf partial;               // statistics partial sums
    int nblocks = 0;
    int nblocks_fused = 0;        // entries of blk_fused (each FusedAtoms * block atoms)
    std::vector<MagStat*> mags;
    std::vector<EnergyStat*> energies;

    // Macrocell dipole field.  hdip ([3][n] R, internal order) is added to H_ext when hdip_on;
    // with dip.on the device solver refreshes it every dip.rate Heun steps.
    struct Dipole {
        int on = 0;
        int nc = 0, nl = 0, rate = 1;
        long long time = 0;       // Heun steps and Metropolis sweeps since the field was set up
        DevBuf start, atoms;      // per cell: its magnetic atoms (internal index), ascending caller index
        DevBuf local_cell, count, volume, tensor, mag, B, Hd, acell, mus;
        DevBuf hd_atom;           // [3][n] doubles, internal order: mu0 Hd for the energies
        std::vector<int> acell_h; // internal order, -1: no field
    } dip;
    int hdip_on = 0;
    DevBuf hdip;
    void drop_graph() {
        if (graph_exec) { (void)hipGraphExecDestroy(graph_exec); graph_exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
        graph_key = -1;
    }
    template <typename R>
    int dipole_refresh() {
        const Spin3<R> S = spins<R>(s_idx);
        hipLaunchKernelGGL((k_dip_cell_mag<R>), dim3((dip.nc + 63) / 64), dim3(64), 0, stream, dip.nc,
                           dip.start.as<int>(), dip.atoms.as<int>(), type.as<uint8_t>(), dip.mus.as<double>(),
                           S.x, S.y, S.z, dip.mag.as<double>());
        hipLaunchKernelGGL(k_dip_cell_field, dim3((dip.nl + 63) / 64, 3), dim3(64), 0, stream, dip.nl, dip.nc,
                           dip.local_cell.as<int>(), dip.count.as<int>(), dip.volume.as<double>(),
                           dip.mag.as<double>(), dip.tensor.as<double>(), dip.B.as<double>(), dip.Hd.as<double>());
        hipLaunchKernelGGL((k_dip_unroll<R>), dim3((n + 255) / 256), dim3(256), 0, stream, n, dip.nc,
                           dip.acell.as<int>(), dip.B.as<double>(), dip.Hd.as<double>(), hdip.as<R>(),
                           dip.hd_atom.as<double>());
        VMX_HIP(hipGetLastError());
        return VMX_OK;
    }

    // Monte Carlo
    bool coloured = false;
    int ncolours = 0;
    std::vector<int> colour_start;
    // constrained Monte Carlo: per colour, per replica, the colour's atoms in caller order (the
    // pairing follows that order); workgroup table {first entry, replica, run length, first pair}
    DevBuf cmc_list, cmc_tab, cmc_frame, cmc_M, cmc_part, cmc_counts, cmc_dm, cmc_range, cmc_mlen, cmc_perm;
    std::vector<int> cmc_mlen_h;           // [colour][replica] run lengths
    DevBuf cmc_list_co, mco_invp, mco_mm;  // colour order: the pair list's atoms, storage -> colour
                                           // position, moments
    std::vector<int> mco_invp_h;
    std::vector<int> cmc_tab_start;        // colour k's workgroups: cmc_tab[cmc_tab_start[k] ...]
                                           // cmc_range[k * (nrep + 1) + r]: replica r's first of them
    std::vector<double> cmc_frame_h;       // [nrep][21]: polar matrix, transpose, constraint direction
    std::vector<char> cmc_frame_set;       // per replica
    bool cmc_frame_dirty = true;
    int cmc_move = 0;                      // 0: cone of the replica's Metropolis width; 1: angle move
    template <typename R>
    int cmc_sweeps(int nsweeps);
    int cmc_permutations(uint64_t sw);
    int colour_order(const std::vector<int>& c, const std::vector<int>& lau, int nc);
    template <typename R>
    SysDev<R> sys_colour_order() const;
    static constexpr int kMcOrderedMinSweeps = 4;   // calls with fewer sweeps stay in storage order
    static constexpr int kMcBlock = 256;   // slots per constrained Monte Carlo workgroup
    std::vector<int> mc_blk_start;         // launch l's workgroups: mc_blk[mc_blk_start[l] ...]
    DevBuf mc_atom, mc_pat, mc_blk, mc_sigma, mc_rej, mc_tot_moves, mc_tot_rej;
    // Metropolis in colour order (stencil rows only): each replica's atoms colour by colour,
    // storage order within a colour.  mco_perm[p]: storage index of colour-order atom p; the
    // stencils, material ids and caller indices (Philox counters) of the colour order; the sweep's
    // workgroups {first, replica, end, 0} per launch; spins copied in for a call's sweeps.
    bool mco_ok = false;
    int mco_npat = 0, mco_full_rows = 0;
    std::vector<int> mco_blk_start;
    DevBuf mco_perm, mco_lidx, mco_type, mco_pat, mco_pat_off, mco_pat_jid, mco_pat_deg, mco_blk, mco_spin;
    std::vector<uint16_t> idx16_;   // host per-edge J ids of the stencil layout (JM_IDX16)

    // graphs: (nsteps) -> executable
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    long long graph_key = -1;     // nsteps, block size, XCD mapping and buffer roles at capture

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int timed_steps = 0;
    PinnedRing ring;

    ~vmx_ctx() {
        for (auto* m : mags) delete m;
        for (auto* e : energies) delete e;
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        fit.drop_graph();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }

    size_t rsize() const { return precision == VMX_FP64 ? sizeof(double) : sizeof(float); }

    int use_device() const {
        if (device >= 0) VMX_HIP(hipSetDevice(device));
        return VMX_OK;
    }

    template <typename R>
    Spin3<R> spins(int idx) const {
        R* base = spin[idx].as<R>();
        return Spin3<R>{base, base + n, base + 2 * (size_t)n};
    }
    // Euler-derivative arrays of the fused step, or null arrays when the corrector recomputes dS
    template <typename R>
    Spin3<R> euler() const {
        if (!store_euler || sizeof(R) == sizeof(float)) return Spin3<R>{nullptr, nullptr, nullptr};
        return arr3<R>(euler_buf);
    }
    // kept external-field arrays, or null arrays when the corrector regenerates the field
    template <typename R>
    Spin3<R> field() const {
        if (!store_field) return Spin3<R>{nullptr, nullptr, nullptr};
        return arr3<R>(field_buf);
    }
    template <typename R>
    Spin3<R> arr3(const DevBuf& b) const {
        R* base = b.as<R>();
        return Spin3<R>{base, base + n, base + 2 * (size_t)n};
    }

    template <typename R>
    SysDev<R> sys() const {
        SysDev<R> s;
        s.n = n; s.nmat = nmat; s.nblocks = nblocks; s.xcd_remap = xcd_remap;
        s.zmax = zmax;
        s.full_rows = full_rows;
        s.npat = npat;
        // the fused kernels stage the stencil offsets in LDS when they take at most 16 KiB and
        // the whole staging area stays within 64 KiB
        s.pat_lds = (tp == TP_PAT && (size_t)npat * zmax * sizeof(int) <= 16384 &&
                     fused_lds_bytes<R>(nmat, npat, zmax, 1) <= 65536) ? 1 : 0;
        s.am = am();
        s.q = q;
        s.rep_m = rep_m.as<int>();
        s.row = row.as<int>(); s.col = col.as<int>();
        s.jidx = jm == JM_UNIFORM ? nullptr : jidx.p;
        s.deg = deg.as<uint8_t>();
        s.pat = pat.as<uint16_t>();
        s.pat_off = pat_off.as<int>();
        s.pat_jid = pat_jid.as<uint16_t>();
        s.pat_deg = pat_deg.as<uint8_t>();
        s.jtab = jtab.as<R>();
        s.type = type.as<uint8_t>();
        s.mat = mat.as<MatConst<R>>();
        s.cond = cond.as<Cond<R>>();
        s.blk = blk.as<int4>();
        s.rep_off = rep_off.as<int>();
        s.rep_key = rep_key.as<uint2>();
        s.step_base = step_base.as<unsigned long long>();
        s.hdip = hdip_on ? hdip.as<R>() : nullptr;
        s.hd = dip.on ? dip.hd_atom.as<double>() : nullptr;
        s.ani = ani; s.fld = fld | (demag_on() ? FLD_DEMAG : 0u);
        s.dt = (R)dt; s.half_dt = (R)half_dt;
        return s;
    }

    long long graph_key_of(int nsteps) const {
        return (((long long)nsteps * 2048 + block) * 2 + xcd_remap) * 16 + s_idx * 4 + o_idx;
    }

    Launch launch() const {
        int g = nblocks;
        if (xcd_remap) g = (g + 7) / 8 * 8;
        return Launch{g, block, stream};
    }
    // the fused kernels take FusedAtoms<R> atoms per thread from their own block table
    int fused_atoms() const {
        return precision == VMX_FP32 ? FusedAtoms<float>::value : FusedAtoms<double>::value;
    }
    Launch launch_fused() const {
        int g = nblocks_fused;
        if (xcd_remap) g = (g + 7) / 8 * 8;
        return Launch{g, block, stream};
    }
    template <typename R>
    SysDev<R> sys_fused() const {
        SysDev<R> s = sys<R>();
        s.blk = blk_fused.as<int4>();
        s.nblocks = nblocks_fused;
        return s;
    }

    // Logical workgroups {first atom, replica, end atom, replica start}.  With a basis-major
    // order (q > 1) each basis block of a replica is cut into chunks of `span` cells and the
    // table runs chunk-major (chunk 0 of every basis, then chunk 1, ...): a lattice cell's q
    // atoms, which are each other's neighbours, land on the same XCD's stretch of the table and
    // share its L2 instead of being fetched into q different L2s.
    static std::vector<int4> block_table(int nrep, const std::vector<int>& off, int q, int span) {
        std::vector<int4> tab;
        for (int r = 0; r < nrep; ++r) {
            const int r0 = off[r], r1 = off[r + 1];
            if (q > 1) {
                const int m = (r1 - r0) / q;
                for (int c0 = 0; c0 < m; c0 += span)
                    for (int b = 0; b < q; ++b) {
                        const int first = r0 + b * m + c0;
                        tab.push_back(make_int4(first, r, first + std::min(span, m - c0), r0));
                    }
            } else {
                for (int a = r0; a < r1; a += span) tab.push_back(make_int4(a, r, std::min(a + span, r1), r0));
            }
        }
        if (tab.empty()) tab.push_back(make_int4(0, 0, 0, 0));
        return tab;
    }
    int build_blocks() {
        const std::vector<int4> tab = block_table(nrep, rep_off_h, q, block);
        const std::vector<int4> tabf = block_table(nrep, rep_off_h, q, block * fused_atoms());
        nblocks = (int)tab.size();
        nblocks_fused = (int)tabf.size();
        // a captured graph holds the tables' addresses: never replay one across a rebuild
        drop_graph();
        VMX_TRY(blk.alloc(tab.size() * sizeof(int4)));
        VMX_HIP(hipMemcpy(blk.p, tab.data(), tab.size() * sizeof(int4), hipMemcpyHostToDevice));
        VMX_TRY(blk_fused.alloc(tabf.size() * sizeof(int4)));
        VMX_HIP(hipMemcpy(blk_fused.p, tabf.data(), tabf.size() * sizeof(int4), hipMemcpyHostToDevice));
        // the table lists each replica's entries together, replica after replica
        std::vector<int> first(nrep + 1, (int)tab.size());
        for (int e = (int)tab.size() - 1; e >= 0; --e)
            if (tab[e].z > tab[e].x) first[tab[e].y] = e;
        for (int r = nrep - 1; r >= 0; --r) first[r] = std::min(first[r], first[r + 1]);
        VMX_TRY(upload(rep_blk, first));
        return VMX_OK;
    }

    template <typename R>
    int upload_materials() {
        std::vector<MatConst<R>> mc(nmat);
        for (int m = 0; m < nmat; ++m) {
            const vmx_material& a = mats[m];
            MatConst<R>& c = mc[m];
            c.one_opa2 = (R)a.one_oneplusalpha_sq;
            c.alpha_opa2 = (R)a.alpha_oneplusalpha_sq;
            c.ku2 = (R)a.ku2;
            for (int k = 0; k < 3; ++k) {
                c.ku[k] = (R)a.ku_vector[k];
                c.e1[k] = (R)a.kc_vector1[k];
                c.e2[k] = (R)a.kc_vector2[k];
                c.e3[k] = (R)a.kc_vector3[k];
            }
            c.kc4 = (R)a.kc4;
            c.mu_s_SI = (R)a.mu_s_SI;
            c.alpha = (R)a.alpha;
        }
        VMX_TRY(mat.alloc(mc.size() * sizeof(MatConst<R>)));
        VMX_HIP(hipMemcpy(mat.p, mc.data(), mc.size() * sizeof(MatConst<R>), hipMemcpyHostToDevice));
        return VMX_OK;
    }

    // Conditions of every (replica, material); temperature rescaling and thermal width as in
    // src/simulate/fields.cpp:258-272, applied field as in fields.cpp:181-210, Metropolis
    // factor as in src/montecarlo/mc.cpp:45-53.
    template <typename R>
    int upload_conditions() {
        std::vector<Cond<R>> cv((size_t)nrep * nmat);
        for (int r = 0; r < nrep; ++r) {
            const double T = T_h[r];
            const double Hx = H_h[3 * r + 0], Hy = H_h[3 * r + 1], Hz = H_h[3 * r + 2];
            for (int m = 0; m < nmat; ++m) {
                const vmx_material& a = mats[m];
                const double alpha = a.temperature_rescaling_alpha;
                const double Tc = a.temperature_rescaling_Tc;
                const double rescaled_temperature = T < Tc ? Tc * pow(T / Tc, alpha) : T;
                const double sqrt_T = sqrt(rescaled_temperature);
                Cond<R>& c = cv[(size_t)r * nmat + m];
                c.sigma = (R)(sqrt_T * a.H_th_sigma);
                double h[3] = {Hx, Hy, Hz};
                for (int k = 0; k < 3; ++k) {
                    double hl = h[k];
                    if (fld & FLD_LOCAL) hl = h[k] + a.applied_field_strength * a.applied_field_unit_vector[k];
                    c.happ[k] = (R)hl;
                    c.hdemag[k] = (R)Hd_h[3 * (size_t)r + k];
                    c.hglob[k] = (R)h[k];
                }
                const double kbt = 9.27400915e-24 / (rescaled_temperature * 1.3806503e-23);
                c.kBTBohr = (R)kbt;
                c.cmc_delta = (R)(rescaled_temperature < 1.0 ? 0.02 : pow(1.0 / kbt, 0.2) * 0.08);
            }
        }
        VMX_TRY(cond.alloc(cv.size() * sizeof(Cond<R>)));
        const size_t b = cv.size() * sizeof(Cond<R>);
        if (b % 4 == 0 && b <= sizeof(ArgWords)) {
            ArgWords v;
            std::memcpy(v.w, cv.data(), b);
            hipLaunchKernelGGL(k_put_words, dim3(1), dim3(kArgWords), 0, stream, cond.as<uint32_t>(), v,
                               (int)(b / 4));
            VMX_HIP(hipGetLastError());
            return VMX_OK;
        }
        if (ring.bytes < b) return fail(VMX_ERR_STATE, "condition table exceeds staging");
        return ring.copy(cond.p, cv.data(), b, stream);
    }
    int upload_conditions_any() {
        return precision == VMX_FP64 ? upload_conditions<double>() : upload_conditions<float>();
    }

    // true when an FP32 step without caller noise takes its normals from the FP64 generator
    template <typename R>
    bool normals_from_fp64(const double* noise) const {
        return fp64_normals && !noise && sizeof(R) == sizeof(float) && (fld & FLD_THERMAL);
    }
    // the FP64 kernels' normals of Philox step `st` into noise_dev ([3][n], internal order)
    int draw_fp64_normals(uint64_t st) {
        VMX_HIP(Ops<double>::thermal_noise(launch(), sys<double>(), st, noise_dev.as<double>()));
        return VMX_OK;
    }

    template <typename R>
    int heun_steps(int nsteps, const double* noise, double* kernel_ms = nullptr);
    template <typename R>
    int midpoint_steps(int nsteps, const double* noise);
    template <typename R>
    int capture_graph(int nsteps);
    template <typename R>
    int stats_update(double* fixed = nullptr);
    int stats_grid(int nbins) const;
    size_t stats_partial_bytes() const;
    double* stats_scratch = nullptr;   // this device's copy of `scratch` (no-allocation update)
    template <typename R>
    int mc_sweeps(int nsweeps);
    int colour();

    // deterministic fitting trajectory and its adjoint
    struct Fit {
        bool ready = false;
        int nsteps = 0, record_every = 1, K = 100, nrec = 0;
        double J = 0, k = 0, alpha = 0, gamma_rel = 1, H[3] = {0, 0, 0};
        bool has_h = false;
        DevBuf ckpt, seg, P, lam, lH1, lPd, lT, H0, sumS, lH0, lSp, dh, nhat, h, acc, mz;
        // backward: dL/dm_z injections per step (inj[q], zero between records) and the slice of
        // the segment being replayed (inj_seg[K]); the graph of one full segment, captured by
        // each backward call (its kernels take that call's buffers and parameters) and replayed
        // for every full segment
        DevBuf inj, inj_seg;
        hipGraph_t graph = nullptr;
        hipGraphExec_t graph_exec = nullptr;
        void drop_graph() {
            if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
            if (graph) (void)hipGraphDestroy(graph);
            graph_exec = nullptr;
            graph = nullptr;
        }
    } fit;
    bool symmetric_checked = false, symmetric = false;
    template <typename R>
    FitPrm<R> fit_prm() const;
    template <typename R>
    int fit_forward(const double* h, const double* spins0, double* mz_out);
    template <typename R>
    int fit_backward(const double* dmz, double* grad, double* dh);
};

// --------------------------------------------------------------------------------------------
template <typename R>
int vmx_ctx::heun_steps(int nsteps, const double* noise, double* kernel_ms) {
    const SysDev<R> s = sys<R>(), sf = sys_fused<R>();
    const Launch L = launch(), Lf = launch_fused();
    uint32_t off0 = 0;
    VMX_TRY(eager_step_off(nsteps, off0));
    const double* nd = nullptr;
    const bool normals64 = normals_from_fp64<R>(noise);
    if (noise || normals64) {
        VMX_TRY(noise_dev.alloc(3 * (size_t)n * sizeof(double)));
        nd = noise_dev.as<double>();
    }
    if (kernel == VMX_KERNEL_PLAIN) {
        VMX_TRY(plainF.alloc(3 * (size_t)n * rsize()));
        VMX_TRY(plainE.alloc(3 * (size_t)n * rsize()));
        VMX_TRY(plainD.alloc(3 * (size_t)n * rsize()));
    }
    // per-kernel timing: one event before each kernel and one after the last
    hipEvent_t kev[6] = {};
    if (kernel_ms) {
        for (int e = 0; e < 6; ++e) VMX_HIP(hipEventCreate(&kev[e]));
        for (int e = 0; e < 5; ++e) kernel_ms[e] = 0.0;
    }
    auto mark = [&](int e) -> hipError_t { return kernel_ms ? hipEventRecord(kev[e], stream) : hipSuccess; };
    // a call of one step (gpu::llg_heun's) is not timed: the two events would cost it more than
    // its own launches
    const bool timed = nsteps > 1 || kernel_ms;
    if (timed) VMX_HIP(hipEventRecord(ev0, stream));
    for (int k = 0; k < nsteps; ++k) {
        if (noise) {
            // the previous step still reads noise_dev: wait, then copy this step's block in
            noise_host.resize(3 * (size_t)n);
            to_internal(noise + (size_t)k * 3 * n, noise_host.data(), 3);
            VMX_HIP(hipStreamSynchronize(stream));
            VMX_HIP(hipMemcpy(noise_dev.p, noise_host.data(), 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
        } else if (normals64) {
            VMX_TRY(draw_fp64_normals(step + (uint64_t)k));
        }
        const int p_idx = 3 - s_idx - o_idx;
        int nk = 0;
        if (kernel == VMX_KERNEL_FUSED) {
            VMX_HIP(mark(0));
            VMX_HIP(Ops<R>::fused_predictor(Lf, et, am(), sf, spins<R>(s_idx), spins<R>(p_idx), euler<R>(),
                                            field<R>(), nd, off0 + (uint32_t)k));
            VMX_HIP(mark(1));
            const int out = corrector_in_place() ? s_idx : o_idx;
            VMX_HIP(Ops<R>::fused_corrector(L, et, am(), s, spins<R>(s_idx), spins<R>(p_idx), spins<R>(out),
                                            euler<R>(), field<R>(), nd, off0 + (uint32_t)k));
            VMX_HIP(mark(2));
            nk = 2;
            if (out == o_idx) std::swap(s_idx, o_idx);
        } else {
            const Spin3<R> F = arr3<R>(plainF), E = arr3<R>(plainE), D = arr3<R>(plainD);
            VMX_HIP(mark(0));
            VMX_HIP(Ops<R>::spin_fields(L, et, am(),s, spins<R>(s_idx), F));
            VMX_HIP(mark(1));
            VMX_HIP(Ops<R>::ext_fields(L, s, E, nd, off0 + (uint32_t)k));
            VMX_HIP(mark(2));
            VMX_HIP(Ops<R>::plain_predictor(L, s, spins<R>(s_idx), F, E, D, spins<R>(p_idx)));
            VMX_HIP(mark(3));
            VMX_HIP(Ops<R>::spin_fields(L, et, am(),s, spins<R>(p_idx), F));
            VMX_HIP(mark(4));
            VMX_HIP(Ops<R>::plain_corrector(L, s, spins<R>(s_idx), spins<R>(p_idx), F, E, D));
            VMX_HIP(mark(5));
            nk = 5;
        }
        // sim::increment_time then dipole::calculate_field: refresh when the step count since
        // set-up reaches a multiple of the update rate (src/dipole/field.cpp:52-58)
        if (dip.on && (dip.time + k + 1) % dip.rate == 0) VMX_TRY(dipole_refresh<R>());
        if (kernel_ms) {
            VMX_HIP(hipEventSynchronize(kev[nk]));
            for (int e = 0; e < nk; ++e) {
                float t = 0.f;
                VMX_HIP(hipEventElapsedTime(&t, kev[e], kev[e + 1]));
                kernel_ms[e] += (double)t / nsteps;
            }
        }
    }
    if (kernel_ms)
        for (int e = 0; e < 6; ++e) (void)hipEventDestroy(kev[e]);
    if (timed) VMX_HIP(hipEventRecord(ev1, stream));
    timed_steps = timed ? nsteps : 0;
    step += (uint64_t)nsteps;   // *step_base stays: the next eager call's offset grows instead
    if (dip.on) dip.time += nsteps;
    return VMX_OK;
}

// sim::LLG_Midpoint (src/simulate/LLGMidpoint.cpp:87-186) with the fused workgroup organisation:
// predictor into the spare buffer, corrector into the output buffer, roles swapped per step.
template <typename R>
int vmx_ctx::midpoint_steps(int nsteps, const double* noise) {
    const SysDev<R> s = sys<R>(), sf = sys_fused<R>();
    const Launch L = launch(), Lf = launch_fused();
    uint32_t off0 = 0;
    VMX_TRY(eager_step_off(nsteps, off0));
    const double* nd = nullptr;
    const bool normals64 = normals_from_fp64<R>(noise);
    if (noise || normals64) {
        VMX_TRY(noise_dev.alloc(3 * (size_t)n * sizeof(double)));
        nd = noise_dev.as<double>();
    }
    const bool timed = nsteps > 1;   // as the Heun step: a one-step call is not timed
    if (timed) VMX_HIP(hipEventRecord(ev0, stream));
    for (int k = 0; k < nsteps; ++k) {
        if (noise) {
            noise_host.resize(3 * (size_t)n);
            to_internal(noise + (size_t)k * 3 * n, noise_host.data(), 3);
            VMX_HIP(hipStreamSynchronize(stream));
            VMX_HIP(hipMemcpy(noise_dev.p, noise_host.data(), 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
        } else if (normals64) {
            VMX_TRY(draw_fp64_normals(step + (uint64_t)k));
        }
        const int p_idx = 3 - s_idx - o_idx;
        VMX_HIP(Ops<R>::mid_predictor(Lf, et, am(), sf, spins<R>(s_idx), spins<R>(p_idx), field<R>(), nd,
                                      off0 + (uint32_t)k));
        // the midpoint corrector gathers only the midpoint: S_next overwrites S
        VMX_HIP(Ops<R>::mid_corrector(L, et, am(), s, spins<R>(s_idx), spins<R>(p_idx), spins<R>(s_idx), field<R>(),
                                      nd, off0 + (uint32_t)k));
        if (dip.on && (dip.time + k + 1) % dip.rate == 0) VMX_TRY(dipole_refresh<R>());
    }
    if (timed) VMX_HIP(hipEventRecord(ev1, stream));
    timed_steps = timed ? nsteps : 0;
    step += (uint64_t)nsteps;   // *step_base stays: the next eager call's offset grows instead
    if (dip.on) dip.time += nsteps;
    return VMX_OK;
}

template <typename R>
int vmx_ctx::capture_graph(int nsteps) {
    if (graph_exec) { VMX_HIP(hipGraphExecDestroy(graph_exec)); graph_exec = nullptr; }
    if (graph) { VMX_HIP(hipGraphDestroy(graph)); graph = nullptr; }
    const SysDev<R> s = sys<R>(), sf = sys_fused<R>();
    const Launch L = launch(), Lf = launch_fused();
    VMX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int si = s_idx, oi = o_idx;
    for (int k = 0; k < nsteps; ++k) {
        const int pi = 3 - si - oi;
        hipError_t e = Ops<R>::fused_predictor(Lf, et, am(), sf, spins<R>(si), spins<R>(pi), euler<R>(), field<R>(),
                                               nullptr, (uint32_t)k);
        const int out = corrector_in_place() ? si : oi;
        if (e == hipSuccess)
            e = Ops<R>::fused_corrector(L, et, am(), s, spins<R>(si), spins<R>(pi), spins<R>(out), euler<R>(),
                                        field<R>(), nullptr, (uint32_t)k);
        if (e != hipSuccess) {
            hipGraph_t g;
            (void)hipStreamEndCapture(stream, &g);
            return fail(VMX_ERR_HIP, "kernel launch during graph capture failed: %s", hipGetErrorString(e));
        }
        if (out == oi) std::swap(si, oi);
    }
    hipLaunchKernelGGL(k_advance_step, dim3(1), dim3(1), 0, stream, step_base.as<unsigned long long>(),
                       (unsigned)nsteps);
    VMX_HIP(hipStreamEndCapture(stream, &graph));
    VMX_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    graph_key = graph_key_of(nsteps);
    return VMX_OK;
}

// Workgroups of the partial-sum kernels.  Each thread walks its atoms with a grid stride, one
// dependent memory round trip per atom, so masks with few bins get eight workgroups per CU
// (2048); masks with many bins keep 512, which bounds the partials k_*_finish reduces.
int vmx_ctx::stats_grid(int nbins) const {
    return std::max(1, std::min((n + 255) / 256, nbins <= 16 ? 2048 : 512));
}
// Bytes of workgroup partials the largest registered statistic needs.
size_t vmx_ctx::stats_partial_bytes() const {
    size_t b = 0;
    for (const MagStat* st : mags) b = std::max(b, (size_t)stats_grid(st->nbins) * 4 * st->nbins * sizeof(double));
    for (const EnergyStat* st : energies)
        b = std::max(b, (size_t)stats_grid(st->nbins) * 4 * st->nbins * sizeof(double));
    return b;
}

// fixed: device memory of at least stats_partial_bytes() for the workgroup partials; null: the
// context's own buffer, grown as needed.
template <typename R>
int vmx_ctx::stats_update(double* fixed) {
    const Spin3<R> S = spins<R>(s_idx);
    double* part = fixed;
    if (!part) {
        VMX_TRY(partial.alloc(stats_partial_bytes()));
        part = partial.as<double>();
    }
    for (MagStat* st : mags) {
        const int grid = stats_grid(st->nbins);
        VMX_HIP(Ops<R>::mag_partial(grid, stream, n, S, st->mask.as<int>(), moments(), st->nbins, part));
        hipLaunchKernelGGL(k_mag_finish, dim3(1), dim3(256), 0, stream, grid, part, st->nbins,
                           st->cur.as<double>(), st->mean.as<double>(), st->empty_dev.as<int>(),
                           (int)st->empty.size(), st->counter.as<long long>());
        VMX_HIP(hipGetLastError());
    }
    if (!energies.empty()) {
        const SysDev<R> s = sys<R>();
        for (EnergyStat* st : energies) {
            const int grid = stats_grid(st->nbins);
            VMX_HIP(Ops<R>::energy_partial(grid, stream, et, am(), s, S, st->mask.as<int>(), moments(),
                                           nrep > 1 ? atom_rep.as<int>() : nullptr, st->nbins, part));
            hipLaunchKernelGGL(k_energy_finish, dim3(1), dim3(256), 0, stream, grid, part, st->nbins,
                               st->cur.as<double>(), st->mean.as<double>(), st->empty_dev.as<int>(),
                               (int)st->empty.size(), st->counter.as<long long>());
            VMX_HIP(hipGetLastError());
        }
    }
    return VMX_OK;
}

// Greedy graph colouring in atom order over the symmetrised neighbour graph, so that no two
// atoms of one colour interact in either direction.
int vmx_ctx::colour() {
    const std::vector<int>& rowh = rowh_;
    const std::vector<int>& colh = colh_;
    // reverse adjacency
    std::vector<int> rrow(n + 1, 0), rcol((size_t)nnz);
    for (int64_t e = 0; e < nnz; ++e) rrow[colh[e] + 1]++;
    for (int i = 0; i < n; ++i) rrow[i + 1] += rrow[i];
    {
        std::vector<int> fill(rrow.begin(), rrow.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int e = rowh[i]; e < rowh[i + 1]; ++e) rcol[fill[colh[e]]++] = i;
    }
    // atoms are visited in the caller's order (o), so the colouring does not depend on the
    // internal storage order; c[] and the lists hold internal indices
    std::vector<int> c(n, -1);
    std::vector<int> seen;  // seen[colour] == o marks colours used around the o-th atom
    int nc = 0;
    for (int o = 0; o < n; ++o) {
        const int i = q > 1 ? pi[o] : o;
        auto mark = [&](int j) {
            if (j != i && c[j] >= 0) {
                if ((int)seen.size() <= c[j]) seen.resize(c[j] + 1, -1);
                seen[c[j]] = o;
            }
        };
        for (int e = rowh[i]; e < rowh[i + 1]; ++e) mark(colh[e]);
        for (int e = rrow[i]; e < rrow[i + 1]; ++e) mark(rcol[e]);
        int k = 0;
        while (k < (int)seen.size() && seen[k] == o) ++k;
        c[i] = k;
        nc = std::max(nc, k + 1);
    }
    // Each colour lists its atoms replica by replica in storage order (moves of one colour commute,
    // so the order is free: storage order keeps a wavefront's atoms, and their neighbours, close
    // in memory), and each replica's run is cut into workgroups of 256 entries.
    std::vector<int> list;
    list.reserve(n);
    colour_start.assign(nc + 1, 0);
    std::vector<int> run_first((size_t)nc * nrep), run_last((size_t)nc * nrep);
    for (int k = 0; k < nc; ++k) {
        colour_start[k] = (int)list.size();
        for (int r = 0; r < nrep; ++r) {
            run_first[(size_t)k * nrep + r] = (int)list.size();
            for (int i = rep_off_h[r]; i < rep_off_h[r + 1]; ++i)
                if (c[i] == k) list.push_back(i);
            run_last[(size_t)k * nrep + r] = (int)list.size();
        }
    }
    colour_start[nc] = (int)list.size();
    // A sweep is nc launches.  Replicas do not interact, so only each replica's own colour order
    // matters: replica r with k_r < nc colours may put its colours in any increasing set of
    // launches.  Replicas with nc colours take launch k for colour k; the others, largest first,
    // take the increasing set that keeps the largest launch smallest (then the loads most even),
    // so that no launch is left with only the few atoms of the last colours of odd lattices.
    std::vector<int> lau((size_t)nc * nrep);
    {
        std::vector<double> load(nc, 0.0);
        std::vector<int> kr(nrep, 0);
        for (int r = 0; r < nrep; ++r)
            for (int k = 0; k < nc; ++k)
                if (run_last[(size_t)k * nrep + r] > run_first[(size_t)k * nrep + r]) kr[r] = k + 1;
        auto size = [&](int k, int r) {
            return (double)(run_last[(size_t)k * nrep + r] - run_first[(size_t)k * nrep + r]);
        };
        auto sets = [&](int k) {   // number of increasing launch sets of size k
            double b = 1.0;
            for (int j = 0; j < k; ++j) b = b * (nc - j) / (j + 1);
            return b;
        };
        std::vector<int> order;
        for (int r = 0; r < nrep; ++r) {
            if (kr[r] == nc || sets(kr[r]) > 4096.0) {
                for (int k = 0; k < nc; ++k) { lau[(size_t)k * nrep + r] = k; load[k] += size(k, r); }
            } else {
                order.push_back(r);
            }
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return rep_off_h[a + 1] - rep_off_h[a] > rep_off_h[b + 1] - rep_off_h[b];
        });
        std::vector<int> pick(nc), best(nc);
        for (int r : order) {
            const int k = kr[r];
            double best_max = 1e300, best_sq = 1e300;
            std::vector<double> trial(nc);
            // increasing launch sets of size k, first in lexicographic order
            for (int j = 0; j < k; ++j) pick[j] = j;
            for (;;) {
                double mx = 0.0, sq = 0.0;
                trial = load;
                for (int j = 0; j < k; ++j) trial[pick[j]] += size(j, r);
                for (double v : trial) { mx = std::max(mx, v); sq += v * v; }
                if (mx < best_max || (mx == best_max && sq < best_sq)) {
                    best_max = mx; best_sq = sq;
                    std::copy(pick.begin(), pick.begin() + k, best.begin());
                }
                int j = k - 1;
                while (j >= 0 && pick[j] == nc - k + j) --j;
                if (j < 0) break;
                ++pick[j];
                for (int q2 = j + 1; q2 < k; ++q2) pick[q2] = pick[q2 - 1] + 1;
            }
            for (int j = 0; j < k; ++j) { lau[(size_t)j * nrep + r] = best[j]; load[best[j]] += size(j, r); }
        }
    }
    // A launch holds at most one colour of each replica; its workgroups go replica by replica, so
    // the costly short runs (the last colours of odd lattices, scattered atoms with few cache
    // hits) are spread over the launch and over the XCDs' contiguous ranges of workgroups.
    // Workgroup w of the sweep owns entries [w * mcb, (w + 1) * mcb) of two flat arrays: the atom
    // (~first atom of the run past the run's end) and, with stencil rows, its stencil id; mc_blk[w]
    // = {replica, its first atom}.  A thread's first loads depend on nothing but its position.
    const int mcb = 256 * (precision == VMX_FP64 ? McAtoms<double>::value : McAtoms<float>::value);
    std::vector<uint16_t> pat_h;
    if (tp == TP_PAT) {
        pat_h.resize(n);
        VMX_HIP(hipMemcpy(pat_h.data(), pat.p, (size_t)n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    }
    std::vector<int> atom_flat;
    std::vector<uint16_t> pat_flat;
    std::vector<int2> tab;
    mc_blk_start.assign(nc + 1, 0);
    for (int l = 0; l < nc; ++l) {
        mc_blk_start[l] = (int)tab.size();
        for (int r = 0; r < nrep; ++r)
            for (int k = 0; k < nc; ++k) {
                if (lau[(size_t)k * nrep + r] != l) continue;
                const int first = run_first[(size_t)k * nrep + r], last = run_last[(size_t)k * nrep + r];
                for (int a = first; a < last; a += mcb) {
                    tab.push_back(make_int2(r, rep_off_h[r]));
                    for (int e = a; e < a + mcb; ++e) {
                        const int i = e < last ? list[e] : list[a];
                        atom_flat.push_back(e < last ? i : ~i);
                        if (tp == TP_PAT) pat_flat.push_back(pat_h[i]);
                    }
                }
            }
    }
    mc_blk_start[nc] = (int)tab.size();
    VMX_TRY(upload(mc_atom, atom_flat));
    if (tp == TP_PAT) VMX_TRY(upload(mc_pat, pat_flat));
    VMX_TRY(upload(mc_blk, tab));
    if (tp == TP_PAT) VMX_TRY(colour_order(c, lau, nc));
    // constrained Monte Carlo: the same colours, each replica's run in caller order, workgroups of
    // 256 slots per run (m/2 slots for a run of even length m, m for odd m)
    std::vector<int> clist, crange((size_t)nc * (nrep + 1));
    std::vector<int>& cmlen = cmc_mlen_h;
    cmlen.assign((size_t)nc * nrep, 0);
    std::vector<int4> ctab;
    clist.reserve(n);
    cmc_tab_start.assign(nc + 1, 0);
    for (int k = 0; k < nc; ++k) {
        cmc_tab_start[k] = (int)ctab.size();
        for (int r = 0; r < nrep; ++r) {
            crange[(size_t)k * (nrep + 1) + r] = (int)ctab.size() - cmc_tab_start[k];
            const int first = (int)clist.size();
            for (int o = rep_off_h[r]; o < rep_off_h[r + 1]; ++o) {
                const int i = q > 1 ? pi[o] : o;
                if (c[i] == k) clist.push_back(i);
            }
            const int m = (int)clist.size() - first;
            cmlen[(size_t)k * nrep + r] = m;
            const int slots = m < 2 ? 0 : (m % 2 == 0 ? m / 2 : m);
            for (int t0 = 0; t0 < slots; t0 += kMcBlock) ctab.push_back(make_int4(first, r, m, t0));
        }
        crange[(size_t)k * (nrep + 1) + nrep] = (int)ctab.size() - cmc_tab_start[k];
    }
    cmc_tab_start[nc] = (int)ctab.size();
    VMX_TRY(upload(cmc_list, clist));
    if (mco_ok) {
        std::vector<int> clist_co(clist.size());
        for (size_t e = 0; e < clist.size(); ++e) clist_co[e] = mco_invp_h[clist[e]];
        VMX_TRY(upload(cmc_list_co, clist_co));
    }
    VMX_TRY(upload(cmc_tab, ctab));
    VMX_TRY(upload(cmc_range, crange));
    VMX_TRY(upload(cmc_mlen, cmlen));
    VMX_TRY(cmc_perm.alloc(2 * (size_t)nc * nrep * sizeof(int4)));
    VMX_TRY(cmc_dm.alloc(3 * std::max<size_t>(ctab.size(), 1) * sizeof(double)));
    ncolours = nc;
    coloured = true;
    return VMX_OK;
}

namespace {
bool build_stencils(int n, int zmax, const std::vector<int>& rowh, const std::vector<int>& colh,
                    const std::vector<uint16_t>& jid, std::vector<uint16_t>& pat, std::vector<int>& off,
                    std::vector<uint16_t>& pjid, std::vector<uint8_t>& pdeg, int max_stencils);
}

// The colour order of the Metropolis sweep (stencil rows): the permutation, its stencils (rows
// keep their neighbour order, so every energy sums in the same order), material ids, caller
// indices and per-launch workgroups.  A table that does not fit the stencil limit leaves the
// storage-order sweep in use.
int vmx_ctx::colour_order(const std::vector<int>& c, const std::vector<int>& lau, int nc) {
    mco_ok = false;
    std::vector<int> perm(n), invp(n), first((size_t)nc * nrep), last((size_t)nc * nrep);
    int p = 0;
    for (int r = 0; r < nrep; ++r)
        for (int k = 0; k < nc; ++k) {
            first[(size_t)k * nrep + r] = p;
            for (int i = rep_off_h[r]; i < rep_off_h[r + 1]; ++i)
                if (c[i] == k) { perm[p] = i; invp[i] = p; ++p; }
            last[(size_t)k * nrep + r] = p;
        }
    std::vector<int> rowc(n + 1), colc((size_t)nnz);
    std::vector<uint16_t> jidc(jm == JM_IDX16 ? (size_t)nnz : 0);
    rowc[0] = 0;
    for (int q2 = 0; q2 < n; ++q2) {
        const int i = perm[q2], b = rowh_[i], d = rowh_[i + 1] - b;
        rowc[q2 + 1] = rowc[q2] + d;
        for (int k = 0; k < d; ++k) {
            colc[rowc[q2] + k] = invp[colh_[b + k]];
            if (jm == JM_IDX16) jidc[rowc[q2] + k] = idx16_[b + k];
        }
    }
    std::vector<uint16_t> patc, pjid;
    std::vector<int> off;
    std::vector<uint8_t> pdeg;
    // the colour order of lattices with an odd period has irregular rows near the seams: up to the
    // full 16-bit range of stencil ids (the table stays in L2: 24-48 bytes per stencil)
    if (!build_stencils(n, zmax, rowc, colc, jidc, patc, off, pjid, pdeg, 65536)) return VMX_OK;
    // material ids in colour order; caller index within the replica (the inverse of the
    // basis-major storage order, as caller_local computes it on the device)
    std::vector<uint8_t> type_s(n), type_c(n);
    VMX_HIP(hipMemcpy(type_s.data(), type.p, (size_t)n, hipMemcpyDeviceToHost));
    std::vector<int> lidx(n);
    for (int r = 0; r < nrep; ++r) {
        const int m = (rep_off_h[r + 1] - rep_off_h[r]) / std::max(q, 1);
        for (int q2 = rep_off_h[r]; q2 < rep_off_h[r + 1]; ++q2) {
            const int i = perm[q2], loc = i - rep_off_h[r];
            lidx[q2] = q <= 1 ? loc : (loc % m) * q + loc / m;
            type_c[q2] = type_s[i];
        }
    }
    const int mcb = 256 * (precision == VMX_FP64 ? McAtoms<double>::value : McAtoms<float>::value);
    std::vector<int4> tab;
    mco_blk_start.assign(nc + 1, 0);
    for (int l = 0; l < nc; ++l) {
        mco_blk_start[l] = (int)tab.size();
        for (int r = 0; r < nrep; ++r)
            for (int k = 0; k < nc; ++k) {
                if (lau[(size_t)k * nrep + r] != l) continue;
                const int a0 = first[(size_t)k * nrep + r], a1 = last[(size_t)k * nrep + r];
                for (int a = a0; a < a1; a += mcb) tab.push_back(make_int4(a, r, std::min(a + mcb, a1), 0));
            }
    }
    mco_blk_start[nc] = (int)tab.size();
    VMX_TRY(upload(mco_perm, perm));
    VMX_TRY(upload(mco_invp, invp));
    mco_invp_h = invp;
    {
        std::vector<double> mm_s(n), mm_c(n);
        VMX_HIP(hipMemcpy(mm_s.data(), mm.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
        for (int q2 = 0; q2 < n; ++q2) mm_c[q2] = mm_s[perm[q2]];
        VMX_TRY(upload(mco_mm, mm_c));
    }
    VMX_TRY(upload(mco_lidx, lidx));
    VMX_TRY(upload(mco_type, type_c));
    VMX_TRY(upload(mco_pat, patc));
    VMX_TRY(upload(mco_pat_off, off));
    VMX_TRY(upload(mco_pat_jid, pjid));
    VMX_TRY(upload(mco_pat_deg, pdeg));
    VMX_TRY(upload(mco_blk, tab));
    VMX_TRY(mco_spin.alloc(3 * (size_t)n * rsize()));
    mco_npat = (int)pdeg.size();
    mco_full_rows = std::all_of(pdeg.begin(), pdeg.end(), [this](uint8_t d) { return d == zmax; }) ? 1 : 0;
    mco_ok = true;
    return VMX_OK;
}

// The system description of the colour order: the storage order's with the colour order's
// stencils and material ids.
template <typename R>
SysDev<R> vmx_ctx::sys_colour_order() const {
    SysDev<R> s = sys<R>();
    s.npat = mco_npat;
    s.full_rows = mco_full_rows;
    s.pat_lds = ((size_t)mco_npat * zmax * sizeof(int) <= 16384 && fused_lds_bytes<R>(nmat, mco_npat, zmax, 1) <= 65536)
                    ? 1 : 0;
    s.pat = mco_pat.as<uint16_t>();
    s.pat_off = mco_pat_off.as<int>();
    s.pat_jid = mco_pat_jid.as<uint16_t>();
    s.pat_deg = mco_pat_deg.as<uint8_t>();
    s.type = mco_type.as<uint8_t>();
    return s;
}

template <typename R>
int vmx_ctx::mc_sweeps(int nsweeps) {
    if (!coloured) VMX_TRY(colour());
    const SysDev<R> s = sys<R>();
    const Spin3<R> S = spins<R>(s_idx);
    // Calls of several sweeps without the dipole field run in colour order: the spins go there
    // once and come back at the end.  Both orders give the same results bit for bit (the same
    // moves, counters and arithmetic; only the memory layout differs).
    const bool ordered = mco_ok && !dip.on && nsweeps >= kMcOrderedMinSweeps;
    const SysDev<R> so = ordered ? sys_colour_order<R>() : s;
    Spin3<R> C{};
    if (ordered) {
        R* base = mco_spin.as<R>();
        C = Spin3<R>{base, base + n, base + 2 * (size_t)n};
        VMX_HIP(Ops<R>::mc_permute(stream, n, mco_perm.as<int>(), S, C, true));
    }
    for (int w = 0; w < nsweeps; ++w) {
        for (int k = 0; k < ncolours; ++k) {   // launch k (each replica's colours in order)
            if (ordered) {
                VMX_HIP(Ops<R>::mc_colour_ordered(stream, et, am(), so, C, mco_lidx.as<int>(), mco_blk.as<int4>(),
                                                  mco_blk_start[k], mco_blk_start[k + 1] - mco_blk_start[k],
                                                  mc_sigma.as<double>(), sweep, mc_rej.as<double>()));
                continue;
            }
            VMX_HIP(Ops<R>::mc_colour(stream, et, am(), s, S, mc_atom.as<int>(), mc_pat.as<uint16_t>(),
                                      mc_blk.as<int2>(), mc_blk_start[k], mc_blk_start[k + 1] - mc_blk_start[k],
                                      mc_sigma.as<double>(), sweep,
                                      mc_rej.as<double>()));
        }
        hipLaunchKernelGGL(k_mc_adapt, dim3((nrep + 3) / 4), dim3(256), 0, stream, nrep, rep_off.as<int>(),
                           mc_sigma.as<double>(), mc_rej.as<double>(), mc_tot_moves.as<double>(),
                           mc_tot_rej.as<double>());
        VMX_HIP(hipGetLastError());
        ++sweep;
        // a sweep advances sim::time like a Heun step (sim.cpp:608-615, 191-200)
        if (dip.on && ++dip.time % dip.rate == 0) VMX_TRY(dipole_refresh<R>());
    }
    if (ordered) VMX_HIP(Ops<R>::mc_permute(stream, n, mco_perm.as<int>(), C, S, false));
    return VMX_OK;
}

// The permutations of every (colour, round, replica) of sweep sw into cmc_perm: for a table of up
// to 1 KiB (a few replicas) formed here with the device's own integer arithmetic and written by a
// one-workgroup launch whose arguments carry it; else by k_cmc_perm, one thread per entry.  The
// device kernel is a chain of dependent divisions (≈ 10 µs), longer than a small system's round.
int vmx_ctx::cmc_permutations(uint64_t sw) {
    const size_t count = 2 * (size_t)ncolours * nrep;
    if (count * sizeof(int4) > sizeof(ArgWords)) {
        VMX_HIP(Ops<float>::cmc_perm(stream, rep_key.as<uint2>(), nrep, ncolours, cmc_mlen.as<int>(), sw,
                                     cmc_perm.as<int4>()));
        return VMX_OK;
    }
    ArgWords v;
    int4* p = reinterpret_cast<int4*>(v.w);
    for (size_t g = 0; g < count; ++g) {
        const int r = (int)(g % nrep), kr = (int)(g / nrep), k = kr >> 1, round = kr & 1;
        const int m = cmc_mlen_h[(size_t)k * nrep + r];
        if (m < 2) { p[g] = make_int4(1, 0, 1, 0); continue; }
        uint64_t a, b;
        cmc_permutation(rep_key_h[r], k, round, sw, m, a, b);
        p[g] = make_int4((int)a, (int)b, (int)mod_inverse((uint32_t)a, (uint32_t)m), 0);
    }
    hipLaunchKernelGGL(k_put_words, dim3(1), dim3(kArgWords), 0, stream, cmc_perm.as<uint32_t>(), v,
                       (int)(count * sizeof(int4) / 4));
    VMX_HIP(hipGetLastError());
    return VMX_OK;
}

// Constrained Monte Carlo sweeps (montecarlo::cmc_step, src/montecarlo/cmc.cpp:282-497): per
// colour two rounds.  Every replica's magnetisation is summed over all spins at the start of each
// sweep (as cmc_step does, cmc.cpp:346-351) and brought forward after each round by the sum of
// the round's accepted changes, so no round but a sweep's first reads every spin for it.
template <typename R>
int vmx_ctx::cmc_sweeps(int nsweeps) {
    if (!coloured) VMX_TRY(colour());
    if (cmc_frame_dirty) {
        std::vector<CmcFrame<R>> f(nrep);
        for (int r = 0; r < nrep; ++r)
            for (int k = 0; k < 9; ++k) {
                f[r].P[k] = (R)cmc_frame_h[21 * (size_t)r + k];
                f[r].Pt[k] = (R)cmc_frame_h[21 * (size_t)r + 9 + k];
                if (k < 3) f[r].v[k] = (R)cmc_frame_h[21 * (size_t)r + 18 + k];
            }
        VMX_TRY(upload(cmc_frame, f));
        cmc_frame_dirty = false;
    }
    VMX_TRY(cmc_M.alloc(3 * (size_t)nrep * sizeof(double)));
    VMX_TRY(cmc_part.alloc(3 * (size_t)nblocks * sizeof(double)));
    const SysDev<R> s = sys<R>();
    const Spin3<R> S = spins<R>(s_idx);
    int last = -1;                                 // the sweep's last colour with pairs
    for (int k = 0; k < ncolours; ++k)
        if (cmc_tab_start[k + 1] > cmc_tab_start[k]) last = k;
    // Calls of several sweeps run on the colour-ordered copy of the spins, as Metropolis does: each
    // pair's atoms come from the colour-ordered list, the moments and tables follow, the Philox
    // counters use the caller's indices, and the magnetisation is summed in storage order through
    // the inverse permutation, so the results are those of storage order bit for bit.
    const bool ordered = mco_ok && !dip.on && nsweeps >= kMcOrderedMinSweeps && last >= 0;
    const SysDev<R> so = ordered ? sys_colour_order<R>() : s;
    Spin3<R> C = S;
    if (ordered) {
        R* base = mco_spin.as<R>();
        C = Spin3<R>{base, base + n, base + 2 * (size_t)n};
        VMX_HIP(Ops<R>::mc_permute(stream, n, mco_perm.as<int>(), S, C, true));
    }
    const int* plist = ordered ? cmc_list_co.as<int>() : cmc_list.as<int>();
    const int* plidx = ordered ? mco_lidx.as<int>() : nullptr;
    const double* pmm = ordered ? mco_mm.as<double>() : mm.as<double>();
    const int* gather = ordered ? mco_invp.as<int>() : nullptr;
    for (int w = 0; w < nsweeps; ++w) {
        if (last < 0) { ++sweep; continue; }
        VMX_TRY(cmc_permutations(sweep));
        VMX_HIP(Ops<R>::replica_moment(stream, s, C, moments(), gather, rep_blk.as<int>(), nrep,
                                       cmc_part.as<double>(), cmc_M.as<double>()));
        for (int k = 0; k <= last; ++k) {
            const int ntab = cmc_tab_start[k + 1] - cmc_tab_start[k];
            if (ntab == 0) continue;
            double* dm = cmc_dm.as<double>() + 3 * (size_t)cmc_tab_start[k];
            for (int round = 0; round < 2; ++round) {
                VMX_HIP(Ops<R>::cmc_pairs(stream, et, am(), so, C, pmm, plist, plidx,
                                          cmc_tab.as<int4>() + cmc_tab_start[k], ntab,
                                          cmc_perm.as<int4>() + (size_t)(2 * k + round) * nrep,
                                          cmc_frame.as<CmcFrame<R>>(), cmc_M.as<double>(), mc_sigma.as<double>(),
                                          cmc_move, round, sweep, cmc_counts.as<double>(), dm));
                if (k < last || round == 0)
                    VMX_HIP(Ops<R>::moment_add(stream, cmc_range.as<int>() + (size_t)k * (nrep + 1), nrep, dm,
                                               cmc_M.as<double>()));
            }
        }
        ++sweep;   // a constrained sweep advances sim::time like a Heun step (sim.cpp:623-628)
    }
    if (ordered) VMX_HIP(Ops<R>::mc_permute(stream, n, mco_perm.as<int>(), C, S, false));
    return VMX_OK;
}

// ============================================================================================
// ---- fitting trajectory and adjoint -------------------------------------------------------------
template <typename R>
FitPrm<R> vmx_ctx::fit_prm() const {
    FitPrm<R> p;
    const double a = -fit.gamma_rel / (1.0 + fit.alpha * fit.alpha);
    p.J = (R)fit.J;
    p.k = (R)fit.k;
    p.a = (R)a;
    p.b = (R)(fit.alpha * a);
    p.dt = (R)dt;
    p.half_dt = (R)half_dt;
    for (int q = 0; q < 3; ++q) {
        p.e[q] = (R)mats[0].ku_vector[q];
        p.H[q] = (R)fit.H[q];
    }
    return p;
}

template <typename R>
int vmx_ctx::fit_forward(const double* h, const double* spins0, double* mz_out) {
    const size_t N3 = 3 * (size_t)n;
    const int nck = fit.nsteps / fit.K + 1;
    VMX_TRY(fit.ckpt.alloc((size_t)nck * N3 * sizeof(R)));
    VMX_TRY(fit.seg.alloc((size_t)(fit.K + 1) * N3 * sizeof(R)));
    VMX_TRY(fit.P.alloc(N3 * sizeof(R)));
    VMX_TRY(fit.h.alloc(N3 * sizeof(R)));
    VMX_TRY(fit.mz.alloc((size_t)4 * std::max(fit.nrec, 1) * sizeof(double)));
    double* st = staging.as<double>();
    VMX_HIP(hipStreamSynchronize(stream));
    fit.has_h = h != nullptr;
    if (h) {
        std::vector<double> hi(N3);
        to_internal(h, hi.data(), 3);
        VMX_HIP(hipMemcpy(st, hi.data(), N3 * sizeof(double), hipMemcpyHostToDevice));
        VMX_HIP(Ops<R>::from_double(stream, (int)N3, st, fit.h.as<R>()));
        VMX_HIP(hipStreamSynchronize(stream));
    }
    {
        std::vector<double> si(N3);
        to_internal(spins0, si.data(), 3);
        VMX_HIP(hipMemcpy(st, si.data(), N3 * sizeof(double), hipMemcpyHostToDevice));
    }
    R* ck = fit.ckpt.as<R>();
    VMX_HIP(Ops<R>::from_double(stream, (int)N3, st, ck));
    VMX_HIP(hipMemsetAsync(fit.mz.p, 0, (size_t)4 * std::max(fit.nrec, 1) * sizeof(double), stream));
    const SysDev<R> s = sys<R>();
    const FitPrm<R> p = fit_prm<R>();
    const Spin3<R> hh = fit.has_h ? arr3<R>(fit.h) : Spin3<R>{nullptr, nullptr, nullptr};
    auto state = [&](R* base) { return Spin3<R>{base, base + n, base + 2 * (size_t)n}; };
    R* seg = fit.seg.as<R>();
    R* cur = seg;
    R* nxt = seg + N3;
    VMX_HIP(hipMemcpyAsync(cur, ck, N3 * sizeof(R), hipMemcpyDeviceToDevice, stream));
    for (int q = 0; q < fit.nsteps; ++q) {
        VMX_HIP(Ops<R>::fit_predictor(stream, am(), s, p, hh, state(cur), arr3<R>(fit.P)));
        VMX_HIP(Ops<R>::fit_corrector(stream, am(), s, p, hh, state(cur), arr3<R>(fit.P), state(nxt)));
        std::swap(cur, nxt);
        if ((q + 1) % fit.K == 0)
            VMX_HIP(hipMemcpyAsync(ck + (size_t)((q + 1) / fit.K) * N3, cur, N3 * sizeof(R), hipMemcpyDeviceToDevice,
                                   stream));
        if ((q + 1) % fit.record_every == 0)
            VMX_HIP(Ops<R>::fit_mz(stream, n, state(cur), fit.mz.as<double>() + 4 * (size_t)((q + 1) / fit.record_every - 1)));
    }
    std::vector<double> mzh((size_t)4 * std::max(fit.nrec, 1));
    VMX_HIP(hipStreamSynchronize(stream));
    VMX_HIP(hipMemcpy(mzh.data(), fit.mz.p, mzh.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < fit.nrec; ++r) mz_out[r] = mzh[4 * (size_t)r] / n;
    fit.ready = true;
    return VMX_OK;
}

template <typename R>
int vmx_ctx::fit_backward(const double* dmz, double* grad, double* dh_out) {
    const size_t N3 = 3 * (size_t)n;
    for (DevBuf* b : {&fit.lam, &fit.lH1, &fit.lPd, &fit.lT, &fit.H0, &fit.sumS, &fit.lH0, &fit.lSp, &fit.dh})
        VMX_TRY(b->alloc(N3 * sizeof(R)));
    VMX_TRY(fit.nhat.alloc((size_t)n * sizeof(R)));
    const int nwg = (n + 255) / 256;   // workgroups of the backward kernels: one slot of 4 sums each
    VMX_TRY(fit.acc.alloc((size_t)4 * nwg * sizeof(double)));
    const SysDev<R> s = sys<R>();
    const FitPrm<R> p = fit_prm<R>();
    const Spin3<R> hh = fit.has_h ? arr3<R>(fit.h) : Spin3<R>{nullptr, nullptr, nullptr};
    auto state = [&](R* base) { return Spin3<R>{base, base + n, base + 2 * (size_t)n}; };
    const int fill_grid = (int)std::min<size_t>((N3 + 255) / 256, 4096);
    hipLaunchKernelGGL((k_fill<R>), dim3(fill_grid), dim3(256), 0, stream, N3, fit.lam.as<R>(), (R)0);
    hipLaunchKernelGGL((k_fill<R>), dim3(fill_grid), dim3(256), 0, stream, N3, fit.dh.as<R>(), (R)0);
    hipLaunchKernelGGL((k_fill<double>), dim3(std::min(nwg, 1024)), dim3(64), 0, stream, (size_t)4 * nwg,
                       fit.acc.as<double>(), 0.0);
    VMX_HIP(hipGetLastError());
    R* ck = fit.ckpt.as<R>();
    R* seg = fit.seg.as<R>();
    const Spin3<R> lam = arr3<R>(fit.lam);
    const int nseg = (fit.nsteps + fit.K - 1) / fit.K;
    // dL/dm_z enters lambda_z after step q when q + 1 is a recorded step (zero elsewhere)
    {
        std::vector<R> inj((size_t)fit.nsteps, R(0));
        for (int q = 0; q < fit.nsteps; ++q)
            if ((q + 1) % fit.record_every == 0) inj[q] = (R)(dmz[(q + 1) / fit.record_every - 1] / n);
        VMX_TRY(upload(fit.inj, inj));
        VMX_TRY(fit.inj_seg.alloc((size_t)fit.K * sizeof(R)));
    }
    // One segment: recompute its states from the checkpoint in seg[0], then step back over them.
    // inj: the segment's injections (inj[t] after local step t).  The same sequence for every
    // full segment, so it is captured once and replayed.
    auto segment = [&](int len, const R* inj) -> hipError_t {
        for (int t = 0; t < len; ++t) {
            R* a = seg + (size_t)t * N3;
            hipError_t e = Ops<R>::fit_predictor(stream, am(), s, p, hh, state(a), arr3<R>(fit.P));
            if (e == hipSuccess) e = Ops<R>::fit_corrector(stream, am(), s, p, hh, state(a), arr3<R>(fit.P),
                                                          state(a + N3));
            if (e != hipSuccess) return e;
        }
        for (int t = len - 1; t >= 0; --t) {
            hipLaunchKernelGGL((k_lam_inject<R>), dim3((n + 255) / 256), dim3(256), 0, stream, n, lam.z, inj + t);
            hipError_t e = hipGetLastError();
            R* a = seg + (size_t)t * N3;
            if (e == hipSuccess) e = Ops<R>::fit_predictor(stream, am(), s, p, hh, state(a), arr3<R>(fit.P));
            if (e == hipSuccess)
                e = Ops<R>::adj_b1(stream, am(), s, p, hh, state(a), arr3<R>(fit.P), lam, arr3<R>(fit.lH1),
                                   arr3<R>(fit.lPd), arr3<R>(fit.lT), arr3<R>(fit.H0), arr3<R>(fit.sumS),
                                   fit.nhat.as<R>(), fit.acc.as<double>());
            if (e == hipSuccess)
                e = Ops<R>::adj_b2(stream, am(), s, p, state(a), arr3<R>(fit.P), arr3<R>(fit.lH1), arr3<R>(fit.lPd),
                                   arr3<R>(fit.lT), arr3<R>(fit.H0), arr3<R>(fit.sumS), fit.nhat.as<R>(),
                                   arr3<R>(fit.lH0), arr3<R>(fit.lSp), arr3<R>(fit.dh), fit.acc.as<double>());
            if (e == hipSuccess) e = Ops<R>::adj_b3(stream, am(), s, p, arr3<R>(fit.lH0), arr3<R>(fit.lSp), lam, (R)0);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    // full segments replay one graph when there are at least two of them
    const int nfull = fit.nsteps / fit.K;
    const bool use_graph = nfull >= 2;
    if (use_graph) {
        fit.drop_graph();
        VMX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const hipError_t e = segment(fit.K, fit.inj_seg.as<R>());
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(stream, &g);
        if (e != hipSuccess || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(VMX_ERR_HIP, "adjoint segment capture failed: %s",
                        hipGetErrorString(e != hipSuccess ? e : ec));
        }
        fit.graph = g;
        VMX_HIP(hipGraphInstantiate(&fit.graph_exec, fit.graph, nullptr, nullptr, 0));
    }
    for (int c = nseg - 1; c >= 0; --c) {
        const int start = c * fit.K, end = std::min(start + fit.K, fit.nsteps);
        VMX_HIP(hipMemcpyAsync(seg, ck + (size_t)c * N3, N3 * sizeof(R), hipMemcpyDeviceToDevice, stream));
        if (use_graph && end - start == fit.K) {
            VMX_HIP(hipMemcpyAsync(fit.inj_seg.p, fit.inj.as<R>() + start, (size_t)fit.K * sizeof(R),
                                   hipMemcpyDeviceToDevice, stream));
            VMX_HIP(hipGraphLaunch(fit.graph_exec, stream));
        } else {
            VMX_HIP(segment(end - start, fit.inj.as<R>() + start));
        }
    }
    VMX_HIP(hipStreamSynchronize(stream));
    fit.drop_graph();   // its replays have finished
    std::vector<double> part((size_t)4 * nwg);
    VMX_HIP(hipMemcpy(part.data(), fit.acc.p, part.size() * sizeof(double), hipMemcpyDeviceToHost));
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int w = 0; w < nwg; ++w)
        for (int k = 0; k < 4; ++k) acc[k] += part[4 * (size_t)w + k];
    // a = -gamma/(1+alpha^2), b = alpha a
    const double al = fit.alpha, g = fit.gamma_rel;
    const double a = -g / (1.0 + al * al);
    const double da_dal = 2.0 * al * g / ((1.0 + al * al) * (1.0 + al * al));
    const double db_dal = a + al * da_dal;
    grad[0] = acc[0];
    grad[1] = acc[1];
    grad[2] = acc[2] * da_dal + acc[3] * db_dal;
    if (dh_out) {
        double* st = staging.as<double>();
        VMX_HIP(Ops<R>::to_double(stream, (int)N3, fit.dh.as<R>(), st));
        VMX_HIP(hipStreamSynchronize(stream));
        std::vector<double> tmp(N3);
        VMX_HIP(hipMemcpy(tmp.data(), st, N3 * sizeof(double), hipMemcpyDeviceToHost));
        to_caller(tmp.data(), dh_out, 3);
    }
    return VMX_OK;
}

// Exchange table compression: exact (bitwise) deduplication of the J values edges refer to.
namespace {

struct KeyHash {
    size_t operator()(const std::string& k) const { return std::hash<std::string>()(k); }
};

int compress_exchange(const vmx_system* sys, const std::vector<int>& iid_csr, int K, int& jm,
                      std::vector<double>& table, std::vector<uint16_t>& idx16, std::vector<int>& idx32) {
    const int64_t nnz = (int64_t)iid_csr.size();
    const double* ex = sys->exchange;
    // fast path: every referenced entry equal to the first one
    bool uniform = true;
    if (nnz > 0) {
        const double* j0 = ex + (size_t)K * iid_csr[0];
        for (int64_t e = 1; e < nnz && uniform; ++e) {
            const double* je = ex + (size_t)K * iid_csr[e];
            if (std::memcmp(j0, je, K * sizeof(double)) != 0) uniform = false;
        }
        if (uniform) {
            table.assign(j0, j0 + K);
            jm = JM_UNIFORM;
            return VMX_OK;
        }
    } else {
        table.assign(K, 0.0);
        jm = JM_UNIFORM;
        return VMX_OK;
    }
    // dedup the interaction ids that are referenced
    std::unordered_map<std::string, int, KeyHash> uniq;
    std::vector<int> remap((size_t)sys->num_interactions, -1);
    bool small = true;
    for (int64_t e = 0; e < nnz; ++e) {
        const int id = iid_csr[e];
        if (remap[id] >= 0) continue;
        std::string key((const char*)(ex + (size_t)K * id), K * sizeof(double));
        auto it = uniq.find(key);
        if (it == uniq.end()) {
            const int u = (int)uniq.size();
            uniq.emplace(std::move(key), u);
            table.insert(table.end(), ex + (size_t)K * id, ex + (size_t)K * id + K);
            remap[id] = u;
            if (u >= 65536) { small = false; break; }
        } else {
            remap[id] = it->second;
        }
    }
    if (small) {
        jm = JM_IDX16;
        idx16.resize(nnz);
        for (int64_t e = 0; e < nnz; ++e) idx16[e] = (uint16_t)remap[iid_csr[e]];
        return VMX_OK;
    }
    jm = JM_IDX32;
    table.assign(ex, ex + (size_t)K * sys->num_interactions);
    idx32 = iid_csr;
    return VMX_OK;
}

// Stencils: a row is (degree, j-i for each slot, J id for each slot).  Rows of a periodic lattice
// repeat a few of them (per basis atom, times the boundary-wrap cases), so a row can be coded as
// a 16-bit stencil id.  Returns false when the rows do not fit `max_pat` stencils.
constexpr int kMaxStencils = 4096;

bool build_stencils(int n, int zmax, const std::vector<int>& rowh, const std::vector<int>& colh,
                    const std::vector<uint16_t>& jid, std::vector<uint16_t>& pat, std::vector<int>& off,
                    std::vector<uint16_t>& pjid, std::vector<uint8_t>& pdeg, int max_stencils) {
    std::unordered_map<std::string, int> ids;
    pat.resize(n);
    std::string key;
    std::vector<int> row_off(zmax);
    std::vector<uint16_t> row_jid(zmax);
    for (int i = 0; i < n; ++i) {
        const int b = rowh[i], d = rowh[i + 1] - rowh[i];
        key.assign(1, (char)d);
        for (int k = 0; k < d; ++k) {
            row_off[k] = colh[b + k] - i;
            row_jid[k] = jid.empty() ? 0 : jid[b + k];
        }
        key.append((const char*)row_off.data(), d * sizeof(int));
        key.append((const char*)row_jid.data(), d * sizeof(uint16_t));
        auto it = ids.find(key);
        int p;
        if (it == ids.end()) {
            p = (int)ids.size();
            if (p >= max_stencils) return false;
            ids.emplace(key, p);
            off.resize((size_t)(p + 1) * zmax, 0);
            pjid.resize((size_t)(p + 1) * zmax, 0);
            pdeg.push_back((uint8_t)d);
            for (int k = 0; k < d; ++k) {
                off[(size_t)p * zmax + k] = row_off[k];
                pjid[(size_t)p * zmax + k] = row_jid[k];
            }
        } else {
            p = it->second;
        }
        pat[i] = (uint16_t)p;
    }
    return true;
}

}  // namespace

// Lattice period of the atom order.  VAMPIRE numbers atoms cell by cell with the unit-cell
// atoms innermost (src/create/cs_create_crystal_structure2.cpp:100-107), so the rows of atoms
// i and i + q (q = atoms per cell) carry the same j - i offsets away from the boundaries.
// Stored as is, a wavefront's 64 consecutive rows mix q different stencils and each neighbour
// slot reads q strided spans of memory; stored basis-major (all basis-0 atoms, then basis-1,
// ...) the 64 rows of a wavefront read one contiguous span per slot.  Returns the q in 2..8
// with the largest share of equal (i, i+q) rows when that share is at least 20% and larger than
// for q = 1; else 1 (keep the order).  q must divide every replica's size.
static int detect_period(int n, const std::vector<int>& rowh, const std::vector<int>& colh,
                         const std::vector<int>& roff) {
    const int stride = std::max(1, n / 200000);   // sample at most ~200k rows
    auto share = [&](int q) {
        int64_t same = 0, tot = 0;
        for (int i = 0; i + q < n; i += stride) {
            ++tot;
            const int a = rowh[i], da = rowh[i + 1] - a, b = rowh[i + q], db = rowh[i + q + 1] - b;
            if (da != db || da == 0) continue;
            bool eq = true;
            for (int k = 0; k < da && eq; ++k) eq = (colh[a + k] - i) == (colh[b + k] - (i + q));
            same += eq;
        }
        return tot ? (double)same / (double)tot : 0.0;
    };
    const double s1 = share(1);
    int best = 1;
    double best_share = s1;
    for (int q = 2; q <= 8; ++q) {
        bool divides = true;
        for (size_t r = 0; r + 1 < roff.size(); ++r) divides &= ((roff[r + 1] - roff[r]) % q) == 0;
        if (!divides) continue;
        const double sq = share(q);
        if (sq > best_share + 1e-12) { best = q; best_share = sq; }
    }
    return (best > 1 && best_share >= 0.2) ? best : 1;
}

// Chooses and uploads the device layout of the neighbour list (VMX_TOPOLOGY_*).
static int build_topology(vmx_ctx* c, int request, const std::vector<int>& rowh, const std::vector<int>& colh,
                          const std::vector<uint16_t>& idx16, const std::vector<int>& idx32) {
    const int n = c->n;
    int zmax = 0;
    for (int i = 0; i < n; ++i) zmax = std::max(zmax, rowh[i + 1] - rowh[i]);
    if (request < VMX_TOPOLOGY_AUTO || request > VMX_TOPOLOGY_STENCIL)
        return fail(VMX_ERR_ARG, "unknown topology %d", request);
    // slot width: the narrowest compiled width that holds the widest row
    int wsel = -1;
    for (int w = 0; w < 4; ++w)
        if (zmax <= kSlotWidths[w]) { wsel = w; break; }
    const bool slotted = wsel >= 0;
    const int W = slotted ? kSlotWidths[wsel] : 0;
    c->zmax = zmax;
    if ((request == VMX_TOPOLOGY_AUTO || request == VMX_TOPOLOGY_STENCIL) && c->jm != JM_IDX32 && slotted) {
        std::vector<uint16_t> pat, pjid;
        std::vector<int> off;
        std::vector<uint8_t> pdeg;
        if (build_stencils(n, W, rowh, colh, c->jm == JM_IDX16 ? idx16 : std::vector<uint16_t>(), pat, off, pjid,
                           pdeg, kMaxStencils)) {
            c->tp = TP_PAT;
            c->wsel = wsel;
            c->zmax = W;
            c->npat = (int)pdeg.size();
            c->full_rows = std::all_of(pdeg.begin(), pdeg.end(), [W](uint8_t d) { return d == W; }) ? 1 : 0;
            VMX_TRY(upload(c->pat, pat));
            VMX_TRY(upload(c->pat_off, off));
            VMX_TRY(upload(c->pat_jid, pjid));
            VMX_TRY(upload(c->pat_deg, pdeg));
            return VMX_OK;
        }
        if (request == VMX_TOPOLOGY_STENCIL)
            return fail(VMX_ERR_ARG, "rows do not fit %d stencils", kMaxStencils);
    } else if (request == VMX_TOPOLOGY_STENCIL) {
        return fail(VMX_ERR_ARG, "stencil layout needs at most 65536 distinct J values and rows of <= %d",
                    kSlotWidths[3]);
    }
    // ELL when padding every row to W slots costs at most half again the CSR entries
    const int64_t nnz = c->nnz;
    const bool ell_ok = slotted && (int64_t)W * n <= nnz + nnz / 2 + n;
    if ((request == VMX_TOPOLOGY_AUTO && ell_ok) || request == VMX_TOPOLOGY_ELL) {
        if (!slotted) return fail(VMX_ERR_ARG, "ELL layout needs rows of at most %d neighbours", kSlotWidths[3]);
        const size_t slots = (size_t)W * n;
        std::vector<int> col(slots, 0);
        std::vector<uint8_t> deg(n);
        std::vector<uint16_t> j16(c->jm == JM_IDX16 ? slots : 0, 0);
        std::vector<int> j32(c->jm == JM_IDX32 ? slots : 0, 0);
        for (int i = 0; i < n; ++i) {
            const int b = rowh[i], d = rowh[i + 1] - rowh[i];
            deg[i] = (uint8_t)d;
            for (int k = 0; k < W; ++k) {
                const size_t e = (size_t)k * n + i;
                col[e] = k < d ? colh[b + k] : i;   // padding: the atom itself (in bounds, cached)
                if (c->jm == JM_IDX16 && k < d) j16[e] = idx16[b + k];
                if (c->jm == JM_IDX32 && k < d) j32[e] = idx32[b + k];
            }
        }
        c->tp = TP_ELL;
        c->wsel = wsel;
        c->zmax = W;
        c->full_rows = std::all_of(deg.begin(), deg.end(), [W](uint8_t d) { return d == W; }) ? 1 : 0;
        VMX_TRY(upload(c->col, col));
        VMX_TRY(upload(c->deg, deg));
        if (c->jm == JM_IDX16) VMX_TRY(upload(c->jidx, j16));
        if (c->jm == JM_IDX32) VMX_TRY(upload(c->jidx, j32));
        return VMX_OK;
    }
    c->tp = TP_CSR;
    c->wsel = 0;
    VMX_TRY(upload(c->row, rowh));
    VMX_TRY(upload(c->col, colh));
    if (c->jm == JM_IDX16) VMX_TRY(upload(c->jidx, idx16));
    if (c->jm == JM_IDX32) VMX_TRY(upload(c->jidx, idx32));
    return VMX_OK;
}

// ============================================================================================
extern "C" {

const char* vmx_last_error(void) { return g_last_error.c_str(); }

const char* vmx_version(void) { return "vmx 0.1 gfx950"; }

int vmx_device_count(int32_t* count) {
    if (!count) return fail(VMX_ERR_ARG, "count is null");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
    return VMX_OK;
}

int vmx_create(const vmx_system* sys, const vmx_options* opt, vmx_ctx** out) {
    g_last_error.clear();
    if (!sys || !opt || !out) return fail(VMX_ERR_ARG, "null argument");
    *out = nullptr;
    const int n = sys->num_atoms;
    if (n <= 0) return fail(VMX_ERR_ARG, "num_atoms must be positive (got %d)", n);
    // spin gathers use 32-bit byte offsets into each component array
    const int64_t max_atoms = (opt->precision == VMX_FP32 ? (int64_t)1 << 30 : (int64_t)1 << 29) - 1;
    if (n > max_atoms)
        return fail(VMX_ERR_ARG, "num_atoms %d exceeds %lld for this precision (shard the system)", n,
                    (long long)max_atoms);
    if (sys->num_materials < 1 || sys->num_materials > 255)
        return fail(VMX_ERR_ARG, "num_materials must be in [1,255] (got %d)", sys->num_materials);
    if (!sys->type || !sys->nbr_start || !sys->nbr_end || !sys->materials || !sys->moment)
        return fail(VMX_ERR_ARG, "type, moment, nbr_start, nbr_end and materials are required");
    if (sys->num_nbr > 0 && (!sys->nbr || !sys->nbr_iid)) return fail(VMX_ERR_ARG, "nbr and nbr_iid are required");
    if (sys->exchange_type < 0 || sys->exchange_type > 2)
        return fail(VMX_ERR_ARG, "exchange_type must be 0, 1 or 2 (got %d)", sys->exchange_type);
    if (sys->num_nbr > 0 && (sys->num_interactions <= 0 || !sys->exchange))
        return fail(VMX_ERR_ARG, "exchange table is required");
    if (opt->precision != VMX_FP64 && opt->precision != VMX_FP32) return fail(VMX_ERR_ARG, "bad precision");
    if (opt->kernel != VMX_KERNEL_PLAIN && opt->kernel != VMX_KERNEL_FUSED) return fail(VMX_ERR_ARG, "bad kernel");
    const int bs = opt->block_size ? opt->block_size : 256;
    if (bs < 64 || bs > 512 || bs % 64) return fail(VMX_ERR_ARG, "block_size must be a multiple of 64 in [64,512]");
    const int nrep = sys->num_replicas < 1 ? 1 : sys->num_replicas;
    std::vector<int> roff(nrep + 1);
    if (nrep == 1) {
        roff[0] = 0; roff[1] = n;
        if (sys->replica_offset && (sys->replica_offset[0] != 0 || sys->replica_offset[1] != n))
            return fail(VMX_ERR_ARG, "replica_offset must be {0, num_atoms} for one replica");
    } else {
        if (!sys->replica_offset) return fail(VMX_ERR_ARG, "replica_offset is required for several replicas");
        for (int r = 0; r <= nrep; ++r) roff[r] = sys->replica_offset[r];
        if (roff[0] != 0 || roff[nrep] != n) return fail(VMX_ERR_ARG, "replica_offset must run from 0 to num_atoms");
        for (int r = 0; r < nrep; ++r)
            if (roff[r + 1] <= roff[r]) return fail(VMX_ERR_ARG, "replica %d is empty or offsets decrease", r);
    }
    const int K = sys->exchange_type == 0 ? 1 : (sys->exchange_type == 1 ? 3 : 9);

    // CSR from inclusive start/end index arrays, with every index checked on the host
    std::vector<int> rowh(n + 1);
    rowh[0] = 0;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const int64_t s0 = sys->nbr_start[i], e0 = sys->nbr_end[i];
        const int64_t cnt = e0 - s0 + 1;
        if (cnt < 0 || (cnt > 0 && (s0 < 0 || e0 >= sys->num_nbr)))
            return fail(VMX_ERR_ARG, "neighbour range of atom %d [%lld,%lld] outside [0,%lld)", i, (long long)s0,
                        (long long)e0, (long long)sys->num_nbr);
        total += cnt;
        if (total > INT32_MAX) return fail(VMX_ERR_ARG, "more than 2^31-1 neighbour entries");
        rowh[i + 1] = (int)total;
    }
    std::vector<int> colh((size_t)total), iidh((size_t)total);
    std::vector<int> atom_rep_h(n);
    for (int r = 0; r < nrep; ++r)
        for (int a = roff[r]; a < roff[r + 1]; ++a) atom_rep_h[a] = r;
    for (int i = 0; i < n; ++i) {
        const int64_t s0 = sys->nbr_start[i];
        const int t = sys->type[i];
        if (t < 0 || t >= sys->num_materials) return fail(VMX_ERR_ARG, "type of atom %d is %d, outside materials", i, t);
        for (int k = rowh[i]; k < rowh[i + 1]; ++k) {
            const int j = sys->nbr[s0 + (k - rowh[i])];
            const int id = sys->nbr_iid[s0 + (k - rowh[i])];
            if (j < 0 || j >= n) return fail(VMX_ERR_ARG, "neighbour %d of atom %d is outside [0,%d)", j, i, n);
            if (id < 0 || id >= sys->num_interactions)
                return fail(VMX_ERR_ARG, "interaction id %d of atom %d is outside the exchange table", id, i);
            if (atom_rep_h[j] != atom_rep_h[i])
                return fail(VMX_ERR_ARG, "atom %d interacts with atom %d of another replica", i, j);
            colh[k] = j;
            iidh[k] = id;
        }
    }

    // Internal atom order: basis-major when the rows repeat with a lattice period q (section
    // comment at detect_period).  Row contents and their order are unchanged.
    int q = 1;
    if (opt->atom_order == VMX_ATOM_ORDER_AUTO) q = detect_period(n, rowh, colh, roff);
    else if (opt->atom_order != VMX_ATOM_ORDER_CALLER) return fail(VMX_ERR_ARG, "unknown atom_order");
    std::vector<int> pi, inv;
    std::vector<int> type_int(sys->type, sys->type + n);
    std::vector<double> mom_int(sys->moment, sys->moment + n);
    if (q > 1) {
        pi.resize(n);
        inv.resize(n);
        for (int r = 0; r < nrep; ++r) {
            const int base = roff[r], nr = roff[r + 1] - roff[r], m = nr / q;
            for (int l = 0; l < nr; ++l) {
                const int p = (l % q) * m + l / q;
                pi[base + l] = base + p;
                inv[base + p] = base + l;
            }
        }
        std::vector<int> rowp(n + 1), colp((size_t)total), iidp((size_t)total);
        rowp[0] = 0;
        for (int p = 0; p < n; ++p) {
            const int i = inv[p], d = rowh[i + 1] - rowh[i];
            rowp[p + 1] = rowp[p] + d;
            for (int k = 0; k < d; ++k) {
                colp[rowp[p] + k] = pi[colh[rowh[i] + k]];
                iidp[rowp[p] + k] = iidh[rowh[i] + k];
            }
            type_int[p] = sys->type[i];
            mom_int[p] = sys->moment[i];
        }
        rowh.swap(rowp);
        colh.swap(colp);
        iidh.swap(iidp);
    }

    std::unique_ptr<vmx_ctx> c(new vmx_ctx());
    c->type_h.assign(sys->type, sys->type + n);
    c->q = q;
    c->pi.swap(pi);
    c->inv.swap(inv);
    c->device = opt->device;
    VMX_TRY(c->use_device());
    c->precision = opt->precision;
    c->kernel = opt->kernel;
    c->block = bs;
    c->seed = opt->seed;
    c->n = n;
    c->nmat = sys->num_materials;
    c->nrep = nrep;
    c->nnz = total;
    c->et = sys->exchange_type;
    c->K = K;
    c->ani = sys->anisotropy_flags & 0x7u;
    c->mats.assign(sys->materials, sys->materials + sys->num_materials);
    c->rep_off_h = roff;
    c->T_h.assign(nrep, 0.0);
    c->H_h.assign(3 * (size_t)nrep, 0.0);
    c->Hd_h.assign(3 * (size_t)nrep, 0.0);
    VMX_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    VMX_HIP(hipEventCreate(&c->ev0));
    VMX_HIP(hipEventCreate(&c->ev1));

    std::vector<double> table;
    std::vector<uint16_t> idx16;
    std::vector<int> idx32;
    VMX_TRY(compress_exchange(sys, iidh, K, c->jm, table, idx16, idx32));

    const size_t rs = c->rsize();
    VMX_TRY(build_topology(c.get(), opt->topology, rowh, colh, idx16, idx32));
    if (c->tp == TP_PAT && c->jm == JM_IDX16) c->idx16_ = idx16;
    c->rowh_.swap(rowh);
    c->colh_.swap(colh);
    VMX_TRY(c->jtab.alloc(table.size() * rs));
    if (c->precision == VMX_FP64) {
        VMX_HIP(hipMemcpy(c->jtab.p, table.data(), table.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> tf(table.begin(), table.end());
        VMX_HIP(hipMemcpy(c->jtab.p, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    std::vector<uint8_t> t8(n);
    for (int i = 0; i < n; ++i) t8[i] = (uint8_t)type_int[i];
    VMX_TRY(c->type.alloc(n));
    VMX_HIP(hipMemcpy(c->type.p, t8.data(), n, hipMemcpyHostToDevice));
    VMX_TRY(c->mm.alloc((size_t)n * sizeof(double)));
    VMX_HIP(hipMemcpy(c->mm.p, mom_int.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    {   // every atom's moment its material's (bit for bit)?  then the statistics read the table
        std::vector<double> tab(c->nmat, 0.0);
        std::vector<char> seen(c->nmat, 0);
        bool by_type = true;
        for (int i = 0; i < n && by_type; ++i) {
            const int t = type_int[i];
            if (!seen[t]) { seen[t] = 1; tab[t] = mom_int[i]; }
            else by_type = std::memcmp(&tab[t], &mom_int[i], sizeof(double)) == 0;
        }
        if (by_type) VMX_TRY(upload(c->mm_tab, tab));
    }
    VMX_TRY(c->rep_off.alloc((nrep + 1) * sizeof(int)));
    VMX_HIP(hipMemcpy(c->rep_off.p, roff.data(), (nrep + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (nrep > 1) {   // replica of every atom (the energy statistics' applied field); replicas keep
                      // their atom ranges in the internal order
        VMX_TRY(upload(c->atom_rep, atom_rep_h));
    }
    VMX_TRY(scratch_on_current_device(&c->stats_scratch));
    {
        std::vector<int> rep_m(nrep);
        for (int r = 0; r < nrep; ++r) rep_m[r] = (roff[r + 1] - roff[r]) / q;
        VMX_TRY(upload(c->rep_m, rep_m));
    }
    std::vector<uint2>& keys = c->rep_key_h;
    keys.resize(nrep);
    for (int r = 0; r < nrep; ++r) {
        const uint64_t k = opt->seed + (uint64_t)r;
        keys[r] = make_uint2((uint32_t)k, (uint32_t)(k >> 32));
    }
    VMX_TRY(c->rep_key.alloc(nrep * sizeof(uint2)));
    VMX_HIP(hipMemcpy(c->rep_key.p, keys.data(), nrep * sizeof(uint2), hipMemcpyHostToDevice));
    VMX_TRY(c->step_base.alloc(sizeof(unsigned long long)));
    VMX_HIP(hipMemset(c->step_base.p, 0, sizeof(unsigned long long)));
    if (c->precision == VMX_FP64) VMX_TRY(c->upload_materials<double>());
    else VMX_TRY(c->upload_materials<float>());
    VMX_TRY(c->build_blocks());
    for (int b = 0; b < 3; ++b) {
        VMX_TRY(c->spin[b].alloc(3 * (size_t)n * rs));
        VMX_HIP(hipMemset(c->spin[b].p, 0, 3 * (size_t)n * rs));
    }
    VMX_TRY(c->staging.alloc(3 * (size_t)n * sizeof(double)));
    c->store_euler = opt->recompute_euler ? 0 : 1;
    c->fp64_normals = (opt->fp64_normals && c->precision == VMX_FP32) ? 1 : 0;
    c->store_field = opt->keep_external_field ? 1 : 0;
    if (c->kernel == VMX_KERNEL_FUSED && c->store_euler && c->precision == VMX_FP64)
        VMX_TRY(c->euler_buf.alloc(3 * (size_t)n * rs));
    if (c->kernel == VMX_KERNEL_FUSED && c->store_field) VMX_TRY(c->field_buf.alloc(3 * (size_t)n * rs));
    const size_t cond_bytes = (size_t)nrep * c->nmat * sizeof(Cond<double>) + 64;
    VMX_TRY(c->ring.init(std::max<size_t>(cond_bytes, 8 * (size_t)nrep + 64)));
    VMX_TRY(c->upload_conditions_any());
    // Monte Carlo state: adaptive sigma starts at 60 (src/montecarlo/data.cpp:46)
    VMX_TRY(c->mc_sigma.alloc(nrep * sizeof(double)));
    VMX_TRY(c->mc_rej.alloc((size_t)nrep * kRejectSlots * kRejectStride * sizeof(double)));
    VMX_TRY(c->mc_tot_moves.alloc(nrep * sizeof(double)));
    VMX_TRY(c->mc_tot_rej.alloc(nrep * sizeof(double)));
    VMX_TRY(c->cmc_counts.alloc((size_t)nrep * kRejectSlots * kRejectStride * sizeof(double)));
    c->cmc_frame_h.assign(21 * (size_t)nrep, 0.0);
    c->cmc_frame_set.assign(nrep, 0);
    {
        std::vector<double> s60(nrep, 60.0);
        VMX_HIP(hipMemcpy(c->mc_sigma.p, s60.data(), nrep * sizeof(double), hipMemcpyHostToDevice));
        VMX_HIP(hipMemset(c->mc_rej.p, 0, (size_t)nrep * kRejectSlots * kRejectStride * sizeof(double)));
        VMX_HIP(hipMemset(c->mc_tot_moves.p, 0, nrep * sizeof(double)));
        VMX_HIP(hipMemset(c->mc_tot_rej.p, 0, nrep * sizeof(double)));
        VMX_HIP(hipMemset(c->cmc_counts.p, 0, (size_t)nrep * kRejectSlots * kRejectStride * sizeof(double)));
    }
    VMX_HIP(hipDeviceSynchronize());
    *out = c.release();
    return VMX_OK;
}

int vmx_get_topology(vmx_ctx* c, int32_t* topology, int32_t* num_stencils) {
    if (!c || !topology) return fail(VMX_ERR_ARG, "null argument");
    *topology = c->tp == TP_PAT ? VMX_TOPOLOGY_STENCIL : (c->tp == TP_ELL ? VMX_TOPOLOGY_ELL : VMX_TOPOLOGY_CSR);
    if (num_stencils) *num_stencils = c->tp == TP_PAT ? c->npat : 0;
    return VMX_OK;
}

int vmx_get_atom_period(vmx_ctx* c, int32_t* period) {
    if (!c || !period) return fail(VMX_ERR_ARG, "null argument");
    *period = c->q;
    return VMX_OK;
}

int vmx_destroy(vmx_ctx* ctx) {
    if (!ctx) return VMX_OK;
    (void)ctx->use_device();
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return VMX_OK;
}

int vmx_set_spins(vmx_ctx* c, const double* sx, const double* sy, const double* sz) {
    if (!c || !sx || !sy || !sz) return fail(VMX_ERR_ARG, "null argument");
    VMX_TRY(c->use_device());
    const size_t nb = (size_t)c->n * sizeof(double);
    double* st = c->staging.as<double>();
    VMX_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> host(3 * (size_t)c->n), perm(3 * (size_t)c->n);
    std::memcpy(host.data(), sx, nb);
    std::memcpy(host.data() + c->n, sy, nb);
    std::memcpy(host.data() + 2 * (size_t)c->n, sz, nb);
    c->to_internal(host.data(), perm.data(), 3);
    VMX_HIP(hipMemcpy(st, perm.data(), 3 * nb, hipMemcpyHostToDevice));
    if (c->precision == VMX_FP64) {
        VMX_HIP(hipMemcpyAsync(c->spin[c->s_idx].p, st, 3 * nb, hipMemcpyDeviceToDevice, c->stream));
    } else {
        VMX_HIP(Ops<float>::from_double(c->stream, 3 * c->n, st, c->spin[c->s_idx].as<float>()));
    }
    VMX_HIP(hipStreamSynchronize(c->stream));
    return VMX_OK;
}

int vmx_get_spins(vmx_ctx* c, double* sx, double* sy, double* sz) {
    if (!c || !sx || !sy || !sz) return fail(VMX_ERR_ARG, "null argument");
    VMX_TRY(c->use_device());
    const size_t nb = (size_t)c->n * sizeof(double);
    double* st = c->staging.as<double>();
    if (c->precision == VMX_FP64) {
        VMX_HIP(hipMemcpyAsync(st, c->spin[c->s_idx].p, 3 * nb, hipMemcpyDeviceToDevice, c->stream));
    } else {
        VMX_HIP(Ops<float>::to_double(c->stream, 3 * c->n, c->spin[c->s_idx].as<float>(), st));
    }
    VMX_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> perm(3 * (size_t)c->n), host(3 * (size_t)c->n);
    VMX_HIP(hipMemcpy(perm.data(), st, 3 * nb, hipMemcpyDeviceToHost));
    c->to_caller(perm.data(), host.data(), 3);
    std::memcpy(sx, host.data(), nb);
    std::memcpy(sy, host.data() + c->n, nb);
    std::memcpy(sz, host.data() + 2 * (size_t)c->n, nb);
    return VMX_OK;
}

int vmx_set_time_step(vmx_ctx* c, double dt, double half_dt) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(VMX_ERR_ARG, "dt must be positive and finite");
    // a recorded graph holds the step's kernel arguments, dt among them
    if (dt != c->dt || half_dt != c->half_dt) c->drop_graph();
    c->dt = dt;
    c->half_dt = half_dt;
    return VMX_OK;
}

int vmx_set_conditions(vmx_ctx* c, int32_t replica, double temperature, const double H[3],
                       uint32_t field_flags) {
    if (!c || !H) return fail(VMX_ERR_ARG, "null argument");
    if (replica < -1 || replica >= c->nrep) return fail(VMX_ERR_ARG, "replica %d outside [-1,%d)", replica, c->nrep);
    if (!(temperature >= 0.0) || !std::isfinite(temperature)) return fail(VMX_ERR_ARG, "temperature must be >= 0");
    VMX_TRY(c->use_device());
    const int r0 = replica < 0 ? 0 : replica, r1 = replica < 0 ? c->nrep : replica + 1;
    // the device table already holds these conditions (every change is uploaded): nothing to do,
    // as for a program that sets the same temperature and field before every step
    bool same = (field_flags & 0x7u) == c->fld;
    for (int r = r0; r < r1 && same; ++r) {
        same = c->T_h[r] == temperature;
        for (int k = 0; k < 3; ++k) same = same && c->H_h[3 * r + k] == H[k];
    }
    if (same) return VMX_OK;
    for (int r = r0; r < r1; ++r) {
        c->T_h[r] = temperature;
        for (int k = 0; k < 3; ++k) c->H_h[3 * r + k] = H[k];
    }
    // temperatures and fields live in the device condition table that a graph reads at replay;
    // the field flags are kernel arguments recorded in the graph
    if ((field_flags & 0x7u) != c->fld) c->drop_graph();
    c->fld = field_flags & 0x7u;
    return c->upload_conditions_any();
}

int vmx_set_demag_field(vmx_ctx* c, int32_t replica, const double H[3]) {
    if (!c || !H) return fail(VMX_ERR_ARG, "null argument");
    if (replica < -1 || replica >= c->nrep) return fail(VMX_ERR_ARG, "replica %d outside [-1,%d)", replica, c->nrep);
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(H[k])) return fail(VMX_ERR_ARG, "demag field component %d is not finite", k);
    VMX_TRY(c->use_device());
    const bool was_on = c->demag_on();
    const int r0 = replica < 0 ? 0 : replica, r1 = replica < 0 ? c->nrep : replica + 1;
    for (int r = r0; r < r1; ++r)
        for (int k = 0; k < 3; ++k) c->Hd_h[3 * (size_t)r + k] = H[k];
    if (c->demag_on() != was_on) c->drop_graph();   // the term's flag is a recorded kernel argument
    return c->upload_conditions_any();
}

int vmx_llg_heun(vmx_ctx* c, int32_t nsteps, const double* noise) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (nsteps < 0) return fail(VMX_ERR_ARG, "nsteps must be >= 0");
    if (!(c->dt > 0.0)) return fail(VMX_ERR_STATE, "time step not set (vmx_set_time_step)");
    if (nsteps == 0) return VMX_OK;
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? c->heun_steps<double>(nsteps, noise) : c->heun_steps<float>(nsteps, noise);
}

int vmx_llg_midpoint(vmx_ctx* c, int32_t nsteps, const double* noise) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (nsteps < 0) return fail(VMX_ERR_ARG, "nsteps must be >= 0");
    if (!(c->dt > 0.0)) return fail(VMX_ERR_STATE, "time step not set (vmx_set_time_step)");
    if (nsteps == 0) return VMX_OK;
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? c->midpoint_steps<double>(nsteps, noise) : c->midpoint_steps<float>(nsteps, noise);
}

int vmx_llg_heun_graph(vmx_ctx* c, int32_t nsteps, int32_t nrepeat) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (nsteps <= 0 || nsteps % 2) return fail(VMX_ERR_ARG, "graph step count must be positive and even");
    if (nrepeat < 0) return fail(VMX_ERR_ARG, "nrepeat must be >= 0");
    if (c->kernel != VMX_KERNEL_FUSED) return fail(VMX_ERR_STATE, "graphs are recorded from the fused kernels");
    if (c->dip.on)
        return fail(VMX_ERR_STATE, "graph replay does not refresh the dipole field; use vmx_llg_heun");
    if (c->fp64_normals && (c->fld & FLD_THERMAL))
        return fail(VMX_ERR_STATE, "graph replay draws the normals in the step kernels; fp64_normals needs "
                                   "vmx_llg_heun");
    if (!(c->dt > 0.0)) return fail(VMX_ERR_STATE, "time step not set (vmx_set_time_step)");
    VMX_TRY(c->use_device());
    if (!c->graph_exec || c->graph_key != c->graph_key_of(nsteps)) {
        VMX_TRY(c->precision == VMX_FP64 ? c->capture_graph<double>(nsteps) : c->capture_graph<float>(nsteps));
    }
    VMX_TRY(c->sync_step_base());   // the graph's kernels read and advance *step_base
    VMX_HIP(hipEventRecord(c->ev0, c->stream));
    for (int r = 0; r < nrepeat; ++r) VMX_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    VMX_HIP(hipEventRecord(c->ev1, c->stream));
    c->timed_steps = nsteps * nrepeat;
    c->step += (uint64_t)nsteps * nrepeat;
    c->dev_step = c->step;
    return VMX_OK;
}

int vmx_get_step(vmx_ctx* c, int64_t* step) {
    if (!c || !step) return fail(VMX_ERR_ARG, "null argument");
    *step = (int64_t)c->step;
    return VMX_OK;
}

int vmx_set_step(vmx_ctx* c, int64_t step) {
    if (!c || step < 0) return fail(VMX_ERR_ARG, "bad argument");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    const unsigned long long v = (unsigned long long)step;
    VMX_HIP(hipMemcpy(c->step_base.p, &v, sizeof v, hipMemcpyHostToDevice));
    c->step = (uint64_t)step;
    c->dev_step = c->step;
    return VMX_OK;
}

int vmx_get_sweep(vmx_ctx* c, int64_t* sweep) {
    if (!c || !sweep) return fail(VMX_ERR_ARG, "null argument");
    *sweep = (int64_t)c->sweep;
    return VMX_OK;
}

int vmx_set_sweep(vmx_ctx* c, int64_t sweep) {
    if (!c || sweep < 0) return fail(VMX_ERR_ARG, "bad argument");
    c->sweep = (uint64_t)sweep;   // a launch argument of each sweep: nothing on the device to update
    return VMX_OK;
}

// ---- checkpoints -------------------------------------------------------------------------
// Layout (little-endian, 8-byte fields):
//   "VMXCHK01" | n | nrep | precision | field flags | Heun step | Metropolis sweep | seed
//   | T[nrep] | H[nrep][3] | sigma[nrep] | moves[nrep] | rejects[nrep] | spins x[n] y[n] z[n]
// Spins are stored in the caller's order as doubles (an FP32 spin converts exactly both ways),
// so a context loaded from a checkpoint continues with the same values, Philox counters and
// Metropolis state as the one that saved it.
namespace {
constexpr char kChkMagic[8] = {'V', 'M', 'X', 'C', 'H', 'K', '0', '1'};
size_t chk_bytes(const vmx_ctx* c) {
    return 8 * 8 + 8 * (size_t)c->nrep * (1 + 3 + 3) + 24 * (size_t)c->n;
}
}  // namespace

int vmx_checkpoint_bytes(vmx_ctx* c, uint64_t* bytes) {
    g_last_error.clear();
    if (!c || !bytes) return fail(VMX_ERR_ARG, "null argument");
    *bytes = chk_bytes(c);
    return VMX_OK;
}

int vmx_checkpoint_save(vmx_ctx* c, void* buf, uint64_t bytes) {
    g_last_error.clear();
    if (!c || !buf) return fail(VMX_ERR_ARG, "null argument");
    if (bytes < chk_bytes(c)) return fail(VMX_ERR_ARG, "checkpoint buffer of %llu bytes, need %llu",
                                          (unsigned long long)bytes, (unsigned long long)chk_bytes(c));
    if (c->hdip_on) return fail(VMX_ERR_STATE, "checkpoints do not hold the dipole field; clear it first");
    VMX_TRY(c->use_device());
    char* p = static_cast<char*>(buf);
    auto put = [&](const void* v, size_t b) { std::memcpy(p, v, b); p += b; };
    auto put64 = [&](uint64_t v) { put(&v, 8); };
    put(kChkMagic, 8);
    put64((uint64_t)c->n);
    put64((uint64_t)c->nrep);
    put64((uint64_t)c->precision);
    put64((uint64_t)c->fld);
    put64(c->step);
    put64(c->sweep);
    put64(c->seed);
    put(c->T_h.data(), 8 * (size_t)c->nrep);
    put(c->H_h.data(), 24 * (size_t)c->nrep);
    std::vector<double> tmp(3 * (size_t)c->n);
    VMX_TRY(vmx_mc_get_state(c, tmp.data(), tmp.data() + c->nrep, tmp.data() + 2 * (size_t)c->nrep));
    put(tmp.data(), 24 * (size_t)c->nrep);
    VMX_TRY(vmx_get_spins(c, tmp.data(), tmp.data() + c->n, tmp.data() + 2 * (size_t)c->n));
    put(tmp.data(), 24 * (size_t)c->n);
    return VMX_OK;
}

int vmx_checkpoint_load(vmx_ctx* c, const void* buf, uint64_t bytes) {
    g_last_error.clear();
    if (!c || !buf) return fail(VMX_ERR_ARG, "null argument");
    if (bytes < chk_bytes(c)) return fail(VMX_ERR_ARG, "checkpoint of %llu bytes, this context needs %llu",
                                          (unsigned long long)bytes, (unsigned long long)chk_bytes(c));
    if (c->hdip_on) return fail(VMX_ERR_STATE, "checkpoints do not hold the dipole field; clear it first");
    const char* p = static_cast<const char*>(buf);
    auto get = [&](void* v, size_t b) { std::memcpy(v, p, b); p += b; };
    auto get64 = [&]() { uint64_t v; get(&v, 8); return v; };
    char magic[8];
    get(magic, 8);
    if (std::memcmp(magic, kChkMagic, 8) != 0) return fail(VMX_ERR_ARG, "not a libvmx checkpoint");
    const uint64_t n = get64(), nrep = get64(), prec = get64(), fld = get64();
    if (n != (uint64_t)c->n || nrep != (uint64_t)c->nrep || prec != (uint64_t)c->precision)
        return fail(VMX_ERR_ARG, "checkpoint of %llu atoms, %llu replicas, precision %llu does not match the context",
                    (unsigned long long)n, (unsigned long long)nrep, (unsigned long long)prec);
    const uint64_t step = get64(), sweep = get64(), seed = get64();
    if (seed != c->seed) return fail(VMX_ERR_ARG, "checkpoint seed %llu differs from the context's %llu",
                                     (unsigned long long)seed, (unsigned long long)c->seed);
    VMX_TRY(c->use_device());
    get(c->T_h.data(), 8 * (size_t)c->nrep);
    get(c->H_h.data(), 24 * (size_t)c->nrep);
    if (((uint32_t)fld & 0x7u) != c->fld) c->drop_graph();
    c->fld = (uint32_t)fld & 0x7u;
    VMX_TRY(c->upload_conditions_any());
    std::vector<double> tmp(3 * (size_t)c->n);
    get(tmp.data(), 24 * (size_t)c->nrep);
    VMX_HIP(hipStreamSynchronize(c->stream));
    const size_t rb = c->nrep * sizeof(double);
    VMX_HIP(hipMemcpy(c->mc_sigma.p, tmp.data(), rb, hipMemcpyHostToDevice));
    VMX_HIP(hipMemcpy(c->mc_tot_moves.p, tmp.data() + c->nrep, rb, hipMemcpyHostToDevice));
    VMX_HIP(hipMemcpy(c->mc_tot_rej.p, tmp.data() + 2 * (size_t)c->nrep, rb, hipMemcpyHostToDevice));
    get(tmp.data(), 24 * (size_t)c->n);
    VMX_TRY(vmx_set_spins(c, tmp.data(), tmp.data() + c->n, tmp.data() + 2 * (size_t)c->n));
    VMX_TRY(vmx_set_step(c, (int64_t)step));
    c->sweep = sweep;
    return VMX_OK;
}

int vmx_mc_sweep(vmx_ctx* c, int32_t nsweeps) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (nsweeps < 0) return fail(VMX_ERR_ARG, "nsweeps must be >= 0");
    if (c->hdip_on && !c->dip.on)
        return fail(VMX_ERR_STATE, "an uploaded dipole field has no demag field for the Metropolis energy; "
                                   "use vmx_set_dipole or clear it");
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? c->mc_sweeps<double>(nsweeps) : c->mc_sweeps<float>(nsweeps);
}

int vmx_mc_get_state(vmx_ctx* c, double* sigma, double* moves, double* rejects) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    const size_t b = c->nrep * sizeof(double);
    if (sigma) VMX_HIP(hipMemcpy(sigma, c->mc_sigma.p, b, hipMemcpyDeviceToHost));
    if (moves) VMX_HIP(hipMemcpy(moves, c->mc_tot_moves.p, b, hipMemcpyDeviceToHost));
    if (rejects) VMX_HIP(hipMemcpy(rejects, c->mc_tot_rej.p, b, hipMemcpyDeviceToHost));
    return VMX_OK;
}

int vmx_mc_set_sigma(vmx_ctx* c, int32_t replica, double sigma) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (replica < -1 || replica >= c->nrep) return fail(VMX_ERR_ARG, "replica out of range");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    const int r0 = replica < 0 ? 0 : replica, r1 = replica < 0 ? c->nrep : replica + 1;
    for (int r = r0; r < r1; ++r)
        VMX_HIP(hipMemcpy(c->mc_sigma.as<double>() + r, &sigma, sizeof(double), hipMemcpyHostToDevice));
    return VMX_OK;
}

int vmx_mc_num_colours(vmx_ctx* c, int32_t* ncolours) {
    if (!c || !ncolours) return fail(VMX_ERR_ARG, "null argument");
    VMX_TRY(c->use_device());
    if (!c->coloured) VMX_TRY(c->colour());
    *ncolours = c->ncolours;
    return VMX_OK;
}

// ---- constrained Monte Carlo ----------------------------------------------------------------
int vmx_cmc_set_constraint(vmx_ctx* c, int32_t replica, double phi, double theta) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (replica < -1 || replica >= c->nrep) return fail(VMX_ERR_ARG, "replica %d outside [-1,%d)", replica, c->nrep);
    if (!std::isfinite(phi) || !std::isfinite(theta)) return fail(VMX_ERR_ARG, "constraint angles must be finite");
    // cmc::polar_rot_matrix (src/montecarlo/cmc.cpp:43-138): P = Y(phi) Z(theta), sums from 0 in
    // the order of vmath::matmul; Pt = transpose; v = (0,0,1) P
    const double pi = 3.14159265358979323846264338327;
    const double dy = (phi / 360.0) * 2.0 * pi, dz = (theta / 360.0) * 2.0 * pi;
    const double sin_y = sin(dy), cos_y = cos(dy), sin_z = sin(dz), cos_z = cos(dz);
    const double Y[9] = {cos_y, 0.0, -sin_y, 0.0, 1.0, 0.0, sin_y, 0.0, cos_y};
    const double Z[9] = {cos_z, sin_z, 0.0, -sin_z, cos_z, 0.0, 0.0, 0.0, 1.0};
    double f[21];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double t = 0.0;
            for (int k = 0; k < 3; ++k) t += Y[3 * a + k] * Z[3 * k + b];
            f[3 * a + b] = t;
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) f[9 + 3 * a + b] = f[3 * b + a];
    const double ref[3] = {0.0, 0.0, 1.0};
    for (int b = 0; b < 3; ++b) {
        double t = 0.0;
        for (int k = 0; k < 3; ++k) t += ref[k] * f[3 * k + b];
        f[18 + b] = t;
    }
    const int r0 = replica < 0 ? 0 : replica, r1 = replica < 0 ? c->nrep : replica + 1;
    for (int r = r0; r < r1; ++r) {
        std::copy(f, f + 21, c->cmc_frame_h.begin() + 21 * (size_t)r);
        c->cmc_frame_set[r] = 1;
    }
    c->cmc_frame_dirty = true;
    return VMX_OK;
}

int vmx_cmc_set_move(vmx_ctx* c, int32_t move) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (move != VMX_CMC_MOVE_CONE && move != VMX_CMC_MOVE_ANGLE) return fail(VMX_ERR_ARG, "unknown move %d", move);
    c->cmc_move = move;
    return VMX_OK;
}

int vmx_cmc_sweep(vmx_ctx* c, int32_t nsweeps) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (nsweeps < 0) return fail(VMX_ERR_ARG, "nsweeps must be >= 0");
    for (int r = 0; r < c->nrep; ++r)
        if (!c->cmc_frame_set[r]) return fail(VMX_ERR_STATE, "no constraint set for replica %d (vmx_cmc_set_constraint)", r);
    if (c->dip.on || c->hdip_on)
        return fail(VMX_ERR_STATE, "constrained Monte Carlo does not refresh a dipole field; clear it first");
    if (nsweeps == 0) return VMX_OK;
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? c->cmc_sweeps<double>(nsweeps) : c->cmc_sweeps<float>(nsweeps);
}

int vmx_cmc_get_counts(vmx_ctx* c, double* counts) {
    if (!c || !counts) return fail(VMX_ERR_ARG, "null argument");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    const size_t per = (size_t)kRejectSlots * kRejectStride;
    std::vector<double> slots((size_t)c->nrep * per);
    VMX_HIP(hipMemcpy(slots.data(), c->cmc_counts.p, slots.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int r = 0; r < c->nrep; ++r)
        for (int k = 0; k < 3; ++k) {
            double t = 0.0;
            for (int q = 0; q < kRejectSlots; ++q) t += slots[r * per + (size_t)q * kRejectStride + k];
            counts[3 * (size_t)r + k] = t;
        }
    return VMX_OK;
}

static int check_mask(vmx_ctx* c, int32_t mask_size, const int32_t* mask, std::vector<int>& empty) {
    if (mask_size < 1 || mask_size > 2048) return fail(VMX_ERR_ARG, "mask_size must be in [1,2048]");
    if (!mask) return fail(VMX_ERR_ARG, "mask is null");
    std::vector<int> count(mask_size, 0);
    for (int i = 0; i < c->n; ++i) {
        if (mask[i] < 0 || mask[i] >= mask_size)
            return fail(VMX_ERR_ARG, "mask id %d of atom %d outside [0,%d)", mask[i], i, mask_size);
        count[mask[i]]++;
    }
    for (int id = 0; id < mask_size - 1; ++id)
        if (count[id] == 0) empty.push_back(id);
    return VMX_OK;
}

int vmx_stats_add_magnetization(vmx_ctx* c, int32_t mask_size, const int32_t* mask, int32_t* handle) {
    if (!c || !handle) return fail(VMX_ERR_ARG, "null argument");
    VMX_TRY(c->use_device());
    std::unique_ptr<MagStat> st(new MagStat());
    VMX_TRY(check_mask(c, mask_size, mask, st->empty));
    st->nbins = mask_size;
    VMX_TRY(st->mask.alloc((size_t)c->n * sizeof(int)));
    VMX_HIP(hipMemcpy(st->mask.p, c->internal_ints(mask).data(), (size_t)c->n * sizeof(int), hipMemcpyHostToDevice));
    VMX_TRY(st->cur.alloc(4 * (size_t)mask_size * sizeof(double)));
    VMX_TRY(st->mean.alloc(4 * (size_t)mask_size * sizeof(double)));
    VMX_HIP(hipMemset(st->cur.p, 0, 4 * (size_t)mask_size * sizeof(double)));
    VMX_HIP(hipMemset(st->mean.p, 0, 4 * (size_t)mask_size * sizeof(double)));
    VMX_TRY(st->empty_dev.alloc(std::max<size_t>(1, st->empty.size()) * sizeof(int)));
    if (!st->empty.empty())
        VMX_HIP(hipMemcpy(st->empty_dev.p, st->empty.data(), st->empty.size() * sizeof(int), hipMemcpyHostToDevice));
    VMX_TRY(st->counter.alloc(sizeof(long long)));
    VMX_HIP(hipMemset(st->counter.p, 0, sizeof(long long)));
    *handle = (int32_t)c->mags.size();
    c->mags.push_back(st.release());
    return VMX_OK;
}

int vmx_stats_add_energy(vmx_ctx* c, int32_t mask_size, const int32_t* mask, int32_t* handle) {
    if (!c || !handle) return fail(VMX_ERR_ARG, "null argument");
    VMX_TRY(c->use_device());
    std::unique_ptr<EnergyStat> st(new EnergyStat());
    VMX_TRY(check_mask(c, mask_size, mask, st->empty));
    st->nbins = mask_size;
    VMX_TRY(st->mask.alloc((size_t)c->n * sizeof(int)));
    VMX_HIP(hipMemcpy(st->mask.p, c->internal_ints(mask).data(), (size_t)c->n * sizeof(int), hipMemcpyHostToDevice));
    VMX_TRY(st->cur.alloc(5 * (size_t)mask_size * sizeof(double)));
    VMX_TRY(st->mean.alloc(5 * (size_t)mask_size * sizeof(double)));
    VMX_HIP(hipMemset(st->cur.p, 0, 5 * (size_t)mask_size * sizeof(double)));
    VMX_HIP(hipMemset(st->mean.p, 0, 5 * (size_t)mask_size * sizeof(double)));
    VMX_TRY(st->empty_dev.alloc(std::max<size_t>(1, st->empty.size()) * sizeof(int)));
    if (!st->empty.empty())
        VMX_HIP(hipMemcpy(st->empty_dev.p, st->empty.data(), st->empty.size() * sizeof(int), hipMemcpyHostToDevice));
    VMX_TRY(st->counter.alloc(sizeof(long long)));
    VMX_HIP(hipMemset(st->counter.p, 0, sizeof(long long)));
    *handle = (int32_t)c->energies.size();
    c->energies.push_back(st.release());
    return VMX_OK;
}

int vmx_stats_update(vmx_ctx* c) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (c->hdip_on && !c->dip.on && !c->energies.empty())
        return fail(VMX_ERR_STATE, "an uploaded dipole field has no demag field for the magnetostatic energy; "
                                   "use vmx_set_dipole or clear it");
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? c->stats_update<double>() : c->stats_update<float>();
}

int vmx_stats_update_noalloc(vmx_ctx* c) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (c->hdip_on && !c->dip.on && !c->energies.empty())
        return fail(VMX_ERR_STATE, "an uploaded dipole field has no demag field for the magnetostatic energy; "
                                   "use vmx_set_dipole or clear it");
    const size_t need = c->stats_partial_bytes();
    if (need > kStatsScratchBytes)
        return fail(VMX_ERR_ARG, "the registered statistics need %zu bytes of workgroup partials; the "
                                 "no-allocation update holds at most %zu (fewer mask bins, or vmx_stats_update)",
                    need, kStatsScratchBytes);
    if (!c->stats_scratch) return fail(VMX_ERR_STATE, "no statistics scratch on this context's device");
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? c->stats_update<double>(c->stats_scratch)
                                    : c->stats_update<float>(c->stats_scratch);
}

int vmx_stats_get_magnetization(vmx_ctx* c, int32_t handle, double* magnetization, double* mean_sum,
                                int64_t* counter) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (handle < 0 || handle >= (int)c->mags.size()) return fail(VMX_ERR_ARG, "bad magnetization handle %d", handle);
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    MagStat* st = c->mags[handle];
    const size_t b = 4 * (size_t)st->nbins * sizeof(double);
    if (magnetization) VMX_HIP(hipMemcpy(magnetization, st->cur.p, b, hipMemcpyDeviceToHost));
    if (mean_sum) VMX_HIP(hipMemcpy(mean_sum, st->mean.p, b, hipMemcpyDeviceToHost));
    if (counter) {
        long long v = 0;
        VMX_HIP(hipMemcpy(&v, st->counter.p, sizeof v, hipMemcpyDeviceToHost));
        *counter = v;
    }
    return VMX_OK;
}

int vmx_stats_get_energy(vmx_ctx* c, int32_t handle, double* energy, double* mean_sum, int64_t* counter) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (handle < 0 || handle >= (int)c->energies.size()) return fail(VMX_ERR_ARG, "bad energy handle %d", handle);
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    EnergyStat* st = c->energies[handle];
    const size_t b = 5 * (size_t)st->nbins * sizeof(double);
    if (energy) VMX_HIP(hipMemcpy(energy, st->cur.p, b, hipMemcpyDeviceToHost));
    if (mean_sum) VMX_HIP(hipMemcpy(mean_sum, st->mean.p, b, hipMemcpyDeviceToHost));
    if (counter) {
        long long v = 0;
        VMX_HIP(hipMemcpy(&v, st->counter.p, sizeof v, hipMemcpyDeviceToHost));
        *counter = v;
    }
    return VMX_OK;
}

int vmx_stats_reset(vmx_ctx* c) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    for (MagStat* st : c->mags) {
        VMX_HIP(hipMemset(st->mean.p, 0, 4 * (size_t)st->nbins * sizeof(double)));
        VMX_HIP(hipMemset(st->counter.p, 0, sizeof(long long)));
    }
    for (EnergyStat* st : c->energies) {
        VMX_HIP(hipMemset(st->mean.p, 0, 5 * (size_t)st->nbins * sizeof(double)));
        VMX_HIP(hipMemset(st->counter.p, 0, sizeof(long long)));
    }
    VMX_HIP(hipDeviceSynchronize());
    return VMX_OK;
}

extern "C++" template <typename R>
static int compute_fields_impl(vmx_ctx* c, const double* noise, double* h_spin, double* h_ext) {
    VMX_TRY(c->plainF.alloc(3 * (size_t)c->n * c->rsize()));
    VMX_TRY(c->plainE.alloc(3 * (size_t)c->n * c->rsize()));
    const SysDev<R> s = c->sys<R>();
    const Launch L = c->launch();
    const double* nd = nullptr;
    if (noise) {
        VMX_TRY(c->noise_dev.alloc(3 * (size_t)c->n * sizeof(double)));
        c->noise_host.resize(3 * (size_t)c->n);
        c->to_internal(noise, c->noise_host.data(), 3);
        VMX_HIP(hipStreamSynchronize(c->stream));
        VMX_HIP(hipMemcpy(c->noise_dev.p, c->noise_host.data(), 3 * (size_t)c->n * sizeof(double),
                          hipMemcpyHostToDevice));
        nd = c->noise_dev.as<double>();
    }
    VMX_HIP(Ops<R>::spin_fields(L, c->et, c->am(), s, c->spins<R>(c->s_idx), c->arr3<R>(c->plainF)));
    VMX_HIP(Ops<R>::ext_fields(L, s, c->arr3<R>(c->plainE), nd, 0u));
    const size_t nb = 3 * (size_t)c->n * sizeof(double);
    double* st = c->staging.as<double>();
    for (int which = 0; which < 2; ++which) {
        double* dst = which == 0 ? h_spin : h_ext;
        if (!dst) continue;
        const DevBuf& src = which == 0 ? c->plainF : c->plainE;
        VMX_HIP(Ops<R>::to_double(c->stream, 3 * c->n, src.as<R>(), st));
        VMX_HIP(hipStreamSynchronize(c->stream));
        std::vector<double> tmp(3 * (size_t)c->n);
        VMX_HIP(hipMemcpy(tmp.data(), st, nb, hipMemcpyDeviceToHost));
        c->to_caller(tmp.data(), dst, 3);
    }
    VMX_HIP(hipStreamSynchronize(c->stream));
    return VMX_OK;
}

int vmx_compute_fields(vmx_ctx* c, const double* noise, double* h_spin, double* h_ext) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? compute_fields_impl<double>(c, noise, h_spin, h_ext)
                                    : compute_fields_impl<float>(c, noise, h_spin, h_ext);
}

extern "C++" template <typename R>
static int thermal_noise_impl(vmx_ctx* c, int64_t step, double* noise) {
    const SysDev<R> s = c->sys<R>();
    double* st = c->staging.as<double>();
    VMX_HIP(Ops<R>::thermal_noise(c->launch(), s, (uint64_t)step, st));
    VMX_HIP(hipStreamSynchronize(c->stream));
    std::vector<double> tmp(3 * (size_t)c->n);
    VMX_HIP(hipMemcpy(tmp.data(), st, 3 * (size_t)c->n * sizeof(double), hipMemcpyDeviceToHost));
    c->to_caller(tmp.data(), noise, 3);
    return VMX_OK;
}

int vmx_thermal_noise(vmx_ctx* c, int64_t step, double* noise) {
    if (!c || !noise || step < 0) return fail(VMX_ERR_ARG, "bad argument");
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? thermal_noise_impl<double>(c, step, noise)
                                    : thermal_noise_impl<float>(c, step, noise);
}

int vmx_synchronize(vmx_ctx* c) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    return VMX_OK;
}

int vmx_last_step_time_ms(vmx_ctx* c, double* ms) {
    if (!c || !ms) return fail(VMX_ERR_ARG, "null argument");
    VMX_TRY(c->use_device());
    *ms = 0.0;
    if (c->timed_steps <= 0) return VMX_OK;
    VMX_HIP(hipEventSynchronize(c->ev1));
    float t = 0.f;
    VMX_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = (double)t / c->timed_steps;
    return VMX_OK;
}

int vmx_time_kernels(vmx_ctx* c, int32_t nsteps, double* avg_ms, int32_t* nkernels) {
    if (!c || !avg_ms || !nkernels) return fail(VMX_ERR_ARG, "null argument");
    if (nsteps <= 0) return fail(VMX_ERR_ARG, "nsteps must be positive");
    if (!(c->dt > 0.0)) return fail(VMX_ERR_STATE, "time step not set (vmx_set_time_step)");
    VMX_TRY(c->use_device());
    *nkernels = c->kernel == VMX_KERNEL_FUSED ? 2 : 5;
    return c->precision == VMX_FP64 ? c->heun_steps<double>(nsteps, nullptr, avg_ms)
                                    : c->heun_steps<float>(nsteps, nullptr, avg_ms);
}

int vmx_fit_forward(vmx_ctx* c, const vmx_fit_params* p, const double* h, const double* spins0, int32_t nsteps,
                    int32_t record_every, int32_t checkpoint_every, double* mz) {
    if (!c || !p || !spins0 || !mz) return fail(VMX_ERR_ARG, "null argument");
    if (c->hdip_on) return fail(VMX_ERR_STATE, "the fitting step has no dipole field; clear it first");
    if (nsteps <= 0 || record_every <= 0 || checkpoint_every <= 0 || nsteps % record_every)
        return fail(VMX_ERR_ARG, "nsteps must be a positive multiple of record_every");
    if (c->nrep != 1 || c->nmat != 1 || c->et != ET_ISO || c->jm != JM_UNIFORM)
        return fail(VMX_ERR_STATE, "fitting needs one replica, one material and uniform isotropic exchange");
    if (!(c->dt > 0.0)) return fail(VMX_ERR_STATE, "time step not set (vmx_set_time_step)");
    if (!c->symmetric_checked) {
        // the adjoint gathers lambda_H over N(j) for the transpose: every bond must be listed both ways
        c->symmetric = true;
        const std::vector<int>& r = c->rowh_;
        const std::vector<int>& col = c->colh_;
        for (int i = 0; i < c->n && c->symmetric; ++i)
            for (int e = r[i]; e < r[i + 1] && c->symmetric; ++e) {
                const int j = col[e];
                bool back = false;
                for (int f = r[j]; f < r[j + 1]; ++f)
                    if (col[f] == i) { back = true; break; }
                c->symmetric = back;
            }
        c->symmetric_checked = true;
    }
    if (!c->symmetric) return fail(VMX_ERR_STATE, "fitting needs a symmetric neighbour list");
    VMX_TRY(c->use_device());
    c->fit.nsteps = nsteps;
    c->fit.record_every = record_every;
    c->fit.K = checkpoint_every;
    c->fit.nrec = nsteps / record_every;
    c->fit.J = p->J;
    c->fit.k = p->ku2;
    c->fit.alpha = p->alpha;
    c->fit.gamma_rel = c->mats[0].gamma_rel;
    for (int q = 0; q < 3; ++q) c->fit.H[q] = p->H[q];
    return c->precision == VMX_FP64 ? c->fit_forward<double>(h, spins0, mz) : c->fit_forward<float>(h, spins0, mz);
}

int vmx_fit_backward(vmx_ctx* c, const double* dL_dmz, double* grad, double* dL_dh) {
    if (!c || !dL_dmz || !grad) return fail(VMX_ERR_ARG, "null argument");
    if (!c->fit.ready) return fail(VMX_ERR_STATE, "vmx_fit_backward needs a vmx_fit_forward first");
    VMX_TRY(c->use_device());
    return c->precision == VMX_FP64 ? c->fit_backward<double>(dL_dmz, grad, dL_dh)
                                    : c->fit_backward<float>(dL_dmz, grad, dL_dh);
}

int vmx_get_launch_config(vmx_ctx* c, int32_t* block_size, int32_t* xcd_remap) {
    g_last_error.clear();
    if (!c || !block_size || !xcd_remap) return fail(VMX_ERR_ARG, "null argument");
    *block_size = c->block;
    *xcd_remap = c->xcd_remap;
    return VMX_OK;
}

int vmx_autotune(vmx_ctx* c, int32_t steps_per_trial, int32_t* chosen) {
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (steps_per_trial <= 0) return fail(VMX_ERR_ARG, "steps_per_trial must be positive");
    if (!(c->dt > 0.0)) return fail(VMX_ERR_STATE, "time step not set (vmx_set_time_step)");
    VMX_TRY(c->use_device());
    // the trials read the dipole field but must not refresh it or advance its clock
    struct PauseDipole {
        vmx_ctx* c;
        int on;
        ~PauseDipole() { c->dip.on = on; }
    } pause{c, c->dip.on};
    c->dip.on = 0;
    // save spins and step counter
    const size_t bytes = 3 * (size_t)c->n * c->rsize();
    DevBuf save;
    VMX_TRY(save.alloc(bytes));
    VMX_HIP(hipMemcpyAsync(save.p, c->spin[c->s_idx].p, bytes, hipMemcpyDeviceToDevice, c->stream));
    const uint64_t step0 = c->step;
    const int s0 = c->s_idx, o0 = c->o_idx;
    const int blocks[] = {64, 128, 256, 512};
    double best = 1e300;
    int best_b = c->block, best_x = c->xcd_remap;
    for (int x = 0; x <= 1; ++x) {
        for (int b : blocks) {
            c->block = b;
            c->xcd_remap = x;
            VMX_TRY(c->build_blocks());
            VMX_TRY(vmx_llg_heun(c, 2, nullptr));   // warm-up
            double ms = 1e300;                       // best of two trials: one is noisy
            for (int t = 0; t < 2; ++t) {
                VMX_TRY(vmx_llg_heun(c, steps_per_trial, nullptr));
                double m = 0.0;
                VMX_TRY(vmx_last_step_time_ms(c, &m));
                if (m > 0.0 && m < ms) ms = m;
            }
            if (ms < best) { best = ms; best_b = b; best_x = x; }
        }
    }
    c->block = best_b;
    c->xcd_remap = best_x;
    VMX_TRY(c->build_blocks());
    c->s_idx = s0; c->o_idx = o0;
    VMX_HIP(hipMemcpyAsync(c->spin[c->s_idx].p, save.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    VMX_HIP(hipStreamSynchronize(c->stream));
    VMX_TRY(vmx_set_step(c, (int64_t)step0));
    if (chosen) *chosen = best_b;
    return VMX_OK;
}

}  // extern "C"

// ---- macrocell dipole field ------------------------------------------------------------------
int vmx_set_dipole(vmx_ctx* c, const vmx_dipole* d) {
    g_last_error.clear();
    if (!c) return fail(VMX_ERR_ARG, "null context");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    c->drop_graph();
    if (!d) {
        c->dip.on = 0;
        c->hdip_on = 0;
        return VMX_OK;
    }
    if (c->nrep != 1) return fail(VMX_ERR_STATE, "the dipole field needs a single replica (this context has %d)", c->nrep);
    const int n = c->n, nc = d->num_cells, nl = d->num_local_cells;
    if (nc <= 0 || nl <= 0 || nl > nc) return fail(VMX_ERR_ARG, "bad cell counts (%d cells, %d local)", nc, nl);
    if (!d->local_cell || !d->atom_cell || !d->volume || !d->tensor)
        return fail(VMX_ERR_ARG, "null dipole array");
    if (d->update_rate < 1) return fail(VMX_ERR_ARG, "update_rate must be >= 1");
    const std::vector<int>& type_h = c->type_h;
    std::vector<int> count(nc, 0);
    std::vector<std::vector<int>> members(nc);
    std::vector<int> acell(n, -1);
    for (int i = 0; i < n; ++i) {                   // caller order: the order cells::mag sums in
        const int cell = d->atom_cell[i];
        if (cell < 0 || cell >= nc) return fail(VMX_ERR_ARG, "atom %d in cell %d of %d", i, cell, nc);
        const int m = type_h[i];
        if (d->magnetic && !d->magnetic[m]) continue;
        const int ii = c->q > 1 ? c->pi[i] : i;
        members[cell].push_back(ii);
        acell[ii] = cell;
        count[cell]++;
    }
    std::vector<char> is_local(nc, 0);
    for (int l = 0; l < nl; ++l) {
        const int cell = d->local_cell[l];
        if (cell < 0 || cell >= nc || is_local[cell]) return fail(VMX_ERR_ARG, "bad or repeated local cell %d", cell);
        is_local[cell] = 1;
    }
    for (int cell = 0; cell < nc; ++cell)
        if (count[cell] > 0 && !is_local[cell])
            return fail(VMX_ERR_ARG, "cell %d holds magnetic atoms but has no tensor row", cell);
    std::vector<int> start(nc + 1, 0), atoms;
    atoms.reserve(n);
    for (int cell = 0; cell < nc; ++cell) {
        atoms.insert(atoms.end(), members[cell].begin(), members[cell].end());
        start[cell + 1] = (int)atoms.size();
    }
    // tensor [6][nl][nc] (row = local cell) to [6][nc][nl]
    const size_t ts = (size_t)nl * nc;
    std::vector<double> tt(6 * ts);
    for (int k = 0; k < 6; ++k)
        for (int l = 0; l < nl; ++l)
            for (int j = 0; j < nc; ++j) tt[k * ts + (size_t)j * nl + l] = d->tensor[k * ts + (size_t)l * nc + j];
    std::vector<double> mus(c->nmat);
    for (int m = 0; m < c->nmat; ++m) mus[m] = c->mats[m].mu_s_SI;
    auto up = [&](DevBuf& b, const void* src, size_t bytes) -> int {
        VMX_TRY(b.alloc(bytes));
        VMX_HIP(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
        return VMX_OK;
    };
    vmx_ctx::Dipole& D = c->dip;
    VMX_TRY(up(D.start, start.data(), start.size() * sizeof(int)));
    VMX_TRY(up(D.atoms, atoms.data(), std::max<size_t>(1, atoms.size()) * sizeof(int)));
    VMX_TRY(up(D.local_cell, d->local_cell, (size_t)nl * sizeof(int)));
    VMX_TRY(up(D.count, count.data(), (size_t)nc * sizeof(int)));
    VMX_TRY(up(D.volume, d->volume, (size_t)nc * sizeof(double)));
    VMX_TRY(up(D.tensor, tt.data(), tt.size() * sizeof(double)));
    VMX_TRY(up(D.acell, acell.data(), (size_t)n * sizeof(int)));
    VMX_TRY(up(D.mus, mus.data(), mus.size() * sizeof(double)));
    for (DevBuf* b : {&D.mag, &D.B, &D.Hd}) {
        VMX_TRY(b->alloc(3 * (size_t)nc * sizeof(double)));
        VMX_HIP(hipMemset(b->p, 0, 3 * (size_t)nc * sizeof(double)));
    }
    VMX_TRY(c->hdip.alloc(3 * (size_t)n * c->rsize()));
    VMX_HIP(hipMemset(c->hdip.p, 0, 3 * (size_t)n * c->rsize()));
    VMX_TRY(D.hd_atom.alloc(3 * (size_t)n * sizeof(double)));
    VMX_HIP(hipMemset(D.hd_atom.p, 0, 3 * (size_t)n * sizeof(double)));
    D.acell_h = acell;
    D.nc = nc; D.nl = nl; D.rate = d->update_rate; D.time = 0;
    D.on = 1;
    c->hdip_on = 1;
    // the field of the current spins at time 0, as dipole::initialize computes it
    VMX_TRY(c->precision == VMX_FP64 ? c->dipole_refresh<double>() : c->dipole_refresh<float>());
    VMX_HIP(hipStreamSynchronize(c->stream));
    return VMX_OK;
}

int vmx_update_dipole(vmx_ctx* c) {
    g_last_error.clear();
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (!c->dip.on) return fail(VMX_ERR_STATE, "no dipole field set (vmx_set_dipole)");
    VMX_TRY(c->use_device());
    VMX_TRY(c->precision == VMX_FP64 ? c->dipole_refresh<double>() : c->dipole_refresh<float>());
    VMX_HIP(hipStreamSynchronize(c->stream));
    return VMX_OK;
}

int vmx_get_dipole_field(vmx_ctx* c, double* bx, double* by, double* bz, double* hx, double* hy, double* hz) {
    g_last_error.clear();
    if (!c) return fail(VMX_ERR_ARG, "null context");
    if (!c->dip.on) return fail(VMX_ERR_STATE, "no dipole field set (vmx_set_dipole)");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    const int n = c->n, nc = c->dip.nc;
    std::vector<double> B(3 * (size_t)nc), H(3 * (size_t)nc);
    VMX_HIP(hipMemcpy(B.data(), c->dip.B.p, B.size() * sizeof(double), hipMemcpyDeviceToHost));
    VMX_HIP(hipMemcpy(H.data(), c->dip.Hd.p, H.size() * sizeof(double), hipMemcpyDeviceToHost));
    double* out[6] = {bx, by, bz, hx, hy, hz};
    for (int i = 0; i < n; ++i) {
        const int cell = c->dip.acell_h[c->q > 1 ? c->pi[i] : i];
        for (int k = 0; k < 6; ++k) {
            if (!out[k]) continue;
            const std::vector<double>& src = k < 3 ? B : H;
            out[k][i] = cell < 0 ? 0.0 : src[(size_t)(k % 3) * nc + cell];
        }
    }
    return VMX_OK;
}

int vmx_set_atom_field(vmx_ctx* c, const double* hx, const double* hy, const double* hz) {
    g_last_error.clear();
    if (!c) return fail(VMX_ERR_ARG, "null context");
    VMX_TRY(c->use_device());
    VMX_HIP(hipStreamSynchronize(c->stream));
    if (!hx && !hy && !hz) {
        c->drop_graph();
        c->dip.on = 0;
        c->hdip_on = 0;
        return VMX_OK;
    }
    if (!hx || !hy || !hz) return fail(VMX_ERR_ARG, "null field component");
    if (c->dip.on) return fail(VMX_ERR_STATE, "the device dipole solver owns the field (vmx_set_dipole(ctx, NULL) first)");
    const int n = c->n;
    std::vector<double> host(3 * (size_t)n), in(3 * (size_t)n);
    std::copy(hx, hx + n, in.begin());
    std::copy(hy, hy + n, in.begin() + n);
    std::copy(hz, hz + n, in.begin() + 2 * (size_t)n);
    c->to_internal(in.data(), host.data(), 3);
    VMX_TRY(c->hdip.alloc(3 * (size_t)n * c->rsize()));
    if (c->precision == VMX_FP64) {
        VMX_HIP(hipMemcpy(c->hdip.p, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(host.begin(), host.end());
        VMX_HIP(hipMemcpy(c->hdip.p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if (!c->hdip_on) c->drop_graph();
    c->hdip_on = 1;
    return VMX_OK;
}
