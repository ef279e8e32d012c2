// vmx_device.hpp — device-side data structures and kernels of the spin-dynamics hot path.
//
// Included by vmx_f64.hip (R = double, compiled with -ffp-contract=off so every multiply and
// add rounds exactly where the reference's C++ rounds) and vmx_f32.hip (R = float, FMA
// contraction allowed).  Only the launchers declared in vmx_internal.hpp leave these units.
//
// Layout in HBM (one context):
//   spins         three SoA arrays of R per buffer (x[N], y[N], z[N]); buffers S, S' and S_next
//   topology      row[N+1] int32 CSR offsets, col[nnz] int32 neighbour ids (row order as given)
//   exchange      deduplicated table of 1/3/9 R per distinct J; per-edge uint16 or int32 index
//                 into it, or no per-edge data at all when every edge carries the same J
//   type          uint8 material id per atom
//   tables        material constants [M], conditions [replicas][M], block table, replica offsets
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vmx {

// ---- plain-old-data shared with the host ------------------------------------------------
enum { ET_ISO = 0, ET_VEC = 1, ET_TEN = 2 };      // exchange type (exchange::exchange_t)
enum { JM_UNIFORM = 0, JM_IDX16 = 1, JM_IDX32 = 2 };  // how an edge finds its J
// How a row finds its neighbours.  All three keep the reference's per-row neighbour order.
//   TP_CSR  row offsets + neighbour ids (the reference layout)
//   TP_ELL  neighbour ids transposed to [slot][atom] with a per-atom degree: a wavefront's 64
//           rows read one coalesced 256-byte line per slot
//   TP_PAT  per-atom stencil id into a small table of (degree, j - i offsets, J ids): rows of a
//           periodic lattice share a handful of stencils, so a row costs 2 bytes of topology
enum { TP_CSR = 0, TP_ELL = 1, TP_PAT = 2 };
// ELL and stencil rows have a compile-time slot width W (rows padded with in-bounds entries);
// CSR rows are walked with a run-time loop.
constexpr int kSlotWidths[4] = {6, 8, 12, 16};
constexpr int kNumModes = 23;
// Access mode (a kernel template parameter): 0..2 CSR x J mode; 3..14 ELL x J mode x width;
// 15..22 stencil x {uniform, 16-bit J id} x width.
constexpr int am_of(int tp, int jm, int wsel) {
    return tp == TP_CSR ? jm : (tp == TP_ELL ? 3 + jm * 4 + wsel : 15 + jm * 4 + wsel);
}
template <int AM>
struct Access {
    static constexpr int TP = AM < 3 ? TP_CSR : (AM < 15 ? TP_ELL : TP_PAT);
    static constexpr int JM = AM < 3 ? AM : (AM < 15 ? (AM - 3) / 4 : (AM - 15) / 4);
    static constexpr int W = AM < 3 ? 0 : kSlotWidths[AM < 15 ? (AM - 3) % 4 : (AM - 15) % 4];
};

enum : uint32_t {
    ANI_UNI2 = 0x1u, ANI_CUB4 = 0x2u, ANI_CUB4R = 0x4u,
    FLD_THERMAL = 0x1u, FLD_APPLIED = 0x2u, FLD_LOCAL = 0x4u,
    FLD_DEMAG = 0x8u          // set by the library while a thin-film demag field is set
};

template <typename R>
struct MatConst {            // per material, set once (mp::material after set_derived_parameters)
    R one_opa2;              // -gamma_rel/(1+alpha^2)
    R alpha_opa2;            // alpha * one_opa2
    R ku2, ku[3];            // uniaxial constant (T) and axis
    R kc4;                   // cubic constant (T)
    R e1[3], e2[3], e3[3];   // rotated cubic basis
    R mu_s_SI;               // for the Monte Carlo energy scale
    R alpha;                 // Gilbert damping (the midpoint integrator's F = H + alpha S x H)
};

template <typename R>
struct Cond {                // per (replica, material); rewritten when conditions change
    R sigma;                 // sqrt(T_rescaled) * H_th_sigma   (src/simulate/fields.cpp:262-271)
    R happ[3];               // applied field of the LLG step: H or H + H_local (fields.cpp:171-210)
    R hglob[3];              // H of the applied-field energy (src/simulate/energy.cpp:94-98)
    R kBTBohr;               // mu_B / (kB T_rescaled)            (src/montecarlo/mc.cpp:45-53)
    R hdemag[3];             // thin-film demag field, added after happ (fields.cpp:219-239)
    R cmc_delta;             // width of the constrained Monte Carlo angle move (cmc.cpp:320-325)
};

template <typename R>
struct SysDev {
    int n;                                   // atoms over all replicas
    int nmat;
    int nblocks;                             // logical workgroups covering all replicas
    int xcd_remap;                           // 1: contiguous logical blocks per XCD
    int zmax;                                // widest row (ELL slots, stencil width)
    int full_rows;                           // ELL / stencil: every row has exactly W neighbours
    int npat;                                // stencils in the table
    int pat_lds;                             // 1: fused kernels stage the stencil offsets in LDS
    int am;                                  // access mode the arrays below were built for
    int q;                                   // basis-major atom order with period q (1: caller's order)
    const int* __restrict__ rep_m;           // [nrep] replica size / q
    const int* __restrict__ row;             // CSR: [n+1]
    const int* __restrict__ col;             // CSR: [nnz]; ELL: [zmax][n]
    const void* __restrict__ jidx;           // per-edge table index (uint16 / int32), CSR or ELL layout
    const uint8_t* __restrict__ deg;         // ELL: [n] row lengths
    const uint16_t* __restrict__ pat;        // PAT: [n] stencil id
    const int* __restrict__ pat_off;         // PAT: [npat][zmax] j - i
    const uint16_t* __restrict__ pat_jid;    // PAT: [npat][zmax] J table index
    const uint8_t* __restrict__ pat_deg;     // PAT: [npat] row lengths
    const R* __restrict__ jtab;              // exchange table
    const uint8_t* __restrict__ type;        // [n]
    const MatConst<R>* __restrict__ mat;     // [nmat]
    const Cond<R>* __restrict__ cond;        // [nrep*nmat]
    const int4* __restrict__ blk;            // [nblocks] {first atom, replica, end atom, replica start}
    const int* __restrict__ rep_off;         // [nrep+1]
    const uint2* __restrict__ rep_key;       // [nrep] Philox key
    const unsigned long long* __restrict__ step_base;  // Heun steps completed before this launch batch
    const R* __restrict__ hdip;              // [3][n] per-atom dipole field (x, y, z blocks), or null
    const double* __restrict__ hd;           // [3][n] per-atom mu0 Hd of the device dipole solver, or null
    uint32_t ani;                            // ANI_* flags
    uint32_t fld;                            // FLD_* flags
    R dt, half_dt;
};

template <typename R>
struct Spin3 {
    R* __restrict__ x;
    R* __restrict__ y;
    R* __restrict__ z;
};

// The spins a corrector reads (S) and writes (O) when the step overwrites S in place: no
// __restrict__, since the two may be the same arrays.  A thread reads S only at its own atoms
// before it stores them (the in-place steps gather their neighbours from the predicted spins).
template <typename R>
struct Spin3Alias {
    R* x;
    R* y;
    R* z;
};
template <typename R>
__host__ __device__ inline Spin3Alias<R> alias(const Spin3<R>& a) {
    return Spin3Alias<R>{a.x, a.y, a.z};
}

// Parameters of the deterministic fitting step (vmx_fit.inc): exchange J (T) of every bond,
// uniaxial k (T) along e, LLG prefactors a = -gamma/(1+alpha^2), b = alpha a, uniform field H.
template <typename R>
struct FitPrm {
    R J, k, a, b, dt, half_dt;
    R e[3];
    R H[3];
};

// ---- Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box with Weyl key schedule ---------------------
struct U4 { uint32_t v[4]; };

__host__ __device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                     uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 product gives both halves (v_mad_u64_u32 instead of mul_hi + mul_lo)
        const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += W0; k1 += W1;
    }
    U4 o; o.v[0] = c0; o.v[1] = c1; o.v[2] = c2; o.v[3] = c3;
    return o;
}

// Stream ids: the fourth counter word separates independent uses of one (atom, step) pair.
enum : uint32_t { STREAM_THERMAL = 0u, STREAM_MC_MOVE = 1u, STREAM_MC_ACCEPT = 2u,
                  // constrained Monte Carlo: a round's pair permutation, and per pair the move
                  // normals (9 + round) and the acceptance uniform (11 + round)
                  STREAM_CMC_PAIR = 8u, STREAM_CMC_MOVE = 9u, STREAM_CMC_ACCEPT = 11u };

// Three standard normals from one Philox block by Box-Muller on (x0,x1) and (x2,x3);
// uniforms are (k + 0.5) * 2^-32, never 0 or 1.
__device__ __forceinline__ void box_muller3(const U4& u, double& g0, double& g1, double& g2) {
    const double s = 2.3283064365386963e-10;  // 2^-32
    const double twopi = 6.283185307179586;
    const double u1 = ((double)u.v[0] + 0.5) * s, u2 = ((double)u.v[1] + 0.5) * s;
    const double u3 = ((double)u.v[2] + 0.5) * s, u4 = ((double)u.v[3] + 0.5) * s;
    const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    double sn, cs;
    sincos(twopi * u2, &sn, &cs);
    g0 = r1 * cs;
    g1 = r1 * sn;
    g2 = r2 * cos(twopi * u4);
}

// FP32 variant on the hardware transcendental units: v_log_f32 is log2, v_sin/v_cos_f32 take
// their argument in revolutions, so sin(2 pi u) is one instruction.  The uniforms are the FP64
// ones, (k + 0.5) 2^-32, rounded to float (relative error 2^-24 at every k): small words keep
// their full resolution, so the tail of the radius sqrt(-2 ln u1) follows the FP64 normals
// instead of jumping in steps of 2^-24 in u1.
__device__ __forceinline__ void box_muller3(const U4& u, float& g0, float& g1, float& g2) {
    const float s = 2.3283064365386963e-10f;  // 2^-32
    const float u1 = ((float)u.v[0] + 0.5f) * s, u2 = ((float)u.v[1] + 0.5f) * s;
    const float u3 = ((float)u.v[2] + 0.5f) * s, u4 = ((float)u.v[3] + 0.5f) * s;
    const float ln2x2 = 1.3862943611198906f;  // 2 ln 2
    const float r1 = __builtin_amdgcn_sqrtf(-ln2x2 * __builtin_amdgcn_logf(u1));
    const float r2 = __builtin_amdgcn_sqrtf(-ln2x2 * __builtin_amdgcn_logf(u3));
    g0 = r1 * __builtin_amdgcn_cosf(u2);
    g1 = r1 * __builtin_amdgcn_sinf(u2);
    g2 = r2 * __builtin_amdgcn_cosf(u4);
}

template <typename R>
__device__ __forceinline__ void thermal_normals(const SysDev<R>& s, int rep, int local, uint64_t step,
                                                R& gx, R& gy, R& gz) {
    const uint2 key = s.rep_key[rep];
    const U4 u = philox4x32_10((uint32_t)local, (uint32_t)step, (uint32_t)(step >> 32),
                               STREAM_THERMAL, key.x, key.y);
    box_muller3(u, gx, gy, gz);
}

// ---- workgroup -> atoms ------------------------------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs.  With xcd_remap the grid is padded to
// a multiple of 8 and hardware block b takes logical block (b%8)*(grid/8) + b/8, so each XCD
// walks one contiguous stretch of the block table (one range of lattice cells, all basis atoms)
// and the neighbour rows it gathers stay in its own L2.
__device__ __forceinline__ int logical_block(int xcd_remap) {
    const int b = (int)blockIdx.x;
    if (!xcd_remap) return b;
    const int per = (int)gridDim.x >> 3;
    return (b & 7) * per + (b >> 3);
}

template <typename R>
__device__ __forceinline__ int caller_local(const SysDev<R>& s, int rep, int p) {
    // internal replica-local index p = b * m + c (basis b, cell c) -> caller's c * q + b
    if (s.q <= 1) return p;
    const int m = s.rep_m[rep];
    int b = 0;
    while (p >= m) { p -= m; ++b; }   // at most q - 1 iterations
    return p * s.q + b;
}

// i: internal atom index; local: the caller's replica-local index (keys the Philox counter)
template <typename R>
__device__ __forceinline__ bool my_atom(const SysDev<R>& s, int& i, int& rep, int& local) {
    const int lb = logical_block(s.xcd_remap);
    if (lb >= s.nblocks) return false;
    const int4 b = s.blk[lb];
    rep = b.y;
    i = b.x + (int)threadIdx.x;
    if (i >= b.z) return false;
    local = caller_local(s, rep, i - b.w);
    return true;
}

// ---- raw buffer access -------------------------------------------------------------------
// A load through a buffer resource takes the array base in SGPRs and a 32-bit byte offset in one
// VGPR: a neighbour costs one shift of its index, shared by its x, y and z loads, instead of a
// 64-bit address per component.  Component arrays stay below 4 GiB (checked at creation).
constexpr int kRawBufferConfig = 0x00020000;   // gfx9 raw buffer word 3: 32-bit data format
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kRawBufferConfig);
}
__device__ __forceinline__ void buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, float& v) {
    v = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off, double& v) {
    v = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// ---- neighbours of atom i, in row order ------------------------------------------------------
// ELL and stencil rows are read in three phases: the row's indices (row_index), then every
// neighbour spin (row_load), then the sum in row order over slots k < degree (row_sum).  Kernels
// that handle several atoms per thread run each phase for all of them before the next, so the
// memory latency of a phase is paid once per thread.  Padding slots load in-bounds entries and
// are not summed; when every row is full (bulk lattices) a uniform branch skips the per-slot
// degree test.
template <int AM>
struct RowIdx {
    static constexpr int W = Access<AM>::W > 0 ? Access<AM>::W : 1;
    int j[W];       // neighbour atom
    int jd[W];      // J table index (16/32-bit J modes)
    int deg;        // row length
};
template <typename R, int AM>
struct RowSpins {
    static constexpr int W = RowIdx<AM>::W;
    R x[W], y[W], z[W];
};

// First load of a row: the stencil id (stencil layout; nothing for ELL).
template <typename R, int AM>
__device__ __forceinline__ int row_head(const SysDev<R>& s, int i) {
    if constexpr (Access<AM>::TP == TP_PAT) return s.pat[i];
    return 0;
}

// off_tab: the stencil offset table, in global memory or staged in LDS.
template <typename R, int AM>
__device__ __forceinline__ void row_index(const SysDev<R>& s, int i, int head, const int* __restrict__ off_tab,
                                          RowIdx<AM>& r) {
    constexpr int TP = Access<AM>::TP, JM = Access<AM>::JM, W = Access<AM>::W;
    static_assert(TP != TP_CSR, "CSR rows are walked, not staged");
    r.deg = W;
    if constexpr (TP == TP_ELL) {
        if (!s.full_rows) r.deg = s.deg[i];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const size_t e = (size_t)k * (size_t)s.n + (size_t)i;
            r.j[k] = s.col[e];
            if (JM == JM_IDX16) r.jd[k] = ((const uint16_t*)s.jidx)[e];
            if (JM == JM_IDX32) r.jd[k] = ((const int*)s.jidx)[e];
        }
    } else {
        const int p = head;
        if (!s.full_rows) r.deg = s.pat_deg[p];
        const int* __restrict__ off = off_tab + p * W;
        if constexpr (W % 4 == 0) {
#pragma unroll
            for (int q = 0; q < W / 4; ++q) {
                const int4 o = reinterpret_cast<const int4*>(off)[q];
                r.j[4 * q + 0] = i + o.x; r.j[4 * q + 1] = i + o.y;
                r.j[4 * q + 2] = i + o.z; r.j[4 * q + 3] = i + o.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < W / 2; ++q) {
                const int2 o = reinterpret_cast<const int2*>(off)[q];
                r.j[2 * q + 0] = i + o.x; r.j[2 * q + 1] = i + o.y;
            }
        }
        if (JM != JM_UNIFORM) {
#pragma unroll
            for (int k = 0; k < W; ++k) r.jd[k] = s.pat_jid[p * W + k];
        }
    }
}

// Buffer resources of one SoA spin buffer (n reals per component).
template <typename R>
struct SpinRsrc {
    __amdgpu_buffer_rsrc_t x, y, z;
    __device__ __forceinline__ SpinRsrc(const SysDev<R>& s, const R* X, const R* Y, const R* Z) {
        const uint32_t bytes = (uint32_t)s.n * (uint32_t)sizeof(R);
        x = make_rsrc(X, bytes); y = make_rsrc(Y, bytes); z = make_rsrc(Z, bytes);
    }
    __device__ __forceinline__ void load(int j, R& vx, R& vy, R& vz) const {
        const uint32_t o = (uint32_t)j * (uint32_t)sizeof(R);
        buf_ld(x, o, vx); buf_ld(y, o, vy); buf_ld(z, o, vz);
    }
};

template <typename R, int AM>
__device__ __forceinline__ void row_load(const SpinRsrc<R>& rs, const RowIdx<AM>& r, RowSpins<R, AM>& v) {
#pragma unroll
    for (int k = 0; k < Access<AM>::W; ++k) rs.load(r.j[k], v.x[k], v.y[k], v.z[k]);
}

template <typename R, int ET, int AM, typename F>
__device__ __forceinline__ void row_sum(const SysDev<R>& s, const RowIdx<AM>& r, const RowSpins<R, AM>& v,
                                        const R* ju, F&& f) {
    constexpr int K = (ET == ET_ISO) ? 1 : (ET == ET_VEC ? 3 : 9);
    constexpr int JM = Access<AM>::JM, W = Access<AM>::W;
    if (s.full_rows) {
#pragma unroll
        for (int k = 0; k < W; ++k) f(v.x[k], v.y[k], v.z[k], JM == JM_UNIFORM ? ju : s.jtab + (size_t)K * r.jd[k]);
    } else {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            if (k < r.deg) f(v.x[k], v.y[k], v.z[k], JM == JM_UNIFORM ? ju : s.jtab + (size_t)K * r.jd[k]);
        }
    }
}

// Calls f(Sx, Sy, Sz, J) for every neighbour of atom i in row order, with its spin and a pointer
// to its 1/3/9 exchange reals.
template <typename R, int ET, int AM, typename F>
__device__ __forceinline__ void for_neighbours(const SysDev<R>& s, const R* __restrict__ X,
                                               const R* __restrict__ Y, const R* __restrict__ Z,
                                               int i, F&& f) {
    constexpr int K = (ET == ET_ISO) ? 1 : (ET == ET_VEC ? 3 : 9);
    constexpr int TP = Access<AM>::TP, JM = Access<AM>::JM;
    const SpinRsrc<R> rs(s, X, Y, Z);
    R ju[K];
    if (JM == JM_UNIFORM) {
#pragma unroll
        for (int k = 0; k < K; ++k) ju[k] = s.jtab[k];
    }
    if constexpr (TP == TP_CSR) {
        const int beg = s.row[i], end = s.row[i + 1];
        for (int nn = beg; nn < end; ++nn) {
            const R* J = ju;
            if (JM == JM_IDX16) J = s.jtab + (size_t)K * ((const uint16_t*)s.jidx)[nn];
            if (JM == JM_IDX32) J = s.jtab + (size_t)K * ((const int*)s.jidx)[nn];
            R sx, sy, sz;
            rs.load(s.col[nn], sx, sy, sz);
            f(sx, sy, sz, J);
        }
    } else {
        RowIdx<AM> r;
        row_index(s, i, row_head<R, AM>(s, i), s.pat_off, r);
        RowSpins<R, AM> v;
        row_load(rs, r, v);
        row_sum<R, ET, AM>(s, r, v, ju, f);
    }
}

// ---- spin-dependent field H_spin(S) at atom i ------------------------------------------------
// Exchange: src/exchange/exchange_fields.cpp:52-78 (isotropic), :85-116 (vectorial),
// :122-162 (tensorial); the sum starts at 0 and adds one neighbour at a time in row order.
// Anisotropy terms are then added in the order of src/anisotropy/fields.cpp:44-59.
template <typename R, int ET>
__device__ __forceinline__ void exchange_term(const R* J, R Sx, R Sy, R Sz, R& hx, R& hy, R& hz) {
    if (ET == ET_ISO) {
        const R Jij = J[0];
        hx += Jij * Sx;
        hy += Jij * Sy;
        hz += Jij * Sz;
    } else if (ET == ET_VEC) {
        hx += J[0] * Sx;
        hy += J[1] * Sy;
        hz += J[2] * Sz;
    } else {
        hx += (J[0] * Sx + J[1] * Sy + J[2] * Sz);
        hy += (J[3] * Sx + J[4] * Sy + J[5] * Sz);
        hz += (J[6] * Sx + J[7] * Sy + J[8] * Sz);
    }
}

// fx..fz = 0 + exchange (the field array starts at zero, fields.cpp:79-81), then anisotropy.
template <typename R>
__device__ __forceinline__ void add_anisotropy(uint32_t ani, const MatConst<R>& mc, R sx, R sy, R sz, R hx, R hy,
                                               R hz, R& fx, R& fy, R& fz) {
    fx = R(0); fy = R(0); fz = R(0);
    fx += hx; fy += hy; fz += hz;
    if (ani) {
        if (ani & ANI_UNI2) {  // uniaxial_second_order.cpp:68-94, scale 2
            const R ex = mc.ku[0], ey = mc.ku[1], ez = mc.ku[2];
            const R sdote = (sx * ex + sy * ey + sz * ez);
            const R k2 = R(2.0) * mc.ku2 * sdote;
            fx += ex * k2; fy += ey * k2; fz += ez * k2;
        }
        if (ani & ANI_CUB4) {  // cubic_fourth_order.cpp:48-68, scale 0.5*4
            const R k4 = R(2.0) * mc.kc4;
            fx += sx * sx * sx * k4;
            fy += sy * sy * sy * k4;
            fz += sz * sz * sz * k4;
        }
        if (ani & ANI_CUB4R) {  // cubic_fourth_order_rotation.cpp:48-95
            const R k4 = R(2.0) * mc.kc4;
            const R d1 = sx * mc.e1[0] + sy * mc.e1[1] + sz * mc.e1[2];
            const R d2 = sx * mc.e2[0] + sy * mc.e2[1] + sz * mc.e2[2];
            const R d3 = sx * mc.e3[0] + sy * mc.e3[1] + sz * mc.e3[2];
            const R d1c = d1 * d1 * d1, d2c = d2 * d2 * d2, d3c = d3 * d3 * d3;
            fx += k4 * (d1c * mc.e1[0] + d2c * mc.e2[0] + d3c * mc.e3[0]);
            fy += k4 * (d1c * mc.e1[1] + d2c * mc.e2[1] + d3c * mc.e3[1]);
            fz += k4 * (d1c * mc.e1[2] + d2c * mc.e2[2] + d3c * mc.e3[2]);
        }
    }
}

template <typename R, int ET, int AM>
__device__ __forceinline__ void spin_field(const SysDev<R>& s, const R* __restrict__ X,
                                           const R* __restrict__ Y, const R* __restrict__ Z,
                                           int i, int m, R sx, R sy, R sz,
                                           R& fx, R& fy, R& fz) {
    R hx = R(0), hy = R(0), hz = R(0);
    for_neighbours<R, ET, AM>(s, X, Y, Z, i, [&](const R Sx, const R Sy, const R Sz, const R* J) {
        exchange_term<R, ET>(J, Sx, Sy, Sz, hx, hy, hz);
    });
    add_anisotropy(s.ani, s.mat[m], sx, sy, sz, hx, hy, hz, fx, fy, fz);
}

// The same field from a staged row, with the atom's material record mc.
// The spin field of a staged row: its neighbours' spins gathered, then summed in row order.  FP64
// rows of eight or more slots go in Parts parts (the first part's loads and sum, then the next
// part's): the same terms in the same order, with a part's gathered values live at once, so the
// FP64 kernels need fewer registers and keep more wavefronts per SIMD in flight.
template <typename R, int ET, int AM, int Parts>
__device__ __forceinline__ void row_gather_field(const SysDev<R>& s, const SpinRsrc<R>& rs, const RowIdx<AM>& r,
                                                 const R* ju, const MatConst<R>& mc, R sx, R sy, R sz, R& fx,
                                                 R& fy, R& fz);

template <typename R, int ET, int AM>
__device__ __forceinline__ void row_field(const SysDev<R>& s, const RowIdx<AM>& r, const RowSpins<R, AM>& v,
                                          const R* ju, const MatConst<R>& mc, R sx, R sy, R sz, R& fx, R& fy,
                                          R& fz) {
    R hx = R(0), hy = R(0), hz = R(0);
    row_sum<R, ET, AM>(s, r, v, ju, [&](const R Sx, const R Sy, const R Sz, const R* J) {
        exchange_term<R, ET>(J, Sx, Sy, Sz, hx, hy, hz);
    });
    add_anisotropy(s.ani, mc, sx, sy, sz, hx, hy, hz, fx, fy, fz);
}

template <typename R, int ET, int AM, int Parts>
__device__ __forceinline__ void row_gather_field(const SysDev<R>& s, const SpinRsrc<R>& rs, const RowIdx<AM>& r,
                                                 const R* ju, const MatConst<R>& mc, R sx, R sy, R sz, R& fx,
                                                 R& fy, R& fz) {
    constexpr int W = Access<AM>::W;
    if constexpr (Parts > 1 && sizeof(R) == 8 && W >= 8) {
        constexpr int K = (ET == ET_ISO) ? 1 : (ET == ET_VEC ? 3 : 9);
        constexpr int JM = Access<AM>::JM, H = (W + Parts - 1) / Parts;
        R hx = R(0), hy = R(0), hz = R(0);
#pragma unroll
        for (int part = 0; part < Parts; ++part) {
            R vx[H], vy[H], vz[H];
#pragma unroll
            for (int k = 0; k < H; ++k)
                if (part * H + k < W) rs.load(r.j[part * H + k], vx[k], vy[k], vz[k]);
#pragma unroll
            for (int k = 0; k < H; ++k) {
                const int slot = part * H + k;
                if (slot < W && (s.full_rows || slot < r.deg))
                    exchange_term<R, ET>(JM == JM_UNIFORM ? ju : s.jtab + (size_t)K * r.jd[slot], vx[k], vy[k],
                                         vz[k], hx, hy, hz);
            }
            if (part + 1 < Parts) __builtin_amdgcn_sched_barrier(0);   // the next part's loads stay here
        }
        add_anisotropy(s.ani, mc, sx, sy, sz, hx, hy, hz, fx, fy, fz);
    } else {
        RowSpins<R, AM> rv;
        row_load(rs, r, rv);
        row_field<R, ET, AM>(s, r, rv, ju, mc, sx, sy, sz, fx, fy, fz);
    }
}

// ---- external field at atom i (thermal + applied + dipole), src/simulate/fields.cpp:120-310 ---
// The array starts at 0; the thermal term overwrites it with N(0,1)*sigma (generate, then *=),
// then the applied field is added, and last the macrocell dipole field when one is set
// (calculate_dipolar_fields, fields.cpp:166, 291-310).
// The unit normals of the thermal field at atom i: injected (noise) or from Philox.
template <typename R>
__device__ __forceinline__ void unit_normals(const SysDev<R>& s, int rep, int i, int local, uint64_t step,
                                             const double* __restrict__ noise, R& gx, R& gy, R& gz) {
    if (noise) {
        gx = (R)noise[i];
        gy = (R)noise[(size_t)s.n + i];
        gz = (R)noise[2 * (size_t)s.n + i];
    } else {
        thermal_normals(s, rep, local, step, gx, gy, gz);
    }
}
// H_ext from the atom's unit normals (read only when the thermal field is on).
template <typename R>
__device__ __forceinline__ void ext_from_normals(const SysDev<R>& s, const Cond<R>& c, int i, R gx, R gy, R gz,
                                                 R& ex, R& ey, R& ez) {
    ex = R(0); ey = R(0); ez = R(0);
    if (s.fld & FLD_THERMAL) {
        ex = gx * c.sigma; ey = gy * c.sigma; ez = gz * c.sigma;
    }
    if (s.fld & FLD_APPLIED) {
        ex += c.happ[0]; ey += c.happ[1]; ez += c.happ[2];
        if (s.fld & FLD_DEMAG) {
            ex += c.hdemag[0]; ey += c.hdemag[1]; ez += c.hdemag[2];
        }
    }
    if (s.hdip) {
        ex += s.hdip[i]; ey += s.hdip[(size_t)s.n + i]; ez += s.hdip[2 * (size_t)s.n + i];
    }
}
template <typename R>
__device__ __forceinline__ void ext_field(const SysDev<R>& s, const Cond<R>& c, int rep, int i, int local,
                                          uint64_t step, const double* __restrict__ noise,
                                          R& ex, R& ey, R& ez) {
    R gx = R(0), gy = R(0), gz = R(0);
    if (s.fld & FLD_THERMAL) unit_normals(s, rep, i, local, step, noise, gx, gy, gz);
    ext_from_normals(s, c, i, gx, gy, gz, ex, ey, ez);
}
template <typename R>
__device__ __forceinline__ void ext_field(const SysDev<R>& s, int rep, int m, int i, int local,
                                          uint64_t step, const double* __restrict__ noise,
                                          R& ex, R& ey, R& ez) {
    ext_field(s, s.cond[rep * s.nmat + m], rep, i, local, step, noise, ex, ey, ez);
}

// ---- LLG right-hand side, src/simulate/LLGHeun.cpp:201-203 (term order kept) --------------
template <typename R>
__device__ __forceinline__ void llg_rhs(R a, R b, R S0, R S1, R S2, R H0, R H1, R H2,
                                        R& d0, R& d1, R& d2) {
    d0 = (a) * (S1 * H2 - S2 * H1) + (b) * (S1 * (S0 * H1 - S1 * H0) - S2 * (S2 * H0 - S0 * H2));
    d1 = (a) * (S2 * H0 - S0 * H2) + (b) * (S2 * (S1 * H2 - S2 * H1) - S0 * (S0 * H1 - S1 * H0));
    d2 = (a) * (S0 * H1 - S1 * H0) + (b) * (S0 * (S2 * H0 - S0 * H2) - S1 * (S1 * H2 - S2 * H1));
}

__device__ __forceinline__ double inv_norm(double x, double y, double z) {
    return 1.0 / sqrt(x * x + y * y + z * z);   // LLGHeun.cpp:216, :269
}
__device__ __forceinline__ float inv_norm(float x, float y, float z) {
    return __builtin_amdgcn_rsqf(x * x + y * y + z * z);
}
// a / b through the hardware reciprocal (about 1 ulp) in FP32, where b is of order one; an exact
// division in FP64.
__device__ __forceinline__ float rcp_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double rcp_div(double a, double b) { return a / b; }

// ======================= fused kernels: two launches per Heun step ========================
// FP32 keeps nothing between the two kernels: its corrector rebuilds A = S + dt/2 dS from S and
// S' (S + dt dS = N S' with N = |S|^2 / (S . S'), dS being orthogonal to S), so the predictor
// writes S' alone.  FP64 keeps dS (or recomputes it) and forms S + dt/2 (dS + dS') in the
// reference's order, which is what makes its results equal to the reference bit for bit.
template <typename R> struct FoldEuler { static constexpr bool value = false; };
template <> struct FoldEuler<float> { static constexpr bool value = true; };

// Atoms per thread of the fused predictors: a workgroup of B threads takes FusedAtoms * B
// consecutive atoms of its block-table entry (thread t: t, t + B, ...).  The loads of every
// atom's own data are issued before the first atom is processed, so each wavefront keeps
// FusedAtoms times more HBM reads in flight.  The Heun corrector, with twice the own data per
// atom, takes one atom per thread from the one-atom table (measured faster on FCC 100^3).
template <typename R> struct FusedAtoms { static constexpr int value = 1; };
template <> struct FusedAtoms<float> { static constexpr int value = 2; };

template <typename R, int ET>
struct UniformJ {            // the exchange reals when every edge carries the same J
    static constexpr int K = (ET == ET_ISO) ? 1 : (ET == ET_VEC ? 3 : 9);
    R v[K];
    __device__ __forceinline__ UniformJ(const SysDev<R>& s, int jm) {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = jm == JM_UNIFORM ? s.jtab[k] : R(0);
    }
};

// Resolves a thread's atoms; false when it has none.  Atoms past the entry's end are replaced by
// its first atom (their loads are harmless) and flagged, so nothing is stored for them.  Every
// thread, with atoms or not, goes on to the workgroup's LDS staging barrier.
template <int U>
__device__ __forceinline__ bool fused_atoms(const int4& b, int (&idx)[U], bool (&ok)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        idx[u] = b.x + (int)threadIdx.x + u * (int)blockDim.x;
        ok[u] = idx[u] < b.z;
        if (!ok[u]) idx[u] = b.x;
    }
    return ok[0];
}

// LDS of the fused kernels, filled once per workgroup by every thread before any returns:
//   [material records [nmat]] [condition records of the workgroup's replica [nmat]]
//   [stencil offsets [npat][W], when the table is small (s.pat_lds)]
// Lanes then read their material, condition and row offsets from LDS (same-address reads are
// broadcast) instead of through the vector memory path, which the neighbour gathers keep busy.
extern __shared__ __align__(16) int vmx_lds[];
template <typename R>
struct FusedLds {
    static constexpr int kMatInts = (int)(sizeof(MatConst<R>) / sizeof(int));
    static constexpr int kCondInts = (int)(sizeof(Cond<R>) / sizeof(int));
    static __device__ __forceinline__ int cond_off(const SysDev<R>& s) { return s.nmat * kMatInts; }
    static __device__ __forceinline__ int pat_off(const SysDev<R>& s) {
        return ((s.nmat * (kMatInts + kCondInts)) + 3) & ~3;   // 16-byte aligned for int4 reads
    }
    static __device__ __forceinline__ const MatConst<R>* mat() { return reinterpret_cast<const MatConst<R>*>(vmx_lds); }
    static __device__ __forceinline__ const Cond<R>* cond(const SysDev<R>& s) {
        return reinterpret_cast<const Cond<R>*>(vmx_lds + cond_off(s));
    }
};
// bytes of dynamic LDS the fused kernels need (host and device agree on this layout)
template <typename R>
__host__ __device__ inline size_t fused_lds_bytes(int nmat, int npat, int zmax, int pat_lds) {
    const size_t rec = (size_t)nmat * ((sizeof(MatConst<R>) + sizeof(Cond<R>)) / sizeof(int));
    const size_t pat = pat_lds ? (size_t)npat * (size_t)zmax : 0;
    return (((rec + 3) & ~(size_t)3) + pat) * sizeof(int);
}
// Thermal unit normals of a thread's atoms, formed before the LDS tables are staged: they depend
// only on the atom and the step, so their arithmetic runs while the own-data loads are in flight.
template <typename R, int U>
__device__ __forceinline__ void early_normals(const SysDev<R>& s, int rep, const int4& b, const int (&idx)[U],
                                              uint64_t step, const double* __restrict__ noise, bool want,
                                              R (&gx)[U], R (&gy)[U], R (&gz)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        gx[u] = R(0); gy[u] = R(0); gz[u] = R(0);
        if (want && (s.fld & FLD_THERMAL))
            unit_normals(s, rep, idx[u], caller_local(s, rep, idx[u] - b.w), step, noise, gx[u], gy[u], gz[u]);
    }
}

template <typename R, int AM>
__device__ __forceinline__ void stage_fused(const SysDev<R>& s, int rep) {
    using L = FusedLds<R>;
    const int* __restrict__ mat = reinterpret_cast<const int*>(s.mat);
    const int* __restrict__ cond = reinterpret_cast<const int*>(s.cond + rep * s.nmat);
    const int nm = s.nmat * L::kMatInts, nc = s.nmat * L::kCondInts;
    const int np = (Access<AM>::TP == TP_PAT && s.pat_lds) ? s.npat * Access<AM>::W : 0;
    const int rec = nm + nc, base = L::pat_off(s), total = rec + np;
    // one pass over the three tables with two loads in flight per thread: every wait for these
    // loads also waits for the own-data loads issued before, so fewer rounds mean fewer waits
    auto src = [&](int k) { return k < nm ? mat[k] : (k < rec ? cond[k - nm] : s.pat_off[k - rec]); };
    auto dst = [&](int k) { return k < rec ? k : base + (k - rec); };
    const int bd = (int)blockDim.x;
    for (int k0 = (int)threadIdx.x; k0 < total; k0 += 2 * bd) {
        const int k1 = k0 + bd;
        const int v0 = src(k0);
        const int v1 = k1 < total ? src(k1) : 0;
        vmx_lds[dst(k0)] = v0;
        if (k1 < total) vmx_lds[dst(k1)] = v1;
    }
    __syncthreads();
}
template <typename R, int AM>
__device__ __forceinline__ void fused_row_index(const SysDev<R>& s, int i, int head, RowIdx<AM>& r) {
    if (Access<AM>::TP == TP_PAT && s.pat_lds)
        row_index(s, i, head, vmx_lds + FusedLds<R>::pat_off(s), r);
    else
        row_index(s, i, head, s.pat_off, r);
}

// Predictor: S' = |S + dt f(S, H_spin(S) + H_ext)|, written to P.  With D.x set the FP64 Euler
// derivative f(S, H) is kept for the corrector, as the reference keeps its euler arrays
// (src/simulate/LLGHeun.cpp:201-211); with E.x set H_ext is kept too, as the
// reference keeps its external-field arrays for both stages (LLGHeun.cpp:187-197, 228-237).
// A thread first issues the loads of its atoms' own data (spin, material, stencil id), which
// come from HBM, and forms their thermal normals while those loads are in flight; it then takes
// its atoms one at a time through row indices, neighbour gathers (mostly L2 hits) and the step.
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(512)
k_fused_predictor(SysDev<R> s, Spin3<R> S, Spin3<R> P, Spin3<R> D, Spin3<R> E, const double* __restrict__ noise,
                  uint32_t step_off) {
    constexpr int U = FusedAtoms<R>::value;
    constexpr bool kStaged = Access<AM>::TP != TP_CSR;
    const int lb = logical_block(s.xcd_remap);
    if (lb >= s.nblocks) return;                 // whole workgroup
    const int4 b = s.blk[lb];
    int idx[U];
    bool ok[U];
    fused_atoms<U>(b, idx, ok);
    const int rep = b.y;
    const SpinRsrc<R> rs(s, S.x, S.y, S.z);
    int m[U], head[U];
    R sx[U], sy[U], sz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {   // own data from HBM, in flight while the LDS tables are staged
        m[u] = s.nmat == 1 ? 0 : (int)s.type[idx[u]];   // one material: no per-atom load
        rs.load(idx[u], sx[u], sy[u], sz[u]);
        if (kStaged) head[u] = row_head<R, AM>(s, idx[u]);
    }
    const uint64_t step = *s.step_base + step_off;
    R gx[U], gy[U], gz[U];
    early_normals<R, U>(s, rep, b, idx, step, noise, true, gx, gy, gz);
    stage_fused<R, AM>(s, rep);
    if (!ok[0]) return;
    const UniformJ<R, ET> ju(s, Access<AM>::JM);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = idx[u];
        const MatConst<R>& mc = FusedLds<R>::mat()[m[u]];
        R fx, fy, fz;
        if constexpr (kStaged) {
            RowIdx<AM> ri;
            fused_row_index(s, i, head[u], ri);
            row_gather_field<R, ET, AM, 2>(s, rs, ri, ju.v, mc, sx[u], sy[u], sz[u], fx, fy, fz);
        } else {
            spin_field<R, ET, AM>(s, S.x, S.y, S.z, i, m[u], sx[u], sy[u], sz[u], fx, fy, fz);
        }
        const Cond<R>& c = FusedLds<R>::cond(s)[m[u]];
        R ex, ey, ez;
        ext_from_normals(s, c, i, gx[u], gy[u], gz[u], ex, ey, ez);
        R d0, d1, d2;
        llg_rhs(mc.one_opa2, mc.alpha_opa2, sx[u], sy[u], sz[u], fx + ex, fy + ey, fz + ez, d0, d1, d2);
        R n0 = sx[u] + d0 * s.dt, n1 = sy[u] + d1 * s.dt, n2 = sz[u] + d2 * s.dt;
        const R r = inv_norm(n0, n1, n2);
        if (ok[u]) {
            if constexpr (!FoldEuler<R>::value) {
                if (D.x) { D.x[i] = d0; D.y[i] = d1; D.z[i] = d2; }
            }
            if (E.x) {
                E.x[i] = ex; E.y[i] = ey; E.z[i] = ez;
            }
            P.x[i] = n0 * r; P.y[i] = n1 * r; P.z[i] = n2 * r;
        }
    }
}

// Corrector: dS' at S' with the same H_ext (read back when the predictor kept it, else
// regenerated from the same Philox counter), then S_next = |S + dt/2 (dS + dS')|
// (LLGHeun.cpp:257-269) with dS read back when the predictor kept it, else recomputed (FP64);
// FP32 rebuilds A = S + dt/2 dS from S and S' and forms |A + dt/2 dS'|.  S_next may overwrite S.
// Own data of all the thread's atoms first, as above.
// Parts the correctors gather their FP64 rows in (both rows of the recomputing variant too, so
// that the register budget of the default path is not set by a whole row of doubles).
template <typename R> struct CorrectorParts { static constexpr int value = 2; };
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(512)
k_fused_corrector(SysDev<R> s, Spin3Alias<R> S, Spin3<R> P, Spin3Alias<R> O, Spin3<R> D, Spin3<R> E,
                  const double* __restrict__ noise, uint32_t step_off) {
    constexpr int U = 1;                          // launched on the one-atom block table
    constexpr bool kStaged = Access<AM>::TP != TP_CSR;
    const int lb = logical_block(s.xcd_remap);
    if (lb >= s.nblocks) return;                 // whole workgroup
    const int4 b = s.blk[lb];
    int idx[U];
    bool ok[U];
    fused_atoms<U>(b, idx, ok);
    const int rep = b.y;
    constexpr bool fold = FoldEuler<R>::value;    // FP32: A rebuilt from S and S'
    const bool keep = !fold && D.x != nullptr;     // FP64: dS kept by the predictor
    const SpinRsrc<R> rp(s, P.x, P.y, P.z);
    int m[U], head[U];
    R px[U], py[U], pz[U];
    R ax[U], ay[U], az[U];          // S
    R dx[U], dy[U], dz[U];          // kept dS (FP64)
    R hx[U], hy[U], hz[U];          // kept H_ext
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = idx[u];
        m[u] = s.nmat == 1 ? 0 : (int)s.type[i];
        rp.load(i, px[u], py[u], pz[u]);
        if (kStaged) head[u] = row_head<R, AM>(s, i);
        ax[u] = S.x[i]; ay[u] = S.y[i]; az[u] = S.z[i];
        if (keep) { dx[u] = D.x[i]; dy[u] = D.y[i]; dz[u] = D.z[i]; }
        if (E.x) { hx[u] = E.x[i]; hy[u] = E.y[i]; hz[u] = E.z[i]; }
    }
    const uint64_t step = *s.step_base + step_off;
    R gx[U], gy[U], gz[U];                        // unless H_ext is read back
    early_normals<R, U>(s, rep, b, idx, step, noise, E.x == nullptr, gx, gy, gz);
    stage_fused<R, AM>(s, rep);      // own data above is in flight meanwhile
    if (!ok[0]) return;
    const UniformJ<R, ET> ju(s, Access<AM>::JM);
    constexpr int parts = CorrectorParts<R>::value;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = idx[u];
        const MatConst<R>& mc = FusedLds<R>::mat()[m[u]];
        RowIdx<AM> ri;
        R fx, fy, fz;
        if constexpr (kStaged) {
            fused_row_index(s, i, head[u], ri);
            row_gather_field<R, ET, AM, parts>(s, rp, ri, ju.v, mc, px[u], py[u], pz[u], fx, fy, fz);
        } else {
            spin_field<R, ET, AM>(s, P.x, P.y, P.z, i, m[u], px[u], py[u], pz[u], fx, fy, fz);
        }
        R ex, ey, ez;
        if (E.x) {
            ex = hx[u]; ey = hy[u]; ez = hz[u];
        } else {
            const Cond<R>& c = FusedLds<R>::cond(s)[m[u]];
            ext_from_normals(s, c, i, gx[u], gy[u], gz[u], ex, ey, ez);
        }
        R h0, h1, h2;
        llg_rhs(mc.one_opa2, mc.alpha_opa2, px[u], py[u], pz[u], fx + ex, fy + ey, fz + ez, h0, h1, h2);
        R n0, n1, n2;
        if (fold) {
            // A = S + dt/2 dS from S and S' = (S + dt dS) / |S + dt dS|: dS is orthogonal to S, so
            // the norm is |S|^2 / (S . S') and A = (S + N S') / 2
            const R N = rcp_div(ax[u] * ax[u] + ay[u] * ay[u] + az[u] * az[u],
                                ax[u] * px[u] + ay[u] * py[u] + az[u] * pz[u]);
            n0 = R(0.5) * (ax[u] + N * px[u]) + s.half_dt * h0;
            n1 = R(0.5) * (ay[u] + N * py[u]) + s.half_dt * h1;
            n2 = R(0.5) * (az[u] + N * pz[u]) + s.half_dt * h2;
        } else {
            R e0, e1, e2;
            if (keep) {
                e0 = dx[u]; e1 = dy[u]; e2 = dz[u];
            } else {
                R gx, gy, gz;
                if constexpr (kStaged) {
                    const SpinRsrc<R> rs(s, S.x, S.y, S.z);
                    row_gather_field<R, ET, AM, parts>(s, rs, ri, ju.v, mc, ax[u], ay[u], az[u], gx, gy, gz);
                } else {
                    spin_field<R, ET, AM>(s, S.x, S.y, S.z, i, m[u], ax[u], ay[u], az[u], gx, gy, gz);
                }
                llg_rhs(mc.one_opa2, mc.alpha_opa2, ax[u], ay[u], az[u], gx + ex, gy + ey, gz + ez, e0, e1, e2);
            }
            n0 = ax[u] + s.half_dt * (e0 + h0);
            n1 = ay[u] + s.half_dt * (e1 + h1);
            n2 = az[u] + s.half_dt * (e2 + h2);
        }
        const R r = inv_norm(n0, n1, n2);
        if (ok[u]) {
            O.x[i] = n0 * r; O.y[i] = n1 * r; O.z[i] = n2 * r;
        }
    }
}

// ======================= LLG midpoint: sim::LLG_Midpoint (src/simulate/LLGMidpoint.cpp) ======
// The step rotates a spin about F = H + alpha M x H in the Cayley form
//   S' = [S (1 - b^2 F.F) + 2 (b (F x S) + b^2 (F.S) F)] / (1 + b^2 F.F),  b = -dt one_opa2 / 2:
// first with M = S and the fields at S, keeping the midpoint (S + S')/2 unnormalised
// (LLGMidpoint.cpp:114-141), then with M the midpoint, the spin field at M and the same external
// field, applied to the initial S (:151-183).  Term order is the reference's.
template <typename R>
__device__ __forceinline__ void cayley(R alpha, R beta, R m0, R m1, R m2, R h0, R h1, R h2, R s0, R s1, R s2,
                                       R& o0, R& o1, R& o2) {
    const R beta2 = beta * beta;
    const R F0 = h0 + alpha * (m1 * h2 - m2 * h1);
    const R F1 = h1 + alpha * (m2 * h0 - m0 * h2);
    const R F2 = h2 + alpha * (m0 * h1 - m1 * h0);
    const R FdotF = F0 * F0 + F1 * F1 + F2 * F2;
    const R beta2FdotS = beta2 * (F0 * s0 + F1 * s1 + F2 * s2);
    const R one_o_one_plus_beta2FdotF = R(1) / (R(1) + beta2 * FdotF);
    const R one_minus_beta2FdotF = R(1) - beta2 * FdotF;
    o0 = one_o_one_plus_beta2FdotF * (s0 * one_minus_beta2FdotF + R(2) * (beta * (F1 * s2 - F2 * s1) + F0 * beta2FdotS));
    o1 = one_o_one_plus_beta2FdotF * (s1 * one_minus_beta2FdotF + R(2) * (beta * (F2 * s0 - F0 * s2) + F1 * beta2FdotS));
    o2 = one_o_one_plus_beta2FdotF * (s2 * one_minus_beta2FdotF + R(2) * (beta * (F0 * s1 - F1 * s0) + F2 * beta2FdotS));
}

template <typename R>
__device__ __forceinline__ R midpoint_beta(const SysDev<R>& s, const MatConst<R>& mc) {
    return R(-1) * s.dt * mc.one_opa2 * R(0.5);      // LLGMidpoint.cpp:118
}

// Midpoint predictor: fields at S, M = (S + S')/2 into P; H_ext into E when kept.  Same workgroup
// organisation as the fused Heun kernels (block table, own data first, LDS tables).
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(512)
k_mid_predictor(SysDev<R> s, Spin3<R> S, Spin3<R> P, Spin3<R> E, const double* __restrict__ noise,
                uint32_t step_off) {
    constexpr int U = FusedAtoms<R>::value;
    constexpr bool kStaged = Access<AM>::TP != TP_CSR;
    const int lb = logical_block(s.xcd_remap);
    if (lb >= s.nblocks) return;
    const int4 b = s.blk[lb];
    int idx[U];
    bool ok[U];
    fused_atoms<U>(b, idx, ok);
    const int rep = b.y;
    const SpinRsrc<R> rs(s, S.x, S.y, S.z);
    int m[U], head[U];
    R sx[U], sy[U], sz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        m[u] = s.nmat == 1 ? 0 : (int)s.type[idx[u]];
        rs.load(idx[u], sx[u], sy[u], sz[u]);
        if (kStaged) head[u] = row_head<R, AM>(s, idx[u]);
    }
    const uint64_t step = *s.step_base + step_off;
    R gx[U], gy[U], gz[U];
    early_normals<R, U>(s, rep, b, idx, step, noise, true, gx, gy, gz);
    stage_fused<R, AM>(s, rep);
    if (!ok[0]) return;
    const UniformJ<R, ET> ju(s, Access<AM>::JM);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = idx[u];
        const MatConst<R>& mc = FusedLds<R>::mat()[m[u]];
        R fx, fy, fz;
        if constexpr (kStaged) {
            RowIdx<AM> ri;
            fused_row_index(s, i, head[u], ri);
            row_gather_field<R, ET, AM, 2>(s, rs, ri, ju.v, mc, sx[u], sy[u], sz[u], fx, fy, fz);
        } else {
            spin_field<R, ET, AM>(s, S.x, S.y, S.z, i, m[u], sx[u], sy[u], sz[u], fx, fy, fz);
        }
        const Cond<R>& c = FusedLds<R>::cond(s)[m[u]];
        R ex, ey, ez;
        ext_from_normals(s, c, i, gx[u], gy[u], gz[u], ex, ey, ez);
        R o0, o1, o2;
        cayley(mc.alpha, midpoint_beta(s, mc), sx[u], sy[u], sz[u], fx + ex, fy + ey, fz + ez, sx[u], sy[u], sz[u],
               o0, o1, o2);
        if (ok[u]) {
            if (E.x) {
                E.x[i] = ex; E.y[i] = ey; E.z[i] = ez;
            }
            P.x[i] = (sx[u] + o0) * R(0.5); P.y[i] = (sy[u] + o1) * R(0.5); P.z[i] = (sz[u] + o2) * R(0.5);
        }
    }
}

// Midpoint corrector: spin field at M (P), the same H_ext (kept in E, or regenerated from the
// same Philox counter), the rotation applied to the initial S, written to O.  One atom per
// thread, like the Heun corrector, which loads as much own data.
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(512)
k_mid_corrector(SysDev<R> s, Spin3Alias<R> S, Spin3<R> P, Spin3Alias<R> O, Spin3<R> E, const double* __restrict__ noise,
                uint32_t step_off) {
    constexpr int U = 1;                          // launched on the one-atom block table
    constexpr bool kStaged = Access<AM>::TP != TP_CSR;
    const int lb = logical_block(s.xcd_remap);
    if (lb >= s.nblocks) return;
    const int4 b = s.blk[lb];
    int idx[U];
    bool ok[U];
    fused_atoms<U>(b, idx, ok);
    const int rep = b.y;
    const SpinRsrc<R> rp(s, P.x, P.y, P.z);
    int m[U], head[U];
    R px[U], py[U], pz[U], ax[U], ay[U], az[U], hx[U], hy[U], hz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = idx[u];
        m[u] = s.nmat == 1 ? 0 : (int)s.type[i];
        rp.load(i, px[u], py[u], pz[u]);
        if (kStaged) head[u] = row_head<R, AM>(s, i);
        ax[u] = S.x[i]; ay[u] = S.y[i]; az[u] = S.z[i];
        if (E.x) { hx[u] = E.x[i]; hy[u] = E.y[i]; hz[u] = E.z[i]; }
    }
    const uint64_t step = *s.step_base + step_off;
    R gx[U], gy[U], gz[U];                        // unless H_ext is read back
    early_normals<R, U>(s, rep, b, idx, step, noise, E.x == nullptr, gx, gy, gz);
    stage_fused<R, AM>(s, rep);
    if (!ok[0]) return;
    const UniformJ<R, ET> ju(s, Access<AM>::JM);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int i = idx[u];
        const MatConst<R>& mc = FusedLds<R>::mat()[m[u]];
        R fx, fy, fz;
        if constexpr (kStaged) {
            RowIdx<AM> ri;
            fused_row_index(s, i, head[u], ri);
            row_gather_field<R, ET, AM, CorrectorParts<R>::value>(s, rp, ri, ju.v, mc, px[u], py[u], pz[u], fx, fy,
                                                                  fz);
        } else {
            spin_field<R, ET, AM>(s, P.x, P.y, P.z, i, m[u], px[u], py[u], pz[u], fx, fy, fz);
        }
        R ex, ey, ez;
        if (E.x) {
            ex = hx[u]; ey = hy[u]; ez = hz[u];
        } else {
            const Cond<R>& c = FusedLds<R>::cond(s)[m[u]];
            ext_from_normals(s, c, i, gx[u], gy[u], gz[u], ex, ey, ez);
        }
        R o0, o1, o2;
        cayley(mc.alpha, midpoint_beta(s, mc), px[u], py[u], pz[u], fx + ex, fy + ey, fz + ez, ax[u], ay[u], az[u],
               o0, o1, o2);
        if (ok[u]) {
            O.x[i] = o0; O.y[i] = o1; O.z[i] = o2;
        }
    }
}

// ======================= plain kernels: the reference dataflow ==============================
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(512)
k_plain_spin_fields(SysDev<R> s, Spin3<R> S, Spin3<R> F) {
    int i, rep, local;
    if (!my_atom(s, i, rep, local)) return;
    const int m = s.type[i];
    R fx, fy, fz;
    spin_field<R, ET, AM>(s, S.x, S.y, S.z, i, m, S.x[i], S.y[i], S.z[i], fx, fy, fz);
    F.x[i] = fx; F.y[i] = fy; F.z[i] = fz;
}

template <typename R>
__global__ void __launch_bounds__(512)
k_plain_ext_fields(SysDev<R> s, Spin3<R> E, const double* __restrict__ noise, uint32_t step_off) {
    int i, rep, local;
    if (!my_atom(s, i, rep, local)) return;
    const uint64_t step = *s.step_base + step_off;
    R ex, ey, ez;
    ext_field(s, rep, (int)s.type[i], i, local, step, noise, ex, ey, ez);
    E.x[i] = ex; E.y[i] = ey; E.z[i] = ez;
}

// Euler predictor (LLGHeun.cpp:188-226): stores dS in D and S' in P.
template <typename R>
__global__ void __launch_bounds__(512)
k_plain_predictor(SysDev<R> s, Spin3<R> S, Spin3<R> F, Spin3<R> E, Spin3<R> D, Spin3<R> P) {
    int i, rep, local;
    if (!my_atom(s, i, rep, local)) return;
    const MatConst<R>& mc = s.mat[s.type[i]];
    const R sx = S.x[i], sy = S.y[i], sz = S.z[i];
    R d0, d1, d2;
    llg_rhs(mc.one_opa2, mc.alpha_opa2, sx, sy, sz, F.x[i] + E.x[i], F.y[i] + E.y[i], F.z[i] + E.z[i],
            d0, d1, d2);
    D.x[i] = d0; D.y[i] = d1; D.z[i] = d2;
    R n0 = sx + d0 * s.dt, n1 = sy + d1 * s.dt, n2 = sz + d2 * s.dt;
    const R r = inv_norm(n0, n1, n2);
    P.x[i] = n0 * r; P.y[i] = n1 * r; P.z[i] = n2 * r;
}

// Heun corrector (LLGHeun.cpp:239-279): S <- |S0 + dt/2 (dS + dS')| in place.
template <typename R>
__global__ void __launch_bounds__(512)
k_plain_corrector(SysDev<R> s, Spin3<R> S, Spin3<R> P, Spin3<R> F, Spin3<R> E, Spin3<R> D) {
    int i, rep, local;
    if (!my_atom(s, i, rep, local)) return;
    const MatConst<R>& mc = s.mat[s.type[i]];
    const R px = P.x[i], py = P.y[i], pz = P.z[i];
    R h0, h1, h2;
    llg_rhs(mc.one_opa2, mc.alpha_opa2, px, py, pz, F.x[i] + E.x[i], F.y[i] + E.y[i], F.z[i] + E.z[i],
            h0, h1, h2);
    R n0 = S.x[i] + s.half_dt * (D.x[i] + h0);
    R n1 = S.y[i] + s.half_dt * (D.y[i] + h1);
    R n2 = S.z[i] + s.half_dt * (D.z[i] + h2);
    const R r = inv_norm(n0, n1, n2);
    S.x[i] = n0 * r; S.y[i] = n1 * r; S.z[i] = n2 * r;
}

// ======================= statistics ==========================================================
// Per-mask sums in FP64.  Each thread keeps a register accumulator for the mask id of its
// current run of atoms and flushes it into the workgroup's LDS histogram when the id changes,
// so a system-wide mask (one id) costs one LDS atomic per thread.  Each workgroup writes its
// histogram to partial[block][nbins*W]; k_stats_finish reduces the partials in block order.

// Fixed-order reduction of the threads' register histograms acc[MAXB][4] (bins < nbins <= MAXB)
// into partial[block][nbins*4]: wavefront shuffles, then the wavefronts in order.  Workgroups of at
// most 256 threads; every thread of the workgroup calls it.
template <int MAXB>
__device__ __forceinline__ void reduce_hist_regs(const double (&acc)[MAXB][4], int nbins,
                                                 double* __restrict__ partial) {
    __shared__ double red[4][4 * MAXB];            // [wavefront][bin * 4 + component]
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double v = acc[b][c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) red[w][4 * b + c] = v;
        }
    }
    __syncthreads();
    const int W = 4 * nbins;
    if ((int)threadIdx.x < W) {
        const int nw = ((int)blockDim.x + 63) >> 6;
        double t = 0.0;
        for (int q = 0; q < nw; ++q) t += red[q][threadIdx.x];
        partial[(size_t)blockIdx.x * W + threadIdx.x] = t;
    }
}

// Each atom's moment (mu_B) for the statistics: per atom, or, when every atom's moment is its
// material's (as VAMPIRE builds them), the material's from a small table through the 1-byte
// material id: the same doubles with 1 byte read per atom instead of 8.
struct Moments {
    const double* __restrict__ mm;
    const uint8_t* __restrict__ type;
    const double* __restrict__ tab;
    __device__ __forceinline__ double at(int i) const { return tab ? tab[type[i]] : mm[i]; }
};

// Masks of at most MAXB bins: each thread keeps every bin's four sums in registers and adds each
// atom to its bin with predicated adds, so the cost does not depend on how often the id changes
// along the atoms (a random alloy's material mask changes at nearly every atom).  The workgroup
// then reduces its threads' sums in a fixed order (wavefront shuffles, then the wavefronts in
// order) into partial[block][nbins*4], the layout k_mag_partial writes.
template <typename R, int MAXB>
__global__ void __launch_bounds__(256)
k_mag_partial_regs(int n, Spin3<R> S, const int* __restrict__ mask, Moments mo, int nbins,
                   double* __restrict__ partial) {
    double acc[MAXB][4];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) { acc[b][0] = 0.0; acc[b][1] = 0.0; acc[b][2] = 0.0; acc[b][3] = 0.0; }
    // with up to four bins, four of the thread's atoms are loaded before any is added (then added
    // in the same order as one at a time), so each wavefront has four atoms' loads in flight; with
    // eight the histogram's 64 registers leave no room, and atoms go one at a time
    constexpr int U = MAXB >= 8 ? 1 : 4;
    auto add = [&](int id, double m, R x, R y, R z) {
        const double v0 = (double)x * m, v1 = (double)y * m, v2 = (double)z * m;
#pragma unroll
        for (int b = 0; b < MAXB; ++b) {
            const bool hit = id == b;
            acc[b][0] += hit ? v0 : 0.0;
            acc[b][1] += hit ? v1 : 0.0;
            acc[b][2] += hit ? v2 : 0.0;
            acc[b][3] += hit ? m : 0.0;
        }
    };
    if constexpr (U == 1) {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            add(MAXB == 1 ? 0 : mask[i], mo.at(i), S.x[i], S.y[i], S.z[i]);
    } else {
        const int stride = (int)(gridDim.x * blockDim.x);
        int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
        for (; (long long)i + (U - 1) * (long long)stride < n; i += U * stride) {
            int id[U];
            double m[U];
            R x[U], y[U], z[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = i + u * stride;
                id[u] = MAXB == 1 ? 0 : mask[k];   // a one-bin mask holds 0 everywhere: not read
                m[u] = mo.at(k);
                x[u] = S.x[k]; y[u] = S.y[k]; z[u] = S.z[k];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) add(id[u], m[u], x[u], y[u], z[u]);
        }
        for (; i < n; i += stride) add(MAXB == 1 ? 0 : mask[i], mo.at(i), S.x[i], S.y[i], S.z[i]);
    }
    reduce_hist_regs<MAXB>(acc, nbins, partial);
}

// W = 4: sum(m s_x), sum(m s_y), sum(m s_z), sum(m)   (magnetization.cpp:141-147)
template <typename R>
__global__ void __launch_bounds__(256)
k_mag_partial(int n, Spin3<R> S, const int* __restrict__ mask, Moments mo, int nbins,
              double* __restrict__ partial) {
    extern __shared__ double hist[];
    for (int k = threadIdx.x; k < 4 * nbins; k += blockDim.x) hist[k] = 0.0;
    __syncthreads();
    int cur = -1;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int id = mask[i];
        if (id != cur) {
            if (cur >= 0) {
                atomicAdd(&hist[4 * cur + 0], a0); atomicAdd(&hist[4 * cur + 1], a1);
                atomicAdd(&hist[4 * cur + 2], a2); atomicAdd(&hist[4 * cur + 3], a3);
            }
            cur = id; a0 = a1 = a2 = a3 = 0.0;
        }
        const double m = mo.at(i);
        a0 += (double)S.x[i] * m;
        a1 += (double)S.y[i] * m;
        a2 += (double)S.z[i] * m;
        a3 += m;
    }
    if (cur >= 0) {
        atomicAdd(&hist[4 * cur + 0], a0); atomicAdd(&hist[4 * cur + 1], a1);
        atomicAdd(&hist[4 * cur + 2], a2); atomicAdd(&hist[4 * cur + 3], a3);
    }
    __syncthreads();
    double* out = partial + (size_t)blockIdx.x * 4 * nbins;
    for (int k = threadIdx.x; k < 4 * nbins; k += blockDim.x) out[k] = hist[k];
}

// Single-spin energies (tesla): exchange (src/exchange/energy.cpp:27-104), anisotropy
// (src/anisotropy/energy.cpp:31-50 with uniaxial_second_order.cpp:104-121,
// cubic_fourth_order.cpp:80-95, cubic_fourth_order_rotation.cpp:104-140) and applied field
// (src/simulate/energy.cpp:94-98).
// One neighbour's term of the exchange energy of spin (sx, sy, sz), subtracted from `energy`.
template <typename R, int ET>
__device__ __forceinline__ void exchange_energy_term(const R* J, R Sx, R Sy, R Sz, R sx, R sy, R sz, R& energy) {
    if (ET == ET_ISO) {
        energy -= J[0] * (Sx * sx + Sy * sy + Sz * sz);
    } else if (ET == ET_VEC) {
        energy -= (J[0] * Sx * sx + J[1] * Sy * sy + J[2] * Sz * sz);
    } else {
        energy -= (J[0] * Sx * sx + J[1] * Sy * sx + J[2] * Sz * sx +
                   J[3] * Sx * sy + J[4] * Sy * sy + J[5] * Sz * sy +
                   J[6] * Sx * sz + J[7] * Sy * sz + J[8] * Sz * sz);
    }
}

template <typename R, int ET, int AM>
__device__ __forceinline__ R exchange_energy(const SysDev<R>& s, const R* __restrict__ X,
                                             const R* __restrict__ Y, const R* __restrict__ Z,
                                             int i, R sx, R sy, R sz) {
    R energy = R(0);
    for_neighbours<R, ET, AM>(s, X, Y, Z, i, [&](const R Sx, const R Sy, const R Sz, const R* J) {
        exchange_energy_term<R, ET>(J, Sx, Sy, Sz, sx, sy, sz, energy);
    });
    return energy;
}

// The same energy from a staged row (neighbour spins already in registers), in row order.
template <typename R, int ET, int AM>
__device__ __forceinline__ R row_exchange_energy(const SysDev<R>& s, const RowIdx<AM>& r, const RowSpins<R, AM>& v,
                                                 const R* ju, R sx, R sy, R sz) {
    R energy = R(0);
    row_sum<R, ET, AM>(s, r, v, ju, [&](const R Sx, const R Sy, const R Sz, const R* J) {
        exchange_energy_term<R, ET>(J, Sx, Sy, Sz, sx, sy, sz, energy);
    });
    return energy;
}

// The exchange energies of a staged row before (o) and after (n) a move, each summed in row order.
// FP64 rows are gathered in Parts parts (a part's loads, then its terms of both sums): the same
// terms in the same order, with a part's neighbour values live at once instead of the whole row's.
template <typename R, int ET, int AM, int Parts>
__device__ __forceinline__ void row_energy_pair(const SysDev<R>& s, const SpinRsrc<R>& rs, const RowIdx<AM>& r,
                                                const R* ju, R ox, R oy, R oz, R nx, R ny, R nz, R& e_old,
                                                R& e_new) {
    constexpr int K = (ET == ET_ISO) ? 1 : (ET == ET_VEC ? 3 : 9);
    constexpr int JM = Access<AM>::JM, W = Access<AM>::W, H = (W + Parts - 1) / Parts;
    e_old = R(0);
    e_new = R(0);
#pragma unroll
    for (int part = 0; part < Parts; ++part) {
        R vx[H], vy[H], vz[H];
#pragma unroll
        for (int k = 0; k < H; ++k)
            if (part * H + k < W) rs.load(r.j[part * H + k], vx[k], vy[k], vz[k]);
#pragma unroll
        for (int k = 0; k < H; ++k) {
            const int slot = part * H + k;
            if (slot < W && (s.full_rows || slot < r.deg)) {
                const R* J = JM == JM_UNIFORM ? ju : s.jtab + (size_t)K * r.jd[slot];
                exchange_energy_term<R, ET>(J, vx[k], vy[k], vz[k], ox, oy, oz, e_old);
                exchange_energy_term<R, ET>(J, vx[k], vy[k], vz[k], nx, ny, nz, e_new);
            }
        }
        if (part + 1 < Parts) __builtin_amdgcn_sched_barrier(0);   // the next part's loads stay here
    }
}

// Anisotropy energy with the material record mc (from LDS or memory).
template <typename R>
__device__ __forceinline__ R anisotropy_energy(uint32_t ani, const MatConst<R>& mc, R sx, R sy, R sz) {
    R energy = R(0);
    if (!ani) return energy;
    if (ani & ANI_UNI2) {
        const R sdote = (sx * mc.ku[0] + sy * mc.ku[1] + sz * mc.ku[2]);
        energy += -mc.ku2 * (sdote * sdote);
    }
    if (ani & ANI_CUB4) {
        const R sx4 = sx * sx * sx * sx, sy4 = sy * sy * sy * sy, sz4 = sz * sz * sz * sz;
        energy += R(-0.5) * mc.kc4 * (sx4 + sy4 + sz4);
    }
    if (ani & ANI_CUB4R) {
        const R d1 = sx * mc.e1[0] + sy * mc.e1[1] + sz * mc.e1[2];
        const R d2 = sx * mc.e2[0] + sy * mc.e2[1] + sz * mc.e2[2];
        const R d3 = sx * mc.e3[0] + sy * mc.e3[1] + sz * mc.e3[2];
        const R d1f = d1 * d1 * d1 * d1, d2f = d2 * d2 * d2 * d2, d3f = d3 * d3 * d3 * d3;
        energy += R(-0.5) * mc.kc4 * (d1f + d2f + d3f);
    }
    return energy;
}

template <typename R>
__device__ __forceinline__ R anisotropy_energy(const SysDev<R>& s, int m, R sx, R sy, R sz) {
    if (!s.ani) return R(0);
    return anisotropy_energy(s.ani, s.mat[m], sx, sy, sz);
}

template <typename R>
__device__ __forceinline__ R applied_energy(const Cond<R>& c, R sx, R sy, R sz) {
    return -(c.hglob[0] * sx + c.hglob[1] * sy + c.hglob[2] * sz);
}

// Magnetostatic single-spin energy -mu0Hd . S (src/dipole/energy.cpp; simulate/energy.cpp:123-125)
template <typename R>
__device__ __forceinline__ R magnetostatic_energy(const SysDev<R>& s, int i, R sx, R sy, R sz) {
    const size_t n = (size_t)s.n;
    return R(-1.0) * ((R)s.hd[i] * sx + (R)s.hd[n + i] * sy + (R)s.hd[2 * n + i] * sz);
}

// W = 4: exchange, anisotropy, applied, magnetostatic, each times m_i (energy.cpp:155-203); the
// last is zero without a device dipole field
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(256)
k_energy_partial(SysDev<R> s, Spin3<R> S, const int* __restrict__ mask, Moments mo,
                 const int* __restrict__ atom_rep, int nbins, double* __restrict__ partial) {
    if (nbins <= 4) {                // few bins: register histograms, as k_mag_partial_regs
        double acc[4][4];
#pragma unroll
        for (int b = 0; b < 4; ++b) { acc[b][0] = 0.0; acc[b][1] = 0.0; acc[b][2] = 0.0; acc[b][3] = 0.0; }
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
            const int id = mask[i];
            const R sx = S.x[i], sy = S.y[i], sz = S.z[i];
            const int m = s.type[i];
            const int rep = atom_rep ? atom_rep[i] : 0;
            const double w = mo.at(i);
            const double e0 = (double)exchange_energy<R, ET, AM>(s, S.x, S.y, S.z, i, sx, sy, sz) * w;
            const double e1 = (double)anisotropy_energy(s, m, sx, sy, sz) * w;
            const double e2 = (double)applied_energy(s.cond[rep * s.nmat + m], sx, sy, sz) * w;
            const double e3 = s.hd ? (double)magnetostatic_energy(s, i, sx, sy, sz) * w : 0.0;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const bool hit = id == b;
                acc[b][0] += hit ? e0 : 0.0;
                acc[b][1] += hit ? e1 : 0.0;
                acc[b][2] += hit ? e2 : 0.0;
                acc[b][3] += hit ? e3 : 0.0;
            }
        }
        reduce_hist_regs<4>(acc, nbins, partial);
        return;
    }
    extern __shared__ double hist[];
    for (int k = threadIdx.x; k < 4 * nbins; k += blockDim.x) hist[k] = 0.0;
    __syncthreads();
    int cur = -1;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
        const int id = mask[i];
        if (id != cur) {
            if (cur >= 0) {
                atomicAdd(&hist[4 * cur + 0], a0); atomicAdd(&hist[4 * cur + 1], a1);
                atomicAdd(&hist[4 * cur + 2], a2); atomicAdd(&hist[4 * cur + 3], a3);
            }
            cur = id; a0 = a1 = a2 = a3 = 0.0;
        }
        const R sx = S.x[i], sy = S.y[i], sz = S.z[i];
        const int m = s.type[i];
        const int rep = atom_rep ? atom_rep[i] : 0;
        const double w = mo.at(i);
        a0 += (double)exchange_energy<R, ET, AM>(s, S.x, S.y, S.z, i, sx, sy, sz) * w;
        a1 += (double)anisotropy_energy(s, m, sx, sy, sz) * w;
        a2 += (double)applied_energy(s.cond[rep * s.nmat + m], sx, sy, sz) * w;
        if (s.hd) a3 += (double)magnetostatic_energy(s, i, sx, sy, sz) * w;
    }
    if (cur >= 0) {
        atomicAdd(&hist[4 * cur + 0], a0); atomicAdd(&hist[4 * cur + 1], a1);
        atomicAdd(&hist[4 * cur + 2], a2); atomicAdd(&hist[4 * cur + 3], a3);
    }
    __syncthreads();
    double* out = partial + (size_t)blockIdx.x * 4 * nbins;
    for (int k = threadIdx.x; k < 4 * nbins; k += blockDim.x) out[k] = hist[k];
}

// ======================= checkerboard Metropolis =============================================
// One launch per colour.  Atoms of a colour have no neighbour of the same colour, so each move
// sees fixed neighbours and the updates of one launch commute.  Move and acceptance follow
// montecarlo::mc_step (src/montecarlo/mc.cpp:59-108) with the adaptive cone of
// mc_moves.cpp:71-87,158-162.

// Rejection counters per replica: a workgroup adds its count to slot (workgroup mod
// kRejectSlots) of its replica, each slot on a cache line of its own (kRejectStride doubles), so
// the thousands of workgroups of one large replica do not all update one line; k_mc_adapt sums
// the slots.
constexpr int kRejectSlots = 64;
constexpr int kRejectStride = 16;

// The workgroup's rejection count (LDS)
__device__ __forceinline__ unsigned int& mc_wg_rejects() {
    __shared__ unsigned int count;
    return count;
}

// Acceptance of a move that raises the energy: exp(-dE mu_B/(kB T)) >= u, u = word / 2^32
// (mc.cpp:93-99).  FP64 in double as the oracle; FP32 on the hardware exp2.
__device__ __forceinline__ bool mc_accept(double x, uint32_t word) {
    return exp(-x) >= (double)word * 2.3283064365386963e-10;
}
__device__ __forceinline__ bool mc_accept(float x, uint32_t word) {
    return __expf(-x) >= (float)word * 2.3283064365386963e-10f;
}

// Atoms per thread of the Metropolis colour kernel: a workgroup takes 256 * McAtoms list entries.
// One atom per thread measured faster than two in both precisions (FP32: 4 % in colour order, 2 %
// in storage order, from the extra resident wavefronts; FP64: 96 VGPRs at two).
template <typename R> struct McAtoms { static constexpr int value = 1; };

// The moves of a thread's U atoms idx[] of replica rep (ok[u]: a real entry; head[u]: its stencil
// id; local[u]: its caller index within the replica, the Philox counter).  The thread forms the
// moves from Philox while its own loads are in flight, gathers the neighbours of all its atoms
// before it stores anything (atoms of one colour do not interact, so no gather can see another
// lane's store), and forms each energy before and after the move from the same registers, in the
// oracle's order.  Every thread of the workgroup reaches the LDS staging and the final barrier.
template <typename R, int ET, int AM, int U>
__device__ __forceinline__ void mc_moves(const SysDev<R>& s, const Spin3<R>& S, int rep, int lb, const int (&idx)[U],
                                         const bool (&ok)[U], const int (&head)[U], const int (&local)[U],
                                         const double* __restrict__ sigma, uint64_t sweep,
                                         double* __restrict__ rejects) {
    constexpr bool kStaged = Access<AM>::TP != TP_CSR;
    int m[U];
    R ox[U], oy[U], oz[U];
    const SpinRsrc<R> rs(s, S.x, S.y, S.z);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        m[u] = s.nmat == 1 ? 0 : (int)s.type[idx[u]];
        rs.load(idx[u], ox[u], oy[u], oz[u]);
    }
    const uint2 key = s.rep_key[rep];
    R gx[U], gy[U], gz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const U4 um = philox4x32_10((uint32_t)local[u], (uint32_t)sweep, (uint32_t)(sweep >> 32), STREAM_MC_MOVE,
                                    key.x, key.y);
        box_muller3(um, gx[u], gy[u], gz[u]);
    }
    const R angle = (R)sigma[rep];
    if (threadIdx.x == 0) mc_wg_rejects() = 0u;
    stage_fused<R, AM>(s, rep);
    const UniformJ<R, ET> ju(s, Access<AM>::JM);
    R nx[U], ny[U], nz[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        nx[u] = ox[u] + gx[u] * angle; ny[u] = oy[u] + gy[u] * angle; nz[u] = oz[u] + gz[u] * angle;
        const R r = inv_norm(nx[u], ny[u], nz[u]);
        nx[u] *= r; ny[u] *= r; nz[u] *= r;
    }
    R x_old[U], x_new[U];
    if constexpr (kStaged && sizeof(R) == 8) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            RowIdx<AM> ri;
            fused_row_index(s, idx[u], head[u], ri);
            row_energy_pair<R, ET, AM, 2>(s, rs, ri, ju.v, ox[u], oy[u], oz[u], nx[u], ny[u], nz[u], x_old[u],
                                          x_new[u]);
        }
    } else if constexpr (kStaged) {
        RowIdx<AM> ri[U];
        RowSpins<R, AM> rv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            fused_row_index(s, idx[u], head[u], ri[u]);
            row_load(rs, ri[u], rv[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            x_old[u] = row_exchange_energy<R, ET, AM>(s, ri[u], rv[u], ju.v, ox[u], oy[u], oz[u]);
            x_new[u] = row_exchange_energy<R, ET, AM>(s, ri[u], rv[u], ju.v, nx[u], ny[u], nz[u]);
        }
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            x_old[u] = exchange_energy<R, ET, AM>(s, S.x, S.y, S.z, idx[u], ox[u], oy[u], oz[u]);
            x_new[u] = exchange_energy<R, ET, AM>(s, S.x, S.y, S.z, idx[u], nx[u], ny[u], nz[u]);
        }
    }
    bool accept[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const MatConst<R>& mc = FusedLds<R>::mat()[m[u]];
        const Cond<R>& c = FusedLds<R>::cond(s)[m[u]];
        // sim::calculate_spin_energy order: exchange, anisotropy, applied, magnetostatic
        // (energy.cpp:147-174)
        R e_old = (R(0) + x_old[u] + anisotropy_energy(s.ani, mc, ox[u], oy[u], oz[u])) +
                  applied_energy(c, ox[u], oy[u], oz[u]);
        R e_new = (R(0) + x_new[u] + anisotropy_energy(s.ani, mc, nx[u], ny[u], nz[u])) +
                  applied_energy(c, nx[u], ny[u], nz[u]);
        if (s.hd) {
            e_old += magnetostatic_energy(s, idx[u], ox[u], oy[u], oz[u]);
            e_new += magnetostatic_energy(s, idx[u], nx[u], ny[u], nz[u]);
        }
        const R de = (e_new - e_old) * mc.mu_s_SI * R(1.07828231e23);
        accept[u] = de < R(0);
        if (!accept[u]) {
            const U4 ua = philox4x32_10((uint32_t)local[u], (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                        STREAM_MC_ACCEPT, key.x, key.y);
            accept[u] = mc_accept(de * c.kBTBohr, ua.v[0]);
        }
    }
    unsigned int wave_rejects = 0u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (ok[u] && accept[u]) {
            S.x[idx[u]] = nx[u]; S.y[idx[u]] = ny[u]; S.z[idx[u]] = nz[u];
        }
        wave_rejects += (unsigned int)__popcll(__ballot(ok[u] && !accept[u]));
    }
    // rejections of the replica: wavefront counts summed in LDS, then one atomic per workgroup
    // (counts are integers, exact in a double in any order)
    if (wave_rejects != 0u && __lane_id() == 0u) atomicAdd(&mc_wg_rejects(), wave_rejects);
    __syncthreads();
    if (threadIdx.x == 0 && mc_wg_rejects() != 0u)
        atomicAdd(&rejects[((size_t)rep * kRejectSlots + (lb & (kRejectSlots - 1))) * kRejectStride],
                  (double)mc_wg_rejects());
}

// Storage-order launch: workgroup w of a sweep takes U * 256 entries (U = McAtoms) of the flat
// launch arrays: atoms of one colour of replica wg[w].x (whose first atom is wg[w].y) and, with
// stencil rows, their stencil ids.  The replica is uniform, its conditions and cone width are read
// once, and the tables are staged in LDS as in the fused LLG kernels.  A thread's entries are t,
// t + 256, ...: their atom and stencil id arrive together, the atom's spin and its stencil offsets
// next, its neighbours last.  Entries past a run's end hold ~(the run's first atom): they load
// harmlessly and store nothing.
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(256)
k_mc_colour(SysDev<R> s, Spin3<R> S, const int* __restrict__ atom, const uint16_t* __restrict__ pat,
            const int2* __restrict__ wg, int wg0, int nwg, const double* __restrict__ sigma, uint64_t sweep,
            double* __restrict__ rejects) {
    constexpr int U = McAtoms<R>::value;
    const int lb = logical_block(s.xcd_remap);
    if (lb >= nwg) return;                         // whole workgroup
    const size_t base = (size_t)(wg0 + lb) * (256 * U) + threadIdx.x;
    int idx[U], head[U], local[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int a = atom[base + u * 256];
        ok[u] = a >= 0;
        idx[u] = ok[u] ? a : ~a;
        head[u] = Access<AM>::TP == TP_PAT ? (int)pat[base + u * 256] : 0;
    }
    const int2 b = wg[wg0 + lb];
#pragma unroll
    for (int u = 0; u < U; ++u) local[u] = caller_local(s, b.x, idx[u] - b.y);
    mc_moves<R, ET, AM, U>(s, S, b.x, lb, idx, ok, head, local, sigma, sweep, rejects);
}

// Colour-order launch: the spins, stencil ids and material ids of s and S are in colour order
// (each replica's atoms colour by colour, storage order within a colour), so workgroup w takes
// the contiguous atoms [wg[w].x, wg[w].z) of replica wg[w].y, all of one colour, and their
// neighbours' loads are contiguous spans too.  lidx: each atom's caller index within its replica.
// Same moves, counters and arithmetic as the storage-order launch.
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(256)
k_mc_colour_ordered(SysDev<R> s, Spin3<R> S, const int* __restrict__ lidx, const int4* __restrict__ wg, int wg0,
                    int nwg, const double* __restrict__ sigma, uint64_t sweep, double* __restrict__ rejects) {
    constexpr int U = McAtoms<R>::value;
    const int lb = logical_block(s.xcd_remap);
    if (lb >= nwg) return;                         // whole workgroup
    const int4 b = wg[wg0 + lb];
    int idx[U], head[U], local[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int t = b.x + (int)threadIdx.x + u * 256;
        ok[u] = t < b.z;
        idx[u] = ok[u] ? t : b.x;
        head[u] = row_head<R, AM>(s, idx[u]);
        local[u] = lidx[idx[u]];
    }
    mc_moves<R, ET, AM, U>(s, S, b.y, lb, idx, ok, head, local, sigma, sweep, rejects);
}

// Spins between storage order and colour order: dst[p] = src[perm[p]] (to colour order) or
// dst[perm[p]] = src[p] (back).
template <typename R>
__global__ void k_mc_permute(int n, const int* __restrict__ perm, Spin3<R> src, Spin3<R> dst, int to_colour) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int i = perm[p];
        if (to_colour) {
            dst.x[p] = src.x[i]; dst.y[p] = src.y[i]; dst.z[p] = src.z[i];
        } else {
            dst.x[i] = src.x[p]; dst.y[i] = src.y[p]; dst.z[i] = src.z[p];
        }
    }
}

// ======================= constrained Monte Carlo ===============================================
// montecarlo::cmc_step (src/montecarlo/cmc.cpp:282-497): pair moves that keep the magnetisation
// along the constraint direction.  The reference draws both sites of a pair at random; here the
// two sites of every pair have the same colour, so they do not interact with each other nor with
// any other pair of the launch, and all pairs of one colour are tried at once.  A round pairs the
// colour's atoms (caller order, m of them) through a random affine permutation
// pi(t) = (a t + b) mod m: pair p joins pi(2p) (spin 1) and pi(2p + 1) (spin 2); two rounds per
// colour make N pair trials per sweep, as the reference's N.  The acceptance's magnetisation is
// the replica's at the start of the round.  Every pair's arithmetic is the reference's.
template <typename R>
struct CmcFrame {            // cmc::polar_matrix, its transpose, cmc::polar_vector
    R P[9], Pt[9], v[3];
};

// The round's permutation: from Philox (colour | round << 16, sweep, STREAM_CMC_PAIR), b = word 1
// mod m, a = 1 + word 0 mod (m - 1), moved up (1 after m - 1) to the first value coprime with m.
__host__ __device__ __forceinline__ void cmc_permutation(uint2 key, int colour, int round, uint64_t sweep, int m,
                                                         uint64_t& a, uint64_t& b) {
    const U4 w = philox4x32_10((uint32_t)colour | ((uint32_t)round << 16), (uint32_t)sweep, (uint32_t)(sweep >> 32),
                               STREAM_CMC_PAIR, key.x, key.y);
    b = (uint64_t)(w.v[1] % (uint32_t)m);
    uint32_t x = m > 1 ? 1u + w.v[0] % (uint32_t)(m - 1) : 1u;
    for (;;) {
        uint32_t p = x, q = (uint32_t)m;
        while (q) { const uint32_t t = p % q; p = q; q = t; }
        if (p == 1u) break;
        x = x + 1u >= (uint32_t)m ? 1u : x + 1u;
    }
    a = x;
}

// a^-1 mod m for 0 < a < m coprime with m < 2^31 (extended Euclid; remainders stay below m and
// the coefficients below m in magnitude, so 32-bit arithmetic holds them)
__host__ __device__ __forceinline__ uint32_t mod_inverse(uint32_t a, uint32_t m) {
    int32_t t = 0, nt = 1;
    uint32_t r = m, nr = a;
    while (nr != 0u) {
        const uint32_t q = r / nr;
        const int32_t t2 = t - (int32_t)q * nt;
        const uint32_t r2 = r - q * nr;
        t = nt; nt = t2; r = nr; nr = r2;
    }
    return (uint32_t)(t < 0 ? t + (int32_t)m : t);
}

// Slots of a round: pair p of the permutation joins entries e = pi(2p) and e + a (mod m), and the
// pairs of a round are independent (disjoint atoms of one colour), so a workgroup may take them
// in any order.  Slot t takes the pair whose first entry is the t-th in list order:
//   m even: the first entries are exactly the entries of b's parity, so slot t is entry 2t + (b & 1)
//           (m/2 slots, all used);
//   m odd:  slot t is entry t when j = (t - b) a^-1 mod m is even and below m - 1, else idle
//           (m slots, (m - 1)/2 used).
// Both spins of the lanes of a wavefront then lie in a short stretch of the list, not at random.
// Returns false for an idle slot; e1 is then the first entry of pair 0 (b), a valid pair.
__device__ __forceinline__ bool cmc_slot(int t, int m, int4 pr, int& e1, int& e2) {
    const uint32_t a = (uint32_t)pr.x, b = (uint32_t)pr.y;
    bool ok;
    if ((m & 1) == 0) {
        ok = t < (m >> 1);
        e1 = ok ? 2 * t + (int)(b & 1u) : (int)b;
    } else {
        const uint32_t d = (uint32_t)t >= b ? (uint32_t)t - b : (uint32_t)t + (uint32_t)m - b;
        const uint32_t j = (uint32_t)(((uint64_t)d * (uint32_t)pr.z) % (uint32_t)m);
        ok = t < m && (j & 1u) == 0u && j < (uint32_t)(m - 1);
        e1 = ok ? t : (int)b;
    }
    const uint32_t f = (uint32_t)e1 + a;
    e2 = (int)(f >= (uint32_t)m ? f - (uint32_t)m : f);
    return ok;
}

template <typename R>
__device__ __forceinline__ void cmc_rotate(const R* M, R x0, R x1, R x2, R (&o)[3]) {   // cmc.cpp:363-365
    o[0] = M[0] * x0 + M[1] * x1 + M[2] * x2;
    o[1] = M[3] * x0 + M[4] * x1 + M[5] * x2;
    o[2] = M[6] * x0 + M[7] * x1 + M[8] * x2;
}

// Single-spin energy of atom i before (o) and after (n) a move, from one gather of its row, in the
// order of sim::calculate_spin_energy (src/simulate/energy.cpp:147-174).
template <typename R, int ET, int AM>
__device__ __forceinline__ void move_energies(const SysDev<R>& s, const Spin3<R>& S, const SpinRsrc<R>& rs, const R* ju,
                                              int i, int head,
                                              const MatConst<R>& mc, const Cond<R>& c, R ox, R oy, R oz, R nx, R ny,
                                              R nz, R& e_old, R& e_new) {
    R x_old, x_new;
    if constexpr (Access<AM>::TP != TP_CSR && sizeof(R) == 8) {
        RowIdx<AM> ri;
        fused_row_index(s, i, head, ri);
        row_energy_pair<R, ET, AM, 2>(s, rs, ri, ju, ox, oy, oz, nx, ny, nz, x_old, x_new);
    } else if constexpr (Access<AM>::TP != TP_CSR) {
        RowIdx<AM> ri;
        RowSpins<R, AM> rv;
        fused_row_index(s, i, head, ri);
        row_load(rs, ri, rv);
        x_old = row_exchange_energy<R, ET, AM>(s, ri, rv, ju, ox, oy, oz);
        x_new = row_exchange_energy<R, ET, AM>(s, ri, rv, ju, nx, ny, nz);
    } else {
        x_old = exchange_energy<R, ET, AM>(s, S.x, S.y, S.z, i, ox, oy, oz);
        x_new = exchange_energy<R, ET, AM>(s, S.x, S.y, S.z, i, nx, ny, nz);
    }
    e_old = (R(0) + x_old + anisotropy_energy(s.ani, mc, ox, oy, oz)) + applied_energy(c, ox, oy, oz);
    e_new = (R(0) + x_new + anisotropy_energy(s.ani, mc, nx, ny, nz)) + applied_energy(c, nx, ny, nz);
}

// The workgroup's outcome counts (LDS): accepted, energy rejects, sphere rejects
__device__ __forceinline__ unsigned int* cmc_wg_counts() {
    __shared__ unsigned int counts[3];
    return counts;
}

// The workgroup's change of magnetisation, one partial per wavefront (LDS)
__device__ __forceinline__ double* cmc_wg_dm() {
    __shared__ double dm[4][3];
    return &dm[0][0];
}

// One round of one colour.  tab[lb] = {first cmc-list entry of the (replica, colour) run,
// replica, run length m, first slot}; a workgroup (256 threads) takes 256 slots (cmc_slot) of one
// run; perm[rep] = {a, b, a^-1 mod m} of the round's permutation (k_cmc_perm).
// M: each replica's magnetisation sum(mu s) at the start of the round (double [nrep][3]);
// mm: moments (mu_B); lidx: with the spins in colour order (list, mm and the system's tables in
// colour order too), each atom's caller index within its replica, the Philox counter (else null:
// computed from the storage index); sigma: cone width per replica (move 0); counts: kRejectSlots slots of three
// outcome counts per replica, each slot on its own cache line; dM[3 lb ..]: the workgroup's
// sum of mu (s_new - s_old) over its accepted pairs, summed in a fixed order (lanes by a
// shuffle tree, then the four wavefronts in order), so that the next round's M follows
// without a pass over all spins.
template <typename R, int ET, int AM>
__global__ void __launch_bounds__(256)
k_cmc_pairs(SysDev<R> s, Spin3<R> S, const double* __restrict__ mm, const int* __restrict__ list,
            const int* __restrict__ lidx, const int4* __restrict__ tab, int ntab, const int4* __restrict__ perm,
            const CmcFrame<R>* __restrict__ frame, const double* __restrict__ M, const double* __restrict__ sigma,
            int move, int round, uint64_t sweep, double* __restrict__ counts, double* __restrict__ dM) {
    const int lb = logical_block(s.xcd_remap);
    if (lb >= ntab) return;                        // whole workgroup
    const int4 b = tab[lb];
    const int rep = b.y, m = b.z;
    int e1, e2;
    const bool ok = cmc_slot(b.w + (int)threadIdx.x, m, perm[rep], e1, e2);
    const uint2 key = s.rep_key[rep];
    const int i1 = list[b.x + e1];
    const int i2 = list[b.x + e2];
    const int m1 = s.nmat == 1 ? 0 : (int)s.type[i1];
    const int m2 = s.nmat == 1 ? 0 : (int)s.type[i2];
    const R a0 = S.x[i1], a1 = S.y[i1], a2 = S.z[i1];      // spin 1 initial
    const R c0 = S.x[i2], c1 = S.y[i2], c2 = S.z[i2];      // spin 2 initial
    const double mu1 = mm[i1], mu2 = mm[i2];
    int h1 = 0, h2 = 0;
    if (Access<AM>::TP != TP_CSR) { h1 = row_head<R, AM>(s, i1); h2 = row_head<R, AM>(s, i2); }
    const int local1 = lidx ? lidx[i1] : caller_local(s, rep, i1 - s.rep_off[rep]);
    const U4 um = philox4x32_10((uint32_t)local1, (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                STREAM_CMC_MOVE + (uint32_t)round, key.x, key.y);
    R g0, g1, g2;
    box_muller3(um, g0, g1, g2);
    if (threadIdx.x < 3) cmc_wg_counts()[threadIdx.x] = 0u;
    stage_fused<R, AM>(s, rep);
    const MatConst<R>& mc1 = FusedLds<R>::mat()[m1];
    const MatConst<R>& mc2 = FusedLds<R>::mat()[m2];
    const Cond<R>& cd1 = FusedLds<R>::cond(s)[m1];
    const Cond<R>& cd2 = FusedLds<R>::cond(s)[m2];
    const CmcFrame<R>& F = frame[rep];
    const R delta = move ? cd1.cmc_delta : (R)sigma[rep];
    // spin 1: the Monte Carlo move (mc_angle, mc_moves.cpp:71-87) seen in the constraint frame
    R s1i_mvd[3], s1f_mvd[3], s2i_mvd[3], s2f_mvd[3];
    cmc_rotate(F.P, a0, a1, a2, s1i_mvd);
    R f0 = a0 + g0 * delta, f1 = a1 + g1 * delta, f2 = a2 + g2 * delta;
    const R r = inv_norm(f0, f1, f2);
    f0 *= r; f1 *= r; f2 *= r;
    cmc_rotate(F.P, f0, f1, f2, s1f_mvd);
    const SpinRsrc<R> rs(s, S.x, S.y, S.z);
    const UniformJ<R, ET> ju(s, Access<AM>::JM);
    R e1o, e1n;
    move_energies<R, ET, AM>(s, S, rs, ju.v, i1, h1, mc1, cd1, a0, a1, a2, f0, f1, f2, e1o, e1n);
    const R delta_energy1 = (e1n - e1o) * mc1.mu_s_SI * R(1.07828231e23);
    // spin 2 compensates the transverse moment in the constraint frame (cmc.cpp:416-423)
    cmc_rotate(F.P, c0, c1, c2, s2i_mvd);
    const R mom_ratio = (R)(mu1 / mu2);
    s2f_mvd[0] = mom_ratio * (s1i_mvd[0] - s1f_mvd[0]) + s2i_mvd[0];
    s2f_mvd[1] = mom_ratio * (s1i_mvd[1] - s1f_mvd[1]) + s2i_mvd[1];
    const bool on_sphere = (s2f_mvd[0] * s2f_mvd[0] + s2f_mvd[1] * s2f_mvd[1]) < R(1);
    // off the sphere the pair is rejected; the lane still runs the arithmetic below on a point
    // of the sphere so every lane stays in step
    s2f_mvd[2] = (s2i_mvd[2] < R(0) ? R(-1) : R(1)) *
                 sqrt(on_sphere ? R(1) - s2f_mvd[0] * s2f_mvd[0] - s2f_mvd[1] * s2f_mvd[1] : R(1));
    R d[3];
    cmc_rotate(F.Pt, s2f_mvd[0], s2f_mvd[1], s2f_mvd[2], d);
    R e2o, e2n;
    move_energies<R, ET, AM>(s, S, rs, ju.v, i2, h2, mc2, cd2, c0, c1, c2, d[0], d[1], d[2], e2o, e2n);
    const R delta_energy2 = (e2n - e2o) * mc2.mu_s_SI * R(1.07828231e23);
    const R delta_energy21 = delta_energy1 * cd1.kBTBohr + delta_energy2 * cd2.kBTBohr;
    const double* Mr = M + 3 * (size_t)rep;
    const double v0 = (double)F.v[0], v1 = (double)F.v[1], v2 = (double)F.v[2];
    const double Mz_old = Mr[0] * v0 + Mr[1] * v1 + Mr[2] * v2;
    const double Mz_new = (Mr[0] + mu1 * (double)f0 + mu2 * (double)d[0] - mu1 * (double)a0 - mu2 * (double)c0) * v0 +
                          (Mr[1] + mu1 * (double)f1 + mu2 * (double)d[1] - mu1 * (double)a1 - mu2 * (double)c1) * v1 +
                          (Mr[2] + mu1 * (double)f2 + mu2 * (double)d[2] - mu1 * (double)a2 - mu2 * (double)c2) * v2;
    const double probability = exp(-(double)delta_energy21) * ((Mz_new / Mz_old) * (Mz_new / Mz_old)) *
                               fabs((double)s2i_mvd[2] / (double)s2f_mvd[2]);
    const U4 ua = philox4x32_10((uint32_t)local1, (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                STREAM_CMC_ACCEPT + (uint32_t)round, key.x, key.y);
    const bool accept = on_sphere && (probability >= (double)ua.v[0] * 2.3283064365386963e-10) && (Mz_new > 0.0);
    if (ok && accept) {
        S.x[i1] = f0; S.y[i1] = f1; S.z[i1] = f2;
        S.x[i2] = d[0]; S.y[i2] = d[1]; S.z[i2] = d[2];
    }
    const int outcome = accept ? 0 : (on_sphere ? 1 : 2);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const unsigned long long hit = __ballot(ok && outcome == k);
        if (hit != 0ull && (int)__lane_id() == __ffsll(hit) - 1) atomicAdd(&cmc_wg_counts()[k], (unsigned)__popcll(hit));
    }
    double dm[3] = {0.0, 0.0, 0.0};
    if (ok && accept) {
        dm[0] = mu1 * (double)f0 + mu2 * (double)d[0] - mu1 * (double)a0 - mu2 * (double)c0;
        dm[1] = mu1 * (double)f1 + mu2 * (double)d[1] - mu1 * (double)a1 - mu2 * (double)c1;
        dm[2] = mu1 * (double)f2 + mu2 * (double)d[2] - mu1 * (double)a2 - mu2 * (double)c2;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        for (int off = 32; off > 0; off >>= 1) dm[q] += __shfl_down(dm[q], off, 64);
        if (__lane_id() == 0) cmc_wg_dm()[3 * (threadIdx.x >> 6) + q] = dm[q];
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const double* w = cmc_wg_dm();
        dM[3 * (size_t)lb + threadIdx.x] = ((w[threadIdx.x] + w[3 + threadIdx.x]) + w[6 + threadIdx.x]) + w[9 + threadIdx.x];
        if (cmc_wg_counts()[threadIdx.x] != 0u)
            atomicAdd(&counts[((size_t)rep * kRejectSlots + (lb & (kRejectSlots - 1))) * kRejectStride + threadIdx.x],
                      (double)cmc_wg_counts()[threadIdx.x]);
    }
}

}  // namespace vmx
