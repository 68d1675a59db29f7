// lb_kernels.hip -- gfx950 kernels of the OpenLUDWIG time step.
//
// One work-item per cell, 256 work-items (4 wave64) per work-group covering
// half of an 8^3 block; a wave is one 8x8 z-plane, so every population load
// and store of a wave is a contiguous 256-byte row of the block-major SoA
// layout f[b][k][c].  The pull for a cell whose source lies in another block
// goes through that block's neighbour-table row, staged in LDS; the centre
// entry of the row is the block itself, so in-block and cross-block pulls are
// one load instruction with a per-lane block index.  Sources with no
// neighbour block fall to the boundary cascade (inlet/outlet equilibrium,
// y/z mirror, coarse-to-fine interpolation), a non-inlined cold path.
//
// Operation order follows the reference kernels statement by statement (see
// the citations), so that with -ffp-contract=off the results equal a plain
// C restatement of them bit for bit.
//
// This file is compiled twice; LBK_NS names the namespace (lbk_strict /
// lbk_fast).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <atomic>
#include <cstdint>
#include <type_traits>

#include "ludwig_hip.h"
#include "lb_kernels.h"

#ifndef LBK_NS
#error "LBK_NS must be defined"
#endif

namespace LBK_NS {
using namespace lbk;

// ---------------------------------------------------------------------------
// D3Q27 lattice (src/collision/lattice.jl:5-56, timestep.jl:161-179):
// k = (cx+1) + 3(cy+1) + 9(cz+1), 0-based; rest = 13; opp(k) = 26-k.
__host__ __device__ constexpr int CX(int k) { return k % 3 - 1; }
__host__ __device__ constexpr int CY(int k) { return (k / 3) % 3 - 1; }
__host__ __device__ constexpr int CZ(int k) { return k / 9 - 1; }
__host__ __device__ constexpr int OPP(int k) { return 26 - k; }
__host__ __device__ constexpr int MIRY(int k) { return k - 6 * CY(k); }
__host__ __device__ constexpr int MIRZ(int k) { return k - 18 * CZ(k); }
__host__ __device__ constexpr float WK(int k) {
    return (CX(k) * CX(k) + CY(k) * CY(k) + CZ(k) * CZ(k)) == 0   ? 8.0f / 27.0f
           : (CX(k) * CX(k) + CY(k) * CY(k) + CZ(k) * CZ(k)) == 1 ? 2.0f / 27.0f
           : (CX(k) * CX(k) + CY(k) * CY(k) + CZ(k) * CZ(k)) == 2 ? 1.0f / 54.0f
                                                                  : 1.0f / 216.0f;
}

__device__ __forceinline__ size_t fidx(int b, int k, int c) { return ((size_t)b * Q + k) * NC + c; }

// Per-launch views of one level at one substep, resolved from LevelArgs and
// CaseState at kernel entry (the buffers of the substep's parity, the seed,
// the inlet velocity).  All fields are uniform over the work-group.
struct CollideArgs {
    const void* f_in;
    void* f_out;
    const float* vel_in;
    float* vel_out;
    float* rho_out;
    const uint8_t* obst;
    const float* sponge;
    const float* wall_dist;
    const uint8_t* bflags;
    const int32_t* nbr;
    const int32_t* borg;
    int32_t nb;
    int32_t nxg, nyg, nzg;
    float tau;
    float u_inlet;
    float c_wale, nu_sgs_bg, inlet_turb;
    float omega_bulk, omega3, omega4, lambda;
    int32_t adaptive, limiter, comp, wall_model, sponge_blend;
    int32_t seed;
    int32_t is_l1;
    int32_t store_rho;  // store rho_out on this substep
};

struct ThermalArgs {
    const void* g_in;  // thermal populations, storage as the flow's (offsets from w_k in 16 bits)
    void* g_out;
    const float* vel;  // the flow step's vel_out
    float* T_out;
    const uint8_t* obst;
    const float* sponge;
    const int32_t* nbr;
    const int32_t* borg;
    int32_t nb;
    int32_t nxg, nyg, nzg;
    float tau_g;
    float u_inlet, T_inlet, T_wall;
    int32_t adiabatic, sponge_blend;
    int32_t is_l1;
    int32_t store_T;  // store T_out on this substep
};

struct GhostArgs {
    void* f;   // input half of substep j (ghost blocks at index >= nb)
    void* f2;  // input half of substep j + 1
    const int32_t* entries;
    const int32_t* parents;  // [n][8] parent cells of each entry (LevelArgs::gpar)
    int64_t n;
    const void* a_new;
    const float* s_new;
    const float* u_new;
    const void* a_old;
    const float* s_old;
    const float* u_old;
    float tau_fine, tau_coarse;
    int32_t temporal, comp, thermal;
};

// The case record, read through the global address space (uniform loads).
__device__ __forceinline__ CaseState case_state(const CaseState* st) {
    static_assert(sizeof(CaseState) % 4 == 0, "CaseState is read as 32-bit words");
    typedef const __attribute__((address_space(1))) uint32_t GW;
    const GW* src = (const GW*)st;
    CaseState out;
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (unsigned i = 0; i < sizeof(CaseState) / 4; i++) dst[i] = src[i];
    return out;
}

__device__ __forceinline__ int64_t substep(const LevelArgs& L, const CaseState& S, int64_t j) {
    return (S.t0 << L.lshift) + j;
}

// Coarse step that substep j of the level belongs to.  An engine step can
// cover several coarse steps of a case; those past the case's last step are
// skipped (the case then idles until it is removed).
__device__ __forceinline__ int64_t coarse_step(const LevelArgs& L, const CaseState& S, int64_t j) {
    return S.t0 + (j >> L.lshift);
}
__device__ __forceinline__ bool live(const LevelArgs& L, const CaseState& S, int64_t j) {
    return S.active != 0 && coarse_step(L, S, j) <= S.n_total;
}
// Inlet velocity of the coarse step: the ramp table (engine, one value per
// batch of S.batch coarse steps, main.jl:214-218) or the value of the call.
__device__ __forceinline__ float inlet_u(const LevelArgs& L, const CaseState& S, int64_t j) {
    if (S.u_tab == nullptr) return S.u;
    const int64_t b = (coarse_step(L, S, j) - 1) / S.batch;
    return ((const __attribute__((address_space(1))) float*)S.u_tab)[b];
}

// Whether substep j stores the level's density (flow) or temperature
// (thermal) output.  A coarser level's are read by the ghost fill of its
// child on every substep; the finest level's only by forces, statistics and
// downloads, which run between coarse steps, so there only the last substep
// of each coarse step stores them (4 of the 125 bytes written per cell in
// FP32).
__device__ __forceinline__ int32_t store_scalars(const LevelArgs& L, int64_t j) {
    const int64_t m = ((int64_t)1 << L.lshift) - 1;
    return (L.leaf == 0 || (j & m) == m) ? 1 : 0;
}

__device__ __forceinline__ CollideArgs collide_view(const LevelArgs& L, const CaseState& S, int64_t j) {
    const int64_t ts = substep(L, S, j);
    const int par = (int)(ts & 1);
    CollideArgs a;
    a.f_in = L.f[par];
    a.f_out = L.f[1 - par];
    a.vel_in = L.vel[par];
    a.vel_out = L.vel[1 - par];
    a.rho_out = L.rho[1 - par];
    a.obst = L.obst;
    a.sponge = L.sponge;
    a.wall_dist = L.wall_dist;
    a.bflags = L.bflags;
    a.nbr = L.nbr;
    a.borg = L.borg;
    a.nb = L.nb;
    a.nxg = L.nxg;
    a.nyg = L.nyg;
    a.nzg = L.nzg;
    a.tau = L.tau;
    a.u_inlet = inlet_u(L, S, j);
    a.c_wale = S.c_wale;
    a.nu_sgs_bg = S.nu_sgs_bg;
    a.inlet_turb = S.inlet_turb;
    a.omega_bulk = S.omega_bulk;
    a.omega3 = S.omega3;
    a.omega4 = S.omega4;
    a.lambda = S.lambda;
    a.adaptive = S.adaptive;
    a.limiter = S.limiter;
    a.comp = S.comp;
    a.wall_model = (S.wall_model == 1 && L.wall_dist != nullptr) ? 1 : 0;
    a.sponge_blend = S.sponge_blend;
    a.seed = (int32_t)(ts % 1000000);
    a.is_l1 = L.lshift == 0 ? 1 : 0;
    a.store_rho = store_scalars(L, j);
    return a;
}

__device__ __forceinline__ ThermalArgs thermal_view(const LevelArgs& L, const CaseState& S, int64_t j) {
    const int par = (int)(substep(L, S, j) & 1);
    ThermalArgs t;
    t.g_in = L.g[par];
    t.g_out = L.g[1 - par];
    t.vel = L.vel[1 - par];
    t.T_out = L.T[1 - par];
    t.obst = L.obst;
    t.sponge = L.sponge;
    t.nbr = L.nbr;
    t.borg = L.borg;
    t.nb = L.nb;
    t.nxg = L.nxg;
    t.nyg = L.nyg;
    t.nzg = L.nzg;
    t.tau_g = L.tau_g;
    t.u_inlet = inlet_u(L, S, j);
    t.T_inlet = S.T_inlet;
    t.T_wall = S.T_wall;
    t.adiabatic = S.adiabatic;
    t.sponge_blend = S.sponge_blend;
    t.is_l1 = L.lshift == 0 ? 1 : 0;
    t.store_T = store_scalars(L, j);
    return t;
}

// Coarse-to-fine source of substeps j and j + 1 (j even, level > 1): the
// parent's last step is substep j >> 1 of the parent level; its output pair
// half is "new", its input half (kept untouched until the parent's next
// step) is "old".  The temporal weight is 0 for the first child substep and
// 0.5 for the second (solver_control.jl:66-96); f and f2 are their input
// halves.
__device__ __forceinline__ GhostArgs ghost_view(const LevelArgs& L, const CaseState& S, int64_t j, int thermal) {
    const int par = (int)(substep(L, S, j) & 1);
    const int ppar = (int)(((S.t0 << (L.lshift - 1)) + (j >> 1)) & 1);
    GhostArgs g;
    g.f = thermal ? (void*)L.g[par] : L.f[par];
    g.f2 = thermal ? (void*)L.g[1 - par] : L.f[1 - par];
    g.entries = L.gent;
    g.n = L.ngent;
    g.parents = L.gpar;
    g.a_new = thermal ? (const void*)L.pg[1 - ppar] : L.pf[1 - ppar];
    g.s_new = thermal ? L.pT[1 - ppar] : L.prho[1 - ppar];
    g.u_new = L.pvel[1 - ppar];
    g.a_old = thermal ? (const void*)L.pg[ppar] : L.pf[ppar];
    g.s_old = thermal ? L.pT[ppar] : L.prho[ppar];
    g.u_old = L.pvel[ppar];
    g.tau_fine = thermal ? L.tau_g : L.tau;
    g.tau_coarse = thermal ? L.ptau_g : L.ptau;
    g.temporal = S.temporal;
    g.comp = S.comp;
    g.thermal = thermal;
    return g;
}

// Global-address-space views of the argument pointers, so that loads and
// stores compile to global_* (not flat_*) instructions.
typedef __attribute__((address_space(1))) float GF;
typedef __attribute__((address_space(1))) uint8_t GU8;
__device__ __forceinline__ const GF* gp(const float* p) { return (const GF*)p; }
__device__ __forceinline__ GF* gp(float* p) { return (GF*)p; }
__device__ __forceinline__ const GU8* gp(const uint8_t* p) { return (const GU8*)p; }
typedef __attribute__((address_space(1))) uint16_t GU16;
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) v4u GV4U;
typedef __attribute__((address_space(1))) char GC;

// Population storage.  ST = 0: FP32.  ST = 1 / 2: the offset f - w_k in
// binary16 / bfloat16, round-to-nearest-even; all arithmetic stays FP32 and a
// stored value decodes to the float w_k + offset.
template <int ST>
struct Codec;
template <>
struct Codec<0> {
    static __device__ __forceinline__ float load(const void* base, size_t i, float) {
        return ((const GF*)base)[i];
    }
    template <class OffT>
    static __device__ __forceinline__ float loadb(const void* base, OffT off, float) {
        return *(const GF*)((const GC*)base + off);
    }
    static __device__ __forceinline__ void store(void* base, size_t i, float f, float) { ((GF*)base)[i] = f; }
    // non-temporal store, for the stream-collide outputs (read again only
    // one substep later, after the whole level has streamed through L2)
    static __device__ __forceinline__ void stream(void* base, size_t i, float f, float) {
        __builtin_nontemporal_store(f, &((GF*)base)[i]);
    }
    // the stored bits of f (for the staged stores of stream_tile)
    static __device__ __forceinline__ uint32_t bits(float f, float) { return __float_as_uint(f); }
};
template <>
struct Codec<1> {
    static __device__ __forceinline__ float load(const void* base, size_t i, float w) {
        const uint16_t u = ((const GU16*)base)[i];
        return w + __half2float(__ushort_as_half(u));
    }
    template <class OffT>
    static __device__ __forceinline__ float loadb(const void* base, OffT off, float w) {
        const uint16_t u = *(const GU16*)((const GC*)base + off);
        return w + __half2float(__ushort_as_half(u));
    }
    static __device__ __forceinline__ void store(void* base, size_t i, float f, float w) {
        ((GU16*)base)[i] = __half_as_ushort(__float2half_rn(f - w));
    }
    static __device__ __forceinline__ void stream(void* base, size_t i, float f, float w) {
        __builtin_nontemporal_store((uint16_t)__half_as_ushort(__float2half_rn(f - w)), &((GU16*)base)[i]);
    }
    static __device__ __forceinline__ uint16_t bits(float f, float w) {
        return (uint16_t)__half_as_ushort(__float2half_rn(f - w));
    }
};
template <>
struct Codec<2> {
    static __device__ __forceinline__ float load(const void* base, size_t i, float w) {
        const uint32_t u = ((const GU16*)base)[i];
        return w + __uint_as_float(u << 16);
    }
    template <class OffT>
    static __device__ __forceinline__ float loadb(const void* base, OffT off, float w) {
        const uint32_t u = *(const GU16*)((const GC*)base + off);
        return w + __uint_as_float(u << 16);
    }
    static __device__ __forceinline__ uint16_t enc(float x) {
        uint32_t b = __float_as_uint(x);
        if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);   // quiet NaN
        b += 0x7fffu + ((b >> 16) & 1u);
        return (uint16_t)(b >> 16);
    }
    static __device__ __forceinline__ void store(void* base, size_t i, float f, float w) {
        ((GU16*)base)[i] = enc(f - w);
    }
    static __device__ __forceinline__ void stream(void* base, size_t i, float f, float w) {
        __builtin_nontemporal_store(enc(f - w), &((GU16*)base)[i]);
    }
    static __device__ __forceinline__ uint16_t bits(float f, float w) { return enc(f - w); }
};

// Staged stores of a wave's stream-collide outputs.  A wave is one z-plane
// (64 consecutive cells) of its block, so row k of its output is 64
// consecutive slots of f[b][k][*].  Each lane writes its cell's stored values
// for ROWS directions into the wave's [ROWS][64] tile in LDS; then each lane
// stores 16 bytes of one row.  With ROWS = 27 the outputs leave in 7 (FP32)
// or 4 (16-bit) vector stores per wave instead of 27 scalar ones; ROWS = 14
// halves the tile for 8 stores in two phases.  The LDS operations of one
// wave complete in order, so the phases need no barrier.
template <int ST>
using Stored = std::conditional_t<ST == 0, uint32_t, uint16_t>;

template <int ST, int ROWS>
__device__ __forceinline__ void stream_tile(void* fo, int b, int c, const float (&A)[27], Stored<ST>* tile) {
    constexpr int EPL = 16 / (int)sizeof(Stored<ST>);  // values per 16-byte store
    constexpr int LPR = 64 / EPL;                      // lanes per row of 64 values
    const int lane = c & 63;
    const int sub = lane % LPR;
    const int c0 = c & ~63;
#pragma unroll
    for (int k0 = 0; k0 < 27; k0 += ROWS) {
        const int nr = 27 - k0 < ROWS ? 27 - k0 : ROWS;
#pragma unroll
        for (int r = 0; r < nr; r++) tile[r * 64 + lane] = Codec<ST>::bits(A[k0 + r], WK(k0 + r));
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int i = 0; i < (nr + EPL - 1) / EPL; i++) {
            const int r = EPL * i + lane / LPR;
            if (r < nr) {
                const v4u v = *(const v4u*)(tile + r * 64 + sub * EPL);
                const size_t slot = fidx(b, k0 + r, c0) + (size_t)(sub * EPL);
                __builtin_nontemporal_store(v, (GV4U*)((GC*)fo + slot * sizeof(Stored<ST>)));
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Equilibria (src/collision/physics_utils.jl:19-82)
__device__ __forceinline__ float feq2(float rho, float ux, float uy, float uz, float w, float cx, float cy, float cz) {
    float cu = cx * ux + cy * uy + cz * uz;
    float usq = ux * ux + uy * uy + uz * uz;
    return rho * w * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * usq);
}
__device__ __forceinline__ float feq4(float rho, float ux, float uy, float uz, float w, float cx, float cy, float cz) {
    float cu = cx * ux + cy * uy + cz * uz;
    float usq = ux * ux + uy * uy + uz * uz;
    float cu2 = cu * cu;
    float cu3 = cu2 * cu;
    float cu4 = cu2 * cu2;
    float usq2 = usq * usq;
    float val = 1.0f + 3.0f * cu + 4.5f * cu2 - 1.5f * usq;
    val += 4.5f * cu3 - 4.5f * cu * usq;
    val += 3.375f * cu4 - 6.75f * cu2 * usq + 1.125f * usq2;
    return rho * w * val;
}
__device__ __forceinline__ float feq_auto(float rho, float ux, float uy, float uz, float w, float cx, float cy,
                                          float cz, int comp) {
    return comp == 1 ? feq4(rho, ux, uy, uz, w, cx, cy, cz) : feq2(rho, ux, uy, uz, w, cx, cy, cz);
}
__device__ __forceinline__ float geq(float T, float ux, float uy, float uz, float w, float cx, float cy, float cz) {
    float cu = cx * ux + cy * uy + cz * uz;
    return w * T * (1.0f + 3.0f * cu);
}

// gpu_hash / gradient_noise (physics_utils.jl:5-16): Int32 arithmetic wraps.
__device__ __forceinline__ float gradient_noise(int gx, int gy, int gz, int seed) {
    uint32_t c = (uint32_t)gx * 374761393u + (uint32_t)gy * 668265263u + (uint32_t)gz * 1274126177u + (uint32_t)seed;
    uint32_t h = c;
    h = (h ^ (h >> 16)) * 0x85ebca6bu;
    h = (h ^ (h >> 13)) * 0xc2b2ae35u;
    h = h ^ (h >> 16);
    return ((float)(h & 0xFFFFu) / 32768.0f) - 1.0f;
}

__device__ __forceinline__ float wk_rt(int k) {
    int d2 = CX(k) * CX(k) + CY(k) * CY(k) + CZ(k) * CZ(k);
    return d2 == 0 ? 8.0f / 27.0f : d2 == 1 ? 2.0f / 27.0f : d2 == 2 ? 1.0f / 54.0f : 1.0f / 216.0f;
}

// Missing-neighbour cascade of stream_collide_cumulant! (kernel_cumulant.jl:117-159)
// for sources outside the domain.  An in-domain source without a block (the
// coarse-to-fine case of :143-156) never reaches here: it is a ghost block
// filled by ghost_kernel.  The last branch (a level-1 source inside the
// domain with no block) cannot occur either, level 1 being the full box; it
// returns w_k as the reference does.
template <int ST>
__device__ __forceinline__ float bc_value(const CollideArgs& a, int k, int b, int c, int sgx, int sgy, int sgz, int gy,
                                          int gz) {
    float cx = (float)CX(k), cy = (float)CY(k), cz = (float)CZ(k), w = wk_rt(k);
    if (sgx < 1) {
        float noise = a.inlet_turb > 0.0f ? gradient_noise(gy, gz, a.seed, 1234) * a.inlet_turb * a.u_inlet : 0.0f;
        float u_inst = a.u_inlet + noise;
        return feq_auto(1.0f, u_inst, 0.0f, 0.0f, w, cx, cy, cz, a.comp);
    }
    if (sgx > a.nxg) return feq_auto(1.0f, a.u_inlet, 0.0f, 0.0f, w, cx, cy, cz, a.comp);
    // the mirrored direction has the same |c|, hence the same weight
    if (sgy < 1 || sgy > a.nyg) return Codec<ST>::load(a.f_in, fidx(b, MIRY(k), c), w);
    if (sgz < 1 || sgz > a.nzg) return Codec<ST>::load(a.f_in, fidx(b, MIRZ(k), c), w);
    return w;
}

__device__ __forceinline__ void chim_fwd(float& fm, float& f0, float& fp, float u) {
    float m0 = fm + f0 + fp;
    float m1 = fp - fm;
    float m2 = fm + fp;
    fm = m0;
    f0 = m1 - u * m0;
    fp = m2 - 2.0f * u * m1 + u * u * m0;
}
__device__ __forceinline__ void chim_bwd(float& k0, float& k1, float& k2, float u) {
    float m0 = k0;
    float m1 = k1 + u * k0;
    float m2 = k2 + 2.0f * u * k1 + u * u * k0;
    float fp = (m2 + m1) * 0.5f;
    float fm = (m2 - m1) * 0.5f;
    float f0 = m0 - m2;
    k0 = fm;
    k1 = f0;
    k2 = fp;
}

// Work-group -> block schedule, two work-groups per block of the work list.
// sched = 1 (the default) maps work-group i to half-tile i, so consecutive
// blocks of the processing order (the storage order) run at the same time,
// spread over the 8 XCDs.  sched = 0 swizzles the work-groups, which are
// dealt round-robin to the XCDs, so that every XCD takes one contiguous run
// of the order (neighbours share an L2, but the level is swept at 8 places
// at once: finest stream-collide of the 52.5M-cell cube case 1612 us against
// 1545 us in order).  Returns false for padding work-groups.
__device__ __forceinline__ bool schedule(const Work* __restrict__ work, int n_tiles, int sched, Work& wk,
                                         int& half) {
    const unsigned G = gridDim.x;            // multiple of 8
    const unsigned w = blockIdx.x;
    const unsigned logical = sched == 1 ? w : (w & 7u) * (G >> 3) + (w >> 3);
    if (logical >= 2u * (unsigned)n_tiles) return false;
    wk = work[logical >> 1];
    half = (int)(logical & 1u);
    return true;
}

// Pull tables.  A wave is one z-plane of its block, so for a pulled direction
// the z-side of the source block is the same for the whole wave and only the
// x/y side varies by lane.  For each of the work-group's 4 waves and each
// entry e = 9(cz+1) + (ox+1) + 3(oy+1) (ox/oy: x/y side of the source block
// seen by the lane), ptab holds a byte offset such that
//     ptab[e] + esz * (c + 512 k - (cx + 8 cy + 64 cz))
// is the source slot (k, wrapped source cell) in that block; the cell index c
// and the direction constant complete it with one add per load.  A missing
// block (no neighbour, outside the domain) maps to an in-bounds slot of the
// cell's own block and sets bit e of the wave's missing mask; the boundary
// cascade replaces the value.  Offsets are computed modulo 2^OffT bits; every final offset is
// in [0, buffer size), which the host guarantees fits OffT.
template <class OffT>
__device__ __forceinline__ void build_pull_table(const int32_t* __restrict__ row, int b, int half, OffT esz,
                                                 OffT* ptab, uint64_t* smask, int tid) {
    const int t = tid;
    bool missing = false;
    if (t < 108) {
        const int w = t / 27, e = t - 27 * w;
        const int cz = e / 9 - 1, j = e - 9 * (cz + 1);
        const int ox = j % 3 - 1, oy = j / 3 - 1;
        const int zw = half * 4 + w;
        const int oz = (cz == 1 && zw == 0) ? -1 : ((cz == -1 && zw == 7) ? 1 : 0);
        const int nbb = row[j + 9 * (oz + 1)];
        const OffT bs = (OffT)(Q * NC);
        OffT v;
        if (nbb >= 0) {
            v = ((OffT)nbb * bs - (OffT)(8 * ox + 64 * oy + 512 * oz)) * esz;
        } else {
            v = ((OffT)b * bs + (OffT)128) * esz;
            missing = true;
        }
        ptab[t] = v;
    }
    // entries 0..107 live in waves 0 and 1: their ballots form a 128-bit mask
    const uint64_t m = __ballot(missing);
    if (t == 0) smask[0] = m;
    if (t == 64) smask[1] = m;
}

// The 27 missing-block bits of wave w (entries 27w .. 27w+26 of the mask).
__device__ __forceinline__ uint32_t wave_miss(const uint64_t* smask, int w) {
    const uint64_t lo = smask[0], hi = smask[1];
    const int o = 27 * w;
    const uint64_t v = o == 0 ? lo : (o < 64 ? (lo >> o) | (hi << (64 - o)) : hi >> (o - 64));
    return (uint32_t)(v & 0x7ffffffu);
}

// Velocity stencil table (WALE neighbours, physics_utils.jl:112-123): per
// wave and face n (+x, -x, +y, -y, +z, -z), the byte offset for lanes whose
// neighbour lies in the next block (x = 7 for +x, ...; for +-z the whole wave
// or none), such that vtab + 4 (c + dx + 8 dy + 64 dz) addresses component 0;
// entry 6 is the own block.  Neighbours outside the level (no block, or a
// ghost block, which holds no velocity) use the cell's own velocity.
template <class OffT>
__device__ __forceinline__ void build_vel_table(const int32_t* __restrict__ row, int nb, int b, int half,
                                                OffT* vtab, int tid) {
    const int t = tid - 128;
    if (t >= 0 && t < 32) {
        const int w = t >> 3, n = t & 7;
        const OffT own = (OffT)b * (OffT)(3 * NC * 4);
        OffT v = own;
        if (n < 6) {
            const int dx = n == 0 ? 1 : (n == 1 ? -1 : 0);
            const int dy = n == 2 ? 1 : (n == 3 ? -1 : 0);
            const int dz = n == 4 ? 1 : (n == 5 ? -1 : 0);
            const int zw = half * 4 + w;
            const bool zedge = (dz == 1 && zw == 7) || (dz == -1 && zw == 0);
            if (dz == 0 || zedge) {
                const int nbb = row[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
                if (nbb >= 0 && nbb < nb)
                    v = (OffT)nbb * (OffT)(3 * NC * 4) - (OffT)(4 * (8 * dx + 64 * dy + 512 * dz));
                else
                    v = own - (OffT)(4 * (dx + 8 * dy + 64 * dz));
            }
        }
        vtab[t] = v;
    }
}

template <class OffT>
__device__ __forceinline__ const GF* gaddr(const void* base, OffT off) {
    return (const GF*)((const GC*)base + off);
}

// Velocity tile of a work-group (half a block, 4 z-planes) for the WALE
// stencil: per component, planes zz = -1..4 (the 4 own planes and the two
// z-face halos) of 10 x 10 cells (x, y = -1..8; of the halo ring only the
// faces are filled and read).  Every lane loads its own cell; wave 0 loads the
// -z face plane, wave 1 the +z plane, wave 2 the two x faces and wave 3 the
// two y faces, each value through the stencil table (the neighbour block's
// cell, or the adjacent own cell where the neighbour is outside the level).
// 6 loads per lane instead of the 18 of a per-lane stencil.  vel_tile_load
// issues the loads; vel_tile_store writes the tile.
constexpr int VT_SIZE = 6 * 100;
__device__ __forceinline__ int vt_index(int d, int zz, int yy, int xx) {
    return d * VT_SIZE + (zz + 1) * 100 + (yy + 1) * 10 + (xx + 1);
}
struct VelTile {
    float o[3], h[3];
    int hz, hy, hx;
};
template <class OffT>
__device__ __forceinline__ VelTile vel_tile_load(const float* vel, const OffT* vtab, int half, int c, int tid) {
    const int lane = tid & 63, wv = tid >> 6;
    VelTile t;
    const GF* p = gaddr(vel, vtab[wv * 8 + 6] + (OffT)(4 * c));
    t.o[0] = p[0], t.o[1] = p[NC], t.o[2] = p[2 * NC];
    int n, ce;  // face and own edge cell of this lane's halo value
    if (wv == 0) {
        n = 5, ce = (half << 8) + lane, t.hz = -1, t.hy = lane >> 3, t.hx = lane & 7;
    } else if (wv == 1) {
        n = 4, ce = (half << 8) + 192 + lane, t.hz = 4, t.hy = lane >> 3, t.hx = lane & 7;
    } else {
        const int side = lane >> 5, s = (lane >> 2) & 7, zl = lane & 3;
        if (wv == 2)
            n = side ? 0 : 1, ce = (half << 8) + zl * 64 + 8 * s + (side ? 7 : 0), t.hz = zl, t.hy = s,
            t.hx = side ? 8 : -1;
        else
            n = side ? 2 : 3, ce = (half << 8) + zl * 64 + 8 * (side ? 7 : 0) + s, t.hz = zl,
            t.hy = side ? 8 : -1, t.hx = s;
    }
    const int dx = n == 0 ? 1 : (n == 1 ? -1 : 0);
    const int dy = n == 2 ? 1 : (n == 3 ? -1 : 0);
    const int dz = n == 4 ? 1 : (n == 5 ? -1 : 0);
    const GF* q = gaddr(vel, vtab[((ce >> 6) & 3) * 8 + n] + (OffT)(4 * (ce + dx + 8 * dy + 64 * dz)));
    t.h[0] = q[0], t.h[1] = q[NC], t.h[2] = q[2 * NC];
    return t;
}
// Second half: the values into the tile, once the caller has issued its
// population loads, so that the wait and the barrier overlap them.
__device__ __forceinline__ void vel_tile_store(const VelTile& t, int c, float* vt, int tid) {
    const int wv = tid >> 6, x = c & 7, y = (c >> 3) & 7;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        vt[vt_index(d, wv, y, x)] = t.o[d];
        vt[vt_index(d, t.hz, t.hy, t.hx)] = t.h[d];
    }
}

// stream_collide_cumulant! (src/collision/kernel_cumulant.jl:29-627), with the
// A/B buffers chosen by the caller and no f_post_collision copy (Bouzidi
// reads f_out before it corrects it, see bouzidi_gather below).
// Occupancy: the collision needs ~160 VGPRs, which allows 3 waves per SIMD.
// Capping the kernel's VGPRs (W waves per SIMD) spills values to scratch,
// which stays in cache, and lets more waves hide the pull's memory latency.
// Boundary blocks (some neighbour outside the domain) need the 27 KB cascade
// buffer in LDS, which limits them to 5 work-groups per CU; they run at 5.
// Interior blocks need no cascade buffer (their LDS holds the output tile and,
// in FP32, the velocity tile: 14-23 KB); W = 5 or 6 is a launch option
// (picked by the autotuner).  The two compile to the same floating-point
// operations (tested bit-identical); a 4-wave build contracts differently and
// is not offered.  Measured on the 52.5M-cell cube case with the staged
// stores: 5 waves is best (6 waves spills more: FP32 +5 %, FP16 +7 %).
//
// collide_tile is one half-block tile (work-list entry wk, half `half`);
// collide_kernel runs one tile per work-group, collide_kernel_persistent
// (LB_SCHED_PERSISTENT) runs a grid of as many work-groups as fit the
// device at once, each looping over the tiles in order.
template <int ST, class OffT, bool BND>
__device__ __forceinline__ void collide_tile(const LevelArgs* __restrict__ las, const CaseState* __restrict__ states,
                                             const Work& wk, int half, int64_t j, int tid) {
    __shared__ OffT ptab[108];
    __shared__ OffT vtab[32];
    __shared__ uint64_t smask[2];
    __shared__ float sfix[BND ? 27 * 256 : 1];
    // interior blocks: outputs leave through stream_tile (the boundary
    // kernel's cascade buffer already takes the LDS budget)
    constexpr bool STAGE = !BND;
    // Interior blocks: the work-group's previous velocity with its face halos
    // goes through LDS for the WALE stencil (vel_tile_load / _store), and in
    // FP32 the output tile is halved so that both fit 5 work-groups per CU.
    // The tile's loads are issued before the pull and stored after it, so
    // its barrier overlaps the pull's loads.  Measured on the 52.5M-cell cube
    // case, finest collide: FP32 -1.9 %, FP16 -3.1 % (with the tile loads
    // issued after the pull, FP16 lost 1.5 % to the barrier).
    constexpr bool VTILE = STAGE;
    constexpr int ROWS = ST == 0 ? 14 : 27;
    __shared__ __attribute__((aligned(16))) Stored<ST> sst[STAGE ? 4 * ROWS * 64 : 8];
    __shared__ float vt[VTILE ? 3 * VT_SIZE : 1];
    const LevelArgs& L = las[wk.li];
    const int b = wk.item;
    constexpr int ESZ = ST == 0 ? 4 : 2;
    const int c = (half << 8) | tid;
    // the neighbour-row loads go out together with the case record's
    build_pull_table<OffT>(wk.nbr_row, b, half, (OffT)ESZ, ptab, smask, tid);
    build_vel_table<OffT>(wk.nbr_row, L.nb, b, half, vtab, tid);
    const CaseState S = case_state(states + wk.ci);
    __syncthreads();
    if (!live(L, S, j)) return;
    const CollideArgs a = collide_view(L, S, j);
    const int x = c & 7, y = (c >> 3) & 7, z = c >> 6;
    const int wv = tid >> 6;
    const int gx = a.borg[b * 3 + 0] + x + 1;
    const int gy = a.borg[b * 3 + 1] + y + 1;
    const int gz = a.borg[b * 3 + 2] + z + 1;
    const size_t i4 = (size_t)b * NC + c;
    bool is_obs = false;
    if constexpr (ST != 0) is_obs = __builtin_nontemporal_load(&gp(a.obst)[i4]) != 0;

    // STEP 1: pull (kernel_cumulant.jl:86-168).  All 27 loads are issued
    // before any is used, each one table read and one add.
    const int oxp = x == 0 ? -1 : 0, oxm = x == 7 ? 1 : 0;
    const int oyp = y == 0 ? -1 : 0, oym = y == 7 ? 1 : 0;
    const OffT cb = (OffT)c * (OffT)ESZ;
    VelTile vtl;
    if constexpr (VTILE) vtl = vel_tile_load<OffT>(a.vel_in, vtab, half, c, tid);
    float A[27];
#pragma unroll
    for (int k = 0; k < 27; k++) {
        const int cx = CX(k), cy = CY(k), cz = CZ(k);
        const int ox = cx == 1 ? oxp : (cx == -1 ? oxm : 0);
        const int oy = cy == 1 ? oyp : (cy == -1 ? oym : 0);
        const OffT off = ptab[wv * 27 + 9 * (cz + 1) + (ox + 1) + 3 * (oy + 1)] + cb +
                         (OffT)((512 * k - (cx + 8 * cy + 64 * cz)) * ESZ);
        A[k] = Codec<ST>::loadb(a.f_in, off, WK(k));
    }
    // WALE stencil of the previous velocity, loaded together with the populations.
    float VN[6][3];
    if constexpr (VTILE) {
        vel_tile_store(vtl, c, vt, tid);
        __syncthreads();
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const int dx = n == 0 ? 1 : (n == 1 ? -1 : 0);
            const int dy = n == 2 ? 1 : (n == 3 ? -1 : 0);
            const int dz = n == 4 ? 1 : (n == 5 ? -1 : 0);
#pragma unroll
            for (int d = 0; d < 3; d++) VN[n][d] = vt[vt_index(d, wv + dz, y + dy, x + dx)];
        }
    } else {
        const OffT own = vtab[wv * 8 + 6];
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const int dx = n == 0 ? 1 : (n == 1 ? -1 : 0);
            const int dy = n == 2 ? 1 : (n == 3 ? -1 : 0);
            const int dz = n == 4 ? 1 : (n == 5 ? -1 : 0);
            const bool edge = n == 0 ? x == 7 : n == 1 ? x == 0 : n == 2 ? y == 7 : n == 3 ? y == 0 : true;
            const OffT off = (edge ? vtab[wv * 8 + n] : own) + (OffT)(4 * (c + dx + 8 * dy + 64 * dz));
            const GF* p = gaddr(a.vel_in, off);
            VN[n][0] = p[0];
            VN[n][1] = p[NC];
            VN[n][2] = p[2 * NC];
        }
    }
    // lanes whose pull crosses into a missing block (domain boundary)
    // single-use inputs: non-temporal loads keep L2 for the neighbour pulls
    // FP32: a block with no solid, sponge or near-wall cell skips those loads
    // (the values they would give: not solid, sponge 0, no wall model); the
    // branch comes after the pull and stencil loads so that it does not hold
    // them back.  Measured on the cube case: FP32 finest collide -2 %; the
    // 16-bit forms, limited by loads in flight rather than bytes, lose 1.4 to
    // 2.3 % to the extra branch and keep the unconditional loads (bfl = 7).
    uint32_t bfl = 7u;
    if constexpr (ST == 0) {
        bfl = a.bflags[b];
        is_obs = (bfl & 4u) ? __builtin_nontemporal_load(&gp(a.obst)[i4]) != 0 : false;
    }
    const float sp = (bfl & 1u) ? __builtin_nontemporal_load(&gp(a.sponge)[i4]) : 0.0f;
    const float wdist =
        (a.wall_model == 1 && (bfl & 2u)) ? __builtin_nontemporal_load(&gp(a.wall_dist)[i4]) : 0.0f;
    if constexpr (BND) {
        uint32_t miss = 0u;
        const uint32_t m = wave_miss(smask, wv);
        if (m != 0u) {
#pragma unroll
            for (int k = 0; k < 27; k++) {
                const int cx = CX(k), cy = CY(k), cz = CZ(k);
                const int ox = cx == 1 ? oxp : (cx == -1 ? oxm : 0);
                const int oy = cy == 1 ? oyp : (cy == -1 ? oym : 0);
                miss |= ((m >> (9 * (cz + 1) + (ox + 1) + 3 * (oy + 1))) & 1u) << k;
            }
        }
        // Boundary cascade for sources in no block: one inlined copy of the
        // cascade in a rolled loop over k, results passed through LDS (a
        // register array cannot be indexed by a run-time k without spilling).
        if (miss != 0u) {
#pragma unroll 1
            for (int k = 0; k < 27; k++)
                if (miss & (1u << k))
                    sfix[k * 256 + tid] =
                        bc_value<ST>(a, k, b, c, gx - CX(k), gy - CY(k), gz - CZ(k), gy, gz);
#pragma unroll
            for (int k = 0; k < 27; k++)
                if (miss & (1u << k)) A[k] = sfix[k * 256 + tid];
        }
    }
    float rho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
#pragma unroll
    for (int k = 0; k < 27; k++) {
        const int cx = CX(k), cy = CY(k), cz = CZ(k);
        const float val = A[k];
        rho += val;
        if (cx == 1) jx += val;
        if (cx == -1) jx += -val;
        if (cy == 1) jy += val;
        if (cy == -1) jy += -val;
        if (cz == 1) jz += val;
        if (cz == -1) jz += -val;
    }

    void* const fo = a.f_out;
    GF* __restrict__ vo = gp(a.vel_out);
    GF* __restrict__ ro = gp(a.rho_out);
    if (is_obs) {
        // simple bounce-back of the streamed values (kernel_cumulant.jl:173-185)
        const size_t iv = (size_t)b * 3 * NC + c;
        vo[iv] = 0.0f;
        vo[iv + NC] = 0.0f;
        vo[iv + 2 * NC] = 0.0f;
        if (a.store_rho) ro[i4] = 1.0f;
        if constexpr (STAGE) {
#pragma unroll
            for (int k = 0; k < 13; k++) {
                const float t = A[k];
                A[k] = A[OPP(k)];
                A[OPP(k)] = t;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 27; k++) Codec<ST>::stream(fo, fidx(b, k, c), A[OPP(k)], WK(k));
            return;
        }
    }
    // (staged form: solid cells skip the collision and join the store)
    if (!is_obs) {

        // STEP 2: macroscopic + sponge + wall model (kernel_cumulant.jl:190-259)
        rho = fmaxf(rho, 0.01f);
        const float inv_rho = 1.0f / rho;
        float ux = jx * inv_rho, uy = jy * inv_rho, uz = jz * inv_rho;
        if (sp > 0.0f) {
            rho = rho * (1.0f - sp) + 1.0f * sp;
            ux = ux * (1.0f - sp) + a.u_inlet * sp;
            uy = uy * (1.0f - sp);
            uz = uz * (1.0f - sp);
            if (a.sponge_blend == 1) {
#pragma unroll
                for (int k = 0; k < 27; k++) {
                    float ft = feq_auto(1.0f, a.u_inlet, 0.0f, 0.0f, WK(k), (float)CX(k), (float)CY(k), (float)CZ(k),
                                        a.comp);
                    A[k] = A[k] * (1.0f - sp) + ft * sp;
                }
            }
        }
        float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f;
        if (a.wall_model == 1) {
            const float d = wdist;
            if (d > 0.0f && d < 10.0f) {
                float umag = sqrtf(ux * ux + uy * uy + uz * uz);
                float nu_visc = (a.tau - 0.5f) / 3.0f;
                if (umag > 1.0e-6f && nu_visc > 1.0e-10f) {
                    float u_tau = umag * powf(nu_visc / (d * umag + 1.0e-10f), 1.0f / 7.0f) *
                                  powf(2.0f * 8.3f, -1.0f / 7.0f);
                    u_tau = fmaxf(u_tau, 1.0e-6f);
                    float y_p = u_tau * d / nu_visc;
                    if (y_p > 11.81f) {
                        float upl = (1.0f / 0.41f) * logf(y_p) + 5.2f;
                        if (upl > 0.1f) {
                            u_tau = u_tau * ((umag / u_tau) / upl);
                            u_tau = fmaxf(u_tau, 1.0e-6f);
                        }
                    }
                    float tau_wall = rho * u_tau * u_tau;
                    float tau_res = rho * nu_visc * (umag / d);
                    if (tau_wall > tau_res) {
                        float fm = (tau_wall - tau_res) / d;
                        Fx = -fm * ux / umag;
                        Fy = -fm * uy / umag;
                        Fz = -fm * uz / umag;
                    }
                }
            }
        }
        const float uxe = ux + 0.5f * Fx * inv_rho;
        const float uye = uy + 0.5f * Fy * inv_rho;
        const float uze = uz + 0.5f * Fz * inv_rho;
        {
            const size_t iv = (size_t)b * 3 * NC + c;
            __builtin_nontemporal_store(ux, &vo[iv]);
            __builtin_nontemporal_store(uy, &vo[iv + NC]);
            __builtin_nontemporal_store(uz, &vo[iv + 2 * NC]);
            if (a.store_rho) __builtin_nontemporal_store(rho, &ro[i4]);
        }

        // STEP 3: WALE on the previous velocity (kernel_cumulant.jl:264-310)
        float om1;
        {
            const float uE = VN[0][0], vE = VN[0][1], wE = VN[0][2], uW = VN[1][0], vW = VN[1][1], wW = VN[1][2];
            const float uN = VN[2][0], vN = VN[2][1], wN = VN[2][2], uS = VN[3][0], vS = VN[3][1], wS = VN[3][2];
            const float uT = VN[4][0], vT = VN[4][1], wT = VN[4][2], uB = VN[5][0], vB = VN[5][1], wB = VN[5][2];
            float g11 = 0.5f * (uE - uW), g12 = 0.5f * (uN - uS), g13 = 0.5f * (uT - uB);
            float g21 = 0.5f * (vE - vW), g22 = 0.5f * (vN - vS), g23 = 0.5f * (vT - vB);
            float g31 = 0.5f * (wE - wW), g32 = 0.5f * (wN - wS), g33 = 0.5f * (wT - wB);
            float q11 = g11 * g11 + g12 * g21 + g13 * g31;
            float q12 = g11 * g12 + g12 * g22 + g13 * g32;
            float q13 = g11 * g13 + g12 * g23 + g13 * g33;
            float q21 = g21 * g11 + g22 * g21 + g23 * g31;
            float q22 = g21 * g12 + g22 * g22 + g23 * g32;
            float q23 = g21 * g13 + g22 * g23 + g23 * g33;
            float q31 = g31 * g11 + g32 * g21 + g33 * g31;
            float q32 = g31 * g12 + g32 * g22 + g33 * g32;
            float q33 = g31 * g13 + g32 * g23 + g33 * g33;
            float tr = q11 + q22 + q33;
            float trt = tr / 3.0f;
            float Sd11 = q11 - trt, Sd22 = q22 - trt, Sd33 = q33 - trt;
            float Sd12 = 0.5f * (q12 + q21), Sd13 = 0.5f * (q13 + q31), Sd23 = 0.5f * (q23 + q32);
            float S12 = 0.5f * (g12 + g21), S13 = 0.5f * (g13 + g31), S23 = 0.5f * (g23 + g32);
            float OP1 = Sd11 * Sd11 + Sd22 * Sd22 + Sd33 * Sd33 + 2.0f * (Sd12 * Sd12 + Sd13 * Sd13 + Sd23 * Sd23);
            float OP2 = g11 * g11 + g22 * g22 + g33 * g33 + 2.0f * (S12 * S12 + S13 * S13 + S23 * S23);
            float nu_eddy = 0.0f;
            if (OP1 > 1.0e-12f) {
                float OP1_32 = OP1 * sqrtf(OP1);
                float OP2_52 = OP2 * OP2 * sqrtf(fmaxf(OP2, 1.0e-12f));
                float den = OP2_52 + OP1 * sqrtf(sqrtf(fmaxf(OP1, 1.0e-12f)));
                if (den > 1.0e-12f) nu_eddy = (a.c_wale * a.c_wale) * OP1_32 / den;
            }
            nu_eddy = fmaxf(nu_eddy, a.nu_sgs_bg);
            float nu_mol = (a.tau - 0.5f) / 3.0f;
            float nu_eff = nu_mol + nu_eddy;
            float tau_eff = 0.5f + 3.0f * nu_eff;
            om1 = 1.0f / fmaxf(tau_eff, 0.500001f);
        }
        const float omb = a.omega_bulk, om3 = a.omega3, om4 = a.omega4;
        float om4d = om4;
        if (a.adaptive == 1) {
            float t = om1 - 2.0f;
            float den = a.lambda * om1 + 1.0e-10f;
            float Av = -(t * t) / den;
            om4d = Av > 2.0f ? 2.0f : (Av < 0.01f ? 0.01f : Av);
        }

        // STEP 4: chimera forward (kernel_cumulant.jl:330-371)
#pragma unroll
        for (int iz = 0; iz < 3; iz++)
#pragma unroll
            for (int iy = 0; iy < 3; iy++) {
                const int bb = 3 * iy + 9 * iz;
                chim_fwd(A[bb], A[bb + 1], A[bb + 2], uxe);
            }
#pragma unroll
        for (int iz = 0; iz < 3; iz++)
#pragma unroll
            for (int ma = 0; ma < 3; ma++) {
                const int i1 = ma + 9 * iz;
                chim_fwd(A[i1], A[i1 + 3], A[i1 + 6], uye);
            }
#pragma unroll
        for (int mb = 0; mb < 3; mb++)
#pragma unroll
            for (int ma = 0; ma < 3; ma++) {
                const int i1 = ma + 3 * mb;
                chim_fwd(A[i1], A[i1 + 9], A[i1 + 18], uze);
            }

        // STEP 5-8: cumulants, relaxation, reconstruction (kernel_cumulant.jl:376-519)
        {
            const float k000 = A[0];
            const float ir = 1.0f / fmaxf(k000, 0.01f);
            const float c200 = A[2] * ir, c020 = A[6] * ir, c002 = A[18] * ir;
            const float c110 = A[4] * ir, c101 = A[10] * ir, c011 = A[12] * ir;
            const float c210 = A[5] * ir, c120 = A[7] * ir, c201 = A[11] * ir;
            const float c021 = A[15] * ir, c102 = A[19] * ir, c012 = A[21] * ir;
            const float c111 = A[13] * ir;
            const float c220n = A[8] * ir, c202n = A[20] * ir, c022n = A[24] * ir;
            const float c211n = A[14] * ir, c121n = A[16] * ir, c112n = A[22] * ir;
            const float C220 = c220n - c200 * c020 - 2.0f * c110 * c110;
            const float C202 = c202n - c200 * c002 - 2.0f * c101 * c101;
            const float C022 = c022n - c020 * c002 - 2.0f * c011 * c011;
            const float C211 = c211n - c200 * c011 - 2.0f * c110 * c101;
            const float C121 = c121n - c020 * c101 - 2.0f * c110 * c011;
            const float C112 = c112n - c002 * c110 - 2.0f * c101 * c011;
            const float cs2 = 0.333333333f;
            const float Dt = c200 + c020 + c002;
            const float Dteq = 3.0f * cs2;
            const float Dts = (1.0f - omb) * Dt + omb * Dteq;
            const float D1 = c200 - c020, D2 = c200 - c002;
            const float D1s = (1.0f - om1) * D1, D2s = (1.0f - om1) * D2;
            const float c200s = (Dts + D1s + D2s) / 3.0f;
            const float c020s = (Dts - 2.0f * D1s + D2s) / 3.0f;
            const float c002s = (Dts + D1s - 2.0f * D2s) / 3.0f;
            float c110s = (1.0f - om1) * c110, c101s = (1.0f - om1) * c101, c011s = (1.0f - om1) * c011;
            if (a.limiter == 1) {
                float b110 = sqrtf(fmaxf(c200s * c020s, 0.0f));
                float b101 = sqrtf(fmaxf(c200s * c002s, 0.0f));
                float b011 = sqrtf(fmaxf(c020s * c002s, 0.0f));
                c110s = c110s > b110 ? b110 : (c110s < -b110 ? -b110 : c110s);
                c101s = c101s > b101 ? b101 : (c101s < -b101 ? -b101 : c101s);
                c011s = c011s > b011 ? b011 : (c011s < -b011 ? -b011 : c011s);
            }
            const float c210s = (1.0f - om3) * c210, c120s = (1.0f - om3) * c120, c201s = (1.0f - om3) * c201;
            const float c021s = (1.0f - om3) * c021, c102s = (1.0f - om3) * c102, c012s = (1.0f - om3) * c012;
            const float c111s = (1.0f - om3) * c111;
            const float C220s = (1.0f - om4d) * C220, C202s = (1.0f - om4d) * C202, C022s = (1.0f - om4d) * C022;
            const float C211s = (1.0f - om4) * C211, C121s = (1.0f - om4) * C121, C112s = (1.0f - om4) * C112;
            const float c220s = C220s + c200s * c020s + 2.0f * c110s * c110s;
            const float c202s = C202s + c200s * c002s + 2.0f * c101s * c101s;
            const float c022s = C022s + c020s * c002s + 2.0f * c011s * c011s;
            const float c211s = C211s + c200s * c011s + 2.0f * c110s * c101s;
            const float c121s = C121s + c020s * c101s + 2.0f * c110s * c011s;
            const float c112s = C112s + c002s * c110s + 2.0f * c101s * c011s;
            const float c221s = c200s * c021s + c020s * c201s + 4.0f * c110s * c111s + 2.0f * c101s * c120s +
                                2.0f * c011s * c210s;
            const float c212s = c200s * c012s + c002s * c210s + 2.0f * c110s * c102s + 4.0f * c101s * c111s +
                                2.0f * c011s * c201s;
            const float c122s = c020s * c102s + c002s * c120s + 2.0f * c110s * c012s + 4.0f * c011s * c111s +
                                2.0f * c101s * c021s;
            float c222s = c200s * C022s + c020s * C202s + c002s * C220s +
                          4.0f * (c110s * C112s + c101s * C121s + c011s * C211s);
            c222s += 2.0f * (c210s * c012s + c201s * c021s + c120s * c102s) + 4.0f * c111s * c111s;
            c222s += c200s * c020s * c002s +
                     2.0f * (c200s * c011s * c011s + c020s * c101s * c101s + c002s * c110s * c110s) +
                     8.0f * c110s * c101s * c011s;
            const float r = k000;
            A[0] = r;
            A[1] = 0.0f; A[3] = 0.0f; A[9] = 0.0f;
            A[2] = r * c200s; A[6] = r * c020s; A[18] = r * c002s;
            A[4] = r * c110s; A[10] = r * c101s; A[12] = r * c011s;
            A[5] = r * c210s; A[7] = r * c120s; A[11] = r * c201s;
            A[15] = r * c021s; A[19] = r * c102s; A[21] = r * c012s;
            A[13] = r * c111s;
            A[8] = r * c220s; A[20] = r * c202s; A[24] = r * c022s;
            A[14] = r * c211s; A[16] = r * c121s; A[22] = r * c112s;
            A[17] = r * c221s; A[23] = r * c212s; A[25] = r * c122s;
            A[26] = r * c222s;
        }

        // STEP 9: chimera backward (kernel_cumulant.jl:525-559)
#pragma unroll
        for (int mb = 0; mb < 3; mb++)
#pragma unroll
            for (int ma = 0; ma < 3; ma++) {
                const int i1 = ma + 3 * mb;
                chim_bwd(A[i1], A[i1 + 9], A[i1 + 18], uze);
            }
#pragma unroll
        for (int iz = 0; iz < 3; iz++)
#pragma unroll
            for (int ma = 0; ma < 3; ma++) {
                const int i1 = ma + 9 * iz;
                chim_bwd(A[i1], A[i1 + 3], A[i1 + 6], uye);
            }
#pragma unroll
        for (int iz = 0; iz < 3; iz++)
#pragma unroll
            for (int iy = 0; iy < 3; iy++) {
                const int bb = 3 * iy + 9 * iz;
                chim_bwd(A[bb], A[bb + 1], A[bb + 2], uxe);
            }

        // positivity limiter (kernel_cumulant.jl:564-599)
        if (a.limiter == 2) {
            float fmin_ = A[0];
#pragma unroll
            for (int k = 1; k < 27; k++) fmin_ = fminf(fmin_, A[k]);
            if (fmin_ < 0.0f) {
                float alpha = 1.0f;
#pragma unroll
                for (int k = 0; k < 27; k++) {
                    float fe = feq_auto(rho, uxe, uye, uze, WK(k), (float)CX(k), (float)CY(k), (float)CZ(k), a.comp);
                    float fn = A[k] - fe;
                    if (fn < 0.0f && fe > 1.0e-15f) {
                        float ak = fe / (fe - fn + 1.0e-15f);
                        alpha = fminf(alpha, ak);
                    }
                }
                alpha = fmaxf(alpha, 0.0f);
                // The equilibria are computed again rather than kept from
                // the loop above: kept, they and their terms stay live across
                // both loops and spill (the interior kernel spilled 63 VGPRs,
                // 8 without them, and 6 waves per SIMD became the faster
                // occupancy for 16-bit storage).  Opaque copies of the inputs
                // stop the compiler from reusing the first loop's values.
                float r2 = rho, ux2 = uxe, uy2 = uye, uz2 = uze;
                asm volatile("" : "+v"(r2), "+v"(ux2), "+v"(uy2), "+v"(uz2));
#pragma unroll
                for (int k = 0; k < 27; k++) {
                    float fe = feq_auto(r2, ux2, uy2, uz2, WK(k), (float)CX(k), (float)CY(k), (float)CZ(k), a.comp);
                    A[k] = fe + alpha * (A[k] - fe);
                }
            }
        }

        // STEP 10: Guo forcing + store (kernel_cumulant.jl:604-623).  With F = 0
        // the forcing term is a signed zero and leaves every population unchanged.
        if (Fx != 0.0f || Fy != 0.0f || Fz != 0.0f) {
            const float pre = 1.0f - 0.5f * om1;
#pragma unroll
            for (int k = 0; k < 27; k++) {
                const float cxf = (float)CX(k), cyf = (float)CY(k), czf = (float)CZ(k);
                float cu = cxf * uxe + cyf * uye + czf * uze;
                float ft = WK(k) * 3.0f *
                           ((cxf - ux + 3.0f * cu * cxf) * Fx + (cyf - uy + 3.0f * cu * cyf) * Fy +
                            (czf - uz + 3.0f * cu * czf) * Fz);
                if constexpr (STAGE)
                    A[k] = A[k] + pre * ft;
                else
                    Codec<ST>::stream(fo, fidx(b, k, c), A[k] + pre * ft, WK(k));
            }
        } else if constexpr (!STAGE) {
#pragma unroll
            for (int k = 0; k < 27; k++) Codec<ST>::stream(fo, fidx(b, k, c), A[k], WK(k));
        }
    }  // !is_obs
    if constexpr (STAGE) stream_tile<ST, ROWS>(fo, b, c, A, sst + wv * (ROWS * 64));
}

template <int ST, class OffT, bool BND, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))
collide_kernel(const LevelArgs* __restrict__ las, const CaseState* __restrict__ states,
               const Work* __restrict__ work, int32_t n_tiles, int64_t j, int32_t sched) {
    Work wk;
    int half;
    if (!schedule(work, n_tiles, sched, wk, half)) return;
    collide_tile<ST, OffT, BND>(las, states, wk, half, j, (int)threadIdx.x);
}

// Persistent form: work-group g takes tiles g, g + G, g + 2G, ... in order.
// All waves of a work-group make the same passes, and a barrier ends each
// pass, so the LDS tables of a tile are not rewritten while any wave still
// reads them.  A tile's early exits are uniform over the work-group (a case
// past its last step) or come after the tile's last barrier (a solid lane's
// bounce-back in the boundary form), so every wave reaches every barrier.
template <int ST, class OffT, bool BND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 5)))
collide_kernel_persistent(const LevelArgs* __restrict__ las, const CaseState* __restrict__ states,
                          const Work* __restrict__ work, int32_t n_tiles, int64_t j) {
    for (unsigned logical = blockIdx.x; logical < 2u * (unsigned)n_tiles; logical += gridDim.x) {
        const Work wk = work[logical >> 1];
        // The work-item index is made opaque on every pass, so that nothing
        // derived from it (lane geometry, table addresses) is hoisted out of
        // the loop and kept live across the whole tile.
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        collide_tile<ST, OffT, BND>(las, states, wk, (int)(logical & 1u), j, tid);
        __syncthreads();
    }
}

inline unsigned padded_groups(int nb) { return ((unsigned)nb * 2u + 7u) & ~7u; }

// Work-groups of a persistent launch: 5 per CU (one wave per SIMD each, at
// 5 waves per SIMD), no more than there are tiles.
inline unsigned persistent_groups(int n_tiles) {
    static std::atomic<int> cus[64] = {};  // CU count per device, queried once
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int n = cus[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev].store(n, std::memory_order_relaxed);
    }
    const unsigned full = 5u * (unsigned)n;
    const unsigned need = padded_groups(n_tiles);
    return need < full ? need : full;
}

template <int ST, class OffT>
void collide_launch(const LevelArgs* las, const CaseState* states, const Work* work, int32_t n_tiles, int64_t j,
                    int boundary, int sched, int waves, hipStream_t s) {
    const dim3 t(256);
    if (sched == LB_SCHED_PERSISTENT) {
        const dim3 g(persistent_groups(n_tiles));
        if (boundary)
            hipLaunchKernelGGL((collide_kernel_persistent<ST, OffT, true>), g, t, 0, s, las, states, work, n_tiles, j);
        else
            hipLaunchKernelGGL((collide_kernel_persistent<ST, OffT, false>), g, t, 0, s, las, states, work, n_tiles, j);
        return;
    }
    const dim3 g(padded_groups(n_tiles));
    if (boundary)
        hipLaunchKernelGGL((collide_kernel<ST, OffT, true, 5>), g, t, 0, s, las, states, work, n_tiles, j, sched);
    else if (waves == 6)
        hipLaunchKernelGGL((collide_kernel<ST, OffT, false, 6>), g, t, 0, s, las, states, work, n_tiles, j, sched);
    else
        hipLaunchKernelGGL((collide_kernel<ST, OffT, false, 5>), g, t, 0, s, las, states, work, n_tiles, j, sched);
}

// boundary = 0 requires every block of the work list to have all 26
// neighbours (blocks or ghost blocks).  waves: occupancy of the interior
// kernel (5 or 6 waves per SIMD).
hipError_t collide(const LevelArgs* las, const CaseState* states, const Work* work, int32_t n_tiles, int64_t j,
                   int storage, int wide, int boundary, int sched, int waves, hipStream_t s) {
    if (n_tiles <= 0) return hipSuccess;
    if (wide) {
        if (storage == 1) collide_launch<1, uint64_t>(las, states, work, n_tiles, j, boundary, sched, waves, s);
        else if (storage == 2) collide_launch<2, uint64_t>(las, states, work, n_tiles, j, boundary, sched, waves, s);
        else collide_launch<0, uint64_t>(las, states, work, n_tiles, j, boundary, sched, waves, s);
    } else {
        if (storage == 1) collide_launch<1, uint32_t>(las, states, work, n_tiles, j, boundary, sched, waves, s);
        else if (storage == 2) collide_launch<2, uint32_t>(las, states, work, n_tiles, j, boundary, sched, waves, s);
        else collide_launch<0, uint32_t>(las, states, work, n_tiles, j, boundary, sched, waves, s);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Thermal DDF (src/collision/kernel_thermal.jl:14-202)
template <int ST>
__device__ __forceinline__ float thermal_bc(const ThermalArgs& a, int k, int b, int c, int sgx, int sgy, int sgz) {
    float cx = (float)CX(k), cy = (float)CY(k), cz = (float)CZ(k), w = wk_rt(k);
    if (sgx < 1) return geq(a.T_inlet, a.u_inlet, 0.0f, 0.0f, w, cx, cy, cz);
    if (sgx > a.nxg) return geq(a.T_inlet, a.u_inlet, 0.0f, 0.0f, w, cx, cy, cz);
    if (sgy < 1 || sgy > a.nyg) return Codec<ST>::load(a.g_in, fidx(b, MIRY(k), c), w);
    if (sgz < 1 || sgz > a.nzg) return Codec<ST>::load(a.g_in, fidx(b, MIRZ(k), c), w);
    // in-domain sources without a block are ghost blocks (thermal ghost fill)
    return w * a.T_inlet;
}

template <int ST, class OffT, bool BND>
__global__ void __launch_bounds__(256) thermal_kernel(const LevelArgs* __restrict__ las,
                                                      const CaseState* __restrict__ states,
                                                      const Work* __restrict__ work, int32_t n_tiles, int64_t j,
                                                      int32_t sched) {
    __shared__ OffT ptab[108];
    __shared__ uint64_t smask[2];
    __shared__ float sfix[BND ? 27 * 256 : 1];
    // interior blocks: outputs leave through stream_tile, as in the collide
    constexpr bool STAGE = !BND;
    __shared__ __attribute__((aligned(16))) Stored<ST> sst[STAGE ? 4 * 27 * 64 : 8];
    Work wk;
    int half;
    if (!schedule(work, n_tiles, sched, wk, half)) return;
    const LevelArgs& L = las[wk.li];
    const int b = wk.item;
    const int c = (half << 8) | threadIdx.x;
    constexpr int ESZ = ST == 0 ? 4 : 2;
    build_pull_table<OffT>(wk.nbr_row, b, half, (OffT)ESZ, ptab, smask, (int)threadIdx.x);
    const CaseState S = case_state(states + wk.ci);
    __syncthreads();
    if (!live(L, S, j)) return;
    const ThermalArgs a = thermal_view(L, S, j);
    const int x = c & 7, y = (c >> 3) & 7, z = c >> 6;
    const int wv = threadIdx.x >> 6;
    const int gx = a.borg[b * 3 + 0] + x + 1;
    const int gy = a.borg[b * 3 + 1] + y + 1;
    const int gz = a.borg[b * 3 + 2] + z + 1;
    const size_t i4 = (size_t)b * NC + c;
    const size_t iv = (size_t)b * 3 * NC + c;
    float ux = gp(a.vel)[iv], uy = gp(a.vel)[iv + NC], uz = gp(a.vel)[iv + 2 * NC];
    const int oxp = x == 0 ? -1 : 0, oxm = x == 7 ? 1 : 0;
    const int oyp = y == 0 ? -1 : 0, oym = y == 7 ? 1 : 0;
    const OffT cb = (OffT)c * (OffT)ESZ;
    float G[27];
#pragma unroll
    for (int k = 0; k < 27; k++) {
        const int cx = CX(k), cy = CY(k), cz = CZ(k);
        const int ox = cx == 1 ? oxp : (cx == -1 ? oxm : 0);
        const int oy = cy == 1 ? oyp : (cy == -1 ? oym : 0);
        const OffT off = ptab[wv * 27 + 9 * (cz + 1) + (ox + 1) + 3 * (oy + 1)] + cb +
                         (OffT)((512 * k - (cx + 8 * cy + 64 * cz)) * ESZ);
        G[k] = Codec<ST>::loadb(a.g_in, off, WK(k));
    }
    const bool is_obs = gp(a.obst)[i4] != 0;
    if constexpr (BND) {
        // boundary blocks only: sources in no block take the boundary cascade
        uint32_t miss = 0u;
        const uint32_t m = wave_miss(smask, wv);
        if (m != 0u) {
#pragma unroll
            for (int k = 0; k < 27; k++) {
                const int cx = CX(k), cy = CY(k), cz = CZ(k);
                const int ox = cx == 1 ? oxp : (cx == -1 ? oxm : 0);
                const int oy = cy == 1 ? oyp : (cy == -1 ? oym : 0);
                miss |= ((m >> (9 * (cz + 1) + (ox + 1) + 3 * (oy + 1))) & 1u) << k;
            }
        }
        if (miss != 0u) {
#pragma unroll 1
            for (int k = 0; k < 27; k++)
                if (miss & (1u << k))
                    sfix[k * 256 + threadIdx.x] = thermal_bc<ST>(a, k, b, c, gx - CX(k), gy - CY(k), gz - CZ(k));
#pragma unroll
            for (int k = 0; k < 27; k++)
                if (miss & (1u << k)) G[k] = sfix[k * 256 + threadIdx.x];
        }
    }
    float Tsum = 0.0f;
#pragma unroll
    for (int k = 0; k < 27; k++) Tsum += G[k];
    if (is_obs) {
#pragma unroll
        for (int k = 0; k < 27; k++) {
            float gopp = Codec<ST>::load(a.g_in, fidx(b, OPP(k), c), WK(k));  // local, unstreamed (kernel_thermal.jl:140)
            float go;
            if (a.adiabatic == 1) {
                go = gopp;
            } else {
                float gw = geq(a.T_wall, 0.0f, 0.0f, 0.0f, WK(k), (float)CX(k), (float)CY(k), (float)CZ(k));
                go = 2.0f * gw - gopp;
            }
            if constexpr (STAGE)
                G[k] = go;
            else
                Codec<ST>::stream(a.g_out, fidx(b, k, c), go, WK(k));
        }
        if (a.store_T) a.T_out[i4] = a.adiabatic == 1 ? Tsum : a.T_wall;
        if constexpr (!STAGE) return;
    } else {
        float Tl = fmaxf(Tsum, 0.01f);
        const float sp = a.sponge[i4];
        if (sp > 0.0f) {
            Tl = Tl * (1.0f - sp) + a.T_inlet * sp;
            ux = ux * (1.0f - sp) + a.u_inlet * sp;
            uy = uy * (1.0f - sp);
            uz = uz * (1.0f - sp);
        }
        if (a.store_T) a.T_out[i4] = Tl;
        const float omg = 1.0f / fmaxf(a.tau_g, 0.500001f);
#pragma unroll
        for (int k = 0; k < 27; k++) {
            float gs = G[k];
            if (sp > 0.0f && a.sponge_blend == 1) {
                float gt = geq(a.T_inlet, a.u_inlet, 0.0f, 0.0f, WK(k), (float)CX(k), (float)CY(k), (float)CZ(k));
                gs = gs * (1.0f - sp) + gt * sp;
            }
            float ge = geq(Tl, ux, uy, uz, WK(k), (float)CX(k), (float)CY(k), (float)CZ(k));
            if constexpr (STAGE)
                G[k] = gs - omg * (gs - ge);
            else
                Codec<ST>::stream(a.g_out, fidx(b, k, c), gs - omg * (gs - ge), WK(k));
        }
    }  // !is_obs
    if constexpr (STAGE) stream_tile<ST, 27>(a.g_out, b, c, G, sst + wv * (27 * 64));
}

// Interior blocks (every neighbour present, ghosts included) run the variant
// without the boundary cascade and its 27 KiB LDS buffer, which would limit
// occupancy to 5 work-groups per CU.
template <int ST, class OffT>
void thermal_launch2(const LevelArgs* las, const CaseState* states, const Work* work, int32_t n_tiles, int64_t j,
                     int boundary, int sched, hipStream_t s) {
    const dim3 g(padded_groups(n_tiles)), t(256);
    if (boundary) hipLaunchKernelGGL((thermal_kernel<ST, OffT, true>), g, t, 0, s, las, states, work, n_tiles, j, sched);
    else hipLaunchKernelGGL((thermal_kernel<ST, OffT, false>), g, t, 0, s, las, states, work, n_tiles, j, sched);
}
template <int ST>
void thermal_launch(const LevelArgs* las, const CaseState* states, const Work* work, int32_t n_tiles, int64_t j,
                    int wide, int boundary, int sched, hipStream_t s) {
    if (wide) thermal_launch2<ST, uint64_t>(las, states, work, n_tiles, j, boundary, sched, s);
    else thermal_launch2<ST, uint32_t>(las, states, work, n_tiles, j, boundary, sched, s);
}
hipError_t thermal(const LevelArgs* las, const CaseState* states, const Work* work, int32_t n_tiles, int64_t j,
                   int storage, int wide, int boundary, int sched, hipStream_t s) {
    if (n_tiles <= 0) return hipSuccess;
    if (storage == 1) thermal_launch<1>(las, states, work, n_tiles, j, wide, boundary, sched, s);
    else if (storage == 2) thermal_launch<2>(las, states, work, n_tiles, j, wide, boundary, sched, s);
    else thermal_launch<0>(las, states, work, n_tiles, j, wide, boundary, sched, s);
    return hipGetLastError();
}

// Parent population component k at the 8 corners (new and, with temporal
// interpolation, old), decoded to FP32.
template <int ST>
__device__ __forceinline__ void ghost_load(const GhostArgs& a, int k, const bool* ok, const int* cb, const int* cc,
                                           bool need_old, float* ln, float* lo) {
    const float w_k = wk_rt(k);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        ln[q] = 0.0f;
        lo[q] = 0.0f;
        if (ok[q]) {
            ln[q] = Codec<ST>::load(a.a_new, fidx(cb[q], k, cc[q]), w_k);
            if (need_old) lo[q] = Codec<ST>::load(a.a_old, fidx(cb[q], k, cc[q]), w_k);
        }
    }
}

// Temporal and trilinear blend of component k for both child substeps,
// rescaled non-equilibrium part, stored into both input halves.
template <int ST>
__device__ __forceinline__ void ghost_blend_store(const GhostArgs& a, int k, const bool* ok, const float* ln,
                                                  const float* lo, bool use_old0, bool use_old1, float wx, float wy,
                                                  float wz, const float* r0, const float* r1, float scale, int gb,
                                                  int c) {
    const float tw0 = 0.0f, tw1 = 0.5f;
    const float w_k = wk_rt(k), cx = (float)CX(k), cy = (float)CY(k), cz = (float)CZ(k);
    float f0[8], f1[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (ok[q]) {
            const float fn = ln[q], fo_ = lo[q];
            f0[q] = use_old0 ? fo_ * (1.0f - tw0) + fn * tw0 : fn;
            f1[q] = use_old1 ? fo_ * (1.0f - tw1) + fn * tw1 : fn;
        } else {
            f0[q] = q == 0 ? w_k : f0[0];
            f1[q] = q == 0 ? w_k : f1[0];
        }
    }
    float rr0, rr1;
    {
        float c00 = f0[0] * (1.0f - wx) + f0[1] * wx;
        float c01 = f0[4] * (1.0f - wx) + f0[5] * wx;
        float c10 = f0[2] * (1.0f - wx) + f0[3] * wx;
        float c11 = f0[6] * (1.0f - wx) + f0[7] * wx;
        float c0 = c00 * (1.0f - wy) + c10 * wy;
        float c1 = c01 * (1.0f - wy) + c11 * wy;
        rr0 = c0 * (1.0f - wz) + c1 * wz;
    }
    {
        float c00 = f1[0] * (1.0f - wx) + f1[1] * wx;
        float c01 = f1[4] * (1.0f - wx) + f1[5] * wx;
        float c10 = f1[2] * (1.0f - wx) + f1[3] * wx;
        float c11 = f1[6] * (1.0f - wx) + f1[7] * wx;
        float c0 = c00 * (1.0f - wy) + c10 * wy;
        float c1 = c01 * (1.0f - wy) + c11 * wy;
        rr1 = c0 * (1.0f - wz) + c1 * wz;
    }
    const float eq0 = a.thermal ? geq(r0[0], r0[1], r0[2], r0[3], w_k, cx, cy, cz)
                                : feq_auto(r0[0], r0[1], r0[2], r0[3], w_k, cx, cy, cz, a.comp);
    const float eq1 = a.thermal ? geq(r1[0], r1[1], r1[2], r1[3], w_k, cx, cy, cz)
                                : feq_auto(r1[0], r1[1], r1[2], r1[3], w_k, cx, cy, cz, a.comp);
    const float neq0 = rr0 - eq0, neq1 = rr1 - eq1;
    Codec<ST>::store(a.f, fidx(gb, k, c), eq0 + neq0 * scale, w_k);
    Codec<ST>::store(a.f2, fidx(gb, k, c), eq1 + neq1 * scale, w_k);
}

// ---------------------------------------------------------------------------
// Ghost-cell interpolation: interpolate_with_rescaling / its thermal twin
// (physics_interpolation.jl:4-127, 135-248) for every direction a ghost cell
// serves.  The two substeps of a child level within one parent step read the
// same parent state and differ only in the temporal weight (0, then 0.5), so
// one pass, run before the first substep, computes both: the tw = 0 values go
// to the ghost slots of the first substep's input half, the tw = 0.5 values to
// the other half, which the first substep does not write (its collide writes
// real blocks only).  The corner geometry and the density/velocity part of
// the trilinear blend depend on the cell only and are computed once; each
// direction then blends its own population component.  Per value, the
// operations are those of the reference function, in its order.
template <int ST>
__global__ void __launch_bounds__(256) ghost_kernel(const LevelArgs* __restrict__ las,
                                                    const CaseState* __restrict__ states,
                                                    const Work* __restrict__ work, int64_t j, int thermal) {
    const Work wk = work[blockIdx.x];
    const LevelArgs& L = las[wk.li];
    const CaseState S = case_state(states + wk.ci);
    if (!live(L, S, j)) return;
    const GhostArgs a = ghost_view(L, S, j, thermal);
    const int64_t i = (int64_t)wk.item + threadIdx.x;
    if (i >= a.n) return;
    const int e0 = a.entries[2 * i];
    const uint32_t mask = (uint32_t)a.entries[2 * i + 1];
    const int gb = e0 >> 9, c = e0 & 511;
    // The fine cell centre lies at (g - 0.5) / 2 in parent cells, g = block
    // origin (a multiple of 8) + local index + 1, so its fractional part is
    // 0.25 for an even local index and 0.75 for an odd one (exactly what
    // (g - 0.5) * 0.5 - floor(...) gives in FP32).  The 8 parent cells come
    // precomputed with the entry.
    const float wx = (c & 1) ? 0.75f : 0.25f;
    const float wy = ((c >> 3) & 1) ? 0.75f : 0.25f;
    const float wz = ((c >> 6) & 1) ? 0.75f : 0.25f;
    int par[8];
    {
        typedef int32_t v4i __attribute__((ext_vector_type(4)));
        const __attribute__((address_space(1))) v4i* pp =
            (const __attribute__((address_space(1))) v4i*)(a.parents + 8 * i);
        const v4i p0 = pp[0], p1 = pp[1];
        par[0] = p0.x, par[1] = p0.y, par[2] = p0.z, par[3] = p0.w;
        par[4] = p1.x, par[5] = p1.y, par[6] = p1.z, par[7] = p1.w;
    }
    // temporal blend with tw = 0 (substep j) and tw = 0.5 (substep j + 1)
    const bool use_old0 = a.temporal == 1 && 0.0f < 0.99f;
    const bool use_old1 = a.temporal == 1 && 0.5f < 0.99f;
    const float tw0 = 0.0f, tw1 = 0.5f;
    int cb[8], cc[8];
    bool ok[8];
    float v0[8][4], v1[8][4];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        ok[q] = par[q] >= 0;
        cb[q] = ok[q] ? (par[q] >> 9) : 0;
        cc[q] = ok[q] ? (par[q] & 511) : 0;
        if (ok[q]) {
            const size_t i4 = (size_t)cb[q] * NC + cc[q];
            const size_t iv = (size_t)cb[q] * 3 * NC + cc[q];
            const float rn = a.s_new[i4], uxn = a.u_new[iv], uyn = a.u_new[iv + NC], uzn = a.u_new[iv + 2 * NC];
            float ro = 0.0f, uxo = 0.0f, uyo = 0.0f, uzo = 0.0f;
            if (use_old0 || use_old1) {
                ro = a.s_old[i4];
                uxo = a.u_old[iv];
                uyo = a.u_old[iv + NC];
                uzo = a.u_old[iv + 2 * NC];
            }
            if (use_old0) {
                v0[q][0] = ro * (1.0f - tw0) + rn * tw0;
                v0[q][1] = uxo * (1.0f - tw0) + uxn * tw0;
                v0[q][2] = uyo * (1.0f - tw0) + uyn * tw0;
                v0[q][3] = uzo * (1.0f - tw0) + uzn * tw0;
            } else {
                v0[q][0] = rn; v0[q][1] = uxn; v0[q][2] = uyn; v0[q][3] = uzn;
            }
            if (use_old1) {
                v1[q][0] = ro * (1.0f - tw1) + rn * tw1;
                v1[q][1] = uxo * (1.0f - tw1) + uxn * tw1;
                v1[q][2] = uyo * (1.0f - tw1) + uyn * tw1;
                v1[q][3] = uzo * (1.0f - tw1) + uzn * tw1;
            } else {
                v1[q][0] = rn; v1[q][1] = uxn; v1[q][2] = uyn; v1[q][3] = uzn;
            }
        } else if (q == 0) {
            v0[0][0] = 1.0f; v0[0][1] = 0.0f; v0[0][2] = 0.0f; v0[0][3] = 0.0f;
            v1[0][0] = 1.0f; v1[0][1] = 0.0f; v1[0][2] = 0.0f; v1[0][3] = 0.0f;
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                v0[q][m] = v0[0][m];
                v1[q][m] = v1[0][m];
            }
        }
    }
    float r0[4], r1[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        {
            float c00 = v0[0][m] * (1.0f - wx) + v0[1][m] * wx;
            float c01 = v0[4][m] * (1.0f - wx) + v0[5][m] * wx;
            float c10 = v0[2][m] * (1.0f - wx) + v0[3][m] * wx;
            float c11 = v0[6][m] * (1.0f - wx) + v0[7][m] * wx;
            float c0 = c00 * (1.0f - wy) + c10 * wy;
            float c1 = c01 * (1.0f - wy) + c11 * wy;
            r0[m] = c0 * (1.0f - wz) + c1 * wz;
        }
        {
            float c00 = v1[0][m] * (1.0f - wx) + v1[1][m] * wx;
            float c01 = v1[4][m] * (1.0f - wx) + v1[5][m] * wx;
            float c10 = v1[2][m] * (1.0f - wx) + v1[3][m] * wx;
            float c11 = v1[6][m] * (1.0f - wx) + v1[7][m] * wx;
            float c0 = c00 * (1.0f - wy) + c10 * wy;
            float c1 = c01 * (1.0f - wy) + c11 * wy;
            r1[m] = c0 * (1.0f - wz) + c1 * wz;
        }
    }
    const float tc = a.tau_coarse - 0.5f;
    const float tf = a.tau_fine - 0.5f;
    float scale = 1.0f;
    if (tc > 1.0e-6f) {
        scale = tf / tc;
        scale = scale > 100.0f ? 100.0f : (scale < 0.01f ? 0.01f : scale);
    }
    const bool need_old = use_old0 || use_old1;
    if constexpr (ST == 0) {
        // FP32: the parent values of the next served direction are loaded
        // while the current one is blended (ghost time -21 % on the cube
        // case).  The 16-bit forms keep the plain loop below, whose higher
        // occupancy hides their short loads better (pipelined: +9 %).
        uint32_t m = mask;
        int k = m ? __builtin_ctz(m) : -1;
        float ln[8], lo[8];
        if (k >= 0) ghost_load<ST>(a, k, ok, cb, cc, need_old, ln, lo);
        while (k >= 0) {
            m &= m - 1u;
            const int kn = m ? __builtin_ctz(m) : -1;
            float nn[8], no[8];
            if (kn >= 0) ghost_load<ST>(a, kn, ok, cb, cc, need_old, nn, no);
            ghost_blend_store<ST>(a, k, ok, ln, lo, use_old0, use_old1, wx, wy, wz, r0, r1, scale, gb, c);
#pragma unroll
            for (int q = 0; q < 8; q++) {
                ln[q] = nn[q];
                lo[q] = no[q];
            }
            k = kn;
        }
    } else {
#pragma unroll 1
        for (int k = 0; k < 27; k++) {
            if (!(mask & (1u << k))) continue;
            float ln[8], lo[8];
            ghost_load<ST>(a, k, ok, cb, cc, need_old, ln, lo);
            ghost_blend_store<ST>(a, k, ok, ln, lo, use_old0, use_old1, wx, wy, wz, r0, r1, scale, gb, c);
        }
    }
}

// 16-bit storage: one work-group per ghost block.  Every interpolation
// corner of the block's cells lies in a 5^3 region of parent cells
// (LevelArgs::gregion), so the work-group first stages that region in LDS:
// the parent density (or temperature) and velocity, new and old, and the
// parent populations of the directions the block serves (gbmask), decoded to
// FP32, 125 cells per direction and time level, the consecutive work-items of
// a load taking consecutive cells of a parent row.  Each work-item then
// interpolates its entries from LDS, with the operations of ghost_kernel in
// its order.  Measured on the 52.5M-cell cube case: FP16 step -1.0 % (A/B on
// one box); in FP32 the per-entry kernel above, which loads the next
// direction's corners while blending the current one, stays faster (fill
// 103 us against 114 us per launch).
constexpr int GREG = 125;  // parent cells of a ghost block's region

template <int ST>
__global__ void __launch_bounds__(256) ghost_kernel_staged(const LevelArgs* __restrict__ las,
                                                    const CaseState* __restrict__ states,
                                                    const Work* __restrict__ work, int64_t j, int thermal) {
    __shared__ float spop[27 * 2 * GREG];  // [direction slot][new, old][region cell]
    __shared__ float smac[8 * GREG];       // [new rho/T, ux, uy, uz, old rho/T, ux, uy, uz][region cell]
    __shared__ int32_t sreg[GREG];
    const Work wk = work[blockIdx.x];
    const LevelArgs& L = las[wk.li];
    const CaseState S = case_state(states + wk.ci);
    if (!live(L, S, j)) return;
    const GhostArgs a = ghost_view(L, S, j, thermal);
    const int g = wk.item;
    const int tid = (int)threadIdx.x;
    const uint32_t bmask = L.gbmask[g];
    // temporal blend with tw = 0 (substep j) and tw = 0.5 (substep j + 1)
    const bool use_old0 = a.temporal == 1 && 0.0f < 0.99f;
    const bool use_old1 = a.temporal == 1 && 0.5f < 0.99f;
    const bool need_old = use_old0 || use_old1;
    const float tw0 = 0.0f, tw1 = 0.5f;
    if (tid < GREG) sreg[tid] = L.gregion[(size_t)g * GREG + tid];
    __syncthreads();
    // stage the parent region
    if (tid < GREG) {
        const int p = sreg[tid];
        float m[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (p >= 0) {
            const size_t i4 = (size_t)(p >> 9) * NC + (p & 511);
            const size_t iv = (size_t)(p >> 9) * 3 * NC + (p & 511);
            m[0] = a.s_new[i4], m[1] = a.u_new[iv], m[2] = a.u_new[iv + NC], m[3] = a.u_new[iv + 2 * NC];
            if (need_old) {
                m[4] = a.s_old[i4];
                m[5] = a.u_old[iv];
                m[6] = a.u_old[iv + NC];
                m[7] = a.u_old[iv + 2 * NC];
            }
        }
#pragma unroll
        for (int q = 0; q < 8; q++) smac[q * GREG + tid] = m[q];
    }
    const int nd = __builtin_popcount(bmask);
    for (int idx = tid; idx < nd * GREG; idx += 256) {
        const int d = idx / GREG, r = idx - d * GREG;
        // the d-th direction of the block's mask
        uint32_t mm = bmask;
        for (int skip = 0; skip < d; skip++) mm &= mm - 1u;
        const int k = __builtin_ctz(mm);
        const int p = sreg[r];
        float vn = 0.0f, vo = 0.0f;
        if (p >= 0) {
            const float w_k = wk_rt(k);
            vn = Codec<ST>::load(a.a_new, fidx(p >> 9, k, p & 511), w_k);
            if (need_old) vo = Codec<ST>::load(a.a_old, fidx(p >> 9, k, p & 511), w_k);
        }
        spop[(2 * d) * GREG + r] = vn;
        spop[(2 * d + 1) * GREG + r] = vo;
    }
    __syncthreads();
    const float tc = a.tau_coarse - 0.5f;
    const float tf = a.tau_fine - 0.5f;
    float scale = 1.0f;
    if (tc > 1.0e-6f) {
        scale = tf / tc;
        scale = scale > 100.0f ? 100.0f : (scale < 0.01f ? 0.01f : scale);
    }
    const int e_end = L.gstart[g + 1];
    for (int i = L.gstart[g] + tid; i < e_end; i += 256) {
        const int e0 = a.entries[2 * i];
        const uint32_t mask = (uint32_t)a.entries[2 * i + 1];
        const int gb = e0 >> 9, c = e0 & 511;
        // The fine cell centre lies at (g - 0.5) / 2 in parent cells, g = block
        // origin (a multiple of 8) + local index + 1, so its fractional part is
        // 0.25 for an even local index and 0.75 for an odd one (exactly what
        // (g - 0.5) * 0.5 - floor(...) gives in FP32).  Its corners in the
        // region: l / 2 and l / 2 + 1, the lower one raised to parent cell 1
        // at the domain's low faces (region cell 0 is parent cell B / 2).
        const float wx = (c & 1) ? 0.75f : 0.25f;
        const float wy = ((c >> 3) & 1) ? 0.75f : 0.25f;
        const float wz = ((c >> 6) & 1) ? 0.75f : 0.25f;
        int rlo[3], rhi[3];
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
            const int l = (c >> (3 * ax)) & 7;
            const int half = L.borg[(size_t)gb * 3 + ax] / 2;
            const int lo = half + (l >> 1);
            rhi[ax] = (l >> 1) + 1;
            rlo[ax] = (lo < 1 ? 1 : lo) - half;
        }
        int rq[8];
        bool ok[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            rq[q] = ((q & 1) ? rhi[0] : rlo[0]) + 5 * ((q & 2) ? rhi[1] : rlo[1]) + 25 * ((q & 4) ? rhi[2] : rlo[2]);
            ok[q] = sreg[rq[q]] >= 0;
        }
        float v0[8][4], v1[8][4];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (ok[q]) {
                const float rn = smac[0 * GREG + rq[q]], uxn = smac[1 * GREG + rq[q]];
                const float uyn = smac[2 * GREG + rq[q]], uzn = smac[3 * GREG + rq[q]];
                const float ro = smac[4 * GREG + rq[q]], uxo = smac[5 * GREG + rq[q]];
                const float uyo = smac[6 * GREG + rq[q]], uzo = smac[7 * GREG + rq[q]];
                if (use_old0) {
                    v0[q][0] = ro * (1.0f - tw0) + rn * tw0;
                    v0[q][1] = uxo * (1.0f - tw0) + uxn * tw0;
                    v0[q][2] = uyo * (1.0f - tw0) + uyn * tw0;
                    v0[q][3] = uzo * (1.0f - tw0) + uzn * tw0;
                } else {
                    v0[q][0] = rn; v0[q][1] = uxn; v0[q][2] = uyn; v0[q][3] = uzn;
                }
                if (use_old1) {
                    v1[q][0] = ro * (1.0f - tw1) + rn * tw1;
                    v1[q][1] = uxo * (1.0f - tw1) + uxn * tw1;
                    v1[q][2] = uyo * (1.0f - tw1) + uyn * tw1;
                    v1[q][3] = uzo * (1.0f - tw1) + uzn * tw1;
                } else {
                    v1[q][0] = rn; v1[q][1] = uxn; v1[q][2] = uyn; v1[q][3] = uzn;
                }
            } else if (q == 0) {
                v0[0][0] = 1.0f; v0[0][1] = 0.0f; v0[0][2] = 0.0f; v0[0][3] = 0.0f;
                v1[0][0] = 1.0f; v1[0][1] = 0.0f; v1[0][2] = 0.0f; v1[0][3] = 0.0f;
            } else {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    v0[q][m] = v0[0][m];
                    v1[q][m] = v1[0][m];
                }
            }
        }
        float r0[4], r1[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            {
                float c00 = v0[0][m] * (1.0f - wx) + v0[1][m] * wx;
                float c01 = v0[4][m] * (1.0f - wx) + v0[5][m] * wx;
                float c10 = v0[2][m] * (1.0f - wx) + v0[3][m] * wx;
                float c11 = v0[6][m] * (1.0f - wx) + v0[7][m] * wx;
                float c0 = c00 * (1.0f - wy) + c10 * wy;
                float c1 = c01 * (1.0f - wy) + c11 * wy;
                r0[m] = c0 * (1.0f - wz) + c1 * wz;
            }
            {
                float c00 = v1[0][m] * (1.0f - wx) + v1[1][m] * wx;
                float c01 = v1[4][m] * (1.0f - wx) + v1[5][m] * wx;
                float c10 = v1[2][m] * (1.0f - wx) + v1[3][m] * wx;
                float c11 = v1[6][m] * (1.0f - wx) + v1[7][m] * wx;
                float c0 = c00 * (1.0f - wy) + c10 * wy;
                float c1 = c01 * (1.0f - wy) + c11 * wy;
                r1[m] = c0 * (1.0f - wz) + c1 * wz;
            }
        }
#pragma unroll 1
        for (uint32_t mm = mask; mm != 0u; mm &= mm - 1u) {
            const int k = __builtin_ctz(mm);
            const int d = __builtin_popcount(bmask & ((1u << k) - 1u));  // k's slot in the staged region
            float ln[8], lo[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                ln[q] = spop[(2 * d) * GREG + rq[q]];
                lo[q] = spop[(2 * d + 1) * GREG + rq[q]];
            }
            ghost_blend_store<ST>(a, k, ok, ln, lo, use_old0, use_old1, wx, wy, wz, r0, r1, scale, gb, c);
        }
    }
}

// One work-group per work entry: (LevelArgs index, first of 256 ghost
// entries) for FP32 storage, (LevelArgs index, ghost block) for 16-bit
// storage.  Thermal populations use the flow's storage.
hipError_t ghost_fill(const LevelArgs* las, const CaseState* states, const Work* work, int32_t n_chunks, int64_t j,
                      int storage, int thermal, hipStream_t s) {
    if (n_chunks <= 0) return hipSuccess;
    const dim3 grid((unsigned)n_chunks), t(256);
    const int st = storage;
    if (st == 1) hipLaunchKernelGGL(ghost_kernel_staged<1>, grid, t, 0, s, las, states, work, j, thermal);
    else if (st == 2) hipLaunchKernelGGL(ghost_kernel_staged<2>, grid, t, 0, s, las, states, work, j, thermal);
    else hipLaunchKernelGGL(ghost_kernel<0>, grid, t, 0, s, las, states, work, j, thermal);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Bouzidi (src/geometry/bouzidi_kernel.jl:13-92) as two sparse passes over
// the link list: gather every corrected value from the uncorrected
// post-collision populations, then scatter.  Each link's destination slot is
// distinct, so the result equals the reference's read-from-f_post_collision
// form without its full-size copy.  __fmul_rn/__fadd_rn keep the two
// products and the sum rounded separately in both builds.
// All three slots of a link have the same |c|, so one weight w_k (from the
// destination's direction) decodes and encodes them.
// The corrected buffer is the substep's output half (f_out, or g_out).
__device__ __forceinline__ void* bouzidi_target(const LevelArgs& L, const CaseState& S, int64_t j, int thermal) {
    const int out = 1 - (int)(substep(L, S, j) & 1);
    return thermal ? (void*)L.g[out] : L.f[out];
}
// Slot index i of a link array (uint32 or int64 entries, uniform per level).
__device__ __forceinline__ size_t link_slot(const BouzidiLinks& l, const void* p, int64_t i) {
    typedef const __attribute__((address_space(1))) uint32_t GU32;
    typedef const __attribute__((address_space(1))) int64_t GI64;
    return l.idx32 ? (size_t)((GU32*)p)[i] : (size_t)((GI64*)p)[i];
}
template <int ST>
__global__ void __launch_bounds__(256) bouzidi_gather(const LevelArgs* __restrict__ las,
                                                      const CaseState* __restrict__ states,
                                                      const Work* __restrict__ work, int64_t j, int thermal) {
    const Work wk = work[blockIdx.x];
    const LevelArgs& L = las[wk.li];
    const CaseState S = case_state(states + wk.ci);
    if (!live(L, S, j)) return;
    const BouzidiLinks& l = L.links;
    const int64_t i = (int64_t)wk.item + threadIdx.x;
    if (i >= l.n) return;
    const void* f = bouzidi_target(L, S, j, thermal);
    const float w = wk_rt((int)((link_slot(l, l.dst, i) >> 9) % 27));
    const float f1 = Codec<ST>::load(f, link_slot(l, l.s1, i), w), f2 = Codec<ST>::load(f, link_slot(l, l.s2, i), w);
    l.tmp[i] = __fadd_rn(__fmul_rn(l.a[i], f1), __fmul_rn(l.b[i], f2));
}
template <int ST>
__global__ void __launch_bounds__(256) bouzidi_scatter(const LevelArgs* __restrict__ las,
                                                       const CaseState* __restrict__ states,
                                                       const Work* __restrict__ work, int64_t j, int thermal) {
    const Work wk = work[blockIdx.x];
    const LevelArgs& L = las[wk.li];
    const CaseState S = case_state(states + wk.ci);
    if (!live(L, S, j)) return;
    const BouzidiLinks& l = L.links;
    const int64_t i = (int64_t)wk.item + threadIdx.x;
    if (i >= l.n) return;
    void* f = bouzidi_target(L, S, j, thermal);
    const size_t d = link_slot(l, l.dst, i);
    const float w = wk_rt((int)((d >> 9) % 27));
    Codec<ST>::store(f, d, l.tmp[i], w);
}

template <int ST>
void bouzidi_launch(const LevelArgs* las, const CaseState* states, const Work* work, int32_t n_chunks, int64_t j,
                    int thermal, hipStream_t s) {
    const dim3 g((unsigned)n_chunks), t(256);
    hipLaunchKernelGGL(bouzidi_gather<ST>, g, t, 0, s, las, states, work, j, thermal);
    hipLaunchKernelGGL(bouzidi_scatter<ST>, g, t, 0, s, las, states, work, j, thermal);
}

// One work-group per work entry (LevelArgs index, first of 256 links).
hipError_t bouzidi(const LevelArgs* las, const CaseState* states, const Work* work, int32_t n_chunks, int64_t j, int storage, int thermal,
                   hipStream_t s) {
    if (n_chunks <= 0) return hipSuccess;
    const int st = storage;
    if (st == 1) bouzidi_launch<1>(las, states, work, n_chunks, j, thermal, s);
    else if (st == 2) bouzidi_launch<2>(las, states, work, n_chunks, j, thermal, s);
    else bouzidi_launch<0>(las, states, work, n_chunks, j, thermal, s);
    return hipGetLastError();
}

// Case state updates.  set_state writes a whole record (single-case stepping,
// once per coarse step); advance moves every case of an engine by its
// step increment and looks up the inlet velocity of the new ramp batch.
__global__ void set_state_kernel(CaseState* st, const CaseState v) {
    if (threadIdx.x == 0) *st = v;
}
hipError_t set_state(CaseState* st, const CaseState& v, hipStream_t s) {
    hipLaunchKernelGGL(set_state_kernel, dim3(1), dim3(64), 0, s, st, v);
    return hipGetLastError();
}
__global__ void advance_kernel(CaseState* st, int32_t n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CaseState& c = st[i];
    if (c.active == 0) return;
    c.t0 += c.t_inc;
    if (c.t0 > c.n_total) c.active = 0;
}
hipError_t advance(CaseState* st, int32_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(advance_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, st, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// init_eq! (main.jl:113-128): f = f_temp = w, rho = 1, vel = 0 on both halves
// of every ping-pong pair (the old-state aliases included).
template <int ST>
__global__ void init_kernel(void* f0, void* f1, float* rho0, float* rho1, float* v0, float* v1) {
    const int b = blockIdx.x >> 1;
    const int c = ((blockIdx.x & 1) << 8) | threadIdx.x;
#pragma unroll
    for (int k = 0; k < 27; k++) {
        Codec<ST>::store(f0, fidx(b, k, c), WK(k), WK(k));
        Codec<ST>::store(f1, fidx(b, k, c), WK(k), WK(k));
    }
    rho0[(size_t)b * NC + c] = 1.0f;
    rho1[(size_t)b * NC + c] = 1.0f;
    size_t iv = (size_t)b * 3 * NC + c;
    for (int d = 0; d < 3; d++) {
        v0[iv + d * NC] = 0.0f;
        v1[iv + d * NC] = 0.0f;
    }
}
hipError_t init_level(void* f0, void* f1, float* rho0, float* rho1, float* v0, float* v1, int32_t nb, int storage,
                      hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    const dim3 grid((unsigned)nb * 2);
    if (storage == 1) hipLaunchKernelGGL(init_kernel<1>, grid, dim3(256), 0, s, f0, f1, rho0, rho1, v0, v1);
    else if (storage == 2) hipLaunchKernelGGL(init_kernel<2>, grid, dim3(256), 0, s, f0, f1, rho0, rho1, v0, v1);
    else hipLaunchKernelGGL(init_kernel<0>, grid, dim3(256), 0, s, f0, f1, rho0, rho1, v0, v1);
    return hipGetLastError();
}

// init_thermal_eq! (main.jl:131-150)
template <int ST>
__global__ void init_thermal_kernel(void* g0, void* g1, float* T0, float* T1, float Tinit) {
    const int b = blockIdx.x >> 1;
    const int c = ((blockIdx.x & 1) << 8) | threadIdx.x;
#pragma unroll
    for (int k = 0; k < 27; k++) {
        float ge = WK(k) * Tinit;
        Codec<ST>::store(g0, fidx(b, k, c), ge, WK(k));
        Codec<ST>::store(g1, fidx(b, k, c), ge, WK(k));
    }
    T0[(size_t)b * NC + c] = Tinit;
    T1[(size_t)b * NC + c] = Tinit;
}
hipError_t init_thermal(void* g0, void* g1, float* T0, float* T1, int32_t nb, float Tinit, int storage,
                        hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    const dim3 grid((unsigned)nb * 2), t(256);
    if (storage == 1) hipLaunchKernelGGL(init_thermal_kernel<1>, grid, t, 0, s, g0, g1, T0, T1, Tinit);
    else if (storage == 2) hipLaunchKernelGGL(init_thermal_kernel<2>, grid, t, 0, s, g0, g1, T0, T1, Tinit);
    else hipLaunchKernelGGL(init_thermal_kernel<0>, grid, t, 0, s, g0, g1, T0, T1, Tinit);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Surface forces (src/forces/surface.jl:143-507, 702-762).
__device__ __forceinline__ bool cell_at(const ForceArgs& a, int gx, int gy, int gz, int& blk, int& c) {
    if (gx < 1 || gy < 1 || gz < 1) return false;
    int bx = (gx - 1) / BS + 1, by = (gy - 1) / BS + 1, bz = (gz - 1) / BS + 1;
    if (bx < 1 || bx > a.gdx || by < 1 || by > a.gdy || bz < 1 || bz > a.gdz) return false;
    int id = a.grid[(bx - 1) + a.gdx * ((by - 1) + a.gdy * (bz - 1))];
    if (id < 0) return false;
    blk = id;
    c = ((gx - 1) % BS) + BS * ((gy - 1) % BS) + BS * BS * ((gz - 1) % BS);
    return true;
}

// trilinear_sample_flow (surface.jl:224-346): corners in the reference order.
__device__ bool trilinear_sample(const ForceArgs& a, float px, float py, float pz, float& ro, float& uo, float& vo,
                                 float& wo, float& wq) {
    float inv_dx = 1.0f / a.dx;
    float fx = px * inv_dx - 0.5f, fy = py * inv_dx - 0.5f, fz = pz * inv_dx - 0.5f;
    int ix0 = (int)floorf(fx) + 1, iy0 = (int)floorf(fy) + 1, iz0 = (int)floorf(fz) + 1;
    float wx = fminf(fmaxf(fx - (float)(ix0 - 1), 0.0f), 1.0f);
    float wy = fminf(fmaxf(fy - (float)(iy0 - 1), 0.0f), 1.0f);
    float wz = fminf(fmaxf(fz - (float)(iz0 - 1), 0.0f), 1.0f);
    float w0x = 1.0f - wx, w1x = wx, w0y = 1.0f - wy, w1y = wy, w0z = 1.0f - wz, w1z = wz;
    float racc = 0.0f, uacc = 0.0f, vacc = 0.0f, wacc_u = 0.0f, wacc = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        int dxi = i & 1, dyi = (i >> 1) & 1, dzi = (i >> 2) & 1;
        float wk = (dxi ? w1x : w0x) * (dyi ? w1y : w0y) * (dzi ? w1z : w0z);
        int blk, c;
        if (cell_at(a, ix0 + dxi, iy0 + dyi, iz0 + dzi, blk, c)) {
            size_t i4 = (size_t)blk * NC + c;
            if (!a.obst[i4]) {
                size_t iv = (size_t)blk * 3 * NC + c;
                racc += wk * a.rho[i4];
                uacc += wk * a.vel[iv];
                vacc += wk * a.vel[iv + NC];
                wacc_u += wk * a.vel[iv + 2 * NC];
                wacc += wk;
            }
        }
    }
    if (wacc > 0.01f) {
        float iw = 1.0f / wacc;
        ro = racc * iw; uo = uacc * iw; vo = vacc * iw; wo = wacc_u * iw; wq = wacc;
        return true;
    }
    ro = 1.0f; uo = 0.0f; vo = 0.0f; wo = 0.0f; wq = 0.0f;
    return false;
}

// map_stresses_kernel! + integrate_forces_kernel! per triangle; the nine sums
// are reduced in FP64 in a fixed tree order per work-group (deterministic).
// With side = -1 it is map_minus_side_only_kernel! (surface.jl:511-662):
// that kernel is this one with the normal negated in the probe, the
// fallback's side test, the wall distance and the stress (each a sign flip,
// exact), so the maps are computed here with n -> -n; the integrated sums
// are then not used.
// One work-group: triangles chunk*256 .. chunk*256+255, partial sums into
// a.partials[chunk].
__device__ void forces_block(const ForceArgs& a, int64_t chunk) {
    __shared__ double red[9][256];
    int64_t i = chunk * 256 + threadIdx.x;
    float d9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < a.n_tri) {
        float tx = a.cx[i] + a.off_x, ty = a.cy[i] + a.off_y, tz = a.cz[i] + a.off_z;
        float n_x = a.side * a.nx[i], n_y = a.side * a.ny[i], n_z = a.side * a.nz[i];
        float brho = 1.0f, bux = 0.0f, buy = 0.0f, buz = 0.0f, bwd = 1.0f, bq = 0.0f;
        bool found = false;
        for (int pi = 1; pi <= a.search_radius; pi++) {
            float dc = (float)pi + 0.5f;
            float px = tx + n_x * dc * a.dx, py = ty + n_y * dc * a.dx, pz = tz + n_z * dc * a.dx;
            float r, u, v, w, wq;
            bool ok = trilinear_sample(a, px, py, pz, r, u, v, w, wq);
            if (ok && wq >= 0.5f) {
                brho = r; bux = u; buy = v; buz = w; bwd = dc; bq = wq; found = true;
                break;
            } else if (ok && wq > bq) {
                brho = r; bux = u; buy = v; buz = w; bwd = dc; bq = wq; found = true;
            }
        }
        if (!found) {
            float gxf = tx / a.dx, gyf = ty / a.dx, gzf = tz / a.dx;
            int g_x = (int)floorf(gxf) + 1, g_y = (int)floorf(gyf) + 1, g_z = (int)floorf(gzf) + 1;
            float best = 1e10f;
            for (int rad = 1; rad <= a.search_radius; rad++) {
                if (found) break;
                for (int dz = -rad; dz <= rad; dz++)
                    for (int dy = -rad; dy <= rad; dy++)
                        for (int ddx = -rad; ddx <= rad; ddx++) {
                            bool shell = (abs(ddx) == rad) || (abs(dy) == rad) || (abs(dz) == rad);
                            if (!shell) continue;
                            int cgx = g_x + ddx, cgy = g_y + dy, cgz = g_z + dz;
                            int blk, c;
                            if (!cell_at(a, cgx, cgy, cgz, blk, c)) continue;
                            size_t i4 = (size_t)blk * NC + c;
                            if (a.obst[i4]) continue;
                            float ccx = ((float)cgx - 0.5f) * a.dx, ccy = ((float)cgy - 0.5f) * a.dx,
                                  ccz = ((float)cgz - 0.5f) * a.dx;
                            float dxp = ccx - tx, dyp = ccy - ty, dzp = ccz - tz;
                            float dn = dxp * n_x + dyp * n_y + dzp * n_z;
                            if (dn <= 0.0f) continue;
                            float dsq = dxp * dxp + dyp * dyp + dzp * dzp;
                            if (dsq < best) {
                                size_t iv = (size_t)blk * 3 * NC + c;
                                best = dsq;
                                brho = a.rho[i4];
                                bux = a.vel[iv]; buy = a.vel[iv + NC]; buz = a.vel[iv + 2 * NC];
                                bwd = dn / a.dx;
                                found = true;
                            }
                        }
            }
        }
        float p = 0.0f, tx_ = 0.0f, ty_ = 0.0f, tz_ = 0.0f;
        if (found) {
            // compute_stress_from_cell (surface.jl:143-190)
            float wd = fmaxf(bwd, 0.5f);
            float pg = (brho - 1.0f) / 3.0f;
            p = pg * a.pressure_scale;
            float und = bux * n_x + buy * n_y + buz * n_z;
            float utx = bux - und * n_x, uty = buy - und * n_y, utz = buz - und * n_z;
            float utm = sqrtf(utx * utx + uty * uty + utz * utz);
            float nu_mol = (a.tau_mol - 0.5f) / 3.0f;
            float nu_sgs = 0.0f;
            if (a.c_wale > 0.0f && wd > 0.01f && utm > 1.0e-10f) {
                float S = utm / wd;
                nu_sgs = (a.c_wale * a.c_wale) * S;
            }
            nu_sgs = fmaxf(nu_sgs, a.nu_sgs_bg);
            float nu_eff = nu_mol + nu_sgs;
            if (utm > 1.0e-10f && wd > 0.01f) {
                float tl = brho * nu_eff * utm / wd;
                float tp = tl * a.stress_scale;
                tx_ = (utx / utm) * tp;
                ty_ = (uty / utm) * tp;
                tz_ = (utz / utm) * tp;
            }
        }
        if (a.p_map) {
            a.p_map[i] = p; a.sx_map[i] = tx_; a.sy_map[i] = ty_; a.sz_map[i] = tz_;
        }
        // integrate_forces_kernel! per-triangle terms (surface.jl:716-748)
        float A = a.area[i];
        float dFpx = -p * n_x * A, dFpy = -p * n_y * A, dFpz = -p * n_z * A;
        float dFvx = tx_ * A, dFvy = ty_ * A, dFvz = tz_ * A;
        float dFx = dFpx + dFvx, dFy = dFpy + dFvy, dFz = dFpz + dFvz;
        float rx = (a.cx[i] + a.off_x) - a.ref_x, ry = (a.cy[i] + a.off_y) - a.ref_y, rz = (a.cz[i] + a.off_z) - a.ref_z;
        d9[0] = dFpx; d9[1] = dFpy; d9[2] = dFpz;
        d9[3] = dFvx; d9[4] = dFvy; d9[5] = dFvz;
        d9[6] = ry * dFz - rz * dFy;
        d9[7] = rz * dFx - rx * dFz;
        d9[8] = rx * dFy - ry * dFx;
    }
#pragma unroll
    for (int j = 0; j < 9; j++) red[j][threadIdx.x] = (double)d9[j];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int j = 0; j < 9; j++) red[j][threadIdx.x] += red[j][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 9) a.partials[(size_t)chunk * 9 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) forces_kernel(const ForceArgs a) { forces_block(a, blockIdx.x); }

hipError_t forces(const ForceArgs& a, int n_groups, double* out9, hipStream_t s) {
    (void)out9;
    if (a.n_tri <= 0) return hipSuccess;
    hipLaunchKernelGGL(forces_kernel, dim3((unsigned)n_groups), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Force samples inside the batched engine's step (accumulate_forces!,
// surface.jl:36-56): coarse step t = t0 + j of the cases whose first level
// runs at this depth is complete at this point of the recursion.  When t is
// a sampling step (t % diag == 0, ramp < t <= n_total) the finest level's
// forces are integrated exactly as lb_forces does -- the density of the
// level's last substep of t, level.vel -- into each case's partials, and
// the reduction adds the nine sums, in chunk order, to the case's
// accumulator (so a case's samples do not depend on the other cases).
__device__ __forceinline__ bool sample_step(const EngineForce& R, const CaseState& S, int64_t j, int64_t* t_out) {
    const int64_t t = S.t0 + j;
    *t_out = t;
    return S.active && t <= S.n_total && t > R.ramp && (t % R.diag) == 0;
}

__global__ void __launch_bounds__(256) engine_forces_kernel(const EngineForce* __restrict__ recs,
                                                            const CaseState* __restrict__ states,
                                                            const Work* __restrict__ work, int64_t j) {
    const Work wk = work[blockIdx.x];
    const EngineForce& R = recs[wk.li];
    int64_t t;
    if (!sample_step(R, states[wk.ci], j, &t)) return;
    ForceArgs a = R.a;
    const int64_t last = (t << R.leaf_shift) + ((int64_t)1 << R.leaf_shift) - 1;
    a.rho = R.rho2[1 - (int)(last & 1)];
    forces_block(a, wk.item);
}

__global__ void engine_forces_reduce(const EngineForce* __restrict__ recs, const CaseState* __restrict__ states,
                                     const int32_t* __restrict__ rec_idx, int32_t n, int64_t j) {
    const int32_t q = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q >= n) return;
    const EngineForce& R = recs[rec_idx[q]];
    int64_t t;
    if (!sample_step(R, states[R.ci], j, &t)) return;
    double s9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t g = 0; g < R.n_groups; g++)
        for (int c = 0; c < 9; c++) s9[c] += R.a.partials[(size_t)g * 9 + c];
    for (int c = 0; c < 9; c++) R.acc[c] += s9[c];
    R.acc[9] += 1.0;
}

hipError_t engine_forces(const EngineForce* recs, const CaseState* states, const Work* work, int32_t n_chunks,
                         const int32_t* rec_idx, int32_t n_recs, int64_t j, hipStream_t s) {
    if (n_chunks <= 0 || n_recs <= 0) return hipSuccess;
    hipLaunchKernelGGL(engine_forces_kernel, dim3((unsigned)n_chunks), dim3(256), 0, s, recs, states, work, j);
    hipLaunchKernelGGL(engine_forces_reduce, dim3((unsigned)((n_recs + 63) / 64)), dim3(64), 0, s, recs, states,
                       rec_idx, n_recs, j);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 diagnostics per block (compute_flow_stats, diagnostics.jl:43-93, plus
// the global mass check): sum rho, KE, min/max rho and max |u| over fluid
// cells, non-finite populations, fluid count, sum of populations.
template <int ST>
__global__ void __launch_bounds__(256) stats_kernel(const void* __restrict__ f, const float* __restrict__ rho,
                                                    const float* __restrict__ vel, const uint8_t* __restrict__ obst,
                                                    int32_t nb, double* partials) {
    __shared__ double red[8][256];
    double s_rho = 0, s_ke = 0, mn = 1e300, mx = -1e300, nonfin = 0, nfl = 0, s_f = 0, vmax = 0;
    for (int64_t blk = blockIdx.x; blk < nb; blk += gridDim.x) {
        for (int c = threadIdx.x; c < NC; c += 256) {
            size_t i4 = (size_t)blk * NC + c;
            double cell_f = 0;
            for (int k = 0; k < 27; k++) {
                float v = Codec<ST>::load(f, fidx((int)blk, k, c), wk_rt(k));
                if (!isfinite(v)) nonfin += 1;
                cell_f += (double)v;
            }
            s_f += cell_f;
            if (!obst[i4]) {
                double r = rho[i4];
                size_t iv = (size_t)blk * 3 * NC + c;
                double u = vel[iv], v = vel[iv + NC], w = vel[iv + 2 * NC];
                const double u2 = u * u + v * v + w * w;
                s_rho += r;
                s_ke += 0.5 * r * u2;
                mn = r < mn ? r : mn;
                mx = r > mx ? r : mx;
                const double um = sqrt(u2);
                vmax = um > vmax ? um : vmax;
                nfl += 1;
            }
        }
    }
    red[0][threadIdx.x] = s_rho; red[1][threadIdx.x] = s_ke; red[2][threadIdx.x] = mn; red[3][threadIdx.x] = mx;
    red[4][threadIdx.x] = nonfin; red[5][threadIdx.x] = nfl; red[6][threadIdx.x] = s_f; red[7][threadIdx.x] = vmax;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
            red[2][threadIdx.x] = fmin(red[2][threadIdx.x], red[2][threadIdx.x + s]);
            red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + s]);
            red[4][threadIdx.x] += red[4][threadIdx.x + s];
            red[5][threadIdx.x] += red[5][threadIdx.x + s];
            red[6][threadIdx.x] += red[6][threadIdx.x + s];
            red[7][threadIdx.x] = fmax(red[7][threadIdx.x], red[7][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 8) partials[(size_t)blockIdx.x * 8 + threadIdx.x] = red[threadIdx.x][0];
}

hipError_t stats(const void* f, const float* rho, const float* vel, const uint8_t* obst, int32_t nb, int storage,
                 double* partials, int n_groups, hipStream_t s) {
    if (nb <= 0) return hipSuccess;
    const dim3 grid((unsigned)n_groups);
    if (storage == 1) hipLaunchKernelGGL(stats_kernel<1>, grid, dim3(256), 0, s, f, rho, vel, obst, nb, partials);
    else if (storage == 2) hipLaunchKernelGGL(stats_kernel<2>, grid, dim3(256), 0, s, f, rho, vel, obst, nb, partials);
    else hipLaunchKernelGGL(stats_kernel<0>, grid, dim3(256), 0, s, f, rho, vel, obst, nb, partials);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Divergence check: non-finite density and velocity values per case.  Row y
// of the grid is one (case, level) entry; its work-groups stride over the
// entry's n cells (rho) and 3n velocity components and add their count to
// counts[entry.slot].
__global__ void __launch_bounds__(256) finite_kernel(const FiniteEntry* __restrict__ ents, int32_t* counts) {
    const FiniteEntry e = ents[blockIdx.y];
    int32_t bad = 0;
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < e.n; i += stride) bad += isfinite(e.rho[i]) ? 0 : 1;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < 3 * e.n; i += stride)
        bad += isfinite(e.vel[i]) ? 0 : 1;
    for (int off = 32; off > 0; off >>= 1) bad += __shfl_down(bad, off, 64);
    if ((threadIdx.x & 63) == 0 && bad != 0) atomicAdd(&counts[e.slot], bad);
}

hipError_t finite_count(const FiniteEntry* ents, int32_t n_ents, int32_t groups_per_entry, int32_t* counts,
                        hipStream_t s) {
    if (n_ents <= 0) return hipSuccess;
    hipLaunchKernelGGL(finite_kernel, dim3((unsigned)groups_per_entry, (unsigned)n_ents), dim3(256), 0, s, ents,
                       counts);
    return hipGetLastError();
}

}  // namespace LBK_NS
