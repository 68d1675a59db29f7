// lb_kernels.h -- argument blocks and launchers of the gfx950 kernels.
//
// lb_kernels.hip is compiled twice: once with -ffp-contract=off into namespace
// lbk_strict (every float operation rounded as written, so results compare
// bit for bit with a plain C restatement) and once with FMA contraction into
// lbk_fast.  The API layer (lb_api.hip) picks one per context.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbk {

constexpr int BS = 8;
constexpr int NC = 512;  // cells per block
constexpr int Q = 27;

// Storage of one level on the device (block-major SoA):
//   f[b][k][c]   c = x + 8y + 64z (0-based local cell), FP32 or 16-bit offsets
//   rho[b][c], vel[b][d][c], obst[b][c], sponge[b][c], wall_dist[b][c]
//   nbr[b][27]   0-based neighbour block, -1 absent (dir as k); ghost blocks
//                (coarse-to-fine interface sources) have indices >= nb
//   borg[b][3]   global 1-based cell coordinate of local (0,0,0) minus 1
//   grid         dense block pointer (max active coords), 0-based ids, -1 absent
//
// Kernels take their arguments from device memory: a LevelArgs record per
// (case, level) and a CaseState per case, named by a work list whose entries
// are (LevelArgs index, item, CaseState index).  One launch can therefore cover one level of
// one case or the same depth of many cases (batched serving); a hipGraph of
// such launches replays without host-side argument changes.

// Per-case state: the step position and the physics parameters of
// lb_step_params.  Substep j of level l in the current step has
// t_sub = t0 * 2^(l-1) + j (the recursion of solver_control.jl:6-197).
struct CaseState {
    int64_t t0;           // coarse step that the current step starts at
    int64_t n_total;      // last coarse step; substeps of later coarse steps are skipped
    float u;              // inlet velocity when u_tab is null
    int32_t active;       // 0: the case's work-groups return at once
    int32_t t_inc;        // coarse steps per engine step (2^(depth offset))
    int32_t batch;        // ramp batch length in coarse steps (u_tab granularity)
    const float* u_tab;   // inlet velocity per ramp batch, or null
    float c_wale, nu_sgs_bg, inlet_turb;
    float omega_bulk, omega3, omega4, lambda;
    int32_t adaptive, limiter, comp, wall_model, sponge_blend;
    float T_inlet, T_wall;
    int32_t adiabatic, temporal;
};

// Bouzidi link list: corrected slot dst = a*f[s1] + b*f[s2], all indices
// into one population buffer (block-major layout).  The slot indices are
// uint32 when idx32 (a buffer below 2^32 slots), else int64.
struct BouzidiLinks {
    int64_t n;
    const void* dst;
    const void* s1;
    const void* s2;
    const float* a;
    const float* b;
    float* tmp;
    int32_t idx32;
    int32_t pad;
};

struct LevelArgs {
    void* f[2];  // ping-pong pairs, indexed by substep parity (input = t_sub & 1)
    float* rho[2];
    float* vel[2];
    void* g[2];  // thermal populations (storage as f) or null
    float* T[2];
    const uint8_t* obst;
    const float* sponge;
    const float* wall_dist;  // null when the level has no wall distances
    // per block: bit 0 some cell has sponge > 0; bit 1 some cell has
    // 0 < wall distance < 10 (the wall model's range); bit 2 some cell is solid
    const uint8_t* bflags;
    const int32_t* nbr;
    const int32_t* borg;
    int32_t nb;
    int32_t nxg, nyg, nzg;
    int32_t lshift;  // level_id - 1
    // 1 for the finest level (no child level reads its density or
    // temperature): those scalar fields are then stored only on the last
    // substep of each coarse step, the only one whose output is read
    int32_t leaf;
    float tau, tau_g;
    // parent level (level_id > 1): its pairs, indexed by the parent's parity
    const void* pf[2];
    const float* prho[2];
    const float* pvel[2];
    const void* pg[2];
    const float* pT[2];
    const int32_t* pgrid;
    int32_t pdx, pdy, pdz;
    float ptau, ptau_g;
    // ghost entries: [n][2] = (ghost block << 9 | cell, 27-bit direction mask)
    const int32_t* gent;
    // FP32 storage, per ghost entry: its 8 parent cells (parent block << 9 |
    // cell, or -1)
    const int32_t* gpar;
    // 16-bit storage (entries grouped by ghost block), per ghost block: its
    // first entry ([ngh + 1]), its 125 parent cells (parent block << 9 |
    // cell, or -1; cell r = rx + 5 ry + 25 rz of the parent region starting
    // at half the block origin), the union of its entries' direction masks
    const int32_t* gstart;
    const int32_t* gregion;
    const uint32_t* gbmask;
    int64_t ngent;
    BouzidiLinks links;
};

// Work list entry.  The neighbour row travels with the entry so that the
// kernel's first dependent loads (LevelArgs, CaseState, neighbour row) are
// issued together.
struct Work {
    int32_t li;              // LevelArgs index
    int32_t item;            // block (stream-collide) or first entry of a 256-entry chunk
    int32_t ci;              // CaseState index
    int32_t pad;
    const int32_t* nbr_row;  // the block's neighbour row (stream-collide), else null
};

struct ForceArgs {
    // finest level
    const float* rho;
    const float* vel;
    const uint8_t* obst;
    const int32_t* grid;
    int32_t gdx, gdy, gdz;
    float dx;
    // mesh
    int64_t n_tri;
    const float* cx;
    const float* cy;
    const float* cz;
    const float* nx;
    const float* ny;
    const float* nz;
    const float* area;
    float off_x, off_y, off_z;
    float pressure_scale, stress_scale, tau_mol;
    int32_t search_radius;
    float c_wale, nu_sgs_bg;
    float ref_x, ref_y, ref_z;
    // outputs
    float* p_map;
    float* sx_map;
    float* sy_map;
    float* sz_map;
    double* partials;  // [n_groups][9]
    float side;        // +1: probe along the outward normal; -1: the minus side
};

// Force sampling of one case attached to the batched engine: the finest
// level's force arguments (a.rho is replaced by the density of the sampled
// step's parity), the sampling cadence and the FP64 accumulator
// acc[0..8] = sums of the nine force/moment components, acc[9] = samples.
struct EngineForce {
    ForceArgs a;
    const float* rho2[2];  // finest level's density pair
    int64_t diag;          // sample every diag coarse steps ...
    int64_t ramp;          // ... after the ramp (t > ramp)
    int32_t ci;            // CaseState index
    int32_t leaf_shift;    // finest level_id - 1
    int32_t n_groups;      // triangle chunks of 256 (partials rows)
    int32_t pad;
    double* acc;
};

// One (case, level) of the divergence check: the level's current density
// (n values) and velocity (3n values); non-finite values are added to
// counts[slot].
struct FiniteEntry {
    const float* rho;
    const float* vel;
    int64_t n;
    int32_t slot;
    int32_t pad;
};

}  // namespace lbk

#define LBK_DECLARE(NS)                                                                                  \
    namespace NS {                                                                                       \
    hipError_t collide(const lbk::LevelArgs* las, const lbk::CaseState* states, const lbk::Work* work,    \
                       int32_t n_tiles, int64_t j, int storage, int wide, int boundary, int sched,       \
                       int waves, hipStream_t s);                                                        \
    hipError_t thermal(const lbk::LevelArgs* las, const lbk::CaseState* states, const lbk::Work* work,    \
                       int32_t n_tiles, int64_t j, int storage, int wide, int boundary, int sched,       \
                       hipStream_t s);                                                                   \
    hipError_t ghost_fill(const lbk::LevelArgs* las, const lbk::CaseState* states, const lbk::Work* work, int32_t n_chunks, int64_t j,  \
                          int storage, int thermal, hipStream_t s);                                      \
    hipError_t bouzidi(const lbk::LevelArgs* las, const lbk::CaseState* states, const lbk::Work* work, int32_t n_chunks, int64_t j,     \
                       int storage, int thermal, hipStream_t s);                                         \
    hipError_t set_state(lbk::CaseState* st, const lbk::CaseState& v, hipStream_t s);                   \
    hipError_t advance(lbk::CaseState* st, int32_t n, hipStream_t s);                             \
    hipError_t init_level(void* f0, void* f1, float* rho0, float* rho1, float* v0, float* v1, int32_t nb,  \
                          int storage, hipStream_t s);                                                   \
    hipError_t init_thermal(void* g0, void* g1, float* T0, float* T1, int32_t nb, float Tinit, int storage, \
                            hipStream_t s);                                                              \
    hipError_t forces(const lbk::ForceArgs& a, int n_groups, double* out9, hipStream_t s);               \
    hipError_t stats(const void* f, const float* rho, const float* vel, const uint8_t* obst, int32_t nb,  \
                     int storage, double* partials, int n_groups, hipStream_t s);                        \
    hipError_t finite_count(const lbk::FiniteEntry* ents, int32_t n_ents, int32_t groups_per_entry,          \
                            int32_t* counts, hipStream_t s);                                             \
    hipError_t engine_forces(const lbk::EngineForce* recs, const lbk::CaseState* states, const lbk::Work* work,  \
                             int32_t n_chunks, const int32_t* rec_idx, int32_t n_recs, int64_t j, hipStream_t s); \
    }

LBK_DECLARE(lbk_strict)
LBK_DECLARE(lbk_fast)
