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
//   f[b][k][c]   c = x + 8y + 64z (0-based local cell), float
//   rho[b][c], vel[b][d][c], obst[b][c], sponge[b][c], wall_dist[b][c]
//   nbr[b][27]   0-based neighbour block, -1 absent (dir as k)
//   borg[b][3]   global 1-based cell coordinate of local (0,0,0) minus 1
//   grid         dense block pointer (max active coords), 0-based ids, -1 absent
struct CollideArgs {
    const void* f_in;        // populations, storage `storage` (0 FP32, 1 FP16 offset, 2 BF16 offset)
    void* f_out;
    int32_t storage;
    const float* vel_in;
    float* vel_out;
    float* rho_out;
    const uint8_t* obst;
    const float* sponge;
    const float* wall_dist;  // may be null when wall_model == 0
    const int32_t* nbr;
    const int32_t* borg;
    const int32_t* order;  // processing order (Morton-sorted block indices)
    int32_t nb;
    int32_t nxg, nyg, nzg;
    float tau;
    float u_inlet;
    float c_wale, nu_sgs_bg, inlet_turb;
    float omega_bulk, omega3, omega4, lambda;
    int32_t adaptive, limiter, comp, wall_model, sponge_blend;
    int32_t seed;
    int32_t is_l1;
};

struct ThermalArgs {
    const float* g_in;
    float* g_out;
    const float* vel;  // the flow step's vel_out
    float* T_out;
    const uint8_t* obst;
    const float* sponge;
    const int32_t* nbr;
    const int32_t* borg;
    const int32_t* order;
    int32_t nb;
    int32_t nxg, nyg, nzg;
    float tau_g;
    float u_inlet, T_inlet, T_wall;
    int32_t adiabatic, sponge_blend;
    int32_t is_l1;
};

// Bouzidi link list: corrected slot dst = a*f[s1] + b*f[s2], all indices
// into one population buffer (block-major layout).
struct BouzidiLinks {
    int64_t n;
    const int64_t* dst;
    const int64_t* s1;
    const int64_t* s2;
    const float* a;
    const float* b;
    float* tmp;
};

// Ghost blocks: in-domain positions next to a level's blocks that the level
// does not hold.  Every link pulled from such a position takes the
// coarse-to-fine interpolation, a function of the source cell and direction
// only, so it is evaluated once per (ghost cell, direction) into the ghost
// block's slots before the level's stream-collide.
struct GhostArgs {
    void* f;                   // populations being pulled (ghost blocks at index >= nb)
    int32_t storage;
    const int32_t* entries;    // [n][2]: (ghost block << 9 | cell, 27-bit direction mask)
    int64_t n;
    const int32_t* borg;       // block origins, ghost blocks included
    const void* a_new;
    const float* s_new;
    const float* u_new;
    const void* a_old;
    const float* s_old;
    const float* u_old;
    const int32_t* pgrid;
    int32_t pdx, pdy, pdz;
    float tau_fine, tau_coarse, tw;
    int32_t temporal, comp, thermal;
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
};

}  // namespace lbk

#define LBK_DECLARE(NS)                                                                                  \
    namespace NS {                                                                                       \
    hipError_t collide(const lbk::CollideArgs& a, hipStream_t s);                                        \
    hipError_t thermal(const lbk::ThermalArgs& a, hipStream_t s);                                        \
    hipError_t bouzidi(const lbk::BouzidiLinks& l, void* f, int storage, hipStream_t s);                 \
    hipError_t ghost_fill(const lbk::GhostArgs& g, hipStream_t s);                                       \
    hipError_t init_level(void* f0, void* f1, float* rho0, float* rho1, float* v0, float* v1, int32_t nb,  \
                          int storage, hipStream_t s);                                                   \
    hipError_t init_thermal(float* g0, float* g1, float* T0, float* T1, int32_t nb, float Tinit,         \
                            hipStream_t s);                                                              \
    hipError_t forces(const lbk::ForceArgs& a, int n_groups, double* out9, hipStream_t s);               \
    hipError_t stats(const void* f, const float* rho, const float* vel, const uint8_t* obst, int32_t nb,  \
                     int storage, double* partials, int n_groups, hipStream_t s);                        \
    }

LBK_DECLARE(lbk_strict)
LBK_DECLARE(lbk_fast)
