/*
 * lb_oracle.c -- CPU restatement of OpenLUDWIG's lattice-Boltzmann time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * openludwig_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only to compare against or to time a CPU
 * baseline.  Nothing in the product links it.
 *
 * It is compiled twice (oracle/build.sh):
 *   liblb_oracle_f32.so  -- REAL=float : the FP32 "mirror".  Same operations in
 *                           the same order as the Julia kernels, no FMA
 *                           contraction (-ffp-contract=off), as Julia emits.
 *   liblb_oracle_f64.so  -- REAL=double: the same algorithm evaluated in FP64
 *                           (the "FP64 oracle" used for error reports).
 *
 * Arrays use the reference's own layout: f[x,y,z,b,k] column-major with
 * 1-based indices, linear (x-1)+8(y-1)+64(z-1)+512(b-1)+512*nb*(k-1)
 * (src/geometry/blocks.jl:131-189); neighbor_table (nb,27) column-major,
 * 1-based, 0 = absent (src/geometry/domain_topology.jl:135-160).
 *
 * Everything the reference does in separate passes is kept separate here on
 * purpose (copy_to_old!, f_post_collision, the Bouzidi kernel), so that the
 * HIP library's fused/aliased design is checked against the reference's
 * literal data flow.
 *
 * Parity pins: see tests/test_oracle_*.py (closed forms, invariants, brute
 * force on tiny inputs).  Rounding of CUDA.jl-compiled code (FMA contraction,
 * pow/log implementations) cannot be pinned without the reference toolchain;
 * those boundaries are absorbed by stated tolerances.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_F64
typedef double real;
#define L(x) (x)
#define RSQRT(x) sqrt(x)
#define RLOG(x) log(x)
#define RPOW(a, b) pow(a, b)
#define RCOS(x) cos(x)
#define RFLOOR(x) floor(x)
#define RMAX(a, b) fmax(a, b)
#define RMIN(a, b) fmin(a, b)
#else
typedef float real;
#define L(x) (x##f)
#define RSQRT(x) sqrtf(x)
#define RLOG(x) logf(x)
#define RPOW(a, b) powf(a, b)
#define RCOS(x) cosf(x)
#define RFLOOR(x) floorf(x)
#define RMAX(a, b) fmaxf(a, b)
#define RMIN(a, b) fminf(a, b)
#endif

#define BS 8

/* ------------------------------------------------------------------------ */
/* D3Q27 lattice: src/collision/lattice.jl:5-56, src/collision/timestep.jl:161-179
 * k (0-based) = (cx+1) + 3(cy+1) + 9(cz+1); rest k = 13.                    */
static int CX[27], CY[27], CZ[27], OPP[27], MY[27], MZ[27];
static real W[27];
static int lattice_ready = 0;

static void build_lattice(void) {
    if (lattice_ready) return;
    int k = 0;
    for (int z = -1; z <= 1; z++)
        for (int y = -1; y <= 1; y++)
            for (int x = -1; x <= 1; x++) {
                CX[k] = x; CY[k] = y; CZ[k] = z;
                int d2 = x * x + y * y + z * z;
#ifdef ORC_F64
                W[k] = d2 == 0 ? 8.0 / 27.0 : d2 == 1 ? 2.0 / 27.0 : d2 == 2 ? 1.0 / 54.0 : 1.0 / 216.0;
#else
                W[k] = d2 == 0 ? 8.0f / 27.0f : d2 == 1 ? 2.0f / 27.0f : d2 == 2 ? 1.0f / 54.0f : 1.0f / 216.0f;
#endif
                k++;
            }
    for (int i = 0; i < 27; i++)
        for (int j = 0; j < 27; j++) {
            if (CX[j] == -CX[i] && CY[j] == -CY[i] && CZ[j] == -CZ[i]) OPP[i] = j;
            if (CX[j] == CX[i] && CY[j] == -CY[i] && CZ[j] == CZ[i]) MY[i] = j;
            if (CX[j] == CX[i] && CY[j] == CY[i] && CZ[j] == -CZ[i]) MZ[i] = j;
        }
    lattice_ready = 1;
}

int orc_lattice(int* cx, int* cy, int* cz, real* w, int* opp, int* my, int* mz) {
    build_lattice();
    for (int k = 0; k < 27; k++) {
        cx[k] = CX[k]; cy[k] = CY[k]; cz[k] = CZ[k]; w[k] = W[k];
        opp[k] = OPP[k]; my[k] = MY[k]; mz[k] = MZ[k];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Equilibria: src/collision/physics_utils.jl:19-82                          */
static inline real feq2(real rho, real ux, real uy, real uz, real w, real cx, real cy, real cz) {
    real cu = cx * ux + cy * uy + cz * uz;
    real usq = ux * ux + uy * uy + uz * uz;
    return rho * w * (L(1.0) + L(3.0) * cu + L(4.5) * cu * cu - L(1.5) * usq);
}

static inline real feq4(real rho, real ux, real uy, real uz, real w, real cx, real cy, real cz) {
    real cu = cx * ux + cy * uy + cz * uz;
    real usq = ux * ux + uy * uy + uz * uz;
    real cu2 = cu * cu;
    real cu3 = cu2 * cu;
    real cu4 = cu2 * cu2;
    real usq2 = usq * usq;
    real val = L(1.0) + L(3.0) * cu + L(4.5) * cu2 - L(1.5) * usq;
    val += L(4.5) * cu3 - L(4.5) * cu * usq;
    val += L(3.375) * cu4 - L(6.75) * cu2 * usq + L(1.125) * usq2;
    return rho * w * val;
}

static inline real feq_auto(real rho, real ux, real uy, real uz, real w, real cx, real cy, real cz, int comp) {
    return comp == 1 ? feq4(rho, ux, uy, uz, w, cx, cy, cz) : feq2(rho, ux, uy, uz, w, cx, cy, cz);
}

static inline real geq(real T, real ux, real uy, real uz, real w, real cx, real cy, real cz) {
    real cu = cx * ux + cy * uy + cz * uz;
    return w * T * (L(1.0) + L(3.0) * cu);
}

/* gpu_hash / gradient_noise: src/collision/physics_utils.jl:5-16 (Int32 wrap) */
static inline uint32_t gpu_hash(int32_t x) {
    uint32_t h = (uint32_t)x;
    h = (h ^ (h >> 16)) * 0x85ebca6bu;
    h = (h ^ (h >> 13)) * 0xc2b2ae35u;
    return h ^ (h >> 16);
}

real orc_gradient_noise(int32_t gx, int32_t gy, int32_t gz, int32_t seed) {
    uint32_t c = (uint32_t)gx * 374761393u + (uint32_t)gy * 668265263u + (uint32_t)gz * 1274126177u + (uint32_t)seed;
    uint32_t h = gpu_hash((int32_t)c);
    return ((real)(h & 0xFFFFu) / L(32768.0)) - L(1.0);
}

/* ------------------------------------------------------------------------ */
/* Level state                                                               */
typedef struct {
    int level_id;
    int nb;
    int *map_x, *map_y, *map_z; /* 1-based block coords                  */
    int* nbt;                   /* (nb,27) col-major, 1-based, 0 = none  */
    int* bptr;                  /* (pdx,pdy,pdz) col-major, 1-based      */
    int pdx, pdy, pdz;
    uint8_t* obst;
    real *sponge, *wall_dist;
    real tau, tau_g;
    int has_temporal; /* f_old allocated (temporal interpolation on)  */
    real *f, *f_temp, *f_post, *f_old;
    real *rho, *rho_old, *vel, *vel_temp, *vel_old;
    int has_thermal;
    real *g, *g_temp, *g_post, *g_old, *T, *T_old;
    /* Bouzidi (sparse cell list, q per listed cell as FP16 bits)          */
    int storage; /* population storage emulated (0 FP32, 1 FP16, 2 BF16)  */
    int nbc;
    int* cb;
    int8_t *cxl, *cyl, *czl;
    uint16_t* q; /* nbc*27, q of listed cell c, direction k at [c*27+k]  */
} orc_level;

typedef struct {
    int n_levels;
    orc_level lv[16];
    int domain_nx, domain_ny, domain_nz;
    int storage;
} orc_ctx;

typedef struct {
    /* every scalar argument of stream_collide_cumulant! that is not a table
     * (src/collision/kernel_cumulant.jl:29-69; timestep.jl:40-86)           */
    float c_wale, nu_sgs_bg, inlet_turbulence;
    int symmetric, wall_model, temporal, sponge_blend;
    float omega_bulk, omega_3, omega_4, lambda_param;
    int adaptive_omega4, limiter, compressible;
    float q_min;
    /* thermal DDF (kernel_thermal.jl:14-40)                                */
    int ddf;
    float T_inlet, T_wall;
    int wall_bc_adiabatic;
} orc_params;

static inline size_t I4(const orc_level* L_, int x, int y, int z, int b) {
    return (size_t)(x - 1) + 8 * (size_t)(y - 1) + 64 * (size_t)(z - 1) + 512 * (size_t)(b - 1);
    (void)L_;
}
static inline size_t I5(const orc_level* L_, int x, int y, int z, int b, int k /*1-based*/) {
    return I4(L_, x, y, z, b) + (size_t)512 * L_->nb * (size_t)(k - 1);
}
static inline int NBT(const orc_level* L_, int b, int dir /*1-based*/) {
    return L_->nbt[(size_t)(b - 1) + (size_t)L_->nb * (dir - 1)];
}

/* f16 bits -> real (exact) */
static real half_to_real(uint16_t h) {
    int s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    double v;
    if (e == 0) v = ldexp((double)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = ldexp((double)(m | 1024), e - 25);
    return (real)(s ? -v : v);
}

/* Population storage of the HIP library (include/ludwig_hip.h LB_STORAGE_*):
 * FP16 / BF16 levels hold f - w_k rounded to nearest even and decode it as
 * w_k + stored, all in FP32.  Every value written to population storage is
 * replaced by what a later read returns.  Float -> binary16 and -> bfloat16
 * follow IEEE 754 roundTiesToEven on the bit patterns.                      */
static float round_half(float d) {
    uint32_t x;
    memcpy(&x, &d, 4);
    uint32_t sign = x & 0x80000000u, ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return d;
    float r;
    if (ax >= 0x47800000u) {
        r = INFINITY; /* |d| >= 2^16 */
    } else if (ax < 0x38800000u) {
        /* binary16 subnormals: integer multiples of 2^-24 */
        r = nearbyintf(fabsf(d) * 16777216.0f) / 16777216.0f;
    } else {
        uint32_t keep = ax & ~0x1fffu, rem = ax & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (ax & 0x2000u))) keep += 0x2000u;
        if (keep >= 0x47800000u) keep = 0x7f800000u; /* rounded past 65504 */
        memcpy(&r, &keep, 4);
    }
    uint32_t rb;
    memcpy(&rb, &r, 4);
    rb |= sign;
    memcpy(&r, &rb, 4);
    return r;
}

static float round_bf16(float d) {
    uint32_t b;
    memcpy(&b, &d, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) return d;
    b += 0x7fffu + ((b >> 16) & 1u);
    b &= 0xffff0000u;
    float r;
    memcpy(&r, &b, 4);
    return r;
}

static inline real store_round(int storage, real f, int k) {
    if (storage == 0) return f;
    float w = (float)W[k];
    float d = (float)f - w;
    float q = storage == 1 ? round_half(d) : round_bf16(d);
    return (real)(w + q);
}

/* Exported for the tests that pin the rounding against numpy / torch. */
double orc_store_round(int storage, int k, double f) {
    build_lattice();
    return (double)store_round(storage, (real)f, k);
}

static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    return p;
}

orc_ctx* orc_create(int n_levels, int domain_nx, int domain_ny, int domain_nz) {
    build_lattice();
    if (n_levels < 1 || n_levels > 16) return NULL;
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    c->n_levels = n_levels;
    c->domain_nx = domain_nx; c->domain_ny = domain_ny; c->domain_nz = domain_nz;
    return c;
}

static void free_level(orc_level* L_) {
    free(L_->map_x); free(L_->map_y); free(L_->map_z); free(L_->nbt); free(L_->bptr);
    free(L_->obst); free(L_->sponge); free(L_->wall_dist);
    free(L_->f); free(L_->f_temp); free(L_->f_post); free(L_->f_old);
    free(L_->rho); free(L_->rho_old); free(L_->vel); free(L_->vel_temp); free(L_->vel_old);
    free(L_->g); free(L_->g_temp); free(L_->g_post); free(L_->g_old); free(L_->T); free(L_->T_old);
    free(L_->cb); free(L_->cxl); free(L_->cyl); free(L_->czl); free(L_->q);
    memset(L_, 0, sizeof(*L_));
}

void orc_destroy(orc_ctx* c) {
    if (!c) return;
    for (int i = 0; i < c->n_levels; i++) free_level(&c->lv[i]);
    free(c);
}

/* Upload one level (0-based index lvl; level_id = lvl+1).  Arrays are in the
 * reference layout.  bq: dense-free sparse q: q of listed cell i, direction k
 * at bq[i*27+k] (FP16 bits).  Allocation mirrors BlockLevel (blocks.jl:102-229). */
int orc_set_level(orc_ctx* c, int lvl, int nb, const int* map_x, const int* map_y, const int* map_z,
                  const int* nbt, const int* bptr, int pdx, int pdy, int pdz, const uint8_t* obst,
                  const float* sponge, const float* wall_dist, float tau, float tau_g, int temporal, int thermal,
                  int nbc, const int* cb, const int8_t* cxl, const int8_t* cyl, const int8_t* czl,
                  const uint16_t* bq) {
    if (!c || lvl < 0 || lvl >= c->n_levels || nb < 0) return -1;
    orc_level* L_ = &c->lv[lvl];
    free_level(L_);
    L_->level_id = lvl + 1;
    L_->storage = c->storage;
    L_->nb = nb;
    size_t n4 = (size_t)512 * nb, n5 = n4 * 27;
    L_->map_x = (int*)xcalloc(nb, sizeof(int)); L_->map_y = (int*)xcalloc(nb, sizeof(int));
    L_->map_z = (int*)xcalloc(nb, sizeof(int));
    memcpy(L_->map_x, map_x, nb * sizeof(int)); memcpy(L_->map_y, map_y, nb * sizeof(int));
    memcpy(L_->map_z, map_z, nb * sizeof(int));
    L_->nbt = (int*)xcalloc((size_t)nb * 27, sizeof(int));
    memcpy(L_->nbt, nbt, (size_t)nb * 27 * sizeof(int));
    L_->pdx = pdx; L_->pdy = pdy; L_->pdz = pdz;
    size_t np = (size_t)pdx * pdy * pdz;
    L_->bptr = (int*)xcalloc(np, sizeof(int));
    memcpy(L_->bptr, bptr, np * sizeof(int));
    L_->obst = (uint8_t*)xcalloc(n4, 1); memcpy(L_->obst, obst, n4);
    L_->sponge = (real*)xcalloc(n4, sizeof(real));
    L_->wall_dist = (real*)xcalloc(n4, sizeof(real));
    for (size_t i = 0; i < n4; i++) { L_->sponge[i] = (real)sponge[i]; L_->wall_dist[i] = (real)wall_dist[i]; }
    L_->tau = (real)tau; L_->tau_g = (real)tau_g;
    L_->has_temporal = temporal && nb > 0;
    L_->f = (real*)xcalloc(n5, sizeof(real));
    L_->f_temp = (real*)xcalloc(n5, sizeof(real));
    L_->rho = (real*)xcalloc(n4, sizeof(real));
    L_->vel = (real*)xcalloc(n4 * 3, sizeof(real));
    L_->vel_temp = (real*)xcalloc(n4 * 3, sizeof(real));
    for (size_t i = 0; i < n4; i++) L_->rho[i] = L(1.0);
    if (L_->has_temporal) {
        L_->f_old = (real*)xcalloc(n5, sizeof(real));
        L_->rho_old = (real*)xcalloc(n4, sizeof(real));
        L_->vel_old = (real*)xcalloc(n4 * 3, sizeof(real));
        for (size_t i = 0; i < n4; i++) L_->rho_old[i] = L(1.0);
    }
    L_->nbc = nbc;
    if (nbc > 0) {
        L_->f_post = (real*)xcalloc(n5, sizeof(real));
        L_->cb = (int*)xcalloc(nbc, sizeof(int)); memcpy(L_->cb, cb, nbc * sizeof(int));
        L_->cxl = (int8_t*)xcalloc(nbc, 1); memcpy(L_->cxl, cxl, nbc);
        L_->cyl = (int8_t*)xcalloc(nbc, 1); memcpy(L_->cyl, cyl, nbc);
        L_->czl = (int8_t*)xcalloc(nbc, 1); memcpy(L_->czl, czl, nbc);
        L_->q = (uint16_t*)xcalloc((size_t)nbc * 27, 2); memcpy(L_->q, bq, (size_t)nbc * 27 * 2);
    }
    L_->has_thermal = thermal && nb > 0;
    if (L_->has_thermal) {
        L_->g = (real*)xcalloc(n5, sizeof(real));
        L_->g_temp = (real*)xcalloc(n5, sizeof(real));
        L_->T = (real*)xcalloc(n4, sizeof(real));
        L_->T_old = (real*)xcalloc(n4, sizeof(real));
        if (L_->has_temporal) L_->g_old = (real*)xcalloc(n5, sizeof(real));
        if (nbc > 0) L_->g_post = (real*)xcalloc(n5, sizeof(real));
    }
    return 0;
}

/* init_eq! / init_thermal_eq!: src/main.jl:113-150 */
int orc_init_equilibrium(orc_ctx* c, float T0f) {
    real T0 = (real)T0f;
    for (int l = 0; l < c->n_levels; l++) {
        orc_level* L_ = &c->lv[l];
        size_t n4 = (size_t)512 * L_->nb;
        for (int k = 0; k < 27; k++)
            for (size_t i = 0; i < n4; i++) {
                L_->f[i + n4 * k] = W[k];
                L_->f_temp[i + n4 * k] = W[k];
                if (L_->has_temporal) L_->f_old[i + n4 * k] = W[k];
            }
        if (L_->has_temporal)
            for (size_t i = 0; i < n4; i++) {
                L_->rho_old[i] = L(1.0);
                L_->vel_old[i] = 0; L_->vel_old[i + n4] = 0; L_->vel_old[i + 2 * n4] = 0;
            }
        if (L_->has_thermal) {
            for (int k = 0; k < 27; k++)
                for (size_t i = 0; i < n4; i++) {
                    real ge = store_round(L_->storage, W[k] * T0, k);
                    L_->g[i + n4 * k] = ge;
                    L_->g_temp[i + n4 * k] = ge;
                    if (L_->g_old) L_->g_old[i + n4 * k] = ge;
                    if (L_->g_post) L_->g_post[i + n4 * k] = ge;
                }
            for (size_t i = 0; i < n4; i++) { L_->T[i] = T0; L_->T_old[i] = T0; }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Parent view passed down the recursion (solver_control.jl:66-96)          */
typedef struct {
    const orc_level* P; /* parent level (tables)                            */
    const real *f_new, *rho_new, *vel_new;
    const real *f_old, *rho_old, *vel_old;
    const real *g_new, *T_new, *gvel_new, *g_old, *T_old, *gvel_old;
    real tau_parent, tau_g_parent;
    int is_l1;
} orc_parent;

/* interpolate_with_rescaling: src/collision/physics_interpolation.jl:4-127
 * Also the thermal variant (:135-248) when thermal != 0.                    */
static real interp_rescale(const orc_parent* pp, int thermal, int fgx, int fgy, int fgz, int k /*0-based*/,
                           real tau_fine, real tw, int use_temporal, int comp) {
    const orc_level* P = pp->P;
    real w_k = W[k], cx = (real)CX[k], cy = (real)CY[k], cz = (real)CZ[k];
    real pxc = ((real)fgx - L(0.5)) * L(0.5);
    real pyc = ((real)fgy - L(0.5)) * L(0.5);
    real pzc = ((real)fgz - L(0.5)) * L(0.5);
    int px0 = (int)RFLOOR(pxc), py0 = (int)RFLOOR(pyc), pz0 = (int)RFLOOR(pzc);
    int px1 = px0 + 1, py1 = py0 + 1, pz1 = pz0 + 1;
    real wx = pxc - (real)px0, wy = pyc - (real)py0, wz = pzc - (real)pz0;
    if (px0 < 1) px0 = 1;
    if (py0 < 1) py0 = 1;
    if (pz0 < 1) pz0 = 1;

    real v[8][5];
    int ok[8];
    int cxs[8] = {px0, px1, px0, px1, px0, px1, px0, px1};
    int cys[8] = {py0, py0, py1, py1, py0, py0, py1, py1};
    int czs[8] = {pz0, pz0, pz0, pz0, pz1, pz1, pz1, pz1};
    const real *a_new = thermal ? pp->g_new : pp->f_new, *a_old = thermal ? pp->g_old : pp->f_old;
    const real *s_new = thermal ? pp->T_new : pp->rho_new, *s_old = thermal ? pp->T_old : pp->rho_old;
    const real *u_new = thermal ? pp->gvel_new : pp->vel_new, *u_old = thermal ? pp->gvel_old : pp->vel_old;
    size_t pn4 = (size_t)512 * P->nb;
    for (int i = 0; i < 8; i++) {
        int pgx = cxs[i], pgy = cys[i], pgz = czs[i];
        int pbx = (pgx - 1) / BS + 1, pby = (pgy - 1) / BS + 1, pbz = (pgz - 1) / BS + 1;
        ok[i] = 0;
        v[i][0] = w_k; v[i][1] = L(1.0); v[i][2] = 0; v[i][3] = 0; v[i][4] = 0;
        if (pbx >= 1 && pbx <= P->pdx && pby >= 1 && pby <= P->pdy && pbz >= 1 && pbz <= P->pdz) {
            int pb = P->bptr[(size_t)(pbx - 1) + (size_t)P->pdx * ((pby - 1) + (size_t)P->pdy * (pbz - 1))];
            if (pb > 0) {
                int plx = (pgx - 1) % BS + 1, ply = (pgy - 1) % BS + 1, plz = (pgz - 1) % BS + 1;
                size_t i4 = I4(P, plx, ply, plz, pb);
                real fn = a_new[i4 + pn4 * k], rn = s_new[i4];
                real uxn = u_new[i4], uyn = u_new[i4 + pn4], uzn = u_new[i4 + 2 * pn4];
                if (use_temporal == 1 && tw < L(0.99)) {
                    real fo = a_old[i4 + pn4 * k], ro = s_old[i4];
                    real uxo = u_old[i4], uyo = u_old[i4 + pn4], uzo = u_old[i4 + 2 * pn4];
                    v[i][0] = fo * (L(1.0) - tw) + fn * tw;
                    v[i][1] = ro * (L(1.0) - tw) + rn * tw;
                    v[i][2] = uxo * (L(1.0) - tw) + uxn * tw;
                    v[i][3] = uyo * (L(1.0) - tw) + uyn * tw;
                    v[i][4] = uzo * (L(1.0) - tw) + uzn * tw;
                } else {
                    v[i][0] = fn; v[i][1] = rn; v[i][2] = uxn; v[i][3] = uyn; v[i][4] = uzn;
                }
                ok[i] = 1;
            }
        }
    }
    /* missing corners fall back to corner 000 (physics_interpolation.jl:89-96) */
    for (int i = 1; i < 8; i++)
        if (!ok[i])
            for (int j = 0; j < 5; j++) v[i][j] = v[0][j];
    real r[5];
    for (int j = 0; j < 5; j++) {
        real c00 = v[0][j] * (L(1.0) - wx) + v[1][j] * wx;
        real c01 = v[4][j] * (L(1.0) - wx) + v[5][j] * wx;
        real c10 = v[2][j] * (L(1.0) - wx) + v[3][j] * wx;
        real c11 = v[6][j] * (L(1.0) - wx) + v[7][j] * wx;
        real c0 = c00 * (L(1.0) - wy) + c10 * wy;
        real c1 = c01 * (L(1.0) - wy) + c11 * wy;
        r[j] = c0 * (L(1.0) - wz) + c1 * wz;
    }
    real eq, tc, tf, scale;
    if (thermal) {
        eq = geq(r[1], r[2], r[3], r[4], w_k, cx, cy, cz);
        tc = pp->tau_g_parent - L(0.5);
    } else {
        eq = feq_auto(r[1], r[2], r[3], r[4], w_k, cx, cy, cz, comp);
        tc = pp->tau_parent - L(0.5);
    }
    real neq = r[0] - eq;
    tf = tau_fine - L(0.5);
    if (tc > L(1.0e-6)) {
        scale = tf / tc;
        if (scale < L(0.01)) scale = L(0.01);
        if (scale > L(100.0)) scale = L(100.0);
    } else {
        scale = L(1.0);
    }
    return eq + neq * scale;
}

/* chimera transforms: kernel_cumulant.jl:4-26 */
static inline void chim_fwd(real* fm, real* f0, real* fp, real u) {
    real m0 = *fm + *f0 + *fp;
    real m1 = *fp - *fm;
    real m2 = *fm + *fp;
    *fm = m0;
    *f0 = m1 - u * m0;
    *fp = m2 - L(2.0) * u * m1 + u * u * m0;
}
static inline void chim_bwd(real* k0, real* k1, real* k2, real u) {
    real m0 = *k0;
    real m1 = *k1 + u * *k0;
    real m2 = *k2 + L(2.0) * u * *k1 + u * u * *k0;
    real fp = (m2 + m1) * L(0.5);
    real fm = (m2 - m1) * L(0.5);
    real f0 = m0 - m2;
    *k0 = fm; *k1 = f0; *k2 = fp;
}

/* get_velocity_neighbor: physics_utils.jl:85-110 */
static inline void vel_nb(const orc_level* L_, const real* vin, int x, int y, int z, int b, int dx, int dy, int dz,
                          real* ux, real* uy, real* uz) {
    size_t n4 = (size_t)512 * L_->nb;
    int nx = x + dx, ny = y + dy, nz = z + dz;
    size_t i4;
    if (nx >= 1 && nx <= BS && ny >= 1 && ny <= BS && nz >= 1 && nz <= BS) {
        i4 = I4(L_, nx, ny, nz, b);
    } else {
        int ox = nx < 1 ? -1 : (nx > BS ? 1 : 0);
        int oy = ny < 1 ? -1 : (ny > BS ? 1 : 0);
        int oz = nz < 1 ? -1 : (nz > BS ? 1 : 0);
        int dir = (ox + 1) + (oy + 1) * 3 + (oz + 1) * 9 + 1;
        int nbi = NBT(L_, b, dir);
        if (nbi > 0) {
            int nnx = nx < 1 ? nx + BS : (nx > BS ? nx - BS : nx);
            int nny = ny < 1 ? ny + BS : (ny > BS ? ny - BS : ny);
            int nnz = nz < 1 ? nz + BS : (nz > BS ? nz - BS : nz);
            i4 = I4(L_, nnx, nny, nnz, nbi);
        } else {
            i4 = I4(L_, x, y, z, b);
        }
    }
    *ux = vin[i4]; *uy = vin[i4 + n4]; *uz = vin[i4 + 2 * n4];
}

#define KAPPA_ L(0.41)

/* stream_collide_cumulant! for one cell: src/collision/kernel_cumulant.jl:29-627 */
static void collide_cell(const orc_level* L_, const orc_parent* pp, const orc_params* prm, real* f_out,
                         const real* f_in, real* vel_out, const real* vel_in, int store_post, real u_inlet, int nxg,
                         int nyg, int nzg, int seed, real tw, int x, int y, int z, int b) {
    size_t n4 = (size_t)512 * L_->nb;
    int bx = L_->map_x[b - 1], by = L_->map_y[b - 1], bz = L_->map_z[b - 1];
    int gx = (bx - 1) * BS + x, gy = (by - 1) * BS + y, gz = (bz - 1) * BS + z;
    size_t i4 = I4(L_, x, y, z, b);
    int is_obs = L_->obst[i4];
    real rho = 0, jx = 0, jy = 0, jz = 0;
    real fs[27];
    int comp = prm->compressible;
    for (int k = 0; k < 27; k++) {
        int cx = CX[k], cy = CY[k], cz = CZ[k];
        int sx = x - cx, sy = y - cy, sz = z - cz;
        real val = 0;
        if (sx >= 1 && sx <= BS && sy >= 1 && sy <= BS && sz >= 1 && sz <= BS) {
            val = f_in[I4(L_, sx, sy, sz, b) + n4 * k];
        } else {
            int ox = sx < 1 ? -1 : (sx > BS ? 1 : 0);
            int oy = sy < 1 ? -1 : (sy > BS ? 1 : 0);
            int oz = sz < 1 ? -1 : (sz > BS ? 1 : 0);
            int dir = (ox + 1) + (oy + 1) * 3 + (oz + 1) * 9 + 1;
            int nbi = NBT(L_, b, dir);
            if (nbi > 0) {
                int nsx = sx < 1 ? sx + BS : (sx > BS ? sx - BS : sx);
                int nsy = sy < 1 ? sy + BS : (sy > BS ? sy - BS : sy);
                int nsz = sz < 1 ? sz + BS : (sz > BS ? sz - BS : sz);
                val = f_in[I4(L_, nsx, nsy, nsz, nbi) + n4 * k];
            } else {
                int sgx = gx - cx, sgy = gy - cy, sgz = gz - cz;
                int inl = sgx < 1, outl = sgx > nxg, ymin = sgy < 1, ymax = sgy > nyg, zmin = sgz < 1,
                    zmax = sgz > nzg;
                if (inl) {
                    real noise = prm->inlet_turbulence > L(0.0)
                                     ? orc_gradient_noise(gy, gz, seed, 1234) * (real)prm->inlet_turbulence * u_inlet
                                     : L(0.0);
                    real u_inst = u_inlet + noise;
                    val = feq_auto(L(1.0), u_inst, 0, 0, W[k], (real)cx, (real)cy, (real)cz, comp);
                } else if (outl) {
                    val = feq_auto(L(1.0), u_inlet, 0, 0, W[k], (real)cx, (real)cy, (real)cz, comp);
                } else if (ymin || ymax) {
                    val = f_in[i4 + n4 * MY[k]];
                } else if (zmin || zmax) {
                    val = f_in[i4 + n4 * MZ[k]];
                } else if (!pp->is_l1) {
                    /* the HIP library stores interpolated values in ghost blocks */
                    val = store_round(L_->storage,
                                      interp_rescale(pp, 0, sgx, sgy, sgz, k, L_->tau, tw, prm->temporal, comp), k);
                } else {
                    val = W[k];
                }
            }
        }
        fs[k] = val;
        rho += val;
        jx += val * (real)CX[k];
        jy += val * (real)CY[k];
        jz += val * (real)CZ[k];
    }

    if (is_obs) {
        vel_out[i4] = 0; vel_out[i4 + n4] = 0; vel_out[i4 + 2 * n4] = 0;
        L_->rho[i4] = L(1.0);
        for (int k = 0; k < 27; k++) {
            real fc = store_round(L_->storage, fs[OPP[k]], k);
            f_out[i4 + n4 * k] = fc;
            if (store_post) L_->f_post[i4 + n4 * k] = fc;
        }
        return;
    }

    rho = RMAX(rho, L(0.01));
    real inv_rho = L(1.0) / rho;
    real ux = jx * inv_rho, uy = jy * inv_rho, uz = jz * inv_rho;
    real sp = L_->sponge[i4];
    if (sp > L(0.0)) {
        real rho_t = L(1.0), ux_t = u_inlet;
        rho = rho * (L(1.0) - sp) + rho_t * sp;
        ux = ux * (L(1.0) - sp) + ux_t * sp;
        uy = uy * (L(1.0) - sp);
        uz = uz * (L(1.0) - sp);
        if (prm->sponge_blend == 1)
            for (int k = 0; k < 27; k++) {
                real ft = feq_auto(rho_t, ux_t, 0, 0, W[k], (real)CX[k], (real)CY[k], (real)CZ[k], comp);
                fs[k] = fs[k] * (L(1.0) - sp) + ft * sp;
            }
    }
    real Fx = 0, Fy = 0, Fz = 0;
    if (prm->wall_model == 1) {
        real d = L_->wall_dist[i4];
        if (d > L(0.0) && d < L(10.0)) {
            real umag = RSQRT(ux * ux + uy * uy + uz * uz);
            real nu_visc = (L_->tau - L(0.5)) / L(3.0);
            if (umag > L(1.0e-6) && nu_visc > L(1.0e-10)) {
                real u_tau = umag * RPOW(nu_visc / (d * umag + L(1.0e-10)), L(1.0) / L(7.0)) *
                             RPOW(L(2.0) * L(8.3), -L(1.0) / L(7.0));
                u_tau = RMAX(u_tau, L(1.0e-6));
                real y_p = u_tau * d / nu_visc;
                if (y_p > L(11.81)) {
                    real upl = (L(1.0) / KAPPA_) * RLOG(y_p) + L(5.2);
                    if (upl > L(0.1)) {
                        u_tau = u_tau * ((umag / u_tau) / upl);
                        u_tau = RMAX(u_tau, L(1.0e-6));
                    }
                }
                real tau_wall = rho * u_tau * u_tau;
                real tau_res = rho * nu_visc * (umag / d);
                if (tau_wall > tau_res) {
                    real fm = (tau_wall - tau_res) / d;
                    Fx = -fm * ux / umag;
                    Fy = -fm * uy / umag;
                    Fz = -fm * uz / umag;
                }
            }
        }
    }
    real uxe = ux + L(0.5) * Fx * inv_rho;
    real uye = uy + L(0.5) * Fy * inv_rho;
    real uze = uz + L(0.5) * Fz * inv_rho;
    vel_out[i4] = ux; vel_out[i4 + n4] = uy; vel_out[i4 + 2 * n4] = uz;
    L_->rho[i4] = rho;

    /* WALE on vel_in (kernel_cumulant.jl:264-310) */
    real uE, vE, wE, uW, vW, wW, uN, vN, wN, uS, vS, wS, uT, vT, wT, uB, vB, wB;
    vel_nb(L_, vel_in, x, y, z, b, 1, 0, 0, &uE, &vE, &wE);
    vel_nb(L_, vel_in, x, y, z, b, -1, 0, 0, &uW, &vW, &wW);
    vel_nb(L_, vel_in, x, y, z, b, 0, 1, 0, &uN, &vN, &wN);
    vel_nb(L_, vel_in, x, y, z, b, 0, -1, 0, &uS, &vS, &wS);
    vel_nb(L_, vel_in, x, y, z, b, 0, 0, 1, &uT, &vT, &wT);
    vel_nb(L_, vel_in, x, y, z, b, 0, 0, -1, &uB, &vB, &wB);
    real g11 = L(0.5) * (uE - uW), g12 = L(0.5) * (uN - uS), g13 = L(0.5) * (uT - uB);
    real g21 = L(0.5) * (vE - vW), g22 = L(0.5) * (vN - vS), g23 = L(0.5) * (vT - vB);
    real g31 = L(0.5) * (wE - wW), g32 = L(0.5) * (wN - wS), g33 = L(0.5) * (wT - wB);
    real q11 = g11 * g11 + g12 * g21 + g13 * g31;
    real q12 = g11 * g12 + g12 * g22 + g13 * g32;
    real q13 = g11 * g13 + g12 * g23 + g13 * g33;
    real q21 = g21 * g11 + g22 * g21 + g23 * g31;
    real q22 = g21 * g12 + g22 * g22 + g23 * g32;
    real q23 = g21 * g13 + g22 * g23 + g23 * g33;
    real q31 = g31 * g11 + g32 * g21 + g33 * g31;
    real q32 = g31 * g12 + g32 * g22 + g33 * g32;
    real q33 = g31 * g13 + g32 * g23 + g33 * g33;
    real tr = q11 + q22 + q33;
    real trt = tr / L(3.0);
    real Sd11 = q11 - trt, Sd22 = q22 - trt, Sd33 = q33 - trt;
    real Sd12 = L(0.5) * (q12 + q21), Sd13 = L(0.5) * (q13 + q31), Sd23 = L(0.5) * (q23 + q32);
    real S12 = L(0.5) * (g12 + g21), S13 = L(0.5) * (g13 + g31), S23 = L(0.5) * (g23 + g32);
    real OP1 = Sd11 * Sd11 + Sd22 * Sd22 + Sd33 * Sd33 + L(2.0) * (Sd12 * Sd12 + Sd13 * Sd13 + Sd23 * Sd23);
    real OP2 = g11 * g11 + g22 * g22 + g33 * g33 + L(2.0) * (S12 * S12 + S13 * S13 + S23 * S23);
    real nu_eddy = 0;
    if (OP1 > L(1.0e-12)) {
        real OP1_32 = OP1 * RSQRT(OP1);
        real OP2_52 = OP2 * OP2 * RSQRT(RMAX(OP2, L(1.0e-12)));
        real den = OP2_52 + OP1 * RSQRT(RSQRT(RMAX(OP1, L(1.0e-12))));
        if (den > L(1.0e-12)) {
            real cw = (real)prm->c_wale;
            nu_eddy = (cw * cw) * OP1_32 / den;
        }
    }
    nu_eddy = RMAX(nu_eddy, (real)prm->nu_sgs_bg);
    real nu_mol = (L_->tau - L(0.5)) / L(3.0);
    real nu_eff = nu_mol + nu_eddy;
    real tau_eff = L(0.5) + L(3.0) * nu_eff;
    real om1 = L(1.0) / RMAX(tau_eff, L(0.500001));
    real omb = (real)prm->omega_bulk, om3 = (real)prm->omega_3, om4 = (real)prm->omega_4;
    real om4d = om4;
    if (prm->adaptive_omega4 == 1) {
        real a = om1 - L(2.0);
        real den = (real)prm->lambda_param * om1 + L(1.0e-10);
        real A_ = -(a * a) / den;
        om4d = A_ > L(2.0) ? L(2.0) : (A_ < L(0.01) ? L(0.01) : A_);
    }

    real A[27];
    for (int k = 0; k < 27; k++) A[k] = fs[k];
    /* forward x, y, z (0-based restatement of :335-371) */
    for (int iz = 0; iz < 3; iz++)
        for (int iy = 0; iy < 3; iy++) {
            int base = 3 * iy + 9 * iz;
            chim_fwd(&A[base], &A[base + 1], &A[base + 2], uxe);
        }
    for (int iz = 0; iz < 3; iz++)
        for (int ma = 0; ma < 3; ma++) {
            int i1 = ma + 9 * iz;
            chim_fwd(&A[i1], &A[i1 + 3], &A[i1 + 6], uye);
        }
    for (int mb = 0; mb < 3; mb++)
        for (int ma = 0; ma < 3; ma++) {
            int i1 = ma + 3 * mb;
            chim_fwd(&A[i1], &A[i1 + 9], &A[i1 + 18], uze);
        }
    /* moments -> cumulants (1-based Julia index j maps to A[j-1]) */
    real k000 = A[0];
    real k200 = A[2], k020 = A[6], k002 = A[18], k110 = A[4], k101 = A[10], k011 = A[12];
    real k210 = A[5], k120 = A[7], k201 = A[11], k021 = A[15], k102 = A[19], k012 = A[21], k111 = A[13];
    real k220 = A[8], k202 = A[20], k022 = A[24], k211 = A[14], k121 = A[16], k112 = A[22];
    real ir = L(1.0) / RMAX(k000, L(0.01));
    real c200 = k200 * ir, c020 = k020 * ir, c002 = k002 * ir;
    real c110 = k110 * ir, c101 = k101 * ir, c011 = k011 * ir;
    real c210 = k210 * ir, c120 = k120 * ir, c201 = k201 * ir, c021 = k021 * ir, c102 = k102 * ir, c012 = k012 * ir;
    real c111 = k111 * ir;
    real c220n = k220 * ir, c202n = k202 * ir, c022n = k022 * ir;
    real c211n = k211 * ir, c121n = k121 * ir, c112n = k112 * ir;
    real C220 = c220n - c200 * c020 - L(2.0) * c110 * c110;
    real C202 = c202n - c200 * c002 - L(2.0) * c101 * c101;
    real C022 = c022n - c020 * c002 - L(2.0) * c011 * c011;
    real C211 = c211n - c200 * c011 - L(2.0) * c110 * c101;
    real C121 = c121n - c020 * c101 - L(2.0) * c110 * c011;
    real C112 = c112n - c002 * c110 - L(2.0) * c101 * c011;
    real cs2 = L(0.333333333);
    real Dt = c200 + c020 + c002;
    real Dteq = L(3.0) * cs2;
    real Dts = (L(1.0) - omb) * Dt + omb * Dteq;
    real D1 = c200 - c020, D2 = c200 - c002;
    real D1s = (L(1.0) - om1) * D1, D2s = (L(1.0) - om1) * D2;
    real c200s = (Dts + D1s + D2s) / L(3.0);
    real c020s = (Dts - L(2.0) * D1s + D2s) / L(3.0);
    real c002s = (Dts + D1s - L(2.0) * D2s) / L(3.0);
    real c110s = (L(1.0) - om1) * c110, c101s = (L(1.0) - om1) * c101, c011s = (L(1.0) - om1) * c011;
    if (prm->limiter == 1) {
        real b110 = RSQRT(RMAX(c200s * c020s, L(0.0)));
        real b101 = RSQRT(RMAX(c200s * c002s, L(0.0)));
        real b011 = RSQRT(RMAX(c020s * c002s, L(0.0)));
        c110s = c110s > b110 ? b110 : (c110s < -b110 ? -b110 : c110s);
        c101s = c101s > b101 ? b101 : (c101s < -b101 ? -b101 : c101s);
        c011s = c011s > b011 ? b011 : (c011s < -b011 ? -b011 : c011s);
    }
    real c210s = (L(1.0) - om3) * c210, c120s = (L(1.0) - om3) * c120, c201s = (L(1.0) - om3) * c201;
    real c021s = (L(1.0) - om3) * c021, c102s = (L(1.0) - om3) * c102, c012s = (L(1.0) - om3) * c012;
    real c111s = (L(1.0) - om3) * c111;
    real C220s = (L(1.0) - om4d) * C220, C202s = (L(1.0) - om4d) * C202, C022s = (L(1.0) - om4d) * C022;
    real C211s = (L(1.0) - om4) * C211, C121s = (L(1.0) - om4) * C121, C112s = (L(1.0) - om4) * C112;
    real c220s = C220s + c200s * c020s + L(2.0) * c110s * c110s;
    real c202s = C202s + c200s * c002s + L(2.0) * c101s * c101s;
    real c022s = C022s + c020s * c002s + L(2.0) * c011s * c011s;
    real c211s = C211s + c200s * c011s + L(2.0) * c110s * c101s;
    real c121s = C121s + c020s * c101s + L(2.0) * c110s * c011s;
    real c112s = C112s + c002s * c110s + L(2.0) * c101s * c011s;
    real c221s = c200s * c021s + c020s * c201s + L(4.0) * c110s * c111s + L(2.0) * c101s * c120s +
                 L(2.0) * c011s * c210s;
    real c212s = c200s * c012s + c002s * c210s + L(2.0) * c110s * c102s + L(4.0) * c101s * c111s +
                 L(2.0) * c011s * c201s;
    real c122s = c020s * c102s + c002s * c120s + L(2.0) * c110s * c012s + L(4.0) * c011s * c111s +
                 L(2.0) * c101s * c021s;
    real c222s = c200s * C022s + c020s * C202s + c002s * C220s +
                 L(4.0) * (c110s * C112s + c101s * C121s + c011s * C211s);
    c222s += L(2.0) * (c210s * c012s + c201s * c021s + c120s * c102s) + L(4.0) * c111s * c111s;
    c222s += c200s * c020s * c002s + L(2.0) * (c200s * c011s * c011s + c020s * c101s * c101s + c002s * c110s * c110s) +
             L(8.0) * c110s * c101s * c011s;
    real r = k000;
    A[0] = r;
    A[1] = 0; A[3] = 0; A[9] = 0;
    A[2] = r * c200s; A[6] = r * c020s; A[18] = r * c002s;
    A[4] = r * c110s; A[10] = r * c101s; A[12] = r * c011s;
    A[5] = r * c210s; A[7] = r * c120s; A[11] = r * c201s;
    A[15] = r * c021s; A[19] = r * c102s; A[21] = r * c012s;
    A[13] = r * c111s;
    A[8] = r * c220s; A[20] = r * c202s; A[24] = r * c022s;
    A[14] = r * c211s; A[16] = r * c121s; A[22] = r * c112s;
    A[17] = r * c221s; A[23] = r * c212s; A[25] = r * c122s;
    A[26] = r * c222s;
    for (int mb = 0; mb < 3; mb++)
        for (int ma = 0; ma < 3; ma++) {
            int i1 = ma + 3 * mb;
            chim_bwd(&A[i1], &A[i1 + 9], &A[i1 + 18], uze);
        }
    for (int iz = 0; iz < 3; iz++)
        for (int ma = 0; ma < 3; ma++) {
            int i1 = ma + 9 * iz;
            chim_bwd(&A[i1], &A[i1 + 3], &A[i1 + 6], uye);
        }
    for (int iz = 0; iz < 3; iz++)
        for (int iy = 0; iy < 3; iy++) {
            int base = 3 * iy + 9 * iz;
            chim_bwd(&A[base], &A[base + 1], &A[base + 2], uxe);
        }
    if (prm->limiter == 2) {
        real fmin_ = A[0];
        for (int k = 1; k < 27; k++) fmin_ = RMIN(fmin_, A[k]);
        if (fmin_ < L(0.0)) {
            real alpha = L(1.0);
            for (int k = 0; k < 27; k++) {
                real fe = feq_auto(rho, uxe, uye, uze, W[k], (real)CX[k], (real)CY[k], (real)CZ[k], comp);
                real fn = A[k] - fe;
                if (fn < L(0.0) && fe > L(1.0e-15)) {
                    real ak = fe / (fe - fn + L(1.0e-15));
                    alpha = RMIN(alpha, ak);
                }
            }
            alpha = RMAX(alpha, L(0.0));
            for (int k = 0; k < 27; k++) {
                real fe = feq_auto(rho, uxe, uye, uze, W[k], (real)CX[k], (real)CY[k], (real)CZ[k], comp);
                A[k] = fe + alpha * (A[k] - fe);
            }
        }
    }
    for (int k = 0; k < 27; k++) {
        real cxf = (real)CX[k], cyf = (real)CY[k], czf = (real)CZ[k];
        real cu = cxf * uxe + cyf * uye + czf * uze;
        real ft = W[k] * L(3.0) *
                  ((cxf - ux + L(3.0) * cu * cxf) * Fx + (cyf - uy + L(3.0) * cu * cyf) * Fy +
                   (czf - uz + L(3.0) * cu * czf) * Fz);
        real fc = store_round(L_->storage, A[k] + (L(1.0) - L(0.5) * om1) * ft, k);
        if (store_post) L_->f_post[i4 + n4 * k] = fc;
        f_out[i4 + n4 * k] = fc;
    }
}

/* bouzidi_correction_kernel_fixed!: src/geometry/bouzidi_kernel.jl:13-92 */
static void bouzidi_apply(const orc_level* L_, real* f_out, const real* f_post, real q_min, int storage) {
    size_t n4 = (size_t)512 * L_->nb;
    for (int i = 0; i < L_->nbc; i++) {
        int b = L_->cb[i], x = L_->cxl[i], y = L_->cyl[i], z = L_->czl[i];
        size_t i4 = I4(L_, x, y, z, b);
        for (int k = 0; k < 27; k++) {
            real q = half_to_real(L_->q[(size_t)i * 27 + k]);
            if (q > q_min && q <= L(1.0)) {
                int ok = OPP[k];
                real fk = f_post[i4 + n4 * k];
                if (q < L(0.5)) {
                    int nx = x + CX[ok], ny = y + CY[ok], nz = z + CZ[ok];
                    real fff = fk;
                    if (nx >= 1 && nx <= BS && ny >= 1 && ny <= BS && nz >= 1 && nz <= BS) {
                        fff = f_post[I4(L_, nx, ny, nz, b) + n4 * k];
                    } else {
                        int ox = nx < 1 ? -1 : (nx > BS ? 1 : 0);
                        int oy = ny < 1 ? -1 : (ny > BS ? 1 : 0);
                        int oz = nz < 1 ? -1 : (nz > BS ? 1 : 0);
                        int dir = (ox + 1) + (oy + 1) * 3 + (oz + 1) * 9 + 1;
                        int nbb = NBT(L_, b, dir);
                        if (nbb > 0) {
                            int nnx = nx < 1 ? nx + BS : (nx > BS ? nx - BS : nx);
                            int nny = ny < 1 ? ny + BS : (ny > BS ? ny - BS : ny);
                            int nnz = nz < 1 ? nz + BS : (nz > BS ? nz - BS : nz);
                            fff = f_post[I4(L_, nnx, nny, nnz, nbb) + n4 * k];
                        }
                    }
                    real c1 = L(2.0) * q;
                    f_out[i4 + n4 * ok] = store_round(storage, c1 * fk + (L(1.0) - c1) * fff, ok);
                } else {
                    real fop = f_post[i4 + n4 * ok];
                    real inv2q = L(1.0) / (L(2.0) * q);
                    real c2 = (L(2.0) * q - L(1.0)) * inv2q;
                    f_out[i4 + n4 * ok] = store_round(storage, inv2q * fk + c2 * fop, ok);
                }
            }
        }
    }
}

/* stream_collide_thermal! for one cell: src/collision/kernel_thermal.jl:14-202 */
static void thermal_cell(const orc_level* L_, const orc_parent* pp, const orc_params* prm, real* g_out,
                         const real* g_in, const real* vel_arr, int store_post, real u_inlet, int nxg, int nyg,
                         int nzg, real tw, int x, int y, int z, int b) {
    size_t n4 = (size_t)512 * L_->nb;
    int bx = L_->map_x[b - 1], by = L_->map_y[b - 1], bz = L_->map_z[b - 1];
    int gx = (bx - 1) * BS + x, gy = (by - 1) * BS + y, gz = (bz - 1) * BS + z;
    size_t i4 = I4(L_, x, y, z, b);
    int is_obs = L_->obst[i4];
    real ux = vel_arr[i4], uy = vel_arr[i4 + n4], uz = vel_arr[i4 + 2 * n4];
    real Tin = (real)prm->T_inlet, Tw = (real)prm->T_wall;
    real Tsum = 0;
    real gs[27];
    for (int k = 0; k < 27; k++) {
        int cx = CX[k], cy = CY[k], cz = CZ[k];
        int sx = x - cx, sy = y - cy, sz = z - cz;
        real val = 0;
        if (sx >= 1 && sx <= BS && sy >= 1 && sy <= BS && sz >= 1 && sz <= BS) {
            val = g_in[I4(L_, sx, sy, sz, b) + n4 * k];
        } else {
            int ox = sx < 1 ? -1 : (sx > BS ? 1 : 0);
            int oy = sy < 1 ? -1 : (sy > BS ? 1 : 0);
            int oz = sz < 1 ? -1 : (sz > BS ? 1 : 0);
            int dir = (ox + 1) + (oy + 1) * 3 + (oz + 1) * 9 + 1;
            int nbi = NBT(L_, b, dir);
            if (nbi > 0) {
                int nsx = sx < 1 ? sx + BS : (sx > BS ? sx - BS : sx);
                int nsy = sy < 1 ? sy + BS : (sy > BS ? sy - BS : sy);
                int nsz = sz < 1 ? sz + BS : (sz > BS ? sz - BS : sz);
                val = g_in[I4(L_, nsx, nsy, nsz, nbi) + n4 * k];
            } else {
                int sgx = gx - cx, sgy = gy - cy, sgz = gz - cz;
                if (sgx < 1) val = geq(Tin, u_inlet, 0, 0, W[k], (real)cx, (real)cy, (real)cz);
                else if (sgx > nxg) val = geq(Tin, u_inlet, 0, 0, W[k], (real)cx, (real)cy, (real)cz);
                else if (sgy < 1 || sgy > nyg) val = g_in[i4 + n4 * MY[k]];
                else if (sgz < 1 || sgz > nzg) val = g_in[i4 + n4 * MZ[k]];
                else if (!pp->is_l1)
                    val = store_round(L_->storage,
                                      interp_rescale(pp, 1, sgx, sgy, sgz, k, L_->tau_g, tw, prm->temporal, 0), k);
                else val = W[k] * Tin;
            }
        }
        Tsum += val;
        if (is_obs) {
            real go;
            if (prm->wall_bc_adiabatic == 1) go = g_in[i4 + n4 * OPP[k]];
            else {
                real gw = geq(Tw, 0, 0, 0, W[k], (real)cx, (real)cy, (real)cz);
                go = L(2.0) * gw - g_in[i4 + n4 * OPP[k]];
            }
            go = store_round(L_->storage, go, k);
            g_out[i4 + n4 * k] = go;
            if (store_post) L_->g_post[i4 + n4 * k] = go;
        } else {
            gs[k] = val;
        }
    }
    if (is_obs) {
        L_->T[i4] = prm->wall_bc_adiabatic == 1 ? Tsum : Tw;
        return;
    }
    real Tl = RMAX(Tsum, L(0.01));
    real sp = L_->sponge[i4];
    if (sp > L(0.0)) {
        Tl = Tl * (L(1.0) - sp) + Tin * sp;
        ux = ux * (L(1.0) - sp) + u_inlet * sp;
        uy = uy * (L(1.0) - sp);
        uz = uz * (L(1.0) - sp);
    }
    L_->T[i4] = Tl;
    real omg = L(1.0) / RMAX(L_->tau_g, L(0.500001));
    for (int k = 0; k < 27; k++) {
        real gst = gs[k];
        if (sp > L(0.0) && prm->sponge_blend == 1) {
            real gt = geq(Tin, u_inlet, 0, 0, W[k], (real)CX[k], (real)CY[k], (real)CZ[k]);
            gst = gst * (L(1.0) - sp) + gt * sp;
        }
        real ge = geq(Tl, ux, uy, uz, W[k], (real)CX[k], (real)CY[k], (real)CZ[k]);
        real gc = store_round(L_->storage, gst - omg * (gst - ge), k);
        if (store_post) L_->g_post[i4 + n4 * k] = gc;
        g_out[i4 + n4 * k] = gc;
    }
}

/* perform_timestep_v2! for one level: src/collision/timestep.jl:1-159 */
static void timestep_level(orc_ctx* c, orc_level* L_, const orc_parent* pp, const orc_params* prm, real* f_out,
                           const real* f_in, real* vel_out, const real* vel_in, real* g_out, const real* g_in,
                           real u_curr, int t_sub, real tw) {
    if (L_->nb == 0) return;
    int scale = 1 << (L_->level_id - 1);
    int nxg = c->domain_nx * scale, nyg = c->domain_ny * scale, nzg = c->domain_nz * scale;
    int store_post = L_->nbc > 0;
    int seed = t_sub % 1000000;
#pragma omp parallel for schedule(static)
    for (int b = 1; b <= L_->nb; b++)
        for (int z = 1; z <= BS; z++)
            for (int y = 1; y <= BS; y++)
                for (int x = 1; x <= BS; x++)
                    collide_cell(L_, pp, prm, f_out, f_in, vel_out, vel_in, store_post, u_curr, nxg, nyg, nzg, seed,
                                 tw, x, y, z, b);
    if (L_->nbc > 0) bouzidi_apply(L_, f_out, L_->f_post, (real)prm->q_min, L_->storage);
    if (prm->ddf && g_out && g_in && L_->has_thermal) {
#pragma omp parallel for schedule(static)
        for (int b = 1; b <= L_->nb; b++)
            for (int z = 1; z <= BS; z++)
                for (int y = 1; y <= BS; y++)
                    for (int x = 1; x <= BS; x++)
                        thermal_cell(L_, pp, prm, g_out, g_in, vel_out, store_post, u_curr, nxg, nyg, nzg, tw, x, y,
                                     z, b);
        if (L_->nbc > 0) bouzidi_apply(L_, g_out, L_->g_post, (real)prm->q_min, L_->storage);
    }
}

/* Optional trace of the recursion: (level 1-based, t_sub, tw) per substep. */
static int* g_trace_lvl = NULL;
static long long* g_trace_t = NULL;
static double* g_trace_tw = NULL;
static int g_trace_cap = 0, g_trace_n = 0;
void orc_trace_set(int* lvl, long long* t, double* tw, int cap) {
    g_trace_lvl = lvl; g_trace_t = t; g_trace_tw = tw; g_trace_cap = cap; g_trace_n = 0;
}
int orc_trace_count(void) { return g_trace_n; }

/* recursive_step! / recursive_step_temporal!: src/solver/solver_control.jl:6-197 */
static void recursive_step(orc_ctx* c, int lvl, int t_sub, const orc_parent* pp, const orc_params* prm, real u,
                           real tw) {
    if (lvl >= c->n_levels) return;
    if (g_trace_cap > 0 && g_trace_n < g_trace_cap) {
        g_trace_lvl[g_trace_n] = lvl + 1; g_trace_t[g_trace_n] = t_sub; g_trace_tw[g_trace_n] = (double)tw;
        g_trace_n++;
    }
    orc_level* L_ = &c->lv[lvl];
    size_t n4 = (size_t)512 * L_->nb, n5 = n4 * 27;
    real *f_in, *f_out, *vin, *vout, *gin = NULL, *gout = NULL;
    if (t_sub % 2 == 0) { f_in = L_->f; f_out = L_->f_temp; vin = L_->vel; vout = L_->vel_temp; }
    else { f_in = L_->f_temp; f_out = L_->f; vin = L_->vel_temp; vout = L_->vel; }
    int ddf_active = prm->ddf && L_->has_thermal;
    if (ddf_active) {
        if (t_sub % 2 == 0) { gin = L_->g; gout = L_->g_temp; }
        else { gin = L_->g_temp; gout = L_->g; }
    }
    int has_children = lvl < c->n_levels - 1;
    if (has_children && prm->temporal && L_->has_temporal) {
        /* copy_to_old! (blocks.jl:241-252) */
        memcpy(L_->f_old, f_in, n5 * sizeof(real));
        memcpy(L_->rho_old, L_->rho, n4 * sizeof(real));
        memcpy(L_->vel_old, vin, 3 * n4 * sizeof(real));
        if (ddf_active && L_->g_old) {
            memcpy(L_->g_old, gin, n5 * sizeof(real));
            memcpy(L_->T_old, L_->T, n4 * sizeof(real));
        }
    }
    orc_parent self;
    const orc_parent* use = pp;
    if (pp == NULL) {
        /* level 1 passes itself as parent (timestep.jl:24-30) */
        memset(&self, 0, sizeof(self));
        self.P = L_;
        self.is_l1 = 1;
        self.tau_parent = L(0.5);
        self.tau_g_parent = L(0.5);
        use = &self;
    }
    timestep_level(c, L_, use, prm, f_out, f_in, vout, vin, gout, gin, u, t_sub, tw);
    if (has_children) {
        orc_parent ch;
        memset(&ch, 0, sizeof(ch));
        ch.P = L_;
        ch.is_l1 = 0;
        ch.f_new = f_out; ch.rho_new = L_->rho; ch.vel_new = vout;
        ch.f_old = L_->f_old; ch.rho_old = L_->rho_old; ch.vel_old = L_->vel_old;
        ch.tau_parent = L_->tau;
        if (ddf_active) {
            ch.g_new = gout; ch.T_new = L_->T; ch.gvel_new = vout;
            ch.g_old = L_->g_old; ch.T_old = L_->T_old; ch.gvel_old = L_->vel_old;
            ch.tau_g_parent = L_->tau_g;
        } else {
            ch.tau_g_parent = L(0.5);
        }
        recursive_step(c, lvl + 1, 2 * t_sub, &ch, prm, u, L(0.0));
        recursive_step(c, lvl + 1, 2 * t_sub + 1, &ch, prm, u, L(0.5));
    }
}

/* execute_timestep_batch!: src/solver/solver_control.jl:199-219 */
int orc_step_batch(orc_ctx* c, int t_start, int batch, float u_curr, const orc_params* prm) {
    if (!c || !prm) return -1;
    for (int t = t_start; t < t_start + batch; t++) recursive_step(c, 0, t, NULL, prm, (real)u_curr, L(0.0));
    return 0;
}

/* Ramp of main.jl:217-218: prog = 0.5(1-cos(Float32(pi)*batch_end/RAMP)) */
float orc_ramp_velocity(float u_target, int batch_end, int ramp_steps) {
    if (batch_end <= ramp_steps) {
        float pf = (float)3.14159265358979323846;
        float a = pf * (float)batch_end;
        a = a / (float)ramp_steps;
        float prog = 0.5f * (1.0f - cosf(a));
        return u_target * prog;
    }
    return u_target;
}

/* Download: which = 0 f (latest), 1 rho, 2 vel (latest), 3 g, 4 T.
 * "latest" uses the parity of the step the level last made (t_last).       */
int orc_get(const orc_ctx* c, int lvl, int which, int t_last, double* out) {
    if (!c || lvl < 0 || lvl >= c->n_levels) return -1;
    const orc_level* L_ = &c->lv[lvl];
    size_t n4 = (size_t)512 * L_->nb, n = 0;
    const real* src = NULL;
    int odd = t_last % 2 != 0;
    switch (which) {
        case 0: src = odd ? L_->f : L_->f_temp; n = n4 * 27; break;
        case 1: src = L_->rho; n = n4; break;
        case 2: src = odd ? L_->vel : L_->vel_temp; n = n4 * 3; break;
        case 3: if (!L_->has_thermal) return -2; src = odd ? L_->g : L_->g_temp; n = n4 * 27; break;
        case 4: if (!L_->has_thermal) return -2; src = L_->T; n = n4; break;
        default: return -3;
    }
    for (size_t i = 0; i < n; i++) out[i] = (double)src[i];
    return 0;
}

/* Overwrite the current (t_last-parity) f buffer, rho, and vel of a level; used
 * by tests that start the oracle and the GPU from the same perturbed state. */
int orc_put(orc_ctx* c, int lvl, int which, int t_last, const double* in) {
    if (!c || lvl < 0 || lvl >= c->n_levels) return -1;
    orc_level* L_ = &c->lv[lvl];
    size_t n4 = (size_t)512 * L_->nb, n = 0;
    real* dst = NULL;
    int odd = t_last % 2 != 0;
    switch (which) {
        case 0: dst = odd ? L_->f : L_->f_temp; n = n4 * 27; break;
        case 1: dst = L_->rho; n = n4; break;
        case 2: dst = odd ? L_->vel : L_->vel_temp; n = n4 * 3; break;
        case 3: if (!L_->has_thermal) return -2; dst = odd ? L_->g : L_->g_temp; n = n4 * 27; break;
        case 4: if (!L_->has_thermal) return -2; dst = L_->T; n = n4; break;
        default: return -3;
    }
    for (size_t i = 0; i < n; i++)
        dst[i] = (which == 0 || which == 3) ? store_round(L_->storage, (real)in[i], (int)(i / n4)) : (real)in[i];
    return 0;
}

/* Emulate the HIP library's population storage on every level. */
int orc_set_storage(orc_ctx* c, int storage) {
    if (!c || storage < 0 || storage > 2) return -1;
    c->storage = storage;
    for (int l = 0; l < c->n_levels; l++) c->lv[l].storage = storage;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Single-cell probes used by closed-form tests of the collision operator.  */
/* Collide one isolated cell whose 27 pulled populations are `fin`, with
 * uniform neighbour velocity `u_nb` (so the WALE gradient is zero).  Returns
 * post-collision populations in fout, rho/u in macro[4].                   */
int orc_collide_single(const double* fin, const double* u_nb, const double* grad, float tau, const orc_params* prm,
                       double sponge, double wall_dist, float u_inlet, double* fout, double* macro) {
    build_lattice();
    orc_level Lv;
    memset(&Lv, 0, sizeof(Lv));
    Lv.level_id = 1;
    Lv.nb = 1;
    int mx = 1, my = 1, mz = 1;
    Lv.map_x = &mx; Lv.map_y = &my; Lv.map_z = &mz;
    int nbt[27];
    for (int d = 0; d < 27; d++) nbt[d] = 1; /* periodic self-neighbour */
    Lv.nbt = nbt;
    int bptr = 1;
    Lv.bptr = &bptr; Lv.pdx = Lv.pdy = Lv.pdz = 1;
    uint8_t obst[512];
    real spg[512], wd[512], rho[512], vin[512 * 3], vout[512 * 3];
    static real fi[512 * 27], fo[512 * 27];
    memset(obst, 0, sizeof(obst));
    for (int i = 0; i < 512; i++) {
        spg[i] = (real)sponge; wd[i] = (real)wall_dist; rho[i] = 1;
        /* previous velocity u_nb + G (x - x0), x0 = the probed cell (4,4,4) */
        double px = (double)(i % 8) - 3.0, py = (double)((i / 8) % 8) - 3.0, pz = (double)(i / 64) - 3.0;
        for (int d = 0; d < 3; d++) {
            double v = u_nb[d];
            if (grad) v += grad[3 * d + 0] * px + grad[3 * d + 1] * py + grad[3 * d + 2] * pz;
            vin[i + 512 * d] = (real)v;
        }
        for (int k = 0; k < 27; k++) fi[i + 512 * k] = (real)fin[k];
    }
    Lv.obst = obst; Lv.sponge = spg; Lv.wall_dist = wd; Lv.rho = rho; Lv.tau = (real)tau;
    orc_parent self;
    memset(&self, 0, sizeof(self));
    self.P = &Lv; self.is_l1 = 1; self.tau_parent = L(0.5);
    /* a uniform periodic block: every pull returns fin[k] */
    collide_cell(&Lv, &self, prm, fo, fi, vout, vin, 0, (real)u_inlet, 1 << 20, 1 << 20, 1 << 20, 0, L(0.0), 4, 4, 4,
                 1);
    size_t i4 = 3 + 8 * 3 + 64 * 3;
    for (int k = 0; k < 27; k++) fout[k] = (double)fo[i4 + 512 * k];
    macro[0] = (double)rho[i4];
    macro[1] = (double)vout[i4]; macro[2] = (double)vout[i4 + 512]; macro[3] = (double)vout[i4 + 1024];
    return 0;
}

/* Chimera round trip (forward then backward) on 27 values. */
int orc_chimera_roundtrip(const double* fin, double ux, double uy, double uz, double* fout) {
    real A[27];
    for (int k = 0; k < 27; k++) A[k] = (real)fin[k];
    for (int iz = 0; iz < 3; iz++)
        for (int iy = 0; iy < 3; iy++) { int b = 3 * iy + 9 * iz; chim_fwd(&A[b], &A[b + 1], &A[b + 2], (real)ux); }
    for (int iz = 0; iz < 3; iz++)
        for (int ma = 0; ma < 3; ma++) { int i1 = ma + 9 * iz; chim_fwd(&A[i1], &A[i1 + 3], &A[i1 + 6], (real)uy); }
    for (int mb = 0; mb < 3; mb++)
        for (int ma = 0; ma < 3; ma++) { int i1 = ma + 3 * mb; chim_fwd(&A[i1], &A[i1 + 9], &A[i1 + 18], (real)uz); }
    for (int mb = 0; mb < 3; mb++)
        for (int ma = 0; ma < 3; ma++) { int i1 = ma + 3 * mb; chim_bwd(&A[i1], &A[i1 + 9], &A[i1 + 18], (real)uz); }
    for (int iz = 0; iz < 3; iz++)
        for (int ma = 0; ma < 3; ma++) { int i1 = ma + 9 * iz; chim_bwd(&A[i1], &A[i1 + 3], &A[i1 + 6], (real)uy); }
    for (int iz = 0; iz < 3; iz++)
        for (int iy = 0; iy < 3; iy++) { int b = 3 * iy + 9 * iz; chim_bwd(&A[b], &A[b + 1], &A[b + 2], (real)ux); }
    for (int k = 0; k < 27; k++) fout[k] = (double)A[k];
    return 0;
}

/* Central moments (forward transform only), for closed-form checks. */
int orc_chimera_forward(const double* fin, double ux, double uy, double uz, double* kout) {
    real A[27];
    for (int k = 0; k < 27; k++) A[k] = (real)fin[k];
    for (int iz = 0; iz < 3; iz++)
        for (int iy = 0; iy < 3; iy++) { int b = 3 * iy + 9 * iz; chim_fwd(&A[b], &A[b + 1], &A[b + 2], (real)ux); }
    for (int iz = 0; iz < 3; iz++)
        for (int ma = 0; ma < 3; ma++) { int i1 = ma + 9 * iz; chim_fwd(&A[i1], &A[i1 + 3], &A[i1 + 6], (real)uy); }
    for (int mb = 0; mb < 3; mb++)
        for (int ma = 0; ma < 3; ma++) { int i1 = ma + 3 * mb; chim_fwd(&A[i1], &A[i1 + 9], &A[i1 + 18], (real)uz); }
    for (int k = 0; k < 27; k++) kout[k] = (double)A[k];
    return 0;
}

double orc_feq(int comp, double rho, double ux, double uy, double uz, int k) {
    build_lattice();
    return (double)feq_auto((real)rho, (real)ux, (real)uy, (real)uz, W[k], (real)CX[k], (real)CY[k], (real)CZ[k],
                            comp);
}

/* Interpolation probe on a parent level already set at index plvl, using its
 * current buffers as "new" and f_old/rho_old/vel_old as "old".  Used for the
 * affine-reproduction and uniform-equilibrium closed forms.                */
double orc_interp_probe(orc_ctx* c, int plvl, int t_last, int fgx, int fgy, int fgz, int k, float tau_fine, float tw,
                        int temporal, int comp) {
    orc_level* P = &c->lv[plvl];
    orc_parent pp;
    memset(&pp, 0, sizeof(pp));
    int odd = t_last % 2 != 0;
    pp.P = P;
    pp.f_new = odd ? P->f : P->f_temp;
    pp.rho_new = P->rho;
    pp.vel_new = odd ? P->vel : P->vel_temp;
    pp.f_old = P->f_old ? P->f_old : pp.f_new;
    pp.rho_old = P->rho_old ? P->rho_old : P->rho;
    pp.vel_old = P->vel_old ? P->vel_old : pp.vel_new;
    pp.tau_parent = P->tau;
    return (double)interp_rescale(&pp, 0, fgx, fgy, fgz, k, (real)tau_fine, (real)tw, temporal, comp);
}


/* ------------------------------------------------------------------------ */
/* Surface forces: map_stresses_kernel! + integrate_forces_kernel!
 * (src/forces/surface.jl:143-507, 702-762) on the finest level, per-triangle
 * terms in FP32 as the reference computes them, summed in FP64 in triangle
 * order (the reference's FP32 atomics have no defined order).
 * Reads level.rho and level.vel -- the vel buffer of even-numbered storage
 * (blocks.jl: `vel`), exactly the arrays surface.jl:784 passes.            */
static int f_cell_at(const orc_level* P, int gx, int gy, int gz, int* b, size_t* i4) {
    if (gx < 1 || gy < 1 || gz < 1) return 0;
    int bx = (gx - 1) / BS + 1, by = (gy - 1) / BS + 1, bz = (gz - 1) / BS + 1;
    if (bx < 1 || bx > P->pdx || by < 1 || by > P->pdy || bz < 1 || bz > P->pdz) return 0;
    int id = P->bptr[(size_t)(bx - 1) + (size_t)P->pdx * ((by - 1) + (size_t)P->pdy * (bz - 1))];
    if (id <= 0) return 0;
    *b = id;
    *i4 = I4(P, (gx - 1) % BS + 1, (gy - 1) % BS + 1, (gz - 1) % BS + 1, id);
    return 1;
}

int orc_forces(const orc_ctx* c, int t_last, const float* cxt, const float* cyt, const float* czt, const float* nxt,
               const float* nyt, const float* nzt, const float* area, long long n_tri, const float* off,
               const float* ref, float dx, float pressure_scale, float stress_scale, int search_radius,
               float c_wale, float nu_bg, double* out9, float* maps) {
    (void)t_last;
    const orc_level* P = &c->lv[c->n_levels - 1];
    size_t n4 = (size_t)512 * P->nb;
    const real* rho_a = P->rho;
    const real* vel_a = P->vel;
    for (int j = 0; j < 9; j++) out9[j] = 0.0;
    float tau_mol = (float)P->tau;
    for (long long i = 0; i < n_tri; i++) {
        float tx = cxt[i] + off[0], ty = cyt[i] + off[1], tz = czt[i] + off[2];
        float n_x = nxt[i], n_y = nyt[i], n_z = nzt[i];
        float brho = 1.0f, bux = 0.0f, buy = 0.0f, buz = 0.0f, bwd = 1.0f, bq = 0.0f;
        int found = 0;
        for (int pi = 1; pi <= search_radius; pi++) {
            float dc = (float)pi + 0.5f;
            float px = tx + n_x * dc * dx, py = ty + n_y * dc * dx, pz = tz + n_z * dc * dx;
            /* trilinear_sample_flow (surface.jl:224-346) */
            float inv_dx = 1.0f / dx;
            float fx = px * inv_dx - 0.5f, fy = py * inv_dx - 0.5f, fz = pz * inv_dx - 0.5f;
            int ix0 = (int)floorf(fx) + 1, iy0 = (int)floorf(fy) + 1, iz0 = (int)floorf(fz) + 1;
            float wx = fx - (float)(ix0 - 1), wy = fy - (float)(iy0 - 1), wz = fz - (float)(iz0 - 1);
            wx = wx < 0.0f ? 0.0f : (wx > 1.0f ? 1.0f : wx);
            wy = wy < 0.0f ? 0.0f : (wy > 1.0f ? 1.0f : wy);
            wz = wz < 0.0f ? 0.0f : (wz > 1.0f ? 1.0f : wz);
            float w0x = 1.0f - wx, w1x = wx, w0y = 1.0f - wy, w1y = wy, w0z = 1.0f - wz, w1z = wz;
            float racc = 0, uacc = 0, vacc = 0, zacc = 0, wacc = 0;
            for (int q = 0; q < 8; q++) {
                int dxi = q & 1, dyi = (q >> 1) & 1, dzi = (q >> 2) & 1;
                float wk = (dxi ? w1x : w0x) * (dyi ? w1y : w0y) * (dzi ? w1z : w0z);
                int b;
                size_t j4;
                if (f_cell_at(P, ix0 + dxi, iy0 + dyi, iz0 + dzi, &b, &j4) && !P->obst[j4]) {
                    racc += wk * (float)rho_a[j4];
                    uacc += wk * (float)vel_a[j4];
                    vacc += wk * (float)vel_a[j4 + n4];
                    zacc += wk * (float)vel_a[j4 + 2 * n4];
                    wacc += wk;
                }
            }
            int ok = wacc > 0.01f;
            float r = 1.0f, u = 0.0f, v = 0.0f, w = 0.0f, wq = 0.0f;
            if (ok) {
                float iw = 1.0f / wacc;
                r = racc * iw; u = uacc * iw; v = vacc * iw; w = zacc * iw; wq = wacc;
            }
            if (ok && wq >= 0.5f) {
                brho = r; bux = u; buy = v; buz = w; bwd = dc; bq = wq; found = 1;
                break;
            } else if (ok && wq > bq) {
                brho = r; bux = u; buy = v; buz = w; bwd = dc; bq = wq; found = 1;
            }
        }
        if (!found) {
            float gxf = tx / dx, gyf = ty / dx, gzf = tz / dx;
            int g_x = (int)floorf(gxf) + 1, g_y = (int)floorf(gyf) + 1, g_z = (int)floorf(gzf) + 1;
            float best = 1e10f;
            for (int rad = 1; rad <= search_radius; rad++) {
                if (found) break;
                for (int dz = -rad; dz <= rad; dz++)
                    for (int dy = -rad; dy <= rad; dy++)
                        for (int ddx = -rad; ddx <= rad; ddx++) {
                            int shell = abs(ddx) == rad || abs(dy) == rad || abs(dz) == rad;
                            if (!shell) continue;
                            int cgx = g_x + ddx, cgy = g_y + dy, cgz = g_z + dz, b;
                            size_t j4;
                            if (!f_cell_at(P, cgx, cgy, cgz, &b, &j4) || P->obst[j4]) continue;
                            float ccx = ((float)cgx - 0.5f) * dx, ccy = ((float)cgy - 0.5f) * dx,
                                  ccz = ((float)cgz - 0.5f) * dx;
                            float dxp = ccx - tx, dyp = ccy - ty, dzp = ccz - tz;
                            float dn = dxp * n_x + dyp * n_y + dzp * n_z;
                            if (dn <= 0.0f) continue;
                            float dsq = dxp * dxp + dyp * dyp + dzp * dzp;
                            if (dsq < best) {
                                best = dsq;
                                brho = (float)rho_a[j4];
                                bux = (float)vel_a[j4]; buy = (float)vel_a[j4 + n4]; buz = (float)vel_a[j4 + 2 * n4];
                                bwd = dn / dx;
                                found = 1;
                            }
                        }
            }
        }
        float p = 0.0f, tx_ = 0.0f, ty_ = 0.0f, tz_ = 0.0f;
        if (found) {
            float wd = bwd > 0.5f ? bwd : 0.5f;
            float pg = (brho - 1.0f) / 3.0f;
            p = pg * pressure_scale;
            float und = bux * n_x + buy * n_y + buz * n_z;
            float utx = bux - und * n_x, uty = buy - und * n_y, utz = buz - und * n_z;
            float utm = sqrtf(utx * utx + uty * uty + utz * utz);
            float nu_mol = (tau_mol - 0.5f) / 3.0f;
            float nu_sgs = 0.0f;
            if (c_wale > 0.0f && wd > 0.01f && utm > 1.0e-10f) {
                float S = utm / wd;
                nu_sgs = (c_wale * c_wale) * S;
            }
            nu_sgs = nu_sgs > nu_bg ? nu_sgs : nu_bg;
            float nu_eff = nu_mol + nu_sgs;
            if (utm > 1.0e-10f && wd > 0.01f) {
                float tl = brho * nu_eff * utm / wd;
                float tp = tl * stress_scale;
                tx_ = (utx / utm) * tp;
                ty_ = (uty / utm) * tp;
                tz_ = (utz / utm) * tp;
            }
        }
        if (maps) {
            maps[i] = p; maps[i + n_tri] = tx_; maps[i + 2 * n_tri] = ty_; maps[i + 3 * n_tri] = tz_;
        }
        float A = area[i];
        float dFpx = -p * n_x * A, dFpy = -p * n_y * A, dFpz = -p * n_z * A;
        float dFvx = tx_ * A, dFvy = ty_ * A, dFvz = tz_ * A;
        float dFx = dFpx + dFvx, dFy = dFpy + dFvy, dFz = dFpz + dFvz;
        float rx = (cxt[i] + off[0]) - ref[0], ry = (cyt[i] + off[1]) - ref[1], rz = (czt[i] + off[2]) - ref[2];
        out9[0] += (double)dFpx; out9[1] += (double)dFpy; out9[2] += (double)dFpz;
        out9[3] += (double)dFvx; out9[4] += (double)dFvy; out9[5] += (double)dFvz;
        out9[6] += (double)(ry * dFz - rz * dFy);
        out9[7] += (double)(rz * dFx - rx * dFz);
        out9[8] += (double)(rx * dFy - ry * dFx);
    }
    return 0;
}
