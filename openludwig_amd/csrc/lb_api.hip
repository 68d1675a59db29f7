// lb_api.hip -- C ABI of libludwig_hip.so (declared in include/ludwig_hip.h).
//
// Host side of the time step: device state per level, the 2:1 multilevel
// recursion of execute_timestep_batch! / recursive_step!
// (src/solver/solver_control.jl:6-219), and the layout conversions between
// the reference's column-major arrays and the device's block-major SoA.
//
// Differences in data flow from the reference, each exact:
//   * f_old / vel_old / g_old are not copies: a level with children keeps its
//     step's input buffers (f_in, vel_in, g_in) untouched until its next step,
//     which comes after both child substeps, so the children read them in
//     place (copy_to_old!, src/geometry/blocks.jl:241-252).  rho and T, single
//     buffers in the reference, are ping-ponged by step parity so the previous
//     step's values survive as rho_old / T_old.
//   * f_post_collision is not allocated: Bouzidi gathers its inputs from the
//     uncorrected f_out before scattering the corrected slots.
//   * No device synchronisation between launches; everything is ordered on
//     the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "ludwig_hip.h"
#include "lb_kernels.h"
#include "lb_sync.h"

namespace {

constexpr int NC = 512;
constexpr int Q = 27;

struct Kern {
    hipError_t (*collide)(const lbk::LevelArgs*, const lbk::CaseState*, const lbk::Work*, int32_t, int64_t, int, int,
                          int, int, int, hipStream_t);
    hipError_t (*thermal)(const lbk::LevelArgs*, const lbk::CaseState*, const lbk::Work*, int32_t, int64_t, int, int,
                          int, int, hipStream_t);
    hipError_t (*ghost_fill)(const lbk::LevelArgs*, const lbk::CaseState*, const lbk::Work*, int32_t, int64_t, int, int,
                             hipStream_t);
    hipError_t (*bouzidi)(const lbk::LevelArgs*, const lbk::CaseState*, const lbk::Work*, int32_t, int64_t, int, int,
                          hipStream_t);
    hipError_t (*set_state)(lbk::CaseState*, const lbk::CaseState&, hipStream_t);
    hipError_t (*advance)(lbk::CaseState*, int32_t, hipStream_t);
    hipError_t (*init_level)(void*, void*, float*, float*, float*, float*, int32_t, int, hipStream_t);
    hipError_t (*init_thermal)(void*, void*, float*, float*, int32_t, float, int, hipStream_t);
    hipError_t (*forces)(const lbk::ForceArgs&, int, double*, hipStream_t);
    hipError_t (*stats)(const void*, const float*, const float*, const uint8_t*, int32_t, int, double*, int,
                        hipStream_t);
    hipError_t (*finite_count)(const lbk::FiniteEntry*, int32_t, int32_t, int32_t*, hipStream_t);
    hipError_t (*engine_forces)(const lbk::EngineForce*, const lbk::CaseState*, const lbk::Work*, int32_t,
                                const int32_t*, int32_t, int64_t, hipStream_t);
};

#define LBK_TABLE(NS)                                                                                     \
    {NS::collide, NS::thermal, NS::ghost_fill, NS::bouzidi, NS::set_state, NS::advance, NS::init_level,  \
     NS::init_thermal, NS::forces, NS::stats, NS::finite_count, NS::engine_forces}
const Kern kStrict = LBK_TABLE(lbk_strict);
const Kern kFast = LBK_TABLE(lbk_fast);

struct Level {
    bool uploaded = false;
    int id = 0;
    int32_t nb = 0;
    double dx = 0;
    float tau = 0, tau_g = 0;
    void* f[2] = {nullptr, nullptr};  // 4 or 2 bytes per population (lb_config.storage)
    float* rho[2] = {nullptr, nullptr};
    float* vel[2] = {nullptr, nullptr};
    void* g[2] = {nullptr, nullptr};  // thermal populations, same storage as f
    float* T[2] = {nullptr, nullptr};
    uint8_t* obst = nullptr;
    float* sponge = nullptr;
    float* wall_dist = nullptr;
    uint8_t* bflags = nullptr;  // per block: any sponge / near-wall / solid cell
    int32_t* nbr = nullptr;
    int32_t* borg = nullptr;
    int32_t* order = nullptr;
    int32_t* grid = nullptr;
    // ghost blocks (coarse-to-fine interface sources)
    int32_t ngh = 0;
    int64_t ngent = 0;
    int32_t* gent = nullptr;
    int32_t* gpar = nullptr;     // FP32: per ghost entry, its 8 parent cells, (block << 9) | cell, or -1
    int32_t* gstart = nullptr;   // 16-bit: per ghost block, its first entry (ngh + 1 values)
    int32_t* gregion = nullptr;  // 16-bit: per ghost block, its 125 parent cells, (block << 9) | cell, or -1
    uint32_t* gbmask = nullptr;  // 16-bit: per ghost block, the directions any of its entries serves
    int32_t pd[3] = {0, 0, 0};
    std::vector<int32_t> grid_host;  // host copy of grid, for the child level's ghost parents
    // Bouzidi link list
    int64_t nlinks = 0;
    // Bouzidi link slots: uint32 (lidx32, a population buffer below 2^32
    // slots) or int64
    void* ldst = nullptr;
    void* ls1 = nullptr;
    void* ls2 = nullptr;
    bool lidx32 = false;
    float* la = nullptr;
    float* lb = nullptr;
    float* ltmp = nullptr;
    bool wide = false;    // population buffer >= 4 GiB: kernels use 64-bit offsets
    bool wide_g = false;  // the same for the thermal populations
    // work lists over this level alone (LevelArgs index = level index):
    // blocks in processing order, ghost-entry chunks, Bouzidi-link chunks
    lbk::Work* wcol = nullptr;
    int32_t n_gchunks = 0;  // ghost fill work items: 256-entry chunks (FP32) or ghost blocks (16-bit)
    lbk::Work* wgh = nullptr;
    int32_t n_lchunks = 0;
    lbk::Work* wlk = nullptr;
    int cur = 0;        // buffer index written by the last step (0 before any step)
    std::vector<int32_t> order_host;  // processing order (engine work lists)
    // Device block i holds the caller's block perm[i] (0-based); the device
    // order is the slab order of lb_level_upload.
    std::vector<int32_t> perm;
    int32_t n_int = 0;                // interior blocks: the first n_int of the order
    // kernel timing: event pairs recorded around collide launches
    // [0] stream-collide launches, [1] thermal launches
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[2];
    size_t ev_used[2] = {0, 0};
};

}  // namespace

struct lb_ctx {
    int device = 0;
    lb_config cfg{};
    hipStream_t stream = nullptr;
    const Kern* k = &kStrict;
    std::vector<Level> levels;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    std::string err;
    // mesh
    int64_t n_tri = 0;
    float* mesh[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    float* maps[4] = {nullptr, nullptr, nullptr, nullptr};
    double* fpart = nullptr;
    double* spart = nullptr;
    int n_stat_groups = 0;
    // divergence check: one entry per level and a counter (allocated at creation)
    lbk::FiniteEntry* fin_ents = nullptr;
    int32_t* fin_count = nullptr;
    bool profile = false;
    // device-resident kernel arguments: the case record and one LevelArgs per level
    lbk::CaseState* st_dev = nullptr;
    lbk::LevelArgs* las_dev = nullptr;
    std::vector<lbk::LevelArgs> las_host;
    lb_engine* engine = nullptr;  // set while attached to a batched engine
    bool dry = false;             // footprint planning: nothing is allocated or copied
};

// One attached case of an engine.
struct EngineCase {
    lb_ctx* c = nullptr;
    int64_t n_total = 0;
    int64_t t0 = 1;          // next coarse step (host copy of the device record)
    int32_t t_inc = 1;
    int32_t batch = 8;
    float* u_tab = nullptr;  // device, owned by the case context
    lb_step_params p{};
    // force sampling inside the engine step (lb_engine_sample_forces)
    double* acc = nullptr;   // device [10], owned by the case context
    bool sample = false;
    int64_t diag = 0, ramp = 0;
    lb_force_params fp{};
};

// Work lists of one depth of the engine's recursion (concatenated over cases).
struct DepthWork {
    lbk::Work *col = nullptr, *colb = nullptr, *gh = nullptr, *lk = nullptr, *tcol = nullptr, *tcolb = nullptr,
              *tgh = nullptr,
              *tlk = nullptr;
    int32_t ncol = 0, ncolb = 0, ngh = 0, nlk = 0, ntcol = 0, ntcolb = 0, ntgh = 0, ntlk = 0;
    int wide = 0, wide_g = 0;
    // force sampling of the cases whose first level runs at this depth:
    // triangle chunks (li = record, item = chunk) and the record indices
    lbk::Work* fw = nullptr;
    int32_t* frec = nullptr;
    int32_t nfw = 0, nfrec = 0;
};

struct lb_engine {
    int device = 0;
    int32_t storage = 0;
    const Kern* k = &kStrict;
    hipStream_t stream = nullptr;
    std::vector<EngineCase> cases;
    std::string err;
    // batched launch data, rebuilt when the case set changes
    bool dirty = true;
    int lmax = 0;
    std::vector<void*> allocs;
    lbk::LevelArgs* las = nullptr;
    lbk::CaseState* states = nullptr;  // one record per attached case, contiguous
    lbk::EngineForce* frecs = nullptr; // force-sampling records
    std::vector<DepthWork> depth;
    int32_t n_active = 0;
    // options
    bool use_graph = true;
    int sched = LB_SCHED_LINEAR;
    int waves = 5;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // divergence check buffers (grown on demand, kept across rebuilds)
    lbk::FiniteEntry* fin_ents = nullptr;
    int32_t* fin_counts = nullptr;
    size_t fin_cap_ents = 0, fin_cap_cases = 0;
};

namespace {

int fail(lb_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                                 \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return fail(ctx, LB_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// Case buffers come from the device's default memory pool, allocated and
// freed in the order of the case's stream: a finished case's buffers go back
// to the pool without the device-wide wait and unmapping of hipFree (about
// 16 ms for the ~100 buffers of a case), and later cases reuse them.  The
// pool keeps up to 64 GiB of freed memory.
void setup_pool(int dev) {
    static std::atomic<bool> done[64];
    if (dev < 0 || dev >= 64 || done[dev].exchange(true)) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = 64ull << 30;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
}

template <class T>
int dalloc(lb_ctx* c, T** p, size_t n) {
    *p = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (c->dry) {  // footprint planning: count only
        c->bytes += (int64_t)bytes;
        return LB_OK;
    }
    setup_pool(c->device);
    // usable from every stream once the case's stream has passed it
    hipError_t e = hipMallocAsync((void**)p, bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(c, LB_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    c->allocs.push_back((void*)*p);
    c->bytes += (int64_t)bytes;
    return LB_OK;
}

template <class T>
int h2d(lb_ctx* c, T* dst, const T* src, size_t n) {
    if (c->dry) return LB_OK;
    hipError_t e = hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(c, LB_EHIP, "hipMemcpy H2D: %s", hipGetErrorString(e));
    return LB_OK;
}

float half_to_float(uint16_t h) {
    int s = (h >> 15) & 1, e = (h >> 10) & 31, m = h & 1023;
    float v;
    if (e == 0) v = std::ldexp((float)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((float)(m | 1024), e - 25);
    return s ? -v : v;
}

inline size_t fidx(int64_t b, int k, int c) { return ((size_t)b * Q + k) * NC + c; }

lbk::Work mkwork(int li, int item, int ci, const int32_t* row) {
    lbk::Work w;
    w.li = li;
    w.item = item;
    w.ci = ci;
    w.pad = 0;
    w.nbr_row = row;
    return w;
}

// Population storage on the host side of transfers.  FP16/BF16 levels hold
// f - w_k rounded to nearest even; the device decodes w_k + stored value.
float wk_host(int k) {
    const int cx = k % 3 - 1, cy = (k / 3) % 3 - 1, cz = k / 9 - 1;
    const int d2 = cx * cx + cy * cy + cz * cz;
    return d2 == 0 ? 8.0f / 27.0f : d2 == 1 ? 2.0f / 27.0f : d2 == 2 ? 1.0f / 54.0f : 1.0f / 216.0f;
}

size_t pop_bytes(int storage) { return storage == LB_STORAGE_FP32 ? 4 : 2; }

uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
    if (ax >= 0x47800000u) return (uint16_t)(sign | 0x7c00u);  // |f| >= 65536 (and inf)
    if (ax < 0x38800000u) {                                      // half subnormal range: multiples of 2^-24
        const float v = std::nearbyint(std::fabs(f) * 16777216.0f);
        return (uint16_t)(sign | (uint32_t)v);
    }
    uint32_t h = ((((ax >> 23) - 112u) << 10) | ((ax >> 13) & 0x3ffu));
    const uint32_t rem = ax & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;  // carry into the exponent gives inf at 65520
    return (uint16_t)(sign | h);
}

uint16_t float_to_bf16(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}

float bf16_to_float(uint16_t h) {
    const uint32_t b = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

float decode_pop(const uint8_t* base, size_t i, int storage, float w) {
    if (storage == LB_STORAGE_FP32) {
        float f;
        std::memcpy(&f, base + i * 4, 4);
        return f;
    }
    uint16_t u;
    std::memcpy(&u, base + i * 2, 2);
    return w + (storage == LB_STORAGE_FP16 ? half_to_float(u) : bf16_to_float(u));
}

void encode_pop(uint8_t* base, size_t i, float f, int storage, float w) {
    if (storage == LB_STORAGE_FP32) {
        std::memcpy(base + i * 4, &f, 4);
        return;
    }
    const uint16_t u = storage == LB_STORAGE_FP16 ? float_to_half(f - w) : float_to_bf16(f - w);
    std::memcpy(base + i * 2, &u, 2);
}

#define RC(expr)                \
    do {                        \
        int rc_ = (expr);       \
        if (rc_ != LB_OK) return rc_; \
    } while (0)

// Build the Bouzidi link list of one level (src/geometry/bouzidi_kernel.jl:35-89):
// link (cell x, direction k) with q_min < q <= 1 corrects slot (x, opp k).
int build_links(lb_ctx* c, Level& L, const lb_level_desc* d) {
    const int64_t nbc = d->n_boundary_cells;
    if (nbc <= 0) return LB_OK;
    if (!d->bouzidi_cell_block || !d->bouzidi_cell_x || !d->bouzidi_cell_y || !d->bouzidi_cell_z ||
        (!d->bouzidi_q_dense && !d->bouzidi_q_cells))
        return fail(c, LB_EINVAL, "level %d: Bouzidi cell list or q map missing", d->level_id);
    const float qmin = c->cfg.q_min_threshold;
    const int64_t nb = d->n_blocks;
    std::vector<int64_t> dst, s1, s2;
    std::vector<float> ca, cb;
    for (int64_t i = 0; i < nbc; i++) {
        int b = d->bouzidi_cell_block[i], x = d->bouzidi_cell_x[i], y = d->bouzidi_cell_y[i],
            z = d->bouzidi_cell_z[i];
        if (b < 1 || b > nb || x < 1 || x > 8 || y < 1 || y > 8 || z < 1 || z > 8)
            return fail(c, LB_EINVAL, "level %d: Bouzidi cell %lld out of range", d->level_id, (long long)i);
        int cl = (x - 1) + 8 * (y - 1) + 64 * (z - 1);
        for (int k = 0; k < Q; k++) {
            // the dense q map is in the caller's block order
            const size_t bref = L.perm.empty() ? (size_t)(b - 1) : (size_t)L.perm[(size_t)b - 1];
            uint16_t bits = d->bouzidi_q_cells ? d->bouzidi_q_cells[i * Q + k]
                                               : d->bouzidi_q_dense[(size_t)cl + 512 * bref + 512 * (size_t)nb * k];
            float q = half_to_float(bits);
            if (!(q > qmin && q <= 1.0f)) continue;
            int ok = 26 - k;
            int cx = k % 3 - 1, cy = (k / 3) % 3 - 1, cz = k / 9 - 1;
            int64_t sA = (int64_t)fidx(b - 1, k, cl);
            int64_t sB;
            float a, bb;
            if (q < 0.5f) {
                int nx = x - cx, ny = y - cy, nz = z - cz;
                if (nx >= 1 && nx <= 8 && ny >= 1 && ny <= 8 && nz >= 1 && nz <= 8) {
                    sB = (int64_t)fidx(b - 1, k, (nx - 1) + 8 * (ny - 1) + 64 * (nz - 1));
                } else {
                    int ox = nx < 1 ? -1 : (nx > 8 ? 1 : 0), oy = ny < 1 ? -1 : (ny > 8 ? 1 : 0),
                        oz = nz < 1 ? -1 : (nz > 8 ? 1 : 0);
                    int dir = (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1);
                    int nbb = d->neighbor_table[(size_t)(b - 1) + (size_t)nb * dir];
                    if (nbb > 0) {
                        int wx = nx < 1 ? nx + 8 : (nx > 8 ? nx - 8 : nx);
                        int wy = ny < 1 ? ny + 8 : (ny > 8 ? ny - 8 : ny);
                        int wz = nz < 1 ? nz + 8 : (nz > 8 ? nz - 8 : nz);
                        sB = (int64_t)fidx(nbb - 1, k, (wx - 1) + 8 * (wy - 1) + 64 * (wz - 1));
                    } else {
                        sB = sA;
                    }
                }
                float c1 = 2.0f * q;
                a = c1;
                bb = 1.0f - c1;
            } else {
                sB = (int64_t)fidx(b - 1, ok, cl);
                float t2 = 2.0f * q;
                float inv = 1.0f / t2;
                a = inv;
                bb = (2.0f * q - 1.0f) * inv;
            }
            dst.push_back((int64_t)fidx(b - 1, ok, cl));
            s1.push_back(sA);
            s2.push_back(sB);
            ca.push_back(a);
            cb.push_back(bb);
        }
    }
    L.nlinks = (int64_t)dst.size();
    if (L.nlinks == 0) return LB_OK;
    size_t n = (size_t)L.nlinks;
    // Link order does not change results (every link is gathered before any
    // is scattered, and each destination slot has one link).  Sorted by
    // destination, a wave's links share (block, direction) rows, so the
    // scatter's stores and the gather's loads of the same row coalesce.
    {
        std::vector<size_t> perm(n);
        for (size_t i = 0; i < n; i++) perm[i] = i;
        std::sort(perm.begin(), perm.end(), [&](size_t u, size_t v) { return dst[u] < dst[v]; });
        auto apply = [&](auto& vec) {
            auto tmp = vec;
            for (size_t i = 0; i < n; i++) vec[i] = tmp[perm[i]];
        };
        apply(dst);
        apply(s1);
        apply(s2);
        apply(ca);
        apply(cb);
    }
    // 32-bit slot indices where the level's population buffer (ghost blocks
    // and padding included) has fewer than 2^32 slots: 12 instead of 24 of
    // the link's 40 bytes of data (gather and scatter read the link data on
    // every substep).  wide_offsets keeps 64-bit indices too.
    const uint64_t slots = ((uint64_t)nb + (uint64_t)L.ngh) * NC * Q + 512;
    L.lidx32 = c->cfg.wide_offsets == 0 && slots < ((uint64_t)1 << 32);
    auto upload_slots = [&](const std::vector<int64_t>& v, void** out) -> int {
        if (L.lidx32) {
            std::vector<uint32_t> v32(v.begin(), v.end());
            uint32_t* p = nullptr;
            RC(dalloc(c, &p, n));
            *out = p;
            return h2d(c, p, v32.data(), n);
        }
        int64_t* p = nullptr;
        RC(dalloc(c, &p, n));
        *out = p;
        return h2d(c, p, v.data(), n);
    };
    RC(upload_slots(dst, &L.ldst));
    RC(upload_slots(s1, &L.ls1));
    RC(upload_slots(s2, &L.ls2));
    RC(dalloc(c, &L.la, n));
    RC(dalloc(c, &L.lb, n));
    RC(dalloc(c, &L.ltmp, n));
    RC(h2d(c, L.la, ca.data(), n));
    RC(h2d(c, L.lb, cb.data(), n));
    return LB_OK;
}

// perform_timestep_v2! + recursive_step! for one level (timestep.jl:1-159,
// solver_control.jl:6-197): substep j of level li in the coarse step held by
// the case record (t_sub = t0 * 2^li + j), then both child substeps.  The
// kernels resolve buffers, parity, seed and temporal weight from (record, j).
// Kernel timing: record the start event of kernel kind `which` (0 collide,
// 1 thermal) and hand back the end event to record after the launch.
int profile_begin(lb_ctx* c, Level& L, int which, hipEvent_t* end) {
    *end = nullptr;
    if (!c->profile) return LB_OK;
    auto& ev = L.ev[which];
    size_t& used = L.ev_used[which];
    if (used == ev.size()) {
        hipEvent_t x, y;
        HIPCHK(c, hipEventCreate(&x));
        HIPCHK(c, hipEventCreate(&y));
        ev.emplace_back(x, y);
    }
    HIPCHK(c, hipEventRecord(ev[used].first, c->stream));
    *end = ev[used].second;
    used++;
    return LB_OK;
}

// Occupancy of the interior stream-collide kernel per population storage:
// 6 waves per SIMD for 16-bit storage, 5 for FP32.  Measured on the 52.5M-cell
// cube case through the engine: FP16 11.56 -> 11.17 ms per coarse step at 6;
// FP32 15.46 ms at either.
int default_waves(int storage) { return storage == LB_STORAGE_FP32 ? 5 : 6; }

int step_level(lb_ctx* c, int li, int64_t j) {
    if (li >= (int)c->levels.size()) return LB_OK;
    Level& L = c->levels[(size_t)li];
    const bool ddf = c->cfg.ddf_enabled && L.g[0] != nullptr;
    const int storage = c->cfg.storage;
    const lbk::LevelArgs* las = c->las_dev;
    const lbk::CaseState* st = c->st_dev;
    if (L.nb > 0) {
        // one ghost pass per parent step fills the inputs of both child substeps
        const bool fill = li > 0 && L.n_gchunks > 0 && (j & 1) == 0;
        if (fill) HIPCHK(c, c->k->ghost_fill(las, st, L.wgh, L.n_gchunks, j, storage, 0, c->stream));
        hipEvent_t e1 = nullptr;
        RC(profile_begin(c, L, 0, &e1));
        HIPCHK(c, c->k->collide(las, st, L.wcol, L.n_int, j, storage, L.wide ? 1 : 0, 0, LB_SCHED_LINEAR,
                                default_waves(storage), c->stream));
        HIPCHK(c, c->k->collide(las, st, L.wcol + L.n_int, L.nb - L.n_int, j, storage, L.wide ? 1 : 0, 1,
                                LB_SCHED_LINEAR, 5, c->stream));
        if (e1) HIPCHK(c, hipEventRecord(e1, c->stream));
        if (L.n_lchunks > 0) HIPCHK(c, c->k->bouzidi(las, st, L.wlk, L.n_lchunks, j, storage, 0, c->stream));
        if (ddf) {
            if (fill) HIPCHK(c, c->k->ghost_fill(las, st, L.wgh, L.n_gchunks, j, storage, 1, c->stream));
            hipEvent_t e2 = nullptr;
            RC(profile_begin(c, L, 1, &e2));
            HIPCHK(c, c->k->thermal(las, st, L.wcol, L.n_int, j, storage, L.wide_g ? 1 : 0, 0, LB_SCHED_LINEAR,
                                    c->stream));
            HIPCHK(c, c->k->thermal(las, st, L.wcol + L.n_int, L.nb - L.n_int, j, storage, L.wide_g ? 1 : 0, 1,
                                    LB_SCHED_LINEAR, c->stream));
            if (e2) HIPCHK(c, hipEventRecord(e2, c->stream));
            if (L.n_lchunks > 0) HIPCHK(c, c->k->bouzidi(las, st, L.wlk, L.n_lchunks, j, storage, 1, c->stream));
        }
    }
    RC(step_level(c, li + 1, 2 * j));
    RC(step_level(c, li + 1, 2 * j + 1));
    return LB_OK;
}

// The case record for coarse step t with the given inlet velocity and the
// physics parameters of one lb_step_batch call.
lbk::CaseState case_record(int64_t t, float u, const lb_step_params* p, const lb_config& cfg) {
    lbk::CaseState s;
    std::memset(&s, 0, sizeof s);
    s.t0 = t;
    s.n_total = INT64_MAX;
    s.u = u;
    s.active = 1;
    s.t_inc = 1;
    s.batch = 1;
    s.u_tab = nullptr;
    s.c_wale = p->c_wale;
    s.nu_sgs_bg = p->nu_sgs_background;
    s.inlet_turb = p->inlet_turbulence;
    s.omega_bulk = p->omega_bulk;
    s.omega3 = p->omega_3;
    s.omega4 = p->omega_4;
    s.lambda = p->lambda_param;
    s.adaptive = p->adaptive_omega4;
    s.limiter = p->limiter;
    s.comp = p->compressible;
    s.wall_model = p->wall_model ? 1 : 0;
    s.sponge_blend = p->sponge_blend_distributions;
    s.T_inlet = p->T_inlet;
    s.T_wall = p->T_wall;
    s.adiabatic = p->wall_bc_adiabatic;
    s.temporal = cfg.temporal_interpolation ? 1 : 0;
    return s;
}

// Output half of every level after coarse step t: the last substep of level
// li is t * 2^li + 2^li - 1.
void set_current(lb_ctx* c, int64_t t) {
    for (size_t li = 0; li < c->levels.size(); li++) {
        const int64_t last = (t << li) + ((int64_t)1 << li) - 1;
        c->levels[li].cur = 1 - (int)(last & 1);
    }
}

// Work list of items 0 .. n - 1 of one level (ghost blocks).
int item_list(lb_ctx* c, int li, int32_t n, lbk::Work** out, int32_t* n_items) {
    *n_items = n;
    if (n == 0) return LB_OK;
    std::vector<lbk::Work> w((size_t)n);
    for (int32_t i = 0; i < n; i++) w[(size_t)i] = mkwork(li, i, 0, nullptr);
    RC(dalloc(c, out, w.size()));
    return h2d(c, *out, w.data(), w.size());
}

// Work list of 256-item chunks of one level (ghost entries, Bouzidi links).
int chunk_list(lb_ctx* c, int li, int64_t n, lbk::Work** out, int32_t* n_chunks) {
    *n_chunks = (int32_t)((n + 255) / 256);
    if (*n_chunks == 0) return LB_OK;
    std::vector<lbk::Work> w((size_t)*n_chunks);
    for (int32_t i = 0; i < *n_chunks; i++) w[(size_t)i] = mkwork(li, i * 256, 0, nullptr);
    RC(dalloc(c, out, w.size()));
    return h2d(c, *out, w.data(), w.size());
}

// LevelArgs of level li (host copy + device record); the parent's pairs come
// from level li - 1, which is uploaded first.
int upload_level_args(lb_ctx* c, int li) {
    const Level& L = c->levels[(size_t)li];
    lbk::LevelArgs& A = c->las_host[(size_t)li];
    std::memset((void*)&A, 0, sizeof A);
    for (int i = 0; i < 2; i++) {
        A.f[i] = L.f[i];
        A.rho[i] = L.rho[i];
        A.vel[i] = L.vel[i];
        A.g[i] = L.g[i];
        A.T[i] = L.T[i];
    }
    A.obst = L.obst;
    A.sponge = L.sponge;
    A.wall_dist = L.wall_dist;
    A.bflags = L.bflags;
    A.nbr = L.nbr;
    A.borg = L.borg;
    A.nb = L.nb;
    const int scale = 1 << (L.id - 1);
    A.nxg = c->cfg.domain_nx * scale;
    A.nyg = c->cfg.domain_ny * scale;
    A.nzg = c->cfg.domain_nz * scale;
    A.lshift = L.id - 1;
    A.leaf = li + 1 == (int)c->levels.size() ? 1 : 0;
    A.tau = L.tau;
    A.tau_g = L.tau_g;
    if (li > 0) {
        const Level& P = c->levels[(size_t)li - 1];
        for (int i = 0; i < 2; i++) {
            A.pf[i] = P.f[i];
            A.prho[i] = P.rho[i];
            A.pvel[i] = P.vel[i];
            A.pg[i] = P.g[i];
            A.pT[i] = P.T[i];
        }
        A.pgrid = P.grid;
        A.pdx = P.pd[0];
        A.pdy = P.pd[1];
        A.pdz = P.pd[2];
        A.ptau = P.tau;
        A.ptau_g = P.tau_g;
    }
    A.gent = L.gent;
    A.gpar = L.gpar;
    A.gstart = L.gstart;
    A.gregion = L.gregion;
    A.gbmask = L.gbmask;
    A.ngent = L.ngent;
    A.links.n = L.nlinks;
    A.links.dst = L.ldst;
    A.links.s1 = L.ls1;
    A.links.s2 = L.ls2;
    A.links.a = L.la;
    A.links.b = L.lb;
    A.links.tmp = L.ltmp;
    A.links.idx32 = L.lidx32 ? 1 : 0;
    return h2d(c, c->las_dev + li, &A, 1);
}

int efail(lb_engine* e, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e) e->err = buf;
    return code;
}

#define EHIPCHK(eng, expr)                                                                                \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return efail(eng, LB_EHIP, "%s: %s", #expr, hipGetErrorString(e_));        \
    } while (0)

void engine_free_lists(lb_engine* e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    for (void* p : e->allocs) (void)hipFree(p);
    e->allocs.clear();
    e->las = nullptr;
    e->states = nullptr;
    e->frecs = nullptr;
    e->depth.clear();
}

template <class T>
int eupload(lb_engine* e, T** dst, const std::vector<T>& v) {
    *dst = nullptr;
    if (v.empty()) return LB_OK;
    EHIPCHK(e, hipMalloc((void**)dst, v.size() * sizeof(T)));
    e->allocs.push_back((void*)*dst);
    EHIPCHK(e, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return LB_OK;
}

lbk::ForceArgs force_args(lb_ctx* c, const lb_force_params* fp, float side);

// The device record of an attached case at its current position.
lbk::CaseState engine_record(const lb_engine* e, const EngineCase& ec) {
    lbk::CaseState r = case_record(ec.t0, 0.0f, &ec.p, ec.c->cfg);
    r.n_total = ec.n_total;
    r.active = ec.t0 <= ec.n_total ? 1 : 0;
    r.t_inc = ec.t_inc;
    r.batch = ec.batch;
    r.u_tab = ec.u_tab;
    (void)e;
    return r;
}

// Rebuild the concatenated LevelArgs, the per-depth work lists and the case
// records after the case set changed.  Cases are aligned on their finest
// level: level l of a case with L levels runs at depth l + (Lmax - L).
int engine_rebuild(lb_engine* e) {
    EHIPCHK(e, hipStreamSynchronize(e->stream));
    engine_free_lists(e);
    e->lmax = 0;
    for (const EngineCase& ec : e->cases) e->lmax = std::max(e->lmax, (int)ec.c->levels.size());
    std::vector<lbk::LevelArgs> las;
    std::vector<lbk::CaseState> sts;
    std::vector<int> base;
    for (EngineCase& ec : e->cases) {
        base.push_back((int)las.size());
        for (const lbk::LevelArgs& a : ec.c->las_host) las.push_back(a);
        ec.t_inc = 1 << (e->lmax - (int)ec.c->levels.size());
        sts.push_back(engine_record(e, ec));
    }
    RC(eupload(e, &e->las, las));
    RC(eupload(e, &e->states, sts));
    // force sampling: one record per sampling case, its triangle chunks
    // listed at the depth of the case's first level
    std::vector<lbk::EngineForce> frecs;
    std::vector<std::vector<lbk::Work>> fw((size_t)e->lmax);
    std::vector<std::vector<int32_t>> frec((size_t)e->lmax);
    for (size_t i = 0; i < e->cases.size(); i++) {
        EngineCase& ec = e->cases[i];
        if (!ec.sample) continue;
        lb_ctx* c = ec.c;
        const Level& fin = c->levels.back();
        lbk::EngineForce r;
        std::memset((void*)&r, 0, sizeof r);
        r.a = force_args(c, &ec.fp, 1.0f);
        r.rho2[0] = fin.rho[0];
        r.rho2[1] = fin.rho[1];
        r.diag = ec.diag;
        r.ramp = ec.ramp;
        r.ci = (int32_t)i;
        r.leaf_shift = (int32_t)c->levels.size() - 1;
        r.n_groups = (int32_t)((c->n_tri + 255) / 256);
        r.acc = ec.acc;
        const int d = e->lmax - (int)c->levels.size() + 1;
        const int32_t ri = (int32_t)frecs.size();
        for (int32_t g = 0; g < r.n_groups; g++) fw[(size_t)d - 1].push_back(mkwork(ri, g, (int32_t)i, nullptr));
        frec[(size_t)d - 1].push_back(ri);
        frecs.push_back(r);
    }
    RC(eupload(e, &e->frecs, frecs));
    e->depth.assign((size_t)e->lmax, DepthWork());
    for (int d = 1; d <= e->lmax; d++) {
        DepthWork& W = e->depth[(size_t)d - 1];
        RC(eupload(e, &W.fw, fw[(size_t)d - 1]));
        RC(eupload(e, &W.frec, frec[(size_t)d - 1]));
        W.nfw = (int32_t)fw[(size_t)d - 1].size();
        W.nfrec = (int32_t)frec[(size_t)d - 1].size();
    }
    for (int d = 1; d <= e->lmax; d++) {
        std::vector<lbk::Work> col, colb, gh, lk, tcol, tcolb, tgh, tlk;
        DepthWork& W = e->depth[(size_t)d - 1];
        for (size_t i = 0; i < e->cases.size(); i++) {
            const lb_ctx* c = e->cases[i].c;
            const int L = (int)c->levels.size();
            const int l = d - (e->lmax - L);
            if (l < 1 || l > L) continue;
            const Level& lev = c->levels[(size_t)l - 1];
            const int li = base[i] + l - 1, ci = (int)i;
            if (lev.nb == 0) continue;
            const bool ddf = c->cfg.ddf_enabled && lev.g[0] != nullptr;
            for (size_t q = 0; q < lev.order_host.size(); q++) {
                const int32_t b = lev.order_host[q];
                const lbk::Work w = mkwork(li, b, ci, lev.nbr + (size_t)b * Q);
                const bool interior = (int32_t)q < lev.n_int;
                (interior ? col : colb).push_back(w);
                if (ddf) (interior ? tcol : tcolb).push_back(w);
            }
            // ghost fill items: chunks of 256 entries (FP32), ghost blocks (16-bit)
            const bool staged = e->storage != LB_STORAGE_FP32;
            const int64_t n_items = staged ? (lev.ngent > 0 ? lev.ngh : 0) : (lev.ngent + 255) / 256;
            if (l > 1)
                for (int64_t q = 0; q < n_items; q++) {
                    const int item = (int)(staged ? q : 256 * q);
                    gh.push_back(mkwork(li, item, ci, nullptr));
                    if (ddf) tgh.push_back(mkwork(li, item, ci, nullptr));
                }
            for (int64_t q = 0; q < lev.nlinks; q += 256) {
                lk.push_back(mkwork(li, (int)q, ci, nullptr));
                if (ddf) tlk.push_back(mkwork(li, (int)q, ci, nullptr));
            }
            if (lev.wide) W.wide = 1;
            if (lev.wide_g) W.wide_g = 1;
        }
        RC(eupload(e, &W.col, col));
        RC(eupload(e, &W.colb, colb));
        RC(eupload(e, &W.gh, gh));
        RC(eupload(e, &W.lk, lk));
        RC(eupload(e, &W.tcol, tcol));
        RC(eupload(e, &W.tcolb, tcolb));
        RC(eupload(e, &W.tgh, tgh));
        RC(eupload(e, &W.tlk, tlk));
        W.ncol = (int32_t)col.size();
        W.ncolb = (int32_t)colb.size();
        W.ngh = (int32_t)gh.size();
        W.nlk = (int32_t)lk.size();
        W.ntcol = (int32_t)tcol.size();
        W.ntcolb = (int32_t)tcolb.size();
        W.ntgh = (int32_t)tgh.size();
        W.ntlk = (int32_t)tlk.size();
    }
    e->n_active = (int32_t)e->cases.size();
    e->dirty = false;
    return LB_OK;
}

// One depth-d operation (substep j) and its subtree, over all cases.
hipError_t engine_op(lb_engine* e, int d, int64_t j) {
    const DepthWork& W = e->depth[(size_t)d - 1];
    const Kern* k = e->k;
    hipError_t r = hipSuccess;
#define ETRY(x)                       \
    do {                              \
        r = (x);                      \
        if (r != hipSuccess) return r; \
    } while (0)
    const lbk::CaseState* st = e->states;
    // one ghost pass per parent step fills the inputs of both child substeps
    const bool fill = (j & 1) == 0;
    if (W.ngh && fill) ETRY(k->ghost_fill(e->las, st, W.gh, W.ngh, j, e->storage, 0, e->stream));
    if (W.ncol)
        ETRY(k->collide(e->las, st, W.col, W.ncol, j, e->storage, W.wide, 0, e->sched, e->waves, e->stream));
    if (W.ncolb)
        ETRY(k->collide(e->las, st, W.colb, W.ncolb, j, e->storage, W.wide, 1, e->sched, e->waves,
                        e->stream));
    if (W.nlk) ETRY(k->bouzidi(e->las, st, W.lk, W.nlk, j, e->storage, 0, e->stream));
    if (W.ntcol || W.ntcolb) {
        // the thermal kernel has no persistent form
        const int tsched = e->sched == LB_SCHED_PERSISTENT ? LB_SCHED_LINEAR : e->sched;
        if (W.ntgh && fill) ETRY(k->ghost_fill(e->las, st, W.tgh, W.ntgh, j, e->storage, 1, e->stream));
        if (W.ntcol) ETRY(k->thermal(e->las, st, W.tcol, W.ntcol, j, e->storage, W.wide_g, 0, tsched, e->stream));
        if (W.ntcolb)
            ETRY(k->thermal(e->las, st, W.tcolb, W.ntcolb, j, e->storage, W.wide_g, 1, tsched, e->stream));
        if (W.ntlk) ETRY(k->bouzidi(e->las, st, W.tlk, W.ntlk, j, e->storage, 1, e->stream));
    }
    if (d < e->lmax) {
        ETRY(engine_op(e, d + 1, 2 * j));
        ETRY(engine_op(e, d + 1, 2 * j + 1));
    }
    // coarse step t0 + j of the cases whose first level runs at depth d is complete
    if (W.nfw) ETRY(k->engine_forces(e->frecs, st, W.fw, W.nfw, W.frec, W.nfrec, j, e->stream));
#undef ETRY
    return r;
}

hipError_t engine_issue_step(lb_engine* e) {
    hipError_t r = engine_op(e, 1, 0);
    if (r != hipSuccess) return r;
    return e->k->advance(e->states, (int32_t)e->cases.size(), e->stream);
}

// Divergence check entry of one level: its density and velocity in buffer
// `cur` (the output of the last completed coarse step).
constexpr int32_t kFiniteGroups = 64;  // work-groups per (case, level) entry
lbk::FiniteEntry finite_entry(const Level& L, int cur, int32_t slot) {
    lbk::FiniteEntry e;
    e.rho = L.rho[cur];
    e.vel = L.vel[cur];
    e.n = (int64_t)L.nb * NC;
    e.slot = slot;
    e.pad = 0;
    return e;
}

int check_ctx(lb_ctx* c) {
    if (!c) return LB_EINVAL;
    if (c->engine) return fail(c, LB_ESTATE, "case is attached to an engine");
    for (size_t i = 0; i < c->levels.size(); i++)
        if (!c->levels[i].uploaded) return fail(c, LB_ESTATE, "level %zu not uploaded", i + 1);
    return LB_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// lb_stats_noalloc: the statistics of lb_stats with their partial rows and
// the reduced result in a fixed device array instead of memory allocated on
// first use, for callers that cannot allocate.  One copy of the array per
// device; its address on each device is taken once, at context creation on
// that device (hipGetSymbolAddress), and only that address is handed to
// kernels and copies.
namespace {
constexpr int kStatsRows = 1024;  // the statistics kernel runs at most this many work-groups
constexpr size_t kScratchDoubles = (size_t)kStatsRows * 8 + 8;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// Address of `scratch` on `device` (the current device), taken once.
void scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) g_scratch_ptr[device] = (double*)p;
    else {
        g_scratch_err[device] = e;
        (void)hipGetLastError();
    }
}

double* scratch_on(int device, hipError_t* err) {
    if (device < 0 || device >= kMaxDevices) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    *err = g_scratch_err[device];
    return g_scratch_ptr[device];
}

// The partial rows reduced in row order, with the operations lb_stats
// applies on the host (so the results are the same to the bit).
__global__ void stats_reduce_kernel(const double* __restrict__ part, int rows, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s0 = 0, s1 = 0, s4 = 0, s5 = 0, s6 = 0, mn = 1e300, mx = -1e300, vmax = 0;
    for (int g = 0; g < rows; g++) {
        const double* q = part + (size_t)g * 8;
        s0 += q[0];
        s1 += q[1];
        mn = q[2] < mn ? q[2] : mn;
        mx = mx < q[3] ? q[3] : mx;
        s4 += q[4];
        s5 += q[5];
        s6 += q[6];
        vmax = vmax < q[7] ? q[7] : vmax;
    }
    out[0] = s0; out[1] = s1; out[2] = mn; out[3] = mx;
    out[4] = s4; out[5] = s5; out[6] = s6; out[7] = vmax;
}
}  // namespace

namespace lbsync {
CaptureLock& capture_mutex() {
    static CaptureLock m;
    return m;
}
}  // namespace lbsync

extern "C" {

const char* lb_last_error(const lb_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

const char* lb_version(void) { return "ludwig_hip 0.1 gfx950"; }

int lb_create(lb_ctx** out, int device, const lb_config* cfg) {
    lbsync::SetupGuard guard;
    if (!out || !cfg) return LB_EINVAL;
    *out = nullptr;
    if (cfg->n_levels < 1 || cfg->n_levels > 16 || cfg->domain_nx < 8 || cfg->domain_ny < 8 || cfg->domain_nz < 8)
        return LB_EINVAL;
    if (cfg->storage != LB_STORAGE_FP32 && cfg->storage != LB_STORAGE_FP16 && cfg->storage != LB_STORAGE_BF16)
        return LB_EINVAL;
    lb_ctx* c = new lb_ctx();
    c->device = device;
    c->cfg = *cfg;
    c->k = cfg->strict_fp ? &kStrict : &kFast;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete c;
        return LB_EHIP;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return LB_EHIP;
    }
    scratch_setup(device);
    c->levels.resize((size_t)cfg->n_levels);
    c->las_host.resize((size_t)cfg->n_levels);
    std::memset((void*)c->las_host.data(), 0, c->las_host.size() * sizeof(lbk::LevelArgs));
    if (dalloc(c, &c->st_dev, 1) != LB_OK || dalloc(c, &c->las_dev, (size_t)cfg->n_levels) != LB_OK ||
        dalloc(c, &c->fin_ents, (size_t)cfg->n_levels) != LB_OK || dalloc(c, &c->fin_count, 1) != LB_OK) {
        lb_destroy(c);
        return LB_ENOMEM;
    }
    *out = c;
    return LB_OK;
}

void lb_destroy(lb_ctx* c) {
    lbsync::SetupGuard guard;
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& L : c->levels)
        for (auto& evk : L.ev)
            for (auto& e : evk) {
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
    if (c->stream) {
        for (void* p : c->allocs) (void)hipFreeAsync(p, c->stream);
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

int lb_level_upload(lb_ctx* c, const lb_level_desc* d) {
    lbsync::SetupGuard guard;
    if (!c || !d) return LB_EINVAL;
    if (d->level_id < 1 || d->level_id > (int)c->levels.size())
        return fail(c, LB_EINVAL, "level_id %d out of range", d->level_id);
    Level& L = c->levels[(size_t)d->level_id - 1];
    if (L.uploaded) return fail(c, LB_ESTATE, "level %d already uploaded", d->level_id);
    if (d->level_id > 1 && !c->levels[(size_t)d->level_id - 2].uploaded)
        return fail(c, LB_ESTATE, "levels must be uploaded in order");
    const int32_t nb = d->n_blocks;
    if (nb < 0 || (nb > 0 && (!d->map_x || !d->map_y || !d->map_z || !d->neighbor_table || !d->obstacle ||
                             !d->sponge || !d->block_pointer)))
        return fail(c, LB_EINVAL, "level %d: missing arrays", d->level_id);
    if (d->ptr_dim_x < 0 || d->ptr_dim_y < 0 || d->ptr_dim_z < 0)
        return fail(c, LB_EINVAL, "level %d: bad block_pointer dims", d->level_id);
    if (!c->dry) HIPCHK(c, hipSetDevice(c->device));
    // Device block order.  The caller's blocks come in the reference's
    // lexicographic (bx, by, bz) order, z fastest.  On the device they are
    // stored, and processed, in 4-layer z-slabs: inside a slab x-rows of
    // 4-block z-columns, x fastest, then y.  The stream-collide pull reads a
    // 4-byte value per cell of each x face of a neighbour block out of
    // 128-byte lines that hold 4 rows of 8 cells (a y face reads one 32-byte
    // run per line, a z face whole lines), so x neighbours should be read
    // right after each other and y and z neighbours soon after, while their
    // lines are still in L2 or the last-level cache; and processing in
    // storage order streams the level through memory.  Measured on the
    // 52.5M-cell cube case (MI355X, FP32), finest stream-collide: the
    // caller's order with the slab work order dealt to XCDs in runs 1680 us;
    // slab storage in order 1545 us (1612 us in XCD runs); x fastest in order
    // 1904 us (the z neighbours fall out of the caches); the caller's order in
    // order 1634 us.  A 2-D Morton order of the columns in place of the
    // x-rows measured the same.  Every per-block input is permuted here and
    // every per-block output is permuted back (lb_download_level,
    // lb_upload_level); per-cell results do not depend on the order.
    std::vector<int32_t> perm((size_t)std::max(nb, 0)), inv((size_t)std::max(nb, 0));
    for (int32_t b = 0; b < nb; b++) perm[(size_t)b] = b;
    auto slab_key = [d](int32_t b) -> uint64_t {
        const uint64_t x = (uint64_t)(d->map_x[b] - 1), y = (uint64_t)(d->map_y[b] - 1), z = (uint64_t)(d->map_z[b] - 1);
        return ((z >> 2) << 48) | (y << 32) | (x << 8) | (z & 3u);
    };
    std::stable_sort(perm.begin(), perm.end(), [&slab_key](int32_t a, int32_t b) { return slab_key(a) < slab_key(b); });
    for (int32_t i = 0; i < nb; i++) inv[(size_t)perm[(size_t)i]] = i;
    std::vector<int32_t> pmx((size_t)std::max(nb, 0)), pmy(pmx.size()), pmz(pmx.size()), pnt((size_t)std::max(nb, 0) * Q);
    std::vector<uint8_t> pobst((size_t)std::max(nb, 0) * NC);
    std::vector<float> psponge(pobst.size()), pwall(d->wall_dist ? pobst.size() : 0);
    std::vector<int32_t> pbp((size_t)d->ptr_dim_x * d->ptr_dim_y * d->ptr_dim_z);
    std::vector<int32_t> pcell((size_t)std::max<int64_t>(d->n_boundary_cells, 0));
    lb_level_desc dp = *d;
    if (nb > 0) {
        for (int32_t i = 0; i < nb; i++) {
            const int32_t b = perm[(size_t)i];
            pmx[(size_t)i] = d->map_x[b];
            pmy[(size_t)i] = d->map_y[b];
            pmz[(size_t)i] = d->map_z[b];
            for (int k = 0; k < Q; k++) {
                const int v = d->neighbor_table[(size_t)b + (size_t)nb * k];
                pnt[(size_t)i + (size_t)nb * k] = (v >= 1 && v <= nb) ? inv[(size_t)v - 1] + 1 : v;
            }
            std::memcpy(&pobst[(size_t)i * NC], d->obstacle + (size_t)b * NC, NC);
            std::memcpy(&psponge[(size_t)i * NC], d->sponge + (size_t)b * NC, NC * sizeof(float));
            if (d->wall_dist) std::memcpy(&pwall[(size_t)i * NC], d->wall_dist + (size_t)b * NC, NC * sizeof(float));
        }
        for (size_t i = 0; i < pbp.size(); i++) {
            const int v = d->block_pointer[i];
            pbp[i] = (v >= 1 && v <= nb) ? inv[(size_t)v - 1] + 1 : v;
        }
        for (size_t i = 0; i < pcell.size() && d->bouzidi_cell_block; i++) {
            const int v = d->bouzidi_cell_block[i];
            pcell[i] = (v >= 1 && v <= nb) ? inv[(size_t)v - 1] + 1 : v;
        }
        dp.map_x = pmx.data();
        dp.map_y = pmy.data();
        dp.map_z = pmz.data();
        dp.neighbor_table = pnt.data();
        dp.block_pointer = pbp.empty() ? d->block_pointer : pbp.data();
        dp.obstacle = pobst.data();
        dp.sponge = psponge.data();
        dp.wall_dist = d->wall_dist ? pwall.data() : nullptr;
        if (d->bouzidi_cell_block) dp.bouzidi_cell_block = pcell.data();
    }
    L.perm = perm;
    d = &dp;
    L.id = d->level_id;
    L.nb = nb;
    L.tau = d->tau;
    L.tau_g = d->tau_g;
    L.dx = d->dx;
    // Host tables: 0-based neighbour rows, block origins, and ghost blocks for
    // in-domain neighbour positions the level does not hold (levels > 1).
    std::vector<int32_t> nbr((size_t)nb * Q), borg((size_t)nb * 3), gent;
    const int scale = 1 << (L.id - 1);
    const int64_t dbx = (int64_t)c->cfg.domain_nx * scale / 8, dby = (int64_t)c->cfg.domain_ny * scale / 8,
                  dbz = (int64_t)c->cfg.domain_nz * scale / 8;
    auto key = [](int64_t x, int64_t y, int64_t z) { return (uint64_t)(x + 1) | ((uint64_t)(y + 1) << 21) | ((uint64_t)(z + 1) << 42); };
    std::unordered_map<uint64_t, int32_t> real, ghost;
    std::vector<int64_t> gpos;
    for (int32_t b = 0; b < nb; b++) real[key(d->map_x[b], d->map_y[b], d->map_z[b])] = b;
    for (int32_t b = 0; b < nb; b++) {
        for (int k = 0; k < Q; k++) {
            int v = d->neighbor_table[(size_t)b + (size_t)nb * k];
            if (v < 0 || v > nb) return fail(c, LB_EINVAL, "level %d: neighbour table entry out of range", L.id);
            int32_t id = v - 1;
            if (id < 0 && L.id > 1) {
                int64_t px = d->map_x[b] + k % 3 - 1, py = d->map_y[b] + (k / 3) % 3 - 1, pz = d->map_z[b] + k / 9 - 1;
                if (px >= 1 && px <= dbx && py >= 1 && py <= dby && pz >= 1 && pz <= dbz) {
                    auto it = ghost.find(key(px, py, pz));
                    if (it == ghost.end()) {
                        it = ghost.emplace(key(px, py, pz), (int32_t)gpos.size() / 3).first;
                        gpos.push_back(px); gpos.push_back(py); gpos.push_back(pz);
                    }
                    id = nb + it->second;
                }
            }
            nbr[(size_t)b * Q + k] = id;
        }
        if (nbr[(size_t)b * Q + 13] != b)
            return fail(c, LB_EINVAL, "level %d: neighbour table centre entry of block %d is not itself", L.id, b + 1);
        borg[(size_t)b * 3 + 0] = (d->map_x[b] - 1) * 8;
        borg[(size_t)b * 3 + 1] = (d->map_y[b] - 1) * 8;
        borg[(size_t)b * 3 + 2] = (d->map_z[b] - 1) * 8;
    }
    const int32_t ngh = (int32_t)(gpos.size() / 3);
    for (int32_t g = 0; g < ngh; g++) {
        const int64_t px = gpos[(size_t)g * 3], py = gpos[(size_t)g * 3 + 1], pz = gpos[(size_t)g * 3 + 2];
        borg.push_back((int32_t)((px - 1) * 8));
        borg.push_back((int32_t)((py - 1) * 8));
        borg.push_back((int32_t)((pz - 1) * 8));
        // ghost cell s serves direction k when s + c_k lies in a block of the level
        uint32_t has_real = 0;  // bit (ox+1) + 3(oy+1) + 9(oz+1): that neighbour block is real
        for (int n = 0; n < Q; n++)
            if (n != 13 && real.count(key(px + n % 3 - 1, py + (n / 3) % 3 - 1, pz + n / 9 - 1)))
                has_real |= 1u << n;
        for (int cl = 0; cl < NC && has_real != 0u; cl++) {
            const int lx = cl & 7, ly = (cl >> 3) & 7, lz = cl >> 6;
            uint32_t mask = 0;
            for (int k = 0; k < Q; k++) {
                const int nx = lx + k % 3 - 1, ny = ly + (k / 3) % 3 - 1, nz = lz + k / 9 - 1;
                const int ox = nx < 0 ? -1 : (nx > 7 ? 1 : 0), oy = ny < 0 ? -1 : (ny > 7 ? 1 : 0),
                          oz = nz < 0 ? -1 : (nz > 7 ? 1 : 0);
                if (ox == 0 && oy == 0 && oz == 0) continue;
                if (has_real & (1u << ((ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)))) mask |= 1u << k;
            }
            if (mask) {
                gent.push_back((int32_t)(((nb + g) << 9) | cl));
                gent.push_back((int32_t)mask);
            }
        }
    }
    L.ngh = ngh;
    L.ngent = (int64_t)gent.size() / 2;
    // The ghost fill's parent tables.  FP32 storage (ghost_kernel): per
    // entry.  16-bit storage (ghost_kernel_staged): per ghost block.
    const bool staged = c->cfg.storage != LB_STORAGE_FP32;
    // The 8 parent cells of every ghost entry, as the interpolation of
    // physics_interpolation.jl:4-127 addresses them: the fine cell centre at
    // (g - 0.5) / 2 in parent cells, its lower corner clamped to the domain,
    // each corner's parent block looked up in the parent grid (-1 where the
    // parent level holds no block).  Static, so the kernel reads them instead
    // of chasing the block origin and the parent grid per entry.
    std::vector<int32_t> gpar(staged ? 0 : (size_t)L.ngent * 8, -1);
    if (!staged && L.ngent > 0) {
        const Level& P = c->levels[(size_t)L.id - 2];
        const int pdx = P.pd[0], pdy = P.pd[1], pdz = P.pd[2];
        for (int64_t e = 0; e < L.ngent; e++) {
            const int gb = gent[(size_t)e * 2] >> 9, cl = gent[(size_t)e * 2] & 511;
            const int fg[3] = {borg[(size_t)gb * 3 + 0] + (cl & 7) + 1, borg[(size_t)gb * 3 + 1] + ((cl >> 3) & 7) + 1,
                               borg[(size_t)gb * 3 + 2] + (cl >> 6) + 1};
            int lo[3], hi[3];
            for (int ax = 0; ax < 3; ax++) {
                const float pc = ((float)fg[ax] - 0.5f) * 0.5f;
                lo[ax] = (int)std::floor(pc);
                hi[ax] = lo[ax] + 1;
                lo[ax] = std::max(1, lo[ax]);
            }
            for (int q = 0; q < 8; q++) {
                const int pgx = (q & 1) ? hi[0] : lo[0], pgy = (q & 2) ? hi[1] : lo[1], pgz = (q & 4) ? hi[2] : lo[2];
                const int pbx = (pgx - 1) / 8 + 1, pby = (pgy - 1) / 8 + 1, pbz = (pgz - 1) / 8 + 1;
                if (pbx < 1 || pbx > pdx || pby < 1 || pby > pdy || pbz < 1 || pbz > pdz) continue;
                const int32_t pb = P.grid_host[(size_t)(pbx - 1) + (size_t)pdx * ((size_t)(pby - 1) + (size_t)pdy * (pbz - 1))];
                if (pb < 0) continue;
                if (pb >= (1 << 22)) return fail(c, LB_EINVAL, "level %d: parent block index %d too large", L.id, pb);
                const int cc = ((pgx - 1) % 8) + 8 * ((pgy - 1) % 8) + 64 * ((pgz - 1) % 8);
                gpar[(size_t)e * 8 + q] = (pb << 9) | cc;
            }
        }
    }
    // Per ghost block, for the staged interpolation of physics_interpolation.jl:
    // 4-127: its entries' range, the union of their served directions, and
    // its parent region.  A fine cell at local index l (0..7) of a block with
    // origin B (a multiple of 8) has its centre at (g - 0.5) / 2 in parent
    // cells, g = B + l + 1; the interpolation's lower corner is
    // max(1, B/2 + l/2) and its upper corner B/2 + l/2 + 1, so every corner
    // of the block lies in the 5^3 parent cells B/2 + r, r = 0..4 per axis.
    // gregion holds (parent block << 9 | cell) for those 125 cells, looked up
    // in the parent grid, or -1 where the parent level holds no block (or the
    // cell is outside it).
    const int32_t ngs = staged && L.ngent > 0 ? ngh : 0;
    std::vector<int32_t> gstart((size_t)ngs + (staged ? 1 : 0), 0), gregion((size_t)ngs * 125, -1);
    std::vector<uint32_t> gbmask((size_t)ngs, 0u);
    for (int64_t e = 0; e < (ngs > 0 ? L.ngent : 0); e++) {
        const int g = (gent[(size_t)e * 2] >> 9) - nb;
        gstart[(size_t)g + 1]++;
        gbmask[(size_t)g] |= (uint32_t)gent[(size_t)e * 2 + 1];
    }
    for (int32_t g = 0; g < ngs; g++) gstart[(size_t)g + 1] += gstart[(size_t)g];
    if (ngs > 0) {
        const Level& P = c->levels[(size_t)L.id - 2];
        const int pdx = P.pd[0], pdy = P.pd[1], pdz = P.pd[2];
        for (int32_t g = 0; g < ngs; g++) {
            int base[3];
            for (int ax = 0; ax < 3; ax++) base[ax] = borg[(size_t)(nb + g) * 3 + ax] / 2;
            for (int r = 0; r < 125; r++) {
                const int pgx = base[0] + r % 5, pgy = base[1] + (r / 5) % 5, pgz = base[2] + r / 25;
                if (pgx < 1 || pgy < 1 || pgz < 1) continue;
                const int pbx = (pgx - 1) / 8 + 1, pby = (pgy - 1) / 8 + 1, pbz = (pgz - 1) / 8 + 1;
                if (pbx > pdx || pby > pdy || pbz > pdz) continue;
                const int32_t pb = P.grid_host[(size_t)(pbx - 1) + (size_t)pdx * ((size_t)(pby - 1) + (size_t)pdy * (pbz - 1))];
                if (pb < 0) continue;
                if (pb >= (1 << 22)) return fail(c, LB_EINVAL, "level %d: parent block index %d too large", L.id, pb);
                const int cc = ((pgx - 1) % 8) + 8 * ((pgy - 1) % 8) + 64 * ((pgz - 1) % 8);
                gregion[(size_t)g * 125 + r] = (pb << 9) | cc;
            }
        }
    }
    const size_t n4 = (size_t)nb * NC, n5 = ((size_t)nb + ngh) * NC * Q;
    L.wide = c->cfg.wide_offsets != 0 || (n5 + 512) * pop_bytes(c->cfg.storage) >= ((size_t)1 << 32);
    L.wide_g = L.wide;  // thermal populations use the same storage
    for (int i = 0; i < 2; i++) {
        // 512 slots of padding: pulls from a missing neighbour block read a
        // harmless slot up to 200 past the cell's own block
        uint8_t* fb = nullptr;
        RC(dalloc(c, &fb, (n5 + 512) * pop_bytes(c->cfg.storage)));
        L.f[i] = fb;
        RC(dalloc(c, &L.rho[i], n4));
        RC(dalloc(c, &L.vel[i], n4 * 3));
        if (c->cfg.ddf_enabled) {
            uint8_t* gb = nullptr;
            RC(dalloc(c, &gb, (n5 + 512) * pop_bytes(c->cfg.storage)));
            L.g[i] = gb;
            RC(dalloc(c, &L.T[i], n4));
        }
    }
    RC(dalloc(c, &L.obst, n4));
    RC(dalloc(c, &L.sponge, n4));
    if (d->wall_dist) RC(dalloc(c, &L.wall_dist, n4));
    RC(dalloc(c, &L.bflags, (size_t)nb));
    RC(dalloc(c, &L.nbr, (size_t)nb * Q));
    RC(dalloc(c, &L.borg, borg.size()));
    RC(dalloc(c, &L.order, (size_t)nb));
    RC(dalloc(c, &L.gent, gent.size()));
    RC(dalloc(c, &L.gpar, gpar.size()));
    if (!gstart.empty()) {
        RC(dalloc(c, &L.gstart, gstart.size()));
        RC(dalloc(c, &L.gregion, std::max<size_t>(gregion.size(), 1)));
        RC(dalloc(c, &L.gbmask, std::max<size_t>(gbmask.size(), 1)));
    }
    const size_t ng = (size_t)d->ptr_dim_x * d->ptr_dim_y * d->ptr_dim_z;
    RC(dalloc(c, &L.grid, ng));
    L.pd[0] = d->ptr_dim_x;
    L.pd[1] = d->ptr_dim_y;
    L.pd[2] = d->ptr_dim_z;
    if (!gent.empty()) RC(h2d(c, L.gent, gent.data(), gent.size()));
    if (!gpar.empty()) RC(h2d(c, L.gpar, gpar.data(), gpar.size()));
    if (!gstart.empty()) RC(h2d(c, L.gstart, gstart.data(), gstart.size()));
    if (!gregion.empty()) RC(h2d(c, L.gregion, gregion.data(), gregion.size()));
    if (!gbmask.empty()) RC(h2d(c, L.gbmask, gbmask.data(), gbmask.size()));
    if (nb > 0) {
        RC(h2d(c, L.nbr, nbr.data(), nbr.size()));
        RC(h2d(c, L.borg, borg.data(), borg.size()));
        // processing order: the storage order, interior blocks (all 26
        // neighbours present, ghosts included) first, then the boundary blocks
        std::vector<int32_t> order;
        order.reserve((size_t)nb);
        for (int pass = 0; pass < 2; pass++)
            for (int32_t b = 0; b < nb; b++) {
                bool interior = true;
                for (int k = 0; k < Q; k++) interior = interior && nbr[(size_t)b * Q + k] >= 0;
                if (interior == (pass == 0)) order.push_back(b);
            }
        L.n_int = 0;
        for (int32_t b : order) {
            bool interior = true;
            for (int k = 0; k < Q; k++) interior = interior && nbr[(size_t)b * Q + k] >= 0;
            if (!interior) break;
            L.n_int++;
        }
        RC(h2d(c, L.order, order.data(), order.size()));
        L.order_host = order;
        std::vector<lbk::Work> wcol((size_t)nb);
        for (int32_t i = 0; i < nb; i++)
            wcol[(size_t)i] = mkwork(L.id - 1, order[(size_t)i], 0, L.nbr + (size_t)order[(size_t)i] * Q);
        RC(dalloc(c, &L.wcol, wcol.size()));
        RC(h2d(c, L.wcol, wcol.data(), wcol.size()));
        RC(h2d(c, L.obst, d->obstacle, n4));
        RC(h2d(c, L.sponge, d->sponge, n4));
        if (d->wall_dist) RC(h2d(c, L.wall_dist, d->wall_dist, n4));
        std::vector<uint8_t> bfl((size_t)nb, 0);
        for (int32_t b = 0; b < nb; b++) {
            uint8_t f = 0;
            for (int i = 0; i < NC; i++) {
                const size_t k = (size_t)b * NC + i;
                if (d->sponge[k] > 0.0f) f |= 1;
                if (d->wall_dist && d->wall_dist[k] > 0.0f && d->wall_dist[k] < 10.0f) f |= 2;
                if (d->obstacle[k]) f |= 4;
            }
            bfl[(size_t)b] = f;
        }
        RC(h2d(c, L.bflags, bfl.data(), bfl.size()));
    }
    if (ng > 0) {
        std::vector<int32_t> grid(ng);
        for (size_t i = 0; i < ng; i++) {
            int v = d->block_pointer[i];
            if (v < 0 || v > nb) return fail(c, LB_EINVAL, "level %d: block_pointer entry out of range", L.id);
            grid[i] = v - 1;
        }
        RC(h2d(c, L.grid, grid.data(), ng));
        L.grid_host = std::move(grid);
    }
    RC(build_links(c, L, d));
    if (staged) RC(item_list(c, L.id - 1, L.ngent > 0 ? ngh : 0, &L.wgh, &L.n_gchunks));
    else RC(chunk_list(c, L.id - 1, L.ngent, &L.wgh, &L.n_gchunks));
    RC(chunk_list(c, L.id - 1, L.nlinks, &L.wlk, &L.n_lchunks));
    RC(upload_level_args(c, L.id - 1));
    L.uploaded = true;
    return LB_OK;
}

int lb_init_equilibrium(lb_ctx* c, float T0) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    HIPCHK(c, hipSetDevice(c->device));
    for (auto& L : c->levels) {
        HIPCHK(c, c->k->init_level(L.f[0], L.f[1], L.rho[0], L.rho[1], L.vel[0], L.vel[1], L.nb, c->cfg.storage,
                                   c->stream));
        if (c->cfg.ddf_enabled)
            HIPCHK(c, c->k->init_thermal(L.g[0], L.g[1], L.T[0], L.T[1], L.nb, T0, c->cfg.storage, c->stream));
        L.cur = 0;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LB_OK;
}

int lb_step_batch(lb_ctx* c, int64_t t_start, int32_t n_steps, float u_curr, const lb_step_params* p) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    if (!p || n_steps < 0 || t_start < 0) return fail(c, LB_EINVAL, "bad step arguments");
    HIPCHK(c, hipSetDevice(c->device));
    for (int64_t t = t_start; t < t_start + n_steps; t++) {
        HIPCHK(c, c->k->set_state(c->st_dev, case_record(t, u_curr, p, c->cfg), c->stream));
        RC(step_level(c, 0, 0));
        set_current(c, t);
    }
    return LB_OK;
}

int lb_synchronize(lb_ctx* c) {
    lbsync::SetupGuard guard;
    if (!c) return LB_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LB_OK;
}

int lb_mesh_upload(lb_ctx* c, int64_t n, const float* cx, const float* cy, const float* cz, const float* nx,
                   const float* ny, const float* nz, const float* areas) {
    lbsync::SetupGuard guard;
    if (!c || n < 0 || (!c->dry && n > 0 && (!cx || !cy || !cz || !nx || !ny || !nz || !areas))) return LB_EINVAL;
    if (c->n_tri != 0) return fail(c, LB_ESTATE, "mesh already uploaded");
    if (!c->dry) HIPCHK(c, hipSetDevice(c->device));
    const float* src[7] = {cx, cy, cz, nx, ny, nz, areas};
    for (int i = 0; i < 7; i++) {
        RC(dalloc(c, &c->mesh[i], (size_t)n));
        if (n > 0) RC(h2d(c, c->mesh[i], src[i], (size_t)n));
    }
    for (int i = 0; i < 4; i++) RC(dalloc(c, &c->maps[i], (size_t)n));
    int groups = (int)((n + 255) / 256);
    RC(dalloc(c, &c->fpart, (size_t)std::max(groups, 1) * 9));
    c->n_tri = n;
    return LB_OK;
}

}  // extern "C"

namespace {
// The force kernel's arguments on the finest level for one side of the surface.
lbk::ForceArgs force_args(lb_ctx* c, const lb_force_params* fp, float side) {
    const Level& L = c->levels.back();
    lbk::ForceArgs a;
    std::memset(&a, 0, sizeof a);
    a.rho = L.rho[L.cur];
    a.vel = L.vel[0];  // level.vel, as surface.jl:784 reads it
    a.obst = L.obst;
    a.grid = L.grid;
    a.gdx = L.pd[0];
    a.gdy = L.pd[1];
    a.gdz = L.pd[2];
    a.dx = (float)L.dx;
    a.n_tri = c->n_tri;
    a.cx = c->mesh[0]; a.cy = c->mesh[1]; a.cz = c->mesh[2];
    a.nx = c->mesh[3]; a.ny = c->mesh[4]; a.nz = c->mesh[5];
    a.area = c->mesh[6];
    a.off_x = (float)fp->mesh_offset[0];
    a.off_y = (float)fp->mesh_offset[1];
    a.off_z = (float)fp->mesh_offset[2];
    const float ps = (float)(fp->rho_physical * fp->velocity_scale * fp->velocity_scale);
    a.pressure_scale = ps;
    a.stress_scale = ps;
    a.tau_mol = L.tau;
    a.search_radius = fp->search_radius;
    a.c_wale = fp->c_wale;
    a.nu_sgs_bg = fp->nu_sgs_background;
    a.ref_x = (float)fp->moment_center[0];
    a.ref_y = (float)fp->moment_center[1];
    a.ref_z = (float)fp->moment_center[2];
    a.p_map = c->maps[0];
    a.sx_map = c->maps[1];
    a.sy_map = c->maps[2];
    a.sz_map = c->maps[3];
    a.partials = c->fpart;
    a.side = side;
    return a;
}
}  // namespace

extern "C" {

int lb_forces(lb_ctx* c, const lb_force_params* fp, double* out, float* p_map, float* sx_map, float* sy_map,
              float* sz_map) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    if (!fp || !out) return fail(c, LB_EINVAL, "null argument");
    if (c->n_tri <= 0) return fail(c, LB_ESTATE, "no mesh uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    const lbk::ForceArgs a = force_args(c, fp, 1.0f);
    const int groups = (int)((c->n_tri + 255) / 256);
    HIPCHK(c, c->k->forces(a, groups, nullptr, c->stream));
    std::vector<double> part((size_t)groups * 9);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->fpart, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (p_map) HIPCHK(c, hipMemcpyAsync(p_map, c->maps[0], (size_t)c->n_tri * 4, hipMemcpyDeviceToHost, c->stream));
    if (sx_map) HIPCHK(c, hipMemcpyAsync(sx_map, c->maps[1], (size_t)c->n_tri * 4, hipMemcpyDeviceToHost, c->stream));
    if (sy_map) HIPCHK(c, hipMemcpyAsync(sy_map, c->maps[2], (size_t)c->n_tri * 4, hipMemcpyDeviceToHost, c->stream));
    if (sz_map) HIPCHK(c, hipMemcpyAsync(sz_map, c->maps[3], (size_t)c->n_tri * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int g = 0; g < groups; g++)
        for (int j = 0; j < 9; j++) s[j] += part[(size_t)g * 9 + j];
    double Fxp = s[0], Fyp = s[1], Fzp = s[2], Fxv = s[3], Fyv = s[4], Fzv = s[5], Mx = s[6], My = s[7], Mz = s[8];
    // integrate_surface_forces! (surface.jl:873-910)
    if (fp->symmetric) {
        Fxp *= 2.0; Fzp *= 2.0; Fxv *= 2.0; Fzv *= 2.0; My *= 2.0;
        Fyp = 0.0; Fyv = 0.0; Mx = 0.0; Mz = 0.0;
    }
    out[LB_F_FXP] = Fxp; out[LB_F_FYP] = Fyp; out[LB_F_FZP] = Fzp;
    out[LB_F_FXV] = Fxv; out[LB_F_FYV] = Fyv; out[LB_F_FZV] = Fzv;
    out[LB_F_MX] = Mx; out[LB_F_MY] = My; out[LB_F_MZ] = Mz;
    out[LB_F_FX] = Fxp + Fxv; out[LB_F_FY] = Fyp + Fyv; out[LB_F_FZ] = Fzp + Fzv;
    double q_inf = 0.5 * fp->rho_ref * fp->u_ref * fp->u_ref;
    double F_ref = q_inf * fp->area_ref;
    double M_ref = F_ref * fp->chord_ref;
    out[LB_F_CD] = out[LB_F_CL] = out[LB_F_CS] = 0.0;
    out[LB_F_CMX] = out[LB_F_CMY] = out[LB_F_CMZ] = 0.0;
    if (F_ref > 1e-10) {
        out[LB_F_CD] = out[LB_F_FX] / F_ref;
        out[LB_F_CL] = out[LB_F_FZ] / F_ref;
        out[LB_F_CS] = out[LB_F_FY] / F_ref;
    }
    if (M_ref > 1e-10) {
        out[LB_F_CMX] = Mx / M_ref;
        out[LB_F_CMY] = My / M_ref;
        out[LB_F_CMZ] = Mz / M_ref;
    }
    return LB_OK;
}

int lb_surface_maps(lb_ctx* c, const lb_force_params* fp, int32_t side, float* p_map, float* sx_map, float* sy_map,
                    float* sz_map) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    if (!fp || !p_map || !sx_map || !sy_map || !sz_map) return fail(c, LB_EINVAL, "null argument");
    if (side != 1 && side != -1) return fail(c, LB_EINVAL, "side must be +1 or -1");
    if (c->n_tri <= 0) return fail(c, LB_ESTATE, "no mesh uploaded");
    HIPCHK(c, hipSetDevice(c->device));
    const lbk::ForceArgs a = force_args(c, fp, (float)side);
    const int groups = (int)((c->n_tri + 255) / 256);
    HIPCHK(c, c->k->forces(a, groups, nullptr, c->stream));
    float* dst[4] = {p_map, sx_map, sy_map, sz_map};
    for (int m = 0; m < 4; m++)
        HIPCHK(c, hipMemcpyAsync(dst[m], c->maps[m], (size_t)c->n_tri * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LB_OK;
}

int lb_download_level(lb_ctx* c, int level_id, int field, float* dst) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    if (!dst || level_id < 1 || level_id > (int)c->levels.size()) return fail(c, LB_EINVAL, "bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->levels[(size_t)level_id - 1];
    const size_t n4 = (size_t)L.nb * NC;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    switch (field) {
        case LB_FIELD_F:
        case LB_FIELD_G: {
            const void* src = field == LB_FIELD_F ? L.f[L.cur] : L.g[L.cur];
            if (!src) return fail(c, LB_EINVAL, "field not allocated");
            const int st = c->cfg.storage;
            std::vector<uint8_t> tmp(n4 * Q * pop_bytes(st));
            HIPCHK(c, hipMemcpy(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost));
            for (int k = 0; k < Q; k++) {
                const float w = wk_host(k);
                for (int64_t b = 0; b < L.nb; b++) {
                    float* out = dst + (size_t)k * n4 + (size_t)L.perm[(size_t)b] * NC;
                    for (int cl = 0; cl < NC; cl++) out[cl] = decode_pop(tmp.data(), fidx(b, k, cl), st, w);
                }
            }
            return LB_OK;
        }
        case LB_FIELD_RHO:
        case LB_FIELD_T: {
            const float* src = field == LB_FIELD_RHO ? L.rho[L.cur] : L.T[L.cur];
            if (!src) return fail(c, LB_EINVAL, "field not allocated");
            std::vector<float> tmp(n4);
            HIPCHK(c, hipMemcpy(tmp.data(), src, n4 * 4, hipMemcpyDeviceToHost));
            for (int64_t b = 0; b < L.nb; b++)
                std::memcpy(dst + (size_t)L.perm[(size_t)b] * NC, tmp.data() + (size_t)b * NC, NC * 4);
            return LB_OK;
        }
        case LB_FIELD_VEL: {
            std::vector<float> tmp(n4 * 3);
            HIPCHK(c, hipMemcpy(tmp.data(), L.vel[L.cur], tmp.size() * 4, hipMemcpyDeviceToHost));
            for (int64_t b = 0; b < L.nb; b++)
                for (int d = 0; d < 3; d++)
                    std::memcpy(dst + (size_t)d * n4 + (size_t)L.perm[(size_t)b] * NC,
                                tmp.data() + ((size_t)b * 3 + d) * NC, NC * 4);
            return LB_OK;
        }
        default:
            return fail(c, LB_EINVAL, "unknown field %d", field);
    }
}

// Overwrite the current state of a level (both halves of the ping-pong pair,
// so the next step reads it whatever its parity).
int lb_upload_level(lb_ctx* c, int level_id, int field, const float* src) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    if (!src || level_id < 1 || level_id > (int)c->levels.size()) return fail(c, LB_EINVAL, "bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    Level& L = c->levels[(size_t)level_id - 1];
    const size_t n4 = (size_t)L.nb * NC;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<float> tmp;
    float* const* pair;
    size_t n;
    switch (field) {
        case LB_FIELD_F:
        case LB_FIELD_G: {
            void* const dev[2] = {field == LB_FIELD_F ? L.f[0] : L.g[0], field == LB_FIELD_F ? L.f[1] : L.g[1]};
            if (!dev[0]) return fail(c, LB_EINVAL, "field not allocated");
            const int st = c->cfg.storage;
            std::vector<uint8_t> raw(n4 * Q * pop_bytes(st));
            for (int k = 0; k < Q; k++) {
                const float w = wk_host(k);
                for (int64_t b = 0; b < L.nb; b++) {
                    const float* in = src + (size_t)k * n4 + (size_t)L.perm[(size_t)b] * NC;
                    for (int cl = 0; cl < NC; cl++) encode_pop(raw.data(), fidx(b, k, cl), in[cl], st, w);
                }
            }
            for (int i = 0; i < 2; i++) HIPCHK(c, hipMemcpy(dev[i], raw.data(), raw.size(), hipMemcpyHostToDevice));
            return LB_OK;
        }
        case LB_FIELD_RHO:
        case LB_FIELD_T:
            pair = field == LB_FIELD_RHO ? L.rho : L.T;
            if (!pair[0]) return fail(c, LB_EINVAL, "field not allocated");
            tmp.resize(n4);
            for (int64_t b = 0; b < L.nb; b++)
                std::memcpy(tmp.data() + (size_t)b * NC, src + (size_t)L.perm[(size_t)b] * NC, NC * 4);
            n = n4;
            break;
        case LB_FIELD_VEL:
            pair = L.vel;
            tmp.resize(n4 * 3);
            for (int64_t b = 0; b < L.nb; b++)
                for (int d = 0; d < 3; d++)
                    std::memcpy(tmp.data() + ((size_t)b * 3 + d) * NC, src + (size_t)d * n4 + (size_t)L.perm[(size_t)b] * NC,
                                NC * 4);
            n = tmp.size();
            break;
        default:
            return fail(c, LB_EINVAL, "unknown field %d", field);
    }
    for (int i = 0; i < 2; i++) HIPCHK(c, hipMemcpy(pair[i], tmp.data(), n * 4, hipMemcpyHostToDevice));
    return LB_OK;
}

int lb_stats(lb_ctx* c, int level_id, double* out) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    if (!out || level_id < 1 || level_id > (int)c->levels.size()) return fail(c, LB_EINVAL, "bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->levels[(size_t)level_id - 1];
    const int groups = std::max(1, std::min<int>(L.nb, kStatsRows));
    if (c->n_stat_groups < groups) {
        RC(dalloc(c, &c->spart, (size_t)kStatsRows * 8));
        c->n_stat_groups = kStatsRows;
    }
    for (int i = 0; i < 8; i++) out[i] = 0;
    if (L.nb == 0) return LB_OK;
    HIPCHK(c, c->k->stats(L.f[L.cur], L.rho[L.cur], L.vel[L.cur], L.obst, L.nb, c->cfg.storage, c->spart, groups,
                          c->stream));
    std::vector<double> part((size_t)groups * 8);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->spart, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double mn = 1e300, mx = -1e300, vmax = 0;
    for (int g = 0; g < groups; g++) {
        const double* q = &part[(size_t)g * 8];
        out[0] += q[0];
        out[1] += q[1];
        mn = std::min(mn, q[2]);
        mx = std::max(mx, q[3]);
        out[4] += q[4];
        out[5] += q[5];
        out[6] += q[6];
        vmax = std::max(vmax, q[7]);
    }
    out[2] = mn;
    out[3] = mx;
    out[7] = vmax;
    return LB_OK;
}

int lb_check_finite(lb_ctx* c, int64_t* n_nonfinite) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<lbk::FiniteEntry> ents;
    for (const Level& L : c->levels)
        if (L.nb > 0) ents.push_back(finite_entry(L, L.cur, 0));
    int32_t count = 0;
    if (!ents.empty()) {
        HIPCHK(c, hipMemcpyAsync(c->fin_ents, ents.data(), ents.size() * sizeof(lbk::FiniteEntry),
                                 hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemsetAsync(c->fin_count, 0, sizeof(int32_t), c->stream));
        HIPCHK(c, c->k->finite_count(c->fin_ents, (int32_t)ents.size(), kFiniteGroups, c->fin_count, c->stream));
        HIPCHK(c, hipMemcpyAsync(&count, c->fin_count, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (n_nonfinite) *n_nonfinite = count;
    if (count > 0) return fail(c, LB_ENONFINITE, "%d non-finite density/velocity values", count);
    return LB_OK;
}

int lb_stats_noalloc(lb_ctx* c, int level_id, double* out) {
    lbsync::SetupGuard guard;
    RC(check_ctx(c));
    if (!out || level_id < 1 || level_id > (int)c->levels.size()) return fail(c, LB_EINVAL, "bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    const Level& L = c->levels[(size_t)level_id - 1];
    const int groups = std::max(1, std::min<int>(L.nb, kStatsRows));
    if (groups > kStatsRows)
        return fail(c, LB_EINVAL, "lb_stats_noalloc: %d partial rows exceed the scratch of %d", groups, kStatsRows);
    for (int i = 0; i < 8; i++) out[i] = 0;
    if (L.nb == 0) return LB_OK;
    hipError_t se = hipSuccess;
    double* s = scratch_on(c->device, &se);
    if (!s) return fail(c, LB_EHIP, "lb_stats_noalloc: no scratch address on device %d (%s)", c->device,
                        hipGetErrorString(se));
    HIPCHK(c, c->k->stats(L.f[L.cur], L.rho[L.cur], L.vel[L.cur], L.obst, L.nb, c->cfg.storage, s, groups,
                          c->stream));
    double* res = s + (size_t)kStatsRows * 8;
    hipLaunchKernelGGL(stats_reduce_kernel, dim3(1), dim3(64), 0, c->stream, s, groups, res);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, res, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LB_OK;
}

int64_t lb_device_bytes(const lb_ctx* c) { return c ? c->bytes : 0; }

int lb_profile_enable(lb_ctx* c, int enable) {
    if (!c) return LB_EINVAL;
    c->profile = enable != 0;
    return LB_OK;
}

int lb_profile_read_kernel(lb_ctx* c, int level_id, int kernel, double* ms, int64_t* launches) {
    lbsync::SetupGuard guard;
    if (!c || !ms || !launches || level_id < 1 || level_id > (int)c->levels.size() || kernel < 0 || kernel > 1)
        return LB_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    Level& L = c->levels[(size_t)level_id - 1];
    double tot = 0.0;
    for (size_t i = 0; i < L.ev_used[kernel]; i++) {
        float t = 0.0f;
        HIPCHK(c, hipEventElapsedTime(&t, L.ev[kernel][i].first, L.ev[kernel][i].second));
        tot += (double)t;
    }
    *ms = tot;
    *launches = (int64_t)L.ev_used[kernel];
    L.ev_used[kernel] = 0;
    return LB_OK;
}

int lb_profile_read(lb_ctx* c, int level_id, double* ms, int64_t* launches) {
    return lb_profile_read_kernel(c, level_id, 0, ms, launches);
}

int lb_engine_create(lb_engine** out, int device, int32_t storage, int32_t strict_fp) {
    lbsync::SetupGuard guard;
    if (!out) return LB_EINVAL;
    *out = nullptr;
    if (storage < LB_STORAGE_FP32 || storage > LB_STORAGE_BF16) return LB_EINVAL;
    lb_engine* e = new lb_engine();
    e->device = device;
    e->storage = storage;
    e->waves = default_waves(storage);
    e->k = strict_fp ? &kStrict : &kFast;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return LB_EHIP;
    }
    *out = e;
    return LB_OK;
}

void lb_engine_destroy(lb_engine* e) {
    lbsync::SetupGuard guard;
    if (!e) return;
    (void)hipSetDevice(e->device);
    (void)hipStreamSynchronize(e->stream);
    for (EngineCase& ec : e->cases) ec.c->engine = nullptr;
    engine_free_lists(e);
    if (e->fin_ents) (void)hipFree(e->fin_ents);
    if (e->fin_counts) (void)hipFree(e->fin_counts);
    (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* lb_engine_last_error(const lb_engine* e) { return e ? e->err.c_str() : "null engine"; }

int lb_engine_add(lb_engine* e, lb_ctx* c, int64_t n_coarse_steps, const float* u_batches, int32_t n_batches,
                  int32_t batch_len, const lb_step_params* p) {
    lbsync::SetupGuard guard;
    if (!e || !c || !u_batches || !p) return LB_EINVAL;
    RC(check_ctx(c));
    if (c->device != e->device) return efail(e, LB_EINVAL, "case is on device %d, engine on %d", c->device, e->device);
    if (c->cfg.storage != e->storage) return efail(e, LB_EINVAL, "case storage differs from the engine's");
    if ((c->k == &kStrict) != (e->k == &kStrict)) return efail(e, LB_EINVAL, "case and engine builds differ");
    if (n_coarse_steps < 1 || batch_len < 1 || (int64_t)n_batches * batch_len < n_coarse_steps)
        return efail(e, LB_EINVAL, "bad step count or ramp table");
    EHIPCHK(e, hipSetDevice(e->device));
    EHIPCHK(e, hipStreamSynchronize(c->stream));
    EngineCase ec;
    ec.c = c;
    ec.n_total = n_coarse_steps;
    ec.t0 = 1;
    ec.batch = batch_len;
    ec.p = *p;
    if (dalloc(c, &ec.u_tab, (size_t)n_batches) != LB_OK) return efail(e, LB_ENOMEM, "%s", c->err.c_str());
    if (h2d(c, ec.u_tab, u_batches, (size_t)n_batches) != LB_OK) return efail(e, LB_EHIP, "%s", c->err.c_str());
    if (dalloc(c, &ec.acc, 10) != LB_OK) return efail(e, LB_ENOMEM, "%s", c->err.c_str());
    EHIPCHK(e, hipMemset(ec.acc, 0, 10 * sizeof(double)));
    c->engine = e;
    e->cases.push_back(ec);
    e->dirty = true;
    return LB_OK;
}

int lb_engine_remove(lb_engine* e, lb_ctx* c) {
    lbsync::SetupGuard guard;
    if (!e || !c) return LB_EINVAL;
    EHIPCHK(e, hipSetDevice(e->device));
    for (size_t i = 0; i < e->cases.size(); i++) {
        if (e->cases[i].c != c) continue;
        EHIPCHK(e, hipStreamSynchronize(e->stream));
        const int64_t done = std::min(e->cases[i].t0 - 1, e->cases[i].n_total);
        if (done >= 1) set_current(c, done);
        c->engine = nullptr;
        e->cases.erase(e->cases.begin() + (long)i);
        e->dirty = true;
        return LB_OK;
    }
    return efail(e, LB_EINVAL, "case is not attached to this engine");
}

int lb_engine_step(lb_engine* e, int32_t n) {
    if (!e || n < 0) return LB_EINVAL;
    EHIPCHK(e, hipSetDevice(e->device));
    if (e->cases.empty() || n == 0) return LB_OK;
    if (e->dirty) {
        lbsync::SetupGuard guard;
        RC(engine_rebuild(e));
    }
    if (e->use_graph && !e->exec) {
        // GPU calls of other threads wait for the capture (lb_sync.h);
        // relaxed mode in addition lets them run host-side work meanwhile
        std::unique_lock<lbsync::CaptureLock> capture(lbsync::capture_mutex());
        EHIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
        const hipError_t r = engine_issue_step(e);
        hipGraph_t g = nullptr;
        const hipError_t r2 = hipStreamEndCapture(e->stream, &g);
        if (r != hipSuccess || r2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return efail(e, LB_EHIP, "graph capture failed: %s", hipGetErrorString(r != hipSuccess ? r : r2));
        }
        e->graph = g;
        EHIPCHK(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    }
    for (int32_t i = 0; i < n; i++) {
        if (e->use_graph) EHIPCHK(e, hipGraphLaunch(e->exec, e->stream));
        else EHIPCHK(e, engine_issue_step(e));
        for (EngineCase& ec : e->cases)
            if (ec.t0 <= ec.n_total) ec.t0 += ec.t_inc;
    }
    return LB_OK;
}

int lb_engine_synchronize(lb_engine* e) {
    lbsync::SetupGuard guard;
    if (!e) return LB_EINVAL;
    EHIPCHK(e, hipSetDevice(e->device));
    EHIPCHK(e, hipStreamSynchronize(e->stream));
    return LB_OK;
}

int lb_engine_case_progress(const lb_engine* e, const lb_ctx* c, int64_t* t_done) {
    if (!e || !c || !t_done) return LB_EINVAL;
    for (const EngineCase& ec : e->cases)
        if (ec.c == c) {
            *t_done = std::min(ec.t0 - 1, ec.n_total);
            return LB_OK;
        }
    return LB_EINVAL;
}

int lb_engine_sample_forces(lb_engine* e, lb_ctx* c, const lb_force_params* fp, int64_t diag_freq,
                            int64_t ramp_steps) {
    lbsync::SetupGuard guard;
    if (!e || !c || !fp || diag_freq < 1 || ramp_steps < 0) return LB_EINVAL;
    for (EngineCase& ec : e->cases) {
        if (ec.c != c) continue;
        if (c->n_tri <= 0) return efail(e, LB_ESTATE, "case has no mesh");
        EHIPCHK(e, hipSetDevice(e->device));
        EHIPCHK(e, hipStreamSynchronize(e->stream));
        ec.sample = true;
        ec.diag = diag_freq;
        ec.ramp = ramp_steps;
        ec.fp = *fp;
        e->dirty = true;
        return LB_OK;
    }
    return efail(e, LB_EINVAL, "case is not attached to this engine");
}

int lb_engine_case_forces(lb_engine* e, const lb_ctx* c, double* sums, int64_t* n_samples) {
    lbsync::SetupGuard guard;
    if (!e || !c || !sums || !n_samples) return LB_EINVAL;
    for (const EngineCase& ec : e->cases) {
        if (ec.c != c) continue;
        EHIPCHK(e, hipSetDevice(e->device));
        EHIPCHK(e, hipStreamSynchronize(e->stream));
        double acc[10];
        EHIPCHK(e, hipMemcpy(acc, ec.acc, sizeof acc, hipMemcpyDeviceToHost));
        for (int i = 0; i < 9; i++) sums[i] = acc[i];
        // integrate_surface_forces! (surface.jl:873-910) doubles the half model
        if (ec.fp.symmetric) {
            sums[0] *= 2.0; sums[2] *= 2.0; sums[3] *= 2.0; sums[5] *= 2.0; sums[7] *= 2.0;
            sums[1] = 0.0; sums[4] = 0.0; sums[6] = 0.0; sums[8] = 0.0;
        }
        *n_samples = (int64_t)acc[9];
        return LB_OK;
    }
    return efail(e, LB_EINVAL, "case is not attached to this engine");
}

int lb_engine_check_finite(lb_engine* e, lb_ctx** diverged, int32_t cap, int32_t* n_diverged) {
    lbsync::SetupGuard guard;
    if (!e || !n_diverged || cap < 0 || (cap > 0 && !diverged)) return LB_EINVAL;
    *n_diverged = 0;
    EHIPCHK(e, hipSetDevice(e->device));
    EHIPCHK(e, hipStreamSynchronize(e->stream));
    const size_t nc = e->cases.size();
    if (nc == 0) return LB_OK;
    std::vector<lbk::FiniteEntry> ents;
    for (size_t i = 0; i < nc; i++) {
        const EngineCase& ec = e->cases[i];
        const int64_t done = std::min(ec.t0 - 1, ec.n_total);
        for (size_t li = 0; li < ec.c->levels.size(); li++) {
            const Level& L = ec.c->levels[li];
            if (L.nb == 0) continue;
            // output buffer of the level's last substep in coarse step `done`
            const int64_t last = (done << li) + ((int64_t)1 << li) - 1;
            ents.push_back(finite_entry(L, done >= 1 ? 1 - (int)(last & 1) : L.cur, (int32_t)i));
        }
    }
    if (ents.size() > e->fin_cap_ents || nc > e->fin_cap_cases) {
        if (e->fin_ents) (void)hipFree(e->fin_ents);
        if (e->fin_counts) (void)hipFree(e->fin_counts);
        e->fin_ents = nullptr;
        e->fin_counts = nullptr;
        e->fin_cap_ents = e->fin_cap_cases = 0;
        const size_t ce = std::max<size_t>(ents.size(), 64), cc = std::max<size_t>(nc, 16);
        EHIPCHK(e, hipMalloc((void**)&e->fin_ents, ce * sizeof(lbk::FiniteEntry)));
        EHIPCHK(e, hipMalloc((void**)&e->fin_counts, cc * sizeof(int32_t)));
        e->fin_cap_ents = ce;
        e->fin_cap_cases = cc;
    }
    std::vector<int32_t> counts(nc, 0);
    EHIPCHK(e, hipMemcpyAsync(e->fin_ents, ents.data(), ents.size() * sizeof(lbk::FiniteEntry), hipMemcpyHostToDevice,
                              e->stream));
    EHIPCHK(e, hipMemsetAsync(e->fin_counts, 0, nc * sizeof(int32_t), e->stream));
    EHIPCHK(e, e->k->finite_count(e->fin_ents, (int32_t)ents.size(), kFiniteGroups, e->fin_counts, e->stream));
    EHIPCHK(e, hipMemcpyAsync(counts.data(), e->fin_counts, nc * sizeof(int32_t), hipMemcpyDeviceToHost, e->stream));
    EHIPCHK(e, hipStreamSynchronize(e->stream));
    int32_t n = 0;
    for (size_t i = 0; i < nc; i++)
        if (counts[i] > 0) {
            if (n < cap) diverged[n] = e->cases[i].c;
            n++;
        }
    *n_diverged = n;
    if (n > 0) return efail(e, LB_ENONFINITE, "%d case(s) hold non-finite density/velocity values", n);
    return LB_OK;
}

int lb_engine_set_options(lb_engine* e, int32_t use_graph, int32_t schedule, int32_t waves) {
    if (!e || (schedule != LB_SCHED_XCD && schedule != LB_SCHED_LINEAR && schedule != LB_SCHED_PERSISTENT) ||
        waves < 5 || waves > 6)
        return LB_EINVAL;
    EHIPCHK(e, hipSetDevice(e->device));
    EHIPCHK(e, hipStreamSynchronize(e->stream));
    e->use_graph = use_graph != 0;
    e->sched = schedule;
    e->waves = waves;
    // the captured graph bakes the schedule in
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    return LB_OK;
}

// Device bytes a case will hold: the allocations of lb_create,
// lb_level_upload (every level) and lb_mesh_upload, counted by running the
// same host code without allocating (no GPU needed), plus the engine's ramp
// table of n_ramp_batches values.  lb_stats adds 65,536 bytes on first use.
int lb_footprint(const lb_config* cfg, const lb_level_desc* levels, int32_t n_levels, int64_t n_tri,
                 int32_t n_ramp_batches, int64_t* bytes) {
    if (!cfg || !levels || !bytes || n_levels != cfg->n_levels || n_levels < 1 || n_levels > 16) return LB_EINVAL;
    lb_ctx c;
    c.dry = true;
    c.cfg = *cfg;
    c.levels.resize((size_t)n_levels);
    c.las_host.resize((size_t)n_levels);
    RC(dalloc(&c, &c.st_dev, 1));
    RC(dalloc(&c, &c.las_dev, (size_t)n_levels));
    RC(dalloc(&c, &c.fin_ents, (size_t)n_levels));
    RC(dalloc(&c, &c.fin_count, 1));
    for (int i = 0; i < n_levels; i++) RC(lb_level_upload(&c, &levels[i]));
    if (n_tri > 0) RC(lb_mesh_upload(&c, n_tri, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr));
    if (n_ramp_batches > 0) {
        float* u = nullptr;
        double* acc = nullptr;
        RC(dalloc(&c, &u, (size_t)n_ramp_batches));
        RC(dalloc(&c, &acc, 10));
    }
    *bytes = c.bytes;
    return LB_OK;
}

int lb_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes) {
    lbsync::SetupGuard guard;
    if (!free_bytes || !total_bytes) return LB_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return LB_EHIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return LB_EHIP;
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return LB_OK;
}

}  // extern "C"
