// lb_setup_gpu.hip -- setup steps on the GPU: the Bouzidi Q-map
// (compute_bouzidi_qmap_sparse, src/geometry/bouzidi_setup.jl:64-167;
// compute_q_for_cell, bouzidi_math.jl:53-102), the SAT voxelization
// (voxelize_blocks!, domain_generation.jl:74-112) and the wall distances
// (compute_wall_distances!, domain_generation.jl:371-434).
//
// Same inputs, outputs and bits as lbs_bouzidi_qmap in the CPU setup library:
// the block -> candidate-triangle lists are built on the host (lb_geom.h), and
// one work-group of 512 work-items (one per cell) per candidate block casts
// the 26 links of every cell against the block's triangles, which it stages in
// LDS in index order.  Per direction the first minimum distance wins, as in
// the reference's loop; FP64 throughout, no FMA contraction (this file is
// compiled with -ffp-contract=off), so q bits and triangle ids equal the CPU's.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "lb_geom.h"
#include "lb_sync.h"
#include "ludwig_hip.h"

namespace {

using namespace lbgeom;

constexpr int TRI_CHUNK = 256;  // triangles staged in LDS per pass (18 KiB)

struct QmapArgs {
    const int32_t* coords;   // (nb, 3) 1-based block coordinates
    const int32_t* blocks;   // candidate block ids (0-based), in block order
    const int64_t* tri_off;  // per candidate block: first entry in tri_idx
    const int32_t* tri_idx;  // concatenated candidate triangle ids (index order)
    const double* tris;      // (ntri, 3, 3) raw vertices
    double ox, oy, oz, dx;
    double dn[27][3];        // unit link directions
    double cmag[27];         // |c_k|
    uint16_t* q;             // (n_cand, 512, 27) FP16 bits, 0 = no link
    int32_t* tri;            // (n_cand, 512, 27) 1-based triangle id, 0 = none
    uint8_t* hit;            // (n_cand, 512) any link
};

__global__ void __launch_bounds__(512) qmap_kernel(const QmapArgs a) {
    __shared__ double sv[TRI_CHUNK][9];
    __shared__ int32_t sid[TRI_CHUNK];
    const int cb = blockIdx.x;
    const int b = a.blocks[cb];
    const int cell = threadIdx.x;
    const int lx = (cell & 7) + 1, ly = ((cell >> 3) & 7) + 1, lz = (cell >> 6) + 1;
    const int64_t bx = a.coords[b * 3], by = a.coords[b * 3 + 1], bz = a.coords[b * 3 + 2];
    const V3 c = {((double)((bx - 1) * BS + lx) - 0.5) * a.dx, ((double)((by - 1) * BS + ly) - 0.5) * a.dx,
                  ((double)((bz - 1) * BS + lz) - 0.5) * a.dx};
    double tmin[27];
    int32_t best[27];
#pragma unroll
    for (int k = 0; k < 27; k++) {
        tmin[k] = std::numeric_limits<double>::infinity();
        best[k] = -1;
    }
    const int64_t t0 = a.tri_off[cb], t1 = a.tri_off[cb + 1];
    for (int64_t base = t0; base < t1; base += TRI_CHUNK) {
        const int n = (int)((t1 - base) < TRI_CHUNK ? (t1 - base) : TRI_CHUNK);
        __syncthreads();
        if (cell < n) {
            const int32_t t = a.tri_idx[base + cell];
            const double* p = a.tris + (int64_t)t * 9;
            sv[cell][0] = p[0] + a.ox; sv[cell][1] = p[1] + a.oy; sv[cell][2] = p[2] + a.oz;
            sv[cell][3] = p[3] + a.ox; sv[cell][4] = p[4] + a.oy; sv[cell][5] = p[5] + a.oz;
            sv[cell][6] = p[6] + a.ox; sv[cell][7] = p[7] + a.oy; sv[cell][8] = p[8] + a.oz;
            sid[cell] = t;
        }
        __syncthreads();
        for (int i = 0; i < n; i++) {
            const V3 v1 = {sv[i][0], sv[i][1], sv[i][2]}, v2 = {sv[i][3], sv[i][4], sv[i][5]},
                     v3 = {sv[i][6], sv[i][7], sv[i][8]};
#pragma unroll
            for (int k = 0; k < 27; k++) {
                if (k == 13) continue;
                const V3 d = {a.dn[k][0], a.dn[k][1], a.dn[k][2]};
                double t;
                if (ray_tri(c, d, v1, v2, v3, &t) && t < tmin[k]) {
                    tmin[k] = t;
                    best[k] = sid[i];
                }
            }
        }
    }
    const size_t o = ((size_t)cb * 512 + cell) * 27;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 27; k++) {
        uint16_t qb = 0;
        int32_t tb = 0;
        if (k != 13 && tmin[k] < std::numeric_limits<double>::infinity()) {
            const double q = tmin[k] / (a.dx * a.cmag[k]);
            if (q > 0.0 && q <= 1.0) {
                qb = double_to_half(q);
                tb = best[k] + 1;
                any = true;
            }
        }
        a.q[o + k] = qb;
        a.tri[o + k] = tb;
    }
    a.hit[(size_t)cb * 512 + cell] = any ? 1 : 0;
}

// voxelize_blocks! (domain_generation.jl:74-112): a cell is solid if any
// candidate triangle of its block passes the SAT test against the cell's box
// (half size 0.75 dx).  One work-group of 512 work-items per candidate block,
// triangles staged in LDS; the result is an OR, so the order of the tests
// does not matter.
struct VoxArgs {
    const int32_t* coords;
    const int32_t* blocks;
    const int64_t* tri_off;
    const int32_t* tri_idx;
    const double* tris;
    double ox, oy, oz, dx;
    uint8_t* obst;  // (nb, 512)
};

__global__ void __launch_bounds__(512) vox_kernel(const VoxArgs a) {
    __shared__ double sv[TRI_CHUNK][9];
    const int cb = blockIdx.x;
    const int b = a.blocks[cb];
    const int cell = threadIdx.x;
    const int lx = (cell & 7) + 1, ly = ((cell >> 3) & 7) + 1, lz = (cell >> 6) + 1;
    const int64_t bx = a.coords[b * 3], by = a.coords[b * 3 + 1], bz = a.coords[b * 3 + 2];
    const V3 c = {((double)((bx - 1) * BS + lx) - 0.5) * a.dx, ((double)((by - 1) * BS + ly) - 0.5) * a.dx,
                  ((double)((bz - 1) * BS + lz) - 0.5) * a.dx};
    const double hb = 0.75 * a.dx;
    bool shell = false;
    const int64_t t0 = a.tri_off[cb], t1 = a.tri_off[cb + 1];
    for (int64_t base = t0; base < t1; base += TRI_CHUNK) {
        const int n = (int)((t1 - base) < TRI_CHUNK ? (t1 - base) : TRI_CHUNK);
        __syncthreads();
        if (cell < n) {
            const double* p = a.tris + (int64_t)a.tri_idx[base + cell] * 9;
            sv[cell][0] = p[0] + a.ox; sv[cell][1] = p[1] + a.oy; sv[cell][2] = p[2] + a.oz;
            sv[cell][3] = p[3] + a.ox; sv[cell][4] = p[4] + a.oy; sv[cell][5] = p[5] + a.oz;
            sv[cell][6] = p[6] + a.ox; sv[cell][7] = p[7] + a.oy; sv[cell][8] = p[8] + a.oz;
        }
        __syncthreads();
        for (int i = 0; i < n && !shell; i++) {
            const V3 v1 = {sv[i][0], sv[i][1], sv[i][2]}, v2 = {sv[i][3], sv[i][4], sv[i][5]},
                     v3 = {sv[i][6], sv[i][7], sv[i][8]};
            shell = tri_box(c, hb, v1, v2, v3);
        }
    }
    if (shell) a.obst[(int64_t)b * 512 + cell] = 1;
}

// compute_wall_distances! (domain_generation.jl:371-434): for a fluid cell
// with a solid cell among its 26 neighbours (across block faces through the
// block grid), the distance to the nearest one, |offset| dx in FP32 (the three
// offset lengths come from the host: sqrt(1), sqrt(2), sqrt(3) as FP32);
// other cells keep their value.  Blocks whose own and neighbour blocks hold no
// solid cell are skipped, as on the CPU.
struct WallArgs {
    const int32_t* coords;
    const int32_t* grid;
    int gdx, gdy, gdz;
    const uint8_t* obst;
    const uint8_t* solid;  // per block: any solid cell
    float len[4];          // len[n] = sqrt(n) as FP32, n = 1..3
    float dxf;
    float* wall;
    unsigned long long* count;
};

__device__ __forceinline__ int32_t grid_at(const WallArgs& a, int64_t bx, int64_t by, int64_t bz) {
    if (bx < 1 || bx > a.gdx || by < 1 || by > a.gdy || bz < 1 || bz > a.gdz) return 0;
    return a.grid[(bx - 1) + (int64_t)a.gdx * ((by - 1) + (int64_t)a.gdy * (bz - 1))];
}

__global__ void __launch_bounds__(512) solid_kernel(const uint8_t* obst, int64_t nb, uint8_t* solid) {
    const int64_t b = blockIdx.x;
    __shared__ int any;
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (obst[b * 512 + threadIdx.x]) any = 1;
    __syncthreads();
    if (threadIdx.x == 0) solid[b] = any ? 1 : 0;
}

__global__ void __launch_bounds__(512) wall_kernel(const WallArgs a) {
    const int64_t b = blockIdx.x;
    const int64_t bx = a.coords[b * 3], by = a.coords[b * 3 + 1], bz = a.coords[b * 3 + 2];
    bool any = false;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dxo = -1; dxo <= 1; dxo++) {
                const int32_t id = (dxo == 0 && dy == 0 && dz == 0) ? (int32_t)(b + 1) : grid_at(a, bx + dxo, by + dy, bz + dz);
                if (id > 0 && a.solid[id - 1]) any = true;
            }
    if (!any) return;  // uniform per work-group
    const int cell = threadIdx.x;
    const int lx = (cell & 7) + 1, ly = ((cell >> 3) & 7) + 1, lz = (cell >> 6) + 1;
    if (a.obst[b * 512 + cell]) return;
    bool near = false;
    float mind = 100.0f;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dxo = -1; dxo <= 1; dxo++) {
                if (dxo == 0 && dy == 0 && dz == 0) continue;
                int nx = lx + dxo, ny = ly + dy, nz = lz + dz;
                int64_t nbi = b;
                if (nx < 1 || nx > BS || ny < 1 || ny > BS || nz < 1 || nz > BS) {
                    const int64_t nbx = bx + (nx < 1 ? -1 : (nx > BS ? 1 : 0));
                    const int64_t nby = by + (ny < 1 ? -1 : (ny > BS ? 1 : 0));
                    const int64_t nbz = bz + (nz < 1 ? -1 : (nz > BS ? 1 : 0));
                    const int32_t id = grid_at(a, nbx, nby, nbz);
                    if (id <= 0) continue;
                    nbi = id - 1;
                    nx = nx < 1 ? nx + BS : (nx > BS ? nx - BS : nx);
                    ny = ny < 1 ? ny + BS : (ny > BS ? ny - BS : ny);
                    nz = nz < 1 ? nz + BS : (nz > BS ? nz - BS : nz);
                }
                if (a.obst[nbi * 512 + (nx - 1) + 8 * (ny - 1) + 64 * (nz - 1)]) {
                    near = true;
                    const float d = a.len[dxo * dxo + dy * dy + dz * dz] * a.dxf;
                    mind = mind < d ? mind : d;
                }
            }
    if (near) {
        a.wall[b * 512 + cell] = mind;
        atomicAdd(a.count, 1ull);
    }
}

thread_local char g_err[512];

int qfail(int code, const char* what, hipError_t e) {
    std::snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
    return code;
}

}  // namespace

extern "C" {

const char* lb_setup_gpu_last_error(void) { return g_err; }

void lb_free(void* p) { std::free(p); }

int lb_bouzidi_qmap_gpu(int device, int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                        const double* tris, int64_t ntri, double dx, const double* offset, int32_t** cells_out,
                        uint16_t** q_out, int32_t** tri_out, int64_t* n_out) {
    if (!coords || !grid || !tris || !offset || !cells_out || !q_out || !tri_out || !n_out || nb < 0 || ntri < 0)
        return LB_EINVAL;
    lbsync::SetupGuard guard;
    *cells_out = nullptr;
    *q_out = nullptr;
    *tri_out = nullptr;
    *n_out = 0;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return qfail(LB_EHIP, "hipSetDevice", e);
    Grid g{gdx, gdy, gdz, grid};
    auto bt = block_tri_map(nb, tris, ntri, g, dx, offset, 1);
    std::vector<int32_t> blocks, idx;
    std::vector<int64_t> off(1, 0);
    for (int64_t b = 0; b < nb; b++) {
        if (bt[(size_t)b].empty()) continue;
        blocks.push_back((int32_t)b);
        idx.insert(idx.end(), bt[(size_t)b].begin(), bt[(size_t)b].end());
        off.push_back((int64_t)idx.size());
    }
    const int64_t nc = (int64_t)blocks.size();
    QmapArgs a;
    std::memset(&a, 0, sizeof a);
    a.ox = offset[0];
    a.oy = offset[1];
    a.oz = offset[2];
    a.dx = dx;
    {
        int k = 0;
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++, k++) {
                    double m = std::sqrt((double)(x * x) + (double)(y * y) + (double)(z * z));
                    a.cmag[k] = m;
                    a.dn[k][0] = m > 0 ? x / m : 0;
                    a.dn[k][1] = m > 0 ? y / m : 0;
                    a.dn[k][2] = m > 0 ? z / m : 0;
                }
    }
    std::vector<uint16_t> hq;
    std::vector<int32_t> ht;
    std::vector<uint8_t> hh;
    if (nc > 0) {
        // a private stream: the call may run beside other work on the device
        hipStream_t st = nullptr;
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
            return qfail(LB_EHIP, "hipStreamCreate", e);
        void *dcoords = nullptr, *dblocks = nullptr, *doff = nullptr, *didx = nullptr, *dtris = nullptr,
             *dq = nullptr, *dt = nullptr, *dh = nullptr;
        const size_t ncell = (size_t)nc * 512;
        struct Buf {
            void** p;
            size_t n;
            const void* src;
        } bufs[] = {{&dcoords, (size_t)nb * 3 * 4, coords}, {&dblocks, (size_t)nc * 4, blocks.data()},
                    {&doff, off.size() * 8, off.data()},      {&didx, std::max<size_t>(idx.size(), 1) * 4, idx.data()},
                    {&dtris, (size_t)ntri * 72, tris},        {&dq, ncell * 27 * 2, nullptr},
                    {&dt, ncell * 27 * 4, nullptr},           {&dh, ncell, nullptr}};
        int rc = LB_OK;
        for (auto& bf : bufs) {
            if ((e = hipMalloc(bf.p, std::max<size_t>(bf.n, 1))) != hipSuccess) {
                rc = qfail(LB_ENOMEM, "hipMalloc", e);
                break;
            }
            if (bf.src && bf.n && (e = hipMemcpyAsync(*bf.p, bf.src, bf.n, hipMemcpyHostToDevice, st)) != hipSuccess) {
                rc = qfail(LB_EHIP, "hipMemcpy", e);
                break;
            }
        }
        if (rc == LB_OK) {
            a.coords = (const int32_t*)dcoords;
            a.blocks = (const int32_t*)dblocks;
            a.tri_off = (const int64_t*)doff;
            a.tri_idx = (const int32_t*)didx;
            a.tris = (const double*)dtris;
            a.q = (uint16_t*)dq;
            a.tri = (int32_t*)dt;
            a.hit = (uint8_t*)dh;
            hipLaunchKernelGGL(qmap_kernel, dim3((unsigned)nc), dim3(512), 0, st, a);
            if ((e = hipGetLastError()) != hipSuccess) {
                rc = qfail(LB_EHIP, "qmap_kernel", e);
            } else {
                hq.resize(ncell * 27);
                ht.resize(ncell * 27);
                hh.resize(ncell);
                if ((e = hipMemcpyAsync(hq.data(), dq, hq.size() * 2, hipMemcpyDeviceToHost, st)) != hipSuccess ||
                    (e = hipMemcpyAsync(ht.data(), dt, ht.size() * 4, hipMemcpyDeviceToHost, st)) != hipSuccess ||
                    (e = hipMemcpyAsync(hh.data(), dh, hh.size(), hipMemcpyDeviceToHost, st)) != hipSuccess)
                    rc = qfail(LB_EHIP, "hipMemcpy D2H", e);
            }
        }
        if ((e = hipStreamSynchronize(st)) != hipSuccess && rc == LB_OK) rc = qfail(LB_EHIP, "qmap_kernel", e);
        for (auto& bf : bufs)
            if (*bf.p) (void)hipFree(*bf.p);
        (void)hipStreamDestroy(st);
        if (rc != LB_OK) return rc;
    }
    // compact in the canonical order: block, then lz, ly, lx
    int64_t n = 0;
    for (uint8_t h : hh) n += h;
    int32_t* co = (int32_t*)std::malloc((size_t)std::max<int64_t>(n, 1) * 4 * sizeof(int32_t));
    uint16_t* qo = (uint16_t*)std::malloc((size_t)std::max<int64_t>(n, 1) * 27 * sizeof(uint16_t));
    int32_t* to = (int32_t*)std::malloc((size_t)std::max<int64_t>(n, 1) * 27 * sizeof(int32_t));
    if (!co || !qo || !to) {
        std::free(co);
        std::free(qo);
        std::free(to);
        std::snprintf(g_err, sizeof g_err, "host allocation failed");
        return LB_ENOMEM;
    }
    int64_t at = 0;
    for (int64_t cb = 0; cb < nc; cb++)
        for (int cell = 0; cell < 512; cell++) {
            const size_t i = (size_t)cb * 512 + cell;
            if (!hh[i]) continue;
            co[at * 4 + 0] = blocks[(size_t)cb] + 1;
            co[at * 4 + 1] = (cell & 7) + 1;
            co[at * 4 + 2] = ((cell >> 3) & 7) + 1;
            co[at * 4 + 3] = (cell >> 6) + 1;
            std::memcpy(qo + at * 27, hq.data() + i * 27, 27 * sizeof(uint16_t));
            std::memcpy(to + at * 27, ht.data() + i * 27, 27 * sizeof(int32_t));
            at++;
        }
    *cells_out = co;
    *q_out = qo;
    *tri_out = to;
    *n_out = n;
    return LB_OK;
}


int lb_voxelize_gpu(int device, int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                    const double* tris, int64_t ntri, double dx, const double* offset, uint8_t* obst) {
    if (!coords || !grid || !tris || !offset || !obst || nb < 0 || ntri < 0) return LB_EINVAL;
    lbsync::SetupGuard guard;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return qfail(LB_EHIP, "hipSetDevice", e);
    Grid g{gdx, gdy, gdz, grid};
    auto bt = block_tri_map(nb, tris, ntri, g, dx, offset, 0);
    std::vector<int32_t> blocks, idx;
    std::vector<int64_t> off(1, 0);
    for (int64_t b = 0; b < nb; b++) {
        if (bt[(size_t)b].empty()) continue;
        blocks.push_back((int32_t)b);
        idx.insert(idx.end(), bt[(size_t)b].begin(), bt[(size_t)b].end());
        off.push_back((int64_t)idx.size());
    }
    const int64_t nc = (int64_t)blocks.size();
    if (nc == 0) return LB_OK;
    hipStream_t st = nullptr;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
        return qfail(LB_EHIP, "hipStreamCreate", e);
    void *dcoords = nullptr, *dblocks = nullptr, *doff = nullptr, *didx = nullptr, *dtris = nullptr, *dob = nullptr;
    struct Buf {
        void** p;
        size_t n;
        const void* src;
    } bufs[] = {{&dcoords, (size_t)nb * 3 * 4, coords}, {&dblocks, (size_t)nc * 4, blocks.data()},
                {&doff, off.size() * 8, off.data()},      {&didx, std::max<size_t>(idx.size(), 1) * 4, idx.data()},
                {&dtris, (size_t)ntri * 72, tris},        {&dob, (size_t)nb * 512, obst}};
    int rc = LB_OK;
    for (auto& bf : bufs) {
        if ((e = hipMalloc(bf.p, std::max<size_t>(bf.n, 1))) != hipSuccess) {
            rc = qfail(LB_ENOMEM, "hipMalloc", e);
            break;
        }
        if (bf.src && bf.n && (e = hipMemcpyAsync(*bf.p, bf.src, bf.n, hipMemcpyHostToDevice, st)) != hipSuccess) {
            rc = qfail(LB_EHIP, "hipMemcpy", e);
            break;
        }
    }
    if (rc == LB_OK) {
        VoxArgs a;
        a.coords = (const int32_t*)dcoords;
        a.blocks = (const int32_t*)dblocks;
        a.tri_off = (const int64_t*)doff;
        a.tri_idx = (const int32_t*)didx;
        a.tris = (const double*)dtris;
        a.ox = offset[0];
        a.oy = offset[1];
        a.oz = offset[2];
        a.dx = dx;
        a.obst = (uint8_t*)dob;
        hipLaunchKernelGGL(vox_kernel, dim3((unsigned)nc), dim3(512), 0, st, a);
        if ((e = hipGetLastError()) != hipSuccess)
            rc = qfail(LB_EHIP, "vox_kernel", e);
        else if ((e = hipMemcpyAsync(obst, dob, (size_t)nb * 512, hipMemcpyDeviceToHost, st)) != hipSuccess)
            rc = qfail(LB_EHIP, "hipMemcpy D2H", e);
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess && rc == LB_OK) rc = qfail(LB_EHIP, "vox_kernel", e);
    for (auto& bf : bufs)
        if (*bf.p) (void)hipFree(*bf.p);
    (void)hipStreamDestroy(st);
    return rc;
}

int lb_wall_distances_gpu(int device, int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy,
                          int gdz, const uint8_t* obst, double dx, float* wall_dist, int64_t* count_out) {
    if (!coords || !grid || !obst || !wall_dist || nb < 0) return LB_EINVAL;
    lbsync::SetupGuard guard;
    if (count_out) *count_out = 0;
    if (nb == 0) return LB_OK;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return qfail(LB_EHIP, "hipSetDevice", e);
    const size_t ng = (size_t)gdx * gdy * gdz;
    hipStream_t st = nullptr;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
        return qfail(LB_EHIP, "hipStreamCreate", e);
    void *dcoords = nullptr, *dgrid = nullptr, *dob = nullptr, *dsolid = nullptr, *dwall = nullptr, *dcnt = nullptr;
    struct Buf {
        void** p;
        size_t n;
        const void* src;
    } bufs[] = {{&dcoords, (size_t)nb * 3 * 4, coords}, {&dgrid, ng * 4, grid},
                {&dob, (size_t)nb * 512, obst},          {&dsolid, (size_t)nb, nullptr},
                {&dwall, (size_t)nb * 512 * 4, wall_dist}, {&dcnt, 8, nullptr}};
    int rc = LB_OK;
    for (auto& bf : bufs) {
        if ((e = hipMalloc(bf.p, std::max<size_t>(bf.n, 1))) != hipSuccess) {
            rc = qfail(LB_ENOMEM, "hipMalloc", e);
            break;
        }
        if (bf.src && bf.n && (e = hipMemcpyAsync(*bf.p, bf.src, bf.n, hipMemcpyHostToDevice, st)) != hipSuccess) {
            rc = qfail(LB_EHIP, "hipMemcpy", e);
            break;
        }
    }
    unsigned long long cnt = 0;
    if (rc == LB_OK && (e = hipMemsetAsync(dcnt, 0, 8, st)) != hipSuccess) rc = qfail(LB_EHIP, "hipMemset", e);
    if (rc == LB_OK) {
        hipLaunchKernelGGL(solid_kernel, dim3((unsigned)nb), dim3(512), 0, st, (const uint8_t*)dob, nb,
                           (uint8_t*)dsolid);
        WallArgs a;
        a.coords = (const int32_t*)dcoords;
        a.grid = (const int32_t*)dgrid;
        a.gdx = gdx;
        a.gdy = gdy;
        a.gdz = gdz;
        a.obst = (const uint8_t*)dob;
        a.solid = (const uint8_t*)dsolid;
        a.len[0] = 0.0f;
        for (int n = 1; n <= 3; n++) a.len[n] = std::sqrt((float)n);
        a.dxf = (float)dx;
        a.wall = (float*)dwall;
        a.count = (unsigned long long*)dcnt;
        hipLaunchKernelGGL(wall_kernel, dim3((unsigned)nb), dim3(512), 0, st, a);
        if ((e = hipGetLastError()) != hipSuccess)
            rc = qfail(LB_EHIP, "wall_kernel", e);
        else if ((e = hipMemcpyAsync(wall_dist, dwall, (size_t)nb * 512 * 4, hipMemcpyDeviceToHost, st)) !=
                     hipSuccess ||
                 (e = hipMemcpyAsync(&cnt, dcnt, 8, hipMemcpyDeviceToHost, st)) != hipSuccess)
            rc = qfail(LB_EHIP, "hipMemcpy D2H", e);
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess && rc == LB_OK) rc = qfail(LB_EHIP, "wall_kernel", e);
    for (auto& bf : bufs)
        if (*bf.p) (void)hipFree(*bf.p);
    (void)hipStreamDestroy(st);
    if (rc == LB_OK && count_out) *count_out = (int64_t)cnt;
    return rc;
}

}  // extern "C"
