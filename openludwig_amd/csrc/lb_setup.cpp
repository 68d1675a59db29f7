// lb_setup.cpp -- host-side producers of the time step's input tables.
//
// These are the per-cell setup loops that OpenLUDWIG runs on the CPU before
// the first step: STL reading and longest-edge subdivision, SAT voxelization,
// flood fill, sponge profile, wall distances and the Bouzidi Q-map.  They are
// restated here in C++ (OpenMP over blocks) because the tables they produce
// must match the reference bit for bit: obstacle flags, sponge/wall-distance
// values, Bouzidi q (FP16 bits), triangle ids and the boundary-cell list.
//
// All arrays follow the reference's layout (1-based block coordinates,
// column-major (8,8,8,nb) cell arrays).  Compiled with -ffp-contract=off: the
// reference's FP64 geometry is plain, uncontracted Julia arithmetic.
#include <algorithm>
#include <array>
#include <cctype>
#include <string>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "ludwig_setup.h"
#include "lb_geom.h"

namespace {

using namespace lbgeom;

thread_local char g_err[512];

}  // namespace

extern "C" {

const char* lbs_last_error(void) { return g_err; }

void lbs_free(void* p) { std::free(p); }

// load_mesh / parse_binary_stl / parse_ascii_stl: src/geometry/geometry.jl:385-520.
// Returns triangles as (n,3,3) doubles (vertex-major), winding-corrected.
int lbs_stl_read(const char* path, double scale, double** tris_out, int64_t* n_out) {
    *tris_out = nullptr;
    *n_out = 0;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        std::snprintf(g_err, sizeof g_err, "STL file not found: %s", path);
        return -1;
    }
    std::fseek(fp, 0, SEEK_END);
    long fsize = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    std::vector<unsigned char> buf((size_t)std::max(fsize, 0L));
    if (fsize > 0 && std::fread(buf.data(), 1, (size_t)fsize, fp) != (size_t)fsize) {
        std::fclose(fp);
        std::snprintf(g_err, sizeof g_err, "read failed: %s", path);
        return -1;
    }
    std::fclose(fp);
    bool binary = true;
    if (fsize < 84) {
        binary = false;
    } else {
        char hdr[6] = {0};
        std::memcpy(hdr, buf.data(), 5);
        for (int i = 0; i < 5; i++) hdr[i] = (char)std::tolower((unsigned char)hdr[i]);
        if (std::strncmp(hdr, "solid", 5) == 0) {
            uint32_t cnt;
            std::memcpy(&cnt, buf.data() + 80, 4);
            if ((uint64_t)fsize != 84ull + (uint64_t)cnt * 50ull) binary = false;
        }
    }
    std::vector<double> out;
    if (binary) {
        uint32_t cnt;
        std::memcpy(&cnt, buf.data() + 80, 4);
        if ((uint64_t)fsize < 84ull + (uint64_t)cnt * 50ull) {
            std::snprintf(g_err, sizeof g_err, "truncated binary STL: %s", path);
            return -1;
        }
        out.resize((size_t)cnt * 9);
        for (uint32_t i = 0; i < cnt; i++) {
            float r[12];
            std::memcpy(r, buf.data() + 84 + (size_t)i * 50, 48);
            double fn[3] = {(double)r[0], (double)r[1], (double)r[2]};
            double v[9];
            for (int j = 0; j < 9; j++) v[j] = (double)r[3 + j] * scale;
            double e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
            double e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
            double cx = e1y * e2z - e1z * e2y;
            double cy = e1z * e2x - e1x * e2z;
            double cz = e1x * e2y - e1y * e2x;
            double* o = &out[(size_t)i * 9];
            if ((cx * fn[0] + cy * fn[1] + cz * fn[2]) < 0.0) {
                double w[9] = {v[0], v[1], v[2], v[6], v[7], v[8], v[3], v[4], v[5]};
                std::memcpy(o, w, sizeof w);
            } else {
                std::memcpy(o, v, sizeof v);
            }
        }
    } else {
        // ASCII: "vertex x y z" lines grouped by "endloop"
        std::string text(buf.begin(), buf.end());
        std::vector<double> cur;
        size_t pos = 0;
        while (pos < text.size()) {
            size_t e = text.find('\n', pos);
            if (e == std::string::npos) e = text.size();
            std::string line = text.substr(pos, e - pos);
            pos = e + 1;
            size_t a = line.find_first_not_of(" \t\r");
            if (a == std::string::npos) continue;
            line = line.substr(a);
            if (line.compare(0, 6, "vertex") == 0) {
                double x, y, z;
                if (std::sscanf(line.c_str() + 6, "%lf %lf %lf", &x, &y, &z) == 3) {
                    cur.push_back(x * scale); cur.push_back(y * scale); cur.push_back(z * scale);
                }
            } else if (line.compare(0, 7, "endloop") == 0) {
                if (cur.size() == 9) out.insert(out.end(), cur.begin(), cur.end());
                cur.clear();
            }
        }
    }
    if (out.empty()) {
        std::snprintf(g_err, sizeof g_err, "No triangles loaded.");
        return -2;
    }
    double* r = (double*)std::malloc(out.size() * sizeof(double));
    std::memcpy(r, out.data(), out.size() * sizeof(double));
    *tris_out = r;
    *n_out = (int64_t)(out.size() / 9);
    return 0;
}

// compute_geometry_properties: src/geometry/geometry.jl:75-103
int lbs_geometry_properties(const double* tris, int64_t n, double* normals, double* areas, double* centers) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        V3 v1 = vtx(tris, i, 0), v2 = vtx(tris, i, 1), v3 = vtx(tris, i, 2);
        V3 cp = cross(sub(v2, v1), sub(v3, v1));
        double area = 0.5 * norm(cp);
        V3 nn = {0.0, 0.0, 0.0};
        if (area > 1e-12) nn = {cp.x / (2.0 * area), cp.y / (2.0 * area), cp.z / (2.0 * area)};
        normals[i * 3 + 0] = nn.x; normals[i * 3 + 1] = nn.y; normals[i * 3 + 2] = nn.z;
        areas[i] = area;
        V3 s = add(add(v1, v2), v3);
        centers[i * 3 + 0] = s.x / 3.0; centers[i * 3 + 1] = s.y / 3.0; centers[i * 3 + 2] = s.z / 3.0;
    }
    return 0;
}

// subdivide_triangle / subdivide_mesh: src/geometry/geometry.jl:120-252.
// Output triangles keep the reference's order: parents in order, children in
// the LIFO worklist order.  parent_index_out[i] = parent triangle (0-based), so
// the caller can inherit normals.  Areas/centres are recomputed by the caller.
int lbs_subdivide(const double* tris, int64_t n, double min_size, double** tris_out, int64_t** parent_out,
                  int64_t* n_out) {
    std::vector<double> out;
    std::vector<int64_t> par;
    const int64_t MAX_TRIANGLES = 50000000;
    std::vector<std::array<V3, 3>> work;
    for (int64_t i = 0; i < n; i++) {
        work.clear();
        work.push_back({vtx(tris, i, 0), vtx(tris, i, 1), vtx(tris, i, 2)});
        while (!work.empty()) {
            auto t = work.back();
            work.pop_back();
            V3 v1 = t[0], v2 = t[1], v3 = t[2];
            double e1 = norm(sub(v2, v1)), e2 = norm(sub(v3, v2)), e3 = norm(sub(v1, v3));
            double longest = std::max(e1, std::max(e2, e3));
            if (longest <= min_size) {
                for (int k = 0; k < 3; k++) {
                    out.push_back(t[k].x); out.push_back(t[k].y); out.push_back(t[k].z);
                }
                par.push_back(i);
            } else if (e1 >= e2 && e1 >= e3) {
                V3 m = {(v1.x + v2.x) / 2.0, (v1.y + v2.y) / 2.0, (v1.z + v2.z) / 2.0};
                work.push_back({t[0], m, t[2]});
                work.push_back({m, t[1], t[2]});
            } else if (e2 >= e1 && e2 >= e3) {
                V3 m = {(v2.x + v3.x) / 2.0, (v2.y + v3.y) / 2.0, (v2.z + v3.z) / 2.0};
                work.push_back({t[1], m, t[0]});
                work.push_back({m, t[2], t[0]});
            } else {
                V3 m = {(v3.x + v1.x) / 2.0, (v3.y + v1.y) / 2.0, (v3.z + v1.z) / 2.0};
                work.push_back({t[2], m, t[1]});
                work.push_back({m, t[0], t[1]});
            }
        }
        if ((int64_t)par.size() > MAX_TRIANGLES) {
            std::snprintf(g_err, sizeof g_err, "Subdivision exceeded %lld triangles", (long long)MAX_TRIANGLES);
            return -3;
        }
    }
    double* r = (double*)std::malloc(std::max<size_t>(out.size(), 1) * sizeof(double));
    int64_t* p = (int64_t*)std::malloc(std::max<size_t>(par.size(), 1) * sizeof(int64_t));
    std::memcpy(r, out.data(), out.size() * sizeof(double));
    std::memcpy(p, par.data(), par.size() * sizeof(int64_t));
    *tris_out = r;
    *parent_out = p;
    *n_out = (int64_t)par.size();
    return 0;
}

// voxelize_blocks!: src/geometry/domain_generation.jl:74-112.
// coords: (nb,3) int32 1-based (bx,by,bz) row-major; grid: dense block pointer
// of the level dims (gdx,gdy,gdz), column-major 1-based ids.  obst: uint8 (512*nb).
int lbs_voxelize(int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                 const double* tris, int64_t ntri, double dx, const double* offset, uint8_t* obst) {
    Grid g{gdx, gdy, gdz, grid};
    auto bt = block_tri_map(nb, tris, ntri, g, dx, offset, 0);
    double hb = 0.75 * dx;
    V3 o = {offset[0], offset[1], offset[2]};
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < nb; i++) {
        const auto& rel = bt[(size_t)i];
        if (rel.empty()) continue;
        int64_t bx = coords[i * 3], by = coords[i * 3 + 1], bz = coords[i * 3 + 2];
        for (int lz = 1; lz <= BS; lz++)
            for (int ly = 1; ly <= BS; ly++)
                for (int lx = 1; lx <= BS; lx++) {
                    V3 c = {((double)((bx - 1) * BS + lx) - 0.5) * dx, ((double)((by - 1) * BS + ly) - 0.5) * dx,
                            ((double)((bz - 1) * BS + lz) - 0.5) * dx};
                    bool shell = false;
                    for (int32_t tid : rel) {
                        V3 v1 = add(vtx(tris, tid, 0), o), v2 = add(vtx(tris, tid, 1), o),
                           v3 = add(vtx(tris, tid, 2), o);
                        if (tri_box(c, hb, v1, v2, v3)) {
                            shell = true;
                            break;
                        }
                    }
                    if (shell) obst[(lx - 1) + 8 * (ly - 1) + 64 * (lz - 1) + 512 * i] = 1;
                }
    }
    return 0;
}

// perform_flood_fill!: src/geometry/domain_generation.jl:114-203.  Cells not
// reachable (6-connectivity) from the non-solid cells of the min-bx blocks
// become solid.  Returns the number of filled cells.
// perform_flood_fill! (domain_generation.jl:114-203): the fluid cells 6-connected
// to the fluid cells of the blocks with the smallest bx are kept; every other
// fluid cell becomes solid.  The result is a set, so it is computed in
// parallel: (1) per block, 6-connected components of its fluid cells; (2) a
// graph of components joined across block faces; (3) reachability from the
// seed components over that graph; (4) unreached fluid cells are filled.
int64_t lbs_flood_fill(int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                       uint8_t* obst) {
    if (nb == 0) return 0;
    Grid g{gdx, gdy, gdz, grid};
    int32_t minbx = coords[0];
    for (int64_t i = 0; i < nb; i++) minbx = std::min(minbx, coords[i * 3]);
    // (1) components inside each block; comp[cell] = local id, -1 for solid
    std::vector<int16_t> comp((size_t)nb * 512, -1);
    std::vector<int32_t> ncomp((size_t)nb, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t b = 0; b < nb; b++) {
        const uint8_t* o = obst + b * 512;
        int16_t* cp = comp.data() + b * 512;
        int16_t next = 0;
        int16_t stack[512];
        for (int c0 = 0; c0 < 512; c0++) {
            if (o[c0] || cp[c0] >= 0) continue;
            int sp = 0;
            stack[sp++] = (int16_t)c0;
            cp[c0] = next;
            while (sp > 0) {
                const int c = stack[--sp];
                const int x = c & 7, y = (c >> 3) & 7, z = c >> 6;
                const int nbr[6] = {x < 7 ? c + 1 : -1, x > 0 ? c - 1 : -1, y < 7 ? c + 8 : -1,
                                    y > 0 ? c - 8 : -1, z < 7 ? c + 64 : -1, z > 0 ? c - 64 : -1};
                for (int n : nbr)
                    if (n >= 0 && !o[n] && cp[n] < 0) {
                        cp[n] = next;
                        stack[sp++] = (int16_t)n;
                    }
            }
            next++;
        }
        ncomp[(size_t)b] = next;
    }
    std::vector<int64_t> base((size_t)nb + 1, 0);
    for (int64_t b = 0; b < nb; b++) base[(size_t)b + 1] = base[(size_t)b] + ncomp[(size_t)b];
    const int64_t nc = base[(size_t)nb];
    // (2) component graph: edges across the +x, +y, +z faces of every block
    std::vector<std::vector<std::pair<int64_t, int64_t>>> edges_t;
#pragma omp parallel
    {
#pragma omp single
        {
#ifdef _OPENMP
            edges_t.resize((size_t)omp_get_num_threads());
#else
            edges_t.resize(1);
#endif
        }
#ifdef _OPENMP
        auto& my = edges_t[(size_t)omp_get_thread_num()];
#else
        auto& my = edges_t[0];
#endif
#pragma omp for schedule(dynamic, 64)
        for (int64_t b = 0; b < nb; b++) {
            const int64_t bx = coords[b * 3], by = coords[b * 3 + 1], bz = coords[b * 3 + 2];
            for (int axis = 0; axis < 3; axis++) {
                const int32_t nbi = g.at(bx + (axis == 0), by + (axis == 1), bz + (axis == 2));
                if (nbi <= 0) continue;
                const int64_t nb1 = nbi - 1;
                int64_t last_a = -1, last_b = -1;
                for (int u = 0; u < 8; u++)
                    for (int v = 0; v < 8; v++) {
                        // cell on this block's + face and the facing cell of the neighbour
                        int ca, cb;
                        if (axis == 0) { ca = 7 + 8 * u + 64 * v; cb = 0 + 8 * u + 64 * v; }
                        else if (axis == 1) { ca = u + 8 * 7 + 64 * v; cb = u + 0 + 64 * v; }
                        else { ca = u + 8 * v + 64 * 7; cb = u + 8 * v; }
                        const int16_t pa = comp[(size_t)b * 512 + ca], pb = comp[(size_t)nb1 * 512 + cb];
                        if (pa < 0 || pb < 0) continue;
                        const int64_t ga = base[(size_t)b] + pa, gb = base[(size_t)nb1] + pb;
                        if (ga == last_a && gb == last_b) continue;
                        last_a = ga;
                        last_b = gb;
                        my.emplace_back(ga, gb);
                    }
            }
        }
    }
    std::vector<int64_t> deg((size_t)nc + 1, 0);
    for (auto& v : edges_t)
        for (auto& e : v) {
            deg[(size_t)e.first + 1]++;
            deg[(size_t)e.second + 1]++;
        }
    for (int64_t i = 0; i < nc; i++) deg[(size_t)i + 1] += deg[(size_t)i];
    std::vector<int64_t> adj((size_t)deg[(size_t)nc]);
    {
        std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
        for (auto& v : edges_t)
            for (auto& e : v) {
                adj[(size_t)fill[(size_t)e.first]++] = e.second;
                adj[(size_t)fill[(size_t)e.second]++] = e.first;
            }
    }
    // (3) reachability from the seed components (blocks with the smallest bx)
    std::vector<uint8_t> reach((size_t)nc, 0);
    std::vector<int64_t> q;
    for (int64_t b = 0; b < nb; b++)
        if (coords[b * 3] == minbx)
            for (int32_t k = 0; k < ncomp[(size_t)b]; k++) {
                reach[(size_t)(base[(size_t)b] + k)] = 1;
                q.push_back(base[(size_t)b] + k);
            }
    for (size_t h = 0; h < q.size(); h++) {
        const int64_t u = q[h];
        for (int64_t i = deg[(size_t)u]; i < deg[(size_t)u + 1]; i++) {
            const int64_t v = adj[(size_t)i];
            if (!reach[(size_t)v]) {
                reach[(size_t)v] = 1;
                q.push_back(v);
            }
        }
    }
    // (4) fill the unreached fluid cells
    int64_t filled = 0;
#pragma omp parallel for schedule(static) reduction(+ : filled)
    for (int64_t b = 0; b < nb; b++)
        for (int c = 0; c < 512; c++) {
            const int16_t p = comp[(size_t)b * 512 + c];
            if (p >= 0 && !reach[(size_t)(base[(size_t)b] + p)]) {
                obst[b * 512 + c] = 1;
                filled++;
            }
        }
    return filled;
}

// smooth_sponge_profile / apply_sponge!: src/geometry/domain_generation.jl:205-289
static double sponge_profile(double x, double th) {
    if (x <= 0.0) return 1.0;
    if (x >= th) return 0.0;
    return 0.5 * (1.0 + std::cos(M_PI * x / th));
}

int lbs_sponge(int64_t nb, const int32_t* coords, double dx, double Lx, double Ly, double Lz, double sponge_thickness,
               int symmetric, float* sponge) {
    double outlet_th = Lx * std::max(sponge_thickness, 0.08);
    double inlet_th = Lx * 0.02;
    double y_th = Ly * sponge_thickness * 0.5;
    double z_th = Lz * sponge_thickness * 0.5;
    double outlet_start = Lx - outlet_th, y_top = Ly - y_th, z_back = Lz - z_th;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; i++) {
        int64_t bx = coords[i * 3], by = coords[i * 3 + 1], bz = coords[i * 3 + 2];
        for (int lz = 1; lz <= BS; lz++)
            for (int ly = 1; ly <= BS; ly++)
                for (int lx = 1; lx <= BS; lx++) {
                    double px = ((double)((bx - 1) * BS + lx) - 0.5) * dx;
                    double py = ((double)((by - 1) * BS + ly) - 0.5) * dx;
                    double pz = ((double)((bz - 1) * BS + lz) - 0.5) * dx;
                    double s = 0.0;
                    if (px > outlet_start) s = std::max(s, sponge_profile(outlet_th - (px - outlet_start), outlet_th) * 1.0);
                    if (px < inlet_th) s = std::max(s, sponge_profile(px, inlet_th) * 0.05);
                    if (!symmetric && py < y_th) s = std::max(s, sponge_profile(py, y_th) * 0.1);
                    if (py > y_top) s = std::max(s, sponge_profile(y_th - (py - y_top), y_th) * 0.1);
                    if (pz < z_th) s = std::max(s, sponge_profile(pz, z_th) * 0.1);
                    if (pz > z_back) s = std::max(s, sponge_profile(z_th - (pz - z_back), z_th) * 0.1);
                    sponge[(lx - 1) + 8 * (ly - 1) + 64 * (lz - 1) + 512 * i] = (float)s;
                }
    }
    return 0;
}

// compute_wall_distances!: src/geometry/domain_generation.jl:371-434.  The
// block lookup (a Dict in the reference) is the dense grid here; blocks are
// only ever looked up at existing coordinates, so the results are identical.
int64_t lbs_wall_distances(int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                           const uint8_t* obst, double dx, float* wall_dist) {
    Grid g{gdx, gdy, gdz, grid};
    float dxf = (float)dx;
    int64_t count = 0;
    // a block whose own cells and 26 neighbour blocks hold no solid cell has
    // no near-wall cell: skip it
    std::vector<uint8_t> solid((size_t)nb, 0);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; b++)
        for (int c = 0; c < 512; c++)
            if (obst[b * 512 + c]) {
                solid[(size_t)b] = 1;
                break;
            }
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int64_t b = 0; b < nb; b++) {
        int64_t bx = coords[b * 3], by = coords[b * 3 + 1], bz = coords[b * 3 + 2];
        bool any = false;
        for (int dz = -1; dz <= 1 && !any; dz++)
            for (int dy = -1; dy <= 1 && !any; dy++)
                for (int dxo = -1; dxo <= 1 && !any; dxo++) {
                    const int32_t id = (dxo == 0 && dy == 0 && dz == 0) ? (int32_t)(b + 1) : g.at(bx + dxo, by + dy, bz + dz);
                    if (id > 0 && solid[(size_t)id - 1]) any = true;
                }
        if (!any) continue;
        for (int lz = 1; lz <= BS; lz++)
            for (int ly = 1; ly <= BS; ly++)
                for (int lx = 1; lx <= BS; lx++) {
                    if (obst[b * 512 + (lx - 1) + 8 * (ly - 1) + 64 * (lz - 1)]) continue;
                    bool near = false;
                    float mind = 100.0f;
                    for (int dz = -1; dz <= 1; dz++)
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dxo = -1; dxo <= 1; dxo++) {
                                if (dxo == 0 && dy == 0 && dz == 0) continue;
                                int nx = lx + dxo, ny = ly + dy, nz = lz + dz;
                                int64_t nbi = b;
                                if (nx < 1 || nx > BS || ny < 1 || ny > BS || nz < 1 || nz > BS) {
                                    int64_t nbx = bx + (nx < 1 ? -1 : (nx > BS ? 1 : 0));
                                    int64_t nby = by + (ny < 1 ? -1 : (ny > BS ? 1 : 0));
                                    int64_t nbz = bz + (nz < 1 ? -1 : (nz > BS ? 1 : 0));
                                    int32_t id = g.at(nbx, nby, nbz);
                                    if (id <= 0) continue;
                                    nbi = id - 1;
                                    nx = nx < 1 ? nx + BS : (nx > BS ? nx - BS : nx);
                                    ny = ny < 1 ? ny + BS : (ny > BS ? ny - BS : ny);
                                    nz = nz < 1 ? nz + BS : (nz > BS ? nz - BS : nz);
                                }
                                if (obst[nbi * 512 + (nx - 1) + 8 * (ny - 1) + 64 * (nz - 1)]) {
                                    near = true;
                                    float d = std::sqrt((float)(dxo * dxo + dy * dy + dz * dz)) * dxf;
                                    mind = std::min(mind, d);
                                }
                            }
                    if (near) {
                        wall_dist[b * 512 + (lx - 1) + 8 * (ly - 1) + 64 * (lz - 1)] = mind;
                        count++;
                    }
                }
    }
    return count;
}

// compute_bouzidi_qmap_sparse / compute_q_for_cell:
// src/geometry/bouzidi_setup.jl:64-167, src/geometry/bouzidi_math.jl:53-102.
// Output (malloc'd, free with lbs_free): boundary cells in canonical order
// (block, lz, ly, lx) -- the reference's single-thread order -- as
// cells[n*4] = (b, lx, ly, lz) 1-based, q_bits[n*27] (FP16), tri[n*27] (1-based
// triangle id, 0 = none).
int lbs_bouzidi_qmap(int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                     const double* tris, int64_t ntri, double dx, const double* offset, int32_t** cells_out,
                     uint16_t** q_out, int32_t** tri_out, int64_t* n_out) {
    Grid g{gdx, gdy, gdz, grid};
    auto bt = block_tri_map(nb, tris, ntri, g, dx, offset, 1);
    V3 o = {offset[0], offset[1], offset[2]};
    double dn[27][3], cmag[27];
    int cxs[27], cys[27], czs[27];
    {
        int k = 0;
        for (int z = -1; z <= 1; z++)
            for (int y = -1; y <= 1; y++)
                for (int x = -1; x <= 1; x++, k++) {
                    cxs[k] = x; cys[k] = y; czs[k] = z;
                    double m = std::sqrt((double)(x * x) + (double)(y * y) + (double)(z * z));
                    cmag[k] = m;
                    dn[k][0] = m > 0 ? x / m : 0; dn[k][1] = m > 0 ? y / m : 0; dn[k][2] = m > 0 ? z / m : 0;
                }
    }
    std::vector<std::vector<int32_t>> cells((size_t)nb);
    std::vector<std::vector<uint16_t>> qs((size_t)nb);
    std::vector<std::vector<int32_t>> ts((size_t)nb);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t b = 0; b < nb; b++) {
        const auto& rel = bt[(size_t)b];
        if (rel.empty()) continue;
        std::vector<std::array<V3, 3>> loc(rel.size());
        for (size_t i = 0; i < rel.size(); i++)
            loc[i] = {add(vtx(tris, rel[i], 0), o), add(vtx(tris, rel[i], 1), o), add(vtx(tris, rel[i], 2), o)};
        int64_t bx = coords[b * 3], by = coords[b * 3 + 1], bz = coords[b * 3 + 2];
        for (int lz = 1; lz <= BS; lz++)
            for (int ly = 1; ly <= BS; ly++)
                for (int lx = 1; lx <= BS; lx++) {
                    V3 c = {((double)((bx - 1) * BS + lx) - 0.5) * dx, ((double)((by - 1) * BS + ly) - 0.5) * dx,
                            ((double)((bz - 1) * BS + lz) - 0.5) * dx};
                    uint16_t qb[27];
                    int32_t tb[27];
                    bool any = false;
                    for (int k = 0; k < 27; k++) {
                        qb[k] = 0;
                        tb[k] = 0;
                        if (cxs[k] == 0 && cys[k] == 0 && czs[k] == 0) continue;
                        V3 d = {dn[k][0], dn[k][1], dn[k][2]};
                        double min_t = std::numeric_limits<double>::infinity();
                        int64_t best = -1;
                        for (size_t i = 0; i < loc.size(); i++) {
                            double t;
                            if (ray_tri(c, d, loc[i][0], loc[i][1], loc[i][2], &t) && t < min_t) {
                                min_t = t;
                                best = rel[i];
                            }
                        }
                        if (min_t < std::numeric_limits<double>::infinity()) {
                            double q = min_t / (dx * cmag[k]);
                            if (q > 0.0 && q <= 1.0) {
                                qb[k] = double_to_half(q);
                                tb[k] = (int32_t)(best + 1);
                                any = true;
                            }
                        }
                    }
                    if (any) {
                        cells[(size_t)b].push_back((int32_t)(b + 1));
                        cells[(size_t)b].push_back(lx);
                        cells[(size_t)b].push_back(ly);
                        cells[(size_t)b].push_back(lz);
                        qs[(size_t)b].insert(qs[(size_t)b].end(), qb, qb + 27);
                        ts[(size_t)b].insert(ts[(size_t)b].end(), tb, tb + 27);
                    }
                }
    }
    int64_t n = 0;
    for (auto& v : cells) n += (int64_t)v.size() / 4;
    int32_t* co = (int32_t*)std::malloc((size_t)std::max<int64_t>(n, 1) * 4 * sizeof(int32_t));
    uint16_t* qo = (uint16_t*)std::malloc((size_t)std::max<int64_t>(n, 1) * 27 * sizeof(uint16_t));
    int32_t* to = (int32_t*)std::malloc((size_t)std::max<int64_t>(n, 1) * 27 * sizeof(int32_t));
    int64_t at = 0;
    for (int64_t b = 0; b < nb; b++) {
        size_t m = cells[(size_t)b].size() / 4;
        if (!m) continue;
        std::memcpy(co + at * 4, cells[(size_t)b].data(), m * 4 * sizeof(int32_t));
        std::memcpy(qo + at * 27, qs[(size_t)b].data(), m * 27 * sizeof(uint16_t));
        std::memcpy(to + at * 27, ts[(size_t)b].data(), m * 27 * sizeof(int32_t));
        at += (int64_t)m;
    }
    *cells_out = co;
    *q_out = qo;
    *tri_out = to;
    *n_out = n;
    return 0;
}

uint16_t lbs_double_to_half(double d) { return double_to_half(d); }

}  // extern "C"
