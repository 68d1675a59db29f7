// lb_geom.h -- FP64 geometry shared by the CPU setup library (lb_setup.cpp,
// g++) and its GPU counterpart (lb_setup_gpu.hip, hipcc): vectors, the
// Moller-Trumbore ray-triangle test, the triangle-box SAT test, Float64 ->
// Float16 rounding and the block -> candidate-triangle map.  Both compile it without FMA contraction,
// so the two produce the same bits.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#if defined(__HIPCC__)
#define LB_HD __host__ __device__
#else
#define LB_HD
#endif

namespace lbgeom {

constexpr int BS = 8;

struct V3 {
    double x, y, z;
};
LB_HD inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LB_HD inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LB_HD inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LB_HD inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LB_HD inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

LB_HD inline V3 vtx(const double* tris, int64_t t, int v) {
    const double* p = tris + t * 9 + v * 3;
    return {p[0], p[1], p[2]};
}

// Julia floor(Int, x) on a finite double.
inline int64_t ifloor(double x) { return (int64_t)std::floor(x); }

// Round a double to IEEE binary16 bits, round-to-nearest-even, directly from
// the double (Julia's Float16(::Float64), src/geometry/bouzidi_setup.jl:128).
LB_HD inline uint16_t double_to_half(double d) {
    uint64_t u;
    __builtin_memcpy(&u, &d, 8);
    uint16_t sign = (uint16_t)((u >> 48) & 0x8000u);
    uint64_t abs = u & 0x7fffffffffffffffull;
    if (abs >= 0x7ff0000000000000ull) {  // inf / nan
        if (abs > 0x7ff0000000000000ull) return sign | 0x7e00u;
        return sign | 0x7c00u;
    }
    int exp = (int)(abs >> 52) - 1023;
    uint64_t mant = abs & 0xfffffffffffffull;
    if (exp > 15) return sign | 0x7c00u;  // overflow -> inf (value >= 65520 rounds up)
    if (exp >= -14) {
        // normal half: 10 mantissa bits; drop 42 bits with RNE
        uint64_t m = mant;
        uint64_t keep = m >> 42;
        uint64_t rem = m & ((1ull << 42) - 1);
        uint64_t half = 1ull << 41;
        uint32_t hexp = (uint32_t)(exp + 15);
        uint32_t hm = (uint32_t)keep;
        if (rem > half || (rem == half && (hm & 1u))) {
            hm++;
            if (hm == 1024u) { hm = 0; hexp++; }
        }
        if (hexp >= 31) return sign | 0x7c00u;
        return sign | (uint16_t)((hexp << 10) | hm);
    }
    // subnormal half: value = m * 2^-24, m in [0, 1023]
    if (exp < -25) {
        // below half of the smallest subnormal (2^-25): rounds to 0 unless > 2^-25
        return sign;
    }
    // full significand with implicit 1: 53 bits, value = sig * 2^(exp-52)
    uint64_t sig = mant | (1ull << 52);
    int shift = (52 - exp) - 24;  // bits to drop so that result is in units of 2^-24
    uint64_t keep = shift >= 64 ? 0 : (sig >> shift);
    uint64_t rem = shift >= 64 ? sig : (sig & ((1ull << shift) - 1));
    uint64_t half = 1ull << (shift - 1);
    if (rem > half || (rem == half && (keep & 1u))) keep++;
    return sign | (uint16_t)keep;  // keep may reach 1024 = smallest normal, encoding still right
}

// ray_triangle_intersection (Moller-Trumbore): src/geometry/bouzidi_math.jl:9-47
LB_HD inline bool ray_tri(V3 o, V3 d, V3 v1, V3 v2, V3 v3, double* t_out) {
    const double EPS = 1e-9;
    V3 e1 = sub(v2, v1), e2 = sub(v3, v1);
    V3 h = cross(d, e2);
    double a = dot(e1, h);
    if (fabs(a) < EPS) return false;
    double f = 1.0 / a;
    V3 s = sub(o, v1);
    double u = f * dot(s, h);
    if (u < 0.0 || u > 1.0) return false;
    V3 q = cross(s, e1);
    double v = f * dot(d, q);
    if (v < 0.0 || u + v > 1.0) return false;
    double t = f * dot(e2, q);
    if (t > EPS) {
        *t_out = t;
        return true;
    }
    return false;
}

LB_HD inline double min3(double a, double b, double c) { return a < b ? (a < c ? a : c) : (b < c ? b : c); }
LB_HD inline double max3(double a, double b, double c) { return a > b ? (a > c ? a : c) : (b > c ? b : c); }

// triangle_intersects_aabb (src/geometry/domain_generation.jl:10-32): the
// separating-axis test of a triangle against the box of half size hb (grown
// by 0.1 %) centred at c; 3 box axes, then the 9 edge cross products (the
// triangle normal axis is not tested, as in the reference).
LB_HD inline bool tri_box(V3 c, double hb, V3 v1, V3 v2, V3 v3) {
    const double h = hb * 1.001;
    const V3 t1 = sub(v1, c), t2 = sub(v2, c), t3 = sub(v3, c);
    if (min3(t1.x, t2.x, t3.x) > h || max3(t1.x, t2.x, t3.x) < -h) return false;
    if (min3(t1.y, t2.y, t3.y) > h || max3(t1.y, t2.y, t3.y) < -h) return false;
    if (min3(t1.z, t2.z, t3.z) > h || max3(t1.z, t2.z, t3.z) < -h) return false;
    const V3 f[3] = {sub(t2, t1), sub(t3, t2), sub(t1, t3)};
    const V3 u[3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const V3 axis = cross(u[i], f[j]);
            if (dot(axis, axis) < 1e-10) continue;
            const double p1 = dot(t1, axis), p2 = dot(t2, axis), p3 = dot(t3, axis);
            const double r = h * std::fabs(axis.x) + h * std::fabs(axis.y) + h * std::fabs(axis.z);
            if (min3(p1, p2, p3) > r || max3(p1, p2, p3) < -r) return false;
        }
    return true;
}

// Dense 1-based block pointer grid helper.
struct Grid {
    int dx, dy, dz;
    const int32_t* p;
    int32_t at(int64_t bx, int64_t by, int64_t bz) const {
        if (bx < 1 || bx > dx || by < 1 || by > dy || bz < 1 || bz > dz) return 0;
        return p[(bx - 1) + (int64_t)dx * ((by - 1) + (int64_t)dy * (bz - 1))];
    }
};

// Block -> triangle lists, triangles in index order (domain_generation.jl:34-72,
// bouzidi_setup.jl:12-54).  `style` 0: voxelize form (min over p+offset, margin
// 2dx); 1: Bouzidi form (min of raw vertices, then +offset, margin 2.5dx).
inline std::vector<std::vector<int32_t>> block_tri_map(int64_t nb, const double* tris, int64_t ntri, const Grid& g,
                                                double dx, const double* off, int style) {
    std::vector<std::vector<int32_t>> bt((size_t)nb);
    double margin = style == 0 ? dx * 2 : dx * 2.5;
    double bsdx = BS * dx;
    for (int64_t t = 0; t < ntri; t++) {
        V3 v[3] = {vtx(tris, t, 0), vtx(tris, t, 1), vtx(tris, t, 2)};
        double mn[3], mx[3];
        if (style == 0) {
            V3 o = {off[0], off[1], off[2]};
            V3 p0 = add(v[0], o);
            mn[0] = mx[0] = p0.x; mn[1] = mx[1] = p0.y; mn[2] = mx[2] = p0.z;
            for (int k = 0; k < 3; k++) {
                V3 pm = add(v[k], o);
                mn[0] = std::min(mn[0], pm.x); mx[0] = std::max(mx[0], pm.x);
                mn[1] = std::min(mn[1], pm.y); mx[1] = std::max(mx[1], pm.y);
                mn[2] = std::min(mn[2], pm.z); mx[2] = std::max(mx[2], pm.z);
            }
        } else {
            double a[3] = {v[0].x, v[0].y, v[0].z}, b[3] = {v[1].x, v[1].y, v[1].z}, c[3] = {v[2].x, v[2].y, v[2].z};
            for (int k = 0; k < 3; k++) {
                mn[k] = std::min(a[k], std::min(b[k], c[k])) + off[k];
                mx[k] = std::max(a[k], std::max(b[k], c[k])) + off[k];
            }
        }
        int64_t lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            lo[k] = ifloor((mn[k] - margin) / bsdx) + 1;
            hi[k] = ifloor((mx[k] + margin) / bsdx) + 1;
        }
        for (int64_t bz = std::max<int64_t>(1, lo[2]); bz <= hi[2]; bz++)
            for (int64_t by = std::max<int64_t>(1, lo[1]); by <= hi[1]; by++)
                for (int64_t bx = std::max<int64_t>(1, lo[0]); bx <= hi[0]; bx++) {
                    int32_t id = g.at(bx, by, bz);
                    if (id > 0 && id <= nb) bt[(size_t)id - 1].push_back((int32_t)t);
                }
    }
    return bt;
}


}  // namespace lbgeom
