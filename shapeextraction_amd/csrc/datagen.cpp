// Synthetic point-cloud generators for the normal-estimation workloads.
//
// The reference ships no data (its default input, src/main.cu:52, lies outside
// the repository), so every workload is generated here from a counter-based RNG:
// point i of a cloud draws only from stream (seed, i), which makes the arrays
// independent of thread count and identical on every machine running this
// library.  All arithmetic is double precision, rounded to float32 once at the
// end; the translation unit is compiled with -ffp-contract=off.
//
// Workloads:
//   sx_gen_sphere      100k points on the unit sphere + isotropic noise
//   sx_gen_scene       200 random primitives (planes, spheres, cylinders) in a
//                      [0,box]^3 scene, noise along the surface normal, uniform
//                      outliers
//   sx_gen_gmm_scene   the same surfaces with density modulated by an
//                      anisotropic Gaussian mixture (bounded max/min density)
//   sx_gen_batch_sizes log-uniform cloud sizes for the batched workload
//   sx_gen_degenerate  lines, corners and thin slabs translated by an offset
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

#include "sx_status.hpp"

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Counter-based stream: value j of stream (seed, stream) is a pure function.
struct Stream {
    uint64_t key;
    uint64_t ctr = 0;
    Stream(uint64_t seed, uint64_t stream)
        : key(mix64(mix64(seed) ^ (stream * 0xD1342543DE82EF95ull + 0x632BE59BD9B4E019ull))) {}
    uint64_t next() { return mix64(key ^ mix64(ctr++)); }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        // Box-Muller; u1 in (0,1].
        double u1 = 1.0 - uniform();
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

struct Vec3 { double x, y, z; };
inline Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 scale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 cross(Vec3 a, Vec3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline Vec3 normalize(Vec3 a) {
    double n = std::sqrt(dot(a, a));
    return n > 0 ? scale(a, 1.0 / n) : Vec3{0, 0, 1};
}

struct Rot { double m[9]; };

// Uniform random rotation from a normalised 4-D Gaussian (a uniform unit quaternion).
Rot random_rotation(Stream& r) {
    double w = r.normal(), x = r.normal(), y = r.normal(), z = r.normal();
    double n = std::sqrt(w * w + x * x + y * y + z * z);
    w /= n; x /= n; y /= n; z /= n;
    Rot R;
    R.m[0] = 1 - 2 * (y * y + z * z); R.m[1] = 2 * (x * y - w * z);     R.m[2] = 2 * (x * z + w * y);
    R.m[3] = 2 * (x * y + w * z);     R.m[4] = 1 - 2 * (x * x + z * z); R.m[5] = 2 * (y * z - w * x);
    R.m[6] = 2 * (x * z - w * y);     R.m[7] = 2 * (y * z + w * x);     R.m[8] = 1 - 2 * (x * x + y * y);
    return R;
}
inline Vec3 apply(const Rot& R, Vec3 v) {
    return {R.m[0] * v.x + R.m[1] * v.y + R.m[2] * v.z,
            R.m[3] * v.x + R.m[4] * v.y + R.m[5] * v.z,
            R.m[6] * v.x + R.m[7] * v.y + R.m[8] * v.z};
}

enum PrimType { kPlane = 0, kSphere = 1, kCylinder = 2 };

struct Prim {
    int type;
    double s;        // size
    Rot R;
    Vec3 c;          // centre
    double area;
};

std::vector<Prim> make_prims(uint64_t seed, int n_prims, double box) {
    std::vector<Prim> prims(n_prims);
    for (int j = 0; j < n_prims; ++j) {
        Stream r(seed, (1ull << 40) + static_cast<uint64_t>(j));
        Prim& p = prims[j];
        p.type = std::min(2, static_cast<int>(r.uniform() * 3.0));
        p.s = 1.0 + 19.0 * r.uniform();
        p.R = random_rotation(r);
        double rb;
        switch (p.type) {
            case kPlane: rb = p.s * std::sqrt(0.5); p.area = p.s * p.s; break;
            case kSphere: rb = 0.5 * p.s; p.area = kPi * p.s * p.s; break;
            default: rb = std::sqrt(p.s * p.s / 16.0 + p.s * p.s / 4.0); p.area = 0.5 * kPi * p.s * p.s; break;
        }
        rb = std::min(rb + 0.05, 0.5 * box);
        p.c = {rb + (box - 2 * rb) * r.uniform(), rb + (box - 2 * rb) * r.uniform(),
               rb + (box - 2 * rb) * r.uniform()};
    }
    return prims;
}

// Largest-remainder apportionment of `total` items proportional to `weights`.
std::vector<int64_t> apportion(const std::vector<double>& weights, int64_t total) {
    const size_t m = weights.size();
    double sum = 0;
    for (double w : weights) sum += w;
    std::vector<int64_t> counts(m);
    std::vector<std::pair<double, size_t>> rem(m);
    int64_t used = 0;
    for (size_t j = 0; j < m; ++j) {
        double exact = static_cast<double>(total) * weights[j] / sum;
        counts[j] = static_cast<int64_t>(std::floor(exact));
        rem[j] = {exact - static_cast<double>(counts[j]), j};
        used += counts[j];
    }
    std::sort(rem.begin(), rem.end(), [](const std::pair<double, size_t>& a, const std::pair<double, size_t>& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    for (int64_t t = 0; t < total - used; ++t) counts[rem[static_cast<size_t>(t) % m].second] += 1;
    return counts;
}

// Local surface frame (point, unit normal) for parameters (a, b) in [0,1)^2.
Vec3 sample_surface(const Prim& p, double a, double b, double noise_n) {
    switch (p.type) {
        case kPlane: {
            Vec3 local{(a - 0.5) * p.s, (b - 0.5) * p.s, noise_n};
            return add(apply(p.R, local), p.c);
        }
        case kSphere: {
            // (a, b) -> uniform direction via z = 1-2a, phi = 2 pi b
            double z = 1.0 - 2.0 * a;
            double rxy = std::sqrt(std::max(0.0, 1.0 - z * z));
            Vec3 dir{rxy * std::cos(2 * kPi * b), rxy * std::sin(2 * kPi * b), z};
            return add(scale(dir, 0.5 * p.s + noise_n), p.c);
        }
        default: {
            double rr = 0.25 * p.s + noise_n;
            double th = 2 * kPi * a;
            Vec3 local{rr * std::cos(th), rr * std::sin(th), (b - 0.5) * p.s};
            return add(apply(p.R, local), p.c);
        }
    }
}

// Fisher-Yates permutation of the three coordinate arrays, driven by one stream.
void shuffle_xyz(uint64_t seed, int64_t n, double* x, double* y, double* z) {
    Stream r(seed, 0xFFFFFFFFFFull);
    for (int64_t i = n - 1; i > 0; --i) {
        int64_t j = static_cast<int64_t>(r.next() % static_cast<uint64_t>(i + 1));
        std::swap(x[i], x[j]);
        std::swap(y[i], y[j]);
        std::swap(z[i], z[j]);
    }
}

void to_float(int64_t n, const double* src, float* dst) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) dst[i] = static_cast<float>(src[i]);
}

}  // namespace

extern "C" {

uint64_t sx_fnv64(const void* data, size_t bytes) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    uint64_t h = 0xCBF29CE484222325ull;
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

int sx_gen_sphere(uint64_t seed, int64_t n, float noise_sigma, float* x, float* y, float* z) {
    if (n < 0 || (n > 0 && (!x || !y || !z))) return sx::fail(SX_EINVAL, "sx_gen_sphere: bad arguments");
    const double sig = noise_sigma;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Stream r(seed, static_cast<uint64_t>(i));
        Vec3 g{r.normal(), r.normal(), r.normal()};
        Vec3 p = normalize(g);
        x[i] = static_cast<float>(p.x + sig * r.normal());
        y[i] = static_cast<float>(p.y + sig * r.normal());
        z[i] = static_cast<float>(p.z + sig * r.normal());
    }
    return SX_OK;
}

int sx_gen_scene(uint64_t seed, int64_t n, int n_prims, float box, float noise_sigma, float outlier_frac,
                 float* x, float* y, float* z) {
    if (n < 0 || n_prims <= 0 || !(box > 0) || outlier_frac < 0 || outlier_frac > 1 ||
        (n > 0 && (!x || !y || !z)))
        return sx::fail(SX_EINVAL, "sx_gen_scene: bad arguments");
    const std::vector<Prim> prims = make_prims(seed, n_prims, box);
    const int64_t n_out = static_cast<int64_t>(std::llround(static_cast<double>(outlier_frac) * n));
    const int64_t n_surf = n - n_out;
    std::vector<double> areas(n_prims);
    for (int j = 0; j < n_prims; ++j) areas[j] = prims[j].area;
    const std::vector<int64_t> counts = apportion(areas, n_surf);
    std::vector<int64_t> start(n_prims + 1, 0);
    for (int j = 0; j < n_prims; ++j) start[j + 1] = start[j] + counts[j];

    std::vector<double> X(n), Y(n), Z(n);
    const double sig = noise_sigma;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Stream r(seed, static_cast<uint64_t>(i));
        Vec3 p;
        if (i < n_surf) {
            int j = static_cast<int>(std::upper_bound(start.begin(), start.end(), i) - start.begin()) - 1;
            double a = r.uniform(), b = r.uniform();
            p = sample_surface(prims[j], a, b, sig * r.normal());
        } else {
            p = {box * r.uniform(), box * r.uniform(), box * r.uniform()};
        }
        X[i] = p.x; Y[i] = p.y; Z[i] = p.z;
    }
    shuffle_xyz(seed, n, X.data(), Y.data(), Z.data());
    to_float(n, X.data(), x);
    to_float(n, Y.data(), y);
    to_float(n, Z.data(), z);
    return SX_OK;
}

int sx_gen_gmm_scene(uint64_t seed, int64_t n, int n_prims, int n_components, float box, float noise_sigma,
                     float max_density_ratio, float* x, float* y, float* z, float* base_fraction_out) {
    if (n < 0 || n_prims <= 0 || n_components <= 0 || !(box > 0) || !(max_density_ratio > 1) ||
        (n > 0 && (!x || !y || !z)))
        return sx::fail(SX_EINVAL, "sx_gen_gmm_scene: bad arguments");
    const std::vector<Prim> prims = make_prims(seed, n_prims, box);
    std::vector<double> areas(n_prims);
    double A = 0;
    for (int j = 0; j < n_prims; ++j) { areas[j] = prims[j].area; A += areas[j]; }
    std::vector<double> cum(n_prims + 1, 0);
    for (int j = 0; j < n_prims; ++j) cum[j + 1] = cum[j] + areas[j] / A;

    // Components: host primitive (area weighted), centre in parameter space,
    // anisotropic scales log-uniform in [0.1, 3] along a random in-surface angle.
    struct Comp { int prim; double a0, b0, sa, sb, ang; };
    std::vector<Comp> comps(n_components);
    double min_area_ratio = 1e300;  // min over j of 2 pi sa sb / (w_j A)
    const double w = 1.0 / n_components;
    for (int c = 0; c < n_components; ++c) {
        Stream r(seed, (2ull << 40) + static_cast<uint64_t>(c));
        double u = r.uniform();
        int j = static_cast<int>(std::upper_bound(cum.begin(), cum.end(), u) - cum.begin()) - 1;
        j = std::min(std::max(j, 0), n_prims - 1);
        Comp& cp = comps[c];
        cp.prim = j;
        cp.a0 = r.uniform();
        cp.b0 = r.uniform();
        cp.sa = std::exp(std::log(0.1) + (std::log(3.0) - std::log(0.1)) * r.uniform());
        cp.sb = std::exp(std::log(0.1) + (std::log(3.0) - std::log(0.1)) * r.uniform());
        cp.ang = 2 * kPi * r.uniform();
        min_area_ratio = std::min(min_area_ratio, 2 * kPi * cp.sa * cp.sb / (w * A));
    }
    // peak/base density = 1 + (1-f)/f / area_ratio  <= R  =>  f >= 1 / (1 + (R-1) * area_ratio)
    const double R = max_density_ratio;
    double f = 1.0 / (1.0 + (R - 1.0) * min_area_ratio);
    f = std::max(f, 0.2);
    if (base_fraction_out) *base_fraction_out = static_cast<float>(f);
    const int64_t n_base = static_cast<int64_t>(std::llround(f * n));

    std::vector<double> X(n), Y(n), Z(n);
    const double sig = noise_sigma;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Stream r(seed, static_cast<uint64_t>(i));
        Vec3 p;
        if (i < n_base) {
            double u = r.uniform();
            int j = static_cast<int>(std::upper_bound(cum.begin(), cum.end(), u) - cum.begin()) - 1;
            j = std::min(std::max(j, 0), n_prims - 1);
            double a = r.uniform(), b = r.uniform();
            p = sample_surface(prims[j], a, b, sig * r.normal());
        } else {
            const Comp& cp = comps[static_cast<size_t>(r.next() % static_cast<uint64_t>(n_components))];
            const Prim& pr = prims[cp.prim];
            // Surface-length units per unit parameter for each primitive type.
            double la, lb;
            switch (pr.type) {
                case kPlane: la = pr.s; lb = pr.s; break;
                case kSphere: la = kPi * 0.5 * pr.s; lb = kPi * pr.s; break;
                default: la = 2 * kPi * 0.25 * pr.s; lb = pr.s; break;
            }
            double a = cp.a0, b = cp.b0;
            for (int tries = 0; tries < 16; ++tries) {
                double ga = cp.sa * r.normal(), gb = cp.sb * r.normal();
                double da = std::cos(cp.ang) * ga - std::sin(cp.ang) * gb;
                double db = std::sin(cp.ang) * ga + std::cos(cp.ang) * gb;
                double ta = cp.a0 + da / la, tb = cp.b0 + db / lb;
                if (pr.type == kSphere) tb = tb - std::floor(tb);          // azimuth wraps
                else if (pr.type == kCylinder) ta = ta - std::floor(ta);   // angle wraps
                if (ta >= 0 && ta < 1 && tb >= 0 && tb < 1) { a = ta; b = tb; break; }
            }
            p = sample_surface(pr, a, b, sig * r.normal());
        }
        X[i] = p.x; Y[i] = p.y; Z[i] = p.z;
    }
    shuffle_xyz(seed, n, X.data(), Y.data(), Z.data());
    to_float(n, X.data(), x);
    to_float(n, Y.data(), y);
    to_float(n, Z.data(), z);
    return SX_OK;
}

int sx_gen_batch_sizes(uint64_t seed, int n_clouds, int64_t lo, int64_t hi, int64_t* sizes) {
    if (n_clouds <= 0 || lo <= 0 || hi < lo || !sizes) return sx::fail(SX_EINVAL, "sx_gen_batch_sizes: bad arguments");
    for (int i = 0; i < n_clouds; ++i) {
        Stream r(seed, static_cast<uint64_t>(i));
        double v = std::exp(std::log(static_cast<double>(lo)) +
                            (std::log(static_cast<double>(hi)) - std::log(static_cast<double>(lo))) * r.uniform());
        sizes[i] = std::min(hi, std::max(lo, static_cast<int64_t>(std::llround(v))));
    }
    return SX_OK;
}

// Degenerate structures.  cls[i] = 0 line, 1 corner, 2 slab.  dir[3i..3i+2] is the
// structure's reference direction: the line direction, the slab normal, or the
// normal of the corner face the point lies on.
int sx_gen_degenerate(uint64_t seed, int64_t n, int n_per_class_structs, float offset, float box,
                      float* x, float* y, float* z, uint8_t* cls, float* dir) {
    if (n < 0 || n_per_class_structs <= 0 || !(box > 0) || (n > 0 && (!x || !y || !z)))
        return sx::fail(SX_EINVAL, "sx_gen_degenerate: bad arguments");
    const int S = n_per_class_structs;
    struct Struct { Rot R; Vec3 c; double L; };
    std::vector<Struct> st(3 * S);
    for (int q = 0; q < 3 * S; ++q) {
        Stream r(seed, (3ull << 40) + static_cast<uint64_t>(q));
        st[q].R = random_rotation(r);
        st[q].L = 5.0 + 15.0 * r.uniform();
        double rb = st[q].L * 1.75;
        st[q].c = {rb + (box - 2 * rb) * r.uniform(), rb + (box - 2 * rb) * r.uniform(),
                   rb + (box - 2 * rb) * r.uniform()};
    }
    const int64_t per_cls = n / 3;
    const double off = offset;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Stream r(seed, static_cast<uint64_t>(i));
        int c = static_cast<int>(std::min<int64_t>(2, i / std::max<int64_t>(1, per_cls)));
        int q = c * S + static_cast<int>(r.next() % static_cast<uint64_t>(S));
        const Struct& s = st[q];
        Vec3 local, ldir;
        if (c == 0) {  // line with tiny isotropic noise
            double t = (r.uniform() - 0.5) * s.L;
            local = {t + 2e-3 * r.normal(), 2e-3 * r.normal(), 2e-3 * r.normal()};
            ldir = {1, 0, 0};
        } else if (c == 1) {  // corner: three orthogonal square faces meeting at the origin
            int face = static_cast<int>(r.next() % 3);
            double u = r.uniform() * s.L, v = r.uniform() * s.L, nn = 1e-3 * r.normal();
            if (face == 0) { local = {nn, u, v}; ldir = {1, 0, 0}; }
            else if (face == 1) { local = {u, nn, v}; ldir = {0, 1, 0}; }
            else { local = {u, v, nn}; ldir = {0, 0, 1}; }
        } else {  // thin slab, thickness / extent = 1e-4
            double u = (r.uniform() - 0.5) * s.L, v = (r.uniform() - 0.5) * s.L;
            double t = (r.uniform() - 0.5) * 1e-4 * s.L;
            local = {u, v, t};
            ldir = {0, 0, 1};
        }
        Vec3 p = add(apply(s.R, local), s.c);
        Vec3 d = apply(s.R, ldir);
        x[i] = static_cast<float>(p.x + off);
        y[i] = static_cast<float>(p.y + off);
        z[i] = static_cast<float>(p.z + off);
        if (cls) cls[i] = static_cast<uint8_t>(c);
        if (dir) { dir[3 * i] = static_cast<float>(d.x); dir[3 * i + 1] = static_cast<float>(d.y); dir[3 * i + 2] = static_cast<float>(d.z); }
    }
    return SX_OK;
}

}  // extern "C"
