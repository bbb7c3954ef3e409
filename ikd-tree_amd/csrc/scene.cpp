// scene.cpp — deterministic synthetic workloads (host code, g++).
//
// A LiDAR-like mapping scene: ground plane z = 0 over [-200,200]^2 plus
// axis-aligned boxes standing on it; a sensor 1.5 m above ground moves along
// a circle of radius 80 m, 1 m of arc per frame, and each frame is a 32-beam
// spinning scan (elevations -15..+15 deg, 1800 azimuths), ray-cast against
// the ground and the boxes, range <= 100 m, range noise N(0, 0.02 m).
//
// Only +, -, *, /, sqrt and integer operations are used (no libm
// transcendentals, no FMA contraction), so every machine produces the same
// bytes: sines and cosines come from a fixed Taylor series, Gaussians from
// the Irwin-Hall sum of 12 uniforms, uniforms from PCG32.
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

struct Pcg32 {
    uint64_t state, inc;
    Pcg32(uint64_t seed, uint64_t stream) {
        state = 0;
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    // u in [0, 1) with 24 random bits
    double uniform() { return (double)(next() >> 8) * (1.0 / 16777216.0); }
    double gauss() {
        double s = 0.0;
        for (int i = 0; i < 12; ++i) s += uniform();
        return s - 6.0;
    }
};

constexpr double kPi = 3.14159265358979323846;

// sin/cos by range reduction to [-pi, pi] and a Taylor series to x^25.
double det_sin(double x) {
    double k = std::floor(x / (2.0 * kPi) + 0.5);
    x = x - k * (2.0 * kPi);
    double term = x, sum = x, x2 = x * x;
    for (int n = 1; n <= 12; ++n) {
        term = -term * x2 / (double)((2 * n) * (2 * n + 1));
        sum = sum + term;
    }
    return sum;
}
double det_cos(double x) { return det_sin(x + 0.5 * kPi); }

}  // namespace

extern "C" {

// Uniform points in [lo, hi)^3 (float32 xyz triples).
void scene_uniform(uint64_t seed, uint64_t stream, int64_t n, double lo, double hi, float* out) {
    Pcg32 r(seed, stream);
    for (int64_t i = 0; i < 3 * n; ++i) out[i] = (float)(lo + (hi - lo) * r.uniform());
}

// Uniform random boxes: centre in [lo, hi)^3, each side in [smin, smax).
void scene_boxes(uint64_t seed, uint64_t stream, int64_t n, double lo, double hi, double smin, double smax,
                 float* out) {
    Pcg32 r(seed, stream);
    for (int64_t i = 0; i < n; ++i) {
        double c[3], s[3];
        for (int a = 0; a < 3; ++a) c[a] = lo + (hi - lo) * r.uniform();
        for (int a = 0; a < 3; ++a) s[a] = smin + (smax - smin) * r.uniform();
        for (int a = 0; a < 3; ++a) {
            out[6 * i + a] = (float)(c[a] - 0.5 * s[a]);
            out[6 * i + 3 + a] = (float)(c[a] + 0.5 * s[a]);
        }
    }
}

// Obstacles of the mapping scene: n boxes (min xyz, max xyz) with xy centre
// uniform in [-200,200]^2, sides uniform in [1,10) m, standing on z = 0, none
// within 3 m of the sensor's circular path of radius `path_radius`.
void scene_obstacles(uint64_t seed, int64_t n, double path_radius, float* out) {
    Pcg32 r(seed, 7);
    int64_t made = 0;
    while (made < n) {
        double cx = -200.0 + 400.0 * r.uniform();
        double cy = -200.0 + 400.0 * r.uniform();
        double sx = 1.0 + 9.0 * r.uniform(), sy = 1.0 + 9.0 * r.uniform(), sz = 1.0 + 9.0 * r.uniform();
        double x0 = cx - 0.5 * sx, x1 = cx + 0.5 * sx, y0 = cy - 0.5 * sy, y1 = cy + 0.5 * sy;
        // distance range from the origin to the rectangle
        double nx = cx < 0 ? std::fmax(x0, std::fmin(0.0, x1)) : std::fmin(x1, std::fmax(0.0, x0));
        double ny = cy < 0 ? std::fmax(y0, std::fmin(0.0, y1)) : std::fmin(y1, std::fmax(0.0, y0));
        double dmin = std::sqrt(nx * nx + ny * ny);
        double fx = std::fmax(std::fabs(x0), std::fabs(x1)), fy = std::fmax(std::fabs(y0), std::fabs(y1));
        double dmax = std::sqrt(fx * fx + fy * fy);
        if (dmin <= path_radius + 3.0 && dmax >= path_radius - 3.0) continue;
        float* b = out + 6 * made;
        b[0] = (float)x0; b[1] = (float)y0; b[2] = 0.0f;
        b[3] = (float)x1; b[4] = (float)y1; b[5] = (float)sz;
        ++made;
    }
}

// Sensor pose of frame f: on the circle, 1 m of arc per frame, 1.5 m high.
void scene_pose(int64_t frame, double path_radius, double* out3) {
    double th = (double)frame / path_radius;
    out3[0] = path_radius * det_cos(th);
    out3[1] = path_radius * det_sin(th);
    out3[2] = 1.5;
}

// One scan: returns the number of points written (<= 32*1800).
int64_t scene_scan(const float* obstacles, int64_t nobs, int64_t frame, uint64_t seed, double path_radius,
                   float* out) {
    double o[3];
    scene_pose(frame, path_radius, o);
    Pcg32 r(seed, 1000 + (uint64_t)frame);
    // obstacles that a 100 m ray can reach
    static thread_local int idx[4096];
    int nidx = 0;
    for (int64_t i = 0; i < nobs && nidx < 4096; ++i) {
        const float* b = obstacles + 6 * i;
        double dx = std::fmax(std::fmax(b[0] - o[0], 0.0), o[0] - b[3]);
        double dy = std::fmax(std::fmax(b[1] - o[1], 0.0), o[1] - b[4]);
        if (dx * dx + dy * dy <= 100.0 * 100.0) idx[nidx++] = (int)i;
    }
    int64_t np = 0;
    const int n_el = 32, n_az = 1800;
    for (int ie = 0; ie < n_el; ++ie) {
        double el = (-15.0 + 30.0 * ie / (n_el - 1)) * kPi / 180.0;
        double ce = det_cos(el), se = det_sin(el);
        for (int ia = 0; ia < n_az; ++ia) {
            double az = 2.0 * kPi * ia / n_az;
            double d[3] = {ce * det_cos(az), ce * det_sin(az), se};
            double tbest = 100.0;
            bool hit = false;
            if (d[2] < 0.0) {
                double tg = -o[2] / d[2];
                if (tg <= tbest) { tbest = tg; hit = true; }
            }
            for (int k = 0; k < nidx; ++k) {
                const float* b = obstacles + 6 * idx[k];
                double t0 = 0.0, t1 = tbest;
                bool ok = true;
                for (int a = 0; a < 3 && ok; ++a) {
                    if (d[a] == 0.0) {
                        if (o[a] < b[a] || o[a] > b[3 + a]) ok = false;
                        continue;
                    }
                    double ta = (b[a] - o[a]) / d[a], tb = (b[3 + a] - o[a]) / d[a];
                    if (ta > tb) { double tmp = ta; ta = tb; tb = tmp; }
                    if (ta > t0) t0 = ta;
                    if (tb < t1) t1 = tb;
                    if (t0 > t1) ok = false;
                }
                if (ok && t0 > 0.0 && t0 < tbest) { tbest = t0; hit = true; }
            }
            double noise = 0.02 * r.gauss();
            if (!hit) continue;
            double rr = tbest + noise;
            out[3 * np] = (float)(o[0] + rr * d[0]);
            out[3 * np + 1] = (float)(o[1] + rr * d[1]);
            out[3 * np + 2] = (float)(o[2] + rr * d[2]);
            ++np;
        }
    }
    return np;
}

}  // extern "C"
