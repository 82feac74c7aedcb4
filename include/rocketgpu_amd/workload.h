// Synthetic workloads for the batched arena step: the procedural arena, seeded per-car control streams, seeded ball
// launches and the body-state digest.  Header-only C++ with no dependencies, so the product library, the tests and
// the oracle's shim around the reference CPU build all see byte-identical inputs.
//
// Procedural arena PA(s, r) (uu, z up, normals facing into the playable space):
//   closed box x in [-4096 s, 4096 s], y in [-5120 s, 5120 s], z in [0, 2044 s]; every edge and corner between
//   floor/walls/ceiling rounded by a fillet of radius r (quarter cylinders, sphere octants in the corners) at 16
//   angular segments; flat faces tessellated on a ~256 s uu grid; two goal recesses 1786 s wide x 642 s high x
//   880 s deep behind the end walls (y = +-5120 s), closed by a floor, ceiling, back wall and side walls.  The
//   floor/end-wall fillet is cut at the goal mouth, and the cut is closed by a fan ("cap").
//   PA(1, 256) has 20,228 triangles.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace rgw {

struct V3 {
    float x, y, z;
};

struct ArenaParams {
    double scale = 1.0;
    double fillet_radius = 256.0;
    int fillet_segments = 16;
    double cell = 256.0;
};

struct ArenaGeometry {
    std::vector<V3> vertices;        // uu
    std::vector<int32_t> indices;    // three per triangle
};

struct SpawnSpot {
    float x, y, z, yaw;
};

struct GoalSpec {
    V3 point, normal, min, max;
};

constexpr int kMaxPads = 48;  // ArenaLayout::boost_pads holds at most kMaxBoostPads (rocketgpu.h:526)

struct ArenaRulesSpec {
    GoalSpec goals[2];               // [0]: Blue's goal (-y), [1]: Orange's goal (+y)
    SpawnSpot kickoff[2][5];
    SpawnSpot respawn[2][4];
    V3 ball_spawn;
    int pad_count;  // at most kMaxPads
    V3 pad_pos[48];
    uint8_t pad_big[48];
};

namespace detail {

struct Builder {
    ArenaGeometry g;
    int add(double x, double y, double z) {
        g.vertices.push_back(V3{static_cast<float>(x), static_cast<float>(y), static_cast<float>(z)});
        return static_cast<int>(g.vertices.size()) - 1;
    }
    // One triangle, wound so that its normal has a positive dot product with `inward`.
    void tri(int a, int b, int c, double ix, double iy, double iz) {
        const V3& pa = g.vertices[a];
        const V3& pb = g.vertices[b];
        const V3& pc = g.vertices[c];
        const double ux = double(pb.x) - pa.x, uy = double(pb.y) - pa.y, uz = double(pb.z) - pa.z;
        const double vx = double(pc.x) - pa.x, vy = double(pc.y) - pa.y, vz = double(pc.z) - pa.z;
        const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
        if (nx * ix + ny * iy + nz * iz < 0.0) {
            const int t = b;
            b = c;
            c = t;
        }
        g.indices.push_back(a);
        g.indices.push_back(b);
        g.indices.push_back(c);
    }
};

// Sample points of [a, b] split into pieces of at most `cell`, endpoints exact.
inline std::vector<double> Samples(double a, double b, double cell) {
    int n = static_cast<int>(std::ceil((b - a) / cell - 1e-9));
    if (n < 1) n = 1;
    std::vector<double> out(static_cast<size_t>(n) + 1);
    for (int i = 0; i <= n; ++i) out[static_cast<size_t>(i)] = i == 0 ? a : i == n ? b : a + (b - a) * i / n;
    return out;
}

// Concatenation of Samples over consecutive breakpoints (shared breakpoints appear once).
inline std::vector<double> Breaks(const std::vector<double>& bp, double cell) {
    std::vector<double> out;
    for (size_t k = 0; k + 1 < bp.size(); ++k) {
        std::vector<double> s = Samples(bp[k], bp[k + 1], cell);
        out.insert(out.end(), s.begin() + (out.empty() ? 0 : 1), s.end());
    }
    return out;
}

}  // namespace detail

inline ArenaGeometry BuildProceduralArena(const ArenaParams& p) {
    const double s = p.scale, r = p.fillet_radius, cell = p.cell * s;
    const int na = p.fillet_segments;
    const double X = 4096.0 * s, Y = 5120.0 * s, Z = 2044.0 * s;
    const double G = 893.0 * s, GH = 642.0 * s, GD = 880.0 * s;
    // Inner (fillet axis) and outer (wall) coordinates, all derived the same way so shared vertices are identical.
    const double Xi = X - r, Yi = Y - r, Zt = Z - r;
    const double Xo = Xi + r, Yo = Yi + r, Zo = Zt + r;
    std::vector<double> ca(static_cast<size_t>(na) + 1), sa(static_cast<size_t>(na) + 1);
    for (int k = 0; k <= na; ++k) {
        const double t = 0.5 * 3.14159265358979323846 * k / na;
        ca[static_cast<size_t>(k)] = k == na ? 0.0 : k == 0 ? 1.0 : std::cos(t);
        sa[static_cast<size_t>(k)] = k == na ? 1.0 : k == 0 ? 0.0 : std::sin(t);
    }
    const std::vector<double> xs = detail::Breaks({-Xi, -G, G, Xi}, cell);
    const std::vector<double> ys = detail::Breaks({-Yi, Yi}, cell);
    const std::vector<double> zs = detail::Breaks({r, GH, Zt}, cell);
    detail::Builder b;

    // A grid patch over samples us x vs of the point function P(u, v); inward direction I(point).
    auto grid = [&](const std::vector<double>& us, const std::vector<double>& vs, auto&& P, auto&& I) {
        std::vector<int> idx(us.size() * vs.size());
        for (size_t i = 0; i < us.size(); ++i)
            for (size_t j = 0; j < vs.size(); ++j) {
                double q[3];
                P(us[i], vs[j], q);
                idx[i * vs.size() + j] = b.add(q[0], q[1], q[2]);
            }
        for (size_t i = 0; i + 1 < us.size(); ++i)
            for (size_t j = 0; j + 1 < vs.size(); ++j) {
                const int a = idx[i * vs.size() + j], c = idx[(i + 1) * vs.size() + j];
                const int d = idx[(i + 1) * vs.size() + j + 1], e = idx[i * vs.size() + j + 1];
                const V3 m = b.g.vertices[static_cast<size_t>(a)];
                double in[3];
                I(m, in);
                b.tri(a, c, d, in[0], in[1], in[2]);
                b.tri(a, d, e, in[0], in[1], in[2]);
            }
    };
    std::vector<double> ks(static_cast<size_t>(na) + 1);
    for (int k = 0; k <= na; ++k) ks[static_cast<size_t>(k)] = k;
    auto K = [&](double k) { return static_cast<size_t>(k); };

    // Flat faces.
    grid(xs, ys, [&](double u, double v, double* q) { q[0] = u; q[1] = v; q[2] = 0.0; },
         [](const V3&, double* in) { in[0] = 0; in[1] = 0; in[2] = 1; });
    grid(xs, ys, [&](double u, double v, double* q) { q[0] = u; q[1] = v; q[2] = Zo; },
         [](const V3&, double* in) { in[0] = 0; in[1] = 0; in[2] = -1; });
    for (int sx : {-1, 1}) {
        grid(ys, zs, [&](double u, double v, double* q) { q[0] = sx * Xo; q[1] = u; q[2] = v; },
             [&](const V3&, double* in) { in[0] = -sx; in[1] = 0; in[2] = 0; });
    }
    const std::vector<double> xl = detail::Breaks({-Xi, -G}, cell), xm = detail::Breaks({-G, G}, cell),
                              xr = detail::Breaks({G, Xi}, cell);
    const std::vector<double> zt = detail::Breaks({GH, Zt}, cell);
    for (int sy : {-1, 1}) {
        auto wall = [&](double u, double v, double* q) { q[0] = u; q[1] = sy * Yo; q[2] = v; };
        auto win = [&](const V3&, double* in) { in[0] = 0; in[1] = -sy; in[2] = 0; };
        grid(xl, zs, wall, win);
        grid(xr, zs, wall, win);
        grid(xm, zt, wall, win);
    }
    // Fillets along y (floor/ceiling meet the side walls): cross-section angle k from the wall (k=0) to the
    // floor or ceiling (k=na).
    for (int sx : {-1, 1})
        for (int top : {0, 1}) {
            const double zc = top ? Zt : r, sz = top ? 1.0 : -1.0;
            grid(ys, ks,
                 [&](double u, double k, double* q) {
                     q[0] = sx * Xi + sx * r * ca[K(k)];
                     q[1] = u;
                     q[2] = zc + sz * r * sa[K(k)];
                 },
                 [&](const V3& m, double* in) { in[0] = sx * Xi - m.x; in[1] = 0; in[2] = zc - m.z; });
        }
    // Fillets along x (floor/ceiling meet the end walls); the floor one is cut at the goal mouth.
    for (int sy : {-1, 1})
        for (int top : {0, 1}) {
            const double zc = top ? Zt : r, sz = top ? 1.0 : -1.0;
            auto P = [&](double u, double k, double* q) {
                q[0] = u;
                q[1] = sy * Yi + sy * r * ca[K(k)];
                q[2] = zc + sz * r * sa[K(k)];
            };
            auto I = [&](const V3& m, double* in) { in[0] = 0; in[1] = sy * Yi - m.y; in[2] = zc - m.z; };
            if (top) {
                grid(xs, ks, P, I);
            } else {
                grid(xl, ks, P, I);
                grid(xr, ks, P, I);
            }
        }
    // Vertical fillets in the four corners.
    for (int sx : {-1, 1})
        for (int sy : {-1, 1}) {
            grid(zs, ks,
                 [&](double u, double k, double* q) {
                     q[0] = sx * Xi + sx * r * ca[K(k)];
                     q[1] = sy * Yi + sy * r * sa[K(k)];
                     q[2] = u;
                 },
                 [&](const V3& m, double* in) { in[0] = sx * Xi - m.x; in[1] = sy * Yi - m.y; in[2] = 0; });
        }
    // Sphere octants where three fillets meet: azimuth j (0: toward the side wall, na: toward the end wall) and
    // elevation k (0: horizontal ring, na: pole on the floor/ceiling).
    for (int sx : {-1, 1})
        for (int sy : {-1, 1})
            for (int top : {0, 1}) {
                const double zc = top ? Zt : r, sz = top ? 1.0 : -1.0;
                std::vector<int> ring((static_cast<size_t>(na) + 1) * static_cast<size_t>(na));
                for (int k = 0; k < na; ++k)
                    for (int j = 0; j <= na; ++j) {
                        const double c = ca[static_cast<size_t>(k)];
                        ring[static_cast<size_t>(k) * (na + 1) + j] =
                            b.add(sx * Xi + sx * r * (c * ca[static_cast<size_t>(j)]),
                                  sy * Yi + sy * r * (c * sa[static_cast<size_t>(j)]),
                                  zc + sz * r * sa[static_cast<size_t>(k)]);
                    }
                const int pole = b.add(sx * Xi + sx * r * 0.0, sy * Yi + sy * r * 0.0, zc + sz * r * 1.0);
                auto in = [&](int v, double* o) {
                    const V3& m = b.g.vertices[static_cast<size_t>(v)];
                    o[0] = sx * Xi - m.x;
                    o[1] = sy * Yi - m.y;
                    o[2] = zc - m.z;
                };
                for (int k = 0; k + 1 < na; ++k)
                    for (int j = 0; j < na; ++j) {
                        const int a = ring[static_cast<size_t>(k) * (na + 1) + j];
                        const int c = ring[static_cast<size_t>(k) * (na + 1) + j + 1];
                        const int d = ring[static_cast<size_t>(k + 1) * (na + 1) + j + 1];
                        const int e = ring[static_cast<size_t>(k + 1) * (na + 1) + j];
                        double o[3];
                        in(a, o);
                        b.tri(a, c, d, o[0], o[1], o[2]);
                        b.tri(a, d, e, o[0], o[1], o[2]);
                    }
                for (int j = 0; j < na; ++j) {
                    const int a = ring[static_cast<size_t>(na - 1) * (na + 1) + j];
                    const int c = ring[static_cast<size_t>(na - 1) * (na + 1) + j + 1];
                    double o[3];
                    in(a, o);
                    b.tri(a, c, pole, o[0], o[1], o[2]);
                }
            }
    // Goal recesses.
    for (int sy : {-1, 1}) {
        const std::vector<double> yg = detail::Breaks({Yi, Yo, Yo + GD}, cell);
        const std::vector<double> yb = detail::Breaks({Yo, Yo + GD}, cell);
        const std::vector<double> zg = detail::Breaks({0.0, GH}, cell);
        grid(xm, yg, [&](double u, double v, double* q) { q[0] = u; q[1] = sy * v; q[2] = 0.0; },
             [](const V3&, double* in) { in[0] = 0; in[1] = 0; in[2] = 1; });
        grid(xm, yb, [&](double u, double v, double* q) { q[0] = u; q[1] = sy * v; q[2] = GH; },
             [](const V3&, double* in) { in[0] = 0; in[1] = 0; in[2] = -1; });
        grid(xm, zg, [&](double u, double v, double* q) { q[0] = u; q[1] = sy * (Yo + GD); q[2] = v; },
             [&](const V3&, double* in) { in[0] = 0; in[1] = -sy; in[2] = 0; });
        for (int sx : {-1, 1}) {
            grid(yb, zg, [&](double u, double v, double* q) { q[0] = sx * G; q[1] = sy * u; q[2] = v; },
                 [&](const V3&, double* in) { in[0] = -sx; in[1] = 0; in[2] = 0; });
            // Cap of the cut floor fillet at x = sx G: fan from the hidden corner (y = sy Yo, z = 0).
            std::vector<int> arc(static_cast<size_t>(na) + 1);
            for (int k = 0; k <= na; ++k)
                arc[static_cast<size_t>(k)] = b.add(sx * G, sy * Yi + sy * r * ca[static_cast<size_t>(k)],
                                                    r + -1.0 * r * sa[static_cast<size_t>(k)]);
            const int corner = b.add(sx * G, sy * Yo, 0.0);
            for (int k = 0; k < na; ++k) {
                b.tri(corner, arc[static_cast<size_t>(k)], arc[static_cast<size_t>(k) + 1], -sx, 0.0, 0.0);
            }
        }
    }
    return b.g;
}

// Goals, spawns, pads and ball spawn of PA(s, r): the official soccar tables with x and y scaled by s.
inline ArenaRulesSpec ProceduralArenaRules(const ArenaParams& p) {
    static const float kKick[5][3] = {{-2048.0f, -2560.0f, 0.25f}, {2048.0f, -2560.0f, 0.75f},
                                      {-256.0f, -3840.0f, 0.5f},   {256.0f, -3840.0f, 0.5f},
                                      {0.0f, -4608.0f, 0.5f}};
    static const float kRespawn[4][2] = {{-2304.0f, -4608.0f}, {-2688.0f, -4608.0f},
                                         {2304.0f, -4608.0f},  {2688.0f, -4608.0f}};
    static const float kPads[34][3] = {
        {0, -4240, 70},     {-1792, -4184, 70}, {1792, -4184, 70}, {-940, -3308, 70},  {940, -3308, 70},
        {0, -2816, 70},     {-3584, -2484, 70}, {3584, -2484, 70}, {-1788, -2302, 70}, {1788, -2302, 70},
        {-2048, -1036, 70}, {0, -1024, 70},     {2048, -1036, 70}, {-1024, 0, 70},     {1024, 0, 70},
        {-2048, 1036, 70},  {0, 1024, 70},      {2048, 1036, 70},  {-1788, 2302, 70},  {1788, 2302, 70},
        {-3584, 2484, 70},  {3584, 2484, 70},   {0, 2816, 70},     {-940, 3308, 70},   {940, 3308, 70},
        {-1792, 4184, 70},  {1792, 4184, 70},   {0, 4240, 70},     {-3584, 0, 73},     {3584, 0, 73},
        {-3072, 4096, 73},  {3072, 4096, 73},   {-3072, -4096, 73}, {3072, -4096, 73}};
    const float pi = 3.14159265358979323846f;
    const float s = static_cast<float>(p.scale);
    ArenaRulesSpec o{};
    const float yo = static_cast<float>(5120.0 * p.scale);
    const float g = static_cast<float>(893.0 * p.scale), gh = static_cast<float>(642.0 * p.scale);
    for (int t = 0; t < 2; ++t) {
        const float sy = t == 0 ? -1.0f : 1.0f;
        o.goals[t].point = V3{0.0f, sy * yo, 0.0f};
        o.goals[t].normal = V3{0.0f, sy, 0.0f};
        o.goals[t].min = V3{-g, -1.0e30f, 0.0f};
        o.goals[t].max = V3{g, 1.0e30f, gh};
    }
    for (int k = 0; k < 5; ++k) {
        o.kickoff[0][k] = SpawnSpot{kKick[k][0] * s, kKick[k][1] * s, 0.0f, pi * kKick[k][2]};
        o.kickoff[1][k] = SpawnSpot{-kKick[k][0] * s, -kKick[k][1] * s, 0.0f, pi * kKick[k][2] + pi};
    }
    for (int k = 0; k < 4; ++k) {
        o.respawn[0][k] = SpawnSpot{kRespawn[k][0] * s, kRespawn[k][1] * s, 0.0f, pi * 0.5f};
        o.respawn[1][k] = SpawnSpot{kRespawn[k][0] * s, -kRespawn[k][1] * s, 0.0f, pi * 0.5f + pi};
    }
    o.ball_spawn = V3{0.0f, 0.0f, 92.75f};
    o.pad_count = 34;
    for (int k = 0; k < 34; ++k) {
        o.pad_pos[k] = V3{kPads[k][0] * s, kPads[k][1] * s, kPads[k][2]};
        o.pad_big[k] = k >= 28 ? 1 : 0;
    }
    return o;
}

// ---------------------------------------------------------------------------------------------------------------
// Seeded streams.

inline uint64_t SplitMix64(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t StreamSeed(uint64_t global_seed, uint64_t arena, uint64_t slot, uint64_t salt) {
    uint64_t s = global_seed ^ (arena * 0xD1B54A32D192ED03ull) ^ ((slot + 1) * 0x8CB92BA72F3D8DD7ull) ^
                 (salt * 0x9E3779B97F4A7C15ull);
    return SplitMix64(s);
}

inline float UnitFloat(uint64_t bits) {  // [0, 1) on a 2^-24 grid
    return static_cast<float>(bits >> 40) * (1.0f / 16777216.0f);
}

// Wire format of one car's input, identical to the reference's CarControls (24 bytes).
struct Controls {
    float throttle, steer, pitch, yaw, roll;
    uint8_t jump, boost, handbrake, pad;
};
static_assert(sizeof(Controls) == 24, "Controls is 24 bytes");

// One car's control stream: every draw is held for a uniform 1..30 ticks.  Axes uniform in [-1, 1];
// jump ~ Bernoulli(0.1), boost ~ Bernoulli(p_boost), handbrake ~ Bernoulli(0.05).
struct ControlStream {
    uint64_t state = 0;
    int hold = 0;
    float p_boost = 0.3f;
    Controls cur{};

    void seed(uint64_t global_seed, uint64_t arena, uint64_t slot, float boost_probability) {
        state = StreamSeed(global_seed, arena, slot, 1);
        hold = 0;
        p_boost = boost_probability;
        cur = Controls{};
    }
    Controls next() {
        if (hold <= 0) {
            float* axes = &cur.throttle;
            for (int a = 0; a < 5; ++a) axes[a] = UnitFloat(SplitMix64(state)) * 2.0f - 1.0f;
            cur.jump = UnitFloat(SplitMix64(state)) < 0.1f ? 1 : 0;
            cur.boost = UnitFloat(SplitMix64(state)) < p_boost ? 1 : 0;
            cur.handbrake = UnitFloat(SplitMix64(state)) < 0.05f ? 1 : 0;
            cur.pad = 0;
            hold = 1 + static_cast<int>(SplitMix64(state) % 30u);
        }
        --hold;
        return cur;
    }
};

// The held controls car `slot` of arena `arena` gets for request number `request` of its client in the serving
// workload: the first draw of a control stream seeded by the request.
inline Controls RequestControls(uint64_t seed, uint64_t request, uint64_t arena, uint64_t slot, float p_boost) {
    ControlStream s;
    s.seed(seed ^ (0x9e3779b97f4a7c15ull * (request + 1)), arena, slot, p_boost);
    return s.next();
}

// A ball launch somewhere in the open part of PA(s, r): position in the inner box, velocity up to 3000 uu/s,
// angular velocity up to 5 rad/s per axis.
struct BallLaunch {
    float pos[3], vel[3], ang_vel[3];
};

inline BallLaunch SeededBallLaunch(uint64_t global_seed, uint64_t arena, double scale) {
    uint64_t st = StreamSeed(global_seed, arena, 0xBA11, 2);
    BallLaunch b{};
    const float ext[3] = {static_cast<float>(3500.0 * scale), static_cast<float>(4500.0 * scale),
                          static_cast<float>(1700.0 * scale)};
    for (int a = 0; a < 3; ++a) b.pos[a] = (UnitFloat(SplitMix64(st)) * 2.0f - 1.0f) * ext[a];
    b.pos[2] = static_cast<float>(150.0 * scale) + (b.pos[2] + ext[2]) * 0.5f;
    for (int a = 0; a < 3; ++a) b.vel[a] = (UnitFloat(SplitMix64(st)) * 2.0f - 1.0f) * 1730.0f;
    for (int a = 0; a < 3; ++a) b.ang_vel[a] = (UnitFloat(SplitMix64(st)) * 2.0f - 1.0f) * 5.0f;
    return b;
}

// Mixed car counts: fraction 55/27/18 % of 1v1/2v2/3v3, decided per arena from the seed (stable for any
// arena count).
inline int MixedCarCount(uint64_t global_seed, uint64_t arena) {
    const float u = UnitFloat(StreamSeed(global_seed, arena, 0xC0, 3));
    return u < 0.55f ? 2 : u < 0.82f ? 4 : 6;
}

// ---------------------------------------------------------------------------------------------------------------
// Body-state digest: FNV-1a 64 over the raw bits of pos, vel, ang_vel (3 floats each) and rot (4 floats) of each
// body, in the order given.

constexpr uint64_t kFnvOffset = 0xcbf29ce484222325ull;
constexpr uint64_t kFnvPrime = 0x100000001b3ull;

inline uint64_t Fnv1a(uint64_t h, const void* data, size_t n) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= kFnvPrime;
    }
    return h;
}

inline uint64_t BodyDigest(uint64_t h, const float pos[3], const float vel[3], const float ang_vel[3],
                           const float rot_wxyz[4]) {
    h = Fnv1a(h, pos, 12);
    h = Fnv1a(h, vel, 12);
    h = Fnv1a(h, ang_vel, 12);
    return Fnv1a(h, rot_wxyz, 16);
}

}  // namespace rgw
