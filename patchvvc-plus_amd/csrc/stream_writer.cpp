/*
 * stream_writer.cpp -- produces patchVVC .frame streams (host, C++).
 *
 * This is the encoder side of the format, restated only as far as a decoder
 * needs inputs: the GoP slice encoder (src/patch/GoP_encoding.cc:97-211) on
 * top of the forward octree and RAHT (src/octree/raht_octree.cc:47-247,
 * src/octree/octree_base.cc:20-31,65-82,164-184), the RLGR encoder
 * (src/common/entropy_codec.cc:22-159), the optional zstd wrap
 * (entropy_codec.cc:304-324) and the .frame writer
 * (src/io/slice_io.cc:258-315).
 *
 * The reference builds its inputs from captured PLY sequences through PCL
 * segmentation, ICP and patch fitting (out of scope here).  In their place
 * a seeded synthetic generator voxelises a humanoid (an ellipsoid torso and
 * six capsules: head, neck, two arms, two legs), cuts each primitive's
 * surface voxels into patches by recursive median splits with lognormal
 * target sizes, and moves every primitive by a small seeded rigid motion per
 * frame.  Per GoP the common geometry of a patch is its voxels at the first
 * frame; the slice motion matrix maps it to the later frame (the role of
 * `mv.inverse()` at GoP_encoding.cc:122); colours are a smooth sinusoidal
 * field sampled at the moved positions plus Gaussian noise, and predictive
 * slices carry the YUV residual against the first frame
 * (GoP_encoding.cc:145-149).
 */
#include "pvvc_internal.h"

#include <zstd.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>

namespace {

/* ------------------------------------------------------------------------ */
/* Forward octree + RAHT                                                    */
/* ------------------------------------------------------------------------ */
struct XYZ {
    float x, y, z;
};
struct YUV {
    float y = 0, u = 0, v = 0;
};

const uint8_t kNodeValue[8] = {0x80, 0x40, 0x20, 0x10, 0x08, 0x04, 0x02, 0x01}; /* octree.h:76 */
const int kNodeWeight[8][2] = {{0, 0}, {2, 3}, {4, 6}, {5, 7}, {8, 12}, {9, 13}, {10, 14}, {11, 15}}; /* octree.h:79 */

/* SubSpaceCenter (octree_base.cc:102-120); `sub` is the child range. */
XYZ sub_space_center(XYZ c, XYZ sub, int pos) {
    float hx = sub.x / 2.0f, hy = sub.y / 2.0f, hz = sub.z / 2.0f;
    return XYZ{(pos & 4) ? c.x - hx : c.x + hx, (pos & 2) ? c.y - hy : c.y + hy, (pos & 1) ? c.z - hz : c.z + hz};
}

/* HaarTransform (octree_base.cc:164-184). */
void haar(int w0, int w1, const YUV& g0, const YUV& g1, YUV& G, YUV& H) {
    if (w0 == 0 && w1 == 0) {
        G = YUV{};
        H = YUV{};
        return;
    }
    float base = std::sqrt(static_cast<float>(w0 + w1));
    float t00 = std::sqrt(static_cast<float>(w0)) / base, t01 = std::sqrt(static_cast<float>(w1)) / base;
    float t10 = -std::sqrt(static_cast<float>(w1)) / base, t11 = std::sqrt(static_cast<float>(w0)) / base;
    G.y = g0.y * t00 + g1.y * t01;
    G.u = g0.u * t00 + g1.u * t01;
    G.v = g0.v * t00 + g1.v * t01;
    H.y = g0.y * t10 + g1.y * t11;
    H.u = g0.u * t10 + g1.u * t11;
    H.v = g0.v * t10 + g1.v * t11;
}

struct FNode {
    uint8_t value = 0;
    std::vector<int> index;
    int weight[16] = {};
    YUV raht[16];
    /* HierarchicalTransform (octree_base.cc:65-82). */
    void hier() {
        YUV H[8];
        for (int i = 7; i > 0; --i) {
            YUV G, h;
            haar(weight[kNodeWeight[i][0]], weight[kNodeWeight[i][1]], raht[kNodeWeight[i][0]], raht[kNodeWeight[i][1]], G, h);
            raht[i] = G;
            H[i] = h;
        }
        raht[0] = raht[1];
        for (int i = 1; i <= 7; ++i) raht[i] = H[i];
    }
};

/* RAHTOctree (raht_octree.cc). */
struct RahtOctree {
    const XYZ* cloud = nullptr;
    int n = 0;
    XYZ center{}, range{};
    int height = 0;
    std::vector<std::vector<FNode>> tree;

    void add_node(std::vector<int>& pts, int h, XYZ c, XYZ r) { /* :115-187 */
        if (pts.empty()) return;
        if (h == height - 1) {
            FNode node;
            node.value = 0xff;
            node.index.swap(pts);
            node.weight[1] = 1;
            tree[h].push_back(std::move(node));
            return;
        }
        FNode node;
        node.index.assign(8, -1);
        XYZ sub{r.x / 2.0f, r.y / 2.0f, r.z / 2.0f};
        std::vector<int> subs[8];
        for (int i : pts) {
            int pos = 0;
            pos |= cloud[i].x > c.x ? 0 : 1;
            pos <<= 1;
            pos |= cloud[i].y > c.y ? 0 : 1;
            pos <<= 1;
            pos |= cloud[i].z > c.z ? 0 : 1;
            subs[pos].push_back(i);
        }
        std::vector<int>().swap(pts);
        for (int i = 0; i < 8; ++i) {
            if (subs[i].empty()) continue;
            node.value |= kNodeValue[i];
            XYZ sc = sub_space_center(c, sub, i);
            node.index[i] = static_cast<int>(tree[h + 1].size());
            add_node(subs[i], h + 1, sc, sub);
        }
        for (int i = 0; i < 8; ++i)
            if (node.index[i] != -1) node.weight[i + 8] = tree[h + 1][node.index[i]].weight[1];
        for (int i = 7; i > 0; --i) node.weight[i] = node.weight[kNodeWeight[i][0]] + node.weight[kNodeWeight[i][1]];
        tree[h].push_back(std::move(node));
    }

    bool make_tree(const XYZ* pts, int npts, float resolution = 1.0f) { /* :62-113 */
        cloud = pts;
        n = npts;
        if (n <= 0) return false;
        float min_x = FLT_MAX, min_y = FLT_MAX, min_z = FLT_MAX;
        float max_x = FLT_TRUE_MIN, max_y = FLT_TRUE_MIN, max_z = FLT_TRUE_MIN;
        for (int i = 0; i < n; ++i) {
            min_x = std::min(min_x, pts[i].x), min_y = std::min(min_y, pts[i].y), min_z = std::min(min_z, pts[i].z);
            max_x = std::max(max_x, pts[i].x), max_y = std::max(max_y, pts[i].y), max_z = std::max(max_z, pts[i].z);
        }
        center.x = max_x / 2.0f + min_x / 2.0f;
        center.y = max_y / 2.0f + min_y / 2.0f;
        center.z = max_z / 2.0f + min_z / 2.0f;
        float max_range = std::max(max_x - min_x, std::max(max_y - min_y, max_z - min_z));
        if (max_range <= 0.0f) max_range = 1.0f;
        int max_height = static_cast<int>(std::ceil(std::log2(max_range)));
        int ths_height = static_cast<int>(std::ceil(std::log2(resolution)));
        height = max_height - ths_height + 1;
        if (height <= 0) return false;
        tree.assign(height, {});
        max_range = std::pow(2.0f, static_cast<float>(max_height));
        range = XYZ{max_range, max_range, max_range};
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        add_node(all, 0, center, range);
        return true;
    }

    /* GetOctree (raht_octree.cc:47-56) with SaveTreeCore (octree_base.cc:20-31). */
    std::vector<uint8_t> get_octree() const {
        std::vector<uint8_t> out;
        float d[6] = {center.x, center.y, center.z, range.x, range.y, range.z};
        for (float f : d) {
            uint32_t w;
            std::memcpy(&w, &f, 4);
            out.push_back(static_cast<uint8_t>(w >> 24));
            out.push_back(static_cast<uint8_t>(w >> 16));
            out.push_back(static_cast<uint8_t>(w >> 8));
            out.push_back(static_cast<uint8_t>(w));
        }
        out.push_back(static_cast<uint8_t>(height));
        for (int i = 0; i < height - 1; ++i)
            for (auto& nd : tree[i]) out.push_back(nd.value);
        return out;
    }

    size_t leaves() const { return tree.empty() ? 0 : tree.back().size(); }

    /* RAHT (raht_octree.cc:189-247): returns the coefficients in stream
     * order (reversed).  `colors` holds one YUV per input point. */
    std::vector<YUV> raht(const YUV* colors) {
        for (auto& nd : tree.back()) {
            YUV t;
            for (int i : nd.index) {
                t.y += colors[i].y;
                t.u += colors[i].u;
                t.v += colors[i].v;
            }
            int c = static_cast<int>(nd.index.size());
            t.y /= c;
            t.u /= c;
            t.v /= c;
            nd.raht[0] = t;
        }
        for (int i = height - 2; i >= 0; --i) {
            for (auto& nd : tree[i]) {
                for (int idx = 0; idx < 8; ++idx)
                    if (nd.index[idx] != -1) nd.raht[idx + 8] = tree[i + 1][nd.index[idx]].raht[0];
                nd.hier();
            }
        }
        std::vector<YUV> res;
        res.push_back(tree.front().front().raht[0]);
        for (int i = 0; i < height - 1; ++i)
            for (auto& nd : tree[i])
                for (int idx = 1; idx < 8; ++idx)
                    if (nd.weight[kNodeWeight[idx][0]] != 0 && nd.weight[kNodeWeight[idx][1]] != 0) res.push_back(nd.raht[idx]);
        std::reverse(res.begin(), res.end());
        return res;
    }
};

/* ------------------------------------------------------------------------ */
/* RLGR encoder (entropy_codec.cc:22-159)                                    */
/* ------------------------------------------------------------------------ */
struct RlgrWriter {
    uint64_t buf = 0;
    int cnt = 0;
    std::vector<uint8_t> out;
    static uint64_t mask(int k) { return (uint64_t{1} << k) - 1; }
    void flush() {
        while (cnt >= 8) {
            cnt -= 8;
            out.push_back(static_cast<uint8_t>((buf >> cnt) & 0xFF));
        }
    }
    void put1(uint8_t x) {
        buf = (buf << 1) | (x & 1);
        ++cnt;
        flush();
    }
    void put(uint64_t x, int bits) {
        if (bits > 56) {
            put(x >> 32, bits - 32);
            put(x & mask(32), 32);
            return;
        }
        buf = (buf << bits) + x;
        cnt += bits;
        flush();
    }
    void put_gr(uint64_t x, int k) {
        uint64_t p = x >> k;
        if (p < 32) {
            put(mask(static_cast<int>(p) + 1) - 1, static_cast<int>(p) + 1);
            put(x & mask(k), k);
        } else {
            put(mask(32), 32);
            put(x, 32);
        }
    }
    static uint64_t zigzag(int64_t x) { return x < 0 ? (static_cast<uint64_t>(-x) << 1) - 1 : static_cast<uint64_t>(x) << 1; }
    void encode(const int64_t* d, int64_t n) {
        const uint64_t L = 4;
        uint64_t u = 0, kP = 0, kRP = 2 * L, m = 0, k = 0, kR = 0, p;
        for (int64_t i = 0; i < n; ++i) {
            u = zigzag(d[i]);
            k = kP / L;
            kR = kRP / L;
            if (k) {
                if (u) {
                    --u;
                    put1(0);
                    put(m, static_cast<int>(k));
                    put_gr(u, static_cast<int>(kR));
                    p = u >> kR;
                    if (p) {
                        kRP += p - 1;
                        kRP = std::min<uint64_t>(kRP, 32 * L);
                    } else {
                        kRP = kRP < 2 ? 0 : kRP - 2;
                    }
                    kP = kP < 1 ? 0 : kP - 1;
                    m = 0;
                } else if (++m == (uint64_t{1} << k)) {
                    put1(1);
                    kP += 2;
                    m = 0;
                }
            } else {
                put_gr(u, static_cast<int>(kR));
                p = u >> kR;
                if (p) {
                    kRP = std::min<uint64_t>(kRP + p - 1, 32 * L);
                } else {
                    kRP = kRP < 2 ? 0 : kRP - 2;
                }
                kP = u ? (kP < 1 ? 0 : kP - 1) : kP + 3;
                m = 0;
            }
        }
        if (n > 0 && k && !u) {
            put1(0);
            put(m, static_cast<int>(kP / L));
        }
        int r = cnt % 8;
        if (r)
            put(0, 8 - r);
        else
            flush();
    }
};

/* ZstdEncoder::Encode (entropy_codec.cc:304-324); empty on error. */
std::vector<uint8_t> zstd_wrap(const std::vector<uint8_t>& src, int level) {
    level = std::min(level, ZSTD_maxCLevel());
    level = std::max(level, ZSTD_minCLevel());
    std::vector<uint8_t> out(ZSTD_compressBound(src.size()));
    size_t r = ZSTD_compress(out.data(), out.size(), src.data(), src.size(), level);
    if (ZSTD_isError(r)) return {};
    out.resize(r);
    return out;
}

/* ColorYUV(r, g, b, false) (include/common/common.h:84-93). */
YUV rgb_to_yuv(float r, float g, float b) {
    YUV c;
    c.y = r * 0.299f + g * 0.587f + b * 0.114f;
    c.u = r * -0.168736f - g * 0.331264f + b * 0.5f + 128;
    c.v = r * 0.5f - g * 0.418688f - b * 0.081312f + 128;
    return c;
}

/* One coded slice. */
struct Slice {
    int timestamp = 0, index = 0;
    uint8_t type = 0, qp = 1;
    uint32_t size = 0;
    float mv[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    std::vector<uint8_t> geo, col;
};

/* Quantise + RLGR + optional zstd of one frame's colours on a built tree
 * (GoP_encoding.cc:172-205).  Fills size/qp/col and the colour zstd bit. */
void encode_colors(RahtOctree& tree, const YUV* colors, uint8_t qp, int zlevel, Slice& s) {
    std::vector<YUV> coef = tree.raht(colors);
    const size_t n = coef.size();
    s.size = static_cast<uint32_t>(n);
    s.qp = qp;
    std::vector<int64_t> q(3 * n);
    for (size_t j = 0; j < n; ++j) {
        q[j] = static_cast<int64_t>(std::round(coef[j].y / qp));
        q[j + n] = static_cast<int64_t>(std::round(coef[j].u / qp));
        q[j + 2 * n] = static_cast<int64_t>(std::round(coef[j].v / qp));
    }
    RlgrWriter w;
    w.encode(q.data(), static_cast<int64_t>(q.size()));
    std::vector<uint8_t> z = zlevel > -1000 ? zstd_wrap(w.out, zlevel) : std::vector<uint8_t>{};
    if (!z.empty() && z.size() < w.out.size()) {
        s.type |= PVVC_TYPE_COLOR_ZSTD;
        s.col = std::move(z);
    } else {
        s.type &= static_cast<uint8_t>(~PVVC_TYPE_COLOR_ZSTD);
        s.col = std::move(w.out);
    }
}

/* .frame image writer (ChangeSliceToFrame, slice_io.cc:258-315); slices must
 * already be sorted by index (:242-243). */
std::vector<uint8_t> frame_image(int timestamp, const std::vector<const Slice*>& sl) {
    std::vector<uint8_t> out;
    auto put = [&](const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        out.insert(out.end(), b, b + n);
    };
    uint32_t S = static_cast<uint32_t>(sl.size());
    put(&timestamp, 4);
    put(&S, 4);
    for (auto* s : sl) put(&s->index, 4);
    for (auto* s : sl) put(&s->type, 1);
    for (auto* s : sl) put(&s->size, 4);
    for (auto* s : sl) put(&s->qp, 1);
    for (auto* s : sl) put(s->mv, 64);
    for (auto* s : sl) {
        uint32_t g = static_cast<uint32_t>(s->geo.size());
        put(&g, 4);
    }
    for (auto* s : sl) put(s->geo.data(), s->geo.size());
    for (auto* s : sl) {
        uint32_t c = static_cast<uint32_t>(s->col.size());
        put(&c, 4);
    }
    for (auto* s : sl) put(s->col.data(), s->col.size());
    return out;
}

/* ------------------------------------------------------------------------ */
/* Synthetic humanoid generator                                             */
/* ------------------------------------------------------------------------ */
uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double u01(uint64_t h) { return (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
uint64_t hash3(uint64_t a, uint64_t b, uint64_t c) { return splitmix(splitmix(splitmix(a) ^ b) ^ c); }
double gauss(uint64_t h) {
    double a = u01(h), b = u01(splitmix(h));
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
}

struct Mat4 {
    double m[16];
    static Mat4 identity() {
        Mat4 r{};
        r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1;
        return r;
    }
    Mat4 operator*(const Mat4& o) const {
        Mat4 r{};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                double s = 0;
                for (int k = 0; k < 4; ++k) s += m[i * 4 + k] * o.m[k * 4 + j];
                r.m[i * 4 + j] = s;
            }
        return r;
    }
    /* inverse of a rigid transform */
    Mat4 rigid_inverse() const {
        Mat4 r = identity();
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[i * 4 + j] = m[j * 4 + i];
        for (int i = 0; i < 3; ++i) r.m[i * 4 + 3] = -(r.m[i * 4] * m[3] + r.m[i * 4 + 1] * m[7] + r.m[i * 4 + 2] * m[11]);
        return r;
    }
    void apply(double x, double y, double z, double* o) const {
        o[0] = m[0] * x + m[1] * y + m[2] * z + m[3];
        o[1] = m[4] * x + m[5] * y + m[6] * z + m[7];
        o[2] = m[8] * x + m[9] * y + m[10] * z + m[11];
    }
};

/* Rotation by `ang` radians about unit axis (ax,ay,az) around point c, then translation t. */
Mat4 rigid(double ax, double ay, double az, double ang, const double* c, const double* t) {
    double s = std::sin(ang), co = std::cos(ang), C = 1 - co;
    double R[9] = {co + ax * ax * C,      ax * ay * C - az * s, ax * az * C + ay * s,
                   ay * ax * C + az * s, co + ay * ay * C,      ay * az * C - ax * s,
                   az * ax * C - ay * s, az * ay * C + ax * s, co + az * az * C};
    Mat4 M = Mat4::identity();
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) M.m[i * 4 + j] = R[i * 3 + j];
        M.m[i * 4 + 3] = c[i] - (R[i * 3] * c[0] + R[i * 3 + 1] * c[1] + R[i * 3 + 2] * c[2]) + t[i];
    }
    return M;
}

struct Prim {
    int kind;      /* 0 ellipsoid, 1 capsule */
    double a[3];   /* ellipsoid centre / capsule end A */
    double b[3];   /* capsule end B */
    double r[3];   /* ellipsoid radii / capsule radius in r[0] */
    double c[3];   /* rotation centre */
    double sdf(const double* p) const {
        if (kind == 0) {
            double q[3], q2[3];
            for (int i = 0; i < 3; ++i) {
                q[i] = (p[i] - a[i]) / r[i];
                q2[i] = (p[i] - a[i]) / (r[i] * r[i]);
            }
            double k0 = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            double k1 = std::sqrt(q2[0] * q2[0] + q2[1] * q2[1] + q2[2] * q2[2]);
            if (k1 < 1e-12) return -std::min(r[0], std::min(r[1], r[2]));
            return k0 * (k0 - 1.0) / k1;
        }
        double pa[3], ba[3];
        for (int i = 0; i < 3; ++i) {
            pa[i] = p[i] - a[i];
            ba[i] = b[i] - a[i];
        }
        double h = (pa[0] * ba[0] + pa[1] * ba[1] + pa[2] * ba[2]) / (ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2]);
        h = std::min(1.0, std::max(0.0, h));
        double d[3] = {pa[0] - ba[0] * h, pa[1] - ba[1] * h, pa[2] - ba[2] * h};
        return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) - r[0];
    }
    void bbox(double* lo, double* hi) const {
        if (kind == 0) {
            double m = std::max(r[0], std::max(r[1], r[2]));
            for (int i = 0; i < 3; ++i) {
                lo[i] = a[i] - m;
                hi[i] = a[i] + m;
            }
        } else {
            for (int i = 0; i < 3; ++i) {
                lo[i] = std::min(a[i], b[i]) - r[0];
                hi[i] = std::max(a[i], b[i]) + r[0];
            }
        }
    }
};

/* Seven primitives of one humanoid in body units (z: feet 0 .. head top ~0.95). */
std::vector<Prim> humanoid(double scale, const double* origin) {
    struct P {
        int kind;
        double a[3], b[3], r[3];
    };
    const P body[7] = {
        {0, {0, 0, 0.55}, {0, 0, 0}, {0.13, 0.085, 0.17}},       /* torso */
        {1, {0, 0, 0.70}, {0, 0, 0.77}, {0.035, 0, 0}},          /* neck */
        {1, {0, 0, 0.83}, {0, 0, 0.88}, {0.065, 0, 0}},          /* head */
        {1, {0.15, 0, 0.68}, {0.33, 0.02, 0.42}, {0.04, 0, 0}},  /* arm L */
        {1, {-0.15, 0, 0.68}, {-0.33, 0.02, 0.42}, {0.04, 0, 0}},/* arm R */
        {1, {0.07, 0, 0.38}, {0.09, 0.01, 0.05}, {0.05, 0, 0}},  /* leg L */
        {1, {-0.07, 0, 0.38}, {-0.09, 0.01, 0.05}, {0.05, 0, 0}},/* leg R */
    };
    std::vector<Prim> out;
    for (const P& p : body) {
        Prim q{};
        q.kind = p.kind;
        for (int i = 0; i < 3; ++i) {
            q.a[i] = origin[i] + p.a[i] * scale;
            q.b[i] = origin[i] + p.b[i] * scale;
            q.r[i] = p.r[i] * scale;
        }
        for (int i = 0; i < 3; ++i) q.c[i] = p.kind == 0 ? q.a[i] : 0.5 * (q.a[i] + q.b[i]);
        out.push_back(q);
    }
    return out;
}

struct SynthCfg {
    int grid_bits, n_frames, gop, qp_i, qp_p, n_humanoids, jitter, patch_min, patch_max, zstd_level, threads;
    double scale, shell, patch_mu, patch_sigma, noise_sigma, max_rot_deg, max_trans;
    uint64_t seed;
    int keep_source; /* keep each frame's source cloud (world xyz + RGB8) */
};

/* Pose of primitive p (global index) at frame t: a seeded rotation of at
 * most max_rot_deg about a random axis through its centre and a
 * translation of at most max_trans voxels per axis. */
Mat4 pose(const SynthCfg& c, const Prim& pr, int p, int t) {
    uint64_t h = hash3(c.seed, 0x5000 + p, t);
    double ax = gauss(h), ay = gauss(splitmix(h + 1)), az = gauss(splitmix(h + 2));
    double nrm = std::sqrt(ax * ax + ay * ay + az * az) + 1e-12;
    double ang = (u01(splitmix(h + 3)) * 2 - 1) * c.max_rot_deg * 3.141592653589793 / 180.0;
    double tr[3];
    for (int i = 0; i < 3; ++i) tr[i] = (u01(splitmix(h + 4 + i)) * 2 - 1) * c.max_trans;
    return rigid(ax / nrm, ay / nrm, az / nrm, ang, pr.c, tr);
}

struct Voxel {
    int32_t x, y, z;
    int prim;
};

/* Surface voxels of the posed primitives: -shell < sdf_union <= jitter term. */
std::vector<Voxel> voxelize(const SynthCfg& c, const std::vector<Prim>& prims, const std::vector<Mat4>& inv_pose) {
    const int G = 1 << c.grid_bits;
    const double t = c.shell;
    std::vector<std::vector<Voxel>> per(prims.size());
    auto sdf_all = [&](const double* w, int& arg) {
        double best = 1e300;
        for (size_t q = 0; q < prims.size(); ++q) {
            double l[3];
            inv_pose[q].apply(w[0], w[1], w[2], l);
            double s = prims[q].sdf(l);
            if (s < best) {
                best = s;
                arg = static_cast<int>(q);
            }
        }
        return best;
    };
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (;;) {
            size_t p = next.fetch_add(1);
            if (p >= prims.size()) break;
            double lo[3], hi[3];
            prims[p].bbox(lo, hi);
            /* posed AABB: rotation <= a few degrees, pad generously */
            double pad = 4 + c.max_trans + 0.06 * std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
            int b0[3], b1[3];
            for (int i = 0; i < 3; ++i) {
                b0[i] = std::max(0, static_cast<int>(std::floor(lo[i] - pad)));
                b1[i] = std::min(G - 1, static_cast<int>(std::ceil(hi[i] + pad)));
            }
            Mat4 fwd = inv_pose[p].rigid_inverse();
            (void)fwd;
            for (int x = b0[0]; x <= b1[0]; ++x)
                for (int y = b0[1]; y <= b1[1]; ++y)
                    for (int z = b0[2]; z <= b1[2]; ++z) {
                        double w[3] = {double(x), double(y), double(z)}, l[3];
                        inv_pose[p].apply(w[0], w[1], w[2], l);
                        double s = prims[p].sdf(l);
                        if (s > 1.5 || s <= -t - 1.5) continue;
                        int arg = 0;
                        double su = sdf_all(w, arg);
                        if (arg != static_cast<int>(p)) continue;
                        double jit = 0;
                        if (c.jitter) jit = (u01(hash3(c.seed ^ 0xABCDEF, (uint64_t(x) << 32) | uint32_t(y), z)) * 2 - 1) * c.jitter;
                        if (su <= jit && su > jit - t) per[p].push_back(Voxel{x, y, z, static_cast<int>(p)});
                    }
        }
    };
    int nt = std::max(1, c.threads);
    std::vector<std::thread> th;
    for (int i = 1; i < nt; ++i) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
    std::vector<Voxel> all;
    for (auto& v : per) all.insert(all.end(), v.begin(), v.end());
    return all;
}

/* Recursive median split with lognormal target sizes. */
void cut(std::vector<Voxel>& v, size_t lo, size_t hi, const SynthCfg& c, uint64_t key,
         std::vector<std::pair<size_t, size_t>>& out) {
    size_t n = hi - lo;
    double target = std::exp(c.patch_mu + c.patch_sigma * gauss(splitmix(key)));
    target = std::min<double>(c.patch_max, std::max<double>(c.patch_min, target));
    if (n <= static_cast<size_t>(target) || n < 2) {
        out.emplace_back(lo, hi);
        return;
    }
    int mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (size_t i = lo; i < hi; ++i) {
        const int q[3] = {v[i].x, v[i].y, v[i].z};
        for (int a = 0; a < 3; ++a) mn[a] = std::min(mn[a], q[a]), mx[a] = std::max(mx[a], q[a]);
    }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (mx[a] - mn[a] > mx[axis] - mn[axis]) axis = a;
    size_t mid = lo + n / 2;
    auto coord = [axis](const Voxel& q) { return axis == 0 ? q.x : (axis == 1 ? q.y : q.z); };
    std::nth_element(v.begin() + lo, v.begin() + mid, v.begin() + hi,
                     [&](const Voxel& a, const Voxel& b) { return coord(a) < coord(b); });
    cut(v, lo, mid, c, splitmix(key * 2 + 1), out);
    cut(v, mid, hi, c, splitmix(key * 2 + 2), out);
}

/* Colour field: smooth sinusoids in world space, wavelength scaled with the grid. */
void field(const SynthCfg& c, double x, double y, double z, double* rgb) {
    double lam = (1 << c.grid_bits) / 16.0;
    const double k = 6.283185307179586 / lam;
    rgb[0] = 128 + 80 * std::sin(k * (0.8 * x + 0.3 * z)) + 30 * std::sin(k * 2.3 * y + 1.0);
    rgb[1] = 128 + 70 * std::sin(k * (0.5 * y - 0.7 * z) + 2.0) + 25 * std::cos(k * 1.7 * x);
    rgb[2] = 128 + 75 * std::cos(k * (0.6 * x + 0.6 * y) + 0.5) + 20 * std::sin(k * 2.9 * z + 3.0);
}

}  // namespace

/* ======================================================================== */
/* C ABI                                                                     */
/* ======================================================================== */
extern "C" {

typedef struct pvvc_synth_config {
    int grid_bits, n_frames, gop, qp_i, qp_p, n_humanoids, jitter, patch_min, patch_max, zstd_level, threads;
    double scale, shell, patch_mu, patch_sigma, noise_sigma, max_rot_deg, max_trans;
    uint64_t seed;
    int keep_source;
} pvvc_synth_config;

struct pvvc_stream {
    std::vector<std::vector<uint8_t>> frames;
    std::vector<uint64_t> points;
    std::vector<int> patches;
    /* source clouds (keep_source): per frame, world xyz f32 x3 and RGB8 x3 in
     * patch order, the cloud the frame was encoded from */
    std::vector<std::vector<float>> src_xyz;
    std::vector<std::vector<uint8_t>> src_rgb;
};

/* Generates n_frames frames; returns 0 or a pvvc_status. */
int pvvc_synth_generate(const pvvc_synth_config* cfg_in, pvvc_stream** out) {
    if (!cfg_in || !out || cfg_in->n_frames <= 0 || cfg_in->gop <= 0 || cfg_in->grid_bits < 4 ||
        cfg_in->grid_bits > 16 || cfg_in->qp_i <= 0 || cfg_in->qp_p <= 0 || cfg_in->qp_i > 255 || cfg_in->qp_p > 255)
        return PVVC_ERR_INVALID_ARG;
    SynthCfg c;
    std::memcpy(&c, cfg_in, sizeof(c));
    static_assert(sizeof(SynthCfg) == sizeof(pvvc_synth_config), "config layout");
    const int G = 1 << c.grid_bits;
    const int H = std::max(1, c.n_humanoids);
    /* Humanoid placement: a row-major grid of H figures on the z=0 plane. */
    int cols = static_cast<int>(std::ceil(std::sqrt(static_cast<double>(H))));
    int rows = (H + cols - 1) / cols;
    double cell = static_cast<double>(G) / std::max(cols, rows);
    double scale = c.scale * cell;
    std::vector<Prim> prims;
    for (int h = 0; h < H; ++h) {
        double origin[3] = {(h % cols + 0.5) * cell, (h / cols + 0.5) * cell, 0.04 * cell};
        if (H == 1) origin[0] = origin[1] = G / 2.0;
        auto hp = humanoid(scale, origin);
        prims.insert(prims.end(), hp.begin(), hp.end());
    }
    const int P = static_cast<int>(prims.size());
    auto* st = new pvvc_stream();
    st->frames.resize(c.n_frames);
    st->points.resize(c.n_frames, 0);
    st->patches.resize(c.n_frames, 0);
    if (c.keep_source) {
        st->src_xyz.resize(c.n_frames);
        st->src_rgb.resize(c.n_frames);
    }
    int nt = std::max(1, c.threads);
    for (int t0 = 0; t0 < c.n_frames; t0 += c.gop) {
        const int F = std::min(c.gop, c.n_frames - t0);
        std::vector<std::vector<Mat4>> poses(F, std::vector<Mat4>(P));
        for (int f = 0; f < F; ++f)
            for (int p = 0; p < P; ++p) poses[f][p] = pose(c, prims[p], p, t0 + f);
        std::vector<Mat4> inv0(P);
        for (int p = 0; p < P; ++p) inv0[p] = poses[0][p].rigid_inverse();
        std::vector<Voxel> vox = voxelize(c, prims, inv0);
        /* group by primitive (voxelize emits them grouped), cut each group */
        std::vector<std::pair<size_t, size_t>> patches;
        size_t i0 = 0;
        while (i0 < vox.size()) {
            size_t i1 = i0;
            while (i1 < vox.size() && vox[i1].prim == vox[i0].prim) ++i1;
            cut(vox, i0, i1, c, hash3(c.seed, 0x77, (uint64_t(t0) << 16) | uint64_t(vox[i0].prim)), patches);
            i0 = i1;
        }
        const int S = static_cast<int>(patches.size());
        std::vector<std::vector<Slice>> coded(S, std::vector<Slice>(F));
        std::vector<size_t> src_off(S + 1, 0);
        for (int s = 0; s < S; ++s) src_off[s + 1] = src_off[s] + (patches[s].second - patches[s].first);
        if (c.keep_source)
            for (int f = 0; f < F; ++f) {
                st->src_xyz[t0 + f].resize(3 * src_off[S]);
                st->src_rgb[t0 + f].resize(3 * src_off[S]);
            }
        std::atomic<int> next{0};
        auto work = [&]() {
            std::vector<XYZ> pts;
            std::vector<YUV> col0, col;
            for (;;) {
                int s = next.fetch_add(1);
                if (s >= S) break;
                size_t lo = patches[s].first, hi = patches[s].second;
                const int n = static_cast<int>(hi - lo);
                const int prim = vox[lo].prim;
                pts.resize(n);
                for (int i = 0; i < n; ++i)
                    pts[i] = XYZ{float(vox[lo + i].x), float(vox[lo + i].y), float(vox[lo + i].z)};
                RahtOctree tree;
                tree.make_tree(pts.data(), n);
                std::vector<uint8_t> geo = tree.get_octree();
                col0.resize(n);
                col.resize(n);
                for (int f = 0; f < F; ++f) {
                    const int t = t0 + f;
                    Slice& sl = coded[s][f];
                    sl.timestamp = t;
                    sl.index = s;
                    Mat4 mv = poses[f][prim] * inv0[prim];
                    for (int k = 0; k < 16; ++k) sl.mv[k] = static_cast<float>(mv.m[k]);
                    if (f == 0)
                        for (int k = 0; k < 16; ++k) sl.mv[k] = (k % 5 == 0) ? 1.0f : 0.0f;
                    for (int i = 0; i < n; ++i) {
                        double w[3], rgb[3];
                        mv.apply(pts[i].x, pts[i].y, pts[i].z, w);
                        field(c, w[0], w[1], w[2], rgb);
                        float ch[3];
                        for (int k = 0; k < 3; ++k) {
                            double v = rgb[k] + c.noise_sigma * gauss(hash3(c.seed ^ 0x1234, (uint64_t(t) << 24) ^ uint64_t(lo + i), k));
                            v = std::min(255.0, std::max(0.0, std::round(v)));
                            ch[k] = static_cast<float>(v);
                        }
                        if (c.keep_source) {
                            const size_t o = 3 * (src_off[s] + i);
                            for (int k = 0; k < 3; ++k) {
                                st->src_xyz[t][o + k] = static_cast<float>(w[k]);
                                st->src_rgb[t][o + k] = static_cast<uint8_t>(ch[k]);
                            }
                        }
                        YUV y = rgb_to_yuv(ch[0], ch[1], ch[2]);
                        if (f == 0) {
                            col0[i] = y;
                            col[i] = y;
                        } else { /* GoP_encoding.cc:145-149 */
                            col[i].y = y.y - col0[i].y;
                            col[i].u = y.u - col0[i].u;
                            col[i].v = y.v - col0[i].v;
                        }
                    }
                    if (f == 0) {
                        sl.type = PVVC_TYPE_VALID;
                        std::vector<uint8_t> z = zstd_wrap(geo, c.zstd_level);
                        if (!z.empty() && z.size() < geo.size()) {
                            sl.type |= PVVC_TYPE_GEO_ZSTD;
                            sl.geo = std::move(z);
                        } else {
                            sl.geo = geo;
                        }
                    } else {
                        sl.type = PVVC_TYPE_VALID | PVVC_TYPE_PREDICT;
                    }
                    encode_colors(tree, col.data(), static_cast<uint8_t>(f == 0 ? c.qp_i : c.qp_p), c.zstd_level, sl);
                }
            }
        };
        std::vector<std::thread> th;
        for (int i = 1; i < nt; ++i) th.emplace_back(work);
        work();
        for (auto& x : th) x.join();
        for (int f = 0; f < F; ++f) {
            std::vector<const Slice*> sl(S);
            uint64_t pts = 0;
            for (int s = 0; s < S; ++s) {
                sl[s] = &coded[s][f];
                pts += coded[s][f].size;
            }
            st->frames[t0 + f] = frame_image(t0 + f, sl);
            st->points[t0 + f] = pts;
            st->patches[t0 + f] = S;
        }
    }
    *out = st;
    return PVVC_OK;
}

int pvvc_stream_frames(const pvvc_stream* s) { return s ? static_cast<int>(s->frames.size()) : 0; }
const uint8_t* pvvc_stream_frame_bytes(const pvvc_stream* s, int i, size_t* nbytes) {
    if (!s || i < 0 || i >= static_cast<int>(s->frames.size())) return nullptr;
    if (nbytes) *nbytes = s->frames[i].size();
    return s->frames[i].data();
}
uint64_t pvvc_stream_frame_points(const pvvc_stream* s, int i) {
    return (s && i >= 0 && i < static_cast<int>(s->points.size())) ? s->points[i] : 0;
}
void pvvc_stream_free(pvvc_stream* s) { delete s; }

/* Source cloud of frame i (generated with keep_source): returns the point
 * count and sets *xyz / *rgb, or 0 when none was kept. */
uint64_t pvvc_stream_frame_source(const pvvc_stream* s, int i, const float** xyz, const uint8_t** rgb) {
    if (!s || i < 0 || i >= static_cast<int>(s->src_xyz.size()) || s->src_xyz[i].empty()) return 0;
    if (xyz) *xyz = s->src_xyz[i].data();
    if (rgb) *rgb = s->src_rgb[i].data();
    return s->src_xyz[i].size() / 3;
}

/* Writes frame i to `path` (the reference names them <seq>_<i>.frame). */
int pvvc_stream_write_frame(const pvvc_stream* s, int i, const char* path) {
    size_t n = 0;
    const uint8_t* b = pvvc_stream_frame_bytes(s, i, &n);
    if (!b || !path) return PVVC_ERR_INVALID_ARG;
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return PVVC_ERR_IO;
    size_t w = std::fwrite(b, 1, n, fp);
    std::fclose(fp);
    return w == n ? PVVC_OK : PVVC_ERR_IO;
}

/* RAHTOctree::MakeTree + GetOctree of one patch; returns the byte count
 * (copies at most cap) and the leaf count, or -1. */
int64_t pvvc_enc_octree(const float* xyz, int n, uint8_t* out, int64_t cap, int* leaves) {
    std::vector<XYZ> p(n);
    for (int i = 0; i < n; ++i) p[i] = XYZ{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    RahtOctree t;
    if (!t.make_tree(p.data(), n)) return -1;
    auto g = t.get_octree();
    if (leaves) *leaves = static_cast<int>(t.leaves());
    if (out) std::memcpy(out, g.data(), std::min<size_t>(g.size(), static_cast<size_t>(cap)));
    return static_cast<int64_t>(g.size());
}

/* Forward RAHT of one patch's YUV colours (n x 3): coefficients in stream
 * order (leaves x 3, row per coefficient).  Returns the leaf count or -1. */
int64_t pvvc_enc_raht(const float* xyz, int n, const float* yuv, float* coef_out) {
    std::vector<XYZ> p(n);
    std::vector<YUV> c(n);
    for (int i = 0; i < n; ++i) {
        p[i] = XYZ{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        c[i].y = yuv[3 * i], c[i].u = yuv[3 * i + 1], c[i].v = yuv[3 * i + 2];
    }
    RahtOctree t;
    if (!t.make_tree(p.data(), n)) return -1;
    auto r = t.raht(c.data());
    for (size_t i = 0; i < r.size() && coef_out; ++i) {
        coef_out[3 * i] = r[i].y;
        coef_out[3 * i + 1] = r[i].u;
        coef_out[3 * i + 2] = r[i].v;
    }
    return static_cast<int64_t>(r.size());
}

/* RLGR encode; returns the byte count (copies at most cap). */
int64_t pvvc_enc_rlgr(const int64_t* sym, int64_t n, uint8_t* out, int64_t cap) {
    RlgrWriter w;
    w.encode(sym, n);
    if (out) std::memcpy(out, w.out.data(), std::min<size_t>(w.out.size(), static_cast<size_t>(cap)));
    return static_cast<int64_t>(w.out.size());
}

/* One patch's colour stream on the CPU, the reference for the GPU encoder:
 * builds the octree of xyz (n points), writes its bytes to geo_out, the
 * per-leaf YUV (points of a leaf averaged, raht_octree.cc:189-200) in leaf
 * order to leaf_yuv (leaves x 3), and the RLGR bytes of the quantised RAHT
 * coefficients (no zstd) to col_out.  Returns the colour byte count, or -1;
 * *leaves and *geo_size receive the leaf count and geometry size. */
int64_t pvvc_enc_patch_colors(const float* xyz, int n, const float* yuv, int qp, uint8_t* geo_out, int64_t geo_cap,
                              int64_t* geo_size, float* leaf_yuv, int* leaves, uint8_t* col_out, int64_t col_cap) {
    if (n <= 0 || qp <= 0) return -1;
    std::vector<XYZ> p(n);
    std::vector<YUV> c(n);
    for (int i = 0; i < n; ++i) {
        p[i] = XYZ{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        c[i] = YUV{yuv[3 * i], yuv[3 * i + 1], yuv[3 * i + 2]};
    }
    RahtOctree t;
    if (!t.make_tree(p.data(), n)) return -1;
    std::vector<uint8_t> geo = t.get_octree();
    if (geo_size) *geo_size = static_cast<int64_t>(geo.size());
    if (geo_out) std::memcpy(geo_out, geo.data(), std::min<size_t>(geo.size(), static_cast<size_t>(geo_cap)));
    const size_t L = t.leaves();
    if (leaves) *leaves = static_cast<int>(L);
    if (leaf_yuv) {
        size_t k = 0;
        for (auto& nd : t.tree.back()) {
            YUV a;
            for (int i : nd.index) {
                a.y += c[i].y;
                a.u += c[i].u;
                a.v += c[i].v;
            }
            const int m = static_cast<int>(nd.index.size());
            leaf_yuv[3 * k] = a.y / m;
            leaf_yuv[3 * k + 1] = a.u / m;
            leaf_yuv[3 * k + 2] = a.v / m;
            ++k;
        }
    }
    Slice sl;
    encode_colors(t, c.data(), static_cast<uint8_t>(qp), -1000, sl);
    if (col_out) std::memcpy(col_out, sl.col.data(), std::min<size_t>(sl.col.size(), static_cast<size_t>(col_cap)));
    return static_cast<int64_t>(sl.col.size());
}

/* The CPU colour encoder on a given octree (its bytes, as in an intra slice):
 * the tree is rebuilt from the occupancy bytes (LoadTreeCore / MakeTree of
 * the decoder, invert_raht_octree.cc:122-254), leaf k holding point k of
 * leaf_yuv (n leaves); then RAHT, quantisation and RLGR as encode_colors.
 * Returns the colour byte count, or -1 (bad geometry / leaf count != n). */
int64_t pvvc_enc_colors_on_tree(const uint8_t* geo, int64_t geo_size, const float* leaf_yuv, int n, int qp,
                                uint8_t* col_out, int64_t col_cap) {
    if (!geo || geo_size < 25 || n <= 0 || qp <= 0) return -1;
    RahtOctree t;
    t.height = geo[24];
    if (t.height <= 0 || t.height > 64) return -1;
    t.tree.assign(t.height, {});
    t.tree[0].emplace_back();
    size_t at = 25;
    for (int l = 0; l + 1 < t.height; ++l) {
        for (auto& nd : t.tree[l]) {
            if (at >= static_cast<size_t>(geo_size)) return -1;
            nd.value = geo[at++];
            if (nd.value == 0) return -1;
            nd.index.assign(8, -1);
            for (int i = 0; i < 8; ++i)
                if (nd.value & kNodeValue[i]) {
                    nd.index[i] = static_cast<int>(t.tree[l + 1].size());
                    t.tree[l + 1].emplace_back();
                }
        }
    }
    if (static_cast<int>(t.tree.back().size()) != n) return -1;
    for (int k = 0; k < n; ++k) {
        FNode& lf = t.tree.back()[k];
        lf.value = 0xff;
        lf.index.assign(1, k);
        lf.weight[1] = 1;
    }
    for (int l = t.height - 2; l >= 0; --l)
        for (auto& nd : t.tree[l]) {
            for (int i = 0; i < 8; ++i)
                if (nd.index[i] != -1) nd.weight[i + 8] = t.tree[l + 1][nd.index[i]].weight[1];
            for (int i = 7; i > 0; --i) nd.weight[i] = nd.weight[kNodeWeight[i][0]] + nd.weight[kNodeWeight[i][1]];
        }
    std::vector<YUV> c(n);
    for (int k = 0; k < n; ++k) c[k] = YUV{leaf_yuv[3 * k], leaf_yuv[3 * k + 1], leaf_yuv[3 * k + 2]};
    Slice sl;
    encode_colors(t, c.data(), static_cast<uint8_t>(qp), -1000, sl);
    if (col_out) std::memcpy(col_out, sl.col.data(), std::min<size_t>(sl.col.size(), static_cast<size_t>(col_cap)));
    return static_cast<int64_t>(sl.col.size());
}

/* Encodes one GoP of one patch: common geometry xyz (n x 3), RGB8 colours
 * for F frames (F x n x 3), F motion matrices (row-major 4x4); writes F
 * slices into a frame-image-ready form via callbacks of the caller:
 * returns slices as a flat buffer: per slice [u8 type][u8 qp][u32 size]
 * [u32 geo][u32 col][geo bytes][col bytes].  Returns total bytes or -1. */
int64_t pvvc_enc_gop_patch(const float* xyz, int n, const uint8_t* rgb, int F, int qp_i, int qp_p, int zstd_level,
                           uint8_t* out, int64_t cap) {
    std::vector<XYZ> p(n);
    for (int i = 0; i < n; ++i) p[i] = XYZ{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    RahtOctree t;
    if (!t.make_tree(p.data(), n)) return -1;
    std::vector<uint8_t> geo = t.get_octree(), res;
    std::vector<YUV> c0(n), c(n);
    for (int f = 0; f < F; ++f) {
        for (int i = 0; i < n; ++i) {
            const uint8_t* q = rgb + (size_t(f) * n + i) * 3;
            YUV y = rgb_to_yuv(q[0], q[1], q[2]);
            if (f == 0) {
                c0[i] = y;
                c[i] = y;
            } else {
                c[i].y = y.y - c0[i].y;
                c[i].u = y.u - c0[i].u;
                c[i].v = y.v - c0[i].v;
            }
        }
        Slice sl;
        if (f == 0) {
            sl.type = PVVC_TYPE_VALID;
            std::vector<uint8_t> z = zstd_level > -1000 ? zstd_wrap(geo, zstd_level) : std::vector<uint8_t>{};
            if (!z.empty() && z.size() < geo.size()) {
                sl.type |= PVVC_TYPE_GEO_ZSTD;
                sl.geo = std::move(z);
            } else {
                sl.geo = geo;
            }
        } else {
            sl.type = PVVC_TYPE_VALID | PVVC_TYPE_PREDICT;
        }
        encode_colors(t, c.data(), static_cast<uint8_t>(f == 0 ? qp_i : qp_p), zstd_level, sl);
        uint32_t g = static_cast<uint32_t>(sl.geo.size()), cc = static_cast<uint32_t>(sl.col.size());
        res.push_back(sl.type);
        res.push_back(sl.qp);
        res.insert(res.end(), reinterpret_cast<uint8_t*>(&sl.size), reinterpret_cast<uint8_t*>(&sl.size) + 4);
        res.insert(res.end(), reinterpret_cast<uint8_t*>(&g), reinterpret_cast<uint8_t*>(&g) + 4);
        res.insert(res.end(), reinterpret_cast<uint8_t*>(&cc), reinterpret_cast<uint8_t*>(&cc) + 4);
        res.insert(res.end(), sl.geo.begin(), sl.geo.end());
        res.insert(res.end(), sl.col.begin(), sl.col.end());
    }
    if (out) std::memcpy(out, res.data(), std::min<size_t>(res.size(), static_cast<size_t>(cap)));
    return static_cast<int64_t>(res.size());
}

}  // extern "C"
