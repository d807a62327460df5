/*
 * pvvc_oracle.cpp -- CPU restatement of patchVVC's per-patch GoP decoder.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * decoder in patchvvc-plus_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function restates one part of the reference CPU decoder and cites the
 * reference file:line it follows (paths relative to the reference repo root).
 * Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math) so that the
 * FP32 operation order is the reference's: x86-64 SSE scalar float, no FMA.
 *
 * The forward path (raht_forward, quantise, orc_color_encode) restates the
 * colour half of the encoder, the checker of the GPU colour encoder.
 *
 * Pinning (see tests/test_oracle.py): RLGR by hand-derived known-answer
 * vectors and encode->decode round trips; the octree by closed-form leaf
 * centres (header (1,1,1)/(2,2,2)/h=2, occupancy FF and 81, h=1) and by a
 * brute-force recursive expansion; the inverse RAHT by the two-leaf closed
 * form, the constant-colour DC identity, and forward->inverse identity; the
 * patch motion by exact closed forms of non-symmetric integer / power-of-two
 * matrices; the container and the .slice files by hand-packed byte
 * fixtures (a non-symmetric motion matrix fixes the storage orders); the quality metric
 * (MSE::Compute) by brute-force neighbours and closed forms.  The reference itself holds no
 * golden vectors for this path.
 *
 * Defined behaviour where the reference is undefined (documented, shared with
 * the HIP decoder):
 *   - RLGR reads past the end of the colour bytes see zero bits and set the
 *     BITSTREAM_OVERRUN status (reference: negative shift, UB).
 *   - ConvertRGB saturates to [0,255] (reference: float->uint8 cast, UB
 *     outside the range); out-of-range channels are counted.
 *   - A failed intra slice invalidates the patch state (reference: partially
 *     updated members).
 *   - Truncated occupancy bytes and tree height 0 are BAD_SLICE (reference:
 *     out-of-bounds vector access); so is an internal occupancy byte of 0,
 *     which the reference writer never emits.
 */
#include <zstd.h>

#include <algorithm>
#include <array>
#include <cerrno>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <queue>
#include <string>
#include <thread>
#include <vector>

namespace {

/* Status codes, shared with include/pvvc_hip.h. */
enum {
    ORC_OK = 0,
    ORC_ERR_BAD_SLICE = 1,
    ORC_ERR_EMPTY_REFERENCE = 2,
    ORC_ERR_UNMATCHED_SIZE = 3,
    ORC_ERR_ZSTD = 4,
    ORC_ERR_BITSTREAM_OVERRUN = 5,
    ORC_ERR_SYMBOL_RANGE = 6,
};

/* ------------------------------------------------------------------------ */
/* RLGR constants: include/common/entropy_codec.h:38-48                     */
/* ------------------------------------------------------------------------ */
constexpr uint64_t L = 4, D1 = 1, D0 = 1, U1 = 2, U0 = 3;
constexpr int HALF = 32;

/* Sign2Unsign / Unsign2Sign: entropy_codec.h:55-81 */
inline uint64_t sign2unsign(int64_t x) {
    uint64_t a;
    if (x < 0) {
        a = static_cast<uint64_t>(-x);
        a = (a << 1) - 1;
    } else {
        a = static_cast<uint64_t>(x);
        a <<= 1;
    }
    return a;
}
inline int64_t unsign2sign(uint64_t x) {
    int64_t a = static_cast<int64_t>(x >> 1);
    return (x & 1) ? -a - 1 : a;
}
inline uint64_t mask(int k) { return (uint64_t{1} << k) - 1; }

/* RLGR encoder: entropy_codec.cc:22-159 (RLGREncoder::Flush/Write/GRWrite/Encode/GetResult). */
struct RlgrEnc {
    uint64_t buf = 0;
    int cnt = 0;
    std::vector<uint8_t> out;
    void flush() { /* :22-29 */
        while (cnt >= 8) {
            cnt -= 8;
            out.push_back(static_cast<uint8_t>((buf >> cnt) & 0xFF));
        }
    }
    void write1(uint8_t x) { /* :31-36 */
        buf <<= 1;
        buf |= (x & 1);
        ++cnt;
        flush();
    }
    void write(uint64_t x, int bits) { /* :38-47 */
        if (bits > 56) {
            write(x >> HALF, bits - HALF);
            write(x & mask(HALF), HALF);
            return;
        }
        buf = (buf << bits) + x;
        cnt += bits;
        flush();
    }
    void gr_write(uint64_t x, int bits) { /* :49-63 */
        uint64_t p = x >> bits;
        if (p < static_cast<uint64_t>(HALF)) {
            write(mask(static_cast<int>(p) + 1) - 1, static_cast<int>(p) + 1);
            write(x & mask(bits), bits);
        } else {
            write(mask(HALF), HALF);
            write(x, HALF);
        }
    }
    void encode(const int64_t* data, int64_t n) { /* :65-148 */
        uint64_t u = 0, kP = 0, kRP = 2 * L, m = 0, k = 0, kR = 0, p = 0;
        for (int64_t i = 0; i < n; ++i) {
            u = sign2unsign(data[i]);
            k = kP / L;
            kR = kRP / L;
            if (k) {
                if (u) {
                    --u;
                    write1(0);
                    write(m, static_cast<int>(k));
                    gr_write(u, static_cast<int>(kR));
                    p = u >> kR;
                    if (p) {
                        kRP += p - 1;
                        kRP = kRP > HALF * L ? HALF * L : kRP;
                    } else {
                        kRP = kRP < 2 ? 0 : kRP - 2;
                    }
                    kP = kP < D1 ? 0 : kP - D1;
                    m = 0;
                } else {
                    ++m;
                    if (m == (uint64_t{1} << k)) {
                        write1(1);
                        kP += U1;
                        m = 0;
                    }
                }
            } else {
                gr_write(u, static_cast<int>(kR));
                p = u >> kR;
                if (p) {
                    kRP = kRP + p - 1;
                    kRP = kRP > HALF * L ? HALF * L : kRP;
                } else {
                    kRP = kRP < 2 ? 0 : kRP - 2;
                }
                if (u) {
                    kP = kP < D0 ? 0 : kP - D0;
                } else {
                    kP += U0;
                }
                m = 0;
            }
        }
        /* :144-147 -- k and u are the values left by the last iteration */
        if (n > 0 && k && !u) {
            write1(0);
            write(m, static_cast<int>(kP / L));
        }
    }
    void finish() { /* GetResult :150-158 */
        int r = cnt % 8;
        if (r)
            write(0, 8 - r);
        else
            flush();
    }
};

/* RLGR decoder: entropy_codec.cc:163-283 (Fill/Read/Read(bits)/GRRead/Decode).
 * Reads past the end see zero bits and set `overrun` (the reference's Read()
 * shifts by a negative count there). */
struct RlgrDec {
    const uint8_t* p = nullptr;
    const uint8_t* end = nullptr;
    uint64_t buf = 0;
    int cnt = 0;
    bool overrun = false;
    bool bad_run = false; /* run exponent k > 30: 1 << k would leave the symbol range */
    void fill() { /* :163-176 */
        while (cnt <= 56) {
            if (p != end) {
                buf = (buf << 8) + *p++;
                cnt += 8;
            } else {
                return;
            }
        }
    }
    /* Pads with zero bytes once the data is exhausted, marking the overrun. */
    void ensure(int bits) {
        fill();
        while (cnt < bits) {
            buf <<= 8;
            cnt += 8;
            overrun = true;
        }
    }
    uint64_t read1() { /* :178-184 */
        if (cnt == 0) fill();
        if (cnt == 0) ensure(1);
        --cnt;
        return (buf >> cnt) & 1;
    }
    uint64_t read(int bits) { /* :186-194 */
        if (bits > 56) {
            uint64_t d = read(bits - HALF) << HALF;
            return d + read(HALF);
        }
        if (bits == 0) { /* (buffer_ >> cnt_) & 0 */
            fill();
            return 0;
        }
        ensure(bits);
        cnt -= bits;
        return (buf >> cnt) & mask(bits);
    }
    uint64_t gr_read(int bits) { /* :196-205 */
        uint64_t q = 0;
        while (read1()) {
            q++;
            if (q >= static_cast<uint64_t>(HALF)) return read(HALF);
        }
        return (q << bits) + read(bits);
    }
    /* Decode :207-283.  Writes at most `size` symbols into out (zeros past a
     * final run are dropped, as the reference's consumers ignore them). */
    int64_t decode(const uint8_t* data, int64_t nbytes, int64_t size, int64_t* out) {
        p = data;
        end = data + nbytes;
        buf = 0;
        cnt = 0;
        overrun = false;
        bad_run = false;
        uint64_t u, kP = 0, kRP = 2 * L, m = 0, k, kR, q;
        int64_t n = 0;
        auto emit = [&](int64_t v) {
            if (n < size) out[n] = v;
            ++n;
        };
        while (n < size) {
            k = kP / L;
            kR = kRP / L;
            if (k) {
                m = 0;
                while (read1()) {
                    m += uint64_t{1} << k;
                    kP += U1;
                    k = kP / L;
                    if (k > 30) {
                        bad_run = true;
                        break;
                    }
                    if (overrun) break;
                }
                if (bad_run) break;
                m += read(static_cast<int>(k));
                if (overrun) break;
                /* emit m zeros; a run past `size` only lengthens the tail the
                 * reference's consumers never read */
                uint64_t room = static_cast<uint64_t>(size - n);
                if (m >= room) {
                    for (int64_t z = n; z < size; ++z) out[z] = 0;
                    n = size;
                    break;
                }
                for (uint64_t z = 0; z < m; ++z) emit(0);
                if (n >= size) break;
                u = gr_read(static_cast<int>(kR));
                emit(unsign2sign(u + 1));
                q = u >> kR;
                if (q) {
                    kRP += q - 1;
                    kRP = kRP > HALF * L ? HALF * L : kRP;
                } else {
                    kRP = kRP < 2 ? 0 : kRP - 2;
                }
                kP = kP < D1 ? 0 : kP - D1;
            } else {
                u = gr_read(static_cast<int>(kR));
                emit(unsign2sign(u));
                q = u >> kR;
                if (q) {
                    kRP = kRP + q - 1;
                    kRP = kRP > HALF * L ? HALF * L : kRP;
                } else {
                    kRP = kRP < 2 ? 0 : kRP - 2;
                }
                if (u) {
                    kP = kP < D0 ? 0 : kP - D0;
                } else {
                    kP += U0;
                }
            }
            if (overrun) break;
        }
        for (int64_t i = n; i < size; ++i) out[i] = 0;
        return std::min(n, size);
    }
};

/* ------------------------------------------------------------------------ */
/* Octree / RAHT: include/octree/octree.h, src/octree/octree_base.cc         */
/* ------------------------------------------------------------------------ */
const uint8_t NodeValue[8] = {0x80, 0x40, 0x20, 0x10, 0x08, 0x04, 0x02, 0x01}; /* octree.h:76 */
const int NodeWeight[8][2] = {{0, 0}, {2, 3}, {4, 6}, {5, 7}, {8, 12}, {9, 13}, {10, 14}, {11, 15}}; /* octree.h:79 */

struct YUV {
    float y = 0, u = 0, v = 0;
};
struct YUV64 {
    double y = 0, u = 0, v = 0;
};
struct XYZ {
    float x = 0, y = 0, z = 0;
};

/* SubSpaceCenter: octree_base.cc:102-120 (the subrange argument is half the parent range). */
XYZ sub_space_center(XYZ c, XYZ r, int pos) {
    float hx = r.x / 2.0f, hy = r.y / 2.0f, hz = r.z / 2.0f;
    XYZ o;
    o.x = (pos & 4) ? c.x - hx : c.x + hx;
    o.y = (pos & 2) ? c.y - hy : c.y + hy;
    o.z = (pos & 1) ? c.z - hz : c.z + hz;
    return o;
}

/* InvertHaarTransform: octree_base.cc:186-206.  FP32, T built as the reference
 * builds it; G = g0*T00 + g1*T01, H = g0*T10 + g1*T11 per channel. */
void inv_haar(int w0, int w1, const YUV& g0, const YUV& g1, YUV& r0, YUV& r1) {
    if (w0 == 0 && w1 == 0) {
        r0 = YUV{};
        r1 = YUV{};
        return;
    }
    float base = std::sqrt(static_cast<float>(w0 + w1));
    float t00 = std::sqrt(static_cast<float>(w0)) / base;
    float t01 = -std::sqrt(static_cast<float>(w1)) / base;
    float t10 = std::sqrt(static_cast<float>(w1)) / base;
    float t11 = std::sqrt(static_cast<float>(w0)) / base;
    YUV G, H;
    G.y = g0.y * t00 + g1.y * t01;
    G.u = g0.u * t00 + g1.u * t01;
    G.v = g0.v * t00 + g1.v * t01;
    H.y = g0.y * t10 + g1.y * t11;
    H.u = g0.u * t10 + g1.u * t11;
    H.v = g0.v * t10 + g1.v * t11;
    r0 = G;
    r1 = H;
}

/* FP64 counterpart of inv_haar for the error oracle (exact-arithmetic
 * reference of the same transform). */
void inv_haar64(int w0, int w1, const YUV64& g0, const YUV64& g1, YUV64& r0, YUV64& r1) {
    if (w0 == 0 && w1 == 0) {
        r0 = YUV64{};
        r1 = YUV64{};
        return;
    }
    double base = std::sqrt(static_cast<double>(w0 + w1));
    double a = std::sqrt(static_cast<double>(w0)) / base;
    double c = std::sqrt(static_cast<double>(w1)) / base;
    r0.y = g0.y * a - g1.y * c;
    r0.u = g0.u * a - g1.u * c;
    r0.v = g0.v * a - g1.v * c;
    r1.y = g0.y * c + g1.y * a;
    r1.u = g0.u * c + g1.u * a;
    r1.v = g0.v * c + g1.v * a;
}

/* OctreeNode_t: octree.h:82-142 */
struct Node {
    uint8_t value = 0;
    std::vector<int> index;
    int weight[16] = {};
    YUV raht[16];
    /* InvertHierarchicalTransform: octree_base.cc:84-100 */
    void inv_hier() {
        YUV H[8];
        for (int i = 0; i < 8; ++i) H[i] = raht[i];
        raht[1] = H[0];
        for (int i = 1; i < 8; ++i) {
            int a = NodeWeight[i][0], b = NodeWeight[i][1];
            YUV g0 = raht[i];
            inv_haar(weight[a], weight[b], g0, H[i], raht[a], raht[b]);
        }
    }
    /* The same transform in FP64 on a shadow array (error oracle only). */
    void inv_hier64(YUV64* r) const {
        YUV64 H[8];
        for (int i = 0; i < 8; ++i) H[i] = r[i];
        r[1] = H[0];
        for (int i = 1; i < 8; ++i) {
            int a = NodeWeight[i][0], b = NodeWeight[i][1];
            YUV64 g0 = r[i];
            inv_haar64(weight[a], weight[b], g0, H[i], r[a], r[b]);
        }
    }
};

/* ZstdDecoder::Decode: entropy_codec.cc:341-360 */
int zstd_unwrap(const uint8_t* src, size_t n, std::vector<uint8_t>& out) {
    unsigned long long sz = ZSTD_getFrameContentSize(src, n);
    if (sz == 0 || sz == ZSTD_CONTENTSIZE_UNKNOWN || sz == ZSTD_CONTENTSIZE_ERROR) return ORC_ERR_ZSTD;
    out.resize(sz);
    size_t r = ZSTD_decompress(out.data(), sz, src, n);
    if (ZSTD_isError(r)) return ORC_ERR_ZSTD;
    out.resize(r);
    return ORC_OK;
}

/* LoadTreeCore: octree_base.cc:33-48 (6 big-endian floats + u8 height). */
void load_tree_core(const uint8_t* p, XYZ& c, XYZ& r, int& h) {
    float d[6];
    for (int i = 0; i < 6; ++i) {
        uint32_t w = 0;
        for (int j = 0; j < 4; ++j) w = (w << 8) | p[i * 4 + j];
        std::memcpy(&d[i], &w, 4);
    }
    c = XYZ{d[0], d[1], d[2]};
    r = XYZ{d[3], d[4], d[5]};
    h = p[24];
}

/* InvertRAHTOctree: src/octree/invert_raht_octree.cc.  One object per patch
 * index, persistent across the GoP (pvvc_decompression.cc:99-101,119-122). */
struct PatchDecoder {
    bool fp64 = true;         /* keep the FP64 shadow transform (error oracle) */
    bool have_intra = false;  /* slice_ holds a valid type bit (:35) */
    uint8_t type = 0;
    uint32_t size = 0;
    XYZ center{}, range{};
    int height = 0;
    std::vector<std::vector<Node>> tree;
    std::vector<std::vector<std::array<YUV64, 16>>> tree64; /* FP64 shadow of raht[] */
    std::vector<XYZ> cloud;
    std::vector<YUV> coefs, ref, colors;
    std::vector<YUV64> coefs64, ref64, colors64;

    /* AddPoints: invert_raht_octree.cc:203-254 (recursive DFS). */
    void add_points(int h, int idx, XYZ c, XYZ r) {
        Node& node = tree[h][idx];
        if (h == height - 1) {
            node.value = 0xff;
            node.weight[1] = 1;
            cloud.push_back(c);
            node.index.push_back(static_cast<int>(cloud.size()));
            return;
        }
        XYZ sub{r.x / 2.0f, r.y / 2.0f, r.z / 2.0f};
        for (int i = 0; i < 8; ++i) {
            if (node.index[i] != -1) add_points(h + 1, node.index[i], sub_space_center(c, sub, i), sub);
        }
        for (int i = 0; i < 8; ++i)
            if (node.index[i] != -1) node.weight[i + 8] = tree[h + 1][node.index[i]].weight[1];
        for (int i = 7; i > 0; --i) node.weight[i] = node.weight[NodeWeight[i][0]] + node.weight[NodeWeight[i][1]];
    }

    /* MakeTree: invert_raht_octree.cc:122-201 */
    int make_tree(const std::vector<uint8_t>& g) {
        if (g.size() < 25) return ORC_ERR_BAD_SLICE;
        load_tree_core(g.data(), center, range, height);
        if (height <= 0) return ORC_ERR_BAD_SLICE;
        tree.assign(height, {});
        size_t it = 25;
        int curr = 1;
        for (int lvl = 0; lvl < height - 1; ++lvl) {
            int next = 0;
            tree[lvl].resize(curr);
            for (int cnt = 0; cnt < curr; ++cnt) {
                if (it >= g.size()) return ORC_ERR_BAD_SLICE; /* reference: UB */
                /* an empty internal node is never written (raht_octree.cc:154-167
                 * sets a bit per non-empty child); it is rejected here */
                if (g[it] == 0) return ORC_ERR_BAD_SLICE;
                Node& nd = tree[lvl][cnt];
                nd.value = g[it];
                for (int i = 0; i < 8; ++i) {
                    if ((g[it] & NodeValue[i]) == 0)
                        nd.index.push_back(-1);
                    else
                        nd.index.push_back(next++);
                }
                ++it;
            }
            curr = next;
        }
        if (static_cast<uint32_t>(curr) != size) return ORC_ERR_UNMATCHED_SIZE; /* :188-190 */
        tree.back().resize(curr);
        add_points(0, 0, center, range);
        return ORC_OK;
    }

    /* InvertRAHT: invert_raht_octree.cc:256-317 */
    void invert_raht(bool predict) {
        std::vector<YUV> c(coefs.rbegin(), coefs.rend()); /* :267 std::reverse */
        size_t it = 0;
        tree.front().front().raht[0] = c[it];
        if (fp64) {
            if (tree64.size() != tree.size()) {
                tree64.assign(tree.size(), {});
                for (size_t l = 0; l < tree.size(); ++l) tree64[l].assign(tree[l].size(), std::array<YUV64, 16>{});
            }
            tree64.front().front()[0] = coefs64[coefs64.size() - 1];
        }
        ++it;
        for (int i = 0; i < height - 1; ++i) {
            for (size_t j = 0; j < tree[i].size(); ++j) {
                Node& node = tree[i][j];
                size_t it0 = it;
                for (int idx = 1; idx < 8; ++idx) {
                    if (node.weight[NodeWeight[idx][0]] != 0 && node.weight[NodeWeight[idx][1]] != 0) {
                        node.raht[idx] = c[it];
                        ++it;
                    }
                }
                node.inv_hier();
                for (int idx = 0; idx < 8; ++idx)
                    if (node.index[idx] != -1) tree[i + 1][node.index[idx]].raht[0] = node.raht[idx + 8];
                if (fp64) {
                    auto& r = tree64[i][j];
                    size_t k = it0;
                    for (int idx = 1; idx < 8; ++idx)
                        if (node.weight[NodeWeight[idx][0]] != 0 && node.weight[NodeWeight[idx][1]] != 0)
                            r[idx] = coefs64[coefs64.size() - 1 - k++];
                    node.inv_hier64(r.data());
                    for (int idx = 0; idx < 8; ++idx)
                        if (node.index[idx] != -1) tree64[i + 1][node.index[idx]][0] = r[idx + 8];
                }
            }
        }
        colors.clear();
        for (auto& node : tree.back()) colors.push_back(node.raht[0]);
        if (fp64) {
            colors64.clear();
            for (auto& r : tree64.back()) colors64.push_back(r[0]);
        }
        if (predict) { /* :304-308 */
            for (size_t i = 0; i < colors.size(); ++i) {
                colors[i].y += ref[i].y;
                colors[i].u += ref[i].u;
                colors[i].v += ref[i].v;
            }
            if (fp64)
                for (size_t i = 0; i < colors64.size(); ++i) {
                    colors64[i].y += ref64[i].y;
                    colors64[i].u += ref64[i].u;
                    colors64[i].v += ref64[i].v;
                }
        } else {
            ref = colors;
            if (fp64) ref64 = colors64;
        }
    }

    /* SetSlice: invert_raht_octree.cc:22-85.  `geo`/`col` are the bytes as
     * stored; with unwrap_zstd the type bits 3/4 select zstd (.slice input,
     * :43-57), without it the bytes are already unwrapped (.frame input). */
    int set_slice(uint8_t t, uint32_t sz, uint8_t qp, const uint8_t* geo, uint32_t geo_size, const uint8_t* col,
                  uint32_t col_size, bool unwrap_zstd) {
        if (!(t & 0x01)) return ORC_ERR_BAD_SLICE;                       /* :25-27 */
        if ((!geo || geo_size == 0) && (!col || col_size == 0)) return ORC_ERR_BAD_SLICE; /* :30-32 */
        bool predict = (t & 0x02) != 0;
        if (!have_intra && predict) return ORC_ERR_EMPTY_REFERENCE; /* :35-40 */
        std::vector<uint8_t> gbytes, cbytes;
        if (!predict) {
            if (unwrap_zstd && (t & 0x08)) {
                if (zstd_unwrap(geo, geo_size, gbytes) != ORC_OK) return fail_intra(predict, ORC_ERR_ZSTD);
            } else {
                gbytes.assign(geo, geo + geo_size);
            }
        }
        if (unwrap_zstd && (t & 0x10)) {
            if (zstd_unwrap(col, col_size, cbytes) != ORC_OK) return fail_intra(predict, ORC_ERR_ZSTD);
        } else if (col) {
            cbytes.assign(col, col + col_size);
        }
        /* RLGR + dequantisation :59-69 */
        std::vector<int64_t> sym(3 * static_cast<size_t>(sz));
        RlgrDec dec;
        dec.decode(cbytes.data(), static_cast<int64_t>(cbytes.size()), 3 * static_cast<int64_t>(sz), sym.data());
        /* The reference detects no bitstream faults: it raises only the
         * structural errors below (tree, size), so those take precedence and
         * a bitstream fault is reported when the structure is sound. */
        int rlgr_st = ORC_OK;
        for (int64_t v : sym)
            if (v > INT32_MAX || v < INT32_MIN) {
                rlgr_st = ORC_ERR_SYMBOL_RANGE;
                break;
            }
        if (rlgr_st == ORC_OK && (dec.overrun || dec.bad_run)) rlgr_st = ORC_ERR_BITSTREAM_OVERRUN;
        coefs.assign(sz, YUV{});
        if (fp64) coefs64.assign(sz, YUV64{});
        for (uint32_t i = 0; i < sz; ++i) {
            int64_t a = sym[i] * qp, b = sym[i + sz] * qp, c = sym[i + 2 * static_cast<size_t>(sz)] * qp;
            coefs[i].y = static_cast<float>(a);
            coefs[i].u = static_cast<float>(b);
            coefs[i].v = static_cast<float>(c);
            if (fp64) {
                coefs64[i].y = static_cast<double>(a);
                coefs64[i].u = static_cast<double>(b);
                coefs64[i].v = static_cast<double>(c);
            }
        }
        if (!predict) { /* :71-78 */
            tree.clear();
            tree64.clear();
            cloud.clear();
            ref.clear();
            ref64.clear();
            size = sz;
            have_intra = false;
            int st = make_tree(gbytes);
            if (st != ORC_OK) return st;
        } else if (sz != cloud.size()) {
            return ORC_ERR_UNMATCHED_SIZE; /* InvertRAHT :264-266 */
        }
        if (rlgr_st != ORC_OK) return fail_intra(predict, rlgr_st);
        type = t;
        invert_raht(predict);
        have_intra = true;
        return ORC_OK;
    }
    int fail_intra(bool predict, int st) {
        if (!predict) {
            have_intra = false;
            cloud.clear();
        }
        return st;
    }
};

/* ConvertRGB: include/common/common.h:154-158.  Saturating (reference: UB
 * outside [0,255]); returns the number of saturated channels. */
int convert_rgb(const YUV& c, uint8_t* rgb, float* rgbf) {
    float r = c.y + 1.4020f * (c.v - 128.0f);
    float g = c.y - 0.3441f * (c.u - 128.0f) - 0.7141f * (c.v - 128.0f);
    float b = c.y + 1.7720f * (c.u - 128.0f);
    float f[3] = {r, g, b};
    int sat = 0;
    for (int i = 0; i < 3; ++i) {
        float q = std::round(f[i]);
        if (q != q) {
            q = 0.0f;
            ++sat;
        } else if (q < 0.0f) {
            q = 0.0f;
            ++sat;
        } else if (q > 255.0f) {
            q = 255.0f;
            ++sat;
        }
        if (rgb) rgb[i] = static_cast<uint8_t>(q);
        if (rgbf) rgbf[i] = f[i];
    }
    return sat;
}

void convert_rgb64(const YUV64& c, double* out) {
    out[0] = c.y + 1.4020 * (c.v - 128.0);
    out[1] = c.y - 0.3441 * (c.u - 128.0) - 0.7141 * (c.v - 128.0);
    out[2] = c.y + 1.7720 * (c.u - 128.0);
}

/* ------------------------------------------------------------------------ */
/* .frame container: reader restates io::LoadFrame (src/cuda/common/        */
/* slice_io.cu:143-275); layout written by ChangeSliceToFrame                */
/* (src/io/slice_io.cc:258-315).                                             */
/* ------------------------------------------------------------------------ */
struct Frame {
    int timestamp = -1;
    uint32_t slice_cnt = 0;
    std::vector<int> index;
    std::vector<uint8_t> type, qp;
    std::vector<uint32_t> size, geo_size, col_size;
    std::vector<float> mv; /* [S][16] row-major */
    std::vector<std::vector<uint8_t>> geo, col;
};

int load_frame(const char* path, Frame& f) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return errno == ENOENT ? 2 : (errno == EACCES ? 3 : 4);
    auto rd = [&](void* dst, size_t sz, size_t n) { return std::fread(dst, sz, n, fp) == n; };
    int rc = 0;
    do {
        if (!rd(&f.timestamp, 4, 1) || !rd(&f.slice_cnt, 4, 1)) { rc = 5; break; }
        uint32_t S = f.slice_cnt;
        f.index.resize(S);
        f.type.resize(S);
        f.size.resize(S);
        f.qp.resize(S);
        f.mv.resize(size_t(S) * 16);
        f.geo_size.resize(S);
        f.col_size.resize(S);
        f.geo.resize(S);
        f.col.resize(S);
        if (S && (!rd(f.index.data(), 4, S) || !rd(f.type.data(), 1, S) || !rd(f.size.data(), 4, S) ||
                  !rd(f.qp.data(), 1, S) || !rd(f.mv.data(), 4, size_t(S) * 16))) { rc = 5; break; }
        if (S && !rd(f.geo_size.data(), 4, S)) { rc = 5; break; }
        for (uint32_t i = 0; i < S && !rc; ++i) {
            if (!f.geo_size[i]) continue;
            std::vector<uint8_t> b(f.geo_size[i]);
            if (!rd(b.data(), 1, b.size())) { rc = 5; break; }
            if (f.type[i] & 0x08) {
                if (zstd_unwrap(b.data(), b.size(), f.geo[i]) != ORC_OK) { rc = 6; break; }
                f.geo_size[i] = static_cast<uint32_t>(f.geo[i].size());
            } else {
                f.geo[i] = std::move(b);
            }
        }
        if (rc) break;
        if (S && !rd(f.col_size.data(), 4, S)) { rc = 5; break; }
        for (uint32_t i = 0; i < S && !rc; ++i) {
            if (!f.col_size[i]) continue;
            std::vector<uint8_t> b(f.col_size[i]);
            if (!rd(b.data(), 1, b.size())) { rc = 5; break; }
            if (f.type[i] & 0x10) {
                if (zstd_unwrap(b.data(), b.size(), f.col[i]) != ORC_OK) { rc = 6; break; }
                f.col_size[i] = static_cast<uint32_t>(f.col[i].size());
            } else {
                f.col[i] = std::move(b);
            }
        }
    } while (false);
    std::fclose(fp);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* .slice files: io::SaveSlice / io::LoadSlice (src/io/slice_io.cc:20-206)   */
/* and io::ChangeSliceToFrame with the Frame(slices) constructor            */
/* (slice_io.cc:208-315, include/common/common.h:378-412).  The reference   */
/* throws; here each exception is a code (the product's numbering):        */
/* WRONG_FILE_FORMAT 1, FILE_NOT_EXIST 2, PERMISSION_DENIED 3,              */
/* UNEXPECTED_FILE_ERROR / INVALID_DIR 4, FILE_READ_ERROR 5, BAD_SLICE 7,   */
/* EMPTY_RESULT (zero blob size) 8, EMPTY_RESULT (no slices) 9,             */
/* FILE_WRITE_ERROR 10.                                                      */
/* ------------------------------------------------------------------------ */
struct OrcSlice {
    int timestamp = -1;
    int index = -1;
    uint8_t type = 0;
    float mv[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}; /* mv[row][col] */
    uint64_t size = 0;
    uint8_t qp = 1;
    bool geometry = false, color = false; /* shared_ptr present */
    std::vector<uint8_t> geo, col;
};

bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

/* LoadSlice, slice_io.cc:124-206 */
int load_slice(const std::string& name, OrcSlice& sl) {
    if (!ends_with(name, ".slice")) return 1;
    std::FILE* fp = std::fopen(name.c_str(), "rb");
    if (!fp) return errno == ENOENT ? 2 : (errno == EACCES ? 3 : 4);
    std::fseek(fp, 0, SEEK_END);
    const long total = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    int rc = 0;
    auto one = [&](void* dst, size_t n) {
        if (rc) return false;
        if (std::fread(dst, 1, n, fp) != n) rc = 5;
        return rc == 0;
    };
    sl = OrcSlice{};
    if (one(&sl.type, 1) && !(sl.type & 0x01)) rc = 7; /* PVVC_SLICE_TYPE_VALID */
    one(&sl.timestamp, sizeof(int));
    one(&sl.index, sizeof(int));
    one(&sl.size, sizeof(uint64_t));
    one(&sl.qp, 1);
    float cm[16];
    if (one(cm, sizeof(cm)))
        for (int k = 0; k < 16; ++k) sl.mv[k % 4][k / 4] = cm[k]; /* Eigen storage is column-major */
    auto blob = [&](std::vector<uint8_t>& v, bool& present) {
        uint64_t n = 0;
        if (!one(&n, sizeof(uint64_t))) return;
        if (n == 0) {
            rc = 8;
            return;
        }
        const long at = std::ftell(fp);
        if (at < 0 || static_cast<uint64_t>(total - at) < n) {
            rc = 5;
            return;
        }
        v.resize(n);
        if (one(v.data(), n)) present = true;
    };
    if (!rc && !(sl.type & 0x02)) blob(sl.geo, sl.geometry); /* not PREDICT */
    if (!rc && !(sl.type & 0x04)) blob(sl.col, sl.color);    /* not SKIP */
    std::fclose(fp);
    return rc;
}

/* ChangeSliceToFrame's writing part, slice_io.cc:238-310 */
int write_frame_of(std::vector<OrcSlice> slices, const std::string& out) {
    if (!ends_with(out, ".frame")) return 1;
    if (slices.empty()) return 9;
    std::stable_sort(slices.begin(), slices.end(),
                     [](const OrcSlice& x, const OrcSlice& y) { return x.index < y.index; });
    std::vector<uint8_t> o;
    auto w = [&](const void* p, size_t n) {
        const auto* b = static_cast<const uint8_t*>(p);
        o.insert(o.end(), b, b + n);
    };
    const int ts = slices.front().timestamp; /* Frame(slices): the first slice's */
    const uint32_t cnt = static_cast<uint32_t>(slices.size());
    w(&ts, sizeof(int));
    w(&cnt, sizeof(uint32_t));
    for (const auto& sl : slices) w(&sl.index, sizeof(int));
    for (const auto& sl : slices) w(&sl.type, 1);
    for (const auto& sl : slices) {
        const uint32_t sz = static_cast<uint32_t>(sl.size);
        w(&sz, sizeof(uint32_t));
    }
    for (const auto& sl : slices) w(&sl.qp, 1);
    for (const auto& sl : slices)
        for (int row = 0; row < 4; ++row)
            for (int col = 0; col < 4; ++col) w(&sl.mv[row][col], sizeof(float));
    for (const auto& sl : slices) {
        const uint32_t g = sl.geometry ? static_cast<uint32_t>(sl.geo.size()) : 0u;
        w(&g, sizeof(uint32_t));
    }
    for (const auto& sl : slices)
        if (sl.geometry) w(sl.geo.data(), sl.geo.size());
    for (const auto& sl : slices) {
        const uint32_t c = sl.color ? static_cast<uint32_t>(sl.col.size()) : 0u;
        w(&c, sizeof(uint32_t));
    }
    for (const auto& sl : slices)
        if (sl.color) w(sl.col.data(), sl.col.size());
    std::FILE* fp = std::fopen(out.c_str(), "wb");
    if (!fp) return errno == ENOENT ? 2 : (errno == EACCES ? 3 : 4);
    const bool ok = std::fwrite(o.data(), 1, o.size(), fp) == o.size();
    return (std::fclose(fp) == 0 && ok) ? 0 : 10;
}

/* ------------------------------------------------------------------------ */
/* Quality metric: symmetric nearest-neighbour MSE, MSE::Compute            */
/* (src/registration/registration_base.cc:168-223).                         */
/*                                                                          */
/* For every point of p, its nearest point in q (squared Euclidean distance, */
/* the value PCL's KdTree::nearestKSearch returns), and the squared Y/U/V    */
/* differences of the pair; then the same from q to p.  Means over each      */
/* cloud.  Defined where the reference is implementation-defined:           */
/*   - equidistant candidates: the lowest index wins (FLANN's order is      */
/*     unspecified);                                                        */
/*   - sums are FP64 over the exact FP64 squares of the FP32 terms; the      */
/*     reference keeps a float running sum (mode bit 1 reproduces that one  */
/*     in query order).                                                     */
/* ------------------------------------------------------------------------ */

/* squared distance, FP32, (dx*dx + dy*dy) + dz*dz (no FMA) */
inline float nn_d2(const float* a, const float* b) {
    const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return (dx * dx + dy * dy) + dz * dz;
}

/* ColorYUV(const PointXYZRGB&), include/common/common.h:95-99 */
inline void nn_yuv(const uint8_t* c, float* o) {
    const float r = c[0], g = c[1], b = c[2];
    o[0] = r * 0.299f + g * 0.587f + b * 0.114f;
    o[1] = r * -0.168736f - g * 0.331264f + b * 0.5f + 128;
    o[2] = r * 0.5f - g * 0.418688f - b * 0.081312f + 128;
}

/* Nearest neighbours of every query in `tgt` by brute force. */
void nn_brute(const float* qry, int64_t nq, const float* tgt, int64_t nt, int64_t* idx, float* d2) {
    for (int64_t i = 0; i < nq; ++i) {
        float best = INFINITY;
        int64_t bi = -1;
        for (int64_t j = 0; j < nt; ++j) {
            const float d = nn_d2(qry + 3 * i, tgt + 3 * j);
            if (d < best) best = d, bi = j;
        }
        idx[i] = bi;
        d2[i] = best;
    }
}

/* Nearest neighbours by a uniform grid of cubic cells (hash of cell ->
 * sorted run) and a search over shells of cells of growing Chebyshev
 * radius, stopped once no unvisited cell can hold a point at or below the
 * best distance.  Same answers as nn_brute (lowest index on ties). */
struct NnGrid {
    double org[3] = {0, 0, 0}, h = 1;
    int64_t dim[3] = {1, 1, 1};
    std::vector<int64_t> order;                 /* target indices sorted by (cell, index) */
    std::vector<uint64_t> key;                  /* cell key of order[i] */
    std::vector<std::pair<uint64_t, int64_t>> head; /* (cell key, first position), sorted */

    void cell_of(const float* p, int64_t* c) const {
        for (int a = 0; a < 3; ++a) {
            int64_t v = static_cast<int64_t>(std::floor((double(p[a]) - org[a]) / h));
            c[a] = std::min(std::max<int64_t>(v, 0), dim[a] - 1);
        }
    }
    static uint64_t pack(const int64_t* c) {
        return (uint64_t(c[0]) << 42) | (uint64_t(c[1]) << 21) | uint64_t(c[2]);
    }
    void build(const float* t, int64_t n) {
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], double(t[3 * i + a])), hi[a] = std::max(hi[a], double(t[3 * i + a]));
        double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-6});
        h = std::max(2.0 * ext / std::sqrt(double(std::max<int64_t>(n, 1))), ext / 1.0e6);
        for (int a = 0; a < 3; ++a) {
            org[a] = lo[a];
            dim[a] = std::min<int64_t>(int64_t((hi[a] - lo[a]) / h) + 1, (int64_t(1) << 21) - 1);
        }
        std::vector<std::pair<uint64_t, int64_t>> kv(n);
        for (int64_t i = 0; i < n; ++i) {
            int64_t c[3];
            cell_of(t + 3 * i, c);
            kv[i] = {pack(c), i};
        }
        std::sort(kv.begin(), kv.end());
        order.resize(n);
        key.resize(n);
        head.clear();
        for (int64_t i = 0; i < n; ++i) {
            order[i] = kv[i].second;
            key[i] = kv[i].first;
            if (i == 0 || kv[i].first != kv[i - 1].first) head.push_back({kv[i].first, i});
        }
    }
    void query(const float* t, const float* q, int64_t& bi, float& best) const {
        int64_t c[3];
        cell_of(q, c);
        best = INFINITY;
        bi = -1;
        const int64_t rmax = std::max({dim[0], dim[1], dim[2]});
        for (int64_t r = 0; r <= rmax; ++r) {
            for (int64_t dx = -r; dx <= r; ++dx)
                for (int64_t dy = -r; dy <= r; ++dy)
                    for (int64_t dz = -r; dz <= r; ++dz) {
                        if (std::max({std::llabs(dx), std::llabs(dy), std::llabs(dz)}) != r) continue;
                        int64_t cc[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
                        if (cc[0] < 0 || cc[1] < 0 || cc[2] < 0 || cc[0] >= dim[0] || cc[1] >= dim[1] || cc[2] >= dim[2])
                            continue;
                        const uint64_t k = pack(cc);
                        auto it = std::lower_bound(head.begin(), head.end(), std::make_pair(k, int64_t(-1)));
                        if (it == head.end() || it->first != k) continue;
                        for (int64_t j = it->second; j < static_cast<int64_t>(key.size()) && key[j] == k; ++j) {
                            const int64_t ti = order[j];
                            const float d = nn_d2(q, t + 3 * ti);
                            if (d < best || (d == best && ti < bi)) best = d, bi = ti;
                        }
                    }
            /* every unvisited point lies outside cells [c-r, c+r]: at least lb away */
            double lb = INFINITY;
            for (int a = 0; a < 3; ++a) {
                const double lo_a = org[a] + double(c[a] - r) * h, hi_a = org[a] + double(c[a] + r + 1) * h;
                if (c[a] - r > 0) lb = std::min(lb, double(q[a]) - lo_a);
                if (c[a] + r < dim[a] - 1) lb = std::min(lb, hi_a - double(q[a]));
            }
            if (lb == INFINITY) break; /* the cube covers the grid */
            lb -= 1e-9 * (std::fabs(double(q[0])) + std::fabs(double(q[1])) + std::fabs(double(q[2])) + h);
            if (lb > 0 && double(best) < lb * lb * (1.0 - 1e-6)) break;
        }
    }
};

void nn_grid(const float* qry, int64_t nq, const float* tgt, int64_t nt, int64_t* idx, float* d2, int threads) {
    NnGrid g;
    g.build(tgt, nt);
    std::atomic<int64_t> next{0};
    auto work = [&]() {
        for (;;) {
            const int64_t b = next.fetch_add(4096);
            if (b >= nq) break;
            for (int64_t i = b; i < std::min(nq, b + 4096); ++i) g.query(tgt, qry + 3 * i, idx[i], d2[i]);
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < std::max(1, threads); ++i) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
}

/* One direction of MSE::Compute: sums over queries (in query order) of the
 * squared distance and the squared Y, U, V differences to the neighbour. */
void mse_sums(const float* qx, const uint8_t* qc, int64_t nq, const uint8_t* tc, const int64_t* idx,
              const float* d2, bool float_running, double* out4) {
    double s[4] = {0, 0, 0, 0};
    float f[4] = {0, 0, 0, 0};
    for (int64_t i = 0; i < nq; ++i) {
        float a[3], b[3];
        nn_yuv(qc + 3 * i, a);
        nn_yuv(tc + 3 * idx[i], b);
        double t[4] = {double(d2[i]), std::pow(double(a[0] - b[0]), 2), std::pow(double(a[1] - b[1]), 2),
                       std::pow(double(a[2] - b[2]), 2)};
        for (int k = 0; k < 4; ++k) {
            s[k] += t[k];
            f[k] = static_cast<float>(double(f[k]) + t[k]); /* float += double */
        }
    }
    (void)qx;
    for (int k = 0; k < 4; ++k) out4[k] = float_running ? double(f[k] / float(nq)) : s[k] / double(nq);
}

}  // namespace

/* ======================================================================== */
/* C ABI for the Python tests (ctypes).                                      */
/* ======================================================================== */
/* ------------------------------------------------------------------------ */
/* Forward path (the colour half of the encoder)                             */
/* ------------------------------------------------------------------------ */

/* HaarTransform: octree_base.cc:164-184.  FP32; T as the reference builds it,
 * G = g0*T00 + g1*T01, H = g0*T10 + g1*T11 per channel (ColorYUV operator*
 * then operator+, common.h:140-152). */
void haar_fwd(int w0, int w1, const YUV& g0, const YUV& g1, YUV& G, YUV& H) {
    if (w0 == 0 && w1 == 0) {
        G = YUV{};
        H = YUV{};
        return;
    }
    float base = std::sqrt(static_cast<float>(w0 + w1));
    float t00 = std::sqrt(static_cast<float>(w0)) / base;
    float t01 = std::sqrt(static_cast<float>(w1)) / base;
    float t10 = -std::sqrt(static_cast<float>(w1)) / base;
    float t11 = std::sqrt(static_cast<float>(w0)) / base;
    G.y = g0.y * t00 + g1.y * t01;
    G.u = g0.u * t00 + g1.u * t01;
    G.v = g0.v * t00 + g1.v * t01;
    H.y = g0.y * t10 + g1.y * t11;
    H.u = g0.u * t10 + g1.u * t11;
    H.v = g0.v * t10 + g1.v * t11;
}

/* HierarchicalTransform: octree_base.cc:65-82.  Pairs i = 7..1 in NodeWeight
 * order; raht[i] takes the low-pass, H[i] the high-pass; then raht[0] is the
 * node's DC and raht[1..7] its high-passes. */
void hier_fwd(Node& nd) {
    YUV H[8];
    for (int i = 7; i > 0; --i) {
        int a = NodeWeight[i][0], b = NodeWeight[i][1];
        YUV G, Hi;
        haar_fwd(nd.weight[a], nd.weight[b], nd.raht[a], nd.raht[b], G, Hi);
        nd.raht[i] = G;
        H[i] = Hi;
    }
    nd.raht[0] = nd.raht[1];
    for (int i = 1; i <= 7; ++i) nd.raht[i] = H[i];
}

/* RAHTOctree::RAHT: raht_octree.cc:189-247, on the tree of intra geometry
 * bytes.  The decoder's MakeTree (invert_raht_octree.cc:122-201) rebuilds the
 * writer's tree from those bytes exactly: the same layers, node order, child
 * indices and weights (AddNode, raht_octree.cc:130-187, appends each layer in
 * the order the decoder's breadth-first expansion reads it).  leaf[k] is leaf
 * k's colour: the writer's average of the points in that voxel (:193-200),
 * done by the caller.  Output: the coefficients in stream order (after the
 * std::reverse of :245), planar Y[n] U[n] V[n].  Returns an ORC status. */
int raht_forward(const uint8_t* geo, uint32_t geo_size, const float* leaf, uint32_t n, std::vector<YUV>& out) {
    PatchDecoder d;
    d.fp64 = false;
    d.size = n;
    std::vector<uint8_t> g(geo, geo + geo_size);
    int st = d.make_tree(g);
    if (st != ORC_OK) return st;
    auto& tree = d.tree;
    const int h = d.height;
    for (size_t k = 0; k < tree.back().size(); ++k) /* :193-200 */
        tree.back()[k].raht[0] = YUV{leaf[3 * k], leaf[3 * k + 1], leaf[3 * k + 2]};
    for (int i = h - 2; i >= 0; --i) { /* :202-213 */
        for (auto& node : tree[i]) {
            for (int idx = 0; idx < 8; ++idx)
                if (node.index[idx] != -1) node.raht[idx + 8] = tree[i + 1][node.index[idx]].raht[0];
            hier_fwd(node);
        }
    }
    std::vector<YUV> res; /* :215-245 */
    res.push_back(tree.front().front().raht[0]);
    for (int i = 0; i < h - 1; ++i)
        for (auto& node : tree[i])
            for (int idx = 1; idx < 8; ++idx)
                if (node.weight[NodeWeight[idx][0]] != 0 && node.weight[NodeWeight[idx][1]] != 0)
                    res.push_back(node.raht[idx]);
    std::reverse(res.begin(), res.end());
    if (res.size() != n) return ORC_ERR_UNMATCHED_SIZE;
    out.swap(res);
    return ORC_OK;
}

/* Quantisation: GoP_encoding.cc:180-186, std::round(c / QP) as int64
 * (FIX_DATA_INT, entropy_codec.h:33); planar Y|U|V. */
void quantise(const std::vector<YUV>& c, uint8_t qp, int64_t* q) {
    const size_t n = c.size();
    for (size_t j = 0; j < n; ++j) {
        q[j] = static_cast<int64_t>(std::round(c[j].y / qp));
        q[j + n] = static_cast<int64_t>(std::round(c[j].u / qp));
        q[j + 2 * n] = static_cast<int64_t>(std::round(c[j].v / qp));
    }
}

extern "C" {

/* RLGR encode of n symbols; returns the byte count (writes at most cap). */
int64_t orc_rlgr_encode(const int64_t* data, int64_t n, uint8_t* out, int64_t cap) {
    RlgrEnc e;
    e.encode(data, n);
    e.finish();
    int64_t sz = static_cast<int64_t>(e.out.size());
    if (out) std::memcpy(out, e.out.data(), static_cast<size_t>(std::min(sz, cap)));
    return sz;
}

/* RLGR decode of nsym symbols; returns 0 or ORC_ERR_BITSTREAM_OVERRUN. */
int orc_rlgr_decode(const uint8_t* data, int64_t nbytes, int64_t nsym, int64_t* out) {
    RlgrDec d;
    d.decode(data, nbytes, nsym, out);
    return (d.overrun || d.bad_run) ? ORC_ERR_BITSTREAM_OVERRUN : ORC_OK;
}

/* Per-patch decoder handle. */
void* orc_patch_new(void) { return new PatchDecoder(); }
void orc_patch_free(void* h) { delete static_cast<PatchDecoder*>(h); }

int orc_patch_set_slice(void* h, uint8_t type, uint32_t size, uint8_t qp, const uint8_t* geo, uint32_t geo_size,
                        const uint8_t* col, uint32_t col_size, int unwrap_zstd) {
    return static_cast<PatchDecoder*>(h)->set_slice(type, size, qp, geo, geo_size, col, col_size, unwrap_zstd != 0);
}

/* GetPatch (invert_raht_octree.cc:87-120): leaf centres, FP32 YUV, RGB8,
 * unrounded FP32 RGB and the FP64 RGB.  Returns the point count or -1. */
int64_t orc_patch_get(void* h, float* xyz, float* yuv, uint8_t* rgb, float* rgbf, double* rgb64) {
    auto* d = static_cast<PatchDecoder*>(h);
    if (!d->have_intra || d->cloud.empty()) return -1;
    size_t n = d->cloud.size();
    for (size_t i = 0; i < n; ++i) {
        if (xyz) {
            xyz[3 * i] = d->cloud[i].x;
            xyz[3 * i + 1] = d->cloud[i].y;
            xyz[3 * i + 2] = d->cloud[i].z;
        }
        if (yuv) {
            yuv[3 * i] = d->colors[i].y;
            yuv[3 * i + 1] = d->colors[i].u;
            yuv[3 * i + 2] = d->colors[i].v;
        }
        convert_rgb(d->colors[i], rgb ? rgb + 3 * i : nullptr, rgbf ? rgbf + 3 * i : nullptr);
        if (rgb64 && d->fp64) convert_rgb64(d->colors64[i], rgb64 + 3 * i);
    }
    return static_cast<int64_t>(n);
}

/* Forward RAHT of leaf colours (n x 3, leaf order) on intra geometry bytes:
 * 3n FP32 coefficients in stream order, planar.  Returns 0 or a status. */
int orc_raht_forward(const uint8_t* geo, uint32_t geo_size, const float* leaf_yuv, uint32_t n, float* coef) {
    std::vector<YUV> c;
    int st = raht_forward(geo, geo_size, leaf_yuv, n, c);
    if (st != ORC_OK) return st;
    for (uint32_t j = 0; j < n; ++j) {
        coef[j] = c[j].y;
        coef[j + n] = c[j].u;
        coef[j + 2 * n] = c[j].v;
    }
    return ORC_OK;
}

/* Inverse RAHT of 3n FP32 coefficients (stream order, planar) on intra
 * geometry bytes: the decoder's InvertRAHT (invert_raht_octree.cc:256-317)
 * without the RLGR and dequantisation steps.  leaf_yuv receives n x 3.
 * Returns 0 or a status. */
int orc_raht_inverse(const uint8_t* geo, uint32_t geo_size, uint32_t n, const float* coef, float* leaf_yuv) {
    PatchDecoder d;
    d.fp64 = false;
    d.size = n;
    std::vector<uint8_t> g(geo, geo + geo_size);
    int st = d.make_tree(g);
    if (st != ORC_OK) return st;
    d.coefs.assign(n, YUV{});
    for (uint32_t j = 0; j < n; ++j) d.coefs[j] = YUV{coef[j], coef[j + n], coef[j + 2 * n]};
    d.invert_raht(false);
    for (uint32_t k = 0; k < n; ++k) {
        leaf_yuv[3 * k] = d.colors[k].y;
        leaf_yuv[3 * k + 1] = d.colors[k].u;
        leaf_yuv[3 * k + 2] = d.colors[k].v;
    }
    return ORC_OK;
}

/* The colour stream of one slice as GoP_encoding.cc:172-192 writes it before
 * the optional zstd wrap: forward RAHT, quantisation, RLGR.  quant (optional)
 * receives the 3n quantised symbols.  Returns the byte count (writes at most
 * cap bytes to out), or -(status). */
int64_t orc_color_encode(const uint8_t* geo, uint32_t geo_size, const float* leaf_yuv, uint32_t n, uint8_t qp,
                         int64_t* quant, uint8_t* out, int64_t cap) {
    if (qp == 0 || n == 0) return -ORC_ERR_BAD_SLICE;
    std::vector<YUV> c;
    int st = raht_forward(geo, geo_size, leaf_yuv, n, c);
    if (st != ORC_OK) return -st;
    std::vector<int64_t> q(3ull * n);
    quantise(c, qp, q.data());
    if (quant) std::copy(q.begin(), q.end(), quant);
    return orc_rlgr_encode(q.data(), static_cast<int64_t>(q.size()), out, cap);
}

/* Tree-only expansion of intra geometry bytes (for known-answer tests).
 * Returns the leaf count, or -(status). */
int64_t orc_tree_expand(const uint8_t* geo, uint32_t geo_size, uint32_t size, float* xyz, int* height) {
    PatchDecoder d;
    d.size = size;
    std::vector<uint8_t> g(geo, geo + geo_size);
    int st = d.make_tree(g);
    if (height) *height = d.height;
    if (st != ORC_OK) return -st;
    for (size_t i = 0; i < d.cloud.size() && xyz; ++i) {
        xyz[3 * i] = d.cloud[i].x;
        xyz[3 * i + 1] = d.cloud[i].y;
        xyz[3 * i + 2] = d.cloud[i].z;
    }
    return static_cast<int64_t>(d.cloud.size());
}

/* .frame loader (io::LoadFrame return codes 1-5, plus 6 for a zstd error). */
void* orc_frame_load(const char* path, int* rc) {
    auto* f = new Frame();
    int r = load_frame(path, *f);
    if (rc) *rc = r;
    if (r) {
        delete f;
        return nullptr;
    }
    return f;
}
void orc_frame_free(void* h) { delete static_cast<Frame*>(h); }
uint32_t orc_frame_slice_cnt(void* h) { return static_cast<Frame*>(h)->slice_cnt; }
int orc_frame_timestamp(void* h) { return static_cast<Frame*>(h)->timestamp; }
/* Copies the per-slice metadata out; geo/col sizes are post-unwrap. */
void orc_frame_meta(void* h, int* index, uint8_t* type, uint32_t* size, uint8_t* qp, float* mv, uint32_t* geo_size,
                    uint32_t* col_size) {
    auto* f = static_cast<Frame*>(h);
    uint32_t S = f->slice_cnt;
    for (uint32_t i = 0; i < S; ++i) {
        if (index) index[i] = f->index[i];
        if (type) type[i] = f->type[i];
        if (size) size[i] = f->size[i];
        if (qp) qp[i] = f->qp[i];
        if (geo_size) geo_size[i] = f->geo_size[i];
        if (col_size) col_size[i] = f->col_size[i];
    }
    if (mv) std::memcpy(mv, f->mv.data(), f->mv.size() * 4);
}
const uint8_t* orc_frame_geo(void* h, uint32_t s) {
    auto* f = static_cast<Frame*>(h);
    return f->geo[s].empty() ? nullptr : f->geo[s].data();
}
const uint8_t* orc_frame_col(void* h, uint32_t s) {
    auto* f = static_cast<Frame*>(h);
    return f->col[s].empty() ? nullptr : f->col[s].data();
}

/* Stream decoder context: PVVCDecompression (pvvc_decompression.cc:86-149):
 * one persistent PatchDecoder per patch index, a thread pool popping patch
 * tasks from a queue, then per-point motion (:141-149). */
struct OrcCtx {
    std::vector<std::unique_ptr<PatchDecoder>> handler;
    bool fp64 = true;
};
void* orc_ctx_new(void) { return new OrcCtx(); }
/* Turns the FP64 shadow transform off (timing the FP32 reference path alone). */
void orc_ctx_set_fp64(void* h, int enable) {
    auto* c = static_cast<OrcCtx*>(h);
    c->fp64 = enable != 0;
    for (auto& d : c->handler) d->fp64 = c->fp64;
}
void orc_ctx_free(void* h) { delete static_cast<OrcCtx*>(h); }

/* Decodes one loaded frame.  Outputs, all optional, are indexed by the exclusive
 * prefix of size[] in slice order: xyz after motion (f32x3), RGB8, unrounded
 * FP32 RGB, FP64 RGB, pre-motion leaf centres.  status[s] per slice.  Returns
 * the total point count; *sat receives the saturated-channel count.
 * `timed_ms` (optional) receives the wall time of the SetSlice+GetPatch region
 * (the reference's timed region, :124-139) and of the full decode incl. motion. */
int64_t orc_decode_frame(void* hctx, void* hframe, int nthreads, float* xyz, uint8_t* rgb, float* rgbf, double* rgb64,
                         float* leaf_xyz, int* status, int64_t* sat, double* timed_ms) {
    auto* ctx = static_cast<OrcCtx*>(hctx);
    auto* f = static_cast<Frame*>(hframe);
    uint32_t S = f->slice_cnt;
    int max_idx = 0;
    for (uint32_t i = 0; i < S; ++i) max_idx = std::max(max_idx, f->index[i]);
    while (ctx->handler.size() < static_cast<size_t>(max_idx) + 1) {
        ctx->handler.emplace_back(new PatchDecoder());
        ctx->handler.back()->fp64 = ctx->fp64;
    }
    std::vector<int64_t> off(S + 1, 0);
    for (uint32_t i = 0; i < S; ++i) off[i + 1] = off[i] + f->size[i];
    std::vector<int> st(S, 0);
    std::atomic<int64_t> satc{0};
    auto t0 = std::chrono::steady_clock::now();
    /* Task :86-103 -- SetSlice + GetPatch per popped patch */
    std::atomic<uint32_t> next{0};
    auto task = [&]() {
        for (;;) {
            uint32_t s = next.fetch_add(1);
            if (s >= S) break;
            PatchDecoder* d = ctx->handler[f->index[s]].get();
            st[s] = d->set_slice(f->type[s], f->size[s], f->qp[s], f->geo[s].data(), f->geo_size[s],
                                 f->col[s].data(), f->col_size[s], false);
            if (st[s] != ORC_OK) continue;
            int64_t base = off[s], ls = 0;
            for (size_t i = 0; i < d->cloud.size(); ++i) {
                int64_t o = base + static_cast<int64_t>(i);
                ls += convert_rgb(d->colors[i], rgb ? rgb + 3 * o : nullptr, rgbf ? rgbf + 3 * o : nullptr);
                if (rgb64 && d->fp64) convert_rgb64(d->colors64[i], rgb64 + 3 * o);
                if (leaf_xyz) {
                    leaf_xyz[3 * o] = d->cloud[i].x;
                    leaf_xyz[3 * o + 1] = d->cloud[i].y;
                    leaf_xyz[3 * o + 2] = d->cloud[i].z;
                }
            }
            satc += ls;
        }
    };
    int nt = std::max(1, nthreads);
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(task);
    task();
    for (auto& t : th) t.join();
    auto t1 = std::chrono::steady_clock::now();
    /* Motion :141-149, on the calling thread as in the reference */
    if (xyz) {
        for (uint32_t s = 0; s < S; ++s) {
            if (st[s] != ORC_OK) continue;
            PatchDecoder* d = ctx->handler[f->index[s]].get();
            const float* m = &f->mv[size_t(s) * 16];
            int64_t base = off[s];
            for (size_t i = 0; i < d->cloud.size(); ++i) {
                const XYZ& p = d->cloud[i];
                int64_t o = base + static_cast<int64_t>(i);
                xyz[3 * o] = p.x * m[0] + p.y * m[1] + p.z * m[2] + m[3];
                xyz[3 * o + 1] = p.x * m[4] + p.y * m[5] + p.z * m[6] + m[7];
                xyz[3 * o + 2] = p.x * m[8] + p.y * m[9] + p.z * m[10] + m[11];
            }
        }
    }
    auto t2 = std::chrono::steady_clock::now();
    if (sat) *sat = satc.load();
    if (status)
        for (uint32_t i = 0; i < S; ++i) status[i] = st[i];
    if (timed_ms) {
        timed_ms[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        timed_ms[1] = std::chrono::duration<double, std::milli>(t2 - t0).count();
    }
    return off[S];
}

/* MSE::Compute (registration_base.cc:168-223).  out8 = {geo, y, u, v} of
 * p->q then {geo, y, u, v} of q->p.  mode bit 0: brute-force neighbours
 * (else the grid); bit 1: the reference's float running sums.  nn_p/nn_q
 * (optional) receive the neighbour index of each point.  An empty cloud
 * gives -1 for its direction and leaves the other at 0 (:170-179). */
int orc_cloud_mse(const float* pxyz, const uint8_t* prgb, int64_t np, const float* qxyz, const uint8_t* qrgb,
                  int64_t nq, int mode, int threads, double* out8, int64_t* nn_p, int64_t* nn_q) {
    for (int k = 0; k < 8; ++k) out8[k] = 0.0;
    if (np <= 0)
        for (int k = 0; k < 4; ++k) out8[k] = -1.0;
    if (nq <= 0)
        for (int k = 4; k < 8; ++k) out8[k] = -1.0;
    if (np <= 0 || nq <= 0) return 0;
    std::vector<int64_t> ip(np), iq(nq);
    std::vector<float> dp(np), dq(nq);
    if (mode & 1) {
        nn_brute(pxyz, np, qxyz, nq, ip.data(), dp.data());
        nn_brute(qxyz, nq, pxyz, np, iq.data(), dq.data());
    } else {
        nn_grid(pxyz, np, qxyz, nq, ip.data(), dp.data(), threads);
        nn_grid(qxyz, nq, pxyz, np, iq.data(), dq.data(), threads);
    }
    mse_sums(pxyz, prgb, np, qrgb, ip.data(), dp.data(), (mode & 2) != 0, out8);
    mse_sums(qxyz, qrgb, nq, prgb, iq.data(), dq.data(), (mode & 2) != 0, out8 + 4);
    if (nn_p) std::memcpy(nn_p, ip.data(), sizeof(int64_t) * np);
    if (nn_q) std::memcpy(nn_q, iq.data(), sizeof(int64_t) * nq);
    return 0;
}

/* SaveSlice (slice_io.cc:20-120): mv_rows is row-major.  Returns the bytes
 * written (the reference's stream_size) or minus a code. */
int64_t orc_save_slice(const char* path, uint8_t type, int timestamp, int index, uint64_t size, uint8_t qp,
                       const float* mv_rows, const uint8_t* geo, uint64_t geo_size, const uint8_t* col,
                       uint64_t col_size) {
    const std::string name = path ? path : "";
    if (!ends_with(name, ".slice")) return -1;
    if (!(type & 0x01)) return -7;
    const bool g = !(type & 0x02), c = !(type & 0x04);
    if ((g && (!geo || !geo_size)) || (c && (!col || !col_size))) return -8;
    std::FILE* fp = std::fopen(name.c_str(), "wb");
    if (!fp) return errno == ENOENT ? -2 : (errno == EACCES ? -3 : -4);
    int64_t n = 0;
    bool ok = true;
    auto w = [&](const void* p, size_t k) {
        ok = ok && std::fwrite(p, 1, k, fp) == k;
        n += static_cast<int64_t>(k);
    };
    w(&type, 1);
    w(&timestamp, sizeof(int));
    w(&index, sizeof(int));
    w(&size, sizeof(uint64_t));
    w(&qp, 1);
    float cm[16];
    for (int k = 0; k < 16; ++k) cm[k] = mv_rows[(k % 4) * 4 + k / 4];
    w(cm, sizeof(cm));
    if (g) {
        w(&geo_size, sizeof(uint64_t));
        w(geo, geo_size);
    }
    if (c) {
        w(&col_size, sizeof(uint64_t));
        w(col, col_size);
    }
    if (std::fclose(fp) != 0 || !ok) return -10;
    return n;
}

/* LoadSlice of each path, then ChangeSliceToFrame's frame written to out. */
int orc_slices_to_frame(const char* const* paths, int n, const char* out) {
    std::vector<OrcSlice> slices;
    for (int i = 0; i < n; ++i) {
        OrcSlice sl;
        if (int rc = load_slice(paths[i] ? paths[i] : "", sl)) return rc;
        slices.push_back(std::move(sl));
    }
    return write_frame_of(std::move(slices), out ? out : "");
}

}  // extern "C"
