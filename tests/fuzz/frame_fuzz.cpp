// Fuzz harness for the host .frame parser (pvvc_parse_frame and the frame
// accessors), built with AddressSanitizer and UBSan by tests/test_frame_fuzz.py:
// truncations, byte mutations and header bit flips of two valid frames, every
// parsed frame's metadata and blobs read back in full.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <fstream>
#include <iterator>
#include <random>
#include <vector>
#include "pvvc_hip.h"
int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: frame_fuzz ITERATIONS FRAME0 FRAME1\n");
        return 2;
    }
    std::vector<std::vector<uint8_t>> base;
    for (const char* p : {argv[2], argv[3]}) {
        std::ifstream f(p, std::ios::binary);
        base.emplace_back(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    }
    std::mt19937 rng(3);
    int n = atoi(argv[1]), ok = 0, bad = 0;
    for (int t = 0; t < n; ++t) {
        std::vector<uint8_t> b = base[t % 2];
        double r = std::uniform_real_distribution<double>(0, 1)(rng);
        if (r < 0.3) b.resize(rng() % (b.size() + 1));
        else if (r < 0.7) { int k = 1 + rng() % 8; for (int i = 0; i < k; ++i) b[(rng() % 4 == 0) ? rng() % b.size() : rng() % std::min<size_t>(b.size(), 4096)] = rng() & 255; }
        else b[rng() % std::min<size_t>(b.size(), 64)] ^= 1u << (rng() % 8);
        pvvc_frame* f = nullptr;
        int rc = pvvc_parse_frame(b.empty() ? nullptr : b.data(), b.size(), &f);
        if (rc == 0 && f) {
            ++ok;
            uint32_t S = pvvc_frame_slice_cnt(f);
            std::vector<int> idx(S); std::vector<uint8_t> ty(S), qp(S); std::vector<uint32_t> sz(S), gs(S), cs(S); std::vector<float> mv(16ull * S);
            pvvc_frame_meta(f, idx.data(), ty.data(), sz.data(), qp.data(), mv.data(), gs.data(), cs.data());
            volatile uint32_t acc = 0;
            for (uint32_t s = 0; s < S; ++s) {
                const uint8_t* g = pvvc_frame_geometry(f, s); const uint8_t* c = pvvc_frame_color(f, s);
                for (uint32_t i = 0; g && i < gs[s]; ++i) acc += g[i];
                for (uint32_t i = 0; c && i < cs[s]; ++i) acc += c[i];
            }
            pvvc_frame_free(f);
        } else ++bad;
    }
    printf("parsed %d rejected %d\n", ok, bad);
    return 0;
}
