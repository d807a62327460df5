/*
 * frame_io.cpp -- host side of the decode path: the .frame container.
 *
 * pvvc_load_frame / pvvc_parse_frame replace io::LoadFrame
 * (src/cuda/common/slice_io.cu:143-275).  The layout is the one
 * io::ChangeSliceToFrame writes (src/io/slice_io.cc:258-315):
 *
 *   i32 timestamp; u32 S; i32 index[S]; u8 type[S]; u32 size[S]; u8 qp[S];
 *   f32 mv[S][16] (row-major); u32 geo_size[S]; geo blobs (non-empty only);
 *   u32 color_size[S]; colour blobs (non-empty only).
 *
 * A blob whose slice type has bit 3 (geometry) or bit 4 (colour) set is a
 * zstd frame and is unwrapped on load, as the reference loader does
 * (slice_io.cu:233-240,258-265).  Unlike the reference (one `new[]` per
 * slice), every unwrapped blob lands in one host arena, 16-byte aligned,
 * so that a frame can be staged to the device with a single copy.
 */
#include "pvvc_internal.h"

#include <zstd.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <regex>

namespace {

/* Bounds-checked little-endian reader over a memory image. */
struct Cursor {
    const uint8_t* p;
    size_t n, at = 0;
    bool take(void* dst, size_t bytes) {
        if (n - at < bytes) return false;
        std::memcpy(dst, p + at, bytes);
        at += bytes;
        return true;
    }
    const uint8_t* span(size_t bytes) {
        if (n - at < bytes) return nullptr;
        const uint8_t* r = p + at;
        at += bytes;
        return r;
    }
};

size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

/* ZstdDecoder::Decode semantics (src/common/entropy_codec.cc:341-360). */
bool zstd_unwrap(const uint8_t* src, size_t n, std::vector<uint8_t>& out) {
    unsigned long long sz = ZSTD_getFrameContentSize(src, n);
    if (sz == 0 || sz == ZSTD_CONTENTSIZE_UNKNOWN || sz == ZSTD_CONTENTSIZE_ERROR) return false;
    out.resize(sz);
    size_t r = ZSTD_decompress(out.data(), sz, src, n);
    if (ZSTD_isError(r)) return false;
    out.resize(r);
    return true;
}

}  // namespace

int pvvc_frame_parse_image(const uint8_t* bytes, size_t nbytes, pvvc_frame& f) {
    Cursor c{bytes, nbytes};
    if (!c.take(&f.timestamp, 4) || !c.take(&f.slice_cnt, 4)) return 5;
    const uint32_t S = f.slice_cnt;
    /* Refuse a count the image cannot hold before allocating for it. */
    if (static_cast<uint64_t>(S) * (4 + 1 + 4 + 1 + 64 + 4 + 4) > nbytes) return 5;
    f.index.resize(S);
    f.type.resize(S);
    f.size.resize(S);
    f.qp.resize(S);
    f.mv.resize(size_t(S) * 16);
    f.geo_size.resize(S);
    f.col_size.resize(S);
    f.geo_off.assign(S, 0);
    f.col_off.assign(S, 0);
    if (!c.take(f.index.data(), 4 * size_t(S)) || !c.take(f.type.data(), S) || !c.take(f.size.data(), 4 * size_t(S)) ||
        !c.take(f.qp.data(), S) || !c.take(f.mv.data(), 64 * size_t(S)))
        return 5;
    std::vector<uint32_t> raw_geo(S), raw_col(S);
    if (!c.take(raw_geo.data(), 4 * size_t(S))) return 5;
    std::vector<const uint8_t*> gsrc(S, nullptr), csrc(S, nullptr);
    for (uint32_t i = 0; i < S; ++i) {
        if (!raw_geo[i]) continue;
        if (!(gsrc[i] = c.span(raw_geo[i]))) return 5;
    }
    if (!c.take(raw_col.data(), 4 * size_t(S))) return 5;
    for (uint32_t i = 0; i < S; ++i) {
        if (!raw_col[i]) continue;
        if (!(csrc[i] = c.span(raw_col[i]))) return 5;
    }
    /* Unwrap into one arena; sizes become the unwrapped sizes (slice_io.cu:239,264). */
    std::vector<std::vector<uint8_t>> tmp_g(S), tmp_c(S);
    size_t total = 0;
    for (uint32_t i = 0; i < S; ++i) {
        if (gsrc[i] && (f.type[i] & PVVC_TYPE_GEO_ZSTD)) {
            if (!zstd_unwrap(gsrc[i], raw_geo[i], tmp_g[i])) return 6;
            f.geo_size[i] = static_cast<uint32_t>(tmp_g[i].size());
        } else {
            f.geo_size[i] = raw_geo[i];
        }
        if (csrc[i] && (f.type[i] & PVVC_TYPE_COLOR_ZSTD)) {
            if (!zstd_unwrap(csrc[i], raw_col[i], tmp_c[i])) return 6;
            f.col_size[i] = static_cast<uint32_t>(tmp_c[i].size());
        } else {
            f.col_size[i] = raw_col[i];
        }
        total += align16(f.geo_size[i]) + align16(f.col_size[i]);
    }
    f.data.assign(total + 16, 0);
    size_t at = 0;
    f.points = 0;
    for (uint32_t i = 0; i < S; ++i) {
        f.geo_off[i] = at;
        if (f.geo_size[i]) std::memcpy(&f.data[at], tmp_g[i].empty() ? gsrc[i] : tmp_g[i].data(), f.geo_size[i]);
        at += align16(f.geo_size[i]);
        f.col_off[i] = at;
        if (f.col_size[i]) std::memcpy(&f.data[at], tmp_c[i].empty() ? csrc[i] : tmp_c[i].data(), f.col_size[i]);
        at += align16(f.col_size[i]);
        f.points += f.size[i];
    }
    return 0;
}

extern "C" {

int pvvc_parse_frame(const uint8_t* bytes, size_t nbytes, pvvc_frame** out) {
    if (!out || (!bytes && nbytes)) return PVVC_ERR_INVALID_ARG;
    auto* f = new pvvc_frame();
    int rc = pvvc_frame_parse_image(bytes, nbytes, *f);
    if (rc) {
        delete f;
        *out = nullptr;
        return rc;
    }
    *out = f;
    return 0;
}

int pvvc_load_frame(const char* path, pvvc_frame** out) {
    if (!out) return PVVC_ERR_INVALID_ARG;
    *out = nullptr;
    /* slice_io.cu:144-148: the name must end in ".frame" */
    static const std::regex name_type{"^.*\\.frame"};
    if (!path || !std::regex_match(path, name_type)) return 1;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        switch (errno) {
            case ENOENT: return 2;
            case EACCES: return 3;
            default: return 4;
        }
    }
    std::vector<uint8_t> img;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), fp)) > 0) img.insert(img.end(), buf, buf + got);
    std::fclose(fp);
    return pvvc_parse_frame(img.data(), img.size(), out);
}

void pvvc_frame_free(pvvc_frame* f) { delete f; }
int pvvc_frame_timestamp(const pvvc_frame* f) { return f ? f->timestamp : -1; }
uint32_t pvvc_frame_slice_cnt(const pvvc_frame* f) { return f ? f->slice_cnt : 0; }
uint64_t pvvc_frame_points(const pvvc_frame* f) { return f ? f->points : 0; }

void pvvc_frame_meta(const pvvc_frame* f, int* index, uint8_t* type, uint32_t* size, uint8_t* qp, float* mv,
                     uint32_t* geometry_size, uint32_t* color_size) {
    if (!f) return;
    const uint32_t S = f->slice_cnt;
    if (index) std::memcpy(index, f->index.data(), 4 * size_t(S));
    if (type) std::memcpy(type, f->type.data(), S);
    if (size) std::memcpy(size, f->size.data(), 4 * size_t(S));
    if (qp) std::memcpy(qp, f->qp.data(), S);
    if (mv) std::memcpy(mv, f->mv.data(), 64 * size_t(S));
    if (geometry_size) std::memcpy(geometry_size, f->geo_size.data(), 4 * size_t(S));
    if (color_size) std::memcpy(color_size, f->col_size.data(), 4 * size_t(S));
}

const uint8_t* pvvc_frame_geometry(const pvvc_frame* f, uint32_t s) {
    if (!f || s >= f->slice_cnt || !f->geo_size[s]) return nullptr;
    return f->data.data() + f->geo_off[s];
}

const uint8_t* pvvc_frame_color(const pvvc_frame* f, uint32_t s) {
    if (!f || s >= f->slice_cnt || !f->col_size[s]) return nullptr;
    return f->data.data() + f->col_off[s];
}

}  // extern "C"
