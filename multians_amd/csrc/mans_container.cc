// mans_container.cc -- the batch container (wire format): packing, validation and the
// batch decoder's stream descriptors for an image resident on the device.  Host code
// only; the layout is described in include/mans_container.h.
#include <cstring>
#include <vector>

#include "../../include/mans_container.h"
#include "mans_internal.h"

namespace {

constexpr char kMagic[8] = {'M', 'A', 'N', 'S', 'C', 'T', 'R', '1'};
constexpr uint64_t kHeader = 128, kTableRec = 16, kStreamRec = 64, kAlign = 256;
constexpr uint32_t kMinChunk = 16, kMaxChunk = 1u << 24;

struct Header {
    char magic[8];
    uint32_t version, header_bytes, n_streams, n_tables;
    uint64_t tables_off, streams_off, image_bytes;
};
struct TableRec {
    uint32_t L, zero;
    uint64_t words_off;
};
struct StreamRec {
    uint64_t units_off, n_units, n_symbols;
    uint32_t table, first_state, first_bit, chunk_units;
    uint64_t sync_off, zero[2];
};
static_assert(sizeof(Header) <= kHeader, "header");
static_assert(sizeof(TableRec) == kTableRec && sizeof(StreamRec) == kStreamRec, "records");

uint64_t align(uint64_t v) { return (v + kAlign - 1u) & ~(kAlign - 1u); }
bool pow2_in(uint64_t v, uint64_t lo, uint64_t hi) { return v >= lo && v <= hi && !(v & (v - 1u)); }
// units stored per stream: whole 16-byte groups (the decoder's loads)
uint64_t stored_units(uint64_t n) { return (n + 3u) & ~3ull; }
// [off, off + len) inside [0, size), overflow-safe
bool inside(uint64_t off, uint64_t len, uint64_t size) { return off <= size && len <= size - off; }

uint64_t sync_count(uint64_t n_units, uint32_t chunk) {
    return chunk ? (n_units + chunk - 1u) / chunk : 0u;
}

int check_inputs(const mans_container_table* t, size_t nt, const mans_container_stream* s,
                 size_t ns) {
    if ((nt && !t) || (ns && !s)) return set_error(MANS_EINVAL, "NULL tables/streams");
    if (nt > 0xFFFFFFFFu || ns > 0xFFFFFFFFu) return set_error(MANS_EINVAL, "too many entries");
    for (size_t i = 0; i < nt; ++i)
        if (!t[i].words || !pow2_in(t[i].L, 2, 16384))
            return set_error(MANS_EINVAL, "table %zu: L=%u (power of two in [2, 16384])", i,
                             t[i].L);
    for (size_t i = 0; i < ns; ++i) {
        const mans_container_stream& x = s[i];
        if (x.table >= nt) return set_error(MANS_EINVAL, "stream %zu: table %u", i, x.table);
        if (x.n_units && !x.units) return set_error(MANS_EINVAL, "stream %zu: units NULL", i);
        if (x.n_units >= 0xFFFFFF00ull) return set_error(MANS_EINVAL, "stream %zu: n_units", i);
        if (x.sync_points && !pow2_in(x.chunk_units, kMinChunk, kMaxChunk))
            return set_error(MANS_EINVAL, "stream %zu: chunk_units=%u", i, x.chunk_units);
        if (x.sync_points) {
            int rc = mans_sync_points_check(x.sync_points, sync_count(x.n_units, x.chunk_units),
                                            t[x.table].L);
            if (rc) return set_error(rc, "stream %zu: %s", i, mans_last_error());
        }
    }
    return MANS_OK;
}

// Offsets of every array in pack order; returns the image size.
uint64_t plan(const mans_container_table* t, size_t nt, const mans_container_stream* s,
              size_t ns, uint64_t* toff, uint64_t* uoff, uint64_t* soff, uint64_t* tables_off,
              uint64_t* streams_off) {
    uint64_t o = kHeader;
    *tables_off = o;
    o += kTableRec * nt;
    *streams_off = o;
    o += kStreamRec * ns;
    o = align(o);
    for (size_t i = 0; i < nt; ++i) {
        if (toff) toff[i] = o;
        o = align(o + 4ull * t[i].L);
    }
    for (size_t i = 0; i < ns; ++i) {
        if (uoff) uoff[i] = o;
        o = align(o + 4ull * stored_units(s[i].n_units));
    }
    for (size_t i = 0; i < ns; ++i) {
        const bool has = s[i].sync_points != nullptr;
        if (soff) soff[i] = has ? o : 0u;
        if (has) o = align(o + sizeof(mans_sync_point) * sync_count(s[i].n_units, s[i].chunk_units));
    }
    return o;
}

// Validated view of an image.
struct View {
    const char* b;
    Header h;
    const TableRec* tabs;
    const StreamRec* strs;
};

int open_view(const void* image, uint64_t bytes, View* v, mans_container_info* info) {
    if (!image) return set_error(MANS_EINVAL, "image is NULL");
    if (bytes < kHeader) return set_error(MANS_ECORRUPT, "image of %llu bytes < header",
                                          (unsigned long long)bytes);
    if (((uintptr_t)image) & 7u) return set_error(MANS_EINVAL, "image must be 8-byte aligned");
    v->b = (const char*)image;
    std::memcpy(&v->h, v->b, sizeof(Header));
    const Header& h = v->h;
    if (std::memcmp(h.magic, kMagic, 8) != 0) return set_error(MANS_ECORRUPT, "bad magic");
    if (h.version != MANS_CONTAINER_VERSION)
        return set_error(MANS_ECORRUPT, "version %u (this library reads %u)", h.version,
                         MANS_CONTAINER_VERSION);
    if (h.header_bytes != kHeader) return set_error(MANS_ECORRUPT, "header_bytes %u", h.header_bytes);
    if (h.image_bytes > bytes || h.image_bytes < kHeader)
        return set_error(MANS_ECORRUPT, "image_bytes %llu, %llu given",
                         (unsigned long long)h.image_bytes, (unsigned long long)bytes);
    const uint64_t size = h.image_bytes;
    if ((h.tables_off & 7u) || (h.streams_off & 7u) ||
        !inside(h.tables_off, kTableRec * (uint64_t)h.n_tables, size) ||
        !inside(h.streams_off, kStreamRec * (uint64_t)h.n_streams, size) ||
        h.tables_off < kHeader || h.streams_off < kHeader)
        return set_error(MANS_ECORRUPT, "directories outside the image");
    v->tabs = (const TableRec*)(v->b + h.tables_off);
    v->strs = (const StreamRec*)(v->b + h.streams_off);
    for (uint32_t i = 0; i < h.n_tables; ++i) {
        const TableRec& t = v->tabs[i];
        if (!pow2_in(t.L, 2, 16384)) return set_error(MANS_ECORRUPT, "table %u: L=%u", i, t.L);
        if ((t.words_off & 15u) || !inside(t.words_off, 4ull * t.L, size))
            return set_error(MANS_ECORRUPT, "table %u outside the image", i);
    }
    uint64_t tu = 0, tsym = 0;
    uint32_t nsync = 0;
    for (uint32_t i = 0; i < h.n_streams; ++i) {
        const StreamRec& s = v->strs[i];
        if (s.table >= h.n_tables) return set_error(MANS_ECORRUPT, "stream %u: table %u", i, s.table);
        const uint32_t L = v->tabs[s.table].L;
        if (s.n_units >= 0xFFFFFF00ull) return set_error(MANS_ECORRUPT, "stream %u: n_units", i);
        if ((s.units_off & 15u) || !inside(s.units_off, 4ull * stored_units(s.n_units), size))
            return set_error(MANS_ECORRUPT, "stream %u: units outside the image", i);
        if (s.n_symbols && !s.n_units)
            return set_error(MANS_ECORRUPT, "stream %u: symbols without units", i);
        if (s.first_bit > 32u) return set_error(MANS_ECORRUPT, "stream %u: first_bit", i);
        if (s.n_symbols && (s.first_state < L || s.first_state >= 2u * L))
            return set_error(MANS_ECORRUPT, "stream %u: first_state %u", i, s.first_state);
        if (s.sync_off) {
            if (!pow2_in(s.chunk_units, kMinChunk, kMaxChunk))
                return set_error(MANS_ECORRUPT, "stream %u: chunk_units %u", i, s.chunk_units);
            const uint64_t n = sync_count(s.n_units, s.chunk_units);
            if ((s.sync_off & 7u) || !inside(s.sync_off, sizeof(mans_sync_point) * n, size))
                return set_error(MANS_ECORRUPT, "stream %u: sync points outside the image", i);
            int rc = mans_sync_points_check((const mans_sync_point*)(v->b + s.sync_off), n, L);
            if (rc) return set_error(rc, "stream %u: %s", i, mans_last_error());
            ++nsync;
        } else if (s.chunk_units) {
            return set_error(MANS_ECORRUPT, "stream %u: chunk_units without sync points", i);
        }
        tu += s.n_units;
        tsym += s.n_symbols;
    }
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->n_streams = h.n_streams;
        info->n_tables = h.n_tables;
        info->n_with_sync_points = nsync;
        info->total_units = tu;
        info->total_symbols = tsym;
        info->image_bytes = size;
    }
    return MANS_OK;
}

}  // namespace

extern "C" int mans_container_bytes(const mans_container_table* tables, size_t n_tables,
                                    const mans_container_stream* streams, size_t n_streams,
                                    uint64_t* bytes) {
    if (!bytes) return set_error(MANS_EINVAL, "bytes is NULL");
    int rc = check_inputs(tables, n_tables, streams, n_streams);
    if (rc) return rc;
    uint64_t a, b;
    *bytes = plan(tables, n_tables, streams, n_streams, nullptr, nullptr, nullptr, &a, &b);
    return MANS_OK;
}

extern "C" int mans_container_pack(const mans_container_table* tables, size_t n_tables,
                                   const mans_container_stream* streams, size_t n_streams,
                                   void* image, uint64_t image_bytes) {
    if (!image) return set_error(MANS_EINVAL, "image is NULL");
    int rc = check_inputs(tables, n_tables, streams, n_streams);
    if (rc) return rc;
    std::vector<uint64_t> toff(n_tables), uoff(n_streams), soff(n_streams);
    uint64_t tables_off, streams_off;
    const uint64_t size = plan(tables, n_tables, streams, n_streams, toff.data(), uoff.data(),
                               soff.data(), &tables_off, &streams_off);
    if (image_bytes < size)
        return set_error(MANS_ENOMEM, "image of %llu bytes < %llu needed",
                         (unsigned long long)image_bytes, (unsigned long long)size);
    char* b = (char*)image;
    std::memset(b, 0, size);
    Header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, 8);
    h.version = MANS_CONTAINER_VERSION;
    h.header_bytes = kHeader;
    h.n_streams = (uint32_t)n_streams;
    h.n_tables = (uint32_t)n_tables;
    h.tables_off = tables_off;
    h.streams_off = streams_off;
    h.image_bytes = size;
    std::memcpy(b, &h, sizeof(h));
    for (size_t i = 0; i < n_tables; ++i) {
        TableRec t{tables[i].L, 0u, toff[i]};
        std::memcpy(b + tables_off + kTableRec * i, &t, sizeof(t));
        std::memcpy(b + toff[i], tables[i].words, 4ull * tables[i].L);
    }
    for (size_t i = 0; i < n_streams; ++i) {
        const mans_container_stream& x = streams[i];
        StreamRec r;
        std::memset(&r, 0, sizeof(r));
        r.units_off = uoff[i];
        r.n_units = x.n_units;
        r.n_symbols = x.n_symbols;
        r.table = x.table;
        r.first_state = x.first_state;
        r.first_bit = x.first_bit;
        r.chunk_units = x.sync_points ? x.chunk_units : 0u;
        r.sync_off = soff[i];
        std::memcpy(b + streams_off + kStreamRec * i, &r, sizeof(r));
        if (x.n_units) std::memcpy(b + uoff[i], x.units, 4ull * x.n_units);
        if (x.sync_points)
            std::memcpy(b + soff[i], x.sync_points,
                        sizeof(mans_sync_point) * sync_count(x.n_units, x.chunk_units));
    }
    return MANS_OK;
}

extern "C" int mans_container_open(const void* image, uint64_t bytes, mans_container_info* info) {
    View v;
    return open_view(image, bytes, &v, info);
}

extern "C" int mans_container_descs(const void* image, uint64_t bytes, const void* device_image,
                                    uint8_t* device_out, uint64_t out_bytes, int use_sync_points,
                                    mans_stream_desc* descs, size_t n_descs,
                                    uint64_t* out_offsets) {
    View v;
    mans_container_info info;
    int rc = open_view(image, bytes, &v, &info);
    if (rc) return rc;
    if (!device_image || (!descs && info.n_streams))
        return set_error(MANS_EINVAL, "NULL argument");
    if (((uintptr_t)device_image) & 255u)
        return set_error(MANS_EINVAL, "device_image must be 256-byte aligned");
    if (n_descs < info.n_streams) return set_error(MANS_EINVAL, "%zu descs < %u streams", n_descs,
                                                   info.n_streams);
    if (info.total_symbols && (!device_out || out_bytes < info.total_symbols))
        return set_error(MANS_EINVAL, "output of %llu bytes < %llu symbols",
                         (unsigned long long)out_bytes, (unsigned long long)info.total_symbols);
    const char* dev = (const char*)device_image;
    uint64_t o = 0;
    for (uint32_t i = 0; i < info.n_streams; ++i) {
        const StreamRec& s = v.strs[i];
        mans_stream_desc& d = descs[i];
        std::memset(&d, 0, sizeof(d));
        d.units = (const uint32_t*)(dev + s.units_off);
        d.n_units = s.n_units;
        d.out = device_out ? device_out + o : nullptr;
        d.n_out = s.n_symbols;
        d.table = (const uint32_t*)(dev + v.tabs[s.table].words_off);
        d.L = v.tabs[s.table].L;
        d.first_state = s.first_state;
        d.first_bit = s.first_bit;
        if (use_sync_points && s.sync_off) {
            d.chunk_units = s.chunk_units;
            d.index_in = (const mans_sync_point*)(dev + s.sync_off);
        }
        if (out_offsets) out_offsets[i] = o;
        o += s.n_symbols;
    }
    if (out_offsets) out_offsets[info.n_streams] = o;
    return MANS_OK;
}

extern "C" const mans_sync_point* mans_container_sync_points(const void* image, uint64_t bytes,
                                                             uint32_t stream, uint64_t* count) {
    // bounds of this one record only (the image was validated by mans_container_open)
    if (count) *count = 0;
    Header h;
    if (!image || bytes < kHeader) {
        set_error(MANS_EINVAL, "image is NULL or short");
        return nullptr;
    }
    const char* b = (const char*)image;
    std::memcpy(&h, b, sizeof(h));
    const uint64_t size = h.image_bytes < bytes ? h.image_bytes : bytes;
    if (std::memcmp(h.magic, kMagic, 8) != 0 || stream >= h.n_streams ||
        !inside(h.streams_off, kStreamRec * ((uint64_t)stream + 1u), size)) {
        set_error(MANS_EINVAL, "stream %u not in this image", stream);
        return nullptr;
    }
    StreamRec s;
    std::memcpy(&s, b + h.streams_off + kStreamRec * stream, sizeof(s));
    if (!s.sync_off) return nullptr;
    const uint64_t n = sync_count(s.n_units, s.chunk_units);
    if (!inside(s.sync_off, sizeof(mans_sync_point) * n, size)) {
        set_error(MANS_ECORRUPT, "stream %u: sync points outside the image", stream);
        return nullptr;
    }
    if (count) *count = n;
    return (const mans_sync_point*)(b + s.sync_off);
}
