/*
 * mans_container.h -- batch container (wire format) of the MI355X tANS decoder library.
 *
 * One contiguous image holds many compressed streams, their decoder tables and,
 * optionally, each stream's sync points (mans_sync_point).  The image is copied to the
 * device as it is (one host-to-device copy); mans_container_descs then points the batch
 * decoder's stream descriptors into that device copy, so a whole container decodes with
 * one mans_decode_batch call, and windows of its streams decode alone from the stored
 * sync points (mans_stream_range).
 *
 * No counterpart in the reference: its demo keeps one stream in CUHDInputBuffer
 * (src/multians_input_buffer.cc:15-30) with the first state and bit passed beside it
 * (src/main.cc:100-123), and its sync points (CUHDSubsequenceSyncPoint,
 * include/cuhd_subsequence_sync_point.h:20-25) live in decoder memory for one call.
 * The per-stream fields are the ones that call takes (units, their count, the number of
 * symbols, table, first state, first bit).
 *
 * Layout (little-endian; offsets from the image start; every array 256-byte aligned):
 *   header, 128 bytes:
 *     char magic[8] = "MANSCTR1"; u32 version = 1; u32 header_bytes = 128;
 *     u32 n_streams; u32 n_tables; u64 tables_off; u64 streams_off; u64 image_bytes;
 *     zeros to 128 bytes
 *   table directory, n_tables x 16 bytes:   u32 L; u32 zero; u64 words_off (L u32,
 *                                           CUHDCodetableItem layout)
 *   stream directory, n_streams x 64 bytes: u64 units_off; u64 n_units; u64 n_symbols;
 *     u32 table; u32 first_state; u32 first_bit; u32 chunk_units (0 = no sync points);
 *     u64 sync_off (mans_sync_point_count(n_units, chunk_units) records, 0 = none);
 *     u64 zero[2]
 *   payload: the tables, the streams' units (decoder order, zero-padded to whole
 *     16-byte groups) and the sync point arrays.
 */
#ifndef MANS_CONTAINER_H_
#define MANS_CONTAINER_H_

#include <stddef.h>
#include <stdint.h>

#include "mans_c_api.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MANS_CONTAINER_VERSION 1u

typedef struct mans_container_table {
    const uint32_t* words;          /* host, L entries */
    uint32_t L;
    uint32_t reserved;
} mans_container_table;

typedef struct mans_container_stream {
    const uint32_t* units;          /* host, n_units (decoder order) */
    uint64_t n_units;
    uint64_t n_symbols;
    uint32_t table;                 /* index into the tables */
    uint32_t first_state;
    uint32_t first_bit;
    uint32_t chunk_units;           /* of sync_points (0 = none) */
    const mans_sync_point* sync_points;   /* host, mans_sync_point_count records, or NULL */
} mans_container_stream;

typedef struct mans_container_info {
    uint32_t n_streams;
    uint32_t n_tables;
    uint32_t n_with_sync_points;
    uint32_t reserved;
    uint64_t total_units;
    uint64_t total_symbols;
    uint64_t image_bytes;
} mans_container_info;

/* Size of the image for these tables and streams. */
int mans_container_bytes(const mans_container_table* tables, size_t n_tables,
                         const mans_container_stream* streams, size_t n_streams, uint64_t* bytes);
/* Writes the image (image_bytes >= mans_container_bytes). */
int mans_container_pack(const mans_container_table* tables, size_t n_tables,
                        const mans_container_stream* streams, size_t n_streams, void* image,
                        uint64_t image_bytes);
/* Validates an image (host copy of `bytes` bytes): header, directories, every table's L
 * and bounds, every stream's fields and bounds, and every sync point array
 * (mans_sync_points_check).  MANS_ECORRUPT names the first fault. */
int mans_container_open(const void* image, uint64_t bytes, mans_container_info* info);
/* Stream descriptors for a validated image whose copy is resident at device_image:
 * units, tables and (with use_sync_points) index_in/chunk_units point into that copy;
 * outputs are packed back to back at device_out (out_bytes >= total_symbols), stream i
 * at out_offsets[i] (n_streams + 1 entries, optional).  index_out is NULL. */
int mans_container_descs(const void* image, uint64_t bytes, const void* device_image,
                         uint8_t* device_out, uint64_t out_bytes, int use_sync_points,
                         mans_stream_desc* descs, size_t n_descs, uint64_t* out_offsets);
/* Host pointer to stream i's sync points inside an image that mans_container_open
 * accepted (NULL if it has none); *count = number of records.  Checks only this
 * stream's record, so it is cheap per window of a partial decode. */
const mans_sync_point* mans_container_sync_points(const void* image, uint64_t bytes,
                                                  uint32_t stream, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* MANS_CONTAINER_H_ */
