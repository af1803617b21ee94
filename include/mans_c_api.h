/*
 * mans_c_api.h -- C ABI of the MI355X tANS decoder library (libmultians_amd.so).
 *
 * Plain pointers and sizes only; no HIP or torch types.  Device pointers are
 * passed as void* / typed pointers into device memory; `stream` is a hipStream_t
 * passed as void* (NULL = the default stream).  Every entry point returns 0 on
 * success or a negative MANS_E* code; mans_last_error() describes the last failure
 * on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   mans_decode            cuhd::CUHDGPUDecoder::decode
 *                          (include/cuhd_gpu_decoder.h:24-35, src/multians_gpu_decoder.cu:575-687)
 *   mans_workspace_bytes   sizing done by cuhd::CUHDGPUDecoderMemory's constructor
 *                          (src/multians_gpu_decoder_memory.cc:13-19)
 *   mans_malloc/mans_free  CUHDGPUMemoryBuffer<T>::allocate/free (zero-filled allocate,
 *                          include/cuhd_gpu_memory_buffer.hxx:35-49)
 *   mans_memcpy_h2d/d2h    CUHDGPUMemoryBuffer<T>::cpy_host_to_device / cpy_device_to_host
 *                          (include/cuhd_gpu_memory_buffer.hxx:52-67)
 *   mans_host_alloc_pinned cudaMallocHost in CUHDGPUDecoderMemory::allocate
 *                          (src/multians_gpu_decoder_memory.cc:45-47)
 *   mans_set_device        cudaSetDevice (src/main.cc:200)
 *   mans_error_string      the message CUERR prints (include/cuhd_cuda_definitions.h:21-28)
 *   mans_build_tables,     ANSTableGenerator::generate_table / generate_encoder_table /
 *   mans_encode            get_decoder_table and ANSEncoder::encode
 *                          (src/ans_table_generator.cc:164-288, src/ans_encoder.cc:13-97):
 *                          host-side producers of the decoder's inputs
 *   mans_decode_batch      no single counterpart: decodes many independent streams (each
 *                          with its own table) in one launch sequence
 */
#ifndef MANS_C_API_H_
#define MANS_C_API_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define MANS_OK 0
#define MANS_EINVAL (-1)      /* bad argument (null pointer, size, L, ...)            */
#define MANS_EHIP (-2)        /* HIP runtime error                                    */
#define MANS_ENOMEM (-3)      /* allocation failed / workspace too small              */
#define MANS_ETABLE (-4)      /* decoder table rejected (see mans_decode)             */
#define MANS_ECORRUPT (-5)    /* stream inconsistent with table/sizes                 */
#define MANS_EUNSUPPORTED (-6)

const char* mans_error_string(int code);
/* Message for the last failing call on this thread (never NULL). */
const char* mans_last_error(void);

/* ---------------------------------------------------- device / memory / copy */
int mans_set_device(int device);
int mans_get_device_count(int* count);
/* Zero-filled device allocation (like CUHDGPUMemoryBuffer<T>::allocate). */
int mans_malloc(void** ptr, size_t bytes);
int mans_free(void* ptr);
int mans_memset(void* dst, int value, size_t bytes, void* stream);
int mans_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int mans_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int mans_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int mans_host_alloc_pinned(void** ptr, size_t bytes);
int mans_host_free_pinned(void* ptr);
int mans_stream_create(void** stream);
int mans_stream_destroy(void* stream);
int mans_sync(void* stream);            /* hipStreamSynchronize (NULL: device) */
int mans_event_create(void** ev);
int mans_event_destroy(void* ev);
int mans_event_record(void* ev, void* stream);
int mans_event_elapsed_ms(void* start, void* stop, float* ms); /* synchronises stop */

/* ------------------------------------------------------------- single stream */
/*
 * Stream format (the reference's CUHDInputBuffer, src/multians_input_buffer.cc:15-30):
 *   units[0..n_units)  u32 in DECODER order (reversed encoder emission); the decoder
 *                      reads zeros beyond n_units, so no padding is required (the
 *                      reference's +4 zero units are accepted and ignored).
 *   first_bit          valid bits in units[0] (left-aligned); decoding starts at bit
 *                      32 - first_bit of units[0], bits are taken LSB-first.
 *   first_state        initial state, in [L, 2L).
 * Table: L u32 entries, reference CUHDCodetableItem layout
 *   (include/cuhd_codetable.h:20-24): bits 0-15 next_state, 16-23 symbol,
 *   24-31 min_num_bits.  Supported: 2 <= L <= 16384 and every entry satisfies
 *   L <= next<<nbits and (next+1)<<nbits <= 2L (always true for power-of-two L and
 *   tables built by ANSTableGenerator); other tables return MANS_ETABLE.
 * Output: out[0..n_out) symbols in decode order (reverse of the encoder input), the
 *   first n_out steps of the decoder started at (first_bit, first_state).
 */
/* Sync point of a stream at the start of sync chunk c (units [c*chunk_units, ...)): the
 * decoder state (x - L) and the bit offset, within unit c*chunk_units, of the first step
 * that starts in that unit, and the number of symbols decoded before that step.  One
 * record per chunk, ceil(n_units / chunk_units) records (mans_sync_point_count).  A
 * decode writes them (index_out) once its sync phase has found the true path; a later
 * decode of the same stream with the same chunk_units starts its speculative chunk
 * decodes there (index_in) instead of at a guessed state, and a window of the output can
 * be decoded alone from the chunk that holds its first symbol (mans_stream_range).
 * No counterpart in the reference, whose sync points (CUHDSubsequenceSyncPoint,
 * include/cuhd_subsequence_sync_point.h:20-25) live in decoder memory for one call. */
typedef struct mans_sync_point {
    uint32_t state;             /* x - L, in [0, L) */
    uint32_t bit;               /* in [0, 31]; chunk 0: 32 - first_bit (0..32) */
    uint64_t symbols_before;
} mans_sync_point;

typedef struct mans_decode_args {
    const uint32_t* units;      /* device */
    uint64_t n_units;
    uint8_t* out;               /* device, n_out bytes */
    uint64_t n_out;
    const uint32_t* table;      /* device, L entries */
    uint32_t L;
    uint32_t first_state;
    uint32_t first_bit;
    uint32_t chunk_units;       /* sync-chunk length in units; 0 = tuned */
    uint32_t checkpoint_units;  /* checkpoint stride in units; 0 = tuned */
    uint32_t flags;             /* MANS_F_* */
    void* workspace;            /* device, >= mans_workspace_bytes(...) */
    uint64_t workspace_bytes;
    void* stream;               /* hipStream_t or NULL */
    void** kernel_events;       /* optional: MANS_N_KERNEL_EVENTS hipEvent_t recorded on
                                   `stream` around the kernels (timing), or NULL */
    /* Optional sync points (device; both need chunk_units, a power of two in
     * [16, 2^24], used as the sync-chunk length as given).  index_in: start the
     * speculative decodes there; the decode still proves every chunk's start, so wrong
     * records cost time, not correctness.  index_out: written with the stream's true
     * sync points (may equal index_in). */
    const mans_sync_point* index_in;
    mans_sync_point* index_out;
    /* Optional (device): a record per 32 units (mans_sync_point_count(n_units, 32)),
     * written by the write pass whatever the sync-chunk length; what
     * MANS_F_TRUST_SYNC_POINTS reads back as index_in with chunk_units 32. */
    mans_sync_point* items_out;
    /* Optional: the reference decoder's own sync state, filled for a caller that inspects
     * it (cuhd::CUHDGPUDecoderMemory's accessors); NULL = none.  Read at enqueue. */
    const struct mans_ref_sync* ref_sync;
} mans_decode_args;

/* The arrays the reference decoder leaves in CUHDGPUDecoderMemory after a decode
 * (src/multians_gpu_decoder_memory.cc:27-44; written at src/multians_gpu_decoder.cu:
 * 335-353, 593-596, 635, 659-670), for subsequences of `subsequence_units` units
 * (n_sub = ceil(n_units / subsequence_units)) and blocks of `threads` subsequences:
 *   sync_points[4j .. 4j+3]  CUHDSubsequenceSyncPoint {unit, bit, output_size, state} of
 *                            subsequence j: the unit (relative to the subsequence's
 *                            first unit) and bit where its last step starts, the number
 *                            of symbols before its first step (the exclusive scan phase 3
 *                            leaves in output_size; low 32 bits, the reference's type), and
 *                            the state x (in [L, 2L)) before its last step.  A step
 *                            belongs to the subsequence holding the unit it starts in;
 *                            zero-bit steps after the last unit belong to the last one.
 *   output_sizes[j]          the same exclusive scan (or NULL)
 *   synced_flags[b]          1 per block when the sync phase converged, else 0 (or NULL)
 * All device memory; only the first n_out steps are described. */
typedef struct mans_ref_sync {
    uint32_t* sync_points;
    uint32_t* output_sizes;
    uint8_t* synced_flags;
    uint32_t subsequence_units;
    uint32_t threads;
} mans_ref_sync;

/* kernel_events[i] is recorded before kernel i (0 init, 1 sync, 2 check, 3 fixup,
 * 4 scan, 5 write); kernel_events[6] after the last one. */
#define MANS_N_KERNEL_EVENTS 7

#define MANS_F_SYNC_ERRORS 1u   /* synchronise and report corrupt-stream errors */
/* Trust index_in at chunk_units 32 (a record per 32 units, as items_out writes them; no
 * index_out): the write pass alone decodes each 32-unit item from its record, with no
 * proof (wrong records give wrong output; writes stay inside out[0..n_out)).  For
 * records this library wrote for the same stream and that the caller kept intact, or a
 * slice of them from mans_stream_range (symbols_before counts from index_in[0]). */
#define MANS_F_TRUST_SYNC_POINTS 2u

/* Upper bound of the workspace any tuning choice needs for n_units. */
int mans_workspace_bytes(uint64_t n_units, uint32_t L, uint64_t* bytes);
/* Workspace for an explicit geometry (chunk_units / checkpoint_units as in the args). */
int mans_workspace_bytes_geometry(uint64_t n_units, uint32_t chunk_units,
                                  uint32_t checkpoint_units, uint64_t* bytes);
/* Enqueue the decode on args->stream.  Returns after enqueueing (like the
 * reference, which returns before its last kernel completes) unless
 * MANS_F_SYNC_ERRORS is set.  Integrity faults found on the device are reported
 * by mans_decode_status(). */
int mans_decode(const mans_decode_args* args);
/* mans_decode without a workspace argument, for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  args->workspace must be NULL and
 * args->workspace_bytes 0: the decode uses a static device array of 64 MiB held by the
 * library on each GPU it runs on (until the process ends).  A decode at the automatic
 * chunk lengths needs a little over 1 byte per unit, so streams of up to about 50 million
 * units (200 MB compressed) fit; a stream whose workspace would exceed 64 MiB is rejected
 * with MANS_ENOMEM before anything is enqueued.
 * Allocates nothing.  Every call on one GPU shares that GPU's array, so calls that overlap
 * on one GPU (other streams, other threads) race; serialise them.  Device-side errors
 * are reported only with MANS_F_SYNC_ERRORS (synchronises args->stream). */
int mans_decode_noalloc(const mans_decode_args* args);
/* Fills args->ref_sync from the workspace of the last decode enqueued with these same
 * arguments (same stream, sizes, geometry and workspace, not reused since), enqueued on
 * args->stream, without decoding again: for a caller that inspects the reference decoder's
 * sync state only sometimes (cuhd::CUHDGPUDecoderMemory's accessors).  Replaces the
 * reading of sync_info / output_sizes / sequence_synced after the reference's decode
 * (src/multians_gpu_decoder.cu:591-596). */
int mans_ref_sync_fill(const mans_decode_args* args);
/* Device-side status word of the last decode that used `workspace`:
 * 0 = ok, else a MANS_E* code (synchronises `stream`). */
int mans_decode_status(void* workspace, void* stream, int* status);

/* --------------------------------------------------------------------- batch */
typedef struct mans_stream_desc {
    const uint32_t* units;      /* device */
    uint64_t n_units;
    uint8_t* out;               /* device */
    uint64_t n_out;
    const uint32_t* table;      /* device, L entries */
    uint32_t L;
    uint32_t first_state;
    uint32_t first_bit;
    uint32_t chunk_units;       /* sync-chunk length; 0 = tuned (then no index_in/out) */
    const mans_sync_point* index_in;   /* optional, as in mans_decode_args */
    mans_sync_point* index_out;        /* optional, as in mans_decode_args */
    mans_sync_point* items_out;        /* optional, as in mans_decode_args */
} mans_stream_desc;

/* mans_batch_opts.flags: decode as one pipeline even when the batch is large enough to
 * be split into concurrent parts (a caller that keeps several batches in flight itself). */
#define MANS_BATCH_ONE_PIPELINE 1u
/* mans_batch_opts.flags: every stream decodes from its trusted index_in, as with
 * MANS_F_TRUST_SYNC_POINTS (also honoured by mans_batch_graph_create). */
#define MANS_BATCH_TRUST_SYNC_POINTS 2u

typedef struct mans_batch_opts {
    void* workspace;            /* device, >= mans_batch_workspace_bytes(...) */
    uint64_t workspace_bytes;
    void* stream;
    uint32_t flags;             /* MANS_BATCH_* */
    uint32_t reserved;
    void** kernel_events;       /* optional, as in mans_decode_args */
    /* Optional (device, n u32, caller order): each stream's own status word, zeroed at the
     * start of the decode; 0 = ok, bit 0 = its table was rejected (nothing of it is
     * written), bit 1 = its sync rounds did not converge (mans_set_sync_round_limit).
     * mans_decode_status reports the OR over the batch. */
    uint32_t* stream_status;
} mans_batch_opts;

typedef struct mans_batch_stats {
    uint64_t total_units;
    uint64_t total_symbols;
    uint32_t n_streams;
    uint32_t n_chunks;
    uint32_t n_items;
    uint32_t status;
} mans_batch_stats;

/* Workspace for mans_decode_batch / mans_batch_graph_create over these streams (split or
 * one pipeline). */
int mans_batch_workspace_bytes(const mans_stream_desc* descs, size_t n, uint64_t* bytes);
/* Decodes every stream of the batch, each with its own table, enqueued on opts->stream.
 * Batches of 512 streams or more run as two or three concurrent pipelines on private
 * streams, joined back into opts->stream (unless MANS_BATCH_ONE_PIPELINE or
 * kernel_events).  The plans are copied from pinned host buffers owned by the library
 * (two per caller stream), so this call must not be captured into a graph; use
 * mans_batch_graph_create for that.  Returns after enqueueing; stats (optional) are
 * host-side totals. */
int mans_decode_batch(const mans_stream_desc* descs, size_t n, const mans_batch_opts* opts,
                      mans_batch_stats* stats);

/* hipGraph form of mans_decode_batch (no reference counterpart: the reference launches
 * its phases one by one from the host).  Plans the batch, uploads the plans into
 * opts->workspace (synchronously), captures the decode's kernel sequence into a graph
 * and instantiates it.  Each mans_graph_launch then decodes the same streams with no
 * host-side planning: the device buffers named in descs and the workspace must stay
 * allocated and unchanged in size; their contents (new units of the same lengths, new
 * tables) may change between launches.  opts->kernel_events is ignored. */
int mans_batch_graph_create(const mans_stream_desc* descs, size_t n, const mans_batch_opts* opts,
                            void** graph);
int mans_graph_launch(void* graph, void* stream);
int mans_graph_destroy(void* graph);

/* Chunk-length rules (the autotuner's output; no reference counterpart: the reference
 * takes one subsequence size from its caller).  A stream uses the first rule of its scope
 * with L <= L_max, symbols per unit (n_out / n_units) <= spu_max and n_units <= units_max;
 * streams no rule matches use the built-in choice.  chunk_units is rounded down to a
 * power of two and must be >= 32.  warm_units: units each chunk's speculative decode
 * starts early (rounded down to the 16-unit checkpoint stride), or MANS_WARM_AUTO for the
 * built-in choice.  scope: MANS_RULE_SINGLE (mans_decode), MANS_RULE_BATCH
 * (mans_decode_batch and the chunk length mans_chunk_units reports), or both (0).
 * n == 0 restores the built-in choice.  The rules are process-wide; set them before
 * decoding. */
#define MANS_WARM_AUTO 0xFFFFFFFFu
#define MANS_RULE_SINGLE 1u
#define MANS_RULE_BATCH 2u
/* Lineage rules (scope exactly MANS_RULE_LINEAGE; chunk_units = D, rounded down to a
 * multiple of 32, or 0 = the chunked pipeline for the streams the rule matches, so that a
 * later rule does not catch them; warm_units unused): a single-stream decode (mans_decode with automatic
 * geometry, no sync points in or out and no ref_sync) that a lineage rule matches finds its
 * true path by lineages -- a guessed-state decode starts every D units and runs until it
 * meets a path another lineage recorded at a 32-unit checkpoint -- and decodes it from the
 * resulting 32-unit item records with the write pass alone.  When the records do not hold
 * (a checkpoint's slot list full, a lineage longer than the cap, maps that do not resolve,
 * a rejected table), the same call runs the chunked pipeline instead, on the device, with
 * no host round trip.  The workspace must hold mans_workspace_bytes(n_units) (the lineage
 * layout is part of that bound); with less, the chunked pipeline runs.  Rules of other
 * scopes never select lineages.  No reference counterpart (the reference synchronises
 * subsequences only, src/multians_gpu_decoder.cu:195-466). */
#define MANS_RULE_LINEAGE 4u
typedef struct mans_geometry_rule {
    uint32_t L_max;
    uint32_t chunk_units;
    double spu_max;
    uint64_t units_max;
    uint32_t warm_units;
    uint32_t scope;
} mans_geometry_rule;
int mans_set_geometry_rules(const mans_geometry_rule* rules, size_t n);
/* Checkpoints one lineage may pass before a lineage decode gives up on lineages and runs
 * the chunked pipeline (0 = the default, 2048; merges take a few dozen).  Process-wide; for
 * tests and tuning.  No reference counterpart. */
int mans_set_lineage_cap(uint32_t checkpoints);
/* The lineage spacing the rules give a single stream of this shape (0 = the chunked
 * pipeline); mans_decode also needs automatic geometry and no sync points in or out for a
 * lineage decode. */
int mans_lineage_spacing(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t* D);
/* The lineage control words of the last lineage decode enqueued on this workspace
 * (synchronises stream): out4[0] 1 when the chunked pipeline ran instead, out4[1] its
 * reason bits (1 a slot list full, 2 the lineage cap, 4 maps unresolved, 8 a missing link,
 * 16 the table), out4[2] lineages run, out4[3] checkpoints they decoded.  No reference
 * counterpart. */
int mans_lineage_stats(void* workspace, void* stream, uint32_t* out4);
/* Polls of the sync kernel's wait for the previous range's last exit (a range of 256
 * chunks starts its first merge walk there).  A wait that runs out leaves that chunk to
 * the fixup rounds (the result is the same; status word 7 counts such waits).  Applies
 * to decodes enqueued (or graphs captured) afterwards.  No counterpart in the reference
 * (its phase-2 kernel re-launches from the host instead of waiting). */
#define MANS_WAIT_POLLS_DEFAULT (1u << 20)
int mans_set_wait_polls(uint32_t polls);
/* Fixup rounds per stream (decoder k_fixup, encoder k_enc_fix) before a stream that is
 * still inconsistent is reported (decoder: mans_decode_status MANS_ECORRUPT; encoder:
 * result status 8) instead of looping on.  0 (the default) = n_chunks + 1, a bound the
 * rounds never reach (each round proves its leftmost listed chunk).  A smaller limit
 * is for tests that force the report.  Applies to work enqueued afterwards.  No
 * counterpart in the reference (its host loop re-launches phase 2 until every flag is
 * set, src/multians_gpu_decoder.cu:617-656, with no bound). */
int mans_set_sync_round_limit(uint32_t rounds);
/* The chunk length the decoder picks for a stream (rules, else built-in). */
int mans_chunk_units(uint64_t n_units, uint64_t n_out, uint32_t L, uint32_t* chunk_units);
/* Sync points of a stream of n_units units with this chunk length (a power of two in
 * [16, 2^24]): ceil(n_units / chunk_units). */
int mans_sync_point_count(uint64_t n_units, uint32_t chunk_units, uint64_t* count);
/* Checks `count` sync points (host copy) of a stream with L states: record 0 at symbol
 * 0, states below L, bits at most 31 (32 in record 0), symbols_before non-decreasing.
 * Decodes never need it (index_in records are masked into range and proven); a caller
 * that takes sync points from storage runs it once before mans_stream_range. */
int mans_sync_points_check(const mans_sync_point* index_host, uint64_t count, uint32_t L);
/* Partial decode.  For the output symbols [first, last) (decode order) of a stream
 * (`full`: units, n_units, n_out, table, L, first_state, first_bit, chunk_units) with
 * sync points `index_host` (host copy, mans_sync_point_count records) and the same
 * records on the device at `index_dev`: fills *sub with the sub-stream that starts at
 * the sync point of the chunk holding symbol `first` and ends at the chunk holding
 * symbol last - 1 (its units, n_units with one unit of look-ahead, n_out, first_state,
 * first_bit, chunk_units, index_in = index_dev + that chunk; out and index_out are
 * NULL: the caller sets out, sub->n_out bytes).  Decoding *sub yields the window at
 * out + *skip.  The sub-stream trusts the record of its first chunk (the ones after it
 * are proven as usual).  Host only; no reference counterpart. */
int mans_stream_range(const mans_stream_desc* full, const mans_sync_point* index_host,
                      const mans_sync_point* index_dev, uint64_t first, uint64_t last,
                      mans_stream_desc* sub, uint64_t* skip);

/* --------------------------------------------------------- host-side codec */
/* Symbols s = min(255, floor(X)), X ~ Exp(lambda); element i depends only on (seed, i). */
int mans_gen_symbols(uint64_t seed, double lambda, uint8_t* out, uint64_t n, int threads);
/* The same symbols, written on the device (out: device, n bytes), enqueued on `stream`.
 * For large test and benchmark streams (generation, tables, encoding and decoding all on
 * the device).  No allocation. */
int mans_gen_symbols_device(uint64_t seed, double lambda, uint8_t* out, uint64_t n, void* stream);
/* hist[256] = occurrences of each byte value in in[0..n) (threads > 1: parallel).
 * Replaces the counting step of ANSTableGenerator::generate_distribution_from_buffer
 * (src/ans_table_generator.cc:61-95). */
int mans_histogram(const uint8_t* in, uint64_t n, uint64_t* hist, int threads);
/* Histogram -> counts summing to L (absent -> 0, present -> >= 1). */
int mans_normalize_counts(const uint64_t* hist, uint32_t nsym, uint32_t L, uint32_t* counts);
/* Encoder table (3 u32 per (symbol, state): next_state, code, length; symbol-major)
 * and decoder table (L u32, reference layout) from counts summing to L. */
int mans_build_tables(const uint32_t* counts, uint32_t nsym, uint32_t L, uint32_t* enc,
                      uint32_t* dec);
uint64_t mans_max_encoded_units(uint64_t n);
/* Encode in[0..n) into units (decoder order); returns the unit count in *n_units. */
int mans_encode(const uint8_t* in, uint64_t n, const uint32_t* enc, uint32_t L,
                uint32_t* units, uint64_t* n_units, uint32_t* first_bit,
                uint32_t* first_state);
/* One call: mans_gen_symbols (1 thread) -> histogram -> mans_normalize_counts ->
 * mans_build_tables -> mans_encode.  symbols_out may be NULL; units_cap must be at
 * least mans_max_encoded_units(n); dec_out receives the L-entry decoder table. */
int mans_make_stream(uint64_t seed, double lambda, uint64_t n, uint32_t L, uint8_t* symbols_out,
                     uint32_t* units_out, uint64_t units_cap, uint64_t* n_units,
                     uint32_t* first_bit, uint32_t* first_state, uint32_t* dec_out);

/* ------------------------------------------------ decoder tables on the device */
/* Builds decoder tables where the symbols already are: histogram of each job's symbols
 * -> counts normalised to L -> spread -> decoder table (L u32, CUHDCodetableItem
 * layout), the same bytes as mans_histogram + mans_normalize_counts +
 * mans_build_tables.  Replaces the host-side ANSTableGenerator path
 * (src/ans_table_generator.cc:61-95, :164-288); the reference has no device form.
 * Enqueued on `stream`; nothing is allocated.  A job that cannot be normalised (no
 * symbols, more distinct symbols than states) leaves its table unwritten and a nonzero
 * code for mans_table_status. */
typedef struct mans_table_job {
    const void* symbols;    /* device, u8[n] */
    uint64_t n;             /* 1 <= n < 2^50 */
    uint32_t L;             /* power of two in [2, 16384] */
    uint32_t reserved;
    void* table;            /* device, u32[L]: decoder table */
    void* counts;           /* device, u32[256] normalised counts, or NULL */
} mans_table_job;
int mans_table_workspace_bytes(size_t n_jobs, uint64_t* bytes);
int mans_build_tables_device(const mans_table_job* jobs, size_t n_jobs, void* workspace,
                             uint64_t workspace_bytes, void* stream);
/* status[j] = 0 (built), 1 (no symbols), 2 (more distinct symbols than L), 3 (failed);
 * waits for the stream. */
int mans_table_status(const void* workspace, size_t n_jobs, void* stream, uint32_t* status);

/* ------------------------------------------------------- encoder on the device */
/* Encodes symbols already on the device into the decoder's stream format, bit-exact
 * with mans_encode (units, first_bit, first_state), which restates ANSEncoder::encode
 * (src/ans_encoder.cc:13-97) and the unit reversal of CUHDInputBuffer
 * (src/multians_input_buffer.cc:29); the reference encodes only on the host.  Each job
 * gives its decoder table (L u32, CUHDCodetableItem layout, as built by
 * mans_build_tables / mans_build_tables_device / ANSTableGenerator): the encoder
 * derives its own tables from it.  Streams are cut into 4,096-symbol chunks encoded in
 * parallel from a guessed state, proven by merge walks and fixup rounds, then written
 * at their exact bit offsets.  Enqueued on `stream`; the plans are copied in first
 * (synchronously).  No allocation. */
typedef struct mans_encode_job {
    const uint8_t* symbols;     /* device, n bytes (when 16-byte aligned, the walks may also
                                   load the rest of the aligned 16-byte block holding the
                                   last symbol; those bytes are never used) */
    uint64_t n;
    const uint32_t* table;      /* device, L entries (decoder table) */
    uint32_t L;                 /* power of two in [2, 16384] */
    uint32_t reserved;
    uint32_t* units;            /* device, units_capacity u32 (>= mans_max_encoded_units(n)) */
    uint64_t units_capacity;
    /* Optional (device, mans_sync_point_count(units_capacity, 32) records): the decoder's
     * item records (one per 32 units, as mans_decode_args.items_out), written by the
     * encoder; with them the new stream decodes by MANS_F_TRUST_SYNC_POINTS at once.
     * Records of items that hold no step are {0, 0, n}. */
    mans_sync_point* items_out;
} mans_encode_job;

typedef struct mans_encode_result {
    uint64_t n_units;
    uint32_t first_bit;
    uint32_t first_state;
    uint32_t status;            /* 0 ok; bits: 1 bad table, 2 symbol without a state,
                                   4 capacity, 8 fixup rounds did not converge */
    uint32_t reserved;
} mans_encode_result;

int mans_encode_workspace_bytes(const mans_encode_job* jobs, size_t n_jobs, uint64_t* bytes);
int mans_encode_device(const mans_encode_job* jobs, size_t n_jobs, void* workspace,
                       uint64_t workspace_bytes, void* stream);
/* Results of the last mans_encode_device with these jobs and workspace (synchronises
 * `stream`). */
int mans_encode_results(const void* workspace, const mans_encode_job* jobs, size_t n_jobs,
                        void* stream, mans_encode_result* results);

/* ------------------------------------------------- two-state interleaved streams
 * The reference decodes one tANS state per stream and names multiple states and
 * multiple tables as future work (README.md:11); this format is the library's own.
 * Two states (slots) share one bit stream: input symbol i is encoded by slot i % 2 with
 * that slot's table (the two tables may be the same); both slots start at x = L; the
 * stream carries both final states.  Units are in decoder order, as for single streams;
 * decode step j yields input symbol n-1-j from slot (n-1-j) % 2.
 * Records, written by the encoder in its one forward pass, cut the stream into intervals
 * of record_units units anchored at the stream's end: record m (m = 0 ..
 * mans_il_record_count - 1) sits at unit boundary n_units - record_units * m and holds,
 * for the first step whose start bit is at or after it, both slots' states (x - L), the
 * bit offset from the boundary and the number of steps before it.  Record 0 is the
 * end. */
typedef struct mans_il_record {
    uint16_t state[2];          /* x - L of slot 0 and slot 1 */
    uint32_t bit;               /* start bit of the step, from the boundary's bit */
    uint64_t before;            /* steps before it (output offset, decode order) */
} mans_il_record;

#define MANS_IL_MIN_RECORD_UNITS 16u
#define MANS_IL_MAX_RECORD_UNITS (1u << 24)
#define MANS_IL_WORKSPACE_BYTES 256u

int mans_il_record_count(uint64_t n_units, uint32_t record_units, uint64_t* count);
/* Host encoder.  enc0 / enc1: encoder tables of mans_build_tables for slot 0 and slot 1
 * (may be the same).  units: mans_max_encoded_units(n) words; records:
 * mans_il_record_count(mans_max_encoded_units(n), record_units) entries suffice
 * (records_cap says how many there are); *n_records gets the count written. */
int mans_encode_interleaved(const uint8_t* in, uint64_t n, const uint32_t* enc0,
                            const uint32_t* enc1, uint32_t L, uint32_t record_units,
                            uint32_t* units, uint64_t* n_units, uint32_t* first_bit,
                            uint32_t* first_states, mans_il_record* records,
                            uint64_t records_cap, uint64_t* n_records);

typedef struct mans_il_args {
    const uint32_t* units;      /* device, 16-byte aligned, readable to n_units rounded up to 4 */
    uint64_t n_units;
    uint8_t* out;               /* device, n_symbols bytes (decode order) */
    uint64_t n_symbols;
    const uint32_t* tables[2];  /* device, L reference CUHDCodetableItem words per slot */
    uint32_t L;
    uint32_t first_states[2];
    uint32_t first_bit;
    uint32_t record_units;      /* in [MANS_IL_MIN_RECORD_UNITS, MANS_IL_MAX_RECORD_UNITS] */
    const mans_il_record* records;   /* device, mans_il_record_count(n_units, record_units) */
    uint32_t flags;             /* MANS_F_SYNC_ERRORS: synchronise and report */
    void* workspace;            /* device, >= MANS_IL_WORKSPACE_BYTES (status word) */
    void* stream;               /* hipStream_t or NULL */
} mans_il_args;

/* Decodes every record interval in parallel and checks that each ends exactly on the next
 * record (both states and the bit position) and the last on the stream's end (both
 * states L, bit 32 n_units).  A mismatch, or records out of order, sets the status to
 * MANS_ECORRUPT and a rejected table to MANS_ETABLE (mans_decode_status on `workspace`).
 * Output bytes stay inside [0, n_symbols) whatever the records hold. */
int mans_decode_interleaved(const mans_il_args* args);

#ifdef __cplusplus
}
#endif

#endif /* MANS_C_API_H_ */
