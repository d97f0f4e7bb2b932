/*
 * dtwax.h -- C ABI of the MI355X (gfx950) sDTW read classifier.
 *
 * Plain C types and pointers only.  Every entry point returns a dtwax_status
 * (0 = success, negative = error; dtwax_last_error() gives the message).
 * Device pointers are addresses returned by dtwax_malloc (or any HIP device
 * allocation in the same process); `stream` is a hipStream_t passed as
 * void* (NULL = the default stream).  Work is asynchronous on `stream`
 * unless a function says otherwise.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * DTWax checkout):
 *   dtwax_distances        FullDTW::distances (-DFP16)   src/include/DTW.hpp:285-295
 *   dtwax_distances_f32    FullDTW::distances (FP32)     src/include/DTW.hpp:285-295,
 *                                                         src/include/common.hpp:47-57
 *   dtwax_score            the batch loop + distances     src/main.cu:211-274
 *   dtwax_normalize        normalizer::normalize          src/include/normalizer.cu:71-114
 *   dtwax_reference_*      load_reference + tile upload   src/include/load_reference.hpp:33-108,
 *                                                         src/main.cu:78-121
 *   dtwax_squiggle_build   load_reference::load_ref_coeffs src/include/load_reference.hpp:33-108
 *   dtwax_ru_*             the Read Until loop            src/ru_main.cu:172-326
 *   dtwax_vbz_decode       the VBZ HDF5 plugin's decode   fast5/nanopolish_fast5_io.cpp:162-176
 *                                                         (vbz_compression v1.0.1,
 *                                                         install_hdf5_vbz.sh:43-58)
 *   dtwax_pod5_*           no counterpart: POD5 input beside the FAST5 reader
 *                                                         (fast5/nanopolish_fast5_io.cpp:72-194)
 *   dtwax_plan_packing     no counterpart: reports this library's packed plan layout
 */
#ifndef DTWAX_H
#define DTWAX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  DTWAX_OK = 0,
  DTWAX_EINVAL = -1,   /* bad argument */
  DTWAX_ENOMEM = -2,   /* device or host allocation failed */
  DTWAX_EDEVICE = -3,  /* HIP runtime error */
  DTWAX_ENOTREADY = -4,/* call order error (e.g. no reference uploaded) */
  DTWAX_EIO = -5       /* file could not be read or parsed */
} dtwax_status;

/* Arithmetic of the DP cell D = (s - q)^2 + min(top, diag). */
typedef enum {
  /* binary16 per op, round-to-nearest-even: t = s-q, t2 = t*t, D = t2+min.
   * The reference's half2 arithmetic (DTW.cu:37-39, common.hpp:32-45). */
  DTWAX_F16 = 0,
  /* binary16 with one rounding of t*t + min (fused multiply-add). */
  DTWAX_F16_FUSED = 1,
  /* binary16 per op, with a per-row drift taken off at every lane boundary
   * (the lane's first row costs rn(t*t - drift), one rounding) and the exact
   * row minimum taken off between bands; the offsets are added back in
   * binary32.  The row minimum is taken over a work item's own
   * columns, so with the reference in chunks (ref_chunk, or automatic
   * chunking of small batches and long references) the subtracted constants
   * and the binary16 rounding depend on the chunking: results are exact for
   * the chunks as computed, and agree with a whole-strand run to within
   * rounding, not bit for bit. */
  DTWAX_F16_RENORM = 2,
  /* binary32 per op (the reference's FP32 build, common.hpp:47-57). */
  DTWAX_F32 = 3
} dtwax_precision;

/* Query normalisation. */
typedef enum {
  DTWAX_NORM_NONE = 0,     /* copy */
  DTWAX_NORM_MEANSTD = 1,  /* (x - mean)/sqrt(var_biased + eps) -- normalizer.cu:86-100 */
  DTWAX_NORM_MEDMAD = 2    /* (x - median)/(1.4826 * MAD) */
} dtwax_norm_mode;

typedef struct dtwax_ctx dtwax_ctx;

/* Per read and strand (0 = fwd, 1 = rev) result. */
typedef struct {
  float score[2];    /* min over the last DP row (strand 1 unused for 1-strand references) */
  int32_t argmin[2]; /* reference column of that min, lowest on ties; -1 if none */
} dtwax_result;

typedef struct {
  int32_t precision;     /* dtwax_precision */
  int32_t segment;       /* query rows per lane (8, 12, 16, 20, 25, 28 or 32) for every read; 0 = automatic:
                            per read for DTWAX_F16 and DTWAX_F16_FUSED (the width with the
                            least padded work, one persistent launch per width used), per
                            read for DTWAX_F16_RENORM (8 up to 1536 rows; longer reads their
                            first 513..1536 rows at 8, the rest at 16: a short first band
                            keeps its precision), one width per batch for DTWAX_F32 */
  int32_t ref_chunk;     /* reference columns per work item; 0 = automatic (whole strand when
                            it fits); chunks overlap by (query length - 1) columns, exact for
                            the no-deletion recurrence */
  float renorm_drift;    /* DTWAX_F16_RENORM: initial per-row drift taken off at lane
                            boundaries (adapted band to band); <= 0 = default 0.01 */
  int32_t compat_quirks; /* must be 0: the reference kernel's seam deviations exist only
                            in its fixed geometry, reached through dtwax_distances */
  int32_t moves;         /* 2 (or 0): D = cost + min(top, diag), the shipped -DNO_REF_DEL
                            build; 3: + the left move D[r][c-1] (reference deletion, the
                            build without -DNO_REF_DEL, DTW.cu:23-27).  3 needs DTWAX_F16
                            or DTWAX_F32 and whole-strand work items (ref_chunk 0 or at
                            least the reference length): a
                            left move reaches back any number of columns, so column
                            chunks with an overlap are no longer exact */
  int32_t waves;         /* resident waves per launch (per width group): 0 = automatic (the
                            kernel's occupancy x compute units, at most one per work
                            item); a smaller value caps it.  Below full occupancy the
                            hardware may place the waves unevenly over the SIMDs (a
                            4096-read batch at 2 waves per SIMD ran 1.3% faster alone
                            and 26% slower right after another kernel), so automatic
                            is the safe choice */
  int32_t packing;       /* DTWAX_F16 / DTWAX_F16_FUSED, two moves: 1 = pack every read whose
                            first band spans at least 4 lanes, 0 = automatic (with segment 0:
                            the reads that gain), -1 = never.  A packed read's rows start
                            with a partial band (its length modulo 64 x S rows, padded to a
                            lane) that shares one band with other reads' partial bands,
                            followed by its full bands.  S: the segment when given (16 to
                            32), else 32 with packing = 1, or the width of 16, 20, 25, 28
                            and 32 whose packing costs least.  Scores are identical to the
                            unpacked layout's */
} dtwax_score_opts;

/* ---- context ----------------------------------------------------------- */
const char* dtwax_last_error(void);
const char* dtwax_version(void);
/* Create a context on HIP device `device`. */
int dtwax_create(int device, dtwax_ctx** out);
void dtwax_destroy(dtwax_ctx* ctx);
/* Device properties used for the roofline: compute units, peak engine clock. */
int dtwax_device_info(dtwax_ctx* ctx, int32_t* n_cu, int32_t* clock_khz, char* name,
                      int32_t name_len);

/* ---- device memory / streams / events (thin HIP wrappers for callers
 *      without a HIP binding of their own) ------------------------------- */
int dtwax_malloc(void** dptr, size_t bytes);
int dtwax_free(void* dptr);
int dtwax_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int dtwax_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int dtwax_stream_create(void** stream);
int dtwax_stream_destroy(void* stream);
int dtwax_stream_sync(void* stream);
int dtwax_device_sync(void);
int dtwax_event_create(void** ev);
int dtwax_event_destroy(void* ev);
int dtwax_event_record(void* ev, void* stream);
int dtwax_event_elapsed_ms(void* start, void* stop, float* ms);

/* ---- reference ---------------------------------------------------------
 * Upload z-scored reference levels: `levels` holds n_strands rows of n_cols
 * floats (host memory; strand 0 = fwd, 1 = rev).  Replaces any previous
 * reference.  Synchronous. */
int dtwax_reference_upload(dtwax_ctx* ctx, const float* levels, int32_t n_strands,
                           int64_t n_cols);
/* Build the z-scored squiggle of both strands from a sequence and a k-mer
 * model (load_reference.hpp:33-108 arithmetic: float sequential sums over the
 * first ref_len k-mers, sd = sqrt(E[x^2]-mu^2), (level-mu)/sd in float).
 * seq: ACGT text (length seq_len); model_means: 4^k levels indexed by the
 * 2-bit k-mer code (A=0, C=1, G=2, T=3, first base most significant);
 * ref_len: k-mers used (<= seq_len-k+1 for all, or the reference's fixed
 * REF_LEN); out: 2*ref_len floats (fwd then reverse complement).  Host-side. */
int dtwax_squiggle_build(const char* seq, int64_t seq_len, int32_t k,
                         const float* model_means, int64_t ref_len, float* out,
                         int64_t* n_skipped);

/* Decode one chunk of a FAST5 Raw/Signal dataset compressed with ONT's VBZ
 * filter (HDF5 filter 32020), for files read without the VBZ plugin.
 * cd_values: the filter's stored values (version, integer_size,
 * delta_zigzag, zstd_level; missing ones default to 0, 2, 1, 1).  The chunk
 * is [uint32 LE decoded byte count][zstd frame (zstd_level > 0) of a
 * stream-VByte block of the zig-zag sample differences]; version 0 blocks
 * have 2-bit keys for 1/2/3/4 bytes, version 1 blocks 2-bit keys for 0/1/2/4
 * bytes or 1-bit keys for 1/2 bytes, recognised from the block (a block
 * that fits none, or both, is an error).  integer_size must be 2.  Writes
 * n = decoded bytes / 2 int16 samples (n <= dst_cap) to dst and sets
 * *n_out.  zstd is loaded at run time (libzstd.so.1 or DTWAX_ZSTD_LIB).
 * Host-side. */
int dtwax_vbz_decode(const void* src, int64_t src_size, const uint32_t* cd_values,
                     int32_t n_cd_values, int16_t* dst, int64_t dst_cap, int64_t* n_out);

/* ---- POD5 input ---------------------------------------------------------
 * No counterpart in the reference, which reads FAST5 only
 * (fast5/nanopolish_fast5_io.cpp:72-194): ONT's POD5 container (a signature,
 * Arrow IPC files for the Reads, Signal and Run Info tables, a flatbuffer
 * footer) read in place from a memory map, without the pod5 or Arrow
 * libraries.  Signal rows are int16 lists or VBZ blobs (zstd frame of a
 * 1-bit-key stream-VByte block of zig-zag differences, sample count in the
 * row's `samples` column); a read's signal is its rows in list order.
 * Arrow body compression: none or ZSTD.  Host-side; a handle is read-only
 * after open and may be shared by threads. */
typedef struct dtwax_pod5 dtwax_pod5;
typedef struct {
  int64_t n_samples;          /* samples over all the read's rows */
  int64_t start_sample;       /* `start` column (0 when absent) */
  int32_t channel;            /* -1 when absent */
  int32_t read_number;        /* -1 when absent */
  int32_t n_rows;             /* signal-table rows holding the read */
  int32_t well;               /* -1 when absent */
  float calibration_offset;   /* pA = (raw + offset) * scale; NaN when absent */
  float calibration_scale;
  char read_id[40];           /* canonical lower-case UUID text, NUL-terminated */
} dtwax_pod5_read_info;
/* Map and index a file: both tables' schemas and batches are checked, every
 * read's row list is checked against the signal table, and a `num_samples`
 * column must agree with the rows' `samples`.  DTWAX_EIO on any mismatch. */
int dtwax_pod5_open(const char* path, dtwax_pod5** out);
int dtwax_pod5_close(dtwax_pod5* f);
int dtwax_pod5_count(const dtwax_pod5* f, int64_t* n_reads, int64_t* n_signal_rows);
/* Metadata of reads [first, first + n) in table order. */
int dtwax_pod5_reads(const dtwax_pod5* f, int64_t first, int64_t n, dtwax_pod5_read_info* out);
/* Samples [skip, skip + take) of the n reads listed in `reads` (table
 * indices), decoded on n_threads host threads (0: up to 16).  take >= 0:
 * read j goes to dst + j*take; take < 0: whole reads from `skip` on, packed
 * in list order.  n_out[j] = samples written for read j (fewer than take when
 * the read is shorter).  Rows wholly outside the window are not decoded. */
int dtwax_pod5_signal(const dtwax_pod5* f, const int64_t* reads, int64_t n, int64_t skip,
                      int64_t take, int16_t* dst, int64_t dst_cap, int64_t* n_out,
                      int32_t n_threads);

/* ---- normalisation -----------------------------------------------------
 * Normalise n_reads device-resident reads (float samples; read i starts at
 * d_offsets[i] and has d_lens[i] samples) into d_out (same layout), with
 * FP64 statistics on the device.  d_stats (optional, 2*n_reads doubles):
 * (centre, scale) per read. */
int dtwax_normalize(dtwax_ctx* ctx, const float* d_raw, const int64_t* d_offsets,
                    const int32_t* d_lens, int32_t n_reads, int32_t mode, double eps,
                    float* d_out, double* d_stats, void* stream);

/* ---- batch scoring -----------------------------------------------------
 * Score n_reads normalised device-resident queries (layout as above) against
 * the uploaded reference; d_out receives one dtwax_result per read.  opts may
 * be NULL (F16, automatic tiling).  Reads of any length >= 1 may be mixed in
 * one call; work is balanced by a persistent kernel over (read, chunk) items
 * ordered by decreasing cost.  Synchronous (copies the lengths to the host to
 * plan, runs, and waits); use a plan for asynchronous repeated scoring. */
int dtwax_score(dtwax_ctx* ctx, const float* d_queries, const int64_t* d_offsets,
                const int32_t* d_lens, int32_t n_reads, const dtwax_score_opts* opts,
                dtwax_result* d_out, void* stream);

/* Plans: the work decomposition for a fixed set of read lengths, built once
 * and reused (dtwax_score builds and frees one per call).  h_lens are host
 * lengths.  dtwax_score_plan is fully asynchronous on `stream`; calls that
 * share a plan must not overlap (the plan owns the work counter and keys). */
typedef struct dtwax_plan dtwax_plan;
int dtwax_plan_create(dtwax_ctx* ctx, const int32_t* h_lens, int32_t n_reads,
                      const dtwax_score_opts* opts, dtwax_plan** out);
void dtwax_plan_destroy(dtwax_plan* plan);
/* cells = sum over reads of len * counted columns * strands (the GCUPS
 * numerator); computed = cells actually evaluated (padding rows, chunk
 * overlap and the 63-step wavefront ramp included); segment = the width of
 * every read, or 0 if the plan mixes widths; grid = the largest launch. */
int dtwax_plan_info(const dtwax_plan* plan, int64_t* cells, int64_t* computed,
                    int32_t* n_items, int32_t* segment, int32_t* ref_chunk, int32_t* grid);
int dtwax_score_plan(dtwax_plan* plan, const float* d_queries, const int64_t* d_offsets,
                     const int32_t* d_lens, dtwax_result* d_out, void* stream);
/* Enable HIP-event timing of the sDTW kernel launches inside
 * dtwax_score_plan; dtwax_plan_kernel_ms then waits for and returns the last
 * call's sDTW time (all of its launches). */
/* Packed layout of a plan (no counterpart in the reference, whose batches
 * give every read whole bands): bands shared by several reads' partial
 * bands, and the reads placed in them (0, 0 when the plan packs nothing). */
int dtwax_plan_packing(const dtwax_plan* plan, int32_t* n_packs, int32_t* n_packed_reads);
int dtwax_plan_set_timing(dtwax_plan* plan, int32_t enable);
int dtwax_plan_kernel_ms(dtwax_plan* plan, float* ms);

/* ---- autotuner -----------------------------------------------------------
 * Picks segment and ref_chunk for a workload (no counterpart in the
 * reference, whose geometry is fixed at compile time, src/include/common.hpp:
 * 60-98) by timing the sDTW kernel on a sample: sample_reads reads (<= 0:
 * 512) spread over the sorted lengths, against the first sample_cols
 * reference columns (<= 0: 2^20), replicated until every candidate has at
 * least 6 work items per resident wave (steady state, as in the full run).
 * It times segment in {0 = per-read widths, 8, 12, 16, 20, 25, 28, 32} ({0, 8} for
 * DTWAX_F16_RENORM) at ref_chunk min(65536, sample_cols); the fastest per
 * counted cell wins, except that segment 0 (the plan's own per-read choice)
 * stays unless a fixed width is at least 2% faster.  ref_chunk is then the plan's own rule for the whole
 * workload at that width (whole strands, or the longest chunks whose band
 * scratch fits the budget: a quarter of free device memory, at most 32 GB),
 * not a sample timing, which would favour short chunks.  A nonzero
 * base->segment or base->ref_chunk fixes that one.  The choice goes into
 * *best (base's other fields kept).
 * d_queries/d_offsets: the workload's device queries (as for dtwax_score);
 * h_lens: host lengths.  The choice is cached per device, options, strand
 * count, reference-size and length-distribution key, in memory and, if
 * cache_path is not NULL, appended to that text file (read on later calls);
 * a cached call times nothing and leaves *n_table = 0.  table (optional)
 * receives up to table_cap timed candidates.  Synchronous. */
typedef struct {
  int32_t segment;
  int32_t ref_chunk;
  float ms;    /* kernel time on the sample */
  float gcups; /* counted cells / kernel time */
} dtwax_tune_entry;

int dtwax_autotune(dtwax_ctx* ctx, const float* d_queries, const int64_t* d_offsets,
                   const int32_t* h_lens, int32_t n_reads, const dtwax_score_opts* base,
                   int32_t sample_reads, int64_t sample_cols, const char* cache_path,
                   dtwax_score_opts* best, dtwax_tune_entry* table, int32_t table_cap,
                   int32_t* n_table);

/* Mirror of FullDTW::distances (src/include/DTW.hpp:285-295) in the
 * reference's shipped build (-DFP16 -DNO_REF_DEL; src/include/common.hpp:
 * 60-98: QUERY_LEN 2048, PREFIX_LEN 1024, SEGMENT 26, WARP 32, REF_LEN 48128,
 * of which 57 tiles x 832 = 47,424 columns are scored).  Same buffers and
 * layouts as the reference's launch (src/main.cu:92-121, 165-204):
 *   ref    REF_LEN half2 {fwd, rev} words in the tiled layout of
 *          src/main.cu:92-109 (h[l*832 + i*32 + j] = flat[l*832 + j*26 + i]);
 *   query  num_entries * 2048 half2 words {q_fwd, q_rev} (the reference
 *          passes {q, q}); its 32-word pad is accepted and never read;
 *   dists  num_entries half2 words {fwd, rev}: min over the last DP row;
 *   thresh accepted and unused, as in the reference (DTW.cu:152-153);
 *   device_last_row  NULL, or num_entries * REF_LEN half2 words: on return
 *          entry [e * REF_LEN + k] holds read e's last DP row (row 2047) at
 *          the scored column whose tiled index is k (tile t < 56 at t*832,
 *          tile 56 at REF_LEN - 832, as DTW.cu:293, 419, 546 store it);
 *          entries of no scored column are left as they were.  The
 *          reference also uses this buffer as its band hand-over scratch;
 *          here that is internal and the buffer is output only.
 * compat_quirks selects the DP: 1 = the reference kernel's own, including
 * its two seam deviations (a +inf diagonal at row 1024, columns 832k and
 * 832k + 26, DTW.cu:179-180, 315, 442; rows 962..992 of either 1024-row pass
 * taking D[r+30][c-1] as their diagonal at columns 832k, k >= 2,
 * DTW.cu:385-388), bit for bit, at the speed of the plain kernel; 0 = the
 * clean recurrence D = (s-q)^2 + min(top, diag) with the same per-op
 * binary16 rounding; 2 = the same output as 1 from a second, independent
 * kernel that follows the reference's tile-by-tile schedule (one wave per
 * read, about half the speed; a cross-check).
 * The call is asynchronous on `stream` and allocates only on first use (the
 * context's untiled-reference buffer, and per-read keys when num_entries
 * grows); it does not touch the reference of dtwax_reference_upload.  Calls
 * on one context share those buffers and must not overlap.  All pointers
 * are device pointers. */
int dtwax_distances(dtwax_ctx* ctx, const void* ref, const void* query, void* dists,
                    int32_t num_entries, uint32_t thresh, void* stream,
                    void* device_last_row, int32_t compat_quirks);

/* The same mirror for the reference's FP32 build (FullDTW::distances
 * instantiated with value_ht = float, common.hpp:47-57, 69-72): ref = REF_LEN
 * = 96256 floats, the concatenation fwd[0:48128) | rev[0:48128) in the tiled
 * layout; 115 tiles (95,680 columns) scored as ONE sequence (the DP crosses
 * the fwd -> rev seam); query = num_entries * 2048 floats (+ pad); dists =
 * num_entries floats; device_last_row = NULL or num_entries * REF_LEN floats
 * (tile 114 at REF_LEN - 832).  Binary32 per op.  compat_quirks as for
 * dtwax_distances: 1 (the reference kernel's bits) on the persistent kernel,
 * 2 the same bits from the tile-by-tile kernel, 0 the clean recurrence (the
 * deviations are the same as in the FP16 build: they come from the
 * schedule, not the type). */
int dtwax_distances_f32(dtwax_ctx* ctx, const void* ref, const void* query, void* dists,
                        int32_t num_entries, float thresh, void* stream,
                        void* device_last_row, int32_t compat_quirks);

/* Allocation-free variant of dtwax_distances (same arguments, layouts,
 * compat_quirks forms and results).  No counterpart in the reference: it is
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture or a tight serving loop.  It calls no hipMalloc or hipFree:
 * its scratch (the untiled reference, a work counter and 2 keys per read) is
 * one file-scope device array of 448 KB on each GPU the library runs on,
 * held until the process ends, whose address is taken when a context is
 * created on that GPU.  At most 16384 entries per call (DTWAX_EINVAL
 * beyond).  All calls on one GPU share that array, which keeps the last
 * call's contents (each call resets what it needs): calls that overlap on
 * one GPU, on any streams or contexts, race. */
int dtwax_distances_noalloc(dtwax_ctx* ctx, const void* ref, const void* query, void* dists,
                            int32_t num_entries, uint32_t thresh, void* stream,
                            void* device_last_row, int32_t compat_quirks);

/* ---- Read Until serving ---------------------------------------------------
 * Replaces the stateless loop of src/ru_main.cu:172-326 (10 FIFO queries of
 * 2048 samples scored from scratch per iteration, scores printed, no
 * decision).  A read arrives as fixed-size chunks of raw samples on a
 * channel.  Per undecided read the engine keeps the last DP row over the
 * whole reference (both strands) in HBM, so chunk k resumes the alignment
 * of chunks 0..k-1: the score after k chunks equals the score of the
 * concatenated query, normalised with the first chunk's statistics, scored
 * from scratch with the same precision.  After each chunk the running score
 * s = min(fwd, rev) decides: s >= threshold is NONTARGET at once (the min
 * over the last row never decreases as rows are added, so under a fixed
 * threshold the decision is final); a read still below it after max_chunks
 * chunks is TARGET.  The length-aware rule (threshold_per_row) compares s
 * with a threshold proportional to the rows scored so far instead.
 * Decision rule as calc_classification_threshold.py:49-57 (target iff
 * score < threshold). */
typedef struct dtwax_ru dtwax_ru;

typedef enum {
  DTWAX_RU_CONTINUE = 0,  /* undecided: send the next chunk */
  DTWAX_RU_NONTARGET = 1, /* score >= threshold: unblock the read */
  DTWAX_RU_TARGET = 2,    /* max_chunks chunks below the threshold: keep it */
  DTWAX_RU_NO_SLOT = 3,   /* new read found every state slot busy: not scored, and
                             later chunks of it are STALE */
  DTWAX_RU_STALE = 4      /* chunk of a read already decided: not scored; score and
                             chunks are the decided read's */
} dtwax_ru_action;

typedef struct {
  int32_t max_reads;     /* state slots = reads undecided at once (0: 512) */
  int32_t chunk_len;     /* samples per chunk, a multiple of 64 (0: 1600) */
  int32_t max_chunks;    /* chunks before a TARGET decision (0: 4) */
  int32_t precision;     /* DTWAX_F16, DTWAX_F16_RENORM (each chunk subtracts a per-row drift
                            at its lane hops and the stored row's minimum at lane 0; the
                            engine adds the amounts back in binary64), DTWAX_F32, or
                            DTWAX_F16_FUSED (one rounding of t*t + min: 3 ops per cell
                            instead of 4, so ~25% more chunks per second when every
                            channel is undecided, at the cost of scores that differ from
                            the reference's half2 arithmetic in the last bits) */
  int32_t segment;       /* rows per lane, dividing chunk_len / 64 (0: largest of 32, 28, 25, 20, 16, 12, 8) */
  int32_t norm;          /* dtwax_norm_mode, statistics of the read's first chunk */
  double eps;            /* DTWAX_NORM_MEANSTD epsilon */
  float threshold;       /* on min(fwd, rev); <= 0: no early NONTARGET */
  int32_t min_ref_chunk; /* shortest reference piece per work item, columns; 0: the engine
                            picks per batch between ~8 pieces per resident wave of at least
                            4 * chunk_len and one to three pieces per SIMD of at least
                            2 * chunk_len, whichever its model of the slowest SIMD favours */
  int64_t hbm_budget;    /* bytes of read state at most (0: max_reads slots) */
  float threshold_per_row; /* > 0: length-aware rule, the threshold after k chunks is
                              threshold_per_row * k * chunk_len (replaces threshold) */
  int32_t n_streams;     /* serving loop: batches in flight at once, each on its own HIP
                            stream (1..8; 0: 2) */
  float renorm_drift;    /* DTWAX_F16_RENORM: initial per-row drift taken off at lane boundaries,
                            adapted chunk to chunk (<= 0: 0.01) */
  int32_t reserved;
} dtwax_ru_opts;

typedef struct {
  int32_t channel; /* >= 0 */
  int32_t number;  /* read number on the channel: a new number ends the previous read */
  int64_t offset;  /* first sample of this chunk in the samples buffer */
} dtwax_ru_chunk;

typedef struct {
  int32_t channel;
  int32_t number;
  int32_t action;     /* dtwax_ru_action */
  int32_t chunks;     /* chunks of the read scored so far */
  float score[2];     /* running fwd, rev score (+inf if never scored) */
  int32_t argmin[2];  /* column of each */
  double latency_us;  /* serving loop only: submission to decision available */
} dtwax_ru_decision;

/* Allocates max_reads (or what hbm_budget allows) state slots of
 * 2 x (n_cols + 448) x 4 bytes (x n_strands for DTWAX_F32) plus staging, on
 * the context's current reference (which must not change while the engine
 * lives). */
int dtwax_ru_create(dtwax_ctx* ctx, const dtwax_ru_opts* opts, dtwax_ru** out);
int dtwax_ru_destroy(dtwax_ru* ru);
int dtwax_ru_info(dtwax_ru* ru, int32_t* slots, int32_t* active, int32_t* segment,
                  int64_t* state_bytes);
/* Batches, chunks scored, cells, and the sum of the sDTW kernels' durations. */
int dtwax_ru_stats(dtwax_ru* ru, int64_t* batches, int64_t* chunks, int64_t* cells,
                   double* gpu_ms);
/* Scores one batch: chunk i is chunk_len host samples at samples + offset.
 * At most one chunk per channel per call (EINVAL otherwise).  Writes one
 * decision per chunk into out[i].  Synchronous; calls are serialised. */
int dtwax_ru_process(dtwax_ru* ru, const dtwax_ru_chunk* chunks, int32_t n, const float* samples,
                     dtwax_ru_decision* out);
/* The read on `channel` ended: free its state. */
int dtwax_ru_release(dtwax_ru* ru, int32_t channel);
/* Serving loop (continuous batching): a worker thread launches, as one
 * batch on a free stream, every chunk that has arrived (the earliest per
 * channel, for channels with no chunk on the GPU), while earlier batches are
 * still running; up to n_streams batches run on the device at once.
 * submit copies chunk_len samples; poll returns up to cap decisions, waiting
 * up to timeout_ms (-1: until one is available).  While the loop runs,
 * dtwax_ru_process is refused and dtwax_ru_release is queued in order with
 * the channel's chunks. */
int dtwax_ru_start(dtwax_ru* ru);
int dtwax_ru_submit(dtwax_ru* ru, int32_t channel, int32_t number, const float* samples);
int dtwax_ru_poll(dtwax_ru* ru, dtwax_ru_decision* out, int32_t cap, int32_t* n_out,
                  int32_t timeout_ms);
int dtwax_ru_stop(dtwax_ru* ru);

/* ---- self-test of the lane-movement primitives the kernels rely on ------
 * Writes, for each of the 64 lanes, the value received by a DPP wave shift
 * right by one (out[0..63]) and left by one (out[64..127]) of the lane
 * index, with out-of-range lanes keeping -1.  Synchronous. */
int dtwax_selftest_lanes(int32_t* out_host);

#ifdef __cplusplus
}
#endif
#endif
