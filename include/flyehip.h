/*
 * flyehip.h - C ABI of the MI355X (gfx950) Flye overlap-detection hot path.
 *
 * One shared library, libflyehip.so, built from the HIP sources under cuflye_amd/csrc with
 * hipcc.  Every entry point takes plain pointers and sizes, returns an int
 * status (FLYEHIP_OK == 0) and never throws across the ABI.  On failure the
 * context keeps a message that flyehip_last_error() returns.  Nothing in the
 * library falls back to the CPU: an entry point that cannot run on the GPU
 * returns an error.
 *
 * Terms follow Flye:
 *   read / target record  - one sequence of the resident read store, numbered
 *                           1..N; its two strands have signed ids +n and -n
 *                           (FastaRecord::Id::signedId()).
 *   query record          - one strand of a sequence whose overlaps are asked
 *                           for (OverlapDetector::getSeqOverlaps' fastaRec).
 *   k-mer match           - Flye's KmerMatch {curPos, extPos, extId, kmerRepr}
 *                           (patches/flye/2.9.6/0001-cuflye-candidate-dump.patch:17-29).
 *   overlap row           - Flye's OverlapRange as dumped by overlap-range-v1
 *                           (docs/abi/overlap-range-v1.md:21-38) plus the chain
 *                           length the divergence was computed from.
 */
#ifndef FLYEHIP_H
#define FLYEHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLYEHIP_ABI_VERSION 4

/* Status codes. */
#define FLYEHIP_OK 0
#define FLYEHIP_ERR_ARG 1          /* invalid argument or unsupported shape   */
#define FLYEHIP_ERR_HIP 2          /* HIP runtime error (message has the name) */
#define FLYEHIP_ERR_BUDGET 3       /* device memory budget would be exceeded   */
#define FLYEHIP_ERR_CAPACITY 4     /* caller buffer too small; *needed is set  */
#define FLYEHIP_ERR_STATE 5        /* call out of order (e.g. no index yet)    */
#define FLYEHIP_ERR_UNSUPPORTED 6  /* Flye option the GPU path does not model  */

typedef struct flyehip_ctx flyehip_ctx;

/* Index parameters.  The reference never pins VertexIndex's thresholds
 * (Flye's src/sequence/vertex_index.* is not part of the reference); this
 * library defines them and the oracle restates the same definition:
 *   count      - occurrences of each canonical k-mer over the forward strand
 *                of every resident read;
 *   solid      - count >= solid_min (0 = first valley of the count histogram);
 *   repetitive - solid and count > rep_max (0 = floor(repeat_kmer_rate * mean
 *                count of solid k-mers)); repetitive k-mers are not indexed
 *                and their query windows go to the filtered-position list;
 *   bucket     - positions of one solid, non-repetitive canonical k-mer in
 *                (read number, position) order, oriented so that the
 *                canonical k-mer is what the stored strand spells there. */
typedef struct {
    int32_t kmer_size;          /* k, 1..31 */
    int32_t solid_min;          /* 0 = automatic */
    int32_t rep_max;            /* 0 = automatic */
    float repeat_kmer_rate;     /* used when rep_max == 0; 10.0 by default */
} flyehip_index_params_t;

/* Index facts after a build. */
typedef struct {
    int32_t kmer_size;
    int32_t solid_min;          /* threshold actually used */
    int32_t rep_max;            /* threshold actually used */
    int32_t minimizer_window;   /* 0: every position indexed */
    int64_t distinct_kmers;     /* distinct canonical k-mers counted */
    int64_t solid_kmers;        /* solid and not repetitive (a minimizer index
                                   buckets only those with a minimizer position) */
    int64_t repetitive_kmers;
    int64_t indexed_positions;  /* total bucket entries */
    int64_t hash_slots;         /* open-addressed table size */
    int64_t device_bytes;       /* device memory held by the index */
} flyehip_index_info_t;

/* getSeqOverlaps parameters (patches/flye/2.9.6/0008-cuflye-overlap-replay-fixture-dump.patch:151-171). */
typedef struct {
    int32_t kmer_size;
    int32_t minimum_overlap;
    int32_t maximum_jump;            /* maxJump, 1500 */
    int32_t max_jump_gap;            /* MAX_GAP, 500 */
    int32_t gap_jump_threshold;      /* GAP_JUMP_THLD, 100 */
    float large_gap_penalty;         /* LG_GAP, 2.0 */
    float small_gap_penalty;         /* SM_GAP, 0.5 */
    float min_kmer_survival_rate;    /* 0.01 */
    int32_t maximum_overhang;
    int32_t check_overhang;
    int32_t force_local;
    int32_t only_max_ext;            /* 1: best proposal per target; 0: containment filter */
    int32_t max_overlaps;            /* 0 = unlimited */
    float max_divergence;
    float sample_rate;               /* VertexIndex::getSampleRate(), 1.0 */
    int32_t keep_alignment;          /* must be 0 */
    int32_t nucl_alignment;          /* must be 0 */
    int32_t partition_bad_mappings;  /* must be 0 */
} flyehip_overlap_params_t;

/* One k-mer match as candidate-record-v1 writes it
 * (docs/abi/candidate-record-v1.md:31-52). */
typedef struct {
    int64_t query_id;    /* signed */
    int64_t target_id;   /* signed; strand '+' iff > 0 */
    uint64_t kmer;       /* Kmer::numRepr() of the query window */
    int32_t query_pos;
    int32_t target_pos;
} flyehip_candidate_t;   /* 32 bytes */

/* One overlap row (overlap-range-v1 fields + chain length). */
typedef struct {
    int64_t cur_id;
    int64_t ext_id;
    int32_t cur_begin;
    int32_t cur_end;
    int32_t cur_len;
    int32_t ext_begin;
    int32_t ext_end;
    int32_t ext_len;
    int32_t score;
    float seq_divergence;
    int32_t chain_length;
    int32_t reserved;
} flyehip_overlap_t;     /* 56 bytes */

/* One edge alignment handed to chainReadAlignments
 * (cuda/cuflye_cuda_read_alignment_chain_replay.cu:37-53;
 *  tools/replay_read_alignment_chains.py:15-29). */
typedef struct {
    int64_t read_id;
    int64_t edge_id;
    int64_t edge_seq_id;
    int32_t read_begin;
    int32_t read_end;
    int32_t read_len;
    int32_t edge_left_node;
    int32_t edge_right_node;
    int32_t edge_begin;
    int32_t edge_end;
    int32_t edge_len;
    int32_t score;
    float seq_divergence;
} flyehip_edge_alignment_t;  /* 64 bytes */

/* ReadAligner chaining parameters (tools/replay_read_alignment_chains.py:64-68). */
typedef struct {
    int32_t maximum_jump;
    int32_t max_read_overlap;
    int32_t minimum_overlap;
    int32_t max_separation;
} flyehip_chain_params_t;

/* One accepted chain segment (read-alignment-v1 row, docs/abi/read-alignment-v1.md:27-47);
 * `input_index` points back into the edge-alignment array of its read. */
typedef struct {
    int32_t read_slot;     /* which read of the batch */
    int32_t chain_id;      /* ordinal of the accepted chain within its read */
    int32_t segment_id;
    int32_t input_index;
    int32_t chain_score;
    int32_t reserved;
} flyehip_chain_segment_t;   /* 24 bytes */

/* Kernels timed in flyehip_stats_t (one HIP kernel each unless noted). */
#define FLYEHIP_K_JOIN_COUNT 0   /* joinCountKernel: windows per query              */
#define FLYEHIP_K_JOIN_FILL 1    /* joinFillKernel: k-mer matches in emission order */
#define FLYEHIP_K_RADIX 2        /* stable (extId, curPos) order of the matches:
                                    matchCountSortKernel (the library radix sort
                                    above its target-id limit)                      */
#define FLYEHIP_K_RANK_KEYS 3    /* tieFlagsKernel + rankKeysKernel                 */
#define FLYEHIP_K_LEAF 4         /* segLeafKernel: exact leaf sorts in LDS          */
#define FLYEHIP_K_HEAP 5         /* segHeapLdsKernel / segHeapGlobalKernel          */
#define FLYEHIP_K_PREFILTER 6    /* chainPrefilterKernel (+ chainRelevanceKernel
                                    where the tie replay runs)                      */
#define FLYEHIP_K_DP 7           /* chainDpKernel<G>                                */
#define FLYEHIP_K_WALK 8         /* chainWalkKernel<size class, onlyMaxExt>         */
#define FLYEHIP_K_ROUND_COUNT 9  /* segCountKernel: partition flags per chunk       */
#define FLYEHIP_K_ROUND_RANK 10  /* segRankKernel: swap partners                    */
#define FLYEHIP_K_ROUND_SWAP 11  /* segSwapKernel                                   */
#define FLYEHIP_K_ROUND_OTHER 12 /* segMedian/Map/SplitKernel and the chunk scans   */
#define FLYEHIP_K_GATHER 13      /* gatherHitsKernel: group-order hits              */
#define FLYEHIP_K_SCORE_KEYS 14  /* chainScoreKeysKernel + chainScorePushKernel     */
#define FLYEHIP_K_CLAIM 15       /* chainClaimKernel<onlyMaxExt>: backtrack .. selection
                                    without the score sort                          */
#define FLYEHIP_K_COUNT 16

/* Timing and counters of the last call (milliseconds on the library stream). */
typedef struct {
    double ms_total;
    double ms_join;          /* window count + fill */
    double ms_sort;          /* k-mer match sort */
    double ms_group;         /* group split, prefilters, ext re-sort */
    double ms_chain;         /* chaining DP */
    double ms_extract;       /* score order, backtrack, filters, compaction */
    int64_t queries;
    int64_t windows;
    int64_t matches;
    int64_t groups;
    int64_t overlaps;
    int64_t batches;
    int64_t device_bytes_peak;
    int64_t tie_queries;     /* queries whose match sort needed tie emulation */
    /* Kernels (FLYEHIP_K_*): GPU time summed over the launches (HIP events
     * on the stream each launch runs on; launches on different streams
     * overlap, so the sum can exceed ms_total), launch count, and algorithmic
     * bytes (the per-unit figures documented with the roofline times the
     * units the launches processed). */
    double kernel_ms[FLYEHIP_K_COUNT];
    int64_t kernel_launches[FLYEHIP_K_COUNT];
    double kernel_bytes[FLYEHIP_K_COUNT];
    int64_t dp_cells;        /* (row, predecessor) pairs the chaining DP evaluated */
    int64_t exact_walk_groups; /* passing groups whose walk needed the exact score sort */
} flyehip_stats_t;

/* ---- context ---------------------------------------------------------- */

/* Opens a context on `device` (CUFLYE_CUDA_DEVICE in patch 0003 file lines
 * 37-47) with a device-memory budget in bytes (0 = 90% of free memory;
 * CUFLYE_CUDA_MEMORY_BUDGET_BYTES). */
int flyehip_open(int device, uint64_t memory_budget_bytes, flyehip_ctx** out_ctx);
int flyehip_close(flyehip_ctx* ctx);
const char* flyehip_last_error(const flyehip_ctx* ctx);
int flyehip_abi_version(void);
int flyehip_get_stats(const flyehip_ctx* ctx, flyehip_stats_t* out);

/* Tuning knobs picked by the autotuner: lanes that chain one target group
 * (1, 4, 8, 16, 32 or 64; 8 by default), the query batch size (2048 by
 * default), and the batches in flight (1 .. 4, 2 by default):
 * flyehip_overlaps* deal batches to that many streams, each driven by its
 * own host thread, so one batch's tail overlaps the next batch (the
 * per-query path always uses one).  0 keeps the current value.  A new batch
 * size or number of batches in flight releases the per-batch device buffers
 * (each batch in flight holds its own; the next call sizes them again). */
int flyehip_set_tuning(flyehip_ctx* ctx, int32_t dp_lanes_per_group,
                       int32_t queries_per_batch, int32_t batches_in_flight);

/* ---- resident read store and k-mer index (VertexIndex) ---------------- */

/* Uploads reads 1..n_reads: bases are ASCII A/C/G/T (other bytes are an
 * error), read i spans bases[offsets[i] .. offsets[i+1]).  The bytes are
 * copied to the device as they are and packed there; the store is kept 2-bit
 * packed in HBM.  Replaces what VertexIndex reads from
 * SequenceContainer (patch 0002 file lines 52-95). */
int flyehip_store_reads(flyehip_ctx* ctx, const uint8_t* bases,
                        const int64_t* offsets, int32_t n_reads);

/* Counts canonical k-mers, fixes the thresholds and builds the bucket index
 * in HBM (VertexIndex::countKmers / buildIndex; CUDA_FLYE_DESIGN.md:81-86). */
int flyehip_build_index(flyehip_ctx* ctx, const flyehip_index_params_t* params,
                        flyehip_index_info_t* out_info);

/* The edge-sequence index of read-to-graph alignment
 * (VertexIndex::buildIndexMinimizers(minCoverage, wndLen) as ReadAligner
 * builds pathsIndex; its parameters are recorded by
 * patches/flye/2.9.6/0040-cuflye-read-to-graph-minimizer-source-pack.patch
 * file lines 119-124 and 264-266, and Flye's quickSeqOverlaps then joins every
 * query k-mer against it, docs/abi/read-to-graph-minimizer-source-pack-v0.md
 * "Full Query-Hit TSV").  Counting and thresholds are flyehip_build_index's;
 * buckets hold only minimizer positions.  Flye's yieldMinimizers is not in
 * the reference, so the selection is this library's: in every run of
 * `window` consecutive k-mer positions of a stored read (all of its
 * positions when it has fewer), the position whose canonical k-mer has the
 * smallest 64-bit mix (the table's splitmix64 finaliser) is a minimizer, the
 * leftmost on equal values; density about 2 / (window + 1).  A k-mer with no
 * minimizer position is not in the index (no hits, not filtered).  For
 * Flye's exact pathsIndex upload it with flyehip_load_index. */
int flyehip_build_index_minimizers(flyehip_ctx* ctx, const flyehip_index_params_t* params,
                                   int32_t window, flyehip_index_info_t* out_info);

/* Uploads Flye's own VertexIndex instead of building one: the backend of
 * patch 0005:245-266 reads Flye's `const VertexIndex&` (iterKmerPos,
 * isRepetitive, kmerFreq; patch 0002 file lines 52-95), so bucket order - and
 * with it the emission order and std::sort's tie order - is Flye's.
 *   kmers[i]            canonical k-mer (Kmer::standardForm, 2 bits per base,
 *                       MSB first) of bucket i, any order, no repeats;
 *   bucket_offsets      n_kmers + 1 offsets into the entries;
 *   entry_ids/entry_pos bucket entries in Flye's order: the signed read id and
 *                       position where the canonical k-mer is spelled (what
 *                       iterKmerPos returns for the canonical k-mer itself);
 *   repetitive          canonical k-mers VertexIndex::isRepetitive holds.
 * The reads must be stored first (flyehip_store_reads: query sequences and
 * the read lengths that orient entries for non-canonical query windows).
 * Every entry must lie inside its read; a k-mer that is not canonical or
 * appears twice is FLYEHIP_ERR_ARG. */
int flyehip_load_index(flyehip_ctx* ctx, int32_t kmer_size,
                       const uint64_t* kmers, int64_t n_kmers,
                       const int64_t* bucket_offsets,
                       const int64_t* entry_ids, const int32_t* entry_pos,
                       const uint64_t* repetitive, int64_t n_repetitive,
                       flyehip_index_info_t* out_info);

/* ---- candidate collection (collectCandidateMatches, patch 0005:245-266) -- */

/* For each query record query_ids[i] (a signed id of a resident read), writes
 * its k-mer matches in Flye's emission order (window position ascending,
 * bucket order inside a window) and its repetitive-window positions.
 * match_offsets[i..i+1] and filtered_offsets[i..i+1] delimit query i's rows
 * (both arrays hold n_queries + 1 entries).  When a capacity is too small the
 * call returns FLYEHIP_ERR_CAPACITY and sets *matches_needed and
 * *filtered_needed; a second call with that capacity succeeds. */
int flyehip_collect_candidates(flyehip_ctx* ctx, const int64_t* query_ids,
                               int32_t n_queries,
                               flyehip_candidate_t* matches, int64_t match_capacity,
                               int64_t* match_offsets, int64_t* matches_needed,
                               int32_t* filtered, int64_t filtered_capacity,
                               int64_t* filtered_offsets, int64_t* filtered_needed);

/* The candidate-record-v1 canonical SHA-256 of the query records' matches
 * (docs/abi/candidate-record-v1.md:31-68; the hash
 * tools/validate_candidate_dump.py:76-134 computes over the dump's records
 * in ascending tuple order), streamed: the join runs in batches of query
 * records in ascending id order, each window's matches are ordered by
 * (target_id, target_pos) and formatted, and the text goes through SHA-256
 * without the dump being sorted or held.  sha256 receives 32 bytes;
 * *records the record count. */
int flyehip_candidates_canonical_sha256(flyehip_ctx* ctx, const int64_t* query_ids,
                                        int32_t n_queries, uint8_t* sha256, int64_t* records);

/* Same, with the matches in Flye's sorted order: std::sort by
 * (internal extId, curPos) with libstdc++'s tie order
 * (patch 0001:88-91); this is the order candidate-record-v1 dumps. */
int flyehip_collect_sorted_candidates(flyehip_ctx* ctx, const int64_t* query_ids,
                                      int32_t n_queries,
                                      flyehip_candidate_t* matches, int64_t match_capacity,
                                      int64_t* match_offsets, int64_t* matches_needed);

/* ---- whole overlap stage (getSeqOverlaps, subsystems a+b) --------------- */

/* Runs k-mer join, match sort, chaining DP and filters for every query record
 * and writes each query's detectedOverlaps in Flye's order.  row_offsets has
 * n_queries + 1 entries.  div_stats (optional, may be NULL) receives, per
 * query, the divergences OvlpDivStats::add would get from its 10 kb windows
 * (patch 0023:245-269), div_offsets delimiting them.  When row_capacity or
 * div_capacity is too small the call returns FLYEHIP_ERR_CAPACITY with
 * *rows_needed and div_offsets[n_queries] set to the counts it needs (a query
 * of length L has at most L / 10000 + 1 values); a second call with those
 * capacities succeeds. */
int flyehip_overlaps(flyehip_ctx* ctx, const int64_t* query_ids, int32_t n_queries,
                     const flyehip_overlap_params_t* params,
                     flyehip_overlap_t* rows, int64_t row_capacity,
                     int64_t* row_offsets, int64_t* rows_needed,
                     float* div_stats, int64_t div_capacity, int64_t* div_offsets);

/* Flye's getSeqOverlaps runs once per query record on a pool of threads
 * (OverlapContainer::findAllOverlaps), while the GPU wants thousands of
 * query records per batch; the cuFlye backend seam is thread-safe behind a
 * mutex (patch 0004:105-109, 252).  The submit / collect pair batches those
 * threads' work:
 *   flyehip_overlaps_submit  queues query records (resident ids, typically
 *                            every read strand the thread pool will ask for)
 *                            and returns at once with a ticket; one library
 *                            thread runs the queued jobs in submission order,
 *                            queries_per_batch x batches_in_flight query
 *                            records at a time, as flyehip_overlaps would;
 *   flyehip_overlaps_collect any thread, any order: blocks until the batch
 *                            holding query_index of the ticket's list is done
 *                            and copies that query's rows (and, with
 *                            with_div_stats, its divStats values).  A small
 *                            capacity returns FLYEHIP_ERR_CAPACITY with the
 *                            counts needed; a failed batch returns its error
 *                            for every query it held;
 *   flyehip_overlaps_release waits for the job to finish and frees its rows.
 * Rows are those flyehip_overlaps returns for the same list.  Close the
 * context only after every collect has returned. */
int flyehip_overlaps_submit(flyehip_ctx* ctx, const int64_t* query_ids, int32_t n_queries,
                            const flyehip_overlap_params_t* params, int32_t with_div_stats,
                            uint64_t* ticket);
int flyehip_overlaps_collect(flyehip_ctx* ctx, uint64_t ticket, int32_t query_index,
                             flyehip_overlap_t* rows, int64_t row_capacity, int64_t* rows_needed,
                             float* div_stats, int64_t div_capacity, int64_t* divs_needed);
int flyehip_overlaps_release(flyehip_ctx* ctx, uint64_t ticket);

/* Same for query sequences that are not resident reads (the serving path and
 * read-to-graph alignment): bases/offsets as in flyehip_store_reads,
 * query_ids are their signed ids (a query never self-matches unless its id
 * names a resident read). */
int flyehip_overlaps_external(flyehip_ctx* ctx, const uint8_t* bases,
                              const int64_t* offsets, const int64_t* query_ids,
                              int32_t n_queries,
                              const flyehip_overlap_params_t* params,
                              flyehip_overlap_t* rows, int64_t row_capacity,
                              int64_t* row_offsets, int64_t* rows_needed);

/* Plain per-query path: the same stages launched once per query record, for
 * the batched-versus-unbatched comparison.  Same outputs as flyehip_overlaps. */
int flyehip_overlaps_per_query(flyehip_ctx* ctx, const int64_t* query_ids,
                               int32_t n_queries,
                               const flyehip_overlap_params_t* params,
                               flyehip_overlap_t* rows, int64_t row_capacity,
                               int64_t* row_offsets, int64_t* rows_needed);

/* ---- chaining of given hit lists (subsystem b alone) -------------------- */

/* Chains n_lists (query, target) hit lists.  List i has hits
 * [list_offsets[i], list_offsets[i+1]) of (cur_pos, ext_pos) pairs already in
 * Flye's group order (curPos ascending, ties as given); cur_ids/ext_ids are
 * signed ids and cur_lens/ext_lens the record lengths.  filtered positions are
 * empty.  Writes each list's overlaps (row_offsets has n_lists + 1 entries). */
int flyehip_chain_lists(flyehip_ctx* ctx, const int32_t* cur_pos,
                        const int32_t* ext_pos, const int64_t* list_offsets,
                        int32_t n_lists, const int64_t* cur_ids,
                        const int64_t* ext_ids, const int32_t* cur_lens,
                        const int32_t* ext_lens,
                        const flyehip_overlap_params_t* params,
                        flyehip_overlap_t* rows, int64_t row_capacity,
                        int64_t* row_offsets, int64_t* rows_needed);

/* ---- read-to-graph chaining (subsystem c) ------------------------------- */

/* For each read of the batch: keeps the edge alignments that pass
 * `ext_len < long_edge || min(cur_range, ext_range) > big_aln`, sorts them by
 * read_begin with std::sort's tie order (patch 0039:72-77, 241-251), and runs
 * ReadAligner::chainReadAlignments (tools/replay_read_alignment_chains.py:64-139).
 * alignments[aln_offsets[r] .. aln_offsets[r+1]) belong to read r.
 * sorted_index (capacity = total alignments) receives, per read, the indices
 * of its kept alignments in sorted order (segments' input_index points into
 * this order); kept_offsets delimits them (n_reads + 1 entries).  Accepted
 * chains are written as segments in acceptance order. */
int flyehip_read_align_chains(flyehip_ctx* ctx,
                              const flyehip_edge_alignment_t* alignments,
                              const int64_t* aln_offsets, int32_t n_reads,
                              int32_t long_edge, int32_t big_aln,
                              const flyehip_chain_params_t* params,
                              int32_t* sorted_index, int64_t* kept_offsets,
                              flyehip_chain_segment_t* segments,
                              int64_t segment_capacity, int64_t* segment_offsets,
                              int64_t* segments_needed);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): flyehip_read_align_chains with every array in device
 * memory the caller owns - d_alignments / d_aln_offsets in, d_sorted_index
 * (capacity aln_offsets[n_reads]), d_kept_offsets and d_segment_offsets
 * (n_reads + 1 each) and d_segments out - and every intermediate buffer in
 * one fixed device array the library holds per device (about 43 MiB), so the
 * call allocates nothing.  Limits: n_reads <= 131072 and
 * aln_offsets[n_reads] <= 1048576 (FLYEHIP_ERR_ARG above them).  Calls that
 * overlap on one device share that array and race; serialise them.  Results
 * equal flyehip_read_align_chains'. */
int flyehip_read_align_chains_noalloc(flyehip_ctx* ctx,
                                      const flyehip_edge_alignment_t* d_alignments,
                                      const int64_t* d_aln_offsets, int32_t n_reads,
                                      int32_t long_edge, int32_t big_aln,
                                      const flyehip_chain_params_t* params,
                                      int32_t* d_sorted_index, int64_t* d_kept_offsets,
                                      flyehip_chain_segment_t* d_segments,
                                      int64_t segment_capacity, int64_t* d_segment_offsets,
                                      int64_t* segments_needed);

/* ---- the divergence logarithm --------------------------------------------- */

/* out[i] = logf(x[i]) exactly as glibc's logf returns it, computed on the GPU
 * the way the chaining kernels compute seq_divergence
 * (tools/replay_overlap_chains.py:87-111, 348-361: float logf of 1 / matchRate).
 * Exact for every float in [1, 2^25), the values 1 / matchRate takes. */
int flyehip_divergence_logf(flyehip_ctx* ctx, const float* x, int64_t n, float* out);

#ifdef __cplusplus
}
#endif

#endif /* FLYEHIP_H */
