/*
 * kc.h — C ABI of the MI355X (gfx950) k-mer spectrum engine.
 *
 * The engine replaces the first-stage k-mer path of spades-rs:
 *   2-bit canonical k-mer extraction from packed reads   (dna.rs, packed_reads.rs)
 *   two-tier Bloom filter insert / query                  (bloom.rs)
 *   Hammer single-base correction against the filter      (hammer.rs)
 *   exact count table, cutoff and solid set               (assemble.rs:122-298, 692-735)
 *
 * Conventions
 *   - Plain pointers and sizes only.  Host buffers are borrowed for the call.
 *   - Every function returns KC_OK (0) or a negative KC_E* status; nothing
 *     throws or aborts across the ABI.  kc_last_error() / kc_filter_last_error()
 *     give a message for the last failure on that object.
 *   - A k-mer is passed as the reference's Kmer256(lo: u128, hi: u128)
 *     (dna.rs:137-138) in four little-endian 64-bit limbs, the limb order of
 *     bloom.rs:62-65:  w[0] = lo bits 0..63, w[1] = lo bits 64..127,
 *     w[2] = hi bits 0..63, w[3] = hi bits 64..127.  For k <= 64 the k-mer
 *     occupies the low 2k bits of lo, first base most significant, hi = 0.
 *   - One object owns one HIP stream on one device.  Calls on one object must be
 *     serialised by the caller; separate objects may be used concurrently.
 *   - Results are deterministic: identical inputs give identical outputs.  The
 *     solid Bloom tier is the one the reference produces in its sequential
 *     order (read 0..n, window 0..L-k, hashes h1, h2, h3), i.e. at `-t 1`.
 */
#ifndef KC_H
#define KC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KC_OK 0
#define KC_EINVAL (-1)       /* bad argument (k out of range, NULL pointer, size mismatch) */
#define KC_ENOMEM (-2)       /* device allocation failed */
#define KC_EHIP (-3)         /* HIP runtime error */
#define KC_ESTATE (-4)       /* call out of order (e.g. export before run) */
#define KC_EUNSUPPORTED (-5) /* valid in the reference but not supported by this build */

#define KC_MAX_K 127u        /* 1 <= k <= 127 (the reference's CLI range is odd k in 11..127) */

typedef struct kc_ctx kc_ctx;
typedef struct kc_filter kc_filter;

/* Options of one spectrum run; mirrors the AssemblerConfig fields the path reads
 * (assemble.rs:35-103): k, bloom_bits, error_correct, min_coverage. */
typedef struct kc_opts {
    int32_t device;        /* HIP device ordinal */
    uint32_t k;            /* k-mer size, 1..KC_MAX_K (CLI: odd, 11..127; main.rs:11-28) */
    uint64_t bloom_bits;   /* capacity bits per tier (TwoTierFilter::new, bloom.rs:26-43) */
    double min_coverage;   /* determine_solid_kmer_cutoff input (assemble.rs:692) */
    int32_t error_correct; /* run Hammer before counting (assemble.rs:188-192) */
    float table_load;      /* device hash-table target load factor, 0 = default (0.5) */
    uint32_t flags;        /* KC_FLAG_* */
} kc_opts;

#define KC_FLAG_TIMING 1u  /* record per-kernel HIP events (kc_kernel_times) */
#define KC_FLAG_EVENTS64 2u /* 64-bit event indices (global window numbers) even when 32 bits would
                              do: the path a library of 2^30 - 1 windows or more takes, selectable
                              on any input (the results are the same) */

/* Summary of the last run. */
typedef struct kc_info {
    uint64_t n_reads;          /* reads uploaded */
    uint64_t n_occ;            /* k-mer occurrences (windows) = sum max(len-k+1, 0) */
    uint64_t n_distinct;       /* distinct canonical k-mers over all windows */
    uint64_t n_counts;         /* entries of the count table (Bloom-solid distinct k-mers) */
    uint64_t n_solid;          /* entries with count >= cutoff (the solid set) */
    uint64_t corrected_reads;  /* reads Hammer modified (error_correct only) */
    uint64_t num_words;        /* 64-bit words per Bloom tier */
    uint64_t table_capacity;   /* device hash-table slots used */
    uint32_t cutoff;           /* determine_solid_kmer_cutoff result */
    uint32_t top_cov;          /* 95th-percentile count used by the cutoff, capped at 1023 */
    double ms_stage2;          /* device time of the Bloom build */
    double ms_stage3;          /* device time of correction + counting + sort + cutoff */
    uint64_t hammer_queries;   /* is_solid tests Hammer made in the last kc_count (error_correct;
                                  all libraries of the context) */
} kc_info;

/* ---- Standalone two-tier filter (bloom.rs) ------------------------------- */

/* TwoTierFilter::new(capacity_bits) (bloom.rs:26-43): max(ceil(bits/64), 1) words per tier. */
int kc_filter_create(kc_filter** out, int32_t device, uint64_t capacity_bits);
void kc_filter_destroy(kc_filter* f);
const char* kc_filter_last_error(const kc_filter* f);
/* TwoTierFilter::memory_usage_bytes (bloom.rs:129-131). */
uint64_t kc_filter_memory_usage_bytes(const kc_filter* f);
uint64_t kc_filter_num_words(const kc_filter* f);
/* TwoTierFilter::insert_kmer256 (bloom.rs:83-107) applied to keys[0], keys[1], ... in that
 * order, after every earlier call.  promoted (nullable) receives each call's return value. */
int kc_filter_insert_kmer256(kc_filter* f, const uint64_t* keys4, size_t n, uint8_t* promoted);
/* TwoTierFilter::is_solid_kmer256 (bloom.rs:111-126) for each key. */
int kc_filter_is_solid_kmer256(kc_filter* f, const uint64_t* keys4, size_t n, uint8_t* out);
/* hamming_distance_2bit (hammer.rs:13-18) for n pairs: out[i] = the number of bases (2-bit
 * groups) in which a[i] and b[i] differ, popcount((d | d >> 1) & 0x5555...) with d = a ^ b. */
int kc_hamming_distance_2bit(int32_t device, const uint64_t* a, const uint64_t* b, size_t n, uint32_t* out);
/* Copy both tiers (num_words u64 each). */
int kc_filter_export(kc_filter* f, uint64_t* seen, uint64_t* solid, size_t num_words);

/* The same correction as kc_filter_correct_reads, for callers that cannot allocate (code
 * that runs under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  No counterpart in the reference.  Every pointer except n_corrected is
 * DEVICE memory on the filter's device: d_data (data_len bytes, PackedReads layout),
 * d_offsets / d_lengths (n_reads each) and d_out (data_len bytes, receives the corrected
 * store).  The call allocates nothing: its scratch (counters, the active-read list) is one
 * static device array per GPU of 512 + 4 * KC_NOALLOC_MAX_READS bytes, held until the
 * process ends and shared by every call on that GPU, so calls that overlap on one GPU
 * race.  More than KC_NOALLOC_MAX_READS reads give KC_EINVAL, as does a read past
 * data_len.  Synchronises the filter's stream before returning. */
#define KC_NOALLOC_MAX_READS (1u << 22)
int kc_filter_correct_reads_noalloc(kc_filter* f, uint32_t k, const uint8_t* d_data, size_t data_len,
                                    const uint64_t* d_offsets, const uint32_t* d_lengths, size_t n_reads,
                                    uint8_t* d_out, uint64_t* n_corrected);
/* ErrorCorrector::new(k).correct_read_in_place (hammer.rs:28-120) for every read of a
 * PackedReads-layout store against this filter; `out` receives the corrected store (same
 * offsets/lengths), n_corrected the number of reads modified (hammer.rs:123-143). */
int kc_filter_correct_reads(kc_filter* f, uint32_t k, const uint8_t* data, size_t data_len, const uint64_t* offsets,
                            const uint32_t* lengths, size_t n_reads, uint8_t* out, uint64_t* n_corrected);

/* ---- Spectrum context (assemble.rs Stage 2 + Stage 3) -------------------- */

int kc_create(kc_ctx** out, const kc_opts* opts);
void kc_destroy(kc_ctx* ctx);
const char* kc_last_error(const kc_ctx* ctx);

/* Upload a PackedReads store (packed_reads.rs:42-59): data = 2-bit bases, 4 per byte,
 * first base in bits 7..6, each read byte-aligned at offsets[i]; lengths[i] bases. */
int kc_reads_upload(kc_ctx* ctx, const uint8_t* data, size_t data_len, const uint64_t* offsets,
                    const uint32_t* lengths, size_t n_reads);

/* Change k for the next run, keeping every library's reads resident on the device: the
 * multi-k loop (multik.rs:109-178) runs k = 21, 33, 55 ... over one upload.  Results of
 * the previous k are dropped. */
int kc_set_k(kc_ctx* ctx, uint32_t k);

/* Stage 2: stream every window of every read into the two-tier filter (assemble.rs:129-161). */
int kc_bloom_build(kc_ctx* ctx);
/* Stage 3: optional Hammer (error_correct), exact counts of Bloom-solid k-mers, sort,
 * cutoff and solid set (assemble.rs:166-292).  Requires kc_bloom_build. */
int kc_count(kc_ctx* ctx, kc_info* info);
/* Stage 2 + Stage 3. */
int kc_run(kc_ctx* ctx, kc_info* info);
/* kc_run of n independent contexts at once, each from its own host thread on its own
 * stream, so that one context's kernels fill the device while another waits on a size
 * read-back or runs a kernel bound by another resource (the multi-k loop without prior
 * contigs: one context per k over the same reads).  infos may be NULL, else n entries.
 * Each context's results are those of kc_run alone.  Returns the first failing context's
 * status (the others still finish); ctxs[i] must be distinct. */
int kc_run_many(kc_ctx* const* ctxs, uint32_t n, kc_info* infos);
int kc_info_get(const kc_ctx* ctx, kc_info* info);

/* Both Bloom tiers after Stage 2 (num_words must equal kc_info.num_words). */
int kc_bloom_export(kc_ctx* ctx, uint64_t* seen, uint64_t* solid, size_t num_words);
/* is_solid_kmer256 (bloom.rs:111-126) against the context's filter. */
int kc_bloom_query(kc_ctx* ctx, const uint64_t* keys4, size_t n, uint8_t* out);
/* The count table sorted ascending by Kmer256 order: keys4 (n x 4 limbs), counts, and
 * solid_mask[i] = counts[i] >= cutoff.  Any of the three may be NULL.  n must equal
 * kc_info.n_counts. */
int kc_counts_export(kc_ctx* ctx, uint64_t* keys4, uint32_t* counts, uint8_t* solid_mask, size_t n);
/* Packed reads as counted (after Hammer when error_correct), same offsets and lengths
 * as uploaded; data_len must equal the uploaded data_len. */
int kc_reads_export(kc_ctx* ctx, uint8_t* data, size_t data_len);

/* Checks of a finished run of library lib that recompute its closed forms on the device,
 * window by window, independently of how the run built them (no counterpart in the
 * reference; for runs too large to compare with a CPU restatement):
 *   seen_words_differing  words of the seen tier that differ from the OR of every window's
 *                         three bits (bloom.rs:95-97): 0 for a correct run;
 *   solid_windows         windows of the counted reads (corrected when error_correct)
 *                         whose canonical k-mer is Bloom-solid; without priors it equals
 *                         sum_counts, the sum of the count table's counts (assemble.rs:196);
 *   unordered_pairs       adjacent entries of the sorted table not strictly ascending: 0;
 *   mask_mismatches       entries whose solid flag differs from count >= cutoff: 0;
 *   checksum              the sum over entries i of a 64-bit mix of (i, k-mer value words,
 *                         count): equal tables give equal checksums. */
typedef struct kc_verify {
    uint64_t windows, solid_windows, sum_counts, seen_words_differing, unordered_pairs, mask_mismatches, checksum;
} kc_verify;
int kc_library_verify(kc_ctx* ctx, uint32_t lib, kc_verify* out);

/* ---- Several libraries in one context ------------------------------------
 * A context may hold many independent libraries (read sets).  Each is counted exactly as
 * if it were alone in its own context (own filter, event order, table, cutoff and solid
 * set), but every stage of kc_run processes all of them with the same kernel launches.
 * kc_reads_upload replaces all libraries by one (library 0, with the context's
 * bloom_bits and min_coverage); the functions without a library argument act on library 0. */

/* Append a library; *lib receives its index.  The reads are copied (borrowed only during
 * the call).  bloom_bits / min_coverage as in kc_opts, for this library. */
int kc_library_add(kc_ctx* ctx, const uint8_t* data, size_t data_len, const uint64_t* offsets,
                   const uint32_t* lengths, size_t n_reads, uint64_t bloom_bits, double min_coverage,
                   uint32_t* lib);
int kc_library_count(const kc_ctx* ctx, uint32_t* n);
int kc_library_info(const kc_ctx* ctx, uint32_t lib, kc_info* info);
int kc_library_bloom_export(kc_ctx* ctx, uint32_t lib, uint64_t* seen, uint64_t* solid, size_t num_words);
int kc_library_bloom_query(kc_ctx* ctx, uint32_t lib, const uint64_t* keys4, size_t n, uint8_t* out);
int kc_library_counts_export(kc_ctx* ctx, uint32_t lib, uint64_t* keys4, uint32_t* counts, uint8_t* solid_mask,
                             size_t n);
int kc_library_reads_export(kc_ctx* ctx, uint32_t lib, uint8_t* data, size_t data_len);

/* Prior / trusted contigs of a library (assemble.rs:244-284: AssemblerConfig::prior_contigs
 * plus the sequences of trusted_contigs files).  n_seqs ASCII sequences, sequence i at
 * seqs[starts[i] .. starts[i] + lengths[i]).  After Stage 3's count, every window whose k
 * bases are all ACGT/acgt adds prior_boost = max((u32)min_coverage * 4, 50) (u32 wrapping)
 * to its canonical k-mer, inserting it when absent; the cutoff and solid set then see the
 * boosted counts.  Replaces the library's previous priors (n_seqs = 0 clears them); the
 * sequences are copied.  Takes effect at the next kc_bloom_build / kc_run, which sizes the
 * table for the prior windows. */
int kc_library_priors_set(kc_ctx* ctx, uint32_t lib, const uint8_t* seqs, size_t seqs_len, const uint64_t* starts,
                          const uint64_t* lengths, size_t n_seqs);
int kc_priors_set(kc_ctx* ctx, const uint8_t* seqs, size_t seqs_len, const uint64_t* starts, const uint64_t* lengths,
                  size_t n_seqs); /* library 0 */

/* ---- FASTQ ingestion on the device (packed_reads.rs) ------------------------
 * File contents (uncompressed) of a paired set -> the three-block PackedReads layout of
 * PackedReads::from_paired_fastq (packed_reads.rs:325-380): FastxReader's record checks,
 * the three N gates (terminal trim, split on internal N into >= 21-base sub-reads, the
 * ambiguous sidecar) and 2-bit packing all run on the GPU.  r2 = NULL ingests one file
 * as append_from_file does (every read through the gates, pe_count 0).  The device
 * handles FASTQ (blank lines between records skipped where a header is due, as
 * FastxReader does) and FASTA (single- or multi-line records, packed_reads.rs:217-235);
 * FASTA with kc_ingest_keep_quals returns KC_EUNSUPPORTED, malformed records KC_EINVAL;
 * kc_fastq_last_error() gives the message. */
typedef struct kc_packed kc_packed;
typedef struct kc_packed_info {
    uint64_t n_reads, data_len, pe_count; /* the PackedReads arrays and its paired block size */
    uint64_t n_ambiguous;                 /* reads routed to the ambiguous sidecar (not packed) */
    uint64_t r1_records, r2_records;      /* records read from each file */
    double ms_device;                     /* device time from the end of the upload to the end of the
                                             last kernel (includes the host reads of sizes between
                                             kernels, not the copies in and out) */
    uint64_t n_quals;                     /* bytes of the quality side array (0 unless kept) */
} kc_packed_info;
/* One-shot form: a context for the call only. */
int kc_fastq_pack(int32_t device, const uint8_t* r1, size_t r1_len, const uint8_t* r2, size_t r2_len,
                  kc_packed** out);
/* A context keeps its stream and device buffers between calls (they only grow), so that
 * repeated calls allocate nothing once the largest input has been seen.  One thread at a
 * time per context. */
typedef struct kc_ingest kc_ingest;
int kc_ingest_create(int32_t device, kc_ingest** out);
void kc_ingest_destroy(kc_ingest* ctx);
int kc_ingest_pack(kc_ingest* ctx, const uint8_t* r1, size_t r1_len, const uint8_t* r2, size_t r2_len,
                   kc_packed** out);
/* Keep a quality side array (no counterpart in the reference, which drops qualities;
 * the learned pass's features read it, kc_library_quals_set): the Phred value
 * (character - 33, 0 below '!') of every packed base, read i's at quals[qoffsets[i] ..]. */
int kc_ingest_keep_quals(kc_ingest* ctx, int32_t keep);
/* The message of the calling thread's last failed ingestion call. */
const char* kc_fastq_last_error(void);
int kc_packed_info_get(const kc_packed* p, kc_packed_info* info);
int kc_packed_export(const kc_packed* p, uint8_t* data, size_t data_len, uint64_t* offsets, uint32_t* lengths,
                     size_t n_reads);
int kc_packed_export_quals(const kc_packed* p, uint8_t* quals, size_t n_quals, uint64_t* qoffsets, size_t n_reads);
void kc_packed_destroy(kc_packed* p);

/* ---- Unique-anchor k-mer indexes over contigs / unitigs ---------------------
 * The maps canonical k-mer -> (sequence, position, is_rc) that the reference builds from
 * ASCII sequences with the counting path's k-mer extraction (windows with a base other
 * than ACGT/acgt are skipped):
 *   KC_ANCHOR_POLISHER     Polisher::build_kmer_index (polisher.rs:33-60): a second
 *                          sighting drops the k-mer, for k < 11 only one in another contig;
 *   KC_ANCHOR_PAIRED_INFO  PairedInfoIndex::build_kmer_index (paired_info.rs:34-70): any
 *                          second sighting drops it;
 *   KC_ANCHOR_SPLITTER     the splitter's anchor step (splitter.rs:57-70): a sighting in
 *                          another contig drops it (the value used there is the contig).
 * The value is the k-mer's first occurrence.  kc_anchor_export gives the anchors sorted
 * by k-mer as Kmer256 limbs; kc_anchor_query looks up canonical k-mers (limbs) and gives
 * the anchor's index in that order, or -1. */
#define KC_ANCHOR_POLISHER 0
#define KC_ANCHOR_PAIRED_INFO 1
#define KC_ANCHOR_SPLITTER 2
typedef struct kc_anchor kc_anchor;
int kc_anchor_build(int32_t device, uint32_t k, int32_t kind, const uint8_t* seqs, size_t seqs_len,
                    const uint64_t* starts, const uint64_t* lengths, size_t n_seqs, kc_anchor** out);
int kc_anchor_count(const kc_anchor* a, uint64_t* n);
int kc_anchor_export(kc_anchor* a, uint64_t* keys4, uint32_t* seq_ids, uint32_t* positions, uint8_t* is_rc,
                     size_t n);
int kc_anchor_query(kc_anchor* a, const uint64_t* keys4, size_t n, int64_t* index);
const char* kc_anchor_last_error(void);
void kc_anchor_destroy(kc_anchor* a);

/* ---- Compacted unitigs of a solid set (graph.rs:65-234) ---------------------
 * CompactedGraph::build over library lib's solid set after kc_count, on the context's
 * device: the maximal non-branching paths of solid canonical k-mers, each as its ASCII
 * sequence (the first k-mer's k bases, then the last base of every further k-mer), its
 * k-mer count and its mean count (the library's counts; graph.rs:159-178).  Unitigs come
 * ordered by their first k-mer's rank in the sorted count table; a path's strand may be
 * the reverse complement of the reference's walk (which follows hash order), and a cycle
 * starts at its smallest k-mer, forward strand (the reference iterating in ascending
 * order, graph.rs:216-230).  Odd k only (KC_EUNSUPPORTED otherwise).  Build errors are
 * reported through kc_last_error(ctx); the result is independent of the context. */
typedef struct kc_unitigs kc_unitigs;
typedef struct kc_unitigs_info {
    uint64_t n_unitigs, n_bases, n_kmers;
    double ms_device;   /* device time of the build */
    double ms_stage[4]; /* of which: decode + index, adjacency + links, ranking, emit */
} kc_unitigs_info;
int kc_unitigs_build(kc_ctx* ctx, uint32_t lib, kc_unitigs** out);
int kc_unitigs_info_get(const kc_unitigs* u, kc_unitigs_info* info);
/* seq: n_bases bytes (unitig i at seq[offsets[i] .. offsets[i+1])); offsets: n_unitigs + 1;
 * kmers, mean_cov: n_unitigs.  Any pointer may be NULL. */
int kc_unitigs_export(kc_unitigs* u, uint8_t* seq, size_t n_bases, uint64_t* offsets, uint64_t* kmers,
                      double* mean_cov, size_t n_unitigs);
const char* kc_unitigs_last_error(const kc_unitigs* u);
void kc_unitigs_destroy(kc_unitigs* u);

/* ---- Tuning and graph replay ---------------------------------------------- */

/* Launch shape of the table insert and the table load factor.  Every setting gives
 * identical results; only the speed differs. */
typedef struct kc_tune {
    uint32_t windows_per_thread; /* 1, 2 or 4: table probes each thread keeps in flight */
    uint32_t blocks_per_cu;      /* 0: one workgroup per 1024-window tile; else CUs x this
                                    workgroups, each looping over tiles.  The binned count:
                                    0 = as many workgroups as are resident at once, else CUs x
                                    this, each looping over bins */
    float table_load;            /* device hash-table target load factor (0: kc_opts.table_load) */
    uint32_t insert_path;        /* Stage 2 count table.  0 (default) or 3 = minimizer-binned:
                                    reads cut into super-k-mers, radix-sorted by bin, each bin
                                    counted in LDS, the table packed; taken for k <= 63 and
                                    reads shorter than 2^24 bases outside graph capture, and
                                    otherwise the table insert below.  1 = device-scope table
                                    atomics per window.  Other values give KC_EINVAL */
    uint32_t setter_path;        /* Bloom first setters.  0 (default) = sorted: three
                                    (entry, setter) items per distinct k-mer radix-sorted by
                                    partition (16384 entries with 32-bit events, 8192 with
                                    64-bit ones), each partition's minima taken in LDS and
                                    written in order; taken in the dense mode outside graph
                                    capture (with 32-bit events also inside it), and otherwise
                                    the atomics below.  1 = three device-scope atomicMin per
                                    distinct k-mer */
    uint32_t kmers_per_bin;      /* binned count: distinct k-mers a bin is sized for (0: 352;
                                    else 64..4096) */
    uint32_t minimizer_len;      /* binned count: m of the canonical m-mers whose minimum bins a
                                    window (0: the default for k; else 1..32, at most k) */
    double ms;                   /* kc_autotune: Stage 2 + 3 device time of the chosen setting */
} kc_tune;

int kc_tune_get(const kc_ctx* ctx, kc_tune* tune);
int kc_tune_set(kc_ctx* ctx, const kc_tune* tune);
/* Time Stage 2 + 3 over the uploaded libraries for a coordinate search (reps timed runs
 * each, after one warm-up) over the knobs of the Stage 2 path the run takes: with the
 * binned count, the distinct k-mers a bin is sized for, the minimizer length and the
 * workgroups per CU; with
 * the table insert, the load factor, the windows per thread and the workgroups per CU;
 * then the first-setter path.  Every setting must give the same per-library summary; the
 * fastest is kept.  *best (nullable) receives it. */
int kc_autotune(kc_ctx* ctx, uint32_t reps, kc_tune* best);

/* Stage 2 + Stage 3 as one HIP graph launch.  The first call (and the first after new
 * reads, priors, tuning or an eager kc_bloom_build / kc_count) runs the stages eagerly
 * to fix every size (table layout, count-table entries), then captures their device work
 * into a graph; later calls replay it.  A replay whose data-dependent sizes differ from
 * the plan (entries, table overflow, read length under error_correct) is redone eagerly
 * and re-captured, so results always equal kc_run's.  The per-read preparation that
 * depends only on the uploaded lengths is done at planning time.  In info, ms_stage2
 * holds the whole replay and ms_stage3 is 0. */
int kc_graph_run(kc_ctx* ctx, kc_info* info);
int kc_graph_release(kc_ctx* ctx);

/* ---- Learned first pass: the network ----------------------------------------
 * A 16 -> 64 -> 64 -> 1 MLP (ReLU, ReLU, sigmoid) that predicts a k-mer's solid label
 * from 16 features.  bf16 weights and inputs, fp32 accumulation and biases; inference
 * on MFMA.  Parameters are kc_mlp_num_params() floats: W1[16][64], b1[64], W2[64][64],
 * b2[64], w3[64], b3.  No counterpart in the reference. */
typedef struct kc_mlp kc_mlp;
int kc_mlp_create(kc_mlp** out, int32_t device, uint64_t seed); /* seeded He initialisation */
void kc_mlp_destroy(kc_mlp* m);
const char* kc_mlp_last_error(const kc_mlp* m);
uint32_t kc_mlp_num_params(void);
int kc_mlp_get_params(kc_mlp* m, float* params, size_t n);
int kc_mlp_set_params(kc_mlp* m, const float* params, size_t n);
/* p[i] = sigmoid(network(X[i])), X row-major n x 16. */
int kc_mlp_forward(kc_mlp* m, const float* X, size_t n, float* p);
/* Gradient of mean_i w_i * BCE(p_i, y_i) with w_i = sample_w[i] (1 when NULL) times
 * pos_weight for y_i = 1, with respect to the parameters (straight through the bf16
 * roundings); loss (nullable) receives the loss. */
int kc_mlp_gradient(kc_mlp* m, const float* X, const uint8_t* y, const float* sample_w, size_t n, float pos_weight,
                    float* grad, double* loss);
/* Adam (0.9, 0.999, 1e-8) over `epochs` shuffled passes of mini-batches; deterministic
 * for a given seed.  last_loss (nullable) receives the last batch's loss. */
int kc_mlp_train(kc_mlp* m, const float* X, const uint8_t* y, const float* sample_w, size_t n, uint32_t epochs,
                 uint32_t batch, float lr, float pos_weight, uint64_t seed, double* last_loss);

/* ---- Learned first pass: features and selection -------------------------------
 * Need kc_count with error_correct off (the table then holds the uncorrected counts). */

/* Per-base Phred qualities of a library's reads (one byte per base; read i's bases at
 * quals[qoffsets[i] .. + length_i]); copied.  n_quals = 0 removes them. */
int kc_library_quals_set(kc_ctx* ctx, uint32_t lib, const uint8_t* quals, size_t n_quals, const uint64_t* qoffsets);
/* 16 features per count-table entry of the library, in the entries' sorted order: its
 * count, the mean quality of its occurrences (0 without qualities), and statistics of the
 * counts of its 3k Hamming-1 neighbours, relative to the library's cutoff; n must equal
 * kc_info.n_counts. */
int kc_library_features(kc_ctx* ctx, uint32_t lib, float* out, size_t n);
/* Features -> network -> label per entry: 1 (solid) when p >= 1 - margin, 0 when
 * p <= margin, else the exact rule count >= cutoff, and fallback[i] = 1 there.  Either
 * output may be NULL; *ms (nullable) receives the device time of the whole pass. */
int kc_library_learned_select(kc_ctx* ctx, uint32_t lib, kc_mlp* mlp, float margin, uint8_t* labels,
                              uint8_t* fallback, size_t n, double* ms);

/* The library-level margin of the learned first pass (no counterpart in the reference):
 * whether the library's cutoff (determine_solid_kmer_cutoff, assemble.rs:692-735) is too
 * fragile to trust a learned selection, so that the caller runs the exact Hammer-on
 * selection for it instead.  Fragile: the valley search finds no valley (cutoff 2 by
 * default), there are at most 100 entries or the 95th-percentile count is under 30, or the
 * search's outcome changes on any of `resamples` Poisson resamplings of the histogram
 * (seeded; the same seed gives the same decision).  Never fragile for min_coverage > 5,
 * where the cutoff is fixed.  kc_library_cutoff_fragile reads the library's count
 * histogram after kc_count; kc_cutoff_is_fragile takes one on the host (hist[c] = entries
 * with count c, the last of n_bins bins those with more) and needs no GPU. */
int kc_library_cutoff_fragile(kc_ctx* ctx, uint32_t lib, uint32_t resamples, uint64_t seed, int32_t* fragile);
int kc_cutoff_is_fragile(const uint64_t* hist, uint32_t n_bins, uint64_t n_counts, double min_coverage,
                         uint32_t resamples, uint64_t seed, int32_t* fragile);

/* Per-kernel device time of the last run (KC_FLAG_TIMING): fills up to cap entries of
 * names (static strings), total ms and launch counts; returns the number of kernels. */
int kc_kernel_times(const kc_ctx* ctx, const char** names, double* ms, uint32_t* launches, int cap);

#ifdef __cplusplus
}
#endif

#endif /* KC_H */
