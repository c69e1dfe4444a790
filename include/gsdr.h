/*
 * gsdr.h — C ABI of the MI355X (gfx950) exact greedy DNA de-redundancy engine.
 *
 * The reference program (JuZhenCS/gene-sequences-de-redundancy, root variant)
 * has no plugin or FFI layer; its path sits behind the stage API of func.h and
 * the `cluster` command line. Each entry point below names the reference
 * interface it replaces. Conventions:
 *   - every call returns an int status (GSDR_OK == 0); nothing throws;
 *   - host buffers are owned by the caller, device memory by the context;
 *   - one context per host thread; a context is not re-entrant;
 *   - calls block, except the serving submit/step/poll family.
 */
#ifndef GSDR_H
#define GSDR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GSDR_OK 0
#define GSDR_E_INVALID 1   /* bad argument or option */
#define GSDR_E_IO 2        /* file could not be read or written */
#define GSDR_E_INPUT 3     /* malformed or unsupported input */
#define GSDR_E_DEVICE 4    /* HIP runtime error */
#define GSDR_E_NOMEM 5     /* allocation failed */
#define GSDR_E_LIMIT 6     /* an implementation size limit was exceeded */
#define GSDR_E_BUSY 7      /* serving: no slot or memory free right now */
#define GSDR_E_NOTFOUND 8  /* serving: unknown job id */

/* Longest gap-free sequence the engine aligns as a representative: 32,767
 * bases, the same bound as the reference's local array (func.cpp:752-754). */
#define GSDR_MAX_GAPFREE 32767

typedef struct gsdr_ctx gsdr_ctx;
typedef struct gsdr_records gsdr_records;
typedef struct gsdr_store gsdr_store;
typedef struct gsdr_server gsdr_server;

enum gsdr_mode {
  GSDR_MODE_PLAIN = 0,   /* one representative per round, kernels launched per round */
  GSDR_MODE_BATCHED = 1  /* persistent rounds, speculative representatives, no host sync per round */
};

typedef struct gsdr_params {
  float threshold;      /* t, parsed as float like std::stof (func.cpp:21); 0.8 <= t < 1 */
  int32_t word_length;  /* k; 0 = from t (func.cpp:36-45), else 4..8 */
  int32_t mode;         /* enum gsdr_mode */
  int32_t spec_reps;    /* batched: representatives evaluated per speculative step (0 = default) */
  int32_t timing;       /* 1 = time every filter/align launch with HIP events */
  int32_t tune;         /* 1 = use the context's knobs from gsdr_autotune / gsdr_set_tuning */
  int32_t drop_filters; /* 1 = the cuda-version's drop mode (cuda-version/func.cu:1068-1097):
                           no composition or shared-word filter, every pair past the length
                           test is aligned; clusters then differ from root's. Default 0. */
  int32_t reserved;
} gsdr_params;

typedef struct gsdr_stats {
  int64_t n_seqs;       /* records clustered */
  int64_t n_bases;      /* raw bases (sum of record data lengths) */
  int64_t reps;         /* representatives written */
  int64_t rounds;       /* host-visible rounds (plain) or persistent steps (batched) */
  int64_t launches;     /* kernel launches issued */
  int64_t pairs;        /* (representative, candidate) pairs considered */
  int64_t rej_length;   /* rejected: min(gap-free lengths) < ceil(L*t)+1 (added, sound) */
  int64_t rej_comp;     /* rejected by the composition filter, incl. the sound tightening */
  int64_t rej_word;     /* rejected by the shared-word filter (func.cpp:544-571) */
  int64_t aligned;      /* pairs that reached the alignment */
  int64_t rej_align;    /* aligned and rejected */
  int64_t accepted;     /* aligned and absorbed into the representative's cluster */
  double cells_band;    /* sum of exact LCS decision-band cells over aligned pairs */
  double cells_exec;    /* bit-cells the alignment kernels actually computed */
  double cells_full;    /* sum of n*m over aligned pairs (the reference computes all) */
  double ms_total;      /* wall time of the clustering call */
  double ms_filter;     /* filter kernel time (timing == 1) */
  double ms_align;      /* align kernel time (timing == 1) */
  int64_t filter_launches;
  int64_t align_launches;
  double filter_bytes;  /* algorithmic bytes read+written by the filter kernels */
  double spec_wasted;   /* batched: pair evaluations discarded by speculation */
  double cells_decided; /* band cells of the rows each exact decision needed (32-row steps) */
  double rej_pigeon;    /* filtered pairs rejected by the exact pigeonhole test (tuning.pigeon = 1);
                           aligned = accepted + rej_align + spec_wasted + rej_pigeon */
  double pigeon_checked; /* filtered pairs the pigeonhole test examined */
  double screened;      /* pairs the word-signature screen rejected (k >= 6; part of rej_word) */
} gsdr_stats;

/* Launch knobs that change speed, never results. */
typedef struct gsdr_tuning {
  int32_t swap_ratio;     /* align with the roles swapped when n_rep >= swap_ratio * m_cand (0 = never) */
  int32_t spec_max;       /* batched: cap on speculative representatives per step (1..64) */
  int32_t filter_blocks;  /* filter workgroups per CU (-1, the default: by the call, 32 where
                             the shapes run concurrently with 4 align workgroups per CU, else 8) */
  int32_t align_blocks;   /* alignment workgroups per CU, per tile shape; 0 = each shape
                             kernel's resident blocks per CU; -1 (default) = by the call: those
                             when its jobs of threshold <= 0.90 hold at least 32 Mi bases and no
                             text needs a tile wider than 64x4; 2 where fork_shapes -1 runs the
                             shapes one after another; else 4 */
  int32_t trials;         /* autotune: clusterings timed */
  int32_t fork_shapes;    /* 1 = tile-shape kernels concurrent on side streams, 0 = one after another,
                             -1 (default) = by the call: one after another where few pairs make a
                             step (no tile wider than 64x4 needed, or every threshold >= 0.99) and
                             the call is not one of many short aligned pairs */
  int32_t list_factor;    /* batched: alignment list capacity per tile shape, in items per
                             sequence (default 64); it also bounds k_spec * candidates */
  int32_t pigeon;         /* 1 = exact pigeonhole prefilter before the alignment (sound; default 0) */
  int32_t screen;         /* 1 = word-signature screen before the filter for k >= 6 (sound; default 1) */
  double ms_default;      /* autotune: clustering time with the default knobs */
  double ms_tuned;        /* autotune: clustering time with the chosen knobs */
  /* Tile shapes per bin of text length (gap-free bases < 512, < 2,048, < 8,192, longer; the
   * shapes, in order: 8x1 16x1 32x1 64x1 64x2 64x3 64x4 64x6 64x8 64x10 64x12 64x16, lanes x
   * 32-bit words per lane). shape_mask[b]: the shapes (bit i = shape i) a pair of bin b whose
   * band a shape's window holds may take, 0 = all; shape_bias[b]: allowed shapes to go past
   * the narrowest that holds the band (0..3). A shape that holds the whole text is always
   * allowed and ends the search. */
  int32_t shape_mask[4];
  int32_t shape_bias[4];
  int32_t win_period;     /* 64-lane tiles: 32-row blocks between window rebases at the DP edge
                             (0 = 2 with 6 or more words per lane, 4 with fewer; else 1, 2, 4, 8 or 16) */
  int32_t split_rows;     /* batched: a pair with at least this many rows (query bases) is aligned by
                             two waves, half of the rows forward and half backward, joined at the
                             middle row, while the split list has room (one pair per block of the
                             split kernel; the rest run on their tile shapes)
                             (0 = never; -1 = by the call: 128 where every threshold of the call
                             is at least 0.99 or no pair needs more than 4 words per lane, never
                             otherwise; else 64..32768) */
} gsdr_tuning;

/* ---- options (host side): replaces checkOption (func.cpp:6-57) and the
 *      Option struct (func.h:11-16) ---- */
typedef struct gsdr_options {
  const char *input;      /* 'i': default "testData.fasta" (func.cpp:8) */
  const char *output;     /* 'o': default "result.fasta" (func.cpp:9) */
  float threshold;        /* 't': std::stof, default 0.95, 0.8 <= t < 1 (func.cpp:10,21,31-35) */
  int32_t word_length;    /* 'w': 4..8, or from t when not given (func.cpp:11,36-51) */
  /* long options of this build (not in the reference) */
  int32_t mode;           /* --mode plain|batched (default batched) */
  int32_t spec_reps;      /* --spec K */
  int32_t device;         /* --device N */
  int32_t drop_filters;   /* --drop 0|1 */
  const char *membership; /* --membership FILE, or NULL */
  const char *report;     /* --report FILE, or NULL */
} gsdr_options;

/* Parses the reference's argv letter/value pairs (`t 0.9 i in.fasta o
 * out.fasta w 6`; `-t` etc. and the long options above are accepted too).
 * Strings point into argv or to static defaults. Where the reference prints
 * its usage text and exits with status 0 (odd argument count, unknown
 * letter, t or w out of range), this returns GSDR_E_INVALID and
 * gsdr_last_error() holds the reference's message lines; the caller prints
 * them and exits 0 to keep the reference's behaviour. */
int gsdr_parse_options(int argc, char **argv, gsdr_options *out);

/* Cutoffs of n sequences of raw lengths raw_len (createCutoff,
 * func.cpp:339-374), in float32 like the reference: bc = ceilf(L*t) and
 * wc = max(L-k+1 - (int)(ceilf(L*(1-t))*k), 1). word_length 0 = from t. */
int gsdr_cutoffs(float threshold, int32_t word_length, const int32_t *raw_len, int64_t n, int32_t *bc_out,
                 int32_t *wc_out);

/* ---- records (host side): replaces readFile/saveFile (func.cpp:70-112,
 *      785-799) and the Read vector (func.h:17-20) ---- */

/* Reads FASTA with the reference's record split (func.cpp:70-104), in file
 * order. */
int gsdr_records_read_fasta(const char *path, gsdr_records **out);
/* Reads FASTA (func.cpp:70-104) straight into the reference's length order
 * (func.cpp:105-107): the same records and order as read_fasta followed by
 * sort, with one copy of the data instead of two. */
int gsdr_records_read_fasta_sorted(const char *path, gsdr_records **out);
/* Builds records from caller buffers (offsets have count+1 entries). */
int gsdr_records_from_buffers(const char *names, const int64_t *name_off, const char *data,
                              const int64_t *data_off, int64_t count, gsdr_records **out);
/* Synthetic workloads: law GSDR_LAW_*; seed; n > 0 overrides the record
 * count; median > 0 overrides the length median. */
enum {
  GSDR_LAW_FAMILIES = 1,         /* 500 parents, 19 mutated children each, L 300..1500 */
  GSDR_LAW_LOGNORMAL = 2,        /* log-normal lengths, 10 mutated children per parent */
  GSDR_LAW_PARETO = 4            /* Pareto(1.5) lengths, singletons and families */
};
int gsdr_records_generate(int32_t law, uint64_t seed, int64_t n, double median,
                          gsdr_records **out);
int gsdr_records_write_fasta(const gsdr_records *r, const char *path, int32_t line_width);
/* Reorders in place into the reference's length order (func.cpp:105-107);
 * perm_out (may be NULL, count entries) gets the original index per slot.
 * Records already in that order (sorted, or read with read_fasta_sorted)
 * stay as they are and perm_out gets the same permutation again. */
int gsdr_records_sort(gsdr_records *r, int64_t *perm_out);
int64_t gsdr_records_count(const gsdr_records *r);
int64_t gsdr_records_bases(const gsdr_records *r);
int gsdr_records_get(const gsdr_records *r, int64_t i, const char **name, int64_t *name_len,
                     const char **data, int64_t *data_len);
/* Writes representatives (cluster[s] == s) like saveFile (func.cpp:785-796). */
int gsdr_records_write_result(const gsdr_records *sorted, const int32_t *cluster, const char *path,
                              int64_t *reps_out);
/* Writes one line per record: sorted index, original index (perm, may be
 * NULL), representative's sorted index, name. */
int gsdr_records_write_membership(const gsdr_records *sorted, const int64_t *perm,
                                  const int32_t *cluster, const char *path);
void gsdr_records_free(gsdr_records *r);

/* ---- device context: replaces selectDevice (func.cpp:58-69) ---- */
int gsdr_open(int32_t device, gsdr_ctx **ctx);
void gsdr_close(gsdr_ctx *ctx);
int gsdr_device_info(gsdr_ctx *ctx, char *name, int32_t name_len, int32_t *cu_count,
                     int64_t *hbm_bytes);

/* ---- sequence store in HBM: replaces copyData and the eight preprocessing
 *      stages (func.cpp:113-446; func.h:52-60). `sorted` must already be in
 *      the reference's length order. ---- */
int gsdr_load(gsdr_ctx *ctx, const gsdr_records *sorted, int32_t word_length, gsdr_store **store);
void gsdr_store_free(gsdr_store *store);
int64_t gsdr_store_footprint(const gsdr_store *store);

/* ---- clustering: replaces clustering(Option&, Data&, Bench&)
 *      (func.cpp:677-784; func.h:61). cluster_out: caller-owned int32[n];
 *      cluster_out[s] == s marks a representative, otherwise it holds the
 *      representative's sorted index. ---- */
int gsdr_cluster(gsdr_ctx *ctx, gsdr_store *store, const gsdr_params *params,
                 int32_t *cluster_out, gsdr_stats *stats);

/* No counterpart in the reference. gsdr_cluster for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, same results, both
 * forms. The launch workspace (alignment lists, counters, job table) lives in
 * a fixed device array of GSDR_NOALLOC_SCRATCH_BYTES per GPU, held from the
 * first gsdr_open on that GPU until the process ends; the call itself
 * allocates nothing (no hipMalloc, hipHostMalloc, events or graphs). Limits:
 * at most GSDR_NOALLOC_MAX_SEQS sequences (GSDR_E_LIMIT beyond), no timing
 * (params.timing must be 0). Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (several contexts or threads) race. */
#define GSDR_NOALLOC_SCRATCH_BYTES (128ull << 20)
#define GSDR_NOALLOC_MAX_SEQS 600000
int gsdr_cluster_noalloc(gsdr_ctx *ctx, gsdr_store *store, const gsdr_params *params, int32_t *cluster_out,
                         gsdr_stats *stats);

/* Many independent jobs in shared persistent launches (no reference
 * counterpart; each job's result equals gsdr_cluster on that job alone). */
int gsdr_cluster_jobs(gsdr_ctx *ctx, int32_t njobs, gsdr_store *const *stores,
                      const gsdr_params *params, int32_t *const *cluster_out, gsdr_stats *stats);
/* Autotuner (no counterpart in the reference): clusters `store` with params
 * (forced to the batched form) under coordinate descent over the knobs of
 * gsdr_tuning, at most max_trials clusterings (0 = 40), and keeps the fastest
 * set in the context; later calls with params.tune = 1 use it. Every trial
 * must return the same clusters, else GSDR_E_DEVICE. out (may be NULL) gets
 * the chosen knobs and timings. */
int gsdr_autotune(gsdr_ctx *ctx, gsdr_store *store, const gsdr_params *params, int32_t max_trials,
                  gsdr_tuning *out);
/* Sets (tuning != NULL) or clears (NULL) the context's knobs. */
int gsdr_set_tuning(gsdr_ctx *ctx, const gsdr_tuning *tuning);
/* The knobs calls with params.tune = 1 use (the defaults when none are set). */
int gsdr_get_tuning(const gsdr_ctx *ctx, gsdr_tuning *out);

/* ---- pair-level entry (diagnostics and parity tests): decides
 *      LCS(text, query) > bc on the device, the test of kernel_dynamic
 *      (func.cpp:572-676). Sequences are given as raw bytes like FASTA data;
 *      pairs holds npairs (text index, query index) pairs. shape < 0 picks the
 *      tile shape per pair; 0..11 forces tile shape index `shape` (or the
 *      narrowest wider one that holds the pair); 12 + s (s = 0..11) runs every
 *      pair with at least 64 query bases as a split pair (two waves, joined at
 *      the middle row) starting from the window of shape s or wider, with at
 *      least two words per lane. ---- */
int gsdr_align_pairs(gsdr_ctx *ctx, const char *data, const int64_t *data_off, int32_t nseq,
                     const int32_t *pairs, const int32_t *bc, int32_t npairs, int32_t shape,
                     int32_t *accept_out, double *cells_exec_out);

/* ---- serving: continuous batching of clustering jobs ---- */
int gsdr_serve_open(gsdr_ctx *ctx, int64_t hbm_budget, int32_t max_slots, gsdr_server **srv);
/* Admits the job if its footprint fits the budget and a slot is free;
 * otherwise returns GSDR_E_BUSY and the caller retries later. */
int gsdr_serve_submit(gsdr_server *srv, const gsdr_records *sorted, const gsdr_params *params,
                      int64_t *job_id);
/* Advances every admitted job by one persistent step. Returns the number
 * of jobs still running in *running. Should a step overflow the shared
 * alignment lists (the width clamp after every step is there to prevent it),
 * the jobs in the device table fail, their memory is released, the call
 * returns GSDR_E_LIMIT and the server goes on with the jobs admitted later. */
int gsdr_serve_step(gsdr_server *srv, int32_t *running);
/* *done = 1 when finished; then cluster_out (n entries) and stats are filled
 * and the job's device memory is released. A job that failed in a step is
 * reported once with *done = 1 and GSDR_E_LIMIT. */
int gsdr_serve_poll(gsdr_server *srv, int64_t job_id, int32_t *cluster_out, gsdr_stats *stats,
                    int32_t *done);
void gsdr_serve_close(gsdr_server *srv);

const char *gsdr_strerror(int32_t status);
/* Message of the last failing call on this thread. */
const char *gsdr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GSDR_H */
