/*
 * sbp_oracle.h -- CPU oracle for the SBP hot path (test infrastructure only).
 *
 * This is a plain-C restatement of vtsynergy/SBP's single-rank hot path
 * (block merge, finetune, MDL, sampling and the agglomerative driver) under
 * a deterministic replay contract.  It is the checker that the HIP product
 * path is compared against, and the timed CPU baseline.  Nothing in the
 * product library links or calls it.
 *
 * Replay contract (the reference draws from random_device/clock-seeded
 * generators and iterates hash maps, so it is not reproducible as shipped):
 *   - Random draws: Philox4x32-10 (Salmon et al., SC'11), key = seed
 *     (lo, hi 32 bits), counter = {unit, epoch, tag, sub}.  One call yields
 *     four 32-bit words w0..w3.
 *       uniform_below(w, n) = (w * n) >> 32          (n < 2^32)
 *       unit_double(a, b)   = (((a << 32) | b) >> 11) * 2^-53
 *       random_block(r, B, w) = uniform_below(w, B-1), +1 if >= r
 *   - Multinomials select by an integer target over weights listed in a
 *     canonical order: block-matrix row cells by ascending column, then
 *     column cells by ascending row.
 *   - Sums of FP64 terms (MDL log-posterior, delta-MDL, Hastings sums) are
 *     accumulated in 128-bit fixed point: each term is scaled by 2^64 and
 *     truncated toward zero; the total is converted back as
 *     (double)hi + (double)lo * 2^-64.  Integer addition is associative, so
 *     the sum does not depend on the order of the terms.  The MH sweep's
 *     sum of accepted delta-MDL is a double sum in visit order, as in the
 *     reference (with degree correction a term can be infinite).
 *   - The snowball samplers' random_integer(lo, hi) calls are
 *     lo + (w0 (hi - lo + 1)) >> 32 over Philox(seed; j, 0, TAG_SNOWBALL, 0),
 *     j counting the calls; reattach's the same over TAG_REATTACH.
 *   - logf is glibc's (the reference's fastlog / nat_log are logf).
 *   - delta-MDL sums the Delta cells and the cells at rows/columns {r, s} term by term, and the remaining
 *     unchanged cells of rows r, s / columns r, s in bulk: such a cell changes only through one block
 *     degree d -> d', so its old-minus-new entropy is M * (fastlog(d') - fastlog(d)), and M summed over
 *     them is a row/column total minus the other cells.  The per-cell form of the reference is kept as
 *     orc_*_delta_mdl_cells and the tests bound the difference by 1e-12 of the sum of |cell terms|.
 *   - Heap ties in the merge carry-out are broken by ascending block id;
 *     the sample extension breaks plurality ties by the smallest block id;
 *     the vertex shuffle is the ascending order of Philox sort keys.
 *
 * What pins each part (tests/test_oracle.py, tests/test_sampling_oracle.py,
 * tests/test_nonparametric_oracle.py, tests/test_degree_corrected_oracle.py):
 * Philox known-answer vectors; the reference's toy-graph known answers for the
 * block matrix, degrees, MDL, null model, vertex / merge delta-MDL, Hastings,
 * max_degree sampling, the nonparametric description length and the
 * degree-corrected sparse entropy; brute-force description-length differences
 * with glibc logf; restatements of the reference's sampling and extension
 * loops; libstdc++'s heap order (degree_product_snowball) against
 * std::priority_queue; spence against the Cephes dilogarithm and log_q
 * against exact partition counts.  The reference's DEGREE_DL and
 * degree-corrected MDL constants are not reproducible from its code and are
 * not used.  The sweeps (async Gibbs, MH, hybrid, finetune_assignment), the
 * merge proposals, both carry-outs with label compaction, and the driver with
 * its golden-section triplet have no reference output under any replay; they
 * are pinned against a literal Python restatement of the reference's loops
 * (tests/reference_loops.py, tests/test_oracle_loops.py), exact on tiny
 * graphs, with deliberate mistakes in each loop that the comparison rejects,
 * and on invariants (rebuild equality after every sweep and merge,
 * thread-count independence, recovering planted partitions).
 *
 * Tags and epochs: every finetune sweep and every merge phase uses its own
 * epoch value, supplied by the driver (the driver numbers them 0, 1, 2, ...
 * in call order).
 */
#ifndef SBP_ORACLE_H
#define SBP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_TAG_FT      0x46540001u  /* finetune proposal: unit = vertex */
#define ORC_TAG_MERGE   0x4d470001u  /* merge proposal: unit = block, sub = proposal index */
#define ORC_TAG_SHUFFLE 0x53480001u  /* vertex shuffle key: unit = vertex */
#define ORC_TAG_EXTEND  0x45580001u  /* sample extension random block: unit = vertex */
#define ORC_TAG_SAMPLE  0x53410001u  /* random sampling key: unit = vertex */
#define ORC_TAG_SNOWBALL 0x53410002u  /* snowball samplers: unit = draw index */
#define ORC_TAG_REATTACH 0x52410001u  /* reattach: unit = draw index */
#define ORC_TAG_SPLIT_SIZE 0x53500001u  /* split size: unit = block, sub = proposal / 4 */
#define ORC_TAG_SPLIT_SIDE 0x53500002u  /* split side: unit = vertex, sub = proposal / 4 */

typedef struct {
    int32_t N;
    int64_t E;
    const int64_t *out_off;  /* N+1 */
    const int32_t *out_adj;  /* E, file order */
    const int64_t *in_off;   /* N+1 */
    const int32_t *in_adj;   /* E, file order */
} orc_graph;

typedef struct orc_bm orc_bm;

typedef struct {
    int32_t algorithm;        /* 0 async_gibbs, 1 metropolis_hastings, 2 hybrid_mcmc */
    int32_t batches;          /* --batches (default 1) */
    int32_t greedy;           /* --greedy (default true) */
    int32_t approximate;      /* --approximate */
    int32_t ordered;          /* --ordered */
    int32_t asynciterations;  /* --asynciterations */
    double  mh_percent;       /* --mh_percent (default 0.075) */
    uint64_t seed;
    int32_t max_iterations;   /* MAX_NUM_ITERATIONS (100) */
    int32_t threads;          /* OpenMP threads for the proposal loops */
    int32_t nonparametric;    /* --nonparametric: entropy::nonparametric description length */
    int32_t degreecorrected;  /* --degreecorrected: with --nonparametric, the degree-corrected terms */
} orc_params;

typedef struct {
    int64_t proposals;        /* vertex-move proposals evaluated */
    int64_t moves;            /* accepted and committed moves */
    int64_t refused;          /* async-Gibbs moves refused at commit (block size 1) */
    int64_t sweeps;
    int64_t merge_proposals;
    int64_t merge_phases;
    double  mdl;
} orc_stats;

/* Nonparametric description length (entropy::nonparametric), directed */
double   orc_lgamma(int64_t x);
double   orc_lbinom(int64_t n, int64_t k);
double   orc_np_sparse_entropy(const orc_bm *bm);
double   orc_np_partition_dl(const orc_bm *bm);
double   orc_np_edges_dl(int64_t B, int64_t E);
double   orc_np_mdl(const orc_bm *bm, const orc_graph *g);
/* Degree correction (args.degreecorrected): spence (src/spence.cpp), log_q (src/entropy.cpp:688-717),
 * get_deg_dl_dist (src/entropy.cpp:719-742) */
double   orc_spence(double x);
double   orc_log_q(uint64_t n, uint64_t k);
double   orc_np_deg_dl(const orc_bm *bm);

/* RNG / math primitives */
void     orc_philox(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]);
uint32_t orc_uniform_below(uint32_t w, uint32_t n);
double   orc_unit_double(uint32_t a, uint32_t b);
void     orc_logf_range(uint32_t first_bits, uint32_t count, float *out);
void     orc_fx_from_double(double d, uint64_t out[2]);
double   orc_fx_to_double(int64_t hi, uint64_t lo);

/* Blockmodel */
orc_bm  *orc_bm_new(const orc_graph *g, int32_t B, const int32_t *assignment);
orc_bm  *orc_bm_copy(const orc_bm *bm);
void     orc_bm_free(orc_bm *bm);
int32_t  orc_bm_num_blocks(const orc_bm *bm);
int64_t  orc_bm_nnz(const orc_bm *bm);
/* Exports row-major CSR (ascending columns) and the per-block arrays. Any pointer may be NULL. */
void     orc_bm_export(const orc_bm *bm, int32_t *row_ptr, int32_t *cols, int32_t *cnts,
                       int64_t *d_out, int64_t *d_in, int64_t *d_tot, int32_t *sizes, int32_t *assignment);
/* model: 0 parametric, 1 nonparametric, 2 nonparametric degree-corrected (degree histograms from g) */
void     orc_bm_set_model(orc_bm *bm, const orc_graph *g, int32_t model);
int      orc_bm_validate(const orc_bm *bm, const orc_graph *g);

/* Entropy */
double   orc_log_posterior(const orc_bm *bm);
double   orc_mdl(const orc_bm *bm, const orc_graph *g);
double   orc_null_mdl_v1(int64_t E);
double   orc_vertex_delta_mdl(const orc_bm *bm, const orc_graph *g, int32_t v, int32_t s);
double   orc_vertex_hastings(const orc_bm *bm, const orc_graph *g, int32_t v, int32_t s);
double   orc_merge_delta_mdl(const orc_bm *bm, int32_t r, int32_t s);
/* The reference's per-cell evaluation order (entropy.cpp); abs_terms = sum of |cell terms|. */
double   orc_vertex_delta_mdl_cells(const orc_bm *bm, const orc_graph *g, int32_t v, int32_t s, double *abs_terms);
double   orc_merge_delta_mdl_cells(const orc_bm *bm, int32_t r, int32_t s, double *abs_terms);

/* Finetune */
int      orc_ft_propose(const orc_bm *bm, const orc_graph *g, uint64_t seed, uint32_t epoch, int32_t v,
                        int32_t greedy, int32_t *s_out, double *dS_out, int32_t *accept_out);
int      orc_async_gibbs_sweep(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t epoch,
                               int32_t *proposed_s, orc_stats *st);
int      orc_mh_sweep(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t epoch, int32_t ordered,
                      double *sum_dS, orc_stats *st);
int      orc_mcmc(orc_bm *bm, const orc_graph *g, const orc_params *p, int32_t golden_ratio_not_reached,
                  uint32_t *epoch, orc_stats *st);
int      orc_finetune_assignment(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t *epoch,
                                 orc_stats *st);

/* Block merge */
int      orc_merge_propose(const orc_bm *bm, uint64_t seed, uint32_t epoch, int32_t *best_s, double *best_dS);
orc_bm  *orc_merge_apply(orc_bm *bm, const orc_graph *g, const int32_t *best_s, const double *best_dS,
                         int64_t num_to_merge, int32_t approximate);
orc_bm  *orc_merge_blocks(orc_bm *bm, const orc_graph *g, const orc_params *p, uint32_t epoch, orc_stats *st);

/* Driver */
int      orc_sbp(const orc_graph *g, const orc_params *p, int32_t *assignment_out, int32_t *num_blocks_out,
                 orc_stats *st);

/* Sampling */
int64_t  orc_sample(const orc_graph *g, int32_t alg, double fraction, uint64_t seed, int32_t *mapping,
                    int32_t *vertices);
float    orc_degree_weight_key(uint32_t w0, int64_t degree);
int64_t  orc_subgraph(const orc_graph *g, const int32_t *mapping, const int32_t *vertices, int32_t n,
                      int64_t *out_off, int32_t *out_adj, int64_t *in_off, int32_t *in_adj);
int      orc_extend(const orc_graph *g, const int32_t *mapping, const int32_t *sample_assignment,
                    int32_t sample_blocks, uint64_t seed, int32_t *assignment_out);
int64_t  orc_detach(const orc_graph *g, int32_t *mapping, int32_t *vertices);
/* divisive::split_communities (src/top_down_sbp.cpp:488-549), random splits, in place on assign (N entries,
 * blocks 0..B-1); nonparametric selects null_mdl_v1's form.  Returns the new block count; delta_out (B
 * entries, may be NULL) gets each block's kept value (DBL_MAX when not split). */
int32_t  orc_split_communities(const orc_graph *g, int32_t *assign, int32_t B, int32_t target, uint64_t seed,
                               uint32_t epoch, int32_t nonparametric, double *delta_out);
int      orc_reattach(const orc_graph *g, const int32_t *mapping, const int32_t *detached_assignment,
                      int32_t detached_blocks, uint64_t seed, int32_t *assignment_out);

#ifdef __cplusplus
}
#endif
#endif
