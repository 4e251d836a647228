/*
 * whale_amd.h -- C ABI of the MI355X (gfx950) amalgamated-likelihood DP.
 *
 * This is the drop-in boundary for Whale.jl's likelihood path.  The reference
 * has no FFI of its own; its boundary is Julia dispatch on
 *     Distributions.logpdf(::WhaleModel, ::CCD)            src/core.jl:39-43
 *     logpdf!(::WhaleModel, ::CCD)                         src/core.jl:29-37
 *     logpdf / logpdf!(::WhaleModel, ::Vector{<:CCD})      src/core.jl:46-64
 *     (m::WhaleModel)(θ) reparameterisation                src/model.jl:147-160
 *     read_ale(path, wm) / CCD(ale, wm, spmap)             src/ccd.jl:100-139
 * A Julia front binds the functions below with `ccall`; the Python front in
 * whale.jl_amd/ binds them with ctypes.  Each function names the reference
 * code it replaces.
 *
 * Conventions
 *   - every function returns 0 (WHALE_OK) or a WHALE_E* code; the message of
 *     the last failure on the calling thread is whale_last_error();
 *   - node arrays are indexed by reference node id - 1 (ids as assigned by
 *     WhaleModel: leaves first, then the other non-WGD nodes in postorder, then
 *     WGD nodes; src/model.jl:121-139);
 *   - host buffers are owned by the caller; device memory is owned by the
 *     handles; calls are synchronous (they return after the device work);
 *   - a zero likelihood yields -Inf, never an error (src/core.jl:15,36);
 *   - a context (and everything created from it) must be used by one host
 *     thread at a time.
 */
#ifndef WHALE_AMD_H
#define WHALE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  WHALE_OK = 0,
  WHALE_EINVAL = 1,  /* bad argument or inconsistent input            */
  WHALE_EPARSE = 2,  /* malformed ALE text (src/ccd.jl:150 assertion) */
  WHALE_EHIP = 3,    /* HIP runtime failure                           */
  WHALE_ENOMEM = 4,  /* device or host allocation failed              */
  WHALE_ESTATE = 5   /* call made in the wrong state                  */
};

/* SamplingCondition (src/condition.jl:1-29) */
enum { WHALE_COND_ROOT = 0, WHALE_COND_NONEXTINCT = 1, WHALE_COND_NONE = 2 };

/* node kinds (src/model.jl:54 iswgd; NewickTree isleaf/isroot) */
enum { WHALE_NODE_LEAF = 0, WHALE_NODE_SPECIATION = 1, WHALE_NODE_WGD = 2, WHALE_NODE_ROOT = 3 };

/* arithmetic of the DP */
enum {
  WHALE_FP64_PLAIN = 0,   /* Float64 as the reference: underflows to -Inf like it      */
  WHALE_FP32_SCALED = 1,  /* FP32 ℓ with per-clade power-of-two exponents              */
  WHALE_FP64_SCALED = 2   /* FP64 ℓ with per-clade exponents (no underflow)            */
};

/* how a batch is launched */
enum { WHALE_LAUNCH_BATCHED = 0, WHALE_LAUNCH_PER_FAMILY = 1 };

typedef struct whale_ctx whale_ctx;
typedef struct whale_model whale_model;
typedef struct whale_ccd whale_ccd;
typedef struct whale_batch whale_batch;

/* Message of the last failed call on this thread ("" if none). */
const char* whale_last_error(void);

/* One HIP device, one stream.  No reference counterpart (the reference runs
 * on Julia threads, src/core.jl:49). */
int whale_ctx_create(int device, whale_ctx** out);
int whale_ctx_destroy(whale_ctx* ctx);
/* The hipStream_t the context launches on (for event timing by callers). */
void* whale_ctx_stream(whale_ctx* ctx);

/* Species-tree model.  Replaces WhaleModel(rates, Ψ, Δt; minn, maxn,
 * condition) (src/model.jl:96-145): the caller has already discretised the
 * tree (n_slices = k_e, slice_len = t_e/k_e as src/model.jl:16-22) and
 * numbered it as the reference does.
 *   parent[i]       parent id-1, or -1 for the root
 *   kind[i]         WHALE_NODE_*
 *   child0/child1   child id-1 (speciation/root: both; WGD: child0; else -1)
 *   n_slices[i]     k_e >= 0 (rows k_e + 1)
 *   slice_len[i]    Δ of the slices of the branch above the node
 *   leaf_prob[i]    leafℙ (1, or 1/m in a MUL group; src/model.jl:110-111)
 *   wgd_index[i]    0-based index into q for WGD nodes (wgdid - 1), else -1
 *   leaf_species[i] species index in [0, n_species) for leaves, else -1
 *                   (the clade id of setclade!, src/model.jl:32-42)
 *   order           the n_nodes ids-1 of wm.order (children before parents)
 */
int whale_model_create(whale_ctx* ctx, int32_t n_nodes, const int32_t* parent, const int32_t* kind,
                       const int32_t* child0, const int32_t* child1, const int32_t* n_slices,
                       const double* slice_len, const double* leaf_prob, const int32_t* wgd_index,
                       const int32_t* leaf_species, int32_t n_species, const int32_t* order,
                       int32_t condition, whale_model** out);
int whale_model_destroy(whale_model* m);

/* Parameters on the natural scale, per node: the getθ of src/rmodels.jl:
 * 31-33, 55-64 already applied by the caller (WGD nodes carry their
 * non-WGD child's λ, μ; NaN where the reference has NaN).  q[w] per WGD
 * index, p[i] sampling probability of leaf i (ignored elsewhere).  Builds the
 * slice tables on the device (setmodel!, src/model.jl:162-191). */
int whale_model_set_params(whale_model* m, const double* lambda, const double* mu, const double* q,
                           int32_t n_q, double eta, const double* p);

/* Slice-table layout and contents (for inspection): node i's rows are
 * [off[i], off[i+1]) of eps/phi/psi (columns 2-4 of Slices.slices). */
int whale_model_table_offsets(const whale_model* m, int32_t* off);
int whale_model_tables(whale_model* m, double* eps, double* phi, double* psi);

/* ALE v0.4 ingest: parse_aleobserve + CCD(ale, wm, spmap)
 * (src/ccd.jl:102-121, 147-248) for one file's text.  species_names[k] maps
 * to species index species_ids[k] (spmap of src/ccd.jl:128). */
int whale_ccd_parse(const char* text, size_t len, const char* const* species_names,
                    const int32_t* species_ids, int32_t n_names, whale_ccd** out);
/* read_ale over many files (src/ccd.jl:126-137 maps CCD(...) over them one by
 * one): whale_ccd_parse of texts[i] (lens[i] bytes) for every i < n, on up to
 * n_threads host threads (<= 0: the hardware threads, at most 16).  On
 * success out[i] holds family i.  If a text does not parse, no handle is
 * returned: *bad (when not NULL) is the lowest such index, and the error
 * message names it. */
int whale_ccd_parse_many(const char* const* texts, const size_t* lens, int32_t n,
                         const char* const* species_names, const int32_t* species_ids, int32_t n_names,
                         int32_t n_threads, whale_ccd** out, int32_t* bad);
int whale_ccd_shape(const whale_ccd* c, int32_t* n_clades, int32_t* n_splits);
/* Clades in the reference's order (by size, then ALE set id; 0-based ids).
 * clade_size[n_clades], leaf_species[n_clades] (-1 for non-leaf clades),
 * split_off[n_clades + 1], g1/g2/p[n_splits] (0-based clade ids). */
int whale_ccd_export(const whale_ccd* c, int32_t* clade_size, int32_t* leaf_species, int32_t* split_off,
                     int32_t* g1, int32_t* g2, double* p);
int whale_ccd_destroy(whale_ccd* c);

/* Device-resident batch of families (the Vector{CCD} of src/core.jl:46-64),
 * flattened: family f owns clades [fam_clade_off[f], fam_clade_off[f+1]);
 * clade c owns splits [split_off[c], split_off[c+1]) of g1/g2/p, with g1/g2
 * family-local clade ids.  Within a family clades must be sorted by size,
 * leaf clades (size 1) carry their species in leaf_species, the last clade is
 * the ubiquitous one, and every split's parts are earlier clades.  Builds the
 * per-family execution plan (compatible clades per branch, stage schedule,
 * split and inverse-split lists) and uploads it. */
int whale_batch_create(whale_model* m, int32_t n_families, const int64_t* fam_clade_off,
                       const int32_t* clade_size, const int32_t* leaf_species, const int64_t* split_off,
                       const int32_t* g1, const int32_t* g2, const double* p, whale_batch** out);
int whale_batch_destroy(whale_batch* b);
/* Workload counts: [families, clades, splits, cells, split_terms, sum_compat,
 * items, records, stages, row_steps, max_stage_items, ell_cells]. */
int whale_batch_stats(const whale_batch* b, int64_t* out12);

/* Number of families the last FP32-scaled call recomputed in FP64 because
 * their FP32 values left the representable range.  No reference counterpart. */
int whale_batch_fallbacks(const whale_batch* b, int32_t* n);

/* Launch shapes of the batched kernels for `mode`: for the forward and the
 * reverse sweep of each of the WHALE_SIZE_CLASSES size classes (rows of 8 values,
 * forward rows first):
 * [families, families per workgroup, slice tables in LDS (0/1), grid,
 *  LDS bytes per workgroup, of which shared by its waves, 0 (reserved),
 *  LDS bytes per family].  No reference counterpart (introspection). */
#define WHALE_SIZE_CLASSES 6
int whale_batch_launches(const whale_batch* b, int32_t mode, int64_t* out);

/* Families whose largest stage does not fit one compute unit's LDS in the
 * FP64-scaled form (the reference evaluates families of any size, src/core.jl:
 * 83-158; example/example-4/cytp450.ale holds 4,917 clades): they run one per
 * workgroup with their stage buffers in a global workspace.  *n receives how
 * many families of the batch take that path.  No reference counterpart. */
int whale_batch_hbm_families(const whale_batch* b, int32_t* n);
/* Lowers the LDS a family's stages may use before it takes that path (bytes;
 * 0 restores the device's limit) and re-plans the batch: with a small limit
 * every family runs with its stage buffers in global memory.  No reference
 * counterpart (testing and tuning). */
int whale_batch_set_lds_limit(whale_batch* b, int64_t bytes);

/* Launch-shape autotuner (no reference counterpart): searches the batch's
 * size-class LDS bounds, the spread of the classes over the streams (one of
 * six layouts) and the waves per compute unit a launch aims for, timing
 * captured evaluations of `mode` (logL + gradient when `gradient` != 0) on
 * the batch itself, `reps` timed runs per candidate (median; a change is kept
 * when at least 1 % faster).  The result is applied to the batch and, when
 * cache_path is not NULL, appended to that text file under a key of the
 * device, the arithmetic and the workload's shape; a later call whose key is
 * in the file applies the cached shape without searching.  Parameters must be
 * set; the evaluations overwrite the batch's last results.
 *   out8 (may be NULL): [5 class bounds in bytes, layout, waves per CU,
 *                        candidates timed (0 when taken from the cache)]
 *   out_ms2 (may be NULL): [ms per evaluation before, after] */
int whale_batch_autotune(whale_batch* b, int32_t mode, int32_t gradient, int32_t reps, const char* cache_path,
                         int64_t* out8, double* out_ms2);
/* Applies / reads a launch shape: [5 class bounds in bytes (increasing,
 * >= 1024), layout 0..5, waves per CU 1..32].  No reference counterpart. */
int whale_batch_set_tuning(whale_batch* b, const int64_t* in7);
int whale_batch_get_tuning(const whale_batch* b, int64_t* out7);
/* Per-family weights of the reverse sweep for the following gradient calls
 * (n_families finite values >= 0; NULL restores weight 1): the gradient
 * becomes that of Σ_f seeds[f]·(logL_f − c), c the conditioning term, i.e.
 * each family's derivative is scaled by its weight and the conditioning
 * term's by the weights' sum.  The value and per-family results are
 * unchanged.  No reference counterpart: the reference differentiates
 * logpdf(MixtureModel, xs) (src/core.jl:66-76) with ForwardDiff, whose
 * derivative for component j is this gradient with the families'
 * responsibilities for j as weights. */
int whale_batch_set_seeds(whale_batch* b, const double* seeds);

/* logpdf of a finite mixture of models over the batch's families, and its
 * gradient (src/core.jl:66-76: Σ_i logsumexp_j (ℓ_ij + log prior_j − c_j)).
 * The K components (1..8) are models with the batch's species-tree structure
 * and their own parameters (set); every size-class launch evaluates all K
 * components (its grid K times the single model's, each workgroup working
 * for one component with that component's tables and workspace), a mixture
 * kernel forms the log-sum-exp and the responsibilities r_ij on the device,
 * and with `gradient` the reverse sweeps of all components run as one launch
 * per class with r_ij as the families' weights.  Not with deterministic
 * gradients, an ℓ budget in force, or families whose stages do not fit LDS.
 *   out_total: the mixture log-likelihood
 *   out_fam (n_families × K, may be NULL): unconditioned ℓ_ij, family-major
 *   out_resp (n_families × K, may be NULL): r_ij
 *   out_grad (K × (2n + n_wgd + 1), with gradient): per component
 *            [dλ (n), dμ (n), dq (n_wgd), dη] on the natural scale
 *   out_dprior (K, may be NULL): d/dprior_j = Σ_i r_ij / prior_j */
int whale_mixture_eval(whale_batch* b, whale_model* const* comps, int32_t K, const double* prior, int32_t mode,
                       int32_t gradient, double* out_total, double* out_fam, double* out_resp, double* out_grad,
                       double* out_dprior);

/* *out = 1 when batched evaluations of `mode` use the pooled launch (one
 * launch whose compute units share their LDS as a pool of pages), else 0.
 * That launch is switched off in this build (layout 6 is rejected), so *out
 * is 0.  No reference counterpart (introspection). */
int whale_batch_pooled(whale_batch* b, int32_t mode, int32_t* out);

/* Memory budget (bytes) for the stored rows of ℓ, with ī and the record
 * coefficients of the reverse sweep; 0 (the default) removes it.  When the
 * whole batch's rows exceed it, evaluations that do not keep ℓ for the caller
 * (whale_logpdf without WHALE_KEEP_ELL, whale_logpdf_grad) run each family's
 * forward DP and reverse sweep in one pass of one wave, whose slot holds that
 * family's rows only while the wave works on it: the rows take (resident
 * waves) x (largest family) instead of growing with the batch, the launches'
 * grids shrinking to fit the budget (one workgroup per launch at the least).
 * Results equal the stored form's.  Calls that keep ℓ (logpdf!, download,
 * backtracking) still store every row.  The reference keeps every ℓ in its
 * CCDs (src/ccd.jl:86); no counterpart. */
int whale_batch_set_ell_budget(whale_batch* b, int64_t bytes);
/* [bytes the stored form needs for the whole batch, bytes of the slots the
 * budgeted form uses (largest planned mode; 0 before its first call), bytes
 * of both the batch holds now]. */
int whale_batch_ell_bytes(const whale_batch* b, int64_t* out3);

/* Deterministic gradients (on != 0): repeated whale_logpdf_grad calls on the
 * same parameters give bit-identical results, as an HMC sampler's reversible
 * trajectories need.  The reverse sweep then runs one family per workgroup
 * with its accumulators in that workgroup's LDS, each family's partial
 * derivatives go to its own slot (plain stores), and the slots are summed in
 * family order; by default the partial derivatives are added atomically in
 * whatever order the workgroups finish, which varies in the last bits
 * (about 1e-9 relative in FP64, 1e-7 in FP32).  The reference's ForwardDiff
 * gradient is deterministic; no counterpart. */
int whale_batch_set_deterministic(whale_batch* b, int32_t on);

/* logpdf of every family (unconditioned, src/core.jl:31-43) into
 * out_family[n_families] (may be NULL) and the conditioned batch value
 * ℓhood(Σ ℓ_i − N·condition) (src/core.jl:54,63) into *out_total. */
int whale_logpdf(whale_batch* b, int32_t mode, int32_t launch, double* out_family, double* out_total);

/* whale_logpdf (batched) that allocates nothing: no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops that forbid hipMalloc).  Its workspace is
 * a fixed device array of WHALE_NOALLOC_SCRATCH_BYTES that every GPU the
 * library runs on holds until the process ends; a batch that needs more
 * (ℓ without stored rows: 8 bytes per cell in the FP64 forms, 12 in the FP32
 * form, plus per-clade exponents and per-family flags) is rejected with
 * WHALE_EINVAL.  d_family (n_families doubles) and d_total (one double) are
 * device memory the caller owns; they receive the per-family values and the
 * conditioned total.  The call is stream-ordered: it returns once the work is
 * enqueued on the context's stream (whale_ctx_stream), and the results are
 * there when that stream reaches this point.  The array is shared by every
 * call on one GPU and keeps the last call's contents (each call clears what
 * must start at zero): calls that overlap on one GPU, from two contexts or two
 * threads, race. */
#define WHALE_NOALLOC_SCRATCH_BYTES (64ull << 20)
int whale_logpdf_noalloc(whale_batch* b, int32_t mode, double* d_family, double* d_total);

/* logpdf plus its gradient with respect to the natural-scale per-node rates
 * (d_lambda, d_mu: n_nodes), q (d_q: n_q) and η (d_eta), by a reverse sweep
 * of the same DP.  The reference obtains this with ForwardDiff
 * (test/runtests.jl:36-38); the branch of getα taken at |λ−μ| ≤ 1e-6 is the
 * one differentiated, as ForwardDiff does. */
int whale_logpdf_grad(whale_batch* b, int32_t mode, int32_t launch, double* out_family, double* out_total,
                      double* d_lambda, double* d_mu, double* d_q, double* d_eta);

/* ℓ of one family from the last whale_logpdf/whale_logpdf_grad call, in the
 * reference layout (logpdf! leaves it in CCD.ℓ, src/core.jl:29, for
 * backtracking, src/track.jl:188-194): for node ids 1..n in turn, the
 * (k_e+1) × #compat column-major matrix ℓ[e].  Values are unscaled doubles.
 * *n_out receives the number of values; nothing is written if cap is too
 * small. */
int whale_download_ell(whale_batch* b, int32_t family, double* out, int64_t cap, int64_t* n_out);

/* Stochastic backtracking of reconciled trees from the ℓ the last call
 * stored (backtrack(wm, ccd) after logpdf!(wm, ccd), src/track.jl:188-414),
 * on the device, one lane per (family, sample): n_samples trees for each of
 * the families [family, family + n_families).  Events are drawn with the
 * reference's weights and in its order; the uniform numbers come from a
 * counter-based generator of (seed, family, sample), so a call is
 * reproducible.  Tree s of family family + i is
 *   counts[i * n_samples + s]   its node count, or -1 if it needs more than cap
 *   nodes[((i * n_samples + s) * cap + k) * 4 + 0..3] for k < count:
 *       clade (0-based, -1 for a loss leaf), species node (id - 1), slice row t
 *       (1-based, as SliceState.t; 0 for a loss), parent node (-1 for the root)
 * in the order the nodes are created (a node is created when the walk enters
 * a new (clade, branch) pair, src/track.jl:172-180).  Needs the last call to
 * have stored ℓ (WHALE_KEEP_ELL or the gradient); families whose likelihood
 * is zero are refused. */
int whale_backtrack(whale_batch* b, int32_t family, int32_t n_families, int32_t n_samples, uint64_t seed,
                    int32_t cap, int32_t* counts, int32_t* nodes);

/* HIP events on the context's stream for callers' own timing: record event
 * `slot` (0..15) now; elapsed ms between two recorded slots (synchronises on
 * the later one).  No reference counterpart. */
int whale_ctx_record(whale_ctx* ctx, int32_t slot);
int whale_ctx_elapsed(whale_ctx* ctx, int32_t from_slot, int32_t to_slot, double* ms);

/* Per-kernel device time of the last call (ms), measured with HIP events on
 * the context's stream when timing is enabled: [tables, forward, backward,
 * gradient tail, total].  No reference counterpart. */
int whale_ctx_set_timing(whale_ctx* ctx, int32_t enabled);
int whale_last_timings(whale_ctx* ctx, double* ms5);

#ifdef __cplusplus
}
#endif
#endif /* WHALE_AMD_H */
